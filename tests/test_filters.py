"""Pins of the product's template/target filters (pytme_amd/filters)
against independent restatements: the step wedge against its geometric
support, the step wedge against the continuous wedge of the same axes, and
the CTF against the analytic zeros of its phase.  CPU only.

The filter arrays enter the GPU scan as inputs on both sides of every parity
test, so these tests are what ties their values to the reference's
definitions (tme/filters/wedge.py:271-507, tme/filters/ctf.py:254-335,
508-655).
"""

import numpy as np
import pytest
from scipy.ndimage import binary_dilation

from oracle import bruteforce as bf
from pytme_amd.filters import CTFReconstructed, WedgeReconstructed
from pytme_amd.filters.ctf import electron_wavelength

TILTS = np.arange(-60, 61, 3.0)


def _step(shape, opening_axis, tilt_axis, angles=TILTS):
    w = WedgeReconstructed(angles=angles, opening_axis=opening_axis, tilt_axis=tilt_axis)
    return w(shape=shape, return_real_fourier=False)["data"]


def _continuous(shape, opening_axis, tilt_axis, angles=(-60, 60)):
    w = WedgeReconstructed(angles=angles, create_continuous_wedge=True, opening_axis=opening_axis,
                           tilt_axis=tilt_axis)
    return w(shape=shape, return_real_fourier=False)["data"]


class TestStepWedge:
    @pytest.mark.parametrize(
        "shape,opening_axis,tilt_axis",
        [((48, 48, 48), 2, 0), ((64, 64, 64), 2, 0), ((48, 48, 48), 0, 2), ((64, 64, 64), 0, 2),
         ((47, 47, 47), 2, 0), ((47, 40, 56), 2, 0), ((48, 41, 56), 0, 2), ((48, 40, 56), 1, 0),
         ((33, 20, 17), 1, 2)],
    )
    def test_equals_geometric_support(self, shape, opening_axis, tilt_axis):
        got = _step(shape, opening_axis, tilt_axis)
        assert set(np.unique(got)) <= {0.0, 1.0}
        want, edge = bf.step_wedge_support(shape, TILTS, opening_axis, tilt_axis)
        # identical except on voxels at distance 1 (to 1e-5) from a line,
        # where float32 matrices decide; those are a tiny minority
        assert np.array_equal((got > 0)[~edge], want[~edge])
        assert edge.mean() < 0.01

    def test_geometric_support_tells_axes_and_angles_apart(self):
        # plausible mistakes give a different support: swapped axes, another
        # tilt range, a coarser step
        got = _step((48, 40, 56), 2, 0) > 0
        for args in ((TILTS, 0, 2), (np.arange(-54, 55, 3.0), 2, 0), (np.arange(-60, 61, 6.0), 2, 0)):
            want, edge = bf.step_wedge_support((48, 40, 56), *args)
            assert np.sum((got != want) & ~edge) > 100

    @pytest.mark.parametrize("shape", [(48, 48, 48), (64, 64, 64), (48, 40, 56)])
    @pytest.mark.parametrize("axes", [(2, 0), (0, 2), (1, 2)])
    def test_matches_continuous_wedge(self, shape, axes):
        # the sampled lines fill the double cone of the continuous wedge with
        # the same axes; on cubes each lies within two voxels of the other
        # (interpolated lines are up to two voxels wide; on other shapes the
        # two use slightly different aspect corrections, P1 being odd)
        step = np.fft.fftshift(_step(shape, *axes) > 0)
        cont = np.fft.fftshift(_continuous(shape, *axes) > 0)
        if len(set(shape)) == 1:
            assert not np.any(step & ~binary_dilation(cont, iterations=2))
            assert not np.any(cont & ~binary_dilation(step, iterations=2))
        # with the axes exchanged the two disagree on a large fraction
        swapped = np.fft.fftshift(_continuous(shape, axes[1], axes[0]) > 0)
        assert np.mean(step != swapped) > 0.2 and np.mean(step != cont) < 0.05

    @pytest.mark.parametrize("axes", [(2, 0), (0, 2)])
    def test_tilt_axis_sampled_opening_axis_missing(self, axes):
        opening, tilt = axes
        n = 64
        c = np.fft.fftshift(_step((n, n, n), opening, tilt) > 0)
        line = [n // 2] * 3
        for i in range(n):  # the whole tilt axis is sampled (|k| <= 0.5)
            line[tilt] = i
            assert c[tuple(line)]
        line = [n // 2] * 3
        for i in list(range(0, n // 2 - 2)) + list(range(n // 2 + 3, n)):
            line[opening] = i  # the opening axis lies in the missing wedge
            assert not c[tuple(line)]
        # without the spherical cutoff: constant along the remaining axis
        w = WedgeReconstructed(angles=TILTS, opening_axis=opening, tilt_axis=tilt, frequency_cutoff=None)
        u = w(shape=(n, n, n))["data"]
        other = 3 - opening - tilt
        assert np.array_equal(np.take(u, 5, axis=other), np.take(u, 40, axis=other))

    def test_half_spectrum_is_the_rfft_crop(self):
        w = WedgeReconstructed(angles=TILTS, opening_axis=2, tilt_axis=0)
        full = w(shape=(48, 40, 56), return_real_fourier=False)["data"]
        half = w(shape=(48, 40, 56), return_real_fourier=True)["data"]
        assert half.shape == (48, 40, 29) and np.array_equal(half, full[..., :29])


class TestCTF:
    @staticmethod
    def _zeros(n_max, dz, cs, lam, amp):
        """k where chi(k) = n pi for chi(u) = pi lam dz u - pi/2 cs lam^3 u^2
        + atan(A / sqrt(1 - A^2)), u = k^2, solved as a quadratic in u
        (smaller root); units of lam, cs and 1/k must agree."""
        phi = np.arctan(amp / np.sqrt(1 - amp * amp))
        a = 0.5 * np.pi * cs * lam**3
        b = -np.pi * lam * dz
        out = []
        for n in range(1, n_max + 1):
            c = n * np.pi - phi
            disc = b * b - 4 * a * c
            if disc < 0:
                break
            u = (-b - np.sqrt(disc)) / (2 * a) if a > 0 else -c / b
            out.append(np.sqrt(u))
        return np.array(out)

    def _check_zeros(self, ctf, zeros, n):
        pos = ctf[n // 2:]
        k = np.arange(pos.size) / n
        sign_change = np.nonzero(np.sign(pos[1:]) != np.sign(pos[:-1]))[0]
        zeros = zeros[zeros < k[-1]]
        assert len(zeros) >= 3 and len(sign_change) == len(zeros)
        for z, i in zip(zeros, sign_change):
            assert k[i] <= z <= k[i + 1]

    @pytest.mark.parametrize("dz,cs", [(5000.0, 2.7e7), (20000.0, 2.7e7), (8000.0, 0.0)])
    def test_zeros_physical_units_at_one_angstrom(self, dz, cs):
        # at 1 A/voxel, k in cycles/voxel is k in 1/A: the physical CTF
        # (lambda(300 kV) in A, Cs in A, defocus in A)
        n = 512
        f = CTFReconstructed(defocus_x=dz, sampling_rate=1.0, acceleration_voltage=300e3, spherical_aberration=cs,
                             amplitude_contrast=0.07, flip_phase=False)
        ctf = np.fft.fftshift(f(shape=(n,))["data"])
        self._check_zeros(ctf, self._zeros(400, dz, cs, electron_wavelength(300e3), 0.07), n)
        if cs > 0:  # the spherical-aberration term moves the zeros
            no_cs = self._zeros(400, dz, 0.0, electron_wavelength(300e3), 0.07)
            with pytest.raises(AssertionError):
                self._check_zeros(ctf, no_cs, n)

    @pytest.mark.parametrize("dz,apix", [(30000.0, 6.8), (5000.0, 1.5), (30000.0, 3.0)])
    def test_zeros_reference_voxel_units(self, dz, apix):
        # the reference expresses lambda and Cs per voxel (divided by the
        # sampling rate) and takes the defocus as given in A
        # (tme/filters/ctf.py:596-645), so its defocus term is that of a
        # defocus dz * apix at 1 A/voxel; the zeros follow that definition
        n = 512
        f = CTFReconstructed(defocus_x=dz, sampling_rate=apix, acceleration_voltage=300e3,
                             spherical_aberration=2.7e7, amplitude_contrast=0.07, flip_phase=False)
        ctf = np.fft.fftshift(f(shape=(n,))["data"])
        lam = electron_wavelength(300e3) / apix
        self._check_zeros(ctf, self._zeros(400, dz, 2.7e7 / apix, lam, 0.07), n)

    def test_radial_dc_flip_and_cutoff(self):
        shape = (40, 48, 56)
        args = dict(defocus_x=30000.0, sampling_rate=6.8, spherical_aberration=2.7e7, amplitude_contrast=0.07)
        raw = CTFReconstructed(flip_phase=False, **args)(shape=shape)["data"]
        flipped = CTFReconstructed(flip_phase=True, **args)(shape=shape)["data"]
        assert np.array_equal(flipped, np.abs(raw))
        # DC: -sin(-phi) = sin(atan(A / sqrt(1 - A^2))) = A
        assert abs(raw[0, 0, 0] - 0.07) < 1e-6
        c = np.fft.fftshift(raw)
        # no astigmatism: the value depends on |k| only (10/40, 12/48 and
        # 14/56 cycles/voxel are all |k| = 0.25)
        k = [np.fft.fftshift(np.fft.fftfreq(n)) for n in shape]
        assert abs(c[shape[0] // 2 + 10, shape[1] // 2, shape[2] // 2]
                   - c[shape[0] // 2, shape[1] // 2 + 12, shape[2] // 2]) < 1e-6
        assert abs(c[shape[0] // 2, shape[1] // 2 + 12, shape[2] // 2]
                   - c[shape[0] // 2, shape[1] // 2, shape[2] // 2 + 14]) < 1e-6
        # zero at and beyond the cutoff |k| >= 0.5 (corners), non-zero inside
        radius = np.sqrt(sum(g * g for g in np.meshgrid(*k, indexing="ij")))
        assert np.all(c[radius >= 0.5] == 0)
