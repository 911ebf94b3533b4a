"""Contrast transfer function weights (``tme/filters/ctf.py:254-335, 508-655``).

chi(k) = pi*lambda*k^2*(dz - Cs*lambda^2*k^2/2) + phase + atan(A/sqrt(1-A^2))
ctf(k) = sin(-chi(k)) for |k| < cutoff, else 0; returned negated and, with
``flip_phase``, as |ctf|.  k is in cycles/voxel; lambda (relativistic
electron wavelength) and Cs are expressed in voxels.
"""

from dataclasses import dataclass
from typing import Dict, Tuple

import numpy as np

from ._utils import crop_real_fourier, fftfreqn, shift_fourier

__all__ = ["CTFReconstructed", "create_ctf", "electron_wavelength"]


def electron_wavelength(acceleration_voltage: float) -> float:
    """Relativistic electron wavelength in Angstrom for a voltage in volts."""
    h = 6.62606896e-34
    e = 1.60217646e-19
    m = 9.10938215e-31
    c = 299792458
    energy = e * acceleration_voltage
    return float(h * c / np.sqrt(energy**2 + 2 * energy * m * c**2) * 1e10)


def create_ctf(
    shape: Tuple[int],
    defocus_x: float,
    acceleration_voltage: float = 300e3,
    phase_shift: float = 0,
    sampling_rate: float = 1,
    spherical_aberration: float = 2.7e7,
    amplitude_contrast: float = 0.07,
    cutoff_frequency: float = 0.5,
    **kwargs,
) -> np.ndarray:
    """Centred CTF (DC at ``shape // 2``) without astigmatism or tilt."""
    if kwargs.get("defocus_y") is not None or kwargs.get("correct_defocus_gradient"):
        raise NotImplementedError("astigmatic / gradient-corrected CTFs are not part of this build")
    lam = electron_wavelength(acceleration_voltage) / sampling_rate
    aberration = (spherical_aberration / sampling_rate) * lam**2
    k = fftfreqn(shape, sampling_rate=True, compute_euclidean_norm=True)
    inside = k < cutoff_frequency
    k = np.square(k, out=k)
    chi = defocus_x - 0.5 * aberration * k
    chi = np.multiply(chi, np.pi * lam, out=chi)
    chi = np.multiply(chi, k, out=chi)
    chi += phase_shift
    chi += np.arctan(np.divide(amplitude_contrast, np.sqrt(1 - np.square(amplitude_contrast))))
    chi = np.sin(-chi, out=chi)
    return np.multiply(chi, inside, out=chi)


@dataclass
class CTFReconstructed:
    """CTF weight of a reconstructed volume (one defocus for all tilts)."""

    shape: Tuple[int] = None
    defocus_x: float = None
    angles: Tuple[float] = None
    opening_axis: int = 2
    tilt_axis: int = 0
    sampling_rate: float = 1
    acceleration_voltage: float = 300e3
    spherical_aberration: float = 2.7e7
    amplitude_contrast: float = 0.07
    phase_shift: float = 0
    defocus_angle: float = 0
    defocus_y: float = None
    flip_phase: bool = True
    return_real_fourier: bool = False

    def __post_init__(self):
        self.defocus_angle = np.radians(self.defocus_angle)
        self.phase_shift = np.radians(self.phase_shift)

    def __call__(self, **kwargs) -> Dict:
        args = dict(vars(self))
        args.update(kwargs)
        ret = self.weight(**args)
        return {
            "data": np.ascontiguousarray(ret, dtype=np.float32),
            "shape": args["shape"],
            "return_real_fourier": args.get("return_real_fourier"),
            "is_multiplicative_filter": True,
        }

    def weight(
        self,
        shape,
        defocus_x,
        amplitude_contrast=0.07,
        phase_shift=0,
        sampling_rate=1,
        spherical_aberration=2.7e7,
        flip_phase=True,
        return_real_fourier=False,
        cutoff_frequency=0.5,
        defocus_y=None,
        **kwargs,
    ) -> np.ndarray:
        stack = create_ctf(
            shape=shape,
            defocus_x=defocus_x,
            defocus_y=defocus_y,
            sampling_rate=np.max(sampling_rate),
            acceleration_voltage=self.acceleration_voltage,
            spherical_aberration=spherical_aberration,
            cutoff_frequency=cutoff_frequency,
            phase_shift=phase_shift,
            amplitude_contrast=amplitude_contrast,
        )
        stack = shift_fourier(stack)
        np.negative(stack, out=stack)
        if flip_phase:
            np.abs(stack, out=stack)
        stack = np.squeeze(stack)
        if return_real_fourier:
            stack = crop_real_fourier(stack)
        return stack
