"""End-to-end scan parity: pytme_amd hip backend vs the FP64 oracle."""

import numpy as np
import pytest
from scipy.ndimage import laplace

from oracle import scan as orc
from tests._parity import assert_scan_parity

pytestmark = pytest.mark.gpu


def _run(score, target, template, mask, rotations, fused, threshold=0.0, template_filter=None,
         target_filter=None, target_splits=None, pad=False, **backend_kwargs):
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.backends import backend as be
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, scan_subsets

    be.change_backend("hip", device="cuda", fused=fused, **backend_kwargs)
    md = MatchingData(target.copy(), template.copy(), template_mask=None if mask is None else mask.copy(),
                      rotations=rotations)
    if template_filter is not None:
        md.template_filter = template_filter
    if target_filter is not None:
        md.target_filter = target_filter
    setup, scoring = MATCHING_EXHAUSTIVE_REGISTER[score]
    return scan_subsets(
        matching_data=md, matching_setup=setup, matching_score=scoring,
        callback_class=MaxScoreOverRotations, callback_class_args={"score_threshold": threshold},
        target_splits=target_splits or {}, pad_target_edges=pad, job_schedule=(1, 1),
        interpolation_order=1, pad_template_filter=False,
    )


class TestReferenceIntegrationCase:
    """tests/test_matching_exhaustive.py:16-97 of the reference, on the GPU."""

    @pytest.mark.parametrize("score", ["CC", "LCC", "FLCSphericalMask", "FLC"])
    @pytest.mark.parametrize("pad_edge", [False, True])
    @pytest.mark.parametrize("fused", [False, True])
    def test_peak(self, score, pad_edge, fused):
        target = np.zeros((80, 80, 80), np.float32)
        target[25:31, 22:28, 12:16] = 1
        template = np.zeros((41, 41, 35), np.float32)
        template[20:26, 25:31, 17:21] = 1
        mask = np.ones_like(template)
        rot = np.eye(3, dtype=np.float32)[None]
        print(f"integration case {score} pad={pad_edge} fused={fused}", flush=True)
        s, off, r, mapping = _run(score, target, template, mask, rot, fused, target_splits={0: 2}, pad=pad_edge)
        ref_s, ref_r = orc.scan_subsets(score, target, template, rot, template_mask=mask, target_splits={0: 2},
                                        pad_target_edges=pad_edge)
        peak = np.unravel_index(np.argmax(s), s.shape)
        assert np.allclose(peak, [25, 17, 12])
        expected = 1.0
        if score == "CC":
            expected = template.sum()
        elif score == "LCC":
            expected = (laplace(template) * laplace(template)).sum()
        assert np.allclose(s[peak], expected, rtol=0.3)
        scale = max(1.0, float(np.abs(ref_s).max()))
        assert abs(s[peak] - ref_s[peak]) <= 1e-4 * scale
        # the box target has zero-variance windows where FP32 cancellation in
        # the local variance dominates (as in the reference's float32 numpy
        # backend); the 1e-4 bound is checked on the noise workloads below
        assert np.abs(s - ref_s).max() <= 5e-3 * scale
        assert 0 in mapping and np.allclose(mapping[0], np.eye(3))


def _workload(seed, shape=(40, 36, 30), T=12, n_rot=12, copies=3):
    from pytme_amd.rotations import uniform_rotations
    from pytme_amd.synthetic import blob_template, embed_copies, threshold_mask

    template = blob_template(T, n_blobs=8, seed=seed)
    mask = threshold_mask(template, dilation=1)
    target, _, _ = embed_copies(shape, template, mask, copies, 0.5, seed)
    rots = uniform_rotations(n_rot, seed=seed + 3, include_identity=True)
    return target, template, mask, rots


@pytest.mark.parametrize("score", ["FLC", "FLCSphericalMask", "CC", "LCC"])
@pytest.mark.parametrize("fused", [False, True])
def test_random_workload_parity(score, fused):
    target, template, mask, rots = _workload(seed=11)
    print(f"random workload {score} fused={fused}", flush=True)
    threshold = -1.0 if score in ("FLC", "FLCSphericalMask") else -1e30
    s, off, r, mapping = _run(score, target, template, mask, rots, fused, threshold=threshold)
    _, _, raws, oc = orc.scan(score, target, template, rots, template_mask=mask, score_threshold=threshold,
                              return_raw=True)
    tol = 1e-4 if score in ("FLC", "FLCSphericalMask") else 1e-4 * max(1.0, max(np.abs(x).max() for x in raws))
    n_diff, err = assert_scan_parity(s, r, raws, oc, threshold, tol=tol, tie_tol=max(2e-5, tol / 5))
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)
    assert len(mapping) == rots.shape[0]


@pytest.mark.parametrize("score", ["FLC", "FLCSphericalMask"])
@pytest.mark.parametrize("fused", [False, True])
def test_filtered_parity(score, fused):
    """Wedge x CTF template filter at template size plus a continuous-wedge
    target filter (the cryo-ET configuration of scripts/match_template.py)."""
    from pytme_amd.filters import Compose, CTFReconstructed, WedgeReconstructed
    from pytme_amd.matching_data import MatchingData

    target, template, mask, rots = _workload(seed=21, T=16)
    if score == "FLCSphericalMask":
        from pytme_amd.synthetic import spherical_mask

        mask = spherical_mask(16, 7)
    wedge = WedgeReconstructed(angles=np.arange(-60, 61, 3.0), opening_axis=2, tilt_axis=0)
    ctf = CTFReconstructed(defocus_x=30000, sampling_rate=6.8, acceleration_voltage=300e3,
                           spherical_aberration=2.7e7, amplitude_contrast=0.07, flip_phase=True)
    comp = Compose((wedge, ctf))
    tcomp = Compose((WedgeReconstructed(angles=(-60, 60), create_continuous_wedge=True, opening_axis=2,
                                        tilt_axis=0),))
    print(f"filtered {score} fused={fused}", flush=True)
    s, off, r, _ = _run(score, target, template, mask, rots, fused, threshold=-1.0, template_filter=comp,
                        target_filter=tcomp)
    conv, fast, ft, shift = MatchingData._fourier_padding(target.shape, template.shape)
    tfilt = comp(shape=template.shape, return_real_fourier=True)["data"]
    gfilt = tcomp(shape=fast, return_real_fourier=True)["data"]
    _, _, raws, oc = orc.scan(score, target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, template_filter=tfilt, target_filter=gfilt)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


@pytest.mark.parametrize("shape", [(128, 128, 200), (128, 128, 128), (128, 64, 256), (128, 256, 200), (256, 128, 256)])
@pytest.mark.parametrize("score", ["FLC", "FLCSphericalMask"])
@pytest.mark.parametrize("fused", [False, True])
def test_specialised_sizes_parity(shape, score, fused):
    """Tomogram sizes that take the compile-time FFT kernels (128/256-point
    axes, half-length 200/128-point C2R) against the FP64 oracle; x lengths
    200 and 256 run the fused scan on split rows (96 + 5 and 128 + 1 bins)
    where the y length has a compile-time kernel (128 and 256 here: pass B
    ends with the x-inverse pre-processing, for x length 200 in radix-4
    units also with the inverse's first stage; mirror pairs for x length
    256), padded rows otherwise (y length 64)."""
    target, template, mask, rots = _workload(seed=31, shape=shape, T=16, n_rot=4, copies=4)
    if score == "FLCSphericalMask":
        from pytme_amd.synthetic import spherical_mask

        mask = spherical_mask(16, 7)
    print(f"specialised {shape} {score} fused={fused}", flush=True)
    s, off, r, _ = _run(score, target, template, mask, rots, fused, threshold=-1.0)
    _, _, raws, oc = orc.scan(score, target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


@pytest.mark.parametrize("score,batch", [("FLC", 1), ("FLC", 2), ("FLC", 3), ("FLCSphericalMask", 1),
                                         ("FLCSphericalMask", 3)])
def test_deferred_store_batches_vs_oracle(score, batch):
    """A 512-point z axis with 16-column tiles runs pass A's kernels that
    hold each output in registers and store it during the next transforms
    (the first rotation of a block stores nothing held, the last output is
    stored after the loop): batches of 1, 2 and 3 rotations over 7 rotations
    (partial last batches; pass A's tail launch splits fewer rotations)
    against the FP64 oracle, for the S1/S2 and X kernels (FLC) and the X
    kernel alone (FLCSphericalMask), on split rows in radix-4 units."""
    target, template, mask, rots = _workload(seed=37, shape=(512, 128, 200), T=16, n_rot=7, copies=3)
    if score == "FLCSphericalMask":
        from pytme_amd.synthetic import spherical_mask

        mask = spherical_mask(16, 7)
    print(f"deferred stores, {score} batch {batch}", flush=True)
    s, off, r, _ = _run(score, target, template, mask, rots, True, threshold=-1.0, rotation_batch=batch)
    _, _, raws, oc = orc.scan(score, target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, workers=8)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_fused_matches_plain_full_size():
    """Size-independent property at the benchmark size: the fused batched scan
    and the per-rotation loop give the same maxima and rotation indices on a
    512x512x200 tomogram with the wedge x CTF filters."""
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, filters, spec = make_workload("flc_512x512x200_t48", copies=20, n_rotations=6)
    outs = []
    for fused in (False, True):
        print(f"full size fused={fused}", flush=True)
        s, off, r, mapping = _run("FLC", target, template, mask, rotations, fused, threshold=0.0,
                                  template_filter=filters[0], target_filter=filters[1])
        outs.append((s, r))
    (s0, r0), (s1, r1) = outs
    assert np.abs(s0 - s1).max() < 1e-5
    agree = (r0 == r1).mean()
    assert agree > 0.9999, agree


def test_c1_full_size_vs_fp64_oracle():
    """The benchmark configuration itself against the FP64 oracle: the
    512x512x200 tomogram, 48^3 template, wedge x CTF template filter and
    continuous-wedge target filter, 3 rotations through the fused path;
    scores within 1e-4 and identical rotation indices except at ties."""
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, filters, spec = make_workload("flc_512x512x200_t48", n_rotations=3)
    print("C1 full size, 3 rotations, fused vs FP64 oracle", flush=True)
    s, off, r, _ = _run("FLC", target, template, mask, rotations, True, threshold=-1.0,
                        template_filter=filters[0], target_filter=filters[1])
    conv, fast, ft, shift = MatchingData._fourier_padding(target.shape, template.shape)
    tfilt = filters[0](shape=template.shape, return_real_fourier=True)["data"]
    gfilt = filters[1](shape=fast, return_real_fourier=True)["data"]
    print("  oracle ...", flush=True)
    _, _, raws, oc = orc.scan("FLC", target, template, rotations, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, template_filter=tfilt, target_filter=gfilt)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


@pytest.mark.parametrize("writer", ["torch", "library"])
@pytest.mark.parametrize("score", ["FLC", "FLCSphericalMask"])
def test_kept_plan_sees_new_target_values(score, writer):
    """The fused plan is kept across calls with the same arrays; it holds a
    row-padded copy of the target spectra, so new values written into the
    caller's arrays at the same addresses must reach the next call, whether
    torch wrote them (version counters) or a library kernel did (the
    backend's spectral multiply by ones)."""
    import torch

    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.analyzer.proxy import SharedAnalyzerProxy
    from pytme_amd.backends import backend as be
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan

    be.change_backend("hip", device="cuda", fused=True)
    target_a, template, mask, rots = _workload(seed=41, n_rot=6)
    target_b, _, _, _ = _workload(seed=42, n_rot=6)
    if score == "FLCSphericalMask":
        from pytme_amd.synthetic import spherical_mask

        mask = spherical_mask(12, 5)
    setup_fn, scoring = MATCHING_EXHAUSTIVE_REGISTER[score]

    def prepare(target):
        md = MatchingData(target.copy(), template.copy(), template_mask=mask.copy(), rotations=rots)
        setup, cb_args, chunks = prepare_scan(md, setup_fn, n_jobs=1, callback_class_args={"score_threshold": -1.0},
                                              pad_target=False, pad_template_filter=False, interpolation_order=1)
        return setup, cb_args, chunks[0]

    def run(setup, cb_args, chunk):
        proxy = SharedAnalyzerProxy(MaxScoreOverRotations, cb_args)
        scoring(rotations=chunk, callback=proxy, **setup)
        s, _, r, _ = proxy.result(**cb_args)
        return be.to_numpy_array(s).copy(), be.to_numpy_array(r).copy()

    setup_a, cb_a, chunk_a = prepare(target_a)
    setup_b, cb_b, chunk_b = prepare(target_b)
    s_b, r_b = run(setup_b, cb_b, chunk_b)
    run(setup_a, cb_a, chunk_a)  # creates (or re-keys) the plan on target a's arrays
    copied = 0
    for key in setup_a:
        if key.startswith("ft_target") or key == "inv_denominator":
            dst, src = be.from_sharedarr(setup_a[key]), be.from_sharedarr(setup_b[key])
            assert torch.is_tensor(dst) and dst.shape == src.shape
            if writer == "torch" or not torch.is_complex(dst):
                dst.copy_(src)
            else:
                version = dst._version
                be.multiply(src, torch.ones_like(src), out=dst)
                assert dst._version == version  # invisible to torch
            copied += 1
    assert copied == 2
    torch.cuda.synchronize()
    from pytme_amd.backends import hip_fused

    handle = hip_fused._PLAN["handle"]
    s_a2, r_a2 = run(setup_a, cb_a, chunk_a)  # same arrays and plan, target b's values
    assert hip_fused._PLAN["handle"] == handle
    assert np.abs(s_a2 - s_b).max() < 1e-6
    assert (r_a2 == r_b).mean() > 0.9999


def test_c0_flcsph_parity():
    """FLCSphericalMask on the 128^3 / 32^3 workload (spherical mask radius
    14, 5 embedded copies at SNR 0.5): fused and plain paths against the FP64
    oracle on a rotation subset."""
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, _, spec = make_workload("flcsph_128_t32", n_rotations=6)
    _, _, raws, oc = orc.scan("FLCSphericalMask", target, template, rotations, template_mask=mask,
                              score_threshold=0.0, return_raw=True)
    for fused in (False, True):
        print(f"C0 FLCSphericalMask fused={fused}", flush=True)
        s, off, r, _ = _run("FLCSphericalMask", target, template, mask, rotations, fused, threshold=0.0)
        n_diff, err = assert_scan_parity(s, r, raws, oc, 0.0)
        print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_c4_fused_matches_plain_full_size():
    """Size-independent property at 1024x1024x256 with the 64^3 template and
    the wedge x CTF filters (20 embedded copies): the fused batched scan and
    the per-rotation loop agree on every maximum and rotation index."""
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, filters, spec = make_workload("flc_1024x1024x256_t64", copies=20,
                                                                     n_rotations=4)
    outs = []
    for fused in (False, True):
        print(f"C4 full size fused={fused}", flush=True)
        s, off, r, mapping = _run("FLC", target, template, mask, rotations, fused, threshold=0.0,
                                  template_filter=filters[0], target_filter=filters[1])
        outs.append((s, r))
    (s0, r0), (s1, r1) = outs
    assert s0.shape == (1024, 1024, 256)
    assert np.abs(s0 - s1).max() < 1e-5
    assert (r0 == r1).mean() > 0.9999


def _vs_oracle(target, entry, result, threshold, workers=1):
    """One template's scan result against the FP64 oracle (template filter
    at template size, target filter at the fast shape)."""
    from pytme_amd.matching_data import MatchingData

    template = entry["template"]
    tfilt = gfilt = None
    if entry.get("template_filter") is not None:
        tfilt = entry["template_filter"](shape=template.shape, return_real_fourier=True)["data"]
    if entry.get("target_filter") is not None:
        fast = MatchingData._fourier_padding(target.shape, template.shape)[1]
        gfilt = entry["target_filter"](shape=fast, return_real_fourier=True)["data"]
    _, _, raws, oc = orc.scan("FLC", target, template, entry["rotations"], template_mask=entry["template_mask"],
                              score_threshold=threshold, return_raw=True, template_filter=tfilt,
                              target_filter=gfilt, workers=workers)
    n_diff, err = assert_scan_parity(result[0], result[2], raws, oc, threshold)
    print(f"  template {template.shape[0]}^3: max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_c2_full_size_vs_fp64_oracle():
    """The multi-template workload at its own size: one 512x512x200
    tomogram with 50 embedded copies of each of four templates (24^3, 32^3,
    48^3, 64^3 blob densities, own thresholded masks, wedge x CTF at 2.0 /
    2.5 / 3.0 / 3.5 um defocus, continuous-wedge target filter), all
    templates' batches in one replayed hipGraph schedule, 2 rotations per
    template; every template's scores within 1e-4 of the FP64 oracle and
    identical rotation indices except at ties."""
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.backends import backend as be
    from pytme_amd.matching_exhaustive import scan_templates
    from pytme_amd.synthetic import make_multi_workload

    target, entries, spec = make_multi_workload(n_rotations=2)
    be.change_backend("hip", device="cuda", fused=True)
    print("C2 full size, 4 templates x 2 rotations, one-graph schedule", flush=True)
    kw = dict(callback_class=MaxScoreOverRotations, callback_class_args={"score_threshold": -1.0},
              interpolation_order=1, pad_template_filter=False)
    fields = ("template", "template_mask", "template_filter", "target_filter", "rotations")
    res = scan_templates(target, [{k: e[k] for k in fields} for e in entries], score="FLC", schedule="graph", **kw)
    for entry, result in zip(entries, res):
        assert result[0].shape == target.shape
        _vs_oracle(target, entry, result, -1.0, workers=8)


@pytest.mark.parametrize("shape", [(1024, 128, 256), (128, 1024, 256), (1024, 48, 256), (128, 1024, 200)])
@pytest.mark.parametrize("fused", [False, True])
def test_1024_point_axes_vs_oracle(shape, fused):
    """The 1024-point kernel instantiations of the large-tomogram scan
    against the FP64 oracle: pass A's z transform at z = 1024 (1024, 128,
    256) and pass B's y transform at y = 1024 (128, 1024, 256, mirror
    pairs), both with the split rows of the 256-point x axis, a y length
    without a compile-time kernel (1024, 48, 256: padded rows), and the
    200-point x axis with y = 1024 (128, 1024, 200: mirror pairs there, the
    radix-4 units being for y lengths up to 512); 32^3 template with a
    threshold mask, 2 rotations."""
    from pytme_amd.synthetic import blob_template, embed_copies, threshold_mask
    from pytme_amd.rotations import uniform_rotations

    template = blob_template(32, n_blobs=12, seed=91)
    mask = threshold_mask(template)
    target, _, _ = embed_copies(shape, template, mask, 3, 0.5, seed=92)
    rots = uniform_rotations(2, seed=93, include_identity=True)
    print(f"1024-point axes {shape} fused={fused}", flush=True)
    s, off, r, _ = _run("FLC", target, template, mask, rots, fused, threshold=-1.0)
    _, _, raws, oc = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, workers=8)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_scan_releases_the_fused_plan():
    """A scan's batch workspace is released when the scan ends (the
    backend's free_cache), so the next scan's batch choice and other HBM
    users see the same free memory as before."""
    import torch

    from pytme_amd.backends import hip_fused

    target, template, mask, rots = _workload(seed=61, n_rot=6)
    _run("FLC", target, template, mask, rots, True, threshold=-1.0)  # settles torch's caches
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    s, off, r, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0)
    torch.cuda.synchronize()
    assert hip_fused._PLAN["handle"] is None
    assert abs(torch.cuda.mem_get_info()[0] - free0) < (256 << 20)


def test_multi_template_schedule():
    """Four templates of uneven size (24^3, 32^3, 48^3, 64^3 blob densities),
    each with its own mask and CTF defocus, against one tomogram: every
    result equals the template's own single-template scan, and each
    template's strongest peaks sit on its embedded copies."""
    from pytme_amd.filters import Compose, CTFReconstructed
    from pytme_amd.matching_exhaustive import scan_templates
    from pytme_amd.rotations import uniform_rotations
    from pytme_amd.synthetic import blob_template, threshold_mask

    rng = np.random.default_rng(41)
    shape = (160, 160, 100)
    target = rng.standard_normal(shape).astype(np.float32)
    entries, centres = [], []
    for k, (T, dz) in enumerate(((24, 20000.0), (32, 25000.0), (48, 30000.0), (64, 35000.0))):
        tmpl = blob_template(T, n_blobs=10, seed=50 + k)
        m = threshold_mask(tmpl)
        cs = []
        for c in range(2):
            lo = np.array([T // 2] * 3)
            hi = np.asarray(shape) - (T - T // 2)
            ctr = np.array([rng.integers(lo[i], hi[i] + 1) for i in range(3)])
            box = tuple(slice(int(x - T // 2), int(x - T // 2) + T) for x in ctr)
            target[box] += 2.0 * tmpl
            cs.append(ctr)
        centres.append(cs)
        ctf = CTFReconstructed(defocus_x=dz, sampling_rate=6.8, acceleration_voltage=300e3,
                               spherical_aberration=2.7e7, amplitude_contrast=0.07, flip_phase=True)
        entries.append(dict(template=tmpl, template_mask=m, template_filter=Compose((ctf,)),
                            rotations=uniform_rotations(8, seed=60 + k, include_identity=True)))
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.backends import backend as be

    be.change_backend("hip", device="cuda", fused=True)
    print("multi-template schedule, 4 templates", flush=True)
    kw = dict(callback_class=MaxScoreOverRotations, callback_class_args={"score_threshold": 0.0},
              job_schedule=(1, 1), interpolation_order=1, pad_template_filter=False)
    multi = scan_templates(target, entries, score="FLC", **kw)
    for k, entry in enumerate(entries):
        single = _run("FLC", target, entry["template"], entry["template_mask"], entry["rotations"], True,
                      threshold=0.0, template_filter=entry["template_filter"])
        assert np.array_equal(multi[k][0], single[0]) and np.array_equal(multi[k][2], single[2])
        _vs_oracle(target, entry, multi[k], 0.0)
        s = multi[k][0]
        for ctr in centres[k]:
            win = tuple(slice(int(x) - 2, int(x) + 3) for x in ctr)
            assert s[win].max() >= np.sort(s.ravel())[-200], (k, ctr)


@pytest.mark.parametrize("counts", [(28, 28, 28, 28), (28, 12, 20, 5)])
def test_multi_template_one_graph_schedule(counts):
    """scan_templates(schedule="graph"): four templates of uneven size whose
    batches of a round run as one replayed hipGraph equal the sequential
    schedule bit for bit.  Batches of 8: with 28 rotations each the rounds
    run eagerly, capture, replay and end with a partial batch; with 28 / 12
    / 20 / 5 rotations templates drop out of later rounds (batch 0)."""
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.backends import backend as be
    from pytme_amd.filters import Compose, CTFReconstructed
    from pytme_amd.matching_exhaustive import scan_templates
    from pytme_amd.rotations import uniform_rotations
    from pytme_amd.synthetic import blob_template, threshold_mask

    rng = np.random.default_rng(43)
    target = rng.standard_normal((96, 80, 72)).astype(np.float32)
    entries = []
    for k, (T, dz) in enumerate(((12, 20000.0), (16, 25000.0), (20, 30000.0), (24, 35000.0))):
        tmpl = blob_template(T, n_blobs=8, seed=70 + k)
        ctf = CTFReconstructed(defocus_x=dz, sampling_rate=6.8, acceleration_voltage=300e3,
                               spherical_aberration=2.7e7, amplitude_contrast=0.07, flip_phase=True)
        entries.append(dict(template=tmpl, template_mask=threshold_mask(tmpl), template_filter=Compose((ctf,)),
                            rotations=uniform_rotations(counts[k], seed=80 + k, include_identity=True)))
    kw = dict(callback_class=MaxScoreOverRotations, callback_class_args={"score_threshold": -1.0},
              interpolation_order=1, pad_template_filter=False)
    be.change_backend("hip", device="cuda", fused=True, rotation_batch=8)
    print(f"one-graph multi-template schedule, rotations {counts}", flush=True)
    graph = scan_templates(target, entries, score="FLC", schedule="graph", **kw)
    seq = scan_templates(target, entries, score="FLC", job_schedule=(1, 1), **kw)
    for k in range(len(entries)):
        assert np.array_equal(graph[k][0], seq[k][0]), k
        assert np.array_equal(graph[k][2], seq[k][2]), k
        assert np.array_equal(graph[k][1], seq[k][1]), k
        assert sorted(graph[k][3]) == sorted(seq[k][3]), k
        _vs_oracle(target, entries[k], graph[k], -1.0)
    with pytest.raises(ValueError):
        scan_templates(target, entries[:1], score="FLC", schedule="graph", target_splits={0: 2}, **kw)


def test_fused_partial_batches_and_plan_reuse():
    """20 rotations in batches of 8 (8, 8, 4): the kept plan, the partial
    last batch and the rotation index assignment against the FP64 oracle."""
    target, template, mask, rots = _workload(seed=51, n_rot=20)
    print("fused partial batches 8+8+4", flush=True)
    s, off, r, mapping = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch=8)
    _, _, raws, oc = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)
    assert len(mapping) == 20


def test_autotuned_batch_matches_fixed_batches():
    """rotation_batch="auto" times the fitting batch sizes on scratch state
    and keeps the fastest; the scan result is bit-identical to fixed batch
    sizes 5 and 24 (per-rotation arithmetic does not depend on the batch) and
    to the FP64 oracle within the parity bound."""
    from pytme_amd.backends import hip_fused

    target, template, mask, rots = _workload(seed=71, n_rot=48)
    print("autotuned vs fixed batches, 48 rotations", flush=True)
    hip_fused.TUNED.clear()
    s_auto, _, r_auto, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch="auto")
    tuned = next(iter(hip_fused.TUNED.values()))
    print(f"  tuned {tuned}", flush=True)
    assert tuned["batch"] in tuned["ms_per_rotation"] and len(tuned["ms_per_rotation"]) >= 2
    assert len(tuned["tiles_ms_per_rotation"]) >= 1 and len(tuned["tiles"]) == 2
    for b in (5, 24):
        s_b, _, r_b, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch=b)
        np.testing.assert_array_equal(s_b, s_auto)
        np.testing.assert_array_equal(r_b, r_auto)
    _, _, raws, oc = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True)
    n_diff, err = assert_scan_parity(s_auto, r_auto, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


@pytest.mark.parametrize("tiles", [(8, 0), (0, 10), (8, 10)])
def test_tile_shapes_give_identical_results(tiles):
    """The tile shapes the autotuner chooses between (pass A 16- or 8-column
    z tiles, pass C 5- or 10-row blocks) change how lines are spread over
    threads, not the arithmetic: the scan state is bit-identical to the
    default tiles on a 128x128x200 target (both pass A and the 200-point
    pass C on split rows in radix-4 units have both variants there)."""
    target, template, mask, rots = _workload(seed=81, shape=(128, 128, 200), n_rot=8)
    print(f"tile shapes {tiles} vs default", flush=True)
    s0, _, r0, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch=4)
    s1, _, r1, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch=4,
                        scan_tiles=tiles)
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(r1, r0)


def test_graph_replays_match_eager_batches():
    """The fused batch replayed as a hipGraph: every replay (batch 3 onwards;
    batch 1 runs eagerly, batch 2 is captured) equals the eager batch, with
    eager torch work and host synchronisation between the replays."""
    import ctypes

    import torch

    from pytme_amd import _lib
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.analyzer.proxy import SharedAnalyzerProxy
    from pytme_amd.backends import backend as be
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan

    be.change_backend("hip", device="cuda", fused=True)
    target, template, mask, rots = _workload(seed=61, shape=(48, 44, 40), n_rot=40)
    setup_fn, _ = MATCHING_EXHAUSTIVE_REGISTER["FLC"]
    md = MatchingData(target, template, template_mask=mask, rotations=rots)
    setup, cb_args, chunks = prepare_scan(md, setup_fn, n_jobs=1, callback_class_args={"score_threshold": -1.0},
                                          pad_target=False, pad_template_filter=False, interpolation_order=1)
    f = lambda k: be.from_sharedarr(setup[k])  # noqa: E731
    templ = f("template").float().contiguous()
    msk = f("template_mask").float().contiguous()
    ft, ft2 = f("ft_target").contiguous(), f("ft_target2").contiguous()
    fast = tuple(int(x) for x in setup["fast_shape"])
    B = 8
    plans = []
    for graph in (0, 1):
        h = ctypes.c_void_p()
        _lib.call("tm_scan_plan_create", 0, _lib.i64_array(fast), _lib.i32_array(tuple(templ.shape)),
                  templ.data_ptr(), msk.data_ptr(), None, ft.data_ptr(), ft2.data_ptr(), None, 0, 0.0, B,
                  ctypes.byref(h))
        _lib.call("tm_scan_plan_use_graph", h.value, graph)
        plans.append(h.value)
    r = chunks[0].float().contiguous()
    idx = torch.arange(r.shape[0], dtype=torch.int32, device="cuda")
    ms = torch.empty(fast, device="cuda")
    mr = torch.empty(fast, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    try:
        for k in range(r.shape[0] // B):
            outs = []
            for plan in plans:
                ms.fill_(-1.0)
                mr.fill_(-1)
                _lib.call("tm_scan_batch", plan, r[k * B].data_ptr(), idx[k * B].data_ptr(), B, ms.data_ptr(),
                          mr.data_ptr(), stream)
                torch.cuda.synchronize()
                outs.append((ms.cpu().numpy().copy(), mr.cpu().numpy().copy()))
            print(f"batch {k}: graph vs eager", flush=True)
            assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1]), k
    finally:
        for plan in plans:
            _lib.call("tm_scan_plan_destroy", plan)


class _ReferenceShapedProxy:
    """The attribute layout of pyTME's SharedAnalyzerProxy
    (tme/analyzer/proxy.py:72-123: _analyzer, _state, _shared) around an
    analyzer named MaxScoreOverRotations."""

    def __init__(self, analyzer_class, params):
        self._analyzer = analyzer_class(**params)
        self._shared = False
        self._state = self._analyzer.init_state()

    def __call__(self, *args, **kwargs):
        self._state = self._analyzer(self._state, *args, **kwargs)

    def result(self, **kwargs):
        return self._analyzer.result(self._state, **kwargs)


def test_installed_pytme_reaches_the_fused_scan(monkeypatch):
    """integration.install on a pyTME-shaped module tree: its registered
    flc_scoring (a stand-in that fails if its loop runs) called the way
    pyTME's scan calls it, with a pyTME-shaped proxy, runs tm_scan_batch and
    matches the FP64 oracle."""
    from pytme_amd import _lib
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.backends import backend as be
    from pytme_amd.integration import install
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan
    from tests.test_host import _stand_in_pytme

    tme, calls = _stand_in_pytme()

    def flc_scoring(**kw):
        raise AssertionError("the reference's per-rotation loop ran")

    tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER["FLC"] = ("flc_setup", flc_scoring)
    install(tme)
    tme.backends.backend.change_backend("hip", device="cuda")
    be.change_backend("hip", device="cuda")
    target, template, mask, rots = _workload(seed=71, n_rot=10)
    md = MatchingData(target.copy(), template.copy(), template_mask=mask.copy(), rotations=rots)
    setup, cb_args, chunks = prepare_scan(md, MATCHING_EXHAUSTIVE_REGISTER["FLC"][0], n_jobs=1,
                                          callback_class_args={"score_threshold": -1.0}, pad_target=False,
                                          pad_template_filter=False, interpolation_order=1)
    proxy = _ReferenceShapedProxy(MaxScoreOverRotations, cb_args)
    seen = []
    real_call = _lib.call
    monkeypatch.setattr(_lib, "call", lambda name, *a: (seen.append(name), real_call(name, *a))[1])
    print("pyTME-shaped registry -> fused scan", flush=True)
    tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER["FLC"][1](rotations=chunks[0], callback=proxy, **setup)
    assert "tm_scan_batch" in seen
    s, off, r, mapping = MaxScoreOverRotations.merge([proxy.result(**cb_args)], **cb_args)
    _, _, raws, oc = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)
    be.free_cache()


@pytest.mark.parametrize("shape,T", [((256, 128, 200), 24), ((512, 256, 256), 32), ((128, 256, 128), 24),
                                     ((1024, 128, 256), 32)])
def test_radix_variant_vs_oracle(shape, T):
    """The fused scan with the second radix plan of the compile-time FFTs
    (tm_scan_plan_set_radix 1: 256 = 4*8*8, 512 = 4*4*4*8, 1024 = 4*4*8*8,
    128 = 4*4*8 for z / y, 100 = 5*5*4, 128 = 4*4*8, 64 = 4*4*4 for the x
    pass) against the FP64 oracle; the two plans round differently, so each
    is checked against the oracle rather than against the other."""
    from pytme_amd.rotations import uniform_rotations
    from pytme_amd.synthetic import blob_template, embed_copies, threshold_mask

    template = blob_template(T, n_blobs=10, seed=95)
    mask = threshold_mask(template)
    target, _, _ = embed_copies(shape, template, mask, 3, 0.5, seed=96)
    rots = uniform_rotations(3, seed=97, include_identity=True)
    _, _, raws, oc = orc.scan("FLC", target, template, rots, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, workers=8)
    for radix in (0, 1):
        print(f"radix plan {radix}, {shape} / {T}^3", flush=True)
        s, off, r, _ = _run("FLC", target, template, mask, rots, True, threshold=-1.0, scan_radix=radix)
        n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
        print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_c1_radix_variant_vs_fp64_oracle():
    """The benchmark configuration with the second radix plan (512 = 4*4*4*8
    for z and y, 100 = 5*5*4 for x) against the FP64 oracle, 2 rotations."""
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.synthetic import make_workload

    target, template, mask, rotations, filters, spec = make_workload("flc_512x512x200_t48", n_rotations=2)
    print("C1 full size, radix plan 1, fused vs FP64 oracle", flush=True)
    s, off, r, _ = _run("FLC", target, template, mask, rotations, True, threshold=-1.0, scan_radix=1,
                        template_filter=filters[0], target_filter=filters[1])
    conv, fast, ft, shift = MatchingData._fourier_padding(target.shape, template.shape)
    tfilt = filters[0](shape=template.shape, return_real_fourier=True)["data"]
    gfilt = filters[1](shape=fast, return_real_fourier=True)["data"]
    _, _, raws, oc = orc.scan("FLC", target, template, rotations, template_mask=mask, score_threshold=-1.0,
                              return_raw=True, template_filter=tfilt, target_filter=gfilt, workers=8)
    n_diff, err = assert_scan_parity(s, r, raws, oc, -1.0)
    print(f"  max err {err:.2e}, tie mismatches {n_diff}", flush=True)


def test_autotuner_reports_the_radix_plan():
    """rotation_batch="auto" times both radix plans at the chosen batch and
    tiles and records the choice with the plans' radices."""
    from pytme_amd.backends import hip_fused

    target, template, mask, rots = _workload(seed=81, shape=(128, 128, 100), T=16, n_rot=16)
    before = set(hip_fused.TUNED)
    _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch="auto")
    added = [k for k in hip_fused.TUNED if k not in before]
    assert added
    rec = hip_fused.TUNED[added[-1]]
    assert rec["radix"] in (0, 1) and set(rec["radix_ms_per_rotation"]) == {0, 1}
    assert rec["radix_plans"]["z"] in ("8x4x4", "4x4x8") and rec["radix_plans"]["y"] in ("8x4x4", "4x4x8")


def test_full_and_partial_batches_keep_replaying():
    """Scans whose rotation count is not a multiple of the batch (20 rotations
    in batches of 8: 8, 8, 4) called repeatedly on a kept plan and the same
    state arrays: after the first call's eager runs and captures, the full
    and the remainder batch both replay (tm_scan_plan_graph_stats), and every
    call equals the same scan on an eager plan."""
    import ctypes

    from pytme_amd import _lib
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.analyzer.proxy import SharedAnalyzerProxy
    from pytme_amd.backends import backend as be
    from pytme_amd.backends import hip_fused
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan

    target, template, mask, rots = _workload(seed=63, shape=(40, 36, 30), n_rot=20)
    eager = _run("FLC", target, template, mask, rots, True, threshold=-1.0, rotation_batch=8, scan_graph=False)
    be.change_backend("hip", device="cuda", fused=True, rotation_batch=8, scan_graph=True)
    setup_fn, scoring = MATCHING_EXHAUSTIVE_REGISTER["FLC"]
    md = MatchingData(target.copy(), template.copy(), template_mask=mask.copy(), rotations=rots)
    setup, cb_args, chunks = prepare_scan(md, setup_fn, n_jobs=1, callback_class_args={"score_threshold": -1.0},
                                          pad_target=False, pad_template_filter=False, interpolation_order=1)
    proxy = SharedAnalyzerProxy(MaxScoreOverRotations, cb_args)
    scores, rot_idx, mapping = proxy.state
    for call in range(3):
        print(f"partial batches, call {call}", flush=True)
        scores.fill_(-1.0)
        rot_idx.fill_(-1)
        mapping.clear()
        scoring(rotations=chunks[0], callback=proxy, **setup)
        s, off, r, _ = MaxScoreOverRotations.merge([proxy.result(**cb_args)], **cb_args)
        assert np.array_equal(s, eager[0]) and np.array_equal(r, eager[2])
    cap, rep, eag = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.call("tm_scan_plan_graph_stats", hip_fused._PLAN["handle"], ctypes.byref(cap), ctypes.byref(rep),
              ctypes.byref(eag))
    # call 0: 8 eager, 8 captured, 4 eager; call 1: 8 and 8 replayed, 4
    # captured; call 2: all three replayed
    assert (eag.value, cap.value, rep.value) == (2, 2, 5), (eag.value, cap.value, rep.value)
    be.free_cache()
