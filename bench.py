"""Throughput benchmark of the exhaustive FLC scan on one MI355X.

Workload (default): cryo-ET FLC with a wedge x CTF template filter and a
continuous-wedge target filter; 512x512x200 float32 tomogram with 200
embedded copies at SNR 0.1, 48^3 blob template, thresholded + dilated mask,
uniform random rotations.  A step scores one batch of rotations over the
whole padded tomogram and folds it into the MaxScoreOverRotations state
(the public scoring function with the analyzer proxy, exactly as ``scan``
runs it).  Setup (target FFTs, filters) is outside the timed region; inputs
are resident in HBM.

Prints one JSON line.  ``value`` is rotations/s of the whole job (one GPU).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP32_PEAK_GFLOPS = 157300.0  # MI355X FP32 (the kernels issue non-packed FP32: half of it is their ceiling)
# rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the default workload (scaled by
# a 1 GiB copy); its traffic_bytes_per_launch line feeds roofline.traffic
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "fused_scan_flc_512x512x200_t48.md")


#: kernels whose per-launch traffic is proportional to the batch (their
#: work is 3 spectra x B rotations of the same lines), so a figure profiled
#: at another batch size scales to the autotuner's choice as an estimate
TRAFFIC_SCALES_WITH_BATCH = ("fused_pass_b", "fused_pass_b_tail")


def _profiled_traffic():
    """{batch: {kernel: HBM bytes per launch}} from the committed counter
    summary (lines ``traffic_bytes_per_launch@B = {json}``)."""
    out = {}
    try:
        with open(TRAFFIC_PROFILE) as fh:
            for line in fh:
                line = line.strip()
                if line.startswith("traffic_bytes_per_launch@"):
                    head, body = line.split("=", 1)
                    out[int(head.split("@", 1)[1])] = json.loads(body)
    except (OSError, ValueError):
        return {}
    return out


def measured_traffic(kernel, workload, batch, path):
    """(measured, estimated, profiled_batch): the HBM bytes per launch of
    `kernel` from the committed counter summary when it was profiled at this
    batch size; otherwise None and, for kernels whose traffic is
    proportional to the batch, an estimate scaled from the nearest profiled
    batch."""
    if workload != "flc_512x512x200_t48" or path != "fused":
        return None, None, None
    prof = _profiled_traffic()
    if batch in prof and kernel in prof[batch]:
        return int(prof[batch][kernel]), None, batch
    have = [b for b in prof if kernel in prof[b]]
    if not have or kernel not in TRAFFIC_SCALES_WITH_BATCH:
        return None, None, None
    b0 = min(have, key=lambda b: abs(b - batch))
    return None, int(round(prof[b0][kernel] * batch / b0)), b0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--batch", type=int, default=None, help="rotations per step")
    p.add_argument("--workload", default="flc_512x512x200_t48")
    p.add_argument("--path", choices=("fused", "plain"), default="fused")
    p.add_argument("--copies", type=int, default=None)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-rotations", type=int, default=None, help="rotations in the CPU sample")
    p.add_argument("--no-compare-plain", action="store_true",
                   help="skip timing the plain per-rotation path on the same GPU")
    p.add_argument("--calibrate-copy", action="store_true",
                   help="run a 1 GiB device copy first (reference dispatch for HBM byte counters)")
    p.add_argument("--no-refinement", action="store_true", help="skip the gradient-refinement measurement")
    p.add_argument("--no-peaks", action="store_true", help="skip the peak-calling measurement")
    p.add_argument("--no-multi-template", action="store_true", help="skip the multi-template schedule")
    p.add_argument("--num-peaks", type=int, default=1000)
    p.add_argument("--min-distance", type=int, default=5)
    p.add_argument("--refine-peaks", type=int, default=4096)
    p.add_argument("--refine-steps", type=int, default=50)
    return p.parse_args()


# ------------------------------------------------------------------ CPU leg
_CPU_ORACLE = None


def _cpu_worker(rotations):
    t0 = time.perf_counter()
    for r in rotations:
        _CPU_ORACLE.raw_scores(r)
    return time.perf_counter() - t0


_CPU_LOCAL = None


def _cpu_local_worker(args):
    """Oracle local-FLC evaluations at perturbed poses for `seconds`."""
    from oracle import refine as orf

    seed, seconds = args
    target, ts, mask, x0 = _CPU_LOCAL
    rng = np.random.default_rng(seed)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        x = x0[n % len(x0)] + rng.normal(0, [1.0, 1.0, 1.0, 3.0, 3.0, 3.0])
        orf.flc_local_score(target, ts, mask, x)
        n += 1
    return n, time.perf_counter() - t0


def cpu_refinement_rate(target, template, mask, starts, seconds=8.0):
    """Host rate of the reference's score evaluation: the oracle's local FLC
    of the 48^3 template (float64 numpy/scipy restatement of
    tme/matching_optimization.py:488-548) at perturbed start poses, one
    process per host core for `seconds`.  Runs before the GPU is touched;
    returns (evaluations/s over all processes, processes)."""
    import multiprocessing as mp

    from oracle import refine as orf

    global _CPU_LOCAL
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    _CPU_LOCAL = (target, orf.prepare_template(template, mask), mask, starts)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        out = pool.map(_cpu_local_worker, [(k, seconds) for k in range(cores)])
    _CPU_LOCAL = None
    return sum(n / t for n, t in out), cores


def cpu_baseline(target, template, mask, tfilt, gfilt, rotations, score, n_rot):
    """The oracle's float32 dataflow (scipy.fft, one thread per process) over
    a bounded rotation sample, one process per host core, forked after the
    target setup (done once).  Runs before the GPU is touched."""
    import multiprocessing as mp

    from oracle.scan import ScanOracle

    global _CPU_ORACLE
    cores = len(os.sched_getaffinity(0))
    cores = max(1, min(cores, 16))
    n_rot = n_rot or cores
    t0 = time.perf_counter()
    _CPU_ORACLE = ScanOracle(score, target, template, mask, tfilt, pad_target=False, dtype=np.float32,
                             workers=1, target_filter=gfilt)
    setup_s = time.perf_counter() - t0
    chunks = [rotations[i::cores] for i in range(cores)]
    chunks = [c for c in chunks if len(c)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(len(chunks)) as pool:
        pool.map(_cpu_worker, chunks)
    wall = time.perf_counter() - t0
    _CPU_ORACLE = None
    return {
        "value": n_rot / wall,
        "unit": "rotations/s",
        "cores": len(chunks),
        "kind": "port",
        "sample": f"{n_rot} rotations of the same workload, float32 numpy/scipy restatement of the "
        f"reference numpy backend, 1 FFT thread per process; target setup {setup_s:.1f}s excluded",
    }


# ------------------------------------------------------------------ GPU leg
def build_scan(workload, target, template, mask, filters, rotations, path, batch):
    import torch

    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.analyzer.proxy import SharedAnalyzerProxy
    from pytme_amd.backends import backend as be
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan

    be.change_backend("hip", device="cuda", fused=(path == "fused"), rotation_batch=batch)
    md = MatchingData(target, template, template_mask=mask, rotations=rotations)
    md.template_filter, md.target_filter = filters
    setup_fn, scoring = MATCHING_EXHAUSTIVE_REGISTER[workload["score"]]
    setup, cb_args, chunks = prepare_scan(
        md, setup_fn, n_jobs=1, callback_class_args={"score_threshold": 0.0}, pad_target=False,
        pad_template_filter=False, interpolation_order=1,
    )
    proxy = SharedAnalyzerProxy(MaxScoreOverRotations, cb_args)
    rot_dev = chunks[0]
    tuned = None
    if batch is None:
        # the backend's autotuner: times the memory-fitting batch sizes on
        # scratch state during a first call (outside the timed region)
        from pytme_amd.backends import hip_fused

        be.change_backend("hip", device="cuda", fused=True, rotation_batch="auto")
        before = set(hip_fused.TUNED)
        scoring(rotations=rot_dev[: min(256, rot_dev.shape[0])],
                callback=SharedAnalyzerProxy(MaxScoreOverRotations, cb_args), **setup)
        # this template's entry: the one the call added, or the cached one of
        # the same fast shape and template shape
        added = [k for k in hip_fused.TUNED if k not in before]
        tshape = tuple(int(x) for x in np.shape(template))
        key = added[-1] if added else [k for k in hip_fused.TUNED if k[2] == tuple(setup["fast_shape"])
                                       and k[3] == tshape][-1]
        tuned = hip_fused.TUNED[key]
        batch = tuned["batch"]
        be.change_backend("hip", device="cuda", fused=True, rotation_batch=batch, scan_tiles=tuned["tiles"],
                          scan_radix=tuned["radix"])
    torch.cuda.synchronize()
    build_scan.callback_args = cb_args
    return scoring, setup, proxy, rot_dev, batch, tuned


def time_steps(scoring, setup, proxy, rot_dev, batch, steps, warmup, timing=False):
    import torch

    from pytme_amd import _lib

    n = rot_dev.shape[0]
    pos = 0

    def step():
        nonlocal pos
        if pos + batch > n:
            pos = 0
        scoring(rotations=rot_dev[pos : pos + batch], callback=proxy, **setup)
        pos += batch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    if timing:
        _lib.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    report = None
    if timing:
        report = _lib.timing_report()
        _lib.timing_enable(False)
    return t1 - t0, report


def bench_refinement(target, template, mask, spec, n_peaks, steps, cpu_rate=None):
    """Batched Adam refinement (gradient kernel) of n_peaks candidate peaks:
    the embedded copies of the tomogram, each start pose perturbed by
    N(0, 4 deg) per zyz angle and N(0, 1.5 voxel) per axis.  Returns the
    refined-peaks/s record (the local FLC over the template box; the target
    stays resident in HBM)."""
    import torch
    from scipy.spatial.transform import Rotation

    from pytme_amd import _lib
    from pytme_amd.refine import LocalFLC, refine_peaks, score_poses

    rng = np.random.default_rng(spec["seed"] + 3)
    T = int(spec["T"])
    centres = np.asarray(spec["centres"], dtype=np.float64)
    angles = Rotation.from_matrix(np.asarray(spec["copy_rotations"], dtype=np.float64)).as_euler("zyz", degrees=True)
    pick = np.arange(n_peaks) % centres.shape[0]
    t_true = centres[pick] - T // 2
    a_true = angles[pick]
    t0 = t_true + rng.normal(0.0, 1.5, size=t_true.shape)
    a0 = a_true + rng.normal(0.0, 4.0, size=a_true.shape)
    local = LocalFLC(target, template, mask, max_poses=n_peaks)
    s0 = score_poses(local, t0, a0).cpu().numpy()
    s_true = score_poses(local, t_true, a_true).cpu().numpy()
    refine_peaks(local, t0[:64], a0[:64], steps=2)  # warm-up
    torch.cuda.synchronize()
    _lib.timing_enable(True)
    t_start = time.perf_counter()
    _, _, s1, info = refine_peaks(local, t0, a0, steps=steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    report = _lib.timing_report()
    _lib.timing_enable(False)
    kern = {k: {"ms_per_launch": round(v[0] / max(v[1], 1), 3), "launches": v[1]} for k, v in report.items()
            if k.startswith("flc_local")}
    cpu = None
    if cpu_rate is not None:
        # evaluations per peak of the reference's optimiser (basinhopping with
        # COBYLA steps, niter 3, translation bounds +-5, scripts/postprocess.py:
        # 588-613) on 4 start poses, through the GPU-backed score object
        from pytme_amd.matching_optimization import FLC, optimize_match

        evals = []
        for i in range(4):
            obj = FLC(target, template, mask)
            optimize_match(obj, optimization_method="basinhopping", maxiter=3,
                           bounds_translation=tuple((x - 5, x + 5) for x in t0[i]), x0=[*t0[i], *a0[i]],
                           rng=np.random.default_rng(200 + i))
            evals.append(obj.evaluations)
        rate, cores = cpu_rate
        per_peak = float(np.mean(evals))
        cpu = {"value": round(rate / per_peak, 4), "unit": "refined peaks/s", "cores": cores, "kind": "port",
               "sample": f"oracle local-FLC evaluations (float64 numpy/scipy, 48^3) on {cores} processes for 8 s: "
                         f"{rate:.1f} evaluations/s; divided by {per_peak:.0f} evaluations per peak, the mean of the "
                         "reference's derivative-free optimiser (basinhopping niter 3, COBYLA) on 4 start poses"}
    return {
        "metric": "refined peaks/s (local FLC, batched Adam on the device; moments kernel and gradient kernel per step)",
        "value": round(n_peaks / elapsed, 1),
        "unit": "refined peaks/s",
        "peaks": n_peaks,
        "template": T,
        "steps": steps,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "score_start_mean": round(float(np.mean(s0)), 4),
        "score_final_mean": round(float(np.mean(s1)), 4),
        "score_embedded_pose_mean": round(float(np.mean(s_true)), 4),
        "score_evaluations_per_peak": info["score_evaluations_per_peak"],
        "kernels": kern,
        "start_poses": "embedded copies, N(0,4 deg) per zyz angle, N(0,1.5 voxel) per axis",
        "cpu_baseline": cpu,
    }


def bench_called_refinement(proxy, cb_args, target, template, mask, n_peaks, steps, min_distance):
    """The device chain of the peak-calling and local-optimisation steps of
    scripts/postprocess.py:482-613 on the scan's own state: roll and crop
    the max-score and rotation-index volumes, PeakCallerMaximumFilter
    (n_peaks, min_distance, boundary distance T/2), each peak's rotation
    from the rotation-index volume and the analyzer's mapping, zyz angles,
    and the batched gradient refinement (50 Adam steps), all on the device;
    refined peaks/s over the whole chain (the rotation table and the
    refinement plan are built before the timed region)."""
    import torch

    from pytme_amd.postprocess import call_and_refine, rotation_table
    from pytme_amd.refine import LocalFLC

    table = rotation_table(proxy.state[2], device="cuda")
    target_dev = torch.as_tensor(target, device="cuda")
    local = LocalFLC(target_dev, template, mask, max_poses=n_peaks)
    kw = dict(num_peaks=n_peaks, min_distance=min_distance, steps=steps, local=local)
    call_and_refine(proxy.state, cb_args, table, target_dev, template, mask, **dict(kw, steps=1))  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    t, a, s, info = call_and_refine(proxy.state, cb_args, table, target_dev, template, mask, **kw)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    n = int(len(s))
    return {
        "metric": "refined peaks/s from called peaks (scan state -> peak calling -> gradient refinement, on the device)",
        "value": round(n / elapsed, 1),
        "unit": "refined peaks/s",
        "peaks": n,
        "steps": steps,
        "ms_chain": round(elapsed * 1e3, 2),
        "score_start_mean": round(float(np.mean(info.get("scan_scores", [0.0]))), 4),
        "local_score_final_mean": round(float(np.mean(s)) if n else 0.0, 4),
        "note": "candidates are the num_peaks best maxima of the partial max-score volume of the timed steps; "
                "scan_scores are filtered-template FLC, local scores the unfiltered local FLC",
    }


def fft_flops(score, F, rotations_per_s):
    """Nominal FFT throughput: 2.5 F log2 F flops per real 3-D transform
    (5 N log2 N per complex transform, the benchFFT convention) on the
    fast shape, 5 transforms per FLC rotation (template and mask forward,
    S1, S2 and X inverse), 2 per FLCSphericalMask / CC / LCC rotation;
    the pruned template transforms are counted at full size."""
    per = float((5 if score == "FLC" else 2) * 2.5 * F * np.log2(F))
    achieved = per * rotations_per_s / 1e9
    return {"convention": "2.5 F log2 F per real 3-D transform, unpruned", "transforms_per_rotation":
            5 if score == "FLC" else 2, "per_rotation_gflop": round(per / 1e9, 3), "achieved": round(achieved, 1),
            "unit": "GFLOP/s", "peak": FP32_PEAK_GFLOPS, "frac": round(achieved / FP32_PEAK_GFLOPS, 4)}


def bench_multi_template(n_rot, steps):
    """The multi-template schedule on its own workload: one 512x512x200
    tomogram with 50 embedded copies of each of four templates of uneven
    size (24^3, 32^3, 48^3, 64^3 blob densities, thresholded and dilated
    masks, wedge x CTF filters at defocus 2.0 / 2.5 / 3.0 / 3.5 um,
    continuous-wedge target filter).  Each template gets its own fused plan
    with an autotuned rotation batch and graph replay; per template,
    ``steps`` batches of the public scoring function are timed after one
    warm-up batch (setup outside the timed region, as for the main record).
    ``value`` is the schedule's rotations/s: all timed rotations over the
    sum of the timed regions."""
    from pytme_amd.synthetic import make_multi_workload

    target, entries, mspec = make_multi_workload(n_rotations=n_rot)
    spec = {"score": mspec["score"]}
    per, total_rot, total_s = {}, 0, 0.0
    group_entries = []
    for entry in entries:
        tmpl, m, rots = entry["template"], entry["template_mask"], entry["rotations"]
        filters = (entry["template_filter"], entry["target_filter"])
        group_entries.append((tmpl, m, filters, rots))
        T = tmpl.shape[0]
        t0 = time.perf_counter()
        scoring, setup, proxy, rot_dev, batch, tuned = build_scan(spec, target, tmpl, m, filters, rots, "fused",
                                                                  None)
        setup_s = time.perf_counter() - t0
        elapsed, _ = time_steps(scoring, setup, proxy, rot_dev, batch, steps, 1)
        per[f"{T}^3"] = {
            "rotations_per_s": round(steps * batch / elapsed, 2),
            "batch": batch,
            "defocus_A": entry["defocus"],
            "autotune_ms_per_rotation": None if tuned is None else tuned["ms_per_rotation"],
            "tiles": None if tuned is None else list(tuned["tiles"]),
            "radix_plans": None if tuned is None else tuned["radix_plans"],
            "setup_s": round(setup_s, 3),
        }
        total_rot += steps * batch
        total_s += elapsed
        del scoring, setup, proxy, rot_dev
    return {
        "metric": "FLC rotations/s of the multi-template schedule (4 templates, 512x512x200 tomogram)",
        "value": round(total_rot / total_s, 2),
        "unit": "rotations/s",
        "workload": "flc_multi_512x512x200 (50 copies of each template at SNR 0.1)",
        "templates": per,
        "steps_per_template": steps,
        "note": "one fused plan per template, batches replayed as graphs; setup_s (target and filter "
                "setup, plan creation, batch autotuning) is outside the timed region",
        "one_graph_schedule": bench_group_schedule(target, spec, group_entries, steps),
    }


def bench_group_schedule(target, spec, entries, rounds):
    """The same four templates as one schedule (``backend.fused_group``, the
    engine of ``scan_templates(schedule="graph")``): all setups first, then
    every round scans one rotation batch of each template as one replayed
    hipGraph.  The first round runs eagerly; the following ``rounds`` are
    timed (host wall clock between stream synchronisations)."""
    from pytme_amd.analyzer import MaxScoreOverRotations
    from pytme_amd.analyzer.proxy import SharedAnalyzerProxy
    from pytme_amd.backends import backend as be
    from pytme_amd.backends import hip_fused
    from pytme_amd.matching_data import MatchingData
    from pytme_amd.matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, prepare_scan

    hip_fused._free_plan()
    be.change_backend("hip", device="cuda", fused=True, rotation_batch="auto")
    setup_fn, scoring = MATCHING_EXHAUSTIVE_REGISTER[spec["score"]]
    free = be.get_available_memory()
    jobs = []
    t0 = time.perf_counter()
    for tmpl, m, filters, rots in entries:
        fast = tuple(target.shape)
        B = hip_fused.choose_batch(fast, tmpl.shape, 0, True, 2 * free // len(entries))
        md = MatchingData(target, tmpl, template_mask=m, rotations=rots[: (rounds + 1) * B])
        md.template_filter, md.target_filter = filters
        setup, dca, chunks = prepare_scan(md, setup_fn, n_jobs=1, callback_class_args={"score_threshold": -1.0},
                                          pad_target=False, pad_template_filter=False, interpolation_order=1)
        jobs.append((setup, chunks[0], SharedAnalyzerProxy(MaxScoreOverRotations, dca)))
    setup_s = time.perf_counter() - t0
    stats = {}
    be.fused_group(scoring, jobs, stats=stats)
    del jobs
    hip_fused._free_plan()
    return {
        "value": round(stats["rotations_timed"] / stats["seconds"], 2),
        "unit": "rotations/s",
        "batches": stats["batches"],
        "rounds_timed": stats["rounds_timed"],
        "rotations_timed": stats["rotations_timed"],
        "setup_s": round(setup_s, 3),
        "note": "all four templates' batches of a round in one replayed hipGraph (tm_scan_group_batch); "
                "each plan's batch fits its share of half the free HBM",
    }


def bench_peaks(scores, num_peaks, min_distance, reps, with_cpu):
    """Peak calling on the scan's max-score volume: PeakCallerMaximumFilter
    (box maximum filter + peak compaction, then one device call for the
    boundary / score mask, top-2k preselection, merge with the state, sort
    and greedy distance filter) then result(); ms per tomogram with the
    kernels' HIP event times, and the oracle on the same volume (one host
    core) as the CPU baseline and parity check."""
    import torch

    from pytme_amd import _lib
    from pytme_amd.analyzer import PeakCallerMaximumFilter

    kw = dict(shape=tuple(scores.shape), num_peaks=num_peaks, min_distance=min_distance)
    pc = PeakCallerMaximumFilter(**kw)
    res = pc.result(pc(pc.init_state(), scores, rotation_matrix=np.eye(3)))  # warm-up
    torch.cuda.synchronize()
    _lib.timing_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        res = pc.result(pc(pc.init_state(), scores, rotation_matrix=np.eye(3)))
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / reps
    report = _lib.timing_report()
    _lib.timing_enable(False)
    kernels = {k: {"us_per_launch": round(v[0] * 1e3 / max(v[1], 1), 1), "launches": v[1],
                   "GB/s": round(v[2] / max(v[0], 1e-9) / 1e6, 1)}
               for k, v in sorted(report.items(), key=lambda kv: -kv[1][0])}
    # one fused pass for edges up to 8 ("peaks_max_filter"), else the z pass
    name = "peaks_max_filter" if "peaks_max_filter" in report else "peaks_max_filter_z"
    roofline = None
    if name in report:
        ms, count, by = report[name]
        achieved = by / (ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "bytes_per_launch": by / max(count, 1)}
    rec = {
        "metric": "peak calling ms per tomogram (PeakCallerMaximumFilter on the max-score volume)",
        "value": round(elapsed * 1e3, 3),
        "unit": "ms",
        "higher_is_better": False,
        "volume": list(scores.shape),
        "num_peaks": num_peaks,
        "min_distance": min_distance,
        "peaks_found": int(res[0].shape[0]),
        "roofline": roofline,
        "kernels": kernels,
    }
    if with_cpu:
        from oracle import peaks as op

        host = scores.cpu().numpy()
        ref = op.PeakCallerOracle("maximum_filter", **kw)
        t0 = time.perf_counter()
        want = ref.result(ref(ref.init_state(), host))
        cpu_s = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(cpu_s * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the same volume, scipy maximum_filter + numpy restatement of the "
                               "reference numpy path"}
        rec["matches_oracle"] = bool(all(np.array_equal(np.asarray(a), np.asarray(b)) for a, b in zip(res, want)))
    return rec


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the scan shards by rotation "
            "or by target split across GPUs, which this build does not launch.",
            file=sys.stderr,
        )
        sys.exit(2)

    from pytme_amd.synthetic import WORKLOADS, make_workload

    spec = WORKLOADS[args.workload]
    target, template, mask, rotations, filters, spec = make_workload(args.workload, copies=args.copies)
    batch = args.batch or (None if args.path == "fused" else 4)

    cpu = cpu_refine = None
    if not args.no_cpu_baseline:
        tfilt = gfilt = None
        if filters[0] is not None:
            from pytme_amd.matching_data import MatchingData

            conv, fast, ft, shift = MatchingData._fourier_padding(target.shape, template.shape)
            tfilt = filters[0](shape=template.shape, return_real_fourier=True)["data"]
            gfilt = filters[1](shape=fast, return_real_fourier=True)["data"]
        cpu_rots = rotations[: args.cpu_rotations or min(16, len(os.sched_getaffinity(0)))]
        cpu = cpu_baseline(target, template, mask, tfilt, gfilt, cpu_rots, spec["score"], len(cpu_rots))
        if not args.no_refinement and args.workload.startswith("flc_512"):
            from scipy.spatial.transform import Rotation

            T = int(spec["T"])
            starts = np.concatenate([np.asarray(spec["centres"][:16], dtype=np.float64) - T // 2,
                                     Rotation.from_matrix(np.asarray(spec["copy_rotations"][:16], dtype=np.float64)
                                                          ).as_euler("zyz", degrees=True)], axis=1)
            cpu_refine = cpu_refinement_rate(target, template, mask, starts)

    import torch

    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    scoring, setup, proxy, rot_dev, batch, tuned = build_scan(spec, target, template, mask, filters, rotations,
                                                              args.path, batch)
    # K steps with per-kernel HIP events (the roofline and kernel figures);
    # events route each batch through eager launches, so the fused path's
    # `value` is a second timing of K steps without them (the production
    # configuration: each batch's kernel sequence replayed as a hipGraph)
    ev_elapsed, report = time_steps(scoring, setup, proxy, rot_dev, batch, args.steps, args.warmup, timing=True)
    n_rot = args.steps * batch
    extra = {"per_kernel_events_rotations_per_s": round(n_rot / ev_elapsed, 2)}
    elapsed = ev_elapsed
    if args.path == "fused":
        elapsed, _ = time_steps(scoring, setup, proxy, rot_dev, batch, args.steps, 1)
    value = n_rot / elapsed

    fast = tuple(int(x) for x in setup["fast_shape"])
    F = int(np.prod(fast))
    C = int(np.prod(fast[:-1])) * (fast[-1] // 2 + 1)
    model_bytes = (112 * C + 16 * F / batch) if spec["score"] == "FLC" else (40 * C + 4 * F + 16 * F / batch)

    dominant = max(report.items(), key=lambda kv: kv[1][0]) if report else None
    roofline = None
    if dominant:
        name, (ms, count, by) = dominant
        achieved = by / (ms * 1e-3) / 1e9
        measured, estimated, prof_batch = measured_traffic(name, args.workload, batch, args.path)
        roofline = {
            "bound": "hbm",
            "kernel": name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured,
            "traffic_source": None if measured is None else
            f"{os.path.relpath(TRAFFIC_PROFILE, ROOT)} (rocprofv3 FETCH_SIZE + WRITE_SIZE at {prof_batch} rotations)",
            "launches": count,
            "avg_launch_us": round(ms * 1e3 / max(count, 1), 2),
            "bytes_per_launch": by / max(count, 1),
        }
        if estimated is not None:
            roofline["traffic_estimated"] = estimated
            roofline["traffic_scaled_from_batch"] = prof_batch
    kernels = {k: {"ms": round(v[0], 3), "launches": v[1], "GB/s": round(v[2] / max(v[0], 1e-9) / 1e6, 1)}
               for k, v in sorted(report.items(), key=lambda kv: -kv[1][0])} if report else {}

    if not args.no_compare_plain and args.path == "fused":
        # the per-rotation loop (flc_scoring without the fused hook) on the
        # same GPU kernels, same workload and analyzer state
        from pytme_amd.backends import backend as be

        be.change_backend("hip", device="cuda", fused=False)
        p_steps = max(4, args.steps // 2)
        p_el, _ = time_steps(scoring, setup, proxy, rot_dev, 4, p_steps, 1)
        plain = p_steps * 4 / p_el
        extra["plain_path"] = {"rotations_per_s": round(plain, 2), "rotations_timed": p_steps * 4,
                               "fused_over_plain": round(value / plain, 3)}
        extra["plain_path_rotations_per_s"] = round(plain, 2)
        be.change_backend("hip", device="cuda", fused=True, rotation_batch=batch,
                          scan_tiles=None if tuned is None else tuned["tiles"],
                          scan_radix=None if tuned is None else tuned["radix"])

    out = {
        "metric": "FLC rotations/s, 512x512x200 tomogram vs 48^3 template" if args.workload.startswith("flc_512")
        else f"{spec['score']} rotations/s ({args.workload})",
        "value": round(value, 3),
        "unit": "rotations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "score": spec["score"],
            "tomogram": list(spec["shape"]),
            "template": spec["T"],
            "fast_shape": list(fast),
            "rotations_per_step": batch,
            "autotune_ms_per_rotation": None if tuned is None else tuned["ms_per_rotation"],
            "autotune_tiles": None if tuned is None else list(tuned["tiles"]),
            "autotune_tiles_ms_per_rotation": None if tuned is None else tuned["tiles_ms_per_rotation"],
            "autotune_radix": None if tuned is None else tuned["radix"],
            "radix_plans": None if tuned is None else tuned["radix_plans"],
            "autotune_radix_ms_per_rotation": None if tuned is None else tuned["radix_ms_per_rotation"],
            "path": args.path,
            "filters": "wedge x CTF template filter, continuous-wedge target filter" if spec["filters"] else None,
        },
        "fft_flops": fft_flops(spec["score"], F, value),
        "whole_step_model_GBps": round(model_bytes * value / 1e9, 1),
        "whole_step_model_frac": round(model_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    if not args.no_peaks and args.path == "fused":
        out["peak_calling"] = bench_peaks(proxy.state[0], args.num_peaks, args.min_distance, 5,
                                          not args.no_cpu_baseline)
    if not args.no_refinement and args.workload.startswith("flc_512") and args.path == "fused":
        out["called_peaks_refinement"] = bench_called_refinement(
            proxy, build_scan.callback_args, target, template, mask, args.refine_peaks, args.refine_steps,
            args.min_distance)
    if not args.no_refinement and args.workload.startswith("flc_512"):
        out["refinement"] = bench_refinement(target, template, mask, spec, args.refine_peaks, args.refine_steps,
                                             cpu_refine)
    if not args.no_multi_template and args.workload.startswith("flc_512") and args.path == "fused":
        out["multi_template"] = bench_multi_template(512, 4)
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
