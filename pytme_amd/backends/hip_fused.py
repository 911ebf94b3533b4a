"""Batched fused scan of the hip backend.

The registered scoring functions (``flc_scoring`` / ``corr_scoring``) offer
their whole rotation set here first.  When the analyzer is an unshared
:class:`~pytme_amd.analyzer.MaxScoreOverRotations` behind a
:class:`~pytme_amd.analyzer.proxy.SharedAnalyzerProxy`, the rotations are
scored in batches by ``tm_scan_batch`` (see ``csrc/fused.hip``), which folds
them straight into the analyzer's max/argmax state: the per-rotation score
volume is never materialised.  Rotation indices are assigned exactly as the
analyzer's ``__call__`` would assign them (call order, or rotation bytes),
so results, tie rule and the rotation mapping are those of the
per-rotation loop.

Anything else (other analyzers, interpolation orders, unusual filter or
numerator shapes) returns ``NotImplemented`` and the caller runs the plain
per-rotation loop on the same GPU kernels.
"""

import ctypes
import time
from typing import Optional

import numpy as np

from .. import _lib

__all__ = ["run", "choose_batch", "autotune_batch", "release", "invalidate", "TUNED"]

#: the kept plan: ``key`` identifies its arguments, ``base`` the same without
#: the batch size, ``keep`` holds its input arrays, ``seen`` the state of the
#: target spectra when the plan last copied them
_PLAN = {"key": None, "base": None, "handle": None, "keep": None, "seen": None}

#: measured batch choices: (device, mode, fast_shape, template_shape, filter,
#: normalize, candidates) -> {"batch": B, "ms_per_rotation": {B: ms}}
TUNED = {}


def _free_plan():
    if _PLAN["handle"] is not None:
        _lib.call("tm_scan_plan_destroy", _PLAN["handle"])
    _PLAN.update(key=None, base=None, handle=None, keep=None, seen=None)


def release():
    """Free the kept fused plan (its batch workspace: tens of GB for a
    512x512x200 target).  The hip backend's ``free_cache`` calls this, so
    :func:`~pytme_amd.matching_exhaustive.scan` releases it when a scan
    ends."""
    _free_plan()


def invalidate():
    """Make the kept plan copy the target spectra again at its next batch.

    The plan keeps a row-padded copy of the spectra it was created with and
    refreshes it when their tensors' version counters changed or any
    library call ran since its last batch.  Writes through another library
    (raw pointers that neither torch nor this library sees) need this call.
    """
    _PLAN["seen"] = None


def _spectra_state(ft_target, ft2):
    return (_lib.CALLS, ft_target._version, None if ft2 is None else ft2._version)


def choose_batch(fast_shape, template_shape, mode: int, with_filter: bool, free_bytes: int,
                 requested: Optional[int] = None, n_rotations: int = 1 << 30) -> int:
    """Rotations per fused launch.

    The per-rotation cost is dominated by the three HBM passes over the
    NQ correlation spectra; larger batches amortise the resident target
    spectra (read once per batch in pass A) and the max/argmax state
    (read-modified-written once per batch in pass C).  The batch is the
    largest of 64, 32, 16, ... that keeps the workspace under a quarter of
    free HBM (512x512x200 FLC, measured: 1325 / 1372 / 1389 / 1401
    rotations/s at 16 / 32 / 48 / 64; the 64-rotation workspace is 42 GB).
    """
    if requested:
        return int(max(1, min(requested, n_rotations)))
    lib = _lib.lib()
    fs = _lib.i64_array(fast_shape)
    ts = _lib.i32_array(template_shape)
    b = 64
    while b > 1 and lib.tm_scan_workspace_bytes(mode, fs, ts, b, int(with_filter)) > free_bytes // 4:
        b //= 2
    return int(max(1, min(b, n_rotations)))


#: tile shapes the autotuner times at the chosen batch: (pass A columns,
#: pass C rows); 0 is the kernels' default, combinations a plan does not
#: support are skipped
TILE_CANDIDATES = ((0, 0), (8, 0), (0, 10), (8, 10))


def _time_plan(t, handle, stream, rots, idx_dev, B, scores, rot, graph):
    """ms per rotation of two timed batches on scratch state, after the
    warm-up batches that settle the launch mode production uses (with graph
    replay: the eager first batch and the captured second one)."""
    _lib.call("tm_scan_plan_use_graph", handle, int(graph))
    warm = 2 if graph else 1
    start, stop = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    for rep in range(warm + 2):
        if rep == warm:
            start.record()
        _lib.call("tm_scan_batch", handle, rots[0].data_ptr(), idx_dev[0].data_ptr(), B, scores.data_ptr(),
                  rot.data_ptr(), stream)
    stop.record()
    stop.synchronize()
    return start.elapsed_time(stop) / (2 * B)


def autotune_batch(backend, key, create, stream, rots, idx_dev, state_shape, candidates):
    """Time each candidate batch size, then the tile shapes of the passes
    (``TILE_CANDIDATES``, ``tm_scan_plan_set_tiles``) at the fastest size,
    on scratch max/argmax state (the caller's state is not touched) with the
    call's own rotations: two timed batches each after the warm-up, in the
    backend's launch mode (graph replay or eager), HIP events on the launch
    stream; keep the fastest per rotation.  ``create(B)``
    returns a plan handle for batch B.  Results are cached in :data:`TUNED`
    (``batch`` and ``tiles``)."""
    t = backend._torch
    if key in TUNED:
        return TUNED[key]["batch"]
    graph = bool(getattr(backend, "scan_graph", True))
    scores = t.zeros(state_shape, dtype=t.float32, device=rots.device)
    rot = t.full(state_shape, -1, dtype=t.int32, device=rots.device)
    n = int(rots.shape[0])
    timings = {}
    for B in candidates:
        if B > n:
            continue
        handle = create(B)
        try:
            timings[B] = _time_plan(t, handle, stream, rots, idx_dev, B, scores, rot, graph)
        finally:
            _lib.call("tm_scan_plan_destroy", handle)
    best = min(timings, key=timings.get) if timings else min(candidates)
    tiles = {}
    if best in timings:
        tiles[TILE_CANDIDATES[0]] = timings[best]
        for combo in TILE_CANDIDATES[1:]:
            handle = create(best)
            try:
                try:
                    _lib.call("tm_scan_plan_set_tiles", handle, *combo)
                except ValueError:
                    continue  # not a tiling of this plan
                tiles[combo] = _time_plan(t, handle, stream, rots, idx_dev, best, scores, rot, graph)
            finally:
                _lib.call("tm_scan_plan_destroy", handle)
    best_tiles = min(tiles, key=tiles.get) if tiles else TILE_CANDIDATES[0]
    # radix plan variants of the compile-time FFTs at the chosen batch and
    # tiles (their results agree to rounding, checked against the oracle)
    radix = {}
    if best in timings:
        radix[0] = tiles.get(best_tiles, timings[best])
        for variant in RADIX_CANDIDATES[1:]:
            handle = create(best)
            try:
                _lib.call("tm_scan_plan_set_tiles", handle, *best_tiles)
                _lib.call("tm_scan_plan_set_radix", handle, variant)
                radix[variant] = _time_plan(t, handle, stream, rots, idx_dev, best, scores, rot, graph)
            finally:
                _lib.call("tm_scan_plan_destroy", handle)
    best_radix = min(radix, key=radix.get) if radix else 0
    TUNED[key] = {
        "batch": int(best),
        "tiles": tuple(int(x) for x in best_tiles),
        "radix": int(best_radix),
        "radix_plans": radix_plans(key[2], int(best_radix)),
        "ms_per_rotation": {int(b): round(v, 5) for b, v in timings.items()},
        "tiles_ms_per_rotation": {f"{a}x{c}": round(v, 5) for (a, c), v in tiles.items()},
        "radix_ms_per_rotation": {int(v): round(ms, 5) for v, ms in radix.items()},
    }
    return int(best)


#: radix plan variants the autotuner times (tm_scan_plan_set_radix)
RADIX_CANDIDATES = (0, 1)


def radix_plans(fast_shape, variant: int):
    """{"z": "8x8x8", "y": ..., "x": ...}: the forward radices of the batch
    passes' compile-time FFTs in a radix plan variant (the x pass transforms
    half the x length); axes without a compile-time plan are omitted."""
    lib = _lib.lib()
    out = {}
    lengths = {"z": int(fast_shape[0]), "y": int(fast_shape[1]), "x": int(fast_shape[2]) // 2}
    compiled = {"z": (128, 256, 512, 1024), "y": (128, 256, 512, 1024), "x": (64, 100, 128)}
    for axis, n in lengths.items():
        if n not in compiled[axis]:
            continue
        buf, cnt = (ctypes.c_int * 16)(), ctypes.c_int()
        _lib.check(lib.tm_fft_radix_plan(n, int(variant), buf, 16, ctypes.byref(cnt)), "tm_fft_radix_plan")
        if cnt.value:
            out[axis] = "x".join(str(buf[i]) for i in range(cnt.value))
    return out


def _proxy_state(callback):
    """(analyzer, state) of an unshared ``MaxScoreOverRotations`` behind the
    scan's callback proxy, or None.

    Two proxies qualify: this package's ``SharedAnalyzerProxy``, and pyTME's
    own (tme/analyzer/proxy.py:72-123 around tme/analyzer/aggregation.py:
    31-431, when an installed pyTME runs on this backend): the latter is
    recognised by its attributes ``_analyzer``, ``_state`` and ``_shared``
    and an analyzer class named ``MaxScoreOverRotations`` whose state is the
    same (scores, rotations, mapping) tuple of backend arrays.
    """
    from ..analyzer.aggregation import MaxScoreOverRotations
    from ..analyzer.proxy import SharedAnalyzerProxy

    if isinstance(callback, SharedAnalyzerProxy):
        analyzer, state, shared = callback.analyzer, callback.state, callback.is_shared
        if type(analyzer) is not MaxScoreOverRotations:
            return None
    elif all(hasattr(callback, a) for a in ("_analyzer", "_state", "_shared")):
        analyzer, state, shared = callback._analyzer, callback._state, callback._shared
        if type(analyzer).__name__ != "MaxScoreOverRotations" or not hasattr(analyzer, "_inversion_mapping"):
            return None
        if not (isinstance(state, tuple) and len(state) == 3 and isinstance(state[2], dict)):
            return None
    else:
        return None
    if shared or getattr(analyzer, "_jax_mode", False):
        return None
    return analyzer, state


def _rotation_indices(job):
    from ..analyzer.aggregation import assign_rotation_indices

    return assign_rotation_indices(job["state"][2], job["rots"], bool(job["analyzer"]._inversion_mapping))


def _job(backend, score: str, kw):
    """Device arrays and plan arguments of one fused scan (``kw`` are the
    scoring function's arguments), or NotImplemented when the fused path
    does not apply."""
    t = backend._torch
    if int(kw.get("interpolation_order", 1)) != 1:
        return NotImplemented
    got = _proxy_state(kw.get("callback"))
    if got is None:
        return NotImplemented
    analyzer, state = got
    max_scores, max_rot, _ = state

    template = backend.from_sharedarr(kw["template"])
    ft_target = backend.from_sharedarr(kw["ft_target"])
    tfilter = backend.from_sharedarr(kw["template_filter"])
    fast_shape = tuple(int(x) for x in kw["fast_shape"])
    rotations = kw["rotations"]
    if template.ndim != 3 or len(fast_shape) != 3:
        return NotImplemented
    tshape = tuple(int(x) for x in template.shape)
    if tuple(max_scores.shape) != fast_shape or not backend._dev_f32(max_scores):
        return NotImplemented
    if max_rot.dtype != t.int32 or not max_rot.is_contiguous():
        return NotImplemented

    filt = None
    if backend.size(tfilter) != 1:
        if tuple(tfilter.shape) != (tshape[0], tshape[1], tshape[2] // 2 + 1):
            return NotImplemented
        filt = tfilter.to(t.float32).contiguous()

    mask = ft2 = inv_den = None
    normalize, mask_sum = 0, 0.0
    if score == "flc":
        mode = 0
        mask = backend.from_sharedarr(kw["template_mask"]).to(t.float32).contiguous()
        ft2 = backend.from_sharedarr(kw["ft_target2"]).contiguous()
        if tuple(mask.shape) != tshape:
            return NotImplemented
    elif score == "corr":
        mode = 1
        numerator = backend.from_sharedarr(kw["numerator"])
        inv = backend.from_sharedarr(kw["inv_denominator"])
        if tuple(numerator.shape) == fast_shape:
            return NotImplemented  # F-shaped numerators are not produced by the registered scores
        if tuple(inv.shape) == fast_shape:
            inv_den = inv.to(t.float32).contiguous()
        tm = kw.get("template_mask")
        if tm is not None:
            mask = backend.from_sharedarr(tm).to(t.float32).contiguous()
            if tuple(mask.shape) != tshape:
                return NotImplemented
            normalize = 1
            mask_sum = float(mask.double().sum())
    else:
        return NotImplemented

    templ = template.to(t.float32).contiguous()
    ft_target = ft_target.contiguous()
    if rotations.ndim != 3 or rotations.shape[0] == 0:
        return NotImplemented
    rots = backend.to_backend_array(rotations).to(t.float32).contiguous()
    return dict(
        mode=mode, fast_shape=fast_shape, tshape=tshape, templ=templ, mask=mask, filt=filt, ft_target=ft_target,
        ft2=ft2, inv_den=inv_den, normalize=normalize, mask_sum=mask_sum, rots=rots, analyzer=analyzer, state=state,
        max_scores=max_scores, max_rot=max_rot,
    )


def _ptr(x):
    return None if x is None else x.data_ptr()


def _create_plan(job, batch):
    handle = ctypes.c_void_p()
    _lib.call(
        "tm_scan_plan_create", job["mode"], _lib.i64_array(job["fast_shape"]), _lib.i32_array(job["tshape"]),
        _ptr(job["templ"]), _ptr(job["mask"]), _ptr(job["filt"]), _ptr(job["ft_target"]), _ptr(job["ft2"]),
        _ptr(job["inv_den"]), job["normalize"], job["mask_sum"], batch, ctypes.byref(handle),
    )
    return handle.value


def run(backend, score: str, **kw):
    """Fused scan of ``kw['rotations']`` into ``kw['callback']`` or NotImplemented."""
    t = backend._torch
    job = _job(backend, score, kw)
    if job is NotImplemented:
        return NotImplemented
    mode, fast_shape, tshape, rots = job["mode"], job["fast_shape"], job["tshape"], job["rots"]
    templ, mask, filt, ft_target, ft2, inv_den = (job[k] for k in ("templ", "mask", "filt", "ft_target", "ft2",
                                                                   "inv_den"))
    normalize, mask_sum = job["normalize"], job["mask_sum"]
    max_scores, max_rot = job["max_scores"], job["max_rot"]
    n = int(rots.shape[0])

    keep = (templ, mask, filt, ft_target, ft2, inv_den)
    base = (
        t.cuda.current_device(), mode, fast_shape, tshape, normalize, mask_sum,
        tuple(0 if x is None else x.data_ptr() for x in keep), bool(getattr(backend, "scan_graph", True)),
    )
    if _PLAN["base"] != base:
        _free_plan()  # before measuring the free memory the new plan may use
    free = backend.get_available_memory()
    auto = backend.rotation_batch == "auto"
    B = choose_batch(fast_shape, tshape, mode, filt is not None, free, None if auto else backend.rotation_batch, n)

    ptr = _ptr
    stream = backend._stream()

    def create(batch):
        return _create_plan(job, batch)

    idx = _rotation_indices(job)
    idx_dev = t.tensor(np.asarray(idx, dtype=np.int32), device=rots.device)
    if auto:
        candidates = tuple(b for b in (B, B // 2, B // 4) if b >= 1)
        tune_key = (t.cuda.current_device(), mode, fast_shape, tshape, filt is not None, normalize, candidates)
        if tune_key not in TUNED:
            _free_plan()  # the candidates' workspaces need the memory
        B = autotune_batch(backend, tune_key, create, stream, rots, idx_dev, fast_shape, candidates)
    key = base + (B,)
    # tile shapes: the backend's, the autotuner's, or the kernels' defaults
    tiles = getattr(backend, "scan_tiles", None)
    if tiles is None and auto:
        tiles = TUNED.get(tune_key, {}).get("tiles")
    tiles = (0, 0) if tiles is None else tiles
    radix = getattr(backend, "scan_radix", None)
    if radix is None:
        radix = TUNED.get(tune_key, {}).get("radix", 0) if auto else 0
    if _PLAN["key"] != key:
        _free_plan()
        _PLAN.update(key=key, base=base, handle=create(B), keep=keep)
        _lib.call("tm_scan_plan_use_graph", _PLAN["handle"], int(getattr(backend, "scan_graph", True)))
    elif _PLAN["seen"] != _spectra_state(ft_target, ft2):
        # the plan holds a row-padded copy of the target spectra: refresh it
        # when the caller's arrays may hold new values (a torch write bumps
        # their version counters; a library call may have written them)
        _lib.call("tm_scan_plan_set_target", _PLAN["handle"], ptr(ft_target), ptr(ft2), stream)
    _lib.call("tm_scan_plan_set_tiles", _PLAN["handle"], *tiles)
    _lib.call("tm_scan_plan_set_radix", _PLAN["handle"], int(radix))

    for start in range(0, n, B):
        nb = min(B, n - start)
        _lib.call(
            "tm_scan_batch", _PLAN["handle"], rots[start].data_ptr(), idx_dev[start].data_ptr(), nb,
            max_scores.data_ptr(), max_rot.data_ptr(), stream,
        )
    _PLAN["seen"] = _spectra_state(ft_target, ft2)
    return kw["callback"]


def run_group(backend, kws, stats=None):
    """Several fused scans (``kws``: one scoring function's arguments per
    template, each with its own analyzer) stepped together: every round runs
    one batch of each template as one replayed hipGraph
    (``tm_scan_group_batch``).  Every template gets a plan whose batch is the
    largest of 64, 32, ... fitting its share of half the free HBM (or the
    backend's fixed ``rotation_batch``) and the backend's or the default tile
    shapes.  Results equal the templates' separate scans.  ``stats`` (a
    dict) receives the batch sizes and the wall time of the rounds after the
    first (the first runs eagerly and settles every launch attribute)."""
    t = backend._torch
    jobs = []
    for i, (score, kw) in enumerate(kws):
        job = _job(backend, score, kw)
        if job is NotImplemented:
            raise ValueError(f"template {i}: the fused scan does not apply to these arguments")
        job["idx_dev"] = t.tensor(np.asarray(_rotation_indices(job), dtype=np.int32), device=job["rots"].device)
        jobs.append(job)
    if not jobs:
        return
    _free_plan()  # the group's workspaces need the memory
    fixed = None if backend.rotation_batch == "auto" else backend.rotation_batch
    free = backend.get_available_memory()
    share = 2 * free // len(jobs)  # choose_batch keeps each plan under a quarter of its share
    tiles = getattr(backend, "scan_tiles", None) or (0, 0)
    stream = backend._stream()
    plans, group = [], None
    try:
        for job in jobs:
            job["B"] = choose_batch(job["fast_shape"], job["tshape"], job["mode"], job["filt"] is not None, share,
                                    fixed, int(job["rots"].shape[0]))
            plans.append(_create_plan(job, job["B"]))
            _lib.call("tm_scan_plan_set_tiles", plans[-1], *tiles)
            _lib.call("tm_scan_plan_set_radix", plans[-1], int(getattr(backend, "scan_radix", None) or 0))
        n = len(plans)
        handles = (ctypes.c_void_p * n)(*plans)
        gh = ctypes.c_void_p()
        _lib.call("tm_scan_group_create", handles, n, ctypes.byref(gh))
        group = gh.value
        rounds = max(-(-int(j["rots"].shape[0]) // j["B"]) for j in jobs)
        t0, timed_rot = None, 0
        for r in range(rounds):
            if r == 1 and stats is not None:
                _lib.call("tm_stream_sync", stream)
                t0 = time.perf_counter()
            rp, ip, nb = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)(), (ctypes.c_int * n)()
            ms, mr = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
            for i, job in enumerate(jobs):
                start = r * job["B"]
                cnt = max(0, min(job["B"], int(job["rots"].shape[0]) - start))
                nb[i] = cnt
                timed_rot += cnt if r >= 1 else 0
                if cnt:
                    rp[i] = job["rots"][start].data_ptr()
                    ip[i] = job["idx_dev"][start].data_ptr()
                    ms[i] = job["max_scores"].data_ptr()
                    mr[i] = job["max_rot"].data_ptr()
            _lib.call("tm_scan_group_batch", group, rp, ip, nb, ms, mr, stream)
        _lib.call("tm_stream_sync", stream)
        if stats is not None:
            stats.update(batches=[j["B"] for j in jobs], rounds=rounds, rounds_timed=max(0, rounds - 1),
                         rotations_timed=timed_rot, seconds=None if t0 is None else time.perf_counter() - t0)
    finally:
        if group is not None:
            _lib.call("tm_scan_group_destroy", group)
        for h in plans:
            _lib.call("tm_scan_plan_destroy", h)
