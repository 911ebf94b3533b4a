"""Exhaustive scan orchestration (mirrors ``tme/matching_exhaustive.py``).

``scan_subsets`` splits target/template, runs ``scan`` per split and merges
the analyzer results by offset.  ``scan`` subsets and pads the data, moves it
to the active backend, computes the Fourier padding, evaluates composable
filters once, runs the score's setup and fans the rotations out to scoring
jobs whose analyzer results it merges.
"""

import sys
import warnings
from functools import wraps
from itertools import product
from math import prod
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from joblib import Parallel, delayed

from .analyzer.proxy import SharedAnalyzerProxy
from .backends import backend as be
from .filters import Compose
from .matching_scores import MATCHING_EXHAUSTIVE_REGISTER
from .matching_utils import split_shape
from .memory import MATCHING_MEMORY_REGISTRY, MatchingMemoryUsage

__all__ = [
    "scan",
    "scan_subsets",
    "prepare_scan",
    "register_matching_exhaustive",
    "MATCHING_EXHAUSTIVE_REGISTER",
    "scan_templates",
]


def _wrap_backend(func):
    """Re-create the active backend inside a (possibly forked) job."""

    @wraps(func)
    def wrapper(*args, backend_name: str, backend_args: Dict, **kwargs):
        from .backends import backend as _be

        _be.change_backend(backend_name, **backend_args)
        return func(*args, **kwargs)

    return wrapper


def _setup_template_filter_apply_target_filter(
    matching_data,
    fast_shape: Tuple[int],
    fast_ft_shape: Tuple[int],
    pad_template_filter: bool = True,
):
    """Evaluate composable filters once; filter the target in place.

    Returns the template filter as a float array: the rfft half spectrum at
    the template shape (``pad_template_filter=False``) or at ``fast_shape``,
    or a size-1 array meaning "no filter".  A target filter multiplies the
    target spectrum once (tme/matching_exhaustive.py:39-119).
    """
    backend_arr = type(be.zeros((1), dtype=be._float_dtype))
    template_filter = be.full(shape=(1,), fill_value=1, dtype=be._float_dtype)
    if isinstance(matching_data.template_filter, backend_arr):
        template_filter = matching_data.template_filter
    filter_template = isinstance(matching_data.template_filter, Compose)
    filter_target = isinstance(matching_data.target_filter, Compose)
    if not filter_template and not filter_target:
        return be.astype(template_filter, be._float_dtype)

    real_shape = tuple(int(x) for x in fast_shape)
    cmpl_shape = tuple(int(x) for x in fast_ft_shape)
    real_template_shape, cmpl_template_shape = real_shape, cmpl_shape
    if filter_template and not pad_template_filter:
        real_template_shape = tuple(int(x) for x in matching_data._output_template_shape)
        cmpl_template_shape = real_template_shape[:-1] + (real_template_shape[-1] // 2 + 1,)

    target_temp = be.topleft_pad(matching_data.target, fast_shape)
    target_temp_ft = be.rfftn(target_temp)
    filter_kwargs = {
        "return_real_fourier": True,
        "shape_is_real_fourier": False,
        "data_rfft": target_temp_ft,
        "batch_dimension": matching_data._target_dim,
    }
    if filter_template:
        template_filter = matching_data.template_filter(shape=real_template_shape, **filter_kwargs)["data"]
        template_filter = be.to_backend_array(template_filter)
        if be.size(template_filter) != prod(cmpl_template_shape):
            raise ValueError("template filter does not match the template's half spectrum")
        template_filter = be._array_backend.reshape(template_filter, cmpl_template_shape)

    if filter_target:
        target_filter = matching_data.target_filter(shape=real_shape, weight_type=None, **filter_kwargs)["data"]
        target_filter = be.astype(be.to_backend_array(target_filter), be._float_dtype)
        target_filter = be._array_backend.reshape(target_filter, cmpl_shape).contiguous()
        target_temp_ft = be.multiply(target_temp_ft, target_filter, out=target_temp_ft)
        target_temp = be.irfftn(target_temp_ft, s=tuple(target_temp.shape))
        matching_data._target = be.topleft_pad(target_temp, matching_data.target.shape)

    return be.astype(be.to_backend_array(template_filter), be._float_dtype).contiguous()


def device_memory_handler(func: Callable):
    """Provide a shared-memory manager (CPU pools) and select the device."""

    @wraps(func)
    def inner(*args, **kwargs):
        gpu_index = kwargs.pop("gpu_index", 0)
        n_jobs = kwargs.get("n_jobs", 4)
        with be.set_device(gpu_index):
            if n_jobs > 1:
                from multiprocessing.managers import SharedMemoryManager

                with SharedMemoryManager() as smh:
                    return func(*args, shm_handler=smh, **kwargs)
            return func(*args, shm_handler=None, **kwargs)

    return inner


def prepare_scan(
    matching_data,
    matching_setup: Callable,
    n_jobs: int = 1,
    callback_class_args: Dict = {},
    pad_target: bool = True,
    pad_template_filter: bool = True,
    interpolation_order: int = None,
    shm_handler=None,
    target_slice=None,
    template_slice=None,
):
    """Setup phase of :func:`scan`: subset, move to the backend, pad, filter
    and run the score's setup.

    Returns ``(setup, default_callback_args, rotation_chunks)`` where
    ``setup`` are the keyword arguments of the scoring function (without
    ``rotations``/``callback``).
    """
    if interpolation_order is None:
        interpolation_order = getattr(be, "default_interpolation_order", 3)
    matching_data, translation_offset = matching_data.subset_by_slice(
        target_slice=target_slice,
        template_slice=template_slice,
        target_pad=matching_data.target_padding(pad_target=pad_target),
    )
    matching_data.to_backend()
    template_shape = matching_data._batch_shape(matching_data.template.shape, matching_data._template_batch)
    conv, fwd, inv, shift = matching_data.fourier_padding(pad_target=pad_target)

    template_filter = _setup_template_filter_apply_target_filter(
        matching_data=matching_data,
        fast_shape=fwd,
        fast_ft_shape=inv,
        pad_template_filter=pad_template_filter,
    )

    default_callback_args = {
        "shape": fwd,
        "offset": translation_offset,
        "fourier_shift": shift,
        "fast_shape": fwd,
        "targetshape": matching_data._output_shape,
        "templateshape": template_shape,
        "convolution_shape": conv,
        "thread_safe": n_jobs > 1,
        "convolution_mode": "valid" if pad_target else "same",
        "shm_handler": shm_handler,
        "only_unique_rotations": True,
        "aggregate_axis": matching_data._batch_axis(matching_data._batch_mask),
        "n_rotations": matching_data.rotations.shape[0],
    }
    callback_class_args = dict(callback_class_args)
    callback_class_args["inversion_mapping"] = n_jobs == 1
    default_callback_args.update(callback_class_args)

    setup = matching_setup(
        matching_data=matching_data,
        template_filter=template_filter,
        fast_shape=fwd,
        fast_ft_shape=inv,
        shm_handler=shm_handler,
    )
    setup["interpolation_order"] = interpolation_order
    setup["template_filter"] = be.to_sharedarr(template_filter, shm_handler)
    rotation_chunks = matching_data._split_rotations_on_jobs(n_jobs)
    matching_data._free_data()
    return setup, default_callback_args, rotation_chunks


@device_memory_handler
def scan(
    matching_data,
    matching_setup: Callable,
    matching_score: Callable,
    n_jobs: int = 4,
    callback_class=None,
    callback_class_args: Dict = {},
    pad_target: bool = True,
    pad_template_filter: bool = True,
    interpolation_order: int = None,
    jobs_per_callback_class: int = 8,
    shm_handler=None,
    target_slice=None,
    template_slice=None,
) -> Optional[Tuple]:
    """Run template matching on one split (tme/matching_exhaustive.py:144-295)."""
    setup, default_callback_args, rotation_chunks = prepare_scan(
        matching_data,
        matching_setup,
        n_jobs=n_jobs,
        callback_class_args=callback_class_args,
        pad_target=pad_target,
        pad_template_filter=pad_template_filter,
        interpolation_order=interpolation_order,
        shm_handler=shm_handler,
        target_slice=target_slice,
        template_slice=template_slice,
    )
    be.free_cache()

    n_callback_classes = max(n_jobs // jobs_per_callback_class, 1)
    callback_classes = [
        SharedAnalyzerProxy(callback_class, default_callback_args, shm_handler=shm_handler if n_jobs > 1 else None)
        if callback_class
        else None
        for _ in range(n_callback_classes)
    ]
    ret = Parallel(n_jobs=n_jobs)(
        delayed(_wrap_backend(matching_score))(
            backend_name=be._backend_name,
            backend_args=be._backend_args,
            rotations=rotation,
            callback=callback_classes[index % n_callback_classes],
            **setup,
        )
        for index, rotation in enumerate(rotation_chunks)
    )
    callbacks = [cb.result(**default_callback_args) for cb in ret[:n_callback_classes] if cb]
    be.free_cache()
    if callback_class:
        ret = callback_class.merge(callbacks, **default_callback_args)
    return ret


def scan_subsets(
    matching_data,
    matching_score: Callable,
    matching_setup: Callable,
    callback_class=None,
    callback_class_args: Dict = {},
    job_schedule: Tuple[int] = (1, 1),
    target_splits: Dict = {},
    template_splits: Dict = {},
    pad_target_edges: bool = False,
    pad_template_filter: bool = True,
    interpolation_order: int = None,
    jobs_per_callback_class: int = 8,
    backend_name: str = None,
    backend_args: Dict = {},
    verbose: bool = False,
    **kwargs,
) -> Optional[Tuple]:
    """Scan every (target split, template split) pair and merge the results.

    ``job_schedule = (outer, inner)``: splits processed in parallel and
    rotation jobs per split (tme/matching_exhaustive.py:298-465).
    ``interpolation_order=None`` selects the active backend's default.
    """
    template_splits = split_shape(matching_data._template.shape, splits=template_splits)
    target_splits = split_shape(matching_data._target.shape, splits=target_splits)
    if len(target_splits) > 1 and not pad_target_edges:
        warnings.warn(
            "Target splitting without padding target edges leads to unreliable "
            "similarity estimates around the split border."
        )
    splits = tuple(product(target_splits, template_splits))
    outer_jobs, inner_jobs = job_schedule
    results = Parallel(n_jobs=outer_jobs, verbose=verbose)(
        [
            delayed(_wrap_backend(scan))(
                backend_name=be._backend_name,
                backend_args=be._backend_args,
                matching_data=matching_data,
                matching_score=matching_score,
                matching_setup=matching_setup,
                n_jobs=inner_jobs,
                callback_class=callback_class,
                callback_class_args=callback_class_args,
                interpolation_order=interpolation_order,
                pad_target=pad_target_edges,
                gpu_index=index % outer_jobs,
                pad_template_filter=pad_template_filter,
                target_slice=target_split,
                template_slice=template_split,
            )
            for index, (target_split, template_split) in enumerate(splits)
        ]
    )
    matching_data._free_data()
    if callback_class is not None:
        return callback_class.merge(results, **callback_class_args)
    return None


def scan_templates(
    target,
    templates: Sequence[Dict],
    score: str = "FLC",
    callback_class=None,
    callback_class_args: Dict = None,
    schedule: str = "sequential",
    **scan_kwargs,
) -> List[Optional[Tuple]]:
    """Several templates of uneven size against one target, one result per
    template (a multi-template schedule; the reference runs one template per
    invocation of scripts/match_template.py, tme/matching_exhaustive.py:
    298-465 each).

    Each entry of ``templates`` is a dict with ``template`` and optional
    ``template_mask``, ``rotations``, ``template_filter`` and
    ``target_filter`` (a ``Compose``).  ``schedule="sequential"``: every
    template runs through :func:`scan_subsets` with its own analyzer (on the
    hip backend each template's fused plan is created once and kept for all
    its rotations); ``scan_kwargs`` are passed to :func:`scan_subsets`.
    ``schedule="graph"`` (hip backend, fused scan, whole target: no
    ``target_splits`` / ``template_splits``): the templates' setups run
    first, then every round scans one rotation batch of each template as one
    replayed hipGraph (``backend.fused_group``); the results equal the
    sequential schedule's.  ``pad_target_edges``, ``pad_template_filter`` and
    ``interpolation_order`` are honoured there, other ``scan_kwargs`` raise
    ``ValueError``.
    """
    from .analyzer import MaxScoreOverRotations
    from .matching_data import MatchingData

    if score not in MATCHING_EXHAUSTIVE_REGISTER:
        raise ValueError(f"score has to be one of {', '.join(MATCHING_EXHAUSTIVE_REGISTER.keys())}")
    setup, scoring = MATCHING_EXHAUSTIVE_REGISTER[score]
    callback_class = MaxScoreOverRotations if callback_class is None else callback_class
    callback_class_args = {} if callback_class_args is None else callback_class_args
    if schedule == "graph":
        return _scan_templates_graph(target, templates, setup, scoring, callback_class, callback_class_args,
                                     **scan_kwargs)
    if schedule != "sequential":
        raise ValueError(f"schedule has to be 'sequential' or 'graph', got {schedule!r}")
    results = []
    for entry in templates:
        if "template" not in entry:
            raise ValueError("every template entry needs a 'template'")
        md = MatchingData(target, entry["template"], template_mask=entry.get("template_mask"),
                          rotations=entry.get("rotations"))
        if entry.get("template_filter") is not None:
            md.template_filter = entry["template_filter"]
        if entry.get("target_filter") is not None:
            md.target_filter = entry["target_filter"]
        results.append(
            scan_subsets(matching_data=md, matching_setup=setup, matching_score=scoring,
                         callback_class=callback_class, callback_class_args=callback_class_args, **scan_kwargs)
        )
    return results


def _template_data(target, entry):
    from .matching_data import MatchingData

    if "template" not in entry:
        raise ValueError("every template entry needs a 'template'")
    md = MatchingData(target, entry["template"], template_mask=entry.get("template_mask"),
                      rotations=entry.get("rotations"))
    if entry.get("template_filter") is not None:
        md.template_filter = entry["template_filter"]
    if entry.get("target_filter") is not None:
        md.target_filter = entry["target_filter"]
    return md


def _scan_templates_graph(target, templates, setup_fn, scoring, callback_class, callback_class_args,
                          pad_target_edges: bool = False, pad_template_filter: bool = True,
                          interpolation_order: int = None, **unsupported):
    """``scan_templates(schedule="graph")``: per template the setup half of
    :func:`scan` (one split, one rotation job), then all templates' scans
    through ``backend.fused_group``, then each analyzer's result merged as
    :func:`scan` and :func:`scan_subsets` merge a single split."""
    bad = sorted(k for k, v in unsupported.items() if v not in (None, {}, (1, 1), False))
    if bad:
        raise ValueError(f"schedule='graph' does not take {', '.join(bad)}")
    if not hasattr(be, "fused_group") or not (getattr(be, "fused", False) and getattr(be, "_gpu", False)):
        raise ValueError("schedule='graph' needs the hip backend with the fused scan on a GPU device")
    jobs, callback_args = [], []
    for entry in templates:
        md = _template_data(target, entry)
        setup, default_callback_args, rotation_chunks = prepare_scan(
            md, setup_fn, n_jobs=1, callback_class_args=callback_class_args, pad_target=pad_target_edges,
            pad_template_filter=pad_template_filter, interpolation_order=interpolation_order,
        )
        jobs.append((setup, rotation_chunks[0], SharedAnalyzerProxy(callback_class, default_callback_args)))
        callback_args.append(default_callback_args)
    be.fused_group(scoring, jobs)
    results = []
    for (_, _, proxy), dca in zip(jobs, callback_args):
        one = callback_class.merge([proxy.result(**dca)], **dca)
        results.append(callback_class.merge([one], **callback_class_args))
    be.free_cache()
    return results


def register_matching_exhaustive(
    matching: str,
    matching_setup: Callable,
    matching_scoring: Callable,
    memory_class: MatchingMemoryUsage,
) -> None:
    """Register a new score.

    Raises ``ValueError`` for a duplicate name or a ``memory_class`` that is
    not a :class:`MatchingMemoryUsage` subclass (:468-501).
    """
    if matching in MATCHING_EXHAUSTIVE_REGISTER:
        raise ValueError(f"A method with name '{matching}' is already registered.")
    if not (isinstance(memory_class, type) and issubclass(memory_class, MatchingMemoryUsage)):
        raise ValueError(f"{memory_class} is not a subclass of {MatchingMemoryUsage}.")
    MATCHING_EXHAUSTIVE_REGISTER[matching] = (matching_setup, matching_scoring)
    MATCHING_MEMORY_REGISTRY[matching] = memory_class
