"""Setup/scoring pairs of the exhaustive scan (mirrors ``tme/matching_scores.py``).

Each score is a ``setup`` that precomputes target-side quantities once and a
``scoring`` loop that, per rotation, rotates the (reversed) template, filters
and normalises it, correlates it with the target through the FFT and hands
the score volume to the analyzer callback.

Registered scores: ``CC``, ``LCC``, ``FLCSphericalMask`` (setup + shared
``corr_scoring``) and ``FLC`` (``flc_setup`` + ``flc_scoring``).  On a
backend that provides ``fused_scoring`` (the hip backend) the scoring
functions first offer the whole rotation set to the batched fused scan; the
per-rotation loop below runs when the fused scan declines (other analyzers,
``fused=False``).
"""

from typing import Callable, Dict, Tuple

import numpy as np
from scipy.ndimage import laplace

from .backends import backend as be
from .matching_utils import (
    _normalize_template_overflow_safe,
    conditional_execute,
    identity,
    normalize_template,
)

__all__ = [
    "cc_setup",
    "lcc_setup",
    "flc_setup",
    "flcSphericalMask_setup",
    "corr_scoring",
    "flc_scoring",
    "MATCHING_EXHAUSTIVE_REGISTER",
]


def _shape_match(shape1: Tuple[int], shape2: Tuple[int]) -> bool:
    return len(shape1) == len(shape2) and tuple(int(x) for x in shape1) == tuple(int(x) for x in shape2)


def _create_filter_func(
    fwd_shape: Tuple[int],
    inv_shape: Tuple[int],
    arr_shape: Tuple[int],
    arr_filter,
    arr_ft_shape: Tuple[int],
    inv_output_shape: Tuple[int],
    real_dtype,
    cmpl_dtype,
    fwd_axes=None,
    inv_axes=None,
    rfftn: Callable = None,
    irfftn: Callable = None,
) -> Callable:
    """Build ``f(arr, ft_temp, filter) -> arr`` applying a Fourier filter.

    A filter of size 1 is the identity.  A filter shaped like the template's
    half spectrum is applied at template size to the top-left template box
    of ``arr`` (reference default ``pad_template_filter=False`` path,
    tme/matching_scores.py:44-150).
    """
    if be.size(arr_filter) == 1:
        return conditional_execute(identity, identity, False)

    filter_shape = tuple(int(x) for x in arr_filter.shape)
    try:
        product_ft_shape = tuple(int(x) for x in np.broadcast_shapes(tuple(arr_ft_shape), filter_shape))
    except ValueError:
        product_ft_shape, inv_output_shape = filter_shape, tuple(arr_shape)

    valid = _shape_match(arr_shape, fwd_shape) and _shape_match(product_ft_shape, inv_shape)
    valid = valid and rfftn is not None and irfftn is not None
    if not valid:
        _fwd_shape = tuple(arr_shape)
        if all(x > y for x, y in zip(arr_shape, product_ft_shape)):
            _fwd_shape = tuple(fwd_shape)
        rfftn, irfftn = be.build_fft(
            fwd_shape=_fwd_shape,
            inv_shape=product_ft_shape,
            real_dtype=real_dtype,
            cmpl_dtype=cmpl_dtype,
            inv_output_shape=inv_output_shape,
            fwd_axes=fwd_axes,
            inv_axes=inv_axes,
        )

    def _apply_filter(template, ft_temp, template_filter):
        ft_temp = rfftn(template, ft_temp)
        ft_temp = be.multiply(ft_temp, template_filter, out=ft_temp)
        return irfftn(ft_temp, template)

    if not _shape_match(arr_ft_shape, filter_shape):
        if any(x > y and y == 1 for x, y in zip(filter_shape, arr_ft_shape)):
            raise NotImplementedError("filters broadcast over a batch dimension are not supported")
        real_subset = tuple(slice(0, int(x)) for x in arr_shape)
        _template = be.zeros(tuple(arr_shape), be._float_dtype)
        _ft_temp = be.zeros(product_ft_shape, be._complex_dtype)

        def _apply_filter_subset(template, ft_temp, template_filter):
            _template[:] = template[real_subset]
            template[real_subset] = _apply_filter(_template, _ft_temp, template_filter)
            return template

        return _apply_filter_subset

    return _apply_filter


# --------------------------------------------------------------------- setups
def cc_setup(matching_data, fast_shape, fast_ft_shape, shm_handler=None, **kwargs) -> Dict:
    """Unnormalised cross-correlation: only the target spectrum is needed."""
    target_pad = be.topleft_pad(matching_data.target, fast_shape)
    return {
        "fast_shape": fast_shape,
        "fast_ft_shape": fast_ft_shape,
        "template": be.to_sharedarr(matching_data.template, shm_handler),
        "ft_target": be.to_sharedarr(be.rfftn(target_pad), shm_handler),
        "inv_denominator": be.to_sharedarr(be.zeros(1, be._float_dtype) + 1, shm_handler),
        "numerator": be.to_sharedarr(be.zeros(1, be._float_dtype), shm_handler),
    }


def lcc_setup(matching_data, **kwargs) -> Dict:
    """CC of the periodic Laplacians of target and template (host Laplace)."""
    target = np.asarray(be.to_numpy_array(matching_data._target), dtype=np.float32)
    template = np.asarray(be.to_numpy_array(matching_data._template), dtype=np.float32)
    matching_data._target = be.to_backend_array(laplace(target, mode="wrap"))
    matching_data._template = be.to_backend_array(laplace(template, mode="wrap"))
    return cc_setup(matching_data=matching_data, **kwargs)


def flc_setup(matching_data, fast_shape, fast_ft_shape, shm_handler=None, **kwargs) -> Dict:
    """FLC: spectra of the target and of the squared target."""
    target_pad = be.topleft_pad(matching_data.target, fast_shape)
    ft_target = be.rfftn(target_pad)
    target_pad = be.square(target_pad, out=target_pad)
    ft_target2 = be.rfftn(target_pad)
    return {
        "fast_shape": fast_shape,
        "fast_ft_shape": fast_ft_shape,
        "template": be.to_sharedarr(matching_data.template, shm_handler),
        "template_mask": be.to_sharedarr(matching_data.template_mask, shm_handler),
        "ft_target": be.to_sharedarr(ft_target, shm_handler),
        "ft_target2": be.to_sharedarr(ft_target2, shm_handler),
    }


def flcSphericalMask_setup(matching_data, fast_shape, fast_ft_shape, shm_handler=None, **kwargs) -> Dict:
    """FLC with a rotation-invariant mask: the denominator is precomputed.

    n = sum(m);  inv_denominator = 1 / (n * sigma_local) with
    sigma_local^2 = (f^2 * m)/n - ((f * m)/n)^2 from two FFT correlations of
    the unrotated mask (tme/matching_scores.py:390-450).
    """
    template_mask = matching_data.template_mask
    n_obs = be.sum(be.astype(template_mask, be._overflow_safe_dtype))
    target_pad = be.topleft_pad(matching_data.target, fast_shape)
    mask_pad = be.topleft_pad(template_mask, fast_shape)
    ft_mask = be.rfftn(mask_pad)
    mask_pad = None

    ft_temp = be.zeros(fast_ft_shape, be._complex_dtype)
    ft_target = be.rfftn(be.square(target_pad))
    ft_temp = be.multiply(ft_target, ft_mask, out=ft_temp)
    temp2 = be.irfftn(ft_temp, s=fast_shape)

    ft_target = be.rfftn(target_pad)
    ft_temp = be.multiply(ft_target, ft_mask, out=ft_temp)
    temp = be.irfftn(ft_temp, s=fast_shape)

    temp2 = be.norm_scores(1, temp2, temp, n_obs, be.eps(be._float_dtype), temp2)
    return {
        "fast_shape": fast_shape,
        "fast_ft_shape": fast_ft_shape,
        "template": be.to_sharedarr(matching_data.template, shm_handler),
        "template_mask": be.to_sharedarr(template_mask, shm_handler),
        "ft_target": be.to_sharedarr(ft_target, shm_handler),
        "inv_denominator": be.to_sharedarr(temp2, shm_handler),
        "numerator": be.to_sharedarr(be.zeros(1, be._float_dtype), shm_handler),
    }


# -------------------------------------------------------------------- scoring
def _fused(score: str, **kwargs):
    hook = getattr(be, "fused_scoring", None)
    if hook is None:
        return NotImplemented
    return hook(score=score, **kwargs)


def corr_scoring(
    template,
    template_filter,
    ft_target,
    inv_denominator,
    numerator,
    fast_shape: Tuple[int],
    fast_ft_shape: Tuple[int],
    rotations,
    callback,
    interpolation_order: int,
    template_mask=None,
):
    """(CC(f, t_R) - numerator) * inv_denominator for every rotation R.

    t_R is the rotated (about the template centre), filtered and - when a
    mask is given - mask-normalised template; CC is the circular correlation
    on ``fast_shape`` computed as irfftn(rfftn(pad(t_R)) * ft_target)
    (tme/matching_scores.py:491-626).
    """
    ret = _fused(
        "corr", template=template, template_filter=template_filter, ft_target=ft_target,
        inv_denominator=inv_denominator, numerator=numerator, fast_shape=fast_shape,
        fast_ft_shape=fast_ft_shape, rotations=rotations, callback=callback,
        interpolation_order=interpolation_order, template_mask=template_mask,
    )
    if ret is not NotImplemented:
        return ret

    template = be.from_sharedarr(template)
    ft_target = be.from_sharedarr(ft_target)
    inv_denominator = be.from_sharedarr(inv_denominator)
    numerator = be.from_sharedarr(numerator)
    template_filter = be.from_sharedarr(template_filter)

    norm_func, norm_template, mask_sum = normalize_template, False, 1
    if template_mask is not None:
        template_mask = be.from_sharedarr(template_mask)
        norm_template, mask_sum = True, be.sum(template_mask)
        if be.datatype_bytes(template_mask.dtype) == 2:
            norm_func = _normalize_template_overflow_safe
            mask_sum = be.sum(be.astype(template_mask, be._overflow_safe_dtype))

    callback_func = conditional_execute(callback, callback is not None)
    norm_template = conditional_execute(norm_func, norm_template)
    norm_numerator = conditional_execute(be.subtract, identity, _shape_match(numerator.shape, fast_shape))
    norm_denominator = conditional_execute(be.multiply, identity, _shape_match(inv_denominator.shape, fast_shape))

    arr = be.zeros(fast_shape, be._float_dtype)
    ft_temp = be.zeros(fast_ft_shape, be._complex_dtype)
    fftargs = {
        "real_dtype": be._float_dtype,
        "cmpl_dtype": be._complex_dtype,
        "inv_output_shape": fast_shape,
        "fwd_axes": None,
        "inv_axes": None,
        "inv_shape": fast_ft_shape,
        "fwd_shape": tuple(arr.shape),
    }
    rfftn, irfftn = be.build_fft(temp_fwd=arr, temp_inv=ft_temp, **fftargs)
    template_filter_func = _create_filter_func(
        arr_shape=template.shape,
        arr_ft_shape=fast_ft_shape,
        arr_filter=template_filter,
        rfftn=rfftn,
        irfftn=irfftn,
        **fftargs,
    )

    center = [(int(x) - 1) / 2 for x in template.shape]
    unpadded = tuple(slice(0, int(x)) for x in template.shape)
    template_rot = be.zeros(template.shape, be._float_dtype)
    for index in range(rotations.shape[0]):
        rotation = rotations[index]
        matrix = be._rigid_transform_matrix(rotation_matrix=rotation, center=center)
        template_rot, _ = be.rigid_transform(
            arr=template, rotation_matrix=matrix, out=template_rot, order=interpolation_order, cache=True
        )
        template_rot = template_filter_func(template_rot, ft_temp, template_filter)
        norm_template(template_rot, template_mask, mask_sum)

        arr = be.fill(arr, 0)
        arr[unpadded] = template_rot
        ft_temp = rfftn(arr, ft_temp)
        ft_temp = be.multiply(ft_target, ft_temp, out=ft_temp)
        arr = irfftn(ft_temp, arr)

        arr = norm_numerator(arr, numerator, out=arr)
        arr = norm_denominator(arr, inv_denominator, out=arr)
        callback_func(arr, rotation_matrix=rotation)
    return callback


def flc_scoring(
    template,
    template_mask,
    ft_target,
    ft_target2,
    template_filter,
    fast_shape: Tuple[int],
    fast_ft_shape: Tuple[int],
    rotations,
    callback,
    interpolation_order: int,
):
    """Fast local correlation with a rotated mask (Roseman / Hrabe FLC).

    Per rotation: rotate template and mask into zeroed ``fast_shape`` arrays,
    n = sum(m_R), filter the template, normalise it inside m_R, then
      S1 = CC(f, m_R),  S2 = CC(f^2, m_R),  X = CC(f, t_R)
      score = X / (n * sqrt(max(S2/n - (S1/n)^2, 0)))   (sigma < eps -> 1)
    (tme/matching_scores.py:629-759).
    """
    ret = _fused(
        "flc", template=template, template_mask=template_mask, ft_target=ft_target, ft_target2=ft_target2,
        template_filter=template_filter, fast_shape=fast_shape, fast_ft_shape=fast_ft_shape,
        rotations=rotations, callback=callback, interpolation_order=interpolation_order,
    )
    if ret is not NotImplemented:
        return ret

    float_dtype, complex_dtype = be._float_dtype, be._complex_dtype
    template = be.from_sharedarr(template)
    template_mask = be.from_sharedarr(template_mask)
    ft_target = be.from_sharedarr(ft_target)
    ft_target2 = be.from_sharedarr(ft_target2)
    template_filter = be.from_sharedarr(template_filter)

    arr = be.zeros(fast_shape, float_dtype)
    temp = be.zeros(fast_shape, float_dtype)
    temp2 = be.zeros(fast_shape, float_dtype)
    ft_temp = be.zeros(fast_ft_shape, complex_dtype)
    ft_denom = be.zeros(fast_ft_shape, complex_dtype)

    fftargs = {
        "real_dtype": float_dtype,
        "cmpl_dtype": complex_dtype,
        "inv_output_shape": fast_shape,
        "fwd_axes": None,
        "inv_axes": None,
        "inv_shape": fast_ft_shape,
        "fwd_shape": tuple(arr.shape),
    }
    rfftn, irfftn = be.build_fft(temp_fwd=arr, temp_inv=ft_temp, **fftargs)
    template_filter_func = _create_filter_func(
        arr_shape=template.shape,
        arr_ft_shape=fast_ft_shape,
        arr_filter=template_filter,
        rfftn=rfftn,
        irfftn=irfftn,
        **fftargs,
    )

    eps = be.eps(float_dtype)
    callback_func = conditional_execute(callback, callback is not None)
    for index in range(rotations.shape[0]):
        rotation = rotations[index]
        arr = be.fill(arr, 0)
        temp = be.fill(temp, 0)
        arr, temp = be.rigid_transform(
            arr=template,
            arr_mask=template_mask,
            rotation_matrix=rotation,
            out=arr,
            out_mask=temp,
            use_geometric_center=True,
            order=interpolation_order,
            cache=True,
        )
        n_obs = be.sum(temp)
        arr = template_filter_func(arr, ft_temp, template_filter)
        arr = normalize_template(arr, temp, n_obs, axis=None)

        ft_temp = rfftn(temp, ft_temp)
        ft_denom = be.multiply(ft_target, ft_temp, out=ft_denom)
        temp = irfftn(ft_denom, temp)
        ft_denom = be.multiply(ft_target2, ft_temp, out=ft_denom)
        temp2 = irfftn(ft_denom, temp2)

        ft_temp = rfftn(arr, ft_temp)
        ft_temp = be.multiply(ft_target, ft_temp, out=ft_temp)
        arr = irfftn(ft_temp, arr)

        arr = be.norm_scores(arr, temp2, temp, n_obs, eps, arr)
        callback_func(arr, rotation_matrix=rotation)
    return callback


MATCHING_EXHAUSTIVE_REGISTER = {
    "CC": (cc_setup, corr_scoring),
    "LCC": (lcc_setup, corr_scoring),
    "FLCSphericalMask": (flcSphericalMask_setup, corr_scoring),
    "FLC": (flc_setup, flc_scoring),
}
