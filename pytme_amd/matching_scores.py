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
from .matching_utils import _normalize_template_overflow_safe, normalize_template

__all__ = [
    "cc_setup",
    "lcc_setup",
    "flc_setup",
    "flcSphericalMask_setup",
    "corr_scoring",
    "flc_scoring",
    "MATCHING_EXHAUSTIVE_REGISTER",
]


def _shape(x) -> Tuple[int, ...]:
    return tuple(int(v) for v in x)


def _half_spectrum(shape) -> Tuple[int, ...]:
    shape = _shape(shape)
    return shape[:-1] + (shape[-1] // 2 + 1,)


class _TemplateFilter:
    """Fourier filter of the rotated template (tme/matching_scores.py:44-150).

    The filter array is a real half spectrum.  Its shape picks the box it
    acts on: the template's own half spectrum filters the template-sized
    top-left box of the array (``pad_template_filter=False``, the CLI
    default), the half spectrum of the array filters the whole array.  A
    filter of size 1 is the identity.  The box is transformed, multiplied
    and transformed back in place: out = irfftn(rfftn(box) * filter).
    """

    def __init__(self, arr_filter, template_shape, arr_shape, real_dtype, cmpl_dtype, plans=None):
        self.identity = be.size(arr_filter) == 1
        if self.identity:
            return
        fshape = _shape(arr_filter.shape)
        if fshape == _half_spectrum(arr_shape):
            box = _shape(arr_shape)
        elif fshape == _half_spectrum(template_shape):
            box = _shape(template_shape)
        else:
            raise NotImplementedError(
                f"template filter of shape {fshape} fits neither the template {_shape(template_shape)} "
                f"nor the array {_shape(arr_shape)} (batch-broadcast filters are not supported)"
            )
        self.whole = box == _shape(arr_shape)
        self.box = tuple(slice(0, n) for n in box)
        if plans is not None and self.whole:
            self.fwd, self.inv = plans
        else:
            self.fwd, self.inv = be.build_fft(fwd_shape=box, inv_shape=fshape, real_dtype=real_dtype,
                                              cmpl_dtype=cmpl_dtype, inv_output_shape=box)
        self.real = None if self.whole else be.zeros(box, real_dtype)
        self.spec = be.zeros(fshape, cmpl_dtype)

    def __call__(self, arr, arr_filter):
        if self.identity:
            return arr
        src = arr if self.whole else self.real
        if not self.whole:
            src[...] = arr[self.box]
        spec = be.multiply(self.fwd(src, self.spec), arr_filter, out=self.spec)
        src = self.inv(spec, src)
        if self.whole:
            return src
        arr[self.box] = src
        return arr


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
    """``f(arr, ft_temp, filter) -> arr`` applying the template filter
    (signature of tme/matching_scores.py:44-57).  ``arr_shape`` is the
    template's shape; the returned function filters the array it is given,
    either whole (filter shaped like ``arr_ft_shape``, using the given
    transforms when they fit ``fwd_shape``) or its template-sized top-left
    box (filter shaped like the template's half spectrum)."""
    full = tuple(int(x) for x in fwd_shape)
    if _shape(arr_filter.shape) != _shape(arr_ft_shape) and be.size(arr_filter) != 1:
        full = _shape(inv_output_shape)
    plans = (rfftn, irfftn) if rfftn is not None and irfftn is not None else None
    flt = _TemplateFilter(arr_filter, arr_shape, full, real_dtype, cmpl_dtype, plans=plans)
    return lambda arr, ft_temp, template_filter: flt(arr, template_filter)


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


class _PlainWorkspace:
    """Buffers and transforms of the per-rotation loop on ``fast_shape``.

    ``forward(real)`` leaves the half spectrum of a zero-padded array in the
    workspace's spectrum buffer; ``correlate(spec, target_spec, out)`` is
    the circular correlation irfftn(spec * target_spec) written to ``out``
    (the template and mask are stored reversed, so a plain product of
    spectra correlates, tme/matching_data.py:698-715).  The inverse
    transform consumes a separate product buffer, so one template spectrum
    can be correlated with several target spectra.
    """

    def __init__(self, fast_shape, fast_ft_shape):
        self.real_dtype, self.cmpl_dtype = be._float_dtype, be._complex_dtype
        self.fast_shape = _shape(fast_shape)
        self.arr = be.zeros(self.fast_shape, self.real_dtype)
        self.spec = be.zeros(_shape(fast_ft_shape), self.cmpl_dtype)
        self.prod = be.zeros(_shape(fast_ft_shape), self.cmpl_dtype)
        self.fwd, self.inv = be.build_fft(
            fwd_shape=self.fast_shape, inv_shape=_shape(fast_ft_shape), real_dtype=self.real_dtype,
            cmpl_dtype=self.cmpl_dtype, inv_output_shape=self.fast_shape, temp_fwd=self.arr, temp_inv=self.spec,
        )

    def volume(self):
        return be.zeros(self.fast_shape, self.real_dtype)

    def template_filter(self, arr_filter, template_shape):
        return _TemplateFilter(arr_filter, template_shape, self.fast_shape, self.real_dtype, self.cmpl_dtype,
                               plans=(self.fwd, self.inv))

    def forward(self, real):
        self.spec = self.fwd(real, self.spec)
        return self.spec

    def correlate(self, spec, target_spec, out):
        self.prod = be.multiply(target_spec, spec, out=self.prod)
        return self.inv(self.prod, out)


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
    (tme/matching_scores.py:491-626).  ``numerator`` / ``inv_denominator``
    take part only when they are ``fast_shape`` volumes (the setups pass
    size-1 placeholders otherwise).
    """
    ret = _fused(
        "corr", template=template, template_filter=template_filter, ft_target=ft_target,
        inv_denominator=inv_denominator, numerator=numerator, fast_shape=fast_shape,
        fast_ft_shape=fast_ft_shape, rotations=rotations, callback=callback,
        interpolation_order=interpolation_order, template_mask=template_mask,
    )
    if ret is not NotImplemented:
        return ret

    template, ft_target, template_filter = (be.from_sharedarr(x) for x in (template, ft_target, template_filter))
    numerator, inv_denominator = be.from_sharedarr(numerator), be.from_sharedarr(inv_denominator)
    work = _PlainWorkspace(fast_shape, fast_ft_shape)
    filt = work.template_filter(template_filter, template.shape)
    offset = numerator if _shape(numerator.shape) == work.fast_shape else None
    scale = inv_denominator if _shape(inv_denominator.shape) == work.fast_shape else None

    normalise = None
    if template_mask is not None:
        template_mask = be.from_sharedarr(template_mask)
        if be.datatype_bytes(template_mask.dtype) == 2:  # 2-byte masks: sums in the overflow-safe type
            n_obs = be.sum(be.astype(template_mask, be._overflow_safe_dtype))
            normalise = lambda t: _normalize_template_overflow_safe(t, template_mask, n_obs)  # noqa: E731
        else:
            n_obs = be.sum(template_mask)
            normalise = lambda t: normalize_template(t, template_mask, n_obs)  # noqa: E731

    centre = [(int(x) - 1) / 2 for x in template.shape]
    box = tuple(slice(0, int(x)) for x in template.shape)
    rotated = be.zeros(template.shape, be._float_dtype)
    scores = work.arr
    for rotation in rotations:
        pose = be._rigid_transform_matrix(rotation_matrix=rotation, center=centre)
        rotated, _ = be.rigid_transform(arr=template, rotation_matrix=pose, out=rotated, order=interpolation_order,
                                        cache=True)
        rotated = filt(rotated, template_filter)
        if normalise is not None:
            normalise(rotated)
        scores = be.fill(scores, 0)
        scores[box] = rotated
        scores = work.correlate(work.forward(scores), ft_target, out=scores)
        if offset is not None:
            scores = be.subtract(scores, offset, out=scores)
        if scale is not None:
            scores = be.multiply(scores, scale, out=scores)
        if callback is not None:
            callback(scores, rotation_matrix=rotation)
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

    template, template_mask, template_filter = (
        be.from_sharedarr(x) for x in (template, template_mask, template_filter)
    )
    ft_target, ft_target2 = be.from_sharedarr(ft_target), be.from_sharedarr(ft_target2)
    work = _PlainWorkspace(fast_shape, fast_ft_shape)
    filt = work.template_filter(template_filter, template.shape)
    t_rot, m_rot, local_sq = work.arr, work.volume(), work.volume()
    eps = be.eps(be._float_dtype)
    for rotation in rotations:
        t_rot, m_rot = be.fill(t_rot, 0), be.fill(m_rot, 0)
        t_rot, m_rot = be.rigid_transform(
            arr=template, arr_mask=template_mask, rotation_matrix=rotation, out=t_rot, out_mask=m_rot,
            use_geometric_center=True, order=interpolation_order, cache=True,
        )
        n_obs = be.sum(m_rot)
        t_rot = normalize_template(filt(t_rot, template_filter), m_rot, n_obs, axis=None)
        # local sums of f and f^2 under the rotated mask, then the overlap
        mask_spec = work.forward(m_rot)
        local_sq = work.correlate(mask_spec, ft_target2, out=local_sq)
        m_rot = work.correlate(mask_spec, ft_target, out=m_rot)
        t_rot = work.correlate(work.forward(t_rot), ft_target, out=t_rot)
        t_rot = be.norm_scores(t_rot, local_sq, m_rot, n_obs, eps, t_rot)
        if callback is not None:
            callback(t_rot, rotation_matrix=rotation)
    return callback


MATCHING_EXHAUSTIVE_REGISTER = {
    "CC": (cc_setup, corr_scoring),
    "LCC": (lcc_setup, corr_scoring),
    "FLCSphericalMask": (flcSphericalMask_setup, corr_scoring),
    "FLC": (flc_setup, flc_scoring),
}
