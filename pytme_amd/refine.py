"""Local FLC refinement of candidate peaks on the GPU, with gradients.

The objective is the reference's derivative-free refinement score
``matching_optimization.FLC`` (tme/matching_optimization.py:488-548) at a
pose x = (translation(3) voxels, zyz Euler angles(3)), with the pose handling
of ``_MatchDensityToDensity.score`` (:198-238): the template corner sits at
v = trunc(t), the template is sampled at R^T (o - c) + c + (t - v),
c = floor(T / 2), with order-1 constant-mode interpolation, standardised
inside the rotated mask and correlated with the target under the box.  The
rotated mask goes into the mask buffer (the reference writes it into the
template buffer, :162-163, which zeroes n).  The reference's sub-voxel
shift is applied as R^T (o - c) + c + s, which moves the template by -s
while the integer part moves it by +v (the template sits at v - s and jumps
two voxels at every integer of t); ``subvoxel="rigid"`` (the default) places
the rotated template centre at t + c instead, R^T (o - c - s) + c, which is
continuous in t, and ``subvoxel="reference"`` keeps the reference's sampling.

``tm_flc_local_fwd`` evaluates the six moments of the score for a batch of
poses, ``tm_flc_local_bwd`` their contraction with dScore/d(moment) into
dScore/d(R^T) and dScore/ds; ``LocalMoments`` wraps both as a
``torch.autograd.Function`` so the score, the Euler (or any other)
parametrisation of R and the optimiser are ordinary torch code in FP64 on a
few numbers per pose.  ``refine_peaks`` runs batched Adam over all peaks.
"""

import ctypes
import math
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib

__all__ = [
    "LocalFLC",
    "LocalMoments",
    "euler_zyz_matrix",
    "score_from_moments",
    "refine_peaks",
]

EPS32 = float(np.finfo(np.float32).eps)


def _torch():
    import torch

    return torch


def standardise(template: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """tme/matching_optimization.py:515-520 (normalize_template with the
    unrotated mask and n = sum(mask), in FP64)."""
    t = np.asarray(template, dtype=np.float64)
    m = np.asarray(mask, dtype=np.float64)
    n = m.sum()
    if n <= 0:
        raise ValueError("the template mask is empty")
    mean = np.sum(t * m) / n
    sd = math.sqrt(max(np.sum(t * t * m) / n - mean * mean, 0.0))
    if sd == 0.0:
        raise ValueError("the template is constant inside its mask")
    return (t - mean) / sd * m


class LocalFLC:
    """A target volume and one template/mask pair on the GPU, ready to
    evaluate the local FLC moments (and their gradient) of pose batches.

    ``target`` is a 3-D float32 array (numpy or a device tensor); the
    template is standardised once inside its mask as the reference's
    ``_post_init`` does.
    """

    def __init__(self, target, template, template_mask=None, max_poses: int = 4096, device="cuda",
                 one_pass_gradient: bool = False):
        torch = _torch()
        template = np.asarray(template, dtype=np.float32)
        if template.ndim != 3:
            raise ValueError("refinement is three-dimensional")
        mask = np.ones_like(template) if template_mask is None else np.asarray(template_mask, dtype=np.float32)
        if mask.shape != template.shape:
            raise ValueError("template and mask shapes differ")
        self.device = torch.device(device)
        if isinstance(target, torch.Tensor):
            self.target = target.to(self.device, torch.float32).contiguous()
        else:
            self.target = torch.as_tensor(np.ascontiguousarray(target, dtype=np.float32), device=self.device)
        if self.target.ndim != 3:
            raise ValueError("the target must be a 3-D volume")
        self.template_shape = tuple(int(x) for x in template.shape)
        self.max_poses = int(max_poses)
        ts = torch.as_tensor(standardise(template, mask).astype(np.float32), device=self.device)
        ms = torch.as_tensor(mask, device=self.device)
        handle = ctypes.c_void_p()
        _lib.call("tm_flc_local_plan_create", ts.data_ptr(), ms.data_ptr(), _lib.i32_array(self.template_shape),
                  self.max_poses, ctypes.byref(handle))
        self._plan = handle.value
        self._tshape = _lib.i64_array(self.target.shape)
        self.evaluations = 0  # poses scored (forward kernel)
        self.gradients = 0  # poses differentiated (backward kernel)
        #: gradients from the backward kernel after the moments kernel
        #: (default: it recomputes the interpolation weights of every
        #: sample; 1.71 + 1.77 ms for 4096 poses of a 48^3 template) or from
        #: the one-pass moments-and-Jacobian kernel (78 sums per thread,
        #: 2 waves/SIMD, 4.02 ms)
        self.one_pass_gradient = bool(one_pass_gradient)

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan:
            try:
                _lib.call("tm_flc_local_plan_destroy", plan)
            except Exception:
                pass
            self._plan = None

    def _pack(self, Rt, s, v):
        torch = _torch()
        n = Rt.shape[0]
        if n < 1 or n > self.max_poses:
            raise ValueError(f"{n} poses outside [1, {self.max_poses}]")
        poses = torch.cat([Rt.reshape(n, 9), s.reshape(n, 3)], 1).to(self.device, torch.float32).contiguous()
        corners = torch.as_tensor(v, dtype=torch.int32, device=self.device).reshape(n, 3).contiguous()
        return poses, corners

    def moments(self, Rt, s, v):
        """(n, 6) float64 moments for R^T (n,3,3), sub-voxel shifts s (n,3)
        and integer corners v (n,3)."""
        torch = _torch()
        poses, corners = self._pack(Rt, s, v)
        out = torch.empty((poses.shape[0], 6), dtype=torch.float64, device=self.device)
        _lib.call("tm_flc_local_fwd", self._plan, self.target.data_ptr(), self._tshape, poses.data_ptr(),
                  corners.data_ptr(), poses.shape[0], out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        self.evaluations += poses.shape[0]
        return out

    def moments_jacobian(self, Rt, s, v):
        """(n, 6) float64 moments and (n, 6, 12) float64 Jacobian
        d(moment)/d(R^T row-major (9), s (3)) from one pass over the samples
        (``tm_flc_local_fwd_jac``)."""
        torch = _torch()
        poses, corners = self._pack(Rt, s, v)
        n = poses.shape[0]
        out = torch.empty((n, 78), dtype=torch.float64, device=self.device)
        _lib.call("tm_flc_local_fwd_jac", self._plan, self.target.data_ptr(), self._tshape, poses.data_ptr(),
                  corners.data_ptr(), n, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
        self.evaluations += n
        self.gradients += n
        return out[:, :6], out[:, 6:].reshape(n, 6, 12)

    def moment_grad(self, Rt, s, v, weights):
        """(n, 12) float64: dScore/d(R^T) (9, row-major) and dScore/ds (3)
        for weights = dScore/d(moment) (n, 6)."""
        torch = _torch()
        poses, corners = self._pack(Rt, s, v)
        w = weights.to(self.device, torch.float64).contiguous()
        out = torch.empty((poses.shape[0], 12), dtype=torch.float64, device=self.device)
        _lib.call("tm_flc_local_bwd", self._plan, self.target.data_ptr(), self._tshape, poses.data_ptr(),
                  corners.data_ptr(), poses.shape[0], w.data_ptr(), out.data_ptr(),
                  torch.cuda.current_stream().cuda_stream)
        self.gradients += poses.shape[0]
        return out


def _make_function():
    torch = _torch()

    class LocalMoments(torch.autograd.Function):
        """moments = LocalMoments.apply(Rt, s, v, local): forward is
        ``tm_flc_local_fwd``, backward ``tm_flc_local_bwd`` (the
        interpolation weights are recomputed, nothing per voxel is stored)."""

        @staticmethod
        def forward(ctx, Rt, s, v, local):
            # a gradient will be asked for: moments and their Jacobian in one
            # pass (the backward is then a 6-term contraction); otherwise the
            # moments alone
            ctx.shapes = (Rt.dtype, s.dtype)
            ctx.jac = False
            if (ctx.needs_input_grad[0] or ctx.needs_input_grad[1]) and local.one_pass_gradient:
                moments, jac = local.moments_jacobian(Rt.detach(), s.detach(), v)
                ctx.save_for_backward(jac)
                ctx.jac = True
                return moments
            ctx.save_for_backward(Rt, s)
            ctx.v, ctx.local = v, local
            return local.moments(Rt.detach(), s.detach(), v)

        @staticmethod
        def backward(ctx, grad_moments):
            if ctx.jac:
                (jac,) = ctx.saved_tensors
                g = torch.einsum("nk,nkj->nj", grad_moments.to(jac.dtype), jac)
            else:
                Rt, s = ctx.saved_tensors
                g = ctx.local.moment_grad(Rt.detach(), s.detach(), ctx.v, grad_moments)
            n = g.shape[0]
            rt_dtype, s_dtype = ctx.shapes
            return g[:, :9].reshape(n, 3, 3).to(rt_dtype), g[:, 9:].to(s_dtype), None, None

    return LocalMoments


_FUNCTION = None


def LocalMoments():
    """The torch.autograd.Function (built on first use so that importing the
    package does not import torch)."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _make_function()
    return _FUNCTION


def euler_zyz_matrix(angles_rad):
    """scipy ``Rotation.from_euler("zyz")`` (extrinsic, tme/rotations.py:171-196)
    as differentiable torch: R = Rz(c) Ry(b) Rz(a) for angles (..., 3)."""
    torch = _torch()
    a, b, c = angles_rad[..., 0], angles_rad[..., 1], angles_rad[..., 2]
    ca, sa, cb, sb, cc, sc = torch.cos(a), torch.sin(a), torch.cos(b), torch.sin(b), torch.cos(c), torch.sin(c)
    r00 = cc * cb * ca - sc * sa
    r01 = -cc * cb * sa - sc * ca
    r02 = cc * sb
    r10 = sc * cb * ca + cc * sa
    r11 = -sc * cb * sa + cc * ca
    r12 = sc * sb
    r20 = -sb * ca
    r21 = sb * sa
    r22 = cb
    return torch.stack([torch.stack([r00, r01, r02], -1), torch.stack([r10, r11, r12], -1),
                        torch.stack([r20, r21, r22], -1)], -2)


def score_from_moments(S, eps: float = EPS32):
    """tme/matching_optimization.py:522-548 in the six moments: the rotated
    template standardised in the rotated mask (normalize_template), then
    overlap / (sd_f * n); sd_f < eps gives 0."""
    torch = _torch()
    n, a, b, c, d, e = S.unbind(-1)
    mu = a / n
    sd_t = torch.sqrt(torch.clamp(b / n - mu * mu, min=0.0))
    overlap = (c - mu * d) / sd_t
    sd_f = torch.sqrt(torch.clamp(e / n - (d / n) ** 2, min=0.0))
    score = overlap / (sd_f * n)
    return torch.where(sd_f < eps, torch.zeros_like(score), score)


def split_translation(t):
    """v = trunc(t) (numpy astype(int), tme/matching_optimization.py:212), s = t - v."""
    torch = _torch()
    v = torch.trunc(t.detach())
    return v.to(torch.int32), t - v


def sampling_shift(Rt, s, subvoxel: str = "rigid"):
    """The shift the kernel adds to R^T (o - c) + c: -R^T s (rigid model) or
    s (the reference's sampling, tme/matching_optimization.py:150-157)."""
    if subvoxel == "rigid":
        return -(Rt @ s.unsqueeze(-1)).squeeze(-1)
    if subvoxel == "reference":
        return s
    raise ValueError("subvoxel is 'rigid' or 'reference'")


def pose_moments(local: "LocalFLC", t, a_rad, subvoxel: str = "rigid"):
    """Differentiable moments (n, 6) of poses t (n,3) voxels, a_rad (n,3)
    zyz radians."""
    v, s = split_translation(t)
    Rt = euler_zyz_matrix(a_rad).transpose(-1, -2)
    return LocalMoments().apply(Rt, sampling_shift(Rt, s, subvoxel), v, local)


def score_poses(local: LocalFLC, translations, angles_deg, subvoxel: str = "rigid"):
    """Scores (n,) float64 of poses (translations (n,3), zyz degrees (n,3))."""
    torch = _torch()
    t = torch.as_tensor(np.asarray(translations), dtype=torch.float64, device=local.device).reshape(-1, 3)
    a = torch.as_tensor(np.asarray(angles_deg), dtype=torch.float64, device=local.device).reshape(-1, 3)
    with torch.no_grad():
        out = []
        for lo in range(0, t.shape[0], local.max_poses):
            hi = min(lo + local.max_poses, t.shape[0])
            v, s = split_translation(t[lo:hi])
            Rt = euler_zyz_matrix(torch.deg2rad(a[lo:hi])).transpose(-1, -2)
            out.append(score_from_moments(local.moments(Rt, sampling_shift(Rt, s, subvoxel), v)))
        return torch.cat(out)


def _refine_fused(local, t0, a0, steps, lr_t, lr_a, bound, bsz, subvoxel):
    """The Adam loop of :func:`refine_peaks` on the device, so no torch work
    runs between the kernels.  Per batch and step: the pose-preparation
    kernel, then either (two-pass, the default) the moments kernel, one
    kernel for the score, best-pose tracking and the weights
    dScore/d(moment), the gradient kernel with those weights and one kernel
    for the chain rule to (t, angles) and the Adam update, or (one-pass) the
    moments-and-Jacobian kernel and one kernel for the score, its gradient,
    best-pose tracking and the Adam update (torch's formulas)."""
    torch = _torch()
    dev = local.device
    n = t0.shape[0]
    if subvoxel not in ("rigid", "reference"):
        raise ValueError("subvoxel is 'rigid' or 'reference'")
    rigid = 1 if subvoxel == "rigid" else 0
    t, a = t0.clone().contiguous(), a0.clone().contiguous()
    t0 = t0.contiguous()
    m1 = torch.zeros((n, 6), dtype=torch.float64, device=dev)
    m2 = torch.zeros_like(m1)
    best_s = torch.full((n,), -math.inf, dtype=torch.float64, device=dev)
    best_t, best_a = t0.clone(), a0.clone().contiguous()
    poses = torch.empty((bsz, 12), dtype=torch.float32, device=dev)
    corners = torch.empty((bsz, 3), dtype=torch.int32, device=dev)
    one_pass = local.one_pass_gradient
    mj = torch.empty((bsz, 78 if one_pass else 6), dtype=torch.float64, device=dev)
    w = torch.empty((bsz, 6), dtype=torch.float64, device=dev)
    g = torch.empty((bsz, 12), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    ptr = lambda x, lo: x.data_ptr() + lo * x.stride(0) * x.element_size()  # noqa: E731
    bnd = -1.0 if bound is None else float(bound)
    for step in range(steps + 1):
        update = int(step < steps)
        lt, la = float(lr_t), float(lr_a)
        for lo in range(0, n, bsz):
            k = min(bsz, n - lo)
            _lib.call("tm_refine_prepare", ptr(t, lo), ptr(a, lo), k, rigid, poses.data_ptr(), corners.data_ptr(),
                      stream)
            if one_pass:
                _lib.call("tm_flc_local_fwd_jac", local._plan, local.target.data_ptr(), local._tshape,
                          poses.data_ptr(), corners.data_ptr(), k, mj.data_ptr(), stream)
                _lib.call("tm_refine_adam", mj.data_ptr(), ptr(t, lo), ptr(a, lo), ptr(t0, lo), ptr(m1, lo),
                          ptr(m2, lo), ptr(best_s, lo), ptr(best_t, lo), ptr(best_a, lo), k, rigid, update,
                          lt, la, 0.9, 0.999, 1e-8, step + 1, bnd, EPS32, stream)
                continue
            _lib.call("tm_flc_local_fwd", local._plan, local.target.data_ptr(), local._tshape, poses.data_ptr(),
                      corners.data_ptr(), k, mj.data_ptr(), stream)
            _lib.call("tm_refine_weights", mj.data_ptr(), ptr(t, lo), ptr(a, lo), ptr(best_s, lo), ptr(best_t, lo),
                      ptr(best_a, lo), w.data_ptr(), k, update, EPS32, stream)
            if not update:
                continue
            _lib.call("tm_flc_local_bwd", local._plan, local.target.data_ptr(), local._tshape, poses.data_ptr(),
                      corners.data_ptr(), k, w.data_ptr(), g.data_ptr(), stream)
            _lib.call("tm_refine_adam_grad", g.data_ptr(), ptr(t, lo), ptr(a, lo), ptr(t0, lo), ptr(m1, lo),
                      ptr(m2, lo), k, rigid, lt, la, 0.9, 0.999, 1e-8, step + 1, bnd, stream)
        local.evaluations += n
        local.gradients += n * update
    return best_t, best_a, best_s


def _poses_on_device(x, dev):
    """(n, 3) float64 device tensor of host or device input (a device tensor
    stays on the device)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64).reshape(-1, 3)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev).reshape(-1, 3)


def refine_peaks(local: LocalFLC, translations, angles_deg, steps: int = 50, lr_translation: float = 0.1,
                 lr_angle_deg: float = 1.0, bounds_translation: Optional[float] = None,
                 batch: Optional[int] = None, subvoxel: str = "rigid",
                 fused: bool = True, return_device: bool = False) -> Tuple[np.ndarray, np.ndarray, np.ndarray, dict]:
    """Gradient ascent (Adam) on the local FLC of every pose at once.

    ``translations`` (n,3) are the template-corner positions the reference's
    caller passes (peak - T/2, scripts/postprocess.py:598-606) and
    ``angles_deg`` (n,3) zyz Euler angles.  Each step evaluates the moments
    and their gradient for all poses and keeps the best pose seen per peak.
    ``fused`` (default) runs the whole step on the device
    (``tm_refine_prepare``, then ``tm_flc_local_fwd``, ``tm_refine_weights``,
    ``tm_flc_local_bwd`` and ``tm_refine_adam_grad``, or with the one-pass
    gradient ``tm_flc_local_fwd_jac`` and ``tm_refine_adam``); otherwise the
    step goes through the autograd function and ``torch.optim.Adam``.
    Returns the best translations, angles (degrees), scores and evaluation
    counts; host arrays, or device tensors with ``return_device`` (the
    poses may be passed as device tensors too: nothing crosses to the host).
    """
    torch = _torch()
    dev = local.device
    t0 = _poses_on_device(translations, dev)
    a0 = torch.deg2rad(_poses_on_device(angles_deg, dev))
    out = (lambda x: x) if return_device else (lambda x: x.cpu().numpy())  # noqa: E731
    n = t0.shape[0]
    bsz = min(batch or local.max_poses, local.max_poses)
    if fused:
        bt, ba, bs = _refine_fused(local, t0, a0, steps, lr_translation, math.radians(lr_angle_deg),
                                   bounds_translation, bsz, subvoxel)
        info = {"score_evaluations_per_peak": steps + 1, "gradient_evaluations_per_peak": steps}
        return out(bt), out(torch.rad2deg(ba)), out(bs), info
    best_t, best_a = t0.clone(), a0.clone()
    best_s = torch.full((n,), -math.inf, dtype=torch.float64, device=dev)
    t = t0.clone().requires_grad_(True)
    a = a0.clone().requires_grad_(True)
    # per-parameter step sizes: translation in voxels, angles in radians
    opt = torch.optim.Adam([{"params": [t], "lr": lr_translation},
                            {"params": [a], "lr": math.radians(lr_angle_deg)}])
    evals = 0
    for step in range(steps + 1):
        opt.zero_grad(set_to_none=True)
        scores = []
        for lo in range(0, n, bsz):
            hi = min(lo + bsz, n)
            with torch.set_grad_enabled(step < steps):  # the last evaluation needs no Jacobian
                sc = score_from_moments(pose_moments(local, t[lo:hi], a[lo:hi], subvoxel))
            scores.append(sc.detach())
            if step < steps:
                (-sc.sum()).backward()
        sc = torch.cat(scores)
        evals += n
        with torch.no_grad():
            # selects instead of masked assignment: no host synchronisation,
            # so the next step's launches queue behind this step's kernels
            better = sc > best_s
            best_s = torch.where(better, sc, best_s)
            best_t = torch.where(better[:, None], t.detach(), best_t)
            best_a = torch.where(better[:, None], a.detach(), best_a)
        if step == steps:
            break
        opt.step()
        if bounds_translation is not None:
            with torch.no_grad():
                t.copy_(torch.clamp(t, t0 - bounds_translation, t0 + bounds_translation))
    info = {"score_evaluations_per_peak": evals // max(n, 1), "gradient_evaluations_per_peak": steps}
    return out(best_t), out(torch.rad2deg(best_a)), out(best_s), info
