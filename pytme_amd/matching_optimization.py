"""Non-exhaustive (local) matching: the reference's refinement API on the GPU.

Re-creates ``tme.matching_optimization`` for the score the hot path needs,
``FLC`` (tme/matching_optimization.py:488-548), with ``create_score_object``
(:1079-1140), ``register_matching_optimization`` (:1053-1076) and
``optimize_match`` (:1143-1312).  Scores are evaluated by the HIP kernel
``tm_flc_local_fwd`` (see :mod:`pytme_amd.refine`); the score object also
offers the gradient (``return_gradient=True``), which the reference's FLC
does not have, and :func:`pytme_amd.refine.refine_peaks` refines many peaks
at once with it.
"""

from typing import Dict, List, Tuple

import numpy as np
from scipy.optimize import LinearConstraint, basinhopping, differential_evolution, minimize

from .refine import LocalFLC, score_poses
from .rotations import euler_to_rotationmatrix

__all__ = ["FLC", "MATCHING_OPTIMIZATION_REGISTER", "register_matching_optimization", "create_score_object",
           "optimize_match"]


class FLC:
    """Local FLC of a template (and mask) against a target at a pose
    x = (translation(3) voxels, zyz Euler angles(3) degrees); see
    tme/matching_optimization.py:488-548 and :54-238 for the constructor.

    ``score(x)`` returns the FLC times ``-1`` when ``negate_score`` (the
    reference's default, for minimisers).  The rotated mask is used as the
    mask (the reference's rotate_array writes it into the template buffer,
    tme/matching_optimization.py:162-163), and the sub-voxel part of the
    translation moves the rotated template rigidly (``subvoxel="rigid"``;
    ``"reference"`` keeps the reference's sampling, see
    :mod:`pytme_amd.refine`).
    """

    def __init__(self, target, template, template_mask=None, target_mask=None, pad_target_edges: bool = False,
                 pad_fourier: bool = False, rotate_mask: bool = True, interpolation_order: int = 1,
                 negate_score: bool = True, return_gradient: bool = False, subvoxel: str = "rigid",
                 **kwargs: Dict):
        if int(interpolation_order) != 1:
            raise NotImplementedError("the hip refinement interpolates with order 1")
        if not rotate_mask:
            raise NotImplementedError("the hip refinement rotates the mask with the template")
        if pad_target_edges or pad_fourier:
            raise NotImplementedError("padding is not used by the local FLC")
        target = np.asarray(target, dtype=np.float32)
        if target_mask is not None:
            target = target * np.asarray(target_mask, dtype=np.float32)  # _post_init, :512-513
        self.eps = float(np.finfo(np.float32).eps)
        self.score_sign = -1 if negate_score else 1
        self.return_gradient = bool(return_gradient)
        self.subvoxel = subvoxel
        self._local = LocalFLC(target, template, template_mask, max_poses=kwargs.get("max_poses", 64))

    @property
    def evaluations(self) -> int:
        return self._local.evaluations

    def score(self, x: Tuple[float]):
        """Score (and, with ``return_gradient``, its gradient d/dx with the
        angles in degrees) of pose x."""
        x = np.asarray(x, dtype=np.float64)
        if not self.return_gradient:
            return float(score_poses(self._local, x[None, :3], x[None, 3:], self.subvoxel)[0]) * self.score_sign
        import torch

        from .refine import pose_moments, score_from_moments

        dev = self._local.device
        t = torch.tensor(x[None, :3], dtype=torch.float64, device=dev, requires_grad=True)
        a = torch.tensor(x[None, 3:], dtype=torch.float64, device=dev, requires_grad=True)
        sc = score_from_moments(pose_moments(self._local, t, torch.deg2rad(a), self.subvoxel))[0] * self.score_sign
        sc.backward()
        grad = np.concatenate([t.grad[0].cpu().numpy(), a.grad[0].cpu().numpy()])
        return float(sc.detach().cpu()), grad

    def score_translation(self, x: Tuple[float]):
        return self.score((*x, *[0 for _ in range(len(x))]))

    def score_angles(self, x: Tuple[float]):
        return self.score((*[0 for _ in range(len(x))], *x))


MATCHING_OPTIMIZATION_REGISTER = {"FLC": FLC}


def register_matching_optimization(match_name: str, match_class: type):
    """tme/matching_optimization.py:1053-1076: add a score class."""
    if match_name in MATCHING_OPTIMIZATION_REGISTER:
        raise ValueError(f"A method with name '{match_name}' is already registered.")
    if not hasattr(match_class, "score") or not callable(getattr(match_class, "score")):
        raise ValueError(f"{match_class} needs a callable 'score' method.")
    MATCHING_OPTIMIZATION_REGISTER[match_name] = match_class


def create_score_object(score: str, **kwargs) -> object:
    """tme/matching_optimization.py:1079-1140."""
    if score not in MATCHING_OPTIMIZATION_REGISTER:
        valid = ", ".join(MATCHING_OPTIMIZATION_REGISTER.keys())
        raise ValueError(f"score has to be one of {valid}")
    return MATCHING_OPTIMIZATION_REGISTER[score](**kwargs)


def optimize_match(score_object: object, bounds_translation: Tuple[Tuple[float]] = None,
                   bounds_rotation: Tuple[Tuple[float]] = None, optimization_method: str = "basinhopping",
                   maxiter: int = 50, x0: Tuple[float] = None,
                   rng=None) -> Tuple[np.ndarray, np.ndarray, float]:
    """Minimise ``score_object.score`` over (translation, zyz angles)
    (tme/matching_optimization.py:1143-1312): scipy basinhopping with COBYLA
    steps, differential evolution or minimize, with the same bounds, linear
    constraints, initial-score fallback and return values
    ``(translation, rotation_matrix, score)``.  ``rng`` (not in the
    reference, default unseeded as there) seeds the random steps of
    basinhopping and differential evolution, for repeatable runs."""
    ndim = 3
    methods = {"differential_evolution": differential_evolution, "basinhopping": basinhopping, "minimize": minimize}
    if optimization_method not in methods:
        raise ValueError(f"{optimization_method} is not supported. Pick from {', '.join(list(methods.keys()))}")
    finfo = np.finfo(np.float32)
    if optimization_method == "differential_evolution" and bounds_translation is None:
        bounds_translation = tuple((finfo.min, finfo.max) for _ in range(ndim))
    if bounds_translation is None and bounds_rotation is not None:
        bounds_translation = tuple((finfo.min, finfo.max) for _ in range(ndim))
    if bounds_rotation is None and bounds_translation is not None:
        bounds_rotation = tuple((-180, 180) for _ in range(ndim))
    bounds, linear_constraint = None, ()
    if bounds_rotation is not None and bounds_translation is not None:
        uncertainty = (*bounds_translation, *bounds_rotation)
        bounds = [b if b != (0, 0) else (-finfo.resolution, finfo.resolution) for b in uncertainty]
        linear_constraint = LinearConstraint(np.eye(len(bounds)), np.min(bounds, axis=1), np.max(bounds, axis=1))
    x0 = np.zeros(2 * ndim) if x0 is None else np.asarray(x0, dtype=np.float64)

    return_gradient = getattr(score_object, "return_gradient", False)
    if optimization_method != "minimize" and return_gradient:
        score_object.return_gradient = False
        return_gradient = False
    initial = score_object.score(x=x0)
    if isinstance(initial, (List, Tuple)):
        initial = initial[0]
    if optimization_method == "basinhopping":
        result = basinhopping(x0=x0, func=score_object.score, niter=maxiter,
                              minimizer_kwargs={"method": "COBYLA", "constraints": linear_constraint}, rng=rng)
    elif optimization_method == "differential_evolution":
        result = differential_evolution(func=score_object.score, bounds=bounds, constraints=linear_constraint,
                                        maxiter=maxiter, rng=rng)
    else:
        result = minimize(x0=x0, fun=score_object.score, jac=return_gradient, bounds=bounds,
                          constraints=linear_constraint, options={"maxiter": maxiter})
    fun = float(result.fun)
    x = np.asarray(result.x, dtype=np.float64)
    if initial < fun:
        x = np.zeros_like(x)
    translation, rotation = x[:ndim], x[ndim:]
    return translation, euler_to_rotationmatrix(rotation), fun
