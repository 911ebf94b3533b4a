"""One-call matching API (mirrors ``tme/cli.py:17-88``)."""

import numpy as np

__all__ = ["match_template"]


def match_template(target, template, template_mask=None, score="FLCSphericalMask", rotations=None):
    """Match ``template`` against ``target`` with ``MaxScoreOverRotations``.

    Uses ``score_threshold=-1``, ``pad_target_edges=True`` and one job, as
    the reference; returns ``(scores, offset, rotation_indices, mapping)``.
    Like the reference, ``template_mask=None`` is passed on explicitly, which
    leaves the template unmasked.
    """
    from .analyzer import MaxScoreOverRotations
    from .matching_data import MatchingData
    from .matching_exhaustive import MATCHING_EXHAUSTIVE_REGISTER, scan_subsets

    target = np.asarray(target)
    if rotations is None:
        rotations = np.eye(target.ndim).reshape(1, target.ndim, target.ndim)
    data = MatchingData(target=target, template=np.asarray(template), template_mask=template_mask, rotations=rotations)
    data._template_mask = template_mask
    setup, scoring = MATCHING_EXHAUSTIVE_REGISTER[score]
    return list(
        scan_subsets(
            matching_data=data,
            matching_score=scoring,
            matching_setup=setup,
            callback_class=MaxScoreOverRotations,
            callback_class_args={"score_threshold": -1},
            pad_target_edges=True,
            job_schedule=(1, 1),
        )
    )
