"""Comparison helpers for HIP-vs-oracle scan parity."""

import numpy as np

SCORE_TOL = 1e-4  # absolute score tolerance vs the FP64 oracle
TIE_TOL = 2e-5  # rotations whose oracle scores differ by less are ties


def assert_scan_parity(scores, rot, raws, orc, threshold, tol=SCORE_TOL, tie_tol=TIE_TOL):
    """scores/rot: HIP result (rolled, cropped).  raws: oracle raw volumes.

    Scores must agree within ``tol`` everywhere.  Rotation indices must be
    identical except at documented ties: where they differ, the oracle score
    of the rotation the HIP path chose must be within ``tie_tol`` of the
    oracle's best (or of the threshold for never-updated voxels).
    """
    best = np.full(orc.fast, threshold, dtype=np.float64)
    ridx = np.full(orc.fast, -1, dtype=np.int32)
    for i, raw in enumerate(raws):
        upd = raw > best
        best[upd] = raw[upd]
        ridx[upd] = i
    ref_s, ref_r = orc.result(best, ridx)
    assert scores.shape == ref_s.shape, (scores.shape, ref_s.shape)
    err = np.abs(scores.astype(np.float64) - ref_s)
    assert err.max() <= tol, f"max |score error| {err.max():.3g} > {tol}"
    diff = rot != ref_r
    n_diff = int(diff.sum())
    if n_diff:
        stack = np.stack([orc.result(raw, ridx)[0] for raw in raws])  # crop each raw the same way
        idx = np.argwhere(diff)
        for p in idx:
            p = tuple(p)
            g = int(rot[p])
            chosen = threshold if g < 0 else stack[(g,) + p]
            assert abs(chosen - ref_s[p]) <= tie_tol, (
                f"rotation mismatch at {p}: hip {g} ({chosen}) vs oracle {int(ref_r[p])} ({ref_s[p]})"
            )
    return n_diff, float(err.max())
