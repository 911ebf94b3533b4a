"""Memory model of the exhaustive scan (subset of ``tme/memory.py``).

``register_matching_exhaustive`` requires a :class:`MatchingMemoryUsage`
subclass per score (tme/matching_exhaustive.py:497-498).  The classes here
count the device arrays the hip backend keeps live for one scan: the target
spectra, the per-rotation work arrays and the analyzer state, in bytes.
"""

from typing import Dict, Tuple

import numpy as np

__all__ = ["MatchingMemoryUsage", "MATCHING_MEMORY_REGISTRY", "estimate_memory_usage"]


class MatchingMemoryUsage:
    """Bytes needed by a score for real shape ``fast_shape``.

    Parameters
    ----------
    fast_shape, ft_shape : tuple of int
        Real and half-spectrum shapes of the padded transforms.
    float_nbytes, complex_nbytes, integer_nbytes : int
        Element sizes.
    """

    #: (real arrays, complex arrays) kept for the whole scan
    base = (0, 0)
    #: (real arrays, complex arrays) allocated per rotation loop
    per_rotation = (0, 0)

    def __init__(self, fast_shape, ft_shape, float_nbytes=4, complex_nbytes=8, integer_nbytes=4):
        self.real_array_size = int(np.prod(fast_shape))
        self.complex_array_size = int(np.prod(ft_shape))
        self.float_nbytes = float_nbytes
        self.complex_nbytes = complex_nbytes
        self.integer_nbytes = integer_nbytes

    def _bytes(self, counts: Tuple[int, int]) -> int:
        r, c = counts
        return r * self.real_array_size * self.float_nbytes + c * self.complex_array_size * self.complex_nbytes

    def base_usage(self) -> int:
        return self._bytes(self.base)

    def per_fork(self) -> int:
        return self._bytes(self.per_rotation)


class CCMemoryUsage(MatchingMemoryUsage):
    base = (1, 1)
    per_rotation = (1, 1)


class LCCMemoryUsage(CCMemoryUsage):
    pass


class FLCSphericalMaskMemoryUsage(MatchingMemoryUsage):
    base = (2, 1)
    per_rotation = (1, 1)


class FLCMemoryUsage(MatchingMemoryUsage):
    base = (0, 2)
    per_rotation = (3, 2)


class MaxScoreOverRotationsMemoryUsage(MatchingMemoryUsage):
    base = (2, 0)


class PeakCallerMaximumFilterMemoryUsage(MatchingMemoryUsage):
    """Peak calling on the max-score volume (tme/memory.py:197-212): the
    volume and the hip backend's filtered copy; the peak bitmask (F/8
    bytes) is folded into the second array."""

    base = (2, 0)


class HipBackendMemoryUsage(MatchingMemoryUsage):
    """Device memory of the hip backend beyond the score's arrays: the fused
    scan's batch workspace (``tm_scan_workspace_bytes``) for the FLC plan
    with 16 rotations per batch, rows padded to 16 bins, and a template of
    edge min(N0, 64) for the pruned intermediates (the memory classes see
    only the transform shapes).  Plays the role of the reference's
    ``CupyBackendMemoryUsage`` (tme/memory.py:216-232) for an HBM budget."""

    batch = 16
    template_edge = 64

    def __init__(self, fast_shape, ft_shape, float_nbytes=4, complex_nbytes=8, integer_nbytes=4):
        super().__init__(fast_shape, ft_shape, float_nbytes, complex_nbytes, integer_nbytes)
        self.fast_shape = tuple(int(x) for x in fast_shape)

    def base_usage(self) -> int:
        n0, n1, n2 = (self.fast_shape + (1, 1, 1))[:3]
        n2s = -(-(n2 // 2 + 1) // 16) * 16
        t0 = min(n0, self.template_edge)
        B, cb = self.batch, self.complex_nbytes
        spectra = 3 * B * n0 * n1 * n2s * cb
        pruned = 2 * B * t0 * (t0 + n1) * n2s * cb
        padded_target = 2 * n0 * n1 * n2s * cb
        return int(spectra + pruned + padded_target)

    def per_fork(self) -> int:
        return 0


MATCHING_MEMORY_REGISTRY: Dict[str, type] = {
    "CC": CCMemoryUsage,
    "LCC": LCCMemoryUsage,
    "FLC": FLCMemoryUsage,
    "FLCSphericalMask": FLCSphericalMaskMemoryUsage,
    "MaxScoreOverRotations": MaxScoreOverRotationsMemoryUsage,
    "PeakCallerMaximumFilter": PeakCallerMaximumFilterMemoryUsage,
    "hip": HipBackendMemoryUsage,
}


def estimate_memory_usage(
    shape1: Tuple[int],
    shape2: Tuple[int],
    matching_method: str,
    ncores: int = 1,
    analyzer_method: str = None,
    backend: str = None,
    float_nbytes: int = 4,
    complex_nbytes: int = 8,
    integer_nbytes: int = 4,
    **kwargs,
) -> int:
    """Bytes for matching ``shape2`` against ``shape1`` with ``ncores``
    loops: the score's arrays, the analyzer's and the backend's
    (tme/memory.py:252-338)."""
    if matching_method not in MATCHING_MEMORY_REGISTRY:
        raise ValueError(f"Supported options are {','.join(MATCHING_MEMORY_REGISTRY)}")
    from .backends.hip_backend import _next_fast_len

    conv = [int(a) + int(b) - 1 for a, b in zip(shape1, shape2)]
    fast = [_next_fast_len(x) for x in conv]
    ft = fast[:-1] + [fast[-1] // 2 + 1]
    args = (fast, ft, float_nbytes, complex_nbytes, integer_nbytes)
    usage = MATCHING_MEMORY_REGISTRY[matching_method](*args)
    total = usage.base_usage() + ncores * usage.per_fork()
    for extra in (analyzer_method, backend):
        if extra in MATCHING_MEMORY_REGISTRY:
            usage = MATCHING_MEMORY_REGISTRY[extra](*args)
            total += usage.base_usage() + ncores * usage.per_fork()
    return int(total)
