"""Filter composition (mirrors ``tme/filters/compose.py:17-59``)."""

from abc import ABC, abstractmethod
from typing import Dict, Tuple

import numpy as np

__all__ = ["Compose", "ComposableFilter"]


class Compose:
    """Apply filters in sequence, multiplying multiplicative filter weights.

    Each transform is called with the accumulated keyword arguments updated
    by the previous transform's metadata; when a transform marks itself
    ``is_multiplicative_filter`` its ``data`` is multiplied into the previous
    ``data``.
    """

    def __init__(self, transforms: Tuple[object]):
        self.transforms = tuple(transforms)

    def __call__(self, **kwargs: Dict) -> Dict:
        if not self.transforms:
            return {}
        meta = self.transforms[0](**kwargs)
        for transform in self.transforms[1:]:
            kwargs.update(meta)
            ret = transform(**kwargs)
            if "data" not in ret:
                continue
            if ret.get("is_multiplicative_filter", False):
                prev = meta.pop("data")
                ret["data"] = np.multiply(np.asarray(ret["data"]), np.asarray(prev))
                ret["merge"] = None
            meta = ret
        return meta


class ComposableFilter(ABC):
    @abstractmethod
    def __call__(self, *args, **kwargs) -> Dict:
        """Return a dict with at least ``data``."""
