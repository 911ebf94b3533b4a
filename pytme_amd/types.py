"""Type aliases shared across the package (mirrors tme/types.py)."""

from typing import Any, Tuple, Union

import numpy as np

NDArray = np.ndarray
BackendArray = Any
Scalar = Union[int, float, complex]
shm_type = Union[Tuple[Any, ...], Any]
CallbackClass = object
MatchingData = object
