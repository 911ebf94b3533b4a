"""MI355X-native exhaustive template matching with pyTME's plugin API.

The package re-creates the API surface of pyTME's exhaustive scan
(backend manager, matching backends, score registry, ``scan`` /
``scan_subsets``, ``MatchingData`` and the ``MaxScoreOverRotations``
analyzer) and provides a ``hip`` backend whose hot path runs hand-written
gfx950 HIP kernels from ``libpytme_hip.so``.
"""

__version__ = "0.1.0"

from .backends import backend  # noqa: F401
from .density import Density  # noqa: F401
from .orientations import Orientations  # noqa: F401
