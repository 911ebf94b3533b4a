"""Score analyzers of the exhaustive scan."""

from .base import AbstractAnalyzer
from .aggregation import MaxScoreOverRotations
from .proxy import SharedAnalyzerProxy, StatelessSharedAnalyzerProxy
from ._utils import cart_to_score, score_to_cart
from .peaks import PeakCaller, PeakCallerMaximumFilter, PeakCallerSort, filter_points, filter_points_indices

__all__ = [
    "AbstractAnalyzer",
    "MaxScoreOverRotations",
    "SharedAnalyzerProxy",
    "StatelessSharedAnalyzerProxy",
    "cart_to_score",
    "score_to_cart",
    "PeakCaller",
    "PeakCallerSort",
    "PeakCallerMaximumFilter",
    "filter_points",
    "filter_points_indices",
]
