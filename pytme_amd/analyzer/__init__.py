"""Score analyzers of the exhaustive scan."""

from .base import AbstractAnalyzer
from .aggregation import MaxScoreOverRotations
from .proxy import SharedAnalyzerProxy, StatelessSharedAnalyzerProxy
from ._utils import cart_to_score, score_to_cart

__all__ = [
    "AbstractAnalyzer",
    "MaxScoreOverRotations",
    "SharedAnalyzerProxy",
    "StatelessSharedAnalyzerProxy",
    "cart_to_score",
    "score_to_cart",
]
