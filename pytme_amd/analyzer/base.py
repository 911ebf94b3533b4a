"""Analyzer protocol (mirrors ``tme/analyzer/base.py:15-127``).

An analyzer is functional: ``init_state()`` creates a state tuple,
``__call__(state, scores, rotation_matrix)`` folds one score volume into it
and returns the new state, ``result(state)`` finalises it and the class
method ``merge(results)`` combines results of independent scans.
"""

from abc import ABC, abstractmethod
from typing import List, Tuple

__all__ = ["AbstractAnalyzer"]


class AbstractAnalyzer(ABC):
    @property
    def shareable(self) -> bool:
        """Whether the state may live in memory shared across processes."""
        return False

    @abstractmethod
    def init_state(self, *args, **kwargs) -> Tuple:
        """Initial state tuple."""

    @abstractmethod
    def __call__(self, state, scores, rotation_matrix, **kwargs) -> Tuple:
        """Fold ``scores`` obtained with ``rotation_matrix`` into ``state``."""

    @abstractmethod
    def result(self, state: Tuple, **kwargs) -> Tuple:
        """Final result of the analysis."""

    @classmethod
    @abstractmethod
    def merge(cls, results: List[Tuple], **kwargs) -> Tuple:
        """Combine results of several analyzers of this class."""
