"""Analyzer proxies: the objects the scan hands to its scoring functions as
``callback`` (public interface of ``tme/analyzer/proxy.py:21-123``).

A proxy owns one analyzer.  In the usual case (one rotation job per
analyzer, always on a GPU device) the analyzer's state tuple stays local and
each call replaces it with the analyzer's return value.  When the scan
gives a shared-memory handler (CPU process pools), the state is published
once through :class:`_SharedState`: its arrays go through the backend's
``to_sharedarr``, its mappings into a manager dict, and every update runs
under one manager lock on the attached arrays, which the analyzer updates
in place.
"""

from multiprocessing import Manager
from typing import Any, Dict, Tuple

from ..backends import backend as be

__all__ = ["StatelessSharedAnalyzerProxy", "SharedAnalyzerProxy"]


def _is_backend_array(item) -> bool:
    # backend arrays have one concrete class (tme/matching_exhaustive.py:46)
    return isinstance(item, type(be.zeros((1,), dtype=be._float_dtype)))


class _SharedState:
    """Publishes an analyzer state across processes and attaches to it.

    Remembers which slots of the state tuple it turned into backend shared
    arrays, so that :meth:`attach` maps exactly those back (the other slots,
    manager dicts and plain values, are usable as they are).  The lock and
    the dict proxies pickle into worker processes; the manager process lives
    as long as a proxy to it does.
    """

    def __init__(self, state: Tuple, shm_handler):
        manager = Manager()
        self.lock = manager.Lock()
        self.array_slots = frozenset(i for i, item in enumerate(state) if _is_backend_array(item))
        published = []
        for i, item in enumerate(state):
            if i in self.array_slots:
                item = be.to_sharedarr(item, shm_handler)
            elif isinstance(item, dict):
                item = manager.dict(item)
            published.append(item)
        self.state = tuple(published)

    def attach(self, state: Tuple) -> Tuple:
        return tuple(be.from_sharedarr(item) if i in self.array_slots else item for i, item in enumerate(state))


class StatelessSharedAnalyzerProxy:
    """Proxy of a functional analyzer; the caller carries the state.

    ``init_state(shm_handler)`` returns a fresh state (published when a
    handler is given); ``proxy(state, ...)`` applies one update and returns
    the analyzer's new state; ``result`` and ``merge`` forward to the
    analyzer.
    """

    def __init__(self, analyzer_class: type, analyzer_params: Dict[str, Any]):
        self._analyzer = analyzer_class(**analyzer_params)
        self._sharing = None  # _SharedState once the state is published

    @property
    def analyzer(self):
        return self._analyzer

    @property
    def is_shared(self) -> bool:
        """True when the state lives in shared memory behind a lock."""
        return self._sharing is not None

    def init_state(self, shm_handler=None, *args, **kwargs) -> Tuple:
        state = self._analyzer.init_state()
        if shm_handler is None:
            self._sharing = None
            return state
        self._sharing = _SharedState(state, shm_handler)
        return self._sharing.state

    def __call__(self, state, *args, **kwargs):
        if self._sharing is None:
            return self._analyzer(state, *args, **kwargs)
        with self._sharing.lock:
            return self._analyzer(self._sharing.attach(state), *args, **kwargs)

    def result(self, state, **kwargs):
        if self._sharing is not None:
            state = self._sharing.attach(state)
        return self._analyzer.result(state, **kwargs)

    def merge(self, *args, **kwargs):
        return self._analyzer.merge(*args, **kwargs)


class SharedAnalyzerProxy(StatelessSharedAnalyzerProxy):
    """Proxy that keeps the state: ``proxy(...)`` updates it, ``result()``
    reads it.  A handler is used only for analyzers that declare their state
    ``shareable``."""

    def __init__(self, analyzer_class: type, analyzer_params: Dict[str, Any], shm_handler=None, **kwargs):
        super().__init__(analyzer_class, analyzer_params)
        self.init_state(shm_handler if self._analyzer.shareable else None)

    def init_state(self, shm_handler=None, *args, **kwargs) -> None:
        self._state = super().init_state(shm_handler)

    @property
    def state(self) -> Tuple:
        return self._state

    def __call__(self, *args, **kwargs) -> None:
        updated = super().__call__(self._state, *args, **kwargs)
        if not self.is_shared:  # shared arrays were updated in place
            self._state = updated

    def result(self, **kwargs):
        return super().result(self._state, **kwargs)
