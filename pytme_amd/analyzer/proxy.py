"""Shared analyzer proxies (mirrors ``tme/analyzer/proxy.py:21-123``).

The proxy owns an analyzer instance and its state.  When a shared-memory
manager is given (CPU process pools) the state arrays move to shared memory
and every update takes a lock; device backends run one rotation job per
analyzer and never share.
"""

from multiprocessing import Manager
from multiprocessing.shared_memory import SharedMemory
from typing import Tuple

from ..backends import backend as be

__all__ = ["StatelessSharedAnalyzerProxy", "SharedAnalyzerProxy"]


class StatelessSharedAnalyzerProxy:
    """Wrap a functional analyzer for (optionally locked) shared updates."""

    def __init__(self, analyzer_class: type, analyzer_params: dict):
        self._shared = False
        self._process = self._direct_call
        self._analyzer = analyzer_class(**analyzer_params)

    def __call__(self, state, *args, **kwargs):
        return self._process(state, *args, **kwargs)

    def init_state(self, shm_handler=None, *args, **kwargs) -> Tuple:
        state = self._analyzer.init_state()
        if shm_handler is not None:
            self._shared = True
            state = self._to_shared(state, shm_handler)
            self._lock = Manager().Lock()
            self._process = self._thread_safe_call
        return state

    def _to_shared(self, state: Tuple, shm_handler):
        backend_arr = type(be.zeros((1), dtype=be._float_dtype))
        out = []
        for v in state:
            if isinstance(v, backend_arr):
                v = be.to_sharedarr(v, shm_handler)
            elif isinstance(v, dict):
                v = Manager().dict(**v)
            out.append(v)
        return tuple(out)

    def _shared_to_object(self, shared):
        if not self._shared:
            return shared
        if isinstance(shared, tuple) and len(shared):
            if isinstance(shared[0], SharedMemory) or len(shared) == 1:
                return be.from_sharedarr(shared)
        return shared

    def _thread_safe_call(self, state, *args, **kwargs):
        with self._lock:
            state = tuple(self._shared_to_object(x) for x in state)
            return self._direct_call(state, *args, **kwargs)

    def _direct_call(self, state, *args, **kwargs):
        return self._analyzer(state, *args, **kwargs)

    def result(self, state, **kwargs):
        if self._shared:
            state = tuple(self._shared_to_object(x) for x in state)
        return self._analyzer.result(state, **kwargs)

    def merge(self, *args, **kwargs):
        return self._analyzer.merge(*args, **kwargs)


class SharedAnalyzerProxy(StatelessSharedAnalyzerProxy):
    """Stateful proxy: keeps the analyzer state and updates it on every call."""

    def __init__(self, analyzer_class: type, analyzer_params: dict, shm_handler: type = None, **kwargs):
        super().__init__(analyzer_class=analyzer_class, analyzer_params=analyzer_params)
        if not self._analyzer.shareable:
            shm_handler = None
        self.init_state(shm_handler)

    def init_state(self, shm_handler=None, *args, **kwargs) -> None:
        self._state = super().init_state(shm_handler, *args, **kwargs)

    @property
    def analyzer(self):
        return self._analyzer

    @property
    def state(self):
        return self._state

    def __call__(self, *args, **kwargs):
        state = super().__call__(self._state, *args, **kwargs)
        if not self._shared:
            self._state = state

    def result(self, **kwargs):
        return super().result(self._state, **kwargs)
