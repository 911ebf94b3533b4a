"""Backend manager (mirrors ``tme/backends/__init__.py:20-172``).

``backend`` is the process-wide manager.  Attribute access on it forwards to
the active backend instance, so score and analyzer code calls ``be.<op>``
without knowing which backend is active.  This package ships one backend,
``hip`` (gfx950 kernels via ``libpytme_hip.so``); further backends may be
registered with :meth:`BackendManager.add_backend`.
"""

from importlib.util import find_spec
from typing import Dict, List

from .matching_backend import MatchingBackend
from .hip_backend import HipBackend

__all__ = ["BackendManager", "MatchingBackend", "HipBackend", "backend"]


class BackendManager:
    """Registry and dispatcher for matching backends.

    The default backend is ``hip``.  It can be constructed without a GPU
    (``device="cpu"``) for host-side logic; its compute primitives then raise
    instead of silently running elsewhere.
    """

    def __init__(self):
        self._BACKEND_REGISTRY = {"hip": HipBackend}
        self._backend = None
        self._backend_name = "hip"
        self._backend_args = {}

    def _active(self):
        if self._backend is None:
            self._backend = self._BACKEND_REGISTRY[self._backend_name](**self._backend_args)
        return self._backend

    def __repr__(self):
        return f"<BackendManager: using {self._backend_name}>"

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(name)
        return getattr(self._active(), name)

    def __dir__(self) -> List:
        attrs = set(dir(self.__class__)) | set(self.__dict__.keys())
        attrs |= set(dir(self._active()))
        return sorted(attrs)

    def add_backend(self, backend_name: str, backend_class: type):
        """Register ``backend_class`` under ``backend_name``.

        Raises
        ------
        ValueError
            If ``backend_class`` does not inherit from :class:`MatchingBackend`
            (reference behaviour, ``tme/backends/__init__.py:105-106``).
        """
        if not (isinstance(backend_class, type) and issubclass(backend_class, MatchingBackend)):
            raise ValueError("backend_class needs to inherit from MatchingBackend.")
        self._BACKEND_REGISTRY[backend_name] = backend_class

    def change_backend(self, backend_name: str, **backend_kwargs: Dict) -> None:
        """Instantiate and activate the backend registered as ``backend_name``.

        Raises
        ------
        NotImplementedError
            If no backend of that name is registered
            (``tme/backends/__init__.py:130-134``).
        """
        if backend_name not in self._BACKEND_REGISTRY:
            available = ", ".join(self.available_backends())
            raise NotImplementedError(f"Available backends are {available} - not {backend_name}.")
        self._backend = self._BACKEND_REGISTRY[backend_name](**backend_kwargs)
        self._backend_name = backend_name
        self._backend_args = backend_kwargs

    def available_backends(self) -> List[str]:
        """Names of registered backends whose dependencies are importable."""
        deps = {"hip": "torch"}
        out = []
        for name in self._BACKEND_REGISTRY:
            dep = deps.get(name)
            if dep is None or find_spec(dep) is not None:
                out.append(name)
        return out


backend = BackendManager()
