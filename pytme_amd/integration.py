"""Plug this package's HIP backend and batched fused scan into an installed,
unmodified pyTME.

pyTME offers two extension points on the scan's path, and :func:`install`
uses only those, so pyTME's own ``scan`` / ``scan_subsets``, setups,
scoring functions, analyzers and CLI run unchanged:

* the backend registry: ``tme.backends.backend.add_backend(name, cls)``
  (tme/backends/__init__.py:82-107) takes a subclass of pyTME's
  ``MatchingBackend``; ``change_backend(name, **kwargs)`` then instantiates
  it (:109-137).  The class registered here is this package's
  ``HipBackend`` combined with pyTME's abstract base class.
* the score registry: ``tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER``
  (tme/matching_scores.py:1183-1193), the dict ``scan_subsets`` and the CLI
  look scores up in (tme/matching_exhaustive.py:24).  Its entries whose
  scoring function is pyTME's ``flc_scoring`` or ``corr_scoring`` are
  replaced, in the same dict, by that function behind a first call to the
  active backend's ``fused_scoring``: on this backend, with pyTME's
  ``MaxScoreOverRotations`` as the (unshared) analyzer, the whole rotation
  set goes through the batched fused scan; anything it does not cover
  (other backends, analyzers, scores, shared CPU state) falls through to
  pyTME's own loop with the same arguments.
"""

import functools
from typing import Dict, Optional

__all__ = ["install", "FUSED_SCORING_FUNCTIONS"]

#: pyTME scoring functions (by name) that the fused scan implements, and
#: the fused hook's score family for each
FUSED_SCORING_FUNCTIONS = {"flc_scoring": "flc", "corr_scoring": "corr"}


def _unsupported(name: str):
    def method(self, *args, **kwargs):
        raise NotImplementedError(f"the hip backend does not implement {name}")

    method.__name__ = name
    return method


def backend_class(reference_base: type) -> type:
    """This package's ``HipBackend`` as a subclass of pyTME's
    ``MatchingBackend`` (``reference_base``).  Abstract methods of that base
    the hip backend has no use for raise ``NotImplementedError`` when
    called, so the class is instantiable."""
    from .backends.hip_backend import HipBackend

    missing = sorted(
        m for m in getattr(reference_base, "__abstractmethods__", ()) if not callable(getattr(HipBackend, m, None))
    )
    namespace = {m: _unsupported(m) for m in missing}
    namespace["__doc__"] = HipBackend.__doc__
    namespace["__module__"] = __name__
    return type("HipBackend", (HipBackend, reference_base), namespace)


def _with_fused_scan(scoring, family: str, manager):
    @functools.wraps(scoring)
    def scoring_with_fused_scan(*args, **kwargs):
        hook = getattr(manager._backend, "fused_scoring", None)
        if hook is not None and not args:
            ret = hook(score=family, **kwargs)
            if ret is not NotImplemented:
                return ret
        return scoring(*args, **kwargs)

    scoring_with_fused_scan.__wrapped_by_pytme_amd__ = True
    return scoring_with_fused_scan


def install(tme=None, name: str = "hip") -> Dict:
    """Register the hip backend under ``name`` and route pyTME's FLC /
    correlation scoring functions through the fused scan.

    ``tme`` is the imported pyTME package (imported here when omitted).
    Calling it again is harmless.  Returns what changed: the backend name
    and class, and the scores whose scoring function now tries the fused
    scan first.
    """
    if tme is None:
        import tme  # noqa: F811 - the user's installed pyTME

    backends = tme.backends
    manager = backends.backend
    cls = backend_class(backends.MatchingBackend)
    manager.add_backend(name, cls)

    # available_backends() lists only names of a fixed dependency table
    # (tme/backends/__init__.py:154-166), which the CLI's --backend choices
    # come from; list this one as well
    listed = manager.available_backends
    if not getattr(listed, "__lists_pytme_amd__", None) == name:

        def available_backends():
            names = list(listed())
            return names if name in names else names + [name]

        available_backends.__lists_pytme_amd__ = name
        manager.available_backends = available_backends

    register = tme.matching_scores.MATCHING_EXHAUSTIVE_REGISTER
    routed = []
    for score, (setup, scoring) in list(register.items()):
        if getattr(scoring, "__wrapped_by_pytme_amd__", False):
            routed.append(score)
            continue
        family = FUSED_SCORING_FUNCTIONS.get(getattr(scoring, "__name__", ""))
        if family is None:
            continue
        register[score] = (setup, _with_fused_scan(scoring, family, manager))
        routed.append(score)
    return {"backend": name, "backend_class": cls, "fused_scores": routed}
