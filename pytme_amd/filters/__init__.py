"""Composable Fourier filters evaluated once per scan (host, numpy).

Only the filters on the cryo-ET scan path are provided: the reconstructed
tilt wedge (step and continuous) and the reconstructed CTF, combined with
:class:`Compose`.  They produce the real half-spectrum weights that the HIP
spectral multiply consumes.
"""

from .compose import Compose, ComposableFilter
from .wedge import WedgeReconstructed
from .ctf import CTFReconstructed, create_ctf

__all__ = ["Compose", "ComposableFilter", "WedgeReconstructed", "CTFReconstructed", "create_ctf"]
