"""MI355X-native OdeVec: a vectorized LSODE-style BDF integrator for many
small stiff chemistry networks, one GPU lane per cell.

* ``odevec_amd.pre_odevec`` -- the preprocessor (networks, orderings,
  packaging, HIP code emission);
* ``odevec_amd.api.Odevec`` -- the host interface mirroring the reference's
  ``type(odevec)`` / ``InitOdevec`` / ``SolveODE`` / ``CloseOdevec``;
* ``odevec_amd._lib`` -- the ctypes binding of ``libodevec.so``.
"""
from . import _lib  # noqa: F401

__all__ = ["Odevec", "networks", "network_info", "kernel_info"]


def __getattr__(name):
    if name == "Odevec":
        from .api import Odevec
        return Odevec
    if name in ("networks", "network_info", "kernel_info"):
        return getattr(_lib, name)
    raise AttributeError(name)
