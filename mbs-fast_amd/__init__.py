"""MI355X (gfx950) physical-optics diffraction engine for MBS-fast.

The compute path is the HIP library built from ``csrc/`` (``libmbs_hip.so``,
C ABI in ``include/mbs_hip.h``).  Python holds only the host-side mirror of the
reference's HandlerPO GPU interface (``handler.py``) and the ctypes binding
(``engine.py``); there is no CPU fallback.  The directory name contains a
hyphen, so it is imported through :func:`load` (or ``importlib``) under the
module name ``mbs_fast_amd``.
"""
from . import abi  # noqa: F401

__all__ = ["abi", "engine", "handler", "synthetic", "cv"]


def __getattr__(name):
    if name in ("engine", "handler", "build", "synthetic", "cv"):
        import importlib
        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(name)
