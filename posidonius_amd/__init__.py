"""posidonius_amd -- MI355X-native WHFast ensemble integrator for Posidonius cases.

The integration step runs only in the HIP library ``libposidonius_hip.so``
(built in-tree by ``__graft_entry__.build()``); this package holds the host
side: the ctypes binding of the C ABI (:mod:`.abi`), the reference's JSON case
format (:mod:`.case`), synthetic ensemble generators (:mod:`.generators`) and
the :class:`Ensemble` handle (:mod:`.ensemble`).
"""
from . import abi, case  # noqa: F401
from .case import EnsembleSpec, read_case  # noqa: F401

__all__ = ["abi", "case", "EnsembleSpec", "read_case"]
