"""Lazy loader of the rans_sst_py extension (reference python/solver_bindings.py:13-27).

Same contract: modules import `_rs` and get the module on first call; the
ImportError is raised only when a solver-backed method is used.  The module
found here is the MI355X build (bayesian-turbulence-modeling_amd/), whose
directory holds the extension next to its engine library librsg.so.
"""
import os
import sys

_rans_sst_py = None
_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _rs():
    global _rans_sst_py
    if _rans_sst_py is None:
        if _PKG not in sys.path:
            sys.path.insert(0, _PKG)
        try:
            import rans_sst_py as _m
        except ImportError:
            raise ImportError(
                "Cannot import rans_sst_py. Build it first:\n"
                "  python -c 'import __graft_entry__ as g; g.build()'\n"
                "(compiles librsg.so for gfx950 and the rans_sst_py module in "
                "bayesian-turbulence-modeling_amd/)") from None
        _rans_sst_py = _m
    return _rans_sst_py
