"""MI355X-native mimetic hybrid pressure solver (opm-porsol IncompFlowSolverHybrid).

The directory name is not a Python identifier; load it with ``load_package``
(in __graft_entry__.py, tests/conftest.py and bench.py), which registers it as
``opm_porsol_amd``.
"""
from . import deck, fields, grids, mhs, output, transport  # noqa: F401
from .mhs import HybridSolver, MhsError  # noqa: F401
from .transport import EulerUpstream  # noqa: F401


def __getattr__(name):
    # torch is imported only when the autograd bridge is asked for
    if name == "autograd":
        import importlib
        return importlib.import_module(__name__ + ".autograd")
    raise AttributeError(name)
