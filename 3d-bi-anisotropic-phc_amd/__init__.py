"""MI355X (gfx950) curl-SVD + bi-anisotropic material matvec path.

Import this directory as a package (its name is not a Python identifier; load it
with ``importlib`` under an alias such as ``phc_amd`` — see ``load_package`` in
``__graft_entry__.py``).
"""
from . import adjoint, driver, jdsira, outputs, structures  # noqa: F401
from .fame import (FameError, FccPlan, NotConverged, Plan, exported_symbols, load_library,  # noqa: F401
                   stream_probe, LIB_PATH)
