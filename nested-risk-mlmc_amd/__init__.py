"""MI355X-native level-l nested Monte Carlo sampler (gfx950 HIP kernels behind a C ABI).

The package directory name is not a Python identifier; load it with
``nested_risk_mlmc_amd = importlib`` helpers such as ``__graft_entry__.package()``,
or put this directory on sys.path and ``import nmlmc``.
"""
from .nmlmc import *  # noqa: F401,F403
from .nmlmc import load, check, NmlmcError  # noqa: F401
