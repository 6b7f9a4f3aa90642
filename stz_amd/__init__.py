"""stz_amd -- MI355X (gfx950) quasi-static STZ elastoplasticity timestep.

The hot path of the reference's ``sb`` simulations (shear_sim's
step_forward_quasistatic with its tgmg multigrid projection) as hand-written
HIP kernels behind the C ABI of include/stz_amd.h.  ``ShearSim`` mirrors the
reference class; ``problems`` holds the workload parameter sets.
"""
from . import problems
from ._lib import StzError, device_count, load
from .ensemble import Ensemble
from .sim import FIELD_NAMES, ShearSim

__all__ = ["ShearSim", "Ensemble", "StzError", "FIELD_NAMES", "problems", "device_count", "load"]
