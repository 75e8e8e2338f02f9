"""torchsim on MI355X: the EPG hot path of torchsim on hand-written gfx950 kernels.

The public API mirrors torchsim's (``mrf_sim``, ``simulators.MRFSimulator``,
``Simulator.simulate/jacobian``); the compute runs in ``libtorchsim_hip.so``.
"""

from . import estimators, model, optim, recon, sequence, simulators
from ._functional import bssfp_sim, fse_sim, mrf_sim
from ._subspace import Subspace
from .optim import crlb

__all__ = ["Subspace", "bssfp_sim", "crlb", "estimators", "fse_sim", "model", "mrf_sim", "optim", "recon",
           "sequence", "simulators"]
