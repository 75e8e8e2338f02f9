"""torchsim on MI355X: the EPG hot path of torchsim on hand-written gfx950 kernels.

The public API mirrors torchsim's (``mrf_sim``, ``simulators.MRFSimulator``,
``Simulator.simulate/jacobian``); the compute runs in ``libtorchsim_hip.so``.
"""

from . import model, sequence, simulators
from ._functional import mrf_sim

__all__ = ["mrf_sim", "model", "sequence", "simulators"]
