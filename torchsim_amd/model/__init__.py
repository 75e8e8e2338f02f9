"""Signal models."""

from ..sequence._events import BALANCED, REFOCUSED, SPOILED, UNBALANCED, EventOperators
from ._simulator import Simulator, SpinPhysics

__all__ = ["Simulator", "SpinPhysics", "EventOperators", "BALANCED", "REFOCUSED", "SPOILED", "UNBALANCED"]
