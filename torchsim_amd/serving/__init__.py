"""Serving: a continuous-batching request loop over the uneven-batch launcher."""

from ._batch import BatchItem, launch_batch
from ._server import FitRequest, Server, SimulateRequest

__all__ = ["BatchItem", "FitRequest", "Server", "SimulateRequest", "launch_batch"]
