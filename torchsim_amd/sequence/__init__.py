"""Event vocabulary, packing and the native engine."""

from ._events import (
    BALANCED,
    REFOCUSED,
    SPOILED,
    UNBALANCED,
    Delay,
    Dephase,
    EventAction,
    EventOperators,
    EventType,
    Excitation,
    FSEReadout,
    Inversion,
    Operator,
    PackedEvents,
    Readout,
    Refocusing,
    RfUse,
    SequenceEvent,
    SPGRReadout,
    Spoil,
    SSFPFidReadout,
    bSSFPReadout,
    compose,
    module,
    pack,
)

__all__ = [
    "BALANCED", "REFOCUSED", "SPOILED", "UNBALANCED", "Delay", "Dephase", "EventAction",
    "EventOperators", "EventType", "Excitation", "FSEReadout", "Inversion", "Operator",
    "PackedEvents", "Readout", "Refocusing", "RfUse", "SequenceEvent", "SPGRReadout", "Spoil",
    "SSFPFidReadout", "bSSFPReadout", "compose", "module", "pack",
]
