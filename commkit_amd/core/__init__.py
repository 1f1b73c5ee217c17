"""Signal records read by the receive tail, and random payloads."""

from .generation import generate
from .signal import Reference, Signal

__all__ = ["Reference", "Signal", "generate"]
