"""Signal records read by the receive tail."""

from .signal import Reference, Signal

__all__ = ["Reference", "Signal"]
