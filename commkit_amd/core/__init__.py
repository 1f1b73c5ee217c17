"""Signal records read by the receive tail."""

from .signal import Signal

__all__ = ["Signal"]
