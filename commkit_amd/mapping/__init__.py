"""Constellations and soft demapping (bit LLRs) on the GPU."""

from .constellation import Constellation
from .llr import compute_llr

__all__ = ["Constellation", "compute_llr"]
