"""LDPC codes, their encoder and their HIP decoders."""

from .base import DecodeResult, decode, encode
from .ldpc import LDPCCode, MinSum, SumProduct

__all__ = ["DecodeResult", "LDPCCode", "MinSum", "SumProduct", "decode", "encode"]
