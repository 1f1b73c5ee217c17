"""LDPC codes, their encoder and their HIP decoders; distribution matching."""

from .base import DecodeResult, decode, encode
from .ldpc import LDPCCode, MinSum, SumProduct
from .shaping import CCDM, PAS

__all__ = ["CCDM", "DecodeResult", "LDPCCode", "MinSum", "PAS", "SumProduct", "decode",
           "encode"]
