"""commkit_amd: the soft-decision receive tail of commkit on MI355X (gfx950).

Bit-LLR demapping (:func:`compute_llr`, :meth:`Constellation.llr`) feeding
QC-LDPC decoding (:func:`decode` with :class:`LDPCCode`, :class:`MinSum`,
:class:`SumProduct`), with commkit's Python signatures and its
array-in, typed-result-out contract.  The arithmetic runs in hand-written
HIP kernels of ``libck_hip.so`` (C ABI in ``include/ck_tail.h``), loaded
with ctypes; device arrays are PyTorch-ROCm tensors (or any DLPack GPU
array), and the results stay on the input's device.
"""

from ._array import PrecisionWarning
from .coding import DecodeResult, LDPCCode, MinSum, SumProduct, decode, encode
from .core import Reference, Signal
from .mapping import Constellation, compute_llr

__all__ = [
    "Constellation",
    "DecodeResult",
    "LDPCCode",
    "MinSum",
    "PrecisionWarning",
    "Reference",
    "Signal",
    "SumProduct",
    "compute_llr",
    "decode",
    "encode",
]
__version__ = "0.1.0"
