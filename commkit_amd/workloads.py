"""Seeded QC-LDPC codes and synthetic coded frames for tests and benchmarks.

Codes are drawn from a seed (``numpy.random.Generator(PCG64(seed))``): a
base matrix of ``mb x nb`` circulants of size ``Z`` with shifts uniform in
``[0, Z)``, redrawn until every block row has at least two and at most 32
blocks (the compressed min-sum state holds 32 edges; the reference requires
at least two bits per check, ``commkit/coding/ldpc.py:106-107``), the column
rule holds, and the parity part (the last ``mb`` block columns) is
invertible over GF(2), so the code encodes by the ``inverse`` rule with the
information bits first and ``k = n - m``.  Two column rules:

- ``density``: each entry present with probability ``p``, every column
  weight at least 2;
- ``weights``: each column weight uniform in ``{3, 4}``.

Frames follow ``commkit/core/generation.py:167-405`` and
``commkit/impairments/noise.py:38-186,638-672``: uniform info bits, the
code's encoder, Gray mapping of MSB-first groups, then circular complex
AWGN with ``Es`` measured per frame as ``mean(|x|^2)``, ``N0 = Es /
10^(EsN0/10)`` and ``sqrt(N0/2)`` per component.  Noise is drawn on the GPU
from torch's Philox generator seeded per call.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache
from typing import Any

import numpy as np

from .coding.ldpc import LDPCCode
from .mapping.constellation import Constellation

__all__ = ["CODES", "seeded_base", "seeded_code", "Frames", "make_frames", "code_by_name",
           "MIXED_CODES", "MIXED_ORDERS", "Mixed", "MixedGroup", "make_mixed", "PasFrames",
           "make_pas_frames"]


@dataclass(frozen=True)
class QcSpec:
    """A seeded quasi-cyclic code: shape, lifting, column rule and seed."""

    name: str
    base_rows: int
    base_cols: int
    lifting: int
    seed: int
    rule: str  # "density" or "weights"
    density: float = 0.0


#: The seeded codes of the workloads (rate = 1 - mb/nb).
CODES: dict[str, QcSpec] = {
    "r1/2-z27": QcSpec("r1/2-z27", 12, 24, 27, 101, "density", 0.4),
    "r2/3-z96": QcSpec("r2/3-z96", 8, 24, 96, 102, "weights"),
    "r5/6-z360": QcSpec("r5/6-z360", 10, 60, 360, 103, "weights"),
    "r4/5-z360": QcSpec("r4/5-z360", 10, 50, 360, 104, "weights"),
}

_MAX_ROW_WEIGHT = 32


def _draw(spec: QcSpec, rng: np.random.Generator) -> np.ndarray:
    mb, nb, z = spec.base_rows, spec.base_cols, spec.lifting
    if spec.rule == "density":
        present = rng.random((mb, nb)) < spec.density
    else:
        present = np.zeros((mb, nb), dtype=bool)
        for j in range(nb):
            w = int(rng.integers(3, 5))
            present[rng.choice(mb, size=w, replace=False), j] = True
    shifts = rng.integers(0, z, size=(mb, nb))
    return np.where(present, shifts, -1).astype(np.int64)


def _acceptable(spec: QcSpec, base: np.ndarray) -> bool:
    rw = (base >= 0).sum(axis=1)
    cw = (base >= 0).sum(axis=0)
    if rw.min() < 2 or rw.max() > _MAX_ROW_WEIGHT:
        return False
    if spec.rule == "density" and cw.min() < 2:
        return False
    return True


def _parity_invertible(base: np.ndarray, z: int) -> bool:
    """Whether the parity part of the lifted H has full rank over GF(2)."""
    from .coding.gf2 import rank

    mb, nb = base.shape
    m = mb * z
    hp = np.zeros((m, m), dtype=np.int8)
    r = np.arange(z)
    for i in range(mb):
        for jj, j in enumerate(range(nb - mb, nb)):
            s = base[i, j]
            if s >= 0:
                hp[i * z + r, jj * z + (r + s) % z] = 1
    return rank(hp) == m


@lru_cache(maxsize=16)
def seeded_base(name: str) -> np.ndarray:
    """The base matrix of a named seeded code (first acceptable draw)."""
    spec = CODES[name]
    rng = np.random.Generator(np.random.PCG64(spec.seed))
    for _ in range(10_000):
        base = _draw(spec, rng)
        if _acceptable(spec, base) and _parity_invertible(base, spec.lifting):
            base.setflags(write=False)
            return base
    raise RuntimeError(f"no acceptable base matrix for {name} in 10000 draws")


@lru_cache(maxsize=16)
def seeded_code(name: str) -> LDPCCode:
    """The named seeded code as an :class:`LDPCCode` (4-cycle warnings silenced)."""
    import logging

    log = logging.getLogger("commkit_amd")
    level = log.level
    log.setLevel(logging.ERROR)
    try:
        code = LDPCCode.quasi_cyclic(seeded_base(name), lifting=CODES[name].lifting)
    finally:
        log.setLevel(level)
    assert code._encoder.kind == "inverse" and code.k == code.num_variables - code.num_checks
    return code


def code_by_name(name: str) -> LDPCCode:
    return seeded_code(name)


@dataclass
class Frames:
    """Synthetic coded frames on one device.

    info (F, k) int8; code_bits (F, n) int8; tx (F, n/km) complex64;
    rx (F, n/km) complex64 received symbols; esn0_db (F,) float64 host;
    noise_var (F,) float64 host (nominal ``10^(-EsN0/10)``, unit-power
    constellation, as the reference's tests pass to compute_llr).
    """

    info: Any
    code_bits: Any
    tx: Any
    rx: Any
    esn0_db: np.ndarray
    noise_var: np.ndarray


def make_frames(code: LDPCCode, constellation: Constellation, esn0_db: Any, num_frames: int,
                *, seed: int = 0, device: Any = None, chunk: int = 1024) -> Frames:
    """``num_frames`` codewords of uniform info bits through Gray mapping and AWGN.

    ``esn0_db`` is a scalar or one value per frame.  Runs on ``device``
    (default: the current GPU) with the library's encoder and mapper.
    """
    import torch

    from .coding.ldpc import _encode

    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    km = constellation.bits_per_symbol
    if code.n % km:
        raise ValueError(f"n = {code.n} is not a multiple of {km} bits per symbol")
    es = np.broadcast_to(np.asarray(esn0_db, dtype=np.float64), (num_frames,)).copy()
    rng = np.random.Generator(np.random.PCG64(seed))
    info = torch.from_numpy(rng.integers(0, 2, size=(num_frames, code.k), dtype=np.int8)).to(dev)
    words = torch.empty((num_frames, code.n), dtype=torch.int8, device=dev)
    for r0 in range(0, num_frames, chunk):
        words[r0 : r0 + chunk] = _encode(code, info[r0 : r0 + chunk])
    tx = constellation.map(words)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed * 1_000_003 + 7)
    es_meas = (tx.abs() ** 2).mean(dim=1, keepdim=True).to(torch.float64)
    n0 = es_meas / torch.from_numpy(10.0 ** (es / 10.0)).to(dev)[:, None]
    std = torch.sqrt(n0 / 2.0).to(torch.float32)
    noise = torch.complex(
        torch.randn(tx.shape, generator=gen, device=dev, dtype=torch.float32),
        torch.randn(tx.shape, generator=gen, device=dev, dtype=torch.float32),
    )
    rx = (tx + noise * std).to(torch.complex64)
    return Frames(info, words, tx, rx, es, 10.0 ** (-es / 10.0))


#: Codes and constellation orders of the mixed workload.
MIXED_CODES = ("r5/6-z360", "r2/3-z96", "r1/2-z27")
MIXED_ORDERS = (4, 16, 64, 256)


@dataclass
class MixedGroup:
    """Frames of one (code, constellation) pair of a mixed workload: rows
    ``[row0, row1)`` of their code's class."""

    code_index: int
    order: int
    row0: int
    row1: int
    frames: Frames


@dataclass
class Mixed:
    """A mixed workload: per code (``MIXED_CODES`` order) its frame count,
    and the (code, constellation) groups that fill its rows."""

    codes: list
    rows: list
    groups: list

    @property
    def info_bits(self) -> int:
        return sum(c.k * r for c, r in zip(self.codes, self.rows))


def make_mixed(num_frames: int = 10_000, *, seed: int = 0, device: Any = None,
               low_db: float = -0.5, high_db: float = 1.0) -> Mixed:
    """Frames of uniformly drawn codes (``MIXED_CODES``) and Gray QAM orders
    (``MIXED_ORDERS``), each at its own Es/N0 uniform in ``[wf + low_db,
    wf + high_db]`` around the pair's measured waterfall.  Within a code the
    frames are ordered by constellation."""
    from .operating import waterfall_db

    rng = np.random.Generator(np.random.PCG64(seed))
    ci = rng.integers(0, len(MIXED_CODES), size=num_frames)
    oi = rng.integers(0, len(MIXED_ORDERS), size=num_frames)
    off = rng.uniform(low_db, high_db, size=num_frames)
    codes = [seeded_code(n) for n in MIXED_CODES]
    rows = [int((ci == i).sum()) for i in range(len(codes))]
    groups = []
    for i, name in enumerate(MIXED_CODES):
        r0 = 0
        for j, order in enumerate(MIXED_ORDERS):
            sel = (ci == i) & (oi == j)
            cnt = int(sel.sum())
            if cnt == 0:
                continue
            es = waterfall_db(name, order) + off[sel]
            fr = make_frames(codes[i], Constellation.qam(order), es, cnt,
                             seed=seed * 1000 + 10 * i + j + 1, device=device)  # fmt: skip
            groups.append(MixedGroup(i, order, r0, r0 + cnt, fr))
            r0 += cnt
    return Mixed(codes, rows, groups)


@dataclass
class PasFrames:
    """Codewords of probabilistic amplitude shaping.

    ``pas`` the :class:`~commkit_amd.coding.PAS` code; info (F, pas.k) int8
    uniform payload bits; sent (F, n) int8 transmitted (label) bits;
    ``perm`` (n,) int64 device: transmitted bit ``t`` is codeword bit
    ``perm[t]``, and ``iperm`` its inverse (codeword LLRs are the received
    bit LLRs gathered by ``iperm``); tx, rx (F, n / k_s) complex64;
    esn0_db, noise_var (F,) float64 host.
    """

    pas: Any
    info: Any
    sent: Any
    perm: Any
    iperm: Any
    tx: Any
    rx: Any
    esn0_db: np.ndarray
    noise_var: np.ndarray


def make_pas_frames(code: LDPCCode, constellation: Constellation, esn0_db: Any,
                    num_frames: int, *, seed: int = 0, device: Any = None) -> PasFrames:
    """Codewords of probabilistic amplitude shaping over ``code``: uniform
    payload bits through the CCDM amplitude matcher, the code and the
    sign/amplitude bit layout (``commkit/coding/shaping.py:391-397``, on the
    GPU), Gray mapping and AWGN as :func:`make_frames`."""
    import torch

    from .coding.shaping import PAS

    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    pas = PAS(code, constellation)
    rng = np.random.Generator(np.random.PCG64(seed))
    info = torch.from_numpy(rng.integers(0, 2, size=(num_frames, pas.k), dtype=np.int8)).to(dev)
    sent = torch.empty((num_frames, code.n), dtype=torch.int8, device=dev)
    for r0 in range(0, num_frames, 1024):
        sent[r0 : r0 + 1024] = pas._encode(info[r0 : r0 + 1024])
    perm, iperm = pas._tables(dev)
    tx = constellation.map(sent)
    es = np.broadcast_to(np.asarray(esn0_db, dtype=np.float64), (num_frames,)).copy()
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed * 1_000_003 + 11)
    es_meas = (tx.abs() ** 2).mean(dim=1, keepdim=True).to(torch.float64)
    n0 = es_meas / torch.from_numpy(10.0 ** (es / 10.0)).to(dev)[:, None]
    std = torch.sqrt(n0 / 2.0).to(torch.float32)
    noise = torch.complex(
        torch.randn(tx.shape, generator=gen, device=dev, dtype=torch.float32),
        torch.randn(tx.shape, generator=gen, device=dev, dtype=torch.float32),
    )
    rx = (tx + noise * std).to(torch.complex64)
    nv = constellation.power() * 10.0 ** (-es / 10.0)
    return PasFrames(pas, info, sent, perm, iperm, tx, rx, es, nv)
