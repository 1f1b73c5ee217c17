"""The cross-lane move rule of the HIP kernels, checked on the CPU.

A DPP move reads its source lane whatever that lane's EXEC bit says, and a
disabled source reads as zero (or as the old value): a move under a mask
narrower than the lanes it reads returns garbage without any error.  The
kernels therefore keep one rule: every lane a move reads belongs to the same
lane group as the reader (the P lanes of one check row in the LDPC
decoders, the G lanes of one block in the CCDM coders), and every branch
around a move is uniform over that group, so the group is either all active
or all inactive.

This file models the DPP controls the sources use (quad_perm, row_shr,
row_mirror, row_half_mirror) and v_permlane32_swap, fails if a source uses a
control it does not model, and checks for every lane map of the kernels
(restated from ``qc_lane`` in ck_ldpc_dev.h and the group index of
ck_ccdm.hip) that each move's source lane lies in the reader's group.
"""

import re
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "commkit_amd" / "csrc"
WAVE = 64


def dpp_source(ctrl: int, lane: int) -> int | None:
    """Lane read by `lane` under DPP control `ctrl`; None for none (bound)."""
    if 0x00 <= ctrl <= 0xFF:  # quad_perm
        return (lane & ~3) | ((ctrl >> (2 * (lane & 3))) & 3)
    row = lane & ~15
    if 0x111 <= ctrl <= 0x11F:  # row_shr:n
        src = lane - (ctrl - 0x110)
        return src if src >= row else None
    if ctrl == 0x140:  # row_mirror
        return row + 15 - (lane & 15)
    if ctrl == 0x141:  # row_half_mirror
        return (lane & ~7) + 7 - (lane & 7)
    raise ValueError(f"DPP control {ctrl:#x} is not modelled")


def permlane32_source(lane: int) -> int:
    """v_permlane32_swap exchanges the two halves of the wave."""
    return lane ^ 32


def _source_controls() -> set[int]:
    """Every DPP control the sources pass, as integers."""
    found = set()
    for path in sorted(CSRC.glob("*.h*")):
        text = path.read_text()
        consts = {m.group(1): int(m.group(2), 16)
                  for m in re.finditer(r"\b(k\w+) = (0x[0-9A-Fa-f]+)", text)}  # fmt: skip
        for m in re.finditer(r"dpp(?:64)?<(\w+)>\(", text):
            arg = m.group(1)
            if arg in consts:
                found.add(consts[arg])
            elif arg.startswith("0x"):
                found.add(int(arg, 16))
            else:
                assert arg == "CTRL", f"{path.name}: unmodelled DPP control {arg}"
        for m in re.finditer(r"mov_dpp\(\w+, ([^,]+),", text):
            arg = m.group(1).strip()
            if arg == "J * 0x55":  # quad_lane<J>, J = 0..3
                found.update(j * 0x55 for j in range(4))
            else:
                found.add(int(arg, 16))
    return found


def test_every_source_control_is_modelled():
    controls = _source_controls()
    # quad permutes (xor 1, xor 2, broadcasts), mirrors and prefix shifts
    assert {0xB1, 0x4E, 0x00, 0x55, 0xAA, 0xFF, 0x140, 0x141, 0x111, 0x112, 0x114,
            0x118} <= controls  # fmt: skip
    for ctrl in controls:
        for lane in range(WAVE):
            dpp_source(ctrl, lane)


def qc_row_of(P: int, thread: int) -> int:
    """The check row (within the workgroup) of a thread: ck_ldpc_dev.h qc_lane."""
    if P == 2:
        return ((thread >> 6) << 5) | (thread & 31)
    return thread // P


@pytest.mark.parametrize("P", [2, 4])
def test_ldpc_moves_stay_in_the_check_row(P):
    """P = 4: lane_xor (quad_perm 1,0,3,2 and 2,3,0,1) and quad_lane<J>;
    P = 2: row_pair (permlane32 swap).  The row's lanes share frame and row,
    hence `done` and `lane_ok`, the conditions around the row updates."""
    threads = 12 * WAVE
    for t in range(threads):
        row = qc_row_of(P, t)
        wave = t & ~(WAVE - 1)
        if P == 4:
            for ctrl in (0xB1, 0x4E, 0x00, 0x55, 0xAA, 0xFF):
                assert qc_row_of(P, wave + dpp_source(ctrl, t % WAVE)) == row
        else:
            assert qc_row_of(P, wave + permlane32_source(t % WAVE)) == row
    # every row has exactly its P lanes
    rows = [qc_row_of(P, t) for t in range(threads)]
    assert all(rows.count(r) == P for r in set(rows))


@pytest.mark.parametrize("G", [2, 4, 8, 16])
def test_ccdm_moves_stay_in_the_block_group(G):
    """ck_ccdm.hip: block b = (workgroup * lanes + thread) / G, position
    j = thread & (G - 1).  The butterfly steps up to G (xor 1, xor 2, half
    mirror, mirror) read lanes of the same block; the prefix keeps a row
    shift by s only where j >= s, whose source is then in the block."""
    butterfly = [0xB1, 0x4E, 0x141, 0x140][: G.bit_length() - 1]
    for lane in range(WAVE):
        block, j = lane // G, lane & (G - 1)
        for ctrl in butterfly:
            assert dpp_source(ctrl, lane) // G == block
        # after the steps up to G every lane has combined all G lanes
        seen = {lane}
        for ctrl in butterfly:
            seen |= {dpp_source(ctrl, s) for s in seen}
        assert seen == set(range(block * G, block * G + G))
        for s in (1, 2, 4, 8):
            if s < G and j >= s:
                src = dpp_source(0x110 + s, lane)
                assert src is not None and src // G == block
