"""Build ``libck_hip.so`` in-tree for gfx950 with hipcc.

``python -m commkit_amd.build`` compiles every ``csrc/*.hip`` to an object
and links the shared library next to this file.  The LDPC translation unit
is compiled with ``-ffp-contract=off`` (min-sum must round every product
and sum on its own to stay bit-identical to the CPU restatement); every
unit carries ``-packed-fp32-ops`` off, so the compiler emits no packed FP32
instructions.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "hip")
LIB = os.path.join(HERE, "libck_hip.so")
ARCH = os.environ.get("CK_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Werror",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]  # fmt: skip

PER_FILE = {
    "ck_ldpc.hip": ["-ffp-contract=off"],
    "ck_batch.hip": ["-ffp-contract=off"],
}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP library needs ROCm's hipcc.")
    return path


def _compile(src: str) -> str:
    name = os.path.basename(src)
    obj = os.path.join(BUILD, name + ".o")
    cmd = [hipcc(), "-c", src, "-o", obj, *COMMON, *PER_FILE.get(name, [])]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {name}:\n{' '.join(cmd)}\n{res.stdout}{res.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    """Compile and link ``libck_hip.so``; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    sources = sorted(
        os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip")
    )
    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp, *objs, f"--offload-arch={ARCH}", "-fPIC"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources for {ARCH}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
