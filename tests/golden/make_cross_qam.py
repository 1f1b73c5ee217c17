"""Writes cross_qam.json: the reference's Gray-labelled cross-QAM points.

The quasi-Gray labelling of cross QAM (commkit/mapping/gray.py:341-425) is
a table builder that no closed form or invariant fixes label by label, so
its output is taken from the reference itself: this script imports
``commkit.mapping.gray`` from the reference tree (in the build container
only; the reference does not travel) and records, for 32- and 128-QAM, the
point of every label before normalisation (odd integers) as [re, im] pairs.
Run: ``python tests/golden/make_cross_qam.py /root/reference``.
"""

import json
import os
import sys


def main(ref: str) -> None:
    sys.dont_write_bytecode = True  # leave the reference tree untouched
    sys.path.insert(0, ref)
    from commkit.mapping import gray  # noqa: E402  (the reference, read-only)

    out = {}
    for order in (32, 128):
        pts = gray._gray_qam_cross(order)
        out[str(order)] = [[float(p.real), float(p.imag)] for p in pts]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cross_qam.json")
    with open(path, "w") as f:
        json.dump(out, f)
        f.write("\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
