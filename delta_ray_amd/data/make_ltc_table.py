"""Convert the fitted LTC matrix table into the binary blob the renderer loads.

The renderer samples points on silhouette edges with linearly transformed
cosines (LTC).  The matrices come from an offline Nelder-Mead fit of the
Blinn-Phong lobe on a 128 x 128 (theta, roughness) grid.  The reference keeps
the fitted values as C++ source text (``ltc.inc``, array ``tabM``, lines
4-16389; lookup in ``material.h:464-489``: entry ``rid + tid * 128``, each a
row-major 3x3 matrix).  Parity needs the very same numbers, so this script
parses that text once and writes them as little-endian float32, 16384 x 9,
into ``ltc_tabM.f32`` beside this file.  A SHA-256 of the blob is written to
``ltc_tabM.sha256`` so a load can verify it.

Only ``tabM`` is used on the hot path; ``tabMinv`` and ``tabAmplitude`` are
not read by the reference renderer and are not converted.

Usage: python make_ltc_table.py [/path/to/ltc.inc]
"""
import hashlib
import os
import re
import sys

import numpy as np

SIZE = 128


def parse_tabM(text):
    start = text.index("tabM[size*size]")
    end = text.index("};", start)
    body = text[text.index("{", start) + 1:end]
    nums = re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?|[-+]?\d+", body)
    arr = np.array([float(x) for x in nums], dtype=np.float32)
    if arr.size != SIZE * SIZE * 9:
        raise ValueError("expected %d values, got %d" % (SIZE * SIZE * 9, arr.size))
    return arr.reshape(SIZE * SIZE, 9)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/ltc.inc"
    with open(src) as f:
        tab = parse_tabM(f.read())
    here = os.path.dirname(os.path.abspath(__file__))
    blob = tab.astype("<f4").tobytes()
    with open(os.path.join(here, "ltc_tabM.f32"), "wb") as f:
        f.write(blob)
    digest = hashlib.sha256(blob).hexdigest()
    with open(os.path.join(here, "ltc_tabM.sha256"), "w") as f:
        f.write(digest + "\n")
    print("wrote", tab.shape, digest)


if __name__ == "__main__":
    main()
