"""Write ``oh_ops.json``: the 48 octahedral operation matrices of the
reference's OHOPS, in its order, read as data from the DATA statement of
``fastoverlap/f90/alignutils.f90`` (column-major RMAT(3,3,48)).  The order
matters: the Oh-cell search keeps the first strictly best operation.

    python tests/golden/make_oh_golden.py [/path/to/reference]
"""

import json
import os
import re
import sys

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oh_ops.json")


def main(ref="/root/reference"):
    path = os.path.join(ref, "fastoverlap", "f90", "alignutils.f90")
    text = open(path).read()
    start = text.index("SUBROUTINE OHOPS")
    block = text[start:text.index("/", text.index("DATA RMAT", start) + len("DATA RMAT /"))]
    body = block[block.index("DATA RMAT"):]
    vals = [float(v.replace("D", "e")) for v in re.findall(r"-?\d+\.\d+D0", body)]
    assert len(vals) == 9 * 48, len(vals)
    mats = []
    for o in range(48):
        col = vals[9 * o:9 * o + 9]  # RMAT(1,1), RMAT(2,1), RMAT(3,1), RMAT(1,2), ...
        mats.append([[col[r + 3 * c] for c in range(3)] for r in range(3)])
    with open(OUT, "w") as f:
        json.dump({"source": "fastoverlap/f90/alignutils.f90 OHOPS DATA RMAT (row-major "
                             "matrices, y = R x per atom)", "ops": mats}, f)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main(*sys.argv[1:])
