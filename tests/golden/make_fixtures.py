"""Extract the reference's own test vectors into small committed fixtures.

Reads, as text, the two 26,183-dimensional count vectors that
src/tests/testmsr.cpp:89-90 (and testmsrcsr.cpp:70-71) define inline, and the
mergetest.cpp:11-29 inputs, and stores them as sparse (index, value) arrays in
tests/golden/testmsr_pair.npz.  Only numbers are stored; no source text.
Run once in a container that has the reference checked out at
/root/reference:  python tests/golden/make_fixtures.py
"""
import os
import re

import numpy as np

REF = "/root/reference/src/tests/testmsr.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "testmsr_pair.npz")


def _vec(line):
    body = line[line.index("{") + 1: line.rindex("}")]
    return np.array([float(t) for t in re.split(r"\s*,\s*", body.strip()) if t], dtype=np.float64)


def main():
    with open(REF) as fh:
        lines = fh.read().splitlines()
    lhs, rhs = _vec(lines[88]), _vec(lines[89])
    assert lhs.shape == rhs.shape
    li, ri = np.nonzero(lhs)[0], np.nonzero(rhs)[0]
    np.savez(OUT, d=np.int64(lhs.size),
             lhs_idx=li.astype(np.uint32), lhs_val=lhs[li].astype(np.float32),
             rhs_idx=ri.astype(np.uint32), rhs_val=rhs[ri].astype(np.float32))
    print(OUT, lhs.size, len(li), len(ri), lhs.sum(), rhs.sum())


if __name__ == "__main__":
    main()
