"""Writes tests/golden/polymul_n2048_k2.npz: the n=2048, k=2 negacyclic product
(the reference's (q120, n=2048) test point, test_bfv_mode.cu:1492-1508).

Inputs come from the Philox uniform sampler (seed 2048, streams 0 and 1).
The output is computed with the oracle's DGT path and accepted only if it
equals the schoolbook negacyclic product residue by residue.
Run from the repository root:  python tests/golden/gen_polymul_n2048_k2.py
"""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
from oracle import oracle as O  # noqa: E402

SEED = 2048


def main():
    P = O.Params(2, 3, 2048, 256)
    a = O.sample_uniform(P.q, P.n, SEED, 0)
    b = O.sample_uniform(P.q, P.n, SEED, 1)
    c = O.unfold(O.polymul_batch(P, O.fold(a, P.Ng)[None], O.fold(b, P.Ng)[None])[0])
    for r, p in enumerate(P.q):
        assert np.array_equal(c[r], O.schoolbook(a[r], b[r], p))
    np.savez_compressed(os.path.join(REPO, "tests", "golden", "polymul_n2048_k2.npz"),
                        seed=np.uint64(SEED), a=a, b=b, c=c)
    print("ok")


if __name__ == "__main__":
    main()
