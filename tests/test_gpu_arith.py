"""The gfx950 forms of the arithmetic primitives (inline asm with carry-outs
and a MODE-register switch) against exact integer and rational arithmetic,
through cupoly_arith_selftest.  Operands include the words that make every
carry of the chains fire (all-ones halves, p - 1, 2^63 - 1) beside random
ones, for the smallest and largest primes of the bucket and one between."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest
import torch

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

M64 = (1 << 64) - 1
KMUL = 0x1D8E4E27C47D124F


def rd(fr):
    """Fraction -> the double below or equal to it (round toward -infinity)."""
    f = float(fr)
    if Fraction(f) > fr:
        f = math.nextafter(f, -math.inf)
    return f


def bits(f):
    return int(np.array([f], dtype=np.float64).view(np.uint64)[0])


def expected(x, y, w, z, p):
    w %= p
    out = [(x * y) >> 64, (z * y) >> 64] * 2
    out += [x * w % p, z * w % p]
    out += [x * (y >> 1) % p, (z % p) * y % p]
    bs = [((y >> 1) ^ ((k * KMUL) & M64)) & ((1 << 63) - 1) for k in range(6)]
    out += [sum(x * b for b in bs) % p, sum(z * b for b in bs) % p]
    r = Fraction(y >> 11, 1 << 63)
    out += [bits(rd(Fraction(x >> 11) * r + Fraction(z >> 11, 1 << 52))),
            bits(rd(Fraction(z >> 11) * r + Fraction(x >> 11, 1 << 52)))]
    return out


def records(p, count, seed):
    rng = random.Random(seed)
    edge = [0, 1, 2, (1 << 32) - 1, 1 << 32, (1 << 32) + 1, p - 1, p, p + 1, (1 << 63) - 1,
            1 << 63, M64, M64 - 1, 0xFFFFFFFF00000000, 0x00000000FFFFFFFF | (1 << 63)]
    recs = [(a, b, c, d) for a in edge for b in edge[::2] for c, d in ((p - 1, a), (a, M64))]
    while len(recs) < count:
        recs.append(tuple(rng.getrandbits(64) for _ in range(4)))
    return recs


@pytest.mark.parametrize("which", [0, 14, 29])
def test_device_primitives_match_exact_arithmetic(gpu, which):
    p = O.tables()["primes"][which]
    eng = cp.CUDAEngine(1, 1, 2048, 2)
    ctx = cp.Context(eng)
    recs = records(p, 3000, which)
    inp = cp.to_device(np.array(recs, dtype=np.uint64))
    out = cp.to_host(cp.arith_selftest(ctx, inp, p))
    torch.cuda.synchronize()
    names = ["mulhi2v x", "mulhi2v z", "mulhi2 x", "mulhi2 z", "shoup2v x", "shoup2v z",
             "mulmod2 x", "mulmod2 z", "acc_mac2 x", "acc_mac2 z", "fma_rd x", "fma_rd z"]
    for i, (x, y, w, z) in enumerate(recs):
        want = expected(x, y, w, z, p)
        got = [int(v) for v in out[i]]
        for j in range(12):
            assert got[j] == want[j], (names[j], hex(x), hex(y), hex(w), hex(z), hex(got[j]),
                                       hex(want[j]))
    ctx.destroy()
    eng.destroy()


def test_selftest_rejects_a_prime_outside_the_bucket(gpu):
    eng = cp.CUDAEngine(1, 1, 2048, 2)
    ctx = cp.Context(eng)
    with pytest.raises(cp.CupolyError):
        cp.arith_selftest(ctx, cp.empty((1, 4)), (1 << 61) - 1)
    ctx.destroy()
    eng.destroy()
