"""Device poly_crt / poly_icrt against Python big integers (exact)."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def fold_ints(vals, primes, ng):
    flat = np.array([[v % p for v in vals] for p in primes], dtype=np.uint64)
    return O.fold(flat, ng)


@pytest.mark.parametrize("k,n", [(1, 2048), (4, 16384), (8, 4096), (14, 1024)])
def test_crt_icrt_match_big_integers(gpu, k, n):
    kl = min(k + 1, 14)
    e = cp.CUDAEngine(k, kl, n, 65537)
    ctx = cp.Context(e)
    P = O.Params(k, kl, n, 65537)
    rng = np.random.default_rng(k * 7 + n)
    limbs = k + 1
    top = 1 << (64 * limbs)
    # crt: arbitrary unsigned integers of `limbs` words, edge values first
    vals = [int.from_bytes(rng.bytes(8 * limbs), "little") for _ in range(n)]
    vals[:6] = [0, 1, top - 1, P.Q - 1, P.Q, P.Q + 1]
    for base in (cp.QBase, cp.BBase, cp.QBBase):
        out = cp.empty((e.get_n_residues(base), P.Ng, 2))
        cp.crt(ctx, out, cp.to_device(cp.ints_to_limbs(vals, limbs)), base=base)
        assert np.array_equal(cp.to_host(out), fold_ints(vals, P.base_primes(base), P.Ng)), base
    # fewer coefficients than n: the rest are zero
    m = n // 3
    out = cp.empty((k, P.Ng, 2))
    cp.crt(ctx, out, cp.to_device(cp.ints_to_limbs(vals[:m], limbs)))
    assert np.array_equal(cp.to_host(out), fold_ints(vals[:m] + [0] * (n - m), P.q, P.Ng))
    # icrt: random residues -> x in [0, Q)
    res = np.stack([rng.integers(0, q, size=(P.Ng, 2), dtype=np.uint64) for q in P.q])
    res[:, 0, 0] = 0
    res[:, 1, 0] = np.array([q - 1 for q in P.q], dtype=np.uint64)  # x = Q - 1
    ints = cp.empty((n, k))
    cp.icrt(ctx, ints, cp.to_device(res))
    got = cp.limbs_to_ints(cp.to_host(ints))
    want = O.crt(P, res)
    assert got == want
    assert got[0] == 0 and got[1] == P.Q - 1
    # round trip
    back = cp.empty((k, P.Ng, 2))
    cp.crt(ctx, back, ints)
    assert np.array_equal(cp.to_host(back), res)


def test_crt_rejects_bad_shapes(gpu):
    e = cp.CUDAEngine(2, 3, 2048, 65537)
    ctx = cp.Context(e)
    with pytest.raises(ValueError):
        cp.crt(ctx, cp.empty((2, 1024, 2)), cp.empty((4096, 2)))
    with pytest.raises(ValueError):
        cp.icrt(ctx, cp.empty((2048, 3)), cp.empty((2, 1024, 2)))
