"""BFV word decomposition (xi, rho) and the dot product against the oracle,
bit for bit; untouched residues keep their contents; the reference's
identity sum_i rho(a)[i] * xi(b)[i] = a * b."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_res(primes, ng, seed, lead=()):
    rng = np.random.default_rng(seed)
    out = np.empty(tuple(lead) + (len(primes), ng, 2), dtype=np.uint64)
    for r, p in enumerate(primes):
        out[..., r, :, :] = rng.integers(0, p, size=tuple(lead) + (ng, 2), dtype=np.uint64)
    return out


@pytest.mark.parametrize("k,n", [(2, 2048), (4, 16384), (8, 4096)])
def test_xi_rho_dot_match_oracle(gpu, k, n):
    e = cp.CUDAEngine(k, k + 1, n, 65537)
    ctx = cp.Context(e)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P.q, P.Ng, 10 + k)
    b = rand_res(P.q, P.Ng, 20 + k)
    sentinel = np.full(a.shape, 7, dtype=np.uint64)
    outs = [cp.to_device(sentinel) for _ in range(k)]
    cp.execute_xi_bfv(ctx, outs, cp.to_device(b))
    want = O.xi_bfv(P, b, [sentinel] * k)
    X = [cp.to_host(o) for o in outs]
    for i in range(k):
        assert np.array_equal(X[i], want[i]), ("xi", i)
    outs = [cp.to_device(sentinel) for _ in range(k)]
    cp.execute_rho_bfv(ctx, outs, cp.to_device(a))
    want = O.rho_bfv(P, a, [sentinel] * k)
    R = [cp.to_host(o) for o in outs]
    for i in range(k):
        assert np.array_equal(R[i], want[i]), ("rho", i)
    # identity through dot on zero-initialised decompositions
    Rz = O.rho_bfv(P, a)
    Xz = O.xi_bfv(P, b)
    out = cp.empty(a.shape)
    cp.dot(ctx, out, [cp.to_device(x) for x in Rz], [cp.to_device(x) for x in Xz])
    assert np.array_equal(cp.to_host(out), O.pointwise(P, a, b, op=2))


@pytest.mark.parametrize("count,base", [(1, 0), (14, 2), (31, 1)])
def test_dot_long_vectors_match_oracle(gpu, count, base):
    """more operands than one launch takes (chunks of 14), every base"""
    k, n = 3, 4096
    e = cp.CUDAEngine(k, k + 1, n, 65537)
    ctx = cp.Context(e)
    P = O.Params(k, k + 1, n, 65537)
    pr = P.base_primes(base)
    A = rand_res(pr, P.Ng, 100 + count, (count,))
    B = rand_res(pr, P.Ng, 200 + count, (count,))
    out = cp.empty(A.shape[1:])
    cp.dot(ctx, out, [cp.to_device(x) for x in A], [cp.to_device(x) for x in B], base=base)
    assert np.array_equal(cp.to_host(out), O.dot(P, list(A), list(B), base=base))
