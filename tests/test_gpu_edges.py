"""Size-independent properties of the batched product at full ring sizes,
extreme residues and empty batches (the oracle is too slow for whole batches
at n = 2^14 and 2^15; these checks need no oracle)."""
import ctypes

import numpy as np
import pytest
import torch

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_engines = {}


def engine(k, n):
    if (k, n) not in _engines:
        e = cp.CUDAEngine(k, k + 1, n, 65537)
        _engines[(k, n)] = (e, cp.Context(e))
    return _engines[(k, n)]


def rand_coefs(q, shape, n, seed):
    rng = np.random.default_rng(seed)
    out = np.empty(tuple(shape) + (len(q), n), dtype=np.uint64)
    for r, p in enumerate(q):
        out[..., r, :] = rng.integers(0, p, size=tuple(shape) + (n,), dtype=np.uint64)
    return out


def negacyclic_shift(a, s, q):
    """a * X^s mod (X^n + 1, q): coefficients move up by s, wrapping with a sign."""
    n = a.shape[-1]
    out = np.empty_like(a)
    for r, p in enumerate(q):
        row = a[r]
        out[r, s:] = row[:n - s]
        wrapped = row[n - s:]
        out[r, :s] = np.where(wrapped == 0, np.uint64(0), np.uint64(p) - wrapped)
    return out


@pytest.mark.parametrize("n,k,batch", [(2048, 2, 8), (16384, 4, 64), (32768, 3, 90)])
def test_products_by_monomials_rotate(gpu, n, k, batch):
    """b = X^s (s per product, including 0, 1, n/2 and n - 1): the product is
    the negacyclic rotation of a.  n = 2^15 with 90 products (270 residue
    products) takes the one-workgroup-per-product kernel."""
    e, ctx = engine(k, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_coefs(P.q, (batch,), n, 900 + n)
    rng = np.random.default_rng(901 + n)
    shifts = [0, 1, n // 2, n - 1] + [int(s) for s in rng.integers(0, n, size=batch - 4)]
    b = np.zeros_like(a)
    for i, s in enumerate(shifts):
        b[i, :, s] = 1
    da = cp.to_device(O.fold(a, P.Ng))
    db = cp.to_device(O.fold(b, P.Ng))
    out = cp.empty(da.shape)
    cp.polymul_batch(ctx, out, da, db)
    h = O.unfold(cp.to_host(out))
    for i, s in enumerate(shifts):
        assert np.array_equal(h[i], negacyclic_shift(a[i], s, P.q)), (i, s)


@pytest.mark.parametrize("n,k", [(2048, 2), (16384, 4), (32768, 3)])
def test_all_minus_one_squared(gpu, n, k):
    """a = b = -1 in every coefficient: c_j = (j + 1) - (n - 1 - j) = 2j + 2 - n."""
    e, ctx = engine(k, n)
    P = O.Params(k, k + 1, n, 65537)
    batch = 96 if n == 32768 else 2
    a = np.empty((batch, k, n), dtype=np.uint64)
    for r, p in enumerate(P.q):
        a[:, r, :] = p - 1
    da = cp.to_device(O.fold(a, P.Ng))
    out = cp.empty(da.shape)
    cp.polymul_batch(ctx, out, da, da.clone())
    h = O.unfold(cp.to_host(out))
    j = np.arange(n, dtype=np.int64)
    for r, p in enumerate(P.q):
        want = np.array([(2 * int(x) + 2 - n) % p for x in j], dtype=np.uint64)
        for i in range(batch):
            assert np.array_equal(h[i, r], want), (i, r)


def test_product_is_linear_at_full_size(gpu):
    """(a + a') * b = a * b + a' * b mod q at n = 2^14, k = 4."""
    n, k, batch = 16384, 4, 16
    e, ctx = engine(k, n)
    P = O.Params(k, k + 1, n, 65537)
    a1 = rand_coefs(P.q, (batch,), n, 11)
    a2 = rand_coefs(P.q, (batch,), n, 12)
    b = rand_coefs(P.q, (batch,), n, 13)
    qv = np.array(P.q, dtype=np.uint64)[None, :, None]
    s = (a1 + a2) % qv  # both < 2^63: no wrap
    outs = []
    for x in (a1, a2, s):
        o = cp.empty((batch, k, P.Ng, 2))
        cp.polymul_batch(ctx, o, cp.to_device(O.fold(x, P.Ng)), cp.to_device(O.fold(b, P.Ng)))
        outs.append(O.unfold(cp.to_host(o)))
    assert np.array_equal((outs[0] + outs[1]) % qv, outs[2])


def test_empty_batches_are_no_ops(gpu):
    n, k = 2048, 2
    e, ctx = engine(k, n)
    ng = n // 2
    a = cp.empty((0, k, ng, 2))
    out = cp.empty((0, k, ng, 2))
    cp.polymul_batch(ctx, out, a, a)
    cp.polymul_plain(ctx, out, a, a)
    cp.sample_muladd_batch(ctx, out, a, 1, 0)
    ct = cp.empty((0, 2, k, ng, 2))
    cp.hom_mul_batch(ctx, cp.empty((0, 3, k, ng, 2)), ct, ct)
    cp.hom_mul_gather(ctx, [], [], [])
    ctx.sync()
    # a non-empty call afterwards still works
    one = np.zeros((1, k, n), dtype=np.uint64)
    one[:, :, 0] = 1
    P = O.Params(k, k + 1, n, 65537)
    x = rand_coefs(P.q, (1,), n, 5)
    o = cp.empty((1, k, ng, 2))
    cp.polymul_batch(ctx, o, cp.to_device(O.fold(x, ng)), cp.to_device(O.fold(one, ng)))
    assert torch.equal(o, cp.to_device(O.fold(x, ng)))


@pytest.mark.parametrize("n,k,batch", [(2048, 2, 4), (16384, 4, 4), (32768, 3, 4),
                                         (32768, 3, 90)])
def test_product_in_place(gpu, n, k, batch):
    """out may be a or b itself (each residue product reads its operands
    before it writes its result; n = 2^15 covers the split and the
    one-workgroup-per-product forms)."""
    e, ctx = engine(k, n)
    P = O.Params(k, k + 1, n, 65537)
    a = O.fold(rand_coefs(P.q, (batch,), n, 31), P.Ng)
    b = O.fold(rand_coefs(P.q, (batch,), n, 32), P.Ng)
    want = cp.empty(a.shape)
    cp.polymul_batch(ctx, want, cp.to_device(a), cp.to_device(b))
    da, db = cp.to_device(a), cp.to_device(b)
    cp.polymul_batch(ctx, da, da, db)  # out is a
    assert torch.equal(da, want)
    da, db = cp.to_device(a), cp.to_device(b)
    cp.polymul_batch(ctx, db, da, db)  # out is b
    assert torch.equal(db, want)
    i = batch - 1
    assert np.array_equal(cp.to_host(want)[i], O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0])


def test_oversized_batch_is_rejected_before_any_launch(gpu):
    """batch * k at or above 2^31 is an argument error (it used to wrap the
    unit count), and no kernel runs: the operand buffers are far too small."""
    e, ctx = engine(2, 2048)
    t = cp.empty((1, 2, 1024, 2))
    p = cp._ptr(t)
    assert cp.lib().cupoly_polymul_batch(ctx.handle, p, p, p, 1 << 30) < 0
    assert "2^31" in cp.lib().cupoly_last_error().decode()
    assert cp.lib().cupoly_sample_muladd_batch(ctx.handle, p, p, 1 << 30, 1, 0) < 0
    assert "2^31" in cp.lib().cupoly_last_error().decode()
    # hom-mul: at most 16383 pairs, in the contiguous and the gathered form
    big = 16384
    assert cp.lib().cupoly_hom_mul_batch(ctx.handle, p, p, p, big) < 0
    assert "16383" in cp.lib().cupoly_last_error().decode()
    tbl = (ctypes.c_void_p * big)(*([p.value] * big))
    assert cp.lib().cupoly_hom_mul_gather(ctx.handle, tbl, tbl, tbl, big) < 0
    assert "16383" in cp.lib().cupoly_last_error().decode()
    ctx.sync()
