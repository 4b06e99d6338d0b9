"""GPU parity: every HIP entry point against the CPU oracle on the same seeded
inputs, bit for bit (integer results; the HPS FP64 steps too, since the oracle
restates them with the same IEEE-754 operations)."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_res(P, primes, shape, seed):
    """uniform residues, GI layout [..., nres, Ng, 2]"""
    rng = np.random.default_rng(seed)
    out = np.empty(tuple(shape) + (len(primes), P.Ng, 2), dtype=np.uint64)
    for r, p in enumerate(primes):
        out[..., r, :, :] = rng.integers(0, p, size=tuple(shape) + (P.Ng, 2), dtype=np.uint64)
    return out


_engines = {}


def engine(k, kl, n, t=65537):
    key = (k, kl, n, t)
    if key not in _engines:
        e = cp.CUDAEngine(k, kl, n, t)
        _engines[key] = (e, cp.Context(e))
    return _engines[key]


def run_dgt(ctx, x, base, direction):
    d = cp.to_device(x)
    cp.execute_dgt(ctx, d, base, direction)
    return cp.to_host(d)


@pytest.mark.parametrize("n", [32, 64, 256, 1024, 2048, 4096, 8192, 16384, 32768, 65536])
def test_dgt_forward_inverse_match_oracle(gpu, n):
    k = 2 if n <= 16384 else 1
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    for base in (cp.QBase, cp.BBase, cp.QBBase):
        x = rand_res(P, P.base_primes(base), (), n + base)
        f = run_dgt(ctx, x, base, cp.FORWARD)
        assert np.array_equal(f, O.dgt(P, x, base, 0)), ("forward", n, base)
        g = run_dgt(ctx, f, base, cp.INVERSE)
        assert np.array_equal(g, x), ("inverse", n, base)


@pytest.mark.parametrize("n,k,batch", [(32, 1, 5), (2048, 2, 3), (4096, 3, 2), (16384, 4, 2),
                                       (32768, 3, 1), (65536, 1, 1)])
def test_polymul_fused_and_plain_match_oracle(gpu, n, k, batch):
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 1 + n)
    b = rand_res(P, P.q, (batch,), 2 + n)
    want = O.polymul_batch(P, a, b)
    da, db = cp.to_device(a), cp.to_device(b)
    out = cp.empty(a.shape)
    cp.polymul_batch(ctx, out, da, db)
    assert np.array_equal(cp.to_host(out), want), "fused"
    assert np.array_equal(cp.to_host(da), a) and np.array_equal(cp.to_host(db), b)
    out2 = cp.empty(a.shape)
    cp.polymul_plain(ctx, out2, da, db)
    assert np.array_equal(cp.to_host(out2), want), "plain"


def test_polymul_n32768_persistent_matches_split_and_oracle(gpu):
    """n = 2^15 batches of at least one residue product per CU take the
    one-workgroup-per-product persistent kernel; smaller ones the two-kernel
    split.  The same 96 products (288 residue products) run both ways must
    agree everywhere, and three of them with the oracle."""
    n, k, batch = 32768, 3, 96
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 77)
    b = rand_res(P, P.q, (batch,), 78)
    da, db = cp.to_device(a), cp.to_device(b)
    whole = cp.empty(a.shape)
    cp.polymul_batch(ctx, whole, da, db)
    parts = cp.empty(a.shape)
    for i in range(0, batch, 16):  # 48 residue products per call
        cp.polymul_batch(ctx, parts[i:i + 16], da[i:i + 16], db[i:i + 16])
    h = cp.to_host(whole)
    assert np.array_equal(h, cp.to_host(parts))
    for i in (0, 41, 95):
        assert np.array_equal(h[i], O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0]), i


def test_polymul_n2048_golden_fixture(gpu):
    import os
    g = np.load(os.path.join(O.REPO, "tests", "golden", "polymul_n2048_k2.npz"))
    e, ctx = engine(2, 3, 2048, 256)
    a = cp.to_device(O.fold(g["a"], 1024)[None])
    b = cp.to_device(O.fold(g["b"], 1024)[None])
    out = cp.empty(a.shape)
    cp.polymul_batch(ctx, out, a, b)
    assert np.array_equal(O.unfold(cp.to_host(out)[0]), g["c"])


@pytest.mark.parametrize("base", [0, 1, 2])
def test_pointwise_ops_match_oracle(gpu, base):
    n = 2048
    e, ctx = engine(3, 4, n)
    P = O.Params(3, 4, n, 65537)
    pr = P.base_primes(base)
    a, b, c = (rand_res(P, pr, (), s) for s in (11, 12, 13))
    da, db, dc = cp.to_device(a), cp.to_device(b), cp.to_device(c)
    out = cp.empty(a.shape)
    for op, fn, args in [(0, cp.execute_add_dgt, (da, db)), (1, cp.execute_sub_dgt, (da, db)),
                         (2, cp.execute_mul_dgt_gi, (da, db)),
                         (3, cp.execute_muladd_dgt, (da, db, dc))]:
        fn(ctx, out, *args, base=base)
        assert np.array_equal(cp.to_host(out), O.pointwise(P, a, b, c, base=base, op=op)), op
    # dr2: ct21 = a0 b0, ct22 = a0 b1 + a1 b0, ct23 = a1 b1
    a1, b1 = rand_res(P, pr, (), 14), rand_res(P, pr, (), 15)
    o1, o2, o3 = cp.empty(a.shape), cp.empty(a.shape), cp.empty(a.shape)
    cp.execute_dr2_dgt(ctx, o1, o2, o3, da, cp.to_device(a1), db, cp.to_device(b1), base=base)
    assert np.array_equal(cp.to_host(o1), O.pointwise(P, a, b, base=base, op=2))
    cross = O.pointwise(P, O.pointwise(P, a, b1, base=base, op=2),
                        O.pointwise(P, a1, b, base=base, op=2), base=base, op=0)
    assert np.array_equal(cp.to_host(o2), cross)
    assert np.array_equal(cp.to_host(o3), O.pointwise(P, a1, b1, base=base, op=2))
    # negate, op_by_int
    cp.execute_polynomial_negate(ctx, out, da, base=base)
    want = np.stack([np.where(a[r] == 0, 0, np.uint64(p) - a[r]) for r, p in enumerate(pr)])
    assert np.array_equal(cp.to_host(out), want)
    x = 123456789
    for op in (cp.ADD, cp.SUB, cp.MUL):
        cp.execute_polynomial_op_by_int(ctx, out, da, x, op, base=base)
        const = np.zeros_like(a)
        const[:, :, 0] = x
        if op == cp.MUL:
            ref = O.pointwise(P, a, const, base=base, op=2)
        else:
            ref = O.pointwise(P, a, const, base=base, op=0 if op == cp.ADD else 1)
        assert np.array_equal(cp.to_host(out), ref), op
    # mul by a uint64 per slot
    u = np.random.default_rng(9).integers(0, 1 << 62, size=(len(pr), P.Ng), dtype=np.uint64)
    cp.execute_mul_dgt_u64(ctx, out, da, cp.to_device(u), base=base)
    g = np.zeros_like(a)
    g[..., 0] = u
    uu = np.stack([g[r] % np.uint64(p) for r, p in enumerate(pr)])
    assert np.array_equal(cp.to_host(out), O.pointwise(P, a, uu, base=base, op=2))


@pytest.mark.parametrize("k,n,t", [(1, 1024, 256), (2, 2048, 256), (4, 16384, 65537),
                                   (8, 4096, 65537)])
def test_hps_kernels_match_oracle(gpu, k, n, t):
    kl = k + 1
    e, ctx = engine(k, kl, n, t)
    P = O.Params(k, kl, n, t)
    xq = rand_res(P, P.q, (), 40 + k)
    # Q -> B (in place in a buffer of max(k, kl) residues)
    buf = np.zeros((max(k, kl), P.Ng, 2), dtype=np.uint64)
    buf[:k] = xq
    d = cp.to_device(buf)
    cp.execute_polynomial_basis_ext_Q_to_B(ctx, d)
    assert np.array_equal(cp.to_host(d)[:kl], O.basis_ext_q_to_b(P, xq))
    # complex scaling from QB residues of random x in [0, QB)
    xb = rand_res(P, P.b, (), 50 + k)
    out = cp.empty((kl, n))
    cp.execute_polynomial_complex_scaling(ctx, out, cp.to_device(xq), cp.to_device(xb))
    cs = O.complex_scaling(P, xq, xb)
    assert np.array_equal(cp.to_host(out), cs)
    # B -> Q on the complex-scaling output
    oq = cp.empty((k, P.Ng, 2))
    cp.execute_polynomial_basis_ext_B_to_Q(ctx, oq, out)
    assert np.array_equal(cp.to_host(oq), O.basis_ext_b_to_q(P, cs))
    d = cp.to_device(xq)
    cp.execute_polynomial_simple_scaling(ctx, d)
    assert np.array_equal(cp.to_host(d)[0], O.simple_scaling(P, xq))


@pytest.mark.parametrize("k", [8, 13])
def test_hps_base_changes_at_extreme_residues(gpu, k):
    """Q -> B and B -> Q with every residue at its largest value (p - 1) in
    half of the slots and 0 in a quarter, up to k = 13, kl = 14 (the largest
    bases): the largest sums the 160-bit accumulation of the base changes
    sees, against the oracle's modular chains."""
    kl, n, t = k + 1, 1024, 65537
    e, ctx = engine(k, kl, n, t)
    P = O.Params(k, kl, n, t)
    half, quarter = P.Ng // 2, P.Ng // 4
    xq = rand_res(P, P.q, (), 70 + k)
    for r, q in enumerate(P.q):
        xq[r, :half] = q - 1
        xq[r, half:half + quarter] = 0
    buf = np.zeros((max(k, kl), P.Ng, 2), dtype=np.uint64)
    buf[:k] = xq
    d = cp.to_device(buf)
    cp.execute_polynomial_basis_ext_Q_to_B(ctx, d)
    assert np.array_equal(cp.to_host(d)[:kl], O.basis_ext_q_to_b(P, xq))
    rng = np.random.default_rng(80 + k)
    xb = np.empty((kl, n), dtype=np.uint64)
    for i, b in enumerate(P.b):
        xb[i] = rng.integers(0, b, size=n, dtype=np.uint64)
        xb[i, : n // 2] = b - 1
        xb[i, n // 2: n // 2 + n // 4] = 0
    oq = cp.empty((k, P.Ng, 2))
    cp.execute_polynomial_basis_ext_B_to_Q(ctx, oq, cp.to_device(xb))
    assert np.array_equal(cp.to_host(oq), O.basis_ext_b_to_q(P, xb))
    # complex scaling at the same extremes, with the exact sums and with the
    # reference's addmod chain forced on every slot: both equal the oracle
    xbq = O.fold(xb, P.Ng)
    want = O.complex_scaling(P, xq, xbq)
    out = cp.empty((kl, n))
    try:
        for chain in (0, 1):
            cp.hom_mul_set_tuning(ctx, 0, chain)
            cp.execute_polynomial_complex_scaling(ctx, out, cp.to_device(xq), cp.to_device(xbq))
            assert np.array_equal(cp.to_host(out), want), chain
    finally:
        cp.hom_mul_set_tuning(ctx, 0, 0)


@pytest.mark.parametrize("k,n,t", [(4, 16384, 65537), (3, 4096, 256), (2, 2048, 40961)])
def test_simple_scaling_recovers_scaled_messages(gpu, k, n, t):
    """Decryption shape: x = floor(Q/t) m + e (|e| small) scaled by t/Q and
    rounded gives m back, for t = 65537 (not a power of two) as for 256."""
    e_, ctx = engine(k, k + 1, n, t)
    P = O.Params(k, k + 1, n, t)
    rng = np.random.default_rng(t + n)
    m = rng.integers(0, t, size=n)
    err = rng.integers(-1000, 1001, size=n)
    delta = P.Q // t
    x = [(delta * int(mi) + int(ei)) % P.Q for mi, ei in zip(m, err)]
    res = O.fold(np.array([[v % q for v in x] for q in P.q], dtype=np.uint64), P.Ng)
    d = cp.to_device(res)
    cp.execute_polynomial_simple_scaling(ctx, d)
    got = O.unfold(cp.to_host(d)[0])
    assert np.array_equal(got, m.astype(np.uint64))
    assert np.array_equal(cp.to_host(d)[0], O.simple_scaling(P, res))


def test_hps_poly_level_compositions(gpu):
    """poly_basis_extension_Q_to_B and poly_complex_scaling_tDivQ on DGT-domain
    operands, k = 4, n = 2^14, t = 65537 (the BFV-size shape)."""
    k, kl, n, t = 4, 5, 16384, 65537
    e, ctx = engine(k, kl, n, t)
    P = O.Params(k, kl, n, t)
    aq = rand_res(P, P.q, (), 61)
    aq_dgt = O.dgt(P, aq, 0, 0)
    ob = cp.empty((kl, P.Ng, 2))
    cp.poly_basis_extension_Q_to_B(ctx, ob, cp.to_device(aq_dgt))
    want_b = O.dgt(P, O.basis_ext_q_to_b(P, aq), 1, 0)
    assert np.array_equal(cp.to_host(ob), want_b)
    oq = cp.empty((k, P.Ng, 2))
    dq = cp.to_device(aq_dgt)
    cp.poly_complex_scaling_tDivQ(ctx, oq, dq, ob)
    want_q = O.dgt(P, O.basis_ext_b_to_q(P, O.complex_scaling(P, aq, O.basis_ext_q_to_b(P, aq))),
                   0, 0)
    assert np.array_equal(cp.to_host(oq), want_q)
    assert np.array_equal(cp.to_host(dq), aq_dgt)  # inputs preserved


SMALL_KIND = {cp.TERNARY: 2, cp.DISCRETE_GAUSSIAN: 3, cp.NARROW: 4, cp.BINARY: 5}


@pytest.mark.parametrize("kind,n", [(cp.UNIFORM, 4096), (cp.TERNARY, 4096),
                                    (cp.DISCRETE_GAUSSIAN, 4096), (cp.NARROW, 4096),
                                    (cp.BINARY, 4096), (cp.ZO, 4096), (cp.HAMMINGWEIGHT, 4096),
                                    (cp.ZO, 65536), (cp.HAMMINGWEIGHT, 32), (cp.ZO, 32),
                                    (cp.ZO, 16384), (cp.HAMMINGWEIGHT, 16384),
                                    (cp.HAMMINGWEIGHT, 32768)])
def test_sampler_matches_oracle_counters(gpu, kind, n):
    k = 3 if n <= 16384 else 1
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    out = cp.empty((k, P.Ng, 2))
    cp.Sampler.sample(ctx, out, kind, 777, 5, base=cp.QBase, coeff_domain=True)
    if kind == cp.UNIFORM:
        want = O.sample_uniform(P.q, n, 777, 5)
    else:
        if kind in (cp.ZO, cp.HAMMINGWEIGHT):
            h = n // 2 if kind == cp.ZO else min(cp.HAMMINGWEIGHT_H, n)
            vals = O.sample_fixed_weight(n, h, 777, 5)
            assert int(np.count_nonzero(vals)) == h
        else:
            vals = O.sample_small(n, SMALL_KIND[kind], 777, 5)
        want = O.small_to_residues(vals, P.q)
    assert np.array_equal(O.unfold(cp.to_host(out)), want)
    # DGT-domain variant = forward DGT of the same draws
    cp.Sampler.sample(ctx, out, kind, 777, 5, base=cp.QBase)
    assert np.array_equal(cp.to_host(out), O.dgt(P, O.fold(want, P.Ng), 0, 0))


def test_large_batch_properties(gpu):
    """n = 2^14, k = 4, 64 products: fused == plain, and a * 1 = a."""
    import torch
    k, n, batch = 4, 16384, 64
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 71)
    b = rand_res(P, P.q, (batch,), 72)
    da, db = cp.to_device(a), cp.to_device(b)
    o1, o2 = cp.empty(a.shape), cp.empty(a.shape)
    cp.polymul_batch(ctx, o1, da, db)
    cp.polymul_plain(ctx, o2, da.clone(), db.clone())
    assert torch.equal(o1, o2)
    one = np.zeros_like(b)
    one[:, :, 0, 0] = 1
    cp.polymul_batch(ctx, o2, da, cp.to_device(one))
    assert torch.equal(o2, da)
    # spot-check three products against the oracle
    h = cp.to_host(o1)
    for i in (0, 31, 63):
        assert np.array_equal(h[i], O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0])


def test_polymul_full_batch_properties(gpu):
    """n = 2^14, k = 4, 1024 products (the benchmarked launch): every fused
    product equals the one-launch-per-op form, a * b = b * a for every
    product, and two products equal the oracle."""
    import torch
    k, n, batch = 4, 16384, 1024
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 81)
    b = rand_res(P, P.q, (batch,), 82)
    da, db = cp.to_device(a), cp.to_device(b)
    o1, o2 = cp.empty(a.shape), cp.empty(a.shape)
    cp.polymul_batch(ctx, o1, da, db)
    cp.polymul_batch(ctx, o2, db, da)
    assert torch.equal(o1, o2)
    cp.polymul_plain(ctx, o2, da.clone(), db.clone())
    assert torch.equal(o1, o2)
    for i in (3, 1000):
        assert np.array_equal(cp.to_host(o1[i]), O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0]), i


@pytest.mark.parametrize("k,n,batch", [(1, 64, 3), (2, 2048, 3), (3, 4096, 2), (5, 8192, 1),
                                         (4, 16384, 2), (2, 32768, 1), (2, 2048, 48),
                                         (8, 32768, 3), (8, 32768, 4), (2, 32768, 13)])
def test_hom_mul_batch_matches_oracle(gpu, k, n, batch):
    """n = 2^14, k = 4, kl = 5, t = 65537 (and other rings): tensor product + HPS
    rescale of ciphertext pairs, bit-exact against the oracle composition.
    Batches under 2^16 output slots take the per-residue HPS kernels, the
    others (48 and 3 pairs among them) the per-slot ones.  At n = 2^15, batches
    with forward residue polynomials for three quarters of the CUs (3 x 4 x 17,
    4 x 4 x 17 and 13 x 4 x 5 on 256 CUs) take the one-workgroup-per-polynomial
    transforms, smaller ones (1 x 4 x 5) the split transforms."""
    e, ctx = engine(k, k + 1, n, 65537)
    P = O.Params(k, k + 1, n, 65537)
    ct1 = rand_res(P, P.q, (batch, 2), 90 + k)
    ct2 = rand_res(P, P.q, (batch, 2), 91 + k)
    out = cp.empty((batch, 3, k, P.Ng, 2))
    cp.hom_mul_batch(ctx, out, cp.to_device(ct1), cp.to_device(ct2))
    h = cp.to_host(out)
    for i in range(batch):
        assert np.array_equal(h[i], O.hom_mul(P, ct1[i], ct2[i])), i


@pytest.mark.parametrize("k,n,batch", [(4, 16384, 2), (2, 2048, 48), (3, 8192, 8),
                                         (8, 32768, 3), (1, 4096, 5)])
def test_hom_mul_scale_forms_agree(gpu, k, n, batch):
    """The t/Q scaling computes each residue as an exact sum reduced once; the
    reference's addmod chain (forced on every slot by scale_chain = 1) must give
    the same words.  Both HPS launch forms (one thread per slot, and one per
    (slot, residue), chosen through wide_slots) are run in both modes; all four
    results are equal and one pair equals the oracle composition."""
    import torch
    e, ctx = engine(k, k + 1, n, 65537)
    P = O.Params(k, k + 1, n, 65537)
    ct1 = cp.to_device(rand_res(P, P.q, (batch, 2), 700 + k))
    ct2 = cp.to_device(rand_res(P, P.q, (batch, 2), 710 + k))
    outs = []
    try:
        for wide_slots in (1, 1 << 40):
            for chain in (0, 1):
                cp.hom_mul_set_tuning(ctx, wide_slots, chain)
                o = cp.empty((batch, 3, k, P.Ng, 2))
                cp.hom_mul_batch(ctx, o, ct1, ct2)
                outs.append(o)
    finally:
        cp.hom_mul_set_tuning(ctx, 0, 0)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    i = batch - 1
    assert np.array_equal(cp.to_host(outs[0][i]),
                          O.hom_mul(P, cp.to_host(ct1[i]), cp.to_host(ct2[i])))


@pytest.mark.parametrize("k,n,batch", [(2, 2048, 52), (2, 2048, 53), (3, 4096, 37),
                                         (3, 8192, 37), (4, 16384, 29)])
def test_hom_mul_fused_and_composed_forms_agree(gpu, k, n, batch):
    """Batches with a (pair, residue) unit per CU (256 CUs: ceil(256 / (k + kl))
    pairs) run every transform of a unit in one workgroup; single pairs take
    the composed transforms.  Both forms give the same words for every pair,
    and two pairs equal the oracle composition.  At n = 2^11 two units share
    a workgroup: 53 pairs (265 units) leave the last workgroup's second row
    idle, which still copies and transforms a clamped unit but stores
    nothing."""
    import torch
    e, ctx = engine(k, k + 1, n, 65537)
    P = O.Params(k, k + 1, n, 65537)
    ct1 = rand_res(P, P.q, (batch, 2), 800 + n)
    ct2 = rand_res(P, P.q, (batch, 2), 801 + n)
    d1, d2 = cp.to_device(ct1), cp.to_device(ct2)
    full = cp.empty((batch, 3, k, P.Ng, 2))
    cp.hom_mul_batch(ctx, full, d1, d2)
    one = cp.empty((1, 3, k, P.Ng, 2))
    for i in range(batch):
        cp.hom_mul_batch(ctx, one, d1[i:i + 1], d2[i:i + 1])
        assert torch.equal(one[0], full[i]), i
    h = cp.to_host(full)
    for i in (0, batch - 1):
        assert np.array_equal(h[i], O.hom_mul(P, ct1[i], ct2[i])), i


def test_hom_mul_full_batch_properties(gpu):
    """n = 2^14, k = 4, 128 pairs (the benchmarked batch): the tensor product is
    symmetric, so swapping the ciphertexts leaves every pair's result unchanged;
    three pairs run alone (a batch of one takes the per-residue HPS kernels,
    the full batch the per-slot ones) give the same residues as in the batch;
    one pair equals the oracle composition."""
    import torch
    k, n, batch = 4, 16384, 128
    e, ctx = engine(k, k + 1, n, 65537)
    P = O.Params(k, k + 1, n, 65537)
    ct1 = rand_res(P, P.q, (batch, 2), 501)
    ct2 = rand_res(P, P.q, (batch, 2), 502)
    d1, d2 = cp.to_device(ct1), cp.to_device(ct2)
    out = cp.empty((batch, 3, k, P.Ng, 2))
    swapped = cp.empty(out.shape)
    cp.hom_mul_batch(ctx, out, d1, d2)
    cp.hom_mul_batch(ctx, swapped, d2, d1)
    assert torch.equal(out, swapped)
    one = cp.empty((1, 3, k, P.Ng, 2))
    for i in (0, 77, 127):
        cp.hom_mul_batch(ctx, one, d1[i:i + 1], d2[i:i + 1])
        assert torch.equal(one[0], out[i]), i
    assert np.array_equal(cp.to_host(out[5]), O.hom_mul(P, ct1[5], ct2[5]))


def test_sample_muladd_full_batch_properties(gpu):
    """n = 2^14, k = 4, batch 1024 (the benchmarked batch): item i of the batch
    uses the counters stream0 + 2i (s) and stream0 + 2i + 1 (e), so three items
    recomputed alone with those counters give the same residues; one item
    equals the oracle.  With a = 0 every item is the spectrum of its Gaussian
    e alone: back in the coefficient domain all 1024 * n coefficients are one
    small integer in every residue, within the sampler's truncation bound."""
    import torch
    k, n, batch, seed, s0 = 4, 16384, 1024, 777, 40
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 611)
    da = cp.to_device(a)
    out = cp.empty(a.shape)
    cp.sample_muladd_batch(ctx, out, da, seed, s0)
    one = cp.empty((1,) + a.shape[1:])
    for i in (0, 513, 1023):
        cp.sample_muladd_batch(ctx, one, da[i:i + 1], seed, s0 + 2 * i)
        assert torch.equal(one[0], out[i]), i
    assert np.array_equal(cp.to_host(out[7]), O.sample_muladd(P, a[7:8], seed, s0 + 14)[0])
    del da, out
    da = cp.empty(a.shape)
    noise = cp.empty(a.shape)
    cp.sample_muladd_batch(ctx, noise, da, seed, s0)
    for i in range(batch):
        cp.execute_dgt(ctx, noise[i], cp.QBase, cp.INVERSE)
    h = cp.to_host(noise)
    q = np.array(P.q, dtype=np.uint64).reshape(1, k, 1, 1)
    centered = np.where(h > q // 2, -((q - h).astype(np.int64)), h.astype(np.int64))
    _, bound, _ = O.cdt()
    assert np.abs(centered).max() <= bound
    assert (centered == centered[:, :1]).all()  # the same integer in every residue
    assert np.count_nonzero(centered) > batch * n // 2  # not all zero


@pytest.mark.parametrize("k,n,batch", [(2, 2048, 5), (4, 16384, 3), (3, 32768, 2)])
def test_sample_muladd_matches_oracle(gpu, k, n, batch):
    """n = 2^14, k = 4 (and other rings): a * s + e with s ternary and e discrete
    Gaussian drawn on the device, bit-exact against the oracle on the same
    Philox counters."""
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 300 + k)
    out = cp.empty(a.shape)
    cp.sample_muladd_batch(ctx, out, cp.to_device(a), 4242, 10)
    assert np.array_equal(cp.to_host(out), O.sample_muladd(P, a, 4242, 10))


@pytest.mark.parametrize("n,k,batch,choices", [
    (2048, 2, 40, [(1, -1), (2, 0), (4, 0), (8, 0), (1, 1)]),
    (4096, 3, 20, [(1, -1), (2, 0), (4, 0), (1, 1)]),
    (8192, 2, 12, [(1, -1), (2, 0), (1, 1)]),
    (16384, 2, 10, [(0, 1), (0, -1)]),
    (32768, 3, 90, [(0, 1), (0, -1)]),
])
def test_polymul_tuning_choices_and_autotune(gpu, n, k, batch, choices):
    """Every launch choice of the batched product (rows per workgroup; the
    persistent streaming form or one-shot workgroups; for n = 2^15 persistent
    or split) gives the same products, equal to the oracle; the autotuner
    picks one of them and leaves the product in out."""
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 500 + n)
    b = rand_res(P, P.q, (batch,), 501 + n)
    da, db = cp.to_device(a), cp.to_device(b)
    want = None
    for rows, persistent in choices:
        cp.polymul_set_tuning(ctx, rows, persistent)
        out = cp.empty(a.shape)
        cp.polymul_batch(ctx, out, da, db)
        h = cp.to_host(out)
        if want is None:
            for i in (0, batch - 1):
                assert np.array_equal(h[i], O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0])
            want = h
        assert np.array_equal(h, want), (rows, persistent)
    out = cp.empty(a.shape)
    pick = cp.polymul_autotune(ctx, out, da, db, reps=2)
    assert (pick["rows"], pick["persistent"]) in choices
    assert np.array_equal(cp.to_host(out), want)
    cp.polymul_set_tuning(ctx, 0, 0)
    with pytest.raises(RuntimeError):
        cp.polymul_set_tuning(ctx, 16, 0)


def test_polymul_persistent_choice_needs_n_2048_or_more(gpu):
    e, ctx = engine(2, 3, 1024)
    with pytest.raises(RuntimeError):
        cp.polymul_set_tuning(ctx, 0, 1)


@pytest.mark.parametrize("n,k,batch", [(16384, 4, 100), (4096, 3, 400), (32768, 3, 230)])
def test_polymul_streaming_many_units_per_workgroup(gpu, n, k, batch):
    """More residue products than the streaming grid holds, so each
    workgroup runs several units, each with its operands copied into LDS
    during the previous transform's last pass (at n = 2^15 the next unit's a
    read during the last inverse pass, 690 units: two or three per
    workgroup): equal to the one-shot (n = 2^15: split) form for every product
    (not a multiple of the grid), and to the oracle for the first and last."""
    e, ctx = engine(k, k + 1, n)
    P = O.Params(k, k + 1, n, 65537)
    a = rand_res(P, P.q, (batch,), 900 + n)
    b = rand_res(P, P.q, (batch,), 901 + n)
    da, db = cp.to_device(a), cp.to_device(b)
    outs = {}
    for persistent in (-1, 1):
        cp.polymul_set_tuning(ctx, 1 if n < 16384 else 0, persistent)
        out = cp.empty(a.shape)
        cp.polymul_batch(ctx, out, da, db)
        outs[persistent] = cp.to_host(out)
    assert np.array_equal(outs[1], outs[-1])
    for i in (0, batch - 1):
        assert np.array_equal(outs[1][i], O.polymul_batch(P, a[i:i + 1], b[i:i + 1])[0])
    # in place: out is a
    cp.polymul_set_tuning(ctx, 0, 1)
    cp.polymul_batch(ctx, da, da, db)
    assert np.array_equal(cp.to_host(da), outs[1])
    cp.polymul_set_tuning(ctx, 0, 0)
