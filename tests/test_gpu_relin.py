"""GPU parity of the BFV relinearization: the device-resident key, the
relinearization and the full multiplication (tensor product + HPS rescale +
relinearization), against the oracle composition, and end to end: a product
of two encryptions, relinearized on the GPU, decrypts to m1 * m2."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_cache = {}


def setup(k, n, t=65537, seed=17):
    """Engine, context, oracle parameters, secret and its key for (k, n)."""
    key = (k, n, t, seed)
    if key not in _cache:
        e = cp.CUDAEngine(k, k + 1, n, t)
        ctx = cp.Context(e)
        P = O.Params(k, k + 1, n, t)
        s_vals = O.sample_small(n, 2, seed, 0)
        s = cp.to_device(O.fold(O.small_to_residues(s_vals, P.q), P.Ng))
        rk = cp.RelinKey(ctx, s, seed, 100)
        _cache[key] = (e, ctx, P, s_vals, rk)
    return _cache[key]


def key_on_host(ctx, rk, P):
    d = cp.empty((P.k, 2, P.k, P.Ng, 2))
    rk.copy_to(ctx, d)
    return cp.to_host(d)


@pytest.mark.parametrize("k,n", [(2, 2048), (3, 8192), (4, 16384), (8, 32768), (2, 64)])
def test_relin_key_matches_oracle(gpu, k, n):
    """The key equals the oracle's (same Philox counters), and each pair
    satisfies b_i + a_i s - rho(s^2)_i = -e_i with e_i the Gaussian draw."""
    e, ctx, P, s_vals, rk = setup(k, n)
    key = key_on_host(ctx, rk, P)
    assert np.array_equal(key, O.relin_key(P, s_vals, 17, 100))
    sh = O.dgt(P, O.fold(O.small_to_residues(s_vals, P.q), P.Ng), 0, 0)
    rho = O.rho_bfv(P, O.pointwise(P, sh, sh, op=2))
    for i in range(k):
        lhs = O.pointwise(P, key[i, 0], O.pointwise(P, key[i, 1], sh, op=2), op=0)
        got_e = O.unfold(O.dgt(P, O.pointwise(P, rho[i], lhs, op=1), 0, 1))
        want_e = O.small_to_residues(O.sample_small(n, 3, 17, 100 + 2 * i + 1), P.q)
        assert np.array_equal(got_e, want_e), i


@pytest.mark.parametrize("k,n,batch", [(2, 2048, 3), (3, 4096, 2), (3, 8192, 3), (4, 16384, 2),
                                         (8, 32768, 1), (2, 64, 3)])
def test_relinearize_matches_oracle(gpu, k, n, batch):
    """Batches this small take the composed form (the fused one is checked
    against it below); bit-exact against the oracle on uniform (y0, y1, y2)."""
    import torch
    e, ctx, P, s_vals, rk = setup(k, n)
    key = key_on_host(ctx, rk, P)
    rng = np.random.default_rng(n + k)
    y = np.empty((batch, 3, k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        y[:, :, j] = rng.integers(0, qj, size=(batch, 3, P.Ng, 2), dtype=np.uint64)
    dy = cp.to_device(y)
    out = cp.empty((batch, 2, k, P.Ng, 2))
    cp.relinearize(ctx, rk, out, dy)
    h = cp.to_host(out)
    for b in range(batch):
        assert np.array_equal(h[b], O.relinearize(P, y[b], key)), b
    # the input is left as it was
    assert torch.equal(dy, cp.to_device(y))


def test_relinearize_n32768_one_workgroup_transforms(gpu):
    """n = 2^15, k = 8, 12 pairs: the digits' forward transforms (768 residue
    polynomials) and the accumulators' inverse transforms (192) cover 5/8 of
    the CUs, so they run one workgroup per residue polynomial
    (k_dgt_single16); the pairs relinearized one at a time take the split
    transforms.  Equal for every pair, and equal to the oracle for two."""
    k, n, batch = 8, 32768, 12
    e, ctx, P, s_vals, rk = setup(k, n)
    rng = np.random.default_rng(9)
    y = np.empty((batch, 3, k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        y[:, :, j] = rng.integers(0, qj, size=(batch, 3, P.Ng, 2), dtype=np.uint64)
    dy = cp.to_device(y)
    out = cp.empty((batch, 2, k, P.Ng, 2))
    cp.relinearize(ctx, rk, out, dy)
    h = cp.to_host(out)
    one = cp.empty((1, 2, k, P.Ng, 2))
    for b in range(batch):
        cp.relinearize(ctx, rk, one, dy[b:b + 1])
        assert np.array_equal(cp.to_host(one)[0], h[b]), b
    key = key_on_host(ctx, rk, P)
    for b in (0, batch - 1):
        assert np.array_equal(h[b], O.relinearize(P, y[b], key)), b


@pytest.mark.parametrize("k,n,batch", [(3, 8192, 4), (4, 16384, 2), (8, 32768, 2), (2, 2048, 24)])
def test_hom_mul_relin_equals_composition(gpu, k, n, batch):
    """hom_mul_relin_batch = relinearize(hom_mul_batch), and the gathered form
    gives the same words; one pair against the oracle composition."""
    import torch
    e, ctx, P, s_vals, rk = setup(k, n)
    rng = np.random.default_rng(3 * n + k)
    ct = np.empty((2, batch, 2, k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        ct[:, :, :, j] = rng.integers(0, qj, size=(2, batch, 2, P.Ng, 2), dtype=np.uint64)
    c1, c2 = cp.to_device(ct[0]), cp.to_device(ct[1])
    full = cp.empty((batch, 2, k, P.Ng, 2))
    cp.hom_mul_relin_batch(ctx, rk, full, c1, c2)
    y = cp.empty((batch, 3, k, P.Ng, 2))
    cp.hom_mul_batch(ctx, y, c1, c2)
    two = cp.empty((batch, 2, k, P.Ng, 2))
    cp.relinearize(ctx, rk, two, y)
    assert torch.equal(full, two)
    outs = [cp.empty((2, k, P.Ng, 2)) for _ in range(batch)]
    cp.hom_mul_relin_gather(ctx, rk, outs, [c1[i] for i in range(batch)],
                            [c2[i] for i in range(batch)])
    assert torch.equal(torch.stack(outs), full)
    key = key_on_host(ctx, rk, P)
    i = batch - 1
    want = O.relinearize(P, O.hom_mul(P, ct[0, i], ct[1, i]), key)
    assert np.array_equal(cp.to_host(full[i]), want)


@pytest.mark.parametrize("k,n", [(3, 4096), (4, 16384), (8, 32768)])
def test_bfv_multiply_relinearize_decrypts(gpu, k, n):
    """Encrypt m1, m2 (textbook BFV under a ternary secret), multiply and
    relinearize on the GPU, decrypt with the secret alone: m1 m2 mod
    (X^n + 1, t)."""
    e, ctx, P, s_vals, rk = setup(k, n)
    t = P.t
    rng = np.random.default_rng(7 * n + k)
    m1, m2 = rng.integers(0, t, n), rng.integers(0, t, n)
    ct1 = O.bfv_encrypt(P, m1, s_vals, 23, 10)
    ct2 = O.bfv_encrypt(P, m2, s_vals, 23, 20)
    out = cp.empty((1, 2, k, P.Ng, 2))
    cp.hom_mul_relin_batch(ctx, rk, out, cp.to_device(ct1[None]), cp.to_device(ct2[None]))
    got = O.bfv_decrypt(P, cp.to_host(out[0]), s_vals)
    assert np.array_equal(got, O.negacyclic_mod_t(m1, m2, t))


@pytest.mark.parametrize("k,n,batch", [(2, 2048, 128), (3, 2048, 87), (3, 8192, 86),
                                         (4, 16384, 64)])
def test_relinearize_fused_and_composed_forms_agree(gpu, k, n, batch):
    """Batches with a (pair, residue) unit per CU (ceil(256 / k) pairs) take the
    one-workgroup-per-unit kernel, single pairs the composed form: the same
    words for every pair, and two pairs equal the oracle.  At n = 2^11 two
    units share a workgroup; 87 pairs at k = 3 (261 units) leave the last
    workgroup's second row idle."""
    import torch
    e, ctx, P, s_vals, rk = setup(k, n)
    rng = np.random.default_rng(11 * n + k)
    y = np.empty((batch, 3, k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        y[:, :, j] = rng.integers(0, qj, size=(batch, 3, P.Ng, 2), dtype=np.uint64)
    dy = cp.to_device(y)
    full = cp.empty((batch, 2, k, P.Ng, 2))
    cp.relinearize(ctx, rk, full, dy)
    one = cp.empty((1, 2, k, P.Ng, 2))
    for i in range(batch):
        cp.relinearize(ctx, rk, one, dy[i:i + 1])
        assert torch.equal(one[0], full[i]), i
    key = key_on_host(ctx, rk, P)
    h = cp.to_host(full)
    for i in (0, batch - 1):
        assert np.array_equal(h[i], O.relinearize(P, y[i], key)), i


@pytest.mark.parametrize("k,n,batch", [(3, 8192, 75), (4, 16384, 58), (2, 2048, 104)])
def test_two_stream_halves_equal_single_pipelines(gpu, k, n, batch):
    """A batch whose halves each have a (pair, residue) unit per CU runs its
    halves on the context's two streams; every pair equals the same pairs run
    as two smaller batches (one pipeline each), in the contiguous, gathered
    and relinearizing forms, and one pair equals the oracle composition."""
    import torch
    e, ctx, P, s_vals, rk = setup(k, n)
    rng = np.random.default_rng(13 * n + batch)
    ct = np.empty((2, batch, 2, k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        ct[:, :, :, j] = rng.integers(0, qj, size=(2, batch, 2, P.Ng, 2), dtype=np.uint64)
    c1, c2 = cp.to_device(ct[0]), cp.to_device(ct[1])
    a = (batch + 1) // 2
    full = cp.empty((batch, 3, k, P.Ng, 2))
    cp.hom_mul_batch(ctx, full, c1, c2)
    parts = cp.empty((batch, 3, k, P.Ng, 2))
    cp.hom_mul_batch(ctx, parts[:a], c1[:a], c2[:a])
    cp.hom_mul_batch(ctx, parts[a:], c1[a:], c2[a:])
    assert torch.equal(full, parts)
    gathered = [cp.empty((3, k, P.Ng, 2)) for _ in range(batch)]
    cp.hom_mul_gather(ctx, gathered, [c1[i] for i in range(batch)], [c2[i] for i in range(batch)])
    assert torch.equal(torch.stack(gathered), full)
    rfull = cp.empty((batch, 2, k, P.Ng, 2))
    cp.hom_mul_relin_batch(ctx, rk, rfull, c1, c2)
    rparts = cp.empty((batch, 2, k, P.Ng, 2))
    cp.hom_mul_relin_batch(ctx, rk, rparts[:a], c1[:a], c2[:a])
    cp.hom_mul_relin_batch(ctx, rk, rparts[a:], c1[a:], c2[a:])
    assert torch.equal(rfull, rparts)
    i = batch - 1  # in the second half
    assert np.array_equal(cp.to_host(full[i]), O.hom_mul(P, ct[0, i], ct[1, i]))
