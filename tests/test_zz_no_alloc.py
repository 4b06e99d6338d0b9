"""The allocation-free hom-mul entry point (cupoly_hom_mul_batch_noalloc)
against cupoly_hom_mul_batch, through the C ABI, bit for bit.

Each shape makes two calls in a row on one context, the second with a smaller
batch and other inputs, so the second runs over a scratch area that still
holds the first call's data.  The shapes cover every launch form of the
pipeline: n = 2^14 (one workgroup per (pair, residue), per-slot and
per-residue HPS kernels), n = 2^13 (the serving shape) and n = 2^15 (the
one-workgroup-per-polynomial transforms at 4 pairs, the split transforms at 1).
A batch whose scratch exceeds the area is refused and leaves the output alone.
"""
import numpy as np
import pytest

import cupoly_amd as cp

pytestmark = pytest.mark.gpu


def _uniform(primes, shape, ng, seed):
    rng = np.random.default_rng(seed)
    out = np.empty(tuple(shape) + (len(primes), ng, 2), dtype=np.uint64)
    for r, p in enumerate(primes):
        out[..., r, :, :] = rng.integers(0, p, size=tuple(shape) + (ng, 2), dtype=np.uint64)
    return out


@pytest.mark.parametrize("k,n,batches", [(4, 16384, (12, 3)), (3, 8192, (20, 2)),
                                         (8, 32768, (4, 1))])
def test_noalloc_equals_hom_mul_batch(gpu, k, n, batches):
    import torch
    e = cp.CUDAEngine(k, k + 1, n, 65537)
    ctx = cp.Context(e)
    try:
        for step, batch in enumerate(batches):
            assert cp.hom_mul_workspace_bytes(e, batch) <= cp.NOALLOC_SCRATCH_BYTES
            ct1 = cp.to_device(_uniform(e.RNSPrimes, (batch, 2), n // 2, 900 + 10 * step + k))
            ct2 = cp.to_device(_uniform(e.RNSPrimes, (batch, 2), n // 2, 901 + 10 * step + k))
            got = cp.empty((batch, 3, k, n // 2, 2))
            want = cp.empty((batch, 3, k, n // 2, 2))
            cp.hom_mul_batch_noalloc(ctx, got, ct1, ct2)
            cp.hom_mul_batch(ctx, want, ct1, ct2)
            torch.cuda.synchronize()
            assert torch.equal(got, want), (k, n, batch)
    finally:
        ctx.destroy()
        e.destroy()


def test_noalloc_refuses_oversized_batch(gpu):
    import torch
    k, n = 4, 16384
    e = cp.CUDAEngine(k, k + 1, n, 65537)
    ctx = cp.Context(e)
    try:
        batch = 1
        while cp.hom_mul_workspace_bytes(e, batch) <= cp.NOALLOC_SCRATCH_BYTES:
            batch += 1
        ct = cp.empty((batch, 2, k, n // 2, 2))
        out = cp.empty((batch, 3, k, n // 2, 2))
        out.fill_(7)
        with pytest.raises(cp.CupolyError) as err:
            cp.hom_mul_batch_noalloc(ctx, out, ct, ct)
        assert err.value.code == -1 and "maximum" in str(err.value)
        torch.cuda.synchronize()
        assert bool((out == 7).all())
    finally:
        ctx.destroy()
        e.destroy()
