"""GPU parity of the mixed-shape multiplication plan: one graph launch runs
requests of several ring shapes, relinearized or not, and every output equals
the per-shape batched call on the same operands, over replays whose operands
are rewritten in place between launches."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_cache = {}


def shape_setup(k, n, t=65537, seed=29):
    key = (k, n, t, seed)
    if key not in _cache:
        e = cp.CUDAEngine(k, k + 1, n, t)
        ctx = cp.Context(e)
        P = O.Params(k, k + 1, n, t)
        s_vals = O.sample_small(n, 2, seed, 0)
        s = cp.to_device(O.fold(O.small_to_residues(s_vals, P.q), P.Ng))
        rk = cp.RelinKey(ctx, s, seed, 300)
        _cache[key] = (e, ctx, P, rk)
    return _cache[key]


def uniform_cts(P, count, rng):
    ct = np.empty((count, 2, P.k, P.Ng, 2), dtype=np.uint64)
    for j, qj in enumerate(P.q):
        ct[:, :, j] = rng.integers(0, qj, size=(count, 2, P.Ng, 2), dtype=np.uint64)
    return ct


@pytest.mark.parametrize("relin,counts", [
    ((True, False, True), (5, 2, 7)),
    ((False, False, False), (5, 2, 7)),
    ((True, True, True), (5, 2, 7)),
    # (3, 8192) x 40 and (2, 2048) x 60 have a (pair, residue) unit per CU:
    # those two run chained, the (8, 32768) pair beside them
    ((True, False, True), (40, 2, 60)),
    ((False, True, False), (40, 2, 60))])
def test_mixed_shape_plan_equals_per_shape_calls(gpu, relin, counts):
    import torch
    shapes = [(3, 8192), (8, 32768), (2, 2048)]
    sets = [shape_setup(k, n) for k, n in shapes]
    engines = [s[0] for s in sets]
    keys = [s[3] if r else None for s, r in zip(sets, relin)]
    rng = np.random.default_rng(sum(relin) + sum(counts))
    # requests interleaved across shapes
    order = [i for i, c in enumerate(counts) for _ in range(c)]
    rng.shuffle(order)
    reqs, dev = [], []
    for si in order:
        P = sets[si][2]
        c1, c2 = (cp.to_device(uniform_cts(P, 1, rng)[0]) for _ in range(2))
        out = cp.empty((2 if relin[si] else 3, P.k, P.Ng, 2))
        reqs.append((si, out, c1, c2))
    origin = cp.Context(engines[0])
    plan = cp.HomMulPlan(origin, engines, reqs, keys=keys)
    for replay in range(3):
        if replay:  # new operands in the same buffers
            for si, out, c1, c2 in reqs:
                P = sets[si][2]
                c1.copy_(cp.to_device(uniform_cts(P, 1, rng)[0]))
                c2.copy_(cp.to_device(uniform_cts(P, 1, rng)[0]))
        torch.cuda.synchronize()
        plan.run()
        origin.sync()
        for si in range(len(shapes)):
            mine = [r for r in reqs if r[0] == si]
            e, ctx, P, rk = sets[si]
            a = torch.stack([r[2] for r in mine])
            b = torch.stack([r[3] for r in mine])
            want = cp.empty((len(mine), 2 if relin[si] else 3, P.k, P.Ng, 2))
            if relin[si]:
                cp.hom_mul_relin_batch(ctx, rk, want, a, b)
            else:
                cp.hom_mul_batch(ctx, want, a, b)
            ctx.sync()
            got = torch.stack([r[1] for r in mine])
            assert torch.equal(got, want), (replay, shapes[si])
    # one request against the oracle composition
    si, out, c1, c2 = reqs[0]
    P = sets[si][2]
    y = O.hom_mul(P, cp.to_host(c1), cp.to_host(c2))
    if relin[si]:
        kd = cp.empty((P.k, 2, P.k, P.Ng, 2))
        sets[si][3].copy_to(sets[si][1], kd)
        y = O.relinearize(P, y, cp.to_host(kd))
    assert np.array_equal(cp.to_host(out), y)
    plan.destroy()
    origin.destroy()


def test_plan_rejects_bad_requests(gpu):
    e, ctx, P, rk = shape_setup(2, 2048)
    c = cp.empty((2, 2, P.Ng, 2))
    with pytest.raises(ValueError, match="shape"):
        cp.HomMulPlan(ctx, [e], [(0, cp.empty((2, 2, P.Ng, 2)), c, c)])  # 3 rows needed
    with pytest.raises(ValueError, match="names no shape"):
        cp.HomMulPlan(ctx, [e], [(1, cp.empty((3, 2, P.Ng, 2)), c, c)], keys=None)
    # the library's own checks: a key of another engine
    e2, ctx2, P2, rk2 = shape_setup(3, 8192)
    with pytest.raises(cp.CupolyError, match="key of another engine"):
        cp.HomMulPlan(ctx, [e], [(0, cp.empty((2, 2, P.Ng, 2)), c, c)], keys=[rk2])
