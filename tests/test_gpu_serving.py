"""Continuous batching: every request's result in the mixed, overlapped run is
bit-identical to the same multiplication (tensor, rescale, relinearization
with the shape's resident key) run alone (batch of one)."""
import pytest
import torch

import cupoly_amd as cp
from cupoly_amd import serving

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape_args,probs,rate", [
    ([(2048, 2), (4096, 3)], [0.7, 0.3], 4000.0),
    ([(8192, 3), (32768, 8)], [0.7, 0.3], 3000.0),
])
def test_serving_results_equal_alone(gpu, shape_args, probs, rate):
    shapes = [serving.Shape(n, k) for n, k in shape_args]
    reqs = serving.make_requests(shapes, probs, 40, rate, seed=5)
    srv = serving.Server(shapes, hbm_budget=8 << 30, max_batch={sh: 8 for sh in shapes})
    stats = srv.run(reqs)
    assert stats["requests"] == 40
    assert stats["batches"] >= 2
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    torch.cuda.synchronize()
    for r in reqs:
        alone = cp.empty(r.out.shape)
        ctx = cp.Context(r.shape.engine)  # torch's current stream: ordered after the inputs
        cp.hom_mul_relin_batch(ctx, r.shape.key, alone[None], r.ct1[None], r.ct2[None])
        torch.cuda.synchronize()
        ctx.destroy()
        assert torch.equal(alone, r.out), r.rid


def test_admission_respects_hbm_budget(gpu):
    shapes = [serving.Shape(2048, 2)]
    reqs = serving.make_requests(shapes, [1.0], 30, 1e6, seed=6)
    sh = shapes[0]
    reserved = len(sh.ctxs) * sh.workspace_bytes(3)
    budget = reserved + 3 * sh.footprint  # at most three outputs admitted at once
    srv = serving.Server(shapes, hbm_budget=budget, max_batch={sh: 3})
    assert srv.reserved == reserved
    stats = srv.run(reqs)
    assert stats["requests"] == 30
    assert stats["reserved_workspace_bytes"] == reserved
    assert reserved < stats["peak_admitted_bytes"] <= budget


def test_hom_mul_reserve_keeps_results(gpu):
    shapes = [serving.Shape(2048, 2, lanes=1)]
    reqs = serving.make_requests(shapes, [1.0], 3, 1e6, seed=7)
    sh = shapes[0]
    fresh = cp.Context(sh.engine)
    cp.hom_mul_reserve(sh.ctx, 8)
    a = cp.empty((3,) + sh.out_shape)
    b = cp.empty((3,) + sh.out_shape)
    ct1 = torch.stack([r.ct1 for r in reqs])
    ct2 = torch.stack([r.ct2 for r in reqs])
    # the stacks and zero fills ran on torch's current stream; the lane's
    # stream and the fresh context's stream wait for them
    torch.cuda.synchronize()
    cp.hom_mul_relin_batch(sh.ctx, sh.key, a, ct1, ct2)
    cp.hom_mul_relin_batch(fresh, sh.key, b, ct1, ct2)
    sh.ctx.sync()
    fresh.sync()
    assert torch.equal(a, b)
    with pytest.raises(RuntimeError):
        cp.hom_mul_reserve(sh.ctx, -1)
    fresh.destroy()
