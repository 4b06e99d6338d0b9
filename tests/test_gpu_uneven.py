"""Uneven batches (mixed ring degree n = 2^11..2^15 and width k = 1..8) in one
persistent launch, against the oracle and against the one-launch-per-op path;
hipGraph replays checked one by one, with eager work in between."""
import numpy as np
import pytest

import cupoly_amd as cp
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DEGREES = (2048, 4096, 8192, 16384, 32768)


def rand_poly(primes, ng, seed):
    rng = np.random.default_rng(seed)
    out = np.empty((len(primes), ng, 2), dtype=np.uint64)
    for r, p in enumerate(primes):
        out[r] = rng.integers(0, p, size=(ng, 2), dtype=np.uint64)
    return out


@pytest.fixture(scope="module")
def setup(gpu):
    import torch
    engines = [cp.CUDAEngine(8, 1, n, 65537) for n in DEGREES]
    stream = torch.cuda.Stream()
    ctx = cp.Context(engines[0], stream=stream)
    return engines, stream, ctx


def make_items(shapes, seed):
    items, host = [], []
    for i, (n, k) in enumerate(shapes):
        P = O.Params(k, 1, n, 65537)
        a = rand_poly(P.q, n // 2, seed + 2 * i)
        b = rand_poly(P.q, n // 2, seed + 2 * i + 1)
        host.append((P, a, b))
        items.append((n, k, cp.empty(a.shape), cp.to_device(a), cp.to_device(b)))
    return items, host


def check(items, want, what):
    for i, (it, w) in enumerate(zip(items, want)):
        assert np.array_equal(cp.to_host(it[2]), w), (what, i, it[0], it[1])


def test_uneven_matches_oracle_tuning_graph_and_plain(setup):
    import torch
    engines, stream, ctx = setup
    shapes = [(2048, 1), (2048, 3), (4096, 2), (8192, 5), (16384, 1), (16384, 4), (32768, 2),
              (32768, 8), (4096, 8), (2048, 7), (8192, 1), (2048, 2), (16384, 6), (4096, 1)]
    items, host = make_items(shapes, 100)
    want = [O.polymul_batch(P, a[None], b[None])[0] for P, a, b in host]
    plan = cp.UnevenPlan(ctx, engines, items)
    print("uneven: default tuning", plan.tuning(), flush=True)
    plan.run()
    stream.synchronize()
    check(items, want, "persistent")

    def zero_outs():
        with torch.cuda.stream(stream):
            for it in items:
                it[2].zero_()

    for grid, order in [(1, 0), (3, 1), (7, 0), (300, 1)]:
        zero_outs()
        plan.set_tuning(grid, order)
        plan.run()
        stream.synchronize()
        check(items, want, "grid %d order %d" % (grid, order))

    plan.set_tuning(5, 0)
    plan.capture()
    for rep in range(3):
        zero_outs()
        plan.run()
        stream.synchronize()
        check(items, want, "graph replay %d" % rep)

    best = plan.autotune(reps=2)
    print("uneven: autotune", best, flush=True)
    assert best["ms"] > 0 and best["order"] in (0, 1) and best["grid"] >= 1
    zero_outs()
    plan.run()
    stream.synchronize()
    check(items, want, "after autotune")

    zero_outs()
    plan.run_plain()
    stream.synchronize()
    check(items, want, "plain")
    plan.destroy()


def test_uneven_large_random_batch_persistent_equals_plain(setup):
    """Many more tiles than workgroups: persistent output == per-op output."""
    engines, stream, ctx = setup
    rng = np.random.default_rng(5)
    shapes = [(int(rng.choice(DEGREES)), int(rng.integers(1, 9))) for _ in range(160)]
    items, _ = make_items(shapes, 1000)
    plan = cp.UnevenPlan(ctx, engines, items)
    plan.set_tuning(256, 0)
    plan.run()
    stream.synchronize()
    got = [cp.to_host(it[2]) for it in items]
    for it in items:
        it[2].zero_()
    plan.run_plain()
    stream.synchronize()
    for i, (it, g) in enumerate(zip(items, got)):
        assert np.array_equal(cp.to_host(it[2]), g), (i, it[0], it[1])
    plan.destroy()


def test_uneven_rejects_bad_input(setup):
    engines, stream, ctx = setup
    x = cp.empty((1, 512, 2))
    with pytest.raises(cp.CupolyError):
        cp.UnevenPlan(ctx, engines, [(1024, 1, x, x, x)])
    y = cp.empty((9, 1024, 2))
    with pytest.raises(cp.CupolyError):
        cp.UnevenPlan(ctx, engines, [(2048, 9, y, y, y)])
    z = cp.empty((1, 2048, 2))
    with pytest.raises(cp.CupolyError):  # no engine for n = 4096 in the list
        cp.UnevenPlan(ctx, engines[:1], [(4096, 1, z, z, z)])
    empty = cp.UnevenPlan(ctx, engines, [])
    empty.run()
    empty.capture()
    empty.run_plain()
    assert empty.tuning()["tiles"] == 0
    empty.destroy()
