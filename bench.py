"""Benchmark: negacyclic poly-muls/s at n = 2^14, k = 4 on one MI355X.

A step is one pass of the hot path over one batch: `--batch` independent
products c = a * b mod (X^16384 + 1, q_i), i < 4, coefficient-domain uint64
residues in and out, inputs resident in HBM.  The batched form is one launch
of the fused forward-DGT / slot-wise product / inverse-DGT kernel per step.
Also reported on the same GPU:
  * plain: the reference's one-launch-per-op order (forward a, forward b,
    mul, inverse per product), same inputs;
  * roofline: algorithmic bytes of the fused kernel (24*k*n per product:
    read a and b, write c; twiddles are resident) over its measured average
    launch time, against 8 TB/s;
  * cpu_baseline: an optimised CPU port of the same algorithm
    (oracle/cpu_port.c, OpenMP over products x residues) on a bounded
    sample of the workload, checked equal to the GPU result.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
# untimed calls before each secondary leg: launch times fall over the first
# ~40 ms of sustained load as the clock settles, then hold
LEG_WARMUP = 30
# 32-bit VALU lane operations per second: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_LANE_OPS_PEAK = 256 * 4 * 16 * 2.4e9


def modmuls_per_product(n, k):
    """63-bit modular products per negacyclic product in the reference's
    algorithm (two forward and one inverse DGT of N_g = n/2 slots, each
    N_g log2 N_g butterfly products plus 5 N_g twist/scale products, and 3 N_g
    for the Gaussian slot-wise product), per residue, times k."""
    ng = n // 2
    return k * (3 * (ng * (ng.bit_length() - 1) + 5 * ng) + 3 * ng)


def product_kernel(n, k, batch):
    """The kernel cupoly_polymul_batch runs for this shape with the built-in
    launch choices."""
    import torch
    logng = (n // 2).bit_length() - 1
    if logng == 13 and batch * k > torch.cuda.get_device_properties(0).multi_processor_count:
        return "k_polymul_stream<13>"
    if logng <= 13:
        rows = 1 if logng >= 10 else max(1, 256 // ((n // 2) // 8))
        return "k_polymul_fused<%d,%d>" % (logng, rows)
    return "k_polymul_16k" if logng == 14 else "k_dgt_cols/k_polymul_blocks"


def bench_edges(n=16384, k=4, reps=100):
    """Per-call device time of the coefficient-edge and key-switching ops at
    (n, k): poly_crt / poly_icrt, Q->t scaling (t = 65537), xi / rho / dot
    (k pairs), and the fixed-weight / narrow / binary samplers; with the HBM
    bytes each moves."""
    import numpy as np
    import torch
    import cupoly_amd as cp

    eng = cp.CUDAEngine(k, k + 1, n, 65537)
    stream = torch.cuda.Stream()
    ctx = cp.Context(eng, stream=stream)
    ng = n // 2
    res = cp.empty((k, ng, 2))
    with torch.cuda.stream(stream):
        cp.Sampler.sample(ctx, res, cp.UNIFORM, 5, 1, coeff_domain=True)
    ints = cp.empty((n, k))
    gi = cp.empty((k, ng, 2))
    outs = [cp.empty((k, ng, 2)) for _ in range(k)]
    A = [cp.empty((k, ng, 2)) for _ in range(k)]
    B = [cp.empty((k, ng, 2)) for _ in range(k)]
    for i in range(k):
        with torch.cuda.stream(stream):
            cp.Sampler.sample(ctx, A[i], cp.UNIFORM, 6, i)
            cp.Sampler.sample(ctx, B[i], cp.UNIFORM, 7, i)
    sc = cp.empty((k, ng, 2))
    stream.synchronize()

    def timed(fn):
        """(GPU time per call from a replayed graph of `reps` captured calls,
        time per call when Python submits them one by one), both in us."""
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        host_us = e0.elapsed_time(e1) * 1e3 / reps
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            for _ in range(reps):
                fn()
        g.replay()  # warm-up replay
        stream.synchronize()
        e0.record(stream)
        g.replay()
        e1.record(stream)
        stream.synchronize()
        gpu_us = e0.elapsed_time(e1) * 1e3 / reps
        del g
        return gpu_us, host_us

    P8 = 8 * k * n  # bytes of one polynomial's residues
    ops = [
        ("icrt", lambda: cp.icrt(ctx, ints, res), P8 + 8 * n * k),
        ("crt", lambda: cp.crt(ctx, gi, ints), 8 * n * k + P8),
        # in place: after the first call residue 0 holds scaled words; the
        # work per call is the same for any residue values
        ("simple_scaling_t65537", lambda: cp.execute_polynomial_simple_scaling(ctx, sc), P8 + 8 * n),
        ("xi", lambda: cp.execute_xi_bfv(ctx, outs, res), 2 * P8),
        ("rho", lambda: cp.execute_rho_bfv(ctx, outs, res), 2 * P8),
        ("dot_k_pairs", lambda: cp.dot(ctx, gi, A, B), (2 * k + 1) * P8),
        ("sample_zo", lambda: cp.Sampler.sample(ctx, gi, cp.ZO, 9, 1, coeff_domain=True), P8),
        ("sample_hamming_weight", lambda: cp.Sampler.sample(ctx, gi, cp.HAMMINGWEIGHT, 9, 2, coeff_domain=True), P8),
        ("sample_narrow", lambda: cp.Sampler.sample(ctx, gi, cp.NARROW, 9, 3, coeff_domain=True), P8),
        ("sample_binary", lambda: cp.Sampler.sample(ctx, gi, cp.BINARY, 9, 4, coeff_domain=True), P8),
    ]
    sc.copy_(res)
    stream.synchronize()
    out = {"n": n, "k": k,
           "unit": "us per call (one polynomial): us = GPU time per call, from a hipGraph "
                   "of %d captured calls replayed back to back; us_host_submitted = the same "
                   "calls submitted one by one from Python (includes its call overhead)" % reps}
    for name, fn, nbytes in ops:
        with torch.cuda.stream(stream):
            us, host_us = timed(fn)
        out[name] = {"us": us, "us_host_submitted": host_us, "bytes": nbytes,
                     "GBs": nbytes / (us * 1e-6) / 1e9}
    # icrt/crt round trip equals the input residues
    with torch.cuda.stream(stream):
        cp.icrt(ctx, ints, res)
        cp.crt(ctx, gi, ints)
    stream.synchronize()
    assert torch.equal(gi, res), "crt(icrt(x)) != x"
    return out


def bench_uneven(count, steps, seed=4096):
    """`count` products, n uniform in {2^11..2^15}, k uniform in 1..8, uniform
    residues from the Philox sampler; persistent one-launch form (graph replay)
    against the one-launch-per-op form on the same inputs, whose outputs must
    be equal."""
    import numpy as np
    import torch
    import cupoly_amd as cp

    degrees = [2048, 4096, 8192, 16384, 32768]
    engines = [cp.CUDAEngine(8, 1, n, 65537) for n in degrees]
    gens = [cp.Context(e) for e in engines]
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice(degrees)), int(rng.integers(1, 9))) for _ in range(count)]
    total = sum(k * n // 2 for n, k in shapes)
    pool = cp.empty((3, total, 2))
    tmp = [cp.empty((8, n // 2, 2)) for n in degrees]
    items, off = [], 0
    for i, (n, k) in enumerate(shapes):
        sz = k * n // 2
        views = [pool[j, off:off + sz].view(k, n // 2, 2) for j in range(3)]
        d = degrees.index(n)
        for j in (1, 2):  # a, b
            cp.Sampler.sample(gens[d], tmp[d], cp.UNIFORM, seed, 2 * i + j - 1, coeff_domain=True)
            views[j].copy_(tmp[d][:k])
        items.append((n, k, views[0], views[1], views[2]))
        off += sz
    torch.cuda.synchronize()
    stream = torch.cuda.Stream()
    ctx = cp.Context(engines[0], stream=stream)
    plan = cp.UnevenPlan(ctx, engines, items)
    tune = plan.autotune(reps=3)
    plan.capture()
    plan.run()  # warm-up replay
    stream.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for s in range(steps):
        plan.run()
    ev1.record(stream)
    stream.synchronize()
    avg_ms = ev0.elapsed_time(ev1) / steps
    ref_out = pool[0].clone()
    # plain: 4 launches per product (transforms a and b in place, so it runs
    # after the persistent timing; the output is checked equal)
    pool[0].zero_()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    plan.run_plain()
    e1.record(stream)
    stream.synchronize()
    plain_ms = e0.elapsed_time(e1)
    assert torch.equal(pool[0], ref_out), "uneven: plain and persistent results differ"
    nbytes = sum(24 * k * n for n, k in shapes)
    info = plan.tuning()
    plan.destroy()
    for g in gens:
        g.destroy()
    del pool, tmp, ref_out
    return {"value": count / (avg_ms * 1e-3), "unit": "poly-muls/s", "count": count,
            "ms_per_batch": avg_ms, "plain_ms_per_batch": plain_ms,
            "plain_value": count / (plain_ms * 1e-3), "speedup_vs_plain": plain_ms / avg_ms,
            "tiles": info["tiles"], "autotune": tune,
            "workload": "negacyclic poly-mul, n uniform in {2^11..2^15}, k uniform in 1..8, "
                        "one persistent launch captured in a hipGraph",
            "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / (avg_ms * 1e-3) / 1e9,
            "hbm_frac": nbytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def bench_mixed(count, reps=10, seed=77):
    """`count` relinearized BFV multiplications, 70% (n=2^13, k=3) and 30%
    (n=2^15, k=8), uniform ciphertexts from the Philox sampler: one
    mixed-shape plan (a single graph launch, one branch per shape) against
    the two per-shape batched calls issued one after the other on one
    stream.  Outputs are checked equal."""
    import torch
    import cupoly_amd as cp

    dims = [(8192, 3), (32768, 8)]
    c0 = int(round(0.7 * count))
    counts = [c0, count - c0]
    engines = [cp.CUDAEngine(k, k + 1, n, 65537) for n, k in dims]
    stream = torch.cuda.Stream()
    ctxs = [cp.Context(e, stream=stream) for e in engines]
    keys, cts, outs, refs = [], [], [], []
    for si, ((n, k), c) in enumerate(zip(dims, ctxs)):
        ng = n // 2
        sk = cp.empty((k, ng, 2))
        with torch.cuda.stream(stream):
            cp.Sampler.sample(c, sk, cp.TERNARY, 4242, si, coeff_domain=True)
        keys.append(cp.RelinKey(c, sk, 4242, 1))
        a = cp.empty((counts[si], 2, k, ng, 2))
        b = cp.empty((counts[si], 2, k, ng, 2))
        with torch.cuda.stream(stream):
            for i in range(counts[si]):
                for j in range(2):
                    cp.Sampler.sample(c, a[i, j], cp.UNIFORM, seed + si, 4 * i + j, coeff_domain=True)
                    cp.Sampler.sample(c, b[i, j], cp.UNIFORM, seed + si, 4 * i + 2 + j,
                                      coeff_domain=True)
        cts.append((a, b))
        outs.append(cp.empty((counts[si], 2, k, ng, 2)))
        refs.append(cp.empty((counts[si], 2, k, ng, 2)))
    stream.synchronize()
    reqs = [(si, outs[si][i], cts[si][0][i], cts[si][1][i])
            for si in range(len(dims)) for i in range(counts[si])]
    origin = cp.Context(engines[0], stream=stream)
    plan = cp.HomMulPlan(origin, engines, reqs, keys=keys)

    def per_shape():
        for si in range(len(dims)):
            cp.hom_mul_relin_batch(ctxs[si], keys[si], refs[si], cts[si][0], cts[si][1])

    plan.run()
    per_shape()  # warm-up: workspaces
    stream.synchronize()
    for si in range(len(dims)):
        assert torch.equal(outs[si], refs[si]), "mixed plan and per-shape calls differ"

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_time(e1) / reps

    plan_ms = timed(plan.run)
    seq_ms = timed(per_shape)
    small = None
    if count > 10:
        # latency case: 7 + 3 requests, neither batch occupies every CU
        sreqs = [r for r in reqs if r[0] == 0][:7] + [r for r in reqs if r[0] == 1][:3]
        splan = cp.HomMulPlan(origin, engines, sreqs, keys=keys)
        sa = [(torch.stack([r[2] for r in sreqs if r[0] == si]),
               torch.stack([r[3] for r in sreqs if r[0] == si]),
               cp.empty((7 if si == 0 else 3,) + tuple(outs[si].shape[1:])))
              for si in range(len(dims))]

        def small_per_shape():
            for si in range(len(dims)):
                cp.hom_mul_relin_batch(ctxs[si], keys[si], sa[si][2], sa[si][0], sa[si][1])

        splan.run()
        small_per_shape()
        stream.synchronize()
        for si in range(len(dims)):
            got = torch.stack([r[1] for r in sreqs if r[0] == si])
            assert torch.equal(got, sa[si][2]), "mixed plan (7 + 3) and per-shape calls differ"
        small = {"count": 10, "ms_per_batch": timed(splan.run),
                 "per_shape_calls_ms": timed(small_per_shape),
                 "mix": "7 x (n=8192, k=3) + 3 x (n=32768, k=8), relinearized"}
        small["speedup_vs_per_shape_calls"] = small["per_shape_calls_ms"] / small["ms_per_batch"]
        splan.destroy()
    nbytes = sum(48 * k * n * c for (n, k), c in zip(dims, counts))
    plan.destroy()
    for key in keys:
        key.destroy()
    for c in ctxs + [origin]:
        c.destroy()
    return {"value": count / (plan_ms * 1e-3), "unit": "hom-mults/s", "count": count,
            "ms_per_batch": plan_ms, "per_shape_calls_ms": seq_ms,
            "speedup_vs_per_shape_calls": seq_ms / plan_ms,
            "mix": "%d x (n=8192, k=3) + %d x (n=32768, k=8), t=65537, relinearized" % tuple(counts),
            "form": "one hipGraph launch, the shapes' pipelines as parallel branches",
            "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / (plan_ms * 1e-3) / 1e9,
            "small": small}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the clock settles over the first ~40 launches of sustained load (launch
    # times fall by ~20% from the third launch to the fortieth): 50 untimed
    # launches (~60 ms) before the timed ones
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--plain-batch", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--hm-batch", type=int, default=128,
                    help="ciphertext pairs per batch for the BFV multiplication core")
    ap.add_argument("--enc-batch", type=int, default=1024,
                    help="instances per batch of the fused a*s+e sampling step")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-lanes", type=int, default=3,
                    help="serving: contexts (streams) per ring shape")
    ap.add_argument("--serve-inflight", type=int, default=2,
                    help="serving: batches queued on the device per lane")
    ap.add_argument("--serve-rate", type=float, default=20000.0,
                    help="offered load of the serving run, requests/s (Poisson)")
    ap.add_argument("--uneven-count", type=int, default=4096,
                    help="products in the uneven batch (n in 2^11..2^15, k in 1..8); 0 skips")
    ap.add_argument("--mixed-count", type=int, default=100,
                    help="relinearized multiplications in the mixed-shape plan (70%% n=2^13 "
                         "k=3, 30%% n=2^15 k=8); 0 skips")
    ap.add_argument("--edges", type=int, default=1,
                    help="1: also time the coefficient-edge / key-switching / sampler ops")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="before timing, copy a buffer of this many MiB once (a plain "
                         "copy kernel of known size for HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: cupoly_amd runs on one GPU only (the path has no multi-GPU "
              "form); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import cupoly_amd as cp

    torch.cuda.set_device(0)
    n, k, B = args.n, args.k, args.batch
    eng = cp.CUDAEngine(k, k + 1, n, 65537)
    ctx = cp.Context(eng)
    ng = n // 2
    # seeded uniform inputs, generated on the device by the Philox sampler
    a = cp.empty((B, k, ng, 2))
    b = cp.empty((B, k, ng, 2))
    for i in range(B):
        cp.Sampler.sample(ctx, a[i], cp.UNIFORM, 1234, 2 * i, coeff_domain=True)
        cp.Sampler.sample(ctx, b[i], cp.UNIFORM, 1234, 2 * i + 1, coeff_domain=True)
    out = cp.empty((B, k, ng, 2))
    torch.cuda.synchronize()
    print("bench: inputs ready (batch %d, n %d, k %d)" % (B, n, k), file=sys.stderr, flush=True)

    if args.calib_copy_mib:
        src = torch.empty(args.calib_copy_mib << 17, dtype=torch.int64, device="cuda")
        src.fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        cp.polymul_batch(ctx, out, a, b)
    torch.cuda.synchronize()
    # two events around the timed launches, on the context's stream (torch's
    # current stream): an event between every two launches costs about 30 us
    # per launch here (each record waits for the kernel's writes to drain)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for s in range(args.steps):
        cp.polymul_batch(ctx, out, a, b)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = B * args.steps / wall
    avg_launch_ms = ev0.elapsed_time(ev1) / args.steps
    bytes_per_launch = 24 * k * n * B
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    print("bench: batched done %.3f ms/step" % ms_per_step, file=sys.stderr, flush=True)

    # plain: one launch per op, same inputs (copies, as the plain path
    # transforms its operands in place)
    PB = min(args.plain_batch, B)
    pa, pb = a[:PB].clone(), b[:PB].clone()
    pout = cp.empty((PB, k, ng, 2))
    cp.polymul_plain(ctx, pout, pa.clone(), pb.clone())  # warm-up
    torch.cuda.synchronize()
    reps = 3
    copies = [(pa.clone(), pb.clone()) for _ in range(reps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for xa, xb in copies:
        cp.polymul_plain(ctx, pout, xa, xb)
    torch.cuda.synchronize()
    plain_rate = PB * reps / (time.perf_counter() - t1)
    assert torch.equal(pout, out[:PB]), "plain and batched results differ"
    # latency of one product alone: plain (the reference's 7 launches: two
    # per transform, one for the slot-wise product) and fused (one launch)
    one_a, one_b, one_o = a[:1].clone(), b[:1].clone(), cp.empty((1, k, ng, 2))
    lat_reps = 20
    torch.cuda.synchronize()
    t6 = time.perf_counter()
    for _ in range(lat_reps):
        cp.polymul_plain(ctx, one_o, one_a, one_b)  # a, b stay transformed: same work
        torch.cuda.synchronize()
    plain_one_ms = (time.perf_counter() - t6) * 1e3 / lat_reps
    t7 = time.perf_counter()
    for _ in range(lat_reps):
        cp.polymul_batch(ctx, one_o, a[:1], b[:1])
        torch.cuda.synchronize()
    fused_one_ms = (time.perf_counter() - t7) * 1e3 / lat_reps
    print("bench: plain done %.0f poly-mul/s" % plain_rate, file=sys.stderr, flush=True)

    # BFV multiplication core (tensor product + HPS Q->B, t/Q, B->Q), t = 65537
    HB = args.hm_batch
    hm = None
    if HB > 0:
        ct1 = cp.empty((HB, 2, k, ng, 2))
        ct2 = cp.empty((HB, 2, k, ng, 2))
        for i in range(HB):
            for j in range(2):
                cp.Sampler.sample(ctx, ct1[i, j], cp.UNIFORM, 99, 4 * i + j, coeff_domain=True)
                cp.Sampler.sample(ctx, ct2[i, j], cp.UNIFORM, 99, 4 * i + 2 + j, coeff_domain=True)
        hout = cp.empty((HB, 3, k, ng, 2))
        # warm-up: allocates the workspace, and the GPU clock settles over the
        # first ~40 ms of sustained load (launch times fall by ~10-15%)
        for _ in range(LEG_WARMUP):
            cp.hom_mul_batch(ctx, hout, ct1, ct2)
        torch.cuda.synchronize()
        hreps = 20
        t3 = time.perf_counter()
        for _ in range(hreps):
            cp.hom_mul_batch(ctx, hout, ct1, ct2)
        torch.cuda.synchronize()
        hm_s = (time.perf_counter() - t3) / hreps
        hm_bytes = 56 * k * n * HB
        hm = {"value": HB / hm_s, "unit": "hom-mults/s", "batch": HB, "ms_per_batch": hm_s * 1e3,
              "workload": "ciphertext tensor product + HPS Q->B (kl=k+1), t/Q scale-and-round "
                          "(t=65537), B->Q; coefficient domain in/out",
              "algorithmic_bytes": hm_bytes, "achieved_GBs": hm_bytes / hm_s / 1e9,
              "hbm_frac": hm_bytes / hm_s / 1e9 / HBM_PEAK_GBS}
        # the full multiplication: + relinearization with a resident key
        sk = cp.empty((k, ng, 2))
        cp.Sampler.sample(ctx, sk, cp.TERNARY, 4242, 0, coeff_domain=True)
        rkey = cp.RelinKey(ctx, sk, 4242, 1)
        rout = cp.empty((HB, 2, k, ng, 2))
        for _ in range(LEG_WARMUP):
            cp.hom_mul_relin_batch(ctx, rkey, rout, ct1, ct2)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        for _ in range(hreps):
            cp.hom_mul_relin_batch(ctx, rkey, rout, ct1, ct2)
        torch.cuda.synchronize()
        hr_s = (time.perf_counter() - t5) / hreps
        hr_bytes = 48 * k * n * HB  # 4 polys in, 2 out; the key is resident
        hm["relinearized"] = {"value": HB / hr_s, "unit": "hom-mults/s", "batch": HB,
                              "ms_per_batch": hr_s * 1e3,
                              "workload": "tensor product + HPS rescale + relinearization "
                                          "(k RNS digits, resident key)",
                              "algorithmic_bytes": hr_bytes,
                              "achieved_GBs": hr_bytes / hr_s / 1e9}
        rkey.destroy()
        del ct1, ct2, hout, rout, sk
        print("bench: hom-mul done %.0f hom-mults/s (%.0f relinearized)"
              % (HB / hm_s, HB / hr_s), file=sys.stderr, flush=True)

    # sampling + encryption-shaped step: a*s + e, s ternary, e Gaussian (sigma 3.2)
    enc = None
    if args.enc_batch > 0:
        EB = args.enc_batch
        ea = cp.empty((EB, k, ng, 2))
        for i in range(EB):
            cp.Sampler.sample(ctx, ea[i], cp.UNIFORM, 7, i)  # a uniform, DGT domain
        eout = cp.empty((EB, k, ng, 2))
        for _ in range(LEG_WARMUP):  # warm-up, as for the hom-mul leg
            cp.sample_muladd_batch(ctx, eout, ea, 31, 0)
        torch.cuda.synchronize()
        ereps = 20
        t4 = time.perf_counter()
        for _ in range(ereps):
            cp.sample_muladd_batch(ctx, eout, ea, 31, 0)
        torch.cuda.synchronize()
        es = (time.perf_counter() - t4) / ereps
        ebytes = 16 * k * n * EB
        enc = {"value": EB / es, "unit": "a*s+e/s", "batch": EB, "ms_per_batch": es * 1e3,
               "workload": "a*s+e in the DGT domain, s ternary, e discrete Gaussian sigma=3.2 "
                           "|e|<=19, Philox counters, n=%d k=%d" % (n, k),
               "algorithmic_bytes": ebytes, "achieved_GBs": ebytes / es / 1e9,
               "hbm_frac": ebytes / es / 1e9 / HBM_PEAK_GBS}
        del ea, eout
        print("bench: a*s+e done %.0f /s" % (EB / es), file=sys.stderr, flush=True)

    # uneven batch: products of random ring degree and width in one persistent
    # launch (hipGraph replay, autotuned grid and tile order) vs one launch per op
    uneven = None
    if args.uneven_count > 0:
        uneven = bench_uneven(args.uneven_count, args.steps)
        print("bench: uneven done %.0f products/s (%.1fx plain)"
              % (uneven["value"], uneven["speedup_vs_plain"]), file=sys.stderr, flush=True)

    mixed = None
    if args.mixed_count > 0:
        mixed = bench_mixed(args.mixed_count)
        print("bench: mixed plan done %.0f hom-mults/s (%.2fx per-shape calls)"
              % (mixed["value"], mixed["speedup_vs_per_shape_calls"]), file=sys.stderr, flush=True)

    edges = bench_edges(n, k) if args.edges else None
    if edges:
        print("bench: edges done (icrt %.1f us, crt %.1f us)" % (edges["icrt"]["us"], edges["crt"]["us"]),
              file=sys.stderr, flush=True)

    # serving: Poisson arrivals, 70% (n=2^13, k=3) / 30% (n=2^15, k=8) BFV
    # multiplications, continuous batching, HBM-footprint admission
    serve = None
    if args.serve_requests > 0:
        from cupoly_amd import serving
        shapes = [serving.Shape(8192, 3, lanes=args.serve_lanes),
                  serving.Shape(32768, 8, lanes=args.serve_lanes)]
        reqs = serving.make_requests(shapes, [0.7, 0.3], args.serve_requests, args.serve_rate,
                                     seed=2024)
        srv = serving.Server(shapes, hbm_budget=32 << 30,
                             max_batch={shapes[0]: 64, shapes[1]: 16},
                             max_inflight=args.serve_inflight)
        serve = srv.run(reqs)
        serve.update({"offered_req_per_s": args.serve_rate, "lanes_per_shape": args.serve_lanes,
                      "inflight_per_lane": args.serve_inflight,
                      "mix": "70% n=8192 k=3 / 30% n=32768 k=8, t=65537, tensor + HPS rescale + "
                             "relinearization with the shape's resident key",
                      "hbm_budget_bytes": 32 << 30})
        del reqs, srv, shapes
        torch.cuda.synchronize()
        print("bench: serving done p99 %.3f ms" % serve["p99_ms"], file=sys.stderr, flush=True)

    # HBM bytes per launch from the committed rocprofv3 PMC summary of this
    # same configuration (FETCH_SIZE / WRITE_SIZE scaled by a copy of known size)
    # and its VALU instruction count (SQ_INSTS_VALU, wave instructions)
    traffic, valu_insts = None, None
    try:
        with open(os.path.join(REPO, "profiles", "polymul_stream_n16384_k4.json")) as f:
            pm = json.load(f)
        if (pm["n"], pm["k"], pm["batch"]) == (n, k, B):
            traffic = pm["hbm_bytes_per_launch"]
            valu_insts = pm.get("sq_insts_valu_per_launch")
    except (OSError, ValueError, KeyError):
        pass
    # the bound of this kernel is integer instruction issue, not HBM: report
    # the modular-product rate and, from the profile's instruction count, the
    # fraction of the VALU issue peak the measured launch time implies
    mm = modmuls_per_product(n, k)
    integer = {"modmuls_per_product": mm, "modmuls_per_s": value * mm,
               "valu_lane_ops_peak": VALU_LANE_OPS_PEAK,
               "valu_insts_per_launch": valu_insts,
               "valu_issue_frac": (valu_insts * 64 / (avg_launch_ms * 1e-3) / VALU_LANE_OPS_PEAK
                                   if valu_insts else None)}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        P = O.Params(k, k + 1, n, 65537)
        S = min(args.cpu_sample, B)
        ha, hb = cp.to_host(a[:S]), cp.to_host(b[:S])
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        hc = O.port_polymul_batch(P, ha, hb)
        assert np.array_equal(hc, cp.to_host(out[:S])), "GPU result differs from the CPU port"
        reps, t2 = 0, time.perf_counter()
        while True:  # repeat the sample until about cpu_seconds of wall time
            O.port_polymul_batch(P, ha, hb)
            reps += 1
            cpu_s = time.perf_counter() - t2
            if cpu_s >= args.cpu_seconds:
                break
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": S * reps / cpu_s, "unit": "poly-muls/s", "cores": threads,
               "kind": "port", "cpu_model": model,
               "implementation": "oracle/cpu_port.c: the same DGT algorithm with precomputed "
                                 "Shoup twiddle and twist tables, no division in its loops, "
                                 "-O3 -march=x86-64-v3, OpenMP over products x residues",
               "sample": "%d products of the same workload (n=%d, k=%d) x %d repetitions, "
                         "%.1f s wall; checked equal to the GPU result" % (S, n, k, reps, cpu_s)}

    line = {
        "metric": "negacyclic poly-muls/s at N=2^14,k=4 batched vs plain; %HBM roofline; p99 ms",
        "value": value, "unit": "poly-muls/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "negacyclic poly-mul mod (X^n+1, q_i), coefficient domain in/out, "
                               "batched fused DGT->product->iDGT, one launch per batch",
                   "n": n, "k": k, "batch": B, "primes": "63-bit bucket q_0..q_%d" % (k - 1),
                   "parallelism": "single GPU"},
        "plain": {"value": plain_rate, "unit": "poly-muls/s", "batch": PB,
                  "launches_per_product": 7,
                  "form": "one product at a time in the reference's order: forward a, "
                          "forward b (column and block kernels each, every residue "
                          "polynomial spread over 2^(log2(n/2)-9) 64-thread workgroups), "
                          "slot-wise product, inverse (two kernels)",
                  "workgroups_per_transform_kernel": k * max(1, (ng >> 9)),
                  "single_product_ms": plain_one_ms,
                  "fused_single_product_ms": fused_one_ms},
        "speedup_vs_plain": value / plain_rate,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": product_kernel(n, k, B), "avg_launch_ms": avg_launch_ms,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     # what actually limits the kernel: 64-bit modular products on
                     # 32-bit integer lanes (the fraction of that peak is below)
                     "binding_resource": "integer VALU issue",
                     "valu_issue_frac": integer["valu_issue_frac"]},
        "integer_issue": integer,
        "cpu_baseline": cpu,
        "hom_mul": hm,
        "sample_muladd": enc,
        "serving": serve,
        "uneven": uneven,
        "mixed_hom_mul": mixed,
        "edges": edges,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
