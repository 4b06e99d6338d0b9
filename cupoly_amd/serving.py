"""Continuous-batching request loop for BFV multiplications on one MI355X.

A request is a pair of ciphertexts of one ring shape to be multiplied: the
tensor product, the HPS rescale and the relinearization with the shape's
key.  Requests arrive over time (their ciphertexts in device memory, as a
receive buffer), and the loop
  * keeps one CUDAEngine and one relinearization key per ring shape resident
    for the whole run (twiddle, twist and HPS tables and the key stay in
    HBM), and `lanes` contexts per shape, each with its own stream and
    workspace, so batches of one shape and of different shapes overlap on the
    device (a batch of one small request fills only a few dozen CUs);
  * sizes every lane's workspace for its largest batch before the clock
    starts (cupoly_hom_mul_reserve) and charges those workspaces to the HBM
    budget up front, so no request waits on an allocation and the budget
    covers the memory the loop really holds;
  * admits arrived requests, in arrival order, while the output buffer each
    needs fits what is left of the budget; the output is allocated at
    admission and its bytes are returned to the budget when the result is
    ready (handed to the caller);
  * whenever a lane of a shape has fewer than `max_inflight` batches on the
    device, launches every admitted request of that shape (up to
    `max_batch`) as one gathered batch (cupoly_hom_mul_relin_gather) on that
    lane, so requests that arrive while others are on the GPU join the next
    batch instead of waiting for a drain;
  * timestamps completions with device events, and reports throughput and
    p50 / p99 latency from arrival to result-ready.

The reference has no serving layer (it is a single-engine library); this is
the caller the batched kernels are built for.
"""
import collections
import time

import numpy as np

import cupoly_amd as cp


class Shape:
    """One ring shape: its engine, lanes (contexts with their own streams) and
    relinearization key (secret drawn ternary from the Philox sampler)."""

    def __init__(self, n, k, kl=None, t=65537, lanes=2, key_seed=4242):
        import torch
        self.n, self.k, self.kl, self.t = n, k, kl or k + 1, t
        self.engine = cp.CUDAEngine(self.k, self.kl, n, t)
        self.streams = [torch.cuda.Stream() for _ in range(lanes)]  # default priority
        self.ctxs = [cp.Context(self.engine, stream=st) for st in self.streams]
        self.stream, self.ctx = self.streams[0], self.ctxs[0]
        ng = n // 2
        self.ct_shape = (2, self.k, ng, 2)
        self.out_shape = (2, self.k, ng, 2)
        # bytes a request holds while admitted: its relinearized output
        self.footprint = 2 * self.k * ng * 16
        s = cp.empty((self.k, ng, 2))
        with torch.cuda.stream(self.stream):
            cp.Sampler.sample(self.ctx, s, cp.TERNARY, key_seed, 0, coeff_domain=True)
        self.key = cp.RelinKey(self.ctx, s, key_seed, 1)
        self.secret = s

    def workspace_bytes(self, batch):
        """Device bytes one lane reserves for batches of up to `batch` requests."""
        return cp.hom_mul_relin_workspace_bytes(self.engine, batch)

    def __repr__(self):
        return "n=%d,k=%d" % (self.n, self.k)


class Request:
    """A multiplication request: its ciphertexts (device receive buffers) and,
    from admission on, its output buffer."""
    __slots__ = ("rid", "shape", "arrival", "ct1", "ct2", "out", "done_ms", "batch")

    def __init__(self, rid, shape, arrival, ct1, ct2):
        self.rid, self.shape, self.arrival = rid, shape, arrival
        self.ct1, self.ct2 = ct1, ct2
        self.out = None
        self.done_ms = None
        self.batch = None


def make_requests(shapes, probs, n_req, rate, seed, device="cuda"):
    """Poisson arrivals at `rate` req/s; shape i with probability probs[i];
    inputs uniform mod q_j from the Philox sampler (stream = request id)."""
    import torch
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    arrivals -= arrivals[0]
    kinds = rng.choice(len(shapes), size=n_req, p=probs)
    gens = {id(sh): cp.Context(sh.engine) for sh in shapes}
    reqs = []
    for i in range(n_req):
        sh = shapes[kinds[i]]
        gen = gens[id(sh)]
        ct1 = cp.empty(sh.ct_shape, device)
        ct2 = cp.empty(sh.ct_shape, device)
        for j in range(2):
            cp.Sampler.sample(gen, ct1[j], cp.UNIFORM, seed, 4 * i + j, coeff_domain=True)
            cp.Sampler.sample(gen, ct2[j], cp.UNIFORM, seed, 4 * i + 2 + j, coeff_domain=True)
        reqs.append(Request(i, sh, float(arrivals[i]), ct1, ct2))
    torch.cuda.synchronize()
    for g in gens.values():
        g.destroy()
    return reqs


class Server:
    def __init__(self, shapes, hbm_budget, max_batch=None, max_inflight=2):
        self.shapes = shapes
        self.hbm_budget = hbm_budget
        max_batch = dict(max_batch or {sh: 64 for sh in shapes})
        self.max_inflight = max_inflight
        # every lane's workspace is held for the whole run: charged first
        self.reserved = sum(len(sh.ctxs) * sh.workspace_bytes(max_batch[sh]) for sh in shapes)
        left = hbm_budget - self.reserved
        # a request larger than what is left could never be admitted, and
        # the loop would wait for it forever
        for sh in shapes:
            if sh.footprint > left:
                raise ValueError("HBM budget of %d bytes is below the footprint of one %r request "
                                 "(%d bytes) plus the reserved lane workspaces (%d bytes)"
                                 % (hbm_budget, sh, sh.footprint, self.reserved))
        # a batch never holds more requests than the budget admits at once
        self.max_batch = {sh: max(1, min(max_batch[sh], left // sh.footprint)) for sh in shapes}
        for sh in shapes:
            for ctx in sh.ctxs:
                cp.hom_mul_reserve(ctx, self.max_batch[sh])

    def run(self, requests):
        if not requests:
            return {"requests": 0, "batches": 0, "req_per_s": 0.0, "p50_ms": None,
                    "p99_ms": None, "mean_ms": None, "max_ms": None,
                    "peak_admitted_bytes": int(self.reserved),
                    "reserved_workspace_bytes": int(self.reserved), "per_shape": {}}
        import torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queues = {sh: collections.deque() for sh in self.shapes}
        # per shape and lane: batches on the device, oldest first (event, [requests])
        inflight = {sh: [collections.deque() for _ in sh.ctxs] for sh in self.shapes}
        used = self.reserved  # bytes held: lane workspaces + admitted outputs
        peak_used = used
        batches = 0
        done = []
        # warm torch's caching allocator with blocks of every output size, so
        # admission inside the timed loop does not reach hipMalloc
        warm = []
        for sh in self.shapes:
            cnt = min(sum(1 for r in requests if r.shape is sh),
                      (self.hbm_budget - self.reserved) // sh.footprint)
            warm += [torch.empty(sh.out_shape, dtype=torch.int64, device="cuda")
                     for _ in range(cnt)]
        del warm
        start = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record()
        while len(done) < len(requests):
            now = time.perf_counter() - t0
            # admission by HBM footprint, in arrival order
            while pending and pending[0].arrival <= now and \
                    used + pending[0].shape.footprint <= self.hbm_budget:
                r = pending.popleft()
                r.out = torch.empty(r.shape.out_shape, dtype=torch.int64, device="cuda")
                used += r.shape.footprint
                queues[r.shape].append(r)
            peak_used = max(peak_used, used)
            # launch: every shape with a lane that has room takes its whole
            # queue (up to max_batch) onto the least-loaded such lane
            for sh in self.shapes:
                q = queues[sh]
                while q:
                    lane = min(range(len(sh.ctxs)), key=lambda i: len(inflight[sh][i]))
                    if len(inflight[sh][lane]) >= self.max_inflight:
                        break
                    take = [q.popleft() for _ in range(min(len(q), self.max_batch[sh]))]
                    # the outputs were allocated on torch's current stream:
                    # the lane's stream waits for that before writing them
                    sh.streams[lane].wait_stream(torch.cuda.current_stream())
                    cp.hom_mul_relin_gather(sh.ctxs[lane], sh.key, [r.out for r in take],
                                            [r.ct1 for r in take], [r.ct2 for r in take])
                    for r in take:  # not reused by the allocator before the lane is done
                        r.out.record_stream(sh.streams[lane])
                    ev = torch.cuda.Event(enable_timing=True)
                    ev.record(sh.streams[lane])
                    for r in take:
                        r.batch = batches
                    batches += 1
                    inflight[sh][lane].append((ev, take))
            # completions (batches of one stream finish in order)
            progressed = False
            for sh in self.shapes:
                for fl in inflight[sh]:
                    while fl and fl[0][0].query():
                        ev, take = fl.popleft()
                        ms = start.elapsed_time(ev)
                        for r in take:
                            r.done_ms = ms
                            used -= sh.footprint
                            done.append(r)
                        progressed = True
            if not progressed and not any(queues[s] for s in self.shapes):
                if pending and pending[0].arrival > time.perf_counter() - t0:
                    time.sleep(min(50e-6, max(0.0, pending[0].arrival - (time.perf_counter() - t0))))
        torch.cuda.synchronize()
        lat = np.array([r.done_ms - 1e3 * r.arrival for r in done])
        span_ms = max(r.done_ms for r in done) - 1e3 * min(r.arrival for r in done)
        per_shape = {}
        for sh in self.shapes:
            ls = np.array([r.done_ms - 1e3 * r.arrival for r in done if r.shape is sh])
            if len(ls):
                per_shape[repr(sh)] = {"count": int(len(ls)), "p50_ms": float(np.percentile(ls, 50)),
                                       "p99_ms": float(np.percentile(ls, 99))}
        return {"requests": len(done), "batches": batches,
                "req_per_s": len(done) / (span_ms * 1e-3),
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
                "peak_admitted_bytes": int(peak_used), "reserved_workspace_bytes": int(self.reserved),
                "per_shape": per_shape}
