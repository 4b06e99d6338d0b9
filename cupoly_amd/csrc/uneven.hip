// Uneven batches in one persistent launch: independent negacyclic products of
// different ring degrees (n = 2^11 .. 2^15) and RNS widths (k = 1 .. 8).
//
// The reference handles one (n, k) per process (CUDAEngine is a singleton,
// cudaengine.h:71-93) and issues 7 launches per product; a mixed batch there
// means re-initialising the engine.  Here every ring degree keeps its own
// resident tables (one engine per degree) and a plan flattens the batch into
// (product, residue) units, grouped into tiles of 8192 Gaussian slots:
//   * n <= 2^14: a tile holds 8192 / (n/2) units of one degree, processed side
//     by side by a 1024-thread workgroup (forward a, forward b, product,
//     inverse, as k_polymul_fused);
//   * n = 2^15 (16384 slots, 256 KiB, more than the LDS): one unit per tile;
//     each thread loads 8 slots of both halves, so the first (last) radix-2
//     stage runs in registers, and the remaining 13 stages run on each
//     8192-slot half in turn through LDS while the other half waits in a
//     private 384 KiB global scratch (polymul_16k_unit).
// A persistent grid of workgroups pulls tiles from an atomic counter until
// the list is exhausted; tile order and grid size are tunables picked by
// cupoly_uneven_autotune on the real batch.  The counter is reset by a small
// kernel (no memset node), so reset + persistent kernel can be captured into
// a hipGraph and replayed.
#include <algorithm>

#include "dgt_dev.hpp"

#define UNEVEN_THREADS 1024
#define UNEVEN_SLOTS 8192  // Gaussian slots per tile
#define UNEVEN_LDS_SLOTS 9216  // its padded LDS array (144 KiB)
#define UNEVEN_NCLASS 5    // log2(n/2) = 10 .. 14

struct cp_unit {
  const cp_gi *a;
  const cp_gi *b;
  cp_gi *c;
  int rid;
  int pad;
};

struct cp_tile {
  int cls;    // log2(n/2) - 10
  int nrows;  // units in this tile
  int unit0;  // first unit
  int pad;
};

// Per degree class: forward twiddles, and the split-basis tables (inverse
// roots, twists, the second operand's twists with 1/2N, constants;
// dgt_dev.hpp).
struct cp_class_tables {
  const uint64_t *twf[UNEVEN_NCLASS];
  const uint64_t *root[UNEVEN_NCLASS];
  const uint64_t *stw[UNEVEN_NCLASS];
  const uint64_t *stwh[UNEVEN_NCLASS];
  const uint64_t *sk[UNEVEN_NCLASS];
  const cp_prime *primes[UNEVEN_NCLASS];
};

struct cupoly_uneven {
  cupoly_ctx *ctx = nullptr;                         // stream owner
  std::vector<cupoly_ctx *> class_ctx;               // one per class (plain path)
  cp_class_tables tabs{};
  std::vector<cp_tile> tiles_in;                     // host copy, input order
  std::vector<cp_unit> units_h;
  std::vector<int> inst_n, inst_k;
  std::vector<const cupoly_gi *> inst_a, inst_b;
  std::vector<cupoly_gi *> inst_c;
  cp_tile *d_tiles = nullptr;
  cp_unit *d_units = nullptr;
  int *d_counter = nullptr;
  cp_gi *d_scratch = nullptr;  // 3 x 8192 slots per workgroup (n = 2^15 tiles only)
  size_t scratch_wgs = 0;
  bool has_16k = false;
  int ntiles = 0;
  int grid = 0;
  int order = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------
// Dynamic LDS of k_uneven (UNEVEN_LDS_SLOTS slots).
extern __shared__ cp_gi uneven_lds[];

// Generic -> global address space.  Pointers read from memory (the work
// list) are generic; typed this way the loads and stores through them are
// global_* instead of flat_*.
template <class T>
__device__ __forceinline__ CP_GLOBAL T *glb(T *ptr) {
  return (CP_GLOBAL T *)ptr;
}
__device__ __forceinline__ cp_unit ld_unit(const cp_unit *units, int i) {
  const CP_GLOBAL cp_unit *q = glb(units) + i;
  cp_unit u;
  u.a = q->a;
  u.b = q->b;
  u.c = q->c;
  u.rid = q->rid;
  u.pad = 0;
  return u;
}
__device__ __forceinline__ cp_prime ld_prime(const cp_prime *primes, int i) {
  const CP_GLOBAL cp_prime *q = glb(primes) + i;
  return cp_prime{q->p};
}

// Tile bodies, one per degree class, inlined into the persistent kernel
// (measured faster than out-of-line calls, whose callee-saved registers cost
// more scratch traffic than the larger shared spill frame).  Tables of the
// tile's ring degree come as separate pointer arguments.
template <int LOGN>
__device__ __forceinline__ void uneven_small(const cp_unit *units, int unit0, int nrows,
                                          const uint64_t *twf_c, const uint64_t *root_c,
                                          const uint64_t *stw_c, const uint64_t *stwh_c,
                                          const uint64_t *sk_c, const cp_prime *primes) {
  constexpr int N = 1 << LOGN, TPR = N / EPT, R = UNEVEN_SLOTS / N;
  static_assert(R * TPR == UNEVEN_THREADS, "tile shape");
  const int row = (int)threadIdx.x / TPR, t = (int)threadIdx.x % TPR;
  cp_gi *lds = uneven_lds + row * lds_row<LOGN>();
  const bool active = row < nrows;
  const cp_unit u = ld_unit(units, unit0 + (active ? row : 0));
  const cp_prime pr = ld_prime(primes, u.rid);
  const uint64_t p = pr.p;
  const CP_GLOBAL uint64_t *twf = glb(twf_c) + (size_t)u.rid * N;
  const CP_GLOBAL uint64_t *root = glb(root_c) + (size_t)u.rid * N * 4;
  const CP_GLOBAL uint64_t *stw = glb(stw_c) + (size_t)u.rid * N * 4;
  const CP_GLOBAL uint64_t *stwh = glb(stwh_c) + (size_t)u.rid * N * 4;
  const CP_GLOBAL uint64_t *sk = glb(sk_c) + (size_t)u.rid * 8;
  const uint64_t J = sk[0], Jp = sk[1];
  const CP_GLOBAL cp_gi *ua = glb(u.a), *ub = glb(u.b);
  CP_GLOBAL cp_gi *uc = glb(u.c);
  MapIdentity id;
  cp_gi xa[EPT], xb[EPT];
  const int ta = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i += 2) {
    const int g0 = first_index<LOGN, 0, LOGN>(ta, i), g1 = first_index<LOGN, 0, LOGN>(ta, i + 1);
    xa[i] = ld_gi_nt(ua + g0);
    xa[i + 1] = ld_gi_nt(ua + g1);
    split_twist2(xa[i], xa[i + 1], stw, g0, g1, J, Jp, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  // b is loaded during a's last forward pass (as in the streaming product)
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xa, lds, t, twf, p, id, [&] {
    const int tq = cp_opaque(t);
#pragma unroll
    for (int i = 0; i < EPT; i++) xb[i] = ld_gi_nt(ub + first_index<LOGN, 0, LOGN>(tq, i));
  });
  const int tb = cp_opaque(t);
#pragma unroll
  for (int i = 0; i < EPT; i += 2) {
    const int g0 = first_index<LOGN, 0, LOGN>(tb, i), g1 = first_index<LOGN, 0, LOGN>(tb, i + 1);
    split_twist2(xb[i], xb[i + 1], stwh, g0, g1, J, Jp, p);
    __builtin_amdgcn_sched_barrier(0);
  }
  dif_passes<LOGN, LOGN, 0, LOGN, 0>(xb, lds, t, twf, p, id);
#pragma unroll
  for (int i = 0; i < EPT; i++) xa[i] = split_mul_paired(xa[i], xb[i], pr);
  gsi_passes<LOGN, LOGN, 0, LOGN, 0, true>(xa, lds, t, root, p, id);
  const int to = cp_opaque(t);
  if (active) {
#pragma unroll
    for (int i = 0; i < EPT; i += 2) {
      split_back_h2(xa[i], xa[i + 1], J, Jp, p);
      st_gi_nt(uc + first_index<LOGN, 0, LOGN>(to, i), xa[i]);
      st_gi_nt(uc + first_index<LOGN, 0, LOGN>(to, i + 1), xa[i + 1]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// n = 2^15: 16384 slots (polymul_16k_unit); the workgroup's scratch holds
// the 3 * 8192 parked slots.
__device__ __forceinline__ void uneven_16k(const cp_unit *units, int unit0, const uint64_t *twf_c,
                                        const uint64_t *root_c, const uint64_t *stw_c,
                                        const uint64_t *stwh_c, const uint64_t *sk_c,
                                        const cp_prime *primes, cp_gi *scratch) {
  constexpr int N = 16384;
  const cp_unit u = ld_unit(units, unit0);
  const cp_prime pr = ld_prime(primes, u.rid);
  polymul_16k_unit(glb(u.c), glb(u.a), glb(u.b), glb(scratch), uneven_lds, pr,
                   glb(twf_c) + (size_t)u.rid * N, glb(root_c) + (size_t)u.rid * N * 4,
                   glb(stw_c) + (size_t)u.rid * N * 4, glb(stwh_c) + (size_t)u.rid * N * 4,
                   sk_c + (size_t)u.rid * 8);
}

__global__ void __launch_bounds__(UNEVEN_THREADS)
    k_uneven(const cp_tile *__restrict__ tiles, const cp_unit *__restrict__ units, int ntiles,
             int *__restrict__ counter, cp_class_tables T, cp_gi *__restrict__ scratch) {
  __shared__ int s_tile;
  cp_gi *my_scratch = scratch + (size_t)blockIdx.x * 3 * 8192;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int ti = s_tile;
    __syncthreads();  // everyone has read s_tile before thread 0 can rewrite it
    if (ti >= ntiles) break;  // uniform: every wave of the workgroup leaves
    const cp_tile tl = tiles[ti];
    switch (tl.cls) {
#define CP_UNEVEN_TABS(c) T.twf[c], T.root[c], T.stw[c], T.stwh[c], T.sk[c], T.primes[c]
      case 0: uneven_small<10>(units, tl.unit0, tl.nrows, CP_UNEVEN_TABS(0)); break;
      case 1: uneven_small<11>(units, tl.unit0, tl.nrows, CP_UNEVEN_TABS(1)); break;
      case 2: uneven_small<12>(units, tl.unit0, tl.nrows, CP_UNEVEN_TABS(2)); break;
      case 3: uneven_small<13>(units, tl.unit0, tl.nrows, CP_UNEVEN_TABS(3)); break;
      default: uneven_16k(units, tl.unit0, CP_UNEVEN_TABS(4), my_scratch); break;
#undef CP_UNEVEN_TABS
    }
    __syncthreads();  // LDS and scratch free for the next tile
  }
}

__global__ void k_reset_counter(int *counter) { *counter = 0; }

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
static int upload_order(cupoly_uneven *u) {
  std::vector<cp_tile> t = u->tiles_in;
  if (u->order == 0)  // most expensive tiles first (longest-processing-time order)
    std::stable_sort(t.begin(), t.end(), [](const cp_tile &x, const cp_tile &y) {
      const double cx = (x.cls == 4 ? 2.6 : 1.0) * (x.cls + 10), cy = (y.cls == 4 ? 2.6 : 1.0) * (y.cls + 10);
      return cx > cy;
    });
  CP_HIP(hipMemcpy(u->d_tiles, t.data(), t.size() * sizeof(cp_tile), hipMemcpyHostToDevice));
  return CUPOLY_OK;
}

static int ensure_scratch(cupoly_uneven *u, int grid) {
  if (!u->has_16k || (size_t)grid <= u->scratch_wgs) return CUPOLY_OK;
  if (u->d_scratch) {
    CP_HIP(hipStreamSynchronize(u->ctx->stream));
    CP_HIP(hipFree(u->d_scratch));
  }
  u->d_scratch = nullptr;
  CP_HIP(hipMalloc((void **)&u->d_scratch, (size_t)grid * 3 * 8192 * sizeof(cp_gi)));
  u->scratch_wgs = grid;
  return CUPOLY_OK;
}

static int launch_uneven(cupoly_uneven *u, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, s, u->d_counter);
  hipLaunchKernelGGL(k_uneven, dim3(u->grid), dim3(UNEVEN_THREADS),
                     UNEVEN_LDS_SLOTS * sizeof(cp_gi), s, u->d_tiles, u->d_units, u->ntiles,
                     u->d_counter, u->tabs, u->d_scratch);
  CP_LAUNCH_CHECK();
  return CUPOLY_OK;
}

extern "C" int cupoly_uneven_create(cupoly_ctx *c, cupoly_engine *const *engines, int n_engines,
                                    const int *n, const int *k, cupoly_gi *const *out,
                                    const cupoly_gi *const *a, const cupoly_gi *const *b,
                                    int count, cupoly_uneven **plan) {
  if (!c || !engines || !n || !k || !out || !a || !b || !plan || count < 0 || n_engines < 1)
    return cp_fail(CUPOLY_ERR_ARG, "uneven_create: bad argument");
  *plan = nullptr;
  cupoly_engine *by_cls[UNEVEN_NCLASS] = {};
  for (int i = 0; i < n_engines; i++) {
    cupoly_engine *e = engines[i];
    if (!e) return cp_fail(CUPOLY_ERR_ARG, "uneven_create: null engine");
    const int cls = e->logng - 10;
    if (cls < 0 || cls >= UNEVEN_NCLASS)
      return cp_fail(CUPOLY_ERR_UNSUPPORTED, "uneven: ring degree outside 2^11..2^15");
    by_cls[cls] = e;
  }
  cupoly_uneven *u = new cupoly_uneven();
  u->ctx = c;
  u->class_ctx.assign(UNEVEN_NCLASS, nullptr);
  for (int cl = 0; cl < UNEVEN_NCLASS; cl++) {
    if (!by_cls[cl]) continue;
    cupoly_engine *e = by_cls[cl];
    u->tabs.twf[cl] = e->d_tw_fwd;
    u->tabs.root[cl] = e->d_split_root;
    u->tabs.stw[cl] = e->d_split_tw;
    u->tabs.stwh[cl] = e->d_split_tw_h;
    u->tabs.sk[cl] = e->d_split_k;
    u->tabs.primes[cl] = e->d_primes;
    int rc = cupoly_ctx_create(e, (void *)c->stream, &u->class_ctx[cl]);
    if (rc) {
      cupoly_uneven_destroy(u);
      return rc;
    }
  }
  // units and tiles
  std::vector<std::vector<int>> per_cls(UNEVEN_NCLASS);
  for (int i = 0; i < count; i++) {
    int lg = 0;
    while ((1 << lg) < n[i] / 2) lg++;
    const int cl = lg - 10;
    if (n[i] < 2048 || (n[i] & (n[i] - 1)) || cl >= UNEVEN_NCLASS || !by_cls[cl]) {
      cupoly_uneven_destroy(u);
      return cp_fail(CUPOLY_ERR_UNSUPPORTED, "uneven: no engine for ring degree " + std::to_string(n[i]));
    }
    if (k[i] < 1 || k[i] > by_cls[cl]->k || !out[i] || !a[i] || !b[i]) {
      cupoly_uneven_destroy(u);
      return cp_fail(CUPOLY_ERR_ARG, "uneven: bad k or null pointer for product " + std::to_string(i));
    }
    u->inst_n.push_back(n[i]);
    u->inst_k.push_back(k[i]);
    u->inst_a.push_back(a[i]);
    u->inst_b.push_back(b[i]);
    u->inst_c.push_back(out[i]);
    const size_t ng = (size_t)n[i] / 2;
    for (int r = 0; r < k[i]; r++) {
      per_cls[cl].push_back((int)u->units_h.size());
      cp_unit un;
      un.a = (const cp_gi *)a[i] + r * ng;
      un.b = (const cp_gi *)b[i] + r * ng;
      un.c = (cp_gi *)out[i] + r * ng;
      un.rid = r;
      un.pad = 0;
      u->units_h.push_back(un);
    }
  }
  // tiles: consecutive units of one class; reorder units so each tile's units
  // are contiguous.  Within a class the units go residue-major, so a tile's
  // rows and consecutive tiles share one residue's tables in L2.
  std::vector<cp_unit> ordered;
  for (int cl = 0; cl < UNEVEN_NCLASS; cl++) {
    const int per_tile = cl == 4 ? 1 : UNEVEN_SLOTS >> (cl + 10);
    auto &ids = per_cls[cl];
    std::stable_sort(ids.begin(), ids.end(),
                     [&](int x, int y) { return u->units_h[x].rid < u->units_h[y].rid; });
    for (size_t s = 0; s < ids.size(); s += per_tile) {
      cp_tile tl;
      tl.cls = cl;
      tl.nrows = (int)std::min<size_t>(per_tile, ids.size() - s);
      tl.unit0 = (int)ordered.size();
      tl.pad = 0;
      for (int r = 0; r < tl.nrows; r++) ordered.push_back(u->units_h[ids[s + r]]);
      u->tiles_in.push_back(tl);
    }
  }
  u->units_h = ordered;
  u->has_16k = !per_cls[4].empty();
  u->ntiles = (int)u->tiles_in.size();
  int dev_cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->e->device) == hipSuccess) dev_cus = prop.multiProcessorCount;
  u->grid = std::max(1, std::min(dev_cus, std::max(1, u->ntiles)));
  hipError_t err = hipSuccess;
  if (u->ntiles) err = hipMalloc((void **)&u->d_tiles, u->tiles_in.size() * sizeof(cp_tile));
  if (err == hipSuccess && !u->units_h.empty())
    err = hipMalloc((void **)&u->d_units, u->units_h.size() * sizeof(cp_unit));
  if (err == hipSuccess) err = hipMalloc((void **)&u->d_counter, sizeof(int));
  if (err == hipSuccess && !u->units_h.empty())
    err = hipMemcpy(u->d_units, u->units_h.data(), u->units_h.size() * sizeof(cp_unit),
                    hipMemcpyHostToDevice);
  if (err != hipSuccess) {
    cupoly_uneven_destroy(u);
    return cp_check_hip(err, "uneven_create");
  }
  int rc;
  if (u->ntiles && ((rc = upload_order(u)) || (rc = ensure_scratch(u, u->grid)))) {
    cupoly_uneven_destroy(u);
    return rc;
  }
  // on the current device, for every plan (a plan may be built on any device)
  err = hipFuncSetAttribute((const void *)k_uneven, hipFuncAttributeMaxDynamicSharedMemorySize,
                            UNEVEN_LDS_SLOTS * sizeof(cp_gi));
  if (err != hipSuccess) {
    cupoly_uneven_destroy(u);
    return cp_check_hip(err, "uneven_create: LDS attribute");
  }
  *plan = u;
  return CUPOLY_OK;
}

extern "C" int cupoly_uneven_set_tuning(cupoly_uneven *u, int grid, int order) {
  if (!u || grid < 1 || (order != 0 && order != 1))
    return cp_fail(CUPOLY_ERR_ARG, "uneven_set_tuning: bad argument");
  int rc;
  if (u->exec) {
    (void)hipGraphExecDestroy(u->exec);
    (void)hipGraphDestroy(u->graph);
    u->exec = nullptr;
    u->graph = nullptr;
  }
  u->grid = grid;
  if (order != u->order) {
    u->order = order;
    if (u->ntiles && (rc = upload_order(u))) return rc;
  }
  return ensure_scratch(u, grid);
}

extern "C" int cupoly_uneven_get_tuning(const cupoly_uneven *u, int *grid, int *order, int *ntiles) {
  if (!u) return cp_fail(CUPOLY_ERR_ARG, "uneven_get_tuning: null");
  if (grid) *grid = u->grid;
  if (order) *order = u->order;
  if (ntiles) *ntiles = u->ntiles;
  return CUPOLY_OK;
}

extern "C" int cupoly_uneven_run(cupoly_uneven *u) {
  if (!u) return cp_fail(CUPOLY_ERR_ARG, "uneven_run: null");
  if (!u->ntiles) return CUPOLY_OK;
  if (u->exec) {
    CP_HIP(hipGraphLaunch(u->exec, u->ctx->stream));
    return CUPOLY_OK;
  }
  return launch_uneven(u, u->ctx->stream);
}

extern "C" int cupoly_uneven_capture(cupoly_uneven *u) {
  if (!u) return cp_fail(CUPOLY_ERR_ARG, "uneven_capture: null");
  if (!u->ntiles) return CUPOLY_OK;
  if (u->exec) return CUPOLY_OK;
  hipStream_t s = u->ctx->stream;
  if (!s) return cp_fail(CUPOLY_ERR_ARG, "uneven_capture: the null stream cannot be captured");
  CP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = launch_uneven(u, s);
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (rc) return rc;
  if (e2 != hipSuccess) return cp_check_hip(e2, "hipStreamEndCapture");
  u->graph = g;
  CP_HIP(hipGraphInstantiate(&u->exec, g, nullptr, nullptr, 0));
  return CUPOLY_OK;
}

// The same products one launch per op in the reference's order (forward a,
// forward b, slot-wise product, inverse), product by product.  a and b are
// transformed in place (left in the DGT domain).
extern "C" int cupoly_uneven_run_plain(cupoly_uneven *u) {
  if (!u) return cp_fail(CUPOLY_ERR_ARG, "uneven_run_plain: null");
  for (size_t i = 0; i < u->inst_n.size(); i++) {
    int lg = 0;
    while ((1 << lg) < u->inst_n[i] / 2) lg++;
    cupoly_ctx *c = u->class_ctx[lg - 10];
    const int k = u->inst_k[i];
    int rc;
    if ((rc = cp_dgt_launch(c, (cp_gi *)u->inst_a[i], 0, k, 1, CUPOLY_FORWARD)) ||
        (rc = cp_dgt_launch(c, (cp_gi *)u->inst_b[i], 0, k, 1, CUPOLY_FORWARD)) ||
        (rc = cp_pointwise_mul_launch(c, (cp_gi *)u->inst_c[i], (const cp_gi *)u->inst_a[i],
                                      (const cp_gi *)u->inst_b[i], k)) ||
        (rc = cp_dgt_launch(c, (cp_gi *)u->inst_c[i], 0, k, 1, CUPOLY_INVERSE)))
      return rc;
  }
  return CUPOLY_OK;
}

extern "C" int cupoly_uneven_autotune(cupoly_uneven *u, int reps, int *best_grid, int *best_order,
                                      float *best_ms) {
  if (!u || reps < 1) return cp_fail(CUPOLY_ERR_ARG, "uneven_autotune: bad argument");
  int cus = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, u->ctx->e->device) == hipSuccess) cus = prop.multiProcessorCount;
  const int grids[] = {cus / 2, cus, 2 * cus, 4 * cus};
  hipEvent_t e0, e1;
  CP_HIP(hipEventCreate(&e0));
  if (hipError_t err = hipEventCreate(&e1); err != hipSuccess) {
    (void)hipEventDestroy(e0);
    return cp_check_hip(err, "hipEventCreate");
  }
  auto hip = [](hipError_t err, const char *what) {
    return err == hipSuccess ? CUPOLY_OK : cp_check_hip(err, what);
  };
  float best = 1e30f;
  const int grid0 = u->grid, order0 = u->order;
  int bg = grid0, bo = order0, rc = CUPOLY_OK;
  for (int order = 0; order < 2 && !rc; order++)
    for (int gi = 0; gi < 4 && !rc; gi++) {
      const int g = std::max(1, std::min(grids[gi], std::max(1, u->ntiles)));
      if ((rc = cupoly_uneven_set_tuning(u, g, order))) break;
      if ((rc = launch_uneven(u, u->ctx->stream))) break;  // warm-up
      if ((rc = hip(hipEventRecord(e0, u->ctx->stream), "hipEventRecord"))) break;
      for (int r = 0; r < reps && !rc; r++) rc = launch_uneven(u, u->ctx->stream);
      if (rc || (rc = hip(hipEventRecord(e1, u->ctx->stream), "hipEventRecord")) ||
          (rc = hip(hipEventSynchronize(e1), "hipEventSynchronize")))
        break;
      float ms = 0;
      if ((rc = hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))) break;
      ms /= reps;
      if (ms < best) {
        best = ms;
        bg = g;
        bo = order;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {  // a failed search leaves the caller's choice in place
    const std::string msg = cupoly_last_error();
    (void)cupoly_uneven_set_tuning(u, grid0, order0);
    return cp_fail(rc, msg);
  }
  if ((rc = cupoly_uneven_set_tuning(u, bg, bo))) return rc;
  if (best_grid) *best_grid = bg;
  if (best_order) *best_order = bo;
  if (best_ms) *best_ms = best;
  return CUPOLY_OK;
}

extern "C" int cupoly_uneven_destroy(cupoly_uneven *u) {
  if (!u) return cp_fail(CUPOLY_ERR_ARG, "uneven_destroy: null");
  if (u->ctx && u->ctx->stream) (void)hipStreamSynchronize(u->ctx->stream);
  if (u->exec) (void)hipGraphExecDestroy(u->exec);
  if (u->graph) (void)hipGraphDestroy(u->graph);
  for (cupoly_ctx *c : u->class_ctx)
    if (c) cupoly_ctx_destroy(c);
  void *ptrs[] = {u->d_tiles, u->d_units, u->d_counter, u->d_scratch};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete u;
  return CUPOLY_OK;
}
