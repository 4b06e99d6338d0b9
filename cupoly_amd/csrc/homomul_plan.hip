// Mixed-shape BFV multiplication batches under one graph launch.
//
// The reference holds one ring shape per process (CUDAEngine's statics,
// include/cuPoly/cuda/cudaengine.h:71-93) and multiplies one ciphertext pair
// per call (polynomial.cu:350-457, 951-999).  A plan here takes a batch of
// requests of several ring shapes (each with its engine and, optionally, its
// relinearization key), groups them by shape, and captures every group's
// pipeline -- Q->B extension, the transforms with the tensor product, t/Q
// scaling with B->Q, and the relinearization -- as one parallel branch of a
// single hipGraph.  A run is one hipGraphLaunch and no launch is issued per
// shape or per stage from the host.  Groups too small to occupy every CU
// start together and overlap on the device; groups that fill the CUs on their
// own are chained one after another (two of them side by side only contend
// for the same CUs and caches, which measured slower than in series), while
// the small groups run beside the chain.
// Everything the branches use (a context per group with its own stream,
// workspace and device pointer tables) is allocated when the plan is made;
// the captured kernels allocate nothing.  The buffers the plan was made
// with must stay valid, and hold the operands of each run.
#include <string>
#include <vector>

#include "internal.hpp"

struct cupoly_hom_mul_plan {
  cupoly_ctx *origin = nullptr;  // the caller's context: runs launch on its stream
  struct group {
    cupoly_ctx *ctx = nullptr;   // own stream and scratch
    const cupoly_relin_key *key = nullptr;
    int count = 0;
    bool fills = false;  // occupies every CU alone: runs in the chain
    void *ws = nullptr;
    cupoly_gi **tables = nullptr;  // device: out[count], ct1[count], ct2[count]
    hipEvent_t done = nullptr;
  };
  std::vector<group> groups;
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static void plan_release(cupoly_hom_mul_plan *p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto &g : p->groups) {
    if (g.ws) (void)hipFree(g.ws);
    if (g.tables) (void)hipFree(g.tables);
    if (g.done) (void)hipEventDestroy(g.done);
    if (g.ctx) (void)cupoly_ctx_destroy(g.ctx);
  }
  if (p->fork) (void)hipEventDestroy(p->fork);
  delete p;
}

extern "C" int cupoly_hom_mul_plan_create(cupoly_ctx *c, cupoly_engine *const *engines,
                                          const cupoly_relin_key *const *keys, int n_shapes,
                                          const int *shape, cupoly_gi *const *out,
                                          const cupoly_gi *const *ct1,
                                          const cupoly_gi *const *ct2, int count,
                                          cupoly_hom_mul_plan **plan) {
  if (!plan) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: null plan");
  *plan = nullptr;
  if (!c || !engines || n_shapes < 1 || !shape || !out || !ct1 || !ct2 || count < 1)
    return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: bad argument");
  std::vector<std::vector<int>> members(n_shapes);
  for (int i = 0; i < count; i++) {
    if (shape[i] < 0 || shape[i] >= n_shapes)
      return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: request " + std::to_string(i) +
                                         " names no shape");
    if (!out[i] || !ct1[i] || !ct2[i])
      return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: null entry " + std::to_string(i));
    members[shape[i]].push_back(i);
  }
  for (int s = 0; s < n_shapes; s++) {
    if (members[s].empty()) continue;
    if (!engines[s]) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: null engine");
    if (engines[s]->device != c->e->device)
      return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: engines on different devices");
    if (keys && keys[s] && cp_relin_key_engine(keys[s]) != engines[s])
      return cp_fail(CUPOLY_ERR_STATE, "hom_mul_plan_create: key of another engine");
    if ((int)members[s].size() > 65535 / 4)
      return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_create: more than 16383 requests of a shape");
  }
  cupoly_hom_mul_plan *p = new cupoly_hom_mul_plan;
  p->origin = c;
  int rc = CUPOLY_OK;
  hipError_t err = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
  if (err != hipSuccess) rc = cp_check_hip(err, "hom_mul_plan_create: event");
  // per shape: a context on its own stream, workspace, device pointer tables
  for (int s = 0; s < n_shapes && !rc; s++) {
    if (members[s].empty()) continue;
    cupoly_hom_mul_plan::group g;
    g.key = keys ? keys[s] : nullptr;
    g.count = (int)members[s].size();
    g.fills = cp_hom_mul_fills(engines[s], g.count);
    p->groups.push_back(g);
    auto &G = p->groups.back();
    if ((rc = cupoly_ctx_create(engines[s], nullptr, &G.ctx))) break;
    if ((rc = cp_hom_mul_prepare(G.ctx, G.count))) break;
    const size_t ws = cp_hom_mul_ws_bytes(engines[s], G.count, G.key != nullptr);
    std::vector<cupoly_gi *> h(3 * (size_t)G.count);
    for (int j = 0; j < G.count; j++) {
      const int i = members[s][j];
      h[j] = out[i];
      h[G.count + j] = (cupoly_gi *)ct1[i];
      h[2 * G.count + j] = (cupoly_gi *)ct2[i];
    }
    err = hipMalloc(&G.ws, ws);
    if (err == hipSuccess) err = hipMalloc((void **)&G.tables, h.size() * sizeof(cupoly_gi *));
    if (err == hipSuccess)
      err = hipMemcpy(G.tables, h.data(), h.size() * sizeof(cupoly_gi *), hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&G.done, hipEventDisableTiming);
    if (err != hipSuccess) rc = cp_check_hip(err, "hom_mul_plan_create: allocation");
  }
  // capture: fork every group's stream from the origin stream (a group that
  // fills the CUs from the previous such group's end instead), run its
  // pipeline there, join back
  if (!rc) {
    err = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) rc = cp_check_hip(err, "hom_mul_plan_create: begin capture");
  }
  if (!rc) {
    int crc = CUPOLY_OK;
    err = hipEventRecord(p->fork, c->stream);
    if (err != hipSuccess) crc = cp_check_hip(err, "hom_mul_plan_create: fork");
    hipEvent_t chain = p->fork;
    for (auto &G : p->groups) {
      if (crc) break;
      cupoly_gi *const *t = G.tables;
      if (hipStreamWaitEvent(G.ctx->stream, G.fills ? chain : p->fork, 0) != hipSuccess)
        crc = cp_fail(CUPOLY_ERR_HIP, "hom_mul_plan_create: fork");
      if (!crc) crc = cp_hom_mul_tables_run(G.ctx, G.key, t, t + G.count, t + 2 * G.count, G.count,
                                            G.ws);
      if (!crc && hipEventRecord(G.done, G.ctx->stream) != hipSuccess)
        crc = cp_fail(CUPOLY_ERR_HIP, "hom_mul_plan_create: join");
      if (G.fills) chain = G.done;
      if (!crc && hipStreamWaitEvent(c->stream, G.done, 0) != hipSuccess)
        crc = cp_fail(CUPOLY_ERR_HIP, "hom_mul_plan_create: join");
    }
    // always end the capture, also after a failed step
    const hipError_t eerr = hipStreamEndCapture(c->stream, &p->graph);
    rc = crc ? crc : (eerr != hipSuccess ? cp_check_hip(eerr, "hom_mul_plan_create: end capture")
                                         : CUPOLY_OK);
  }
  if (!rc) {
    err = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (err != hipSuccess) rc = cp_check_hip(err, "hom_mul_plan_create: instantiate");
  }
  if (rc) {
    plan_release(p);
    return rc;
  }
  *plan = p;
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_plan_run(cupoly_hom_mul_plan *plan) {
  if (!plan || !plan->exec) return cp_fail(CUPOLY_ERR_ARG, "hom_mul_plan_run: no plan");
  CP_HIP(hipGraphLaunch(plan->exec, plan->origin->stream));
  return CUPOLY_OK;
}

extern "C" int cupoly_hom_mul_plan_destroy(cupoly_hom_mul_plan *plan) {
  if (!plan) return CUPOLY_OK;
  // the graph may still run on the origin stream
  const hipError_t err = hipStreamSynchronize(plan->origin->stream);
  plan_release(plan);
  return err == hipSuccess ? CUPOLY_OK : cp_check_hip(err, "hom_mul_plan_destroy");
}
