// Batched evaluation of many independent systems (MOIL contexts of uneven
// size, e.g. the slices of a chain/path minimisation or a set of separate
// boxes): one call evaluates all of them.
//
// The whole batch step is captured once into a hipGraph: the systems'
// eforce sequences are spread round-robin over `nstreams` internal streams
// (default priority) forked from and joined back into a capture stream, so
// independent small systems run concurrently and the step is replayed
// with one graph launch instead of ~25 launches per system.  The graph is
// recaptured when a system's lists were rebuilt (grid sizes and buffers
// change) or the caller passes different device pointers.  Nothing inside
// the captured region is a memset: the per-evaluation clears of the
// kernels are done by kernels, and a context's one-time accumulator clear
// is run eagerly before capture.  Work is ordered after the caller's
// stream and the caller's stream after the batch through events.
#include <hip/hip_runtime.h>

#include <vector>

#include "ctx.h"

struct moil_batch {
  std::vector<moil_ctx *> ctx;
  std::vector<hipStream_t> streams;
  hipStream_t cap = nullptr;                  // capture / launch stream
  std::vector<hipEvent_t> ev_fork, ev_join;   // per stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<long> gen;                      // list generation captured
  std::vector<const void *> key;              // pointers captured
};

static void batch_drop_graph(moil_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

static void batch_free(moil_batch *b) {
  if (!b) return;
  batch_drop_graph(b);
  for (hipEvent_t e : b->ev_fork) hipEventDestroy(e);
  for (hipEvent_t e : b->ev_join) hipEventDestroy(e);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_out) hipEventDestroy(b->ev_out);
  for (hipStream_t s : b->streams) hipStreamDestroy(s);
  if (b->cap) hipStreamDestroy(b->cap);
  delete b;
}

extern "C" int32_t moil_gpu_batch_create(moil_ctx **ctxs, int32_t n, int32_t nstreams,
                                         moil_batch **out) {
  if (!ctxs || n < 1 || !out) {
    set_error("batch_create: null argument or empty batch");
    return MOIL_GPU_EINVAL;
  }
  for (int i = 0; i < n; i++)
    if (!ctxs[i]) {
      set_error("batch_create: null context");
      return MOIL_GPU_EINVAL;
    }
  if (nstreams < 1) nstreams = 1;
  if (nstreams > n) nstreams = n;
  moil_batch *b = new moil_batch();
  b->ctx.assign(ctxs, ctxs + n);
  bool ok = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) == hipSuccess;
  for (int k = 0; ok && k < nstreams; k++) {
    hipStream_t s;
    hipEvent_t f, j;
    ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
    if (ok) b->streams.push_back(s);
    ok = ok && hipEventCreateWithFlags(&f, hipEventDisableTiming) == hipSuccess;
    if (ok) b->ev_fork.push_back(f);
    ok = ok && hipEventCreateWithFlags(&j, hipEventDisableTiming) == hipSuccess;
    if (ok) b->ev_join.push_back(j);
  }
  if (!ok) {
    batch_free(b);
    set_error("batch_create: stream/event creation failed");
    return MOIL_GPU_EHIP;
  }
  *out = b;
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_batch_destroy(moil_batch *b) {
  if (b) {
    hipStreamSynchronize(b->cap);
    batch_free(b);
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_eforce_dev(moil_ctx *c, const double *d_coor, double *d_dpot,
                                       double *d_energies, void *stream);

extern "C" int32_t moil_gpu_batch_eforce_dev(moil_batch *b, const double *const *d_coor,
                                             double *const *d_dpot, double *const *d_energies,
                                             void *stream) {
  if (!b || !d_coor || !d_dpot || !d_energies) {
    set_error("batch_eforce: null argument");
    return MOIL_GPU_EINVAL;
  }
  const int n = (int)b->ctx.size();
  for (int i = 0; i < n; i++) {
    if (!b->ctx[i]->have_lists) {
      set_error("batch_eforce: a system has no lists (build_lists first)");
      return MOIL_GPU_ESTATE;
    }
    if (!d_coor[i] || !d_dpot[i] || !d_energies[i]) {
      set_error("batch_eforce: null device pointer");
      return MOIL_GPU_EINVAL;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  // (re)capture when lists or pointers changed
  bool fresh = b->exec != nullptr && (int)b->gen.size() == n;
  for (int i = 0; fresh && i < n; i++)
    fresh = b->gen[i] == b->ctx[i]->list_gen && b->key[3 * i] == d_coor[i] &&
            b->key[3 * i + 1] == d_dpot[i] && b->key[3 * i + 2] == d_energies[i];
  HIP_CHECK(hipEventRecord(b->ev_in, st));
  HIP_CHECK(hipStreamWaitEvent(b->cap, b->ev_in, 0));
  if (!fresh) {
    batch_drop_graph(b);
    // one-time accumulator clears (memsets) stay out of the graph
    for (int i = 0; i < n; i++)
      if (b->ctx[i]->facc_dirty) {
        HIP_CHECK(hipMemsetAsync(b->ctx[i]->d_facc, 0, sizeof(long long) * 3 * b->ctx[i]->npt,
                                 b->cap));
        b->ctx[i]->facc_dirty = false;
      }
    for (int i = 0; i < n; i++) b->ctx[i]->timing = 0;
    HIP_CHECK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    const int ns = (int)b->streams.size();
    int32_t rc = MOIL_GPU_OK;
    for (int k = 0; k < ns; k++) {
      if (hipEventRecord(b->ev_fork[k], b->cap) != hipSuccess ||
          hipStreamWaitEvent(b->streams[k], b->ev_fork[k], 0) != hipSuccess)
        rc = MOIL_GPU_EHIP;
    }
    for (int i = 0; rc == MOIL_GPU_OK && i < n; i++)
      rc = moil_gpu_eforce_dev(b->ctx[i], d_coor[i], d_dpot[i], d_energies[i],
                               b->streams[i % ns]);
    for (int k = 0; k < ns; k++) {
      if (hipEventRecord(b->ev_join[k], b->streams[k]) != hipSuccess ||
          hipStreamWaitEvent(b->cap, b->ev_join[k], 0) != hipSuccess)
        rc = MOIL_GPU_EHIP;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap, &g);
    if (rc != MOIL_GPU_OK) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    HIP_CHECK(e);
    b->graph = g;
    HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->gen.resize(n);
    b->key.resize(3 * n);
    for (int i = 0; i < n; i++) {
      b->gen[i] = b->ctx[i]->list_gen;
      b->key[3 * i] = d_coor[i];
      b->key[3 * i + 1] = d_dpot[i];
      b->key[3 * i + 2] = d_energies[i];
    }
  }
  HIP_CHECK(hipGraphLaunch(b->exec, b->cap));
  HIP_CHECK(hipEventRecord(b->ev_out, b->cap));
  HIP_CHECK(hipStreamWaitEvent(st, b->ev_out, 0));
  return MOIL_GPU_OK;
}
