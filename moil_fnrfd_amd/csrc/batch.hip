// Batched evaluation of many independent systems (MOIL contexts of uneven
// size, e.g. the slices of a chain/path minimisation, chain/sds.f:71-140,
// or a set of separate boxes): one call evaluates all of them.
//
// Every kernel of the evaluation runs once for the whole batch: a
// multi-system launch covers the workgroups of all systems, each
// workgroup finding its system through a table of per-system argument
// blocks and workgroup offsets (kargs.h, multi.h).  The 3-D transforms
// run as one batched rocFFT plan per distinct grid size, on transform
// buffers owned by the batch (the systems' spreading writes into them and
// the gather reads from them).  The step -- about twenty kernel launches
// plus two transforms per grid size, independent of the number of systems
// -- runs as two captured hipGraphs replayed side by side: the reciprocal
// sums on one stream (the transform groups spread over up to three more
// streams forked and joined inside the capture), the real-space part on
// another (submitted first), joined by an event before the two assembly
// launches.  The graphs are recaptured when a system's lists were rebuilt
// (chunk counts and list buffers change), its parameters changed, or the
// caller passes different device pointers.  Nothing captured is a memset:
// the per-evaluation clears are done by the kernels themselves, and a
// context's one-time accumulator clear runs eagerly before capture.  The
// caller's stream is ordered before and after the batch through events.
// Systems are laid out in the grids in decreasing atom count, so the
// largest systems' workgroups start first.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "kargs.h"

#define RF_TRY(x)                                            \
  do {                                                       \
    if ((x) != rocfft_status_success) {                      \
      set_error(std::string("batch: rocFFT failure at ") +  \
                __FILE__ + ":" + std::to_string(__LINE__));  \
      return MOIL_GPU_EHIP;                                  \
    }                                                        \
  } while (0)

#define CHECK_RC_B(x)                  \
  do {                                 \
    const int32_t r_ = (x);            \
    if (r_ != MOIL_GPU_OK) return r_;  \
  } while (0)

namespace {

// one batched transform pair over the systems sharing a grid size
struct FftGroup {
  int n1 = 0, n2 = 0, n3 = 0;
  std::vector<int> members;  // system indices
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info_f = nullptr, info_b = nullptr;
  void *work = nullptr;
  void *qreal = nullptr, *qcplx = nullptr;
  size_t G = 0, H = 0;
};

// per-kernel table of argument blocks + workgroup offsets
template <class A>
struct Tab {
  std::vector<A> args;
  std::vector<int32_t> bs{0};
  size_t off_a = 0, off_b = 0;  // byte offsets in the device blob
  void add(const A &a, int n) {
    args.push_back(a);
    bs.push_back(bs.back() + std::max(n, 0));
  }
  int total() const { return bs.back(); }
  int nsys() const { return (int)args.size(); }
  void clear() {
    args.clear();
    bs.assign(1, 0);
  }
};

struct Tables {
  Tab<TspreadArgs> tspread;
  Tab<TsumArgs> tsum;
  Tab<ConvArgs> conv;
  Tab<GatherArgs> gather;
  Tab<IntraArgs> intra;
  Tab<BondedArgs> bonded;
  Tab<PackArgs> pack;
  Tab<WaterArgs> water;
  Tab<WReduceArgs> wred;
  Tab<AtomArgs> atoms;
  Tab<AReduceArgs> ared;
  Tab<FinArgs> fin;
  Tab<EnArgs> en;
};

}  // namespace

struct moil_batch {
  std::vector<moil_ctx *> ctx;
  std::vector<int> order;  // systems by decreasing atom count
  bool f64 = false;
  // real space on cap, the reciprocal sums on cap2: each captured into a
  // graph of its own (one stream each), forked and joined by events
  hipStream_t cap = nullptr, cap2 = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_pme = nullptr;
  hipGraph_t graph = nullptr, graph2 = nullptr;
  hipGraphExec_t exec = nullptr, exec2 = nullptr;
  std::vector<long> gen;           // list generation captured
  std::vector<long> pgen;          // parameter generation captured
  std::vector<const void *> key;   // pointers captured
  std::vector<FftGroup> groups;
  std::vector<int> group_of, slot_of;  // per system: its group, index in it
  Tables tab;
  char *d_blob = nullptr;
  size_t blob_cap = 0;
  // eager: the step's kernels are launched directly on the two streams
  // instead of replaying graphs (for callers that rebuild lists before
  // every evaluation, where each call would otherwise recapture)
  bool eager = false;
  bool have_tables = false;  // eager: tables match gen/pgen/key
  // batched list builds: status words (device, pinned host), fork/join
  // events, allocated on first use
  int32_t *d_stat = nullptr, *h_stat = nullptr;
  hipEvent_t ev_lists = nullptr;
  std::vector<hipEvent_t> ev_built;
  // the transform groups run side by side: group k on stream k % (1 +
  // aux.size()) (0 = the reciprocal-sum stream), forked and joined by
  // events around each direction
  std::vector<hipStream_t> aux;
  hipEvent_t ev_ffork = nullptr, ev_bfork = nullptr;
  std::vector<hipEvent_t> ev_fdone, ev_bdone;
};

void batch_set_eager(moil_batch *b, bool on) {
  b->eager = on;
  b->have_tables = false;
}

// a replay of the old graph may still be in flight on the capture stream:
// it ends before its exec (and the tables it reads) are replaced
static void batch_drop_graph(moil_batch *b) {
  if (b->exec && b->cap) hipStreamSynchronize(b->cap);
  if (b->exec2 && b->cap2) hipStreamSynchronize(b->cap2);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  if (b->exec2) hipGraphExecDestroy(b->exec2);
  if (b->graph2) hipGraphDestroy(b->graph2);
  b->exec = b->exec2 = nullptr;
  b->graph = b->graph2 = nullptr;
}

static void batch_free(moil_batch *b) {
  if (!b) return;
  batch_drop_graph(b);
  for (FftGroup &g : b->groups) {
    if (g.fwd) rocfft_plan_destroy(g.fwd);
    if (g.bwd) rocfft_plan_destroy(g.bwd);
    if (g.info_f) rocfft_execution_info_destroy(g.info_f);
    if (g.info_b) rocfft_execution_info_destroy(g.info_b);
    hipFree(g.work);
    hipFree(g.qreal);
    hipFree(g.qcplx);
  }
  hipFree(b->d_blob);
  for (hipStream_t a : b->aux)
    if (a) hipStreamSynchronize(a);
  hipFree(b->d_stat);
  if (b->h_stat) hipHostFree(b->h_stat);
  for (hipStream_t a : b->aux)
    if (a) hipStreamDestroy(a);
  if (b->ev_ffork) hipEventDestroy(b->ev_ffork);
  if (b->ev_bfork) hipEventDestroy(b->ev_bfork);
  for (hipEvent_t e : b->ev_fdone)
    if (e) hipEventDestroy(e);
  for (hipEvent_t e : b->ev_bdone)
    if (e) hipEventDestroy(e);
  if (b->ev_lists) hipEventDestroy(b->ev_lists);
  for (hipEvent_t e : b->ev_built)
    if (e) hipEventDestroy(e);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_out) hipEventDestroy(b->ev_out);
  if (b->ev_pme) hipEventDestroy(b->ev_pme);
  if (b->cap) hipStreamDestroy(b->cap);
  if (b->cap2) hipStreamDestroy(b->cap2);
  delete b;
}

// batched transforms, one group per grid size
static int32_t batch_fft_groups(moil_batch *b) {
  std::map<std::tuple<int, int, int>, int> idx;
  const int n = (int)b->ctx.size();
  b->group_of.assign(n, -1);
  b->slot_of.assign(n, -1);
  for (int i = 0; i < n; i++) {
    const moil_ctx *c = b->ctx[i];
    auto k = std::make_tuple(c->nf1, c->nf2, c->nf3);
    auto it = idx.find(k);
    int g;
    if (it == idx.end()) {
      g = (int)b->groups.size();
      idx[k] = g;
      b->groups.emplace_back();
      b->groups[g].n1 = c->nf1;
      b->groups[g].n2 = c->nf2;
      b->groups[g].n3 = c->nf3;
    } else {
      g = it->second;
    }
    b->group_of[i] = g;
    b->slot_of[i] = (int)b->groups[g].members.size();
    b->groups[g].members.push_back(i);
  }
  const size_t rs = b->f64 ? sizeof(double) : sizeof(float);
  const rocfft_precision prec = b->f64 ? rocfft_precision_double : rocfft_precision_single;
  RF_TRY(rocfft_setup());
  for (FftGroup &g : b->groups) {
    const size_t m = g.members.size();
    g.G = (size_t)g.n1 * g.n2 * g.n3;
    g.H = (size_t)(g.n1 / 2 + 1) * g.n2 * g.n3;
    HIP_CHECK(hipMalloc(&g.qreal, rs * g.G * m));
    HIP_CHECK(hipMalloc(&g.qcplx, 2 * rs * g.H * m));
    size_t len[3] = {(size_t)g.n1, (size_t)g.n2, (size_t)g.n3};
    // default layouts: contiguous transforms, distance G (real) and H
    // (complex) between the members of the batch
    RF_TRY(rocfft_plan_create(&g.fwd, rocfft_placement_notinplace,
                              rocfft_transform_type_real_forward, prec, 3, len, m, nullptr));
    RF_TRY(rocfft_plan_create(&g.bwd, rocfft_placement_notinplace,
                              rocfft_transform_type_real_inverse, prec, 3, len, m, nullptr));
    size_t w1 = 0, w2 = 0;
    RF_TRY(rocfft_plan_get_work_buffer_size(g.fwd, &w1));
    RF_TRY(rocfft_plan_get_work_buffer_size(g.bwd, &w2));
    const size_t w = std::max(w1, w2);
    if (w) HIP_CHECK(hipMalloc(&g.work, w));
    RF_TRY(rocfft_execution_info_create(&g.info_f));
    RF_TRY(rocfft_execution_info_create(&g.info_b));
    if (w) {
      RF_TRY(rocfft_execution_info_set_work_buffer(g.info_f, g.work, w));
      RF_TRY(rocfft_execution_info_set_work_buffer(g.info_b, g.work, w));
    }
  }
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_batch_create(moil_ctx **ctxs, int32_t n, int32_t nstreams,
                                         moil_batch **out) {
  (void)nstreams;  // one launch per kernel covers the batch
  if (!ctxs || n < 1 || !out) {
    set_error("batch_create: null argument or empty batch");
    return MOIL_GPU_EINVAL;
  }
  for (int i = 0; i < n; i++)
    if (!ctxs[i]) {
      set_error("batch_create: null context");
      return MOIL_GPU_EINVAL;
    }
  for (int i = 1; i < n; i++)
    if (ctxs[i]->par.precision != ctxs[0]->par.precision) {
      set_error("batch_create: all systems of a batch must use the same precision");
      return MOIL_GPU_EINVAL;
    }
  moil_batch *b = new moil_batch();
  b->ctx.assign(ctxs, ctxs + n);
  b->f64 = ctxs[0]->par.precision != MOIL_PREC_MIXED;
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return ctxs[x]->npt > ctxs[y]->npt; });
  bool ok = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->cap2, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_pme, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    batch_free(b);
    set_error("batch_create: stream/event creation failed");
    return MOIL_GPU_EHIP;
  }
  const int32_t rc = batch_fft_groups(b);
  if (rc != MOIL_GPU_OK) {
    batch_free(b);
    return rc;
  }
  // up to three more streams for the transform groups
  const int ng = (int)b->groups.size();
  const int naux = std::min(ng - 1, 3);
  bool ok2 = hipEventCreateWithFlags(&b->ev_ffork, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b->ev_bfork, hipEventDisableTiming) == hipSuccess;
  b->aux.assign(std::max(naux, 0), nullptr);
  for (hipStream_t &a : b->aux)
    ok2 = ok2 && hipStreamCreateWithFlags(&a, hipStreamNonBlocking) == hipSuccess;
  b->ev_fdone.assign(ng, nullptr);
  b->ev_bdone.assign(ng, nullptr);
  for (int k = 0; k < ng; k++)
    ok2 = ok2 && hipEventCreateWithFlags(&b->ev_fdone[k], hipEventDisableTiming) == hipSuccess &&
          hipEventCreateWithFlags(&b->ev_bdone[k], hipEventDisableTiming) == hipSuccess;
  if (!ok2) {
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
    hipStreamSynchronize(b->cap2);
    batch_free(b);
  }
  return MOIL_GPU_OK;
}

template <class A>
static void blob_layout(Tab<A> &t, size_t &off) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  t.off_a = off;
  off = up(off + sizeof(A) * std::max<size_t>(t.args.size(), 1));
  t.off_b = off;
  off = up(off + sizeof(int32_t) * t.bs.size());
}
template <class A>
static void blob_fill(const Tab<A> &t, std::vector<char> &h) {
  if (!t.args.empty()) memcpy(h.data() + t.off_a, t.args.data(), sizeof(A) * t.args.size());
  memcpy(h.data() + t.off_b, t.bs.data(), sizeof(int32_t) * t.bs.size());
}
#define TAB_PTRS(b, t) (decltype(b->tab.t.args.data()))(b->d_blob + b->tab.t.off_a), \
                       (const int32_t *)(b->d_blob + b->tab.t.off_b), b->tab.t.nsys(), b->tab.t.total()

// argument tables of every kernel for the current lists and pointers
static int32_t batch_tables(moil_batch *b, const double *const *d_coor, double *const *d_dpot,
                            double *const *d_energies) {
  Tables &T = b->tab;
  T = Tables();
  const size_t rs = b->f64 ? sizeof(double) : sizeof(float);
  for (int i : b->order) {
    moil_ctx *c = b->ctx[i];
    EvalPlan p;
    eval_plan(c, d_coor[i], d_dpot[i], d_energies[i], &p);
    if (p.pme) {
      const FftGroup &g = b->groups[b->group_of[i]];
      const size_t k = (size_t)b->slot_of[i];
      void *qr = (char *)g.qreal + rs * g.G * k;
      void *qc = (char *)g.qcplx + 2 * rs * g.H * k;
      p.tsum.grid = qr;
      p.conv.S = qc;
      p.gather.grid = qr;
      T.tspread.add(p.tspread, p.n_tile);
      T.tsum.add(p.tsum, p.n_tsum);
      T.conv.add(p.conv, p.n_conv);
      T.gather.add(p.gather, p.n_tile);
    }
    T.intra.add(p.intra, p.n_intra);
    T.bonded.add(p.bonded, p.n_bonded);
    T.pack.add(p.pack, p.n_pack);
    T.water.add(p.water, p.n_water);
    T.wred.add(p.wred, p.n_wred);
    T.atoms.add(p.atoms, p.n_atoms);
    T.ared.add(p.ared, p.n_ared);
    T.fin.add(p.fin, p.n_fin);
    T.en.add(p.en, EN_COUNT);
  }
  size_t off = 0;
  blob_layout(T.tspread, off); blob_layout(T.tsum, off); blob_layout(T.conv, off);
  blob_layout(T.gather, off); blob_layout(T.intra, off); blob_layout(T.bonded, off);
  blob_layout(T.pack, off); blob_layout(T.water, off); blob_layout(T.wred, off);
  blob_layout(T.atoms, off); blob_layout(T.ared, off); blob_layout(T.fin, off);
  blob_layout(T.en, off);
  std::vector<char> h(off, 0);
  blob_fill(T.tspread, h); blob_fill(T.tsum, h); blob_fill(T.conv, h);
  blob_fill(T.gather, h); blob_fill(T.intra, h); blob_fill(T.bonded, h);
  blob_fill(T.pack, h); blob_fill(T.water, h); blob_fill(T.wred, h);
  blob_fill(T.atoms, h); blob_fill(T.ared, h); blob_fill(T.fin, h);
  blob_fill(T.en, h);
  if (off > b->blob_cap) {
    HIP_CHECK(hipStreamSynchronize(b->cap));
    hipFree(b->d_blob);
    b->d_blob = nullptr;
    b->blob_cap = 0;
    HIP_CHECK(hipMalloc(&b->d_blob, off));
    b->blob_cap = off;
  }
  // the previous graph (which reads the blob) has finished: the capture
  // stream was synchronised or every replay precedes this copy on it
  HIP_CHECK(hipMemcpyAsync(b->d_blob, h.data(), off, hipMemcpyHostToDevice, b->cap));
  HIP_CHECK(hipStreamSynchronize(b->cap));
  return MOIL_GPU_OK;
}

// one direction of every group's transform, the groups spread over st and
// the auxiliary streams (forked from st, joined back into st)
static int32_t batch_transforms(moil_batch *b, hipStream_t st, bool forward) {
  const int ns = 1 + (int)b->aux.size();
  hipEvent_t fork = forward ? b->ev_ffork : b->ev_bfork;
  std::vector<hipEvent_t> &done = forward ? b->ev_fdone : b->ev_bdone;
  if (ns > 1) HIP_CHECK(hipEventRecord(fork, st));
  for (size_t k = 0; k < b->groups.size(); k++) {
    FftGroup &g = b->groups[k];
    const int si = (int)(k % ns);
    hipStream_t s = si == 0 ? st : b->aux[si - 1];
    if (si != 0 && k < (size_t)ns) HIP_CHECK(hipStreamWaitEvent(s, fork, 0));
    rocfft_execution_info info = forward ? g.info_f : g.info_b;
    void *in[1] = {forward ? g.qreal : g.qcplx}, *out[1] = {forward ? g.qcplx : g.qreal};
    RF_TRY(rocfft_execution_info_set_stream(info, s));
    RF_TRY(rocfft_execute(forward ? g.fwd : g.bwd, in, out, info));
    // the last group of each auxiliary stream signals its completion
    if (si != 0 && k + ns >= b->groups.size()) HIP_CHECK(hipEventRecord(done[k], s));
  }
  for (size_t k = 0; k < b->groups.size(); k++) {
    const int si = (int)(k % ns);
    if (si != 0 && k + ns >= b->groups.size()) HIP_CHECK(hipStreamWaitEvent(st, done[k], 0));
  }
  return MOIL_GPU_OK;
}

// the batch step: every kernel once over all systems, transforms per
// group; the reciprocal sums (spreading, transforms, convolution, gather)
// and the real-space part are independent until the force assembly
static int32_t batch_enqueue_pme(moil_batch *b, hipStream_t st) {
  const bool f = b->f64;
  if (f) {
    launch_k_pme_tspread_f64_multi(TAB_PTRS(b, tspread), st);
    launch_k_pme_tsum_f64_multi(TAB_PTRS(b, tsum), st);
  } else {
    launch_k_pme_tspread_f32_multi(TAB_PTRS(b, tspread), st);
    launch_k_pme_tsum_f32_multi(TAB_PTRS(b, tsum), st);
  }
  HIP_CHECK(hipGetLastError());
  if (b->tab.tspread.nsys() > 0) CHECK_RC_B(batch_transforms(b, st, true));
  if (f) launch_k_conv_f64_multi(TAB_PTRS(b, conv), st);
  else launch_k_conv_f32_multi(TAB_PTRS(b, conv), st);
  if (b->tab.tspread.nsys() > 0) CHECK_RC_B(batch_transforms(b, st, false));
  if (f) launch_k_gather_f64_multi(TAB_PTRS(b, gather), st);
  else launch_k_gather_f32_multi(TAB_PTRS(b, gather), st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

static int32_t batch_enqueue_rs(moil_batch *b, hipStream_t st) {
  const bool f = b->f64;
  launch_k_water_intra_multi(TAB_PTRS(b, intra), st);
  launch_k_bonded_multi(TAB_PTRS(b, bonded), st);
  if (f) {
    launch_k_pack_water_f64_multi(TAB_PTRS(b, pack), st);
    launch_k_water_f64_multi(TAB_PTRS(b, water), st);
  } else {
    launch_k_pack_water_f32_multi(TAB_PTRS(b, pack), st);
    launch_k_water_f32_multi(TAB_PTRS(b, water), st);
  }
  launch_k_water_reduce_multi(TAB_PTRS(b, wred), st);
  if (f) launch_k_atoms_f64_multi(TAB_PTRS(b, atoms), st);
  else launch_k_atoms_f32_multi(TAB_PTRS(b, atoms), st);
  launch_k_atoms_reduce_multi(TAB_PTRS(b, ared), st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

static int32_t batch_enqueue_fin(moil_batch *b, hipStream_t st) {
  launch_k_finalize_forces_multi(TAB_PTRS(b, fin), st);
  launch_k_energies_multi(TAB_PTRS(b, en), st);
  HIP_CHECK(hipGetLastError());
  return MOIL_GPU_OK;
}

// one stream's work captured into a graph and instantiated
static int32_t batch_capture(hipStream_t cs, int32_t (*enq)(moil_batch *, hipStream_t),
                             moil_batch *b, hipGraph_t *graph, hipGraphExec_t *exec) {
  HIP_CHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  const int32_t rc = enq(b, cs);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &g);
  if (rc != MOIL_GPU_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  HIP_CHECK(e);
  *graph = g;
  HIP_CHECK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return MOIL_GPU_OK;
}

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
  bool fresh = (b->eager ? b->have_tables : b->exec != nullptr) && (int)b->gen.size() == n &&
               (int)b->pgen.size() == n;
  for (int i = 0; fresh && i < n; i++)
    fresh = b->gen[i] == b->ctx[i]->list_gen && b->pgen[i] == b->ctx[i]->param_gen &&
            b->key[3 * i] == d_coor[i] &&
            b->key[3 * i + 1] == d_dpot[i] && b->key[3 * i + 2] == d_energies[i];
  HIP_CHECK(hipEventRecord(b->ev_in, st));
  HIP_CHECK(hipStreamWaitEvent(b->cap, b->ev_in, 0));
  HIP_CHECK(hipStreamWaitEvent(b->cap2, b->ev_in, 0));
  if (!fresh) {
    batch_drop_graph(b);
    if (b->eager) {
      // the previous call's kernels read the argument tables about to be
      // replaced
      HIP_CHECK(hipStreamSynchronize(b->cap));
      HIP_CHECK(hipStreamSynchronize(b->cap2));
      b->have_tables = false;
    }
    // one-time accumulator clears (memsets) stay out of the graph
    for (int i = 0; i < n; i++)
      if (b->ctx[i]->facc_dirty) {
        HIP_CHECK(hipMemsetAsync(b->ctx[i]->d_facc, 0, sizeof(long long) * 3 * b->ctx[i]->npt,
                                 b->cap));
        b->ctx[i]->facc_dirty = false;
      }
    CHECK_RC_B(batch_tables(b, d_coor, d_dpot, d_energies));
    if (b->eager) {
      b->have_tables = true;
    } else {
      CHECK_RC_B(batch_capture(b->cap, batch_enqueue_rs, b, &b->graph, &b->exec));
      CHECK_RC_B(batch_capture(b->cap2, batch_enqueue_pme, b, &b->graph2, &b->exec2));
    }
    b->gen.resize(n);
    b->pgen.resize(n);
    b->key.resize(3 * n);
    for (int i = 0; i < n; i++) {
      b->gen[i] = b->ctx[i]->list_gen;
      b->pgen[i] = b->ctx[i]->param_gen;
      b->key[3 * i] = d_coor[i];
      b->key[3 * i + 1] = d_dpot[i];
      b->key[3 * i + 2] = d_energies[i];
    }
  }
  // the real-space part (a few launches) is submitted first so that it
  // starts while the reciprocal part's many transform launches are still
  // being submitted
  if (b->eager) {
    CHECK_RC_B(batch_enqueue_rs(b, b->cap));
    CHECK_RC_B(batch_enqueue_pme(b, b->cap2));
  } else {
    HIP_CHECK(hipGraphLaunch(b->exec, b->cap));
    HIP_CHECK(hipGraphLaunch(b->exec2, b->cap2));
  }
  HIP_CHECK(hipEventRecord(b->ev_pme, b->cap2));
  HIP_CHECK(hipStreamWaitEvent(b->cap, b->ev_pme, 0));
  CHECK_RC_B(batch_enqueue_fin(b, b->cap));
  HIP_CHECK(hipEventRecord(b->ev_out, b->cap));
  HIP_CHECK(hipStreamWaitEvent(st, b->ev_out, 0));
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_batch_set_graph(moil_batch *b, int32_t on) {
  if (!b) {
    set_error("batch_set_graph: null batch");
    return MOIL_GPU_EINVAL;
  }
  if (b->cap) hipStreamSynchronize(b->cap);
  if (b->cap2) hipStreamSynchronize(b->cap2);
  batch_drop_graph(b);
  batch_set_eager(b, on == 0);
  b->gen.clear();  // next evaluation rebuilds its tables
  return MOIL_GPU_OK;
}

extern "C" int32_t moil_gpu_batch_build_lists_dev(moil_batch *b, double *const *d_coor,
                                                  int32_t squeeze, void *stream) {
  if (!b || !d_coor) {
    set_error("batch_build_lists: null argument");
    return MOIL_GPU_EINVAL;
  }
  const int n = (int)b->ctx.size();
  for (int i = 0; i < n; i++)
    if (!d_coor[i]) {
      set_error("batch_build_lists: null device pointer");
      return MOIL_GPU_EINVAL;
    }
  if (!b->d_stat) {
    HIP_CHECK(hipMalloc(&b->d_stat, sizeof(int32_t) * ST_COUNT * n));
    HIP_CHECK(hipHostMalloc(&b->h_stat, sizeof(int32_t) * ST_COUNT * n, hipHostMallocDefault));
    HIP_CHECK(hipEventCreateWithFlags(&b->ev_lists, hipEventDisableTiming));
    b->ev_built.assign(n, nullptr);
    for (int i = 0; i < n; i++)
      HIP_CHECK(hipEventCreateWithFlags(&b->ev_built[i], hipEventDisableTiming));
  }
  return build_lists_many(b->ctx.data(), d_coor, n, squeeze != 0, (hipStream_t)stream, b->d_stat,
                          b->h_stat, b->ev_lists, b->ev_built.data());
}
