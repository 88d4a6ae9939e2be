// Force evaluation that allocates nothing (moil_gpu_eforce_noalloc_dev).
//
// The evaluation of moil_gpu_eforce_dev launched directly (never captured
// or replayed as a graph; the reciprocal sum forks to the context's second
// stream and joins back through the context's own events, created at
// init), with its reduction scratch in a file-scope device array instead
// of the context's buffers: the per-workgroup energy partial sums, the
// per-term totals, the last-workgroup ticket of the energy reduction and
// the fixed-point overflow word.  Nothing is allocated or freed during the
// call; the context's own buffers (lists, accumulators, grids, transform
// plans) were allocated by moil_gpu_init and moil_gpu_build_lists*, and
// no timing events are taken.
//
// Layout of the array (SCRATCH_BYTES):
//   [0, 8)        ticket, overflow word (unsigned int each), cleared by
//                 every call before its kernels run
//   [256, 352)    per-term totals, double[EN_COUNT]
//   [512, ...)    partial sums, double[EN_COUNT][need], need = the largest
//                 number of partial sums of one term in this evaluation
// Every call on one GPU uses that GPU's one copy of the array: calls that
// overlap in time on one GPU (from two threads, or on two streams without
// an order between them) race on it.
#include <mutex>

#include "kargs.h"

#define SCRATCH_BYTES (8u << 20)
#define SCRATCH_TOT 256
#define SCRATCH_PART 512
#define SCRATCH_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// the array's device address on each GPU (hipGetSymbolAddress with that
// GPU current); the host-side name `scratch` is not a device address
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[SCRATCH_MAX_DEVICES];

int32_t scratch_attach(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
    set_error("device " + std::to_string(dev) + " beyond the evaluation scratch table (" +
              std::to_string(SCRATCH_MAX_DEVICES) + " devices)");
    return MOIL_GPU_EHIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return MOIL_GPU_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress of the evaluation scratch failed: ") +
              hipGetErrorString(e));
    return MOIL_GPU_EHIP;
  }
  g_scratch_ptr[dev] = p;
  return MOIL_GPU_OK;
}

static void *scratch_of(int dev) {
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[dev];
}

// ticket and overflow word to zero; with n > 0 also the force accumulator
// (left dirty by an evaluation that was not completed)
__global__ void __launch_bounds__(256) k_scratch_clear(unsigned int *hdr, long long *facc,
                                                       size_t n) {
  const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (g < 2) hdr[g] = 0u;
  for (size_t i = g; i < n; i += (size_t)gridDim.x * 256) facc[i] = 0;
}

// the context's reduction buffers, pointed at the scratch for one call and
// put back at its end (also on an error return)
struct ScratchSwap {
  moil_ctx *c;
  double *epart, *etot;
  unsigned int *ticket;
  int *ovf;
  int max_parts, timing;
  ScratchSwap(moil_ctx *c_, unsigned char *base, int need) : c(c_) {
    epart = c->d_epart; etot = c->d_etot; ticket = c->d_ticket; ovf = c->d_ovf;
    max_parts = c->max_parts; timing = c->timing;
    c->d_ticket = reinterpret_cast<unsigned int *>(base);
    c->d_ovf = reinterpret_cast<int *>(base) + 1;
    c->d_etot = reinterpret_cast<double *>(base + SCRATCH_TOT);
    c->d_epart = reinterpret_cast<double *>(base + SCRATCH_PART);
    c->max_parts = need;
    c->timing = 0;  // no timing events
  }
  ~ScratchSwap() {
    c->d_epart = epart; c->d_etot = etot; c->d_ticket = ticket; c->d_ovf = ovf;
    c->max_parts = max_parts; c->timing = timing;
  }
};

extern "C" int32_t moil_gpu_eforce_noalloc_dev(moil_ctx *c, const double *d_coor, double *d_dpot,
                                               double *d_energies, void *stream) {
  if (!c || !d_coor || !d_dpot) {
    set_error("null argument");
    return MOIL_GPU_EINVAL;
  }
  if (!c->have_lists) {
    set_error("eforce called before build_lists");
    return MOIL_GPU_ESTATE;
  }
  unsigned char *base = static_cast<unsigned char *>(scratch_of(c->device));
  if (!base) {
    set_error("evaluation scratch not attached to device " + std::to_string(c->device));
    return MOIL_GPU_ESTATE;
  }
  // partial sums per energy term: the real-space chunks, the reciprocal
  // convolution blocks, the bonded and intra-water blocks
  const int need =
      std::max(std::max(c->nwchunk + c->nachunk, pme_conv_blocks(c)), bonded_max_blocks(c));
  const size_t maxp = (SCRATCH_BYTES - SCRATCH_PART) / (sizeof(double) * EN_COUNT);
  if ((size_t)need > maxp) {
    set_error("eforce_noalloc: the evaluation needs " + std::to_string(need) +
              " partial sums per energy term; the static scratch (" +
              std::to_string(SCRATCH_BYTES >> 20) + " MiB) holds at most " +
              std::to_string(maxp) + " (use moil_gpu_eforce_dev)");
    return MOIL_GPU_ECAPACITY;
  }
  hipStream_t st = pick_stream(c, stream);
  const size_t nf = c->facc_dirty ? 3 * (size_t)c->npt : 0;
  const int nblk = nf ? (int)std::min<size_t>(1024, (nf + 255) / 256) : 1;
  k_scratch_clear<<<nblk, 256, 0, st>>>(reinterpret_cast<unsigned int *>(base), c->d_facc, nf);
  HIP_CHECK(hipGetLastError());
  c->facc_dirty = false;
  ScratchSwap swap(c, base, need);
  EvalPlan p;
  eval_plan(c, d_coor, d_dpot, d_energies, &p);
  c->facc_dirty = true;
  const int32_t rc = eforce_eager(c, p, st);
  if (rc != MOIL_GPU_OK) return rc;
  c->facc_dirty = false;
  return MOIL_GPU_OK;
}
