// grad_noalloc.hip — ctc_grad_noalloc: the loss adjoint of ctc_grad for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops).  Its only scratch memory is the file-scope device array `scratch`
// below (partial sums of the loss, kGradScratchDoubles doubles per GPU, held
// until the process ends); every call on one GPU shares it, so calls that
// overlap on one GPU race.  The call allocates nothing, issues no memset and
// does not synchronise: it enqueues fill, walk, regulariser and reduction
// kernels on the caller's stream and writes the loss to a device double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "ctc_internal.h"
#include "kernels_launch.h"
#include "model_util.h"

namespace {

// 2 groups of partial sums per mode (walk, regulariser) x 4 modes x 16384
// blocks: 1 MiB
constexpr uint64_t kGradMaxBlocks = 16384;
constexpr uint64_t kGradScratchDoubles = 2 * CTC_MAX_MODES * kGradMaxBlocks + 2 * CTC_MAX_MODES;
constexpr unsigned kRegBlocks = 512;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kGradScratchDoubles];

std::mutex g_lock;
double *g_scratch[kMaxDevices] = {nullptr};

// device address of `scratch` on the current device, looked up once per device
ctc_status scratch_ptr(int device, double **out) {
  if (device < 0 || device >= kMaxDevices) return ctc::fail(CTC_ERR_INVALID_ARG, "device %d", device);
  std::lock_guard<std::mutex> g(g_lock);
  if (!g_scratch[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return ctc::fail(CTC_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch[device] = (double *)p;
  }
  *out = g_scratch[device];
  return CTC_OK;
}

// loss = 0.5 * (squared residuals of the mode-0 walk) + the regularisers
__global__ void k_grad_loss(const double *sums, int N, double *loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double l = 0.5 * sums[0];
    for (int m = 0; m < N; ++m) l += sums[N + m];
    *loss = l;
  }
}

}  // namespace

using namespace ctc;

extern "C" ctc_status ctc_grad_noalloc(ctc_model *m, const float *const *factors, float *const *grads,
                                       double lambda, int reg_per_occurrence, double *loss_device, void *stream) {
  if (!m || !factors || !grads) return fail(CTC_ERR_INVALID_ARG, "ctc_grad_noalloc: null argument");
  CTC_HIP(hipSetDevice(m->ctx->device));
  double *part = nullptr;
  CTC_TRY(scratch_ptr(m->ctx->device, &part));
  const uint32_t N = m->nmodes, R = m->rank;
  hipStream_t st = stream ? (hipStream_t)stream : m->ctx->stream;
  float *F[CTC_MAX_MODES] = {nullptr, nullptr, nullptr, nullptr};
  for (uint32_t md = 0; md < N; ++md) {
    if (!factors[md] || !grads[md]) return fail(CTC_ERR_INVALID_ARG, "ctc_grad_noalloc: null factor/grad");
    if (!m->ciss_ready[md])
      return fail(CTC_ERR_INVALID_ARG,
                  "ctc_grad_noalloc: the CISS of mode %u is not built (call ctc_model_prepare with ALS or CCD++ "
                  "options, or ctc_grad once, before)", md);
    F[md] = (float *)factors[md];
  }
  // walkers of `span` nonzeros, span chosen so that each walk has at most
  // kGradMaxBlocks blocks of partial sums
  const uint64_t per_block = mttkrp_walkers_per_block((int)R, 256);
  uint64_t span = 256;
  while ((m->train_nnz + span - 1) / span > kGradMaxBlocks * per_block) span *= 2;
  const uint64_t nb = std::max<uint64_t>(mttkrp_grid(m->train_nnz, (int)R, (uint32_t)span, 256), kRegBlocks);
  if (nb > kGradMaxBlocks || 2 * N * nb + 2 * N > kGradScratchDoubles)
    return fail(CTC_ERR_INVALID_ARG, "ctc_grad_noalloc: %llu partial sums exceed the %llu of its fixed scratch array",
                (unsigned long long)(2 * N * nb), (unsigned long long)kGradScratchDoubles);
  double *sums = part + 2 * N * nb;
  // unused tails of the walk groups must read as zero: fill kernel, no memset
  CTC_HIP(launch_fill(st, (float *)part, 2 * (2 * N * nb), 0.f));
  for (uint32_t md = 0; md < N; ++md) {
    CTC_HIP(launch_fill(st, grads[md], m->dims[md] * R, 0.f));
    unsigned grid = 0;
    CTC_HIP(launch_mttkrp(st, ciss_args(m->ciss[md]), rel_factors(F, N, md), (int)N, (int)R, (uint32_t)span, true,
                          grads[md], part + (size_t)md * nb, &grid, 256));
    const DevCiss &d = m->ciss[md];
    const uint64_t nrows = reg_per_occurrence ? d.dlength : m->dims[md];
    CTC_HIP(launch_reg_rows(st, d.directory, d.rowptr, nrows, (int)R, (float)lambda, reg_per_occurrence, F[md],
                            grads[md], kRegBlocks, part + (size_t)(N + md) * nb));
  }
  CTC_HIP(launch_reduce_parts(st, part, (uint32_t)nb, 2 * N, sums));
  if (loss_device) {
    hipLaunchKernelGGL(k_grad_loss, dim3(1), dim3(64), 0, st, sums, (int)N, loss_device);
    CTC_HIP(hipGetLastError());
  }
  return CTC_OK;
}
