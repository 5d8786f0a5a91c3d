// qesmc_run_static: qesmc_run for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  qesmc_run allocates the convergence rows (kErrSlots words per
// iteration) when a run asks for more iterations than any earlier run of the
// context; this variant takes them from a fixed, file-scope device array
// instead and calls no allocating routine.  Everything else (kernels,
// resident arrays, results) is qesmc_run's.
//
// The array holds kStaticMaxIter + 1 rows (16.8 MB), one copy per device the
// library runs on, for the life of the process.  Calls on one device share
// it: two qesmc_run_static calls that overlap on one device (contexts on
// different streams or threads) race on it.  The rows are cleared by a fill
// kernel at the start of every call.
#include <mutex>

#include "qesmc_internal.h"

namespace {

constexpr int kStaticMaxIter = 65536;
constexpr long long kStaticWords = (long long)(kStaticMaxIter + 1) * qesmc::kErrSlots;
constexpr int kMaxDevices = 64;

// The convergence rows of qesmc_run_static.  In host code this name is only
// a stand-in for the symbol: its device address (per device) comes from
// hipGetSymbolAddress and is the only thing ever passed to kernels or copies.
__device__ __align__(256) unsigned scratch[kStaticWords];

std::mutex g_mu;
unsigned *g_scratch[kMaxDevices] = {};

}  // namespace

namespace qesmc {

// The device address of this device's copy of `scratch`, looked up once per
// device with that device current (called by qesmc_create, which has made
// it current, and again by qesmc_run_static as a guard).
int static_scratch_init(int device)
{
  if (device < 0 || device >= kMaxDevices) return QESMC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return QESMC_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return e != hipSuccess ? map_hip_error(e) : QESMC_ERR_HIP;
  g_scratch[device] = (unsigned *)p;
  return QESMC_OK;
}

}  // namespace qesmc

extern "C" {

int qesmc_run_static(qesmc_ctx *x, const qesmc_params *pp, qesmc_result *res)
{
  if (!x || !pp) return QESMC_ERR_ARG;
  if (pp->itermax > kStaticMaxIter) return QESMC_ERR_LIMIT;
  if (pp->path == QESMC_PATH_FUSED && (!x->d_lc2[0] || !x->d_lc2[1] || !x->d_cur || !x->d_canon8))
    return QESMC_ERR_STATE;
  // refine_every allocates its base and R copies on first use
  if (pp->refine_every != 0) return QESMC_ERR_UNSUPPORTED;
  if (hipSetDevice(x->device) != hipSuccess) return QESMC_ERR_NO_DEVICE;
  int rc = qesmc::static_scratch_init(x->device);
  if (rc != QESMC_OK) return rc;
  unsigned *rows = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    rows = g_scratch[x->device];
  }
  // the context's own rows are set aside for the call: run_prepare sees
  // enough capacity and allocates nothing, and every kernel of the run takes
  // the scratch rows as its convergence buffer
  unsigned *own = x->d_err;
  const int own_cap = x->err_cap;
  x->d_err = rows;
  x->err_cap = kStaticMaxIter + 1;
  x->no_alloc_run = true;
  rc = qesmc::run_solve(x, pp, res);
  x->no_alloc_run = false;
  x->d_err = own;
  x->err_cap = own_cap;
  return rc;
}

}  // extern "C"
