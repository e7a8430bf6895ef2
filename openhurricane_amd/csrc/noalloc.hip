// noalloc.hip -- ohc_chem_solve_update_noalloc: solveUpdate for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops).  The integrator's work-queue head, its counters and the
// longest-first schedule (per-cell cost, cell order, counting-sort buckets)
// live in the static device array `scratch` below instead of buffers the
// library allocates.  Its device address is looked up once per device with
// hipGetSymbolAddress; host code uses that address only.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "ctx_internal.hpp"
#include "sched.hpp"

namespace {

constexpr int64_t kMaxCells = OHC_NOALLOC_MAX_CELLS;
constexpr size_t kCtlWords = 16;                  // [0] queue head, [1..8] counters
constexpr size_t kBinWords = ohcs::kBins;         // 2 * kBins unsigned ints
constexpr size_t kCellWords = (size_t)kMaxCells;  // cost[] and order[], int each
constexpr size_t kWords = kCtlWords + kBinWords + kCellWords;

}  // namespace

static __device__ __align__(256) unsigned long long scratch[kWords];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];      // device address of scratch, per device
int64_t g_last_n[kMaxDevices];  // n of the last scheduled call, per device (0: none)

// Queue head and counters to their start values (dt_min = +inf), by a
// kernel rather than a memset so that the call can be captured in a graph.
__global__ void noalloc_reset(unsigned long long *ctl) {
    const int i = threadIdx.x;
    if (i < (int)kCtlWords) ctl[i] = (i == 7) ? 0x7ff0000000000000ull : 0ull;
}

int scratch_ptr(ohc_ctx *c, unsigned long long **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device < 0 || c->device >= kMaxDevices)
        return ohci::fail(c, OHC_ERR_ARG, "device index out of range");
    if (!g_ptr[c->device]) {
        void *p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return ohci::fail(c, OHC_ERR_DEVICE, "no device address for the scratch array");
        g_ptr[c->device] = p;
        g_last_n[c->device] = 0;
    }
    *out = (unsigned long long *)g_ptr[c->device];
    return OHC_OK;
}

}  // namespace

extern "C" int ohc_chem_solve_update_noalloc(ohc_ctx *c, int64_t n, double *T, const double *p,
                                             const double *rho, double *Y, int64_t ldY,
                                             const double *dt, double dtFactor, double *subDt,
                                             ohc_stats *stats) {
    int rc = ohci::check_ctx(c);
    if (rc) return rc;
    if (!c->opts.device_pointers)
        return ohci::fail(c, OHC_ERR_ARG,
                          "ohc_chem_solve_update_noalloc needs a context with device_pointers = 1");
    if (n < 0) return ohci::fail(c, OHC_ERR_ARG, "n < 0");
    if (n > kMaxCells)
        return ohci::fail(c, OHC_ERR_ARG,
                          "n = " + std::to_string(n) +
                              " exceeds OHC_NOALLOC_MAX_CELLS (" + std::to_string(kMaxCells) +
                              "); split the cells into calls of at most that many");
    if (n == 0) {
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->dt_min = 1.0e15;
        }
        return OHC_OK;
    }
    if (ldY < n) return ohci::fail(c, OHC_ERR_ARG, "leading dimension smaller than n");
    if (!T || !p || !rho || !Y || !dt || !subDt)
        return ohci::fail(c, OHC_ERR_ARG, "null array argument");
    if (!(dtFactor > 0)) return ohci::fail(c, OHC_ERR_ARG, "dtFactor must be > 0");
    if (!c->lane && !c->wave)
        return ohci::fail(c, OHC_ERR_UNSUPPORTED, "no device integrator for this mechanism");
    unsigned long long *base = nullptr;
    if ((rc = scratch_ptr(c, &base))) return rc;
    unsigned long long *ctl = base;
    unsigned *bins = (unsigned *)(base + kCtlWords);
    int *cost = (int *)(base + kCtlWords + kBinWords);
    int *order = cost + kMaxCells;

    ohcd::LaneParams P;
    ohci::base_params(c, P);
    P.n = n;
    P.T = T;
    P.p = p;
    P.rho = rho;
    P.Y = Y;
    P.dt = dt;
    P.ldY = ldY;
    P.dtFactor = dtFactor;
    P.subDt = subDt;
    P.Tout = T;
    P.Yout = Y;
    P.mode = 1;
    P.counter = ctl;
    P.stats = ctl + 1;
    bool ordered = false;
    if (c->opts.schedule) {
        // the costs in scratch belong to the last scheduled call on this
        // device; they order this call only when it had the same n
        {
            std::lock_guard<std::mutex> lk(g_mu);
            ordered = g_last_n[c->device] == n;
            g_last_n[c->device] = n;
        }
        if (ordered) HIPCHK(c, ohcs::order_by_cost(cost, order, bins, n, c->stream));
        P.order = ordered ? order : nullptr;
        P.cost = cost;
    }
    hipLaunchKernelGGL(noalloc_reset, dim3(1), dim3(64), 0, c->stream, ctl);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, ohci::launch_integrator(c, P));
    if (!stats) return OHC_OK;
    unsigned long long h[8];
    HIPCHK(c, hipMemcpyAsync(h, ctl + 1, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ohci::stats_from_words(c, h, stats);
}
