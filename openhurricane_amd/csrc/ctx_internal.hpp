// ctx_internal.hpp -- definitions shared by the translation units behind the
// C-ABI (mechanism and context objects, error reporting, host staging, the
// integrator driver).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/ohc.h"
#include "lane_kernels.hpp"
#include "mech.hpp"
#include "wave_kernels.hpp"

struct ohc_mech {
    ohc::Mechanism m;
};

struct ohc_ctx {
    ohc::Mechanism mech;
    ohc_opts opts;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    const ohcd::LaneEntry *lane = nullptr;  // specialised kernels for this mechanism
    bool lane_transport = false;
    bool wave = false;          // generic wavefront-per-cell kernels (run-time tables)
    ohcd::WMech wm{};           // device tables of the generic kernels
    void *d_wm = nullptr;       // one allocation holding them
    int grid = 0;  // persistent integrator grid (single-wavefront blocks)
    int cus = 0;         // compute units of the device
    int max_per_cu = 0;  // resident integrator blocks per CU the kernel allows
    unsigned long long *d_ctl = nullptr;  // [0] queue head, [1..8] stats
    std::vector<void *> stage;            // host-pointer staging buffers
    std::vector<size_t> stage_sz;
    // longest-first schedule: per-cell cost of the last call and its order
    int64_t sched_n = 0;
    bool sched_ready = false;
    int *d_cost = nullptr, *d_order = nullptr;
    unsigned *d_bins = nullptr;  // counting-sort buckets (2 * ohcs::kBins)
    bool in_slices = false;  // inside a sliced host-pointer call (one slice at a time)
    std::string err, kname;
};

namespace ohci {

int fail(ohc_ctx *c, int code, const std::string &msg);

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ohci::fail(ctx, OHC_ERR_DEVICE,                                          \
                              std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// Device staging for host pointers: slot i holds at least bytes.
inline int stage_buf(ohc_ctx *c, size_t slot, size_t bytes, void **out) {
    if (c->stage.size() <= slot) {
        c->stage.resize(slot + 1, nullptr);
        c->stage_sz.resize(slot + 1, 0);
    }
    if (c->stage_sz[slot] < bytes) {
        if (c->stage[slot]) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipFree(c->stage[slot]));
        }
        c->stage[slot] = nullptr;
        c->stage_sz[slot] = 0;
        HIPCHK(c, hipMalloc(&c->stage[slot], bytes));
        c->stage_sz[slot] = bytes;
    }
    *out = c->stage[slot];
    return OHC_OK;
}

// Bring one array to the device (host mode) or pass it through (device mode).
// A block of rows (species fields with leading dimension ld) is copied row by
// row, n elements each, so that the ld - n elements between two rows of the
// caller's array are never written.
struct Arr {
    const void *host = nullptr;
    void *dev = nullptr;
    size_t bytes = 0;
    size_t rows = 1, width = 0, pitch = 0;  // width/pitch in bytes when rows > 1
};

class Staging {
   public:
    explicit Staging(ohc_ctx *c) : c_(c) {}
    // in: copy host->device; out: copy device->host at finish()
    int map(const void *p, size_t bytes, bool in, bool out, void **dev) {
        if (!p) {
            *dev = nullptr;
            return OHC_OK;
        }
        if (c_->opts.device_pointers) {
            *dev = const_cast<void *>(p);
            return OHC_OK;
        }
        void *d = nullptr;
        int rc = stage_buf(c_, slot_++, bytes, &d);
        if (rc) return rc;
        if (in) HIPCHK(c_, hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, c_->stream));
        if (out) arrs_.push_back(Arr{p, d, bytes, 1, bytes, bytes});
        *dev = d;
        return OHC_OK;
    }
    // rows x n elements of elem bytes with leading dimension ld (>= n)
    int map_rows(const void *p, int64_t rows, int64_t n, int64_t ld, size_t elem, bool in,
                 bool out, void **dev) {
        const size_t bytes = elem * (size_t)(ld * (rows - 1) + n);
        if (ld == n || rows <= 1 || !p || c_->opts.device_pointers)
            return map(p, bytes, in, out, dev);
        void *d = nullptr;
        int rc = stage_buf(c_, slot_++, bytes, &d);
        if (rc) return rc;
        const size_t w = elem * (size_t)n, pitch = elem * (size_t)ld;
        if (in)
            HIPCHK(c_, hipMemcpy2DAsync(d, pitch, p, pitch, w, (size_t)rows,
                                        hipMemcpyHostToDevice, c_->stream));
        if (out) arrs_.push_back(Arr{p, d, bytes, (size_t)rows, w, pitch});
        *dev = d;
        return OHC_OK;
    }
    int finish() {
        if (c_->opts.device_pointers) return OHC_OK;
        for (auto &a : arrs_) {
            if (a.rows > 1)
                HIPCHK(c_, hipMemcpy2DAsync(const_cast<void *>(a.host), a.pitch, a.dev, a.pitch,
                                            a.width, a.rows, hipMemcpyDeviceToHost, c_->stream));
            else
                HIPCHK(c_, hipMemcpyAsync(const_cast<void *>(a.host), a.dev, a.bytes,
                                          hipMemcpyDeviceToHost, c_->stream));
        }
        HIPCHK(c_, hipStreamSynchronize(c_->stream));
        return OHC_OK;
    }

   private:
    ohc_ctx *c_;
    size_t slot_ = 0;
    std::vector<Arr> arrs_;
};

int check_ctx(ohc_ctx *c);
// Memory-driven slicing of host-pointer calls: true (and the cells per slice
// in *m) when n cells of bytes_per_cell staged bytes each are to be done in
// slices -- opts.max_slice_cells, or as many cells as half the free device
// memory holds (turbulentSpeciesSolverCUDATran.cpp:196-222 splits into
// nSlice = used / available + 1).  Device-pointer calls are never sliced.
bool host_slices(ohc_ctx *c, int64_t n, size_t bytes_per_cell, int64_t *m);
template <class T>
T *off(T *p, int64_t s) {
    return p ? p + s : nullptr;
}
int read_stats(ohc_ctx *c, ohc_stats *st);
// Counters of one integrator launch (the 8 stats words the kernels write:
// cells, steps, attempts, Newton iterations, RHS evaluations, max cell steps,
// dt_min bits, failed cells) into st (nullable); failed cells -> error.
int stats_from_words(ohc_ctx *c, const unsigned long long *words, ohc_stats *st);
// Integrator parameters that come from the context options (pointers, n and
// the queue left zero).
void base_params(const ohc_ctx *c, ohcd::LaneParams &P);
// Launch the context's integrator kernel (lane or generic) on its stream.
hipError_t launch_integrator(const ohc_ctx *c, const ohcd::LaneParams &P);
// chemistrySource::solve (mode 0) / solveUpdate (mode 1) over n cells; with a
// device list of cell indices only those nlist cells are integrated.
int chem_integrate(ohc_ctx *c, int mode, int64_t n, const double *T, double *Tio,
                   const double *p, const double *rho, const double *Y, double *Yio,
                   int64_t ldY, const double *dt, double dtFactor, double *subDt, double *R,
                   int64_t ldR, ohc_stats *stats, const int *list, int64_t nlist);

}  // namespace ohci
