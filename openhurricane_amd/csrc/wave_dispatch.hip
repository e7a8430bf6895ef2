// wave_dispatch.hip -- instantiations of the generic (run-time table) kernels
// for padded ODE sizes N and their launch helpers.
#include "wave_dispatch.hpp"

#include "wave_kernels.hpp"

namespace ohcd {

namespace {
template <int N>
struct WaveN {
    static hipError_t integrate(const WaveParams &p, int grid, hipStream_t s) {
        if (p.L.flow == 0) hipLaunchKernelGGL((ie_wave_kernel<N, 0>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((ie_wave_kernel<N, 1>), dim3(grid), dim3(64), 0, s, p);
        return hipGetLastError();
    }
    static hipError_t eval(const EvalParams &p, const WMech &m, int exact, hipStream_t s) {
        const dim3 g((unsigned)p.n);
        if (p.flow == 0) hipLaunchKernelGGL((eval_wave_kernel<N, 0>), g, dim3(64), 0, s, p, m, exact);
        else hipLaunchKernelGGL((eval_wave_kernel<N, 1>), g, dim3(64), 0, s, p, m, exact);
        return hipGetLastError();
    }
    static const void *fn() { return (const void *)&ie_wave_kernel<N, 0>; }
};

struct Entry {
    int N;
    hipError_t (*integrate)(const WaveParams &, int, hipStream_t);
    hipError_t (*eval)(const EvalParams &, const WMech &, int, hipStream_t);
    const void *(*fn)();
};

#define OHC_WAVE(n) Entry{n, &WaveN<n>::integrate, &WaveN<n>::eval, &WaveN<n>::fn}
const Entry kWave[] = {OHC_WAVE(12), OHC_WAVE(16), OHC_WAVE(24), OHC_WAVE(33),
                       OHC_WAVE(40), OHC_WAVE(48), OHC_WAVE(55), OHC_WAVE(64)};
#undef OHC_WAVE

const Entry *pick(int n) {
    for (const Entry &e : kWave)
        if (e.N >= n) return &e;
    return nullptr;
}
}  // namespace

int wave_size(int n) {
    const Entry *e = pick(n);
    return e ? e->N : -1;
}

hipError_t wave_integrate(const WaveParams &p, int grid, hipStream_t s) {
    const Entry *e = pick(p.M.ns + 1);
    return e ? e->integrate(p, grid, s) : hipErrorInvalidValue;
}

hipError_t wave_eval(const EvalParams &p, const WMech &m, int exact, hipStream_t s) {
    const Entry *e = pick(m.ns + 1);
    return e ? e->eval(p, m, exact, s) : hipErrorInvalidValue;
}

const void *wave_integrate_fn(int n) {
    const Entry *e = pick(n);
    return e ? e->fn() : nullptr;
}

}  // namespace ohcd
