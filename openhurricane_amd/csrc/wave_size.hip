// wave_size.hip -- the generic kernels for one padded ODE size OHC_WAVE_N
// (compiled once per size; see wave_dispatch.hip).
#include "wave_kernels.hpp"

#ifndef OHC_WAVE_N
#error "compile with -DOHC_WAVE_N=<size>"
#endif

#define OHC_CAT2(a, b) a##b
#define OHC_CAT(a, b) OHC_CAT2(a, b)

namespace ohcd {

hipError_t OHC_CAT(wave_integrate_, OHC_WAVE_N)(const WaveParams &p, int grid, hipStream_t s) {
    constexpr int N = OHC_WAVE_N;
    if (p.L.integrator == 4) {
        if (p.L.flow == 0) hipLaunchKernelGGL((mts_wave_kernel<N, 0>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((mts_wave_kernel<N, 1>), dim3(grid), dim3(64), 0, s, p);
    } else if (p.L.integrator == 3) {
        if (p.L.flow == 0) hipLaunchKernelGGL((ie_wave_kernel<N, 0, 3>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((ie_wave_kernel<N, 1, 3>), dim3(grid), dim3(64), 0, s, p);
    } else if (p.L.integrator == 2) {
        if (p.L.flow == 0) hipLaunchKernelGGL((ie_wave_kernel<N, 0, 2>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((ie_wave_kernel<N, 1, 2>), dim3(grid), dim3(64), 0, s, p);
    } else if (p.L.prec == 1) {
        if (p.L.flow == 0) hipLaunchKernelGGL((ie_wave_kernel<N, 0, 0, float>), dim3(grid), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((ie_wave_kernel<N, 1, 0, float>), dim3(grid), dim3(64), 0, s, p);
    } else if (p.L.flow == 0) {
        hipLaunchKernelGGL((ie_wave_kernel<N, 0>), dim3(grid), dim3(64), 0, s, p);
    } else {
        hipLaunchKernelGGL((ie_wave_kernel<N, 1>), dim3(grid), dim3(64), 0, s, p);
    }
    return hipGetLastError();
}

hipError_t OHC_CAT(wave_eval_, OHC_WAVE_N)(const EvalParams &p, const WMech &m, int exact,
                                            hipStream_t s) {
    constexpr int N = OHC_WAVE_N;
    const dim3 g((unsigned)p.n);
    if (p.prec == 1) {
        if (p.flow == 0) hipLaunchKernelGGL((eval_wave_kernel<N, 0, float>), g, dim3(64), 0, s, p, m, exact);
        else hipLaunchKernelGGL((eval_wave_kernel<N, 1, float>), g, dim3(64), 0, s, p, m, exact);
    } else if (p.flow == 0) {
        hipLaunchKernelGGL((eval_wave_kernel<N, 0>), g, dim3(64), 0, s, p, m, exact);
    } else {
        hipLaunchKernelGGL((eval_wave_kernel<N, 1>), g, dim3(64), 0, s, p, m, exact);
    }
    return hipGetLastError();
}

hipError_t OHC_CAT(wave_source_, OHC_WAVE_N)(const SourceParams &p, const WMech &m,
                                              hipStream_t s) {
    hipLaunchKernelGGL((source_wave_kernel<OHC_WAVE_N>), dim3((unsigned)p.n), dim3(64), 0, s, p, m);
    return hipGetLastError();
}

hipError_t OHC_CAT(wave_kim_, OHC_WAVE_N)(const KimParams &p, const WMech &m, int fp32,
                                           hipStream_t s) {
    const dim3 g((unsigned)p.n);
    if (fp32) hipLaunchKernelGGL((kim_wave_kernel<OHC_WAVE_N, float>), g, dim3(64), 0, s, p, m);
    else hipLaunchKernelGGL((kim_wave_kernel<OHC_WAVE_N, double>), g, dim3(64), 0, s, p, m);
    return hipGetLastError();
}

hipError_t OHC_CAT(wave_lu_, OHC_WAVE_N)(const LuParams &p, hipStream_t s) {
    hipLaunchKernelGGL((lu_wave_kernel<OHC_WAVE_N>), dim3((unsigned)p.nmat), dim3(64), 0, s, p);
    return hipGetLastError();
}

const void *OHC_CAT(wave_fn_, OHC_WAVE_N)() {
    return (const void *)&ie_wave_kernel<OHC_WAVE_N, 0>;
}

}  // namespace ohcd
