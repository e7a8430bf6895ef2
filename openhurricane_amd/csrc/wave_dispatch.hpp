// wave_dispatch.hpp -- host interface of the generic wavefront-per-cell kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "lane_kernels.hpp"

namespace ohcd {

struct WMech;
struct WaveParams;
struct KimParams;
struct LuParams;

// Padded ODE size the generic kernels use for n = ns + 1 unknowns (-1: too large).
int wave_size(int n);
hipError_t wave_integrate(const WaveParams &p, int grid, hipStream_t s);
hipError_t wave_eval(const EvalParams &p, const WMech &m, int exact, hipStream_t s);
hipError_t wave_source(const SourceParams &p, const WMech &m, hipStream_t s);
hipError_t wave_kim(const KimParams &p, const WMech &m, int fp32, hipStream_t s);
// batched dense LU solves of n x n systems, n <= 64
hipError_t wave_lu(const LuParams &p, hipStream_t s);
const void *wave_integrate_fn(int n);
// mixture thermo / transport and T from h for run-time tables
hipError_t wave_props(const PropParams &p, const WMech &m, hipStream_t s);
hipError_t wave_tfromh(const ThParams &p, const WMech &m, hipStream_t s);

}  // namespace ohcd
