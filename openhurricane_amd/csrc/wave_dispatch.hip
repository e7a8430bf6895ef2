// wave_dispatch.hip -- dispatch of the generic (run-time table) kernels over
// the padded ODE sizes N.  Each size is instantiated in its own translation
// unit (wave_size.hip compiled with -DOHC_WAVE_N=N), so they build in parallel.
#include "wave_dispatch.hpp"

#include "wave_kernels.hpp"

namespace ohcd {

#define OHC_WAVE_DECL(n)                                                           \
    hipError_t wave_integrate_##n(const WaveParams &, int, hipStream_t);             \
    hipError_t wave_eval_##n(const EvalParams &, const WMech &, int, hipStream_t);  \
    hipError_t wave_source_##n(const SourceParams &, const WMech &, hipStream_t);    \
    const void *wave_fn_##n();
OHC_WAVE_DECL(12)
OHC_WAVE_DECL(16)
OHC_WAVE_DECL(24)
OHC_WAVE_DECL(33)
OHC_WAVE_DECL(40)
OHC_WAVE_DECL(48)
OHC_WAVE_DECL(55)
OHC_WAVE_DECL(64)
#undef OHC_WAVE_DECL

namespace {
struct Entry {
    int N;
    hipError_t (*integrate)(const WaveParams &, int, hipStream_t);
    hipError_t (*eval)(const EvalParams &, const WMech &, int, hipStream_t);
    hipError_t (*source)(const SourceParams &, const WMech &, hipStream_t);
    const void *(*fn)();
};

#define OHC_WAVE(n) Entry{n, &wave_integrate_##n, &wave_eval_##n, &wave_source_##n, &wave_fn_##n}
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

hipError_t wave_source(const SourceParams &p, const WMech &m, hipStream_t s) {
    const Entry *e = pick(m.ns + 1);
    return e ? e->source(p, m, s) : hipErrorInvalidValue;
}

const void *wave_integrate_fn(int n) {
    const Entry *e = pick(n);
    return e ? e->fn() : nullptr;
}

}  // namespace ohcd
