// lane_kernels.hpp -- persistent lane-per-cell kernels specialised on a
// compile-time mechanism M (see chem_device.hpp).
//
// ie_lane_kernel restates, per cell and in FP64, the reference's plain stiff
// integration path:
//   chemistrySource::solve(dt, dtFactor) / solveUpdate(dt, dtFactor)
//       src/speciesTransport/chemistrySource/src/chemistrySource.cpp:1494-1609, :1300-1325
//   ODESolver<chemistrySource>::solve
//       src/speciesTransport/chemistryODESolvers/include/ODESolver.inl:74-90
//   ImplicitEuler::{solve(t0,tn,dt0,y), adaptiveSolve, solve, NewtonIteration, maxIncrement}
//       src/Hurricane/mathematicalTools/ODEs/src/ImplicitEuler.cpp:40-342
//   ODEsSolver::maxError   src/Hurricane/mathematicalTools/ODEs/src/ODEsSolver.cpp:37-46
//   chemistrySource::checkNewY   chemistrySource.cpp:1152-1164
//
// Work distribution: a grid of single-wavefront blocks sized to the device
// stays resident; every lane owns one cell at a time and runs the integrator
// as a state machine whose only expensive transition -- "evaluate f and J at
// the Newton iterate, factor I - dt J, solve" -- is executed by all lanes of
// the wavefront together.  A lane whose cell finishes takes the next cell
// index from a global counter (one atomic per wavefront per refill), so cells
// of very different stiffness never leave lanes idle while work remains.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "chem_device.hpp"

namespace ohcd {

struct LaneParams {
    int64_t n;
    const double *T, *p, *rho, *Y, *dt;
    int64_t ldY;
    double dtFactor;
    double *subDt;
    double *R;
    int64_t ldR;
    double *Tout, *Yout;  // update mode: written in place (may alias T / Y)
    int mode;             // 0: chemistrySource::solve (R), 1: solveUpdate
    int flow;             // 0 constant pressure, 1 constant volume
    int max_steps;
    double atol, rtol, atol_inc, rtol_inc, safe, minsc, maxsc, alpha;
    double Tthr, Talpha, Tbeta, dt_max;
    long long max_cell_iters;     // safety cap on Newton iterations per cell (termination)
    const int *order;             // queue order (a permutation of 0..n-1) or null
    int *cost;                    // per-cell Newton iterations of this call, or null
    unsigned long long *counter;  // work-queue head, zeroed before launch
    unsigned long long *stats;    // see ohc_stats: [0..5] counters, [6] dt_min bits, [7] failures
    int exact_dkp;                // 1: exact d(Kp)/dT in the Jacobian
    int prec;                     // 1: FP32 rates / Gibbs / equilibrium (split precision)
    int integrator;               // 0: ImplicitEuler; 1: explicit Euler (reference GPU path);
                                  // 2: ImplicitTrapezoidal
};

enum : int { PH_NEED = 0, PH_NEWTON = 1, PH_ERR0 = 2, PH_DONE = 3, PH_F0 = 4 };

__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ long long wave_sum(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ long long wave_max(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        long long u = __shfl_xor(v, o);
        v = u > v ? u : v;
    }
    return v;
}

// METH 0: ImplicitEuler (ImplicitEuler.cpp:40-342).  METH 2: ImplicitTrapezoidal
// (ImplicitTrapezoidal.cpp:99-335): Newton on y - y0 - h/2 (f0 + f(y)) with
// I - h/2 J; every converged attempt evaluates f(y) for its error estimate
// (first step y - y0 - h f(y), later the second difference of f times
// h^3/12), and that f(y) becomes the next step's f0 when the attempt is
// accepted (the same state, so f0 is not evaluated again).
template <class M, int FLOW, class F = double, int METH = 0>
__global__ __launch_bounds__(64) void ie_lane_kernel(const LaneParams P) {
    constexpr int NS = M::NS, N = NS + 1;
    constexpr bool TR = METH == 2;
    const int lane = threadIdx.x;
    // Cold per-lane vectors live in LDS (one slot per lane, conflict-free):
    // the step-start state y0 and the previous step-start state y_{n-1}
    // (ImplicitTrapezoidal: s_yl holds f(yk) of the current attempt, s_f0
    // f(y0), s_fl f at the previous step start).
    __shared__ double s_y[N][64], s_yl[N][64], s_yk[N][64];
    __shared__ double s_f0[TR ? N : 1][64], s_fl[TR ? N : 1][64];
    // cold scalars of the state machine and the per-lane counters
    enum { C_DTC, C_LEFT, C_SUB, C_TN, C_TRY, C_LTS0, C_LTS1, C_PR, C_T, C_DT, C_DTL, C_LERR,
           C_NCTL };
    enum { S_CELLS, S_STEPS, S_ATT, S_NEWT, S_RHS, S_MAXS, S_FAIL, S_CSTEPS, S_CITER, S_CELL,
           S_NSTEP, S_NCOUNT, S_CHK, S_LAST, S_F0OK, S_FKOK, S_NST };
    __shared__ double s_c[C_NCTL][64];
    __shared__ long long s_n[S_NST][64];
    __shared__ double s_dtmin[64];
    // integrator parameters, uniform across the block (broadcast LDS reads keep
    // them out of the scalar register file)
    enum { Q_ATOL, Q_RTOL, Q_ATOLI, Q_RTOLI, Q_SAFE, Q_MINSC, Q_MAXSC, Q_ALPHA, Q_TTHR, Q_TALPHA,
           Q_TBETA, Q_DTMAX, Q_DTF, Q_ERRG, Q_NPAR };
    __shared__ double s_par[Q_NPAR];
    if (lane == 0) {
        s_par[Q_ATOL] = P.atol;
        s_par[Q_RTOL] = P.rtol;
        s_par[Q_ATOLI] = P.atol_inc;
        s_par[Q_RTOLI] = P.rtol_inc;
        s_par[Q_SAFE] = P.safe;
        s_par[Q_MINSC] = P.minsc;
        s_par[Q_MAXSC] = P.maxsc;
        s_par[Q_ALPHA] = P.alpha;
        s_par[Q_TTHR] = P.Tthr;
        s_par[Q_TALPHA] = P.Talpha;
        s_par[Q_TBETA] = P.Tbeta;
        s_par[Q_DTMAX] = P.dt_max;
        s_par[Q_DTF] = P.dtFactor;
        s_par[Q_ERRG] = pow(P.maxsc / P.safe, -1.0 / P.alpha);
    }
    __syncthreads();
#define PAR(i) s_par[i]
#define Y0(i) s_y[i][lane]
#define YL(i) s_yl[i][lane]
#define YK(i) s_yk[i][lane]
#define CT(i) s_c[i][lane]
#define CNT(i) s_n[i][lane]
#define F0(i) s_f0[i][lane]
#define FL(i) s_fl[i][lane]

    // The Newton iterate also lives in LDS so that only f and J occupy
    // registers through the evaluation and the factorisation.
    double f[N], J[N][N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        YK(i) = 0.0;
        Y0(i) = 0.0;
        YL(i) = 0.0;
    }
    YK(NS) = 1000.0;
    Y0(NS) = 1000.0;

    // All loop-carried lane state except the phase lives in LDS: the loop body
    // then holds only f, J and transients in registers.
    int phase = PH_NEED;
#pragma unroll
    for (int i = 0; i < C_NCTL; ++i) CT(i) = 0.0;
#pragma unroll
    for (int i = 0; i < S_NST; ++i) CNT(i) = 0;
    s_dtmin[lane] = __builtin_huge_val();
    long long &cell = CNT(S_CELL);
    long long &nstep = CNT(S_NSTEP), &ncount = CNT(S_NCOUNT), &chk = CNT(S_CHK),
              &last = CNT(S_LAST);
    double &pr = CT(C_PR), &t = CT(C_T), &dt = CT(C_DT), &dtl = CT(C_DTL),
           &lastErr = CT(C_LERR);

    // ---- state machine pieces (lane-divergent, cheap) ----
    auto finish_cell = [&]() __attribute__((always_inline)) {
        const double subDt = CT(C_SUB);
        P.subDt[cell] = subDt;
        if (P.mode == 0) {
            const double rho = P.rho[cell], rdt = 1.0 / CT(C_DTC);
            sfor<0, NS>([&](auto k) {
                const double c0 = rho * P.Y[k * P.ldY + cell] / M::W[k];
                P.R[k * P.ldR + cell] = (Y0(k) - c0) * rdt * M::W[k];
            });
        } else {
            // chemistrySource::ciToYi(ci, yi) (chemistrySource.cpp:1222-1240)
            double yy[NS], rs = 0.0, ys = 0.0;
            sfor<0, NS>([&](auto k) {
                yy[k] = fmax(0.0, Y0(k));
                rs += yy[k] * M::W[k];
            });
            const double rr = 1.0 / fmax(rs, kTiny);
            sfor<0, NS>([&](auto k) {
                yy[k] = yy[k] * M::W[k] * rr;
                ys += yy[k];
            });
            const double ri = 1.0 / ys;
            sfor<0, NS>([&](auto k) { P.Yout[k * P.ldY + cell] = ys != 1.0 ? yy[k] * ri : yy[k]; });
            P.Tout[cell] = Y0(NS);
        }
        CNT(S_CELLS) += 1;
        if (P.cost) P.cost[cell] = (int)CNT(S_CITER);
        s_dtmin[lane] = fmin(s_dtmin[lane], subDt);
        CNT(S_MAXS) = CNT(S_CSTEPS) > CNT(S_MAXS) ? CNT(S_CSTEPS) : CNT(S_MAXS);
        phase = PH_NEED;
    };
    auto start_attempt = [&]() __attribute__((always_inline)) {
        CT(C_LTS0) = dtl;
        CNT(S_ATT) += 1;
#pragma unroll
        for (int i = 0; i < N; ++i) YK(i) = Y0(i);
        ncount = 0;
        lastErr = 0.0;
        phase = PH_NEWTON;
        if constexpr (TR) {
            CNT(S_FKOK) = 0;
            if (!CNT(S_F0OK)) phase = PH_F0;  // f(y0) first (the iterate is y0)
        }
    };
    // ImplicitEuler::solve(t0, tn, dt0, y) loop head for step nstep; false when
    // the step budget is exhausted (tn then becomes the time reached).
    auto start_step = [&]() __attribute__((always_inline)) -> bool {
        const double tn = CT(C_TN);
        if (nstep >= P.max_steps) {
            CT(C_TN) = t;
            return false;
        }
        CT(C_TRY) = dt;
        if ((t + dt - tn) * (t + dt) > 0) {
            dt = tn - t;
            last = true;
        } else if (t + dt == tn) {
            last = true;
        }
        dtl = dt;  // adaptiveSolve entry
        CT(C_LTS1) = CT(C_LTS0);
        chk = 0;
        start_attempt();
        return true;
    };
    auto start_ode = [&]() __attribute__((always_inline)) {
        CT(C_TN) = CT(C_LEFT);
        t = 0.0;
        dt = 0.2 * CT(C_SUB);
        last = false;
        nstep = 0;
        return start_step();
    };
    // ImplicitEuler::maxIncrement of the error vector e (held in f)
    auto max_increment = [&](const double(&e)[N]) __attribute__((always_inline)) {
        double m = 0.0;
        bool nan_y = false;
        sfor<0, N>([&](auto i) {
            const double yki = YK(i);
            nan_y |= isnan(yki);
            const double tol = PAR(Q_ATOLI) + PAR(Q_RTOLI) * fmax(fabs(Y0(i)), fabs(yki));
            m = fmax(m, fabs(e[i]) * frcp(tol));
        });
        return nan_y ? 1.5 : m;
    };
    // adaptiveSolve after one ImplicitEuler::solve attempt returned `error`,
    // followed by the ImplicitEuler::solve and chemistrySource::solve loops.
    enum { A_RETRY, A_FAIL, A_NEXT, A_ODE_END };
    auto attempt_done = [&](double error) __attribute__((always_inline)) {
        int action;
        const double y0T = Y0(NS);
        const bool ok = !((YK(NS) - y0T) > PAR(Q_TALPHA) * y0T);
        if (!ok) dtl *= PAR(Q_TBETA);
        if (!ok && chk < 100) {
            chk++;
            action = dtl <= kVeryTiny ? A_FAIL : A_RETRY;
        } else if (error > 1.0) {
            dtl *= fmax(PAR(Q_SAFE) * pow(error, -PAR(Q_ALPHA)), PAR(Q_MINSC));
            action = dtl <= kVeryTiny ? A_FAIL : A_RETRY;
        } else {
            t += dtl;
            if constexpr (TR) {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    FL(i) = F0(i);
                    F0(i) = YL(i);  // f(yk) of the accepted attempt = f at the new y0
                    Y0(i) = YK(i);
                }
                CNT(S_F0OK) = CNT(S_FKOK);
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    YL(i) = Y0(i);
                    Y0(i) = YK(i);
                }
            }
            if (error > PAR(Q_ERRG))
                dt = fmin(fmax(PAR(Q_SAFE) * pow(error, -PAR(Q_ALPHA)), PAR(Q_MINSC)), PAR(Q_MAXSC)) * dtl;
            else
                dt = fmax(PAR(Q_SAFE) * PAR(Q_MAXSC) * dtl, dtl);
            CNT(S_STEPS) += 1;
            CNT(S_CSTEPS) += 1;
            const double tn = CT(C_TN);
            if ((t - tn) * tn >= 0) {
                if (nstep > 0 && last) CT(C_SUB) = fmin(CT(C_TRY), tn);
                action = A_ODE_END;
            } else {
                nstep++;
                action = A_NEXT;
            }
        }
        if (action == A_RETRY) {
            start_attempt();
        } else if (action == A_NEXT) {
            if (!start_step()) action = A_ODE_END;
        }
        if (action == A_ODE_END) {
            // chemistrySource::solve: while (leftTime > 0) { dt = leftTime; solve(c, dt, subDt); leftTime -= dt; }
            const double left = CT(C_LEFT) - CT(C_TN);
            CT(C_LEFT) = left;
            if (left > 0.0) start_ode();
            else finish_cell();
        } else if (action == A_FAIL) {
            CNT(S_FAIL) += 1;
            finish_cell();
        }
    };
    // start a fresh cell (chemistrySource::solve / solveUpdate loop body)
    auto start_cell = [&]() __attribute__((always_inline)) {
        const double T0 = P.T[cell];
        const double dtc = P.dt[cell] * PAR(Q_DTF);
        CT(C_DTC) = dtc;
        const double dtI = P.subDt[cell] > 0.0 ? P.subDt[cell] : dtc;
        if (T0 < PAR(Q_TTHR)) {
            P.subDt[cell] = dtc;
            if (P.cost) P.cost[cell] = 0;
            if (P.mode == 0) sfor<0, NS>([&](auto k) { P.R[k * P.ldR + cell] = 0.0; });
            return;  // stays PH_NEED
        }
        const double rho = P.rho[cell];
        pr = P.flow == 0 ? P.p[cell] : rho;
        sfor<0, NS>([&](auto k) { Y0(k) = rho * P.Y[k * P.ldY + cell] / M::W[k]; });
        Y0(NS) = T0;
        CT(C_SUB) = fmin(dtc, fmin(dtI, PAR(Q_DTMAX)));
        CT(C_LEFT) = dtc;
        CNT(S_CSTEPS) = 0;
        CNT(S_CITER) = 0;
        CNT(S_F0OK) = 0;
        start_ode();
    };

    for (;;) {
        // ---- refill lanes that need a cell (one atomic per wavefront) ----
        for (;;) {
            const unsigned long long need = __ballot(phase == PH_NEED);
            if (need == 0ull) break;
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.counter, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            if (phase == PH_NEED) {
                const unsigned long long below = lane ? (need & ((1ull << lane) - 1ull)) : 0ull;
                const unsigned long long mine = base + (unsigned long long)__popcll(below);
                if (mine >= (unsigned long long)P.n) {
                    phase = PH_DONE;
                } else {
                    cell = P.order ? (long long)P.order[mine] : (long long)mine;
                    start_cell();
                }
            }
        }
        if (!__any(phase == PH_NEWTON || phase == PH_ERR0 || phase == PH_F0)) break;

        // ---- the wavefront-wide heavy step: f and J at the Newton iterate ----
        // The fences keep the compiler from hoisting control-state loads above
        // (or sinking them below) the evaluation and factorisation, so only f,
        // J and their transients are live there.
        asm volatile("" ::: "memory");
        __builtin_amdgcn_sched_barrier(0);
        {
            double yk[N];
#pragma unroll
            for (int i = 0; i < N; ++i) yk[i] = YK(i);
            eval_fJ<M, true, FLOW, F, true>(yk, pr, f, J, P.exact_dkp != 0);
        }

        bool have = false;
        double aerr = 0.0;
        if constexpr (TR) {
            if (phase == PH_F0) {
                // f at the step start (adaptiveSolve's DyDt(t, y)), then Newton
                CNT(S_RHS) += 1;
                sfor<0, N>([&](auto i) { F0(i) = f[i]; });
                CNT(S_F0OK) = 1;
                phase = PH_NEWTON;
            } else if (phase == PH_ERR0) {
                // ImplicitTrapezoidal::solve error (ImplicitTrapezoidal.cpp:99-142)
                CNT(S_RHS) += 1;
                sfor<0, N>([&](auto i) { YL(i) = f[i]; });
                CNT(S_FKOK) = 1;
                if (nstep == 0) {
                    sfor<0, N>([&](auto i) { f[i] = YK(i) - Y0(i) - dtl * f[i]; });
                } else {
                    const double dtn = CT(C_LTS0), dtn1 = CT(C_LTS1);
                    const double c1 = 2.0 / (dtn * (dtn + dtn1));
                    const double c2 = 2.0 / (dtn1 * (dtn + dtn1));
                    const double LTEF = dtn * dtn * dtn / 12.0;
                    sfor<0, N>([&](auto i) {
                        const double f0 = F0(i);
                        f[i] = (c1 * (f[i] - f0) - c2 * (f0 - FL(i))) * LTEF;
                    });
                }
                have = true;
                aerr = max_increment(f);
            }
        } else if (phase == PH_ERR0) {
            // first step: e = y - y0 - dt f(y) (ImplicitEuler.cpp:111-120)
            CNT(S_RHS) += 1;
            sfor<0, N>([&](auto i) { f[i] = YK(i) - Y0(i) - dtl * f[i]; });
            have = true;
            aerr = max_increment(f);
        }
        if (__any(phase == PH_NEWTON)) {
            // (I - dt J) delta = -yk + y0 + dt f  (ImplicitEuler.cpp:186-198)
            const double h = TR ? 0.5 * dtl : dtl;
            sfor<0, N>([&](auto i) {
                sfor<0, N>([&](auto j) { J[i][j] = -h * J[i][j]; });
                J[i][i] += 1.0;
                if constexpr (TR) f[i] = -YK(i) + Y0(i) + h * (F0(i) + f[i]);
                else f[i] = -YK(i) + Y0(i) + h * f[i];
            });
            lu_solve<N>(J, f);
        }
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::: "memory");

        // ---- lane-local transitions ----
        if (phase == PH_NEWTON) {
            // NewtonIteration body after the solve (ImplicitEuler.cpp:200-249); delta in f
            CNT(S_NEWT) += 1;
            CNT(S_CITER) += 1;
            bool stop = false;
            sfor<0, N>([&](auto i) {
                if (!stop) {
                    if (bad(f[i])) stop = true;
                    else YK(i) += f[i];
                }
            });
            if (stop) {
                have = true;
                aerr = 1000.0;
            } else {
                double maxErr = 0.0;
                sfor<0, N>([&](auto i) {
                    const double tol = PAR(Q_ATOL) + PAR(Q_RTOL) * fabs(YK(i));
                    maxErr = fmax(maxErr, fabs(f[i]) * frcp(tol));
                });
                const bool diverged =
                    ncount > 0 && (maxErr / fmax(lastErr, kTiny)) > 2 && ncount > 5;
                if (diverged) {
                    have = true;
                    aerr = 1000.0;
                } else {
                    lastErr = maxErr;
                    ncount++;
                    if (isnan(maxErr)) {
                        have = true;
                        aerr = 1000.0;
                    } else if (ncount > 1000) {
                        have = true;
                        aerr = maxErr;
                    } else if (maxErr > 1.0) {
                        // iterate again at the new yk
                    } else if (TR || nstep == 0) {
                        phase = PH_ERR0;
                    } else {
                        // second-difference LTE (ImplicitEuler.cpp:122-139)
                        const double dtn = CT(C_LTS0), dtn1 = CT(C_LTS1);
                        const double f1 = 2.0 / (dtn * (dtn + dtn1));
                        const double f2 = 2.0 / (dtn1 * (dtn + dtn1));
                        const double LTEF = dtn * dtn * 0.5;
                        sfor<0, N>([&](auto i) {
                            const double y0 = Y0(i);
                            f[i] = (f1 * (YK(i) - y0) - f2 * (y0 - YL(i))) * LTEF;
                        });
                        have = true;
                        aerr = max_increment(f);
                    }
                }
            }
        }
        if (have) {
            attempt_done(aerr);
        } else if (phase == PH_NEWTON && CNT(S_CITER) > P.max_cell_iters) {
            // no legitimate cell gets here; the cap guarantees every wavefront drains
            CNT(S_FAIL) += 1;
            finish_cell();
        }
    }
    // ---- counters: one atomic per wavefront ----
    const long long s0 = wave_sum(CNT(S_CELLS)), s1 = wave_sum(CNT(S_STEPS)),
                    s2 = wave_sum(CNT(S_ATT)), s3 = wave_sum(CNT(S_NEWT)),
                    s4 = wave_sum(CNT(S_RHS)), s5 = wave_max(CNT(S_MAXS)),
                    s7 = wave_sum(CNT(S_FAIL));
    const double dmin = wave_min(s_dtmin[lane]);
#undef Y0
#undef YL
#undef YK
#undef CT
#undef CNT
#undef PAR
#undef F0
#undef FL
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)s0);
        atomicAdd(&P.stats[1], (unsigned long long)s1);
        atomicAdd(&P.stats[2], (unsigned long long)s2);
        atomicAdd(&P.stats[3], (unsigned long long)s3);
        atomicAdd(&P.stats[4], (unsigned long long)s4);
        atomicMax(&P.stats[5], (unsigned long long)s5);
        if (dmin < __builtin_huge_val())
            atomicMin(&P.stats[6], (unsigned long long)__double_as_longlong(dmin));
        atomicAdd(&P.stats[7], (unsigned long long)s7);
    }
}

// ---- the reference GPU integrator (comparison mode) ----
// Explicit Euler with a step-doubling error estimate, per cell as
// solveChemODEs (CS/src/CUDAChemistrySourceODEs.cu:588-640) drives
// EulerCUDA::solve / adaptiveSolve (CS/src/EulerCUDA.cu:34-62, :130-224) with
// ODEsSolverCUDA::maxError (CS/src/ODEsSolverCUDA.cu:34-48): an attempt of
// size h from y0 takes two half steps (y_h) and one full step (y_f) with
// f0 = f(y0); error = max |2 (y_f - y_h)| / (atol + rtol max(|y_f|, |y_h|)).
// Every cell is integrated (no temperature threshold) starting from subDt.
// Lane-per-cell state machine: each loop trip evaluates f once per lane, at
// y0 (start of a step) or at the half-step point of the current attempt.
constexpr int kEEMaxTinyRejects = 1000;  // rejected attempts at h = 1e-15 before a cell fails

template <class M, int FLOW>
__global__ __launch_bounds__(64) void ee_lane_kernel(const LaneParams P) {
    constexpr int NS = M::NS, N = NS + 1;
    enum { E_NEED, E_F0, E_MID, E_DONE };
    const int lane = threadIdx.x;
    // Every loop-carried lane value except the phase lives in LDS (one slot
    // per lane): the loop body then holds only y, f and transients in
    // registers.
    __shared__ double s_y0[N][64], s_f0[N][64];
    enum { C_PR, C_T, C_DT, C_DTL, C_TRY, C_TN, C_LEFT, C_SUB, C_DTC, C_DTMIN, C_NCTL };
    enum { S_CELLS, S_STEPS, S_ATT, S_RHS, S_MAXS, S_FAIL, S_CELL, S_CSTEPS, S_NSTEP, S_LAST,
           S_REJ, S_NST };
    __shared__ double s_c[C_NCTL][64];
    __shared__ long long s_n[S_NST][64];
    enum { Q_ATOL, Q_RTOL, Q_DTF, Q_ERRG, Q_NPAR };
    __shared__ double s_par[Q_NPAR];
    if (lane == 0) {
        s_par[Q_ATOL] = P.atol;
        s_par[Q_RTOL] = P.rtol;
        s_par[Q_DTF] = P.dtFactor;
        s_par[Q_ERRG] = pow(10.0 / 0.8, -1.0 / 0.25);
    }
#define CT(i) s_c[i][lane]
#define CNT(i) s_n[i][lane]
#pragma unroll
    for (int i = 0; i < C_NCTL; ++i) CT(i) = 0.0;
#pragma unroll
    for (int i = 0; i < S_NST; ++i) CNT(i) = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        s_y0[i][lane] = 0.0;
        s_f0[i][lane] = 0.0;
    }
    s_y0[NS][lane] = 1000.0;
    CT(C_DTMIN) = __builtin_huge_val();
    __syncthreads();
    int phase = E_NEED;

    auto finish_cell = [&](bool failed) __attribute__((always_inline)) {
        const long long cell = CNT(S_CELL);
        P.subDt[cell] = CT(C_SUB);
        if (P.mode == 0) {
            const double rho = P.rho[cell], rdt = 1.0 / CT(C_DTC);
            sfor<0, NS>([&](auto k) {
                const double c0 = rho * P.Y[k * P.ldY + cell] / M::W[k];
                P.R[k * P.ldR + cell] = (s_y0[k][lane] - c0) * rdt * M::W[k];
            });
        } else {
            // device ciToYi (CUDAChemistrySourceODEs.cu:545-586): Y = c W / sum c W
            double r = 0.0;
            sfor<0, NS>([&](auto k) { r += s_y0[k][lane] * M::W[k]; });
            const double rr = 1.0 / r;
            sfor<0, NS>([&](auto k) { P.Yout[k * P.ldY + cell] = s_y0[k][lane] * M::W[k] * rr; });
            P.Tout[cell] = s_y0[NS][lane];
        }
        const long long cs = CNT(S_CSTEPS);
        if (P.cost) P.cost[cell] = (int)(cs < 0x7fffffffll ? cs : 0x7fffffffll);
        CNT(S_CELLS) += 1;
        if (failed) CNT(S_FAIL) += 1;
        if (cs > CNT(S_MAXS)) CNT(S_MAXS) = cs;
        CT(C_DTMIN) = fmin(CT(C_DTMIN), CT(C_SUB));
        phase = E_NEED;
    };
    // EulerCUDA::solve loop head; false when the step budget is exhausted
    auto start_step = [&]() __attribute__((always_inline)) -> bool {
        const double t = CT(C_T), tn = CT(C_TN);
        double dt = CT(C_DT);
        if (CNT(S_NSTEP) >= P.max_steps) {
            CT(C_TN) = t;
            return false;
        }
        CT(C_TRY) = dt;
        if ((t + dt - tn) * (t + dt) > 0) {
            dt = tn - t;
            CNT(S_LAST) = 1;
        } else if (t + dt == tn) {
            CNT(S_LAST) = 1;
        }
        CT(C_DT) = dt;
        CT(C_DTL) = dt;
        CNT(S_REJ) = 0;
        phase = E_F0;
        return true;
    };
    // end of one EulerCUDA::solve call over [0, tn]: the cell either has time
    // left (the budget ran out first) or is finished
    auto ode_end = [&]() __attribute__((always_inline)) {
        const double left = CT(C_LEFT) - CT(C_TN);
        CT(C_LEFT) = left;
        if (left > 0.0) {
            CT(C_TN) = left;
            CT(C_T) = 0.0;
            CT(C_DT) = CT(C_SUB);
            CNT(S_LAST) = 0;
            CNT(S_NSTEP) = 0;
            // no step possible (a zero budget): give up on the cell
            if (!start_step()) finish_cell(true);
        } else {
            finish_cell(false);
        }
    };

    for (;;) {
        for (;;) {
            const unsigned long long need = __ballot(phase == E_NEED);
            if (need == 0ull) break;
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.counter, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            if (phase == E_NEED) {
                const unsigned long long below = lane ? (need & ((1ull << lane) - 1ull)) : 0ull;
                const unsigned long long mine = base + (unsigned long long)__popcll(below);
                if (mine >= (unsigned long long)P.n) {
                    phase = E_DONE;
                } else {
                    const long long cell = P.order ? (long long)P.order[mine] : (long long)mine;
                    CNT(S_CELL) = cell;
                    const double rho = P.rho[cell];
                    CT(C_PR) = FLOW == 0 ? P.p[cell] : rho;
                    const double dtc = P.dt[cell] * s_par[Q_DTF];
                    const double s0 = P.subDt[cell];
                    const double sub = s0 > 0.0 ? s0 : dtc;
                    CT(C_DTC) = dtc;
                    CT(C_SUB) = sub;
                    sfor<0, NS>([&](auto k) { s_y0[k][lane] = rho * P.Y[k * P.ldY + cell] / M::W[k]; });
                    s_y0[NS][lane] = P.T[cell];
                    CNT(S_CSTEPS) = 0;
                    CT(C_LEFT) = dtc;
                    CT(C_TN) = dtc;
                    CT(C_T) = 0.0;
                    CT(C_DT) = sub;
                    CNT(S_LAST) = 0;
                    CNT(S_NSTEP) = 0;
                    if (!start_step()) finish_cell(true);
                }
            }
        }
        if (!__any(phase == E_F0 || phase == E_MID)) break;
        // one f evaluation per lane: at y0 (E_F0) or at the half-step point
        double y[N], f[N], Jd[N][N];
        {
            const double hh = phase == E_MID ? 0.5 * CT(C_DTL) : 0.0;
            sfor<0, N>([&](auto i) { y[i] = s_y0[i][lane] + hh * s_f0[i][lane]; });
        }
        eval_fJ<M, false, FLOW>(y, CT(C_PR), f, Jd);
        if (phase == E_F0) {
            CNT(S_RHS) += 1;
            sfor<0, N>([&](auto i) { s_f0[i][lane] = f[i]; });
            phase = E_MID;
        } else if (phase == E_MID) {
            CNT(S_RHS) += 1;
            CNT(S_ATT) += 1;
            const double dtl = CT(C_DTL), atol = s_par[Q_ATOL], rtol = s_par[Q_RTOL];
            double e = 0.0;
            sfor<0, N>([&](auto i) {
                const double y0 = s_y0[i][lane], f0 = s_f0[i][lane];
                const double yh = y[i] + 0.5 * dtl * f[i];
                const double yf = y0 + dtl * f0;
                const double err = 2.0 / (1.0 - 2.0) * (yh - yf);
                const double tol = atol + rtol * fmax(fabs(yf), fabs(yh));
                e = fmax(e, fabs(err) / tol);
                f[i] = yh;  // keep the two-half-step state
            });
            if (e > 1.0) {
                double d = dtl * fmax(0.8 * pow(e, -0.25), 0.25);
                if (d <= kTiny) {
                    d = kTiny;
                    CNT(S_REJ) += 1;
                }
                CT(C_DTL) = d;
                // the attempt at y0 is repeated with the smaller step (f0 kept)
                if (CNT(S_REJ) > kEEMaxTinyRejects) finish_cell(true);
            } else {
                const double t = CT(C_T) + dtl, tn = CT(C_TN);
                CT(C_T) = t;
                CT(C_DT) = e > s_par[Q_ERRG] ? fmin(fmax(0.8 * pow(e, -0.25), 0.25), 10.0) * dtl
                                              : 0.8 * 10.0 * dtl;
                sfor<0, N>([&](auto i) { s_y0[i][lane] = f[i]; });
                CNT(S_STEPS) += 1;
                const long long cs = CNT(S_CSTEPS) + 1;
                CNT(S_CSTEPS) = cs;
                if (cs > P.max_cell_iters) {
                    finish_cell(true);  // termination guard for every wavefront
                } else if ((t - tn) * tn >= 0) {
                    if (CNT(S_NSTEP) > 0 && CNT(S_LAST)) CT(C_SUB) = fmin(CT(C_TRY), tn);
                    ode_end();
                } else {
                    CNT(S_NSTEP) += 1;
                    if (!start_step()) ode_end();
                }
            }
        }
    }
    const long long s0 = wave_sum(CNT(S_CELLS)), s1 = wave_sum(CNT(S_STEPS)),
                    s2 = wave_sum(CNT(S_ATT)), s4 = wave_sum(CNT(S_RHS)),
                    s5 = wave_max(CNT(S_MAXS)), s7 = wave_sum(CNT(S_FAIL));
    const double dmin = wave_min(CT(C_DTMIN));
#undef CT
#undef CNT
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)s0);
        atomicAdd(&P.stats[1], (unsigned long long)s1);
        atomicAdd(&P.stats[2], (unsigned long long)s2);
        atomicAdd(&P.stats[4], (unsigned long long)s4);
        atomicMax(&P.stats[5], (unsigned long long)s5);
        if (dmin < __builtin_huge_val())
            atomicMin(&P.stats[6], (unsigned long long)__double_as_longlong(dmin));
        atomicAdd(&P.stats[7], (unsigned long long)s7);
    }
}

// MTS_whichGroup: 1 + decades of tn / tci (1 when tci >= tn)
__device__ __forceinline__ int mts_group(double tn, double tci) {
    if (!(tn > tci)) return 1;
    const int g = (int)log10(tn / tci);
    return (g > 0 ? g : 0) + 1;
}

// ---- MTS explicit Euler (the reference GPU path's MTSEuler solver type:
// solveChemODEsMTS, CUDAChemistrySourceODEs.cu:684-778, timescale /
// MTS_whichGroup / MTS_GroupDT :199-309): the algorithm of mts_wave_kernel
// with one cell per lane and one RHS evaluation per lane per loop trip.  The
// per-cell state lives in LDS (one slot per lane), like ee_lane_kernel's.
template <class M, int FLOW>
__global__ __launch_bounds__(64) void mts_lane_kernel(const LaneParams P) {
    constexpr int NS = M::NS, N = NS + 1;
    enum { T_NEED, T_STEP, T_DONE };
    const int lane = threadIdx.x;
    __shared__ double s_y[N][64], s_tci[NS][64];
    __shared__ int s_ig[NS][64];
    enum { C_PR, C_TN, C_DID, C_DTG, C_DTMIN, C_NCTL };
    enum { S_CELLS, S_STEPS, S_MAXS, S_FAIL, S_CELL, S_NST, S_IM, S_NM, S_NCNT };
    __shared__ double s_c[C_NCTL][64];
    __shared__ long long s_n[S_NCNT][64];
#define CT(i) s_c[i][lane]
#define CNT(i) s_n[i][lane]
#pragma unroll
    for (int i = 0; i < C_NCTL; ++i) CT(i) = 0.0;
#pragma unroll
    for (int i = 0; i < S_NCNT; ++i) CNT(i) = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) s_y[i][lane] = 0.0;
    s_y[NS][lane] = 1000.0;
    CT(C_DTMIN) = __builtin_huge_val();
    __syncthreads();
    int phase = T_NEED;
    auto finish_cell = [&](bool failed) __attribute__((always_inline)) {
        const long long cell = CNT(S_CELL);
        const double tn = CT(C_TN);
        if (P.mode == 0) {
            const double rho = P.rho[cell], rdt = 1.0 / tn;
            sfor<0, NS>([&](auto k) {
                const double c0 = rho * P.Y[k * P.ldY + cell] / M::W[k];
                P.R[k * P.ldR + cell] = (s_y[k][lane] - c0) * rdt * M::W[k];
            });
        } else {
            // device ciToYi (CUDAChemistrySourceODEs.cu:545-586): Y = c W / sum c W
            double r = 0.0;
            sfor<0, NS>([&](auto k) { r += s_y[k][lane] * M::W[k]; });
            const double rr = 1.0 / r;
            sfor<0, NS>([&](auto k) { P.Yout[k * P.ldY + cell] = s_y[k][lane] * M::W[k] * rr; });
            P.Tout[cell] = s_y[NS][lane];
        }
        const long long cs = CNT(S_NST);
        if (P.cost) P.cost[cell] = (int)(cs < 0x7fffffffll ? cs : 0x7fffffffll);
        CNT(S_CELLS) += 1;
        CNT(S_STEPS) += cs;
        if (failed) CNT(S_FAIL) += 1;
        if (cs > CNT(S_MAXS)) CNT(S_MAXS) = cs;
        CT(C_DTMIN) = fmin(CT(C_DTMIN), P.subDt[cell]);
        phase = T_NEED;
    };
    // the next group with species after im (Nm when none), and its step
    auto next_group = [&](long long im) __attribute__((always_inline)) {
        long long nx = CNT(S_NM);
        sfor<0, NS>([&](auto k) {
            const long long g = s_ig[k][lane];
            if (g > im && g < nx) nx = g;
        });
        double dtG = __builtin_huge_val();
        sfor<0, NS>([&](auto k) {
            if (s_ig[k][lane] == nx) dtG = fmin(dtG, s_tci[k][lane]);
        });
        CNT(S_IM) = nx;
        CT(C_DTG) = dtG;
        return nx < CNT(S_NM);
    };
    for (;;) {
        for (;;) {
            const unsigned long long need = __ballot(phase == T_NEED);
            if (need == 0ull) break;
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(P.counter, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            if (phase == T_NEED) {
                const unsigned long long below = lane ? (need & ((1ull << lane) - 1ull)) : 0ull;
                const unsigned long long mine = base + (unsigned long long)__popcll(below);
                if (mine >= (unsigned long long)P.n) {
                    phase = T_DONE;
                } else {
                    const long long cell = P.order ? (long long)P.order[mine] : (long long)mine;
                    CNT(S_CELL) = cell;
                    const double rho = P.rho[cell];
                    const double pr = FLOW == 0 ? P.p[cell] : rho;
                    const double tn = P.dt[cell] * P.dtFactor;
                    CT(C_PR) = pr;
                    CT(C_TN) = tn;
                    CT(C_DID) = 0.0;
                    CNT(S_NST) = 0;
                    double y[N], f[N], J[N][N];
                    sfor<0, NS>([&](auto k) { y[k] = rho * P.Y[k * P.ldY + cell] / M::W[k]; });
                    y[NS] = P.T[cell];
                    sfor<0, N>([&](auto i) { s_y[i][lane] = y[i]; });
                    // timescale / MTS_whichGroup / MTS_GroupDT
                    eval_fJ<M, true, FLOW, double, false, true>(y, pr, f, J);
                    double tmin = __builtin_huge_val();
                    sfor<0, NS>([&](auto k) {
                        const double t = fmin(200.0, 1.0 / (fabs(J[k][k]) + 1e-20));
                        s_tci[k][lane] = t;
                        tmin = fmin(tmin, t);
                    });
                    const int Nm = tn > tmin ? (int)fmax(0.0, floor(log10(tn / tmin))) + 1 : 1;
                    CNT(S_NM) = Nm;
                    sfor<0, NS>([&](auto k) { s_ig[k][lane] = Nm - mts_group(tn, s_tci[k][lane]); });
                    if (next_group(-1) && tn > 0.0) phase = T_STEP;
                    else finish_cell(false);
                }
            }
        }
        if (!__any(phase == T_STEP)) break;
        double y[N], f[N], Jd[N][N];
        sfor<0, N>([&](auto i) { y[i] = s_y[i][lane]; });
        eval_fJ<M, false, FLOW>(y, CT(C_PR), f, Jd);
        if (phase == T_STEP) {
            const long long im = CNT(S_IM), Nm = CNT(S_NM);
            if (CNT(S_NST) >= P.max_steps) {
                finish_cell(true);
            } else {
                const double tn = CT(C_TN), did = CT(C_DID);
                double dt = 0.8 * CT(C_DTG);
                if (did + dt > tn) dt = tn - did;
                const double atol = P.atol, rtol = P.rtol;
                double e = 0.0;
                sfor<0, N>([&](auto i) {
                    double fi = f[i];
                    if constexpr (i < NS) {
                        if (s_ig[i][lane] < im) fi = 0.0;  // faster groups are frozen
                    }
                    const double err = dt * fi, ay0 = fabs(y[i]);
                    const double yn = y[i] + err;
                    s_y[i][lane] = yn;
                    if constexpr (i < NS) {
                        const double tol = atol + rtol * fmax(fabs(yn), ay0);
                        if (s_ig[i][lane] == im) e = fmax(e, fabs(err) / tol);
                    }
                });
                const double nd = did + dt;
                CT(C_DID) = nd;
                CNT(S_NST) += 1;
                if (!(nd < tn)) finish_cell(false);
                else if (im != Nm - 1 && e < 1.0 && !next_group(im)) finish_cell(false);
            }
        }
    }
    const long long s0 = wave_sum(CNT(S_CELLS)), s1 = wave_sum(CNT(S_STEPS)),
                    s5 = wave_max(CNT(S_MAXS)), s7 = wave_sum(CNT(S_FAIL));
    const double dmin = wave_min(CT(C_DTMIN));
#undef CT
#undef CNT
    if (lane == 0) {
        atomicAdd(&P.stats[0], (unsigned long long)s0);
        atomicAdd(&P.stats[1], (unsigned long long)s1);
        atomicAdd(&P.stats[2], (unsigned long long)s1);
        atomicAdd(&P.stats[4], (unsigned long long)s1);
        atomicMax(&P.stats[5], (unsigned long long)s5);
        if (dmin < __builtin_huge_val() && dmin >= 0.0)
            atomicMin(&P.stats[6], (unsigned long long)__double_as_longlong(dmin));
        atomicAdd(&P.stats[7], (unsigned long long)s7);
    }
}

// ---- RHS / Jacobian evaluation at given states (chemistrySource::DyDt / jacobian) ----
struct EvalParams {
    int64_t n;
    const double *y;
    int64_t ldy;
    const double *pr;
    double *f;
    int64_t ldf;
    double *J;
    int64_t ldJ;
    int flow;
    int exact;  // 1: exact d(Kp)/dT
    int prec;   // 1: split precision
};

template <class M, class F = double>
__global__ __launch_bounds__(64) void eval_lane_kernel(const EvalParams P) {
    constexpr int NS = M::NS, N = NS + 1;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n) return;
    double y[N], f[N], J[N][N];
    sfor<0, N>([&](auto i) { y[i] = P.y[i * P.ldy + cell]; });
    if (P.flow == 0) eval_fJ<M, true, 0, F>(y, P.pr[cell], f, J, P.exact != 0);
    else eval_fJ<M, true, 1, F>(y, P.pr[cell], f, J, P.exact != 0);
    sfor<0, N>([&](auto i) { P.f[i * P.ldf + cell] = f[i]; });
    if (P.J)
        sfor<0, N>([&](auto i) {
            sfor<0, N>([&](auto j) { P.J[(i * N + j) * P.ldJ + cell] = J[i][j]; });
        });
}

// ---- mixture thermo + transport (transportList::muKappa, Df; thermoList cp, ha) ----
struct PropParams {
    int64_t n;
    const double *T, *p, *Y;
    int64_t ldY;
    double *mu, *kappa, *Dk;
    int64_t ldD;
    double *cp, *ha, *hk;
    int prec;  // 1: species transport and mixture rules in FP32
    double *cv, *ea;  // thermoList::cv_p, ea_p [J/kg/K, J/kg] (nullable)
};

// F = float evaluates the kinetic-theory species properties, Wilke's mixing
// rule and the binary / mixture-averaged diffusivities in FP32; thermo (cp,
// h) and the mole fractions stay FP64.
template <class M, class F = double>
__global__ __launch_bounds__(256) void props_lane_kernel(const PropParams P) {
    constexpr int NS = M::NS;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n) return;
    const double T = P.T[cell], p = P.p[cell];
    double yi[NS], xi[NS], cpi[NS], mui[NS], kap[NS];
    sfor<0, NS>([&](auto k) { yi[k] = P.Y[k * P.ldY + cell]; });
    const TFac tf = tfac(T);
    // species cp0 and ha0 [J/kg] (JANAF.cpp:155-189)
    double cpm = 0.0, ham = 0.0;
    sfor<0, NS>([&](auto k) {
        Thermo<M> th;
        species_thermo<M, k>(T, tf.lnT, tf.invTm, th);
        cpi[k] = th.Cp[k] / M::W[k];
        const double h = th.H[k] / M::W[k];
        cpm += cpi[k] * yi[k];
        ham += h * yi[k];
        if (P.hk) P.hk[k * P.ldD + cell] = h;
    });
    if (P.cp) P.cp[cell] = cpm;
    if (P.ha) P.ha[cell] = ham;
    if (P.cv || P.ea) {
        // perfect gas: cv = cp - Rm, ea = ha - p / rho = ha - Rm T
        // (thermoList.inl:145-156, :289-298; speciesList.cpp:83-86, :106-116)
        double Wm;
        if constexpr (NS == 1) {
            Wm = M::W[0];
        } else {
            double s = 0.0;
            sfor<0, NS>([&](auto k) { s += yi[k] / M::W[k]; });
            Wm = 1.0 / fmax(s, kVeryTiny);
        }
        const double Rm = kRu / Wm;
        if (P.cv) P.cv[cell] = cpm - Rm;
        if (P.ea) P.ea[cell] = ham - Rm * T;
    }
    // Yi2Xi (speciesList.cpp:208-222)
    if constexpr (NS == 1) {
        xi[0] = 1.0;
    } else {
        double wm = 0.0, xt = 0.0;
        sfor<0, NS>([&](auto k) { wm += yi[k] * (1.0 / M::W[k]); });
        wm = 1.0 / fmax(wm, kVeryTiny);
        sfor<0, NS>([&](auto k) {
            xi[k] = fmax(0.0, fmin(1.0, yi[k] * wm / M::W[k]));
            xt += xi[k];
        });
        if (xt != 1.0) {
            const double r = 1.0 / xt;
            sfor<0, NS>([&](auto k) { xi[k] *= r; });
        }
    }
    // kineticTheory::mu / kappa (kineticTheory.cpp:37-57, kineticTheory.inl:32-34)
    const F TF = (F)T;
    F muF[NS], kaF[NS], xF[NS];
    sfor<0, NS>([&](auto k) {
        xF[k] = (F)xi[k];
        const F Ts = TF * (F)(1.0 / M::EKB[k]);
        const F om = (F)1.147 * exp((F)-0.145 * log(Ts)) + (F)1.0 / ((Ts + (F)0.5) * (Ts + (F)0.5));
        muF[k] = (F)2.6693e-6 * sqrt((F)M::W[k] * TF) / ((F)(M::SIGMA[k] * M::SIGMA[k]) * om);
        kaF[k] = muF[k] * (F)M::Ri[k] * ((F)cpi[k] / (F)M::Ri[k] + (F)1.25);
    });
    if (P.mu || P.kappa) {
        F mum = 0, km = 0;
        if constexpr (NS == 1) {
            mum = muF[0];
            km = kaF[0];
        } else {
            // Wilke (transportList.cpp:309-333)
            sfor<0, NS>([&](auto i) {
                F phi = 0;
                sfor<0, NS>([&](auto j) {
                    constexpr double Wi = M::W[i], Wj = M::W[j];
                    constexpr F phiP = (F)(1.0 / ohcd::csqrt(8.0) / ohcd::csqrt(1.0 + Wi / Wj));
                    constexpr F wr = (F)ohcd::csqrt(Wj / Wi);
                    const F tt = (F)1.0 + sqrt(muF[i] / muF[j] * wr);
                    phi += phiP * xF[j] * (tt * tt);
                });
                const F rphi = (F)1.0 / phi;
                mum += muF[i] * xF[i] * rphi;
                km += kaF[i] * xF[i] * rphi;
            });
        }
        if (P.mu) P.mu[cell] = (double)mum;
        if (P.kappa) P.kappa[cell] = (double)km;
    }
    if (P.Dk) {
        // kineticTheory::Di (kineticTheory.cpp:118-168)
        const F Tv1p5 = TF * sqrt(TF), rp = (F)(1.0 / p);
        sfor<0, NS>([&](auto i) {
            F sum = 0;
            sfor<0, NS>([&](auto j) {
                if constexpr (i != j) {
                    constexpr double Wi = M::W[i], Wj = M::W[j];
                    constexpr double sij = 0.5 * (M::SIGMA[i] + M::SIGMA[j]);
                    constexpr F cst = (F)(1.858e-7 * ohcd::csqrt(1.0 / Wi + 1.0 / Wj) / (sij * sij) * kPatm);
                    constexpr F rek = (F)(1.0 / ohcd::csqrt(M::EKB[i] * M::EKB[j]));
                    const F Ts = TF * rek;
                    const F Od = exp((F)-0.145 * log(Ts)) + (F)1.0 / ((Ts + (F)0.5) * (Ts + (F)0.5));
                    const F Dij = Tv1p5 * cst * rp / Od;
                    sum += xF[j] / Dij;
                }
            });
            double Dim = xi[i] < 1.0 - kVeryTiny ? (1.0 - xi[i]) / (double)sum : 0.0;
            if (isinf(Dim)) Dim = 0.0;
            P.Dk[i * P.ldD + cell] = Dim;
        });
    }
}

// ---- source terms: calculateSourceTerms / calculateSourceTermsImp ----
// R_k = W_k omega_k(rho Y / W, T) and, with diag, the diagonal
// d(omega_k)/d(c_k) plus the temperature coupling at constant internal energy
// d(omega_k)/dT * W_k (-e_k) / (rho cv) (chemistrySource.cpp:1371-1450,
// :319-480 with the diagonal of reactionTypes::DqDci, :447-584).  With
// withoutLast the last species is eliminated through sum Y = 1
// (omegaCoupled): the diagonal of species i < ns-1 becomes
// J_ii - J_i,last W_i / W_last + dw_i/dT W_i (e_last - e_i) / (rho cv).
struct SourceParams {
    int64_t n;
    const double *T, *p, *rho, *Y;
    int64_t ldY;
    double *R;
    int64_t ldR;
    double *D;  // nullable
    int64_t ldD;
    int without_last;
    int exact;
};

template <class M>
__global__ __launch_bounds__(64) void source_lane_kernel(const SourceParams P) {
    constexpr int NS = M::NS, N = NS + 1;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n) return;
    const double T = P.T[cell], rho = P.rho[cell];
    double y[N], f[N], J[N][N], Yk[NS];
    sfor<0, NS>([&](auto k) {
        Yk[k] = P.Y[k * P.ldY + cell];
        y[k] = rho * Yk[k] / M::W[k];
    });
    y[NS] = T;
    eval_fJ<M, true, 0>(y, P.p[cell], f, J, P.exact != 0);
    sfor<0, NS>([&](auto k) { P.R[k * P.ldR + cell] = f[k] * M::W[k]; });
    if (!P.D) return;
    // e_k = ha_k - R_k T per unit mass; cv = sum Y_k (cp_k - R_k)
    double ea[NS], cv = 0.0;
    sfor<0, NS>([&](auto k) {
        double Cp, H, dCp;
        species_CH<M, k>(T, Cp, H, dCp);
        ea[k] = H / M::W[k] - M::Ri[k] * T;
        cv += (Cp / M::W[k] - M::Ri[k]) * Yk[k];
    });
    const double b = 1.0 / (rho * cv);
    if (!P.without_last) {
        sfor<0, NS>([&](auto i) {
            P.D[i * P.ldD + cell] = J[i][i] + J[i][NS] * M::W[i] * (-ea[i]) * b;
        });
    } else {
        constexpr int L = NS - 1;
        sfor<0, L>([&](auto i) {
            P.D[i * P.ldD + cell] = J[i][i] - J[i][L] * (M::W[i] / M::W[L]) +
                                    J[i][NS] * M::W[i] * (ea[L] - ea[i]) * b;
        });
    }
}

// ---- thermoList::THa_p Newton inversion (thermoList.cpp:504-533) ----
struct ThParams {
    int64_t n;
    const double *h, *Y;  // h: target ha, or ea when energy = 1 [J/kg]
    int64_t ldY;
    double *T;
    int *iflag;
    double Tlow, Thigh;
    int energy;  // 0: THa (F = ha, dF/dT = cp); 1: TEa (F = ea = ha - Rm T, dF/dT = cv = cp - Rm)
};

// Rm = Ru / W of a mass-fraction vector (speciesList::Rm / MWbyYi,
// speciesList.cpp:83-86, :106-116), species in order.
template <class M>
__device__ __forceinline__ double mix_Rm(const double *yi) {
    if constexpr (M::NS == 1) {
        return kRu / M::W[0];
    } else {
        double s = 0.0;
        sfor<0, M::NS>([&](auto k) { s += yi[k] / M::W[k]; });
        return kRu / (1.0 / fmax(s, kVeryTiny));
    }
}

template <class M>
__global__ __launch_bounds__(256) void tfromh_lane_kernel(const ThParams P) {
    constexpr int NS = M::NS;
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= P.n) return;
    double yi[NS];
    sfor<0, NS>([&](auto k) { yi[k] = P.Y[k * P.ldY + cell]; });
    const double hstar = P.h[cell];
    const double T0 = P.T[cell], Ttol = T0 * 1.0e-4;
    const double Rm = P.energy ? mix_Rm<M>(yi) : 0.0;
    double Test = T0, Tnew = T0;
    int iter = 0, flag = 1;
    do {
        Test = Tnew;
        const TFac tf = tfac(Test);
        double hm = 0.0, cpm = 0.0;
        sfor<0, NS>([&](auto k) {
            Thermo<M> th;
            species_thermo<M, k>(Test, tf.lnT, tf.invTm, th);
            hm += th.H[k] / M::W[k] * yi[k];
            cpm += th.Cp[k] / M::W[k] * yi[k];
        });
        if (P.energy) {
            hm -= Rm * Test;
            cpm -= Rm;
        }
        double Tn = Test - (hm - hstar) / cpm;
        if (Tn < P.Tlow) {
            flag = 0;
            Tn = P.Tlow;
        } else if (Tn > P.Thigh) {
            flag = 2;
            Tn = P.Thigh;
        } else {
            flag = 1;
        }
        Tnew = Tn;
        if (iter++ > 100) break;
    } while (fabs(Tnew - Test) > Ttol);
    P.T[cell] = Tnew;
    if (P.iflag) P.iflag[cell] = flag;
}

// Host-side entry of one compiled mechanism.
struct LaneEntry {
    const char *name;
    unsigned long long hash, thash;
    int ns;
    hipError_t (*integrate)(const LaneParams &, int grid, hipStream_t);
    hipError_t (*eval)(const EvalParams &, hipStream_t);
    hipError_t (*props)(const PropParams &, hipStream_t);
    hipError_t (*tfromh)(const ThParams &, hipStream_t);
    const void *integrate_fn;
    hipError_t (*source)(const SourceParams &, hipStream_t);
};

template <class M>
struct LaneLaunch {
    static hipError_t integrate(const LaneParams &p, int grid, hipStream_t s) {
        if (p.integrator == 4) {
            if (p.flow == 0) hipLaunchKernelGGL((mts_lane_kernel<M, 0>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((mts_lane_kernel<M, 1>), dim3(grid), dim3(64), 0, s, p);
        } else if (p.integrator == 1) {
            if (p.flow == 0) hipLaunchKernelGGL((ee_lane_kernel<M, 0>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((ee_lane_kernel<M, 1>), dim3(grid), dim3(64), 0, s, p);
        } else if (p.integrator == 2) {
            if (p.flow == 0) hipLaunchKernelGGL((ie_lane_kernel<M, 0, double, 2>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((ie_lane_kernel<M, 1, double, 2>), dim3(grid), dim3(64), 0, s, p);
        } else if (p.prec == 0) {
            if (p.flow == 0) hipLaunchKernelGGL((ie_lane_kernel<M, 0>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((ie_lane_kernel<M, 1>), dim3(grid), dim3(64), 0, s, p);
        } else {
            if (p.flow == 0) hipLaunchKernelGGL((ie_lane_kernel<M, 0, float>), dim3(grid), dim3(64), 0, s, p);
            else hipLaunchKernelGGL((ie_lane_kernel<M, 1, float>), dim3(grid), dim3(64), 0, s, p);
        }
        return hipGetLastError();
    }
    static hipError_t eval(const EvalParams &p, hipStream_t s) {
        const int g = (int)((p.n + 63) / 64);
        if (p.prec == 0) hipLaunchKernelGGL((eval_lane_kernel<M>), dim3(g), dim3(64), 0, s, p);
        else hipLaunchKernelGGL((eval_lane_kernel<M, float>), dim3(g), dim3(64), 0, s, p);
        return hipGetLastError();
    }
    static hipError_t props(const PropParams &p, hipStream_t s) {
        const int g = (int)((p.n + 255) / 256);
        if (p.prec == 0) hipLaunchKernelGGL((props_lane_kernel<M>), dim3(g), dim3(256), 0, s, p);
        else hipLaunchKernelGGL((props_lane_kernel<M, float>), dim3(g), dim3(256), 0, s, p);
        return hipGetLastError();
    }
    static hipError_t tfromh(const ThParams &p, hipStream_t s) {
        const int g = (int)((p.n + 255) / 256);
        hipLaunchKernelGGL(tfromh_lane_kernel<M>, dim3(g), dim3(256), 0, s, p);
        return hipGetLastError();
    }
    static hipError_t source(const SourceParams &p, hipStream_t s) {
        const int g = (int)((p.n + 63) / 64);
        hipLaunchKernelGGL(source_lane_kernel<M>, dim3(g), dim3(64), 0, s, p);
        return hipGetLastError();
    }
    static LaneEntry entry(const char *name) {
        return LaneEntry{name,    M::HASH, M::THASH, M::NS,
                         &integrate, &eval,  &props,   &tfromh,
                         (const void *)&ie_lane_kernel<M, 0>, &source};
    }
};

}  // namespace ohcd

// Registry of compiled mechanisms (defined in the generated registry TU).
extern "C" const ohcd::LaneEntry *ohc_lane_entries(int *count);
