/*
 * ohc.h -- C-ABI of the MI355X per-cell chemistry engine (libohc.so).
 *
 * This is the drop-in boundary for OpenHurricane's per-cell chemistry hot path.
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (OHC_OK = 0) and never aborts the process; the text of the last
 * error is read with ohc_last_error(ctx) (ctx may be NULL for errors raised
 * before a context exists).
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * OpenHurricane source tree):
 *   ohc_mech_load / ohc_mech_from_tables
 *       mixture::cuParsingReaction / cuAddThermoToCUDA + cuChemInterface
 *       (src/speciesTransport/species/src/mixture.cpp:400-528,
 *        src/speciesTransport/reactions/cuda/include/cuChemInterface.hpp:61-310)
 *       and the CHEMKIN reader chemkinFileRead
 *       (src/speciesTransport/chemistryRead/src/chemkinFileRead.cpp:1349-1700).
 *   ohc_ctx_create
 *       chemistrySource::chemistrySource + ODEsSolver::ODEsSolver +
 *       ImplicitEuler::ImplicitEuler controller keys
 *       (src/speciesTransport/chemistrySource/src/chemistrySource.cpp:41-85,
 *        src/Hurricane/mathematicalTools/ODEs/src/ODEsSolver.cpp:80-92,
 *        src/Hurricane/mathematicalTools/ODEs/src/ImplicitEuler.cpp:68-84).
 *   ohc_chem_solve
 *       real chemistrySource::solve(const realArray& dt, const real dtFactor)
 *       (chemistrySource.hpp:622, chemistrySource.cpp:1494-1556).
 *   ohc_chem_solve_update
 *       void chemistrySource::solveUpdate(const realArray& dt, const real dtFactor)
 *       (chemistrySource.hpp:628, chemistrySource.cpp:1558-1609).
 *   ohc_chem_source
 *       chemistrySource::calculateSourceTerms / calculateSourceTermsImp
 *       (chemistrySource.hpp:597-598, chemistrySource.cpp:1371-1450).
 *   ohc_chem_source_kim
 *       CUDAChemistrySource::calculateSourceTermsImp / ...ImpAsync / ...ImpAsyncHybrid
 *       (src/speciesTransport/chemistrySource/src/CUDAChemistrySource.cpp:152-250,
 *        CUDAChemSourceKimDiagonalKernel.cu:39-235).
 *   ohc_eval_rhs / ohc_eval_jacobian
 *       chemistrySource::DyDt / chemistrySource::jacobian
 *       (chemistrySource.cpp:863-907, :967-1055).
 *   ohc_lu_solve
 *       SquareMatrix::LUDecomposeDPivoting + LUBacksubstituteDPivoting
 *       (SquareMatrix.hpp:185-238, LUDecompose.hpp:68-96).
 *   ohc_thermo_T_from_h
 *       thermoList::THa_p / THa_rho (src/speciesTransport/thermo/include/thermoList.inl:315-340,
 *       thermoList.cpp:504-533).
 *   ohc_thermo_T_from_e
 *       thermoList::TEa_rho / TEa_p (thermoList.inl:343-370 -> thermoList.cpp:504-533 with
 *       ea_rho / ea_p and cv_rho / cv_p, thermoList.inl:145-170, :289-311); caller
 *       rhoThermo::TEa (src/speciesTransport/basic/src/rhoThermo.cpp:87-93).
 *   ohc_mix_thermo
 *       thermoList::cp_p, cv_p, ha_p, ea_p (thermoList.cpp:165-290, thermoList.inl:43-311).
 *   ohc_mix_props
 *       mixture::gasProperties / transportList::muKappa / transportList::Df and
 *       CUDATransport::calcTransportPropertiesAsync
 *       (src/speciesTransport/transport/src/transportList.cpp:286-427,
 *        src/speciesTransport/transport/include/transportListCUDA.hpp:121-128).
 *
 * Data layout.  Cell fields are structure-of-arrays: scalars are arrays of
 * n doubles; species fields are one block with leading dimension ld
 * (species k of cell i is at Y[k*ld + i], ld >= n), the contiguous form of the
 * reference's PtrList<cellRealArray> yi_[isp][cell].  Units follow the
 * reference: T [K], p [Pa], rho [kg/m^3], Y mass fractions, dt [s],
 * R [kg/m^3/s], concentrations [kmol/m^3].
 *
 * Pointers are host pointers unless ohc_opts.device_pointers is set, in which
 * case every array argument must be device memory of the context's device
 * and the call is asynchronous on the context stream (ohc_sync waits).
 */
#ifndef OHC_H
#define OHC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OHC_OK 0
#define OHC_ERR_ARG 1      /* invalid argument                          */
#define OHC_ERR_PARSE 2    /* mechanism / thermo / transport file error */
#define OHC_ERR_DEVICE 3   /* HIP runtime error                         */
#define OHC_ERR_UNSUPPORTED 4
#define OHC_ERR_NUMERIC 5  /* integration failed (step underflow, NaN)  */

#define OHC_MAXP 6         /* max distinct species on one side of a reaction */
#define OHC_MAXNS 64       /* max species handled by the device kernels      */

/* reaction reversibility (reactionTypes::types) */
#define OHC_RXN_IRREVERSIBLE 0
#define OHC_RXN_REVERSIBLE 1      /* kr = kf / Kc                           */
#define OHC_RXN_REV_PARAMS 2      /* kr from explicit REV Arrhenius params  */

/* rate-constant law (reactionRateTypes) */
#define OHC_RATE_ARRHENIUS 0
#define OHC_RATE_THIRDBODY 1      /* ArrheniusThirdBodyRR                   */
#define OHC_RATE_FALLOFF 2        /* unimolecularFallOffRR                  */
#define OHC_RATE_CHEMACT 3        /* chemicallyActicatedBimolecularRR       */
/* Read from CHEMKIN text and carried in the tables, but not evaluated on the
 * GPU: ohc_ctx_create refuses a mechanism that holds them, as the reference's
 * GPU table builder does ("Unsupported reaction rate type",
 * src/speciesTransport/reactions/reactionTypes/src/reversibleReactions.cpp:186,
 * irreversibleReactions.cpp:188, reversibleReactionWithRevPars.cpp:109). */
#define OHC_RATE_PLOG 4           /* logInterpolationRR (PLOG)              */
#define OHC_RATE_CHEB 5           /* ChebyshevPolynomialsRR (CHEB/TCHEB/PCHEB) */

/* fall-off broadening function (fallOffFunctions) */
#define OHC_FO_LINDEMANN 0
#define OHC_FO_TROE3 1
#define OHC_FO_TROE4 2
#define OHC_FO_SRI 3

/* Flat mechanism tables (SoA by species / by reaction).
 * All arrays are owned by the caller of ohc_mech_from_tables (copied in) or by
 * the ohc_mech returned by ohc_mech_get_tables (valid while it lives). */
typedef struct ohc_mech_tables {
    int32_t ns, nr, ne;
    /* species */
    const double *W;        /* [ns] molecular weight kg/kmol                     */
    const double *Tlow;     /* [ns] NASA-7 validity range and switch temperature */
    const double *Thigh;
    const double *Tcommon;
    const double *nasa_lo;  /* [ns*7] raw (dimensionless) NASA-7 a0..a6, T<Tcommon */
    const double *nasa_hi;  /* [ns*7] raw NASA-7 a0..a6, T>=Tcommon              */
    const double *elem;     /* [ns*ne] atoms of element e in species k            */
    const double *ekb;      /* [ns] Lennard-Jones well depth eps/k [K]           */
    const double *sigma;    /* [ns] Lennard-Jones diameter [Angstrom]            */
    /* reactions */
    const int32_t *nf, *nb;       /* [nr] distinct forward / backward participants */
    const int32_t *f_idx, *b_idx; /* [nr*OHC_MAXP] species index                  */
    const double *f_nu, *b_nu;    /* [nr*OHC_MAXP] stoichiometric coefficient     */
    const double *f_ord, *b_ord;  /* [nr*OHC_MAXP] concentration exponent         */
    const int32_t *rev;           /* [nr] OHC_RXN_*                               */
    const int32_t *kind;          /* [nr] OHC_RATE_*                              */
    const double *arr;            /* [nr*3] A, beta, Ta (SI kmol-m^3-s-K): kf, or
                                     k_inf for FALLOFF and CHEMACT                  */
    const double *arr_lo;         /* [nr*3] k0 for FALLOFF and CHEMACT             */
    const double *arr_rev;        /* [nr*3] REV parameters (rev == REV_PARAMS)     */
    const int32_t *fo_type;       /* [nr] OHC_FO_*                                */
    const double *fo;             /* [nr*5] Troe alpha,T***,T*,T** / SRI a,b,c,d,e */
    const double *eff;            /* [nr*ns] third-body efficiencies              */
    /* Per-reaction data of the PLOG and CHEB laws: reaction r owns
     * aux[aux_off[r] .. aux_off[r+1]) (aux_off has nr+1 entries; both may be
     * NULL when no reaction is PLOG or CHEB).
     *   PLOG: 4 values per pressure point, ascending in p: p [Pa], A [SI
     *         kmol-m^3-s], beta, Ta [K] (logInterpolationRR, PLOG lines
     *         chemkinFileRead.cpp:606-644, units :1595-1601).
     *   CHEB: N, M, Tmin [K], Tmax [K], Pmin [Pa], Pmax [Pa], then the N*M
     *         coefficients a(n,m) row by row (ChebyshevPolynomialsRR, CHEB /
     *         TCHEB / PCHEB lines chemkinFileRead.cpp:645-755, units
     *         :1623-1627; limits default to 300-2500 K and 0.001-100 atm). */
    const int32_t *aux_off;
    const double *aux;
} ohc_mech_tables;

/* Integrator and context options; defaults are the reference defaults. */
typedef struct ohc_opts {
    int32_t flow_type;      /* 0 ConstantPressure (default), 1 ConstantVolume     */
    int32_t max_steps;      /* solver/maxstep (10000)                             */
    double atol, rtol;      /* solver/atol, rtol (1e-15, 1e-4): Newton test       */
    double atol_inc;        /* ImplicitEuler/ATolIncrement (1e-7): step error     */
    double rtol_inc;        /* ImplicitEuler/RTolIncrement (1e-3)                 */
    double safe_scale;      /* solver/safescale (0.8)                             */
    double min_scale;       /* solver/minscale (0.25)                             */
    double max_scale;       /* solver/maxScale (2.0)                              */
    double alpha_inc;       /* solver/alphainc (0.25)                             */
    double T_threshold;     /* solveOptions/TemperatureThreshold (200 K)          */
    double T_alpha;         /* solveOptions/TemperatureAlpha (0.2)                */
    double T_beta;          /* solveOptions/TemperatureBeta (0.25)                */
    double dt_max;          /* chemicalMaxStep (1e15)                             */
    double T_low, T_high;   /* thermoList TLow / THigh for T(h) limits (50, 5000) */
    int32_t precision;      /* 0: FP64 everywhere; 1: split -- Gibbs energies,
                               equilibrium and Arrhenius rate constants, species
                               transport and mixing rules in FP32; concentration
                               products, net rates, Jacobian, LU, T(h) Newton,
                               mole fractions and mass sums in FP64 (compiled
                               mechanisms only)                                    */
    int32_t device_pointers;/* 1: array arguments are device pointers             */
    int32_t kernel;         /* 0 auto (specialised if compiled for this
                               mechanism, else generic), 1 lane-per-cell
                               specialised, 2 generic wavefront-per-cell (any
                               mechanism with ns <= 63 and <= 3 distinct species
                               per reaction side)                                  */
    int32_t waves_per_cu;   /* persistent-kernel occupancy target (0 = auto)      */
    int32_t schedule;       /* 0: cells in index order; 1: longest-first, ordered
                               by the per-cell Newton-iteration counts of the
                               previous call with the same n (first call: index
                               order).  Any order gives the same results.         */
    int32_t jacobian;       /* 0: the reference's analytic Jacobian, including its
                               sign convention for dKp/dT (reactionTypes.cpp:
                               198-215); 1: exact dKp/dT -- same converged
                               states within the Newton tolerance, far fewer
                               Newton iterations when heat release couples
                               strongly to the rates                              */
    int32_t integrator;     /* 0: ImplicitEuler (chemistrySource, the CPU path);
                               1: the reference GPU path -- explicit Euler with a
                               step-doubling error estimate (EulerCUDA), every
                               cell integrated, subDt the starting step, Y from
                               c without clipping; compiled mechanisms only;
                               2: ImplicitTrapezoidal (ImplicitTrapezoidal.cpp:
                               99-335), the reference's second-order implicit
                               alternative; FP64;
                               3: BDF23 at its default order 2 (BDF23.cpp:
                               54-509); generic kernels, FP64;
                               4: the reference GPU path's multi-time-scale
                               explicit Euler (MTSEuler, solveChemODEsMTS,
                               CUDAChemistrySourceODEs.cu:684-778): species
                               grouped by the decade of their time scale,
                               group steps of 0.8 x its smallest time scale,
                               atol/rtol the group's increment test; every
                               cell integrated, subDt left as it is, Y from c
                               without clipping; FP64                          */
    int64_t max_slice_cells;/* host-pointer calls: cells staged on the device at
                               a time (memory-driven slicing); 0 = as many as
                               half the free device memory holds.  Results do
                               not depend on it; the longest-first schedule
                               is not used within a sliced call              */
} ohc_opts;

/* Per-call counters (summed over cells) for roofline accounting. */
typedef struct ohc_stats {
    int64_t cells;          /* cells integrated (T >= threshold)                  */
    int64_t steps;          /* accepted sub-steps                                 */
    int64_t attempts;       /* sub-step attempts (accepted + rejected)            */
    int64_t newton_iters;   /* Jacobian + LU + back-solve count                   */
    int64_t rhs_evals;      /* extra RHS evaluations (first-step error estimate)  */
    int64_t max_cell_steps; /* largest accepted sub-step count of one cell        */
    double dt_min;          /* min over cells of the suggested sub-step           */
} ohc_stats;

typedef struct ohc_mech ohc_mech;
typedef struct ohc_ctx ohc_ctx;

int ohc_opts_default(ohc_opts *opts);

/* Mechanism.  chemkin: CHEMKIN-II reaction file; thermo: NASA-7 THERMO file
 * (may be NULL when the CHEMKIN file carries a THERMO section); transport:
 * CHEMKIN transport file (may be NULL: transport entry points then fail). */
int ohc_mech_load(const char *chemkin, const char *thermo, const char *transport,
                  ohc_mech **out);
int ohc_mech_from_tables(const ohc_mech_tables *t, ohc_mech **out);
int ohc_mech_get_tables(const ohc_mech *m, ohc_mech_tables *out);
int ohc_mech_species_name(const ohc_mech *m, int32_t k, const char **name);
/* Seeded synthetic element-balanced mechanism (C/H/O/N) of ns species and nr
 * reactions: A log-uniform, beta, Ea uniform, frac_rev reversible,
 * frac_tb third-body, random smooth NASA-7 and transport data. */
int ohc_mech_synthetic(int32_t ns, int32_t nr, uint64_t seed, double frac_rev,
                       double frac_tb, ohc_mech **out);
/* 64-bit content hash of the tables (selects specialised kernels). */
int ohc_mech_hash(const ohc_mech *m, uint64_t *hash);
int ohc_mech_free(ohc_mech *m);

int ohc_ctx_create(const ohc_mech *m, const ohc_opts *opts, int32_t device, ohc_ctx **out);
int ohc_ctx_free(ohc_ctx *ctx);
/* Use a caller stream (hipStream_t passed as void*) instead of the context's. */
int ohc_ctx_set_stream(ohc_ctx *ctx, void *stream);
/* Autotuner (no counterpart in the reference, whose GPU path uses a static
 * grid, src/speciesTransport/chemistrySource/include/nThreadsAndBlocks.inl:66-107):
 * times two flow steps of solveUpdate on the given sample of cells (private
 * copies; the arguments are not modified) for persistent-grid sizes CUs x w,
 * w from the resident blocks per CU the kernel allows downwards until a size
 * is clearly slower, and keeps the fastest w for this context (also into
 * *waves_per_cu, the timed step into *best_ms; both nullable).  The sample
 * should be large enough that no single cell's integration dominates the
 * step.  Arrays follow the context's pointer mode. */
int ohc_ctx_autotune(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                     const double *rho, const double *Y, int64_t ldY, const double *dt,
                     double dtFactor, int32_t *waves_per_cu, double *best_ms);
int ohc_sync(ohc_ctx *ctx);
const char *ohc_last_error(const ohc_ctx *ctx);
/* Name of the integrator kernel the context selected ("lane:<mech>" or "generic"). */
const char *ohc_ctx_kernel_name(const ohc_ctx *ctx);

/* chemistrySource::solve(dt, dtFactor).  For each cell: T < T_threshold ->
 * R = 0 and subDt = dt*dtFactor; otherwise integrate [0, dt*dtFactor] at
 * constant p (or rho) from c = rho*Y/W, write R_k = (c_k - c0_k) W_k / (dt*dtFactor)
 * and the suggested sub-step into subDt.  subDt is read as the starting
 * sub-step (dtInit); a value <= 0 means dt*dtFactor (the reference's first
 * call).  stats (host, nullable) receives counters and min subDt. */
int ohc_chem_solve(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                   const double *rho, const double *Y, int64_t ldY, const double *dt,
                   double dtFactor, double *subDt, double *R, int64_t ldR,
                   ohc_stats *stats);

/* chemistrySource::solveUpdate(dt, dtFactor): Y and T updated in place. */
int ohc_chem_solve_update(ohc_ctx *ctx, int64_t n, double *T, const double *p,
                          const double *rho, double *Y, int64_t ldY, const double *dt,
                          double dtFactor, double *subDt, ohc_stats *stats);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  ohc_chem_solve_update on device arrays (the context must
 * have device_pointers = 1) that allocates nothing: the integrator's work-queue
 * head, its counters and the longest-first schedule live in a static device
 * array of the library (33.6 MB on each device it runs on, held until the
 * process ends) instead of buffers the library allocates.  n <= 
 * OHC_NOALLOC_MAX_CELLS, larger n is rejected with OHC_ERR_ARG.  With
 * opts.schedule the cells are taken longest-first by the Newton counts of the
 * previous call on this device when it had the same n (else in index order);
 * results do not depend on the order.  All calls on one device share the
 * array: calls that overlap in time on one device race (serialise them, e.g.
 * on one stream).  stats (host, nullable): when given, the call synchronises
 * the stream and reports failed cells as OHC_ERR_NUMERIC; without it the call
 * is asynchronous and performs no host synchronisation. */
#define OHC_NOALLOC_MAX_CELLS 4194304
int ohc_chem_solve_update_noalloc(ohc_ctx *ctx, int64_t n, double *T, const double *p,
                                  const double *rho, double *Y, int64_t ldY, const double *dt,
                                  double dtFactor, double *subDt, ohc_stats *stats);

/* calculateSourceTerms(false) (diagJ == NULL) and calculateSourceTermsImp
 * (diagJ != NULL, diagonal d(omega_k)/d(c_k) incl. the temperature coupling
 * term; withoutLastSpec selects the omegaCoupled last-species substitution). */
int ohc_chem_source(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                    const double *rho, const double *Y, int64_t ldY, double *R,
                    int64_t ldR, double *diagJ, int64_t ldD, int32_t withoutLastSpec);

/* The GPU path's calculateSourceTermsImp / calculateSourceTermsImpAsync /
 * calculateSourceTermsImpAsyncHybrid (the "Kim" diagonal kernel): R_k =
 * omega_k W_k (FP64, ns rows of ldR) and, into diag (nullable, ns-1 rows of
 * ldD), the mass-action diagonal d(omega_k)/d(c_k) of species k < ns-1 with the
 * last species eliminated by mass conservation (d c_ns / d c_k = -W_k / W_ns);
 * the third-body and fall-off factors of kf and kr are held fixed and there is
 * no temperature coupling.  diag_fp32 = 0: diag is double*; 1: float* (the
 * Hybrid variant).  Needs the run-time device tables (ns <= 63, at most 3
 * distinct species per reaction side). */
int ohc_chem_source_kim(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                        const double *rho, const double *Y, int64_t ldY, double *R, int64_t ldR,
                        void *diag, int64_t ldD, int32_t diag_fp32);

/* RHS and analytic Jacobian at ODE states y = [c_1..c_ns, T] (ld = row stride of
 * the (ns+1) x n block); J is (ns+1)^2 x n, entry (i,j) of cell c at
 * J[(i*(ns+1)+j)*ldJ + c].  p (or rho0 for ConstantVolume) per cell. */
int ohc_eval_rhs(ohc_ctx *ctx, int64_t n, const double *y, int64_t ldy,
                 const double *p, double *f, int64_t ldf);
int ohc_eval_jacobian(ohc_ctx *ctx, int64_t n, const double *y, int64_t ldy,
                      const double *p, double *f, int64_t ldf, double *J, int64_t ldJ);

/* Batched dense solves A_m x_m = b_m of nmat n x n systems (1 <= n <= 64)
 * with partial pivoting: SquareMatrix::LUDecomposeDPivoting followed by
 * LUBacksubstituteDPivoting (src/Hurricane/dataStructure/containers/Matrices/
 * include/SquareMatrix.hpp:185-238, src/Hurricane/mathematicalTools/
 * LinearEquations/include/LUDecompose.hpp:68-96), the factorisation the
 * integrators run on I - h J.  Entry (i,j) of system m at A[(i*n+j)*ldA + m]
 * (the ohc_eval_jacobian layout), b_i and x_i at b[i*ldb + m], x[i*ldx + m].
 * The pivot of each column is the largest |a_ik| to 2^-20 relative (lowest
 * row among those); a column whose remaining rows are all NaN gives an
 * all-NaN x.  Independent of the context's mechanism. */
int ohc_lu_solve(ohc_ctx *ctx, int64_t nmat, int32_t n, const double *A, int64_t ldA,
                 const double *b, int64_t ldb, double *x, int64_t ldx);

/* thermoList::THa_p: T from absolute enthalpy h [J/kg] at pressure p and
 * composition Y, Newton from T_inout (tol 1e-4*T0, 100 iterations, limited to
 * [T_low, T_high]); iflag (nullable) receives 0/1/2 as thermoList::limit. */
int ohc_thermo_T_from_h(ohc_ctx *ctx, int64_t n, const double *h, const double *p,
                        const double *Y, int64_t ldY, double *T_inout, int32_t *iflag);

/* thermoList::TEa_p / TEa_rho: T from absolute internal energy e [J/kg]
 * (e = ha - p/rho = ha - Rm T for the perfect gas), Newton on e(T) with
 * de/dT = cv = cp - Rm, otherwise as ohc_thermo_T_from_h.  p_or_rho is not
 * needed by the perfect-gas equation of state and may be NULL. */
int ohc_thermo_T_from_e(ohc_ctx *ctx, int64_t n, const double *e, const double *p_or_rho,
                        const double *Y, int64_t ldY, double *T_inout, int32_t *iflag);

/* Mixture thermo at (T, p, Y): cp, cv [J/kg/K], ha, ea [J/kg]; any output may
 * be NULL.  cv = cp - Rm, ea = ha - Rm T (perfect gas). */
int ohc_mix_thermo(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                   const double *Y, int64_t ldY, double *cp, double *cv, double *ha,
                   double *ea);

/* Mixture thermo + transport at (T, p, Y): any output may be NULL.
 * mu [kg/m/s] and kappa [W/m/K] by kinetic theory + Wilke; Dk [m^2/s]
 * mixture-averaged (ld ldD); cp [J/kg/K]; ha [J/kg]; hk species h_a [J/kg] (ld ldD). */
int ohc_mix_props(ohc_ctx *ctx, int64_t n, const double *T, const double *p,
                  const double *Y, int64_t ldY, double *mu, double *kappa, double *Dk,
                  int64_t ldD, double *cp, double *ha, double *hk);

/* ---- chemistry surrogate (no counterpart in the reference) ------------------
 * A multilayer perceptron that maps one cell's state to its chemistry
 * increment over the flow time step it was trained for:
 *   inputs  [T, ln p, ln max(Y_k, 1e-20)]          (ns + 2)
 *   outputs [T1 - T0, ln Y1_k - ln Y0_k]            (ns + 1)
 * with `depth` hidden layers of `width` GELU units, BF16 weights and
 * activations, FP32 accumulation on the matrix cores.  It is trained on state
 * pairs produced by ohc_chem_solve_update itself.  The surrogate belongs to a
 * context created with device_pointers = 1; every array argument below is
 * device memory. */
typedef struct ohc_surrogate ohc_surrogate;

int ohc_surrogate_create(ohc_ctx *ctx, int32_t width, int32_t depth, uint64_t seed,
                         ohc_surrogate **out);
int ohc_surrogate_free(ohc_surrogate *s);
int ohc_surrogate_num_params(const ohc_surrogate *s, int64_t *n);
/* FP32 master parameters (host arrays): per layer W [out][in] then b [out] */
int ohc_surrogate_get_params(ohc_surrogate *s, float *params);
int ohc_surrogate_set_params(ohc_surrogate *s, const float *params);
/* Host arrays: feature mean/std and the accepted feature box [lo, hi] (NULL:
 * unbounded), target mean/std, per-output loss weights (NULL: 1), and the
 * flow time step the pairs were generated with. */
int ohc_surrogate_set_norm(ohc_surrogate *s, const float *in_mean, const float *in_std,
                           const float *in_lo, const float *in_hi, const float *out_mean,
                           const float *out_std, const float *out_weight, double dt_train);
/* Target transform (host array of ns+1 scales, NULL: none): output k is
 * trained on z = sign(d) ln(1 + |d| / delta_k) of the increment d, and the
 * inference inverts it, so that small increments are resolved as finely as
 * large ones.  Set before computing the training targets. */
int ohc_surrogate_set_target_scale(ohc_surrogate *s, const float *delta);
/* Inference precision: 0 BF16 (default); 1 FP8 E4M3 weights (per-row scale)
 * and activations for the hidden-to-hidden layers on the f8f6f4 matrix
 * instructions, BF16 first and last layers.  Training stays BF16. */
int ohc_surrogate_set_precision(ohc_surrogate *s, int32_t precision);
/* Gate on a second network of the same context (NULL: none): a cell passes
 * only when both networks agree on dT within tol * T and on the ln Y
 * increment of every species with Y >= 1e-4 before or after the step within
 * tol; it then takes the mean of the two increments.  The peer must outlive
 * the calls. */
int ohc_surrogate_set_peer(ohc_surrogate *s, ohc_surrogate *peer, double tol);
/* Raw features X [n][ns+2] of states (T0, p, Y0) and, when T1/Y1 are given,
 * targets Yt [n][ns+1] of the pairs (T0, Y0) -> (T1, Y1): the increments, or
 * their transform when target scales are set. */
int ohc_surrogate_features(ohc_surrogate *s, int64_t n, const double *T0, const double *p,
                           const double *Y0, int64_t ldY, const double *T1, const double *Y1,
                           float *X, float *Yt);
/* One Adam step (lr) on a batch of raw features / targets; the batch loss
 * (weighted mean-square error of the normalised targets) into *loss (host,
 * nullable -- a non-null loss synchronises the stream). */
int ohc_surrogate_train_step(ohc_surrogate *s, int64_t n, const float *X, const float *Yt,
                             float lr, float *loss);
/* Loss gradient with respect to every parameter (host array, parameter
 * layout) for a batch -- the backward kernels without the update. */
int ohc_surrogate_gradient(ohc_surrogate *s, int64_t n, const float *X, const float *Yt,
                           float *grad, float *loss);
/* Predicted increments dout [n][ns+1] (dT, d ln Y; the target transform
 * inverted) at states (T, p, Y). */
int ohc_surrogate_predict(ohc_surrogate *s, int64_t n, const double *T, const double *p,
                          const double *Y, int64_t ldY, float *dout);
/* solveUpdate with the surrogate: every cell is advanced by the network unless
 * its gate fails -- time step differs from the training step, a feature lies
 * outside the accepted box, the new T leaves [T_low, T_high], the relative
 * change of any element's mass exceeds gate_tol, or the peer network (if
 * set) disagrees -- and the failed cells
 * (count in *n_fallback) are integrated by the plain integrator, whose
 * counters go to stats. */
int ohc_chem_solve_update_surrogate(ohc_ctx *ctx, ohc_surrogate *s, int64_t n, double *T,
                                    const double *p, const double *rho, double *Y, int64_t ldY,
                                    const double *dt, double dtFactor, double *subDt,
                                    double gate_tol, int64_t *n_fallback, ohc_stats *stats);
int ohc_surrogate_save(const ohc_surrogate *s, const char *path);
int ohc_surrogate_load(ohc_surrogate *s, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* OHC_H */
