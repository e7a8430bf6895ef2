/*
 * ohc_oracle.h -- FP64 CPU oracle of OpenHurricane's per-cell chemistry path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Product code never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the OpenHurricane source tree).  Parity pins (what checks each part
 * independently of this restatement) are listed in tests/test_oracle_*.py:
 *   thermo        closed forms + cp = dh/dT, g = h - Ts, s' = cp/T
 *   rates / Kc    closed-form Arrhenius, detailed balance, van 't Hoff
 *   omega / RHS   mass and element conservation, dh/dt = 0 at constant p
 *   Jacobian      central finite differences of the RHS
 *   LU            numpy.linalg.solve, pivot-requiring matrices
 *   Kim diagonal  J_kk - (W_k / W_ns) J_k,ns of the analytic Jacobian; the
 *                 FP32 variant within float rounding of the FP64 one
 *   MTS           time scales = 1 / |Jacobian diagonal| (mass-action part);
 *                 closed form of an isomerisation A -> B; one-group cells =
 *                 forward Euler, error O(tn^2) against Radau
 *   integrator    scipy Radau at rtol 1e-12 (accuracy), conservation; the
 *                 trapezoidal rule at second order; BDF23 between step
 *                 tolerances 1e-3 and 1e-4 (and its underflow at 1e-8)
 *   T(h), T(e)    round trips h(T(h)) = h, e(T(e)) = e, and limiter flags
 *   cv, ea        cv = de/dT by finite differences, ea = ha - p/rho (perfect gas)
 *   transport     single-species and binary closed forms, symmetry
 */
#ifndef OHC_ORACLE_H
#define OHC_ORACLE_H

#include "../include/ohc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mech orc_mech;

orc_mech *orc_mech_new(const ohc_mech_tables *t);
void orc_mech_free(orc_mech *m);

/* Molar species thermo at T (J/kmol/K, J/kmol): Cp, dCp/dT, H, G, dG/dT, S. */
void orc_species_thermo(const orc_mech *m, double T, double *Cp, double *dCpdT, double *H,
                        double *G, double *dGdT, double *S);
/* Per-reaction kf, kr, qf, qr at (p, T, c) with c clamped at 0 as the reference. */
void orc_rates(const orc_mech *m, double p, double T, const double *c, double *kf,
               double *kr, double *qf, double *qr);
/* Kc and dKc/dT of each reaction (reference sign convention). */
void orc_equilibrium(const orc_mech *m, double T, double *Kc, double *dKcdT);
/* Net production rates omega_k [kmol/m^3/s]. */
void orc_omega(const orc_mech *m, double p, double T, const double *c, double *w);

/* RHS f = dy/dt and Jacobian J = df/dy ((ns+1)^2 row-major) at y = [c, T].
 * pr = p for ConstantPressure, rho0 for ConstantVolume.  exact_dkp = 1
 * replaces the reference's dKp/dT sign by the thermodynamic one (test aid). */
void orc_dydt(const orc_mech *m, int flow_type, double pr, const double *y, double *f);
void orc_jacobian(const orc_mech *m, int flow_type, double pr, const double *y, double *f,
                  double *J, int exact_dkp);

/* Doolittle LU with partial pivoting + back substitution, in place on A (n x n
 * row-major) and b.  Returns the permutation sign. */
int orc_lu_solve(int n, double *A, double *b);

/* chemistrySource::solve(dt, dtFactor) / solveUpdate(dt, dtFactor) over n
 * cells with nthreads host threads (static contiguous ranges).  Arguments as
 * ohc_chem_solve / ohc_chem_solve_update; per-cell counters are written to
 * cell_steps (accepted sub-steps, nullable).  Returns 0 or OHC_ERR_NUMERIC. */
int orc_chem_solve(const orc_mech *m, const ohc_opts *o, long n, const double *T,
                   const double *p, const double *rho, const double *Y, long ldY,
                   const double *dt, double dtFactor, double *subDt, double *R, long ldR,
                   int nthreads, ohc_stats *stats, int *cell_steps);
int orc_chem_solve_update(const orc_mech *m, const ohc_opts *o, long n, double *T,
                          const double *p, const double *rho, double *Y, long ldY,
                          const double *dt, double dtFactor, double *subDt, int nthreads,
                          ohc_stats *stats, int *cell_steps);

/* MTS explicit Euler (integrator 4): species time scales at (T, c) and one
 * cell advanced over tn (y = [c, T] in place); *Nm = number of time-scale
 * groups, *steps = explicit steps.  Returns 0 or OHC_ERR_NUMERIC. */
void orc_mts_timescales(const orc_mech *m, double T, const double *c, double *tci);
int orc_mts_cell(const orc_mech *m, const ohc_opts *o, double pr, double tn, double *y,
                 int *Nm, int64_t *steps);

/* calculateSourceTerms (diagJ NULL) / calculateSourceTermsImp. */
void orc_chem_source(const orc_mech *m, const ohc_opts *o, long n, const double *T,
                     const double *p, const double *rho, const double *Y, long ldY, double *R,
                     long ldR, double *diagJ, long ldD, int withoutLastSpec);

/* The GPU path's calculateSourceTermsImp (Kim diagonal): R [ns] in FP64 and
 * the mass-action diagonal with the last species substituted [ns-1], as
 * double (fp32 = 0) or float (fp32 = 1, the Hybrid variant) into diag. */
void orc_chem_source_kim(const orc_mech *m, long n, const double *T, const double *p,
                         const double *rho, const double *Y, long ldY, double *R, long ldR,
                         void *diag, long ldD, int fp32);

/* thermoList::THa_p over n cells. */
void orc_T_from_h(const orc_mech *m, const ohc_opts *o, long n, const double *h,
                  const double *p, const double *Y, long ldY, double *T, int *iflag);

/* thermoList::TEa_p / TEa_rho over n cells (internal energy e [J/kg]). */
void orc_T_from_e(const orc_mech *m, const ohc_opts *o, long n, const double *e,
                  const double *p_or_rho, const double *Y, long ldY, double *T, int *iflag);

/* Mixture cp_p, cv_p [J/kg/K], ha_p, ea_p [J/kg] (any NULL). */
void orc_mix_thermo(const orc_mech *m, long n, const double *T, const double *p,
                    const double *Y, long ldY, double *cp, double *cv, double *ha, double *ea);

/* Mixture cp, ha, species ha, mu, kappa and mixture-averaged D_k (any NULL). */
void orc_mix_props(const orc_mech *m, long n, const double *T, const double *p,
                   const double *Y, long ldY, double *mu, double *kappa, double *Dk, long ldD,
                   double *cp, double *ha, double *hk);

#ifdef __cplusplus
}
#endif
#endif
