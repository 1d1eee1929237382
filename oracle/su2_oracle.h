/*
 * su2_oracle.h -- FP64 CPU restatement of the SU(2) pure-gauge link-field path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product path (masterarbeit_amd/) never links or calls it.
 *
 * Every function restates the reference algorithm in the reference's own
 * floating-point operation order (compiled with -ffp-contract=off), and cites
 * the reference file:line it follows (paths relative to the reference's
 * LGT_Code/ directory).  Field layout is the reference's: one contiguous
 * array of VOL*4 quaternions, site-major then direction, each quaternion
 * {c0,c1,c2,c3} with U = c0 + i(c1 s1 + c2 s2 + c3 s3)   (include/gauge.h:21-24,
 * modules/admin/init.c:842-868).  Site index n = x3 + L3*(x2 + L2*(x1 + L1*x0)),
 * x0 = t  (modules/admin/init.c:768-811).
 *
 * Functions without a reference counterpart (Philox, checkerboard sub-sweeps,
 * overrelaxation, Haar start, the APE adjoint) state what pins them in
 * tests/ next to their declaration.
 */
#ifndef SU2_ORACLE_H
#define SU2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double c0, c1, c2, c3; } oq_t;

typedef struct {
    int L[4];          /* extents, L[0] = time */
    int64_t vol;
    int64_t *nb;       /* [vol][8]: +mu at [mu], -mu at [mu+4] */
} olat_t;

/* ---- lattice ---- */
int  or_lat_init(olat_t *lat, const int L[4]);
void or_lat_free(olat_t *lat);
int64_t or_lat_vol(const int L[4]);

/* ---- quaternion algebra (include/gauge.h:57-178) ---- */
void or_qmul(const double *v, const double *w, double *out);      /* v*w     */
void or_qmul_dag(const double *v, const double *w, double *out);  /* v*w^+   */

/* ---- uniform sources ---- */
/* Buffer source: returns buf[pos++]; records an error when exhausted. */
typedef struct {
    const double *buf;
    int64_t n, pos;
    int exhausted;
} or_ubuf_t;

/* ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- Kennedy-Pendleton heatbath on one link (modules/update/heatbath.c) ---- */
/* Draws X from p(X) ~ exp(alpha x0) on S^3 with uniforms from ub; returns the
 * number of trials (heatbath.c:76-121). */
int  or_kp_draw(double alpha, or_ubuf_t *ub, double X[4]);
/* Reference localHeatbathUpdate for a given staple A (heatbath.c:124-149):
 * writes the new link, returns 1/count (0 on the degenerate branch). */
double or_heatbath_link(const double A[4], double beta, or_ubuf_t *ub, double Unew[4]);
/* Reference su2RandomMatrix (modules/random/random_su3.c:141-155). */
void or_su2_random_matrix(or_ubuf_t *ub, double U[4]);
/* Batch form for the per-link harness: n links, staples[n][4], uniforms
 * consumed sequentially, outputs links[n][4] and trial counts. */
int  or_heatbath_batch(int64_t n, const double *staples, double beta,
                       const double *uniforms, int64_t nuni,
                       double *links, int *counts, int64_t *consumed);

/* ---- staples / sweeps ---- */
void or_staple(const olat_t *lat, const double *U, int64_t n, int mu, double A[4]);   /* metro.c:34-63 */
void or_project_link(double *u);                                                   /* sun_vfunc.c:69-79 */
void or_project_field(const olat_t *lat, double *U);                               /* sun_vfunc.c:82-97 */
/* Lexicographic heatbath sweep with uniforms from a buffer, followed by a
 * projection of the whole field (update.c:102-176 with stype=0, utype=1).
 * Returns the mean acceptance 1/count like the reference; -1 on buffer
 * exhaustion. */
double or_hb_sweep_lex(const olat_t *lat, double *U, double beta, const double *uniforms,
                       int64_t nuni, int64_t *consumed);

/* Checkerboard heatbath sub-sweep of direction mu, parity p, driven by the
 * engine's Philox streams (key = seed, counter = (cb site, mu|p<<2|kind<<3,
 * sweep, draw)).  fp64 uniforms.  Returns the sum of 1/count. */
double or_hb_subsweep_philox(const olat_t *lat, double *U, int mu, int parity, double beta,
                             uint64_t seed, uint32_t sweep);
/* Overrelaxation sub-sweep: U' = V^+ U^+ V^+ with V = A/|A|, followed by the
 * link projection.  Leaves links with |A| <= DBL_EPSILON unchanged. */
void or_or_subsweep(const olat_t *lat, double *U, int mu, int parity);
void or_or_link(const double A[4], const double U[4], double Unew[4]);
/* Full engine sweep: HB over (mu,p) in order mu=0..3, p=0,1, then n_or OR
 * sweeps in the same order; with n_or == 0 a projection pass follows. */
double or_engine_sweep(const olat_t *lat, double *U, double beta, int n_or, uint64_t seed,
                       uint32_t sweep);
/* Metropolis with one hit (metro.c:65-109, proposal expx of exp_fct.c:137-151):
 * X_i = 1 - 2 r_i (3 uniforms), U' = expx(eps, X) U = [((1-z) c, c eps X)] U with
 * z = eps^2 |X|^2 / 4, c = 1/(1+z); accepted iff r_4 <= exp(beta (Re tr U'A -
 * Re tr UA) / 2).  Writes the new (or old) link, returns 1 if accepted. */
int or_metropolis_link(const double A[4], double beta, double eps, or_ubuf_t *ub, const double U[4],
                       double Unew[4]);
/* Batch form for the per-link harness: 4 uniforms per link. */
int or_metropolis_batch(int64_t n, const double *staples, const double *links_in, double beta, double eps,
                        const double *uniforms, double *links, int *accepted);
/* Checkerboard Metropolis sub-sweep driven by the engine's Philox streams
 * (kind 2); returns the number of accepted links. */
int64_t or_metro_subsweep_philox(const olat_t *lat, double *U, int mu, int parity, double beta, double eps,
                                 uint64_t seed, uint32_t sweep);
/* Engine Metropolis update: nsweeps x (mu = 0..3, p = 0,1) sub-sweeps from
 * sweep index sweep0, then one projection of the field (update.c:166);
 * returns the mean acceptance. */
double or_metropolis_update(const olat_t *lat, double *U, int nsweeps, double beta, double eps, uint64_t seed,
                            uint32_t sweep0);
/* Engine hot start: Haar-uniform links from Philox (Shoemake's method). */
void or_hot_start(const olat_t *lat, double *U, uint64_t seed);
void or_cold_start(const olat_t *lat, double *U);
/* uniform helpers shared with the engine's definition */
double or_u53(uint32_t hi, uint32_t lo);

/* ---- observables ---- */
double or_plaquette(const olat_t *lat, const double *U);            /* plaquette.c:67-101 (reference order) */
double or_plaquette_acc(const olat_t *lat, const double *U);        /* same terms, compensated sum */
double or_plaquette_site(const olat_t *lat, const double *U, int64_t n);

/* APE smearing, Jacobi form U <- proj(U + sp*S(zu)); mode 0 = spatial
 * (smearing.c:107-138), 1 = all four directions (smearing.c:74-105),
 * 2 = temporal (smearing.c:140-168). */
void or_ape(const olat_t *lat, double *U, int mode, int nsteps, double sp);
void or_ape_staple(const olat_t *lat, const double *Z, int64_t n, int mu, int dt, double S[4]);

/* Wilson loops (wilson.c:21-56, 268-324).  w[i] for (R[i],T[i]).  The
 * reference sequential sum (seq=1) or a compensated sum (seq=0). */
double or_wloop_site(const olat_t *lat, const double *U, int64_t n, int d1, int d2, int l1, int l2);
int  or_wilson_loops(const olat_t *lat, const double *U, int nloop, const int *R, const int *T,
                     int seq, double *w);

/* Topological charge: improved clover (topologicalcharge.c:164-199) and the
 * 1x1 clover (topologicalcharge.c:391-409).  seq=1 reference sequential sum,
 * seq=0 compensated; the density q(x) may be written to qx (nullable). */
double or_qtop_improved(const olat_t *lat, const double *U, int seq, double *qx, double *abs_sum);
double or_qtop_clover(const olat_t *lat, const double *U, int seq, double *qx, double *abs_sum);

/* ---- adjoint of APE smearing ----
 * L = sum_x sum_{d=1..3} 2*Re[ W_{R x T}(x; 0, d) ]_0 on the field smeared
 * by nsteps APE steps of the given mode (0 spatial, 1 all, 2 temporal, as
 * or_ape).  Writes dL/dU (4 real components per link, Euclidean gradient)
 * for the thin field into G (same layout as U); returns L. */
double or_wloop_objective(const olat_t *lat, const double *U, int R, int T);
double or_ape_vjp(const olat_t *lat, const double *U, int mode, int nsteps, double sp, int R, int T,
                  double *G);

#ifdef __cplusplus
}
#endif
#endif
