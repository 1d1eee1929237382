/*
 * lgt_su2.h -- C ABI of the MI355X SU(2) pure-gauge link-field engine.
 *
 * One engine = one lattice resident in HBM on the current HIP device, in one
 * of two precisions (FP64 links, or FP32 links with FP64 global sums), plus
 * one HIP stream on which every call runs.  All calls return an int status
 * (LGT_OK or a negative LGT_E* code; lgt_last_error() gives the message);
 * results are complete and copied to host memory when a call returns.  Host
 * arrays are plain C arrays; no call takes or returns device pointers.
 *
 * Field exchange uses the reference's memory order: VOL*4 quaternions, site
 * major then direction, each {c0,c1,c2,c3} with U = c0 + i(c1 s1+c2 s2+c3 s3);
 * site n = x3 + L3*(x2 + L2*(x1 + L1*x0)), x0 = t
 * (LGT_Code/include/gauge.h:21-24, LGT_Code/modules/admin/init.c:768-868).
 *
 * Each entry point cites the reference function whose role it takes.  The
 * reference keeps its state in globals (pu, neib, runParams, measParams,
 * LGT_Code/include/headers.h:345-413); here the same state lives in the
 * engine handle and is passed explicitly.  lgt_compat.h exports the
 * reference's own names on top of this ABI.
 */
#ifndef LGT_SU2_H
#define LGT_SU2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lgt_engine lgt_engine;

/* status codes */
#define LGT_OK 0
#define LGT_EINVAL (-1)   /* bad argument (extent, size, mode, loop list) */
#define LGT_ENOMEM (-2)   /* device or host allocation failed */
#define LGT_EDEVICE (-3)  /* HIP runtime error */

/* precision */
#define LGT_FP64 0        /* FP64 links and arithmetic */
#define LGT_FP32 1        /* FP32 links and arithmetic, FP64 global sums */

/* APE smearing modes (LGT_Code/modules/obs/smearing.c:74-168) */
#define LGT_APE_SPATIAL 0  /* smearing_APE_spatial: links 1..3, 4 spatial staples */
#define LGT_APE_ALL 1      /* smearing_APE_all: links 0..3, 6 staples */
#define LGT_APE_TEMPORAL 2 /* smearing_APE_temporal: link 0, 6 staples */

/* topological charge definitions (LGT_Code/modules/obs/topologicalcharge.c) */
#define LGT_Q_IMPROVED 0   /* topologicalCharge(): 1x1+2x2+3x3 clover, :164-199 */
#define LGT_Q_CLOVER 1     /* meas_topologicalcharge(): 1x1 clover, :391-409 */

/* engine options (lgt_set_option) */
#define LGT_OPT_TILE_ORDER 0 /* order of the 256-site tiles: -1 (default) = tuned;
                                else flags 1 = each XCD takes a contiguous share of the
                                tiles, 2 = time coordinate fastest, 4 = time fastest within
                                groups of spatial tiles (2 and 4 apply when a time slice
                                holds whole tiles), 8 = the resulting order reversed.
                                Results do not depend on it. */
#define LGT_OPT_PROFILE 1    /* 1: time every launch with HIP events */
#define LGT_OPT_GROUP 2      /* spatial tiles per group of the grouped tile order: 0 (default)
                                = tuned for the lattice; else a divisor of the tiles in one
                                time slice.  Results do not depend on it. */
#define LGT_OPT_ALTERNATE 3  /* 1 (default): consecutive update, smearing and adjoint
                                launches walk the tiles in opposite orders, so each starts
                                on what the previous one touched last; 0: every launch in
                                the same order.  Results do not depend on it. */

/* kernel classes of the launch profile (lgt_profile_read) */
#define LGT_K_HEATBATH 0
#define LGT_K_OVERRELAX 1
#define LGT_K_PROJECT 2
#define LGT_K_PLAQUETTE 3
#define LGT_K_APE 4
#define LGT_K_WLOOP 5
#define LGT_K_QTOP 6
#define LGT_K_REDUCE 7
#define LGT_K_FIELD 8      /* cold/hot start, layout conversion */
#define LGT_K_APE_SAVE 9   /* smearing steps recomputed by the adjoint */
#define LGT_K_VJP 10       /* adjoint steps of the smearing */
#define LGT_K_WLOOP_GRAD 11
#define LGT_K_METROPOLIS 12
#define LGT_K_COUNT 13

/* ---- life cycle ---------------------------------------------------------- */

/* Allocates the field on the current device.  extents[0] = time; every extent
 * even and >= 2, extents[3] >= 4 (LGT_Code/include/headers.h:58-68; here
 * runtime values instead of compile-time LENGT..LENGS3).  The seed keys the
 * engine's counter-based RNG.  The field starts cold.
 * Replaces initArrayOfNeighbours + initGaugeField's allocation
 * (LGT_Code/modules/admin/init.c:741-869). */
int lgt_create(const int extents[4], int precision, uint64_t seed, lgt_engine **out);
/* releaseGaugeField (LGT_Code/modules/admin/init.c:922-925) */
int lgt_destroy(lgt_engine *e);

const char *lgt_strerror(int status);
/* message of the last failed call on this host thread */
const char *lgt_last_error(void);

int lgt_set_option(lgt_engine *e, int option, int value);
int lgt_get_extents(const lgt_engine *e, int extents[4], int *precision);
/* bytes of device memory held by the engine */
int64_t lgt_device_bytes(const lgt_engine *e);

/* ---- field --------------------------------------------------------------- */

/* initGaugeField(flag) (LGT_Code/modules/admin/init.c:831-869): 0 = cold
 * (all links 1), 1 = hot.  The hot start is Haar-uniform on S^3 (the
 * reference's su2RandomMatrix, random_su3.c:141-155, is not), drawn from the
 * engine's RNG in FP64 and rounded to the storage precision. */
int lgt_init_field(lgt_engine *e, int start);
/* Field exchange in the reference layout (pu, LGT_Code/include/headers.h:402);
 * FP32 engines round on upload and widen on download.  The conversion of
 * readConfig/writeConfig (LGT_Code/modules/io/config_IO.c:40-117,537-621)
 * between file and pu. */
int lgt_upload(lgt_engine *e, const double *aos);
int lgt_download(lgt_engine *e, double *aos);
/* lgt_upload, and plaq = the plaquette of the links as given (FP64, before an
 * FP32 engine rounds them): the value readConfig tests against the file's
 * (LGT_Code/modules/io/config_IO.c:609-614, plaquette.c:67-101), which an
 * FP32 engine's lgt_plaquette of the rounded links would miss by ~1e-10. */
int lgt_upload_plaquette(lgt_engine *e, const double *aos, double *plaq);

/* writeConfig / readConfig (LGT_Code/modules/io/config_IO.c:40-117,537-621):
 * the reference's binary configuration file, little-endian: int32 DIM = 4,
 * int32 SUN = 2, int32 extents[4] (t, x, y, z), double plaquette, then per
 * site (z fastest) and direction the link's c0..c3 as doubles.
 * lgt_write_config stores the engine's field with its plaquette.
 * lgt_read_config checks DIM, SUN and the extents (LGT_EINVAL otherwise),
 * loads the field and recomputes the plaquette of the links as read (FP64,
 * also in FP32 engines, see lgt_upload_plaquette); it returns LGT_EINVAL
 * ("plaquette test failed", the field stays loaded) when it differs from the
 * file's by more than sqrt(6 VOL) DBL_EPSILON, the reference's test.
 * plaq_file / plaq_calc (nullable) return the two values. */
int lgt_write_config(lgt_engine *e, const char *path);
int lgt_read_config(lgt_engine *e, const char *path, double *plaq_file, double *plaq_calc);
/* readConfigForThermalisation (LGT_Code/modules/io/config_IO.c:348-535): loads
 * a configuration file of half the engine's extents into the 16 sublattices
 * of the engine's lattice, mirrored as the reference does (in a mirrored
 * direction j a link along j moves from x to L_j - 1 - x and is daggered, a
 * link across j from x to L_j - x, x = 0 to L_j / 2), then the plaquette test
 * of lgt_read_config against the file's plaquette.  Extents must be even. */
int lgt_read_config_doubled(lgt_engine *e, const char *path, double *plaq_file, double *plaq_calc);

/* ---- update -------------------------------------------------------------- */

/* gaugefieldUpdate(iup, nup, utype=1, stype) (LGT_Code/modules/update/
 * update.c:102-176): nsweeps sweeps at coupling beta.  Each sweep is one
 * checkerboard Kennedy-Pendleton heatbath sweep (sub-sweeps mu = 0..3,
 * parity 0,1, heatbath.c:76-149 per link) followed by n_or overrelaxation
 * sweeps in the same order (U' = V^+ U^+ V^+, V = A/|A|, links re-projected
 * as they are stored).  With n_or == 0 the whole field is projected after
 * every sweep (update.c:166); FP32 engines instead store each heatbath link
 * projected, so later sub-sweeps of the same sweep see projected links
 * (a difference at FP32 rounding level).  Sweep k draws its uniforms from Philox4x32-10
 * keyed by the seed with counter (cb site, mu|parity<<2, sweep index, draw).
 * acceptance (nullable) = mean over links and sweeps of 1/trials, as the
 * reference logs it (update.c:159,170-173). */
int lgt_update(lgt_engine *e, int nsweeps, int n_or, double beta, double *acceptance);
/* gaugefieldUpdate(iup, nup, utype=0, stype) (update.c:102-176 with
 * localMetropolisUpdate, LGT_Code/modules/update/metro.c:65-109): nsweeps
 * checkerboard sweeps (sub-sweeps mu = 0..3, parity 0,1) of one Metropolis hit
 * per link with the proposal expx(eps, X) U, X_i = 1 - 2 r_i (exp_fct.c:137-151;
 * eps = runParams.eps), then one projection of the field (update.c:166).
 * Uniforms: Philox4x32-10, counter (cb site, mu|parity<<2|2<<3, sweep index,
 * draw), 4 per link.  acceptance (nullable) = accepted fraction. */
int lgt_metropolis_update(lgt_engine *e, int nsweeps, double beta, double eps, double *acceptance);
/* single sub-sweeps (no projection), for replays of a sweep */
int lgt_heatbath_subsweep(lgt_engine *e, int mu, int parity, double beta, double *acc_sum);
int lgt_overrelax_subsweep(lgt_engine *e, int mu, int parity);
/* project_gfield_to_sun (LGT_Code/modules/admin/sun_vfunc.c:82-97) */
int lgt_project(lgt_engine *e);

/* ---- single links, for programs that update link by link ----------------
 * site is the lexicographic index n = x3 + L3*(x2 + L2*(x1 + L1*x0)).
 * staples(n, dir, &A) (LGT_Code/modules/update/metro.c:34-63): A = the
 * staple of link (site, mu) of the current field, in the reference's term
 * order (FP64 engines bit-identical; FP32 engines widen their FP32 sum). */
int lgt_staple(lgt_engine *e, int64_t site, int mu, double A[4]);
/* localHeatbathUpdate(n, dir, m) (LGT_Code/modules/update/heatbath.c:124-149)
 * fed with injected uniforms: the staple of link (site, mu), then the
 * Kennedy-Pendleton draw (FP64, heatbath.c:76-121) or, when |A| <=
 * DBL_EPSILON, su2RandomMatrix (random_su3.c:141-155), consuming
 * uniforms[0 .. n_uniforms) in the reference's order (3 on the degenerate
 * branch; else 4 per trial, then 2).  If they suffice, the new link
 * U = X V^+ replaces the link on the device (rounded to the storage
 * precision), link (nullable) receives it, *trials = the trial count (0: the
 * degenerate branch) and *needed = the uniforms consumed.  If they run out,
 * nothing changes, *trials = -1 and *needed = how many the next call needs
 * at least (those read so far plus the next trial, direction or degenerate
 * link): a caller that draws exactly that many more from its stream
 * reproduces the reference's draws. */
int lgt_local_heatbath(lgt_engine *e, int64_t site, int mu, double beta, const double *uniforms, int n_uniforms,
                       double link[4], int *trials, int *needed);
/* No reference counterpart: the random gauge transformation
 * U'_mu(x) = g(x) U_mu(x) g(x+mu)^+ of the field, in place, with g(x)
 * Haar-uniform from the engine's counter-based RNG under `seed` (computed in
 * FP64, links rounded to the storage precision).  The plaquette, Wilson loops
 * and charges are invariant under it and APE smearing is covariant: a check
 * that costs one pass over the field at any lattice size.  g_out (nullable)
 * receives g(x) for every site in lexicographic order (VOL quaternions). */
int lgt_gauge_transform(lgt_engine *e, uint64_t seed, double *g_out);
/* sweep counter of the RNG (advanced by lgt_update; set to resume a chain) */
int lgt_set_sweep_index(lgt_engine *e, uint32_t sweep);
uint32_t lgt_get_sweep_index(const lgt_engine *e);

/* ---- observables ---------------------------------------------------------- */

/* plaquette() (LGT_Code/modules/obs/plaquette.c:67-101): mean of
 * (1/2) Re tr U_p over the 6 VOL plaquettes, FP64 sum. */
int lgt_plaquette(lgt_engine *e, double *plaq);
/* gaugeAction() (plaquette.c:104-110): beta * 6 VOL * (1 - plaq) */
int lgt_gauge_action(lgt_engine *e, double beta, double *action);

/* smearing_APE_spatial / _all / _temporal(ns, sp, u) (LGT_Code/modules/obs/
 * smearing.c:74-168): nsteps Jacobi steps U <- proj[U + sp * staples].
 * sp is the reference's raw parameter; the usual alpha form
 * U <- proj[(1-alpha) U + alpha/n S] is sp = alpha / (n (1 - alpha)) with
 * n = 4 (spatial) or 6 (all, temporal). */
int lgt_ape(lgt_engine *e, int mode, int nsteps, double sp);

/* measureWilsonLoop(w) (LGT_Code/modules/obs/wilson.c:268-324) with the loop
 * list given explicitly (measParams.rExtents/tExtents, wilson.c:274-285):
 * w[i] = mean over sites x and spatial d of (1/2) Re tr W_{R[i] x T[i]}(x; t, d).
 * Requires 1 <= T[i] <= extents[0]-1 (wilson.c:290-291) and 1 <= R[i] <= 4096
 * (a spatial line longer than the lattice wraps around it).
 * Device memory: the smearing buffer (one field) holds the spatial lines; a
 * list with more than one R also allocates a second line field (three
 * directions), kept for later calls. */
int lgt_wilson_loops(lgt_engine *e, int nloop, const int *R, const int *T, double *w);

/* topologicalCharge() / meas_topologicalcharge(q) (LGT_Code/modules/obs/
 * topologicalcharge.c:164-199, 391-409).  q: the charge; abs_sum (nullable):
 * the same sum over |q(x)|, the scale of its rounding error; density
 * (nullable): q(x) per site in lexicographic order, VOL doubles. */
int lgt_qtop(lgt_engine *e, int kind, double *q, double *abs_sum, double *density);

/* ---- adjoint of APE smearing (no reference counterpart) -------------------
 * A forward/backward pair for the vector-Jacobian product of N APE steps of
 * a mode (LGT_APE_SPATIAL, LGT_APE_ALL, LGT_APE_TEMPORAL: smearing_APE_*,
 * LGT_Code/modules/obs/smearing.c:74-168).
 *
 * lgt_ape_vjp_forward smears the current field like lgt_ape(mode, nsteps,
 * sp) and keeps a checkpoint of the links the mode changes every `segment`
 * steps (0: ceil(sqrt(nsteps))), and every step of the last segment.
 * lgt_ape_vjp_backward then maps a cotangent dL/dU of the smeared field to
 * dL/dU of the thin field, recomputing each lower segment's steps from its
 * checkpoint (storing every step's |U + sp S|) and running the adjoint steps
 * backwards through it.  segment = nsteps keeps every step's output, so
 * nothing is recomputed (device memory: nsteps fields of the changed links
 * and their norms).  Cotangents and gradients
 * have 4 real components per link, Euclidean (<a,b> = a0 b0 + a.b), in the
 * field layout of lgt_upload; links the mode does not smear keep their
 * cotangent and add what they contribute to the smeared links' staples (the
 * temporal links of spatial smearing, in no staple, pass it through).
 * cotangent may be NULL to use the one left on the device by lgt_wloop_grad.
 * The smeared field stays current; it must not be changed between the two
 * calls. */
int lgt_ape_vjp_forward(lgt_engine *e, int mode, int nsteps, double sp, int segment);
int lgt_ape_vjp_backward(lgt_engine *e, const double *cotangent, double *grad);
/* The R x T loops of measureWilsonLoop (LGT_Code/modules/obs/wilson.c:21-56)
 * summed without normalisation, L = sum_x sum_{d=1..3} 2 Re[W(x; t, d)]_0,
 * returned as an unevaluated double-double sum[0] + sum[1].  R and T as in
 * lgt_wilson_loops. */
int lgt_wloop_sum(lgt_engine *e, int R, int T, double sum[2]);
/* dL/dU of that L for every link of the current field; kept on the device as
 * the pending cotangent of lgt_ape_vjp_backward, and copied to grad (field
 * layout) unless grad is NULL. */
int lgt_wloop_grad(lgt_engine *e, int R, int T, double *grad);

/* ---- per-link harnesses (injected inputs, FP64) ---------------------------
 * The per-link heatbath of localHeatbathUpdate after the staple
 * (LGT_Code/modules/update/heatbath.c:133-146), fed with given uniforms: link i
 * has staple staples[4i..4i+3] and consumes uniforms[offsets[i] ..
 * offsets[i]+lens[i]) in the reference's order (4 per trial, then 2; 3 on the
 * degenerate branch).  counts[i] = trials (0: degenerate, -1: ran out). */
int lgt_heatbath_harness(int n, const double *staples, double beta, const double *uniforms,
                         const int *offsets, const int *lens, double *links, int *counts);
/* The overrelaxation map on given staples and links (no reference
 * counterpart; the update above adds it to the heatbath). */
int lgt_overrelax_harness(int n, const double *staples, const double *links_in, double *links_out);
/* One Metropolis hit (metro.c:80-105 after the staple) on given staples and
 * links with the uniforms uniforms[4i..4i+3] (3 for the proposal, 1 for the
 * acceptance, the reference's order); accepted[i] = 0 or 1. */
int lgt_metropolis_harness(int n, const double *staples, const double *links_in, double beta, double eps,
                           const double *uniforms, double *links_out, int *accepted);

/* ---- allocation-free plaquette --------------------------------------------
 * No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the plaquette of lgt_upload_plaquette (plaquette.c:67-101 on an
 * FP64 field in the reference layout) with the field already in device memory
 * and the result written to device memory, enqueued on `stream` (a
 * hipStream_t; NULL: the null stream) without any allocation, and no engine.
 * field_dev: VOL*16 doubles (site major, then direction, c0..c3); plaq_dev:
 * one double; both device pointers owned by the caller, and the only
 * exception to "no device pointers" in this ABI.  extents as lgt_create, with
 * VOL <= LGT_NOALLOC_MAX_VOL (LGT_EINVAL otherwise).  The partial sums live in
 * a static device array of (LGT_NOALLOC_MAX_VOL/256 + 1) * 16 bytes = 5.3 MB
 * that every device the library runs on keeps until the process ends, shared
 * by all calls on that device: calls that overlap on one device (other
 * streams, other host threads) race, so serialise them.  Returns after
 * enqueueing; the result is complete when `stream` reaches it, and equals
 * lgt_upload_plaquette's for the same links bit for bit. */
#define LGT_NOALLOC_MAX_VOL (96LL * 96 * 96 * 96)
int lgt_plaquette_aos_noalloc(const int extents[4], const double *field_dev, double *plaq_dev, void *stream);

/* ---- timing --------------------------------------------------------------- */
/* With LGT_OPT_PROFILE on, every launch is bracketed by HIP events on the
 * engine's stream; this returns the summed kernel time and launch count of
 * one kernel class since the last reset. */
int lgt_profile_read(lgt_engine *e, int kernel_class, double *ms_total, int64_t *launches);
int lgt_profile_reset(lgt_engine *e);
/* waits for the engine's stream */
int lgt_sync(lgt_engine *e);

#ifdef __cplusplus
}
#endif
#endif
