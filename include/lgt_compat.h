/*
 * lgt_compat.h -- the reference's own C entry points for the SU(2) hot path,
 * exported by masterarbeit_amd/liblgt_compat.so on top of the engine
 * (lgt_su2.h).  A program built against the reference (LGT_Code/main/qcd.c,
 * main_meas/meas.c, main_qtop/qtop.c) links this library instead of the
 * reference's update.c, heatbath.c, metro.c, plaquette.c, smearing.c,
 * wilson.c and topologicalcharge.c and runs unchanged on one GPU.
 *
 * The reference keeps its state in globals defined by the main program
 * (LGT_Code/include/headers.h:376-413: pu, latParams, runParams, measParams)
 * and reports errors through error(), which logs and exits
 * (LGT_Code/modules/io/IO_utils.c:61-80).  This library reads those globals
 * and calls the program's logging(), error(), getTime() and ranlxd(), so the
 * declarations below repeat the reference's struct layouts member for member
 * (DIM = 4, SUN = 2); only their layout matters, not these names' origin.
 *
 * Coherence of pu with the device field:
 *  - eager (default): every entry point uploads the field it reads and
 *    downloads the field it changes, so pu is current after every call;
 *  - lazy: the device field is authoritative; entry points upload pu only when
 *    the program changed it and download nothing.  The program is linked with
 *    -Wl,--wrap=initGaugeField,--wrap=readConfig,--wrap=readConfigForThermalisation,
 *    --wrap=writeConfig,--wrap=copyGaugeField,--wrap=releaseGaugeField (and
 *    --wrap=measure2ptFunctions where the fermion modules are linked) so
 *    that its own field-lifecycle functions (modules.h:119-142) pass through
 *    the __wrap_ functions below, which download before pu is read and mark
 *    pu changed after it is written.  The first intercepted call selects lazy
 *    mode.  A program that writes pu by other means calls
 *    lgt_compat_host_modified(); one that reads pu by other means calls
 *    lgt_compat_sync_host() first.
 *
 * Environment: LGT_PRECISION=fp64|fp32 (default fp64), LGT_N_OR=<n> (number
 * of overrelaxation sweeps after each heatbath sweep, default 0 as in the
 * reference), LGT_COMPAT_SYNC=eager|lazy (overrides the mode).  The engine's
 * RNG seed is drawn from the program's own ranlxd stream when the engine is
 * created (first call).
 */
#ifndef LGT_COMPAT_H
#define LGT_COMPAT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- layouts of the reference's types (include/gauge.h:21-24, headers.h:345-374) ---- */
typedef struct { double c0, c1, c2, c3; } su2mat;

typedef struct {
    int idForOutputFilesName, numConfs, numThermConfs, decorSteps, writeConfsFreq, numSweeps;
    double beta, eps, mass;
    int mwil, mcorrs, mtopcharge;
} runParameters;

typedef struct {
    int linearExtent[4];
    int volumeOtherDirs[4];
    int volumeFactor[4];
} latticeParameters;

typedef struct {
    int imeas;
    int ts, tf, dt;
    int rs, rf, dr;
    int mwil_mode;
    int *tExtents, *rExtents;
    int ismt, isms, ism_tc;
    double smpart, smpars, smpar_tc;
    int nmax;
    double eps;
    int stype;
    int source[4];
    int n2pt;
    int *g1, *g2;
} measParameters;

/* ---- globals and services the host program provides ---- */
extern su2mat *pu[][4];                 /* headers.h:402 (sun_mat *pu[VOL][DIM]) */
extern runParameters runParams;         /* headers.h:382 */
extern latticeParameters latParams;     /* headers.h:383, set by initArrayOfNeighbours */
extern measParameters measParams;       /* headers.h:384 */
extern void logging(char *format, ...);                   /* IO_utils.c:47-59 */
extern void error(int test, char *name, char *format, ...); /* IO_utils.c:61-80 */
extern double getTime(void);                              /* utils.c:56-78 */
extern void ranlxd(double r[], int n);                    /* ranlxd.c:571-586 */

/* ---- exported entry points (LGT_Code/include/modules.h) ---- */
/* modules.h:99, update.c:102-176: utype 1 = heatbath (+ LGT_N_OR overrelaxation
 * sweeps), 0 = Metropolis with runParams.eps; the GPU sweep is checkerboard
 * whatever stype says (stype 1, random site order, logs a warning once). */
void gaugefieldUpdate(int iup, int nup, int utype, int stype);
/* modules.h:89, metro.c:34-63: the staple of pu's link (n, dir), computed on the device */
void staples(uint64_t n, int dir, su2mat *stap);
/* modules.h:95, heatbath.c:124-149: one link's heatbath on the device with
 * uniforms drawn from the program's ranlxd in the reference's pattern; the new
 * link goes to the device field and to pu[n][dir].  One device round trip per
 * trial: for whole sweeps gaugefieldUpdate is the fast path. */
double localHeatbathUpdate(uint64_t n, int dir, int m);
/* modules.h:27, plaquette.c:67-101; when pu is uploaded by the call, the FP64
 * plaquette of pu's links as given (so readConfig's test holds in FP32 storage) */
double plaquette(void);
double gaugeAction(void);                                 /* modules.h:28, plaquette.c:104-110 */
void smearing_APE_all(int ns, double sp, su2mat *u[][4]);      /* modules.h:32, smearing.c:74-105 */
void smearing_APE_spatial(int ns, double sp, su2mat *u[][4]);  /* modules.h:33, smearing.c:107-138 */
void smearing_APE_temporal(int ns, double sp, su2mat *u[][4]); /* modules.h:34, smearing.c:140-168 */
void measureWilsonLoop(double *w);                        /* modules.h:38, wilson.c:268-324 */
double topologicalCharge(void);                           /* modules.h:42, topologicalcharge.c:164-199 */
void meas_topologicalcharge(double *q);                   /* modules.h:43, topologicalcharge.c:391-409 */
void project_gfield_to_sun(su2mat *u[][4]);               /* modules.h:176, sun_vfunc.c:82-97 */

/* no reference counterpart: releases the engine (also registered with atexit) */
void lgt_compat_release(void);
/* no reference counterpart: pu was written by the program outside the
 * intercepted functions (lazy mode uploads it at the next entry point) */
void lgt_compat_host_modified(void);
/* no reference counterpart: make pu current (lazy mode downloads if needed) */
void lgt_compat_sync_host(void);
/* no reference counterpart: field transfers so far (host to device, back) */
void lgt_compat_transfers(long long *uploads, long long *downloads);

/* no reference counterpart: select lazy coherence (called by the
 * interceptors of liblgt_compat_wrap.a; LGT_COMPAT_SYNC=eager overrides) */
void lgt_compat_lazy(void);

/* interceptors of the program's field-lifecycle functions for lazy mode, in
 * liblgt_compat_wrap.a (link the program with it and -Wl,--wrap=<name>) */
void __wrap_initGaugeField(int start);
void __wrap_releaseGaugeField(void);
void __wrap_readConfig(char *name);
void __wrap_readConfigForThermalisation(char *name);
void __wrap_writeConfig(char *name);
void __wrap_copyGaugeField(su2mat *u1[][4], su2mat *u2[][4]);
void __wrap_measure2ptFunctions(int *source, int stype, int n2pt, int *g1, int *g2);

#ifdef __cplusplus
}
#endif
#endif
