/*
 * lgt_compat.c -- the reference's entry points (include/lgt_compat.h) on top of
 * the GPU engine (include/lgt_su2.h).  C host code.
 *
 * Coherence of the program's host field pu with the device field:
 *   eager (default): every call uploads the field it reads and downloads the
 *     field it changes, so pu is always current;
 *   lazy: the device field is authoritative between calls.  An entry point
 *     uploads pu only when the host copy changed since the last transfer and
 *     downloads nothing; the field-lifecycle functions of the program
 *     (initGaugeField, readConfig, readConfigForThermalisation, writeConfig,
 *     copyGaugeField, releaseGaugeField) are intercepted with the linker's
 *     --wrap (lgt_compat_wrap.c, a static archive linked into the program)
 *     and download before the program reads pu (writeConfig, a copy out of
 *     pu) or mark the host copy changed after it writes pu.  Lazy mode
 *     starts with the first intercepted call (a program linked with --wrap
 *     always calls initGaugeField first) or with LGT_COMPAT_SYNC=lazy;
 *     LGT_COMPAT_SYNC=eager keeps eager mode.
 */
#include "../../include/lgt_compat.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lgt_su2.h"

static lgt_engine *g_eng = NULL;
static int g_L[4];
static int64_t g_vol = 0;
static double *g_buf = NULL;
static int g_nor = 0;
static int g_prec = LGT_FP64;
static double *g_uni = NULL;     /* uniforms of localHeatbathUpdate */
static int g_uni_cap = 0;
/* FP64 plaquette of pu's links at their last upload, while the device field
 * still holds exactly those links */
static int g_plaq_valid = 0;
static double g_plaq = 0.0;

/* coherence state */
static int g_mode = -1;          /* -1: not chosen yet, 0: eager, 1: lazy */
static int g_host_dirty = 1;     /* pu changed since the last transfer */
static int g_dev_dirty = 0;      /* device field newer than pu */
static long long g_uploads = 0, g_downloads = 0;

static char NAME[] = "lgt_compat [lgt_compat.c]";

static void check(int rc, const char *what)
{
    if (rc != LGT_OK) {
        char fmt[] = "%s failed: %s (%s)";
        error(1, NAME, fmt, what, lgt_strerror(rc), lgt_last_error());
    }
}

static int lazy(void)
{
    if (g_mode < 0) {
        const char *m = getenv("LGT_COMPAT_SYNC");
        g_mode = (m && !strcmp(m, "lazy")) ? 1 : 0;
    }
    return g_mode == 1;
}

void lgt_compat_lazy(void)
{
    if (g_mode < 0) {
        const char *m = getenv("LGT_COMPAT_SYNC");
        g_mode = (m && !strcmp(m, "eager")) ? 0 : 1;
    }
}

void lgt_compat_release(void)
{
    if (g_eng) lgt_destroy(g_eng);
    g_eng = NULL;
    free(g_buf);
    g_buf = NULL;
    free(g_uni);
    g_uni = NULL;
    g_uni_cap = 0;
    g_plaq_valid = 0;
    g_host_dirty = 1;
    g_dev_dirty = 0;
}

static void engine(void)
{
    if (g_eng) return;
    g_vol = 1;
    for (int d = 0; d < 4; d++) {
        g_L[d] = latParams.linearExtent[d];
        g_vol *= g_L[d];
    }
    if (g_vol <= 0) {
        char fmt[] = "lattice extents unset: call initArrayOfNeighbours() first";
        error(1, NAME, fmt);
    }
    int prec = LGT_FP64;
    const char *p = getenv("LGT_PRECISION");
    if (p && (!strcmp(p, "fp32") || !strcmp(p, "FP32"))) prec = LGT_FP32;
    g_prec = prec;
    const char *n = getenv("LGT_N_OR");
    g_nor = n ? atoi(n) : 0;
    if (g_nor < 0) g_nor = 0;
    /* seed from the program's own RANLUX stream (rlxd_init(2, seed) in main) */
    double r[2];
    ranlxd(r, 2);
    uint64_t seed = ((uint64_t)(r[0] * 4294967296.0) << 32) | (uint64_t)(r[1] * 4294967296.0);
    check(lgt_create(g_L, prec, seed, &g_eng), "lgt_create");
    g_buf = (double *)malloc(sizeof(double) * 16 * (size_t)g_vol);
    if (!g_buf) {
        char fmt[] = "cannot allocate the %lld-site transfer buffer";
        error(1, NAME, fmt, (long long)g_vol);
    }
    g_host_dirty = 1;
    g_dev_dirty = 0;
    atexit(lgt_compat_release);
}

/* The reference's field table points into one allocation, link (s, d) at
 * base + 4 s + d (init.c:842-869); such a field is handed to the engine in
 * place.  Returns its base, or NULL when the table points anywhere else (the
 * links are then packed through the transfer buffer). */
static double *in_place(su2mat *u[][4])
{
    double *base = (double *)u[0][0];
    for (int64_t s = 0; s < g_vol; s++)
        for (int d = 0; d < 4; d++)
            if ((double *)u[s][d] != base + 16 * s + 4 * d) return NULL;
    return base;
}

/* uploads u; plaq (nullable) receives the plaquette of u's links as they
 * are, in FP64 (lgt_upload_plaquette) */
static void upload_plaq(su2mat *u[][4], double *plaq)
{
    engine();
    double *f = in_place(u);
    if (!f) {
        for (int64_t s = 0; s < g_vol; s++)
            for (int d = 0; d < 4; d++) memcpy(g_buf + 16 * s + 4 * d, u[s][d], 4 * sizeof(double));
        f = g_buf;
    }
    if (plaq) {
        check(lgt_upload_plaquette(g_eng, f, plaq), "lgt_upload_plaquette");
        g_plaq = *plaq;
        g_plaq_valid = (u == pu);
    } else {
        check(lgt_upload(g_eng, f), "lgt_upload");
        g_plaq_valid = 0;
    }
    g_uploads++;
}

static void upload(su2mat *u[][4]) { upload_plaq(u, NULL); }

static void download(su2mat *u[][4])
{
    double *f = in_place(u);
    check(lgt_download(g_eng, f ? f : g_buf), "lgt_download");
    if (!f)
        for (int64_t s = 0; s < g_vol; s++)
            for (int d = 0; d < 4; d++) memcpy(u[s][d], g_buf + 16 * s + 4 * d, 4 * sizeof(double));
    g_downloads++;
}

/* the device holds pu's current links */
static void pu_on_device(void)
{
    engine();
    if (!lazy() || g_host_dirty) {
        upload(pu);
        g_host_dirty = 0;
    }
}

/* pu holds the device's current links */
static void pu_on_host(void)
{
    if (g_eng && g_dev_dirty) {
        download(pu);
        g_dev_dirty = 0;
    }
}

/* the device field (pu's links) was changed on the device */
static void device_changed(void)
{
    g_plaq_valid = 0;
    if (lazy()) {
        g_dev_dirty = 1;
    } else {
        download(pu);
    }
}

/* a field other than pu goes through the device: pu's device copy is lost */
static void run_on_other_field(su2mat *u[][4], int (*op)(int, double), int a, double b)
{
    pu_on_host();
    upload(u);
    check(op(a, b), "lgt call");
    download(u);
    g_host_dirty = 1;
}

void lgt_compat_host_modified(void)
{
    g_plaq_valid = 0;
    g_host_dirty = 1;
    g_dev_dirty = 0;
}

void lgt_compat_sync_host(void) { pu_on_host(); }

void lgt_compat_transfers(long long *uploads, long long *downloads)
{
    if (uploads) *uploads = g_uploads;
    if (downloads) *downloads = g_downloads;
}

/* ---- the reference's entry points ---- */
void gaugefieldUpdate(int iup, int nup, int utype, int stype)
{
    char m1[] = "Unknown update type! Only Metropolis (SIM_TYPE=0) and heatbath (SIM_TYPE=1) implemented!";
    char m2[] = "Unknown sweep type! Only sequential (SWEEP_TYPE=0) and random (SWEEP_TYPE=1) sweeps implemented!";
    error(utype != 0 && utype != 1, NAME, m1);
    error(stype != 0 && stype != 1, NAME, m2);
    static int warned = 0;
    if (stype != 0 && !warned) {
        /* update.c:122-153 visits random sites in that mode; every GPU sweep
         * visits the links in checkerboard order (same equilibrium) */
        char w[] = "\nWarning: gaugefieldUpdate: sweep type %d (random site order) runs as the GPU's "
                   "checkerboard sweep\n";
        logging(w, stype);
        warned = 1;
    }
    double t1 = getTime();
    pu_on_device();
    double acc = 0.0;
    if (utype == 1)
        check(lgt_update(g_eng, nup, g_nor, runParams.beta, &acc), "lgt_update");
    else
        check(lgt_metropolis_update(g_eng, nup, runParams.beta, runParams.eps, &acc), "lgt_metropolis_update");
    device_changed();
    double t2 = getTime();
    char fmt[] = "\nUpdate %d done:\n(algorithm %d; sweep %d; acc=%.3f)\nupdate took %.9f seconds\n";
    logging(fmt, iup, utype, stype, acc, t2 - t1);
}

/* While the device field holds pu's links as the program last handed them
 * over, the plaquette is that of pu's FP64 links (what the reference computes,
 * and what readConfig tests against the file's, config_IO.c:609-614), also in
 * FP32 storage, where the plaquette of the rounded links would fail that
 * test; once the device has changed the field it is the device field's (FP64
 * arithmetic in both storages). */
double plaquette(void)
{
    double p = 0.0;
    engine();
    if (!lazy() || g_host_dirty) {
        upload_plaq(pu, &p);
        g_host_dirty = 0;
        return p;
    }
    if (g_plaq_valid) return g_plaq;
    check(lgt_plaquette(g_eng, &p), "lgt_plaquette");
    return p;
}

double gaugeAction(void)
{
    double p = plaquette();
    return runParams.beta * 6 * (double)g_vol * (1. - p);
}

static int ape_op(int mode_ns, double sp)
{
    return lgt_ape(g_eng, mode_ns & 3, mode_ns >> 2, sp);
}

static void ape(int mode, int ns, double sp, su2mat *u[][4])
{
    if (u == pu) {
        pu_on_device();
        check(lgt_ape(g_eng, mode, ns, sp), "lgt_ape");
        device_changed();
    } else {
        engine();
        run_on_other_field(u, ape_op, mode | (ns << 2), sp);
    }
}

void smearing_APE_all(int ns, double sp, su2mat *u[][4]) { ape(LGT_APE_ALL, ns, sp, u); }
void smearing_APE_spatial(int ns, double sp, su2mat *u[][4]) { ape(LGT_APE_SPATIAL, ns, sp, u); }
void smearing_APE_temporal(int ns, double sp, su2mat *u[][4]) { ape(LGT_APE_TEMPORAL, ns, sp, u); }

/* loop list as measureWilsonLoop builds it (wilson.c:274-285) */
void measureWilsonLoop(double *w)
{
    int nloop;
    if (measParams.mwil_mode <= 0) {
        int nt = ((measParams.tf - measParams.ts) / measParams.dt) + 1;
        int nr = ((measParams.rf - measParams.rs) / measParams.dr) + 1;
        nloop = nt * nr;
    } else {
        nloop = measParams.mwil_mode;
    }
    for (int i = 0; i < nloop; i++) w[i] = 0.;
    engine();
    char m[] = "Lattice too small to measure loops!";
    char nm[] = "meas_wilson";
    error(latParams.linearExtent[0] - 1 < measParams.tf, nm, m);
    pu_on_device();
    check(lgt_wilson_loops(g_eng, nloop, measParams.rExtents, measParams.tExtents, w), "lgt_wilson_loops");
}

double topologicalCharge(void)
{
    char m1[] = "calculating topological charge ..\n";
    char m2[] = "measurement done (took %.3e sec)\n";
    double t1 = getTime(), q = 0.0;
    logging(m1);
    pu_on_device();
    check(lgt_qtop(g_eng, LGT_Q_IMPROVED, &q, NULL, NULL), "lgt_qtop");
    logging(m2, getTime() - t1);
    return q;
}

void meas_topologicalcharge(double *q)
{
    pu_on_device();
    check(lgt_qtop(g_eng, LGT_Q_CLOVER, q, NULL, NULL), "lgt_qtop");
}

/* staples(n, dir, &A) (metro.c:34-63) of pu's link (n, dir), on the device */
void staples(uint64_t n, int dir, su2mat *stap)
{
    char m[] = "staples: link (%llu, %d) outside the lattice";
    engine();
    error(n >= (uint64_t)g_vol || dir < 0 || dir > 3, NAME, m, (unsigned long long)n, dir);
    pu_on_device();
    double A[4];
    check(lgt_staple(g_eng, (int64_t)n, dir, A), "lgt_staple");
    stap->c0 = A[0]; stap->c1 = A[1]; stap->c2 = A[2]; stap->c3 = A[3];
}

/* localHeatbathUpdate(n, dir, m) (heatbath.c:124-149): the new link is drawn
 * on the device from uniforms of the program's ranlxd stream, drawn exactly
 * as the reference draws them (3 on the degenerate branch; else 4 per trial,
 * then 2), and written to the device field and to pu[n][dir].  Returns
 * 1/trials, 0 on the degenerate branch. */
double localHeatbathUpdate(uint64_t n, int dir, int m)
{
    (void)m;
    char m1[] = "localHeatbathUpdate: link (%llu, %d) outside the lattice";
    char m2[] = "localHeatbathUpdate: cannot allocate %d uniforms";
    engine();
    error(n >= (uint64_t)g_vol || dir < 0 || dir > 3, NAME, m1, (unsigned long long)n, dir);
    pu_on_device();
    int have = 0, trials = -1, needed = 0;
    double link[4];
    for (;;) {
        check(lgt_local_heatbath(g_eng, (int64_t)n, dir, runParams.beta, g_uni, have, link, &trials, &needed),
              "lgt_local_heatbath");
        if (trials >= 0) break;
        if (needed > g_uni_cap) {
            int cap = needed > 2 * g_uni_cap ? needed + 64 : 2 * g_uni_cap;
            double *u = (double *)realloc(g_uni, sizeof(double) * (size_t)cap);
            error(u == NULL, NAME, m2, cap);
            g_uni = u;
            g_uni_cap = cap;
        }
        ranlxd(g_uni + have, needed - have);
        have = needed;
    }
    g_plaq_valid = 0;
    /* pu[n][dir] now holds the device's link: eager mode stays current, and
     * lazy mode keeps its marks (the device field is current) */
    pu[n][dir]->c0 = link[0]; pu[n][dir]->c1 = link[1]; pu[n][dir]->c2 = link[2]; pu[n][dir]->c3 = link[3];
    return trials > 0 ? 1.0 / trials : 0.0;
}

static int project_op(int unused_a, double unused_b)
{
    (void)unused_a;
    (void)unused_b;
    return lgt_project(g_eng);
}

void project_gfield_to_sun(su2mat *u[][4])
{
    if (u == pu) {
        pu_on_device();
        check(lgt_project(g_eng), "lgt_project");
        device_changed();
    } else {
        engine();
        run_on_other_field(u, project_op, 0, 0.0);
    }
}
