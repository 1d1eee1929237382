/*
 * compat_driver.c -- a reference-style host program (the role of
 * LGT_Code/main/qcd.c and main_meas/meas.c) that owns the globals and the
 * logging/error/timer/RNG services and calls the reference's entry points,
 * which liblgt_compat.so serves from the GPU engine.  Test infrastructure.
 *
 * usage: compat_driver L0 L1 L2 L3 field.bin outdir [perlink]
 *
 * With "perlink" it runs the reference's sweep loop itself (update.c:122-166:
 * lexicographic sites, directions 0..3, localHeatbathUpdate per link, then
 * project_gfield_to_sun) through the per-link entry points, after the
 * staples of a few links, with ranlxd drawing from a 64-bit LCG that the test
 * reproduces.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../../include/lgt_compat.h"

#define MAXVOL (16 * 16 * 16 * 16)

/* the globals of a reference main program (headers.h:376-413 with MAIN_C) */
su2mat *pu[MAXVOL][4];
runParameters runParams;
latticeParameters latParams;
measParameters measParams;
static char LOG_FILE[512];

void logging(char *format, ...)
{
    FILE *f = fopen(LOG_FILE, "ab");
    va_list ap;
    va_start(ap, format);
    vfprintf(f, format, ap);
    va_end(ap);
    fclose(f);
}

void error(int test, char *name, char *format, ...)
{
    if (!test) return;
    va_list ap;
    fprintf(stderr, "\nError in %s:\n", name);
    va_start(ap, format);
    vfprintf(stderr, format, ap);
    va_end(ap);
    fprintf(stderr, "\nProgram aborted!!!\n");
    exit(-1);
}

double getTime(void)
{
    struct timeval t;
    gettimeofday(&t, 0);
    return t.tv_sec + 1e-6 * t.tv_usec;
}

/* deterministic stand-ins for the program's RANLUX stream: 0.25, 0.5, 0.75,
 * ...; or, once g_lcg is set, x <- 6364136223846793005 x + 1442695040888963407
 * (from x = 12345), r = (x >> 11) / 2^53 */
static int g_lcg = 0;
static unsigned long long g_lcg_x = 12345ULL;
static long long g_draws = 0;

void ranlxd(double r[], int n)
{
    static int k = 0;
    for (int i = 0; i < n; i++) {
        if (g_lcg) {
            g_lcg_x = 6364136223846793005ULL * g_lcg_x + 1442695040888963407ULL;
            r[i] = (double)(g_lcg_x >> 11) * (1.0 / 9007199254740992.0);
        } else {
            r[i] = 0.25 * (double)((k++ % 3) + 1);
        }
        g_draws++;
    }
}

/* field lifecycle (compat_lifecycle.c, a separate object as in the reference) */
long driver_vol = 0;
void initGaugeField(int start);
void releaseGaugeField(void);
void readConfig(char *name);
void writeConfig(char *name);

static void write_field(const char *dir, const char *name)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    writeConfig(path);
}

/* the reference's sweep, link by link through the exported entry points */
static int perlink(const char *field, const char *out)
{
    char path[1024];
    readConfig((char *)field);
    runParams.beta = 3.09;
    (void)plaquette();              /* creates the engine: its seed draws come first */
    snprintf(path, sizeof(path), "%s/staples.bin", out);
    FILE *f = fopen(path, "wb");
    for (long n = 0; n < driver_vol; n += 37)
        for (int d = 0; d < 4; d++) {
            su2mat A;
            staples((uint64_t)n, d, &A);
            fwrite(&A, sizeof(double), 4, f);
        }
    fclose(f);
    g_lcg = 1;
    const long long d0 = g_draws;
    double nsum = 0.0;
    for (long n = 0; n < driver_vol; n++)
        for (int d = 0; d < 4; d++) nsum += localHeatbathUpdate((uint64_t)n, d, 1);
    project_gfield_to_sun(pu);
    write_field(out, "perlink_sweep.bin");
    snprintf(path, sizeof(path), "%s/summary.txt", out);
    f = fopen(path, "w");
    fprintf(f, "acc %.17g\nuniforms %lld\n", nsum / (4.0 * (double)driver_vol), g_draws - d0);
    fclose(f);
    releaseGaugeField();
    lgt_compat_release();
    return 0;
}

int main(int argc, char **argv)
{
    if (argc != 7 && !(argc == 8 && !strcmp(argv[7], "perlink"))) {
        fprintf(stderr, "usage: %s L0 L1 L2 L3 field.bin outdir [perlink]\n", argv[0]);
        return 2;
    }
    long vol = 1;
    for (int d = 0; d < 4; d++) {
        latParams.linearExtent[d] = atoi(argv[1 + d]);
        vol *= latParams.linearExtent[d];
    }
    if (vol > MAXVOL) return 2;
    const char *out = argv[6];
    snprintf(LOG_FILE, sizeof(LOG_FILE), "%s/log.txt", out);
    driver_vol = vol;
    initGaugeField(0);
    if (argc == 8) return perlink(argv[5], out);
    readConfig(argv[5]);
    runParams.beta = 3.09;

    char path[1024];
    snprintf(path, sizeof(path), "%s/summary.txt", out);
    FILE *s = fopen(path, "w");
    fprintf(s, "plaq %.17g\n", plaquette());
    fprintf(s, "action %.17g\n", gaugeAction());

    /* loop grid as readInputFileForMeas builds it: R outer, T inner (inp_IO.c:319-334) */
    measParams.ts = 1; measParams.tf = 3; measParams.dt = 1;
    measParams.rs = 1; measParams.rf = 2; measParams.dr = 1;
    measParams.mwil_mode = 0;
    int nloop = 6, R[6], T[6], k = 0;
    for (int r = 1; r <= 2; r++)
        for (int t = 1; t <= 3; t++) { R[k] = r; T[k] = t; k++; }
    measParams.rExtents = R;
    measParams.tExtents = T;
    double w[6];
    measureWilsonLoop(w);
    for (int i = 0; i < nloop; i++) fprintf(s, "WIL %d %d\t%.17g\n", R[i], T[i], w[i]);
    fprintf(s, "QTOP %.17g\n", topologicalCharge());
    double qc;
    meas_topologicalcharge(&qc);
    fprintf(s, "QCLOV %.17g\n", qc);

    smearing_APE_spatial(3, 0.25, pu);
    write_field(out, "ape_spatial.bin");
    smearing_APE_all(2, 1.0 / 6.0, pu);
    write_field(out, "ape_all.bin");
    smearing_APE_temporal(1, 0.3, pu);
    write_field(out, "ape_temporal.bin");

    readConfig(argv[5]);
    gaugefieldUpdate(1, 2, 1, 0);
    write_field(out, "update.bin");
    project_gfield_to_sun(pu);
    write_field(out, "project.bin");
    fprintf(s, "plaq_after %.17g\n", plaquette());

    long long up = 0, down = 0;
    lgt_compat_transfers(&up, &down);
    fprintf(s, "uploads %lld\ndownloads %lld\n", up, down);
    fclose(s);

    /* Metropolis (SIM_TYPE=0) with the program's proposal width */
    runParams.eps = 0.4;
    readConfig(argv[5]);
    gaugefieldUpdate(2, 2, 0, 0);
    write_field(out, "metro.bin");
    releaseGaugeField();
    lgt_compat_release();
    return 0;
}
