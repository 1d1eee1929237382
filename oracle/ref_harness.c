/*
 * ref_harness.c -- ctypes-facing wrapper around the reference's own C
 * modules, compiled from their sources where they lie (see oracle/Makefile).
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement and to write
 * the golden fixtures under tests/golden/.  It defines the reference's global
 * state (MAIN_C, include/headers.h:376-413) and forwards to the reference
 * entry points unchanged; the lattice size is fixed at compile time by the
 * generated headers.h, exactly as in the reference build.
 */
#define MAIN_C

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "ranlxd.h"
#include "headers.h"
#include "modules.h"

static int ref_ready = 0;

int ref_extent(int d) { return latParams.linearExtent[d]; }
long ref_vol(void) { return (long)VOL; }

/* Minimal set-up as in main/qcd.c:35-44, without input files. */
int ref_setup(double beta, int seed)
{
    runParams.numConfs = 1;
    runParams.numThermConfs = 0;
    runParams.decorSteps = 0;
    runParams.writeConfsFreq = 1;
    runParams.beta = beta;
    snprintf(LOG_FILE, sizeof(LOG_FILE), "/dev/null");
    rlxd_init(2, seed);
    initProgram(1);
    if (!ref_ready) {
        initArrayOfNeighbours();
        initGaugeField(0);
        ref_ready = 1;
    }
    return 0;
}

void ref_set_beta(double beta) { runParams.beta = beta; }

void ref_set_field(const double *aos)
{
    for (uint64_t n = 0; n < VOL; n++)
        for (int d = 0; d < DIM; d++) {
            const double *a = aos + 16 * n + 4 * d;
            pu[n][d]->c0 = a[0]; pu[n][d]->c1 = a[1]; pu[n][d]->c2 = a[2]; pu[n][d]->c3 = a[3];
        }
}

void ref_get_field(double *aos)
{
    for (uint64_t n = 0; n < VOL; n++)
        for (int d = 0; d < DIM; d++) {
            double *a = aos + 16 * n + 4 * d;
            a[0] = pu[n][d]->c0; a[1] = pu[n][d]->c1; a[2] = pu[n][d]->c2; a[3] = pu[n][d]->c3;
        }
}

void ref_get_link(long n, int mu, double *a)
{
    a[0] = pu[n][mu]->c0; a[1] = pu[n][mu]->c1; a[2] = pu[n][mu]->c2; a[3] = pu[n][mu]->c3;
}

void ref_su2_random(double *a)
{
    su2mat u;
    su2RandomMatrix(&u);
    a[0] = u.c0; a[1] = u.c1; a[2] = u.c2; a[3] = u.c3;
}

void ref_init_field(int flag) { releaseGaugeField(); initGaugeField(flag); }

/* RNG access: state save/restore lets a caller replay exactly the uniforms a
 * reference routine consumed. */
int ref_rng_state_size(void) { return rlxd_size(); }
void ref_rng_get(int *state) { rlxd_get(state); }
void ref_rng_reset(int *state) { rlxd_reset(state); }
void ref_rng_draw(double *r, int n) { ranlxd(r, n); }

void ref_staple(long n, int mu, double *A)
{
    su2mat s;
    staples((uint64_t)n, mu, &s);
    A[0] = s.c0; A[1] = s.c1; A[2] = s.c2; A[3] = s.c3;
}

double ref_heatbath_link(long n, int mu) { return localHeatbathUpdate((uint64_t)n, mu, 1); }

double ref_metropolis_link(long n, int mu, double eps)
{
    runParams.eps = eps;
    return localMetropolisUpdate((uint64_t)n, mu, 1);
}

void ref_update(int nup) { gaugefieldUpdate(1, nup, 1, 0); }

void ref_project(void) { project_gfield_to_sun(pu); }

double ref_plaquette(void) { return plaquette(); }
double ref_gauge_action(void) { return gaugeAction(); }

void ref_ape(int mode, int ns, double sp)
{
    if (mode == 0) smearing_APE_spatial(ns, sp, pu);
    else if (mode == 1) smearing_APE_all(ns, sp, pu);
    else smearing_APE_temporal(ns, sp, pu);
}

/* measureWilsonLoop in list mode (mwil_mode = nloop, wilson.c:274-285). */
void ref_wilson(int nloop, int *R, int *T, double *w)
{
    int tmax = 0;
    for (int i = 0; i < nloop; i++) if (T[i] > tmax) tmax = T[i];
    measParams.mwil_mode = nloop;
    measParams.tExtents = T;
    measParams.rExtents = R;
    measParams.tf = tmax;
    measureWilsonLoop(w);
}

double ref_qtop_improved(void) { return topologicalCharge(); }

double ref_qtop_clover(void)
{
    double q;
    meas_topologicalcharge(&q);
    return q;
}

void ref_write_config(char *path) { writeConfig(path); }
void ref_read_config(char *path) { readConfig(path); }
void ref_read_config_doubled(char *path) { readConfigForThermalisation(path); }
