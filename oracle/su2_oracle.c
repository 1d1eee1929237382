/*
 * su2_oracle.c -- FP64 CPU restatement of the SU(2) link-field path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or shipped):
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Compiled with -O2 -ffp-contract=off so that every expression rounds like the
 * reference's gcc -O3 x86-64 build (no FMA).  Citations are relative to the
 * reference's LGT_Code/ directory.
 */
#include "su2_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* PI = 2 asin(1) (modules/admin/init.c:63) equals M_PI in double. */
static const double OR_PI = 3.14159265358979323846;

/* ------------------------------------------------------------------------ */
/* lattice geometry (modules/admin/init.c:741-829)                          */
/* ------------------------------------------------------------------------ */

int64_t or_lat_vol(const int L[4])
{
    return (int64_t)L[0] * L[1] * L[2] * L[3];
}

int or_lat_init(olat_t *lat, const int L[4])
{
    int64_t stride[4];
    for (int d = 0; d < 4; d++) {
        if (L[d] < 2 || (L[d] & 1)) return -1;
        lat->L[d] = L[d];
    }
    lat->vol = or_lat_vol(L);
    stride[3] = 1;
    stride[2] = L[3];
    stride[1] = (int64_t)L[2] * L[3];
    stride[0] = (int64_t)L[1] * L[2] * L[3];
    lat->nb = (int64_t *)malloc(sizeof(int64_t) * 8 * lat->vol);
    if (!lat->nb) return -2;
    for (int64_t n = 0; n < lat->vol; n++) {
        int64_t x[4], r = n;
        for (int d = 0; d < 4; d++) { x[d] = r / stride[d]; r %= stride[d]; }
        for (int d = 0; d < 4; d++) {
            int64_t up = (x[d] + 1) % L[d], dn = (x[d] + L[d] - 1) % L[d];
            lat->nb[8 * n + d] = n + (up - x[d]) * stride[d];
            lat->nb[8 * n + d + 4] = n + (dn - x[d]) * stride[d];
        }
    }
    return 0;
}

void or_lat_free(olat_t *lat)
{
    free(lat->nb);
    lat->nb = NULL;
}

#define NB(lat, n, d) ((lat)->nb[8 * (int64_t)(n) + (d)])
#define LINK(U, n, d) ((U) + 16 * (int64_t)(n) + 4 * (d))
#define CLINK(U, n, d) ((const double *)(U) + 16 * (int64_t)(n) + 4 * (d))

/* ------------------------------------------------------------------------ */
/* quaternion algebra: same expression trees as include/gauge.h              */
/* ------------------------------------------------------------------------ */

/* v*w  (gauge.h:83-91) */
void or_qmul(const double *v, const double *w, double *o)
{
    double c0 = v[0] * w[0] - v[1] * w[1] - v[2] * w[2] - v[3] * w[3];
    double c1 = v[0] * w[1] + v[1] * w[0] - v[2] * w[3] + v[3] * w[2];
    double c2 = v[0] * w[2] + v[1] * w[3] + v[2] * w[0] - v[3] * w[1];
    double c3 = v[0] * w[3] - v[1] * w[2] + v[2] * w[1] + v[3] * w[0];
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

/* v*w^+  (gauge.h:94-102) */
void or_qmul_dag(const double *v, const double *w, double *o)
{
    double c0 = v[0] * w[0] + v[1] * w[1] + v[2] * w[2] + v[3] * w[3];
    double c1 = -v[0] * w[1] + v[1] * w[0] + v[2] * w[3] - v[3] * w[2];
    double c2 = -v[0] * w[2] - v[1] * w[3] + v[2] * w[0] + v[3] * w[1];
    double c3 = -v[0] * w[3] + v[1] * w[2] - v[2] * w[1] + v[3] * w[0];
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

/* u = v^+  (gauge.h:57-61) */
static void qdag(const double *v, double *o)
{
    o[0] = v[0]; o[1] = -v[1]; o[2] = -v[2]; o[3] = -v[3];
}

static void qadd(const double *a, const double *b, double *o)
{
    o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2]; o[3] = a[3] + b[3];
}

static void qunit(double *o) { o[0] = 1.; o[1] = 0.; o[2] = 0.; o[3] = 0.; }
static void qzero(double *o) { o[0] = 0.; o[1] = 0.; o[2] = 0.; o[3] = 0.; }

/* sqrt(|det v|) accumulated as in gauge.h:147-152 */
static double qdet_sqrt(const double *v)
{
    double u = v[0] * v[0];
    u += v[1] * v[1];
    u += v[2] * v[2];
    u += v[3] * v[3];
    return sqrt(fabs(u));
}

/* project_to_su2 (sun_vfunc.c:69-79): divide by sqrt(det). */
void or_project_link(double *u)
{
    double zw = qdet_sqrt(u);
    u[0] /= zw; u[1] /= zw; u[2] /= zw; u[3] /= zw;
}

void or_project_field(const olat_t *lat, double *U)
{
    for (int64_t n = 0; n < lat->vol; n++)
        for (int d = 0; d < 4; d++) or_project_link(LINK(U, n, d));
}

/* ------------------------------------------------------------------------ */
/* compensated summation (Neumaier)                                         */
/* ------------------------------------------------------------------------ */
typedef struct { double s, c; } nsum_t;

static void nsum_add(nsum_t *a, double x)
{
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}

static double nsum_val(const nsum_t *a) { return a->s + a->c; }

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, Salmon et al., "Parallel random numbers: as easy as       */
/* 1, 2, 3", SC'11.  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key          */
/* increments 0x9E3779B9 / 0xBB67AE85, ten rounds.                          */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53-bit uniform in [0,1) from two 32-bit words (engine definition). */
double or_u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1p-53;
}

/* Per-link FP64 uniform stream of the engine: draw j of the link comes from
 * Philox call j/2, words (0,1) for even j and (2,3) for odd j. */
typedef struct {
    uint32_t ctr[4], key[2];
    uint32_t w[4];
    int have;    /* uniforms left in w: 0, 1 or 2 */
} olink_rng_t;

enum { OR_KIND_HB = 0, OR_KIND_HOT = 1, OR_KIND_METRO = 2 };

static void lrng_init(olink_rng_t *r, uint64_t seed, uint32_t site, uint32_t tag, uint32_t sweep)
{
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32);
    r->ctr[0] = site; r->ctr[1] = tag; r->ctr[2] = sweep; r->ctr[3] = 0;
    r->have = 0;
}

static double lrng_next(olink_rng_t *r)
{
    if (r->have == 0) {
        or_philox4x32_10(r->ctr, r->key, r->w);
        r->ctr[3]++;
        r->have = 2;
        return or_u53(r->w[0], r->w[1]);
    }
    r->have = 0;
    return or_u53(r->w[2], r->w[3]);
}

/* A uniform source is either a buffer or a per-link Philox stream. */
typedef struct {
    or_ubuf_t *buf;
    olink_rng_t *rng;
} usrc_t;

static double usrc_next(usrc_t *s)
{
    if (s->rng) return lrng_next(s->rng);
    if (s->buf->pos >= s->buf->n) { s->buf->exhausted = 1; return 0.5; }
    return s->buf->buf[s->buf->pos++];
}

/* ------------------------------------------------------------------------ */
/* Kennedy-Pendleton heatbath (modules/update/heatbath.c:76-149)            */
/* ------------------------------------------------------------------------ */
static int kp_draw(double alpha, usrc_t *src, double X[4])
{
    int count = 0;
    double ud[4], delta, x0;
    do {
        ud[0] = usrc_next(src); ud[1] = usrc_next(src);
        ud[2] = usrc_next(src); ud[3] = usrc_next(src);
        double X1 = -log(1.0 - ud[0]) / alpha;                 /* heatbath.c:79 */
        double X2 = -log(1.0 - ud[1]) / alpha;                 /* heatbath.c:80 */
        double C = pow(cos(2.0 * OR_PI * ud[2]), 2.0);         /* heatbath.c:81 */
        double A = X1 * C;
        delta = X2 + A;
        count++;
    } while (pow(ud[3], 2.0) > 1.0 - (0.5 * delta));           /* heatbath.c:87 */
    x0 = 1.0 - delta;
    double a0 = usrc_next(src), a1 = usrc_next(src);           /* heatbath.c:91 */
    double phi = 2.0 * OR_PI * a0;
    double theta = acos(1.0 - (2.0 * a1));
    double length = sqrt(1.0 - pow(x0, 2.0));
    X[0] = x0;
    X[1] = length * sin(theta) * cos(phi);
    X[2] = length * sin(theta) * sin(phi);
    X[3] = length * cos(theta);
    return count;
}

int or_kp_draw(double alpha, or_ubuf_t *ub, double X[4])
{
    usrc_t s = { ub, NULL };
    return kp_draw(alpha, &s, X);
}

/* su2RandomMatrix (modules/random/random_su3.c:141-155) */
static void su2_random_matrix(usrc_t *src, double U[4])
{
    double r0 = usrc_next(src), r1 = usrc_next(src), r2 = usrc_next(src);
    double twopi = 2. * OR_PI;                                  /* random_su3.c:50-54 */
    U[0] = 1 - 2. * r0;
    U[1] = 1 - 2. * r1;
    double zw = sqrt(1. - U[1] * U[1]);
    U[2] = zw * cos(twopi * r2);
    U[3] = zw * sin(twopi * r2);
    zw = sqrt(1. - U[0] * U[0]);
    U[1] *= zw; U[2] *= zw; U[3] *= zw;
}

void or_su2_random_matrix(or_ubuf_t *ub, double U[4])
{
    usrc_t s = { ub, NULL };
    su2_random_matrix(&s, U);
}

/* localHeatbathUpdate on a given staple (heatbath.c:124-149). */
static double heatbath_link(const double Ain[4], double beta, usrc_t *src, double Unew[4])
{
    double A[4] = { Ain[0], Ain[1], Ain[2], Ain[3] }, X[4];
    double sqrt_det = qdet_sqrt(A);
    if (sqrt_det <= DBL_EPSILON) {
        su2_random_matrix(src, Unew);
        return 0.0;
    }
    A[0] /= sqrt_det; A[1] /= sqrt_det; A[2] /= sqrt_det; A[3] /= sqrt_det;
    double alpha = sqrt_det * beta;
    int count = kp_draw(alpha, src, X);
    or_qmul_dag(X, A, Unew);                                    /* U = X V^+ */
    return 1.0 / count;
}

double or_heatbath_link(const double A[4], double beta, or_ubuf_t *ub, double Unew[4])
{
    usrc_t s = { ub, NULL };
    return heatbath_link(A, beta, &s, Unew);
}

int or_heatbath_batch(int64_t n, const double *staples, double beta, const double *uniforms,
                      int64_t nuni, double *links, int *counts, int64_t *consumed)
{
    or_ubuf_t ub = { uniforms, nuni, 0, 0 };
    usrc_t s = { &ub, NULL };
    for (int64_t i = 0; i < n; i++) {
        double r = heatbath_link(staples + 4 * i, beta, &s, links + 4 * i);
        counts[i] = (r > 0.0) ? (int)lround(1.0 / r) : 0;
    }
    if (consumed) *consumed = ub.pos;
    return ub.exhausted ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* staple of the update (modules/update/metro.c:34-63)                      */
/* ------------------------------------------------------------------------ */
void or_staple(const olat_t *lat, const double *U, int64_t n, int mu, double A[4])
{
    double un0[4], un1[4], un2[4], un3[4];
    qzero(un3);
    for (int k = 0; k < 4; k++) {
        if (k == mu) continue;
        int64_t n1 = NB(lat, n, mu), n2 = NB(lat, n, k);
        or_qmul_dag(CLINK(U, n1, k), CLINK(U, n2, mu), un1);
        or_qmul_dag(un1, CLINK(U, n, k), un0);
        qadd(un3, un0, un2);
        n2 = NB(lat, n, k + 4);
        n1 = NB(lat, n2, mu);
        qdag(CLINK(U, n1, k), un0);
        or_qmul_dag(un0, CLINK(U, n2, mu), un1);
        or_qmul(un1, CLINK(U, n2, k), un0);
        qadd(un2, un0, un3);
    }
    memcpy(A, un3, sizeof(un3));
}

/* gaugefieldUpdate with utype=1, stype=0 (update.c:102-176): lexicographic
 * Gauss-Seidel sweep, then one projection of the whole field. */
double or_hb_sweep_lex(const olat_t *lat, double *U, double beta, const double *uniforms,
                       int64_t nuni, int64_t *consumed)
{
    or_ubuf_t ub = { uniforms, nuni, 0, 0 };
    usrc_t s = { &ub, NULL };
    double nsum = 0;
    for (int64_t n = 0; n < lat->vol; n++) {
        for (int mu = 0; mu < 4; mu++) {
            double A[4], Un[4];
            or_staple(lat, U, n, mu, A);
            nsum += heatbath_link(A, beta, &s, Un);
            memcpy(LINK(U, n, mu), Un, sizeof(Un));
        }
    }
    or_project_field(lat, U);
    if (consumed) *consumed = ub.pos;
    if (ub.exhausted) return -1.0;
    return (nsum / (double)(lat->vol * 4));
}

/* ------------------------------------------------------------------------ */
/* checkerboard sub-sweeps of the engine                                     */
/* ------------------------------------------------------------------------ */
/* cb index of lexicographic site n (L3 even): n/2; parity = (t+x+y+z)&1 */
static int site_parity(const olat_t *lat, int64_t n)
{
    int64_t r = n;
    int64_t z = r % lat->L[3]; r /= lat->L[3];
    int64_t y = r % lat->L[2]; r /= lat->L[2];
    int64_t x = r % lat->L[1];
    int64_t t = r / lat->L[1];
    return (int)((t + x + y + z) & 1);
}

double or_hb_subsweep_philox(const olat_t *lat, double *U, int mu, int parity, double beta,
                             uint64_t seed, uint32_t sweep)
{
    double acc = 0.0;
    for (int64_t n = 0; n < lat->vol; n++) {
        if (site_parity(lat, n) != parity) continue;
        olink_rng_t rng;
        lrng_init(&rng, seed, (uint32_t)(n / 2), (uint32_t)(mu | (parity << 2) | (OR_KIND_HB << 3)),
                  sweep);
        usrc_t s = { NULL, &rng };
        double A[4], Un[4];
        or_staple(lat, U, n, mu, A);
        acc += heatbath_link(A, beta, &s, Un);
        memcpy(LINK(U, n, mu), Un, sizeof(Un));
    }
    return acc;
}

/* Overrelaxation: the reflection U' = V^+ U^+ V^+, V = A/|A|, keeps
 * Re tr(U A) and the Haar measure; the result is projected. */
void or_or_link(const double A[4], const double U[4], double Unew[4])
{
    double a = qdet_sqrt(A);
    if (a <= DBL_EPSILON) { memcpy(Unew, U, 4 * sizeof(double)); return; }
    double V[4] = { A[0] / a, A[1] / a, A[2] / a, A[3] / a };
    double Vd[4], Ud[4], W[4];
    qdag(V, Vd);
    qdag(U, Ud);
    or_qmul(Vd, Ud, W);
    or_qmul_dag(W, V, Unew);
    or_project_link(Unew);
}

void or_or_subsweep(const olat_t *lat, double *U, int mu, int parity)
{
    for (int64_t n = 0; n < lat->vol; n++) {
        if (site_parity(lat, n) != parity) continue;
        double A[4], Un[4];
        or_staple(lat, U, n, mu, A);
        or_or_link(A, CLINK(U, n, mu), Un);
        memcpy(LINK(U, n, mu), Un, sizeof(Un));
    }
}

double or_engine_sweep(const olat_t *lat, double *U, double beta, int n_or, uint64_t seed,
                       uint32_t sweep)
{
    double acc = 0.0;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++) acc += or_hb_subsweep_philox(lat, U, mu, p, beta, seed, sweep);
    for (int k = 0; k < n_or; k++)
        for (int mu = 0; mu < 4; mu++)
            for (int p = 0; p < 2; p++) or_or_subsweep(lat, U, mu, p);
    if (n_or == 0) or_project_field(lat, U);
    return acc / (double)(lat->vol * 4);
}

/* ------------------------------------------------------------------------ */
/* Metropolis (modules/update/metro.c:65-109, exp_fct.c:137-151)            */
/* ------------------------------------------------------------------------ */
static int metropolis_link(const double A[4], double beta, double eps, usrc_t *src, const double U[4],
                           double Unew[4])
{
    double zw4[4], lpl_old, lpl, X[3], etx[4], upr[4];
    or_qmul(U, A, zw4);                                   /* sun_mul(zw, upr, stap), :87 */
    lpl_old = 2. * zw4[0];                                /* su2_trace, gauge.h:177-178 */
    for (int i = 0; i < 3; i++) X[i] = (1. - 2. * usrc_next(src));   /* proposeNewLink, :65-76 */
    double zw = eps * eps * (X[0] * X[0] + X[1] * X[1] + X[2] * X[2]) / 4.;   /* expx */
    etx[0] = 1. - zw;
    zw = 1. / (1. + zw);
    etx[0] *= zw;
    etx[1] = zw * eps * X[0];
    etx[2] = zw * eps * X[1];
    etx[3] = zw * eps * X[2];
    or_qmul(etx, U, upr);                                 /* su2_mat_mul(uz, etx, *u) */
    or_qmul(upr, A, zw4);
    lpl = 2. * zw4[0];
    double dact = exp(beta * (lpl - lpl_old) / 2.);       /* :93, SUN = 2 */
    double r = usrc_next(src);
    if (r <= dact) {
        memcpy(Unew, upr, sizeof(upr));
        return 1;
    }
    memcpy(Unew, U, 4 * sizeof(double));
    return 0;
}

int or_metropolis_link(const double A[4], double beta, double eps, or_ubuf_t *ub, const double U[4],
                       double Unew[4])
{
    usrc_t s = { ub, NULL };
    return metropolis_link(A, beta, eps, &s, U, Unew);
}

int or_metropolis_batch(int64_t n, const double *staples, const double *links_in, double beta, double eps,
                        const double *uniforms, double *links, int *accepted)
{
    for (int64_t i = 0; i < n; i++) {
        or_ubuf_t ub = { uniforms + 4 * i, 4, 0, 0 };
        usrc_t s = { &ub, NULL };
        accepted[i] = metropolis_link(staples + 4 * i, beta, eps, &s, links_in + 4 * i, links + 4 * i);
    }
    return 0;
}

int64_t or_metro_subsweep_philox(const olat_t *lat, double *U, int mu, int parity, double beta, double eps,
                                 uint64_t seed, uint32_t sweep)
{
    int64_t acc = 0;
    for (int64_t n = 0; n < lat->vol; n++) {
        if (site_parity(lat, n) != parity) continue;
        olink_rng_t rng;
        lrng_init(&rng, seed, (uint32_t)(n / 2), (uint32_t)(mu | (parity << 2) | (OR_KIND_METRO << 3)), sweep);
        usrc_t s = { NULL, &rng };
        double A[4], Un[4];
        or_staple(lat, U, n, mu, A);
        acc += metropolis_link(A, beta, eps, &s, CLINK(U, n, mu), Un);
        memcpy(LINK(U, n, mu), Un, sizeof(Un));
    }
    return acc;
}

double or_metropolis_update(const olat_t *lat, double *U, int nsweeps, double beta, double eps, uint64_t seed,
                            uint32_t sweep0)
{
    int64_t acc = 0;
    for (int s = 0; s < nsweeps; s++)
        for (int mu = 0; mu < 4; mu++)
            for (int p = 0; p < 2; p++) acc += or_metro_subsweep_philox(lat, U, mu, p, beta, eps, seed, sweep0 + s);
    or_project_field(lat, U);
    return nsweeps > 0 ? (double)acc / (double)(lat->vol * 4 * nsweeps) : 0.0;
}

/* Haar-uniform start (Shoemake, "Uniform random rotations", Graphics Gems
 * III, 1992): q = (sqrt(r1) cos 2pi r3, sqrt(1-r1) sin 2pi r2,
 * sqrt(1-r1) cos 2pi r2, sqrt(r1) sin 2pi r3). */
void or_hot_start(const olat_t *lat, double *U, uint64_t seed)
{
    for (int64_t n = 0; n < lat->vol; n++) {
        int p = site_parity(lat, n);
        for (int mu = 0; mu < 4; mu++) {
            olink_rng_t rng;
            lrng_init(&rng, seed, (uint32_t)(n / 2), (uint32_t)(mu | (p << 2) | (OR_KIND_HOT << 3)), 0);
            double r1 = lrng_next(&rng), r2 = lrng_next(&rng), r3 = lrng_next(&rng);
            double s1 = sqrt(1.0 - r1), s2 = sqrt(r1);
            double t1 = 2.0 * OR_PI * r2, t2 = 2.0 * OR_PI * r3;
            double *u = LINK(U, n, mu);
            u[0] = s2 * cos(t2);
            u[1] = s1 * sin(t1);
            u[2] = s1 * cos(t1);
            u[3] = s2 * sin(t2);
        }
    }
}

void or_cold_start(const olat_t *lat, double *U)
{
    for (int64_t n = 0; n < lat->vol; n++)
        for (int mu = 0; mu < 4; mu++) qunit(LINK(U, n, mu));
}

/* ------------------------------------------------------------------------ */
/* plaquette (modules/obs/plaquette.c:67-101)                               */
/* ------------------------------------------------------------------------ */
static double plaq_term(const olat_t *lat, const double *U, int64_t ii, int jj, int kk)
{
    double un0[4], un1[4], un2[4];
    memcpy(un0, CLINK(U, ii, jj), sizeof(un0));
    int64_t n = NB(lat, ii, jj);
    memcpy(un1, CLINK(U, n, kk), sizeof(un1));
    or_qmul(un0, un1, un2);
    n = NB(lat, ii, kk);
    qdag(CLINK(U, n, jj), un0);
    or_qmul(un2, un0, un1);
    qdag(CLINK(U, ii, kk), un2);
    or_qmul(un1, un2, un0);
    return 2. * un0[0];
}

double or_plaquette_site(const olat_t *lat, const double *U, int64_t n)
{
    double s = 0.0;
    for (int jj = 0; jj < 3; jj++)
        for (int kk = jj + 1; kk < 4; kk++) s += plaq_term(lat, U, n, jj, kk);
    return s;
}

double or_plaquette(const olat_t *lat, const double *U)
{
    double plaq = 0.;
    for (int64_t ii = 0; ii < lat->vol; ii++)
        for (int jj = 0; jj < 3; jj++)
            for (int kk = jj + 1; kk < 4; kk++) plaq += plaq_term(lat, U, ii, jj, kk);
    plaq /= (double)(2 * 6 * lat->vol);
    return plaq;
}

double or_plaquette_acc(const olat_t *lat, const double *U)
{
    nsum_t a = { 0, 0 };
    for (int64_t ii = 0; ii < lat->vol; ii++)
        for (int jj = 0; jj < 3; jj++)
            for (int kk = jj + 1; kk < 4; kk++) nsum_add(&a, plaq_term(lat, U, ii, jj, kk));
    return nsum_val(&a) / (double)(2 * 6 * lat->vol);
}

/* ------------------------------------------------------------------------ */
/* APE smearing (modules/obs/smearing.c:42-168)                             */
/* ------------------------------------------------------------------------ */
void or_ape_staple(const olat_t *lat, const double *Z, int64_t n, int mu, int dt, double S[4])
{
    double un0[4], un1[4], un2[4], un3[4];
    qzero(un3);
    for (int k = 0; k < 4; k++) {
        if (k == mu || k == dt) continue;
        int64_t n1 = NB(lat, n, mu), n2 = NB(lat, n, k);
        or_qmul(CLINK(Z, n, k), CLINK(Z, n2, mu), un1);
        or_qmul_dag(un1, CLINK(Z, n1, k), un0);
        qadd(un3, un0, un2);
        n2 = NB(lat, n, k + 4);
        n1 = NB(lat, n2, mu);
        qdag(CLINK(Z, n2, k), un0);
        or_qmul(un0, CLINK(Z, n2, mu), un1);
        or_qmul(un1, CLINK(Z, n1, k), un0);
        qadd(un2, un0, un3);
    }
    memcpy(S, un3, sizeof(un3));
}

void or_ape(const olat_t *lat, double *U, int mode, int nsteps, double sp)
{
    size_t bytes = sizeof(double) * 16 * (size_t)lat->vol;
    double *Z = (double *)malloc(bytes);
    if (!Z) return;
    int d_lo = (mode == 0) ? 1 : 0, d_hi = (mode == 2) ? 0 : 3;
    for (int s = 0; s < nsteps; s++) {
        memcpy(Z, U, bytes);                                      /* copyGaugeField(u, zu) */
        for (int64_t n = 0; n < lat->vol; n++) {
            for (int j = d_lo; j <= d_hi; j++) {
                double S[4];
                or_ape_staple(lat, Z, n, j, (mode == 1) ? j : 0, S);
                S[0] *= sp; S[1] *= sp; S[2] *= sp; S[3] *= sp;
                double *u = LINK(U, n, j);
                qadd(u, S, u);
                or_project_link(u);
            }
        }
    }
    free(Z);
}

/* ------------------------------------------------------------------------ */
/* Wilson loops (modules/obs/wilson.c:21-56, 268-324)                       */
/* ------------------------------------------------------------------------ */
double or_wloop_site(const olat_t *lat, const double *U, int64_t n, int d1, int d2, int l1, int l2)
{
    double u0[4], u1[4];
    int64_t nn = n;
    qunit(u0);
    for (int i = 0; i < l1; i++) { memcpy(u1, u0, sizeof(u0)); or_qmul(u1, CLINK(U, nn, d1), u0); nn = NB(lat, nn, d1); }
    for (int i = 0; i < l2; i++) { memcpy(u1, u0, sizeof(u0)); or_qmul(u1, CLINK(U, nn, d2), u0); nn = NB(lat, nn, d2); }
    for (int i = 0; i < l1; i++) { memcpy(u1, u0, sizeof(u0)); nn = NB(lat, nn, d1 + 4); or_qmul_dag(u1, CLINK(U, nn, d1), u0); }
    for (int i = 0; i < l2; i++) { memcpy(u1, u0, sizeof(u0)); nn = NB(lat, nn, d2 + 4); or_qmul_dag(u1, CLINK(U, nn, d2), u0); }
    return 2. * u0[0];
}

int or_wilson_loops(const olat_t *lat, const double *U, int nloop, const int *R, const int *T,
                    int seq, double *w)
{
    for (int i = 0; i < nloop; i++)
        if (T[i] > lat->L[0] - 1 || T[i] < 0 || R[i] < 0) return -1;      /* wilson.c:290-291 */
    if (nloop <= 0) return -1;
    nsum_t *acc = (nsum_t *)calloc((size_t)nloop, sizeof(nsum_t));
    if (!acc) return -2;
    for (int i = 0; i < nloop; i++) w[i] = 0.;
    for (int64_t n = 0; n < lat->vol; n++)
        for (int d = 1; d < 4; d++)
            for (int i = 0; i < nloop; i++) {
                double v = or_wloop_site(lat, U, n, 0, d, T[i], R[i]);
                if (seq) w[i] += v; else nsum_add(&acc[i], v);
            }
    for (int i = 0; i < nloop; i++) {
        if (!seq) w[i] = nsum_val(&acc[i]);
        w[i] /= (double)(3 * lat->vol * 2);
    }
    free(acc);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* improved clover topological charge (topologicalcharge.c:18-199)          */
/* ------------------------------------------------------------------------ */
/* computeLine (topologicalcharge.c:18-44) */
static void compute_line(const olat_t *lat, const double *U, double *u, int64_t *n, int dir, int l, int back)
{
    double t1[4], t2[4];
    memcpy(t1, u, sizeof(t1));
    if (!back) {
        for (int i = 0; i < l; i++) { memcpy(t2, t1, sizeof(t1)); or_qmul(t2, CLINK(U, *n, dir), t1); *n = NB(lat, *n, dir); }
    } else {
        for (int i = 0; i < l; i++) { *n = NB(lat, *n, dir + 4); memcpy(t2, t1, sizeof(t1)); or_qmul_dag(t2, CLINK(U, *n, dir), t1); }
    }
    memcpy(u, t1, sizeof(t1));
}

/* computeCloverAverages with e1 == e2 == m (topologicalcharge.c:46-110) */
static void clover_average(const olat_t *lat, const double *U, int64_t n, int d1, int d2, int m, double out[4])
{
    double sum[4], t[4];
    int64_t n1 = n;
    qzero(sum);
    /* +d1,+d2 */
    qunit(t);
    compute_line(lat, U, t, &n1, d1, m, 0); compute_line(lat, U, t, &n1, d2, m, 0);
    compute_line(lat, U, t, &n1, d1, m, 1); compute_line(lat, U, t, &n1, d2, m, 1);
    qadd(sum, t, sum);
    /* +d1,-d2 */
    qunit(t);
    compute_line(lat, U, t, &n1, d2, m, 1); compute_line(lat, U, t, &n1, d1, m, 0);
    compute_line(lat, U, t, &n1, d2, m, 0); compute_line(lat, U, t, &n1, d1, m, 1);
    qadd(sum, t, sum);
    /* -d1,+d2 */
    qunit(t);
    compute_line(lat, U, t, &n1, d2, m, 0); compute_line(lat, U, t, &n1, d1, m, 1);
    compute_line(lat, U, t, &n1, d2, m, 1); compute_line(lat, U, t, &n1, d1, m, 0);
    qadd(sum, t, sum);
    /* -d1,-d2 */
    qunit(t);
    compute_line(lat, U, t, &n1, d1, m, 1); compute_line(lat, U, t, &n1, d2, m, 1);
    compute_line(lat, U, t, &n1, d1, m, 0); compute_line(lat, U, t, &n1, d2, m, 0);
    qadd(sum, t, sum);
    sum[0] /= 4.; sum[1] /= 4.; sum[2] /= 4.; sum[3] /= 4.;
    memcpy(out, sum, sizeof(sum));
}

/* hermitianTraceless (topologicalcharge.c:112-121): X = (X - X^+)/2 */
static void anti_hermitian(double *X)
{
    double Xd[4];
    qdag(X, Xd);
    X[0] -= Xd[0]; X[1] -= Xd[1]; X[2] -= Xd[2]; X[3] -= Xd[3];
    X[0] /= 2.; X[1] /= 2.; X[2] /= 2.; X[3] /= 2.;
}

static void improved_field_strength(const olat_t *lat, const double *U, int64_t x, int d1, int d2, double F[4])
{
    const double k1 = 1.5, k2 = -0.15, k3 = 1. / 90.;
    double U1[4], U2[4], U3[4];
    clover_average(lat, U, x, d1, d2, 1, U1);
    clover_average(lat, U, x, d1, d2, 2, U2);
    clover_average(lat, U, x, d1, d2, 3, U3);
    for (int c = 0; c < 4; c++) { U1[c] *= k1; U2[c] *= k2; U3[c] *= k3; }
    qadd(U1, U2, F);
    qadd(F, U3, F);
    anti_hermitian(F);
}

/* computeCloverProducts (topologicalcharge.c:123-162) */
static double clover_products_improved(const olat_t *lat, const double *U, int64_t x, int d1, int d2, int d3, int d4)
{
    double F[4], G[4], P[4];
    improved_field_strength(lat, U, x, d1, d2, F);
    improved_field_strength(lat, U, x, d3, d4, G);
    or_qmul(F, G, P);
    return 2. * P[0];
}

/* site term of topologicalCharge (topologicalcharge.c:175-178) */
static double qdens_improved(const olat_t *lat, const double *U, int64_t n)
{
    return 8. * clover_products_improved(lat, U, n, 0, 1, 2, 3) +
           8. * clover_products_improved(lat, U, n, 0, 2, 3, 1) +
           8. * clover_products_improved(lat, U, n, 0, 3, 1, 2);
}

double or_qtop_improved(const olat_t *lat, const double *U, int seq, double *qx, double *abs_sum)
{
    double q = 0.0;
    nsum_t a = { 0, 0 }, b = { 0, 0 };
    for (int64_t n = 0; n < lat->vol; n++) {
        double v = qdens_improved(lat, U, n);
        if (qx) qx[n] = v;
        if (seq) q += v; else nsum_add(&a, v);
        nsum_add(&b, fabs(v));
    }
    if (!seq) q = nsum_val(&a);
    if (abs_sum) *abs_sum = nsum_val(&b) / (32. * OR_PI * OR_PI);
    q /= (32. * OR_PI * OR_PI);
    return q;
}

/* ------------------------------------------------------------------------ */
/* 1x1 clover charge (topologicalcharge.c:201-409)                          */
/* ------------------------------------------------------------------------ */
static void U_plpl(const olat_t *lat, const double *U, int64_t ii, int jj, int kk, double out[4])
{
    double un0[4], un1[4], un2[4];
    memcpy(un0, CLINK(U, ii, jj), sizeof(un0));
    int64_t n = NB(lat, ii, jj);
    memcpy(un1, CLINK(U, n, kk), sizeof(un1));
    or_qmul(un0, un1, un2);
    n = NB(lat, ii, kk);
    qdag(CLINK(U, n, jj), un0);
    or_qmul(un2, un0, un1);
    qdag(CLINK(U, ii, kk), un2);
    or_qmul(un1, un2, un0);
    memcpy(out, un0, sizeof(un0));
}

static void U_mipl(const olat_t *lat, const double *U, int64_t ii, int jj, int kk, double out[4])
{
    double un0[4], un1[4], un2[4];
    int64_t n = NB(lat, ii, jj + 4);
    qdag(CLINK(U, n, jj), un0);
    memcpy(un1, CLINK(U, n, kk), sizeof(un1));
    or_qmul(un0, un1, un2);
    n = NB(lat, n, kk);
    or_qmul(un2, CLINK(U, n, jj), un1);
    qdag(CLINK(U, ii, kk), un2);
    or_qmul(un1, un2, un0);
    memcpy(out, un0, sizeof(un0));
}

static void U_mimi(const olat_t *lat, const double *U, int64_t ii, int jj, int kk, double out[4])
{
    double un0[4], un1[4], un2[4];
    int64_t n = NB(lat, ii, jj + 4);
    qdag(CLINK(U, n, jj), un0);
    n = NB(lat, n, kk + 4);
    or_qmul_dag(un0, CLINK(U, n, kk), un2);
    or_qmul(un2, CLINK(U, n, jj), un1);
    n = NB(lat, ii, kk + 4);
    or_qmul(un1, CLINK(U, n, kk), un0);
    memcpy(out, un0, sizeof(un0));
}

static void U_plmi(const olat_t *lat, const double *U, int64_t ii, int jj, int kk, double out[4])
{
    double un0[4], un1[4], un2[4];
    memcpy(un0, CLINK(U, ii, jj), sizeof(un0));
    int64_t n = NB(lat, ii, jj);
    n = NB(lat, n, kk + 4);
    or_qmul_dag(un0, CLINK(U, n, kk), un2);
    n = NB(lat, ii, kk + 4);
    qdag(CLINK(U, n, jj), un0);
    or_qmul(un2, un0, un1);
    or_qmul(un1, CLINK(U, n, kk), un0);
    memcpy(out, un0, sizeof(un0));
}

/* clover sum of the four 1x1 leaves, accumulated as in compute_clover_products */
static void clover_1x1(const olat_t *lat, const double *U, int64_t x, int d1, int d2, double C[4])
{
    double p0[4], p1[4], un[4];
    qzero(p0); qzero(p1);
    U_plpl(lat, U, x, d1, d2, un); qadd(un, p0, p1);
    U_mipl(lat, U, x, d2, d1, un); qadd(un, p1, p0);
    U_mimi(lat, U, x, d1, d2, un); qadd(un, p0, p1);
    U_plmi(lat, U, x, d2, d1, un); qadd(un, p1, p0);
    C[0] = p0[0] / 4.; C[1] = p0[1] / 4.; C[2] = p0[2] / 4.; C[3] = p0[3] / 4.;
}

/* compute_clover_products (topologicalcharge.c:318-389) */
static double clover_products_1x1(const olat_t *lat, const double *U, int64_t x, int d1, int d2, int d3, int d4)
{
    double C[4], D[4], Cd[4], Dd[4], P0[4], P1[4], R[4];
    clover_1x1(lat, U, x, d1, d2, C);
    clover_1x1(lat, U, x, d3, d4, D);
    qdag(C, Cd);
    for (int c = 0; c < 4; c++) P0[c] = C[c] - Cd[c];
    qdag(D, Dd);
    for (int c = 0; c < 4; c++) P1[c] = D[c] - Dd[c];
    for (int c = 0; c < 4; c++) { P0[c] /= 2.; P1[c] /= 2.; }
    or_qmul(P0, P1, R);
    return 2. * R[0];
}

static double qdens_clover(const olat_t *lat, const double *U, int64_t n)
{
    return -clover_products_1x1(lat, U, n, 0, 1, 3, 2) - clover_products_1x1(lat, U, n, 1, 2, 3, 0) -
           clover_products_1x1(lat, U, n, 2, 0, 3, 1);
}

double or_qtop_clover(const olat_t *lat, const double *U, int seq, double *qx, double *abs_sum)
{
    double q = 0.0;
    nsum_t a = { 0, 0 }, b = { 0, 0 };
    for (int64_t n = 0; n < lat->vol; n++) {
        double v = qdens_clover(lat, U, n);
        if (qx) qx[n] = v;
        if (seq) q += v; else nsum_add(&a, v);
        nsum_add(&b, fabs(v));
    }
    if (!seq) q = nsum_val(&a);
    if (abs_sum) *abs_sum = nsum_val(&b) / (4. * OR_PI * OR_PI);
    return q / (4. * OR_PI * OR_PI);
}

/* ------------------------------------------------------------------------ */
/* adjoint of spatial APE smearing (no reference counterpart)               */
/*                                                                          */
/* Gradients are Euclidean in the 4 real components of every link:         */
/* <a,b> = a0 b0 + a.b = Re tr(a b^+)/2.  Right multiplication by q has      */
/* adjoint right multiplication by q^+, left multiplication likewise, so     */
/* for P = A B C with upstream H: dA = H (B C)^+, dB = A^+ H C^+,            */
/* dC = (A B)^+ H.  This file scatters every staple term's three factor      */
/* gradients (the engine gathers them per link instead).                    */
/* ------------------------------------------------------------------------ */
static void qsub_scaled(double *a, const double *u, double s)
{
    for (int c = 0; c < 4; c++) a[c] -= s * u[c];
}

static double qdot4(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

static void qacc(double *a, const double *b, double s)
{
    for (int c = 0; c < 4; c++) a[c] += s * b[c];
}

/* triple products with daggers: flags bit0/1/2 dagger the 1st/2nd/3rd factor */
static void qmul3(const double *a, const double *b, const double *c, int dag, double *o)
{
    double A[4], B[4], Cc[4], t[4];
    memcpy(A, a, sizeof(A)); memcpy(B, b, sizeof(B)); memcpy(Cc, c, sizeof(Cc));
    if (dag & 1) qdag(a, A);
    if (dag & 2) qdag(b, B);
    if (dag & 4) qdag(c, Cc);
    or_qmul(A, B, t);
    or_qmul(t, Cc, o);
}

/* smeared directions j and staple directions k of the three modes
 * (smearing.c:74-168): spatial j = 1..3, k in 1..3 \ {j}; all j = 0..3,
 * k != j; temporal j = 0, k = 1..3 */
static int ape_lo(int mode) { return mode == 0 ? 1 : 0; }
static int ape_hi(int mode) { return mode == 2 ? 0 : 3; }
static int ape_k_ok(int mode, int j, int k) { return k != j && (mode == 1 || k != 0); }

/* one APE step of the given mode that also keeps |M| of every smeared link */
static void ape_step_save(const olat_t *lat, int mode, const double *Z, double *Uo, double *Mn, double sp)
{
    memcpy(Uo, Z, sizeof(double) * 16 * (size_t)lat->vol);
    for (int64_t n = 0; n < lat->vol; n++)
        for (int j = ape_lo(mode); j <= ape_hi(mode); j++) {
            double S[4], M[4];
            or_ape_staple(lat, Z, n, j, (mode == 1) ? j : 0, S);
            for (int c = 0; c < 4; c++) M[c] = CLINK(Z, n, j)[c] + sp * S[c];
            double m = sqrt(qdot4(M, M));
            Mn[4 * n + j] = m;
            for (int c = 0; c < 4; c++) LINK(Uo, n, j)[c] = M[c] / m;
        }
}

/* backward of one step: Gout = dL/dU' -> Gin = dL/dZ.  Links that the mode
 * does not smear pass their cotangent through and collect what they
 * contribute to the staples of the smeared ones. */
static void ape_step_backward(const olat_t *lat, int mode, const double *Z, const double *Uo, const double *Mn,
                              double sp, const double *Gout, double *Gin)
{
    const int64_t V = lat->vol;
    double *H = (double *)calloc((size_t)16 * V, sizeof(double));
    memcpy(Gin, Gout, sizeof(double) * 16 * (size_t)V);
    for (int64_t n = 0; n < V; n++)
        for (int j = ape_lo(mode); j <= ape_hi(mode); j++) {
            double g[4];
            memcpy(g, CLINK(Gout, n, j), sizeof(g));
            qsub_scaled(g, CLINK(Uo, n, j), qdot4(g, CLINK(Uo, n, j)));
            for (int c = 0; c < 4; c++) {
                g[c] /= Mn[4 * n + j];
                H[16 * n + 4 * j + c] = g[c];
                LINK(Gin, n, j)[c] = g[c];                    /* dM/dZ = 1 */
            }
        }
    for (int64_t x = 0; x < V; x++)
        for (int j = ape_lo(mode); j <= ape_hi(mode); j++) {
            double h[4], t[4];
            for (int c = 0; c < 4; c++) h[c] = sp * H[16 * x + 4 * j + c];
            for (int k = 0; k < 4; k++) {
                if (!ape_k_ok(mode, j, k)) continue;
                int64_t xk = NB(lat, x, k), xj = NB(lat, x, j);
                /* forward term Z_k(x) Z_j(x+k) Z_k(x+j)^+ */
                qmul3(h, CLINK(Z, xj, k), CLINK(Z, xk, j), 4, t);          /* h Z_k(x+j) Z_j(x+k)^+ */
                qacc(LINK(Gin, x, k), t, 1.0);
                qmul3(CLINK(Z, x, k), h, CLINK(Z, xj, k), 1, t);          /* Z_k(x)^+ h Z_k(x+j) */
                qacc(LINK(Gin, xk, j), t, 1.0);
                qmul3(h, CLINK(Z, x, k), CLINK(Z, xk, j), 1, t);          /* h^+ Z_k(x) Z_j(x+k) */
                qacc(LINK(Gin, xj, k), t, 1.0);
                /* backward term Z_k(x-k)^+ Z_j(x-k) Z_k(x-k+j) */
                int64_t xm = NB(lat, x, k + 4), xmj = NB(lat, xm, j);
                qmul3(CLINK(Z, xm, j), CLINK(Z, xmj, k), h, 4, t);        /* Z_j(x-k) Z_k(x-k+j) h^+ */
                qacc(LINK(Gin, xm, k), t, 1.0);
                qmul3(CLINK(Z, xm, k), h, CLINK(Z, xmj, k), 4, t);        /* Z_k(x-k) h Z_k(x-k+j)^+ */
                qacc(LINK(Gin, xm, j), t, 1.0);
                qmul3(CLINK(Z, xm, j), CLINK(Z, xm, k), h, 1, t);         /* Z_j(x-k)^+ Z_k(x-k) h */
                qacc(LINK(Gin, xmj, k), t, 1.0);
            }
        }
    free(H);
}

/* the R x T loops of measureWilsonLoop summed without normalisation:
 * L = sum_x sum_{d=1..3} 2 [W(x; 0, d, T, R)]_0  (wilson.c:21-56) */
double or_wloop_objective(const olat_t *lat, const double *U, int R, int T)
{
    nsum_t a = { 0, 0 };
    for (int64_t n = 0; n < lat->vol; n++)
        for (int d = 1; d < 4; d++) nsum_add(&a, or_wloop_site(lat, U, n, 0, d, T, R));
    return nsum_val(&a);
}

/* dL/dU of the loop sum: for the link A at position k of a loop P = X A Y,
 * d(2 Re[P]_0)/dA = 2 (Y X)^+, and for a daggered link 2 Y X. */
static void wloop_grad(const olat_t *lat, const double *U, int R, int T, double *G)
{
    int m = 2 * (R + T);
    int64_t *site = (int64_t *)malloc(sizeof(int64_t) * m);
    int *dir = (int *)malloc(sizeof(int) * m), *dag = (int *)malloc(sizeof(int) * m);
    double *pre = (double *)malloc(sizeof(double) * 4 * (m + 1)), *suf = (double *)malloc(sizeof(double) * 4 * (m + 1));
    memset(G, 0, sizeof(double) * 16 * (size_t)lat->vol);
    for (int64_t x = 0; x < lat->vol; x++)
        for (int d = 1; d < 4; d++) {
            int64_t nn = x;
            int k = 0;
            for (int i = 0; i < T; i++) { site[k] = nn; dir[k] = 0; dag[k++] = 0; nn = NB(lat, nn, 0); }
            for (int i = 0; i < R; i++) { site[k] = nn; dir[k] = d; dag[k++] = 0; nn = NB(lat, nn, d); }
            for (int i = 0; i < T; i++) { nn = NB(lat, nn, 4); site[k] = nn; dir[k] = 0; dag[k++] = 1; }
            for (int i = 0; i < R; i++) { nn = NB(lat, nn, d + 4); site[k] = nn; dir[k] = d; dag[k++] = 1; }
            qunit(pre);
            for (int i = 0; i < m; i++) {
                double A[4];
                if (dag[i]) qdag(CLINK(U, site[i], dir[i]), A);
                else memcpy(A, CLINK(U, site[i], dir[i]), sizeof(A));
                or_qmul(pre + 4 * i, A, pre + 4 * (i + 1));
            }
            qunit(suf + 4 * m);
            for (int i = m - 1; i >= 0; i--) {
                double A[4];
                if (dag[i]) qdag(CLINK(U, site[i], dir[i]), A);
                else memcpy(A, CLINK(U, site[i], dir[i]), sizeof(A));
                or_qmul(A, suf + 4 * (i + 1), suf + 4 * i);
            }
            for (int i = 0; i < m; i++) {
                double YX[4], g[4];
                or_qmul(suf + 4 * (i + 1), pre + 4 * i, YX);
                if (dag[i]) memcpy(g, YX, sizeof(g));
                else qdag(YX, g);
                qacc(LINK(G, site[i], dir[i]), g, 2.0);
            }
        }
    free(site); free(dir); free(dag); free(pre); free(suf);
}

double or_ape_vjp(const olat_t *lat, const double *U, int mode, int nsteps, double sp, int R, int T, double *G)
{
    size_t fb = sizeof(double) * 16 * (size_t)lat->vol;
    double *fields = (double *)malloc(fb * (size_t)(nsteps + 1));
    double *mn = (double *)malloc(sizeof(double) * 4 * (size_t)lat->vol * (nsteps > 0 ? nsteps : 1));
    double *g2 = (double *)malloc(fb);
    memcpy(fields, U, fb);
    for (int s = 0; s < nsteps; s++)
        ape_step_save(lat, mode, fields + 16 * lat->vol * s, fields + 16 * lat->vol * (s + 1),
                      mn + 4 * lat->vol * s, sp);
    const double *Uf = fields + 16 * lat->vol * nsteps;
    double L = or_wloop_objective(lat, Uf, R, T);
    wloop_grad(lat, Uf, R, T, G);
    for (int s = nsteps - 1; s >= 0; s--) {
        ape_step_backward(lat, mode, fields + 16 * lat->vol * s, fields + 16 * lat->vol * (s + 1),
                          mn + 4 * lat->vol * s, sp, G, g2);
        memcpy(G, g2, fb);
    }
    free(fields); free(mn); free(g2);
    return L;
}
