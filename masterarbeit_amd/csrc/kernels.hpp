// kernels.hpp -- HIP kernels of the SU(2) link-field engine, templated on the
// storage/arithmetic type T (double: all-FP64, compiled without FMA
// contraction; float: FP32 links and arithmetic, FP64 accumulations).
//
// Every kernel processes one or both checkerboard planes with one thread per
// link or site; blocks are 256 threads (four 64-wide waves).  Reductions are
// deterministic: per-thread double-double sums, wave shuffles, a fixed-order
// combination of the waves of a block, one partial per block, and a
// fixed-order final pass.
#pragma once

#include "su2_device.hpp"

namespace lgt {

constexpr int BLOCK = 256;
constexpr int NWAVES = BLOCK / 64;
constexpr double PI_D = 3.14159265358979323846;   // = 2 asin(1), LGT_Code/modules/admin/init.c:63

// Tile order of the 256-site blocks of a checkerboard plane.
//  bit 0: workgroups are dispatched round-robin over the 8 XCDs of the chip;
//         XCD k takes the k-th contiguous eighth of the tile sequence, so that
//         tiles that share neighbours share one L2.
//  bit 1: time fastest: consecutive tiles are the same spatial tile at
//         consecutive t, so the +-t neighbours (a whole time slice away in
//         memory) are read while still in L2.  Needs the cb sites of one time
//         slice to be a multiple of 256; the host clears the bit otherwise.
//  bit 2: grouped: time fastest within groups of g.gs consecutive spatial
//         tiles, group after group, so that the +-x neighbours (a few tiles
//         away in the same time slice) and the +-t neighbours (g.gs tiles
//         away) both stay close in the sequence.
// Returns the first cb index of the block's tile.
__device__ __forceinline__ uint32_t xcd_remap(const Geo& g, uint32_t b, uint32_t nb)
{
    const uint32_t l = ((g.order & 1) && (nb & 7u) == 0) ? (b & 7u) * (nb >> 3) + (b >> 3) : b;
    return (g.order & 8) ? nb - 1u - l : l;
}

__device__ __forceinline__ uint32_t tile_map(const Geo& g, uint32_t b)
{
    if (g.order & 4) {
        uint32_t grp = fdiv(g.dgt, b);
        uint32_t r = b - grp * ((uint32_t)g.gs * (uint32_t)g.L[0]);
        uint32_t t = fdiv(g.dgs, r);
        uint32_t s = grp * (uint32_t)g.gs + (r - t * (uint32_t)g.gs);
        return t * (g.vh / (uint32_t)g.L[0]) + s * 256u;
    }
    if (g.order & 2) {
        uint32_t s = fdiv(g.dt, b);
        uint32_t t = b - s * (uint32_t)g.L[0];
        return t * (g.vh / (uint32_t)g.L[0]) + s * 256u;
    }
    return b * 256u;
}

__device__ __forceinline__ uint32_t tile_base(const Geo& g)
{
    return tile_map(g, xcd_remap(g, blockIdx.x, gridDim.x));
}

// Launches over both parities (grid x = 2 * tiles): the two parities of a
// tile are consecutive blocks of the (remapped) sequence, so they run side by
// side on one XCD and read the neighbours they share (the other parity's
// links of the same few rows) through one L2 instead of in two passes over
// the field.  Sets p and returns the first cb index of the tile.
__device__ __forceinline__ uint32_t tile_base_pair(const Geo& g, int& p)
{
    uint32_t b = xcd_remap(g, blockIdx.x, gridDim.x);
    p = (int)(b & 1u);
    return tile_map(g, b >> 1);
}

// Both parities of a tile in one block (waves alternate between the parities,
// consecutive blocks take the two halves of the tile), so that the links the
// two parities share are read through one CU's L1.  Sets p and returns the
// thread's cb index.
__device__ __forceinline__ uint32_t tile_site_mixed(const Geo& g, int& p)
{
    const uint32_t b = xcd_remap(g, blockIdx.x, gridDim.x);
    p = (int)((threadIdx.x >> 6) & 1u);
    return tile_map(g, b >> 1) + (b & 1u) * 128u + ((threadIdx.x >> 7) << 6) + (threadIdx.x & 63u);
}

// keeps q's value in registers at this point of the program: what computes it
// is not moved past the following statements
template <typename T>
__device__ __forceinline__ void pin(Q<T>& q)
{
    asm volatile("" : "+v"(q.c0), "+v"(q.c1), "+v"(q.c2), "+v"(q.c3));
}

// the same for a site: the addresses derived from it are computed after this
// point, not hoisted ahead of earlier work
__device__ __forceinline__ void pin(Site& s)
{
    asm volatile("" : "+v"(s.c[0]), "+v"(s.c[1]), "+v"(s.c[2]), "+v"(s.c[3]));
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ DD wave_sum_dd(DD v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        DD o;
        o.hi = __shfl_xor(v.hi, off);
        o.lo = __shfl_xor(v.lo, off);
        v = dd_add(v, o);
    }
    return v;
}

// block-wide double-double sum over NW waves, valid in thread 0
template <int NW = NWAVES>
__device__ __forceinline__ DD block_sum_dd(DD v)
{
    __shared__ DD ws[NW];
    v = wave_sum_dd(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) ws[wid] = v;
    __syncthreads();
    DD s = DD{0.0, 0.0};
    if (threadIdx.x == 0)
        for (int w = 0; w < NW; w++) s = dd_add(s, ws[w]);
    __syncthreads();
    return s;
}

// out[row] = fixed-order dd sum of in[row*ncols .. row*ncols+ncols), one
// block of REDUCE_THREADS per row.  Thread c sums columns c, c + RT, c + 2 RT,
// ... in that order, loading eight of them ahead of their sums: with one load
// in flight per thread, a 48^4 plaquette's 20736 partials took 26-37 us, each
// thread waiting on 81 loads in turn.
constexpr int REDUCE_THREADS = 1024;
static __global__ void __launch_bounds__(REDUCE_THREADS) k_reduce_rows(const DD* __restrict__ in, int ncols,
                                                                      DD* __restrict__ out)
{
    constexpr int RT = REDUCE_THREADS, AHEAD = 8;
    const DD* row = in + (size_t)blockIdx.x * ncols;
    DD a = DD{0.0, 0.0};
    int c = threadIdx.x;
    for (; c + (AHEAD - 1) * RT < ncols; c += AHEAD * RT) {
        DD v[AHEAD];
#pragma unroll
        for (int k = 0; k < AHEAD; k++) v[k] = row[c + k * RT];
#pragma unroll
        for (int k = 0; k < AHEAD; k++) a = dd_add(a, v[k]);
    }
    for (; c < ncols; c += RT) a = dd_add(a, row[c]);
    a = block_sum_dd<RT / 64>(a);
    if (threadIdx.x == 0) out[blockIdx.x] = a;
}

// ---------------------------------------------------------------------------
// field initialisation and host-layout conversion
// ---------------------------------------------------------------------------
// grid: x over cb sites, y = mu*2 + parity
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_cold(Planes<T> P, Geo g)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int mu = blockIdx.y >> 1, p = blockIdx.y & 1;
    if (i < g.vh) qstore_fill<T>(P.p[mu][p], i, qunit<T>());
}

// Haar-uniform links (Shoemake's method) from the link's Philox stream,
// always evaluated in FP64 so both storage precisions start from the same
// configuration.
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_hot(Planes<T> P, Geo g, uint64_t seed)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int mu = blockIdx.y >> 1, p = blockIdx.y & 1;
    if (i >= g.vh) return;
    RngSrc64 src(seed, i, (uint32_t)(mu | (p << 2) | (KIND_HOT << 3)), 0u);
    double r1 = src.next(), r2 = src.next(), r3 = src.next();
    double s1 = sqrt(1.0 - r1), s2 = sqrt(r1);
    double t1 = 2.0 * PI_D * r2, t2 = 2.0 * PI_D * r3;
    Q<T> q;
    q.c0 = (T)(s2 * cos(t2));
    q.c1 = (T)(s1 * sin(t1));
    q.c2 = (T)(s1 * cos(t1));
    q.c3 = (T)(s2 * sin(t2));
    qstore_fill<T>(P.p[mu][p], i, q);
}

// site of lexicographic index n (n = z + L3(y + L2(x + L1 t)))
__device__ __forceinline__ Site lex_site(const Geo& g, uint32_t n)
{
    Site s;
    s.c[3] = (int)(n % (uint32_t)g.L[3]);
    n /= (uint32_t)g.L[3];
    s.c[2] = (int)(n % (uint32_t)g.L[2]);
    n /= (uint32_t)g.L[2];
    s.c[1] = (int)(n % (uint32_t)g.L[1]);
    s.c[0] = (int)(n / (uint32_t)g.L[1]);
    return s;
}

// lexicographic site of cb index i, parity p
__device__ __forceinline__ uint64_t lex_index(const Geo& g, const Site& s)
{
    return (((uint64_t)s.c[0] * g.L[1] + s.c[1]) * g.L[2] + s.c[2]) * g.L[3] + s.c[3];
}

// Reference AoS FP64 [VOL][4][4] <-> planes, one thread per lexicographic
// site n: its record (4 links, 128 bytes) is read or written whole, so a wave
// moves 8 KB of contiguous AoS, and its links are at cb index n/2 of the
// parity planes (n even and n + 1 hold one site of each parity).  One thread
// per (link, parity) wrote 32-byte pieces 256 bytes apart instead (the FP32
// download's conversion ran at 1.2 TB/s).
// grid: x over lexicographic sites (VOL / BLOCK blocks)
__device__ __forceinline__ int lex_parity(const Geo& g, uint32_t n)
{
    const Site s = cb_site(g, n >> 1, 0);   // t, x, y of n (and z/2 in c[3] >> 1)
    const int z = (s.c[3] & ~1) | (int)(n & 1u);
    return (s.c[0] + s.c[1] + s.c[2] + z) & 1;
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_from_aos(Planes<T> P, Geo g, const double* __restrict__ aos)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= 2 * g.vh) return;
    const int p = lex_parity(g, n);
    const d2v* a = (const d2v*)(aos + 16 * (size_t)n);
    d2v v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = a[k];
    // both parity planes from p[mu][0] (adjacent planes): a plane pointer
    // chosen by the lane's parity would be read from the kernel arguments with
    // a vector load, and every store would wait on the loads in flight
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
        qstore<T>(P.p[mu][0], (uint32_t)p * g.vh + (n >> 1),
                  Q<T>{(T)v[2 * mu].x, (T)v[2 * mu].y, (T)v[2 * mu + 1].x, (T)v[2 * mu + 1].y});
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_to_aos(Planes<T> P, Geo g, double* __restrict__ aos)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= 2 * g.vh) return;
    const int p = lex_parity(g, n);
    d2v* a = (d2v*)(aos + 16 * (size_t)n);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
        const Q<T> q = qload<T>(P.p[mu][0], (uint32_t)p * g.vh + (n >> 1));   // adjacent parity planes
        d2v lo, hi;
        lo.x = (double)q.c0; lo.y = (double)q.c1;
        hi.x = (double)q.c2; hi.y = (double)q.c3;
        a[2 * mu] = lo;
        a[2 * mu + 1] = hi;
    }
}

// ---------------------------------------------------------------------------
// random gauge transformation U'_mu(x) = g(x) U_mu(x) g(x+mu)^+ (no reference
// counterpart; observables are invariant under it and APE smearing is
// covariant, which tests check at any lattice size).  g(x) is Haar-uniform
// (Shoemake's method) from the Philox stream of site x under the given seed,
// in FP64; every link depends on its own old value only, so the transform
// runs in place.
// ---------------------------------------------------------------------------
__device__ __forceinline__ Q<double> gauge_element(const Geo& g, const Site& s, uint64_t seed)
{
    RngSrc64 src(seed, cb_index(g, s), (uint32_t)(parity(s) << 2 | (KIND_GAUGE << 3)), 0u);
    double r1 = src.next(), r2 = src.next(), r3 = src.next();
    double s1 = sqrt(1.0 - r1), s2 = sqrt(r1);
    double t1 = 2.0 * PI_D * r2, t2 = 2.0 * PI_D * r3;
    return Q<double>{s2 * cos(t2), s1 * sin(t1), s1 * cos(t1), s2 * sin(t2)};
}

// grid: x over cb sites, y = mu*2 + parity
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_gauge(Planes<T> P, Geo g, uint64_t seed)
{
    const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    const int mu = blockIdx.y >> 1, p = blockIdx.y & 1;
    if (i >= g.vh) return;
    const Site s = cb_site(g, i, p);
    const Q<double> gx = gauge_element(g, s, seed), gy = gauge_element(g, shift(g, s, mu, 1), seed);
    const Q<T> u = qload<T>(P.p[mu][p], i);
    const Q<double> v = qmul_dag(qmul(gx, Q<double>{u.c0, u.c1, u.c2, u.c3}), gy);
    qstore<T>(P.p[mu][p], i, Q<T>{(T)v.c0, (T)v.c1, (T)v.c2, (T)v.c3});
}

// g(x) of every lexicographic site (VOL quaternions, for checks)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_gauge_elements(Geo g, uint64_t seed, double* __restrict__ out)
{
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= 2 * g.vh) return;
    const Q<double> q = gauge_element(g, lex_site(g, n), seed);
    out[4 * (size_t)n] = q.c0; out[4 * (size_t)n + 1] = q.c1; out[4 * (size_t)n + 2] = q.c2; out[4 * (size_t)n + 3] = q.c3;
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_project(Planes<T> P, Geo g)
{
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int mu = blockIdx.y >> 1, p = blockIdx.y & 1;
    if (i >= g.vh) return;
    qstore<T>(P.p[mu][p], i, qproject(qload<T>(P.p[mu][p], i)));
}

// ---------------------------------------------------------------------------
// staple of the update (LGT_Code/modules/update/metro.c:34-63)
//   A = sum_{k != mu} [U_k(n+mu) U_mu(n+k)^+ U_k(n)^+ + U_k(n-k+mu)^+ U_mu(n-k)^+ U_k(n-k)]
// Site s has parity p; n+mu, n+k, n-k have parity 1-p, n-k+mu has parity p.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Q<T> lk(const Planes<T>& P, const Geo& g, int mu, int par, const Site& s)
{
    return qload<T>(P.p[mu][par], cb_index(g, s));
}

template <typename T, int MU>
__device__ __forceinline__ Q<T> update_staple(const Planes<T>& P, const Geo& g, const Site& s, int p)
{
    const int q = p ^ 1;
    Q<T> acc = qzero<T>();
    Site smu = shift(g, s, MU, 1);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k == MU) continue;
        Site sk = shift(g, s, k, 1);
        Q<T> un1 = qmul_dag(lk(P, g, k, q, smu), lk(P, g, MU, q, sk));
        Q<T> un0 = qmul_dag(un1, lk(P, g, k, p, s));
        Q<T> un2 = qadd(acc, un0);
        Site skm = shift(g, s, k, -1);
        Site skmmu = shift(g, skm, MU, 1);
        un0 = qdag(lk(P, g, k, p, skmmu));
        un1 = qmul_dag(un0, lk(P, g, MU, q, skm));
        un0 = qmul(un1, lk(P, g, k, q, skm));
        acc = qadd(un2, un0);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// Kennedy-Pendleton heatbath (LGT_Code/modules/update/heatbath.c:76-149)
// ---------------------------------------------------------------------------
// FP64: the reference formula term by term (pow(x,2) evaluated as x*x).
// The trial loop is bounded: a source that runs dry or KP_MAX_TRIALS
// rejections (only for alpha -> 0) end it with accepted = false.
constexpr int KP_MAX_TRIALS = 10000;

// one trial of the rejection loop (heatbath.c:84-95): 4 uniforms; returns
// whether it is accepted, with its delta
template <class Src>
__device__ __forceinline__ bool kp_trial64(double alpha, Src& src, double& delta)
{
    double u0 = src.next(), u1 = src.next(), u2 = src.next();
    double u3 = src.next();
    double X1 = -log(1.0 - u0) / alpha;
    double X2 = -log(1.0 - u1) / alpha;
    double cc = cos(2.0 * PI_D * u2);
    double C = cc * cc;
    double A = X1 * C;
    delta = X2 + A;
    return !(u3 * u3 > 1.0 - (0.5 * delta));
}

// the accepted x0 = 1 - delta and its direction on S^2 (heatbath.c:97-107):
// 2 uniforms
template <class Src>
__device__ __forceinline__ Q<double> kp_finish64(double delta, Src& src)
{
    Q<double> X;
    double x0 = 1.0 - delta;
    double a0 = src.next(), a1 = src.next();
    double phi = 2.0 * PI_D * a0;
    double theta = acos(1.0 - (2.0 * a1));
    double length = sqrt(1.0 - x0 * x0);
    X.c0 = x0;
    X.c1 = length * sin(theta) * cos(phi);
    X.c2 = length * sin(theta) * sin(phi);
    X.c3 = length * cos(theta);
    return X;
}

template <class Src>
__device__ __forceinline__ int kp_draw64(double alpha, Src& src, Q<double>& X, bool& accepted)
{
    int count = 0;
    double delta;
    do {
        accepted = kp_trial64(alpha, src, delta);
        count++;
    } while (!accepted && count < KP_MAX_TRIALS && !src.dry());
    if (!accepted) return count;
    X = kp_finish64(delta, src);
    return count;
}

// su2RandomMatrix (LGT_Code/modules/random/random_su3.c:141-155)
template <class Src>
__device__ __forceinline__ Q<double> random_su2_64(Src& src)
{
    double r0 = src.next(), r1 = src.next(), r2 = src.next();
    const double twopi = 2. * PI_D;
    Q<double> u;
    u.c0 = 1 - 2. * r0;
    u.c1 = 1 - 2. * r1;
    double zw = sqrt(1. - u.c1 * u.c1);
    u.c2 = zw * cos(twopi * r2);
    u.c3 = zw * sin(twopi * r2);
    zw = sqrt(1. - u.c0 * u.c0);
    u.c1 *= zw; u.c2 *= zw; u.c3 *= zw;
    return u;
}

// localHeatbathUpdate for a given staple (heatbath.c:124-149); returns the
// trial count (0 on the degenerate branch).  When the bounded trial loop ends
// without acceptance (alpha -> 0, where the heatbath distribution becomes
// uniform on SU(2)) the link is drawn uniformly like the degenerate branch.
template <class Src>
__device__ __forceinline__ int heatbath_link64(Q<double> A, double beta, Src& src, Q<double>& Unew)
{
    double sqrt_det = qdet_sqrt(A);
    if (sqrt_det <= __DBL_EPSILON__) {
        Unew = random_su2_64(src);
        return 0;
    }
    A = qdiv(A, sqrt_det);
    double alpha = sqrt_det * beta;
    Q<double> X;
    bool ok;
    int count = kp_draw64(alpha, src, X, ok);
    Unew = ok ? qmul_dag(X, A) : random_su2_64(src);
    return count;
}

// FP32: one Philox call per trial (four 24-bit uniforms), hardware
// transcendentals (v_log_f32 = log2, v_cos_f32/v_sin_f32 take revolutions),
// 1 - x0^2 as delta(2 - delta) and sin(acos(c)) as 2 sqrt(a(1-a)).
__device__ __forceinline__ Q<float> random_su2_32(LinkRng& rng)
{
    U32x4 w = rng.next();
    float r0 = u24(w.x), r1 = u24(w.y), r2 = u24(w.z);
    Q<float> u;
    u.c0 = 1.0f - 2.0f * r0;
    u.c1 = 1.0f - 2.0f * r1;
    float zw = sqrtf(fmaxf(0.0f, 1.0f - u.c1 * u.c1));
    u.c2 = zw * __builtin_amdgcn_cosf(r2);
    u.c3 = zw * __builtin_amdgcn_sinf(r2);
    zw = sqrtf(fmaxf(0.0f, 1.0f - u.c0 * u.c0));
    u.c1 *= zw; u.c2 *= zw; u.c3 *= zw;
    return u;
}

__device__ __forceinline__ int heatbath_link32(Q<float> A, float beta, LinkRng& rng, Q<float>& Unew)
{
    float n2 = A.c0 * A.c0 + A.c1 * A.c1 + A.c2 * A.c2 + A.c3 * A.c3;
    if (n2 <= 1e-30f) {
        Unew = random_su2_32(rng);
        return 0;
    }
    float rn = rsqrtf(n2);
    float a = n2 * rn;                      // |A|
    Q<float> V = qscale(A, rn);
    const float kln2 = 0.69314718055994531f;
    float s = -kln2 / (a * beta);           // -1/alpha * ln 2
    int count = 0;
    float delta;
    bool accepted;
    do {
        U32x4 w = rng.next();
        float l0 = __builtin_amdgcn_logf(1.0f - u24(w.x));   // log2
        float l1 = __builtin_amdgcn_logf(1.0f - u24(w.y));
        float c = __builtin_amdgcn_cosf(u24(w.z));
        float u3 = u24(w.w);
        delta = s * (l1 + l0 * (c * c));
        count++;
        accepted = !(u3 * u3 > 1.0f - 0.5f * delta);
    } while (!accepted && count < KP_MAX_TRIALS);
    if (!accepted) {
        Unew = random_su2_32(rng);
        return count;
    }
    U32x4 w = rng.next();
    float a0 = u24(w.x), a1 = u24(w.y);
    float ct = 1.0f - 2.0f * a1;
    float st = 2.0f * sqrtf(a1 * (1.0f - a1));
    float len = sqrtf(fmaxf(0.0f, delta * (2.0f - delta)));
    Q<float> X;
    X.c0 = 1.0f - delta;
    X.c1 = len * st * __builtin_amdgcn_cosf(a0);
    X.c2 = len * st * __builtin_amdgcn_sinf(a0);
    X.c3 = len * ct;
    Unew = qmul_dag(X, V);
    return count;
}

template <typename T> struct HB;
template <> struct HB<double> {
    __device__ static __forceinline__ int update(const Q<double>& A, double beta, uint64_t seed, uint32_t i,
                                                 uint32_t tag, uint32_t sweep, Q<double>& Un)
    {
        RngSrc64 src(seed, i, tag, sweep);
        return heatbath_link64(A, beta, src, Un);
    }
};
template <> struct HB<float> {
    __device__ static __forceinline__ int update(const Q<float>& A, float beta, uint64_t seed, uint32_t i,
                                                 uint32_t tag, uint32_t sweep, Q<float>& Un)
    {
        LinkRng rng(seed, i, tag, sweep);
        return heatbath_link32(A, beta, rng, Un);
    }
};

// heatbath sub-sweep of direction MU, parity p.  acc_partial[tile] receives
// the block's sum of 1/count (nullable).  PROJ stores the new link projected
// onto SU(2), which stands in for a separate projection pass after the sweep.
template <typename T, int MU, bool PROJ>
__global__ void __launch_bounds__(BLOCK) k_heatbath(Planes<T> P, Geo g, int p, T beta, uint64_t seed, uint32_t sweep,
                                                    DD* __restrict__ acc_partial)
{
    uint32_t base = tile_base(g);
    uint32_t tile = base / BLOCK;
    uint32_t i = base + threadIdx.x;
    int count = 0;
    if (i < g.vh) {
        Site s = cb_site(g, i, p);
        Q<T> A = update_staple<T, MU>(P, g, s, p);
        Q<T> Un;
        count = HB<T>::update(A, beta, seed, i, (uint32_t)(MU | (p << 2) | (KIND_HB << 3)), sweep, Un);
        qstore<T>(P.p[MU][p], i, PROJ ? qproject(Un) : Un);
    }
    if (acc_partial) {
        DD s = block_sum_dd(DD{count > 0 ? 1.0 / count : 0.0, 0.0});
        if (threadIdx.x == 0) acc_partial[tile] = s;
    }
}

// ---------------------------------------------------------------------------
// overrelaxation: U' = V^+ U^+ V^+, V = A/|A|, projected on store
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Q<T> overrelax_link(const Q<T>& A, const Q<T>& U);

template <> __device__ __forceinline__ Q<double> overrelax_link(const Q<double>& A, const Q<double>& U)
{
    double a = qdet_sqrt(A);
    if (a <= __DBL_EPSILON__) return U;
    Q<double> V = qdiv(A, a);
    Q<double> W = qmul(qdag(V), qdag(U));
    return qproject(qmul_dag(W, V));
}

template <> __device__ __forceinline__ Q<float> overrelax_link(const Q<float>& A, const Q<float>& U)
{
    float n2 = A.c0 * A.c0 + A.c1 * A.c1 + A.c2 * A.c2 + A.c3 * A.c3;
    if (n2 <= 1e-30f) return qproject(U);
    // V^+ U^+ V^+ = (V U V)^+; |V|=1 is not needed: the projection removes |A|^2
    Q<float> W = qmul(qdag(A), qdag(U));
    return qproject(qmul_dag(W, A));
}

template <typename T, int MU>
__global__ void __launch_bounds__(BLOCK) k_overrelax(Planes<T> P, Geo g, int p)
{
    uint32_t i = tile_base(g) + threadIdx.x;
    if (i >= g.vh) return;
    Site s = cb_site(g, i, p);
    Q<T> A = update_staple<T, MU>(P, g, s, p);
    // the link's own old value is read once by this launch and by no staple:
    // in FP32 the nontemporal hint keeps it from displacing gathered
    // neighbours in L2 (-4 % at 48^4; in FP64 it measured slower)
    Q<T> U;
    if constexpr (sizeof(T) == 4) {
        typename Store<T>::V v = __builtin_nontemporal_load(&P.p[MU][p][i]);
        U = Q<T>{v.x, v.y, v.z, v.w};
    } else {
        U = qload<T>(P.p[MU][p], i);
    }
    qstore<T>(P.p[MU][p], i, overrelax_link(A, U));
}

// ---------------------------------------------------------------------------
// Metropolis, one hit (LGT_Code/modules/update/metro.c:65-109 with the proposal
// expx of exp_fct.c:137-151): X_i = 1 - 2 r_i, U' = expx(eps, X) U, accepted
// iff r_4 <= exp(beta (Re tr U'A - Re tr UA) / 2).  The FP64 form keeps the
// reference's operation order; returns 1 if accepted.
// ---------------------------------------------------------------------------
template <typename T, class Src>
__device__ __forceinline__ int metropolis_link(const Q<T>& A, T beta, T eps, Src& src, const Q<T>& U, Q<T>& Un)
{
    T lpl_old = T(2) * qmul(U, A).c0;
    T X1 = (T(1) - T(2) * (T)src.next());
    T X2 = (T(1) - T(2) * (T)src.next());
    T X3 = (T(1) - T(2) * (T)src.next());
    T zw = eps * eps * (X1 * X1 + X2 * X2 + X3 * X3) / T(4);
    Q<T> etx;
    etx.c0 = T(1) - zw;
    zw = T(1) / (T(1) + zw);
    etx.c0 *= zw;
    etx.c1 = zw * eps * X1;
    etx.c2 = zw * eps * X2;
    etx.c3 = zw * eps * X3;
    Q<T> upr = qmul(etx, U);
    T lpl = T(2) * qmul(upr, A).c0;
    T dact = exp(beta * (lpl - lpl_old) / T(2));
    T r = (T)src.next();
    const int acc = (r <= dact) ? 1 : 0;
    Un = acc ? upr : U;
    return acc;
}

// FP32 uniforms: one Philox call gives the hit's four 24-bit uniforms
struct RngSrc32 {
    U32x4 w;
    int k;
    __device__ __forceinline__ RngSrc32(uint64_t seed, uint32_t site, uint32_t tag, uint32_t sweep) : k(0)
    {
        LinkRng rng(seed, site, tag, sweep);
        w = rng.next();
    }
    __device__ __forceinline__ float next()
    {
        uint32_t v = (k == 0) ? w.x : (k == 1) ? w.y : (k == 2) ? w.z : w.w;
        k++;
        return u24(v);
    }
};

template <typename T> struct MetroSrc;
template <> struct MetroSrc<double> { typedef RngSrc64 type; };
template <> struct MetroSrc<float> { typedef RngSrc32 type; };

// Metropolis sub-sweep of direction MU, parity p; acc_partial[tile] receives
// the block's number of accepted links (nullable).
template <typename T, int MU>
__global__ void __launch_bounds__(BLOCK) k_metropolis(Planes<T> P, Geo g, int p, T beta, T eps, uint64_t seed,
                                                      uint32_t sweep, DD* __restrict__ acc_partial)
{
    uint32_t base = tile_base(g);
    uint32_t tile = base / BLOCK;
    uint32_t i = base + threadIdx.x;
    int acc = 0;
    if (i < g.vh) {
        Site s = cb_site(g, i, p);
        Q<T> A = update_staple<T, MU>(P, g, s, p);
        Q<T> U = qload<T>(P.p[MU][p], i), Un;
        typename MetroSrc<T>::type src(seed, i, (uint32_t)(MU | (p << 2) | (KIND_METRO << 3)), sweep);
        acc = metropolis_link<T>(A, beta, eps, src, U, Un);
        qstore<T>(P.p[MU][p], i, Un);
    }
    if (acc_partial) {
        DD s = block_sum_dd(DD{(double)acc, 0.0});
        if (threadIdx.x == 0) acc_partial[tile] = s;
    }
}

// ---------------------------------------------------------------------------
// per-link harnesses with injected inputs (FP64)
// ---------------------------------------------------------------------------
static __global__ void __launch_bounds__(BLOCK) k_hb_harness(int n, const double* __restrict__ staples, double beta,
                                                      const double* __restrict__ uniforms,
                                                      const int* __restrict__ offsets, const int* __restrict__ lens,
                                                      double* __restrict__ links, int* __restrict__ counts)
{
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    BufSrc src{uniforms + offsets[i], lens[i], 0};
    Q<double> A{staples[4 * i], staples[4 * i + 1], staples[4 * i + 2], staples[4 * i + 3]};
    Q<double> U;
    int c = heatbath_link64(A, beta, src, U);
    links[4 * i] = U.c0; links[4 * i + 1] = U.c1; links[4 * i + 2] = U.c2; links[4 * i + 3] = U.c3;
    counts[i] = (src.pos > src.n) ? -1 : c;
}

static __global__ void __launch_bounds__(BLOCK) k_metro_harness(int n, const double* __restrict__ staples,
                                                         const double* __restrict__ links_in, double beta, double eps,
                                                         const double* __restrict__ uniforms,
                                                         double* __restrict__ links, int* __restrict__ accepted)
{
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    BufSrc src{uniforms + 4 * (size_t)i, 4, 0};
    Q<double> A{staples[4 * i], staples[4 * i + 1], staples[4 * i + 2], staples[4 * i + 3]};
    Q<double> U{links_in[4 * i], links_in[4 * i + 1], links_in[4 * i + 2], links_in[4 * i + 3]}, R;
    accepted[i] = metropolis_link<double>(A, beta, eps, src, U, R);
    links[4 * i] = R.c0; links[4 * i + 1] = R.c1; links[4 * i + 2] = R.c2; links[4 * i + 3] = R.c3;
}

static __global__ void __launch_bounds__(BLOCK) k_or_harness(int n, const double* __restrict__ staples,
                                                      const double* __restrict__ links_in, double* __restrict__ links)
{
    int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    Q<double> A{staples[4 * i], staples[4 * i + 1], staples[4 * i + 2], staples[4 * i + 3]};
    Q<double> U{links_in[4 * i], links_in[4 * i + 1], links_in[4 * i + 2], links_in[4 * i + 3]};
    Q<double> R = overrelax_link(A, U);
    links[4 * i] = R.c0; links[4 * i + 1] = R.c1; links[4 * i + 2] = R.c2; links[4 * i + 3] = R.c3;
}

template <typename T>
__device__ __forceinline__ Q<double> widen(const Q<T>& q) { return Q<double>{q.c0, q.c1, q.c2, q.c3}; }

// ---------------------------------------------------------------------------
// one link of the device field: staples(n, dir, &A) (LGT_Code/modules/update/
// metro.c:34-63) and localHeatbathUpdate(n, dir, m) (heatbath.c:124-149) at a
// lexicographic site n, for programs that call them link by link.  One
// thread; FP32 fields compute the staple in FP32 and run the FP64 sampler.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ Q<T> update_staple_mu(const Planes<T>& P, const Geo& g, const Site& s, int p, int mu)
{
    switch (mu) {
    case 0: return update_staple<T, 0>(P, g, s, p);
    case 1: return update_staple<T, 1>(P, g, s, p);
    case 2: return update_staple<T, 2>(P, g, s, p);
    default: return update_staple<T, 3>(P, g, s, p);
    }
}

// out[0..3] = the staple of link (n, mu)
template <typename T>
__global__ void __launch_bounds__(64) k_staple_link(Planes<T> P, Geo g, uint32_t n, int mu, double* __restrict__ out)
{
    if (threadIdx.x != 0) return;
    const Site s = lex_site(g, n);
    const Q<double> A = widen(update_staple_mu(P, g, s, parity(s), mu));
    out[0] = A.c0; out[1] = A.c1; out[2] = A.c2; out[3] = A.c3;
}

// The heatbath of link (n, mu) fed with the uniforms u[0 .. nu), consumed in
// the reference's order (3 on the degenerate branch; else 4 per trial, then
// 2).  When they last: the new link is stored and also written to link_out,
// status[0] = trials (0: degenerate branch), status[1] = uniforms consumed.
// When they run out: nothing is stored, status[0] = -1 and status[1] = the
// number of uniforms the next attempt needs at least (the uniforms read so
// far plus those of the next step: a trial, the direction or the degenerate
// link), so a caller can draw exactly what the reference would draw.
template <typename T>
__global__ void __launch_bounds__(64) k_local_heatbath(Planes<T> P, Geo g, uint32_t n, int mu, double beta,
                                                        const double* __restrict__ u, int nu,
                                                        double* __restrict__ link_out, int* __restrict__ status)
{
    if (threadIdx.x != 0) return;
    const Site s = lex_site(g, n);
    const int p = parity(s);
    Q<double> A = widen(update_staple_mu(P, g, s, p, mu));
    BufSrc src{u, nu, 0};
    Q<double> Un;
    int trials = 0;
    double sqrt_det = qdet_sqrt(A);
    if (sqrt_det <= __DBL_EPSILON__) {
        if (nu < 3) {
            status[0] = -1;
            status[1] = 3;
            return;
        }
        Un = random_su2_64(src);
    } else {
        A = qdiv(A, sqrt_det);
        const double alpha = sqrt_det * beta;
        double delta = 0.0;
        bool accepted = false;
        while (!accepted) {
            if (trials == KP_MAX_TRIALS) break;
            if (src.pos + 4 > nu) {
                status[0] = -1;
                status[1] = src.pos + 4;
                return;
            }
            accepted = kp_trial64(alpha, src, delta);
            trials++;
        }
        const int rest = accepted ? 2 : 3;   // the direction, or a uniform link after KP_MAX_TRIALS
        if (src.pos + rest > nu) {
            status[0] = -1;
            status[1] = src.pos + rest;
            return;
        }
        Un = accepted ? qmul_dag(kp_finish64(delta, src), A) : random_su2_64(src);
    }
    const Q<T> st{(T)Un.c0, (T)Un.c1, (T)Un.c2, (T)Un.c3};
    qstore<T>(P.p[mu][p], cb_index(g, s), st);
    const Q<double> w = widen(st);
    link_out[0] = w.c0; link_out[1] = w.c1; link_out[2] = w.c2; link_out[3] = w.c3;
    status[0] = trials;
    status[1] = src.pos;
}

// ---------------------------------------------------------------------------
// plaquette (LGT_Code/modules/obs/plaquette.c:67-101)
// ---------------------------------------------------------------------------

// The plaquette is evaluated in FP64 for both storages: an FP32 field's
// plaquette is then the one the reference computes from the same links (as
// widened into a configuration file), so files written from FP32 storage pass
// readConfig's plaquette test (config_IO.c:609-614).
template <typename T>
__device__ __forceinline__ double plaq_term(const Planes<T>& P, const Geo& g, const Site& s, int jj, int kk)
{
#pragma clang fp contract(off)
    Q<double> un0 = widen(link(P, g, jj, s));
    Q<double> un1 = widen(link(P, g, kk, shift(g, s, jj, 1)));
    Q<double> un2 = qmul(un0, un1);
    un0 = qdag(widen(link(P, g, jj, shift(g, s, kk, 1))));
    un1 = qmul(un2, un0);
    un2 = qdag(widen(link(P, g, kk, s)));
    un0 = qmul(un1, un2);
    return 2.0 * un0.c0;
}

// The plaquette of a field in the reference's FP64 layout ([VOL][4][4], site
// major) as given, before it is rounded to the storage precision: the value
// readConfig compares with the file's (config_IO.c:609-614).  One thread per
// lexicographic site; partial[block].
__device__ __forceinline__ Q<double> aos_link(const double* __restrict__ aos, const Site& s, const Geo& g, int mu)
{
    const double* q = aos + 16 * (size_t)((((uint32_t)s.c[0] * g.L[1] + s.c[1]) * g.L[2] + s.c[2]) * (uint32_t)g.L[3] +
                                          (uint32_t)s.c[3]) + 4 * mu;
    return Q<double>{q[0], q[1], q[2], q[3]};
}

static __global__ void __launch_bounds__(BLOCK) k_plaquette_aos(const double* __restrict__ aos, Geo g,
                                                                 DD* __restrict__ partial)
{
#pragma clang fp contract(off)
    const uint32_t n = blockIdx.x * BLOCK + threadIdx.x;
    DD acc = DD{0.0, 0.0};
    if (n < 2 * g.vh) {
        const Site s = lex_site(g, n);
#pragma unroll
        for (int jj = 0; jj < 3; jj++)
#pragma unroll
            for (int kk = jj + 1; kk < 4; kk++) {
                Q<double> un0 = aos_link(aos, s, g, jj);
                Q<double> un1 = aos_link(aos, shift(g, s, jj, 1), g, kk);
                Q<double> un2 = qmul(un0, un1);
                un0 = qdag(aos_link(aos, shift(g, s, kk, 1), g, jj));
                un1 = qmul(un2, un0);
                un2 = qdag(aos_link(aos, s, g, kk));
                un0 = qmul(un1, un2);
                acc = dd_add(acc, 2.0 * un0.c0);
            }
    }
    acc = block_sum_dd(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// grid: x over (tile, parity) pairs; partial[p*tiles + tile]
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_plaquette(Planes<T> P, Geo g, DD* __restrict__ partial)
{
    int p;
    uint32_t base = tile_base_pair(g, p);
    uint32_t tile = base / BLOCK;
    uint32_t i = base + threadIdx.x;
    DD acc = DD{0.0, 0.0};
    if (i < g.vh) {
        Site s = cb_site(g, i, p);
#pragma unroll
        for (int jj = 0; jj < 3; jj++)
#pragma unroll
            for (int kk = jj + 1; kk < 4; kk++) acc = dd_add(acc, plaq_term(P, g, s, jj, kk));
    }
    acc = block_sum_dd(acc);
    if (threadIdx.x == 0) partial[p * (gridDim.x >> 1) + tile] = acc;
}

// ---------------------------------------------------------------------------
// APE smearing (LGT_Code/modules/obs/smearing.c:42-168), Jacobi form
//   U'_mu(n) = proj[ Z_mu(n) + sp * S_mu(n) ],  S from the step-start field Z
//   S = sum_{k != mu, dt} [Z_k(n) Z_mu(n+k) Z_k(n+mu)^+ + Z_k(n-k)^+ Z_mu(n-k) Z_k(n-k+mu)]
// MODE 0: spatial (mu = 1..3, dt = 0), 1: all (mu = 0..3, dt = mu),
// 2: temporal (mu = 0, dt = 0).
// ---------------------------------------------------------------------------
// Smeared links of site s (the directions MODE smears) from its +-1
// neighbours spn[b] = s + b, smn[b] = s - b.  All links of the site are
// computed together, so the gathers that several directions share (Z_k(n),
// Z_a(n+b), Z_a(n-b), Z_a(n-a+b)) are loaded once: 24 loads for the 3 spatial
// links instead of 36, 44 instead of 72 for all four.  Each staple is still
// summed in the reference's order (k ascending, forward then backward term).
// |M| = |Z + sp S| goes to nrm (for the adjoint).
template <typename T, int MODE>
__device__ __forceinline__ void ape_site(const Planes<T>& Z, const Geo& g, const Site& s, const Site* spn,
                                         const Site* smn, T sp, Q<T>* res, T* nrm)
{
    constexpr int MU0 = (MODE == 0) ? 1 : 0, MU1 = (MODE == 2) ? 0 : 3;
    // directions that enter the staples
    constexpr int K0 = (MODE == 1) ? 0 : 1;
    Q<T> z0[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
        if (a < K0 && !(a >= MU0 && a <= MU1)) continue;
        z0[a] = link(Z, g, a, s);
    }
#pragma unroll
    for (int mu = MU0; mu <= MU1; mu++) {
        Q<T> acc = qzero<T>();
        const int dt = (MODE == 1) ? mu : 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k == mu || k == dt) continue;
            // forward: Z_k(n) Z_mu(n+k) Z_k(n+mu)^+
            Q<T> un1 = qmul(z0[k], link(Z, g, mu, spn[k]));
            Q<T> un0 = qmul_dag(un1, link(Z, g, k, spn[mu]));
            Q<T> un2 = qadd(acc, un0);
            // backward: Z_k(n-k)^+ Z_mu(n-k) Z_k(n-k+mu)
            un0 = qdag(link(Z, g, k, smn[k]));
            un1 = qmul(un0, link(Z, g, mu, smn[k]));
            un0 = qmul(un1, link(Z, g, k, shift(g, smn[k], mu, 1)));
            acc = qadd(un2, un0);
        }
        Q<T> S = qscale(acc, sp);
        Q<T> M = qadd(z0[mu], S);
        nrm[mu] = qdet_sqrt(M);
        res[mu] = qproject(M);
    }
}

template <typename T> struct NormPlanes { T* p[4][2]; };

// One thread per site.  grid: x over (tile, parity) pairs.  With Mn's planes
// of the smeared directions set, |M| of every smeared link is stored too.
// FP64 spatial smearing puts both parities of a tile in one block
// (tile_site_mixed: 347 instead of 363 us at 48^4; FP32 changes by 1%, FP64
// 4D smearing runs 31% slower with it).
template <typename T, int MODE>
__global__ void __launch_bounds__(BLOCK) k_ape(Planes<T> Z, Planes<T> O, Geo g, T sp, NormPlanes<T> Mn)
{
    constexpr int MU0 = (MODE == 0) ? 1 : 0, MU1 = (MODE == 2) ? 0 : 3;
    constexpr bool MIXED = sizeof(T) == 8 && MODE == 0;
    int p;
    uint32_t i = MIXED ? tile_site_mixed(g, p) : tile_base_pair(g, p) + threadIdx.x;
    // p is the same for the whole wave; in a scalar register the output plane
    // pointers O.p[mu][p] come from scalar loads (a per-lane index makes them
    // vector loads, and the stores wait on every load in flight)
    p = __builtin_amdgcn_readfirstlane(p);
    if (i >= g.vh) return;
    const Site s = cb_site(g, i, p);
    Site spn[4], smn[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        spn[b] = shift(g, s, b, 1);
        smn[b] = shift(g, s, b, -1);
    }
    Q<T> res[4];
    T nrm[4];
    ape_site<T, MODE>(Z, g, s, spn, smn, sp, res, nrm);
    // stores last: no store separates the loads the directions share
#pragma unroll
    for (int mu = MU0; mu <= MU1; mu++) {
        if (Mn.p[MU0][0]) Mn.p[mu][p][i] = nrm[mu];
        qstore<T>(O.p[mu][p], i, res[mu]);
    }
}

// One thread per pair of sites n, n + z (z even), the two sites of one cb
// index: most links the two sites share (at n + z, at n, at n - b + z) are
// loaded once, 42 loads per pair in spatial smearing instead of 48.  Pays in
// FP32 spatial smearing only (147 instead of 162 us at 48^4): with FP64 links
// or all four directions the two sites' loads take up to 256 VGPRs and the
// kernel runs slower than k_ape.  (Passing precomputed link offsets instead
// of sites merges all 9 shared loads, 39 per pair, but runs 6% slower.)
// grid: x over tiles (both parities of a tile in one block).  O's parity
// planes must be adjacent (p[mu][1] == p[mu][0] + VOL/2).
template <typename T, int MODE>
__global__ void __launch_bounds__(BLOCK) k_ape_zpair(Planes<T> Z, Planes<T> O, Geo g, T sp, NormPlanes<T> Mn)
{
    constexpr int MU0 = (MODE == 0) ? 1 : 0, MU1 = (MODE == 2) ? 0 : 3;
    uint32_t i = tile_base(g) + threadIdx.x;
    if (i >= g.vh) return;
    Site s0 = cb_site(g, i, 0);
    s0.c[3] &= ~1;   // the even-z site of cb index i
    Site sp0[4], sm0[4], sp1[4], sm1[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        sp0[b] = shift(g, s0, b, 1);
        sm0[b] = shift(g, s0, b, -1);
    }
    // s1 = s0 + z and its neighbours, built from s0's so that the compiler
    // sees the shared links' addresses as equal
    const Site s1 = sp0[3];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        sp1[b] = shift(g, s1, b, 1);
        sm1[b] = (b == 3) ? s0 : shift(g, sm0[b], 3, 1);
    }
    Q<T> r0[4], r1[4];
    T m0[4], m1[4];
    ape_site<T, MODE>(Z, g, s0, sp0, sm0, sp, r0, m0);
    ape_site<T, MODE>(Z, g, s1, sp1, sm1, sp, r1, m1);
    const int q0 = parity(s0);
#pragma unroll
    for (int mu = MU0; mu <= MU1; mu++) {
        if (Mn.p[MU0][0]) {
            (q0 ? Mn.p[mu][1] : Mn.p[mu][0])[i] = m0[mu];
            (q0 ? Mn.p[mu][0] : Mn.p[mu][1])[i] = m1[mu];
        }
        qstore<T>(O.p[mu][0], i + (uint32_t)q0 * g.vh, r0[mu]);
        qstore<T>(O.p[mu][0], i + (uint32_t)(q0 ^ 1) * g.vh, r1[mu]);
    }
}

// ---------------------------------------------------------------------------
// Wilson loops W(R,T) in the (t, d) planes, d = 1..3 (LGT_Code/modules/obs/
// wilson.c:21-56,268-324), fused over all T of one R:
//   W(x) = Re tr[ L_T(x) S_R(x+T t) L_T(x+R d)^+ S_R(x)^+ ]
//        = Re tr[ M S_R(x+T t) ],  M = L_T(x+R d)^+ S_R(x)^+ L_T(x)
// with M carried from T to T by one link on each side (two products per T),
// and the spatial lines S_R read from a line field built beforehand
// (k_sline), one load per step instead of R.  Results per (group, d, parity,
// tile) are reduced in double-double; every wave adds its shuffle-reduced
// value to its own LDS slot.
// ---------------------------------------------------------------------------
struct WGroup { int R, Tmax, slot_off, pad; };
constexpr int WMAXT = 128;

// Line field: S.p[d] holds S_R(y) = U_d(y) U_d(y+d) ... U_d(y+(R-1)d) for
// d = 1..3, the product taken left to right as wilson.c:21-56 builds the
// line.  Advances it in place from R0 (0: empty) to R1 > R0 (each site reads
// and writes only its own line).
// grid: x over (tile, parity) pairs, y = d - 1
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_sline(Planes<T> P, Planes<T> S, Geo g, int R0, int R1)
{
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    const int d = 1 + (int)blockIdx.y;
    const Site y = cb_site(g, i, p);
    Q<T> line;
    int j = R0;
    if (R0 == 0) {
        line = link(P, g, d, y);
        j = 1;
    } else {
        line = qload<T>(S.p[d][p], i);
    }
    Site yj = shift(g, y, d, j % g.L[d]);
    for (; j < R1; j++) {
        line = qmul(line, link(P, g, d, yj));
        yj = shift(g, yj, d, 1);
    }
    typename Store<T>::V v;
    v.x = line.c0; v.y = line.c1; v.z = line.c2; v.w = line.c3;
    S.p[d][p][i] = v;   // read back by the loops right after: a cached store
}

// Wave sums of eight per-lane values at once.  Three exchange stages (lane
// masks 32, 16, 8) each keep the half of a lane's values that its lane bit
// chooses and add the partner's copy of that half; three more (4, 2, 1) add
// the lanes that hold the same value.  Lane l ends with the double-double sum
// over the wave of v[(l >> 3) & 7], in a fixed order.  One wave_sum_dd per
// value costs six double-double additions and shuffles each; this costs ten
// for all eight.
__device__ __forceinline__ DD shfl_xor_dd(const DD& v, int m)
{
    return DD{__shfl_xor(v.hi, m), __shfl_xor(v.lo, m)};
}

__device__ __forceinline__ DD wave_sum8_dd(const double (&v)[8], int lane)
{
    const bool b5 = (lane & 32) != 0, b4 = (lane & 16) != 0, b3 = (lane & 8) != 0;
    DD u[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double mine = b5 ? v[4 + k] : v[k];
        const double other = __shfl_xor(b5 ? v[k] : v[4 + k], 32);
        u[k] = dd_add(DD{mine, 0.0}, other);
    }
    DD w[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const DD mine = b4 ? u[2 + k] : u[k];
        w[k] = dd_add(mine, shfl_xor_dd(b4 ? u[k] : u[2 + k], 16));
    }
    DD x = dd_add(b3 ? w[1] : w[0], shfl_xor_dd(b3 ? w[0] : w[1], 8));
#pragma unroll
    for (int m = 4; m > 0; m >>= 1) x = dd_add(x, shfl_xor_dd(x, m));
    return x;
}

// One group (one R, all its T) per launch; S holds the lines S_R.  With
// Rn > 0 the launch also writes the next group's lines S_Rn of its own sites
// into Sn (a second line field: S is still read by other blocks), so the line
// field is built by k_sline only before the first group.  A block
// covers one 256-site tile with WL<T>::THREADS lanes, each taking WL<T>::SITES
// sites of it (stride THREADS): the lane sums its sites' values per T in
// double, and the wave sums of eight T at a time are taken together
// (wave_sum8_dd).
// grid: x over (tile, parity) pairs, y = d - 1; WL<T>::THREADS per block
template <typename T> struct WL {
    // sites per lane; 48^4, 160 loops with the batched wave sums and one
    // carried product per site: FP32 1, 2, 4 sites -> 9.0, 7.0, 8.4 ms;
    // FP64 -> 13.1, 13.2, 18.9 ms
    static constexpr int SITES = 2;
    static constexpr int THREADS = BLOCK / SITES;
    static constexpr int NW = THREADS / 64;
};

template <typename T>
__global__ void __launch_bounds__(WL<T>::THREADS) k_wloops(Planes<T> P, Planes<T> S, Geo g,
                                                          const WGroup* __restrict__ groups, int group,
                                                          const int* __restrict__ slots, DD* __restrict__ partial,
                                                          int ncols, Planes<T> Sn, int Rn)
{
    constexpr int NS = WL<T>::SITES, NT = WL<T>::THREADS, NW = WL<T>::NW;
    __shared__ DD wacc[NW][WMAXT];
    int p;
    const uint32_t base = tile_base_pair(g, p);
    const uint32_t tile = base / BLOCK;
    const WGroup grp = groups[group];
    const int d = 1 + (int)blockIdx.y;
    const int R = grp.R;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // the lines' top ends x + t0 t and x + R d + t0 t as time coordinate and cb
    // index, advanced one time step per T (t0: their time offset, mod L0)
    int tT[NS], tR[NS];
    uint32_t cbT[NS], cbR[NS];
    // M = B^+ L_T(x) with B = S_R(x) L_T(x + R d): the loop is
    // L_T(x) S_R(x + T t) B^+, so Re tr(loop) = Re tr(M S_R(x + T t)), and one
    // step in T is M <- U_0(x + R d + T t)^+ M U_0(x + T t)
    Q<T> M[NS];
    bool valid[NS];
    const int pR = p ^ (R & 1);   // parity of x + R d
    const uint32_t st0 = g.stride[0], wrap0 = (uint32_t)(g.L[0] - 1) * st0;
    constexpr uint32_t LB = (uint32_t)sizeof(typename Store<T>::V);
#pragma unroll
    for (int j = 0; j < NS; j++) {
        const uint32_t i = base + (uint32_t)(threadIdx.x + NT * j);
        valid[j] = i < g.vh;
        const Site s0 = cb_site(g, valid[j] ? i : 0u, p);
        const Site sR = shift(g, s0, d, R % g.L[d]);
        tT[j] = s0.c[0];
        cbT[j] = valid[j] ? i : 0u;
        tR[j] = sR.c[0];
        cbR[j] = cb_index(g, sR);
        const Q<T> S0 = link(S, g, d, s0);
        M[j] = qmul_c(qdag(qmul(S0, link(P, g, 0, sR))), link(P, g, 0, s0));
        if (Rn > 0 && valid[j]) {
            // the next group's line S_Rn(x) = S_R(x) U_d(x+Rd) ... U_d(x+(Rn-1)d),
            // the products in k_sline's order
            Q<T> line = S0;
            Site y = sR;
            for (int r = R; r < Rn; r++) {
                line = qmul(line, link(P, g, d, y));
                y = shift(g, y, d, 1);
            }
            typename Store<T>::V v;
            v.x = line.c0; v.y = line.c1; v.z = line.c2; v.w = line.c3;
            Sn.p[d][p][cbT[j]] = v;   // read by the next launch: a cached store
        }
    }
    int t0 = 0;
    // eight T per batch; the lanes' sums of a batch are reduced together
    for (int Tb = 1; Tb <= grp.Tmax; Tb += 8) {
        double acc[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int Tt = Tb + k;
            acc[k] = 0.0;
            if (Tt > grp.Tmax) continue;
            const int slot = slots[grp.slot_off + Tt - 1];
            const uint32_t odd = (uint32_t)(t0 & 1);
#pragma unroll
            for (int j = 0; j < NS; j++) {
                if (Tt > 1) {
                    const Q<T> Ub = link_at(P, 0, (((uint32_t)pR ^ odd) * g.vh + cbR[j]) * LB);
                    M[j] = qmul_c(qmul_c(qdag(Ub), M[j]), link_at(P, 0, (((uint32_t)p ^ odd) * g.vh + cbT[j]) * LB));
                }
                // one step in t
                cbT[j] += (tT[j] + 1 == g.L[0]) ? (uint32_t)0 - wrap0 : st0;
                tT[j] = (tT[j] + 1 == g.L[0]) ? 0 : tT[j] + 1;
                cbR[j] += (tR[j] + 1 == g.L[0]) ? (uint32_t)0 - wrap0 : st0;
                tR[j] = (tR[j] + 1 == g.L[0]) ? 0 : tR[j] + 1;
                if (slot >= 0) {
                    const Q<T> St = link_at(S, d, (((uint32_t)p ^ odd ^ 1u) * g.vh + cbT[j]) * LB);
                    const T w = T(2) * qdot_c(M[j], qdag(St));   // 2 Re[M St]_0
                    acc[k] += valid[j] ? (double)w : 0.0;
                }
            }
            t0 = (t0 + 1 == g.L[0]) ? 0 : t0 + 1;
        }
        const DD v = wave_sum8_dd(acc, lane);
        const int Tt = Tb + (lane >> 3);
        if ((lane & 7) == 0 && Tt <= grp.Tmax) wacc[wid][Tt - 1] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < grp.Tmax; t += NT) {
        const int slot = slots[grp.slot_off + t];
        if (slot < 0) continue;
        DD a = wacc[0][t];
        for (int w = 1; w < NW; w++) a = dd_add(a, wacc[w][t]);
        const int col = ((d - 1) * 2 + p) * (gridDim.x >> 1) + tile;
        partial[(size_t)slot * ncols + col] = a;
    }
}

// ---------------------------------------------------------------------------
// topological charges (LGT_Code/modules/obs/topologicalcharge.c): the 1x1
// clover here, the improved clover from line fields below
// ---------------------------------------------------------------------------
// 1x1 leaves (topologicalcharge.c:201-290)
template <typename T>
__device__ __forceinline__ Q<T> U_plpl(const Planes<T>& P, const Geo& g, const Site& x, int jj, int kk)
{
    Q<T> un0 = link(P, g, jj, x);
    Q<T> un1 = link(P, g, kk, shift(g, x, jj, 1));
    Q<T> un2 = qmul(un0, un1);
    un0 = qdag(link(P, g, jj, shift(g, x, kk, 1)));
    un1 = qmul(un2, un0);
    un2 = qdag(link(P, g, kk, x));
    return qmul(un1, un2);
}

template <typename T>
__device__ __forceinline__ Q<T> U_mipl(const Planes<T>& P, const Geo& g, const Site& x, int jj, int kk)
{
    Site n = shift(g, x, jj, -1);
    Q<T> un0 = qdag(link(P, g, jj, n));
    Q<T> un1 = link(P, g, kk, n);
    Q<T> un2 = qmul(un0, un1);
    n = shift(g, n, kk, 1);
    un1 = qmul(un2, link(P, g, jj, n));
    un2 = qdag(link(P, g, kk, x));
    return qmul(un1, un2);
}

template <typename T>
__device__ __forceinline__ Q<T> U_mimi(const Planes<T>& P, const Geo& g, const Site& x, int jj, int kk)
{
    Site n = shift(g, x, jj, -1);
    Q<T> un0 = qdag(link(P, g, jj, n));
    n = shift(g, n, kk, -1);
    Q<T> un2 = qmul_dag(un0, link(P, g, kk, n));
    Q<T> un1 = qmul(un2, link(P, g, jj, n));
    n = shift(g, x, kk, -1);
    return qmul(un1, link(P, g, kk, n));
}

template <typename T>
__device__ __forceinline__ Q<T> U_plmi(const Planes<T>& P, const Geo& g, const Site& x, int jj, int kk)
{
    Q<T> un0 = link(P, g, jj, x);
    Site n = shift(g, shift(g, x, jj, 1), kk, -1);
    Q<T> un2 = qmul_dag(un0, link(P, g, kk, n));
    n = shift(g, x, kk, -1);
    un0 = qdag(link(P, g, jj, n));
    Q<T> un1 = qmul(un2, un0);
    return qmul(un1, link(P, g, kk, n));
}

template <typename T>
__device__ __forceinline__ Q<T> clover_1x1(const Planes<T>& P, const Geo& g, const Site& x, int d1, int d2)
{
    Q<T> p = U_plpl(P, g, x, d1, d2);
    p = qadd(U_mipl(P, g, x, d2, d1), p);
    p = qadd(U_mimi(P, g, x, d1, d2), p);
    p = qadd(U_plmi(P, g, x, d2, d1), p);
    return qdiv(p, T(4));
}

template <typename T>
__device__ __forceinline__ T clover_products_1x1(const Planes<T>& P, const Geo& g, const Site& x, int d1, int d2,
                                                 int d3, int d4)
{
    Q<T> C = clover_1x1(P, g, x, d1, d2);
    Q<T> D = clover_1x1(P, g, x, d3, d4);
    Q<T> Cd = qdag(C), Dd = qdag(D);
    Q<T> P0{C.c0 - Cd.c0, C.c1 - Cd.c1, C.c2 - Cd.c2, C.c3 - Cd.c3};
    Q<T> P1{D.c0 - Dd.c0, D.c1 - Dd.c1, D.c2 - Dd.c2, D.c3 - Dd.c3};
    P0 = qdiv(P0, T(2));
    P1 = qdiv(P1, T(2));
    return T(2) * qmul(P0, P1).c0;
}

// ---------------------------------------------------------------------------
// The improved clover from straight-line fields.  A leaf of side m is four
// straight lines of m links, Lm_mu(y) = U_mu(y) U_mu(y+mu) ... U_mu(y+(m-1)mu):
// three products of four loaded lines instead of 4m - 1 products of 4m
// loaded links, and the four leaves of one (plane, m) share their lines: 12
// loads, 216 per site for the six F, against 576 link loads walking the
// leaves link by link.  The lines of two and three links are built once per
// charge (k_qlines).  The reference multiplies each leaf's 4m links
// sequentially (computeLine, topologicalcharge.c:18-44); grouping them by
// lines changes the rounding of q(x) by a few units in the last place.
// ---------------------------------------------------------------------------
// L2_mu(x) = U_mu(x) U_mu(x+mu), L3_mu(x) = L2_mu(x) U_mu(x+2mu).  One thread
// per pair of sites x, x+mu with x_mu even: the four links U_mu(x .. x+3mu)
// serve the lines of both sites (two loads per site instead of three).
// Threads run over the lattice halved along mu, z fastest, so a wave covers
// consecutive sites of one row.  grid: x over VOL/2 threads, y = mu
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_qlines(Planes<T> U, Planes<T> L2, Planes<T> L3, Geo g)
{
    const uint32_t j = blockIdx.x * BLOCK + threadIdx.x;
    const int mu = blockIdx.y;
    if (j >= g.vh) return;
    int h[4];
    for (int d = 0; d < 4; d++) h[d] = (d == mu) ? g.L[d] / 2 : g.L[d];
    Site s;
    uint32_t r = j;
    s.c[3] = (int)(r % (uint32_t)h[3]);
    r /= (uint32_t)h[3];
    s.c[2] = (int)(r % (uint32_t)h[2]);
    r /= (uint32_t)h[2];
    s.c[1] = (int)(r % (uint32_t)h[1]);
    s.c[0] = (int)(r / (uint32_t)h[1]);
    s.c[mu] *= 2;
    const Site s1 = shift(g, s, mu, 1), s2 = shift(g, s1, mu, 1), s3 = shift(g, s2, mu, 1);
    const Q<T> u0 = link(U, g, mu, s), u1 = link(U, g, mu, s1), u2 = link(U, g, mu, s2), u3 = link(U, g, mu, s3);
    const int p = parity(s);
    const uint32_t i0 = cb_index(g, s), i1 = cb_index(g, s1);
    const Q<T> a = qmul(u0, u1), b = qmul(u1, u2);
    // both parity planes from p[mu][0] (adjacent planes), not a plane pointer
    // chosen by the lane's parity (read from the kernel arguments per lane)
    const uint32_t o0 = (uint32_t)p * g.vh + i0, o1 = (uint32_t)(p ^ 1) * g.vh + i1;
    qstore<T>(L2.p[mu][0], o0, a);
    qstore<T>(L3.p[mu][0], o0, qmul(a, u2));
    qstore<T>(L2.p[mu][0], o1, b);
    qstore<T>(L3.p[mu][0], o1, qmul(b, u3));
}

// the part of the cb index that coordinate c along D contributes (z: the cb
// index holds z/2)
template <int D>
__device__ __forceinline__ uint32_t cb_part(const Geo& g, int c)
{
    return (D == 3) ? (uint32_t)(c >> 1) : (uint32_t)c * g.stride[D];
}

// c mod L[D] for -2 L < c < 2 L: two wraps, since a shift of m <= 3 links
// exceeds one period on lattices of extent 2
template <int D>
__device__ __forceinline__ int wrap_coord(const Geo& g, int c)
{
#pragma unroll
    for (int k = 0; k < 2; k++) {
        c = (c >= g.L[D]) ? c - g.L[D] : c;
        c = (c < 0) ? c + g.L[D] : c;
    }
    return c;
}

// The vector part (c1, c2, c3) of a quaternion: all that the field strength
// keeps.  AH(X) = (X - X^+)/2 (topologicalcharge.c:112-121) zeroes c0 and
// returns the vector part of X exactly ((c + c)/2 = c), so the clover sums,
// their scalings and the last product of each leaf form only c1..c3, each
// with the reference's operations, and Re tr(F G) = 2 (F G)_0 = -2 f.g.
template <typename T> struct V3 { T c1, c2, c3; };

// vector part of v w (gauge.h:83-91)
template <typename T>
__device__ __forceinline__ V3<T> vmul(const Q<T>& v, const Q<T>& w)
{
    return V3<T>{v.c0 * w.c1 + v.c1 * w.c0 - v.c2 * w.c3 + v.c3 * w.c2,
                 v.c0 * w.c2 + v.c1 * w.c3 + v.c2 * w.c0 - v.c3 * w.c1,
                 v.c0 * w.c3 - v.c1 * w.c2 + v.c2 * w.c1 + v.c3 * w.c0};
}

// vector part of v w^+ (gauge.h:94-102)
template <typename T>
__device__ __forceinline__ V3<T> vmul_dag(const Q<T>& v, const Q<T>& w)
{
    return V3<T>{-v.c0 * w.c1 + v.c1 * w.c0 + v.c2 * w.c3 - v.c3 * w.c2,
                 -v.c0 * w.c2 - v.c1 * w.c3 + v.c2 * w.c0 + v.c3 * w.c1,
                 -v.c0 * w.c3 + v.c1 * w.c2 - v.c2 * w.c1 + v.c3 * w.c0};
}

template <typename T>
__device__ __forceinline__ V3<T> vadd(const V3<T>& a, const V3<T>& b) { return V3<T>{a.c1 + b.c1, a.c2 + b.c2, a.c3 + b.c3}; }

template <typename T>
__device__ __forceinline__ void pin(V3<T>& v)
{
    asm volatile("" : "+v"(v.c1), "+v"(v.c2), "+v"(v.c3));
}

// Index offsets of a site x along each direction d: moving by a*m links along
// d (a = -1, 0, +1) adds delta[d][a+1] to the cb index (0 for a = 0); the
// parity changes iff m is odd.
struct LineGeo {
    int32_t delta[4][3];
};

template <int D>
__device__ __forceinline__ void line_offsets(const Geo& g, const Site& s, int m, LineGeo& lg)
{
    const int c = s.c[D];
    const uint32_t here = cb_part<D>(g, c);
    lg.delta[D][0] = (int32_t)(cb_part<D>(g, wrap_coord<D>(g, c - m)) - here);
    lg.delta[D][1] = 0;
    lg.delta[D][2] = (int32_t)(cb_part<D>(g, wrap_coord<D>(g, c + m)) - here);
}

// the line along `dir` at x + (ia-1) m D1 + (ib-1) m D2
template <typename T, int D1, int D2>
__device__ __forceinline__ Q<T> line_at(const Planes<T>& Lm, const Geo& g, const LineGeo& lg, uint32_t i, int p,
                                        int odd, int dir, int ia, int ib)
{
    const uint32_t par = (uint32_t)(p ^ ((ia + ib) & odd));
    const uint32_t cb = i + (uint32_t)lg.delta[D1][ia] + (uint32_t)lg.delta[D2][ib];
    return link_at(Lm, dir, (par * g.vh + cb) * (uint32_t)sizeof(typename Store<T>::V));
}

// computeCloverAverages(u, x, D1, D2, m, m) (topologicalcharge.c:46-110) from
// the lines Lm of side m (m = 1: the links): the four leaves in the
// reference's order and orientation, summed in its order, / 4 (vector part).
// S[d] = Lm_d(x) and B[d] = Lm_d(x - m d), which every plane containing d
// uses, are loaded once per site and m by the caller.
template <typename T, int D1, int D2>
__device__ __forceinline__ V3<T> clover_lines(const Planes<T>& Lm, const Geo& g, const LineGeo& lg, uint32_t i,
                                              int p, int odd, const Q<T>* S, const Q<T>* B)
{
    auto ld = [&](int dir, int ia, int ib) { return line_at<T, D1, D2>(Lm, g, lg, i, p, odd, dir, ia, ib); };
    Q<T> t;
    V3<T> sum;
    // +D1,+D2: L1(x) L2(x+m D1) L1(x+m D2)^+ L2(x)^+
    t = qmul(S[D1], ld(D2, 2, 1));
    t = qmul_dag(t, ld(D1, 1, 2));
    sum = vmul_dag(t, S[D2]);
    // +D1,-D2: -D2, +D1, +D2, -D1
    t = qmul(qdag(B[D2]), ld(D1, 1, 0));
    t = qmul(t, ld(D2, 2, 0));
    sum = vadd(sum, vmul_dag(t, S[D1]));
    // -D1,+D2: +D2, -D1, -D2, +D1
    t = qmul_dag(S[D2], ld(D1, 0, 2));
    t = qmul_dag(t, ld(D2, 0, 1));
    sum = vadd(sum, vmul(t, B[D1]));
    // -D1,-D2: -D1, -D2, +D1, +D2
    t = qmul_dag(qdag(B[D1]), ld(D2, 0, 0));
    t = qmul(t, ld(D1, 0, 0));
    sum = vadd(sum, vmul(t, B[D2]));
    return V3<T>{sum.c1 / T(4), sum.c2 / T(4), sum.c3 / T(4)};
}

// F += k C_m for the six planes of the charge, in the order (0,1), (2,3),
// (0,2), (3,1), (0,3), (1,2); m = 1 starts the sums (F_{D1 D2} =
// AH(1.5 C1 - 0.15 C2 + C3/90), topologicalcharge.c:123-162, summed in its order)
template <typename T>
__device__ __forceinline__ void add_cloverlines(const Planes<T>& Lm, const Geo& g, const Site& s, uint32_t i, int p,
                                                int m, T k, V3<T>* F)
{
    LineGeo lg;
    line_offsets<0>(g, s, m, lg);
    line_offsets<1>(g, s, m, lg);
    line_offsets<2>(g, s, m, lg);
    line_offsets<3>(g, s, m, lg);
    const int odd = m & 1;
    const uint32_t sz = (uint32_t)sizeof(typename Store<T>::V);
    Q<T> S[4], B[4];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        S[d] = link_at(Lm, d, ((uint32_t)p * g.vh + i) * sz);
        B[d] = link_at(Lm, d, ((uint32_t)(p ^ odd) * g.vh + i + (uint32_t)lg.delta[d][0]) * sz);
    }
    auto acc = [&](int f, const V3<T>& C) {
        const V3<T> u{C.c1 * k, C.c2 * k, C.c3 * k};
        F[f] = (m == 1) ? u : vadd(F[f], u);
        pin(F[f]);
    };
    acc(0, clover_lines<T, 0, 1>(Lm, g, lg, i, p, odd, S, B));
    acc(1, clover_lines<T, 2, 3>(Lm, g, lg, i, p, odd, S, B));
    acc(2, clover_lines<T, 0, 2>(Lm, g, lg, i, p, odd, S, B));
    acc(3, clover_lines<T, 3, 1>(Lm, g, lg, i, p, odd, S, B));
    acc(4, clover_lines<T, 0, 3>(Lm, g, lg, i, p, odd, S, B));
    acc(5, clover_lines<T, 1, 2>(Lm, g, lg, i, p, odd, S, B));
}

// 2 (F G)_0 of two field strengths with c0 = 0 (gauge.h:83-91, 177-178)
template <typename T>
__device__ __forceinline__ T tr_FF(const V3<T>& f, const V3<T>& g)
{
    return T(2) * (-f.c1 * g.c1 - f.c2 * g.c2 - f.c3 * g.c3);
}

// the improved charge density from the line fields; grid and partials as k_qtop
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_qtop_lines(Planes<T> P, Planes<T> L2, Planes<T> L3, Geo g,
                                                      DD* __restrict__ partial, double* __restrict__ qx)
{
    int p;
    uint32_t base = tile_base_pair(g, p);
    uint32_t tile = base / BLOCK;
    uint32_t i = base + threadIdx.x;
    double q = 0.0;
    if (i < g.vh) {
        Site s = cb_site(g, i, p);
        V3<T> F[6];
        add_cloverlines(P, g, s, i, p, 1, T(1.5), F);
        pin(s);
        add_cloverlines(L2, g, s, i, p, 2, T(-0.15), F);
        pin(s);
        add_cloverlines(L3, g, s, i, p, 3, T(1.) / T(90.), F);
        T v = T(8.) * tr_FF(F[0], F[1]);
        v = v + T(8.) * tr_FF(F[2], F[3]);
        v = v + T(8.) * tr_FF(F[4], F[5]);
        q = (double)v;
        if (qx) qx[lex_index(g, s)] = q;
    }
    DD a = block_sum_dd(DD{q, 0.0});
    DD b = block_sum_dd(DD{fabs(q), 0.0});
    if (threadIdx.x == 0) {
        size_t k = (size_t)p * (gridDim.x >> 1) + tile;
        partial[k] = a;
        partial[(size_t)gridDim.x + k] = b;
    }
}

// the 1x1 clover charge density, -cp(0132) - cp(1230) - cp(2031)
// (topologicalcharge.c:391-409).  grid: x over (tile, parity) pairs.
// partial[k] = sum q and partial[ncols + k] = sum |q| with k = p*tiles + tile,
// ncols = 2*tiles; qx: density in lexicographic order (nullable).
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_qtop_clover(Planes<T> P, Geo g, DD* __restrict__ partial,
                                                       double* __restrict__ qx)
{
    int p;
    uint32_t base = tile_base_pair(g, p);
    uint32_t tile = base / BLOCK;
    uint32_t i = base + threadIdx.x;
    double q = 0.0;
    if (i < g.vh) {
        Site s = cb_site(g, i, p);
        const T v = -clover_products_1x1(P, g, s, 0, 1, 3, 2) - clover_products_1x1(P, g, s, 1, 2, 3, 0) -
                    clover_products_1x1(P, g, s, 2, 0, 3, 1);
        q = (double)v;
        if (qx) qx[lex_index(g, s)] = q;
    }
    DD a = block_sum_dd(DD{q, 0.0});
    DD b = block_sum_dd(DD{fabs(q), 0.0});
    if (threadIdx.x == 0) {
        size_t k = (size_t)p * (gridDim.x >> 1) + tile;
        partial[k] = a;
        partial[(size_t)gridDim.x + k] = b;
    }
}

}  // namespace lgt
