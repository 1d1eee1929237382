// su2_device.hpp -- SU(2) quaternion algebra, checkerboard geometry and the
// counter-based RNG shared by every kernel of the engine.
//
// Quaternion convention of the reference: U = c0 + i(c1 s1 + c2 s2 + c3 s3),
// Re tr U = 2 c0 (LGT_Code/include/gauge.h:21-24,177-178,190-198).  Products are
// written with the reference's left-to-right term order (gauge.h:83-102) so
// that the FP64 build, compiled without FMA contraction, rounds exactly like
// the reference on the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lgt {

// ---------------------------------------------------------------------------
// storage vectors: one link = one 16 B (FP32) or 32 B (FP64) vector
// ---------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

template <typename T> struct Store;
template <> struct Store<float> { typedef f4v V; };
template <> struct Store<double> { typedef d4v V; };

template <typename T> struct Q { T c0, c1, c2, c3; };

template <typename T>
__device__ __forceinline__ Q<T> qload(const typename Store<T>::V* __restrict__ p, uint32_t i)
{
    typename Store<T>::V v = p[i];
    return Q<T>{v.x, v.y, v.z, v.w};
}

// Links are stored with the nontemporal hint: no kernel reads back what it
// writes, so the written lines need not displace the gathered neighbours in L2
// (measured on MI355X at 48^4, FP32: heatbath -5 %, APE -6 %).
template <typename T>
__device__ __forceinline__ void qstore(typename Store<T>::V* __restrict__ p, uint32_t i, const Q<T>& q)
{
    typename Store<T>::V v;
    v.x = q.c0; v.y = q.c1; v.z = q.c2; v.w = q.c3;
    __builtin_nontemporal_store(v, &p[i]);
}

// The same with a plain store, for kernels that write a field without reading
// it (cold and hot start): FP64 links are two 16-byte stores per lane, each
// covering half of every cache line, and with the nontemporal hint those
// half-line writes to lines not in L2 ran at a third of the copy rate
// (48^4 FP64 cold start 320 us, 105 us with plain stores; FP32 unchanged).
template <typename T>
__device__ __forceinline__ void qstore_fill(typename Store<T>::V* __restrict__ p, uint32_t i, const Q<T>& q)
{
    typename Store<T>::V v;
    v.x = q.c0; v.y = q.c1; v.z = q.c2; v.w = q.c3;
    p[i] = v;
}

// v*w  (gauge.h:83-91)
template <typename T>
__device__ __forceinline__ Q<T> qmul(const Q<T>& v, const Q<T>& w)
{
    Q<T> o;
    o.c0 = v.c0 * w.c0 - v.c1 * w.c1 - v.c2 * w.c2 - v.c3 * w.c3;
    o.c1 = v.c0 * w.c1 + v.c1 * w.c0 - v.c2 * w.c3 + v.c3 * w.c2;
    o.c2 = v.c0 * w.c2 + v.c1 * w.c3 + v.c2 * w.c0 - v.c3 * w.c1;
    o.c3 = v.c0 * w.c3 - v.c1 * w.c2 + v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

// v*w^+  (gauge.h:94-102)
template <typename T>
__device__ __forceinline__ Q<T> qmul_dag(const Q<T>& v, const Q<T>& w)
{
    Q<T> o;
    o.c0 = v.c0 * w.c0 + v.c1 * w.c1 + v.c2 * w.c2 + v.c3 * w.c3;
    o.c1 = -v.c0 * w.c1 + v.c1 * w.c0 + v.c2 * w.c3 - v.c3 * w.c2;
    o.c2 = -v.c0 * w.c2 - v.c1 * w.c3 + v.c2 * w.c0 + v.c3 * w.c1;
    o.c3 = -v.c0 * w.c3 + v.c1 * w.c2 - v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

// Products for code with no reference rounding to reproduce (the adjoint of
// the smearing, the Wilson loops' carried product): FMA contraction in either
// precision.  v*w, v*w^+, <a,b>, and a + v*w, a + v*w^+ with the sum in each
// component's FMA chain.
template <typename T>
__device__ __forceinline__ Q<T> qmul_c(const Q<T>& v, const Q<T>& w)
{
#pragma clang fp contract(fast)
    Q<T> o;
    o.c0 = v.c0 * w.c0 - v.c1 * w.c1 - v.c2 * w.c2 - v.c3 * w.c3;
    o.c1 = v.c0 * w.c1 + v.c1 * w.c0 - v.c2 * w.c3 + v.c3 * w.c2;
    o.c2 = v.c0 * w.c2 + v.c1 * w.c3 + v.c2 * w.c0 - v.c3 * w.c1;
    o.c3 = v.c0 * w.c3 - v.c1 * w.c2 + v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

template <typename T>
__device__ __forceinline__ Q<T> qmul_dag_c(const Q<T>& v, const Q<T>& w)
{
#pragma clang fp contract(fast)
    Q<T> o;
    o.c0 = v.c0 * w.c0 + v.c1 * w.c1 + v.c2 * w.c2 + v.c3 * w.c3;
    o.c1 = -v.c0 * w.c1 + v.c1 * w.c0 + v.c2 * w.c3 - v.c3 * w.c2;
    o.c2 = -v.c0 * w.c2 - v.c1 * w.c3 + v.c2 * w.c0 + v.c3 * w.c1;
    o.c3 = -v.c0 * w.c3 + v.c1 * w.c2 - v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

template <typename T>
__device__ __forceinline__ T qdot_c(const Q<T>& a, const Q<T>& b)
{
#pragma clang fp contract(fast)
    return a.c0 * b.c0 + a.c1 * b.c1 + a.c2 * b.c2 + a.c3 * b.c3;
}

template <typename T>
__device__ __forceinline__ Q<T> qmul_add(const Q<T>& a, const Q<T>& v, const Q<T>& w)
{
#pragma clang fp contract(fast)
    Q<T> o;
    o.c0 = a.c0 + v.c0 * w.c0 - v.c1 * w.c1 - v.c2 * w.c2 - v.c3 * w.c3;
    o.c1 = a.c1 + v.c0 * w.c1 + v.c1 * w.c0 - v.c2 * w.c3 + v.c3 * w.c2;
    o.c2 = a.c2 + v.c0 * w.c2 + v.c1 * w.c3 + v.c2 * w.c0 - v.c3 * w.c1;
    o.c3 = a.c3 + v.c0 * w.c3 - v.c1 * w.c2 + v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

template <typename T>
__device__ __forceinline__ Q<T> qmul_dag_add(const Q<T>& a, const Q<T>& v, const Q<T>& w)
{
#pragma clang fp contract(fast)
    Q<T> o;
    o.c0 = a.c0 + v.c0 * w.c0 + v.c1 * w.c1 + v.c2 * w.c2 + v.c3 * w.c3;
    o.c1 = a.c1 - v.c0 * w.c1 + v.c1 * w.c0 + v.c2 * w.c3 - v.c3 * w.c2;
    o.c2 = a.c2 - v.c0 * w.c2 - v.c1 * w.c3 + v.c2 * w.c0 + v.c3 * w.c1;
    o.c3 = a.c3 - v.c0 * w.c3 + v.c1 * w.c2 - v.c2 * w.c1 + v.c3 * w.c0;
    return o;
}

// v^+ * w, written as the reference writes it: dagger first, then product
template <typename T>
__device__ __forceinline__ Q<T> qdag(const Q<T>& v) { return Q<T>{v.c0, -v.c1, -v.c2, -v.c3}; }

template <typename T>
__device__ __forceinline__ Q<T> qadd(const Q<T>& a, const Q<T>& b)
{
    return Q<T>{a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2, a.c3 + b.c3};
}

template <typename T>
__device__ __forceinline__ Q<T> qscale(const Q<T>& a, T s) { return Q<T>{a.c0 * s, a.c1 * s, a.c2 * s, a.c3 * s}; }

template <typename T>
__device__ __forceinline__ Q<T> qdiv(const Q<T>& a, T s) { return Q<T>{a.c0 / s, a.c1 / s, a.c2 / s, a.c3 / s}; }

template <typename T> __device__ __forceinline__ Q<T> qunit() { return Q<T>{T(1), T(0), T(0), T(0)}; }
template <typename T> __device__ __forceinline__ Q<T> qzero() { return Q<T>{T(0), T(0), T(0), T(0)}; }

// Euclidean 4-vector dot product = Re tr(a b^+)/2
template <typename T>
__device__ __forceinline__ T qdot(const Q<T>& a, const Q<T>& b)
{
    return a.c0 * b.c0 + a.c1 * b.c1 + a.c2 * b.c2 + a.c3 * b.c3;
}

// sqrt(|det v|) accumulated as in gauge.h:147-152
template <typename T>
__device__ __forceinline__ T qdet_sqrt(const Q<T>& v)
{
    T u = v.c0 * v.c0;
    u += v.c1 * v.c1;
    u += v.c2 * v.c2;
    u += v.c3 * v.c3;
    return sqrt(fabs(u));
}

// project_to_su2 (LGT_Code/modules/admin/sun_vfunc.c:69-79): U / sqrt(det U).
// The FP64 engine divides like the reference; FP32 storage multiplies by the
// reciprocal square root.
template <typename T> __device__ __forceinline__ Q<T> qproject(const Q<T>& u);
template <> __device__ __forceinline__ Q<double> qproject(const Q<double>& u)
{
    double zw = qdet_sqrt(u);
    return qdiv(u, zw);
}
template <> __device__ __forceinline__ Q<float> qproject(const Q<float>& u)
{
    float n2 = u.c0 * u.c0 + u.c1 * u.c1 + u.c2 * u.c2 + u.c3 * u.c3;
    return qscale(u, rsqrtf(n2));
}

// ---------------------------------------------------------------------------
// checkerboard geometry
//
// Site (t,x,y,z), z fastest, lexicographic n = z + L3(y + L2(x + L1 t)) as in
// the reference (LGT_Code/modules/admin/init.c:768-811).  With L3 even the
// pair (n, n+1), n even, holds one site of each parity, so cb = n/2 maps each
// parity one-to-one onto [0, VOL/2).  Every direction mu and parity p owns a
// contiguous plane of VOL/2 links.
// ---------------------------------------------------------------------------
struct FastDiv {
    // q = (n * m) >> (32 + s) exactly for 0 <= n < 2^31 (m = ceil(2^(32+s)/d))
    uint64_t m;
    uint32_t s, d;
};

__host__ inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    uint32_t s = 0;
    while ((1u << s) < d) s++;
    f.s = s;
    f.d = d;
    unsigned __int128 num = ((unsigned __int128)1) << (32 + s);
    f.m = (uint64_t)((num + d - 1) / d);
    return f;
}

__device__ __forceinline__ uint32_t fdiv(const FastDiv& f, uint32_t n)
{
    return (uint32_t)(((uint64_t)n * f.m) >> (32 + f.s));
}

struct Geo {
    int L[4];          // t, x, y, z extents
    int hz;            // L3 / 2
    uint32_t vh;       // VOL / 2
    uint32_t stride[4];// cb index strides of t, x, y (z handled apart)
    FastDiv dz, dy, dx;// divisions by hz, L2, L1
    FastDiv dt;        // division by L0 (tile order)
    int order;         // tile order flags: 1 = XCD-contiguous, 2 = time fastest, 4 = grouped
    int gs;            // spatial tiles per group of the grouped order (0: not available)
    FastDiv dgs, dgt;  // divisions by gs and gs * L0
};

struct Site { int c[4]; };

__device__ __forceinline__ Site cb_site(const Geo& g, uint32_t i, int p)
{
    Site s;
    uint32_t r = fdiv(g.dz, i);
    int zh = (int)(i - r * (uint32_t)g.hz);
    uint32_t r2 = fdiv(g.dy, r);
    s.c[2] = (int)(r - r2 * (uint32_t)g.L[2]);
    uint32_t t = fdiv(g.dx, r2);
    s.c[1] = (int)(r2 - t * (uint32_t)g.L[1]);
    s.c[0] = (int)t;
    s.c[3] = 2 * zh + ((s.c[0] + s.c[1] + s.c[2] + p) & 1);
    return s;
}

__device__ __forceinline__ uint32_t cb_index(const Geo& g, const Site& s)
{
    return (((uint32_t)s.c[0] * g.L[1] + s.c[1]) * g.L[2] + s.c[2]) * (uint32_t)g.hz + (uint32_t)(s.c[3] >> 1);
}

__device__ __forceinline__ int parity(const Site& s) { return (s.c[0] + s.c[1] + s.c[2] + s.c[3]) & 1; }

// shift by step (|step| <= L[mu]) along mu with periodic wrap
__device__ __forceinline__ Site shift(const Geo& g, Site s, int mu, int step)
{
    int c = s.c[mu] + step;
    c = (c >= g.L[mu]) ? c - g.L[mu] : c;
    c = (c < 0) ? c + g.L[mu] : c;
    s.c[mu] = c;
    return s;
}

// link planes of one field: p[mu][parity] -> VOL/2 links.  The two parity
// planes of a direction are adjacent, p[mu][1] == p[mu][0] + VOL/2, so a link
// of any parity is one 32-bit byte offset from p[mu][0] (the host keeps
// 2 * VOL/2 * sizeof(link) < 2^32).
template <typename T> struct Planes {
    typename Store<T>::V* p[4][2];
};

// the link at byte offset off from the start of direction mu's planes
template <typename T>
__device__ __forceinline__ Q<T> link_at(const Planes<T>& P, int mu, uint32_t off)
{
    const typename Store<T>::V* q = (const typename Store<T>::V*)((const char*)P.p[mu][0] + off);
    typename Store<T>::V v = *q;
    return Q<T>{v.x, v.y, v.z, v.w};
}

template <typename T>
__device__ __forceinline__ Q<T> link(const Planes<T>& P, const Geo& g, int mu, const Site& s)
{
    return link_at(P, mu, ((uint32_t)parity(s) * g.vh + cb_index(g, s)) * (uint32_t)sizeof(typename Store<T>::V));
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): multipliers 0xD2511F53 /
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85, ten rounds.
// ---------------------------------------------------------------------------
struct U32x4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        // one 32x32->64 multiply-add (v_mad_u64_u32) yields both halves
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        U32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
    }
    return c;
}

// 53-bit uniform in [0,1) from two words; 24-bit uniform from one word
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return (double)m * 0x1p-53;
}
__device__ __forceinline__ float u24(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// counter layout of a link stream: (cb site, mu | parity<<2 | kind<<3, sweep, draw)
enum { KIND_HB = 0, KIND_HOT = 1, KIND_METRO = 2, KIND_GAUGE = 3 };

struct LinkRng {
    U32x4 ctr;
    uint32_t k0, k1;
    __device__ __forceinline__ LinkRng(uint64_t seed, uint32_t site, uint32_t tag, uint32_t sweep)
    {
        ctr.x = site; ctr.y = tag; ctr.z = sweep; ctr.w = 0;
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
    }
    __device__ __forceinline__ U32x4 next()
    {
        U32x4 r = philox4x32_10(ctr, k0, k1);
        ctr.w++;
        return r;
    }
};

// FP64 uniform stream of a link: draws 2j, 2j+1 come from Philox call j.
struct RngSrc64 {
    LinkRng rng;
    double buf;
    bool have;
    __device__ __forceinline__ RngSrc64(uint64_t seed, uint32_t site, uint32_t tag, uint32_t sweep)
        : rng(seed, site, tag, sweep), buf(0.0), have(false) {}
    __device__ __forceinline__ double next()
    {
        if (have) { have = false; return buf; }
        U32x4 w = rng.next();
        buf = u53(w.z, w.w);
        have = true;
        return u53(w.x, w.y);
    }
    __device__ __forceinline__ bool dry() const { return false; }
};

// uniforms injected from memory (test harness of the per-link update)
struct BufSrc {
    const double* p;
    int n, pos;
    __device__ __forceinline__ double next()
    {
        double v = (pos < n) ? p[pos] : 0.5;
        pos++;
        return v;
    }
    __device__ __forceinline__ bool dry() const { return pos >= n; }
};

// ---------------------------------------------------------------------------
// double-double accumulation (error-free TwoSum, Knuth)
// ---------------------------------------------------------------------------
struct DD { double hi, lo; };

__device__ __forceinline__ DD dd_add(DD a, double b)
{
    double s = a.hi + b;
    double bb = s - a.hi;
    double err = (a.hi - (s - bb)) + (b - bb);
    double lo = a.lo + err;
    double hi = s + lo;
    lo = lo - (hi - s);
    return DD{hi, lo};
}

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    DD r = dd_add(a, b.hi);
    return dd_add(r, b.lo);
}

}  // namespace lgt
