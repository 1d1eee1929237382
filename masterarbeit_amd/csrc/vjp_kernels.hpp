// vjp_kernels.hpp -- adjoint (vector-Jacobian product) of N-step APE
// smearing (spatial, all directions, temporal) and the gradient of the
// Wilson-loop objective
//   L = sum_x sum_{d=1..3} 2 Re[W_{R x T}(x; t, d)]_0 .
// No reference counterpart.  Gradients are Euclidean in the 4 real link
// components, <a,b> = a0 b0 + a.b = Re tr(a b^+)/2; right (left)
// multiplication by q has adjoint right (left) multiplication by q^+.
//
// One forward step (Jacobi) smears the links a of the mode's set D:
//   M_a(y) = Z_a(y) + sp S_a(y),   U'_a(y) = M_a(y) / |M_a(y)|,
// the staple S_a summing over the mode's staple directions K(a) (ApeDirs).
// Backward of that step for a cotangent G' of U':
//   H_a(y)  = (G'_a(y) - <G'_a(y), U'_a(y)> U'_a(y)) / |M_a(y)|,  a in D
//   H_a(y)  = G'_a(y),                                           a not in D
//   G_a(y)  = H_a(y) + sp * sum_{b != a} (terms below)            (gather)
// where Z_a(y) enters a staple: as a side link of a staple of U'_b (b in D,
// a in K(b)): t1, t3, t4, t6; as the middle link of a staple of U'_a at a
// neighbour (a in D, b in K(a)): t2, t5:
//   t1 = H_b(y) Z_a(y+b) Z_b(y+a)^+          t4 = Z_b(y) Z_a(y+b) H_b(y+a)^+
//   t2 = Z_b(y-b)^+ H_a(y-b) Z_b(y-b+a)      t5 = Z_b(y) H_a(y+b) Z_b(y+a)^+
//   t3 = H_b(y-b)^+ Z_a(y-b) Z_b(y-b+a)      t6 = Z_b(y-b)^+ Z_a(y-b) H_b(y-b+a)
// Links in no staple of the mode (the temporal ones of spatial APE) keep
// their cotangent through every step.
#pragma once

#include "kernels.hpp"

namespace lgt {


// acc + (u w^+ + v x^+) and acc + (u w + v x), both products folded into the
// sum's FMA chains
template <typename T>
__device__ __forceinline__ Q<T> add_terms(const Q<T>& acc, const Q<T>& u, const Q<T>& w, const Q<T>& v, const Q<T>& x)
{
    return qmul_dag_add(qmul_dag_add(acc, u, w), v, x);
}

template <typename T>
__device__ __forceinline__ Q<T> add_terms_fwd(const Q<T>& acc, const Q<T>& u, const Q<T>& w, const Q<T>& v, const Q<T>& x)
{
    return qmul_add(qmul_add(acc, u, w), v, x);
}

// Directions of the APE modes (LGT_Code/modules/obs/smearing.c:74-168): the
// smeared links j in [LO, HI]; the staple directions k of j; the directions
// whose cotangent the adjoint carries, [CLO, 3] (all that a staple touches).
template <int MODE> struct ApeDirs {
    static constexpr int LO = (MODE == 0) ? 1 : 0, HI = (MODE == 2) ? 0 : 3, CLO = (MODE == 0) ? 1 : 0;
    static constexpr bool smeared(int j) { return j >= LO && j <= HI; }
    static constexpr bool staple(int j, int k) { return smeared(j) && k != j && (MODE == 1 || k != 0); }
};

// (x - c u) / m for the four components of x and u, with r = 1/m.  FP32
// corrects each product x r by one FMA step (q + (x - q m) r, within an ulp of
// the quotient): the plain products let the 400-step FP32 gradient drift half
// as far again from the FP64 one.
template <typename T>
__device__ __forceinline__ Q<T> tangent_div(const Q<T>& x, T c, const Q<T>& u, T m, T r)
{
#pragma clang fp contract(fast)
    const T v[4] = {x.c0 - c * u.c0, x.c1 - c * u.c1, x.c2 - c * u.c2, x.c3 - c * u.c3};
    T o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const T q = v[k] * r;
        o[k] = (sizeof(T) == 4) ? fma(fma(-q, m, v[k]), r, q) : q;
    }
    return Q<T>{o[0], o[1], o[2], o[3]};
}

// H = (G - <G,U> U) / |M| on the smeared links, H = G on the other carried ones
// grid: x over cb sites, y = (direction - CLO) * 2 + parity
template <typename T, int MODE>
__global__ void __launch_bounds__(BLOCK) k_vjp_tangent(Planes<T> U, Planes<T> G, NormPlanes<T> Mn, Planes<T> H, Geo g)
{
#pragma clang fp contract(fast)
    typedef ApeDirs<MODE> D;
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.vh) return;
    int p = blockIdx.y & 1;
    int mu = D::CLO + (int)(blockIdx.y >> 1);
    Q<T> gr = qload<T>(G.p[mu][p], i);
    if (D::smeared(mu)) {
        Q<T> u = qload<T>(U.p[mu][p], i);
        const T c = qdot(gr, u), m = Mn.p[mu][p][i];
        gr = tangent_div(gr, c, u, m, T(1) / m);
    }
    qstore<T>(H.p[mu][p], i, gr);
}

// One backward step.  Z = input field of the step, Hin = tangent cotangent of
// its output (H above).  Writes G = dL/dZ for the carried links, or, when
// Mprev is given (|M| of the step before), directly the tangent cotangent of
// that earlier step: (G - <G,Z> Z)/Mprev on the smeared links (Z is that
// step's output), G on the others.
// One thread handles the carried links of its site and loads each (Z, H) link
// they need once (spatial: 48), in batches: the links at y; for every pair
// {a,b} the four at y+a and y+b, which close the forward staples of both G_a
// and G_b; for every b those around y-b, which close the backward staples of
// the G_a, a != b.  A batch's loads are issued together before its products
// (its latency is exposed once) and only one batch is live at a time
// (spatial FP64: 169 VGPRs, two waves per SIMD).  A term the mode does not
// have is neither loaded nor computed; the spatial mode's arithmetic is the
// same operation for operation as with every term present.  Products and
// running sums contract to FMAs in both precisions (qmul_c, qmul_add): the
// adjoint has no reference rounding to reproduce.
// grid: x over (tile, parity) pairs
template <typename T, int MODE>
__global__ void __launch_bounds__(BLOCK) k_vjp_step(Planes<T> Z, Planes<T> Hin, NormPlanes<T> Mprev, int has_prev,
                                                    Planes<T> Out, Geo g, T sp)
{
#pragma clang fp contract(fast)
    typedef ApeDirs<MODE> D;
    constexpr int C0 = D::CLO;
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    const Site y = cb_site(g, i, p);
    Q<T> Zy[4], Hy[4], acc[4];
#pragma unroll
    for (int a = C0; a < 4; a++) {
        Zy[a] = link(Z, g, a, y);
        Hy[a] = link(Hin, g, a, y);
        acc[a] = qzero<T>();
    }
    // forward staples: for each pair {a,b} the links at y+a and y+b
#pragma unroll
    for (int a = C0; a < 4; a++)
#pragma unroll
        for (int b = a + 1; b < 4; b++) {
            const bool side_ab = D::staple(b, a), mid_ab = D::staple(a, b);   // terms of G_a
            const bool side_ba = D::staple(a, b), mid_ba = D::staple(b, a);   // terms of G_b
            if (!(side_ab || mid_ab || side_ba || mid_ba)) continue;
#pragma unroll
            for (int c = C0; c < 4; c++) pin(acc[c]);
            asm volatile("" ::: "memory");
            const Site ya = shift(g, y, a, 1), yb = shift(g, y, b, 1);
            const Q<T> Zab = link(Z, g, a, yb), Hab = link(Hin, g, a, yb);   // Z_a(y+b), H_a(y+b)
            const Q<T> Zba = link(Z, g, b, ya), Hba = link(Hin, g, b, ya);   // Z_b(y+a), H_b(y+a)
            asm volatile("" ::: "memory");
            if (sizeof(T) == 8 && side_ab && mid_ab && side_ba && mid_ba) {
                // t1 + t5 + t4 of G_a = H_b(y) W + Z_b(y) V and of G_b = H_a(y) W^+ + Z_a(y) V^+ with
                // W = Z_a(y+b) Z_b(y+a)^+, V = H_a(y+b) Z_b(y+a)^+ + Z_a(y+b) H_b(y+a)^+
                // (G_b's middle-link terms sum to V^+): 7 products instead of 10.  FP64
                // only: in FP32 this grouping raised the gradient's rounding error by half
                // (48^4 APE-400, against FP64) for a 2% faster step
                const Q<T> W = qmul_dag_c(Zab, Zba);
                const Q<T> V = qmul_dag_add(qmul_dag_c(Hab, Zba), Zab, Hba);
                acc[a] = qmul_add(qmul_add(acc[a], Hy[b], W), Zy[b], V);
                acc[b] = qmul_add(qmul_add(acc[b], Hy[a], qdag(W)), Zy[a], qdag(V));
                continue;
            }
            // t1 + t5 + t4 of G_a = (H_b(y) Z_a(y+b) + Z_b(y) H_a(y+b)) Z_b(y+a)^+ + Z_b(y) Z_a(y+b) H_b(y+a)^+
            if (side_ab || mid_ab) {
                const Q<T> up = (side_ab && mid_ab) ? qmul_add(qmul_c(Hy[b], Zab), Zy[b], Hab)
                                                   : (side_ab ? qmul_c(Hy[b], Zab) : qmul_c(Zy[b], Hab));
                acc[a] = side_ab ? add_terms(acc[a], up, Zba, qmul_c(Zy[b], Zab), Hba) : qmul_dag_add(acc[a], up, Zba);
            }
            // the same for G_b, a and b exchanged
            if (side_ba || mid_ba) {
                const Q<T> up = (side_ba && mid_ba) ? qmul_add(qmul_c(Hy[a], Zba), Zy[a], Hba)
                                                   : (side_ba ? qmul_c(Hy[a], Zba) : qmul_c(Zy[a], Hba));
                acc[b] = side_ba ? add_terms(acc[b], up, Zab, qmul_c(Zy[a], Zba), Hab) : qmul_dag_add(acc[b], up, Zab);
            }
        }
    // backward staples: for each b the links around y-b
#pragma unroll
    for (int b = C0; b < 4; b++) {
        bool any = false;
#pragma unroll
        for (int a = C0; a < 4; a++) any = any || (a != b && (D::staple(b, a) || D::staple(a, b)));
        if (!any) continue;
        // the previous batch's products are complete before this batch's loads
        // are issued (its links are dead by then: one batch live at a time)
#pragma unroll
        for (int a = C0; a < 4; a++) pin(acc[a]);
        asm volatile("" ::: "memory");
        const Site ym = shift(g, y, b, -1);
        const Q<T> Zb = link(Z, g, b, ym), Hb = link(Hin, g, b, ym);   // Z_b(y-b), H_b(y-b)
        Q<T> C[4][2], E[4][2];   // C[a] = (Z_a(y-b), H_a(y-b)), E[a] = (Z_b(y-b+a), H_b(y-b+a))
#pragma unroll
        for (int a = C0; a < 4; a++) {
            if (a == b || !(D::staple(b, a) || D::staple(a, b))) continue;
            const Site yma = shift(g, ym, a, 1);
            C[a][0] = link(Z, g, a, ym);
            C[a][1] = link(Hin, g, a, ym);
            E[a][0] = link(Z, g, b, yma);
            E[a][1] = link(Hin, g, b, yma);
        }
        asm volatile("" ::: "memory");
        const Q<T> Zbd = qdag(Zb), Hbd = qdag(Hb);
#pragma unroll
        for (int a = C0; a < 4; a++) {
            const bool side = D::staple(b, a), mid = D::staple(a, b);
            if (a == b || !(side || mid)) continue;
            // t2 + t3 + t6 = (Z_b(y-b)^+ H_a(y-b) + H_b(y-b)^+ Z_a(y-b)) Z_b(y-b+a) + Z_b(y-b)^+ Z_a(y-b) H_b(y-b+a)
            const Q<T> lo = (side && mid) ? qmul_add(qmul_c(Zbd, C[a][1]), Hbd, C[a][0])
                                          : (mid ? qmul_c(Zbd, C[a][1]) : qmul_c(Hbd, C[a][0]));
            acc[a] = side ? add_terms_fwd(acc[a], lo, E[a][0], qmul_c(Zbd, C[a][0]), E[a][1]) : qmul_add(acc[a], lo, E[a][0]);
        }
    }
#pragma unroll
    for (int a = C0; a < 4; a++) {
        Q<T> G = qadd(Hy[a], qscale(acc[a], sp));
        if (has_prev && D::smeared(a)) {
            const Q<T>& z = Zy[a];
            const T c = qdot(G, z), m = Mprev.p[a][p][i];
            G = tangent_div(G, c, z, m, T(1) / m);
        }
        qstore<T>(Out.p[a][p], i, G);
    }
}

// ---------------------------------------------------------------------------
// gradient of L for every link, gathered per link over the loops that hold it
// (the loop path: T steps +t, R steps +d, T steps -t, R steps -d from x).
// For a link A at some position of a loop P = X A Y, d(2 Re[P]_0)/dA is
// 2 (Y X)^+, and 2 Y X where the loop runs the link backwards; Y X is the
// open path from the end of the link around the loop to its start.  That
// path is five straight lines, read from line fields built beforehand:
//   line(d, n)(z) = U_d(z) U_d(z+d) ... U_d(z+(n-1)d),
// d = 0 for n = 1..T, d = 1..3 for n = 1..R, one field of VOL links each, at
// index (d == 0 ? n - 1 : T + (d-1) R + n - 1) laid out like a link direction
// (parity planes of VOL/2).
// ---------------------------------------------------------------------------
template <typename T> struct Lines {
    typename Store<T>::V* base;
    int R, Tl;   // spatial and temporal line lengths
};

// the n-link line along d from z (n = 0: the identity)
template <typename T>
__device__ __forceinline__ Q<T> line(const Lines<T>& Ls, const Geo& g, int d, int n, const Site& z)
{
    if (n == 0) return qunit<T>();
    const size_t idx = (d == 0) ? (size_t)(n - 1) : (size_t)(Ls.Tl + (d - 1) * Ls.R + n - 1);
    const typename Store<T>::V* f = Ls.base + idx * 2 * (size_t)g.vh;
    const typename Store<T>::V v = f[(uint32_t)parity(z) * g.vh + cb_index(g, z)];
    return Q<T>{v.x, v.y, v.z, v.w};
}

// shift by any number of steps along mu (periodic)
__device__ __forceinline__ Site shiftn(const Geo& g, Site s, int mu, int n)
{
    n %= g.L[mu];
    if (n < 0) n += g.L[mu];
    return shift(g, s, mu, n);
}

// all lines of one direction from one site: grid x over (tile, parity)
// pairs, y = d; lengths 1..T (d = 0) or 1..R
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_lines(Planes<T> U, Lines<T> Ls, Geo g)
{
    int p;
    const uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    const int d = (int)blockIdx.y;
    const int nmax = (d == 0) ? Ls.Tl : Ls.R;
    const Site z = cb_site(g, i, p);
    Site w = z;
    Q<T> acc = link(U, g, d, w);
    for (int n = 1; n <= nmax; n++) {
        if (n > 1) acc = qmul(acc, link(U, g, d, w));
        const size_t idx = (d == 0) ? (size_t)(n - 1) : (size_t)(Ls.Tl + (d - 1) * Ls.R + n - 1);
        typename Store<T>::V v;
        v.x = acc.c0; v.y = acc.c1; v.z = acc.c2; v.w = acc.c3;
        Ls.base[idx * 2 * (size_t)g.vh + (size_t)p * g.vh + i] = v;
        w = shift(g, w, d, 1);
    }
}

// grid: x over (tile, parity) pairs, y = mu
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_wloop_grad(Lines<T> Ls, Geo g, int R, int Tt, Planes<T> G)
{
#pragma clang fp contract(fast)
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    const int mu = blockIdx.y;
    const Site y = cb_site(g, i, p);
    Q<T> grad = qzero<T>();
    if (mu == 0) {
        const Site yt = shift(g, y, 0, 1);
        for (int d = 1; d < 4; d++) {
            for (int k = 0; k < Tt; k++) {
                // left edge, position k (loop from x = y - k t): the link runs forward
                const Site x = shiftn(g, y, 0, -k);
                Q<T> t = line(Ls, g, 0, Tt - 1 - k, yt);
                t = qmul(t, line(Ls, g, d, R, shiftn(g, x, 0, Tt)));
                t = qmul_dag(t, line(Ls, g, 0, Tt, shiftn(g, x, d, R)));
                t = qmul_dag(t, line(Ls, g, d, R, x));
                t = qmul(t, line(Ls, g, 0, k, x));
                grad = qadd(grad, qdag(t));
                // right edge, position k: the loop runs the link backwards
                const Site z1 = shiftn(g, y, 0, -(Tt - 1 - k));
                const Site z2 = shiftn(g, z1, d, -R);
                t = qdag(line(Ls, g, 0, Tt - 1 - k, z1));
                t = qmul_dag(t, line(Ls, g, d, R, z2));
                t = qmul(t, line(Ls, g, 0, Tt, z2));
                t = qmul(t, line(Ls, g, d, R, shiftn(g, z2, 0, Tt)));
                t = qmul_dag(t, line(Ls, g, 0, k, yt));
                grad = qadd(grad, t);
            }
        }
    } else {
        const int d = mu;
        const Site yd = shift(g, y, d, 1);
        for (int k = 0; k < R; k++) {
            // top edge, position k: forward
            const Site b = shiftn(g, shiftn(g, y, d, -k), 0, -Tt);
            Q<T> t = line(Ls, g, d, R - 1 - k, yd);
            t = qmul_dag(t, line(Ls, g, 0, Tt, shiftn(g, b, d, R)));
            t = qmul_dag(t, line(Ls, g, d, R, b));
            t = qmul(t, line(Ls, g, 0, Tt, b));
            t = qmul(t, line(Ls, g, d, k, shiftn(g, y, d, -k)));
            grad = qadd(grad, qdag(t));
            // bottom edge, position k: backwards
            const Site c = shiftn(g, y, d, -(R - 1 - k));
            t = qdag(line(Ls, g, d, R - 1 - k, c));
            t = qmul(t, line(Ls, g, 0, Tt, c));
            t = qmul(t, line(Ls, g, d, R, shiftn(g, c, 0, Tt)));
            t = qmul_dag(t, line(Ls, g, 0, Tt, shiftn(g, y, d, k + 1)));
            t = qmul_dag(t, line(Ls, g, d, k, yd));
            grad = qadd(grad, t);
        }
    }
    qstore<T>(G.p[mu][p], i, qscale(grad, T(2)));
}

}  // namespace lgt
