// vjp_kernels.hpp -- adjoint (vector-Jacobian product) of N-step spatial APE
// smearing and the gradient of the Wilson-loop objective
//   L = sum_x sum_{d=1..3} 2 Re[W_{R x T}(x; t, d)]_0 .
// No reference counterpart.  Gradients are Euclidean in the 4 real link
// components, <a,b> = a0 b0 + a.b = Re tr(a b^+)/2; right (left)
// multiplication by q has adjoint right (left) multiplication by q^+.
//
// One forward step (Jacobi, spatial links a = 1..3):
//   M_a(y) = Z_a(y) + sp S_a(y),   U'_a(y) = M_a(y) / |M_a(y)|
// Backward of that step for a cotangent G' of U':
//   H_a(y)  = (G'_a(y) - <G'_a(y), U'_a(y)> U'_a(y)) / |M_a(y)|      (tangent)
//   G_a(y)  = H_a(y) + sp * sum_{b != a} (t1 + ... + t6)            (gather)
// with, for every spatial b != a, the six places where Z_a(y) enters a staple:
//   t1 = H_b(y) Z_a(y+b) Z_b(y+a)^+          t4 = Z_b(y) Z_a(y+b) H_b(y+a)^+
//   t2 = Z_b(y-b)^+ H_a(y-b) Z_b(y-b+a)      t5 = Z_b(y) H_a(y+b) Z_b(y+a)^+
//   t3 = H_b(y-b)^+ Z_a(y-b) Z_b(y-b+a)      t6 = Z_b(y-b)^+ Z_a(y-b) H_b(y-b+a)
// Temporal links are not used by spatial staples: their cotangent passes
// through every step unchanged.
#pragma once

#include "kernels.hpp"

namespace lgt {


// H = (G - <G,U> U) / |M| on the spatial links
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_vjp_tangent(Planes<T> U, Planes<T> G, NormPlanes<T> Mn, Planes<T> H, Geo g)
{
#pragma clang fp contract(fast)
    uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.vh) return;
    int p = blockIdx.y & 1;
    int mu = 1 + (blockIdx.y >> 1);
    Q<T> u = qload<T>(U.p[mu][p], i), gr = qload<T>(G.p[mu][p], i);
    T c = qdot(gr, u), m = Mn.p[mu][p][i];
    Q<T> h{(gr.c0 - c * u.c0) / m, (gr.c1 - c * u.c1) / m, (gr.c2 - c * u.c2) / m, (gr.c3 - c * u.c3) / m};
    qstore<T>(H.p[mu][p], i, h);
}

// One backward step.  Z = input field of the step, Hin = tangent cotangent of
// its output (H above).  Writes G = dL/dZ for the spatial links, or, when
// Mprev is given (|M| of the step before), directly the tangent cotangent of
// that earlier step: (G - <G,Z> Z)/Mprev  (Z is that step's output).
// One thread handles the three spatial links of its site and loads each of
// the 48 (Z, H) links they need once, in batches: the six at y; for every
// pair {a,b} the four at y+a and y+b, which close the forward staples of both
// G_a and G_b; for every b the ten around y-b, which close the backward
// staples of the two G_a with a != b.  A batch's loads are issued together
// before its products (its latency is exposed once) and only one batch is
// live at a time (FP64: 166 VGPRs, three waves per SIMD).
// grid: x over (tile, parity) pairs
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_vjp_step(Planes<T> Z, Planes<T> Hin, NormPlanes<T> Mprev, int has_prev,
                                                    Planes<T> Out, Geo g, T sp)
{
#pragma clang fp contract(fast)
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    const Site y = cb_site(g, i, p);
    Q<T> Zy[4], Hy[4], acc[4];
#pragma unroll
    for (int a = 1; a < 4; a++) {
        Zy[a] = link(Z, g, a, y);
        Hy[a] = link(Hin, g, a, y);
        acc[a] = qzero<T>();
    }
    // batches 1-3: for each pair {a,b} the forward staples through y+a, y+b
#pragma unroll
    for (int a = 1; a < 4; a++)
#pragma unroll
        for (int b = a + 1; b < 4; b++) {
#pragma unroll
            for (int c = 1; c < 4; c++) pin(acc[c]);
            asm volatile("" ::: "memory");
            const Site ya = shift(g, y, a, 1), yb = shift(g, y, b, 1);
            const Q<T> Zab = link(Z, g, a, yb), Hab = link(Hin, g, a, yb);   // Z_a(y+b), H_a(y+b)
            const Q<T> Zba = link(Z, g, b, ya), Hba = link(Hin, g, b, ya);   // Z_b(y+a), H_b(y+a)
            asm volatile("" ::: "memory");
            // t1 + t5 + t4 of G_a = (H_b(y) Z_a(y+b) + Z_b(y) H_a(y+b)) Z_b(y+a)^+ + Z_b(y) Z_a(y+b) H_b(y+a)^+
            Q<T> up = qadd(qmul(Hy[b], Zab), qmul(Zy[b], Hab));
            acc[a] = qadd(acc[a], qadd(qmul_dag(up, Zba), qmul_dag(qmul(Zy[b], Zab), Hba)));
            // the same for G_b, a and b exchanged
            up = qadd(qmul(Hy[a], Zba), qmul(Zy[a], Hba));
            acc[b] = qadd(acc[b], qadd(qmul_dag(up, Zab), qmul_dag(qmul(Zy[a], Zba), Hab)));
        }
    // batches 4-6: for each b the backward staples through y-b
#pragma unroll
    for (int b = 1; b < 4; b++) {
        // the previous batch's products are complete before this batch's loads
        // are issued (its links are dead by then: one batch live at a time)
#pragma unroll
        for (int a = 1; a < 4; a++) pin(acc[a]);
        asm volatile("" ::: "memory");
        const Site ym = shift(g, y, b, -1);
        const Q<T> Zb = link(Z, g, b, ym), Hb = link(Hin, g, b, ym);   // Z_b(y-b), H_b(y-b)
        Q<T> C[4][2], D[4][2];   // C[a] = (Z_a(y-b), H_a(y-b)), D[a] = (Z_b(y-b+a), H_b(y-b+a))
#pragma unroll
        for (int a = 1; a < 4; a++) {
            if (a == b) continue;
            const Site yma = shift(g, ym, a, 1);
            C[a][0] = link(Z, g, a, ym);
            C[a][1] = link(Hin, g, a, ym);
            D[a][0] = link(Z, g, b, yma);
            D[a][1] = link(Hin, g, b, yma);
        }
        asm volatile("" ::: "memory");
        const Q<T> Zbd = qdag(Zb), Hbd = qdag(Hb);
#pragma unroll
        for (int a = 1; a < 4; a++) {
            if (a == b) continue;
            // t2 + t3 + t6 = (Z_b(y-b)^+ H_a(y-b) + H_b(y-b)^+ Z_a(y-b)) Z_b(y-b+a) + Z_b(y-b)^+ Z_a(y-b) H_b(y-b+a)
            const Q<T> lo = qadd(qmul(Zbd, C[a][1]), qmul(Hbd, C[a][0]));
            const Q<T> zl = qmul(Zbd, C[a][0]);
            acc[a] = qadd(acc[a], qadd(qmul(lo, D[a][0]), qmul(zl, D[a][1])));
        }
    }
#pragma unroll
    for (int a = 1; a < 4; a++) {
        Q<T> G = qadd(Hy[a], qscale(acc[a], sp));
        if (has_prev) {
            const Q<T>& z = Zy[a];
            T c = qdot(G, z), m = Mprev.p[a][p][i];
            G = Q<T>{(G.c0 - c * z.c0) / m, (G.c1 - c * z.c1) / m, (G.c2 - c * z.c2) / m, (G.c3 - c * z.c3) / m};
        }
        qstore<T>(Out.p[a][p], i, G);
    }
}

// ---------------------------------------------------------------------------
// gradient of L for every link, gathered per link over the loops that hold it
// (the loop path: T steps +t, R steps +d, T steps -t, R steps -d from x).
// For a link A at some position of a loop P = X A Y, d(2 Re[P]_0)/dA is
// 2 (Y X)^+, and 2 Y X where the loop runs the link backwards; Y X is the
// open path from the end of the link around the loop to its start.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void seg(const Planes<T>& U, const Geo& g, Q<T>& acc, Site& s, int dir, int n, bool back)
{
    for (int k = 0; k < n; k++) {
        if (!back) {
            acc = qmul(acc, link(U, g, dir, s));
            s = shift(g, s, dir, 1);
        } else {
            s = shift(g, s, dir, -1);
            acc = qmul_dag(acc, link(U, g, dir, s));
        }
    }
}

// grid: x over (tile, parity) pairs, y = mu
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_wloop_grad(Planes<T> U, Geo g, int R, int Tt, Planes<T> G)
{
#pragma clang fp contract(fast)
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    int mu = blockIdx.y;
    Site y = cb_site(g, i, p);
    Q<T> grad = qzero<T>();
    if (mu == 0) {
        for (int d = 1; d < 4; d++) {
            for (int k = 0; k < Tt; k++) {
                // left edge, position k: forward link
                Q<T> acc = qunit<T>();
                Site s = shift(g, y, 0, 1);
                seg(U, g, acc, s, 0, Tt - 1 - k, false);
                seg(U, g, acc, s, d, R, false);
                seg(U, g, acc, s, 0, Tt, true);
                seg(U, g, acc, s, d, R, true);
                seg(U, g, acc, s, 0, k, false);
                grad = qadd(grad, qdag(acc));
                // right edge, position k: the loop runs the link backwards
                acc = qunit<T>();
                s = y;
                seg(U, g, acc, s, 0, Tt - 1 - k, true);
                seg(U, g, acc, s, d, R, true);
                seg(U, g, acc, s, 0, Tt, false);
                seg(U, g, acc, s, d, R, false);
                seg(U, g, acc, s, 0, k, true);
                grad = qadd(grad, acc);
            }
        }
    } else {
        const int d = mu;
        for (int k = 0; k < R; k++) {
            // top edge, position k: forward link
            Q<T> acc = qunit<T>();
            Site s = shift(g, y, d, 1);
            seg(U, g, acc, s, d, R - 1 - k, false);
            seg(U, g, acc, s, 0, Tt, true);
            seg(U, g, acc, s, d, R, true);
            seg(U, g, acc, s, 0, Tt, false);
            seg(U, g, acc, s, d, k, false);
            grad = qadd(grad, qdag(acc));
            // bottom edge, position k: backwards
            acc = qunit<T>();
            s = y;
            seg(U, g, acc, s, d, R - 1 - k, true);
            seg(U, g, acc, s, 0, Tt, false);
            seg(U, g, acc, s, d, R, false);
            seg(U, g, acc, s, 0, Tt, true);
            seg(U, g, acc, s, d, k, true);
            grad = qadd(grad, acc);
        }
    }
    qstore<T>(G.p[mu][p], i, qscale(grad, T(2)));
}

}  // namespace lgt
