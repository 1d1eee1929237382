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
// One thread handles the three spatial links of its site, so the gathers the
// directions share are loaded once (51 instead of 81).
// grid: x over (tile, parity) pairs
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_vjp_step(Planes<T> Z, Planes<T> Hin, NormPlanes<T> Mprev, int has_prev,
                                                    Planes<T> Out, Geo g, T sp)
{
#pragma clang fp contract(fast)
    int p;
    uint32_t i = tile_base_pair(g, p) + threadIdx.x;
    if (i >= g.vh) return;
    Site y = cb_site(g, i, p);
    Site yp[4], ym[4];
#pragma unroll
    for (int b = 1; b < 4; b++) {
        yp[b] = shift(g, y, b, 1);
        ym[b] = shift(g, y, b, -1);
    }
    Q<T> res[4];
#pragma unroll
    for (int a = 1; a < 4; a++) {
        Q<T> acc = qzero<T>();
        // the 24 gathers of both b != a are issued before any product: the
        // latency of one batch is exposed once instead of per term
        const int b0 = (a == 1) ? 2 : 1, b1 = (a == 3) ? 2 : 3;
        Q<T> L[2][12];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int b = j ? b1 : b0;
            Site ymba = shift(g, ym[b], a, 1);
            L[j][0] = link(Z, g, b, y);       L[j][1] = link(Hin, g, b, y);
            L[j][2] = link(Z, g, a, yp[b]);   L[j][3] = link(Hin, g, a, yp[b]);
            L[j][4] = link(Z, g, b, yp[a]);   L[j][5] = link(Hin, g, b, yp[a]);
            L[j][6] = link(Z, g, b, ym[b]);   L[j][7] = link(Hin, g, b, ym[b]);
            L[j][8] = link(Z, g, a, ym[b]);   L[j][9] = link(Hin, g, a, ym[b]);
            L[j][10] = link(Z, g, b, ymba);   L[j][11] = link(Hin, g, b, ymba);
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < 2; j++) {
            // forward staple: t1 + t5 = (H_b(y) Z_a(y+b) + Z_b(y) H_a(y+b)) Z_b(y+a)^+,
            //                 t4 = Z_b(y) Z_a(y+b) H_b(y+a)^+   (5 products, not 6)
            Q<T> up = qadd(qmul(L[j][1], L[j][2]), qmul(L[j][0], L[j][3]));
            Q<T> zz = qmul(L[j][0], L[j][2]);
            acc = qadd(acc, qmul_dag(up, L[j][4]));
            acc = qadd(acc, qmul_dag(zz, L[j][5]));
            // backward staple: t2 + t3 = (Z_b(y-b)^+ H_a(y-b) + H_b(y-b)^+ Z_a(y-b)) Z_b(y-b+a),
            //                  t6 = Z_b(y-b)^+ Z_a(y-b) H_b(y-b+a)
            Q<T> lo = qadd(qmul(qdag(L[j][6]), L[j][9]), qmul(qdag(L[j][7]), L[j][8]));
            Q<T> zl = qmul(qdag(L[j][6]), L[j][8]);
            acc = qadd(acc, qmul(lo, L[j][10]));
            acc = qadd(acc, qmul(zl, L[j][11]));
        }
        Q<T> G = qadd(link(Hin, g, a, y), qscale(acc, sp));
        if (has_prev) {
            Q<T> z = link(Z, g, a, y);
            T c = qdot(G, z), m = Mprev.p[a][p][i];
            G = Q<T>{(G.c0 - c * z.c0) / m, (G.c1 - c * z.c1) / m, (G.c2 - c * z.c2) / m, (G.c3 - c * z.c3) / m};
        }
        res[a] = G;
    }
#pragma unroll
    for (int a = 1; a < 4; a++) qstore<T>(Out.p[a][p], i, res[a]);
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
