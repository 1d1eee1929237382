// launch_impl.hpp -- definitions of Launch<T>; included once per precision.
#pragma once

#include "kernels.hpp"
#include "launch.hpp"
#include "vjp_kernels.hpp"

namespace lgt {

template <typename T>
static Planes<T> planes(const RawPlanes& R)
{
    Planes<T> P;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++) P.p[mu][p] = (typename Store<T>::V*)R.p[mu][p];
    return P;
}

template <typename T>
hipError_t Launch<T>::cold(hipStream_t s, const Geo& g, const RawPlanes& P)
{
    hipLaunchKernelGGL(k_cold<T>, dim3(grid_x(g), 8), dim3(BLOCK), 0, s, planes<T>(P), g);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::hot(hipStream_t s, const Geo& g, const RawPlanes& P, uint64_t seed)
{
    hipLaunchKernelGGL(k_hot<T>, dim3(grid_x(g), 8), dim3(BLOCK), 0, s, planes<T>(P), g, seed);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::from_aos(hipStream_t s, const Geo& g, const RawPlanes& P, const double* aos)
{
    hipLaunchKernelGGL(k_from_aos<T>, dim3((2 * g.vh + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, planes<T>(P), g, aos);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::to_aos(hipStream_t s, const Geo& g, const RawPlanes& P, double* aos)
{
    hipLaunchKernelGGL(k_to_aos<T>, dim3((2 * g.vh + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, planes<T>(P), g, aos);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::project(hipStream_t s, const Geo& g, const RawPlanes& P)
{
    hipLaunchKernelGGL(k_project<T>, dim3(grid_x(g), 8), dim3(BLOCK), 0, s, planes<T>(P), g);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::gauge(hipStream_t s, const Geo& g, const RawPlanes& P, uint64_t seed)
{
    hipLaunchKernelGGL(k_gauge<T>, dim3(grid_x(g), 8), dim3(BLOCK), 0, s, planes<T>(P), g, seed);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::gauge_elements(hipStream_t s, const Geo& g, uint64_t seed, double* out)
{
    hipLaunchKernelGGL(k_gauge_elements<T>, dim3((2 * g.vh + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, g, seed, out);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::heatbath(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity, double beta,
                               uint64_t seed, uint32_t sweep, void* acc_partial, bool project)
{
    dim3 grid(grid_x(g)), block(BLOCK);
    Planes<T> Pl = planes<T>(P);
    DD* ap = (DD*)acc_partial;
    auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, grid, block, 0, s, Pl, g, parity, (T)beta, seed, sweep, ap); };
    if (project) {
        // only FP32 fields fold the projection into the heatbath
        if constexpr (sizeof(T) == 4) {
            switch (mu) {
            case 0: go(k_heatbath<T, 0, true>); break;
            case 1: go(k_heatbath<T, 1, true>); break;
            case 2: go(k_heatbath<T, 2, true>); break;
            default: go(k_heatbath<T, 3, true>); break;
            }
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    switch (mu) {
    case 0: go(k_heatbath<T, 0, false>); break;
    case 1: go(k_heatbath<T, 1, false>); break;
    case 2: go(k_heatbath<T, 2, false>); break;
    default: go(k_heatbath<T, 3, false>); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::overrelax(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity)
{
    dim3 grid(grid_x(g)), block(BLOCK);
    Planes<T> Pl = planes<T>(P);
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_overrelax<T, 0>), grid, block, 0, s, Pl, g, parity); break;
    case 1: hipLaunchKernelGGL((k_overrelax<T, 1>), grid, block, 0, s, Pl, g, parity); break;
    case 2: hipLaunchKernelGGL((k_overrelax<T, 2>), grid, block, 0, s, Pl, g, parity); break;
    default: hipLaunchKernelGGL((k_overrelax<T, 3>), grid, block, 0, s, Pl, g, parity); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::metropolis(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity, double beta,
                                 double eps, uint64_t seed, uint32_t sweep, void* acc_partial)
{
    dim3 grid(grid_x(g)), block(BLOCK);
    Planes<T> Pl = planes<T>(P);
    DD* ap = (DD*)acc_partial;
    const T b = (T)beta, e = (T)eps;
    switch (mu) {
    case 0: hipLaunchKernelGGL((k_metropolis<T, 0>), grid, block, 0, s, Pl, g, parity, b, e, seed, sweep, ap); break;
    case 1: hipLaunchKernelGGL((k_metropolis<T, 1>), grid, block, 0, s, Pl, g, parity, b, e, seed, sweep, ap); break;
    case 2: hipLaunchKernelGGL((k_metropolis<T, 2>), grid, block, 0, s, Pl, g, parity, b, e, seed, sweep, ap); break;
    default: hipLaunchKernelGGL((k_metropolis<T, 3>), grid, block, 0, s, Pl, g, parity, b, e, seed, sweep, ap); break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::staple_link(hipStream_t s, const Geo& g, const RawPlanes& P, uint32_t n, int mu, double* out)
{
    hipLaunchKernelGGL(k_staple_link<T>, dim3(1), dim3(64), 0, s, planes<T>(P), g, n, mu, out);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::local_heatbath(hipStream_t s, const Geo& g, const RawPlanes& P, uint32_t n, int mu, double beta,
                                     const double* u, int nu, double* link_out, int* status)
{
    hipLaunchKernelGGL(k_local_heatbath<T>, dim3(1), dim3(64), 0, s, planes<T>(P), g, n, mu, beta, u, nu, link_out,
                       status);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::plaquette(hipStream_t s, const Geo& g, const RawPlanes& P, void* partial)
{
    hipLaunchKernelGGL(k_plaquette<T>, dim3(2 * grid_x(g)), dim3(BLOCK), 0, s, planes<T>(P), g, (DD*)partial);
    return hipGetLastError();
}

// FP32 spatial smearing runs the site-pair kernel (the one case it is faster
// in) when the output's parity planes are adjacent; returns false where
// k_ape serves
template <typename T>
static bool ape_zpair(hipStream_t s, const Geo& g, const Planes<T>& Z, const Planes<T>& O, int mode, double sp,
                      const NormPlanes<T>& mn)
{
    if constexpr (sizeof(T) == 4) {
        bool adjacent = true;
        for (int mu = 1; mu < 4; mu++) adjacent = adjacent && O.p[mu][1] == O.p[mu][0] + g.vh;
        if (mode == 0 && adjacent) {
            hipLaunchKernelGGL((k_ape_zpair<T, 0>), dim3(grid_x(g)), dim3(BLOCK), 0, s, Z, O, g, (T)sp, mn);
            return true;
        }
    }
    return false;
}

template <typename T>
hipError_t Launch<T>::ape(hipStream_t s, const Geo& g, const RawPlanes& in, const RawPlanes& out, int mode, double sp)
{
    dim3 grid(2 * grid_x(g)), block(BLOCK);
    Planes<T> Z = planes<T>(in), O = planes<T>(out);
    NormPlanes<T> none{};
    if (ape_zpair<T>(s, g, Z, O, mode, sp, none)) return hipGetLastError();
    if (mode == 0) hipLaunchKernelGGL((k_ape<T, 0>), grid, block, 0, s, Z, O, g, (T)sp, none);
    else if (mode == 1) hipLaunchKernelGGL((k_ape<T, 1>), grid, block, 0, s, Z, O, g, (T)sp, none);
    else hipLaunchKernelGGL((k_ape<T, 2>), grid, block, 0, s, Z, O, g, (T)sp, none);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::sline(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& S, int R0, int R1)
{
    dim3 grid(2 * grid_x(g), 3), block(BLOCK);
    hipLaunchKernelGGL(k_sline<T>, grid, block, 0, s, planes<T>(P), planes<T>(S), g, R0, R1);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::wloops(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& S,
                             const WGroupHost* groups, int group, const int* slots, void* partial, int ncols,
                             const RawPlanes& Sn, int Rn)
{
    dim3 grid(2 * grid_x(g), 3), block(WL<T>::THREADS);
    hipLaunchKernelGGL(k_wloops<T>, grid, block, 0, s, planes<T>(P), planes<T>(S), g, (const WGroup*)groups, group,
                       slots, (DD*)partial, ncols, planes<T>(Sn), Rn);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::qlines(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& L2,
                             const RawPlanes& L3)
{
    hipLaunchKernelGGL(k_qlines<T>, dim3(grid_x(g), 4), dim3(BLOCK), 0, s, planes<T>(P), planes<T>(L2), planes<T>(L3),
                       g);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::qtop(hipStream_t s, const Geo& g, const RawPlanes& P, int kind, const RawPlanes& L2,
                           const RawPlanes& L3, void* partial, double* density)
{
    dim3 grid(2 * grid_x(g)), block(BLOCK);
    if (kind == 0)
        hipLaunchKernelGGL(k_qtop_lines<T>, grid, block, 0, s, planes<T>(P), planes<T>(L2), planes<T>(L3), g,
                           (DD*)partial, density);
    else hipLaunchKernelGGL(k_qtop_clover<T>, grid, block, 0, s, planes<T>(P), g, (DD*)partial, density);
    return hipGetLastError();
}

template <typename T>
static NormPlanes<T> norms(const RawNorms& R)
{
    NormPlanes<T> N;
    for (int mu = 0; mu < 4; mu++)
        for (int p = 0; p < 2; p++) N.p[mu][p] = (T*)R.p[mu][p];
    return N;
}

template <typename T>
hipError_t Launch<T>::ape_save(hipStream_t s, const Geo& g, const RawPlanes& in, const RawPlanes& out,
                               const RawNorms& mn, int mode, double sp)
{
    dim3 grid(2 * grid_x(g)), block(BLOCK);
    Planes<T> Z = planes<T>(in), O = planes<T>(out);
    if (ape_zpair<T>(s, g, Z, O, mode, sp, norms<T>(mn))) return hipGetLastError();
    if (mode == 0) hipLaunchKernelGGL((k_ape<T, 0>), grid, block, 0, s, Z, O, g, (T)sp, norms<T>(mn));
    else if (mode == 1) hipLaunchKernelGGL((k_ape<T, 1>), grid, block, 0, s, Z, O, g, (T)sp, norms<T>(mn));
    else hipLaunchKernelGGL((k_ape<T, 2>), grid, block, 0, s, Z, O, g, (T)sp, norms<T>(mn));
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::vjp_tangent(hipStream_t s, const Geo& g, const RawPlanes& U, const RawPlanes& G,
                                  const RawNorms& mn, const RawPlanes& H, int mode)
{
    const dim3 block(BLOCK);
    const Planes<T> Up = planes<T>(U), Gp = planes<T>(G), Hp = planes<T>(H);
    const NormPlanes<T> M = norms<T>(mn);
    if (mode == 0) hipLaunchKernelGGL((k_vjp_tangent<T, 0>), dim3(grid_x(g), 6), block, 0, s, Up, Gp, M, Hp, g);
    else if (mode == 1) hipLaunchKernelGGL((k_vjp_tangent<T, 1>), dim3(grid_x(g), 8), block, 0, s, Up, Gp, M, Hp, g);
    else hipLaunchKernelGGL((k_vjp_tangent<T, 2>), dim3(grid_x(g), 8), block, 0, s, Up, Gp, M, Hp, g);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::vjp_step(hipStream_t s, const Geo& g, const RawPlanes& Z, const RawPlanes& Hin,
                               const RawNorms* mprev, const RawPlanes& out, int mode, double sp)
{
    RawNorms none{};
    const dim3 grid(2 * grid_x(g)), block(BLOCK);
    const Planes<T> Zp = planes<T>(Z), Hp = planes<T>(Hin), Op = planes<T>(out);
    const NormPlanes<T> M = norms<T>(mprev ? *mprev : none);
    const int hp = mprev ? 1 : 0;
    if (mode == 0) hipLaunchKernelGGL((k_vjp_step<T, 0>), grid, block, 0, s, Zp, Hp, M, hp, Op, g, (T)sp);
    else if (mode == 1) hipLaunchKernelGGL((k_vjp_step<T, 1>), grid, block, 0, s, Zp, Hp, M, hp, Op, g, (T)sp);
    else hipLaunchKernelGGL((k_vjp_step<T, 2>), grid, block, 0, s, Zp, Hp, M, hp, Op, g, (T)sp);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::lines(hipStream_t s, const Geo& g, const RawPlanes& U, void* base, int R, int T_)
{
    Lines<T> Ls{(typename Store<T>::V*)base, R, T_};
    hipLaunchKernelGGL(k_lines<T>, dim3(2 * grid_x(g), 4), dim3(BLOCK), 0, s, planes<T>(U), Ls, g);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::wloop_grad(hipStream_t s, const Geo& g, void* lines, int R, int T_, const RawPlanes& G)
{
    Lines<T> Ls{(typename Store<T>::V*)lines, R, T_};
    hipLaunchKernelGGL(k_wloop_grad<T>, dim3(2 * grid_x(g), 4), dim3(BLOCK), 0, s, Ls, g, R, T_, planes<T>(G));
    return hipGetLastError();
}

}  // namespace lgt
