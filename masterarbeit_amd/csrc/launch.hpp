// launch.hpp -- host-side launchers of the engine kernels, one explicit
// instantiation per storage precision (lgt_f64.hip, lgt_f32.hip).  All
// launches are asynchronous on the caller's stream; device pointers only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "su2_device.hpp"

namespace lgt {

// raw plane pointers of one field: p[mu][parity]
struct RawPlanes { void* p[4][2]; };
// raw planes of one scalar per link (|M| of a smearing step)
struct RawNorms { void* p[4][2]; };

struct WGroupHost { int R, Tmax, slot_off, pad; };

// longest temporal extent of one Wilson-loop launch (LDS slots per wave)
constexpr int WMAXT_HOST = 128;

template <typename T> struct Launch {
    static hipError_t cold(hipStream_t s, const Geo& g, const RawPlanes& P);
    static hipError_t hot(hipStream_t s, const Geo& g, const RawPlanes& P, uint64_t seed);
    static hipError_t from_aos(hipStream_t s, const Geo& g, const RawPlanes& P, const double* aos);
    static hipError_t to_aos(hipStream_t s, const Geo& g, const RawPlanes& P, double* aos);
    static hipError_t project(hipStream_t s, const Geo& g, const RawPlanes& P);
    // random gauge transformation in place; gauge_elements: g(x) per lexicographic site
    static hipError_t gauge(hipStream_t s, const Geo& g, const RawPlanes& P, uint64_t seed);
    static hipError_t gauge_elements(hipStream_t s, const Geo& g, uint64_t seed, double* out);
    // acc_partial: nullable, receives ceil(vh/256) block sums of 1/count;
    // project (FP32 only): store the new links projected onto SU(2)
    static hipError_t heatbath(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity, double beta,
                               uint64_t seed, uint32_t sweep, void* acc_partial, bool project = false);
    static hipError_t overrelax(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity);
    // acc_partial: nullable, receives ceil(vh/256) block sums of accepted links
    static hipError_t metropolis(hipStream_t s, const Geo& g, const RawPlanes& P, int mu, int parity, double beta,
                                 double eps, uint64_t seed, uint32_t sweep, void* acc_partial);
    // partial: 2*ceil(vh/256) double-doubles
    static hipError_t plaquette(hipStream_t s, const Geo& g, const RawPlanes& P, void* partial);
    static hipError_t ape(hipStream_t s, const Geo& g, const RawPlanes& in, const RawPlanes& out, int mode, double sp);
    // groups/slots: device tables; partial: nloop rows of 6*ceil(vh/256) columns
    static hipError_t sline(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& S, int R0, int R1);
    // Rn > 0: also writes the next group's lines S_Rn into Sn
    static hipError_t wloops(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& S,
                             const WGroupHost* groups, int group, const int* slots, void* partial, int ncols,
                             const RawPlanes& Sn, int Rn);
    // partial: 2 rows of 2*ceil(vh/256) columns (sum q, sum |q|); the improved
    // charge (kind 0) reads the line fields L2, L3 that qlines builds from P
    static hipError_t qlines(hipStream_t s, const Geo& g, const RawPlanes& P, const RawPlanes& L2,
                             const RawPlanes& L3);
    static hipError_t qtop(hipStream_t s, const Geo& g, const RawPlanes& P, int kind, const RawPlanes& L2,
                           const RawPlanes& L3, void* partial, double* density);
    // adjoint of APE, mode 0 spatial / 1 all / 2 temporal (vjp_kernels.hpp)
    static hipError_t ape_save(hipStream_t s, const Geo& g, const RawPlanes& in, const RawPlanes& out,
                               const RawNorms& mn, int mode, double sp);
    static hipError_t vjp_tangent(hipStream_t s, const Geo& g, const RawPlanes& U, const RawPlanes& G,
                                  const RawNorms& mn, const RawPlanes& H, int mode);
    static hipError_t vjp_step(hipStream_t s, const Geo& g, const RawPlanes& Z, const RawPlanes& Hin,
                               const RawNorms* mprev, const RawPlanes& out, int mode, double sp);
    // the straight-line fields of the loop gradient (lines_bytes of them, vjp_kernels.hpp)
    static hipError_t lines(hipStream_t s, const Geo& g, const RawPlanes& U, void* base, int R, int Tt);
    static hipError_t wloop_grad(hipStream_t s, const Geo& g, void* lines, int R, int Tt, const RawPlanes& G);
    // one link at lexicographic site n: its staple (out: 4 doubles), and its
    // heatbath from nu injected uniforms (link_out: 4 doubles, status: 2 ints)
    static hipError_t staple_link(hipStream_t s, const Geo& g, const RawPlanes& P, uint32_t n, int mu, double* out);
    static hipError_t local_heatbath(hipStream_t s, const Geo& g, const RawPlanes& P, uint32_t n, int mu, double beta,
                                     const double* u, int nu, double* link_out, int* status);
};

// precision-independent launchers (lgt_f64.hip)
hipError_t launch_reduce_rows(hipStream_t s, const void* in, int nrows, int ncols, void* out);
hipError_t launch_hb_harness(hipStream_t s, int n, const double* staples, double beta, const double* uniforms,
                             const int* offsets, const int* lens, double* links, int* counts);
hipError_t launch_or_harness(hipStream_t s, int n, const double* staples, const double* links_in, double* links);
// plaquette of an FP64 field in the reference layout; partial: ceil(VOL/256) double-doubles
hipError_t launch_plaquette_aos(hipStream_t s, const Geo& g, const double* aos, void* partial);
hipError_t launch_metro_harness(hipStream_t s, int n, const double* staples, const double* links_in, double beta,
                                double eps, const double* uniforms, double* links, int* accepted);

inline uint32_t grid_x(const Geo& g) { return (g.vh + 255u) / 256u; }

// records the message lgt_last_error returns (no allocation); returns code
int set_error(int code, const char* fmt, ...);

extern template struct Launch<double>;
extern template struct Launch<float>;

}  // namespace lgt
