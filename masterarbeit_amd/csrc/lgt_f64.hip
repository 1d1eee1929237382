// lgt_f64.hip -- FP64 instantiation of the engine kernels.  Compiled with
// -ffp-contract=off: every product and sum rounds separately, in the order of
// the reference's FP64 C code, so per-site values equal the oracle's bit for
// bit on the same inputs.
#include "launch_impl.hpp"

namespace lgt {

template struct Launch<double>;

hipError_t launch_reduce_rows(hipStream_t s, const void* in, int nrows, int ncols, void* out)
{
    hipLaunchKernelGGL(k_reduce_rows, dim3(nrows), dim3(REDUCE_THREADS), 0, s, (const DD*)in, ncols, (DD*)out);
    return hipGetLastError();
}

hipError_t launch_hb_harness(hipStream_t s, int n, const double* staples, double beta, const double* uniforms,
                             const int* offsets, const int* lens, double* links, int* counts)
{
    hipLaunchKernelGGL(k_hb_harness, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, n, staples, beta, uniforms,
                       offsets, lens, links, counts);
    return hipGetLastError();
}

hipError_t launch_plaquette_aos(hipStream_t s, const Geo& g, const double* aos, void* partial)
{
    hipLaunchKernelGGL(k_plaquette_aos, dim3((2 * g.vh + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, aos, g, (DD*)partial);
    return hipGetLastError();
}

hipError_t launch_or_harness(hipStream_t s, int n, const double* staples, const double* links_in, double* links)
{
    hipLaunchKernelGGL(k_or_harness, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, n, staples, links_in, links);
    return hipGetLastError();
}

hipError_t launch_metro_harness(hipStream_t s, int n, const double* staples, const double* links_in, double beta,
                                double eps, const double* uniforms, double* links, int* accepted)
{
    hipLaunchKernelGGL(k_metro_harness, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, n, staples, links_in, beta,
                       eps, uniforms, links, accepted);
    return hipGetLastError();
}

}  // namespace lgt
