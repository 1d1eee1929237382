// lgt_noalloc.hip -- lgt_plaquette_aos_noalloc (include/lgt_su2.h): the
// plaquette of lgt_upload_plaquette for callers that cannot allocate device
// memory.  The field and the result are device arrays the caller owns; the
// partial sums live in a file-scope device array of fixed size (`scratch`),
// whose device address is looked up once per device.  Compiled like
// lgt_f64.hip (no FMA contraction): the same kernels on the same input give
// the same bits as lgt_upload_plaquette.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lgt_su2.h"
#include "kernels.hpp"
#include "launch.hpp"

// one double-double partial per 256 lexicographic sites, plus the total:
// (96^4 / 256 + 1) * 16 bytes = 5.3 MB on every device the library runs on
constexpr int64_t NOALLOC_PARTIALS = (int64_t)LGT_NOALLOC_MAX_VOL / lgt::BLOCK;
static __device__ __align__(256) lgt::DD scratch[NOALLOC_PARTIALS + 1];

namespace {

using namespace lgt;

// plaq_dev[0] = (sum.hi + sum.lo) / (12 VOL), as lgt_upload_plaquette rounds it
static __global__ void __launch_bounds__(64) k_plaquette_mean(const DD* __restrict__ sum, double denom,
                                                              double* __restrict__ out)
{
    if (threadIdx.x == 0) out[0] = (sum->hi + sum->lo) / denom;
}

constexpr int MAX_DEVICES = 64;
std::mutex g_mutex;
DD* g_scratch[MAX_DEVICES] = {};   // device address of `scratch` on each device

// the device address of this device's copy of `scratch`, looked up on the
// first call on the device (the host-side name is not a device address)
hipError_t scratch_of(int dev, DD** out)
{
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess) return err;
        g_scratch[dev] = (DD*)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int lgt_plaquette_aos_noalloc(const int extents[4], const double* field_dev, double* plaq_dev, void* stream)
{
    if (!extents || !field_dev || !plaq_dev) return lgt::set_error(LGT_EINVAL, "lgt_plaquette_aos_noalloc: null argument");
    int64_t vol = 1;
    for (int d = 0; d < 4; d++) {
        if (extents[d] < 2 || (extents[d] & 1) || extents[d] > 4096)
            return lgt::set_error(LGT_EINVAL, "lgt_plaquette_aos_noalloc: extent[%d] = %d must be even and in [2, 4096]", d,
                             extents[d]);
        vol *= extents[d];
    }
    if (extents[3] < 4) return lgt::set_error(LGT_EINVAL, "lgt_plaquette_aos_noalloc: extent[3] = %d must be >= 4", extents[3]);
    if (vol > (int64_t)LGT_NOALLOC_MAX_VOL)
        return lgt::set_error(LGT_EINVAL, "lgt_plaquette_aos_noalloc: volume %lld exceeds the scratch limit %lld (96^4)",
                         (long long)vol, (long long)LGT_NOALLOC_MAX_VOL);
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return lgt::set_error(LGT_EDEVICE, "hipGetDevice: %s", hipGetErrorString(err));
    if (dev < 0 || dev >= MAX_DEVICES) return lgt::set_error(LGT_EDEVICE, "lgt_plaquette_aos_noalloc: device %d", dev);
    DD* sc = nullptr;
    err = scratch_of(dev, &sc);
    if (err != hipSuccess) return lgt::set_error(LGT_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(err));

    Geo g{};
    for (int d = 0; d < 4; d++) g.L[d] = extents[d];
    g.hz = extents[3] / 2;
    g.vh = (uint32_t)(vol / 2);
    const int nb = (int)((vol + BLOCK - 1) / BLOCK);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_plaquette_aos, dim3(nb), dim3(BLOCK), 0, s, field_dev, g, sc);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(REDUCE_THREADS), 0, s, (const DD*)sc, nb, sc + NOALLOC_PARTIALS);
    hipLaunchKernelGGL(k_plaquette_mean, dim3(1), dim3(64), 0, s, (const DD*)(sc + NOALLOC_PARTIALS),
                       (double)(2 * 6 * vol), plaq_dev);
    err = hipGetLastError();
    if (err != hipSuccess) return lgt::set_error(LGT_EDEVICE, "lgt_plaquette_aos_noalloc: launch: %s", hipGetErrorString(err));
    return LGT_OK;
}
