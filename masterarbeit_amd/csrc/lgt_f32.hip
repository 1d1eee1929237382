// lgt_f32.hip -- FP32-storage instantiation of the engine kernels: FP32 link
// arithmetic (contraction to FMA allowed), FP64 global accumulations.
#include "launch_impl.hpp"

namespace lgt {

template struct Launch<float>;

}  // namespace lgt
