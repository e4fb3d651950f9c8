// oxr_sim_f32_simple.hip -- production simulate kernels for geometries whose leaves are all
// spheres, zspheres, cylinders or boxes (single-precision roots, no list nodes): the other
// leaves' code is left out (intersect_leaf<FP64, SIMPLE>), traversal mode tm + 16
#include "oxr_sim.h"

namespace oxl {
template hipError_t launch_simulate_simple<16>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int);
template const void* persist_kernel_simple<16>(int);
} // namespace oxl
