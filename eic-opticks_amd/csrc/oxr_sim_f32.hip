// oxr_sim_f32.hip -- simulate kernels with single-precision roots
#include "oxr_sim.h"

namespace oxl {
template hipError_t launch_simulate_plain<false, false>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, int, int);
template hipError_t launch_simulate_persist<false, false>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int);
} // namespace oxl
