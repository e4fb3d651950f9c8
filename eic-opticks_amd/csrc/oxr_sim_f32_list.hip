// oxr_sim_f32_list.hip -- simulate kernels with single-precision roots and list-node support
#include "oxr_sim.h"

namespace oxl {
template hipError_t launch_simulate_plain<false, true>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, int, int);
template hipError_t launch_simulate_persist<false, true>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int);
} // namespace oxl
