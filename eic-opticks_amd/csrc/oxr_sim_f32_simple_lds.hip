// oxr_sim_f32_simple_lds.hip -- the simple-leaf production kernels that also read the node
// and itra arrays from a copy in LDS (stage_csg), traversal mode tm + 48; launched when the
// engine makes that copy (Params::csg_node4 > 0)
#include "oxr_sim.h"

namespace oxl {
template hipError_t launch_simulate_simple<48>(hipStream_t, const oxd::Params&, dim3, dim3, size_t, bool, int);
template const void* persist_kernel_simple<48>(int);
} // namespace oxl
