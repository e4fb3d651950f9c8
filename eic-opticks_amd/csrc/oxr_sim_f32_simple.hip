// oxr_sim_f32_simple.hip -- production simulate kernels for geometries whose leaves are all
// spheres, zspheres, cylinders or boxes (single-precision roots, no list nodes): the other
// leaves' code is left out (intersect_leaf<FP64, SIMPLE>), traversal mode tm + 16
#include "oxr_sim.h"

namespace oxl {

hipError_t launch_simulate_simple(hipStream_t s, const oxd::Params& P, dim3 g, dim3 b, size_t sm, bool persist, int tm) {
#define OXR_S(T)                                                                                        \
    do {                                                                                                \
        if (persist) hipLaunchKernelGGL((oxd::k_simulate_persist<false, false, false, 16 + T>), g, b, sm, s, P); \
        else hipLaunchKernelGGL((oxd::k_simulate<false, false, false, false, 16 + T>), g, b, sm, s, P);  \
    } while (0)
    switch (tm & 15) {
        case 0: OXR_S(0); break;   case 1: OXR_S(1); break;   case 2: OXR_S(2); break;   case 3: OXR_S(3); break;
        case 4: OXR_S(4); break;   case 5: OXR_S(5); break;   case 6: OXR_S(6); break;   case 7: OXR_S(7); break;
        case 8: OXR_S(8); break;   case 9: OXR_S(9); break;   case 10: OXR_S(10); break; case 11: OXR_S(11); break;
        case 12: OXR_S(12); break; case 13: OXR_S(13); break; case 14: OXR_S(14); break; default: OXR_S(15); break;
    }
#undef OXR_S
    return hipGetLastError();
}

const void* persist_kernel_simple(int tm) {
#define OXR_K(T) return reinterpret_cast<const void*>(&oxd::k_simulate_persist<false, false, false, 16 + T>)
    switch (tm & 15) {
        case 0: OXR_K(0);   case 1: OXR_K(1);   case 2: OXR_K(2);   case 3: OXR_K(3);
        case 4: OXR_K(4);   case 5: OXR_K(5);   case 6: OXR_K(6);   case 7: OXR_K(7);
        case 8: OXR_K(8);   case 9: OXR_K(9);   case 10: OXR_K(10); case 11: OXR_K(11);
        case 12: OXR_K(12); case 13: OXR_K(13); case 14: OXR_K(14); default: OXR_K(15);
    }
#undef OXR_K
}

} // namespace oxl
