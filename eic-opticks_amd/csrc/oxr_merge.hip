// oxr_merge.hip -- device hit merging, the role of SPM::merge_partial_select
// (sysrap/SPM.cu:135-260) behind QEvt::gatherHitMerged / gatherHitLiteMerged /
// FinalMerge (qudarap/QEvt.cc:1034-1086,1159-1300).
//
//   select  items whose flagmask shares a bit with the mask, order kept
//           (select_pred: sphoton.h:300-304, sphotonlite.h:93-97)
//   key     (identity << 48) | unsigned(time / window), 64-bit arithmetic
//           (key_functor: sphoton.h:277-286, sphotonlite.h:68-77)
//   sort    stable radix sort of (key, item index)
//   fold    every run of equal keys into one item (reduce_op: sphoton.h:288-298,
//           sphotonlite.h:79-90): sphoton keeps the item with the earliest time (the
//           earlier one on equal times), sphotonlite keeps the run's first item with the
//           run's minimum time; flagmasks are OR-ed and 16-bit hitcounts summed.
// A window of 0 returns the selection unmerged (NOMERGE_TIME_WINDOW, SPM.hh:25); an
// already selected input skips the selection (ALREADY_HITMASK_SELECTED, SPM.hh:23). The fold is a left
// fold in sorted order; the reduce operations are associative, so this equals any
// grouping a parallel reduce-by-key would use.
#include "oxr_kernels.h"
#include <hipcub/hipcub.hpp>

namespace oxd {

// 32-bit words per item and the word offsets the merge reads
template <bool LITE> struct MW;
template <> struct MW<false> { static constexpr int n = 16, time = 3, fm = 15; };   // sphoton
template <> struct MW<true> { static constexpr int n = 4, time = 1, fm = 3; };      // sphotonlite

template <bool LITE>
OXD uint32_t merge_identity(const uint32_t* it) {
    return LITE ? (it[0] & 0xffffu) : (it[13] & 0x00ffffffu);   // identity(), get_identity()
}
template <bool LITE>
OXD uint32_t merge_hitcount(const uint32_t* it) { return (LITE ? it[0] : it[7]) >> 16; }

// unsigned(time / window) with the saturating float->unsigned conversion
OXD uint32_t time_bucket(float t, float tw) {
    float q = t / tw;
    if (!(q > 0.f)) return 0u;
    if (q >= 4294967296.f) return 0xffffffffu;
    return uint32_t(q);
}

template <bool LITE>
__global__ void __launch_bounds__(256) k_merge_flags(const uint32_t* __restrict__ in, long long n, unsigned mask,
                                                     uint8_t* __restrict__ flag) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (in[MW<LITE>::n * i + MW<LITE>::fm] & mask) != 0u ? 1 : 0;
}

template <bool LITE>
__global__ void __launch_bounds__(256) k_merge_keys(const uint32_t* __restrict__ in, const uint32_t* __restrict__ sel,
                                                    long long m, float tw, unsigned long long* __restrict__ key,
                                                    uint32_t* __restrict__ val) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    uint32_t i = sel ? sel[j] : uint32_t(j);
    const uint32_t* it = in + size_t(MW<LITE>::n) * i;
    key[j] = ((unsigned long long)merge_identity<LITE>(it) << 48) |
             (unsigned long long)time_bucket(__uint_as_float(it[MW<LITE>::time]), tw);
    val[j] = i;
}

template <bool LITE>
__global__ void __launch_bounds__(256) k_merge_gather(const uint32_t* __restrict__ in, const uint32_t* __restrict__ sel,
                                                      long long m, uint32_t* __restrict__ out) {
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const uint32_t* it = in + size_t(MW<LITE>::n) * (sel ? sel[j] : uint32_t(j));
    uint32_t* o = out + size_t(MW<LITE>::n) * j;
#pragma unroll
    for (int w = 0; w < MW<LITE>::n; w++) o[w] = it[w];
}

// one lane per run of equal keys
template <bool LITE>
__global__ void __launch_bounds__(256) k_merge_fold(const uint32_t* __restrict__ in, const uint32_t* __restrict__ order,
                                                    const int* __restrict__ start, const int* __restrict__ count,
                                                    int nruns, uint32_t* __restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    const int s0 = start[r], c = count[r];
    uint32_t rep = order[s0];
    const uint32_t* a = in + size_t(MW<LITE>::n) * rep;
    float t = __uint_as_float(a[MW<LITE>::time]);
    uint32_t fm = a[MW<LITE>::fm];
    uint32_t hc = merge_hitcount<LITE>(a);
    for (int k = 1; k < c; k++) {
        uint32_t bi = order[s0 + k];
        const uint32_t* b = in + size_t(MW<LITE>::n) * bi;
        float tb = __uint_as_float(b[MW<LITE>::time]);
        if (LITE) {
            t = fminf(t, tb);                       // r = a; r.time = fminf(a.time, b.time)
        } else {
            bool a_first = fminf(t, tb) == t;       // r = a_first ? a : b
            if (!a_first) { rep = bi; t = tb; }
        }
        fm |= b[MW<LITE>::fm];
        hc = (hc + merge_hitcount<LITE>(b)) & 0xffffu;
    }
    const uint32_t* it = in + size_t(MW<LITE>::n) * rep;
    uint32_t* o = out + size_t(MW<LITE>::n) * r;
#pragma unroll
    for (int w = 0; w < MW<LITE>::n; w++) o[w] = it[w];
    o[MW<LITE>::fm] = fm;
    if (LITE) {
        o[0] = (hc << 16) | (it[0] & 0xffffu);     // set_hitcount_identity(hc, a.identity())
        o[MW<LITE>::time] = __float_as_uint(t);
    } else {
        o[7] = (it[7] & 0xffffu) | (hc << 16);       // set_hitcount
    }
}

} // namespace oxd

namespace oxl {
using namespace oxd;

namespace {
struct Tmp {
    void* p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
    hipError_t get(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
};
inline unsigned nblk(long long n) { return unsigned((n + 255) / 256); }
}

template <bool LITE>
static hipError_t merge_t(hipStream_t s, const uint32_t* in, long long n, unsigned mask, bool select, float tw,
                          uint32_t* out, long long* num_out) {
    *num_out = 0;
    if (n <= 0) return hipSuccess;
    if (n > 0x7fffffffLL) return hipErrorInvalidValue;
    const int N = int(n);
    hipError_t e;
    Tmp sel, flag, nsel, tmp;
    int m = N;
    const uint32_t* psel = nullptr;
    if (select) {
        if ((e = flag.get(size_t(N))) != hipSuccess) return e;
        if ((e = sel.get(size_t(N) * 4)) != hipSuccess) return e;
        if ((e = nsel.get(sizeof(int))) != hipSuccess) return e;
        hipLaunchKernelGGL(k_merge_flags<LITE>, dim3(nblk(N)), dim3(256), 0, s, in, n, mask, (uint8_t*)flag.p);
        hipcub::CountingInputIterator<uint32_t> idx(0u);
        size_t bytes = 0;
        if ((e = hipcub::DeviceSelect::Flagged(nullptr, bytes, idx, (uint8_t*)flag.p, (uint32_t*)sel.p, (int*)nsel.p, N, s)) != hipSuccess) return e;
        if ((e = tmp.get(bytes)) != hipSuccess) return e;
        if ((e = hipcub::DeviceSelect::Flagged(tmp.p, bytes, idx, (uint8_t*)flag.p, (uint32_t*)sel.p, (int*)nsel.p, N, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(&m, nsel.p, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        psel = (const uint32_t*)sel.p;
    }
    if (m == 0) return hipSuccess;
    if (tw == 0.f) {
        hipLaunchKernelGGL(k_merge_gather<LITE>, dim3(nblk(m)), dim3(256), 0, s, in, psel, (long long)m, out);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        *num_out = m;
        return hipStreamSynchronize(s);
    }
    Tmp k0, k1, v0, v1, uk, cnt, st, nr, tmp2, tmp3, tmp4;
    if ((e = k0.get(size_t(m) * 8)) != hipSuccess || (e = k1.get(size_t(m) * 8)) != hipSuccess ||
        (e = v0.get(size_t(m) * 4)) != hipSuccess || (e = v1.get(size_t(m) * 4)) != hipSuccess ||
        (e = uk.get(size_t(m) * 8)) != hipSuccess || (e = cnt.get(size_t(m) * 4)) != hipSuccess ||
        (e = st.get(size_t(m) * 4)) != hipSuccess || (e = nr.get(sizeof(int))) != hipSuccess)
        return e;
    auto* K0 = (unsigned long long*)k0.p; auto* K1 = (unsigned long long*)k1.p;
    auto* V0 = (uint32_t*)v0.p; auto* V1 = (uint32_t*)v1.p;
    hipLaunchKernelGGL(k_merge_keys<LITE>, dim3(nblk(m)), dim3(256), 0, s, in, psel, (long long)m, tw, K0, V0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t b2 = 0, b3 = 0, b4 = 0;
    if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, b2, K0, K1, V0, V1, m, 0, 64, s)) != hipSuccess) return e;
    if ((e = tmp2.get(b2)) != hipSuccess) return e;
    if ((e = hipcub::DeviceRadixSort::SortPairs(tmp2.p, b2, K0, K1, V0, V1, m, 0, 64, s)) != hipSuccess) return e;
    if ((e = hipcub::DeviceRunLengthEncode::Encode(nullptr, b3, K1, (unsigned long long*)uk.p, (int*)cnt.p, (int*)nr.p, m, s)) != hipSuccess) return e;
    if ((e = tmp3.get(b3)) != hipSuccess) return e;
    if ((e = hipcub::DeviceRunLengthEncode::Encode(tmp3.p, b3, K1, (unsigned long long*)uk.p, (int*)cnt.p, (int*)nr.p, m, s)) != hipSuccess) return e;
    int nruns = 0;
    if ((e = hipMemcpyAsync(&nruns, nr.p, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(nullptr, b4, (int*)cnt.p, (int*)st.p, nruns, s)) != hipSuccess) return e;
    if ((e = tmp4.get(b4)) != hipSuccess) return e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(tmp4.p, b4, (int*)cnt.p, (int*)st.p, nruns, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_merge_fold<LITE>, dim3(nblk(nruns)), dim3(256), 0, s, in, V1, (const int*)st.p,
                       (const int*)cnt.p, nruns, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    *num_out = nruns;
    return hipStreamSynchronize(s);
}

hipError_t merge(hipStream_t s, bool lite, const void* in, long long n, unsigned select_mask, bool select,
                 float time_window, void* out, long long* num_out) {
    if (lite) return merge_t<true>(s, (const uint32_t*)in, n, select_mask, select, time_window, (uint32_t*)out, num_out);
    return merge_t<false>(s, (const uint32_t*)in, n, select_mask, select, time_window, (uint32_t*)out, num_out);
}

} // namespace oxl
