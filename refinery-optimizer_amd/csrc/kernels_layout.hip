// gfx950 kernel that lays a tile set's matrix copy out in gather order on
// the device: every chunk of every tile is sorted by the index of the
// vector entry it gathers (ties in storage order), its index words become
// packed (local << log2 tn | position) words, and the value arrays are
// permuted alike.  One workgroup per tile sorts each chunk in LDS with a
// bitonic network; the packed word is the sort key, and it is unique (the
// position bits), so the result is the same total order as a stable sort by
// gathered index.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "kernels_common.hpp"

namespace mpdhg {
namespace dev {

constexpr int kSortBlock = 256;

template <int TN>
__global__ __launch_bounds__(kSortBlock) void k_gather_order(const Tile *__restrict__ tiles, int ntiles,
                                                             const int *__restrict__ src_idx,
                                                             int *__restrict__ dst_idx,
                                                             const float *__restrict__ s32,
                                                             float *__restrict__ d32,
                                                             const double *__restrict__ s64,
                                                             double *__restrict__ d64) {
    __shared__ uint32_t key[TN];
    constexpr int kBits = ilog2(TN);
    constexpr uint32_t kMask = static_cast<uint32_t>(TN) - 1u;
    const int t = static_cast<int>(blockIdx.x);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const int tid = static_cast<int>(threadIdx.x);
    if (dense_group<TN>(tl.end - tl.begin, tl.lo, tl.hi)) {
        // A dense row-group tile stays in storage order (the split rule's
        // kernels stream its rows from their own first nonzero): the packed
        // words keep each entry's own position, as an identity sort would.
        for (int k = tl.lo + tid; k < tl.hi; k += kSortBlock) {
            const int p = (k - tl.lo) & static_cast<int>(kMask);
            dst_idx[k] = static_cast<int>((static_cast<uint32_t>(src_idx[k] - tl.base) << kBits) |
                                          static_cast<uint32_t>(p));
            if (d32) d32[k] = s32[k];
            if (d64) d64[k] = s64[k];
        }
        return;
    }
    for (int c0 = tl.lo; c0 < tl.hi; c0 += TN) {
        const int n = min(tl.hi - c0, TN);
        int np2 = 1;
        while (np2 < n) np2 <<= 1;
        for (int p = tid; p < np2; p += kSortBlock)
            key[p] = p < n ? ((static_cast<uint32_t>(src_idx[c0 + p] - tl.base) << kBits) |
                              static_cast<uint32_t>(p))
                           : 0xFFFFFFFFu;
        __syncthreads();
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = tid; i < (np2 >> 1); i += kSortBlock) {
                    const int lo = 2 * i - (i & (stride - 1));
                    const int hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    const uint32_t a = key[lo], b = key[hi];
                    if ((a > b) == up) {
                        key[lo] = b;
                        key[hi] = a;
                    }
                }
                __syncthreads();
            }
        }
        for (int p = tid; p < n; p += kSortBlock) {
            const uint32_t w = key[p];
            const int pos = static_cast<int>(w & kMask);
            dst_idx[c0 + p] = static_cast<int>(w);
            if (d32) d32[c0 + p] = s32[c0 + pos];
            if (d64) d64[c0 + p] = s64[c0 + pos];
        }
        __syncthreads();
    }
}

}  // namespace dev

namespace launch {

void gather_order(hipStream_t s, const TileSet &ts, const int *src_idx, int *dst_idx, const float *s32,
                  float *d32, const double *s64, double *d64) {
    if (!ts.count) return;
    const dim3 g(static_cast<unsigned>(ts.count)), b(dev::kSortBlock);
    switch (ts.tn) {
        case 1024:
            hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_gather_order<1024>), g, b, 0, s, ts.tiles, ts.count,
                               src_idx, dst_idx, s32, d32, s64, d64);
            break;
        case 2048:
            hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_gather_order<2048>), g, b, 0, s, ts.tiles, ts.count,
                               src_idx, dst_idx, s32, d32, s64, d64);
            break;
        case 4096:
            hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_gather_order<4096>), g, b, 0, s, ts.tiles, ts.count,
                               src_idx, dst_idx, s32, d32, s64, d64);
            break;
        default:
            hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_gather_order<8192>), g, b, 0, s, ts.tiles, ts.count,
                               src_idx, dst_idx, s32, d32, s64, d64);
            break;
    }
}

}  // namespace launch
}  // namespace mpdhg
