// Device helpers shared by the tile kernels (HIP translation units only).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "launch.hpp"

namespace mpdhg {
namespace dev {

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remap so that each XCD walks a contiguous range of tiles: tiles
// that touch neighbouring rows/columns then share that XCD's L2.  The grid
// is padded to 8 * chunk; the map is a bijection on [0, 8 * chunk).
__device__ __forceinline__ int xcd_tile(unsigned b, int chunk) {
    return static_cast<int>(b & 7u) * chunk + static_cast<int>(b >> 3);
}

// s + v * x with one rounding, in the accumulator type A.
template <typename A, typename V>
__device__ __forceinline__ A fma_acc(V v, V x, A s);
template <>
__device__ __forceinline__ float fma_acc<float, float>(float v, float x, float s) {
    return __builtin_fmaf(v, x, s);
}
template <>
__device__ __forceinline__ double fma_acc<double, float>(float v, float x, double s) {
    return __builtin_fma(static_cast<double>(v), static_cast<double>(x), s);
}
template <>
__device__ __forceinline__ double fma_acc<double, double>(double v, double x, double s) {
    return __builtin_fma(v, x, s);
}

template <typename V, typename X>
struct Pair {
    V v;
    X x;
};

// LDS bytes a tile kernel with chunk size TN reserves for the staged
// (value, x) pairs: one chunk of 8-byte pairs (FP32 value and x); 16-byte
// pairs (FP64) are staged in two passes over the same chunk.
constexpr int stage_bytes(int tn) { return tn * 8; }

// The row-pointer reads of a tile: issued first, so that a kernel can put
// its other independent loads (epilogue operands, instance state) in flight
// before it consumes them.
struct TileRows {
    int k0, k1, lo, hi, base, rows, first;
};

__device__ __forceinline__ TileRows tile_rows(const Tile &tl, const int *__restrict__ ptr) {
    TileRows t;
    const int r = tl.begin + static_cast<int>(threadIdx.x);
    t.k0 = 0;
    t.k1 = 0;
    if (r < tl.end) {
        t.k0 = ptr[r];
        t.k1 = ptr[r + 1];
    }
    t.lo = tl.lo;
    t.hi = tl.hi;
    t.base = tl.base;
    t.rows = tl.end - tl.begin;
    t.first = tl.begin;
    return t;
}

// A tile of one dense row group: at most kGroup rows holding more nonzeros
// than one chunk (the tile planner gives such a group a tile of its own).
template <int TN>
__device__ __forceinline__ bool dense_group(int rows, int lo, int hi) {
    return rows <= kGroup && hi - lo > TN;
}

// Refreshes the run hints of a tile list from the instance statuses (after
// every kernel that may stop an instance).
template <typename I>
__global__ __launch_bounds__(256) void k_tile_run(Tile *__restrict__ tiles, int ntiles,
                                                  const I *__restrict__ st) {
    const int t = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (t < ntiles) tiles[t].run = st[tiles[t].inst].status == kRunning ? 1 : 0;
}

// Row sums of one tile: sum_k val[k] * xs[col(k)] for every row of the
// tile, in storage order (k ascending), one row per lane of a workgroup of
// B threads.  Each chunk of TN
// nonzeros is streamed with coalesced loads of the index words and values,
// the gathers are issued back to back, and the (value, x) pairs are parked
// in LDS at their storage positions (the position is k - chunk start in
// storage order, G = false, or the low log2 TN bits of the index word in
// gather order, G = true); each lane then walks its own row from LDS.  A
// row longer than a chunk is continued across chunks by the same lane, so
// the summation order never depends on the layout or on TN.  `lds` must
// hold stage_bytes(TN).  The layout is a template parameter (one per tile
// set) so that each instantiation carries one addressing path: two paths in
// one kernel cost a third more registers and occupancy with it.  NT reads
// the index and value stream with non-temporal loads; it is a template
// parameter because a uniform branch between the two load forms is merged
// by the compiler into one form, without the hint.
// `init` starts each lane's accumulator: a row split into column bands
// (products over one band of columns per launch) continues its sequential
// FMA chain from the sum the previous band stored, so the summation order is
// still the row's storage order.
// Row sums of a dense row-group tile (dense_group) with the split rule of
// kSplitRow.  The tile's nonzeros are in storage order (the gather-order
// layout leaves dense tiles unsorted), so a pair's LDS slot is its offset in
// the chunk; chunks hold CS pairs, one staging pass.
// - A split row is streamed on its own, in chunks that start at its first
//   nonzero (CS divides kSplitBlockNnz or is a multiple of it, so neither a
//   segment nor a block straddles two chunks): its segments are summed one
//   per lane, and its owner lane folds the segment sums into block sums and
//   the block sums into the row sum, in order.  With `pre` (and rowseg) the
//   block sums come from the pre-pass (k_split_seg, the same blocks) and the
//   owner lane only adds them in order.
// - A run of consecutive other rows is streamed together, each lane walking
//   its own row as in tile_dot_rows.
template <int B, int TN, bool G, typename V, typename X, typename A, bool NT>
__device__ A dense_rows(const TileRows &tr, const int *__restrict__ idx, const V *__restrict__ val,
                        const X *__restrict__ xs, void *lds_raw, A init, const A *__restrict__ pre,
                        const int *__restrict__ rowseg) {
    constexpr int CS = stage_bytes(TN) / static_cast<int>(sizeof(Pair<V, X>));
    static_assert(CS % kSplitSeg == 0 && (kSplitBlockNnz % CS == 0 || CS % kSplitBlockNnz == 0), "aligned chunks");
    static_assert(CS / kSplitSeg <= B && CS % B == 0, "one segment per lane");
    constexpr int kBits = ilog2(TN);
    constexpr int PT = CS / B;
    constexpr int kSegsPerBlock = kSplitBlockNnz / kSplitSeg;
    __shared__ int rb[kGroup + 1];
    Pair<V, X> *lds = static_cast<Pair<V, X> *>(lds_raw);
    A *seg = static_cast<A *>(lds_raw);  // reused once a chunk's pairs are consumed
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < tr.rows) {
        rb[tid] = tr.k0;
        if (tid == tr.rows - 1) rb[tr.rows] = tr.k1;
    }
    __syncthreads();
    // stages chunk [c0, c1) of the stream as (value, x) pairs at offsets k - c0
    auto stage = [&](int c0, int c1) {
        const uint32_t last = static_cast<uint32_t>(c1 - 1 - c0);
        const int *__restrict__ ib = idx + c0;
        const V *__restrict__ vb = val + c0;
        int id[PT];
        V v[PT];
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const uint32_t o = min(static_cast<uint32_t>(tid + q * B), last);
            if constexpr (NT) {
                id[q] = __builtin_nontemporal_load(ib + o);
                v[q] = __builtin_nontemporal_load(vb + o);
            } else {
                id[q] = ib[o];
                v[q] = vb[o];
            }
        }
        X g[PT];
        if (G) {
            const X *__restrict__ xb = xs + tr.base;
#pragma unroll
            for (int q = 0; q < PT; ++q) g[q] = xb[static_cast<uint32_t>(id[q]) >> kBits];
        } else {
#pragma unroll
            for (int q = 0; q < PT; ++q) g[q] = xs[static_cast<uint32_t>(id[q])];
        }
#pragma unroll
        for (int q = 0; q < PT; ++q) lds[min(static_cast<uint32_t>(tid + q * B), last)] = Pair<V, X>{v[q], g[q]};
        __syncthreads();
    };
    A sum = init;
    for (int r = 0; r < tr.rows;) {
        const int k0 = rb[r], k1 = rb[r + 1];
        if (k1 - k0 > kSplitRow) {
            if (pre) {
                // all lanes stage the block sums in LDS with coalesced loads; the
                // owner lane adds them in order from there
                const A *__restrict__ sp = pre + rowseg[tr.first + r];
                const int nblk = (k1 - k0 + kSplitBlockNnz - 1) / kSplitBlockNnz;
                constexpr int cap = stage_bytes(TN) / static_cast<int>(sizeof(A));
                for (int t0 = 0; t0 < nblk; t0 += cap) {
                    const int cnt = min(cap, nblk - t0);
                    for (int t = tid; t < cnt; t += B) seg[t] = sp[t0 + t];
                    __syncthreads();
                    if (tid == r)
                        for (int t = 0; t < cnt; ++t) sum = sum + seg[t];
                    __syncthreads();
                }
            } else {
                A blk = A(0);
                int done = 0;  // segments of the row summed so far
                for (int c0 = k0; c0 < k1; c0 += CS) {
                    const int c1 = min(k1, c0 + CS), len = c1 - c0;
                    stage(c0, c1);
                    const int nseg = (len + kSplitSeg - 1) / kSplitSeg;
                    A sg = A(0);
                    if (tid < nseg) {
                        const int a = tid * kSplitSeg, e = min(len, a + kSplitSeg);
                        for (int k = a; k < e; ++k) sg = fma_acc<A, V>(lds[k].v, lds[k].x, sg);
                    }
                    __syncthreads();
                    if (tid < nseg) seg[tid] = sg;
                    __syncthreads();
                    if (tid == r) {
                        for (int t = 0; t < nseg; ++t) {
                            blk = blk + seg[t];
                            if (++done % kSegsPerBlock == 0) {
                                sum = sum + blk;
                                blk = A(0);
                            }
                        }
                        if (c1 == k1 && done % kSegsPerBlock != 0) sum = sum + blk;
                    }
                    __syncthreads();
                }
            }
            ++r;
            continue;
        }
        // a run of consecutive rows of at most kSplitRow nonzeros
        int r2 = r + 1;
        while (r2 < tr.rows && rb[r2 + 1] - rb[r2] <= kSplitRow) ++r2;
        const int lo = rb[r], hi = rb[r2];
        const bool mine = tid >= r && tid < r2;
        for (int c0 = lo; c0 < hi; c0 += CS) {
            const int c1 = min(hi, c0 + CS);
            stage(c0, c1);
            if (mine) {
                const int a = max(tr.k0, c0) - c0, e = min(tr.k1, c1) - c0;
                int k = a;
                for (; k + 4 <= e; k += 4) {
                    const Pair<V, X> p0 = lds[k], p1 = lds[k + 1], p2 = lds[k + 2], p3 = lds[k + 3];
                    sum = fma_acc<A, V>(p0.v, p0.x, sum);
                    sum = fma_acc<A, V>(p1.v, p1.x, sum);
                    sum = fma_acc<A, V>(p2.v, p2.x, sum);
                    sum = fma_acc<A, V>(p3.v, p3.x, sum);
                }
                for (; k < e; ++k) sum = fma_acc<A, V>(lds[k].v, lds[k].x, sum);
            }
            __syncthreads();
        }
        r = r2;
    }
    return sum;
}

// SPLIT: apply the split rule of kSplitRow (enhanced mode's FP32 products);
// dense row-group tiles then take dense_rows.  Without it every row is one
// sequential chain (reference mode, FP64 checks).
template <int B, int TN, bool G, typename V, typename X, typename A, bool NT = false, bool SPLIT = false>
__device__ __forceinline__ A tile_dot_rows(const TileRows &tr, const int *__restrict__ idx,
                                           const V *__restrict__ val, const X *__restrict__ xs,
                                           void *lds_raw, A init = A(0), const A *pre = nullptr,
                                           const int *rowseg = nullptr) {
    static_assert(valid_schedule(B, TN), "tile schedule");
    if constexpr (SPLIT) {
        if (dense_group<TN>(tr.rows, tr.lo, tr.hi))
            return dense_rows<B, TN, G, V, X, A, NT>(tr, idx, val, xs, lds_raw, init, pre, rowseg);
    }
    constexpr int kBits = ilog2(TN);
    constexpr uint32_t kMask = static_cast<uint32_t>(TN) - 1u;
    constexpr int S = stage_bytes(TN) / static_cast<int>(sizeof(Pair<V, X>));  // pairs per pass
    constexpr int NP = TN / S;                                                 // passes per chunk
    static_assert(S * NP == TN, "staging must tile the chunk");
    constexpr int PT = TN / B;
    Pair<V, X> *lds = static_cast<Pair<V, X> *>(lds_raw);
    const int tid = static_cast<int>(threadIdx.x);
    const int k0 = tr.k0, k1 = tr.k1, lo = tr.lo, hi = tr.hi;
    A sum = init;
    for (int c0 = lo; c0 < hi; c0 += TN) {
        const int c1 = min(hi, c0 + TN);
        // Lanes past the chunk end re-read its last entry and rewrite the same
        // pair to the same LDS slot: no per-entry branches, every load issued.
        // Addresses as a uniform base plus an unsigned 32-bit offset, so that
        // every load needs one offset register rather than a 64-bit address.
        // A stream larger than the caches is read once per product:
        // non-temporal loads keep it from displacing the gathered vectors in
        // L2.  A stream that fits stays cached for the next product.
        const int *__restrict__ ib = idx + c0;
        const V *__restrict__ vb = val + c0;
        const uint32_t last = static_cast<uint32_t>(c1 - 1 - c0);
        int id[PT];
        V v[PT];
        if constexpr (NT) {
#pragma unroll
            for (int q = 0; q < PT; ++q) {
                const uint32_t o = min(static_cast<uint32_t>(tid + q * B), last);
                id[q] = __builtin_nontemporal_load(ib + o);
                v[q] = __builtin_nontemporal_load(vb + o);
            }
        } else {
#pragma unroll
            for (int q = 0; q < PT; ++q) {
                const uint32_t o = min(static_cast<uint32_t>(tid + q * B), last);
                id[q] = ib[o];
                v[q] = vb[o];
            }
        }
        X g[PT];
        if (G) {
            const X *__restrict__ xb = xs + tr.base;
#pragma unroll
            for (int q = 0; q < PT; ++q) g[q] = xb[static_cast<uint32_t>(id[q]) >> kBits];
        } else {
#pragma unroll
            for (int q = 0; q < PT; ++q) g[q] = xs[static_cast<uint32_t>(id[q])];
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
#pragma unroll
            for (int q = 0; q < PT; ++q) {
                const int d = G ? static_cast<int>(static_cast<uint32_t>(id[q]) & kMask)
                                : static_cast<int>(min(static_cast<uint32_t>(tid + q * B), last));
                if (NP == 1 || d / S == p) lds[d - p * S] = Pair<V, X>{v[q], g[q]};
            }
            __syncthreads();
            const int a = max(k0, c0 + p * S), b = min(k1, min(c1, c0 + (p + 1) * S));
            // four LDS reads in flight ahead of the (sequential) FMA chain
            const Pair<V, X> *row = lds - (c0 + p * S);
            int k = a;
            for (; k + 4 <= b; k += 4) {
                const Pair<V, X> p0 = row[k], p1 = row[k + 1], p2 = row[k + 2], p3 = row[k + 3];
                sum = fma_acc<A, V>(p0.v, p0.x, sum);
                sum = fma_acc<A, V>(p1.v, p1.x, sum);
                sum = fma_acc<A, V>(p2.v, p2.x, sum);
                sum = fma_acc<A, V>(p3.v, p3.x, sum);
            }
            for (; k < b; ++k) sum = fma_acc<A, V>(row[k].v, row[k].x, sum);
            __syncthreads();
        }
    }
    return sum;
}

template <int B, int TN, bool G, typename V, typename X, typename A>
__device__ __forceinline__ A tile_dot_t(const Tile &tl, const int *__restrict__ ptr,
                                        const int *__restrict__ idx, const V *__restrict__ val,
                                        const X *__restrict__ xs, void *lds_raw) {
    return tile_dot_rows<B, TN, G, V, X, A>(tile_rows(tl, ptr), idx, val, xs, lds_raw);
}

template <int B, int TN, bool G, typename A>
__device__ __forceinline__ A tile_dot(const Tile &tl, const int *__restrict__ ptr,
                                      const int *__restrict__ idx, const float *__restrict__ val,
                                      const float *__restrict__ xs, void *lds) {
    return tile_dot_t<B, TN, G, float, float, A>(tl, ptr, idx, val, xs, lds);
}

// Block-wide max of non-negative floats (NaN operands are dropped, as
// fmaxf does) over NT threads; every lane receives the result.
template <int NT>
__device__ __forceinline__ float block_max(float v, float *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) r = fmaxf(r, red[q]);
    return r;
}

template <int NT>
__device__ __forceinline__ double block_max_d(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) r = fmax(r, red[q]);
    return r;
}

// Block-wide FP64 sum over NT threads in a fixed tree order; every lane
// receives the result.
template <int NT>
__device__ __forceinline__ double block_sum_d(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) r += red[q];
    return r;
}

// FP64 sum / max over the kGroup lanes of a row group, in a fixed
// butterfly order; every lane of the group receives the result.
__device__ __forceinline__ double group_sum_d(double v) {
#pragma unroll
    for (int off = kGroup / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double group_max_d(double v) {
#pragma unroll
    for (int off = kGroup / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Stores a group partial: tiles start on group boundaries, so lanes
// [kGroup g, kGroup (g + 1)) of a tile hold exactly group tl.grp + g (rows
// past the tile end add 0).  Every lane of the workgroup must take part in
// the group reduction before.
__device__ __forceinline__ void put_group(double *part, const Tile &tl, double v) {
    const int l = static_cast<int>(threadIdx.x);
    if ((l & (kGroup - 1)) == 0 && tl.begin + l < tl.end) part[tl.grp + l / kGroup] = v;
}

// sum_{g in [b, e)} f(g) in FP64, in an order fixed by e - b alone: lane l
// adds f(b + l), f(b + l + kInstBlock), ... in ascending order, then the
// lanes are combined by the fixed block tree.  Run by the kInstBlock threads
// of a per-instance kernel over an instance's groups; with per-group
// partials formed by one wave each, the result does not depend on how the
// rows were tiled.  `red` holds kInstBlock / 64 doubles.
// A lane's walk issues the loads of kCanonUnroll steps before it adds them,
// in the same order, so that one instance's thousands of groups do not cost
// one memory latency per step.
constexpr int kCanonUnroll = 8;
template <typename F>
__device__ __forceinline__ double canon_sum(int b, int e, F f, double *red) {
    constexpr int U = kCanonUnroll;
    double acc = 0.0;
    int g = b + static_cast<int>(threadIdx.x);
    for (; g + (U - 1) * kInstBlock < e; g += U * kInstBlock) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = f(g + u * kInstBlock);
#pragma unroll
        for (int u = 0; u < U; ++u) acc += v[u];
    }
    for (; g < e; g += kInstBlock) acc += f(g);
    return block_sum_d<kInstBlock>(acc, red);
}
template <typename F>
__device__ __forceinline__ double canon_max(int b, int e, F f, double *red) {
    constexpr int U = kCanonUnroll;
    double acc = 0.0;
    int g = b + static_cast<int>(threadIdx.x);
    for (; g + (U - 1) * kInstBlock < e; g += U * kInstBlock) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = f(g + u * kInstBlock);
#pragma unroll
        for (int u = 0; u < U; ++u) acc = fmax(acc, v[u]);
    }
    for (; g < e; g += kInstBlock) acc = fmax(acc, f(g));
    return block_max_d<kInstBlock>(acc, red);
}

// canon_sum / canon_max of NQ quantities at once (group partials src[q][g];
// is_max[q] picks the max): lane l runs through its groups b + l,
// b + l + kInstBlock, ... once, carrying every quantity in exactly the order
// canon_sum / canon_max would, and one block pass combines the wave partials
// of all quantities with the same shuffle tree and the same sequential sum
// over the waves.  out[q] equals the separate call's result bit for bit;
// the quantities' loads interleave and the block synchronises twice instead
// of twice per quantity.  `red` holds NQ * kInstBlock / 64 doubles.
template <int NQ>
__device__ __forceinline__ void canon_multi(int b, int e, const double *const (&src)[NQ],
                                            const bool (&is_max)[NQ], double (&out)[NQ], double *red) {
    constexpr int U = kCanonUnroll;  // steps whose loads are issued together (NQ x U in flight)
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    int g = b + static_cast<int>(threadIdx.x);
    for (; g + (U - 1) * kInstBlock < e; g += U * kInstBlock) {
        double v[U][NQ];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQ; ++q) v[u][q] = src[q][g + u * kInstBlock];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = is_max[q] ? fmax(acc[q], v[u][q]) : acc[q] + v[u][q];
    }
    for (; g < e; g += kInstBlock) {
        double v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[q] = src[q][g];
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] = is_max[q] ? fmax(acc[q], v[q]) : acc[q] + v[q];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(acc[q], off, 64);
            acc[q] = is_max[q] ? fmax(acc[q], o) : acc[q] + o;
        }
    }
    constexpr int W = kInstBlock / 64;
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) red[q * W + w] = acc[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double r = red[q * W];
#pragma unroll
        for (int k = 1; k < W; ++k) r = is_max[q] ? fmax(r, red[q * W + k]) : r + red[q * W + k];
        out[q] = r;
    }
}

}  // namespace dev

// Launch-argument prefix of a tile kernel over a tile set (in launcher
// bodies, with the stream in `s`): one workgroup of ts.block threads per tile.
#define MPDHG_TILE_GRID(ts) dim3((ts).grid), dim3((ts).block), 0, s, (ts).tiles, (ts).count, (ts).chunk

// Launches KERNEL<block, tn, gather> for the tile set's schedule and layout
// (every pair of kSchedules).
#define MPDHG_LAUNCH_G(ts, B_, TN_, KERNEL, ...)                                                   \
    do {                                                                                           \
        if ((ts).gather)                                                                           \
            hipLaunchKernelGGL(HIP_KERNEL_NAME(KERNEL<B_, TN_, true>), MPDHG_TILE_GRID(ts), __VA_ARGS__); \
        else                                                                                       \
            hipLaunchKernelGGL(HIP_KERNEL_NAME(KERNEL<B_, TN_, false>), MPDHG_TILE_GRID(ts), __VA_ARGS__); \
    } while (0)
#define MPDHG_LAUNCH_TN(ts, KERNEL, ...)                                                           \
    do {                                                                                           \
        const int b_ = (ts).block, t_ = (ts).tn;                                                   \
        if (b_ == 128 && t_ == 1024) MPDHG_LAUNCH_G(ts, 128, 1024, KERNEL, __VA_ARGS__);           \
        else if (b_ == 128 && t_ == 2048) MPDHG_LAUNCH_G(ts, 128, 2048, KERNEL, __VA_ARGS__);      \
        else if (b_ == 256 && t_ == 1024) MPDHG_LAUNCH_G(ts, 256, 1024, KERNEL, __VA_ARGS__);      \
        else if (b_ == 256 && t_ == 2048) MPDHG_LAUNCH_G(ts, 256, 2048, KERNEL, __VA_ARGS__);      \
        else if (b_ == 256 && t_ == 4096) MPDHG_LAUNCH_G(ts, 256, 4096, KERNEL, __VA_ARGS__);      \
        else if (b_ == 512 && t_ == 4096) MPDHG_LAUNCH_G(ts, 512, 4096, KERNEL, __VA_ARGS__);      \
        else MPDHG_LAUNCH_G(ts, 512, 8192, KERNEL, __VA_ARGS__);                                   \
    } while (0)
static_assert(kNumSchedules == 7, "MPDHG_LAUNCH_TN dispatches every schedule");

// The same launch with HIP events e0 / e1 recorded at the kernel's own start
// and end (hipExtLaunchKernelGGL): the elapsed time between them is the
// kernel's execution alone, the figure a kernel trace reports.
#define MPDHG_LAUNCH_G_EV(ts, e0, e1, B_, TN_, KERNEL, ...)                                        \
    do {                                                                                           \
        if ((ts).gather)                                                                           \
            hipExtLaunchKernelGGL(KERNEL<B_, TN_, true>, dim3((ts).grid), dim3((ts).block), 0, s,  \
                                  e0, e1, 0, (ts).tiles, (ts).count, (ts).chunk, __VA_ARGS__);      \
        else                                                                                       \
            hipExtLaunchKernelGGL(KERNEL<B_, TN_, false>, dim3((ts).grid), dim3((ts).block), 0, s, \
                                  e0, e1, 0, (ts).tiles, (ts).count, (ts).chunk, __VA_ARGS__);      \
    } while (0)
#define MPDHG_LAUNCH_TN_EV(ts, e0, e1, KERNEL, ...)                                                \
    do {                                                                                           \
        const int b_ = (ts).block, t_ = (ts).tn;                                                   \
        if (b_ == 128 && t_ == 1024) MPDHG_LAUNCH_G_EV(ts, e0, e1, 128, 1024, KERNEL, __VA_ARGS__); \
        else if (b_ == 128 && t_ == 2048) MPDHG_LAUNCH_G_EV(ts, e0, e1, 128, 2048, KERNEL, __VA_ARGS__); \
        else if (b_ == 256 && t_ == 1024) MPDHG_LAUNCH_G_EV(ts, e0, e1, 256, 1024, KERNEL, __VA_ARGS__); \
        else if (b_ == 256 && t_ == 2048) MPDHG_LAUNCH_G_EV(ts, e0, e1, 256, 2048, KERNEL, __VA_ARGS__); \
        else if (b_ == 256 && t_ == 4096) MPDHG_LAUNCH_G_EV(ts, e0, e1, 256, 4096, KERNEL, __VA_ARGS__); \
        else if (b_ == 512 && t_ == 4096) MPDHG_LAUNCH_G_EV(ts, e0, e1, 512, 4096, KERNEL, __VA_ARGS__); \
        else MPDHG_LAUNCH_G_EV(ts, e0, e1, 512, 8192, KERNEL, __VA_ARGS__);                          \
    } while (0)

}  // namespace mpdhg
