// Device helpers shared by the tile kernels (HIP translation units only).
#pragma once

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

// LDS bytes every tile kernel reserves for the staged (value, x) pairs.
constexpr int kStageBytes = kTileNnz * 8;

// Row sums of one tile: sum_k val[k] * xs[idx[k]] for every row of the tile,
// in storage order (k ascending), one row per lane.  The tile's nonzeros
// are streamed in chunks: coalesced loads of idx/val, the gathers issued
// back to back, then (val, x) pairs parked in LDS; each lane then walks its
// own row from LDS.  A row longer than a chunk is continued across chunks
// by the same lane, so the summation order never changes.  `lds` must hold
// kStageBytes.
// The row-pointer reads of a tile: issued first, so that a kernel can put
// its other independent loads (epilogue operands, instance state) in flight
// before it consumes them.
struct TileRows {
    int k0, k1, lo, hi;
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
    t.lo = ptr[tl.begin];
    t.hi = ptr[tl.end];
    return t;
}

template <typename V, typename X, typename A>
__device__ __forceinline__ A tile_dot_rows(const TileRows &tr, const int *__restrict__ idx,
                                           const V *__restrict__ val, const X *__restrict__ xs,
                                           void *lds_raw) {
    constexpr int CH = kStageBytes / static_cast<int>(sizeof(Pair<V, X>));
    constexpr int PT = CH / kBlock;
    Pair<V, X> *lds = static_cast<Pair<V, X> *>(lds_raw);
    const int tid = static_cast<int>(threadIdx.x);
    const int k0 = tr.k0, k1 = tr.k1, lo = tr.lo, hi = tr.hi;
    A sum = A(0);
    for (int c0 = lo; c0 < hi; c0 += CH) {
        const int c1 = min(hi, c0 + CH);
        int id[PT];
        V v[PT];
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int k = c0 + tid + q * kBlock;
            if (k < c1) {
                id[q] = idx[k];
                v[q] = val[k];
            }
        }
        X g[PT];
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int k = c0 + tid + q * kBlock;
            if (k < c1) g[q] = xs[id[q]];
        }
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int k = c0 + tid + q * kBlock;
            if (k < c1) lds[k - c0] = Pair<V, X>{v[q], g[q]};
        }
        __syncthreads();
        const int a = max(k0, c0), b = min(k1, c1);
        for (int k = a; k < b; ++k) {
            const Pair<V, X> p = lds[k - c0];
            sum = fma_acc<A, V>(p.v, p.x, sum);
        }
        __syncthreads();
    }
    return sum;
}

template <typename V, typename X, typename A>
__device__ __forceinline__ A tile_dot_t(const Tile &tl, const int *__restrict__ ptr,
                                        const int *__restrict__ idx, const V *__restrict__ val,
                                        const X *__restrict__ xs, void *lds_raw) {
    return tile_dot_rows<V, X, A>(tile_rows(tl, ptr), idx, val, xs, lds_raw);
}

template <typename A>
__device__ __forceinline__ A tile_dot(const Tile &tl, const int *__restrict__ ptr,
                                      const int *__restrict__ idx, const float *__restrict__ val,
                                      const float *__restrict__ xs, void *lds) {
    return tile_dot_t<float, float, A>(tl, ptr, idx, val, xs, lds);
}

// Block-wide max of non-negative floats (NaN operands are dropped, as
// fmaxf does); every lane receives the result.
__device__ __forceinline__ float block_max(float v, float *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int q = 1; q < kBlock / 64; ++q) r = fmaxf(r, red[q]);
    return r;
}

__device__ __forceinline__ double block_max_d(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int q = 1; q < kBlock / 64; ++q) r = fmax(r, red[q]);
    return r;
}

// Block-wide FP64 sum in a fixed tree order (deterministic for a given
// block shape); every lane receives the result.
__device__ __forceinline__ double block_sum_d(double v, double *red) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = static_cast<int>(threadIdx.x) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int q = 1; q < kBlock / 64; ++q) r += red[q];
    return r;
}

}  // namespace dev
}  // namespace mpdhg
