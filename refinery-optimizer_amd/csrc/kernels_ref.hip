// gfx950 kernels of the reference-mode PDHG (the algorithm of
// core/src/gpu/cuda_pdhg.cu:14-151 driven by gpu_solver.cpp:227-271),
// restructured for CDNA4: each workgroup owns a tile of rows, streams the
// tile's values and indices from HBM with coalesced loads, gathers the
// vector entries, stages (value, x) pairs in LDS and lets one lane sum one
// row in storage order with fmaf -- the reference's per-row operation
// order, so results are bit-identical to it -- then applies the fused
// update epilogue.  Every launch covers all running instances of a batch.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "kernels_common.hpp"

namespace mpdhg {
namespace dev {

// cuda_pdhg.cu:33-49 (A^T y) fused with :54-77 (projected primal step).
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_ref_primal(
    const Tile *__restrict__ tiles, int ntiles, int chunk, const RefInst *__restrict__ st,
    const int *__restrict__ cptr, const int *__restrict__ cidx, const float *__restrict__ cval,
    const float *__restrict__ y, const float *__restrict__ c, const float *__restrict__ l,
    const float *__restrict__ u, const float *__restrict__ tau, float *__restrict__ x,
    float *__restrict__ xbar) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const float g = tile_dot<B, TN, G, float>(tl, cptr, cidx, cval, y, lds);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) {
        const float xo = x[j];
        const float grad = c[j] + g;
        float xn = __builtin_fmaf(-tau[j], grad, xo);
        const float lo = l[j], hi = u[j];
        if (xn < lo) xn = lo;
        if (xn > hi) xn = hi;
        x[j] = xn;
        xbar[j] = 2.0f * xn - xo;
    }
}

// cuda_pdhg.cu:14-30 (A xbar) fused with :82-103 (Moreau dual step).
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_ref_dual(
    const Tile *__restrict__ tiles, int ntiles, int chunk, const RefInst *__restrict__ st,
    const int *__restrict__ rptr, const int *__restrict__ ridx, const float *__restrict__ rval,
    const float *__restrict__ xbar, const float *__restrict__ rl, const float *__restrict__ ru,
    const float *__restrict__ sigma, float *__restrict__ y) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const float h = tile_dot<B, TN, G, float>(tl, rptr, ridx, rval, xbar, lds);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) {
        const float sg = sigma[i];
        const float w = h + (y[i] / sg);
        float p = w;
        const float lo = rl[i], hi = ru[i];
        if (p < lo) p = lo;
        if (p > hi) p = hi;
        y[i] = sg * (w - p);
    }
}

// cuda_pdhg.cu:14-30 (A xbar) over one band of columns: band pass `first`
// starts each row's sum at 0, later passes continue it from hpart; every pass
// but the `last` stores the running sums in hpart, and the last one applies
// the Moreau dual step (:82-103) of k_ref_dual.  Each band's nonzeros of a
// row are the row's next columns in storage order, so the sums equal
// k_ref_dual's bit for bit.  rptr / ridx / rval: the band's own CSR.
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_ref_dual_band(
    const Tile *__restrict__ tiles, int ntiles, int chunk, const RefInst *__restrict__ st,
    const int *__restrict__ rptr, const int *__restrict__ ridx, const float *__restrict__ rval,
    const float *__restrict__ xbar, float *__restrict__ hpart, int first, int last,
    const float *__restrict__ rl, const float *__restrict__ ru, const float *__restrict__ sigma,
    float *__restrict__ y) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const TileRows tr = tile_rows(tl, rptr);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    const bool mine = i < tl.end;
    const float h0 = (mine && !first) ? hpart[i] : 0.0f;
    const float h = tile_dot_rows<B, TN, G, float, float, float>(tr, ridx, rval, xbar, lds, h0);
    if (!mine) return;
    if (!last) {
        hpart[i] = h;
        return;
    }
    const float sg = sigma[i];
    const float w = h + (y[i] / sg);
    float p = w;
    const float lo = rl[i], hi = ru[i];
    if (p < lo) p = lo;
    if (p > hi) p = hi;
    y[i] = sg * (w - p);
}

// cuda_pdhg.cu:114-129 row residuals of A x, max-reduced per tile (the
// reference copies m floats to the host and reduces there, :249-255).
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_ref_rowres(
    const Tile *__restrict__ tiles, int ntiles, int chunk, const RefInst *__restrict__ st,
    const int *__restrict__ rptr, const int *__restrict__ ridx, const float *__restrict__ rval,
    const float *__restrict__ x, const float *__restrict__ rl, const float *__restrict__ ru,
    float *__restrict__ part) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    // block reductions run after the row sums, in the then idle staging area
    float *const red = reinterpret_cast<float *>(lds);
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const float ax = tile_dot<B, TN, G, float>(tl, rptr, ridx, rval, x, lds);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    float v = 0.0f;
    if (i < tl.end) {
        const float lo = rl[i], hi = ru[i];
        if (lo > -1e15f && ax < lo) v = fmaxf(v, lo - ax);
        if (hi < 1e15f && ax > hi) v = fmaxf(v, ax - hi);
    }
    v = block_max<B>(v, red);
    if (threadIdx.x == 0) part[t] = v;
}

// cuda_pdhg.cu:131-151 column fixed-point residuals from A^T y, max-reduced per tile.
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_ref_colres(
    const Tile *__restrict__ tiles, int ntiles, int chunk, const RefInst *__restrict__ st,
    const int *__restrict__ cptr, const int *__restrict__ cidx, const float *__restrict__ cval,
    const float *__restrict__ y, const float *__restrict__ x, const float *__restrict__ c,
    const float *__restrict__ l, const float *__restrict__ u, const float *__restrict__ tau,
    float *__restrict__ part) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    // block reductions run after the row sums, in the then idle staging area
    float *const red = reinterpret_cast<float *>(lds);
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const float g = tile_dot<B, TN, G, float>(tl, cptr, cidx, cval, y, lds);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    float v = 0.0f;
    if (j < tl.end) {
        const float xj = x[j];
        const float grad = c[j] + g;
        const float step = tau[j];
        float proj = __builtin_fmaf(-step, grad, xj);
        const float lo = l[j], hi = u[j];
        if (proj < lo) proj = lo;
        if (proj > hi) proj = hi;
        v = fabsf(xj - proj) / fmaxf(1e-4f, step);
    }
    v = block_max<B>(v, red);
    if (threadIdx.x == 0) part[t] = v;
}

// gpu_solver.cpp:252-270 per instance: P, D, best-x bookkeeping, stop test.
__global__ __launch_bounds__(kInstBlock) void k_ref_finalize(RefInst *__restrict__ st, int ninst,
                                                             const float *__restrict__ prow,
                                                             const float *__restrict__ pcol,
                                                             const int *__restrict__ itb, int len,
                                                             int max_iter, float tol_p, float tol_d) {
    __shared__ float red[kInstBlock / 64];
    const int q = blockIdx.x;
    if (q >= ninst) return;
    RefInst &s = st[q];
    if (s.status != kRunning) {
        if (threadIdx.x == 0) s.copy_best = 0;
        return;
    }
    float mp = 0.0f, md = 0.0f;
    for (int t = s.row_tile_begin + static_cast<int>(threadIdx.x); t < s.row_tile_end; t += kInstBlock)
        mp = fmaxf(mp, prow[t]);
    for (int t = s.col_tile_begin + static_cast<int>(threadIdx.x); t < s.col_tile_end; t += kInstBlock)
        md = fmaxf(md, pcol[t]);
    mp = block_max<kInstBlock>(mp, red);
    md = block_max<kInstBlock>(md, red);
    const int iter = *itb + len;  // the iteration this check follows
    if (threadIdx.x == 0) {
        const float p = mp / s.max_abs_b;
        const float d = md / s.max_abs_c;
        const float comb = (p < d) ? d : p;
        s.checks += 1;
        s.last_p = p;
        s.last_d = d;
        s.copy_best = 0;
        if (comb < s.best_res) {
            s.best_res = comb;
            s.copy_best = 1;
        }
        if (p <= tol_p && d <= tol_d) {
            s.status = 0;  // Optimal
            s.iterations = iter;
        } else if (iter == max_iter) {
            s.status = 4;  // IterationLimit
            s.iterations = max_iter + 1;
        }
    }
}

// gpu_solver.cpp:261-264 best-x copy, device-side, for flagged instances.
__global__ __launch_bounds__(kBlockMax) void k_ref_copy_best(const Tile *__restrict__ tiles,
                                                          int ntiles, int chunk,
                                                          const RefInst *__restrict__ st,
                                                          const float *__restrict__ x,
                                                          float *__restrict__ xbest) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!st[tl.inst].copy_best) return;
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) xbest[j] = x[j];
}

}  // namespace dev

// ---- launchers (host) ------------------------------------------------------
namespace launch {

void ref_primal(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *c, const float *l,
                const float *u, const float *tau, float *x, float *xbar, hipEvent_t e0, hipEvent_t e1) {
    if (e0)
        MPDHG_LAUNCH_TN_EV(ts, e0, e1, dev::k_ref_primal, st, cptr, cidx, cval, y, c, l, u, tau, x, xbar);
    else
        MPDHG_LAUNCH_TN(ts, dev::k_ref_primal, st, cptr, cidx, cval, y, c, l, u, tau, x, xbar);
}

void ref_dual(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
              const int *ridx, const float *rval, const float *xbar, const float *rl,
              const float *ru, const float *sigma, float *y, hipEvent_t e0, hipEvent_t e1) {
    if (e0)
        MPDHG_LAUNCH_TN_EV(ts, e0, e1, dev::k_ref_dual, st, rptr, ridx, rval, xbar, rl, ru, sigma, y);
    else
        MPDHG_LAUNCH_TN(ts, dev::k_ref_dual, st, rptr, ridx, rval, xbar, rl, ru, sigma, y);
}

void ref_dual_band(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
                   const int *ridx, const float *rval, const float *xbar, float *hpart, int first,
                   int last, const float *rl, const float *ru, const float *sigma, float *y,
                   hipEvent_t e0, hipEvent_t e1) {
    if (e0 || e1)
        MPDHG_LAUNCH_TN_EV(ts, e0, e1, dev::k_ref_dual_band, st, rptr, ridx, rval, xbar, hpart, first,
                           last, rl, ru, sigma, y);
    else
        MPDHG_LAUNCH_TN(ts, dev::k_ref_dual_band, st, rptr, ridx, rval, xbar, hpart, first, last, rl,
                        ru, sigma, y);
}

void ref_rowres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
                const int *ridx, const float *rval, const float *x, const float *rl,
                const float *ru, float *part) {
    MPDHG_LAUNCH_TN(ts, dev::k_ref_rowres, st, rptr, ridx, rval, x, rl, ru, part);
}

void ref_colres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *x,
                const float *c, const float *l, const float *u, const float *tau, float *part) {
    MPDHG_LAUNCH_TN(ts, dev::k_ref_colres, st, cptr, cidx, cval, y, x, c, l, u, tau, part);
}

void ref_finalize(hipStream_t s, RefInst *st, int ninst, const float *prow, const float *pcol,
                  const int *itb, int len, int max_iter, float tol_p, float tol_d) {
    hipLaunchKernelGGL(dev::k_ref_finalize, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, prow,
                       pcol, itb, len, max_iter, tol_p, tol_d);
}

void ref_copy_best(hipStream_t s, const TileSet &ts, const RefInst *st, const float *x,
                   float *xbest) {
    hipLaunchKernelGGL(dev::k_ref_copy_best, dim3(ts.grid), dim3(ts.block), 0, s, ts.tiles,
                       ts.count, ts.chunk, st, x, xbest);
}

void tile_run(hipStream_t s, const TileSet &ts, const RefInst *st) {
    if (ts.count)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_tile_run<RefInst>), dim3((ts.count + 255) / 256),
                           dim3(256), 0, s, ts.tiles, ts.count, st);
}

}  // namespace launch
}  // namespace mpdhg
