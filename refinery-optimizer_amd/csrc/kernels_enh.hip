// gfx950 kernels of the enhanced mode: scaled, reflected restarted Halpern
// PDHG with FP32 iterates, FP64 checks and on-device restart / refinement
// decisions.  Same tile machinery as the reference mode (one workgroup per
// tile of rows, LDS-staged (value, x) pairs, one lane per row); every launch
// covers all running instances of a batch and each tile reads its
// instance's scalars from the device-resident EnhInst array, so the host
// never has to synchronise inside a check cycle.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "kernels_common.hpp"

namespace mpdhg {
namespace dev {

// Gains of the primal-weight controller (see k_enh_finalize).
constexpr double kPwP = 0.6, kPwI = 0.05;

__device__ __forceinline__ void halpern_weights(int it, int last_restart, float &a, float &b) {
    const int k = it - last_restart;
    a = static_cast<float>(k + 1) / static_cast<float>(k + 2);
    b = 1.0f / static_cast<float>(k + 2);
}

// ---- power iteration on A^T A (step size) ----------------------------------
__global__ __launch_bounds__(kBlockMax) void k_pow_scale(const Tile *__restrict__ tiles, int ntiles,
                                                      int chunk, const EnhInst *__restrict__ st,
                                                      const float *__restrict__ z,
                                                      float *__restrict__ v) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const double n2 = st[tl.inst].pow_norm2;
    const float inv = n2 > 0.0 ? static_cast<float>(1.0 / sqrt(n2)) : 0.0f;
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) v[j] = z[j] * inv;
}

template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_pow_rows(const Tile *__restrict__ tiles, int ntiles,
                                                     int chunk, const int *__restrict__ rptr,
                                                     const int *__restrict__ ridx,
                                                     const float *__restrict__ rval,
                                                     const float *__restrict__ v,
                                                     float *__restrict__ w, const float *__restrict__ pre,
                                                     const int *__restrict__ rowseg) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const float s = tile_dot_rows<B, TN, G, float, float, float, false, true>(tile_rows(tl, rptr), ridx, rval, v, lds,
                                                                              0.0f, pre, rowseg);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) w[i] = s;
}

template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_pow_cols(const Tile *__restrict__ tiles, int ntiles,
                                                     int chunk, const int *__restrict__ cptr,
                                                     const int *__restrict__ cidx,
                                                     const float *__restrict__ cval,
                                                     const float *__restrict__ w,
                                                     float *__restrict__ z,
                                                     double *__restrict__ gpow, const float *__restrict__ pre,
                                                     const int *__restrict__ rowseg) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const float s = tile_dot_rows<B, TN, G, float, float, float, false, true>(tile_rows(tl, cptr), cidx, cval, w, lds,
                                                                              0.0f, pre, rowseg);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    double q = 0.0;
    if (j < tl.end) {
        z[j] = s;
        q = static_cast<double>(s) * static_cast<double>(s);
    }
    put_group(gpow, tl, group_sum_d(q));
}

__global__ __launch_bounds__(kInstBlock) void k_pow_fin(EnhInst *__restrict__ st, int ninst,
                                                        const double *__restrict__ gpow) {
    __shared__ double red[kInstBlock / 64];
    const int q = blockIdx.x;
    if (q >= ninst) return;
    EnhInst &s = st[q];
    const double acc = canon_sum(s.cgrp_begin, s.cgrp_end, [&](int g) { return gpow[g]; }, red);
    if (threadIdx.x == 0) s.pow_norm2 = acc;
}

// The iteration's arithmetic per column and per row.
struct PrimalStep {
    float v;     // x+ = proj(x - tau (c + A^T y))
    float xbar;  // 2 x+ - x
    float x;     // the Halpern update: a ((1 + rho) x+ - rho x) + b x0
};
__device__ __forceinline__ PrimalStep primal_step(float g, float xo, float x0j, float cj, float lo, float hi,
                                                  float tau, float rho, float a, float b) {
    const float rho1 = 1.0f + rho;
    const float tt = cj + g;
    float v = __builtin_fmaf(-tau, tt, xo);
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    const float r = __builtin_fmaf(rho1, v, -(rho * xo));
    return PrimalStep{v, 2.0f * v - xo, __builtin_fmaf(a, r, b * x0j)};
}
struct DualStep {
    float yn;  // y+, the shifted dual prox
    float y;   // the Halpern update
};
// y+ = sigma (w - clamp(w, rl, ru)) - ybar, w = h + (y + ybar) / sigma:
// y + sigma (h - bound) on an active bound, -ybar in between.
__device__ __forceinline__ DualStep dual_step(float h, float yo, float y0i, float yb, float lo, float hi,
                                              float sigma, float inv_sigma, float rho, float a, float b) {
    const float rho1 = 1.0f + rho;
    const float w = __builtin_fmaf(yo + yb, inv_sigma, h);
    float yn = -yb;
    if (w < lo) yn = __builtin_fmaf(sigma, h - lo, yo);
    else if (w > hi) yn = __builtin_fmaf(sigma, h - hi, yo);
    const float r = __builtin_fmaf(rho1, yn, -(rho * yo));
    return DualStep{yn, __builtin_fmaf(a, r, b * y0i)};
}

// ---- K1: A^T y fused with the projected primal step and Halpern update ----
struct PrimalArgs {
    const int *cptr, *cidx;
    const float *cval;
    const float *y, *yp;
    const float *c, *l, *u;
    float *x, *x0, *xp, *xbar;
    const double *xbase, *dc;
    double *xchk;
    double *gc;  // column-group partials, quantity-major (kGc*), gcn groups
    int gcn;
    const int *itb;  // first iteration of the current check cycle (device)
    float rho;
    const float *seg;   // split columns' segment sums, or null
    const int *rowseg;  // per column: offset of its first segment sum
};

// Iteration it = *A.itb + off: the iteration number lives on the device, so
// one recorded check cycle (a HIP graph) replays for every cycle.
template <int B, int TN, bool G, bool NT>
__device__ __forceinline__ void enh_primal_body(const Tile *__restrict__ tiles, int ntiles, int chunk,
                                                const EnhInst *__restrict__ st, const PrimalArgs &A,
                                                int off, int check) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run) return;
    // Put every load that does not depend on the instance state in flight
    // before the state is consumed: row pointers, the nonzero stream and
    // epilogue operands.
    const TileRows tr = tile_rows(tl, A.cptr);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    const bool mine = j < tl.end;
    float xo = 0.0f, x0j = 0.0f, cj = 0.0f, lo = 0.0f, hi = 0.0f;
    if (mine) {
        xo = A.x[j];
        x0j = A.x0[j];
        cj = A.c[j];
        lo = A.l[j];
        hi = A.u[j];
    }
    const EnhInst &s = st[tl.inst];
    const int it = *A.itb + off;
    const bool restart = (s.restart_at == it);
    const float g = tile_dot_rows<B, TN, G, float, float, float, NT, true>(tr, A.cidx, A.cval, restart ? A.yp : A.y,
                                                                           lds, 0.0f, A.seg, A.rowseg);
    if (s.status != kRunning) return;  // a stale run hint: store nothing
    float a, b;
    halpern_weights(it, s.last_restart, a, b);
    double d2 = 0.0, e2 = 0.0;
    if (mine) {
        if (restart) {  // z+ of the last check becomes iterate and anchor
            xo = A.xp[j];
            x0j = xo;
            A.x0[j] = x0j;
        }
        const PrimalStep ps = primal_step(g, xo, x0j, cj, lo, hi, s.tau, A.rho, a, b);
        const float v = ps.v;
        A.xbar[j] = ps.xbar;
        A.x[j] = ps.x;
        if (check) {
            A.xp[j] = v;
            const double d = static_cast<double>(xo) - static_cast<double>(v);
            const double e = static_cast<double>(v) - static_cast<double>(x0j);
            d2 = d * d;
            e2 = e * e;
            A.xchk[j] = A.xbase[j] + A.dc[j] * static_cast<double>(v) * s.inv_ap;
        }
    }
    if (check) {
        put_group(A.gc + kGcDx2 * A.gcn, tl, group_sum_d(d2));
        put_group(A.gc + kGcDxa2 * A.gcn, tl, group_sum_d(e2));
    }
}

// ---- K2: A xbar fused with the shifted dual prox and Halpern update -------
struct DualArgs {
    const int *rptr, *ridx;
    const float *rval;
    const float *xbar;
    const float *rl, *ru, *ybar;
    float *y, *y0, *yp;
    const double *ybase, *dr;
    double *ychk;
    double *gr;  // row-group partials, quantity-major (kGr*), grn groups
    int grn;
    const int *itb;
    float rho;
    const float *seg;   // split rows' segment sums, or null
    const int *rowseg;  // per row: offset of its first segment sum
};

// One pass of a column-banded A xbar: band b holds each row's nonzeros in
// columns [cb_b, cb_{b+1}) as a CSR of its own (A.rptr / A.ridx / A.rval of
// the pass).  The first pass starts each row's sum at zero, later passes
// continue the FMA chain from the running sum in hpart, and every pass but
// the last stores its running sums and ends there, so the last pass sees the
// unbanded product bit for bit.
struct BandPass {
    float *hpart;
    int first, last;
};

template <int B, int TN, bool G, bool NT, bool BAND>
__device__ __forceinline__ void enh_dual_body(const Tile *__restrict__ tiles, int ntiles, int chunk,
                                              const EnhInst *__restrict__ st, const DualArgs &A,
                                              int off, int check, BandPass bp) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run) return;
    const TileRows tr = tile_rows(tl, A.rptr);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    const bool mine = i < tl.end;
    const bool epilogue = !BAND || bp.last;
    float yo = 0.0f, y0i = 0.0f, yb = 0.0f, lo = 0.0f, hi = 0.0f, h0 = 0.0f;
    if (mine && epilogue) {
        yo = A.y[i];
        y0i = A.y0[i];
        yb = A.ybar[i];
        lo = A.rl[i];
        hi = A.ru[i];
    }
    if (BAND && mine && !bp.first) h0 = bp.hpart[i];
    const EnhInst &s = st[tl.inst];
    const int it = *A.itb + off;
    const bool restart = (s.restart_at == it);
    const float h = tile_dot_rows<B, TN, G, float, float, float, NT, true>(tr, A.ridx, A.rval, A.xbar, lds, h0,
                                                                           A.seg, A.rowseg);
    if (BAND && !bp.last) {
        if (mine) bp.hpart[i] = h;
        return;
    }
    if (s.status != kRunning) return;  // a stale run hint: store nothing
    float a, b;
    halpern_weights(it, s.last_restart, a, b);
    double d2 = 0.0, e2 = 0.0;
    if (mine) {
        if (restart) {
            yo = A.yp[i];
            y0i = yo;
            A.y0[i] = y0i;
        }
        const DualStep ds = dual_step(h, yo, y0i, yb, lo, hi, s.sigma, s.inv_sigma, A.rho, a, b);
        const float yn = ds.yn;
        A.y[i] = ds.y;
        if (check) {
            A.yp[i] = yn;
            const double d = static_cast<double>(yo) - static_cast<double>(yn);
            const double e = static_cast<double>(yn) - static_cast<double>(y0i);
            d2 = d * d;
            e2 = e * e;
            A.ychk[i] = A.ybase[i] + A.dr[i] * static_cast<double>(yn) * s.inv_ad;
        }
    }
    if (check) {
        put_group(A.gr + kGrDy2 * A.grn, tl, group_sum_d(d2));
        put_group(A.gr + kGrDya2 * A.grn, tl, group_sum_d(e2));
    }
}

// The fused kernels, with the matrix stream read through the caches (k_enh_*)
// or non-temporally (k_enh_*_nt, for streams larger than the caches).
#define MPDHG_ENH_KERNEL(NAME, BODY, ARGS, NT)                                                   \
    template <int B, int TN, bool G>                                                             \
    __global__ __launch_bounds__(B) void NAME(const Tile *__restrict__ tiles, int ntiles,       \
                                              int chunk, const EnhInst *__restrict__ st, ARGS A, \
                                              int off, int check) {                              \
        BODY<B, TN, G, NT>(tiles, ntiles, chunk, st, A, off, check);                             \
    }
MPDHG_ENH_KERNEL(k_enh_primal, enh_primal_body, PrimalArgs, false)
MPDHG_ENH_KERNEL(k_enh_primal_nt, enh_primal_body, PrimalArgs, true)
#undef MPDHG_ENH_KERNEL
#define MPDHG_ENH_DUAL(NAME, NT, BAND)                                                           \
    template <int B, int TN, bool G>                                                             \
    __global__ __launch_bounds__(B) void NAME(const Tile *__restrict__ tiles, int ntiles,       \
                                              int chunk, const EnhInst *__restrict__ st,         \
                                              DualArgs A, int off, int check, BandPass bp) {     \
        enh_dual_body<B, TN, G, NT, BAND>(tiles, ntiles, chunk, st, A, off, check, bp);          \
    }
MPDHG_ENH_DUAL(k_enh_dual, false, false)
MPDHG_ENH_DUAL(k_enh_dual_nt, true, false)
MPDHG_ENH_DUAL(k_enh_dual_band, false, true)
MPDHG_ENH_DUAL(k_enh_dual_band_nt, true, true)
#undef MPDHG_ENH_DUAL

// ---- split rows: block sums, over many workgroups --------------------------
// One workgroup per SplitChunk (one block of a split row): the block's
// (value, x) pairs are staged in LDS in storage order, lane t sums segment t
// (kSplitSeg nonzeros, a sequential FMA chain from zero), and lane 0 adds the
// segment sums in order: exactly the block sums dense_rows forms.
__global__ __launch_bounds__(kSplitBlock) void k_split_seg(const SplitChunk *__restrict__ items, int n,
                                                           const int *__restrict__ idx,
                                                           const float *__restrict__ val,
                                                           const float *__restrict__ x,
                                                           const float *__restrict__ xalt,
                                                           const EnhInst *__restrict__ st,
                                                           const int *__restrict__ itb, int off,
                                                           float *__restrict__ blk) {
    constexpr int PT = kSplitBlockNnz / kSplitBlock;
    constexpr int NS = kSplitBlockNnz / kSplitSeg;
    static_assert(NS <= kSplitBlock, "one segment per lane");
    __shared__ Pair<float, float> lds[kSplitBlockNnz];
    __shared__ float sg[NS];
    const int i = static_cast<int>(blockIdx.x);
    if (i >= n) return;
    const SplitChunk c = items[i];
    if (st) {
        const EnhInst &s = st[c.inst];
        if (s.status != kRunning) return;
        if (xalt && s.restart_at == *itb + off) x = xalt;
    }
    const int tid = static_cast<int>(threadIdx.x), len = c.c1 - c.c0;
    const uint32_t last = static_cast<uint32_t>(len - 1);
    int id[PT];
    float v[PT];
#pragma unroll
    for (int q = 0; q < PT; ++q) {
        const uint32_t o = min(static_cast<uint32_t>(tid + q * kSplitBlock), last);
        id[q] = idx[c.c0 + static_cast<int>(o)];
        v[q] = val[c.c0 + static_cast<int>(o)];
    }
    float g[PT];
#pragma unroll
    for (int q = 0; q < PT; ++q) g[q] = x[id[q]];
#pragma unroll
    for (int q = 0; q < PT; ++q)
        lds[min(static_cast<uint32_t>(tid + q * kSplitBlock), last)] = Pair<float, float>{v[q], g[q]};
    __syncthreads();
    const int nseg = (len + kSplitSeg - 1) / kSplitSeg;
    if (tid < nseg) {
        const int a = tid * kSplitSeg, b = min(len, a + kSplitSeg);
        float acc = 0.0f;
        for (int k = a; k < b; ++k) acc = __builtin_fmaf(lds[k].v, lds[k].x, acc);
        sg[tid] = acc;
    }
    __syncthreads();
    if (tid == 0) {
        float bs = 0.0f;
        for (int t = 0; t < nseg; ++t) bs = bs + sg[t];
        blk[c.slot] = bs;
    }
}

// ---- FP64 checks on the original (unscaled) LP ----------------------------
__device__ __forceinline__ bool fin(double v) { return isfinite(v); }

struct CheckRowArgs {
    const int *rptr, *ridx;
    const double *rval;
    const double *xchk, *ychk;
    const double *rl, *ru;
    double *ax;
    double *gr;  // row-group partials, quantity-major (kGr*), grn groups
    int grn;
};

template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_enh_chk_rows(const Tile *__restrict__ tiles, int ntiles,
                                                         int chunk, const EnhInst *__restrict__ st,
                                                         CheckRowArgs A) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const double axi = tile_dot_rows<B, TN, G, double, double, double, false, true>(tile_rows(tl, A.rptr), A.ridx,
                                                                                    A.rval, A.xchk, lds);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    double viol = 0.0, sviol = 0.0, dobj = 0.0;
    if (i < tl.end) {
        const double lo = A.rl[i], hi = A.ru[i];
        if (fin(lo) && axi < lo) viol = lo - axi;
        if (fin(hi) && axi > hi) viol = axi - hi;
        A.ax[i] = axi;
        const double yv = A.ychk[i];
        double yi = yv;
        if (!fin(hi) && yi > 0.0) yi = 0.0;
        if (!fin(lo) && yi < 0.0) yi = 0.0;
        sviol = fabs(yv - yi);
        if (yi > 0.0) dobj = -hi * yi;
        else if (yi < 0.0) dobj = -lo * yi;
    }
    put_group(A.gr + kGrRowViol * A.grn, tl, group_max_d(viol));
    put_group(A.gr + kGrSignViol * A.grn, tl, group_max_d(sviol));
    put_group(A.gr + kGrDobj * A.grn, tl, group_sum_d(dobj));
}

struct CheckColArgs {
    const int *cptr, *cidx;
    const double *cval;
    const double *ychk, *xchk;
    const double *c, *l, *u;
    double *lam;
    double *gc;  // column-group partials, quantity-major (kGc*), gcn groups
    int gcn;
};

template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_enh_chk_cols(const Tile *__restrict__ tiles, int ntiles,
                                                         int chunk, const EnhInst *__restrict__ st,
                                                         CheckColArgs A) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!tl.run || st[tl.inst].status != kRunning) return;
    const double aty = tile_dot_rows<B, TN, G, double, double, double, false, true>(tile_rows(tl, A.cptr), A.cidx,
                                                                                    A.cval, A.ychk, lds);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    double dv = 0.0, cv = 0.0, pobj = 0.0, dobj = 0.0;
    if (j < tl.end) {
        const double lo = A.l[j], hi = A.u[j], xj = A.xchk[j], cj = A.c[j];
        const double lam = cj + aty;
        A.lam[j] = lam;
        double p = lam;
        if (!fin(lo) && p > 0.0) p = 0.0;
        if (!fin(hi) && p < 0.0) p = 0.0;
        dv = fabs(lam - p);
        if (p > 0.0) dobj = lo * p;
        else if (p < 0.0) dobj = hi * p;
        if (fin(lo) && xj < lo) cv = lo - xj;
        if (fin(hi) && xj > hi) cv = xj - hi;
        pobj = cj * xj;
    }
    put_group(A.gc + kGcDualViol * A.gcn, tl, group_max_d(dv));
    put_group(A.gc + kGcColViol * A.gcn, tl, group_max_d(cv));
    put_group(A.gc + kGcPobj * A.gcn, tl, group_sum_d(pobj));
    put_group(A.gc + kGcDobj * A.gcn, tl, group_sum_d(dobj));
}

// ---- per-instance decisions -------------------------------------------------
struct FinalizeArgs {
    const double *gc;  // column-group partials (kGc*), gcn groups
    int gcn;
    const double *gr;  // row-group partials (kGr*), grn groups
    int grn;
    int32_t *status_out;
    double refine_ratio;
    int max_refinements;
    const int *itb;
};

__device__ __forceinline__ void set_steps(EnhInst &s) {
    s.tau = static_cast<float>(s.eta / s.omega);
    s.sigma = static_cast<float>(s.eta * s.omega);
    s.inv_sigma = static_cast<float>(1.0 / (s.eta * s.omega));
}

__global__ __launch_bounds__(kInstBlock) void k_enh_finalize(EnhInst *__restrict__ st, int ninst,
                                                             FinalizeArgs A, int off) {
    __shared__ double red[6 * (kInstBlock / 64)];
    const int q = blockIdx.x;
    if (q >= ninst) return;
    EnhInst &s = st[q];
    if (s.status != kRunning) {
        if (threadIdx.x == 0) {
            s.copy_best = 0;
            s.refine = 0;
        }
        return;
    }
    // per instance over its groups, in a fixed order
    const int cb = s.cgrp_begin, ce = s.cgrp_end, rb = s.rgrp_begin, re = s.rgrp_end;
    const double *gc = A.gc, *gr = A.gr;
    const int gcn = A.gcn, grn = A.grn;
    // one fused pass per group kind; each quantity in canon_sum / canon_max's order
    const double *const csrc[6] = {gc + kGcDx2 * gcn, gc + kGcDxa2 * gcn, gc + kGcPobj * gcn,
                                   gc + kGcDobj * gcn, gc + kGcDualViol * gcn, gc + kGcColViol * gcn};
    const bool cmax[6] = {false, false, false, false, true, true};
    double cq[6];
    canon_multi<6>(cb, ce, csrc, cmax, cq, red);
    const double *const rsrc[5] = {gr + kGrDy2 * grn, gr + kGrDya2 * grn, gr + kGrDobj * grn,
                                   gr + kGrRowViol * grn, gr + kGrSignViol * grn};
    const bool rmax[5] = {false, false, false, true, true};
    double rq[5];
    canon_multi<5>(rb, re, rsrc, rmax, rq, red);
    const double dx2 = cq[0], dxa2 = cq[1], pobj = cq[2], dobjc = cq[3], dv = cq[4], cv = cq[5];
    const double dy2 = rq[0], dya2 = rq[1], dobjr = rq[2], rv = rq[3], sv = rq[4];
    if (threadIdx.x != 0) return;

    const int done = *A.itb + off + 1;
    const double dobj = dobjr + dobjc;
    const double ep = fmax(rv, cv) / (1.0 + s.bnorm);
    const double ed = fmax(dv, sv) / (1.0 + s.cnorm);
    const double eg = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    // fmax drops NaN operands; a NaN anywhere must surface as NaN here
    const double kkt = (isnan(ep) || isnan(ed) || isnan(eg)) ? ep + ed + eg : fmax(ep, fmax(ed, eg));
    s.checks += 1;
    s.eps_p = ep;
    s.eps_d = ed;
    s.eps_gap = eg;
    s.kkt = kkt;
    s.pobj = pobj;
    s.dobj = dobj;
    s.copy_best = 0;
    s.refine = 0;
    if (isnan(kkt)) {  // NaN/Inf reached the iterates: stop instead of propagating
        s.status = 6;  // NumericalFailure
        s.iterations = done;
        A.status_out[q] = 6;
        return;
    }
    if (!(kkt >= s.best_kkt)) {
        s.best_kkt = kkt;
        s.best_eps_p = ep;
        s.best_eps_d = ed;
        s.best_eps_gap = eg;
        s.best_pobj = pobj;
        s.best_dobj = dobj;
        s.copy_best = 1;
    }
    if (kkt <= s.tol) {
        s.status = 0;
        s.iterations = done;
        A.status_out[q] = 0;
        return;
    }
    if (done >= s.max_iter) {
        s.status = 4;
        s.iterations = done;
        A.status_out[q] = 4;
        return;
    }
    // the refinement baseline is the first check's KKT; a warm start with its
    // primal weight holds the largest finite double instead, so that its first
    // restart refines
    if (s.round_start_kkt == HUGE_VAL) s.round_start_kkt = kkt;
    const double r = sqrt(s.omega * dx2 + dy2 / s.omega);
    const int kc = done - s.last_restart;
    if (s.r_anchor < 0.0) {
        s.r_anchor = r;
        s.r_prev = r;
        return;
    }
    const bool rs = (r <= 0.2 * s.r_anchor) || (r <= 0.8 * s.r_anchor && r > s.r_prev) ||
                    (static_cast<double>(kc) >= 0.36 * static_cast<double>(done));
    s.r_prev = r;
    if (!rs) return;
    // restart to z+ = this check's point.  Primal weight: a PI controller on
    // the log-balance e = log(omega |dx| / |dy|) of the primal and dual moves
    // since the last restart (Lu & Yang, cuPDLPx, 2025): log omega -= kP e +
    // kI sum(e).  (The PDLP smoothing omega <- sqrt(omega |dy| / |dx|) is the
    // P term alone with kP = 1/2.)
    {
        const double ddx = sqrt(dxa2), ddy = sqrt(dya2);
        if (ddx > 1e-10 && ddy > 1e-10) {
            const double e = log(s.omega * ddx / ddy);
            s.pw_int += e;
            s.omega = exp(log(s.omega) - (kPwP * e + kPwI * s.pw_int));
        }
    }
    s.last_restart = done;
    s.restarts += 1;
    s.r_anchor = r;
    if (s.rounds < A.max_refinements && kkt <= A.refine_ratio * s.round_start_kkt) {
        // Refinement at the restart: the same point, re-centred in FP64 as the
        // residual LP (x = x~ + D_c d / alpha_p, y = y~ + D_r e / alpha_d); the
        // refine kernels build it from this check's buffers.  The primal weight
        // is rescaled so that the effective weight is unchanged.
        const double dp = fmax(ep, eg), dd = fmax(ed, eg);
        double ap = fmin(1.0 / fmax(dp, 1e-300), 1e12), ad = fmin(1.0 / fmax(dd, 1e-300), 1e12);
        if (ap < 1.0) ap = 1.0;
        if (ad < 1.0) ad = 1.0;
        s.omega *= (ad / ap) / (s.alpha_d / s.alpha_p);
        s.alpha_p = ap;
        s.alpha_d = ad;
        s.inv_ap = 1.0 / ap;
        s.inv_ad = 1.0 / ad;
        s.rounds += 1;
        s.round_start_kkt = kkt;
        s.refine = 1;
    } else {
        s.restart_at = done;
    }
    set_steps(s);
}

// ---- best-point copies and residual-LP construction -----------------------
__global__ __launch_bounds__(kBlockMax) void k_enh_best_cols(const Tile *__restrict__ tiles, int ntiles,
                                                          int chunk, const EnhInst *__restrict__ st,
                                                          const double *__restrict__ xchk,
                                                          const double *__restrict__ lam,
                                                          double *__restrict__ xbest,
                                                          double *__restrict__ lambest) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!st[tl.inst].copy_best) return;
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) {
        xbest[j] = xchk[j];
        lambest[j] = lam[j];
    }
}

__global__ __launch_bounds__(kBlockMax) void k_enh_best_rows(const Tile *__restrict__ tiles, int ntiles,
                                                          int chunk, const EnhInst *__restrict__ st,
                                                          const double *__restrict__ ychk,
                                                          const double *__restrict__ ax,
                                                          double *__restrict__ ybest,
                                                          double *__restrict__ axbest) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (!st[tl.inst].copy_best) return;
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) {
        ybest[i] = ychk[i];
        axbest[i] = ax[i];
    }
}

struct RefineColArgs {
    const double *xchk, *lam, *dc, *l64, *u64;
    double *xbase;
    float *c, *l, *u, *x, *x0, *xp;
};

// Residual LP around this check's point (x~, y~): costs alpha_d * lambda~,
// bounds alpha_p * ([l, u] - x~), row bounds alpha_p * ([rl, ru] - A x~), dual
// offset alpha_d * y~, all in the scaled space; the correction restarts at 0.
__global__ __launch_bounds__(kBlockMax) void k_enh_refine_cols(const Tile *__restrict__ tiles,
                                                            int ntiles, int chunk,
                                                            const EnhInst *__restrict__ st,
                                                            RefineColArgs A) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const EnhInst &s = st[tl.inst];
    if (!s.refine) return;
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) {
        const double xb = A.xchk[j], d = A.dc[j];
        A.xbase[j] = xb;
        A.c[j] = static_cast<float>(d * (s.alpha_d * A.lam[j]));
        const float lo = static_cast<float>(s.alpha_p * (A.l64[j] - xb) / d);
        const float hi = static_cast<float>(s.alpha_p * (A.u64[j] - xb) / d);
        A.l[j] = lo;
        A.u[j] = hi;
        float v = 0.0f;
        if (v < lo) v = lo;
        if (v > hi) v = hi;
        A.x[j] = v;
        A.x0[j] = v;
        A.xp[j] = v;
    }
}

struct RefineRowArgs {
    const double *ychk, *ax, *dr, *rl64, *ru64;
    double *ybase;
    float *rl, *ru, *ybar, *y, *y0, *yp;
};

__global__ __launch_bounds__(kBlockMax) void k_enh_refine_rows(const Tile *__restrict__ tiles,
                                                            int ntiles, int chunk,
                                                            const EnhInst *__restrict__ st,
                                                            RefineRowArgs A) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const EnhInst &s = st[tl.inst];
    if (!s.refine) return;
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) {
        const double yb = A.ychk[i], ab = A.ax[i], d = A.dr[i];
        A.ybase[i] = yb;
        A.rl[i] = static_cast<float>(s.alpha_p * (A.rl64[i] - ab) * d);
        A.ru[i] = static_cast<float>(s.alpha_p * (A.ru64[i] - ab) * d);
        A.ybar[i] = static_cast<float>(s.alpha_d * yb / d);
        A.y[i] = 0.0f;
        A.y0[i] = 0.0f;
        A.yp[i] = 0.0f;
    }
}

// End of a check cycle of `len` iterations: the next cycle starts at *itb.
__global__ void k_enh_advance(int *__restrict__ itb, int len) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *itb += len;
}

}  // namespace dev

// ---- launchers --------------------------------------------------------------
namespace launch {

void pow_scale(hipStream_t s, const TileSet &cts, const EnhInst *st, const float *z, float *v) {
    if (cts.count) hipLaunchKernelGGL(dev::k_pow_scale, MPDHG_TILE_GRID(cts), st, z, v);
}
void pow_rows(hipStream_t s, const TileSet &rts, const int *rptr, const int *ridx,
              const float *rval, const float *v, float *w, const float *pre, const int *rowseg) {
    if (rts.count) MPDHG_LAUNCH_TN(rts, dev::k_pow_rows, rptr, ridx, rval, v, w, pre, rowseg);
}
void pow_cols(hipStream_t s, const TileSet &cts, const int *cptr, const int *cidx,
              const float *cval, const float *w, float *z, double *gpow, const float *pre,
              const int *rowseg) {
    if (cts.count) MPDHG_LAUNCH_TN(cts, dev::k_pow_cols, cptr, cidx, cval, w, z, gpow, pre, rowseg);
}
void pow_fin(hipStream_t s, EnhInst *st, int ninst, const double *gpow) {
    hipLaunchKernelGGL(dev::k_pow_fin, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, gpow);
}
void enh_primal(hipStream_t s, const TileSet &cts, const EnhInst *st, const EnhPrimal &a, int off,
                int check, hipEvent_t e0, hipEvent_t e1) {
    dev::PrimalArgs A{a.cptr, a.cidx, a.cval, a.y, a.yp, a.c, a.l, a.u, a.x, a.x0, a.xp,
                      a.xbar, a.xbase, a.dc, a.xchk, a.gc, a.gcn, a.itb, a.rho, a.seg, a.rowseg};
    if (!cts.count) return;
    if (e0 && cts.stream_nt)
        MPDHG_LAUNCH_TN_EV(cts, e0, e1, dev::k_enh_primal_nt, st, A, off, check);
    else if (e0)
        MPDHG_LAUNCH_TN_EV(cts, e0, e1, dev::k_enh_primal, st, A, off, check);
    else if (cts.stream_nt)
        MPDHG_LAUNCH_TN(cts, dev::k_enh_primal_nt, st, A, off, check);
    else
        MPDHG_LAUNCH_TN(cts, dev::k_enh_primal, st, A, off, check);
}
void enh_dual(hipStream_t s, const TileSet &rts, const EnhInst *st, const EnhDual &a, int off,
              int check, hipEvent_t e0, hipEvent_t e1) {
    dev::DualArgs A{a.rptr, a.ridx, a.rval, a.xbar, a.rl, a.ru, a.ybar, a.y, a.y0, a.yp,
                    a.ybase, a.dr, a.ychk, a.gr, a.grn, a.itb, a.rho, a.seg, a.rowseg};
    const dev::BandPass bp{nullptr, 1, 1};
    if (!rts.count) return;
    if (e0 && rts.stream_nt)
        MPDHG_LAUNCH_TN_EV(rts, e0, e1, dev::k_enh_dual_nt, st, A, off, check, bp);
    else if (e0)
        MPDHG_LAUNCH_TN_EV(rts, e0, e1, dev::k_enh_dual, st, A, off, check, bp);
    else if (rts.stream_nt)
        MPDHG_LAUNCH_TN(rts, dev::k_enh_dual_nt, st, A, off, check, bp);
    else
        MPDHG_LAUNCH_TN(rts, dev::k_enh_dual, st, A, off, check, bp);
}
void enh_dual_band(hipStream_t s, const TileSet &bts, const EnhInst *st, const EnhDual &a,
                   const int *rptr, const int *ridx, const float *rval, float *hpart, int first,
                   int last, int off, int check, hipEvent_t e0, hipEvent_t e1) {
    dev::DualArgs A{rptr, ridx, rval, a.xbar, a.rl, a.ru, a.ybar, a.y, a.y0, a.yp,
                    a.ybase, a.dr, a.ychk, a.gr, a.grn, a.itb, a.rho, nullptr, nullptr};
    const dev::BandPass bp{hpart, first, last};
    if (!bts.count) return;
    if ((e0 || e1) && bts.stream_nt)
        MPDHG_LAUNCH_TN_EV(bts, e0, e1, dev::k_enh_dual_band_nt, st, A, off, check, bp);
    else if (e0 || e1)
        MPDHG_LAUNCH_TN_EV(bts, e0, e1, dev::k_enh_dual_band, st, A, off, check, bp);
    else if (bts.stream_nt)
        MPDHG_LAUNCH_TN(bts, dev::k_enh_dual_band_nt, st, A, off, check, bp);
    else
        MPDHG_LAUNCH_TN(bts, dev::k_enh_dual_band, st, A, off, check, bp);
}
void split_segments(hipStream_t s, const SplitChunk *items, int n, const int *idx, const float *val,
                    const float *x, const float *xalt, const EnhInst *st, const int *itb, int off,
                    float *blk) {
    if (n <= 0) return;
    hipLaunchKernelGGL(dev::k_split_seg, dim3(static_cast<unsigned>(n)), dim3(kSplitBlock), 0, s, items, n, idx,
                       val, x, xalt, st, itb, off, blk);
}
void enh_advance(hipStream_t s, int *itb, int len) {
    hipLaunchKernelGGL(dev::k_enh_advance, dim3(1), dim3(64), 0, s, itb, len);
}
void enh_check(hipStream_t s, const TileSet &rts, const TileSet &cts, EnhInst *st, int ninst,
               const EnhCheck &a, int off) {
    dev::CheckRowArgs R{a.rptr, a.ridx, a.rval64, a.xchk, a.ychk, a.rl64, a.ru64, a.ax, a.gr,
                        a.grn};
    dev::CheckColArgs Cc{a.cptr, a.cidx, a.cval64, a.ychk, a.xchk, a.c64, a.l64, a.u64, a.lam,
                         a.gc, a.gcn};
    if (rts.count) MPDHG_LAUNCH_TN(rts, dev::k_enh_chk_rows, st, R);
    if (cts.count) MPDHG_LAUNCH_TN(cts, dev::k_enh_chk_cols, st, Cc);
    dev::FinalizeArgs F{a.gc, a.gcn, a.gr, a.grn, a.status_out, a.refine_ratio, a.max_refinements,
                        a.itb};
    hipLaunchKernelGGL(dev::k_enh_finalize, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, F, off);
    if (cts.count)
        hipLaunchKernelGGL(dev::k_enh_best_cols, MPDHG_TILE_GRID(cts), st, a.xchk, a.lam, a.xbest,
                           a.lambest);
    if (rts.count)
        hipLaunchKernelGGL(dev::k_enh_best_rows, MPDHG_TILE_GRID(rts), st, a.ychk, a.ax, a.ybest,
                           a.axbest);
    if (a.max_refinements > 0) {
        dev::RefineColArgs RC{a.xchk, a.lam, a.dc, a.l64, a.u64, a.xbase,
                              a.cs, a.ls, a.us, a.x, a.x0, a.xp};
        dev::RefineRowArgs RR{a.ychk, a.ax, a.dr, a.rl64, a.ru64, a.ybase,
                              a.rls, a.rus, a.ybar, a.y, a.y0, a.yp};
        if (cts.count) hipLaunchKernelGGL(dev::k_enh_refine_cols, MPDHG_TILE_GRID(cts), st, RC);
        if (rts.count) hipLaunchKernelGGL(dev::k_enh_refine_rows, MPDHG_TILE_GRID(rts), st, RR);
    }
    tile_run(s, rts, st);
    tile_run(s, cts, st);
}

void tile_run(hipStream_t s, const TileSet &ts, const EnhInst *st) {
    if (ts.count)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(dev::k_tile_run<EnhInst>), dim3((ts.count + 255) / 256),
                           dim3(256), 0, s, ts.tiles, ts.count, st);
}


}  // namespace launch
}  // namespace mpdhg
