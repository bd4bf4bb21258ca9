// gfx950 kernels of the adjoint (implicit differentiation of the PDHG fixed
// point on its active set).  At a solution with basic columns B (strictly
// inside their bounds) and active rows R (nonzero duals), the fixed-point
// condition of the projected PDHG step restricted to (B, R) is
//     x_B - tau (c_B + A_RB^T y_R) = x_B  and  y_R + sigma (A_R: x - b_R) = y_R,
// so its transposed linearisation for L = w^T x reduces to A_RB^T mu = w_B
// (dL/db_R = mu, dL/dc = 0, dL/dl_j = dL/du_j = w_j - A_:j^T mu on the
// columns held at a bound).  The system is solved per instance by CGLS in the
// Ruiz/Pock-Chambolle-scaled space (min-norm least squares, so degenerate
// active sets with |R| != |B| are handled), in FP64, with the same tile
// machinery as the forward kernels: one masked SpMV over the CSC tiles and
// one over the CSR tiles per CGLS step, each leaving per-group partials of
// |q|^2 or |s|^2, and per-instance scalar kernels that sum the partials in a
// fixed order.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "kernels_common.hpp"

namespace mpdhg {
namespace dev {

using launch::AdjColArgs;
using launch::AdjRowArgs;

// q_j = mB_j dc_j sum_i A_ij pt_i (pt = masked, D_r-scaled row vector);
// also rt_j for the next row SpMV after the residual update r -= alpha q.
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_adj_cols(const Tile *__restrict__ tiles, int ntiles,
                                                     int chunk, const AdjInst *__restrict__ st,
                                                     AdjColArgs A) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    if (st[tl.inst].done) return;
    const double s = tile_dot_rows<B, TN, G, double, double, double, false, true>(tile_rows(tl, A.cptr), A.cidx, A.cval, A.pt, lds);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    double q2 = 0.0;
    if (j < tl.end) {
        const double q = A.mb[j] * A.dc[j] * s;
        A.q[j] = q;
        q2 = q * q;
    }
    put_group(A.gq, tl, group_sum_d(q2));
}

// alpha = gamma / |q|^2 per instance.
__global__ __launch_bounds__(kInstBlock) void k_adj_alpha(AdjInst *__restrict__ st, int ninst,
                                                          const double *__restrict__ gq) {
    __shared__ double red[kInstBlock / 64];
    const int q = blockIdx.x;
    if (q >= ninst) return;
    AdjInst &s = st[q];
    if (s.done) return;
    const double acc = canon_sum(s.cgrp_begin, s.cgrp_end, [&](int g) { return gq[g]; }, red);
    if (threadIdx.x == 0) s.alpha = acc > 0.0 ? s.gamma / acc : 0.0;
}

// r -= alpha q; rt = mB dc r (column side, elementwise).
__global__ __launch_bounds__(kBlockMax) void k_adj_upd_cols(const Tile *__restrict__ tiles, int ntiles,
                                                         int chunk, const AdjInst *__restrict__ st,
                                                         AdjColArgs A) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const AdjInst &s = st[tl.inst];
    if (s.done) return;
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) {
        const double r = A.r[j] - s.alpha * A.q[j];
        A.r[j] = r;
        A.rt[j] = A.mb[j] * A.dc[j] * r;
    }
}

// s_i = mR_i dr_i sum_j A_ij rt_j ; mu += alpha p ; group partials of |s|^2.
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_adj_rows(const Tile *__restrict__ tiles, int ntiles,
                                                     int chunk, const AdjInst *__restrict__ st,
                                                     AdjRowArgs A) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const AdjInst &s = st[tl.inst];
    if (s.done) return;
    const double v = tile_dot_rows<B, TN, G, double, double, double, false, true>(tile_rows(tl, A.rptr), A.ridx, A.rval, A.rt, lds);
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    double s2 = 0.0;
    if (i < tl.end) {
        const double sv = A.mr[i] * A.dr[i] * v;
        A.s[i] = sv;
        A.mu[i] += s.alpha * A.p[i];
        s2 = sv * sv;
    }
    put_group(A.gs, tl, group_sum_d(s2));
}

// gamma' = |s|^2, beta = gamma'/gamma, stopping test, per instance.
__global__ __launch_bounds__(kInstBlock) void k_adj_beta(AdjInst *__restrict__ st, int ninst,
                                                         const double *__restrict__ gs, double tol,
                                                         int max_iter) {
    __shared__ double red[kInstBlock / 64];
    const int q = blockIdx.x;
    if (q >= ninst) return;
    AdjInst &s = st[q];
    if (s.done) return;
    const double acc = canon_sum(s.rgrp_begin, s.rgrp_end, [&](int g) { return gs[g]; }, red);
    if (threadIdx.x != 0) return;
    if (s.iters == 0) s.gamma0 = acc;
    s.beta = s.gamma > 0.0 ? acc / s.gamma : 0.0;
    s.gamma = acc;
    s.iters += 1;
    if (acc <= tol * tol * s.gamma0 || acc == 0.0 || s.iters > max_iter || !(acc == acc)) s.done = 1;
}

// p = s + beta p ; pt = mR dr p (row side, elementwise).
__global__ __launch_bounds__(kBlockMax) void k_adj_upd_rows(const Tile *__restrict__ tiles, int ntiles,
                                                         int chunk, const AdjInst *__restrict__ st,
                                                         AdjRowArgs A) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const AdjInst &s = st[tl.inst];
    if (s.done) return;
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) {
        const double p = A.s[i] + s.beta * A.p[i];
        A.p[i] = p;
        A.pt[i] = A.mr[i] * A.dr[i] * p;
    }
}

// mu (scaled) -> pt_i = dr_i mu_i, the multiplier of the original rows.
__global__ __launch_bounds__(kBlockMax) void k_adj_unscale(const Tile *__restrict__ tiles, int ntiles,
                                                        int chunk, AdjRowArgs A) {
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const int i = tl.begin + static_cast<int>(threadIdx.x);
    if (i < tl.end) A.pt[i] = A.dr[i] * A.mu[i];
}

// Final column pass in the original space: q_j = sum_i A_ij pt_i = (A^T mu)_j
// with the unscaled multipliers, for the bound gradients.
template <int B, int TN, bool G>
__global__ __launch_bounds__(B) void k_adj_atmu(const Tile *__restrict__ tiles, int ntiles,
                                                     int chunk, AdjColArgs A) {
    __shared__ __align__(16) unsigned char lds[stage_bytes(TN)];
    const int t = xcd_tile(blockIdx.x, chunk);
    if (t >= ntiles) return;
    const Tile tl = tiles[t];
    const double s = tile_dot_rows<B, TN, G, double, double, double, false, true>(tile_rows(tl, A.cptr), A.cidx, A.cval, A.pt, lds);
    const int j = tl.begin + static_cast<int>(threadIdx.x);
    if (j < tl.end) A.q[j] = s;
}

}  // namespace dev

namespace launch {


void adj_step(hipStream_t s, const TileSet &rts, const TileSet &cts, AdjInst *st, int ninst,
              const AdjColArgs &c, const AdjRowArgs &r, double tol, int max_iter) {
    if (cts.count) MPDHG_LAUNCH_TN(cts, dev::k_adj_cols, st, c);
    hipLaunchKernelGGL(dev::k_adj_alpha, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, c.gq);
    if (cts.count) hipLaunchKernelGGL(dev::k_adj_upd_cols, MPDHG_TILE_GRID(cts), st, c);
    if (rts.count) MPDHG_LAUNCH_TN(rts, dev::k_adj_rows, st, r);
    hipLaunchKernelGGL(dev::k_adj_beta, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, r.gs, tol, max_iter);
    if (rts.count) hipLaunchKernelGGL(dev::k_adj_upd_rows, MPDHG_TILE_GRID(rts), st, r);
}

void adj_start(hipStream_t s, const TileSet &rts, AdjInst *st, int ninst, const AdjRowArgs &r,
               double tol, int max_iter) {
    // s = K^T r0 with alpha = 0 (mu unchanged), gamma0 = |s|^2, p = s
    if (rts.count) MPDHG_LAUNCH_TN(rts, dev::k_adj_rows, st, r);
    hipLaunchKernelGGL(dev::k_adj_beta, dim3(ninst), dim3(kInstBlock), 0, s, st, ninst, r.gs, tol, max_iter);
    if (rts.count) hipLaunchKernelGGL(dev::k_adj_upd_rows, MPDHG_TILE_GRID(rts), st, r);
}

void adj_atmu(hipStream_t s, const TileSet &rts, const TileSet &cts, const AdjColArgs &c,
              const AdjRowArgs &r) {
    if (rts.count) hipLaunchKernelGGL(dev::k_adj_unscale, MPDHG_TILE_GRID(rts), r);
    if (cts.count) MPDHG_LAUNCH_TN(cts, dev::k_adj_atmu, c);
}


}  // namespace launch
}  // namespace mpdhg
