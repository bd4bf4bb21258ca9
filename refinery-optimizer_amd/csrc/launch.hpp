// Host-side launchers of the gfx950 kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>

#include "engine_dev.hpp"

namespace mpdhg {

// A device tile table with its padded, XCD-remapped launch grid.
struct TileSet {
    Tile *tiles = nullptr;  // run hints are refreshed in place (tile_run)
    int count = 0;
    int chunk = 0;        // ceil(count / 8)
    int grid = 0;         // 8 * chunk
    int block = 256;      // threads per workgroup (rows per tile at most)
    int tn = 4096;        // nonzeros per LDS chunk; (block, tn) is one of kSchedules
    bool gather = false;  // every tile's chunks in gather order (Tile::base >= 0)
    bool stream_nt = false;  // index/value stream read non-temporally (fused enhanced kernels)
};

namespace launch {

// reference mode
void ref_primal(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *c, const float *l,
                const float *u, const float *tau, float *x, float *xbar, hipEvent_t e0 = nullptr,
                hipEvent_t e1 = nullptr);
void ref_dual(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
              const int *ridx, const float *rval, const float *xbar, const float *rl,
              const float *ru, const float *sigma, float *y, hipEvent_t e0 = nullptr,
              hipEvent_t e1 = nullptr);
// A xbar over one column band (k_ref_dual_band): first / last band pass of
// the product; e0 / e1 (either may be null) at the kernel's own start / end.
void ref_dual_band(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
                   const int *ridx, const float *rval, const float *xbar, float *hpart, int first,
                   int last, const float *rl, const float *ru, const float *sigma, float *y,
                   hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void ref_rowres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
                const int *ridx, const float *rval, const float *x, const float *rl,
                const float *ru, float *part);
void ref_colres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *x,
                const float *c, const float *l, const float *u, const float *tau, float *part);
// The check after iteration *itb + len (itb: iterations before the cycle).
void ref_finalize(hipStream_t s, RefInst *st, int ninst, const float *prow, const float *pcol,
                  const int *itb, int len, int max_iter, float tol_p, float tol_d);
void ref_copy_best(hipStream_t s, const TileSet &ts, const RefInst *st, const float *x,
                   float *xbest);
// Tile::run = (instance status == kRunning) for every tile of the set.
void tile_run(hipStream_t s, const TileSet &ts, const RefInst *st);
void tile_run(hipStream_t s, const TileSet &ts, const EnhInst *st);

// Gather order of a tile set's matrix copy, built on the device (see
// build_dir): dst = src with every chunk sorted by gathered index.  d32 /
// d64 may be null (value copies not wanted).
void gather_order(hipStream_t s, const TileSet &ts, const int *src_idx, int *dst_idx, const float *s32,
                  float *d32, const double *s64, double *d64);

// enhanced mode
struct EnhPrimal {
    const int *cptr, *cidx;
    const float *cval;
    const float *y, *yp;
    const float *c, *l, *u;
    float *x, *x0, *xp, *xbar;
    const double *xbase, *dc;
    double *xchk;
    double *gc;  // column-group partials (kGc* quantities x gcn groups)
    int gcn;
    const int *itb;  // device: first iteration of the current check cycle
    float rho;
    const float *seg = nullptr;    // split columns' block sums (k_split_seg), or null
    const int *rowseg = nullptr;   // per column: offset of its first block sum
};
struct EnhDual {
    const int *rptr, *ridx;
    const float *rval;
    const float *xbar;
    const float *rl, *ru, *ybar;
    float *y, *y0, *yp;
    const double *ybase, *dr;
    double *ychk;
    double *gr;  // row-group partials (kGr* quantities x grn groups)
    int grn;
    const int *itb;
    float rho;
    const float *seg = nullptr;    // split rows' block sums (k_split_seg), or null
    const int *rowseg = nullptr;   // per row: offset of its first block sum
};
struct EnhCheck {
    const int *rptr, *ridx, *cptr, *cidx;
    const double *rval64, *cval64;
    double *xchk, *ychk, *ax, *lam;
    const double *rl64, *ru64, *c64, *l64, *u64, *dc, *dr;
    double *gc;  // column-group partials (kGc* quantities x gcn groups)
    int gcn;
    double *gr;  // row-group partials (kGr* quantities x grn groups)
    int grn;
    int32_t *status_out;
    double refine_ratio;
    int max_refinements;
    double *xbest, *lambest, *ybest, *axbest, *xbase, *ybase;
    float *cs, *ls, *us, *x, *x0, *xp, *rls, *rus, *ybar, *y, *y0, *yp;
    const int *itb;
};

void pow_scale(hipStream_t s, const TileSet &cts, const EnhInst *st, const float *z, float *v);
// pre / rowseg (optional): the split rows' block sums formed from v by
// split_segments (else the tiles form them).
void pow_rows(hipStream_t s, const TileSet &rts, const int *rptr, const int *ridx,
              const float *rval, const float *v, float *w, const float *pre = nullptr,
              const int *rowseg = nullptr);
// z = A^T w, with the column-group partials of |z|^2 in gpow.
void pow_cols(hipStream_t s, const TileSet &cts, const int *cptr, const int *cidx,
              const float *cval, const float *w, float *z, double *gpow, const float *pre = nullptr,
              const int *rowseg = nullptr);
// pow_norm2 = |z|^2 per instance, summed over its column groups in a fixed order.
void pow_fin(hipStream_t s, EnhInst *st, int ninst, const double *gpow);
// Iteration *a.itb + off of the current check cycle.  e0 / e1 (optional):
// events recorded at the kernel's own start and end.
void enh_primal(hipStream_t s, const TileSet &cts, const EnhInst *st, const EnhPrimal &a, int off,
                int check, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void enh_dual(hipStream_t s, const TileSet &rts, const EnhInst *st, const EnhDual &a, int off,
              int check, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// One pass of a column-banded A xbar (k_enh_dual_band): the band's own CSR
// replaces a.rptr / a.ridx / a.rval; the last pass runs the dual update.
void enh_dual_band(hipStream_t s, const TileSet &bts, const EnhInst *st, const EnhDual &a,
                   const int *rptr, const int *ridx, const float *rval, float *hpart, int first,
                   int last, int off, int check, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Block sums of the split rows (k_split_seg) into blk: each block gathers
// from x, or from xalt where its instance restarts at iteration *itb + off
// (the primal's A^T y reads y+ then); blocks of stopped instances are
// skipped.  With st null every block is formed from x (power iteration).
void split_segments(hipStream_t s, const SplitChunk *items, int n, const int *idx, const float *val,
                    const float *x, const float *xalt, const EnhInst *st, const int *itb, int off,
                    float *blk);
// *itb += len (end of a check cycle of len iterations).
void enh_advance(hipStream_t s, int *itb, int len);
// Set-up of the enhanced mode on the device (kernels_scale.hip), for a whole
// batch: the CSR copy of the FP64 values and its index words (ridx,
// rval64[perm[k]] = cval64[k]), then ruiz Ruiz max-norm passes and one
// Pock-Chambolle pass giving D_r (dr) and D_c (dc), then the scaled FP32 values
// (float)((dr_i a) dc_j) in both orders.  rs / cs are scratch of rows / cols
// doubles.  The sums run per row and column in storage order, so D_r and D_c
// equal scale_instance's bit for bit.
void scale_batch(hipStream_t s, int rows, int cols, const int *cptr, const int *cidx, const int *rptr,
                 const int *perm, const double *cval64, double *rval64, int *ridx, int ruiz, double *dr,
                 double *dc, double *rs, double *cs, float *cval32, float *rval32);
// FP64 check, per-instance decisions, best-point copies and (when enabled)
// residual-LP construction, in that stream order.
void enh_check(hipStream_t s, const TileSet &rts, const TileSet &cts, EnhInst *st, int ninst,
               const EnhCheck &a, int off);

// adjoint: CGLS on A_RB^T mu = w_B in the scaled space
struct AdjColArgs {
    const int *cptr, *cidx;
    const double *cval;
    const double *pt;    // gathered: mR * dr * p (or dr * mu for the final pass)
    const double *mb, *dc;
    double *q, *r, *rt;
    double *gq;  // column-group partials of |q|^2
};
struct AdjRowArgs {
    const int *rptr, *ridx;
    const double *rval;
    const double *rt;    // gathered: mB * dc * r
    const double *mr, *dr;
    double *s, *mu, *p, *pt;
    double *gs;  // row-group partials of |s|^2
};
void adj_start(hipStream_t s, const TileSet &rts, AdjInst *st, int ninst, const AdjRowArgs &r,
               double tol, int max_iter);
void adj_step(hipStream_t s, const TileSet &rts, const TileSet &cts, AdjInst *st, int ninst,
              const AdjColArgs &c, const AdjRowArgs &r, double tol, int max_iter);
void adj_atmu(hipStream_t s, const TileSet &rts, const TileSet &cts, const AdjColArgs &c,
              const AdjRowArgs &r);

}  // namespace launch
}  // namespace mpdhg
