// Host-side launchers of the gfx950 kernels (defined in the .hip files).
#pragma once

#include <hip/hip_runtime_api.h>

#include "engine_dev.hpp"

namespace mpdhg {

// A device tile table with its padded, XCD-remapped launch grid.
struct TileSet {
    const Tile *tiles = nullptr;
    int count = 0;
    int chunk = 0;  // ceil(count / 8)
    int grid = 0;   // 8 * chunk
};

namespace launch {

// reference mode
void ref_primal(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *c, const float *l,
                const float *u, const float *tau, float *x, float *xbar);
void ref_dual(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
              const int *ridx, const float *rval, const float *xbar, const float *rl,
              const float *ru, const float *sigma, float *y);
void ref_rowres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *rptr,
                const int *ridx, const float *rval, const float *x, const float *rl,
                const float *ru, float *part);
void ref_colres(hipStream_t s, const TileSet &ts, const RefInst *st, const int *cptr,
                const int *cidx, const float *cval, const float *y, const float *x,
                const float *c, const float *l, const float *u, const float *tau, float *part);
void ref_finalize(hipStream_t s, RefInst *st, int ninst, const float *prow, const float *pcol,
                  int iter, int max_iter, float tol_p, float tol_d);
void ref_copy_best(hipStream_t s, const TileSet &ts, const RefInst *st, const float *x,
                   float *xbest);

}  // namespace launch
}  // namespace mpdhg
