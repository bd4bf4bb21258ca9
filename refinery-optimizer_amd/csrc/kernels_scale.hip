// gfx950 kernels of the enhanced mode's set-up: the CSR copy of the FP64
// values, the Ruiz / Pock-Chambolle scaling passes and the scaled FP32 value
// copies, for a whole batch at once (rows, columns and nonzeros concatenated;
// every index below is global).  Each row or column is one lane walking its
// nonzeros in storage order, which is the order the host restatement
// (scale_instance) and the CPU port add them in, so D_r, D_c and the scaled
// values are the same bits.
#include <hip/hip_runtime.h>

#include "engine_dev.hpp"
#include "launch.hpp"

namespace mpdhg {
namespace dev {

constexpr int kScaleBlock = 256;

// CSR copy: the CSC entry k of column j goes to CSR slot perm[k].
__global__ __launch_bounds__(kScaleBlock) void k_scale_expand(int cols, const int *__restrict__ cptr,
                                                              const int *__restrict__ perm,
                                                              const double *__restrict__ cval64,
                                                              double *__restrict__ rval64, int *__restrict__ ridx) {
    const int j = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (j >= cols) return;
    for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
        const int d = perm[k];
        rval64[d] = cval64[k];
        ridx[d] = j;
    }
}

__global__ __launch_bounds__(kScaleBlock) void k_scale_fill(int n, double *__restrict__ v, double x) {
    const int i = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (i < n) v[i] = x;
}

// Column norms of |D_r A D_c|: the maximum (Ruiz) or, with l1, the sum in
// storage order (Pock-Chambolle, alpha = 1); |a| dr_i dc_j is formed as
// (|a| dr_i) dc_j, as on the host.
__global__ __launch_bounds__(kScaleBlock) void k_scale_cols(int cols, const int *__restrict__ cptr,
                                                            const int *__restrict__ cidx,
                                                            const double *__restrict__ cval64,
                                                            const double *__restrict__ dr,
                                                            const double *__restrict__ dc, double *__restrict__ cs,
                                                            int l1) {
    const int j = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (j >= cols) return;
    const double dj = dc[j];
    double acc = 0.0;
    for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
        const double a = fabs(cval64[k]) * dr[cidx[k]] * dj;
        if (l1) acc += a;
        else if (a > acc) acc = a;
    }
    cs[j] = acc;
}

// Row norms, over each row's entries in column order (the order in which the
// host's column sweep reaches them).
__global__ __launch_bounds__(kScaleBlock) void k_scale_rows(int rows, const int *__restrict__ rptr,
                                                            const int *__restrict__ ridx,
                                                            const double *__restrict__ rval64,
                                                            const double *__restrict__ dr,
                                                            const double *__restrict__ dc, double *__restrict__ rs,
                                                            int l1) {
    const int i = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (i >= rows) return;
    const double di = dr[i];
    double acc = 0.0;
    for (int k = rptr[i]; k < rptr[i + 1]; ++k) {
        const double a = fabs(rval64[k]) * di * dc[ridx[k]];
        if (l1) acc += a;
        else if (a > acc) acc = a;
    }
    rs[i] = acc;
}

// d /= sqrt(norm) where the norm is positive (both sides of a pass use the
// scales from before the pass, so this runs after both norm kernels).
__global__ __launch_bounds__(kScaleBlock) void k_scale_apply(int n, double *__restrict__ d,
                                                             const double *__restrict__ norm) {
    const int i = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (i < n && norm[i] > 0.0) d[i] /= sqrt(norm[i]);
}

// The scaled FP32 values (float)((dr_i a) dc_j), in CSC order and, through
// perm, in CSR order.
__global__ __launch_bounds__(kScaleBlock) void k_scale_values(int cols, const int *__restrict__ cptr,
                                                              const int *__restrict__ cidx,
                                                              const int *__restrict__ perm,
                                                              const double *__restrict__ cval64,
                                                              const double *__restrict__ dr,
                                                              const double *__restrict__ dc,
                                                              float *__restrict__ cval32, float *__restrict__ rval32) {
    const int j = static_cast<int>(blockIdx.x) * kScaleBlock + static_cast<int>(threadIdx.x);
    if (j >= cols) return;
    const double dj = dc[j];
    for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
        const float s = static_cast<float>(dr[cidx[k]] * cval64[k] * dj);
        cval32[k] = s;
        rval32[perm[k]] = s;
    }
}

}  // namespace dev

namespace launch {

namespace {
dim3 lanes(int64_t n) { return dim3(static_cast<unsigned>((n + dev::kScaleBlock - 1) / dev::kScaleBlock)); }
}  // namespace

void scale_batch(hipStream_t s, int rows, int cols, const int *cptr, const int *cidx, const int *rptr,
                 const int *perm, const double *cval64, double *rval64, int *ridx, int ruiz, double *dr,
                 double *dc, double *rs, double *cs, float *cval32, float *rval32) {
    if (cols > 0)
        hipLaunchKernelGGL(dev::k_scale_expand, lanes(cols), dim3(dev::kScaleBlock), 0, s, cols, cptr, perm, cval64,
                           rval64, ridx);
    if (rows > 0) hipLaunchKernelGGL(dev::k_scale_fill, lanes(rows), dim3(dev::kScaleBlock), 0, s, rows, dr, 1.0);
    if (cols > 0) hipLaunchKernelGGL(dev::k_scale_fill, lanes(cols), dim3(dev::kScaleBlock), 0, s, cols, dc, 1.0);
    for (int pass = 0; pass <= ruiz; ++pass) {
        const int l1 = pass == ruiz;  // ruiz max-norm passes, then one Pock-Chambolle pass
        if (cols > 0)
            hipLaunchKernelGGL(dev::k_scale_cols, lanes(cols), dim3(dev::kScaleBlock), 0, s, cols, cptr, cidx, cval64,
                               dr, dc, cs, l1);
        if (rows > 0)
            hipLaunchKernelGGL(dev::k_scale_rows, lanes(rows), dim3(dev::kScaleBlock), 0, s, rows, rptr, ridx, rval64,
                               dr, dc, rs, l1);
        if (rows > 0) hipLaunchKernelGGL(dev::k_scale_apply, lanes(rows), dim3(dev::kScaleBlock), 0, s, rows, dr, rs);
        if (cols > 0) hipLaunchKernelGGL(dev::k_scale_apply, lanes(cols), dim3(dev::kScaleBlock), 0, s, cols, dc, cs);
    }
    if (cols > 0)
        hipLaunchKernelGGL(dev::k_scale_values, lanes(cols), dim3(dev::kScaleBlock), 0, s, cols, cptr, cidx, perm,
                           cval64, dr, dc, cval32, rval32);
}

}  // namespace launch
}  // namespace mpdhg
