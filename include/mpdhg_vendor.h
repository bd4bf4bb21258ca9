/*
 * mpdhg_vendor.h -- bench-only vendor baseline (libmpdhg_vendor.so).
 *
 * The reference's bench/cusparse_pdhg_baseline.cu:60-255 runs a fixed number
 * of reference PDHG iterations (diagonal steps 0.95 / max(1, ||a||_1), the
 * updates of cuda_pdhg.cu:54-77 and :82-103) with the vendor sparse library
 * doing both products, and prints microseconds per iteration and the final
 * objective.  This library is the ROCm counterpart with rocSPARSE: a
 * comparison point for the engine's fused kernels, not part of the product
 * (libmpdhg.so does not link it).  Its summation order is rocSPARSE's, so
 * its iterates equal the engine's only up to FP32 rounding.
 */
#ifndef MPDHG_VENDOR_H
#define MPDHG_VENDOR_H

#include "mpdhg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPDHG_VENDOR_TRANSPOSE_OP 0  /* A^T y as the transposed product of the CSR (the reference's way) */
#define MPDHG_VENDOR_EXPLICIT_CSC 1  /* A^T y as a product with an explicit CSR of A^T */

typedef struct mpdhg_vendor_result {
    double us_per_iter;   /* wall time of the iteration loop / iterations (HIP events) */
    double us_spmv_t;     /* A^T y alone, mean over 20 launches */
    double us_spmv;       /* A xbar alone, mean over 20 launches */
    double objective;     /* sum_j c_j x_j of the final x */
} mpdhg_vendor_result;

/* bench/cusparse_pdhg_baseline.cu:60-255 main(): `iterations` reference PDHG
 * iterations from x = clamp(0, l, u), y = 0 on device 0.  x (optional, [n])
 * receives the final primal iterate.  Returns 0 or a negative MPDHG_E_* code
 * (message in mpdhg_vendor_last_error()). */
int mpdhg_vendor_pdhg(const mpdhg_lp *lp, int iterations, int transpose_mode, double *x,
                      mpdhg_vendor_result *out);
const char *mpdhg_vendor_last_error(void);

/* Streaming ceilings of device 0 (no counterpart in the reference: the
 * practical HBM bandwidth the fused kernels' roofline fractions are read
 * against).  Over a buffer of `bytes` (a multiple of 16), float4 kernels with
 * grid-stride loops at 1024 ... 16384 workgroups of 256 threads; each figure
 * is the best grid's bytes moved / time, over `reps` timed launches after one
 * untimed one (HIP events). */
typedef struct mpdhg_vendor_stream_result {
    double read_gbs;     /* reads through the caches (a sum kept in registers) */
    double read_nt_gbs;  /* non-temporal reads */
    double copy_gbs;     /* copy: bytes read + bytes written */
} mpdhg_vendor_stream_result;
int mpdhg_vendor_stream(size_t bytes, int reps, mpdhg_vendor_stream_result *out);

/* Gather cost in the fused tile kernels' memory pattern (no counterpart in
 * the reference): 8 x per_xcd chunks of 8192 (int32 index, FP32 value) words,
 * one 512-thread workgroup per chunk reading 16 of each per lane
 * non-temporally, chunk c gathering x[base_c + floor(k stride)] (sorted, as a
 * gather-order chunk whose entries are `stride` vector entries apart), each
 * XCD over its own contiguous chunks and its own eighth of x.  Mean µs per
 * launch over `reps` launches after one untimed: the stream alone, the stream
 * with the gathers summed in registers, and with the (value, x) pairs staged
 * in LDS and walked 16 per lane. */
typedef struct mpdhg_vendor_gather_result {
    int64_t entries;    /* nonzeros per launch: 8 per_xcd 8192 */
    double stream_us;   /* index and value stream only */
    double gather_us;   /* + x[index] gathers */
    double staged_us;   /* + LDS pairs and the row walk */
    double stream_gbs;  /* 8 entries / stream_us */
} mpdhg_vendor_gather_result;
int mpdhg_vendor_gather(int per_xcd, double stride, int reps, mpdhg_vendor_gather_result *out);

#ifdef __cplusplus
}
#endif

#endif /* MPDHG_VENDOR_H */
