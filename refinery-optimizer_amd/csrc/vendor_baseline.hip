// Bench-only vendor baseline (include/mpdhg_vendor.h): the reference PDHG
// iteration with rocSPARSE doing A^T y and A xbar, as
// bench/cusparse_pdhg_baseline.cu:60-255 does with cuSPARSE.  Host
// preparation reuses the engine's validated int32 structure.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "engine.hpp"
#include "mpdhg_vendor.h"

namespace {

thread_local std::string g_err;

struct Fail {
    int code;
    std::string msg;
};

void hip_ok(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Fail{e == hipErrorOutOfMemory ? MPDHG_E_NOMEM : MPDHG_E_HIP,
                   std::string(what) + ": " + hipGetErrorString(e)};
}

void sp_ok(rocsparse_status s, const char *what) {
    if (s != rocsparse_status_success)
        throw Fail{MPDHG_E_HIP, std::string(what) + ": rocsparse status " + std::to_string(static_cast<int>(s))};
}

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) { hip_ok(hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc"); }
    Dev(const std::vector<T> &h, hipStream_t s) : Dev(h.size()) {
        if (!h.empty()) hip_ok(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
    }
    ~Dev() { if (p) (void)hipFree(p); }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

// cuda_pdhg.cu:54-77 (with the contraction nvcc applies to x - tau * grad).
__global__ __launch_bounds__(256) void k_primal(int n, const float *__restrict__ tau,
                                                const float *__restrict__ c,
                                                const float *__restrict__ aty,
                                                const float *__restrict__ l,
                                                const float *__restrict__ u, float *__restrict__ x,
                                                float *__restrict__ xbar) {
    const int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= n) return;
    const float xo = x[j];
    float xn = __builtin_fmaf(-tau[j], c[j] + aty[j], xo);
    const float lo = l[j], hi = u[j];
    if (xn < lo) xn = lo;
    if (xn > hi) xn = hi;
    x[j] = xn;
    xbar[j] = 2.0f * xn - xo;
}

// cuda_pdhg.cu:82-103.
__global__ __launch_bounds__(256) void k_dual(int m, const float *__restrict__ sigma,
                                              const float *__restrict__ ax,
                                              const float *__restrict__ rl,
                                              const float *__restrict__ ru, float *__restrict__ y) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= m) return;
    const float sg = sigma[i];
    const float w = ax[i] + (y[i] / sg);
    float p = w;
    const float lo = rl[i], hi = ru[i];
    if (p < lo) p = lo;
    if (p > hi) p = hi;
    y[i] = sg * (w - p);
}

// One rocSPARSE product y = op(A) x with its buffer, prepared once.
struct SpMV {
    rocsparse_handle h;
    rocsparse_spmat_descr mat;
    rocsparse_dnvec_descr in, out;
    rocsparse_operation op;
    void *buf = nullptr;
    size_t bytes = 0;
    const float one = 1.0f, zero = 0.0f;

    SpMV(rocsparse_handle h_, rocsparse_spmat_descr mat_, rocsparse_operation op_, int64_t nin,
         float *xin, int64_t nout, float *yout)
        : h(h_), mat(mat_), op(op_) {
        sp_ok(rocsparse_create_dnvec_descr(&in, nin, xin, rocsparse_datatype_f32_r), "dnvec");
        sp_ok(rocsparse_create_dnvec_descr(&out, nout, yout, rocsparse_datatype_f32_r), "dnvec");
        run(rocsparse_spmv_stage_buffer_size);
        hip_ok(hipMalloc(&buf, std::max<size_t>(bytes, 4)), "hipMalloc spmv buffer");
        run(rocsparse_spmv_stage_preprocess);
    }
    void run(rocsparse_spmv_stage stage) {
        sp_ok(rocsparse_spmv(h, op, &one, mat, in, &zero, out, rocsparse_datatype_f32_r,
                             rocsparse_spmv_alg_default, stage, &bytes, buf),
              "rocsparse_spmv");
    }
    void compute() { run(rocsparse_spmv_stage_compute); }
    SpMV(const SpMV &) = delete;
    SpMV &operator=(const SpMV &) = delete;
    ~SpMV() {
        (void)rocsparse_destroy_dnvec_descr(in);
        (void)rocsparse_destroy_dnvec_descr(out);
        if (buf) (void)hipFree(buf);
    }
};

int run_baseline(const mpdhg_lp *lp, int iters, int mode, double *xout, mpdhg_vendor_result *res) {
    if (!lp || !res || iters < 0 || (mode != MPDHG_VENDOR_TRANSPOSE_OP && mode != MPDHG_VENDOR_EXPLICIT_CSC))
        throw Fail{MPDHG_E_ARG, "bad arguments"};
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Fail{MPDHG_E_NODEV, "no HIP device visible"};
    mpdhg::HostLp h;
    mpdhg::prepare_structure(*lp, h);
    const int m = h.m, n = h.n, nnz = h.nnz;
    // host preparation as bench/cusparse_pdhg_baseline.cu:80-124
    std::vector<float> cv(static_cast<size_t>(nnz)), rv(static_cast<size_t>(nnz));
    for (int k = 0; k < nnz; ++k) {
        cv[static_cast<size_t>(k)] = static_cast<float>(lp->vals[k]);
        rv[static_cast<size_t>(h.csc2csr[static_cast<size_t>(k)])] = cv[static_cast<size_t>(k)];
    }
    std::vector<float> c(static_cast<size_t>(n)), l(static_cast<size_t>(n)), u(static_cast<size_t>(n)),
        tau(static_cast<size_t>(n)), x0(static_cast<size_t>(n));
    std::vector<float> rl(static_cast<size_t>(m)), ru(static_cast<size_t>(m)), sigma(static_cast<size_t>(m));
    for (int j = 0; j < n; ++j) {
        c[static_cast<size_t>(j)] = static_cast<float>(lp->c[j]);
        l[static_cast<size_t>(j)] = static_cast<float>(lp->col_lower[j]);
        u[static_cast<size_t>(j)] = static_cast<float>(lp->col_upper[j]);
        float s = 0.0f;
        for (int k = h.cptr[static_cast<size_t>(j)]; k < h.cptr[static_cast<size_t>(j) + 1]; ++k) s += std::abs(cv[static_cast<size_t>(k)]);
        tau[static_cast<size_t>(j)] = 0.95f / std::max(1.0f, s);
        x0[static_cast<size_t>(j)] = std::max(l[static_cast<size_t>(j)], std::min(u[static_cast<size_t>(j)], 0.0f));
    }
    for (int i = 0; i < m; ++i) {
        rl[static_cast<size_t>(i)] = static_cast<float>(lp->row_lower[i]);
        ru[static_cast<size_t>(i)] = static_cast<float>(lp->row_upper[i]);
        float s = 0.0f;
        for (int k = h.rptr[static_cast<size_t>(i)]; k < h.rptr[static_cast<size_t>(i) + 1]; ++k) s += std::abs(rv[static_cast<size_t>(k)]);
        sigma[static_cast<size_t>(i)] = 0.95f / std::max(1.0f, s);
    }
    hip_ok(hipSetDevice(0), "hipSetDevice");
    hipStream_t st;
    hip_ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "stream");
    struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{st};
    Dev<int> drp(h.rptr, st), dci(h.ridx, st), dcp(h.cptr, st), dri(h.cidx, st);
    Dev<float> drv(rv, st), dcv(cv, st), dc(c, st), dl(l, st), du(u, st), dtau(tau, st), drl(rl, st),
        dru(ru, st), dsig(sigma, st), dx(x0, st), dxbar(x0, st);
    Dev<float> dy(static_cast<size_t>(m)), daty(static_cast<size_t>(n)), dax(static_cast<size_t>(m));
    hip_ok(hipMemsetAsync(dy.p, 0, static_cast<size_t>(m) * sizeof(float), st), "memset");

    rocsparse_handle hd;
    {
        const hipError_t pending = hipGetLastError();
        const rocsparse_status hs = rocsparse_create_handle(&hd);
        if (hs != rocsparse_status_success)
            throw Fail{MPDHG_E_HIP, std::string("rocsparse_create_handle: status ") +
                                        std::to_string(static_cast<int>(hs)) + " (HIP error pending before: " +
                                        hipGetErrorName(pending) + ", after: " + hipGetErrorName(hipGetLastError()) + ")"};
    }
    struct HandleGuard { rocsparse_handle h; ~HandleGuard() { (void)rocsparse_destroy_handle(h); } } hg{hd};
    sp_ok(rocsparse_set_stream(hd, st), "set_stream");
    rocsparse_spmat_descr A, At = nullptr;
    sp_ok(rocsparse_create_csr_descr(&A, m, n, nnz, drp.p, dci.p, drv.p, rocsparse_indextype_i32,
                                     rocsparse_indextype_i32, rocsparse_index_base_zero,
                                     rocsparse_datatype_f32_r), "csr descr");
    struct MatGuard { rocsparse_spmat_descr a, b; ~MatGuard() { (void)rocsparse_destroy_spmat_descr(a); if (b) (void)rocsparse_destroy_spmat_descr(b); } } mg{A, nullptr};
    if (mode == MPDHG_VENDOR_EXPLICIT_CSC) {
        sp_ok(rocsparse_create_csr_descr(&At, n, m, nnz, dcp.p, dri.p, dcv.p, rocsparse_indextype_i32,
                                         rocsparse_indextype_i32, rocsparse_index_base_zero,
                                         rocsparse_datatype_f32_r), "csr descr (A^T)");
        mg.b = At;
    }
    SpMV fwd(hd, A, rocsparse_operation_none, n, dxbar.p, m, dax.p);
    SpMV bwd = (mode == MPDHG_VENDOR_EXPLICIT_CSC)
                   ? SpMV(hd, At, rocsparse_operation_none, m, dy.p, n, daty.p)
                   : SpMV(hd, A, rocsparse_operation_transpose, m, dy.p, n, daty.p);
    const dim3 gn((n + 255) / 256), gm((m + 255) / 256), blk(256);
    hipEvent_t e0, e1;
    hip_ok(hipEventCreate(&e0), "event");
    hip_ok(hipEventCreate(&e1), "event");
    struct EvGuard { hipEvent_t a, b; ~EvGuard() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    hip_ok(hipStreamSynchronize(st), "sync");
    hip_ok(hipEventRecord(e0, st), "record");
    for (int it = 0; it < iters; ++it) {  // bench/cusparse_pdhg_baseline.cu:214-228
        bwd.compute();
        if (n) hipLaunchKernelGGL(k_primal, gn, blk, 0, st, n, dtau.p, dc.p, daty.p, dl.p, du.p, dx.p, dxbar.p);
        fwd.compute();
        if (m) hipLaunchKernelGGL(k_dual, gm, blk, 0, st, m, dsig.p, dax.p, drl.p, dru.p, dy.p);
    }
    hip_ok(hipGetLastError(), "launch");
    hip_ok(hipEventRecord(e1, st), "record");
    hip_ok(hipEventSynchronize(e1), "sync");
    float ms = 0.0f;
    hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    res->us_per_iter = iters ? 1e3 * ms / iters : 0.0;
    std::vector<float> xh(static_cast<size_t>(n));
    if (n) hip_ok(hipMemcpyAsync(xh.data(), dx.p, static_cast<size_t>(n) * sizeof(float), hipMemcpyDeviceToHost, st), "download");
    hip_ok(hipStreamSynchronize(st), "sync");
    double obj = 0.0;
    for (int j = 0; j < n; ++j) {
        obj += lp->c[j] * static_cast<double>(xh[static_cast<size_t>(j)]);
        if (xout) xout[j] = xh[static_cast<size_t>(j)];
    }
    res->objective = obj;
    // the two products alone (state is not used afterwards)
    auto time_op = [&](SpMV &op) {
        hip_ok(hipEventRecord(e0, st), "record");
        for (int r = 0; r < 20; ++r) op.compute();
        hip_ok(hipEventRecord(e1, st), "record");
        hip_ok(hipEventSynchronize(e1), "sync");
        float t = 0.0f;
        hip_ok(hipEventElapsedTime(&t, e0, e1), "elapsed");
        return 1e3 * t / 20.0;
    };
    res->us_spmv_t = time_op(bwd);
    res->us_spmv = time_op(fwd);
    return MPDHG_OK;
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Reads a[0, n) with a grid-stride loop; the sum is stored only when it
// takes a value it never takes (the buffer is zero), so the loads stay.
template <bool NT>
__global__ __launch_bounds__(256) void k_stream_read(const f4 *__restrict__ a, size_t n, float *__restrict__ out) {
    float s = 0.0f;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256ull) {
        f4 v;
        if constexpr (NT) v = __builtin_nontemporal_load(a + i);
        else v = a[i];
        s += v.x + v.y + v.z + v.w;
    }
    if (s == 1.0f) out[0] = s;
}

__global__ __launch_bounds__(256) void k_stream_copy(const f4 *__restrict__ a, f4 *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256ull)
        b[i] = a[i];
}

int run_stream(size_t bytes, int reps, mpdhg_vendor_stream_result *out) {
    if (!out || bytes < 16 || bytes % 16 != 0 || reps < 1) throw Fail{MPDHG_E_ARG, "bytes must be a positive multiple of 16, reps >= 1"};
    hip_ok(hipSetDevice(0), "hipSetDevice");
    const size_t n = bytes / 16;
    Dev<f4> a(n), b(n);
    Dev<float> o(1);
    hip_ok(hipMemset(a.p, 0, bytes), "hipMemset");
    hip_ok(hipMemset(b.p, 0, bytes), "hipMemset");
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (auto &v : ev.e) hip_ok(hipEventCreate(&v), "hipEventCreate");
    double best[3] = {0.0, 0.0, 0.0};
    for (int grid : {1024, 2048, 4096, 8192, 16384}) {
        for (int kind = 0; kind < 3; ++kind) {
            auto launch = [&] {
                if (kind == 0) hipLaunchKernelGGL(k_stream_read<false>, dim3(grid), dim3(256), 0, 0, a.p, n, o.p);
                else if (kind == 1) hipLaunchKernelGGL(k_stream_read<true>, dim3(grid), dim3(256), 0, 0, a.p, n, o.p);
                else hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, 0, a.p, b.p, n);
            };
            launch();
            hip_ok(hipGetLastError(), "stream launch");
            hip_ok(hipEventRecord(ev.e[0], 0), "hipEventRecord");
            for (int r = 0; r < reps; ++r) launch();
            hip_ok(hipEventRecord(ev.e[1], 0), "hipEventRecord");
            hip_ok(hipEventSynchronize(ev.e[1]), "hipEventSynchronize");
            float ms = 0.0f;
            hip_ok(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "hipEventElapsedTime");
            const double moved = (kind == 2 ? 2.0 : 1.0) * static_cast<double>(bytes) * reps;
            best[kind] = std::max(best[kind], moved / (1e-3 * ms) / 1e9);
        }
    }
    out->read_gbs = best[0];
    out->read_nt_gbs = best[1];
    out->copy_gbs = best[2];
    return MPDHG_OK;
}

// The fused tile kernels' memory pattern without their arithmetic: one
// workgroup of kGatherBlock lanes per chunk of kGatherChunk (index, value)
// words, each lane loading kGatherChunk / kGatherBlock of each non-temporally
// (lane t takes entries t, t + kGatherBlock, ...); block b runs on XCD b % 8
// and takes chunk (b % 8) per + b / 8, so each XCD walks a contiguous eighth
// of the chunks.  MODE 0 streams only; MODE 1 also gathers x[index] and sums
// value * x in registers; MODE 2 parks the (value, x) pairs in LDS and each
// lane walks 16 consecutive pairs, as the kernels' row walk does.  Every mode
// holds the 64 KB LDS stage (two workgroups per CU, as the dual kernel).
constexpr int kGatherBlock = 512, kGatherChunk = 8192, kGatherPt = kGatherChunk / kGatherBlock;

template <int MODE>
__global__ __launch_bounds__(kGatherBlock) void k_gather_probe(const int *__restrict__ idx,
                                                              const float *__restrict__ val,
                                                              const float *__restrict__ x, int per,
                                                              float *__restrict__ out) {
    __shared__ __align__(16) float2 lds[kGatherChunk];
    const int c = static_cast<int>(blockIdx.x & 7u) * per + static_cast<int>(blockIdx.x >> 3);
    const int *__restrict__ ib = idx + static_cast<size_t>(c) * kGatherChunk;
    const float *__restrict__ vb = val + static_cast<size_t>(c) * kGatherChunk;
    const int tid = static_cast<int>(threadIdx.x);
    int id[kGatherPt];
    float v[kGatherPt];
#pragma unroll
    for (int q = 0; q < kGatherPt; ++q) {
        id[q] = __builtin_nontemporal_load(ib + tid + q * kGatherBlock);
        v[q] = __builtin_nontemporal_load(vb + tid + q * kGatherBlock);
    }
    float s = 0.0f;
    if constexpr (MODE == 2) {
#pragma unroll
        for (int q = 0; q < kGatherPt; ++q) lds[tid + q * kGatherBlock] = make_float2(v[q], x[id[q]]);
        __syncthreads();
        const float2 *row = lds + tid * kGatherPt;
        for (int k = 0; k < kGatherPt; ++k) s = __builtin_fmaf(row[k].x, row[k].y, s);
    } else {
#pragma unroll
        for (int q = 0; q < kGatherPt; ++q)
            s = MODE == 1 ? __builtin_fmaf(v[q], x[id[q]], s) : s + v[q] * static_cast<float>(id[q] & 7);
        lds[tid * 8] = make_float2(s, s);
        __syncthreads();
        s = lds[((tid + 1) % kGatherBlock) * 8].x;
    }
    out[blockIdx.x * kGatherBlock + tid] = s;
}

int run_gather(int per_xcd, double stride, int reps, mpdhg_vendor_gather_result *out) {
    if (!out || per_xcd < 1 || per_xcd > 4096 || !(stride > 0.0) || stride > 64.0 || reps < 1)
        throw Fail{MPDHG_E_ARG, "per_xcd in [1, 4096], stride in (0, 64], reps >= 1"};
    hip_ok(hipSetDevice(0), "hipSetDevice");
    const int grid = 8 * per_xcd;
    const size_t n = static_cast<size_t>(grid) * kGatherChunk;
    // Chunk c of XCD w = c / per_xcd gathers x[base + floor(k stride)], k < kGatherChunk,
    // with base inside XCD w's own eighth of x: consecutive chunks advance by
    // half their span, as the sorted chunks of consecutive tiles do.
    const int64_t span = static_cast<int64_t>(kGatherChunk * stride) + 1;
    const int64_t eighth = std::max<int64_t>(2 * span, 1 << 20);
    const int64_t nx = 8 * eighth;
    std::vector<int> h(n);
    const int64_t step = static_cast<int64_t>(kGatherChunk * stride * 0.5) + 1;
    for (int c = 0; c < grid; ++c) {
        const int w = c / per_xcd, k8 = c % per_xcd;
        const int64_t base = w * eighth + (k8 * step) % (eighth - span);
        for (int k = 0; k < kGatherChunk; ++k)
            h[static_cast<size_t>(c) * kGatherChunk + k] = static_cast<int>(base + static_cast<int64_t>(k * stride));
    }
    Dev<int> idx(h, 0);
    Dev<float> val(n), x(static_cast<size_t>(nx)), o(static_cast<size_t>(grid) * kGatherBlock);
    hip_ok(hipMemset(val.p, 0, n * sizeof(float)), "hipMemset");
    hip_ok(hipMemset(x.p, 0, static_cast<size_t>(nx) * sizeof(float)), "hipMemset");
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events() { for (auto v : e) if (v) (void)hipEventDestroy(v); }
    } ev;
    for (auto &v : ev.e) hip_ok(hipEventCreate(&v), "hipEventCreate");
    double us[3];
    for (int mode = 0; mode < 3; ++mode) {
        auto launch = [&] {
            if (mode == 0) hipLaunchKernelGGL(k_gather_probe<0>, dim3(grid), dim3(kGatherBlock), 0, 0, idx.p, val.p, x.p, per_xcd, o.p);
            else if (mode == 1) hipLaunchKernelGGL(k_gather_probe<1>, dim3(grid), dim3(kGatherBlock), 0, 0, idx.p, val.p, x.p, per_xcd, o.p);
            else hipLaunchKernelGGL(k_gather_probe<2>, dim3(grid), dim3(kGatherBlock), 0, 0, idx.p, val.p, x.p, per_xcd, o.p);
        };
        launch();
        hip_ok(hipGetLastError(), "gather probe launch");
        hip_ok(hipEventRecord(ev.e[0], 0), "hipEventRecord");
        for (int r = 0; r < reps; ++r) launch();
        hip_ok(hipEventRecord(ev.e[1], 0), "hipEventRecord");
        hip_ok(hipEventSynchronize(ev.e[1]), "hipEventSynchronize");
        float ms = 0.0f;
        hip_ok(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]), "hipEventElapsedTime");
        us[mode] = 1e3 * ms / reps;
    }
    out->entries = static_cast<int64_t>(n);
    out->stream_us = us[0];
    out->gather_us = us[1];
    out->staged_us = us[2];
    out->stream_gbs = 8.0 * static_cast<double>(n) / (1e-6 * us[0]) / 1e9;
    return MPDHG_OK;
}

}  // namespace

extern "C" {

int mpdhg_vendor_gather(int per_xcd, double stride, int reps, mpdhg_vendor_gather_result *out) {
    try {
        return run_gather(per_xcd, stride, reps, out);
    } catch (const Fail &f) {
        g_err = f.msg;
        return f.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MPDHG_E_STATE;
    }
}

int mpdhg_vendor_stream(size_t bytes, int reps, mpdhg_vendor_stream_result *out) {
    try {
        return run_stream(bytes, reps, out);
    } catch (const Fail &f) {
        g_err = f.msg;
        return f.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MPDHG_E_STATE;
    }
}

int mpdhg_vendor_pdhg(const mpdhg_lp *lp, int iterations, int transpose_mode, double *x,
                      mpdhg_vendor_result *out) {
    try {
        return run_baseline(lp, iterations, transpose_mode, x, out);
    } catch (const Fail &f) {
        g_err = f.msg;
        return f.code;
    } catch (const mpdhg::Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MPDHG_E_STATE;
    }
}

const char *mpdhg_vendor_last_error(void) { return g_err.c_str(); }

}  // extern "C"
