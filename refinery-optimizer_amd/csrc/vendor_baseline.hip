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

}  // namespace

extern "C" {

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
