// Host engine: device memory, per-instance host preparation, tile plans.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_dev.hpp"
#include "launch.hpp"
#include "mpdhg.h"

namespace mpdhg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what);
#define MPDHG_HIP(call) ::mpdhg::hip_check((call), #call)

// Owning device buffer.
template <typename T>
class DBuf {
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { resize(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void resize(size_t n) {
        release();
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T));
        if (e != hipSuccess) {
            p_ = nullptr;
            throw Error(MPDHG_E_NOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        n_ = n;
    }
    void upload(const std::vector<T> &h, hipStream_t s) {
        if (h.size() > n_) resize(h.size());
        if (!h.empty())
            MPDHG_HIP(hipMemcpyAsync(p_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(std::vector<T> &h, size_t n, hipStream_t s) const {
        h.resize(n);
        if (n) MPDHG_HIP(hipMemcpyAsync(h.data(), p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }

  private:
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *p_ = nullptr;
    size_t n_ = 0;
};

// One LP validated and converted to int32 CSC (the layout every kernel
// reads), with its CSR built by counting sort (gpu_solver.cpp:92-109).
struct HostLp {
    int m = 0, n = 0, nnz = 0;
    const mpdhg_lp *src = nullptr;
    std::vector<int> cptr, cidx, rptr, ridx, csc2csr;  // csc2csr[k]: CSR slot of CSC entry k
};
void prepare_structure(const mpdhg_lp &lp, HostLp &h);

// Concatenation offsets of a batch.
struct BatchLayout {
    std::vector<int64_t> row_off, col_off, nnz_off;
    int64_t rows = 0, cols = 0, nnz = 0;
};

// Greedy tile plan over one instance's compressed structure.
void plan_tiles(const std::vector<int> &ptr, int count, int64_t global_off, int inst,
                std::vector<Tile> &out);
TileSet make_tileset(const DBuf<Tile> &d, int count);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> events;
    ~Context();
};

void solve_reference(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                     mpdhg_result *res);
void solve_enhanced(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                    mpdhg_result *res);
void solve_adjoint(Context &ctx, const mpdhg_lp *lps, int count, mpdhg_adjoint_io *io,
                   double active_tol, double tol, int max_iter);

// Ruiz equilibration (infinity norm, `ruiz` passes) then one Pock-Chambolle
// pass (alpha = 1): D_r A D_c has || . ||_2 <= 1.
void scale_instance(const mpdhg_lp &lp, const HostLp &h, int ruiz, double *dr, double *dc);

// Runs f(i) for i in [0, count) on up to hardware_concurrency threads.
void parallel_for_each(int count, const std::function<void(int)> &f);

// FP64 post-processing on the host for reference-mode results: A x,
// c + A^T y and the relative KKT of (x, y).
void host_kkt(const mpdhg_lp &lp, const HostLp &h, const std::vector<double> &x,
              const std::vector<double> &y, mpdhg_result &r);

double now_sec();

}  // namespace mpdhg
