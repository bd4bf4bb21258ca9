// Host engine: device memory, per-instance host preparation, tile plans.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
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

// Pinned host memory (hipHostMalloc), for asynchronous device-to-host polls.
struct Pinned {
    void *p = nullptr;
    explicit Pinned(size_t bytes) { MPDHG_HIP(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault)); }
    ~Pinned() { if (p) (void)hipHostFree(p); }
    Pinned(const Pinned &) = delete;
    Pinned &operator=(const Pinned &) = delete;
};

// One LP validated and converted to int32 CSC (the layout every kernel
// reads), with its CSR built by counting sort (gpu_solver.cpp:92-109).
struct HostLp {
    int m = 0, n = 0, nnz = 0;
    const mpdhg_lp *src = nullptr;
    std::vector<int> cptr, cidx, rptr, ridx, csc2csr;  // csc2csr[k]: CSR slot of CSC entry k
};
void prepare_structure(const mpdhg_lp &lp, HostLp &h);

// Concatenation offsets of a batch: rows, columns, nonzeros, and the
// kGroup-row groups of the rows and of the columns.
struct BatchLayout {
    std::vector<int64_t> row_off, col_off, nnz_off, rgrp_off, cgrp_off;
    int64_t rows = 0, cols = 0, nnz = 0, rgroups = 0, cgroups = 0;
};
// Layout of the instances H[act[a]]; throws if the batch overflows int32 indexing.
BatchLayout make_layout(const std::vector<HostLp> &H, const std::vector<int> &act);

// Greedy tile plan over one instance's compressed structure: whole
// kGroup-row groups, at most `block` rows and at most tn nonzeros per tile
// (a denser group alone, streamed in several chunks).  global_off / nz_off /
// grp_off: the instance's first row, first nonzero and first group in the
// batch.
void plan_tiles(const std::vector<int> &ptr, int count, int64_t global_off, int64_t nz_off,
                int64_t grp_off, int inst, int block, int tn, std::vector<Tile> &out);
TileSet make_tileset(const DBuf<Tile> &d, int count, Schedule sc, bool gather);

// Schedule of a tile set over `rows` rows holding `nnz` nonzeros, by rule:
// the largest workgroup that still gives at least kTilesWanted tiles (a
// full GPU), with the smallest chunk holding `block` rows of average
// length.  forced_block / forced_tn (mpdhg_config::spmv_block / spmv_chunk)
// override either when nonzero; throws if they name no schedule.
constexpr int64_t kTilesWanted = 1024;
Schedule choose_schedule(int64_t nnz, int64_t rows, int forced_block, int forced_tn);

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

// SpMV layout choice for a solve (mpdhg_config::spmv_layout).
enum class Layout { Auto = 0, Storage = 1 };

namespace detail {
template <typename V>
void permute_chunk(std::vector<V> &vals, int c0, const std::vector<uint32_t> &key, uint32_t mask,
                   std::vector<V> &tmp) {
    const int nc = static_cast<int>(key.size());
    tmp.assign(vals.begin() + c0, vals.begin() + c0 + nc);
    for (int q = 0; q < nc; ++q) vals[static_cast<size_t>(c0 + q)] = tmp[key[static_cast<size_t>(q)] & mask];
}
}  // namespace detail

// Puts every chunk (tn nonzeros) of every tile of a tile set into gather
// order (see Tile) when the layout is Auto and every instance gathers from a
// vector shorter than gather_local_limit(tn): the global indices in `idx`
// (storage order) become packed (local << log2 tn | position) words sorted
// ascending, each value array is permuted alike, and Tile::base is set to
// the vector's offset.  Otherwise the whole set keeps storage order
// (base = -1).  Returns whether the set is in gather order.  vec_off /
// vec_len: offset and length of each instance's gathered vector.
template <typename... Vs>
bool gather_order(Layout layout, int tn, const std::vector<int> &ptr, std::vector<Tile> &tiles,
                  const std::vector<int64_t> &vec_off, const std::vector<int64_t> &vec_len,
                  std::vector<int> &idx, std::vector<Vs> &...vals) {
    bool all = (layout == Layout::Auto);
    for (int64_t len : vec_len) all = all && (len < gather_local_limit(tn));
    if (!all) {
        for (Tile &tl : tiles) tl.base = -1;
        return false;
    }
    constexpr int kGroup = 64;  // tiles per work item
    const int bits = ilog2(tn);
    const uint32_t mask = static_cast<uint32_t>(tn) - 1u;
    const int nt = static_cast<int>(tiles.size());
    parallel_for_each((nt + kGroup - 1) / kGroup, [&](int g) {
        std::vector<uint32_t> key;
        std::tuple<std::vector<Vs>...> tmp;
        for (int t = g * kGroup; t < std::min(nt, (g + 1) * kGroup); ++t) {
            Tile &tl = tiles[static_cast<size_t>(t)];
            const size_t a = static_cast<size_t>(tl.inst);
            const int64_t off = vec_off[a];
            tl.base = static_cast<int32_t>(off);
            const int lo = ptr[static_cast<size_t>(tl.begin)], hi = ptr[static_cast<size_t>(tl.end)];
            for (int c0 = lo; c0 < hi; c0 += tn) {
                const int nc = std::min(hi, c0 + tn) - c0;
                key.resize(static_cast<size_t>(nc));
                for (int q = 0; q < nc; ++q)
                    key[static_cast<size_t>(q)] =
                        (static_cast<uint32_t>(idx[static_cast<size_t>(c0 + q)] - off) << bits) |
                        static_cast<uint32_t>(q);
                std::sort(key.begin(), key.end());
                for (int q = 0; q < nc; ++q) idx[static_cast<size_t>(c0 + q)] = static_cast<int>(key[static_cast<size_t>(q)]);
                std::apply([&](auto &...tv) { (detail::permute_chunk(vals, c0, key, mask, tv), ...); }, tmp);
            }
        }
    });
    return true;
}

// FP64 post-processing on the host for reference-mode results: A x,
// c + A^T y and the relative KKT of (x, y).
void host_kkt(const mpdhg_lp &lp, const HostLp &h, const std::vector<double> &x,
              const std::vector<double> &y, mpdhg_result &r);

double now_sec();

}  // namespace mpdhg
