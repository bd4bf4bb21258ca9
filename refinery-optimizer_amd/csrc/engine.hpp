// Host engine: device memory, per-instance host preparation, tile plans.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
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

// A bump allocator over caller-provided device memory.  While one is set on
// the calling thread (ArenaScope), every DBuf takes its memory from it and
// never frees it, and the solves avoid everything else that allocates:
// pinned host buffers, events, HIP graphs and the schedule search.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes) {
        const size_t off = (used + 255) & ~static_cast<size_t>(255);
        if (off > cap || bytes > cap - off)
            throw Error(MPDHG_E_NOMEM, "scratch exhausted: " + std::to_string(off + bytes) +
                                           " bytes needed, " + std::to_string(cap) + " held");
        used = off + bytes;
        return base + off;
    }
};
Arena *&thread_arena();
inline bool arena_active() { return thread_arena() != nullptr; }
struct ArenaScope {
    explicit ArenaScope(Arena *a) { thread_arena() = a; }
    ~ArenaScope() { thread_arena() = nullptr; }
    ArenaScope(const ArenaScope &) = delete;
    ArenaScope &operator=(const ArenaScope &) = delete;
};

// Owning device buffer (or, under an ArenaScope, a span of the arena).
template <typename T>
class DBuf {
  public:
    DBuf() = default;
    explicit DBuf(size_t n) { resize(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_), owned_(o.owned_) { o.p_ = nullptr; o.n_ = 0; }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_;
            n_ = o.n_;
            owned_ = o.owned_;
            o.p_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void resize(size_t n) {
        release();
        if (n == 0) n = 1;
        if (Arena *a = thread_arena()) {
            p_ = static_cast<T *>(a->take(n * sizeof(T)));
            n_ = n;
            owned_ = false;
            return;
        }
        owned_ = true;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T));
        if (e != hipSuccess) {
            p_ = nullptr;
            throw Error(MPDHG_E_NOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                           " bytes failed: " + hipGetErrorString(e));
        }
        n_ = n;
    }
    void upload(const T *h, size_t n, hipStream_t s) {
        if (n > n_) resize(n);
        if (n) MPDHG_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T> &h, hipStream_t s) { upload(h.data(), h.size(), s); }
    void download(std::vector<T> &h, size_t n, hipStream_t s) const {
        h.resize(n);
        if (n) MPDHG_HIP(hipMemcpyAsync(h.data(), p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }

  private:
    void release() {
        if (p_ && owned_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *p_ = nullptr;
    size_t n_ = 0;
    bool owned_ = true;
};

// Pinned host memory (hipHostMalloc), for asynchronous device-to-host polls;
// ordinary host memory under an ArenaScope, whose solves poll synchronously.
struct Pinned {
    void *p = nullptr;
    std::vector<unsigned char> plain;
    explicit Pinned(size_t bytes) {
        if (arena_active()) {
            plain.resize(bytes ? bytes : 1);
            p = plain.data();
        } else {
            MPDHG_HIP(hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
        }
    }
    ~Pinned() {
        if (p && plain.empty()) (void)hipHostFree(p);
    }
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
// XCDs of an MI355X: each walks a contiguous eighth of a tile list.
constexpr size_t kXcds = 8;
// Layout order of a batch's LPs, stratified by size: the LPs (indices into H)
// ranked by nonzeros are dealt round-robin into kXcds bins, laid out bin
// after bin, so that every XCD's eighth of the tile list holds the same mix
// of LP sizes whatever the caller's order (LPs that stop together would
// otherwise leave whole XCDs idle).  No result depends on the order.
void stratify_by_size(std::vector<int> &act, const std::vector<HostLp> &H);

// prepare_structure of every LP of a batch, on the host threads; an invalid
// LP raises the error the lowest-numbered invalid LP gives.
void prepare_all(const mpdhg_lp *lps, int count, std::vector<HostLp> &H);
// LPs of at least this many nonzeros get their CSR built on all host threads
// (the same CSR), one LP after the other.
constexpr int kParallelCsrNnz = 1 << 20;

// Host staging array whose elements start uninitialised (no zero fill), for
// the batch arrays that the per-instance preparation writes in full: their
// pages are first touched by the threads that fill them.
template <typename T>
class HostArray {
  public:
    void resize(size_t n) {
        p_.reset(n ? new T[n] : nullptr);
        n_ = n;
    }
    T *data() { return p_.get(); }
    const T *data() const { return p_.get(); }
    size_t size() const { return n_; }
    T &operator[](size_t i) { return p_[i]; }
    const T &operator[](size_t i) const { return p_[i]; }

  private:
    std::unique_ptr<T[]> p_;
    size_t n_ = 0;
};

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

// Column bands of a reference-mode CSR product (mpdhg_config::spmv_bands):
// by rule, bands of at most kBandBytes of the gathered vector once it
// exceeds kBandMinBytes (an XCD's L2 holds 4 MB).  Measured on the 12M-
// nonzero LP (xbar 4.8 MB): 2 bands 65.7 us per product, none 75.6, 3 bands
// 67.4, 4 bands 71.5 (each pass rereads the row pointers and the running sums).
constexpr int kMaxBands = 16;
constexpr int64_t kBandBytes = 3 << 20;
constexpr int64_t kBandMinBytes = 3 << 20;
int band_count(int forced, int64_t vec_len);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> events;
    ~Context();
};

// warm: one start per LP, or null (every LP cold).
void solve_reference(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                     const mpdhg_warm_start *warm, mpdhg_result *res);
void solve_enhanced(Context &ctx, const mpdhg_lp *lps, int count, const mpdhg_config &cfg,
                    const mpdhg_warm_start *warm, mpdhg_result *res);
// mpdhg_config::vertex_check: the FP64 crossover from each result's point
// (host threads), filling vertex_status / vertex_objective / vertex_rel_error.
// x[q] is LP q's PDHG point.
void vertex_check(const mpdhg_lp *lps, int count, const std::vector<const double *> &x, int max_rows,
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

// One direction of a batch (CSR rows for A x, or CSC columns for A^T y) as
// the tile planner sees it: per instance its local row pointers and row
// count, and its offsets in the batch (first row, first nonzero, first row
// group, and the offset and length of the vector its rows gather from).
struct DirPlan {
    std::vector<const std::vector<int> *> ptr;
    std::vector<int> count;
    std::vector<int64_t> row_off, nnz_off, grp_off, vec_off, vec_len;
    int64_t rows = 0, nnz = 0;
};
DirPlan dir_plan(const std::vector<HostLp> &H, const std::vector<int> &act, const BatchLayout &L,
                 bool csr);

// The matrix of one direction on the device in storage order: gathered
// (global) index per nonzero and the value copies the solve uses.
struct DirSource {
    const int *idx = nullptr;
    const float *v32 = nullptr;
    const double *v64 = nullptr;
};

// A direction laid out for one schedule: tile records (also kept on the
// host, for compaction) and the matrix copy the tile kernels read.  In
// gather order the copy is owned here (built on the device from the
// source); in storage order the source arrays are used as they are.
struct DevDir {
    Schedule sc{0, 0};
    bool gather = false;
    std::vector<Tile> host;
    std::vector<int32_t> first;  // tiles [first[a], first[a + 1]) belong to instance a
    DBuf<Tile> tiles;
    DBuf<int> own_idx;
    DBuf<float> own32;
    DBuf<double> own64;
    const int *idx = nullptr;
    const float *v32 = nullptr;
    const double *v64 = nullptr;
    TileSet set() const { return make_tileset(tiles, static_cast<int>(host.size()), sc, gather); }
};

// Plans the tiles of schedule sc and lays the source out for them: gather
// order (each chunk sorted on the device by gathered index, packed index
// words, values permuted alike) when `layout` is Auto and every instance's
// vector is shorter than gather_local_limit(sc.tn), else storage order.
// want32 / want64: which value copies the solve reads.
void build_dir(const DirPlan &P, Schedule sc, Layout layout, const DirSource &src, bool want32,
               bool want64, hipStream_t st, DevDir &out);

// Schedule search of one direction: builds every compiled schedule allowed
// by forced_block / forced_tn and times `reps` launches of the FP32 SpMV of
// this direction (y = A v for CSR, z = A^T w for CSC) on each, keeping the
// fastest in `out` (with the value copies of want32 / want64).  ptr: the
// direction's global row pointers on the device; vin / vout / gpart:
// scratch of the gathered length, the row count and the row-group count.
// Returns the milliseconds spent; `trials` receives (schedule, microseconds
// per launch) for each candidate.
struct TuneTrial {
    Schedule sc;
    bool gather;
    double us;
};
// Whether a solve searches its schedules (mpdhg_config::spmv_tune: 0 never,
// 1 when the batch holds at least kTuneMinNnz nonzeros, 2 always; never
// when both spmv_block and spmv_chunk are fixed).
constexpr int64_t kTuneMinNnz = 256 * 1024;
bool tune_wanted(const mpdhg_config &cfg, int64_t nnz);
double tune_dir(const DirPlan &P, bool csr, Layout layout, const DirSource &src, const int *ptr,
                int forced_block, int forced_tn, bool want32, bool want64, float *vin, float *vout,
                double *gpart, hipStream_t st, DevDir &out, std::vector<TuneTrial> &trials);

// Column bands of one LP's CSR product (for an LP whose gathered vector
// exceeds an XCD's L2): band b holds each row's nonzeros in columns
// [cb_b, cb_{b+1}), cb_b = cols * b / bands, a contiguous piece of the row in
// storage order, as a CSR of its own; its tiles gather from that slice of the
// vector (the slice start is the tiles' gather base, so the packed index word
// of the gather-order layout only has to span the band).
struct ColumnBands {
    struct Band {
        std::vector<int> rptr, ridx;
        std::vector<float> rval;
        DBuf<int> d_rptr, d_ridx;
        DBuf<float> d_rval;
        DevDir dir;
    };
    std::vector<Band> bd;
    std::vector<TileSet> sets;  // one tile set per band
    DBuf<float> hpart;          // running row sums between band passes
    int count() const { return static_cast<int>(bd.size()); }
};
// Splits the rows x cols CSR (rptr / ridx / rval, column indices ascending
// within each row) into `bands` > 1 bands and uploads them.  The first band's
// schedule is searched when `tune` (vin / vout: scratch vectors of cols and
// rows floats) or else chosen by rule, and every band uses it.  Returns the
// search time in ms.
double build_bands(int bands, int rows, int64_t cols, const int *rptr, const int *ridx, const float *rval,
                   const mpdhg_config &cfg, Layout lay, bool tune, float *vin, float *vout,
                   hipStream_t st, ColumnBands &out);

// FP64 post-processing on the host for reference-mode results: A x,
// c + A^T y and the relative KKT of (x, y).
void host_kkt(const mpdhg_lp &lp, const HostLp &h, const std::vector<double> &x,
              const std::vector<double> &y, mpdhg_result &r);

double now_sec();

// The no-allocation solve's device memory (noalloc.hip).  noalloc_setup
// looks up, with `device` current, the device address of that device's copy
// of the library's scratch array (once per device; called at context
// creation); noalloc_scratch returns it, or throws if the lookup failed.
void noalloc_setup(int device);
void *noalloc_scratch(int device);
// Upper bound of the scratch bytes a solve of one LP of this size needs
// (every device buffer of the mode, with the gather-order copies when
// `gather`; tiles counted as one per 16-row group), checked before the solve.
size_t noalloc_bytes(int mode, int64_t m, int64_t n, int64_t nnz, bool gather);

}  // namespace mpdhg
