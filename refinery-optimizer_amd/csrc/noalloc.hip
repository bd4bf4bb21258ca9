// The device memory of the no-allocation solve (mpdhg_solve_noalloc): one
// file-scope array per GPU, held from the library's load until the process
// ends.  Host code never uses the array's name as an address: at set-up
// (mpdhg_create, with the context's device current) its device address is
// looked up once per device and kept, and the solve hands that address to
// the bump arena every buffer of the solve is carved from.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"
#include "mpdhg.h"

static __device__ __align__(256) unsigned char scratch[MPDHG_NOALLOC_SCRATCH_BYTES];

namespace mpdhg {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
bool g_done[kMaxDevices];
void *g_ptr[kMaxDevices];
std::string g_err[kMaxDevices];

}  // namespace

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_done[device] = true;
    if (e != hipSuccess || !p) {
        g_ptr[device] = nullptr;
        g_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    } else {
        g_ptr[device] = p;
    }
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw Error(MPDHG_E_ARG, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) throw Error(MPDHG_E_STATE, "no-allocation scratch not set up on this device");
    if (!g_ptr[device]) throw Error(MPDHG_E_HIP, g_err[device]);
    return g_ptr[device];
}

}  // namespace mpdhg

namespace mpdhg {

// Every device buffer a one-LP solve makes, in bytes, each rounded up to the
// arena's 256-byte alignment (a buffer of 0 elements still takes one).
// Tiles hold whole 16-row groups, so a direction has at most one tile per
// group.  Mirrors the buffers of solve_reference / solve_enhanced and
// build_dir; the arena still throws if a solve ever needed more.
size_t noalloc_bytes(int mode, int64_t m, int64_t n, int64_t nnz, bool gather) {
    const size_t R = static_cast<size_t>(m), C = static_cast<size_t>(n), Z = static_cast<size_t>(nnz);
    const size_t TR = (R + kGroup - 1) / kGroup, TC = (C + kGroup - 1) / kGroup;
    size_t total = 0;
    auto buf = [&](size_t count, size_t elem) {
        const size_t b = (count ? count : 1) * elem;
        total += (b + 255) & ~static_cast<size_t>(255);
    };
    auto bufs = [&](int k, size_t count, size_t elem) {
        for (int i = 0; i < k; ++i) buf(count, elem);
    };
    buf(C + 1, 4);  // column and row pointers
    buf(R + 1, 4);
    bufs(2, Z, 4);  // storage-order index words, both directions
    bufs(2, Z, 4);  // FP32 values, both directions
    buf(TR, sizeof(Tile));
    buf(TC, sizeof(Tile));
    buf(1, 4);  // iteration base
    if (mode == MPDHG_MODE_REFERENCE) {
        bufs(4, C, 4);  // c, l, u, tau
        bufs(3, R, 4);  // rl, ru, sigma
        bufs(3, C, 4);  // x, xbar, xbest
        buf(R, 4);      // y
        buf(1, sizeof(RefInst));
        buf(TR, 4);  // per-tile residual maxima
        buf(TC, 4);
        if (gather) bufs(2, Z, 8);  // per direction: packed index word and FP32 value
        return total;
    }
    bufs(2, Z, 8);  // FP64 values, both directions
    buf(Z, 4);      // CSC -> CSR permutation (set-up)
    buf(R, 8);      // row and column norms of the scaling passes (set-up)
    buf(C, 8);
    bufs(3, C, 4);  // scaled c, l, u
    bufs(2, R, 4);  // scaled rl, ru
    bufs(4, C, 8);  // c, l, u, D_c in FP64
    bufs(3, R, 8);  // rl, ru, D_r in FP64
    bufs(4, C, 4);  // x, x0, x+, xbar
    bufs(4, R, 4);  // y, y0, y+, ybar
    bufs(5, C, 8);  // xbase, xchk, lambda, xbest, lambda best
    bufs(5, R, 8);  // ybase, ychk, A x, ybest, A x best
    buf(1, sizeof(EnhInst));
    buf(static_cast<size_t>(kGcCount) * TC, 8);  // per-group partials
    buf(static_cast<size_t>(kGrCount) * TR, 8);
    buf(1, 4);  // status word
    if (gather) bufs(2, Z, 16);  // per direction: packed index word, FP32 and FP64 value
    return total;
}

}  // namespace mpdhg
