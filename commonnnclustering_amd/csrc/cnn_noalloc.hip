// Labelling from a caller CSR without any device allocation.
//
// cnn_labels_from_csr_noalloc is cnn_labels_from_csr for callers that cannot
// allocate device memory (code under stream capture, serving loops with a fixed
// memory plan): every temporary of the labelling pass comes from one
// file-scope device array, `scratch`, instead of the context's arena.  The
// reference has no counterpart; its semantics are those of cnn_labels_from_csr
// (FitterExtBFS._fit, _fit.pyx:408-491).
//
// Each device holds its own copy of `scratch` for the life of the process.
// Host code never uses the array's name as an address: its device address is
// looked up once per device with hipGetSymbolAddress when the first context on
// that device is created (that device current), and only that address is
// passed to kernels.  Calls on one device share the array, so calls that
// overlap on one device (several host threads, several streams) race; every
// buffer the pass reads is cleared or fully written by the call itself, so
// calls that follow one another need nothing from the caller.
#include <mutex>

#include "cnn_internal.hpp"

namespace cnn {

constexpr int64_t kNoAllocMaxPoints = CNN_NOALLOC_MAX_POINTS;
// The pass needs at most 26 bytes per point at a time (core flags 1, snapshot
// ranks 8, parents 4, component minima 4 and sizes 4, first-occurrence flags 1
// and their scan 4; the root counts, 4, are released before the minima), plus
// the statistics counters (32 KiB), the caller-CSR check (one word) and the
// scan's temporary storage (a few KiB), each block rounded up to 256 bytes:
// about 26 MiB at n = 2^20.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
// Looked up once per device; throws on failure.
static void *scratch_address(int device) {
    if (device < 0 || device >= kMaxDevices) throw InvalidArg{"device index beyond the supported range"};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        CNN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipFailure{hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)"};
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

// Called by cnn_ctx_create with `device` current.
void noalloc_setup(int device) { (void)scratch_address(device); }

struct FixedScope {
    cnn_ctx *ctx;
    FixedWorkspace ws;
    FixedScope(cnn_ctx *c, void *base, size_t size) : ctx(c) {
        ws.base = static_cast<char *>(base);
        ws.size = size;
        ctx->fixed = &ws;
    }
    ~FixedScope() { ctx->fixed = nullptr; }
};

}  // namespace cnn

using namespace cnn;

extern "C" {

int cnn_labels_from_csr_noalloc(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n,
                                int32_t c_sim, int32_t m_core, uint8_t *consider, int64_t current_start,
                                int32_t sim_mode, int64_t *labels, cnn_label_stats *stats_out) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n < 0) throw InvalidArg{"negative n"};
        if (n > kNoAllocMaxPoints)
            throw InvalidArg{"n = " + std::to_string(n) + " exceeds the no-allocation limit of " +
                             std::to_string(kNoAllocMaxPoints) + " points; use cnn_labels_from_csr"};
        if (n > 0 && (!row_ptr || !col || !labels)) throw InvalidArg{"null buffer"};
        if (sim_mode == CNN_SIM_BITSET)
            throw InvalidArg{"CNN_SIM_BITSET needs per-row word lists (O(nnz) memory); the no-allocation entry "
                             "point tests with the sorted-list merge (CNN_SIM_AUTO or CNN_SIM_MERGE)"};
        if (sim_mode != CNN_SIM_AUTO && sim_mode != CNN_SIM_MERGE) throw InvalidArg{"unknown sim_mode"};
        CNN_HIP(hipSetDevice(ctx->device));
        void *base = scratch_address(ctx->device);
        FixedScope scope(ctx, base, kNoAllocScratchBytes);
        validate_csr(ctx, row_ptr, col, n);
        LabelRequest r{row_ptr, col, n, nullptr, c_sim, m_core, consider, current_start, CNN_SIM_MERGE,
                       false, false, 2, -1, labels};
        labels_run(ctx, r, stats_out);
        return (int)CNN_OK;
    });
}

}  // extern "C"
