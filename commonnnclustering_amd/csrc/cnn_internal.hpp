// Internal declarations shared by the HIP translation units of libcnn_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cnn_gpu.h"

namespace cnn {

constexpr int kTile = 64;        // points per tile: one wave64 lane per row
constexpr int kWave = 64;
constexpr int kSpatialMaxD = 8;  // d <= 8: spatial (Morton) order + tile-box culling
constexpr int kMfmaMinD = 16;    // d >= 16: FP32 Gram form on MFMA in CNN_F32_BAND
constexpr int kBatchWorkers = 6;    // datasets of a batch fitted at once (own stream each)
constexpr int64_t kPivotMinN = 2048;  // d > 8: smallest n ordered by nearest-pivot cells

struct Error {
    int code;
    std::string msg;
};

// Tuned launch shapes of one problem class (cnn_tune_params); defaults when
// the table holds no entry for the class.
struct TuneParams {
    int mfma_col_chunk = 64;
    int wide_col_chunk = 32;
    int emit_waves = 8;
    int root_passes = 2;
    int64_t words_min_fill = 8;
};
TuneParams tune_lookup(int d, int64_t n, int precision);

// Device-memory arena with best-fit reuse.  Blocks stay allocated until trim();
// all users enqueue on the context stream, so reuse is stream-ordered.
class Arena {
public:
    explicit Arena(int device) : device_(device) {}
    ~Arena() { trim(); }
    void *get(size_t bytes);
    void put(void *p);
    void trim();
    int64_t bytes() const { return total_; }

private:
    struct Block { void *p; size_t size; bool used; };
    int device_;
    std::vector<Block> blocks_;
    int64_t total_ = 0;
};

// A fixed device buffer handed out front to back (entry points that may not
// allocate).  Blocks are 256-byte aligned; releasing the most recent block
// makes its bytes available again (Scratch buffers die in reverse order).
struct FixedWorkspace {
    char *base = nullptr;
    size_t size = 0;
    size_t used = 0;
    size_t peak = 0;
};

}  // namespace cnn

struct cnn_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    cnn::Arena *arena = nullptr;
    std::string last_error;
    hipEvent_t ev[12] = {};
    int cu_count = 256;
    // contexts (own stream and arena each) that fit independent datasets of a
    // batch concurrently; created on first use, destroyed with this context
    std::vector<cnn_ctx *> workers;
    // when set, every Scratch buffer comes from this fixed buffer instead of
    // the arena (no device allocation)
    cnn::FixedWorkspace *fixed = nullptr;
    // page-locked host staging for small device-to-host reads: copies into it
    // are asynchronous, so several reads share one synchronisation (a copy to
    // pageable memory is a round trip of its own)
    static constexpr size_t kPinnedBytes = (size_t)1 << 16;
    uint8_t *pinned = nullptr;
};

struct cnn_graph {
    int device = 0;
    int64_t n = 0;               // rows of the working order (packed batches include padding rows)
    int64_t n_points = 0;        // caller points
    int32_t d = 0;
    int32_t dtype = CNN_F32;
    int32_t precision = CNN_PLAIN_F64;
    double radius = 0.0;
    double r2 = 0.0;
    double *bounds = nullptr;    // periodic metric: period per dimension (device, d entries)
    cnn::Arena *arena = nullptr; // memory owner when the graph lives inside one call (else hipMalloc)
    int32_t layout = 0;          // 0 identity, 1 spatial order
    int32_t dp = 0;              // row stride of xw32 (d, or d padded to 8 for the MFMA form)
    int64_t n_tiles = 0;
    int64_t nnz = 0;
    // device buffers owned by the graph
    int32_t *perm = nullptr;     // working -> caller index (nullptr when identity)
    double *xw64 = nullptr;      // working-order coordinates, FP64, n_tiles*64 rows
    float *xw32 = nullptr;       // working-order coordinates, FP32 (band precision)
    int64_t *row_ptr = nullptr;  // working CSR
    int32_t *col = nullptr;
    int64_t *tp_ptr = nullptr;   // candidate column tiles per row tile (spatial layout, culled dense layout)
    int32_t *tp_col = nullptr;
    int32_t *selfpos = nullptr;  // per working row: position of the row's own index in its list
    // single-batch builds with tile-pair lists: per candidate tile pair k (rows of tile R = the
    // pairs tp_ptr[R] .. tp_ptr[R+1]), mask[64k + r] holds row 64R+r's bits over
    // the 64 points of column tile tp_col[k], valid where pair_nz[k] != 0
    uint64_t *mask = nullptr;
    uint8_t *pair_nz = nullptr;
    int64_t n_tile_pairs = 0;
    int32_t *diag = nullptr;     // spatial layout: per row tile, its own pair's place in its list (nullable)
    // spatial one-batch builds: the row tiles, heavy ones first (nullable)
    int32_t *tile_order = nullptr;
    int64_t bytes = 0;
    cnn_graph_stats stats{};
    cnn::TuneParams tune{};      // launch shapes of the graph's class
    // a graph that keeps its words emits its lists (col, selfpos) only when a
    // caller needs them; this lock makes that happen once
    std::mutex lists_mu;
};

namespace cnn {

struct InvalidArg { std::string msg; };
struct OutOfMemory { std::string msg; };
struct Capacity { std::string msg; };

[[noreturn]] inline void fixed_exhausted(size_t bytes, const FixedWorkspace &f) {
    throw Capacity{"fixed workspace exhausted: " + std::to_string(bytes) + " bytes requested, " +
                   std::to_string(f.size - f.used) + " of " + std::to_string(f.size) + " free"};
}

// RAII scratch buffer from the context arena (or from the context's fixed
// workspace when one is set).
template <typename T>
struct Scratch {
    cnn_ctx *ctx;
    T *p = nullptr;
    size_t fixed_from = 0, fixed_to = 0;  // block within ctx->fixed (fixed_to == 0: arena)
    Scratch(cnn_ctx *c, size_t count) : ctx(c) {
        if (count == 0) count = 1;
        if (ctx->fixed) {
            FixedWorkspace &f = *ctx->fixed;
            const size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
            if (bytes > f.size - f.used) fixed_exhausted(bytes, f);
            fixed_from = f.used;
            fixed_to = f.used + bytes;
            f.used = fixed_to;
            if (f.used > f.peak) f.peak = f.used;
            p = reinterpret_cast<T *>(f.base + fixed_from);
            return;
        }
        p = static_cast<T *>(ctx->arena->get(count * sizeof(T)));
    }
    ~Scratch() {
        if (!p) return;
        if (fixed_to) {
            if (ctx->fixed && ctx->fixed->used == fixed_to) ctx->fixed->used = fixed_from;
            return;
        }
        ctx->arena->put(p);
    }
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
};

struct HipFailure {
    hipError_t err;
    const char *what;
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw HipFailure{e, what};
}

#define CNN_HIP(call) ::cnn::hip_check((call), #call)


// Convert exceptions raised inside an entry point into a status code.
template <typename F>
int guarded(cnn_ctx *ctx, F &&f) {
    try {
        return f();
    } catch (const HipFailure &h) {
        if (ctx) {
            ctx->last_error = std::string(h.what) + ": " + hipGetErrorString(h.err);
            (void)hipGetLastError();
        }
        return h.err == hipErrorOutOfMemory ? CNN_ERR_OOM : CNN_ERR_HIP;
    } catch (const InvalidArg &e) {
        if (ctx) ctx->last_error = e.msg;
        return CNN_ERR_INVALID;
    } catch (const OutOfMemory &e) {
        if (ctx) ctx->last_error = e.msg;
        return CNN_ERR_OOM;
    } catch (const Capacity &e) {
        if (ctx) ctx->last_error = e.msg;
        return CNN_ERR_CAPACITY;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return CNN_ERR_HIP;
    } catch (...) {
        if (ctx) ctx->last_error = "unknown failure";
        return CNN_ERR_HIP;
    }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// --------------------------------------------------------------- host helpers
void exclusive_scan_i32_to_i64(cnn_ctx *ctx, const int32_t *in, int64_t *out, int64_t n);
void exclusive_scan_i64(cnn_ctx *ctx, const int64_t *in, int64_t *out, int64_t n);
void inclusive_scan_u8_to_i32(cnn_ctx *ctx, const uint8_t *in, int32_t *out, int64_t n);
void sort_pairs_u64_i32(cnn_ctx *ctx, uint64_t *keys, int32_t *vals, int64_t n, int end_bit);
// Stable sort of vals by keys[0, end_bit); only vals come back in key order
// (keys are left in an unspecified order).
void sort_pairs_u32_i32(cnn_ctx *ctx, uint32_t *keys, int32_t *vals, int64_t n, int end_bit);
void sort_pairs_i64_i32(cnn_ctx *ctx, int64_t *keys, int32_t *vals, int64_t n);
void segmented_sort_i32(cnn_ctx *ctx, int32_t *keys, int64_t nnz, const int64_t *offsets,
                        int64_t n_segments, int end_bit);
void fill_i32(cnn_ctx *ctx, int32_t *p, int32_t v, int64_t n);
void fill_u8(cnn_ctx *ctx, uint8_t *p, uint8_t v, int64_t n);
void fill_i64(cnn_ctx *ctx, int64_t *p, int64_t v, int64_t n);

template <typename T>
T read_scalar(cnn_ctx *ctx, const T *dev) {
    T *h = (T *)ctx->pinned;
    CNN_HIP(hipMemcpyAsync(h, dev, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    CNN_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
}

// Small device-to-host reads gathered in the context's page-locked staging
// buffer: get() enqueues a copy and returns where its value will be, wait()
// synchronises once; the values are valid after wait() until the next use of
// the staging buffer (read_scalar included).
struct PinnedReads {
    cnn_ctx *ctx;
    size_t off = 0;
    explicit PinnedReads(cnn_ctx *c) : ctx(c) {}
    template <typename T>
    T *get(const T *dev, size_t count = 1) {
        off = (off + 15) & ~(size_t)15;
        if (off + count * sizeof(T) > cnn_ctx::kPinnedBytes) throw HipFailure{hipErrorInvalidValue, "staging buffer"};
        T *h = (T *)(ctx->pinned + off);
        CNN_HIP(hipMemcpyAsync(h, dev, count * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
        off += count * sizeof(T);
        return h;
    }
    void wait() { CNN_HIP(hipStreamSynchronize(ctx->stream)); }
};

// dev[0] = v, enqueued as a kernel (the value travels as an argument).
void set_i64(cnn_ctx *ctx, int64_t *dev, int64_t v);

// --------------------------------------------------------------- graph / labels
// transient: the graph is freed before the calling entry point returns, on the
// same context; its memory then comes from the context's arena.
cnn_graph *graph_build(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype,
                       double r2, int32_t precision, const double *bounds_host = nullptr,
                       bool transient = false);
void graph_free(cnn_graph *g);
void graph_export(cnn_ctx *ctx, const cnn_graph *g, int64_t *row_ptr, int32_t *col);
// Throws InvalidArg unless row_ptr (n+1, from 0) is non-decreasing and every
// row of col is strictly ascending within [0, n).
void validate_csr(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n);
cnn_graph *graph_rethreshold(cnn_ctx *ctx, const cnn_graph *g, double r2, int32_t precision);
// Emits the neighbour lists (col, selfpos) of a graph that holds only its row
// words; no-op when they exist.  Safe to call from several contexts at once;
// the lists are complete when it returns.
void graph_lists(cnn_ctx *ctx, const cnn_graph *g);
int64_t csr_filter(cnn_ctx *ctx, const double *xw64, int64_t n, int d, const int64_t *rp, const int32_t *col,
                   double r2, int64_t *out_rp, int32_t **out_col);
// Row words of a spatial graph (tile pairs tp_ptr / tp_col, pair-major words)
// re-thresholded exactly at r2 <= the words' radius: dst / dst_nz in the same
// layout, per-row degrees in deg.  Enqueued only (no host synchronisation).
void words_filter(cnn_ctx *ctx, const cnn_graph *g, const uint64_t *src, const uint8_t *src_nz, double r2,
                  uint64_t *dst, uint8_t *dst_nz, int32_t *deg);
// Per-row degrees of a graph (from its row offsets); enqueued only.
void words_degrees(cnn_ctx *ctx, const cnn_graph *g, int32_t *deg);

// Host-side description of a packed batch graph.
struct PackedHost {
    std::vector<int64_t> tile_start;  // first tile of each dataset (n_sets + 1 entries)
    std::vector<int32_t> tile_ds;     // dataset of each tile
};
// Many datasets (rows offsets[b] .. offsets[b+1] of X, own r2[b]) in one graph:
// each dataset is Morton-ordered and padded to whole tiles (padding rows have
// perm -1 and NaN coordinates); candidate tile pairs never cross datasets.
cnn_graph *graph_build_packed(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d,
                              int32_t dtype, const double *r2, int32_t precision, PackedHost &ph);

struct LabelRequest {
    const int64_t *row_ptr;
    const int32_t *col;
    int64_t n;
    const int32_t *perm;   // working -> caller index, nullptr for identity
    int32_t c_sim;
    int32_t m_core;
    uint8_t *consider;     // caller order, nullable
    int64_t current_start;
    int32_t sim_mode;
    bool sort;             // sort_by_size on the fresh labels
    bool write_all;        // write every label (0 for noise) instead of members only
    int64_t member_cutoff;
    int64_t max_clusters;
    int64_t *labels;       // caller order
    // optional tile structure of the working order (rows of tile t are 64t..64t+63)
    const int64_t *tp_ptr = nullptr;  // candidate column tiles per row tile
    const int32_t *tp_col = nullptr;
    int64_t n_tiles = 0;              // 0: no tile structure (caller CSR)
    int64_t n_tile_pairs = 0;         // tp_ptr[n_tiles], known on the host
    const int32_t *selfpos = nullptr; // per row: position of the row's own index in its list (nullable)
    const uint64_t *mask = nullptr;   // the lists as row words per candidate tile pair (nullable)
    const uint8_t *pair_nz = nullptr;
    const int32_t *diag = nullptr;    // per row tile: its own pair's place in its list (nullable)
    const int32_t *tile_order = nullptr;  // the row tiles, heavy ones first (nullable: ascending)
    // lists as per-row word lists (nullable): row w's words are wword[wptr[w] ..
    // wptr[w+1]), word t holding the row's neighbours among the 64 points of
    // column tile wtile[t], tiles ascending
    const int64_t *wptr = nullptr;
    const int32_t *wtile = nullptr;
    const uint64_t *wword = nullptr;
    // packed batches: per-tile cutoffs and the per-dataset cluster ordering
    const int32_t *tile_csim = nullptr;
    const int32_t *tile_mcore = nullptr;
    const struct PackedSort *packed = nullptr;
    // tuned switches (graph requests take the graph's; caller lists the d = 0 class)
    int root_passes = 2;
    int64_t words_min_fill = 8;
};

// Sort key layout of a packed batch: dataset << ds_shift | (size_max - size)
// << size_shift | smallest original index (relative to ds_offset when set).
struct PackedSort {
    int n_sets;
    int ds_shift;
    int size_shift;
    uint32_t size_max;
    int key_bits;
    const int32_t *tile_ds;     // device, per tile
    const int64_t *ds_offset;   // device, per dataset (nullable)
};
void labels_run(cnn_ctx *ctx, const LabelRequest &req, cnn_label_stats *st);
// Request over a resident graph (lists and/or words, tile structure, order);
// emits the lists first when the similarity form needs them (sim_mode MERGE,
// or a graph without words).  The caller fills in cutoffs and outputs.
LabelRequest graph_label_request(cnn_ctx *ctx, const cnn_graph *g, int32_t sim_mode);
// Looks up this device's copy of the no-allocation workspace (device current).
void noalloc_setup(int device);
int64_t sort_by_size_generic(cnn_ctx *ctx, int64_t *labels, int64_t n, int64_t member_cutoff,
                             int64_t max_clusters);

// Labelling of several cutoff pairs of one graph at once (launch y of every
// kernel takes pair y), enqueued without host synchronisation or memset, so
// a stream capture can record it.  Buffers hold n_c slices each.
struct MultiLabelBufs {
    int n_c = 0;
    int64_t n = 0, n_tiles = 0;
    uint8_t *core = nullptr;
    int *parent = nullptr;
    uint64_t *rank = nullptr;
    int32_t *cnt = nullptr;
    int32_t *hq = nullptr;
    unsigned int *hq_len = nullptr;
    uint64_t *tmax = nullptr, *cand_max = nullptr, *tile_common = nullptr;
    int32_t *cmin = nullptr, *csize = nullptr, *newlab = nullptr;
    // sort_by_size: the clusters of each pair (keys, roots) and their count
    uint64_t *ckey = nullptr;
    int32_t *cidx = nullptr;
    unsigned int *ccount = nullptr;                    // n_c
    const int32_t *csim = nullptr, *mcore = nullptr;  // device, n_c each
};
struct MultiLabelSizes {
    int64_t per_point = 0, per_tile = 0;
};
// The row words the pass reads: a spatial graph's tile-pair words at one radius.
struct MultiLabelSource {
    const uint64_t *mask;
    const uint8_t *pair_nz;
    const int64_t *tp_ptr;
    const int32_t *tp_col;
    const int32_t *deg;
    const int32_t *perm;
    int64_t n_tile_pairs;  // tp_ptr[n_tiles], known on the host
};
void multi_labels_bytes(int64_t n, int64_t n_tiles, int n_c, MultiLabelSizes &sz);
// labels: n_c blocks of n (block y = pair y), sorted by size; counters: 8 per
// pair ([2] core points, [3] components, [7] clusters); tstats: 4 * 1024 per pair.
void labels_enqueue_multi(cnn_ctx *ctx, const MultiLabelBufs &b, const MultiLabelSource &src, int root_passes,
                          int64_t member_cutoff, int64_t max_clusters, int64_t *labels, unsigned long long *counters,
                          unsigned long long *tstats);
void fill_u64(cnn_ctx *ctx, unsigned long long *p, unsigned long long v, int64_t n);

}  // namespace cnn

// --------------------------------------------------------------- device helpers
#if defined(__HIPCC__)
namespace cnn {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t x) { return __popcll(x); }

// Inclusive prefix sum across the 64 lanes of a wave.
// Inclusive prefix sum over the 64 lanes with DPP moves (no LDS round trips):
// shifts by 1, 2, 4, 8 inside each row of 16 lanes (out-of-row sources read
// 0), then the last lane of row 0 is added to row 1 and row 3 (row_bcast:15)
// and the last lane of row 1 to rows 2 and 3 (row_bcast:31).  Every lane of
// the wave must be active.
__device__ __forceinline__ int wave_incl_scan(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return x;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ long long wave_sum64(long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T uniform(T v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int64_t uniform64(int64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int atomic_load_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Value of lane ^ S (S = 1..16) without LDS address traffic: DPP quad
// permutes for 1 and 2, half-row mirror (^7) then quad reverse (^3) for 4,
// row rotate by 8 for 8, and gfx950's v_permlane16_swap for 16 (it swaps the
// odd 16-lane rows of its first operand with the even rows of its second).
template <int S>
__device__ __forceinline__ uint32_t lane_xor32(uint32_t v) {
    if constexpr (S == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    } else if constexpr (S == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    } else if constexpr (S == 4) {
        const int m7 = __builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false);
        return (uint32_t)__builtin_amdgcn_mov_dpp(m7, 0x1B, 0xF, 0xF, false);
    } else if constexpr (S == 8) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false);
    } else {
        static_assert(S == 16, "lane_xor32 covers 1, 2, 4, 8 and 16");
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return (threadIdx.x & 16) ? r[0] : r[1];
    }
}

// One butterfly stage of transpose_bits for S < 32 on the two 32-bit halves
// of the lane's word.  The half of each 32-bit half that the partner lane
// (lane ^ S) needs is packed into one 32-bit payload (the low half's share in
// the M positions, the high half's in ~M), exchanged once, and inserted.
template <int S, uint32_t M>
__device__ __forceinline__ void transpose_stage(uint32_t &lo, uint32_t &hi) {
    const bool up = (threadIdx.x & S) != 0;
    const uint32_t sa = up ? 0u : (uint32_t)S, sb = (uint32_t)S - sa;
    const uint32_t keep = up ? ~M : M;
    const uint32_t p = ((lo >> sa) & M) | ((hi << sb) & ~M);
    const uint32_t q = lane_xor32<S>(p);
    lo = (lo & keep) | ((q << sa) & ~keep);
    hi = (hi & keep) | ((q >> sb) & ~keep);
}

// The same transpose as transpose_bits, by six 64-bit lane permutes through
// the LDS crossbar: slower, but it keeps no lane-dependent constants live, for
// kernels whose occupancy is set by their register count.
__device__ __forceinline__ uint64_t transpose_bits_shfl(uint64_t x) {
    const int lane = threadIdx.x & 63;
    const uint64_t masks[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                               0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        const int sft = 32 >> t;
        const uint64_t m = masks[t];
        const uint64_t y = __shfl_xor(x, sft, 64);
        x = (lane & sft) ? ((x & ~m) | ((y >> sft) & m)) : ((x & m) | ((y & m) << sft));
    }
    return x;
}

// Bit transpose across the wave: lane j receives the word whose bit i is bit j
// of lane i's input.  Each stage swaps the off-diagonal blocks of width
// 32, 16, ..., 1 between partner lanes (lane ^ s); the stages act on
// independent (lane bit, word bit) pairs, so their order is free.  Width 32 is
// one v_permlane32_swap of the low halves' upper row with the high halves'
// lower row; the others are transpose_stage.
__device__ __forceinline__ uint64_t transpose_bits(uint64_t x) {
    const auto r = __builtin_amdgcn_permlane32_swap((uint32_t)x, (uint32_t)(x >> 32), false, false);
    uint32_t lo = r[0], hi = r[1];
    transpose_stage<16, 0x0000FFFFu>(lo, hi);
    transpose_stage<8, 0x00FF00FFu>(lo, hi);
    transpose_stage<4, 0x0F0F0F0Fu>(lo, hi);
    transpose_stage<2, 0x33333333u>(lo, hi);
    transpose_stage<1, 0x55555555u>(lo, hi);
    return ((uint64_t)hi << 32) | lo;
}

// |A & B| >= c for sorted, duplicate-free lists, evaluated by the whole wave:
// 64-wide windows of both lists advance like the two-pointer merge; each lane
// looks its B element up in the A window by a shuffle binary search.  Every
// common element <= min(last A, last B) lies in both windows, so each is
// counted once.  Wave-uniform result.
__device__ inline bool wave_sim(const int32_t *__restrict__ A, int64_t na, const int32_t *__restrict__ B, int64_t nb,
                         int c) {
    if (c <= 0) return true;
    if (na < c || nb < c) return false;
    const int lane = __lane_id();
    int64_t ia = 0, ib = 0;
    int count = 0;
    while (ia < na && ib < nb) {
        const int va = (int)min<int64_t>(64, na - ia);
        const int vb = (int)min<int64_t>(64, nb - ib);
        const int a = lane < va ? A[ia + lane] : INT_MAX;
        const int b = lane < vb ? B[ib + lane] : INT_MAX;
        const int alast = __shfl(a, va - 1, 64);
        const int blast = __shfl(b, vb - 1, 64);
        const int m = min(alast, blast);
        int lo = 0, hi = va;
#pragma unroll
        for (int step = 0; step < 7; ++step) {
            const int mid = (lo + hi) >> 1;
            const int am = __shfl(a, min(mid, 63), 64);
            if (lo < hi) {
                if (am < b) lo = mid + 1;
                else hi = mid;
            }
        }
        const int found = __shfl(a, min(lo, 63), 64);
        const bool hit = (lane < vb) && (b <= m) && (lo < va) && (found == b);
        count += __popcll(__ballot(hit));
        if (count >= c) return true;
        ia += __popcll(__ballot(lane < va && a <= m));
        ib += __popcll(__ballot(lane < vb && b <= m));
        if ((int64_t)count + min(na - ia, nb - ib) < (int64_t)c) return false;
    }
    return false;
}

// |A ∩ B| of two strictly ascending lists (the count SimilarityChecker._get
// returns, _types.pyx:2612-2634): the merge of wave_sim without its early exits.
// Every lane of the wave must call it with the same arguments.
__device__ inline int wave_common(const int32_t *__restrict__ A, int64_t na, const int32_t *__restrict__ B,
                                  int64_t nb) {
    const int lane = __lane_id();
    int64_t ia = 0, ib = 0;
    int count = 0;
    while (ia < na && ib < nb) {
        const int va = (int)min<int64_t>(64, na - ia);
        const int vb = (int)min<int64_t>(64, nb - ib);
        const int a = lane < va ? A[ia + lane] : INT_MAX;
        const int b = lane < vb ? B[ib + lane] : INT_MAX;
        const int alast = __shfl(a, va - 1, 64);
        const int blast = __shfl(b, vb - 1, 64);
        const int m = min(alast, blast);
        int lo = 0, hi = va;
#pragma unroll
        for (int step = 0; step < 7; ++step) {
            const int mid = (lo + hi) >> 1;
            const int am = __shfl(a, min(mid, 63), 64);
            if (lo < hi) {
                if (am < b) lo = mid + 1;
                else hi = mid;
            }
        }
        const int found = __shfl(a, min(lo, 63), 64);
        const bool hit = (lane < vb) && (b <= m) && (lo < va) && (found == b);
        count += __popcll(__ballot(hit));
        ia += __popcll(__ballot(lane < va && a <= m));
        ib += __popcll(__ballot(lane < vb && b <= m));
    }
    return count;
}

// A row's neighbour set as a list of 64-bit words over column tiles: word t
// holds the row's neighbours among points 64*tile[t] .. 64*tile[t] + 63 (bit b
// = point 64*tile[t] + b), tiles strictly ascending.  Words are strided
// (stride 64 for the per-tile-pair masks of a spatial graph, 1 for per-row
// lists); nz (nullable) flags the words that were stored -- an unflagged word
// is empty.
struct WordRow {
    const int32_t *tile;
    const uint64_t *word;
    const uint8_t *nz;
    int64_t len;
    int stride;
};

__device__ __forceinline__ uint64_t word_at(const WordRow &r, int64_t t) {
    if (r.nz && r.nz[t] == 0) return 0ull;
    return r.word[t * (int64_t)r.stride];
}

// |A & B| >= c from two rows' word lists, by the whole wave: the words of
// column tiles present in both lists are and-ed and their bits counted.
// 64-wide windows of both tile lists advance like the two-pointer merge of
// wave_sim, each lane looking its B tile up in the A window by a shuffle binary
// search; rows sharing one tile list (the rows of one row tile of a spatial
// graph) pair their words lane by lane.  Early exit once c common neighbours
// are counted.  Wave-uniform result; every lane must call it.
__device__ inline bool wave_sim_words(const WordRow &A, const WordRow &B, int c) {
    if (c <= 0) return true;
    const int lane = __lane_id();
    int count = 0;
    if (A.tile == B.tile && A.len == B.len) {
        // one list: the next window's pair flags load with this window's words
        int z = lane < A.len ? (A.nz ? (int)A.nz[lane] : 1) : 0;
        for (int64_t t0 = 0; t0 < A.len; t0 += 64) {
            const int64_t t = t0 + lane, t2 = t + 64;
            const int z2 = t2 < A.len ? (A.nz ? (int)A.nz[t2] : 1) : 0;
            int cnt = 0;
            if (z) cnt = __popcll(A.word[t * (int64_t)A.stride] & B.word[t * (int64_t)B.stride]);
            count += wave_sum(cnt);
            if (count >= c) return true;
            z = z2;
        }
        return false;
    }
    // windows of 64 tiles of each list; the next windows' tile ids and pair
    // flags are loaded together with this window's words, and a matched A
    // position's flag comes from the lane that holds it: one memory round trip
    // per window, words only for pairs whose flags are both set
    int64_t ia = 0, ib = 0;
    int va = (int)min<int64_t>(64, A.len), vb = (int)min<int64_t>(64, B.len);
    int ta = lane < va ? A.tile[lane] : INT_MAX;
    int tb = lane < vb ? B.tile[lane] : INT_MAX;
    int za = lane < va ? (A.nz ? (int)A.nz[lane] : 1) : 0;
    int zb = lane < vb ? (B.nz ? (int)B.nz[lane] : 1) : 0;
    while (ia < A.len && ib < B.len) {
        const int alast = __shfl(ta, va - 1, 64);
        const int blast = __shfl(tb, vb - 1, 64);
        const int m = min(alast, blast);
        int lo = 0, hi = va;
#pragma unroll
        for (int step = 0; step < 7; ++step) {
            const int mid = (lo + hi) >> 1;
            const int am = __shfl(ta, min(mid, 63), 64);
            if (lo < hi) {
                if (am < tb) lo = mid + 1;
                else hi = mid;
            }
        }
        const int found = __shfl(ta, min(lo, 63), 64);
        const int zfound = __shfl(za, min(lo, 63), 64);
        const bool hit = lane < vb && tb <= m && lo < va && found == tb && zfound != 0 && zb != 0;
        const int64_t pa = ia + lo, pb = ib + lane;
        const int64_t ia2 = ia + __popcll(__ballot(lane < va && ta <= m));
        const int64_t ib2 = ib + __popcll(__ballot(lane < vb && tb <= m));
        const int va2 = (int)max<int64_t>(0, min<int64_t>(64, A.len - ia2));
        const int vb2 = (int)max<int64_t>(0, min<int64_t>(64, B.len - ib2));
        const int ta2 = lane < va2 ? A.tile[ia2 + lane] : INT_MAX;
        const int tb2 = lane < vb2 ? B.tile[ib2 + lane] : INT_MAX;
        const int za2 = lane < va2 ? (A.nz ? (int)A.nz[ia2 + lane] : 1) : 0;
        const int zb2 = lane < vb2 ? (B.nz ? (int)B.nz[ib2 + lane] : 1) : 0;
        int cnt = 0;
        if (hit) cnt = __popcll(A.word[pa * (int64_t)A.stride] & B.word[pb * (int64_t)B.stride]);
        count += wave_sum(cnt);
        if (count >= c) return true;
        ia = ia2;
        ib = ib2;
        va = va2;
        vb = vb2;
        ta = ta2;
        tb = tb2;
        za = za2;
        zb = zb2;
    }
    return false;
}

}  // namespace cnn
#endif
