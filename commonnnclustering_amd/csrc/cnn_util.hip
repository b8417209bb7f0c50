// Context, workspace arena, and library primitives (scan, radix sort, fills).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>

#include "cnn_internal.hpp"

namespace cnn {

void *Arena::get(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    int best = -1;
    for (size_t i = 0; i < blocks_.size(); ++i) {
        const Block &b = blocks_[i];
        if (!b.used && b.size >= bytes && b.size <= 2 * bytes + (1u << 20)) {
            if (best < 0 || b.size < blocks_[best].size) best = (int)i;
        }
    }
    if (best >= 0) {
        blocks_[best].used = true;
        return blocks_[best].p;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // release idle blocks and retry once
        for (size_t i = 0; i < blocks_.size();) {
            if (!blocks_[i].used) {
                (void)hipFree(blocks_[i].p);
                total_ -= (int64_t)blocks_[i].size;
                blocks_.erase(blocks_.begin() + i);
            } else {
                ++i;
            }
        }
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw OutOfMemory{"device allocation of " + std::to_string(bytes) + " bytes failed"};
        }
    }
    blocks_.push_back({p, bytes, true});
    total_ += (int64_t)bytes;
    return p;
}

void Arena::put(void *p) {
    for (auto &b : blocks_)
        if (b.p == p) { b.used = false; return; }
}

void Arena::trim() {
    for (size_t i = 0; i < blocks_.size();) {
        if (!blocks_[i].used) {
            (void)hipFree(blocks_[i].p);
            total_ -= (int64_t)blocks_[i].size;
            blocks_.erase(blocks_.begin() + i);
        } else {
            ++i;
        }
    }
}

// ------------------------------------------------------------------ fills
template <typename T>
__global__ void k_fill(T *p, T v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static void fill(cnn_ctx *ctx, T *p, T v, int64_t n) {
    if (n <= 0) return;
    int64_t blocks = std::min<int64_t>(ceil_div(n, 256), 8192);
    k_fill<T><<<dim3((unsigned)blocks), dim3(256), 0, ctx->stream>>>(p, v, n);
    CNN_HIP(hipGetLastError());
}

void fill_i32(cnn_ctx *ctx, int32_t *p, int32_t v, int64_t n) { fill(ctx, p, v, n); }

__global__ void k_set_i64(int64_t *p, int64_t v) { *p = v; }
void set_i64(cnn_ctx *ctx, int64_t *dev, int64_t v) {
    k_set_i64<<<1, 1, 0, ctx->stream>>>(dev, v);
    CNN_HIP(hipGetLastError());
}
void fill_u8(cnn_ctx *ctx, uint8_t *p, uint8_t v, int64_t n) { fill(ctx, p, v, n); }
void fill_i64(cnn_ctx *ctx, int64_t *p, int64_t v, int64_t n) { fill(ctx, p, v, n); }
void fill_u64(cnn_ctx *ctx, unsigned long long *p, unsigned long long v, int64_t n) { fill(ctx, p, v, n); }

// ------------------------------------------------------------------ scans / sorts
void exclusive_scan_i32_to_i64(cnn_ctx *ctx, const int32_t *in, int64_t *out, int64_t n) {
    size_t tmp = 0;
    CNN_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, int64_t(0), (size_t)n,
                                    rocprim::plus<int64_t>(), ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::exclusive_scan(t.p, tmp, in, out, int64_t(0), (size_t)n,
                                    rocprim::plus<int64_t>(), ctx->stream));
}

void exclusive_scan_i64(cnn_ctx *ctx, const int64_t *in, int64_t *out, int64_t n) {
    size_t tmp = 0;
    CNN_HIP(rocprim::exclusive_scan(nullptr, tmp, in, out, int64_t(0), (size_t)n,
                                    rocprim::plus<int64_t>(), ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::exclusive_scan(t.p, tmp, in, out, int64_t(0), (size_t)n,
                                    rocprim::plus<int64_t>(), ctx->stream));
}

void inclusive_scan_u8_to_i32(cnn_ctx *ctx, const uint8_t *in, int32_t *out, int64_t n) {
    size_t tmp = 0;
    auto it = rocprim::make_transform_iterator(in, [] __device__(uint8_t v) { return (int32_t)v; });
    CNN_HIP(rocprim::inclusive_scan(nullptr, tmp, it, out, (size_t)n, rocprim::plus<int32_t>(),
                                    ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::inclusive_scan(t.p, tmp, it, out, (size_t)n, rocprim::plus<int32_t>(),
                                    ctx->stream));
}

void sort_pairs_u64_i32(cnn_ctx *ctx, uint64_t *keys, int32_t *vals, int64_t n, int end_bit) {
    if (n <= 1) return;
    Scratch<uint64_t> k2(ctx, n);
    Scratch<int32_t> v2(ctx, n);
    size_t tmp = 0;
    CNN_HIP(rocprim::radix_sort_pairs(nullptr, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit,
                                      ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::radix_sort_pairs(t.p, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit,
                                      ctx->stream));
    CNN_HIP(hipMemcpyAsync(keys, k2.p, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
    CNN_HIP(hipMemcpyAsync(vals, v2.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream));
}

void sort_pairs_u32_i32(cnn_ctx *ctx, uint32_t *keys, int32_t *vals, int64_t n, int end_bit) {
    if (n <= 1) return;
    Scratch<uint32_t> k2(ctx, n);
    Scratch<int32_t> v2(ctx, n);
    size_t tmp = 0;
    CNN_HIP(rocprim::radix_sort_pairs(nullptr, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit,
                                      ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::radix_sort_pairs(t.p, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, end_bit,
                                      ctx->stream));
    CNN_HIP(hipMemcpyAsync(vals, v2.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream));
}

void sort_pairs_i64_i32(cnn_ctx *ctx, int64_t *keys, int32_t *vals, int64_t n) {
    if (n <= 1) return;
    Scratch<int64_t> k2(ctx, n);
    Scratch<int32_t> v2(ctx, n);
    size_t tmp = 0;
    CNN_HIP(rocprim::radix_sort_pairs(nullptr, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, 64,
                                      ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::radix_sort_pairs(t.p, tmp, keys, k2.p, vals, v2.p, (size_t)n, 0, 64,
                                      ctx->stream));
    CNN_HIP(hipMemcpyAsync(keys, k2.p, n * sizeof(int64_t), hipMemcpyDeviceToDevice, ctx->stream));
    CNN_HIP(hipMemcpyAsync(vals, v2.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream));
}

void segmented_sort_i32(cnn_ctx *ctx, int32_t *keys, int64_t nnz, const int64_t *offsets,
                        int64_t n_segments, int end_bit) {
    if (nnz <= 1 || n_segments <= 0) return;
    Scratch<int32_t> k2(ctx, nnz);
    size_t tmp = 0;
    CNN_HIP(rocprim::segmented_radix_sort_keys(nullptr, tmp, keys, k2.p, (size_t)nnz,
                                               (unsigned)n_segments, offsets, offsets + 1, 0,
                                               end_bit, ctx->stream));
    Scratch<char> t(ctx, tmp);
    CNN_HIP(rocprim::segmented_radix_sort_keys(t.p, tmp, keys, k2.p, (size_t)nnz,
                                               (unsigned)n_segments, offsets, offsets + 1, 0,
                                               end_bit, ctx->stream));
    CNN_HIP(hipMemcpyAsync(keys, k2.p, nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, ctx->stream));
}

// ------------------------------------------------------------------ tuning table
struct TuneEntry {
    int32_t d_lo, d_hi;
    int64_t n_lo, n_hi;
    int32_t precision;
    TuneParams p;
};
static std::mutex g_tune_mu;
static std::vector<TuneEntry> g_tune;

TuneParams tune_lookup(int d, int64_t n, int precision) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (auto it = g_tune.rbegin(); it != g_tune.rend(); ++it)
        if (d >= it->d_lo && d <= it->d_hi && n >= it->n_lo && n < it->n_hi && precision == it->precision)
            return it->p;
    return TuneParams{};
}

}  // namespace cnn

// ================================================================== C-ABI: context
using namespace cnn;

extern "C" {

int cnn_ctx_create(int device, cnn_ctx **out) {
    if (!out) return CNN_ERR_INVALID;
    *out = nullptr;
    cnn_ctx *ctx = new cnn_ctx();
    int rc = guarded(ctx, [&]() {
        int count = 0;
        CNN_HIP(hipGetDeviceCount(&count));
        if (device < 0 || device >= count) throw InvalidArg{"device index out of range"};
        ctx->device = device;
        CNN_HIP(hipSetDevice(device));
        CNN_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
        ctx->stream = ctx->own_stream;
        for (auto &e : ctx->ev) CNN_HIP(hipEventCreate(&e));
        hipDeviceProp_t prop;
        CNN_HIP(hipGetDeviceProperties(&prop, device));
        ctx->cu_count = prop.multiProcessorCount;
        ctx->arena = new Arena(device);
        CNN_HIP(hipHostMalloc((void **)&ctx->pinned, cnn_ctx::kPinnedBytes, hipHostMallocDefault));
        noalloc_setup(device);
        return (int)CNN_OK;
    });
    if (rc != CNN_OK) {
        static thread_local std::string err;
        err = ctx->last_error;
        delete ctx;
        return rc;
    }
    *out = ctx;
    return CNN_OK;
}

int cnn_ctx_destroy(cnn_ctx *ctx) {
    if (!ctx) return CNN_ERR_INVALID;
    for (cnn_ctx *w : ctx->workers) cnn_ctx_destroy(w);
    ctx->workers.clear();
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx->arena;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CNN_OK;
}

const char *cnn_last_error(const cnn_ctx *ctx) {
    if (!ctx) return "null context";
    return ctx->last_error.c_str();
}

int cnn_ctx_set_stream(cnn_ctx *ctx, void *stream) {
    if (!ctx) return CNN_ERR_INVALID;
    ctx->stream = (hipStream_t)stream;  // NULL: the default (null) stream
    return CNN_OK;
}

void *cnn_ctx_get_stream(cnn_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int cnn_ctx_synchronize(cnn_ctx *ctx) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        CNN_HIP(hipStreamSynchronize(ctx->stream));
        return (int)CNN_OK;
    });
}

int cnn_ctx_workspace_bytes(const cnn_ctx *ctx, int64_t *bytes) {
    if (!ctx || !bytes) return CNN_ERR_INVALID;
    *bytes = ctx->arena->bytes();
    return CNN_OK;
}

int cnn_ctx_trim(cnn_ctx *ctx) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        CNN_HIP(hipSetDevice(ctx->device));
        CNN_HIP(hipStreamSynchronize(ctx->stream));
        ctx->arena->trim();
        return (int)CNN_OK;
    });
}

int cnn_tune_set(int32_t d_lo, int32_t d_hi, int64_t n_lo, int64_t n_hi, int32_t precision,
                 const cnn_tune_params *p) {
    if (!p || d_lo < 0 || d_hi < d_lo || n_lo < 0 || n_hi <= n_lo) return CNN_ERR_INVALID;
    if (precision != CNN_PLAIN_F64 && precision != CNN_F32_BAND) return CNN_ERR_INVALID;
    if (p->mfma_col_chunk < 8 || p->mfma_col_chunk > 1024 || p->wide_col_chunk < 1 || p->wide_col_chunk > 1024 ||
        (p->emit_waves != 4 && p->emit_waves != 8) || p->root_passes < 0 || p->root_passes > 8 ||
        p->words_min_fill < 0 || p->words_min_fill > (1 << 30))
        return CNN_ERR_INVALID;
    TuneEntry e{d_lo, d_hi, n_lo, n_hi, precision, {}};
    e.p.mfma_col_chunk = p->mfma_col_chunk;
    e.p.wide_col_chunk = p->wide_col_chunk;
    e.p.emit_waves = p->emit_waves;
    e.p.root_passes = p->root_passes;
    e.p.words_min_fill = p->words_min_fill;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.push_back(e);
    return CNN_OK;
}

int cnn_tune_get(int32_t d, int64_t n, int32_t precision, cnn_tune_params *out) {
    if (!out) return CNN_ERR_INVALID;
    const TuneParams t = tune_lookup(d, n, precision);
    out->mfma_col_chunk = t.mfma_col_chunk;
    out->wide_col_chunk = t.wide_col_chunk;
    out->emit_waves = t.emit_waves;
    out->root_passes = t.root_passes;
    out->words_min_fill = (int32_t)t.words_min_fill;
    return CNN_OK;
}

int cnn_tune_clear(void) {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
    return CNN_OK;
}

const char *cnn_version(void) { return "commonnnclustering_amd 0.1.0 (gfx950)"; }

}  // extern "C"
