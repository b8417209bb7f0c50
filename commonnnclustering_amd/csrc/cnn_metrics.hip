// Neighbourhoods from a precomputed distance matrix: MetricExtPrecomputed
// (_types.pyx:2105-2127) under NeighboursGetterExtBruteForce._get
// (_types.pyx:1692-1710): N(i) = { j : D[i, j] <= r }, ascending j, with the
// radius used as given (_adjust_radius returns it unchanged).
//
// The decisions of each row are packed into 64-bit words, one per 64 columns
// (mask[i][c], bit b = column 64c + b); the CSR is emitted from the words.  The
// labelling treats the neighbour relation as symmetric (components of the
// passing edges); for an asymmetric matrix the reference's BFS follows
// directed edges, so pairs whose two decisions differ are counted (a bit
// transpose of each 64x64 block against its mirror block) and reported.
#include <algorithm>

#include "cnn_internal.hpp"

namespace cnn {

// One wave per row: the words of row i.
template <typename T>
__global__ __launch_bounds__(256) void k_pre_masks(const T *__restrict__ D, int64_t n, int64_t n_words, double r,
                                                   uint64_t *__restrict__ mask, int32_t *__restrict__ deg) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const T *row = D + i * n;
    int count = 0;
    for (int64_t c = 0; c < n_words; ++c) {
        const int64_t j = c * 64 + lane;
        const double v = (double)row[j < n ? j : n - 1];  // clamped: no load past the row
        const bool in = j < n && v <= r;
        const uint64_t w = __ballot(in);
        if (lane == 0) mask[i * n_words + c] = w;
        count += __popcll(w);
    }
    if (lane == 0) deg[i] = count;
}

// Lane l of a wave: transposed bits of the 64x64 block (row block R, column
// block C) against the block (C, R); counts the pairs {i, j} whose two
// decisions differ.
__global__ __launch_bounds__(256) void k_pre_symmetry(const uint64_t *__restrict__ mask, int64_t n, int64_t n_words,
                                                      unsigned long long *__restrict__ n_asym) {
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= n_words * n_words) return;
    const int64_t R = item / n_words, C = item - R * n_words;
    if (C < R) return;
    const int64_t i = R * 64 + lane;   // row of block (R, C)
    const int64_t j = C * 64 + lane;   // row of block (C, R)
    const uint64_t a = i < n ? mask[i * n_words + C] : 0ull;
    const uint64_t b = j < n ? mask[j * n_words + R] : 0ull;
    // lane l of the transpose: bit k = bit l of lane k's word, i.e. D[64R + k, 64C + l] <= r
    const uint64_t at = transpose_bits(a);
    // a diagonal block holds both orders of each pair: count bits k < l only
    const uint64_t keep = R == C ? ((1ull << lane) - 1ull) : ~0ull;
    const int diff = __popcll((at ^ b) & keep);
    const int total = wave_sum(diff);
    if (lane == 0 && total) atomicAdd(n_asym, (unsigned long long)total);
}

// One wave per row: the set bits of its words, in ascending column order.
__global__ __launch_bounds__(256) void k_pre_emit(const uint64_t *__restrict__ mask, int64_t n, int64_t n_words,
                                                  const int64_t *__restrict__ row_ptr, int32_t *__restrict__ col) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    int64_t pos = row_ptr[i];
    for (int64_t c = 0; c < n_words; ++c) {
        const uint64_t w = mask[i * n_words + c];
        if ((w >> lane) & 1ull) {
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(w >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)w, 0u));
            col[pos + rank] = (int32_t)(c * 64 + lane);
        }
        pos += __popcll(w);
    }
}

}  // namespace cnn

using namespace cnn;

extern "C" {

int cnn_precomputed_csr(cnn_ctx *ctx, const void *D, int64_t n, int32_t dtype, double r, int64_t *row_ptr,
                        int32_t *col, int64_t col_cap, int64_t *nnz_out, int64_t *asymmetric_out) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n < 0) throw InvalidArg{"negative n"};
        if (dtype != CNN_F32 && dtype != CNN_F64) throw InvalidArg{"dtype must be CNN_F32 or CNN_F64"};
        if (!row_ptr) throw InvalidArg{"null row_ptr"};
        if (n > 0 && !D) throw InvalidArg{"null distance matrix"};
        if (n >= (int64_t)INT32_MAX) throw InvalidArg{"n exceeds the int32 index range"};
        CNN_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        if (nnz_out) *nnz_out = 0;
        if (asymmetric_out) *asymmetric_out = 0;
        if (n == 0) {
            fill_i64(ctx, row_ptr, 0, 1);
            CNN_HIP(hipStreamSynchronize(s));
            return (int)CNN_OK;
        }
        const int64_t n_words = ceil_div(n, 64);
        Scratch<uint64_t> mask(ctx, n * n_words);
        Scratch<int32_t> deg(ctx, n);
        const unsigned wblocks = (unsigned)ceil_div(n, 4);
        if (dtype == CNN_F32)
            k_pre_masks<float><<<wblocks, 256, 0, s>>>((const float *)D, n, n_words, r, mask.p, deg.p);
        else
            k_pre_masks<double><<<wblocks, 256, 0, s>>>((const double *)D, n, n_words, r, mask.p, deg.p);
        CNN_HIP(hipGetLastError());
        exclusive_scan_i32_to_i64(ctx, deg.p, row_ptr, n);
        const int64_t last_excl = read_scalar(ctx, row_ptr + n - 1);
        const int32_t last_deg = read_scalar(ctx, deg.p + n - 1);
        const int64_t nnz = last_excl + last_deg;
        CNN_HIP(hipMemcpyAsync(row_ptr + n, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
        Scratch<unsigned long long> asym(ctx, 1);
        CNN_HIP(hipMemsetAsync(asym.p, 0, sizeof(unsigned long long), s));
        k_pre_symmetry<<<(unsigned)ceil_div(n_words * n_words, 4), 256, 0, s>>>(mask.p, n, n_words, asym.p);
        CNN_HIP(hipGetLastError());
        unsigned long long n_asym = 0;
        CNN_HIP(hipMemcpyAsync(&n_asym, asym.p, sizeof(n_asym), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipStreamSynchronize(s));
        if (nnz_out) *nnz_out = nnz;
        if (asymmetric_out) *asymmetric_out = (int64_t)n_asym;
        if (nnz > col_cap) throw Capacity{"col_cap " + std::to_string(col_cap) + " < nnz " + std::to_string(nnz)};
        if (nnz > 0) {
            if (!col) throw InvalidArg{"null col"};
            k_pre_emit<<<wblocks, 256, 0, s>>>(mask.p, n, n_words, row_ptr, col);
            CNN_HIP(hipGetLastError());
        }
        CNN_HIP(hipStreamSynchronize(s));
        return (int)CNN_OK;
    });
}

}  // extern "C"
