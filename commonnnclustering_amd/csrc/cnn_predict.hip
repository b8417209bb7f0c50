// Label prediction for a second point set: PredictorFirstmatch.predict
// (_fit.pyx:1241-1291) called by Clustering.predict (cluster.pyx:560-653).
//
// For every considered point p of Y, the reference points within r of it
// (NeighboursGetterExtBruteForce._get_other, _types.pyx:1712-1731: same metric,
// operands swapped, which leaves every squared difference unchanged) are
// visited in ascending reference index; p needs more than m_core of them.  The
// first member whose label is allowed, whose own neighbourhood among the
// reference points has more than m_core entries and with which p shares at
// least c_sim reference neighbours gives p its label and clears p's consider
// flag.  Other points of Y keep label and flag.
//
// Device procedure: one radius graph over the concatenation [X; Y] (same
// kernels and precision forms as a fit), then every list is cut to its
// reference points (N_X(m) for reference rows, the cross list for rows of Y),
// and one wave per row of Y walks its candidates in ascending reference index:
// each step takes the smallest index above the last one tested (a wave
// minimum) and runs the wave-wide merge test against that member's list.
#include <algorithm>
#include <memory>

#include "cnn_internal.hpp"

namespace cnn {

__device__ __forceinline__ int32_t caller_of(const int32_t *perm, int64_t w) {
    return perm ? perm[w] : (int32_t)w;
}

// Entries whose point is a reference point (caller index < n_x), per row:
// counted (FILL = false) or written in list order (FILL = true).
template <bool FILL>
__global__ __launch_bounds__(256) void k_keep_ref(const int64_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col, const int32_t *__restrict__ perm,
                                                  int64_t n, int64_t n_x, int32_t *__restrict__ cnt,
                                                  const int64_t *__restrict__ out_ptr, int32_t *__restrict__ out_col) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    int64_t pos = FILL ? out_ptr[w] : 0;
    int kept = 0;
    for (int64_t t0 = b; t0 < e; t0 += 64) {
        const int64_t t = t0 + lane;
        const int32_t j = t < e ? col[t] : 0;
        const bool keep = t < e && caller_of(perm, j) < n_x;
        const uint64_t m = __ballot(keep);
        if (FILL && keep) {
            const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            out_col[pos + rank] = j;
        }
        pos += __popcll(m);
        kept += __popcll(m);
    }
    if (!FILL && lane == 0) cnt[w] = kept;
}

// Per working row of a reference point: its label, and whether it may give
// its label (label allowed and more than m_core reference neighbours).
__global__ void k_ref_members(const int64_t *__restrict__ rp, const int32_t *__restrict__ perm, int64_t n,
                              int64_t n_x, int32_t m_core, const int64_t *__restrict__ labels_x,
                              const uint8_t *__restrict__ allowed, int64_t n_allowed, int64_t *__restrict__ lab,
                              uint8_t *__restrict__ ok) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int32_t o = caller_of(perm, w);
    bool good = false;
    int64_t l = 0;
    if (o < n_x) {
        l = labels_x[o];
        good = (rp[w + 1] - rp[w]) > (int64_t)m_core && l >= 0 && l < n_allowed && allowed[l] != 0;
    }
    lab[w] = l;
    ok[w] = good ? 1 : 0;
}

// One wave per working row of a point of Y.
__global__ __launch_bounds__(256) void k_predict_rows(const int64_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ perm, int64_t n, int64_t n_x,
                                                      int32_t c_sim, int32_t m_core, const int64_t *__restrict__ lab,
                                                      const uint8_t *__restrict__ ok, int64_t *__restrict__ labels_y,
                                                      uint8_t *__restrict__ consider_y,
                                                      unsigned long long *__restrict__ tests) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int32_t o = caller_of(perm, w);
    if (o < n_x) return;
    const int64_t q = (int64_t)o - n_x;
    if (!consider_y[q]) return;
    const int64_t b = rp[w], e = rp[w + 1];
    if (e - b <= (int64_t)m_core) return;
    unsigned n_tests = 0;
    int64_t last = -1;  // reference index of the last member tested
    for (;;) {
        // next member: the smallest reference index above `last` among the usable ones
        uint64_t best = ~0ull;  // (reference index << 32) | working row
        for (int64_t t0 = b; t0 < e; t0 += 64) {
            const int64_t t = t0 + lane;
            if (t < e) {
                const int32_t j = col[t];
                const int32_t cj = caller_of(perm, j);
                if (ok[j] && (int64_t)cj > last) {
                    const uint64_t key = ((uint64_t)(uint32_t)cj << 32) | (uint32_t)j;
                    best = key < best ? key : best;
                }
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const uint64_t other = __shfl_xor(best, off, 64);
            best = other < best ? other : best;
        }
        if (best == ~0ull) break;
        const int32_t jm = (int32_t)(uint32_t)best;
        ++n_tests;
        if (wave_sim(col + b, e - b, col + rp[jm], rp[jm + 1] - rp[jm], c_sim)) {
            if (lane == 0) {
                labels_y[q] = lab[jm];
                consider_y[q] = 0;
            }
            break;
        }
        last = (int64_t)(best >> 32);
    }
    if (lane == 0 && n_tests) atomicAdd(tests, (unsigned long long)n_tests);
}

}  // namespace cnn

using namespace cnn;

extern "C" {

int cnn_predict(cnn_ctx *ctx, const void *X, int64_t n_x, const void *Y, int64_t n_y, int32_t d, int32_t dtype,
                double r2, int32_t precision, int32_t c_sim, int32_t m_core, const int64_t *labels_x,
                const uint8_t *allowed, int64_t n_allowed, int64_t *labels_y, uint8_t *consider_y,
                int64_t *tests_out) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n_x < 0 || n_y < 0) throw InvalidArg{"negative size"};
        if (d < 1) throw InvalidArg{"d must be >= 1"};
        if (dtype != CNN_F32 && dtype != CNN_F64) throw InvalidArg{"dtype must be CNN_F32 or CNN_F64"};
        if (n_allowed < 0) throw InvalidArg{"negative n_allowed"};
        if (n_y > 0 && (!Y || !labels_y || !consider_y)) throw InvalidArg{"null buffer for Y"};
        if (n_x > 0 && (!X || !labels_x)) throw InvalidArg{"null buffer for X"};
        if (n_allowed > 0 && !allowed) throw InvalidArg{"null allowed-label table"};
        if (n_x + n_y >= (int64_t)INT32_MAX) throw InvalidArg{"n_x + n_y exceeds the int32 index range"};
        CNN_HIP(hipSetDevice(ctx->device));
        if (tests_out) *tests_out = 0;
        if (n_y == 0 || n_x == 0) return (int)CNN_OK;  // nothing to predict from
        hipStream_t s = ctx->stream;
        const size_t es = dtype == CNN_F32 ? sizeof(float) : sizeof(double);
        const int64_t n = n_x + n_y;
        // [X; Y] in one buffer: one graph holds N_X of both sets' points
        Scratch<uint8_t> Z(ctx, (size_t)n * d * es);
        CNN_HIP(hipMemcpyAsync(Z.p, X, (size_t)n_x * d * es, hipMemcpyDeviceToDevice, s));
        CNN_HIP(hipMemcpyAsync(Z.p + (size_t)n_x * d * es, Y, (size_t)n_y * d * es, hipMemcpyDeviceToDevice, s));
        cnn_graph *g = graph_build(ctx, Z.p, n, d, dtype, r2, precision, nullptr, true);
        std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
        graph_lists(ctx, g);
        const unsigned wblocks = (unsigned)ceil_div(n, 4);
        Scratch<int32_t> cnt(ctx, n);
        Scratch<int64_t> rp(ctx, n + 1);
        k_keep_ref<false><<<wblocks, 256, 0, s>>>(g->row_ptr, g->col, g->perm, n, n_x, cnt.p, nullptr, nullptr);
        CNN_HIP(hipGetLastError());
        exclusive_scan_i32_to_i64(ctx, cnt.p, rp.p, n);
        const int64_t last_excl = read_scalar(ctx, rp.p + n - 1);
        const int32_t last_cnt = read_scalar(ctx, cnt.p + n - 1);
        const int64_t nnz = last_excl + last_cnt;
        CNN_HIP(hipMemcpyAsync(rp.p + n, &nnz, sizeof(int64_t), hipMemcpyHostToDevice, s));
        Scratch<int32_t> col(ctx, std::max<int64_t>(nnz, 1));
        k_keep_ref<true><<<wblocks, 256, 0, s>>>(g->row_ptr, g->col, g->perm, n, n_x, nullptr, rp.p, col.p);
        CNN_HIP(hipGetLastError());
        Scratch<int64_t> lab(ctx, n);
        Scratch<uint8_t> ok(ctx, n);
        k_ref_members<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(rp.p, g->perm, n, n_x, m_core, labels_x, allowed,
                                                                 n_allowed, lab.p, ok.p);
        CNN_HIP(hipGetLastError());
        Scratch<unsigned long long> tests(ctx, 1);
        CNN_HIP(hipMemsetAsync(tests.p, 0, sizeof(unsigned long long), s));
        k_predict_rows<<<wblocks, 256, 0, s>>>(rp.p, col.p, g->perm, n, n_x, c_sim, m_core, lab.p, ok.p, labels_y,
                                               consider_y, tests.p);
        CNN_HIP(hipGetLastError());
        unsigned long long nt = 0;
        CNN_HIP(hipMemcpyAsync(&nt, tests.p, sizeof(nt), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipStreamSynchronize(s));
        if (tests_out) *tests_out = (int64_t)nt;
        return (int)CNN_OK;
    });
}

}  // extern "C"
