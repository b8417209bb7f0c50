// Phase N: radius neighbour graph.
//
// Replaces NeighboursGetterExtBruteForce._get (_types.pyx:1692-1710) with the
// reduced Euclidean metric (_types.pyx:2253-2267, r2 = r*r from :2286-2287).
//
// Layout in HBM (working order = spatial Morton order for d <= 8, identity
// otherwise):
//   xw64[n_pad][d]   FP64 coordinates in working order (n_pad = 64 * n_tiles)
//   xw32[n_pad][d]   FP32 coordinates (CNN_F32_BAND only)
//   tp_ptr[R], tp_col candidate column tiles of row tile R (box-culled; implicit
//                     "all tiles" for the dense layout)
//   mask[k][64]      one 64-bit word per (tile pair k, row): bit b set when
//                     column C*64+b is a neighbour of the row
//   row_ptr[n+1] (int64), col[nnz] (int32): CSR in working order, rows ascending
//
// The distance decision is exact: plain FP64 kernels evaluate
// t = a_k - b_k; s = s + t*t in ascending k with separately rounded operations
// (this file is compiled with -ffp-contract=off); the FP32 kernels decide pairs
// outside a guard band around r2 and recompute the band in that FP64 order.
#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <cfloat>
#include <cmath>
#include <memory>
#include <utility>

#include "cnn_internal.hpp"

namespace cnn {

// ------------------------------------------------------------------ prep kernels

template <typename Tin>
__global__ void k_bbox_partial(const Tin *__restrict__ X, int64_t n, int d, double *__restrict__ part,
                               double *__restrict__ part_norm) {
    // part[block][k][0..1] = min/max of finite values of dim k; part_norm[block] = max ||x||^2
    __shared__ double smin[256], smax[256];
    __shared__ double snorm[256];
    const int tid = threadIdx.x;
    double nmax = 0.0;
    for (int k = 0; k < d; ++k) {
        double lo = INFINITY, hi = -INFINITY;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + tid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            double v = (double)X[i * d + k];
            if (isfinite(v)) { lo = fmin(lo, v); hi = fmax(hi, v); }
        }
        smin[tid] = lo; smax[tid] = hi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) { smin[tid] = fmin(smin[tid], smin[tid + s]); smax[tid] = fmax(smax[tid], smax[tid + s]); }
            __syncthreads();
        }
        if (tid == 0) { part[((int64_t)blockIdx.x * d + k) * 2] = smin[0]; part[((int64_t)blockIdx.x * d + k) * 2 + 1] = smax[0]; }
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + tid; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int k = 0; k < d; ++k) { double v = (double)X[i * d + k]; s += v * v; }
        if (isfinite(s)) nmax = fmax(nmax, s);
    }
    snorm[tid] = nmax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) snorm[tid] = fmax(snorm[tid], snorm[tid + s]);
        __syncthreads();
    }
    if (tid == 0) part_norm[blockIdx.x] = snorm[0];
}

__device__ __forceinline__ uint64_t spread2(uint64_t x) {  // 32 bits -> even bits
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

__device__ __forceinline__ uint64_t spread3(uint64_t x) {  // 21 bits -> every third bit
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

struct KeySpec {
    int nk;         // dims in the key (1..3)
    int bits;       // bits per dimension (the key holds nk * bits)
    int dim[3];
    double lo[3];
    double scale[3];
    double qmax;
};

template <typename Tin, typename K>
__global__ void k_morton(const Tin *__restrict__ X, int64_t n, int d, KeySpec ks, K *__restrict__ keys,
                         int32_t *__restrict__ vals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t q[3] = {0, 0, 0};
    for (int t = 0; t < ks.nk; ++t) {
        double v = (double)X[i * d + ks.dim[t]];
        double f = (v - ks.lo[t]) * ks.scale[t];
        if (!(f >= 0.0)) f = 0.0;  // NaN / negative
        if (f > ks.qmax) f = ks.qmax;
        q[t] = (uint64_t)f;
    }
    uint64_t key;
    if (ks.nk == 1) key = q[0];
    else if (ks.nk == 2) key = spread2(q[0]) | (spread2(q[1]) << 1);
    else key = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    keys[i] = (K)key;  // K = uint32_t only when the key has at most 32 bits
    vals[i] = (int32_t)i;
}

// Working order for d > 8 (all tile pairs are evaluated, in any order): points
// grouped by their nearest of L pivots (the points at caller indices j*n/L),
// measured in FP32 over the first min(d, 64) dimensions, the cells themselves
// grouped by the nearest of S super-pivots (pivots s * L / S) so that the
// cells of one cluster are adjacent.  Only the order comes from this;
// clustered data then fills tiles from one cluster each, so most tile pairs
// lie entirely outside the radius, and the exact pivot-bound test below
// (k_dense_live) removes them before the distance kernels run.  Key = (super
// cell, pivot index); the sort is stable (caller order within a cell).
constexpr int kPivotDims = 64;
constexpr int kSuperPivots = 16;

typedef float v16f __attribute__((ext_vector_type(16)));  // a 32x32 FP32 MFMA accumulator

// super_of[p] = nearest super-pivot of pivot p (one thread per pivot, L <= 256).
template <typename Tin>
__global__ __launch_bounds__(256) void k_pivot_super(const Tin *__restrict__ X, int64_t n, int d, int L, int S,
                                                     int32_t *__restrict__ super_of) {
    __shared__ float ss[kSuperPivots][kPivotDims];
    const int dd = min(d, kPivotDims);
    for (int q = threadIdx.x; q < S * kPivotDims; q += 256) {
        const int sp = q / kPivotDims, k = q % kPivotDims;
        const int64_t row = ((int64_t)(sp * L / S)) * n / L;
        ss[sp][k] = k < dd ? (float)X[row * d + k] : 0.0f;
    }
    __syncthreads();
    const int p = threadIdx.x;
    if (p >= L) return;
    float x[kPivotDims];
    const int64_t row = (int64_t)p * n / L;
#pragma unroll
    for (int k = 0; k < kPivotDims; ++k) x[k] = k < dd ? (float)X[row * d + k] : 0.0f;
    float best = INFINITY;
    int bs = 0;
    for (int sp = 0; sp < S; ++sp) {
        float acc = 0.0f;  // dimensions past dd are zero on both sides
#pragma unroll
        for (int k = 0; k < kPivotDims; ++k) {
            const float t = x[k] - ss[sp][k];
            acc = __builtin_fmaf(t, t, acc);
        }
        if (acc < best) {
            best = acc;
            bs = sp;
        }
    }
    super_of[p] = bs;
}

// Pivot table of the nearest-pivot order: Lp >= L rows of 64 floats (the
// pivot's first min(d, 64) coordinates, zero-padded) and |p|^2 / 2 per row
// (+inf for rows past L, which are never nearest).  One wave per row.
template <typename Tin>
__global__ __launch_bounds__(64) void k_pivot_table(const Tin *__restrict__ X, int64_t n, int d, int L,
                                                    float *__restrict__ piv, float *__restrict__ phalf) {
    const int lane = threadIdx.x & 63;
    const int64_t j = blockIdx.x;
    const int dd = min(d, kPivotDims);
    const float v = (j < L && lane < dd) ? (float)X[(j * n / L) * d + lane] : 0.0f;
    piv[j * kPivotDims + lane] = v;
    float s = v * v;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) phalf[j] = j < L ? 0.5f * s : INFINITY;
}

// Nearest-pivot cells on MFMA: per wave 32 points against the pivots in blocks
// of 32, x.p accumulated in FP32 by v_mfma_f32_32x32x2f32 over the first
// min(d, 64) dimensions (lane l supplies point l%32 and pivot l%32 at
// dimension 32(l/32) + s of step s: each lane reads 32 contiguous floats,
// float4 at a time), and the cell is argmin_p |p|^2/2 - x.p, ties to the
// smaller pivot.  Only the working order depends on the cells.
template <typename Tin, typename K>
__global__ __launch_bounds__(256) void k_pivot_cells_mfma(const Tin *__restrict__ X, int64_t n, int d, int L,
                                                          const float *__restrict__ piv,
                                                          const float *__restrict__ phalf,
                                                          const int32_t *__restrict__ super_of,
                                                          K *__restrict__ keys, int32_t *__restrict__ vals) {
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    const int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    if (base >= n) return;
    const int dd = min(d, kPivotDims);
    const int64_t i = base + r;
    const int kb = (kPivotDims / 2) * h;  // this lane's dimensions: kb .. kb + 31
    float a[kPivotDims / 2];
    if (std::is_same<Tin, float>::value && (d & 3) == 0 && ((uintptr_t)X & 15) == 0) {
        // rows start 16-byte aligned: float4 loads (dd is a multiple of 4 as well)
        const float4 *xr = (const float4 *)((const float *)X + i * d + kb);
#pragma unroll
        for (int q = 0; q < kPivotDims / 8; ++q) {
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i < n && kb + 4 * q < dd) v = xr[q];
            a[4 * q] = v.x;
            a[4 * q + 1] = v.y;
            a[4 * q + 2] = v.z;
            a[4 * q + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int s = 0; s < kPivotDims / 2; ++s) {
            const int k = kb + s;
            a[s] = (i < n && k < dd) ? (float)X[i * d + k] : 0.0f;
        }
    }
    float best[16];
    int bidx[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        best[v] = INFINITY;
        bidx[v] = 0;
    }
    const int nb = (L + 31) / 32;
    for (int blk = 0; blk < nb; ++blk) {
        const int j = blk * 32 + r;
        const float4 *__restrict__ pj = (const float4 *)(piv + (int64_t)j * kPivotDims + kb);
        float b[kPivotDims / 2];
#pragma unroll
        for (int q = 0; q < kPivotDims / 8; ++q) {
            const float4 v = pj[q];
            b[4 * q] = v.x;
            b[4 * q + 1] = v.y;
            b[4 * q + 2] = v.z;
            b[4 * q + 3] = v.w;
        }
        const float hn = phalf[j];
        v16f c = v16f{0};
#pragma unroll
        for (int s = 0; s < kPivotDims / 2; ++s) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], b[s], c, 0, 0, 0);
        // element v of lane l: point 8(v/4) + 4h + v%4, pivot blk*32 + l%32
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const float sc = hn - c[v];
            if (sc < best[v]) {
                best[v] = sc;
                bidx[v] = j;
            }
        }
    }
    // the nearest over the 32 lanes that hold the same points
#pragma unroll
    for (int v = 0; v < 16; ++v) {
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) {
            const float ob = __shfl_xor(best[v], off, 64);
            const int oi = __shfl_xor(bidx[v], off, 64);
            if (ob < best[v] || (ob == best[v] && oi < bidx[v])) {
                best[v] = ob;
                bidx[v] = oi;
            }
        }
    }
    if (r == 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int64_t ii = base + 8 * (v >> 2) + 4 * h + (v & 3);
            if (ii < n) {
                keys[ii] = (K)(((uint32_t)super_of[bidx[v]] << 8) | (uint32_t)bidx[v]);
                vals[ii] = (int32_t)ii;
            }
        }
    }
}

template <typename Tin>
__global__ void k_gather(const Tin *__restrict__ X, const int32_t *__restrict__ perm, int64_t n, int64_t n_pad,
                         int d, double *__restrict__ xw64, float *__restrict__ xw32) {
    const int64_t total = n_pad * d;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        int64_t w = idx / d;
        int k = (int)(idx - w * d);
        if (w < n) {
            int64_t src = perm ? (int64_t)perm[w] : w;
            Tin v = X[src * d + k];
            xw64[idx] = (double)v;
            if (xw32) xw32[idx] = (float)v;
        } else {
            xw64[idx] = 0.0;
            if (xw32) xw32[idx] = 0.0f;
        }
    }
}

// Exact culling of dense tile pairs.  For kCullPivots pivots p (the centroids
// of tiles spread evenly over the working order: each tile holds points of one
// nearest-pivot cell, so its centroid sits near a cluster's centre, where
// distances to the cluster's points vary least) every tile keeps the range
// [lo_p, hi_p] of its real
// points' Euclidean distances to p; by the triangle inequality no point of R
// is within r of a point of C when lo_p(C) - hi_p(R) > r or lo_p(R) - hi_p(C) > r
// for some p.  Distances are FP64 (relative error below (d + 3) * 2^-53) and
// stored in FP32 rounded outward; the test adds a relative margin of 1e-6,
// which covers both the rounding of the stored bounds and of the FP32
// subtraction, and the reference's own rounding of the squared distance, so a
// culled pair never holds a pair with s_ij <= r^2.  A tile with a real point
// whose distance to p is not finite (non-finite or huge coordinates) gets the
// range [0, inf] for p, which culls nothing through p.
constexpr int kCullPivots = 64;
constexpr int kCullDims = 32;  // dimensions staged per step

// Pivot p = FP64 mean of the real rows of tile p * T / kCullPivots (one
// workgroup per pivot, lanes over dimensions).  Any point serves as a pivot:
// the bounds below hold for every p.
__global__ __launch_bounds__(64) void k_cull_pivots(const double *__restrict__ xw64, int64_t n, int d, int64_t T,
                                                    double *__restrict__ piv) {
    const int64_t t = (int64_t)blockIdx.x * T / kCullPivots;
    const int rows = (int)min<int64_t>(64, n - t * 64);
    for (int k = threadIdx.x; k < d; k += 64) {
        double sum = 0.0;
        for (int r = 0; r < rows; ++r) sum += xw64[(t * 64 + r) * d + k];
        piv[(int64_t)blockIdx.x * d + k] = sum / (double)rows;
    }
}

// home[R] = the pivot with the smallest hi for tile R: the one nearest to all of
// its points, tried first by the pair test.
__global__ __launch_bounds__(256) void k_tile_pivot_bounds(const double *__restrict__ xw64, int64_t n, int d,
                                                           int64_t T, const double *__restrict__ piv,
                                                           float2 *__restrict__ bnd, int32_t *__restrict__ home) {
    __shared__ float s_best[4];
    __shared__ int s_bp[4];
    static_assert(kCullPivots % 4 == 0, "pivots per wave");
    constexpr int kPerWave = kCullPivots / 4;
    __shared__ double sX[kCullDims][65];            // the tile's rows, dimension-major (padded)
    __shared__ double sP[kCullPivots][kCullDims];  // the pivots
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t R = blockIdx.x;
    double acc[kPerWave];
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) acc[j] = 0.0;
    for (int k0 = 0; k0 < d; k0 += kCullDims) {
        const int kc = min(kCullDims, d - k0);
        __syncthreads();
        for (int q = threadIdx.x; q < 64 * kc; q += 256) {
            const int r = q / kc, kk = q - r * kc;
            sX[kk][r] = xw64[(R * 64 + r) * d + k0 + kk];
        }
        for (int q = threadIdx.x; q < kCullPivots * kc; q += 256) {
            const int p = q / kc, kk = q - p * kc;
            sP[p][kk] = piv[(int64_t)p * d + k0 + kk];
        }
        __syncthreads();
        for (int kk = 0; kk < kc; ++kk) {
            const double x = sX[kk][lane];
#pragma unroll
            for (int j = 0; j < kPerWave; ++j) {
                const double t = x - sP[wave + 4 * j][kk];
                acc[j] = __builtin_fma(t, t, acc[j]);
            }
        }
    }
    const bool real = R * 64 + lane < n;
    float best = INFINITY;
    int bp = wave;
#pragma unroll
    for (int j = 0; j < kPerWave; ++j) {
        const double dist = sqrt(acc[j]);
        const bool fin = dist < INFINITY;  // false for +inf and NaN
        double lo = (real && fin) ? dist : INFINITY, hi = (real && fin) ? dist : 0.0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, off, 64));
            hi = fmax(hi, __shfl_xor(hi, off, 64));
        }
        const bool bad = __ballot(real && !fin) != 0ull;
        const float2 b = bad ? make_float2(0.0f, INFINITY) : make_float2(__double2float_rd(lo), __double2float_ru(hi));
        if (b.y < best) {
            best = b.y;
            bp = wave + 4 * j;
        }
        if (lane == 0) bnd[(int64_t)(wave + 4 * j) * T + R] = b;
    }
    if (lane == 0) {
        s_best[wave] = best;
        s_bp[wave] = bp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int h = s_bp[0];
        float bb = s_best[0];
        for (int w = 1; w < 4; ++w)
            if (s_best[w] < bb) {
                bb = s_best[w];
                h = s_bp[w];
            }
        home[R] = h;
    }
}

// Live bits of the dense tile pairs: one wave per (row tile R, 64 column tiles);
// counts[0] = live pairs, counts[1] = live pairs with C >= R.
__device__ __forceinline__ bool pivot_cuts(const float2 *__restrict__ bnd, int64_t T, int p, int64_t R, int64_t C,
                                           float rc) {
    const float2 br = bnd[(int64_t)p * T + R], bc = bnd[(int64_t)p * T + C];
    const float lb = fmaxf(bc.x - br.y, br.x - bc.y);
    return lb > rc + 1e-6f * (br.y + bc.y);
}

// The home pivots of R and C are tried first (they cut most pairs of
// different clusters), then every pivot.  The test is symmetric in R and C,
// so only the words of the 64 x 64 blocks on or right of the diagonal block
// are evaluated here (k_live_mirror fills the others); counts[1] = live pairs
// with C >= R.
__global__ __launch_bounds__(256) void k_dense_live(const float2 *__restrict__ bnd, const int32_t *__restrict__ home,
                                                    int64_t T, int64_t W, float rc, uint64_t *__restrict__ live,
                                                    unsigned long long *__restrict__ counts) {
    const int lane = threadIdx.x & 63;
    unsigned long long n_up = 0;  // this wave's count: one atomic at the end
    for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < T * W; gw += (int64_t)gridDim.x * 4) {
        const int64_t R = gw / W, cw = gw - R * W, C = cw * 64 + lane;
        if (cw < (R >> 6)) continue;  // below the diagonal block (wave-uniform)
        bool cut = false;
        if (C < T && C != R) {
            cut = pivot_cuts(bnd, T, home[R], R, C, rc) || pivot_cuts(bnd, T, home[C], R, C, rc);
            for (int p = 0; p < kCullPivots && !cut; ++p) cut = pivot_cuts(bnd, T, p, R, C, rc);
        }
        const bool lv = C < T && !cut;
        const uint64_t w = __ballot(lv);
        n_up += __popcll(__ballot(lv && C >= R));
        if (lane == 0) live[gw] = w;
    }
    if (lane == 0 && n_up) atomicAdd(&counts[1], n_up);
}

// The words below the diagonal blocks: block (a, b), b < a (rows 64a.., word
// b) is the bit transpose of block (b, a).  One wave per block.
__global__ __launch_bounds__(256) void k_live_mirror(uint64_t *__restrict__ live, int64_t T, int64_t W) {
    const int lane = threadIdx.x & 63;
    const int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (blk >= W * W) return;
    const int64_t a = blk / W, b = blk - a * W;
    if (b >= a) return;
    const int64_t C = b * 64 + lane;  // row of block (b, a)
    const uint64_t src = C < T ? live[C * W + a] : 0ull;
    const uint64_t t = transpose_bits_shfl(src);  // lane l: row 64a + l over columns 64b..
    const int64_t R = a * 64 + lane;
    if (R < T) live[R * W + b] = t;
}

// Per row tile (one wave): live pairs in all, and per word the live bits of
// the row's earlier words.
__global__ __launch_bounds__(256) void k_live_rows(const uint64_t *__restrict__ live, int64_t T, int64_t W,
                                                   int32_t *__restrict__ pre, int32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t R = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (R >= T) return;
    int run = 0;
    for (int64_t w0 = 0; w0 < W; w0 += 64) {
        const int64_t w = w0 + lane;
        const int c = w < W ? __popcll(live[R * W + w]) : 0;
        const int incl = wave_incl_scan(c);
        if (w < W) pre[R * W + w] = run + incl - c;
        run += __shfl(incl, 63, 64);
    }
    if (lane == 0) cnt[R] = run;
}

// Explicit lists of the live pairs (column tiles ascending) and the mirror of
// every pair: the index of (C, R) in C's list.
__global__ __launch_bounds__(256) void k_live_lists(const uint64_t *__restrict__ live, int64_t T, int64_t W,
                                                    const int32_t *__restrict__ pre, const int64_t *__restrict__ tp_ptr,
                                                    int32_t *__restrict__ tp_col, int64_t *__restrict__ mirror) {
    const int lane = threadIdx.x & 63;
    const int64_t R = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (R >= T) return;
    const int64_t rw = R >> 6;
    const uint64_t rbelow = (1ull << (R & 63)) - 1ull;
    for (int64_t w = lane; w < W; w += 64) {
        uint64_t bits = live[R * W + w];
        int64_t k = tp_ptr[R] + pre[R * W + w];
        while (bits) {
            const int b = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            const int64_t C = w * 64 + b;
            tp_col[k] = (int32_t)C;
            mirror[k] = tp_ptr[C] + pre[C * W + rw] + __popcll(live[C * W + rw] & rbelow);
            ++k;
        }
    }
}

#ifndef CNN_ROW_GROUPS
#define CNN_ROW_GROUPS 8
#endif
// Row groups of a tile for the culling of candidate tile pairs: kRowGroups
// groups of kGroupRows consecutive rows (a Morton-ordered tile's groups are
// sub-boxes of its box); per candidate pair a bit mask of the groups of R
// within reach of tile C.  Finer groups skip more rows of a pair.
constexpr int kRowGroups = CNN_ROW_GROUPS;
constexpr int kGroupRows = 64 / kRowGroups;
static_assert(kRowGroups >= 1 && kRowGroups <= 16 && 64 % kRowGroups == 0, "row groups: 1..16, dividing 64");
using GroupMask = uint16_t;

// Per-tile bounding boxes over all d dims (finite coordinates of valid rows).
// Also the boxes of the tile's row groups (rows kGroupRows*g ..), used to
// skip row groups of a candidate tile pair (glo/ghi nullable).
__global__ void k_tile_boxes(const double *__restrict__ xw64, int64_t n, int d, int64_t n_tiles,
                             double *__restrict__ blo, double *__restrict__ bhi, double *__restrict__ glo,
                             double *__restrict__ ghi) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (t >= n_tiles) return;
    const int64_t i = t * 64 + lane;
    for (int k = 0; k < d; ++k) {
        double v = (i < n) ? xw64[i * d + k] : NAN;
        double lo = isfinite(v) ? v : INFINITY;
        double hi = isfinite(v) ? v : -INFINITY;
#pragma unroll
        for (int off = 1; off < kGroupRows; off <<= 1) {
            lo = fmin(lo, __shfl_xor(lo, off, 64));
            hi = fmax(hi, __shfl_xor(hi, off, 64));
        }
        if (glo && (lane % kGroupRows) == 0) {
            glo[(t * kRowGroups + lane / kGroupRows) * d + k] = lo;
            ghi[(t * kRowGroups + lane / kGroupRows) * d + k] = hi;
        }
#pragma unroll
        for (int off = kGroupRows; off < 64; off <<= 1) {
            lo = fmin(lo, __shfl_xor(lo, off, 64));
            hi = fmax(hi, __shfl_xor(hi, off, 64));
        }
        if (lane == 0) { blo[t * d + k] = lo; bhi[t * d + k] = hi; }
    }
}

// Boxes of super-tiles (64 consecutive tiles = 4096 points of the working order).
__global__ void k_super_boxes(const double *__restrict__ blo, const double *__restrict__ bhi, int d, int64_t n_tiles,
                              int64_t n_super, double *__restrict__ slo, double *__restrict__ shi) {
    const int lane = threadIdx.x & 63;
    const int64_t S = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (S >= n_super) return;
    const int64_t t = S * 64 + lane;
    for (int k = 0; k < d; ++k) {
        double lo = (t < n_tiles) ? blo[t * d + k] : INFINITY;
        double hi = (t < n_tiles) ? bhi[t * d + k] : -INFINITY;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, off, 64));
            hi = fmax(hi, __shfl_xor(hi, off, 64));
        }
        if (lane == 0) { slo[S * d + k] = lo; shi[S * d + k] = hi; }
    }
}

__device__ __forceinline__ bool boxes_close(const double *__restrict__ alo, const double *__restrict__ ahi,
                                            const double *__restrict__ blo, const double *__restrict__ bhi, int d,
                                            double thr) {
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
        double g = fmax(blo[k] - ahi[k], alo[k] - bhi[k]);
        g = fmax(g, 0.0);
        s += g * g;
    }
    return s <= thr;
}

// Box test against R's box held in registers (d <= kSpatialMaxD): all of the
// other box's coordinates are loaded before any is used.
struct RowBox {
    double lo[kSpatialMaxD], hi[kSpatialMaxD];
    int d;
    double thr;
    __device__ __forceinline__ bool close(const double *__restrict__ olo, const double *__restrict__ ohi) const {
        double l[kSpatialMaxD], h[kSpatialMaxD];
#pragma unroll
        for (int k = 0; k < kSpatialMaxD; ++k) {
            l[k] = k < d ? olo[k] : 0.0;
            h[k] = k < d ? ohi[k] : 0.0;
        }
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < kSpatialMaxD; ++k) {
            double g = fmax(l[k] - hi[k], lo[k] - h[k]);
            g = fmax(g, 0.0);
            s += k < d ? g * g : 0.0;
        }
        return s <= thr;
    }
};

// Candidate column tiles of row tile R, by one wave: super-tiles are culled
// first (lane per super-tile), then the 64 tiles of every surviving super-tile
// (lane per tile), in ascending order, kSuperPerStep surviving super-tiles
// per step (their loads in flight together).  Each candidate C at list
// position pos (0-based) is handed to emit(pos, C); returns the count.
constexpr int kSuperPerStep = 8;
template <typename Emit>
__device__ __forceinline__ int64_t tile_pair_search(const double *__restrict__ blo, const double *__restrict__ bhi,
                                                    const double *__restrict__ slo, const double *__restrict__ shi,
                                                    int d, int64_t n_tiles, int64_t n_super, double thr_all,
                                                    const double *__restrict__ tile_thr,
                                                    const int32_t *__restrict__ seg_lo,
                                                    const int32_t *__restrict__ seg_hi, int64_t R, Emit emit) {
    const int lane = threadIdx.x & 63;
    RowBox rb;
    rb.d = d;
    rb.thr = tile_thr ? tile_thr[R] : thr_all;
#pragma unroll
    for (int k = 0; k < kSpatialMaxD; ++k) {
        rb.lo[k] = k < d ? blo[R * d + k] : 0.0;
        rb.hi[k] = k < d ? bhi[R * d + k] : 0.0;
    }
    const int64_t clo = seg_lo ? seg_lo[R] : 0, chi = seg_hi ? seg_hi[R] : n_tiles;
    int64_t out = 0;
    for (int64_t S0 = clo / 64; S0 * 64 < chi; S0 += 64) {
        const int64_t Sl = S0 + lane;
        const bool sok = (Sl < n_super) && (Sl * 64 < chi) && rb.close(slo + Sl * d, shi + Sl * d);
        uint64_t sbits = __ballot(sok);
        while (sbits) {
            int64_t Sq[kSuperPerStep];
            int m = 0;
            while (sbits && m < kSuperPerStep) {
                Sq[m++] = S0 + __ffsll((unsigned long long)sbits) - 1;
                sbits &= sbits - 1;
            }
            bool ok[kSuperPerStep];
#pragma unroll
            for (int q = 0; q < kSuperPerStep; ++q) {
                const int64_t C = (q < m ? Sq[q] : 0) * 64 + lane;
                ok[q] = q < m && C >= clo && C < chi && rb.close(blo + C * d, bhi + C * d);
            }
#pragma unroll
            for (int q = 0; q < kSuperPerStep; ++q) {
                if (q >= m) break;
                const int64_t C = Sq[q] * 64 + lane;
                const uint64_t b = __ballot(ok[q]);
                if (ok[q]) emit(out + __popcll(b & ((1ull << lane) - 1ull)), C);
                out += __popcll(b);
            }
        }
    }
    return out;
}

// List entries a row tile keeps in the first search pass: at most
// kPairBucketMax, fewer when T * kPairBucketMax * 4 bytes would pass
// kPairBucketBytes (row tiles whose lists do not fit search again).
constexpr int kPairBucketMax = 512;
constexpr int64_t kPairBucketBytes = int64_t(512) << 20;

// First search pass: the count of every row tile, and its list in its bucket
// of B entries (complete when the count fits) with diag[R], R's place in its
// own list.
__global__ void k_tile_pairs(const double *__restrict__ blo, const double *__restrict__ bhi,
                             const double *__restrict__ slo, const double *__restrict__ shi, int d, int64_t n_tiles,
                             int64_t n_super, double thr_all, const double *__restrict__ tile_thr,
                             const int32_t *__restrict__ seg_lo, const int32_t *__restrict__ seg_hi,
                             int32_t *__restrict__ cnt, int32_t *__restrict__ bucket, int B,
                             int32_t *__restrict__ diag) {
    const int64_t R = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (R >= n_tiles) return;
    int32_t *mine = bucket + R * B;
    const int64_t total = tile_pair_search(blo, bhi, slo, shi, d, n_tiles, n_super, thr_all, tile_thr, seg_lo, seg_hi,
                                           R, [&](int64_t pos, int64_t C) {
                                               if (pos < B) mine[pos] = (int32_t)C;
                                               if (C == R) diag[R] = (int32_t)pos;
                                           });
    if ((threadIdx.x & 63) == 0) cnt[R] = (int32_t)total;
}

// Second pass: each row tile's list moves from its bucket to its place in
// tp_col; a row tile whose list overflowed its bucket searches again, writing
// in place.
__global__ void k_tile_pairs_place(const double *__restrict__ blo, const double *__restrict__ bhi,
                                   const double *__restrict__ slo, const double *__restrict__ shi, int d,
                                   int64_t n_tiles, int64_t n_super, double thr_all,
                                   const double *__restrict__ tile_thr, const int32_t *__restrict__ seg_lo,
                                   const int32_t *__restrict__ seg_hi, const int32_t *__restrict__ cnt,
                                   const int32_t *__restrict__ bucket, int B, const int64_t *__restrict__ tp_ptr,
                                   int32_t *__restrict__ tp_col) {
    const int lane = threadIdx.x & 63;
    const int64_t R = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (R >= n_tiles) return;
    const int64_t c = cnt[R], base = tp_ptr[R];
    if (c <= B) {
        // 512 entries a step, their loads in flight together
        for (int64_t t0 = 0; t0 < c; t0 += 512) {
            int32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t t = t0 + 64 * u + lane;
                v[u] = t < c ? bucket[R * B + t] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int64_t t = t0 + 64 * u + lane;
                if (t < c) tp_col[base + t] = v[u];
            }
        }
        return;
    }
    (void)tile_pair_search(blo, bhi, slo, shi, d, n_tiles, n_super, thr_all, tile_thr, seg_lo, seg_hi, R,
                           [&](int64_t pos, int64_t C) { tp_col[base + pos] = (int32_t)C; });
}

// ------------------------------------------------------------------ distance kernels

struct BandParams {
    double r2;
    float lo;    // FP32 d2 < lo: neighbour for certain
    float hi;    // FP32 d2 > hi: not a neighbour for certain; otherwise recompute in FP64
};

struct DistArgs {
    const double *xw64;
    const float *xw32;
    int64_t n;
    int d;
    int64_t R0;              // first row tile of this batch
    int64_t n_tiles;
    const int64_t *tp_ptr;   // explicit lists (nullptr for dense)
    const int32_t *tp_col;
    int64_t k_base;          // mask index of the batch's first tile pair
    uint64_t *mask;
    int32_t *deg;
    BandParams bp;
    unsigned long long *band_stats;  // [0] hits, [1] flips, [2] max err bits (double as u64)
    int dp;                  // padded width of the augmented Gram vectors (8 per group)
    int mfma_k;              // MFMA steps (two dimensions each): ceil((d + 2) / 2)
    // symmetric evaluation: only tile pairs with C >= R are computed; the
    // transposed words go to the mirror pair (C, R)
    int sym;
    const int64_t *mirror;   // explicit lists: index of (C, R) for each k (nullptr for dense)
    // augmented Gram form: row-role [x, 1, c] and column-role [x, c, 1] vectors
    // whose dot product is T = (r2 - d2)/2, stored in the packed MFMA operand
    // layout (see packed_at); e = error margin per point
    const float *xra;
    const float *xca;
    const float *emarg;
    // column chunking of the dense wide FP64 kernel (tiles per workgroup)
    int64_t col_chunk;
    // packed batches: radius and FP32 band per row tile (nullptr: bp for all)
    const double *tile_r2;
    const float *tile_lo;
    const float *tile_hi;
    const int32_t *tile_valid;  // packed batches: real rows per tile (the first ones); nullptr: rows < n
    // MFMA form: tile pairs with guard-band elements, settled by a second
    // kernel -- a record (k, R, 64 per-lane band flags) per pair while the list
    // has room, else the pair alone (k, R) in the redo list
    unsigned long long *band_cnt;
    int64_t *band_list;         // 66 int64 per record
    int64_t band_cap;           // records
    unsigned long long *redo_cnt;
    int64_t *redo;              // (k, R) per entry
    uint8_t *pair_nz;           // per pair of the batch: 1 if its words hold any bit (nullable)
    const GroupMask *pair_groups; // per pair (global index): row groups of R in reach of C
    const double *bounds;       // periodic metric: period per dimension (<= 0: not periodic); nullptr: none
    // dense layout with pivot culling: bit C % 64 of live[R * live_words + C / 64]
    // is set when tile pair (R, C) may hold a neighbour (nullptr: every pair)
    // (the live pairs are also the explicit lists tp_ptr / tp_col; live_pre
    // holds, per row tile and word, the live bits of the row's earlier words)
    const uint64_t *live;
    const int32_t *live_pre;
    int64_t live_words;
    int deal_rr;  // dense MFMA kernel: work items dealt round-robin over the XCDs
    // spatial layout, one batch: the row tile of workgroup b is row_order[b]
    // (heavy tiles first), nullptr: R0 + b
    const int32_t *row_order;
};

// Row tile of a workgroup of the row-tile kernels.
__device__ __forceinline__ int64_t block_row_tile(const DistArgs &a) {
    return a.row_order ? (int64_t)a.row_order[blockIdx.x] : a.R0 + blockIdx.x;
}

__device__ __forceinline__ bool pair_live(const DistArgs &a, int64_t R, int64_t C) {
    return !a.live || ((a.live[R * a.live_words + (C >> 6)] >> (C & 63)) & 1ull);
}

// Mask index of tile pair (R, C) for kernels that walk column ranges: all pairs
// (R * T + C), or the position of C in R's list of live pairs.
__device__ __forceinline__ int64_t pair_index(const DistArgs &a, int64_t R, int64_t C) {
    if (!a.live) return R * a.n_tiles + C;
    const int64_t w = R * a.live_words + (C >> 6);
    const uint64_t below = a.live[w] & ((1ull << (C & 63)) - 1ull);
    return a.tp_ptr[R] + a.live_pre[w] + __popcll(below);
}

__device__ __forceinline__ double r2_of(const DistArgs &a, int64_t R) { return a.tile_r2 ? a.tile_r2[R] : a.bp.r2; }

__device__ __forceinline__ void tile_range(const DistArgs &a, int64_t R, int64_t &kb, int64_t &ke) {
    if (a.tp_ptr) {
        kb = a.tp_ptr[R];
        ke = a.tp_ptr[R + 1];
    } else {
        kb = R * a.n_tiles;
        ke = kb + a.n_tiles;
    }
}

__device__ __forceinline__ int64_t tile_col(const DistArgs &a, int64_t k, int64_t kb) {
    return a.tp_ptr ? (int64_t)a.tp_col[k] : (k - kb);
}

__device__ __forceinline__ uint64_t valid_cols(int64_t n, int64_t C) {
    int64_t v = n - C * 64;
    if (v >= 64) return ~0ull;
    if (v <= 0) return 0ull;
    return (1ull << v) - 1ull;
}

// Real rows of tile C as a bit mask, and whether working row i is real.
__device__ __forceinline__ uint64_t tile_mask(const DistArgs &a, int64_t C) {
    if (!a.tile_valid) return valid_cols(a.n, C);
    const int v = a.tile_valid[C];
    return v >= 64 ? ~0ull : ((1ull << v) - 1ull);
}

__device__ __forceinline__ bool row_real(const DistArgs &a, int64_t i) {
    return a.tile_valid ? (int)(i & 63) < a.tile_valid[i >> 6] : i < a.n;
}

// Exact FP64 reduced distance in reference order (rows of the working layout).
__device__ __forceinline__ double exact_d2(const double *__restrict__ xw64, int d, int64_t i, int64_t j) {
    const double *a = xw64 + i * d;
    const double *b = xw64 + j * d;
    double s = 0.0;
    for (int k = 0; k < d; ++k) {
        double t = a[k] - b[k];
        s = s + t * t;
    }
    return s;
}

// Exact FP64 reduced distance in reference order, computed by the whole wave:
// lane k forms the rounded square of dimension k's rounded difference, then the
// squares are added one after another in ascending k (a chain of readlanes),
// exactly the reference's sequence.  Wave-uniform result; all lanes active.
__device__ __forceinline__ double exact_d2_wave(const double *__restrict__ xw64, int d, int64_t gi, int64_t gj) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int k0 = 0; k0 < d; k0 += 64) {
        const int k = k0 + lane;
        double sq = 0.0;
        if (k < d) {
            const double t = xw64[gi * d + k] - xw64[gj * d + k];
            sq = t * t;
        }
        const int m = min(64, d - k0);
        const uint64_t bits = __double_as_longlong(sq);
        for (int q = 0; q < m; ++q) {
            const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)bits, q);
            const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(bits >> 32), q);
            s = s + __longlong_as_double(((int64_t)hi << 32) | lo);
        }
    }
    return s;
}

__device__ __forceinline__ void atomic_max_double(unsigned long long *p, double v) {
    // v >= 0: the IEEE bit pattern orders like the value
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// First tile pair of row tile R to evaluate: all pairs, or those with C >= R.
__device__ __forceinline__ int64_t first_pair(const DistArgs &a, int64_t R, int64_t kb, int64_t ke) {
    if (!a.sym) return kb;
    if (!a.tp_ptr) return kb + R;
    int64_t lo = kb, hi = ke;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a.tp_col[mid] < R) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Store a row word (and, for an off-diagonal symmetric pair, the column word
// of the mirror pair).  Returns the row word's popcount.  With per-pair flags,
// the words of a pair without any bit are not stored: readers skip flagged-
// empty pairs (the mirror of an empty pair is empty too).  `force` stores them
// anyway, for pairs whose guard-band bits a later kernel ors in.
// Mask index of the mirror pair (C, R) of pair k = (R, C).
__device__ __forceinline__ int64_t mirror_of(const DistArgs &a, int64_t k, int64_t R, int64_t C) {
    return a.tp_ptr ? a.mirror[k] : C * a.n_tiles + R;
}

// km: the mirror's index when the caller has it (loaded ahead), else -1.
// LOW_REGS: transpose with transpose_bits_shfl (callers at the edge of an
// occupancy step).
template <bool LOW_REGS = false>
__device__ __forceinline__ int store_pair(const DistArgs &a, int64_t k, int64_t R, int64_t i, int64_t C, uint64_t m,
                                          bool have_col, uint64_t colword, bool force = false, int64_t km = -1) {
    m &= tile_mask(a, C);
    if (!row_real(a, i)) m = 0;
    const int lane = threadIdx.x & 63;
    const bool any = __ballot(m != 0ull) != 0ull;
    const bool mirrored = a.sym && C != R;
    if (mirrored && km < 0) km = mirror_of(a, k, R, C);
    if (a.pair_nz && lane == 0) {
        a.pair_nz[k - a.k_base] = any ? 1 : 0;
        if (mirrored) a.pair_nz[km - a.k_base] = any ? 1 : 0;
    }
    if (a.pair_nz && !any && !force) return 0;
    a.mask[(k - a.k_base) * 64 + lane] = m;
    if (mirrored) {
        if (!have_col) colword = LOW_REGS ? transpose_bits_shfl(m) : transpose_bits(m);
        else {
            colword &= tile_mask(a, R);                   // padding rows of R
            if (!row_real(a, C * 64 + lane)) colword = 0;  // padding columns of C
        }
        a.mask[(km - a.k_base) * 64 + lane] = colword;
        const int cd = __popcll(colword);
        if (cd) atomicAdd(&a.deg[C * 64 + lane], cd);  // column side of the symmetric pair
    }
    return __popcll(m);
}

// Row degrees of this workgroup's tile: written directly (all pairs evaluated
// here) or added to the column contributions of other tiles (symmetric).
__device__ __forceinline__ void finish_deg(const DistArgs &a, int64_t i, int mydeg) {
    __shared__ int sdeg[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    sdeg[wave][lane] = mydeg;
    __syncthreads();
    if (wave == 0 && i < a.n) {
        const int total = sdeg[0][lane] + sdeg[1][lane] + sdeg[2][lane] + sdeg[3][lane];
        if (a.sym) atomicAdd(&a.deg[i], total);
        else a.deg[i] = total;
    }
}

// Plain FP64, runtime d: 64 per-column accumulators in registers, dims in
// chunks of 8 (the per-pair summation stays in ascending k).
__global__ __launch_bounds__(256) void k_dist_f64_generic(DistArgs a) {
    const int64_t R = a.R0 + blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = R * 64 + lane;
    const int d = a.d;
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    const double r2 = r2_of(a, R);
    int mydeg = 0;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        if (!pair_live(a, R, C)) continue;
        const double *__restrict__ xc = a.xw64 + C * 64 * d;
        double acc[64];
#pragma unroll
        for (int jj = 0; jj < 64; ++jj) acc[jj] = 0.0;
        for (int q = 0; q < d; q += 8) {
            double xi[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) xi[kk] = (q + kk < d) ? a.xw64[i * d + q + kk] : 0.0;
            const int kn = min(8, d - q);
#pragma unroll
            for (int jj = 0; jj < 64; ++jj) {
                const double *__restrict__ xj = xc + (int64_t)jj * d + q;
                double s = acc[jj];
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    if (kk < kn) {
                        double t = xi[kk] - xj[kk];
                        s = s + t * t;
                    }
                }
                acc[jj] = s;
            }
        }
        uint64_t m = 0;
#pragma unroll
        for (int jj = 0; jj < 64; ++jj) m |= (uint64_t)(acc[jj] <= r2) << jj;
        mydeg += store_pair<true>(a, k, R, i, C, m, false, 0);
    }
    finish_deg(a, i, mydeg);
}

// Plain FP64 with the periodic reduced metric (MetricExtEuclideanPeriodicReduced
// ._calc_distance, _types.pyx:2296-2317): per dimension the absolute
// difference, reduced modulo a positive period and replaced by period - x when
// above half of it, squared and added in ascending k, each step rounded as
// there (fmod is exact).  Dense layout, symmetric evaluation; the column
// point's coordinates are wave-uniform loads.
__global__ __launch_bounds__(256) void k_dist_f64_periodic(DistArgs a) {
    const int64_t R = a.R0 + blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = R * 64 + lane;
    const int d = a.d;
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    const double r2 = a.bp.r2;
    const double *__restrict__ xi = a.xw64 + i * d;
    int mydeg = 0;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        const double *__restrict__ xc = a.xw64 + C * 64 * d;
        uint64_t m = 0;
        for (int jj = 0; jj < 64; ++jj) {
            const double *__restrict__ xj = xc + (int64_t)jj * d;
            double total = 0.0;
            for (int q = 0; q < d; ++q) {
                double dist = fabs(xi[q] - xj[q]);
                const double bound = a.bounds[q];
                if (bound > 0.0) {
                    dist = fmod(dist, bound);
                    if (dist > bound / 2.0) dist = bound - dist;
                }
                total = total + dist * dist;
            }
            m |= (uint64_t)(total <= r2) << jj;
        }
        mydeg += store_pair<true>(a, k, R, i, C, m, false, 0);
    }
    finish_deg(a, i, mydeg);
}

// Resolve the band bits of one row: pairs whose FP32 value lies in [lo, hi]
// (or is NaN) are recomputed in FP64 in reference order.
__device__ __forceinline__ uint64_t band_fix(const DistArgs &a, double r2, int64_t i, int64_t C, uint64_t m,
                                             uint64_t band, unsigned &hits, unsigned &flips, double &maxerr,
                                             const float *__restrict__ xc32) {
    band &= tile_mask(a, C);
    if (!row_real(a, i)) band = 0;
    while (band) {
        const int jj = __ffsll((unsigned long long)band) - 1;
        band &= band - 1;
        {
            const double e = exact_d2(a.xw64, a.d, i, C * 64 + jj);
            const bool in = e <= r2;
            m |= (uint64_t)in << jj;
            // recompute the FP32 value for the statistics
            float s32 = 0.0f;
            for (int k = 0; k < a.d; ++k) {
                float t = a.xw32[i * a.d + k] - xc32[jj * a.d + k];
                s32 = __builtin_fmaf(t, t, s32);
            }
            ++hits;
            flips += (in != ((double)s32 <= r2));
            const double err = fabs((double)s32 - e);
            if (err == err) maxerr = fmax(maxerr, err);
        }
    }
    return m;
}

// Resolve FP32 decisions from 64 accumulators (runtime-d kernel).
__device__ __forceinline__ uint64_t band_resolve(const DistArgs &a, double r2, float lo, float hi, int64_t i,
                                                 int64_t C, const float *acc, unsigned &hits, unsigned &flips,
                                                 double &maxerr, const float *__restrict__ xc32) {
    uint64_t m = 0, band = 0;
#pragma unroll
    for (int jj = 0; jj < 64; ++jj) {
        const float s = acc[jj];
        const bool acc_ok = s < lo;
        m |= (uint64_t)acc_ok << jj;
        band |= (uint64_t)(!acc_ok && !(s > hi)) << jj;
    }
    if (band) m = band_fix(a, r2, i, C, m, band, hits, flips, maxerr, xc32);
    return m;
}

// Band statistics of the workgroup, added with one set of atomics per
// workgroup (the counters are single addresses: per-wave atomics there
// serialize).  Every thread of the workgroup must call it.
__device__ __forceinline__ void flush_band(const DistArgs &a, unsigned hits, unsigned flips, double maxerr) {
    __shared__ unsigned long long s_h[16], s_f[16];
    __shared__ double s_m[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, n_waves = (blockDim.x + 63) >> 6;
    unsigned long long h = (unsigned long long)wave_sum((int)hits);
    unsigned long long f = (unsigned long long)wave_sum((int)flips);
    double mx = maxerr;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (lane == 0) {
        s_h[wave] = h;
        s_f[wave] = f;
        s_m[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < n_waves; ++w) {
            h += s_h[w];
            f += s_f[w];
            mx = fmax(mx, s_m[w]);
        }
        if (h) {
            atomicAdd(&a.band_stats[0], h);
            atomicAdd(&a.band_stats[1], f);
            atomic_max_double(&a.band_stats[2], mx);
        }
    }
}

// Column-lane kernels (d <= 8).  Lane j holds column point j of the tile pair
// in registers; the row tile's coordinates sit in LDS and the wave walks its
// 64 rows, reading each row as a broadcast.  Lane j accumulates its column
// word (bit i = pair (row i, column j)) with one compare and an or/select per
// row; one butterfly transpose per tile pair turns the 64 column words into
// the row words, and the column words themselves are the mirror pair's row
// words.  The arithmetic is the reference metric's (difference, square, sum in
// ascending dimension), so every decision equals the row-lane kernels'.

// Calls f(std::integral_constant<int, I>) for I = 0 .. N-1 in order.
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// A zero the compiler cannot see through: indexing the LDS rows with it inside
// the tile-pair loop keeps the row reads in the loop (otherwise all 64 rows
// are hoisted into registers).
__device__ __forceinline__ int opaque_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// Set bit I of (hi:lo) when `on`.
template <int I>
__device__ __forceinline__ void set_bit_if(uint32_t &lo, uint32_t &hi, bool on) {
    if constexpr (I < 32) lo = on ? (lo | (1u << I)) : lo;
    else hi = on ? (hi | (1u << (I - 32))) : hi;
}

// LDS row stride: rows padded to 16-byte multiples, so each row is one or two
// 16-byte LDS reads at an immediate offset from one base register.
template <typename T, int D>
constexpr int lds_row_stride() {
    return (int)(((D * sizeof(T) + 15) / 16) * 16 / sizeof(T));
}

template <typename T, int D>
__device__ __forceinline__ void load_row_tile(T *srow, const T *__restrict__ xw, int64_t R) {
    constexpr int DP = lds_row_stride<T, D>();
    for (int t = threadIdx.x; t < 64 * DP; t += 256) {
        const int r = t / DP, q = t - (t / DP) * DP;
        srow[t] = q < D ? xw[(R * 64 + r) * D + q] : (T)0;
    }
    __syncthreads();
}

template <int D>
__global__ __launch_bounds__(256) void k_dist_f64_cols(DistArgs a) {
    constexpr int DP = lds_row_stride<double, D>();
    __shared__ double srow[64 * DP];
    const int64_t R = block_row_tile(a);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i_lane = R * 64 + lane;
    load_row_tile<double, D>(srow, a.xw64, R);
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    const double r2 = r2_of(a, R);
    int mydeg = 0;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        double xc[D];
#pragma unroll
        for (int q = 0; q < D; ++q) xc[q] = a.xw64[(C * 64 + lane) * D + q];
        uint32_t clo = 0, chi = 0;
        const double *rows = srow + opaque_zero();
        const int groups = a.pair_groups ? a.pair_groups[k] : (1 << kRowGroups) - 1;  // row groups in reach of C
        static_for<0, kRowGroups>([&](auto gc) {
            constexpr int g = decltype(gc)::value;
            if (groups & (1 << g)) {
                static_for<kGroupRows * g, kGroupRows * g + kGroupRows>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    double t = rows[i * DP] - xc[0];
                    double s = t * t;  // 0 + t*t == t*t exactly (t*t is never -0)
#pragma unroll
                    for (int q = 1; q < D; ++q) {
                        t = rows[i * DP + q] - xc[q];
                        s = s + t * t;
                    }
                    set_bit_if<i>(clo, chi, s <= r2);
                });
            }
        });
        const uint64_t colword = ((uint64_t)chi << 32) | clo;
        mydeg += store_pair(a, k, R, i_lane, C, __ballot(colword != 0ull) ? transpose_bits_shfl(colword) : 0ull, true, colword);
    }
    finish_deg(a, i_lane, mydeg);
}

// Transpose for k_dist_f32_cols<D>: the DPP / permlane form where it costs
// no occupancy step (D <= 4), the register-light permute form above that.
template <int D>
__device__ __forceinline__ uint64_t transpose_cols(uint64_t x) {
    if constexpr (D <= 4) return transpose_bits(x);
    else return transpose_bits_shfl(x);
}

template <int D>
__device__ __forceinline__ float f32_d2_row(const float *srow, int i, const float *xc) {
    constexpr int DP = lds_row_stride<float, D>();
    float s = 0.0f;
#pragma unroll
    for (int q = 0; q < D; ++q) {
        const float t = srow[i * DP + q] - xc[q];
        s = __builtin_fmaf(t, t, s);
    }
    return s;
}

template <int D>
__global__ __launch_bounds__(256) void k_dist_f32_cols(DistArgs a) {
    constexpr int DP = lds_row_stride<float, D>();
    __shared__ float srow[64 * DP];
    const int64_t R = block_row_tile(a);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i_lane = R * 64 + lane;
    load_row_tile<float, D>(srow, a.xw32, R);
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    const float lo = a.tile_lo ? a.tile_lo[R] : a.bp.lo, hi = a.tile_hi ? a.tile_hi[R] : a.bp.hi;
    const double r2 = r2_of(a, R);
    int mydeg = 0;
    unsigned hits = 0, flips = 0;
    double maxerr = 0.0;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        const int64_t j = C * 64 + lane;
        float xc[D];
#pragma unroll
        for (int q = 0; q < D; ++q) xc[q] = a.xw32[j * D + q];
        const uint64_t cmask = tile_mask(a, C);
        uint32_t clo = 0, chi = 0;
        uint64_t band_cols = 0;  // columns with an FP32 value inside the guard band (any row)
        const float *rows = srow + opaque_zero();
        const int groups = a.pair_groups ? a.pair_groups[k] : (1 << kRowGroups) - 1;  // row groups in reach of C
        static_for<0, kRowGroups>([&](auto gc) {
            constexpr int g = decltype(gc)::value;
            if (groups & (1 << g)) {
                static_for<kGroupRows * g, kGroupRows * g + kGroupRows>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    const float s = f32_d2_row<D>(rows, i, xc);
                    const bool in = s < lo;
                    band_cols |= ~(__ballot(in) | __ballot(s > hi));
                    set_bit_if<i>(clo, chi, in);
                });
            }
        });
        uint64_t colword = ((uint64_t)chi << 32) | clo;
        if (band_cols & cmask) {
            // rare: those columns recompute their rows and settle the band pairs in FP64
            const bool col_band = ((band_cols & cmask) >> lane) & 1ull;
            for (int i = 0; i < 64; ++i) {
                if (!(groups & (1 << (i / kGroupRows)))) continue;
                const float s = f32_d2_row<D>(rows, i, xc);
                const int64_t gi = R * 64 + i;
                if (col_band && !(s < lo) && !(s > hi) && row_real(a, gi)) {
                    const double e = exact_d2(a.xw64, D, gi, j);
                    const bool inx = e <= r2;
                    if (inx) colword |= 1ull << i;
                    ++hits;
                    flips += (inx != ((double)s <= r2));
                    const double err = fabs((double)s - e);
                    if (err == err) maxerr = fmax(maxerr, err);
                }
            }
        }
        mydeg += store_pair(a, k, R, i_lane, C, __ballot(colword != 0ull) ? transpose_cols<D>(colword) : 0ull, true, colword);
    }
    flush_band(a, hits, flips, maxerr);
    finish_deg(a, i_lane, mydeg);
}

// FP32 direct form, runtime d.
__global__ __launch_bounds__(256) void k_dist_f32_generic(DistArgs a) {
    const int64_t R = a.R0 + blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = R * 64 + lane;
    const int d = a.d;
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    const float lo = a.tile_lo ? a.tile_lo[R] : a.bp.lo, hi = a.tile_hi ? a.tile_hi[R] : a.bp.hi;
    const double r2 = r2_of(a, R);
    int mydeg = 0;
    unsigned hits = 0, flips = 0;
    double maxerr = 0.0;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        if (!pair_live(a, R, C)) continue;
        const float *__restrict__ xc = a.xw32 + C * 64 * d;
        float acc[64];
#pragma unroll
        for (int jj = 0; jj < 64; ++jj) acc[jj] = 0.0f;
        for (int q = 0; q < d; q += 8) {
            float xi[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) xi[kk] = (q + kk < d) ? a.xw32[i * d + q + kk] : 0.0f;
            const int kn = min(8, d - q);
#pragma unroll
            for (int jj = 0; jj < 64; ++jj) {
                const float *__restrict__ xj = xc + (int64_t)jj * d + q;
                float s = acc[jj];
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    if (kk < kn) {
                        float t = xi[kk] - xj[kk];
                        s = __builtin_fmaf(t, t, s);
                    }
                }
                acc[jj] = s;
            }
        }
        uint64_t m = band_resolve(a, r2, lo, hi, i, C, acc, hits, flips, maxerr, xc);
        mydeg += store_pair<true>(a, k, R, i, C, m, false, 0);
    }
    flush_band(a, hits, flips, maxerr);
    finish_deg(a, i, mydeg);
}

// ------------------------------------------------------------------ FP32 Gram form on MFMA (d >= 16)


// Augmented FP32 rows for the Gram form.  With w = |x|^2 (sequential FMA),
// margin e = ka * w + kb and c = r2/4 - w/2:
//   row role    [x_1 .. x_d, 1, c]
//   column role [x_1 .. x_d, c, 1]
// so row_i . col_j = x_i.x_j + c_i + c_j = (r2 - d2_ij)/2.
//
// Packed operand layout (G = dp/8 groups of 8 dimensions per point): the
// float4 at ((t*G + g)*2 + h)*64 + r holds dimensions 8g + 2j + h (j = 0..3)
// of point 64t + r.  MFMA step m = 4g + j multiplies dimension 2m (lanes
// 0-31, k-slot 0) and 2m + 1 (lanes 32-63, k-slot 1), so one wave-wide float4
// load of a 32-row block is 512 contiguous bytes per lane half and feeds four
// steps; dimensions past d + 1 are zero and steps past mfma_k are skipped.
__device__ __forceinline__ int64_t packed_at(int64_t i, int k, int G) {
    const int m = k >> 1;
    return (((i >> 6) * G + (m >> 2)) * 2 + (k & 1)) * 256 + (i & 63) * 4 + (m & 3);
}

__global__ void k_augment_rows(const float *__restrict__ xw32, int64_t n, int64_t n_pad, int d, float r2q, float ka,
                               float kb, float *__restrict__ emarg, float *__restrict__ caug) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    float w = 0.0f;
    if (i < n) {
        const float *x = xw32 + i * d;
        for (int k = 0; k < d; ++k) w = __builtin_fmaf(x[k], x[k], w);
    }
    // non-finite or huge points (|T| could overflow in FP32): margin +inf, so
    // every element of their tile pairs goes to the FP64 settling
    emarg[i] = (w <= 1e36f) ? __builtin_fmaf(ka, w, kb) : INFINITY;
    caug[i] = r2q - 0.5f * w;
}

// One thread per packed float4 (n_pad / 64 * G * 128 of them, in layout order).
__global__ void k_augment_pack(const float *__restrict__ xw32, int64_t n, int64_t n_pad, int d, int G,
                               const float *__restrict__ caug, float4 *__restrict__ xra, float4 *__restrict__ xca) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_pad * 2 * G) return;
    const int r = (int)(idx & 63), h = (int)((idx >> 6) & 1);
    const int64_t tg = idx >> 7;
    const int g = (int)(tg % G);
    const int64_t i = (tg / G) * 64 + r;
    const float c = caug[i];
    float vr[4], vc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int k = 8 * g + 2 * j + h;
        const float x = (k < d && i < n) ? xw32[i * d + k] : 0.0f;
        vr[j] = k < d ? x : k == d ? 1.0f : k == d + 1 ? c : 0.0f;
        vc[j] = k < d ? x : k == d ? c : k == d + 1 ? 1.0f : 0.0f;
    }
    xra[idx] = make_float4(vr[0], vr[1], vr[2], vr[3]);
    xca[idx] = make_float4(vc[0], vc[1], vc[2], vc[3]);
}

__device__ __forceinline__ void mark_pair_nz(const DistArgs &a, int64_t k, int64_t R, int64_t C) {
    a.pair_nz[k - a.k_base] = 1;
    if (a.sym && C != R) a.pair_nz[mirror_of(a, k, R, C) - a.k_base] = 1;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// T = row_i . col_j = (r2 - d2)/2 for a 64x64 tile pair, accumulated in FP32
// on v_mfma_f32_32x32x2f32 as a 2x2 grid of 32x32 blocks; lane l supplies row
// l%32 at k-slot l/32 from the packed layout.  Accumulator element v of lane l
// is T[8(v/4) + 4(l/32) + v%4][l%32].
struct MfmaTile {
    v16f c00, c01, c10, c11;
};

__device__ __forceinline__ void mfma_zero(MfmaTile &t) {
    t.c00 = v16f{0};
    t.c01 = v16f{0};
    t.c10 = v16f{0};
    t.c11 = v16f{0};
}

__device__ __forceinline__ void mfma_step(MfmaTile &t, float a0, float a1, float b0, float b1) {
    t.c00 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, t.c00, 0, 0, 0);
    t.c01 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, t.c01, 0, 0, 0);
    t.c10 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, t.c10, 0, 0, 0);
    t.c11 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, t.c11, 0, 0, 0);
}

// The steps of one group of four (all four, or the first m of the last group).
__device__ __forceinline__ void mfma_group(MfmaTile &t, const float4 &A0, const float4 &A1, const float4 &B0,
                                           const float4 &B1, int m) {
    mfma_step(t, A0.x, A1.x, B0.x, B1.x);
    if (m > 1) mfma_step(t, A0.y, A1.y, B0.y, B1.y);
    if (m > 2) mfma_step(t, A0.z, A1.z, B0.z, B1.z);
    if (m > 3) mfma_step(t, A0.w, A1.w, B0.w, B1.w);
}

// Both operands from global memory (the general form, any d); the next
// group's operands are loaded while the current group's MFMAs run.
__device__ __forceinline__ MfmaTile mfma_tile(const DistArgs &a, int64_t R, int64_t C) {
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, h = lane >> 5;
    const int G = a.dp >> 3;
    const float4 *__restrict__ pa = (const float4 *)a.xra + R * G * 128 + h * 64 + r;
    const float4 *__restrict__ pb = (const float4 *)a.xca + C * G * 128 + h * 64 + r;
    MfmaTile t;
    mfma_zero(t);
    float4 A0 = pa[0], A1 = pa[32], B0 = pb[0], B1 = pb[32];
    for (int g = 0; g < G; ++g) {
        const int gn = (g + 1 < G) ? g + 1 : g;  // the last group reloads itself (in bounds, unused)
        const float4 nA0 = pa[gn * 128], nA1 = pa[gn * 128 + 32];
        const float4 nB0 = pb[gn * 128], nB1 = pb[gn * 128 + 32];
        mfma_group(t, A0, A1, B0, B1, a.mfma_k - 4 * g);
        A0 = nA0;
        A1 = nA1;
        B0 = nB0;
        B1 = nB1;
    }
    return t;
}

// Words of the tile pair from per-lane row bits: lane (r, h) sets, for its
// rows 8(v/4) + v%4 (+32 for the second row block, shifted by 4h), the bits of
// its two columns r (blocks c00, c10) and 32 + r (c01, c11); the lane halves
// are merged with one exchange and lane j ends up with column j's word.
__device__ __forceinline__ uint64_t merge_col_words(uint32_t a_lo, uint32_t a_hi, uint32_t b_lo, uint32_t b_hi) {
    const int h = (threadIdx.x & 63) >> 5;
    const uint64_t wa = ((((uint64_t)a_hi) << 32) | a_lo) << (4 * h);
    const uint64_t wb = ((((uint64_t)b_hi) << 32) | b_lo) << (4 * h);
    const uint64_t fa = wa | __shfl_xor(wa, 32, 64);
    const uint64_t fb = wb | __shfl_xor(wb, 32, 64);
    return h == 0 ? fa : fb;
}

// Band flags of a lane use the layout of its column words: bits 0-31 column
// l%32 (blocks c00, c10), bits 32-63 column 32 + l%32 (c01, c11); within a
// half, bit q is accumulator slot v = 4(q/8) + q%4 of the first row block
// (q/4 even) or of the second (q/4 odd, rows + 32), i.e. row
// 8(q/8) + 4(l/32) + q%4 (+32).
__device__ __forceinline__ void band_element(int lane, int e, int &row, int &col) {
    const int q = e & 31;
    row = 8 * (q >> 3) + 4 * (lane >> 5) + (q & 3) + (((q >> 2) & 1) ? 32 : 0);
    col = (lane & 31) + (e >= 32 ? 32 : 0);
}

// The column-word halves (a_lo | a_hi << 4, b_lo | b_hi << 4) as one flag word.
__device__ __forceinline__ uint64_t pack_halves(uint32_t a_lo, uint32_t a_hi, uint32_t b_lo, uint32_t b_hi) {
    return (uint64_t)(a_lo | (a_hi << 4)) | ((uint64_t)(b_lo | (b_hi << 4)) << 32);
}

// Guard-band elements of the listed tile pairs, settled one per thread.  The
// records (k, R, 64 per-lane flag words) are flattened into a list of
// elements (record << 12 | lane << 6 | bit): k_band_count counts each
// record's elements, an exclusive scan places them, k_band_list writes them,
// and k_band_settle decides each in FP64 in reference order and ors its bit
// into the stored words of the pair and of its mirror (and counts it in both
// degrees).  The FP32 statistics recompute the Gram value from the augmented
// vectors (sequential FMAs; statistics only).
__global__ __launch_bounds__(256) void k_band_count(const int64_t *__restrict__ band_list, int64_t n_rec,
                                                    int32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t rec = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (rec >= n_rec) return;
    if (band_list[rec * 66] < 0) {  // a reserved slot left empty
        if (lane == 0) cnt[rec] = 0;
        return;
    }
    const int c = wave_sum(__popcll((uint64_t)band_list[rec * 66 + 2 + lane]));
    if (lane == 0) cnt[rec] = c;
}

__global__ __launch_bounds__(256) void k_band_list(const int64_t *__restrict__ band_list, int64_t n_rec,
                                                   const int64_t *__restrict__ off, int64_t *__restrict__ elem) {
    const int lane = threadIdx.x & 63;
    const int64_t rec = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (rec >= n_rec || band_list[rec * 66] < 0) return;
    uint64_t f = (uint64_t)band_list[rec * 66 + 2 + lane];
    const int c = __popcll(f);
    int64_t o = off[rec] + wave_incl_scan(c) - c;
    while (f) {
        const int b = __ffsll((unsigned long long)f) - 1;
        f &= f - 1;
        elem[o++] = (rec << 12) | ((int64_t)lane << 6) | b;
    }
}

// exact_d2 with the row reads grouped: eight dimensions of both rows loaded
// together as 16-byte pairs (rows start 16-byte aligned when d is even), so a
// thread's accesses to a row's cache lines come close together; the sum is
// the same sequence of separately rounded steps in ascending k.
__device__ __forceinline__ double exact_d2_grouped(const double *__restrict__ xw64, int d, int64_t i, int64_t j) {
    const double *a = xw64 + i * d;
    const double *b = xw64 + j * d;
    double s = 0.0;
    int k = 0;
    if ((d & 1) == 0) {
        for (; k + 8 <= d; k += 8) {
            const double2 *pa = (const double2 *)(a + k), *pb = (const double2 *)(b + k);
            const double2 a0 = pa[0], a1 = pa[1], a2 = pa[2], a3 = pa[3];
            const double2 b0 = pb[0], b1 = pb[1], b2 = pb[2], b3 = pb[3];
            const double av[8] = {a0.x, a0.y, a1.x, a1.y, a2.x, a2.y, a3.x, a3.y};
            const double bv[8] = {b0.x, b0.y, b1.x, b1.y, b2.x, b2.y, b3.x, b3.y};
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double t = av[q] - bv[q];
                s = s + t * t;
            }
        }
    }
    for (; k < d; ++k) {
        const double t = a[k] - b[k];
        s = s + t * t;
    }
    return s;
}

// FP32 Gram value of points gi, gj from the packed augmented vectors (four
// dimensions per float4; statistics only).
__device__ __forceinline__ float gram_value(const DistArgs &a, int64_t gi, int64_t gj) {
    const int G = a.dp >> 3;
    const float4 *pa = (const float4 *)a.xra + (gi >> 6) * G * 128 + (gi & 63);
    const float4 *pb = (const float4 *)a.xca + (gj >> 6) * G * 128 + (gj & 63);
    float prod = 0.0f;
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float4 x = pa[(g * 2 + h) * 64], y = pb[(g * 2 + h) * 64];
            prod = __builtin_fmaf(x.x, y.x, prod);
            prod = __builtin_fmaf(x.y, y.y, prod);
            prod = __builtin_fmaf(x.z, y.z, prod);
            prod = __builtin_fmaf(x.w, y.w, prod);
        }
    }
    return prod;
}

__global__ __launch_bounds__(256) void k_band_settle(DistArgs a, const int64_t *__restrict__ elem, int64_t n_el) {
    unsigned hits = 0, flips = 0;
    double maxerr = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_el; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = elem[t];
        const int64_t rec = v >> 12;
        const int le = (int)((v >> 6) & 63), e = (int)(v & 63);
        const int64_t k = a.band_list[rec * 66], R = a.band_list[rec * 66 + 1];
        int64_t kb, ke;
        tile_range(a, R, kb, ke);
        const int64_t C = tile_col(a, k, kb);
        int row, col;
        band_element(le, e, row, col);
        const int64_t gi = R * 64 + row, gj = C * 64 + col;
        const double ex = exact_d2_grouped(a.xw64, a.d, gi, gj);
        const bool in = ex <= a.bp.r2;
        const double d2 = a.bp.r2 - 2.0 * (double)gram_value(a, gi, gj);
        ++hits;
        flips += in != (d2 <= a.bp.r2);
        const double err = fabs(d2 - ex);
        if (err == err) maxerr = fmax(maxerr, err);
        if (in) {
            atomicOr(&a.mask[(k - a.k_base) * 64 + row], 1ull << col);
            atomicAdd(&a.deg[gi], 1);
            if (a.pair_nz) a.pair_nz[k - a.k_base] = 1;
            if (a.sym && C != R) {
                const int64_t km = mirror_of(a, k, R, C);
                atomicOr(&a.mask[(km - a.k_base) * 64 + col], 1ull << row);
                atomicAdd(&a.deg[gj], 1);
                if (a.pair_nz) a.pair_nz[km - a.k_base] = 1;
            }
        }
    }
    flush_band(a, hits, flips, maxerr);
}

// Decision margin of a tile pair: NaN (every element fails both the
// certain-in and the certain-out test and is settled in FP64) when a point of
// either tile is non-finite or huge, whose margin is +inf.
__device__ __forceinline__ float pair_margin(float e_row, float e_col) {
    const float E = e_row + e_col;
    return E < INFINITY ? E : NAN;
}

// Decides the elements of one tile pair from its FP32 values.  With E = (largest
// margin of the row tile) + (largest margin of the column tile): T >= E is a
// neighbour for certain, T < -E is not.  Elements in between (and NaN) are left
// out here and the pair is queued (a record of per-lane band flags, or the pair
// alone when the record list is full) for settling in FP64 in reference order.
// No LDS: each lane sets its rows' bits in its column words, one butterfly
// transpose gives the row words, and the column words are the mirror pair's
// row words.  Stores both and returns the row word's popcount.
// Band record slots of a wave, reserved kBandSlotChunk at a time with one
// atomic (the counter is a single address); slots left at the wave's end are
// marked empty (k = -1) by band_slots_release.
constexpr int kBandSlotChunk = 8;
struct BandSlots {
    unsigned long long next = 0, end = 0;
};

__device__ __forceinline__ void band_slots_release(const DistArgs &a, BandSlots &bs) {
    if ((threadIdx.x & 63) == 0)
        for (unsigned long long q = bs.next; q < bs.end && (int64_t)q < a.band_cap; ++q) a.band_list[q * 66] = -1;
    bs.next = bs.end;
}

__device__ __forceinline__ int mfma_settle_store(const DistArgs &a, const MfmaTile &t, int64_t k, int64_t R,
                                                 int64_t C, int64_t i, float E, BandSlots &bs, int64_t km = -1) {
    const int lane = threadIdx.x & 63;
    const float nE = -E;
    // Fast path: the largest element is below -E.  fmaxf skips NaN elements, but
    // those arise only from non-finite or huge inputs, whose pair margin is NaN,
    // which fails this test and every certain-in / certain-out test below.
    float mx = -INFINITY;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        mx = fmaxf(fmaxf(mx, t.c00[v]), t.c01[v]);
        mx = fmaxf(fmaxf(mx, t.c10[v]), t.c11[v]);
    }
    uint64_t word = 0, colword = 0;
    bool queued = false;  // guard-band elements settled later: the words must be stored
    if (__ballot(!(mx < nE))) {
        // per element: in = (T >= E), not-out = !(T < -E) (true for NaN); the
        // band is not-out and not in
        uint32_t a_lo = 0, a_hi = 0, b_lo = 0, b_hi = 0;
        uint32_t na_lo = 0, na_hi = 0, nb_lo = 0, nb_hi = 0;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const uint32_t bit = 1u << (8 * (v >> 2) + (v & 3));
            a_lo = t.c00[v] >= E ? (a_lo | bit) : a_lo;
            a_hi = t.c10[v] >= E ? (a_hi | bit) : a_hi;
            b_lo = t.c01[v] >= E ? (b_lo | bit) : b_lo;
            b_hi = t.c11[v] >= E ? (b_hi | bit) : b_hi;
            na_lo = t.c00[v] < nE ? na_lo : (na_lo | bit);
            na_hi = t.c10[v] < nE ? na_hi : (na_hi | bit);
            nb_lo = t.c01[v] < nE ? nb_lo : (nb_lo | bit);
            nb_hi = t.c11[v] < nE ? nb_hi : (nb_hi | bit);
        }
        uint64_t flags = pack_halves(na_lo, na_hi, nb_lo, nb_hi) & ~pack_halves(a_lo, a_hi, b_lo, b_hi);
        if (__ballot(flags != 0ull)) {
            // only real rows and columns
            const uint64_t cm = valid_cols(a.n, C);
            const bool ca = (cm >> (lane & 31)) & 1ull, cb = (cm >> (32 + (lane & 31))) & 1ull;
            flags &= (ca ? 0x00000000ffffffffull : 0ull) | (cb ? 0xffffffff00000000ull : 0ull);
            if (R * 64 + 64 > a.n) {
                for (int e = 0; e < 64; ++e) {
                    int row, col;
                    band_element(lane, e, row, col);
                    if (R * 64 + row >= a.n) flags &= ~(1ull << e);
                }
            }
        }
        if (__ballot(flags != 0ull)) {
            queued = true;
            if (bs.next == bs.end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(a.band_cnt, (unsigned long long)kBandSlotChunk);
                base = __shfl(base, 0, 64);
                bs.next = base;
                bs.end = base + kBandSlotChunk;
            }
            const unsigned long long slot = bs.next++;
            if ((int64_t)slot < a.band_cap) {
                int64_t *p = a.band_list + (int64_t)slot * 66;
                if (lane == 0) {
                    p[0] = k;
                    p[1] = R;
                }
                p[2 + lane] = (int64_t)flags;
            } else if (lane == 0) {
                const unsigned long long rs = atomicAdd(a.redo_cnt, 1ull);
                a.redo[2 * rs] = k;
                a.redo[2 * rs + 1] = R;
            }
        }
        colword = merge_col_words(a_lo, a_hi, b_lo, b_hi);
        word = transpose_bits(colword);
    }
    return store_pair(a, k, R, i, C, word, true, colword, queued, km);
}

// General MFMA form (any d, explicit or dense pair lists): one wave per 64x64
// tile pair, the 4 waves of a workgroup share the row tile; both operands are
// read from global memory.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_dist_mfma(DistArgs a) {
    const int64_t R = a.R0 + blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i = R * 64 + lane;
    const float e_row = wave_max(a.emarg[R * 64 + lane]);
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t k0 = uniform64(first_pair(a, R, kb, ke));
    int mydeg = 0;
    BandSlots bs;
    for (int64_t k = k0 + wave; k < ke; k += 4) {
        const int64_t C = uniform64(tile_col(a, k, kb));
        if (!pair_live(a, R, C)) continue;
        const float e_col = a.emarg[C * 64 + lane];
        const MfmaTile t = mfma_tile(a, R, C);
        mydeg += mfma_settle_store(a, t, k, R, C, i, pair_margin(e_row, wave_max(e_col)), bs);
    }
    band_slots_release(a, bs);
    finish_deg(a, i, mydeg);
}

// Row blocks of the dense LDS MFMA kernel with work in column chunk cc: all of
// them, or (symmetric evaluation, C >= R only) those starting before the
// chunk's end.
__host__ __device__ inline int64_t mfma_chunk_blocks(int sym, int64_t cc, int64_t chunk, int64_t n_tiles,
                                                     int64_t R0, int64_t n_rb) {
    if (!sym) return n_rb;
    const int64_t c_hi = cc * chunk + chunk < n_tiles ? cc * chunk + chunk : n_tiles;
    const int64_t b = (c_hi - R0 + 3) / 4;
    return b < 0 ? 0 : (b < n_rb ? b : n_rb);
}

// Dense MFMA form for d + 2 <= 8G.  A workgroup takes 4 consecutive row tiles
// (one per wave, its operands held in registers for the whole launch) and a
// chunk of column tiles; each column tile is staged once in LDS (double
// buffered: the next tile's loads are in flight during the current tile's
// MFMAs) and read by all 4 waves.  The work items (column chunk, row block)
// that hold any pair are numbered chunk-major and dealt to the 8 XCDs in
// contiguous ranges of equal length (workgroup b runs on XCD b % 8), so every
// XCD gets the same work and workgroups that share an L2 read the same column
// tiles.  Degrees are added atomically (several chunks contribute to a row).
constexpr int kMfmaMaxChunk = 1024;  // largest column chunk of the tuning table
template <int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_dist_mfma_lds(
    DistArgs a, int64_t n_rt, int64_t n_rb, int64_t n_cc, int64_t per_xcd, int64_t n_items) {
    constexpr int kTileF4 = G * 128;  // float4 per packed tile
    constexpr int kPer = (kTileF4 + 255) / 256;
    __shared__ float4 sB[2][kTileF4];
    __shared__ float sE[2];  // largest margin of the staged column tile
    __shared__ int32_t sCol[kMfmaMaxChunk];  // live column tiles of the item (culled layout)
    __shared__ int sNcol;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 31, h = lane >> 5;
    // items in contiguous ranges per XCD, or dealt round-robin (deal_rr)
    const int64_t item = a.deal_rr ? (int64_t)blockIdx.x : (int64_t)(blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (item >= n_items) return;
    int64_t cc = 0, rb = item;
    for (; cc < n_cc; ++cc) {
        const int64_t nb = mfma_chunk_blocks(a.sym, cc, a.col_chunk, a.n_tiles, a.R0, n_rb);
        if (rb < nb) break;
        rb -= nb;
    }
    if (cc >= n_cc) return;  // not reached: item < n_items
    const int64_t R_lo = a.R0 + 4 * rb;
    const int64_t R = R_lo + wave;
    const bool rvalid = R < a.R0 + n_rt;
    const int64_t c_lo = cc * a.col_chunk;
    const int64_t c_hi = min(a.n_tiles, c_lo + a.col_chunk);
    const int64_t c_begin = a.sym ? max(c_lo, R_lo) : c_lo;
    if (c_begin >= c_hi) return;  // the whole workgroup
    const int tail = a.mfma_k - 4 * (G - 1);
    // the column tiles to visit: all of [c_begin, c_hi), or (culled layout)
    // those with a live pair for any of the 4 row tiles
    int n_col = (int)(c_hi - c_begin);
    if (a.live) {
        if (wave == 0) {
            const int64_t r_end = a.R0 + n_rt;
            int cnt = 0;
            for (int64_t cw = c_begin; cw < c_hi; cw += 64) {
                const int64_t C = cw + lane;
                bool any = false;
                if (C < c_hi) {
                    for (int q = 0; q < 4; ++q) {
                        const int64_t Rq = R_lo + q;
                        if (Rq < r_end && (!a.sym || C >= Rq) && pair_live(a, Rq, C)) any = true;
                    }
                }
                const uint64_t m = __ballot(any);
                const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                if (any) sCol[cnt + rank] = (int32_t)C;
                cnt += __popcll(m);
            }
            if (lane == 0) sNcol = cnt;
        }
        __syncthreads();
        n_col = sNcol;
        if (n_col == 0) return;  // the whole workgroup
    }
    auto col_at = [&](int idx) -> int64_t { return a.live ? (int64_t)sCol[idx] : c_begin + idx; };
    const int64_t C0 = col_at(0);

    float4 A0[G], A1[G];
    float e_row = 0.0f;
    if (rvalid) {
        const float4 *__restrict__ pa = (const float4 *)a.xra + R * kTileF4 + h * 64 + r;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            A0[g] = pa[g * 128];
            A1[g] = pa[g * 128 + 32];
        }
        e_row = wave_max(a.emarg[R * 64 + lane]);
    } else {
#pragma unroll
        for (int g = 0; g < G; ++g) A0[g] = A1[g] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const float4 *__restrict__ xb = (const float4 *)a.xca;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int q = threadIdx.x + 256 * u;
        if (q < kTileF4) sB[0][q] = xb[C0 * kTileF4 + q];
    }
    if (wave == 0) {
        const float e = wave_max(a.emarg[C0 * 64 + lane]);
        if (lane == 0) sE[0] = e;
    }
    __syncthreads();
    const int64_t i = R * 64 + lane;
    int mydeg = 0;
    BandSlots bs;
    for (int idx = 0; idx < n_col; ++idx) {
        const int64_t C = col_at(idx);
        const int buf = idx & 1;
        const bool more = idx + 1 < n_col;
        // the next tile's operands and margins (the last tile reloads itself, in
        // bounds, unused): loads in flight during this tile's MFMAs, no branch
        const int64_t Cn = more ? col_at(idx + 1) : C;
        // named registers, not an array: a private array here is placed in LDS or
        // scratch memory by the compiler, which waits for the loads right away
        static_assert(kPer <= 5, "staging registers");
        const float4 *src = xb + Cn * kTileF4;
        float4 n0, n1, n2, n3, n4;
        n0 = src[min((int)threadIdx.x, kTileF4 - 1)];
        if constexpr (kPer > 1) n1 = src[min((int)threadIdx.x + 256, kTileF4 - 1)];
        if constexpr (kPer > 2) n2 = src[min((int)threadIdx.x + 512, kTileF4 - 1)];
        if constexpr (kPer > 3) n3 = src[min((int)threadIdx.x + 768, kTileF4 - 1)];
        if constexpr (kPer > 4) n4 = src[min((int)threadIdx.x + 1024, kTileF4 - 1)];
        const float e_nx = a.emarg[Cn * 64 + lane];
        const bool work = rvalid && (!a.sym || C >= R) && pair_live(a, R, C);
        // the pair's mask index and its mirror's, loaded during the MFMAs
        int64_t kp = 0, km = -1;
        if (work) {
            kp = pair_index(a, R, C);
            if (a.sym && C != R) km = mirror_of(a, kp, R, C);
        }
        MfmaTile t;
        mfma_zero(t);
        if (work) {
            const float4 *sb = &sB[buf][h * 64 + r];
            float4 B0 = sb[0], B1 = sb[32];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                float4 nB0 = B0, nB1 = B1;
                if (g + 1 < G) {
                    nB0 = sb[(g + 1) * 128];
                    nB1 = sb[(g + 1) * 128 + 32];
                }
                mfma_group(t, A0[g], A1[g], B0, B1, g + 1 < G ? 4 : tail);
                B0 = nB0;
                B1 = nB1;
            }
        }
        const float e_col = sE[buf];
        // Stage the next tile between the MFMAs and the epilogue: the loads have
        // had the MFMAs' time to arrive, and no store of this tile is in flight
        // yet (waiting for the loads would also wait for those).  The empty asm
        // keeps the compiler from moving the LDS stores up next to their loads.
        asm volatile("" ::: "memory");
        if (more) {
            float4 *dst = sB[buf ^ 1];
            const int q = threadIdx.x;
            dst[q] = n0;  // kTileF4 >= 384
            if constexpr (kPer > 1) { if (q + 256 < kTileF4) dst[q + 256] = n1; }
            if constexpr (kPer > 2) { if (q + 512 < kTileF4) dst[q + 512] = n2; }
            if constexpr (kPer > 3) { if (q + 768 < kTileF4) dst[q + 768] = n3; }
            if constexpr (kPer > 4) { if (q + 1024 < kTileF4) dst[q + 1024] = n4; }
            if (wave == 0) {
                const float e = wave_max(e_nx);
                if (lane == 0) sE[buf ^ 1] = e;
            }
        }
        asm volatile("" ::: "memory");
        if (work) mydeg += mfma_settle_store(a, t, kp, R, C, i, pair_margin(e_row, e_col), bs, km);
        __syncthreads();
    }
    band_slots_release(a, bs);
    if (rvalid && mydeg && i < a.n) atomicAdd(&a.deg[i], mydeg);
}

// Band elements of the queued tile pairs (one wave per pair): the FP32 values
// are recomputed exactly as in k_dist_mfma; every element with -E <= T < E is
// decided in FP64 in reference order (one wave-wide evaluation per element),
// and its bits are added to the stored words (and to the degrees) of the pair
// and of its mirror.
__global__ __launch_bounds__(64) void k_dist_mfma_band(DistArgs a, int64_t n_redo) {
    const int64_t slot = blockIdx.x;
    if (slot >= n_redo) return;
    const int lane = threadIdx.x & 63;
    const int r = lane & 31, h = lane >> 5;
    const int64_t k = a.redo[2 * slot], R = a.redo[2 * slot + 1];
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    const int64_t C = uniform64(tile_col(a, k, kb));
    const float e_row = wave_max(a.emarg[R * 64 + lane]);
    const float E = pair_margin(e_row, wave_max(a.emarg[C * 64 + lane]));
    const float nE = -E;
    const MfmaTile t = mfma_tile(a, R, C);
    unsigned hits = 0, flips = 0;
    double maxerr = 0.0;
    const uint64_t cm = valid_cols(a.n, C);
    const bool col_a = (cm >> r) & 1ull, col_b = (cm >> (32 + r)) & 1ull;
    uint32_t a_lo = 0, a_hi = 0, b_lo = 0, b_hi = 0;
    // element (row 4h + rr + rb, column cb + r) of lane (r, h) with value x,
    // settled by its own lane: the exact FP64 distance in reference order
    auto settle = [&](float x, int rr, int rb, int cb, bool col_ok, uint32_t &part) {
        const bool band = !(x >= E) && !(x < nE) && col_ok && (R * 64 + 4 * h + rr + rb) < a.n;
        if (band) {
            const int64_t gi = R * 64 + 4 * h + rr + rb;
            const int64_t gj = C * 64 + cb + r;
            const double e = exact_d2(a.xw64, a.d, gi, gj);
            const bool in = e <= a.bp.r2;
            if (in) part |= 1u << rr;
            ++hits;
            const double d2 = a.bp.r2 - 2.0 * (double)x;
            flips += (in != (d2 <= a.bp.r2));
            const double err = fabs(d2 - e);
            if (err == err) maxerr = fmax(maxerr, err);
        }
    };
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int rr = 8 * (v >> 2) + (v & 3);
        settle(t.c00[v], rr, 0, 0, col_a, a_lo);
        settle(t.c10[v], rr, 32, 0, col_a, a_hi);
        settle(t.c01[v], rr, 0, 32, col_b, b_lo);
        settle(t.c11[v], rr, 32, 32, col_b, b_hi);
    }
    const uint64_t colword = merge_col_words(a_lo, a_hi, b_lo, b_hi);
    const uint64_t word = transpose_bits(colword);  // lane i: row i's extra bits
    const int64_t i = R * 64 + lane;
    const bool any_new = __ballot(word != 0ull) != 0ull;
    if (a.pair_nz && lane == 0 && any_new) mark_pair_nz(a, k, R, C);
    if (word) {
        atomicOr(&a.mask[(k - a.k_base) * 64 + lane], (unsigned long long)word);
        atomicAdd(&a.deg[i], __popcll(word));
    }
    if (a.sym && C != R && colword) {
        const int64_t km = a.tp_ptr ? a.mirror[k] : C * a.n_tiles + R;
        atomicOr(&a.mask[(km - a.k_base) * 64 + lane], (unsigned long long)colword);
        atomicAdd(&a.deg[C * 64 + lane], __popcll(colword));
    }
    flush_band(a, hits, flips, maxerr);
}

// Plain FP64 for the dense layout with 4 < d <= KMAX, on a copy of the
// coordinates zero-padded to KMAX dims: the padded terms add (0 - 0)^2 = +0 at
// the end of the reference-order sum, which leaves every s_ij unchanged.  The
// row point lives in VGPRs, column points stream through wave-uniform (scalar)
// loads.  A workgroup covers 4 consecutive row tiles (one per wave) and a chunk
// of column tiles, so the 4 waves read the same column rows (scalar-cache
// reuse); the chunking balances the triangle of the symmetric evaluation.
// Degrees are accumulated atomically (several chunks contribute to a row).
template <int KMAX>
__global__ __launch_bounds__(256) void k_dist_f64_wide(DistArgs a, const double *__restrict__ xp, int64_t n_chunks,
                                                       int64_t n_rt) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t rg = blockIdx.x / n_chunks, cc = blockIdx.x % n_chunks;
    const int64_t R = a.R0 + rg * 4 + wave;  // row tiles R0 .. R0 + n_rt - 1 of this batch
    const int64_t T = a.n_tiles;
    if (R >= a.R0 + n_rt) return;
    const int64_t i = R * 64 + lane;
    double xi[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) xi[k] = xp[i * KMAX + k];
    const int64_t c_lo = cc * a.col_chunk, c_hi = min(T, (cc + 1) * a.col_chunk);
    const int64_t c_begin = a.sym ? max(c_lo, R) : c_lo;
    int mydeg = 0;
    for (int64_t C = c_begin; C < c_hi; ++C) {
        if (!pair_live(a, R, C)) continue;
        const double *__restrict__ xc = xp + uniform64(C) * 64 * KMAX;
        uint64_t m = 0;
#pragma unroll 2
        for (int jj = 0; jj < 64; jj += 2) {
            const double *__restrict__ x0 = xc + jj * KMAX;
            const double *__restrict__ x1 = x0 + KMAX;
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) {
                const double t0 = xi[k] - x0[k];
                const double t1 = xi[k] - x1[k];
                s0 = s0 + t0 * t0;
                s1 = s1 + t1 * t1;
            }
            m |= (uint64_t)(s0 <= a.bp.r2) << jj;
            m |= (uint64_t)(s1 <= a.bp.r2) << (jj + 1);
        }
        mydeg += store_pair<true>(a, pair_index(a, R, C), R, i, C, m, false, 0);
    }
    if (i < a.n && mydeg) atomicAdd(&a.deg[i], mydeg);
}

__global__ void k_pad_rows(const double *__restrict__ xw64, int64_t n_pad, int d, int kmax, double *__restrict__ xp) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_pad * kmax) return;
    const int64_t w = idx / kmax;
    const int k = (int)(idx - w * kmax);
    xp[idx] = k < d ? xw64[w * d + k] : 0.0;
}

// Pairs per wave of k_pair_info in one-sided builds, and the count of such
// chunks of each row tile (its pairs from its own on).
constexpr int64_t kPairChunk = 128;
__global__ void k_pair_chunks(const int64_t *__restrict__ tp_ptr, const int32_t *__restrict__ diag, int64_t T,
                              int32_t *__restrict__ cnt) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= T) return;
    const int64_t kS = tp_ptr[R] + (diag[R] >= 0 ? diag[R] : 0);
    cnt[R] = (int32_t)((tp_ptr[R + 1] - kS + kPairChunk - 1) / kPairChunk);
}

// Mirror index of every candidate pair (R, C): position of R in C's list.
// Per candidate pair k = (R, C), lane-parallel over the pairs of a row tile:
// mirror[k] = index of (C, R) in C's list, and gmask[k] = the row groups of R
// whose boxes are within reach of tile C (bit g for group g).
__global__ __launch_bounds__(256) void k_pair_info(const int64_t *__restrict__ tp_ptr,
                                                   const int32_t *__restrict__ tp_col, int64_t n_tiles,
                                                   int64_t *__restrict__ mirror, const double *__restrict__ glo,
                                                   const double *__restrict__ ghi, const double *__restrict__ blo,
                                                   const double *__restrict__ bhi, int d, double thr_all,
                                                   const double *__restrict__ tile_thr,
                                                   GroupMask *__restrict__ gmask,
                                                   unsigned long long *__restrict__ rows_evaluated,
                                                   const int32_t *__restrict__ diag, int upper_only,
                                                   const int64_t *__restrict__ chunk_off,
                                                   const int32_t *__restrict__ chunk_cnt) {
    // rows_evaluated[2R]: rows of R evaluated against C over the pairs with
    // C >= R (one-sided evaluation), rows_evaluated[2R + 1]: over all pairs.
    // upper_only (the graph is evaluated one-sided): pairs with C < R are
    // never read, so they get neither mirror nor groups, and rows_evaluated
    // [2R + 1] is not filled.
    // the group boxes of R, shared by all of R's pairs: staged once per wave.
    // Without chunks two waves per row tile (pairs 64h + lane, step 128):
    // twice the loads in flight of one wave's dependent chain.  With chunks
    // (one-sided), wave gw takes chunk gw of kPairChunk pairs from C = R on,
    // numbered row by row (chunk_off: exclusive scan of chunk_cnt), so a row
    // tile with thousands of candidates is spread over many waves.
    __shared__ double sgl[4][kRowGroups * kSpatialMaxD], sgh[4][kRowGroups * kSpatialMaxD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t gw = (int64_t)blockIdx.x * 4 + wave;
    int64_t R, part;
    if (chunk_off) {
        int64_t lo = 0, hi = n_tiles;  // the last row tile whose chunks start at or before gw
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (chunk_off[mid] <= gw) lo = mid;
            else hi = mid;
        }
        R = lo;
        part = gw - chunk_off[R];
        if (part >= chunk_cnt[R]) R = n_tiles;  // past the last chunk
    } else {
        R = gw >> 1;
        part = gw & 1;
    }
    if (R < n_tiles)
        for (int t = lane; t < kRowGroups * d; t += 64) {
            sgl[wave][t] = glo[R * kRowGroups * d + t];
            sgh[wave][t] = ghi[R * kRowGroups * d + t];
        }
    __syncthreads();
    if (R >= n_tiles) return;
    const double thr = tile_thr ? tile_thr[R] : thr_all;
    unsigned long long rows = 0, rows_all = 0;
    const int64_t kR = tp_ptr[R];
    const int32_t dR = diag ? diag[R] : -1;
    const int64_t kS = upper_only && dR >= 0 ? kR + dR : kR;  // the pairs with C >= R start at the diagonal
    const int64_t kE = tp_ptr[R + 1];
    const int64_t k_lo = chunk_off ? kS + kPairChunk * part : kS + 64 * part;
    const int64_t k_hi = chunk_off ? std::min<int64_t>(kE, kS + kPairChunk * (part + 1)) : kE;
    const int64_t k_step = chunk_off ? 64 : 128;
    for (int64_t k = k_lo + lane; k < k_hi; k += k_step) {
        const int64_t C = tp_col[k];
        if (upper_only && C < R) continue;
        const int64_t cb = tp_ptr[C], ce = tp_ptr[C + 1];
        // Only pairs with C > R read their mirror: the one-sided evaluation
        // visits C >= R and writes the words of (C, R) too.  The others keep
        // their own index.
        // Neighbouring tiles have similar lists: R sits in C's list about as far
        // from C's own place as C sits from R's in R's list.  Look there first,
        // 16 entries at once; on a miss, search the side of the window that R
        // is on, below C's own place (R < C).
        int64_t m = C > R ? -1 : k;
        const int32_t dC = diag ? diag[C] : -1;
        int64_t lo = cb, hi = ce;  // R's place in C's list is in [lo, hi)
        if (dC >= 0) hi = cb + dC;
        if (m < 0 && dR >= 0 && dC >= 0 && ce - cb >= 16) {
            const int64_t est = dC - ((k - kR) - dR);
            const int64_t w0 = cb + std::min<int64_t>(std::max<int64_t>(est - 8, 0), ce - cb - 16);
            int32_t v[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) v[q] = tp_col[w0 + q];
            int below = 0;
#pragma unroll
            for (int q = 0; q < 16; ++q) below += v[q] < (int32_t)R;
            if (below < 16 && v[below] == (int32_t)R) m = w0 + below;
            else if (below == 0) hi = std::min(hi, w0);
            else if (below == 16) lo = std::max(lo, w0 + 16);
        }
        if (m < 0) {
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (tp_col[mid] < R) lo = mid + 1;
                else hi = mid;
            }
            m = lo;
        }
        mirror[k] = m;
        double cl[kSpatialMaxD], ch[kSpatialMaxD];  // box of C, in registers
#pragma unroll
        for (int q = 0; q < kSpatialMaxD; ++q) {
            cl[q] = q < d ? blo[C * d + q] : 0.0;
            ch[q] = q < d ? bhi[C * d + q] : 0.0;
        }
        uint32_t gm = 0;
        for (int gi = 0; gi < kRowGroups; ++gi) {
            // boxes_close(group gi of R, C)
            double sum = 0.0;
#pragma unroll
            for (int q = 0; q < kSpatialMaxD; ++q) {
                if (q < d) {
                    double g = fmax(cl[q] - sgh[wave][gi * d + q], sgl[wave][gi * d + q] - ch[q]);
                    g = fmax(g, 0.0);
                    sum += g * g;
                }
            }
            gm |= (uint32_t)(sum <= thr) << gi;
        }
        gmask[k] = (GroupMask)gm;
        const unsigned long long r = (unsigned long long)(__popc(gm) * kGroupRows);
        rows_all += r;
        if (C >= R) rows += r;
    }
    rows = (unsigned long long)wave_sum64((long long)rows);
    rows_all = (unsigned long long)wave_sum64((long long)rows_all);
    if (lane == 0) {  // (zeroed beforehand; the two waves of R add)
        if (rows) atomicAdd(&rows_evaluated[2 * R], rows);
        if (rows_all) atomicAdd(&rows_evaluated[2 * R + 1], rows_all);
    }
}

// Sort keys that put the heavy row tiles first, heaviest first (rows to
// evaluate, the one-sided count rows[2R], above kHeavyRowFactor times the
// mean of total[0] / T): a workgroup that starts late on a heavy tile would
// run on alone after the others.  The other tiles share the largest key and
// keep their order in the stable sort, so that neighbouring workgroups still
// read the same column tiles through L2.
constexpr unsigned long long kHeavyRowFactor = 2;
__global__ void k_row_order_keys(const unsigned long long *__restrict__ rows, int64_t T, uint32_t *__restrict__ key,
                                 int32_t *__restrict__ order, const unsigned long long *__restrict__ total) {
    const int64_t R = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (R >= T) return;
    const unsigned long long v = rows[2 * R];
    const bool heavy = v * (unsigned long long)T > kHeavyRowFactor * total[0];
    key[R] = heavy ? ~(uint32_t)(v < 0xffffffffull ? v : 0xffffffffull) : 0xffffffffu;
    order[R] = (int32_t)R;
}

// One workgroup: out[0] = sum of v[2i], out[1] = sum of v[2i + 1], i < n.
__global__ __launch_bounds__(256) void k_sum_pairs_u64(const unsigned long long *__restrict__ v, int64_t n,
                                                       unsigned long long *__restrict__ out) {
    __shared__ unsigned long long s0[4], s1[4];
    unsigned long long a = 0, b = 0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        a += v[2 * i];
        b += v[2 * i + 1];
    }
    a = (unsigned long long)wave_sum64((long long)a);
    b = (unsigned long long)wave_sum64((long long)b);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = s0[0] + s0[1] + s0[2] + s0[3];
        out[1] = s1[0] + s1[1] + s1[2] + s1[3];
    }
}

// ------------------------------------------------------------------ CSR emit

// Coalesced CSR emit: one workgroup per row tile.  The row tile's non-empty
// tile pairs (all pairs when no flags are kept) are gathered in ascending
// order, 64 at a time, and their words staged in LDS (pair-major, rows padded
// to 65 words).  For each of its 16 rows a wave scans the popcounts of the
// row's 64 words (lane = word), then walks the non-zero words: the word is
// broadcast to the wave, lane b holds bit b, and the set bits go out as one
// contiguous store, each at its rank among the word's set bits (mbcnt) after
// the word's prefix.  Rows come out sorted: candidate tiles ascend, then bits.
// Position of the t-th (from 0) set bit of w; t < popcount(w).
__device__ __forceinline__ int select_bit(uint64_t w, int t) {
    int pos = 0;
    uint32_t x = (uint32_t)w;
    int c = __popc(x);
    if (t >= c) {
        t -= c;
        pos = 32;
        x = (uint32_t)(w >> 32);
    }
#pragma unroll
    for (int width = 16; width >= 2; width >>= 1) {
        c = __popc(x & ((1u << width) - 1u));
        if (t >= c) {
            t -= c;
            pos += width;
            x >>= width;
        }
    }
    return pos + (t >= (int)(x & 1u) ? 1 : 0);
}

// Waves per emit workgroup (a template parameter, 4 or 8, chosen per problem
// class by the tuning table): the 33 KB of staged words per workgroup bound
// the resident workgroups per CU, so more waves per workgroup means more
// waves in flight to cover the latency of the per-word chains.
template <bool OUTPUT_PARALLEL, int kEmitWaves>
__global__ __launch_bounds__(64 * kEmitWaves) void k_emit_rows(DistArgs a, const int64_t *__restrict__ row_ptr,
                                                               int32_t *__restrict__ col,
                                                               int32_t *__restrict__ selfpos) {
    constexpr int kRows = 64 / kEmitWaves;  // rows per wave
    __shared__ uint64_t sm[64 * 65];
    __shared__ int64_t s_k[64];
    __shared__ int s_nk;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t R = a.R0 + blockIdx.x;
    int64_t kb, ke;
    tile_range(a, R, kb, ke);
    // lane rr < kRows holds the write position of row wave * kRows + rr
    const int64_t wl = R * 64 + wave * kRows + lane;
    int64_t mypos = (lane < kRows && wl < a.n) ? row_ptr[wl] : 0;
    const int64_t myrow0 = mypos;  // the row's first entry
    int64_t next = kb;  // gathering cursor (wave 0)
    for (;;) {
        __syncthreads();
        if (wave == 0) {
            // windows of 512 pairs, 8 consecutive flags per lane (all loads of a
            // window in flight together)
            int cnt = 0;
            while (cnt < 64 && next < ke) {
                const int64_t b0 = next + (int64_t)lane * 8;
                uint32_t nzb = 0;  // bit q: pair b0 + q holds a word with a bit
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int64_t kk = b0 + q;
                    const bool nz = kk < ke && (a.pair_nz == nullptr || a.pair_nz[kk - a.k_base] != 0);
                    nzb |= (uint32_t)nz << q;
                }
                const int c = __popc(nzb);
                const int incl = wave_incl_scan(c);
                const int excl = incl - c;
                const int total = __shfl(incl, 63, 64);
                const int room = 64 - cnt;
                uint32_t m = nzb;
                for (int rank = excl; m && rank < room; ++rank) {
                    const int q = __ffs(m) - 1;
                    m &= m - 1;
                    s_k[cnt + rank] = b0 + q;
                }
                if (total > room) {
                    // the window holds more than fits: resume after the last pair
                    // taken (in the lane whose ranks hold room - 1)
                    int64_t last = 0;
                    const bool here = excl < room && incl >= room;
                    if (here) {
                        uint32_t mm = nzb;
                        for (int t = 0; t < room - 1 - excl; ++t) mm &= mm - 1;
                        last = b0 + (__ffs(mm) - 1);
                    }
                    const int src = __ffsll((unsigned long long)__ballot(here)) - 1;
                    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)last, src);
                    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)last >> 32), src);
                    next = (int64_t)(((uint64_t)hi << 32) | lo) + 1;
                    cnt = 64;
                } else {
                    cnt += total;
                    next = min(next + 512, ke);
                }
            }
            if (lane == 0) s_nk = cnt;
        }
        __syncthreads();
        const int nk = s_nk;
        if (nk == 0) break;
#pragma unroll 4
        for (int st = 0; st < 64 / kEmitWaves; ++st) {
            const int idx = (int)threadIdx.x + 64 * kEmitWaves * st;
            const int p = idx >> 6, r = idx & 63;
            sm[p * 65 + r] = p < nk ? a.mask[(s_k[p] - a.k_base) * 64 + r] : 0ull;
        }
        __syncthreads();
        const int cbase = lane < nk ? (int)tile_col(a, s_k[lane], kb) * 64 : 0;
        for (int rr = 0; rr < kRows; ++rr) {
            const int r = wave * kRows + rr;
            const uint64_t v = sm[lane * 65 + r];
            const int pc = __popcll(v);
            const int incl = wave_incl_scan(pc);
            const int total = __shfl(incl, 63, 64);
            if (total == 0) continue;
            // the row's first entry as a wave-uniform (scalar) pointer: stores
            // then take a scalar base plus a 32-bit lane offset
            const int64_t base_v = __shfl(mypos, rr, 64);
            const uint32_t blo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base_v);
            const uint32_t bhi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)base_v >> 32));
            int32_t *__restrict__ rowp = col + (int64_t)(((uint64_t)bhi << 32) | blo);
            const int excl = incl - pc;
            const int64_t row = R * 64 + r;  // also the row's own column index
            // entries of the row written by earlier chunks: self positions count from the row start
            const int done = __builtin_amdgcn_readfirstlane((int)(base_v - __shfl(myrow0, rr, 64)));
            if constexpr (OUTPUT_PARALLEL) {
                // output-parallel: lane l writes entry E + l; its word is the first
                // with inclusive count > E + l, its bit the rank within that word
                for (int E = 0; E < total; E += 64) {
                    const int e = E + lane;
                    int lo = 0, hi = 63;
#pragma unroll
                    for (int step = 0; step < 6; ++step) {
                        const int mid = (lo + hi) >> 1;
                        if (__shfl(incl, mid, 64) > e) hi = mid;
                        else lo = mid + 1;
                    }
                    const int q = min(lo, 63);
                    const int t = e - __shfl(excl, q, 64);
                    const uint64_t wq = __shfl(v, q, 64);
                    const int cq = __shfl(cbase, q, 64);
                    if (e < total) {
                        const int v = cq + select_bit(wq, t);
                        rowp[(uint32_t)e] = v;
                        if (v == (int)row) selfpos[row] = done + e;
                    }
                }
                if (lane == rr) mypos += total;
                continue;
            }
            uint64_t nz = __ballot(v != 0ull);
            int eq = 0;  // entries of this chunk's earlier words: a running scalar sum
            while (nz) {
                const int q = __ffsll((unsigned long long)nz) - 1;
                nz ^= 1ull << q;
                const uint32_t wlo = __builtin_amdgcn_readlane((uint32_t)v, q);
                const uint32_t whi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), q);
                const int cq = __builtin_amdgcn_readlane(cbase, q);
                const uint64_t w = ((uint64_t)whi << 32) | wlo;
                // lanes whose bit is set: w (uniform) becomes the exec mask directly
                if (__builtin_amdgcn_inverse_ballot_w64(w)) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi(whi, __builtin_amdgcn_mbcnt_lo(wlo, 0u));
                    // 32-bit byte offset from the scalar row base (a chunk holds < 2^12 entries)
                    const uint32_t boff = (uint32_t)(eq + (int)rank) << 2;
                    *(int32_t *)((char *)rowp + boff) = cq + lane;
                    if (cq + lane == (int)row) selfpos[row] = done + eq + (int)rank;
                }
                eq += __popcll(w);
            }
            if (lane == rr) mypos += total;
        }
    }
}

// Emission of the row tiles R0 .. R0 + n_rt - 1 with `waves` waves per workgroup (4 or 8).
static void launch_emit(hipStream_t s, const DistArgs &a, int64_t n_rt, bool output_parallel, int waves,
                        const int64_t *row_ptr, int32_t *col, int32_t *selfpos) {
    if (n_rt <= 0) return;
    const unsigned g = (unsigned)n_rt;
    if (waves == 4) {
        if (output_parallel) k_emit_rows<true, 4><<<g, 256, 0, s>>>(a, row_ptr, col, selfpos);
        else k_emit_rows<false, 4><<<g, 256, 0, s>>>(a, row_ptr, col, selfpos);
    } else {
        if (output_parallel) k_emit_rows<true, 8><<<g, 512, 0, s>>>(a, row_ptr, col, selfpos);
        else k_emit_rows<false, 8><<<g, 512, 0, s>>>(a, row_ptr, col, selfpos);
    }
    CNN_HIP(hipGetLastError());
}

__global__ void k_deg_from_rowptr_perm(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ perm,
                                       int64_t n, int64_t *__restrict__ deg_o) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    deg_o[perm[w]] = row_ptr[w + 1] - row_ptr[w];
}

// Scatter working rows into caller order with caller column indices.
__global__ __launch_bounds__(256) void k_scatter_rows(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ perm, int64_t n,
                                                      const int64_t *__restrict__ row_ptr_o, int32_t *__restrict__ col_o) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    const int64_t o = row_ptr_o[perm[w]];
    for (int64_t t = b + lane; t < e; t += 64) col_o[o + (t - b)] = perm[col[t]];
}

__global__ void k_set_last(int64_t *p, int64_t idx, const int64_t *prev_excl, const int64_t *prev_val) {
    p[idx] = prev_excl[idx - 1] + prev_val[idx - 1];
}

// p[idx] = p[idx - 1] + cnt[idx - 1]: the total after an exclusive scan of counts.
__global__ void k_scan_total(int64_t *p, int64_t idx, const int32_t *cnt) { p[idx] = p[idx - 1] + cnt[idx - 1]; }

// ------------------------------------------------------------------ host side

static double gamma_m(int m, double u) { return m * u / (1.0 - m * u); }

// Guard band for the FP32 direct form.  |d2_fp32 - s64| <= E where, near the
// threshold (magnitudes D <= 2 r2):
//   FP32 accumulation of d rounded squares of rounded differences: gamma_{d+2}(u32) * D
//   the reference's own FP64 rounding:                           gamma_{2d+2}(u64) * D
//   rounding f64 coordinates to f32 (f64 input only):            2 u32 M sqrt(D) + u32^2 M^2
// with M the largest point norm times two.  A safety factor 2 is applied.
static BandParams make_band_direct(double r2, int d, bool rounded_coords, double max_norm) {
    const double u32 = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    const double D = 2.0 * r2;
    double E = (gamma_m(d + 2, u32) + gamma_m(2 * d + 2, u64)) * D;
    if (rounded_coords) {
        const double M = 2.0 * max_norm;
        E += 2.0 * u32 * M * std::sqrt(D) + u32 * u32 * M * M;
    }
    E = 2.0 * E + 1e-300;
    BandParams bp;
    bp.r2 = r2;
    double lo = r2 - E, hi = r2 + E;
    float flo = (float)lo;
    if ((double)flo > lo) flo = std::nextafter(flo, -INFINITY);
    float fhi = (hi >= (double)FLT_MAX) ? INFINITY : (float)hi;
    if ((double)fhi < hi) fhi = std::nextafter(fhi, INFINITY);
    if (!(lo > 0.0)) flo = -1.0f;  // nothing certain: every pair at or below hi is rechecked
    if (!(r2 == r2)) { flo = -1.0f; fhi = -2.0f; }  // NaN radius: no neighbours at all
    bp.lo = flo;
    bp.hi = fhi;
    return bp;
}

// Margin of the augmented Gram form.  T_hat = row_i . col_j is an FP32 sum of
// dp products in any order (MFMA), so with Sabs = sum |x_ik x_jk| + |c_i| + |c_j|
//   |T_hat - (x_i.x_j + c_i_hat + c_j_hat)| <= gamma_dp Sabs,
//   Sabs <= (w_i + w_j) + r2/2,
// and c_hat = r2/4 - w_hat/2 carries gamma_d w/2 from the norm plus u r2/4 +
// u |c| from rounding r2/4 and the subtraction.  In T units (T = (r2 - d2)/2)
// the reference's FP64 rounding adds gamma_{2d+2}(u64) r2 near the threshold,
// and rounding f64 coordinates to f32 adds u (w_i + w_j) + u r2.  Hence
//   e_i = S (A w_i + B r2 / 2),  A = gamma_dp + gamma_d/2 + 2u (+u),
//                                B = gamma_dp/2 + 2u + gamma64 (+u)
// gives e_i + e_j >= the total error with S = 1; S = 1 + 2 gamma_d + 8u also
// covers evaluating e in FP32 (w_hat, ka, the FMA) and the FP32 sum E = max e
// over the row tile + max e over the column tile >= e_i + e_j.  So T_hat >= E
// proves d2 <= r2 and T_hat < -E proves d2 > r2 in the reference's FP64.
static void make_band_gram(int d, int dp, bool rounded_coords, double r2, float &ka, float &kb) {
    const double u = std::ldexp(1.0, -24), u64 = std::ldexp(1.0, -53);
    double A = gamma_m(dp, u) + 0.5 * gamma_m(d, u) + 2.0 * u;
    double B = 0.5 * gamma_m(dp, u) + 2.0 * u + gamma_m(2 * d + 2, u64);
    if (rounded_coords) {
        A += u;
        B += u;
    }
    const double S = 1.0 + 2.0 * gamma_m(d, u) + 8.0 * u;
    ka = (float)(S * A);
    kb = (float)(S * B * 0.5 * r2 + 1e-30);
}

// Padded width of the wide FP64 kernel (dense layout), 0 if unused.
static int wide_kmax(const DistArgs &a) {
    if ((a.tp_ptr && !a.live) || a.d <= 8 || a.d > 64) return 0;
    return a.d <= 8 ? 8 : a.d <= 16 ? 16 : a.d <= 32 ? 32 : 64;
}

// Events 6 and 7 bracket the main distance kernel (without guard-band settling).
static void launch_dist(cnn_ctx *ctx, const DistArgs &a, int64_t n_rt, bool band, int64_t batch_pairs,
                        int mfma_col_chunk) {
    dim3 grid((unsigned)n_rt), block(256);
    hipStream_t s = ctx->stream;
    CNN_HIP(hipEventRecord(ctx->ev[6], s));
    if (a.bounds) {
        k_dist_f64_periodic<<<grid, block, 0, s>>>(a);
        CNN_HIP(hipGetLastError());
        CNN_HIP(hipEventRecord(ctx->ev[7], s));
        return;
    }
    if (band && a.d >= kMfmaMinD) {
        // guard-band elements are listed (and, when the list is full, whole tile pairs,
        // at most once each) and settled afterwards
        int64_t cap = std::min<int64_t>(std::max<int64_t>(batch_pairs / 4, 1 << 14), (int64_t)1 << 21);
        if (const char *ev = std::getenv("CNN_HIP_BAND_RECORDS")) cap = std::max<int64_t>(1, std::atoll(ev));
        Scratch<int64_t> cnt(ctx, 2);
        Scratch<int64_t> list(ctx, 66 * cap);
        Scratch<int64_t> redo(ctx, 2 * std::max<int64_t>(batch_pairs, 1));
        fill_i64(ctx, cnt.p, 0, 2);
        DistArgs b = a;
        b.band_cnt = (unsigned long long *)cnt.p;
        b.band_list = list.p;
        b.band_cap = cap;
        b.redo_cnt = (unsigned long long *)(cnt.p + 1);
        b.redo = redo.p;
        const int G = a.dp / 8;
        if ((!a.tp_ptr || a.live) && !a.tile_valid && G >= 3 && G <= 10) {
            // dense pairs, row operands in registers, column tiles shared through LDS
            b.col_chunk = mfma_col_chunk;
            // culled items range from no work to the full chunk: round-robin
            // dealing balances them (measured 5.5 ms against 9.2 ms for
            // contiguous ranges at n = 2e5, d = 64)
            b.deal_rr = a.live ? 1 : 0;
            const int64_t n_rb = ceil_div(n_rt, 4), n_cc = ceil_div(a.n_tiles, b.col_chunk);
            int64_t n_items = 0;
            for (int64_t cc = 0; cc < n_cc; ++cc)
                n_items += mfma_chunk_blocks(b.sym, cc, b.col_chunk, b.n_tiles, b.R0, n_rb);
            const int64_t per_xcd = ceil_div(n_items, 8);
            if (8 * per_xcd > (int64_t)INT32_MAX) throw HipFailure{hipErrorInvalidValue, "MFMA grid too large"};
            const dim3 g8((unsigned)std::max<int64_t>(8 * per_xcd, 1));
            switch (G) {
#define CNN_MFMA_LDS(GG) \
    case GG: k_dist_mfma_lds<GG><<<g8, block, 0, s>>>(b, n_rt, n_rb, n_cc, per_xcd, n_items); break;
                CNN_MFMA_LDS(3)
                CNN_MFMA_LDS(4)
                CNN_MFMA_LDS(5)
                CNN_MFMA_LDS(6)
                CNN_MFMA_LDS(7)
                CNN_MFMA_LDS(8)
                CNN_MFMA_LDS(9)
                default: k_dist_mfma_lds<10><<<g8, block, 0, s>>>(b, n_rt, n_rb, n_cc, per_xcd, n_items); break;
#undef CNN_MFMA_LDS
            }
        } else {
            k_dist_mfma<<<grid, block, 0, s>>>(b);
        }
        CNN_HIP(hipGetLastError());
        CNN_HIP(hipEventRecord(ctx->ev[7], s));
        PinnedReads rd(ctx);
        const int64_t *counts = rd.get(cnt.p, 2);
        rd.wait();
        const int64_t n_elems = std::min(counts[0], cap), n_redo = counts[1];
        if (n_redo > batch_pairs) throw HipFailure{hipErrorUnknown, "band pair queue overflow"};
        if (n_elems > 0) {
            Scratch<int32_t> rcnt(ctx, n_elems);
            Scratch<int64_t> roff(ctx, n_elems);
            k_band_count<<<(unsigned)ceil_div(n_elems, 4), 256, 0, s>>>(list.p, n_elems, rcnt.p);
            CNN_HIP(hipGetLastError());
            exclusive_scan_i32_to_i64(ctx, rcnt.p, roff.p, n_elems);
            const int64_t n_el = read_scalar(ctx, roff.p + (n_elems - 1)) + read_scalar(ctx, rcnt.p + (n_elems - 1));
            if (n_el > 0) {
                Scratch<int64_t> elem(ctx, n_el);
                k_band_list<<<(unsigned)ceil_div(n_elems, 4), 256, 0, s>>>(list.p, n_elems, roff.p, elem.p);
                k_band_settle<<<(unsigned)std::min<int64_t>(ceil_div(n_el, 256), 4096), 256, 0, s>>>(b, elem.p, n_el);
                CNN_HIP(hipGetLastError());
            }
        }
        if (n_redo > 0) k_dist_mfma_band<<<(unsigned)n_redo, 64, 0, s>>>(b, n_redo);
    } else if (!band && wide_kmax(a) > 0) {
        const int kmax = wide_kmax(a);
        const int64_t n_pad = a.n_tiles * 64;
        Scratch<double> xp(ctx, n_pad * kmax);
        k_pad_rows<<<(unsigned)ceil_div(n_pad * kmax, 256), 256, 0, s>>>(a.xw64, n_pad, a.d, kmax, xp.p);
        const int64_t n_rg = ceil_div(n_rt, 4);
        const int64_t n_chunks = ceil_div(a.n_tiles, a.col_chunk);
        dim3 g2((unsigned)(n_rg * n_chunks));
        if (kmax == 8) k_dist_f64_wide<8><<<g2, block, 0, s>>>(a, xp.p, n_chunks, n_rt);
        else if (kmax == 16) k_dist_f64_wide<16><<<g2, block, 0, s>>>(a, xp.p, n_chunks, n_rt);
        else if (kmax == 32) k_dist_f64_wide<32><<<g2, block, 0, s>>>(a, xp.p, n_chunks, n_rt);
        else k_dist_f64_wide<64><<<g2, block, 0, s>>>(a, xp.p, n_chunks, n_rt);
    } else if (!band) {
        switch (a.d) {
            case 1: k_dist_f64_cols<1><<<grid, block, 0, s>>>(a); break;
            case 2: k_dist_f64_cols<2><<<grid, block, 0, s>>>(a); break;
            case 3: k_dist_f64_cols<3><<<grid, block, 0, s>>>(a); break;
            case 4: k_dist_f64_cols<4><<<grid, block, 0, s>>>(a); break;
            case 5: k_dist_f64_cols<5><<<grid, block, 0, s>>>(a); break;
            case 6: k_dist_f64_cols<6><<<grid, block, 0, s>>>(a); break;
            case 7: k_dist_f64_cols<7><<<grid, block, 0, s>>>(a); break;
            case 8: k_dist_f64_cols<8><<<grid, block, 0, s>>>(a); break;
            default: k_dist_f64_generic<<<grid, block, 0, s>>>(a); break;
        }
    } else {
        switch (a.d) {
            case 1: k_dist_f32_cols<1><<<grid, block, 0, s>>>(a); break;
            case 2: k_dist_f32_cols<2><<<grid, block, 0, s>>>(a); break;
            case 3: k_dist_f32_cols<3><<<grid, block, 0, s>>>(a); break;
            case 4: k_dist_f32_cols<4><<<grid, block, 0, s>>>(a); break;
            case 5: k_dist_f32_cols<5><<<grid, block, 0, s>>>(a); break;
            case 6: k_dist_f32_cols<6><<<grid, block, 0, s>>>(a); break;
            case 7: k_dist_f32_cols<7><<<grid, block, 0, s>>>(a); break;
            case 8: k_dist_f32_cols<8><<<grid, block, 0, s>>>(a); break;
            default: k_dist_f32_generic<<<grid, block, 0, s>>>(a); break;
        }
    }
    CNN_HIP(hipGetLastError());
    if (!(band && a.d >= kMfmaMinD)) CNN_HIP(hipEventRecord(ctx->ev[7], s));
}

// Graph memory: from the building context's arena for a graph that lives only
// inside one call on that context (stream-ordered reuse, no hipMalloc/hipFree
// per fit), else hipMalloc.
static void *graph_mem(cnn_graph *g, size_t bytes) {
    void *p = nullptr;
    if (g->arena) {
        p = g->arena->get(bytes);
    } else {
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
    }
    if (!p) throw OutOfMemory{"graph allocation of " + std::to_string(bytes) + " bytes failed"};
    return p;
}

static void graph_mem_free(cnn_graph *g, void *p) {
    if (!p) return;
    if (g->arena) g->arena->put(p);
    else (void)hipFree(p);
}

template <typename T>
static T *dev_alloc(cnn_graph *g, int64_t count) {
    size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    void *p = graph_mem(g, bytes);
    g->bytes += (int64_t)bytes;
    return (T *)p;
}

void graph_free(cnn_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    graph_mem_free(g, g->perm);
    graph_mem_free(g, g->xw64);
    graph_mem_free(g, g->xw32);
    graph_mem_free(g, g->row_ptr);
    graph_mem_free(g, g->col);
    graph_mem_free(g, g->tp_ptr);
    graph_mem_free(g, g->tp_col);
    graph_mem_free(g, g->diag);
    graph_mem_free(g, g->tile_order);
    graph_mem_free(g, g->selfpos);
    graph_mem_free(g, g->mask);
    graph_mem_free(g, g->pair_nz);
    graph_mem_free(g, g->bounds);
    delete g;
}

struct GraphGuard {
    cnn_graph *g;
    ~GraphGuard() { if (g) graph_free(g); }
};

// Per-tile parameters of a packed batch (device arrays, one entry per tile).
struct PackedTiles {
    const int32_t *tile_valid = nullptr;
    const double *tile_r2 = nullptr;
    const double *tile_thr = nullptr;
    const int32_t *seg_lo = nullptr;
    const int32_t *seg_hi = nullptr;
    const float *tile_lo = nullptr;
    const float *tile_hi = nullptr;
};

// Bounding box of all finite coordinates and the largest squared norm.
static void bbox(cnn_ctx *ctx, const void *X, int64_t n, int d, int dtype, std::vector<double> &lo,
                 std::vector<double> &hi, double &max_norm2) {
    hipStream_t s = ctx->stream;
    const int nb = (int)std::min<int64_t>(256, ceil_div(n, 256));
    std::vector<double> part((size_t)nb * std::max(d, 1) * 2), pnorm(nb);
    Scratch<double> dpart(ctx, part.size()), dnorm(ctx, nb);
    if (dtype == CNN_F32)
        k_bbox_partial<float><<<nb, 256, 0, s>>>((const float *)X, n, d, dpart.p, dnorm.p);
    else
        k_bbox_partial<double><<<nb, 256, 0, s>>>((const double *)X, n, d, dpart.p, dnorm.p);
    CNN_HIP(hipGetLastError());
    if ((part.size() + nb) * sizeof(double) + 32 <= cnn_ctx::kPinnedBytes) {
        PinnedReads rd(ctx);
        const double *hp = rd.get(dpart.p, part.size());
        const double *hn = rd.get(dnorm.p, (size_t)nb);
        rd.wait();
        std::copy(hp, hp + part.size(), part.begin());
        std::copy(hn, hn + nb, pnorm.begin());
    } else {
        CNN_HIP(hipMemcpyAsync(part.data(), dpart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipMemcpyAsync(pnorm.data(), dnorm.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipStreamSynchronize(s));
    }
    lo.assign(d, INFINITY);
    hi.assign(d, -INFINITY);
    max_norm2 = 0.0;
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < d; ++k) {
            lo[k] = std::fmin(lo[k], part[((size_t)b * d + k) * 2]);
            hi[k] = std::fmax(hi[k], part[((size_t)b * d + k) * 2 + 1]);
        }
        max_norm2 = std::fmax(max_norm2, pnorm[b]);
    }
}

// Morton key over the (up to) three widest dims.  Bits per dimension: about 2^(3 nk) cells per point (ceil(log2 n) / nk + 3),
// so points rarely share a cell and the order is that of a full-resolution
// key, while the radix sort runs over fewer key bits.
static KeySpec make_keyspec(int d, const std::vector<double> &lo, const std::vector<double> &hi, int max_bits,
                            int64_t n) {
    KeySpec ks{};
    std::vector<int> dims(d);
    for (int k = 0; k < d; ++k) dims[k] = k;
    auto ext = [&](int k) { return (hi[k] >= lo[k]) ? hi[k] - lo[k] : 0.0; };
    std::stable_sort(dims.begin(), dims.end(), [&](int x, int y) { return ext(x) > ext(y); });
    ks.nk = std::min(d, 3);
    int lg = 1;
    while (lg < 62 && ((int64_t)1 << lg) < n) ++lg;
    const int bits = std::min({max_bits / ks.nk, ks.nk == 1 ? 62 : (ks.nk == 2 ? 32 : 21), (lg + ks.nk - 1) / ks.nk + 3});
    ks.bits = bits;
    ks.qmax = std::ldexp(1.0, bits) - 1.0;
    for (int t = 0; t < ks.nk; ++t) {
        int k = dims[t];
        ks.dim[t] = k;
        double e = ext(k);
        ks.lo[t] = std::isfinite(lo[k]) ? lo[k] : 0.0;
        ks.scale[t] = (e > 0.0 && std::isfinite(e)) ? ks.qmax / e : 0.0;
    }
    return ks;
}

// Candidate tile pairs, distance masks and CSR for a graph whose working
// coordinates (xw64, xw32), perm and n_tiles are set.  pk: per-tile
// parameters of a packed batch (nullptr: one dataset with radius g->r2).
static void build_pairs_and_csr(cnn_ctx *ctx, cnn_graph *g, bool band, int dtype, double max_norm2,
                                const PackedTiles *pk, const BandParams *bp_in) {
    hipStream_t s = ctx->stream;
    const int64_t n = g->n;
    const int d = g->d;
    const int64_t T = g->n_tiles;
    const int64_t n_pad = T * 64;
    const bool gram = band && d >= kMfmaMinD;
    const int dp = gram ? ((d + 2 + 7) / 8) * 8 : d;  // augmented MFMA rows: d + 2 dims, padded to 8
    std::unique_ptr<Scratch<float>> xra, xca, emarg;
    if (gram) {
        float gka = 0.0f, gkb = 0.0f;
        make_band_gram(d, dp, dtype == CNN_F64, g->r2, gka, gkb);
        xra.reset(new Scratch<float>(ctx, n_pad * dp));
        xca.reset(new Scratch<float>(ctx, n_pad * dp));
        emarg.reset(new Scratch<float>(ctx, n_pad));
        Scratch<float> caug(ctx, n_pad);
        k_augment_rows<<<(unsigned)ceil_div(n_pad, 256), 256, 0, s>>>(g->xw32, n, n_pad, d, (float)(0.25 * g->r2), gka,
                                                                      gkb, emarg->p, caug.p);
        k_augment_pack<<<(unsigned)ceil_div(n_pad * 2 * (dp / 8), 256), 256, 0, s>>>(
            g->xw32, n, n_pad, d, dp / 8, caug.p, (float4 *)xra->p, (float4 *)xca->p);
        CNN_HIP(hipGetLastError());
    }

    // row-tile batches are bounded by the mask budget (512 B per tile pair);
    // one batch is evaluated one-sided (C >= R)
    int64_t mask_budget = (int64_t)8 << 30;
    if (const char *ev = std::getenv("CNN_HIP_MASK_BUDGET")) mask_budget = std::max<int64_t>(512, std::atoll(ev));
    const int64_t max_pairs_per_batch = std::max<int64_t>(mask_budget / 512, T);

    // candidate tile pairs
    int64_t *tp_ptr = nullptr;
    int32_t *tp_col = nullptr;
    int64_t total_pairs = T * T;
    const double thr = g->r2 * (1.0 + 1e-6);
    std::unique_ptr<Scratch<double>> glo, ghi;
    std::unique_ptr<Scratch<GroupMask>> gmask;
    unsigned long long rows_evaluated[2] = {0, 0};  // spatial layout: rows evaluated (C >= R, all pairs)
    std::unique_ptr<Scratch<unsigned long long>> rows_eval;  // (read with the final statistics)
    std::unique_ptr<Scratch<int64_t>> mirror;
    if (g->layout == 1) {
        Scratch<double> blo(ctx, T * d), bhi(ctx, T * d);
        glo.reset(new Scratch<double>(ctx, T * kRowGroups * d));
        ghi.reset(new Scratch<double>(ctx, T * kRowGroups * d));
        k_tile_boxes<<<(unsigned)ceil_div(T, 4), 256, 0, s>>>(g->xw64, n, d, T, blo.p, bhi.p, glo->p, ghi->p);
        CNN_HIP(hipGetLastError());
        const int64_t NS = ceil_div(T, 64);
        Scratch<double> slo(ctx, NS * d), shi(ctx, NS * d);
        k_super_boxes<<<(unsigned)ceil_div(NS, 4), 256, 0, s>>>(blo.p, bhi.p, d, T, NS, slo.p, shi.p);
        CNN_HIP(hipGetLastError());
        Scratch<int32_t> cnt(ctx, T);
        const double *tthr = pk ? pk->tile_thr : nullptr;
        const int32_t *sl = pk ? pk->seg_lo : nullptr, *sh = pk ? pk->seg_hi : nullptr;
        // one search: counts and lists (in per-tile buckets), then the lists
        // move to their places once the counts are scanned
        // kept with the graph: the label pass's tile hooking starts at it
        int32_t *diag = g->diag = dev_alloc<int32_t>(g, T);
        fill_i32(ctx, diag, -1, T);
        {
            int B = (int)std::max<int64_t>(
                16, std::min<int64_t>(kPairBucketMax, kPairBucketBytes / (4 * std::max<int64_t>(T, 1))));
            if (const char *ev = std::getenv("CNN_HIP_PAIR_BUCKET")) B = std::max(1, std::atoi(ev));  // tests
            Scratch<int32_t> bucket(ctx, T * B);
            k_tile_pairs<<<(unsigned)ceil_div(T, 4), 256, 0, s>>>(blo.p, bhi.p, slo.p, shi.p, d, T, NS, thr, tthr,
                                                                  sl, sh, cnt.p, bucket.p, B, diag);
            CNN_HIP(hipGetLastError());
            tp_ptr = g->tp_ptr = dev_alloc<int64_t>(g, T + 1);
            exclusive_scan_i32_to_i64(ctx, cnt.p, tp_ptr, T);
            k_scan_total<<<1, 1, 0, s>>>(tp_ptr, T, cnt.p);
            total_pairs = read_scalar(ctx, tp_ptr + T);
            tp_col = g->tp_col = dev_alloc<int32_t>(g, total_pairs);
            k_tile_pairs_place<<<(unsigned)ceil_div(T, 4), 256, 0, s>>>(blo.p, bhi.p, slo.p, shi.p, d, T, NS, thr,
                                                                        tthr, sl, sh, cnt.p, bucket.p, B, tp_ptr,
                                                                        tp_col);
            CNN_HIP(hipGetLastError());
        }
        gmask.reset(new Scratch<GroupMask>(ctx, std::max<int64_t>(total_pairs, 1)));
        mirror.reset(new Scratch<int64_t>(ctx, std::max<int64_t>(total_pairs, 1)));
        Scratch<unsigned long long> rows_part(ctx, 2 * T);
        rows_eval.reset(new Scratch<unsigned long long>(ctx, 2));
        fill_u64(ctx, rows_part.p, 0ull, 2 * T);
        if (total_pairs <= max_pairs_per_batch) {
            // one-sided: chunks of kPairChunk pairs from each row tile's own pair on
            Scratch<int32_t> ccnt(ctx, T);
            Scratch<int64_t> coff(ctx, T);
            k_pair_chunks<<<(unsigned)ceil_div(T, 256), 256, 0, s>>>(tp_ptr, diag, T, ccnt.p);
            CNN_HIP(hipGetLastError());
            exclusive_scan_i32_to_i64(ctx, ccnt.p, coff.p, T);
            const int64_t waves = T + total_pairs / kPairChunk + 1;  // bounds the chunks
            k_pair_info<<<(unsigned)ceil_div(waves, 4), 256, 0, s>>>(tp_ptr, tp_col, T, mirror->p, glo->p, ghi->p,
                                                                   blo.p, bhi.p, d, thr, tthr, gmask->p, rows_part.p,
                                                                   diag, 1, coff.p, ccnt.p);
        } else {
            k_pair_info<<<(unsigned)ceil_div(2 * T, 4), 256, 0, s>>>(tp_ptr, tp_col, T, mirror->p, glo->p, ghi->p,
                                                                   blo.p, bhi.p, d, thr, tthr, gmask->p, rows_part.p,
                                                                   diag, 0, nullptr, nullptr);
        }
        CNN_HIP(hipGetLastError());
        k_sum_pairs_u64<<<1, 256, 0, s>>>(rows_part.p, T, rows_eval->p);
        CNN_HIP(hipGetLastError());
        if (total_pairs <= max_pairs_per_batch) {
            // kept with the graph: the label pass's per-tile kernels take the same order
            g->tile_order = dev_alloc<int32_t>(g, T);
            Scratch<uint32_t> okey(ctx, T);
            k_row_order_keys<<<(unsigned)ceil_div(T, 256), 256, 0, s>>>(rows_part.p, T, okey.p, g->tile_order,
                                                                       rows_eval->p);
            CNN_HIP(hipGetLastError());
            sort_pairs_u32_i32(ctx, okey.p, g->tile_order, T, 32);
        }
    }
    // dense layout in nearest-pivot order: tile pairs out of reach by the pivot
    // bounds are culled (never evaluated, stored as empty pairs)
    std::unique_ptr<Scratch<uint64_t>> live;
    std::unique_ptr<Scratch<int32_t>> live_pre;
    int64_t live_words = 0;
    unsigned long long live_count[2] = {0, 0};
    if (g->layout == 0 && g->perm && !g->bounds && !pk && T >= 2) {
        Scratch<float2> bnd(ctx, (int64_t)kCullPivots * T);
        Scratch<double> piv(ctx, (int64_t)kCullPivots * d);
        Scratch<int32_t> home(ctx, T);
        k_cull_pivots<<<kCullPivots, 64, 0, s>>>(g->xw64, n, d, T, piv.p);
        k_tile_pivot_bounds<<<(unsigned)T, 256, 0, s>>>(g->xw64, n, d, T, piv.p, bnd.p, home.p);
        CNN_HIP(hipGetLastError());
        live_words = ceil_div(T, 64);
        live.reset(new Scratch<uint64_t>(ctx, T * live_words));
        Scratch<unsigned long long> lcnt(ctx, 2);
        fill_u64(ctx, lcnt.p, 0ull, 2);
        const double rr = std::sqrt(g->r2) * (1.0 + 1e-6);
        float rc = (float)rr;
        if ((double)rc < rr) rc = std::nextafter(rc, INFINITY);
        k_dense_live<<<(unsigned)std::min<int64_t>(ceil_div(T * live_words, 4), 2048), 256, 0, s>>>(
            bnd.p, home.p, T, live_words, rc, live->p, lcnt.p);
        k_live_mirror<<<(unsigned)ceil_div(live_words * live_words, 4), 256, 0, s>>>(live->p, T, live_words);
        CNN_HIP(hipGetLastError());
        {
            PinnedReads rd(ctx);
            const unsigned long long *lc = rd.get(lcnt.p, 2);
            rd.wait();
            live_count[1] = lc[1];
            live_count[0] = 2 * lc[1] - (unsigned long long)T;  // the diagonal pairs are live and counted once
        }
        // the live pairs become the graph's explicit tile-pair lists: the
        // words of culled pairs are never stored, and the label pass can test
        // similarities on the words as for the spatial layout
        live_pre.reset(new Scratch<int32_t>(ctx, T * live_words));
        Scratch<int32_t> cnt(ctx, T);
        k_live_rows<<<(unsigned)ceil_div(T, 4), 256, 0, s>>>(live->p, T, live_words, live_pre->p, cnt.p);
        CNN_HIP(hipGetLastError());
        tp_ptr = g->tp_ptr = dev_alloc<int64_t>(g, T + 1);
        exclusive_scan_i32_to_i64(ctx, cnt.p, tp_ptr, T);
        k_scan_total<<<1, 1, 0, s>>>(tp_ptr, T, cnt.p);
        total_pairs = (int64_t)live_count[0];
        tp_col = g->tp_col = dev_alloc<int32_t>(g, std::max<int64_t>(total_pairs, 1));
        mirror.reset(new Scratch<int64_t>(ctx, std::max<int64_t>(total_pairs, 1)));
        k_live_lists<<<(unsigned)ceil_div(T, 4), 256, 0, s>>>(live->p, T, live_words, live_pre->p, tp_ptr, tp_col,
                                                             mirror->p);
        CNN_HIP(hipGetLastError());
    }
    // row-tile batches bounded by the mask budget (max_pairs_per_batch); the
    // row offsets of the pair lists are read back only when several batches
    // are needed
    const bool one_batch = total_pairs <= max_pairs_per_batch;
    std::vector<int64_t> h_tp_ptr;
    if (tp_ptr && !one_batch) {
        h_tp_ptr.resize(T + 1);
        CNN_HIP(hipMemcpyAsync(h_tp_ptr.data(), tp_ptr, (T + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipStreamSynchronize(s));
    }
    auto pair_begin = [&](int64_t R) -> int64_t {
        if (R <= 0) return 0;
        if (R >= T) return total_pairs;
        return tp_ptr ? h_tp_ptr[R] : R * T;
    };
    CNN_HIP(hipEventRecord(ctx->ev[1], s));

    BandParams bp;
    if (bp_in) {
        bp = *bp_in;
    } else if (band) {
        bp = make_band_direct(g->r2, d, dtype == CNN_F64, std::sqrt(max_norm2));
    } else {
        bp.r2 = g->r2;
        bp.lo = 0.0f;
        bp.hi = 0.0f;
    }
    Scratch<unsigned long long> bstats(ctx, 4);
    fill_u64(ctx, bstats.p, 0ull, 4);
    Scratch<int32_t> deg(ctx, n);

    std::vector<std::pair<int64_t, int64_t>> batches;
    if (one_batch) batches.push_back({0, T});
    for (int64_t R = one_batch ? T : 0; R < T;) {
        int64_t R1 = R + 1;
        while (R1 < T && pair_begin(R1 + 1) - pair_begin(R) <= max_pairs_per_batch) ++R1;
        batches.push_back({R, R1});
        R = R1;
    }
    int64_t max_batch_pairs = 0;
    for (auto &b : batches) max_batch_pairs = std::max(max_batch_pairs, pair_begin(b.second) - pair_begin(b.first));
    // symmetric evaluation needs every mask of the graph in one buffer
    const int sym = batches.size() == 1 ? 1 : 0;
    // a single-batch build with explicit tile-pair lists (spatial, or dense
    // after culling) keeps its words: the label pass walks a row's lists tile
    // by tile from them
    uint64_t *mask_p = nullptr;
    uint8_t *pair_nz_p = nullptr;
    std::unique_ptr<Scratch<uint64_t>> mask_s;
    std::unique_ptr<Scratch<uint8_t>> pair_nz_s;
    if (sym && tp_ptr) {
        mask_p = g->mask = dev_alloc<uint64_t>(g, std::max<int64_t>(max_batch_pairs, 1) * 64);
        pair_nz_p = g->pair_nz = dev_alloc<uint8_t>(g, std::max<int64_t>(max_batch_pairs, 1));
    } else {
        mask_s.reset(new Scratch<uint64_t>(ctx, std::max<int64_t>(max_batch_pairs, 1) * 64));
        pair_nz_s.reset(new Scratch<uint8_t>(ctx, std::max<int64_t>(max_batch_pairs, 1)));
        mask_p = mask_s->p;
        pair_nz_p = pair_nz_s->p;
    }
    float ms_dist = 0.0f, ms_emit = 0.0f, ms_kernel = 0.0f;
    bool emitted = false;  // the last batch emitted lists (events 4, 5 hold its emission)
    auto add_batch_times = [&]() {
        float t1 = 0, t2 = 0, t3 = 0;
        CNN_HIP(hipEventElapsedTime(&t1, ctx->ev[2], ctx->ev[3]));
        CNN_HIP(hipEventElapsedTime(&t3, ctx->ev[6], ctx->ev[7]));
        if (emitted) CNN_HIP(hipEventElapsedTime(&t2, ctx->ev[4], ctx->ev[5]));
        ms_dist += t1;
        ms_emit += t2;
        ms_kernel += t3;
    };
    int64_t nnz_total = 0;
    int64_t col_cap = 0;
    int32_t *col = nullptr;
    // A single-batch graph with explicit tile-pair lists whose words take no
    // more memory than its lists would (512 bytes per tile pair against 4 per
    // entry) keeps only the words: similarities are tested on them and the
    // lists are emitted when a caller needs them (graph_lists).
    bool words_only = false;
    const bool words_deferred = sym && tp_ptr && !pk && g->mask && batches.size() == 1;
    for (auto &b : batches) {
        DistArgs a{};
        a.xw64 = g->xw64;
        a.xw32 = g->xw32;
        a.n = n;
        a.d = d;
        a.R0 = b.first;
        a.n_tiles = T;
        a.tp_ptr = tp_ptr;
        a.tp_col = tp_col;
        a.k_base = pair_begin(b.first);
        a.mask = mask_p;
        a.deg = deg.p;
        a.bp = bp;
        a.band_stats = bstats.p;
        a.dp = dp;
        a.mfma_k = (d + 3) / 2;
        a.sym = sym;
        a.mirror = mirror ? mirror->p : nullptr;
        a.xra = xra ? xra->p : nullptr;
        a.xca = xca ? xca->p : nullptr;
        a.emarg = emarg ? emarg->p : nullptr;
        a.col_chunk = g->tune.wide_col_chunk;
        a.pair_nz = pair_nz_p;
        a.pair_groups = gmask ? gmask->p : nullptr;
        a.row_order = g->tile_order && b.first == 0 && b.second == T ? g->tile_order : nullptr;
        a.bounds = g->bounds;
        if (live) {
            a.live = live->p;
            a.live_pre = live_pre->p;
            a.live_words = live_words;
        }
        if (pk) {
            a.tile_r2 = pk->tile_r2;
            a.tile_lo = pk->tile_lo;
            a.tile_hi = pk->tile_hi;
            a.tile_valid = pk->tile_valid;
        }
        fill_i32(ctx, deg.p, 0, n);
        CNN_HIP(hipEventRecord(ctx->ev[2], s));
        launch_dist(ctx, a, b.second - b.first, band, pair_begin(b.second) - pair_begin(b.first),
                    g->tune.mfma_col_chunk);
        CNN_HIP(hipEventRecord(ctx->ev[3], s));
        // row_ptr for this batch's rows: exclusive scan + running offset
        const int64_t row_lo = b.first * 64, row_hi = std::min<int64_t>(b.second * 64, n);
        const int64_t rows = row_hi - row_lo;
        exclusive_scan_i32_to_i64(ctx, deg.p + row_lo, g->row_ptr + row_lo, rows);
        if (words_deferred) {
            // one batch that may keep its words: row_ptr[n] is completed on the
            // device and the entry count read with the final statistics (no
            // synchronisation here); the lists are emitted after it only if
            // they are the smaller form
            k_scan_total<<<1, 1, 0, s>>>(g->row_ptr, n, deg.p);
            CNN_HIP(hipGetLastError());
            words_only = true;
            emitted = false;
            continue;
        }
        PinnedReads rd(ctx);
        const int64_t *last_excl = rd.get(g->row_ptr + row_hi - 1);
        const int32_t *last_deg = rd.get(deg.p + row_hi - 1);
        rd.wait();
        const int64_t batch_total = *last_excl + *last_deg;
        if (nnz_total) {
            // shift by the entries of earlier batches
            std::vector<int64_t> hbuf(rows);
            CNN_HIP(hipMemcpyAsync(hbuf.data(), g->row_ptr + row_lo, rows * sizeof(int64_t), hipMemcpyDeviceToHost, s));
            CNN_HIP(hipStreamSynchronize(s));
            for (auto &v : hbuf) v += nnz_total;
            CNN_HIP(hipMemcpyAsync(g->row_ptr + row_lo, hbuf.data(), rows * sizeof(int64_t), hipMemcpyHostToDevice, s));
            CNN_HIP(hipStreamSynchronize(s));
        }
        const int64_t need = nnz_total + batch_total;
        words_only = sym && tp_ptr && !pk && g->mask && (int64_t)512 * total_pairs <= (int64_t)4 * need;
        if (words_only) {  // (one batch: its times are read after the final synchronisation)
            emitted = false;
            nnz_total = need;
            continue;
        }
        // every row with a finite point holds itself; rows without one (degree 0) are never core
        if (!g->selfpos) g->selfpos = dev_alloc<int32_t>(g, std::max<int64_t>(n, 1));
        if (need > col_cap || !col) {  // (a graph without any entry still holds its empty lists)
            const int64_t new_cap = batches.size() == 1 ? need : std::max<int64_t>(need, col_cap + col_cap / 2);
            int32_t *nc = nullptr;
            try {
                nc = (int32_t *)graph_mem(g, (size_t)std::max<int64_t>(new_cap, 1) * sizeof(int32_t));
            } catch (const OutOfMemory &) {
                throw OutOfMemory{"neighbour list allocation failed (" + std::to_string(need) + " entries)"};
            }
            if (col && nnz_total) CNN_HIP(hipMemcpyAsync(nc, col, nnz_total * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
            if (col) {
                CNN_HIP(hipStreamSynchronize(s));
                graph_mem_free(g, col);
                g->bytes -= col_cap * (int64_t)sizeof(int32_t);
            }
            col = nc;
            col_cap = new_cap;
            g->bytes += col_cap * (int64_t)sizeof(int32_t);
            g->col = col;
        }
        CNN_HIP(hipEventRecord(ctx->ev[4], s));
        // word-serial emission for the spatial layout (words hold many bits each),
        // output-parallel for the dense layout (few bits per word)
        if (rows > 0)
            launch_emit(s, a, b.second - b.first, g->layout != 1, g->tune.emit_waves, g->row_ptr, col, g->selfpos);
        CNN_HIP(hipEventRecord(ctx->ev[5], s));
        emitted = true;
        if (batches.size() > 1) {  // the events are recorded again by the next batch
            CNN_HIP(hipEventSynchronize(ctx->ev[5]));
            add_batch_times();
        }
        nnz_total = need;
    }
    if (!words_only && g->mask) {
        // the lists are the smaller form: drop the words
        graph_mem_free(g, g->mask);
        graph_mem_free(g, g->pair_nz);
        g->bytes -= std::max<int64_t>(max_batch_pairs, 1) * 64 * (int64_t)sizeof(uint64_t) +
                    std::max<int64_t>(max_batch_pairs, 1);
        g->mask = nullptr;
        g->pair_nz = nullptr;
    }
    if (!words_deferred) set_i64(ctx, g->row_ptr + n, nnz_total);
    PinnedReads rd(ctx);
    const unsigned long long *hb = rd.get(bstats.p, 4);
    const unsigned long long *hre = rows_eval ? rd.get(rows_eval->p, 2) : nullptr;
    const int64_t *hnnz = words_deferred ? rd.get(g->row_ptr + n) : nullptr;
    rd.wait();
    if (hnnz) nnz_total = *hnnz;
    if (hre) {
        rows_evaluated[0] = hre[0];
        rows_evaluated[1] = hre[1];
    }
    if (batches.size() == 1) add_batch_times();
    g->nnz = nnz_total;
    float ms_prep = 0.0f;
    CNN_HIP(hipEventElapsedTime(&ms_prep, ctx->ev[0], ctx->ev[1]));
    g->n_tile_pairs = total_pairs;
    cnn_graph_stats &st = g->stats;
    st.tile_pairs = total_pairs;
    // lists are symmetric and hold the diagonal: the symmetric kernels compute C >= R only
    st.tile_pairs_computed = sym ? (total_pairs + T) / 2 : total_pairs;
    // point pairs computed: row groups out of reach are skipped in the spatial layout
    st.pairs_evaluated = g->layout == 1 ? (int64_t)rows_evaluated[sym ? 0 : 1] * 64
                                        : st.tile_pairs_computed * 64 * 64;
    st.band_hits = (int64_t)hb[0];
    st.band_flips = (int64_t)hb[1];
    st.max_abs_err_band = __builtin_bit_cast(double, (uint64_t)hb[2]);
    st.nnz = nnz_total;
    st.ms_prep = ms_prep;
    st.ms_distance = ms_dist;
    st.ms_emit = ms_emit;
    st.ms_distance_kernel = ms_kernel;
    st.distance_launches = (int32_t)batches.size();
    st.layout = g->layout;
    if (words_deferred && (int64_t)512 * total_pairs > (int64_t)4 * nnz_total) {
        // the lists are the smaller form: emit them from the words, then drop the words
        graph_lists(ctx, g);
        graph_mem_free(g, g->mask);
        graph_mem_free(g, g->pair_nz);
        g->bytes -= std::max<int64_t>(max_batch_pairs, 1) * 64 * (int64_t)sizeof(uint64_t) +
                    std::max<int64_t>(max_batch_pairs, 1);
        g->mask = nullptr;
        g->pair_nz = nullptr;
    }
}

static void check_build_args(const void *X, int64_t n, int32_t d, int32_t dtype, int32_t precision) {
    if (n < 0 || d < 0) throw InvalidArg{"negative size"};
    if (n > 0 && d > 0 && !X) throw InvalidArg{"X is null"};
    if (n >= (int64_t)INT32_MAX - 64 * 4096) throw InvalidArg{"n exceeds the int32 index range"};
    if (dtype != CNN_F32 && dtype != CNN_F64) throw InvalidArg{"dtype must be CNN_F32 or CNN_F64"};
    if (precision != CNN_PLAIN_F64 && precision != CNN_F32_BAND) throw InvalidArg{"unknown precision"};
}

cnn_graph *graph_build(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, double r2,
                       int32_t precision, const double *bounds_host, bool transient) {
    check_build_args(X, n, d, dtype, precision);
    if (bounds_host) precision = CNN_PLAIN_F64;  // the periodic metric has the FP64 form only
    CNN_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    GraphGuard guard{new cnn_graph()};
    cnn_graph *g = guard.g;
    if (transient) g->arena = ctx->arena;
    g->device = ctx->device;
    g->n = n;
    g->n_points = n;
    g->d = d;
    g->dtype = dtype;
    g->precision = precision;
    g->r2 = r2;
    g->radius = std::sqrt(r2);
    g->n_tiles = ceil_div(n, 64);
    g->tune = tune_lookup(d, n, precision);
    const int64_t n_pad = g->n_tiles * 64;
    const bool band = precision == CNN_F32_BAND;

    CNN_HIP(hipEventRecord(ctx->ev[0], s));
    g->row_ptr = dev_alloc<int64_t>(g, n + 1);
    if (n == 0) {
        fill_i64(ctx, g->row_ptr, 0, 1);
        g->col = dev_alloc<int32_t>(g, 1);
        CNN_HIP(hipStreamSynchronize(s));
        g->stats = cnn_graph_stats{};
        return std::exchange(guard.g, nullptr);
    }
    // ordering: Morton order over the (up to) three widest dims for small d,
    // nearest-pivot cells for larger d; the caller's order (all tile pairs)
    // for the periodic metric, whose neighbours wrap around the box
    if (bounds_host) {
        g->bounds = dev_alloc<double>(g, d);
        CNN_HIP(hipMemcpyAsync(g->bounds, bounds_host, d * sizeof(double), hipMemcpyHostToDevice, s));
        CNN_HIP(hipStreamSynchronize(s));
    }
    g->layout = (!bounds_host && d >= 1 && d <= kSpatialMaxD && n > 128) ? 1 : 0;
    // the box feeds the Morton keys and the direct FP32 form's band (largest
    // norm); the plain FP64 and MFMA forms of the dense layout need neither
    std::vector<double> lo, hi;
    double max_norm2 = 0.0;
    if (g->layout == 1 || (band && d < kMfmaMinD)) bbox(ctx, X, n, d, dtype, lo, hi, max_norm2);
    if (g->layout == 1) {
        const KeySpec ks = make_keyspec(d, lo, hi, 64, n);
        g->perm = dev_alloc<int32_t>(g, n);
        const unsigned blocks = (unsigned)ceil_div(n, 256);
        const int key_bits = ks.nk * ks.bits;
        if (key_bits <= 32) {
            // 32-bit keys: half the key traffic, and the sort's 32-bit configuration
            Scratch<uint32_t> keys(ctx, n);
            if (dtype == CNN_F32)
                k_morton<float, uint32_t><<<blocks, 256, 0, s>>>((const float *)X, n, d, ks, keys.p, g->perm);
            else
                k_morton<double, uint32_t><<<blocks, 256, 0, s>>>((const double *)X, n, d, ks, keys.p, g->perm);
            CNN_HIP(hipGetLastError());
            sort_pairs_u32_i32(ctx, keys.p, g->perm, n, key_bits);
        } else {
            Scratch<uint64_t> keys(ctx, n);
            if (dtype == CNN_F32)
                k_morton<float, uint64_t><<<blocks, 256, 0, s>>>((const float *)X, n, d, ks, keys.p, g->perm);
            else
                k_morton<double, uint64_t><<<blocks, 256, 0, s>>>((const double *)X, n, d, ks, keys.p, g->perm);
            CNN_HIP(hipGetLastError());
            sort_pairs_u64_i32(ctx, keys.p, g->perm, n, key_bits);
        }
    } else if (!bounds_host && d > kSpatialMaxD && n >= kPivotMinN) {
        const int L = (int)std::min<int64_t>(256, n / 64);
        const int S = std::min(kSuperPivots, L);
        g->perm = dev_alloc<int32_t>(g, n);
        Scratch<uint32_t> keys(ctx, n);
        Scratch<int32_t> super_of(ctx, L);
        const unsigned blocks = (unsigned)ceil_div(n, 256);
        const int Lp = (L + 31) / 32 * 32;
        Scratch<float> piv(ctx, (int64_t)Lp * kPivotDims), phalf(ctx, Lp);
        const unsigned mblocks = (unsigned)ceil_div(n, 128);  // 32 points per wave
        if (dtype == CNN_F32) {
            k_pivot_super<float><<<1, 256, 0, s>>>((const float *)X, n, d, L, S, super_of.p);
            k_pivot_table<float><<<(unsigned)Lp, 64, 0, s>>>((const float *)X, n, d, L, piv.p, phalf.p);
            k_pivot_cells_mfma<float, uint32_t><<<mblocks, 256, 0, s>>>((const float *)X, n, d, L, piv.p, phalf.p,
                                                                       super_of.p, keys.p, g->perm);
        } else {
            k_pivot_super<double><<<1, 256, 0, s>>>((const double *)X, n, d, L, S, super_of.p);
            k_pivot_table<double><<<(unsigned)Lp, 64, 0, s>>>((const double *)X, n, d, L, piv.p, phalf.p);
            k_pivot_cells_mfma<double, uint32_t><<<mblocks, 256, 0, s>>>((const double *)X, n, d, L, piv.p,
                                                                        phalf.p, super_of.p, keys.p, g->perm);
        }
        CNN_HIP(hipGetLastError());
        sort_pairs_u32_i32(ctx, keys.p, g->perm, n, 12);
    }
    g->xw64 = dev_alloc<double>(g, n_pad * std::max(d, 1));
    if (band) g->xw32 = dev_alloc<float>(g, n_pad * std::max(d, 1));
    g->dp = d;
    {
        const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n_pad * std::max(d, 1), 256), 16384);
        if (dtype == CNN_F32)
            k_gather<float><<<blocks, 256, 0, s>>>((const float *)X, g->perm, n, n_pad, d, g->xw64, g->xw32);
        else
            k_gather<double><<<blocks, 256, 0, s>>>((const double *)X, g->perm, n, n_pad, d, g->xw64, g->xw32);
        CNN_HIP(hipGetLastError());
    }
    build_pairs_and_csr(ctx, g, band, dtype, max_norm2, nullptr, nullptr);
    return std::exchange(guard.g, nullptr);
}

// ------------------------------------------------------------------ packed batches

template <typename Tin>
__global__ void k_morton_packed(const Tin *__restrict__ X, int64_t n, int d, KeySpec ks, const int64_t *__restrict__ offs,
                                int n_sets, int shift, uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int lo = 0, hi = n_sets;  // dataset: last b with offs[b] <= p
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (offs[mid] <= p) lo = mid;
        else hi = mid;
    }
    uint64_t q[3] = {0, 0, 0};
    for (int t = 0; t < ks.nk; ++t) {
        double v = (double)X[p * d + ks.dim[t]];
        double f = (v - ks.lo[t]) * ks.scale[t];
        if (!(f >= 0.0)) f = 0.0;
        if (f > ks.qmax) f = ks.qmax;
        q[t] = (uint64_t)f;
    }
    uint64_t key;
    if (ks.nk == 1) key = q[0];
    else if (ks.nk == 2) key = spread2(q[0]) | (spread2(q[1]) << 1);
    else key = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    keys[p] = ((uint64_t)lo << shift) | key;
    vals[p] = (int32_t)p;
}

// Sorted position -> padded working index: dataset b occupies tiles
// tstart[b] .. tstart[b+1]; its points keep the sorted order.
__global__ void k_place_packed(const uint64_t *__restrict__ keys, const int32_t *__restrict__ vals, int64_t n,
                               int shift, const int64_t *__restrict__ offs, const int64_t *__restrict__ tstart,
                               int32_t *__restrict__ perm) {
    const int64_t sidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= n) return;
    const int b = (int)(keys[sidx] >> shift);
    perm[tstart[b] * 64 + (sidx - offs[b])] = vals[sidx];
}

template <typename Tin>
__global__ void k_gather_packed(const Tin *__restrict__ X, const int32_t *__restrict__ perm, int64_t n_pad, int d,
                                double *__restrict__ xw64, float *__restrict__ xw32) {
    const int64_t total = n_pad * d;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = idx / d;
        const int k = (int)(idx - w * d);
        const int32_t src = perm[w];
        if (src >= 0) {
            const Tin v = X[(int64_t)src * d + k];
            xw64[idx] = (double)v;
            if (xw32) xw32[idx] = (float)v;
        } else {  // padding rows: NaN is never within any radius, not even of itself
            xw64[idx] = NAN;
            if (xw32) xw32[idx] = NAN;
        }
    }
}

cnn_graph *graph_build_packed(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d,
                              int32_t dtype, const double *r2, int32_t precision, PackedHost &ph) {
    const int64_t N = offsets[n_sets];
    check_build_args(X, N, d, dtype, precision);
    if (d < 1 || d > kSpatialMaxD) throw InvalidArg{"packed batches need 1 <= d <= 8"};
    CNN_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    GraphGuard guard{new cnn_graph()};
    cnn_graph *g = guard.g;
    const bool band = precision == CNN_F32_BAND;
    // tiles per dataset
    ph.tile_start.assign(n_sets + 1, 0);
    double r2max = 0.0;
    for (int b = 0; b < n_sets; ++b) {
        ph.tile_start[b + 1] = ph.tile_start[b] + ceil_div(offsets[b + 1] - offsets[b], 64);
        r2max = std::fmax(r2max, r2[b]);
    }
    const int64_t T = ph.tile_start[n_sets];
    if (T * 64 >= (int64_t)INT32_MAX - 64 * 4096) throw InvalidArg{"padded batch exceeds the int32 index range"};
    g->device = ctx->device;
    g->n = T * 64;
    g->n_points = N;
    g->d = d;
    g->dtype = dtype;
    g->precision = precision;
    g->tune = tune_lookup(d, N, precision);
    g->r2 = r2max;
    g->radius = std::sqrt(r2max);
    g->n_tiles = T;
    g->layout = 1;
    g->dp = d;
    const int64_t n_pad = T * 64;
    CNN_HIP(hipEventRecord(ctx->ev[0], s));
    g->row_ptr = dev_alloc<int64_t>(g, n_pad + 1);
    if (N == 0 || T == 0) {
        fill_i64(ctx, g->row_ptr, 0, n_pad + 1);
        g->col = dev_alloc<int32_t>(g, 1);
        CNN_HIP(hipStreamSynchronize(s));
        return std::exchange(guard.g, nullptr);
    }
    std::vector<double> lo, hi;
    double max_norm2 = 0.0;
    bbox(ctx, X, N, d, dtype, lo, hi, max_norm2);
    int ds_bits = 1;
    while ((1 << ds_bits) < n_sets) ++ds_bits;
    int64_t n_max = 0;
    for (int32_t b = 0; b < n_sets; ++b) n_max = std::max(n_max, offsets[b + 1] - offsets[b]);
    const KeySpec ks = make_keyspec(d, lo, hi, 63 - ds_bits, n_max);
    const int key_bits = (int)std::log2(ks.qmax + 1.0) * ks.nk;
    Scratch<int64_t> doffs(ctx, n_sets + 1), dtstart(ctx, n_sets + 1);
    CNN_HIP(hipMemcpyAsync(doffs.p, offsets, (n_sets + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(dtstart.p, ph.tile_start.data(), (n_sets + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    {
        Scratch<uint64_t> keys(ctx, N);
        Scratch<int32_t> vals(ctx, N);
        const unsigned blocks = (unsigned)ceil_div(N, 256);
        if (dtype == CNN_F32)
            k_morton_packed<float><<<blocks, 256, 0, s>>>((const float *)X, N, d, ks, doffs.p, n_sets, key_bits, keys.p, vals.p);
        else
            k_morton_packed<double><<<blocks, 256, 0, s>>>((const double *)X, N, d, ks, doffs.p, n_sets, key_bits, keys.p, vals.p);
        CNN_HIP(hipGetLastError());
        sort_pairs_u64_i32(ctx, keys.p, vals.p, N, std::min(64, key_bits + ds_bits));
        g->perm = dev_alloc<int32_t>(g, n_pad);
        fill_i32(ctx, g->perm, -1, n_pad);
        k_place_packed<<<blocks, 256, 0, s>>>(keys.p, vals.p, N, key_bits, doffs.p, dtstart.p, g->perm);
        CNN_HIP(hipGetLastError());
    }
    g->xw64 = dev_alloc<double>(g, n_pad * d);
    if (band) g->xw32 = dev_alloc<float>(g, n_pad * d);
    {
        const unsigned blocks = (unsigned)std::min<int64_t>(ceil_div(n_pad * d, 256), 16384);
        if (dtype == CNN_F32)
            k_gather_packed<float><<<blocks, 256, 0, s>>>((const float *)X, g->perm, n_pad, d, g->xw64, g->xw32);
        else
            k_gather_packed<double><<<blocks, 256, 0, s>>>((const double *)X, g->perm, n_pad, d, g->xw64, g->xw32);
        CNN_HIP(hipGetLastError());
    }
    // per-tile parameters
    std::vector<double> h_r2(T), h_thr(T);
    std::vector<int32_t> h_lo(T), h_hi(T), h_valid(T);
    std::vector<float> h_blo(T), h_bhi(T);
    ph.tile_ds.resize(T);
    for (int b = 0; b < n_sets; ++b) {
        const BandParams bp = band ? make_band_direct(r2[b], d, dtype == CNN_F64, std::sqrt(max_norm2))
                                   : BandParams{r2[b], 0.0f, 0.0f};
        for (int64_t t = ph.tile_start[b]; t < ph.tile_start[b + 1]; ++t) {
            h_r2[t] = r2[b];
            h_thr[t] = r2[b] * (1.0 + 1e-6);
            h_lo[t] = (int32_t)ph.tile_start[b];
            h_hi[t] = (int32_t)ph.tile_start[b + 1];
            h_blo[t] = bp.lo;
            h_bhi[t] = bp.hi;
            h_valid[t] = (int32_t)std::min<int64_t>(64, offsets[b + 1] - offsets[b] - (t - ph.tile_start[b]) * 64);
            ph.tile_ds[t] = b;
        }
    }
    Scratch<double> d_r2(ctx, T), d_thr(ctx, T);
    Scratch<int32_t> d_lo(ctx, T), d_hi(ctx, T), d_valid(ctx, T);
    Scratch<float> d_blo(ctx, T), d_bhi(ctx, T);
    CNN_HIP(hipMemcpyAsync(d_r2.p, h_r2.data(), T * sizeof(double), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_thr.p, h_thr.data(), T * sizeof(double), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_lo.p, h_lo.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_hi.p, h_hi.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_valid.p, h_valid.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_blo.p, h_blo.data(), T * sizeof(float), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_bhi.p, h_bhi.data(), T * sizeof(float), hipMemcpyHostToDevice, s));
    PackedTiles pk;
    pk.tile_valid = d_valid.p;
    pk.tile_r2 = d_r2.p;
    pk.tile_thr = d_thr.p;
    pk.seg_lo = d_lo.p;
    pk.seg_hi = d_hi.p;
    pk.tile_lo = band ? d_blo.p : nullptr;
    pk.tile_hi = band ? d_bhi.p : nullptr;
    BandParams bp0{r2max, 0.0f, 0.0f};
    build_pairs_and_csr(ctx, g, band, dtype, max_norm2, &pk, &bp0);
    return std::exchange(guard.g, nullptr);
}

void graph_lists(cnn_ctx *ctx, const cnn_graph *gc) {
    cnn_graph *g = const_cast<cnn_graph *>(gc);
    std::lock_guard<std::mutex> lock(g->lists_mu);
    if (g->col) return;
    if (!g->mask || !g->pair_nz || !g->tp_ptr) throw InvalidArg{"graph holds neither neighbour lists nor row words"};
    CNN_HIP(hipSetDevice(g->device));
    hipStream_t s = ctx->stream;
    const int64_t n = g->n, T = g->n_tiles;
    int32_t *col = dev_alloc<int32_t>(g, std::max<int64_t>(g->nnz, 1));
    int32_t *selfpos = nullptr;
    try {
        selfpos = dev_alloc<int32_t>(g, std::max<int64_t>(n, 1));
    } catch (...) {
        graph_mem_free(g, col);
        throw;
    }
    DistArgs a{};
    a.n = n;
    a.d = g->d;
    a.n_tiles = T;
    a.tp_ptr = g->tp_ptr;
    a.tp_col = g->tp_col;
    a.mask = g->mask;
    a.pair_nz = g->pair_nz;
    CNN_HIP(hipEventRecord(ctx->ev[4], s));
    launch_emit(s, a, T, g->layout != 1, g->tune.emit_waves, g->row_ptr, col, selfpos);
    CNN_HIP(hipEventRecord(ctx->ev[5], s));
    // complete before any stream (of any context) reads them
    CNN_HIP(hipEventSynchronize(ctx->ev[5]));
    float ms = 0.0f;
    CNN_HIP(hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]));
    g->stats.ms_emit += ms;
    g->selfpos = selfpos;
    g->col = col;
}

void graph_export(cnn_ctx *ctx, const cnn_graph *g, int64_t *row_ptr, int32_t *col) {
    hipStream_t s = ctx->stream;
    const int64_t n = g->n;
    if (n > 0) graph_lists(ctx, g);
    if (n == 0) {
        fill_i64(ctx, row_ptr, 0, 1);
        return;
    }
    if (!g->perm) {
        CNN_HIP(hipMemcpyAsync(row_ptr, g->row_ptr, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
        if (g->nnz)
            CNN_HIP(hipMemcpyAsync(col, g->col, g->nnz * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        return;
    }
    Scratch<int64_t> deg_o(ctx, n);
    k_deg_from_rowptr_perm<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(g->row_ptr, g->perm, n, deg_o.p);
    CNN_HIP(hipGetLastError());
    exclusive_scan_i64(ctx, deg_o.p, row_ptr, n);
    k_set_last<<<1, 1, 0, s>>>(row_ptr, n, row_ptr, deg_o.p);
    CNN_HIP(hipGetLastError());
    k_scatter_rows<<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(g->row_ptr, g->col, g->perm, n, row_ptr, col);
    CNN_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 31 && ((int64_t)1 << bits) < n) ++bits;
    segmented_sort_i32(ctx, col, g->nnz, row_ptr, n, bits);
}


// ------------------------------------------------------------------ re-threshold

// Keep the entries of a resident graph whose exact FP64 distance (reference
// order) is within the new radius.  N_r'(i) is a subset of N_r(i) for r' <= r.
template <bool FILL>
__global__ __launch_bounds__(256) void k_rethreshold(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                                                     const double *__restrict__ xw64, int64_t n, int d, double r2,
                                                     int32_t *__restrict__ deg, const int64_t *__restrict__ out_ptr,
                                                     int32_t *__restrict__ out_col) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    int64_t out = FILL ? out_ptr[w] : 0;
    for (int64_t t0 = b; t0 < e; t0 += 64) {
        const int64_t t = t0 + lane;
        bool keep = false;
        int32_t j = 0;
        if (t < e) {
            j = col[t];
            keep = exact_d2(xw64, d, w, j) <= r2;
        }
        const uint64_t bits = __ballot(keep);
        if (FILL && keep) out_col[out + __popcll(bits & ((1ull << lane) - 1ull))] = j;
        out += __popcll(bits);
    }
    if (!FILL && lane == 0) deg[w] = (int32_t)out;
}

__global__ void k_f64_to_f32(const double *__restrict__ x, int64_t count, float *__restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (float)x[i];
}

// The graph at r2 built afresh from src's working order and coordinates (no
// ordering step): tile pairs, words and lists as a new build at r2 makes them,
// so the lists are exactly those of a fresh graph.  Used for graphs that hold
// no lists (emitting them only to filter them would cost more than the build)
// and for d > 8, where filtering costs d FP64 steps per list entry while the
// build evaluates tile pairs in FP32 / MFMA.  Working FP32 coordinates are the
// FP64 ones rounded, which equals rounding the caller's values (the FP64
// copies are exact).
static cnn_graph *graph_rebuild(cnn_ctx *ctx, const cnn_graph *src, double r2, int32_t precision) {
    hipStream_t s = ctx->stream;
    GraphGuard guard{new cnn_graph()};
    cnn_graph *g = guard.g;
    g->device = src->device;
    g->n = src->n;
    g->n_points = src->n_points;
    g->d = src->d;
    g->dtype = src->dtype;
    g->precision = precision;
    g->r2 = r2;
    g->radius = std::sqrt(r2);
    g->layout = src->layout;
    g->n_tiles = src->n_tiles;
    g->tune = tune_lookup(src->d, src->n, precision);
    const int64_t n = g->n, n_pad = g->n_tiles * 64;
    const int d = g->d;
    const bool band = precision == CNN_F32_BAND;
    CNN_HIP(hipEventRecord(ctx->ev[0], s));
    g->row_ptr = dev_alloc<int64_t>(g, n + 1);
    if (src->perm) {
        g->perm = dev_alloc<int32_t>(g, n);
        CNN_HIP(hipMemcpyAsync(g->perm, src->perm, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    if (src->bounds) {
        g->bounds = dev_alloc<double>(g, d);
        CNN_HIP(hipMemcpyAsync(g->bounds, src->bounds, d * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    g->xw64 = dev_alloc<double>(g, n_pad * d);
    CNN_HIP(hipMemcpyAsync(g->xw64, src->xw64, n_pad * d * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (band) {
        g->xw32 = dev_alloc<float>(g, n_pad * d);
        if (src->xw32 && src->dp == d) {
            CNN_HIP(hipMemcpyAsync(g->xw32, src->xw32, n_pad * d * sizeof(float), hipMemcpyDeviceToDevice, s));
        } else {
            k_f64_to_f32<<<(unsigned)std::min<int64_t>(ceil_div(n_pad * d, 256), 16384), 256, 0, s>>>(
                g->xw64, n_pad * d, g->xw32);
            CNN_HIP(hipGetLastError());
        }
    }
    g->dp = d;
    std::vector<double> lo, hi;
    double max_norm2 = 0.0;
    if (band && d < kMfmaMinD) bbox(ctx, g->xw64, n, d, CNN_F64, lo, hi, max_norm2);
    build_pairs_and_csr(ctx, g, band, g->dtype, max_norm2, nullptr, nullptr);
    return std::exchange(guard.g, nullptr);
}

cnn_graph *graph_rethreshold(cnn_ctx *ctx, const cnn_graph *src, double r2, int32_t precision) {
    if (precision != CNN_PLAIN_F64 && precision != CNN_F32_BAND) throw InvalidArg{"unknown precision"};
    if (!(r2 <= src->r2)) throw InvalidArg{"re-threshold radius must not exceed the graph radius"};
    CNN_HIP(hipSetDevice(ctx->device));
    if (src->n > 0 && src->d > 0 && !src->bounds && (!src->col || src->d > kSpatialMaxD))
        return graph_rebuild(ctx, src, r2, precision);
    if (src->n > 0) graph_lists(ctx, src);
    hipStream_t s = ctx->stream;
    GraphGuard guard{new cnn_graph()};
    cnn_graph *g = guard.g;
    g->device = src->device;
    g->n = src->n;
    g->d = src->d;
    g->dtype = src->dtype;
    g->precision = precision;
    g->r2 = r2;
    g->radius = std::sqrt(r2);
    g->layout = src->layout;
    g->n_tiles = src->n_tiles;
    g->tune = src->tune;
    const int64_t n = g->n, n_pad = g->n_tiles * 64;
    const int dd = std::max(g->d, 1);
    g->row_ptr = dev_alloc<int64_t>(g, n + 1);
    if (src->perm) {
        g->perm = dev_alloc<int32_t>(g, n);
        CNN_HIP(hipMemcpyAsync(g->perm, src->perm, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    g->xw64 = dev_alloc<double>(g, n_pad * dd);
    CNN_HIP(hipMemcpyAsync(g->xw64, src->xw64, n_pad * dd * sizeof(double), hipMemcpyDeviceToDevice, s));
    g->dp = src->dp;
    if (src->xw32) {
        const int64_t words = n_pad * std::max(src->dp, 1);
        g->xw32 = dev_alloc<float>(g, words);
        CNN_HIP(hipMemcpyAsync(g->xw32, src->xw32, words * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
    if (src->tp_ptr) {
        // candidate tiles at r stay valid (a superset) at r' <= r
        const int64_t T = g->n_tiles;
        const int64_t pairs = src->n_tile_pairs;
        g->n_tile_pairs = pairs;
        g->tp_ptr = dev_alloc<int64_t>(g, T + 1);
        g->tp_col = dev_alloc<int32_t>(g, pairs);
        CNN_HIP(hipMemcpyAsync(g->tp_ptr, src->tp_ptr, (T + 1) * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
        CNN_HIP(hipMemcpyAsync(g->tp_col, src->tp_col, pairs * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        if (src->diag) {
            g->diag = dev_alloc<int32_t>(g, T);
            CNN_HIP(hipMemcpyAsync(g->diag, src->diag, T * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        }
    }
    if (n == 0) {
        fill_i64(ctx, g->row_ptr, 0, 1);
        g->col = dev_alloc<int32_t>(g, 1);
        CNN_HIP(hipStreamSynchronize(s));
        return std::exchange(guard.g, nullptr);
    }
    CNN_HIP(hipEventRecord(ctx->ev[2], s));
    Scratch<int32_t> deg(ctx, n);
    k_rethreshold<false><<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(src->row_ptr, src->col, src->xw64, n, g->d, r2, deg.p,
                                                                  nullptr, nullptr);
    CNN_HIP(hipGetLastError());
    exclusive_scan_i32_to_i64(ctx, deg.p, g->row_ptr, n);
    const int64_t last_excl = read_scalar(ctx, g->row_ptr + (n - 1));
    const int32_t last_deg = read_scalar(ctx, deg.p + (n - 1));
    g->nnz = last_excl + last_deg;
    set_i64(ctx, g->row_ptr + n, g->nnz);
    g->col = dev_alloc<int32_t>(g, g->nnz);
    k_rethreshold<true><<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(src->row_ptr, src->col, src->xw64, n, g->d, r2, nullptr,
                                                                 g->row_ptr, g->col);
    CNN_HIP(hipGetLastError());
    CNN_HIP(hipEventRecord(ctx->ev[3], s));
    CNN_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    CNN_HIP(hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]));
    g->stats = cnn_graph_stats{};
    g->stats.nnz = g->nnz;
    g->stats.pairs_evaluated = src->nnz;
    g->stats.ms_distance = ms;
    g->stats.layout = g->layout;
    return std::exchange(guard.g, nullptr);
}

// Row words at a smaller radius from row words of the same spatial graph's
// tile pairs (the graph's own, or an earlier filter's): every set bit is
// re-evaluated exactly in FP64 in reference order, so the result is the graph
// re-thresholded at r2 (N_r'(i) is a subset of N_r(i) for r' <= r) in the same
// pair-major layout.  One wave per tile pair (its row tile found by a binary
// search over tp_ptr), lane = row; degrees are added to deg (zeroed first).
__global__ __launch_bounds__(256) void k_word_filter(const double *__restrict__ xw64, int d, int64_t n_tiles,
                                                     int64_t n_pairs, const int64_t *__restrict__ tp_ptr,
                                                     const int32_t *__restrict__ tp_col,
                                                     const uint64_t *__restrict__ src, const uint8_t *__restrict__ src_nz,
                                                     double r2, uint64_t *__restrict__ dst, uint8_t *__restrict__ dst_nz,
                                                     int32_t *__restrict__ deg) {
    const int lane = threadIdx.x & 63;
    const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= n_pairs) return;
    if (src_nz[k] == 0) {
        if (lane == 0) dst_nz[k] = 0;
        return;
    }
    int64_t lo = 0, hi = n_tiles;  // row tile R: tp_ptr[R] <= k < tp_ptr[R + 1]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (tp_ptr[mid] <= k) lo = mid;
        else hi = mid;
    }
    const int64_t R = lo, C = tp_col[k];
    const int64_t w = R * 64 + lane;
    uint64_t v = src[k * 64 + lane], keep = 0ull;
    while (v) {
        const int b = __ffsll((unsigned long long)v) - 1;
        v &= v - 1;
        if (exact_d2(xw64, d, w, C * 64 + b) <= r2) keep |= 1ull << b;
    }
    const bool any = __ballot(keep != 0ull) != 0ull;
    if (lane == 0) dst_nz[k] = any ? 1 : 0;
    if (any) dst[k * 64 + lane] = keep;
    const int c = __popcll(keep);
    if (c) atomicAdd(&deg[w], c);
}

// Row degrees of a graph from its row offsets.
__global__ void k_deg_of_rows(const int64_t *__restrict__ row_ptr, int64_t n, int32_t *__restrict__ deg) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < n) deg[w] = (int32_t)(row_ptr[w + 1] - row_ptr[w]);
}

void words_filter(cnn_ctx *ctx, const cnn_graph *g, const uint64_t *src, const uint8_t *src_nz, double r2,
                  uint64_t *dst, uint8_t *dst_nz, int32_t *deg) {
    hipStream_t s = ctx->stream;
    fill_i32(ctx, deg, 0, g->n);
    if (g->n_tile_pairs > 0)
        k_word_filter<<<(unsigned)ceil_div(g->n_tile_pairs, 4), 256, 0, s>>>(
            g->xw64, g->d, g->n_tiles, g->n_tile_pairs, g->tp_ptr, g->tp_col, src, src_nz, r2, dst, dst_nz, deg);
    CNN_HIP(hipGetLastError());
}

void words_degrees(cnn_ctx *ctx, const cnn_graph *g, int32_t *deg) {
    if (g->n > 0) k_deg_of_rows<<<(unsigned)ceil_div(g->n, 256), 256, 0, ctx->stream>>>(g->row_ptr, g->n, deg);
    CNN_HIP(hipGetLastError());
}

// Neighbour lists at r2 from lists at a larger radius (rows of the working
// order): each entry is re-evaluated exactly in FP64.  out_rp has n+1
// entries; out_col is allocated here (hipMalloc) and owned by the caller.
int64_t csr_filter(cnn_ctx *ctx, const double *xw64, int64_t n, int d, const int64_t *rp, const int32_t *col,
                   double r2, int64_t *out_rp, int32_t **out_col) {
    hipStream_t s = ctx->stream;
    *out_col = nullptr;
    if (n == 0) {
        fill_i64(ctx, out_rp, 0, 1);
        CNN_HIP(hipMalloc(out_col, sizeof(int32_t)));
        return 0;
    }
    Scratch<int32_t> deg(ctx, n);
    k_rethreshold<false><<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(rp, col, xw64, n, d, r2, deg.p, nullptr, nullptr);
    CNN_HIP(hipGetLastError());
    exclusive_scan_i32_to_i64(ctx, deg.p, out_rp, n);
    const int64_t last_excl = read_scalar(ctx, out_rp + (n - 1));
    const int32_t last_deg = read_scalar(ctx, deg.p + (n - 1));
    const int64_t nnz = last_excl + last_deg;
    set_i64(ctx, out_rp + n, nnz);
    hipError_t e = hipMalloc(out_col, (size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out_col = nullptr;
        throw OutOfMemory{"neighbour list allocation failed (" + std::to_string(nnz) + " entries)"};
    }
    k_rethreshold<true><<<(unsigned)ceil_div(n, 4), 256, 0, s>>>(rp, col, xw64, n, d, r2, nullptr, out_rp, *out_col);
    CNN_HIP(hipGetLastError());
    CNN_HIP(hipStreamSynchronize(s));  // nnz lives on this stack frame
    return nnz;
}

}  // namespace cnn

// ================================================================== C-ABI: phase N
using namespace cnn;

extern "C" {

int cnn_graph_build(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, double r2,
                    int32_t precision, cnn_graph **out) {
    if (!ctx || !out) return CNN_ERR_INVALID;
    *out = nullptr;
    return guarded(ctx, [&]() {
        *out = graph_build(ctx, X, n, d, dtype, r2, precision);
        return (int)CNN_OK;
    });
}

int cnn_graph_build_periodic(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, double r2,
                             const double *bounds, cnn_graph **out) {
    if (!ctx || !out) return CNN_ERR_INVALID;
    *out = nullptr;
    return guarded(ctx, [&]() {
        if (!bounds) throw InvalidArg{"null bounds"};
        *out = graph_build(ctx, X, n, d, dtype, r2, CNN_PLAIN_F64, bounds);
        return (int)CNN_OK;
    });
}

int cnn_graph_destroy(cnn_graph *g) {
    if (!g) return CNN_ERR_INVALID;
    graph_free(g);
    return CNN_OK;
}

int cnn_graph_nnz(const cnn_graph *g, int64_t *nnz) {
    if (!g || !nnz) return CNN_ERR_INVALID;
    *nnz = g->nnz;
    return CNN_OK;
}

int cnn_graph_stats_get(const cnn_graph *g, cnn_graph_stats *out) {
    if (!g || !out) return CNN_ERR_INVALID;
    *out = g->stats;
    return CNN_OK;
}

int cnn_graph_bytes(const cnn_graph *g, int64_t *bytes) {
    if (!g || !bytes) return CNN_ERR_INVALID;
    *bytes = g->bytes;
    return CNN_OK;
}

int cnn_graph_export_csr(cnn_ctx *ctx, const cnn_graph *g, int64_t *row_ptr, int32_t *col, int64_t col_cap) {
    if (!ctx || !g || !row_ptr) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (col_cap < g->nnz) throw Capacity{"col capacity " + std::to_string(col_cap) + " < nnz " + std::to_string(g->nnz)};
        if (g->nnz > 0 && !col) throw InvalidArg{"col is null"};
        CNN_HIP(hipSetDevice(ctx->device));
        graph_export(ctx, g, row_ptr, col);
        return (int)CNN_OK;
    });
}

int cnn_radius_csr(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, double r2,
                   int32_t precision, int64_t *row_ptr, int32_t *col, int64_t col_cap, int64_t *nnz_out,
                   cnn_graph_stats *stats_out) {
    if (!ctx || !row_ptr || !nnz_out) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        cnn_graph *g = graph_build(ctx, X, n, d, dtype, r2, precision, nullptr, true);
        GraphGuard guard{g};
        *nnz_out = g->nnz;
        if (stats_out) *stats_out = g->stats;
        if (col_cap < g->nnz) throw Capacity{"col capacity " + std::to_string(col_cap) + " < nnz " + std::to_string(g->nnz)};
        if (g->nnz > 0 && !col) throw InvalidArg{"col is null"};
        graph_export(ctx, g, row_ptr, col);
        CNN_HIP(hipStreamSynchronize(ctx->stream));
        return (int)CNN_OK;
    });
}

int cnn_graph_rethreshold(cnn_ctx *ctx, const cnn_graph *g, double r2, int32_t precision,
                          cnn_graph **out) {
    if (!ctx || !g || !out) return CNN_ERR_INVALID;
    *out = nullptr;
    return guarded(ctx, [&]() {
        if (g->bounds) throw InvalidArg{"a periodic-metric graph cannot be re-thresholded; build it again"};
        *out = graph_rethreshold(ctx, g, r2, precision);
        return (int)CNN_OK;
    });
}

}  // extern "C"
