// Phases S and U: common-neighbour similarity, union-find labelling, sort_by_size.
//
// Replaces FitterExtBFS._fit (_fit.pyx:408-491) with its connected-components
// equivalent: core(i) = consider(i) && |N(i)| > m_core (_types.pyx:1561-1564);
// edge(i, j) = core(i) && core(j) && j in N(i) && |N(i) & N(j)| >= c_sim
// (SimilarityCheckerExtScreensorted._check, _types.pyx:2654-2706, early exit at
// c_sim).  The BFS opens clusters in ascending init-point order and never
// reaches an earlier cluster through a valid edge, so its labels are the
// components of (core points, edges) numbered by smallest point index; a
// union-find whose links always point from the larger to the smaller root
// yields the same partition regardless of scheduling.
//
// Work plan on the GPU (all on the context stream):
//   k_init      core flags, parent[i] = i
//   k_hook      each core row tests its nearest smaller-index core neighbours
//               and hooks its parent to the first that passes (a forest, no
//               atomics); links most of every dense region cheaply
//   k_compress  every parent becomes its root
//   k_hook_roots roots re-point at a smaller neighbouring root on a pass
//   k_snapshot  rank[i] = (component size << 32 | root) of core points
//   k_round_b   every edge between different snapshot components is tested by
//               the endpoint of lower rank, unless the endpoints were joined
//               meanwhile; rows that own no edge are skipped by a per-tile bound
//   k_compress, component statistics, label assignment / sort_by_size
#include <algorithm>
#include <climits>
#include <memory>

#include "cnn_internal.hpp"

namespace cnn {

// ------------------------------------------------------------------ union-find

__device__ __forceinline__ int uf_find(int *parent, int x) {
    int p = atomic_load_relaxed(&parent[x]);
    while (p != x) {
        int gp = atomic_load_relaxed(&parent[p]);
        if (gp == p) return p;
        __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // path halving
        x = gp;
        p = atomic_load_relaxed(&parent[x]);
    }
    return x;
}

// Link the larger root below the smaller one (parent[x] <= x always holds).
__device__ __forceinline__ void uf_union(int *parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        int old = atomicCAS(&parent[a], a, b);
        if (old == a) return;
    }
}

// ------------------------------------------------------------------ similarity

constexpr int kStatSlots = 1024;  // striped counters: no same-address atomic hot spot

struct LabelArgs {
    const int64_t *row_ptr;
    const int32_t *col;
    int64_t n;
    int c_sim;
    int m_core;
    uint8_t *core;
    int *parent;
    unsigned long long *stats;  // [slot][0] tests, [slot][1] passed
    const uint64_t *rank;       // snapshot rank of core points: (component size << 32) | root; 0 otherwise
    const uint64_t *cand_max;   // per row tile: max rank over its candidate tiles (nullable)
    // per tile: the rank every point of the tile has (0: no core point; all
    // ones: ranks differ), so list entries in such tiles need no rank gather
    const uint64_t *tile_rank;
    const int32_t *selfpos;     // per row: position of its own index in its list (nullable)
    const int32_t *tile_csim;   // packed batches: c_sim / m_core per row tile (nullable)
    const int32_t *tile_mcore;
    // the lists as row words: per candidate tile pair of a spatial graph
    // (mask / pair_nz over tp_ptr / tp_col), or per row (wptr / wtile / wword)
    const uint64_t *mask;
    const uint8_t *pair_nz;
    const int64_t *tp_ptr;
    const int32_t *tp_col;
    const int64_t *wptr;
    const int32_t *wtile;
    const uint64_t *wword;
    const int32_t *diag;  // per row tile: its own pair's place in its list (nullable)
    // per-row degrees (nullable: row_ptr differences); word-only passes need no row_ptr
    const int32_t *deg;
    // several cutoff pairs in one launch, one per blockIdx.y (nullable: one
    // pair, c_sim / m_core): launch y takes csim_y[y] / mcore_y[y] and its own
    // slices of the per-point (stride n), per-tile (stride tiles) and
    // statistics arrays (see slice)
    const int32_t *csim_y;
    const int32_t *mcore_y;
    int64_t tiles;
    // rows k_round_tile hands to k_round_queue (nullable: it tests them itself):
    // {row, first pair, end pair (both relative to the row tile's first pair),
    // component the row has joined or -1}
    int4 *rq;
    unsigned long long *rq_len;
    int64_t rq_cap;
    // the row tiles in the order the per-tile kernels take them: heavy ones
    // first (nullable: ascending)
    const int32_t *tile_order;
};

// The cutoff pair of this launch's blockIdx.y and its slices of the
// per-pair arrays; unchanged for a single pair.
__device__ __forceinline__ LabelArgs slice(LabelArgs a) {
    if (!a.csim_y) return a;
    const int64_t y = blockIdx.y;
    a.core += y * a.n;
    a.parent += y * a.n;
    a.rank += y * a.n;
    a.stats += y * 4 * kStatSlots;
    if (a.cand_max) a.cand_max += y * a.tiles;
    if (a.tile_rank) a.tile_rank += y * a.tiles;
    a.c_sim = a.csim_y[y];
    a.m_core = a.mcore_y[y];
    return a;
}

__device__ __forceinline__ int64_t deg_of(const LabelArgs &a, int64_t w) {
    return a.deg ? (int64_t)a.deg[w] : a.row_ptr[w + 1] - a.row_ptr[w];
}

// Row w's words: its row of the candidate tile pairs of its row tile (spatial
// graph masks), or its own word list.
__device__ __forceinline__ WordRow word_row(const LabelArgs &a, int64_t w) {
    WordRow r;
    if (a.mask) {
        const int64_t R = w >> 6;
        const int64_t kb = a.tp_ptr[R];
        r.tile = a.tp_col + kb;
        r.word = a.mask + kb * 64 + (w & 63);
        r.nz = a.pair_nz + kb;
        r.len = a.tp_ptr[R + 1] - kb;
        r.stride = 64;
    } else {
        const int64_t b = a.wptr[w];
        r.tile = a.wtile + b;
        r.word = a.wword + b;
        r.nz = nullptr;
        r.len = a.wptr[w + 1] - b;
        r.stride = 1;
    }
    return r;
}

__device__ __forceinline__ int csim_of(const LabelArgs &a, int64_t w) {
    return a.tile_csim ? a.tile_csim[w >> 6] : a.c_sim;
}

__device__ __forceinline__ void add_stats(const LabelArgs &a, unsigned tests, unsigned passed, unsigned rows = 0) {
    if ((threadIdx.x & 63) == 0 && (tests || rows)) {
        const unsigned slot = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kStatSlots - 1);
        if (tests) {
            atomicAdd(&a.stats[4 * slot], (unsigned long long)tests);
            atomicAdd(&a.stats[4 * slot + 1], (unsigned long long)passed);
        }
        if (rows) atomicAdd(&a.stats[4 * slot + 2], (unsigned long long)rows);
    }
}

__global__ void k_init(LabelArgs a, const uint8_t *__restrict__ consider, const int32_t *__restrict__ perm) {
    a = slice(a);
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.n) return;
    const int64_t deg = deg_of(a, w);
    const int m_core = a.tile_mcore ? a.tile_mcore[w >> 6] : a.m_core;
    const int32_t o = perm ? perm[w] : (int32_t)w;  // -1: padding row of a packed batch
    bool ok = deg > (int64_t)m_core && o >= 0;
    if (consider && ok) ok = consider[o] != 0;
    a.core[w] = ok ? 1 : 0;
    a.parent[w] = (int)w;
}

// Test edge (w, j) unless the endpoints are already joined; link on a pass.
// Returns true when w and j are in one component afterwards.
__device__ __forceinline__ bool test_and_link(const LabelArgs &a, int w, int j, int64_t bw, int64_t ew, int c,
                                              unsigned &tests, unsigned &passed) {
    const int rw = uf_find(a.parent, w);
    const int rj = uf_find(a.parent, j);
    if (rw == rj) return true;
    const int64_t bj = a.row_ptr[j], ej = a.row_ptr[j + 1];
    const bool pass = wave_sim(a.col + bw, ew - bw, a.col + bj, ej - bj, c);
    ++tests;
    if (pass) {
        ++passed;
        if (__lane_id() == 0) uf_union(a.parent, w, j);
    }
    return pass;
}

// The same on row words (bitset similarity); A is w's word row, dw its degree.
__device__ __forceinline__ bool test_and_link_words(const LabelArgs &a, int w, int j, const WordRow &A, int64_t dw,
                                                    int c, unsigned &tests, unsigned &passed) {
    const int rw = uf_find(a.parent, w);
    const int rj = uf_find(a.parent, j);
    if (rw == rj) return true;
    const int64_t dj = deg_of(a, j);
    const bool pass = dw >= c && dj >= c && wave_sim_words(A, word_row(a, j), c);
    ++tests;
    if (pass) {
        ++passed;
        if (__lane_id() == 0) uf_union(a.parent, w, j);
    }
    return pass;
}

__device__ __forceinline__ uint64_t lane_bit_mask(int lane) { return 1ull << lane; }

// Sampling round (hooking): each core row tests its nearest smaller-index core
// neighbours in list order (the previous points of the working order, which
// are spatially closest) and points its parent at the first one that passes.
// Every row writes only its own parent entry and parents have smaller
// indices, so the hooks form a forest whose trees are the components of the
// passing edges tested here -- no atomics, no contention.
// Hooking of one core row on row words: its smaller-index neighbours in
// descending order, at most max_tests tests; skip_prev: the test against w - 1
// was already made (k_hook_tile).
__device__ __forceinline__ void hook_row_words(const LabelArgs &a, int64_t w, int max_tests, bool skip_prev) {
    const int lane = threadIdx.x & 63;
    const int64_t dw = deg_of(a, w);
    const bool tested_prev = skip_prev;
    // the row's words from its own tile down, 64 tiles a step (lane l holds
    // the l-th tile below the window's top), bits from the highest down:
    // smaller indices, nearest first
    const WordRow A = word_row(a, w);
    const int Rw = (int)(w >> 6);
    int64_t top = 0;  // tiles up to and including the row's own tile
    if (a.mask && a.diag && a.diag[Rw] >= 0) {
        top = a.diag[Rw] + 1;  // recorded when the lists were built
    } else {
        for (int64_t t0 = 0; t0 < A.len; t0 += 64) {
            const int64_t t = t0 + lane;
            const uint64_t le = __ballot(t < A.len && A.tile[t] <= Rw);
            top += __popcll(le);
            if (__popcll(le) < 64) break;
        }
    }
    const int c = csim_of(a, w);
    unsigned tests = 0, passed = 0;
    int attempts = 0;
    bool done = false;
    for (int64_t t1 = top; t1 > 0 && attempts < max_tests && !done; t1 -= 64) {
        const int64_t t = t1 - 1 - lane;
        uint64_t v = 0ull;
        int C = 0;
        if (t >= 0) {
            C = A.tile[t];
            v = word_at(A, t);
            if (C == Rw) v &= lane_bit_mask((int)(w & 63)) - 1ull;  // smaller indices only
        }
        uint64_t tiles = __ballot(v != 0ull);
        while (tiles && attempts < max_tests && !done) {
            const int src = __ffsll((unsigned long long)tiles) - 1;
            tiles &= tiles - 1;
            const uint32_t vlo = __builtin_amdgcn_readlane((uint32_t)v, src);
            const uint32_t vhi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
            const int Cs = __builtin_amdgcn_readlane(C, src);
            const bool bit = ((lane < 32 ? vlo >> lane : vhi >> (lane - 32)) & 1u) != 0u;
            uint64_t bits = __ballot(bit && a.core[Cs * 64 + lane]);
            while (bits && attempts < max_tests) {
                const int hb = 63 - __clzll((long long)bits);
                bits &= ~lane_bit_mask(hb);
                const int jj = Cs * 64 + hb;
                if (tested_prev && jj == (int)w - 1) continue;
                ++attempts;
                ++tests;
                const int64_t dj = deg_of(a, jj);
                if (dw >= c && dj >= c && wave_sim_words(A, word_row(a, jj), c)) {
                    ++passed;
                    if (lane == 0) a.parent[w] = jj;
                    done = true;
                    break;
                }
            }
        }
    }
    add_stats(a, tests, passed);
}

template <bool WORDS>
__global__ __launch_bounds__(256) void k_hook(LabelArgs a, int max_tests) {
    a = slice(a);
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= a.n || !a.core[w]) return;
    const int64_t b = a.row_ptr[w], e = a.row_ptr[w + 1];
    if constexpr (WORDS) {
        hook_row_words(a, w, max_tests, false);
        return;
    }
    // first entry >= w: the row's own entry, recorded when the lists were
    // emitted, or a 64-way search (each step probes 64 evenly spaced entries at
    // once and keeps the gap that holds the boundary)
    int64_t lo = a.selfpos ? min(b + (int64_t)a.selfpos[w], e) : b, hi = a.selfpos ? lo : e;
    while (lo < hi) {
        const int64_t step = (hi - lo + 63) >> 6;
        const int64_t pos = lo + lane * step;
        const bool below = pos < hi && a.col[pos] < (int32_t)w;
        const int cnt = __popcll(__ballot(below));  // probes 0 .. cnt-1 lie below w
        const int64_t nlo = cnt > 0 ? lo + (int64_t)(cnt - 1) * step + 1 : lo;
        const int64_t nhi = min(hi, lo + (int64_t)cnt * step);
        lo = nlo;
        hi = nhi;
    }
    unsigned tests = 0, passed = 0;
    int attempts = 0;
    // entries [b, lo) are the smaller indices, nearest first when walked backwards
    for (int64_t t1 = lo; t1 > b && attempts < max_tests; t1 -= 64) {
        const int64_t t = t1 - 1 - lane;
        const int j = t >= b ? a.col[t] : -1;
        uint64_t bits = __ballot(t >= b && a.core[j]);
        bool done = false;
        while (bits && attempts < max_tests) {
            const int src = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            const int jj = __shfl(j, src, 64);
            if (jj >= (int)w) continue;  // parents point to smaller indices only: the hooks stay a forest
            const int64_t bj = a.row_ptr[jj], ej = a.row_ptr[jj + 1];
            ++attempts;
            ++tests;
            if (wave_sim(a.col + b, e - b, a.col + bj, ej - bj, csim_of(a, w))) {
                ++passed;
                if (lane == 0) a.parent[w] = jj;
                done = true;
                break;
            }
        }
        if (done) break;
    }
    add_stats(a, tests, passed);
}

// Hooking of adjacent rows on a spatial graph's tile-pair words: one wave per
// row tile, lane r = row w = 64R + r.  Rows of one tile share the tile's
// candidate list, and pair k's 64 row words are contiguous (lane r reads row
// r's word), so one pass over the tile's non-empty pairs counts
// |N(w) & N(w - 1)| for all 63 adjacent row pairs at once (w - 1's word comes
// from the lane below).  A core row whose predecessor is a core neighbour with
// at least c_sim common neighbours hooks to it (to the first row of its run of
// hooked rows, which is connected to it through them).  The pass stops once
// every row still in question has reached c_sim.  Writes only parent[w] of
// hooked rows, to a smaller index: a forest, as the per-row hooking.
__global__ __launch_bounds__(256) void k_hook_tile(LabelArgs a, int64_t n_tiles, int32_t *__restrict__ queue,
                                                   unsigned int *__restrict__ queue_len) {
    a = slice(a);
    queue += (int64_t)blockIdx.y * a.n;
    queue_len += blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int64_t gR = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gR >= n_tiles) return;
    const int64_t R = a.tile_order ? (int64_t)a.tile_order[gR] : gR;
    const int64_t w = R * 64 + lane;
    const bool core_w = w < a.n && a.core[w] != 0;
    const bool core_p = __shfl_up((int)core_w, 1, 64) != 0;
    const int c = csim_of(a, R * 64);
    // rows whose test is still open: core, predecessor core within the tile
    const bool open = lane > 0 && core_w && core_p;
    bool hooked = false;
    unsigned tests = 0, passed = 0;
    if (__ballot(open) != 0ull) {
        const int64_t kb = a.tp_ptr[R], ke = a.tp_ptr[R + 1];
        // the row tile's own pair first (recorded when the lists were built,
        // or found by a ballot per 64 of its ascending column tiles): it tells
        // every row whether w - 1 is a neighbour, so the walk below can stop
        // as soon as the open rows with that neighbour have c common neighbours
        int64_t kd = -1;
        if (a.diag && a.diag[R] >= 0) kd = kb + a.diag[R];
        for (int64_t k0 = kb; k0 < ke && kd < 0; k0 += 64) {
            const int64_t kk = k0 + lane;
            const uint64_t hit = __ballot(kk < ke && a.tp_col[kk] == (int)R);
            if (hit) kd = k0 + __ffsll((unsigned long long)hit) - 1;
        }
        int cnt = 0;
        bool nbr = false;  // w - 1 in N(w): bit lane-1 of the row's word in its own tile
        if (kd >= 0 && a.pair_nz[kd] != 0) {
            const uint64_t v = a.mask[kd * 64 + lane];
            const uint64_t pv = __shfl_up(v, 1, 64);
            cnt = __popcll(v & pv);
            nbr = lane > 0 && ((v >> (lane - 1)) & 1ull) != 0ull;
        }
        bool done = __ballot(open && nbr && cnt < c) == 0ull;
        for (int64_t k0 = kb; k0 < ke && !done; k0 += 64) {
            // the window's other non-empty pairs, then their words 8 at a time (loads in flight together)
            const int64_t kk = k0 + lane;
            const bool nzl = kk < ke && kk != kd && a.pair_nz[kk] != 0;
            uint64_t nzm = __ballot(nzl);
            while (nzm && !done) {
                int idx[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    idx[u] = nzm ? __ffsll((unsigned long long)nzm) - 1 : -1;
                    nzm &= nzm - 1;
                }
                uint64_t v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = idx[u] >= 0 ? a.mask[(k0 + idx[u]) * 64 + lane] : 0ull;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const uint64_t pv = __shfl_up(v[u], 1, 64);
                    cnt += __popcll(v[u] & pv);
                }
                if (__ballot(open && nbr && cnt < c) == 0ull) done = true;
            }
        }
        hooked = open && nbr && cnt >= c;
        // a run of hooked rows hangs from the unhooked row before it (lane 0
        // is never hooked): depth one instead of a chain through w - 1
        const uint64_t below = ~__ballot(hooked) & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
        if (hooked) a.parent[w] = (int)(R * 64 + (63 - __clzll((long long)below)));
        tests = (unsigned)__popcll(__ballot(open && nbr));
        passed = (unsigned)__popcll(__ballot(hooked));
    }
    // core rows left to the per-row hooking
    const bool left = core_w && !hooked;
    const uint64_t lm = __ballot(left);
    if (lm) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(queue_len, (unsigned)__popcll(lm));
        base = __shfl(base, 0, 64);
        if (left) queue[base + __popcll(lm & ((1ull << lane) - 1ull))] = (int32_t)w;
    }
    add_stats(a, tests, passed);
}

// Per-row hooking of the rows k_hook_tile left (a queue; grid-stride, one wave
// per row).
__global__ __launch_bounds__(256) void k_hook_queue(LabelArgs a, const int32_t *__restrict__ queue,
                                                    const unsigned int *__restrict__ queue_len, int max_tests) {
    a = slice(a);
    queue += (int64_t)blockIdx.y * a.n;
    queue_len += blockIdx.y;
    const unsigned nq = *queue_len;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); q < (int64_t)nq; q += waves) {
        const int64_t w = queue[q];
        hook_row_words(a, w, max_tests, (w & 63) != 0);
    }
}

// Root hooking: a core point that is still a root tests core neighbours whose
// current root is smaller than itself (either side of it in the list) and
// re-points itself at that root on the first pass.  Only roots write their own
// entry and only to smaller values, so the parents stay a forest.
template <bool WORDS>
__device__ void hook_root(const LabelArgs &a, int64_t w, int max_tests) {
    const int lane = threadIdx.x & 63;
    unsigned tests = 0, passed = 0;
    int attempts = 0;
    if constexpr (WORDS) {
        const int64_t dw = deg_of(a, w);
        const WordRow A = word_row(a, w);
        const int c = csim_of(a, w);
        bool done = false;
        for (int64_t t0 = 0; t0 < A.len && attempts < max_tests && !done; t0 += 64) {
            const int64_t t = t0 + lane;
            uint64_t v = 0ull;
            int C = 0;
            if (t < A.len) {
                v = word_at(A, t);
                C = A.tile[t];
            }
            uint64_t tiles = __ballot(v != 0ull);
            while (tiles && attempts < max_tests && !done) {
                const int src = __ffsll((unsigned long long)tiles) - 1;
                tiles &= tiles - 1;
                const uint32_t vlo = __builtin_amdgcn_readlane((uint32_t)v, src);
                const uint32_t vhi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
                const int Cs = __builtin_amdgcn_readlane(C, src);
                const bool bit = ((lane < 32 ? vlo >> lane : vhi >> (lane - 32)) & 1u) != 0u;
                const int j = Cs * 64 + lane;
                // the parent loads with the core flag (parents are flattened)
                const int pj = bit ? a.parent[j] : INT_MAX;
                const int rj = (bit && a.core[j]) ? pj : INT_MAX;
                uint64_t bits = __ballot(rj < (int)w);
                while (bits && attempts < max_tests) {
                    const int src2 = __ffsll((unsigned long long)bits) - 1;
                    bits &= bits - 1;
                    const int jj = Cs * 64 + src2;
                    const int rr = __shfl(rj, src2, 64);
                    ++attempts;
                    ++tests;
                    const int64_t dj = deg_of(a, jj);
                    if (dw >= c && dj >= c && wave_sim_words(A, word_row(a, jj), c)) {
                        ++passed;
                        if (lane == 0) a.parent[w] = rr;
                        done = true;
                        break;
                    }
                    bits &= ~__ballot(rj == rr);  // one test per neighbouring component
                }
            }
        }
        add_stats(a, tests, passed);
        return;
    }
    const int64_t b = a.row_ptr[w], e = a.row_ptr[w + 1];
    for (int64_t t0 = b; t0 < e && attempts < max_tests; t0 += 64) {
        const int64_t t = t0 + lane;
        const int j = t < e ? a.col[t] : -1;
        const int pj = t < e ? a.parent[j] : INT_MAX;  // with the core flag (parents are flattened)
        const int rj = (t < e && a.core[j]) ? pj : INT_MAX;
        uint64_t bits = __ballot(rj < (int)w);
        bool done = false;
        while (bits && attempts < max_tests) {
            const int src = __ffsll((unsigned long long)bits) - 1;
            bits &= bits - 1;
            const int jj = __shfl(j, src, 64);
            const int rr = __shfl(rj, src, 64);
            const int64_t bj = a.row_ptr[jj], ej = a.row_ptr[jj + 1];
            ++attempts;
            ++tests;
            if (wave_sim(a.col + b, e - b, a.col + bj, ej - bj, csim_of(a, w))) {
                ++passed;
                if (lane == 0) a.parent[w] = rr;
                done = true;
                break;
            }
            bits &= ~__ballot(rj == rr);  // one test per neighbouring component
        }
        if (done) break;
    }
    add_stats(a, tests, passed);
}

template <bool WORDS>
__global__ __launch_bounds__(256) void k_hook_roots(LabelArgs a, int max_tests) {
    a = slice(a);
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= a.n || !a.core[w] || a.parent[w] != (int)w) return;
    hook_root<WORDS>(a, w, max_tests);
}

// The same for the roots k_compress_roots listed, one wave each (grid-stride):
// no wave is started for the points that are not roots.
template <bool WORDS>
__global__ __launch_bounds__(256) void k_hook_roots_list(LabelArgs a, const int32_t *__restrict__ roots,
                                                         const unsigned int *__restrict__ n_roots, int max_tests) {
    const int64_t nr = *n_roots;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); q < nr; q += waves) {
        const int64_t w = roots[q];
        if (a.parent[w] == (int)w) hook_root<WORDS>(a, w, max_tests);
    }
}

// Row w's part of a full round on row words, over its candidate tiles
// [t_from, t_to) (clamped to the row's tiles): 64 tiles a step (ascending
// column tile, so entries come in list order), tiles whose points share a
// rank that is not above w's, or that w has joined, are skipped without
// reading their bits; tests are bitset intersections of the two rows' words.
// resolved: the snapshot component w is known to have joined (-1: none).
__device__ void round_row_words(const LabelArgs &a, int w, uint64_t rw, int c, int64_t t_from, int64_t t_to,
                                int resolved, unsigned &tests, unsigned &passed) {
    const int lane = threadIdx.x & 63;
    const WordRow A = word_row(a, w);
    const int64_t dw = deg_of(a, w);
    t_to = min(t_to, A.len);
    for (int64_t t0 = t_from; t0 < t_to; t0 += 64) {
        const int64_t t = t0 + lane;
        uint64_t v = 0ull, tr = 0ull;
        int C = 0;
        if (t < t_to) {
            // the tile id with the pair's flag, the tile's rank with the word
            C = A.tile[t];
            v = word_at(A, t);
            tr = a.tile_rank ? a.tile_rank[C] : ~0ull;
        }
        uint64_t tiles = __ballot(v != 0ull && (tr == ~0ull || tr > rw));
        while (tiles) {
            const int src = __ffsll((unsigned long long)tiles) - 1;
            tiles &= tiles - 1;
            const uint32_t tlo = __builtin_amdgcn_readlane((uint32_t)tr, src);
            const uint32_t thi = __builtin_amdgcn_readlane((uint32_t)(tr >> 32), src);
            const uint64_t trs = ((uint64_t)thi << 32) | tlo;
            if (trs != ~0ull && (int)tlo == resolved) continue;
            const uint32_t vlo = __builtin_amdgcn_readlane((uint32_t)v, src);
            const uint32_t vhi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
            const bool bit = ((lane < 32 ? vlo >> lane : vhi >> (lane - 32)) & 1u) != 0u;
            const int j = __builtin_amdgcn_readlane(C, src) * 64 + lane;
            uint64_t rj = 0ull;
            if (bit) rj = trs == ~0ull ? a.rank[j] : trs;
            const int kj = (int)(uint32_t)rj;
            uint64_t bits = __ballot(bit && rj > rw && kj != resolved);
            while (bits) {
                const int jsrc = __ffsll((unsigned long long)bits) - 1;
                bits &= bits - 1;
                const int jj = __builtin_amdgcn_readlane(j, jsrc);
                const int kjj = __builtin_amdgcn_readlane(kj, jsrc);
                if (test_and_link_words(a, w, jj, A, dw, c, tests, passed)) {
                    resolved = kjj;
                    bits &= ~__ballot(kj == kjj);
                }
            }
        }
    }
}

// Full round.  Every edge between two different snapshot components is owned
// by the endpoint whose component ranks lower (rank = size << 32 | root, all
// distinct), and only the owner examines it.  A row whose rank is at least the
// largest rank among its candidate tiles owns no edge and is skipped without
// reading its list.  Candidates are filtered with the read-only snapshot (one
// 8-byte gather per entry); a stale snapshot only yields redundant candidates,
// which test_and_link's find() check discards.
template <bool WORDS>
__global__ __launch_bounds__(256) void k_round_b(LabelArgs a) {
    a = slice(a);
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= a.n) return;
    const uint64_t rw = a.rank[w];
    if (rw == 0) return;
    if (a.cand_max && rw >= a.cand_max[w >> 6]) return;
    const int c = csim_of(a, w);
    unsigned tests = 0, passed = 0;
    int resolved = -1;  // last snapshot component found joined to w
    if constexpr (WORDS) {
        round_row_words(a, (int)w, rw, c, 0, INT64_MAX, resolved, tests, passed);
        add_stats(a, tests, passed, 1);
        return;
    }
    // four 64-entry windows per step: their loads and rank lookups are issued
    // together, so a row's scan waits on one memory round trip per 256 entries
    const int64_t b = a.row_ptr[w], e = a.row_ptr[w + 1];
    constexpr int kWin = 4;
    for (int64_t t0 = b; t0 < e; t0 += 64 * kWin) {
        int j[kWin];
        uint64_t rj[kWin];
#pragma unroll
        for (int u = 0; u < kWin; ++u) {
            const int64_t t = t0 + 64 * u + lane;
            j[u] = t < e ? a.col[t] : -1;
        }
#pragma unroll
        for (int u = 0; u < kWin; ++u) {
            uint64_t tr = 0ull;
            if (j[u] >= 0) tr = a.tile_rank ? a.tile_rank[j[u] >> 6] : ~0ull;
            rj[u] = tr;
        }
#pragma unroll
        for (int u = 0; u < kWin; ++u)
            if (rj[u] == ~0ull) rj[u] = a.rank[j[u]];
#pragma unroll
        for (int u = 0; u < kWin; ++u) {
            const int kj = (int)(uint32_t)rj[u];
            const bool cand = rj[u] > rw && kj != resolved;
            uint64_t bits = __ballot(cand);
            while (bits) {
                const int src = __ffsll((unsigned long long)bits) - 1;
                bits &= bits - 1;
                const int jj = __shfl(j[u], src, 64);
                const int kjj = __shfl(kj, src, 64);
                if (test_and_link(a, (int)w, jj, b, e, c, tests, passed)) {
                    // w has joined snapshot component kjj: its other entries are joined too
                    resolved = kjj;
                    bits &= ~__ballot(kj == kjj);
                }
            }
        }
    }
    add_stats(a, tests, passed, 1);
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int src) {
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, src);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// Columns (lanes) of a column tile whose rank, held by lane b for column b,
// lies above rw and belongs to no component numbered `resolved`: one pass per
// distinct rank of the tile (a ballot each), the per-row test folded in.
__device__ __forceinline__ uint64_t columns_above(uint64_t rk, uint64_t rw, int resolved) {
    uint64_t rem = ~0ull, bits = 0ull;
    while (rem) {
        const uint64_t val = readlane_u64(rk, __ffsll((unsigned long long)rem) - 1);
        const uint64_t m = __ballot(rk == val);
        rem &= ~m;
        if (val > rw && (int)(uint32_t)val != resolved) bits |= m;
    }
    return bits;
}

// Full round on a spatial graph's tile-pair words, `split` waves per row tile
// (lane r owns row w = 64R + r; wave q of the tile takes the q-th of `split`
// equal parts of its candidate pairs), with k_round_b's ownership and skip
// rules: an edge is examined by the endpoint of lower snapshot rank, rows at
// or above the tile's candidate maximum own none, and a row drops every column
// of a component it has joined (per wave: the skip is only a saving, every
// test re-checks with two finds).  The pairs are walked once for all 64 rows
// -- pair k's 64 row words are one contiguous 512-byte load, 8 pairs in
// flight -- and the rank filter and the already-joined check run for all rows
// in parallel.  An edge whose endpoints are still apart needs a wave-wide bitset
// test (test_and_link_words): the wave runs kInlineTests of them itself, then
// hands each row that still has one, with the rest of the wave's pairs, to
// k_round_queue (one wave per row there, as k_round_b), so no wave runs the
// tests of many rows one after another; when the queue is full (or absent)
// the wave tests every edge itself.
__global__ __launch_bounds__(256) void k_round_tile(LabelArgs a, int64_t n_tiles, int split) {
    a = slice(a);
    const int lane = threadIdx.x & 63;
    const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= n_tiles * split) return;
    const int64_t gR = g / split;
    const int q = (int)(g - gR * split);
    const int64_t R = a.tile_order ? (int64_t)a.tile_order[gR] : gR;
    const int64_t w = R * 64 + lane;
    const uint64_t rw = w < a.n ? a.rank[w] : 0ull;
    const uint64_t cmax = a.cand_max ? a.cand_max[R] : ~0ull;
    const bool active = rw != 0ull && rw < cmax;
    const uint64_t act = __ballot(active);
    if (act == 0ull) return;
    // a column tile whose points share a rank not above every active row's has no candidate
    uint64_t rmin = active ? rw : ~0ull;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = __shfl_xor(rmin, off, 64);
        rmin = o < rmin ? o : rmin;
    }
    const int c = csim_of(a, R * 64);
    unsigned tests = 0, passed = 0;
    int resolved = -1;  // last snapshot component this row was found joined to
    bool live = active;  // false once the row is queued
    bool rq_full = false;
    // wave-wide tests a wave runs itself before it queues open rows: a passing
    // test usually joins the components the other rows' open edges connect
    constexpr unsigned kInlineTests = 2;
    const int64_t kb0 = a.tp_ptr[R], ke0 = a.tp_ptr[R + 1];
    const int64_t per = (ke0 - kb0 + split - 1) / split;
    const int64_t kb = kb0 + q * per, ke = min(ke0, kb + per);
    constexpr int kInFlight = 8;
    for (int64_t k0 = kb; k0 < ke; k0 += 64) {
        const int64_t kk = k0 + lane;
        int C = 0;
        uint64_t tr = 0ull;
        bool want = false;
        if (kk < ke && a.pair_nz[kk] != 0) {
            C = a.tp_col[kk];
            tr = a.tile_rank ? a.tile_rank[C] : ~0ull;
            want = tr == ~0ull || tr > rmin;
        }
        uint64_t wm = __ballot(want);
        while (wm) {
            int idx[kInFlight];
#pragma unroll
            for (int u = 0; u < kInFlight; ++u) {
                idx[u] = wm ? __ffsll((unsigned long long)wm) - 1 : -1;
                wm &= wm - 1;
            }
            uint64_t v[kInFlight], rk[kInFlight];
#pragma unroll
            for (int u = 0; u < kInFlight; ++u) {
                v[u] = 0ull;
                rk[u] = 0ull;
                if (idx[u] >= 0) {
                    const int64_t Cu = __builtin_amdgcn_readlane(C, idx[u]);
                    if (live) v[u] = a.mask[(k0 + idx[u]) * 64 + lane];
                    if (readlane_u64(tr, idx[u]) == ~0ull && Cu * 64 + lane < a.n) rk[u] = a.rank[Cu * 64 + lane];
                }
            }
#pragma unroll
            for (int u = 0; u < kInFlight; ++u) {
                if (!live) v[u] = 0ull;  // queued by an earlier pair of this batch
                if (idx[u] < 0 || __ballot(v[u] != 0ull) == 0ull) continue;
                const int Cu = __builtin_amdgcn_readlane(C, idx[u]);
                const uint64_t tru = readlane_u64(tr, idx[u]);
                const bool common = tru != ~0ull;
                uint64_t bits;
                if (common) bits = (tru > rw && (int)(uint32_t)tru != resolved) ? v[u] : 0ull;
                else bits = columns_above(rk[u], rw, resolved) & v[u];
                while (__ballot(bits != 0ull)) {
                    const int col = bits ? __ffsll((unsigned long long)bits) - 1 : 0;
                    const uint64_t rb = common ? tru : __shfl(rk[u], col, 64);
                    int j = -1;
                    bool joined = false;
                    if (bits) {
                        bits &= bits - 1;
                        j = Cu * 64 + col;
                        joined = uf_find(a.parent, (int)w) == uf_find(a.parent, j);
                    }
                    bool changed = joined;
                    uint64_t need = __ballot(j >= 0 && !joined);
                    while (need) {
                        if (tests >= kInlineTests && a.rq && !rq_full) {
                            // queue the rows still open, from this pair on
                            unsigned long long base = 0;
                            if (lane == 0) base = atomicAdd(a.rq_len, (unsigned long long)__popcll(need));
                            base = __shfl(base, 0, 64);
                            const bool mine = ((need >> lane) & 1ull) != 0ull;
                            const unsigned long long slot = base + __popcll(need & ((1ull << lane) - 1ull));
                            const bool queued = mine && slot < (unsigned long long)a.rq_cap;
                            if (queued) {
                                a.rq[slot] = make_int4((int)w, (int)(k0 + idx[u] - kb0), (int)(ke - kb0), resolved);
                                live = false;
                                bits = 0ull;
                            }
                            rq_full = __ballot(mine && !queued) != 0ull;
                            need &= ~__ballot(queued);
                            if (!need) break;
                        }
                        const int src = __ffsll((unsigned long long)need) - 1;
                        need &= need - 1;
                        const int ws = (int)(R * 64 + src);
                        const int js = __builtin_amdgcn_readlane(j, src);
                        if (test_and_link_words(a, ws, js, word_row(a, ws), deg_of(a, ws), c, tests, passed) &&
                            lane == src)
                            changed = true;
                    }
                    // a joined component: its remaining columns are joined as well
                    if (changed) resolved = (int)(uint32_t)rb;
                    if (__ballot(changed)) {
                        if (common) {
                            if (changed) bits = 0ull;
                        } else {
                            bits &= columns_above(rk[u], rw, resolved);
                        }
                    }
                }
            }
        }
    }
    add_stats(a, tests, passed, q == 0 ? (unsigned)__popcll(act) : 0u);
}

// The rows k_round_tile queued, one wave each (grid-stride): the row's
// candidate pairs from the queued one to the end of its wave's part, with the
// per-row walk of k_round_b.
__global__ __launch_bounds__(256) void k_round_queue(LabelArgs a) {
    a = slice(a);
    const int64_t nq = (int64_t)min<unsigned long long>(*a.rq_len, (unsigned long long)a.rq_cap);
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); q < nq; q += waves) {
        const int4 e = a.rq[q];
        unsigned tests = 0, passed = 0;
        round_row_words(a, e.x, a.rank[e.x], csim_of(a, e.x), e.y, e.z, e.w, tests, passed);
        add_stats(a, tests, passed);
    }
}

// Waves per row tile for k_round_tile: about 64 candidate pairs each.
static int round_split(int64_t n_tile_pairs, int64_t n_tiles) {
    if (n_tiles <= 0) return 1;
    const int64_t per_tile = (n_tile_pairs + n_tiles - 1) / n_tiles;
    return (int)std::min<int64_t>(32, std::max<int64_t>(1, (per_tile + 63) / 64));
}

// Largest snapshot rank among the core points of each tile (0: none), and the
// tile's common rank: the rank of all its points when they share one (a
// non-core point has rank 0), all ones when they differ.
__global__ void k_tile_rank_max(const uint64_t *__restrict__ rank, int64_t n, int64_t n_tiles,
                                uint64_t *__restrict__ tmax, uint64_t *__restrict__ tcommon) {
    rank += (int64_t)blockIdx.y * n;  // cutoff pair blockIdx.y (one pair: 0)
    tmax += (int64_t)blockIdx.y * n_tiles;
    tcommon += (int64_t)blockIdx.y * n_tiles;
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= n_tiles) return;
    const int64_t w = t * 64 + lane;
    const bool real = w < n;
    const uint64_t mine = real ? rank[w] : 0ull;
    uint64_t v = mine;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    const bool same = __ballot(real && mine != v) == 0ull;  // every real point has the largest rank
    if (lane == 0) {
        tmax[t] = v;
        tcommon[t] = same ? v : ~0ull;
    }
}

// Largest rank over the candidate column tiles of each row tile, one
// workgroup per row tile: a tile with thousands of candidates takes a few
// steps of 256 instead of a long serial walk by one wave.
__global__ __launch_bounds__(256) void k_cand_max(const uint64_t *__restrict__ tmax, const int64_t *__restrict__ tp_ptr,
                                                  const int32_t *__restrict__ tp_col, int64_t n_tiles,
                                                  uint64_t *__restrict__ cmax) {
    tmax += (int64_t)blockIdx.y * n_tiles;
    cmax += (int64_t)blockIdx.y * n_tiles;
    __shared__ uint64_t sv[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t R = blockIdx.x;
    if (R >= n_tiles) return;  // the whole workgroup
    const int64_t kb = tp_ptr ? tp_ptr[R] : 0, ke = tp_ptr ? tp_ptr[R + 1] : n_tiles;
    uint64_t v = 0;
    for (int64_t k = kb + threadIdx.x; k < ke; k += 256) {
        const int64_t C = tp_ptr ? (int64_t)tp_col[k] : k;
        const uint64_t x = tmax[C];
        v = x > v ? x : v;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if (lane == 0) sv[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t m = sv[0];
        for (int q = 1; q < 4; ++q) m = sv[q] > m ? sv[q] : m;
        cmax[R] = m;
    }
}

__global__ void k_snapshot(const uint8_t *core, const int *parent, const int32_t *cnt, int64_t n, uint64_t *rank) {
    const int64_t off = (int64_t)blockIdx.y * n;
    core += off;
    parent += off;
    cnt += off;
    rank += off;
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int r = parent[w];
    rank[w] = core[w] ? (((uint64_t)(uint32_t)cnt[r] << 32) | (uint32_t)r) : 0ull;
}

// Flatten: every entry points at its root afterwards.  The walk only reads
// (entries of other threads may change meanwhile, but only towards their
// roots) and each thread writes only its own entry, so no write can leave an
// entry below its root.
__global__ void k_compress(int *parent, int64_t n) {
    parent += (int64_t)blockIdx.y * n;
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    int x = (int)w;
    int p = atomic_load_relaxed(&parent[x]);
    while (p != x) {
        x = p;
        p = atomic_load_relaxed(&parent[x]);
    }
    __hip_atomic_store(&parent[w], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// k_compress that also lists the core points that are roots afterwards
// (one atomic per wave), for k_hook_roots_list.
__global__ void k_compress_roots(int *parent, const uint8_t *__restrict__ core, int64_t n, int32_t *__restrict__ roots,
                                 unsigned int *__restrict__ n_roots) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool root = false;
    if (w < n) {
        int x = (int)w;
        int p = atomic_load_relaxed(&parent[x]);
        while (p != x) {
            x = p;
            p = atomic_load_relaxed(&parent[x]);
        }
        __hip_atomic_store(&parent[w], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        root = x == (int)w && core[w] != 0;
    }
    const uint64_t m = __ballot(root);
    if (m) {
        const int lane = threadIdx.x & 63;
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(n_roots, (unsigned)__popcll(m));
        base = __shfl(base, 0, 64);
        if (root) roots[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)w;
    }
}

// Lanes sharing a key are combined before one atomic per distinct key (points
// of a component are mostly adjacent in the working order).
__device__ __forceinline__ void wave_add_by_key(int32_t *cnt, int32_t *mins, bool active, int key, int val) {
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int k = __shfl(key, leader, 64);
        const bool mine = active && key == k;
        const uint64_t match = __ballot(mine);
        if (mins) {
            int v = mine ? val : INT_MAX;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
            if (__lane_id() == leader) atomicMin(&mins[k], v);
        }
        if (__lane_id() == leader) atomicAdd(&cnt[k], __popcll(match));
        pending &= ~match;
    }
}

// Per-key sums over consecutive points: a wave walks kRunSteps steps of 64
// points and keeps one run (key, count, minimum) while every active lane of a
// step has the run's key, so a large component costs one atomic per run, not
// one per wave (same-address atomics serialize).
constexpr int kRunSteps = 16;

struct KeyRun {
    int key = -1;
    int cnt = 0;
    int mn = INT_MAX;
};

__device__ __forceinline__ void run_flush(KeyRun &run, int32_t *cnt, int32_t *mins) {
    if (run.key >= 0 && __lane_id() == 0) {
        atomicAdd(&cnt[run.key], run.cnt);
        if (mins) atomicMin(&mins[run.key], run.mn);
    }
    run.key = -1;
    run.cnt = 0;
    run.mn = INT_MAX;
}

__device__ __forceinline__ void run_add(KeyRun &run, int32_t *cnt, int32_t *mins, bool active, int key, int val) {
    const uint64_t act = __ballot(active);
    if (!act) return;
    const int k = __shfl(key, __ffsll((unsigned long long)act) - 1, 64);
    if (__ballot(active && key != k) != 0ull) {  // mixed keys: one atomic per key
        wave_add_by_key(cnt, mins, active, key, val);
        return;
    }
    if (k != run.key) {
        run_flush(run, cnt, mins);
        run.key = k;
    }
    run.cnt += __popcll(act);
    if (mins) {
        int v = active ? val : INT_MAX;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
        run.mn = min(run.mn, v);
    }
}

__device__ __forceinline__ int64_t run_base() {
    return ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 * kRunSteps);
}

__global__ __launch_bounds__(256) void k_root_count(const uint8_t *core, const int *parent, int64_t n,
                                                    int32_t *cnt) {
    const int64_t off = (int64_t)blockIdx.y * n;
    core += off;
    parent += off;
    cnt += off;
    const int64_t base = run_base();
    KeyRun run;
    for (int it = 0; it < kRunSteps; ++it) {
        const int64_t w = base + it * 64 + __lane_id();
        const bool active = w < n && core[w];
        run_add(run, cnt, nullptr, active, active ? parent[w] : 0, 0);
    }
    run_flush(run, cnt, nullptr);
}

__global__ void k_count_roots(const int32_t *cnt, int64_t n, unsigned long long *roots) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t b = __ballot(w < n && cnt[w] > 0);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(roots, (unsigned long long)__popcll(b));
}

// ------------------------------------------------------------------ labels

__global__ __launch_bounds__(256) void k_comp_stats(const uint8_t *core, const int *parent, const int32_t *perm,
                                                    int64_t n, int32_t *cmin, int32_t *csize,
                                                    unsigned long long *counters) {
    const int64_t off = (int64_t)blockIdx.y * n;
    core += off;
    parent += off;
    cmin += off;
    csize += off;
    counters += 8 * blockIdx.y;
    const int64_t base = run_base();
    KeyRun run;
    unsigned long long nc = 0, nr = 0;
    for (int it = 0; it < kRunSteps; ++it) {
        const int64_t w = base + it * 64 + __lane_id();
        bool is_core = false, is_root = false;
        int r = 0, orig = 0;
        if (w < n && core[w]) {
            is_core = true;
            r = parent[w];
            is_root = r == (int)w;
            orig = perm ? perm[w] : (int32_t)w;
        }
        run_add(run, csize, cmin, is_core, r, orig);
        nc += __popcll(__ballot(is_core));
        nr += __popcll(__ballot(is_root));
    }
    run_flush(run, csize, cmin);
    if (__lane_id() == 0) {
        if (nc) atomicAdd(&counters[0], nc);
        if (nr) atomicAdd(&counters[1], nr);
    }
}

__global__ void k_collect_clusters(const uint8_t *core, const int *parent, const int32_t *cmin, const int32_t *csize,
                                   int64_t n, int64_t member_cutoff, unsigned long long *K, uint64_t *keys,
                                   int32_t *vals) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n || !core[w] || parent[w] != (int)w) return;
    const int32_t sz = csize[w];
    if ((int64_t)sz < member_cutoff) return;
    const unsigned long long idx = atomicAdd(K, 1ull);
    keys[idx] = ((uint64_t)(uint32_t)(0x7fffffff - sz) << 32) | (uint32_t)cmin[w];
    vals[idx] = (int32_t)w;
}

__global__ void k_rank_clusters(const int32_t *vals, int64_t K, int64_t max_clusters, int32_t *newlab) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K) return;
    newlab[vals[t]] = (max_clusters < 0 || t < max_clusters) ? (int32_t)(t + 1) : 0;
}

// sort_by_size without the host: the K cluster keys (K read on the device;
// they are distinct) are sorted by one workgroup in LDS (bitonic, over the
// next power of two at or above K) and ranked as k_rank_clusters does.  More
// than kSmallClusters clusters: *big = 1 and nothing is written (the host
// sorts them after its next read of the counters).
constexpr int kSmallClusters = 1024;
__global__ __launch_bounds__(kSmallClusters) void k_rank_clusters_small(const uint64_t *__restrict__ keys,
                                                                        const int32_t *__restrict__ vals,
                                                                        const unsigned long long *__restrict__ Kp,
                                                                        int64_t max_clusters,
                                                                        int32_t *__restrict__ newlab,
                                                                        unsigned long long *__restrict__ big) {
    __shared__ uint64_t sk[kSmallClusters];
    __shared__ int32_t sv[kSmallClusters];
    const int64_t K = (int64_t)*Kp;
    const int t = threadIdx.x;
    if (K > kSmallClusters) {
        if (t == 0) *big = 1ull;
        return;
    }
    int n2 = 1;
    while (n2 < K) n2 <<= 1;
    sk[t] = t < K ? keys[t] : ~0ull;
    sv[t] = t < K ? vals[t] : -1;
    __syncthreads();
    for (int size = 2; size <= n2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int q = t ^ stride;
            if (q > t && q < n2) {
                const bool up = (t & size) == 0;
                const uint64_t ka = sk[t], kb = sk[q];
                if ((ka > kb) == up) {
                    sk[t] = kb;
                    sk[q] = ka;
                    const int32_t v = sv[t];
                    sv[t] = sv[q];
                    sv[q] = v;
                }
            }
            __syncthreads();
        }
    }
    if (t < K) newlab[sv[t]] = (max_clusters < 0 || t < max_clusters) ? (int32_t)(t + 1) : 0;
}

// Packed batches: clusters of all datasets sorted together by
// (dataset, size descending, smallest original index); ranks restart at 1 in
// every dataset.
struct PackedKey {
    int ds_shift, size_shift;
    uint32_t size_max;
    const int32_t *tile_ds;
    const int64_t *ds_offset;  // cmin is stored relative to the dataset's first row when set
};

__global__ void k_collect_clusters_packed(const uint8_t *core, const int *parent, const int32_t *cmin,
                                          const int32_t *csize, int64_t n, int64_t member_cutoff, PackedKey pk,
                                          unsigned long long *K, uint64_t *keys, int32_t *vals) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n || !core[w] || parent[w] != (int)w) return;
    const int32_t sz = csize[w];
    if ((int64_t)sz < member_cutoff) return;
    const int ds = pk.tile_ds[w >> 6];
    const uint64_t cm = (uint64_t)(cmin[w] - (pk.ds_offset ? pk.ds_offset[ds] : 0));
    const unsigned long long idx = atomicAdd(K, 1ull);
    keys[idx] = ((uint64_t)ds << pk.ds_shift) | ((uint64_t)(pk.size_max - (uint32_t)sz) << pk.size_shift) | cm;
    vals[idx] = (int32_t)w;
}

__global__ void k_first_of_set(const uint64_t *keys, int64_t K, int ds_shift, int32_t *first) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K) return;
    const uint64_t ds = keys[t] >> ds_shift;
    if (t == 0 || (keys[t - 1] >> ds_shift) != ds) first[ds] = (int32_t)t;
}

__global__ void k_rank_clusters_packed(const uint64_t *keys, const int32_t *vals, int64_t K, int ds_shift,
                                       const int32_t *first, int64_t max_clusters, int32_t *newlab,
                                       unsigned long long *kept) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    if (t < K) {
        const int64_t r = t - first[keys[t] >> ds_shift] + 1;
        keep = max_clusters < 0 || r <= max_clusters;
        newlab[vals[t]] = keep ? (int32_t)r : 0;
    }
    const uint64_t b = __ballot(keep);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(kept, (unsigned long long)__popcll(b));
}

__global__ void k_write_sorted(const uint8_t *core, const int *parent, const int32_t *newlab, const int32_t *perm,
                               int64_t n, int64_t *labels) {
    const int64_t off = (int64_t)blockIdx.y * n;
    core += off;
    parent += off;
    newlab += off;
    labels += off;
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int64_t o = perm ? perm[w] : w;
    if (o >= 0) labels[o] = core[w] ? (int64_t)newlab[parent[w]] : 0;
}

__global__ void k_mark_first(const uint8_t *core, const int *parent, const int32_t *cmin, int64_t n,
                             uint8_t *isfirst) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n || !core[w] || parent[w] != (int)w) return;
    isfirst[cmin[w]] = 1;
}

__global__ void k_write_unsorted(const uint8_t *core, const int *parent, const int32_t *cmin, const int32_t *rank_incl,
                                 const int32_t *perm, int64_t n, int64_t current_start, bool write_all,
                                 int64_t *labels) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int64_t o = perm ? perm[w] : w;
    if (o < 0) return;
    if (core[w]) labels[o] = current_start + (int64_t)rank_incl[cmin[parent[w]]] - 1;
    else if (write_all) labels[o] = 0;
}

// ------------------------------------------------------------------ per-row word lists from lists

// Rows of a CSR (strictly ascending) as word lists: one word per distinct
// column tile (index >> 6) of a row.  Count per row, one wave per row.
__global__ __launch_bounds__(256) void k_row_word_count(const int64_t *__restrict__ row_ptr,
                                                        const int32_t *__restrict__ col, int64_t n,
                                                        int32_t *__restrict__ cnt) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    int total = 0;
    // four windows of 64 entries per step, their loads issued together
    for (int64_t t0 = b; t0 < e; t0 += 256) {
        int32_t cur[4], prev[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = t0 + 64 * u + lane;
            cur[u] = t < e ? col[t] : 0;
            prev[u] = (t < e && t > b) ? col[t - 1] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = t0 + 64 * u + lane;
            const bool head = t < e && (t == b || (prev[u] >> 6) != (cur[u] >> 6));
            total += __popcll(__ballot(head));
        }
    }
    if (lane == 0) cnt[w] = total;
}

// Entries and words of every stride-th row, summed into sums[0] / sums[1]:
// the automatic similarity choice estimates the lists' fill from them.
__global__ __launch_bounds__(256) void k_row_word_sample(const int64_t *__restrict__ row_ptr,
                                                         const int32_t *__restrict__ col, int64_t n, int64_t stride,
                                                         unsigned long long *__restrict__ sums) {
    const int lane = threadIdx.x & 63;
    const int64_t w = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * stride;
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    int total = 0;
    for (int64_t t0 = b; t0 < e; t0 += 64) {
        const int64_t t = t0 + lane;
        bool head = false;
        if (t < e) head = t == b || (col[t - 1] >> 6) != (col[t] >> 6);
        total += __popcll(__ballot(head));
    }
    if (lane == 0 && total) {
        atomicAdd(&sums[0], (unsigned long long)(e - b));
        atomicAdd(&sums[1], (unsigned long long)total);
    }
}

// Fill: the entries of one word are consecutive in the row, so each window of
// 64 entries ors its bits per segment (a segmented scan over the lanes, the
// open segment of the previous window carried in) and the segment's last
// entry writes the word and its tile.
__global__ __launch_bounds__(256) void k_row_word_fill(const int64_t *__restrict__ row_ptr,
                                                       const int32_t *__restrict__ col, int64_t n,
                                                       const int64_t *__restrict__ wptr, int32_t *__restrict__ wtile,
                                                       uint64_t *__restrict__ wword) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    const int64_t base = wptr[w];
    int64_t run = 0;          // words started in earlier windows
    uint64_t carry = 0ull;    // bits of the word still open at the end of the previous window
    const uint64_t upto_mask = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    for (int64_t t0 = b; t0 < e; t0 += 64) {
        const int64_t t = t0 + lane;
        const bool valid = t < e;
        int tile = 0;
        uint64_t x = 0ull;
        bool head = false, tail = false;
        if (valid) {
            const int v = col[t];
            tile = v >> 6;
            x = 1ull << (v & 63);
            head = t == b || (col[t - 1] >> 6) != tile;
            tail = t + 1 == e || (col[t + 1] >> 6) != tile;
        }
        const uint64_t hm = __ballot(head);
        const uint64_t upto = hm & upto_mask;
        const int start = upto ? 63 - __clzll((long long)upto) : -1;  // this lane's word starts here
        if (start < 0) x |= carry;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t y = __shfl_up(x, off, 64);
            if (lane >= off && lane - off >= start) x |= y;
        }
        const int64_t idx = base + run + __popcll(upto) - 1;
        if (valid && tail) {
            wtile[idx] = tile;
            wword[idx] = x;
        }
        const bool open = __shfl((int)(valid && !tail), 63, 64) != 0;
        const uint64_t last = __shfl(x, 63, 64);
        carry = open ? last : 0ull;
        run += __popcll(hm);
    }
}

static inline unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)std::max<int64_t>(1, ceil_div(n, per)); }


void labels_run(cnn_ctx *ctx, const LabelRequest &req, cnn_label_stats *st) {
    hipStream_t s = ctx->stream;
    const int64_t n = req.n;
    if (req.c_sim < 0 || req.m_core < 0) throw InvalidArg{"c_sim and m_core must be >= 0"};
    if (req.sim_mode != CNN_SIM_AUTO && req.sim_mode != CNN_SIM_MERGE && req.sim_mode != CNN_SIM_BITSET)
        throw InvalidArg{"unknown sim_mode"};
    cnn_label_stats local{};
    if (n == 0) {
        if (st) *st = local;
        return;
    }
    Scratch<uint8_t> core(ctx, n);
    Scratch<uint64_t> rank(ctx, n);
    Scratch<int> parent(ctx, n);
    // the pass's small counters in one buffer, zeroed by one fill kernel:
    // counters (8), striped test statistics (4 per slot), then one 8-byte slot
    // each for the hook queue length, the round queue length and the root count
    constexpr int64_t kSmallSlots = 8 + 4 * kStatSlots + 3;
    Scratch<unsigned long long> small(ctx, kSmallSlots);
    fill_u64(ctx, small.p, 0ull, kSmallSlots);
    unsigned long long *const counters_p = small.p, *const tstats_p = small.p + 8;
    unsigned long long *const slot_p = small.p + 8 + 4 * kStatSlots;
    LabelArgs a{req.row_ptr, req.col, n, req.c_sim, req.m_core, core.p, parent.p, tstats_p, rank.p, nullptr,
                nullptr, req.selfpos, req.tile_csim, req.tile_mcore};

    CNN_HIP(hipEventRecord(ctx->ev[0], s));
    // similarity form: bitset intersections of row words (the spatial graph's
    // tile-pair masks, or per-row word lists built here from the lists), or the
    // sorted-list merge
    const bool tile_words = req.mask && req.pair_nz && req.tp_ptr;
    bool words = false;
    std::unique_ptr<Scratch<int64_t>> wptr;
    std::unique_ptr<Scratch<int32_t>> wtile;
    std::unique_ptr<Scratch<uint64_t>> wword;
    if (req.sim_mode != CNN_SIM_MERGE && tile_words) {
        words = true;
        a.mask = req.mask;
        a.pair_nz = req.pair_nz;
        a.tp_ptr = req.tp_ptr;
        a.tp_col = req.tp_col;
        a.diag = req.diag;
        a.tile_order = req.tile_order;
    } else if (req.sim_mode != CNN_SIM_MERGE && req.col) {
        bool build = req.sim_mode == CNN_SIM_BITSET;
        if (!build && req.words_min_fill > 0 && req.words_min_fill <= (int64_t)1 << 24) {
            // automatic choice: the fill of every stride-th row decides, the
            // stride at least 16 and large enough to sample about 2^21 entries
            const int64_t nnz = read_scalar(ctx, req.row_ptr + n);
            const int64_t stride = std::max<int64_t>(16, ceil_div(nnz, (int64_t)1 << 21));
            Scratch<unsigned long long> sums(ctx, 2);
            fill_u64(ctx, sums.p, 0, 2);
            const int64_t rows = ceil_div(n, stride);
            k_row_word_sample<<<blocks_for(rows, 4), 256, 0, s>>>(req.row_ptr, req.col, n, stride, sums.p);
            CNN_HIP(hipGetLastError());
            PinnedReads rd(ctx);
            const unsigned long long *h = rd.get(sums.p, 2);
            rd.wait();
            build = h[1] > 0 && h[0] >= (unsigned long long)req.words_min_fill * h[1];
        }
        if (build) {
            wptr.reset(new Scratch<int64_t>(ctx, n + 1));
            Scratch<int32_t> wcnt(ctx, n);
            k_row_word_count<<<blocks_for(n, 4), 256, 0, s>>>(req.row_ptr, req.col, n, wcnt.p);
            CNN_HIP(hipGetLastError());
            exclusive_scan_i32_to_i64(ctx, wcnt.p, wptr->p, n);
            PinnedReads rd(ctx);
            const int64_t *last_excl = rd.get(wptr->p + (n - 1));
            const int32_t *last_cnt = rd.get(wcnt.p + (n - 1));
            rd.wait();
            const int64_t n_words = *last_excl + *last_cnt;
            set_i64(ctx, wptr->p + n, n_words);
            wtile.reset(new Scratch<int32_t>(ctx, n_words));
            wword.reset(new Scratch<uint64_t>(ctx, n_words));
            k_row_word_fill<<<blocks_for(n, 4), 256, 0, s>>>(req.row_ptr, req.col, n, wptr->p, wtile->p, wword->p);
            CNN_HIP(hipGetLastError());
            words = true;
            a.wptr = wptr->p;
            a.wtile = wtile->p;
            a.wword = wword->p;
        }
    }
    if (!words && !req.col) throw InvalidArg{"the sorted-list merge needs the neighbour lists"};

    k_init<<<blocks_for(n), 256, 0, s>>>(a, req.consider, req.perm);
    CNN_HIP(hipGetLastError());
    CNN_HIP(hipEventRecord(ctx->ev[8], s));
    if (words && a.mask) {
        // adjacent rows per row tile, then the rows left, one wave each
        Scratch<int32_t> hq(ctx, n);
        unsigned int *const hq_len = (unsigned int *)(slot_p + 0);  // zeroed with the counters
        k_hook_tile<<<blocks_for(req.n_tiles, 4), 256, 0, s>>>(a, req.n_tiles, hq.p, hq_len);
        CNN_HIP(hipGetLastError());
        const unsigned qblocks = (unsigned)std::min<int64_t>(blocks_for(n, 4), 4096);
        k_hook_queue<<<qblocks, 256, 0, s>>>(a, hq.p, hq_len, 2);
    } else if (words) {
        k_hook<true><<<blocks_for(n, 4), 256, 0, s>>>(a, 2);
    } else {
        k_hook<false><<<blocks_for(n, 4), 256, 0, s>>>(a, 2);
    }
    CNN_HIP(hipGetLastError());
    CNN_HIP(hipEventRecord(ctx->ev[9], s));
    // flatten (every parent becomes its root), then let roots re-point to smaller
    // neighbouring roots root_passes times, flattening after each pass; each
    // flattening before a pass lists the core roots, and only they get a wave
    if (req.root_passes <= 0) {
        k_compress<<<blocks_for(n), 256, 0, s>>>(parent.p, n);
    } else {
        Scratch<int32_t> roots(ctx, n);
        unsigned int *const n_roots = (unsigned int *)(slot_p + 2);  // zeroed with the counters
        const unsigned rblocks = (unsigned)std::min<int64_t>(blocks_for(n, 4), 4096);
        for (int pass = 0; pass < req.root_passes; ++pass) {
            if (pass > 0) fill_u64(ctx, slot_p + 2, 0ull, 1);
            k_compress_roots<<<blocks_for(n), 256, 0, s>>>(parent.p, core.p, n, roots.p, n_roots);
            if (words) k_hook_roots_list<true><<<rblocks, 256, 0, s>>>(a, roots.p, n_roots, 4);
            else k_hook_roots_list<false><<<rblocks, 256, 0, s>>>(a, roots.p, n_roots, 4);
        }
        k_compress<<<blocks_for(n), 256, 0, s>>>(parent.p, n);
    }
    CNN_HIP(hipGetLastError());
    {
        Scratch<int32_t> cnt(ctx, n);
        fill_i32(ctx, cnt.p, 0, n);
        k_root_count<<<blocks_for(n, 256 * kRunSteps), 256, 0, s>>>(core.p, parent.p, n, cnt.p);
        k_count_roots<<<blocks_for(n), 256, 0, s>>>(cnt.p, n, counters_p + 6);
        k_snapshot<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, cnt.p, n, rank.p);
        CNN_HIP(hipGetLastError());
    }
    std::unique_ptr<Scratch<uint64_t>> cand_max, tile_common;
    if (req.n_tiles > 0) {
        Scratch<uint64_t> tmax(ctx, req.n_tiles);
        cand_max.reset(new Scratch<uint64_t>(ctx, req.n_tiles));
        tile_common.reset(new Scratch<uint64_t>(ctx, req.n_tiles));
        k_tile_rank_max<<<blocks_for(req.n_tiles, 4), 256, 0, s>>>(rank.p, n, req.n_tiles, tmax.p,
                                                                   tile_common->p);
        k_cand_max<<<blocks_for(req.n_tiles, 1), 256, 0, s>>>(tmax.p, req.tp_ptr, req.tp_col, req.n_tiles,
                                                              cand_max->p);
        CNN_HIP(hipGetLastError());
        a.cand_max = cand_max->p;
        a.tile_rank = tile_common->p;
    }
    CNN_HIP(hipEventRecord(ctx->ev[10], s));
    if (words && a.mask && req.n_tiles > 0) {
        // rows with tests to run go to a queue of at most kRoundQueueMax
        // entries (beyond it the tile kernel tests them itself)
        constexpr int64_t kRoundQueueMax = int64_t(1) << 22;
        const int split = round_split(req.n_tile_pairs, req.n_tiles);
        const int64_t cap = std::min<int64_t>(std::max<int64_t>(n, 64), kRoundQueueMax);
        Scratch<int4> rq(ctx, cap);
        a.rq = rq.p;
        a.rq_len = slot_p + 1;  // zeroed with the counters
        a.rq_cap = cap;
        k_round_tile<<<blocks_for(req.n_tiles * split, 4), 256, 0, s>>>(a, req.n_tiles, split);
        k_round_queue<<<(unsigned)std::min<int64_t>(blocks_for(n, 4), 4096), 256, 0, s>>>(a);
        a.rq = nullptr;
        a.rq_len = nullptr;
        a.rq_cap = 0;
    } else if (words) k_round_b<true><<<blocks_for(n, 4), 256, 0, s>>>(a);
    else k_round_b<false><<<blocks_for(n, 4), 256, 0, s>>>(a);
    CNN_HIP(hipEventRecord(ctx->ev[11], s));
    CNN_HIP(hipGetLastError());
    k_compress<<<blocks_for(n), 256, 0, s>>>(parent.p, n);
    CNN_HIP(hipEventRecord(ctx->ev[1], s));

    Scratch<int32_t> cmin(ctx, n), csize(ctx, n);
    fill_i32(ctx, cmin.p, INT_MAX, n);
    fill_i32(ctx, csize.p, 0, n);
    k_comp_stats<<<blocks_for(n, 256 * kRunSteps), 256, 0, s>>>(core.p, parent.p, req.perm, n, cmin.p, csize.p,
                                                                counters_p + 2);
    CNN_HIP(hipGetLastError());
    // the cluster keys outlive the sort block: a fit with more clusters than
    // k_rank_clusters_small takes sorts them again after the final read
    std::unique_ptr<Scratch<uint64_t>> skeys;
    std::unique_ptr<Scratch<int32_t>> svals, snewlab;
    bool big_sort = false;
    if (req.sort) {
        skeys.reset(new Scratch<uint64_t>(ctx, n));
        svals.reset(new Scratch<int32_t>(ctx, n));
        snewlab.reset(new Scratch<int32_t>(ctx, n));
        Scratch<uint64_t> &keys = *skeys;
        Scratch<int32_t> &vals = *svals, &newlab = *snewlab;
        fill_i32(ctx, newlab.p, 0, n);
        if (req.packed) {
            // (only the packed keys: their top bits are the dataset, which
            // k_first_of_set and k_rank_clusters_packed use as an index)
            const PackedSort &ps = *req.packed;
            PackedKey pk{ps.ds_shift, ps.size_shift, ps.size_max, ps.tile_ds, ps.ds_offset};
            k_collect_clusters_packed<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, cmin.p, csize.p, n,
                                                                    req.member_cutoff, pk, counters_p + 5, keys.p,
                                                                    vals.p);
            CNN_HIP(hipGetLastError());
            const int64_t K = (int64_t)read_scalar(ctx, counters_p + 5);
            sort_pairs_u64_i32(ctx, keys.p, vals.p, K, ps.key_bits);
            if (K > 0) {
                Scratch<int32_t> first(ctx, ps.n_sets);
                k_first_of_set<<<blocks_for(K), 256, 0, s>>>(keys.p, K, ps.ds_shift, first.p);
                k_rank_clusters_packed<<<blocks_for(K), 256, 0, s>>>(keys.p, vals.p, K, ps.ds_shift, first.p,
                                                                     req.max_clusters, newlab.p, counters_p + 7);
                CNN_HIP(hipGetLastError());
            }
            k_write_sorted<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, newlab.p, req.perm, n, req.labels);
            CNN_HIP(hipGetLastError());
            local.n_clusters = (int64_t)read_scalar(ctx, counters_p + 7);
        } else {
            k_collect_clusters<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, cmin.p, csize.p, n,
                                                             req.member_cutoff, counters_p + 5, keys.p, vals.p);
            // up to kSmallClusters clusters sorted on the device; counters[4]
            // flags more, which the host sorts after the final read below
            k_rank_clusters_small<<<1, kSmallClusters, 0, s>>>(keys.p, vals.p, counters_p + 5, req.max_clusters,
                                                               newlab.p, counters_p + 4);
            k_write_sorted<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, newlab.p, req.perm, n, req.labels);
            CNN_HIP(hipGetLastError());
            big_sort = true;
        }
    } else {
        Scratch<uint8_t> isfirst(ctx, n);
        Scratch<int32_t> rank(ctx, n);
        fill_u8(ctx, isfirst.p, 0, n);
        k_mark_first<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, cmin.p, n, isfirst.p);
        CNN_HIP(hipGetLastError());
        inclusive_scan_u8_to_i32(ctx, isfirst.p, rank.p, n);
        k_write_unsorted<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, cmin.p, rank.p, req.perm, n, req.current_start,
                                                       req.write_all, req.labels);
        CNN_HIP(hipGetLastError());
    }
    if (req.consider) fill_u8(ctx, req.consider, 0, n);
    CNN_HIP(hipEventRecord(ctx->ev[2], s));
    unsigned long long h[8];
    unsigned long long ht[4 * kStatSlots];
    {
        PinnedReads rd(ctx);
        const unsigned long long *hp = rd.get(counters_p, 8);
        const unsigned long long *htp = rd.get(tstats_p, 4 * kStatSlots);
        rd.wait();
        std::copy(hp, hp + 8, h);
        std::copy(htp, htp + 4 * kStatSlots, ht);
    }
    if (big_sort) {
        const int64_t K = (int64_t)h[5];
        if (h[4]) {
            // more clusters than one workgroup sorts: sort them here and write the labels again
            fill_i32(ctx, snewlab->p, 0, n);
            sort_pairs_u64_i32(ctx, skeys->p, svals->p, K, 63);
            k_rank_clusters<<<blocks_for(K), 256, 0, s>>>(svals->p, K, req.max_clusters, snewlab->p);
            k_write_sorted<<<blocks_for(n), 256, 0, s>>>(core.p, parent.p, snewlab->p, req.perm, n, req.labels);
            CNN_HIP(hipGetLastError());
            CNN_HIP(hipStreamSynchronize(s));
        }
        local.n_clusters = (req.max_clusters >= 0) ? std::min<int64_t>(K, req.max_clusters) : K;
    }
    h[0] = h[1] = 0;
    unsigned long long rows = 0;
    for (int i = 0; i < kStatSlots; ++i) {
        h[0] += ht[4 * i];
        h[1] += ht[4 * i + 1];
        rows += ht[4 * i + 2];
    }
    local.sampled_components = (int64_t)h[6];
    local.rows_scanned = (int64_t)rows;
    float t1 = 0, t2 = 0;
    CNN_HIP(hipEventElapsedTime(&t1, ctx->ev[0], ctx->ev[1]));
    CNN_HIP(hipEventElapsedTime(&t2, ctx->ev[1], ctx->ev[2]));
    local.tests = (int64_t)h[0];
    local.tests_passed = (int64_t)h[1];
    local.n_core = (int64_t)h[2];
    local.n_components = (int64_t)h[3];
    if (!req.sort) local.n_clusters = local.n_components;
    local.ms_similarity = t1;
    local.ms_labels = t2;
    CNN_HIP(hipEventElapsedTime(&local.ms_hook, ctx->ev[8], ctx->ev[9]));
    CNN_HIP(hipEventElapsedTime(&local.ms_round, ctx->ev[10], ctx->ev[11]));
    local.sim_form = words ? (tile_words ? CNN_SIM_FORM_TILE_WORDS : CNN_SIM_FORM_ROW_WORDS) : CNN_SIM_FORM_MERGE;
    if (st) *st = local;
}

LabelRequest graph_label_request(cnn_ctx *ctx, const cnn_graph *g, int32_t sim_mode) {
    if (g->n > 0 && !g->col && (sim_mode == CNN_SIM_MERGE || !g->mask)) graph_lists(ctx, g);
    LabelRequest r{};
    r.row_ptr = g->row_ptr;
    r.col = g->col;
    r.n = g->n;
    r.perm = g->perm;
    r.sim_mode = sim_mode;
    r.current_start = 1;
    r.member_cutoff = 2;
    r.max_clusters = -1;
    r.tp_ptr = g->tp_ptr;
    r.tp_col = g->tp_col;
    r.n_tiles = g->n_tiles;
    r.selfpos = g->selfpos;
    r.mask = g->mask;
    r.pair_nz = g->pair_nz;
    r.diag = g->diag;
    r.tile_order = g->tile_order;
    r.n_tile_pairs = g->n_tile_pairs;
    r.root_passes = g->tune.root_passes;
    r.words_min_fill = g->tune.words_min_fill;
    return r;
}

// ------------------------------------------------------------------ several cutoff pairs, no host synchronisation

// sort_by_size of every cutoff pair without a library sort (a captured graph
// may hold no memset node, and library sorts record some): the clusters of
// pair y (roots of components with at least member_cutoff points) are
// collected with their keys (size descending, smallest original index
// ascending; all distinct), and each cluster's rank is the number of keys
// below its own, counted against LDS-staged chunks of the pair's keys --
// K^2 / 256 comparisons per block for K clusters.
__global__ void k_collect_multi(const uint8_t *core, const int *parent, const int32_t *cmin, const int32_t *csize,
                                int64_t n, int64_t member_cutoff, uint64_t *ckey, int32_t *cidx,
                                unsigned int *ccount) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t y = blockIdx.y, off = y * n;
    const bool cluster = w < n && core[off + w] && parent[off + w] == (int)w &&
                         (int64_t)csize[off + w] >= member_cutoff;
    const uint64_t b = __ballot(cluster);
    if (!b) return;
    const int lane = threadIdx.x & 63;
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(&ccount[y], (unsigned)__popcll(b));
    base = __shfl(base, 0, 64);
    if (cluster) {
        const unsigned pos = base + (unsigned)__popcll(b & ((1ull << lane) - 1ull));
        ckey[off + pos] = ((uint64_t)(uint32_t)(0x7fffffff - csize[off + w]) << 32) | (uint32_t)cmin[off + w];
        cidx[off + pos] = (int32_t)w;
    }
}

__global__ __launch_bounds__(256) void k_rank_count(const uint64_t *ckey, const int32_t *cidx,
                                                    const unsigned int *ccount, int64_t n, int64_t max_clusters,
                                                    int32_t *newlab, unsigned long long *counters) {
    __shared__ uint64_t chunk[256];
    const int64_t y = blockIdx.y, off = y * n;
    const int64_t K = ccount[y];
    const int64_t x = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if ((int64_t)blockIdx.x * 256 >= K) return;  // whole block: uniform
    const uint64_t mine = x < K ? ckey[off + x] : ~0ull;
    int64_t rank = 0;
    for (int64_t base = 0; base < K; base += 256) {
        chunk[threadIdx.x] = base + threadIdx.x < K ? ckey[off + base + threadIdx.x] : ~0ull;
        __syncthreads();
        const int m = (int)min<int64_t>(256, K - base);
        for (int t = 0; t < m; ++t) rank += chunk[t] < mine ? 1 : 0;
        __syncthreads();
    }
    const bool keep = x < K && (max_clusters < 0 || rank < max_clusters);
    if (keep) newlab[off + cidx[off + x]] = (int32_t)(rank + 1);
    const uint64_t b = __ballot(keep);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&counters[8 * y + 7], (unsigned long long)__popcll(b));
}

void multi_labels_bytes(int64_t n, int64_t n_tiles, int n_c, MultiLabelSizes &sz) {
    sz.per_point = n * n_c;
    sz.per_tile = std::max<int64_t>(n_tiles, 1) * n_c;
}

void labels_enqueue_multi(cnn_ctx *ctx, const MultiLabelBufs &b, const MultiLabelSource &src, int root_passes,
                          int64_t member_cutoff, int64_t max_clusters, int64_t *labels, unsigned long long *counters,
                          unsigned long long *tstats) {
    hipStream_t s = ctx->stream;
    const int64_t n = b.n, T = b.n_tiles;
    const unsigned nc = (unsigned)b.n_c;
    if (n == 0 || nc == 0) return;
    fill_u64(ctx, counters, 0, 8 * (int64_t)nc);
    fill_u64(ctx, tstats, 0, 4 * (int64_t)kStatSlots * nc);
    fill_i32(ctx, (int32_t *)b.hq_len, 0, nc);
    fill_i32(ctx, (int32_t *)b.ccount, 0, nc);
    fill_i32(ctx, b.cnt, 0, n * nc);
    fill_i32(ctx, b.cmin, INT_MAX, n * nc);
    fill_i32(ctx, b.csize, 0, n * nc);
    fill_i32(ctx, b.newlab, 0, n * nc);
    LabelArgs a{};
    a.n = n;
    a.core = b.core;
    a.parent = b.parent;
    a.stats = tstats;
    a.rank = b.rank;
    a.mask = src.mask;
    a.pair_nz = src.pair_nz;
    a.tp_ptr = src.tp_ptr;
    a.tp_col = src.tp_col;
    a.deg = src.deg;
    a.csim_y = b.csim;
    a.mcore_y = b.mcore;
    a.tiles = T;
    auto g2 = [&](int64_t items, int per) { return dim3(blocks_for(items, per), nc); };
    k_init<<<g2(n, 256), 256, 0, s>>>(a, nullptr, src.perm);
    k_hook_tile<<<g2(T, 4), 256, 0, s>>>(a, T, b.hq, b.hq_len);
    k_hook_queue<<<dim3((unsigned)std::min<int64_t>(blocks_for(n, 4), 4096), nc), 256, 0, s>>>(a, b.hq, b.hq_len,
                                                                                                   2);
    k_compress<<<g2(n, 256), 256, 0, s>>>(b.parent, n);
    for (int pass = 0; pass < root_passes; ++pass) {
        k_hook_roots<true><<<g2(n, 4), 256, 0, s>>>(a, 4);
        k_compress<<<g2(n, 256), 256, 0, s>>>(b.parent, n);
    }
    CNN_HIP(hipGetLastError());
    k_root_count<<<g2(n, 256 * kRunSteps), 256, 0, s>>>(b.core, b.parent, n, b.cnt);
    k_snapshot<<<g2(n, 256), 256, 0, s>>>(b.core, b.parent, b.cnt, n, b.rank);
    k_tile_rank_max<<<g2(T, 4), 256, 0, s>>>(b.rank, n, T, b.tmax, b.tile_common);
    k_cand_max<<<g2(T, 1), 256, 0, s>>>(b.tmax, src.tp_ptr, src.tp_col, T, b.cand_max);
    CNN_HIP(hipGetLastError());
    a.cand_max = b.cand_max;
    a.tile_rank = b.tile_common;
    k_round_b<true><<<g2(n, 4), 256, 0, s>>>(a);
    k_compress<<<g2(n, 256), 256, 0, s>>>(b.parent, n);
    k_comp_stats<<<g2(n, 256 * kRunSteps), 256, 0, s>>>(b.core, b.parent, src.perm, n, b.cmin, b.csize, counters);
    k_collect_multi<<<g2(n, 256), 256, 0, s>>>(b.core, b.parent, b.cmin, b.csize, n, member_cutoff, b.ckey,
                                               b.cidx, b.ccount);
    k_rank_count<<<g2(n, 256), 256, 0, s>>>(b.ckey, b.cidx, b.ccount, n, max_clusters, b.newlab, counters);
    k_write_sorted<<<g2(n, 256), 256, 0, s>>>(b.core, b.parent, b.newlab, src.perm, n, labels);
    CNN_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ caller CSR validation

// Flags rows that are not strictly ascending or hold indices outside [0, n),
// and offsets that decrease.  Precondition (caller contract): col holds at
// least row_ptr[n] entries and row_ptr is n+1 long.
__global__ void k_validate_csr(const int64_t *__restrict__ row_ptr, const int32_t *__restrict__ col, int64_t n,
                               int64_t nnz_cap, unsigned long long *bad) {
    int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const int64_t b = row_ptr[w], e = row_ptr[w + 1];
    if (b < 0 || e < b || e > nnz_cap) { atomicOr(bad, 1ull); return; }
    int32_t prev = -1;
    for (int64_t t = b; t < e; ++t) {
        const int32_t j = col[t];
        if (j <= prev || (int64_t)j >= n) { atomicOr(bad, 2ull); return; }
        prev = j;
    }
}

void validate_csr(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n) {
    if (n == 0) return;
    hipStream_t s = ctx->stream;
    const int64_t first = read_scalar(ctx, row_ptr);
    const int64_t nnz = read_scalar(ctx, row_ptr + n);
    if (first != 0 || nnz < 0) throw InvalidArg{"row_ptr must start at 0 and end at nnz >= 0"};
    Scratch<unsigned long long> bad(ctx, 1);
    CNN_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), s));
    k_validate_csr<<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(row_ptr, col, n, nnz, bad.p);
    CNN_HIP(hipGetLastError());
    const unsigned long long flags = read_scalar(ctx, bad.p);
    if (flags & 1ull) throw InvalidArg{"row_ptr is not non-decreasing within [0, nnz]"};
    if (flags & 2ull) throw InvalidArg{"neighbour rows must be strictly ascending with indices in [0, n)"};
}

// ------------------------------------------------------------------ generic sort_by_size

__global__ void k_iota(int32_t *v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

__global__ void k_heads(const int64_t *keys, int64_t n, uint8_t *head) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

__global__ void k_segments(const int64_t *keys, const int32_t *vals, const int32_t *segid, int64_t n,
                           int32_t *seg_start, int32_t *seg_end) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int sid = segid[t] - 1;
    if (t == 0 || keys[t] != keys[t - 1]) seg_start[sid] = (int32_t)t;
    if (t == n - 1 || keys[t + 1] != keys[t]) seg_end[sid] = (int32_t)t;
}

__global__ void k_seg_collect(const int64_t *keys, const int32_t *vals, const int32_t *seg_start,
                              const int32_t *seg_end, int64_t n_seg, unsigned long long *K, uint64_t *okeys,
                              int32_t *ovals) {
    int64_t sid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (sid >= n_seg) return;
    const int32_t b = seg_start[sid];
    if (keys[b] == 0) return;
    const int32_t sz = seg_end[sid] - b + 1;
    const unsigned long long idx = atomicAdd(K, 1ull);
    okeys[idx] = ((uint64_t)(uint32_t)(0x7fffffff - sz) << 32) | (uint32_t)vals[b];
    ovals[idx] = (int32_t)sid;
}

__global__ void k_seg_rank(const uint64_t *okeys, const int32_t *ovals, int64_t K, int64_t member_cutoff,
                           int64_t max_clusters, int32_t *seglab) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= K) return;
    const int64_t sz = 0x7fffffff - (int64_t)(okeys[t] >> 32);
    const bool keep = (max_clusters < 0 || t < max_clusters) && sz >= member_cutoff;
    seglab[ovals[t]] = keep ? (int32_t)(t + 1) : 0;
}

__global__ void k_seg_write(const int32_t *vals, const int32_t *segid, const int32_t *seglab, int64_t n,
                            int64_t *labels) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    labels[vals[t]] = seglab[segid[t] - 1];
}

int64_t sort_by_size_generic(cnn_ctx *ctx, int64_t *labels, int64_t n, int64_t member_cutoff, int64_t max_clusters) {
    hipStream_t s = ctx->stream;
    if (n == 0) return 0;
    if (n >= (int64_t)INT32_MAX) throw InvalidArg{"n exceeds the int32 index range"};
    Scratch<int64_t> keys(ctx, n);
    Scratch<int32_t> vals(ctx, n), segid(ctx, n);
    Scratch<uint8_t> head(ctx, n);
    CNN_HIP(hipMemcpyAsync(keys.p, labels, n * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    k_iota<<<blocks_for(n), 256, 0, s>>>(vals.p, n);
    CNN_HIP(hipGetLastError());
    sort_pairs_i64_i32(ctx, keys.p, vals.p, n);  // stable: equal labels keep ascending index
    k_heads<<<blocks_for(n), 256, 0, s>>>(keys.p, n, head.p);
    inclusive_scan_u8_to_i32(ctx, head.p, segid.p, n);
    const int64_t n_seg = read_scalar(ctx, segid.p + (n - 1));
    Scratch<int32_t> seg_start(ctx, n_seg), seg_end(ctx, n_seg), seglab(ctx, n_seg);
    k_segments<<<blocks_for(n), 256, 0, s>>>(keys.p, vals.p, segid.p, n, seg_start.p, seg_end.p);
    fill_i32(ctx, seglab.p, 0, n_seg);
    Scratch<unsigned long long> K(ctx, 1);
    CNN_HIP(hipMemsetAsync(K.p, 0, sizeof(unsigned long long), s));
    Scratch<uint64_t> okeys(ctx, n_seg);
    Scratch<int32_t> ovals(ctx, n_seg);
    k_seg_collect<<<blocks_for(n_seg), 256, 0, s>>>(keys.p, vals.p, seg_start.p, seg_end.p, n_seg, K.p, okeys.p,
                                                    ovals.p);
    CNN_HIP(hipGetLastError());
    const int64_t k = (int64_t)read_scalar(ctx, K.p);
    sort_pairs_u64_i32(ctx, okeys.p, ovals.p, k, 63);
    if (k > 0) k_seg_rank<<<blocks_for(k), 256, 0, s>>>(okeys.p, ovals.p, k, member_cutoff, max_clusters, seglab.p);
    k_seg_write<<<blocks_for(n), 256, 0, s>>>(vals.p, segid.p, seglab.p, n, labels);
    CNN_HIP(hipGetLastError());
    // number of survivors
    int64_t kept = 0;
    if (k > 0) {
        std::vector<uint64_t> hk(k);
        CNN_HIP(hipMemcpyAsync(hk.data(), okeys.p, k * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        CNN_HIP(hipStreamSynchronize(s));
        for (int64_t t = 0; t < k; ++t) {
            const int64_t sz = 0x7fffffff - (int64_t)(hk[t] >> 32);
            if ((max_clusters < 0 || t < max_clusters) && sz >= member_cutoff) ++kept;
        }
    } else {
        CNN_HIP(hipStreamSynchronize(s));
    }
    return kept;
}

}  // namespace cnn

// ================================================================== C-ABI: phase S+U
using namespace cnn;

extern "C" {

int cnn_labels_from_csr(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n, int32_t c_sim,
                        int32_t m_core, uint8_t *consider, int64_t current_start, int32_t sim_mode, int64_t *labels,
                        cnn_label_stats *stats_out) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n < 0) throw InvalidArg{"negative n"};
        if (n > 0 && (!row_ptr || !col || !labels)) throw InvalidArg{"null buffer"};
        if (n >= (int64_t)INT32_MAX) throw InvalidArg{"n exceeds the int32 index range"};
        CNN_HIP(hipSetDevice(ctx->device));
        validate_csr(ctx, row_ptr, col, n);
        LabelRequest r{row_ptr, col, n, nullptr, c_sim, m_core, consider, current_start, sim_mode,
                       false, false, 2, -1, labels};
        const TuneParams t = tune_lookup(0, n, CNN_PLAIN_F64);
        r.root_passes = t.root_passes;
        r.words_min_fill = t.words_min_fill;
        labels_run(ctx, r, stats_out);
        return (int)CNN_OK;
    });
}

int cnn_graph_labels(cnn_ctx *ctx, const cnn_graph *g, int32_t c_sim, int32_t m_core, uint8_t *consider,
                     int64_t current_start, int32_t sim_mode, int64_t *labels, cnn_label_stats *stats_out) {
    if (!ctx || !g) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (g->n > 0 && !labels) throw InvalidArg{"null labels"};
        CNN_HIP(hipSetDevice(ctx->device));
        if (sim_mode != CNN_SIM_AUTO && sim_mode != CNN_SIM_MERGE && sim_mode != CNN_SIM_BITSET)
            throw InvalidArg{"unknown sim_mode"};
        LabelRequest r = graph_label_request(ctx, g, sim_mode);
        r.c_sim = c_sim;
        r.m_core = m_core;
        r.consider = consider;
        r.current_start = current_start;
        r.labels = labels;
        labels_run(ctx, r, stats_out);
        return (int)CNN_OK;
    });
}

int cnn_sort_by_size(cnn_ctx *ctx, int64_t *labels, int64_t n, int64_t member_cutoff, int64_t max_clusters,
                     int64_t *n_clusters_out) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n < 0) throw InvalidArg{"negative n"};
        if (n > 0 && !labels) throw InvalidArg{"null labels"};
        CNN_HIP(hipSetDevice(ctx->device));
        int64_t k = sort_by_size_generic(ctx, labels, n, member_cutoff, max_clusters);
        if (n_clusters_out) *n_clusters_out = k;
        return (int)CNN_OK;
    });
}

int cnn_graph_fit(cnn_ctx *ctx, const cnn_graph *g, int32_t c_sim, int32_t m_core, int32_t sort,
                  int64_t member_cutoff, int64_t max_clusters, int64_t *labels, cnn_label_stats *lstats) {
    if (!ctx || !g) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (g->n > 0 && !labels) throw InvalidArg{"null labels"};
        CNN_HIP(hipSetDevice(ctx->device));
        LabelRequest r = graph_label_request(ctx, g, CNN_SIM_AUTO);
        r.c_sim = c_sim;
        r.m_core = m_core;
        r.sort = sort != 0;
        r.write_all = true;
        r.member_cutoff = member_cutoff;
        r.max_clusters = max_clusters;
        r.labels = labels;
        labels_run(ctx, r, lstats);
        return (int)CNN_OK;
    });
}

int cnn_fit(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, double r2, int32_t c_sim,
            int32_t m_core, int32_t precision, int32_t sort, int64_t member_cutoff, int64_t max_clusters,
            int64_t *labels, cnn_graph **graph_out, cnn_graph_stats *gstats, cnn_label_stats *lstats) {
    if (!ctx) return CNN_ERR_INVALID;
    if (graph_out) *graph_out = nullptr;
    return guarded(ctx, [&]() {
        if (n > 0 && !labels) throw InvalidArg{"null labels"};
        CNN_HIP(hipSetDevice(ctx->device));
        cnn_graph *g = graph_build(ctx, X, n, d, dtype, r2, precision, nullptr, graph_out == nullptr);
        std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
        if (gstats) *gstats = g->stats;
        LabelRequest r = graph_label_request(ctx, g, CNN_SIM_AUTO);
        r.c_sim = c_sim;
        r.m_core = m_core;
        r.sort = sort != 0;
        r.write_all = true;
        r.member_cutoff = member_cutoff;
        r.max_clusters = max_clusters;
        r.labels = labels;
        labels_run(ctx, r, lstats);
        if (graph_out) *graph_out = hold.release();
        return (int)CNN_OK;
    });
}

}  // extern "C"
