// Similarity-weighted neighbour graph and its maximum spanning forest, for
// the MST cluster hierarchy of HierarchicalFitterMSTPrim._fit
// (_fit.pyx:820-915).
//
// Edge weight: w(a, b) = |N(a) ∩ N(b)|, the count SimilarityChecker._get
// returns (_types.pyx:2612-2634), for every b in N(a) (self-counting lists).
// The weights are computed on the device, one wave per row a, one wave-wide
// sorted-list merge per entry.
//
// Spanning forest: Prim's procedure of the reference, run on the host over the
// weighted lists.  Every not yet visited point, in ascending index, starts a
// tree: the edges from the newly visited point to unvisited members are
// queued, and the queue's top edge (a, b) whose b is unvisited joins the tree.
// The reference queue (PriorityQueueMaxHeap, _types.pyx:2825-2851) orders
// entries by (-weight, a, b); each (a, b) is queued at most once (when a is
// visited), so the order is total and any queue with the same order pops the
// same sequence (here: the equivalent eager form, one heap entry per reached
// point).  The tree edges are returned sorted by that same order, the order in
// which the reference pops its tree queue (_fit.pyx:937-939).
#include <algorithm>
#include <functional>
#include <vector>

#include "cnn_internal.hpp"

namespace cnn {

// One wave per row a: weight[t] = |N(a) ∩ N(col[t])| for every entry t of row a.
__global__ __launch_bounds__(256) void k_edge_common(const int64_t *__restrict__ row_ptr,
                                                     const int32_t *__restrict__ col, int64_t n,
                                                     int32_t *__restrict__ weight) {
    const int lane = threadIdx.x & 63;
    const int64_t a = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= n) return;
    const int64_t ba = row_ptr[a], ea = row_ptr[a + 1];
    for (int64_t t = ba; t < ea; ++t) {
        const int32_t b = col[t];
        const int64_t bb = row_ptr[b], eb = row_ptr[b + 1];
        const int c = wave_common(col + ba, ea - ba, col + bb, eb - bb);
        if (lane == 0) weight[t] = c;
    }
}

static void edge_weights(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n, int32_t *weight) {
    if (n == 0) return;
    k_edge_common<<<(unsigned)ceil_div(n, 4), 256, 0, ctx->stream>>>(row_ptr, col, n, weight);
    CNN_HIP(hipGetLastError());
}

struct QueuedEdge {
    int32_t w, a, b;
};

// (-w, a, b) ascending: true when x pops before y.
static bool pops_before(const QueuedEdge &x, const QueuedEdge &y) {
    if (x.w != y.w) return x.w > y.w;
    if (x.a != y.a) return x.a < y.a;
    return x.b < y.b;
}

// Prim's spanning forest over host lists; returns tree edges sorted in pop order.
//
// The reference pops queued edges lazily in (-w, a, b) order, skipping edges
// whose b is already visited.  Its next tree edge is therefore, over all
// unvisited b, the smallest (-w(a, b), a, b) with a visited -- so keeping, per
// unvisited b, only its best (-w, a) and taking the smallest (-w, a, b) gives
// the same sequence ("eager" Prim).  An indexed binary heap holds one entry
// per reached, unvisited point: at most n entries instead of up to nnz.
class VertexHeap {
  public:
    explicit VertexHeap(int64_t n) : pos_((size_t)n, -1), key_((size_t)n, 0), from_((size_t)n, 0) {}
    bool empty() const { return heap_.empty(); }
    // offer edge (a, b) of weight w; keeps b's best (-w, a)
    void offer(int32_t b, int32_t w, int32_t a) {
        const uint64_t k = pack(w, a);
        if (pos_[(size_t)b] < 0) {
            key_[(size_t)b] = k;
            from_[(size_t)b] = a;
            pos_[(size_t)b] = (int64_t)heap_.size();
            heap_.push_back(b);
            up(heap_.size() - 1);
        } else if (k < key_[(size_t)b]) {
            key_[(size_t)b] = k;
            from_[(size_t)b] = a;
            up((size_t)pos_[(size_t)b]);
        }
    }
    QueuedEdge pop() {
        const int32_t b = heap_[0];
        const QueuedEdge e{(int32_t)(0x7fffffff - (int32_t)(key_[(size_t)b] >> 32)), from_[(size_t)b], b};
        const int32_t last = heap_.back();
        heap_.pop_back();
        pos_[(size_t)b] = -2;  // visited: never offered again
        if (!heap_.empty()) {
            heap_[0] = last;
            pos_[(size_t)last] = 0;
            down(0);
        }
        return e;
    }
    bool reached_or_visited(int32_t b) const { return pos_[(size_t)b] != -1; }
    bool visited(int32_t b) const { return pos_[(size_t)b] == -2; }

  private:
    static uint64_t pack(int32_t w, int32_t a) {
        return ((uint64_t)(uint32_t)(0x7fffffff - w) << 32) | (uint32_t)a;  // ascending = (-w, a)
    }
    bool less(int32_t x, int32_t y) const {
        const uint64_t kx = key_[(size_t)x], ky = key_[(size_t)y];
        return kx != ky ? kx < ky : x < y;  // ties: smaller b
    }
    void place(size_t i, int32_t v) {
        heap_[i] = v;
        pos_[(size_t)v] = (int64_t)i;
    }
    void up(size_t i) {
        const int32_t v = heap_[i];
        while (i > 0) {
            const size_t p = (i - 1) / 2;
            if (!less(v, heap_[p])) break;
            place(i, heap_[p]);
            i = p;
        }
        place(i, v);
    }
    void down(size_t i) {
        const int32_t v = heap_[i];
        const size_t n = heap_.size();
        for (;;) {
            size_t c = 2 * i + 1;
            if (c >= n) break;
            if (c + 1 < n && less(heap_[c + 1], heap_[c])) ++c;
            if (!less(heap_[c], v)) break;
            place(i, heap_[c]);
            i = c;
        }
        place(i, v);
    }
    std::vector<int64_t> pos_;    // heap slot; -1 not reached, -2 visited
    std::vector<uint64_t> key_;   // (-w, a) of the best offer
    std::vector<int32_t> from_;   // a of the best offer
    std::vector<int32_t> heap_;
};

static std::vector<QueuedEdge> prim_forest(const std::vector<int64_t> &rp, const std::vector<int32_t> &col,
                                           const std::vector<int32_t> &w, int64_t n) {
    std::vector<QueuedEdge> tree;
    tree.reserve(n > 0 ? (size_t)(n - 1) : 0);
    VertexHeap heap(n);
    auto offer_from = [&](int32_t a) {
        for (int64_t t = rp[a]; t < rp[a + 1]; ++t) {
            const int32_t b = col[t];
            if (!heap.visited(b)) heap.offer(b, w[t], a);
        }
    };
    for (int64_t p = 0; p < n; ++p) {
        if (heap.reached_or_visited((int32_t)p)) continue;
        // a new tree: p is visited without an edge
        heap.offer((int32_t)p, 0, 0);
        heap.pop();
        offer_from((int32_t)p);
        while (!heap.empty()) {
            const QueuedEdge e = heap.pop();
            tree.push_back(e);
            offer_from(e.b);
        }
    }
    std::sort(tree.begin(), tree.end(), pops_before);
    return tree;
}

}  // namespace cnn

using namespace cnn;

extern "C" {

int cnn_csr_edge_similarity(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n, int32_t *weight) {
    if (!ctx) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n < 0) throw InvalidArg{"negative n"};
        if (n >= (int64_t)INT32_MAX) throw InvalidArg{"n exceeds the int32 index range"};
        if (n > 0 && !row_ptr) throw InvalidArg{"null row_ptr"};
        CNN_HIP(hipSetDevice(ctx->device));
        validate_csr(ctx, row_ptr, col, n);
        const int64_t nnz = n ? read_scalar(ctx, row_ptr + n) : 0;
        if (nnz > 0 && (!col || !weight)) throw InvalidArg{"null col or weight"};
        edge_weights(ctx, row_ptr, col, n, weight);
        CNN_HIP(hipStreamSynchronize(ctx->stream));
        return (int)CNN_OK;
    });
}

int cnn_similarity_mst(cnn_ctx *ctx, const int64_t *row_ptr, const int32_t *col, int64_t n, int64_t *tree_a,
                       int64_t *tree_b, int64_t *tree_w, int64_t cap, int64_t *n_edges_out) {
    if (!ctx || !n_edges_out) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        *n_edges_out = 0;
        if (n < 0) throw InvalidArg{"negative n"};
        if (n >= (int64_t)INT32_MAX) throw InvalidArg{"n exceeds the int32 index range"};
        if (n > 0 && !row_ptr) throw InvalidArg{"null row_ptr"};
        if (n == 0) return (int)CNN_OK;
        CNN_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        validate_csr(ctx, row_ptr, col, n);
        const int64_t nnz = read_scalar(ctx, row_ptr + n);
        if (nnz > 0 && !col) throw InvalidArg{"null col"};
        Scratch<int32_t> wd(ctx, (size_t)std::max<int64_t>(nnz, 1));
        edge_weights(ctx, row_ptr, col, n, wd.p);
        std::vector<int64_t> rp((size_t)n + 1);
        std::vector<int32_t> hc((size_t)nnz), hw((size_t)nnz);
        CNN_HIP(hipMemcpyAsync(rp.data(), row_ptr, (n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
        if (nnz) {
            CNN_HIP(hipMemcpyAsync(hc.data(), col, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
            CNN_HIP(hipMemcpyAsync(hw.data(), wd.p, nnz * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        }
        CNN_HIP(hipStreamSynchronize(s));
        const std::vector<QueuedEdge> tree = prim_forest(rp, hc, hw, n);
        *n_edges_out = (int64_t)tree.size();
        if ((int64_t)tree.size() > cap)
            throw Capacity{"cap " + std::to_string(cap) + " < tree edges " + std::to_string(tree.size())};
        if (!tree.empty() && (!tree_a || !tree_b || !tree_w)) throw InvalidArg{"null tree output"};
        for (size_t k = 0; k < tree.size(); ++k) {
            tree_a[k] = tree[k].a;
            tree_b[k] = tree[k].b;
            tree_w[k] = tree[k].w;
        }
        return (int)CNN_OK;
    });
}

}  // extern "C"
