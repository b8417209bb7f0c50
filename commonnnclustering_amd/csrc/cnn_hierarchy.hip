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
// same sequence (here: buckets by weight, a heap on (a, b) in each).  The
// tree edges are returned sorted by that same order, the order in which the
// reference pops its tree queue (_fit.pyx:937-939).
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

// Max-weight queue of edges as buckets by weight (the weights are counts no
// larger than a list length), each bucket a binary min-heap of a << 32 | b:
// pops in (-w, a, b) order, as a single heap with that comparison would.
class BucketQueue {
  public:
    bool empty() const { return count_ == 0; }
    void push(int32_t w, int32_t a, int32_t b) {
        if ((size_t)w >= buckets_.size()) buckets_.resize((size_t)w + 1);
        std::vector<uint64_t> &h = buckets_[(size_t)w];
        h.push_back(((uint64_t)(uint32_t)a << 32) | (uint32_t)b);
        std::push_heap(h.begin(), h.end(), std::greater<uint64_t>());
        if (w > top_) top_ = w;
        ++count_;
    }
    QueuedEdge pop() {
        while (buckets_[(size_t)top_].empty()) --top_;
        std::vector<uint64_t> &h = buckets_[(size_t)top_];
        std::pop_heap(h.begin(), h.end(), std::greater<uint64_t>());
        const uint64_t key = h.back();
        h.pop_back();
        --count_;
        return QueuedEdge{top_, (int32_t)(key >> 32), (int32_t)(uint32_t)key};
    }

  private:
    std::vector<std::vector<uint64_t>> buckets_;
    int32_t top_ = -1;
    size_t count_ = 0;
};

// Prim's spanning forest over host lists; returns tree edges sorted in pop order.
static std::vector<QueuedEdge> prim_forest(const std::vector<int64_t> &rp, const std::vector<int32_t> &col,
                                           const std::vector<int32_t> &w, int64_t n) {
    std::vector<uint8_t> visited((size_t)n, 0);
    std::vector<QueuedEdge> tree;
    tree.reserve(n > 0 ? (size_t)(n - 1) : 0);
    BucketQueue queue;
    auto push_from = [&](int32_t a) {
        for (int64_t t = rp[a]; t < rp[a + 1]; ++t) {
            const int32_t b = col[t];
            if (!visited[b]) queue.push(w[t], a, b);
        }
    };
    for (int64_t p = 0; p < n; ++p) {
        if (visited[p]) continue;
        visited[p] = 1;
        push_from((int32_t)p);
        while (!queue.empty()) {
            const QueuedEdge e = queue.pop();
            if (visited[e.b]) continue;
            visited[e.b] = 1;
            tree.push_back(e);
            push_from(e.b);
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
