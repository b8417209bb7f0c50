// Batched fits over many independent datasets (cnn_fit_batch).
//
// The reference fits one dataset per Clustering.fit call (cluster.pyx:256-392);
// parameter scans loop over such calls (HierarchicalFitterRepeat._fit,
// _fit.pyx:1058-1187).  Here each dataset is a row range of one device buffer
// and is fitted with its own radius and cutoffs.
#include <algorithm>
#include <memory>
#include <vector>

#include <atomic>
#include <exception>
#include <mutex>
#include <thread>

#include "cnn_internal.hpp"

using namespace cnn;

// Batches of low-dimensional datasets (1 <= d <= 8) are fitted in one working
// order (fit_packed); the others go through the per-dataset loop below.
constexpr bool kPackedBatches = true;

// Bits needed to hold values 0 .. v.
static int bits_for(int64_t v) {
    int b = 0;
    while (b < 63 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

static void add_graph_stats(cnn_graph_stats &gs, const cnn_graph_stats &o) {
    gs.pairs_evaluated += o.pairs_evaluated;
    gs.band_hits += o.band_hits;
    gs.band_flips += o.band_flips;
    gs.max_abs_err_band = std::max(gs.max_abs_err_band, o.max_abs_err_band);
    gs.tile_pairs += o.tile_pairs;
    gs.nnz += o.nnz;
    gs.ms_prep += o.ms_prep;
    gs.ms_distance += o.ms_distance;
    gs.ms_emit += o.ms_emit;
    gs.ms_distance_kernel += o.ms_distance_kernel;
    gs.distance_launches += o.distance_launches;
    gs.tile_pairs_computed += o.tile_pairs_computed;
    gs.layout = o.layout;
}

static void add_label_stats(cnn_label_stats &ls, const cnn_label_stats &o) {
    ls.n_core += o.n_core;
    ls.n_components += o.n_components;
    ls.n_clusters += o.n_clusters;
    ls.tests += o.tests;
    ls.tests_passed += o.tests_passed;
    ls.sampled_components += o.sampled_components;
    ls.rows_scanned += o.rows_scanned;
    ls.ms_similarity += o.ms_similarity;
    ls.ms_labels += o.ms_labels;
    ls.sim_form = std::max(ls.sim_form, o.sim_form);
    ls.ms_hook += o.ms_hook;
    ls.ms_round += o.ms_round;
}

// All datasets in one working order (1 <= d <= 8): one graph build, one
// labelling pass, one segmented sort_by_size.
static void fit_packed(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d, int32_t dtype,
                       const double *r2, const int32_t *c_sim, const int32_t *m_core, int32_t precision,
                       int64_t member_cutoff, int64_t max_clusters, int64_t *labels, int size_bits, int ds_bits,
                       cnn_graph_stats &gs, cnn_label_stats &ls) {
    hipStream_t s = ctx->stream;
    PackedHost ph;
    cnn_graph *g = graph_build_packed(ctx, X, offsets, n_sets, d, dtype, r2, precision, ph);
    std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
    gs = g->stats;
    const int64_t T = g->n_tiles;
    if (T == 0) return;
    std::vector<int32_t> h_c(T), h_m(T);
    for (int64_t t = 0; t < T; ++t) {
        h_c[t] = c_sim[ph.tile_ds[t]];
        h_m[t] = m_core[ph.tile_ds[t]];
    }
    Scratch<int32_t> d_c(ctx, T), d_m(ctx, T), d_ds(ctx, T);
    Scratch<int64_t> d_off(ctx, n_sets + 1);
    CNN_HIP(hipMemcpyAsync(d_c.p, h_c.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_m.p, h_m.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_ds.p, ph.tile_ds.data(), T * sizeof(int32_t), hipMemcpyHostToDevice, s));
    CNN_HIP(hipMemcpyAsync(d_off.p, offsets, (n_sets + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    PackedSort ps{n_sets, 2 * size_bits, size_bits, (uint32_t)(((uint64_t)1 << size_bits) - 1), ds_bits + 2 * size_bits,
                  d_ds.p, d_off.p};
    LabelRequest r = graph_label_request(ctx, g, CNN_SIM_AUTO);
    r.sort = true;
    r.write_all = true;
    r.member_cutoff = member_cutoff;
    r.max_clusters = max_clusters;
    r.labels = labels;
    r.tile_csim = d_c.p;
    r.tile_mcore = d_m.p;
    r.packed = &ps;
    labels_run(ctx, r, &ls);
    // the host vectors above must outlive the asynchronous copies
    CNN_HIP(hipStreamSynchronize(s));
}

// Runs item(wc, i) for i = 0 .. count-1 on up to kBatchWorkers worker contexts
// (own stream and arena each) from as many host threads, items taken in the
// given order from a shared counter.  Work already queued on the caller's
// stream is finished first; each item finishes its own work before the next
// one starts on that worker, so all of it is done on return.  The first
// exception of any item is rethrown after every thread has stopped.
template <typename F>
static void run_on_workers(cnn_ctx *ctx, int32_t count, const std::vector<int32_t> &order, F &&item) {
    if (count <= 0) return;
    CNN_HIP(hipStreamSynchronize(ctx->stream));
    const int n_workers = std::min<int>(kBatchWorkers, count);
    while ((int)ctx->workers.size() < n_workers) {
        cnn_ctx *w = nullptr;
        if (cnn_ctx_create(ctx->device, &w) != CNN_OK) throw HipFailure{hipErrorUnknown, "worker context creation failed"};
        ctx->workers.push_back(w);
    }
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    std::mutex mu;
    std::exception_ptr first_error;
    auto work = [&](cnn_ctx *wc) {
        try {
            CNN_HIP(hipSetDevice(wc->device));
            for (;;) {
                const int32_t k = next.fetch_add(1);
                if (k >= count || failed.load()) break;
                item(wc, order[k], mu);
                CNN_HIP(hipStreamSynchronize(wc->stream));
            }
        } catch (...) {
            failed.store(true);
            std::lock_guard<std::mutex> lock(mu);
            if (!first_error) first_error = std::current_exception();
        }
    };
    std::vector<std::thread> threads;
    for (int k = 1; k < n_workers; ++k) threads.emplace_back(work, ctx->workers[k]);
    work(ctx->workers[0]);
    for (auto &t : threads) t.join();
    if (first_error) std::rethrow_exception(first_error);
}

// Datasets fitted independently, several at once (largest first), so the
// device overlaps the kernels, copies and host round trips of different
// datasets.
static void fit_each(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d, int32_t dtype,
                     const double *r2, const int32_t *c_sim, const int32_t *m_core, int32_t precision,
                     int64_t member_cutoff, int64_t max_clusters, int64_t *labels, cnn_graph_stats &gs,
                     cnn_label_stats &ls) {
    std::vector<int32_t> order(std::max(n_sets, 0));
    for (int32_t b = 0; b < n_sets; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return offsets[x + 1] - offsets[x] > offsets[y + 1] - offsets[y];
    });
    const size_t esize = dtype == CNN_F64 ? sizeof(double) : sizeof(float);
    run_on_workers(ctx, n_sets, order, [&](cnn_ctx *wc, int32_t b, std::mutex &mu) {
        const int64_t n = offsets[b + 1] - offsets[b];
        const char *xb = (const char *)X + (size_t)offsets[b] * d * esize;
        cnn_graph *g = graph_build(wc, xb, n, d, dtype, r2[b], precision, nullptr, true);
        std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
        LabelRequest r = graph_label_request(wc, g, CNN_SIM_AUTO);
        r.c_sim = c_sim[b];
        r.m_core = m_core[b];
        r.sort = true;
        r.write_all = true;
        r.member_cutoff = member_cutoff;
        r.max_clusters = max_clusters;
        r.labels = labels + offsets[b];
        cnn_label_stats one{};
        labels_run(wc, r, &one);
        std::lock_guard<std::mutex> lock(mu);
        add_graph_stats(gs, g->stats);
        add_label_stats(ls, one);
    });
}

// ------------------------------------------------------------------ captured scans

// A parameter scan of a spatial graph that keeps its row words, recorded once
// into a hipGraph: per radius (descending) the words are re-thresholded exactly
// (k_word_filter, from the previous radius's words), then one launch per
// labelling kernel covers every cutoff pair of that radius (launch y = pair
// y).  Every buffer is allocated before the capture and no step reads a value
// back to the host, so the graph replays the whole scan with one launch.
struct cnn_scan {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {};
    std::vector<void *> bufs;
    int32_t n_r = 0, n_c = 0;
    int64_t n = 0;
    unsigned long long *counters = nullptr;  // n_r * n_c * 8
    unsigned long long *tstats = nullptr;    // n_r * n_c * 4 * 1024
};

namespace {

constexpr int kScanStatSlots = 4 * 1024;

void scan_free(cnn_scan *sc) {
    if (!sc) return;
    (void)hipSetDevice(sc->device);
    if (sc->exec) (void)hipGraphExecDestroy(sc->exec);
    if (sc->graph) (void)hipGraphDestroy(sc->graph);
    for (auto &e : sc->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *p : sc->bufs) (void)hipFree(p);
    delete sc;
}

template <typename T>
T *scan_alloc(cnn_scan *sc, int64_t count) {
    void *p = nullptr;
    const size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        throw OutOfMemory{"scan buffer of " + std::to_string(bytes) + " bytes"};
    }
    sc->bufs.push_back(p);
    return (T *)p;
}

cnn_scan *scan_create(cnn_ctx *ctx, const cnn_graph *g, const double *r2, int32_t n_r, const int32_t *c_sim,
                      const int32_t *m_core, int32_t n_c, int64_t member_cutoff, int64_t max_clusters,
                      int64_t *labels) {
    if (!g->mask || !g->pair_nz || !g->tp_ptr)
        throw Capacity{"captured scans need a graph that keeps its row words"};
    for (int32_t i = 0; i < n_r; ++i)
        if (!(r2[i] <= g->r2)) throw InvalidArg{"scan radii must not exceed the graph radius"};
    for (int32_t ci = 0; ci < n_c; ++ci)
        if (c_sim[ci] < 0 || m_core[ci] < 0) throw InvalidArg{"c_sim and m_core must be >= 0"};
    std::unique_ptr<cnn_scan, void (*)(cnn_scan *)> sc(new cnn_scan(), scan_free);
    sc->device = ctx->device;
    sc->n_r = n_r;
    sc->n_c = n_c;
    const int64_t n = g->n, T = g->n_tiles, P = g->n_tile_pairs;
    sc->n = n;
    for (auto &e : sc->ev) CNN_HIP(hipEventCreate(&e));
    std::vector<int32_t> order(n_r);
    for (int32_t i = 0; i < n_r; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return r2[x] > r2[y]; });
    // buffers: words of two radii (rotating), degrees, the labelling slices of n_c pairs
    bool smaller = false;
    for (int32_t i = 0; i < n_r; ++i) smaller = smaller || r2[i] < g->r2;
    uint64_t *wmask[2] = {nullptr, nullptr};
    uint8_t *wnz[2] = {nullptr, nullptr};
    int32_t *wdeg[2] = {nullptr, nullptr};
    if (smaller)
        for (int k = 0; k < 2; ++k) {
            wmask[k] = scan_alloc<uint64_t>(sc.get(), P * 64);
            wnz[k] = scan_alloc<uint8_t>(sc.get(), P);
            wdeg[k] = scan_alloc<int32_t>(sc.get(), n);
        }
    int32_t *own_deg = scan_alloc<int32_t>(sc.get(), n);
    MultiLabelSizes sz;
    multi_labels_bytes(n, T, n_c, sz);
    MultiLabelBufs b;
    b.n_c = n_c;
    b.n = n;
    b.n_tiles = T;
    const int64_t m = sz.per_point, mt = sz.per_tile;
    b.core = scan_alloc<uint8_t>(sc.get(), m);
    b.parent = scan_alloc<int>(sc.get(), m);
    b.rank = scan_alloc<uint64_t>(sc.get(), m);
    b.cnt = scan_alloc<int32_t>(sc.get(), m);
    b.hq = scan_alloc<int32_t>(sc.get(), m);
    b.hq_len = scan_alloc<unsigned int>(sc.get(), n_c);
    b.tmax = scan_alloc<uint64_t>(sc.get(), mt);
    b.cand_max = scan_alloc<uint64_t>(sc.get(), mt);
    b.tile_common = scan_alloc<uint64_t>(sc.get(), mt);
    b.cmin = scan_alloc<int32_t>(sc.get(), m);
    b.csize = scan_alloc<int32_t>(sc.get(), m);
    b.newlab = scan_alloc<int32_t>(sc.get(), m);
    b.ckey = scan_alloc<uint64_t>(sc.get(), m);
    b.cidx = scan_alloc<int32_t>(sc.get(), m);
    b.ccount = scan_alloc<unsigned int>(sc.get(), n_c);
    int32_t *d_csim = scan_alloc<int32_t>(sc.get(), n_c);
    int32_t *d_mcore = scan_alloc<int32_t>(sc.get(), n_c);
    CNN_HIP(hipMemcpy(d_csim, c_sim, n_c * sizeof(int32_t), hipMemcpyHostToDevice));
    CNN_HIP(hipMemcpy(d_mcore, m_core, n_c * sizeof(int32_t), hipMemcpyHostToDevice));
    b.csim = d_csim;
    b.mcore = d_mcore;
    sc->counters = scan_alloc<unsigned long long>(sc.get(), (int64_t)n_r * n_c * 8);
    sc->tstats = scan_alloc<unsigned long long>(sc.get(), (int64_t)n_r * n_c * kScanStatSlots);

    // record the scan on a private stream
    hipStream_t cs = nullptr;
    CNN_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipStream_t saved = ctx->stream;
    ctx->stream = cs;
    bool capturing = false;
    try {
        CNN_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        capturing = true;
        const uint64_t *cur_mask = g->mask;
        const uint8_t *cur_nz = g->pair_nz;
        const int32_t *cur_deg = own_deg;
        words_degrees(ctx, g, own_deg);
        double cur_r2 = g->r2;
        int next = 0;
        for (int32_t oi = 0; oi < n_r; ++oi) {
            const int32_t ri = order[oi];
            if (r2[ri] < cur_r2) {
                words_filter(ctx, g, cur_mask, cur_nz, r2[ri], wmask[next], wnz[next], wdeg[next]);
                cur_mask = wmask[next];
                cur_nz = wnz[next];
                cur_deg = wdeg[next];
                cur_r2 = r2[ri];
                next ^= 1;
            }
            MultiLabelSource src{cur_mask, cur_nz, g->tp_ptr, g->tp_col, cur_deg, g->perm, g->n_tile_pairs};
            labels_enqueue_multi(ctx, b, src, g->tune.root_passes, member_cutoff, max_clusters,
                                 labels + (int64_t)ri * n_c * n, sc->counters + (int64_t)ri * n_c * 8,
                                 sc->tstats + (int64_t)ri * n_c * kScanStatSlots);
        }
        capturing = false;
        CNN_HIP(hipStreamEndCapture(cs, &sc->graph));
        CNN_HIP(hipGraphInstantiate(&sc->exec, sc->graph, nullptr, nullptr, 0));
    } catch (...) {
        if (capturing) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(cs, &partial);
            if (partial) (void)hipGraphDestroy(partial);
        }
        ctx->stream = saved;
        (void)hipStreamDestroy(cs);
        (void)hipGetLastError();
        throw;
    }
    ctx->stream = saved;
    CNN_HIP(hipStreamDestroy(cs));
    return sc.release();
}

void scan_launch(cnn_ctx *ctx, cnn_scan *sc, cnn_label_stats *lstats, float *ms_out) {
    hipStream_t s = ctx->stream;
    CNN_HIP(hipEventRecord(sc->ev[0], s));
    CNN_HIP(hipGraphLaunch(sc->exec, s));
    CNN_HIP(hipEventRecord(sc->ev[1], s));
    CNN_HIP(hipEventSynchronize(sc->ev[1]));
    float ms = 0.0f;
    CNN_HIP(hipEventElapsedTime(&ms, sc->ev[0], sc->ev[1]));
    if (ms_out) *ms_out = ms;
    if (!lstats) return;
    const int64_t pairs = (int64_t)sc->n_r * sc->n_c;
    std::vector<unsigned long long> hc(pairs * 8), ht(pairs * kScanStatSlots);
    CNN_HIP(hipMemcpy(hc.data(), sc->counters, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    CNN_HIP(hipMemcpy(ht.data(), sc->tstats, ht.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < pairs; ++p) {
        cnn_label_stats st{};
        const unsigned long long *c = hc.data() + p * 8, *t = ht.data() + p * kScanStatSlots;
        for (int i = 0; i < kScanStatSlots / 4; ++i) {
            st.tests += (int64_t)t[4 * i];
            st.tests_passed += (int64_t)t[4 * i + 1];
            st.rows_scanned += (int64_t)t[4 * i + 2];
        }
        st.n_core = (int64_t)c[2];
        st.n_components = (int64_t)c[3];
        st.n_clusters = (int64_t)c[7];
        st.ms_similarity = ms / (float)pairs;  // the graph's time, shared evenly
        st.sim_form = CNN_SIM_FORM_TILE_WORDS;
        lstats[p] = st;
    }
}

}  // namespace

extern "C" {

int cnn_graph_scan_create(cnn_ctx *ctx, const cnn_graph *g, const double *r2, int32_t n_r, const int32_t *c_sim,
                          const int32_t *m_core, int32_t n_c, int64_t member_cutoff, int64_t max_clusters,
                          int64_t *labels, cnn_scan **out) {
    if (!ctx || !g || !out || n_r <= 0 || n_c <= 0) return CNN_ERR_INVALID;
    *out = nullptr;
    return guarded(ctx, [&]() {
        if (!r2 || !c_sim || !m_core) throw InvalidArg{"null parameter array"};
        if (g->n > 0 && !labels) throw InvalidArg{"labels is null"};
        CNN_HIP(hipSetDevice(ctx->device));
        *out = scan_create(ctx, g, r2, n_r, c_sim, m_core, n_c, member_cutoff, max_clusters, labels);
        return (int)CNN_OK;
    });
}

int cnn_scan_launch(cnn_ctx *ctx, cnn_scan *scan, cnn_label_stats *lstats) {
    if (!ctx || !scan) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        CNN_HIP(hipSetDevice(ctx->device));
        scan_launch(ctx, scan, lstats, nullptr);
        return (int)CNN_OK;
    });
}

int cnn_scan_destroy(cnn_scan *scan) {
    if (!scan) return CNN_ERR_INVALID;
    scan_free(scan);
    return CNN_OK;
}

int cnn_fit_batch(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d, int32_t dtype,
                  const double *r2, const int32_t *c_sim, const int32_t *m_core, int32_t precision,
                  int64_t member_cutoff, int64_t max_clusters, int64_t *labels, cnn_graph_stats *gstats,
                  cnn_label_stats *lstats) {
    if (!ctx || !offsets || n_sets < 0) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n_sets > 0 && (!r2 || !c_sim || !m_core)) throw InvalidArg{"null parameter array"};
        if (offsets[0] < 0) throw InvalidArg{"offsets must start at a non-negative row"};
        int64_t n_max = 0;
        for (int32_t b = 0; b < n_sets; ++b) {
            if (offsets[b + 1] < offsets[b]) throw InvalidArg{"offsets must be non-decreasing"};
            if (c_sim[b] < 0 || m_core[b] < 0) throw InvalidArg{"c_sim and m_core must be >= 0"};
            n_max = std::max(n_max, offsets[b + 1] - offsets[b]);
        }
        if (n_sets > 0 && !labels && offsets[n_sets] > offsets[0]) throw InvalidArg{"labels is null"};
        CNN_HIP(hipSetDevice(ctx->device));
        cnn_graph_stats gs{};
        cnn_label_stats ls{};
        const int size_bits = bits_for(n_max);
        const int ds_bits = bits_for(std::max(n_sets - 1, 0));
        const bool packed = kPackedBatches && n_sets > 0 && d >= 1 && d <= kSpatialMaxD &&
                            ds_bits + 2 * size_bits <= 64 && offsets[0] == 0;
        if (packed) {
            fit_packed(ctx, X, offsets, n_sets, d, dtype, r2, c_sim, m_core, precision, member_cutoff,
                       max_clusters, labels, size_bits, ds_bits, gs, ls);
        } else {
            fit_each(ctx, X, offsets, n_sets, d, dtype, r2, c_sim, m_core, precision, member_cutoff, max_clusters,
                     labels, gs, ls);
        }
        if (gstats) *gstats = gs;
        if (lstats) *lstats = ls;
        return (int)CNN_OK;
    });
}

// Parameter scan on one dataset: the graph is built once at the largest
// radius; each smaller radius filters the previous (larger) lists exactly in
// FP64, and every cutoff pair runs the labelling pass on the lists of its
// radius.  Radii are visited in descending order; results land at the
// caller's (radius, cutoff) positions.
int cnn_fit_scan(cnn_ctx *ctx, const void *X, int64_t n, int32_t d, int32_t dtype, const double *r2, int32_t n_r,
                 const int32_t *c_sim, const int32_t *m_core, int32_t n_c, int32_t precision, int64_t member_cutoff,
                 int64_t max_clusters, int64_t *labels, cnn_graph_stats *gstats, cnn_label_stats *lstats) {
    if (!ctx || n_r < 0 || n_c < 0) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n_r == 0 || n_c == 0) return (int)CNN_OK;
        if (!r2 || !c_sim || !m_core) throw InvalidArg{"null parameter array"};
        if (n > 0 && !labels) throw InvalidArg{"labels is null"};
        for (int32_t ci = 0; ci < n_c; ++ci)
            if (c_sim[ci] < 0 || m_core[ci] < 0) throw InvalidArg{"c_sim and m_core must be >= 0"};
        std::vector<int32_t> order(n_r);
        for (int32_t i = 0; i < n_r; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return r2[x] > r2[y]; });
        CNN_HIP(hipSetDevice(ctx->device));
        cnn_graph *g = graph_build(ctx, X, n, d, dtype, r2[order[0]], precision, nullptr, true);
        std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
        if (gstats) *gstats = g->stats;
        if (g->mask && n > 0) {
            // the graph keeps its row words: the whole scan as one captured graph
            std::unique_ptr<cnn_scan, int (*)(cnn_scan *)> sc(
                scan_create(ctx, g, r2, n_r, c_sim, m_core, n_c, member_cutoff, max_clusters, labels),
                cnn_scan_destroy);
            scan_launch(ctx, sc.get(), lstats, nullptr);
            return (int)CNN_OK;
        }
        // smaller radii filter the lists of the largest
        bool smaller = false;
        for (int32_t i = 0; i < n_r; ++i) smaller = smaller || r2[i] < r2[order[0]];
        if (smaller && n > 0) graph_lists(ctx, g);
        if (gstats) *gstats = g->stats;
        // lists of the radius being labelled; the graph's own for the largest
        const int64_t *rp = g->row_ptr;
        const int32_t *col = g->col;
        int64_t *own_rp = nullptr;
        int32_t *own_col = nullptr;
        auto release = [&]() {
            if (own_rp) (void)hipFree(own_rp);
            if (own_col) (void)hipFree(own_col);
            own_rp = nullptr;
            own_col = nullptr;
        };
        try {
            double cur_r2 = r2[order[0]];
            for (int32_t oi = 0; oi < n_r; ++oi) {
                const int32_t ri = order[oi];
                if (r2[ri] < cur_r2) {
                    int64_t *nrp = nullptr;
                    int32_t *ncol = nullptr;
                    if (hipMalloc(&nrp, (size_t)(n + 1) * sizeof(int64_t)) != hipSuccess) {
                        (void)hipGetLastError();
                        throw OutOfMemory{"row offsets for the scan"};
                    }
                    try {
                        csr_filter(ctx, g->xw64, n, d, rp, col, r2[ri], nrp, &ncol);
                    } catch (...) {
                        (void)hipFree(nrp);
                        throw;
                    }
                    release();
                    own_rp = nrp;
                    own_col = ncol;
                    rp = own_rp;
                    col = own_col;
                    cur_r2 = r2[ri];
                }
                // the cutoffs of one radius read the same lists: labelled concurrently
                std::vector<int32_t> cis(n_c);
                for (int32_t ci = 0; ci < n_c; ++ci) cis[ci] = ci;
                run_on_workers(ctx, n_c, cis, [&](cnn_ctx *wc, int32_t ci, std::mutex &) {
                    // self positions and row words describe the graph's own lists, not lists
                    // filtered to a smaller r
                    const bool own = rp == g->row_ptr;
                    LabelRequest req = graph_label_request(wc, g, CNN_SIM_AUTO);
                    if (!own) {
                        req.row_ptr = rp;
                        req.col = col;
                        req.selfpos = nullptr;
                        req.mask = nullptr;
                        req.pair_nz = nullptr;
                    }
                    req.c_sim = c_sim[ci];
                    req.m_core = m_core[ci];
                    req.sort = true;
                    req.write_all = true;
                    req.member_cutoff = member_cutoff;
                    req.max_clusters = max_clusters;
                    req.labels = labels + ((int64_t)ri * n_c + ci) * n;
                    cnn_label_stats one{};
                    labels_run(wc, req, &one);
                    if (lstats) lstats[(int64_t)ri * n_c + ci] = one;
                });
            }
        } catch (...) {
            release();
            throw;
        }
        release();
        return (int)CNN_OK;
    });
}

}  // extern "C"
