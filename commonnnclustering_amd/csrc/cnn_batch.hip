// Batched fits over many independent datasets (cnn_fit_batch).
//
// The reference fits one dataset per Clustering.fit call (cluster.pyx:256-392);
// parameter scans loop over such calls (HierarchicalFitterRepeat._fit,
// _fit.pyx:1058-1187).  Here each dataset is a row range of one device buffer
// and is fitted with its own radius and cutoffs.
#include <algorithm>
#include <memory>

#include "cnn_internal.hpp"

using namespace cnn;

extern "C" {

int cnn_fit_batch(cnn_ctx *ctx, const void *X, const int64_t *offsets, int32_t n_sets, int32_t d, int32_t dtype,
                  const double *r2, const int32_t *c_sim, const int32_t *m_core, int32_t precision,
                  int64_t member_cutoff, int64_t max_clusters, int64_t *labels, cnn_graph_stats *gstats,
                  cnn_label_stats *lstats) {
    if (!ctx || !offsets || n_sets < 0) return CNN_ERR_INVALID;
    return guarded(ctx, [&]() {
        if (n_sets > 0 && (!r2 || !c_sim || !m_core)) throw InvalidArg{"null parameter array"};
        for (int32_t b = 0; b < n_sets; ++b)
            if (offsets[b + 1] < offsets[b] || offsets[b] < 0) throw InvalidArg{"offsets must be non-decreasing"};
        CNN_HIP(hipSetDevice(ctx->device));
        const size_t esize = dtype == CNN_F64 ? sizeof(double) : sizeof(float);
        cnn_graph_stats gs{};
        cnn_label_stats ls{};
        for (int32_t b = 0; b < n_sets; ++b) {
            const int64_t n = offsets[b + 1] - offsets[b];
            const char *xb = (const char *)X + (size_t)offsets[b] * d * esize;
            cnn_graph *g = graph_build(ctx, xb, n, d, dtype, r2[b], precision);
            std::unique_ptr<cnn_graph, void (*)(cnn_graph *)> hold(g, graph_free);
            LabelRequest r{g->row_ptr, g->col, n, g->perm, c_sim[b], m_core[b], nullptr, 1, CNN_SIM_AUTO,
                           true, true, member_cutoff, max_clusters, labels + offsets[b],
                           g->tp_ptr, g->tp_col, g->n_tiles};
            cnn_label_stats one{};
            labels_run(ctx, r, &one);
            gs.pairs_evaluated += g->stats.pairs_evaluated;
            gs.band_hits += g->stats.band_hits;
            gs.band_flips += g->stats.band_flips;
            gs.max_abs_err_band = std::max(gs.max_abs_err_band, g->stats.max_abs_err_band);
            gs.tile_pairs += g->stats.tile_pairs;
            gs.nnz += g->stats.nnz;
            gs.ms_prep += g->stats.ms_prep;
            gs.ms_distance += g->stats.ms_distance;
            gs.ms_emit += g->stats.ms_emit;
            gs.distance_launches += g->stats.distance_launches;
            ls.n_core += one.n_core;
            ls.n_components += one.n_components;
            ls.n_clusters += one.n_clusters;
            ls.tests += one.tests;
            ls.tests_passed += one.tests_passed;
            ls.ms_similarity += one.ms_similarity;
            ls.ms_labels += one.ms_labels;
        }
        if (gstats) *gstats = gs;
        if (lstats) *lstats = ls;
        return (int)CNN_OK;
    });
}

}  // extern "C"
