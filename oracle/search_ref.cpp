// search_ref.cpp -- C-ABI driver around the reference's own compiled graph
// search, used to pin the oracle's query-time search loop (ro_search in
// rabitq_oracle.c). TEST INFRASTRUCTURE ONLY: tests/ load it from
// oracle/_ref; the product never links it. Nothing of the reference is
// copied here; this file only calls its functions.
//
// Two entry points, both over a space made by the reference bridge
// (rabitq_space_new, compiled in through REF_BRIDGE):
//   * sr_index_*: a VamanaIndex loaded from a VMNARBQ1 file with
//     setPaperPrune(true, eps0) and the external residual file opened,
//     answering searchKnnPlainThenResidualRerank
//     (Ours/core/hnswlib/vamana_index.h:180, 530-636): top-k ids,
//     reranked estimates and the search metrics.
//   * sr_pool: the traversal itself, vamana::search_vamana_paper_active
//     (Ours/core/hnswlib/vamana_search.h:216-364), driven with an evaluator
//     that makes the same space calls as VamanaIndex's QueryEvaluator
//     (vamana_index.h:78-133; that class is private to the index): the
//     final pool (ids, distances) and the SearchStats counters.
#include REF_BRIDGE

#include <cstdio>
#include <string>
#include <vector>

namespace {

struct SrIndex {
    hnswlib::RaBitQSpace *space;
    hnswlib::VamanaIndex *index;
};

// The space calls of QueryEvaluator: full distance from the compact record,
// the 1-bit gate from the sidecar, the remaining-bit distance after it.
struct SidecarEvaluator {
    hnswlib::RaBitQSpace &space;
    const uint8_t *compact;
    size_t compact_stride;
    const uint8_t *msb;
    size_t msb_stride;
    const hnswlib::PaperPruneFactors<float> *factors;
    const void *prepared;

    float distance(hnswlib::vamana::NodeId id) {
        return space.query_distance(prepared, compact + static_cast<size_t>(id) * compact_stride);
    }
    hnswlib::PaperPruneEstimate<float> paper_estimate(hnswlib::vamana::NodeId id, float epsilon0) {
        return space.compute_paper_prune_estimate_sidecar(prepared, msb + static_cast<size_t>(id) * msb_stride,
                                                          factors[id], epsilon0);
    }
    float distance_with_paper(hnswlib::vamana::NodeId id, float short_ip) {
        return space.query_distance_with_paper_msb(prepared, compact + static_cast<size_t>(id) * compact_stride,
                                                   short_ip);
    }
    void prefetch(hnswlib::vamana::NodeId) const {}
};

}  // namespace

extern "C" {

// Counters of vamana::SearchStats / RaBitQSearchMetrics, in this order.
struct SrStats {
    uint64_t visited_nodes, distance_computations, hops, paper_checked, paper_would_prune, paper_not_pruned,
        paper_msb_kernel_calls, paper_remaining_kernel_calls;
};

// VamanaIndex over the space `space` (not owned) loaded from index_path, its
// external residual records (n rows) opened from residual_path, paper prune
// on with eps0. nullptr when the reference rejects any step.
void *sr_index_open(void *space, const char *index_path, const char *residual_path, size_t n, float eps0) {
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        auto *index = new hnswlib::VamanaIndex(typed, 0);
        index->loadIndex(index_path);
        typed->openExternalResidualStorage(residual_path, n);
        index->setPaperPrune(true, eps0);
        return new SrIndex{typed, index};
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sr_index_open: %s\n", e.what());
        return nullptr;
    }
}

void sr_index_close(void *h) {
    auto *s = static_cast<SrIndex *>(h);
    if (!s) return;
    try {
        s->space->closeExternalResidualStorage();
    } catch (...) {
    }
    delete s->index;
    delete s;
}

// searchKnnPlainThenResidualRerank for one query: up to k (estimate, label)
// pairs in ascending (estimate, label) order; *count receives how many.
int sr_index_knn(void *h, const float *query, size_t k, size_t l_search, size_t rerank, uint64_t *labels_out,
                 float *d_out, size_t *count, SrStats *st) {
    auto *s = static_cast<SrIndex *>(h);
    try {
        hnswlib::RaBitQSearchMetrics m;
        auto result = s->index->searchKnnPlainThenResidualRerank(query, k, l_search, rerank, &m, true);
        const size_t got = result.size();
        for (size_t i = got; i-- > 0;) {  // the max-heap pops in descending order
            d_out[i] = result.top().first;
            labels_out[i] = static_cast<uint64_t>(result.top().second);
            result.pop();
        }
        *count = got;
        if (st) {
            st->visited_nodes = m.visited_nodes;
            st->distance_computations = m.distance_computations;
            st->hops = m.hops;
            st->paper_checked = m.paper_checked;
            st->paper_would_prune = m.paper_would_prune;
            st->paper_not_pruned = m.paper_not_pruned;
            st->paper_msb_kernel_calls = m.paper_msb_kernel_calls;
            st->paper_remaining_kernel_calls = m.paper_remaining_kernel_calls;
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sr_index_knn: %s\n", e.what());
        return -1;
    }
}

// search_vamana_paper_active from start node 0 with beam width 1 over the
// graph (degrees[n], edges[n][edge_stride]); the final pool (at most L
// entries) in pool order.
int sr_pool(void *space, const float *query, const uint8_t *compact, size_t compact_stride, const uint8_t *msb,
            size_t msb_stride, const void *factors, const uint32_t *degrees, const uint32_t *edges,
            size_t edge_stride, size_t n, size_t l_value, float eps0, uint32_t *pool_ids, float *pool_d,
            size_t *pool_len, SrStats *st) {
    try {
        auto *typed = static_cast<hnswlib::RaBitQSpace *>(space);
        std::vector<std::vector<hnswlib::vamana::NodeId>> graph(n);
        for (size_t i = 0; i < n; ++i)
            graph[i].assign(edges + i * edge_stride, edges + i * edge_stride + degrees[i]);
        const void *prepared = typed->prepare_query(query);
        SidecarEvaluator ev{*typed, compact, compact_stride, msb, msb_stride,
                            static_cast<const hnswlib::PaperPruneFactors<float> *>(factors), prepared};
        hnswlib::vamana::SearchStats stats;
        hnswlib::vamana::SearchScratch scratch;
        const std::vector<hnswlib::vamana::SearchCandidate> &pool = hnswlib::vamana::search_vamana_paper_active(
            graph, 0, n, l_value, 1, ev, eps0, 0, scratch, &stats);
        for (size_t i = 0; i < pool.size(); ++i) {
            pool_ids[i] = pool[i].id;
            pool_d[i] = pool[i].distance;
        }
        *pool_len = pool.size();
        typed->release_query(prepared);
        if (st) {
            st->visited_nodes = stats.visited_nodes;
            st->distance_computations = stats.distance_computations;
            st->hops = stats.hops;
            st->paper_checked = stats.paper_checked;
            st->paper_would_prune = stats.paper_would_prune;
            st->paper_not_pruned = stats.paper_not_pruned;
            st->paper_msb_kernel_calls = stats.paper_msb_kernel_calls;
            st->paper_remaining_kernel_calls = stats.paper_remaining_kernel_calls;
        }
        return 0;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sr_pool: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"
