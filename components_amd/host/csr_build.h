// Host CSR construction shared by the METIS reader and the generators.
#pragma once

#include <cstdint>
#include <vector>

#include "kacomp.h"

namespace kc {

// Undirected edge list -> symmetric CSR without self-loops or duplicate edges;
// every row sorted ascending.  Runs on all OpenMP threads.
void build_symmetric_csr(uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t count, kc_csr *out);

// Allocates out->offsets / out->targets with malloc (freed by kc_csr_free).
bool alloc_csr(kc_csr *out, uint64_t n, uint64_t m2);

// Rows [a, z) of a CSR whose targets are ids relative to a (all < z - a).
// Returns true and one edge u -> v (relative ids) whose reverse v -> u is
// missing, or false when every edge appears in both rows.  Sorted copies of
// the rows, then a binary search per edge; parallel over rows when the range
// is large.
bool find_one_way_edge(const uint64_t *offsets, const uint32_t *targets, uint64_t a, uint64_t z, uint64_t &u,
                       uint64_t &v);
// The same over the graphs of a packed batch (graph g owns the rows
// [vtx_off[g], vtx_off[g+1]) and local targets), parallel over the graphs;
// reports the first graph with a one-way edge.
bool find_one_way_edge_batch(uint32_t graphs, const uint64_t *vtx_off, const uint64_t *offsets,
                             const uint32_t *targets, uint32_t &graph, uint64_t &u, uint64_t &v);

}  // namespace kc
