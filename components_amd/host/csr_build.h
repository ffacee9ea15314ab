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

}  // namespace kc
