// igc_internal.hpp -- declarations shared by the translation units of libigc.so.
#pragma once
#include "../../include/igc.h"

#include <cstddef>
#include <cstdint>

namespace igc {

// Records a message for igc_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Reorders every row of a host CSR from ascending to the neighbour order of
// the reference's T-thread FillAdjList merge (src/SparseMat.cpp:599-609):
// within row x, neighbour y is keyed by (min(x,y) mod T, min(x,y), max(x,y)).
void reorder_threads(igc_csr &csr, uint32_t T);

// Host memory for arrays handed to the caller (released with host_free by
// igc_csr_free).  Blocks of 4 MiB and more are 2 MiB aligned, advised for
// transparent huge pages, and recycled through a bounded cache of blocks whose
// pages are already faulted in; host_cache_trim() frees the cached blocks.
void *host_alloc(size_t bytes);
void host_free(void *p);
void host_cache_trim();

} // namespace igc
