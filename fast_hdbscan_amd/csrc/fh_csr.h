// Boruvka on a CSR graph (fh_csr.hip).
#pragma once

#include <stdint.h>

namespace fh {

// core_graph.boruvka_mst on the given device (HOST arrays in and out).
int csr_boruvka(int device, int64_t n, const int32_t *indptr, const int32_t *indices, const float *weights,
                double *edges, int64_t *n_edges, int64_t *point_components, int64_t *n_components);

}  // namespace fh
