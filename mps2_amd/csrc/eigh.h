// Batched symmetric eigen-decomposition of Gram matrices on the device
// (Householder tridiagonalisation, tridiagonal divide and conquer, blocked
// back-transformation).  Used by the sector SVD.
#pragma once

#include <memory>
#include <vector>

#include "runtime.h"

namespace mps2 {

struct EighState;

struct EighResult {
  std::vector<std::vector<double>> evals;  // per sector, ascending
  std::shared_ptr<EighState> st;           // device data for eigh_vectors
};

// G holds the matrices packed back to back (sector q: n_q x n_q row-major at
// goffs[q]); it is overwritten.
EighResult eigh_batched(Ctx *c, double *G, const std::vector<int64_t> &goffs, const std::vector<int> &ns);

// For every sector q: E + eoff[q] (n_q x cols[q].size(), row-major) = the
// eigenvectors whose indices (into the ascending order) are cols[q], in that
// order.  Sectors with no columns are skipped.
void eigh_vectors_all(Ctx *c, const EighResult &r, const std::vector<std::vector<int32_t>> &cols, double *E,
                      const std::vector<int64_t> &eoff);

}  // namespace mps2
