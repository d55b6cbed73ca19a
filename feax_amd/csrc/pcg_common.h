// Internal interface between the PCG driver (krylov.hip) and the
// matrix-free HEX8 operator (hex8.hip).
#pragma once

#include <cstdint>

#include "hex8_common.h"

namespace fx {

// Device-resident PCG scalars (one per solve).
struct PcgState {
  double alpha, beta, gamma, rs, atol2, bb;
  int k, done, maxiter, nparts;
  double g[2];  // gamma of even / odd iterations (CSR iterations read g[k % 2])
};

// J_bc(u) v of the HEX8 mesh (fx_hex8_jvp_f64) as a PCG operator.
struct Hex8Op {
  int num_nodes;
  const double *x, *y, *z;
  const int *cells, *inc_ptr, *inc;
  MatParams mat;
  const double* param;
  const double* u;        // linearisation point (Neo-Hookean), else nullptr
  const uint8_t* bc_mask;
};

// Blocks of one operator launch (= partial sums it writes when st != nullptr).
int hex8_op_blocks(int num_nodes);

// out = J_bc v.  With st: skip when st->done, and write the block partials
// of v . out to partials[0 .. hex8_op_blocks).
void launch_hex8_op(hipStream_t s, const Hex8Op& op, const double* v, double* out, const PcgState* st,
                    double* partials);

}  // namespace fx
