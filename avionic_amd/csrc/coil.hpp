// Coil-sensitivity estimation (coil.hip): the steps of CoilConstruction
// (CUDA/src/coil_construction.cpp) on device images [coils][height][width].
#pragma once

#include <vector>

#include "../../include/avionic.h"
#include "solver.hpp"

namespace avionic {

struct CoilCtx {
  int W, H, C;
  long N;  // W * H
  hipStream_t s;
};

// Cartesian 2D+t or non-Cartesian operators; kdata is the operator's dense
// k-space [frames][coils][...].
void coil_time_averaged(const Operator &op, const float2 *kdata, float2 *u, float2 *crec,
                        bool apply_phase, hipStream_t s);
// The all-frames, one-coil gridding operator of a non-Cartesian operator
// (noncartesian.cpp); nullptr for any other operator.
Operator *make_noncartesian_all_frames(const Operator &op);
void coil_b1_from_uh1(const CoilCtx &k, const float2 *u, const float2 *crec, double mu, float2 *b1);
int coil_find_max_sum(const CoilCtx &k, const float2 *crec);
void coil_utgv2(const CoilCtx &k, const float2 *b1, const float2 *crec, const std::vector<int> &inds,
                const avionic_coil_params &p, float2 *x_inout);
void coil_b1_recon(const CoilCtx &k, const float2 *u0, const float2 *crec, int coils,
                   unsigned maxit, double mu, const avionic_coil_params &p, float2 *x1);
void coil_construction_from(const CoilCtx &k, const float2 *crec, const float2 *u0c,
                            const avionic_coil_params &p, float2 *u, float2 *b1,
                            std::vector<int> *order);
void coil_default_params(avionic_coil_params *p);

}  // namespace avionic
