// mcpm_plane.h — plane-field hooks of the context used by the field
// surrogate (mcpm_fno.hip); not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "mcpm_kernels.h"

struct PlaneField {
  double *n_node;  // (nx+1)(ny+1) nodal density
  double *pot;     // (nx+1)(ny+1) nodal field
  mcpm::FemParams fem;
  hipStream_t s;   // the context's stream
};

// Deposit the nodal density of a screened-Poisson plane context (the first
// part of compute_SP_2D, Main_Functions.cpp:2268-2364) and describe its field.
int ctx_plane_nodes(mcpm_ctx *c, PlaneField *pf);

// Conjugate gradients on the Q1 system (K + lambda^2 M) pot = M n.
// warm = 0 starts from zero, warm = 1 from the current pot.  mode 0 solves
// to the context's tolerance; mode 1 stops at once when the starting
// residual |M n - A pot| / |M n| is at most accept; mode 2 only measures it.
// res0 gets the starting residual when warm (else -1), iters the iterations.
int ctx_plane_cg(mcpm_ctx *c, int warm, int mode, double accept, double *res0, int *iters);

// Core subtraction (FEM.cpp:497-554) and the kick of the next push.
int ctx_plane_finish(mcpm_ctx *c);

// Statistics of the last surrogate field, kept by the context.
int ctx_set_surrogate_stats(mcpm_ctx *c, const mcpm_fno_field_stats *st);

// Field of one step through the surrogate attached by mcpm_set_field_surrogate.
int fno_field_step(mcpm_fno *f, mcpm_ctx *c, int64_t step, double accept, int64_t every);
