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

// The Q1 system (K + lambda^2 M) pot = M n.  warm = 1: measure the starting
// residual |M n - A pot| / |M n| of the current pot into res0 (else -1);
// mode 2 stops there, mode 1 stops there when it is at most accept.  Else
// solve: by FFT on periodic power-of-two planes, otherwise by conjugate
// gradients to the context's tolerance (from the current pot when warm);
// iters gets the CG iterations, solved 1 when a solve ran.
int ctx_plane_cg(mcpm_ctx *c, int warm, int mode, double accept, double *res0, int *iters,
                 int *solved);

// Core subtraction (FEM.cpp:497-554) and the kick of the next push.
int ctx_plane_finish(mcpm_ctx *c);

// Statistics of the last surrogate field, kept by the context.
int ctx_set_surrogate_stats(mcpm_ctx *c, const mcpm_fno_field_stats *st);

// Field of one step through the surrogate attached by mcpm_set_field_surrogate.
int fno_field_step(mcpm_fno *f, mcpm_ctx *c, int64_t step, double accept, int64_t every);
