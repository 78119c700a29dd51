// mcpm_internal.h — context accessors shared by the library's host sources
// (not part of the C ABI).
#pragma once
#include <stdint.h>

#include "mcpm.h"

const mcpm_params *ctx_params(const mcpm_ctx *c);
// cell_poly_velocity calls since the last mcpm_reset_samples
int64_t ctx_sample_count(const mcpm_ctx *c);
