// mcpm_device.h — device-side building blocks shared by the slab kernels:
// particle record layout, cell binning, the contact value Y(n), launch
// parameter block and wavefront helpers.  gfx950 only (64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mcpm.h"
#include "mcpm_rng.h"

#pragma clang fp contract(off)

namespace mcpm {

// One particle: slab coordinate y (the reference's x2) and the three
// velocity components U1, U2, U3.  Packed so that a random gather of one
// particle (the collision partner) touches one 16-byte (FP32) or 32-byte
// (FP64) span instead of four separate arrays.
template <typename real>
struct alignas(4 * sizeof(real)) Rec {
  using real_type = real;
  static constexpr int dims = 1;
  real y, u1, u2, u3;
};

// One particle of the 2D problems (the reference's x1, x2, U1, U2, U3),
// padded to 8 components so that a record is one aligned 32/64-byte span.
template <typename real>
struct alignas(8 * sizeof(real)) Rec2 {
  using real_type = real;
  static constexpr int dims = 2;
  real x, y, u1, u2, u3, p0, p1, p2;
};

// A record read once by a streaming pass, with the non-temporal hint: the L2
// keeps such lines only briefly and leaves its space to the lines a scatter
// is still filling.
template <typename R>
__device__ __forceinline__ R load_stream(const R *p) {
  typedef uint32_t w4 __attribute__((ext_vector_type(4)));
  static_assert(sizeof(R) % 16 == 0, "records are whole 16-byte words");
  w4 v[sizeof(R) / 16];
  const w4 *s = reinterpret_cast<const w4 *>(p);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(R) / 16); ++i) v[i] = __builtin_nontemporal_load(s + i);
  R r;
  __builtin_memcpy(&r, v, sizeof(R));
  return r;
}

// What the push's kick reads: the per-cell slab force F2 (1D) or the nodal
// potential and its coefficient (2D, compute_SP_2D / compute_Vlasov_2D).
struct Kick {
  const double *F2;
  const double *xnode;
  double coef;
  double *pm;  // measure_pressure face sums (slab), or null when not sampling
  unsigned long long *esc;  // re-inserted particles of the evaporation boundary, or null
};

// Constants every kernel reads, passed by value.  n_cells is the total
// number of cells (nx * ny); 1D slabs have nx = 1.
struct KParams {
  int32_t n_cells;
  int32_t boundary;     // along y
  int32_t y_model;
  int32_t dims;
  int32_t nx, ny;       // cells along x (1 in 1D) and y
  int32_t boundary_x;   // along x (2D)
  int32_t reserved;
  double dx, length_x;  // 2D
  double dy, length, wall_lo, wall_hi;
  double T_lo, T_hi, U_lo, U_hi;
  double sigma, mass, kb, b, Fn, volume, n_ratio, dt;
  double coll_pref;  // 4 pi sigma^2, multiplied left to right as in collisions.cpp:691
  uint32_t key0, key1;  // Philox key from the seed
  // Batched replicas (nrep > 0): independent slabs laid end to end, replica
  // r owning global cells [rcell_off[r], rcell_off[r + 1]) and particles
  // [rpart_off[r], rpart_off[r + 1]).  rkp[r] holds the replica's own
  // constants (cells, dy, volume, Fn, Philox key) with nrep = 0; kernels
  // work in the replica's local cell numbers and add rcell_off[r].
  int32_t nrep;
  int32_t reserved2;
  const KParams *rkp;
  const int32_t *cell_rep;    // replica of each global cell
  const int32_t *rcell_off;   // nrep + 1
  const int64_t *rpart_off;   // nrep + 1
  const uint32_t *rprio;      // candidate priority key of the step, per replica
  const uint32_t *rcand_off;  // first candidate of each replica in this step, nrep + 1
  // replica of particle b << REP_BUCKET_SHIFT for every b (the search of
  // rep_of_particle starts there)
  const int32_t *rpart_bucket;
  // per-cell particle weights (cells[num].n_ratio) when they differ from
  // cell to cell, else null and every cell has n_ratio
  const int32_t *cell_ratio;
};

// The weight of cell c's particles in its density n = n_ratio Fn N / V.
__host__ __device__ __forceinline__ double cell_weight(const KParams &kp, int32_t c) {
  return kp.cell_ratio ? (double)kp.cell_ratio[c] : kp.n_ratio;
}

// Replica of particle p (batched contexts): the last r with rpart_off[r] <= p.
constexpr int REP_BUCKET_SHIFT = 14;
__device__ __forceinline__ int32_t rep_of_particle(const KParams &kp, int64_t p) {
  if (kp.rpart_bucket) {
    // from the replica of the bucket's first particle, forward past the
    // replicas that end at or before p (one or two for replicas of more
    // than 2^14 particles)
    int32_t r = kp.rpart_bucket[p >> REP_BUCKET_SHIFT];
    while (kp.rpart_off[r + 1] <= p) ++r;
    return r;
  }
  int32_t a = 0, b = kp.nrep;
  while (b - a > 1) {
    const int32_t m = (a + b) >> 1;
    if (kp.rpart_off[m] <= p) a = m;
    else b = m;
  }
  return a;
}
// Replica of candidate k of the step: the last r with rcand_off[r] <= k.
__device__ __forceinline__ int32_t rep_of_cand(const KParams &kp, uint32_t k) {
  int32_t a = 0, b = kp.nrep;
  while (b - a > 1) {
    const int32_t m = (a + b) >> 1;
    if (kp.rcand_off[m] <= k) a = m;
    else b = m;
  }
  return a;
}

// Parameters of the Fokker-Planck and DSMC velocity models.
struct FPParams {
  int32_t model;
  int32_t reserved;
  double mu, T0, visp, crref;
};

// Hardware places workgroup b on XCD b % 8.  Renumber blocks so that each
// XCD runs a contiguous range of logical blocks: neighbouring cells (and the
// operations drawn in them) then share one XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t per = nb / 8u;
  if (per == 0u || b >= per * 8u) return b;
  return (b % 8u) * per + b / 8u;
}

// Per-cell constants of the FP_dense position drift, see k_fp_closures.
constexpr int DRIFT_N = 12;

// x^a through the deterministic log/exp (exact shortcuts for a = 0, 1/2, 1);
// used for mu(T) (scale_mu) and the VHS cross-section.
__host__ __device__ __forceinline__ double det_pow(double x, double a) {
  if (a == 0.5) return __builtin_sqrt(x);
  if (a == 1.0) return x;
  if (a == 0.0) return 1.0;
  return mcpm_exp(a * mcpm_log(x));
}

__host__ __device__ __forceinline__ double contact_y(double n, double b, int model) {
  // increase_collision_rate, Main_Functions.cpp:1906-1915
  double nb = n * b;
  if (model == MCPM_Y_CARNAHAN_STARLING) {
    double eta = nb / 4.0;
    double om = 1.0 - eta;
    return (1.0 - eta / 2.0) / (om * om * om);
  }
  return 1.0 + 0.625 * n * b + 0.2869 * (nb * nb) + 0.115 * (nb * nb * nb);
}

// Cell of slab coordinate y: floor(y / dy) (Main_Functions.cpp:2236), held
// inside [0, n_cells).  Identical arithmetic in the oracle.
__host__ __device__ __forceinline__ int32_t cell_of(double y, double dy, int32_t nc) {
  double q = mcpm_floor(y / dy);
  if (!(q >= 0.0)) return 0;  // also catches NaN
  if (q >= (double)nc) return nc - 1;
  return (int32_t)q;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// The valid lanes of the wavefront whose key equals this lane's (empty for
// an invalid lane), from one ballot per key bit: every key lies in
// [0, 2^kbits).  Its cost does not depend on how many distinct keys the
// wavefront holds.
__device__ __forceinline__ uint64_t match_key(uint32_t key, bool valid, int kbits) {
  uint64_t m = __ballot(valid);
  for (int b = 0; b < kbits; ++b) {
    const bool bit = (key >> b) & 1u;
    const uint64_t on = __ballot(bit);
    m &= bit ? on : ~on;
  }
  return valid ? m : 0ull;
}
// Bits that hold every bin index of [0, nb).
__host__ __device__ __forceinline__ int key_bits(int nb) {
  int k = 0;
  while (k < 31 && (1 << k) < nb) ++k;
  return k;
}

__device__ __forceinline__ void atomic_max_pos_double(double *addr, double v) {
  // For non-negative doubles the IEEE bit pattern orders like the value.  A
  // relaxed device-scope read (served by L2) filters most calls: the stored
  // maximum only grows, so a stale read is never above the true value and
  // cannot skip a needed update.
  if (v > 0.0 &&
      v > __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax((unsigned long long *)addr, (unsigned long long)mcpm_dbits(v));
}

// Sum of v over the workgroup in one fixed order: each wavefront folds its
// lanes pairwise (lane l adds lane l + off, off = 32, 16, ..., 1), then the
// wavefront totals are added in turn to 0; sh holds K x (blockDim / 64)
// doubles.  Every thread must call it.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *sh /* [K][blockDim/64] */) {
  const int lane = lane_id();
  const int wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane == 0) sh[k * nw + wid] = x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double x = 0.0;
    for (int w = 0; w < nw; ++w) x += sh[k * nw + w];
    v[k] = x;
  }
  __syncthreads();
}

}  // namespace mcpm
