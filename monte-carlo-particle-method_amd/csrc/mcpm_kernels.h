// mcpm_kernels.h — kernel declarations and the device-side data structures
// shared between mcpm_kernels.hip and the host layer (mcpm_api.hip).
#pragma once
#include "mcpm_device.h"

namespace mcpm {

// push modes
enum : int { PUSH_KICK = 1, PUSH_MOVE = 2, PUSH_CHECK = 4 };

// device error bits -> status codes
enum : uint32_t { ERRBIT_DOMAIN = 1u, ERRBIT_NAN = 2u, ERRBIT_BC = 4u, ERRBIT_CAPACITY = 8u };

// layout of the u32 device counter block
enum : int {
  CNT_OPS = 0,     // valid candidate operations of the current step
  CNT_ACC = 2,     // accepted collisions (u64, 8-byte aligned)
  CNT_LOG = 4,     // entries in the debug collision log
  CNT_LIST = 6,    // CNT_LIST + r: length of the list processed by round r (r >= 1)
  MAX_ROUNDS = 512,
  CNT_TOTAL = CNT_LIST + MAX_ROUNDS + 2
};

// Candidate operation {i, j, priority, t} as one uint4; i = OP_NONE marks a
// candidate without a partner.
constexpr uint32_t OP_NONE = 0xFFFFFFFFu;

// Debug log of committed operations (null pointers when disabled).
struct OpLog {
  uint32_t *p, *i, *j;
  uint8_t *a;
};

// Work arrays and host-prepared factors of the 1D field solve.
struct FieldArrays {
  double *nj, *rhs, *tB, *y, *x;   // faces
  double *mL, *invd, *mU;           // LU factors: forward -l_k, 1/d_k, -c_k/d_k
  double *sm_z;                     // Sherman-Morrison vector (periodic)
  double *gl_r, *gl_dr, *gl_cos, *gl_sin, *gl_dphi;  // 50-point tables (Dirichlet)
  double sm_vlast, sm_den;
};

// collision models sharing the reservation rounds
enum : int { COLL_ESMC = 0, COLL_DSMC = 1 };

template <typename real, bool SCATTER>
__global__ void k_push(const Rec<real> *in, Rec<real> *out, int64_t n, int32_t chunk,
                       KParams kp, const double *F2, const double *drift, int mode,
                       uint32_t step_lo, uint32_t step_hi, uint32_t *table,
                       const uint32_t *start, uint32_t *err);
__global__ void k_scan_a(const uint32_t *table, uint32_t *partial, int nchunks, int nc, int seg);
__global__ void k_scan_b(uint32_t *partial, int nseg, int nc, uint32_t *cnt, uint32_t *start,
                         double *nden, double *yc, KParams kp);
__global__ void k_scan_c(uint32_t *table, const uint32_t *partial, int nchunks, int nc, int seg);
__global__ void k_mcand(const uint32_t *cnt, const double *omax, const double *nden,
                        uint32_t *mcand, int nc, KParams kp, int model, uint32_t step_lo,
                        uint32_t step_hi);
__global__ void k_excl_scan(const uint32_t *in, uint32_t *out, int nc);
template <typename real, int MODEL>
__global__ void k_gen(const Rec<real> *P, const uint32_t *cnt, const uint32_t *start,
                      const uint32_t *mcand, const uint32_t *cstart, KParams kp,
                      uint32_t step_lo, uint32_t step_hi, uint32_t prio_key, uint4 *ops,
                      uint32_t *res0, uint32_t *counters);
template <typename real, int MODEL>
__global__ void k_round(int r, uint32_t count0, Rec<real> *P, const uint4 *in, uint4 *out,
                        uint32_t *res_cur, uint32_t *res_next, uint32_t *counters,
                        const double *nden, const double *yc, const double *omax,
                        double *omega_new, KParams kp, uint32_t step_lo, uint32_t step_hi,
                        FPParams fpp, OpLog log);
__global__ void k_omega_finish(double *omax, double *omega_new, const uint32_t *cnt, int nc,
                               int model);
__global__ void k_zero_u32(uint32_t *p, int n);
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v);
__global__ void k_field(const double *nden, FieldArrays fa, double *F2, KParams kp, double ast,
                        double bst, double eqdist);
template <typename real>
__global__ void k_moments(const Rec<real> *P, const uint32_t *cnt, const uint32_t *start,
                          double *out, KParams kp);
template <typename real>
__global__ void k_sample(const Rec<real> *P, const uint32_t *cnt, const uint32_t *start,
                         double *acc);
__global__ void k_fp_closures(const double *mom, const uint32_t *cnt, const double *nden,
                              const double *yc, double *clos, double *drift, KParams kp,
                              FPParams fpp, int64_t step, uint32_t *err);
template <typename real>
__global__ void k_fp_velocity(Rec<real> *P, const uint32_t *cnt, const uint32_t *start,
                              const double *mom, const double *clos, const double *yc,
                              KParams kp, FPParams fpp, uint32_t slo, uint32_t shi,
                              uint32_t *err);
__global__ void k_vlasov_faces(const double *nden, const double *tab, double *V, KParams kp);
__global__ void k_vlasov_force(const double *V, double *F2, KParams kp, double phi);
template <typename real>
__global__ void k_kde(const Rec<real> *P, const uint32_t *cnt, const uint32_t *start, double *ns,
                      double *d2, double *d3, KParams kp);
__global__ void k_lr2_force(const double *d2, double *F2, KParams kp, double pref);

}  // namespace mcpm
