// mcpm_geom.h — geometry of the 1D slab (Rec) and the 2D plane (Rec2):
// cell keys, the ESMC partner cell and midpoint cell, and the per-axis
// boundary treatment of the push.  Overloads on the record type let one set
// of kernels serve both.
#pragma once
#include "mcpm_device.h"

#pragma clang fp contract(off)

namespace mcpm {

// ---- cell keys (cell_update binning, Main_Functions.cpp:2231-2266) -------
template <typename real>
__host__ __device__ __forceinline__ int32_t cell_key(const Rec<real> &r, const KParams &kp) {
  return cell_of((double)r.y, kp.dy, kp.n_cells);
}
template <typename real>
__host__ __device__ __forceinline__ int32_t cell_key(const Rec2<real> &r, const KParams &kp) {
  int32_t cx = cell_of((double)r.x, kp.dx, kp.nx);
  int32_t cy = cell_of((double)r.y, kp.dy, kp.ny);
  return cy * kp.nx + cx;  // num = k N0 + j (Main_Functions.cpp:2263)
}

// Sort keys carried by the two-digit sort: the cell in the low bits; on the
// plane the top two bits hold the particle's quarter of its cell, q =
// (x >= centre_x) + 2 (y >= centre_y), the nodal deposit's index
// (Main_Functions.cpp:2268-2342).  The sort bins on the cell bits only.
constexpr uint32_t KEY_QUARTER_SHIFT = 30;
constexpr uint32_t KEY_CELL_MASK = (1u << KEY_QUARTER_SHIFT) - 1u;
template <typename real>
__host__ __device__ __forceinline__ uint32_t key_quarter_bits(const Rec<real> &, int32_t,
                                                              const KParams &) {
  return 0u;
}
template <typename real>
__host__ __device__ __forceinline__ uint32_t key_quarter_bits(const Rec2<real> &r, int32_t ckey,
                                                              const KParams &kp) {
  const int32_t cx = ckey % kp.nx, cy = ckey / kp.nx;
  const double ccx = 0.5 * ((double)cx * kp.dx + (double)(cx + 1) * kp.dx);
  const double ccy = 0.5 * ((double)cy * kp.dy + (double)(cy + 1) * kp.dy);
  const uint32_t q = ((double)r.x < ccx ? 0u : 1u) + ((double)r.y < ccy ? 0u : 2u);
  return q << KEY_QUARTER_SHIFT;
}

__host__ __device__ __forceinline__ double wrap_coord(double v, double L) {
  if (v < 0.0) v += L;
  else if (v >= L) v -= L;
  return v;
}

// Cell index along one axis of a position that may lie outside the box:
// false when outside a walled axis (the candidate has no partner,
// collisions.cpp:668-671), wrapped on a periodic axis.
__host__ __device__ __forceinline__ bool axis_cell(double v, double h, double L, int32_t n,
                                                   bool periodic, int32_t *out) {
  if (periodic) {
    *out = cell_of(wrap_coord(v, L), h, n);
    return true;
  }
  double q = mcpm_floor(v / h);
  if (!(q >= 0.0) || q >= (double)n) return false;
  *out = (int32_t)q;
  return true;
}

// ---- partner cell J at r_i + sigma dir (collisions.cpp:661-671) -----------
template <typename real>
__host__ __device__ __forceinline__ bool partner_cell(const Rec<real> &ri, double d0, double d1,
                                                      const KParams &kp, int32_t *J) {
  (void)d0;
  return axis_cell((double)ri.y + kp.sigma * d1, kp.dy, kp.length, kp.n_cells,
                   kp.boundary == MCPM_BC_PERIODIC, J);
}
template <typename real>
__host__ __device__ __forceinline__ bool partner_cell(const Rec2<real> &ri, double d0, double d1,
                                                      const KParams &kp, int32_t *J) {
  int32_t jx, jy;
  if (!axis_cell((double)ri.x + kp.sigma * d0, kp.dx, kp.length_x, kp.nx,
                 kp.boundary_x == MCPM_BC_PERIODIC, &jx))
    return false;
  if (!axis_cell((double)ri.y + kp.sigma * d1, kp.dy, kp.length, kp.ny,
                 kp.boundary == MCPM_BC_PERIODIC, &jy))
    return false;
  *J = jy * kp.nx + jx;  // J = ind[1] N0 + ind[0] (collisions.cpp:666)
  return true;
}

// Collision_ESMC_new4's partner cell on the slab (collisions.cpp:781-792):
// floor((y_i + sigma dir_y) / dy) clamped into the slab's cells.
template <typename real>
__host__ __device__ __forceinline__ int32_t partner_cell_clamped(const Rec<real> &ri, double d1,
                                                                 const KParams &kp) {
  double q = mcpm_floor(((double)ri.y + kp.sigma * d1) / kp.dy);
  if (q < 0.0) return 0;
  if (q > (double)(kp.n_cells - 1)) return kp.n_cells - 1;
  return (int32_t)q;
}

// ---- midpoint cell of a pair (collisions.cpp:682-686) ----------------------
__host__ __device__ __forceinline__ double axis_mid(double a, double b, double L, bool periodic) {
  if (!periodic) return (a + b) / 2.0;
  double d = b - a;
  if (d > 0.5 * L) d -= L;
  else if (d < -0.5 * L) d += L;
  return wrap_coord(a + d / 2.0, L);
}
template <typename real>
__host__ __device__ __forceinline__ int32_t mid_cell(const Rec<real> &a, const Rec<real> &b,
                                                     const KParams &kp) {
  return cell_of(axis_mid((double)a.y, (double)b.y, kp.length, kp.boundary == MCPM_BC_PERIODIC),
                 kp.dy, kp.n_cells);
}
template <typename real>
__host__ __device__ __forceinline__ int32_t mid_cell(const Rec2<real> &a, const Rec2<real> &b,
                                                     const KParams &kp) {
  double xm = axis_mid((double)a.x, (double)b.x, kp.length_x, kp.boundary_x == MCPM_BC_PERIODIC);
  double ym = axis_mid((double)a.y, (double)b.y, kp.length, kp.boundary == MCPM_BC_PERIODIC);
  return cell_of(ym, kp.dy, kp.ny) * kp.nx + cell_of(xm, kp.dx, kp.nx);
}

// Specular walls at 0 and L on the walled axes of the plane, as
// BC_specular_reflection does it (Main_Functions.cpp:6399-6525): every pass
// tests each walled axis once, reflecting when x - L > -1e-16 (x -= 2|x - L|)
// or else when x < 1e-16 (x += 2|x|) and reversing that velocity component;
// passes repeat while a coordinate lies outside [0, L].  False when 20
// passes were not enough ("TOO MANY BC OPERATIONS", :6415-6420).
__host__ __device__ __forceinline__ bool reflect_once(double *v, double *u, double L) {
  if (*v - L > -1e-16) {
    *u = -*u;
    *v = *v - 2.0 * __builtin_fabs(*v - L);
  } else if (*v < 1e-16) {
    *u = -*u;
    *v = *v + 2.0 * __builtin_fabs(*v);
  }
  return *v - L > 0.0 || *v < 0.0;  // still outside
}
__host__ __device__ __forceinline__ bool specular_plane(double *x, double *u1, bool wall_x,
                                                        double Lx, double *y, double *u2,
                                                        bool wall_y, double Ly) {
  for (int pass = 0; pass < 20; ++pass) {
    bool out = false;
    if (wall_x) out |= reflect_once(x, u1, Lx);
    if (wall_y) out |= reflect_once(y, u2, Ly);
    if (!out) return true;
  }
  return false;
}

}  // namespace mcpm
