/*
 * mcpm_oracle.cpp — CPU oracle for the slab particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the
 * smoke check and the CPU-baseline leg of bench.py load liboracle.so and
 * compare the GPU library against it.  Nothing in the product (libmcpm.so,
 * its Python binding, the C++ host layer) links, loads or calls it.
 *
 * It restates, sequentially and in plain loops, the reference algorithms of
 * mohsensadr/Monte-Carlo-Particle-Method for the 1D slab step:
 *   cell_update binning + stable index lists   Main_Functions.cpp:2138-2376
 *   cell_update moments                         Main_Functions.cpp:2461-2593
 *   increase_collision_rate (Y)                 Main_Functions.cpp:1906-1915
 *   Collision_ESMC_new3                         collisions.cpp:583-720
 *     (1D partner-cell rule of new4,             collisions.cpp:784-792)
 *   poisson_solver + Thompson_3d                Main_Functions.cpp:3285-3581,
 *                                               Math_Func.cpp:28-52
 *   Position_FP_ideal_gas (slab axis)           Fokker_Planck.cpp:948-969
 *   BC_thermal_wall_3D (slab axis)              Main_Functions.cpp:4917-5248
 *   BC_specular_reflection (slab axis)          Main_Functions.cpp:6399-6525
 *   cell_poly_velocity                          Main_Functions.cpp:1918-1968
 *   omega_max initialisation                    main.cpp:250-269
 *   Collision_DSMC_VHS / _HS                    collisions.cpp:2063-2141, 1989-2061
 *   Collision_CBA_VHS / _HS                     collisions.cpp:2144-2248, 2251-2346
 *   random_generator + Velocity_FP_parallel /   Fokker_Planck.cpp:20-145,
 *     Velocity_FP_linear                        1065-1182, 1186-1287
 * with these deliberate, documented substitutions:
 *   - random numbers: Philox4x32-10 addressed by (purpose, step, cell or
 *     particle, draw) from mcpm_rng.h instead of random_device-seeded
 *     mt19937 / rand();
 *   - candidate order: new3 draws each candidate's cell with probability
 *     proportional to the cell's remaining quota (collisions.cpp:631-646),
 *     i.e. it processes a uniformly random interleaving of all cells'
 *     candidates.  Here that interleaving is a keyed bijection of the
 *     candidate index (mcpm_priority), sorted ascending, so it is a pure
 *     function of (seed, step);
 *   - particle order: the per-cell index lists become a physical stable
 *     counting sort, so indices_inside[c][u] is start[c] + u;
 *   - periodic boundaries (cyclic tridiagonal by Sherman-Morrison, wrapped
 *     partner cell and minimum-image midpoint) for the periodic slab, which
 *     the reference's 1D path does not offer.
 *
 * The plane (problem injection) adds:
 *   cell_update 2D binning and nodal deposit   Main_Functions.cpp:2231-2364
 *   Collision_ESMC_new3 2D partner cell        collisions.cpp:658-687
 *   BC_specular_reflection (x and y)           Main_Functions.cpp:6399-6525
 *   compute_SP_2D                              FEM.cpp:249-610
 *     (assembled matrix with Dirichlet rows, solved here by banded LU
 *      instead of PARDISO, FEM.cpp:450-458)
 *   compute_Vlasov_2D                          FEM.cpp:613-726
 * and, as an extension, periodic planes (dense LU, wrapped neighbours).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "mcpm.h"
#include "mcpm_rng.h"

namespace {

const double PI = 3.14159265358979323846;

double orc_y(double n, const mcpm_params &p) {
  // increase_collision_rate, Main_Functions.cpp:1906-1915
  double nb = n * p.b;
  if (p.y_model == MCPM_Y_CARNAHAN_STARLING) {
    double eta = nb / 4.0;
    double om = 1.0 - eta;
    return (1.0 - eta / 2.0) / (om * om * om);
  }
  return 1.0 + 0.625 * n * p.b + 0.2869 * (nb * nb) + 0.115 * (nb * nb * nb);
}

int orc_cell(double y, double dy, int nc) {
  double q = mcpm_floor(y / dy);
  if (!(q >= 0.0)) return 0;
  if (q >= (double)nc) return nc - 1;
  return (int)q;
}

double dn_xi(int i, double xi, double eta);
double dn_eta(int i, double xi, double eta);

struct Oracle {
  mcpm_params p;
  int nc = 0;
  double dy = 0.0;
  int64_t n = 0;
  // record layout: slab (y, u1, u2, u3), plane (x, y, u1, u2, u3)
  int dims = 1, nx = 1, ny = 1, S = 4, U = 1;
  double dx = 1.0;
  std::vector<double> rec;  // S per particle, stored through `real` rounding
  // the plane: per-cell corner copies of the nodal density (cells[c].n_n)
  // and the nodal field of the last solve
  std::vector<double> n_n, pot;
  double kick_coef = 0.0;
  std::vector<int> cnt, start;
  std::vector<double> nden, yc, omax, F2, samples;
  // measure_pressure face sums (slab)
  std::vector<double> pm;
  bool pressure_on = false;
  int64_t escaped = 0;  // particles re-inserted by the REINJECT boundary
  bool kick_pending = false;
  bool sorted = false;
  uint32_t err = 0;
  // collision log of the last call, in processing order
  std::vector<uint32_t> log_p, log_i, log_j;
  std::vector<uint8_t> log_a;
  int64_t last_cand = 0, last_acc = 0;
  // Fokker-Planck closures of the last orc_closures call
  std::vector<double> mom, clos;
  bool have_closures = false;
  // per-cell weights cells[num].n_ratio (empty: p.n_ratio everywhere) and
  // the particles' weights, carried by the sort while track_w
  std::vector<int32_t> cr, pw;
  bool track_w = false, w_pending = false;
};

template <typename real>
double rnd(double v) {
  return (double)(real)v;
}

// Sums of per-particle terms in one fixed order, the order of a 256-thread
// workgroup: term k goes to lane k mod 256, each lane adds its terms in the
// order they arrive; each group of 64 lanes then folds pairwise (lane l adds
// lane l + off for off = 32, 16, ..., 1) and the four group totals are
// added in turn to 0.  The reference sums serially (e.g. cell_update,
// Main_Functions.cpp:2483-2593); this is the same sum in another
// association, fixed so that sums over a cell come out bit for bit the same
// on every run.
template <int K>
struct WgSum {
  static const int NT = 256;
  std::vector<double> lane;  // NT x K; lanes at or past `used` hold 0
  int used = 0;
  WgSum() : lane((size_t)NT * K) {}
  void add(int l, const double (&v)[K]) {
    double *a = &lane[(size_t)l * K];
    for (; used <= l; ++used)
      for (int q = 0; q < K; ++q) lane[(size_t)used * K + q] = 0.0;
    for (int q = 0; q < K; ++q) a[q] += v[q];
  }
  // lane l exists (holds its 0 start value) without a term
  void touch(int l) {
    for (; used <= l; ++used)
      for (int q = 0; q < K; ++q) lane[(size_t)used * K + q] = 0.0;
  }
  void total(double (&out)[K]) const {
    if (used == 0) {
      for (int q = 0; q < K; ++q) out[q] = 0.0;
      return;
    }
    for (int q = 0; q < K; ++q) {
      double t = 0.0;
      for (int w = 0; w < NT / 64; ++w) {
        double x[64];
        for (int l = 0; l < 64; ++l) {
          const int g = 64 * w + l;
          x[l] = g < used ? lane[(size_t)g * K + q] : 0.0;
        }
        for (int off = 32; off > 0; off >>= 1)
          for (int l = 0; l < off; ++l) x[l] = x[l] + x[l + off];
        t += x[0];
      }
      out[q] = t;
    }
  }
};

// Cell of particle k: floor(y/dy) on the slab, k N0 + j on the plane
// (Main_Functions.cpp:2233-2263).
int orc_key(const Oracle &o, int64_t k) {
  const double *r = &o.rec[(size_t)o.S * k];
  if (o.dims == 2) return orc_cell(r[1], o.dy, o.ny) * o.nx + orc_cell(r[0], o.dx, o.nx);
  return orc_cell(r[0], o.dy, o.nc);
}

// One coordinate of the point r_i + sigma dir: its cell on a periodic axis
// after wrapping; on a walled axis false when outside (collisions.cpp:668-671).
bool orc_axis(double v, double h, double L, int n, bool periodic, int *out) {
  if (periodic) {
    if (v < 0.0) v += L;
    else if (v >= L) v -= L;
    *out = orc_cell(v, h, n);
    return true;
  }
  double q = mcpm_floor(v / h);
  if (!(q >= 0.0) || q >= (double)n) return false;
  *out = (int)q;
  return true;
}

bool orc_partner(const Oracle &o, int64_t i, double d0, double d1, int *J) {
  const double *r = &o.rec[(size_t)o.S * i];
  const mcpm_params &p = o.p;
  if (o.dims == 2) {  // J = ind[1] N0 + ind[0] (collisions.cpp:660-666)
    int jx, jy;
    if (!orc_axis(r[0] + p.sigma * d0, o.dx, p.length_x, o.nx, p.boundary_x == MCPM_BC_PERIODIC,
                  &jx))
      return false;
    if (!orc_axis(r[1] + p.sigma * d1, o.dy, p.length, o.ny, p.boundary == MCPM_BC_PERIODIC, &jy))
      return false;
    *J = jy * o.nx + jx;
    return true;
  }
  return orc_axis(r[0] + p.sigma * d1, o.dy, p.length, o.nc, p.boundary == MCPM_BC_PERIODIC, J);
}

double orc_mid_axis(double a, double b, double L, bool periodic) {
  if (!periodic) return (a + b) / 2.0;  // collisions.cpp:682-683
  double d = b - a;
  if (d > 0.5 * L) d -= L;
  else if (d < -0.5 * L) d += L;
  double m = a + d / 2.0;
  if (m < 0.0) m += L;
  else if (m >= L) m -= L;
  return m;
}

int orc_mid(const Oracle &o, int64_t i, int64_t j) {
  const double *ri = &o.rec[(size_t)o.S * i], *rj = &o.rec[(size_t)o.S * j];
  const mcpm_params &p = o.p;
  if (o.dims == 2) {
    double xm = orc_mid_axis(ri[0], rj[0], p.length_x, p.boundary_x == MCPM_BC_PERIODIC);
    double ym = orc_mid_axis(ri[1], rj[1], p.length, p.boundary == MCPM_BC_PERIODIC);
    return orc_cell(ym, o.dy, o.ny) * o.nx + orc_cell(xm, o.dx, o.nx);  // IJmid (:686)
  }
  return orc_cell(orc_mid_axis(ri[0], rj[0], p.length, p.boundary == MCPM_BC_PERIODIC), o.dy,
                  o.nc);
}

// n = n_ratio Fn weight / volume (Main_Functions.cpp:2588, 2601) with the
// cell's own n_ratio where the cells have weights, and Y(n).
void orc_densities(Oracle &o) {
  for (int c = 0; c < o.nc; ++c) {
    const double w = o.cr.empty() ? o.p.n_ratio : (double)o.cr[c];
    o.nden[c] = w * o.p.Fn * (double)o.cnt[c] / o.p.cell_volume;
    o.yc[c] = orc_y(o.nden[c], o.p);
  }
}

// Stable counting sort of the particles by cell (cell_update binning and
// index lists, Main_Functions.cpp:2231-2376).
void orc_sort(Oracle &o, bool check) {
  const int nc = o.nc;
  std::vector<int> key(o.n);
  std::fill(o.cnt.begin(), o.cnt.end(), 0);
  for (int64_t k = 0; k < o.n; ++k) {
    const double *r = &o.rec[(size_t)o.S * k];
    double y = o.dims == 2 ? r[1] : r[0];
    if (check && (!(y >= 0.0) || !(y <= o.p.length))) o.err |= 1u;
    if (check && o.dims == 2 && (!(r[0] >= 0.0) || !(r[0] <= o.p.length_x))) o.err |= 1u;
    key[k] = orc_key(o, k);
    o.cnt[key[k]]++;
  }
  int run = 0;
  for (int c = 0; c < nc; ++c) {
    o.start[c] = run;
    run += o.cnt[c];
  }
  o.start[nc] = run;
  std::vector<int> fill(o.start.begin(), o.start.end() - 1);
  std::vector<double> out(o.rec.size());
  std::vector<int32_t> wout(o.track_w ? (size_t)o.n : 0);
  for (int64_t k = 0; k < o.n; ++k) {
    int d = fill[key[k]]++;
    memcpy(&out[(size_t)o.S * d], &o.rec[(size_t)o.S * k], o.S * sizeof(double));
    if (o.track_w) wout[d] = o.pw[k];
  }
  o.rec.swap(out);
  if (o.track_w) o.pw.swap(wout);
  orc_densities(o);
  if (!o.cr.empty() && o.w_pending) {
    // the particles take their cells' weights (main.cpp:239-242)
    o.pw.assign((size_t)o.n, 0);
    for (int c = 0; c < nc; ++c)
      for (int u = 0; u < o.cnt[c]; ++u) o.pw[(size_t)o.start[c] + u] = o.cr[c];
    o.track_w = true;
    o.w_pending = false;
  }
  o.sorted = true;
}

// measure_pressure (Main_Functions.cpp:13-34): cells num in [min, max) of
// the old and new cell indices take Fn m |U2|; on a periodic slab the cells
// of the unwrapped flight modulo nc; indices outside the slab are skipped.
void orc_pressure_faces(Oracle &o, double y_old, double y_new, double u2, bool periodic) {
  double qa = floor(y_old / o.dy), qb = floor(y_new / o.dy);
  if (qa == qb || !(fabs(qa) < 1e12) || !(fabs(qb) < 1e12)) return;
  long long a = (long long)qa, b = (long long)qb;
  long long lo = std::min(a, b), hi = std::max(a, b);
  if (hi - lo > o.nc) hi = lo + o.nc;
  for (long long num = lo; num < hi; ++num) {
    long long c = num;
    if (periodic) c = ((c % o.nc) + o.nc) % o.nc;
    else if (c < 0 || c >= o.nc) continue;
    o.pm[c] += o.p.Fn * o.p.mass * fabs(u2);
  }
}

// Kick, stream and boundary of one particle (see header list).
template <typename real>
void orc_push_one(Oracle &o, int64_t slot, bool kick, bool move, uint32_t slo, uint32_t shi) {
  const mcpm_params &p = o.p;
  double *r = &o.rec[4 * (size_t)slot];
  double y = r[0];
  if (kick) {
    int c = orc_cell(y, o.dy, o.nc);
    r[2] = rnd<real>(r[2] + o.F2[c] * p.delta_t);  // U2 += F2 dt (Main_Functions.cpp:3576)
  }
  if (!move) return;
  double v = r[2];
  const double y_start = y;
  bool drifted = false;
  if (p.model == MCPM_MODEL_FP_DENSE && o.have_closures) {
    // DFP_extra_streaming, slab axis (Fokker_Planck.cpp:932-937)
    int c = orc_cell(y, o.dy, o.nc);
    const double *m = &o.mom[(size_t)c * MCPM_NMOM];
    const double *cl = &o.clos[(size_t)c * MCPM_NCLOS];
    if (o.cnt[c] > 1) {  // closures exist for cells with two or more particles
      double e = (r[1] - m[0]) * (r[1] - m[0]) + (r[2] - m[1]) * (r[2] - m[1]) +
                 (r[3] - m[2]) * (r[3] - m[2]);
      double q1 = m[10] + m[15] + m[17];  // M3[1] + M3[6] + M3[8]
      double dt = p.delta_t;
      y = y + dt * ((r[1] - m[0]) * cl[13] + (r[2] - m[1]) * cl[14] + (r[3] - m[2]) * cl[15] +
                    cl[20] * (e - 3.0 * p.kb * m[29] / p.mass)) +
          cl[22] * ((r[2] - m[1]) * e - q1) * dt;
      y = y - dt * cl[24];
      drifted = true;
    }
  }
  const double y_old = y;  // x2_old (store_old_positions, main.cpp:484)
  y = y + v * p.delta_t;  // x2 += dt U2 (Fokker_Planck.cpp:963)
  // speed used for the time spent past a wall: the displacement this step
  const double v_eff = drifted ? (y - y_start) / p.delta_t : v;
  if (p.boundary == MCPM_BC_PERIODIC) {
    if (o.pressure_on) orc_pressure_faces(o, y_old, y, v, true);
    if (y < 0.0) y += p.length;
    else if (y >= p.length) y -= p.length;
    real yr = (real)y;
    if (!(yr < (real)p.length)) yr = (real)0;
    if (!(yr >= (real)0)) yr = (real)0;
    r[0] = (double)yr;
    return;
  }
  if (p.boundary == MCPM_BC_SPECULAR) {
    int guard = 0;
    while ((y < p.wall_lo || y > p.wall_hi) && guard < 20) {
      if (y > p.wall_hi) y = 2.0 * p.wall_hi - y;
      else y = 2.0 * p.wall_lo - y;
      v = -v;
      ++guard;
    }
    if (guard >= 20) o.err |= 4u;
    r[0] = rnd<real>(y);
    r[2] = rnd<real>(v);
    if (o.pressure_on) orc_pressure_faces(o, y_old, r[0], r[2], false);
    return;
  }
  if (p.boundary == MCPM_BC_REINJECT) {
    // BC_evaporation (Main_Functions.cpp:5444-5617): a particle past y > L - eps
    // (:5519-5546) or y < eps (:5547-5574) is put back at
    // L/2 + uniform(-1, 1) L/4 with three normal velocity components at
    // gas.T (here T_wall_lo), and box.out counts it.
    uint32_t k = 0;
    while ((y < p.wall_lo || y > p.wall_hi) && k < 4) {
      uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
      const uint32_t c3 = mcpm_ctr3(MCPM_TAG_REINJECT, shi, (uint32_t)((uint64_t)slot >> 32));
      mcpm_u32x4 a = mcpm_philox((uint32_t)slot, 2u * k, slo, c3, key0, key1);
      mcpm_u32x4 b = mcpm_philox((uint32_t)slot, 2u * k + 1u, slo, c3, key0, key1);
      y = 0.5 * p.length + (2.0 * mcpm_u01(a.v[0]) - 1.0) * (0.25 * p.length);
      double z0, z1, z2, z3;
      mcpm_normal2(a.v[1], a.v[2], &z0, &z1);
      mcpm_normal2(b.v[0], b.v[1], &z2, &z3);
      const double sd = sqrt(p.kb * p.T_wall_lo / p.mass);
      r[1] = rnd<real>(sd * z0);
      r[2] = rnd<real>(sd * z1);
      r[3] = rnd<real>(sd * z2);
      ++o.escaped;
      ++k;
    }
    r[0] = rnd<real>(y);
    if (o.pressure_on && k == 0) orc_pressure_faces(o, y_old, r[0], r[2], false);
    return;
  }
  // diffuse walls (BC_thermal_wall_3D, upper/lower wall branches :5044-5127)
  double dt_left = p.delta_t;
  double v_prev = v_eff;
  uint32_t bounce = 0;
  while (y < p.wall_lo || y > p.wall_hi) {
    if (bounce >= 20) {
      o.err |= 4u;
      y = y < p.wall_lo ? p.wall_lo : p.wall_hi;
      break;
    }
    bool upper = y > p.wall_hi;
    double yw = upper ? p.wall_hi : p.wall_lo;
    double T = upper ? p.T_wall_hi : p.T_wall_lo;
    double Uw = upper ? p.U_wall_hi : p.U_wall_lo;
    uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
    mcpm_u32x4 w = mcpm_philox((uint32_t)slot, bounce, slo,
                               mcpm_ctr3(MCPM_TAG_WALL, shi, (uint32_t)((uint64_t)slot >> 32)),
                               key0, key1);
    double z0, z1;
    mcpm_normal2(w.v[0], w.v[1], &z0, &z1);
    double sd = sqrt(p.kb * T / p.mass);
    double vn = sqrt(2.0 * p.kb * T / p.mass) * sqrt(-mcpm_log(mcpm_u01(w.v[2])));
    double u1 = sd * z0 + Uw;
    double u3 = sd * z1;
    double u2 = upper ? -vn : vn;
    double dtr = fabs((y - yw) / v_prev);
    if (dtr > dt_left) dtr = dt_left;
    if (dt_left > 1e-15) y = yw + dtr * u2;
    else y = yw;
    dt_left = dt_left - dtr;
    v_prev = u2;
    r[1] = rnd<real>(u1);
    r[2] = rnd<real>(u2);
    r[3] = rnd<real>(u3);
    ++bounce;
  }
  r[0] = rnd<real>(y);
  if (o.pressure_on) orc_pressure_faces(o, y_old, r[0], r[2], false);
}

// One coordinate's periodic wrap, stored through `real` (the value must stay
// below L after rounding).
template <typename real>
double wrap_real(double v, double L) {
  if (v < 0.0) v += L;
  else if (v >= L) v -= L;
  real vr = (real)v;
  if (!(vr < (real)L)) vr = (real)0;
  if (!(vr >= (real)0)) vr = (real)0;
  return (double)vr;
}

// The plane's kick and move for particle k:
//   kick: U -= grad(Q1 interpolant of the nodal field) coef / m dt
//         (FEM.cpp:568-582, 709-721; coef = ast or phi)
//   move: x += U1 dt, y += U2 dt (Position_FP_ideal_gas, Fokker_Planck.cpp:948-969)
//         then BC_specular_reflection on the walled axes
//         (Main_Functions.cpp:6399-6525) or a periodic wrap.
// Force per unit mass times -1 at (x, y) in cell num: the gradient of the
// bilinear interpolant of the nodal field pot, times coef / m
// (FEM.cpp:568-580).
void fem_force(const Oracle &o, int num, double x, double y, const double *pot, double coef,
               double *F1o, double *F2o) {
  const mcpm_params &p = o.p;
  {
    const int j = num % o.nx, kk = num / o.nx;
    // cells[num] geometry (myIO.cpp:1455-1479)
    const double d0 = (double)j * o.dx, d1 = (double)(j + 1) * o.dx;
    const double d2 = (double)kk * o.dy, d3 = (double)(kk + 1) * o.dy;
    const double cdx = d1 - d0, cdy = d3 - d2;
    const double ccx = 0.5 * (d0 + d1), ccy = 0.5 * (d2 + d3);
    const double xn0 = ccx - cdx / 2.0, yn0 = ccy - cdy / 2.0;
    const int id[4] = {kk * (o.nx + 1) + j, kk * (o.nx + 1) + j + 1, (kk + 1) * (o.nx + 1) + j,
                       (kk + 1) * (o.nx + 1) + j + 1};
    const double xi = (x - xn0) / cdx, eta = (y - yn0) / cdy;
    double F1 = 0.0, F2 = 0.0;
    for (int q = 0; q < 4; ++q) {
      F1 += dn_xi(q, xi, eta) * pot[id[q]];
      F2 += dn_eta(q, xi, eta) * pot[id[q]];
    }
    *F1o = F1 / p.mass * coef / cdx;
    *F2o = F2 / p.mass * coef / cdy;
  }
}

template <typename real>
void orc_push2_one(Oracle &o, int64_t k, bool kick, bool move) {
  const mcpm_params &p = o.p;
  double *r = &o.rec[(size_t)o.S * k];
  if (kick) {
    double F1, F2;
    fem_force(o, orc_key(o, k), r[0], r[1], o.pot.data(), o.kick_coef, &F1, &F2);
    r[2] = rnd<real>(r[2] - F1 * p.delta_t);
    r[3] = rnd<real>(r[3] - F2 * p.delta_t);
  }
  if (!move) return;
  double x = r[0] + r[2] * p.delta_t, y = r[1] + r[3] * p.delta_t;
  double u1 = r[2], u2 = r[3];
  const bool wx = p.boundary_x != MCPM_BC_PERIODIC, wy = p.boundary != MCPM_BC_PERIODIC;
  if (wx || wy) {
    const double Lx = p.length_x, Ly = p.length;
    const double tol1 = 1e-16, tol2 = -1e-16;
    bool done = false;
    for (int pass = 0; pass < 20 && !done; ++pass) {
      if (wx) {
        if (x - Lx > tol2) {
          u1 = -u1;
          x = x - 2.0 * fabs(x - Lx);
        } else if (x < tol1) {
          u1 = -u1;
          x = x + 2.0 * fabs(x);
        }
      }
      if (wy) {
        if (y - Ly > tol2) {
          u2 = -u2;
          y = y - 2.0 * fabs(y - Ly);
        } else if (y < tol1) {
          u2 = -u2;
          y = y + 2.0 * fabs(y);
        }
      }
      done = true;
      if (wx && (x - Lx > 0.0 || x < 0.0)) done = false;
      if (wy && (y - Ly > 0.0 || y < 0.0)) done = false;
    }
    if (!done) o.err |= 4u;
  }
  r[2] = rnd<real>(u1);
  r[3] = rnd<real>(u2);
  r[0] = wx ? rnd<real>(x) : wrap_real<real>(x, p.length_x);
  r[1] = wy ? rnd<real>(y) : wrap_real<real>(y, p.length);
}

template <typename real>
void orc_push(Oracle &o, bool kick, bool move, int64_t step) {
  uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  const int U = o.U;
  for (int64_t k = 0; k < o.n; ++k) {
    double *r = &o.rec[(size_t)o.S * k];
    if (r[U] != r[U] || r[U + 1] != r[U + 1] || r[U + 2] != r[U + 2]) o.err |= 2u;
    if (o.dims == 2) orc_push2_one<real>(o, k, kick, move);
    else orc_push_one<real>(o, k, kick, move, slo, shi);
  }
}

// LSD radix sort of (priority, payload) pairs.
void radix_sort_pairs(std::vector<uint64_t> &a) {
  std::vector<uint64_t> tmp(a.size());
  for (int shift = 32; shift < 64; shift += 8) {
    size_t count[257] = {0};
    for (uint64_t v : a) count[((v >> shift) & 0xFF) + 1]++;
    for (int b = 0; b < 256; ++b) count[b + 1] += count[b];
    for (uint64_t v : a) tmp[count[(v >> shift) & 0xFF]++] = v;
    a.swap(tmp);
  }
}

struct Cand {
  uint32_t c, t, i, j;
};

// Collision_ESMC_new3 (collisions.cpp:583-720), candidates in priority order.
// With new4, Collision_ESMC_new4 (collisions.cpp:722-837) instead: cells
// from the last to the first, each cell's candidates in turn (:758), only
// cells of two or more particles draw (:763-765), the partner cell
// floor((y_i + sigma dir_y) / dy) clamped into the slab (:781-792; wrapped on
// a periodic slab), acceptance only with g.dir > 0 (:819).  Its log carries
// each candidate's rank in that order as the priority.
template <typename real>
void orc_collide(Oracle &o, int64_t step, mcpm_collision_stats *st, bool new4 = false) {
  const mcpm_params &p = o.p;
  const int nc = o.nc;
  uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
  // candidate counts (collisions.cpp:617-629)
  std::vector<uint32_t> mc(nc), cs(nc + 1);
  uint64_t M = 0;
  for (int c = 0; c < nc; ++c) {
    mcpm_u32x4 w = mcpm_philox((uint32_t)c, 0u, slo, mcpm_ctr3(MCPM_TAG_MCAND, shi, 0u), key0, key1);
    double m = mcpm_floor(0.5 * o.omax[c] * (1.0 * (double)o.cnt[c]) + mcpm_u01(w.v[0]));
    mc[c] = m > 0.0 ? (uint32_t)m : 0u;
    if (new4 && o.cnt[c] < 2) mc[c] = 0u;
    cs[c] = (uint32_t)M;
    M += mc[c];
  }
  cs[nc] = (uint32_t)M;
  uint32_t pkey = mcpm_philox(0u, 0u, slo, mcpm_ctr3(MCPM_TAG_PRIO, shi, 0u), key0, key1).v[0];
  std::vector<Cand> cand;
  std::vector<uint64_t> order;
  cand.reserve(M);
  order.reserve(M);
  for (int c = 0; c < nc; ++c) {
    if (o.cnt[c] == 0) continue;
    for (uint32_t t = 0; t < mc[c]; ++t) {
      mcpm_u32x4 w = mcpm_philox(t, (uint32_t)c, slo, mcpm_ctr3(MCPM_TAG_CAND, shi, 0u), key0, key1);
      uint32_t i = (uint32_t)o.start[c] + mcpm_uint_below(w.v[0], (uint32_t)o.cnt[c]);
      double d0, d1, d2;
      mcpm_iso_dir(w.v[1], w.v[2], &d0, &d1, &d2);
      int J;  // cell at r_i + sigma dir (collisions.cpp:658-671)
      if (new4 && p.boundary != MCPM_BC_PERIODIC) {
        double q = mcpm_floor((o.rec[(size_t)o.S * i] + p.sigma * d1) / o.dy);
        J = q < 0.0 ? 0 : (q > (double)(nc - 1) ? nc - 1 : (int)q);
      } else if (!orc_partner(o, i, d0, d1, &J)) {
        continue;
      }
      if (o.cnt[J] == 0) continue;
      uint32_t j = (uint32_t)o.start[J] + mcpm_uint_below(w.v[3], (uint32_t)o.cnt[J]);
      if (i == j) continue;  // collisions.cpp:677-678
      uint32_t pr = new4 ? ((uint32_t)M - cs[c + 1]) + t : mcpm_priority(cs[c] + t, pkey);
      order.push_back(((uint64_t)pr << 32) | (uint64_t)cand.size());
      cand.push_back(Cand{(uint32_t)c, t, i, j});
    }
  }
  radix_sort_pairs(order);
  std::vector<double> omega(nc, 0.0);
  o.log_p.resize(order.size());
  o.log_i.resize(order.size());
  o.log_j.resize(order.size());
  o.log_a.resize(order.size());
  int64_t acc = 0;
  const double pref = 4.0 * PI * p.sigma * p.sigma;
  for (size_t q = 0; q < order.size(); ++q) {
    const Cand &cd = cand[order[q] & 0xFFFFFFFFu];
    mcpm_u32x4 w = mcpm_philox(cd.t, cd.c, slo, mcpm_ctr3(MCPM_TAG_CAND, shi, 0u), key0, key1);
    mcpm_u32x4 w2 = mcpm_philox(cd.t, cd.c, slo, mcpm_ctr3(MCPM_TAG_CAND, shi, 1u), key0, key1);
    double d0, d1, d2;
    mcpm_iso_dir(w.v[1], w.v[2], &d0, &d1, &d2);
    double *ri = &o.rec[(size_t)o.S * cd.i] + o.U - 1;  // ri[1..3] = U1..U3
    double *rj = &o.rec[(size_t)o.S * cd.j] + o.U - 1;
    int cm = orc_mid(o, cd.i, cd.j);
    int J = orc_key(o, cd.j);
    double Ymid = o.yc[cm];  // collisions.cpp:687
    double g0 = ri[1] - rj[1], g1 = ri[2] - rj[2], g2 = ri[3] - rj[3];
    double gd = g0 * d0 + g1 * d1 + g2 * d2;
    double pv = pref * gd * Ymid * o.nden[J] * p.delta_t;  // collisions.cpp:691
    if (pv - omega[cd.c] > 0.0) omega[cd.c] = pv;          // :692-693
    bool a = pv / o.omax[cd.c] > mcpm_u01(w2.v[0]);        // :694
    if (new4 && !(gd > 0.0)) a = false;                      // :819
    if (a) {
      ri[1] = rnd<real>(ri[1] - gd * d0);
      ri[2] = rnd<real>(ri[2] - gd * d1);
      ri[3] = rnd<real>(ri[3] - gd * d2);
      rj[1] = rnd<real>(rj[1] + gd * d0);
      rj[2] = rnd<real>(rj[2] + gd * d1);
      rj[3] = rnd<real>(rj[3] + gd * d2);
      ++acc;
    }
    o.log_p[q] = (uint32_t)(order[q] >> 32);
    o.log_i[q] = cd.i;
    o.log_j[q] = cd.j;
    o.log_a[q] = a ? 1 : 0;
  }
  for (int c = 0; c < nc; ++c)
    if (omega[c] > o.omax[c]) o.omax[c] = omega[c];  // collisions.cpp:714-716
  o.last_cand = (int64_t)M;
  o.last_acc = acc;
  if (st) {
    st->candidates = (int64_t)M;
    st->valid = (int64_t)order.size();
    st->accepted = acc;
    st->rounds = 0;
    st->revisits = 0;
  }
}

// Collision_ESMC_new3 as the reference runs it (collisions.cpp:583-720), for
// ensemble comparisons with the canonical order: one std::mt19937 and the
// standard library's distributions; candidate counts floor(0.5 omega_max N_c +
// U); each candidate's cell drawn by z = uniform_int(0, total) against the
// running quota sums, which are then decremented from that cell on (the
// reference's own weighting, off-by-one included); i uniform in the cell,
// theta = 2 pi u, phi = acos(2 v - 1), dir = (cos theta sin phi,
// sin theta sin phi, cos phi); the partner cell and the rest as the canonical
// path.  The generator is seeded from (seed, step) so runs are repeatable.
template <typename real>
void orc_collide_mt(Oracle &o, int64_t step, mcpm_collision_stats *st) {
  const mcpm_params &p = o.p;
  const int nc = o.nc;
  std::seed_seq sq{(uint32_t)p.seed, (uint32_t)((uint64_t)p.seed >> 32), (uint32_t)step,
                   (uint32_t)((uint64_t)step >> 32)};
  std::mt19937 gen(sq);
  std::uniform_real_distribution<> dis_r(0.0, 1.0), dis_u(0.0, 1.0), dis_v(0.0, 1.0);
  std::vector<int> sum_M(nc);
  std::vector<double> omega(nc, 0.0);
  int M_cand = 0;
  for (int c = 0; c < nc; ++c) {
    int mtemp = (int)floor(0.5 * o.omax[c] * (1.0 * o.cnt[c]) + dis_r(gen));
    sum_M[c] = M_cand + mtemp;
    M_cand = sum_M[c];
  }
  int64_t valid = 0, acc = 0;
  const double pref = 4.0 * PI * p.sigma * p.sigma;
  for (int m = 0; m < M_cand; ++m) {
    if (sum_M[nc - 1] <= 0) continue;
    std::uniform_int_distribution<> d1(0, sum_M[nc - 1]);
    const int z = d1(gen);
    int I = -1;
    for (int c = 0; c < nc; ++c)
      if (z <= sum_M[c]) {
        I = c;
        break;
      }
    if (I < 0) continue;
    for (int c = I; c < nc; ++c) sum_M[c]--;
    if (o.cnt[I] <= 0) continue;
    std::uniform_int_distribution<> d2(0, o.cnt[I] - 1);
    const int64_t i = o.start[I] + d2(gen);
    const double u = dis_u(gen), v = dis_v(gen);
    const double theta = 2.0 * PI * u, phi = acos(2.0 * v - 1.0);
    const double d0 = cos(theta) * sin(phi), dd1 = sin(theta) * sin(phi), d2v = cos(phi);
    int J;
    if (!orc_partner(o, i, d0, dd1, &J) || o.cnt[J] <= 0) continue;
    std::uniform_int_distribution<> d3(0, o.cnt[J] - 1);
    const int64_t j = o.start[J] + d3(gen);
    if (i == j) continue;
    ++valid;
    double *ri = &o.rec[(size_t)o.S * i] + o.U - 1, *rj = &o.rec[(size_t)o.S * j] + o.U - 1;
    const double Ymid = o.yc[orc_mid(o, i, j)];
    const double g0 = ri[1] - rj[1], g1 = ri[2] - rj[2], g2 = ri[3] - rj[3];
    const double gd = g0 * d0 + g1 * dd1 + g2 * d2v;
    const double pv = pref * gd * Ymid * o.nden[J] * p.delta_t;
    if (pv - omega[I] > 0.0) omega[I] = pv;
    if (pv / o.omax[I] > dis_r(gen)) {
      ri[1] = rnd<real>(ri[1] - gd * d0);
      ri[2] = rnd<real>(ri[2] - gd * dd1);
      ri[3] = rnd<real>(ri[3] - gd * d2v);
      rj[1] = rnd<real>(rj[1] + gd * d0);
      rj[2] = rnd<real>(rj[2] + gd * dd1);
      rj[3] = rnd<real>(rj[3] + gd * d2v);
      ++acc;
    }
  }
  for (int c = 0; c < nc; ++c) o.omax[c] = std::max(omega[c], o.omax[c]);
  if (st) {
    st->candidates = M_cand;
    st->valid = valid;
    st->accepted = acc;
    st->rounds = 0;
    st->revisits = 0;
  }
}

// Thompson_3d (Math_Func.cpp:28-52): a main diagonal, b lower, c upper.
void thomas(std::vector<double> a, const std::vector<double> &b, const std::vector<double> &c,
            std::vector<double> &x, std::vector<double> rhs) {
  int n = (int)a.size();
  for (int i = 1; i < n; ++i) {
    double F = (1.0 / a[i - 1]) * b[i - 1];
    a[i] = a[i] - c[i - 1] * F;
    rhs[i] = rhs[i] - F * rhs[i - 1];
  }
  x.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double r = rhs[i];
    if (i != n - 1) r = r - c[i] * x[i + 1];
    x[i] = r / a[i];
  }
}

void gl50(double *x, double *w) {
  // 50-point Gauss-Legendre rule rounded to 8 decimals, the precision of the
  // reference's table (Main_Functions.cpp:3328-3345).
  const int n = 50;
  for (int i = 0; i < n; ++i) {
    double z = cos(PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    x[n - 1 - i] = nearbyint(z * 1e8) / 1e8;
    w[n - 1 - i] = nearbyint(2.0 / ((1.0 - z * z) * dp * dp) * 1e8) / 1e8;
  }
}

// Forward or backward substitution x_k = A_k x_{k-1} + B_k (x_{-1} = 0) in
// the association of a 1024-thread blocked scan: thread t composes the maps
// of elements [tE, tE + E) (E = ceil(n / 1024)) into one affine map, an
// inclusive Hillis-Steele scan over the threads composes the prefixes, and
// each thread then replays its elements from its predecessor's value.  With
// reverse the element order is n-1, ..., 0.  Mathematically this is the
// serial recurrence of Thompson_3d (Math_Func.cpp:28-52); the blocked
// association is what makes the result bit for bit reproducible against the
// HIP path.
void affine_scan(const std::vector<double> &A, const std::vector<double> &B,
                 std::vector<double> &out, int n, bool reverse) {
  const int nt = 1024;
  const int E = (n + nt - 1) / nt;
  std::vector<double> sA(nt), sB(nt);
  for (int t = 0; t < nt; ++t) {
    const int s0 = t * E, s1 = std::min(n, s0 + E);
    double a = 1.0, b = 0.0;
    for (int s = s0; s < s1; ++s) {
      const int k = reverse ? (n - 1 - s) : s;
      a = A[k] * a;
      b = A[k] * b + B[k];
    }
    sA[t] = a;
    sB[t] = b;
  }
  for (int off = 1; off < nt; off <<= 1) {
    std::vector<double> nA = sA, nB = sB;
    for (int t = off; t < nt; ++t) {
      nA[t] = sA[t] * sA[t - off];
      nB[t] = sA[t] * sB[t - off] + sB[t];
    }
    sA.swap(nA);
    sB.swap(nB);
  }
  out.assign(n, 0.0);
  for (int t = 0; t < nt; ++t) {
    const int s0 = t * E, s1 = std::min(n, s0 + E);
    double v = t == 0 ? 0.0 : sB[t - 1];
    for (int s = s0; s < s1; ++s) {
      const int k = reverse ? (n - 1 - s) : s;
      v = A[k] * v + B[k];
      out[k] = v;
    }
  }
}

// poisson_solver (Main_Functions.cpp:3285-3581) without the kick.  The face
// system (2 + (dy bst)^2) x_f - x_{f-1} - x_{f+1} = dy^2 n_f (:3375-3377)
// with Dirichlet rows x_0 = V_lo, x_nc = V_hi (:3448-3454), or, periodic,
// the cyclic system by Sherman-Morrison.  Thompson_3d's elimination is
// written as its LU factors (d_k = a_k - u_{k-1} l_{k-1} / d_{k-1}) and the
// two substitutions run through affine_scan.  The Dirichlet values sum the
// 50 x 50 Gauss-Legendre points (:3322-3442) as 1024 partial sums (point q
// to partial q mod 1024) folded pairwise (partial t adds t + off for
// off = 512, ..., 1).  exp is the shared deterministic mcpm_exp.
void orc_field(Oracle &o) {
  const mcpm_params &p = o.p;
  const int nc = o.nc;
  const double dy = o.dy;
  const bool periodic = p.boundary == MCPM_BC_PERIODIC;
  const int nf = periodic ? nc : nc + 1;
  std::vector<double> nj(nf), rhs(nf), x;
  for (int i = 0; i < nf; ++i) {
    if (periodic) nj[i] = (o.nden[i] + o.nden[(i + nc - 1) % nc]) / 2.0;
    else if (i == 0) nj[i] = o.nden[0];
    else if (i == nc) nj[i] = o.nden[nc - 1];
    else nj[i] = (o.nden[i] + o.nden[i - 1]) / 2.0;
    rhs[i] = nj[i] * (dy * dy);
  }
  // matrix: diagonal a, lo[k] = A[k+1][k], up[k] = A[k][k+1]
  const double diag = 2.0 + pow(dy * p.bst, 2.0);
  std::vector<double> a(nf, diag), lo(nf, -1.0), up(nf, -1.0);
  double vlast = 0.0;
  if (!periodic) {
    a[0] = 1.0;
    up[0] = 0.0;
    a[nf - 1] = 1.0;
    lo[nf - 2] = 0.0;
    double xi[50], wi[50], rr[50], dr[50], ph[50], dph[50], cph[50], sph[50];
    gl50(xi, wi);
    double bb = 3.0 * p.eqdist, aa = 0.0, cc = 0.0, dd = PI;
    for (int i = 0; i < 50; ++i) {
      rr[i] = (bb - aa) / 2.0 * xi[i] + (aa + bb) / 2.0;
      dr[i] = wi[i] * (bb - aa) / 2.0;
      ph[i] = (dd - cc) / 2.0 * xi[i] + (cc + dd) / 2.0;
      dph[i] = wi[i] * (dd - cc) / 2.0;
      cph[i] = cos(ph[i]);
      sph[i] = sin(ph[i]);
    }
    const int NT = 1024;
    std::vector<double> part[2] = {std::vector<double>(NT, 0.0), std::vector<double>(NT, 0.0)};
    for (int q = 0; q < 2500; ++q) {
      const int i = q / 50, ii = q % 50;
      const double radial = mcpm_exp(-p.bst * rr[i]) / 2.0 * rr[i] * dr[i];
      for (int side = 0; side < 2; ++side) {
        double y = (side == 0 ? 0.0 : (double)nc * dy) + rr[i] * cph[ii];
        double qq = mcpm_floor(y / dy);
        double n;
        if (qq < 0.0) n = o.nden[0];
        else if (qq > (double)(nc - 1)) n = o.nden[nc - 1];
        else n = o.nden[(int)qq];
        part[side][q % NT] += radial * n * sph[ii] * dph[ii];
      }
    }
    for (int off = NT / 2; off > 0; off >>= 1)
      for (int t = 0; t < off; ++t) {
        part[0][t] += part[0][t + off];
        part[1][t] += part[1][t + off];
      }
    rhs[0] = part[0][0];
    rhs[nf - 1] = part[1][0];
  } else {
    // cyclic corners A[0][n-1] = beta = -1, A[n-1][0] = alpha = -1
    double gamma = -a[0], alpha = -1.0, beta = -1.0;
    a[0] -= gamma;
    a[nf - 1] -= alpha * beta / gamma;
    vlast = beta / gamma;
  }
  std::vector<double> mL(nf, 0.0), d(nf), invd(nf), mU(nf, 0.0);
  d[0] = a[0];
  for (int k = 1; k < nf; ++k) {
    double l = lo[k - 1] / d[k - 1];
    d[k] = a[k] - up[k - 1] * l;
    mL[k] = -l;
  }
  for (int k = 0; k < nf; ++k) {
    invd[k] = 1.0 / d[k];
    mU[k] = (k < nf - 1) ? -up[k] / d[k] : 0.0;
  }
  std::vector<double> yv, tB(nf);
  affine_scan(mL, rhs, yv, nf, false);
  for (int k = 0; k < nf; ++k) tB[k] = yv[k] * invd[k];
  affine_scan(mU, tB, x, nf, true);
  if (periodic) {
    // Sherman-Morrison vector z = A'^-1 u, u = (gamma, 0, ..., 0, alpha),
    // by the serial substitutions
    double gamma = -(2.0 + pow(dy * p.bst, 2.0));
    std::vector<double> u(nf, 0.0), w(nf), z(nf);
    u[0] = gamma;
    u[nf - 1] = -1.0;
    w[0] = u[0];
    for (int k = 1; k < nf; ++k) w[k] = u[k] + mL[k] * w[k - 1];
    z[nf - 1] = w[nf - 1] * invd[nf - 1];
    for (int k = nf - 2; k >= 0; --k) z[k] = w[k] * invd[k] + mU[k] * z[k + 1];
    const double den = 1.0 + z[0] + vlast * z[nf - 1];
    const double corr = (x[0] + vlast * x[nf - 1]) / den;
    for (int k = 0; k < nf; ++k) x[k] = x[k] - corr * z[k];
  }
  double core = 1.0 - (p.sigma * p.bst + 1.0) * mcpm_exp(-p.bst * p.sigma);  // :3539-3543
  core = core / p.bst / p.bst;
  for (int i = 0; i < nf; ++i) x[i] = x[i] - nj[i] * core;
  for (int cc = 0; cc < nc; ++cc) {
    int f1 = periodic ? (cc + 1) % nc : cc + 1;
    double F = -(x[f1] - x[cc]) / dy;  // :3571-3572
    o.F2[cc] = F / p.mass * p.ast;
  }
  o.kick_pending = true;
}

// Cell density at a slab position for the quadratures: the cell's value,
// the boundary cell's value outside the slab (Main_Functions.cpp:3757-3784),
// the wrapped cell in a periodic box.
double orc_n_at(const Oracle &o, double y) {
  double q = floor(y / o.dy);
  if (o.p.boundary == MCPM_BC_PERIODIC) {
    long long k = (long long)q % o.nc;
    if (k < 0) k += o.nc;
    return o.nden[(size_t)k];
  }
  if (q < 0.0) return o.nden[0];
  if (q > (double)(o.nc - 1)) return o.nden[o.nc - 1];
  return o.nden[(size_t)q];
}

// vlasov_taking_the_integral (Main_Functions.cpp:3582-3942).
void orc_vlasov(Oracle &o) {
  const mcpm_params &p = o.p;
  const bool periodic = p.boundary == MCPM_BC_PERIODIC;
  const int nf = periodic ? o.nc : o.nc + 1;
  double xi[50], wi[50], rr[50], dr[50], ph[50], dph[50];
  gl50(xi, wi);
  const double b = 3.0 * p.eqdist, a = 1.0 * p.eqdist, c = 0.0, d = PI;
  for (int i = 0; i < 50; ++i) {
    rr[i] = (b - a) / 2.0 * xi[i] + (a + b) / 2.0;
    dr[i] = wi[i] * (b - a) / 2.0;
    ph[i] = (d - c) / 2.0 * xi[i] + (c + d) / 2.0;
    dph[i] = wi[i] * (d - c) / 2.0;
  }
  double rad[50];
  for (int i = 0; i < 50; ++i) {
    double q = p.sigma / rr[i];
    double q2 = q * q;
    rad[i] = -(-(q2 * q2 * q2)) * rr[i] * rr[i] * (2.0 * PI) * dr[i];
  }
  // each face's 2500 points summed in WgSum's order (point 50 i + ii)
  std::vector<double> V(nf, 0.0);
  for (int fc = 0; fc < nf; ++fc) {
    WgSum<1> ws;
    for (int q = 0; q < 2500; ++q) {
      const int i = q / 50, ii = q % 50;
      double y = (double)fc * o.dy + rr[i] * cos(ph[ii]);
      double v[1] = {-rad[i] * orc_n_at(o, y) * sin(ph[ii]) * dph[ii]};
      ws.add(q % 256, v);
    }
    double t[1];
    ws.total(t);
    V[fc] = t[0];
  }
  for (int num = 0; num < o.nc; ++num) {
    int f1 = periodic ? (num + 1) % o.nc : num + 1;
    double F = -(V[f1] - V[num]) / o.dy;
    o.F2[num] = F / p.mass * p.phi;
  }
  o.kick_pending = true;
}

// Gaussian kernel density estimate of cell_update (Main_Functions.cpp:
// 2378-2427): h = dy/2, cut-off 2 dy, neighbour cells k-2..k+2 (clamped
// at walls with the particle shifted by the clamped distance, wrapped in a
// periodic box).
void orc_kde(Oracle &o, std::vector<double> &ns, std::vector<double> &d2, std::vector<double> &d3) {
  const int nc = o.nc;
  const double r_cut = 2.0 * o.dy, h = r_cut / 4.0;
  const double sqrt_pi = sqrt(2.0 * PI);
  const int search = (int)mcpm_floor(r_cut / o.dy);
  const double mb = o.p.Fn;
  const double ih2 = 1.0 / h / h;
  const double s2h2 = ih2 * ih2;
  ns.assign(nc, 0.0);
  d2.assign(nc, 0.0);
  d3.assign(nc, 0.0);
  // each cell's sums in WgSum's order, lane i mod 256 for the i-th particle
  // of each neighbour cell in turn
  for (int k = 0; k < nc; ++k) {
    const double r1 = (k + 0.5) * o.dy;
    WgSum<3> ws;
    for (int kk = k - search; kk <= k + search; ++kk) {
      int kkk;
      double shift;
      if (o.p.boundary == MCPM_BC_PERIODIC) {
        kkk = ((kk % nc) + nc) % nc;
        shift = (double)(kk - kkk) * o.dy;
      } else {
        kkk = kk < 0 ? 0 : (kk > nc - 1 ? nc - 1 : kk);
        shift = (kk - kkk) * o.dy;
      }
      for (int i = 0; i < o.cnt[kkk]; ++i) {
        double r2 = o.rec[4 * (size_t)(o.start[kkk] + i)] + shift;
        double r1r2 = r2 - r1;
        double rr = fabs(r2 - r1);
        double v[3] = {0.0, 0.0, 0.0};
        if (rr < r_cut) {
          double ex = mcpm_exp(-rr * rr / h / h / 2.0);
          v[0] = mb / sqrt_pi / h * ex;
          double W = 1.0 / sqrt_pi / h * ex;
          v[1] = mb * W * r1r2 * (-1.0 / h / h);
          v[2] = mb * W * s2h2 * (3.0 * r1r2 - 1.0 / h / h * r1r2 * r1r2 * r1r2);
          ws.add(i % 256, v);
        } else {
          ws.touch(i % 256);
        }
      }
    }
    double t[3];
    ws.total(t);
    ns[k] = t[0];
    d2[k] = t[1];
    d3[k] = t[2];
  }
}

// Density-expansion force (long_range == 2, Main_Functions.cpp:1632-1653).
void orc_lr2(Oracle &o) {
  std::vector<double> ns, d2, d3;
  orc_kde(o, ns, d2, d3);
  const double s6 = o.p.sigma * pow(2.0, 1.0 / 6.0);
  const double pref = 4.0 * PI * (s6 * s6 * s6) * o.p.phi / o.p.mass * 2.0;
  for (int c = 0; c < o.nc; ++c) {
    bool wall_cell = o.p.boundary != MCPM_BC_PERIODIC && (c == 0 || c == o.nc - 1);
    o.F2[c] = wall_cell ? 0.0 : -(d2[c] * pref);
  }
  o.kick_pending = true;
}

void orc_moments(const Oracle &o, double *out) {
  // two-pass moments of cell_update (Main_Functions.cpp:2483-2593), each
  // sum in the workgroup order of WgSum
  for (int c = 0; c < o.nc; ++c) {
    double *m = out + (size_t)c * MCPM_NMOM;
    for (int k = 0; k < MCPM_NMOM; ++k) m[k] = 0.0;
    int N = o.cnt[c];
    double U[3] = {0, 0, 0};
    {
      WgSum<3> ws;
      for (int k = 0; k < N; ++k) {
        const double *r = &o.rec[(size_t)o.S * (o.start[c] + k)] + o.U - 1;
        double v[3] = {r[1], r[2], r[3]};
        ws.add(k % 256, v);
      }
      ws.total(U);
    }
    if (N > 0)
      for (int q = 0; q < 3; ++q) U[q] /= (double)N;
    // a[0..25] as below, a[26..28] the peculiar-velocity sums
    double a[29];
    {
      WgSum<29> ws;
      for (int k = 0; k < N; ++k) {
        const double *r = &o.rec[(size_t)o.S * (o.start[c] + k)] + o.U - 1;
        double p1 = r[1] - U[0], p2 = r[2] - U[1], p3 = r[3] - U[2];
        double e = p1 * p1 + p2 * p2 + p3 * p3;
        double v[29] = {p1 * p1, p1 * p2, p1 * p3, p2 * p2, p2 * p3, p3 * p3,
                        p1 * p1 * p1, p1 * p1 * p2, p1 * p1 * p3, p1 * p2 * p2,
                        p1 * p2 * p3, p1 * p3 * p3, p2 * p2 * p2, p2 * p2 * p3,
                        p2 * p3 * p3, p3 * p3 * p3,
                        p1 * p1 * e, p1 * p2 * e, p1 * p3 * e,
                        p2 * p2 * e, p2 * p3 * e, p3 * p3 * e,
                        p1 * e * e, p2 * e * e, p3 * e * e, e, p1, p2, p3};
        ws.add(k % 256, v);
      }
      ws.total(a);
    }
    m[0] = U[0]; m[1] = U[1]; m[2] = U[2];
    for (int q = 0; q < 26; ++q) m[3 + q] = N > 0 ? a[q] / (double)N : 0.0;
    m[29] = o.p.mass * m[28] / (3.0 * o.p.kb);  // T = m DM2 / (3 kB)
    m[30] = (double)N;
    for (int q = 0; q < 3; ++q) m[31 + q] = N > 0 ? a[26 + q] / (double)N : 0.0;
  }
}

// ------------------------------------------------------------------------
// Fokker-Planck model
// ------------------------------------------------------------------------
// mu(T) = mu0 (T/T0)^visp, scale_mu (Main_Functions.cpp:1970-1973).
double orc_pow(double x, double a) {
  if (a == 0.5) return sqrt(x);
  if (a == 1.0) return x;
  if (a == 0.0) return 1.0;
  return mcpm_exp(a * mcpm_log(x));
}

// Gauss_Jordan (Math_Func.cpp:165-232): forward elimination with scaled
// partial pivoting (row scale = largest magnitude in the row), then back
// substitution.  A is n x n row-major and is overwritten, as is rhs.
void orc_gauss_jordan(double *A, double *rhs, int n, double *x) {
  for (int i = 0; i < n - 1; ++i) {
    int best = 0;
    double bz = -1.0;
    for (int j = i; j < n; ++j) {
      const double *row = A + (size_t)j * n;
      int l = 0;
      double b = fabs(row[0]);
      for (int k = 0; k < n; ++k)
        if (fabs(b) < fabs(row[k])) {
          b = fabs(row[k]);
          l = k;
        }
      double z = fabs(row[i]) / fabs(row[l]);
      if (j == i || fabs(bz) < fabs(z)) {
        if (j == i || z > bz) {
          bz = z;
          best = j - i;
        }
      }
    }
    int k = i + best;
    if (k != i) {
      for (int j = 0; j < n; ++j) {
        double t = A[(size_t)i * n + j];
        A[(size_t)i * n + j] = A[(size_t)k * n + j];
        A[(size_t)k * n + j] = t;
      }
      double t = rhs[i];
      rhs[i] = rhs[k];
      rhs[k] = t;
    }
    for (int j = i + 1; j < n; ++j) {
      double F = (1.0 / A[(size_t)i * n + i]) * A[(size_t)j * n + i];
      for (int kk = 0; kk < n; ++kk) A[(size_t)j * n + kk] = A[(size_t)j * n + kk] - F * A[(size_t)i * n + kk];
      rhs[j] = rhs[j] - F * rhs[i];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double a = 0.0;
    for (int j = i + 1; j < n; ++j) a = a + A[(size_t)i * n + j] * x[j];
    x[i] = (rhs[i] - a) / A[(size_t)i * n + i];
  }
}

// Per-cell closures of cell_update and the FP model, written to
// clos[c * MCPM_NCLOS + k]: 0 Lambda, 1-6 c_ij, 7-9 gamma_i, 10-18 alpha,
// 19-21 beta, 22 gamma_st, 23-25 phi, 26 div U, 27 tau.
void orc_closures(Oracle &o, int64_t step) {
  const mcpm_params &p = o.p;
  const int nc = o.nc;
  std::vector<double> mom((size_t)nc * MCPM_NMOM);
  orc_moments(o, mom.data());
  o.mom = mom;
  o.clos.assign((size_t)nc * MCPM_NCLOS, 0.0);
  const bool dense = p.model == MCPM_MODEL_FP_DENSE;
  const bool fp = dense || p.model == MCPM_MODEL_FP_IDEAL;
  for (int c = 0; c < nc; ++c) {
    const double *m = &mom[(size_t)c * MCPM_NMOM];
    double *cl = &o.clos[(size_t)c * MCPM_NCLOS];
    const double *P = m + 3, *M3 = m + 9, *M4 = m + 19, *M5 = m + 25;
    const double DM2 = m[28], T = m[29], N = m[30];
    const double n = o.nden[c];
    if (!(N > 1e-14)) continue;
    const double Q[3] = {M3[0] + M3[3] + M3[5], M3[1] + M3[6] + M3[8], M3[2] + M3[7] + M3[9]};
    const double DM4 = M4[0] + M4[3] + M4[5];
    const double Yn = orc_y(n, p);
    // gamma_st (Main_Functions.cpp:2602-2605), Lambda (:2718-2728)
    cl[22] = -0.001 * (n * p.b * Yn / (p.kb * T / p.mass));
    double mu = p.mu * orc_pow(T / p.T0, p.visp);
    double tau = 2.0 * mu / (n * p.kb * T);
    double D3 = DM2 / 3.0;
    double L = (P[0] - D3) * ((P[3] - D3) * (P[5] - D3) - P[4] * P[4]);
    L = L - P[1] * (P[1] * (P[5] - D3) - P[4] * P[2]);
    L = L + P[2] * (P[1] * P[4] - P[2] * (P[3] - D3));
    double D2 = DM2 * DM2;
    L = -L / (tau * (D2 * D2));
    cl[0] = L;
    cl[27] = tau;
    if (!fp || o.cnt[c] < 2) continue;
    // alpha_calculation (Fokker_Planck.cpp:200-587), FP_dense after step 0
    if (dense && step > 0) {
      double A[144] = {0}, rhs[12], x[12] = {0};
      double trace = P[0] + P[3] + P[5];
      double dev[6] = {P[0], P[1], P[2], P[3], P[4], P[5]};
      dev[0] -= (1.0 / 3.0) * trace;
      dev[3] -= (1.0 / 3.0) * trace;
      dev[5] -= (1.0 / 3.0) * trace;
      double a = Q[0], b = Q[1], cc = Q[2];
      double d = DM4 - DM2 * DM2;
#define AA(i, j) A[(i) * 12 + (j)]
      AA(0, 0) = P[0]; AA(0, 1) = P[1]; AA(0, 2) = P[2]; AA(0, 9) = a;
      AA(1, 0) = P[1]; AA(1, 1) = P[3]; AA(1, 2) = P[4]; AA(1, 9) = a;
      AA(2, 0) = P[2]; AA(2, 1) = P[4]; AA(2, 2) = P[5]; AA(2, 9) = a;
      AA(3, 3) = P[0]; AA(3, 4) = P[1]; AA(3, 5) = P[2]; AA(3, 10) = b;
      AA(4, 3) = P[1]; AA(4, 4) = P[3]; AA(4, 5) = P[4]; AA(4, 10) = b;
      AA(5, 3) = P[2]; AA(5, 4) = P[4]; AA(5, 5) = P[5]; AA(5, 10) = b;
      AA(6, 6) = P[0]; AA(6, 7) = P[1]; AA(6, 8) = P[2]; AA(6, 11) = cc;
      AA(7, 6) = P[1]; AA(7, 7) = P[3]; AA(7, 8) = P[4]; AA(7, 11) = cc;
      AA(8, 6) = P[2]; AA(8, 7) = P[4]; AA(8, 8) = P[5]; AA(8, 11) = cc;
      AA(9, 0) = a; AA(9, 1) = b; AA(9, 2) = cc; AA(9, 9) = d;
      AA(10, 3) = a; AA(10, 4) = b; AA(10, 5) = cc; AA(10, 10) = d;
      AA(11, 6) = a; AA(11, 7) = b; AA(11, 8) = cc; AA(11, 11) = d;
#undef AA
      double nbY = n * p.b * Yn;
      double p_crr = 1.0 + nbY;
      double mu_crr = 1.0 + 2.0 * nbY / 5.0;
      double k_crr = 1.0 + 3.0 * nbY / 5.0;
      double s2 = p.sigma * p.sigma;
      double w = (4.0 / 9.0) * sqrt(PI) * Yn * (n * n) * (s2 * s2) * sqrt(p.mass * p.kb * T);
      // div U from the neighbouring cells' mean slab velocity (:459-468)
      double divU;
      auto Uy = [&](int cc2) { return mom[(size_t)cc2 * MCPM_NMOM + 1]; };
      if (p.boundary == MCPM_BC_PERIODIC) {
        divU = (Uy((c + 1) % nc) - Uy((c + nc - 1) % nc)) / o.dy / 2.0;
      } else if (c == 0) {
        divU = (Uy(1 % nc) - Uy(0)) / o.dy;
      } else if (c == nc - 1) {
        divU = (Uy(c) - Uy(c - 1)) / o.dy;
      } else {
        divU = (Uy(c + 1) - Uy(c - 1)) / o.dy / 2.0;
      }
      cl[26] = divU;
      double cmu = 3.0 * 1.002 / 5.0 * (nbY * nbY) / (1 + 2.0 * nbY / 5.0);
      double ck = 1.002 * (nbY * nbY) / (1 + 3.0 * nbY / 5.0) / 5.0;
      double gst = cl[22];
      double bulk = w * divU / (p.mass * n);
      rhs[0] = (mu_crr - 1.0) * dev[0] + (p_crr - 1.0) * trace / 3.0 - bulk + cmu * dev[0] - gst * M4[0];
      rhs[1] = (mu_crr - 1.0) * dev[1] + cmu * dev[1] - gst * M4[1];
      rhs[2] = (mu_crr - 1.0) * dev[2] + cmu * dev[2] - gst * M4[2];
      rhs[3] = (mu_crr - 1.0) * dev[1] + cmu * dev[1] - gst * M4[1];
      rhs[4] = (mu_crr - 1.0) * dev[3] + (p_crr - 1.0) * trace / 3.0 - bulk + cmu * dev[3] - gst * M4[3];
      rhs[5] = (mu_crr - 1.0) * dev[4] + cmu * dev[4] - gst * M4[4];
      rhs[6] = (mu_crr - 1.0) * dev[2] + cmu * dev[2] - gst * M4[2];
      rhs[7] = (mu_crr - 1.0) * dev[4] + cmu * dev[4] - gst * M4[4];
      rhs[8] = (mu_crr - 1.0) * dev[5] + (p_crr - 1.0) * trace / 3.0 - bulk + cmu * dev[5] - gst * M4[5];
      rhs[9] = (k_crr - 1.0) * a + ck * a - gst * M5[0] + gst * (DM2 * Q[0]);
      rhs[10] = (k_crr - 1.0) * b + ck * b - gst * M5[1] + gst * (DM2 * Q[1]);
      rhs[11] = (k_crr - 1.0) * cc + ck * cc - gst * M5[2] + gst * (DM2 * Q[2]);
      orc_gauss_jordan(A, rhs, 12, x);
      for (int q = 0; q < 9; ++q) cl[10 + q] = x[q];
      for (int q = 0; q < 3; ++q) cl[19 + q] = x[9 + q];
      if (x[0] != x[0]) o.err |= 2u;  // Fokker_Planck.cpp:557-571
    }
    // c_gamma_calculation (Fokker_Planck.cpp:590-804)
    {
      double A[81] = {0}, rhs[9], x[9] = {0};
#define AB(i, j) A[(i) * 9 + (j)]
      AB(0, 0) = 2.0 * P[0]; AB(0, 1) = 2.0 * P[1]; AB(0, 2) = 2.0 * P[2]; AB(0, 6) = 2.0 * Q[0];
      AB(1, 0) = P[1]; AB(1, 1) = P[0] + P[3]; AB(1, 2) = P[4]; AB(1, 3) = P[1]; AB(1, 4) = P[2];
      AB(1, 6) = Q[1]; AB(1, 7) = Q[0];
      AB(2, 0) = P[2]; AB(2, 1) = P[4]; AB(2, 2) = P[0] + P[5]; AB(2, 4) = P[1]; AB(2, 5) = P[2];
      AB(2, 6) = Q[2]; AB(2, 8) = Q[0];
      AB(3, 1) = 2.0 * P[1]; AB(3, 3) = 2.0 * P[3]; AB(3, 4) = 2.0 * P[4]; AB(3, 7) = 2.0 * Q[1];
      AB(4, 1) = P[2]; AB(4, 2) = P[1]; AB(4, 3) = P[4]; AB(4, 4) = P[3] + P[5]; AB(4, 5) = P[4];
      AB(4, 7) = Q[2]; AB(4, 8) = Q[1];
      AB(5, 2) = 2.0 * P[2]; AB(5, 4) = 2.0 * P[4]; AB(5, 5) = 2.0 * P[5]; AB(5, 8) = 2.0 * Q[2];
      AB(6, 0) = Q[0] + 2.0 * M3[0]; AB(6, 1) = Q[1] + 4.0 * M3[1]; AB(6, 2) = Q[2] + 4.0 * M3[2];
      AB(6, 3) = 2.0 * M3[3]; AB(6, 4) = 4.0 * M3[4]; AB(6, 5) = 2.0 * M3[5];
      AB(6, 6) = 2.0 * M4[0] - 2.0 * P[0] * DM2; AB(6, 7) = 2.0 * M4[1] - 2.0 * P[1] * DM2;
      AB(6, 8) = 2.0 * M4[2] - 2.0 * P[2] * DM2;
      AB(7, 0) = 2.0 * M3[1]; AB(7, 1) = Q[0] + 4.0 * M3[3]; AB(7, 2) = 4.0 * M3[4];
      AB(7, 3) = Q[1] + 2.0 * M3[6]; AB(7, 4) = Q[2] + 4.0 * M3[7]; AB(7, 5) = 2.0 * M3[8];
      AB(7, 6) = 2.0 * M4[1] - 2.0 * P[1] * DM2; AB(7, 7) = 2.0 * M4[3] - 2.0 * P[3] * DM2;
      AB(7, 8) = 2.0 * M4[4] - 2.0 * P[4] * DM2;
      AB(8, 0) = 2.0 * M3[2]; AB(8, 1) = 4.0 * M3[4]; AB(8, 2) = Q[0] + 4.0 * M3[5];
      AB(8, 3) = 2.0 * M3[7]; AB(8, 4) = Q[1] + 4.0 * M3[8]; AB(8, 5) = Q[2] + 2.0 * M3[9];
      AB(8, 6) = 2.0 * M4[2] - 2.0 * P[2] * DM2; AB(8, 7) = 2.0 * M4[4] - 2.0 * P[4] * DM2;
      AB(8, 8) = 2.0 * M4[5] - 2.0 * P[5] * DM2;
      AB(6, 6) += DM4 - DM2 * DM2;
      AB(7, 7) += DM4 - DM2 * DM2;
      AB(8, 8) += DM4 - DM2 * DM2;
#undef AB
      const double Pr = 2.0 / 3.0;
      const double Yc = dense ? Yn : 1.0;
      for (int q = 0; q < 6; ++q) rhs[q] = -2.0 * L * M4[q];
      for (int q = 0; q < 3; ++q) rhs[6 + q] = Yc * (3.0 / tau - 2.0 * Pr / tau) * Q[q];
      rhs[6] += L * (-3.0 * M5[0] + DM2 * Q[0] + 2.0 * (P[0] * Q[0] + P[1] * Q[1] + P[2] * Q[2]));
      rhs[7] += L * (-3.0 * M5[1] + DM2 * Q[1] + 2.0 * (P[1] * Q[0] + P[3] * Q[1] + P[4] * Q[2]));
      rhs[8] += L * (-3.0 * M5[2] + DM2 * Q[2] + 2.0 * (P[2] * Q[0] + P[4] * Q[1] + P[5] * Q[2]));
      orc_gauss_jordan(A, rhs, 9, x);
      for (int q = 0; q < 6; ++q) cl[1 + q] = x[q];
      for (int q = 0; q < 3; ++q) cl[7 + q] = x[6 + q];
    }
    // phi = alpha . Mp (Main_Functions.cpp:2734-2746)
    const double *al = cl + 10, *mp = m + 31;
    cl[23] = al[0] * mp[0] + al[1] * mp[1] + al[2] * mp[2];
    cl[24] = al[3] * mp[0] + al[4] * mp[1] + al[5] * mp[2];
    cl[25] = al[6] * mp[0] + al[7] * mp[1] + al[8] * mp[2];
  }
  o.have_closures = true;
}

// random_generator (Fokker_Planck.cpp:20-145) with Philox normals: per cell
// subtract the mean, divide by the rms (cells with more than one particle).
void orc_noise(const Oracle &o, int64_t step, std::vector<double> &xi) {
  const mcpm_params &p = o.p;
  uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
  xi.assign(3 * (size_t)o.n, 0.0);
  for (int64_t k = 0; k < o.n; ++k) {
    mcpm_u32x4 w = mcpm_philox((uint32_t)k, 0u, slo,
                               mcpm_ctr3(MCPM_TAG_FPNOISE, shi, (uint32_t)((uint64_t)k >> 32)),
                               key0, key1);
    double z0, z1, z2, z3;
    mcpm_normal2(w.v[0], w.v[1], &z0, &z1);
    mcpm_normal2(w.v[2], w.v[3], &z2, &z3);
    xi[3 * k] = z0;
    xi[3 * k + 1] = z1;
    xi[3 * k + 2] = z2;
  }
  // per-cell sums in WgSum's order
  for (int c = 0; c < o.nc; ++c) {
    int N = o.cnt[c];
    if (N < 2) continue;
    double s[3], s2[3];
    {
      WgSum<3> ws;
      for (int k = 0; k < N; ++k) {
        const double *x = &xi[3 * (size_t)(o.start[c] + k)];
        double v[3] = {x[0], x[1], x[2]};
        ws.add(k % 256, v);
      }
      ws.total(s);
    }
    for (int q = 0; q < 3; ++q) s[q] = s[q] / (1.0 * N);
    {
      WgSum<3> ws;
      for (int k = 0; k < N; ++k) {
        double *x = &xi[3 * (size_t)(o.start[c] + k)];
        double v[3];
        for (int q = 0; q < 3; ++q) {
          x[q] = x[q] - s[q];
          v[q] = x[q] * x[q];
        }
        ws.add(k % 256, v);
      }
      ws.total(s2);
    }
    for (int q = 0; q < 3; ++q) s2[q] = sqrt(s2[q] / (1.0 * N));
    for (int k = 0; k < N; ++k)
      for (int q = 0; q < 3; ++q) xi[3 * (size_t)(o.start[c] + k) + q] /= s2[q];
  }
}

// Velocity_FP_parallel (Fokker_Planck.cpp:1065-1182); with model FP_LINEAR
// Velocity_FP_linear (Fokker_Planck.cpp:1186-1287): tau = 2 mu / (n kB T)
// (Fokker_Planck.cpp:1237-1238, Y = 1), no drift, the same rescale.
template <typename real>
void orc_fp_velocity(Oracle &o, int64_t step) {
  const mcpm_params &p = o.p;
  const bool dense = p.model == MCPM_MODEL_FP_DENSE;
  const bool linear = p.model == MCPM_MODEL_FP_LINEAR;
  std::vector<double> xi;
  orc_noise(o, step, xi);
  for (int c = 0; c < o.nc; ++c) {
    const int N = o.cnt[c];
    if (N == 0) continue;
    const double *m = &o.mom[(size_t)c * MCPM_NMOM];
    const double *cl = &o.clos[(size_t)c * MCPM_NCLOS];
    const double U[3] = {m[0], m[1], m[2]};
    const double T = m[29];
    if (N < 2) {
      for (int k = 0; k < N; ++k) {
        double *r = &o.rec[4 * (size_t)(o.start[c] + k)];
        r[1] = rnd<real>(U[0]);
        r[2] = rnd<real>(U[1]);
        r[3] = rnd<real>(U[2]);
      }
      continue;
    }
    double tau = cl[27];
    if (dense) tau = tau / o.yc[c];
    // exp through the shared deterministic routine, as on the GPU
    const double A1 = mcpm_exp(-p.delta_t / tau);
    const double A3 = sqrt((1.0 - mcpm_exp(-2.0 * p.delta_t / tau)) * p.kb * T / (1.0 * p.mass));
    const double A2 = (1.0 - mcpm_exp(-p.delta_t / tau)) * tau;
    const double L = cl[0];
    const double *cc = cl + 1, *g = cl + 7;
    const double Q[3] = {m[9] + m[12] + m[14], m[10] + m[15] + m[17], m[11] + m[16] + m[18]};
    const double e0 = 3.0 * p.kb * T / p.mass;
    // the rescale's two sums in WgSum's order
    WgSum<2> ws;
    std::vector<double> mp(3 * (size_t)N);
    for (int k = 0; k < N; ++k) {
      const double *r = &o.rec[4 * (size_t)(o.start[c] + k)];
      double a = r[1] - U[0], b = r[2] - U[1], d = r[3] - U[2];
      double e = a * a + b * b + d * d;
      double d0 = cc[0] * a + cc[1] * b + cc[2] * d;
      d0 += g[0] * (e - e0);
      d0 += L * (a * e - Q[0]);
      double d1 = cc[1] * a + cc[3] * b + cc[4] * d;
      d1 += g[1] * (e - e0);
      d1 += L * (b * e - Q[1]);
      double d2 = cc[2] * a + cc[4] * b + cc[5] * d;
      d2 += g[2] * (e - e0);
      d2 += L * (d * e - Q[2]);
      const double *x = &xi[3 * (size_t)(o.start[c] + k)];
      double n0 = A1 * a + A2 * d0 + A3 * x[0];
      double n1 = A1 * b + A2 * d1 + A3 * x[1];
      double n2 = A1 * d + A2 * d2 + A3 * x[2];
      if (linear) {
        // Velocity_FP_linear: Mp = A1 Mp + A3 xi (Fokker_Planck.cpp:1262-1264)
        n0 = A1 * a + A3 * x[0];
        n1 = A1 * b + A3 * x[1];
        n2 = A1 * d + A3 * x[2];
      }
      if (n1 != n1) o.err |= 2u;  // Fokker_Planck.cpp:1137-1149
      mp[3 * k] = n0;
      mp[3 * k + 1] = n1;
      mp[3 * k + 2] = n2;
      double fr[2] = {e, n0 * n0 + n1 * n1 + n2 * n2};
      ws.add(k % 256, fr);
    }
    double fr[2];
    ws.total(fr);
    double s = sqrt(fabs(fr[0] / fr[1]));
    for (int k = 0; k < N; ++k) {
      double *r = &o.rec[4 * (size_t)(o.start[c] + k)];
      r[1] = rnd<real>(mp[3 * k] * s + U[0]);
      r[2] = rnd<real>(mp[3 * k + 1] * s + U[1]);
      r[3] = rnd<real>(mp[3 * k + 2] * s + U[2]);
    }
  }
}

// The CBA displacement (collisions.cpp:2221-2238): the pair moves by -+sigma
// along the change of its relative velocity, read back from the stored
// velocities; x1 is the plane's x, x2 the slab or plane y.  A pair whose
// stored relative velocity did not change stays in place (the reference
// would divide 0 by 0).
template <typename real>
void orc_cba_displace(Oracle &o, uint32_t i, uint32_t j, const double pre[3]) {
  double *pi = &o.rec[(size_t)o.S * i], *pj = &o.rec[(size_t)o.S * j];
  const double *ui = pi + o.U - 1, *uj = pj + o.U - 1;
  double post[3] = {ui[1] - uj[1], ui[2] - uj[2], ui[3] - uj[3]};
  double e0 = post[0] - pre[0], e1 = post[1] - pre[1], e2 = post[2] - pre[2];
  double nrm = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  if (!(nrm > 0.0)) return;
  double d0 = e0 * (o.p.sigma / nrm), d1 = e1 * (o.p.sigma / nrm);
  if (o.dims == 2) {
    pi[0] = rnd<real>(pi[0] + d0);
    pj[0] = rnd<real>(pj[0] - d0);
    pi[1] = rnd<real>(pi[1] + d1);
    pj[1] = rnd<real>(pj[1] - d1);
  } else {
    pi[0] = rnd<real>(pi[0] + d1);
    pj[0] = rnd<real>(pj[0] - d1);
  }
}

// Collision_DSMC_VHS (collisions.cpp:2063-2141): pairs within a cell,
// candidates in index order, VHS cross-section, isotropic scattering;
// crm <- max(1.1 crm_observed, 0.9 crm) per cell with two or more
// particles.  variant DSMC_HS (collisions.cpp:1989-2061): sigma_T = Vr and
// crm updated in every cell; CBA_VHS (collisions.cpp:2144-2248): the count
// times Y(n_c), the displacement, crm updated in every cell.  Cells with
// fewer than two particles draw nothing (the reference's pair loop would not
// end for one particle).  Candidate k (global, cell order) of cell c draws
// from Philox blocks keyed (k, c).
template <typename real>
void orc_dsmc(Oracle &o, int64_t step, mcpm_collision_stats *st, int variant) {
  const mcpm_params &p = o.p;
  uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
  int64_t M = 0, valid = 0, acc = 0;
  o.log_p.clear();
  o.log_i.clear();
  o.log_j.clear();
  o.log_a.clear();
  uint32_t kglob = 0;
  for (int c = 0; c < o.nc; ++c) {
    const int Nc = o.cnt[c];
    if (Nc < 2) {
      if (variant != MCPM_MODEL_DSMC_VHS) o.omax[c] = std::max(0.0, 0.9 * o.omax[c]);
      continue;
    }
    double Mc = PI * p.sigma * p.sigma * ((double)Nc * o.nden[c]) * o.omax[c] * p.delta_t / 2.0;
    if (variant == MCPM_MODEL_CBA_VHS) Mc = Mc * o.yc[c];
    mcpm_u32x4 w0 = mcpm_philox((uint32_t)c, 0u, slo, mcpm_ctr3(MCPM_TAG_MCAND, shi, 0u), key0, key1);
    double mf = mcpm_floor(Mc + mcpm_u01(w0.v[0]));
    uint32_t m = mf > 0.0 ? (uint32_t)mf : 0u;
    M += m;
    double crm = 0.0;
    for (uint32_t t = 0; t < m; ++t) {
      const uint32_t k = kglob + t;
      mcpm_u32x4 w = mcpm_philox(k, (uint32_t)c, slo, mcpm_ctr3(MCPM_TAG_DSMC, shi, 0u), key0, key1);
      uint32_t li = mcpm_uint_below(w.v[0], (uint32_t)Nc);
      uint32_t lj = mcpm_uint_below(w.v[1], (uint32_t)(Nc - 1));
      if (lj >= li) lj += 1;
      uint32_t i = (uint32_t)o.start[c] + li, j = (uint32_t)o.start[c] + lj;
      double *ri = &o.rec[(size_t)o.S * i] + o.U - 1, *rj = &o.rec[(size_t)o.S * j] + o.U - 1;
      double g0 = ri[1] - rj[1], g1 = ri[2] - rj[2], g2 = ri[3] - rj[3];
      double Vr = sqrt(g0 * g0 + g1 * g1 + g2 * g2);
      double sig = variant == MCPM_MODEL_DSMC_HS
                       ? Vr
                       : orc_pow(Vr, 2.0 - 2.0 * p.visp) * orc_pow(p.crref, 2.0 * p.visp - 1.0);
      if (sig - crm > 0.0) crm = sig;
      bool a = sig / o.omax[c] > mcpm_u01(w.v[2]);
      if (a) {
        double ct = 2.0 * mcpm_u01(w.v[3]) - 1.0;
        double stt = sqrt(1.0 - ct * ct);
        mcpm_u32x4 w2 = mcpm_philox(k, (uint32_t)c, slo, mcpm_ctr3(MCPM_TAG_DSMC, shi, 1u), key0, key1);
        double sp, cp;
        mcpm_sincos2pi(mcpm_u01(w2.v[0]), &sp, &cp);
        double v0 = Vr * ct, v1 = Vr * stt * cp, v2 = Vr * stt * sp;
        double c0 = 0.5 * (ri[1] + rj[1]), c1 = 0.5 * (ri[2] + rj[2]), c2 = 0.5 * (ri[3] + rj[3]);
        ri[1] = rnd<real>(c0 + 0.5 * v0);
        ri[2] = rnd<real>(c1 + 0.5 * v1);
        ri[3] = rnd<real>(c2 + 0.5 * v2);
        rj[1] = rnd<real>(c0 - 0.5 * v0);
        rj[2] = rnd<real>(c1 - 0.5 * v1);
        rj[3] = rnd<real>(c2 - 0.5 * v2);
        if (variant == MCPM_MODEL_CBA_VHS) {
          const double pre[3] = {g0, g1, g2};
          orc_cba_displace<real>(o, i, j, pre);
        }
        ++acc;
      }
      ++valid;
      o.log_p.push_back(k);
      o.log_i.push_back(i);
      o.log_j.push_back(j);
      o.log_a.push_back(a ? 1 : 0);
    }
    kglob += m;
    o.omax[c] = std::max(1.1 * crm, 0.9 * o.omax[c]);
  }
  if (st) {
    st->candidates = M;
    st->valid = valid;
    st->accepted = acc;
    st->rounds = 0;
    st->revisits = 0;
  }
}

// Collision_CBA_HS (collisions.cpp:2251-2346), cell after cell, candidate
// after candidate: Vr_max = sqrt(pi kB T / m), M = Fn N^2 pi sigma^2 Vr_max
// dt / (2 V_c) Y with Y = (1 - b rho / 8) / (1 - b rho / 4)^3, rho = n m; the
// loop bound floor(M) is re-read after every candidate, and an accepted pair
// faster than Vr_max (by more than 1e-14) raises Vr_max and M.  Candidate t
// of cell c draws (i, j, r, alpha2) from one Philox block keyed (t, c), i and
// j with replacement; r is both the acceptance draw and alpha1.
template <typename real>
void orc_cba_hs(Oracle &o, int64_t step, mcpm_collision_stats *st) {
  const mcpm_params &p = o.p;
  uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
  std::vector<double> mom((size_t)o.nc * MCPM_NMOM);
  orc_moments(o, mom.data());
  int64_t M = 0, acc = 0;
  for (int c = 0; c < o.nc; ++c) {
    const int Nc = o.cnt[c];
    if (Nc < 2) continue;
    const double T = mom[(size_t)c * MCPM_NMOM + 29];
    const double rho = o.nden[c] * p.mass;
    double vmax = sqrt(PI * p.kb * T / p.mass);
    const double Vc = p.cell_volume;
    double Mc = p.Fn * Nc * Nc * PI * p.sigma * p.sigma * vmax * p.delta_t / (2.0 * Vc);
    const double q = 1.0 - p.b * rho / 4.0;
    const double Y = (1.0 - p.b * rho / 8.0) / (q * q * q);
    Mc = Mc * Y;
    for (uint32_t t = 0; t < mcpm_floor(Mc); ++t) {
      mcpm_u32x4 w = mcpm_philox(t, (uint32_t)c, slo, mcpm_ctr3(MCPM_TAG_CBA, shi, 0u), key0, key1);
      uint32_t i = (uint32_t)o.start[c] + mcpm_uint_below(w.v[0], (uint32_t)Nc);
      uint32_t j = (uint32_t)o.start[c] + mcpm_uint_below(w.v[1], (uint32_t)Nc);
      double r = mcpm_u01(w.v[2]);
      ++M;
      double *ri = &o.rec[(size_t)o.S * i] + o.U - 1, *rj = &o.rec[(size_t)o.S * j] + o.U - 1;
      double Vr = sqrt((ri[1] - rj[1]) * (ri[1] - rj[1]) + (ri[2] - rj[2]) * (ri[2] - rj[2]) +
                       (ri[3] - rj[3]) * (ri[3] - rj[3]));
      if (!(Vr / vmax > r)) continue;
      if (Vr - vmax > 1e-14) {
        vmax = Vr;
        Mc = p.Fn * Nc * Nc * PI * p.sigma * p.sigma * vmax * p.delta_t / (2.0 * Vc);
        Mc = Mc * Y;
      }
      const double pre[3] = {ri[1] - rj[1], ri[2] - rj[2], ri[3] - rj[3]};
      double alpha2 = mcpm_u01(w.v[3]);
      double qq = 1.0 - 2 * r;
      double ct = qq, stt = sqrt(1.0 - qq * qq);
      double sp, cp;
      mcpm_sincos2pi(alpha2, &sp, &cp);
      double v2 = Vr * stt * cp, v1 = Vr * stt * sp, v0 = Vr * ct;
      double c0 = 0.5 * (ri[1] + rj[1]), c1 = 0.5 * (ri[2] + rj[2]), c2 = 0.5 * (ri[3] + rj[3]);
      ri[1] = rnd<real>(c0 + 0.5 * v0);
      ri[2] = rnd<real>(c1 + 0.5 * v1);
      ri[3] = rnd<real>(c2 + 0.5 * v2);
      rj[1] = rnd<real>(c0 - 0.5 * v0);
      rj[2] = rnd<real>(c1 - 0.5 * v1);
      rj[3] = rnd<real>(c2 - 0.5 * v2);
      orc_cba_displace<real>(o, i, j, pre);
      ++acc;
    }
  }
  if (st) {
    st->candidates = M;
    st->valid = M;
    st->accepted = acc;
    st->rounds = 0;
    st->revisits = 0;
  }
}

// cell_poly_velocity (Main_Functions.cpp:1918-1968): per cell the count and
// the raw velocity sums of orders 1-3 and the fourth-order products with
// |U|^2.  pow(U, k) is written as k-fold products, the third-order terms
// U1^m U2^n U3^l (m = 0..3, n = 0..3-m, l = 3-m-n, :1940-1948) as the
// products below; each sum over the cell in WgSum's order.
void orc_sample(Oracle &o) {
  for (int c = 0; c < o.nc; ++c) {
    double *s = &o.samples[(size_t)c * MCPM_NSAMP];
    const int N = o.cnt[c];
    WgSum<25> ws;
    for (int k = 0; k < N; ++k) {
      const double *r = &o.rec[(size_t)o.S * (o.start[c] + k)] + o.U - 1;
      const double u = r[1], v = r[2], w = r[3];
      const double e = u * u + v * v + w * w;
      double t[25] = {u, v, w,
                      u * u, u * v, u * w, v * v, v * w, w * w,
                      w * w * w, v * w * w, v * v * w, v * v * v,
                      u * w * w, u * v * w, u * v * v,
                      u * u * w, u * u * v, u * u * u,
                      u * u * e, u * v * e, u * w * e, v * v * e, v * w * e, w * w * e};
      ws.add(k % 256, t);
    }
    double a[25];
    ws.total(a);
    s[0] += (double)N;
    for (int q = 0; q < 25; ++q) s[1 + q] += a[q];
  }
}

// ------------------------------------------------------------------------
// The plane's mean field (FEM.cpp)

// Derivatives of the bilinear shape functions on the unit cell (FEM.cpp:22-49).
double dn_xi(int i, double xi, double eta) {
  (void)xi;
  if (i == 0) return eta - 1.0;
  if (i == 1) return 1.0 - eta;
  if (i == 2) return -eta;
  return eta;
}
double dn_eta(int i, double xi, double eta) {
  (void)eta;
  if (i == 0) return xi - 1.0;
  if (i == 1) return -xi;
  if (i == 2) return -xi + 1;
  return xi;
}

// Full-precision Gauss-Legendre rule (nodes ascending), the reference's
// 100-point table of compute_Vlasov_2D (FEM.cpp:656-657).
void gl_full(int n, double *x, double *w) {
  // Newton iteration in long double, so the double nodes and weights are
  // correctly rounded (the weights are sensitive to the last bits of the
  // nodes near the ends of the interval)
  for (int i = 0; i < n; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double pm1 = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p0 = 1.0L, p1 = z;
      for (int k = 2; k <= n; ++k) {
        long double p2 = ((2.0L * k - 1.0L) * z * p1 - (k - 1.0L) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      long double dp = n * (z * p1 - p0) / (z * z - 1.0L);
      long double dz = p1 / dp;
      z -= dz;
      pm1 = p0;
      if (fabsl(dz) < 1e-19L) break;
    }
    x[n - 1 - i] = (double)z;
    w[n - 1 - i] = (double)(2.0L * ((1.0L - z) * (1.0L + z)) / ((n * pm1) * (n * pm1)));
  }
}

// Geometry of cell (j, k) as make_cells computes it (myIO.cpp:1455-1479).
struct CellGeo {
  double x_n[4], y_n[4], dx, dy;
  int id[4];
};
CellGeo cell_geo(const Oracle &o, int j, int k) {
  CellGeo g;
  const double d0 = (double)j * o.dx, d1 = (double)(j + 1) * o.dx;
  const double d2 = (double)k * o.dy, d3 = (double)(k + 1) * o.dy;
  g.dx = d1 - d0;
  g.dy = d3 - d2;
  const double cx = 0.5 * (d0 + d1), cy = 0.5 * (d2 + d3);
  g.x_n[0] = cx - g.dx / 2.0; g.x_n[1] = cx + g.dx / 2.0;
  g.x_n[2] = cx - g.dx / 2.0; g.x_n[3] = cx + g.dx / 2.0;
  g.y_n[0] = cy - g.dy / 2.0; g.y_n[1] = cy - g.dy / 2.0;
  g.y_n[2] = cy + g.dy / 2.0; g.y_n[3] = cy + g.dy / 2.0;
  g.id[0] = k * (o.nx + 1) + j;
  g.id[1] = k * (o.nx + 1) + j + 1;
  g.id[2] = (k + 1) * (o.nx + 1) + j;
  g.id[3] = (k + 1) * (o.nx + 1) + j + 1;
  return g;
}

// Nodal density deposit of cell_update (Main_Functions.cpp:2268-2364): each
// particle credits Fn/V to the corner of its own cell nearest to it and to
// the same node's copies in the three neighbouring cells; afterwards the
// copies on each wall are doubled.  Counts are kept as integers and scaled
// once (the reference adds Fn/V one particle at a time).  Periodic axes
// (an extension) wrap the neighbours and skip the doubling.
void orc_deposit(Oracle &o) {
  const int nx = o.nx, ny = o.ny;
  const bool px = o.p.boundary_x == MCPM_BC_PERIODIC, py = o.p.boundary == MCPM_BC_PERIODIC;
  std::vector<int64_t> k4((size_t)4 * nx * ny, 0);
  auto credit = [&](int j, int k, int corner) {
    if (px) j = (j + nx) % nx;
    else if (j < 0 || j > nx - 1) return;
    if (py) k = (k + ny) % ny;
    else if (k < 0 || k > ny - 1) return;
    k4[(size_t)4 * (k * nx + j) + corner]++;
  };
  for (int64_t i = 0; i < o.n; ++i) {
    const int num = orc_key(o, i);
    const int j = num % nx, k = num / nx;
    const double *r = &o.rec[(size_t)o.S * i];
    const double ccx = 0.5 * ((double)j * o.dx + (double)(j + 1) * o.dx);
    const double ccy = 0.5 * ((double)k * o.dy + (double)(k + 1) * o.dy);
    if (r[0] < ccx) {
      if (r[1] < ccy) {
        credit(j, k, 0); credit(j - 1, k, 1); credit(j, k - 1, 2); credit(j - 1, k - 1, 3);
      } else {
        credit(j, k, 2); credit(j - 1, k, 3); credit(j, k + 1, 0); credit(j - 1, k + 1, 1);
      }
    } else {
      if (r[1] < ccy) {
        credit(j, k, 1); credit(j + 1, k, 0); credit(j, k - 1, 3); credit(j + 1, k - 1, 2);
      } else {
        credit(j, k, 3); credit(j + 1, k, 2); credit(j, k + 1, 1); credit(j + 1, k + 1, 0);
      }
    }
  }
  const double fnv = o.p.Fn / o.p.cell_volume;
  o.n_n.assign((size_t)4 * nx * ny, 0.0);
  for (size_t q = 0; q < k4.size(); ++q) o.n_n[q] = (double)k4[q] * fnv;
  if (!py) {
    for (int jj = 0; jj < nx; ++jj) {
      o.n_n[4 * (size_t)jj + 0] *= 2.0;
      o.n_n[4 * (size_t)jj + 1] *= 2.0;
      o.n_n[4 * (size_t)((ny - 1) * nx + jj) + 2] *= 2.0;
      o.n_n[4 * (size_t)((ny - 1) * nx + jj) + 3] *= 2.0;
    }
  }
  if (!px) {
    for (int kk = 0; kk < ny; ++kk) {
      o.n_n[4 * (size_t)(kk * nx) + 0] *= 2.0;
      o.n_n[4 * (size_t)(kk * nx) + 2] *= 2.0;
      o.n_n[4 * (size_t)(kk * nx + nx - 1) + 1] *= 2.0;
      o.n_n[4 * (size_t)(kk * nx + nx - 1) + 3] *= 2.0;
    }
  }
}

// The node value seen by node (i, j): corner k of its owner cell (ii, jj).
void owner(const Oracle &o, int i, int j, int *ii, int *jj, int *k) {
  *ii = i < o.nx ? i : i - 1;
  *jj = j < o.ny ? j : j - 1;
  *k = (i == o.nx ? 1 : 0) + (j == o.ny ? 2 : 0);
}

std::vector<double> node_values(const Oracle &o) {
  std::vector<double> v((size_t)(o.nx + 1) * (o.ny + 1));
  for (int j = 0; j <= o.ny; ++j)
    for (int i = 0; i <= o.nx; ++i) {
      int ii, jj, k;
      owner(o, i, j, &ii, &jj, &k);
      v[(size_t)j * (o.nx + 1) + i] = o.n_n[4 * (size_t)(jj * o.nx + ii) + k];
    }
  return v;
}

// Quadrature sum around (xx, yy) (Compute_SP_using_Green_Function,
// FEM.cpp:196-247, and vlasov_taking_the_integral_2D, FEM.cpp:157-195):
// rad[i] is the radial factor with its weight; cells clamp to the mesh.
double node_integral(const Oracle &o, double xx, double yy, int nq, const double *rr,
                     const double *rad, const double *phi, const double *dphi,
                     const double *theta, const double *dtheta) {
  const bool px = o.p.boundary_x == MCPM_BC_PERIODIC, py = o.p.boundary == MCPM_BC_PERIODIC;
  // cos(phi), sin(phi), sin(theta) tabulated: the same libm values the
  // reference evaluates inside the loops
  std::vector<double> cph(nq), sph(nq), sth(nq);
  for (int k = 0; k < nq; ++k) {
    cph[k] = cos(phi[k]);
    sph[k] = sin(phi[k]);
    sth[k] = sin(theta[k]);
  }
  double V = 0.0;
  for (int i = 0; i < nq; ++i) {
    for (int ii = 0; ii < nq; ++ii) {
      for (int iii = 0; iii < nq; ++iii) {
        double x = xx + rr[i] * cph[ii] * sth[iii];
        double y = yy + rr[i] * sph[ii] * sth[iii];
        int j, k;
        if (px) {
          if (x < 0.0) x += o.p.length_x;
          else if (x >= o.p.length_x) x -= o.p.length_x;
          j = orc_cell(x, o.dx, o.nx);
        } else {
          j = (int)floor(x / o.dx);
          if (j < 0) j = 0;
          else if (j > o.nx - 1) j = o.nx - 1;
        }
        if (py) {
          if (y < 0.0) y += o.p.length;
          else if (y >= o.p.length) y -= o.p.length;
          k = orc_cell(y, o.dy, o.ny);
        } else {
          k = (int)floor(y / o.dy);
          if (k < 0) k = 0;
          else if (k > o.ny - 1) k = o.ny - 1;
        }
        const double n = o.nden[k * o.nx + j];
        V += rad[i] * n * dphi[ii] * sth[iii] * dtheta[iii];
      }
    }
  }
  return V;
}

struct Quad {
  int nq;
  std::vector<double> rr, rad, phi, dphi, theta, dtheta;
};

// r in [a, 3 sigma], azimuth in [0, 2 pi], polar angle in [0, pi].
Quad make_quad(const Oracle &o, bool yukawa) {
  Quad q;
  q.nq = yukawa ? 50 : 100;
  std::vector<double> xi(q.nq), wi(q.nq);
  if (yukawa) gl50(xi.data(), wi.data());
  else gl_full(q.nq, xi.data(), wi.data());
  const double aa = yukawa ? 0.0 : o.p.sigma, bb = 3.0 * o.p.sigma;
  const double cc = 0.0, dd = 2 * PI, ee = 0.0, ff = PI;
  for (int i = 0; i < q.nq; ++i) {
    const double r = (bb - aa) / 2.0 * xi[i] + (aa + bb) / 2.0;
    const double dr = wi[i] * (bb - aa) / 2.0;
    q.rr.push_back(r);
    if (yukawa) q.rad.push_back(exp(-o.p.bst * r) / (4.0 * PI) * r * dr);   // FEM.cpp:216
    else q.rad.push_back(pow(o.p.sigma / r, 6.0) * r * r * dr);            // FEM.cpp:166
    q.phi.push_back((dd - cc) / 2.0 * xi[i] + (cc + dd) / 2.0);
    q.dphi.push_back(wi[i] * (dd - cc) / 2.0);
    q.theta.push_back((ff - ee) / 2.0 * xi[i] + (ff + ee) / 2.0);
    q.dtheta.push_back(wi[i] * (ff - ee) / 2.0);
  }
  return q;
}

double quad_at(const Oracle &o, const Quad &q, double x, double y) {
  return node_integral(o, x, y, q.nq, q.rr.data(), q.rad.data(), q.phi.data(), q.dphi.data(),
                       q.theta.data(), q.dtheta.data());
}

// Banded LU without pivoting; A holds row r's entries A[r][c] at
// band[r * (2h+1) + (c - r + h)].
void band_solve(std::vector<double> &band, int n, int h, std::vector<double> &b) {
  const int w = 2 * h + 1;
  auto at = [&](int r, int c) -> double & { return band[(size_t)r * w + (c - r + h)]; };
  for (int k = 0; k < n; ++k) {
    const double piv = at(k, k);
    for (int r = k + 1; r <= std::min(n - 1, k + h); ++r) {
      const double l = at(r, k) / piv;
      if (l == 0.0) continue;
      for (int c = k; c <= std::min(n - 1, k + h); ++c) at(r, c) -= l * at(k, c);
      b[r] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double t = b[k];
    for (int c = k + 1; c <= std::min(n - 1, k + h); ++c) t -= at(k, c) * b[c];
    b[k] = t / at(k, k);
  }
}

// Dense Gaussian elimination with partial pivoting (periodic planes).
void dense_solve(std::vector<double> &A, int n, std::vector<double> &b) {
  for (int k = 0; k < n; ++k) {
    int pr = k;
    for (int r = k + 1; r < n; ++r)
      if (fabs(A[(size_t)r * n + k]) > fabs(A[(size_t)pr * n + k])) pr = r;
    if (pr != k) {
      for (int c = 0; c < n; ++c) std::swap(A[(size_t)k * n + c], A[(size_t)pr * n + c]);
      std::swap(b[k], b[pr]);
    }
    for (int r = k + 1; r < n; ++r) {
      const double l = A[(size_t)r * n + k] / A[(size_t)k * n + k];
      if (l == 0.0) continue;
      for (int c = k; c < n; ++c) A[(size_t)r * n + c] -= l * A[(size_t)k * n + c];
      b[r] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double t = b[k];
    for (int c = k + 1; c < n; ++c) t -= A[(size_t)k * n + c] * b[c];
    b[k] = t / A[(size_t)k * n + k];
  }
}

// compute_SP_2D (FEM.cpp:249-554) up to the nodal field.
void orc_sp2d(Oracle &o) {
  orc_deposit(o);
  const mcpm_params &p = o.p;
  const int nx = o.nx, ny = o.ny, nn = (nx + 1) * (ny + 1);
  const double dudv[4][4] = {{2.0 / 3.0, -1.0 / 6.0, -1.0 / 6.0, -1.0 / 3.0},
                             {-1.0 / 6.0, 2.0 / 3.0, -1.0 / 3.0, -1.0 / 6.0},
                             {-1.0 / 6.0, -1.0 / 3.0, 2.0 / 3.0, -1.0 / 6.0},
                             {-1.0 / 3.0, -1.0 / 6.0, -1.0 / 6.0, 2.0 / 3.0}};  // FEM.cpp:73-88
  const double uv[4][4] = {{1.0 / 9.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0},
                           {1.0 / 18.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 18.0},
                           {1.0 / 18.0, 1.0 / 36.0, 1.0 / 9.0, 1.0 / 18.0},
                           {1.0 / 36.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 9.0}};  // FEM.cpp:93-108
  const double lambda2 = pow(p.bst, 2.0);
  const double V = p.cell_volume;
  const bool periodic = p.boundary == MCPM_BC_PERIODIC;
  std::vector<double> x(nn, 0.0);
  if (!periodic) {
    const int h = nx + 2, w = 2 * h + 1;
    std::vector<double> band((size_t)nn * w, 0.0), b(nn, 0.0);
    auto A = [&](int r, int c) -> double & { return band[(size_t)r * w + (c - r + h)]; };
    for (int k = 0; k < ny; ++k)
      for (int j = 0; j < nx; ++j) {  // FEM.cpp:283-305
        const int num = k * nx + j;
        const CellGeo g = cell_geo(o, j, k);
        double v[4];
        for (int a = 0; a < 4; ++a) {
          v[a] = 0.0;
          for (int c = 0; c < 4; ++c) v[a] += o.n_n[4 * (size_t)num + c] * uv[a][c];
        }
        for (int a = 0; a < 4; ++a) {
          for (int c = 0; c < 4; ++c) {
            A(g.id[a], g.id[c]) += dudv[a][c];
            A(g.id[a], g.id[c]) += lambda2 * uv[a][c] * V;
          }
          b[g.id[a]] += v[a] * V;
        }
      }
    // Dirichlet rows from the Green's function, in the reference's wall order
    // (lower, upper, left, right; FEM.cpp:396-448)
    const Quad q = make_quad(o, true);
    auto dirichlet = [&](int j, int k, int corner) {
      const CellGeo g = cell_geo(o, j, k);
      const int r = g.id[corner];
      for (int c = std::max(0, r - h); c <= std::min(nn - 1, r + h); ++c) A(r, c) = 0.0;
      A(r, r) = 1.0;
      b[r] = quad_at(o, q, g.x_n[corner], g.y_n[corner]);
    };
    for (int jj = 0; jj < nx; ++jj) { dirichlet(jj, 0, 0); dirichlet(jj, 0, 1); }
    for (int jj = 0; jj < nx; ++jj) { dirichlet(jj, ny - 1, 2); dirichlet(jj, ny - 1, 3); }
    for (int jj = 0; jj < ny; ++jj) { dirichlet(0, jj, 0); dirichlet(0, jj, 2); }
    for (int jj = 0; jj < ny; ++jj) { dirichlet(nx - 1, jj, 1); dirichlet(nx - 1, jj, 3); }
    band_solve(band, nn, h, b);
    x = b;
  } else {
    // periodic plane: unique nodes (i mod nx, j mod ny)
    const int nu = nx * ny;
    std::vector<double> A((size_t)nu * nu, 0.0), b(nu, 0.0);
    auto uid = [&](int id) {
      const int i = id % (nx + 1), j = id / (nx + 1);
      return (j % ny) * nx + (i % nx);
    };
    for (int k = 0; k < ny; ++k)
      for (int j = 0; j < nx; ++j) {
        const int num = k * nx + j;
        const CellGeo g = cell_geo(o, j, k);
        double v[4];
        for (int a = 0; a < 4; ++a) {
          v[a] = 0.0;
          for (int c = 0; c < 4; ++c) v[a] += o.n_n[4 * (size_t)num + c] * uv[a][c];
        }
        for (int a = 0; a < 4; ++a) {
          for (int c = 0; c < 4; ++c) {
            A[(size_t)uid(g.id[a]) * nu + uid(g.id[c])] += dudv[a][c];
            A[(size_t)uid(g.id[a]) * nu + uid(g.id[c])] += lambda2 * uv[a][c] * V;
          }
          b[uid(g.id[a])] += v[a] * V;
        }
      }
    dense_solve(A, nu, b);
    for (int id = 0; id < nn; ++id) x[id] = b[uid(id)];
  }
  // core subtraction (FEM.cpp:497-554)
  double dummy = 1.0 - (p.sigma * p.bst + 1.0) * exp(-p.bst * p.sigma);
  dummy = dummy / p.bst / p.bst;
  double dummy2 = (-1.0 / 6.0 * pow(p.bst, 3) * pow(p.sigma, 3) -
                   1.0 / 2.0 * pow(p.bst, 2) * pow(p.sigma, 2) - p.bst * p.sigma +
                   exp(p.bst * p.sigma) - 1) *
                  exp(-p.bst * p.sigma) / pow(p.bst, 4);
  dummy2 = dummy2 / pow(o.dx, 2);
  if (!periodic) {
    for (int j = 0; j < ny + 1; ++j)
      for (int i = 0; i < nx + 1; ++i) {
        const int id = j * (nx + 1) + i;
        int ii = i, jj = j, k = 0;
        int ipm[4] = {0, -1, 0, 0}, jpm[4] = {0, 0, 0, -1}, kpm[4] = {1, 0, 2, 0};
        if (i == nx) { ii = i - 1; jj = j; k = 1; kpm[0] = k; }
        if (j == ny) { ii = i; jj = j - 1; k = 2; kpm[2] = k; }
        if (i == nx && j == ny) { ii = i - 1; jj = j - 1; k = 3; kpm[0] = k; kpm[2] = k; }
        if (i == 0) { ipm[1] = 0; jpm[1] = 0; kpm[1] = 0; }
        if (j == 0) { ipm[3] = 0; jpm[3] = 0; kpm[3] = 0; }
        int num = jj * nx + ii;
        const double d2 = o.n_n[4 * (size_t)num + k] * dummy;
        double d3 = -4.0 * o.n_n[4 * (size_t)num + k] * dummy2;
        for (int ll = 0; ll < 4; ++ll) {
          num = (jj + jpm[ll]) * nx + ii + ipm[ll];
          d3 += o.n_n[4 * (size_t)num + kpm[ll]] * dummy2;
        }
        x[id] += -(d2 + d3);
      }
  } else {
    const std::vector<double> nv = node_values(o);
    auto nat = [&](int i, int j) {
      i = (i % nx + nx) % nx;
      j = (j % ny + ny) % ny;
      return nv[(size_t)j * (nx + 1) + i];
    };
    for (int j = 0; j < ny + 1; ++j)
      for (int i = 0; i < nx + 1; ++i) {
        const double nc = nat(i, j);
        const double d2 = nc * dummy;
        double d3 = -4.0 * nc * dummy2;
        d3 += nat(i + 1, j) * dummy2;
        d3 += nat(i - 1, j) * dummy2;
        d3 += nat(i, j + 1) * dummy2;
        d3 += nat(i, j - 1) * dummy2;
        x[(size_t)j * (nx + 1) + i] += -(d2 + d3);
      }
  }
  o.pot = x;
  o.kick_coef = p.ast;
  o.kick_pending = true;
}

// compute_Vlasov_2D (FEM.cpp:613-706) up to the nodal field.
void orc_vlasov2d(Oracle &o) {
  const Quad q = make_quad(o, false);
  o.pot.assign((size_t)(o.nx + 1) * (o.ny + 1), 0.0);
  for (int j = 0; j < o.ny + 1; ++j)
    for (int i = 0; i < o.nx + 1; ++i) {
      int ii, jj, k;
      owner(o, i, j, &ii, &jj, &k);
      const CellGeo g = cell_geo(o, ii, jj);
      o.pot[(size_t)j * (o.nx + 1) + i] = quad_at(o, q, g.x_n[k], g.y_n[k]);
    }
  o.kick_coef = o.p.phi;
  o.kick_pending = true;
}

int take_err(Oracle &o) {
  uint32_t e = o.err;
  o.err = 0;
  if (e & 1u) return MCPM_ERR_DOMAIN;
  if (e & 2u) return MCPM_ERR_NAN;
  if (e & 4u) return MCPM_ERR_BC;
  return MCPM_OK;
}

}  // namespace

extern "C" {

void *orc_create(const mcpm_params *p) {
  if (!p || p->n_cells < 1) return nullptr;
  Oracle *o = new Oracle();
  o->p = *p;
  if (o->p.dims == 0) o->p.dims = 1;
  o->dims = o->p.dims;
  if (o->dims == 2) {
    if (p->n_cells_x < 1) {
      delete o;
      return nullptr;
    }
    o->nx = p->n_cells_x;
    o->ny = p->n_cells;
    o->dx = p->length_x / (double)p->n_cells_x;
    o->S = 5;
    o->U = 2;
  } else {
    o->ny = p->n_cells;
  }
  o->nc = o->nx * o->ny;
  o->dy = p->length / (double)p->n_cells;
  o->cnt.assign(o->nc, 0);
  o->start.assign(o->nc + 1, 0);
  o->nden.assign(o->nc, 0.0);
  o->yc.assign(o->nc, 1.0);
  o->omax.assign(o->nc, 0.0);
  o->F2.assign(o->nc, 0.0);
  o->samples.assign((size_t)o->nc * MCPM_NSAMP, 0.0);
  o->pm.assign(o->nc, 0.0);
  return o;
}

void orc_destroy(void *h) { delete (Oracle *)h; }

int orc_upload(void *h, int64_t n, const double *y, const double *u1, const double *u2,
               const double *u3) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1) return MCPM_ERR_ARG;
  o.n = n;
  o.rec.assign(4 * (size_t)n, 0.0);
  bool f32 = o.p.precision == MCPM_FP32;
  for (int64_t k = 0; k < n; ++k) {
    double v[4] = {y[k], u1[k], u2[k], u3[k]};
    for (int q = 0; q < 4; ++q) o.rec[4 * k + q] = f32 ? (double)(float)v[q] : v[q];
  }
  o.sorted = false;
  o.kick_pending = false;
  if (!o.cr.empty()) {
    o.track_w = false;
    o.w_pending = true;
  }
  return MCPM_OK;
}

int orc_upload_xy(void *h, int64_t n, const double *x, const double *y, const double *u1,
                  const double *u2, const double *u3) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 2) return MCPM_ERR_ARG;
  o.n = n;
  o.rec.assign(5 * (size_t)n, 0.0);
  bool f32 = o.p.precision == MCPM_FP32;
  for (int64_t k = 0; k < n; ++k) {
    double v[5] = {x[k], y[k], u1[k], u2[k], u3[k]};
    for (int q = 0; q < 5; ++q) o.rec[5 * k + q] = f32 ? (double)(float)v[q] : v[q];
  }
  o.sorted = false;
  o.kick_pending = false;
  return MCPM_OK;
}

// initialization (Main_Functions.cpp:4222-4815) with the library's Philox
// blocks: (i, TAG_INIT, 0) position, (i, TAG_INIT, 1) two Box-Muller pairs;
// slab and closed box: uniform along streaming axes, L/2 elsewhere;
// injection: the last Nh particles in the centred disc (r = R sqrt(u),
// theta = 2 pi u'); velocities U0 + sqrt(kB T_init / m) z.
int orc_cell_update(void *h);
int orc_initialize(void *h, const mcpm_settings *s) {
  Oracle &o = *(Oracle *)h;
  const mcpm_params &p = o.p;
  const bool disc = s->problem == 12;
  if (disc != (o.dims == 2) || s->n_particles < 0 || s->n_drop > s->n_particles)
    return MCPM_ERR_ARG;
  const int64_t n = s->n_particles;
  const bool f32 = p.precision == MCPM_FP32;
  auto rnd_t = [&](double v) { return f32 ? (double)(float)v : v; };
  const uint32_t key0 = (uint32_t)p.seed, key1 = (uint32_t)(p.seed >> 32);
  const double Lx = o.dims == 2 ? p.length_x : s->box_len[0], Ly = p.length;
  const double sd = sqrt(p.kb * s->T_init / p.mass);
  o.n = n;
  o.rec.assign((size_t)o.S * n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t lo = (uint32_t)i, hi = (uint32_t)((uint64_t)i >> 32);
    mcpm_u32x4 w = mcpm_philox(lo, hi, 0u, mcpm_ctr3(MCPM_TAG_INIT, 0u, 0u), key0, key1);
    mcpm_u32x4 w1 = mcpm_philox(lo, hi, 0u, mcpm_ctr3(MCPM_TAG_INIT, 0u, 1u), key0, key1);
    double x, y;
    if (disc && i >= n - s->n_drop) {
      double r = (s->drop_dia / 2.0) * sqrt(mcpm_u01(w.v[0]));
      double sn, cs;
      mcpm_sincos2pi(mcpm_u01(w.v[1]), &sn, &cs);
      x = 0.5 * Lx + r * cs;
      y = 0.5 * Ly + r * sn;
    } else if (disc) {
      x = mcpm_u01(w.v[0]) * Lx;
      y = mcpm_u01(w.v[1]) * Ly;
    } else {
      x = s->direction[0] ? mcpm_u01(w.v[0]) * Lx : 0.5 * Lx;
      y = s->direction[1] ? mcpm_u01(w.v[1]) * Ly : 0.5 * Ly;
    }
    double z0, z1, z2, z3;
    mcpm_normal2(w1.v[0], w1.v[1], &z0, &z1);
    mcpm_normal2(w1.v[2], w1.v[3], &z2, &z3);
    double *r = &o.rec[(size_t)o.S * i];
    double yr = rnd_t(y);
    if (!(yr < rnd_t(Ly))) yr = 0.0;
    if (o.dims == 2) {
      double xr = rnd_t(x);
      if (!(xr < rnd_t(Lx))) xr = 0.0;
      r[0] = xr;
      r[1] = yr;
    } else {
      r[0] = yr;
    }
    double *u = r + o.U;
    u[0] = rnd_t(s->U0[0] + sd * z0);
    u[1] = rnd_t(s->U0[1] + sd * z1);
    u[2] = rnd_t(s->U0[2] + sd * z2);
  }
  o.sorted = false;
  o.kick_pending = false;
  o.have_closures = false;
  if (!o.cr.empty()) {
    o.track_w = false;
    o.w_pending = true;
  }
  return orc_cell_update(h);
}

int orc_download_xy(void *h, double *x, double *y, double *u1, double *u2, double *u3) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 2) return MCPM_ERR_ARG;
  double *dst[5] = {x, y, u1, u2, u3};
  for (int64_t k = 0; k < o.n; ++k)
    for (int q = 0; q < 5; ++q)
      if (dst[q]) dst[q][k] = o.rec[5 * k + q];
  return MCPM_OK;
}

int orc_get_nodes(void *h, double *n_node, double *pot) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 2 || !o.sorted) return MCPM_ERR_ARG;
  const size_t nn = (size_t)(o.nx + 1) * (o.ny + 1);
  if (n_node) {
    orc_deposit(o);
    std::vector<double> v = node_values(o);
    memcpy(n_node, v.data(), nn * sizeof(double));
  }
  if (pot) {
    if (o.pot.size() != nn) o.pot.assign(nn, 0.0);
    memcpy(pot, o.pot.data(), nn * sizeof(double));
  }
  return MCPM_OK;
}

/* The plane's nodal field set from outside (a test hands the oracle the
 * field another solver produced, to compare the kick alone); the next push
 * kicks with coef (ast for the screened Poisson field, phi for the direct
 * integral). */
int orc_set_nodes(void *h, const double *pot, double coef) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 2 || !pot) return MCPM_ERR_ARG;
  const size_t nn = (size_t)(o.nx + 1) * (o.ny + 1);
  o.pot.assign(pot, pot + nn);
  o.kick_coef = coef;
  o.kick_pending = true;
  return MCPM_OK;
}

void orc_gl_full(int n, double *x, double *w) { gl_full(n, x, w); }

/* Probe of the plane's force interpolation at (x, y) for a given nodal field. */
int orc_fem_force(void *h, double x, double y, const double *pot, double coef, double *F1,
                  double *F2) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 2) return MCPM_ERR_ARG;
  const int num = orc_cell(y, o.dy, o.ny) * o.nx + orc_cell(x, o.dx, o.nx);
  fem_force(o, num, x, y, pot, coef, F1, F2);
  return MCPM_OK;
}

int orc_download(void *h, double *y, double *u1, double *u2, double *u3) {
  Oracle &o = *(Oracle *)h;
  const int y0 = o.dims == 2 ? 1 : 0;
  for (int64_t k = 0; k < o.n; ++k) {
    const double *r = &o.rec[(size_t)o.S * k] + y0;
    if (y) y[k] = r[0];
    if (u1) u1[k] = r[1];
    if (u2) u2[k] = r[2];
    if (u3) u3[k] = r[3];
  }
  return MCPM_OK;
}

int orc_cell_update(void *h) {
  Oracle &o = *(Oracle *)h;
  orc_sort(o, true);
  return take_err(o);
}

int orc_collide_esmc(void *h, int64_t step, mcpm_collision_stats *st) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) orc_collide<float>(o, step, st);
  else orc_collide<double>(o, step, st);
  return MCPM_OK;
}

int orc_collide_esmc_new4(void *h, int64_t step, mcpm_collision_stats *st) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.dims == 2) return MCPM_ERR_ARG;
  if (o.p.precision == MCPM_FP32) orc_collide<float>(o, step, st, true);
  else orc_collide<double>(o, step, st, true);
  return MCPM_OK;
}

int orc_poisson_solver(void *h) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.dims == 2) {
    orc_sp2d(o);
    return MCPM_OK;
  }
  orc_field(o);
  return MCPM_OK;
}

int orc_kick(void *h) {
  Oracle &o = *(Oracle *)h;
  if (!o.kick_pending) return MCPM_OK;
  if (o.p.precision == MCPM_FP32) orc_push<float>(o, true, false, 0);
  else orc_push<double>(o, true, false, 0);
  o.kick_pending = false;
  orc_sort(o, false);
  return take_err(o);
}

int orc_stream_bc(void *h, int64_t step, int32_t with_kick) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  bool kick = with_kick && o.kick_pending;
  if (o.p.precision == MCPM_FP32) orc_push<float>(o, kick, true, step);
  else orc_push<double>(o, kick, true, step);
  o.kick_pending = false;
  orc_sort(o, false);
  return take_err(o);
}

int orc_vlasov_integral(void *h) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.dims == 2) {
    orc_vlasov2d(o);
    return MCPM_OK;
  }
  orc_vlasov(o);
  return MCPM_OK;
}

int orc_density_expansion(void *h) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  orc_lr2(o);
  return MCPM_OK;
}

int orc_get_kde(void *h, double *ns, double *d2, double *d3) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  std::vector<double> a, b, c;
  orc_kde(o, a, b, c);
  for (int k = 0; k < o.nc; ++k) {
    if (ns) ns[k] = a[k];
    if (d2) d2[k] = b[k];
    if (d3) d3[k] = c[k];
  }
  return MCPM_OK;
}

int orc_fp_closures(void *h, int64_t step) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  orc_closures(o, step);
  return take_err(o);
}

int orc_velocity_fp(void *h, int64_t step) {
  Oracle &o = *(Oracle *)h;
  if (!o.have_closures) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) orc_fp_velocity<float>(o, step);
  else orc_fp_velocity<double>(o, step);
  return take_err(o);
}

int orc_get_closures(void *h, double *out) {
  Oracle &o = *(Oracle *)h;
  if (!o.have_closures) return MCPM_ERR_STATE;
  memcpy(out, o.clos.data(), o.clos.size() * sizeof(double));
  return MCPM_OK;
}

static int orc_dsmc_any(void *h, int64_t step, mcpm_collision_stats *st, int variant) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) orc_dsmc<float>(o, step, st, variant);
  else orc_dsmc<double>(o, step, st, variant);
  return MCPM_OK;
}

int orc_collide_dsmc_vhs(void *h, int64_t step, mcpm_collision_stats *st) {
  return orc_dsmc_any(h, step, st, MCPM_MODEL_DSMC_VHS);
}

int orc_collide_dsmc_hs(void *h, int64_t step, mcpm_collision_stats *st) {
  return orc_dsmc_any(h, step, st, MCPM_MODEL_DSMC_HS);
}

int orc_collide_cba_vhs(void *h, int64_t step, mcpm_collision_stats *st) {
  return orc_dsmc_any(h, step, st, MCPM_MODEL_CBA_VHS);
}

int orc_collide_esmc_mt(void *h, int64_t step, mcpm_collision_stats *st) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) orc_collide_mt<float>(o, step, st);
  else orc_collide_mt<double>(o, step, st);
  return MCPM_OK;
}

int orc_collide_cba_hs(void *h, int64_t step, mcpm_collision_stats *st) {
  Oracle &o = *(Oracle *)h;
  if (!o.sorted) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) orc_cba_hs<float>(o, step, st);
  else orc_cba_hs<double>(o, step, st);
  return MCPM_OK;
}

// velocity_update dispatch (Main_Functions.cpp:1545-1600) + long range +
// stream and boundary, in the order of main.cpp:389-669.
int orc_step(void *h, int64_t step, mcpm_collision_stats *st) {
  Oracle &o = *(Oracle *)h;
  int r = MCPM_OK;
  if (st) memset(st, 0, sizeof *st);
  switch (o.p.model) {
    case MCPM_MODEL_ESMC: r = orc_collide_esmc(h, step, st); break;
    case MCPM_MODEL_ESMC_NEW4: r = orc_collide_esmc_new4(h, step, st); break;
    case MCPM_MODEL_FP_DENSE:
    case MCPM_MODEL_FP_IDEAL:
    case MCPM_MODEL_FP_LINEAR:
      r = orc_fp_closures(h, step);
      if (!r) r = orc_velocity_fp(h, step);
      break;
    case MCPM_MODEL_DSMC_VHS: r = orc_collide_dsmc_vhs(h, step, st); break;
    case MCPM_MODEL_DSMC_HS: r = orc_collide_dsmc_hs(h, step, st); break;
    case MCPM_MODEL_CBA_VHS: r = orc_collide_cba_vhs(h, step, st); break;
    case MCPM_MODEL_CBA_HS: r = orc_collide_cba_hs(h, step, st); break;
    default: break;
  }
  if (r) return r;
  if (o.p.long_range == MCPM_LR_SCREENED_POISSON) r = orc_poisson_solver(h);
  else if (o.p.long_range == MCPM_LR_DIRECT) r = orc_vlasov_integral(h);
  else if (o.p.long_range == MCPM_LR_DENSITY_EXPANSION) orc_lr2(o);
  if (r) return r;
  r = orc_stream_bc(h, step, 1);
  o.have_closures = false;
  return r;
}

int orc_moments(void *h, double *out) {
  orc_moments(*(Oracle *)h, out);
  return MCPM_OK;
}

int orc_get_escaped(void *h, int64_t *count) {
  *count = ((Oracle *)h)->escaped;
  return MCPM_OK;
}

// thermostat, inverted problem (Main_Functions.cpp:1760-1830): cells by cell
// centre, lambda = sqrt(T_target / cells[num].T), U = lambda (U - U_space) +
// U_space for every particle of the cell (cells with T = 0 are skipped).
/* Per-cell weights (cells[num].n_ratio); null: p.n_ratio everywhere. */
int orc_set_cell_ratio(void *h, const int32_t *ratio) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1) return MCPM_ERR_ARG;
  if (!ratio) {
    o.cr.clear();
    o.track_w = o.w_pending = false;
    if (o.sorted) orc_densities(o);
    return MCPM_OK;
  }
  for (int c = 0; c < o.nc; ++c)
    if (ratio[c] < 1) return MCPM_ERR_ARG;
  o.cr.assign(ratio, ratio + o.nc);
  o.track_w = false;
  o.w_pending = true;
  if (o.sorted) {
    o.pw.assign((size_t)o.n, 0);
    for (int c = 0; c < o.nc; ++c)
      for (int u = 0; u < o.cnt[c]; ++u) o.pw[(size_t)o.start[c] + u] = o.cr[c];
    o.track_w = true;
    o.w_pending = false;
    orc_densities(o);
  }
  return MCPM_OK;
}

int orc_set_particle_ratio(void *h, const int32_t *w) {
  Oracle &o = *(Oracle *)h;
  if (o.cr.empty()) return MCPM_ERR_STATE;
  for (int64_t k = 0; k < o.n; ++k)
    if (w[k] < 1) return MCPM_ERR_ARG;
  o.pw.assign(w, w + o.n);
  o.track_w = true;
  o.w_pending = false;
  return MCPM_OK;
}

int orc_get_particle_ratio(void *h, int32_t *w) {
  Oracle &o = *(Oracle *)h;
  if (!o.track_w) return MCPM_ERR_STATE;
  for (int64_t k = 0; k < o.n; ++k) w[k] = o.pw[k];
  return MCPM_OK;
}

/* inverted_addremove_particles (Main_Functions.cpp:1206-1450) on the sorted
 * slab, cell by cell.  Per particle: heavier than its cell -> it splits
 * (:1254-1259); lighter -> deleted when the draw (p, 0, TAG_ADDREM, 0) is
 * below 1 - 1/n_ratio (:1262-1268), else it takes the cell's weight
 * (:1270-1273).  Per cell with splits: the reference's mean velocity, the
 * last kept particle's velocity times its weight over the cell weight
 * (:1302-1313), sig^2 = sum of weight |U - mean|^2 over the kept / weight
 * (:1314-1325, summed in WgSum's order), h = 1.06 sig (n_rem +
 * n_ratio)^(-1/5) (:1343) by the deterministic exp and log; copy j of a
 * split particle at its position with U + h (z0, z1, z2) from the block
 * (p, j, TAG_ADDREM, 1) (:1344-1352).  Output per cell: the kept particles
 * in order, then the copies by parent and j, all with the cell's weight
 * (:1435-1441). */
typedef struct orc_addremove_stats {
  int64_t before, split, created, deleted, after;
} orc_addremove_stats;

}  // extern "C"

template <typename real>
int orc_addremove_t(Oracle &o, int32_t nr, int64_t step, orc_addremove_stats *st) {
  const int nc = o.nc;
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  const uint32_t key0 = (uint32_t)o.p.seed, key1 = (uint32_t)(o.p.seed >> 32);
  const double pdel = 1.0 - 1.0 / ((double)nr * 1.0);
  std::vector<uint8_t> fl((size_t)o.n, 0);
  std::vector<int> ns(nc, 0), nk(nc, 0), nd(nc, 0);
  std::vector<double> hc(nc, 0.0);
  int64_t total = 0;
  for (int c = 0; c < nc; ++c) {
    const int N = o.cnt[c], s0 = o.start[c];
    const int32_t wc = o.cr[c];
    uint64_t W = 0;
    int last = -1;
    for (int u = 0; u < N; ++u) {
      const uint32_t p = (uint32_t)(s0 + u);
      int32_t w = o.pw[p];
      uint8_t f = 0;
      if (w > wc) {
        f = 1;
        ++ns[c];
      } else if (w < wc) {
        mcpm_u32x4 r = mcpm_philox(p, 0u, slo, mcpm_ctr3(MCPM_TAG_ADDREM, shi, 0u), key0, key1);
        if (mcpm_u01(r.v[0]) < pdel) {
          f = 2;
          ++nd[c];
        } else {
          w = wc;
          ++nk[c];
        }
      } else {
        ++nk[c];
      }
      fl[p] = f;
      if (f != 2) {
        W += (uint64_t)(uint32_t)w;
        last = u;
      }
    }
    if (ns[c] > 0) {
      const uint32_t pl = (uint32_t)(s0 + last);
      const double *rl = &o.rec[4 * (size_t)pl];
      const double wl = (double)std::max(o.pw[pl], wc);
      const double Wd = 1.0 * (double)W;
      const double m1 = (rl[1] * wl) / Wd, m2 = (rl[2] * wl) / Wd, m3 = (rl[3] * wl) / Wd;
      WgSum<1> ws;
      for (int u = 0; u < N; ++u) {
        const uint32_t p = (uint32_t)(s0 + u);
        if (fl[p] == 2) {
          ws.touch(u % 256);
          continue;
        }
        const double *r = &o.rec[4 * (size_t)p];
        const double a = r[1] - m1, b = r[2] - m2, d = r[3] - m3;
        const double e = a * a + b * b + d * d;
        const double v[1] = {e * (double)std::max(o.pw[p], wc)};
        ws.add(u % 256, v);
      }
      double S[1];
      ws.total(S);
      const double sig = sqrt(S[0] / Wd);
      hc[c] = 1.06 * sig * mcpm_exp((-1.0 / 5.0) * mcpm_log(1.0 * (double)(nk[c] + nr)));
    }
    total += ns[c] + nk[c] + (int64_t)ns[c] * (nr - 1);
  }
  if (total > o.p.capacity) return MCPM_ERR_CAPACITY;
  std::vector<double> out(4 * (size_t)total);
  std::vector<int32_t> wout((size_t)total);
  int64_t d = 0;
  int64_t split = 0, del = 0;
  for (int c = 0; c < nc; ++c) {
    const int N = o.cnt[c], s0 = o.start[c];
    const int32_t wc = o.cr[c];
    const int64_t c0 = d;
    for (int u = 0; u < N; ++u) {
      const size_t p = (size_t)(s0 + u);
      if (fl[p] == 2) continue;
      memcpy(&out[4 * (size_t)d], &o.rec[4 * p], 4 * sizeof(double));
      wout[(size_t)d] = wc;
      ++d;
    }
    for (int u = 0; u < N; ++u) {
      const uint32_t p = (uint32_t)(s0 + u);
      if (fl[p] != 1) continue;
      const double *r = &o.rec[4 * (size_t)p];
      for (int32_t j = 0; j < nr - 1; ++j) {
        mcpm_u32x4 q =
            mcpm_philox(p, (uint32_t)j, slo, mcpm_ctr3(MCPM_TAG_ADDREM, shi, 1u), key0, key1);
        double z0, z1, z2, z3;
        mcpm_normal2(q.v[0], q.v[1], &z0, &z1);
        mcpm_normal2(q.v[2], q.v[3], &z2, &z3);
        double *n = &out[4 * (size_t)d];
        n[0] = r[0];
        n[1] = rnd<real>(r[1] + hc[c] * z0);
        n[2] = rnd<real>(r[2] + hc[c] * z1);
        n[3] = rnd<real>(r[3] + hc[c] * z2);
        wout[(size_t)d] = wc;
        ++d;
      }
    }
    o.start[c] = (int)c0;
    o.cnt[c] = (int)(d - c0);
    split += ns[c];
    del += nd[c];
  }
  o.start[nc] = (int)d;
  if (st) {
    st->before = o.n;
    st->split = split;
    st->created = split * (int64_t)(nr - 1);
    st->deleted = del;
    st->after = total;
  }
  o.rec.swap(out);
  o.pw.swap(wout);
  o.n = total;
  orc_densities(o);
  o.have_closures = false;
  return MCPM_OK;
}

extern "C" {

int orc_addremove(void *h, int32_t nr, int64_t step, orc_addremove_stats *st) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1 || nr < 1) return MCPM_ERR_ARG;
  if (o.cr.empty() || !o.track_w || !o.sorted) return MCPM_ERR_STATE;
  if (o.p.precision == MCPM_FP32) return orc_addremove_t<float>(o, nr, step, st);
  return orc_addremove_t<double>(o, nr, step, st);
}

int orc_thermostat(void *h, double zone_lo, double T_lo, double zone_hi, double T_hi) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1) return MCPM_ERR_ARG;
  std::vector<double> mom((size_t)o.nc * MCPM_NMOM);
  orc_moments(o, mom.data());
  const bool f32 = o.p.precision == MCPM_FP32;
  for (int c = 0; c < o.nc; ++c) {
    const double centre = ((double)c + 0.5) * o.dy;
    double Tt;
    if (centre < zone_lo) Tt = T_lo;
    else if (centre > o.p.length - zone_hi) Tt = T_hi;
    else continue;
    const double *m = &mom[(size_t)c * MCPM_NMOM];
    if (!(m[29] > 0.0)) continue;
    const double lam = sqrt(Tt / m[29]);
    for (int u = 0; u < o.cnt[c]; ++u) {
      double *r = &o.rec[4 * (size_t)(o.start[c] + u)];
      for (int d = 0; d < 3; ++d) {
        const double v = lam * (r[1 + d] - m[d]) + m[d];
        r[1 + d] = f32 ? (double)(float)v : v;
      }
    }
  }
  return MCPM_OK;
}

int orc_sample(void *h) {
  orc_sample(*(Oracle *)h);
  return MCPM_OK;
}

int orc_get_samples(void *h, double *out) {
  Oracle &o = *(Oracle *)h;
  memcpy(out, o.samples.data(), o.samples.size() * sizeof(double));
  return MCPM_OK;
}

int orc_reset_samples(void *h) {
  Oracle &o = *(Oracle *)h;
  std::fill(o.samples.begin(), o.samples.end(), 0.0);
  std::fill(o.pm.begin(), o.pm.end(), 0.0);
  return MCPM_OK;
}

int orc_measure_pressure(void *h, int32_t enable) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1) return MCPM_ERR_ARG;
  o.pressure_on = enable != 0;
  return MCPM_OK;
}

// PostProc's ptot and p_rep columns (myIO.cpp:855, 890) from the moments' T.
int orc_get_pressure(void *h, double *p_m, double *p_rep, double *p_tot) {
  Oracle &o = *(Oracle *)h;
  if (o.dims != 1) return MCPM_ERR_ARG;
  std::vector<double> mom((size_t)o.nc * MCPM_NMOM);
  orc_moments(o, mom.data());
  for (int c = 0; c < o.nc; ++c) {
    const double n = o.nden[c], T = mom[(size_t)c * MCPM_NMOM + 29];
    const double pr = n * o.p.kb * T * (1.0 + n * o.p.b * orc_y(n, o.p));
    if (p_m) p_m[c] = o.pm[c];
    if (p_rep) p_rep[c] = pr;
    if (p_tot) p_tot[c] = pr - 2.0 * o.p.b * n * n * o.p.phi;
  }
  return MCPM_OK;
}

int orc_set_omega_max(void *h, const double *om) {
  Oracle &o = *(Oracle *)h;
  for (int c = 0; c < o.nc; ++c) o.omax[c] = om[c];
  return MCPM_OK;
}

int orc_init_omega_max(void *h, double n_gas, int32_t times_y) {
  // main.cpp:250-269
  Oracle &o = *(Oracle *)h;
  std::vector<double> mom((size_t)o.nc * MCPM_NMOM);
  orc_moments(o, mom.data());
  double y = times_y ? orc_y(n_gas, o.p) : 1.0;
  for (int c = 0; c < o.nc; ++c) {
    double T = mom[(size_t)c * MCPM_NMOM + 29];
    double crm = 10.0 * sqrt(o.p.kb * T / o.p.mass);
    double om = 4.0 * PI * o.p.sigma * o.p.sigma * crm * o.nden[c] * o.p.delta_t;
    o.omax[c] = om * y;
  }
  return MCPM_OK;
}

int orc_get_cells(void *h, int32_t *count, int32_t *start, double *n, double *omax, double *F2) {
  Oracle &o = *(Oracle *)h;
  for (int c = 0; c < o.nc; ++c) {
    if (count) count[c] = o.cnt[c];
    if (start) start[c] = o.start[c];
    if (n) n[c] = o.nden[c];
    if (omax) omax[c] = o.omax[c];
    if (F2) F2[c] = o.F2[c];
  }
  return MCPM_OK;
}

int orc_get_collision_log(void *h, int64_t cap, uint32_t *prio, uint32_t *i, uint32_t *j,
                          uint8_t *acc, int64_t *n_out) {
  Oracle &o = *(Oracle *)h;
  int64_t n = (int64_t)o.log_p.size();
  *n_out = n;
  if (n > cap) return MCPM_ERR_CAPACITY;
  for (int64_t k = 0; k < n; ++k) {
    if (prio) prio[k] = o.log_p[k];
    if (i) i[k] = o.log_i[k];
    if (j) j[k] = o.log_j[k];
    if (acc) acc[k] = o.log_a[k];
  }
  return MCPM_OK;
}

/* Philox and deterministic-math probes for the known-answer tests. */
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                uint32_t *out) {
  mcpm_u32x4 r = mcpm_philox(c0, c1, c2, c3, k0, k1);
  for (int q = 0; q < 4; ++q) out[q] = r.v[q];
}
double orc_det_log(double x) { return mcpm_log(x); }
double orc_det_exp(double x) { return mcpm_exp(x); }
void orc_det_sincos2pi(double u, double *s, double *c) { mcpm_sincos2pi(u, s, c); }
uint32_t orc_priority(uint32_t k, uint32_t key) { return mcpm_priority(k, key); }
double orc_contact_y(const mcpm_params *p, double n) { return orc_y(n, *p); }
void orc_gl50(double *x, double *w) { gl50(x, w); }
void orc_gauss_jordan(int n, const double *A, const double *rhs, double *x) {
  std::vector<double> a(A, A + (size_t)n * n), r(rhs, rhs + n);
  for (int k = 0; k < n; ++k) x[k] = 0.0;
  orc_gauss_jordan(a.data(), r.data(), n, x);
}
void orc_thomas(int n, const double *a, const double *b, const double *c, const double *rhs,
                double *x) {
  std::vector<double> A(a, a + n), B(b, b + n), C(c, c + n), R(rhs, rhs + n), X;
  thomas(A, B, C, X, R);
  for (int k = 0; k < n; ++k) x[k] = X[k];
}

}  // extern "C"
