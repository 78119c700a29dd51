// mcpm_dropin.cpp — the reference's C++ operators (include/mcpm_dropin.h)
// over the C ABI of libmcpm.so.
//
// One device context serves the caller's arrays.  It is created on the
// first call from GAS / BOX / CELLS, holds the particles resident, and
// tracks each particle's index in the caller's arrays (mcpm_track_ids).
// Every entry point
//   1. takes in what the caller changed since the last return: the
//      particle arrays (compared with a copy of what was last written back)
//      and the per-cell omega_max (crm for DSMC);
//   2. runs the operator on the device;
//   3. writes back what the operator changed: the particles at their own
//      indices and, after an operation that bins particles, index[] and the
//      cells' lists and fields.
// The library keeps the caller's particle order; its own (physical) order is
// the stable cell order the GPU works in, so cells[c].indices_inside[u] is
// the caller index of the particle at position start[c] + u.
#include "mcpm_dropin.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cstdint>
#include <vector>

namespace {

struct State {
  mcpm_ctx *ctx = nullptr;
  const CELLS *cells = nullptr;  // the cell array the context was made for
  mcpm_params p{};
  int64_t n = 0;
  int nc = 0;
  bool plane = false;
  bool dsmc = false;             // omega_max holds crm (DSMC, CBA models)
  // what was last written to the caller, in caller order
  std::vector<double> sx, sy, su1, su2, su3, som;
  std::vector<int32_t> ids;      // caller index of each device position
  std::vector<int32_t> cnt, start;
  bool have_shadow = false;
  int status = MCPM_OK;
};

State g;

[[noreturn]] void fail(int st, const char *what) {
  g.status = st;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, mcpm_status_string(st));
  throw mcpm_dropin_error(st, buf);
}

void check(int st, const char *what) {
  if (st != MCPM_OK) fail(st, what);
}

int model_of(const std::string &m) {
  if (m == "ESMC") return MCPM_MODEL_ESMC;
  if (m == "FP_dense") return MCPM_MODEL_FP_DENSE;
  if (m == "FP_ideal_gas") return MCPM_MODEL_FP_IDEAL;
  if (m == "FP_linear") return MCPM_MODEL_FP_LINEAR;
  if (m == "DSMC_VHS") return MCPM_MODEL_DSMC_VHS;
  if (m == "DSMC_HS") return MCPM_MODEL_DSMC_HS;
  if (m == "CBA_VHS") return MCPM_MODEL_CBA_VHS;
  if (m == "CBA_HS") return MCPM_MODEL_CBA_HS;
  if (m == "none") return MCPM_MODEL_NONE;
  return -1;
}

// mcpm_params from the reference's structs: the fields readSettingsFile
// fills (myIO.cpp:1756-3016) and the boundary apply_BC picks for the
// problem (Main_Functions.cpp:1451-1492).
mcpm_params params_of(const GAS *gas, const BOX *box, const CELLS *cells) {
  mcpm_params p;
  memset(&p, 0, sizeof p);
  const int mdl = model_of(gas->model);
  if (mdl < 0) fail(MCPM_ERR_ARG, "gas.model");
  p.precision = MCPM_FP64;
  p.y_model = MCPM_Y_REFERENCE;
  p.long_range = gas->long_range;
  p.model = mdl;
  p.capacity = gas->N > 0 ? (int64_t)gas->N : 1;
  const char *seed = getenv("MCPM_SEED");
  p.seed = seed ? strtoull(seed, nullptr, 10) : 1u;
  p.sigma = gas->sigma;
  p.mass = gas->m;
  p.kb = gas->kb;
  p.b = gas->b;
  p.Fn = gas->Fn;
  p.cell_volume = cells[0].volume;
  p.n_ratio = cells[0].n_ratio > 0 ? (double)cells[0].n_ratio : 1.0;
  p.delta_t = gas->delta_t;
  p.ast = gas->ast;
  p.bst = gas->bst;
  p.eqdist = gas->eqdist;
  p.mu = gas->mu;
  p.T0 = gas->T0;
  p.visp = gas->visp;
  p.crref = gas->crref;
  p.phi = gas->phi;
  p.T_wall_lo = box->T_wall_1;
  p.T_wall_hi = box->T_wall_2;
  p.U_wall_lo = box->U_wall_1 < 0.0 ? 0.0 : box->U_wall_1;  // -1 marks "unset"
  p.U_wall_hi = box->U_wall_2 < 0.0 ? 0.0 : box->U_wall_2;
  p.n_cells = box->N[1];
  p.length = box->Len[1];
  if (box->problem == MCPM_REF_INJECTION) {
    p.dims = 2;
    p.n_cells_x = box->N[0];
    p.length_x = box->Len[0];
    p.boundary = MCPM_BC_SPECULAR;  // BC_specular_reflection (Main_Functions.cpp:1485-1486)
    p.boundary_x = MCPM_BC_SPECULAR;
  } else if (box->problem == MCPM_REF_CLOSED_BOX || box->problem == MCPM_REF_COUETTE_FLOW) {
    if (box->N[0] != 1 || box->N[2] != 1) fail(MCPM_ERR_ARG, "box.N (the slab bins along y)");
    p.dims = 1;
    p.n_cells_x = 1;
    p.length_x = box->Len[0];
    p.boundary_x = MCPM_BC_SPECULAR;
    p.boundary = box->problem == MCPM_REF_COUETTE_FLOW ? MCPM_BC_DIFFUSE : MCPM_BC_SPECULAR;
  } else {
    fail(MCPM_ERR_ARG, "box.problem");
  }
  // walls of BC_thermal_wall_3D (Main_Functions.cpp:5088, 5126)
  if (p.boundary == MCPM_BC_DIFFUSE) {
    p.wall_lo = gas->sigma + 1e-15;
    p.wall_hi = box->Len[1] - gas->sigma - 1e-15;
  } else {
    p.wall_lo = 0.0;
    p.wall_hi = box->Len[1];
  }
  return p;
}

bool same(const std::vector<double> &v, const double *a, int64_t n) {
  return a && (int64_t)v.size() == n && memcmp(v.data(), a, (size_t)n * sizeof(double)) == 0;
}

// (Re)create the context when the caller's problem changed.
void ensure(GAS *gas, BOX *box, CELLS *cells) {
  const int64_t n = gas->N;
  if (g.ctx && g.cells == cells && g.n == n) return;
  if (g.ctx) mcpm_dropin_reset();
  g.p = params_of(gas, box, cells);
  check(mcpm_create(&g.p, &g.ctx), "mcpm_create");
  check(mcpm_track_ids(g.ctx, 1), "mcpm_track_ids");
  g.cells = cells;
  g.n = n;
  g.plane = g.p.dims == 2;
  g.nc = g.p.n_cells * (g.plane ? g.p.n_cells_x : 1);
  g.dsmc = g.p.model >= MCPM_MODEL_DSMC_VHS && g.p.model <= MCPM_MODEL_CBA_HS &&
           g.p.model != MCPM_MODEL_NONE;
  g.have_shadow = false;
  g.status = MCPM_OK;
}

// Step 1: upload the particles if they are not what was last written back
// (then the device bins them again), and the per-cell omega_max / crm.
void take_in(const double *x1, const double *x2, const double *U1, const double *U2,
             const double *U3, const CELLS *cells) {
  const int64_t n = g.n;
  bool fresh = !g.have_shadow;
  if (!fresh) {
    if (x2 && !same(g.sy, x2, n)) fresh = true;
    if (g.plane && x1 && !same(g.sx, x1, n)) fresh = true;
    if (U1 && !same(g.su1, U1, n)) fresh = true;
    if (U2 && !same(g.su2, U2, n)) fresh = true;
    if (U3 && !same(g.su3, U3, n)) fresh = true;
  }
  if (fresh) {
    // arrays the call does not pass are taken from what was last written
    auto pick = [&](const double *a, const std::vector<double> &v) -> const double * {
      return a ? a : (g.have_shadow ? v.data() : nullptr);
    };
    const double *px = pick(x1, g.sx), *py = pick(x2, g.sy), *p1 = pick(U1, g.su1),
                 *p2 = pick(U2, g.su2), *p3 = pick(U3, g.su3);
    if (!py || !p1 || !p2 || !p3 || (g.plane && !px))
      fail(MCPM_ERR_STATE, "first call without the particle arrays");
    if (g.plane) check(mcpm_upload_xy(g.ctx, n, px, py, p1, p2, p3), "mcpm_upload_xy");
    else check(mcpm_upload(g.ctx, n, py, p1, p2, p3), "mcpm_upload");
    check(mcpm_cell_update(g.ctx), "mcpm_cell_update");
  }
  std::vector<double> om(g.nc);
  for (int c = 0; c < g.nc; ++c) om[c] = g.dsmc ? cells[c].crm : cells[c].omega_max;
  if (fresh || g.som != om) {
    check(mcpm_set_omega_max(g.ctx, om.data()), "mcpm_set_omega_max");
    g.som = om;
  }
}

// Step 3: the particles back to the caller's indices; with binned, index[]
// and every cell's list, count and density as cell_update leaves them.
void write_back(double *x1, double *x2, double *U1, double *U2, double *U3, int *index,
                CELLS *cells, bool binned) {
  const int64_t n = g.n;
  std::vector<double> a(n), b(n), c(n), d(n), e(g.plane ? n : 0);
  if (g.plane) check(mcpm_download_xy(g.ctx, e.data(), a.data(), b.data(), c.data(), d.data()),
                     "mcpm_download_xy");
  else check(mcpm_download(g.ctx, a.data(), b.data(), c.data(), d.data()), "mcpm_download");
  g.ids.resize(n);
  check(mcpm_download_ids(g.ctx, g.ids.data()), "mcpm_download_ids");
  g.sy.resize(n);
  g.su1.resize(n);
  g.su2.resize(n);
  g.su3.resize(n);
  if (g.plane) g.sx.resize(n);
  for (int64_t k = 0; k < n; ++k) {
    const int32_t q = g.ids[k];
    g.sy[q] = a[k];
    g.su1[q] = b[k];
    g.su2[q] = c[k];
    g.su3[q] = d[k];
    if (g.plane) g.sx[q] = e[k];
  }
  if (x2) memcpy(x2, g.sy.data(), n * sizeof(double));
  if (g.plane && x1) memcpy(x1, g.sx.data(), n * sizeof(double));
  if (U1) memcpy(U1, g.su1.data(), n * sizeof(double));
  if (U2) memcpy(U2, g.su2.data(), n * sizeof(double));
  if (U3) memcpy(U3, g.su3.data(), n * sizeof(double));
  g.have_shadow = true;
  std::vector<double> nd(g.nc), om(g.nc), f2(g.nc);
  g.cnt.resize(g.nc);
  g.start.resize(g.nc);
  check(mcpm_get_cells(g.ctx, g.cnt.data(), g.start.data(), nd.data(), om.data(), f2.data()),
        "mcpm_get_cells");
  for (int cc = 0; cc < g.nc; ++cc) {
    if (g.dsmc) cells[cc].crm = om[cc];
    else cells[cc].omega_max = om[cc];
    cells[cc].F2 = f2[cc];
  }
  g.som = om;
  if (!binned) return;
  // cell_update's lists (Main_Functions.cpp:2367-2376): freed and allocated
  // again, one entry per particle of the cell
  for (int cc = 0; cc < g.nc; ++cc) {
    CELLS &ce = cells[cc];
    free(ce.indices_inside);
    ce.num_inside = g.cnt[cc];
    ce.indices_inside = (int *)malloc((size_t)(ce.num_inside > 0 ? ce.num_inside : 1) * sizeof(int));
    for (int u = 0; u < ce.num_inside; ++u) {
      const int32_t q = g.ids[(size_t)g.start[cc] + u];
      ce.indices_inside[u] = q;
      if (index) index[q] = cc;
    }
    ce.n = nd[cc];
    ce.weight = (double)g.cnt[cc];
  }
}

// The moments and closures cell_update leaves in each cell
// (Main_Functions.cpp:2461-2746).
void write_moments(CELLS *cells, int64_t step) {
  std::vector<double> m((size_t)g.nc * MCPM_NMOM);
  check(mcpm_moments(g.ctx, m.data()), "mcpm_moments");
  const bool fp = !g.plane && (g.p.model == MCPM_MODEL_FP_DENSE || g.p.model == MCPM_MODEL_FP_IDEAL ||
                               g.p.model == MCPM_MODEL_FP_LINEAR);
  std::vector<double> cl;
  if (!g.plane) {
    // Lambda and gamma_st for every model, the FP closures for the FP models
    check(mcpm_fp_closures(g.ctx, step), "mcpm_fp_closures");
    cl.resize((size_t)g.nc * MCPM_NCLOS);
    check(mcpm_get_closures(g.ctx, cl.data()), "mcpm_get_closures");
  }
  for (int c = 0; c < g.nc; ++c) {
    CELLS &ce = cells[c];
    const double *o = &m[(size_t)c * MCPM_NMOM];
    for (int k = 0; k < 3; ++k) ce.U_space[k] = o[k];
    for (int k = 0; k < 6; ++k) ce.PIJ[k] = o[3 + k];
    for (int k = 0; k < 10; ++k) ce.M3[k] = o[9 + k];
    for (int k = 0; k < 6; ++k) ce.M4[k] = o[19 + k];
    for (int k = 0; k < 3; ++k) ce.M5[k] = o[25 + k];
    ce.DM2 = o[28];
    ce.T = o[29];
    ce.weight = o[30];
    for (int k = 0; k < 3; ++k) ce.Mp[k] = o[31 + k];
    ce.DM4 = ce.M4[0] + ce.M4[3] + ce.M4[5];
    ce.Q[0] = ce.M3[0] + ce.M3[3] + ce.M3[5];
    ce.Q[1] = ce.M3[1] + ce.M3[6] + ce.M3[8];
    ce.Q[2] = ce.M3[2] + ce.M3[7] + ce.M3[9];
    if (cl.empty()) continue;
    const double *q = &cl[(size_t)c * MCPM_NCLOS];
    if (ce.weight > 1e-14) {
      ce.Lambda = q[0];
      ce.gamma_st = q[22];
    }
    if (!fp) continue;
    for (int k = 0; k < 6; ++k) ce.c[k] = q[1 + k];
    for (int k = 0; k < 3; ++k) ce.gamma[k] = q[7 + k];
    for (int k = 0; k < 9; ++k) ce.alpha[k] = q[10 + k];
    for (int k = 0; k < 3; ++k) ce.beta[k] = q[19 + k];
    for (int k = 0; k < 3; ++k) ce.phi[k] = q[23 + k];
  }
}

void collide(int (*op)(mcpm_ctx *, int64_t, mcpm_collision_stats *), const char *what,
             double *U1, double *U2, double *U3, double *x1, double *x2, GAS *gas, BOX *box,
             CELLS *cells) {
  ensure(gas, box, cells);
  take_in(x1, x2, U1, U2, U3, cells);
  mcpm_collision_stats st;
  check(op(g.ctx, box->step, &st), what);
  write_back(x1, x2, U1, U2, U3, nullptr, cells, false);
}

// A field solve and its kick (poisson_solver, vlasov_taking_the_integral,
// compute_SP_2D, compute_Vlasov_2D all kick inside).
void field_and_kick(int (*op)(mcpm_ctx *), const char *what, double *x1, double *x2,
                    double *U1, double *U2, GAS *gas, BOX *box, CELLS *cells) {
  ensure(gas, box, cells);
  take_in(x1, x2, U1, U2, nullptr, cells);
  check(op(g.ctx), what);
  check(mcpm_kick(g.ctx), "mcpm_kick");
  write_back(x1, x2, U1, U2, nullptr, nullptr, cells, false);
}

}  // namespace

void cell_update(double *x1, double *x2, double *U1, double *U2, double *U3, GAS *gas,
                 CELLS *cells, BOX *box, int *index, double **T, double *rho, int *color) {
  (void)T;
  (void)rho;
  (void)color;
  ensure(gas, box, cells);
  take_in(x1, x2, U1, U2, U3, cells);
  check(mcpm_cell_update(g.ctx), "mcpm_cell_update");
  write_back(x1, x2, U1, U2, U3, index, cells, true);
  write_moments(cells, box->step);
  g.status = MCPM_OK;
}

void Collision_ESMC_new3(double *U1, double *U2, double *U3, double *x1, double *x2, GAS *gas,
                         BOX *box, CELLS *cells, int *index) {
  (void)index;
  collide(mcpm_collide_esmc, "Collision_ESMC_new3", U1, U2, U3, x1, x2, gas, box, cells);
  g.status = MCPM_OK;
}

void Collision_ESMC_new4(double *U1, double *U2, double *U3, double *x1, double *x2, GAS *gas,
                         BOX *box, CELLS *cells, int *index) {
  (void)index;
  collide(mcpm_collide_esmc_new4, "Collision_ESMC_new4", U1, U2, U3, x1, x2, gas, box, cells);
  g.status = MCPM_OK;
}

void Collision_DSMC_VHS(double *U1, double *U2, double *U3, double *x1, double *x2, GAS *gas,
                        BOX *box, CELLS *cells, Collision *DSMC) {
  (void)DSMC;
  collide(mcpm_collide_dsmc_vhs, "Collision_DSMC_VHS", U1, U2, U3, x1, x2, gas, box, cells);
  g.status = MCPM_OK;
}

void poisson_solver(double *U2, double *x1, double *x2, double *x3, GAS *gas, BOX *box,
                    CELLS *cells, int *index) {
  (void)x3;
  (void)index;
  field_and_kick(mcpm_poisson_solver, "poisson_solver", x1, x2, nullptr, U2, gas, box, cells);
  g.status = MCPM_OK;
}

void vlasov_taking_the_integral(double *U2, GAS *gas, BOX *box, CELLS *cells, int *index,
                                double *x1, double *x2, double *x3) {
  (void)x3;
  (void)index;
  field_and_kick(mcpm_vlasov_integral, "vlasov_taking_the_integral", x1, x2, nullptr, U2, gas,
                 box, cells);
  g.status = MCPM_OK;
}

void compute_SP_2D(GAS *gas, BOX *box, CELLS *cells, double *x1, double *x2, double *U1,
                   double *U2, int *index) {
  (void)index;
  field_and_kick(mcpm_poisson_solver, "compute_SP_2D", x1, x2, U1, U2, gas, box, cells);
  g.status = MCPM_OK;
}

void compute_Vlasov_2D(GAS *gas, BOX *box, CELLS *cells, double *x1, double *x2, double *U1,
                       double *U2, int *index) {
  (void)index;
  field_and_kick(mcpm_vlasov_integral, "compute_Vlasov_2D", x1, x2, U1, U2, gas, box, cells);
  g.status = MCPM_OK;
}

void stream_particles(double *U1, double *U2, double *U3, double *x1, double *x2, double *x3,
                      double *x1_old, double *x2_old, double *x3_old, GAS *gas, BOX *box,
                      CELLS *cells, int *index) {
  (void)x1_old;
  (void)x2_old;
  (void)x3_old;
  ensure(gas, box, cells);
  take_in(x1, x2, U1, U2, U3, cells);
  check(mcpm_stream_bc(g.ctx, box->step, 0), "stream_particles");
  write_back(x1, x2, U1, U2, U3, index, cells, true);
  if (!g.plane) {
    // the slab's other axes: free flight of the enabled directions on the
    // host (Position_FP_ideal_gas, Fokker_Planck.cpp:948-969)
    const double dt = gas->delta_t;
    for (int64_t k = 0; k < g.n; ++k) {
      if (box->direction[0] == 1 && x1) x1[k] = x1[k] + dt * U1[k];
      if (box->direction[2] == 1 && x3) x3[k] = x3[k] + dt * U3[k];
    }
  }
  g.status = MCPM_OK;
}

void apply_BC(double *U1, double *U2, double *U3, double *x1, double *x2, double *x3,
              double *x1_old, double *x2_old, double *x3_old, GAS *gas, BOX *box, CELLS *cells,
              int *index, int *flag, double *xi_1, double *xi_2, double *xi_3, double *xi_1f,
              double *xi_2f, double *xi_3f, double *Mp1, double *Mp2, double *Mp3, double *T) {
  (void)x3; (void)x1_old; (void)x2_old; (void)x3_old; (void)index; (void)flag;
  (void)xi_1; (void)xi_2; (void)xi_3; (void)xi_1f; (void)xi_2f; (void)xi_3f;
  (void)Mp1; (void)Mp2; (void)Mp3; (void)T;
  ensure(gas, box, cells);
  take_in(x1, x2, U1, U2, U3, cells);
  g.status = MCPM_OK;
}

double increase_collision_rate(double n, GAS *gas) {
  mcpm_params p;
  memset(&p, 0, sizeof p);
  p.y_model = MCPM_Y_REFERENCE;
  p.b = gas->b;
  return mcpm_contact_y(&p, n);
}

void random_generator(double *xi_1, double *xi_2, double *xi_3, double *xi_1f, double *xi_2f,
                      double *xi_3f, CELLS *cells, BOX *box, GAS *gas, int *index) {
  (void)xi_1; (void)xi_2; (void)xi_3; (void)xi_1f; (void)xi_2f; (void)xi_3f;
  (void)cells; (void)box; (void)gas; (void)index;
  g.status = MCPM_OK;
}

void c_gamma_calculation(double *M1, double *M2, double *M3, GAS *gas, CELLS *cell, BOX *box) {
  ensure(gas, box, cell);
  take_in(nullptr, nullptr, M1, M2, M3, cell);
  write_moments(cell, box->step);
  g.status = MCPM_OK;
}

void alpha_calculation(double *M1, double *M2, double *M3, GAS *gas, CELLS *cells, BOX *box) {
  c_gamma_calculation(M1, M2, M3, gas, cells, box);
}

void Velocity_FP_parallel(double *M1, double *M2, double *M3, GAS *gas, CELLS *cells, BOX *box,
                          int *index, double *xi_1, double *xi_2, double *xi_3, double *Mp1,
                          double *Mp2, double *Mp3) {
  (void)index; (void)xi_1; (void)xi_2; (void)xi_3; (void)Mp1; (void)Mp2; (void)Mp3;
  ensure(gas, box, cells);
  take_in(nullptr, nullptr, M1, M2, M3, cells);
  int st = mcpm_velocity_fp(g.ctx, box->step);
  if (st == MCPM_ERR_STATE) {  // no closures yet: those of the current state
    check(mcpm_fp_closures(g.ctx, box->step), "mcpm_fp_closures");
    st = mcpm_velocity_fp(g.ctx, box->step);
  }
  check(st, "Velocity_FP_parallel");
  write_back(nullptr, nullptr, M1, M2, M3, nullptr, cells, false);
  g.status = MCPM_OK;
}

void DFP_extra_streaming(double *U1, double *U2, double *U3, double *x1, double *x2, GAS *gas,
                         BOX *box, CELLS *cells, int *index, double dt) {
  (void)U1; (void)U2; (void)U3; (void)x1; (void)x2; (void)gas; (void)box; (void)cells;
  (void)index; (void)dt;
  g.status = MCPM_OK;
}

int mcpm_dropin_status() { return g.status; }

void mcpm_dropin_reset() {
  if (g.ctx) mcpm_destroy(g.ctx);
  g = State();
}
