// mcpm_settings.cpp — the reference's inputs and outputs around the step:
//   mcpm_read_settings    readSettingsFile (myIO.cpp:1756-3016): keywords,
//                         defaults and the derived gas/box constants for the
//                         problems this library runs (closed_box,
//                         couette_flow, injection)
//   mcpm_write_postproc   write_post_processing's PostProc file
//                         (myIO.cpp:597-897)
// Host code only; the initial state (initialization) is drawn on the device
// in mcpm_api.hip.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "mcpm.h"
#include "mcpm_internal.h"

namespace {

const double PI = 3.14159265358979323846;

// Every keyword readSettingsFile accepts (myIO.cpp:1897-2125) with the
// number of values it reads; "model" and "output" read one word.
struct Key {
  const char *name;
  int count;
};
const Key KEYS[] = {
    {"N", 1},          {"gp_nn", 1},      {"N13", 1},         {"sigc", 1},
    {"alphac", 1},     {"sigh", 1},       {"alphah", 1},      {"n", 1},
    {"nc", 1},         {"nh", 1},         {"nvc", 1},         {"nvh", 1},
    {"nv1", 1},        {"nv2", 1},        {"nv3", 1},         {"Tc", 1},
    {"Th", 1},         {"Tv1", 1},        {"Tv2", 1},         {"Tv3", 1},
    {"LcN", 1},        {"LhN", 1},        {"LvN", 1},         {"LvN1", 1},
    {"LvN3", 1},       {"Lv1", 1},        {"Lv3", 1},         {"Nh", 1},
    {"Nv1", 1},        {"Nv2", 1},        {"Nv3", 1},         {"long_range", 1},
    {"box_len", 3},    {"cone_len", 4},   {"x2_prob", 4},     {"ghost", 1},
    {"thermc", 1},     {"thermh", 1},     {"cone_num_cells", 5}, {"cone_angles", 2},
    {"flatnose", 8},   {"flatratio", 2},  {"box_delta_dim", 3}, {"direction", 3},
    {"U0", 3},         {"sigma", 1},      {"weight", 1},      {"dcones_n", 1},
    {"mass", 1},       {"steps", 1},      {"dt", 1},          {"init_step", 1},
    {"model", 1},      {"output", 1},     {"movie", 1},       {"every", 1},
    {"num_realizations", 1}, {"after", 1}, {"reset_every", 1}, {"frac_mft", 1},
    {"T", 1},          {"epsilon", 1},    {"kb", 1},          {"num_thr", 1},
    {"mu", 1},         {"visp", 1},       {"nualpha", 1},     {"phi", 1},
    {"phi0", 1},       {"factor", 1},     {"n_ratio", 1},     {"problem", 1},
    {"U_wall_1", 1},   {"U_wall_2", 1},   {"U_wall_3", 1},    {"U_wall_4", 1},
    {"U_wall_5", 1},   {"U_wall_6", 1},   {"U_wall_7", 1},    {"T_wall_1", 1},
    {"T_wall_2", 1},   {"T_wall_3", 1},   {"T_wall_4", 1},    {"T_wall_5", 1},
    {"T_wall_6", 1},   {"T_wall_7", 1},   {"n_wall_4", 1},    {"n_wall_5", 1},
    {"n_wall_6", 1},   {"n_wall_7", 1},   {"reset", 1},       {"drop_dia", 1},
    {"num_cells", 3},
};

// Integer-typed fields read through operator>> into an int (a value such as
// "1e5" then reads as 1, as in the reference).
bool int_key(const std::string &k) {
  static const char *ints[] = {"N",     "gp_nn",     "N13",    "long_range", "cone_num_cells",
                               "direction", "steps", "init_step", "movie", "every",
                               "num_realizations", "after", "reset_every", "num_thr",
                               "problem", "reset", "num_cells", "Nh", "Nv1", "Nv2", "Nv3",
                               "LcN", "LhN", "LvN"};
  for (const char *s : ints)
    if (k == s) return true;
  return false;
}

struct Parsed {
  std::map<std::string, std::vector<double>> num;
  std::string model = "ESMC", output = "data.txt";
  double get(const char *k, double dflt, int idx = 0) const {
    auto it = num.find(k);
    if (it == num.end() || (int)it->second.size() <= idx) return dflt;
    return it->second[idx];
  }
  bool has(const char *k) const { return num.count(k) != 0; }
};

int parse_file(const char *path, Parsed *out) {
  std::ifstream f(path);
  if (!f) return MCPM_ERR_ARG;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line[0] == '#') continue;  // comment (myIO.cpp:1892-1893)
    std::istringstream iss(line);
    std::string key;
    if (!(iss >> key)) continue;  // blank line
    int count = -1;
    for (const Key &k : KEYS)
      if (key == k.name) count = k.count;
    if (count < 0) return MCPM_ERR_ARG;  // "Unknown keyword in the settings file"
    if (key == "model") {
      iss >> out->model;
      continue;
    }
    if (key == "output") {
      iss >> out->output;
      continue;
    }
    std::vector<double> v;
    const bool is_int = int_key(key);
    for (int q = 0; q < count; ++q) {
      if (is_int) {
        long long x = 0;
        iss >> x;  // a failed read stores 0, as operator>> does
        v.push_back((double)x);
      } else {
        double x = 0.0;
        iss >> x;
        v.push_back(x);
      }
    }
    out->num[key] = v;
  }
  return MCPM_OK;
}

double reference_y(double n, double b) {
  // increase_collision_rate (Main_Functions.cpp:1906-1915)
  const double nb = n * b;
  return 1.0 + 0.625 * n * b + 0.2869 * pow(nb, 2.0) + 0.115 * pow(nb, 3.0);
}

}  // namespace

extern "C" int mcpm_read_settings(const char *path, mcpm_settings *s) {
  if (!path || !s) return MCPM_ERR_ARG;
  Parsed P;
  int st = parse_file(path, &P);
  if (st) return st;
  memset(s, 0, sizeof *s);
  // defaults of readSettingsFile (myIO.cpp:1760-1858) and after the parse
  // (myIO.cpp:2137-2186)
  const int problem = (int)P.get("problem", -1) < 0 ? 1 : (int)P.get("problem", -1);
  if (problem != 1 && problem != 2 && problem != 12) return MCPM_ERR_ARG;
  int N3[3] = {(int)P.get("num_cells", -1, 0), (int)P.get("num_cells", -1, 1),
               (int)P.get("num_cells", -1, 2)};
  double Len[3] = {P.get("box_len", -1.0, 0), P.get("box_len", -1.0, 1),
                   P.get("box_len", -1.0, 2)};
  if (Len[0] < 0.0) Len[0] = Len[1] = Len[2] = 1.0;
  if (N3[0] < 1 || N3[1] < 1 || N3[2] < 1) return MCPM_ERR_ARG;
  double kb = P.get("kb", -1.0);
  if (kb < 0.0) kb = 1.38064852 * 1e-23;
  double T = P.get("T", -1.0);
  if (T < 0.0) T = 1.0 / kb;
  double sigma = P.get("sigma", -1.0);
  if (sigma < 0.0) sigma = 1;
  double n = P.get("n", -1.0);
  if (n < 0.0) n = 1e-3;
  double m = P.get("mass", -1.0);
  if (m < 0.0) m = 1.0;
  int64_t N = (int64_t)P.get("N", -1);
  double U0[3] = {P.get("U0", -1.0, 0), P.get("U0", -1.0, 1), P.get("U0", -1.0, 2)};
  if (U0[0] < 0.0) U0[0] = U0[1] = U0[2] = 1.0;
  int steps = (int)P.get("steps", -1);
  if (steps < 0) steps = 10;
  int every = (int)P.get("every", -1);
  if (every < 0) every = 1;
  int after = (int)P.get("after", -1);
  if (after < 0) after = 1;
  int reset_every = (int)P.get("reset_every", -1);
  if (reset_every < 0) reset_every = 1;
  int nreal = (int)P.get("num_realizations", -1);
  if (nreal < 0) nreal = 1;
  int init_step = (int)P.get("init_step", -1);
  if (init_step < 0) init_step = 0;
  double Tw[4];
  for (int q = 0; q < 4; ++q) {
    char k[16];
    snprintf(k, sizeof k, "T_wall_%d", q + 1);
    Tw[q] = P.get(k, -1.0);
    if (Tw[q] < 0.0) Tw[q] = 1.0;
  }
  const double Uw1 = P.get("U_wall_1", -1.0), Uw2 = P.get("U_wall_2", -1.0);
  double mu = P.get("mu", -1.0);
  const double visp = P.get("visp", 0.5), nualpha = P.get("nualpha", 1.0);
  const double T0 = 273.0;  // myIO.cpp:2235
  double epsilon = P.get("epsilon", -1.0);
  if (epsilon < 0.0) epsilon = 119.8 * kb;
  // sigma and crref from the viscosity (myIO.cpp:2240-2263)
  const double thermal = sqrt(kb * T / m);
  double crref = thermal;
  if (fabs(sigma - 1.0) < 1e-14) {
    crref = thermal;
  } else if (fabs(visp - 0.5) < 1e-13) {
    sigma = 1.016 * 5.0 * sqrt(m * kb * T0 / PI);
    sigma = sigma / (16.0 * mu);
    sigma = sqrt(sigma);
    crref = thermal;
  } else {
    sigma = 1.016 * 5.0 * (nualpha + 1.0) * (nualpha + 2.0) * sqrt(m * kb * T0 / PI);
    sigma = sigma / ((4.0 * nualpha) * (5.0 - 2.0 * visp) * (7.0 - 2.0 * visp) * mu);
    sigma = sqrt(sigma);
    const double numer = m * pow(4.0, visp - 1) * 5.0 * (nualpha + 1.0) * (nualpha + 2.0) * sqrt(PI);
    const double denom = PI * pow(sigma, 2.0) * nualpha * (5.0 - 2.0 * visp) * (7.0 - 2.0 * visp) *
                         2.0 * tgamma(5.0 / 2.0 - visp);
    crref = (numer / denom) * pow(kb * 273.0 / m, visp) / mu;
    crref = pow(crref, 1.0 / (visp * 2.0 - 1.0));
  }
  double delta[3] = {P.get("box_delta_dim", -1.0, 0), P.get("box_delta_dim", -1.0, 1),
                     P.get("box_delta_dim", -1.0, 2)};
  if (delta[0] < 0.0)
    for (int q = 0; q < 3; ++q) delta[q] = Len[q] / N3[q];
  double Fn = n * (Len[0] * Len[1] * Len[2]) / (double)N;
  double eqdist = pow(2.0, 1.0 / 6.0) * sigma;  // myIO.cpp:2820-2822
  int64_t Nh = 0;
  double drop_dia = P.get("drop_dia", 0.0);
  double nc = P.get("nc", 0.0), nh = P.get("nh", 0.0);
  const double Tc = P.get("Tc", 0.0);  // Th is read but used by none of these problems
  if (problem == 12) {
    // injection (myIO.cpp:2729-2761)
    sigma = 3.405e-10;
    eqdist = sigma;
    nc = nc / pow(sigma, 3.0);
    nh = nh / pow(sigma, 3.0);
    N = N * N3[0] * N3[1];
    Len[0] = Len[0] * sigma;
    Len[1] = Len[1] * sigma;
    T = Tc;
    n = nc;
    Fn = n * Len[0] * Len[1] * Len[2] / (1.0 * N);
    drop_dia = drop_dia * sigma;
    Nh = (int64_t)floor((nh - nc) * PI * drop_dia * drop_dia / 4.0 / Fn);
    N = N + Nh;
    for (int q = 0; q < 3; ++q) delta[q] = Len[q] / N3[q];
  }
  if (N < 1) return MCPM_ERR_ARG;
  // the common tail (myIO.cpp:2828-2921)
  const double b = 2 * PI * pow(sigma, 3) / (3.0);
  mu = 5.0 / (16.0 * sigma * sigma) * sqrt(m * kb * T0 / acos(-1.0));
  double lambda = 4.0 * nualpha * (5.0 - 2.0 * visp) * (7.0 - 2.0 * visp);
  lambda = lambda / (5.0 * (nualpha + 1.0) * (nualpha + 2.0));
  lambda = lambda * sqrt(m / (2.0 * PI * kb * T)) * mu / (m * n);
  const double Kn = lambda / Len[1];
  double dt = P.get("dt", -1.0);
  const std::string &model = P.model;
  if (dt < 0.0) {
    const double frac = P.get("frac_mft", -1.0);
    const double f = frac < 0.0 ? 0.5 : frac;
    const double vmax = std::max(sqrt(kb * Tw[0] / m), std::max(Uw1, Uw2));
    if (model == "DSMC_VHS" || model == "CBA_VHS" || model == "ESMC" || model == "CBA_HS" ||
        model == "MD" || model == "FP_dense")
      dt = f * std::min(lambda, Len[1] / N3[1]) / vmax;
    else
      dt = f * Len[1] / N3[1] / vmax;
  }
  double phi = P.get("phi", -1.0);
  phi = phi < 0.0 ? -4.0 * epsilon : -phi * epsilon;  // myIO.cpp:2915-2920
  // the model and the long-range force
  int32_t mdl;
  if (model == "ESMC") mdl = MCPM_MODEL_ESMC;
  else if (model == "FP_dense") mdl = MCPM_MODEL_FP_DENSE;
  else if (model == "FP_ideal_gas") mdl = MCPM_MODEL_FP_IDEAL;
  else if (model == "FP_linear") mdl = MCPM_MODEL_FP_LINEAR;
  else if (model == "DSMC_VHS") mdl = MCPM_MODEL_DSMC_VHS;
  else if (model == "DSMC_HS") mdl = MCPM_MODEL_DSMC_HS;
  else if (model == "CBA_VHS") mdl = MCPM_MODEL_CBA_VHS;
  else if (model == "CBA_HS") mdl = MCPM_MODEL_CBA_HS;
  else if (model == "none") mdl = MCPM_MODEL_NONE;
  else return MCPM_ERR_ARG;
  const int lr = (int)P.get("long_range", 0);
  if (lr < 0 || lr > 3) return MCPM_ERR_ARG;

  mcpm_params &p = s->params;
  p.precision = MCPM_FP64;
  p.y_model = MCPM_Y_REFERENCE;
  p.long_range = lr;
  p.model = mdl;
  p.capacity = N;
  p.seed = 1;
  p.sigma = sigma;
  p.mass = m;
  p.kb = kb;
  p.b = b;
  p.Fn = Fn;
  p.cell_volume = delta[0] * delta[1] * delta[2];
  p.n_ratio = 1.0;
  p.delta_t = dt;
  p.ast = -1.64835851e-28;  // myIO.cpp:2891-2892
  p.bst = 6.91304716e+09;
  p.eqdist = eqdist;
  p.mu = mu;
  p.T0 = T0;
  p.visp = visp;
  p.crref = crref;
  p.phi = phi;
  p.T_wall_lo = Tw[0];
  p.T_wall_hi = Tw[1];
  p.U_wall_lo = Uw1 < 0.0 ? 0.0 : Uw1;
  p.U_wall_hi = Uw2 < 0.0 ? 0.0 : Uw2;
  p.n_cells = N3[1];
  p.length = Len[1];
  if (problem == 12) {
    p.dims = 2;
    p.n_cells_x = N3[0];
    p.length_x = Len[0];
    p.boundary = MCPM_BC_SPECULAR;    // BC_specular_reflection (Main_Functions.cpp:1485-1486)
    p.boundary_x = MCPM_BC_SPECULAR;
  } else {
    if (N3[0] != 1 || N3[2] != 1) return MCPM_ERR_ARG;  // the slab: cells along y only
    p.dims = 1;
    p.n_cells_x = 1;
    p.length_x = Len[0];
    p.boundary_x = MCPM_BC_SPECULAR;
    p.boundary = problem == 2 ? MCPM_BC_DIFFUSE : MCPM_BC_SPECULAR;  // apply_BC (:1458-1461)
  }
  // walls of BC_thermal_wall_3D (Main_Functions.cpp:5088, 5126); the specular
  // slab reflects at the box faces
  if (p.boundary == MCPM_BC_DIFFUSE) {
    p.wall_lo = sigma + 1e-15;
    p.wall_hi = Len[1] - sigma - 1e-15;
  } else {
    p.wall_lo = 0.0;
    p.wall_hi = Len[1];
  }

  s->n_particles = N;
  s->n_drop = Nh;
  s->problem = problem;
  s->num_steps = steps;
  s->every = every;
  s->after = after;
  s->reset_every = reset_every;
  s->num_realizations = nreal;
  s->init_step = init_step;
  for (int q = 0; q < 3; ++q) {
    s->direction[q] = (int32_t)P.get("direction", 0, q);
    s->num_cells[q] = N3[q];
    s->box_len[q] = Len[q];
    s->U0[q] = U0[q];
  }
  s->n = n;
  s->T = T;
  // initialization's Maxwellian temperature (Main_Functions.cpp:4791-4792)
  s->T_init = std::max(T, std::max(Tw[0] + Tw[1], Tw[2] + Tw[3]) / 2.0);
  s->drop_dia = drop_dia;
  s->epsilon = epsilon;
  s->lambda = lambda;
  s->Kn = Kn;
  snprintf(s->model, sizeof s->model, "%s", P.model.c_str());
  snprintf(s->output, sizeof s->output, "%s", P.output.c_str());
  return MCPM_OK;
}

// ---------------------------------------------------------------------------
// PostProc writer
// ---------------------------------------------------------------------------
namespace {

// The header of write_post_processing (myIO.cpp:619-744).
const char *POST_NAMES[] = {
    "#variables=", "1-x1", "x2", "dx", "dy", "dz", "tavg", "dt", "U_space[0]", "U_space[1]",
    "10-U_space[2]", "T", "n", "sum_weight", "sum_M[0]", "sum_M[1]", "sum_M[2]", "sum_MM[0]",
    "sum_MM[1]", "sum_MM[2]", "20-sum_MM[3]", "sum_MM[4]", "sum_MM[5]", "sum_MMM[0]",
    "sum_MMM[1]", "sum_MMM[2]", "sum_MMM[3]", "sum_MMM[4]", "sum_MMM[5]", "sum_MMM[6]",
    "30-sum_MMM[7]", "sum_MMM[8]", "sum_MMM[9]", "sum_T_f[0]", "sum_T_f[1]", "sum_T_f[2]",
    "sum_T_f[3]", "sum_M_n[0]", "sum_M_n[1]", "sum_M_n[2]", "40-sum_M_n[3]", "sum_M1_f[0]",
    "sum_M1_f[1]", "sum_M1_f[2]", "sum_M1_f[3]", "sum_M1_f[4]", "sum_M1_f[5]", "sum_M2_f[0]",
    "sum_M2_f[1]", "sum_M2_f[2]", "50-sum_M2_f[3]", "sum_M2_f[4]", "sum_M2_f[5]", "sum_M3_f[0]",
    "sum_M3_f[1]", "sum_M3_f[2]", "sum_M3_f[3]", "sum_N[2]", "sum_N[3]", "sum_M_np[0]",
    "60-sum_M_np[1]", "sum_M_np[2]", "sum_M_np[3]", "sum_M_np[4]", "sum_M_np[5]",
    "sum_MM_f_1[0]", "sum_MM_f_1[1]", "sum_MM_f_1[2]", "sum_MM_f_1[3]", "sum_MM_f_1[4]",
    "70-sum_MM_f_1[5]", "sum_MM_f_2[0]", "sum_MM_f_2[1]", "sum_MM_f_2[2]", "sum_MM_f_2[3]",
    "sum_MM_f_2[4]", "sum_MM_f_2[5]", "sum_MM_f_3[0]", "sum_MM_f_3[1]", "sum_MM_f_3[2]",
    "80-sum_MM_f_3[3]", "sum_MM_f_3[4]", "sum_MM_f_3[5]", "volume", "n_smooth", "N", "ptot",
    "Fn", "sum_MMMM[0]", "sum_MMMM[1]", "90-sum_MMMM[2]", "sum_MMMM[3]", "sum_MMMM[4]",
    "sum_MMMM[5]", "model", "xx", "PIJ[3]", "Q[1]", "F2", "divU", "100-UU", "UU0", "UUR",
    "UU0R", "alpha1", "alpha2", "alpha3", "alpha4", "alpha5", "alpha6", "110-alpha7", "alpha8",
    "alpha9", "beta1", "beta2", "beta3", "gamma_st", "phi[0]", "phi[1]", "phi[2]", "p_m",
    "p_rep"};
constexpr int POST_NCOL = (int)(sizeof POST_NAMES / sizeof POST_NAMES[0]) - 1;

const char *lr_name(int lr) {
  // myIO.cpp:606-613
  return lr == 1 ? "direct" : lr == 2 ? "denexp" : lr == 3 ? "screened-Poisson" : "noatt";
}

}  // namespace

extern "C" int mcpm_write_postproc(mcpm_ctx *c, const mcpm_settings *s, int64_t step,
                                   const char *dir) {
  if (!c || !s || !dir) return MCPM_ERR_ARG;
  const mcpm_params &p = *ctx_params(c);
  const int nx = p.dims == 2 ? p.n_cells_x : 1, ny = p.n_cells, nc = nx * ny;
  std::vector<double> mom((size_t)nc * MCPM_NMOM), smp((size_t)nc * MCPM_NSAMP);
  std::vector<double> nd(nc), F2(nc), ns(nc, 0.0), pm(nc, 0.0), cl((size_t)nc * MCPM_NCLOS, 0.0);
  std::vector<int32_t> cnt(nc);
  int st = mcpm_moments(c, mom.data());
  if (!st) st = mcpm_get_samples(c, smp.data());
  if (!st) st = mcpm_get_cells(c, cnt.data(), nullptr, nd.data(), nullptr, F2.data());
  if (st) return st;
  if (p.dims == 1) {
    if ((st = mcpm_get_kde(c, ns.data(), nullptr, nullptr))) return st;
    if ((st = mcpm_get_pressure(c, pm.data(), nullptr, nullptr))) return st;
  }
  if (mcpm_get_closures(c, cl.data()) != MCPM_OK) std::fill(cl.begin(), cl.end(), 0.0);
  char path[512];
  snprintf(path, sizeof path, "%s/%s_%s_%lld.dat", dir, s->output, lr_name(p.long_range),
           (long long)step);
  FILE *fp = fopen(path, "w");
  if (!fp) return MCPM_ERR_ARG;
  fprintf(fp, "%10s%10s", POST_NAMES[0], POST_NAMES[1]);
  for (int k = 2; k <= POST_NCOL; ++k) fprintf(fp, "%20s", POST_NAMES[k]);
  fprintf(fp, "\n");
  const double dy = p.length / ny;
  const double dx = p.dims == 2 ? p.length_x / nx : s->box_len[0];
  const double dz = s->box_len[2];
  const double t_post = (double)ctx_sample_count(c) * p.delta_t;
  for (int i = 0; i < nx; ++i) {
    for (int j = 0; j < ny; ++j) {
      const int num = j * nx + i;
      const double *m = &mom[(size_t)num * MCPM_NMOM];
      const double *sm = &smp[(size_t)num * MCPM_NSAMP];
      const double *q = &cl[(size_t)num * MCPM_NCLOS];
      const double n = nd[num], T = m[29];
      const double prep = n * p.kb * T * (1.0 + n * p.b * reference_y(n, p.b));
      double v[POST_NCOL];
      int k = 0;
      v[k++] = p.dims == 2 ? (i + 0.5) * dx : 0.5 * s->box_len[0];
      v[k++] = (j + 0.5) * dy;
      v[k++] = dx;
      v[k++] = dy;
      v[k++] = dz;
      v[k++] = t_post;
      v[k++] = p.delta_t;
      v[k++] = m[0];
      v[k++] = m[1];
      v[k++] = m[2];
      v[k++] = T;
      v[k++] = n;
      for (int t = 0; t < 20; ++t) v[k++] = sm[t];  // sum_weight, sum_M, sum_MM, sum_MMM
      for (int t = 0; t < 50; ++t) v[k++] = 0.0;    // face sums of the inverted problem
      v[k++] = p.cell_volume;
      v[k++] = ns[num];
      v[k++] = (double)cnt[num];  // printed as an integer
      v[k++] = prep - 2.0 * p.b * n * n * p.phi;
      v[k++] = p.n_ratio * p.Fn;
      for (int t = 20; t < 26; ++t) v[k++] = sm[t];  // sum_MMMM
      v[k++] = 0.0;                                  // model (printed as an integer)
      v[k++] = 0.0;                                  // xx
      v[k++] = m[6];                                 // PIJ[3]
      v[k++] = m[10] + m[15] + m[17];                // Q[1]
      v[k++] = p.dims == 1 ? F2[num] : 0.0;
      v[k++] = q[26];                                // div U
      for (int t = 0; t < 4; ++t) v[k++] = 0.0;      // UU, UU0, UUR, UU0R
      for (int t = 10; t < 26; ++t) v[k++] = q[t];   // alpha 1-9, beta 1-3, gamma_st, phi
      v[k++] = pm[num];
      v[k++] = prep;
      if (k != POST_NCOL) {
        fclose(fp);
        return MCPM_ERR_STATE;
      }
      for (int t = 0; t < POST_NCOL; ++t) {
        if (t == 84 || t == 93) fprintf(fp, "%20d", (int)v[t]);  // N, model
        else fprintf(fp, "%20e", v[t]);
      }
      fprintf(fp, "\n");
    }
    fprintf(fp, "\n");
  }
  fclose(fp);
  return MCPM_OK;
}
