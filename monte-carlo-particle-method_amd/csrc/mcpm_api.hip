// mcpm_api.hip — host layer of libmcpm.so: the context, device memory,
// launch sequencing and the extern "C" entry points declared in mcpm.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "mcpm_internal.h"
#include "mcpm_kernels.h"
#include "mcpm_plane.h"

using namespace mcpm;

namespace {

const double PI = 3.14159265358979323846;

enum {
  T_COLLIDE = 0,
  T_SORT = 1,
  T_GEN = 2,
  T_FIELD = 3,
  T_STEP = 4,
  T_FPVEL = 5,   // Fokker-Planck velocity update (k_fp_velocity)
  T_FPCLOS = 6,  // Fokker-Planck moments and closures
  T_NCLASS = 7
};

}  // namespace

struct mcpm_ctx {
  mcpm_params p;
  KParams kp;
  int device = 0;
  hipStream_t s = nullptr;
  int nc = 0;
  int64_t n = 0;
  bool sorted = false;
  bool kick_pending = false;
  size_t rec = 0;
  void *P[2] = {nullptr, nullptr};
  int cur = 0;
  // the slab's sorted positions y (record precision), written by the sorts
  // beside the records; valid while no other kernel has moved or reordered
  // a particle since
  void *ys = nullptr;
  bool ys_valid = false;
  // per-cell weights (device, nc; null: every cell has p.n_ratio), the
  // particles' weights (ping-pong like P, carried by the sorts while
  // track_w), and the work arrays of mcpm_addremove
  uint32_t *scan_tmp = nullptr;  // tile sums and offsets of the long scans
  int scan_tiles = 0;             // the most tiles they hold
  int32_t *cratio = nullptr;
  int32_t *pw[2] = {nullptr, nullptr};
  bool track_w = false;
  bool w_pending = false;  // weights from the cells at the next sort
  uint8_t *ar_fl = nullptr;
  uint32_t *ar_cst = nullptr, *ar_tot = nullptr, *ar_ostart = nullptr;
  double *ar_h = nullptr;
  // windowed sort (particles a few cells from where they were): chunk bounds
  // and replicas (host-built when the particle counts change), windows, table
  // (chunks from WSPAN-cell blocks of the previous sort's cells, built on the
  // device per sort; the blocks are fixed by the mesh)
  uint32_t *wbeg = nullptr, *rchunk = nullptr, *wpieces = nullptr, *wpbase = nullptr;
  int32_t *wrep = nullptr, *wlo = nullptr, *wlen = nullptr;
  int32_t *wbcell = nullptr, *wbrep = nullptr, *wrblk = nullptr;
  uint32_t *wtab = nullptr, *wflag = nullptr;  // wflag[0] overflow, wflag[1] chunk count
  int64_t wcap = 0;        // chunks the arrays hold
  int nwblk = 0;           // cell blocks
  bool order_sorted = false;  // the particles are in the cell order of their positions
  int w_fallbacks = 0;     // windowed sorts redone by the full histogram in a row
  bool w_enabled = true;
  int wdelta = WDELTA;     // slack cells of the windows
  int64_t w_runs = 0, w_redone = 0;
  // particle ids carried through every sort when tracking is on: ids[cur][k]
  // is the upload slot of the particle at position k
  uint32_t *ids[2] = {nullptr, nullptr};
  bool track_ids = false;
  // sort
  int64_t nchunks_cap = 0;
  int seg = 32;
  uint32_t *table = nullptr, *partial = nullptr;
  // more cells than one histogram holds: two digits of the cell key,
  // low = key % kdiv (nb_lo bins), high = key / kdiv (nb_hi bins)
  bool two_digit = false;
  int32_t kdiv = 0;
  int nb_lo = 0, nb_hi = 0;
  uint32_t *dcnt = nullptr, *dstart = nullptr, *keys = nullptr, *keys_tmp = nullptr;
  // cells
  uint32_t *cnt = nullptr, *start = nullptr, *mcand = nullptr, *cstart = nullptr;
  double *nden = nullptr, *yc = nullptr, *omax = nullptr, *omega_new = nullptr, *F2 = nullptr;
  double *moments = nullptr, *samples = nullptr;
  // measure_pressure face sums, then the two PostProc pressure columns (3 x nc)
  double *press = nullptr;
  bool pressure_on = false;
  int64_t n_samples = 0;  // cell_poly_velocity calls since the last reset
  uint32_t *counters = nullptr, *err = nullptr;
  unsigned long long *esc = nullptr;  // re-inserted particles (REINJECT boundary)
  double *energy_part = nullptr;      // per-workgroup sums of the energy check
  // reservations: the priority tables of even and odd rounds (capacity words
  // each), then the shared-particle marks (capacity bytes), one allocation
  uint32_t *rtab = nullptr;
  uint8_t *cflag = nullptr;
  uint32_t round_base = 0;  // rounds run so far: round marks are its low byte
  // collisions
  uint4 *opA = nullptr, *opB = nullptr;  // one allocation: ping-pong op lists,
  uint4 *opBand = nullptr;               // then the priority-band lists
  uint32_t *gen_table = nullptr;         // per-chunk list sizes of the candidate draw
  uint32_t *gen_cells = nullptr;         // each chunk's first and last cell (in gen_table)
  uint32_t gen_tstride = 0;              // gen_table[q gen_tstride + chunk]: list q's counts
  int32_t gen_reach = -1;                // cells a partner can lie from its candidate's (-1: any)
  int nbands = 12;
  // round termination: round r + 1 writes the losers left after round r into
  // pinned host memory, which the host polls without touching the stream
  uint32_t *hleft = nullptr, *dleft = nullptr;
  int round_grid = 1024;
  int round_unroll = 0;    // k_round's operations per thread: 0 = from the band size, 2 or 4
  int sort_chunk = 0;
  OpLog log{};                           // one allocation: debug log
  int64_t op_cap = 0;
  bool debug_log = false;
  int64_t last_ops = 0, last_log = 0;
  // Fokker-Planck / DSMC
  FPParams fpp{};
  double *clos = nullptr, *drift = nullptr;
  bool have_closures = false;
  // direct and density-expansion long range
  double *lr_tab = nullptr, *lr_V = nullptr;
  double *kde = nullptr;  // n_smooth, dn2, dn3
  // batched replicas (kp.nrep > 0): per-replica constants on the host and
  // device, cell/particle/candidate offsets, per-replica field factors
  std::vector<KParams> rkp_host;
  std::vector<int32_t> rcell_host;
  std::vector<int64_t> rpart_host;
  std::vector<double> rep_ngas;
  KParams *rkp = nullptr;
  int32_t *cell_rep = nullptr, *rcell_off = nullptr;
  int64_t *rpart_off = nullptr;
  int32_t *rpart_bucket = nullptr;  // replica of every 2^REP_BUCKET_SHIFT-th particle
  uint32_t *rprio = nullptr, *rcand_off = nullptr;
  FieldArrays *rfa = nullptr;
  // field
  FieldArrays fa{};
  double *field_buf = nullptr;
  // the plane: nodal density, nodal field, CG work arrays, quadrature tables
  FemParams fem{};
  uint32_t *q4 = nullptr, *fem_iters = nullptr;
  bool keys_current = false;  // two-digit sort keys (with quarters) match the positions
  double *cg_part = nullptr, *cg_st = nullptr;  // multi-workgroup CG partials and scalars
  double *fem_buf = nullptr;  // n_node, pot, r, p, q, then the two tables
  double *n_node = nullptr, *pot = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_q = nullptr;
  double *tab_sp = nullptr, *tab_vl = nullptr;
  // periodic power-of-two planes: the solve by FFT (row and column twiddles,
  // cosines of the two axes, one complex work array)
  bool fft_ok = false;
  int fft_lgx = 0, fft_lgy = 0;
  double2 *fft_buf = nullptr, *fft_twx = nullptr, *fft_twy = nullptr;
  double *fft_cosx = nullptr, *fft_cosy = nullptr;
  double kick_coef = 0.0;
  // field surrogate of a periodic screened-Poisson plane (mcpm_set_field_surrogate)
  mcpm_fno *fno = nullptr;
  double fno_accept = 1e-3;
  int64_t fno_every = 50;
  int64_t cur_step = 0;
  mcpm_fno_field_stats fno_last{};
  // graph-replayed steps: the run state on the device, the captured step for
  // each parity of cur, and what the capture assumed
  StepDev *sd = nullptr;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int64_t g_cap = 0, g_n = -1;
  int64_t last_m = 0;  // candidates of the last collision call
  int g_rounds = 0;
  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[T_NCLASS];
  std::vector<hipEvent_t> ev_pool;
};

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? MCPM_OK : MCPM_ERR_HIP; }

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "mcpm: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(e_), \
              __FILE__, __LINE__);                                                 \
      return MCPM_ERR_HIP;                                                         \
    }                                                                              \
  } while (0)

#define LAUNCHCHK()                                                                \
  do {                                                                             \
    hipError_t e_ = hipGetLastError();                                             \
    if (e_ != hipSuccess) {                                                        \
      fprintf(stderr, "mcpm: launch failed: %s (%s:%d)\n", hipGetErrorString(e_), \
              __FILE__, __LINE__);                                                 \
      return MCPM_ERR_HIP;                                                         \
    }                                                                              \
  } while (0)

// Call FN<R>(args) with R the context's particle record: the slab's Rec or
// the plane's Rec2, FP32 or FP64.
#define MCPM_DISPATCH(c, FN, ...)                                                    \
  ((c)->p.dims == 2                                                                  \
       ? ((c)->p.precision == MCPM_FP32 ? FN<Rec2<float>>(__VA_ARGS__)                \
                                        : FN<Rec2<double>>(__VA_ARGS__))              \
       : ((c)->p.precision == MCPM_FP32 ? FN<Rec<float>>(__VA_ARGS__)                 \
                                        : FN<Rec<double>>(__VA_ARGS__)))

template <typename T>
int dalloc(T **ptr, size_t count) {
  *ptr = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void **)ptr, count * sizeof(T)));
  return MCPM_OK;
}

hipEvent_t ev_get(mcpm_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct Timer {
  mcpm_ctx *c;
  int cls;
  hipEvent_t a = nullptr;
  Timer(mcpm_ctx *c_, int cls_) : c(c_), cls(cls_) {
    if (c->timing) {
      a = ev_get(c);
      if (a) hipEventRecord(a, c->s);
    }
  }
  ~Timer() {
    if (c->timing && a) {
      hipEvent_t b = ev_get(c);
      if (b) {
        hipEventRecord(b, c->s);
        c->ev[cls].push_back({a, b});
      }
    }
  }
};

// Gauss-Legendre nodes and weights on [-1,1], ascending, by Newton
// iteration on P_n (full double precision: the reference's 100-point table
// of compute_Vlasov_2D, FEM.cpp:656-657).
void gauss_legendre(int n, double *x, double *w) {
  // long double Newton iteration: the weights near the interval ends are
  // sensitive to the last bits of the nodes
  for (int i = 0; i < n; ++i) {
    long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    long double pnm1 = 0.0L;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int k = 1; k <= n; ++k) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * k - 1.0L) * z * p2 - (k - 1.0L) * p3) / k;
      }
      // p1 = P_n(z), p2 = P_{n-1}(z)
      long double pp = n * (z * p1 - p2) / (z * z - 1.0L);
      long double z1 = z;
      z = z1 - p1 / pp;
      pnm1 = p2;
      if (fabsl(z - z1) < 1e-19L) break;
    }
    // ascending order: node i counts from -1; w = 2 (1 - z^2) / (n P_{n-1})^2
    x[n - 1 - i] = (double)z;
    w[n - 1 - i] = (double)(2.0L * ((1.0L - z) * (1.0L + z)) / ((n * pnm1) * (n * pnm1)));
  }
}

// The same rule rounded to 8 decimals: the reference tabulates its 50-point
// rule so (Main_Functions.cpp:3328-3345, FEM.cpp:366-380) and parity needs
// the same numbers.
void gauss_legendre_8dec(int n, double *x, double *w) {
  gauss_legendre(n, x, w);
  for (int i = 0; i < n; ++i) {
    x[i] = nearbyint(x[i] * 1e8) / 1e8;
    w[i] = nearbyint(w[i] * 1e8) / 1e8;
  }
}

// Factor the constant tridiagonal matrix of the 1D field solve.
// LU factors of one slab's face system (and, periodic, its Sherman-Morrison
// vector), copied to the device buffer b of 9 (nf + 1) doubles; fa's face
// pointers are set into b.
int field_factors(int nc, double dy, bool periodic, double bst, double *b, FieldArrays *fa) {
  const int nf = periodic ? nc : nc + 1;
  const double diag = 2.0 + pow(dy * bst, 2.0);  // Main_Functions.cpp:3375-3377
  std::vector<double> a(nf, diag), lo(nf, -1.0), up(nf, -1.0);  // lo[k]=A[k+1][k], up[k]=A[k][k+1]
  double vlast = 0.0, den = 1.0;
  std::vector<double> z;
  if (!periodic) {
    a[0] = 1.0;
    up[0] = 0.0;  // c[0] = 0 (:3448-3449)
    a[nf - 1] = 1.0;
    lo[nf - 2] = 0.0;  // b[nf-2] = 0 (:3452-3454)
  } else {
    // cyclic: corners A[0][n-1] = beta = -1, A[n-1][0] = alpha = -1
    double gamma = -a[0], alpha = -1.0, beta = -1.0;
    a[0] -= gamma;
    a[nf - 1] -= alpha * beta / gamma;
    vlast = beta / gamma;
    (void)alpha;
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
  if (periodic) {
    double gamma = -(2.0 + pow(dy * bst, 2.0));
    double alpha = -1.0;
    std::vector<double> u(nf, 0.0), yv(nf), zz(nf);
    u[0] = gamma;
    u[nf - 1] = alpha;
    yv[0] = u[0];
    for (int k = 1; k < nf; ++k) yv[k] = u[k] + mL[k] * yv[k - 1];
    zz[nf - 1] = yv[nf - 1] * invd[nf - 1];
    for (int k = nf - 2; k >= 0; --k) zz[k] = yv[k] * invd[k] + mU[k] * zz[k + 1];
    z = zz;
    den = 1.0 + zz[0] + vlast * zz[nf - 1];
  }
  const size_t nfa = (size_t)nf + 1;
  fa->nj = b; b += nfa;
  fa->rhs = b; b += nfa;
  fa->tB = b; b += nfa;
  fa->y = b; b += nfa;
  fa->x = b; b += nfa;
  fa->mL = b; b += nfa;
  fa->invd = b; b += nfa;
  fa->mU = b; b += nfa;
  fa->sm_z = b; b += nfa;
  fa->sm_vlast = vlast;
  fa->sm_den = den;
  HIPCHK(hipMemcpy(fa->mL, mL.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(fa->invd, invd.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(fa->mU, mU.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  if (periodic) HIPCHK(hipMemcpy(fa->sm_z, z.data(), nf * sizeof(double), hipMemcpyHostToDevice));
  return MCPM_OK;
}

// The 1D field set-up: face-system factors (one slab, or every replica of a
// batch) and the 50-point Dirichlet quadrature tables.
int build_field(mcpm_ctx *c) {
  const bool periodic = c->p.boundary == MCPM_BC_PERIODIC;
  const int nrep = c->kp.nrep;
  std::vector<size_t> off(std::max(nrep, 1) + 1, 0);
  for (int r = 0; r < std::max(nrep, 1); ++r) {
    const int n = nrep ? c->rkp_host[r].n_cells : c->nc;
    off[r + 1] = off[r] + 9 * ((size_t)(periodic ? n : n + 1) + 1);
  }
  const size_t nbuf = off.back();
  int st = dalloc(&c->field_buf, nbuf + 250);
  if (st) return st;
  std::vector<FieldArrays> rfa(std::max(nrep, 1));
  for (int r = 0; r < std::max(nrep, 1); ++r) {
    const int n = nrep ? c->rkp_host[r].n_cells : c->nc;
    const double dy = nrep ? c->rkp_host[r].dy : c->kp.dy;
    if ((st = field_factors(n, dy, periodic, c->p.bst, c->field_buf + off[r], &rfa[r]))) return st;
  }
  double *g = c->field_buf + nbuf;
  double xi[50], wi[50], r[50], dr[50], ph[50], dph[50], cph[50], sph[50];
  gauss_legendre_8dec(50, xi, wi);
  const double bb = 3.0 * c->p.eqdist, aa = 0.0, cc = 0.0, dd = PI;
  for (int i = 0; i < 50; ++i) {
    r[i] = (bb - aa) / 2.0 * xi[i] + (aa + bb) / 2.0;
    dr[i] = wi[i] * (bb - aa) / 2.0;
    ph[i] = (dd - cc) / 2.0 * xi[i] + (cc + dd) / 2.0;
    dph[i] = wi[i] * (dd - cc) / 2.0;
    cph[i] = cos(ph[i]);
    sph[i] = sin(ph[i]);
  }
  for (FieldArrays &fa : rfa) {
    fa.gl_r = g;
    fa.gl_dr = g + 50;
    fa.gl_cos = g + 100;
    fa.gl_sin = g + 150;
    fa.gl_dphi = g + 200;
  }
  HIPCHK(hipMemcpy(g, r, sizeof r, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g + 50, dr, sizeof dr, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g + 100, cph, sizeof cph, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g + 150, sph, sizeof sph, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g + 200, dph, sizeof dph, hipMemcpyHostToDevice));
  c->fa = rfa[0];
  if (nrep) {
    if ((st = dalloc(&c->rfa, (size_t)nrep))) return st;
    HIPCHK(hipMemcpy(c->rfa, rfa.data(), nrep * sizeof(FieldArrays), hipMemcpyHostToDevice));
  }
  return MCPM_OK;
}

// Quadrature table of the direct Sutherland integral
// (vlasov_taking_the_integral, Main_Functions.cpp:3705-3729): r in
// [eqdist, 3 eqdist], polar angle in [0, pi], 8-decimal 50-point rule.
int build_vlasov_table(mcpm_ctx *c) {
  double xi[50], wi[50], tab[250];
  gauss_legendre_8dec(50, xi, wi);
  const double b = 3.0 * c->p.eqdist, a = 1.0 * c->p.eqdist, cc = 0.0, d = PI;
  for (int i = 0; i < 50; ++i) {
    double rr = (b - a) / 2.0 * xi[i] + (a + b) / 2.0;
    double dr = wi[i] * (b - a) / 2.0;
    double ph = (d - cc) / 2.0 * xi[i] + (cc + d) / 2.0;
    double q = c->p.sigma / rr;
    double q2 = q * q;
    tab[i] = rr;
    tab[50 + i] = -(-(q2 * q2 * q2)) * rr * rr * (2.0 * PI) * dr;
    tab[100 + i] = cos(ph);
    tab[150 + i] = sin(ph);
    tab[200 + i] = wi[i] * (d - cc) / 2.0;
  }
  int st = dalloc(&c->lr_tab, 250 + (size_t)c->nc + 1);
  if (st) return st;
  c->lr_V = c->lr_tab + 250;
  HIPCHK(hipMemcpy(c->lr_tab, tab, sizeof tab, hipMemcpyHostToDevice));
  return MCPM_OK;
}

// The plane's field set-up: deposit and solver constants, node arrays and
// the quadrature tables of the wall Green's function (50^3 points, r in
// [0, 3 sigma], FEM.cpp:344-394) and of the direct integral (100^3 points,
// r in [sigma, 3 sigma], FEM.cpp:640-672).
int build_fem(mcpm_ctx *c) {
  const mcpm_params &p = c->p;
  FemParams &f = c->fem;
  f.nx = p.n_cells_x;
  f.ny = p.n_cells;
  f.per_x = p.boundary_x == MCPM_BC_PERIODIC;
  f.per_y = p.boundary == MCPM_BC_PERIODIC;
  f.fnv = p.Fn / p.cell_volume;
  f.lam2V = pow(p.bst, 2.0) * p.cell_volume;
  f.volume = p.cell_volume;
  // FEM.cpp:497-502
  const double ls = p.sigma * p.bst;
  double d = 1.0 - (ls + 1.0) * exp(-p.bst * p.sigma);
  f.core0 = p.bst > 0.0 ? d / p.bst / p.bst : 0.0;
  double d2 = (-1.0 / 6.0 * pow(p.bst, 3) * pow(p.sigma, 3) - 1.0 / 2.0 * pow(p.bst, 2) * pow(p.sigma, 2) -
               p.bst * p.sigma + exp(p.bst * p.sigma) - 1) *
              exp(-p.bst * p.sigma) / pow(p.bst, 4);
  f.core2 = p.bst > 0.0 ? d2 / pow(c->kp.dx, 2) : 0.0;
  f.tol2 = 1e-26;
  f.maxit = 20000;
  f.nq = 0;
  const size_t nn = (size_t)(f.nx + 1) * (f.ny + 1);
  int st = dalloc(&c->fem_buf, nn * 5 + 2 * QT_N * 128 + 2 * CG_GRID + 4);
  if (st) return st;
  auto pow2 = [](int v) { return v >= 4 && v <= 4096 && (v & (v - 1)) == 0; };
  if (f.per_x && f.per_y && pow2(f.nx) && pow2(f.ny) && p.long_range == MCPM_LR_SCREENED_POISSON) {
    const int nx = f.nx, ny = f.ny;
    if ((st = dalloc(&c->fft_buf, (size_t)nx * ny))) return st;
    if ((st = dalloc(&c->fft_twx, (size_t)nx / 2))) return st;
    if ((st = dalloc(&c->fft_twy, (size_t)ny / 2))) return st;
    if ((st = dalloc(&c->fft_cosx, (size_t)nx))) return st;
    if ((st = dalloc(&c->fft_cosy, (size_t)ny))) return st;
    // angles reduced exactly in integers, then FP64 cos / sin
    auto tables = [&](int n, double2 *tw, double *cs) -> int {
      std::vector<double2> t(n / 2);
      std::vector<double> co(n);
      for (int k = 0; k < n; ++k) {
        const double a = 2.0 * PI * (double)k / (double)n;
        co[k] = std::cos(a);
        if (k < n / 2) t[k] = make_double2(std::cos(a), -std::sin(a));
      }
      HIPCHK(hipMemcpy(tw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(cs, co.data(), co.size() * sizeof(double), hipMemcpyHostToDevice));
      return MCPM_OK;
    };
    if ((st = tables(nx, c->fft_twx, c->fft_cosx))) return st;
    if ((st = tables(ny, c->fft_twy, c->fft_cosy))) return st;
    c->fft_lgx = __builtin_ctz((unsigned)nx);
    c->fft_lgy = __builtin_ctz((unsigned)ny);
    c->fft_ok = true;
  }
  c->n_node = c->fem_buf;
  c->pot = c->n_node + nn;
  c->cg_r = c->pot + nn;
  c->cg_p = c->cg_r + nn;
  c->cg_q = c->cg_p + nn;
  c->tab_sp = c->cg_q + nn;
  c->tab_vl = c->tab_sp + QT_N * 128;
  c->cg_part = c->tab_vl + QT_N * 128;
  c->cg_st = c->cg_part + 2 * CG_GRID;
  if ((st = dalloc(&c->q4, (size_t)c->nc * 4 + 2))) return st;
  c->fem_iters = c->q4 + (size_t)c->nc * 4;
  HIPCHK(hipMemset(c->fem_buf, 0, nn * 5 * sizeof(double)));
  for (int which = 0; which < 2; ++which) {
    const int nq = which == 0 ? 50 : 100;
    std::vector<double> xi(nq), wi(nq), t((size_t)QT_N * nq);
    if (which == 0) gauss_legendre_8dec(nq, xi.data(), wi.data());
    else gauss_legendre(nq, xi.data(), wi.data());
    const double aa = which == 0 ? 0.0 : p.sigma, bb = 3.0 * p.sigma;
    const double cc = 0.0, dd = 2 * PI, ee = 0.0, ff = PI;
    for (int k = 0; k < nq; ++k) {
      const double rr = (bb - aa) / 2.0 * xi[k] + (aa + bb) / 2.0;
      const double dr = wi[k] * (bb - aa) / 2.0;
      const double ph = (dd - cc) / 2.0 * xi[k] + (cc + dd) / 2.0;
      const double th = (ff - ee) / 2.0 * xi[k] + (ff + ee) / 2.0;
      t[QT_R * nq + k] = rr;
      t[QT_RAD * nq + k] = which == 0 ? exp(-p.bst * rr) / (4.0 * PI) * rr * dr      // FEM.cpp:216
                                      : pow(p.sigma / rr, 6.0) * rr * rr * dr;      // FEM.cpp:166
      t[QT_CPH * nq + k] = cos(ph);
      t[QT_SPH * nq + k] = sin(ph);
      t[QT_DPH * nq + k] = wi[k] * (dd - cc) / 2.0;
      t[QT_STH * nq + k] = sin(th);
      t[QT_DTH * nq + k] = wi[k] * (ff - ee) / 2.0;
    }
    HIPCHK(hipMemcpy(which == 0 ? c->tab_sp : c->tab_vl, t.data(), t.size() * sizeof(double),
                     hipMemcpyHostToDevice));
  }
  return MCPM_OK;
}

int check_err(mcpm_ctx *c) {
  uint32_t e = 0;
  HIPCHK(hipMemcpyAsync(&e, c->err, sizeof e, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  if (e == 0) return MCPM_OK;
  HIPCHK(hipMemsetAsync(c->err, 0, sizeof(uint32_t), c->s));
  if (e & ERRBIT_DOMAIN) return MCPM_ERR_DOMAIN;
  if (e & ERRBIT_NAN) return MCPM_ERR_NAN;
  if (e & ERRBIT_BC) return MCPM_ERR_BC;
  if (e & ERRBIT_ORDER) return MCPM_ERR_STATE;
  return MCPM_ERR_CAPACITY;
}

int chunk_for(int64_t n, int fixed) {
  if (fixed > 0) return fixed;
  // about 2048 chunks (one wavefront each), between 1024 and 32768 particles
  int64_t ch = (n + 2047) / 2048;
  ch = (ch + 255) / 256 * 256;
  if (ch < 1024) ch = 1024;
  if (ch > 32768) ch = 32768;
  return (int)ch;
}

// Whether the windowed sort applies: a slab (or batched slabs) whose
// particles are still in the cell order of their positions, moved by a push
// that adds nothing per step (no pressure face sums, no re-insertion), dense
// enough that a chunk spans a few cells.
// Exclusive scan of n counts into out[0..n] (out[n] the total): one
// workgroup for short arrays, tiles over many workgroups for long ones.
int excl_scan(mcpm_ctx *c, const uint32_t *in, uint32_t *out, int n,
              const StepDev *sd = nullptr) {
  if (n <= 16384) {
    hipLaunchKernelGGL(k_excl_scan, dim3(1), dim3(1024), 0, c->s, in, out, n, sd);
  } else {
    const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt > c->scan_tiles) return MCPM_ERR_CAPACITY;
    uint32_t *tsum = c->scan_tmp, *toff = c->scan_tmp + nt + 1;
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)nt), dim3(1024), 0, c->s, in, n, tsum, sd);
    hipLaunchKernelGGL(k_excl_scan, dim3(1), dim3(1024), 0, c->s, tsum, toff, nt, sd);
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)nt), dim3(1024), 0, c->s, in, out, n,
                       (const uint32_t *)toff, nt, sd);
  }
  LAUNCHCHK();
  return MCPM_OK;
}

// After a sort: particles whose weights are still to be set (new particles
// with per-cell weights in force) take their cells' (main.cpp:239-242).
int weights_after_sort(mcpm_ctx *c) {
  if (!c->kp.cell_ratio || !c->w_pending) return MCPM_OK;
  hipLaunchKernelGGL(k_w_from_cells, dim3((unsigned)c->nc), dim3(256), 0, c->s, c->start, c->nc,
                     (const int32_t *)c->cratio, c->pw[c->cur]);
  LAUNCHCHK();
  c->w_pending = false;
  c->track_w = true;
  return MCPM_OK;
}

// The sorted y array serves only the candidate draw: a sort under a model
// that draws no candidates leaves it unwritten (and not valid).
template <typename real>
real *ys_out(const mcpm_ctx *c) {
  switch (c->p.model) {
    case MCPM_MODEL_FP_DENSE:
    case MCPM_MODEL_FP_IDEAL:
    case MCPM_MODEL_FP_LINEAR:
    case MCPM_MODEL_NONE:
    case MCPM_MODEL_CBA_HS: return nullptr;
    default: return (real *)c->ys;
  }
}

bool window_ok(const mcpm_ctx *c, int mode) {
  if (!c->w_enabled || c->p.dims == 2 || !c->wbeg || !c->order_sorted || c->w_fallbacks >= 3)
    return false;
  if (c->p.boundary == MCPM_BC_REINJECT) return false;
  if (c->pressure_on && (mode & PUSH_MOVE)) return false;
  if (c->p.model == MCPM_MODEL_CBA_VHS || c->p.model == MCPM_MODEL_CBA_HS) return false;
  return c->n >= 16 * (int64_t)c->nc;
}

// The launches of a windowed sort (with a captured step's state when sd is
// given: its step number, and the flag cleared by a kernel, not a memset).
template <typename R>
int wsort_launch(mcpm_ctx *c, int mode, int64_t step, const Kick &kick, const double *drift,
                 const StepDev *sd) {
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  const R *in = (const R *)c->P[c->cur];
  R *out = (R *)c->P[1 - c->cur];
  const uint32_t *iid = c->track_ids ? c->ids[c->cur] : nullptr;
  uint32_t *oid = c->track_ids ? c->ids[1 - c->cur] : nullptr;
  const int nblk = c->nwblk;
  const int nrep1 = std::max(c->kp.nrep, 1) + 1;
  uint32_t *nchp = c->wflag + 1;
  if (sd) hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, c->s, c->wflag, (int64_t)1, 0u);
  else HIPCHK(hipMemsetAsync(c->wflag, 0, sizeof(uint32_t), c->s));
  // the chunks: pieces of the cell blocks of the previous sort (start[])
  hipLaunchKernelGGL(k_wblocks, dim3((unsigned)((nblk + 255) / 256)), dim3(256), 0, c->s,
                     c->wbcell, nblk, c->start, c->wpieces, sd);
  if (int st = excl_scan(c, c->wpieces, c->wpbase, nblk, sd)) return st;
  hipLaunchKernelGGL(k_wfill, dim3((unsigned)(nblk + (nrep1 + 255) / 256)), dim3(256), 0, c->s,
                     c->wbcell, c->wbrep, nblk, c->start, c->wpbase, c->wrblk, nrep1,
                     (uint32_t)c->n, c->wbeg, c->wrep, nchp, c->rchunk, sd);
  // the most chunks there can be; waves past the count stop at once
  const unsigned gw = (unsigned)((c->wcap + 3) / 4);
  using real = typename R::real_type;
  // the lean push: every slab periodic, no position drift, no face sums or
  // re-insertion counts to keep
  bool lean = drift == nullptr && kick.pm == nullptr && kick.esc == nullptr;
  if (c->kp.nrep) {
    for (const KParams &q : c->rkp_host) lean = lean && q.boundary == MCPM_BC_PERIODIC;
  } else {
    lean = lean && c->kp.boundary == MCPM_BC_PERIODIC;
  }
  auto count_k = lean ? k_wpush<R, false, 1> : k_wpush<R, false, 0>;
  auto scatter_k = lean ? k_wpush<R, true, 1> : k_wpush<R, true, 0>;
  hipLaunchKernelGGL(count_k, dim3(gw), dim3(256), 0, c->s, in, nullptr, c->kp, kick,
                     drift, mode, slo, shi, c->wbeg, c->wrep, nchp, c->wlo, c->wlen, c->wtab,
                     nullptr, c->err, c->wflag, nullptr, nullptr, c->wdelta, (real *)nullptr,
                     nullptr, nullptr, sd);
  // about 32k wavefronts: one cell each on short cell arrays, up to 16
  const int cpw = std::max(1, std::min(16, c->nc / 32768));
  hipLaunchKernelGGL(k_wscan, dim3((unsigned)((c->nc + 4 * cpw - 1) / (4 * cpw))),
                     dim3(256), 0, c->s, c->wlo,
                     c->wlen, c->wtab, nchp, c->rchunk, c->kp, c->cnt, c->nden, c->yc, cpw, sd);
  if (int st = excl_scan(c, c->cnt, c->start, c->nc, sd)) return st;
  hipLaunchKernelGGL(scatter_k, dim3(gw), dim3(256), 0, c->s, in, out, c->kp, kick,
                     drift, mode, slo, shi, c->wbeg, c->wrep, nchp, c->wlo, c->wlen, c->wtab,
                     c->start, c->err, c->wflag, iid, oid, c->wdelta, ys_out<real>(c),
                     c->track_w ? c->pw[c->cur] : nullptr,
                     c->track_w ? c->pw[1 - c->cur] : nullptr, sd);
  LAUNCHCHK();
  return MCPM_OK;
}

// 1: a particle left its chunk's window (sort again by the full histogram).
template <typename R>
int run_wsort_t(mcpm_ctx *c, int mode, int64_t step, const Kick &kick, const double *drift) {
  using real = typename R::real_type;
  if (int st = wsort_launch<R>(c, mode, step, kick, drift, nullptr)) return st;
  uint32_t flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, c->wflag, sizeof flag, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  ++c->w_runs;
  if (flag) {
    // wider windows for the next sorts, up to the table's width
    ++c->w_redone;
    if (c->wdelta < WDELTA_MAX) c->wdelta = std::min(2 * c->wdelta, (int)WDELTA_MAX);
    else ++c->w_fallbacks;
    return 1;
  }
  c->w_fallbacks = 0;
  c->cur = 1 - c->cur;
  c->sorted = true;
  c->order_sorted = true;
  c->ys_valid = ys_out<real>(c) != nullptr;
  return weights_after_sort(c);
}

template <typename R>
int run_sort_t(mcpm_ctx *c, int mode, int64_t step) {
  using real = typename R::real_type;
  c->keys_current = false;
  c->ys_valid = false;
  if (c->n == 0) {
    HIPCHK(hipMemsetAsync(c->cnt, 0, c->nc * sizeof(uint32_t), c->s));
    HIPCHK(hipMemsetAsync(c->start, 0, (c->nc + 1) * sizeof(uint32_t), c->s));
    c->sorted = true;
    return MCPM_OK;
  }
  Timer tm(c, T_SORT);
  const int nc = c->nc;
  const int chunk = chunk_for(c->n, c->sort_chunk);
  const int64_t nch = (c->n + chunk - 1) / chunk;
  if (nch > c->nchunks_cap) return MCPM_ERR_CAPACITY;
  const int nseg = (int)((nch + c->seg - 1) / c->seg);
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  const Kick kick{c->F2, c->pot, c->kick_coef, nullptr};
  // the face sums are taken once, in the scatter pass, of a moving push
  Kick kick_scatter = kick;
  if (c->pressure_on && (mode & PUSH_MOVE)) kick_scatter.pm = c->press;
  if (mode & PUSH_MOVE) kick_scatter.esc = c->esc;
  const double *drift = (c->p.model == MCPM_MODEL_FP_DENSE && c->have_closures &&
                         (mode & PUSH_MOVE))
                            ? c->drift
                            : nullptr;
  if constexpr (R::dims == 1) {
    if (window_ok(c, mode)) {
      const int w = run_wsort_t<R>(c, mode, step, kick, drift);
      if (w <= 0) return w;  // done, or an error; 1: the full histogram below
    }
  }
  // one stable counting-sort pass over nb bins: histogram, scans, scatter
  auto pass = [&](const R *in, R *out, int pmode, const Kick &ks, int nb, int kmode,
                  uint32_t *cnt, uint32_t *start, double *nden, double *yc, uint32_t *okey,
                  const uint32_t *ikey, const uint32_t *iid, uint32_t *oid, real *oys,
                  const int32_t *iw, int32_t *ow) -> int {
    const size_t lds = (size_t)nb * sizeof(uint32_t);
    hipLaunchKernelGGL((k_push<R, false>), dim3((unsigned)nch), dim3(64 * PUSH_WAVES), lds,
                       c->s, in, nullptr, c->n, chunk, c->kp, kick, drift, pmode, slo, shi, c->table,
                       nullptr, c->err, nb, kmode, c->kdiv, nullptr, ikey, nullptr, nullptr,
                       (real *)nullptr, nullptr, nullptr);
    LAUNCHCHK();
    dim3 gs((nb + 255) / 256, nseg);
    hipLaunchKernelGGL(k_scan_a, gs, dim3(256), 0, c->s, c->table, c->partial, (int)nch, nb,
                       c->seg);
    hipLaunchKernelGGL(k_scan_bins, dim3((nb + 255) / 256), dim3(256), 0, c->s, c->partial,
                       nseg, nb, cnt, nden, yc, c->kp);
    if (int st = excl_scan(c, cnt, start, nb)) return st;
    hipLaunchKernelGGL(k_scan_c, gs, dim3(256), 0, c->s, c->table, c->partial, (int)nch, nb,
                       c->seg);
    hipLaunchKernelGGL((k_push<R, true>), dim3((unsigned)nch), dim3(64 * PUSH_WAVES), lds,
                       c->s, in, out,
                       c->n, chunk, c->kp, ks, drift, pmode, slo, shi, c->table, start,
                       c->err, nb, kmode, c->kdiv, okey, ikey, iid, oid, oys, iw, ow);
    LAUNCHCHK();
    return MCPM_OK;
  };
  const R *in = (const R *)c->P[c->cur];
  R *out = (R *)c->P[1 - c->cur];
  uint32_t *id_in = c->track_ids ? c->ids[c->cur] : nullptr;
  uint32_t *id_out = c->track_ids ? c->ids[1 - c->cur] : nullptr;
  int32_t *w_in = c->track_w ? c->pw[c->cur] : nullptr;
  int32_t *w_out = c->track_w ? c->pw[1 - c->cur] : nullptr;
  int st;
  if (!c->two_digit) {
    if ((st = pass(in, out, mode, kick_scatter, nc, 0, c->cnt, c->start, c->nden, c->yc,
                   nullptr, nullptr, id_in, id_out, ys_out<real>(c), w_in, w_out)))
      return st;
  } else {
    // low digit with the push (keys carried to the destinations), then the
    // high digit of those keys back into the first buffer; cells from the
    // sorted keys
    if ((st = pass(in, out, mode, kick_scatter, c->nb_lo, 1, c->dcnt, c->dstart, nullptr,
                   nullptr, c->keys_tmp, nullptr, id_in, id_out, nullptr, w_in, w_out)))
      return st;
    if ((st = pass(out, (R *)in, 0, kick, c->nb_hi, 2, c->dcnt, c->dstart, nullptr, nullptr,
                   c->keys, c->keys_tmp, id_out, id_in, nullptr, w_out, w_in)))
      return st;
    hipLaunchKernelGGL(k_fill_u32, dim3(64), dim3(256), 0, c->s, c->start, (int64_t)nc + 1, 0u);
    hipLaunchKernelGGL(k_cell_bounds, dim3(2048), dim3(256), 0, c->s, c->keys, c->n, nc,
                       c->start, c->err);
    hipLaunchKernelGGL(k_cell_fill, dim3((nc + 255) / 256), dim3(256), 0, c->s, c->start, nc,
                       c->cnt, c->nden, c->yc, c->kp);
    LAUNCHCHK();
    c->sorted = true;
    c->order_sorted = true;
    c->keys_current = true;
    return weights_after_sort(c);
  }
  c->cur = 1 - c->cur;
  c->sorted = true;
  c->order_sorted = true;
  c->ys_valid = ys_out<real>(c) != nullptr;
  return weights_after_sort(c);
}

int run_sort(mcpm_ctx *c, int mode, int64_t step) {
  return MCPM_DISPATCH(c, run_sort_t, c, mode, step);
}

int ensure_ops(mcpm_ctx *c, int64_t m) {
  if (m <= c->op_cap) return MCPM_OK;
  int64_t cap = m + m / 4 + 1024;
  HIPCHK(hipStreamSynchronize(c->s));
  // captured steps hold the old lists' addresses
  for (auto &g : c->gexec)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  if (c->opA) HIPCHK(hipFree(c->opA));
  if (c->log.p) HIPCHK(hipFree(c->log.p));
  c->opA = c->opB = c->opBand = nullptr;
  c->log = OpLog{};
  c->op_cap = 0;
  uint4 *blk = nullptr;
  int st = dalloc(&blk, (size_t)cap * 3);
  if (st) return st;
  c->opA = blk;
  c->opB = blk + cap;
  c->opBand = blk + 2 * cap;
  if (c->gen_table) HIPCHK(hipFree(c->gen_table));
  c->gen_table = nullptr;
  const size_t nch_cap = (size_t)cap / GEN_CHUNK + 2;
  const size_t tstride = (nch_cap + 7) / 8 * 8;
  if ((st = dalloc(&c->gen_table, tstride * NLIST + 2 * nch_cap))) return st;
  c->gen_cells = c->gen_table + tstride * NLIST;
  c->gen_tstride = (uint32_t)tstride;
  uint32_t *lg = nullptr;
  st = dalloc(&lg, (size_t)cap * 3 + ((size_t)cap + 3) / 4);
  if (st) return st;
  c->log.p = lg;
  c->log.i = lg + cap;
  c->log.j = lg + 2 * cap;
  c->log.a = (uint8_t *)(lg + 3 * cap);
  c->op_cap = cap;
  return MCPM_OK;
}

// k_round's operations per thread: the knob's value, else (ESMC) 4 when a
// band's operations fit in one pass of the round grid at 2 per thread (the
// whole grid is then resident, and more operations per thread put more loads
// in flight) and 2 when they do not (fewer registers, more wavefronts per
// SIMD); DSMC runs all its operations in one band: 4.
int round_unroll_for(const mcpm_ctx *c, int model, int64_t m, int nbands, unsigned grid) {
  if (c->round_unroll) return c->round_unroll;
  if (model == COLL_DSMC) return 4;
  return m / std::max(nbands, 1) > (int64_t)grid * 256 * 2 ? 2 : 4;
}
template <typename R, int MODEL>
decltype(&k_round<R, MODEL, true, 4>) round_kernel(bool plain, int u) {
  if (plain) return u == 2 ? k_round<R, MODEL, true, 2> : k_round<R, MODEL, true, 4>;
  return u == 2 ? k_round<R, MODEL, false, 2> : k_round<R, MODEL, false, 4>;
}

// variant: for COLL_DSMC the mcpm_model run (DSMC_VHS, DSMC_HS, CBA_VHS).
template <typename R, int MODEL>
int collide_t(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats, int variant) {
  const int nc = c->nc;
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  R *P = (R *)c->P[c->cur];
  FPParams fpp = c->fpp;
  fpp.model = variant;
  hipLaunchKernelGGL(k_mcand, dim3((nc + 255) / 256), dim3(256), 0, c->s, c->cnt, c->omax,
                     c->nden, c->yc, c->mcand, nc, c->kp, (int)MODEL, variant, slo, shi,
                     (const StepDev *)nullptr);
  if (int st = excl_scan(c, c->mcand, c->cstart, nc)) return st;
  hipLaunchKernelGGL(k_zero_u32, dim3((CNT_TOTAL + 255) / 256), dim3(256), 0, c->s,
                     c->counters, (int)CNT_TOTAL);
  LAUNCHCHK();
  uint32_t M = 0;
  HIPCHK(hipMemcpyAsync(&M, c->cstart + nc, sizeof M, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  c->last_m = M;
  int st = ensure_ops(c, (int64_t)M);
  if (st) return st;
  mcpm_u32x4 pk = mcpm_philox(0u, 0u, slo, mcpm_ctr3(MCPM_TAG_PRIO, shi, 0u), c->kp.key0,
                              c->kp.key1);
  if (c->kp.nrep) {
    hipLaunchKernelGGL(k_rep_step, dim3((c->kp.nrep + 256) / 256), dim3(256), 0, c->s, c->kp,
                       c->cstart, c->rprio, c->rcand_off, slo, shi, (const StepDev *)nullptr);
    LAUNCHCHK();
  }
  // ESMC priorities spread over 32 bits and enter the rounds in bands; DSMC
  // operations of different cells never conflict, one band keeps every cell
  // busy from the first round
  // (new4's priorities are ranks in its cell order: its bands are rank ranges)
  const int nbands = MODEL != COLL_DSMC ? std::max(1, std::min<int>(c->nbands, MAX_BANDS)) : 1;
  const uint32_t band_mul = MODEL == COLL_ESMC ? (uint32_t)nbands : 0u;
  // a multiple of 8 blocks so that every XCD gets the same share
  unsigned grid = (unsigned)std::min<int64_t>(std::max<int64_t>(8, ((int64_t)M + 255) / 256),
                                              c->round_grid);
  grid = (grid + 7u) / 8u * 8u;
  // round r's priority table is rtab + (r & 1) capacity; its shared-particle
  // mark is the low byte of round_base, which advances once per round (and
  // across calls), so consecutive rounds never share a mark
  uint32_t *const tabs[2] = {c->rtab, c->rtab + (size_t)c->p.capacity};
  const int unroll = round_unroll_for(c, MODEL, M, nbands, grid);
  if (M > 0) {
    Timer tm(c, T_GEN);
    const uint32_t nchunks = (M + GEN_CHUNK - 1) / GEN_CHUNK;
    hipLaunchKernelGGL(k_cand_count, dim3(nchunks), dim3(256), 0, c->s, M, pk.v[0], (int)MODEL,
                       band_mul, nbands, c->gen_table, c->gen_tstride, c->kp, c->cstart, nc,
                       c->gen_cells,
                       (const StepDev *)nullptr);
    hipLaunchKernelGGL(k_cand_scan, dim3(nbands * NREG), dim3(1024), 0, c->s, c->gen_table,
                       c->gen_tstride, nchunks, nbands * NREG, c->counters,
                       (const uint32_t *)nullptr, (const StepDev *)nullptr);
    const typename R::real_type *ys =
        c->ys_valid ? (const typename R::real_type *)c->ys : nullptr;
    hipLaunchKernelGGL((k_gen<R, MODEL>), dim3(nchunks), dim3(256), 0, c->s, P, ys, c->cnt,
                       c->start,
                       c->cstart, nc, M, c->kp, slo, shi, pk.v[0], band_mul, nbands,
                       c->gen_table, c->gen_tstride, (const uint32_t *)c->gen_cells, c->gen_reach,
                       c->opBand,
                       tabs[0], c->counters, (const StepDev *)nullptr);
    LAUNCHCHK();
  }
  int rounds = 0;
  const uint32_t seg = (M + NREG - 1) / NREG;  // loser-list segment of one region
  {
    OpLog lg = c->debug_log ? c->log : OpLog{};
    // hleft[q] = losers after round q - 1, written by round q; unset = ~0
    volatile uint32_t *hl = c->hleft;
    for (int q = 0; q <= MAX_ROUNDS; ++q) hl[q] = 0xFFFFFFFFu;
    constexpr int LAG = 6;  // rounds the host may run ahead of the last count seen
    int polled = 1;         // next count to examine: losers after round polled - 1
    bool done = false;
    for (int r = 0; r < MAX_ROUNDS && !done; ++r) {
      {
        Timer tm(c, T_COLLIDE);
        const uint4 *in = (const uint4 *)((r & 1) ? c->opB : c->opA);
        const uint32_t mark = c->round_base & 0xFFu;
        hipLaunchKernelGGL((k_check<MODEL>), dim3(grid), dim3(256), 0, c->s, r, nbands, seg, in,
                           (const uint4 *)c->opBand, tabs[r & 1], c->cflag, mark,
                           c->counters, c->dleft, (const StepDev *)nullptr);
        // no replicas and no log: the kernel without their code (fewer registers)
        auto round_k = round_kernel<R, MODEL>(c->kp.nrep == 0 && !lg.p, unroll);
        hipLaunchKernelGGL(round_k, dim3(grid), dim3(256), 0, c->s, r, nbands, seg,
                           P, in, (const uint4 *)c->opBand, (r & 1) ? c->opA : c->opB,
                           (const uint32_t *)tabs[r & 1], tabs[(r + 1) & 1],
                           (const uint8_t *)c->cflag, mark, c->counters, c->nden,
                           c->yc, c->omax, c->omega_new, c->kp, slo, shi, fpp, lg,
                           (const StepDev *)nullptr);
      }
      c->round_base += 1u;
      LAUNCHCHK();
      rounds = r + 1;
      // examine the counts published so far (round q publishes hleft[q]);
      // wait only when the host is LAG rounds ahead
      int spins = 0;
      while (polled <= r) {
        uint32_t v = hl[polled];
        if (v == 0xFFFFFFFFu) {
          if (r - polled < LAG) break;
          // spin on the host word; ask the runtime only now and then whether
          // the stream has stopped (an error, or every launched round ran)
          if (++spins < 4096) {
            __builtin_ia32_pause();
            continue;
          }
          spins = 0;
          hipError_t q = hipStreamQuery(c->s);
          if (q != hipErrorNotReady && q != hipSuccess) return MCPM_ERR_HIP;
          if (q == hipSuccess && hl[polled] == 0xFFFFFFFFu) return MCPM_ERR_HIP;
          continue;
        }
        // every band admitted and processed (band b is processed in round b)
        if (polled >= nbands && v == 0) {
          done = true;
          break;
        }
        ++polled;
      }
    }
  }
  hipLaunchKernelGGL(k_omega_finish, dim3((nc + 255) / 256), dim3(256), 0, c->s, c->omax,
                     c->omega_new, c->cnt, nc, (int)MODEL, variant, (const StepDev *)nullptr);
  LAUNCHCHK();
  // the counters up to the loser counts of the last round
  std::vector<uint32_t> hc(CNT_LIST + (size_t)(rounds + 1) * NREG);
  HIPCHK(hipMemcpyAsync(hc.data(), c->counters, hc.size() * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  uint32_t left = 0;
  for (int x = 0; x < NREG; ++x) left += hc[CNT_LIST + (size_t)rounds * NREG + x];
  if (left != 0) return MCPM_ERR_CAPACITY;  // more rounds than MAX_ROUNDS
  int64_t revisits = 0;
  for (size_t q = CNT_LIST + NREG; q < hc.size(); ++q) revisits += hc[q];
  c->last_ops = hc[CNT_OPS];
  c->last_log = hc[CNT_LOG];
  if (stats) {
    stats->candidates = M;
    stats->valid = hc[CNT_OPS];
    uint64_t acc;
    memcpy(&acc, &hc[CNT_ACC], sizeof acc);
    stats->accepted = (int64_t)acc;
    stats->rounds = rounds;
    stats->revisits = (int32_t)std::min<int64_t>(revisits, INT32_MAX);
  }
  return MCPM_OK;
}

template <typename R>
int collide_esmc_t(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  return collide_t<R, COLL_ESMC>(c, step, stats, MCPM_MODEL_ESMC);
}
template <typename R>
int collide_esmc4_t(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  return collide_t<R, COLL_ESMC4>(c, step, stats, MCPM_MODEL_ESMC_NEW4);
}
template <typename R>
int collide_dsmc_t(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats, int variant) {
  if (variant == MCPM_MODEL_CBA_VHS) c->ys_valid = false;  // CBA displaces the pairs
  return collide_t<R, COLL_DSMC>(c, step, stats, variant);
}

template <typename R>
int launch_moments(mcpm_ctx *c);

// Collision_CBA_HS: moments (T), then one thread per cell.
template <typename R>
int collide_cba_hs_t(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  c->ys_valid = false;  // CBA displaces the pairs
  int st = launch_moments<R>(c);
  if (st) return st;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->s, c->counters, (int)CNT_LIST);
  Timer tm(c, T_COLLIDE);
  hipLaunchKernelGGL((k_cba_hs<R>), dim3((c->nc + 63) / 64), dim3(64), 0, c->s,
                     (R *)c->P[c->cur], c->cnt, c->start, c->moments, c->nden, c->kp, slo, shi,
                     c->counters, c->err);
  LAUNCHCHK();
  if ((st = check_err(c))) return st;
  uint32_t hc[CNT_LIST];
  HIPCHK(hipMemcpyAsync(hc, c->counters, sizeof hc, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  c->last_ops = 0;
  c->last_log = 0;
  if (stats) {
    stats->candidates = hc[CNT_OPS];
    stats->valid = hc[CNT_OPS];
    uint64_t acc;
    memcpy(&acc, &hc[CNT_ACC], sizeof acc);
    stats->accepted = (int64_t)acc;
    stats->rounds = 0;
    stats->revisits = 0;
  }
  return MCPM_OK;
}

template <typename R>
int launch_moments(mcpm_ctx *c) {
  hipLaunchKernelGGL((k_moments<R>), dim3(c->nc), dim3(256), 0, c->s, (const R *)c->P[c->cur],
                     c->cnt, c->start, c->moments, c->kp);
  LAUNCHCHK();
  return MCPM_OK;
}

template <typename R>
int launch_sample(mcpm_ctx *c) {
  hipLaunchKernelGGL((k_sample<R>), dim3(c->nc), dim3(256), 0, c->s, (const R *)c->P[c->cur],
                     c->cnt, c->start, c->samples);
  LAUNCHCHK();
  return MCPM_OK;
}

// Nodal density of the plane (cell_update's deposit) from the sorted particles.
template <typename R>
int launch_nodes(mcpm_ctx *c) {
  if constexpr (sizeof(R) == 8 * sizeof(typename R::real_type)) {
    if (c->two_digit && c->keys_current)  // the sort keys carry each particle's quarter
      hipLaunchKernelGGL(k_quadrants_keys, dim3((c->nc + 3) / 4), dim3(256), 0, c->s, c->keys,
                         c->cnt, c->start, c->q4, c->nc);
    else
      hipLaunchKernelGGL((k_quadrants<R>), dim3(c->nc), dim3(256), 0, c->s,
                         (const R *)c->P[c->cur], c->cnt, c->start, c->q4, c->kp);
    const int nn = (c->fem.nx + 1) * (c->fem.ny + 1);
    hipLaunchKernelGGL(k_nodes, dim3((nn + 255) / 256), dim3(256), 0, c->s, c->q4, c->n_node,
                       c->fem);
    LAUNCHCHK();
    return MCPM_OK;
  } else {
    return MCPM_ERR_ARG;
  }
}

// The FEM solve of a large plane over CG_GRID workgroups.  Iterations are
// launched in batches; between batches the host reads the convergence flag.
constexpr int CG_ONE_WG_NODES = 129 * 129;

int fem_cg_grid(mcpm_ctx *c, const FemParams &f, int warm = 0, int mode = 0,
                double accept = 0.0, double *res0 = nullptr, int *iters = nullptr) {
  const dim3 g(CG_GRID), b(256);
  if (!warm) hipLaunchKernelGGL(k_cgm_zero, g, b, 0, c->s, c->pot, f);
  hipLaunchKernelGGL(k_cgm_start, g, b, 0, c->s, c->n_node, c->pot, c->cg_r, c->cg_p, f,
                     c->cg_part);
  hipLaunchKernelGGL(k_cgm_start2, dim3(1), b, 0, c->s, c->cg_st, c->cg_part, CG_GRID, f,
                     c->fem_iters);
  LAUNCHCHK();
  if (res0) *res0 = -1.0;
  if (iters) *iters = 0;
  if (warm) {
    // starting residual |M n - A pot| / |M n| (one host read)
    double sc[4];
    HIPCHK(hipMemcpyAsync(sc, c->cg_st, sizeof sc, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    const double r = sc[3] > 0.0 ? std::sqrt(sc[0] / sc[3]) : (sc[0] > 0.0 ? INFINITY : 0.0);
    if (res0) *res0 = r;
    if (mode == 2 || (mode == 1 && r <= accept)) {
      hipLaunchKernelGGL(k_cgm_seams, g, b, 0, c->s, c->pot, f);
      LAUNCHCHK();
      return MCPM_OK;
    }
  }
  int k = 0;
  uint32_t flag[2] = {0u, 0u};
  while (k < f.maxit) {
    const int batch = std::min(32, f.maxit - k);
    for (int t = 0; t < batch; ++t, ++k) {
      hipLaunchKernelGGL(k_cgm_aq, g, b, 0, c->s, c->cg_p, c->cg_q, f, c->cg_part,
                         c->fem_iters);
      hipLaunchKernelGGL(k_cgm_xr, g, b, 0, c->s, c->pot, c->cg_r, c->cg_p, c->cg_q, f,
                         c->cg_part, c->cg_st, k, c->fem_iters);
      hipLaunchKernelGGL(k_cgm_p, g, b, 0, c->s, c->cg_p, c->cg_r, f, c->cg_part, c->cg_st, k,
                         c->fem_iters);
    }
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(flag, c->fem_iters, sizeof flag, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    if (flag[1]) break;
  }
  if (iters) *iters = (int)flag[0];
  hipLaunchKernelGGL(k_cgm_seams, g, b, 0, c->s, c->pot, f);
  LAUNCHCHK();
  return MCPM_OK;
}

// The periodic power-of-two plane's Q1 solve by FFT (k_fft_rows / k_fft_cols):
// the exact solution of the system the conjugate gradients iterate on.
int fem_fft(mcpm_ctx *c) {
  const FemParams &f = c->fem;
  const int cols = std::max(1, std::min(4, 4096 / f.ny));
  hipLaunchKernelGGL((k_fft_rows<false>), dim3(f.ny), dim3(256), (size_t)f.nx * sizeof(double2),
                     c->s, c->n_node, c->fft_buf, c->pot, c->fft_twx, c->fft_lgx, f);
  hipLaunchKernelGGL(k_fft_cols, dim3(f.nx / cols), dim3(256),
                     (size_t)cols * f.ny * sizeof(double2), c->s, c->fft_buf, c->fft_twy,
                     c->fft_lgy, c->fft_cosx, c->fft_cosy, cols, f);
  hipLaunchKernelGGL((k_fft_rows<true>), dim3(f.ny), dim3(256), (size_t)f.nx * sizeof(double2),
                     c->s, c->n_node, c->fft_buf, c->pot, c->fft_twx, c->fft_lgx, f);
  LAUNCHCHK();
  return MCPM_OK;
}

// compute_SP_2D (FEM.cpp:249-610) up to the nodal field; the kick is the push's.
int sp_2d(mcpm_ctx *c) {
  int st = MCPM_DISPATCH(c, launch_nodes, c);
  if (st) return st;
  FemParams f = c->fem;
  const int nn = (f.nx + 1) * (f.ny + 1);
  if (!f.per_x) {
    f.nq = 50;
    const int nb = 2 * (f.nx + 1) + 2 * (f.ny - 1);
    hipLaunchKernelGGL(k_node_integral, dim3(nb), dim3(256), 0, c->s, c->nden, c->tab_sp, c->pot,
                       c->kp, f, 1);
  }
  if (c->fft_ok) {
    int st2 = fem_fft(c);
    if (st2) return st2;
  } else if (nn <= CG_ONE_WG_NODES) {
    hipLaunchKernelGGL(k_fem_cg, dim3(1), dim3(1024), 0, c->s, c->n_node, c->pot, c->cg_r,
                       c->cg_p, c->cg_q, f, c->fem_iters);
  } else {
    int st2 = fem_cg_grid(c, f);
    if (st2) return st2;
  }
  hipLaunchKernelGGL(k_fem_core, dim3((nn + 255) / 256), dim3(256), 0, c->s, c->n_node, c->pot,
                     f);
  LAUNCHCHK();
  c->kick_coef = c->p.ast;
  return MCPM_OK;
}

// compute_Vlasov_2D (FEM.cpp:613-726) up to the nodal field.
int vlasov_2d(mcpm_ctx *c) {
  FemParams f = c->fem;
  f.nq = 100;
  const int nn = (f.nx + 1) * (f.ny + 1);
  hipLaunchKernelGGL(k_node_integral, dim3(nn), dim3(256), 0, c->s, c->nden, c->tab_vl, c->pot,
                     c->kp, f, 0);
  LAUNCHCHK();
  c->kick_coef = c->p.phi;
  return MCPM_OK;
}

double contact_y_host(const mcpm_params *p, double n) { return contact_y(n, p->b, p->y_model); }

// ---- graph-replayed ESMC steps ---------------------------------------------
// What a captured step covers: Collision_ESMC_new3 on a slab (one or batched
// replicas), no mean field or the screened-Poisson one, the windowed sort,
// nothing that needs the host between launches (pressure face sums,
// re-insertion, particle ids or weights, the debug log).
bool graph_ok(const mcpm_ctx *c) {
  if (c->p.dims != 1 || c->p.model != MCPM_MODEL_ESMC) return false;
  if (c->p.long_range != MCPM_LR_NONE && c->p.long_range != MCPM_LR_SCREENED_POISSON) return false;
  if (c->pressure_on || c->track_ids || c->track_w || c->kp.cell_ratio || c->debug_log) return false;
  if (c->n <= 0 || !c->sorted || !c->order_sorted) return false;
  return window_ok(c, PUSH_MOVE);
}

void graph_release(mcpm_ctx *c) {
  for (auto &g : c->gexec)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
  c->g_n = -1;
}

// One ESMC step (mcpm_step's order: collisions, field, kick + stream +
// boundary + sort) as launches whose step-dependent values all come from
// c->sd: the candidate count from cstart[nc] (grids sized for c->g_cap
// candidates), the step and its priority key, the round tags; a fixed
// g_rounds grid rounds, then the tail kernel.  The particles are read from
// P[cur] and the sort writes P[1 - cur].
template <typename R>
int graph_step_launch(mcpm_ctx *c) {
  const int nc = c->nc;
  StepDev *sd = c->sd;
  const StepDev *csd = sd;
  R *P = (R *)c->P[c->cur];
  FPParams fpp = c->fpp;
  fpp.model = MCPM_MODEL_ESMC;
  const uint32_t cap = (uint32_t)c->g_cap;
  hipLaunchKernelGGL(k_mcand, dim3((nc + 255) / 256), dim3(256), 0, c->s, c->cnt, c->omax,
                     c->nden, c->yc, c->mcand, nc, c->kp, (int)COLL_ESMC, (int)MCPM_MODEL_ESMC,
                     0u, 0u, csd);
  if (int st = excl_scan(c, c->mcand, c->cstart, nc, csd)) return st;
  hipLaunchKernelGGL(k_zero_u32, dim3((CNT_TOTAL + 255) / 256), dim3(256), 0, c->s, c->counters,
                     (int)CNT_TOTAL);
  hipLaunchKernelGGL(k_step_mid, dim3(1), dim3(64), 0, c->s, sd, (const uint32_t *)c->cstart, nc,
                     cap, c->kp.key0, c->kp.key1);
  if (c->kp.nrep)
    hipLaunchKernelGGL(k_rep_step, dim3((c->kp.nrep + 256) / 256), dim3(256), 0, c->s, c->kp,
                       c->cstart, c->rprio, c->rcand_off, 0u, 0u, csd);
  const int nbands = std::max(1, std::min<int>(c->nbands, MAX_BANDS));
  const uint32_t band_mul = (uint32_t)nbands;
  const unsigned nch_cap = (unsigned)((cap + GEN_CHUNK - 1) / GEN_CHUNK);
  uint32_t *const tabs[2] = {c->rtab, c->rtab + (size_t)c->p.capacity};
  hipLaunchKernelGGL(k_cand_count, dim3(nch_cap), dim3(256), 0, c->s, 0u, 0u, (int)COLL_ESMC,
                     band_mul, nbands, c->gen_table, c->gen_tstride, c->kp, c->cstart, nc,
                     c->gen_cells, csd);
  hipLaunchKernelGGL(k_cand_scan, dim3(nbands * NREG), dim3(1024), 0, c->s, c->gen_table,
                     c->gen_tstride, 0u, nbands * NREG, c->counters,
                     (const uint32_t *)c->cstart + nc, csd);
  using real = typename R::real_type;
  hipLaunchKernelGGL((k_gen<R, COLL_ESMC>), dim3(nch_cap), dim3(256), 0, c->s, P,
                     (const real *)c->ys, c->cnt, c->start, c->cstart, nc, 0u, c->kp, 0u, 0u, 0u,
                     band_mul, nbands, c->gen_table, c->gen_tstride, (const uint32_t *)c->gen_cells,
                     c->gen_reach,
                     c->opBand, tabs[0], c->counters, csd);
  LAUNCHCHK();
  const uint32_t seg = (cap + NREG - 1) / NREG;
  const unsigned grid = (unsigned)((std::max(8, c->round_grid) + 7) / 8 * 8);
  const int unroll = round_unroll_for(c, COLL_ESMC, c->last_m, nbands, grid);
  const OpLog nolog{};
  for (int r = 0; r < c->g_rounds; ++r) {
    const uint4 *in = (const uint4 *)((r & 1) ? c->opB : c->opA);
    hipLaunchKernelGGL((k_check<COLL_ESMC>), dim3(grid), dim3(256), 0, c->s, r, nbands, seg, in,
                       (const uint4 *)c->opBand, tabs[r & 1], c->cflag, 0u,
                       c->counters, (uint32_t *)nullptr, csd);
    auto round_k = round_kernel<R, COLL_ESMC>(c->kp.nrep == 0, unroll);
    hipLaunchKernelGGL(round_k, dim3(grid), dim3(256), 0, c->s, r, nbands, seg, P,
                       in, (const uint4 *)c->opBand, (r & 1) ? c->opA : c->opB,
                       (const uint32_t *)tabs[r & 1], tabs[(r + 1) & 1],
                       (const uint8_t *)c->cflag, 0u, c->counters, c->nden, c->yc, c->omax,
                       c->omega_new, c->kp, 0u, 0u, fpp, nolog, csd);
  }
  hipLaunchKernelGGL((k_round_tail<R, COLL_ESMC>), dim3(1), dim3(1024), 0, c->s, c->g_rounds,
                     nbands, seg, P, c->opA, c->opB, tabs[0], tabs[1], c->cflag, c->counters,
                     (const double *)c->nden, (const double *)c->yc, (const double *)c->omax,
                     c->omega_new, c->kp, fpp, sd);
  hipLaunchKernelGGL(k_omega_finish, dim3((nc + 255) / 256), dim3(256), 0, c->s, c->omax,
                     c->omega_new, c->cnt, nc, (int)COLL_ESMC, (int)MCPM_MODEL_ESMC, csd);
  const bool field = c->p.long_range == MCPM_LR_SCREENED_POISSON;
  if (field)
    hipLaunchKernelGGL(k_field, dim3(c->kp.nrep ? c->kp.nrep : 1), dim3(1024), 0, c->s, c->nden,
                       c->fa, c->F2, c->kp, c->p.ast, c->p.bst, c->p.eqdist,
                       (const FieldArrays *)c->rfa, csd);
  LAUNCHCHK();
  const Kick kick{c->F2, c->pot, c->kick_coef, nullptr};
  if (int st = wsort_launch<R>(c, PUSH_MOVE | (field ? PUSH_KICK : 0), 0, kick, nullptr, csd))
    return st;
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, c->s, sd, (const uint32_t *)c->counters,
                     (const uint32_t *)c->cstart, nc, (const uint32_t *)c->wflag);
  LAUNCHCHK();
  return MCPM_OK;
}

// Capture the step for both parities of cur (the sort swaps the particle
// buffers every step, so replays alternate between the two graphs).
template <typename R>
int graph_capture_t(mcpm_ctx *c, int64_t cap, int rounds) {
  graph_release(c);
  cap = (cap + NREG - 1) / NREG * NREG;
  if (cap > (int64_t)UINT32_MAX - GEN_CHUNK) return MCPM_ERR_ARG;
  if (int st = ensure_ops(c, cap)) return st;
  if (!c->sd) {
    if (int st = dalloc(&c->sd, 1)) return st;
  }
  c->g_cap = cap;
  c->g_rounds = rounds;
  const int cur0 = c->cur;
  for (int par = 0; par < 2; ++par) {
    c->cur = par;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      c->cur = cur0;
      return MCPM_ERR_HIP;
    }
    const int st = graph_step_launch<R>(c);
    e = hipStreamEndCapture(c->s, &g);
    if (st == MCPM_OK && e == hipSuccess) e = hipGraphInstantiate(&c->gexec[par], g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    if (st || e != hipSuccess) {
      c->cur = cur0;
      graph_release(c);
      return st ? st : MCPM_ERR_HIP;
    }
  }
  c->cur = cur0;
  c->g_n = c->n;
  return MCPM_OK;
}

int graph_capture(mcpm_ctx *c, int64_t cap, int rounds) {
  return c->p.precision == MCPM_FP32 ? graph_capture_t<Rec<float>>(c, cap, rounds)
                                     : graph_capture_t<Rec<double>>(c, cap, rounds);
}

template <typename R>
int ys_fill_t(mcpm_ctx *c) {
  using real = typename R::real_type;
  hipLaunchKernelGGL(k_ys_from_records<R>, dim3(1024), dim3(256), 0, c->s,
                     (const R *)c->P[c->cur], c->n, (real *)c->ys);
  LAUNCHCHK();
  c->ys_valid = true;
  return MCPM_OK;
}

}  // namespace

const mcpm_params *ctx_params(const mcpm_ctx *c) { return &c->p; }
int64_t ctx_sample_count(const mcpm_ctx *c) { return c->n_samples; }

int ctx_plane_nodes(mcpm_ctx *c, PlaneField *pf) {
  if (!c || !pf) return MCPM_ERR_ARG;
  if (c->p.dims != 2 || c->p.long_range != MCPM_LR_SCREENED_POISSON) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  int st = MCPM_DISPATCH(c, launch_nodes, c);
  if (st) return st;
  pf->n_node = c->n_node;
  pf->pot = c->pot;
  pf->fem = c->fem;
  pf->s = c->s;
  return MCPM_OK;
}

int ctx_plane_cg(mcpm_ctx *c, int warm, int mode, double accept, double *res0, int *iters,
                 int *solved) {
  if (!c || c->p.dims != 2 || !c->fem.per_x || !c->fem.per_y) return MCPM_ERR_ARG;
  if (iters) *iters = 0;
  if (res0) *res0 = -1.0;
  if (solved) *solved = 0;
  double r = -1.0;
  if (warm) {
    // the starting residual; mode 2, or mode 1 within `accept`, keeps the guess
    int st = fem_cg_grid(c, c->fem, 1, 2, 0.0, &r, nullptr);
    if (st) return st;
    if (res0) *res0 = r;
    if (mode == 2 || (mode == 1 && r <= accept)) return MCPM_OK;
  }
  if (solved) *solved = 1;
  if (c->fft_ok) return fem_fft(c);  // the exact solve of a periodic power-of-two plane
  return fem_cg_grid(c, c->fem, warm, 0, 0.0, nullptr, iters);
}

int ctx_plane_finish(mcpm_ctx *c) {
  const int nn = (c->fem.nx + 1) * (c->fem.ny + 1);
  hipLaunchKernelGGL(k_fem_core, dim3((nn + 255) / 256), dim3(256), 0, c->s, c->n_node, c->pot,
                     c->fem);
  LAUNCHCHK();
  c->kick_coef = c->p.ast;
  c->kick_pending = true;
  return MCPM_OK;
}

int mcpm_set_field_surrogate(mcpm_ctx *c, mcpm_fno *f, double accept_residual, int64_t every) {
  if (!c) return MCPM_ERR_ARG;
  if (f && (c->p.dims != 2 || c->p.long_range != MCPM_LR_SCREENED_POISSON || !c->fem.per_x ||
            !c->fem.per_y || every < 0))
    return MCPM_ERR_ARG;
  c->fno = f;
  c->fno_accept = accept_residual;
  c->fno_every = every;
  c->fno_last = mcpm_fno_field_stats{};
  return MCPM_OK;
}

int ctx_set_surrogate_stats(mcpm_ctx *c, const mcpm_fno_field_stats *st) {
  c->fno_last = *st;
  return MCPM_OK;
}

int mcpm_get_field_surrogate_stats(mcpm_ctx *c, mcpm_fno_field_stats *out) {
  if (!c || !out) return MCPM_ERR_ARG;
  *out = c->fno_last;
  return MCPM_OK;
}

template <typename R>
int launch_init(mcpm_ctx *c, const InitParams &ip) {
  if (ip.n > 0)
    hipLaunchKernelGGL((k_init<R>), dim3((unsigned)((ip.n + 255) / 256)), dim3(256), 0, c->s,
                       (R *)c->P[c->cur], ip, c->kp);
  if (c->track_ids && ip.n > 0)
    hipLaunchKernelGGL(k_iota_u32, dim3(1024), dim3(256), 0, c->s, c->ids[c->cur], ip.n);
  c->order_sorted = false;
  c->ys_valid = false;
  if (c->kp.cell_ratio) {
    c->track_w = false;
    c->w_pending = true;
  }
  LAUNCHCHK();
  return MCPM_OK;
}

// inverted_addremove_particles on the sorted slab (mcpm_addremove.hip).
template <typename R>
int addremove_t(mcpm_ctx *c, int32_t nr, int64_t step, mcpm_addremove_stats *stats) {
  if constexpr (R::dims != 1) {
    return MCPM_ERR_ARG;
  } else {
    using real = typename R::real_type;
    const int nc = c->nc;
    const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
    const R *P = (const R *)c->P[c->cur];
    R *out = (R *)c->P[1 - c->cur];
    uint32_t *tot = c->ar_tot, *ostart = c->ar_ostart;
    if (nc > 0)
      hipLaunchKernelGGL((k_ar_classify<R>), dim3((unsigned)nc), dim3(256), 0, c->s, P, c->cnt,
                         c->start, (const int32_t *)c->pw[c->cur], (const int32_t *)c->cratio,
                         nr, c->kp, slo, shi, c->ar_fl, c->ar_cst, c->ar_h, tot);
    if (int st = excl_scan(c, tot, ostart, nc)) return st;
    LAUNCHCHK();
    uint32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, ostart + nc, sizeof total, hipMemcpyDeviceToHost, c->s));
    std::vector<uint32_t> cst;
    if (stats) {
      cst.resize(4 * (size_t)nc);
      HIPCHK(hipMemcpyAsync(cst.data(), c->ar_cst, cst.size() * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, c->s));
    }
    HIPCHK(hipStreamSynchronize(c->s));
    if ((int64_t)total > c->p.capacity) return MCPM_ERR_CAPACITY;  // nothing changed yet
    if (nc > 0)
      hipLaunchKernelGGL((k_ar_scatter<R>), dim3((unsigned)nc), dim3(256), 0, c->s, P, out,
                         c->cnt, c->start, (const uint32_t *)ostart, (const uint8_t *)c->ar_fl,
                         (const uint32_t *)c->ar_cst, (const double *)c->ar_h,
                         (const int32_t *)c->cratio, nr, c->kp, slo, shi, c->pw[1 - c->cur],
                         (real *)c->ys);
    HIPCHK(hipMemcpyAsync(c->start, ostart, ((size_t)nc + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, c->s));
    hipLaunchKernelGGL(k_cell_fill, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->s,
                       c->start, nc, c->cnt, c->nden, c->yc, c->kp);
    LAUNCHCHK();
    if (stats) {
      int64_t split = 0, del = 0;
      for (int q = 0; q < nc; ++q) {
        split += cst[4 * (size_t)q];
        del += cst[4 * (size_t)q + 2];
      }
      stats->before = c->n;
      stats->split = split;
      stats->created = split * (int64_t)(nr - 1);
      stats->deleted = del;
      stats->after = (int64_t)total;
    }
    c->n = (int64_t)total;
    c->cur = 1 - c->cur;
    c->ys_valid = c->ys != nullptr;
    c->keys_current = false;
    c->have_closures = false;
    return MCPM_OK;
  }
}

extern "C" {

const char *mcpm_status_string(int s) {
  switch (s) {
    case MCPM_OK: return "ok";
    case MCPM_ERR_ARG: return "invalid argument";
    case MCPM_ERR_HIP: return "HIP runtime error";
    case MCPM_ERR_CAPACITY: return "capacity exceeded";
    case MCPM_ERR_DOMAIN: return "particle outside the domain";
    case MCPM_ERR_NAN: return "non-finite velocity";
    case MCPM_ERR_BC: return "too many boundary interactions in one step";
    case MCPM_ERR_STATE: return "call out of order";
    default: return "unknown status";
  }
}

static int create_ctx(const mcpm_params *p_in, int32_t nrep, const mcpm_replica *reps,
                      mcpm_ctx **out) {
  if (!p_in || !out) return MCPM_ERR_ARG;
  *out = nullptr;
  mcpm_params pv = *p_in;
  if (pv.dims == 0) pv.dims = 1;
  const mcpm_params *p = &pv;
  if (p->dims != 1 && p->dims != 2) return MCPM_ERR_ARG;
  const int64_t nx = p->dims == 2 ? p->n_cells_x : 1;
  if (p->n_cells < 1 || nx < 1 || nx * p->n_cells > MAX_CELLS) return MCPM_ERR_ARG;
  if (p->dims == 2) {
    // the plane: specular walls or periodic axes, hard-sphere or DSMC
    // collisions, direct or screened-Poisson attraction (FEM.cpp)
    if (!(p->length_x > 0.0)) return MCPM_ERR_ARG;
    if (p->boundary == MCPM_BC_DIFFUSE || p->boundary_x == MCPM_BC_DIFFUSE) return MCPM_ERR_ARG;
    if (p->boundary < 0 || p->boundary > 2 || p->boundary_x < 0 || p->boundary_x > 2)
      return MCPM_ERR_ARG;
    if (p->model == MCPM_MODEL_FP_DENSE || p->model == MCPM_MODEL_FP_IDEAL ||
        p->model == MCPM_MODEL_FP_LINEAR)
      return MCPM_ERR_ARG;
    if (p->long_range == MCPM_LR_DENSITY_EXPANSION) return MCPM_ERR_ARG;
    if (p->long_range != MCPM_LR_NONE && (nx < 2 || p->n_cells < 2)) return MCPM_ERR_ARG;
    if (p->long_range == MCPM_LR_SCREENED_POISSON && p->boundary != p->boundary_x)
      return MCPM_ERR_ARG;
  }
  if (p->capacity < 1 || p->capacity > 2000000000LL) return MCPM_ERR_ARG;
  if (p->precision != MCPM_FP32 && p->precision != MCPM_FP64) return MCPM_ERR_ARG;
  if (p->boundary < 0 || p->boundary > MCPM_BC_REINJECT) return MCPM_ERR_ARG;
  if (p->boundary == MCPM_BC_REINJECT && (p->dims != 1 || !(p->T_wall_lo > 0.0) ||
                                          !(p->wall_hi > p->wall_lo)))
    return MCPM_ERR_ARG;
  if (!(p->length > 0.0) || !(p->delta_t > 0.0) || !(p->cell_volume > 0.0)) return MCPM_ERR_ARG;
  if (p->long_range == MCPM_LR_SCREENED_POISSON && !(p->bst > 0.0)) return MCPM_ERR_ARG;
  if (p->long_range < MCPM_LR_NONE || p->long_range > MCPM_LR_SCREENED_POISSON)
    return MCPM_ERR_ARG;
  if (p->model < MCPM_MODEL_ESMC || p->model > MCPM_MODEL_ESMC_NEW4) return MCPM_ERR_ARG;
  if (p->model == MCPM_MODEL_ESMC_NEW4 && p->dims == 2) return MCPM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MCPM_ERR_HIP;
  mcpm_ctx *c = new mcpm_ctx();
  c->p = *p;
  c->nc = (int)(nx * p->n_cells);
  HIPCHK(hipGetDevice(&c->device));
  HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  KParams &k = c->kp;
  k.n_cells = c->nc;
  k.boundary = p->boundary;
  k.y_model = p->y_model;
  k.dims = p->dims;
  k.nx = (int32_t)nx;
  k.ny = p->n_cells;
  k.boundary_x = p->dims == 2 ? p->boundary_x : MCPM_BC_PERIODIC;
  k.reserved = 0;
  k.length_x = p->dims == 2 ? p->length_x : 1.0;
  k.dx = k.length_x / (double)nx;
  k.dy = p->length / (double)p->n_cells;
  k.length = p->length;
  k.wall_lo = p->wall_lo;
  k.wall_hi = p->wall_hi;
  k.T_lo = p->T_wall_lo;
  k.T_hi = p->T_wall_hi;
  k.U_lo = p->U_wall_lo;
  k.U_hi = p->U_wall_hi;
  k.sigma = p->sigma;
  k.mass = p->mass;
  k.kb = p->kb;
  k.b = p->b;
  k.Fn = p->Fn;
  k.volume = p->cell_volume;
  k.n_ratio = p->n_ratio;
  k.dt = p->delta_t;
  k.coll_pref = 4.0 * PI * p->sigma * p->sigma;
  k.key0 = (uint32_t)p->seed;
  k.key1 = (uint32_t)(p->seed >> 32);
  k.nrep = 0;
  k.cell_ratio = nullptr;
  if (nrep > 0) {
    // per-replica constants: own cells, dy, volume, Fn and key; the rest shared
    c->rkp_host.resize(nrep);
    c->rcell_host.assign(nrep + 1, 0);
    c->rpart_host.assign(nrep + 1, 0);
    std::vector<int32_t> crep(c->nc);
    for (int r = 0; r < nrep; ++r) {
      KParams q = k;
      q.n_cells = reps[r].n_cells;
      q.ny = reps[r].n_cells;
      q.dy = p->length / (double)reps[r].n_cells;
      q.volume = reps[r].cell_volume;
      q.Fn = reps[r].Fn;
      q.key0 = (uint32_t)reps[r].seed;
      q.key1 = (uint32_t)(reps[r].seed >> 32);
      c->rkp_host[r] = q;
      c->rcell_host[r + 1] = c->rcell_host[r] + reps[r].n_cells;
      for (int cc = c->rcell_host[r]; cc < c->rcell_host[r + 1]; ++cc) crep[cc] = r;
    }
    int st0;
    if ((st0 = dalloc(&c->rkp, (size_t)nrep)) || (st0 = dalloc(&c->cell_rep, (size_t)c->nc)) ||
        (st0 = dalloc(&c->rcell_off, (size_t)nrep + 1)) ||
        (st0 = dalloc(&c->rpart_off, (size_t)nrep + 1)) || (st0 = dalloc(&c->rprio, (size_t)nrep)) ||
        (st0 = dalloc(&c->rcand_off, (size_t)nrep + 1)) ||
        (st0 = dalloc(&c->rpart_bucket, (size_t)(p->capacity >> REP_BUCKET_SHIFT) + 2))) {
      mcpm_destroy(c);
      return st0;
    }
    HIPCHK(hipMemcpy(c->rkp, c->rkp_host.data(), nrep * sizeof(KParams), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->cell_rep, crep.data(), c->nc * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->rcell_off, c->rcell_host.data(), (nrep + 1) * sizeof(int32_t),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemset(c->rpart_off, 0, (nrep + 1) * sizeof(int64_t)));
    k.nrep = nrep;
    k.rkp = c->rkp;
    k.cell_rep = c->cell_rep;
    k.rcell_off = c->rcell_off;
    k.rpart_off = c->rpart_off;
    k.rprio = c->rprio;
    k.rcand_off = c->rcand_off;
    k.rpart_bucket = nullptr;  // set by the first upload
  }
  // a partner point lies within sigma of its candidate's particle: on the
  // slab at most ceil(sigma / dy) + 1 cells from the candidate's cell (the
  // draw keeps those cells' counts at hand; the plane looks them up)
  if (p->dims != 2) {
    double r = std::ceil(p->sigma / k.dy) + 1.0;
    for (const KParams &q : c->rkp_host) r = std::max(r, std::ceil(p->sigma / q.dy) + 1.0);
    c->gen_reach = r < (double)GEN_WIN ? (int32_t)r : -1;
  }
  c->fpp.model = p->model;
  c->fpp.reserved = 0;
  c->fpp.mu = p->mu;
  c->fpp.T0 = p->T0;
  c->fpp.visp = p->visp;
  c->fpp.crref = p->crref;
  c->rec = (p->dims == 2 ? 8 : 4) * (p->precision == MCPM_FP32 ? sizeof(float) : sizeof(double));
  const int nc = c->nc;
  int st = MCPM_OK;
  auto fail = [&](int s2) {
    mcpm_destroy(c);
    return s2;
  };
  if ((st = dalloc((char **)&c->P[0], c->rec * (size_t)p->capacity))) return fail(st);
  if ((st = dalloc((char **)&c->P[1], c->rec * (size_t)p->capacity))) return fail(st);
  if (p->dims != 2 &&
      (st = dalloc((char **)&c->ys, (c->rec / 4) * (size_t)p->capacity)))
    return fail(st);
  // chunks of at least 1024 particles; the per-chunk histogram table is kept
  // under 512 MB (chunk_for's automatic size makes about 2048 chunks)
  // more cells than one LDS histogram holds: sort by two balanced digits
  int nbins = nc;
  if (nc > SORT_BINS) {
    c->two_digit = true;
    int32_t d = (int32_t)std::ceil(std::sqrt((double)nc));
    c->kdiv = d;
    c->nb_lo = d;
    c->nb_hi = (nc + d - 1) / d;
    nbins = std::max(c->nb_lo, c->nb_hi);
    if ((st = dalloc(&c->dcnt, nbins))) return fail(st);
    if ((st = dalloc(&c->dstart, nbins + 1))) return fail(st);
    if ((st = dalloc(&c->keys, (size_t)p->capacity))) return fail(st);
    if ((st = dalloc(&c->keys_tmp, (size_t)p->capacity))) return fail(st);
  }
  c->nchunks_cap = (p->capacity + 1023) / 1024;
  {
    const int64_t tab_cap = ((int64_t)512 << 20) / (4 * (int64_t)nbins);
    if (c->nchunks_cap > tab_cap) c->nchunks_cap = tab_cap;
    if (c->nchunks_cap < 2100) c->nchunks_cap = 2100;
  }
  if ((st = dalloc(&c->table, (size_t)c->nchunks_cap * nbins))) return fail(st);
  // windowed sort (slabs): blocks of WSPAN cells of one replica, and at most
  // one partial chunk per block beyond capacity / WCH
  if (p->dims != 2) {
    std::vector<int32_t> bcell, brep, rblk;
    for (int r = 0; r < std::max(nrep, 1); ++r) {
      const int32_t a = nrep ? c->rcell_host[r] : 0, z = nrep ? c->rcell_host[r + 1] : nc;
      rblk.push_back((int32_t)brep.size());
      for (int32_t k = a; k < z; k += WSPAN) {
        bcell.push_back(k);
        brep.push_back(r);
      }
    }
    rblk.push_back((int32_t)brep.size());
    bcell.push_back(nc);
    c->nwblk = (int)brep.size();
    c->wcap = (p->capacity + WCH - 1) / WCH + c->nwblk + 1;
    if ((st = dalloc(&c->wbeg, (size_t)c->wcap + 1))) return fail(st);
    if ((st = dalloc(&c->wrep, (size_t)c->wcap))) return fail(st);
    if ((st = dalloc(&c->wlo, (size_t)c->wcap))) return fail(st);
    if ((st = dalloc(&c->wlen, (size_t)c->wcap))) return fail(st);
    if ((st = dalloc(&c->wtab, (size_t)c->wcap * WMAX))) return fail(st);
    if ((st = dalloc(&c->rchunk, rblk.size()))) return fail(st);
    if ((st = dalloc(&c->wpieces, (size_t)c->nwblk + 1))) return fail(st);
    if ((st = dalloc(&c->wpbase, (size_t)c->nwblk + 1))) return fail(st);
    if ((st = dalloc(&c->wbcell, bcell.size()))) return fail(st);
    if ((st = dalloc(&c->wbrep, brep.size() + 1))) return fail(st);
    if ((st = dalloc(&c->wrblk, rblk.size()))) return fail(st);
    if ((st = dalloc(&c->wflag, 2))) return fail(st);
    HIPCHK(hipMemcpy(c->wbcell, bcell.data(), bcell.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wbrep, brep.data(), brep.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->wrblk, rblk.data(), rblk.size() * 4, hipMemcpyHostToDevice));
  }
  if ((st = dalloc(&c->partial, (size_t)((c->nchunks_cap + c->seg - 1) / c->seg + 1) * nbins)))
    return fail(st);
  if ((st = dalloc(&c->cnt, nc))) return fail(st);
  if ((st = dalloc(&c->start, nc + 1))) return fail(st);
  if ((st = dalloc(&c->mcand, nc))) return fail(st);
  if ((st = dalloc(&c->cstart, nc + 1))) return fail(st);
  if ((st = dalloc(&c->nden, nc))) return fail(st);
  if ((st = dalloc(&c->yc, nc))) return fail(st);
  if ((st = dalloc(&c->omax, nc))) return fail(st);
  if ((st = dalloc(&c->omega_new, nc))) return fail(st);
  if ((st = dalloc(&c->F2, nc))) return fail(st);
  if ((st = dalloc(&c->moments, (size_t)nc * MCPM_NMOM))) return fail(st);
  if ((st = dalloc(&c->samples, (size_t)nc * MCPM_NSAMP))) return fail(st);
  if ((st = dalloc(&c->press, (size_t)nc * 3))) return fail(st);
  if ((st = dalloc(&c->counters, CNT_TOTAL))) return fail(st);
  if ((st = dalloc(&c->clos, (size_t)nc * MCPM_NCLOS))) return fail(st);
  if ((st = dalloc(&c->drift, (size_t)nc * DRIFT_N))) return fail(st);
  if ((st = dalloc(&c->err, 1))) return fail(st);
  if ((st = dalloc(&c->esc, 1))) return fail(st);
  if ((st = dalloc(&c->energy_part, 4 * 1024))) return fail(st);
  if (hipHostMalloc((void **)&c->hleft, (size_t)(MAX_ROUNDS + 2) * sizeof(uint32_t),
                    hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void **)&c->dleft, c->hleft, 0) != hipSuccess)
    return fail(MCPM_ERR_HIP);
  if ((st = dalloc(&c->rtab, 2 * (size_t)p->capacity + ((size_t)p->capacity + 3) / 4)))
    return fail(st);
  c->cflag = (uint8_t *)(c->rtab + 2 * (size_t)p->capacity);
  c->scan_tiles = (int)(((int64_t)c->nc + SCAN_TILE - 1) / SCAN_TILE) + 1;
  if ((st = dalloc(&c->scan_tmp, 2 * ((size_t)c->scan_tiles + 1)))) return fail(st);
  if (hipMemset(c->omax, 0, nc * sizeof(double)) != hipSuccess ||
      hipMemset(c->omega_new, 0, nc * sizeof(double)) != hipSuccess ||
      hipMemset(c->F2, 0, nc * sizeof(double)) != hipSuccess ||
      hipMemset(c->samples, 0, (size_t)nc * MCPM_NSAMP * sizeof(double)) != hipSuccess ||
      hipMemset(c->press, 0, (size_t)nc * 3 * sizeof(double)) != hipSuccess ||
      hipMemset(c->err, 0, sizeof(uint32_t)) != hipSuccess ||
      hipMemset(c->esc, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(c->cnt, 0, nc * sizeof(uint32_t)) != hipSuccess ||
      hipMemset(c->start, 0, (nc + 1) * sizeof(uint32_t)) != hipSuccess ||
      hipMemset(c->cflag, 0, (size_t)p->capacity) != hipSuccess)
    return fail(MCPM_ERR_HIP);
  if (p->dims == 2) {
    if ((st = build_fem(c))) return fail(st);
  } else {
    if (p->long_range == MCPM_LR_SCREENED_POISSON)
      if ((st = build_field(c))) return fail(st);
    if ((st = build_vlasov_table(c))) return fail(st);
  }
  if ((st = dalloc(&c->kde, (size_t)nc * 3))) return fail(st);
  if (hipDeviceSynchronize() != hipSuccess) return fail(MCPM_ERR_HIP);
  *out = c;
  return MCPM_OK;
}

int mcpm_create(const mcpm_params *p, mcpm_ctx **out) { return create_ctx(p, 0, nullptr, out); }

int mcpm_create_batch(const mcpm_params *common, int32_t n_rep, const mcpm_replica *reps,
                      mcpm_ctx **out) {
  if (!common || !reps || !out || n_rep < 1 || n_rep > 65536) return MCPM_ERR_ARG;
  *out = nullptr;
  // independent slabs of one length: ESMC or free flight, no field or the
  // screened-Poisson field
  if (common->dims > 1 || common->boundary == MCPM_BC_REINJECT) return MCPM_ERR_ARG;
  if (common->model != MCPM_MODEL_ESMC && common->model != MCPM_MODEL_NONE) return MCPM_ERR_ARG;
  if (common->long_range != MCPM_LR_NONE && common->long_range != MCPM_LR_SCREENED_POISSON)
    return MCPM_ERR_ARG;
  mcpm_params p = *common;
  int64_t cells = 0, cap = 0;
  for (int r = 0; r < n_rep; ++r) {
    if (reps[r].n_cells < 1 || reps[r].capacity < 0 || !(reps[r].cell_volume > 0.0) ||
        !(reps[r].Fn > 0.0))
      return MCPM_ERR_ARG;
    cells += reps[r].n_cells;
    cap += reps[r].capacity;
  }
  if (cells > MAX_CELLS || cap < 1) return MCPM_ERR_ARG;
  p.dims = 1;
  p.n_cells = (int32_t)cells;
  p.capacity = cap;
  return create_ctx(&p, n_rep, reps, out);
}

int mcpm_init_omega_max_batch(mcpm_ctx *c, const double *n_gas, int32_t times_y) {
  if (!c || !n_gas) return MCPM_ERR_ARG;
  if (c->kp.nrep == 0) return MCPM_ERR_STATE;
  std::vector<double> mom((size_t)c->nc * MCPM_NMOM), nd(c->nc), om(c->nc);
  int st = mcpm_moments(c, mom.data());
  if (st) return st;
  st = mcpm_get_cells(c, nullptr, nullptr, nd.data(), nullptr, nullptr);
  if (st) return st;
  for (int r = 0; r < c->kp.nrep; ++r) {
    const double y = times_y ? contact_y_host(&c->p, n_gas[r]) : 1.0;
    for (int k = c->rcell_host[r]; k < c->rcell_host[r + 1]; ++k) {
      double T = mom[(size_t)k * MCPM_NMOM + 29];
      double crm = 10.0 * sqrt(c->p.kb * T / c->p.mass);                          // main.cpp:259
      double o = 4.0 * PI * c->p.sigma * c->p.sigma * crm * nd[k] * c->p.delta_t;  // :260
      om[k] = o * y;                                                                  // :268-269
    }
  }
  return mcpm_set_omega_max(c, om.data());
}

int mcpm_destroy(mcpm_ctx *c) {
  if (!c) return MCPM_OK;
  if (c->s) hipStreamSynchronize(c->s);
  void *bufs[] = {c->rkp, c->cell_rep, c->rcell_off, c->rpart_off, c->rprio, c->rcand_off,
                  c->rpart_bucket,
                  c->rfa,
                  c->P[0], c->P[1], c->ys, c->scan_tmp, c->cratio, c->pw[0], c->ar_fl, c->ar_cst,
                  c->ar_h, c->table, c->partial, c->dcnt, c->dstart, c->keys,
                  c->keys_tmp,
                  c->cnt, c->start, c->mcand,
                  c->cstart, c->nden, c->yc, c->omax, c->omega_new, c->F2, c->moments,
                  c->samples, c->press, c->counters, c->err, c->esc, c->energy_part, c->rtab, c->opA,
                  c->log.p, c->field_buf, c->clos, c->drift, c->lr_tab, c->kde, c->gen_table,
                  c->fem_buf, c->q4, c->fft_buf, c->fft_twx, c->fft_twy,
                  c->fft_cosx, c->fft_cosy, c->ids[0], c->wbeg, c->wrep, c->wlo,
                  c->wlen, c->rchunk, c->wtab, c->wflag, c->wpieces, c->wpbase, c->wbcell,
                  c->wbrep, c->wrblk};
  for (void *b : bufs)
    if (b) hipFree(b);
  graph_release(c);
  if (c->sd) hipFree(c->sd);
  for (int k = 0; k < T_NCLASS; ++k)
    for (auto &pr : c->ev[k]) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  for (auto e : c->ev_pool) hipEventDestroy(e);
  if (c->hleft) hipHostFree(c->hleft);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return MCPM_OK;
}

int mcpm_sync(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  return check_err(c);
}

}  // extern "C"

namespace {
// Pack host SoA arrays into the context's records: (y, u1, u2, u3) for the
// slab, (x, y, u1, u2, u3, 0, 0, 0) for the plane.
template <typename real>
void pack(std::vector<char> &host, int64_t n, int stride, const double *const *src, int nsrc) {
  real *h = (real *)host.data();
  for (int64_t k = 0; k < n; ++k) {
    for (int q = 0; q < nsrc; ++q) h[stride * k + q] = (real)src[q][k];
    for (int q = nsrc; q < stride; ++q) h[stride * k + q] = (real)0;
  }
}
template <typename real>
void unpack(const std::vector<char> &host, int64_t n, int stride, double *const *dst, int ndst) {
  const real *h = (const real *)host.data();
  for (int64_t k = 0; k < n; ++k)
    for (int q = 0; q < ndst; ++q)
      if (dst[q]) dst[q][k] = (double)h[stride * k + q];
}

int upload_any(mcpm_ctx *c, int64_t n, const double *const *src, int nsrc) {
  if (n < 0) return MCPM_ERR_ARG;
  if (n > c->p.capacity) return MCPM_ERR_CAPACITY;
  for (int q = 0; q < nsrc; ++q)
    if (n > 0 && !src[q]) return MCPM_ERR_ARG;
  std::vector<char> host(c->rec * (size_t)n);
  const int stride = c->p.dims == 2 ? 8 : 4;
  if (c->p.precision == MCPM_FP32) pack<float>(host, n, stride, src, nsrc);
  else pack<double>(host, n, stride, src, nsrc);
  HIPCHK(hipStreamSynchronize(c->s));
  if (n > 0)
    HIPCHK(hipMemcpy(c->P[c->cur], host.data(), host.size(), hipMemcpyHostToDevice));
  c->n = n;
  c->sorted = false;
  c->kick_pending = false;
  c->order_sorted = false;
  c->ys_valid = false;
  if (c->kp.cell_ratio) {
    c->track_w = false;
    c->w_pending = true;
  }
  if (c->track_ids && n > 0) {
    hipLaunchKernelGGL(k_iota_u32, dim3(1024), dim3(256), 0, c->s, c->ids[c->cur], n);
    LAUNCHCHK();
  }
  return MCPM_OK;
}

int download_any(mcpm_ctx *c, double *const *dst, int ndst) {
  const int64_t n = c->n;
  std::vector<char> host(c->rec * (size_t)(n > 0 ? n : 1));
  HIPCHK(hipStreamSynchronize(c->s));
  if (n > 0)
    HIPCHK(hipMemcpy(host.data(), c->P[c->cur], c->rec * (size_t)n, hipMemcpyDeviceToHost));
  const int stride = c->p.dims == 2 ? 8 : 4;
  if (c->p.precision == MCPM_FP32) unpack<float>(host, n, stride, dst, ndst);
  else unpack<double>(host, n, stride, dst, ndst);
  return MCPM_OK;
}
}  // namespace

namespace {

template <typename R>
int launch_thermostat(mcpm_ctx *c, double zone_lo, double T_lo, double zone_hi, double T_hi) {
  if constexpr (sizeof(R) == 4 * sizeof(typename R::real_type)) {
    int st = launch_moments<R>(c);
    if (st) return st;
    if (c->n > 0)
      hipLaunchKernelGGL((k_thermostat<R>), dim3((unsigned)((c->n + 255) / 256)), dim3(256), 0,
                         c->s, (R *)c->P[c->cur], c->n, c->moments, c->kp, zone_lo, T_lo, zone_hi,
                         T_hi);
    LAUNCHCHK();
    return MCPM_OK;
  } else {
    return MCPM_ERR_ARG;
  }
}


template <typename R>
int launch_energy(mcpm_ctx *c, double *out) {
  std::vector<double> part(4 * (size_t)ENERGY_GRID, 0.0);
  if (c->n > 0) {
    hipLaunchKernelGGL((k_energy<R>), dim3(ENERGY_GRID), dim3(256), 0, c->s,
                       (const R *)c->P[c->cur], c->n, c->energy_part);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(part.data(), c->energy_part, part.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
  }
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < ENERGY_GRID; ++b)
    for (int k = 0; k < 4; ++k) s[k] += part[4 * (size_t)b + k];
  const double mf = c->p.mass * c->p.Fn;
  out[0] = 0.5 * mf * s[0];
  for (int k = 1; k < 4; ++k) out[k] = mf * s[k];
  return MCPM_OK;
}

}  // namespace

extern "C" {

int mcpm_upload(mcpm_ctx *c, int64_t n, const double *y, const double *u1, const double *u2,
                const double *u3) {
  if (!c || c->p.dims == 2) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // a batch needs its per-replica counts
  const double *src[4] = {y, u1, u2, u3};
  return upload_any(c, n, src, 4);
}

int mcpm_upload_batch(mcpm_ctx *c, const int64_t *counts, const double *y, const double *u1,
                      const double *u2, const double *u3) {
  if (!c || !counts) return MCPM_ERR_ARG;
  const int nrep = c->kp.nrep;
  if (nrep == 0) return MCPM_ERR_STATE;
  int64_t n = 0;
  for (int r = 0; r < nrep; ++r) {
    if (counts[r] < 0) return MCPM_ERR_ARG;
    c->rpart_host[r] = n;
    n += counts[r];
  }
  c->rpart_host[nrep] = n;
  if (n > c->p.capacity) return MCPM_ERR_CAPACITY;
  HIPCHK(hipMemcpy(c->rpart_off, c->rpart_host.data(), (nrep + 1) * sizeof(int64_t),
                   hipMemcpyHostToDevice));
  // the replica of each bucket's first particle (the last r with
  // rpart_off[r] <= b << REP_BUCKET_SHIFT)
  std::vector<int32_t> bucket((size_t)(c->p.capacity >> REP_BUCKET_SHIFT) + 2, 0);
  int32_t r = 0;
  for (size_t b = 0; b < bucket.size(); ++b) {
    const int64_t p0 = (int64_t)b << REP_BUCKET_SHIFT;
    while (r + 1 < nrep && c->rpart_host[r + 1] <= p0) ++r;
    bucket[b] = r;
  }
  HIPCHK(hipMemcpy(c->rpart_bucket, bucket.data(), bucket.size() * sizeof(int32_t),
                   hipMemcpyHostToDevice));
  c->kp.rpart_bucket = c->rpart_bucket;
  const double *src[4] = {y, u1, u2, u3};
  return upload_any(c, n, src, 4);
}


int mcpm_upload_xy(mcpm_ctx *c, int64_t n, const double *x, const double *y, const double *u1,
                   const double *u2, const double *u3) {
  if (!c || c->p.dims != 2) return MCPM_ERR_ARG;
  const double *src[5] = {x, y, u1, u2, u3};
  return upload_any(c, n, src, 5);
}

int mcpm_download(mcpm_ctx *c, double *y, double *u1, double *u2, double *u3) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims == 2) {
    double *dst[5] = {nullptr, y, u1, u2, u3};
    return download_any(c, dst, 5);
  }
  double *dst[4] = {y, u1, u2, u3};
  return download_any(c, dst, 4);
}

int mcpm_download_xy(mcpm_ctx *c, double *x, double *y, double *u1, double *u2, double *u3) {
  if (!c || c->p.dims != 2) return MCPM_ERR_ARG;
  double *dst[5] = {x, y, u1, u2, u3};
  return download_any(c, dst, 5);
}

int64_t mcpm_particle_count(const mcpm_ctx *c) { return c ? c->n : -1; }

int mcpm_cell_update(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  int st = run_sort(c, PUSH_CHECK, 0);
  if (st) return st;
  return check_err(c);
}

int mcpm_collide_esmc(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  return MCPM_DISPATCH(c, collide_esmc_t, c, step, stats);
}

int mcpm_collide_esmc_new4(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims == 2 || c->kp.nrep) return MCPM_ERR_ARG;  // the reference's 1D operator
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  return MCPM_DISPATCH(c, collide_esmc4_t, c, step, stats);
}

int mcpm_collide_dsmc_vhs(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  return MCPM_DISPATCH(c, collide_dsmc_t, c, step, stats, (int)MCPM_MODEL_DSMC_VHS);
}

int mcpm_collide_dsmc_hs(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  return MCPM_DISPATCH(c, collide_dsmc_t, c, step, stats, (int)MCPM_MODEL_DSMC_HS);
}

int mcpm_collide_cba_vhs(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  c->keys_current = false;  // the collisions displace particles
  return MCPM_DISPATCH(c, collide_dsmc_t, c, step, stats, (int)MCPM_MODEL_CBA_VHS);
}

int mcpm_collide_cba_hs(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->n == 0) {
    if (stats) memset(stats, 0, sizeof *stats);
    return MCPM_OK;
  }
  c->keys_current = false;  // the collisions displace particles
  return MCPM_DISPATCH(c, collide_cba_hs_t, c, step, stats);
}

int mcpm_fp_closures(mcpm_ctx *c, int64_t step) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->p.dims != 1) return MCPM_ERR_ARG;
  Timer tm(c, T_FPCLOS);
  int st = MCPM_DISPATCH(c, launch_moments, c);
  if (st) return st;
  hipLaunchKernelGGL(k_fp_closures, dim3(c->nc), dim3(64), 0, c->s, c->moments,
                     c->cnt, c->nden, c->yc, c->clos, c->drift, c->kp, c->fpp, step, c->err);
  LAUNCHCHK();
  c->have_closures = true;
  return MCPM_OK;
}

int mcpm_velocity_fp(mcpm_ctx *c, int64_t step) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->have_closures || c->p.dims != 1) return MCPM_ERR_STATE;
  const uint32_t slo = (uint32_t)step, shi = (uint32_t)((uint64_t)step >> 32);
  Timer tm(c, T_FPVEL);
  if (c->p.precision == MCPM_FP32)
    hipLaunchKernelGGL((k_fp_velocity<float, FP_CACHE_K>), dim3(c->nc), dim3(256), 0, c->s,
                       (Rec<float> *)c->P[c->cur], c->cnt, c->start, c->moments, c->clos, c->yc,
                       c->kp, c->fpp, slo, shi, c->err);
  else
    hipLaunchKernelGGL((k_fp_velocity<double, FP_CACHE_K>), dim3(c->nc), dim3(256), 0, c->s,
                       (Rec<double> *)c->P[c->cur], c->cnt, c->start, c->moments, c->clos, c->yc,
                       c->kp, c->fpp, slo, shi, c->err);
  LAUNCHCHK();
  return MCPM_OK;
}

int mcpm_get_closures(mcpm_ctx *c, double *out) {
  if (!c || !out) return MCPM_ERR_ARG;
  if (!c->have_closures) return MCPM_ERR_STATE;
  HIPCHK(hipMemcpyAsync(out, c->clos, (size_t)c->nc * MCPM_NCLOS * sizeof(double),
                        hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return MCPM_OK;
}

int mcpm_poisson_solver(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  if (c->p.long_range != MCPM_LR_SCREENED_POISSON) return MCPM_ERR_ARG;
  Timer tm(c, T_FIELD);
  if (c->p.dims == 2 && c->fno) {
    int st = fno_field_step(c->fno, c, c->cur_step, c->fno_accept, c->fno_every);
    if (st) return st;
    c->kick_pending = true;
    return MCPM_OK;
  }
  if (c->p.dims == 2) {
    int st = sp_2d(c);
    if (st) return st;
    c->kick_pending = true;
    return MCPM_OK;
  }
  hipLaunchKernelGGL(k_field, dim3(c->kp.nrep ? c->kp.nrep : 1), dim3(1024), 0, c->s, c->nden,
                     c->fa, c->F2, c->kp, c->p.ast, c->p.bst, c->p.eqdist,
                     (const FieldArrays *)c->rfa, (const StepDev *)nullptr);
  LAUNCHCHK();
  c->kick_pending = true;
  return MCPM_OK;
}

int mcpm_vlasov_integral(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  Timer tm(c, T_FIELD);
  if (c->p.dims == 2) {
    int st = vlasov_2d(c);
    if (st) return st;
    c->kick_pending = true;
    return MCPM_OK;
  }
  const int nf = c->p.boundary == MCPM_BC_PERIODIC ? c->nc : c->nc + 1;
  hipLaunchKernelGGL(k_vlasov_faces, dim3(nf), dim3(256), 0, c->s, c->nden, c->lr_tab, c->lr_V,
                     c->kp);
  hipLaunchKernelGGL(k_vlasov_force, dim3((c->nc + 255) / 256), dim3(256), 0, c->s, c->lr_V,
                     c->F2, c->kp, c->p.phi);
  LAUNCHCHK();
  c->kick_pending = true;
  return MCPM_OK;
}

static int run_kde(mcpm_ctx *c) {
  if (c->p.dims != 1) return MCPM_ERR_ARG;
  if (c->p.precision == MCPM_FP32)
    hipLaunchKernelGGL((k_kde<float>), dim3(c->nc), dim3(256), 0, c->s,
                       (const Rec<float> *)c->P[c->cur], c->cnt, c->start, c->kde,
                       c->kde + c->nc, c->kde + 2 * c->nc, c->kp);
  else
    hipLaunchKernelGGL((k_kde<double>), dim3(c->nc), dim3(256), 0, c->s,
                       (const Rec<double> *)c->P[c->cur], c->cnt, c->start, c->kde,
                       c->kde + c->nc, c->kde + 2 * c->nc, c->kp);
  LAUNCHCHK();
  return MCPM_OK;
}

int mcpm_density_expansion(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  Timer tm(c, T_FIELD);
  int st = run_kde(c);
  if (st) return st;
  const double s6 = c->p.sigma * pow(2.0, 1.0 / 6.0);
  const double pref = 4.0 * PI * (s6 * s6 * s6) * c->p.phi / c->p.mass * 2.0;
  hipLaunchKernelGGL(k_lr2_force, dim3((c->nc + 255) / 256), dim3(256), 0, c->s,
                     c->kde + c->nc, c->F2, c->kp, pref);
  LAUNCHCHK();
  c->kick_pending = true;
  return MCPM_OK;
}

int mcpm_get_kde(mcpm_ctx *c, double *ns, double *d2, double *d3) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (!c->sorted) return MCPM_ERR_STATE;
  int st = run_kde(c);
  if (st) return st;
  HIPCHK(hipStreamSynchronize(c->s));
  const size_t nb = (size_t)c->nc * sizeof(double);
  if (ns) HIPCHK(hipMemcpy(ns, c->kde, nb, hipMemcpyDeviceToHost));
  if (d2) HIPCHK(hipMemcpy(d2, c->kde + c->nc, nb, hipMemcpyDeviceToHost));
  if (d3) HIPCHK(hipMemcpy(d3, c->kde + 2 * c->nc, nb, hipMemcpyDeviceToHost));
  return MCPM_OK;
}

int mcpm_kick(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  if (!c->kick_pending) return MCPM_OK;
  int st = run_sort(c, PUSH_KICK, 0);
  c->kick_pending = false;
  return st;
}

int mcpm_stream_bc(mcpm_ctx *c, int64_t step, int32_t with_kick) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  int mode = PUSH_MOVE;
  if (with_kick && c->kick_pending) mode |= PUSH_KICK;
  int st = run_sort(c, mode, step);
  c->kick_pending = false;
  return st;
}

int mcpm_step(mcpm_ctx *c, int64_t step, mcpm_collision_stats *stats) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  Timer tm(c, T_STEP);
  int st = MCPM_OK;
  if (stats) memset(stats, 0, sizeof *stats);
  // velocity_update dispatch (Main_Functions.cpp:1545-1600)
  switch (c->p.model) {
    case MCPM_MODEL_ESMC: st = mcpm_collide_esmc(c, step, stats); break;
    case MCPM_MODEL_ESMC_NEW4: st = mcpm_collide_esmc_new4(c, step, stats); break;
    case MCPM_MODEL_FP_DENSE:
    case MCPM_MODEL_FP_IDEAL:
    case MCPM_MODEL_FP_LINEAR:
      st = mcpm_fp_closures(c, step);
      if (!st) st = mcpm_velocity_fp(c, step);
      break;
    case MCPM_MODEL_DSMC_VHS: st = mcpm_collide_dsmc_vhs(c, step, stats); break;
    case MCPM_MODEL_DSMC_HS: st = mcpm_collide_dsmc_hs(c, step, stats); break;
    case MCPM_MODEL_CBA_VHS: st = mcpm_collide_cba_vhs(c, step, stats); break;
    case MCPM_MODEL_CBA_HS: st = mcpm_collide_cba_hs(c, step, stats); break;
    default: break;
  }
  if (st) return st;
  // long-range attraction (Main_Functions.cpp:1612-1668)
  c->cur_step = step;
  if (c->p.long_range == MCPM_LR_SCREENED_POISSON) st = mcpm_poisson_solver(c);
  else if (c->p.long_range == MCPM_LR_DIRECT) st = mcpm_vlasov_integral(c);
  else if (c->p.long_range == MCPM_LR_DENSITY_EXPANSION) st = mcpm_density_expansion(c);
  if (st) return st;
  st = mcpm_stream_bc(c, step, 1);
  c->have_closures = false;
  return st;
}

int mcpm_graph_capture(mcpm_ctx *c, int64_t op_capacity, int32_t rounds) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 1 || c->p.model != MCPM_MODEL_ESMC ||
      (c->p.long_range != MCPM_LR_NONE && c->p.long_range != MCPM_LR_SCREENED_POISSON))
    return MCPM_ERR_ARG;
  if (!graph_ok(c)) return MCPM_ERR_STATE;
  const int nbands = std::max(1, std::min<int>(c->nbands, MAX_BANDS));
  if (rounds <= 0) rounds = nbands + 6;
  if (rounds <= nbands || rounds >= MAX_ROUNDS) return MCPM_ERR_ARG;
  if (op_capacity <= 0) op_capacity = std::max<int64_t>(2 * c->last_m, c->n) + 4096;
  return graph_capture(c, op_capacity, rounds);
}

int mcpm_graph_release(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  graph_release(c);
  c->g_cap = 0;
  return MCPM_OK;
}

int mcpm_graph_steps(mcpm_ctx *c, int64_t step0, int64_t nsteps, mcpm_collision_stats *stats) {
  if (!c || step0 < 0 || nsteps < 0) return MCPM_ERR_ARG;
  if (stats) memset(stats, 0, sizeof *stats);
  if (c->g_cap <= 0) return MCPM_ERR_STATE;  // nothing captured
  if (!graph_ok(c)) return MCPM_ERR_STATE;
  // lists reallocated or other particles since the capture
  if (!c->gexec[0] || !c->gexec[1] || c->n != c->g_n)
    if (int st = graph_capture(c, c->g_cap, c->g_rounds)) return st;
  if (!c->ys_valid) {
    const int st = c->p.precision == MCPM_FP32 ? ys_fill_t<Rec<float>>(c) : ys_fill_t<Rec<double>>(c);
    if (st) return st;
  }
  int64_t done = 0, rsum = 0, tail = 0;
  mcpm_collision_stats tot{};
  StepDev h{};
  h.lo = (uint32_t)step0;
  h.hi = (uint32_t)((uint64_t)step0 >> 32);
  h.mark = c->round_base;
  h.wdelta = (uint32_t)c->wdelta;
  // replays queued between looks at the run state: the replays behind a
  // step that stops run as no-ops (about 0.2 ms each), so a run starts with
  // a short batch, doubles it after every batch that ran through, up to 64,
  // and starts short again after a stop
  constexpr int64_t BATCH0 = 8, BATCH_MAX = 64;
  int64_t batch = BATCH0;
  bool eager = false;
  while (done < nsteps) {
    if (eager) {
      // the windowed sort no longer applies: the remaining steps one by one
      mcpm_collision_stats s1{};
      if (int st = mcpm_step(c, step0 + done, &s1)) return st;
      tot.candidates += s1.candidates;
      tot.valid += s1.valid;
      tot.accepted += s1.accepted;
      rsum += s1.rounds;
      ++done;
      continue;
    }
    const StepDev up = h;
    HIPCHK(hipMemcpyAsync(c->sd, &up, sizeof up, hipMemcpyHostToDevice, c->s));
    const int cur0 = c->cur;
    const int64_t todo = std::min(batch, nsteps - done);
    for (int64_t k = 0; k < todo; ++k) HIPCHK(hipGraphLaunch(c->gexec[(cur0 + k) & 1], c->s));
    HIPCHK(hipMemcpyAsync(&h, c->sd, sizeof h, hipMemcpyDeviceToHost, c->s));
    HIPCHK(hipStreamSynchronize(c->s));
    c->cur = (cur0 + (int)(h.steps & 1u)) & 1;
    done += h.steps;
    tot.candidates += (int64_t)h.cand;
    tot.valid += (int64_t)h.valid;
    tot.accepted += (int64_t)h.acc;
    rsum += (int64_t)h.rsum;
    tail += (int64_t)h.tail_ops;
    c->round_base = h.mark;
    int64_t next = (int64_t)(((uint64_t)h.hi << 32) | h.lo);
    batch = h.stop ? BATCH0 : std::min(2 * batch, BATCH_MAX);
    if (h.stop & STOP_ROUNDS) return MCPM_ERR_CAPACITY;
    if (h.stop & STOP_WINDOW) {
      // step `next` ran its collisions and field; a particle left its sort
      // window, so its push and sort run here (full histogram; the failed
      // windowed sort left the cell bounds unusable), and the windows widen
      // as after an eager step's failed windowed sort: the graphs read the
      // width from the run state, so they replay as captured
      c->kick_pending = c->p.long_range == MCPM_LR_SCREENED_POISSON;
      const bool we = c->w_enabled;
      c->w_enabled = false;
      const int sst = mcpm_stream_bc(c, next, 1);
      c->w_enabled = we;
      if (sst) return sst;
      ++c->w_runs;
      ++c->w_redone;
      if (c->wdelta < WDELTA_MAX) c->wdelta = std::min(2 * c->wdelta, (int)WDELTA_MAX);
      else ++c->w_fallbacks;
      ++done;
      ++next;
      if (!graph_ok(c)) {
        eager = true;
      } else {
        if (!c->gexec[0] || !c->gexec[1])
          if (int st = graph_capture(c, c->g_cap, c->g_rounds)) return st;
        // the two-digit sort (many cells) leaves no sorted y for the draw
        if (!c->ys_valid) {
          const int st = c->p.precision == MCPM_FP32 ? ys_fill_t<Rec<float>>(c)
                                                     : ys_fill_t<Rec<double>>(c);
          if (st) return st;
        }
      }
    } else if (h.stop & STOP_CAPACITY) {
      // nothing of step `next` ran: longer operation lists
      if (int st = graph_capture(c, 2 * c->g_cap, c->g_rounds)) return st;
    }
    h = StepDev{};
    h.lo = (uint32_t)next;
    h.hi = (uint32_t)((uint64_t)next >> 32);
    h.mark = c->round_base;
    h.wdelta = (uint32_t)c->wdelta;
  }
  if (done > 0) {
    // as after as many mcpm_step calls
    c->sorted = true;
    c->order_sorted = true;
    c->ys_valid = true;
    c->kick_pending = false;
    c->keys_current = false;
    c->have_closures = false;
  }
  if (stats) {
    *stats = tot;
    stats->rounds = (int32_t)std::min<int64_t>(rsum, INT32_MAX);
    stats->revisits = (int32_t)std::min<int64_t>(tail, INT32_MAX);
  }
  return check_err(c);
}

int mcpm_moments(mcpm_ctx *c, double *out) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  int st = MCPM_DISPATCH(c, launch_moments, c);
  if (st || !out) return st;
  HIPCHK(hipMemcpyAsync(out, c->moments, (size_t)c->nc * MCPM_NMOM * sizeof(double),
                        hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return MCPM_OK;
}

int mcpm_sample(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  int st = MCPM_DISPATCH(c, launch_sample, c);
  if (!st) ++c->n_samples;
  return st;
}

int mcpm_get_nodes(mcpm_ctx *c, double *n_node, double *potential) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 2) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  const size_t nn = (size_t)(c->fem.nx + 1) * (c->fem.ny + 1);
  if (n_node) {
    int st = MCPM_DISPATCH(c, launch_nodes, c);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(n_node, c->n_node, nn * sizeof(double), hipMemcpyDeviceToHost, c->s));
  }
  if (potential)
    HIPCHK(hipMemcpyAsync(potential, c->pot, nn * sizeof(double), hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return MCPM_OK;
}

int mcpm_get_samples(mcpm_ctx *c, double *out) {
  if (!c || !out) return MCPM_ERR_ARG;
  HIPCHK(hipMemcpyAsync(out, c->samples, (size_t)c->nc * MCPM_NSAMP * sizeof(double),
                        hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return MCPM_OK;
}

int mcpm_reset_samples(mcpm_ctx *c) {
  if (!c) return MCPM_ERR_ARG;
  HIPCHK(hipMemsetAsync(c->samples, 0, (size_t)c->nc * MCPM_NSAMP * sizeof(double), c->s));
  HIPCHK(hipMemsetAsync(c->press, 0, (size_t)c->nc * sizeof(double), c->s));
  c->n_samples = 0;
  return MCPM_OK;
}

int mcpm_initialize(mcpm_ctx *c, const mcpm_settings *s) {
  if (!c || !s) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  if (s->n_particles < 0 || s->n_particles > c->p.capacity) return MCPM_ERR_CAPACITY;
  if (s->n_drop < 0 || s->n_drop > s->n_particles) return MCPM_ERR_ARG;
  const bool disc = s->problem == 12;
  if (disc != (c->p.dims == 2)) return MCPM_ERR_ARG;
  InitParams ip{};
  ip.n = s->n_particles;
  ip.n_drop = disc ? s->n_drop : 0;
  ip.disc = disc ? 1 : 0;
  for (int q = 0; q < 3; ++q) {
    ip.dirs[q] = s->direction[q];
    ip.len[q] = s->box_len[q];
    ip.U0[q] = s->U0[q];
  }
  // the context's own geometry along its axes
  ip.len[1] = c->p.length;
  if (c->p.dims == 2) ip.len[0] = c->p.length_x;
  ip.drop_r = s->drop_dia / 2.0;
  ip.sd = sqrt(c->p.kb * s->T_init / c->p.mass);
  int st = MCPM_DISPATCH(c, launch_init, c, ip);
  if (st) return st;
  c->n = s->n_particles;
  c->sorted = false;
  c->have_closures = false;
  c->kick_pending = false;
  return mcpm_cell_update(c);
}

int mcpm_measure_pressure(mcpm_ctx *c, int32_t enable) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 1) return MCPM_ERR_ARG;
  c->pressure_on = enable != 0;
  return MCPM_OK;
}

int mcpm_get_escaped(mcpm_ctx *c, int64_t *count) {
  if (!c || !count) return MCPM_ERR_ARG;
  unsigned long long v = 0;
  HIPCHK(hipMemcpyAsync(&v, c->esc, sizeof v, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  *count = (int64_t)v;
  return MCPM_OK;
}

int mcpm_energy_momentum(mcpm_ctx *c, double *out) {
  if (!c || !out) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_ARG;
  return MCPM_DISPATCH(c, launch_energy, c, out);
}

int mcpm_track_ids(mcpm_ctx *c, int32_t enable) {
  if (!c) return MCPM_ERR_ARG;
  if (!enable) {
    c->track_ids = false;
    return MCPM_OK;
  }
  if (!c->ids[0]) {
    int st = dalloc(&c->ids[0], 2 * (size_t)c->p.capacity);
    if (st) return st;
    c->ids[1] = c->ids[0] + c->p.capacity;
  }
  c->track_ids = true;
  if (c->n > 0) {
    hipLaunchKernelGGL(k_iota_u32, dim3(1024), dim3(256), 0, c->s, c->ids[c->cur], c->n);
    LAUNCHCHK();
  }
  return MCPM_OK;
}

int mcpm_download_ids(mcpm_ctx *c, int32_t *ids) {
  if (!c || (!ids && c->n > 0)) return MCPM_ERR_ARG;
  if (!c->track_ids) return MCPM_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->s));
  if (c->n > 0)
    HIPCHK(hipMemcpy(ids, c->ids[c->cur], (size_t)c->n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return MCPM_OK;
}

int mcpm_set_cell_ratio(mcpm_ctx *c, const int32_t *ratio) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 1 || c->kp.nrep) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  if (!ratio) {  // every cell back to p.n_ratio
    c->kp.cell_ratio = nullptr;
    c->track_w = false;
    c->w_pending = false;
  } else {
    const int nc = c->nc;
    for (int q = 0; q < nc; ++q)
      if (ratio[q] < 1) return MCPM_ERR_ARG;
    int st;
    if (!c->cratio) {
      const size_t cap = (size_t)c->p.capacity;
      if ((st = dalloc(&c->cratio, (size_t)nc))) return st;
      if ((st = dalloc(&c->pw[0], 2 * cap))) return st;
      c->pw[1] = c->pw[0] + cap;
      if ((st = dalloc(&c->ar_fl, cap))) return st;
      if ((st = dalloc(&c->ar_cst, 4 * (size_t)nc + 2 * ((size_t)nc + 1)))) return st;
      c->ar_tot = c->ar_cst + 4 * (size_t)nc;
      c->ar_ostart = c->ar_tot + ((size_t)nc + 1);
      if ((st = dalloc(&c->ar_h, (size_t)nc))) return st;
    }
    HIPCHK(hipMemcpy(c->cratio, ratio, (size_t)nc * sizeof(int32_t), hipMemcpyHostToDevice));
    c->kp.cell_ratio = c->cratio;
    c->w_pending = true;
    c->track_w = false;
    if (c->sorted) {
      // every particle takes its cell's weight, the densities the new weights
      if ((st = weights_after_sort(c))) return st;
      hipLaunchKernelGGL(k_cell_fill, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->s,
                         c->start, nc, c->cnt, c->nden, c->yc, c->kp);
      LAUNCHCHK();
    }
  }
  return MCPM_OK;
}

int mcpm_set_particle_ratio(mcpm_ctx *c, const int32_t *w) {
  if (!c || (!w && c->n > 0)) return MCPM_ERR_ARG;
  if (!c->kp.cell_ratio) return MCPM_ERR_STATE;
  for (int64_t q = 0; q < c->n; ++q)
    if (w[q] < 1) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  if (c->n > 0)
    HIPCHK(hipMemcpy(c->pw[c->cur], w, (size_t)c->n * sizeof(int32_t), hipMemcpyHostToDevice));
  c->track_w = true;
  c->w_pending = false;
  return MCPM_OK;
}

int mcpm_get_particle_ratio(mcpm_ctx *c, int32_t *w) {
  if (!c || (!w && c->n > 0)) return MCPM_ERR_ARG;
  if (!c->track_w) return MCPM_ERR_STATE;
  HIPCHK(hipStreamSynchronize(c->s));
  if (c->n > 0)
    HIPCHK(hipMemcpy(w, c->pw[c->cur], (size_t)c->n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return MCPM_OK;
}

int mcpm_addremove(mcpm_ctx *c, int32_t n_ratio, int64_t step, mcpm_addremove_stats *stats) {
  if (!c || n_ratio < 1) return MCPM_ERR_ARG;
  if (c->p.dims != 1 || c->kp.nrep || c->track_ids) return MCPM_ERR_ARG;
  if (!c->kp.cell_ratio || !c->track_w || !c->sorted) return MCPM_ERR_STATE;
  return MCPM_DISPATCH(c, addremove_t, c, n_ratio, step, stats);
}

int mcpm_thermostat(mcpm_ctx *c, double zone_lo, double T_lo, double zone_hi, double T_hi) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 1 || c->kp.nrep) return MCPM_ERR_ARG;
  if (!(T_lo >= 0.0) || !(T_hi >= 0.0)) return MCPM_ERR_ARG;
  if (!c->sorted) return MCPM_ERR_STATE;
  return MCPM_DISPATCH(c, launch_thermostat, c, zone_lo, T_lo, zone_hi, T_hi);
}

int mcpm_get_pressure(mcpm_ctx *c, double *p_m, double *p_rep, double *p_tot) {
  if (!c) return MCPM_ERR_ARG;
  if (c->p.dims != 1) return MCPM_ERR_ARG;
  const size_t nc = (size_t)c->nc;
  if (p_rep || p_tot) {
    if (!c->sorted) return MCPM_ERR_STATE;
    int st = MCPM_DISPATCH(c, launch_moments, c);
    if (st) return st;
    hipLaunchKernelGGL(k_pressure, dim3((c->nc + 255) / 256), dim3(256), 0, c->s, c->moments,
                       c->nden, c->yc, c->press + nc, c->press + 2 * nc, c->kp, c->p.phi);
    LAUNCHCHK();
  }
  if (p_m)
    HIPCHK(hipMemcpyAsync(p_m, c->press, nc * sizeof(double), hipMemcpyDeviceToHost, c->s));
  if (p_rep)
    HIPCHK(hipMemcpyAsync(p_rep, c->press + nc, nc * sizeof(double), hipMemcpyDeviceToHost,
                          c->s));
  if (p_tot)
    HIPCHK(hipMemcpyAsync(p_tot, c->press + 2 * nc, nc * sizeof(double),
                          hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return MCPM_OK;
}

int mcpm_init_omega_max(mcpm_ctx *c, double n_gas, int32_t times_y) {
  if (!c) return MCPM_ERR_ARG;
  if (c->kp.nrep) return MCPM_ERR_STATE;  // not on batched replicas
  std::vector<double> mom((size_t)c->nc * MCPM_NMOM), nd(c->nc), om(c->nc);
  int st = mcpm_moments(c, mom.data());
  if (st) return st;
  st = mcpm_get_cells(c, nullptr, nullptr, nd.data(), nullptr, nullptr);
  if (st) return st;
  const double y = times_y ? contact_y_host(&c->p, n_gas) : 1.0;
  for (int k = 0; k < c->nc; ++k) {
    double T = mom[(size_t)k * MCPM_NMOM + 29];
    double crm = 10.0 * sqrt(c->p.kb * T / c->p.mass);               // main.cpp:259
    double o = 4.0 * PI * c->p.sigma * c->p.sigma * crm * nd[k] * c->p.delta_t;  // :260
    om[k] = o * y;                                                       // :268-269
  }
  return mcpm_set_omega_max(c, om.data());
}

int mcpm_set_omega_max(mcpm_ctx *c, const double *om) {
  if (!c || !om) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  HIPCHK(hipMemcpy(c->omax, om, c->nc * sizeof(double), hipMemcpyHostToDevice));
  // the running maximum of the next collision call starts afresh
  HIPCHK(hipMemset(c->omega_new, 0, c->nc * sizeof(double)));
  return MCPM_OK;
}

int mcpm_get_cells(mcpm_ctx *c, int32_t *count, int32_t *start, double *n, double *omega_max,
                   double *F2) {
  if (!c) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  const size_t nc = c->nc;
  if (count) HIPCHK(hipMemcpy(count, c->cnt, nc * 4, hipMemcpyDeviceToHost));
  if (start) HIPCHK(hipMemcpy(start, c->start, nc * 4, hipMemcpyDeviceToHost));
  if (n) HIPCHK(hipMemcpy(n, c->nden, nc * 8, hipMemcpyDeviceToHost));
  if (omega_max) HIPCHK(hipMemcpy(omega_max, c->omax, nc * 8, hipMemcpyDeviceToHost));
  if (F2) HIPCHK(hipMemcpy(F2, c->F2, nc * 8, hipMemcpyDeviceToHost));
  return MCPM_OK;
}

double mcpm_contact_y(const mcpm_params *p, double n) { return p ? contact_y_host(p, n) : 0.0; }

int mcpm_debug_collision_log(mcpm_ctx *c, int32_t enable) {
  if (!c) return MCPM_ERR_ARG;
  c->debug_log = enable != 0;
  return MCPM_OK;
}

int mcpm_get_collision_log(mcpm_ctx *c, int64_t cap, uint32_t *prio, uint32_t *i, uint32_t *j,
                           uint8_t *accepted, int64_t *n_out) {
  if (!c || !n_out) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  int64_t n = c->last_log;
  *n_out = n;
  if (n > cap) return MCPM_ERR_CAPACITY;
  if (n == 0) return MCPM_OK;
  if (prio) HIPCHK(hipMemcpy(prio, c->log.p, n * 4, hipMemcpyDeviceToHost));
  if (i) HIPCHK(hipMemcpy(i, c->log.i, n * 4, hipMemcpyDeviceToHost));
  if (j) HIPCHK(hipMemcpy(j, c->log.j, n * 4, hipMemcpyDeviceToHost));
  if (accepted) HIPCHK(hipMemcpy(accepted, c->log.a, n, hipMemcpyDeviceToHost));
  return MCPM_OK;
}

int mcpm_set_tuning(mcpm_ctx *c, int32_t knob, int64_t v) {
  if (!c) return MCPM_ERR_ARG;
  switch (knob) {
    case MCPM_TUNE_BANDS:
      if (v < 1 || v > MAX_BANDS) return MCPM_ERR_ARG;
      c->nbands = (int)v;
      return MCPM_OK;
    case MCPM_TUNE_ROUND_GRID:
      if (v < 8 || v > 65536 || v % 8) return MCPM_ERR_ARG;
      c->round_grid = (int)v;
      return MCPM_OK;
    case MCPM_TUNE_SORT_CHUNK: {
      if (v != 0 && (v < 1024 || v > 32768 || v % 256)) return MCPM_ERR_ARG;
      // the chunk table holds nchunks_cap chunks
      if (v != 0 && (c->p.capacity + v - 1) / v > c->nchunks_cap) return MCPM_ERR_ARG;
      c->sort_chunk = (int)v;
      return MCPM_OK;
    }
    case MCPM_TUNE_SORT_WINDOW:
      if (v != 0 && v != 1) return MCPM_ERR_ARG;
      c->w_enabled = v == 1;
      c->w_fallbacks = 0;
      return MCPM_OK;
    case MCPM_TUNE_ROUND_UNROLL:
      if (v != 0 && v != 2 && v != 4) return MCPM_ERR_ARG;
      c->round_unroll = (int)v;
      return MCPM_OK;
    default: return MCPM_ERR_ARG;
  }
}

int mcpm_get_tuning(mcpm_ctx *c, int32_t knob, int64_t *v) {
  if (!c || !v) return MCPM_ERR_ARG;
  switch (knob) {
    case MCPM_TUNE_BANDS: *v = c->nbands; return MCPM_OK;
    case MCPM_TUNE_ROUND_GRID: *v = c->round_grid; return MCPM_OK;
    case MCPM_TUNE_SORT_CHUNK: *v = c->sort_chunk; return MCPM_OK;
    case MCPM_TUNE_SORT_WINDOW: *v = c->w_enabled ? 1 : 0; return MCPM_OK;
    case MCPM_TUNE_SORT_WINDOW_RUNS: *v = c->w_runs; return MCPM_OK;
    case MCPM_TUNE_SORT_WINDOW_REDONE: *v = c->w_redone; return MCPM_OK;
    case MCPM_TUNE_ROUND_UNROLL: *v = c->round_unroll; return MCPM_OK;
    default: return MCPM_ERR_ARG;
  }
}

int mcpm_timing_enable(mcpm_ctx *c, int32_t enable) {
  if (!c) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  for (int k = 0; k < T_NCLASS; ++k) {
    for (auto &pr : c->ev[k]) {
      c->ev_pool.push_back(pr.first);
      c->ev_pool.push_back(pr.second);
    }
    c->ev[k].clear();
  }
  c->timing = enable != 0;
  return MCPM_OK;
}

int mcpm_timing_get(mcpm_ctx *c, int32_t which, double *avg_ms, int64_t *calls) {
  if (!c || which < 0 || which >= T_NCLASS) return MCPM_ERR_ARG;
  HIPCHK(hipStreamSynchronize(c->s));
  double tot = 0.0;
  for (auto &pr : c->ev[which]) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
    tot += ms;
  }
  int64_t k = (int64_t)c->ev[which].size();
  if (avg_ms) *avg_ms = k ? tot / (double)k : 0.0;
  if (calls) *calls = k;
  return MCPM_OK;
}

}  // extern "C"
