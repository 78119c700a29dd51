// mcpm_kernels.h — kernel declarations and the device-side data structures
// shared between mcpm_kernels.hip and the host layer (mcpm_api.hip).
#pragma once
#include "mcpm_device.h"

namespace mcpm {

// push modes
enum : int { PUSH_KICK = 1, PUSH_MOVE = 2, PUSH_CHECK = 4 };

// device error bits -> status codes
enum : uint32_t {
  ERRBIT_DOMAIN = 1u,
  ERRBIT_NAN = 2u,
  ERRBIT_BC = 4u,
  ERRBIT_CAPACITY = 8u,
  ERRBIT_ORDER = 16u  // sorted keys out of order (an internal error)
};

// layout of the u32 device counter block
enum : int {
  CNT_OPS = 0,     // valid candidate operations of the current step
  CNT_ACC = 2,     // accepted collisions (u64, 8-byte aligned)
  CNT_LOG = 4,     // entries in the debug collision log
  CNT_SCANNED = 6, // candidate lists whose chunk counts k_cand_scan has scanned
  NREG = 8,        // operation regions: contiguous candidate ranges, one per XCD
  CNT_LIST = 8,    // CNT_LIST + r NREG + x: losers of round r - 1 in region x (r >= 1)
  MAX_ROUNDS = 4096,
  MAX_BANDS = 16,
  CNT_BAND = CNT_LIST + (MAX_ROUNDS + 2) * NREG,  // [b NREG + x]: operations of band b, region x
  CNT_BSTART = CNT_BAND + MAX_BANDS * NREG,       // start of list [b NREG + x] in the band lists
  // the draw's valid-operation count, spread over NSPREAD slots one cache
  // line apart (one address serializes its atomics: ~30k per draw at 20M
  // particles took 175 us), folded into CNT_OPS by round 0's k_check
  NSPREAD = 64,
  SPREAD_STRIDE = 32,
  CNT_SPREAD = (CNT_BSTART + MAX_BANDS * NREG + 1 + SPREAD_STRIDE - 1) / SPREAD_STRIDE * SPREAD_STRIDE,
  CNT_TOTAL = CNT_SPREAD + NSPREAD * SPREAD_STRIDE
};

// A shared-particle mark of the reservation rounds covers a group of
// 2^MARK_SHIFT consecutive particles (neighbours in cell order).  A mark on
// a group whose particle has a single operation only makes that operation
// read its particle's priority word, which holds its own priority; the
// smaller table stays in an XCD's L2 while the rounds gather it.
constexpr int MARK_SHIFT = 5;

// State of a run of graph-replayed steps, in device memory.  The kernels of
// a captured step read the step number, its priority key and its round tags
// from here instead of from their arguments, and return at once while stop
// is set: a step that cannot finish on the device (its candidates exceed
// the operation lists, a particle left its sort window, rounds ran out)
// stops every later kernel of the run, and the host completes or reports it.
struct StepDev {
  uint32_t lo, hi;   // step number
  uint32_t pk;       // the step's candidate priority key (k_step_mid)
  uint32_t mark;     // tag of the step's round 0 (round r: mark + r, low byte)
  uint32_t stop;     // STOP_* bits
  uint32_t rounds;   // rounds the step used (k_round_tail)
  uint32_t steps;    // steps completed in the run
  uint32_t wdelta;   // slack cells of the sort windows (widened by the host between runs)
  unsigned long long cand, valid, acc, rsum, tail_ops;  // totals over the completed steps
};
enum : uint32_t { STOP_WINDOW = 1u, STOP_CAPACITY = 2u, STOP_ROUNDS = 4u };
__device__ __forceinline__ bool sd_stopped(const StepDev *sd) {
  return sd != nullptr && *(volatile const uint32_t *)&sd->stop != 0u;
}

// Region of candidate k of M: the candidates are in cell order, so a region
// is a run of neighbouring cells.
__host__ __device__ __forceinline__ uint32_t op_region(uint32_t k, uint32_t M) {
  return (uint32_t)(((uint64_t)k * NREG) / (M ? M : 1u));
}

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

// collision models sharing the reservation rounds: ESMC in the canonical
// order of Collision_ESMC_new3, DSMC in cell order, ESMC in the order of
// Collision_ESMC_new4 (cells descending, each cell's candidates in turn)
enum : int { COLL_ESMC = 0, COLL_DSMC = 1, COLL_ESMC4 = 2 };

// R is Rec<float|double> (1D slab) or Rec2<float|double> (2D plane).
template <typename R, bool SCATTER>
__global__ void k_push(const R *in, R *out, int64_t n, int32_t chunk, KParams kp, Kick kick,
                       const double *drift, int mode,
                       uint32_t step_lo, uint32_t step_hi, uint32_t *table,
                       const uint32_t *start, uint32_t *err, int nb, int kmode, int32_t kdiv,
                       uint32_t *okey, const uint32_t *ikey, const uint32_t *iid,
                       uint32_t *oid, typename R::real_type *oys, const int32_t *iw,
                       int32_t *ow);
// The windowed sort of particles that move a few cells per step: WCH
// particles per chunk, histograms over at most WMAX cells, WDELTA cells of
// slack either side of the chunk's old cells.
// A chunk covers at most WSPAN old cells; its window adds wdelta cells each
// side, WDELTA at first and up to WDELTA_MAX after particles left it.
constexpr int WCH = 1024;
constexpr int WMAX = 128;
constexpr int WDELTA = 4;
constexpr int WSPAN = 32;
constexpr int WDELTA_MAX = (WMAX - WSPAN) / 2;
static_assert(WSPAN + 2 * WDELTA_MAX <= WMAX, "window of a chunk");
// PK: 0 the push of any slab, 1 of a periodic slab without position drift
// and face sums.
template <typename R, bool SCATTER, int PK>
__global__ void k_wpush(const R *in, R *out, KParams kp, Kick kick, const double *drift, int mode,
                        uint32_t step_lo, uint32_t step_hi, const uint32_t *wbeg,
                        const int32_t *wrep, const uint32_t *nchp, int32_t *wlo, int32_t *wlen,
                        uint32_t *wtab, const uint32_t *start, uint32_t *err, uint32_t *wflag,
                        const uint32_t *iid, uint32_t *oid, int wdelta,
                        typename R::real_type *oys, const int32_t *iw, int32_t *ow,
                        const StepDev *sd);
__global__ void k_wblocks(const int32_t *bcell, int nblk, const uint32_t *start,
                          uint32_t *pieces, const StepDev *sd);
__global__ void k_wfill(const int32_t *bcell, const int32_t *brep, int nblk,
                        const uint32_t *start, const uint32_t *pbase, const int32_t *rblk,
                        int nrep1, uint32_t n, uint32_t *wbeg, int32_t *wrep, uint32_t *nchp,
                        uint32_t *rchunk, const StepDev *sd);
// cpw: consecutive cells per wavefront (more than one on long cell arrays,
// where one wavefront per cell ran generation after generation of waves)
__global__ void k_wscan(const int32_t *wlo, const int32_t *wlen, uint32_t *wtab,
                        const uint32_t *nchp, const uint32_t *rchunk, KParams kp, uint32_t *cnt,
                        double *nden, double *yc, int cpw, const StepDev *sd);
// Most bins of one sort pass (one LDS histogram of 64 KB per chunk); more
// cells sort by two digits of the cell key, up to MAX_CELLS.
constexpr int SORT_BINS = 16384;
// Wavefronts of one sort workgroup (they would share its chunk's histogram;
// one measured fastest: the push is bound by its double-precision arithmetic).
constexpr int PUSH_WAVES = 1;
constexpr int64_t MAX_CELLS = (int64_t)1 << 22;
__global__ void k_cell_bounds(const uint32_t *key, int64_t n, int nc, uint32_t *start,
                              uint32_t *err);
__global__ void k_cell_fill(const uint32_t *start, int nc, uint32_t *cnt, double *nden,
                            double *yc, KParams kp);
__global__ void k_scan_a(const uint32_t *table, uint32_t *partial, int nchunks, int nc, int seg);
__global__ void k_scan_bins(uint32_t *partial, int nseg, int nc, uint32_t *cnt, double *nden,
                            double *yc, KParams kp);
__global__ void k_scan_c(uint32_t *table, const uint32_t *partial, int nchunks, int nc, int seg);
// variant: the mcpm_model of a COLL_DSMC collision (DSMC_VHS, DSMC_HS, CBA_VHS).
__global__ void k_mcand(const uint32_t *cnt, const double *omax, const double *nden,
                        const double *yc, uint32_t *mcand, int nc, KParams kp, int model,
                        int variant, uint32_t step_lo, uint32_t step_hi, const StepDev *sd);
__global__ void k_excl_scan(const uint32_t *in, uint32_t *out, int nc, const StepDev *sd);
// the same scan over several workgroups for long arrays (SCAN_TILE each)
constexpr int SCAN_TILE = 4096;
__global__ void k_tile_sums(const uint32_t *in, int n, uint32_t *tsum, const StepDev *sd);
__global__ void k_tile_scan(const uint32_t *in, uint32_t *out, int n, const uint32_t *toff,
                            int nt, const StepDev *sd);
// Priority band of an operation: the top bits of the priority scaled to
// nbands (band_mul = nbands for priorities spread over 32 bits, 0 for one band).
__device__ __forceinline__ uint32_t op_band(uint32_t pr, uint32_t band_mul, int nbands) {
  uint32_t b = (uint32_t)(((uint64_t)pr * band_mul) >> 32);
  return b < (uint32_t)nbands ? b : (uint32_t)nbands - 1u;
}
// Collision_ESMC_new4 processes the cells in descending order and each
// cell's candidates in turn (collisions.cpp:758-835): the rank of candidate t
// of cell c in that order is its priority, and its band the rank's share of
// the M candidates.
__device__ __forceinline__ uint32_t rank_new4(uint32_t c, uint32_t t, const uint32_t *cstart,
                                              uint32_t M) {
  return (M - cstart[c + 1]) + t;
}
__device__ __forceinline__ uint32_t band_new4(uint32_t rank, uint32_t M, int nbands) {
  uint32_t b = (uint32_t)(((uint64_t)rank * (uint64_t)nbands) / (uint64_t)(M ? M : 1u));
  return b < (uint32_t)nbands ? b : (uint32_t)nbands - 1u;
}
// Candidates are generated in chunks of GEN_CHUNK consecutive candidate
// indices (cell order).  A candidate's list (priority band, region) depends
// only on its index, so the lists are sized before anything is drawn.
constexpr uint32_t GEN_CHUNK = 2048;
constexpr int NLIST = MAX_BANDS * NREG;
// table[q tstride + chunk]: the chunk's candidates in list q, then (after
// k_cand_scan) their offset in the list.
__global__ void k_cand_count(uint32_t M, uint32_t prio_key, int model, uint32_t band_mul,
                             int nbands, uint32_t *table, uint32_t tstride, KParams kp,
                             const uint32_t *cstart, int nc, uint32_t *gcells, const StepDev *sd);
// cells of one candidate chunk whose candidate starts the draw keeps in LDS,
// and partner cells whose counts and starts it keeps there
constexpr int GEN_CELLS = 512;
constexpr int GEN_WIN = 1024;
__global__ void k_cand_scan(uint32_t *table, uint32_t tstride, uint32_t nchunks, int nl,
                            uint32_t *counters, const uint32_t *Mp, const StepDev *sd);
template <typename R, int MODEL>
__global__ void k_gen(const R *P, const typename R::real_type *ys, const uint32_t *cnt,
                      const uint32_t *start,
                      const uint32_t *cstart, int nc, uint32_t M, KParams kp, uint32_t step_lo,
                      uint32_t step_hi, uint32_t prio_key, uint32_t band_mul, int nbands,
                      const uint32_t *table, uint32_t tstride, const uint32_t *gcells,
                      int32_t reach,
                      uint4 *bands, uint32_t *tab0, uint32_t *counters, const StepDev *sd);
// A reservation round: k_check resolves the particles that several of the
// round's operations share, k_round commits the winners and admits the next
// round's operations.
template <int MODEL>
__global__ void k_check(int r, int nbands, uint32_t seg, const uint4 *in, const uint4 *bands,
                        uint32_t *tab, uint8_t *cflag, uint32_t tag8, uint32_t *counters,
                        uint32_t *host_left, const StepDev *sd);
// PLAIN: no batched replicas and no debug log (their code left out).  U:
// operations per thread per tile (2 or 4): more of them put more loads in
// flight per wavefront, fewer leave room for more wavefronts.
template <typename R, int MODEL, bool PLAIN, int U>
__global__ void k_round(int r, int nbands, uint32_t seg, R *P, const uint4 *in,
                        const uint4 *bands, uint4 *out, const uint32_t *tab_cur,
                        uint32_t *tab_next, const uint8_t *cflag, uint32_t tag8,
                        uint32_t *counters,
                        const double *nden, const double *yc, const double *omax,
                        double *omega_new, KParams kp, uint32_t step_lo, uint32_t step_hi,
                        FPParams fpp, OpLog log, const StepDev *sd);
// The rounds left after a captured step's grid rounds, in one workgroup:
// round r0 (>= nbands: every band admitted) onwards until no operation is
// left; the step's round count goes to sd->rounds.
template <typename R, int MODEL>
__global__ void k_round_tail(int r0, int nbands, uint32_t seg, R *P, uint4 *opA, uint4 *opB,
                             uint32_t *tab0, uint32_t *tab1, uint8_t *cflag, uint32_t *counters,
                             const double *nden, const double *yc, const double *omax,
                             double *omega_new, KParams kp, FPParams fpp, StepDev *sd);
__global__ void k_omega_finish(double *omax, double *omega_new, const uint32_t *cnt, int nc,
                               int model, int variant, const StepDev *sd);
// Graph-run bookkeeping: after the candidate scan (capacity check, priority
// key), and at the end of a step (totals, window check, next step).
__global__ void k_step_mid(StepDev *sd, const uint32_t *cstart, int nc, uint32_t cap,
                           uint32_t key0, uint32_t key1);
__global__ void k_step_end(StepDev *sd, const uint32_t *counters, const uint32_t *cstart, int nc,
                           const uint32_t *wflag);
// y of every record into the sorted-y array (the draw's positions)
template <typename R>
__global__ void k_ys_from_records(const R *P, int64_t n, typename R::real_type *ys);
// Most candidates one cell may run in one Collision_CBA_HS call.
constexpr uint32_t CBA_MAX_CAND = 1u << 26;
template <typename R>
__global__ void k_cba_hs(R *P, const uint32_t *cnt, const uint32_t *start, const double *mom,
                         const double *nden, KParams kp, uint32_t step_lo, uint32_t step_hi,
                         uint32_t *counters, uint32_t *err);
// Batched replicas, per step: each replica's candidate priority key and its
// first candidate (rcand_off[nrep] = all candidates).
__global__ void k_rep_step(KParams kp, const uint32_t *cstart, uint32_t *rprio,
                           uint32_t *rcand_off, uint32_t step_lo, uint32_t step_hi,
                           const StepDev *sd);
__global__ void k_zero_u32(uint32_t *p, int n);
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v);
__global__ void k_iota_u32(uint32_t *p, int64_t n);
__global__ void k_field(const double *nden, FieldArrays fa, double *F2, KParams kp, double ast,
                        double bst, double eqdist, const FieldArrays *rfa, const StepDev *sd);
template <typename R>
__global__ void k_energy(const R *P, int64_t n, double *part);
__global__ void k_energy_soa(const double *u1, const double *u2, const double *u3, int64_t n,
                             double *part);
__global__ void k_energy_finish(const double *part, int nblk, double mf, double *out);
// workgroups of the energy check (its per-workgroup partials: 4 doubles each)
constexpr int ENERGY_GRID = 1024;
template <typename R>
__global__ void k_thermostat(R *P, int64_t n, const double *mom, KParams kp, double zone_lo,
                             double T_lo, double zone_hi, double T_hi);
template <typename R>
__global__ void k_moments(const R *P, const uint32_t *cnt, const uint32_t *start,
                          double *out, KParams kp);
template <typename R>
__global__ void k_sample(const R *P, const uint32_t *cnt, const uint32_t *start,
                         double *acc);
// The initial state of initialization (Main_Functions.cpp:4222-4815).
struct InitParams {
  int64_t n, n_drop;        // particles; the last n_drop fill the disc
  int32_t dirs[3];          // streaming axes (box.direction)
  int32_t disc;             // injection: background + centred disc
  double len[3];            // box.Len
  double drop_r;            // disc radius
  double U0[3], sd;         // Maxwellian mean and standard deviation
};
template <typename R>
__global__ void k_init(R *P, InitParams ip, KParams kp);
__global__ void k_pressure(const double *mom, const double *nden, const double *yc,
                           double *p_rep, double *p_tot, KParams kp, double phi);
__global__ void k_fp_closures(const double *mom, const uint32_t *cnt, const double *nden,
                              const double *yc, double *clos, double *drift, KParams kp,
                              FPParams fpp, int64_t step, uint32_t *err);
// particles per thread whose noise the FP velocity kernel keeps in registers
// (cells up to 256 x this)
constexpr int FP_CACHE_K = 12;
template <typename real, int KMAX>
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

// ---- particle add/remove (mcpm_addremove.hip) --------------------------------
enum : uint8_t { AR_KEEP = 0, AR_SPLIT = 1, AR_DELETE = 2 };
template <typename R>
__global__ void k_ar_classify(const R *P, const uint32_t *cnt, const uint32_t *start,
                              const int32_t *pw, const int32_t *cr, int32_t nr, KParams kp,
                              uint32_t slo, uint32_t shi, uint8_t *fl, uint32_t *cst,
                              double *ch, uint32_t *tot);
template <typename R>
__global__ void k_ar_scatter(const R *P, R *out, const uint32_t *cnt, const uint32_t *start,
                             const uint32_t *ostart, const uint8_t *fl, const uint32_t *cst,
                             const double *ch, const int32_t *cr, int32_t nr, KParams kp,
                             uint32_t slo, uint32_t shi, int32_t *ow,
                             typename R::real_type *oys);
__global__ void k_w_from_cells(const uint32_t *start, int nc, const int32_t *cr, int32_t *w);

// ---- the plane's mean field (mcpm_fem.hip) ----------------------------------
// Constants of the nodal deposit and the Q1 screened-Poisson solve.
struct FemParams {
  int32_t nx, ny;          // cells per axis; nodes (nx+1)(ny+1), node (i,j) at j(nx+1)+i
  int32_t per_x, per_y;    // periodic axes (1) or specular walls (0)
  double fnv;              // Fn / V: one particle's share of a node density
  double lam2V;            // lambda^2 V, the mass-matrix factor (FEM.cpp:301)
  double volume;           // V, the right-hand side factor (FEM.cpp:303)
  double core0, core2;     // core subtraction constants dummy, dummy2 (FEM.cpp:497-502)
  double tol2;             // CG stop: |r|^2 <= tol2 |b|^2
  int32_t maxit;
  int32_t nq;              // quadrature points per dimension of the node integral
};
// Quadrature table layout of k_node_integral, nq entries each: r, radial
// weight, cos(azimuth), sin(azimuth), azimuth weight, sin(polar), polar weight.
enum : int { QT_R = 0, QT_RAD = 1, QT_CPH = 2, QT_SPH = 3, QT_DPH = 4, QT_STH = 5, QT_DTH = 6,
             QT_N = 7 };
template <typename R>
__global__ void k_quadrants(const R *P, const uint32_t *cnt, const uint32_t *start, uint32_t *q4,
                            KParams kp);
__global__ void k_nodes(const uint32_t *q4, double *n_node, FemParams f);
template <bool INVERSE>
__global__ void k_fft_rows(const double *n_node, double2 *buf, double *pot, const double2 *twx,
                           int lgx, FemParams f);
__global__ void k_fft_cols(double2 *buf, const double2 *twy, int lgy, const double *cosx,
                           const double *cosy, int cols, FemParams f);
__global__ void k_quadrants_keys(const uint32_t *keys, const uint32_t *cnt, const uint32_t *start,
                                 uint32_t *q4, int nc);
__global__ void k_node_integral(const double *nden, const double *tab, double *pot, KParams kp,
                                FemParams f, int boundary_only);
// Multi-workgroup CG (large planes): CG_GRID workgroups of 256 threads;
// part holds 2 CG_GRID partial sums, st three scalars, flag two words.
constexpr int CG_GRID = 512;
__global__ void k_cgm_zero(double *pot, FemParams f);
__global__ void k_cgm_start(const double *n_node, const double *pot, double *r, double *p,
                            FemParams f, double *part);
__global__ void k_cgm_start2(double *st, const double *part, int G, FemParams f,
                             uint32_t *flag);
__global__ void k_cgm_aq(const double *p, double *q, FemParams f, double *part,
                         const uint32_t *flag);
__global__ void k_cgm_xr(double *pot, double *r, const double *p, const double *q, FemParams f,
                         double *part, const double *st, int k, const uint32_t *flag);
__global__ void k_cgm_p(double *p, const double *r, FemParams f, const double *part, double *st,
                        int k, uint32_t *flag);
__global__ void k_cgm_seams(double *pot, FemParams f);
__global__ void k_fem_cg(const double *n_node, double *pot, double *r, double *p, double *q,
                         FemParams f, uint32_t *iters);
__global__ void k_fem_core(const double *n_node, double *pot, FemParams f);

}  // namespace mcpm
