// mcpm_kernels.hip — HIP kernels of the 1D slab step for gfx950.
//
//   k_push<real, SCATTER>   kick + stream + boundary + cell key; per-chunk
//                           LDS histogram (SCATTER=false) or stable scatter
//                           into the other buffer (SCATTER=true)
//   k_scan_*                per-cell column scans of the chunk histograms
//   k_mcand, k_gen, k_round ESMC collisions by deterministic reservations
//   k_field                 1D screened-Poisson solve in one workgroup
//   k_moments, k_sample     per-cell FP64 moments / time-average sums
//
// Every kernel is written for 64-lane wavefronts.
#include "mcpm_kernels.h"
#include "mcpm_geom.h"

#pragma clang fp contract(off)

namespace mcpm {

// ---------------------------------------------------------------------------
// kick + stream + boundary of one particle
// ---------------------------------------------------------------------------
// measure_pressure (Main_Functions.cpp:13-34): every face between the cell of
// x2_old (after the drift, before streaming) and the cell of x2 (after the
// boundary) adds Fn m |U2| to the cell below it.  Periodic slab: the faces of
// the unwrapped flight, wrapped into [0, nc) (the reference's slab has walls).
// Faces outside the slab are dropped instead of written out of bounds.
__device__ __forceinline__ void pressure_faces(double *pm, const KParams &kp, double y_old,
                                               double y_new, double u2, bool periodic) {
  const double qa = mcpm_floor(y_old / kp.dy), qb = mcpm_floor(y_new / kp.dy);
  if (!(qa != qb) || !(__builtin_fabs(qa) < 1e12) || !(__builtin_fabs(qb) < 1e12)) return;
  const int64_t a = (int64_t)qa, b = (int64_t)qb;
  const int64_t nc = kp.n_cells;
  int64_t lo = a < b ? a : b, hi = a < b ? b : a;
  if (hi - lo > nc) hi = lo + nc;
  const double f = kp.Fn * kp.mass * __builtin_fabs(u2);
  for (int64_t k = lo; k < hi; ++k) {
    int64_t c = k;
    if (periodic) {
      c %= nc;
      if (c < 0) c += nc;
    } else if (c < 0 || c >= nc) {
      continue;
    }
    atomicAdd(&pm[c], f);
  }
}

// Position_FP_ideal_gas (Fokker_Planck.cpp:948-969) for the slab axis, the
// kick U2 += F2 dt of poisson_solver (Main_Functions.cpp:3574-3577), and the
// slab-axis part of BC_thermal_wall_3D (Main_Functions.cpp:4917-5248):
// a particle past a wall is put on the wall, re-emitted with a half-range
// Maxwellian at the wall temperature, and flies the time it spent past the
// wall with the new velocity.  Draws: Philox(slot, bounce, step, WALL).
// With have_f2 the caller has read kick.F2 at the particle's cell into f2.
// BC >= 0 fixes the boundary at compile time (kp.boundary must be it), and
// DRIFT = false / FACES = false drop the position drift and the face sums:
// the same operations for the cases kept, with the other paths (and the
// registers they hold) compiled out.
template <typename real, int BC = -1, bool DRIFT = true, bool FACES = true>
__device__ __forceinline__ Rec<real> push_one(Rec<real> r, uint64_t slot, const KParams &kp,
                                              const Kick &kick,
                                              const double *__restrict__ drift, int mode,
                                              uint32_t step_lo, uint32_t step_hi,
                                              uint32_t *err, double f2 = 0.0,
                                              bool have_f2 = false) {
  double y = (double)r.y;
  if (mode & PUSH_KICK) {
    if (!have_f2) f2 = kick.F2[cell_of(y, kp.dy, kp.n_cells)];
    r.u2 = (real)((double)r.u2 + f2 * kp.dt);
  }
  if (!(mode & PUSH_MOVE)) return r;
  double v = (double)r.u2;
  const double y_start = y;
  bool drifted = false;
  if (DRIFT && drift) {
    // DFP_extra_streaming, slab axis (Fokker_Planck.cpp:932-937)
    const double *d = drift + (size_t)cell_of(y, kp.dy, kp.n_cells) * DRIFT_N;
    if (d[11] != 0.0) {
      double a = (double)r.u1 - d[0], b = (double)r.u2 - d[1], c3 = (double)r.u3 - d[2];
      double e = a * a + b * b + c3 * c3;
      double dt = kp.dt;
      y = y + dt * (a * d[4] + b * d[5] + c3 * d[6] + d[7] * (e - 3.0 * kp.kb * d[3] / kp.mass)) +
          d[8] * (b * e - d[9]) * dt;
      y = y - dt * d[10];
      drifted = true;
    }
  }
  const double y_old = y;  // store_old_positions after DFP_extra_streaming
  y = y + v * kp.dt;
  // speed used for the time spent past a wall: the displacement this step
  const double v_eff = drifted ? (y - y_start) / kp.dt : v;
  const int bc = BC >= 0 ? BC : kp.boundary;
  if (bc == MCPM_BC_PERIODIC) {
    if (FACES && kick.pm) pressure_faces(kick.pm, kp, y_old, y, v, true);
    if (y < 0.0) y += kp.length;
    else if (y >= kp.length) y -= kp.length;
    real yr = (real)y;
    if (!(yr < (real)kp.length)) yr = (real)0;
    if (!(yr >= (real)0)) yr = (real)0;
    r.y = yr;
    return r;
  }
  if (bc == MCPM_BC_SPECULAR) {
    // BC_specular_reflection (Main_Functions.cpp:6399-6525), slab axis.
    int guard = 0;
    while ((y < kp.wall_lo || y > kp.wall_hi) && guard < 20) {
      if (y > kp.wall_hi) y = 2.0 * kp.wall_hi - y;
      else y = 2.0 * kp.wall_lo - y;
      v = -v;
      ++guard;
    }
    if (guard >= 20) atomicOr(err, ERRBIT_BC);
    r.y = (real)y;
    r.u2 = (real)v;
    if (FACES && kick.pm) pressure_faces(kick.pm, kp, y_old, (double)r.y, (double)r.u2, false);
    return r;
  }
  if (bc == MCPM_BC_REINJECT) {
    // BC_evaporation (Main_Functions.cpp:5444-5617), slab axis: a particle
    // past a wall leaves the system and is put back at L/2 + U(-1,1) L/4 with
    // a Maxwellian velocity at T_lo (gas.T).  Draws: Philox(slot, 2k | 2k+1,
    // step, REINJECT).  The re-inserted position is inside, so k ends at 1.
    uint32_t k = 0;
    while ((y < kp.wall_lo || y > kp.wall_hi) && k < 4) {
      const uint32_t ctr3 = mcpm_ctr3(MCPM_TAG_REINJECT, step_hi, (uint32_t)(slot >> 32));
      mcpm_u32x4 a = mcpm_philox((uint32_t)slot, 2u * k, step_lo, ctr3, kp.key0, kp.key1);
      mcpm_u32x4 b = mcpm_philox((uint32_t)slot, 2u * k + 1u, step_lo, ctr3, kp.key0, kp.key1);
      y = 0.5 * kp.length + (2.0 * mcpm_u01(a.v[0]) - 1.0) * (0.25 * kp.length);
      double z0, z1, z2, z3;
      mcpm_normal2(a.v[1], a.v[2], &z0, &z1);
      mcpm_normal2(b.v[0], b.v[1], &z2, &z3);
      const double sd = __builtin_sqrt(kp.kb * kp.T_lo / kp.mass);
      r.u1 = (real)(sd * z0);
      r.u2 = (real)(sd * z1);
      r.u3 = (real)(sd * z2);
      if (FACES && kick.esc) atomicAdd(kick.esc, 1ull);
      ++k;
    }
    r.y = (real)y;
    if (FACES && kick.pm && k == 0)
      pressure_faces(kick.pm, kp, y_old, (double)r.y, (double)r.u2, false);
    return r;
  }
  // diffuse walls
  double dt_left = kp.dt;
  double v_prev = v_eff;
  uint32_t bounce = 0;
  while (y < kp.wall_lo || y > kp.wall_hi) {
    if (bounce >= 20) {  // "TOO MANY BC OPERATIONS", Main_Functions.cpp:4958-4965
      atomicOr(err, ERRBIT_BC);
      y = y < kp.wall_lo ? kp.wall_lo : kp.wall_hi;
      break;
    }
    bool upper = y > kp.wall_hi;
    double yw = upper ? kp.wall_hi : kp.wall_lo;
    double T = upper ? kp.T_hi : kp.T_lo;
    double Uw = upper ? kp.U_hi : kp.U_lo;
    mcpm_u32x4 w = mcpm_philox((uint32_t)slot, bounce, step_lo,
                               mcpm_ctr3(MCPM_TAG_WALL, step_hi, (uint32_t)(slot >> 32)),
                               kp.key0, kp.key1);
    double z0, z1;
    mcpm_normal2(w.v[0], w.v[1], &z0, &z1);
    double sd = __builtin_sqrt(kp.kb * T / kp.mass);
    double vn = __builtin_sqrt(2.0 * kp.kb * T / kp.mass) *
                __builtin_sqrt(-mcpm_log(mcpm_u01(w.v[2])));
    double u1 = sd * z0 + Uw;
    double u3 = sd * z1;
    double u2 = upper ? -vn : vn;
    double dtr = __builtin_fabs((y - yw) / v_prev);
    if (dtr > dt_left) dtr = dt_left;
    if (dt_left > 1e-15) y = yw + dtr * u2;
    else y = yw;
    dt_left = dt_left - dtr;
    v_prev = u2;
    r.u1 = (real)u1;
    r.u2 = (real)u2;
    r.u3 = (real)u3;
    ++bounce;
  }
  r.y = (real)y;
  if (FACES && kick.pm) pressure_faces(kick.pm, kp, y_old, (double)r.y, (double)r.u2, false);
  return r;
}

// Periodic wrap of one coordinate into [0, L) (stored type rounding included).
template <typename real>
__device__ __forceinline__ real wrap_store(double v, double L) {
  if (v < 0.0) v += L;
  else if (v >= L) v -= L;
  real vr = (real)v;
  if (!(vr < (real)L)) vr = (real)0;
  if (!(vr >= (real)0)) vr = (real)0;
  return vr;
}

// Derivatives of the bilinear shape functions (dNxi, dNeta, FEM.cpp:22-49)
// for the nodes 0 (lower left), 1 (lower right), 2 (upper left), 3.
__device__ __forceinline__ double dn_xi(int i, double xi, double eta) {
  (void)xi;
  return i == 0 ? eta - 1.0 : (i == 1 ? 1.0 - eta : (i == 2 ? -eta : eta));
}
__device__ __forceinline__ double dn_eta(int i, double xi, double eta) {
  (void)eta;
  return i == 0 ? xi - 1.0 : (i == 1 ? -xi : (i == 2 ? -xi + 1 : xi));
}

// 2D push: the FEM force kick of compute_SP_2D / compute_Vlasov_2D at the
// particle (FEM.cpp:568-582, 709-721: U -= grad(Q1 interpolant) coef / m dt),
// free flight in x and y, specular walls at 0 and L or periodic wrap per axis.
template <typename real>
__device__ __forceinline__ Rec2<real> push_one(Rec2<real> r, uint64_t slot, const KParams &kp,
                                               const Kick &kick,
                                               const double *__restrict__ drift, int mode,
                                               uint32_t step_lo, uint32_t step_hi,
                                               uint32_t *err, double f2 = 0.0,
                                               bool have_f2 = false) {
  (void)slot; (void)drift; (void)step_lo; (void)step_hi; (void)f2; (void)have_f2;
  if (mode & PUSH_KICK) {
    const int32_t cx = cell_of((double)r.x, kp.dx, kp.nx);
    const int32_t cy = cell_of((double)r.y, kp.dy, kp.ny);
    // node coordinates as make_cells builds them (myIO.cpp:1455-1479)
    const double x0 = (double)cx * kp.dx, x1 = (double)(cx + 1) * kp.dx;
    const double y0 = (double)cy * kp.dy, y1 = (double)(cy + 1) * kp.dy;
    const double dxc = x1 - x0, dyc = y1 - y0;
    const double xn0 = 0.5 * (x0 + x1) - dxc / 2.0;
    const double yn0 = 0.5 * (y0 + y1) - dyc / 2.0;
    const double xi = ((double)r.x - xn0) / dxc;
    const double eta = ((double)r.y - yn0) / dyc;
    const int32_t w = kp.nx + 1;
    const int32_t id[4] = {cy * w + cx, cy * w + cx + 1, (cy + 1) * w + cx, (cy + 1) * w + cx + 1};
    double F1 = 0.0, F2 = 0.0;
    for (int q = 0; q < 4; ++q) {
      F1 += dn_xi(q, xi, eta) * kick.xnode[id[q]];
      F2 += dn_eta(q, xi, eta) * kick.xnode[id[q]];
    }
    F1 = F1 / kp.mass * kick.coef / dxc;
    F2 = F2 / kp.mass * kick.coef / dyc;
    r.u1 = (real)((double)r.u1 - F1 * kp.dt);
    r.u2 = (real)((double)r.u2 - F2 * kp.dt);
  }
  if (!(mode & PUSH_MOVE)) return r;
  // Position_FP_ideal_gas for both axes (Fokker_Planck.cpp:948-969)
  double x = (double)r.x + (double)r.u1 * kp.dt;
  double y = (double)r.y + (double)r.u2 * kp.dt;
  const bool px = kp.boundary_x == MCPM_BC_PERIODIC, py = kp.boundary == MCPM_BC_PERIODIC;
  double u1 = (double)r.u1, u2 = (double)r.u2;
  if (!(px && py)) {
    if (!specular_plane(&x, &u1, !px, kp.length_x, &y, &u2, !py, kp.length))
      atomicOr(err, ERRBIT_BC);
    r.u1 = (real)u1;
    r.u2 = (real)u2;
  }
  r.x = px ? wrap_store<real>(x, kp.length_x) : (real)x;
  r.y = py ? wrap_store<real>(y, kp.length) : (real)y;
  return r;
}

template <typename real>
__device__ __forceinline__ bool in_domain(const Rec<real> &r, const KParams &kp) {
  double yy = (double)r.y;
  return yy >= 0.0 && yy <= kp.length;
}
template <typename real>
__device__ __forceinline__ bool in_domain(const Rec2<real> &r, const KParams &kp) {
  double xx = (double)r.x, yy = (double)r.y;
  return yy >= 0.0 && yy <= kp.length && xx >= 0.0 && xx <= kp.length_x;
}

// ---------------------------------------------------------------------------
// Stable counting sort, pass 1 and pass 3.  One workgroup of PUSH_WAVES
// wavefronts per chunk of consecutive particles; the chunk's cell histogram
// lives in LDS, shared by the wavefronts.
//   SCATTER=false: count particles per new cell -> table[chunk][cell]
//   SCATTER=true : running[cell] = start[cell] + table[chunk][cell] (exclusive
//                  over chunks) and each particle lands at running + rank, the
//                  rank taken among equal keys in the wavefront with ballots,
//                  the wavefronts of a tile in turn.
// Particles of a chunk are processed in index order, so equal keys keep their
// relative order: the sort is stable.
// The bins are the cells (kmode 0) or, for more cells than one LDS histogram
// holds, the low (kmode 1: key % kdiv) then the high digit (kmode 2:
// key / kdiv) of the cell key in two stable passes (least significant digit
// first), which together sort by the key.  okey, when given, receives each
// particle's cell key at its destination; ikey, when given, supplies the keys
// instead of the records (the second digit pass, whose input is in low-digit
// order: a batched replica cannot be told from a particle's index there).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t sort_bin(int32_t key, int kmode, int32_t kdiv) {
  return kmode == 0 ? key : (kmode == 1 ? key % kdiv : key / kdiv);
}

template <typename R, bool SCATTER>
__global__ __launch_bounds__(64 * PUSH_WAVES) void k_push(
    const R *__restrict__ in, R *__restrict__ out, int64_t n, int32_t chunk, KParams kp,
    Kick kick, const double *__restrict__ drift, int mode, uint32_t step_lo, uint32_t step_hi,
    uint32_t *__restrict__ table, const uint32_t *__restrict__ start, uint32_t *err, int nb,
    int kmode, int32_t kdiv, uint32_t *__restrict__ okey, const uint32_t *__restrict__ ikey,
    const uint32_t *__restrict__ iid, uint32_t *__restrict__ oid,
    typename R::real_type *__restrict__ oys, const int32_t *__restrict__ iw,
    int32_t *__restrict__ ow) {
  extern __shared__ uint32_t lds_cnt[];
  constexpr int TILE = 64 * PUSH_WAVES;  // particles of one tile, wavefront w the w-th 64
  const int nc = nb;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t b = blockIdx.x;
  const int64_t beg = b * (int64_t)chunk;
  int64_t end = beg + chunk;
  if (end > n) end = n;
  uint32_t *trow = table + (size_t)b * (size_t)nc;
  if (SCATTER) {
    for (int c = threadIdx.x; c < nc; c += TILE) lds_cnt[c] = start[c] + trow[c];
  } else {
    for (int c = threadIdx.x; c < nc; c += TILE) lds_cnt[c] = 0u;
  }
  __syncthreads();
  const uint64_t lt = lanemask_lt();
  const int kbits = key_bits(nb);
  // U tiles per batch; the next batch's records are loaded while this
  // batch is processed.  Loads complete in issue order, so the slab kick's
  // field values, the one other read of the common path, are read first.
  constexpr int U = 4;
  const bool pre = R::dims == 1 && (mode & PUSH_KICK) && kp.nrep == 0 && !ikey;
  const int64_t mine0 = (int64_t)wave * 64 + lane;  // this lane's offset inside a tile
  // (reads past the chunk's end repeat its last record, never used)
  R nx[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t p = beg + u * TILE + mine0;
    nx[u] = load_stream(in + (p < end ? p : end - 1));
  }
  for (int64_t t0 = beg; t0 < end; t0 += (int64_t)TILE * U) {
    R rr[U];
    double f2v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      rr[u] = nx[u];
      f2v[u] = 0.0;
    }
    if (pre) {  // uniform: one block of U reads, one wait
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int32_t c =
            t0 + u * TILE + mine0 < end ? cell_of((double)rr[u].y, kp.dy, kp.n_cells) : 0;
        f2v[u] = kick.F2[c];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t p = t0 + (int64_t)TILE * U + u * TILE + mine0;
      nx[u] = load_stream(in + (p < end ? p : end - 1));
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t p = t0 + u * TILE + mine0;
      const bool valid = p < end;
      int32_t key = 0, ckey = 0;
      uint32_t okv = 0;  // the key carried to the destination (plane: with its quarter)
      R r = rr[u];
      if (valid) {
        if (!(r.u1 == r.u1) || !(r.u2 == r.u2) || !(r.u3 == r.u3)) atomicOr(err, ERRBIT_NAN);
        if (ikey) {
          okv = ikey[p];
          ckey = (int32_t)(okv & KEY_CELL_MASK);
        } else if (kp.nrep == 0) {
          if ((mode & PUSH_CHECK) && !in_domain(r, kp)) atomicOr(err, ERRBIT_DOMAIN);
          r = push_one(r, (uint64_t)p, kp, kick, drift, mode, step_lo, step_hi, err, f2v[u],
                       pre);
          ckey = cell_key(r, kp);
          okv = (uint32_t)ckey | key_quarter_bits(r, ckey, kp);
        } else {
          // a batched replica: its own geometry, Philox key, particle slots
          // and cells counted from its first
          const int32_t q = rep_of_particle(kp, p);
          const KParams &kq = kp.rkp[q];
          const int32_t off = kp.rcell_off[q];
          Kick kk = kick;
          kk.F2 = kick.F2 + off;
          if (kk.pm) kk.pm = kick.pm + off;
          if ((mode & PUSH_CHECK) && !in_domain(r, kq)) atomicOr(err, ERRBIT_DOMAIN);
          r = push_one(r, (uint64_t)(p - kp.rpart_off[q]), kq, kk, drift, mode, step_lo,
                       step_hi, err);
          ckey = off + cell_key(r, kq);
          okv = (uint32_t)ckey;
        }
        key = sort_bin(ckey, kmode, kdiv);
      }
      if (SCATTER) {
        // the tile's wavefronts take turns, in particle order: equal keys
        // land in index order (the sort is stable)
        for (int w = 0; w < PUSH_WAVES; ++w) {
          if (wave == w) {
            // rank among the lanes below with the same key, and how many
            // lanes share it: one ballot per key bit, no memory access
            const uint64_t eq = match_key((uint32_t)key, valid, kbits);
            const uint32_t rk = (uint32_t)__popcll(eq & lt), tot = (uint32_t)__popcll(eq);
            // one counter read per lane (lanes of one key read the same
            // word), one store per lane; the last lane of each key advances
            // its counter (distinct keys: no two lanes write one word)
            uint32_t base = 0;
            if (valid) base = lds_cnt[key];
            __builtin_amdgcn_wave_barrier();
            if (valid) {
              const uint32_t dst = base + rk;
              if (dst < (uint32_t)n) {
                out[dst] = r;
                if (okey) okey[dst] = okv;
                if (oid) oid[dst] = iid[p];
                if (oys) oys[dst] = r.y;
                if (ow) ow[dst] = iw[p];
              } else {
                atomicOr(err, ERRBIT_ORDER);  // the two passes disagreed on a key
              }
              if (rk + 1 == tot) lds_cnt[key] = base + tot;
            }
            __builtin_amdgcn_wave_barrier();
          }
          if constexpr (PUSH_WAVES > 1) __syncthreads();
        }
      } else {
        // counts only: the order of the additions does not matter.  The
        // first two distinct keys are counted with one ballot each (on
        // coarse cells they hold most of the wavefront); every lane left
        // adds itself, one LDS atomic per lane instead of one ballot round
        // per distinct key (on fine cells a wavefront's particles spread
        // over tens of cells).
        uint64_t active = __ballot(valid);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
          if (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const int lkey = __builtin_amdgcn_readlane(key, leader);
            const uint64_t m = __ballot(valid && key == lkey);
            if (lane == leader) atomicAdd(&lds_cnt[lkey], (uint32_t)__popcll(m));
            active &= ~m;
          }
        }
        if ((active >> lane) & 1ull) atomicAdd(&lds_cnt[key], 1u);
      }
    }
  }
  if (!SCATTER) {
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += TILE) trow[c] = lds_cnt[c];
  }
}

// ---------------------------------------------------------------------------
// The same stable counting sort when particles move only a few cells per
// step (the slab and the batched slabs): chunk b of consecutive particles
// (WCH of them, never two replicas in one chunk) was sorted by cell in the
// previous step, so its particles' new cells lie within WDELTA cells of the
// old cells of its first and last particle.  Its histogram is kept only over
// that window of at most WMAX cells, lo = old cell of the first particle -
// WDELTA (in the replica's own, unwrapped cell numbers: a periodic slab's
// cells -1, -2, ... are its last ones), so the table is chunks x WMAX instead
// of chunks x cells and a chunk can be small: one wavefront per chunk of
// 1024 particles instead of one per several thousand.  A particle outside
// its chunk's window (a long flight, a wide chunk, input out of cell order)
// sets *wflag and ends its wavefront's count pass; the scatter pass then
// stores nothing and the caller sorts again by the full histogram, from the
// same input.
//   SCATTER=false: window and histogram of each chunk -> wlo, wlen, wtab
//   SCATTER=true : wtab holds each window cell's offset among the chunks
//                  (k_wscan), start[] the cells' starts; each particle lands
//                  at start + offset + its stable rank in the chunk
// ---------------------------------------------------------------------------
// Window offset of a particle of new local cell key in a window [lo, lo + len)
// of a slab of nc cells, -1 when outside.
__device__ __forceinline__ int32_t window_offset(int32_t key, int32_t lo, int32_t len, int32_t nc,
                                                 bool periodic) {
  int32_t d = key - lo;
  if (periodic) {
    if (d < 0) d += nc;
    else if (d >= len) d -= nc;
  }
  return (d >= 0 && d < len) ? d : -1;
}

template <typename R, bool SCATTER, int PK>
__global__ __launch_bounds__(256) void k_wpush(
    const R *__restrict__ in, R *__restrict__ out, KParams kp, Kick kick,
    const double *__restrict__ drift, int mode, uint32_t step_lo, uint32_t step_hi,
    const uint32_t *__restrict__ wbeg, const int32_t *__restrict__ wrep,
    const uint32_t *__restrict__ nchp, int32_t *__restrict__ wlo, int32_t *__restrict__ wlen,
    uint32_t *__restrict__ wtab, const uint32_t *__restrict__ start, uint32_t *__restrict__ err,
    uint32_t *__restrict__ wflag, const uint32_t *__restrict__ iid, uint32_t *__restrict__ oid,
    int wdelta, typename R::real_type *__restrict__ oys, const int32_t *__restrict__ iw,
    int32_t *__restrict__ ow, const StepDev *__restrict__ sd) {
  __shared__ uint32_t lcnt[4][WMAX];
  if (sd_stopped(sd)) return;
  if (sd) {
    step_lo = sd->lo;
    step_hi = sd->hi;
    wdelta = (int)sd->wdelta;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nch = (int)*nchp;
  const int b = blockIdx.x * 4 + wave;  // one chunk per wavefront, no workgroup barrier
  if (b >= nch) return;
  // a particle left its window in the count pass: this sort is redone anyway
  if (SCATTER && *(volatile const uint32_t *)wflag != 0u) return;
  uint32_t *cnt = lcnt[wave];
  const uint32_t beg = wbeg[b], end = wbeg[b + 1], ntot = wbeg[nch];
  // the first tiles' loads go out with the window's (every chunk holds at
  // least one particle: a block without particles has no pieces)
  constexpr int U = 4;  // tiles of 64 in flight
  R nx[U];
  if (beg < end) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t p = beg + u * 64 + lane;
      nx[u] = in[p < end ? p : end - 1];
    }
  }
  // the chunk's replica: its constants, force, cell offset and slot base
  const int32_t q = kp.nrep ? wrep[b] : 0;
  const KParams &kq = kp.nrep ? kp.rkp[q] : kp;
  const int32_t off = kp.nrep ? kp.rcell_off[q] : 0;
  const int64_t slot0 = kp.nrep ? kp.rpart_off[q] : 0;
  Kick kk = kick;
  if (kp.nrep) kk.F2 = kick.F2 + off;
  const int32_t nc = kq.n_cells;
  const bool periodic = kq.boundary == MCPM_BC_PERIODIC;
  int32_t lo, len;
  if (SCATTER) {
    lo = wlo[b];
    len = wlen[b];
  } else {
    const int32_t kf = beg < end ? cell_key(in[beg], kq) : 0;
    const int32_t kl = beg < end ? cell_key(in[end - 1], kq) : 0;
    lo = kf - wdelta;
    len = kl - kf + 2 * wdelta + 1;
    if (len > WMAX || len < 1 || (periodic && len > nc)) {
      if (lane == 0) atomicOr(wflag, 1u);
      len = 0;
    }
    if (lane == 0) {
      wlo[b] = lo;
      wlen[b] = len;
    }
  }
  if (len == 0) return;
  // every particle counted once: otherwise a chunk was missed (input out of
  // cell order) and the offsets are not a permutation
  if (SCATTER && b == 0 && lane == 0 && start[kp.n_cells] != ntot) atomicOr(wflag, 1u);
  for (int w = lane; w < len; w += 64) {
    if (SCATTER) {
      int32_t c = lo + w;
      c = ((c % nc) + nc) % nc;  // cells beyond a wall's end are empty windows
      cnt[w] = start[off + c] + wtab[(size_t)b * WMAX + w];
    } else {
      cnt[w] = 0u;
    }
  }
  __builtin_amdgcn_wave_barrier();
  const uint64_t lt = lanemask_lt();
  for (uint32_t t0 = beg; t0 < end; t0 += 64 * U) {
    R rr[U];
#pragma unroll
    for (int u = 0; u < U; ++u) rr[u] = nx[u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t p = t0 + 64 * U + u * 64 + lane;
      nx[u] = in[p < end ? p : end - 1];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t p = t0 + u * 64 + lane;
      bool valid = p < end;
      int32_t d = 0;
      R r = rr[u];
      if (valid) {
        if (!(r.u1 == r.u1) || !(r.u2 == r.u2) || !(r.u3 == r.u3)) atomicOr(err, ERRBIT_NAN);
        if ((mode & PUSH_CHECK) && !in_domain(r, kq)) atomicOr(err, ERRBIT_DOMAIN);
        // PK: the push of any slab (0), or of a periodic slab without drift
        // and face sums (1: far fewer registers, more waves per SIMD)
        const uint64_t slot = (uint64_t)((int64_t)p - slot0);
        if constexpr (PK == 1)
          r = push_one<typename R::real_type, MCPM_BC_PERIODIC, false, false>(
              r, slot, kq, kk, drift, mode, step_lo, step_hi, err);
        else
          r = push_one(r, slot, kq, kk, drift, mode, step_lo, step_hi, err);
        d = window_offset(cell_key(r, kq), lo, len, nc, periodic);
        if (d < 0) valid = false;
      }
      // a particle outside the window: the sort is redone by the full
      // histogram, so the count pass stops here (one flag store per wavefront)
      if (!SCATTER && __ballot(p < end && d < 0)) {
        if (lane == 0) atomicOr(wflag, 1u);
        return;
      }
      if (SCATTER) {
        // stable rank among the lanes with the same window cell, as k_push
        const uint64_t eq = match_key((uint32_t)d, valid, key_bits(WMAX));
        const uint32_t rk = (uint32_t)__popcll(eq & lt), tot = (uint32_t)__popcll(eq);
        uint32_t base = 0;
        if (valid) base = cnt[d];
        __builtin_amdgcn_wave_barrier();
        if (valid) {
          const uint32_t dst = base + rk;
          if (dst < ntot) {
            out[dst] = r;
            if (oid) oid[dst] = iid[p];
            if (oys) oys[dst] = r.y;
            if (ow) ow[dst] = iw[p];
          } else {
            atomicOr(wflag, 1u);  // counts out of step with the particles
          }
          if (rk + 1 == tot) cnt[d] = base + tot;
        }
        __builtin_amdgcn_wave_barrier();
      } else {
        uint64_t active = __ballot(valid);
#pragma unroll
        for (int it = 0; it < 2; ++it) {
          if (active) {
            const int leader = __ffsll((unsigned long long)active) - 1;
            const int ld = __builtin_amdgcn_readlane(d, leader);
            const uint64_t m = __ballot(valid && d == ld);
            if (lane == leader) atomicAdd(&cnt[ld], (uint32_t)__popcll(m));
            active &= ~m;
          }
        }
        if ((active >> lane) & 1ull) atomicAdd(&cnt[d], 1u);
      }
    }
  }
  if (!SCATTER) {
    __builtin_amdgcn_wave_barrier();
    for (int w = lane; w < len; w += 64) wtab[(size_t)b * WMAX + w] = cnt[w];
  }
}

// The chunks of a windowed sort, from the cell bounds of the previous sort:
// block j of WSPAN consecutive cells of one replica (cells [bcell[j],
// bcell[j + 1])) holds the particles [start[bcell[j]], start[bcell[j + 1]]),
// cut into pieces of at most WCH.  A chunk then spans at most WSPAN old cells
// however sparse the gas.  k_wblocks counts each block's pieces, k_excl_scan
// numbers them, k_wfill writes the pieces' bounds and replicas, the total
// (*nch, wbeg[*nch] = n) and the first piece of each replica.
__global__ void k_wblocks(const int32_t *__restrict__ bcell, int nblk,
                          const uint32_t *__restrict__ start, uint32_t *__restrict__ pieces,
                          const StepDev *__restrict__ sd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nblk || sd_stopped(sd)) return;
  const uint32_t a = start[bcell[j]], z = start[bcell[j + 1]];
  pieces[j] = (z - a + WCH - 1) / WCH;
}
// One workgroup per block (blockIdx.x < nblk), its threads over the pieces;
// the workgroups past nblk write the replicas' first pieces.
__global__ void k_wfill(const int32_t *__restrict__ bcell, const int32_t *__restrict__ brep,
                        int nblk, const uint32_t *__restrict__ start,
                        const uint32_t *__restrict__ pbase, const int32_t *__restrict__ rblk,
                        int nrep1, uint32_t n, uint32_t *__restrict__ wbeg,
                        int32_t *__restrict__ wrep, uint32_t *__restrict__ nchp,
                        uint32_t *__restrict__ rchunk, const StepDev *__restrict__ sd) {
  if (sd_stopped(sd)) return;
  const int j = blockIdx.x;
  if (j < nblk) {
    const uint32_t a = start[bcell[j]], z = start[bcell[j + 1]];
    const uint32_t k0 = pbase[j], np = pbase[j + 1] - k0;
    for (uint32_t t = threadIdx.x; t < np; t += blockDim.x) {
      wbeg[k0 + t] = a + t * WCH;
      wrep[k0 + t] = brep[j];
    }
    (void)z;
    if (j == 0 && threadIdx.x == 0) {
      *nchp = pbase[nblk];
      wbeg[pbase[nblk]] = n;
    }
  } else {
    const int r = (j - nblk) * blockDim.x + threadIdx.x;
    if (r < nrep1) rchunk[r] = pbase[rblk[r]];
  }
}

// First i in [a, z) with pred(i) (pred false then true along i), or z: the
// wavefront tests 64 evenly spread indices per round, so a range of N takes
// about log64(N) rounds of one load per lane.
template <typename Pred>
__device__ __forceinline__ int wave_lower_bound(int a, int z, Pred pred) {
  const int lane = (int)lane_id();
  while (z > a) {
    const int step = (z - a + 63) / 64;
    const int i = a + lane * step;
    const bool t = i >= z || pred(i);
    const uint64_t m = __ballot(t);
    if (m == 0) {
      a = a + 63 * step + 1;
      continue;
    }
    const int f = __ffsll((unsigned long long)m) - 1;
    if (f == 0) return a;
    const int hi = a + f * step;
    a = a + (f - 1) * step + 1;
    z = hi < z ? hi : z;
  }
  return z;
}

// The same from a lower bound a that is usually close: the 64 indices from a
// in one load per lane, the search beyond them only when none holds.
template <typename Pred>
__device__ __forceinline__ int wave_first_from(int a, int z, Pred pred) {
  const int i = a + (int)lane_id();
  const uint64_t m = __ballot(i < z && pred(i));
  if (m) return a + __ffsll((unsigned long long)m) - 1;
  return a + 64 < z ? wave_lower_bound(a + 64, z, pred) : z;
}

// Per cell (cpw consecutive cells per wavefront): the chunks whose windows hold it, in
// chunk order (a contiguous run for each of the cell's unwrapped numbers
// k - nc, k, k + nc of a periodic slab, in that order), their counts
// replaced by the running offset; the cell's count, density and contact
// value.  wlo and wlo + wlen ascend along a replica's chunks when the input
// was in cell order; out of order the counts are only wrong, never out of
// bounds, and the chunks' own checks have raised *wflag.
__global__ __launch_bounds__(256) void k_wscan(const int32_t *__restrict__ wlo,
                                               const int32_t *__restrict__ wlen,
                                               uint32_t *__restrict__ wtab,
                                               const uint32_t *__restrict__ nchp,
                                               const uint32_t *__restrict__ rchunk, KParams kp,
                                               uint32_t *__restrict__ cnt,
                                               double *__restrict__ nden,
                                               double *__restrict__ yc, int cpw,
                                               const StepDev *__restrict__ sd) {
  if (sd_stopped(sd)) return;
  const int lane = threadIdx.x & 63;
  const int k0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * cpw;
  const int k1 = k0 + cpw < kp.n_cells ? k0 + cpw : kp.n_cells;
  // the previous cell's replica and chunk bounds (image 0): both bounds only
  // move forward from one cell to the next of a replica
  int32_t pq = -1;
  int pblo = 0, pbhi = 0;
  for (int k = k0; k < k1; ++k) {
    const int32_t q = kp.nrep ? kp.cell_rep[k] : 0;
    const KParams &kq = kp.nrep ? kp.rkp[q] : kp;
    const int32_t off = kp.nrep ? kp.rcell_off[q] : 0;
    const int32_t nc = kq.n_cells;
    const int32_t kl = k - off;
    const int b0 = kp.nrep ? (int)rchunk[q] : 0, b1 = kp.nrep ? (int)rchunk[q + 1] : (int)*nchp;
    const bool periodic = kq.boundary == MCPM_BC_PERIODIC;
    uint32_t run = 0;
    for (int im = periodic ? -1 : 0; im <= (periodic ? 1 : 0); ++im) {
      // every window lies in [-WDELTA_MAX, nc + WDELTA_MAX): the images
      // below 0 and past nc - 1 hold chunks only for cells near the ends
      if ((im < 0 && kl < nc - WMAX) || (im > 0 && kl >= WMAX)) continue;
      const int32_t u = kl + im * nc;
      // first chunk whose window ends past u, first whose window starts past u
      auto ends_past = [&](int m) { return wlo[m] + wlen[m] > u; };
      auto starts_past = [&](int m) { return wlo[m] > u; };
      int blo, bhi;
      if (im == 0 && q == pq) {
        blo = wave_first_from(pblo, b1, ends_past);
        bhi = wave_first_from(pbhi > blo ? pbhi : blo, b1, starts_past);
      } else {
        blo = wave_lower_bound(b0, b1, ends_past);
        bhi = wave_lower_bound(blo, b1, starts_past);
      }
      if (im == 0) {
        pq = q;
        pblo = blo;
        pbhi = bhi;
      }
      for (int bb = blo; bb < bhi; bb += 64) {
        const int b = bb + lane;
        uint32_t v = 0;
        int32_t w = -1;
        if (b < bhi) {
          w = u - wlo[b];
          if (w >= 0 && w < wlen[b]) v = wtab[(size_t)b * WMAX + w];
          else w = -1;
        }
        uint32_t x = v;  // inclusive scan over the wavefront
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o);
          if (lane >= o) x += y;
        }
        if (w >= 0) wtab[(size_t)b * WMAX + w] = run + x - v;
        run += __shfl(x, 63);
      }
    }
    if (lane == 0) {
      cnt[k] = run;
      const double nd = cell_weight(kq, k) * kq.Fn * (double)run / kq.volume;  // Main_Functions.cpp:2601
      nden[k] = nd;
      yc[k] = contact_y(nd, kq.b, kq.y_model);
    }
  }
}

// Column scans of table[chunk][cell] in segments of SEG chunks.
__global__ void k_scan_a(const uint32_t *__restrict__ table, uint32_t *__restrict__ partial,
                         int nchunks, int nc, int seg) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int s = blockIdx.y;
  if (c >= nc) return;
  int b0 = s * seg, b1 = min(nchunks, b0 + seg);
  uint32_t sum = 0;
  for (int b = b0; b < b1; ++b) sum += table[(size_t)b * nc + c];
  partial[(size_t)s * nc + c] = sum;
}

// Per bin (one thread each): exclusive scan over the segments of chunks, the
// bin's count and, for cells, the density and contact value; the starts are
// then one exclusive scan of the counts (k_excl_scan).
__global__ void k_scan_bins(uint32_t *__restrict__ partial, int nseg, int nc,
                            uint32_t *__restrict__ cnt, double *__restrict__ nden,
                            double *__restrict__ yc, KParams kp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  uint32_t run = 0;
  for (int s = 0; s < nseg; ++s) {
    const uint32_t v = partial[(size_t)s * nc + c];
    partial[(size_t)s * nc + c] = run;
    run += v;
  }
  cnt[c] = run;
  if (nden) {
    const KParams &kq = kp.nrep ? kp.rkp[kp.cell_rep[c]] : kp;
    double nd = cell_weight(kq, c) * kq.Fn * (double)run / kq.volume;  // Main_Functions.cpp:2601
    nden[c] = nd;
    yc[c] = contact_y(nd, kq.b, kq.y_model);
  }
}

// Cells of the two-digit sort: start[c] = first particle of cell c from the
// sorted keys (every c in [0, nc] written once when the keys ascend, which the
// sort guarantees), then counts, densities and contact values as k_scan_bins.
// start is zeroed first and a descent raises ERRBIT_ORDER, so even keys out
// of order leave every start in [0, n] and every count within the array.
__global__ void k_cell_bounds(const uint32_t *__restrict__ key, int64_t n, int nc,
                              uint32_t *__restrict__ start, uint32_t *__restrict__ err) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t prev = i > 0 ? (int64_t)(key[i - 1] & KEY_CELL_MASK) : -1;
    const int64_t cur = i < n ? (int64_t)(key[i] & KEY_CELL_MASK) : (int64_t)nc;
    if (cur < prev || cur > nc) atomicOr(err, ERRBIT_ORDER);
    const int64_t hi = cur < nc ? cur : (int64_t)nc;
    for (int64_t c = prev + 1; c <= hi; ++c) start[c] = (uint32_t)i;
  }
}

__global__ void k_cell_fill(const uint32_t *__restrict__ start, int nc,
                            uint32_t *__restrict__ cnt, double *__restrict__ nden,
                            double *__restrict__ yc, KParams kp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint32_t run = start[c + 1] >= start[c] ? start[c + 1] - start[c] : 0u;
  cnt[c] = run;
  const KParams &kq = kp.nrep ? kp.rkp[kp.cell_rep[c]] : kp;
  double nd = cell_weight(kq, c) * kq.Fn * (double)run / kq.volume;  // Main_Functions.cpp:2601
  nden[c] = nd;
  yc[c] = contact_y(nd, kq.b, kq.y_model);
}

__global__ void k_scan_c(uint32_t *__restrict__ table, const uint32_t *__restrict__ partial,
                         int nchunks, int nc, int seg) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int s = blockIdx.y;
  if (c >= nc) return;
  int b0 = s * seg, b1 = min(nchunks, b0 + seg);
  uint32_t run = partial[(size_t)s * nc + c];
  for (int b = b0; b < b1; ++b) {
    uint32_t v = table[(size_t)b * nc + c];
    table[(size_t)b * nc + c] = run;
    run += v;
  }
}

// ---------------------------------------------------------------------------
// ESMC collisions (Collision_ESMC_new3, collisions.cpp:583-720)
// ---------------------------------------------------------------------------
// Candidate count per cell, collisions.cpp:617-629:
//   m_c = floor(0.5 * omega_max_c * N_c + U)
// DSMC (collisions.cpp:2087-2090): M_c = pi sigma^2 N_c n_c crm dt / 2 for
// cells with two or more particles, crm kept in the omax array; CBA_VHS
// multiplies by the contact value Y(n_c) (collisions.cpp:2171-2172).
__global__ void k_mcand(const uint32_t *__restrict__ cnt, const double *__restrict__ omax,
                        const double *__restrict__ nden, const double *__restrict__ yc,
                        uint32_t *__restrict__ mcand, int nc, KParams kp, int model, int variant,
                        uint32_t step_lo, uint32_t step_hi, const StepDev *__restrict__ sd) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc || sd_stopped(sd)) return;
  if (sd) {
    step_lo = sd->lo;
    step_hi = sd->hi;
  }
  // batched replicas draw with their own key and local cell numbers
  const int32_t rq = kp.nrep ? kp.cell_rep[c] : 0;
  const KParams &kq = kp.nrep ? kp.rkp[rq] : kp;
  const uint32_t cl = (uint32_t)(kp.nrep ? c - kp.rcell_off[rq] : c);
  mcpm_u32x4 w = mcpm_philox(cl, 0u, step_lo, mcpm_ctr3(MCPM_TAG_MCAND, step_hi, 0u),
                             kq.key0, kq.key1);
  double m;
  if (model == COLL_DSMC) {
    if (cnt[c] < 2) {
      mcand[c] = 0u;
      return;
    }
    double Mc = 3.14159265358979323846 * kp.sigma * kp.sigma * ((double)cnt[c] * nden[c]) *
                omax[c] * kp.dt / 2.0;
    if (variant == MCPM_MODEL_CBA_VHS) Mc = Mc * yc[c];
    m = mcpm_floor(Mc + mcpm_u01(w.v[0]));
  } else {
    // Collision_ESMC_new4 draws only in cells with two or more particles
    // (collisions.cpp:763)
    if (model == COLL_ESMC4 && cnt[c] < 2) {
      mcand[c] = 0u;
      return;
    }
    m = mcpm_floor(0.5 * omax[c] * (1.0 * (double)cnt[c]) + mcpm_u01(w.v[0]));
  }
  mcand[c] = m > 0.0 ? (uint32_t)m : 0u;
}

// Exclusive scan of a per-cell u32 array (one workgroup); total in out[nc].
__global__ __launch_bounds__(1024) void k_excl_scan(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, int nc,
                                                    const StepDev *__restrict__ sd) {
  __shared__ uint32_t sh[1024];
  if (sd_stopped(sd)) return;
  const int tid = threadIdx.x;
  const int per = (nc + 1023) / 1024;
  const int c0 = tid * per;
  uint32_t tot = 0;
  for (int k = 0; k < per; ++k) {
    int c = c0 + k;
    if (c < nc) tot += in[c];
  }
  sh[tid] = tot;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    uint32_t v = (tid >= off) ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += v;
    __syncthreads();
  }
  uint32_t run = (tid == 0) ? 0u : sh[tid - 1];
  for (int k = 0; k < per; ++k) {
    int c = c0 + k;
    if (c < nc) {
      out[c] = run;
      run += in[c];
    }
  }
  if (tid == 1023) out[nc] = sh[1023];
}

// Exclusive scan of long arrays over many workgroups: SCAN_TILE elements per
// workgroup, 4 consecutive ones per thread.  k_tile_sums sums each tile,
// k_excl_scan scans the tile sums into tile offsets, k_tile_scan scans each
// tile from its offset.  Integer sums: the same result as one serial scan.
__global__ __launch_bounds__(1024) void k_tile_sums(const uint32_t *__restrict__ in, int n,
                                                    uint32_t *__restrict__ tsum,
                                                    const StepDev *__restrict__ sd) {
  __shared__ uint32_t sw[16];
  if (sd_stopped(sd)) return;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + 4 * (int64_t)threadIdx.x;
  uint32_t v = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (base + q < n) v += in[base + q];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (lane_id() == 0) sw[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; ++w) t += sw[w];
    tsum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void k_tile_scan(const uint32_t *__restrict__ in,
                                                    uint32_t *__restrict__ out, int n,
                                                    const uint32_t *__restrict__ toff, int nt,
                                                    const StepDev *__restrict__ sd) {
  __shared__ uint32_t sw[16];
  if (sd_stopped(sd)) return;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + 4 * (int64_t)threadIdx.x;
  uint32_t e[4], v = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    e[q] = base + q < n ? in[base + q] : 0u;
    v += e[q];
  }
  // exclusive prefix of the thread sums over the workgroup
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if ((int)lane_id() >= o) x += y;
  }
  if (lane_id() == 63) sw[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t run = toff[blockIdx.x] + x - v;
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) run += sw[w];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (base + q < n) out[base + q] = run;
    run += e[q];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = toff[nt];
}

// The candidate draw: i in cell c, isotropic direction, the partner cell J
// at y_i + sigma dir_y (collisions.cpp:648-671; 1D rule of new4,
// collisions.cpp:784-792), j in J, j != i (collisions.cpp:675-679).

// clamp (Collision_ESMC_new4, collisions.cpp:784-792): a partner point
// beyond a wall takes the wall's cell instead of leaving the candidate
// without a partner.
// Particle i's record as the partner-cell rule reads it: on the slab only y,
// from the sorted positions ys when the sort wrote them.
template <typename R>
__device__ __forceinline__ R position_of(const R *P, const typename R::real_type *ys,
                                         uint32_t i) {
  if constexpr (R::dims == 1) {
    if (ys) {
      R r;
      r.y = ys[i];
      r.u1 = r.u2 = r.u3 = 0;
      return r;
    }
  }
  return P[i];
}

// DSMC pair draw (collisions.cpp:2093-2098): two distinct particles of cell c.
// The draws of candidate k (global index) are keyed by (k, c): the commit
// finds both in the operation record, whatever the particles' positions.
__device__ __forceinline__ void draw_dsmc_pair(uint32_t c, uint32_t k, const uint32_t *cnt,
                                               const uint32_t *start, const KParams &kp,
                                               uint32_t step_lo, uint32_t step_hi, uint32_t *oi,
                                               uint32_t *oj) {
  mcpm_u32x4 w = mcpm_philox(k, c, step_lo, mcpm_ctr3(MCPM_TAG_DSMC, step_hi, 0u), kp.key0,
                             kp.key1);
  uint32_t Nc = cnt[c];
  uint32_t li = mcpm_uint_below(w.v[0], Nc);
  uint32_t lj = mcpm_uint_below(w.v[1], Nc - 1u);
  if (lj >= li) lj += 1u;
  *oi = start[c] + li;
  *oj = start[c] + lj;
}

// First candidate of region x: the smallest k with op_region(k, M) >= x.
__device__ __forceinline__ uint32_t region_start(uint32_t x, uint32_t M) {
  if (x >= NREG) return 0xFFFFFFFFu;
  return (uint32_t)(((uint64_t)x * M + NREG - 1) / NREG);
}
// op_region(k, M) for k >= region_start(x0, M), walking up from region x0
// (no division by M; a chunk usually stays in x0 or x0 + 1).
__device__ __forceinline__ uint32_t region_from(uint32_t k, uint32_t x0, uint32_t M) {
  uint32_t x = x0;
  while (k >= region_start(x + 1u, M)) ++x;
  return x;
}

// Priority of candidate k: ESMC the keyed bijection of the index (new3's
// random interleave of all cells' candidates), DSMC the index itself (the
// reference's in-cell order).
__device__ __forceinline__ uint32_t cand_priority(uint32_t k, uint32_t key, int esmc) {
  return esmc ? mcpm_priority(k, key) : k;
}
// The same for candidate k of replica rq of a batched context: the replica's
// own candidate number and key, as if it ran alone.  Priorities of different
// replicas may coincide; their operations never share a particle.
__device__ __forceinline__ uint32_t cand_priority_rep(const KParams &kp, int32_t rq, uint32_t k,
                                                      int esmc) {
  return cand_priority(k - kp.rcand_off[rq], kp.rprio[rq], esmc);
}

// Size of every (band, region) list per chunk of candidates.
__global__ __launch_bounds__(256) void k_cand_count(uint32_t M, uint32_t prio_key, int model,
                                                    uint32_t band_mul, int nbands,
                                                    uint32_t *__restrict__ table,
                                                    uint32_t tstride, KParams kp,
                                                    const uint32_t *__restrict__ cstart,
                                                    int nc, uint32_t *__restrict__ gcells,
                                                    const StepDev *__restrict__ sd) {
  __shared__ uint32_t scount[NLIST];
  __shared__ uint32_t scell;
  if (sd) {
    // a captured step: the candidates and key of this step, chunks past them idle
    if (sd_stopped(sd)) return;
    M = cstart[nc];
    prio_key = sd->pk;
    if (blockIdx.x * GEN_CHUNK >= M) return;
  }
  const int nl = nbands * NREG;
  if ((int)threadIdx.x < NLIST) scount[threadIdx.x] = 0;
  __syncthreads();  // the counts are zero before any thread adds to them
  const uint32_t lo = blockIdx.x * GEN_CHUNK;
  // the chunk's first and last cell, for the draw: the last c with
  // cstart[c] <= lo (resp. hi - 1), searched by wavefront 0 over cstart[0..nc]
  // (cstart[0] = 0 <= lo < M = cstart[nc])
  if (threadIdx.x < 64) {
    const uint32_t last = (lo + GEN_CHUNK < M ? lo + GEN_CHUNK : M) - 1u;
    const int a = wave_lower_bound(1, nc + 1, [&](int cc) { return cstart[cc] > lo; }) - 1;
    const int b = wave_lower_bound(a + 1, nc + 1, [&](int cc) { return cstart[cc] > last; }) - 1;
    if (threadIdx.x == 0) {
      scell = (uint32_t)a;
      gcells[2 * blockIdx.x] = (uint32_t)a;
      gcells[2 * blockIdx.x + 1] = (uint32_t)b;
    }
  }
  // only new4's ranks and the replicas' priorities need the first cell here
  if (model == COLL_ESMC4 || kp.nrep) __syncthreads();
  const int esmc = model != COLL_DSMC;
  const uint32_t x0 = op_region(lo, M);
  uint32_t c = model == COLL_ESMC4 ? scell : 0u;
  // the replica of this thread's candidates: candidates run in cell order,
  // replica after replica, so it walks forward from the replica of the
  // chunk's first cell (rep_of_cand's answer, without its search)
  int32_t rq = kp.nrep ? kp.cell_rep[scell] : 0;
  for (uint32_t b0 = 0; b0 < GEN_CHUNK; b0 += blockDim.x) {
    const uint32_t k = lo + b0 + threadIdx.x;
    if (k >= M) break;
    uint32_t band;
    if (model == COLL_ESMC4) {
      while (cstart[c + 1] <= k) ++c;
      band = band_new4(rank_new4(c, k - cstart[c], cstart, M), M, nbands);
    } else {
      uint32_t pr;
      if (kp.nrep) {
        while (kp.rcand_off[rq + 1] <= k) ++rq;
        pr = cand_priority_rep(kp, rq, k, esmc);
      } else {
        pr = cand_priority(k, prio_key, esmc);
      }
      band = op_band(pr, band_mul, nbands);
    }
    const uint32_t q = band * NREG + region_from(k, x0, M);
    atomicAdd(&scount[q], 1u);
  }
  __syncthreads();
  if ((int)threadIdx.x < nl) table[(size_t)threadIdx.x * tstride + blockIdx.x] = scount[threadIdx.x];
}

// Exclusive scan of list q's chunk counts (one workgroup per list, the
// list's counts contiguous: table[q tstride + chunk], eight per thread per
// pass); the list's total goes to the band counters, and the last list to
// finish writes where every (band, region) list starts.
__global__ __launch_bounds__(1024) void k_cand_scan(uint32_t *__restrict__ table,
                                                    uint32_t tstride, uint32_t nchunks, int nl,
                                                    uint32_t *__restrict__ counters,
                                                    const uint32_t *__restrict__ Mp,
                                                    const StepDev *__restrict__ sd) {
  constexpr int PER = 8;
  __shared__ uint32_t sw[16];
  __shared__ uint32_t s_last;
  const int q = blockIdx.x;
  if (q >= nl || sd_stopped(sd)) return;
  if (Mp) nchunks = (*Mp + GEN_CHUNK - 1) / GEN_CHUNK;
  uint32_t *row = table + (size_t)q * tstride;
  const int lane = (int)lane_id(), wave = (int)(threadIdx.x >> 6);
  const int nw = (int)(blockDim.x >> 6);
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < nchunks; b0 += blockDim.x * PER) {
    const uint32_t e0 = b0 + threadIdx.x * PER;
    uint32_t v[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      v[i] = e0 + i < nchunks ? row[e0 + i] : 0u;
      sum += v[i];
    }
    uint32_t x = sum;  // inclusive scan of the threads' sums in the wavefront
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) sw[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
      if (w < wave) wbase += sw[w];
      tot += sw[w];
    }
    uint32_t run = carry + wbase + x - sum;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (e0 + i < nchunks) row[e0 + i] = run;
      run += v[i];
    }
    carry += tot;
    __syncthreads();  // sw is written again by the next pass
  }
  if (threadIdx.x == 0) {
    counters[CNT_BAND + q] = carry;
    __threadfence();
    s_last = atomicAdd(&counters[CNT_SCANNED], 1u) == (uint32_t)nl - 1u;
  }
  __syncthreads();
  // the last list: every list's total is in; their exclusive prefix, two
  // lists per lane (nl <= NLIST = 128)
  if (s_last && threadIdx.x < 64) {
    __threadfence();
    const int q0 = 2 * lane, q1 = 2 * lane + 1;
    const uint32_t t0 = q0 < nl ? __hip_atomic_load(&counters[CNT_BAND + q0], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT)
                                : 0u;
    const uint32_t t1 = q1 < nl ? __hip_atomic_load(&counters[CNT_BAND + q1], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT)
                                : 0u;
    const uint32_t pair = t0 + t1;
    uint32_t x = pair;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    const uint32_t before = x - pair;
    if (q0 < nl) counters[CNT_BSTART + q0] = before;
    if (q1 < nl) counters[CNT_BSTART + q1] = before + t0;
    if (lane == 63) counters[CNT_BSTART + nl] = x;
  }
}

// Candidate generation, one chunk of GEN_CHUNK candidate indices per
// workgroup (cell order, so a chunk covers one or a few cells).  Each
// candidate k of cell c (t = k - cstart[c]) is drawn, given its priority
// and written to its (band, region) list as {i, j, priority, t}, or as
// OP_NONE when it has no partner; band 0 enters the first round: its
// priority is stored at both of its particles in tab0.
template <typename R, int MODEL>
__global__ __launch_bounds__(256) void k_gen(const R *__restrict__ P,
                                             const typename R::real_type *__restrict__ ys,
                                             const uint32_t *__restrict__ cnt,
                                             const uint32_t *__restrict__ start,
                                             const uint32_t *__restrict__ cstart, int nc,
                                             uint32_t M, KParams kp, uint32_t step_lo,
                                             uint32_t step_hi, uint32_t prio_key,
                                             uint32_t band_mul, int nbands,
                                             const uint32_t *__restrict__ table,
                                             uint32_t tstride,
                                             const uint32_t *__restrict__ gcells, int32_t reach,
                                             uint4 *__restrict__ bands,
                                             uint32_t *__restrict__ tab0,
                                             uint32_t *__restrict__ counters,
                                             const StepDev *__restrict__ sd) {
  __shared__ uint32_t spos[NLIST];
  __shared__ uint32_t s_cs[GEN_CELLS + 1];
  __shared__ uint32_t s_wn[GEN_WIN], s_ws[GEN_WIN];
  const int nl = nbands * NREG;
  // consecutive chunks (one cell's candidates, drawing from its particles) on
  // one XCD, whose L2 then serves their reads of the cell's records
  uint32_t nchunks = gridDim.x;
  if (sd) {
    // a captured step: its candidates, step and key; the grid is sized for
    // the most candidates and the chunks past this step's idle
    if (sd_stopped(sd)) return;
    M = cstart[nc];
    step_lo = sd->lo;
    step_hi = sd->hi;
    prio_key = sd->pk;
    nchunks = (M + GEN_CHUNK - 1) / GEN_CHUNK;
    if (blockIdx.x >= nchunks) return;
  }
  const uint32_t bx = xcd_block(blockIdx.x, nchunks);
  const uint32_t lo = bx * GEN_CHUNK;
  // the chunk's cells c0..c1 (their candidate starts in LDS when they fit)
  // and the cells its partners can lie in, [wlo, wlo + wn) (their counts
  // and starts in LDS when they fit); lookups outside fall back to memory
  const uint32_t c0 = gcells[2 * bx], c1 = gcells[2 * bx + 1];
  // a chunk inside one replica (the usual case): that replica's number for
  // every candidate, instead of a lookup per candidate
  const int32_t rq0 = kp.nrep ? kp.cell_rep[c0] : 0;
  const bool one_rep = kp.nrep && kp.cell_rep[c1] == rq0;
  auto rep_of = [&](uint32_t cc) { return one_rep ? rq0 : kp.cell_rep[cc]; };
  const int ncs = (int)(c1 - c0) + 2;
  const bool lcs = ncs <= GEN_CELLS + 1;
  int32_t wlo = (int32_t)c0 - reach, whi = (int32_t)c1 + reach + 1;
  if (wlo < 0) wlo = 0;
  if (whi > nc) whi = nc;
  const int32_t wn = whi - wlo;
  const bool lw = reach >= 0 && wn <= GEN_WIN;
  if ((int)threadIdx.x < nl)  // list start + this chunk's offset inside the list
    spos[threadIdx.x] = counters[CNT_BSTART + threadIdx.x] +
                        table[(size_t)threadIdx.x * tstride + bx];
  if (lcs)
    for (int q = threadIdx.x; q < ncs; q += blockDim.x) s_cs[q] = cstart[c0 + q];
  if (lw)
    for (int q = threadIdx.x; q < wn; q += blockDim.x) {
      s_wn[q] = cnt[wlo + q];
      s_ws[q] = start[wlo + q];
    }
  __syncthreads();
  auto cs = [&](uint32_t cc) { return lcs ? s_cs[cc - c0] : cstart[cc]; };
  auto ccnt = [&](int32_t J) {
    const uint32_t o = (uint32_t)(J - wlo);
    return (lw && o < (uint32_t)wn) ? s_wn[o] : cnt[J];
  };
  auto cstr = [&](int32_t J) {
    const uint32_t o = (uint32_t)(J - wlo);
    return (lw && o < (uint32_t)wn) ? s_ws[o] : start[J];
  };
  const bool clamp = MODEL == COLL_ESMC4 && kp.boundary != MCPM_BC_PERIODIC;
  const uint32_t x0 = op_region(lo, M);
  uint32_t c = c0;  // this thread's cell, walking forward with its candidates
  uint32_t nvalid = 0;
  // U candidates per thread at a time: their position loads are issued
  // together; the passes of 256 U candidates tile the chunk exactly
  constexpr int U = 4;
  static_assert(GEN_CHUNK % (256 * U) == 0, "candidate passes must tile the chunk");
  const uint32_t kend = lo + GEN_CHUNK < M ? lo + GEN_CHUNK : M;
  for (uint32_t b0 = 0; b0 < GEN_CHUNK; b0 += blockDim.x * U) {
    uint32_t kk[U], cc[U], tt[U], ii[U];
    bool live[U], drawn[U];
    double d0[U], d1[U];
    uint32_t w3[U];
    R ri[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t k = lo + b0 + u * blockDim.x + threadIdx.x;
      kk[u] = k;
      live[u] = k < kend;
      drawn[u] = false;
      ii[u] = 0;
      tt[u] = 0;
      cc[u] = c;
      if (!live[u]) continue;
      // empty cells have cstart[c] == cstart[c + 1]; the chunk's candidates
      // lie in its cells c0..c1
      while (c < c1 && cs(c + 1) <= k) ++c;
      cc[u] = c;
      tt[u] = k - cs(c);
      const uint32_t Nc = ccnt((int32_t)c);
      if (MODEL == COLL_DSMC || Nc == 0) continue;
      // the candidate's draws: i in c, the direction, j's uniform
      // (collisions.cpp:648-679); a replica's own key and local cell
      const int32_t rqc = kp.nrep ? rep_of(c) : 0;
      const KParams &kq = kp.nrep ? kp.rkp[rqc] : kp;
      const uint32_t cl = kp.nrep ? c - (uint32_t)kp.rcell_off[rqc] : c;
      const mcpm_u32x4 w =
          mcpm_philox(tt[u], cl, step_lo, mcpm_ctr3(MCPM_TAG_CAND, step_hi, 0u), kq.key0, kq.key1);
      ii[u] = cstr((int32_t)c) + mcpm_uint_below(w.v[0], Nc);
      double e2;
      mcpm_iso_dir(w.v[1], w.v[2], &d0[u], &d1[u], &e2);
      w3[u] = w.v[3];
      drawn[u] = true;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (drawn[u]) ri[u] = position_of(P, ys, ii[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!live[u]) continue;
      const uint32_t k = kk[u], cu = cc[u], t = tt[u];
      uint32_t i = ii[u], j = 0;
      bool ok = false;
      if (MODEL == COLL_DSMC) {
        if (ccnt((int32_t)cu) > 0) {
          draw_dsmc_pair(cu, k, cnt, start, kp, step_lo, step_hi, &i, &j);
          ok = true;
        }
      } else if (drawn[u]) {
        int32_t J;
        bool inside;
        if (kp.nrep == 0) {
          if (clamp) {
            if constexpr (R::dims == 1) {
              J = partner_cell_clamped(ri[u], d1[u], kp);
              inside = true;
            } else {
              inside = false;
            }
          } else {
            inside = partner_cell(ri[u], d0[u], d1[u], kp, &J);
          }
        } else {
          const int32_t rq = rep_of(cu);
          inside = partner_cell(ri[u], d0[u], d1[u], kp.rkp[rq], &J);
          J += kp.rcell_off[rq];
        }
        if (inside) {
          const uint32_t NJ = ccnt(J);
          if (NJ > 0) {
            j = cstr(J) + mcpm_uint_below(w3[u], NJ);
            ok = j != i;
          }
        }
      }
      uint32_t pr, band;
      if (MODEL == COLL_ESMC4) {
        pr = (M - cs(cu + 1)) + t;  // rank_new4
        band = band_new4(pr, M, nbands);
      } else {
        pr = kp.nrep ? cand_priority_rep(kp, rep_of(cu), k, MODEL == COLL_ESMC)
                     : cand_priority(k, prio_key, MODEL == COLL_ESMC);
        band = op_band(pr, band_mul, nbands);
      }
      const uint32_t q = band * NREG + region_from(k, x0, M);
      if (ok) {
        ++nvalid;
        if (band == 0) {
          tab0[i] = pr;
          tab0[j] = pr;
        }
      }
      const uint32_t pos = atomicAdd(&spos[q], 1u);
      // the fourth word: the candidate's index in its cell (ESMC) or its cell (DSMC)
      const uint32_t w4 = MODEL == COLL_DSMC ? cu : t;
      bands[pos] = ok ? make_uint4(i, j, pr, w4) : make_uint4(OP_NONE, 0u, 0u, 0u);
    }
  }
  for (int off = 32; off > 0; off >>= 1) nvalid += __shfl_down(nvalid, off);
  if (lane_id() == 0 && nvalid)
    atomicAdd(&counters[CNT_SPREAD + (bx % NSPREAD) * SPREAD_STRIDE], nvalid);
}

// The collision itself, collisions.cpp:680-716: Y at the midpoint cell,
// p = 4 pi sigma^2 (g.dir) Y n_J dt, omega_I = max(omega_I, p),
// accept iff p / omega_max_I > U, hard-sphere exchange along dir.
// The cell I of the candidate is the cell of particle i (i was drawn in it).
// REP = false: the context holds no batched replicas (kp.nrep == 0), which
// keeps the replica lookup out of the code.
template <typename R, bool REP = true>
__device__ __forceinline__ bool commit_op(uint32_t i, uint32_t j, uint32_t t, R a, R bb, R *P,
                                          const double *nden, const double *yc,
                                          const double *omax, double *omega_new,
                                          const KParams &kp, uint32_t step_lo,
                                          uint32_t step_hi, bool new4 = false) {
  using T = typename R::real_type;
  // a batched replica: its geometry and key, cells counted from its first
  const bool rep = REP && kp.nrep;
  const int32_t rq = rep ? rep_of_particle(kp, (int64_t)i) : 0;
  const KParams &kq = rep ? kp.rkp[rq] : kp;
  const int32_t off = rep ? kp.rcell_off[rq] : 0;
  const uint32_t cl = (uint32_t)cell_key(a, kq);
  const uint32_t c = cl + (uint32_t)off;
  mcpm_u32x4 w = mcpm_philox(t, cl, step_lo, mcpm_ctr3(MCPM_TAG_CAND, step_hi, 0u), kq.key0,
                             kq.key1);
  mcpm_u32x4 w2 = mcpm_philox(t, cl, step_lo, mcpm_ctr3(MCPM_TAG_CAND, step_hi, 1u), kq.key0,
                              kq.key1);
  double d0, d1, d2;
  mcpm_iso_dir(w.v[1], w.v[2], &d0, &d1, &d2);
  int32_t cm = off + mid_cell(a, bb, kq);
  int32_t J = off + cell_key(bb, kq);
  double Ymid = yc[cm];
  double g0 = (double)a.u1 - (double)bb.u1;
  double g1 = (double)a.u2 - (double)bb.u2;
  double g2 = (double)a.u3 - (double)bb.u3;
  double gd = g0 * d0 + g1 * d1 + g2 * d2;
  double pv = kp.coll_pref * gd * Ymid * nden[J] * kp.dt;
  // the step's running maximum starts at omega_max (k_omega_finish), so only
  // a p above omega_max can raise it
  const double om = omax[c];
  if (pv > om) atomic_max_pos_double(&omega_new[c], pv);
  // new4 also requires an approaching pair, g.dir > 0 (collisions.cpp:819)
  if (pv / om > mcpm_u01(w2.v[0]) && (!new4 || gd > 0.0)) {
    a.u1 = (T)((double)a.u1 - gd * d0);
    a.u2 = (T)((double)a.u2 - gd * d1);
    a.u3 = (T)((double)a.u3 - gd * d2);
    bb.u1 = (T)((double)bb.u1 + gd * d0);
    bb.u2 = (T)((double)bb.u2 + gd * d1);
    bb.u3 = (T)((double)bb.u3 + gd * d2);
    P[i] = a;
    P[j] = bb;
    return true;
  }
  return false;
}

// The CBA displacement (collisions.cpp:2221-2238, 2325-2340): both particles
// move by -+sigma along the change of their relative velocity, read back from
// the stored (rounded) post-collision velocities.  The reference divides by
// the norm of that change unguarded; a pair whose stored relative velocity
// did not change is left in place instead of moved to NaN.
template <typename R>
__device__ __forceinline__ void cba_displace(R &a, R &bb, double pre0, double pre1, double pre2,
                                             double sigma) {
  using real = typename R::real_type;
  const double e0 = ((double)a.u1 - (double)bb.u1) - pre0;
  const double e1 = ((double)a.u2 - (double)bb.u2) - pre1;
  const double e2 = ((double)a.u3 - (double)bb.u3) - pre2;
  const double nrm = __builtin_sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  if (!(nrm > 0.0)) return;
  const double d0 = e0 * (sigma / nrm), d1 = e1 * (sigma / nrm);
  if constexpr (R::dims == 2) {
    a.x = (real)((double)a.x + d0);
    bb.x = (real)((double)bb.x - d0);
  }
  a.y = (real)((double)a.y + d1);
  bb.y = (real)((double)bb.y - d1);
}

// Collision_DSMC_VHS commit (collisions.cpp:2099-2131): VHS cross-section
// sigma_T ~ Vr^(2-2 visp) crref^(2 visp - 1), accept iff sigma_T / crm > U,
// isotropic post-collision relative velocity about the centre of mass.
// fpp.model selects the variant: DSMC_HS weighs by sigma_T = Vr
// (collisions.cpp:2026), CBA_VHS adds the displacement.  Candidate k of
// cell c (both from the operation record).
template <typename R>
__device__ __forceinline__ bool commit_dsmc(uint32_t i, uint32_t j, uint32_t k, uint32_t c, R a,
                                            R bb, R *P, const double *omax, double *omega_new,
                                            const KParams &kp, const FPParams &fpp,
                                            uint32_t step_lo, uint32_t step_hi) {
  using real = typename R::real_type;
  mcpm_u32x4 w = mcpm_philox(k, c, step_lo, mcpm_ctr3(MCPM_TAG_DSMC, step_hi, 0u), kp.key0,
                             kp.key1);
  double g0 = (double)a.u1 - (double)bb.u1;
  double g1 = (double)a.u2 - (double)bb.u2;
  double g2 = (double)a.u3 - (double)bb.u3;
  double Vr = __builtin_sqrt(g0 * g0 + g1 * g1 + g2 * g2);
  double sig = fpp.model == MCPM_MODEL_DSMC_HS
                   ? Vr
                   : det_pow(Vr, 2.0 - 2.0 * fpp.visp) * det_pow(fpp.crref, 2.0 * fpp.visp - 1.0);
  atomic_max_pos_double(&omega_new[c], sig);
  if (!(sig / omax[c] > mcpm_u01(w.v[2]))) return false;
  mcpm_u32x4 w2 = mcpm_philox(k, c, step_lo, mcpm_ctr3(MCPM_TAG_DSMC, step_hi, 1u), kp.key0,
                              kp.key1);
  double ct = 2.0 * mcpm_u01(w.v[3]) - 1.0;
  double st = __builtin_sqrt(1.0 - ct * ct);
  double sp, cp;
  mcpm_sincos2pi(mcpm_u01(w2.v[0]), &sp, &cp);
  double v0 = Vr * ct, v1 = Vr * st * cp, v2 = Vr * st * sp;
  double c0 = 0.5 * ((double)a.u1 + (double)bb.u1);
  double c1 = 0.5 * ((double)a.u2 + (double)bb.u2);
  double c2 = 0.5 * ((double)a.u3 + (double)bb.u3);
  a.u1 = (real)(c0 + 0.5 * v0);
  a.u2 = (real)(c1 + 0.5 * v1);
  a.u3 = (real)(c2 + 0.5 * v2);
  bb.u1 = (real)(c0 - 0.5 * v0);
  bb.u2 = (real)(c1 - 0.5 * v1);
  bb.u3 = (real)(c2 - 0.5 * v2);
  if (fpp.model == MCPM_MODEL_CBA_VHS) cba_displace(a, bb, g0, g1, g2, kp.sigma);
  P[i] = a;
  P[j] = bb;
  return true;
}

// Deterministic reservations without atomics on the common path.  An
// operation commits when it holds the smallest priority on both of its
// particles among the operations of its round, which is exactly when every
// earlier-priority operation sharing a particle has committed: the result
// equals processing all operations one by one in priority order.
// Operations enter by priority band: round r takes the losers of round r-1
// plus band r.  Every operation of a lower priority than an admitted one is
// admitted too, so the smallest priority among a round's operations is the
// smallest among all unresolved ones; late operations no longer lose round
// after round to early ones.
// A round is two launches over its operations:
//  * admission (the round before, or candidate generation for round 0):
//    each operation stores its priority at both particles of tab (plain
//    stores; one of the values stored at a particle survives);
//  * k_check: an operation that does not find its own priority at a
//    particle lowers the word with atomicMin and marks the particle's group
//    (2^MARK_SHIFT consecutive particles share a mark byte of cflag) with
//    the round's tag.  After it tab holds, at every particle, the smallest
//    priority of the round's operations on it, and a particle of an
//    unmarked group has exactly one operation on it;
//  * k_round: an operation wins when each of its particles is in an
//    unmarked group or holds its priority.  Winners commit; losers store
//    their priority into the next round's table and are compacted into the
//    next list, and band r+1 is admitted there.
// Every word a round reads was stored in that round, so the tables are never
// cleared; a stale mark equal to the round's tag, or one that another
// particle of the group set, only sends an operation to read its word, which
// then holds its own priority.  Atomics are left to the particles that
// several operations share.
// Region x of the operations is processed by the workgroups with
// blockIdx % 8 == x, which the hardware places on XCD x: each XCD keeps
// working on the same run of cells, whose particles and reservations stay in
// its L2.  Losers of region x are kept in segment x of the output list.
template <int MODEL>
__global__ __launch_bounds__(256) void k_check(int r, int nbands, uint32_t seg,
                                               const uint4 *__restrict__ in,
                                               const uint4 *__restrict__ bands,
                                               uint32_t *__restrict__ tab,
                                               uint8_t *__restrict__ cflag, uint32_t tag8,
                                               uint32_t *__restrict__ counters,
                                               uint32_t *__restrict__ host_left,
                                               const StepDev *__restrict__ sd) {
  if (sd_stopped(sd)) return;
  if (sd) tag8 = sd->mark + (uint32_t)r;
  // round 0: the draw's valid-operation count, from its slots
  if (r == 0 && blockIdx.x == 0 && threadIdx.x < NSPREAD) {
    static_assert(NSPREAD == 64, "one wavefront sums the slots");
    uint32_t v = counters[CNT_SPREAD + threadIdx.x * SPREAD_STRIDE];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (threadIdx.x == 0) counters[CNT_OPS] = v;
  }
  // publish the losers left after round r - 1 (final now) to the host
  if (host_left && blockIdx.x == 0 && threadIdx.x == 0 && r > 0) {
    uint32_t left = 0;
    for (int q = 0; q < NREG; ++q) left += counters[CNT_LIST + r * NREG + q];
    __hip_atomic_store(&host_left[r], left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  const uint32_t x = blockIdx.x % NREG;
  const uint32_t lb = blockIdx.x / NREG, nbx = gridDim.x / NREG;
  const int qB = r * NREG + (int)x;
  const bool hasB = r < nbands;
  const uint32_t countB = hasB ? counters[CNT_BAND + qB] : 0u;
  const uint32_t sB = hasB ? counters[CNT_BSTART + qB] : 0u;
  const uint32_t countA = (r == 0) ? 0u : counters[CNT_LIST + r * NREG + x];
  const uint32_t count = countA + countB;
  const uint4 *inA = in + (size_t)x * seg;
  const uint4 *inB = bands + sB;
  const uint8_t mark = (uint8_t)tag8;
  constexpr int U = 4;
  const uint32_t tile = blockDim.x * U;
  for (uint32_t base = lb * tile; base < count; base += nbx * tile) {
    uint4 rec[U];
    uint32_t ti[U], tj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t idx = base + u * blockDim.x + threadIdx.x;
      rec[u] = make_uint4(OP_NONE, 0u, 0u, 0u);
      if (idx < countA) rec[u] = inA[idx];
      else if (idx < count) rec[u] = inB[idx - countA];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      ti[u] = tj[u] = 0u;
      if (rec[u].x != OP_NONE) {
        ti[u] = tab[rec[u].x];
        tj[u] = tab[rec[u].y];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (rec[u].x == OP_NONE) continue;
      if (ti[u] != rec[u].z) {
        atomicMin(&tab[rec[u].x], rec[u].z);
        cflag[rec[u].x >> MARK_SHIFT] = mark;
      }
      if (tj[u] != rec[u].z) {
        atomicMin(&tab[rec[u].y], rec[u].z);
        cflag[rec[u].y >> MARK_SHIFT] = mark;
      }
    }
  }
}

template <typename R, int MODEL, bool PLAIN, int U>
__global__ __launch_bounds__(256) void k_round(int r, int nbands, uint32_t seg,
                                               R *__restrict__ P,
                                               const uint4 *__restrict__ in,
                                               const uint4 *__restrict__ bands,
                                               uint4 *__restrict__ out,
                                               const uint32_t *__restrict__ tab_cur,
                                               uint32_t *__restrict__ tab_next,
                                               const uint8_t *__restrict__ cflag, uint32_t tag8,
                                               uint32_t *__restrict__ counters,
                                               const double *__restrict__ nden,
                                               const double *__restrict__ yc,
                                               const double *__restrict__ omax,
                                               double *__restrict__ omega_new, KParams kp,
                                               uint32_t step_lo, uint32_t step_hi,
                                               FPParams fpp, OpLog log,
                                               const StepDev *__restrict__ sd) {
  __shared__ uint32_t s_lose, s_base, s_acc;
  if (sd_stopped(sd)) return;
  if (sd) {
    tag8 = sd->mark + (uint32_t)r;
    step_lo = sd->lo;
    step_hi = sd->hi;
  }
  const uint32_t x = blockIdx.x % NREG;
  const uint32_t lb = blockIdx.x / NREG, nbx = gridDim.x / NREG;
  // band r and band r+1 of region x in the band-major lists
  const int qB = r * NREG + (int)x, qC = (r + 1) * NREG + (int)x;
  const bool hasB = r < nbands, hasC = r + 1 < nbands;
  const uint32_t countB = hasB ? counters[CNT_BAND + qB] : 0u;
  const uint32_t countC = hasC ? counters[CNT_BAND + qC] : 0u;
  const uint32_t sB = hasB ? counters[CNT_BSTART + qB] : 0u;
  const uint32_t sC = hasC ? counters[CNT_BSTART + qC] : 0u;
  const uint32_t countA = (r == 0) ? 0u : counters[CNT_LIST + r * NREG + x];
  const uint32_t count = countA + countB;
  const uint4 *inA = in + (size_t)x * seg;
  const uint4 *inB = bands + sB;
  const uint4 *inC = bands + sC;
  uint4 *outx = out + (size_t)x * seg;
  uint32_t *next_count = &counters[CNT_LIST + (r + 1) * NREG + x];
  const uint32_t stride = nbx * blockDim.x;
  const uint8_t mark = (uint8_t)tag8;
  // admit band r+1 first: its stores overlap with the work below; the lists
  // also hold partnerless candidates as OP_NONE
  for (uint32_t idx = lb * blockDim.x + threadIdx.x; idx < countC; idx += stride) {
    const uint4 rec = inC[idx];
    if (rec.x == OP_NONE) continue;
    tab_next[rec.x] = rec.z;
    tab_next[rec.y] = rec.z;
  }
  // U operations per thread per tile: their loads are issued together
  const uint32_t tile = blockDim.x * U;
  uint32_t acc_local = 0;
  for (uint32_t base = lb * tile; base < count; base += nbx * tile) {
    uint4 rec[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t idx = base + u * blockDim.x + threadIdx.x;
      rec[u] = make_uint4(OP_NONE, 0u, 0u, 0u);
      if (idx < countA) rec[u] = inA[idx];
      else if (idx < count) rec[u] = inB[idx - countA];
    }
    bool fi[U], fj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      fi[u] = fj[u] = false;
      if (rec[u].x != OP_NONE) {
        fi[u] = cflag[rec[u].x >> MARK_SHIFT] == mark;
        fj[u] = cflag[rec[u].y >> MARK_SHIFT] == mark;
      }
    }
    R pa[U], pb[U];
    bool win[U], lose[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool valid = rec[u].x != OP_NONE;
      // a marked particle is shared: its word holds the round's smallest priority
      const bool oki = !fi[u] || (valid && tab_cur[rec[u].x] == rec[u].z);
      const bool okj = !fj[u] || (valid && tab_cur[rec[u].y] == rec[u].z);
      win[u] = valid && oki && okj;
      lose[u] = valid && !win[u];
      if (win[u]) {
        pa[u] = P[rec[u].x];
        pb[u] = P[rec[u].y];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = rec[u].x, j = rec[u].y, pr = rec[u].z;
      if (win[u]) {
        bool acc = MODEL == COLL_DSMC
                       ? commit_dsmc<R>(i, j, pr, rec[u].w, pa[u], pb[u], P, omax, omega_new,
                                        kp, fpp, step_lo, step_hi)
                       : commit_op<R, !PLAIN>(i, j, rec[u].w, pa[u], pb[u], P, nden, yc, omax,
                                      omega_new, kp, step_lo, step_hi, MODEL == COLL_ESMC4);
        acc_local += acc ? 1u : 0u;
        if (!PLAIN && log.p) {
          uint32_t q = atomicAdd(&counters[CNT_LOG], 1u);
          log.p[q] = pr;
          log.i[q] = i;
          log.j[q] = j;
          log.a[q] = acc ? 1 : 0;
        }
      } else if (lose[u]) {
        tab_next[i] = pr;
        tab_next[j] = pr;
      }
    }
    // compact the tile's losers: ranks in the wavefront, one LDS atomic per
    // wavefront, one global atomic per workgroup
    uint64_t lm[U];
    uint32_t wtot = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      lm[u] = __ballot(lose[u]);
      wtot += (uint32_t)__popcll(lm[u]);
    }
    if (threadIdx.x == 0) s_lose = 0;
    __syncthreads();
    uint32_t wbase = 0;
    if (lane_id() == 0 && wtot) wbase = atomicAdd(&s_lose, wtot);
    wbase = __shfl(wbase, 0);
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_lose ? atomicAdd(next_count, s_lose) : 0u;
    __syncthreads();
    uint32_t off = s_base + wbase;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (lose[u]) outx[off + (uint32_t)__popcll(lm[u] & lanemask_lt())] = rec[u];
      off += (uint32_t)__popcll(lm[u]);
    }
  }
  // accepted count, one atomic per workgroup
  for (int off = 32; off > 0; off >>= 1) acc_local += __shfl_down(acc_local, off);
  if (threadIdx.x == 0) s_acc = 0;
  __syncthreads();
  if (lane_id() == 0 && acc_local) atomicAdd(&s_acc, acc_local);
  __syncthreads();
  if (threadIdx.x == 0 && s_acc)
    atomicAdd((unsigned long long *)&counters[CNT_ACC], (unsigned long long)s_acc);
}

// The rounds of a captured step after its grid rounds, in one workgroup
// (by then few operations are left, and a fixed count of grid rounds per
// step leaves no host in the loop).  Round r reads the losers of round r - 1
// (region x at in + x seg, their count in the counters); each round is the
// same two phases as k_check and k_round, separated by workgroup barriers,
// so it commits exactly the operations a grid round would.  Every round
// commits at least the operation of smallest priority, so the loop ends; it
// stops the run (STOP_ROUNDS) at the round limit.
template <typename R, int MODEL>
__global__ __launch_bounds__(1024) void k_round_tail(int r0, int nbands, uint32_t seg,
                                                     R *__restrict__ P,
                                                     uint4 *__restrict__ opA,
                                                     uint4 *__restrict__ opB,
                                                     uint32_t *__restrict__ tab0,
                                                     uint32_t *__restrict__ tab1,
                                                     uint8_t *__restrict__ cflag,
                                                     uint32_t *__restrict__ counters,
                                                     const double *__restrict__ nden,
                                                     const double *__restrict__ yc,
                                                     const double *__restrict__ omax,
                                                     double *__restrict__ omega_new, KParams kp,
                                                     FPParams fpp, StepDev *__restrict__ sd) {
  __shared__ uint32_t s_cnt[NREG], s_out[NREG], s_pre[NREG + 1];
  __shared__ uint32_t s_acc;
  if (sd_stopped(sd)) return;
  const uint32_t step_lo = sd->lo, step_hi = sd->hi;
  const uint32_t tid = threadIdx.x;
  uint32_t acc_local = 0, visits = 0;
  // the losers of the last grid round; the tail's own rounds keep their
  // counts in LDS
  if (tid < NREG) s_cnt[tid] = counters[CNT_LIST + r0 * NREG + (int)tid];
  __syncthreads();
  int r = r0;
  for (;;) {
    if (tid == 0) {
      uint32_t run = 0;
      for (int x = 0; x < NREG; ++x) {
        s_pre[x] = run;
        run += s_cnt[x];
      }
      s_pre[NREG] = run;
    }
    if (tid < NREG) s_out[tid] = 0u;
    __syncthreads();
    const uint32_t total = s_pre[NREG];
    if (total == 0u) break;
    if (r >= MAX_ROUNDS) {
      if (tid == 0) atomicOr(&sd->stop, STOP_ROUNDS);
      return;
    }
    visits += total;
    const uint4 *in = (r & 1) ? opB : opA;
    uint4 *out = (r & 1) ? opA : opB;
    uint32_t *tab = (r & 1) ? tab1 : tab0;
    uint32_t *tabn = (r & 1) ? tab0 : tab1;
    const uint8_t mark = (uint8_t)(sd->mark + (uint32_t)r);
    // the round's operations as one index over the regions' segments, so a
    // thread's loads do not wait region after region
    auto region_of = [&](uint32_t g) {
      int x = 0;
      while (g >= s_pre[x + 1]) ++x;
      return x;
    };
    // phase 1 (k_check): the smallest priority at every shared particle
    for (uint32_t g = tid; g < total; g += blockDim.x) {
      const int x = region_of(g);
      const uint4 rec = in[(size_t)x * seg + (g - s_pre[x])];
      if (rec.x == OP_NONE) continue;
      if (tab[rec.x] != rec.z) {
        atomicMin(&tab[rec.x], rec.z);
        cflag[rec.x >> MARK_SHIFT] = mark;
      }
      if (tab[rec.y] != rec.z) {
        atomicMin(&tab[rec.y], rec.z);
        cflag[rec.y >> MARK_SHIFT] = mark;
      }
    }
    __syncthreads();
    // phase 2 (k_round): winners commit, losers enter round r + 1
    for (uint32_t g = tid; g < total; g += blockDim.x) {
      const int x = region_of(g);
      const uint4 rec = in[(size_t)x * seg + (g - s_pre[x])];
      if (rec.x == OP_NONE) continue;
      const bool oki = cflag[rec.x >> MARK_SHIFT] != mark || tab[rec.x] == rec.z;
      const bool okj = cflag[rec.y >> MARK_SHIFT] != mark || tab[rec.y] == rec.z;
      if (oki && okj) {
        const R a = P[rec.x], bb = P[rec.y];
        const bool acc =
            MODEL == COLL_DSMC
                ? commit_dsmc<R>(rec.x, rec.y, rec.z, rec.w, a, bb, P, omax, omega_new, kp,
                                 fpp, step_lo, step_hi)
                : commit_op<R>(rec.x, rec.y, rec.w, a, bb, P, nden, yc, omax, omega_new, kp,
                               step_lo, step_hi, MODEL == COLL_ESMC4);
        acc_local += acc ? 1u : 0u;
      } else {
        tabn[rec.x] = rec.z;
        tabn[rec.y] = rec.z;
        out[(size_t)x * seg + atomicAdd(&s_out[x], 1u)] = rec;
      }
    }
    __syncthreads();
    if (tid < NREG) {
      counters[CNT_LIST + (r + 1) * NREG + (int)tid] = s_out[tid];
      s_cnt[tid] = s_out[tid];
    }
    __syncthreads();
    ++r;
  }
  for (int off = 32; off > 0; off >>= 1) acc_local += __shfl_down(acc_local, off);
  if (tid == 0) s_acc = 0;
  __syncthreads();
  if (lane_id() == 0 && acc_local) atomicAdd(&s_acc, acc_local);
  __syncthreads();
  if (tid == 0) {
    if (s_acc) atomicAdd((unsigned long long *)&counters[CNT_ACC], (unsigned long long)s_acc);
    // rounds of the step: the grid rounds up to the first one that found
    // nothing left after every band was admitted, or all of them and the tail's
    int used = r;
    if (r == r0) {
      for (int q = nbands; q < r0; ++q) {
        uint32_t left = 0;
        for (int x = 0; x < NREG; ++x) left += counters[CNT_LIST + q * NREG + x];
        if (left == 0u) {
          used = q;
          break;
        }
      }
    }
    sd->rounds = (uint32_t)used;
    sd->tail_ops += visits;
  }
}

// After the candidate scan of a captured step: stop before anything is
// drawn when the candidates exceed the operation lists; the step's priority
// key (the same Philox block the host derives for an eager step).
__global__ void k_step_mid(StepDev *__restrict__ sd, const uint32_t *__restrict__ cstart, int nc,
                           uint32_t cap, uint32_t key0, uint32_t key1) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || sd->stop) return;
  if (cstart[nc] > cap) {
    sd->stop |= STOP_CAPACITY;
    return;
  }
  sd->pk = mcpm_philox(0u, 0u, sd->lo, mcpm_ctr3(MCPM_TAG_PRIO, sd->hi, 0u), key0, key1).v[0];
}

// End of a captured step: the step's counts into the totals; a windowed sort
// that a particle left stops the run with the step's collisions and field
// done and its push and sort to be redone by the host; otherwise the next
// step.
__global__ void k_step_end(StepDev *__restrict__ sd, const uint32_t *__restrict__ counters,
                           const uint32_t *__restrict__ cstart, int nc,
                           const uint32_t *__restrict__ wflag) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || sd->stop) return;
  const unsigned long long acc = *(const unsigned long long *)&counters[CNT_ACC];
  sd->cand += cstart[nc];
  sd->valid += counters[CNT_OPS];
  sd->acc += acc;
  sd->rsum += sd->rounds;
  sd->mark += sd->rounds + 1u;
  if (wflag[0] != 0u) {
    sd->stop |= STOP_WINDOW;
    return;
  }
  const uint64_t st = (((uint64_t)sd->hi << 32) | sd->lo) + 1u;
  sd->lo = (uint32_t)st;
  sd->hi = (uint32_t)(st >> 32);
  sd->steps += 1u;
}

template <typename R>
__global__ void k_ys_from_records(const R *__restrict__ P, int64_t n,
                                  typename R::real_type *__restrict__ ys) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    ys[k] = P[k].y;
}

__global__ void k_omega_finish(double *__restrict__ omax, double *__restrict__ omega_new,
                               const uint32_t *__restrict__ cnt, int nc, int model,
                               int variant, const StepDev *__restrict__ sd) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc || sd_stopped(sd)) return;
  double o = omega_new[c];
  if (model == COLL_DSMC) {
    // crm <- max(1.1 crm_seen, 0.9 crm) (collisions.cpp:2136), in cells with
    // two or more particles for DSMC_VHS, in every cell for DSMC_HS and
    // CBA_VHS (collisions.cpp:2059, 2247)
    if (cnt[c] > 1 || variant != MCPM_MODEL_DSMC_VHS) {
      double a = 1.1 * o, b = 0.9 * omax[c];
      omax[c] = a > b ? a : b;
    }
    omega_new[c] = 0.0;
  } else {
    // collisions.cpp:714-716; the next step's maximum starts from the new
    // omega_max (the same max, and commits below it skip the atomic)
    if (o > omax[c]) omax[c] = o;
    omega_new[c] = omax[c];
  }
}

// initialization (Main_Functions.cpp:4222-4815) for particle i: Philox block
// (i, TAG_INIT, 0) for the position, (i, TAG_INIT, 1) and (i, TAG_INIT, 2)
// for two Box-Muller pairs of which three normals make the velocity.
// Positions uniform along streaming axes, L/2 elsewhere (:4657-4671);
// injection (:4640-4655): the last n_drop particles at r = R sqrt(u),
// theta = 2 pi u' about the centre.  Velocities U0 + sd z (:4790-4800).  A
// position that rounds up to L in the stored type is put at 0.
template <typename R>
__global__ void k_init(R *__restrict__ P, InitParams ip, KParams kp) {
  using real = typename R::real_type;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ip.n) return;
  const uint32_t lo = (uint32_t)i, hi = (uint32_t)((uint64_t)i >> 32);
  const mcpm_u32x4 w = mcpm_philox(lo, hi, 0u, mcpm_ctr3(MCPM_TAG_INIT, 0u, 0u), kp.key0, kp.key1);
  const mcpm_u32x4 w1 = mcpm_philox(lo, hi, 0u, mcpm_ctr3(MCPM_TAG_INIT, 0u, 1u), kp.key0, kp.key1);
  double x, y;
  if (ip.disc && i >= ip.n - ip.n_drop) {
    const double r = ip.drop_r * __builtin_sqrt(mcpm_u01(w.v[0]));
    double sn, cs;
    mcpm_sincos2pi(mcpm_u01(w.v[1]), &sn, &cs);
    x = 0.5 * ip.len[0] + r * cs;
    y = 0.5 * ip.len[1] + r * sn;
  } else if (ip.disc) {
    x = mcpm_u01(w.v[0]) * ip.len[0];
    y = mcpm_u01(w.v[1]) * ip.len[1];
  } else {
    x = ip.dirs[0] ? mcpm_u01(w.v[0]) * ip.len[0] : 0.5 * ip.len[0];
    y = ip.dirs[1] ? mcpm_u01(w.v[1]) * ip.len[1] : 0.5 * ip.len[1];
  }
  double z0, z1, z2, z3;
  mcpm_normal2(w1.v[0], w1.v[1], &z0, &z1);
  mcpm_normal2(w1.v[2], w1.v[3], &z2, &z3);
  R r{};
  real yr = (real)y;
  if (!(yr < (real)ip.len[1])) yr = (real)0;
  r.y = yr;
  if constexpr (R::dims == 2) {
    real xr = (real)x;
    if (!(xr < (real)ip.len[0])) xr = (real)0;
    r.x = xr;
  }
  r.u1 = (real)(ip.U0[0] + ip.sd * z0);
  r.u2 = (real)(ip.U0[1] + ip.sd * z1);
  r.u3 = (real)(ip.U0[2] + ip.sd * z2);
  P[i] = r;
}

// The pressure columns of PostProc (myIO.cpp:855, 890): n kB T (1 + n b Y(n))
// and that minus the mean-field 2 b n^2 phi, from the moments' T and the cell
// density.
__global__ void k_pressure(const double *__restrict__ mom, const double *__restrict__ nden,
                           const double *__restrict__ yc, double *__restrict__ p_rep,
                           double *__restrict__ p_tot, KParams kp, double phi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= kp.n_cells) return;
  const double n = nden[c], T = mom[(size_t)c * MCPM_NMOM + 29];
  const double pe = n * kp.kb * T * (1.0 + n * kp.b * yc[c]);
  p_rep[c] = pe;
  p_tot[c] = pe - 2.0 * kp.b * n * n * phi;
}

__global__ void k_rep_step(KParams kp, const uint32_t *__restrict__ cstart,
                           uint32_t *__restrict__ rprio, uint32_t *__restrict__ rcand_off,
                           uint32_t step_lo, uint32_t step_hi, const StepDev *__restrict__ sd) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > kp.nrep || sd_stopped(sd)) return;
  if (sd) {
    step_lo = sd->lo;
    step_hi = sd->hi;
  }
  rcand_off[r] = cstart[kp.rcell_off[r]];
  if (r < kp.nrep) {
    const KParams &kq = kp.rkp[r];
    rprio[r] = mcpm_philox(0u, 0u, step_lo, mcpm_ctr3(MCPM_TAG_PRIO, step_hi, 0u), kq.key0,
                           kq.key1).v[0];
  }
}

__global__ void k_zero_u32(uint32_t *p, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}
__global__ void k_iota_u32(uint32_t *p, int64_t n) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = (uint32_t)k;
}
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; k < n; k += (int64_t)gridDim.x * blockDim.x) p[k] = v;
}

// ---------------------------------------------------------------------------
// 1D screened-Poisson mean field (poisson_solver, Main_Functions.cpp:3285-3581)
// ---------------------------------------------------------------------------
// x(y) solves (-d2/dy2 + bst^2) x = n on the faces, discretised as
// (2 + (dy bst)^2) x_f - x_{f-1} - x_{f+1} = dy^2 n_f, with n_f the mean of
// the two adjacent cells.  Walls: Dirichlet values from the truncated Yukawa
// integral (50x50 Gauss-Legendre, the reference's 8-digit table).  Periodic:
// cyclic system by Sherman-Morrison.  The constant tridiagonal matrix is
// factored once on the host; each step runs the forward and backward
// substitutions as block-wide scans of affine maps x -> A x + B.
// F2_c = -(x_{c+1} - x_c) / dy * ast / m after subtracting the analytic core
// part n_f (1 - (1 + bst sigma) e^{-bst sigma}) / bst^2 (:3539-3577).

__device__ void block_affine_scan(const double *__restrict__ A, const double *__restrict__ B,
                                  double *__restrict__ out, int n, bool reverse, double *sA,
                                  double *sB) {
  const int tid = threadIdx.x;
  const int nt = blockDim.x;
  const int E = (n + nt - 1) / nt;
  const int s0 = tid * E;
  const int s1 = min(n, s0 + E);
  double a = 1.0, b = 0.0;
  for (int s = s0; s < s1; ++s) {
    int k = reverse ? (n - 1 - s) : s;
    a = A[k] * a;
    b = A[k] * b + B[k];
  }
  sA[tid] = a;
  sB[tid] = b;
  __syncthreads();
  for (int off = 1; off < nt; off <<= 1) {
    double pa = 1.0, pb = 0.0;
    bool has = tid >= off;
    if (has) {
      pa = sA[tid - off];
      pb = sB[tid - off];
    }
    __syncthreads();
    if (has) {
      double ca = sA[tid], cb = sB[tid];
      sA[tid] = ca * pa;
      sB[tid] = ca * pb + cb;
    }
    __syncthreads();
  }
  double v = (tid == 0) ? 0.0 : sB[tid - 1];
  __syncthreads();
  for (int s = s0; s < s1; ++s) {
    int k = reverse ? (n - 1 - s) : s;
    v = A[k] * v + B[k];
    out[k] = v;
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_field(const double *__restrict__ nden,
                                                FieldArrays fa0, double *__restrict__ F2,
                                                KParams kp0, double ast, double bst,
                                                double eqdist,
                                                const FieldArrays *__restrict__ rfa,
                                                const StepDev *__restrict__ sd) {
  __shared__ double sA[1024];
  __shared__ double sB[1024];
  __shared__ double red[2][1024];
  if (sd_stopped(sd)) return;
  // batched replicas: workgroup b solves replica b's slab
  const bool batch = kp0.nrep > 0;
  const KParams &kp = batch ? kp0.rkp[blockIdx.x] : kp0;
  const FieldArrays fa = batch ? rfa[blockIdx.x] : fa0;
  if (batch) {
    nden += kp0.rcell_off[blockIdx.x];
    F2 += kp0.rcell_off[blockIdx.x];
  }
  const int tid = threadIdx.x;
  const int nc = kp.n_cells;
  const bool periodic = kp.boundary == MCPM_BC_PERIODIC;
  const int nf = periodic ? nc : nc + 1;
  const double dy = kp.dy;
  // face densities (Main_Functions.cpp:3352-3372)
  for (int f = tid; f < nf; f += blockDim.x) {
    double nj;
    if (periodic) nj = (nden[f] + nden[(f + nc - 1) % nc]) / 2.0;
    else if (f == 0) nj = nden[0];
    else if (f == nc) nj = nden[nc - 1];
    else nj = (nden[f] + nden[f - 1]) / 2.0;
    fa.nj[f] = nj;
    fa.rhs[f] = nj * (dy * dy);
  }
  if (!periodic) {
    // Dirichlet values: truncated Yukawa integrals at both walls
    // (Main_Functions.cpp:3322-3442).
    double s_lo = 0.0, s_hi = 0.0;
    for (int q = tid; q < 2500; q += blockDim.x) {
      int ir = q / 50, ip = q % 50;
      double rr = fa.gl_r[ir], dr = fa.gl_dr[ir];
      double dph = fa.gl_dphi[ip];
      double sp = fa.gl_sin[ip], cp = fa.gl_cos[ip];
      double radial = mcpm_exp(-bst * rr) / 2.0 * rr * dr;
      for (int side = 0; side < 2; ++side) {
        double yf = (side == 0 ? 0.0 : (double)nc * dy) + rr * cp;
        double qq = mcpm_floor(yf / dy);
        double n;
        if (qq < 0.0) n = nden[0];
        else if (qq > (double)(nc - 1)) n = nden[nc - 1];
        else n = nden[(int)qq];
        double term = radial * n * sp * dph;
        if (side == 0) s_lo += term; else s_hi += term;
      }
    }
    red[0][tid] = s_lo;
    red[1][tid] = s_hi;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if (tid < off) {
        red[0][tid] += red[0][tid + off];
        red[1][tid] += red[1][tid + off];
      }
      __syncthreads();
    }
    if (tid == 0) {
      fa.rhs[0] = red[0][0];
      fa.rhs[nf - 1] = red[1][0];
    }
    __syncthreads();
  }
  __syncthreads();
  // forward: y_k = -l_k y_{k-1} + r_k
  for (int f = tid; f < nf; f += blockDim.x) fa.tB[f] = fa.rhs[f];
  __syncthreads();
  block_affine_scan(fa.mL, fa.tB, fa.y, nf, false, sA, sB);
  // backward: x_k = (y_k - c_k x_{k+1}) / d_k
  for (int f = tid; f < nf; f += blockDim.x) fa.tB[f] = fa.y[f] * fa.invd[f];
  __syncthreads();
  block_affine_scan(fa.mU, fa.tB, fa.x, nf, true, sA, sB);
  if (periodic) {
    __shared__ double corr;
    if (tid == 0) {
      double vy = fa.x[0] + fa.sm_vlast * fa.x[nf - 1];
      corr = vy / fa.sm_den;
    }
    __syncthreads();
    for (int f = tid; f < nf; f += blockDim.x) fa.x[f] = fa.x[f] - corr * fa.sm_z[f];
    __syncthreads();
  }
  // core subtraction and per-cell force
  double core = 1.0 - (kp.sigma * bst + 1.0) * mcpm_exp(-bst * kp.sigma);
  core = core / bst / bst;
  for (int f = tid; f < nf; f += blockDim.x) fa.y[f] = fa.x[f] - fa.nj[f] * core;
  __syncthreads();
  for (int c = tid; c < nc; c += blockDim.x) {
    int f1 = periodic ? (c + 1) % nc : c + 1;
    double F = -(fa.y[f1] - fa.y[c]) / dy;
    F2[c] = F / kp.mass * ast;
  }
}


// ---------------------------------------------------------------------------
// Per-cell moments (cell_update, Main_Functions.cpp:2461-2593) and the
// time-average sums of cell_poly_velocity (Main_Functions.cpp:1918-1968).
// One workgroup per cell; FP64 accumulation; fixed-order tree reductions.
// ---------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void k_moments(const R *__restrict__ P,
                                                 const uint32_t *__restrict__ cnt,
                                                 const uint32_t *__restrict__ start,
                                                 double *__restrict__ out, KParams kp) {
  __shared__ double sh[29 * 4];
  const int c = blockIdx.x;
  const uint32_t N = cnt[c];
  const uint32_t s = start[c];
  double m1[3] = {0.0, 0.0, 0.0};
  for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
    R r = P[s + k];
    m1[0] += (double)r.u1;
    m1[1] += (double)r.u2;
    m1[2] += (double)r.u3;
  }
  block_sum<3>(m1, sh);
  double U0 = 0.0, U1 = 0.0, U2 = 0.0;
  if (N > 0) {
    U0 = m1[0] / (double)N;
    U1 = m1[1] / (double)N;
    U2 = m1[2] / (double)N;
  }
  double a[29];
#pragma unroll
  for (int q = 0; q < 29; ++q) a[q] = 0.0;
  for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
    R r = P[s + k];
    double p1 = (double)r.u1 - U0, p2 = (double)r.u2 - U1, p3 = (double)r.u3 - U2;
    a[26] += p1; a[27] += p2; a[28] += p3;
    double e = p1 * p1 + p2 * p2 + p3 * p3;
    a[0] += p1 * p1; a[1] += p1 * p2; a[2] += p1 * p3;
    a[3] += p2 * p2; a[4] += p2 * p3; a[5] += p3 * p3;
    a[6] += p1 * p1 * p1; a[7] += p1 * p1 * p2; a[8] += p1 * p1 * p3;
    a[9] += p1 * p2 * p2; a[10] += p1 * p2 * p3; a[11] += p1 * p3 * p3;
    a[12] += p2 * p2 * p2; a[13] += p2 * p2 * p3; a[14] += p2 * p3 * p3;
    a[15] += p3 * p3 * p3;
    a[16] += p1 * p1 * e; a[17] += p1 * p2 * e; a[18] += p1 * p3 * e;
    a[19] += p2 * p2 * e; a[20] += p2 * p3 * e; a[21] += p3 * p3 * e;
    a[22] += p1 * e * e; a[23] += p2 * e * e; a[24] += p3 * e * e;
    a[25] += e;
  }
  block_sum<29>(a, sh);
  if (threadIdx.x == 0) {
    double *o = out + (size_t)c * MCPM_NMOM;
    double w = (double)N;
    o[0] = U0; o[1] = U1; o[2] = U2;
    for (int q = 0; q < 26; ++q) o[3 + q] = (N > 0) ? a[q] / w : 0.0;
    double DM2 = (N > 0) ? a[25] / w : 0.0;
    o[29] = kp.mass * DM2 / (3.0 * kp.kb);
    o[30] = w;
    for (int q = 0; q < 3; ++q) o[31 + q] = (N > 0) ? a[26 + q] / w : 0.0;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_sample(const R *__restrict__ P,
                                                const uint32_t *__restrict__ cnt,
                                                const uint32_t *__restrict__ start,
                                                double *__restrict__ acc) {
  __shared__ double sh[25 * 4];
  const int c = blockIdx.x;
  const uint32_t N = cnt[c];
  const uint32_t s = start[c];
  double a[25];
#pragma unroll
  for (int q = 0; q < 25; ++q) a[q] = 0.0;
  for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
    R r = P[s + k];
    double u = (double)r.u1, v = (double)r.u2, w = (double)r.u3;
    double e = u * u + v * v + w * w;
    a[0] += u; a[1] += v; a[2] += w;
    a[3] += u * u; a[4] += u * v; a[5] += u * w; a[6] += v * v; a[7] += v * w; a[8] += w * w;
    // (m,n,l) for m=0..3, n=0..3-m, l=3-m-n (Main_Functions.cpp:1940-1948)
    a[9] += w * w * w;  a[10] += v * w * w; a[11] += v * v * w; a[12] += v * v * v;
    a[13] += u * w * w; a[14] += u * v * w; a[15] += u * v * v;
    a[16] += u * u * w; a[17] += u * u * v; a[18] += u * u * u;
    a[19] += u * u * e; a[20] += u * v * e; a[21] += u * w * e;
    a[22] += v * v * e; a[23] += v * w * e; a[24] += w * w * e;
  }
  block_sum<25>(a, sh);
  if (threadIdx.x == 0) {
    double *o = acc + (size_t)c * MCPM_NSAMP;
    o[0] += (double)N;
    for (int q = 0; q < 25; ++q) o[1 + q] += a[q];
  }
}

// ---------------------------------------------------------------------------
// explicit instantiations: 1D (Rec) and 2D (Rec2) records, FP32 and FP64
// ---------------------------------------------------------------------------
// Collision_CBA_HS (collisions.cpp:2251-2346), one thread per cell.  The
// number of candidates grows whenever an accepted pair is faster than
// Vr_max, so the candidates of a cell run one after the other, as in the
// reference; cells are independent (each touches only its own particles).
// Candidate t of cell c draws (i, j, r, alpha2) from one Philox block keyed
// (t, c); i and j are drawn with replacement (collisions.cpp:2282-2283) and
// r serves both the acceptance and the polar angle (collisions.cpp:2298).
// T is the cell temperature of the moments, n the cell number density.  A
// cell whose bound passes CBA_MAX_CAND (non-finite or absurd velocities)
// stops there and raises the capacity error.
template <typename R>
__global__ __launch_bounds__(64) void k_cba_hs(R *__restrict__ P,
                                               const uint32_t *__restrict__ cnt,
                                               const uint32_t *__restrict__ start,
                                               const double *__restrict__ mom,
                                               const double *__restrict__ nden, KParams kp,
                                               uint32_t step_lo, uint32_t step_hi,
                                               uint32_t *__restrict__ counters,
                                               uint32_t *__restrict__ err) {
  using real = typename R::real_type;
  constexpr double PI = 3.14159265358979323846;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t ncand = 0, nacc = 0;
  const uint32_t N = c < kp.n_cells ? cnt[c] : 0u;
  if (N >= 2) {
    // a single particle has T = 0 and so no candidates
    const uint32_t s = start[c];
    const double T = mom[(size_t)c * MCPM_NMOM + 29];
    const double rho = nden[c] * kp.mass;
    double vmax = __builtin_sqrt(PI * kp.kb * T / kp.mass);
    const double q = 1.0 - kp.b * rho / 4.0;
    const double Y = (1.0 - kp.b * rho / 8.0) / (q * q * q);
    auto mcand = [&](double v) {
      return kp.Fn * (double)N * (double)N * PI * kp.sigma * kp.sigma * v * kp.dt /
             (2.0 * kp.volume) * Y;
    };
    double lim = mcpm_floor(mcand(vmax));
    for (uint32_t t = 0; (double)t < lim; ++t) {
      if (!(lim <= (double)CBA_MAX_CAND)) {
        atomicOr(err, ERRBIT_CAPACITY);
        break;
      }
      const mcpm_u32x4 w = mcpm_philox(t, (uint32_t)c, step_lo,
                                       mcpm_ctr3(MCPM_TAG_CBA, step_hi, 0u), kp.key0, kp.key1);
      const uint32_t i = s + mcpm_uint_below(w.v[0], N);
      const uint32_t j = s + mcpm_uint_below(w.v[1], N);
      const double r = mcpm_u01(w.v[2]);
      ++ncand;
      R a = P[i], bb = P[j];
      const double g0 = (double)a.u1 - (double)bb.u1;
      const double g1 = (double)a.u2 - (double)bb.u2;
      const double g2 = (double)a.u3 - (double)bb.u3;
      const double Vr = __builtin_sqrt(g0 * g0 + g1 * g1 + g2 * g2);
      if (!(Vr / vmax > r)) continue;
      if (Vr - vmax > 1e-14) {
        vmax = Vr;
        lim = mcpm_floor(mcand(vmax));
      }
      const double ct = 1.0 - 2.0 * r;
      const double st = __builtin_sqrt(1.0 - ct * ct);
      double sp, cp;
      mcpm_sincos2pi(mcpm_u01(w.v[3]), &sp, &cp);
      const double v2 = Vr * st * cp, v1 = Vr * st * sp, v0 = Vr * ct;
      const double c0 = 0.5 * ((double)a.u1 + (double)bb.u1);
      const double c1 = 0.5 * ((double)a.u2 + (double)bb.u2);
      const double c2 = 0.5 * ((double)a.u3 + (double)bb.u3);
      a.u1 = (real)(c0 + 0.5 * v0);
      a.u2 = (real)(c1 + 0.5 * v1);
      a.u3 = (real)(c2 + 0.5 * v2);
      bb.u1 = (real)(c0 - 0.5 * v0);
      bb.u2 = (real)(c1 - 0.5 * v1);
      bb.u3 = (real)(c2 - 0.5 * v2);
      cba_displace(a, bb, g0, g1, g2, kp.sigma);
      P[i] = a;
      P[j] = bb;
      ++nacc;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    ncand += __shfl_down(ncand, off);
    nacc += __shfl_down(nacc, off);
  }
  if (lane_id() == 0 && ncand) {
    atomicAdd(&counters[CNT_OPS], ncand);
    atomicAdd((unsigned long long *)&counters[CNT_ACC], (unsigned long long)nacc);
  }
}

#define MCPM_INST_ROUND(R, M, PL, U)                                                       \
  template __global__ void k_round<R, M, PL, U>(                                           \
      int, int, uint32_t, R *, const uint4 *, const uint4 *, uint4 *, const uint32_t *,    \
      uint32_t *, const uint8_t *, uint32_t, uint32_t *, const double *, const double *,   \
      const double *, double *, KParams, uint32_t, uint32_t, FPParams, OpLog,              \
      const StepDev *);
#define MCPM_INST_COLL(R, M)                                                               \
  template __global__ void k_gen<R, M>(const R *, const R::real_type *, const uint32_t *,  \
                                       const uint32_t *,                                   \
                                       const uint32_t *, int, uint32_t, KParams, uint32_t, \
                                       uint32_t, uint32_t, uint32_t, int, const uint32_t *,\
                                       uint32_t, const uint32_t *, int32_t,                \
                                       uint4 *, uint32_t *, uint32_t *, const StepDev *);  \
  MCPM_INST_ROUND(R, M, false, 2)                                                          \
  MCPM_INST_ROUND(R, M, false, 4)                                                          \
  MCPM_INST_ROUND(R, M, true, 2)                                                           \
  MCPM_INST_ROUND(R, M, true, 4)

template __global__ void k_check<COLL_ESMC>(int, int, uint32_t, const uint4 *, const uint4 *,
                                             uint32_t *, uint8_t *, uint32_t, uint32_t *,
                                             uint32_t *, const StepDev *);
template __global__ void k_check<COLL_DSMC>(int, int, uint32_t, const uint4 *, const uint4 *,
                                             uint32_t *, uint8_t *, uint32_t, uint32_t *,
                                             uint32_t *, const StepDev *);
template __global__ void k_check<COLL_ESMC4>(int, int, uint32_t, const uint4 *, const uint4 *,
                                              uint32_t *, uint8_t *, uint32_t, uint32_t *,
                                              uint32_t *, const StepDev *);

#define MCPM_INST_WP(R, S, L)                                                              \
  template __global__ void k_wpush<R, S, L>(const R *, R *, KParams, Kick, const double *,  \
                                            int, uint32_t, uint32_t, const uint32_t *,      \
                                            const int32_t *, const uint32_t *, int32_t *,   \
                                            int32_t *, uint32_t *, const uint32_t *,        \
                                            uint32_t *, uint32_t *, const uint32_t *,       \
                                            uint32_t *, int, R::real_type *,                \
                                            const int32_t *, int32_t *, const StepDev *);
#define MCPM_INST_W(R)                                                                     \
  MCPM_INST_WP(R, false, 0)                                                                \
  MCPM_INST_WP(R, true, 0)                                                                 \
  MCPM_INST_WP(R, false, 1)                                                                \
  MCPM_INST_WP(R, true, 1)                                                                 \
  template __global__ void k_round_tail<R, COLL_ESMC>(int, int, uint32_t, R *, uint4 *,     \
                                                      uint4 *,                              \
                                                      uint32_t *, uint32_t *, uint8_t *,    \
                                                      uint32_t *, const double *,           \
                                                      const double *, const double *,       \
                                                      double *, KParams, FPParams,          \
                                                      StepDev *);                           \
  template __global__ void k_ys_from_records<R>(const R *, int64_t, R::real_type *);

#define MCPM_INST(R)                                                                       \
  template __global__ void k_push<R, false>(const R *, R *, int64_t, int32_t, KParams,     \
                                            Kick, const double *, int, uint32_t, uint32_t, \
                                            uint32_t *, const uint32_t *, uint32_t *, int, \
                                            int, int32_t, uint32_t *, const uint32_t *,    \
                                            const uint32_t *, uint32_t *, R::real_type *,  \
                                            const int32_t *, int32_t *);                   \
  template __global__ void k_push<R, true>(const R *, R *, int64_t, int32_t, KParams,      \
                                           Kick, const double *, int, uint32_t, uint32_t,  \
                                           uint32_t *, const uint32_t *, uint32_t *, int,  \
                                           int, int32_t, uint32_t *, const uint32_t *,     \
                                           const uint32_t *, uint32_t *, R::real_type *,   \
                                           const int32_t *, int32_t *);                    \
  MCPM_INST_COLL(R, COLL_ESMC)                                                             \
  MCPM_INST_COLL(R, COLL_DSMC)                                                             \
  MCPM_INST_COLL(R, COLL_ESMC4)                                                            \
  template __global__ void k_init<R>(R *, InitParams, KParams);                           \
  template __global__ void k_cba_hs<R>(R *, const uint32_t *, const uint32_t *,            \
                                       const double *, const double *, KParams, uint32_t,  \
                                       uint32_t, uint32_t *, uint32_t *);                  \
  template __global__ void k_moments<R>(const R *, const uint32_t *, const uint32_t *,     \
                                        double *, KParams);                                \
  template __global__ void k_sample<R>(const R *, const uint32_t *, const uint32_t *,      \
                                       double *);

MCPM_INST(Rec<float>)
MCPM_INST(Rec<double>)
MCPM_INST_W(Rec<float>)
MCPM_INST_W(Rec<double>)
MCPM_INST(Rec2<float>)
MCPM_INST(Rec2<double>)

// thermostat, the inverted problem's branch (Main_Functions.cpp:1760-1830):
// U = lambda (U - U_cell) + U_cell with lambda = sqrt(T_target / T_cell) for
// the particles of cells whose centre lies below zone_lo (T_lo) or above
// length - zone_hi (T_hi); cells with T_cell = 0 keep their velocities.
// mom: the per-cell moments of these positions (U at 0..2, T at 29).
template <typename R>
__global__ __launch_bounds__(256) void k_thermostat(R *__restrict__ P, int64_t n,
                                                    const double *__restrict__ mom, KParams kp,
                                                    double zone_lo, double T_lo, double zone_hi,
                                                    double T_hi) {
  using real = typename R::real_type;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R r = P[i];
  const int32_t c = cell_of((double)r.y, kp.dy, kp.n_cells);
  const double centre = ((double)c + 0.5) * kp.dy;
  double Tt;
  if (centre < zone_lo) Tt = T_lo;
  else if (centre > kp.length - zone_hi) Tt = T_hi;
  else return;
  const double *m = mom + (size_t)c * MCPM_NMOM;
  if (!(m[29] > 0.0)) return;
  const double lam = __builtin_sqrt(Tt / m[29]);
  r.u1 = (real)(lam * ((double)r.u1 - m[0]) + m[0]);
  r.u2 = (real)(lam * ((double)r.u2 - m[1]) + m[1]);
  r.u3 = (real)(lam * ((double)r.u3 - m[2]) + m[2]);
  P[i] = r;
}
// FP64 sums over the particles for the energy and momentum check: per
// workgroup {sum |U|^2, sum U1, sum U2, sum U3} into part[4 blockIdx.x],
// particle i (grid-stride) read by load(i, &u1, &u2, &u3).
template <typename Load>
__device__ __forceinline__ void energy_partial(Load load, int64_t n, double *__restrict__ part) {
  __shared__ double sh[4][4];
  double e = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a, b, c;
    load(i, &a, &b, &c);
    e += a * a + b * b + c * c;
    m1 += a;
    m2 += b;
    m3 += c;
  }
  double v[4] = {e, m1, m2, m3};
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o);
    if (l == 0) sh[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    part[4 * (size_t)blockIdx.x + k] = (sh[0][k] + sh[1][k]) + (sh[2][k] + sh[3][k]);
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_energy(const R *__restrict__ P, int64_t n,
                                                double *__restrict__ part) {
  energy_partial(
      [&](int64_t i, double *a, double *b, double *c) {
        const R r = P[i];
        *a = (double)r.u1;
        *b = (double)r.u2;
        *c = (double)r.u3;
      },
      n, part);
}

// The same sums over three FP64 velocity arrays (the reference's U1, U2, U3).
__global__ __launch_bounds__(256) void k_energy_soa(const double *__restrict__ u1,
                                                    const double *__restrict__ u2,
                                                    const double *__restrict__ u3, int64_t n,
                                                    double *__restrict__ part) {
  energy_partial(
      [&](int64_t i, double *a, double *b, double *c) {
        *a = u1[i];
        *b = u2[i];
        *c = u3[i];
      },
      n, part);
}

// out[0] = (1/2) mf sum_b part[4b], out[k] = mf sum_b part[4b + k]: the
// partials added in workgroup order, one thread per sum.
__global__ void k_energy_finish(const double *__restrict__ part, int nblk, double mf,
                                double *__restrict__ out) {
  const int k = threadIdx.x;
  if (blockIdx.x != 0 || k >= 4) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[4 * (size_t)b + k];
  out[k] = k == 0 ? 0.5 * mf * s : mf * s;
}
template __global__ void k_energy<Rec<float>>(const Rec<float> *, int64_t, double *);
template __global__ void k_energy<Rec<double>>(const Rec<double> *, int64_t, double *);
template __global__ void k_energy<Rec2<float>>(const Rec2<float> *, int64_t, double *);
template __global__ void k_energy<Rec2<double>>(const Rec2<double> *, int64_t, double *);
template __global__ void k_thermostat<Rec<float>>(Rec<float> *, int64_t, const double *, KParams,
                                                  double, double, double, double);
template __global__ void k_thermostat<Rec<double>>(Rec<double> *, int64_t, const double *,
                                                   KParams, double, double, double, double);

}  // namespace mcpm
