// mcpm_fem.hip — the plane's mean field (the reference's FEM.cpp):
//   k_quadrants      particles per cell quadrant (the deposit's case split)
//   k_nodes          nodal density n_n (cell_update, Main_Functions.cpp:2268-2364)
//   k_node_integral  Gauss-Legendre sums around a node: the Yukawa Green's
//                    function on the walls (Compute_SP_using_Green_Function,
//                    FEM.cpp:196-247) and the direct Sutherland integral
//                    (vlasov_taking_the_integral_2D, FEM.cpp:157-195)
//   k_fem_cg         Q1 stiffness + lambda^2 V mass system (compute_SP_2D,
//                    FEM.cpp:249-458) by conjugate gradients, one workgroup
//   k_cgm_*          the same solve over a grid of workgroups (large planes)
//   k_fem_core       core subtraction with its 5-point correction (FEM.cpp:497-554)
// The kick from the nodal field is part of the push (mcpm_kernels.hip).
#include "mcpm_geom.h"
#include "mcpm_kernels.h"

#pragma clang fp contract(off)

namespace mcpm {

namespace {

// Node coordinate along one axis as make_cells builds it (myIO.cpp:1455-1479):
// node i < n is corner 0 of cell i (centre - width/2), node n is the far
// corner of cell n-1 (centre + width/2).
__device__ __forceinline__ double node_coord(int i, int n, double h) {
  const int ii = i < n ? i : n - 1;
  const double lo = (double)ii * h, hi = (double)(ii + 1) * h;
  const double cc = 0.5 * (lo + hi), w = hi - lo;
  return i < n ? cc - w / 2.0 : cc + w / 2.0;
}

__device__ __forceinline__ int wrap_index(int i, int n) {
  return i < 0 ? i + n : (i >= n ? i - n : i);
}

// Sum of one double over a workgroup of up to 1024 threads; every thread
// gets the total.  The order is fixed, so results are reproducible.
__device__ double wg_sum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int k = 0; k < nw; ++k) t += sh[k];
  return t;
}

// Boundary node b of the walled plane: the bottom row, the top row, then the
// left and right columns without their corners.
__device__ __forceinline__ void boundary_node(int b, int nx, int ny, int *i, int *j) {
  const int row = nx + 1;
  if (b < row) { *i = b; *j = 0; return; }
  b -= row;
  if (b < row) { *i = b; *j = ny; return; }
  b -= row;
  if (b < ny - 1) { *i = 0; *j = b + 1; return; }
  b -= ny - 1;
  *i = nx;
  *j = b + 1;
}

__device__ __forceinline__ bool fem_unknown(int i, int j, const FemParams &f) {
  if (f.per_x) {
    if (i >= f.nx) return false;
  } else if (i == 0 || i == f.nx) {
    return false;
  }
  if (f.per_y) {
    if (j >= f.ny) return false;
  } else if (j == 0 || j == f.ny) {
    return false;
  }
  return true;
}

// 9-point stencils of the assembled Q1 matrices for square cells
// (compute_intdudvdx / compute_intuvdx, FEM.cpp:51-112): stiffness 8/3 at
// the node, -1/3 at the edge and diagonal neighbours; mass 4/9, 1/9, 1/36.
struct Stencil {
  double c, e, d;
};

__device__ __forceinline__ double apply9(const double *v, int i, int j, const FemParams &f,
                                         const Stencil &s) {
  const int w = f.nx + 1;
  const int il = f.per_x ? wrap_index(i - 1, f.nx) : i - 1;
  const int ir = f.per_x ? wrap_index(i + 1, f.nx) : i + 1;
  const int jd = f.per_y ? wrap_index(j - 1, f.ny) : j - 1;
  const int ju = f.per_y ? wrap_index(j + 1, f.ny) : j + 1;
  const double edge = v[j * w + il] + v[j * w + ir] + v[jd * w + i] + v[ju * w + i];
  const double diag = v[jd * w + il] + v[jd * w + ir] + v[ju * w + il] + v[ju * w + ir];
  return s.c * v[j * w + i] + s.e * edge + s.d * diag;
}

}  // namespace

// ---------------------------------------------------------------------------
// Deposit: which quarter of its cell each particle sits in.  Quarter q =
// (x >= centre_x) + 2 (y >= centre_y) touches corner q of the cell, the
// node the reference credits first (Main_Functions.cpp:2268-2342).
template <typename R>
__global__ __launch_bounds__(256) void k_quadrants(const R *__restrict__ P,
                                                   const uint32_t *__restrict__ cnt,
                                                   const uint32_t *__restrict__ start,
                                                   uint32_t *__restrict__ q4, KParams kp) {
  __shared__ uint32_t sq[4];
  const int c = blockIdx.x;
  const int cx = c % kp.nx, cy = c / kp.nx;
  const double ccx = 0.5 * ((double)cx * kp.dx + (double)(cx + 1) * kp.dx);
  const double ccy = 0.5 * ((double)cy * kp.dy + (double)(cy + 1) * kp.dy);
  if (threadIdx.x < 4) sq[threadIdx.x] = 0;
  __syncthreads();
  uint32_t n[4] = {0, 0, 0, 0};
  const uint32_t N = cnt[c], s = start[c];
  for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
    R r = P[s + k];
    const int q = ((double)r.x < ccx ? 0 : 1) + ((double)r.y < ccy ? 0 : 2);
    n[q]++;
  }
  for (int q = 0; q < 4; ++q) {
    uint32_t v = n[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&sq[q], v);
  }
  __syncthreads();
  if (threadIdx.x < 4) q4[4 * c + threadIdx.x] = sq[threadIdx.x];
}

// The same counts from the sort keys of a two-digit sort, whose top bits
// carry each particle's quarter (key_quarter_bits): one wavefront per cell
// reads 4 bytes per particle instead of the record.
__global__ __launch_bounds__(256) void k_quadrants_keys(const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ cnt,
                                                        const uint32_t *__restrict__ start,
                                                        uint32_t *__restrict__ q4, int nc) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (c >= nc) return;
  uint32_t n[4] = {0, 0, 0, 0};
  const uint32_t N = cnt[c], s = start[c];
  for (uint32_t k = l; k < N; k += 64) {
    const uint32_t q = keys[s + k] >> KEY_QUARTER_SHIFT;
    n[0] += q == 0u;
    n[1] += q == 1u;
    n[2] += q == 2u;
    n[3] += q == 3u;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t v = n[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (l == q) q4[4 * c + q] = v;
  }
}

// Node density: each cell around node (i,j) contributes its quarter that
// touches the node; Fn / V per particle, times 2 per wall the node lies on
// (Main_Functions.cpp:2344-2363).  A periodic axis wraps instead.
__global__ void k_nodes(const uint32_t *__restrict__ q4, double *__restrict__ n_node,
                        FemParams f) {
  const int w = f.nx + 1;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= w * (f.ny + 1)) return;
  const int i = id % w, j = id / w;
  uint32_t S = 0;
  for (int b = 0; b < 2; ++b) {
    int cy = j - 1 + b;
    if (f.per_y) cy = wrap_index(cy, f.ny);
    else if (cy < 0 || cy >= f.ny) continue;
    for (int a = 0; a < 2; ++a) {
      int cx = i - 1 + a;
      if (f.per_x) cx = wrap_index(cx, f.nx);
      else if (cx < 0 || cx >= f.nx) continue;
      S += q4[4 * (cy * f.nx + cx) + (1 - a) + 2 * (1 - b)];
    }
  }
  double wgt = 1.0;
  if (!f.per_x && (i == 0 || i == f.nx)) wgt *= 2.0;
  if (!f.per_y && (j == 0 || j == f.ny)) wgt *= 2.0;
  n_node[id] = (double)S * f.fnv * wgt;
}

// Gauss-Legendre sum around one node per workgroup:
//   V = sum_{r, azimuth, polar} rad(r) n(cell(x + r cos(az) sin(pol),
//                                        y + r sin(az) sin(pol))) w_az sin(pol) w_pol
// with the cell index clamped to the mesh on a walled axis (FEM.cpp:172-189)
// and wrapped on a periodic one.  boundary_only: the wall nodes of the
// screened-Poisson Dirichlet rows; otherwise every node.
__global__ __launch_bounds__(256) void k_node_integral(const double *__restrict__ nden,
                                                       const double *__restrict__ tab,
                                                       double *__restrict__ pot, KParams kp,
                                                       FemParams f, int boundary_only) {
  __shared__ double sh[16];
  __shared__ double st[QT_N * 128];
  const int nq = f.nq;
  for (int k = threadIdx.x; k < QT_N * nq; k += blockDim.x) st[k] = tab[k];
  __syncthreads();
  int i, j;
  if (boundary_only) {
    boundary_node(blockIdx.x, f.nx, f.ny, &i, &j);
  } else {
    i = blockIdx.x % (f.nx + 1);
    j = blockIdx.x / (f.nx + 1);
  }
  const double xx = node_coord(i, f.nx, kp.dx);
  const double yy = node_coord(j, f.ny, kp.dy);
  const double *R_ = st + QT_R * nq, *RAD = st + QT_RAD * nq, *CPH = st + QT_CPH * nq,
               *SPH = st + QT_SPH * nq, *DPH = st + QT_DPH * nq, *STH = st + QT_STH * nq,
               *DTH = st + QT_DTH * nq;
  double V = 0.0;
  for (int pr = threadIdx.x; pr < nq * nq; pr += blockDim.x) {
    const int a = pr / nq, b = pr % nq;
    const double rc = R_[a] * CPH[b], rs = R_[a] * SPH[b];
    const double rad = RAD[a], dph = DPH[b];
    double part = 0.0;
    for (int c = 0; c < nq; ++c) {
      double x = xx + rc * STH[c];
      double y = yy + rs * STH[c];
      int cx, cy;
      if (f.per_x) {
        cx = cell_of(wrap_coord(x, kp.length_x), kp.dx, f.nx);
      } else {
        double q = mcpm_floor(x / kp.dx);
        cx = q < 0.0 ? 0 : (q > (double)(f.nx - 1) ? f.nx - 1 : (int)q);
      }
      if (f.per_y) {
        cy = cell_of(wrap_coord(y, kp.length), kp.dy, f.ny);
      } else {
        double q = mcpm_floor(y / kp.dy);
        cy = q < 0.0 ? 0 : (q > (double)(f.ny - 1) ? f.ny - 1 : (int)q);
      }
      const double n = nden[cy * f.nx + cx];
      part += rad * n * dph * STH[c] * DTH[c];
    }
    V += part;
  }
  V = wg_sum(V, sh);
  if (threadIdx.x == 0) pot[j * (f.nx + 1) + i] = V;
}

// Conjugate gradients on the unknown nodes of
//   (K + lambda^2 V M) x = V M n_n,  x = wall values on walled boundaries,
// the system compute_SP_2D assembles (FEM.cpp:283-305) and PARDISO factors
// (FEM.cpp:450-458).  The known wall values move to the right-hand side.
// One workgroup of 1024 threads runs every iteration; pot holds the wall
// values on entry and the solution on exit (duplicated on periodic seams).
__global__ __launch_bounds__(1024) void k_fem_cg(const double *__restrict__ n_node,
                                                 double *__restrict__ pot,
                                                 double *__restrict__ r, double *__restrict__ p,
                                                 double *__restrict__ q, FemParams f,
                                                 uint32_t *iters) {
  __shared__ double sh[16];
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  const Stencil A{8.0 / 3.0 + 4.0 * f.lam2V / 9.0, -1.0 / 3.0 + f.lam2V / 9.0,
                  -1.0 / 3.0 + f.lam2V / 36.0};
  const Stencil M{4.0 / 9.0 * f.volume, 1.0 / 9.0 * f.volume, 1.0 / 36.0 * f.volume};
  for (int id = threadIdx.x; id < nn; id += blockDim.x) {
    const int i = id % w, j = id / w;
    if (fem_unknown(i, j, f)) pot[id] = 0.0;
  }
  __syncthreads();
  double rr = 0.0, bb = 0.0;
  for (int id = threadIdx.x; id < nn; id += blockDim.x) {
    const int i = id % w, j = id / w;
    double rv = 0.0, bv = 0.0;
    if (fem_unknown(i, j, f)) {
      bv = apply9(n_node, i, j, f, M);
      rv = bv - apply9(pot, i, j, f, A);
    }
    r[id] = rv;
    p[id] = rv;
    rr += rv * rv;
    bb += bv * bv;
  }
  rr = wg_sum(rr, sh);
  bb = wg_sum(bb, sh);
  const double stop = f.tol2 * bb;
  int it = 0;
  while (it < f.maxit && rr > stop) {
    __syncthreads();  // p complete
    double pq = 0.0;
    for (int id = threadIdx.x; id < nn; id += blockDim.x) {
      const int i = id % w, j = id / w;
      double qv = 0.0;
      if (fem_unknown(i, j, f)) qv = apply9(p, i, j, f, A);
      q[id] = qv;
      pq += p[id] * qv;
    }
    pq = wg_sum(pq, sh);
    const double alpha = rr / pq;
    double rn = 0.0;
    for (int id = threadIdx.x; id < nn; id += blockDim.x) {
      const int i = id % w, j = id / w;
      if (fem_unknown(i, j, f)) {
        pot[id] += alpha * p[id];
        const double rv = r[id] - alpha * q[id];
        r[id] = rv;
        rn += rv * rv;
      }
    }
    rn = wg_sum(rn, sh);
    const double beta = rn / rr;
    rr = rn;
    __syncthreads();  // every thread has read p for this iteration
    for (int id = threadIdx.x; id < nn; id += blockDim.x) p[id] = r[id] + beta * p[id];
    ++it;
  }
  __syncthreads();
  // periodic seams: node nx repeats node 0, row ny repeats row 0
  for (int id = threadIdx.x; id < nn; id += blockDim.x) {
    const int i = id % w, j = id / w;
    if ((f.per_x && i == f.nx) || (f.per_y && j == f.ny)) {
      const int i0 = (f.per_x && i == f.nx) ? 0 : i;
      const int j0 = (f.per_y && j == f.ny) ? 0 : j;
      pot[id] = pot[j0 * w + i0];
    }
  }
  if (threadIdx.x == 0) *iters = (uint32_t)it;
}

// The same conjugate gradients over a grid of CG_GRID workgroups for large
// planes: three launches per iteration (q = A p with partial p.q; x and r
// with partial r.r; the new direction), every workgroup reducing the
// partials of the previous launch in one fixed order, so all of them see the
// same scalars and the solve is reproducible.  Iteration k reads rr from
// st[k & 1] and writes the next one to st[(k + 1) & 1]; flag[0] counts
// iterations, flag[1] = 1 once converged (later launches return at once).
namespace {

__device__ double grid_partials_sum(const double *part, int G, double *sh) {
  double v = 0.0;
  for (int b = threadIdx.x; b < G; b += blockDim.x) v += part[b];
  return wg_sum(v, sh);
}

}  // namespace

__global__ __launch_bounds__(256) void k_cgm_zero(double *__restrict__ pot, FemParams f) {
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x)
    if (fem_unknown(id % w, id / w, f)) pot[id] = 0.0;
}

__global__ __launch_bounds__(256) void k_cgm_start(const double *__restrict__ n_node,
                                                   const double *__restrict__ pot,
                                                   double *__restrict__ r,
                                                   double *__restrict__ p, FemParams f,
                                                   double *__restrict__ part) {
  __shared__ double sh[16];
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  const Stencil A{8.0 / 3.0 + 4.0 * f.lam2V / 9.0, -1.0 / 3.0 + f.lam2V / 9.0,
                  -1.0 / 3.0 + f.lam2V / 36.0};
  const Stencil M{4.0 / 9.0 * f.volume, 1.0 / 9.0 * f.volume, 1.0 / 36.0 * f.volume};
  double rr = 0.0, bb = 0.0;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x) {
    const int i = id % w, j = id / w;
    double rv = 0.0, bv = 0.0;
    if (fem_unknown(i, j, f)) {
      bv = apply9(n_node, i, j, f, M);
      rv = bv - apply9(pot, i, j, f, A);
    }
    r[id] = rv;
    p[id] = rv;
    rr += rv * rv;
    bb += bv * bv;
  }
  rr = wg_sum(rr, sh);
  bb = wg_sum(bb, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(256) void k_cgm_start2(double *__restrict__ st,
                                                    const double *__restrict__ part, int G,
                                                    FemParams f, uint32_t *__restrict__ flag) {
  __shared__ double sh[16];
  const double rr = grid_partials_sum(part, G, sh);
  const double bb = grid_partials_sum(part + G, G, sh);
  if (threadIdx.x == 0) {
    st[0] = rr;
    st[2] = f.tol2 * bb;
    st[3] = bb;
    flag[0] = 0u;
    flag[1] = (rr > f.tol2 * bb && f.maxit > 0) ? 0u : 1u;
  }
}

__global__ __launch_bounds__(256) void k_cgm_aq(const double *__restrict__ p,
                                                double *__restrict__ q, FemParams f,
                                                double *__restrict__ part,
                                                const uint32_t *__restrict__ flag) {
  __shared__ double sh[16];
  if (flag[1]) return;
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  const Stencil A{8.0 / 3.0 + 4.0 * f.lam2V / 9.0, -1.0 / 3.0 + f.lam2V / 9.0,
                  -1.0 / 3.0 + f.lam2V / 36.0};
  double pq = 0.0;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x) {
    const int i = id % w, j = id / w;
    double qv = 0.0;
    if (fem_unknown(i, j, f)) qv = apply9(p, i, j, f, A);
    q[id] = qv;
    pq += p[id] * qv;
  }
  pq = wg_sum(pq, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = pq;
}

__global__ __launch_bounds__(256) void k_cgm_xr(double *__restrict__ pot, double *__restrict__ r,
                                                const double *__restrict__ p,
                                                const double *__restrict__ q, FemParams f,
                                                double *__restrict__ part,
                                                const double *__restrict__ st, int k,
                                                const uint32_t *__restrict__ flag) {
  __shared__ double sh[16];
  if (flag[1]) return;
  const int G = gridDim.x;
  const double alpha = st[k & 1] / grid_partials_sum(part, G, sh);
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  double rn = 0.0;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x) {
    if (fem_unknown(id % w, id / w, f)) {
      pot[id] += alpha * p[id];
      const double rv = r[id] - alpha * q[id];
      r[id] = rv;
      rn += rv * rv;
    }
  }
  rn = wg_sum(rn, sh);
  if (threadIdx.x == 0) part[G + blockIdx.x] = rn;
}

__global__ __launch_bounds__(256) void k_cgm_p(double *__restrict__ p,
                                               const double *__restrict__ r, FemParams f,
                                               const double *__restrict__ part,
                                               double *__restrict__ st, int k,
                                               uint32_t *__restrict__ flag) {
  __shared__ double sh[16];
  if (flag[1]) return;
  const int G = gridDim.x;
  const double rn = grid_partials_sum(part + G, G, sh);
  const double beta = rn / st[k & 1];
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x)
    p[id] = r[id] + beta * p[id];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[(k + 1) & 1] = rn;
    flag[0] = (uint32_t)(k + 1);
    if (!(rn > st[2]) || k + 1 >= f.maxit) flag[1] = 1u;
  }
}

// Periodic seams after the solve: node nx repeats node 0, row ny row 0.
__global__ __launch_bounds__(256) void k_cgm_seams(double *__restrict__ pot, FemParams f) {
  const int w = f.nx + 1, nn = w * (f.ny + 1);
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < nn; id += gridDim.x * blockDim.x) {
    const int i = id % w, j = id / w;
    if ((f.per_x && i == f.nx) || (f.per_y && j == f.ny)) {
      const int i0 = (f.per_x && i == f.nx) ? 0 : i;
      const int j0 = (f.per_y && j == f.ny) ? 0 : j;
      pot[id] = pot[j0 * w + i0];
    }
  }
}

// ---------------------------------------------------------------------------
// Periodic planes of power-of-two size: the Q1 system (K + lambda^2 V M) x =
// V M n is diagonal in the plane's DFT, with
//   x_hat = V Mu / (K + lam2V Mu) n_hat,  Mu = (2 + cx)(2 + cy) / 9,
//   K = 8/3 - (2/3)(cx + cy) - (4/3) cx cy,  c = cos(2 pi k / N),
// the symbols of the 9-point stencils the conjugate-gradient kernels apply.
// Three launches: FFT of every row (from the nodal density), FFT of every
// column, multiply, inverse FFT of the column (one kernel), inverse FFT of
// every row (into the nodal field, seams included).  Radix-2, in LDS, FP64.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place radix-2 FFT of x[0..n) in LDS (input already in bit-reversed
// order); tw[k] = exp(-2 pi i k / n), k < n/2; inverse uses the conjugates.
// Threads t0, t0 + nt, ... of the workgroup share the butterflies.
__device__ void lds_fft(double2 *x, int n, const double2 *__restrict__ tw, bool inverse, int t0,
                        int nt) {
  for (int len = 2; len <= n; len <<= 1) {
    const int half = len >> 1, step = n / len;
    for (int j = t0; j < n / 2; j += nt) {
      const int pos = j % half, i0 = (j / half) * len + pos, i1 = i0 + half;
      double2 w = tw[pos * step];
      if (inverse) w.y = -w.y;
      const double2 a = x[i0], b = cmul(w, x[i1]);
      x[i0] = make_double2(a.x + b.x, a.y + b.y);
      x[i1] = make_double2(a.x - b.x, a.y - b.y);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int bitrev(int i, int lg) {
  return (int)(__builtin_bitreverse32((uint32_t)i) >> (32 - lg));
}

}  // namespace

// Row pass.  Forward: rows of the nodal density (seams skipped) -> buf.
// Inverse: buf rows -> real part / (N0 N1) into the nodal field, seams copied.
template <bool INVERSE>
__global__ __launch_bounds__(256) void k_fft_rows(const double *__restrict__ n_node,
                                                  double2 *__restrict__ buf,
                                                  double *__restrict__ pot,
                                                  const double2 *__restrict__ twx, int lgx,
                                                  FemParams f) {
  extern __shared__ double2 fx[];
  const int j = blockIdx.x, n0 = f.nx, w = f.nx + 1;
  for (int i = threadIdx.x; i < n0; i += blockDim.x) {
    const double2 v = INVERSE ? buf[(size_t)j * n0 + i]
                              : make_double2(n_node[(size_t)j * w + i], 0.0);
    fx[bitrev(i, lgx)] = v;
  }
  __syncthreads();
  lds_fft(fx, n0, twx, INVERSE, threadIdx.x, blockDim.x);
  if (!INVERSE) {
    for (int i = threadIdx.x; i < n0; i += blockDim.x) buf[(size_t)j * n0 + i] = fx[i];
  } else {
    const double inv = 1.0 / ((double)f.nx * (double)f.ny);
    for (int i = threadIdx.x; i <= n0; i += blockDim.x) {
      const double v = fx[i == n0 ? 0 : i].x * inv;
      pot[(size_t)j * w + i] = v;
      if (j == 0) pot[(size_t)f.ny * w + i] = v;  // seam row ny repeats row 0
    }
  }
}

// Column pass: FFT of `cols` columns per workgroup, multiply by the symbol,
// inverse FFT, back into buf.  cosx/cosy: cos(2 pi k / N) per axis.
__global__ __launch_bounds__(256) void k_fft_cols(double2 *__restrict__ buf,
                                                  const double2 *__restrict__ twy, int lgy,
                                                  const double *__restrict__ cosx,
                                                  const double *__restrict__ cosy, int cols,
                                                  FemParams f) {
  extern __shared__ double2 fy[];
  const int n0 = f.nx, n1 = f.ny;
  const int c0 = blockIdx.x * cols;
  const int per = blockDim.x / cols;  // threads per column
  const int col = threadIdx.x / per, t0 = threadIdx.x % per;
  double2 *x = fy + (size_t)col * n1;
  const int i = c0 + col;
  for (int jj = t0; jj < n1; jj += per) x[bitrev(jj, lgy)] = buf[(size_t)jj * n0 + i];
  __syncthreads();
  lds_fft(x, n1, twy, false, t0, per);
  const double cx = cosx[i];
  for (int jj = t0; jj < n1; jj += per) {
    const double cy = cosy[jj];
    const double mu = (2.0 + cx) * (2.0 + cy) / 9.0;
    const double K = 8.0 / 3.0 - 2.0 / 3.0 * (cx + cy) - 4.0 / 3.0 * cx * cy;
    const double s = f.volume * mu / (K + f.lam2V * mu);
    x[jj] = make_double2(x[jj].x * s, x[jj].y * s);
  }
  __syncthreads();
  // inverse: reorder to bit-reversed positions first
  for (int jj = t0; jj < n1; jj += per) {
    const int r = bitrev(jj, lgy);
    if (r > jj) {
      const double2 a = x[jj];
      x[jj] = x[r];
      x[r] = a;
    }
  }
  __syncthreads();
  lds_fft(x, n1, twy, true, t0, per);
  for (int jj = t0; jj < n1; jj += per) buf[(size_t)jj * n0 + i] = x[jj];
}

template __global__ void k_fft_rows<false>(const double *, double2 *, double *, const double2 *,
                                           int, FemParams);
template __global__ void k_fft_rows<true>(const double *, double2 *, double *, const double2 *,
                                          int, FemParams);

// Core subtraction x -= n dummy + (Laplacian-like sum) dummy2
// (FEM.cpp:497-554).  On walled planes the neighbour set is the
// reference's: node (i,j) is corner k of its owner cell (ii,jj), and the
// four terms read corner kpm[l] of cell (ii+ipm[l], jj+jpm[l]), which on
// the last row and column are not the geometric neighbours; this is kept.
// Periodic planes use the wrapped 5-point neighbours.
__global__ void k_fem_core(const double *__restrict__ n_node, double *__restrict__ pot,
                           FemParams f) {
  const int w = f.nx + 1;
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= w * (f.ny + 1)) return;
  const int i = id % w, j = id / w;
  double d2, d3;
  if (f.per_x || f.per_y) {
    const int i0 = (f.per_x && i == f.nx) ? 0 : i, j0 = (f.per_y && j == f.ny) ? 0 : j;
    const int il = wrap_index(i0 - 1, f.nx), ir = wrap_index(i0 + 1, f.nx);
    const int jd = wrap_index(j0 - 1, f.ny), ju = wrap_index(j0 + 1, f.ny);
    const double nc = n_node[j0 * w + i0];
    d2 = nc * f.core0;
    d3 = -4.0 * nc * f.core2;
    d3 += n_node[j0 * w + ir] * f.core2;
    d3 += n_node[j0 * w + il] * f.core2;
    d3 += n_node[ju * w + i0] * f.core2;
    d3 += n_node[jd * w + i0] * f.core2;
  } else {
    int ipm[4] = {0, -1, 0, 0}, jpm[4] = {0, 0, 0, -1}, kpm[4] = {1, 0, 2, 0};
    int ii = i, jj = j, k = 0;
    if (i == f.nx) { ii = i - 1; k = 1; kpm[0] = k; }
    if (j == f.ny) { jj = j - 1; k = (i == f.nx) ? 3 : 2; kpm[2] = k; }
    if (i == f.nx && j == f.ny) kpm[0] = 3;
    if (i == 0) { ipm[1] = 0; jpm[1] = 0; kpm[1] = 0; }
    if (j == 0) { ipm[3] = 0; jpm[3] = 0; kpm[3] = 0; }
    // corner q of cell (cx, cy) is node (cx + (q & 1), cy + (q >> 1))
    const double nk = n_node[(jj + (k >> 1)) * w + ii + (k & 1)];
    d2 = nk * f.core0;
    d3 = -4.0 * nk * f.core2;
    for (int l = 0; l < 4; ++l) {
      const int cx = ii + ipm[l], cy = jj + jpm[l];
      d3 += n_node[(cy + (kpm[l] >> 1)) * w + cx + (kpm[l] & 1)] * f.core2;
    }
  }
  pot[id] += -(d2 + d3);
}

template __global__ void k_quadrants<Rec2<float>>(const Rec2<float> *, const uint32_t *,
                                                  const uint32_t *, uint32_t *, KParams);
template __global__ void k_quadrants<Rec2<double>>(const Rec2<double> *, const uint32_t *,
                                                   const uint32_t *, uint32_t *, KParams);

}  // namespace mcpm
