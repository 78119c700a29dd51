// mcpm_longrange.hip — the direct and density-expansion mean-field forces
// of the slab (long_range 1 and 2 of velocity_update) for gfx950.
//
//   k_vlasov_faces  one workgroup per face: the 50 x 50 Gauss-Legendre
//                   Sutherland quadrature of vlasov_taking_the_integral
//   k_vlasov_force  F2 per cell from the face potentials
//   k_kde           one workgroup per cell: Gaussian kernel density n_smooth
//                   and its derivatives dn2, dn3 over cells k-2..k+2
//   k_lr2_force     F2 = -dn2 * prefactor (density expansion)
#include "mcpm_kernels.h"

#pragma clang fp contract(off)

namespace mcpm {

// Density of the cell holding slab position y; boundary cells beyond the
// walls (Main_Functions.cpp:3757-3784), wrapped in a periodic box.
__device__ __forceinline__ double density_at(const double *nden, double y, const KParams &kp) {
  double q = mcpm_floor(y / kp.dy);
  const int nc = kp.n_cells;
  if (kp.boundary == MCPM_BC_PERIODIC) {
    long long k = (long long)q % nc;
    if (k < 0) k += nc;
    return nden[k];
  }
  if (q < 0.0) return nden[0];
  if (q > (double)(nc - 1)) return nden[nc - 1];
  return nden[(int)q];
}

template <int K>
__device__ __forceinline__ void lr_block_sum(double (&v)[K], double *sh) {
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

// tab: rr[50], radial[50] = (sigma/r)^6 r^2 2 pi dr, cos(phi)[50],
// sin(phi)[50], dphi[50] (host-built, vlasov_taking_the_integral:3708-3729).
__global__ __launch_bounds__(256) void k_vlasov_faces(const double *__restrict__ nden,
                                                      const double *__restrict__ tab,
                                                      double *__restrict__ V, KParams kp) {
  __shared__ double sh[4];
  const int f = blockIdx.x;
  const double *rr = tab, *rad = tab + 50, *cph = tab + 100, *sph = tab + 150, *dph = tab + 200;
  double acc[1] = {0.0};
  for (int q = threadIdx.x; q < 2500; q += blockDim.x) {
    int i = q / 50, ii = q % 50;
    double y = (double)f * kp.dy + rr[i] * cph[ii];
    acc[0] += -rad[i] * density_at(nden, y, kp) * sph[ii] * dph[ii];
  }
  lr_block_sum<1>(acc, sh);
  if (threadIdx.x == 0) V[f] = acc[0];
}

__global__ void k_vlasov_force(const double *__restrict__ V, double *__restrict__ F2, KParams kp,
                               double phi) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.n_cells;
  if (c >= nc) return;
  int f1 = kp.boundary == MCPM_BC_PERIODIC ? (c + 1) % nc : c + 1;
  double F = -(V[f1] - V[c]) / kp.dy;  // Main_Functions.cpp:3915-3916
  F2[c] = F / kp.mass * phi;
}

template <typename real>
__global__ __launch_bounds__(256) void k_kde(const Rec<real> *__restrict__ P,
                                             const uint32_t *__restrict__ cnt,
                                             const uint32_t *__restrict__ start,
                                             double *__restrict__ ns, double *__restrict__ d2,
                                             double *__restrict__ d3, KParams kp) {
  __shared__ double sh[3 * 4];
  const int k = blockIdx.x;
  const int nc = kp.n_cells;
  const double r_cut = 2.0 * kp.dy, h = r_cut / 4.0;
  const double sqrt_pi = __builtin_sqrt(2.0 * 3.14159265358979323846);
  const int search = (int)mcpm_floor(r_cut / kp.dy);
  const double mb = kp.Fn;
  const double ih2 = 1.0 / h / h;
  const double s2h2 = ih2 * ih2;
  const double r1 = (k + 0.5) * kp.dy;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int kk = k - search; kk <= k + search; ++kk) {
    int kkk;
    double shift;
    if (kp.boundary == MCPM_BC_PERIODIC) {
      kkk = ((kk % nc) + nc) % nc;
      shift = (double)(kk - kkk) * kp.dy;
    } else {
      kkk = kk < 0 ? 0 : (kk > nc - 1 ? nc - 1 : kk);
      shift = (kk - kkk) * kp.dy;
    }
    const uint32_t n = cnt[kkk], s = start[kkk];
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
      double r2 = (double)P[s + i].y + shift;
      double r1r2 = r2 - r1;
      double rr = __builtin_fabs(r2 - r1);
      if (rr < r_cut) {
        double ex = mcpm_exp(-rr * rr / h / h / 2.0);
        acc[0] += mb / sqrt_pi / h * ex;
        double W = 1.0 / sqrt_pi / h * ex;
        acc[1] += mb * W * r1r2 * (-1.0 / h / h);
        acc[2] += mb * W * s2h2 * (3.0 * r1r2 - 1.0 / h / h * r1r2 * r1r2 * r1r2);
      }
    }
  }
  lr_block_sum<3>(acc, sh);
  if (threadIdx.x == 0) {
    ns[k] = acc[0];
    d2[k] = acc[1];
    d3[k] = acc[2];
  }
}

__global__ void k_lr2_force(const double *__restrict__ d2, double *__restrict__ F2, KParams kp,
                            double pref) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.n_cells;
  if (c >= nc) return;
  bool wall_cell = kp.boundary != MCPM_BC_PERIODIC && (c == 0 || c == nc - 1);
  F2[c] = wall_cell ? 0.0 : -(d2[c] * pref);  // Main_Functions.cpp:1648-1651
}

template __global__ void k_kde<float>(const Rec<float> *, const uint32_t *, const uint32_t *,
                                      double *, double *, double *, KParams);
template __global__ void k_kde<double>(const Rec<double> *, const uint32_t *, const uint32_t *,
                                       double *, double *, double *, KParams);

}  // namespace mcpm
