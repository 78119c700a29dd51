// mcpm_fp.hip — Fokker-Planck model kernels for gfx950.
//
//   k_fp_closures  one wavefront per cell: Lambda, gamma_st, alpha/beta
//                  (12x12), c/gamma (9x9) by Gauss-Jordan with scaled
//                  pivoting (the elimination spread over the lanes), phi,
//                  and the per-cell constants of the position drift
//   k_fp_velocity  one workgroup per cell: Philox Gaussian noise normalised
//                  per cell, exact OU step with the closure drift, per-cell
//                  kinetic-energy rescale.  The noise is regenerated from its
//                  counter in each pass instead of being stored.
#include "mcpm_kernels.h"

#pragma clang fp contract(off)

namespace mcpm {

// Gauss_Jordan (Math_Func.cpp:165-232) on an n x n row-major matrix in LDS,
// by the one wavefront of the workgroup.  The elimination's element updates
// are independent of each other and run on the lanes; the pivot choice and
// the back substitution run in the reference's sequence in every lane (the
// pivot) or in lane 0 (the substitution), so each value comes out of the
// same operations as in a serial elimination.  x is left in LDS.
template <int n>
__device__ void wave_gauss_jordan(double *A, double *rhs, double *x, double *zb, double *fb) {
  const int lane = (int)lane_id();
  for (int i = 0; i < n - 1; ++i) {
    // each row's scaled pivot candidate |A[j][i]| / max_k |A[j][k]|
    if (lane >= i && lane < n) {
      const double *row = A + lane * n;
      int l = 0;
      double b = __builtin_fabs(row[0]);
      for (int k = 0; k < n; ++k)
        if (__builtin_fabs(b) < __builtin_fabs(row[k])) {
          b = __builtin_fabs(row[k]);
          l = k;
        }
      zb[lane] = __builtin_fabs(row[i]) / __builtin_fabs(row[l]);
    }
    __syncthreads();
    int best = 0;
    double bz = -1.0;
    for (int j = i; j < n; ++j) {
      const double z = zb[j];
      if (j == i || z > bz) {
        bz = z;
        best = j - i;
      }
    }
    const int k = i + best;
    if (k != i) {
      if (lane < n) {
        const double t = A[i * n + lane];
        A[i * n + lane] = A[k * n + lane];
        A[k * n + lane] = t;
      }
      if (lane == 0) {
        const double t = rhs[i];
        rhs[i] = rhs[k];
        rhs[k] = t;
      }
    }
    __syncthreads();
    if (lane > i && lane < n) fb[lane] = (1.0 / A[i * n + i]) * A[lane * n + i];
    __syncthreads();
    for (int e = lane; e < (n - 1 - i) * n; e += 64) {
      const int j = i + 1 + e / n, kk = e % n;
      A[j * n + kk] = A[j * n + kk] - fb[j] * A[i * n + kk];
    }
    if (lane > i && lane < n) rhs[lane] = rhs[lane] - fb[lane] * rhs[i];
    __syncthreads();
  }
  if (lane == 0)
    for (int i = n - 1; i >= 0; --i) {
      double a = 0.0;
      for (int j = i + 1; j < n; ++j) a = a + A[i * n + j] * x[j];
      x[i] = (rhs[i] - a) / A[i * n + i];
    }
  __syncthreads();
}

// Closures of one cell.  Layout of clos: 0 Lambda, 1-6 c_ij, 7-9 gamma,
// 10-18 alpha, 19-21 beta, 22 gamma_st, 23-25 phi, 26 div U, 27 tau.
// drift (DRIFT_N per cell): U0 U1 U2 T alpha3 alpha4 alpha5 beta1 gamma_st
// Q1 phi1 valid.
__global__ __launch_bounds__(64) void k_fp_closures(const double *__restrict__ mom,
                                                   const uint32_t *__restrict__ cnt,
                                                   const double *__restrict__ nden,
                                                   const double *__restrict__ yc,
                                                   double *__restrict__ clos,
                                                   double *__restrict__ drift, KParams kp,
                                                   FPParams fpp, int64_t step, uint32_t *err) {
  // one wavefront per cell: every lane forms the cell's scalars, the lanes
  // share the two eliminations, lane 0 writes the cell's results
  __shared__ double sA[144], sR[12], sX[12], sZ[12], sF[12];
  const int c = blockIdx.x;
  const int nc = kp.n_cells;
  if (c >= nc) return;
  const bool w0 = lane_id() == 0;
  const double *m = mom + (size_t)c * MCPM_NMOM;
  double *cl = clos + (size_t)c * MCPM_NCLOS;
  double *dr = drift + (size_t)c * DRIFT_N;
  // the cell's results, kept by lane 0 and written at the end
  double out[MCPM_NCLOS];
  for (int q = 0; q < MCPM_NCLOS; ++q) out[q] = 0.0;
  auto finish = [&](bool with_drift, const double *d) {
    if (!w0) return;
    for (int q = 0; q < MCPM_NCLOS; ++q) cl[q] = out[q];
    for (int q = 0; q < DRIFT_N; ++q) dr[q] = with_drift ? d[q] : 0.0;
  };
  const double *P = m + 3, *M3 = m + 9, *M4 = m + 19, *M5 = m + 25;
  const double DM2 = m[28], T = m[29], N = m[30];
  const double n = nden[c];
  if (!(N > 1e-14)) {
    finish(false, nullptr);
    return;
  }
  const double Q[3] = {M3[0] + M3[3] + M3[5], M3[1] + M3[6] + M3[8], M3[2] + M3[7] + M3[9]};
  const double DM4 = M4[0] + M4[3] + M4[5];
  const double Yn = yc[c];
  const double mass = kp.mass, kb = kp.kb;
  out[22] = -0.001 * (n * kp.b * Yn / (kb * T / mass));
  double mu = fpp.mu * det_pow(T / fpp.T0, fpp.visp);
  double tau = 2.0 * mu / (n * kb * T);
  double D3 = DM2 / 3.0;
  double L = (P[0] - D3) * ((P[3] - D3) * (P[5] - D3) - P[4] * P[4]);
  L = L - P[1] * (P[1] * (P[5] - D3) - P[4] * P[2]);
  L = L + P[2] * (P[1] * P[4] - P[2] * (P[3] - D3));
  double D2 = DM2 * DM2;
  L = -L / (tau * (D2 * D2));
  out[0] = L;
  out[27] = tau;
  const bool dense = fpp.model == MCPM_MODEL_FP_DENSE;
  const bool fp = dense || fpp.model == MCPM_MODEL_FP_IDEAL;
  if (!fp || cnt[c] < 2) {
    finish(false, nullptr);
    return;
  }
  if (dense && step > 0) {
    double *A = sA;
    double rhs[12];
    for (int q = lane_id(); q < 144; q += 64) A[q] = 0.0;
    if ((int)lane_id() < 12) sX[lane_id()] = 0.0;
    __syncthreads();
    double trace = P[0] + P[3] + P[5];
    double dev[6] = {P[0], P[1], P[2], P[3], P[4], P[5]};
    dev[0] -= (1.0 / 3.0) * trace;
    dev[3] -= (1.0 / 3.0) * trace;
    dev[5] -= (1.0 / 3.0) * trace;
    double a = Q[0], b = Q[1], cc = Q[2];
    double d = DM4 - DM2 * DM2;
    // block rows of alpha_calculation (Fokker_Planck.cpp:277-370)
    for (int blk = 0; blk < 3 && w0; ++blk) {
      const double qv = blk == 0 ? a : (blk == 1 ? b : cc);
      const int r0 = 3 * blk;
      A[(r0 + 0) * 12 + r0 + 0] = P[0]; A[(r0 + 0) * 12 + r0 + 1] = P[1]; A[(r0 + 0) * 12 + r0 + 2] = P[2];
      A[(r0 + 1) * 12 + r0 + 0] = P[1]; A[(r0 + 1) * 12 + r0 + 1] = P[3]; A[(r0 + 1) * 12 + r0 + 2] = P[4];
      A[(r0 + 2) * 12 + r0 + 0] = P[2]; A[(r0 + 2) * 12 + r0 + 1] = P[4]; A[(r0 + 2) * 12 + r0 + 2] = P[5];
      A[(r0 + 0) * 12 + 9 + blk] = qv;
      A[(r0 + 1) * 12 + 9 + blk] = qv;
      A[(r0 + 2) * 12 + 9 + blk] = qv;
      A[(9 + blk) * 12 + r0 + 0] = a;
      A[(9 + blk) * 12 + r0 + 1] = b;
      A[(9 + blk) * 12 + r0 + 2] = cc;
      A[(9 + blk) * 12 + 9 + blk] = d;
    }
    double nbY = n * kp.b * Yn;
    double p_crr = 1.0 + nbY;
    double mu_crr = 1.0 + 2.0 * nbY / 5.0;
    double k_crr = 1.0 + 3.0 * nbY / 5.0;
    double s2 = kp.sigma * kp.sigma;
    double w = (4.0 / 9.0) * __builtin_sqrt(3.14159265358979323846) * Yn * (n * n) * (s2 * s2) *
               __builtin_sqrt(mass * kb * T);
    auto Uy = [&](int c2) { return mom[(size_t)c2 * MCPM_NMOM + 1]; };
    double divU;
    if (kp.boundary == MCPM_BC_PERIODIC) divU = (Uy((c + 1) % nc) - Uy((c + nc - 1) % nc)) / kp.dy / 2.0;
    else if (c == 0) divU = (Uy(1 % nc) - Uy(0)) / kp.dy;
    else if (c == nc - 1) divU = (Uy(c) - Uy(c - 1)) / kp.dy;
    else divU = (Uy(c + 1) - Uy(c - 1)) / kp.dy / 2.0;
    out[26] = divU;
    double cmu = 3.0 * 1.002 / 5.0 * (nbY * nbY) / (1 + 2.0 * nbY / 5.0);
    double ck = 1.002 * (nbY * nbY) / (1 + 3.0 * nbY / 5.0) / 5.0;
    double gst = out[22];
    double bulk = w * divU / (mass * n);
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
    if (w0)
      for (int q = 0; q < 12; ++q) sR[q] = rhs[q];
    __syncthreads();
    wave_gauss_jordan<12>(A, sR, sX, sZ, sF);
    for (int q = 0; q < 9; ++q) out[10 + q] = sX[q];
    for (int q = 0; q < 3; ++q) out[19 + q] = sX[9 + q];
    if (w0 && sX[0] != sX[0]) atomicOr(err, ERRBIT_NAN);
    __syncthreads();  // sX is read before the next elimination writes it
  }
  {
    double *A = sA;
    double rhs[9];
    for (int q = lane_id(); q < 81; q += 64) A[q] = 0.0;
    if ((int)lane_id() < 9) sX[lane_id()] = 0.0;
    __syncthreads();
    if (w0) {
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
    }
    const double Pr = 2.0 / 3.0;
    const double Yc = dense ? Yn : 1.0;
    for (int q = 0; q < 6; ++q) rhs[q] = -2.0 * L * M4[q];
    for (int q = 0; q < 3; ++q) rhs[6 + q] = Yc * (3.0 / tau - 2.0 * Pr / tau) * Q[q];
    rhs[6] += L * (-3.0 * M5[0] + DM2 * Q[0] + 2.0 * (P[0] * Q[0] + P[1] * Q[1] + P[2] * Q[2]));
    rhs[7] += L * (-3.0 * M5[1] + DM2 * Q[1] + 2.0 * (P[1] * Q[0] + P[3] * Q[1] + P[4] * Q[2]));
    rhs[8] += L * (-3.0 * M5[2] + DM2 * Q[2] + 2.0 * (P[2] * Q[0] + P[4] * Q[1] + P[5] * Q[2]));
    if (w0)
      for (int q = 0; q < 9; ++q) sR[q] = rhs[q];
    __syncthreads();
    wave_gauss_jordan<9>(A, sR, sX, sZ, sF);
    for (int q = 0; q < 6; ++q) out[1 + q] = sX[q];
    for (int q = 0; q < 3; ++q) out[7 + q] = sX[6 + q];
  }
  const double *al = out + 10, *mp = m + 31;
  out[23] = al[0] * mp[0] + al[1] * mp[1] + al[2] * mp[2];
  out[24] = al[3] * mp[0] + al[4] * mp[1] + al[5] * mp[2];
  out[25] = al[6] * mp[0] + al[7] * mp[1] + al[8] * mp[2];
  double d[DRIFT_N];
  if (dense) {
    d[0] = m[0]; d[1] = m[1]; d[2] = m[2]; d[3] = T;
    d[4] = out[13]; d[5] = out[14]; d[6] = out[15]; d[7] = out[20];
    d[8] = out[22]; d[9] = M3[1] + M3[6] + M3[8]; d[10] = out[24]; d[11] = 1.0;
  }
  finish(dense, d);
}

__device__ __forceinline__ void fp_noise(uint64_t slot, const KParams &kp, uint32_t slo,
                                         uint32_t shi, double *z) {
  mcpm_u32x4 w = mcpm_philox((uint32_t)slot, 0u, slo,
                             mcpm_ctr3(MCPM_TAG_FPNOISE, shi, (uint32_t)(slot >> 32)), kp.key0,
                             kp.key1);
  double z3;
  mcpm_normal2(w.v[0], w.v[1], &z[0], &z[1]);
  mcpm_normal2(w.v[2], w.v[3], &z[2], &z3);
}

template <int K>
__device__ __forceinline__ void fp_block_sum(double (&v)[K], double *sh) {
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

// random_generator + Velocity_FP_parallel for one cell per workgroup.
// Cells of at most 256 KMAX particles: a thread keeps the noise of its
// particles (k = tid + 256 j) in registers, so the noise is drawn once
// instead of once per pass.  Larger cells draw it again in each of the four
// passes.  The same operations in the same order either way.
template <typename real, int KMAX>
__global__ __launch_bounds__(256) void k_fp_velocity(Rec<real> *__restrict__ P,
                                                     const uint32_t *__restrict__ cnt,
                                                     const uint32_t *__restrict__ start,
                                                     const double *__restrict__ mom,
                                                     const double *__restrict__ clos,
                                                     const double *__restrict__ yc, KParams kp,
                                                     FPParams fpp, uint32_t slo, uint32_t shi,
                                                     uint32_t *err) {
  __shared__ double sh[3 * 4];
  const int c = blockIdx.x;
  const uint32_t N = cnt[c];
  if (N == 0) return;
  const uint32_t s = start[c];
  const double *m = mom + (size_t)c * MCPM_NMOM;
  const double U0 = m[0], U1 = m[1], U2 = m[2];
  if (N < 2) {
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x) {
      Rec<real> r = P[s + k];
      r.u1 = (real)U0;
      r.u2 = (real)U1;
      r.u3 = (real)U2;
      P[s + k] = r;
    }
    return;
  }
  const double *cl = clos + (size_t)c * MCPM_NCLOS;
  const double T = m[29];
  constexpr int KC = KMAX > 0 ? KMAX : 1;
  const bool cached = KMAX > 0 && N <= 256u * (uint32_t)KC && blockDim.x == 256;
  double zc[KC][3];
  // noise of particle k (the j-th of this thread when cached)
  auto noise = [&](uint32_t k, int j, double *z) {
    if (cached) {
      z[0] = zc[0][0];
      z[1] = zc[0][1];
      z[2] = zc[0][2];
#pragma unroll
      for (int q = 1; q < KC; ++q)
        if (q == j) {
          z[0] = zc[q][0];
          z[1] = zc[q][1];
          z[2] = zc[q][2];
        }
    } else {
      fp_noise(s + k, kp, slo, shi, z);
    }
  };
  // noise: per-cell mean and rms (Fokker_Planck.cpp:104-138)
  double acc[3] = {0.0, 0.0, 0.0};
  {
    int j = 0;
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x, ++j) {
      double z[3];
      fp_noise(s + k, kp, slo, shi, z);
      if (cached) {
#pragma unroll
        for (int q = 0; q < KC; ++q)
          if (q == j) {
            zc[q][0] = z[0];
            zc[q][1] = z[1];
            zc[q][2] = z[2];
          }
      }
      acc[0] += z[0]; acc[1] += z[1]; acc[2] += z[2];
    }
  }
  fp_block_sum<3>(acc, sh);
  const double mean0 = acc[0] / (1.0 * N), mean1 = acc[1] / (1.0 * N), mean2 = acc[2] / (1.0 * N);
  double sq[3] = {0.0, 0.0, 0.0};
  {
    int j = 0;
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x, ++j) {
      double z[3];
      noise(k, j, z);
      double a = z[0] - mean0, b = z[1] - mean1, d = z[2] - mean2;
      sq[0] += a * a; sq[1] += b * b; sq[2] += d * d;
    }
  }
  fp_block_sum<3>(sq, sh);
  const double rms0 = __builtin_sqrt(sq[0] / (1.0 * N));
  const double rms1 = __builtin_sqrt(sq[1] / (1.0 * N));
  const double rms2 = __builtin_sqrt(sq[2] / (1.0 * N));
  // OU coefficients (Fokker_Planck.cpp:1099-1112)
  double tau = cl[27];
  if (fpp.model == MCPM_MODEL_FP_DENSE) tau = tau / yc[c];
  const double A1 = mcpm_exp(-kp.dt / tau);
  const double A3 = __builtin_sqrt((1.0 - mcpm_exp(-2.0 * kp.dt / tau)) * kp.kb * T / (1.0 * kp.mass));
  const double A2 = (1.0 - mcpm_exp(-kp.dt / tau)) * tau;
  const double L = cl[0];
  const double *cc = cl + 1, *g = cl + 7;
  const double Q0 = m[9] + m[12] + m[14], Q1 = m[10] + m[15] + m[17], Q2 = m[11] + m[16] + m[18];
  const double e0 = 3.0 * kp.kb * T / kp.mass;
  // Velocity_FP_linear (Fokker_Planck.cpp:1240-1264): the same step without drift
  const bool linear = fpp.model == MCPM_MODEL_FP_LINEAR;
  auto update = [&](const Rec<real> &r, double &n0, double &n1, double &n2, uint32_t k,
                    int j) {
    double a = (double)r.u1 - U0, b = (double)r.u2 - U1, d = (double)r.u3 - U2;
    double e = a * a + b * b + d * d;
    if (linear) {
      double z[3];
      noise(k, j, z);
      n0 = A1 * a + A3 * ((z[0] - mean0) / rms0);
      n1 = A1 * b + A3 * ((z[1] - mean1) / rms1);
      n2 = A1 * d + A3 * ((z[2] - mean2) / rms2);
      return e;
    }
    double d0 = cc[0] * a + cc[1] * b + cc[2] * d;
    d0 += g[0] * (e - e0);
    d0 += L * (a * e - Q0);
    double d1 = cc[1] * a + cc[3] * b + cc[4] * d;
    d1 += g[1] * (e - e0);
    d1 += L * (b * e - Q1);
    double d2 = cc[2] * a + cc[4] * b + cc[5] * d;
    d2 += g[2] * (e - e0);
    d2 += L * (d * e - Q2);
    double z[3];
    noise(k, j, z);
    double x0 = (z[0] - mean0) / rms0, x1 = (z[1] - mean1) / rms1, x2 = (z[2] - mean2) / rms2;
    n0 = A1 * a + A2 * d0 + A3 * x0;
    n1 = A1 * b + A2 * d1 + A3 * x1;
    n2 = A1 * d + A2 * d2 + A3 * x2;
    return e;
  };
  auto record = [&](uint32_t k, int j) {
    (void)j;
    return P[s + k];
  };
  double fr[2] = {0.0, 0.0};
  {
    int j = 0;
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x, ++j) {
      Rec<real> r = record(k, j);
      double n0, n1, n2;
      double e = update(r, n0, n1, n2, k, j);
      if (n1 != n1) atomicOr(err, ERRBIT_NAN);  // Fokker_Planck.cpp:1137-1149
      fr[0] += e;
      fr[1] += n0 * n0 + n1 * n1 + n2 * n2;
    }
  }
  fp_block_sum<2>(fr, sh);
  const double sc = __builtin_sqrt(__builtin_fabs(fr[0] / fr[1]));
  {
    int j = 0;
    for (uint32_t k = threadIdx.x; k < N; k += blockDim.x, ++j) {
      Rec<real> r = record(k, j);
      double n0, n1, n2;
      update(r, n0, n1, n2, k, j);
      r.u1 = (real)(n0 * sc + U0);
      r.u2 = (real)(n1 * sc + U1);
      r.u3 = (real)(n2 * sc + U2);
      P[s + k] = r;
    }
  }
}

#define MCPM_INST_FPV(real, K)                                                                   \
  template __global__ void k_fp_velocity<real, K>(Rec<real> *, const uint32_t *, const uint32_t *, \
                                                  const double *, const double *, const double *,  \
                                                  KParams, FPParams, uint32_t, uint32_t, uint32_t *);
MCPM_INST_FPV(float, FP_CACHE_K)
MCPM_INST_FPV(double, FP_CACHE_K)

}  // namespace mcpm
