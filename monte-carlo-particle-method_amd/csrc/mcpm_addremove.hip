// mcpm_addremove.hip — particle weights that differ from cell to cell:
// the splitting and removal of inverted_addremove_particles
// (Main_Functions.cpp:1206-1450) on the sorted slab, on the device.
//
// Every particle carries an integer weight (its n_ratio), every cell one
// (cells[num].n_ratio).  A particle heavier than its cell splits into
// n_ratio particles: it stays, and n_ratio - 1 copies at its position take
// its velocity plus a kernel-density perturbation h N(0,1) per component,
// h = 1.06 sig (n_rem + n_ratio)^(-1/5) (:1337-1349).  A lighter one is
// deleted with probability 1 - 1/n_ratio (:1256-1268), otherwise it takes the
// cell's weight.  Afterwards every particle has its cell's weight (:1435-1441).
//
// The output stays in cell order: cell c's kept particles in their order,
// then its new particles by parent and copy, which is the order the
// reference's index lists give the cell at its next binning (kept particles
// in array order, the appended new ones after them).  Random draws are
// Philox blocks keyed by (particle position, copy, step): (p, 0, TAG_ADDREM,
// sub 0) the deletion draw, (p, j, TAG_ADDREM, sub 1) copy j's three normals.
#include "mcpm_kernels.h"

namespace mcpm {

// Classification, one workgroup of 256 per cell: per particle its flag, per
// cell the split
// parents, the particles kept besides them, the deleted ones, the split
// bandwidth h and the cell's particle total after the operation.
template <typename R>
__global__ __launch_bounds__(256) void k_ar_classify(const R *__restrict__ P,
                                                     const uint32_t *__restrict__ cnt,
                                                     const uint32_t *__restrict__ start,
                                                     const int32_t *__restrict__ pw,
                                                     const int32_t *__restrict__ cr, int32_t nr,
                                                     KParams kp, uint32_t slo, uint32_t shi,
                                                     uint8_t *__restrict__ fl,
                                                     uint32_t *__restrict__ cst,
                                                     double *__restrict__ ch,
                                                     uint32_t *__restrict__ tot) {
  __shared__ uint32_t s_n[3];
  __shared__ unsigned long long s_w;
  __shared__ int32_t s_last;
  __shared__ double sh[4];
  const int c = blockIdx.x;
  const uint32_t N = cnt[c], s0 = start[c];
  const int32_t wc = cr[c];
  if (threadIdx.x == 0) {
    s_n[0] = s_n[1] = s_n[2] = 0u;
    s_w = 0ull;
    s_last = -1;
  }
  __syncthreads();
  // ratio = 1.0 - 1.0 / (n_ratio * 1.0) (:1251)
  const double pdel = 1.0 - 1.0 / ((double)nr * 1.0);
  uint32_t ns = 0, nk = 0, nd = 0;
  unsigned long long wsum = 0ull;
  int32_t last = -1;
  for (uint32_t u = threadIdx.x; u < N; u += blockDim.x) {
    const uint32_t p = s0 + u;
    int32_t w = pw[p];
    uint8_t f = AR_KEEP;
    if (w > wc) {
      f = AR_SPLIT;
      ++ns;
    } else if (w < wc) {
      const mcpm_u32x4 r =
          mcpm_philox(p, 0u, slo, mcpm_ctr3(MCPM_TAG_ADDREM, shi, 0u), kp.key0, kp.key1);
      if (mcpm_u01(r.v[0]) < pdel) {
        f = AR_DELETE;
        ++nd;
      } else {
        w = wc;  // it takes the cell's weight (:1270-1273)
        ++nk;
      }
    } else {
      ++nk;
    }
    fl[p] = f;
    if (f != AR_DELETE) {
      wsum += (unsigned long long)(uint32_t)w;
      last = (int32_t)u;
    }
  }
  if (ns) atomicAdd(&s_n[0], ns);
  if (nk) atomicAdd(&s_n[1], nk);
  if (nd) atomicAdd(&s_n[2], nd);
  if (wsum) atomicAdd(&s_w, wsum);
  if (last >= 0) atomicMax(&s_last, last);
  __syncthreads();
  const uint32_t Ns = s_n[0], Nk = s_n[1];
  double h = 0.0;
  if (Ns > 0) {
    // mean velocity as the reference forms it (:1302-1313): the last
    // particle kept in the cell, times its weight, over the cell's weight
    // (a kept particle's weight is its own or, if lighter, the cell's)
    const uint32_t pl = s0 + (uint32_t)s_last;
    const R rl = P[pl];
    const double wl = (double)(pw[pl] > wc ? pw[pl] : wc);
    const double W = 1.0 * (double)s_w;
    const double m1 = ((double)rl.u1 * wl) / W;
    const double m2 = ((double)rl.u2 * wl) / W;
    const double m3 = ((double)rl.u3 * wl) / W;
    // sig = sqrt(sum of weight x |U - mean|^2 over the kept / weight) (:1314-1325)
    double v[1] = {0.0};
    for (uint32_t u = threadIdx.x; u < N; u += blockDim.x) {
      const uint32_t p = s0 + u;
      if (fl[p] == AR_DELETE) continue;
      const R r = P[p];
      const double a = (double)r.u1 - m1, b = (double)r.u2 - m2, d = (double)r.u3 - m3;
      const double e = a * a + b * b + d * d;
      v[0] += e * (double)(pw[p] > wc ? pw[p] : wc);
    }
    block_sum<1>(v, sh);
    const double sig = __builtin_sqrt(v[0] / W);
    // h = 1.06 sig (n_rem + n_ratio)^(-1/5) (:1343), the power by the
    // deterministic exp and log
    h = 1.06 * sig * mcpm_exp((-1.0 / 5.0) * mcpm_log(1.0 * (double)(Nk + (uint32_t)nr)));
  }
  if (threadIdx.x == 0) {
    cst[4 * c + 0] = Ns;
    cst[4 * c + 1] = Nk;
    cst[4 * c + 2] = s_n[2];
    cst[4 * c + 3] = 0u;
    ch[c] = h;
    tot[c] = Ns + Nk + Ns * (uint32_t)(nr - 1);
  }
}

// Compaction and insertion, one workgroup of 256 per cell: kept particles
// to ostart[c] + their rank among the kept, copy j of split parent q to
// ostart[c] + kept + q (n_ratio - 1) + j; every output particle takes the
// cell's weight.
template <typename R>
__global__ __launch_bounds__(256) void k_ar_scatter(const R *__restrict__ P, R *__restrict__ out,
                                                    const uint32_t *__restrict__ cnt,
                                                    const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ ostart,
                                                    const uint8_t *__restrict__ fl,
                                                    const uint32_t *__restrict__ cst,
                                                    const double *__restrict__ ch,
                                                    const int32_t *__restrict__ cr, int32_t nr,
                                                    KParams kp, uint32_t slo, uint32_t shi,
                                                    int32_t *__restrict__ ow,
                                                    typename R::real_type *__restrict__ oys) {
  using real = typename R::real_type;
  __shared__ uint32_t s_wk[4], s_ws[4];
  const int c = blockIdx.x;
  const uint32_t N = cnt[c], s0 = start[c], o0 = ostart[c];
  const uint32_t kept = cst[4 * c + 0] + cst[4 * c + 1];
  const double h = ch[c];
  const int32_t wc = cr[c];
  const uint32_t nco = (uint32_t)(nr - 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // kept and split particles of the earlier tiles (the same in every thread)
  uint32_t ck = 0u, cs = 0u;
  for (uint32_t base = 0; base < N; base += blockDim.x) {
    const uint32_t u = base + threadIdx.x;
    const bool live = u < N;
    const uint8_t f = live ? fl[s0 + u] : AR_DELETE;
    const bool k = f != AR_DELETE, sp = f == AR_SPLIT;
    const uint64_t mk = __ballot(k), ms = __ballot(sp);
    if (lane == 0) {
      s_wk[wave] = (uint32_t)__popcll(mk);
      s_ws[wave] = (uint32_t)__popcll(ms);
    }
    __syncthreads();
    uint32_t rk = ck, rs = cs, tk = 0u, ts = 0u;
    for (int w = 0; w < 4; ++w) {
      if (w < wave) {
        rk += s_wk[w];
        rs += s_ws[w];
      }
      tk += s_wk[w];
      ts += s_ws[w];
    }
    rk += (uint32_t)__popcll(mk & lanemask_lt());
    rs += (uint32_t)__popcll(ms & lanemask_lt());
    ck += tk;
    cs += ts;
    if (k) {
      const uint32_t p = s0 + u;
      const R r = P[p];
      const uint32_t d = o0 + rk;
      out[d] = r;
      ow[d] = wc;
      if (oys) oys[d] = r.y;
      if (sp) {
        for (uint32_t j = 0; j < nco; ++j) {
          const mcpm_u32x4 q =
              mcpm_philox(p, j, slo, mcpm_ctr3(MCPM_TAG_ADDREM, shi, 1u), kp.key0, kp.key1);
          double z0, z1, z2, z3;
          mcpm_normal2(q.v[0], q.v[1], &z0, &z1);
          mcpm_normal2(q.v[2], q.v[3], &z2, &z3);
          R n = r;
          n.u1 = (real)((double)r.u1 + h * z0);
          n.u2 = (real)((double)r.u2 + h * z1);
          n.u3 = (real)((double)r.u3 + h * z2);
          const uint32_t dn = o0 + kept + rs * nco + j;
          out[dn] = n;
          ow[dn] = wc;
          if (oys) oys[dn] = n.y;
        }
      }
    }
    __syncthreads();  // every read of s_wk, s_ws done before the next tile writes them
  }
}

// Every particle's weight from its cell (main.cpp:239-242).
__global__ void k_w_from_cells(const uint32_t *__restrict__ start, int nc,
                               const int32_t *__restrict__ cr, int32_t *__restrict__ w) {
  const int c = blockIdx.x;
  if (c >= nc) return;
  const int32_t wc = cr[c];
  for (uint32_t p = start[c] + threadIdx.x; p < start[c + 1]; p += blockDim.x) w[p] = wc;
}

template __global__ void k_ar_classify<Rec<float>>(const Rec<float> *, const uint32_t *,
                                                   const uint32_t *, const int32_t *,
                                                   const int32_t *, int32_t, KParams, uint32_t,
                                                   uint32_t,
                                                   uint8_t *, uint32_t *, double *, uint32_t *);
template __global__ void k_ar_classify<Rec<double>>(const Rec<double> *, const uint32_t *,
                                                    const uint32_t *, const int32_t *,
                                                    const int32_t *, int32_t, KParams, uint32_t,
                                                    uint32_t,
                                                    uint8_t *, uint32_t *, double *, uint32_t *);
template __global__ void k_ar_scatter<Rec<float>>(const Rec<float> *, Rec<float> *,
                                                  const uint32_t *, const uint32_t *,
                                                  const uint32_t *, const uint8_t *,
                                                  const uint32_t *, const double *,
                                                  const int32_t *, int32_t, KParams, uint32_t,
                                                  uint32_t, int32_t *, float *);
template __global__ void k_ar_scatter<Rec<double>>(const Rec<double> *, Rec<double> *,
                                                   const uint32_t *, const uint32_t *,
                                                   const uint32_t *, const uint8_t *,
                                                   const uint32_t *, const double *,
                                                   const int32_t *, int32_t, KParams, uint32_t,
                                                   uint32_t, int32_t *, double *);

}  // namespace mcpm
