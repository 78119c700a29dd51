/*
 * mcpm_rng.h — counter-based random numbers and deterministic elementary
 * functions shared, source for source, by the HIP kernels and the CPU oracle.
 *
 * Why this header exists: the reference draws every random number from a
 * std::mt19937 seeded by std::random_device on each call
 * (collisions.cpp:584-590, Main_Functions.cpp:4919-4924) or from rand()
 * (random_generator.cpp:5-29), so no run of it can be reproduced.  Parity
 * between the GPU path and the CPU oracle instead rests on:
 *   - Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 *     as easy as 1, 2, 3", SC'11), keyed by the run seed and addressed by a
 *     counter built from (purpose, step, cell/particle, draw index);
 *   - elementary functions (sin/cos of 2*pi*u, log, exp) written here from
 *     textbook series with Cody-Waite style range reduction, using only
 *     IEEE-754 +,-,*,/ and sqrt, which are correctly rounded on x86-64 SSE2
 *     and on gfx950.  Compiled with -ffp-contract=off on both sides they give
 *     identical bits, which libm and the device math library do not promise.
 *
 * Everything is plain C99/C++ plus an optional __host__ __device__ marker.
 */
#ifndef MCPM_RNG_H
#define MCPM_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MCPM_HD __host__ __device__ __forceinline__
#else
#define MCPM_HD static inline
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                       */
/* ------------------------------------------------------------------ */
typedef struct { uint32_t v[4]; } mcpm_u32x4;

MCPM_HD uint32_t mcpm_mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

/* One Philox4x32 bijection with 10 rounds. */
MCPM_HD mcpm_u32x4 mcpm_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                               uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    /* one 32 x 32 -> 64-bit product per multiplier (a single
       multiply-add instruction on the GPU instead of a low and a high one) */
    const uint64_t p0 = (uint64_t)M0 * (uint64_t)c0, p1 = (uint64_t)M1 * (uint64_t)c2;
    uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  mcpm_u32x4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

/* Stream tags: the fourth counter word is (tag << 24) | (step >> 32 bits
 * folded) so that every random purpose has a disjoint counter space. */
enum {
  MCPM_TAG_MCAND   = 1,  /* per-cell candidate-count rounding draw         */
  MCPM_TAG_CAND    = 2,  /* per-candidate draws (i, direction, j, accept)  */
  MCPM_TAG_PRIO    = 3,  /* per-step key of the candidate priority bijection */
  MCPM_TAG_WALL    = 4,  /* diffuse-wall re-emission draws                  */
  MCPM_TAG_INIT    = 5,  /* initial positions and velocities               */
  MCPM_TAG_FPNOISE = 6,  /* Fokker-Planck Gaussian noise                    */
  MCPM_TAG_DSMC    = 7,  /* plain DSMC / CBA_VHS draws                      */
  MCPM_TAG_NUDGE   = 8,  /* cell-face tie-break nudge                       */
  MCPM_TAG_CBA     = 9,  /* CBA_HS per-candidate draws                      */
  MCPM_TAG_REINJECT = 10, /* evaporation boundary: re-insertion draws       */
  MCPM_TAG_ADDREM  = 11   /* particle add/remove: deletion and copy draws   */
};

MCPM_HD uint32_t mcpm_ctr3(uint32_t tag, uint32_t step_hi, uint32_t sub) {
  return (tag << 24) ^ (step_hi << 8) ^ sub;
}

/* ------------------------------------------------------------------ */
/* 32-bit keyed bijection used as the canonical candidate priority.     */
/* MurmurHash3's fmix32 finaliser (public domain, A. Appleby): each of  */
/* its steps (xor-shift, odd multiply) is invertible on 2^32.            */
/* ------------------------------------------------------------------ */
MCPM_HD uint32_t mcpm_fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85EBCA6Bu;
  h ^= h >> 13; h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}
MCPM_HD uint32_t mcpm_priority(uint32_t cand_index, uint32_t step_key) {
  return mcpm_fmix32(cand_index ^ step_key);
}

/* ------------------------------------------------------------------ */
/* Uniform variates                                                    */
/* ------------------------------------------------------------------ */
/* (w + 0.5) * 2^-32: in the open interval (0,1), exact in double. */
MCPM_HD double mcpm_u01(uint32_t w) {
  return ((double)w + 0.5) * 2.3283064365386962890625e-10;
}
/* Uniform integer in [0, n) by multiply-high (bias below n / 2^32). */
MCPM_HD uint32_t mcpm_uint_below(uint32_t w, uint32_t n) {
  return (uint32_t)(((uint64_t)w * (uint64_t)n) >> 32);
}

/* ------------------------------------------------------------------ */
/* Deterministic elementary functions (double).                         */
/* ------------------------------------------------------------------ */
MCPM_HD double mcpm_floor(double x) {
  /* IEEE floor is exact: the host's floor and the device's v_floor_f64 agree
     on every input. */
  return __builtin_floor(x);
}
MCPM_HD double mcpm_rint(double x) { return mcpm_floor(x + 0.5); }

MCPM_HD uint64_t mcpm_dbits(double x) { union { double d; uint64_t u; } c; c.d = x; return c.u; }
MCPM_HD double mcpm_bitsd(uint64_t u) { union { double d; uint64_t u; } c; c.u = u; return c.d; }

/* sin and cos of 2*pi*u.  Reduce u to r in [-1/8, 1/8] turns plus a
   quadrant, then Taylor series in theta = 2*pi*r, |theta| <= pi/4, carried
   to theta^17 (sin) and theta^18 (cos): truncation below 1e-19. */
MCPM_HD void mcpm_sincos2pi(double u, double *s, double *c) {
  double x = u - mcpm_floor(u);          /* [0,1) */
  double q = mcpm_rint(4.0 * x);         /* 0..4 quarter turns */
  double r = x - 0.25 * q;               /* [-1/8, 1/8] */
  int iq = ((int)q) & 3;
  const double TWO_PI = 6.283185307179586476925286766559;
  double t = TWO_PI * r;
  double t2 = t * t;
  double sp = 1.0 / 355687428096000.0;                /* 1/17! */
  sp = sp * t2 - 1.0 / 1307674368000.0;               /* 1/15! */
  sp = sp * t2 + 1.0 / 6227020800.0;                  /* 1/13! */
  sp = sp * t2 - 1.0 / 39916800.0;                    /* 1/11! */
  sp = sp * t2 + 1.0 / 362880.0;                      /* 1/9!  */
  sp = sp * t2 - 1.0 / 5040.0;                        /* 1/7!  */
  sp = sp * t2 + 1.0 / 120.0;                         /* 1/5!  */
  sp = sp * t2 - 1.0 / 6.0;                           /* 1/3!  */
  double sn = t + t * (t2 * sp);
  double cp = 1.0 / 6402373705728000.0;               /* 1/18! */
  cp = cp * t2 - 1.0 / 20922789888000.0;              /* 1/16! */
  cp = cp * t2 + 1.0 / 87178291200.0;                 /* 1/14! */
  cp = cp * t2 - 1.0 / 479001600.0;                   /* 1/12! */
  cp = cp * t2 + 1.0 / 3628800.0;                     /* 1/10! */
  cp = cp * t2 - 1.0 / 40320.0;                       /* 1/8!  */
  cp = cp * t2 + 1.0 / 720.0;                         /* 1/6!  */
  cp = cp * t2 - 1.0 / 24.0;                          /* 1/4!  */
  cp = cp * t2 + 0.5;                                 /* 1/2!  */
  double cs = 1.0 - t2 * cp;
  double so, co;
  switch (iq) {
    case 0: so = sn; co = cs; break;
    case 1: so = cs; co = -sn; break;
    case 2: so = -sn; co = -cs; break;
    default: so = -cs; co = sn; break;
  }
  *s = so; *c = co;
}

/* Natural log for x > 0 (finite, normal).  x = m * 2^e with m in
   [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1), |s| < 0.1716,
   series to s^27; ln 2 split in a 32-bit-exact head and a tail. */
MCPM_HD double mcpm_log(double x) {
  uint64_t b = mcpm_dbits(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double m = mcpm_bitsd((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 27.0;
  p = p * s2 + 1.0 / 25.0;
  p = p * s2 + 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  double lm = 2.0 * s + 2.0 * s * (s2 * p);
  const double LN2_HI = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
  const double LN2_LO = 1.90821492927058770002e-10;
  double de = (double)e;
  return de * LN2_HI + (lm + de * LN2_LO);
}

/* exp(x) for |x| < 700.  x = k ln2 + r, |r| <= ln2/2 (Cody-Waite), Taylor
   series of exp(r) to r^16, then scale by 2^k through the exponent field. */
MCPM_HD double mcpm_exp(double x) {
  const double INV_LN2 = 1.44269504088896338700e+00;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double kd = mcpm_rint(x * INV_LN2);
  double r = (x - kd * LN2_HI) - kd * LN2_LO;
  double p = 1.0 / 20922789888000.0;                  /* 1/16! */
  p = p * r + 1.0 / 1307674368000.0;
  p = p * r + 1.0 / 87178291200.0;
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  double er = 1.0 + (r + r * (r * p));
  int k = (int)kd;
  /* split the scaling in two so that k down to -1074 stays representable */
  int k1 = k / 2, k2 = k - k1;
  double s1 = mcpm_bitsd((uint64_t)(k1 + 1023) << 52);
  double s2v = mcpm_bitsd((uint64_t)(k2 + 1023) << 52);
  return (er * s1) * s2v;
}

/* Standard normal pair by Box-Muller from two 32-bit words. */
MCPM_HD void mcpm_normal2(uint32_t w0, uint32_t w1, double *z0, double *z1) {
  double u1 = mcpm_u01(w0);
  double u2 = mcpm_u01(w1);
  double rad = __builtin_sqrt(-2.0 * mcpm_log(u1));
  double s, c;
  mcpm_sincos2pi(u2, &s, &c);
  *z0 = rad * c;
  *z1 = rad * s;
}

/* Isotropic unit vector as in the reference (collisions.cpp:651-657):
   theta = 2 pi u, cos(phi) = 2 v - 1, sin(phi) = sqrt(1 - cos^2 phi). */
MCPM_HD void mcpm_iso_dir(uint32_t wu, uint32_t wv, double *d0, double *d1, double *d2) {
  double u = mcpm_u01(wu);
  double v = mcpm_u01(wv);
  double cph = 2.0 * v - 1.0;
  double sph = __builtin_sqrt(1.0 - cph * cph);
  double st, ct;
  mcpm_sincos2pi(u, &st, &ct);
  *d0 = ct * sph;
  *d1 = st * sph;
  *d2 = cph;
}

#endif /* MCPM_RNG_H */
