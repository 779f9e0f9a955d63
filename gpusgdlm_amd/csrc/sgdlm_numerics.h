/*
 * sgdlm_numerics.h -- deterministic scalar numerics shared by the HIP kernels
 * and the FP64 host oracle.
 *
 * Everything here is plain IEEE-754 binary64 arithmetic written out operation
 * by operation (no libm calls, no fused multiply-add unless written as an
 * explicit fma, which is correctly rounded on both x86-64 and gfx950), so a
 * function evaluated on the host and on the GPU returns the same bits.  All
 * translation units that include this file must be compiled with
 * -ffp-contract=off.
 *
 * Contents
 *   - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 *     "Parallel random numbers: as easy as 1, 2, 3", SC'11).  Replaces the
 *     cuRAND XORWOW stream of rSGDLM/src/curand_manager.cuh:28-56.
 *   - sg_log / sg_exp / sg_sincospi: argument reduction + truncated series
 *     (atanh series for log, Taylor series for exp, sin(pi r), cos(pi r)).
 *   - Box-Muller normal pair.
 *   - Marsaglia-Tsang gamma sampler (ACM TOMS 26(3), 2000), parameterised as
 *     in rSGDLM/src/kernel_functions.cuh:568-634 (shape n/2, rate n*s/2),
 *     with the shape<1 boost G(a) = G(a+1) * U^(1/a).
 *   - Beta draw of the multi-step evolution, kernel_functions.cuh:643-798.
 *   - Per-series Cholesky with the operation order of
 *     kernel_functions.cuh:844-902 and the theta transform of
 *     SGDLM.cu:659-673 + kernel_functions.cuh:71-83.
 *
 * The header is valid C99 (for the oracle) and HIP C++.
 */
#ifndef SGDLM_NUMERICS_H
#define SGDLM_NUMERICS_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SG_HD static __host__ __device__ inline
#else
#define SG_HD static inline
#endif

/* ------------------------------------------------------------------------ */
/* RNG stream layout.                                                        */
/*   key     = (seed low 32 bits, seed high 32 bits)                         */
/*   counter = (x = slot, y = series j, z = draw i, w = epoch<<8 | purpose)  */
/* ------------------------------------------------------------------------ */
enum {
  SG_PURPOSE_LAMBDA = 1, /* gamma draw of lambda_ij (VB and forecast)      */
  SG_PURPOSE_THETA = 2,  /* standard normals z for theta_ij                 */
  SG_PURPOSE_NU = 3,     /* observation noise nu_ij of a forecast           */
  SG_PURPOSE_BETA_A = 4, /* G_a of the evolution beta draw                  */
  SG_PURPOSE_BETA_B = 5, /* G_b of the evolution beta draw                  */
  SG_PURPOSE_OMEGA = 6   /* state innovations omega_ij of the evolution     */
};

#define SG_GAMMA_MAX_ATTEMPTS 256u
#define SG_BOOST_SLOT 0xFFFFFFFEu

typedef struct {
  uint32_t v[4];
} sg_u32x4;

typedef union {
  double d;
  uint64_t u;
} sg_bits64;

SG_HD sg_u32x4 sg_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                uint32_t c3, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  int r;
  sg_u32x4 out;
  for (r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * (uint64_t)c0;
    uint64_t p1 = (uint64_t)M1 * (uint64_t)c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

/* One Philox block addressed by the stream layout above. */
SG_HD sg_u32x4 sg_rng_block(uint64_t seed, uint32_t epoch, uint32_t purpose,
                            uint32_t draw, uint32_t series, uint32_t slot) {
  return sg_philox4x32_10(slot, series, draw, (epoch << 8) | (purpose & 0xFFu),
                          (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* Uniform in (0,1): (k + 1/2) * 2^-52 with a 52-bit integer k built from two
 * 32-bit words; exact in binary64, never 0 or 1. */
SG_HD double sg_u01(uint32_t hi, uint32_t lo) {
  uint64_t k = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return ((double)k + 0.5) * 0x1.0p-52;
}

/* ---------------------------- elementary functions ---------------------- */

/* Natural log for finite x > 0.  x = 2^e * f with f in (sqrt(2)/2, sqrt(2)],
 * log f = 2 atanh(s), s = (f-1)/(f+1), atanh series to s^23 (|s| <= 0.1716,
 * truncation < 1e-18 relative). */
SG_HD double sg_log(double x) {
  sg_bits64 b;
  int e;
  double f, s, z, R, lm, de;
  if (!(x > 0.0)) {
    return (x == 0.0) ? -1.0 / 0.0 : (x - x) / 0.0; /* -inf or NaN */
  }
  if (x == 1.0 / 0.0)
    return x;
  b.d = x;
  e = 0;
  if ((b.u >> 52) == 0) { /* subnormal */
    b.d = x * 0x1.0p54;
    e = -54;
  }
  e += (int)((b.u >> 52) & 0x7FFu) - 1023;
  b.u = (b.u & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  f = b.d; /* [1,2) */
  if (f > 0x1.6a09e667f3bcdp+0) {
    f = f * 0.5;
    e += 1;
  }
  f = f - 1.0;
  s = f / (2.0 + f);
  z = s * s;
  R = 0x1.642c8590b2164p-4;
  R = 0x1.8618618618618p-4 + z * R;
  R = 0x1.af286bca1af28p-4 + z * R;
  R = 0x1.e1e1e1e1e1e1ep-4 + z * R;
  R = 0x1.1111111111111p-3 + z * R;
  R = 0x1.3b13b13b13b14p-3 + z * R;
  R = 0x1.745d1745d1746p-3 + z * R;
  R = 0x1.c71c71c71c71cp-3 + z * R;
  R = 0x1.2492492492492p-2 + z * R;
  R = 0x1.999999999999ap-2 + z * R;
  R = 0x1.5555555555555p-1 + z * R;
  R = z * R;
  lm = 2.0 * s + s * R;
  de = (double)e;
  return de * 0x1.62e42fee00000p-1 + (lm + de * 0x1.a39ef35793c76p-33);
}

/* 2^k as a double for -1074 <= k <= 1023, built from bits in at most two
 * exact steps. */
SG_HD double sg_pow2i(int k) {
  sg_bits64 b;
  if (k >= -1022) {
    b.u = (uint64_t)(k + 1023) << 52;
    return b.d;
  }
  b.u = (uint64_t)(k + 54 + 1023) << 52;
  return b.d * 0x1.0p-54;
}

/* exp(x): k = round(x / ln2), r = x - k ln2 (two-part ln2), Taylor to r^13
 * (|r| <= 0.347, truncation < 5e-18 relative). */
SG_HD double sg_exp(double x) {
  double kd, r, p;
  int k;
  if (x != x)
    return x;
  if (x > 709.78)
    return 1.0 / 0.0;
  if (x < -745.2)
    return 0.0;
  kd = x * 0x1.71547652b82fep+0 + 0.5;
  /* floor without libm: truncate, then fix negatives */
  k = (int)kd;
  if ((double)k > kd)
    k -= 1;
  kd = (double)k;
  r = (x - kd * 0x1.62e42fee00000p-1) - kd * 0x1.a39ef35793c76p-33;
  p = 0x1.6124613a86d09p-33;
  p = 0x1.1eed8eff8d898p-29 + r * p;
  p = 0x1.ae64567f544e4p-26 + r * p;
  p = 0x1.27e4fb7789f5cp-22 + r * p;
  p = 0x1.71de3a556c734p-19 + r * p;
  p = 0x1.a01a01a01a01ap-16 + r * p;
  p = 0x1.a01a01a01a01ap-13 + r * p;
  p = 0x1.6c16c16c16c17p-10 + r * p;
  p = 0x1.1111111111111p-7 + r * p;
  p = 0x1.5555555555555p-5 + r * p;
  p = 0x1.5555555555555p-3 + r * p;
  p = 0x1.0p-1 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  if (k > 1023) { /* only reachable for x within rounding of the limit */
    return (p * 0x1.0p+1000) * sg_pow2i(k - 1000);
  }
  return p * sg_pow2i(k);
}

/* sin(pi x) and cos(pi x) for 0 <= x <= 2.  x = q/2 + r with |r| <= 1/4 (the
 * subtraction is exact), then the Taylor series of sin(pi r) to r^17 and of
 * cos(pi r) to r^18 (truncation < 3e-18). */
SG_HD void sg_sincospi(double x, double *s_out, double *c_out) {
  double t = x * 2.0 + 0.5;
  int q = (int)t; /* t >= 0.5, so truncation == floor */
  double r = x - 0.5 * (double)q;
  double z = r * r;
  double sp, cp;
  sp = 0x1.aaec32af93359p-21;
  sp = -0x1.6fadb9f155744p-16 + z * sp;
  sp = 0x1.e8f434d018d63p-12 + z * sp;
  sp = -0x1.e3074fde8871fp-8 + z * sp;
  sp = 0x1.50783487ee782p-4 + z * sp;
  sp = -0x1.32d2cce62bd86p-1 + z * sp;
  sp = 0x1.466bc6775aae2p+1 + z * sp;
  sp = -0x1.4abbce625be53p+2 + z * sp;
  sp = 0x1.921fb54442d18p+1 + z * sp;
  sp = r * sp;
  cp = -0x1.2a0c591af8314p-23;
  cp = 0x1.20c62c2f2d7f5p-18 + z * cp;
  cp = -0x1.b6e24f44b128fp-14 + z * cp;
  cp = 0x1.f9d38a3763cc3p-10 + z * cp;
  cp = -0x1.a6d1f2a204a8cp-6 + z * cp;
  cp = 0x1.e1f506891babbp-3 + z * cp;
  cp = -0x1.55d3c7e3cbffap+0 + z * cp;
  cp = 0x1.03c1f081b5ac4p+2 + z * cp;
  cp = -0x1.3bd3cc9be45dep+2 + z * cp;
  cp = 1.0 + z * cp;
  switch (q & 3) {
  case 0:
    *s_out = sp;
    *c_out = cp;
    break;
  case 1:
    *s_out = cp;
    *c_out = -sp;
    break;
  case 2:
    *s_out = -sp;
    *c_out = -cp;
    break;
  default:
    *s_out = -cp;
    *c_out = sp;
    break;
  }
}

/* ------------------------------- samplers ------------------------------- */

/* Box-Muller: two independent N(0,1) from one Philox block. */
SG_HD void sg_normal_pair(sg_u32x4 b, double *z0, double *z1) {
  double u1 = sg_u01(b.v[0], b.v[1]);
  double u2 = sg_u01(b.v[2], b.v[3]);
  double rad = __builtin_sqrt(-2.0 * sg_log(u1));
  double sn, cs;
  sg_sincospi(2.0 * u2, &sn, &cs);
  *z0 = rad * cs;
  *z1 = rad * sn;
}

/* Gamma(shape, rate) by Marsaglia-Tsang, one (draw, series, purpose) stream:
 * attempt t uses slot 2t for the normal and slot 2t+1 for the uniform; the
 * shape<1 boost uses slot SG_BOOST_SLOT.  Returns NaN if no attempt is
 * accepted within SG_GAMMA_MAX_ATTEMPTS (only for invalid parameters). */
SG_HD double sg_gamma(double shape, double rate, uint64_t seed, uint32_t epoch,
                      uint32_t purpose, uint32_t draw, uint32_t series) {
  double a = shape;
  double boost = 1.0;
  double d, c, g;
  uint32_t t;
  if (!(shape > 0.0) || !(rate > 0.0)) {
    return 0.0 / 0.0;
  }
  if (a < 1.0) {
    sg_u32x4 bb = sg_rng_block(seed, epoch, purpose, draw, series, SG_BOOST_SLOT);
    double ub = sg_u01(bb.v[0], bb.v[1]);
    boost = sg_exp(sg_log(ub) / a);
    a = a + 1.0;
  }
  d = a - 1.0 / 3.0;
  c = 1.0 / __builtin_sqrt(9.0 * d);
  g = 0.0 / 0.0;
  for (t = 0; t < SG_GAMMA_MAX_ATTEMPTS; ++t) {
    double x, z1, v, w, x2, u;
    sg_u32x4 bn = sg_rng_block(seed, epoch, purpose, draw, series, 2u * t);
    sg_normal_pair(bn, &x, &z1);
    w = 1.0 + c * x;
    v = w * w * w;
    if (v <= 0.0)
      continue;
    x2 = x * x;
    {
      sg_u32x4 bu = sg_rng_block(seed, epoch, purpose, draw, series, 2u * t + 1u);
      u = sg_u01(bu.v[0], bu.v[1]);
    }
    if (u < 1.0 - 0.0331 * x2 * x2 ||
        sg_log(u) < 0.5 * x2 + d * (1.0 - v + sg_log(v))) {
      g = d * v;
      break;
    }
  }
  return (g * boost) / rate;
}

/* Evolution multiplier eta = G_a / (G_a + G_b) / beta with
 * G_a ~ Gamma(beta r / 2, 1), G_b ~ Gamma((1 - beta) r / 2, 1); eta = 1 when
 * beta == 1 (kernel_functions.cuh:655,737,754-787). */
SG_HD double sg_beta_eta(double r, double beta, uint64_t seed, uint32_t epoch,
                         uint32_t draw, uint32_t series) {
  double ga, gb;
  if (beta == 1.0)
    return 1.0;
  ga = sg_gamma(0.5 * beta * r, 1.0, seed, epoch, SG_PURPOSE_BETA_A, draw, series);
  gb = sg_gamma(0.5 * (1.0 - beta) * r, 1.0, seed, epoch, SG_PURPOSE_BETA_B,
                draw, series);
  return (ga / (ga + gb)) / beta;
}

/* Standard normal number k of the (draw, series, purpose) stream: slot k/2,
 * cosine branch for even k, sine branch for odd k. */
SG_HD double sg_normal_k(uint64_t seed, uint32_t epoch, uint32_t purpose,
                         uint32_t draw, uint32_t series, uint32_t k) {
  double z0, z1;
  sg_u32x4 b = sg_rng_block(seed, epoch, purpose, draw, series, k >> 1);
  sg_normal_pair(b, &z0, &z1);
  return (k & 1u) ? z1 : z0;
}

/* Upper Cholesky factor U (C = U^T U) of the p x p upper-left block of the
 * column-major max_p x max_p matrix C (leading dimension ld), written to U
 * with zeros outside the upper triangle of the block.  Operation order of
 * kernel_functions.cuh:844-902; a non-positive pivot gives a zero row. */
SG_HD void sg_cholesky_upper(int p, int max_p, int ld, const double *C,
                             double *U) {
  int n, k, l;
  for (n = 0; n < p; ++n) {
    for (k = 0; k < n; ++k)
      U[k * ld + n] = 0.0;
    for (k = n; k < p; ++k)
      U[k * ld + n] = C[k * ld + n];
    for (k = p; k < max_p; ++k)
      U[k * ld + n] = 0.0;
  }
  for (n = p; n < max_p; ++n)
    for (k = 0; k < max_p; ++k)
      U[k * ld + n] = 0.0;
  for (n = 0; n < p; ++n) {
    double diag, a;
    if (!(U[n * ld + n] > 0.0)) {
      /* zero (or invalid) pivot of a positive semi-definite block, e.g. the
       * evolution variance W = C (1/Delta - 1) with Delta = 1: the row of U
       * is zero and contributes nothing (the reference computes 0 * (1/0) =
       * NaN here). */
      for (k = n; k < p; ++k)
        U[k * ld + n] = 0.0;
      continue;
    }
    diag = __builtin_sqrt(U[n * ld + n]);
    U[n * ld + n] = diag;
    a = 1.0 / diag;
    for (k = n + 1; k < p; ++k) {
      U[k * ld + n] = U[k * ld + n] * a;
      U[k * ld + k] = U[k * ld + k] - U[k * ld + n] * U[k * ld + n];
    }
    for (k = n + 1; k < p; ++k)
      for (l = k + 1; l < p; ++l)
        U[l * ld + k] = U[l * ld + k] - U[k * ld + n] * U[l * ld + n];
  }
}

/* One naive-posterior draw of series j: lambda ~ Gamma(n/2, n s/2) and
 * theta_k = m_k + (U^T z)_k / sqrt(s lambda) for k < p, with z ~ N(0, I_p)
 * (SGDLM.cu:638-673).  U is the upper Cholesky factor of C in the layout of
 * sg_cholesky_upper.  Writes theta[k * ts], k < p (ts = element stride, so a
 * caller may keep the accumulators in strided scratch memory; the arithmetic
 * does not depend on it). */
SG_HD double sg_draw_lambda_theta(int p, int ld, const double *mj,
                                  const double *Uj, double n, double s,
                                  uint64_t seed, uint32_t epoch, uint32_t draw,
                                  uint32_t series, double *theta, int ts) {
  double lam = sg_gamma(0.5 * n, 0.5 * n * s, seed, epoch, SG_PURPOSE_LAMBDA,
                        draw, series);
  double scale = __builtin_sqrt(s * lam);
  int k, q;
  /* z_k for k = 2q, 2q+1 come from slot q; compute the transform column by
   * column: theta_k = sum_{q<=k} U[k*ld+q] z_q. */
  for (k = 0; k < p; ++k)
    theta[k * ts] = 0.0;
  for (q = 0; q < p; q += 2) {
    double z0, z1;
    sg_normal_pair(sg_rng_block(seed, epoch, SG_PURPOSE_THETA, draw, series,
                                (uint32_t)(q >> 1)),
                   &z0, &z1);
    for (k = q; k < p; ++k) {
      double t = theta[k * ts] + Uj[k * ld + q] * z0;
      if (q + 1 <= k)
        t = t + Uj[k * ld + q + 1] * z1;
      theta[k * ts] = t;
    }
  }
  for (k = 0; k < p; ++k)
    theta[k * ts] = mj[k] + theta[k * ts] / scale;
  return lam;
}

#endif /* SGDLM_NUMERICS_H */
