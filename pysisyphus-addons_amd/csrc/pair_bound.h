// Rigorous upper bound of the Schwarz norm of a shell pair, and the test
// that drops a pair before its (ab|ab) block is evaluated when no triplet of
// it can pass the screening.
//
// Q_ab = (sum over the (ab|ab) block of squares)^(1/4).  With phi_mn the
// contracted spherical products of the pair and |f|_C the norm of the
// Coulomb inner product, |(mn|kl)| <= |phi_mn|_C |phi_kl|_C, hence
//   Q_ab <= (sum_mn |phi_mn|_C^2)^(1/2) <= sqrt(na nb) max_mn |phi_mn|_C.
// Pointwise |x^i y^j z^k| <= r^(i+j+k), so a real solid harmonic of degree l
// is bounded by S_l r^l (S_l: the largest sum of |coefficients| of one of
// them), and with p = a + b, K = exp(-ab/p R_AB^2) and the product centre P,
// r_A <= r_P + |PA|, r_B <= r_P + |PB|, n = la + lb, w = la / n:
//   r_A^la r_B^lb exp(-a r_A^2 - b r_B^2)
//     <= K (r_P + |PA|)^la (r_P + |PB|)^lb exp(-p r_P^2) <= K C exp(-p/2 r_P^2),
//   C = c(|PA|, la, w p/2) c(|PB|, lb, (1 - w) p/2),
//   c(d, l, q) = max_s (s + d)^l exp(-q s^2), at s = (sqrt(d^2 + 2l/q) - d)/2.
// The Coulomb kernel is positive, so |f| <= h pointwise gives |f|_C <= |h|_C,
// the norm is subadditive over the primitive pairs, and
//   |K C exp(-q r^2)|_C^2 = (K C)^2 2 pi^(5/2) / (q^2 sqrt(2q)),  q = p/2.
// An auxiliary shell is bounded the same way (one centre: C = max_s s^l
// exp(-beta/2 s^2) = (l/beta)^(l/2) exp(-l/2)), Q_C <= sqrt(nc) max |phi|_C.
//
// A pair is dropped when bound(Q_ab) F < thresh / 2 with F bounding every
// factor its triplets' estimates multiply Q_ab by: max_C Q_C (Schwarz and the
// QVL near field), and under QVL also pi sqrt(2) beta_C / ((a_min +
// b_min)^(1/4) R^(l_C+1)) with R > ext_C (the far field), bounded by
// pi sqrt(2) max_C beta_C / ext_C^(l_C+1).  Such a pair keeps no triplet with
// its exact Q_ab either; the factor 1/2 covers rounding.
#pragma once

#include "hermite.h"

namespace dfhip {

struct PairSkip {
  double thresh = 0.0;            // screening threshold; 0: no pair is dropped
  double qc = 0.0;                // upper bound of max_C Q_C
  double qvl = 0.0;               // QVL: pi sqrt(2) max_C beta_C / ext_C^(l_C+1); Schwarz: 0
  double sl[LMAX_SPH + 1] = {};   // S_l
};

// max_s (s + d)^l exp(-q s^2) over s >= 0 (1 for l = 0).
__host__ __device__ inline double poly_gauss_max(double d, int l, double q) {
  if (l == 0) return 1.0;
  const double s = 0.5 * (sqrt(d * d + 2.0 * l / q) - d);
  return pow(s + d, (double)l) * exp(-q * s * s);
}

// Contribution of primitive pair (a, b) with coefficients ca, cb at squared
// distance R2 to the bound of max_mn |phi_mn|_C (without S_la S_lb).
__host__ __device__ inline double pair_bound_term(double a, double b, double ca, double cb, double R2, int la, int lb) {
  const double p = a + b, R = sqrt(R2);
  const double K = exp(-a * b / p * R2);
  const int n = la + lb;
  const double w = n > 0 ? (double)la / n : 0.5, q = 0.5 * p;
  const double C = poly_gauss_max(b * R / p, la, w * q) * poly_gauss_max(a * R / p, lb, (1.0 - w) * q);
  return fabs(ca * cb) * K * C * sqrt(2.0 * pow(PI, 2.5) / (q * q * sqrt(2.0 * q)));
}

// Upper bound of Q_ab of shell pair (s1, s2).
__host__ __device__ inline double schwarz_pair_bound(const DevShells& s, int s1, int s2, const double* sl) {
  const int la = s.L[s1], lb = s.L[s2];
  const double dx = s.xyz[3 * s1] - s.xyz[3 * s2], dy = s.xyz[3 * s1 + 1] - s.xyz[3 * s2 + 1],
               dz = s.xyz[3 * s1 + 2] - s.xyz[3 * s2 + 2];
  const double R2 = dx * dx + dy * dy + dz * dz;
  const int a0 = s.pptr[s1], b0 = s.pptr[s2];
  double u = 0.0;
  for (int i = 0; i < s.nprim[s1]; i++)
    for (int j = 0; j < s.nprim[s2]; j++)
      u += pair_bound_term(s.exps[a0 + i], s.exps[b0 + j], s.coefs[a0 + i], s.coefs[b0 + j], R2, la, lb);
  return sqrt((double)((2 * la + 1) * (2 * lb + 1))) * sl[la] * sl[lb] * u;
}

// The factor F of the drop test for pair (s1, s2).
__host__ __device__ inline double pair_skip_factor(const DevShells& s, int s1, int s2, const PairSkip& sk) {
  double amin = s.exps[s.pptr[s1]], bmin = s.exps[s.pptr[s2]];
  for (int i = 1; i < s.nprim[s1]; i++) amin = fmin(amin, s.exps[s.pptr[s1] + i]);
  for (int j = 1; j < s.nprim[s2]; j++) bmin = fmin(bmin, s.exps[s.pptr[s2] + j]);
  return sk.qvl > 0.0 ? fmax(sk.qc, sk.qvl / pow(amin + bmin, 0.25)) : sk.qc;
}

__host__ __device__ inline bool pair_negligible(const DevShells& s, int s1, int s2, const PairSkip& sk) {
  if (!(sk.thresh > 0.0)) return false;
  return schwarz_pair_bound(s, s1, s2, sk.sl) * pair_skip_factor(s, s1, s2, sk) < 0.5 * sk.thresh;
}

// The same test evaluated by one wave (lanes over the primitive pairs); the
// result is uniform across the wave.
__device__ inline bool pair_negligible_wave(const DevShells& s, int s1, int s2, const PairSkip& sk, int lane) {
  if (!(sk.thresh > 0.0)) return false;
  const int la = s.L[s1], lb = s.L[s2], nb = s.nprim[s2], npp = s.nprim[s1] * nb;
  const double dx = s.xyz[3 * s1] - s.xyz[3 * s2], dy = s.xyz[3 * s1 + 1] - s.xyz[3 * s2 + 1],
               dz = s.xyz[3 * s1 + 2] - s.xyz[3 * s2 + 2];
  const double R2 = dx * dx + dy * dy + dz * dz;
  const int a0 = s.pptr[s1], b0 = s.pptr[s2];
  double u = 0.0;
  for (int e = lane; e < npp; e += 64) {
    const int i = e / nb, j = e - i * nb;
    u += pair_bound_term(s.exps[a0 + i], s.exps[b0 + j], s.coefs[a0 + i], s.coefs[b0 + j], R2, la, lb);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) u += __shfl_xor(u, off);
  const double B = sqrt((double)((2 * la + 1) * (2 * lb + 1))) * sk.sl[la] * sk.sl[lb] * u;
  return B * pair_skip_factor(s, s1, s2, sk) < 0.5 * sk.thresh;
}

}  // namespace dfhip
