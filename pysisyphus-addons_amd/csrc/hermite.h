// Device building blocks of the McMurchie-Davidson integrals.
//
// Conventions (shared by the 2c2e, 3c2e and 4c2e kernels):
//  * A bra shell pair (a at A, b at B) is expanded per primitive pair ij in
//    Hermite Gaussians Lambda_tuv(p, P) with the matrix
//        Et[mu nu][tuv] = c_i c_j sum_{cart a, cart b} S_a[mu][a] S_b[nu][b]
//                         E^{ax bx}_t E^{ay by}_u E^{az bz}_v
//    where S are the normalized real-solid-harmonic coefficients (m = -l..l)
//    and E the Hermite expansion coefficients (E^{00}_0 = exp(-mu X_AB^2)).
//  * An auxiliary (single-centre) spherical Gaussian S_m(r_C) e^{-g r_C^2}
//    equals (2g)^-l S_m(d/dC) e^{-g r_C^2} (Hobson's theorem), i.e. only the
//    Hermite functions of order exactly l, with the solid-harmonic
//    coefficients themselves.
//  * Hermite Coulomb integrals are evaluated in closed form,
//        R_tuv = sum_{ijk} c_{t,i} c_{u,j} c_{v,k} (2X)^{t-2i} (2Y)^{u-2j}
//                (2Z)^{v-2k} (-alpha)^N F_N(alpha R^2),  N = t+u+v-i-j-k,
//    with c_{t,i} = t!/(i!(t-2i)!), since R_tuv = d^t/dX^t d^u/dY^u
//    d^v/dZ^v F_0(alpha (X^2+Y^2+Z^2)).  Every R_tuv is independent, so a
//    whole Hermite block is one parallel step.
//  * F_N(T): 7-term Taylor expansion around the nearest grid point of a
//    table F_n(k h) (h = 0.05, T < 40), else F_0 = sqrt(pi/T)/2 and the
//    upward recursion F_{n+1} = ((2n+1) F_n - e^{-T}) / (2T).
#pragma once

#include "dfhip_common.h"

namespace dfhip {

constexpr double PI = 3.141592653589793238462643383279502884;

__device__ inline double boys_f(int N, double T, const double* __restrict__ tab) {
  if (T < BOYS_TMAX) {
    int k = __double2int_rn(T * (1.0 / BOYS_H));
    double d = k * BOYS_H - T;  // expansion in (-dT)
    const double* f = tab + (size_t)k * BOYS_NORD + N;
    // sum_j F_{N+j}(T_k) d^j / j!
    double r = f[6] * (1.0 / 720.0);
    r = fma(r, d, f[5] * (1.0 / 120.0));
    r = fma(r, d, f[4] * (1.0 / 24.0));
    r = fma(r, d, f[3] * (1.0 / 6.0));
    r = fma(r, d, f[2] * 0.5);
    r = fma(r, d, f[1]);
    r = fma(r, d, f[0]);
    return r;
  }
  double eT = exp(-T);
  double inv2T = 0.5 / T;
  double F = 0.5 * sqrt(PI / T);
  for (int n = 0; n < N; n++) F = ((2 * n + 1) * F - eT) * inv2T;
  return F;
}

// Hermite expansion coefficients of one Cartesian direction for a primitive
// pair: E[(i*(LB+1)+j)*(LB+LA+1) + t], i <= LA, j <= LB, t <= i+j.  Serial
// (one lane per direction).
template <int LA, int LB>
__device__ inline void herm_E_dir(double a, double b, double A, double B, double* E) {
  constexpr int NT = LA + LB + 1;
  const double p = a + b, oo2p = 0.5 / p;
  const double P = (a * A + b * B) / p;
  const double XPA = P - A, XPB = P - B, X = A - B;
  for (int k = 0; k < (LA + 1) * (LB + 1) * NT; k++) E[k] = 0.0;
  E[0] = exp(-a * b / p * X * X);
  for (int i = 0; i <= LA; i++)
    for (int j = 0; j <= LB; j++) {
      if (i == 0 && j == 0) continue;
      double* e = E + (i * (LB + 1) + j) * NT;
      const double* s = (i > 0) ? E + ((i - 1) * (LB + 1) + j) * NT : E + (i * (LB + 1) + j - 1) * NT;
      const double xp = (i > 0) ? XPA : XPB;
      const int smax = i + j - 1;
      for (int t = 0; t <= i + j; t++) {
        double v = 0.0;
        if (t > 0) v += oo2p * s[t - 1];
        if (t <= smax) v += xp * s[t];
        if (t + 1 <= smax) v += (t + 1) * s[t + 1];
        e[t] = v;
      }
    }
}

// One element of the spherical bra matrix Et[mu*nsph(LB)+nu][h].
template <int LA, int LB>
__device__ inline double bra_element(const SphTable* __restrict__ sph, int mu, int nu, int t, int u, int v,
                                     const double* Ex, const double* Ey, const double* Ez) {
  constexpr int NT = LA + LB + 1;
  double val = 0.0;
  const int a0 = sph->start[LA][mu], a1 = sph->start[LA][mu + 1];
  const int b0 = sph->start[LB][nu], b1 = sph->start[LB][nu + 1];
  for (int ia = a0; ia < a1; ia++) {
    const SphTerm ta = sph->terms[ia];
    for (int ib = b0; ib < b1; ib++) {
      const SphTerm tb = sph->terms[ib];
      if (t > ta.lx + tb.lx || u > ta.ly + tb.ly || v > ta.lz + tb.lz) continue;
      val += ta.c * tb.c * Ex[(ta.lx * (LB + 1) + tb.lx) * NT + t] * Ey[(ta.ly * (LB + 1) + tb.ly) * NT + u] *
             Ez[(ta.lz * (LB + 1) + tb.lz) * NT + v];
    }
  }
  return val;
}

// (t,u,v) of Hermite index h (inverse of herm_index), by search over orders.
__device__ inline void herm_tuv(int h, int& t, int& u, int& v) {
  int s = 0;
  while (nherm(s) <= h) s++;
  int r = h - (s == 0 ? 0 : nherm(s - 1));
  // r = (s-t)(s-t+1)/2 + v with v <= s-t
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= r) k++;
  t = s - k;
  v = r - k * (k + 1) / 2;
  u = s - t - v;
}

// Closed-form Hermite Coulomb integral from per-pair tables
//   pw[d*(L+1) + a] = (2 X_d)^a,  g[N] = scale (-alpha)^N F_N.
__device__ inline double herm_R(int t, int u, int v, int L, const double* pw, const double* g) {
  const double* px = pw;
  const double* py = pw + (L + 1);
  const double* pz = pw + 2 * (L + 1);
  double sum = 0.0;
  double ci = 1.0;
  for (int i = 0; 2 * i <= t; i++) {
    double cj = 1.0;
    const double xi = ci * px[t - 2 * i];
    for (int j = 0; 2 * j <= u; j++) {
      double ck = 1.0;
      const double xij = xi * cj * py[u - 2 * j];
      const int n0 = t + u + v - i - j;
      double inner = 0.0;
      for (int k = 0; 2 * k <= v; k++) {
        inner += ck * pz[v - 2 * k] * g[n0 - k];
        ck *= (double)((v - 2 * k) * (v - 2 * k - 1)) / (k + 1);
      }
      sum += xij * inner;
      cj *= (double)((u - 2 * j) * (u - 2 * j - 1)) / (j + 1);
    }
    ci *= (double)((t - 2 * i) * (t - 2 * i - 1)) / (i + 1);
  }
  return sum;
}

// R_tuv in closed form with the coefficient table hc[t*LH + i] = t!/(i!(t-2i)!).
__device__ inline double herm_R_tab(int t, int u, int v, const double* __restrict__ pw, int L,
                                    const double* __restrict__ g, const double* __restrict__ hc, int LH) {
  const double* px = pw;
  const double* py = pw + (L + 1);
  const double* pz = pw + 2 * (L + 1);
  double sum = 0.0;
  for (int i = 0; 2 * i <= t; i++) {
    const double xi = hc[t * LH + i] * px[t - 2 * i];
    for (int j = 0; 2 * j <= u; j++) {
      const double xij = xi * hc[u * LH + j] * py[u - 2 * j];
      const double* gg = g + (t + u + v - i - j);
      double inner = 0.0;
      for (int k = 0; 2 * k <= v; k++) inner = fma(hc[v * LH + k] * pz[v - 2 * k], gg[-k], inner);
      sum = fma(xij, inner, sum);
    }
  }
  return sum;
}

// herm_E_dir computed in registers (fully unrolled) and stored once.
// bra_element with the solid-harmonic terms of LA and LB staged in LDS:
// terms of (LA, mu) are ta[sa[mu] .. sa[mu+1]), likewise for LB.
template <int LA, int LB>
__device__ inline double bra_element_lds(const int* sa, const SphTerm* ta, const int* sb, const SphTerm* tb, int mu,
                                         int nu, int t, int u, int v, const double* Ex, const double* Ey,
                                         const double* Ez) {
  constexpr int NT = LA + LB + 1;
  double val = 0.0;
  const int b0 = sb[nu], b1 = sb[nu + 1];
  for (int ia = sa[mu]; ia < sa[mu + 1]; ia++) {
    const SphTerm A = ta[ia];
    if (t > A.lx + LB || u > A.ly + LB || v > A.lz + LB) continue;
    const double* ex = Ex + A.lx * (LB + 1) * NT + t;
    const double* ey = Ey + A.ly * (LB + 1) * NT + u;
    const double* ez = Ez + A.lz * (LB + 1) * NT + v;
    double s = 0.0;
    for (int ib = b0; ib < b1; ib++) {
      const SphTerm B = tb[ib];
      if (t > A.lx + B.lx || u > A.ly + B.ly || v > A.lz + B.lz) continue;
      s = fma(B.c, ex[B.lx * NT] * ey[B.ly * NT] * ez[B.lz * NT], s);
    }
    val = fma(A.c, s, val);
  }
  return val;
}

// QVL screening estimate (common/screening.f90:267-299): with the pair
// record pr = [centre x, y, z, extent, (min a + min b)^(1/4), Q_ab] and the aux
// record xr = [extent, beta, Q_C], R = |centre - C|; for R <= ext_ab + ext_C
// the Schwarz product Q_ab Q_C, else Q_ab pi sqrt(2) beta / ((..)^(1/4) R^(L+1)).
// The triplet is kept unless the estimate is below thresh.  Evaluated in
// the reference's operation order without contraction, so host and device
// take the same decisions.
__host__ __device__ inline double qvl_estimate(const double* pr, const double* xr, double Cx, double Cy, double Cz,
                                               int L) {
#pragma clang fp contract(off)
  const double dx = pr[0] - Cx, dy = pr[1] - Cy, dz = pr[2] - Cz;
  const double R = sqrt(dx * dx + dy * dy + dz * dz);
  if (R <= pr[3] + xr[0]) return pr[5] * xr[2];
  double Rp = R;
  for (int i = 0; i < L; i++) Rp = Rp * R;
  return pr[5] * PI * sqrt(2.0) * xr[1] / (pr[4] * Rp);
}

__host__ __device__ inline bool qvl_keep(const double* pr, const double* xr, double Cx, double Cy, double Cz, int L,
                                         double thresh) {
  return !(qvl_estimate(pr, xr, Cx, Cy, Cz, L) < thresh);
}

}  // namespace dfhip
