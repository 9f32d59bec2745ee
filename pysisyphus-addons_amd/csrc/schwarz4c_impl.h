// Diagonal 4c2e blocks (ab|ab) -> Schwarz norms Q_ab = sqrt(||(ab|ab)||_F)
// (common/integrals/intor.f90:131-160).
//
// One workgroup = one shell pair x one chunk of ket functions.  For every
// primitive quartet (ij|kl): spherical Hermite matrices of bra (Et) and ket
// (Ft, chunk rows only), the closed-form R_tuv up to 2(la+lb), then
//   V[h][n2] = sum_h2 (-1)^{|h2|} R_{h+h2} Ft[n2][h2],  acc[n1][n2] += Et[n1] . V[.][n2].
// Each workgroup writes the sum of squares of its chunk; the host-side
// finalize sums the chunks and takes sqrt(sqrt(.)).
#pragma once

#include "hermite.h"

namespace dfhip {

template <int LA, int LB>
struct Cls4 {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L);
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NT = 128;
  static constexpr int BUDGET = 20000;
  static constexpr int FIXED = 6 * ESZ + NAB * H + 4 * (L + 1) + NR + NT;
  // rows of Ft and columns of V per ket function
  static constexpr int PER_N2 = 2 * H;
  static constexpr int NB2_MAX = (BUDGET - FIXED) / PER_N2;
  static constexpr int NB2 = NB2_MAX >= NAB ? NAB : NB2_MAX;
  static constexpr int NCH = (NAB + NB2 - 1) / NB2;
  static constexpr int LDS = FIXED + NB2 * PER_N2;
  static constexpr int OPT = (NAB * NB2 + NT - 1) / NT;
  static_assert(NB2 >= 1, "4c2e class does not fit in LDS");
};

template <int LA, int LB>
__global__ __launch_bounds__(128) void schwarz4c_kernel(DevShells bra, const int* __restrict__ s1v,
                                                        const int* __restrict__ s2v, const SphTable* __restrict__ sph,
                                                        const double* __restrict__ boys, double* __restrict__ part) {
  using K = Cls4<LA, LB>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NAB = K::NAB, NT = K::NT, NB2 = K::NB2;
  __shared__ double lds[K::LDS];
  const int pidx = blockIdx.x, chunk = blockIdx.y;
  const int s1 = s1v[pidx], s2 = s2v[pidx];
  const int n2lo = chunk * NB2, n2n = min(NB2, NAB - n2lo);
  const int tid = threadIdx.x;

  double* Ebra = lds;
  double* Eket = lds + 3 * K::ESZ;
  double* Et = lds + 6 * K::ESZ;
  double* gN = Et + NAB * H;
  double* pw = gN + (L + 1);
  double* R = pw + 3 * (L + 1);
  double* red = R + NR;
  double* Ft = red + NT;
  double* V = Ft + NB2 * H;

  double A[3], B[3];
  for (int d = 0; d < 3; d++) {
    A[d] = bra.xyz[3 * s1 + d];
    B[d] = bra.xyz[3 * s2 + d];
  }
  const int npa = bra.nprim[s1], pa0 = bra.pptr[s1], npb = bra.nprim[s2], pb0 = bra.pptr[s2];
  const int npp = npa * npb;

  double acc[K::OPT];
#pragma unroll
  for (int r = 0; r < K::OPT; r++) acc[r] = 0.0;

  for (int ij = 0; ij < npp; ij++) {
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j];
    const double cab = bra.coefs[pa0 + i] * bra.coefs[pb0 + j];
    const double p = a + b;
    for (int kl = 0; kl < npp; kl++) {
      const int k = kl / npb, l = kl - (kl / npb) * npb;
      const double c = bra.exps[pa0 + k], d = bra.exps[pb0 + l];
      const double ccd = bra.coefs[pa0 + k] * bra.coefs[pb0 + l];
      const double q = c + d;
      __syncthreads();
      if (tid < 3) herm_E_dir<LA, LB>(a, b, A[tid], B[tid], Ebra + tid * K::ESZ);
      else if (tid < 6) herm_E_dir<LA, LB>(c, d, A[tid - 3], B[tid - 3], Eket + (tid - 3) * K::ESZ);
      __syncthreads();
      double PQ[3];
      for (int z = 0; z < 3; z++) PQ[z] = (a * A[z] + b * B[z]) / p - (c * A[z] + d * B[z]) / q;
      const int nE = NAB * H, nF = n2n * H;
      for (int t0 = tid; t0 < nE + nF + 4 * (L + 1); t0 += NT) {
        if (t0 < nE) {
          const int n1 = t0 / H, h = t0 - n1 * H;
          int t, u, v;
          herm_tuv(h, t, u, v);
          Et[t0] = cab * bra_element<LA, LB>(sph, n1 / K::NB, n1 % K::NB, t, u, v, Ebra, Ebra + K::ESZ,
                                             Ebra + 2 * K::ESZ);
        } else if (t0 < nE + nF) {
          const int r = t0 - nE, n2 = r / H, h = r - n2 * H;
          const int nn = n2lo + n2;
          int t, u, v;
          herm_tuv(h, t, u, v);
          const double sgn = ((t + u + v) & 1) ? -1.0 : 1.0;
          Ft[r] = sgn * ccd *
                  bra_element<LA, LB>(sph, nn / K::NB, nn % K::NB, t, u, v, Eket, Eket + K::ESZ, Eket + 2 * K::ESZ);
        } else {
          const int qd = t0 - nE - nF;
          const double alpha = p * q / (p + q);
          if (qd <= L) {
            const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
            double scale = 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q));
            for (int e = 0; e < qd; e++) scale *= -alpha;
            gN[qd] = scale * boys_f(qd, T, boys);
          } else {
            const int qq = qd - (L + 1), dd = qq / (L + 1), e = qq - dd * (L + 1);
            const double x = 2.0 * PQ[dd];
            double v = 1.0;
            for (int z = 0; z < e; z++) v *= x;
            pw[dd * (L + 1) + e] = v;
          }
        }
      }
      __syncthreads();
      for (int h = tid; h < NR; h += NT) {
        int t, u, v;
        herm_tuv(h, t, u, v);
        R[h] = herm_R(t, u, v, L, pw, gN);
      }
      __syncthreads();
      for (int r = tid; r < H * n2n; r += NT) {
        const int h = r / n2n, n2 = r - h * n2n;
        int t, u, v;
        herm_tuv(h, t, u, v);
        const double* f = Ft + n2 * H;
        double s = 0.0;
        for (int h2 = 0; h2 < H; h2++) {
          int t2, u2, v2;
          herm_tuv(h2, t2, u2, v2);
          s += R[herm_index(t + t2, u + u2, v + v2)] * f[h2];
        }
        V[r] = s;
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < K::OPT; r++) {
        const int o = tid + r * NT;
        if (o < NAB * n2n) {
          const int n1 = o / n2n, n2 = o - n1 * n2n;
          const double* e = Et + n1 * H;
          double s = 0.0;
          for (int h = 0; h < H; h++) s += e[h] * V[h * n2n + n2];
          acc[r] += s;
        }
      }
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int r = 0; r < K::OPT; r++) ss += acc[r] * acc[r];
  __syncthreads();
  red[tid] = ss;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) part[(size_t)pidx * K::NCH + chunk] = red[0];
}

struct Schwarz4Launch {
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, hipStream_t);
  int nch;
};

template <int LA, int LB>
void launch_schwarz4c(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph,
                      const double* boys, double* part, hipStream_t st) {
  using K = Cls4<LA, LB>;
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz4c_kernel<LA, LB>), dim3(npairs, K::NCH), dim3(K::NT), 0, st, bra, s1, s2, sph, boys,
                     part);
}

template <int LA, int LB>
Schwarz4Launch make_schwarz4() {
  return Schwarz4Launch{&launch_schwarz4c<LA, LB>, Cls4<LA, LB>::NCH};
}

using Schwarz4Table = Schwarz4Launch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip
