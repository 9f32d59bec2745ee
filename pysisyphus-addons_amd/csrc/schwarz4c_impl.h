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

namespace dfhip {

// Faster variant for shell pairs whose primitive-pair count fits the LDS
// cache: the spherical Hermite matrices Et[ij] of all primitive pairs are
// built once per shell pair, the R-gather uses a precomputed index table,
// and both small GEMMs per primitive quartet are register-blocked by 4.
template <int LA, int LB>
struct Cls4c {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L);
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NT = 256;
  static constexpr int NQ = (NAB + 3) / 4;              // n2 quads
  static constexpr int HP = H + 1;                       // padded Et row
  static constexpr int BUDGET = 20000;                   // doubles
  static constexpr int FIXED = 4 * (L + 1) + NR + H * NQ * 4 + NT + (H * H + 3) / 4 + H;
  static constexpr int PERPP = NAB * HP + 3 * ESZ;
  static constexpr int NPPC_FIT = (BUDGET - FIXED) / PERPP;
  static constexpr int NPPC = NPPC_FIT < 9 ? NPPC_FIT : 9;  // cached primitive pairs
  static constexpr int LDS = FIXED + (NPPC > 0 ? NPPC : 0) * PERPP;
  static constexpr int NOUT = NQ * NAB;                  // (n2 quad, n1) work units
  static constexpr int OPT = (NOUT + NT - 1) / NT;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz4c_cached_kernel(DevShells bra, const int* __restrict__ s1v,
                                                               const int* __restrict__ s2v,
                                                               const SphTable* __restrict__ sph,
                                                               const double* __restrict__ boys,
                                                               double* __restrict__ part) {
  using K = Cls4c<LA, LB>;
  constexpr int L = K::L, H = K::H, HP = K::HP, NR = K::NR, NAB = K::NAB, NT = K::NT, NQ = K::NQ;
  __shared__ double lds[K::LDS];
  const int pidx = blockIdx.x;
  const int s1 = s1v[pidx], s2 = s2v[pidx];
  const int tid = threadIdx.x;
  const int npa = bra.nprim[s1], pa0 = bra.pptr[s1], npb = bra.nprim[s2], pb0 = bra.pptr[s2];
  const int npp = npa * npb;  // host guarantees npp <= NPPC

  double* gN = lds;
  double* pw = gN + (L + 1);
  double* R = pw + 3 * (L + 1);
  double* V = R + NR;                       // [h][4*NQ]
  double* red = V + H * NQ * 4;
  short* idx2 = (short*)(red + NT);         // H*H
  double* sg = (double*)(idx2 + 4 * ((H * H + 3) / 4));  // H
  double* Etab = sg + H;                    // NPPC * 3 * ESZ
  double* Et = Etab + K::NPPC * 3 * K::ESZ; // NPPC * NAB * HP

  double A[3], B[3];
  for (int d = 0; d < 3; d++) {
    A[d] = bra.xyz[3 * s1 + d];
    B[d] = bra.xyz[3 * s2 + d];
  }
  for (int e = tid; e < H * H; e += NT) {
    const int h = e / H, h2 = e - h * H;
    int t, u, v, t2, u2, v2;
    herm_tuv(h, t, u, v);
    herm_tuv(h2, t2, u2, v2);
    idx2[e] = (short)herm_index(t + t2, u + u2, v + v2);
  }
  for (int h = tid; h < H; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    sg[h] = ((t + u + v) & 1) ? -1.0 : 1.0;
  }
  for (int e = tid; e < npp * 3; e += NT) {
    const int ij = e / 3, d = e - ij * 3;
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    herm_E_dir<LA, LB>(bra.exps[pa0 + i], bra.exps[pb0 + j], A[d], B[d], Etab + (ij * 3 + d) * K::ESZ);
  }
  __syncthreads();
  for (int e = tid; e < npp * NAB * H; e += NT) {
    const int ij = e / (NAB * H), r = e - ij * (NAB * H), n = r / H, h = r - n * H;
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    int t, u, v;
    herm_tuv(h, t, u, v);
    const double* E = Etab + ij * 3 * K::ESZ;
    Et[(ij * NAB + n) * HP + h] = bra.coefs[pa0 + i] * bra.coefs[pb0 + j] *
                                  bra_element<LA, LB>(sph, n / K::NB, n % K::NB, t, u, v, E, E + K::ESZ,
                                                      E + 2 * K::ESZ);
  }

  double acc[K::OPT][4];
#pragma unroll
  for (int r = 0; r < K::OPT; r++)
#pragma unroll
    for (int q = 0; q < 4; q++) acc[r][q] = 0.0;

  for (int ij = 0; ij < npp; ij++) {
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], p = a + b;
    for (int kl = 0; kl < npp; kl++) {
      const int k = kl / npb, l = kl - (kl / npb) * npb;
      const double c = bra.exps[pa0 + k], d = bra.exps[pb0 + l], q = c + d;
      __syncthreads();  // Et complete / previous quartet done with R and V
      if (tid < 4 * (L + 1)) {
        double PQ[3];
        for (int z = 0; z < 3; z++) PQ[z] = (a * A[z] + b * B[z]) / p - (c * A[z] + d * B[z]) / q;
        const double alpha = p * q / (p + q);
        if (tid <= L) {
          const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
          double scale = 2.0 * PI * PI * sqrt(PI) / (p * q * sqrt(p + q));
          for (int e = 0; e < tid; e++) scale *= -alpha;
          gN[tid] = scale * boys_f(tid, T, boys);
        } else {
          const int qq = tid - (L + 1), dd = qq / (L + 1), e = qq - dd * (L + 1);
          const double x = 2.0 * PQ[dd];
          double v = 1.0;
          for (int z = 0; z < e; z++) v *= x;
          pw[dd * (L + 1) + e] = v;
        }
      }
      __syncthreads();
      for (int h = tid; h < NR; h += NT) {
        int t, u, v;
        herm_tuv(h, t, u, v);
        R[h] = herm_R(t, u, v, L, pw, gN);
      }
      __syncthreads();
      // V[h][n2] = sum_h2 sg[h2] R[h + h2] Et[kl][n2][h2], 4 consecutive n2 per item
      const double* Ek = Et + kl * NAB * HP;
      for (int e = tid; e < H * NQ; e += NT) {
        const int h = e / NQ, nq = e - h * NQ;
        const short* ix = idx2 + h * H;
        double s0 = 0.0, s1v_ = 0.0, s2v_ = 0.0, s3 = 0.0;
        const int n0 = 4 * nq;
        const double* e0 = Ek + min(n0, NAB - 1) * HP;
        const double* e1 = Ek + min(n0 + 1, NAB - 1) * HP;
        const double* e2 = Ek + min(n0 + 2, NAB - 1) * HP;
        const double* e3 = Ek + min(n0 + 3, NAB - 1) * HP;
        for (int h2 = 0; h2 < H; h2++) {
          const double r = sg[h2] * R[ix[h2]];
          s0 = fma(r, e0[h2], s0);
          s1v_ = fma(r, e1[h2], s1v_);
          s2v_ = fma(r, e2[h2], s2v_);
          s3 = fma(r, e3[h2], s3);
        }
        double* vv = V + h * (4 * NQ) + n0;
        vv[0] = s0;
        vv[1] = s1v_;
        vv[2] = s2v_;
        vv[3] = s3;
      }
      __syncthreads();
      // acc[n1][n2..n2+3] += Et[ij][n1] . V[.][n2..n2+3]
      const double* Ei = Et + ij * NAB * HP;
#pragma unroll
      for (int r = 0; r < K::OPT; r++) {
        const int o = tid + r * NT;
        if (o < K::NOUT) {
          const int n1 = o / NQ, nq = o - n1 * NQ;
          const double* e = Ei + n1 * HP;
          const double* vv = V + 4 * nq;
          for (int h = 0; h < H; h++) {
            const double x = e[h];
            acc[r][0] = fma(x, vv[h * 4 * NQ + 0], acc[r][0]);
            acc[r][1] = fma(x, vv[h * 4 * NQ + 1], acc[r][1]);
            acc[r][2] = fma(x, vv[h * 4 * NQ + 2], acc[r][2]);
            acc[r][3] = fma(x, vv[h * 4 * NQ + 3], acc[r][3]);
          }
        }
      }
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int r = 0; r < K::OPT; r++) {
    const int o = tid + r * NT;
    if (o < K::NOUT) {
      const int nq = o - (o / NQ) * NQ;
#pragma unroll
      for (int z = 0; z < 4; z++)
        if (4 * nq + z < NAB) ss += acc[r][z] * acc[r][z];
    }
  }
  __syncthreads();
  red[tid] = ss;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid == 0) part[pidx] = red[0];
}

template <int LA, int LB>
void launch_schwarz4c_cached(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph,
                             const double* boys, double* part, hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz4c_cached_kernel<LA, LB>), dim3(npairs), dim3(Cls4c<LA, LB>::NT), 0, st, bra, s1, s2,
                     sph, boys, part);
}

struct Schwarz4CachedLaunch {
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, hipStream_t);
  int nppc;  // largest primitive-pair count handled; 0 = not available
};

template <int LA, int LB>
Schwarz4CachedLaunch make_schwarz4_cached() {
  if constexpr (Cls4c<LA, LB>::NPPC >= 1)
    return Schwarz4CachedLaunch{&launch_schwarz4c_cached<LA, LB>, Cls4c<LA, LB>::NPPC};
  else
    return Schwarz4CachedLaunch{nullptr, 0};
}

using Schwarz4CachedTable = Schwarz4CachedLaunch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip
