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
#include "pair_bound.h"

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
                                                        const double* __restrict__ boys, double* __restrict__ part,
                                                        const PairSkip sk) {
  using K = Cls4<LA, LB>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NAB = K::NAB, NT = K::NT, NB2 = K::NB2;
  __shared__ double lds[K::LDS];
  const int pidx = blockIdx.x, chunk = blockIdx.y;
  const int s1 = s1v[pidx], s2 = s2v[pidx];
  const int n2lo = chunk * NB2, n2n = min(NB2, NAB - n2lo);
  const int tid = threadIdx.x;
  // a pair whose bound on Q_ab keeps no triplet: zero sum, no quartets
  if (pair_negligible_wave(bra, s1, s2, sk, tid & 63)) {
    if (tid == 0) part[(size_t)pidx * K::NCH + chunk] = 0.0;
    return;
  }

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
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, const PairSkip&,
             hipStream_t);
  int nch;
};

template <int LA, int LB>
void launch_schwarz4c(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph,
                      const double* boys, double* part, const PairSkip& sk, hipStream_t st) {
  using K = Cls4<LA, LB>;
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz4c_kernel<LA, LB>), dim3(npairs, K::NCH), dim3(K::NT), 0, st, bra, s1, s2, sph, boys,
                     part, sk);
}

template <int LA, int LB>
Schwarz4Launch make_schwarz4() {
  return Schwarz4Launch{&launch_schwarz4c<LA, LB>, Cls4<LA, LB>::NCH};
}

using Schwarz4Table = Schwarz4Launch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip

namespace dfhip {

// Faster variant for shell pairs whose primitive-pair count fits the LDS
// cache.  PPW shell pairs share a workgroup.  The spherical Hermite matrices
// Et[ij] of all primitive pairs are built once per shell pair.  For each bra
// primitive pair ij, all ket primitive pairs kl are processed together:
//     Vs = sum_kl M(ij,kl) Et[kl]^T,   M[h][h2] = (-1)^{|h2|} R_{h+h2}(P_ij - P_kl)
//     acc += Et[ij] Vs
// so the second product runs once per ij instead of once per quartet.  The
// ket sum is split over NG groups of kl (partial Vp summed into Vs), and both
// products use 4x4 register blocks.
template <int LA, int LB>
struct Cls4c {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L);
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NT = (NAB >= 25) ? 512 : 256;
  static constexpr int NG = 4;
  static constexpr int NQ = (NAB + 3) / 4, NA4 = 4 * NQ;  // n quads
  static constexpr int HQ = (H + 3) / 4;                   // h quads
  static constexpr int HP = H + 1;                         // padded Et row
  static constexpr int BUDGET = 20000;                     // doubles
  static constexpr int NPPC_MAX = 9;
  // per cached primitive pair: Et rows + (union of E tables | gN, pw, R)
  static constexpr int PER_KL = 4 * (L + 1) + NR;
  static constexpr int UNION = 3 * ESZ > PER_KL ? 3 * ESZ : PER_KL;
  static constexpr int PERPP = NAB * HP + UNION;
  static constexpr int SHARED = NT + (H * H + 3) / 4 + H;
  static constexpr int PERPAIR0 = NG * H * NA4 + H * NA4;
  static constexpr int NPPC_FIT = (BUDGET - SHARED - PERPAIR0) / PERPP;
  static constexpr int NPPC = NPPC_FIT < NPPC_MAX ? NPPC_FIT : NPPC_MAX;
  static constexpr int PERPAIR = PERPAIR0 + (NPPC > 0 ? NPPC : 0) * PERPP;
  // shell pairs per workgroup: fill ~64 KB of LDS, at most 8
  static constexpr int PPW0 = (8192 - SHARED) / PERPAIR;
  static constexpr int PPW = PPW0 < 1 ? 1 : (PPW0 > 8 ? 8 : PPW0);
  static constexpr int LDS = SHARED + PPW * PERPAIR;
  static constexpr int NBLK = NQ * NQ;                     // 4x4 output blocks per pair
  static constexpr int OPT = (PPW * NBLK + NT - 1) / NT;
};

template <int LA, int LB>
__global__ __launch_bounds__((Cls4c<LA, LB>::NT)) void schwarz4c_cached_kernel(
    DevShells bra, const int* __restrict__ s1v, const int* __restrict__ s2v, int npairs,
    const SphTable* __restrict__ sph, const double* __restrict__ boys, double* __restrict__ part) {
  using K = Cls4c<LA, LB>;
  constexpr int L = K::L, H = K::H, HP = K::HP, NR = K::NR, NAB = K::NAB, NT = K::NT, NQ = K::NQ;
  constexpr int NA4 = K::NA4, HQ = K::HQ, NG = K::NG, NPPC = K::NPPC, PPW = K::PPW, UNION = K::UNION;
  __shared__ double lds[K::LDS];
  __shared__ int pinfo[PPW][4];  // s1, s2, npp, npb
  const int tid = threadIdx.x;
  const int pbase = blockIdx.x * PPW;
  const int nloc = min(PPW, npairs - pbase);

  double* red = lds;                                    // NT
  short* idx2 = (short*)(red + NT);                     // H*H
  double* sg = (double*)(idx2 + 4 * ((H * H + 3) / 4));  // H
  double* pairmem = sg + H;                             // PPW * PERPAIR
  // per pair q: Vp [NG][H][NA4], Vs [H][NA4], Et [NPPC][NAB][HP], scr [NPPC][UNION]
  auto Vp = [&](int q) { return pairmem + q * K::PERPAIR; };
  auto Vs = [&](int q) { return pairmem + q * K::PERPAIR + NG * H * NA4; };
  auto Etp = [&](int q) { return pairmem + q * K::PERPAIR + K::PERPAIR0; };
  auto scrp = [&](int q) { return pairmem + q * K::PERPAIR + K::PERPAIR0 + NPPC * NAB * HP; };

  if (tid < nloc) {
    const int s1 = s1v[pbase + tid], s2 = s2v[pbase + tid];
    pinfo[tid][0] = s1;
    pinfo[tid][1] = s2;
    pinfo[tid][2] = bra.nprim[s1] * bra.nprim[s2];
    pinfo[tid][3] = bra.nprim[s2];
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
  __syncthreads();
  int nppmax = 0;
  for (int q = 0; q < nloc; q++) nppmax = max(nppmax, pinfo[q][2]);
  // E tables of every primitive pair of every shell pair
  for (int e = tid; e < nloc * NPPC * 3; e += NT) {
    const int q = e / (NPPC * 3), r = e - q * (NPPC * 3), ij = r / 3, d = r - ij * 3;
    if (ij >= pinfo[q][2]) continue;
    const int s1 = pinfo[q][0], s2 = pinfo[q][1], npb = pinfo[q][3];
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    herm_E_dir<LA, LB>(bra.exps[bra.pptr[s1] + i], bra.exps[bra.pptr[s2] + j], bra.xyz[3 * s1 + d],
                       bra.xyz[3 * s2 + d], scrp(q) + ij * UNION + d * K::ESZ);
  }
  __syncthreads();
  for (int e = tid; e < nloc * NPPC * NAB * H; e += NT) {
    const int q = e / (NPPC * NAB * H), r0 = e - q * (NPPC * NAB * H);
    const int ij = r0 / (NAB * H), r = r0 - ij * (NAB * H), n = r / H, h = r - n * H;
    if (ij >= pinfo[q][2]) continue;
    const int s1 = pinfo[q][0], s2 = pinfo[q][1], npb = pinfo[q][3];
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    int t, u, v;
    herm_tuv(h, t, u, v);
    const double* E = scrp(q) + ij * UNION;
    Etp(q)[(ij * NAB + n) * HP + h] = bra.coefs[bra.pptr[s1] + i] * bra.coefs[bra.pptr[s2] + j] *
                                      bra_element<LA, LB>(sph, n / K::NB, n % K::NB, t, u, v, E, E + K::ESZ,
                                                          E + 2 * K::ESZ);
  }

  double acc[K::OPT][16];
#pragma unroll
  for (int r = 0; r < K::OPT; r++)
#pragma unroll
    for (int z = 0; z < 16; z++) acc[r][z] = 0.0;

  for (int ij = 0; ij < nppmax; ij++) {
    __syncthreads();  // Et complete (first ij); previous ij done with scr, Vp, Vs
    // scaled Boys values and powers of 2(P-Q) for every ket primitive pair
    for (int e = tid; e < nloc * NPPC * 4 * (L + 1); e += NT) {
      const int q = e / (NPPC * 4 * (L + 1)), r0 = e - q * (NPPC * 4 * (L + 1));
      const int kl = r0 / (4 * (L + 1)), q4 = r0 - kl * (4 * (L + 1));
      const int npp = pinfo[q][2];
      if (ij >= npp || kl >= npp) continue;
      const int s1 = pinfo[q][0], s2 = pinfo[q][1], npb = pinfo[q][3];
      const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
      const int i = ij / npb, j = ij - (ij / npb) * npb, k = kl / npb, l = kl - (kl / npb) * npb;
      const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], p = a + b;
      const double c = bra.exps[pa0 + k], d = bra.exps[pb0 + l], qq_ = c + d;
      double* gN = scrp(q) + kl * UNION;
      double* pw = gN + (L + 1);
      double PQ[3];
      for (int z = 0; z < 3; z++) {
        const double Az = bra.xyz[3 * s1 + z], Bz = bra.xyz[3 * s2 + z];
        PQ[z] = (a * Az + b * Bz) / p - (c * Az + d * Bz) / qq_;
      }
      const double alpha = p * qq_ / (p + qq_);
      if (q4 <= L) {
        const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
        double scale = 2.0 * PI * PI * sqrt(PI) / (p * qq_ * sqrt(p + qq_));
        for (int z = 0; z < q4; z++) scale *= -alpha;
        gN[q4] = scale * boys_f(q4, T, boys);
      } else {
        const int qq = q4 - (L + 1), dd = qq / (L + 1), ee = qq - dd * (L + 1);
        const double x = 2.0 * PQ[dd];
        double v = 1.0;
        for (int z = 0; z < ee; z++) v *= x;
        pw[dd * (L + 1) + ee] = v;
      }
    }
    __syncthreads();
    for (int e = tid; e < nloc * NPPC * NR; e += NT) {
      const int q = e / (NPPC * NR), r0 = e - q * (NPPC * NR), kl = r0 / NR, h = r0 - kl * NR;
      const int npp = pinfo[q][2];
      if (ij >= npp || kl >= npp) continue;
      double* gN = scrp(q) + kl * UNION;
      int t, u, v;
      herm_tuv(h, t, u, v);
      gN[4 * (L + 1) + h] = herm_R(t, u, v, L, gN + (L + 1), gN);
    }
    __syncthreads();
    // Vp[g][h][n2] = sum_{kl in g} sum_h2 sg[h2] R_kl[h+h2] Et[kl][n2][h2]  (4x4 blocks)
    for (int e = tid; e < nloc * NG * HQ * NQ; e += NT) {
      const int q = e / (NG * HQ * NQ), r0 = e - q * (NG * HQ * NQ);
      const int g = r0 / (HQ * NQ), r = r0 - g * (HQ * NQ), hq = r / NQ, nq = r - hq * NQ;
      const int npp = pinfo[q][2];
      if (ij >= npp) continue;
      double c16[16];
#pragma unroll
      for (int z = 0; z < 16; z++) c16[z] = 0.0;
      int hrow[4], nrow[4];
#pragma unroll
      for (int z = 0; z < 4; z++) {
        hrow[z] = min(4 * hq + z, H - 1);
        nrow[z] = min(4 * nq + z, NAB - 1);
      }
      for (int kl = g; kl < npp; kl += NG) {
        const double* Rk = scrp(q) + kl * UNION + 4 * (L + 1);
        const double* Ek = Etp(q) + kl * NAB * HP;
        for (int h2 = 0; h2 < H; h2++) {
          const double s2g = sg[h2];
          double rr[4], ee[4];
#pragma unroll
          for (int z = 0; z < 4; z++) {
            rr[z] = s2g * Rk[idx2[hrow[z] * H + h2]];
            ee[z] = Ek[nrow[z] * HP + h2];
          }
#pragma unroll
          for (int x = 0; x < 4; x++)
#pragma unroll
            for (int y = 0; y < 4; y++) c16[x * 4 + y] = fma(rr[x], ee[y], c16[x * 4 + y]);
        }
      }
      double* vp = Vp(q);
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const int h = 4 * hq + x;
        if (h < H) {
#pragma unroll
          for (int y = 0; y < 4; y++) vp[(g * H + h) * NA4 + 4 * nq + y] = c16[x * 4 + y];
        }
      }
    }
    __syncthreads();
    for (int e = tid; e < nloc * H * NA4; e += NT) {
      const int q = e / (H * NA4), r = e - q * (H * NA4);
      if (ij >= pinfo[q][2]) continue;
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < NG; g++) v += Vp(q)[g * H * NA4 + r];
      Vs(q)[r] = v;
    }
    __syncthreads();
    // acc[n1][n2] += sum_h Et[ij][n1][h] Vs[h][n2]  (4x4 blocks)
#pragma unroll
    for (int r = 0; r < K::OPT; r++) {
      const int o = tid + r * NT;
      if (o < nloc * K::NBLK) {
        const int q = o / K::NBLK, ob = o - q * K::NBLK;
        if (ij < pinfo[q][2]) {
          const int q1 = ob / NQ, q2 = ob - q1 * NQ;
          const double* Ei = Etp(q) + ij * NAB * HP;
          const double* vs = Vs(q);
          int n1[4];
#pragma unroll
          for (int z = 0; z < 4; z++) n1[z] = min(4 * q1 + z, NAB - 1);
          for (int h = 0; h < H; h++) {
            double ei[4], vv[4];
#pragma unroll
            for (int z = 0; z < 4; z++) {
              ei[z] = Ei[n1[z] * HP + h];
              vv[z] = vs[h * NA4 + 4 * q2 + z];
            }
#pragma unroll
            for (int x = 0; x < 4; x++)
#pragma unroll
              for (int y = 0; y < 4; y++) acc[r][x * 4 + y] = fma(ei[x], vv[y], acc[r][x * 4 + y]);
          }
        }
      }
    }
  }
  // per-pair sums of squares
  for (int q = 0; q < nloc; q++) {
    double ss = 0.0;
#pragma unroll
    for (int r = 0; r < K::OPT; r++) {
      const int o = tid + r * NT;
      if (o < nloc * K::NBLK && o / K::NBLK == q) {
        const int ob = o - q * K::NBLK, q1 = ob / NQ, q2 = ob - q1 * NQ;
#pragma unroll
        for (int x = 0; x < 4; x++)
#pragma unroll
          for (int y = 0; y < 4; y++)
            if (4 * q1 + x < NAB && 4 * q2 + y < NAB) ss += acc[r][x * 4 + y] * acc[r][x * 4 + y];
      }
    }
    __syncthreads();
    red[tid] = ss;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    if (tid == 0) part[pbase + q] = red[0];
  }
}

template <int LA, int LB>
void launch_schwarz4c_cached(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph,
                             const double* boys, double* part, hipStream_t st) {
  using K = Cls4c<LA, LB>;
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz4c_cached_kernel<LA, LB>), dim3((npairs + K::PPW - 1) / K::PPW), dim3(K::NT), 0, st,
                     bra, s1, s2, npairs, sph, boys, part);
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
