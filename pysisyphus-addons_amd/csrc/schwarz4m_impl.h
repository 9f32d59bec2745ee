// Schwarz norms Q_ab = sqrt(||(ab|ab)||_F) on the FP64 matrix cores
// (common/integrals/intor.f90:131-160).
//
// One workgroup = one shell pair with npp <= NPPM primitive pairs.  With the
// spherical Hermite matrices Et_pq (NAB x H, contraction coefficients folded
// in) and M_{pq,rs}[h][h2] = (-1)^{|h2|} R_{h+h2}(P_pq - P_rs), the block is
//     X = sum_{pq,rs} Et_pq M_{pq,rs} Et_rs^T.
// Because M_{rs,pq} = M_{pq,rs}^T, X = Z + Z^T with
//     Z = sum_pq Et_pq V_pq,   V_pq = sum_{rs >= pq} w M_{pq,rs} Et_rs^T,
// w = 1/2 on the diagonal, so only the npp(npp+1)/2 quartets pq <= rs are
// evaluated.  Both products run on v_mfma_f64_16x16x4:
//   * V: rows (pq, h), columns n2, reduction over (rs, h2); the A operand is
//     gathered from the R table of the quartet, the B operand is the table
//     F[(rs, h2)][n2] = Et_rs[n2][h2];
//   * Z: the V accumulator tile feeds the B operand directly: register r of
//     lane l holds V[l/16 + 4 r][l % 16], which is exactly the B fragment of
//     a K=4 step over the rows {4r, 4r+1, 4r+2, 4r+3} of the tile.
// Row tiles are dealt to the waves in snake order (the triangle makes low
// tiles longer); each wave adds its partial Z into LDS, and the workgroup
// writes sum_{n1,n2} (Z + Z^T)^2; the host takes the fourth root.
#pragma once

#include "gen3c_common.h"
#include "hermite.h"
#include "pair_bound.h"

namespace dfhip {

typedef double dbl4 __attribute__((ext_vector_type(4)));

template <int LA, int LB>
struct Cls4m {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L);
  // R stride per quartet: R[NR] = 0 pads h2 >= H; odd, so that lanes on
  // consecutive quartets hit distinct banks
  static constexpr int NR1 = (NR + 1) | 1;
  // odd levels of the R recursion (orders <= L-1) live in F's space before F is built
  static constexpr int NRT = nherm(L - 1) | 1;
  static constexpr int HK = (H + 3) / 4 * 4, NS = HK / 4;              // K steps per ket primitive pair
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NCT = (NAB + 15) / 16, ZL = 16 * NCT;
  static constexpr int NPPM = 9, NQM = NPPM * (NPPM + 1) / 2;
  static constexpr int NRTL = (NPPM * H + 15) / 16;  // row tiles of V
  static constexpr int NW = NRTL >= 16 ? 8 : (NRTL >= 8 ? 4 : (NRTL >= 4 ? 2 : 1));
  static constexpr int NT = 64 * NW;
  static constexpr int LH = L / 2 + 1;
  // F rows read: < 16 NRTL + HK, columns < ZL past the row start
  static constexpr int FSZ0 = (16 * NRTL + HK) * NAB + ZL;
  static constexpr int FSZ = FSZ0 > NQM * NRT ? FSZ0 : NQM * NRT;
  static constexpr int RSZ = NQM * NR1;
  // scaled Boys values and (2 PQ_d)^e per quartet; odd stride
  static constexpr int GW = 4 * (L + 1) + 1;
  static constexpr int SCR0 = NPPM * 3 * ESZ + NQM * GW;
  static constexpr int SCR = SCR0 > ZL * ZL ? SCR0 : ZL * ZL;
  static constexpr int LDS = FSZ + RSZ + SCR + NR;  // + tuv and rec
  static constexpr bool OK = LDS <= 20000;
};

template <int LA, int LB>
__device__ inline void herm_E_dir_reg(double a, double b, double A, double B, double* __restrict__ out) {
  constexpr int NT = LA + LB + 1, ESZ = (LA + 1) * (LB + 1) * NT;
  const double p = a + b, oo2p = 0.5 / p;
  const double P = (a * A + b * B) / p;
  const double XPA = P - A, XPB = P - B, X = A - B;
  double E[ESZ];
#pragma unroll
  for (int k = 0; k < ESZ; k++) E[k] = 0.0;
  E[0] = exp(-a * b / p * X * X);
#pragma unroll
  for (int i = 0; i <= LA; i++)
#pragma unroll
    for (int j = 0; j <= LB; j++) {
      if (i == 0 && j == 0) continue;
      const int e = (i * (LB + 1) + j) * NT;
      const int s = (i > 0) ? ((i - 1) * (LB + 1) + j) * NT : (i * (LB + 1) + j - 1) * NT;
      const double xp = (i > 0) ? XPA : XPB;
      const int smax = i + j - 1;
#pragma unroll
      for (int t = 0; t <= i + j; t++) {
        double v = 0.0;
        if (t > 0) v += oo2p * E[s + t - 1];
        if (t <= smax) v += xp * E[s + t];
        if (t + 1 <= smax) v += (t + 1) * E[s + t + 1];
        E[e + t] = v;
      }
    }
#pragma unroll
  for (int k = 0; k < ESZ; k++) out[k] = E[k];
}

template <int LA, int LB>
__global__ __launch_bounds__((Cls4m<LA, LB>::NT)) void schwarz4m_kernel(DevShells bra, const int* __restrict__ s1v,
                                                                      const int* __restrict__ s2v,
                                                                      const SphTable* __restrict__ sph,
                                                                      const double* __restrict__ boys,
                                                                      double* __restrict__ part, const PairSkip sk) {
  using K = Cls4m<LA, LB>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NR1 = K::NR1, NAB = K::NAB, NB = K::NB, NT = K::NT, NW = K::NW;
  constexpr int NCT = K::NCT, ZL = K::ZL, NS = K::NS, GW = K::GW, ESZ = K::ESZ;
  __shared__ double F[K::FSZ];
  __shared__ double R[K::RSZ];
  __shared__ double scr[K::SCR];
  __shared__ int tuv[NR];
  __shared__ int rec[NR];  // recursion of index h: parents h1, h2, direction, coefficient
  __shared__ double red[NW];
  __shared__ int sa[2 * LA + 2], sb[2 * LB + 2];
  __shared__ SphTerm ta[nsph(LA) * ncart(LA)], tb[nsph(LB) * ncart(LB)];

  const int tid = threadIdx.x, pair = blockIdx.x;
  const int s1 = s1v[pair], s2 = s2v[pair];
  // a pair whose bound on Q_ab keeps no triplet: zero sum, no quartets
  if (pair_negligible_wave(bra, s1, s2, sk, tid & 63)) {
    if (tid == 0) part[pair] = 0.0;
    return;
  }
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb, nq = npp * (npp + 1) / 2;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  double* Esc = scr;                    // [pq][d][ESZ]
  double* G = scr + K::NPPM * 3 * ESZ;  // [q][GW]: scale (-alpha)^n F_n, then (2 PQ_d)^e

  for (int h = tid; h < NR; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
  }
  {
    const int a0 = sph->start[LA][0], b0 = sph->start[LB][0];
    const int na = sph->start[LA][2 * LA + 1] - a0, nb = sph->start[LB][2 * LB + 1] - b0;
    if (tid < 2 * LA + 2) sa[tid] = sph->start[LA][tid] - a0;
    if (tid < 2 * LB + 2) sb[tid] = sph->start[LB][tid] - b0;
    for (int e = tid; e < na; e += NT) ta[e] = sph->terms[a0 + e];
    for (int e = tid; e < nb; e += NT) tb[e] = sph->terms[b0 + e];
  }
  // R^(n)_tuv = 2 (t-1) R^(n+1)_{t-2,u,v} + 2 PQ_x R^(n+1)_{t-1,u,v} along the first
  // non-zero index (in the scale of G: R^(n)_000 = g_n)
  for (int h = 1 + tid; h < NR; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    int dir, h1, h2, cf;
    if (t > 0) {
      dir = 0, cf = t - 1, h1 = herm_index(t - 1, u, v), h2 = t > 1 ? herm_index(t - 2, u, v) : 0;
    } else if (u > 0) {
      dir = 1, cf = u - 1, h1 = herm_index(t, u - 1, v), h2 = u > 1 ? herm_index(t, u - 2, v) : 0;
    } else {
      dir = 2, cf = v - 1, h1 = herm_index(t, u, v - 1), h2 = v > 1 ? herm_index(t, u, v - 2) : 0;
    }
    rec[h] = h1 | (h2 << 10) | (dir << 20) | (cf << 22);
  }
  for (int e = tid; e < npp * 3 + nq; e += NT) {
    if (e < npp * 3) {
      const int pq = e / 3, d = e - 3 * pq, i = pq / npb, j = pq - i * npb;
      herm_E_dir_reg<LA, LB>(bra.exps[pa0 + i], bra.exps[pb0 + j], bra.xyz[3 * s1 + d], bra.xyz[3 * s2 + d],
                             Esc + (pq * 3 + d) * ESZ);
    } else {
      const int q = e - npp * 3;
      int rs = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
      while (rs * (rs + 1) / 2 > q) rs--;
      while ((rs + 1) * (rs + 2) / 2 <= q) rs++;
      const int pq = q - rs * (rs + 1) / 2;
      const int i = pq / npb, j = pq - i * npb, k = rs / npb, l = rs - k * npb;
      const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], c = bra.exps[pa0 + k], d = bra.exps[pb0 + l];
      const double p = a + b, qq = c + d;
      double PQ[3];
#pragma unroll
      for (int z = 0; z < 3; z++) {
        const double Az = bra.xyz[3 * s1 + z], Bz = bra.xyz[3 * s2 + z];
        PQ[z] = (a * Az + b * Bz) / p - (c * Az + d * Bz) / qq;
      }
      const double alpha = p * qq / (p + qq);
      const double T = alpha * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
      double Fb[L + 1];
      boys_array<L>(T, boys, Fb);
      double* g = G + q * GW;
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * qq * sqrt(p + qq));
#pragma unroll
      for (int n = 0; n <= L; n++) {
        g[n] = sc * Fb[n];
        sc *= -alpha;
      }
#pragma unroll
      for (int z = 0; z < 3; z++) {
        const double x = 2.0 * PQ[z];
        double v = 1.0;
#pragma unroll
        for (int n = 0; n <= L; n++) {
          g[(L + 1) * (1 + z) + n] = v;
          v *= x;
        }
      }
    }
  }
  __syncthreads();
  // R^(n) by levels n = L .. 0 (L is even: level n in R if n is even, else in
  // F's space), quartets on lanes; every entry of a level is independent
  {
    double* Rt = F;
#pragma unroll
    for (int n = L; n >= 0; n--) {
      const double* So = (n % 2 == 0) ? Rt : R;  // level n+1
      double* Sn = (n % 2 == 0) ? R : Rt;
      const int so = (n % 2 == 0) ? K::NRT : NR1, sn = (n % 2 == 0) ? NR1 : K::NRT;
      const int nent = nherm(L - n);
      for (int e = tid; e < nent * K::NQM; e += NT) {
        const int h = e / K::NQM, q = e - h * K::NQM;  // compile-time divisor
        if (q >= nq) continue;
        const double* g = G + q * GW;
        double val;
        if (h == 0) {
          val = g[n];
        } else {
          const int rc = rec[h];
          const int h1 = rc & 1023, h2 = (rc >> 10) & 1023, dir = (rc >> 20) & 3, cf = rc >> 22;
          const double* s0 = So + q * so;
          val = fma((double)(2 * cf), s0[h2], g[(L + 1) * (1 + dir) + 1] * s0[h1]);
        }
        Sn[q * sn + h] = val;
      }
      __syncthreads();
    }
    for (int q = tid; q < nq; q += NT) R[q * NR1 + NR] = 0.0;  // the h2 >= H pad
  }
  // F[(pq H + h) NAB + n]: Hermite index on lanes, so that lanes share the
  // solid-harmonic term loops of (mu, nu); rows past npp H stay zero
  for (int e = tid; e < K::FSZ; e += NT) {
    int f = e;
    if (e < npp * H * NAB) {
      const int pq = e / (H * NAB), r = e - pq * (H * NAB), n = r / H, h = r - n * H;
      f = (pq * H + h) * NAB + n;
      const int i = pq / npb, j = pq - i * npb, tv = tuv[h];
      const double* E = Esc + pq * 3 * ESZ;
      F[f] = bra.coefs[pa0 + i] * bra.coefs[pb0 + j] *
             bra_element_lds<LA, LB>(sa, ta, sb, tb, n / NB, n % NB, tv & 255, (tv >> 8) & 255, tv >> 16, E, E + ESZ,
                                     E + 2 * ESZ);
    } else {
      F[f] = 0.0;
    }
  }
  __syncthreads();
  for (int e = tid; e < ZL * ZL; e += NT) scr[e] = 0.0;  // Z (E tables and G are dead)

  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int nrt = (npp * H + 15) / 16;
  dbl4 zacc[NCT][NCT];
#pragma unroll
  for (int a = 0; a < NCT; a++)
#pragma unroll
    for (int j = 0; j < NCT; j++) zacc[a][j] = dbl4{0.0, 0.0, 0.0, 0.0};
  const int nround = (nrt + NW - 1) / NW;
  for (int m = 0; m < nround; m++) {
    const int i = (m & 1) ? (m + 1) * NW - 1 - wave : m * NW + wave;
    if (i >= nrt) continue;
    const int row = 16 * i + lr;
    const int pqr = row / H, hr = row - pqr * H;
    const int tvr = tuv[hr];
    const int tr = tvr & 255, ur = (tvr >> 8) & 255, vr = tvr >> 16;
    int roff[NS];
    double sgn[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int h2 = 4 * s + lk;
      if (h2 < H) {
        const int tv2 = tuv[h2];
        const int t2 = tv2 & 255, u2 = (tv2 >> 8) & 255, v2 = tv2 >> 16;
        roff[s] = herm_index(tr + t2, ur + u2, vr + v2);
        sgn[s] = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
      } else {
        roff[s] = NR;
        sgn[s] = 0.0;
      }
    }
    dbl4 vacc[NCT];
#pragma unroll
    for (int j = 0; j < NCT; j++) vacc[j] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int rs = (16 * i) / H; rs < npp; rs++) {
      const double w = (pqr < npp && rs >= pqr) ? (rs == pqr ? 0.5 : 1.0) : 0.0;
      const double* Rq = R + (rs * (rs + 1) / 2 + min(pqr, rs)) * NR1;
      const double* Fb = F + (rs * H + lk) * NAB + lr;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const double av = (w * sgn[s]) * Rq[roff[s]];
#pragma unroll
        for (int j = 0; j < NCT; j++)
          vacc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Fb[4 * s * NAB + 16 * j], vacc[j], 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double* Fz = F + (16 * i + lk + 4 * r) * NAB + lr;
#pragma unroll
      for (int a = 0; a < NCT; a++) {
        const double ea = Fz[16 * a];
#pragma unroll
        for (int j = 0; j < NCT; j++) zacc[a][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ea, vacc[j][r], zacc[a][j], 0, 0, 0);
      }
    }
  }
  __syncthreads();
  // the waves add their partial Z one after another (a fixed order, so Q_ab
  // is reproducible bit for bit)
  for (int w = 0; w < NW; w++) {
    if (wave == w) {
#pragma unroll
      for (int a = 0; a < NCT; a++)
#pragma unroll
        for (int j = 0; j < NCT; j++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            double* z = &scr[(16 * a + lk + 4 * r) * ZL + 16 * j + lr];
            if (w == 0) *z = zacc[a][j][r];
            else *z += zacc[a][j][r];
          }
    }
    __syncthreads();
  }
  double ss = 0.0;
  for (int e = tid; e < NAB * NAB; e += NT) {
    const int n1 = e / NAB, n2 = e - n1 * NAB;
    const double x = scr[n1 * ZL + n2] + scr[n2 * ZL + n1];
    ss = fma(x, x, ss);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ss += __shfl_down(ss, off);
  if (lane == 0) red[wave] = ss;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; w++) s += red[w];
    part[pair] = s;
  }
}

template <int LA, int LB>
void launch_schwarz4m(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph,
                      const double* boys, double* part, const PairSkip& sk, hipStream_t st) {
  using K = Cls4m<LA, LB>;
  if (npairs <= 0) return;
  hipLaunchKernelGGL((schwarz4m_kernel<LA, LB>), dim3(npairs), dim3(K::NT), 0, st, bra, s1, s2, sph, boys, part, sk);
}

struct Schwarz4MLaunch {
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, const PairSkip&,
             hipStream_t);
  int nppm;  // largest primitive-pair count handled; 0 = class not instantiated
};

template <int LA, int LB>
Schwarz4MLaunch make_schwarz4m() {
  if constexpr (Cls4m<LA, LB>::OK)
    return Schwarz4MLaunch{&launch_schwarz4m<LA, LB>, Cls4m<LA, LB>::NPPM};
  else
    return Schwarz4MLaunch{nullptr, 0};
}

using Schwarz4MTable = Schwarz4MLaunch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip
