// Schwarz norms on the FP64 matrix cores for the classes whose tables do not
// fit in LDS at once (common/integrals/intor.f90:131-160).  Same algebra as
// schwarz4m_impl.h (X = Z + Z^T, Z = sum_pq Et_pq V_pq, V_pq = sum_{rs >= pq}
// w M_{pq,rs} Et_rs^T, w = 1/2 on the diagonal):
//   * schwarz4s_prep_kernel writes, per pair, the spherical Hermite table
//     F[(pq, h)][n] = Et_pq[n][h] (contraction coefficients folded in) to
//     global scratch (pairs in batches that fit the scratch);
//   * schwarz4s_kernel, one workgroup of 4 waves per pair, walks the bra
//     primitive pairs pq.  V_pq (H x NAB) lives in the matrix-core
//     accumulators: each wave owns a set of 16-row tiles and all 16-column
//     tiles of them.  For every ket primitive pair rs >= pq the workgroup
//     builds R of the quartet in LDS (downward recursion by levels from the
//     scaled Boys values, which are computed for all quartets first), then
//     runs the V product: the A operand is gathered from R once per row tile
//     and K step and serves every column tile, the B operand comes from F in
//     global memory one K step ahead.  Z += Et_pq^T V_pq feeds V's
//     accumulators as the B operand (register r of lane l holds
//     V[l/16 + 4 r][l % 16] of a tile, the B fragment of the K step over its
//     rows 4r .. 4r+3); the waves add their Z tiles into LDS one after
//     another, so Q_ab is reproducible bit for bit.
#pragma once

#include "schwarz4m_impl.h"

namespace dfhip {

template <int LA, int LB>
struct Cls4s {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L);
  static constexpr int NRO = nherm(L - 1);  // entries of the odd levels of the R recursion
  static constexpr int HK = (H + 3) / 4 * 4, NS = HK / 4;
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NCT = (NAB + 15) / 16;   // column tiles of V and Z
  static constexpr int NRT1 = (H + 15) / 16;    // row tiles of V_pq
  static constexpr int NW = 4, NT = 64 * NW;
  static constexpr int TPW = (NRT1 + NW - 1) / NW;  // row tiles per wave
  static constexpr int ZC = NCT < 2 ? NCT : 2;       // Z column tiles per pass (accumulator registers)
  static constexpr int NPPS = 9, NQS = NPPS * (NPPS + 1) / 2;
  static constexpr int GW = (L + 4) | 1;  // per quartet: g_0..g_L, then 2 PQ_x, 2 PQ_y, 2 PQ_z (odd stride)
  // LDS of the main kernel in doubles: R (even levels, pad), R (odd levels), Z, G
  static constexpr int LDS = (NR + 1) + NRO + NAB * NAB + NQS * GW;
  static constexpr bool OK = TPW * NCT <= 24 && LDS * 8 + (NR + H) * 4 <= 96 * 1024;
  // global scratch per pair (doubles): the F rows of NPPS primitive pairs
  static constexpr int64_t PERPAIR = (int64_t)NPPS * H * NAB + 16;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz4s_prep_kernel(DevShells bra, const int* __restrict__ s1v,
                                                             const int* __restrict__ s2v,
                                                             const SphTable* __restrict__ sph,
                                                             double* __restrict__ scr, const PairSkip sk) {
  using K = Cls4s<LA, LB>;
  constexpr int H = K::H, NAB = K::NAB, NB = K::NB, ESZ = K::ESZ;
  constexpr int NT = 256;
  __shared__ double Esc[K::NPPS * 3 * ESZ];
  __shared__ int tuv[H];
  __shared__ int sa[2 * LA + 2], sb[2 * LB + 2];
  __shared__ SphTerm ta[nsph(LA) * ncart(LA)], tb[nsph(LB) * ncart(LB)];
  const int tid = threadIdx.x, pair = blockIdx.x;
  const int s1 = s1v[pair], s2 = s2v[pair];
  if (pair_negligible_wave(bra, s1, s2, sk, tid & 63)) return;  // the main kernel drops it too
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  double* F = scr + (size_t)pair * K::PERPAIR;

  for (int h = tid; h < H; h += NT) {
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
  for (int e = tid; e < npp * 3; e += NT) {
    const int pq = e / 3, d = e - 3 * pq, i = pq / npb, j = pq - i * npb;
    herm_E_dir<LA, LB>(bra.exps[pa0 + i], bra.exps[pb0 + j], bra.xyz[3 * s1 + d], bra.xyz[3 * s2 + d],
                       Esc + (pq * 3 + d) * ESZ);
  }
  __syncthreads();
  for (int64_t e = tid; e < (int64_t)npp * H * NAB; e += NT) {
    const int pq = (int)(e / (H * NAB)), r = (int)(e - (int64_t)pq * (H * NAB)), h = r / NAB, n = r - h * NAB;
    const int i = pq / npb, j = pq - i * npb, tv = tuv[h];
    const double* E = Esc + pq * 3 * ESZ;
    F[e] = bra.coefs[pa0 + i] * bra.coefs[pb0 + j] *
           bra_element_lds<LA, LB>(sa, ta, sb, tb, n / NB, n % NB, tv & 255, (tv >> 8) & 255, tv >> 16, E, E + ESZ,
                                   E + 2 * ESZ);
  }
}

template <int LA, int LB>
__global__ __launch_bounds__(256, 2) void schwarz4s_kernel(DevShells bra, const int* __restrict__ s1v,
                                                        const int* __restrict__ s2v,
                                                        const double* __restrict__ boys,
                                                        const double* __restrict__ scr, double* __restrict__ part,
                                                        const PairSkip sk) {
  using K = Cls4s<LA, LB>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NAB = K::NAB, NT = K::NT, NW = K::NW, GW = K::GW;
  constexpr int NCT = K::NCT, NS = K::NS, TPW = K::TPW, NRT1 = K::NRT1, ZC = K::ZC;
  __shared__ double Rq[NR + 1];  // even levels of the current quartet, finally R; Rq[NR] = 0 pads h >= H
  __shared__ double Ro[K::NRO > 0 ? K::NRO : 1];  // odd levels
  __shared__ double Z[NAB * NAB];
  __shared__ double G[K::NQS * GW];
  __shared__ int tuv[H];
  __shared__ int rec[NR];  // recursion of index h: parents h1, h2, direction, coefficient
  __shared__ double red[NW];
  const int tid = threadIdx.x, pair = blockIdx.x;
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int s1 = s1v[pair], s2 = s2v[pair];
  // a pair whose bound on Q_ab keeps no triplet: zero sum, no quartets
  if (pair_negligible_wave(bra, s1, s2, sk, tid & 63)) {
    if (tid == 0) part[pair] = 0.0;
    return;
  }
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb, nq = npp * (npp + 1) / 2;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  const double* F = scr + (size_t)pair * K::PERPAIR;

  for (int h = tid; h < H; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
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
    rec[h] = h1 | (h2 << 11) | (dir << 22) | (cf << 24);
  }
  for (int e = tid; e < NAB * NAB; e += NT) Z[e] = 0.0;
  if (tid == 0) Rq[NR] = 0.0;
  // scaled Boys values and 2 PQ of every quartet q = rs (rs + 1) / 2 + pq, pq <= rs
  for (int q = tid; q < nq; q += NT) {
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
    for (int z = 0; z < 3; z++) g[L + 1 + z] = 2.0 * PQ[z];
  }
  __syncthreads();
  // this lane's rows h1 = 16 t + l%16 of its row tiles t = wave + NW tt
  int tr[TPW], ur[TPW], vr[TPW];
  bool rok[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; tt++) {
    const int h1 = 16 * (wave + NW * tt) + lr;
    rok[tt] = h1 < H;
    const int tv = tuv[min(h1, H - 1)];
    tr[tt] = tv & 255;
    ur[tt] = (tv >> 8) & 255;
    vr[tt] = tv >> 16;
  }

  for (int pq = 0; pq < npp; pq++) {
    dbl4 acc[TPW][NCT];
#pragma unroll
    for (int tt = 0; tt < TPW; tt++)
#pragma unroll
      for (int ch = 0; ch < NCT; ch++) acc[tt][ch] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int rs = pq; rs < npp; rs++) {
      const double* g = G + (rs * (rs + 1) / 2 + pq) * GW;
      // R of the quartet by levels n = L .. 0 (L is even: level 0 lands in Rq)
      for (int n = L; n >= 0; n--) {
        const double* So = (n % 2 == 0) ? Ro : Rq;  // level n + 1
        double* Sn = (n % 2 == 0) ? Rq : Ro;
        const int nent = nherm(L - n);
        for (int h = tid; h < nent; h += NT) {
          double val;
          if (h == 0) {
            val = g[n];
          } else {
            const int rc = rec[h];
            const int h1 = rc & 2047, h2 = (rc >> 11) & 2047, dir = (rc >> 22) & 3, cf = rc >> 24;
            val = fma((double)(2 * cf), So[h2], g[L + 1 + dir] * So[h1]);
          }
          Sn[h] = val;
        }
        __syncthreads();
      }
      const double w = rs == pq ? 0.5 : 1.0;
      const double* Frs = F + (size_t)rs * H * NAB;
      // B of K step s: F rows h2 = 4 s + l/16, columns 16 ch + l%16
      double bcur[NCT], bnext[NCT];
      auto loadb = [&](int s, double* b) {
        const int h2 = 4 * s + lk;
#pragma unroll
        for (int ch = 0; ch < NCT; ch++) {
          const int n = 16 * ch + lr;
          b[ch] = (h2 < H && n < NAB) ? Frs[(size_t)h2 * NAB + n] : 0.0;
        }
      };
      loadb(0, bcur);
      for (int s = 0; s < NS; s++) {
        if (s + 1 < NS) loadb(s + 1, bnext);
        const int h2 = 4 * s + lk;
        int t2 = 0, u2 = 0, v2 = 0;
        double sg = 0.0;
        if (h2 < H) {
          const int tv2 = tuv[h2];
          t2 = tv2 & 255;
          u2 = (tv2 >> 8) & 255;
          v2 = tv2 >> 16;
          sg = ((t2 + u2 + v2) & 1) ? -w : w;
        }
#pragma unroll
        for (int tt = 0; tt < TPW; tt++) {
          if (wave + NW * tt < NRT1) {  // wave-uniform
            const int ro = (rok[tt] && h2 < H) ? herm_index(tr[tt] + t2, ur[tt] + u2, vr[tt] + v2) : NR;
            const double av = sg * Rq[ro];
#pragma unroll
            for (int ch = 0; ch < NCT; ch++)
              acc[tt][ch] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bcur[ch], acc[tt][ch], 0, 0, 0);
          }
        }
#pragma unroll
        for (int ch = 0; ch < NCT; ch++) bcur[ch] = bnext[ch];
      }
      __syncthreads();  // R of this quartet read by all waves before the next quartet's levels
    }
    // Z[n1][n2] += sum_h Et_pq[n1][h] V_pq[h][n2], one n1 tile and ZC n2
    // tiles at a time (few accumulator registers besides V's)
    const double* Fpq = F + (size_t)pq * H * NAB;
    for (int a0 = 0; a0 < NCT; a0++) {
#pragma unroll
      for (int cz = 0; cz < NCT; cz += ZC) {
        dbl4 zacc[ZC];
#pragma unroll
        for (int z = 0; z < ZC; z++) zacc[z] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int tt = 0; tt < TPW; tt++) {
          const int t = wave + NW * tt;
          if (t < NRT1) {  // wave-uniform
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int row = 16 * t + 4 * r + lk, n1 = 16 * a0 + lr;
              const double ea = (row < H && n1 < NAB) ? Fpq[(size_t)row * NAB + n1] : 0.0;
#pragma unroll
              for (int z = 0; z < ZC; z++)
                if (cz + z < NCT)
                  zacc[z] = __builtin_amdgcn_mfma_f64_16x16x4f64(ea, acc[tt][cz + z][r], zacc[z], 0, 0, 0);
            }
          }
        }
        // the waves add their partial tiles one after another (a fixed order)
        for (int wv = 0; wv < NW; wv++) {
          if (wave == wv) {
#pragma unroll
            for (int z = 0; z < ZC; z++)
#pragma unroll
              for (int r = 0; r < 4; r++) {
                const int n1 = 16 * a0 + lk + 4 * r, n2 = 16 * (cz + z) + lr;
                if (cz + z < NCT && n1 < NAB && n2 < NAB) Z[n1 * NAB + n2] += zacc[z][r];
              }
          }
          __syncthreads();
        }
      }
    }
  }
  double ss = 0.0;
  for (int e = tid; e < NAB * NAB; e += NT) {
    const int n1 = e / NAB, n2 = e - n1 * NAB;
    const double x = Z[n1 * NAB + n2] + Z[n2 * NAB + n1];
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

// Runs the class over npairs pairs in batches that fit scratch (nscr doubles).
template <int LA, int LB>
void launch_schwarz4s(DevShells bra, const int* s1, const int* s2, int npairs, const SphTable* sph, const double* boys,
                      double* part, double* scr, size_t nscr, const PairSkip& sk, hipStream_t st) {
  using K = Cls4s<LA, LB>;
  if (npairs <= 0) return;
  const int batch = (int)std::min<int64_t>(npairs, std::max<int64_t>(1, (int64_t)(nscr / K::PERPAIR)));
  for (int p0 = 0; p0 < npairs; p0 += batch) {
    const int nb = std::min(batch, npairs - p0);
    hipLaunchKernelGGL((schwarz4s_prep_kernel<LA, LB>), dim3(nb), dim3(256), 0, st, bra, s1 + p0, s2 + p0, sph, scr,
                       sk);
    hipLaunchKernelGGL((schwarz4s_kernel<LA, LB>), dim3(nb), dim3(K::NT), 0, st, bra, s1 + p0, s2 + p0, boys,
                       (const double*)scr, part + p0, sk);
  }
}

struct Schwarz4SLaunch {
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, double*, size_t,
             const PairSkip&, hipStream_t);
  int npps;          // largest primitive-pair count handled; 0 = class not instantiated
  int64_t perpair;   // scratch doubles per pair
};

template <int LA, int LB>
Schwarz4SLaunch make_schwarz4s() {
  if constexpr (Cls4s<LA, LB>::OK)
    return Schwarz4SLaunch{&launch_schwarz4s<LA, LB>, Cls4s<LA, LB>::NPPS, Cls4s<LA, LB>::PERPAIR};
  else
    return Schwarz4SLaunch{nullptr, 0, 0};
}

using Schwarz4STable = Schwarz4SLaunch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip
