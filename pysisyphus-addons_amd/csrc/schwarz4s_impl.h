// Schwarz norms on the FP64 matrix cores for the classes whose tables do not
// fit in LDS at once (common/integrals/intor.f90:131-160).  Same algebra as
// schwarz4m_impl.h (X = Z + Z^T, Z = sum_pq Et_pq V_pq, V_pq = sum_{rs >= pq}
// w M_{pq,rs} Et_rs^T), streamed:
//   * schwarz4s_prep_kernel writes, per pair, the spherical Hermite table
//     F[(pq, h)][n] and the R tables of the quartets pq <= rs to global
//     scratch (pairs processed in batches that fit the scratch);
//   * schwarz4s_kernel (one workgroup of 8 waves per pair) loops over
//     16-column chunks of V and, inside, over the ket primitive pair rs,
//     staging F_rs (that chunk) and R_{., rs} in LDS; the V tiles of all
//     rows stay in registers across rs; the Z product reads its A operand
//     (Et) from the global table and accumulates the full Z in LDS.
#pragma once

#include "schwarz4m_impl.h"

namespace dfhip {

template <int LA, int LB>
struct Cls4s {
  static constexpr int LAB = LA + LB, L = 2 * LAB;
  static constexpr int NAB = nsph(LA) * nsph(LB), NB = nsph(LB);
  static constexpr int H = nherm(LAB), NR = nherm(L), NR1 = NR + 1;
  static constexpr int NRL = nherm(L - 1) | 1;  // odd levels of the R recursion, per quartet
  static constexpr int HK = (H + 3) / 4 * 4, NS = HK / 4;
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NCT = (NAB + 15) / 16, ZL = 16 * NCT;  // n1 tiles of Z
  static constexpr int NCH = NCT;                               // 16-column chunks of V
  static constexpr int NPPS = 9, NQS = NPPS * (NPPS + 1) / 2;
  static constexpr int NRT = (NPPS * H + 15) / 16;
  static constexpr int NW = NRT > 48 ? 16 : 8, NT = 64 * NW;  // at most 6 row tiles per wave
  static constexpr int MAXT = (NRT + NW - 1) / NW;
  static constexpr int GW = (L + 4) | 1;  // per quartet: g_0..g_L, then 2 PQ_x, 2 PQ_y, 2 PQ_z (odd stride)
  // LDS of the main kernel: B chunk [HK][16], R of one ket pair [NPPS][NR1], Z [NAB][NAB]
  static constexpr int LDS = HK * 16 + NPPS * NR1 + NAB * NAB;
  static constexpr bool OK = LDS <= 19500;
  // global scratch per pair (doubles): F rows padded by HK - H + 16 for the
  // last ket pair's padded K steps and the Z operand's last row tile
  static constexpr int64_t FPAIR = (int64_t)(NPPS * H + HK + 16) * NAB + 16;
  // the F region holds the odd levels of the R recursion before F is built
  static constexpr int64_t FREG = FPAIR > (int64_t)NQS * NRL ? FPAIR : (int64_t)NQS * NRL;
  static constexpr int64_t RPAIR = (int64_t)NQS * NR1;
  static constexpr int64_t PERPAIR = FREG + RPAIR;
  // LDS of the prep kernel
  static constexpr int PREP_LDS = NPPS * 3 * ESZ + NQS * GW + NR;
};

template <int LA, int LB>
__global__ __launch_bounds__(256) void schwarz4s_prep_kernel(DevShells bra, const int* __restrict__ s1v,
                                                             const int* __restrict__ s2v,
                                                             const SphTable* __restrict__ sph,
                                                             const double* __restrict__ boys, double* __restrict__ scr) {
  using K = Cls4s<LA, LB>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NR1 = K::NR1, NAB = K::NAB, NB = K::NB, ESZ = K::ESZ, GW = K::GW;
  constexpr int NT = 256;
  __shared__ double Esc[K::NPPS * 3 * ESZ];
  __shared__ double G[K::NQS * GW];
  __shared__ int tuv[NR];
  __shared__ int rec[NR];  // recursion of index h: parents h1, h2, direction, coefficient
  __shared__ int sa[2 * LA + 2], sb[2 * LB + 2];
  __shared__ SphTerm ta[nsph(LA) * ncart(LA)], tb[nsph(LB) * ncart(LB)];
  const int tid = threadIdx.x, pair = blockIdx.x;
  const int s1 = s1v[pair], s2 = s2v[pair];
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb, nq = npp * (npp + 1) / 2;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  double* F = scr + (size_t)pair * K::PERPAIR;
  double* R = F + K::FREG;

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
    rec[h] = h1 | (h2 << 11) | (dir << 22) | (cf << 24);
  }
  for (int e = tid; e < npp * 3 + nq; e += NT) {
    if (e < npp * 3) {
      const int pq = e / 3, d = e - 3 * pq, i = pq / npb, j = pq - i * npb;
      herm_E_dir<LA, LB>(bra.exps[pa0 + i], bra.exps[pb0 + j], bra.xyz[3 * s1 + d], bra.xyz[3 * s2 + d],
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
      for (int z = 0; z < 3; z++) g[L + 1 + z] = 2.0 * PQ[z];
    }
  }
  __syncthreads();
  // R^(n) by levels n = L .. 0 in the pair's scratch (L is even: level n in R
  // if n is even, else in the F region, which is built afterwards); every
  // entry of a level is independent.  Levels n >= 1 are stored entry-major
  // ([h][quartet], the quartets of one entry on neighbouring lanes, so that
  // the stores of a wave are contiguous); level 0 goes quartet-major
  // ([quartet][h], the layout the main kernel stages), with the entries of
  // one quartet on neighbouring lanes.  Global memory written before
  // __syncthreads() is visible to the whole workgroup after it.
  {
    constexpr int NQS = K::NQS;
    double* Rt = F;
    for (int n = L; n >= 1; n--) {
      const double* So = (n % 2 == 0) ? Rt : R;  // level n+1, [h][q]
      double* Sn = (n % 2 == 0) ? R : Rt;
      const int nent = nherm(L - n);
      for (int e = tid; e < nent * NQS; e += NT) {
        const int h = e / NQS, q = e - h * NQS;  // compile-time divisor
        if (q >= nq) continue;
        const double* g = G + q * GW;
        double val;
        if (h == 0) {
          val = g[n];
        } else {
          const int rc = rec[h];
          const int h1 = rc & 2047, h2 = (rc >> 11) & 2047, dir = (rc >> 22) & 3, cf = rc >> 24;
          val = fma((double)(2 * cf), So[h2 * NQS + q], g[L + 1 + dir] * So[h1 * NQS + q]);
        }
        Sn[h * NQS + q] = val;
      }
      __syncthreads();
    }
    // level 0 from level 1 (in Rt); the entries NR of every quartet are the h2 >= H pad
    for (int e = tid; e < nq * NR1; e += NT) {
      const int q = e / NR1, h = e - q * NR1;  // compile-time divisor
      const double* g = G + q * GW;
      double val;
      if (h == 0) {
        val = g[0];
      } else if (h == NR) {
        val = 0.0;
      } else {
        const int rc = rec[h];
        const int h1 = rc & 2047, h2 = (rc >> 11) & 2047, dir = (rc >> 22) & 3, cf = rc >> 24;
        val = fma((double)(2 * cf), Rt[h2 * NQS + q], g[L + 1 + dir] * Rt[h1 * NQS + q]);
      }
      R[(size_t)q * NR1 + h] = val;
    }
    __syncthreads();
  }
  // F rows beyond npp H (padding read by the main kernel) are zero
  for (int64_t e = tid; e < K::FPAIR; e += NT) {
    double val = 0.0;
    if (e < (int64_t)npp * H * NAB) {
      const int pq = (int)(e / (H * NAB)), r = (int)(e - (int64_t)pq * (H * NAB)), h = r / NAB, n = r - h * NAB;
      const int i = pq / npb, j = pq - i * npb, tv = tuv[h];
      const double* E = Esc + pq * 3 * ESZ;
      val = bra.coefs[pa0 + i] * bra.coefs[pb0 + j] *
            bra_element_lds<LA, LB>(sa, ta, sb, tb, n / NB, n % NB, tv & 255, (tv >> 8) & 255, tv >> 16, E, E + ESZ,
                                    E + 2 * ESZ);
    }
    F[e] = val;
  }
}

template <int LA, int LB>
__global__ __launch_bounds__((Cls4s<LA, LB>::NT)) void schwarz4s_kernel(DevShells bra, const int* __restrict__ s1v,
                                                        const int* __restrict__ s2v,
                                                        const double* __restrict__ scr, double* __restrict__ part) {
  using K = Cls4s<LA, LB>;
  constexpr int H = K::H, HK = K::HK, NR = K::NR, NR1 = K::NR1, NAB = K::NAB, NT = K::NT, NW = K::NW;
  constexpr int NCT = K::NCT, NS = K::NS, MAXT = K::MAXT;
  __shared__ double Bs[HK * 16];
  __shared__ double Rs[K::NPPS * NR1];
  __shared__ double Z[NAB * NAB];
  __shared__ int tuv[H];
  __shared__ double red[NW];
  const int tid = threadIdx.x, pair = blockIdx.x;
  const int s1 = s1v[pair], s2 = s2v[pair];
  const int npp = bra.nprim[s1] * bra.nprim[s2];
  const double* F = scr + (size_t)pair * K::PERPAIR;
  const double* R = F + K::FREG;
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int nrt = (npp * H + 15) / 16;

  for (int h = tid; h < H; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
  }
  for (int e = tid; e < NAB * NAB; e += NT) Z[e] = 0.0;

  for (int ch = 0; ch < K::NCH; ch++) {
    dbl4 vacc[MAXT];
#pragma unroll
    for (int t = 0; t < MAXT; t++) vacc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    for (int rs = 0; rs < npp; rs++) {
      __syncthreads();  // previous stage consumed (and tuv, Z cleared on the first pass)
      // B chunk: Bs[h2][j] = F[(rs H + h2)][16 ch + j], zero past H and NAB
      for (int e = tid; e < HK * 16; e += NT) {
        const int h2 = e >> 4, j = e & 15, n = 16 * ch + j;
        Bs[e] = (h2 < H && n < NAB) ? F[(size_t)(rs * H + h2) * NAB + n] : 0.0;
      }
      // R of the quartets (pq, rs), pq <= rs; entry NR is the zero pad
      for (int e = tid; e < (rs + 1) * NR1; e += NT) {
        const int pq = e / NR1, h = e - pq * NR1;
        Rs[e] = R[(size_t)(rs * (rs + 1) / 2 + pq) * NR1 + h];
      }
      __syncthreads();
      // per row tile (snake order): weight, R base and Hermite indices of the lane's row
      int rb[MAXT], tr[MAXT], ur[MAXT], vr[MAXT];
      double wt[MAXT];
      bool on[MAXT];
#pragma unroll
      for (int t = 0; t < MAXT; t++) {
        const int i = (t & 1) ? (t + 1) * NW - 1 - wave : t * NW + wave;
        on[t] = i < nrt && (16 * i) / H <= rs;  // wave-uniform
        const int row = 16 * i + lr;
        const int pqr = row / H, hr = row - pqr * H;
        wt[t] = (pqr < npp && rs >= pqr) ? (rs == pqr ? 0.5 : 1.0) : 0.0;
        rb[t] = min(pqr, rs) * NR1;
        const int tv = tuv[min(hr, H - 1)];
        tr[t] = tv & 255;
        ur[t] = (tv >> 8) & 255;
        vr[t] = tv >> 16;
      }
      for (int s = 0; s < NS; s++) {
        const int h2 = 4 * s + lk;
        int t2 = 0, u2 = 0, v2 = 0;
        double sg = 0.0;
        if (h2 < H) {
          const int tv2 = tuv[h2];
          t2 = tv2 & 255;
          u2 = (tv2 >> 8) & 255;
          v2 = tv2 >> 16;
          sg = ((t2 + u2 + v2) & 1) ? -1.0 : 1.0;
        }
        const double b = Bs[h2 * 16 + lr];
#pragma unroll
        for (int t = 0; t < MAXT; t++) {
          if (!on[t]) continue;
          const int ro = h2 < H ? herm_index(tr[t] + t2, ur[t] + u2, vr[t] + v2) : NR;
          vacc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(sg * wt[t] * Rs[rb[t] + ro], b, vacc[t], 0, 0, 0);
        }
      }
    }
    // Z[n1][16 ch + n2] += sum_rows Et[n1][row] V[row][n2]; register r of lane l holds V row l/16 + 4 r.
    // Two n1 tiles at a time keep the accumulators small.
    for (int a0 = 0; a0 < NCT; a0 += 2) {
      dbl4 zacc[2] = {dbl4{0.0, 0.0, 0.0, 0.0}, dbl4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
      for (int t = 0; t < MAXT; t++) {
        const int m = t;
        const int i = (m & 1) ? (m + 1) * NW - 1 - wave : m * NW + wave;
        if (i >= nrt) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const double* Fz = F + (size_t)(16 * i + lk + 4 * r) * NAB;
#pragma unroll
          for (int z = 0; z < 2; z++) {
            const int n1 = 16 * (a0 + z) + lr;
            const double ea = n1 < NAB ? Fz[n1] : 0.0;
            zacc[z] = __builtin_amdgcn_mfma_f64_16x16x4f64(ea, vacc[t][r], zacc[z], 0, 0, 0);
          }
        }
        __builtin_amdgcn_sched_barrier(0);  // keeps the loads of later tiles from piling up in registers
      }
      // the waves add their partial tiles one after another (a fixed order,
      // so Q_ab is reproducible bit for bit)
      for (int w = 0; w < NW; w++) {
        if (wave == w) {
#pragma unroll
          for (int z = 0; z < 2; z++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int n1 = 16 * (a0 + z) + lk + 4 * r, n2 = 16 * ch + lr;
              if (n1 < NAB && n2 < NAB) Z[n1 * NAB + n2] += zacc[z][r];
            }
        }
        __syncthreads();
      }
    }
  }
  __syncthreads();
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
                      double* part, double* scr, size_t nscr, hipStream_t st) {
  using K = Cls4s<LA, LB>;
  if (npairs <= 0) return;
  const int batch = (int)std::min<int64_t>(npairs, std::max<int64_t>(1, (int64_t)(nscr / K::PERPAIR)));
  for (int p0 = 0; p0 < npairs; p0 += batch) {
    const int nb = std::min(batch, npairs - p0);
    hipLaunchKernelGGL((schwarz4s_prep_kernel<LA, LB>), dim3(nb), dim3(256), 0, st, bra, s1 + p0, s2 + p0, sph, boys,
                       scr);
    hipLaunchKernelGGL((schwarz4s_kernel<LA, LB>), dim3(nb), dim3(K::NT), 0, st, bra, s1 + p0, s2 + p0,
                       (const double*)scr, part + p0);
  }
}

struct Schwarz4SLaunch {
  void (*fn)(DevShells, const int*, const int*, int, const SphTable*, const double*, double*, double*, size_t,
             hipStream_t);
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
