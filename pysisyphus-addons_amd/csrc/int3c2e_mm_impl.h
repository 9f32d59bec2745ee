// 3c2e classes on the FP64 matrix cores (the classes without a generated
// register kernel: an f or g orbital shell, or a g or h aux shell).
//
// The bra contraction dominates these classes: per primitive pair ij and aux
// shell C, out[mu nu][m] += sum_h Et_ij[mu nu][h] W_ijC[h][m] with
// Et (nsph(a) nsph(b) x nherm(la+lb)) up to 81 x 165.  Rebuilding Et per
// chunk of aux shells (what the LDS kernel of int3c2e_impl.h does, with only
// a few aux shells per chunk at these sizes) costs as much as the product, so
//   * int3c2e_et_kernel writes, once per call, Et of every primitive pair of
//     the pairs that have screened-in triplets to a global table, h-major and
//     padded to MFMA fragments: EtG[ij][h < HK][mu nu < NABP];
//   * int3c2e_mm_kernel takes one shell pair x NC aux shells of one L per
//     workgroup: per primitive pair and aux primitive each wave builds R_tuv
//     of its own aux shells by levels in LDS (synchronized within the wave
//     only, so one wave's levels run beside the other waves' products), the
//     workgroup builds the Hobson-transformed W[h][(C, m)] of all NC shells,
//     then the waves run the (NAB x H) . (H x NC nsph(lc)) product on the
//     matrix cores with A fragments read straight from the global Et table
//     and B fragments from LDS; the accumulators stay in registers over all
//     primitive pairs.  Two workgroup barriers per primitive pair.
// Workgroups of 4 waves, two per CU for T = double (the FP64 path,
// v_mfma_f64_16x16x4) and three for T = float (the mixed path's
// low-precision window, v_mfma_f32_16x16x4f32, half the LDS): with T = float
// the Et table, R_tuv, the Hobson step, the product and its accumulation over
// the primitive pairs run in FP32, while the Boys values and prefactors (the
// G table) stay FP64; the block is stored as FP64.
#pragma once

#include "gen3c_common.h"
#include "int3c2e_impl.h"

namespace dfhip {

typedef double dbl4m __attribute__((ext_vector_type(4)));
typedef float flt4m __attribute__((ext_vector_type(4)));

// The 16x16x4 matrix-core product per arithmetic type.  A fragment: lane l
// holds A[l % 16][l / 16]; B fragment: B[l / 16][l % 16]; accumulator
// register r of lane l: column l % 16, row l / 16 + 4 r for FP64
// (v_mfma_f64_16x16x4f64) and 4 (l / 16) + r for FP32
// (v_mfma_f32_16x16x4f32).
template <typename T>
struct MMOp;
template <>
struct MMOp<double> {
  typedef dbl4m v4;
  static __device__ __forceinline__ v4 mfma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __host__ __device__ constexpr int row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct MMOp<float> {
  typedef flt4m v4;
  static __device__ __forceinline__ v4 mfma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __host__ __device__ constexpr int row(int lk, int r) { return 4 * lk + r; }
};

// LDS writes of a wave visible to all its lanes (LDS operations of one wave
// complete in order; the fences keep the compiler from moving accesses across)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LA, int LB, int LC, typename T = double>
struct Cls3m {
  static constexpr int LAB = LA + LB, L = LAB + LC;
  static constexpr int NA = nsph(LA), NB = nsph(LB), NAB = NA * NB, NCC = nsph(LC);
  static constexpr int H = nherm(LAB), NR = nherm(L), NRLO = (LC == 0) ? 0 : nherm(LC - 1);
  static constexpr int HK = (H + 3) / 4 * 4, NABP = (NAB + 15) / 16 * 16, MT = NABP / 16;
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int GW = L + 4;  // per aux shell: scaled Boys values g_0..g_L, then 2 PC_x, 2 PC_y, 2 PC_z
  // LDS stride per aux shell: odd, so lanes on consecutive aux shells read distinct banks
  static constexpr int GWP = GW | 1;
  // R^(n)_tuv by levels n = L-1 .. 0 in two buffers: even levels in R (all NR
  // indices), odd levels in Rt (orders <= L-1); odd strides per aux shell
  static constexpr int NRP = NR | 1, NRT = nherm(L - 1) | 1;
  // Workgroups of 4 waves; OCC of them share a CU: two for FP64, three for
  // FP32, whose tables take half the LDS (measured on the 200-atom g/h
  // workload: one 8-wave FP64 workgroup per CU took 3.70 s of 3c2e time per
  // step, two 4-wave ones 2.99 s; three for FP32 saved another 5 %, three for
  // FP64 nothing).  One workgroup's barriers overlap another's products.
  static constexpr int OCC = sizeof(T) == sizeof(double) ? 2 : 3;
  static constexpr int NW = 4, NT = 64 * NW;
  static constexpr int BUD = 160 * 1024 / OCC - 6 * 1024;  // LDS bytes for R, Rt, G and W
  // G of up to NPG primitive pairs is built once per workgroup (uncontracted aux)
  static constexpr int NPG = 9;
  static constexpr int nc_fit(int nc) {
    return (int)((nc * (NRP + NRT) + HK * ((nc * NCC + 15) / 16 * 16)) * sizeof(T) + nc * NPG * GWP * 8) <= BUD;
  }
  // accumulator tiles per wave at nc aux shells: the MT x NTL tiles in
  // row-major order, an equal run of at most 8 per wave
  static constexpr int npw_of(int nc) { return (MT * ((nc * NCC + 15) / 16) + NW - 1) / NW; }
  static constexpr int nc_max() {
    int nc = 32;
    while (nc > 1 && (!nc_fit(nc) || npw_of(nc) > 8)) nc--;
    // a multiple of the waves, so that every wave owns as many shells in the
    // R levels (200-atom workload: 3c2e phase 2.99 -> 2.90 s)
    if (nc >= NW) nc -= nc % NW;
    return nc;
  }
  static constexpr int NC = nc_max();
  static constexpr int NCOL = NC * NCC, NTL = (NCOL + 15) / 16, NCOLP = NTL * 16;
  // tiles per wave and the most row tiles one wave's run touches
  static constexpr int TPW = (MT * NTL + NW - 1) / NW;
  static constexpr int RPM = (TPW + NTL - 2) / NTL + 1;
  static constexpr int NCW = (NC + NW - 1) / NW;  // aux shells per wave in the R levels
  static constexpr int KS = HK / 4;
  // Hobson step: one thread per (h, m) and group of aux shells
  static constexpr int MAXT = sph_max_terms(LC);
  static constexpr int CG = NT / (H * NCC) >= 1 ? NT / (H * NCC) : 1;
  static constexpr int LH = L / 2 + 1;
  static constexpr int64_t ETPP = (int64_t)HK * NABP;  // Et table doubles per primitive pair
  // rounds of independent entries per lane in flight in the R levels, and
  // the unrolling of the Hobson step's loop over aux shells (measured on the
  // 200-atom g/h workload; compile-time trip counts over all NC shells were
  // slower there, the last chunks of a pair hold fewer)
  static constexpr int UR = 8;
  static constexpr bool OK = nc_fit(NC);
};

// R levels of one wave's aux shells (c = wave, wave + NW, ...; at most NCW
// of them), level N down to 0.  Level N covers the orders t+u+v <= L-N (level
// 0 only those >= LC, which the Hobson step reads), from level N+1 in the
// other buffer:
//   R^(N)_tuv = 2 (t-1) R^(N+1)_{t-2,u,v} + 2 PC_x R^(N+1)_{t-1,u,v}
// along the first non-zero index (rec[h]: parents, direction, coefficient;
// rec[0] is a harmless dummy), R^(N)_000 = g_N, in the scale of the G table.
// Every level has a compile-time size, so its rounds of 64 entries are
// straight-line code: UR rounds at a time issue their reads together, entries
// past the end or of shells >= nc read clamped indices and skip the store.
template <int L, int NW, int NCW, int NRP, int NRT, int GWP, int NRLO, int UR, typename T>
struct MMLevels {
  template <int N>
  static __device__ __forceinline__ void run(T* __restrict__ R, T* __restrict__ Rt, const int* __restrict__ rec,
                                             const double* __restrict__ gb, int lane, int wave, int nc) {
    static_assert(N >= 0 && N < L, "R levels run from L-1 down to 0");
    constexpr bool EVEN = N % 2 == 0;
    const T* So = EVEN ? Rt : R;  // level N+1
    T* Sn = EVEN ? R : Rt;
    constexpr int SO = EVEN ? NRT : NRP, SN = EVEN ? NRP : NRT;
    constexpr int HB = N == 0 ? NRLO : 0, NE = nherm(L - N) - HB, TOT = NE * NCW, NRND = (TOT + 63) / 64;
#pragma unroll
    for (int r0 = 0; r0 < NRND; r0 += UR) {
      T val[UR];
      int dst[UR];
      bool live[UR];
#pragma unroll
      for (int u = 0; u < UR; u++) {
        const int r = r0 + u;
        if (r >= NRND) break;
        const int k0 = lane + 64 * r;
        const bool ok = r < NRND - 1 || k0 < TOT;
        const int k = ok ? k0 : TOT - 1;
        const int e = k / NCW, cl = k - e * NCW, h = HB + e;  // compile-time divisor
        const int c = wave + NW * cl;
        live[u] = ok && c < nc;
        const int cc = c < nc ? c : 0;
        const double* gc = gb + cc * GWP;
        const int rc = rec[h];
        const int h1 = rc & 1023, h2 = (rc >> 10) & 1023, dir = (rc >> 20) & 3, cf = rc >> 22;
        const T* sc0 = So + cc * SO;
        const T v = fma((T)(2 * cf), sc0[h2], (T)gc[L + 1 + dir] * sc0[h1]);
        val[u] = h == 0 ? (T)gc[N] : v;
        dst[u] = cc * SN + h;
      }
#pragma unroll
      for (int u = 0; u < UR; u++) {
        if (r0 + u >= NRND) break;
        if (live[u]) Sn[dst[u]] = val[u];
      }
    }
    wave_lds_sync();
    if constexpr (N > 0) run<N - 1>(R, Rt, rec, gb, lane, wave, nc);
  }
};

// Et tables of the listed pairs: EtG[off[p] + ij*ETPP + h*NABP + mu*NB + nu],
// contraction coefficients folded in, zero padding; evaluated in FP64 and
// stored as TO (double, or float for the FP32 window of the mixed path).
template <int LA, int LB, typename TO>
__global__ __launch_bounds__(256) void int3c2e_et_kernel(DevShells bra, const int* __restrict__ s1v,
                                                         const int* __restrict__ s2v,
                                                         const int* __restrict__ list,
                                                         const int64_t* __restrict__ off,
                                                         const SphTable* __restrict__ sph, TO* __restrict__ EtG) {
  using K = Cls3m<LA, LB, 0>;
  constexpr int H = K::H, NAB = K::NAB, NABP = K::NABP, NB = K::NB, ESZ = K::ESZ, NT = 256;
  constexpr int NPPM = 16;  // primitive pairs per pass
  __shared__ double E[NPPM * 3 * ESZ];
  __shared__ int tuv[H];
  __shared__ int sa[2 * LA + 2], sb[2 * LB + 2];
  __shared__ SphTerm ta[nsph_terms(LA)], tb[nsph_terms(LB)];
  const int tid = threadIdx.x;
  const int p = list[blockIdx.x];
  const int s1 = s1v[p], s2 = s2v[p];
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  TO* out = EtG + off[p];
  for (int h = tid; h < H; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
  }
  {
    const int a0 = sph->start[LA][0], b0 = sph->start[LB][0];
    if (tid < 2 * LA + 2) sa[tid] = sph->start[LA][tid] - a0;
    if (tid < 2 * LB + 2) sb[tid] = sph->start[LB][tid] - b0;
    for (int e = tid; e < nsph_terms(LA); e += NT) ta[e] = sph->terms[a0 + e];
    for (int e = tid; e < nsph_terms(LB); e += NT) tb[e] = sph->terms[b0 + e];
  }
  for (int q0 = 0; q0 < npp; q0 += NPPM) {
    const int nq = min(NPPM, npp - q0);
    __syncthreads();
    for (int e = tid; e < nq * 3; e += NT) {
      const int ij = q0 + e / 3, d = e % 3, i = ij / npb, j = ij - i * npb;
      herm_E_dir<LA, LB>(bra.exps[pa0 + i], bra.exps[pb0 + j], bra.xyz[3 * s1 + d], bra.xyz[3 * s2 + d],
                         E + e * ESZ);
    }
    __syncthreads();
    for (int64_t e = tid; e < (int64_t)nq * K::ETPP; e += NT) {
      const int q = (int)(e / K::ETPP), r = (int)(e - (int64_t)q * K::ETPP), h = r / NABP, n = r - h * NABP;
      const int ij = q0 + q, i = ij / npb, j = ij - i * npb;
      double val = 0.0;
      if (h < H && n < NAB) {
        const int tv = tuv[h];
        const double* Eq = E + q * 3 * ESZ;
        val = bra.coefs[pa0 + i] * bra.coefs[pb0 + j] *
              bra_element_lds<LA, LB>(sa, ta, sb, tb, n / NB, n % NB, tv & 255, (tv >> 8) & 255, tv >> 16, Eq,
                                      Eq + ESZ, Eq + 2 * ESZ);
      }
      out[(int64_t)ij * K::ETPP + r] = (TO)val;
    }
  }
}

// One shell pair x NC aux shells of one L per workgroup, the list window
// [nbeg, nsurv) of the pair (nbeg null: from the start), T the arithmetic
// type of R_tuv, the Hobson step and the product (see the file head).
template <int LA, int LB, int LC, typename T>
__global__ __launch_bounds__((Cls3m<LA, LB, LC, T>::NT), (Cls3m<LA, LB, LC, T>::OCC)) void int3c2e_mm_kernel(DevShells bra, DevShells ket, PairArgs pa, AuxArgs xa,
                                                         const T* __restrict__ EtG,
                                                         const int64_t* __restrict__ etoff,
                                                         const SphTable* __restrict__ sph,
                                                         const double* __restrict__ boys, double* __restrict__ out,
                                                         int64_t ldo) {
  using K = Cls3m<LA, LB, LC, T>;
  using Op = MMOp<T>;
  constexpr int L = K::L, H = K::H, HK = K::HK, NR = K::NR, NCC = K::NCC, NAB = K::NAB, NABP = K::NABP;
  constexpr int NT = K::NT, NC = K::NC, NCOLP = K::NCOLP, GWP = K::GWP;
  constexpr int NRP = K::NRP, NRT = K::NRT;
  constexpr int MT = K::MT, KS = K::KS, NTL = K::NTL, NW = K::NW, NCW = K::NCW;
  __shared__ double G[K::NPG * NC * GWP];  // per primitive pair and aux shell: scaled Boys values, then 2 PC_d
  __shared__ T R[NC * NRP];       // per aux shell: R^(n)_tuv of the even levels, finally R_tuv
  __shared__ T Rt[NC * NRT];      // per aux shell: R^(n)_tuv of the odd levels
  __shared__ T W[HK * NCOLP];     // W[h][c NCC + m]
  __shared__ double cxyz[NC * 3];
  __shared__ int cshell[NC], ckeep[NC];
  __shared__ int tuv[NR];
  __shared__ int rec[NR];  // recursion of index h: parents h1, h2, direction, coefficient
  __shared__ int sc[2 * LC + 2];
  __shared__ SphTerm tc[nsph_terms(LC)];

  // XCD-aware order: workgroups go round-robin to the 8 XCDs by linear id, so
  // slot blockIdx.x % 8 of a row of the grid keeps one pair and walks its aux
  // chunks (blockIdx.x / 8): the chunks of a pair share its Et table in one L2
  const int pidx = (int)blockIdx.y * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int chunk = (int)(blockIdx.x >> 3);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  __shared__ int cidx[NC];  // list entry of each aux shell of the chunk
  int nc;
  if (pa.kidx) {
    // chunk = NC consecutive kept entries of the window [cbeg, cend); the
    // dropped entries from this chunk's first kept one (the window start for
    // chunk 0) to the next chunk's first kept one (the window end for the
    // last) get zero blocks from this workgroup
    const int cbeg = pa.nbeg ? max(xa.lo, pa.nbeg[pidx]) : xa.lo;
    if (cbeg >= cend) return;
    const int* kl = pa.kidx + pa.koff[pidx];
    const int nk = (int)(pa.koff[pidx + 1] - pa.koff[pidx]);
    auto lower = [&](int v) {
      int lo = 0, hi = nk;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (kl[mid] < v) lo = mid + 1;
        else hi = mid;
      }
      return lo;
    };
    const int ja = pa.kja ? pa.kja[pidx] : lower(cbeg), jb = pa.kja ? pa.kjb[pidx] : lower(cend);
    const int j0 = ja + chunk * NC;
    if (j0 >= jb && chunk > 0) return;
    const int j1 = min(jb, j0 + NC);
    nc = max(0, j1 - j0);
    // zero blocks of the dropped entries in [zb, ze): the range's columns are
    // contiguous (the list's columns are), so consecutive threads store
    // consecutive columns of each block row
    const int zb = chunk == 0 ? cbeg : kl[j0], ze = j1 < jb ? kl[j1] : cend;
    const int span = (ze - zb) * NCC;
    double* zout = out + pa.row[pidx] * ldo + xa.col[zb];
    for (int q = tid; q < span; q += NT) {
      const int i = zb + q / NCC;
      int lo = j0, hi = j1;  // is entry i one of the chunk's kept entries?
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (kl[mid] < i) lo = mid + 1;
        else hi = mid;
      }
      if (lo < j1 && kl[lo] == i) continue;
      for (int munu = 0; munu < NAB; munu++) zout[munu * ldo + q] = 0.0;
    }
    for (int c = tid; c < nc; c += NT) cidx[c] = kl[j0 + c];
    if (nc == 0) return;
  } else {
    const int cchunk = xa.lo + chunk * NC;
    const int c0 = pa.nbeg ? max(cchunk, pa.nbeg[pidx]) : cchunk;
    if (c0 >= min(cend, cchunk + NC)) return;
    nc = min(cchunk + NC, cend) - c0;
    for (int c = tid; c < nc; c += NT) cidx[c] = c0 + c;
  }
  __syncthreads();
  const int s1 = pa.s1[pidx], s2 = pa.s2[pidx];
  const T* Et0 = EtG + etoff[pidx];

  for (int h = tid; h < NR; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
  }
  // R^(n)_tuv = 2 (t-1) R^(n+1)_{t-2,u,v} + 2 PC_x R^(n+1)_{t-1,u,v} along the first
  // non-zero index (here in the scale of the G table: R^(n)_000 = g_n)
  if (tid == 0) rec[0] = 0;  // dummy parents of h = 0 (its value is g_n)
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
  {
    const int x0 = sph->start[LC][0];
    if (tid < 2 * LC + 2) sc[tid] = sph->start[LC][tid] - x0;
    for (int e = tid; e < nsph_terms(LC); e += NT) tc[e] = sph->terms[x0 + e];
  }
  for (int c = tid; c < nc; c += NT) {
    const int sh = xa.shell[cidx[c]];
    cshell[c] = sh;
    cxyz[3 * c + 0] = ket.xyz[3 * sh + 0];
    cxyz[3 * c + 1] = ket.xyz[3 * sh + 1];
    cxyz[3 * c + 2] = ket.xyz[3 * sh + 2];
    // the compact lists hold kept entries only (the test of qvl_compact)
    ckeep[c] = pa.kidx ? 1
                       : triplet_kept(pa, xa, pidx, sh, LC, ket.xyz[3 * sh + 0], ket.xyz[3 * sh + 1], ket.xyz[3 * sh + 2]);
  }
  for (int e = tid; e < HK * NCOLP; e += NT) W[e] = (T)0;  // padding rows and columns stay zero

  double A[3], B[3];
  for (int d = 0; d < 3; d++) {
    A[d] = bra.xyz[3 * s1 + d];
    B[d] = bra.xyz[3 * s2 + d];
  }
  const int npb = bra.nprim[s2], npp = bra.nprim[s1] * npb;
  const int pa0 = bra.pptr[s1], pb0 = bra.pptr[s2];
  const double lc_sign = (LC & 1) ? -1.0 : 1.0;

  // G of one (primitive pair, aux primitive, aux shell) in FP64, in two
  // parts so that every lane of a loop takes the same branch: gboys the
  // scaled Boys values g_n = s (-alpha)^n F_n(T) (F_L from the table, the
  // rest by downward recursion), gpc the 2 PC_d
  auto gboys = [&](int c, const double* P, double p, int kp, double* Gd) {
    const double PC0 = P[0] - cxyz[3 * c + 0], PC1 = P[1] - cxyz[3 * c + 1], PC2 = P[2] - cxyz[3 * c + 2];
    double* gc = Gd + c * GWP;
    const int sh = cshell[c];
    const bool live = kp < ket.nprim[sh] && ckeep[c];
    double F[L + 1];
    double scale = 0.0, alpha = 0.0;
    if (live) {
      const double g = ket.exps[ket.pptr[sh] + kp], dk = ket.coefs[ket.pptr[sh] + kp];
      alpha = p * g / (p + g);
      const double Tb = alpha * (PC0 * PC0 + PC1 * PC1 + PC2 * PC2);
      scale = 2.0 * PI * PI * sqrt(PI) / (p * g * sqrt(p + g)) * dk * lc_sign;
      const double inv2g = 0.5 / g;
#pragma unroll
      for (int e = 0; e < LC; e++) scale *= inv2g;
      boys_array<L>(Tb, boys, F);
    } else {
#pragma unroll
      for (int n = 0; n <= L; n++) F[n] = 0.0;
    }
#pragma unroll
    for (int n = 0; n <= L; n++) {
      gc[n] = scale * F[n];
      scale *= -alpha;
    }
  };
  auto gpc = [&](int c, int d, const double* P, double* Gd) {
    Gd[c * GWP + L + 1 + d] = 2.0 * (P[d] - cxyz[3 * c + d]);
  };
  auto seed = [&](int c, const double* gb) { (L % 2 == 0 ? R + c * NRP : Rt + c * NRT)[0] = (T)gb[c * GWP + L]; };
  // uncontracted aux shells and at most NPG primitive pairs: G of every
  // primitive pair at once (9x the parallelism, one wait on the Boys table)
  const bool pre = xa.kmax == 1 && npp <= K::NPG;
  if (pre) {
    __syncthreads();  // cshell, cxyz and ckeep of the prologue are read below
    for (int k = tid; k < npp * nc; k += NT) {
      const int ij = k / nc, c = k - ij * nc;
      const int i = ij / npb, j = ij - i * npb;
      const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], p = a + b;
      double P[3];
      for (int d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;
      gboys(c, P, p, 0, G + ij * NC * GWP);
    }
    for (int k = tid; k < npp * nc * 3; k += NT) {
      const int ij = k / (nc * 3), r = k - ij * (nc * 3), c = r / 3, d = r - c * 3;
      const int i = ij / npb, j = ij - i * npb;
      const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], p = a + b;
      double P[3];
      for (int e = 0; e < 3; e++) P[e] = (a * A[e] + b * B[e]) / p;
      gpc(c, d, P, G + ij * NC * GWP);
    }
  }

  typename Op::v4 acc[K::TPW];
#pragma unroll
  for (int t = 0; t < K::TPW; t++) acc[t] = typename Op::v4{(T)0, (T)0, (T)0, (T)0};

  // A fragments of one chunk of 8 K steps for the row tiles
  // mt0 .. mt0 + RPM - 1 (clamped), double-buffered across chunks.  The Et
  // table is read-only: without the opaque offset the compiler hoists every
  // chunk's loads above the barriers of the table phase and spills them
  T ak[2][K::RPM][8];
  auto load_a = [&](T (&dst)[K::RPM][8], const T* Etb, int ch, int mt0) {
    int zo = 0;
    asm volatile("" : "+s"(zo));
#pragma unroll
    for (int r = 0; r < K::RPM; r++) {
      const int mtr = min(mt0 + r, MT - 1);
      const T* ea = Etb + zo + lk * NABP + 16 * mtr + lr;
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (8 * ch + s < KS) dst[r][s] = ea[4 * (8 * ch + s) * NABP];
    }
  };
  __syncthreads();  // prologue tables, G, zeroed W
  for (int ij = 0; ij < npp; ij++) {
    const int i = ij / npb, j = ij - i * npb;
    const double a = bra.exps[pa0 + i], b = bra.exps[pb0 + j], p = a + b;
    double P[3];
    for (int d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;
    const T* Et = Et0 + (int64_t)ij * K::ETPP;
    for (int kp = 0; kp < xa.kmax; kp++) {
      const double* gb = pre ? G + ij * NC * GWP : G;
      if (!pre) {
        // G of this (primitive pair, aux primitive) for all shells: every
        // wave is past the previous Hobson step, so G is free
        for (int c = tid; c < nc; c += NT) gboys(c, P, p, kp, G);
        for (int k = tid; k < nc * 3; k += NT) gpc(k / 3, k % 3, P, G);
        __syncthreads();
      }
      // R levels of the wave's own aux shells, synchronized within the wave
      // only: a wave starts them as soon as its product of the previous
      // primitive pair is issued, beside the other waves' products
      if (lane < NCW) {
        const int c = wave + NW * lane;
        if (c < nc) seed(c, gb);
      }
      wave_lds_sync();
      // (L = 0: the seed g_0 is R itself)
      if constexpr (L > 0)
        MMLevels<L, NW, NCW, NRP, NRT, GWP, K::NRLO, K::UR, T>::template run<L - 1>(R, Rt, rec, gb, lane, wave, nc);
      // R of every shell is there, and every wave's previous product is done
      // with W
      __syncthreads();
      load_a(ak[0], Et, 0, wave * K::TPW / NTL);
      // W[h][c NCC + m] = sum_q c_q R_c[(t, u, v) + l_q] over the solid-harmonic terms
      // q of (LC, m): the term indices are resolved once per (h, m) and reused
      // for every aux shell of the thread's group
      for (int k = tid; k < H * NCC * K::CG; k += NT) {
        const int hm = k % (H * NCC), cg = k / (H * NCC), h = hm / NCC, m = hm - h * NCC;
        const int tv = tuv[h];
        const int t = tv & 255, u = (tv >> 8) & 255, v = tv >> 16;
        const int q0 = sc[m], nq = sc[m + 1] - q0;
        T cq[K::MAXT];
        int iq[K::MAXT];
#pragma unroll
        for (int z = 0; z < K::MAXT; z++) {
          const SphTerm tm = tc[z < nq ? q0 + z : q0];
          cq[z] = z < nq ? (T)tm.c : (T)0;
          iq[z] = herm_index(t + tm.lx, u + tm.ly, v + tm.lz);
        }
#pragma unroll(K::UR)
        for (int c = cg; c < nc; c += K::CG) {
          const T* Rc = R + c * NRP;
          T s = (T)0;
#pragma unroll
          for (int z = 0; z < K::MAXT; z++) s = fma(cq[z], Rc[iq[z]], s);
          W[h * NCOLP + c * NCC + m] = s;
        }
      }
      __syncthreads();
      // acc[tile] += Et[16 mt.., h] W[h, 16 nt..] over all primitive pairs:
      // the wave's run of tiles [tb, te), in chunks of 8 K steps: the A
      // fragments of the next chunk (every row tile the run touches) are
      // requested before this chunk's products, chunk 0 before the
      // Hobson step; one accumulation chain per tile (in the arithmetic type)
      constexpr int TPW = K::TPW, RPM = K::RPM, NCH = (KS + 7) / 8;
      const int tb = wave * TPW, te = min(tb + TPW, MT * NTL), mt0 = tb / NTL;
#pragma unroll
      for (int ch = 0; ch < NCH; ch++) {
        if (ch + 1 < NCH) load_a(ak[(ch + 1) & 1], Et, ch + 1, mt0);
        const int s0 = 8 * ch;
#pragma unroll
        for (int q = 0; q < TPW; q++) {
          const int tau = tb + q;
          if (tau >= te) break;
          const int mtq = tau / NTL, nt = tau - mtq * NTL;
          const T* wb = W + lk * NCOLP + 16 * nt + lr;
          typename Op::v4 c4 = acc[q];
#pragma unroll
          for (int s = 0; s < 8; s++) {
            if (s0 + s >= KS) break;
            T a = ak[ch & 1][0][s];
#pragma unroll
            for (int r = 1; r < RPM; r++) a = (mtq == mt0 + r) ? ak[ch & 1][r][s] : a;
            c4 = Op::mfma(a, wb[4 * (s0 + s) * NCOLP], c4);
          }
          acc[q] = c4;
        }
      }
    }
  }
  const int64_t row0 = pa.row[pidx];
  const int tb = wave * K::TPW, te = min(tb + K::TPW, MT * NTL);
#pragma unroll
  for (int q = 0; q < K::TPW; q++) {
    const int tau = tb + q;
    if (tau >= te) break;
    const int mtq = tau / NTL, nt = tau - mtq * NTL;
    const int col = 16 * nt + lr, c = col / NCC, m = col - c * NCC;
    if (c >= nc) continue;
    const int64_t ocol = xa.col[cidx[c]] + m;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int munu = 16 * mtq + Op::row(lk, r);
      if (munu < NAB) out[(row0 + munu) * ldo + ocol] = (double)acc[q][r];
    }
  }
}

template <typename T>
struct Int3MFn {
  typedef void (*type)(DevShells, DevShells, PairArgs, AuxArgs, const T*, const int64_t*, const SphTable*,
                       const double*, double*, int64_t, hipStream_t);
};

struct Int3MLaunch {
  Int3MFn<double>::type fn;
  int nc;
  Int3MFn<float>::type fn32;  // FP32 window of the mixed path
  int nc32;
};

template <int LA, int LB, int LC, typename T>
void launch_int3c2e_mm(DevShells bra, DevShells ket, PairArgs pa, AuxArgs xa, const T* EtG, const int64_t* etoff,
                       const SphTable* sph, const double* boys, double* out, int64_t ldo, hipStream_t st) {
  using K = Cls3m<LA, LB, LC, T>;
  const int nchunks = (xa.hi - xa.lo + K::NC - 1) / K::NC;
  if (pa.npairs <= 0 || nchunks <= 0) return;
  hipLaunchKernelGGL((int3c2e_mm_kernel<LA, LB, LC, T>), dim3(8 * nchunks, (pa.npairs + 7) / 8), dim3(K::NT), 0,
                     st, bra, ket, pa, xa, EtG, etoff, sph, boys, out, ldo);
}

template <int LA, int LB, int LC>
Int3MLaunch make_int3m() {
  Int3MLaunch r{nullptr, 0, nullptr, 0};
  if constexpr (Cls3m<LA, LB, LC, double>::OK) {
    r.fn = &launch_int3c2e_mm<LA, LB, LC, double>;
    r.nc = Cls3m<LA, LB, LC, double>::NC;
  }
  if constexpr (Cls3m<LA, LB, LC, float>::OK) {
    r.fn32 = &launch_int3c2e_mm<LA, LB, LC, float>;
    r.nc32 = Cls3m<LA, LB, LC, float>::NC;
  }
  return r;
}

template <typename TO>
struct EtFn {
  typedef void (*type)(DevShells, const int*, const int*, const int*, const int64_t*, int, const SphTable*, TO*,
                       hipStream_t);
};

struct EtLaunch {
  EtFn<double>::type fn;
  EtFn<float>::type fn32;
  int64_t etpp;  // elements per primitive pair
};

template <int LA, int LB, typename TO>
void launch_et(DevShells bra, const int* s1, const int* s2, const int* list, const int64_t* off, int nlist,
               const SphTable* sph, TO* EtG, hipStream_t st) {
  if (nlist <= 0) return;
  hipLaunchKernelGGL((int3c2e_et_kernel<LA, LB, TO>), dim3(nlist), dim3(256), 0, st, bra, s1, s2, list, off, sph,
                     EtG);
}

template <int LA, int LB>
EtLaunch make_et() {
  return EtLaunch{&launch_et<LA, LB, double>, &launch_et<LA, LB, float>, Cls3m<LA, LB, 0>::ETPP};
}

using Int3MTable = Int3MLaunch[LMAX_ORB + 1][LMAX_ORB + 1][LMAX_AUX + 1];
using EtTable = EtLaunch[LMAX_ORB + 1][LMAX_ORB + 1];

}  // namespace dfhip
