// Screened 3c2e (ab|C) blocks, written to a dense (pair function x aux
// function) matrix in HBM.  Also used for the 2c2e metric (P|Q) with a
// single-centre "pair" (s2 < 0).
//
// One workgroup = one shell pair (a, b) x a chunk of NC auxiliary shells of
// one angular momentum.  Per primitive pair ij the spherical bra matrix Et
// (nsph(a) nsph(b) x nherm(la+lb)) is built once in LDS and reused by every
// auxiliary shell of the chunk; per auxiliary primitive the closed-form
// R_tuv and the Hobson-transformed W[C][tuv][m] are built in LDS; the
// contracted block accumulates in registers as acc += Et . W.
#pragma once

#include "hermite.h"

namespace dfhip {

struct PairArgs {
  const int* s1;       // first shell (L = LA) of each pair of the class
  const int* s2;       // second shell (L = LB), or -1 for a single-centre bra
  const int64_t* row;  // output row of the pair's first function
  const int* nsurv;    // screened-in prefix length of the aux list, per pair
  int npairs;
  const double* qvl = nullptr;  // QVL pair records (8 doubles per pair), or null for Schwarz
  const int* nbeg = nullptr;    // first aux-list entry per pair (generated and LDS kernels; null = 0)
  // QVL only, matrix-core kernels: the kept list entries of each pair's
  // prefix, kidx[koff[p] .. koff[p + 1]) ascending (null: every prefix entry
  // is computed, the dropped ones as zeros)
  const int64_t* koff = nullptr;
  const int* kidx = nullptr;
  // optional, with kidx: the kept-list window of each pair for this launch
  // (qvl_windows), instead of two searches per workgroup
  const int* kja = nullptr;
  const int* kjb = nullptr;
};

struct AuxArgs {
  const int* shell;    // aux shells of one L in processing order
  const int64_t* col;  // output column of each listed aux shell
  int lo, hi;          // list window processed by this launch
  int kmax;            // largest primitive count in the list
  const double* qvl = nullptr;  // QVL aux records (4 doubles per aux shell), used with PairArgs::qvl
  double thresh = 0.0;
};

// Whether triplet (pair pidx, aux shell sh at C) survives QVL screening; always
// true under Schwarz screening, whose screened-in set is the nsurv prefix.
__device__ inline bool triplet_kept(const PairArgs& pa, const AuxArgs& xa, int pidx, int sh, int LC, double Cx,
                                    double Cy, double Cz) {
  return pa.qvl == nullptr || qvl_keep(pa.qvl + 8 * (size_t)pidx, xa.qvl + 4 * (size_t)sh, Cx, Cy, Cz, LC, xa.thresh);
}

// Number of terms of all real solid harmonics of angular momentum L.
__host__ __device__ constexpr int nsph_terms(int L) {
  return L == 0 ? 1 : L == 1 ? 3 : L == 2 ? 8 : L == 3 ? 16 : L == 4 ? 28 : L == 5 ? 46 : 70;
}
// Largest number of terms of one real solid harmonic (L, m).
__host__ __device__ constexpr int sph_max_terms(int L) {
  return L <= 1 ? 1 : L <= 3 ? 3 : L <= 5 ? 6 : 10;
}

template <int LA, int LB, int LC>
struct Cls3 {
  static constexpr int LAB = LA + LB, L = LAB + LC;
  static constexpr int NA = nsph(LA), NB = nsph(LB), NAB = NA * NB, NCC = nsph(LC);
  static constexpr int H = nherm(LAB), NR = nherm(L), NRLO = (LC == 0) ? 0 : nherm(LC - 1);
  static constexpr int ESZ = (LA + 1) * (LB + 1) * (LAB + 1);
  static constexpr int NT = 128;
  static constexpr int FIXED = 3 * ESZ + NAB * H;
  static constexpr int PERC = 4 + 4 * (L + 1) + NR + H * NCC;
  static constexpr int SMALL = 8192, LARGE = 19300;  // doubles (64 KB / 151 KB; ~5 KB more for tables)
  static constexpr int NC_LDS = (SMALL - FIXED) / PERC >= 2 ? (SMALL - FIXED) / PERC : (LARGE - FIXED) / PERC;
  static constexpr int NC_OUT = (16 * NT) / (NAB * NCC) > 1 ? (16 * NT) / (NAB * NCC) : 1;
  static constexpr int NC0 = NC_LDS < NC_OUT ? NC_LDS : NC_OUT;
  static constexpr int NC = NC0 > 128 ? 128 : (NC0 < 1 ? 1 : NC0);
  static constexpr int LDS = FIXED + NC * PERC;
  static constexpr int OPT = (NAB * NC * NCC + NT - 1) / NT;
  static_assert(NC_LDS >= 1, "class does not fit in LDS");
  static_assert(LDS <= LARGE, "LDS budget");
};

template <int LA, int LB, int LC>
__global__ __launch_bounds__(128) void int3c2e_kernel(DevShells bra, DevShells ket, PairArgs pa, AuxArgs xa,
                                                      const SphTable* __restrict__ sph,
                                                      const double* __restrict__ boys, double* __restrict__ out,
                                                      int64_t ldo) {
  using K = Cls3<LA, LB, LC>;
  constexpr int L = K::L, H = K::H, NR = K::NR, NCC = K::NCC, NAB = K::NAB, NT = K::NT;
  __shared__ double lds[K::LDS];
  __shared__ int cshell[K::NC];
  __shared__ int ckeep[K::NC];
  // tables: packed (t,u,v) per Hermite index, t!/(i!(t-2i)!), solid-harmonic terms of LA, LB, LC
  constexpr int LH = L / 2 + 1;
  __shared__ int tuv[NR];
  __shared__ double hc[(L + 1) * LH];
  __shared__ int sa[2 * LA + 2], sb[2 * LB + 2], sc[2 * LC + 2];
  __shared__ SphTerm ta[nsph_terms(LA)], tb[nsph_terms(LB)], tc[nsph_terms(LC)];

  const int pidx = blockIdx.x;
  const int s1 = pa.s1[pidx], s2 = pa.s2[pidx];
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  // this workgroup's chunk of the list, cut to the pair's window [nbeg, nsurv)
  const int cchunk = xa.lo + (int)blockIdx.y * K::NC;
  const int c0 = pa.nbeg ? max(cchunk, pa.nbeg[pidx]) : cchunk;
  if (c0 >= min(cend, cchunk + K::NC)) return;
  const int nc = min(cchunk + K::NC, cend) - c0;
  const int tid = threadIdx.x;

  double* Ex = lds;
  double* Et = lds + 3 * K::ESZ;
  double* cxyz = Et + NAB * H;           // NC*4
  double* gN = cxyz + 4 * K::NC;         // NC*(L+1)
  double* pw = gN + K::NC * (L + 1);     // NC*3*(L+1)
  double* R = pw + 3 * K::NC * (L + 1);  // NC*NR
  double* W = R + K::NC * NR;            // NC*H*NCC

  for (int h = tid; h < NR; h += NT) {
    int t, u, v;
    herm_tuv(h, t, u, v);
    tuv[h] = t | (u << 8) | (v << 16);
  }
  for (int e = tid; e < (L + 1) * LH; e += NT) {
    const int t = e / LH, i = e - t * LH;
    double c = 0.0;
    if (2 * i <= t) {
      c = 1.0;
      for (int z = 0; z < i; z++) c *= (double)((t - 2 * z) * (t - 2 * z - 1)) / (z + 1);
    }
    hc[e] = c;
  }
  {
    const int a0 = sph->start[LA][0], b0 = sph->start[LB][0], x0 = sph->start[LC][0];
    if (tid < 2 * LA + 2) sa[tid] = sph->start[LA][tid] - a0;
    if (tid < 2 * LB + 2) sb[tid] = sph->start[LB][tid] - b0;
    if (tid < 2 * LC + 2) sc[tid] = sph->start[LC][tid] - x0;
    for (int e = tid; e < nsph_terms(LA); e += NT) ta[e] = sph->terms[a0 + e];
    for (int e = tid; e < nsph_terms(LB); e += NT) tb[e] = sph->terms[b0 + e];
    for (int e = tid; e < nsph_terms(LC); e += NT) tc[e] = sph->terms[x0 + e];
  }
  for (int c = tid; c < nc; c += NT) {
    const int sh = xa.shell[c0 + c];
    cshell[c] = sh;
    cxyz[4 * c + 0] = ket.xyz[3 * sh + 0];
    cxyz[4 * c + 1] = ket.xyz[3 * sh + 1];
    cxyz[4 * c + 2] = ket.xyz[3 * sh + 2];
    ckeep[c] = triplet_kept(pa, xa, pidx, sh, LC, ket.xyz[3 * sh + 0], ket.xyz[3 * sh + 1], ket.xyz[3 * sh + 2]);
  }

  double A[3], B[3];
  for (int d = 0; d < 3; d++) A[d] = bra.xyz[3 * s1 + d];
  const int npa = bra.nprim[s1], pa0 = bra.pptr[s1];
  int npb = 1, pb0 = 0;
  if (s2 >= 0) {
    for (int d = 0; d < 3; d++) B[d] = bra.xyz[3 * s2 + d];
    npb = bra.nprim[s2];
    pb0 = bra.pptr[s2];
  } else {
    for (int d = 0; d < 3; d++) B[d] = A[d];
  }

  double acc[K::OPT];
#pragma unroll
  for (int r = 0; r < K::OPT; r++) acc[r] = 0.0;

  const int nW = nc * H * NCC;
  const double lc_sign = (LC & 1) ? -1.0 : 1.0;

  for (int ij = 0; ij < npa * npb; ij++) {
    const int i = ij / npb, j = ij - (ij / npb) * npb;
    const double a = bra.exps[pa0 + i], ca = bra.coefs[pa0 + i];
    const double b = (s2 >= 0) ? bra.exps[pb0 + j] : 0.0;
    const double cb = (s2 >= 0) ? bra.coefs[pb0 + j] : 1.0;
    const double p = a + b;
    double P[3];
    for (int d = 0; d < 3; d++) P[d] = (a * A[d] + b * B[d]) / p;

    __syncthreads();  // previous primitive pair is done with Et and W
    if (tid < 3) herm_E_dir<LA, LB>(a, b, A[tid], B[tid], Ex + tid * K::ESZ);
    for (int k = tid; k < nW; k += NT) W[k] = 0.0;
    __syncthreads();
    const double cab = ca * cb;
    for (int k = tid; k < NAB * H; k += NT) {
      const int munu = k / H, h = k - (k / H) * H;
      const int mu = munu / K::NB, nu = munu - (munu / K::NB) * K::NB;
      const int tv = tuv[h];
      Et[k] = cab * bra_element_lds<LA, LB>(sa, ta, sb, tb, mu, nu, tv & 255, (tv >> 8) & 255, tv >> 16, Ex,
                                            Ex + K::ESZ, Ex + 2 * K::ESZ);
    }

    for (int kp = 0; kp < xa.kmax; kp++) {
      // per-C scaled Boys values and powers of 2(P-C)
      for (int k = tid; k < nc * 4 * (L + 1); k += NT) {
        const int c = k / (4 * (L + 1)), q = k - c * (4 * (L + 1));
        const int sh = cshell[c];
        const bool live = kp < ket.nprim[sh] && ckeep[c];
        const double g = live ? ket.exps[ket.pptr[sh] + kp] : 1.0;
        const double PC0 = P[0] - cxyz[4 * c + 0], PC1 = P[1] - cxyz[4 * c + 1], PC2 = P[2] - cxyz[4 * c + 2];
        if (q <= L) {
          double val = 0.0;
          if (live) {
            const double dk = ket.coefs[ket.pptr[sh] + kp];
            const double alpha = p * g / (p + g);
            const double T = alpha * (PC0 * PC0 + PC1 * PC1 + PC2 * PC2);
            double scale = 2.0 * PI * PI * sqrt(PI) / (p * g * sqrt(p + g)) * dk * lc_sign;
            const double inv2g = 0.5 / g;
            for (int e = 0; e < LC; e++) scale *= inv2g;
            for (int e = 0; e < q; e++) scale *= -alpha;
            val = scale * boys_f(q, T, boys);
          }
          gN[c * (L + 1) + q] = val;
        } else {
          const int qq = q - (L + 1), d = qq / (L + 1), e = qq - d * (L + 1);
          const double x = 2.0 * (d == 0 ? PC0 : (d == 1 ? PC1 : PC2));
          double v = 1.0;
          for (int z = 0; z < e; z++) v *= x;
          pw[c * 3 * (L + 1) + d * (L + 1) + e] = v;
        }
      }
      __syncthreads();
      for (int k = tid; k < nc * (NR - K::NRLO); k += NT) {
        const int c = k / (NR - K::NRLO), h = K::NRLO + (k - c * (NR - K::NRLO));
        const int tv = tuv[h];
        R[c * NR + h] = herm_R_tab(tv & 255, (tv >> 8) & 255, tv >> 16, pw + c * 3 * (L + 1), L, gN + c * (L + 1), hc, LH);
      }
      __syncthreads();
      for (int k = tid; k < nW; k += NT) {
        const int c = k / (H * NCC), rr = k - c * (H * NCC), h = rr / NCC, m = rr - (rr / NCC) * NCC;
        const int tv = tuv[h];
        const int t = tv & 255, u = (tv >> 8) & 255, v = tv >> 16;
        double s = 0.0;
        const double* Rc = R + c * NR;
        for (int q = sc[m]; q < sc[m + 1]; q++) {
          const SphTerm tm = tc[q];
          s = fma(tm.c, Rc[herm_index(t + tm.lx, u + tm.ly, v + tm.lz)], s);
        }
        W[k] += s;
      }
      __syncthreads();
    }

    const int ncol = nc * NCC;
#pragma unroll
    for (int r = 0; r < K::OPT; r++) {
      const int o = tid + r * NT;
      if (o < NAB * ncol) {
        const int munu = o / ncol, col = o - (o / ncol) * ncol;
        const int c = col / NCC, m = col - (col / NCC) * NCC;
        const double* e = Et + munu * H;
        const double* w = W + c * H * NCC + m;
        double s = 0.0;
        for (int h = 0; h < H; h++) s += e[h] * w[h * NCC];
        acc[r] += s;
      }
    }
  }

  const int ncol = nc * NCC;
  const int64_t row0 = pa.row[pidx];
#pragma unroll
  for (int r = 0; r < K::OPT; r++) {
    const int o = tid + r * NT;
    if (o < NAB * ncol) {
      const int munu = o / ncol, col = o - (o / ncol) * ncol;
      const int c = col / NCC, m = col - (col / NCC) * NCC;
      out[(row0 + munu) * ldo + xa.col[c0 + c] + m] = acc[r];
    }
  }
}

struct Int3Launch {
  void (*fn)(DevShells, DevShells, PairArgs, AuxArgs, const SphTable*, const double*, double*, int64_t,
             hipStream_t);
  int nc;
  int lds_bytes;
};

template <int LA, int LB, int LC>
void launch_int3c2e(DevShells bra, DevShells ket, PairArgs pa, AuxArgs xa, const SphTable* sph,
                    const double* boys, double* out, int64_t ldo, hipStream_t st) {
  using K = Cls3<LA, LB, LC>;
  const int nchunks = (xa.hi - xa.lo + K::NC - 1) / K::NC;
  if (pa.npairs <= 0 || nchunks <= 0) return;
  dim3 grid(pa.npairs, nchunks);
  hipLaunchKernelGGL((int3c2e_kernel<LA, LB, LC>), grid, dim3(K::NT), 0, st, bra, ket, pa, xa, sph, boys, out,
                     ldo);
}

template <int LA, int LB, int LC>
Int3Launch make_int3() {
  using K = Cls3<LA, LB, LC>;
  return Int3Launch{&launch_int3c2e<LA, LB, LC>, K::NC, (int)(K::LDS * sizeof(double))};
}

using Int3Table = Int3Launch[LMAX_AUX + 1][LMAX_ORB + 1][LMAX_AUX + 1];

}  // namespace dfhip
