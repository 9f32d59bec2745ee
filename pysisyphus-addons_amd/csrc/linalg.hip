// Dense FP64 kernels of the density-fitting path:
//   * pair-packing of the (transition) densities,
//   * the contraction d(P, I) = sum_k T(k, P) Dp(k, I) (split-K GEMM),
//   * the blocked Cholesky of the metric bordered by d, which yields
//     L (dpotrf('L'), common/linalg.f90:78), B^T = (L^-1 d)^T
//     (dtrtri + dtrmm, diabatization/intor.f90:198) and -B^T B,
//   * the 8-fold Coulomb-tensor assembly (diabatization/intor.f90:241-269),
//   * Schwarz norms of the auxiliary shells from the metric's diagonal blocks,
//   * a Boys-function probe.
#include "hermite.h"
#include "linalg.h"

namespace dfhip {

// ---------------------------------------------------------------------------
// Density packing: Dp[(row(p) + mu*n2 + nu) * ldd + I] =
//   D_I(f1+mu, f2+nu) + D_I(f2+nu, f1+mu)   (s1 != s2)
//   D_I(f1+mu, f2+nu)                        (s1 == s2)
// (the symmetry use of diabatization/intor.f90:162-169).
// out[c * nr + r] = in[r * nc + c] (an nr x nc row-major matrix turned):
// 32 x 64 tiles through LDS, reads and writes along rows.
constexpr int TR_R = 32, TR_C = 64;
__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ in, int64_t nr, int64_t nc,
                                                        double* __restrict__ out) {
  __shared__ double tile[TR_R][TR_C + 1];
  const int64_t c0 = (int64_t)blockIdx.x * TR_C, r0 = (int64_t)blockIdx.y * TR_R;
  const int tid = threadIdx.x;
  for (int e = tid; e < TR_R * TR_C; e += 256) {
    const int r = e / TR_C, c = e % TR_C;
    if (r0 + r < nr && c0 + c < nc) tile[r][c] = in[(r0 + r) * nc + c0 + c];
  }
  __syncthreads();
  for (int e = tid; e < TR_R * TR_C; e += 256) {
    const int c = e / TR_R, r = e % TR_R;
    if (r0 + r < nr && c0 + c < nc) out[(c0 + c) * nr + r0 + r] = tile[r][c];
  }
}

void transpose(const double* in, int64_t nr, int64_t nc, double* out, hipStream_t st) {
  if (nr <= 0 || nc <= 0) return;
  const dim3 grid((unsigned)((nc + TR_C - 1) / TR_C), (unsigned)((nr + TR_R - 1) / TR_R));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, st, in, nr, nc, out);
}

// One thread per packed entry: x over densities, y over rows (pair
// functions), PK_ROWS rows per thread; the pair of a row comes from rowpair.
constexpr int PK_ROWS = 4;
__global__ __launch_bounds__(256) void pack_density_kernel(const int* __restrict__ rowpair,
                                                           const int* __restrict__ s1v, const int* __restrict__ s2v,
                                                           const int64_t* __restrict__ rowv,
                                                           const int* __restrict__ Lv, const int* __restrict__ f0v,
                                                           const double* __restrict__ D, int64_t sI, int64_t si,
                                                           int64_t sj, int ndens, int64_t npf,
                                                           double* __restrict__ Dp, int ldd) {
  const int64_t rb = ((int64_t)blockIdx.x * PK_ROWS) * blockDim.y + threadIdx.y;
#pragma unroll
  for (int k = 0; k < PK_ROWS; k++) {
    const int64_t r = rb + (int64_t)k * blockDim.y;
    if (r >= npf) return;
    const int p = rowpair[r];
    const int s1 = s1v[p], s2 = s2v[p];
    const int n2 = 2 * Lv[s2] + 1;
    const int mn = (int)(r - rowv[p]);
    const int mu = mn / n2, nu = mn - mu * n2;
    const int64_t i = f0v[s1] + mu, j = f0v[s2] + nu;
    const double* __restrict__ Dij = D + i * si + j * sj;
    const double* __restrict__ Dji = D + j * si + i * sj;
    for (int I = threadIdx.x; I < ndens; I += blockDim.x) {
      double v = Dij[I * sI];
      if (s1 != s2) v += Dji[I * sI];
      Dp[r * ldd + I] = v;
    }
  }
}

void pack_density(const int* rowpair, const int* s1, const int* s2, const int64_t* row, int64_t npf, const int* L,
                  const int* f0, const double* D, int64_t sI, int64_t si, int64_t sj, int ndens, double* Dp, int ldd,
                  hipStream_t st) {
  if (npf <= 0 || ndens <= 0) return;
  const int bx = std::min(ndens, 64), by = std::max(1, 256 / bx);
  const int64_t nblk = (npf + (int64_t)by * PK_ROWS - 1) / ((int64_t)by * PK_ROWS);
  hipLaunchKernelGGL(pack_density_kernel, dim3((unsigned)nblk), dim3(bx, by), 0, st, rowpair, s1, s2, row, L, f0, D,
                     sI, si, sj, ndens, npf, Dp, ldd);
}

// The same packing for densities whose density index is the slowest
// (D[I*sI + i*nbf + j], the C layout): a tiled transpose.  One workgroup =
// one row shell s x 32 densities x 64 columns j.  For each row i = f(s) + mu
// the tile D[I0..I0+32][i][j0..j0+64] is read along j (coalesced), turned in
// LDS and written along I into the packed row of the element's pair:
//   pass 0: elements whose pair has s as first shell, Dp = D (stored);
//   pass 1: elements whose pair has s as second shell (s1 != s2), Dp += D^T.
// Every packed entry gets its pass-0 value from one thread and, for s1 != s2,
// its pass-1 addition from one thread of the next launch, so the result does
// not depend on the schedule.  pairof[s*nsh + t] = pair index of shells {s, t}
// with bit 30 set when the pair is oriented (s, t); fshell[j] = shell of
// function j.
constexpr int PK_I = 32, PK_J = 64;
__global__ __launch_bounds__(256) void pack_density_t_kernel(const int* __restrict__ pairof,
                                                             const int* __restrict__ fshell,
                                                             const int64_t* __restrict__ rowv,
                                                             const int* __restrict__ Lv, const int* __restrict__ f0v,
                                                             int nsh, int nbf, const double* __restrict__ D,
                                                             int64_t sI, int ndens, double* __restrict__ Dp, int ldd,
                                                             int pass) {
  __shared__ double tile[PK_I][PK_J + 1];
  const int s = blockIdx.x, I0 = blockIdx.y * PK_I, j0 = blockIdx.z * PK_J;
  const int ns = 2 * Lv[s] + 1, fs = f0v[s];
  const int tid = threadIdx.x;
  const int nI = min(PK_I, ndens - I0), nj = min(PK_J, nbf - j0);
  for (int mu = 0; mu < ns; mu++) {
    __syncthreads();
    for (int e = tid; e < PK_I * PK_J; e += 256) {
      const int I = e / PK_J, j = e % PK_J;
      tile[I][j] = (I < nI && j < nj) ? D[(int64_t)(I0 + I) * sI + (int64_t)(fs + mu) * nbf + j0 + j] : 0.0;
    }
    __syncthreads();
    // column j of the tile -> its pair, packed position; 32 densities per column
    for (int e = tid; e < PK_I * PK_J; e += 256) {
      const int j = e / PK_I, I = e % PK_I;
      if (j >= nj || I >= nI) continue;
      const int col = j0 + j, t = fshell[col];
      const int code = pairof[s * nsh + t];
      const bool first = (code >> 30) & 1;
      const int p = code & ((1 << 30) - 1);
      const int nt = 2 * Lv[t] + 1, nu = col - f0v[t];
      if (pass == 0 && first) {
        Dp[(rowv[p] + mu * nt + nu) * ldd + I0 + I] = tile[I][j];
      } else if (pass == 1 && !first) {  // pair (t, s): position (nu, mu)
        double* d = &Dp[(rowv[p] + nu * ns + mu) * ldd + I0 + I];
        *d += tile[I][j];
      }
    }
  }
}

void pack_density_t(const int* pairof, const int* fshell, const int64_t* row, const int* L, const int* f0, int nsh,
                    int nbf, const double* D, int64_t sI, int ndens, double* Dp, int ldd, hipStream_t st) {
  if (nsh <= 0 || nbf <= 0 || ndens <= 0) return;
  const dim3 grid(nsh, (ndens + PK_I - 1) / PK_I, (nbf + PK_J - 1) / PK_J);
  for (int pass = 0; pass < 2; pass++)
    hipLaunchKernelGGL(pack_density_t_kernel, grid, dim3(256), 0, st, pairof, fshell, row, L, f0, nsh, nbf, D, sI,
                       ndens, Dp, ldd, pass);
}

// One thread per pair: its class by a scan of the class starts, then its
// shells by the class's index order (triangular for la == lb).
__global__ void gen_pairs_kernel(PairClassTable t, const int* __restrict__ shflat, int* __restrict__ s1,
                                 int* __restrict__ s2, int64_t* __restrict__ prow0) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= t.begin[t.ncls]) return;
  int k = 0;
  while (k + 1 < t.ncls && p >= t.begin[k + 1]) k++;
  const int i = p - t.begin[k];
  int ia, ib;
  if (t.la[k] != t.lb[k]) {
    ia = i / t.n_b[k];
    ib = i - ia * t.n_b[k];
  } else {
    ia = (int)((sqrt(8.0 * i + 1.0) - 1.0) * 0.5);
    while ((int64_t)(ia + 1) * (ia + 2) / 2 <= i) ia++;
    while ((int64_t)ia * (ia + 1) / 2 > i) ia--;
    ib = i - ia * (ia + 1) / 2;
  }
  s1[p] = shflat[t.off_a[k] + ia];
  s2[p] = shflat[t.off_b[k] + ib];
  if (prow0) prow0[p] = t.row0[k] + (int64_t)i * t.w[k];
}

void gen_pairs(const PairClassTable& t, const int* shflat, int* s1, int* s2, int64_t* prow0, hipStream_t st) {
  const int np = t.begin[t.ncls];
  if (np > 0)
    hipLaunchKernelGGL(gen_pairs_kernel, dim3((np + 255) / 256), dim3(256), 0, st, t, shflat, s1, s2, prow0);
}

// rowpair[r] = p for the rows r = row[p] .. row[p+1]-1 of pair p (the last
// pair ends at npf), and with dprow rowperm[r] = dprow[p] + r - row[p]: one
// workgroup per pair.
__global__ void fill_rowpair_kernel(const int64_t* __restrict__ row, int npairs, int64_t npf,
                                    int* __restrict__ rowpair, const int64_t* __restrict__ dprow,
                                    int* __restrict__ rowperm) {
  const int p = blockIdx.x;
  const int64_t r0 = row[p], r1 = p + 1 < npairs ? row[p + 1] : npf;
  const int64_t d0 = dprow ? dprow[p] : 0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    rowpair[r] = p;
    if (dprow) rowperm[r] = (int)(d0 + (r - r0));
  }
}

void fill_rowpair(const int64_t* row, int npairs, int64_t npf, int* rowpair, const int64_t* dprow, int* rowperm,
                  hipStream_t st) {
  if (npairs > 0)
    hipLaunchKernelGGL(fill_rowpair_kernel, dim3(npairs), dim3(64), 0, st, row, npairs, npf, rowpair, dprow,
                       rowperm);
}

// ---------------------------------------------------------------------------
// part[s][P][I] = sum_{k in split s} T[k][P] * Dp[rowperm[k]][I]
// T: npf x ncol (row-major, ldt), Dp: npf x ndens (row-major, ldd).
// Only rows that can be non-zero for a column tile are visited: the rows of
// the tile are the concatenation of nseg row segments [seg_lo, seg_hi) per
// column tile (host-computed from the screened prefixes), split evenly over
// the split index s.
// Output tile per workgroup: GP = 128 columns P x 16 IPT densities, 256
// threads with 8 x IPT each (IPT from the number of densities, so that 36
// densities take one tile of 48), K step 16.
constexpr int GP = 128, GK = 16;

template <int IPT>
__global__ __launch_bounds__(256) void gemm_tn_splitk_kernel(const double* __restrict__ T, int64_t ldt,
                                                             const double* __restrict__ Dp, int ldd, int ncol,
                                                             int ndens, int nseg, const int64_t* __restrict__ seg,
                                                             int nsplit, const ScreenMask mk,
                                                             double* __restrict__ part) {
  constexpr int GI = 16 * IPT;
  __shared__ double sT[GK][GP + 1];
  __shared__ double sD[GK][GI + 1];
  const int tcol = blockIdx.x;
  const int tp = blockIdx.x * GP, ti = blockIdx.y * GI, s = blockIdx.z;
  const int64_t* sg = seg + (size_t)tcol * (2 * nseg + 1);  // [total, lo0, hi0, lo1, hi1, ...]
  const int64_t tot = sg[0];
  const int64_t chunk = ((tot + nsplit - 1) / nsplit + GK - 1) / GK * GK;
  int64_t q0 = (int64_t)s * chunk, q1 = min(tot, q0 + chunk);
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;  // outputs P = tp + ty*8 + a, I = ti + tx*IPT + b
  // this lane always loads column tp + (tid % GP) of T: its screened prefix test
  const int mcol = tp + (tid % GP);
  const bool colok = mcol < ncol;
  const int* __restrict__ ns_col = colok ? mk.nsurv + (size_t)mk.colL[mcol] * mk.npairs : mk.nsurv;
  const int ia_col = colok ? mk.colia[mcol] : 0;
  double acc[8][IPT] = {};
  // walk the segments overlapping [q0, q1) of the concatenated row range
  int64_t base = 0;
  for (int g = 0; g < nseg && q0 < q1; g++) {
    const int64_t lo = sg[1 + 2 * g], hi = sg[2 + 2 * g], len = hi - lo;
    if (q0 >= base + len) {
      base += len;
      continue;
    }
    const int64_t a0 = lo + (q0 - base), a1 = lo + min(len, q1 - base);
    // K blocks of GK rows; the global loads of block kb + GK are issued
    // before the FMAs of block kb (unconditional loads from clamped, in-bounds
    // addresses, then selects)
    const int mcolc = colok ? mcol : ncol - 1;
    constexpr int NZ = GK * GP / 256, NZD = (GK * GI + 255) / 256;
    double tv[NZ], dv[NZD];
    int ns[NZ];
    auto load = [&](int64_t kb) {
#pragma unroll
      for (int z = 0; z < NZ; z++) {
        const int e = tid + 256 * z, kk = e / GP;
        const int64_t kc = min(kb + kk, a1 - 1);
        tv[z] = T[kc * ldt + mcolc];
        ns[z] = ns_col[mk.rowpair[kc]];
      }
#pragma unroll
      for (int z = 0; z < NZD; z++) {
        const int e = min(tid + 256 * z, GK * GI - 1), kk = e / GI, c = e - kk * GI;
        const int64_t kc = min(kb + kk, a1 - 1);
        dv[z] = Dp[(int64_t)mk.rowperm[kc] * ldd + min(ti + c, ndens - 1)];
      }
    };
    load(a0);
    for (int64_t kb = a0; kb < a1; kb += GK) {
#pragma unroll
      for (int z = 0; z < NZ; z++) {
        const int e = tid + 256 * z, kk = e / GP, c = e - kk * GP;
        const int64_t k = kb + kk;
        // entries of screened-out triplets are never written: mask them here
        const bool live = k < a1 && colok && ia_col < ns[z];
        sT[kk][c] = live ? tv[z] : 0.0;
      }
#pragma unroll
      for (int z = 0; z < NZD; z++) {
        const int e = tid + 256 * z, kk = e / GI, c = e - kk * GI;
        if (e < GK * GI) sD[kk][c] = (kb + kk < a1 && ti + c < ndens) ? dv[z] : 0.0;
      }
      __syncthreads();
      if (kb + GK < a1) load(kb + GK);
#pragma unroll
      for (int kk = 0; kk < GK; kk++) {
        double a[8], b[IPT];
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = sT[kk][ty * 8 + q];
#pragma unroll
        for (int q = 0; q < IPT; q++) b[q] = sD[kk][tx * IPT + q];
#pragma unroll
        for (int x = 0; x < 8; x++)
#pragma unroll
          for (int y = 0; y < IPT; y++) acc[x][y] = fma(a[x], b[y], acc[x][y]);
      }
      __syncthreads();
    }
    q0 = base + len;
    base += len;
  }
  for (int x = 0; x < 8; x++)
    for (int y = 0; y < IPT; y++) {
      const int P = tp + ty * 8 + x, I = ti + tx * IPT + y;
      if (P < ncol && I < ndens) part[((size_t)s * ncol + P) * ndens + I] = acc[x][y];
    }
}

// The same contraction on the FP64 matrix cores (v_mfma_f64_16x16x4):
// output tile MP = 128 columns P x 32 NIT densities per workgroup of 8 waves;
// wave w takes the column tiles 2 (w % 4), 2 (w % 4) + 1 and the NIT density
// tiles from NIT (w / 4) on, 2 NIT accumulator tiles held over the whole K
// range.  K steps of MK = 16 rows staged in LDS as in gemm_tn_splitk_kernel
// (the same masks, the next block's global loads issued before the current
// block's products); per 4-row step a wave reads 2 + NIT fragments from LDS
// for 2 NIT products, where the VALU kernel reads 8 + IPT values per 8 IPT
// FMAs.  A[m][k] = T[k][P], lane l (m = l%16, k = l/16); B[k][n] = Dp[k][I],
// lane l (k = l/16, n = l%16); register r of lane l holds the output
// (P row l/16 + 4 r, density l%16) of its tile.
typedef double gemm_dbl4 __attribute__((ext_vector_type(4)));
constexpr int MP = 128, MK = 16;

template <int NIT>
__global__ __launch_bounds__(512) void gemm_tn_mfma_kernel(const double* __restrict__ T, int64_t ldt,
                                                           const double* __restrict__ Dp, int ldd, int ncol,
                                                           int ndens, int nseg, const int64_t* __restrict__ seg,
                                                           int nsplit, const ScreenMask mk,
                                                           double* __restrict__ part) {
  constexpr int MI = 32 * NIT, NT = 512;
  __shared__ double sT[MK][MP + 1];
  __shared__ double sD[MK][MI + 1];
  const int tcol = blockIdx.x;
  const int tp = blockIdx.x * MP, ti = blockIdx.y * MI, s = blockIdx.z;
  const int64_t* sg = seg + (size_t)tcol * (2 * nseg + 1);  // [total, lo0, hi0, lo1, hi1, ...]
  const int64_t tot = sg[0];
  const int64_t chunk = ((tot + nsplit - 1) / nsplit + MK - 1) / MK * MK;
  int64_t q0 = (int64_t)s * chunk, q1 = min(tot, q0 + chunk);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int pt0 = 2 * (wave & 3), it0 = NIT * (wave >> 2);
  // this lane always loads column tp + (tid % MP) of T: its screened prefix test
  const int mcol = tp + (tid % MP);
  const bool colok = mcol < ncol;
  const int* __restrict__ ns_col = colok ? mk.nsurv + (size_t)mk.colL[mcol] * mk.npairs : mk.nsurv;
  const int ia_col = colok ? mk.colia[mcol] : 0;
  gemm_dbl4 acc[2][NIT];
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int i = 0; i < NIT; i++) acc[j][i] = gemm_dbl4{0.0, 0.0, 0.0, 0.0};
  int64_t base = 0;
  for (int g = 0; g < nseg && q0 < q1; g++) {
    const int64_t lo = sg[1 + 2 * g], hi = sg[2 + 2 * g], len = hi - lo;
    if (q0 >= base + len) {
      base += len;
      continue;
    }
    const int64_t a0 = lo + (q0 - base), a1 = lo + min(len, q1 - base);
    const int mcolc = colok ? mcol : ncol - 1;
    constexpr int NZ = MK * MP / NT, NZD = (MK * MI + NT - 1) / NT;
    double tv[NZ], dv[NZD];
    int ns[NZ];
    auto load = [&](int64_t kb) {
#pragma unroll
      for (int z = 0; z < NZ; z++) {
        const int e = tid + NT * z, kk = e / MP;
        const int64_t kc = min(kb + kk, a1 - 1);
        tv[z] = T[kc * ldt + mcolc];
        ns[z] = ns_col[mk.rowpair[kc]];
      }
#pragma unroll
      for (int z = 0; z < NZD; z++) {
        const int e = min(tid + NT * z, MK * MI - 1), kk = e / MI, c = e - kk * MI;
        const int64_t kc = min(kb + kk, a1 - 1);
        dv[z] = Dp[(int64_t)mk.rowperm[kc] * ldd + min(ti + c, ndens - 1)];
      }
    };
    load(a0);
    for (int64_t kb = a0; kb < a1; kb += MK) {
#pragma unroll
      for (int z = 0; z < NZ; z++) {
        const int e = tid + NT * z, kk = e / MP, c = e - kk * MP;
        const int64_t k = kb + kk;
        // entries of screened-out triplets are never written: mask them here
        const bool live = k < a1 && colok && ia_col < ns[z];
        sT[kk][c] = live ? tv[z] : 0.0;
      }
#pragma unroll
      for (int z = 0; z < NZD; z++) {
        const int e = tid + NT * z, kk = e / MI, c = e - kk * MI;
        if (e < MK * MI) sD[kk][c] = (kb + kk < a1 && ti + c < ndens) ? dv[z] : 0.0;
      }
      __syncthreads();
      if (kb + MK < a1) load(kb + MK);
#pragma unroll
      for (int k4 = 0; k4 < MK / 4; k4++) {
        const int kr = 4 * k4 + lk;
        double a[2], b[NIT];
#pragma unroll
        for (int j = 0; j < 2; j++) a[j] = sT[kr][16 * (pt0 + j) + lr];
#pragma unroll
        for (int i = 0; i < NIT; i++) b[i] = sD[kr][16 * (it0 + i) + lr];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int i = 0; i < NIT; i++) acc[j][i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[j], b[i], acc[j][i], 0, 0, 0);
      }
      __syncthreads();
    }
    q0 = base + len;
    base += len;
  }
#pragma unroll
  for (int j = 0; j < 2; j++)
#pragma unroll
    for (int i = 0; i < NIT; i++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int P = tp + 16 * (pt0 + j) + lk + 4 * r, I = ti + 16 * (it0 + i) + lr;
        if (P < ncol && I < ndens) part[((size_t)s * ncol + P) * ndens + I] = acc[j][i][r];
      }
}

// Output-tile codes of the contraction: 1-5 and 7 = densities per thread of
// the VALU kernel (tiles of 16 x code densities), 100 + NIT (NIT = 2..7) =
// the matrix-core kernel with tiles of 32 NIT densities.  The default: VALU
// up to 64 densities, the matrix cores above.
int contract_ipt(int ndens) {
  if (ndens > 64) return 100 + std::min(7, (ndens + 31) / 32);
  return ndens <= 16 ? 1 : ndens <= 32 ? 2 : ndens <= 48 ? 3 : 4;
}
bool contract_ipt_ok(int ipt) { return (ipt >= 1 && ipt <= 5) || ipt == 7 || (ipt >= 102 && ipt <= 107); }
static int contract_width(int ipt) { return ipt > 100 ? 32 * (ipt - 100) : 16 * ipt; }

int64_t contract_ntiles(int64_t ncol, int ndens, int ipt) {
  const int code = contract_ipt_ok(ipt) ? ipt : contract_ipt(ndens);
  const int gi = contract_width(code);
  static_assert(GP == MP, "both kernels tile 128 columns");
  return ((ncol + GP - 1) / GP) * ((ndens + gi - 1) / gi);
}

// d[orig(P) * sP + I * sI] += sum_s part[s][P][I]   (fixed summation order)
__global__ void reduce_unpermute_kernel(const double* __restrict__ part, int nsplit, int ncol, int ndens,
                                        const int* __restrict__ orig, double* __restrict__ d, int64_t sP,
                                        int64_t sI) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)ncol * ndens) return;
  const int P = (int)(e / ndens), I = (int)(e - (int64_t)P * ndens);
  double s = 0.0;
  for (int z = 0; z < nsplit; z++) s += part[((size_t)z * ncol + P) * ndens + I];
  d[orig[P] * sP + I * sI] += s;
}

void contract_T(const double* T, int64_t ldt, const double* Dp, int ldd, int ncol, int ndens, int nseg,
                const int64_t* seg, const ScreenMask& mk, double* part, int nsplit, const int* orig, double* d,
                int64_t sP, int64_t sI, hipStream_t st, int ipt_req) {
  if (ncol <= 0) return;
  const int ipt = contract_ipt_ok(ipt_req) ? ipt_req : contract_ipt(ndens), gi = contract_width(ipt);
  dim3 grid((ncol + GP - 1) / GP, (ndens + gi - 1) / gi, nsplit);
#define DFHIP_MFMA(N)                                                                                          \
  case 100 + N:                                                                                                \
    hipLaunchKernelGGL(gemm_tn_mfma_kernel<N>, grid, dim3(512), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg, \
                       nsplit, mk, part);                                                                      \
    break;
  switch (ipt) {
    DFHIP_MFMA(2)
    DFHIP_MFMA(3)
    DFHIP_MFMA(4)
    DFHIP_MFMA(5)
    DFHIP_MFMA(6)
    DFHIP_MFMA(7)
#undef DFHIP_MFMA
    case 1:
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<1>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
      break;
    case 2:
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<2>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
      break;
    case 3:
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<3>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
      break;
    case 4:
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<4>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
      break;
    case 5:
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<5>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
      break;
    default:  // 7: tiles of 112 densities
      hipLaunchKernelGGL(gemm_tn_splitk_kernel<7>, grid, dim3(256), 0, st, T, ldt, Dp, ldd, ncol, ndens, nseg, seg,
                         nsplit, mk, part);
  }
  const int64_t n = (int64_t)ncol * ndens;
  hipLaunchKernelGGL(reduce_unpermute_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, nsplit, ncol,
                     ndens, orig, d, sP, sI);
}

int contract_tile() { return GP; }


// ---------------------------------------------------------------------------
// Blocked right-looking Cholesky of a matrix bordered by extra rows,
//     [ M ]      M = (P|Q), n x n
//     [ C ]      C: rows below M
// stored lower-column-major in A (lda).  Only the first n columns are
// factored: afterwards the leading block holds L (M = L L^T, dpotrf('L') of
// common/linalg.f90:78) and the border rows hold C L^-T.  With C = I they
// hold L^-T, the transpose of the inverse that the reference forms with
// dtrtri (diabatization/intor.f90:198).
constexpr int CB = 64;
static_assert(CB == 64, "the staging loops of the Cholesky kernels index with shifts by 6 and 7");

// One block column of the Cholesky factorization: the 64x64 diagonal block
// A11 at k0 and 64 rows of the panel A21 below it, as the 128x64 stack
// [A11; A21] factored right-looking in 4-column steps:
//   L11 L11^T = A11,   L21 = A21 L11^-T.
// Every workgroup factors A11 itself (the same arithmetic, so the same
// result) instead of waiting for a separate diagonal kernel.  L11 is not
// stored: nothing after this step reads it (the trailing updates use L21, the
// results are the border rows and the trailing block), and leaving A11 as it
// is keeps it intact for the workgroups still loading it.  512 threads: thread (ti, tk) keeps the 4x4 tile of rows 4ti..
// (ti < 16: A11, lower tiles only; ti >= 16: the panel rows) and columns 4tk..
// in registers.  Per step jb: the owner of diagonal tile jb factors it and
// publishes it; the tiles of column block jb solve against it and publish
// their rows; the tiles right of jb take the rank-4 update.  Rows and
// columns past nb of a partial last block are padded with the identity.
__global__ __launch_bounds__(512) void chol_panel_kernel(double* __restrict__ A, int64_t lda, int64_t ntot, int k0,
                                                         int nb, int* __restrict__ info) {
  __shared__ double Lb[16];                     // factored diagonal tile: L (lower) and 1/L[c][c] on the diagonal
  __shared__ __align__(16) double P[128][4];    // solved rows of the current column block
  __shared__ double S[128][CB + 1];             // [A11; A21] staged for column-wise (coalesced) loads and stores
  const int tid = threadIdx.x, ti = tid >> 4, tk = tid & 15;
  const int64_t r0 = k0 + nb + (int64_t)blockIdx.x * CB;  // first panel row of this workgroup
  const int nr = (int)max((int64_t)0, min((int64_t)CB, ntot - r0));
  const bool diag = ti < 16;
  const bool active = diag ? ti >= tk : true;
  // lanes along the rows of a column: stage row r (< 64: A11, else A21), column c
  for (int e = tid; e < 128 * CB; e += 512) {
    const int r = e & 127, c = e >> 7;
    const int64_t row = r < CB ? k0 + min(r, nb - 1) : (nr > 0 ? r0 + min(r - CB, nr - 1) : k0);
    S[r][c] = A[row + lda * (k0 + min(c, nb - 1))];
  }
  __syncthreads();
  double a[4][4];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int c = 4 * tk + y;
      const int r = diag ? 4 * ti + x : 4 * (ti - 16) + x;
      const bool ok = c < nb && (diag ? r < nb : r < nr);
      a[x][y] = ok ? S[diag ? r : CB + r][c] : ((diag && r == c) ? 1.0 : 0.0);
    }

  for (int jb = 0; jb < 16; jb++) {
    if (diag && ti == jb && tk == jb) {
      double L[4][4], dinv[4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        double d = a[c][c];
#pragma unroll
        for (int z = 0; z < c; z++) d -= L[c][z] * L[c][z];
        if (!(d > 0.0)) {
          if (blockIdx.x == 0 && 4 * jb + c < nb) atomicCAS(info, 0, k0 + 4 * jb + c + 1);
          d = 1.0;
        }
        L[c][c] = sqrt(d);
        dinv[c] = 1.0 / L[c][c];
#pragma unroll
        for (int x = c + 1; x < 4; x++) {
          double v = a[x][c];
#pragma unroll
          for (int z = 0; z < c; z++) v -= L[x][z] * L[c][z];
          L[x][c] = v * dinv[c];
        }
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
          a[x][y] = (y <= x) ? L[x][y] : 0.0;
          Lb[4 * x + y] = (y < x) ? L[x][y] : (y == x ? dinv[x] : 0.0);
        }
    }
    __syncthreads();
    if (tk == jb && ti > jb) {
      double L[16];
#pragma unroll
      for (int z = 0; z < 16; z++) L[z] = Lb[z];
      // X L^T = a:  X[x][y] = (a[x][y] - sum_{z<y} X[x][z] L[y][z]) / L[y][y]
#pragma unroll
      for (int x = 0; x < 4; x++) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
          double v = a[x][y];
#pragma unroll
          for (int z = 0; z < y; z++) v -= a[x][z] * L[4 * y + z];
          a[x][y] = v * L[5 * y];
        }
        *reinterpret_cast<double4*>(&P[4 * ti + x][0]) = make_double4(a[x][0], a[x][1], a[x][2], a[x][3]);
      }
    }
    __syncthreads();
    if (active && tk > jb) {
      double pr[4][4], pc[4][4];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const double4 u = *reinterpret_cast<const double4*>(&P[4 * ti + x][0]);
        const double4 v = *reinterpret_cast<const double4*>(&P[4 * tk + x][0]);
        pr[x][0] = u.x; pr[x][1] = u.y; pr[x][2] = u.z; pr[x][3] = u.w;
        pc[x][0] = v.x; pc[x][1] = v.y; pc[x][2] = v.z; pc[x][3] = v.w;
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
          double v = a[x][y];
#pragma unroll
          for (int z = 0; z < 4; z++) v = fma(-pr[x][z], pc[y][z], v);
          a[x][y] = v;
        }
    }
  }
  // store the panel rows L21, through S for column-wise stores
  if (!diag) {
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 4; y++) S[CB + 4 * (ti - 16) + x][4 * tk + y] = a[x][y];
  }
  __syncthreads();
  for (int e = tid; e < CB * CB; e += 512) {
    const int r = e & (CB - 1), c = e >> 6;
    if (c < nb && r < nr) A[(r0 + r) + lda * (k0 + c)] = S[CB + r][c];
  }
}

// A22 -= L21 L21^T on the lower part of the trailing matrix: rows
// k0+nb .. nrow_end-1, columns k0+nb .. ncol_end-1, one 64x64 tile per
// workgroup (grid: row tiles x column tiles; tiles above the diagonal exit).
__global__ __launch_bounds__(256) void chol_update_kernel(double* __restrict__ A, int64_t lda, int64_t nrow_end,
                                                          int k0, int nb, int64_t ncol_end) {
  __shared__ double Pr[CB][CB + 1];
  __shared__ double Pc[CB][CB + 1];
  const int64_t m0 = k0 + nb;
  const int tr = blockIdx.x, tc = blockIdx.y;
  if (tc > tr) return;
  const int64_t ntot = nrow_end;
  const int64_t r0 = m0 + (int64_t)tr * CB, c0 = m0 + (int64_t)tc * CB;
  const int nr = (int)min((int64_t)CB, ntot - r0), ncl = (int)min((int64_t)CB, ncol_end - c0);
  const int tid = threadIdx.x;
  {
    double vr[CB * CB / 256], vc[CB * CB / 256];
#pragma unroll
    for (int z = 0; z < CB * CB / 256; z++) {
      const int e = tid + 256 * z, r = e % CB, l = e / CB;
      const int ll = min(l, nb - 1);
      vr[z] = A[(r0 + min(r, nr - 1)) + lda * (k0 + ll)];
      vc[z] = A[(c0 + min(r, ncl - 1)) + lda * (k0 + ll)];
    }
#pragma unroll
    for (int z = 0; z < CB * CB / 256; z++) {
      const int e = tid + 256 * z, r = e % CB, l = e / CB;
      Pr[r][l] = (r < nr && l < nb) ? vr[z] : 0.0;
      Pc[r][l] = (r < ncl && l < nb) ? vc[z] : 0.0;
    }
  }
  __syncthreads();
  // the 64 x 64 product Pr Pc^T on the matrix cores: wave w owns rows
  // 32 (w / 2) .. + 31 and columns 32 (w % 2) .. + 31 as 2 x 2 tiles of
  // v_mfma_f64_16x16x4 (A: row lane % 16, k lane / 16; B: column lane % 16,
  // k lane / 16; D: row lane / 16 + 4 i, column lane % 16); the padding
  // columns l >= nb of Pr and Pc are zero
  const int wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const int mr0 = 32 * (wave >> 1), nc0 = 32 * (wave & 1);
  gemm_dbl4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) acc[x][y] = gemm_dbl4{0.0, 0.0, 0.0, 0.0};
  for (int k4 = 0; k4 < (nb + 3) / 4; k4++) {
    const int kk = 4 * k4 + lk;
    double a[2], b[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      a[x] = Pr[mr0 + 16 * x + lr][kk];
      b[x] = Pc[nc0 + 16 * x + lr][kk];
    }
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
  }
  // A22 -= acc through LDS, so that the read-modify-write runs along columns
  __syncthreads();
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int i = 0; i < 4; i++) Pr[mr0 + 16 * x + lk + 4 * i][nc0 + 16 * y + lr] = acc[x][y][i];
  __syncthreads();
  for (int e = tid; e < CB * CB; e += 256) {
    const int rr = e & (CB - 1), cc = e >> 6;
    const int64_t r = r0 + rr, c = c0 + cc;
    if (r < ntot && c < ncol_end && r >= c) A[r + lda * c] -= Pr[rr][cc];
  }
}

void chol_update(double* A, int64_t lda, int64_t nrow_end, int k0, int nb, int64_t ncol_end, hipStream_t st) {
  const int64_t m0 = k0 + nb;
  if (nrow_end <= m0 || ncol_end <= m0) return;
  const unsigned ntr = (unsigned)((nrow_end - m0 + CB - 1) / CB), ntc = (unsigned)((ncol_end - m0 + CB - 1) / CB);
  hipLaunchKernelGGL(chol_update_kernel, dim3(ntr, ntc), dim3(256), 0, st, A, lda, nrow_end, k0, nb, ncol_end);
}

// Cholesky of M bordered by the identity, A = [[M], [I]] (column-major, n
// columns, rows 0..2n-1 of lda): afterwards the lower triangle of the leading
// block holds L (M = L L^T; the diagonal blocks keep their last update) and
// the border rows hold L^-T: A[n + i + lda j] = (L^-1)[j][i] for j >= i.
// Identity row i is still untouched while the factorization is left of column
// i (its entries there are zero), so step k0 takes the border rows only up to
// n + k0 + nb.
void cholesky_inverse_rows(double* A, int64_t lda, int64_t n, int* info, hipStream_t st) {
  for (int k0 = 0; k0 < n; k0 += CB) {
    const int nb = (int)std::min<int64_t>(CB, n - k0);
    const int64_t rend = n + k0 + nb;
    const int64_t m = rend - k0 - nb;
    const unsigned npanel = (unsigned)std::max<int64_t>(1, (m + CB - 1) / CB);
    hipLaunchKernelGGL(chol_panel_kernel, dim3(npanel), dim3(512), 0, st, A, lda, rend, k0, nb, info);
    chol_update(A, lda, rend, k0, nb, n, st);
  }
}

// The metric computed on one side only: for functions u != v the block of
// the shell with the smaller L as bra (equal L: the smaller function index)
// was written at M[bra * ldm + ket]; copy it to M[ket * ldm + bra].  32 x 32
// tiles: the tile at (a0, b0) reads its mirror tile along rows into LDS.
__global__ __launch_bounds__(256) void symmetrize_metric_kernel(double* __restrict__ M, int64_t ldm, int64_t n,
                                                                const int* __restrict__ fL) {
  __shared__ double t[32][33];
  const int64_t a0 = (int64_t)blockIdx.y * 32, b0 = (int64_t)blockIdx.x * 32;
  const int tid = threadIdx.x;
  // mirror tile: element (a, b) of this tile is M[b * ldm + a]
  for (int e = tid; e < 1024; e += 256) {
    const int bi = e >> 5, ai = e & 31;
    const int64_t a = a0 + ai, b = b0 + bi;
    if (a < n && b < n) t[bi][ai] = M[b * ldm + a];
  }
  __syncthreads();
  for (int e = tid; e < 1024; e += 256) {
    const int ai = e >> 5, bi = e & 31;
    const int64_t a = a0 + ai, b = b0 + bi;
    if (a >= n || b >= n || a == b) continue;
    const int La = fL[a], Lb = fL[b];
    // (a, b) was computed when a is the bra: smaller L, or equal L and a < b
    const bool computed = La < Lb || (La == Lb && a < b);
    if (!computed) M[a * ldm + b] = t[bi][ai];
  }
}

void symmetrize_metric(double* M, int64_t ldm, int64_t n, const int* fL, hipStream_t st) {
  if (n <= 0) return;
  const dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32));
  hipLaunchKernelGGL(symmetrize_metric_kernel, grid, dim3(256), 0, st, M, ldm, n, fL);
}

// A[n + i + lda i] = 1 for i < n (the identity border of cholesky_inverse_rows).
__global__ void set_identity_rows_kernel(double* __restrict__ A, int64_t lda, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[n + i + lda * i] = 1.0;
}

void set_identity_rows(double* A, int64_t lda, int64_t n, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(set_identity_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, lda, n);
}

// BT[I + ndens j] = sum_{i <= j} DT[I + ldd i] LT[i + ldl j]: the rows of
// B^T = D^T L^-T with D^T (ndens x n, column-major, ldd) and the upper
// triangular L^-T (ldl).  64x64 output tiles (64 densities x 64 columns j),
// 4x4 per thread, K split into chunks of kc (grid z); each chunk writes its
// partial tile, a second kernel adds the chunks in a fixed order.
__global__ __launch_bounds__(256) void apply_linv_kernel(const double* __restrict__ DT, int64_t ldd,
                                                         const double* __restrict__ LT, int64_t ldl, int64_t n,
                                                         int ndens, int64_t kc, double* __restrict__ part) {
  __shared__ double sD[16][64 + 1];  // [i][I]
  __shared__ double sL[16][64 + 1];  // [i][j]
  const int64_t j0 = (int64_t)blockIdx.x * 64;
  const int I0 = blockIdx.y * 64;
  const int64_t kb = (int64_t)blockIdx.z * kc, ke = min(min(n, j0 + 64), kb + kc);
  if (kb >= ke) return;  // past the triangle of this column tile: the reduction skips the chunk
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;  // I = I0 + 4 ty + x, j = j0 + 4 tx + y
  double acc[4][4] = {};
  for (int64_t k0 = kb; k0 < ke; k0 += 16) {
    for (int e = tid; e < 16 * 64; e += 256) {
      const int kk = e >> 6, c = e & 63;
      const int64_t k = k0 + kk;
      sD[kk][c] = (k < ke && I0 + c < ndens) ? DT[(I0 + c) + ldd * k] : 0.0;
      sL[kk][c] = (k < ke && j0 + c < n && k <= j0 + c) ? LT[k + ldl * (j0 + c)] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        a[q] = sD[kk][4 * ty + q];
        b[q] = sL[kk][4 * tx + q];
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
    __syncthreads();
  }
  double* pz = part + (size_t)blockIdx.z * ndens * n;
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int I = I0 + 4 * ty + x;
      const int64_t j = j0 + 4 * tx + y;
      if (I < ndens && j < n) pz[I + (int64_t)ndens * j] = acc[x][y];
    }
}

__global__ void apply_linv_reduce_kernel(const double* __restrict__ part, int64_t n, int ndens, int64_t kc,
                                         double* __restrict__ BT) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ndens) return;
  const int64_t j = e / ndens, kend = min(n, (j / 64) * 64 + 64);
  double s = 0.0;
  for (int64_t z = 0; z * kc < kend; z++) s += part[(size_t)z * ndens * n + e];
  BT[e] = s;
}

// K chunk of apply_linv: 64 at naux <= 2048 (a few hundred workgroups of
// four K steps for the small triangle), at most 32 chunks above
int64_t apply_linv_chunk(int64_t n) { return std::max<int64_t>(64, ((n + 31) / 32 + 15) / 16 * 16); }
int64_t apply_linv_part_size(int64_t n, int ndens) {
  const int64_t kc = apply_linv_chunk(n);
  return ((n + kc - 1) / kc) * ndens * n;
}

void apply_linv(const double* DT, int64_t ldd, const double* LT, int64_t ldl, int64_t n, int ndens, double* part,
                double* BT, hipStream_t st) {
  if (n <= 0 || ndens <= 0) return;
  const int64_t kc = apply_linv_chunk(n), nkc = (n + kc - 1) / kc;
  hipLaunchKernelGGL(apply_linv_kernel, dim3((unsigned)((n + 63) / 64), (ndens + 63) / 64, (unsigned)nkc), dim3(256),
                     0, st, DT, ldd, LT, ldl, n, ndens, kc, part);
  const int64_t tot = n * ndens;
  hipLaunchKernelGGL(apply_linv_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part, n, ndens,
                     kc, BT);
}

// G[I + ndens K] = -sum_j BT[I + ndens j] BT[K + ndens j] (the packed Coulomb
// tensor with a minus sign, as the trailing block of the bordered
// factorization): partial sums over chunks of gj columns j (about 512
// workgroups in all), then a reduction over the chunks in a fixed order.
int64_t gram_jchunk(int64_t n, int ndens) {
  const int64_t t = (ndens + 63) / 64, want = std::max<int64_t>(1, 512 / (t * t));
  return std::max<int64_t>(16, ((n + want - 1) / want + 15) / 16 * 16);
}

__global__ __launch_bounds__(256) void gram_part_kernel(const double* __restrict__ BT, int64_t n, int ndens,
                                                        int64_t gj, double* __restrict__ part) {
  __shared__ double sA[16][64 + 1], sB[16][64 + 1];
  const int I0 = blockIdx.x * 64, K0 = blockIdx.y * 64;
  const int64_t jc = (int64_t)blockIdx.z * gj, jend = min(n, jc + gj);
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double acc[4][4] = {};
  for (int64_t k0 = jc; k0 < jend; k0 += 16) {
    for (int e = tid; e < 16 * 64; e += 256) {
      const int kk = e >> 6, c = e & 63;
      const int64_t k = k0 + kk;
      sA[kk][c] = (k < jend && I0 + c < ndens) ? BT[(I0 + c) + (int64_t)ndens * k] : 0.0;
      sB[kk][c] = (k < jend && K0 + c < ndens) ? BT[(K0 + c) + (int64_t)ndens * k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      double a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        a[q] = sA[kk][4 * ty + q];
        b[q] = sB[kk][4 * tx + q];
      }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int I = I0 + 4 * ty + x, K = K0 + 4 * tx + y;
      if (I < ndens && K < ndens) part[((size_t)blockIdx.z * ndens + K) * ndens + I] = acc[x][y];
    }
}

__global__ void gram_reduce_kernel(const double* __restrict__ part, int nchunk, int ndens, double* __restrict__ G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ndens * ndens) return;
  double s = 0.0;
  for (int z = 0; z < nchunk; z++) s += part[(size_t)z * ndens * ndens + e];
  G[e] = -s;
}

int gram_chunks(int64_t n, int ndens) {
  const int64_t gj = gram_jchunk(n, ndens);
  return (int)std::max<int64_t>(1, (n + gj - 1) / gj);
}

void gram_neg(const double* BT, int64_t n, int ndens, double* part, double* G, hipStream_t st) {
  if (ndens <= 0) return;
  const int nchunk = gram_chunks(n, ndens);
  const unsigned t = (unsigned)((ndens + 63) / 64);
  hipLaunchKernelGGL(gram_part_kernel, dim3(t, t, nchunk), dim3(256), 0, st, BT, n, ndens, gram_jchunk(n, ndens),
                     part);
  hipLaunchKernelGGL(gram_reduce_kernel, dim3((ndens * ndens + 255) / 256), dim3(256), 0, st, part, nchunk, ndens, G);
}

// B[P + n I] = BT[I + ndens P]
__global__ void bt_to_b_kernel(const double* __restrict__ BT, int64_t n, int ndens, double* __restrict__ B) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ndens) return;
  const int64_t P = e % n, I = e / n;
  B[P + n * I] = BT[I + ndens * P];
}

void bt_to_b(const double* BT, int64_t n, int ndens, double* B, hipStream_t st) {
  const int64_t tot = n * ndens;
  if (tot > 0)
    hipLaunchKernelGGL(bt_to_b_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, BT, n, ndens, B);
}

// ---------------------------------------------------------------------------
// J assembly: the trailing block of the bordered factorization holds
// G(I1, I2) = -B(:,I1) . B(:,I2) for I2 <= I1; J is written to the 8 index
// permutations of the column-major nstates^4 tensor.
__device__ inline void unpack_pair(int q, int& i, int& j) {
  int a = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
  while ((a + 1) * (a + 2) / 2 <= q) a++;
  while (a * (a + 1) / 2 > q) a--;
  i = a;
  j = q - a * (a + 1) / 2;
}

__global__ void assemble_j_kernel(const double* __restrict__ A, int64_t lda, int64_t n, int npairs, int nstates,
                                  double* __restrict__ J) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  int I1, I2;
  unpack_pair(q, I1, I2);
  const double v = -A[(n + I1) + lda * (n + I2)];  // G(I1, I2) at row n + I1, column n + I2
  int i, j, k, l;
  unpack_pair(I1, i, j);
  unpack_pair(I2, k, l);
  const int64_t ns = nstates;
  auto at = [&](int a, int b, int c, int d) -> double& { return J[a + ns * (b + ns * (c + ns * d))]; };
  at(i, j, k, l) = v;
  at(j, i, k, l) = v;
  at(i, j, l, k) = v;
  at(j, i, l, k) = v;
  at(k, l, i, j) = v;
  at(k, l, j, i) = v;
  at(l, k, i, j) = v;
  at(l, k, j, i) = v;
}

void assemble_j(const double* A, int64_t lda, int64_t n, int ndens, int nstates, double* J, hipStream_t st) {
  const int npairs = ndens * (ndens + 1) / 2;
  if (npairs > 0)
    hipLaunchKernelGGL(assemble_j_kernel, dim3((npairs + 127) / 128), dim3(128), 0, st, A, lda, n, npairs, nstates, J);
}

// ---------------------------------------------------------------------------
// Q_C = sqrt(||(C|C)||_F) from the diagonal blocks of the metric (row-major =
// column-major, symmetric).
// compact: the diagonal blocks sit at M[(f0 + i) n + j] (rows of the shell's
// functions, columns m) instead of on the diagonal of the full metric
__global__ void schwarz_aux_kernel(const double* __restrict__ M, int64_t n, const int* __restrict__ L,
                                   const int* __restrict__ f0, int nsh, int compact, double* __restrict__ q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nsh) return;
  const int nc = 2 * L[c] + 1;
  const int64_t o = f0[c], oc = compact ? 0 : o;
  double s = 0.0;
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < nc; j++) {
      const double v = M[(o + i) * n + oc + j];
      s += v * v;
    }
  q[c] = sqrt(sqrt(s));
}

void schwarz_aux_from_metric(const double* M, int64_t n, const int* L, const int* f0, int nsh, double* q,
                             hipStream_t st, bool compact) {
  if (nsh > 0)
    hipLaunchKernelGGL(schwarz_aux_kernel, dim3((nsh + 127) / 128), dim3(128), 0, st, M, n, L, f0, nsh,
                       compact ? 1 : 0, q);
}

// ---------------------------------------------------------------------------
__global__ void boys_probe_kernel(const int* __restrict__ nv, const double* __restrict__ T, int64_t npts,
                                  const double* __restrict__ tab, double* __restrict__ F) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npts) F[i] = boys_f(nv[i], T[i], tab);
}

void boys_probe(const int* n, const double* T, int64_t npts, const double* tab, double* F, hipStream_t st) {
  if (npts > 0)
    hipLaunchKernelGGL(boys_probe_kernel, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, n, T, npts, tab, F);
}

}  // namespace dfhip
