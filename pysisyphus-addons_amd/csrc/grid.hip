// Densities of Cartesian AO density matrices on grid points
// (grid/grid.f90:112-286), in three kernels per batch of point blocks:
//   * grid_ao_kernel: the Cartesian AO values chi_c(r) of every point
//     (grid.f90:17-110, cart_gto3d_rel), one thread per point and shell
//     slice, stored block by block with the points of a block contiguous;
//   * grid_mean_kernel: the mean |chi_c| over the points of each block
//     (grid.f90:222-229), summed in point order;
//   * grid_density_kernel: rho_j(i) = sum_nu chi_nu(i) S_j(i, nu) with
//     S_j(i, nu) = sum_{mu >= nu} M_j(mu, nu) chi_mu(i), M_j the density
//     matrix with the factor 2 off the diagonal and the pairs the block's
//     means screen out set to zero (grid.f90:231-271).  S is a product of the
//     block's AO values (points x functions) with M on the FP64 matrix cores
//     (v_mfma_f64_16x16x4f64): one workgroup per (block, density), each wave
//     walks 16-column tiles of nu and the mu >= nu rows below them, and a
//     4-row slice of M that the screen clears entirely is skipped.
#include <hip/hip_runtime.h>

#include "grid.h"

namespace dfhip {

namespace {

typedef double gdbl4 __attribute__((ext_vector_type(4)));

constexpr double GRID_CUTOFF = -36.0;  // grid.f90:9: primitives with -a r^2 below it are skipped
constexpr int SH_SLICE = 8;            // shells per thread of the AO kernel

// (2l - 1)!! for l = 0..4
__device__ inline double dfact_odd(int l) { return l <= 1 ? 1.0 : l == 2 ? 3.0 : l == 3 ? 15.0 : 105.0; }

__global__ __launch_bounds__(256) void grid_ao_kernel(GridShells sh, const double* __restrict__ pts, int64_t p0,
                                                      int64_t np, int bs, double* __restrict__ X) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= np) return;
  const int64_t gp = p0 + p;
  const int64_t blk = p / bs;
  const int i = (int)(p - blk * bs);
  const double rx = pts[3 * gp], ry = pts[3 * gp + 1], rz = pts[3 * gp + 2];
  const int s0 = blockIdx.y * SH_SLICE, s1 = min(sh.n, s0 + SH_SLICE);
  double* Xb = X + (size_t)blk * sh.ncart * bs + i;
  for (int s = s0; s < s1; s++) {
    const double dx = rx - sh.xyz[3 * s], dy = ry - sh.xyz[3 * s + 1], dz = rz - sh.xyz[3 * s + 2];
    const double dist = dx * dx + dy * dy + dz * dz;
    const int pp = sh.pptr[s], npr = sh.nprim[s], L = sh.L[s];
    double rad = 0.0;
    for (int k = 0; k < npr; k++) {
      const double arg = -sh.exps[pp + k] * dist;
      if (arg < GRID_CUTOFF) continue;
      rad += sh.coefs[pp + k] * exp(arg);
    }
    double px[5], py[5], pz[5];
    px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
    for (int e = 1; e < 5; e++) {
      px[e] = px[e - 1] * dx;
      py[e] = py[e - 1] * dy;
      pz[e] = pz[e - 1] * dz;
    }
    // Cartesian order of the shell: lx = L..0, then ly = L - lx..0 (xx, xy, xz, yy, yz, zz for d);
    // each component normalized by 1/sqrt((2lx-1)!! (2ly-1)!! (2lz-1)!!) (grid.f90:10-13)
    double* out = Xb + (size_t)sh.c0[s] * bs;
    int c = 0;
    for (int lx = L; lx >= 0; lx--)
      for (int ly = L - lx; ly >= 0; ly--, c++) {
        const int lz = L - lx - ly;
        const double nrm = 1.0 / sqrt(dfact_odd(lx) * dfact_odd(ly) * dfact_odd(lz));
        out[(size_t)c * bs] = nrm * px[lx] * py[ly] * pz[lz] * rad;
      }
  }
}

__global__ void grid_mean_kernel(const double* __restrict__ X, int nblk, int ncart, int bs, int64_t np,
                                 double* __restrict__ means) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nblk * ncart) return;
  const int64_t b = e / ncart;
  const int cur = (int)min((int64_t)bs, np - b * bs);
  const double* x = X + (size_t)e * bs;
  double s = 0.0;
  for (int i = 0; i < cur; i++) s += fabs(x[i]);
  means[e] = s / (double)cur;
}

__global__ __launch_bounds__(256) void grid_density_kernel(const double* __restrict__ X,
                                                           const double* __restrict__ means,
                                                           const double* __restrict__ Dd, int ndens, int ncart,
                                                           int bs, int64_t p0, int64_t np, double thresh,
                                                           double* __restrict__ rho) {
  constexpr int NIT = GRID_GROUP / 16;
  __shared__ double red[4][GRID_GROUP];
  const int b = blockIdx.x, j = blockIdx.y;
  const int cur = (int)min((int64_t)bs, np - (int64_t)b * bs);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  const double* Xb = X + (size_t)b * ncart * bs;
  const double* mb = means + (size_t)b * ncart;
  const double* Dj = Dd + (size_t)j * ncart * ncart;
  const int ntile = (ncart + 15) / 16;
  for (int g0 = 0; g0 < cur; g0 += GRID_GROUP) {
    const int gn = min(GRID_GROUP, cur - g0);
    double racc[NIT][4];
#pragma unroll
    for (int t = 0; t < NIT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) racc[t][r] = 0.0;
    for (int tile = wave; tile < ntile; tile += 4) {
      const int n0 = 16 * tile, nu = n0 + lr;
      const bool nuok = nu < ncart;
      const double mnu = nuok ? mb[nu] : 0.0;
      gdbl4 acc[NIT];
#pragma unroll
      for (int t = 0; t < NIT; t++) acc[t] = gdbl4{0.0, 0.0, 0.0, 0.0};
      for (int k0 = n0 & ~3; k0 < ncart; k0 += 4) {
        // B operand M(mu, nu), lane (k = lk, n = lr); the screen of grid.f90:250-256:
        // |((f D) mean_nu) mean_mu| >= thresh
        const int mu = k0 + lk;
        double bv = 0.0;
        if (mu < ncart && nuok && mu >= nu) {
          const double dd = (mu == nu ? 1.0 : 2.0) * Dj[(size_t)mu * ncart + nu];
          if (fabs(dd * mnu * mb[mu]) >= thresh) bv = dd;
        }
        if (__ballot(bv != 0.0) == 0) continue;  // the whole 4 x 16 slice is screened out
        const double* xm = Xb + (size_t)min(mu, ncart - 1) * bs + g0;
#pragma unroll
        for (int t = 0; t < NIT; t++) {
          if (16 * t >= gn) break;  // uniform
          const int row = 16 * t + lr;
          const double a = (row < gn && mu < ncart) ? xm[row] : 0.0;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
        }
      }
      // register r of lane l holds S(row 16 t + l/16 + 4 r, column n0 + l%16)
      const double* xn = Xb + (size_t)min(nu, ncart - 1) * bs + g0;
#pragma unroll
      for (int t = 0; t < NIT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = 16 * t + lk + 4 * r;
          if (row < gn && nuok) racc[t][r] = fma(acc[t][r], xn[row], racc[t][r]);
        }
    }
    // sum over the 16 columns of each lane group, then over the waves in order
#pragma unroll
    for (int t = 0; t < NIT; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        double v = racc[t][r];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        if (lr == 0) red[wave][16 * t + lk + 4 * r] = v;
      }
    __syncthreads();
    for (int i = tid; i < gn; i += 256) {
      const double s = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
      rho[(size_t)(p0 + (int64_t)b * bs + g0 + i) * ndens + j] = s;
    }
    __syncthreads();
  }
}

__global__ void grid_transpose_kernel(const double* __restrict__ D, int ndens, int n, double* __restrict__ Dd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n * n;
  if (e >= nn * ndens) return;
  const int j = (int)(e / nn);
  const int64_t r = e - (int64_t)j * nn;
  const int mu = (int)(r / n), nu = (int)(r - (int64_t)mu * n);
  Dd[e] = D[j + (int64_t)ndens * (mu + (int64_t)n * nu)];
}

}  // namespace

void grid_ao_values(const GridShells& sh, const double* pts, int64_t p0, int64_t np, int bs, double* X,
                    hipStream_t st) {
  if (np <= 0) return;
  const dim3 grid((unsigned)((np + 255) / 256), (unsigned)((sh.n + SH_SLICE - 1) / SH_SLICE));
  hipLaunchKernelGGL(grid_ao_kernel, grid, dim3(256), 0, st, sh, pts, p0, np, bs, X);
}

void grid_block_means(const double* X, int nblk, int ncart, int bs, int64_t np, double* means, hipStream_t st) {
  const int64_t n = (int64_t)nblk * ncart;
  if (n <= 0) return;
  hipLaunchKernelGGL(grid_mean_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, X, nblk, ncart, bs, np,
                     means);
}

void grid_block_densities(const double* X, const double* means, const double* Dd, int ndens, int ncart, int nblk,
                          int bs, int64_t p0, int64_t np, double thresh, double* rho, hipStream_t st) {
  if (nblk <= 0 || ndens <= 0) return;
  hipLaunchKernelGGL(grid_density_kernel, dim3(nblk, ndens), dim3(256), 0, st, X, means, Dd, ndens, ncart, bs, p0, np,
                     thresh, rho);
}

void grid_transpose_densities(const double* D, int ndens, int n, double* Dd, hipStream_t st) {
  const int64_t m = (int64_t)ndens * n * n;
  if (m <= 0) return;
  hipLaunchKernelGGL(grid_transpose_kernel, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, D, ndens, n, Dd);
}

}  // namespace dfhip
