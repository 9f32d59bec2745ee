// Densities on grid points from Cartesian AO density matrices (grid.hip);
// the GPU form of grid/grid.f90:112-286.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfhip {

// Shell table on the device for the AO values: centre xyz[3 s], angular
// momentum L[s] (0..4), primitives pptr[s] .. pptr[s] + nprim[s] - 1 of
// exps/coefs (coefficients normalized as in norm_cgto_coeffs), first
// Cartesian function c0[s].
struct GridShells {
  const double* xyz;
  const double* exps;
  const double* coefs;
  const int* L;
  const int* nprim;
  const int* pptr;
  const int* c0;
  int n;
  int ncart;
};

// Largest points per block the density kernel holds in registers at once
// (larger blocks are walked in groups of this many points).
constexpr int GRID_GROUP = 64;

// AO values of the points p0 .. p0 + np - 1 (pts: x, y, z per point) in
// blocks of bs points: X[(b * ncart + c) * bs + i] for point p0 + b * bs + i
// (p0 is a multiple of bs; rows past the last point are not written).
void grid_ao_values(const GridShells& sh, const double* pts, int64_t p0, int64_t np, int bs, double* X,
                    hipStream_t st);
// means[b * ncart + c] = sum_i |X[(b ncart + c) bs + i]| / (points of block b),
// summed over i in order.
void grid_block_means(const double* X, int nblk, int ncart, int bs, int64_t np, double* means, hipStream_t st);
// rho[(p0 + b bs + i) * ndens + j] = sum over nu <= mu of f D_j(mu, nu)
// chi_mu chi_nu (f = 2 off the diagonal), the pair kept when
// |((f D_j(mu, nu)) mean_nu) mean_mu| >= thresh with the means of block b.
// Dd[(j * ncart + mu) * ncart + nu] = D_j(mu, nu).
void grid_block_densities(const double* X, const double* means, const double* Dd, int ndens, int ncart, int nblk,
                          int bs, int64_t p0, int64_t np, double thresh, double* rho, hipStream_t st);
// Dd[(j * n + mu) * n + nu] = D[j + ndens (mu + n nu)] (the Fortran (ndens, n, n) array)
void grid_transpose_densities(const double* D, int ndens, int n, double* Dd, hipStream_t st);

}  // namespace dfhip
