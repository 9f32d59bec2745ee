// Shared definitions of the density-fitting HIP library (host and device).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dfhip {

// Largest shell angular momenta the kernels are instantiated for.
constexpr int LMAX_ORB = 4;   // orbital basis up to g
constexpr int LMAX_AUX = 5;   // auxiliary basis up to h
constexpr int LMAX_SPH = 6;   // largest L with solid-harmonic terms in the tables
constexpr int LMAX_R = 4 * LMAX_ORB;  // largest Hermite order (gg|gg)

// Boys table: F_n(k h) for k = 0..BOYS_NGRID-1, n = 0..BOYS_NORD-1.
constexpr double BOYS_H = 0.05;
constexpr double BOYS_TMAX = 40.0;
constexpr int BOYS_NGRID = 808;
constexpr int BOYS_NTAYLOR = 7;
constexpr int BOYS_NORD = LMAX_R + BOYS_NTAYLOR + 1;

__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
__host__ __device__ constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }
__host__ __device__ constexpr int nsph(int L) { return 2 * L + 1; }

// Hermite index of (t,u,v): all triples of order s = t+u+v come after those of
// lower order; inside an order, t descends, then u descends.
__host__ __device__ inline int herm_index(int t, int u, int v) {
  int s = t + u + v;
  return s * (s + 1) * (s + 2) / 6 + (s - t) * (s - t + 1) / 2 + v;
}

// Solid-harmonic term: coefficient of x^lx y^ly z^lz in the normalized real
// solid harmonic (L, m), scaled by 1/sqrt((2L-1)!!).
struct SphTerm {
  double c;
  int lx, ly, lz;
};

constexpr int SPH_MAX_TERMS = 512;

struct SphTable {
  int start[LMAX_SPH + 1][2 * LMAX_SPH + 2];  // terms of (L, m) are [start[L][m], start[L][m+1])
  SphTerm terms[SPH_MAX_TERMS];
};

// Flattened basis in device memory.  Shell s has nprim[s] primitives at
// prim offset pptr[s], centre xyz[3s..3s+2], spherical function offset f0[s].
struct DevShells {
  int nshells;
  const int* L;
  const int* nprim;
  const int* pptr;
  const int* f0;
  const double* xyz;
  const double* exps;
  const double* coefs;
};

}  // namespace dfhip
