// Shared pieces of the generated register-resident 3c2e kernels.
#pragma once

#include "int3c2e_impl.h"

namespace dfhip {

constexpr int GEN3_LA_MAX = 2;  // must match gen3c.py
constexpr int GEN3_LC_MAX = 3;

// F_0..F_L(T): 7-term Taylor expansion of F_L around the nearest table point
// and downward recursion F_n = (2T F_{n+1} + e^{-T}) / (2n+1) for T < 40;
// sqrt(pi/T)/2 and upward recursion otherwise.
template <int L>
__device__ __forceinline__ void boys_array(double T, const double* __restrict__ tab, double* F) {
  if (T < BOYS_TMAX) {
    F[L] = boys_f(L, T, tab);
    if (L > 0) {
      const double eT = exp(-T), T2 = 2.0 * T;
#pragma unroll
      for (int n = L - 1; n >= 0; n--) F[n] = fma(T2, F[n + 1], eT) * (1.0 / (2 * n + 1));
    }
  } else {
    const double eT = exp(-T), i2T = 0.5 / T;
    F[0] = 0.5 * sqrt(PI / T);
#pragma unroll
    for (int n = 0; n < L; n++) F[n + 1] = ((2 * n + 1) * F[n] - eT) * i2T;
  }
}

// Index triples of the E-products of a bra record, in the record order
// (Cartesian a, Cartesian b, t, u, v; gen3c.py record_layout): entry k is
// c_i c_j E_x[ix[k]] E_y[iy[k]] E_z[iz[k]], each E_d laid out as
// [(a_d (LB+1) + b_d) (LA+LB+1) + t].
template <int LA, int LB>
struct BraRecMap {
  static constexpr int NT = LA + LB + 1;
  static constexpr int count() {
    int n = 0;
    for (int ax = LA; ax >= 0; ax--)
      for (int ay = LA - ax; ay >= 0; ay--)
        for (int bx = LB; bx >= 0; bx--)
          for (int by = LB - bx; by >= 0; by--)
            n += (ax + bx + 1) * (ay + by + 1) * (LA - ax - ay + LB - bx - by + 1);
    return n;
  }
  static constexpr int N = count();
  unsigned short ix[N], iy[N], iz[N];
  constexpr BraRecMap() : ix(), iy(), iz() {
    int k = 0;
    for (int ax = LA; ax >= 0; ax--)
      for (int ay = LA - ax; ay >= 0; ay--) {
        const int az = LA - ax - ay;
        for (int bx = LB; bx >= 0; bx--)
          for (int by = LB - bx; by >= 0; by--) {
            const int bz = LB - bx - by;
            for (int t = 0; t <= ax + bx; t++)
              for (int u = 0; u <= ay + by; u++)
                for (int v = 0; v <= az + bz; v++) {
                  ix[k] = (unsigned short)((ax * (LB + 1) + bx) * NT + t);
                  iy[k] = (unsigned short)((ay * (LB + 1) + by) * NT + u);
                  iz[k] = (unsigned short)((az * (LB + 1) + bz) * NT + v);
                  k++;
                }
          }
      }
  }
};

template <int LA, int LB>
__constant__ constexpr BraRecMap<LA, LB> kBraRecMap{};

// Per-primitive-pair bra records: [p, Px, Py, Pz, c_i c_j E^{ax bx}_t E^{ay by}_u E^{az bz}_v ...]
// in the order of BraRecMap.  One workgroup per pair: up to BRA_NB primitive
// pairs at a time put their E tables in LDS, then the lanes write the records
// entry by entry (consecutive lanes, consecutive doubles).
constexpr int BRA_NB = 16;
constexpr int BRA_THREADS = 128;

template <int LA, int LB>
__global__ __launch_bounds__(BRA_THREADS) void bra_records_kernel(DevShells bra, const int* __restrict__ s1v,
                                                                  const int* __restrict__ s2v,
                                                                  const int64_t* __restrict__ recoff, int rs,
                                                                  double* __restrict__ recs) {
  constexpr int ESZ = (LA + 1) * (LB + 1) * (LA + LB + 1);
  constexpr int ES = 3 * ESZ + 1;  // E_x, E_y, E_z, then c_i c_j
  constexpr int N = BraRecMap<LA, LB>::N;
  __shared__ double Es[BRA_NB * ES];
  const int pidx = blockIdx.x;
  const int s1 = s1v[pidx], s2 = s2v[pidx];
  const int npa = bra.nprim[s1], pa0 = bra.pptr[s1];
  const int npb = s2 >= 0 ? bra.nprim[s2] : 1, pb0 = s2 >= 0 ? bra.pptr[s2] : 0;
  const int npp = npa * npb;
  double* __restrict__ r0 = recs + recoff[pidx];
  for (int ij0 = 0; ij0 < npp; ij0 += BRA_NB) {
    const int nb = min(BRA_NB, npp - ij0);
    if ((int)threadIdx.x < nb) {
      const int ij = ij0 + threadIdx.x;
      const int i = ij / npb, j = ij - (ij / npb) * npb;
      double A[3], B[3];
      for (int d = 0; d < 3; d++) {
        A[d] = bra.xyz[3 * s1 + d];
        B[d] = s2 >= 0 ? bra.xyz[3 * s2 + d] : A[d];
      }
      const double a = bra.exps[pa0 + i];
      const double b = s2 >= 0 ? bra.exps[pb0 + j] : 0.0;
      const double cab = bra.coefs[pa0 + i] * (s2 >= 0 ? bra.coefs[pb0 + j] : 1.0);
      const double p = a + b;
      double E[3][ESZ];
      for (int d = 0; d < 3; d++) herm_E_dir<LA, LB>(a, b, A[d], B[d], E[d]);
      double* El = Es + threadIdx.x * ES;
      for (int d = 0; d < 3; d++)
        for (int e = 0; e < ESZ; e++) El[d * ESZ + e] = E[d][e];
      El[3 * ESZ] = cab;
      double* r = r0 + (int64_t)ij * rs;
      r[0] = p;
      for (int d = 0; d < 3; d++) r[1 + d] = (a * A[d] + b * B[d]) / p;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nb * N; e += BRA_THREADS) {
      const int l = e / N, k = e - l * N;
      const double* El = Es + l * ES;
      r0[(int64_t)(ij0 + l) * rs + 4 + k] = El[3 * ESZ] * El[kBraRecMap<LA, LB>.ix[k]] *
                                            El[ESZ + kBraRecMap<LA, LB>.iy[k]] *
                                            El[2 * ESZ + kBraRecMap<LA, LB>.iz[k]];
    }
    __syncthreads();
  }
}

using Gen3Kernel = void (*)(const double*, const float*, const int64_t*, const int*, DevShells, PairArgs, AuxArgs,
                            const double*, double*, int64_t);

struct Gen3Launch {
  Gen3Kernel kernel;    // FP64
  Gen3Kernel kernel32;  // FP32 integral arithmetic, FP64 Boys values and primitive-pair sums
  int groups;           // aux-component groups (grid z) of kernel
  int groups32;         // and of kernel32
  int rs;
};


using Gen3Table = Gen3Launch[GEN3_LA_MAX + 1][GEN3_LA_MAX + 1][GEN3_LC_MAX + 1];

}  // namespace dfhip
