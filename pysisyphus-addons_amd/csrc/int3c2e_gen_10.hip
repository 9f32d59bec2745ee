// Generated by gen3c.py -- 3c2e classes (10|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (10|0): L = 1, record 10 doubles, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_100(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
    const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa, const double* __restrict__ boys,
    double* __restrict__ out, int64_t ldo) {
  const int pidx = blockIdx.x;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int c = xa.lo + (int)blockIdx.y * 64 + (int)threadIdx.x;
  if (xa.lo + (int)blockIdx.y * 64 >= cend) return;
  const bool active = c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 0, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (blockIdx.z) {
  case 0: {
    double acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      double F[2];
      boys_array<1>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      double W[4];
#pragma unroll
      for (int q = 0; q < 4; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      {
      const double y0_0_0 = py0 * z0_0;
      { const double r = px0 * y0_0_0;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y0_0_1 = py0 * z0_1;
      { const double r = px1 * y0_0_1;
        W[1] = fma(1.0, r, W[1]);
      }
      }
      {
      const double y1_0_1 = py1 * z0_1;
      { const double r = px0 * y1_0_1;
        W[2] = fma(1.0, r, W[2]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      {
      const double y0_1_1 = py0 * z1_1;
      { const double r = px0 * y0_1_1;
        W[3] = fma(1.0, r, W[3]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[7];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[8];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[9];
        acc[2] = fma(e, W[3], acc[2]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[1];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[2];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[0];
    }
    break;
  }
  default: break;
  }
}

// class (10|1): L = 2, record 10 doubles, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_101(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
    const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa, const double* __restrict__ boys,
    double* __restrict__ out, int64_t ldo) {
  const int pidx = blockIdx.x;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int c = xa.lo + (int)blockIdx.y * 64 + (int)threadIdx.x;
  if (xa.lo + (int)blockIdx.y * 64 >= cend) return;
  const bool active = c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 1, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (blockIdx.z) {
  case 0: {
    double acc[9];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
      double F[3];
      boys_array<2>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      double W[12];
#pragma unroll
      for (int q = 0; q < 12; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      {
      const double y0_0_1 = py0 * z0_1;
      { const double r = px1 * y0_0_1;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y0_0_2 = py0 * z0_2;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[5] = fma(1.0, r, W[5]);
      }
      }
      {
      const double y1_0_1 = py1 * z0_1;
      { const double r = px0 * y1_0_1;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[3] = fma(1.0, r, W[3]);
        W[8] = fma(1.0, r, W[8]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[6] = fma(1.0, r, W[6]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      {
      const double y0_1_1 = py0 * z1_1;
      { const double r = px0 * y0_1_1;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[4] = fma(1.0, r, W[4]);
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[9] = fma(1.0, r, W[9]);
        W[7] = fma(1.0, r, W[7]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[3], acc[0]);
        acc[1] = fma(e, W[4], acc[1]);
        acc[2] = fma(e, W[5], acc[2]);
      }
      { const double e = rec[6];
        acc[3] = fma(e, W[0], acc[3]);
        acc[4] = fma(e, W[1], acc[4]);
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[7];
        acc[3] = fma(e, W[6], acc[3]);
        acc[4] = fma(e, W[7], acc[4]);
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[8];
        acc[6] = fma(e, W[0], acc[6]);
        acc[7] = fma(e, W[1], acc[7]);
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[9];
        acc[6] = fma(e, W[9], acc[6]);
        acc[7] = fma(e, W[10], acc[7]);
        acc[8] = fma(e, W[11], acc[8]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[4];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[5];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[2];
    }
    break;
  }
  default: break;
  }
}

// class (10|2): L = 3, record 10 doubles, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_102(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
    const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa, const double* __restrict__ boys,
    double* __restrict__ out, int64_t ldo) {
  const int pidx = blockIdx.x;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int c = xa.lo + (int)blockIdx.y * 64 + (int)threadIdx.x;
  if (xa.lo + (int)blockIdx.y * 64 >= cend) return;
  const bool active = c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 2, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (blockIdx.z) {
  case 0: {
    double acc[15];
#pragma unroll
    for (int q = 0; q < 15; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[4];
      boys_array<3>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      {
      const double y0_0_2 = py0 * z0_2;
      const double y0_0_1 = py0 * z0_1;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[2] = fma(-0.2886751345948129, r, W[2]);
        W[4] = fma(0.5, r, W[4]);
      }
      const double y0_0_3 = py0 * z0_3;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[7] = fma(-0.2886751345948129, r, W[7]);
        W[9] = fma(0.5, r, W[9]);
      }
      }
      {
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_0_3 = py1 * z0_3;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[5] = fma(1.0, r, W[5]);
        W[12] = fma(-0.2886751345948129, r, W[12]);
        W[14] = fma(0.5, r, W[14]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[2] = fma(-0.2886751345948129, r, W[2]);
        W[4] = fma(-0.5, r, W[4]);
      }
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[10] = fma(1.0, r, W[10]);
        W[7] = fma(-0.2886751345948129, r, W[7]);
        W[9] = fma(-0.5, r, W[9]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[12] = fma(-0.2886751345948129, r, W[12]);
        W[14] = fma(-0.5, r, W[14]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      {
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y0_1_3 = py0 * z1_3;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[17] = fma(-0.2886751345948129, r, W[17]);
        W[8] = fma(1.0, r, W[8]);
        W[19] = fma(0.5, r, W[19]);
      }
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[15] = fma(1.0, r, W[15]);
        W[6] = fma(1.0, r, W[6]);
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[11] = fma(1.0, r, W[11]);
        W[17] = fma(-0.2886751345948129, r, W[17]);
        W[19] = fma(-0.5, r, W[19]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[2] = fma(0.5773502691896258, r, W[2]);
      }
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[7] = fma(0.5773502691896258, r, W[7]);
        W[18] = fma(1.0, r, W[18]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[16] = fma(1.0, r, W[16]);
        W[12] = fma(0.5773502691896258, r, W[12]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[17] = fma(0.5773502691896258, r, W[17]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
        acc[3] = fma(e, W[3], acc[3]);
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[5], acc[0]);
        acc[1] = fma(e, W[6], acc[1]);
        acc[2] = fma(e, W[7], acc[2]);
        acc[3] = fma(e, W[8], acc[3]);
        acc[4] = fma(e, W[9], acc[4]);
      }
      { const double e = rec[6];
        acc[5] = fma(e, W[0], acc[5]);
        acc[6] = fma(e, W[1], acc[6]);
        acc[7] = fma(e, W[2], acc[7]);
        acc[8] = fma(e, W[3], acc[8]);
        acc[9] = fma(e, W[4], acc[9]);
      }
      { const double e = rec[7];
        acc[5] = fma(e, W[10], acc[5]);
        acc[6] = fma(e, W[11], acc[6]);
        acc[7] = fma(e, W[12], acc[7]);
        acc[8] = fma(e, W[13], acc[8]);
        acc[9] = fma(e, W[14], acc[9]);
      }
      { const double e = rec[8];
        acc[10] = fma(e, W[0], acc[10]);
        acc[11] = fma(e, W[1], acc[11]);
        acc[12] = fma(e, W[2], acc[12]);
        acc[13] = fma(e, W[3], acc[13]);
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[9];
        acc[10] = fma(e, W[15], acc[10]);
        acc[11] = fma(e, W[16], acc[11]);
        acc[12] = fma(e, W[17], acc[12]);
        acc[13] = fma(e, W[18], acc[13]);
        acc[14] = fma(e, W[19], acc[14]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[6];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[11];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[4];
    }
    break;
  }
  default: break;
  }
}

// class (10|3): L = 4, record 10 doubles, 2 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_103(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
    const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa, const double* __restrict__ boys,
    double* __restrict__ out, int64_t ldo) {
  const int pidx = blockIdx.x;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int c = xa.lo + (int)blockIdx.y * 64 + (int)threadIdx.x;
  if (xa.lo + (int)blockIdx.y * 64 >= cend) return;
  const bool active = c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 3, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (blockIdx.z) {
  case 0: {
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[5];
      boys_array<4>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      double W[16];
#pragma unroll
      for (int q = 0; q < 16; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      {
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[0] = fma(0.6123724356957945, r, W[0]);
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[4] = fma(0.6123724356957945, r, W[4]);
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      }
      {
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[8] = fma(0.6123724356957945, r, W[8]);
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[0] = fma(-0.20412414523193148, r, W[0]);
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[4] = fma(-0.20412414523193148, r, W[4]);
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[8] = fma(-0.20412414523193148, r, W[8]);
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[3] = fma(-0.38729833462074165, r, W[3]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[7] = fma(-0.38729833462074165, r, W[7]);
      }
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[12] = fma(0.6123724356957945, r, W[12]);
        W[5] = fma(1.0, r, W[5]);
        W[14] = fma(-0.15811388300841894, r, W[14]);
        W[11] = fma(-0.38729833462074165, r, W[11]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[3] = fma(-0.38729833462074165, r, W[3]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[9] = fma(1.0, r, W[9]);
        W[7] = fma(-0.38729833462074165, r, W[7]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[12] = fma(-0.20412414523193148, r, W[12]);
        W[14] = fma(-0.15811388300841894, r, W[14]);
        W[11] = fma(-0.38729833462074165, r, W[11]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      {
      const double y0_2_4 = py0 * z2_4;
      const double y0_2_3 = py0 * z2_3;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[15] = fma(-0.38729833462074165, r, W[15]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[2] = fma(0.6324555320336758, r, W[2]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[13] = fma(1.0, r, W[13]);
        W[6] = fma(0.6324555320336758, r, W[6]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[10] = fma(0.6324555320336758, r, W[10]);
        W[15] = fma(-0.38729833462074165, r, W[15]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[3] = fma(0.2581988897471611, r, W[3]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[7] = fma(0.2581988897471611, r, W[7]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[14] = fma(0.6324555320336758, r, W[14]);
        W[11] = fma(0.2581988897471611, r, W[11]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[15] = fma(0.2581988897471611, r, W[15]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
        acc[3] = fma(e, W[3], acc[3]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[4], acc[0]);
        acc[1] = fma(e, W[5], acc[1]);
        acc[2] = fma(e, W[6], acc[2]);
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[6];
        acc[4] = fma(e, W[0], acc[4]);
        acc[5] = fma(e, W[1], acc[5]);
        acc[6] = fma(e, W[2], acc[6]);
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[7];
        acc[4] = fma(e, W[8], acc[4]);
        acc[5] = fma(e, W[9], acc[5]);
        acc[6] = fma(e, W[10], acc[6]);
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[8];
        acc[8] = fma(e, W[0], acc[8]);
        acc[9] = fma(e, W[1], acc[9]);
        acc[10] = fma(e, W[2], acc[10]);
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[9];
        acc[8] = fma(e, W[12], acc[8]);
        acc[9] = fma(e, W[13], acc[9]);
        acc[10] = fma(e, W[14], acc[10]);
        acc[11] = fma(e, W[15], acc[11]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[5];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[3];
    }
    break;
  }
  case 1: {
    double acc[9];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[5];
      boys_array<4>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      double W[12];
#pragma unroll
      for (int q = 0; q < 12; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      {
      const double y0_0_3 = py0 * z0_3;
      const double y0_0_2 = py0 * z0_2;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[0] = fma(-0.15811388300841894, r, W[0]);
        W[2] = fma(0.20412414523193148, r, W[2]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[3] = fma(-0.15811388300841894, r, W[3]);
        W[5] = fma(0.20412414523193148, r, W[5]);
      }
      }
      {
      const double y1_0_4 = py1 * z0_4;
      const double y1_0_3 = py1 * z0_3;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[6] = fma(-0.15811388300841894, r, W[6]);
        W[8] = fma(0.20412414523193148, r, W[8]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[0] = fma(-0.15811388300841894, r, W[0]);
        W[2] = fma(-0.6123724356957945, r, W[2]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[3] = fma(-0.15811388300841894, r, W[3]);
        W[5] = fma(-0.6123724356957945, r, W[5]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[6] = fma(-0.15811388300841894, r, W[6]);
        W[8] = fma(-0.6123724356957945, r, W[8]);
      }
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[1] = fma(0.5, r, W[1]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[4] = fma(0.5, r, W[4]);
        W[11] = fma(0.20412414523193148, r, W[11]);
      }
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[7] = fma(0.5, r, W[7]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[1] = fma(-0.5, r, W[1]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[4] = fma(-0.5, r, W[4]);
        W[11] = fma(-0.6123724356957945, r, W[11]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[7] = fma(-0.5, r, W[7]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      {
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[0] = fma(0.6324555320336758, r, W[0]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[3] = fma(0.6324555320336758, r, W[3]);
        W[10] = fma(0.5, r, W[10]);
      }
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[6] = fma(0.6324555320336758, r, W[6]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[10] = fma(-0.5, r, W[10]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[9] = fma(0.6324555320336758, r, W[9]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[3], acc[0]);
        acc[1] = fma(e, W[4], acc[1]);
        acc[2] = fma(e, W[5], acc[2]);
      }
      { const double e = rec[6];
        acc[3] = fma(e, W[0], acc[3]);
        acc[4] = fma(e, W[1], acc[4]);
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[7];
        acc[3] = fma(e, W[6], acc[3]);
        acc[4] = fma(e, W[7], acc[4]);
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[8];
        acc[6] = fma(e, W[0], acc[6]);
        acc[7] = fma(e, W[1], acc[7]);
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[9];
        acc[6] = fma(e, W[9], acc[6]);
        acc[7] = fma(e, W[10], acc[7]);
        acc[8] = fma(e, W[11], acc[8]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[4];
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[5];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[6];
    out[(row0 + 1) * ldo + col + 5] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 6] = 1.0 * acc[8];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[0];
    out[(row0 + 2) * ldo + col + 5] = 1.0 * acc[1];
    out[(row0 + 2) * ldo + col + 6] = 1.0 * acc[2];
    }
    break;
  }
  default: break;
  }
}

void register_gen3_10(Gen3Table& t) {
  t[1][0][0] = Gen3Launch{&int3c2e_gen_100, 1, 10};
  t[1][0][1] = Gen3Launch{&int3c2e_gen_101, 1, 10};
  t[1][0][2] = Gen3Launch{&int3c2e_gen_102, 1, 10};
  t[1][0][3] = Gen3Launch{&int3c2e_gen_103, 2, 10};
}

}  // namespace dfhip
