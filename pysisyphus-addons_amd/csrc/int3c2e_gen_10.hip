// Generated by gen3c.py -- 3c2e classes (10|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (10|0), FP64: L = 1, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_100(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 0, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
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
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
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

// class (10|0), FP32: L = 1, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_100(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 0, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const float* __restrict__ recp32 = recs32 + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 10;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[3];
#pragma unroll
      for (int q = 0; q < 3; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      double F[2];
      boys_array<1>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      float W[4];
#pragma unroll
      for (int q = 0; q < 4; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      {
      const float y0_0_0 = py0 * z0_0;
      { const float r = px0 * y0_0_0;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y0_0_1 = py0 * z0_1;
      { const float r = px1 * y0_0_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      }
      {
      const float y1_0_1 = py1 * z0_1;
      { const float r = px0 * y1_0_1;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      {
      const float y0_1_1 = py0 * z1_1;
      { const float r = px0 * y0_1_1;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[1], accf[0]);
      }
      { const float e = rf[6];
        accf[1] = fmaf(e, W[0], accf[1]);
      }
      { const float e = rf[7];
        accf[1] = fmaf(e, W[2], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[3], accf[2]);
      }
      }
#pragma unroll
      for (int q = 0; q < 3; q++) acc[q] += (double)accf[q];
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

// class (10|1), FP64: L = 2, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_101(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 1, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
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
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
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

// class (10|1), FP32: L = 2, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_101(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 1, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const float* __restrict__ recp32 = recs32 + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[9];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 10;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[9];
#pragma unroll
      for (int q = 0; q < 9; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
      double F[3];
      boys_array<2>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      float W[12];
#pragma unroll
      for (int q = 0; q < 12; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      {
      const float y0_0_1 = py0 * z0_1;
      { const float r = px1 * y0_0_1;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      const float y0_0_2 = py0 * z0_2;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      }
      {
      const float y1_0_1 = py1 * z0_1;
      { const float r = px0 * y1_0_1;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[3] = fmaf(1.0f, r, W[3]);
        W[8] = fmaf(1.0f, r, W[8]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[6] = fmaf(1.0f, r, W[6]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      {
      const float y0_1_1 = py0 * z1_1;
      { const float r = px0 * y0_1_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[4] = fmaf(1.0f, r, W[4]);
        W[11] = fmaf(1.0f, r, W[11]);
      }
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[9] = fmaf(1.0f, r, W[9]);
        W[7] = fmaf(1.0f, r, W[7]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[10] = fmaf(1.0f, r, W[10]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
        accf[2] = fmaf(e, W[2], accf[2]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[3], accf[0]);
        accf[1] = fmaf(e, W[4], accf[1]);
        accf[2] = fmaf(e, W[5], accf[2]);
      }
      { const float e = rf[6];
        accf[3] = fmaf(e, W[0], accf[3]);
        accf[4] = fmaf(e, W[1], accf[4]);
        accf[5] = fmaf(e, W[2], accf[5]);
      }
      { const float e = rf[7];
        accf[3] = fmaf(e, W[6], accf[3]);
        accf[4] = fmaf(e, W[7], accf[4]);
        accf[5] = fmaf(e, W[8], accf[5]);
      }
      { const float e = rf[8];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
        accf[8] = fmaf(e, W[2], accf[8]);
      }
      { const float e = rf[9];
        accf[6] = fmaf(e, W[9], accf[6]);
        accf[7] = fmaf(e, W[10], accf[7]);
        accf[8] = fmaf(e, W[11], accf[8]);
      }
      }
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q] += (double)accf[q];
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

// class (10|2), FP64: L = 3, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_102(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 2, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
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
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
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

// class (10|2), FP32: L = 3, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_102(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 2, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const float* __restrict__ recp32 = recs32 + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[15];
#pragma unroll
    for (int q = 0; q < 15; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 10;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[15];
#pragma unroll
      for (int q = 0; q < 15; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[4];
      boys_array<3>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      float W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      {
      const float y0_0_2 = py0 * z0_2;
      const float y0_0_1 = py0 * z0_1;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
        W[4] = fmaf(0.5f, r, W[4]);
      }
      const float y0_0_3 = py0 * z0_3;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[7] = fmaf(-0.2886751345948129f, r, W[7]);
        W[9] = fmaf(0.5f, r, W[9]);
      }
      }
      {
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y1_0_3 = py1 * z0_3;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[5] = fmaf(1.0f, r, W[5]);
        W[12] = fmaf(-0.2886751345948129f, r, W[12]);
        W[14] = fmaf(0.5f, r, W[14]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
        W[4] = fmaf(-0.5f, r, W[4]);
      }
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[10] = fmaf(1.0f, r, W[10]);
        W[7] = fmaf(-0.2886751345948129f, r, W[7]);
        W[9] = fmaf(-0.5f, r, W[9]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[12] = fmaf(-0.2886751345948129f, r, W[12]);
        W[14] = fmaf(-0.5f, r, W[14]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      {
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y0_1_3 = py0 * z1_3;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[17] = fmaf(-0.2886751345948129f, r, W[17]);
        W[8] = fmaf(1.0f, r, W[8]);
        W[19] = fmaf(0.5f, r, W[19]);
      }
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[15] = fmaf(1.0f, r, W[15]);
        W[6] = fmaf(1.0f, r, W[6]);
        W[13] = fmaf(1.0f, r, W[13]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[11] = fmaf(1.0f, r, W[11]);
        W[17] = fmaf(-0.2886751345948129f, r, W[17]);
        W[19] = fmaf(-0.5f, r, W[19]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[2] = fmaf(0.5773502691896258f, r, W[2]);
      }
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[7] = fmaf(0.5773502691896258f, r, W[7]);
        W[18] = fmaf(1.0f, r, W[18]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[16] = fmaf(1.0f, r, W[16]);
        W[12] = fmaf(0.5773502691896258f, r, W[12]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[17] = fmaf(0.5773502691896258f, r, W[17]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
        accf[4] = fmaf(e, W[4], accf[4]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[5], accf[0]);
        accf[1] = fmaf(e, W[6], accf[1]);
        accf[2] = fmaf(e, W[7], accf[2]);
        accf[3] = fmaf(e, W[8], accf[3]);
        accf[4] = fmaf(e, W[9], accf[4]);
      }
      { const float e = rf[6];
        accf[5] = fmaf(e, W[0], accf[5]);
        accf[6] = fmaf(e, W[1], accf[6]);
        accf[7] = fmaf(e, W[2], accf[7]);
        accf[8] = fmaf(e, W[3], accf[8]);
        accf[9] = fmaf(e, W[4], accf[9]);
      }
      { const float e = rf[7];
        accf[5] = fmaf(e, W[10], accf[5]);
        accf[6] = fmaf(e, W[11], accf[6]);
        accf[7] = fmaf(e, W[12], accf[7]);
        accf[8] = fmaf(e, W[13], accf[8]);
        accf[9] = fmaf(e, W[14], accf[9]);
      }
      { const float e = rf[8];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
        accf[14] = fmaf(e, W[4], accf[14]);
      }
      { const float e = rf[9];
        accf[10] = fmaf(e, W[15], accf[10]);
        accf[11] = fmaf(e, W[16], accf[11]);
        accf[12] = fmaf(e, W[17], accf[12]);
        accf[13] = fmaf(e, W[18], accf[13]);
        accf[14] = fmaf(e, W[19], accf[14]);
      }
      }
#pragma unroll
      for (int q = 0; q < 15; q++) acc[q] += (double)accf[q];
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

// class (10|3), FP64: L = 4, record 10 values, 2 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_103(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 2;
  const int pidx = (xs / 2) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 3, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
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
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
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
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
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

// class (10|3), FP32: L = 4, record 10 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_103(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 1;
  const int pidx = (xs / 1) * 8 + (int)(blockIdx.x & 7);
  if (pidx >= pa.npairs) return;
  const int cend = min(xa.hi, pa.nsurv[pidx]);
  const int cbeg = pa.nbeg ? pa.nbeg[pidx] : 0;
  const int c0 = xa.lo + (int)blockIdx.y * 64, c = c0 + (int)threadIdx.x;
  if (c0 >= cend || c0 + 64 <= cbeg) return;
  const bool active = c >= cbeg && c < cend;
  const int sh = active ? xa.shell[c] : xa.shell[xa.lo];
  const double Cx = ket.xyz[3 * sh], Cy = ket.xyz[3 * sh + 1], Cz = ket.xyz[3 * sh + 2];
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 3, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const float* __restrict__ recp32 = recs32 + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[21];
#pragma unroll
    for (int q = 0; q < 21; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 10;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 10;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[21];
#pragma unroll
      for (int q = 0; q < 21; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[5];
      boys_array<4>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      float W[28];
#pragma unroll
      for (int q = 0; q < 28; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      {
      const float y0_0_3 = py0 * z0_3;
      const float y0_0_2 = py0 * z0_2;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
        W[6] = fmaf(0.20412414523193148f, r, W[6]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[11] = fmaf(-0.15811388300841894f, r, W[11]);
        W[13] = fmaf(0.20412414523193148f, r, W[13]);
      }
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[0] = fmaf(0.6123724356957945f, r, W[0]);
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[7] = fmaf(0.6123724356957945f, r, W[7]);
        W[9] = fmaf(-0.15811388300841894f, r, W[9]);
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[20] = fmaf(0.20412414523193148f, r, W[20]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
        W[6] = fmaf(-0.6123724356957945f, r, W[6]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[14] = fmaf(0.6123724356957945f, r, W[14]);
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
        W[11] = fmaf(-0.15811388300841894f, r, W[11]);
        W[13] = fmaf(-0.6123724356957945f, r, W[13]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[0] = fmaf(-0.20412414523193148f, r, W[0]);
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[7] = fmaf(-0.20412414523193148f, r, W[7]);
        W[9] = fmaf(-0.15811388300841894f, r, W[9]);
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[20] = fmaf(-0.6123724356957945f, r, W[20]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[14] = fmaf(-0.20412414523193148f, r, W[14]);
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      {
      const float y0_1_3 = py0 * z1_3;
      const float y0_1_2 = py0 * z1_2;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
        W[5] = fmaf(0.5f, r, W[5]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[10] = fmaf(-0.38729833462074165f, r, W[10]);
        W[25] = fmaf(-0.15811388300841894f, r, W[25]);
        W[12] = fmaf(0.5f, r, W[12]);
        W[27] = fmaf(0.20412414523193148f, r, W[27]);
      }
      }
      {
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[21] = fmaf(0.6123724356957945f, r, W[21]);
        W[8] = fmaf(1.0f, r, W[8]);
        W[23] = fmaf(-0.15811388300841894f, r, W[23]);
        W[17] = fmaf(-0.38729833462074165f, r, W[17]);
        W[19] = fmaf(0.5f, r, W[19]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
        W[5] = fmaf(-0.5f, r, W[5]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[15] = fmaf(1.0f, r, W[15]);
        W[10] = fmaf(-0.38729833462074165f, r, W[10]);
        W[25] = fmaf(-0.15811388300841894f, r, W[25]);
        W[12] = fmaf(-0.5f, r, W[12]);
        W[27] = fmaf(-0.6123724356957945f, r, W[27]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[21] = fmaf(-0.20412414523193148f, r, W[21]);
        W[23] = fmaf(-0.15811388300841894f, r, W[23]);
        W[17] = fmaf(-0.38729833462074165f, r, W[17]);
        W[19] = fmaf(-0.5f, r, W[19]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      {
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[4] = fmaf(0.6324555320336758f, r, W[4]);
      }
      const float y0_2_4 = py0 * z2_4;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[24] = fmaf(-0.38729833462074165f, r, W[24]);
        W[11] = fmaf(0.6324555320336758f, r, W[11]);
        W[26] = fmaf(0.5f, r, W[26]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[2] = fmaf(0.6324555320336758f, r, W[2]);
      }
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[22] = fmaf(1.0f, r, W[22]);
        W[9] = fmaf(0.6324555320336758f, r, W[9]);
        W[18] = fmaf(0.6324555320336758f, r, W[18]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[16] = fmaf(0.6324555320336758f, r, W[16]);
        W[24] = fmaf(-0.38729833462074165f, r, W[24]);
        W[26] = fmaf(-0.5f, r, W[26]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[3] = fmaf(0.2581988897471611f, r, W[3]);
      }
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[10] = fmaf(0.2581988897471611f, r, W[10]);
        W[25] = fmaf(0.6324555320336758f, r, W[25]);
      }
      }
      {
      const float y1_3_4 = py1 * z3_4;
      { const float r = px0 * y1_3_4;
        W[23] = fmaf(0.6324555320336758f, r, W[23]);
        W[17] = fmaf(0.2581988897471611f, r, W[17]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      {
      const float y0_4_4 = py0 * z4_4;
      { const float r = px0 * y0_4_4;
        W[24] = fmaf(0.2581988897471611f, r, W[24]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
        accf[4] = fmaf(e, W[4], accf[4]);
        accf[5] = fmaf(e, W[5], accf[5]);
        accf[6] = fmaf(e, W[6], accf[6]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[7], accf[0]);
        accf[1] = fmaf(e, W[8], accf[1]);
        accf[2] = fmaf(e, W[9], accf[2]);
        accf[3] = fmaf(e, W[10], accf[3]);
        accf[4] = fmaf(e, W[11], accf[4]);
        accf[5] = fmaf(e, W[12], accf[5]);
        accf[6] = fmaf(e, W[13], accf[6]);
      }
      { const float e = rf[6];
        accf[7] = fmaf(e, W[0], accf[7]);
        accf[8] = fmaf(e, W[1], accf[8]);
        accf[9] = fmaf(e, W[2], accf[9]);
        accf[10] = fmaf(e, W[3], accf[10]);
        accf[11] = fmaf(e, W[4], accf[11]);
        accf[12] = fmaf(e, W[5], accf[12]);
        accf[13] = fmaf(e, W[6], accf[13]);
      }
      { const float e = rf[7];
        accf[7] = fmaf(e, W[14], accf[7]);
        accf[8] = fmaf(e, W[15], accf[8]);
        accf[9] = fmaf(e, W[16], accf[9]);
        accf[10] = fmaf(e, W[17], accf[10]);
        accf[11] = fmaf(e, W[18], accf[11]);
        accf[12] = fmaf(e, W[19], accf[12]);
        accf[13] = fmaf(e, W[20], accf[13]);
      }
      { const float e = rf[8];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
        accf[20] = fmaf(e, W[6], accf[20]);
      }
      { const float e = rf[9];
        accf[14] = fmaf(e, W[21], accf[14]);
        accf[15] = fmaf(e, W[22], accf[15]);
        accf[16] = fmaf(e, W[23], accf[16]);
        accf[17] = fmaf(e, W[24], accf[17]);
        accf[18] = fmaf(e, W[25], accf[18]);
        accf[19] = fmaf(e, W[26], accf[19]);
        accf[20] = fmaf(e, W[27], accf[20]);
      }
      }
#pragma unroll
      for (int q = 0; q < 21; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[9];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[10];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[11];
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[12];
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[13];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[15];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[17];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[18];
    out[(row0 + 1) * ldo + col + 5] = 1.0 * acc[19];
    out[(row0 + 1) * ldo + col + 6] = 1.0 * acc[20];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 2) * ldo + col + 5] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 6] = 1.0 * acc[6];
    }
    break;
  }
  default: break;
  }
}

void register_gen3_10(Gen3Table& t) {
  t[1][0][0] = Gen3Launch{&int3c2e_gen_100, &int3c2e_gen32_100, 1, 1, 10};
  t[1][0][1] = Gen3Launch{&int3c2e_gen_101, &int3c2e_gen32_101, 1, 1, 10};
  t[1][0][2] = Gen3Launch{&int3c2e_gen_102, &int3c2e_gen32_102, 1, 1, 10};
  t[1][0][3] = Gen3Launch{&int3c2e_gen_103, &int3c2e_gen32_103, 2, 1, 10};
}

}  // namespace dfhip
