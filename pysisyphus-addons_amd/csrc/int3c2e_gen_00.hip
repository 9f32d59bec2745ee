// Generated by gen3c.py -- 3c2e classes (00|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (00|0), FP64: L = 0, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_000(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[1];
#pragma unroll
    for (int q = 0; q < 1; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      double F[1];
      boys_array<0>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      double W[1];
#pragma unroll
      for (int q = 0; q < 1; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      {
      const double y0_0_0 = py0 * z0_0;
      { const double r = px0 * y0_0_0;
        W[0] = fma(1.0, r, W[0]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    }
    break;
  }
  default: break;
  }
}

// class (00|0), FP32: L = 0, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_000(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[1];
#pragma unroll
    for (int q = 0; q < 1; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 5;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[1];
#pragma unroll
      for (int q = 0; q < 1; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      double F[1];
      boys_array<0>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      float W[1];
#pragma unroll
      for (int q = 0; q < 1; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      {
      const float y0_0_0 = py0 * z0_0;
      { const float r = px0 * y0_0_0;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
      }
      }
#pragma unroll
      for (int q = 0; q < 1; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    }
    break;
  }
  default: break;
  }
}

// class (00|1), FP64: L = 1, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_001(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
      double F[2];
      boys_array<1>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      double W[3];
#pragma unroll
      for (int q = 0; q < 3; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      {
      const double y0_0_1 = py0 * z0_1;
      { const double r = px1 * y0_0_1;
        W[2] = fma(1.0, r, W[2]);
      }
      }
      {
      const double y1_0_1 = py1 * z0_1;
      { const double r = px0 * y1_0_1;
        W[0] = fma(1.0, r, W[0]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      {
      const double y0_1_1 = py0 * z1_1;
      { const double r = px0 * y0_1_1;
        W[1] = fma(1.0, r, W[1]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    }
    break;
  }
  default: break;
  }
}

// class (00|1), FP32: L = 1, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_001(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[3];
#pragma unroll
    for (int q = 0; q < 3; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 5;
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
      { const double i2g = 0.5 / gam; sc *= -i2g; }
      double F[2];
      boys_array<1>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      float W[3];
#pragma unroll
      for (int q = 0; q < 3; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      {
      const float y0_0_1 = py0 * z0_1;
      { const float r = px1 * y0_0_1;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      }
      {
      const float y1_0_1 = py1 * z0_1;
      { const float r = px0 * y1_0_1;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      {
      const float y0_1_1 = py0 * z1_1;
      { const float r = px0 * y0_1_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
        accf[2] = fmaf(e, W[2], accf[2]);
      }
      }
#pragma unroll
      for (int q = 0; q < 3; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    }
    break;
  }
  default: break;
  }
}

// class (00|2), FP64: L = 2, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_002(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[5];
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[5];
#pragma unroll
      for (int q = 0; q < 5; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      {
      const double y0_0_2 = py0 * z0_2;
      const double y0_0_1 = py0 * z0_1;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[2] = fma(-0.2886751345948129, r, W[2]);
        W[4] = fma(0.5, r, W[4]);
      }
      }
      {
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[0] = fma(1.0, r, W[0]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[2] = fma(-0.2886751345948129, r, W[2]);
        W[4] = fma(-0.5, r, W[4]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      {
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[3] = fma(1.0, r, W[3]);
      }
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[1] = fma(1.0, r, W[1]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[2] = fma(0.5773502691896258, r, W[2]);
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
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    }
    break;
  }
  default: break;
  }
}

// class (00|2), FP32: L = 2, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_002(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[5];
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 5;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[5];
#pragma unroll
      for (int q = 0; q < 5; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      float W[5];
#pragma unroll
      for (int q = 0; q < 5; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      {
      const float y0_0_2 = py0 * z0_2;
      const float y0_0_1 = py0 * z0_1;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
        W[4] = fmaf(0.5f, r, W[4]);
      }
      }
      {
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
        W[4] = fmaf(-0.5f, r, W[4]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      {
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[2] = fmaf(0.5773502691896258f, r, W[2]);
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
      }
#pragma unroll
      for (int q = 0; q < 5; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    }
    break;
  }
  default: break;
  }
}

// class (00|3), FP64: L = 3, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_003(const double* __restrict__ recs, const float* __restrict__ recs32,
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
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[7];
#pragma unroll
    for (int q = 0; q < 7; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
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
      double W[7];
#pragma unroll
      for (int q = 0; q < 7; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      {
      const double y0_0_3 = py0 * z0_3;
      const double y0_0_2 = py0 * z0_2;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[4] = fma(-0.15811388300841894, r, W[4]);
        W[6] = fma(0.20412414523193148, r, W[6]);
      }
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[0] = fma(0.6123724356957945, r, W[0]);
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[4] = fma(-0.15811388300841894, r, W[4]);
        W[6] = fma(-0.6123724356957945, r, W[6]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[0] = fma(-0.20412414523193148, r, W[0]);
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[3] = fma(-0.38729833462074165, r, W[3]);
        W[5] = fma(0.5, r, W[5]);
      }
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[1] = fma(1.0, r, W[1]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[3] = fma(-0.38729833462074165, r, W[3]);
        W[5] = fma(-0.5, r, W[5]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      {
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[4] = fma(0.6324555320336758, r, W[4]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[2] = fma(0.6324555320336758, r, W[2]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[3] = fma(0.2581988897471611, r, W[3]);
      }
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
        acc[1] = fma(e, W[1], acc[1]);
        acc[2] = fma(e, W[2], acc[2]);
        acc[3] = fma(e, W[3], acc[3]);
        acc[4] = fma(e, W[4], acc[4]);
        acc[5] = fma(e, W[5], acc[5]);
        acc[6] = fma(e, W[6], acc[6]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[5];
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[6];
    }
    break;
  }
  default: break;
  }
}

// class (00|3), FP32: L = 3, record 5 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_003(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[7];
#pragma unroll
    for (int q = 0; q < 7; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 5;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 5;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[7];
#pragma unroll
      for (int q = 0; q < 7; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
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
      float W[7];
#pragma unroll
      for (int q = 0; q < 7; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      {
      const float y0_0_3 = py0 * z0_3;
      const float y0_0_2 = py0 * z0_2;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
        W[6] = fmaf(0.20412414523193148f, r, W[6]);
      }
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[0] = fmaf(0.6123724356957945f, r, W[0]);
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
        W[6] = fmaf(-0.6123724356957945f, r, W[6]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[0] = fmaf(-0.20412414523193148f, r, W[0]);
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      {
      const float y0_1_3 = py0 * z1_3;
      const float y0_1_2 = py0 * z1_2;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
        W[5] = fmaf(0.5f, r, W[5]);
      }
      }
      {
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
        W[5] = fmaf(-0.5f, r, W[5]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      {
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[4] = fmaf(0.6324555320336758f, r, W[4]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[2] = fmaf(0.6324555320336758f, r, W[2]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[3] = fmaf(0.2581988897471611f, r, W[3]);
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
      }
#pragma unroll
      for (int q = 0; q < 7; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[0];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[1];
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[2];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[5];
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[6];
    }
    break;
  }
  default: break;
  }
}

void register_gen3_00(Gen3Table& t) {
  t[0][0][0] = Gen3Launch{&int3c2e_gen_000, &int3c2e_gen32_000, 1, 1, 5};
  t[0][0][1] = Gen3Launch{&int3c2e_gen_001, &int3c2e_gen32_001, 1, 1, 5};
  t[0][0][2] = Gen3Launch{&int3c2e_gen_002, &int3c2e_gen32_002, 1, 1, 5};
  t[0][0][3] = Gen3Launch{&int3c2e_gen_003, &int3c2e_gen32_003, 1, 1, 5};
}

}  // namespace dfhip
