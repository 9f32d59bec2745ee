// Generated by gen3c.py -- 3c2e classes (21|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (21|0), FP64: L = 3, record 112 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_210(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
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
      const double y0_0_0 = py0 * z0_0;
      { const double r = px0 * y0_0_0;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y0_0_1 = py0 * z0_1;
      { const double r = px1 * y0_0_1;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y0_0_2 = py0 * z0_2;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y0_0_3 = py0 * z0_3;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y1_0_1 = py1 * z0_1;
      { const double r = px0 * y1_0_1;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y1_0_3 = py1 * z0_3;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      {
      const double y0_1_1 = py0 * z1_1;
      { const double r = px0 * y0_1_1;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y0_1_3 = py0 * z1_3;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[19] = fma(1.0, r, W[19]);
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
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|0), FP32: L = 3, record 112 values, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_210(const double* __restrict__ recs, const float* __restrict__ recs32,
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
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[18];
#pragma unroll
      for (int q = 0; q < 18; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
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
      const float y0_0_0 = py0 * z0_0;
      { const float r = px0 * y0_0_0;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y0_0_1 = py0 * z0_1;
      { const float r = px1 * y0_0_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y0_0_2 = py0 * z0_2;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[4] = fmaf(1.0f, r, W[4]);
      }
      const float y0_0_3 = py0 * z0_3;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[10] = fmaf(1.0f, r, W[10]);
      }
      }
      {
      const float y1_0_1 = py1 * z0_1;
      { const float r = px0 * y1_0_1;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y1_0_3 = py1 * z0_3;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[11] = fmaf(1.0f, r, W[11]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[13] = fmaf(1.0f, r, W[13]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[16] = fmaf(1.0f, r, W[16]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      {
      const float y0_1_1 = py0 * z1_1;
      { const float r = px0 * y0_1_1;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[6] = fmaf(1.0f, r, W[6]);
      }
      const float y0_1_3 = py0 * z1_3;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[12] = fmaf(1.0f, r, W[12]);
      }
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[8] = fmaf(1.0f, r, W[8]);
      }
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[14] = fmaf(1.0f, r, W[14]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[17] = fmaf(1.0f, r, W[17]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[15] = fmaf(1.0f, r, W[15]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[18] = fmaf(1.0f, r, W[18]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[19] = fmaf(1.0f, r, W[19]);
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
        accf[0] = fmaf(e, W[4], accf[0]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[10], accf[0]);
      }
      { const float e = rf[8];
        accf[1] = fmaf(e, W[0], accf[1]);
      }
      { const float e = rf[9];
        accf[1] = fmaf(e, W[2], accf[1]);
      }
      { const float e = rf[10];
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[11];
        accf[1] = fmaf(e, W[5], accf[1]);
      }
      { const float e = rf[12];
        accf[1] = fmaf(e, W[4], accf[1]);
      }
      { const float e = rf[13];
        accf[1] = fmaf(e, W[11], accf[1]);
      }
      { const float e = rf[14];
        accf[2] = fmaf(e, W[0], accf[2]);
      }
      { const float e = rf[15];
        accf[2] = fmaf(e, W[3], accf[2]);
      }
      { const float e = rf[16];
        accf[2] = fmaf(e, W[1], accf[2]);
      }
      { const float e = rf[17];
        accf[2] = fmaf(e, W[6], accf[2]);
      }
      { const float e = rf[18];
        accf[2] = fmaf(e, W[4], accf[2]);
      }
      { const float e = rf[19];
        accf[2] = fmaf(e, W[12], accf[2]);
      }
      { const float e = rf[20];
        accf[3] = fmaf(e, W[0], accf[3]);
      }
      { const float e = rf[21];
        accf[3] = fmaf(e, W[2], accf[3]);
      }
      { const float e = rf[22];
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[23];
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[24];
        accf[3] = fmaf(e, W[4], accf[3]);
      }
      { const float e = rf[25];
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[26];
        accf[4] = fmaf(e, W[0], accf[4]);
      }
      { const float e = rf[27];
        accf[4] = fmaf(e, W[2], accf[4]);
      }
      { const float e = rf[28];
        accf[4] = fmaf(e, W[7], accf[4]);
      }
      { const float e = rf[29];
        accf[4] = fmaf(e, W[1], accf[4]);
      }
      { const float e = rf[30];
        accf[4] = fmaf(e, W[5], accf[4]);
      }
      { const float e = rf[31];
        accf[4] = fmaf(e, W[13], accf[4]);
      }
      { const float e = rf[32];
        accf[5] = fmaf(e, W[0], accf[5]);
      }
      { const float e = rf[33];
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[34];
        accf[5] = fmaf(e, W[2], accf[5]);
      }
      { const float e = rf[35];
        accf[5] = fmaf(e, W[8], accf[5]);
      }
      { const float e = rf[36];
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[37];
        accf[5] = fmaf(e, W[6], accf[5]);
      }
      { const float e = rf[38];
        accf[5] = fmaf(e, W[5], accf[5]);
      }
      { const float e = rf[39];
        accf[5] = fmaf(e, W[14], accf[5]);
      }
      { const float e = rf[40];
        accf[6] = fmaf(e, W[0], accf[6]);
      }
      { const float e = rf[41];
        accf[6] = fmaf(e, W[3], accf[6]);
      }
      { const float e = rf[42];
        accf[6] = fmaf(e, W[1], accf[6]);
      }
      { const float e = rf[43];
        accf[6] = fmaf(e, W[6], accf[6]);
      }
      { const float e = rf[44];
        accf[6] = fmaf(e, W[4], accf[6]);
      }
      { const float e = rf[45];
        accf[6] = fmaf(e, W[12], accf[6]);
      }
      { const float e = rf[46];
        accf[7] = fmaf(e, W[0], accf[7]);
      }
      { const float e = rf[47];
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[48];
        accf[7] = fmaf(e, W[2], accf[7]);
      }
      { const float e = rf[49];
        accf[7] = fmaf(e, W[8], accf[7]);
      }
      { const float e = rf[50];
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[51];
        accf[7] = fmaf(e, W[6], accf[7]);
      }
      { const float e = rf[52];
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[53];
        accf[7] = fmaf(e, W[14], accf[7]);
      }
      { const float e = rf[54];
        accf[8] = fmaf(e, W[0], accf[8]);
      }
      { const float e = rf[55];
        accf[8] = fmaf(e, W[3], accf[8]);
      }
      { const float e = rf[56];
        accf[8] = fmaf(e, W[9], accf[8]);
      }
      { const float e = rf[57];
        accf[8] = fmaf(e, W[1], accf[8]);
      }
      { const float e = rf[58];
        accf[8] = fmaf(e, W[6], accf[8]);
      }
      { const float e = rf[59];
        accf[8] = fmaf(e, W[15], accf[8]);
      }
      { const float e = rf[60];
        accf[9] = fmaf(e, W[0], accf[9]);
      }
      { const float e = rf[61];
        accf[9] = fmaf(e, W[2], accf[9]);
      }
      { const float e = rf[62];
        accf[9] = fmaf(e, W[7], accf[9]);
      }
      { const float e = rf[63];
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[64];
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[65];
        accf[9] = fmaf(e, W[13], accf[9]);
      }
      { const float e = rf[66];
        accf[10] = fmaf(e, W[0], accf[10]);
      }
      { const float e = rf[67];
        accf[10] = fmaf(e, W[2], accf[10]);
      }
      { const float e = rf[68];
        accf[10] = fmaf(e, W[7], accf[10]);
      }
      { const float e = rf[69];
        accf[10] = fmaf(e, W[16], accf[10]);
      }
      { const float e = rf[70];
        accf[11] = fmaf(e, W[0], accf[11]);
      }
      { const float e = rf[71];
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[72];
        accf[11] = fmaf(e, W[2], accf[11]);
      }
      { const float e = rf[73];
        accf[11] = fmaf(e, W[8], accf[11]);
      }
      { const float e = rf[74];
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[75];
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[76];
        accf[12] = fmaf(e, W[0], accf[12]);
      }
      { const float e = rf[77];
        accf[12] = fmaf(e, W[3], accf[12]);
      }
      { const float e = rf[78];
        accf[12] = fmaf(e, W[2], accf[12]);
      }
      { const float e = rf[79];
        accf[12] = fmaf(e, W[8], accf[12]);
      }
      { const float e = rf[80];
        accf[12] = fmaf(e, W[1], accf[12]);
      }
      { const float e = rf[81];
        accf[12] = fmaf(e, W[6], accf[12]);
      }
      { const float e = rf[82];
        accf[12] = fmaf(e, W[5], accf[12]);
      }
      { const float e = rf[83];
        accf[12] = fmaf(e, W[14], accf[12]);
      }
      { const float e = rf[84];
        accf[13] = fmaf(e, W[0], accf[13]);
      }
      { const float e = rf[85];
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[86];
        accf[13] = fmaf(e, W[2], accf[13]);
      }
      { const float e = rf[87];
        accf[13] = fmaf(e, W[8], accf[13]);
      }
      { const float e = rf[88];
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[89];
        accf[13] = fmaf(e, W[17], accf[13]);
      }
      { const float e = rf[90];
        accf[14] = fmaf(e, W[0], accf[14]);
      }
      { const float e = rf[91];
        accf[14] = fmaf(e, W[3], accf[14]);
      }
      { const float e = rf[92];
        accf[14] = fmaf(e, W[9], accf[14]);
      }
      { const float e = rf[93];
        accf[14] = fmaf(e, W[2], accf[14]);
      }
      { const float e = rf[94];
        accf[14] = fmaf(e, W[8], accf[14]);
      }
      { const float e = rf[95];
        accf[14] = fmaf(e, W[18], accf[14]);
      }
      { const float e = rf[96];
        accf[15] = fmaf(e, W[0], accf[15]);
      }
      { const float e = rf[97];
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[98];
        accf[15] = fmaf(e, W[9], accf[15]);
      }
      { const float e = rf[99];
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[100];
        accf[15] = fmaf(e, W[6], accf[15]);
      }
      { const float e = rf[101];
        accf[15] = fmaf(e, W[15], accf[15]);
      }
      { const float e = rf[102];
        accf[16] = fmaf(e, W[0], accf[16]);
      }
      { const float e = rf[103];
        accf[16] = fmaf(e, W[3], accf[16]);
      }
      { const float e = rf[104];
        accf[16] = fmaf(e, W[9], accf[16]);
      }
      { const float e = rf[105];
        accf[16] = fmaf(e, W[2], accf[16]);
      }
      { const float e = rf[106];
        accf[16] = fmaf(e, W[8], accf[16]);
      }
      { const float e = rf[107];
        accf[16] = fmaf(e, W[18], accf[16]);
      }
      { const float e = rf[108];
        accf[17] = fmaf(e, W[0], accf[17]);
      }
      { const float e = rf[109];
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[110];
        accf[17] = fmaf(e, W[9], accf[17]);
      }
      { const float e = rf[111];
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      }
#pragma unroll
      for (int q = 0; q < 18; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|1), FP64: L = 4, record 112 values, 3 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_211(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 3;
  const int pidx = (xs / 3) * 8 + (int)(blockIdx.x & 7);
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
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      {
      }
      {
      const double y1_0_1 = py1 * z0_1;
      { const double r = px0 * y1_0_1;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_0_3 = py1 * z0_3;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      {
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      {
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[19] = fma(1.0, r, W[19]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 1: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      {
      }
      {
      }
      {
      }
      {
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
      const double y0_1_1 = py0 * z1_1;
      { const double r = px0 * y0_1_1;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y0_1_3 = py0 * z1_3;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[19] = fma(1.0, r, W[19]);
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
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 2: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      {
      const double y0_0_1 = py0 * z0_1;
      { const double r = px1 * y0_0_1;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y0_0_2 = py0 * z0_2;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y0_0_3 = py0 * z0_3;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y1_0_3 = py1 * z0_3;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[16] = fma(1.0, r, W[16]);
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
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y0_1_3 = py0 * z1_3;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      {
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[19] = fma(1.0, r, W[19]);
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
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|1), FP32: L = 4, record 112 values, 2 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_211(const double* __restrict__ recs, const float* __restrict__ recs32,
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
  const bool live = active && triplet_kept(pa, xa, pidx, sh, 1, Cx, Cy, Cz);
  const int nprim_c = live ? ket.nprim[sh] : 0, pc0 = ket.pptr[sh];
  const int npp = npp_v[pidx];
  const double* __restrict__ recp = recs + recoff[pidx];
  const float* __restrict__ recp32 = recs32 + recoff[pidx];
  const int64_t row0 = pa.row[pidx];
  const int64_t col = active ? xa.col[c] : 0;
  switch (grp) {
  case 0: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      {
      }
      {
      const float y1_0_1 = py1 * z0_1;
      { const float r = px0 * y1_0_1;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      const float y1_0_3 = py1 * z0_3;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[8] = fmaf(1.0f, r, W[8]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[20] = fmaf(1.0f, r, W[20]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[4] = fmaf(1.0f, r, W[4]);
      }
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[10] = fmaf(1.0f, r, W[10]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[22] = fmaf(1.0f, r, W[22]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[14] = fmaf(1.0f, r, W[14]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[26] = fmaf(1.0f, r, W[26]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[32] = fmaf(1.0f, r, W[32]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      {
      const float y0_1_1 = py0 * z1_1;
      { const float r = px0 * y0_1_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y0_1_3 = py0 * z1_3;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[21] = fmaf(1.0f, r, W[21]);
      }
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[6] = fmaf(1.0f, r, W[6]);
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[12] = fmaf(1.0f, r, W[12]);
        W[11] = fmaf(1.0f, r, W[11]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[24] = fmaf(1.0f, r, W[24]);
        W[23] = fmaf(1.0f, r, W[23]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[16] = fmaf(1.0f, r, W[16]);
        W[15] = fmaf(1.0f, r, W[15]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[28] = fmaf(1.0f, r, W[28]);
        W[27] = fmaf(1.0f, r, W[27]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[34] = fmaf(1.0f, r, W[34]);
        W[33] = fmaf(1.0f, r, W[33]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[13] = fmaf(1.0f, r, W[13]);
      }
      const float y0_2_4 = py0 * z2_4;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[25] = fmaf(1.0f, r, W[25]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[18] = fmaf(1.0f, r, W[18]);
        W[17] = fmaf(1.0f, r, W[17]);
      }
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[30] = fmaf(1.0f, r, W[30]);
        W[29] = fmaf(1.0f, r, W[29]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[36] = fmaf(1.0f, r, W[36]);
        W[35] = fmaf(1.0f, r, W[35]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[19] = fmaf(1.0f, r, W[19]);
      }
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[31] = fmaf(1.0f, r, W[31]);
      }
      }
      {
      const float y1_3_4 = py1 * z3_4;
      { const float r = px0 * y1_3_4;
        W[38] = fmaf(1.0f, r, W[38]);
        W[37] = fmaf(1.0f, r, W[37]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      {
      const float y0_4_4 = py0 * z4_4;
      { const float r = px0 * y0_4_4;
        W[39] = fmaf(1.0f, r, W[39]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 1: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[18];
#pragma unroll
      for (int q = 0; q < 18; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      float W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      {
      const float y0_0_1 = py0 * z0_1;
      { const float r = px1 * y0_0_1;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y0_0_2 = py0 * z0_2;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y0_0_3 = py0 * z0_3;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[4] = fmaf(1.0f, r, W[4]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[10] = fmaf(1.0f, r, W[10]);
      }
      }
      {
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      const float y1_0_3 = py1 * z0_3;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[11] = fmaf(1.0f, r, W[11]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[13] = fmaf(1.0f, r, W[13]);
      }
      }
      {
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[16] = fmaf(1.0f, r, W[16]);
      }
      }
      {
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      {
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y0_1_3 = py0 * z1_3;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[6] = fmaf(1.0f, r, W[6]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[12] = fmaf(1.0f, r, W[12]);
      }
      }
      {
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[8] = fmaf(1.0f, r, W[8]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[14] = fmaf(1.0f, r, W[14]);
      }
      }
      {
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[17] = fmaf(1.0f, r, W[17]);
      }
      }
      {
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      {
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y0_2_4 = py0 * z2_4;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[15] = fmaf(1.0f, r, W[15]);
      }
      }
      {
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[18] = fmaf(1.0f, r, W[18]);
      }
      }
      {
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      {
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[19] = fmaf(1.0f, r, W[19]);
      }
      }
      {
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[1], accf[0]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[4], accf[0]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[10], accf[0]);
      }
      { const float e = rf[8];
        accf[1] = fmaf(e, W[0], accf[1]);
      }
      { const float e = rf[9];
        accf[1] = fmaf(e, W[2], accf[1]);
      }
      { const float e = rf[10];
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[11];
        accf[1] = fmaf(e, W[5], accf[1]);
      }
      { const float e = rf[12];
        accf[1] = fmaf(e, W[4], accf[1]);
      }
      { const float e = rf[13];
        accf[1] = fmaf(e, W[11], accf[1]);
      }
      { const float e = rf[14];
        accf[2] = fmaf(e, W[0], accf[2]);
      }
      { const float e = rf[15];
        accf[2] = fmaf(e, W[3], accf[2]);
      }
      { const float e = rf[16];
        accf[2] = fmaf(e, W[1], accf[2]);
      }
      { const float e = rf[17];
        accf[2] = fmaf(e, W[6], accf[2]);
      }
      { const float e = rf[18];
        accf[2] = fmaf(e, W[4], accf[2]);
      }
      { const float e = rf[19];
        accf[2] = fmaf(e, W[12], accf[2]);
      }
      { const float e = rf[20];
        accf[3] = fmaf(e, W[0], accf[3]);
      }
      { const float e = rf[21];
        accf[3] = fmaf(e, W[2], accf[3]);
      }
      { const float e = rf[22];
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[23];
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[24];
        accf[3] = fmaf(e, W[4], accf[3]);
      }
      { const float e = rf[25];
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[26];
        accf[4] = fmaf(e, W[0], accf[4]);
      }
      { const float e = rf[27];
        accf[4] = fmaf(e, W[2], accf[4]);
      }
      { const float e = rf[28];
        accf[4] = fmaf(e, W[7], accf[4]);
      }
      { const float e = rf[29];
        accf[4] = fmaf(e, W[1], accf[4]);
      }
      { const float e = rf[30];
        accf[4] = fmaf(e, W[5], accf[4]);
      }
      { const float e = rf[31];
        accf[4] = fmaf(e, W[13], accf[4]);
      }
      { const float e = rf[32];
        accf[5] = fmaf(e, W[0], accf[5]);
      }
      { const float e = rf[33];
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[34];
        accf[5] = fmaf(e, W[2], accf[5]);
      }
      { const float e = rf[35];
        accf[5] = fmaf(e, W[8], accf[5]);
      }
      { const float e = rf[36];
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[37];
        accf[5] = fmaf(e, W[6], accf[5]);
      }
      { const float e = rf[38];
        accf[5] = fmaf(e, W[5], accf[5]);
      }
      { const float e = rf[39];
        accf[5] = fmaf(e, W[14], accf[5]);
      }
      { const float e = rf[40];
        accf[6] = fmaf(e, W[0], accf[6]);
      }
      { const float e = rf[41];
        accf[6] = fmaf(e, W[3], accf[6]);
      }
      { const float e = rf[42];
        accf[6] = fmaf(e, W[1], accf[6]);
      }
      { const float e = rf[43];
        accf[6] = fmaf(e, W[6], accf[6]);
      }
      { const float e = rf[44];
        accf[6] = fmaf(e, W[4], accf[6]);
      }
      { const float e = rf[45];
        accf[6] = fmaf(e, W[12], accf[6]);
      }
      { const float e = rf[46];
        accf[7] = fmaf(e, W[0], accf[7]);
      }
      { const float e = rf[47];
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[48];
        accf[7] = fmaf(e, W[2], accf[7]);
      }
      { const float e = rf[49];
        accf[7] = fmaf(e, W[8], accf[7]);
      }
      { const float e = rf[50];
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[51];
        accf[7] = fmaf(e, W[6], accf[7]);
      }
      { const float e = rf[52];
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[53];
        accf[7] = fmaf(e, W[14], accf[7]);
      }
      { const float e = rf[54];
        accf[8] = fmaf(e, W[0], accf[8]);
      }
      { const float e = rf[55];
        accf[8] = fmaf(e, W[3], accf[8]);
      }
      { const float e = rf[56];
        accf[8] = fmaf(e, W[9], accf[8]);
      }
      { const float e = rf[57];
        accf[8] = fmaf(e, W[1], accf[8]);
      }
      { const float e = rf[58];
        accf[8] = fmaf(e, W[6], accf[8]);
      }
      { const float e = rf[59];
        accf[8] = fmaf(e, W[15], accf[8]);
      }
      { const float e = rf[60];
        accf[9] = fmaf(e, W[0], accf[9]);
      }
      { const float e = rf[61];
        accf[9] = fmaf(e, W[2], accf[9]);
      }
      { const float e = rf[62];
        accf[9] = fmaf(e, W[7], accf[9]);
      }
      { const float e = rf[63];
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[64];
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[65];
        accf[9] = fmaf(e, W[13], accf[9]);
      }
      { const float e = rf[66];
        accf[10] = fmaf(e, W[0], accf[10]);
      }
      { const float e = rf[67];
        accf[10] = fmaf(e, W[2], accf[10]);
      }
      { const float e = rf[68];
        accf[10] = fmaf(e, W[7], accf[10]);
      }
      { const float e = rf[69];
        accf[10] = fmaf(e, W[16], accf[10]);
      }
      { const float e = rf[70];
        accf[11] = fmaf(e, W[0], accf[11]);
      }
      { const float e = rf[71];
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[72];
        accf[11] = fmaf(e, W[2], accf[11]);
      }
      { const float e = rf[73];
        accf[11] = fmaf(e, W[8], accf[11]);
      }
      { const float e = rf[74];
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[75];
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[76];
        accf[12] = fmaf(e, W[0], accf[12]);
      }
      { const float e = rf[77];
        accf[12] = fmaf(e, W[3], accf[12]);
      }
      { const float e = rf[78];
        accf[12] = fmaf(e, W[2], accf[12]);
      }
      { const float e = rf[79];
        accf[12] = fmaf(e, W[8], accf[12]);
      }
      { const float e = rf[80];
        accf[12] = fmaf(e, W[1], accf[12]);
      }
      { const float e = rf[81];
        accf[12] = fmaf(e, W[6], accf[12]);
      }
      { const float e = rf[82];
        accf[12] = fmaf(e, W[5], accf[12]);
      }
      { const float e = rf[83];
        accf[12] = fmaf(e, W[14], accf[12]);
      }
      { const float e = rf[84];
        accf[13] = fmaf(e, W[0], accf[13]);
      }
      { const float e = rf[85];
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[86];
        accf[13] = fmaf(e, W[2], accf[13]);
      }
      { const float e = rf[87];
        accf[13] = fmaf(e, W[8], accf[13]);
      }
      { const float e = rf[88];
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[89];
        accf[13] = fmaf(e, W[17], accf[13]);
      }
      { const float e = rf[90];
        accf[14] = fmaf(e, W[0], accf[14]);
      }
      { const float e = rf[91];
        accf[14] = fmaf(e, W[3], accf[14]);
      }
      { const float e = rf[92];
        accf[14] = fmaf(e, W[9], accf[14]);
      }
      { const float e = rf[93];
        accf[14] = fmaf(e, W[2], accf[14]);
      }
      { const float e = rf[94];
        accf[14] = fmaf(e, W[8], accf[14]);
      }
      { const float e = rf[95];
        accf[14] = fmaf(e, W[18], accf[14]);
      }
      { const float e = rf[96];
        accf[15] = fmaf(e, W[0], accf[15]);
      }
      { const float e = rf[97];
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[98];
        accf[15] = fmaf(e, W[9], accf[15]);
      }
      { const float e = rf[99];
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[100];
        accf[15] = fmaf(e, W[6], accf[15]);
      }
      { const float e = rf[101];
        accf[15] = fmaf(e, W[15], accf[15]);
      }
      { const float e = rf[102];
        accf[16] = fmaf(e, W[0], accf[16]);
      }
      { const float e = rf[103];
        accf[16] = fmaf(e, W[3], accf[16]);
      }
      { const float e = rf[104];
        accf[16] = fmaf(e, W[9], accf[16]);
      }
      { const float e = rf[105];
        accf[16] = fmaf(e, W[2], accf[16]);
      }
      { const float e = rf[106];
        accf[16] = fmaf(e, W[8], accf[16]);
      }
      { const float e = rf[107];
        accf[16] = fmaf(e, W[18], accf[16]);
      }
      { const float e = rf[108];
        accf[17] = fmaf(e, W[0], accf[17]);
      }
      { const float e = rf[109];
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[110];
        accf[17] = fmaf(e, W[9], accf[17]);
      }
      { const float e = rf[111];
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      }
#pragma unroll
      for (int q = 0; q < 18; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|2), FP64: L = 5, record 112 values, 5 m-group(s)
__global__ __launch_bounds__(64, 3) void int3c2e_gen_212(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 5;
  const int pidx = (xs / 5) * 8 + (int)(blockIdx.x & 7);
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
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      {
      }
      {
      const double y1_0_2 = py1 * z0_2;
      { const double r = px1 * y1_0_2;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_0_3 = py1 * z0_3;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[16] = fma(1.0, r, W[16]);
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
      const double z1_5 = pz1 * g5;
      {
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      {
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      {
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[19] = fma(1.0, r, W[19]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 1: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      {
      }
      {
      const double y1_1_2 = py1 * z1_2;
      { const double r = px0 * y1_1_2;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      {
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      {
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[19] = fma(1.0, r, W[19]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 2: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      {
      const double y0_0_2 = py0 * z0_2;
      const double y0_0_1 = py0 * z0_1;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[0] = fma(-0.2886751345948129, r, W[0]);
      }
      const double y0_0_3 = py0 * z0_3;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[1] = fma(-0.2886751345948129, r, W[1]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[4] = fma(-0.2886751345948129, r, W[4]);
      }
      const double y0_0_5 = py0 * z0_5;
      { const double r = px5 * y0_0_5 + 20.0 * px3 * y0_0_4 + 60.0 * px1 * y0_0_3;
        W[10] = fma(-0.2886751345948129, r, W[10]);
      }
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[2] = fma(-0.2886751345948129, r, W[2]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[5] = fma(-0.2886751345948129, r, W[5]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[11] = fma(-0.2886751345948129, r, W[11]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[0] = fma(-0.2886751345948129, r, W[0]);
      }
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[1] = fma(-0.2886751345948129, r, W[1]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[7] = fma(-0.2886751345948129, r, W[7]);
        W[4] = fma(-0.2886751345948129, r, W[4]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[13] = fma(-0.2886751345948129, r, W[13]);
        W[10] = fma(-0.2886751345948129, r, W[10]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[2] = fma(-0.2886751345948129, r, W[2]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[5] = fma(-0.2886751345948129, r, W[5]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[16] = fma(-0.2886751345948129, r, W[16]);
        W[11] = fma(-0.2886751345948129, r, W[11]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[7] = fma(-0.2886751345948129, r, W[7]);
      }
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[13] = fma(-0.2886751345948129, r, W[13]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[16] = fma(-0.2886751345948129, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[3] = fma(-0.2886751345948129, r, W[3]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[6] = fma(-0.2886751345948129, r, W[6]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[12] = fma(-0.2886751345948129, r, W[12]);
      }
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[8] = fma(-0.2886751345948129, r, W[8]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[14] = fma(-0.2886751345948129, r, W[14]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[3] = fma(-0.2886751345948129, r, W[3]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[6] = fma(-0.2886751345948129, r, W[6]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[17] = fma(-0.2886751345948129, r, W[17]);
        W[12] = fma(-0.2886751345948129, r, W[12]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[8] = fma(-0.2886751345948129, r, W[8]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[14] = fma(-0.2886751345948129, r, W[14]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[17] = fma(-0.2886751345948129, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      {
      const double y0_2_2 = py0 * z2_2;
      { const double r = px0 * y0_2_2;
        W[0] = fma(0.5773502691896258, r, W[0]);
      }
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[1] = fma(0.5773502691896258, r, W[1]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[9] = fma(-0.2886751345948129, r, W[9]);
        W[4] = fma(0.5773502691896258, r, W[4]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[15] = fma(-0.2886751345948129, r, W[15]);
        W[10] = fma(0.5773502691896258, r, W[10]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[2] = fma(0.5773502691896258, r, W[2]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[5] = fma(0.5773502691896258, r, W[5]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[18] = fma(-0.2886751345948129, r, W[18]);
        W[11] = fma(0.5773502691896258, r, W[11]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[9] = fma(-0.2886751345948129, r, W[9]);
        W[7] = fma(0.5773502691896258, r, W[7]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[15] = fma(-0.2886751345948129, r, W[15]);
        W[13] = fma(0.5773502691896258, r, W[13]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[18] = fma(-0.2886751345948129, r, W[18]);
        W[16] = fma(0.5773502691896258, r, W[16]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[3] = fma(0.5773502691896258, r, W[3]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[6] = fma(0.5773502691896258, r, W[6]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[19] = fma(-0.2886751345948129, r, W[19]);
        W[12] = fma(0.5773502691896258, r, W[12]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[8] = fma(0.5773502691896258, r, W[8]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[14] = fma(0.5773502691896258, r, W[14]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[19] = fma(-0.2886751345948129, r, W[19]);
        W[17] = fma(0.5773502691896258, r, W[17]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[9] = fma(0.5773502691896258, r, W[9]);
      }
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[15] = fma(0.5773502691896258, r, W[15]);
      }
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[18] = fma(0.5773502691896258, r, W[18]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      const double y0_5_5 = py0 * z5_5;
      { const double r = px0 * y0_5_5;
        W[19] = fma(0.5773502691896258, r, W[19]);
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
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 3: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      {
      const double y0_1_2 = py0 * z1_2;
      { const double r = px1 * y0_1_2;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y0_1_3 = py0 * z1_3;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      {
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      {
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[19] = fma(1.0, r, W[19]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 3] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 3] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 3] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 3] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 3] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 3] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 4: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      {
      const double y0_0_2 = py0 * z0_2;
      const double y0_0_1 = py0 * z0_1;
      { const double r = px2 * y0_0_2 + 2.0 * px0 * y0_0_1;
        W[0] = fma(0.5, r, W[0]);
      }
      const double y0_0_3 = py0 * z0_3;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[1] = fma(0.5, r, W[1]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[4] = fma(0.5, r, W[4]);
      }
      const double y0_0_5 = py0 * z0_5;
      { const double r = px5 * y0_0_5 + 20.0 * px3 * y0_0_4 + 60.0 * px1 * y0_0_3;
        W[10] = fma(0.5, r, W[10]);
      }
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[2] = fma(0.5, r, W[2]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[5] = fma(0.5, r, W[5]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[11] = fma(0.5, r, W[11]);
      }
      }
      {
      const double y2_0_2 = py2 * z0_2 + 2.0 * py0 * z0_1;
      { const double r = px0 * y2_0_2;
        W[0] = fma(-0.5, r, W[0]);
      }
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[1] = fma(-0.5, r, W[1]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[7] = fma(0.5, r, W[7]);
        W[4] = fma(-0.5, r, W[4]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[13] = fma(0.5, r, W[13]);
        W[10] = fma(-0.5, r, W[10]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[2] = fma(-0.5, r, W[2]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[5] = fma(-0.5, r, W[5]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[16] = fma(0.5, r, W[16]);
        W[11] = fma(-0.5, r, W[11]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[7] = fma(-0.5, r, W[7]);
      }
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[13] = fma(-0.5, r, W[13]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[16] = fma(-0.5, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[3] = fma(0.5, r, W[3]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[6] = fma(0.5, r, W[6]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[12] = fma(0.5, r, W[12]);
      }
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[8] = fma(0.5, r, W[8]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[14] = fma(0.5, r, W[14]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[3] = fma(-0.5, r, W[3]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[6] = fma(-0.5, r, W[6]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[17] = fma(0.5, r, W[17]);
        W[12] = fma(-0.5, r, W[12]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[8] = fma(-0.5, r, W[8]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[14] = fma(-0.5, r, W[14]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[17] = fma(-0.5, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      {
      const double y0_2_4 = py0 * z2_4;
      const double y0_2_3 = py0 * z2_3;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[9] = fma(0.5, r, W[9]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[15] = fma(0.5, r, W[15]);
      }
      }
      {
      const double y1_2_5 = py1 * z2_5;
      const double y1_2_4 = py1 * z2_4;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[18] = fma(0.5, r, W[18]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[9] = fma(-0.5, r, W[9]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[15] = fma(-0.5, r, W[15]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[18] = fma(-0.5, r, W[18]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      {
      const double y0_3_5 = py0 * z3_5;
      const double y0_3_4 = py0 * z3_4;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[19] = fma(0.5, r, W[19]);
      }
      }
      {
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[19] = fma(-0.5, r, W[19]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 4] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 4] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 4] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 4] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 4] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 4] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|2), FP32: L = 5, record 112 values, 3 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_212(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 3;
  const int pidx = (xs / 3) * 8 + (int)(blockIdx.x & 7);
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      {
      }
      {
      const float y1_0_2 = py1 * z0_2;
      { const float r = px1 * y1_0_2;
        W[0] = fmaf(1.0f, r, W[0]);
      }
      const float y1_0_3 = py1 * z0_3;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[2] = fmaf(1.0f, r, W[2]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[8] = fmaf(1.0f, r, W[8]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[20] = fmaf(1.0f, r, W[20]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[4] = fmaf(1.0f, r, W[4]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[10] = fmaf(1.0f, r, W[10]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[22] = fmaf(1.0f, r, W[22]);
      }
      }
      {
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[14] = fmaf(1.0f, r, W[14]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[26] = fmaf(1.0f, r, W[26]);
      }
      }
      {
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[32] = fmaf(1.0f, r, W[32]);
      }
      }
      {
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      {
      }
      {
      const float y1_1_2 = py1 * z1_2;
      { const float r = px0 * y1_1_2;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[6] = fmaf(1.0f, r, W[6]);
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[12] = fmaf(1.0f, r, W[12]);
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[24] = fmaf(1.0f, r, W[24]);
        W[21] = fmaf(1.0f, r, W[21]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[16] = fmaf(1.0f, r, W[16]);
        W[11] = fmaf(1.0f, r, W[11]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[28] = fmaf(1.0f, r, W[28]);
        W[23] = fmaf(1.0f, r, W[23]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[15] = fmaf(1.0f, r, W[15]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[34] = fmaf(1.0f, r, W[34]);
        W[27] = fmaf(1.0f, r, W[27]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[33] = fmaf(1.0f, r, W[33]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      {
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[18] = fmaf(1.0f, r, W[18]);
        W[13] = fmaf(1.0f, r, W[13]);
      }
      const float y1_2_5 = py1 * z2_5;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[30] = fmaf(1.0f, r, W[30]);
        W[25] = fmaf(1.0f, r, W[25]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[17] = fmaf(1.0f, r, W[17]);
      }
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[36] = fmaf(1.0f, r, W[36]);
        W[29] = fmaf(1.0f, r, W[29]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[35] = fmaf(1.0f, r, W[35]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      {
      }
      {
      const float y1_3_4 = py1 * z3_4;
      { const float r = px0 * y1_3_4;
        W[19] = fmaf(1.0f, r, W[19]);
      }
      const float y1_3_5 = py1 * z3_5;
      { const float r = px1 * y1_3_5;
        W[38] = fmaf(1.0f, r, W[38]);
        W[31] = fmaf(1.0f, r, W[31]);
      }
      }
      {
      const float y2_3_5 = py2 * z3_5 + 2.0f * py0 * z3_4;
      { const float r = px0 * y2_3_5;
        W[37] = fmaf(1.0f, r, W[37]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      {
      }
      {
      const float y1_4_5 = py1 * z4_5;
      { const float r = px0 * y1_4_5;
        W[39] = fmaf(1.0f, r, W[39]);
      }
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 1: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      {
      const float y0_0_2 = py0 * z0_2;
      const float y0_0_1 = py0 * z0_1;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[0] = fmaf(-0.2886751345948129f, r, W[0]);
      }
      const float y0_0_3 = py0 * z0_3;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[8] = fmaf(-0.2886751345948129f, r, W[8]);
      }
      const float y0_0_5 = py0 * z0_5;
      { const float r = px5 * y0_0_5 + 20.0f * px3 * y0_0_4 + 60.0f * px1 * y0_0_3;
        W[20] = fmaf(-0.2886751345948129f, r, W[20]);
      }
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[4] = fmaf(-0.2886751345948129f, r, W[4]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[10] = fmaf(-0.2886751345948129f, r, W[10]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[22] = fmaf(-0.2886751345948129f, r, W[22]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[0] = fmaf(-0.2886751345948129f, r, W[0]);
      }
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[2] = fmaf(-0.2886751345948129f, r, W[2]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[14] = fmaf(-0.2886751345948129f, r, W[14]);
        W[8] = fmaf(-0.2886751345948129f, r, W[8]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[26] = fmaf(-0.2886751345948129f, r, W[26]);
        W[20] = fmaf(-0.2886751345948129f, r, W[20]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[4] = fmaf(-0.2886751345948129f, r, W[4]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[10] = fmaf(-0.2886751345948129f, r, W[10]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[32] = fmaf(-0.2886751345948129f, r, W[32]);
        W[22] = fmaf(-0.2886751345948129f, r, W[22]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[14] = fmaf(-0.2886751345948129f, r, W[14]);
      }
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[26] = fmaf(-0.2886751345948129f, r, W[26]);
      }
      }
      {
      const float y5_0_5 = py5 * z0_5 + 20.0f * py3 * z0_4 + 60.0f * py1 * z0_3;
      { const float r = px0 * y5_0_5;
        W[32] = fmaf(-0.2886751345948129f, r, W[32]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      {
      const float y0_1_2 = py0 * z1_2;
      { const float r = px1 * y0_1_2;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y0_1_3 = py0 * z1_3;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[6] = fmaf(-0.2886751345948129f, r, W[6]);
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[12] = fmaf(-0.2886751345948129f, r, W[12]);
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y0_1_5 = py0 * z1_5;
      { const float r = px4 * y0_1_5 + 12.0f * px2 * y0_1_4 + 12.0f * px0 * y0_1_3;
        W[24] = fmaf(-0.2886751345948129f, r, W[24]);
        W[21] = fmaf(1.0f, r, W[21]);
      }
      }
      {
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[16] = fmaf(-0.2886751345948129f, r, W[16]);
        W[11] = fmaf(1.0f, r, W[11]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[28] = fmaf(-0.2886751345948129f, r, W[28]);
        W[23] = fmaf(1.0f, r, W[23]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[6] = fmaf(-0.2886751345948129f, r, W[6]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[12] = fmaf(-0.2886751345948129f, r, W[12]);
        W[15] = fmaf(1.0f, r, W[15]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[34] = fmaf(-0.2886751345948129f, r, W[34]);
        W[24] = fmaf(-0.2886751345948129f, r, W[24]);
        W[27] = fmaf(1.0f, r, W[27]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[16] = fmaf(-0.2886751345948129f, r, W[16]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[28] = fmaf(-0.2886751345948129f, r, W[28]);
        W[33] = fmaf(1.0f, r, W[33]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[34] = fmaf(-0.2886751345948129f, r, W[34]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      {
      const float y0_2_2 = py0 * z2_2;
      { const float r = px0 * y0_2_2;
        W[0] = fmaf(0.5773502691896258f, r, W[0]);
      }
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[2] = fmaf(0.5773502691896258f, r, W[2]);
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y0_2_4 = py0 * z2_4;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[18] = fmaf(-0.2886751345948129f, r, W[18]);
        W[8] = fmaf(0.5773502691896258f, r, W[8]);
        W[13] = fmaf(1.0f, r, W[13]);
      }
      const float y0_2_5 = py0 * z2_5;
      { const float r = px3 * y0_2_5 + 6.0f * px1 * y0_2_4;
        W[30] = fmaf(-0.2886751345948129f, r, W[30]);
        W[20] = fmaf(0.5773502691896258f, r, W[20]);
        W[25] = fmaf(1.0f, r, W[25]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[4] = fmaf(0.5773502691896258f, r, W[4]);
      }
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[10] = fmaf(0.5773502691896258f, r, W[10]);
        W[17] = fmaf(1.0f, r, W[17]);
      }
      const float y1_2_5 = py1 * z2_5;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[36] = fmaf(-0.2886751345948129f, r, W[36]);
        W[22] = fmaf(0.5773502691896258f, r, W[22]);
        W[29] = fmaf(1.0f, r, W[29]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[18] = fmaf(-0.2886751345948129f, r, W[18]);
        W[14] = fmaf(0.5773502691896258f, r, W[14]);
      }
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[30] = fmaf(-0.2886751345948129f, r, W[30]);
        W[26] = fmaf(0.5773502691896258f, r, W[26]);
        W[35] = fmaf(1.0f, r, W[35]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[36] = fmaf(-0.2886751345948129f, r, W[36]);
        W[32] = fmaf(0.5773502691896258f, r, W[32]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[6] = fmaf(0.5773502691896258f, r, W[6]);
      }
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[12] = fmaf(0.5773502691896258f, r, W[12]);
        W[19] = fmaf(1.0f, r, W[19]);
      }
      const float y0_3_5 = py0 * z3_5;
      { const float r = px2 * y0_3_5 + 2.0f * px0 * y0_3_4;
        W[38] = fmaf(-0.2886751345948129f, r, W[38]);
        W[24] = fmaf(0.5773502691896258f, r, W[24]);
        W[31] = fmaf(1.0f, r, W[31]);
      }
      }
      {
      const float y1_3_4 = py1 * z3_4;
      { const float r = px0 * y1_3_4;
        W[16] = fmaf(0.5773502691896258f, r, W[16]);
      }
      const float y1_3_5 = py1 * z3_5;
      { const float r = px1 * y1_3_5;
        W[28] = fmaf(0.5773502691896258f, r, W[28]);
        W[37] = fmaf(1.0f, r, W[37]);
      }
      }
      {
      const float y2_3_5 = py2 * z3_5 + 2.0f * py0 * z3_4;
      { const float r = px0 * y2_3_5;
        W[38] = fmaf(-0.2886751345948129f, r, W[38]);
        W[34] = fmaf(0.5773502691896258f, r, W[34]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      {
      const float y0_4_4 = py0 * z4_4;
      { const float r = px0 * y0_4_4;
        W[18] = fmaf(0.5773502691896258f, r, W[18]);
      }
      const float y0_4_5 = py0 * z4_5;
      { const float r = px1 * y0_4_5;
        W[30] = fmaf(0.5773502691896258f, r, W[30]);
        W[39] = fmaf(1.0f, r, W[39]);
      }
      }
      {
      const float y1_4_5 = py1 * z4_5;
      { const float r = px0 * y1_4_5;
        W[36] = fmaf(0.5773502691896258f, r, W[36]);
      }
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      {
      const float y0_5_5 = py0 * z5_5;
      { const float r = px0 * y0_5_5;
        W[38] = fmaf(0.5773502691896258f, r, W[38]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 3] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 3] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 3] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 3] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 3] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 3] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 2: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[18];
#pragma unroll
      for (int q = 0; q < 18; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
      double F[6];
      boys_array<5>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      float W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      {
      const float y0_0_2 = py0 * z0_2;
      const float y0_0_1 = py0 * z0_1;
      { const float r = px2 * y0_0_2 + 2.0f * px0 * y0_0_1;
        W[0] = fmaf(0.5f, r, W[0]);
      }
      const float y0_0_3 = py0 * z0_3;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[1] = fmaf(0.5f, r, W[1]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[4] = fmaf(0.5f, r, W[4]);
      }
      const float y0_0_5 = py0 * z0_5;
      { const float r = px5 * y0_0_5 + 20.0f * px3 * y0_0_4 + 60.0f * px1 * y0_0_3;
        W[10] = fmaf(0.5f, r, W[10]);
      }
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[2] = fmaf(0.5f, r, W[2]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[5] = fmaf(0.5f, r, W[5]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[11] = fmaf(0.5f, r, W[11]);
      }
      }
      {
      const float y2_0_2 = py2 * z0_2 + 2.0f * py0 * z0_1;
      { const float r = px0 * y2_0_2;
        W[0] = fmaf(-0.5f, r, W[0]);
      }
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[1] = fmaf(-0.5f, r, W[1]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[7] = fmaf(0.5f, r, W[7]);
        W[4] = fmaf(-0.5f, r, W[4]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[13] = fmaf(0.5f, r, W[13]);
        W[10] = fmaf(-0.5f, r, W[10]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[2] = fmaf(-0.5f, r, W[2]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[5] = fmaf(-0.5f, r, W[5]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[16] = fmaf(0.5f, r, W[16]);
        W[11] = fmaf(-0.5f, r, W[11]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[7] = fmaf(-0.5f, r, W[7]);
      }
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[13] = fmaf(-0.5f, r, W[13]);
      }
      }
      {
      const float y5_0_5 = py5 * z0_5 + 20.0f * py3 * z0_4 + 60.0f * py1 * z0_3;
      { const float r = px0 * y5_0_5;
        W[16] = fmaf(-0.5f, r, W[16]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      {
      const float y0_1_3 = py0 * z1_3;
      const float y0_1_2 = py0 * z1_2;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[3] = fmaf(0.5f, r, W[3]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[6] = fmaf(0.5f, r, W[6]);
      }
      const float y0_1_5 = py0 * z1_5;
      { const float r = px4 * y0_1_5 + 12.0f * px2 * y0_1_4 + 12.0f * px0 * y0_1_3;
        W[12] = fmaf(0.5f, r, W[12]);
      }
      }
      {
      const float y1_1_4 = py1 * z1_4;
      const float y1_1_3 = py1 * z1_3;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[8] = fmaf(0.5f, r, W[8]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[14] = fmaf(0.5f, r, W[14]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[3] = fmaf(-0.5f, r, W[3]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[6] = fmaf(-0.5f, r, W[6]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[17] = fmaf(0.5f, r, W[17]);
        W[12] = fmaf(-0.5f, r, W[12]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[8] = fmaf(-0.5f, r, W[8]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[14] = fmaf(-0.5f, r, W[14]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[17] = fmaf(-0.5f, r, W[17]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      {
      const float y0_2_4 = py0 * z2_4;
      const float y0_2_3 = py0 * z2_3;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[9] = fmaf(0.5f, r, W[9]);
      }
      const float y0_2_5 = py0 * z2_5;
      { const float r = px3 * y0_2_5 + 6.0f * px1 * y0_2_4;
        W[15] = fmaf(0.5f, r, W[15]);
      }
      }
      {
      const float y1_2_5 = py1 * z2_5;
      const float y1_2_4 = py1 * z2_4;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[18] = fmaf(0.5f, r, W[18]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[9] = fmaf(-0.5f, r, W[9]);
      }
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[15] = fmaf(-0.5f, r, W[15]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[18] = fmaf(-0.5f, r, W[18]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      {
      const float y0_3_5 = py0 * z3_5;
      const float y0_3_4 = py0 * z3_4;
      { const float r = px2 * y0_3_5 + 2.0f * px0 * y0_3_4;
        W[19] = fmaf(0.5f, r, W[19]);
      }
      }
      {
      }
      {
      const float y2_3_5 = py2 * z3_5 + 2.0f * py0 * z3_4;
      { const float r = px0 * y2_3_5;
        W[19] = fmaf(-0.5f, r, W[19]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      {
      }
      {
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[1], accf[0]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[4], accf[0]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[10], accf[0]);
      }
      { const float e = rf[8];
        accf[1] = fmaf(e, W[0], accf[1]);
      }
      { const float e = rf[9];
        accf[1] = fmaf(e, W[2], accf[1]);
      }
      { const float e = rf[10];
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[11];
        accf[1] = fmaf(e, W[5], accf[1]);
      }
      { const float e = rf[12];
        accf[1] = fmaf(e, W[4], accf[1]);
      }
      { const float e = rf[13];
        accf[1] = fmaf(e, W[11], accf[1]);
      }
      { const float e = rf[14];
        accf[2] = fmaf(e, W[0], accf[2]);
      }
      { const float e = rf[15];
        accf[2] = fmaf(e, W[3], accf[2]);
      }
      { const float e = rf[16];
        accf[2] = fmaf(e, W[1], accf[2]);
      }
      { const float e = rf[17];
        accf[2] = fmaf(e, W[6], accf[2]);
      }
      { const float e = rf[18];
        accf[2] = fmaf(e, W[4], accf[2]);
      }
      { const float e = rf[19];
        accf[2] = fmaf(e, W[12], accf[2]);
      }
      { const float e = rf[20];
        accf[3] = fmaf(e, W[0], accf[3]);
      }
      { const float e = rf[21];
        accf[3] = fmaf(e, W[2], accf[3]);
      }
      { const float e = rf[22];
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[23];
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[24];
        accf[3] = fmaf(e, W[4], accf[3]);
      }
      { const float e = rf[25];
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[26];
        accf[4] = fmaf(e, W[0], accf[4]);
      }
      { const float e = rf[27];
        accf[4] = fmaf(e, W[2], accf[4]);
      }
      { const float e = rf[28];
        accf[4] = fmaf(e, W[7], accf[4]);
      }
      { const float e = rf[29];
        accf[4] = fmaf(e, W[1], accf[4]);
      }
      { const float e = rf[30];
        accf[4] = fmaf(e, W[5], accf[4]);
      }
      { const float e = rf[31];
        accf[4] = fmaf(e, W[13], accf[4]);
      }
      { const float e = rf[32];
        accf[5] = fmaf(e, W[0], accf[5]);
      }
      { const float e = rf[33];
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[34];
        accf[5] = fmaf(e, W[2], accf[5]);
      }
      { const float e = rf[35];
        accf[5] = fmaf(e, W[8], accf[5]);
      }
      { const float e = rf[36];
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[37];
        accf[5] = fmaf(e, W[6], accf[5]);
      }
      { const float e = rf[38];
        accf[5] = fmaf(e, W[5], accf[5]);
      }
      { const float e = rf[39];
        accf[5] = fmaf(e, W[14], accf[5]);
      }
      { const float e = rf[40];
        accf[6] = fmaf(e, W[0], accf[6]);
      }
      { const float e = rf[41];
        accf[6] = fmaf(e, W[3], accf[6]);
      }
      { const float e = rf[42];
        accf[6] = fmaf(e, W[1], accf[6]);
      }
      { const float e = rf[43];
        accf[6] = fmaf(e, W[6], accf[6]);
      }
      { const float e = rf[44];
        accf[6] = fmaf(e, W[4], accf[6]);
      }
      { const float e = rf[45];
        accf[6] = fmaf(e, W[12], accf[6]);
      }
      { const float e = rf[46];
        accf[7] = fmaf(e, W[0], accf[7]);
      }
      { const float e = rf[47];
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[48];
        accf[7] = fmaf(e, W[2], accf[7]);
      }
      { const float e = rf[49];
        accf[7] = fmaf(e, W[8], accf[7]);
      }
      { const float e = rf[50];
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[51];
        accf[7] = fmaf(e, W[6], accf[7]);
      }
      { const float e = rf[52];
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[53];
        accf[7] = fmaf(e, W[14], accf[7]);
      }
      { const float e = rf[54];
        accf[8] = fmaf(e, W[0], accf[8]);
      }
      { const float e = rf[55];
        accf[8] = fmaf(e, W[3], accf[8]);
      }
      { const float e = rf[56];
        accf[8] = fmaf(e, W[9], accf[8]);
      }
      { const float e = rf[57];
        accf[8] = fmaf(e, W[1], accf[8]);
      }
      { const float e = rf[58];
        accf[8] = fmaf(e, W[6], accf[8]);
      }
      { const float e = rf[59];
        accf[8] = fmaf(e, W[15], accf[8]);
      }
      { const float e = rf[60];
        accf[9] = fmaf(e, W[0], accf[9]);
      }
      { const float e = rf[61];
        accf[9] = fmaf(e, W[2], accf[9]);
      }
      { const float e = rf[62];
        accf[9] = fmaf(e, W[7], accf[9]);
      }
      { const float e = rf[63];
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[64];
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[65];
        accf[9] = fmaf(e, W[13], accf[9]);
      }
      { const float e = rf[66];
        accf[10] = fmaf(e, W[0], accf[10]);
      }
      { const float e = rf[67];
        accf[10] = fmaf(e, W[2], accf[10]);
      }
      { const float e = rf[68];
        accf[10] = fmaf(e, W[7], accf[10]);
      }
      { const float e = rf[69];
        accf[10] = fmaf(e, W[16], accf[10]);
      }
      { const float e = rf[70];
        accf[11] = fmaf(e, W[0], accf[11]);
      }
      { const float e = rf[71];
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[72];
        accf[11] = fmaf(e, W[2], accf[11]);
      }
      { const float e = rf[73];
        accf[11] = fmaf(e, W[8], accf[11]);
      }
      { const float e = rf[74];
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[75];
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[76];
        accf[12] = fmaf(e, W[0], accf[12]);
      }
      { const float e = rf[77];
        accf[12] = fmaf(e, W[3], accf[12]);
      }
      { const float e = rf[78];
        accf[12] = fmaf(e, W[2], accf[12]);
      }
      { const float e = rf[79];
        accf[12] = fmaf(e, W[8], accf[12]);
      }
      { const float e = rf[80];
        accf[12] = fmaf(e, W[1], accf[12]);
      }
      { const float e = rf[81];
        accf[12] = fmaf(e, W[6], accf[12]);
      }
      { const float e = rf[82];
        accf[12] = fmaf(e, W[5], accf[12]);
      }
      { const float e = rf[83];
        accf[12] = fmaf(e, W[14], accf[12]);
      }
      { const float e = rf[84];
        accf[13] = fmaf(e, W[0], accf[13]);
      }
      { const float e = rf[85];
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[86];
        accf[13] = fmaf(e, W[2], accf[13]);
      }
      { const float e = rf[87];
        accf[13] = fmaf(e, W[8], accf[13]);
      }
      { const float e = rf[88];
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[89];
        accf[13] = fmaf(e, W[17], accf[13]);
      }
      { const float e = rf[90];
        accf[14] = fmaf(e, W[0], accf[14]);
      }
      { const float e = rf[91];
        accf[14] = fmaf(e, W[3], accf[14]);
      }
      { const float e = rf[92];
        accf[14] = fmaf(e, W[9], accf[14]);
      }
      { const float e = rf[93];
        accf[14] = fmaf(e, W[2], accf[14]);
      }
      { const float e = rf[94];
        accf[14] = fmaf(e, W[8], accf[14]);
      }
      { const float e = rf[95];
        accf[14] = fmaf(e, W[18], accf[14]);
      }
      { const float e = rf[96];
        accf[15] = fmaf(e, W[0], accf[15]);
      }
      { const float e = rf[97];
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[98];
        accf[15] = fmaf(e, W[9], accf[15]);
      }
      { const float e = rf[99];
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[100];
        accf[15] = fmaf(e, W[6], accf[15]);
      }
      { const float e = rf[101];
        accf[15] = fmaf(e, W[15], accf[15]);
      }
      { const float e = rf[102];
        accf[16] = fmaf(e, W[0], accf[16]);
      }
      { const float e = rf[103];
        accf[16] = fmaf(e, W[3], accf[16]);
      }
      { const float e = rf[104];
        accf[16] = fmaf(e, W[9], accf[16]);
      }
      { const float e = rf[105];
        accf[16] = fmaf(e, W[2], accf[16]);
      }
      { const float e = rf[106];
        accf[16] = fmaf(e, W[8], accf[16]);
      }
      { const float e = rf[107];
        accf[16] = fmaf(e, W[18], accf[16]);
      }
      { const float e = rf[108];
        accf[17] = fmaf(e, W[0], accf[17]);
      }
      { const float e = rf[109];
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[110];
        accf[17] = fmaf(e, W[9], accf[17]);
      }
      { const float e = rf[111];
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      }
#pragma unroll
      for (int q = 0; q < 18; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 4] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 4] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 4] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 4] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 4] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 4] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|3), FP64: L = 6, record 112 values, 7 m-group(s)
__global__ __launch_bounds__(64, 3) void int3c2e_gen_213(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 7;
  const int pidx = (xs / 7) * 8 + (int)(blockIdx.x & 7);
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
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[0] = fma(0.6123724356957945, r, W[0]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[1] = fma(0.6123724356957945, r, W[1]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[4] = fma(0.6123724356957945, r, W[4]);
      }
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[10] = fma(0.6123724356957945, r, W[10]);
      }
      }
      {
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[2] = fma(0.6123724356957945, r, W[2]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[5] = fma(0.6123724356957945, r, W[5]);
      }
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[11] = fma(0.6123724356957945, r, W[11]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[0] = fma(-0.20412414523193148, r, W[0]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[1] = fma(-0.20412414523193148, r, W[1]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[7] = fma(0.6123724356957945, r, W[7]);
        W[4] = fma(-0.20412414523193148, r, W[4]);
      }
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[13] = fma(0.6123724356957945, r, W[13]);
        W[10] = fma(-0.20412414523193148, r, W[10]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[2] = fma(-0.20412414523193148, r, W[2]);
      }
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[5] = fma(-0.20412414523193148, r, W[5]);
      }
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[16] = fma(0.6123724356957945, r, W[16]);
        W[11] = fma(-0.20412414523193148, r, W[11]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[7] = fma(-0.20412414523193148, r, W[7]);
      }
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[13] = fma(-0.20412414523193148, r, W[13]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[16] = fma(-0.20412414523193148, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[3] = fma(0.6123724356957945, r, W[3]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[6] = fma(0.6123724356957945, r, W[6]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[12] = fma(0.6123724356957945, r, W[12]);
      }
      }
      {
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[8] = fma(0.6123724356957945, r, W[8]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[14] = fma(0.6123724356957945, r, W[14]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[3] = fma(-0.20412414523193148, r, W[3]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[6] = fma(-0.20412414523193148, r, W[6]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[17] = fma(0.6123724356957945, r, W[17]);
        W[12] = fma(-0.20412414523193148, r, W[12]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[8] = fma(-0.20412414523193148, r, W[8]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[14] = fma(-0.20412414523193148, r, W[14]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[17] = fma(-0.20412414523193148, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      }
      {
      const double y1_2_5 = py1 * z2_5;
      const double y1_2_4 = py1 * z2_4;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[9] = fma(0.6123724356957945, r, W[9]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[15] = fma(0.6123724356957945, r, W[15]);
      }
      }
      {
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[18] = fma(0.6123724356957945, r, W[18]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[9] = fma(-0.20412414523193148, r, W[9]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[15] = fma(-0.20412414523193148, r, W[15]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[18] = fma(-0.20412414523193148, r, W[18]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      }
      {
      const double y1_3_6 = py1 * z3_6;
      const double y1_3_5 = py1 * z3_5;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[19] = fma(0.6123724356957945, r, W[19]);
      }
      }
      {
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[19] = fma(-0.20412414523193148, r, W[19]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      }
      {
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 1: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      }
      {
      const double y1_1_3 = py1 * z1_3;
      { const double r = px1 * y1_1_3;
        W[0] = fma(1.0, r, W[0]);
      }
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[1] = fma(1.0, r, W[1]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[4] = fma(1.0, r, W[4]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[10] = fma(1.0, r, W[10]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[2] = fma(1.0, r, W[2]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[5] = fma(1.0, r, W[5]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[11] = fma(1.0, r, W[11]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[7] = fma(1.0, r, W[7]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[13] = fma(1.0, r, W[13]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[16] = fma(1.0, r, W[16]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[3] = fma(1.0, r, W[3]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[6] = fma(1.0, r, W[6]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[12] = fma(1.0, r, W[12]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[8] = fma(1.0, r, W[8]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[14] = fma(1.0, r, W[14]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[17] = fma(1.0, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[15] = fma(1.0, r, W[15]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[18] = fma(1.0, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      }
      {
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[19] = fma(1.0, r, W[19]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 2: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      }
      {
      const double y1_0_3 = py1 * z0_3;
      const double y1_0_2 = py1 * z0_2;
      { const double r = px2 * y1_0_3 + 2.0 * px0 * y1_0_2;
        W[0] = fma(-0.15811388300841894, r, W[0]);
      }
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[1] = fma(-0.15811388300841894, r, W[1]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[4] = fma(-0.15811388300841894, r, W[4]);
      }
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      {
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[5] = fma(-0.15811388300841894, r, W[5]);
      }
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[11] = fma(-0.15811388300841894, r, W[11]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[0] = fma(-0.15811388300841894, r, W[0]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[1] = fma(-0.15811388300841894, r, W[1]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[7] = fma(-0.15811388300841894, r, W[7]);
        W[4] = fma(-0.15811388300841894, r, W[4]);
      }
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[13] = fma(-0.15811388300841894, r, W[13]);
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[5] = fma(-0.15811388300841894, r, W[5]);
      }
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[16] = fma(-0.15811388300841894, r, W[16]);
        W[11] = fma(-0.15811388300841894, r, W[11]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[7] = fma(-0.15811388300841894, r, W[7]);
      }
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[13] = fma(-0.15811388300841894, r, W[13]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[16] = fma(-0.15811388300841894, r, W[16]);
      }
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[3] = fma(-0.15811388300841894, r, W[3]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[12] = fma(-0.15811388300841894, r, W[12]);
      }
      }
      {
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[8] = fma(-0.15811388300841894, r, W[8]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[14] = fma(-0.15811388300841894, r, W[14]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[3] = fma(-0.15811388300841894, r, W[3]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[17] = fma(-0.15811388300841894, r, W[17]);
        W[12] = fma(-0.15811388300841894, r, W[12]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[8] = fma(-0.15811388300841894, r, W[8]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[14] = fma(-0.15811388300841894, r, W[14]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[17] = fma(-0.15811388300841894, r, W[17]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[0] = fma(0.6324555320336758, r, W[0]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[1] = fma(0.6324555320336758, r, W[1]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[4] = fma(0.6324555320336758, r, W[4]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[15] = fma(-0.15811388300841894, r, W[15]);
        W[10] = fma(0.6324555320336758, r, W[10]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[2] = fma(0.6324555320336758, r, W[2]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[5] = fma(0.6324555320336758, r, W[5]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[11] = fma(0.6324555320336758, r, W[11]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[7] = fma(0.6324555320336758, r, W[7]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[15] = fma(-0.15811388300841894, r, W[15]);
        W[13] = fma(0.6324555320336758, r, W[13]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[16] = fma(0.6324555320336758, r, W[16]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[3] = fma(0.6324555320336758, r, W[3]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[6] = fma(0.6324555320336758, r, W[6]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[12] = fma(0.6324555320336758, r, W[12]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[8] = fma(0.6324555320336758, r, W[8]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[14] = fma(0.6324555320336758, r, W[14]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[17] = fma(0.6324555320336758, r, W[17]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[9] = fma(0.6324555320336758, r, W[9]);
      }
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[15] = fma(0.6324555320336758, r, W[15]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[18] = fma(0.6324555320336758, r, W[18]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      }
      {
      const double y1_5_6 = py1 * z5_6;
      { const double r = px0 * y1_5_6;
        W[19] = fma(0.6324555320336758, r, W[19]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 3: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[0] = fma(-0.38729833462074165, r, W[0]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[1] = fma(-0.38729833462074165, r, W[1]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[4] = fma(-0.38729833462074165, r, W[4]);
      }
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[10] = fma(-0.38729833462074165, r, W[10]);
      }
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[2] = fma(-0.38729833462074165, r, W[2]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[5] = fma(-0.38729833462074165, r, W[5]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[11] = fma(-0.38729833462074165, r, W[11]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[0] = fma(-0.38729833462074165, r, W[0]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[1] = fma(-0.38729833462074165, r, W[1]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[7] = fma(-0.38729833462074165, r, W[7]);
        W[4] = fma(-0.38729833462074165, r, W[4]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[13] = fma(-0.38729833462074165, r, W[13]);
        W[10] = fma(-0.38729833462074165, r, W[10]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[2] = fma(-0.38729833462074165, r, W[2]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[5] = fma(-0.38729833462074165, r, W[5]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[16] = fma(-0.38729833462074165, r, W[16]);
        W[11] = fma(-0.38729833462074165, r, W[11]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[7] = fma(-0.38729833462074165, r, W[7]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[13] = fma(-0.38729833462074165, r, W[13]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[16] = fma(-0.38729833462074165, r, W[16]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      const double y0_2_4 = py0 * z2_4;
      const double y0_2_3 = py0 * z2_3;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[3] = fma(-0.38729833462074165, r, W[3]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[6] = fma(-0.38729833462074165, r, W[6]);
      }
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[12] = fma(-0.38729833462074165, r, W[12]);
      }
      }
      {
      const double y1_2_5 = py1 * z2_5;
      const double y1_2_4 = py1 * z2_4;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[8] = fma(-0.38729833462074165, r, W[8]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[14] = fma(-0.38729833462074165, r, W[14]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[3] = fma(-0.38729833462074165, r, W[3]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[6] = fma(-0.38729833462074165, r, W[6]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[17] = fma(-0.38729833462074165, r, W[17]);
        W[12] = fma(-0.38729833462074165, r, W[12]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[8] = fma(-0.38729833462074165, r, W[8]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[14] = fma(-0.38729833462074165, r, W[14]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[17] = fma(-0.38729833462074165, r, W[17]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[0] = fma(0.2581988897471611, r, W[0]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[1] = fma(0.2581988897471611, r, W[1]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[9] = fma(-0.38729833462074165, r, W[9]);
        W[4] = fma(0.2581988897471611, r, W[4]);
      }
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[15] = fma(-0.38729833462074165, r, W[15]);
        W[10] = fma(0.2581988897471611, r, W[10]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[2] = fma(0.2581988897471611, r, W[2]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[5] = fma(0.2581988897471611, r, W[5]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[18] = fma(-0.38729833462074165, r, W[18]);
        W[11] = fma(0.2581988897471611, r, W[11]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[9] = fma(-0.38729833462074165, r, W[9]);
        W[7] = fma(0.2581988897471611, r, W[7]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[15] = fma(-0.38729833462074165, r, W[15]);
        W[13] = fma(0.2581988897471611, r, W[13]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[18] = fma(-0.38729833462074165, r, W[18]);
        W[16] = fma(0.2581988897471611, r, W[16]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[3] = fma(0.2581988897471611, r, W[3]);
      }
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[6] = fma(0.2581988897471611, r, W[6]);
      }
      const double y0_4_6 = py0 * z4_6;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[19] = fma(-0.38729833462074165, r, W[19]);
        W[12] = fma(0.2581988897471611, r, W[12]);
      }
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[8] = fma(0.2581988897471611, r, W[8]);
      }
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[14] = fma(0.2581988897471611, r, W[14]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[19] = fma(-0.38729833462074165, r, W[19]);
        W[17] = fma(0.2581988897471611, r, W[17]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      const double y0_5_5 = py0 * z5_5;
      { const double r = px0 * y0_5_5;
        W[9] = fma(0.2581988897471611, r, W[9]);
      }
      const double y0_5_6 = py0 * z5_6;
      { const double r = px1 * y0_5_6;
        W[15] = fma(0.2581988897471611, r, W[15]);
      }
      }
      {
      const double y1_5_6 = py1 * z5_6;
      { const double r = px0 * y1_5_6;
        W[18] = fma(0.2581988897471611, r, W[18]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      const double y0_6_6 = py0 * z6_6;
      { const double r = px0 * y0_6_6;
        W[19] = fma(0.2581988897471611, r, W[19]);
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
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 3] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 3] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 3] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 3] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 3] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 3] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 4: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      const double y0_0_3 = py0 * z0_3;
      const double y0_0_2 = py0 * z0_2;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[0] = fma(-0.15811388300841894, r, W[0]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[1] = fma(-0.15811388300841894, r, W[1]);
      }
      const double y0_0_5 = py0 * z0_5;
      { const double r = px5 * y0_0_5 + 20.0 * px3 * y0_0_4 + 60.0 * px1 * y0_0_3;
        W[4] = fma(-0.15811388300841894, r, W[4]);
      }
      const double y0_0_6 = py0 * z0_6;
      { const double r = px6 * y0_0_6 + 30.0 * px4 * y0_0_5 + 180.0 * px2 * y0_0_4 + 120.0 * px0 * y0_0_3;
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      {
      const double y1_0_4 = py1 * z0_4;
      const double y1_0_3 = py1 * z0_3;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[5] = fma(-0.15811388300841894, r, W[5]);
      }
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[11] = fma(-0.15811388300841894, r, W[11]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[0] = fma(-0.15811388300841894, r, W[0]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[1] = fma(-0.15811388300841894, r, W[1]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[7] = fma(-0.15811388300841894, r, W[7]);
        W[4] = fma(-0.15811388300841894, r, W[4]);
      }
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[13] = fma(-0.15811388300841894, r, W[13]);
        W[10] = fma(-0.15811388300841894, r, W[10]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[2] = fma(-0.15811388300841894, r, W[2]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[5] = fma(-0.15811388300841894, r, W[5]);
      }
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[16] = fma(-0.15811388300841894, r, W[16]);
        W[11] = fma(-0.15811388300841894, r, W[11]);
      }
      }
      {
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[7] = fma(-0.15811388300841894, r, W[7]);
      }
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[13] = fma(-0.15811388300841894, r, W[13]);
      }
      }
      {
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[16] = fma(-0.15811388300841894, r, W[16]);
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
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      const double y0_1_4 = py0 * z1_4;
      const double y0_1_3 = py0 * z1_3;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[3] = fma(-0.15811388300841894, r, W[3]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[12] = fma(-0.15811388300841894, r, W[12]);
      }
      }
      {
      const double y1_1_5 = py1 * z1_5;
      const double y1_1_4 = py1 * z1_4;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[8] = fma(-0.15811388300841894, r, W[8]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[14] = fma(-0.15811388300841894, r, W[14]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[3] = fma(-0.15811388300841894, r, W[3]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[6] = fma(-0.15811388300841894, r, W[6]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[17] = fma(-0.15811388300841894, r, W[17]);
        W[12] = fma(-0.15811388300841894, r, W[12]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[8] = fma(-0.15811388300841894, r, W[8]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[14] = fma(-0.15811388300841894, r, W[14]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[17] = fma(-0.15811388300841894, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[0] = fma(0.6324555320336758, r, W[0]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[1] = fma(0.6324555320336758, r, W[1]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[4] = fma(0.6324555320336758, r, W[4]);
      }
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[15] = fma(-0.15811388300841894, r, W[15]);
        W[10] = fma(0.6324555320336758, r, W[10]);
      }
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[2] = fma(0.6324555320336758, r, W[2]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[5] = fma(0.6324555320336758, r, W[5]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[11] = fma(0.6324555320336758, r, W[11]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[9] = fma(-0.15811388300841894, r, W[9]);
        W[7] = fma(0.6324555320336758, r, W[7]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[15] = fma(-0.15811388300841894, r, W[15]);
        W[13] = fma(0.6324555320336758, r, W[13]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[16] = fma(0.6324555320336758, r, W[16]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[3] = fma(0.6324555320336758, r, W[3]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[6] = fma(0.6324555320336758, r, W[6]);
      }
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[12] = fma(0.6324555320336758, r, W[12]);
      }
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[8] = fma(0.6324555320336758, r, W[8]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[14] = fma(0.6324555320336758, r, W[14]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[17] = fma(0.6324555320336758, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[9] = fma(0.6324555320336758, r, W[9]);
      }
      const double y0_4_6 = py0 * z4_6;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[15] = fma(0.6324555320336758, r, W[15]);
      }
      }
      {
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[18] = fma(0.6324555320336758, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      const double y0_5_6 = py0 * z5_6;
      { const double r = px1 * y0_5_6;
        W[19] = fma(0.6324555320336758, r, W[19]);
      }
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 4] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 4] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 4] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 4] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 4] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 4] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 5: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      {
      }
      }
      {
      const double z1_1 = pz1 * g1;
      const double z1_2 = pz1 * g2;
      const double z1_3 = pz1 * g3;
      const double z1_4 = pz1 * g4;
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      const double y0_1_3 = py0 * z1_3;
      const double y0_1_2 = py0 * z1_2;
      { const double r = px2 * y0_1_3 + 2.0 * px0 * y0_1_2;
        W[0] = fma(0.5, r, W[0]);
      }
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[1] = fma(0.5, r, W[1]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[4] = fma(0.5, r, W[4]);
      }
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[10] = fma(0.5, r, W[10]);
      }
      }
      {
      const double y1_1_4 = py1 * z1_4;
      const double y1_1_3 = py1 * z1_3;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[2] = fma(0.5, r, W[2]);
      }
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[5] = fma(0.5, r, W[5]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[11] = fma(0.5, r, W[11]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[0] = fma(-0.5, r, W[0]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[1] = fma(-0.5, r, W[1]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[7] = fma(0.5, r, W[7]);
        W[4] = fma(-0.5, r, W[4]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[13] = fma(0.5, r, W[13]);
        W[10] = fma(-0.5, r, W[10]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[2] = fma(-0.5, r, W[2]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[5] = fma(-0.5, r, W[5]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[16] = fma(0.5, r, W[16]);
        W[11] = fma(-0.5, r, W[11]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[7] = fma(-0.5, r, W[7]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[13] = fma(-0.5, r, W[13]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[16] = fma(-0.5, r, W[16]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      const double y0_2_4 = py0 * z2_4;
      const double y0_2_3 = py0 * z2_3;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[3] = fma(0.5, r, W[3]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[6] = fma(0.5, r, W[6]);
      }
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[12] = fma(0.5, r, W[12]);
      }
      }
      {
      const double y1_2_5 = py1 * z2_5;
      const double y1_2_4 = py1 * z2_4;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[8] = fma(0.5, r, W[8]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[14] = fma(0.5, r, W[14]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[3] = fma(-0.5, r, W[3]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[6] = fma(-0.5, r, W[6]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[17] = fma(0.5, r, W[17]);
        W[12] = fma(-0.5, r, W[12]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[8] = fma(-0.5, r, W[8]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[14] = fma(-0.5, r, W[14]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[17] = fma(-0.5, r, W[17]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      const double y0_3_5 = py0 * z3_5;
      const double y0_3_4 = py0 * z3_4;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[9] = fma(0.5, r, W[9]);
      }
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[15] = fma(0.5, r, W[15]);
      }
      }
      {
      const double y1_3_6 = py1 * z3_6;
      const double y1_3_5 = py1 * z3_5;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[18] = fma(0.5, r, W[18]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[9] = fma(-0.5, r, W[9]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[15] = fma(-0.5, r, W[15]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[18] = fma(-0.5, r, W[18]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      const double y0_4_6 = py0 * z4_6;
      const double y0_4_5 = py0 * z4_5;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[19] = fma(0.5, r, W[19]);
      }
      }
      {
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[19] = fma(-0.5, r, W[19]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 5] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 5] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 5] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 5] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 5] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 5] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 5] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 5] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 5] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 5] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 5] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 5] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 5] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 5] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  case 6: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const double g0 = (sc * F[0]);
      sc *= ma; const double g1 = (sc * F[1]);
      sc *= ma; const double g2 = (sc * F[2]);
      sc *= ma; const double g3 = (sc * F[3]);
      sc *= ma; const double g4 = (sc * F[4]);
      sc *= ma; const double g5 = (sc * F[5]);
      sc *= ma; const double g6 = (sc * F[6]);
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      double W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      {
      const double y0_0_3 = py0 * z0_3;
      const double y0_0_2 = py0 * z0_2;
      { const double r = px3 * y0_0_3 + 6.0 * px1 * y0_0_2;
        W[0] = fma(0.20412414523193148, r, W[0]);
      }
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[1] = fma(0.20412414523193148, r, W[1]);
      }
      const double y0_0_5 = py0 * z0_5;
      { const double r = px5 * y0_0_5 + 20.0 * px3 * y0_0_4 + 60.0 * px1 * y0_0_3;
        W[4] = fma(0.20412414523193148, r, W[4]);
      }
      const double y0_0_6 = py0 * z0_6;
      { const double r = px6 * y0_0_6 + 30.0 * px4 * y0_0_5 + 180.0 * px2 * y0_0_4 + 120.0 * px0 * y0_0_3;
        W[10] = fma(0.20412414523193148, r, W[10]);
      }
      }
      {
      const double y1_0_4 = py1 * z0_4;
      const double y1_0_3 = py1 * z0_3;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[2] = fma(0.20412414523193148, r, W[2]);
      }
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[5] = fma(0.20412414523193148, r, W[5]);
      }
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[11] = fma(0.20412414523193148, r, W[11]);
      }
      }
      {
      const double y2_0_3 = py2 * z0_3 + 2.0 * py0 * z0_2;
      { const double r = px1 * y2_0_3;
        W[0] = fma(-0.6123724356957945, r, W[0]);
      }
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[1] = fma(-0.6123724356957945, r, W[1]);
      }
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[7] = fma(0.20412414523193148, r, W[7]);
        W[4] = fma(-0.6123724356957945, r, W[4]);
      }
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[13] = fma(0.20412414523193148, r, W[13]);
        W[10] = fma(-0.6123724356957945, r, W[10]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[2] = fma(-0.6123724356957945, r, W[2]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[5] = fma(-0.6123724356957945, r, W[5]);
      }
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[16] = fma(0.20412414523193148, r, W[16]);
        W[11] = fma(-0.6123724356957945, r, W[11]);
      }
      }
      {
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[7] = fma(-0.6123724356957945, r, W[7]);
      }
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[13] = fma(-0.6123724356957945, r, W[13]);
      }
      }
      {
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[16] = fma(-0.6123724356957945, r, W[16]);
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
      const double z1_5 = pz1 * g5;
      const double z1_6 = pz1 * g6;
      {
      const double y0_1_4 = py0 * z1_4;
      const double y0_1_3 = py0 * z1_3;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[3] = fma(0.20412414523193148, r, W[3]);
      }
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[6] = fma(0.20412414523193148, r, W[6]);
      }
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[12] = fma(0.20412414523193148, r, W[12]);
      }
      }
      {
      const double y1_1_5 = py1 * z1_5;
      const double y1_1_4 = py1 * z1_4;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[8] = fma(0.20412414523193148, r, W[8]);
      }
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[14] = fma(0.20412414523193148, r, W[14]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[3] = fma(-0.6123724356957945, r, W[3]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[6] = fma(-0.6123724356957945, r, W[6]);
      }
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[17] = fma(0.20412414523193148, r, W[17]);
        W[12] = fma(-0.6123724356957945, r, W[12]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[8] = fma(-0.6123724356957945, r, W[8]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[14] = fma(-0.6123724356957945, r, W[14]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[17] = fma(-0.6123724356957945, r, W[17]);
      }
      }
      {
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      {
      const double y0_2_5 = py0 * z2_5;
      const double y0_2_4 = py0 * z2_4;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[9] = fma(0.20412414523193148, r, W[9]);
      }
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[15] = fma(0.20412414523193148, r, W[15]);
      }
      }
      {
      const double y1_2_6 = py1 * z2_6;
      const double y1_2_5 = py1 * z2_5;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[18] = fma(0.20412414523193148, r, W[18]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[9] = fma(-0.6123724356957945, r, W[9]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[15] = fma(-0.6123724356957945, r, W[15]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[18] = fma(-0.6123724356957945, r, W[18]);
      }
      }
      {
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      {
      const double y0_3_6 = py0 * z3_6;
      const double y0_3_5 = py0 * z3_5;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[19] = fma(0.20412414523193148, r, W[19]);
      }
      }
      {
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[19] = fma(-0.6123724356957945, r, W[19]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      {
      }
      {
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      }
      }
      { const double e = rec[4];
        acc[0] = fma(e, W[0], acc[0]);
      }
      { const double e = rec[5];
        acc[0] = fma(e, W[1], acc[0]);
      }
      { const double e = rec[6];
        acc[0] = fma(e, W[4], acc[0]);
      }
      { const double e = rec[7];
        acc[0] = fma(e, W[10], acc[0]);
      }
      { const double e = rec[8];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[14];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[15];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[16];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[20];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[21];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[22];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[23];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[24];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[26];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[27];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[28];
        acc[4] = fma(e, W[7], acc[4]);
      }
      { const double e = rec[29];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[30];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[31];
        acc[4] = fma(e, W[13], acc[4]);
      }
      { const double e = rec[32];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[33];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[34];
        acc[5] = fma(e, W[2], acc[5]);
      }
      { const double e = rec[35];
        acc[5] = fma(e, W[8], acc[5]);
      }
      { const double e = rec[36];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[37];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[38];
        acc[5] = fma(e, W[5], acc[5]);
      }
      { const double e = rec[39];
        acc[5] = fma(e, W[14], acc[5]);
      }
      { const double e = rec[40];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[41];
        acc[6] = fma(e, W[3], acc[6]);
      }
      { const double e = rec[42];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[43];
        acc[6] = fma(e, W[6], acc[6]);
      }
      { const double e = rec[44];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[45];
        acc[6] = fma(e, W[12], acc[6]);
      }
      { const double e = rec[46];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[47];
        acc[7] = fma(e, W[3], acc[7]);
      }
      { const double e = rec[48];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[49];
        acc[7] = fma(e, W[8], acc[7]);
      }
      { const double e = rec[50];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[51];
        acc[7] = fma(e, W[6], acc[7]);
      }
      { const double e = rec[52];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[53];
        acc[7] = fma(e, W[14], acc[7]);
      }
      { const double e = rec[54];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[55];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[56];
        acc[8] = fma(e, W[9], acc[8]);
      }
      { const double e = rec[57];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[58];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[59];
        acc[8] = fma(e, W[15], acc[8]);
      }
      { const double e = rec[60];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[61];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[62];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[63];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[64];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[65];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[66];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[67];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[68];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[69];
        acc[10] = fma(e, W[16], acc[10]);
      }
      { const double e = rec[70];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[71];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[72];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[73];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[74];
        acc[11] = fma(e, W[7], acc[11]);
      }
      { const double e = rec[75];
        acc[11] = fma(e, W[17], acc[11]);
      }
      { const double e = rec[76];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[77];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[78];
        acc[12] = fma(e, W[2], acc[12]);
      }
      { const double e = rec[79];
        acc[12] = fma(e, W[8], acc[12]);
      }
      { const double e = rec[80];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[81];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[82];
        acc[12] = fma(e, W[5], acc[12]);
      }
      { const double e = rec[83];
        acc[12] = fma(e, W[14], acc[12]);
      }
      { const double e = rec[84];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[85];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[86];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[87];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[88];
        acc[13] = fma(e, W[7], acc[13]);
      }
      { const double e = rec[89];
        acc[13] = fma(e, W[17], acc[13]);
      }
      { const double e = rec[90];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[91];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[92];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[93];
        acc[14] = fma(e, W[2], acc[14]);
      }
      { const double e = rec[94];
        acc[14] = fma(e, W[8], acc[14]);
      }
      { const double e = rec[95];
        acc[14] = fma(e, W[18], acc[14]);
      }
      { const double e = rec[96];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[97];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[98];
        acc[15] = fma(e, W[9], acc[15]);
      }
      { const double e = rec[99];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[100];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[101];
        acc[15] = fma(e, W[15], acc[15]);
      }
      { const double e = rec[102];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[103];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[104];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[105];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[106];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[107];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[108];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[109];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[110];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[111];
        acc[17] = fma(e, W[19], acc[17]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 6] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 6] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 6] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 6] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 6] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 6] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 6] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 6] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 6] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 6] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 6] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 6] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 6] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 6] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

// class (21|3), FP32: L = 6, record 112 values, 4 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen32_213(const double* __restrict__ recs, const float* __restrict__ recs32,
    const int64_t* __restrict__ recoff, const int* __restrict__ npp_v, DevShells ket, PairArgs pa, AuxArgs xa,
    const double* __restrict__ boys, double* __restrict__ out, int64_t ldo) {
  const int xs = (int)(blockIdx.x >> 3), grp = xs % 4;
  const int pidx = (xs / 4) * 8 + (int)(blockIdx.x & 7);
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      sc *= ma; const float g6 = (float)(sc * F[6]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float px6 = px5 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float py6 = py5 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      const float pz6 = pz5 * pz1;
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      const float z0_6 = pz0 * g6;
      {
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[0] = fmaf(0.6123724356957945f, r, W[0]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[2] = fmaf(0.6123724356957945f, r, W[2]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[8] = fmaf(0.6123724356957945f, r, W[8]);
      }
      const float y1_0_6 = py1 * z0_6;
      { const float r = px5 * y1_0_6 + 20.0f * px3 * y1_0_5 + 60.0f * px1 * y1_0_4;
        W[20] = fmaf(0.6123724356957945f, r, W[20]);
      }
      }
      {
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[4] = fmaf(0.6123724356957945f, r, W[4]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[10] = fmaf(0.6123724356957945f, r, W[10]);
      }
      const float y2_0_6 = py2 * z0_6 + 2.0f * py0 * z0_5;
      { const float r = px4 * y2_0_6 + 12.0f * px2 * y2_0_5 + 12.0f * px0 * y2_0_4;
        W[22] = fmaf(0.6123724356957945f, r, W[22]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[0] = fmaf(-0.20412414523193148f, r, W[0]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[2] = fmaf(-0.20412414523193148f, r, W[2]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[14] = fmaf(0.6123724356957945f, r, W[14]);
        W[8] = fmaf(-0.20412414523193148f, r, W[8]);
      }
      const float y3_0_6 = py3 * z0_6 + 6.0f * py1 * z0_5;
      { const float r = px3 * y3_0_6 + 6.0f * px1 * y3_0_5;
        W[26] = fmaf(0.6123724356957945f, r, W[26]);
        W[20] = fmaf(-0.20412414523193148f, r, W[20]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[4] = fmaf(-0.20412414523193148f, r, W[4]);
      }
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[10] = fmaf(-0.20412414523193148f, r, W[10]);
      }
      const float y4_0_6 = py4 * z0_6 + 12.0f * py2 * z0_5 + 12.0f * py0 * z0_4;
      { const float r = px2 * y4_0_6 + 2.0f * px0 * y4_0_5;
        W[32] = fmaf(0.6123724356957945f, r, W[32]);
        W[22] = fmaf(-0.20412414523193148f, r, W[22]);
      }
      }
      {
      const float y5_0_5 = py5 * z0_5 + 20.0f * py3 * z0_4 + 60.0f * py1 * z0_3;
      { const float r = px0 * y5_0_5;
        W[14] = fmaf(-0.20412414523193148f, r, W[14]);
      }
      const float y5_0_6 = py5 * z0_6 + 20.0f * py3 * z0_5 + 60.0f * py1 * z0_4;
      { const float r = px1 * y5_0_6;
        W[26] = fmaf(-0.20412414523193148f, r, W[26]);
      }
      }
      {
      const float y6_0_6 = py6 * z0_6 + 30.0f * py4 * z0_5 + 180.0f * py2 * z0_4 + 120.0f * py0 * z0_3;
      { const float r = px0 * y6_0_6;
        W[32] = fmaf(-0.20412414523193148f, r, W[32]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      const float z1_6 = pz1 * g6;
      {
      }
      {
      const float y1_1_3 = py1 * z1_3;
      { const float r = px1 * y1_1_3;
        W[1] = fmaf(1.0f, r, W[1]);
      }
      const float y1_1_4 = py1 * z1_4;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[6] = fmaf(0.6123724356957945f, r, W[6]);
        W[3] = fmaf(1.0f, r, W[3]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[12] = fmaf(0.6123724356957945f, r, W[12]);
        W[9] = fmaf(1.0f, r, W[9]);
      }
      const float y1_1_6 = py1 * z1_6;
      { const float r = px4 * y1_1_6 + 12.0f * px2 * y1_1_5 + 12.0f * px0 * y1_1_4;
        W[24] = fmaf(0.6123724356957945f, r, W[24]);
        W[21] = fmaf(1.0f, r, W[21]);
      }
      }
      {
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[5] = fmaf(1.0f, r, W[5]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[16] = fmaf(0.6123724356957945f, r, W[16]);
        W[11] = fmaf(1.0f, r, W[11]);
      }
      const float y2_1_6 = py2 * z1_6 + 2.0f * py0 * z1_5;
      { const float r = px3 * y2_1_6 + 6.0f * px1 * y2_1_5;
        W[28] = fmaf(0.6123724356957945f, r, W[28]);
        W[23] = fmaf(1.0f, r, W[23]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[6] = fmaf(-0.20412414523193148f, r, W[6]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[12] = fmaf(-0.20412414523193148f, r, W[12]);
        W[15] = fmaf(1.0f, r, W[15]);
      }
      const float y3_1_6 = py3 * z1_6 + 6.0f * py1 * z1_5;
      { const float r = px2 * y3_1_6 + 2.0f * px0 * y3_1_5;
        W[34] = fmaf(0.6123724356957945f, r, W[34]);
        W[24] = fmaf(-0.20412414523193148f, r, W[24]);
        W[27] = fmaf(1.0f, r, W[27]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[16] = fmaf(-0.20412414523193148f, r, W[16]);
      }
      const float y4_1_6 = py4 * z1_6 + 12.0f * py2 * z1_5 + 12.0f * py0 * z1_4;
      { const float r = px1 * y4_1_6;
        W[28] = fmaf(-0.20412414523193148f, r, W[28]);
        W[33] = fmaf(1.0f, r, W[33]);
      }
      }
      {
      const float y5_1_6 = py5 * z1_6 + 20.0f * py3 * z1_5 + 60.0f * py1 * z1_4;
      { const float r = px0 * y5_1_6;
        W[34] = fmaf(-0.20412414523193148f, r, W[34]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      const float z2_6 = pz2 * g6 + 2.0f * pz0 * g5;
      {
      }
      {
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[7] = fmaf(1.0f, r, W[7]);
      }
      const float y1_2_5 = py1 * z2_5;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[18] = fmaf(0.6123724356957945f, r, W[18]);
        W[13] = fmaf(1.0f, r, W[13]);
      }
      const float y1_2_6 = py1 * z2_6;
      { const float r = px3 * y1_2_6 + 6.0f * px1 * y1_2_5;
        W[30] = fmaf(0.6123724356957945f, r, W[30]);
        W[25] = fmaf(1.0f, r, W[25]);
      }
      }
      {
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[17] = fmaf(1.0f, r, W[17]);
      }
      const float y2_2_6 = py2 * z2_6 + 2.0f * py0 * z2_5;
      { const float r = px2 * y2_2_6 + 2.0f * px0 * y2_2_5;
        W[36] = fmaf(0.6123724356957945f, r, W[36]);
        W[29] = fmaf(1.0f, r, W[29]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[18] = fmaf(-0.20412414523193148f, r, W[18]);
      }
      const float y3_2_6 = py3 * z2_6 + 6.0f * py1 * z2_5;
      { const float r = px1 * y3_2_6;
        W[30] = fmaf(-0.20412414523193148f, r, W[30]);
        W[35] = fmaf(1.0f, r, W[35]);
      }
      }
      {
      const float y4_2_6 = py4 * z2_6 + 12.0f * py2 * z2_5 + 12.0f * py0 * z2_4;
      { const float r = px0 * y4_2_6;
        W[36] = fmaf(-0.20412414523193148f, r, W[36]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      const float z3_6 = pz3 * g6 + 6.0f * pz1 * g5;
      {
      }
      {
      const float y1_3_5 = py1 * z3_5;
      { const float r = px1 * y1_3_5;
        W[19] = fmaf(1.0f, r, W[19]);
      }
      const float y1_3_6 = py1 * z3_6;
      { const float r = px2 * y1_3_6 + 2.0f * px0 * y1_3_5;
        W[38] = fmaf(0.6123724356957945f, r, W[38]);
        W[31] = fmaf(1.0f, r, W[31]);
      }
      }
      {
      const float y2_3_6 = py2 * z3_6 + 2.0f * py0 * z3_5;
      { const float r = px1 * y2_3_6;
        W[37] = fmaf(1.0f, r, W[37]);
      }
      }
      {
      const float y3_3_6 = py3 * z3_6 + 6.0f * py1 * z3_5;
      { const float r = px0 * y3_3_6;
        W[38] = fmaf(-0.20412414523193148f, r, W[38]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      const float z4_6 = pz4 * g6 + 12.0f * pz2 * g5 + 12.0f * pz0 * g4;
      {
      }
      {
      const float y1_4_6 = py1 * z4_6;
      { const float r = px1 * y1_4_6;
        W[39] = fmaf(1.0f, r, W[39]);
      }
      }
      {
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      const float z5_6 = pz5 * g6 + 20.0f * pz3 * g5 + 60.0f * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const float z6_6 = pz6 * g6 + 30.0f * pz4 * g5 + 180.0f * pz2 * g4 + 120.0f * pz0 * g3;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 0] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 1] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 1] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 0] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 1] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 1] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 0] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 1] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 1] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 0] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 1: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      sc *= ma; const float g6 = (float)(sc * F[6]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float px6 = px5 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float py6 = py5 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      const float pz6 = pz5 * pz1;
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      const float z0_6 = pz0 * g6;
      {
      }
      {
      const float y1_0_3 = py1 * z0_3;
      const float y1_0_2 = py1 * z0_2;
      { const float r = px2 * y1_0_3 + 2.0f * px0 * y1_0_2;
        W[0] = fmaf(-0.15811388300841894f, r, W[0]);
      }
      const float y1_0_4 = py1 * z0_4;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[8] = fmaf(-0.15811388300841894f, r, W[8]);
      }
      const float y1_0_6 = py1 * z0_6;
      { const float r = px5 * y1_0_6 + 20.0f * px3 * y1_0_5 + 60.0f * px1 * y1_0_4;
        W[20] = fmaf(-0.15811388300841894f, r, W[20]);
      }
      }
      {
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[10] = fmaf(-0.15811388300841894f, r, W[10]);
      }
      const float y2_0_6 = py2 * z0_6 + 2.0f * py0 * z0_5;
      { const float r = px4 * y2_0_6 + 12.0f * px2 * y2_0_5 + 12.0f * px0 * y2_0_4;
        W[22] = fmaf(-0.15811388300841894f, r, W[22]);
      }
      }
      {
      const float y3_0_3 = py3 * z0_3 + 6.0f * py1 * z0_2;
      { const float r = px0 * y3_0_3;
        W[0] = fmaf(-0.15811388300841894f, r, W[0]);
      }
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[14] = fmaf(-0.15811388300841894f, r, W[14]);
        W[8] = fmaf(-0.15811388300841894f, r, W[8]);
      }
      const float y3_0_6 = py3 * z0_6 + 6.0f * py1 * z0_5;
      { const float r = px3 * y3_0_6 + 6.0f * px1 * y3_0_5;
        W[26] = fmaf(-0.15811388300841894f, r, W[26]);
        W[20] = fmaf(-0.15811388300841894f, r, W[20]);
      }
      }
      {
      const float y4_0_4 = py4 * z0_4 + 12.0f * py2 * z0_3 + 12.0f * py0 * z0_2;
      { const float r = px0 * y4_0_4;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
      }
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[10] = fmaf(-0.15811388300841894f, r, W[10]);
      }
      const float y4_0_6 = py4 * z0_6 + 12.0f * py2 * z0_5 + 12.0f * py0 * z0_4;
      { const float r = px2 * y4_0_6 + 2.0f * px0 * y4_0_5;
        W[32] = fmaf(-0.15811388300841894f, r, W[32]);
        W[22] = fmaf(-0.15811388300841894f, r, W[22]);
      }
      }
      {
      const float y5_0_5 = py5 * z0_5 + 20.0f * py3 * z0_4 + 60.0f * py1 * z0_3;
      { const float r = px0 * y5_0_5;
        W[14] = fmaf(-0.15811388300841894f, r, W[14]);
      }
      const float y5_0_6 = py5 * z0_6 + 20.0f * py3 * z0_5 + 60.0f * py1 * z0_4;
      { const float r = px1 * y5_0_6;
        W[26] = fmaf(-0.15811388300841894f, r, W[26]);
      }
      }
      {
      const float y6_0_6 = py6 * z0_6 + 30.0f * py4 * z0_5 + 180.0f * py2 * z0_4 + 120.0f * py0 * z0_3;
      { const float r = px0 * y6_0_6;
        W[32] = fmaf(-0.15811388300841894f, r, W[32]);
      }
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      const float z1_6 = pz1 * g6;
      {
      const float y0_1_3 = py0 * z1_3;
      const float y0_1_2 = py0 * z1_2;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[1] = fmaf(-0.38729833462074165f, r, W[1]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
      }
      const float y0_1_5 = py0 * z1_5;
      { const float r = px4 * y0_1_5 + 12.0f * px2 * y0_1_4 + 12.0f * px0 * y0_1_3;
        W[9] = fmaf(-0.38729833462074165f, r, W[9]);
      }
      const float y0_1_6 = py0 * z1_6;
      { const float r = px5 * y0_1_6 + 20.0f * px3 * y0_1_5 + 60.0f * px1 * y0_1_4;
        W[21] = fmaf(-0.38729833462074165f, r, W[21]);
      }
      }
      {
      const float y1_1_4 = py1 * z1_4;
      const float y1_1_3 = py1 * z1_3;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[6] = fmaf(-0.15811388300841894f, r, W[6]);
        W[5] = fmaf(-0.38729833462074165f, r, W[5]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[12] = fmaf(-0.15811388300841894f, r, W[12]);
        W[11] = fmaf(-0.38729833462074165f, r, W[11]);
      }
      const float y1_1_6 = py1 * z1_6;
      { const float r = px4 * y1_1_6 + 12.0f * px2 * y1_1_5 + 12.0f * px0 * y1_1_4;
        W[24] = fmaf(-0.15811388300841894f, r, W[24]);
        W[23] = fmaf(-0.38729833462074165f, r, W[23]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[1] = fmaf(-0.38729833462074165f, r, W[1]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[3] = fmaf(-0.38729833462074165f, r, W[3]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
        W[15] = fmaf(-0.38729833462074165f, r, W[15]);
        W[9] = fmaf(-0.38729833462074165f, r, W[9]);
      }
      const float y2_1_6 = py2 * z1_6 + 2.0f * py0 * z1_5;
      { const float r = px3 * y2_1_6 + 6.0f * px1 * y2_1_5;
        W[28] = fmaf(-0.15811388300841894f, r, W[28]);
        W[27] = fmaf(-0.38729833462074165f, r, W[27]);
        W[21] = fmaf(-0.38729833462074165f, r, W[21]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[6] = fmaf(-0.15811388300841894f, r, W[6]);
        W[5] = fmaf(-0.38729833462074165f, r, W[5]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[12] = fmaf(-0.15811388300841894f, r, W[12]);
        W[11] = fmaf(-0.38729833462074165f, r, W[11]);
      }
      const float y3_1_6 = py3 * z1_6 + 6.0f * py1 * z1_5;
      { const float r = px2 * y3_1_6 + 2.0f * px0 * y3_1_5;
        W[34] = fmaf(-0.15811388300841894f, r, W[34]);
        W[24] = fmaf(-0.15811388300841894f, r, W[24]);
        W[33] = fmaf(-0.38729833462074165f, r, W[33]);
        W[23] = fmaf(-0.38729833462074165f, r, W[23]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
        W[15] = fmaf(-0.38729833462074165f, r, W[15]);
      }
      const float y4_1_6 = py4 * z1_6 + 12.0f * py2 * z1_5 + 12.0f * py0 * z1_4;
      { const float r = px1 * y4_1_6;
        W[28] = fmaf(-0.15811388300841894f, r, W[28]);
        W[27] = fmaf(-0.38729833462074165f, r, W[27]);
      }
      }
      {
      const float y5_1_6 = py5 * z1_6 + 20.0f * py3 * z1_5 + 60.0f * py1 * z1_4;
      { const float r = px0 * y5_1_6;
        W[34] = fmaf(-0.15811388300841894f, r, W[34]);
        W[33] = fmaf(-0.38729833462074165f, r, W[33]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      const float z2_6 = pz2 * g6 + 2.0f * pz0 * g5;
      {
      const float y0_2_4 = py0 * z2_4;
      const float y0_2_3 = py0 * z2_3;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[7] = fmaf(-0.38729833462074165f, r, W[7]);
      }
      const float y0_2_5 = py0 * z2_5;
      { const float r = px3 * y0_2_5 + 6.0f * px1 * y0_2_4;
        W[13] = fmaf(-0.38729833462074165f, r, W[13]);
      }
      const float y0_2_6 = py0 * z2_6;
      { const float r = px4 * y0_2_6 + 12.0f * px2 * y0_2_5 + 12.0f * px0 * y0_2_4;
        W[25] = fmaf(-0.38729833462074165f, r, W[25]);
      }
      }
      {
      const float y1_2_3 = py1 * z2_3;
      { const float r = px0 * y1_2_3;
        W[0] = fmaf(0.6324555320336758f, r, W[0]);
      }
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[2] = fmaf(0.6324555320336758f, r, W[2]);
      }
      const float y1_2_5 = py1 * z2_5;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[8] = fmaf(0.6324555320336758f, r, W[8]);
        W[17] = fmaf(-0.38729833462074165f, r, W[17]);
      }
      const float y1_2_6 = py1 * z2_6;
      { const float r = px3 * y1_2_6 + 6.0f * px1 * y1_2_5;
        W[30] = fmaf(-0.15811388300841894f, r, W[30]);
        W[20] = fmaf(0.6324555320336758f, r, W[20]);
        W[29] = fmaf(-0.38729833462074165f, r, W[29]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[4] = fmaf(0.6324555320336758f, r, W[4]);
        W[7] = fmaf(-0.38729833462074165f, r, W[7]);
      }
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[10] = fmaf(0.6324555320336758f, r, W[10]);
        W[13] = fmaf(-0.38729833462074165f, r, W[13]);
      }
      const float y2_2_6 = py2 * z2_6 + 2.0f * py0 * z2_5;
      { const float r = px2 * y2_2_6 + 2.0f * px0 * y2_2_5;
        W[36] = fmaf(-0.15811388300841894f, r, W[36]);
        W[22] = fmaf(0.6324555320336758f, r, W[22]);
        W[35] = fmaf(-0.38729833462074165f, r, W[35]);
        W[25] = fmaf(-0.38729833462074165f, r, W[25]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[14] = fmaf(0.6324555320336758f, r, W[14]);
        W[17] = fmaf(-0.38729833462074165f, r, W[17]);
      }
      const float y3_2_6 = py3 * z2_6 + 6.0f * py1 * z2_5;
      { const float r = px1 * y3_2_6;
        W[30] = fmaf(-0.15811388300841894f, r, W[30]);
        W[26] = fmaf(0.6324555320336758f, r, W[26]);
        W[29] = fmaf(-0.38729833462074165f, r, W[29]);
      }
      }
      {
      const float y4_2_6 = py4 * z2_6 + 12.0f * py2 * z2_5 + 12.0f * py0 * z2_4;
      { const float r = px0 * y4_2_6;
        W[36] = fmaf(-0.15811388300841894f, r, W[36]);
        W[32] = fmaf(0.6324555320336758f, r, W[32]);
        W[35] = fmaf(-0.38729833462074165f, r, W[35]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      const float z3_6 = pz3 * g6 + 6.0f * pz1 * g5;
      {
      const float y0_3_3 = py0 * z3_3;
      { const float r = px0 * y0_3_3;
        W[1] = fmaf(0.2581988897471611f, r, W[1]);
      }
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[3] = fmaf(0.2581988897471611f, r, W[3]);
      }
      const float y0_3_5 = py0 * z3_5;
      { const float r = px2 * y0_3_5 + 2.0f * px0 * y0_3_4;
        W[19] = fmaf(-0.38729833462074165f, r, W[19]);
        W[9] = fmaf(0.2581988897471611f, r, W[9]);
      }
      const float y0_3_6 = py0 * z3_6;
      { const float r = px3 * y0_3_6 + 6.0f * px1 * y0_3_5;
        W[31] = fmaf(-0.38729833462074165f, r, W[31]);
        W[21] = fmaf(0.2581988897471611f, r, W[21]);
      }
      }
      {
      const float y1_3_4 = py1 * z3_4;
      { const float r = px0 * y1_3_4;
        W[6] = fmaf(0.6324555320336758f, r, W[6]);
        W[5] = fmaf(0.2581988897471611f, r, W[5]);
      }
      const float y1_3_5 = py1 * z3_5;
      { const float r = px1 * y1_3_5;
        W[12] = fmaf(0.6324555320336758f, r, W[12]);
        W[11] = fmaf(0.2581988897471611f, r, W[11]);
      }
      const float y1_3_6 = py1 * z3_6;
      { const float r = px2 * y1_3_6 + 2.0f * px0 * y1_3_5;
        W[38] = fmaf(-0.15811388300841894f, r, W[38]);
        W[24] = fmaf(0.6324555320336758f, r, W[24]);
        W[37] = fmaf(-0.38729833462074165f, r, W[37]);
        W[23] = fmaf(0.2581988897471611f, r, W[23]);
      }
      }
      {
      const float y2_3_5 = py2 * z3_5 + 2.0f * py0 * z3_4;
      { const float r = px0 * y2_3_5;
        W[16] = fmaf(0.6324555320336758f, r, W[16]);
        W[19] = fmaf(-0.38729833462074165f, r, W[19]);
        W[15] = fmaf(0.2581988897471611f, r, W[15]);
      }
      const float y2_3_6 = py2 * z3_6 + 2.0f * py0 * z3_5;
      { const float r = px1 * y2_3_6;
        W[28] = fmaf(0.6324555320336758f, r, W[28]);
        W[31] = fmaf(-0.38729833462074165f, r, W[31]);
        W[27] = fmaf(0.2581988897471611f, r, W[27]);
      }
      }
      {
      const float y3_3_6 = py3 * z3_6 + 6.0f * py1 * z3_5;
      { const float r = px0 * y3_3_6;
        W[38] = fmaf(-0.15811388300841894f, r, W[38]);
        W[34] = fmaf(0.6324555320336758f, r, W[34]);
        W[37] = fmaf(-0.38729833462074165f, r, W[37]);
        W[33] = fmaf(0.2581988897471611f, r, W[33]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      const float z4_6 = pz4 * g6 + 12.0f * pz2 * g5 + 12.0f * pz0 * g4;
      {
      const float y0_4_4 = py0 * z4_4;
      { const float r = px0 * y0_4_4;
        W[7] = fmaf(0.2581988897471611f, r, W[7]);
      }
      const float y0_4_5 = py0 * z4_5;
      { const float r = px1 * y0_4_5;
        W[13] = fmaf(0.2581988897471611f, r, W[13]);
      }
      const float y0_4_6 = py0 * z4_6;
      { const float r = px2 * y0_4_6 + 2.0f * px0 * y0_4_5;
        W[39] = fmaf(-0.38729833462074165f, r, W[39]);
        W[25] = fmaf(0.2581988897471611f, r, W[25]);
      }
      }
      {
      const float y1_4_5 = py1 * z4_5;
      { const float r = px0 * y1_4_5;
        W[18] = fmaf(0.6324555320336758f, r, W[18]);
        W[17] = fmaf(0.2581988897471611f, r, W[17]);
      }
      const float y1_4_6 = py1 * z4_6;
      { const float r = px1 * y1_4_6;
        W[30] = fmaf(0.6324555320336758f, r, W[30]);
        W[29] = fmaf(0.2581988897471611f, r, W[29]);
      }
      }
      {
      const float y2_4_6 = py2 * z4_6 + 2.0f * py0 * z4_5;
      { const float r = px0 * y2_4_6;
        W[36] = fmaf(0.6324555320336758f, r, W[36]);
        W[39] = fmaf(-0.38729833462074165f, r, W[39]);
        W[35] = fmaf(0.2581988897471611f, r, W[35]);
      }
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      const float z5_6 = pz5 * g6 + 20.0f * pz3 * g5 + 60.0f * pz1 * g4;
      {
      const float y0_5_5 = py0 * z5_5;
      { const float r = px0 * y0_5_5;
        W[19] = fmaf(0.2581988897471611f, r, W[19]);
      }
      const float y0_5_6 = py0 * z5_6;
      { const float r = px1 * y0_5_6;
        W[31] = fmaf(0.2581988897471611f, r, W[31]);
      }
      }
      {
      const float y1_5_6 = py1 * z5_6;
      { const float r = px0 * y1_5_6;
        W[38] = fmaf(0.6324555320336758f, r, W[38]);
        W[37] = fmaf(0.2581988897471611f, r, W[37]);
      }
      }
      }
      {
      const float z6_6 = pz6 * g6 + 30.0f * pz4 * g5 + 180.0f * pz2 * g4 + 120.0f * pz0 * g3;
      {
      const float y0_6_6 = py0 * z6_6;
      { const float r = px0 * y0_6_6;
        W[39] = fmaf(0.2581988897471611f, r, W[39]);
      }
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 2] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 2] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 3] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 3] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 2] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 3] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 3] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 2] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 3] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 2] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 3] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 2] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 2: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[36];
#pragma unroll
      for (int q = 0; q < 36; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      sc *= ma; const float g6 = (float)(sc * F[6]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float px6 = px5 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float py6 = py5 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      const float pz6 = pz5 * pz1;
      float W[40];
#pragma unroll
      for (int q = 0; q < 40; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      const float z0_6 = pz0 * g6;
      {
      const float y0_0_3 = py0 * z0_3;
      const float y0_0_2 = py0 * z0_2;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[0] = fmaf(-0.15811388300841894f, r, W[0]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y0_0_5 = py0 * z0_5;
      { const float r = px5 * y0_0_5 + 20.0f * px3 * y0_0_4 + 60.0f * px1 * y0_0_3;
        W[8] = fmaf(-0.15811388300841894f, r, W[8]);
      }
      const float y0_0_6 = py0 * z0_6;
      { const float r = px6 * y0_0_6 + 30.0f * px4 * y0_0_5 + 180.0f * px2 * y0_0_4 + 120.0f * px0 * y0_0_3;
        W[20] = fmaf(-0.15811388300841894f, r, W[20]);
      }
      }
      {
      const float y1_0_4 = py1 * z0_4;
      const float y1_0_3 = py1 * z0_3;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[10] = fmaf(-0.15811388300841894f, r, W[10]);
      }
      const float y1_0_6 = py1 * z0_6;
      { const float r = px5 * y1_0_6 + 20.0f * px3 * y1_0_5 + 60.0f * px1 * y1_0_4;
        W[22] = fmaf(-0.15811388300841894f, r, W[22]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[0] = fmaf(-0.15811388300841894f, r, W[0]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[2] = fmaf(-0.15811388300841894f, r, W[2]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[14] = fmaf(-0.15811388300841894f, r, W[14]);
        W[8] = fmaf(-0.15811388300841894f, r, W[8]);
      }
      const float y2_0_6 = py2 * z0_6 + 2.0f * py0 * z0_5;
      { const float r = px4 * y2_0_6 + 12.0f * px2 * y2_0_5 + 12.0f * px0 * y2_0_4;
        W[26] = fmaf(-0.15811388300841894f, r, W[26]);
        W[20] = fmaf(-0.15811388300841894f, r, W[20]);
      }
      }
      {
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[4] = fmaf(-0.15811388300841894f, r, W[4]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[10] = fmaf(-0.15811388300841894f, r, W[10]);
      }
      const float y3_0_6 = py3 * z0_6 + 6.0f * py1 * z0_5;
      { const float r = px3 * y3_0_6 + 6.0f * px1 * y3_0_5;
        W[32] = fmaf(-0.15811388300841894f, r, W[32]);
        W[22] = fmaf(-0.15811388300841894f, r, W[22]);
      }
      }
      {
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[14] = fmaf(-0.15811388300841894f, r, W[14]);
      }
      const float y4_0_6 = py4 * z0_6 + 12.0f * py2 * z0_5 + 12.0f * py0 * z0_4;
      { const float r = px2 * y4_0_6 + 2.0f * px0 * y4_0_5;
        W[26] = fmaf(-0.15811388300841894f, r, W[26]);
      }
      }
      {
      const float y5_0_6 = py5 * z0_6 + 20.0f * py3 * z0_5 + 60.0f * py1 * z0_4;
      { const float r = px1 * y5_0_6;
        W[32] = fmaf(-0.15811388300841894f, r, W[32]);
      }
      }
      {
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      const float z1_6 = pz1 * g6;
      {
      const float y0_1_3 = py0 * z1_3;
      const float y0_1_2 = py0 * z1_2;
      { const float r = px2 * y0_1_3 + 2.0f * px0 * y0_1_2;
        W[1] = fmaf(0.5f, r, W[1]);
      }
      const float y0_1_4 = py0 * z1_4;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[6] = fmaf(-0.15811388300841894f, r, W[6]);
        W[3] = fmaf(0.5f, r, W[3]);
      }
      const float y0_1_5 = py0 * z1_5;
      { const float r = px4 * y0_1_5 + 12.0f * px2 * y0_1_4 + 12.0f * px0 * y0_1_3;
        W[12] = fmaf(-0.15811388300841894f, r, W[12]);
        W[9] = fmaf(0.5f, r, W[9]);
      }
      const float y0_1_6 = py0 * z1_6;
      { const float r = px5 * y0_1_6 + 20.0f * px3 * y0_1_5 + 60.0f * px1 * y0_1_4;
        W[24] = fmaf(-0.15811388300841894f, r, W[24]);
        W[21] = fmaf(0.5f, r, W[21]);
      }
      }
      {
      const float y1_1_4 = py1 * z1_4;
      const float y1_1_3 = py1 * z1_3;
      { const float r = px2 * y1_1_4 + 2.0f * px0 * y1_1_3;
        W[5] = fmaf(0.5f, r, W[5]);
      }
      const float y1_1_5 = py1 * z1_5;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
        W[11] = fmaf(0.5f, r, W[11]);
      }
      const float y1_1_6 = py1 * z1_6;
      { const float r = px4 * y1_1_6 + 12.0f * px2 * y1_1_5 + 12.0f * px0 * y1_1_4;
        W[28] = fmaf(-0.15811388300841894f, r, W[28]);
        W[23] = fmaf(0.5f, r, W[23]);
      }
      }
      {
      const float y2_1_3 = py2 * z1_3 + 2.0f * py0 * z1_2;
      { const float r = px0 * y2_1_3;
        W[1] = fmaf(-0.5f, r, W[1]);
      }
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[6] = fmaf(-0.15811388300841894f, r, W[6]);
        W[3] = fmaf(-0.5f, r, W[3]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[12] = fmaf(-0.15811388300841894f, r, W[12]);
        W[15] = fmaf(0.5f, r, W[15]);
        W[9] = fmaf(-0.5f, r, W[9]);
      }
      const float y2_1_6 = py2 * z1_6 + 2.0f * py0 * z1_5;
      { const float r = px3 * y2_1_6 + 6.0f * px1 * y2_1_5;
        W[34] = fmaf(-0.15811388300841894f, r, W[34]);
        W[24] = fmaf(-0.15811388300841894f, r, W[24]);
        W[27] = fmaf(0.5f, r, W[27]);
        W[21] = fmaf(-0.5f, r, W[21]);
      }
      }
      {
      const float y3_1_4 = py3 * z1_4 + 6.0f * py1 * z1_3;
      { const float r = px0 * y3_1_4;
        W[5] = fmaf(-0.5f, r, W[5]);
      }
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[16] = fmaf(-0.15811388300841894f, r, W[16]);
        W[11] = fmaf(-0.5f, r, W[11]);
      }
      const float y3_1_6 = py3 * z1_6 + 6.0f * py1 * z1_5;
      { const float r = px2 * y3_1_6 + 2.0f * px0 * y3_1_5;
        W[28] = fmaf(-0.15811388300841894f, r, W[28]);
        W[33] = fmaf(0.5f, r, W[33]);
        W[23] = fmaf(-0.5f, r, W[23]);
      }
      }
      {
      const float y4_1_5 = py4 * z1_5 + 12.0f * py2 * z1_4 + 12.0f * py0 * z1_3;
      { const float r = px0 * y4_1_5;
        W[15] = fmaf(-0.5f, r, W[15]);
      }
      const float y4_1_6 = py4 * z1_6 + 12.0f * py2 * z1_5 + 12.0f * py0 * z1_4;
      { const float r = px1 * y4_1_6;
        W[34] = fmaf(-0.15811388300841894f, r, W[34]);
        W[27] = fmaf(-0.5f, r, W[27]);
      }
      }
      {
      const float y5_1_6 = py5 * z1_6 + 20.0f * py3 * z1_5 + 60.0f * py1 * z1_4;
      { const float r = px0 * y5_1_6;
        W[33] = fmaf(-0.5f, r, W[33]);
      }
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      const float z2_6 = pz2 * g6 + 2.0f * pz0 * g5;
      {
      const float y0_2_3 = py0 * z2_3;
      { const float r = px1 * y0_2_3;
        W[0] = fmaf(0.6324555320336758f, r, W[0]);
      }
      const float y0_2_4 = py0 * z2_4;
      { const float r = px2 * y0_2_4 + 2.0f * px0 * y0_2_3;
        W[2] = fmaf(0.6324555320336758f, r, W[2]);
        W[7] = fmaf(0.5f, r, W[7]);
      }
      const float y0_2_5 = py0 * z2_5;
      { const float r = px3 * y0_2_5 + 6.0f * px1 * y0_2_4;
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[8] = fmaf(0.6324555320336758f, r, W[8]);
        W[13] = fmaf(0.5f, r, W[13]);
      }
      const float y0_2_6 = py0 * z2_6;
      { const float r = px4 * y0_2_6 + 12.0f * px2 * y0_2_5 + 12.0f * px0 * y0_2_4;
        W[30] = fmaf(-0.15811388300841894f, r, W[30]);
        W[20] = fmaf(0.6324555320336758f, r, W[20]);
        W[25] = fmaf(0.5f, r, W[25]);
      }
      }
      {
      const float y1_2_4 = py1 * z2_4;
      { const float r = px1 * y1_2_4;
        W[4] = fmaf(0.6324555320336758f, r, W[4]);
      }
      const float y1_2_5 = py1 * z2_5;
      { const float r = px2 * y1_2_5 + 2.0f * px0 * y1_2_4;
        W[10] = fmaf(0.6324555320336758f, r, W[10]);
        W[17] = fmaf(0.5f, r, W[17]);
      }
      const float y1_2_6 = py1 * z2_6;
      { const float r = px3 * y1_2_6 + 6.0f * px1 * y1_2_5;
        W[36] = fmaf(-0.15811388300841894f, r, W[36]);
        W[22] = fmaf(0.6324555320336758f, r, W[22]);
        W[29] = fmaf(0.5f, r, W[29]);
      }
      }
      {
      const float y2_2_4 = py2 * z2_4 + 2.0f * py0 * z2_3;
      { const float r = px0 * y2_2_4;
        W[7] = fmaf(-0.5f, r, W[7]);
      }
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[18] = fmaf(-0.15811388300841894f, r, W[18]);
        W[14] = fmaf(0.6324555320336758f, r, W[14]);
        W[13] = fmaf(-0.5f, r, W[13]);
      }
      const float y2_2_6 = py2 * z2_6 + 2.0f * py0 * z2_5;
      { const float r = px2 * y2_2_6 + 2.0f * px0 * y2_2_5;
        W[30] = fmaf(-0.15811388300841894f, r, W[30]);
        W[26] = fmaf(0.6324555320336758f, r, W[26]);
        W[35] = fmaf(0.5f, r, W[35]);
        W[25] = fmaf(-0.5f, r, W[25]);
      }
      }
      {
      const float y3_2_5 = py3 * z2_5 + 6.0f * py1 * z2_4;
      { const float r = px0 * y3_2_5;
        W[17] = fmaf(-0.5f, r, W[17]);
      }
      const float y3_2_6 = py3 * z2_6 + 6.0f * py1 * z2_5;
      { const float r = px1 * y3_2_6;
        W[36] = fmaf(-0.15811388300841894f, r, W[36]);
        W[32] = fmaf(0.6324555320336758f, r, W[32]);
        W[29] = fmaf(-0.5f, r, W[29]);
      }
      }
      {
      const float y4_2_6 = py4 * z2_6 + 12.0f * py2 * z2_5 + 12.0f * py0 * z2_4;
      { const float r = px0 * y4_2_6;
        W[35] = fmaf(-0.5f, r, W[35]);
      }
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      const float z3_6 = pz3 * g6 + 6.0f * pz1 * g5;
      {
      const float y0_3_4 = py0 * z3_4;
      { const float r = px1 * y0_3_4;
        W[6] = fmaf(0.6324555320336758f, r, W[6]);
      }
      const float y0_3_5 = py0 * z3_5;
      { const float r = px2 * y0_3_5 + 2.0f * px0 * y0_3_4;
        W[12] = fmaf(0.6324555320336758f, r, W[12]);
        W[19] = fmaf(0.5f, r, W[19]);
      }
      const float y0_3_6 = py0 * z3_6;
      { const float r = px3 * y0_3_6 + 6.0f * px1 * y0_3_5;
        W[38] = fmaf(-0.15811388300841894f, r, W[38]);
        W[24] = fmaf(0.6324555320336758f, r, W[24]);
        W[31] = fmaf(0.5f, r, W[31]);
      }
      }
      {
      const float y1_3_5 = py1 * z3_5;
      { const float r = px1 * y1_3_5;
        W[16] = fmaf(0.6324555320336758f, r, W[16]);
      }
      const float y1_3_6 = py1 * z3_6;
      { const float r = px2 * y1_3_6 + 2.0f * px0 * y1_3_5;
        W[28] = fmaf(0.6324555320336758f, r, W[28]);
        W[37] = fmaf(0.5f, r, W[37]);
      }
      }
      {
      const float y2_3_5 = py2 * z3_5 + 2.0f * py0 * z3_4;
      { const float r = px0 * y2_3_5;
        W[19] = fmaf(-0.5f, r, W[19]);
      }
      const float y2_3_6 = py2 * z3_6 + 2.0f * py0 * z3_5;
      { const float r = px1 * y2_3_6;
        W[38] = fmaf(-0.15811388300841894f, r, W[38]);
        W[34] = fmaf(0.6324555320336758f, r, W[34]);
        W[31] = fmaf(-0.5f, r, W[31]);
      }
      }
      {
      const float y3_3_6 = py3 * z3_6 + 6.0f * py1 * z3_5;
      { const float r = px0 * y3_3_6;
        W[37] = fmaf(-0.5f, r, W[37]);
      }
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      const float z4_6 = pz4 * g6 + 12.0f * pz2 * g5 + 12.0f * pz0 * g4;
      {
      const float y0_4_5 = py0 * z4_5;
      { const float r = px1 * y0_4_5;
        W[18] = fmaf(0.6324555320336758f, r, W[18]);
      }
      const float y0_4_6 = py0 * z4_6;
      { const float r = px2 * y0_4_6 + 2.0f * px0 * y0_4_5;
        W[30] = fmaf(0.6324555320336758f, r, W[30]);
        W[39] = fmaf(0.5f, r, W[39]);
      }
      }
      {
      const float y1_4_6 = py1 * z4_6;
      { const float r = px1 * y1_4_6;
        W[36] = fmaf(0.6324555320336758f, r, W[36]);
      }
      }
      {
      const float y2_4_6 = py2 * z4_6 + 2.0f * py0 * z4_5;
      { const float r = px0 * y2_4_6;
        W[39] = fmaf(-0.5f, r, W[39]);
      }
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      const float z5_6 = pz5 * g6 + 20.0f * pz3 * g5 + 60.0f * pz1 * g4;
      {
      const float y0_5_6 = py0 * z5_6;
      { const float r = px1 * y0_5_6;
        W[38] = fmaf(0.6324555320336758f, r, W[38]);
      }
      }
      {
      }
      }
      {
      const float z6_6 = pz6 * g6 + 30.0f * pz4 * g5 + 180.0f * pz2 * g4 + 120.0f * pz0 * g3;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[2], accf[0]);
        accf[1] = fmaf(e, W[3], accf[1]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[8], accf[0]);
        accf[1] = fmaf(e, W[9], accf[1]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[20], accf[0]);
        accf[1] = fmaf(e, W[21], accf[1]);
      }
      { const float e = rf[8];
        accf[2] = fmaf(e, W[0], accf[2]);
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[9];
        accf[2] = fmaf(e, W[4], accf[2]);
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[10];
        accf[2] = fmaf(e, W[2], accf[2]);
        accf[3] = fmaf(e, W[3], accf[3]);
      }
      { const float e = rf[11];
        accf[2] = fmaf(e, W[10], accf[2]);
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[12];
        accf[2] = fmaf(e, W[8], accf[2]);
        accf[3] = fmaf(e, W[9], accf[3]);
      }
      { const float e = rf[13];
        accf[2] = fmaf(e, W[22], accf[2]);
        accf[3] = fmaf(e, W[23], accf[3]);
      }
      { const float e = rf[14];
        accf[4] = fmaf(e, W[0], accf[4]);
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[15];
        accf[4] = fmaf(e, W[6], accf[4]);
        accf[5] = fmaf(e, W[7], accf[5]);
      }
      { const float e = rf[16];
        accf[4] = fmaf(e, W[2], accf[4]);
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[17];
        accf[4] = fmaf(e, W[12], accf[4]);
        accf[5] = fmaf(e, W[13], accf[5]);
      }
      { const float e = rf[18];
        accf[4] = fmaf(e, W[8], accf[4]);
        accf[5] = fmaf(e, W[9], accf[5]);
      }
      { const float e = rf[19];
        accf[4] = fmaf(e, W[24], accf[4]);
        accf[5] = fmaf(e, W[25], accf[5]);
      }
      { const float e = rf[20];
        accf[6] = fmaf(e, W[0], accf[6]);
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[21];
        accf[6] = fmaf(e, W[4], accf[6]);
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[22];
        accf[6] = fmaf(e, W[2], accf[6]);
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[23];
        accf[6] = fmaf(e, W[10], accf[6]);
        accf[7] = fmaf(e, W[11], accf[7]);
      }
      { const float e = rf[24];
        accf[6] = fmaf(e, W[8], accf[6]);
        accf[7] = fmaf(e, W[9], accf[7]);
      }
      { const float e = rf[25];
        accf[6] = fmaf(e, W[22], accf[6]);
        accf[7] = fmaf(e, W[23], accf[7]);
      }
      { const float e = rf[26];
        accf[8] = fmaf(e, W[0], accf[8]);
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[27];
        accf[8] = fmaf(e, W[4], accf[8]);
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[28];
        accf[8] = fmaf(e, W[14], accf[8]);
        accf[9] = fmaf(e, W[15], accf[9]);
      }
      { const float e = rf[29];
        accf[8] = fmaf(e, W[2], accf[8]);
        accf[9] = fmaf(e, W[3], accf[9]);
      }
      { const float e = rf[30];
        accf[8] = fmaf(e, W[10], accf[8]);
        accf[9] = fmaf(e, W[11], accf[9]);
      }
      { const float e = rf[31];
        accf[8] = fmaf(e, W[26], accf[8]);
        accf[9] = fmaf(e, W[27], accf[9]);
      }
      { const float e = rf[32];
        accf[10] = fmaf(e, W[0], accf[10]);
        accf[11] = fmaf(e, W[1], accf[11]);
      }
      { const float e = rf[33];
        accf[10] = fmaf(e, W[6], accf[10]);
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[34];
        accf[10] = fmaf(e, W[4], accf[10]);
        accf[11] = fmaf(e, W[5], accf[11]);
      }
      { const float e = rf[35];
        accf[10] = fmaf(e, W[16], accf[10]);
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[36];
        accf[10] = fmaf(e, W[2], accf[10]);
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[37];
        accf[10] = fmaf(e, W[12], accf[10]);
        accf[11] = fmaf(e, W[13], accf[11]);
      }
      { const float e = rf[38];
        accf[10] = fmaf(e, W[10], accf[10]);
        accf[11] = fmaf(e, W[11], accf[11]);
      }
      { const float e = rf[39];
        accf[10] = fmaf(e, W[28], accf[10]);
        accf[11] = fmaf(e, W[29], accf[11]);
      }
      { const float e = rf[40];
        accf[12] = fmaf(e, W[0], accf[12]);
        accf[13] = fmaf(e, W[1], accf[13]);
      }
      { const float e = rf[41];
        accf[12] = fmaf(e, W[6], accf[12]);
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[42];
        accf[12] = fmaf(e, W[2], accf[12]);
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[43];
        accf[12] = fmaf(e, W[12], accf[12]);
        accf[13] = fmaf(e, W[13], accf[13]);
      }
      { const float e = rf[44];
        accf[12] = fmaf(e, W[8], accf[12]);
        accf[13] = fmaf(e, W[9], accf[13]);
      }
      { const float e = rf[45];
        accf[12] = fmaf(e, W[24], accf[12]);
        accf[13] = fmaf(e, W[25], accf[13]);
      }
      { const float e = rf[46];
        accf[14] = fmaf(e, W[0], accf[14]);
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[47];
        accf[14] = fmaf(e, W[6], accf[14]);
        accf[15] = fmaf(e, W[7], accf[15]);
      }
      { const float e = rf[48];
        accf[14] = fmaf(e, W[4], accf[14]);
        accf[15] = fmaf(e, W[5], accf[15]);
      }
      { const float e = rf[49];
        accf[14] = fmaf(e, W[16], accf[14]);
        accf[15] = fmaf(e, W[17], accf[15]);
      }
      { const float e = rf[50];
        accf[14] = fmaf(e, W[2], accf[14]);
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[51];
        accf[14] = fmaf(e, W[12], accf[14]);
        accf[15] = fmaf(e, W[13], accf[15]);
      }
      { const float e = rf[52];
        accf[14] = fmaf(e, W[10], accf[14]);
        accf[15] = fmaf(e, W[11], accf[15]);
      }
      { const float e = rf[53];
        accf[14] = fmaf(e, W[28], accf[14]);
        accf[15] = fmaf(e, W[29], accf[15]);
      }
      { const float e = rf[54];
        accf[16] = fmaf(e, W[0], accf[16]);
        accf[17] = fmaf(e, W[1], accf[17]);
      }
      { const float e = rf[55];
        accf[16] = fmaf(e, W[6], accf[16]);
        accf[17] = fmaf(e, W[7], accf[17]);
      }
      { const float e = rf[56];
        accf[16] = fmaf(e, W[18], accf[16]);
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      { const float e = rf[57];
        accf[16] = fmaf(e, W[2], accf[16]);
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[58];
        accf[16] = fmaf(e, W[12], accf[16]);
        accf[17] = fmaf(e, W[13], accf[17]);
      }
      { const float e = rf[59];
        accf[16] = fmaf(e, W[30], accf[16]);
        accf[17] = fmaf(e, W[31], accf[17]);
      }
      { const float e = rf[60];
        accf[18] = fmaf(e, W[0], accf[18]);
        accf[19] = fmaf(e, W[1], accf[19]);
      }
      { const float e = rf[61];
        accf[18] = fmaf(e, W[4], accf[18]);
        accf[19] = fmaf(e, W[5], accf[19]);
      }
      { const float e = rf[62];
        accf[18] = fmaf(e, W[14], accf[18]);
        accf[19] = fmaf(e, W[15], accf[19]);
      }
      { const float e = rf[63];
        accf[18] = fmaf(e, W[2], accf[18]);
        accf[19] = fmaf(e, W[3], accf[19]);
      }
      { const float e = rf[64];
        accf[18] = fmaf(e, W[10], accf[18]);
        accf[19] = fmaf(e, W[11], accf[19]);
      }
      { const float e = rf[65];
        accf[18] = fmaf(e, W[26], accf[18]);
        accf[19] = fmaf(e, W[27], accf[19]);
      }
      { const float e = rf[66];
        accf[20] = fmaf(e, W[0], accf[20]);
        accf[21] = fmaf(e, W[1], accf[21]);
      }
      { const float e = rf[67];
        accf[20] = fmaf(e, W[4], accf[20]);
        accf[21] = fmaf(e, W[5], accf[21]);
      }
      { const float e = rf[68];
        accf[20] = fmaf(e, W[14], accf[20]);
        accf[21] = fmaf(e, W[15], accf[21]);
      }
      { const float e = rf[69];
        accf[20] = fmaf(e, W[32], accf[20]);
        accf[21] = fmaf(e, W[33], accf[21]);
      }
      { const float e = rf[70];
        accf[22] = fmaf(e, W[0], accf[22]);
        accf[23] = fmaf(e, W[1], accf[23]);
      }
      { const float e = rf[71];
        accf[22] = fmaf(e, W[6], accf[22]);
        accf[23] = fmaf(e, W[7], accf[23]);
      }
      { const float e = rf[72];
        accf[22] = fmaf(e, W[4], accf[22]);
        accf[23] = fmaf(e, W[5], accf[23]);
      }
      { const float e = rf[73];
        accf[22] = fmaf(e, W[16], accf[22]);
        accf[23] = fmaf(e, W[17], accf[23]);
      }
      { const float e = rf[74];
        accf[22] = fmaf(e, W[14], accf[22]);
        accf[23] = fmaf(e, W[15], accf[23]);
      }
      { const float e = rf[75];
        accf[22] = fmaf(e, W[34], accf[22]);
        accf[23] = fmaf(e, W[35], accf[23]);
      }
      { const float e = rf[76];
        accf[24] = fmaf(e, W[0], accf[24]);
        accf[25] = fmaf(e, W[1], accf[25]);
      }
      { const float e = rf[77];
        accf[24] = fmaf(e, W[6], accf[24]);
        accf[25] = fmaf(e, W[7], accf[25]);
      }
      { const float e = rf[78];
        accf[24] = fmaf(e, W[4], accf[24]);
        accf[25] = fmaf(e, W[5], accf[25]);
      }
      { const float e = rf[79];
        accf[24] = fmaf(e, W[16], accf[24]);
        accf[25] = fmaf(e, W[17], accf[25]);
      }
      { const float e = rf[80];
        accf[24] = fmaf(e, W[2], accf[24]);
        accf[25] = fmaf(e, W[3], accf[25]);
      }
      { const float e = rf[81];
        accf[24] = fmaf(e, W[12], accf[24]);
        accf[25] = fmaf(e, W[13], accf[25]);
      }
      { const float e = rf[82];
        accf[24] = fmaf(e, W[10], accf[24]);
        accf[25] = fmaf(e, W[11], accf[25]);
      }
      { const float e = rf[83];
        accf[24] = fmaf(e, W[28], accf[24]);
        accf[25] = fmaf(e, W[29], accf[25]);
      }
      { const float e = rf[84];
        accf[26] = fmaf(e, W[0], accf[26]);
        accf[27] = fmaf(e, W[1], accf[27]);
      }
      { const float e = rf[85];
        accf[26] = fmaf(e, W[6], accf[26]);
        accf[27] = fmaf(e, W[7], accf[27]);
      }
      { const float e = rf[86];
        accf[26] = fmaf(e, W[4], accf[26]);
        accf[27] = fmaf(e, W[5], accf[27]);
      }
      { const float e = rf[87];
        accf[26] = fmaf(e, W[16], accf[26]);
        accf[27] = fmaf(e, W[17], accf[27]);
      }
      { const float e = rf[88];
        accf[26] = fmaf(e, W[14], accf[26]);
        accf[27] = fmaf(e, W[15], accf[27]);
      }
      { const float e = rf[89];
        accf[26] = fmaf(e, W[34], accf[26]);
        accf[27] = fmaf(e, W[35], accf[27]);
      }
      { const float e = rf[90];
        accf[28] = fmaf(e, W[0], accf[28]);
        accf[29] = fmaf(e, W[1], accf[29]);
      }
      { const float e = rf[91];
        accf[28] = fmaf(e, W[6], accf[28]);
        accf[29] = fmaf(e, W[7], accf[29]);
      }
      { const float e = rf[92];
        accf[28] = fmaf(e, W[18], accf[28]);
        accf[29] = fmaf(e, W[19], accf[29]);
      }
      { const float e = rf[93];
        accf[28] = fmaf(e, W[4], accf[28]);
        accf[29] = fmaf(e, W[5], accf[29]);
      }
      { const float e = rf[94];
        accf[28] = fmaf(e, W[16], accf[28]);
        accf[29] = fmaf(e, W[17], accf[29]);
      }
      { const float e = rf[95];
        accf[28] = fmaf(e, W[36], accf[28]);
        accf[29] = fmaf(e, W[37], accf[29]);
      }
      { const float e = rf[96];
        accf[30] = fmaf(e, W[0], accf[30]);
        accf[31] = fmaf(e, W[1], accf[31]);
      }
      { const float e = rf[97];
        accf[30] = fmaf(e, W[6], accf[30]);
        accf[31] = fmaf(e, W[7], accf[31]);
      }
      { const float e = rf[98];
        accf[30] = fmaf(e, W[18], accf[30]);
        accf[31] = fmaf(e, W[19], accf[31]);
      }
      { const float e = rf[99];
        accf[30] = fmaf(e, W[2], accf[30]);
        accf[31] = fmaf(e, W[3], accf[31]);
      }
      { const float e = rf[100];
        accf[30] = fmaf(e, W[12], accf[30]);
        accf[31] = fmaf(e, W[13], accf[31]);
      }
      { const float e = rf[101];
        accf[30] = fmaf(e, W[30], accf[30]);
        accf[31] = fmaf(e, W[31], accf[31]);
      }
      { const float e = rf[102];
        accf[32] = fmaf(e, W[0], accf[32]);
        accf[33] = fmaf(e, W[1], accf[33]);
      }
      { const float e = rf[103];
        accf[32] = fmaf(e, W[6], accf[32]);
        accf[33] = fmaf(e, W[7], accf[33]);
      }
      { const float e = rf[104];
        accf[32] = fmaf(e, W[18], accf[32]);
        accf[33] = fmaf(e, W[19], accf[33]);
      }
      { const float e = rf[105];
        accf[32] = fmaf(e, W[4], accf[32]);
        accf[33] = fmaf(e, W[5], accf[33]);
      }
      { const float e = rf[106];
        accf[32] = fmaf(e, W[16], accf[32]);
        accf[33] = fmaf(e, W[17], accf[33]);
      }
      { const float e = rf[107];
        accf[32] = fmaf(e, W[36], accf[32]);
        accf[33] = fmaf(e, W[37], accf[33]);
      }
      { const float e = rf[108];
        accf[34] = fmaf(e, W[0], accf[34]);
        accf[35] = fmaf(e, W[1], accf[35]);
      }
      { const float e = rf[109];
        accf[34] = fmaf(e, W[6], accf[34]);
        accf[35] = fmaf(e, W[7], accf[35]);
      }
      { const float e = rf[110];
        accf[34] = fmaf(e, W[18], accf[34]);
        accf[35] = fmaf(e, W[19], accf[35]);
      }
      { const float e = rf[111];
        accf[34] = fmaf(e, W[38], accf[34]);
        accf[35] = fmaf(e, W[39], accf[35]);
      }
      }
#pragma unroll
      for (int q = 0; q < 36; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[9];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[10];
    out[(row0 + 1) * ldo + col + 5] = 1.0 * acc[11];
    out[(row0 + 2) * ldo + col + 4] = 1.0 * acc[6];
    out[(row0 + 2) * ldo + col + 5] = 1.0 * acc[7];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[26];
    out[(row0 + 3) * ldo + col + 5] = 1.0 * acc[27];
    out[(row0 + 4) * ldo + col + 4] = 1.0 * acc[28];
    out[(row0 + 4) * ldo + col + 5] = 1.0 * acc[29];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[24];
    out[(row0 + 5) * ldo + col + 5] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 6) * ldo + col + 5] = -0.2886751345948129 * acc[3] + -0.2886751345948129 * acc[21] + 0.5773502691896258 * acc[33];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 7) * ldo + col + 5] = -0.2886751345948129 * acc[5] + -0.2886751345948129 * acc[23] + 0.5773502691896258 * acc[35];
    out[(row0 + 8) * ldo + col + 4] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[18] + 0.5773502691896258 * acc[30];
    out[(row0 + 8) * ldo + col + 5] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 9) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 9) * ldo + col + 5] = 1.0 * acc[15];
    out[(row0 + 10) * ldo + col + 4] = 1.0 * acc[16];
    out[(row0 + 10) * ldo + col + 5] = 1.0 * acc[17];
    out[(row0 + 11) * ldo + col + 4] = 1.0 * acc[12];
    out[(row0 + 11) * ldo + col + 5] = 1.0 * acc[13];
    out[(row0 + 12) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 12) * ldo + col + 5] = 0.5 * acc[3] + -0.5 * acc[21];
    out[(row0 + 13) * ldo + col + 4] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 13) * ldo + col + 5] = 0.5 * acc[5] + -0.5 * acc[23];
    out[(row0 + 14) * ldo + col + 4] = 0.5 * acc[0] + -0.5 * acc[18];
    out[(row0 + 14) * ldo + col + 5] = 0.5 * acc[1] + -0.5 * acc[19];
    }
    break;
  }
  case 3: {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 112;
      const double p = rec[0];
      const float* __restrict__ rf = recp32 + (size_t)ij * 112;
      const double PCxd = rec[1] - Cx, PCyd = rec[2] - Cy, PCzd = rec[3] - Cz;
      const float PCx = (float)PCxd, PCy = (float)PCyd, PCz = (float)PCzd;
      float accf[18];
#pragma unroll
      for (int q = 0; q < 18; q++) accf[q] = 0.0f;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCxd * PCxd + PCyd * PCyd + PCzd * PCzd);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[7];
      boys_array<6>(T, boys, F);
      const double ma = -alpha;
      const float g0 = (float)(sc * F[0]);
      sc *= ma; const float g1 = (float)(sc * F[1]);
      sc *= ma; const float g2 = (float)(sc * F[2]);
      sc *= ma; const float g3 = (float)(sc * F[3]);
      sc *= ma; const float g4 = (float)(sc * F[4]);
      sc *= ma; const float g5 = (float)(sc * F[5]);
      sc *= ma; const float g6 = (float)(sc * F[6]);
      const float px0 = 1.0f, px1 = 2.0f * PCx;
      const float px2 = px1 * px1;
      const float px3 = px2 * px1;
      const float px4 = px3 * px1;
      const float px5 = px4 * px1;
      const float px6 = px5 * px1;
      const float py0 = 1.0f, py1 = 2.0f * PCy;
      const float py2 = py1 * py1;
      const float py3 = py2 * py1;
      const float py4 = py3 * py1;
      const float py5 = py4 * py1;
      const float py6 = py5 * py1;
      const float pz0 = 1.0f, pz1 = 2.0f * PCz;
      const float pz2 = pz1 * pz1;
      const float pz3 = pz2 * pz1;
      const float pz4 = pz3 * pz1;
      const float pz5 = pz4 * pz1;
      const float pz6 = pz5 * pz1;
      float W[20];
#pragma unroll
      for (int q = 0; q < 20; q++) W[q] = 0.0f;
      {
      const float z0_0 = pz0 * g0;
      const float z0_1 = pz0 * g1;
      const float z0_2 = pz0 * g2;
      const float z0_3 = pz0 * g3;
      const float z0_4 = pz0 * g4;
      const float z0_5 = pz0 * g5;
      const float z0_6 = pz0 * g6;
      {
      const float y0_0_3 = py0 * z0_3;
      const float y0_0_2 = py0 * z0_2;
      { const float r = px3 * y0_0_3 + 6.0f * px1 * y0_0_2;
        W[0] = fmaf(0.20412414523193148f, r, W[0]);
      }
      const float y0_0_4 = py0 * z0_4;
      { const float r = px4 * y0_0_4 + 12.0f * px2 * y0_0_3 + 12.0f * px0 * y0_0_2;
        W[1] = fmaf(0.20412414523193148f, r, W[1]);
      }
      const float y0_0_5 = py0 * z0_5;
      { const float r = px5 * y0_0_5 + 20.0f * px3 * y0_0_4 + 60.0f * px1 * y0_0_3;
        W[4] = fmaf(0.20412414523193148f, r, W[4]);
      }
      const float y0_0_6 = py0 * z0_6;
      { const float r = px6 * y0_0_6 + 30.0f * px4 * y0_0_5 + 180.0f * px2 * y0_0_4 + 120.0f * px0 * y0_0_3;
        W[10] = fmaf(0.20412414523193148f, r, W[10]);
      }
      }
      {
      const float y1_0_4 = py1 * z0_4;
      const float y1_0_3 = py1 * z0_3;
      { const float r = px3 * y1_0_4 + 6.0f * px1 * y1_0_3;
        W[2] = fmaf(0.20412414523193148f, r, W[2]);
      }
      const float y1_0_5 = py1 * z0_5;
      { const float r = px4 * y1_0_5 + 12.0f * px2 * y1_0_4 + 12.0f * px0 * y1_0_3;
        W[5] = fmaf(0.20412414523193148f, r, W[5]);
      }
      const float y1_0_6 = py1 * z0_6;
      { const float r = px5 * y1_0_6 + 20.0f * px3 * y1_0_5 + 60.0f * px1 * y1_0_4;
        W[11] = fmaf(0.20412414523193148f, r, W[11]);
      }
      }
      {
      const float y2_0_3 = py2 * z0_3 + 2.0f * py0 * z0_2;
      { const float r = px1 * y2_0_3;
        W[0] = fmaf(-0.6123724356957945f, r, W[0]);
      }
      const float y2_0_4 = py2 * z0_4 + 2.0f * py0 * z0_3;
      { const float r = px2 * y2_0_4 + 2.0f * px0 * y2_0_3;
        W[1] = fmaf(-0.6123724356957945f, r, W[1]);
      }
      const float y2_0_5 = py2 * z0_5 + 2.0f * py0 * z0_4;
      { const float r = px3 * y2_0_5 + 6.0f * px1 * y2_0_4;
        W[7] = fmaf(0.20412414523193148f, r, W[7]);
        W[4] = fmaf(-0.6123724356957945f, r, W[4]);
      }
      const float y2_0_6 = py2 * z0_6 + 2.0f * py0 * z0_5;
      { const float r = px4 * y2_0_6 + 12.0f * px2 * y2_0_5 + 12.0f * px0 * y2_0_4;
        W[13] = fmaf(0.20412414523193148f, r, W[13]);
        W[10] = fmaf(-0.6123724356957945f, r, W[10]);
      }
      }
      {
      const float y3_0_4 = py3 * z0_4 + 6.0f * py1 * z0_3;
      { const float r = px1 * y3_0_4;
        W[2] = fmaf(-0.6123724356957945f, r, W[2]);
      }
      const float y3_0_5 = py3 * z0_5 + 6.0f * py1 * z0_4;
      { const float r = px2 * y3_0_5 + 2.0f * px0 * y3_0_4;
        W[5] = fmaf(-0.6123724356957945f, r, W[5]);
      }
      const float y3_0_6 = py3 * z0_6 + 6.0f * py1 * z0_5;
      { const float r = px3 * y3_0_6 + 6.0f * px1 * y3_0_5;
        W[16] = fmaf(0.20412414523193148f, r, W[16]);
        W[11] = fmaf(-0.6123724356957945f, r, W[11]);
      }
      }
      {
      const float y4_0_5 = py4 * z0_5 + 12.0f * py2 * z0_4 + 12.0f * py0 * z0_3;
      { const float r = px1 * y4_0_5;
        W[7] = fmaf(-0.6123724356957945f, r, W[7]);
      }
      const float y4_0_6 = py4 * z0_6 + 12.0f * py2 * z0_5 + 12.0f * py0 * z0_4;
      { const float r = px2 * y4_0_6 + 2.0f * px0 * y4_0_5;
        W[13] = fmaf(-0.6123724356957945f, r, W[13]);
      }
      }
      {
      const float y5_0_6 = py5 * z0_6 + 20.0f * py3 * z0_5 + 60.0f * py1 * z0_4;
      { const float r = px1 * y5_0_6;
        W[16] = fmaf(-0.6123724356957945f, r, W[16]);
      }
      }
      {
      }
      }
      {
      const float z1_1 = pz1 * g1;
      const float z1_2 = pz1 * g2;
      const float z1_3 = pz1 * g3;
      const float z1_4 = pz1 * g4;
      const float z1_5 = pz1 * g5;
      const float z1_6 = pz1 * g6;
      {
      const float y0_1_4 = py0 * z1_4;
      const float y0_1_3 = py0 * z1_3;
      { const float r = px3 * y0_1_4 + 6.0f * px1 * y0_1_3;
        W[3] = fmaf(0.20412414523193148f, r, W[3]);
      }
      const float y0_1_5 = py0 * z1_5;
      { const float r = px4 * y0_1_5 + 12.0f * px2 * y0_1_4 + 12.0f * px0 * y0_1_3;
        W[6] = fmaf(0.20412414523193148f, r, W[6]);
      }
      const float y0_1_6 = py0 * z1_6;
      { const float r = px5 * y0_1_6 + 20.0f * px3 * y0_1_5 + 60.0f * px1 * y0_1_4;
        W[12] = fmaf(0.20412414523193148f, r, W[12]);
      }
      }
      {
      const float y1_1_5 = py1 * z1_5;
      const float y1_1_4 = py1 * z1_4;
      { const float r = px3 * y1_1_5 + 6.0f * px1 * y1_1_4;
        W[8] = fmaf(0.20412414523193148f, r, W[8]);
      }
      const float y1_1_6 = py1 * z1_6;
      { const float r = px4 * y1_1_6 + 12.0f * px2 * y1_1_5 + 12.0f * px0 * y1_1_4;
        W[14] = fmaf(0.20412414523193148f, r, W[14]);
      }
      }
      {
      const float y2_1_4 = py2 * z1_4 + 2.0f * py0 * z1_3;
      { const float r = px1 * y2_1_4;
        W[3] = fmaf(-0.6123724356957945f, r, W[3]);
      }
      const float y2_1_5 = py2 * z1_5 + 2.0f * py0 * z1_4;
      { const float r = px2 * y2_1_5 + 2.0f * px0 * y2_1_4;
        W[6] = fmaf(-0.6123724356957945f, r, W[6]);
      }
      const float y2_1_6 = py2 * z1_6 + 2.0f * py0 * z1_5;
      { const float r = px3 * y2_1_6 + 6.0f * px1 * y2_1_5;
        W[17] = fmaf(0.20412414523193148f, r, W[17]);
        W[12] = fmaf(-0.6123724356957945f, r, W[12]);
      }
      }
      {
      const float y3_1_5 = py3 * z1_5 + 6.0f * py1 * z1_4;
      { const float r = px1 * y3_1_5;
        W[8] = fmaf(-0.6123724356957945f, r, W[8]);
      }
      const float y3_1_6 = py3 * z1_6 + 6.0f * py1 * z1_5;
      { const float r = px2 * y3_1_6 + 2.0f * px0 * y3_1_5;
        W[14] = fmaf(-0.6123724356957945f, r, W[14]);
      }
      }
      {
      const float y4_1_6 = py4 * z1_6 + 12.0f * py2 * z1_5 + 12.0f * py0 * z1_4;
      { const float r = px1 * y4_1_6;
        W[17] = fmaf(-0.6123724356957945f, r, W[17]);
      }
      }
      {
      }
      }
      {
      const float z2_2 = pz2 * g2 + 2.0f * pz0 * g1;
      const float z2_3 = pz2 * g3 + 2.0f * pz0 * g2;
      const float z2_4 = pz2 * g4 + 2.0f * pz0 * g3;
      const float z2_5 = pz2 * g5 + 2.0f * pz0 * g4;
      const float z2_6 = pz2 * g6 + 2.0f * pz0 * g5;
      {
      const float y0_2_5 = py0 * z2_5;
      const float y0_2_4 = py0 * z2_4;
      { const float r = px3 * y0_2_5 + 6.0f * px1 * y0_2_4;
        W[9] = fmaf(0.20412414523193148f, r, W[9]);
      }
      const float y0_2_6 = py0 * z2_6;
      { const float r = px4 * y0_2_6 + 12.0f * px2 * y0_2_5 + 12.0f * px0 * y0_2_4;
        W[15] = fmaf(0.20412414523193148f, r, W[15]);
      }
      }
      {
      const float y1_2_6 = py1 * z2_6;
      const float y1_2_5 = py1 * z2_5;
      { const float r = px3 * y1_2_6 + 6.0f * px1 * y1_2_5;
        W[18] = fmaf(0.20412414523193148f, r, W[18]);
      }
      }
      {
      const float y2_2_5 = py2 * z2_5 + 2.0f * py0 * z2_4;
      { const float r = px1 * y2_2_5;
        W[9] = fmaf(-0.6123724356957945f, r, W[9]);
      }
      const float y2_2_6 = py2 * z2_6 + 2.0f * py0 * z2_5;
      { const float r = px2 * y2_2_6 + 2.0f * px0 * y2_2_5;
        W[15] = fmaf(-0.6123724356957945f, r, W[15]);
      }
      }
      {
      const float y3_2_6 = py3 * z2_6 + 6.0f * py1 * z2_5;
      { const float r = px1 * y3_2_6;
        W[18] = fmaf(-0.6123724356957945f, r, W[18]);
      }
      }
      {
      }
      }
      {
      const float z3_3 = pz3 * g3 + 6.0f * pz1 * g2;
      const float z3_4 = pz3 * g4 + 6.0f * pz1 * g3;
      const float z3_5 = pz3 * g5 + 6.0f * pz1 * g4;
      const float z3_6 = pz3 * g6 + 6.0f * pz1 * g5;
      {
      const float y0_3_6 = py0 * z3_6;
      const float y0_3_5 = py0 * z3_5;
      { const float r = px3 * y0_3_6 + 6.0f * px1 * y0_3_5;
        W[19] = fmaf(0.20412414523193148f, r, W[19]);
      }
      }
      {
      }
      {
      const float y2_3_6 = py2 * z3_6 + 2.0f * py0 * z3_5;
      { const float r = px1 * y2_3_6;
        W[19] = fmaf(-0.6123724356957945f, r, W[19]);
      }
      }
      {
      }
      }
      {
      const float z4_4 = pz4 * g4 + 12.0f * pz2 * g3 + 12.0f * pz0 * g2;
      const float z4_5 = pz4 * g5 + 12.0f * pz2 * g4 + 12.0f * pz0 * g3;
      const float z4_6 = pz4 * g6 + 12.0f * pz2 * g5 + 12.0f * pz0 * g4;
      {
      }
      {
      }
      {
      }
      }
      {
      const float z5_5 = pz5 * g5 + 20.0f * pz3 * g4 + 60.0f * pz1 * g3;
      const float z5_6 = pz5 * g6 + 20.0f * pz3 * g5 + 60.0f * pz1 * g4;
      {
      }
      {
      }
      }
      {
      const float z6_6 = pz6 * g6 + 30.0f * pz4 * g5 + 180.0f * pz2 * g4 + 120.0f * pz0 * g3;
      {
      }
      }
      { const float e = rf[4];
        accf[0] = fmaf(e, W[0], accf[0]);
      }
      { const float e = rf[5];
        accf[0] = fmaf(e, W[1], accf[0]);
      }
      { const float e = rf[6];
        accf[0] = fmaf(e, W[4], accf[0]);
      }
      { const float e = rf[7];
        accf[0] = fmaf(e, W[10], accf[0]);
      }
      { const float e = rf[8];
        accf[1] = fmaf(e, W[0], accf[1]);
      }
      { const float e = rf[9];
        accf[1] = fmaf(e, W[2], accf[1]);
      }
      { const float e = rf[10];
        accf[1] = fmaf(e, W[1], accf[1]);
      }
      { const float e = rf[11];
        accf[1] = fmaf(e, W[5], accf[1]);
      }
      { const float e = rf[12];
        accf[1] = fmaf(e, W[4], accf[1]);
      }
      { const float e = rf[13];
        accf[1] = fmaf(e, W[11], accf[1]);
      }
      { const float e = rf[14];
        accf[2] = fmaf(e, W[0], accf[2]);
      }
      { const float e = rf[15];
        accf[2] = fmaf(e, W[3], accf[2]);
      }
      { const float e = rf[16];
        accf[2] = fmaf(e, W[1], accf[2]);
      }
      { const float e = rf[17];
        accf[2] = fmaf(e, W[6], accf[2]);
      }
      { const float e = rf[18];
        accf[2] = fmaf(e, W[4], accf[2]);
      }
      { const float e = rf[19];
        accf[2] = fmaf(e, W[12], accf[2]);
      }
      { const float e = rf[20];
        accf[3] = fmaf(e, W[0], accf[3]);
      }
      { const float e = rf[21];
        accf[3] = fmaf(e, W[2], accf[3]);
      }
      { const float e = rf[22];
        accf[3] = fmaf(e, W[1], accf[3]);
      }
      { const float e = rf[23];
        accf[3] = fmaf(e, W[5], accf[3]);
      }
      { const float e = rf[24];
        accf[3] = fmaf(e, W[4], accf[3]);
      }
      { const float e = rf[25];
        accf[3] = fmaf(e, W[11], accf[3]);
      }
      { const float e = rf[26];
        accf[4] = fmaf(e, W[0], accf[4]);
      }
      { const float e = rf[27];
        accf[4] = fmaf(e, W[2], accf[4]);
      }
      { const float e = rf[28];
        accf[4] = fmaf(e, W[7], accf[4]);
      }
      { const float e = rf[29];
        accf[4] = fmaf(e, W[1], accf[4]);
      }
      { const float e = rf[30];
        accf[4] = fmaf(e, W[5], accf[4]);
      }
      { const float e = rf[31];
        accf[4] = fmaf(e, W[13], accf[4]);
      }
      { const float e = rf[32];
        accf[5] = fmaf(e, W[0], accf[5]);
      }
      { const float e = rf[33];
        accf[5] = fmaf(e, W[3], accf[5]);
      }
      { const float e = rf[34];
        accf[5] = fmaf(e, W[2], accf[5]);
      }
      { const float e = rf[35];
        accf[5] = fmaf(e, W[8], accf[5]);
      }
      { const float e = rf[36];
        accf[5] = fmaf(e, W[1], accf[5]);
      }
      { const float e = rf[37];
        accf[5] = fmaf(e, W[6], accf[5]);
      }
      { const float e = rf[38];
        accf[5] = fmaf(e, W[5], accf[5]);
      }
      { const float e = rf[39];
        accf[5] = fmaf(e, W[14], accf[5]);
      }
      { const float e = rf[40];
        accf[6] = fmaf(e, W[0], accf[6]);
      }
      { const float e = rf[41];
        accf[6] = fmaf(e, W[3], accf[6]);
      }
      { const float e = rf[42];
        accf[6] = fmaf(e, W[1], accf[6]);
      }
      { const float e = rf[43];
        accf[6] = fmaf(e, W[6], accf[6]);
      }
      { const float e = rf[44];
        accf[6] = fmaf(e, W[4], accf[6]);
      }
      { const float e = rf[45];
        accf[6] = fmaf(e, W[12], accf[6]);
      }
      { const float e = rf[46];
        accf[7] = fmaf(e, W[0], accf[7]);
      }
      { const float e = rf[47];
        accf[7] = fmaf(e, W[3], accf[7]);
      }
      { const float e = rf[48];
        accf[7] = fmaf(e, W[2], accf[7]);
      }
      { const float e = rf[49];
        accf[7] = fmaf(e, W[8], accf[7]);
      }
      { const float e = rf[50];
        accf[7] = fmaf(e, W[1], accf[7]);
      }
      { const float e = rf[51];
        accf[7] = fmaf(e, W[6], accf[7]);
      }
      { const float e = rf[52];
        accf[7] = fmaf(e, W[5], accf[7]);
      }
      { const float e = rf[53];
        accf[7] = fmaf(e, W[14], accf[7]);
      }
      { const float e = rf[54];
        accf[8] = fmaf(e, W[0], accf[8]);
      }
      { const float e = rf[55];
        accf[8] = fmaf(e, W[3], accf[8]);
      }
      { const float e = rf[56];
        accf[8] = fmaf(e, W[9], accf[8]);
      }
      { const float e = rf[57];
        accf[8] = fmaf(e, W[1], accf[8]);
      }
      { const float e = rf[58];
        accf[8] = fmaf(e, W[6], accf[8]);
      }
      { const float e = rf[59];
        accf[8] = fmaf(e, W[15], accf[8]);
      }
      { const float e = rf[60];
        accf[9] = fmaf(e, W[0], accf[9]);
      }
      { const float e = rf[61];
        accf[9] = fmaf(e, W[2], accf[9]);
      }
      { const float e = rf[62];
        accf[9] = fmaf(e, W[7], accf[9]);
      }
      { const float e = rf[63];
        accf[9] = fmaf(e, W[1], accf[9]);
      }
      { const float e = rf[64];
        accf[9] = fmaf(e, W[5], accf[9]);
      }
      { const float e = rf[65];
        accf[9] = fmaf(e, W[13], accf[9]);
      }
      { const float e = rf[66];
        accf[10] = fmaf(e, W[0], accf[10]);
      }
      { const float e = rf[67];
        accf[10] = fmaf(e, W[2], accf[10]);
      }
      { const float e = rf[68];
        accf[10] = fmaf(e, W[7], accf[10]);
      }
      { const float e = rf[69];
        accf[10] = fmaf(e, W[16], accf[10]);
      }
      { const float e = rf[70];
        accf[11] = fmaf(e, W[0], accf[11]);
      }
      { const float e = rf[71];
        accf[11] = fmaf(e, W[3], accf[11]);
      }
      { const float e = rf[72];
        accf[11] = fmaf(e, W[2], accf[11]);
      }
      { const float e = rf[73];
        accf[11] = fmaf(e, W[8], accf[11]);
      }
      { const float e = rf[74];
        accf[11] = fmaf(e, W[7], accf[11]);
      }
      { const float e = rf[75];
        accf[11] = fmaf(e, W[17], accf[11]);
      }
      { const float e = rf[76];
        accf[12] = fmaf(e, W[0], accf[12]);
      }
      { const float e = rf[77];
        accf[12] = fmaf(e, W[3], accf[12]);
      }
      { const float e = rf[78];
        accf[12] = fmaf(e, W[2], accf[12]);
      }
      { const float e = rf[79];
        accf[12] = fmaf(e, W[8], accf[12]);
      }
      { const float e = rf[80];
        accf[12] = fmaf(e, W[1], accf[12]);
      }
      { const float e = rf[81];
        accf[12] = fmaf(e, W[6], accf[12]);
      }
      { const float e = rf[82];
        accf[12] = fmaf(e, W[5], accf[12]);
      }
      { const float e = rf[83];
        accf[12] = fmaf(e, W[14], accf[12]);
      }
      { const float e = rf[84];
        accf[13] = fmaf(e, W[0], accf[13]);
      }
      { const float e = rf[85];
        accf[13] = fmaf(e, W[3], accf[13]);
      }
      { const float e = rf[86];
        accf[13] = fmaf(e, W[2], accf[13]);
      }
      { const float e = rf[87];
        accf[13] = fmaf(e, W[8], accf[13]);
      }
      { const float e = rf[88];
        accf[13] = fmaf(e, W[7], accf[13]);
      }
      { const float e = rf[89];
        accf[13] = fmaf(e, W[17], accf[13]);
      }
      { const float e = rf[90];
        accf[14] = fmaf(e, W[0], accf[14]);
      }
      { const float e = rf[91];
        accf[14] = fmaf(e, W[3], accf[14]);
      }
      { const float e = rf[92];
        accf[14] = fmaf(e, W[9], accf[14]);
      }
      { const float e = rf[93];
        accf[14] = fmaf(e, W[2], accf[14]);
      }
      { const float e = rf[94];
        accf[14] = fmaf(e, W[8], accf[14]);
      }
      { const float e = rf[95];
        accf[14] = fmaf(e, W[18], accf[14]);
      }
      { const float e = rf[96];
        accf[15] = fmaf(e, W[0], accf[15]);
      }
      { const float e = rf[97];
        accf[15] = fmaf(e, W[3], accf[15]);
      }
      { const float e = rf[98];
        accf[15] = fmaf(e, W[9], accf[15]);
      }
      { const float e = rf[99];
        accf[15] = fmaf(e, W[1], accf[15]);
      }
      { const float e = rf[100];
        accf[15] = fmaf(e, W[6], accf[15]);
      }
      { const float e = rf[101];
        accf[15] = fmaf(e, W[15], accf[15]);
      }
      { const float e = rf[102];
        accf[16] = fmaf(e, W[0], accf[16]);
      }
      { const float e = rf[103];
        accf[16] = fmaf(e, W[3], accf[16]);
      }
      { const float e = rf[104];
        accf[16] = fmaf(e, W[9], accf[16]);
      }
      { const float e = rf[105];
        accf[16] = fmaf(e, W[2], accf[16]);
      }
      { const float e = rf[106];
        accf[16] = fmaf(e, W[8], accf[16]);
      }
      { const float e = rf[107];
        accf[16] = fmaf(e, W[18], accf[16]);
      }
      { const float e = rf[108];
        accf[17] = fmaf(e, W[0], accf[17]);
      }
      { const float e = rf[109];
        accf[17] = fmaf(e, W[3], accf[17]);
      }
      { const float e = rf[110];
        accf[17] = fmaf(e, W[9], accf[17]);
      }
      { const float e = rf[111];
        accf[17] = fmaf(e, W[19], accf[17]);
      }
      }
#pragma unroll
      for (int q = 0; q < 18; q++) acc[q] += (double)accf[q];
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[4];
    out[(row0 + 1) * ldo + col + 6] = 1.0 * acc[5];
    out[(row0 + 2) * ldo + col + 6] = 1.0 * acc[3];
    out[(row0 + 3) * ldo + col + 6] = 1.0 * acc[13];
    out[(row0 + 4) * ldo + col + 6] = 1.0 * acc[14];
    out[(row0 + 5) * ldo + col + 6] = 1.0 * acc[12];
    out[(row0 + 6) * ldo + col + 6] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[10] + 0.5773502691896258 * acc[16];
    out[(row0 + 7) * ldo + col + 6] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[11] + 0.5773502691896258 * acc[17];
    out[(row0 + 8) * ldo + col + 6] = -0.2886751345948129 * acc[0] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[15];
    out[(row0 + 9) * ldo + col + 6] = 1.0 * acc[7];
    out[(row0 + 10) * ldo + col + 6] = 1.0 * acc[8];
    out[(row0 + 11) * ldo + col + 6] = 1.0 * acc[6];
    out[(row0 + 12) * ldo + col + 6] = 0.5 * acc[1] + -0.5 * acc[10];
    out[(row0 + 13) * ldo + col + 6] = 0.5 * acc[2] + -0.5 * acc[11];
    out[(row0 + 14) * ldo + col + 6] = 0.5 * acc[0] + -0.5 * acc[9];
    }
    break;
  }
  default: break;
  }
}

void register_gen3_21(Gen3Table& t) {
  t[2][1][0] = Gen3Launch{&int3c2e_gen_210, &int3c2e_gen32_210, 1, 1, 112};
  t[2][1][1] = Gen3Launch{&int3c2e_gen_211, &int3c2e_gen32_211, 3, 2, 112};
  t[2][1][2] = Gen3Launch{&int3c2e_gen_212, &int3c2e_gen32_212, 5, 3, 112};
  t[2][1][3] = Gen3Launch{&int3c2e_gen_213, &int3c2e_gen32_213, 7, 4, 112};
}

}  // namespace dfhip
