// Generated by gen3c.py -- 3c2e classes (21|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (21|0): L = 3, record 112 doubles, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_210(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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

// class (21|1): L = 4, record 112 doubles, 3 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_211(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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

// class (21|2): L = 5, record 112 doubles, 5 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_212(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
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

// class (21|3): L = 6, record 112 doubles, 7 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_213(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
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

void register_gen3_21(Gen3Table& t) {
  t[2][1][0] = Gen3Launch{&int3c2e_gen_210, 1, 112};
  t[2][1][1] = Gen3Launch{&int3c2e_gen_211, 3, 112};
  t[2][1][2] = Gen3Launch{&int3c2e_gen_212, 5, 112};
  t[2][1][3] = Gen3Launch{&int3c2e_gen_213, 7, 112};
}

}  // namespace dfhip
