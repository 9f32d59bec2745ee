// Generated by gen3c.py -- 3c2e classes (22|lc), lc <= 3.
#include "gen3c_common.h"

namespace dfhip {

// class (22|0): L = 4, record 340 doubles, 1 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_220(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
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
      const double y0_0_4 = py0 * z0_4;
      { const double r = px4 * y0_0_4 + 12.0 * px2 * y0_0_3 + 12.0 * px0 * y0_0_2;
        W[20] = fma(1.0, r, W[20]);
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
      const double y1_0_4 = py1 * z0_4;
      { const double r = px3 * y1_0_4 + 6.0 * px1 * y1_0_3;
        W[21] = fma(1.0, r, W[21]);
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
      const double y2_0_4 = py2 * z0_4 + 2.0 * py0 * z0_3;
      { const double r = px2 * y2_0_4 + 2.0 * px0 * y2_0_3;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y3_0_3 = py3 * z0_3 + 6.0 * py1 * z0_2;
      { const double r = px0 * y3_0_3;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[30] = fma(1.0, r, W[30]);
      }
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
      const double y0_1_4 = py0 * z1_4;
      { const double r = px3 * y0_1_4 + 6.0 * px1 * y0_1_3;
        W[22] = fma(1.0, r, W[22]);
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
      const double y1_1_4 = py1 * z1_4;
      { const double r = px2 * y1_1_4 + 2.0 * px0 * y1_1_3;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y2_1_3 = py2 * z1_3 + 2.0 * py0 * z1_2;
      { const double r = px0 * y2_1_3;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_2_3 = py0 * z2_3;
      { const double r = px1 * y0_2_3;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y1_2_3 = py1 * z2_3;
      { const double r = px0 * y1_2_3;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[32] = fma(1.0, r, W[32]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      {
      const double y0_3_3 = py0 * z3_3;
      { const double r = px0 * y0_3_3;
        W[19] = fma(1.0, r, W[19]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[33] = fma(1.0, r, W[33]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 0] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 0] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 0] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 0] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 0] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 0] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 0] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 0] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 0] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  default: break;
  }
}

// class (22|1): L = 5, record 340 doubles, 3 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_221(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
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
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[20] = fma(1.0, r, W[20]);
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
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[21] = fma(1.0, r, W[21]);
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
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y4_0_4 = py4 * z0_4 + 12.0 * py2 * z0_3 + 12.0 * py0 * z0_2;
      { const double r = px0 * y4_0_4;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[30] = fma(1.0, r, W[30]);
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
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[22] = fma(1.0, r, W[22]);
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
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[32] = fma(1.0, r, W[32]);
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
        W[19] = fma(1.0, r, W[19]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[33] = fma(1.0, r, W[33]);
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
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 0] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 0] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 0] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 0] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 0] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 0] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 0] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 0] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 0] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 1: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
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
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[20] = fma(1.0, r, W[20]);
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
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[21] = fma(1.0, r, W[21]);
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
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y3_1_4 = py3 * z1_4 + 6.0 * py1 * z1_3;
      { const double r = px0 * y3_1_4;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[30] = fma(1.0, r, W[30]);
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
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[22] = fma(1.0, r, W[22]);
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
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y2_2_4 = py2 * z2_4 + 2.0 * py0 * z2_3;
      { const double r = px0 * y2_2_4;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_3_4 = py0 * z3_4;
      { const double r = px1 * y0_3_4;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y1_3_4 = py1 * z3_4;
      { const double r = px0 * y1_3_4;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[32] = fma(1.0, r, W[32]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      {
      const double y0_4_4 = py0 * z4_4;
      { const double r = px0 * y0_4_4;
        W[19] = fma(1.0, r, W[19]);
      }
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[33] = fma(1.0, r, W[33]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      {
      const double y0_5_5 = py0 * z5_5;
      { const double r = px0 * y0_5_5;
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 1] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 1] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 1] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 1] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 1] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 1] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 1] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 1] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 1] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 1] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 1] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 1] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 1] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 1] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 2: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
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
      const double y0_0_5 = py0 * z0_5;
      { const double r = px5 * y0_0_5 + 20.0 * px3 * y0_0_4 + 60.0 * px1 * y0_0_3;
        W[20] = fma(1.0, r, W[20]);
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
      const double y1_0_5 = py1 * z0_5;
      { const double r = px4 * y1_0_5 + 12.0 * px2 * y1_0_4 + 12.0 * px0 * y1_0_3;
        W[21] = fma(1.0, r, W[21]);
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
      const double y2_0_5 = py2 * z0_5 + 2.0 * py0 * z0_4;
      { const double r = px3 * y2_0_5 + 6.0 * px1 * y2_0_4;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y3_0_4 = py3 * z0_4 + 6.0 * py1 * z0_3;
      { const double r = px1 * y3_0_4;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y3_0_5 = py3 * z0_5 + 6.0 * py1 * z0_4;
      { const double r = px2 * y3_0_5 + 2.0 * px0 * y3_0_4;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[30] = fma(1.0, r, W[30]);
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
      const double y0_1_5 = py0 * z1_5;
      { const double r = px4 * y0_1_5 + 12.0 * px2 * y0_1_4 + 12.0 * px0 * y0_1_3;
        W[22] = fma(1.0, r, W[22]);
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
      const double y1_1_5 = py1 * z1_5;
      { const double r = px3 * y1_1_5 + 6.0 * px1 * y1_1_4;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y2_1_4 = py2 * z1_4 + 2.0 * py0 * z1_3;
      { const double r = px1 * y2_1_4;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y2_1_5 = py2 * z1_5 + 2.0 * py0 * z1_4;
      { const double r = px2 * y2_1_5 + 2.0 * px0 * y2_1_4;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_2_4 = py0 * z2_4;
      { const double r = px2 * y0_2_4 + 2.0 * px0 * y0_2_3;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y1_2_4 = py1 * z2_4;
      { const double r = px1 * y1_2_4;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[32] = fma(1.0, r, W[32]);
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
        W[19] = fma(1.0, r, W[19]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[33] = fma(1.0, r, W[33]);
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
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 2] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 2] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 2] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 2] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 2] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 2] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 2] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 2] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 2] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 2] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 2] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 2] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 2] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  default: break;
  }
}

// class (22|2): L = 6, record 340 doubles, 5 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_222(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
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
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[20] = fma(1.0, r, W[20]);
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
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[21] = fma(1.0, r, W[21]);
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
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y4_0_5 = py4 * z0_5 + 12.0 * py2 * z0_4 + 12.0 * py0 * z0_3;
      { const double r = px1 * y4_0_5;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[30] = fma(1.0, r, W[30]);
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
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[22] = fma(1.0, r, W[22]);
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
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_2_5 = py1 * z2_5;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[32] = fma(1.0, r, W[32]);
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
        W[19] = fma(1.0, r, W[19]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[33] = fma(1.0, r, W[33]);
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
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 0] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 0] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 0] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 0] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 0] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 0] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 0] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 0] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 0] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 1: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
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
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[20] = fma(1.0, r, W[20]);
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
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[21] = fma(1.0, r, W[21]);
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
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[30] = fma(1.0, r, W[30]);
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
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[22] = fma(1.0, r, W[22]);
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
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[32] = fma(1.0, r, W[32]);
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
        W[19] = fma(1.0, r, W[19]);
      }
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[33] = fma(1.0, r, W[33]);
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
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 1] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 1] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 1] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 1] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 1] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 1] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 1] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 1] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 1] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 1] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 1] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 1] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 1] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 1] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 2: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
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
      const double y0_0_6 = py0 * z0_6;
      { const double r = px6 * y0_0_6 + 30.0 * px4 * y0_0_5 + 180.0 * px2 * y0_0_4 + 120.0 * px0 * y0_0_3;
        W[20] = fma(-0.2886751345948129, r, W[20]);
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
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[21] = fma(-0.2886751345948129, r, W[21]);
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
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[23] = fma(-0.2886751345948129, r, W[23]);
        W[20] = fma(-0.2886751345948129, r, W[20]);
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
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[26] = fma(-0.2886751345948129, r, W[26]);
        W[21] = fma(-0.2886751345948129, r, W[21]);
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
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[30] = fma(-0.2886751345948129, r, W[30]);
        W[23] = fma(-0.2886751345948129, r, W[23]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[16] = fma(-0.2886751345948129, r, W[16]);
      }
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[26] = fma(-0.2886751345948129, r, W[26]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[30] = fma(-0.2886751345948129, r, W[30]);
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
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[22] = fma(-0.2886751345948129, r, W[22]);
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
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[24] = fma(-0.2886751345948129, r, W[24]);
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
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[27] = fma(-0.2886751345948129, r, W[27]);
        W[22] = fma(-0.2886751345948129, r, W[22]);
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
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[31] = fma(-0.2886751345948129, r, W[31]);
        W[24] = fma(-0.2886751345948129, r, W[24]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[17] = fma(-0.2886751345948129, r, W[17]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[27] = fma(-0.2886751345948129, r, W[27]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[31] = fma(-0.2886751345948129, r, W[31]);
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
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[25] = fma(-0.2886751345948129, r, W[25]);
        W[20] = fma(0.5773502691896258, r, W[20]);
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
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[28] = fma(-0.2886751345948129, r, W[28]);
        W[21] = fma(0.5773502691896258, r, W[21]);
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
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[32] = fma(-0.2886751345948129, r, W[32]);
        W[25] = fma(-0.2886751345948129, r, W[25]);
        W[23] = fma(0.5773502691896258, r, W[23]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[18] = fma(-0.2886751345948129, r, W[18]);
        W[16] = fma(0.5773502691896258, r, W[16]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[28] = fma(-0.2886751345948129, r, W[28]);
        W[26] = fma(0.5773502691896258, r, W[26]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[32] = fma(-0.2886751345948129, r, W[32]);
        W[30] = fma(0.5773502691896258, r, W[30]);
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
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[29] = fma(-0.2886751345948129, r, W[29]);
        W[22] = fma(0.5773502691896258, r, W[22]);
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
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[33] = fma(-0.2886751345948129, r, W[33]);
        W[24] = fma(0.5773502691896258, r, W[24]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[19] = fma(-0.2886751345948129, r, W[19]);
        W[17] = fma(0.5773502691896258, r, W[17]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[29] = fma(-0.2886751345948129, r, W[29]);
        W[27] = fma(0.5773502691896258, r, W[27]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[33] = fma(-0.2886751345948129, r, W[33]);
        W[31] = fma(0.5773502691896258, r, W[31]);
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
        W[9] = fma(0.5773502691896258, r, W[9]);
      }
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[15] = fma(0.5773502691896258, r, W[15]);
      }
      const double y0_4_6 = py0 * z4_6;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[34] = fma(-0.2886751345948129, r, W[34]);
        W[25] = fma(0.5773502691896258, r, W[25]);
      }
      }
      {
      const double y1_4_5 = py1 * z4_5;
      { const double r = px0 * y1_4_5;
        W[18] = fma(0.5773502691896258, r, W[18]);
      }
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[28] = fma(0.5773502691896258, r, W[28]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[34] = fma(-0.2886751345948129, r, W[34]);
        W[32] = fma(0.5773502691896258, r, W[32]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      {
      const double y0_5_5 = py0 * z5_5;
      { const double r = px0 * y0_5_5;
        W[19] = fma(0.5773502691896258, r, W[19]);
      }
      const double y0_5_6 = py0 * z5_6;
      { const double r = px1 * y0_5_6;
        W[29] = fma(0.5773502691896258, r, W[29]);
      }
      }
      {
      const double y1_5_6 = py1 * z5_6;
      { const double r = px0 * y1_5_6;
        W[33] = fma(0.5773502691896258, r, W[33]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      {
      const double y0_6_6 = py0 * z6_6;
      { const double r = px0 * y0_6_6;
        W[34] = fma(0.5773502691896258, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 2] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 2] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 2] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 2] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 2] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 2] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 2] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 2] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 2] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 2] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 2] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 2] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 2] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 3: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
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
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[20] = fma(1.0, r, W[20]);
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
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[21] = fma(1.0, r, W[21]);
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
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y3_1_5 = py3 * z1_5 + 6.0 * py1 * z1_4;
      { const double r = px1 * y3_1_5;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[30] = fma(1.0, r, W[30]);
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
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[22] = fma(1.0, r, W[22]);
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
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px1 * y2_2_5;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[31] = fma(1.0, r, W[31]);
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
        W[9] = fma(1.0, r, W[9]);
      }
      const double y0_3_5 = py0 * z3_5;
      { const double r = px2 * y0_3_5 + 2.0 * px0 * y0_3_4;
        W[15] = fma(1.0, r, W[15]);
      }
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y1_3_5 = py1 * z3_5;
      { const double r = px1 * y1_3_5;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y1_3_6 = py1 * z3_6;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[32] = fma(1.0, r, W[32]);
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
        W[19] = fma(1.0, r, W[19]);
      }
      const double y0_4_6 = py0 * z4_6;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[33] = fma(1.0, r, W[33]);
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
        W[34] = fma(1.0, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 3] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 3] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 3] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 3] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 3] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 3] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 3] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 3] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 3] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 3] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 3] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 3] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 3] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 3] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 3] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 3] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 3] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 4: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= i2g * i2g; }
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
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
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
      const double y0_0_6 = py0 * z0_6;
      { const double r = px6 * y0_0_6 + 30.0 * px4 * y0_0_5 + 180.0 * px2 * y0_0_4 + 120.0 * px0 * y0_0_3;
        W[20] = fma(0.5, r, W[20]);
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
      const double y1_0_6 = py1 * z0_6;
      { const double r = px5 * y1_0_6 + 20.0 * px3 * y1_0_5 + 60.0 * px1 * y1_0_4;
        W[21] = fma(0.5, r, W[21]);
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
      const double y2_0_6 = py2 * z0_6 + 2.0 * py0 * z0_5;
      { const double r = px4 * y2_0_6 + 12.0 * px2 * y2_0_5 + 12.0 * px0 * y2_0_4;
        W[23] = fma(0.5, r, W[23]);
        W[20] = fma(-0.5, r, W[20]);
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
      const double y3_0_6 = py3 * z0_6 + 6.0 * py1 * z0_5;
      { const double r = px3 * y3_0_6 + 6.0 * px1 * y3_0_5;
        W[26] = fma(0.5, r, W[26]);
        W[21] = fma(-0.5, r, W[21]);
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
      const double y4_0_6 = py4 * z0_6 + 12.0 * py2 * z0_5 + 12.0 * py0 * z0_4;
      { const double r = px2 * y4_0_6 + 2.0 * px0 * y4_0_5;
        W[30] = fma(0.5, r, W[30]);
        W[23] = fma(-0.5, r, W[23]);
      }
      }
      {
      const double y5_0_5 = py5 * z0_5 + 20.0 * py3 * z0_4 + 60.0 * py1 * z0_3;
      { const double r = px0 * y5_0_5;
        W[16] = fma(-0.5, r, W[16]);
      }
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[26] = fma(-0.5, r, W[26]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[30] = fma(-0.5, r, W[30]);
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
      const double y0_1_6 = py0 * z1_6;
      { const double r = px5 * y0_1_6 + 20.0 * px3 * y0_1_5 + 60.0 * px1 * y0_1_4;
        W[22] = fma(0.5, r, W[22]);
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
      const double y1_1_6 = py1 * z1_6;
      { const double r = px4 * y1_1_6 + 12.0 * px2 * y1_1_5 + 12.0 * px0 * y1_1_4;
        W[24] = fma(0.5, r, W[24]);
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
      const double y2_1_6 = py2 * z1_6 + 2.0 * py0 * z1_5;
      { const double r = px3 * y2_1_6 + 6.0 * px1 * y2_1_5;
        W[27] = fma(0.5, r, W[27]);
        W[22] = fma(-0.5, r, W[22]);
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
      const double y3_1_6 = py3 * z1_6 + 6.0 * py1 * z1_5;
      { const double r = px2 * y3_1_6 + 2.0 * px0 * y3_1_5;
        W[31] = fma(0.5, r, W[31]);
        W[24] = fma(-0.5, r, W[24]);
      }
      }
      {
      const double y4_1_5 = py4 * z1_5 + 12.0 * py2 * z1_4 + 12.0 * py0 * z1_3;
      { const double r = px0 * y4_1_5;
        W[17] = fma(-0.5, r, W[17]);
      }
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[27] = fma(-0.5, r, W[27]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[31] = fma(-0.5, r, W[31]);
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
        W[9] = fma(0.5, r, W[9]);
      }
      const double y0_2_5 = py0 * z2_5;
      { const double r = px3 * y0_2_5 + 6.0 * px1 * y0_2_4;
        W[15] = fma(0.5, r, W[15]);
      }
      const double y0_2_6 = py0 * z2_6;
      { const double r = px4 * y0_2_6 + 12.0 * px2 * y0_2_5 + 12.0 * px0 * y0_2_4;
        W[25] = fma(0.5, r, W[25]);
      }
      }
      {
      const double y1_2_5 = py1 * z2_5;
      const double y1_2_4 = py1 * z2_4;
      { const double r = px2 * y1_2_5 + 2.0 * px0 * y1_2_4;
        W[18] = fma(0.5, r, W[18]);
      }
      const double y1_2_6 = py1 * z2_6;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[28] = fma(0.5, r, W[28]);
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
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[32] = fma(0.5, r, W[32]);
        W[25] = fma(-0.5, r, W[25]);
      }
      }
      {
      const double y3_2_5 = py3 * z2_5 + 6.0 * py1 * z2_4;
      { const double r = px0 * y3_2_5;
        W[18] = fma(-0.5, r, W[18]);
      }
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[28] = fma(-0.5, r, W[28]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[32] = fma(-0.5, r, W[32]);
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
        W[19] = fma(0.5, r, W[19]);
      }
      const double y0_3_6 = py0 * z3_6;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[29] = fma(0.5, r, W[29]);
      }
      }
      {
      const double y1_3_6 = py1 * z3_6;
      const double y1_3_5 = py1 * z3_5;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[33] = fma(0.5, r, W[33]);
      }
      }
      {
      const double y2_3_5 = py2 * z3_5 + 2.0 * py0 * z3_4;
      { const double r = px0 * y2_3_5;
        W[19] = fma(-0.5, r, W[19]);
      }
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[29] = fma(-0.5, r, W[29]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[33] = fma(-0.5, r, W[33]);
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
        W[34] = fma(0.5, r, W[34]);
      }
      }
      {
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[34] = fma(-0.5, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 4] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 4] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 4] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 4] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 4] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 4] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 4] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 4] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 4] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 4] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 4] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 4] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 4] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 4] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 4] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 4] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 4] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  default: break;
  }
}

// class (22|3): L = 7, record 340 doubles, 7 m-group(s)
__global__ __launch_bounds__(64) void int3c2e_gen_223(const double* __restrict__ recs, const int64_t* __restrict__ recoff,
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
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double y1_0_7 = py1 * z0_7;
      { const double r = px6 * y1_0_7 + 30.0 * px4 * y1_0_6 + 180.0 * px2 * y1_0_5 + 120.0 * px0 * y1_0_4;
        W[20] = fma(0.6123724356957945, r, W[20]);
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
      const double y2_0_7 = py2 * z0_7 + 2.0 * py0 * z0_6;
      { const double r = px5 * y2_0_7 + 20.0 * px3 * y2_0_6 + 60.0 * px1 * y2_0_5;
        W[21] = fma(0.6123724356957945, r, W[21]);
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
      const double y3_0_7 = py3 * z0_7 + 6.0 * py1 * z0_6;
      { const double r = px4 * y3_0_7 + 12.0 * px2 * y3_0_6 + 12.0 * px0 * y3_0_5;
        W[23] = fma(0.6123724356957945, r, W[23]);
        W[20] = fma(-0.20412414523193148, r, W[20]);
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
      const double y4_0_7 = py4 * z0_7 + 12.0 * py2 * z0_6 + 12.0 * py0 * z0_5;
      { const double r = px3 * y4_0_7 + 6.0 * px1 * y4_0_6;
        W[26] = fma(0.6123724356957945, r, W[26]);
        W[21] = fma(-0.20412414523193148, r, W[21]);
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
      const double y5_0_7 = py5 * z0_7 + 20.0 * py3 * z0_6 + 60.0 * py1 * z0_5;
      { const double r = px2 * y5_0_7 + 2.0 * px0 * y5_0_6;
        W[30] = fma(0.6123724356957945, r, W[30]);
        W[23] = fma(-0.20412414523193148, r, W[23]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[16] = fma(-0.20412414523193148, r, W[16]);
      }
      const double y6_0_7 = py6 * z0_7 + 30.0 * py4 * z0_6 + 180.0 * py2 * z0_5 + 120.0 * py0 * z0_4;
      { const double r = px1 * y6_0_7;
        W[26] = fma(-0.20412414523193148, r, W[26]);
      }
      }
      {
      const double y7_0_7 = py7 * z0_7 + 42.0 * py5 * z0_6 + 420.0 * py3 * z0_5 + 840.0 * py1 * z0_4;
      { const double r = px0 * y7_0_7;
        W[30] = fma(-0.20412414523193148, r, W[30]);
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
      const double z1_7 = pz1 * g7;
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[22] = fma(0.6123724356957945, r, W[22]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[24] = fma(0.6123724356957945, r, W[24]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[27] = fma(0.6123724356957945, r, W[27]);
        W[22] = fma(-0.20412414523193148, r, W[22]);
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
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[31] = fma(0.6123724356957945, r, W[31]);
        W[24] = fma(-0.20412414523193148, r, W[24]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[17] = fma(-0.20412414523193148, r, W[17]);
      }
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[27] = fma(-0.20412414523193148, r, W[27]);
      }
      }
      {
      const double y6_1_7 = py6 * z1_7 + 30.0 * py4 * z1_6 + 180.0 * py2 * z1_5 + 120.0 * py0 * z1_4;
      { const double r = px0 * y6_1_7;
        W[31] = fma(-0.20412414523193148, r, W[31]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[25] = fma(0.6123724356957945, r, W[25]);
      }
      }
      {
      const double y2_2_6 = py2 * z2_6 + 2.0 * py0 * z2_5;
      const double y2_2_5 = py2 * z2_5 + 2.0 * py0 * z2_4;
      { const double r = px2 * y2_2_6 + 2.0 * px0 * y2_2_5;
        W[18] = fma(0.6123724356957945, r, W[18]);
      }
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[28] = fma(0.6123724356957945, r, W[28]);
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
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[32] = fma(0.6123724356957945, r, W[32]);
        W[25] = fma(-0.20412414523193148, r, W[25]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[18] = fma(-0.20412414523193148, r, W[18]);
      }
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[28] = fma(-0.20412414523193148, r, W[28]);
      }
      }
      {
      const double y5_2_7 = py5 * z2_7 + 20.0 * py3 * z2_6 + 60.0 * py1 * z2_5;
      { const double r = px0 * y5_2_7;
        W[32] = fma(-0.20412414523193148, r, W[32]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
      {
      }
      {
      const double y1_3_6 = py1 * z3_6;
      const double y1_3_5 = py1 * z3_5;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[19] = fma(0.6123724356957945, r, W[19]);
      }
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[29] = fma(0.6123724356957945, r, W[29]);
      }
      }
      {
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[33] = fma(0.6123724356957945, r, W[33]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[19] = fma(-0.20412414523193148, r, W[19]);
      }
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[29] = fma(-0.20412414523193148, r, W[29]);
      }
      }
      {
      const double y4_3_7 = py4 * z3_7 + 12.0 * py2 * z3_6 + 12.0 * py0 * z3_5;
      { const double r = px0 * y4_3_7;
        W[33] = fma(-0.20412414523193148, r, W[33]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
      {
      }
      {
      const double y1_4_7 = py1 * z4_7;
      const double y1_4_6 = py1 * z4_6;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[34] = fma(0.6123724356957945, r, W[34]);
      }
      }
      {
      }
      {
      const double y3_4_7 = py3 * z4_7 + 6.0 * py1 * z4_6;
      { const double r = px0 * y3_4_7;
        W[34] = fma(-0.20412414523193148, r, W[34]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      }
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      }
      {
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 0] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 0] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 0] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 0] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 0] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 0] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 0] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 0] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 0] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 0] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 0] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 0] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 0] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 0] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 0] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 0] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 0] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 0] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 0] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 0] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 0] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 0] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 0] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 0] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 0] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 1: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double z1_7 = pz1 * g7;
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[20] = fma(1.0, r, W[20]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[21] = fma(1.0, r, W[21]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[23] = fma(1.0, r, W[23]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[16] = fma(1.0, r, W[16]);
      }
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[26] = fma(1.0, r, W[26]);
      }
      }
      {
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[30] = fma(1.0, r, W[30]);
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
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[22] = fma(1.0, r, W[22]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[24] = fma(1.0, r, W[24]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[17] = fma(1.0, r, W[17]);
      }
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[27] = fma(1.0, r, W[27]);
      }
      }
      {
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[31] = fma(1.0, r, W[31]);
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
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
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
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[25] = fma(1.0, r, W[25]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[18] = fma(1.0, r, W[18]);
      }
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[28] = fma(1.0, r, W[28]);
      }
      }
      {
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[32] = fma(1.0, r, W[32]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
      {
      }
      {
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[19] = fma(1.0, r, W[19]);
      }
      const double y1_4_7 = py1 * z4_7;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[29] = fma(1.0, r, W[29]);
      }
      }
      {
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[33] = fma(1.0, r, W[33]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      }
      {
      const double y1_5_7 = py1 * z5_7;
      { const double r = px1 * y1_5_7;
        W[34] = fma(1.0, r, W[34]);
      }
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      }
      {
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 1] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 1] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 1] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 1] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 1] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 1] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 1] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 1] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 1] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 1] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 1] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 1] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 1] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 1] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 1] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 1] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 1] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 1] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 1] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 1] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 1] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 1] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 1] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 1] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 1] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 2: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double y1_0_7 = py1 * z0_7;
      { const double r = px6 * y1_0_7 + 30.0 * px4 * y1_0_6 + 180.0 * px2 * y1_0_5 + 120.0 * px0 * y1_0_4;
        W[20] = fma(-0.15811388300841894, r, W[20]);
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
      const double y2_0_7 = py2 * z0_7 + 2.0 * py0 * z0_6;
      { const double r = px5 * y2_0_7 + 20.0 * px3 * y2_0_6 + 60.0 * px1 * y2_0_5;
        W[21] = fma(-0.15811388300841894, r, W[21]);
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
      const double y3_0_7 = py3 * z0_7 + 6.0 * py1 * z0_6;
      { const double r = px4 * y3_0_7 + 12.0 * px2 * y3_0_6 + 12.0 * px0 * y3_0_5;
        W[23] = fma(-0.15811388300841894, r, W[23]);
        W[20] = fma(-0.15811388300841894, r, W[20]);
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
      const double y4_0_7 = py4 * z0_7 + 12.0 * py2 * z0_6 + 12.0 * py0 * z0_5;
      { const double r = px3 * y4_0_7 + 6.0 * px1 * y4_0_6;
        W[26] = fma(-0.15811388300841894, r, W[26]);
        W[21] = fma(-0.15811388300841894, r, W[21]);
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
      const double y5_0_7 = py5 * z0_7 + 20.0 * py3 * z0_6 + 60.0 * py1 * z0_5;
      { const double r = px2 * y5_0_7 + 2.0 * px0 * y5_0_6;
        W[30] = fma(-0.15811388300841894, r, W[30]);
        W[23] = fma(-0.15811388300841894, r, W[23]);
      }
      }
      {
      const double y6_0_6 = py6 * z0_6 + 30.0 * py4 * z0_5 + 180.0 * py2 * z0_4 + 120.0 * py0 * z0_3;
      { const double r = px0 * y6_0_6;
        W[16] = fma(-0.15811388300841894, r, W[16]);
      }
      const double y6_0_7 = py6 * z0_7 + 30.0 * py4 * z0_6 + 180.0 * py2 * z0_5 + 120.0 * py0 * z0_4;
      { const double r = px1 * y6_0_7;
        W[26] = fma(-0.15811388300841894, r, W[26]);
      }
      }
      {
      const double y7_0_7 = py7 * z0_7 + 42.0 * py5 * z0_6 + 420.0 * py3 * z0_5 + 840.0 * py1 * z0_4;
      { const double r = px0 * y7_0_7;
        W[30] = fma(-0.15811388300841894, r, W[30]);
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
      const double z1_7 = pz1 * g7;
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[22] = fma(-0.15811388300841894, r, W[22]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[24] = fma(-0.15811388300841894, r, W[24]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[27] = fma(-0.15811388300841894, r, W[27]);
        W[22] = fma(-0.15811388300841894, r, W[22]);
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
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[31] = fma(-0.15811388300841894, r, W[31]);
        W[24] = fma(-0.15811388300841894, r, W[24]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[17] = fma(-0.15811388300841894, r, W[17]);
      }
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[27] = fma(-0.15811388300841894, r, W[27]);
      }
      }
      {
      const double y6_1_7 = py6 * z1_7 + 30.0 * py4 * z1_6 + 180.0 * py2 * z1_5 + 120.0 * py0 * z1_4;
      { const double r = px0 * y6_1_7;
        W[31] = fma(-0.15811388300841894, r, W[31]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[25] = fma(-0.15811388300841894, r, W[25]);
        W[20] = fma(0.6324555320336758, r, W[20]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[28] = fma(-0.15811388300841894, r, W[28]);
        W[21] = fma(0.6324555320336758, r, W[21]);
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
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[32] = fma(-0.15811388300841894, r, W[32]);
        W[25] = fma(-0.15811388300841894, r, W[25]);
        W[23] = fma(0.6324555320336758, r, W[23]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[16] = fma(0.6324555320336758, r, W[16]);
      }
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[28] = fma(-0.15811388300841894, r, W[28]);
        W[26] = fma(0.6324555320336758, r, W[26]);
      }
      }
      {
      const double y5_2_7 = py5 * z2_7 + 20.0 * py3 * z2_6 + 60.0 * py1 * z2_5;
      { const double r = px0 * y5_2_7;
        W[32] = fma(-0.15811388300841894, r, W[32]);
        W[30] = fma(0.6324555320336758, r, W[30]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
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
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[29] = fma(-0.15811388300841894, r, W[29]);
        W[22] = fma(0.6324555320336758, r, W[22]);
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
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[33] = fma(-0.15811388300841894, r, W[33]);
        W[24] = fma(0.6324555320336758, r, W[24]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[17] = fma(0.6324555320336758, r, W[17]);
      }
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[29] = fma(-0.15811388300841894, r, W[29]);
        W[27] = fma(0.6324555320336758, r, W[27]);
      }
      }
      {
      const double y4_3_7 = py4 * z3_7 + 12.0 * py2 * z3_6 + 12.0 * py0 * z3_5;
      { const double r = px0 * y4_3_7;
        W[33] = fma(-0.15811388300841894, r, W[33]);
        W[31] = fma(0.6324555320336758, r, W[31]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
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
      const double y1_4_7 = py1 * z4_7;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[34] = fma(-0.15811388300841894, r, W[34]);
        W[25] = fma(0.6324555320336758, r, W[25]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[18] = fma(0.6324555320336758, r, W[18]);
      }
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[28] = fma(0.6324555320336758, r, W[28]);
      }
      }
      {
      const double y3_4_7 = py3 * z4_7 + 6.0 * py1 * z4_6;
      { const double r = px0 * y3_4_7;
        W[34] = fma(-0.15811388300841894, r, W[34]);
        W[32] = fma(0.6324555320336758, r, W[32]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      }
      {
      const double y1_5_6 = py1 * z5_6;
      { const double r = px0 * y1_5_6;
        W[19] = fma(0.6324555320336758, r, W[19]);
      }
      const double y1_5_7 = py1 * z5_7;
      { const double r = px1 * y1_5_7;
        W[29] = fma(0.6324555320336758, r, W[29]);
      }
      }
      {
      const double y2_5_7 = py2 * z5_7 + 2.0 * py0 * z5_6;
      { const double r = px0 * y2_5_7;
        W[33] = fma(0.6324555320336758, r, W[33]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      }
      {
      const double y1_6_7 = py1 * z6_7;
      { const double r = px0 * y1_6_7;
        W[34] = fma(0.6324555320336758, r, W[34]);
      }
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 2] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 2] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 2] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 2] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 2] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 2] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 2] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 2] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 2] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 2] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 2] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 2] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 2] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 2] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 2] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 2] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 2] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 2] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 2] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 2] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 2] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 2] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 2] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 2] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 2] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 3: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double z1_7 = pz1 * g7;
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
      const double y0_1_7 = py0 * z1_7;
      { const double r = px6 * y0_1_7 + 30.0 * px4 * y0_1_6 + 180.0 * px2 * y0_1_5 + 120.0 * px0 * y0_1_4;
        W[20] = fma(-0.38729833462074165, r, W[20]);
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[21] = fma(-0.38729833462074165, r, W[21]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[23] = fma(-0.38729833462074165, r, W[23]);
        W[20] = fma(-0.38729833462074165, r, W[20]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[26] = fma(-0.38729833462074165, r, W[26]);
        W[21] = fma(-0.38729833462074165, r, W[21]);
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
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[30] = fma(-0.38729833462074165, r, W[30]);
        W[23] = fma(-0.38729833462074165, r, W[23]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[16] = fma(-0.38729833462074165, r, W[16]);
      }
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[26] = fma(-0.38729833462074165, r, W[26]);
      }
      }
      {
      const double y6_1_7 = py6 * z1_7 + 30.0 * py4 * z1_6 + 180.0 * py2 * z1_5 + 120.0 * py0 * z1_4;
      { const double r = px0 * y6_1_7;
        W[30] = fma(-0.38729833462074165, r, W[30]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y0_2_7 = py0 * z2_7;
      { const double r = px5 * y0_2_7 + 20.0 * px3 * y0_2_6 + 60.0 * px1 * y0_2_5;
        W[22] = fma(-0.38729833462074165, r, W[22]);
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[24] = fma(-0.38729833462074165, r, W[24]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[27] = fma(-0.38729833462074165, r, W[27]);
        W[22] = fma(-0.38729833462074165, r, W[22]);
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
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[31] = fma(-0.38729833462074165, r, W[31]);
        W[24] = fma(-0.38729833462074165, r, W[24]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[17] = fma(-0.38729833462074165, r, W[17]);
      }
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[27] = fma(-0.38729833462074165, r, W[27]);
      }
      }
      {
      const double y5_2_7 = py5 * z2_7 + 20.0 * py3 * z2_6 + 60.0 * py1 * z2_5;
      { const double r = px0 * y5_2_7;
        W[31] = fma(-0.38729833462074165, r, W[31]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
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
      const double y0_3_7 = py0 * z3_7;
      { const double r = px4 * y0_3_7 + 12.0 * px2 * y0_3_6 + 12.0 * px0 * y0_3_5;
        W[25] = fma(-0.38729833462074165, r, W[25]);
        W[20] = fma(0.2581988897471611, r, W[20]);
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
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[28] = fma(-0.38729833462074165, r, W[28]);
        W[21] = fma(0.2581988897471611, r, W[21]);
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
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[32] = fma(-0.38729833462074165, r, W[32]);
        W[25] = fma(-0.38729833462074165, r, W[25]);
        W[23] = fma(0.2581988897471611, r, W[23]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[18] = fma(-0.38729833462074165, r, W[18]);
        W[16] = fma(0.2581988897471611, r, W[16]);
      }
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[28] = fma(-0.38729833462074165, r, W[28]);
        W[26] = fma(0.2581988897471611, r, W[26]);
      }
      }
      {
      const double y4_3_7 = py4 * z3_7 + 12.0 * py2 * z3_6 + 12.0 * py0 * z3_5;
      { const double r = px0 * y4_3_7;
        W[32] = fma(-0.38729833462074165, r, W[32]);
        W[30] = fma(0.2581988897471611, r, W[30]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
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
      const double y0_4_7 = py0 * z4_7;
      { const double r = px3 * y0_4_7 + 6.0 * px1 * y0_4_6;
        W[29] = fma(-0.38729833462074165, r, W[29]);
        W[22] = fma(0.2581988897471611, r, W[22]);
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
      const double y1_4_7 = py1 * z4_7;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[33] = fma(-0.38729833462074165, r, W[33]);
        W[24] = fma(0.2581988897471611, r, W[24]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[19] = fma(-0.38729833462074165, r, W[19]);
        W[17] = fma(0.2581988897471611, r, W[17]);
      }
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[29] = fma(-0.38729833462074165, r, W[29]);
        W[27] = fma(0.2581988897471611, r, W[27]);
      }
      }
      {
      const double y3_4_7 = py3 * z4_7 + 6.0 * py1 * z4_6;
      { const double r = px0 * y3_4_7;
        W[33] = fma(-0.38729833462074165, r, W[33]);
        W[31] = fma(0.2581988897471611, r, W[31]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      const double y0_5_5 = py0 * z5_5;
      { const double r = px0 * y0_5_5;
        W[9] = fma(0.2581988897471611, r, W[9]);
      }
      const double y0_5_6 = py0 * z5_6;
      { const double r = px1 * y0_5_6;
        W[15] = fma(0.2581988897471611, r, W[15]);
      }
      const double y0_5_7 = py0 * z5_7;
      { const double r = px2 * y0_5_7 + 2.0 * px0 * y0_5_6;
        W[34] = fma(-0.38729833462074165, r, W[34]);
        W[25] = fma(0.2581988897471611, r, W[25]);
      }
      }
      {
      const double y1_5_6 = py1 * z5_6;
      { const double r = px0 * y1_5_6;
        W[18] = fma(0.2581988897471611, r, W[18]);
      }
      const double y1_5_7 = py1 * z5_7;
      { const double r = px1 * y1_5_7;
        W[28] = fma(0.2581988897471611, r, W[28]);
      }
      }
      {
      const double y2_5_7 = py2 * z5_7 + 2.0 * py0 * z5_6;
      { const double r = px0 * y2_5_7;
        W[34] = fma(-0.38729833462074165, r, W[34]);
        W[32] = fma(0.2581988897471611, r, W[32]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      const double y0_6_6 = py0 * z6_6;
      { const double r = px0 * y0_6_6;
        W[19] = fma(0.2581988897471611, r, W[19]);
      }
      const double y0_6_7 = py0 * z6_7;
      { const double r = px1 * y0_6_7;
        W[29] = fma(0.2581988897471611, r, W[29]);
      }
      }
      {
      const double y1_6_7 = py1 * z6_7;
      { const double r = px0 * y1_6_7;
        W[33] = fma(0.2581988897471611, r, W[33]);
      }
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
      {
      const double y0_7_7 = py0 * z7_7;
      { const double r = px0 * y0_7_7;
        W[34] = fma(0.2581988897471611, r, W[34]);
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 3] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 3] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 3] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 3] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 3] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 3] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 3] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 3] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 3] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 3] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 3] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 3] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 3] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 3] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 3] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 3] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 3] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 3] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 3] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 3] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 3] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 3] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 3] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 3] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 3] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 4: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double y0_0_7 = py0 * z0_7;
      { const double r = px7 * y0_0_7 + 42.0 * px5 * y0_0_6 + 420.0 * px3 * y0_0_5 + 840.0 * px1 * y0_0_4;
        W[20] = fma(-0.15811388300841894, r, W[20]);
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
      const double y1_0_7 = py1 * z0_7;
      { const double r = px6 * y1_0_7 + 30.0 * px4 * y1_0_6 + 180.0 * px2 * y1_0_5 + 120.0 * px0 * y1_0_4;
        W[21] = fma(-0.15811388300841894, r, W[21]);
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
      const double y2_0_7 = py2 * z0_7 + 2.0 * py0 * z0_6;
      { const double r = px5 * y2_0_7 + 20.0 * px3 * y2_0_6 + 60.0 * px1 * y2_0_5;
        W[23] = fma(-0.15811388300841894, r, W[23]);
        W[20] = fma(-0.15811388300841894, r, W[20]);
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
      const double y3_0_7 = py3 * z0_7 + 6.0 * py1 * z0_6;
      { const double r = px4 * y3_0_7 + 12.0 * px2 * y3_0_6 + 12.0 * px0 * y3_0_5;
        W[26] = fma(-0.15811388300841894, r, W[26]);
        W[21] = fma(-0.15811388300841894, r, W[21]);
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
      const double y4_0_7 = py4 * z0_7 + 12.0 * py2 * z0_6 + 12.0 * py0 * z0_5;
      { const double r = px3 * y4_0_7 + 6.0 * px1 * y4_0_6;
        W[30] = fma(-0.15811388300841894, r, W[30]);
        W[23] = fma(-0.15811388300841894, r, W[23]);
      }
      }
      {
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[16] = fma(-0.15811388300841894, r, W[16]);
      }
      const double y5_0_7 = py5 * z0_7 + 20.0 * py3 * z0_6 + 60.0 * py1 * z0_5;
      { const double r = px2 * y5_0_7 + 2.0 * px0 * y5_0_6;
        W[26] = fma(-0.15811388300841894, r, W[26]);
      }
      }
      {
      const double y6_0_7 = py6 * z0_7 + 30.0 * py4 * z0_6 + 180.0 * py2 * z0_5 + 120.0 * py0 * z0_4;
      { const double r = px1 * y6_0_7;
        W[30] = fma(-0.15811388300841894, r, W[30]);
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
      const double z1_7 = pz1 * g7;
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
      const double y0_1_7 = py0 * z1_7;
      { const double r = px6 * y0_1_7 + 30.0 * px4 * y0_1_6 + 180.0 * px2 * y0_1_5 + 120.0 * px0 * y0_1_4;
        W[22] = fma(-0.15811388300841894, r, W[22]);
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[24] = fma(-0.15811388300841894, r, W[24]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[27] = fma(-0.15811388300841894, r, W[27]);
        W[22] = fma(-0.15811388300841894, r, W[22]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[31] = fma(-0.15811388300841894, r, W[31]);
        W[24] = fma(-0.15811388300841894, r, W[24]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[17] = fma(-0.15811388300841894, r, W[17]);
      }
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[27] = fma(-0.15811388300841894, r, W[27]);
      }
      }
      {
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[31] = fma(-0.15811388300841894, r, W[31]);
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
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y0_2_7 = py0 * z2_7;
      { const double r = px5 * y0_2_7 + 20.0 * px3 * y0_2_6 + 60.0 * px1 * y0_2_5;
        W[25] = fma(-0.15811388300841894, r, W[25]);
        W[20] = fma(0.6324555320336758, r, W[20]);
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[28] = fma(-0.15811388300841894, r, W[28]);
        W[21] = fma(0.6324555320336758, r, W[21]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[32] = fma(-0.15811388300841894, r, W[32]);
        W[25] = fma(-0.15811388300841894, r, W[25]);
        W[23] = fma(0.6324555320336758, r, W[23]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[18] = fma(-0.15811388300841894, r, W[18]);
        W[16] = fma(0.6324555320336758, r, W[16]);
      }
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[28] = fma(-0.15811388300841894, r, W[28]);
        W[26] = fma(0.6324555320336758, r, W[26]);
      }
      }
      {
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[32] = fma(-0.15811388300841894, r, W[32]);
        W[30] = fma(0.6324555320336758, r, W[30]);
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
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
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
      const double y0_3_7 = py0 * z3_7;
      { const double r = px4 * y0_3_7 + 12.0 * px2 * y0_3_6 + 12.0 * px0 * y0_3_5;
        W[29] = fma(-0.15811388300841894, r, W[29]);
        W[22] = fma(0.6324555320336758, r, W[22]);
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
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[33] = fma(-0.15811388300841894, r, W[33]);
        W[24] = fma(0.6324555320336758, r, W[24]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[19] = fma(-0.15811388300841894, r, W[19]);
        W[17] = fma(0.6324555320336758, r, W[17]);
      }
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[29] = fma(-0.15811388300841894, r, W[29]);
        W[27] = fma(0.6324555320336758, r, W[27]);
      }
      }
      {
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[33] = fma(-0.15811388300841894, r, W[33]);
        W[31] = fma(0.6324555320336758, r, W[31]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
      {
      const double y0_4_5 = py0 * z4_5;
      { const double r = px1 * y0_4_5;
        W[9] = fma(0.6324555320336758, r, W[9]);
      }
      const double y0_4_6 = py0 * z4_6;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[15] = fma(0.6324555320336758, r, W[15]);
      }
      const double y0_4_7 = py0 * z4_7;
      { const double r = px3 * y0_4_7 + 6.0 * px1 * y0_4_6;
        W[34] = fma(-0.15811388300841894, r, W[34]);
        W[25] = fma(0.6324555320336758, r, W[25]);
      }
      }
      {
      const double y1_4_6 = py1 * z4_6;
      { const double r = px1 * y1_4_6;
        W[18] = fma(0.6324555320336758, r, W[18]);
      }
      const double y1_4_7 = py1 * z4_7;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[28] = fma(0.6324555320336758, r, W[28]);
      }
      }
      {
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[34] = fma(-0.15811388300841894, r, W[34]);
        W[32] = fma(0.6324555320336758, r, W[32]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      const double y0_5_6 = py0 * z5_6;
      { const double r = px1 * y0_5_6;
        W[19] = fma(0.6324555320336758, r, W[19]);
      }
      const double y0_5_7 = py0 * z5_7;
      { const double r = px2 * y0_5_7 + 2.0 * px0 * y0_5_6;
        W[29] = fma(0.6324555320336758, r, W[29]);
      }
      }
      {
      const double y1_5_7 = py1 * z5_7;
      { const double r = px1 * y1_5_7;
        W[33] = fma(0.6324555320336758, r, W[33]);
      }
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      const double y0_6_7 = py0 * z6_7;
      { const double r = px1 * y0_6_7;
        W[34] = fma(0.6324555320336758, r, W[34]);
      }
      }
      {
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 4] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 4] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 4] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 4] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 4] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 4] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 4] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 4] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 4] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 4] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 4] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 4] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 4] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 4] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 4] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 4] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 4] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 4] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 4] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 4] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 4] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 4] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 4] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 4] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 4] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 5: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double z1_7 = pz1 * g7;
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
      const double y0_1_7 = py0 * z1_7;
      { const double r = px6 * y0_1_7 + 30.0 * px4 * y0_1_6 + 180.0 * px2 * y0_1_5 + 120.0 * px0 * y0_1_4;
        W[20] = fma(0.5, r, W[20]);
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[21] = fma(0.5, r, W[21]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[23] = fma(0.5, r, W[23]);
        W[20] = fma(-0.5, r, W[20]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[26] = fma(0.5, r, W[26]);
        W[21] = fma(-0.5, r, W[21]);
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
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[30] = fma(0.5, r, W[30]);
        W[23] = fma(-0.5, r, W[23]);
      }
      }
      {
      const double y5_1_6 = py5 * z1_6 + 20.0 * py3 * z1_5 + 60.0 * py1 * z1_4;
      { const double r = px0 * y5_1_6;
        W[16] = fma(-0.5, r, W[16]);
      }
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[26] = fma(-0.5, r, W[26]);
      }
      }
      {
      const double y6_1_7 = py6 * z1_7 + 30.0 * py4 * z1_6 + 180.0 * py2 * z1_5 + 120.0 * py0 * z1_4;
      { const double r = px0 * y6_1_7;
        W[30] = fma(-0.5, r, W[30]);
      }
      }
      }
      {
      const double z2_2 = pz2 * g2 + 2.0 * pz0 * g1;
      const double z2_3 = pz2 * g3 + 2.0 * pz0 * g2;
      const double z2_4 = pz2 * g4 + 2.0 * pz0 * g3;
      const double z2_5 = pz2 * g5 + 2.0 * pz0 * g4;
      const double z2_6 = pz2 * g6 + 2.0 * pz0 * g5;
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y0_2_7 = py0 * z2_7;
      { const double r = px5 * y0_2_7 + 20.0 * px3 * y0_2_6 + 60.0 * px1 * y0_2_5;
        W[22] = fma(0.5, r, W[22]);
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
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[24] = fma(0.5, r, W[24]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[27] = fma(0.5, r, W[27]);
        W[22] = fma(-0.5, r, W[22]);
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
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[31] = fma(0.5, r, W[31]);
        W[24] = fma(-0.5, r, W[24]);
      }
      }
      {
      const double y4_2_6 = py4 * z2_6 + 12.0 * py2 * z2_5 + 12.0 * py0 * z2_4;
      { const double r = px0 * y4_2_6;
        W[17] = fma(-0.5, r, W[17]);
      }
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[27] = fma(-0.5, r, W[27]);
      }
      }
      {
      const double y5_2_7 = py5 * z2_7 + 20.0 * py3 * z2_6 + 60.0 * py1 * z2_5;
      { const double r = px0 * y5_2_7;
        W[31] = fma(-0.5, r, W[31]);
      }
      }
      }
      {
      const double z3_3 = pz3 * g3 + 6.0 * pz1 * g2;
      const double z3_4 = pz3 * g4 + 6.0 * pz1 * g3;
      const double z3_5 = pz3 * g5 + 6.0 * pz1 * g4;
      const double z3_6 = pz3 * g6 + 6.0 * pz1 * g5;
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
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
      const double y0_3_7 = py0 * z3_7;
      { const double r = px4 * y0_3_7 + 12.0 * px2 * y0_3_6 + 12.0 * px0 * y0_3_5;
        W[25] = fma(0.5, r, W[25]);
      }
      }
      {
      const double y1_3_6 = py1 * z3_6;
      const double y1_3_5 = py1 * z3_5;
      { const double r = px2 * y1_3_6 + 2.0 * px0 * y1_3_5;
        W[18] = fma(0.5, r, W[18]);
      }
      const double y1_3_7 = py1 * z3_7;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[28] = fma(0.5, r, W[28]);
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
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[32] = fma(0.5, r, W[32]);
        W[25] = fma(-0.5, r, W[25]);
      }
      }
      {
      const double y3_3_6 = py3 * z3_6 + 6.0 * py1 * z3_5;
      { const double r = px0 * y3_3_6;
        W[18] = fma(-0.5, r, W[18]);
      }
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[28] = fma(-0.5, r, W[28]);
      }
      }
      {
      const double y4_3_7 = py4 * z3_7 + 12.0 * py2 * z3_6 + 12.0 * py0 * z3_5;
      { const double r = px0 * y4_3_7;
        W[32] = fma(-0.5, r, W[32]);
      }
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
      {
      const double y0_4_6 = py0 * z4_6;
      const double y0_4_5 = py0 * z4_5;
      { const double r = px2 * y0_4_6 + 2.0 * px0 * y0_4_5;
        W[19] = fma(0.5, r, W[19]);
      }
      const double y0_4_7 = py0 * z4_7;
      { const double r = px3 * y0_4_7 + 6.0 * px1 * y0_4_6;
        W[29] = fma(0.5, r, W[29]);
      }
      }
      {
      const double y1_4_7 = py1 * z4_7;
      const double y1_4_6 = py1 * z4_6;
      { const double r = px2 * y1_4_7 + 2.0 * px0 * y1_4_6;
        W[33] = fma(0.5, r, W[33]);
      }
      }
      {
      const double y2_4_6 = py2 * z4_6 + 2.0 * py0 * z4_5;
      { const double r = px0 * y2_4_6;
        W[19] = fma(-0.5, r, W[19]);
      }
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[29] = fma(-0.5, r, W[29]);
      }
      }
      {
      const double y3_4_7 = py3 * z4_7 + 6.0 * py1 * z4_6;
      { const double r = px0 * y3_4_7;
        W[33] = fma(-0.5, r, W[33]);
      }
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      const double y0_5_7 = py0 * z5_7;
      const double y0_5_6 = py0 * z5_6;
      { const double r = px2 * y0_5_7 + 2.0 * px0 * y0_5_6;
        W[34] = fma(0.5, r, W[34]);
      }
      }
      {
      }
      {
      const double y2_5_7 = py2 * z5_7 + 2.0 * py0 * z5_6;
      { const double r = px0 * y2_5_7;
        W[34] = fma(-0.5, r, W[34]);
      }
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      }
      {
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 5] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 5] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 5] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 5] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 5] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 5] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 5] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 5] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 5] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 5] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 5] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 5] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 5] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 5] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 5] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 5] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 5] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 5] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 5] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 5] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 5] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 5] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 5] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 5] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 5] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  case 6: {
    double acc[36];
#pragma unroll
    for (int q = 0; q < 36; q++) acc[q] = 0.0;
    for (int ij = 0; ij < npp; ij++) {
      const double* __restrict__ rec = recp + (size_t)ij * 340;
      const double p = rec[0];
      const double PCx = rec[1] - Cx, PCy = rec[2] - Cy, PCz = rec[3] - Cz;
      for (int kp = 0; kp < nprim_c; kp++) {
      const double gam = ket.exps[pc0 + kp];
      const double alpha = p * gam / (p + gam);
      const double T = alpha * (PCx * PCx + PCy * PCy + PCz * PCz);
      double sc = 2.0 * PI * PI * sqrt(PI) / (p * gam * sqrt(p + gam)) * ket.coefs[pc0 + kp];
      { const double i2g = 0.5 / gam; sc *= -i2g * i2g * i2g; }
      double F[8];
      boys_array<7>(T, boys, F);
      const double ma = -alpha;
      const double g0 = sc * F[0];
      sc *= ma; const double g1 = sc * F[1];
      sc *= ma; const double g2 = sc * F[2];
      sc *= ma; const double g3 = sc * F[3];
      sc *= ma; const double g4 = sc * F[4];
      sc *= ma; const double g5 = sc * F[5];
      sc *= ma; const double g6 = sc * F[6];
      sc *= ma; const double g7 = sc * F[7];
      const double px0 = 1.0, px1 = 2.0 * PCx;
      const double px2 = px1 * px1;
      const double px3 = px2 * px1;
      const double px4 = px3 * px1;
      const double px5 = px4 * px1;
      const double px6 = px5 * px1;
      const double px7 = px6 * px1;
      const double py0 = 1.0, py1 = 2.0 * PCy;
      const double py2 = py1 * py1;
      const double py3 = py2 * py1;
      const double py4 = py3 * py1;
      const double py5 = py4 * py1;
      const double py6 = py5 * py1;
      const double py7 = py6 * py1;
      const double pz0 = 1.0, pz1 = 2.0 * PCz;
      const double pz2 = pz1 * pz1;
      const double pz3 = pz2 * pz1;
      const double pz4 = pz3 * pz1;
      const double pz5 = pz4 * pz1;
      const double pz6 = pz5 * pz1;
      const double pz7 = pz6 * pz1;
      double W[35];
#pragma unroll
      for (int q = 0; q < 35; q++) W[q] = 0.0;
      {
      const double z0_0 = pz0 * g0;
      const double z0_1 = pz0 * g1;
      const double z0_2 = pz0 * g2;
      const double z0_3 = pz0 * g3;
      const double z0_4 = pz0 * g4;
      const double z0_5 = pz0 * g5;
      const double z0_6 = pz0 * g6;
      const double z0_7 = pz0 * g7;
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
      const double y0_0_7 = py0 * z0_7;
      { const double r = px7 * y0_0_7 + 42.0 * px5 * y0_0_6 + 420.0 * px3 * y0_0_5 + 840.0 * px1 * y0_0_4;
        W[20] = fma(0.20412414523193148, r, W[20]);
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
      const double y1_0_7 = py1 * z0_7;
      { const double r = px6 * y1_0_7 + 30.0 * px4 * y1_0_6 + 180.0 * px2 * y1_0_5 + 120.0 * px0 * y1_0_4;
        W[21] = fma(0.20412414523193148, r, W[21]);
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
      const double y2_0_7 = py2 * z0_7 + 2.0 * py0 * z0_6;
      { const double r = px5 * y2_0_7 + 20.0 * px3 * y2_0_6 + 60.0 * px1 * y2_0_5;
        W[23] = fma(0.20412414523193148, r, W[23]);
        W[20] = fma(-0.6123724356957945, r, W[20]);
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
      const double y3_0_7 = py3 * z0_7 + 6.0 * py1 * z0_6;
      { const double r = px4 * y3_0_7 + 12.0 * px2 * y3_0_6 + 12.0 * px0 * y3_0_5;
        W[26] = fma(0.20412414523193148, r, W[26]);
        W[21] = fma(-0.6123724356957945, r, W[21]);
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
      const double y4_0_7 = py4 * z0_7 + 12.0 * py2 * z0_6 + 12.0 * py0 * z0_5;
      { const double r = px3 * y4_0_7 + 6.0 * px1 * y4_0_6;
        W[30] = fma(0.20412414523193148, r, W[30]);
        W[23] = fma(-0.6123724356957945, r, W[23]);
      }
      }
      {
      const double y5_0_6 = py5 * z0_6 + 20.0 * py3 * z0_5 + 60.0 * py1 * z0_4;
      { const double r = px1 * y5_0_6;
        W[16] = fma(-0.6123724356957945, r, W[16]);
      }
      const double y5_0_7 = py5 * z0_7 + 20.0 * py3 * z0_6 + 60.0 * py1 * z0_5;
      { const double r = px2 * y5_0_7 + 2.0 * px0 * y5_0_6;
        W[26] = fma(-0.6123724356957945, r, W[26]);
      }
      }
      {
      const double y6_0_7 = py6 * z0_7 + 30.0 * py4 * z0_6 + 180.0 * py2 * z0_5 + 120.0 * py0 * z0_4;
      { const double r = px1 * y6_0_7;
        W[30] = fma(-0.6123724356957945, r, W[30]);
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
      const double z1_7 = pz1 * g7;
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
      const double y0_1_7 = py0 * z1_7;
      { const double r = px6 * y0_1_7 + 30.0 * px4 * y0_1_6 + 180.0 * px2 * y0_1_5 + 120.0 * px0 * y0_1_4;
        W[22] = fma(0.20412414523193148, r, W[22]);
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
      const double y1_1_7 = py1 * z1_7;
      { const double r = px5 * y1_1_7 + 20.0 * px3 * y1_1_6 + 60.0 * px1 * y1_1_5;
        W[24] = fma(0.20412414523193148, r, W[24]);
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
      const double y2_1_7 = py2 * z1_7 + 2.0 * py0 * z1_6;
      { const double r = px4 * y2_1_7 + 12.0 * px2 * y2_1_6 + 12.0 * px0 * y2_1_5;
        W[27] = fma(0.20412414523193148, r, W[27]);
        W[22] = fma(-0.6123724356957945, r, W[22]);
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
      const double y3_1_7 = py3 * z1_7 + 6.0 * py1 * z1_6;
      { const double r = px3 * y3_1_7 + 6.0 * px1 * y3_1_6;
        W[31] = fma(0.20412414523193148, r, W[31]);
        W[24] = fma(-0.6123724356957945, r, W[24]);
      }
      }
      {
      const double y4_1_6 = py4 * z1_6 + 12.0 * py2 * z1_5 + 12.0 * py0 * z1_4;
      { const double r = px1 * y4_1_6;
        W[17] = fma(-0.6123724356957945, r, W[17]);
      }
      const double y4_1_7 = py4 * z1_7 + 12.0 * py2 * z1_6 + 12.0 * py0 * z1_5;
      { const double r = px2 * y4_1_7 + 2.0 * px0 * y4_1_6;
        W[27] = fma(-0.6123724356957945, r, W[27]);
      }
      }
      {
      const double y5_1_7 = py5 * z1_7 + 20.0 * py3 * z1_6 + 60.0 * py1 * z1_5;
      { const double r = px1 * y5_1_7;
        W[31] = fma(-0.6123724356957945, r, W[31]);
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
      const double z2_7 = pz2 * g7 + 2.0 * pz0 * g6;
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
      const double y0_2_7 = py0 * z2_7;
      { const double r = px5 * y0_2_7 + 20.0 * px3 * y0_2_6 + 60.0 * px1 * y0_2_5;
        W[25] = fma(0.20412414523193148, r, W[25]);
      }
      }
      {
      const double y1_2_6 = py1 * z2_6;
      const double y1_2_5 = py1 * z2_5;
      { const double r = px3 * y1_2_6 + 6.0 * px1 * y1_2_5;
        W[18] = fma(0.20412414523193148, r, W[18]);
      }
      const double y1_2_7 = py1 * z2_7;
      { const double r = px4 * y1_2_7 + 12.0 * px2 * y1_2_6 + 12.0 * px0 * y1_2_5;
        W[28] = fma(0.20412414523193148, r, W[28]);
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
      const double y2_2_7 = py2 * z2_7 + 2.0 * py0 * z2_6;
      { const double r = px3 * y2_2_7 + 6.0 * px1 * y2_2_6;
        W[32] = fma(0.20412414523193148, r, W[32]);
        W[25] = fma(-0.6123724356957945, r, W[25]);
      }
      }
      {
      const double y3_2_6 = py3 * z2_6 + 6.0 * py1 * z2_5;
      { const double r = px1 * y3_2_6;
        W[18] = fma(-0.6123724356957945, r, W[18]);
      }
      const double y3_2_7 = py3 * z2_7 + 6.0 * py1 * z2_6;
      { const double r = px2 * y3_2_7 + 2.0 * px0 * y3_2_6;
        W[28] = fma(-0.6123724356957945, r, W[28]);
      }
      }
      {
      const double y4_2_7 = py4 * z2_7 + 12.0 * py2 * z2_6 + 12.0 * py0 * z2_5;
      { const double r = px1 * y4_2_7;
        W[32] = fma(-0.6123724356957945, r, W[32]);
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
      const double z3_7 = pz3 * g7 + 6.0 * pz1 * g6;
      {
      const double y0_3_6 = py0 * z3_6;
      const double y0_3_5 = py0 * z3_5;
      { const double r = px3 * y0_3_6 + 6.0 * px1 * y0_3_5;
        W[19] = fma(0.20412414523193148, r, W[19]);
      }
      const double y0_3_7 = py0 * z3_7;
      { const double r = px4 * y0_3_7 + 12.0 * px2 * y0_3_6 + 12.0 * px0 * y0_3_5;
        W[29] = fma(0.20412414523193148, r, W[29]);
      }
      }
      {
      const double y1_3_7 = py1 * z3_7;
      const double y1_3_6 = py1 * z3_6;
      { const double r = px3 * y1_3_7 + 6.0 * px1 * y1_3_6;
        W[33] = fma(0.20412414523193148, r, W[33]);
      }
      }
      {
      const double y2_3_6 = py2 * z3_6 + 2.0 * py0 * z3_5;
      { const double r = px1 * y2_3_6;
        W[19] = fma(-0.6123724356957945, r, W[19]);
      }
      const double y2_3_7 = py2 * z3_7 + 2.0 * py0 * z3_6;
      { const double r = px2 * y2_3_7 + 2.0 * px0 * y2_3_6;
        W[29] = fma(-0.6123724356957945, r, W[29]);
      }
      }
      {
      const double y3_3_7 = py3 * z3_7 + 6.0 * py1 * z3_6;
      { const double r = px1 * y3_3_7;
        W[33] = fma(-0.6123724356957945, r, W[33]);
      }
      }
      {
      }
      }
      {
      const double z4_4 = pz4 * g4 + 12.0 * pz2 * g3 + 12.0 * pz0 * g2;
      const double z4_5 = pz4 * g5 + 12.0 * pz2 * g4 + 12.0 * pz0 * g3;
      const double z4_6 = pz4 * g6 + 12.0 * pz2 * g5 + 12.0 * pz0 * g4;
      const double z4_7 = pz4 * g7 + 12.0 * pz2 * g6 + 12.0 * pz0 * g5;
      {
      const double y0_4_7 = py0 * z4_7;
      const double y0_4_6 = py0 * z4_6;
      { const double r = px3 * y0_4_7 + 6.0 * px1 * y0_4_6;
        W[34] = fma(0.20412414523193148, r, W[34]);
      }
      }
      {
      }
      {
      const double y2_4_7 = py2 * z4_7 + 2.0 * py0 * z4_6;
      { const double r = px1 * y2_4_7;
        W[34] = fma(-0.6123724356957945, r, W[34]);
      }
      }
      {
      }
      }
      {
      const double z5_5 = pz5 * g5 + 20.0 * pz3 * g4 + 60.0 * pz1 * g3;
      const double z5_6 = pz5 * g6 + 20.0 * pz3 * g5 + 60.0 * pz1 * g4;
      const double z5_7 = pz5 * g7 + 20.0 * pz3 * g6 + 60.0 * pz1 * g5;
      {
      }
      {
      }
      {
      }
      }
      {
      const double z6_6 = pz6 * g6 + 30.0 * pz4 * g5 + 180.0 * pz2 * g4 + 120.0 * pz0 * g3;
      const double z6_7 = pz6 * g7 + 30.0 * pz4 * g6 + 180.0 * pz2 * g5 + 120.0 * pz0 * g4;
      {
      }
      {
      }
      }
      {
      const double z7_7 = pz7 * g7 + 42.0 * pz5 * g6 + 420.0 * pz3 * g5 + 840.0 * pz1 * g4;
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
        acc[0] = fma(e, W[20], acc[0]);
      }
      { const double e = rec[9];
        acc[1] = fma(e, W[0], acc[1]);
      }
      { const double e = rec[10];
        acc[1] = fma(e, W[2], acc[1]);
      }
      { const double e = rec[11];
        acc[1] = fma(e, W[1], acc[1]);
      }
      { const double e = rec[12];
        acc[1] = fma(e, W[5], acc[1]);
      }
      { const double e = rec[13];
        acc[1] = fma(e, W[4], acc[1]);
      }
      { const double e = rec[14];
        acc[1] = fma(e, W[11], acc[1]);
      }
      { const double e = rec[15];
        acc[1] = fma(e, W[10], acc[1]);
      }
      { const double e = rec[16];
        acc[1] = fma(e, W[21], acc[1]);
      }
      { const double e = rec[17];
        acc[2] = fma(e, W[0], acc[2]);
      }
      { const double e = rec[18];
        acc[2] = fma(e, W[3], acc[2]);
      }
      { const double e = rec[19];
        acc[2] = fma(e, W[1], acc[2]);
      }
      { const double e = rec[20];
        acc[2] = fma(e, W[6], acc[2]);
      }
      { const double e = rec[21];
        acc[2] = fma(e, W[4], acc[2]);
      }
      { const double e = rec[22];
        acc[2] = fma(e, W[12], acc[2]);
      }
      { const double e = rec[23];
        acc[2] = fma(e, W[10], acc[2]);
      }
      { const double e = rec[24];
        acc[2] = fma(e, W[22], acc[2]);
      }
      { const double e = rec[25];
        acc[3] = fma(e, W[0], acc[3]);
      }
      { const double e = rec[26];
        acc[3] = fma(e, W[2], acc[3]);
      }
      { const double e = rec[27];
        acc[3] = fma(e, W[7], acc[3]);
      }
      { const double e = rec[28];
        acc[3] = fma(e, W[1], acc[3]);
      }
      { const double e = rec[29];
        acc[3] = fma(e, W[5], acc[3]);
      }
      { const double e = rec[30];
        acc[3] = fma(e, W[13], acc[3]);
      }
      { const double e = rec[31];
        acc[3] = fma(e, W[4], acc[3]);
      }
      { const double e = rec[32];
        acc[3] = fma(e, W[11], acc[3]);
      }
      { const double e = rec[33];
        acc[3] = fma(e, W[23], acc[3]);
      }
      { const double e = rec[34];
        acc[4] = fma(e, W[0], acc[4]);
      }
      { const double e = rec[35];
        acc[4] = fma(e, W[3], acc[4]);
      }
      { const double e = rec[36];
        acc[4] = fma(e, W[2], acc[4]);
      }
      { const double e = rec[37];
        acc[4] = fma(e, W[8], acc[4]);
      }
      { const double e = rec[38];
        acc[4] = fma(e, W[1], acc[4]);
      }
      { const double e = rec[39];
        acc[4] = fma(e, W[6], acc[4]);
      }
      { const double e = rec[40];
        acc[4] = fma(e, W[5], acc[4]);
      }
      { const double e = rec[41];
        acc[4] = fma(e, W[14], acc[4]);
      }
      { const double e = rec[42];
        acc[4] = fma(e, W[4], acc[4]);
      }
      { const double e = rec[43];
        acc[4] = fma(e, W[12], acc[4]);
      }
      { const double e = rec[44];
        acc[4] = fma(e, W[11], acc[4]);
      }
      { const double e = rec[45];
        acc[4] = fma(e, W[24], acc[4]);
      }
      { const double e = rec[46];
        acc[5] = fma(e, W[0], acc[5]);
      }
      { const double e = rec[47];
        acc[5] = fma(e, W[3], acc[5]);
      }
      { const double e = rec[48];
        acc[5] = fma(e, W[9], acc[5]);
      }
      { const double e = rec[49];
        acc[5] = fma(e, W[1], acc[5]);
      }
      { const double e = rec[50];
        acc[5] = fma(e, W[6], acc[5]);
      }
      { const double e = rec[51];
        acc[5] = fma(e, W[15], acc[5]);
      }
      { const double e = rec[52];
        acc[5] = fma(e, W[4], acc[5]);
      }
      { const double e = rec[53];
        acc[5] = fma(e, W[12], acc[5]);
      }
      { const double e = rec[54];
        acc[5] = fma(e, W[25], acc[5]);
      }
      { const double e = rec[55];
        acc[6] = fma(e, W[0], acc[6]);
      }
      { const double e = rec[56];
        acc[6] = fma(e, W[2], acc[6]);
      }
      { const double e = rec[57];
        acc[6] = fma(e, W[1], acc[6]);
      }
      { const double e = rec[58];
        acc[6] = fma(e, W[5], acc[6]);
      }
      { const double e = rec[59];
        acc[6] = fma(e, W[4], acc[6]);
      }
      { const double e = rec[60];
        acc[6] = fma(e, W[11], acc[6]);
      }
      { const double e = rec[61];
        acc[6] = fma(e, W[10], acc[6]);
      }
      { const double e = rec[62];
        acc[6] = fma(e, W[21], acc[6]);
      }
      { const double e = rec[63];
        acc[7] = fma(e, W[0], acc[7]);
      }
      { const double e = rec[64];
        acc[7] = fma(e, W[2], acc[7]);
      }
      { const double e = rec[65];
        acc[7] = fma(e, W[7], acc[7]);
      }
      { const double e = rec[66];
        acc[7] = fma(e, W[1], acc[7]);
      }
      { const double e = rec[67];
        acc[7] = fma(e, W[5], acc[7]);
      }
      { const double e = rec[68];
        acc[7] = fma(e, W[13], acc[7]);
      }
      { const double e = rec[69];
        acc[7] = fma(e, W[4], acc[7]);
      }
      { const double e = rec[70];
        acc[7] = fma(e, W[11], acc[7]);
      }
      { const double e = rec[71];
        acc[7] = fma(e, W[23], acc[7]);
      }
      { const double e = rec[72];
        acc[8] = fma(e, W[0], acc[8]);
      }
      { const double e = rec[73];
        acc[8] = fma(e, W[3], acc[8]);
      }
      { const double e = rec[74];
        acc[8] = fma(e, W[2], acc[8]);
      }
      { const double e = rec[75];
        acc[8] = fma(e, W[8], acc[8]);
      }
      { const double e = rec[76];
        acc[8] = fma(e, W[1], acc[8]);
      }
      { const double e = rec[77];
        acc[8] = fma(e, W[6], acc[8]);
      }
      { const double e = rec[78];
        acc[8] = fma(e, W[5], acc[8]);
      }
      { const double e = rec[79];
        acc[8] = fma(e, W[14], acc[8]);
      }
      { const double e = rec[80];
        acc[8] = fma(e, W[4], acc[8]);
      }
      { const double e = rec[81];
        acc[8] = fma(e, W[12], acc[8]);
      }
      { const double e = rec[82];
        acc[8] = fma(e, W[11], acc[8]);
      }
      { const double e = rec[83];
        acc[8] = fma(e, W[24], acc[8]);
      }
      { const double e = rec[84];
        acc[9] = fma(e, W[0], acc[9]);
      }
      { const double e = rec[85];
        acc[9] = fma(e, W[2], acc[9]);
      }
      { const double e = rec[86];
        acc[9] = fma(e, W[7], acc[9]);
      }
      { const double e = rec[87];
        acc[9] = fma(e, W[16], acc[9]);
      }
      { const double e = rec[88];
        acc[9] = fma(e, W[1], acc[9]);
      }
      { const double e = rec[89];
        acc[9] = fma(e, W[5], acc[9]);
      }
      { const double e = rec[90];
        acc[9] = fma(e, W[13], acc[9]);
      }
      { const double e = rec[91];
        acc[9] = fma(e, W[26], acc[9]);
      }
      { const double e = rec[92];
        acc[10] = fma(e, W[0], acc[10]);
      }
      { const double e = rec[93];
        acc[10] = fma(e, W[3], acc[10]);
      }
      { const double e = rec[94];
        acc[10] = fma(e, W[2], acc[10]);
      }
      { const double e = rec[95];
        acc[10] = fma(e, W[8], acc[10]);
      }
      { const double e = rec[96];
        acc[10] = fma(e, W[7], acc[10]);
      }
      { const double e = rec[97];
        acc[10] = fma(e, W[17], acc[10]);
      }
      { const double e = rec[98];
        acc[10] = fma(e, W[1], acc[10]);
      }
      { const double e = rec[99];
        acc[10] = fma(e, W[6], acc[10]);
      }
      { const double e = rec[100];
        acc[10] = fma(e, W[5], acc[10]);
      }
      { const double e = rec[101];
        acc[10] = fma(e, W[14], acc[10]);
      }
      { const double e = rec[102];
        acc[10] = fma(e, W[13], acc[10]);
      }
      { const double e = rec[103];
        acc[10] = fma(e, W[27], acc[10]);
      }
      { const double e = rec[104];
        acc[11] = fma(e, W[0], acc[11]);
      }
      { const double e = rec[105];
        acc[11] = fma(e, W[3], acc[11]);
      }
      { const double e = rec[106];
        acc[11] = fma(e, W[9], acc[11]);
      }
      { const double e = rec[107];
        acc[11] = fma(e, W[2], acc[11]);
      }
      { const double e = rec[108];
        acc[11] = fma(e, W[8], acc[11]);
      }
      { const double e = rec[109];
        acc[11] = fma(e, W[18], acc[11]);
      }
      { const double e = rec[110];
        acc[11] = fma(e, W[1], acc[11]);
      }
      { const double e = rec[111];
        acc[11] = fma(e, W[6], acc[11]);
      }
      { const double e = rec[112];
        acc[11] = fma(e, W[15], acc[11]);
      }
      { const double e = rec[113];
        acc[11] = fma(e, W[5], acc[11]);
      }
      { const double e = rec[114];
        acc[11] = fma(e, W[14], acc[11]);
      }
      { const double e = rec[115];
        acc[11] = fma(e, W[28], acc[11]);
      }
      { const double e = rec[116];
        acc[12] = fma(e, W[0], acc[12]);
      }
      { const double e = rec[117];
        acc[12] = fma(e, W[3], acc[12]);
      }
      { const double e = rec[118];
        acc[12] = fma(e, W[1], acc[12]);
      }
      { const double e = rec[119];
        acc[12] = fma(e, W[6], acc[12]);
      }
      { const double e = rec[120];
        acc[12] = fma(e, W[4], acc[12]);
      }
      { const double e = rec[121];
        acc[12] = fma(e, W[12], acc[12]);
      }
      { const double e = rec[122];
        acc[12] = fma(e, W[10], acc[12]);
      }
      { const double e = rec[123];
        acc[12] = fma(e, W[22], acc[12]);
      }
      { const double e = rec[124];
        acc[13] = fma(e, W[0], acc[13]);
      }
      { const double e = rec[125];
        acc[13] = fma(e, W[3], acc[13]);
      }
      { const double e = rec[126];
        acc[13] = fma(e, W[2], acc[13]);
      }
      { const double e = rec[127];
        acc[13] = fma(e, W[8], acc[13]);
      }
      { const double e = rec[128];
        acc[13] = fma(e, W[1], acc[13]);
      }
      { const double e = rec[129];
        acc[13] = fma(e, W[6], acc[13]);
      }
      { const double e = rec[130];
        acc[13] = fma(e, W[5], acc[13]);
      }
      { const double e = rec[131];
        acc[13] = fma(e, W[14], acc[13]);
      }
      { const double e = rec[132];
        acc[13] = fma(e, W[4], acc[13]);
      }
      { const double e = rec[133];
        acc[13] = fma(e, W[12], acc[13]);
      }
      { const double e = rec[134];
        acc[13] = fma(e, W[11], acc[13]);
      }
      { const double e = rec[135];
        acc[13] = fma(e, W[24], acc[13]);
      }
      { const double e = rec[136];
        acc[14] = fma(e, W[0], acc[14]);
      }
      { const double e = rec[137];
        acc[14] = fma(e, W[3], acc[14]);
      }
      { const double e = rec[138];
        acc[14] = fma(e, W[9], acc[14]);
      }
      { const double e = rec[139];
        acc[14] = fma(e, W[1], acc[14]);
      }
      { const double e = rec[140];
        acc[14] = fma(e, W[6], acc[14]);
      }
      { const double e = rec[141];
        acc[14] = fma(e, W[15], acc[14]);
      }
      { const double e = rec[142];
        acc[14] = fma(e, W[4], acc[14]);
      }
      { const double e = rec[143];
        acc[14] = fma(e, W[12], acc[14]);
      }
      { const double e = rec[144];
        acc[14] = fma(e, W[25], acc[14]);
      }
      { const double e = rec[145];
        acc[15] = fma(e, W[0], acc[15]);
      }
      { const double e = rec[146];
        acc[15] = fma(e, W[3], acc[15]);
      }
      { const double e = rec[147];
        acc[15] = fma(e, W[2], acc[15]);
      }
      { const double e = rec[148];
        acc[15] = fma(e, W[8], acc[15]);
      }
      { const double e = rec[149];
        acc[15] = fma(e, W[7], acc[15]);
      }
      { const double e = rec[150];
        acc[15] = fma(e, W[17], acc[15]);
      }
      { const double e = rec[151];
        acc[15] = fma(e, W[1], acc[15]);
      }
      { const double e = rec[152];
        acc[15] = fma(e, W[6], acc[15]);
      }
      { const double e = rec[153];
        acc[15] = fma(e, W[5], acc[15]);
      }
      { const double e = rec[154];
        acc[15] = fma(e, W[14], acc[15]);
      }
      { const double e = rec[155];
        acc[15] = fma(e, W[13], acc[15]);
      }
      { const double e = rec[156];
        acc[15] = fma(e, W[27], acc[15]);
      }
      { const double e = rec[157];
        acc[16] = fma(e, W[0], acc[16]);
      }
      { const double e = rec[158];
        acc[16] = fma(e, W[3], acc[16]);
      }
      { const double e = rec[159];
        acc[16] = fma(e, W[9], acc[16]);
      }
      { const double e = rec[160];
        acc[16] = fma(e, W[2], acc[16]);
      }
      { const double e = rec[161];
        acc[16] = fma(e, W[8], acc[16]);
      }
      { const double e = rec[162];
        acc[16] = fma(e, W[18], acc[16]);
      }
      { const double e = rec[163];
        acc[16] = fma(e, W[1], acc[16]);
      }
      { const double e = rec[164];
        acc[16] = fma(e, W[6], acc[16]);
      }
      { const double e = rec[165];
        acc[16] = fma(e, W[15], acc[16]);
      }
      { const double e = rec[166];
        acc[16] = fma(e, W[5], acc[16]);
      }
      { const double e = rec[167];
        acc[16] = fma(e, W[14], acc[16]);
      }
      { const double e = rec[168];
        acc[16] = fma(e, W[28], acc[16]);
      }
      { const double e = rec[169];
        acc[17] = fma(e, W[0], acc[17]);
      }
      { const double e = rec[170];
        acc[17] = fma(e, W[3], acc[17]);
      }
      { const double e = rec[171];
        acc[17] = fma(e, W[9], acc[17]);
      }
      { const double e = rec[172];
        acc[17] = fma(e, W[19], acc[17]);
      }
      { const double e = rec[173];
        acc[17] = fma(e, W[1], acc[17]);
      }
      { const double e = rec[174];
        acc[17] = fma(e, W[6], acc[17]);
      }
      { const double e = rec[175];
        acc[17] = fma(e, W[15], acc[17]);
      }
      { const double e = rec[176];
        acc[17] = fma(e, W[29], acc[17]);
      }
      { const double e = rec[177];
        acc[18] = fma(e, W[0], acc[18]);
      }
      { const double e = rec[178];
        acc[18] = fma(e, W[2], acc[18]);
      }
      { const double e = rec[179];
        acc[18] = fma(e, W[7], acc[18]);
      }
      { const double e = rec[180];
        acc[18] = fma(e, W[1], acc[18]);
      }
      { const double e = rec[181];
        acc[18] = fma(e, W[5], acc[18]);
      }
      { const double e = rec[182];
        acc[18] = fma(e, W[13], acc[18]);
      }
      { const double e = rec[183];
        acc[18] = fma(e, W[4], acc[18]);
      }
      { const double e = rec[184];
        acc[18] = fma(e, W[11], acc[18]);
      }
      { const double e = rec[185];
        acc[18] = fma(e, W[23], acc[18]);
      }
      { const double e = rec[186];
        acc[19] = fma(e, W[0], acc[19]);
      }
      { const double e = rec[187];
        acc[19] = fma(e, W[2], acc[19]);
      }
      { const double e = rec[188];
        acc[19] = fma(e, W[7], acc[19]);
      }
      { const double e = rec[189];
        acc[19] = fma(e, W[16], acc[19]);
      }
      { const double e = rec[190];
        acc[19] = fma(e, W[1], acc[19]);
      }
      { const double e = rec[191];
        acc[19] = fma(e, W[5], acc[19]);
      }
      { const double e = rec[192];
        acc[19] = fma(e, W[13], acc[19]);
      }
      { const double e = rec[193];
        acc[19] = fma(e, W[26], acc[19]);
      }
      { const double e = rec[194];
        acc[20] = fma(e, W[0], acc[20]);
      }
      { const double e = rec[195];
        acc[20] = fma(e, W[3], acc[20]);
      }
      { const double e = rec[196];
        acc[20] = fma(e, W[2], acc[20]);
      }
      { const double e = rec[197];
        acc[20] = fma(e, W[8], acc[20]);
      }
      { const double e = rec[198];
        acc[20] = fma(e, W[7], acc[20]);
      }
      { const double e = rec[199];
        acc[20] = fma(e, W[17], acc[20]);
      }
      { const double e = rec[200];
        acc[20] = fma(e, W[1], acc[20]);
      }
      { const double e = rec[201];
        acc[20] = fma(e, W[6], acc[20]);
      }
      { const double e = rec[202];
        acc[20] = fma(e, W[5], acc[20]);
      }
      { const double e = rec[203];
        acc[20] = fma(e, W[14], acc[20]);
      }
      { const double e = rec[204];
        acc[20] = fma(e, W[13], acc[20]);
      }
      { const double e = rec[205];
        acc[20] = fma(e, W[27], acc[20]);
      }
      { const double e = rec[206];
        acc[21] = fma(e, W[0], acc[21]);
      }
      { const double e = rec[207];
        acc[21] = fma(e, W[2], acc[21]);
      }
      { const double e = rec[208];
        acc[21] = fma(e, W[7], acc[21]);
      }
      { const double e = rec[209];
        acc[21] = fma(e, W[16], acc[21]);
      }
      { const double e = rec[210];
        acc[21] = fma(e, W[30], acc[21]);
      }
      { const double e = rec[211];
        acc[22] = fma(e, W[0], acc[22]);
      }
      { const double e = rec[212];
        acc[22] = fma(e, W[3], acc[22]);
      }
      { const double e = rec[213];
        acc[22] = fma(e, W[2], acc[22]);
      }
      { const double e = rec[214];
        acc[22] = fma(e, W[8], acc[22]);
      }
      { const double e = rec[215];
        acc[22] = fma(e, W[7], acc[22]);
      }
      { const double e = rec[216];
        acc[22] = fma(e, W[17], acc[22]);
      }
      { const double e = rec[217];
        acc[22] = fma(e, W[16], acc[22]);
      }
      { const double e = rec[218];
        acc[22] = fma(e, W[31], acc[22]);
      }
      { const double e = rec[219];
        acc[23] = fma(e, W[0], acc[23]);
      }
      { const double e = rec[220];
        acc[23] = fma(e, W[3], acc[23]);
      }
      { const double e = rec[221];
        acc[23] = fma(e, W[9], acc[23]);
      }
      { const double e = rec[222];
        acc[23] = fma(e, W[2], acc[23]);
      }
      { const double e = rec[223];
        acc[23] = fma(e, W[8], acc[23]);
      }
      { const double e = rec[224];
        acc[23] = fma(e, W[18], acc[23]);
      }
      { const double e = rec[225];
        acc[23] = fma(e, W[7], acc[23]);
      }
      { const double e = rec[226];
        acc[23] = fma(e, W[17], acc[23]);
      }
      { const double e = rec[227];
        acc[23] = fma(e, W[32], acc[23]);
      }
      { const double e = rec[228];
        acc[24] = fma(e, W[0], acc[24]);
      }
      { const double e = rec[229];
        acc[24] = fma(e, W[3], acc[24]);
      }
      { const double e = rec[230];
        acc[24] = fma(e, W[2], acc[24]);
      }
      { const double e = rec[231];
        acc[24] = fma(e, W[8], acc[24]);
      }
      { const double e = rec[232];
        acc[24] = fma(e, W[1], acc[24]);
      }
      { const double e = rec[233];
        acc[24] = fma(e, W[6], acc[24]);
      }
      { const double e = rec[234];
        acc[24] = fma(e, W[5], acc[24]);
      }
      { const double e = rec[235];
        acc[24] = fma(e, W[14], acc[24]);
      }
      { const double e = rec[236];
        acc[24] = fma(e, W[4], acc[24]);
      }
      { const double e = rec[237];
        acc[24] = fma(e, W[12], acc[24]);
      }
      { const double e = rec[238];
        acc[24] = fma(e, W[11], acc[24]);
      }
      { const double e = rec[239];
        acc[24] = fma(e, W[24], acc[24]);
      }
      { const double e = rec[240];
        acc[25] = fma(e, W[0], acc[25]);
      }
      { const double e = rec[241];
        acc[25] = fma(e, W[3], acc[25]);
      }
      { const double e = rec[242];
        acc[25] = fma(e, W[2], acc[25]);
      }
      { const double e = rec[243];
        acc[25] = fma(e, W[8], acc[25]);
      }
      { const double e = rec[244];
        acc[25] = fma(e, W[7], acc[25]);
      }
      { const double e = rec[245];
        acc[25] = fma(e, W[17], acc[25]);
      }
      { const double e = rec[246];
        acc[25] = fma(e, W[1], acc[25]);
      }
      { const double e = rec[247];
        acc[25] = fma(e, W[6], acc[25]);
      }
      { const double e = rec[248];
        acc[25] = fma(e, W[5], acc[25]);
      }
      { const double e = rec[249];
        acc[25] = fma(e, W[14], acc[25]);
      }
      { const double e = rec[250];
        acc[25] = fma(e, W[13], acc[25]);
      }
      { const double e = rec[251];
        acc[25] = fma(e, W[27], acc[25]);
      }
      { const double e = rec[252];
        acc[26] = fma(e, W[0], acc[26]);
      }
      { const double e = rec[253];
        acc[26] = fma(e, W[3], acc[26]);
      }
      { const double e = rec[254];
        acc[26] = fma(e, W[9], acc[26]);
      }
      { const double e = rec[255];
        acc[26] = fma(e, W[2], acc[26]);
      }
      { const double e = rec[256];
        acc[26] = fma(e, W[8], acc[26]);
      }
      { const double e = rec[257];
        acc[26] = fma(e, W[18], acc[26]);
      }
      { const double e = rec[258];
        acc[26] = fma(e, W[1], acc[26]);
      }
      { const double e = rec[259];
        acc[26] = fma(e, W[6], acc[26]);
      }
      { const double e = rec[260];
        acc[26] = fma(e, W[15], acc[26]);
      }
      { const double e = rec[261];
        acc[26] = fma(e, W[5], acc[26]);
      }
      { const double e = rec[262];
        acc[26] = fma(e, W[14], acc[26]);
      }
      { const double e = rec[263];
        acc[26] = fma(e, W[28], acc[26]);
      }
      { const double e = rec[264];
        acc[27] = fma(e, W[0], acc[27]);
      }
      { const double e = rec[265];
        acc[27] = fma(e, W[3], acc[27]);
      }
      { const double e = rec[266];
        acc[27] = fma(e, W[2], acc[27]);
      }
      { const double e = rec[267];
        acc[27] = fma(e, W[8], acc[27]);
      }
      { const double e = rec[268];
        acc[27] = fma(e, W[7], acc[27]);
      }
      { const double e = rec[269];
        acc[27] = fma(e, W[17], acc[27]);
      }
      { const double e = rec[270];
        acc[27] = fma(e, W[16], acc[27]);
      }
      { const double e = rec[271];
        acc[27] = fma(e, W[31], acc[27]);
      }
      { const double e = rec[272];
        acc[28] = fma(e, W[0], acc[28]);
      }
      { const double e = rec[273];
        acc[28] = fma(e, W[3], acc[28]);
      }
      { const double e = rec[274];
        acc[28] = fma(e, W[9], acc[28]);
      }
      { const double e = rec[275];
        acc[28] = fma(e, W[2], acc[28]);
      }
      { const double e = rec[276];
        acc[28] = fma(e, W[8], acc[28]);
      }
      { const double e = rec[277];
        acc[28] = fma(e, W[18], acc[28]);
      }
      { const double e = rec[278];
        acc[28] = fma(e, W[7], acc[28]);
      }
      { const double e = rec[279];
        acc[28] = fma(e, W[17], acc[28]);
      }
      { const double e = rec[280];
        acc[28] = fma(e, W[32], acc[28]);
      }
      { const double e = rec[281];
        acc[29] = fma(e, W[0], acc[29]);
      }
      { const double e = rec[282];
        acc[29] = fma(e, W[3], acc[29]);
      }
      { const double e = rec[283];
        acc[29] = fma(e, W[9], acc[29]);
      }
      { const double e = rec[284];
        acc[29] = fma(e, W[19], acc[29]);
      }
      { const double e = rec[285];
        acc[29] = fma(e, W[2], acc[29]);
      }
      { const double e = rec[286];
        acc[29] = fma(e, W[8], acc[29]);
      }
      { const double e = rec[287];
        acc[29] = fma(e, W[18], acc[29]);
      }
      { const double e = rec[288];
        acc[29] = fma(e, W[33], acc[29]);
      }
      { const double e = rec[289];
        acc[30] = fma(e, W[0], acc[30]);
      }
      { const double e = rec[290];
        acc[30] = fma(e, W[3], acc[30]);
      }
      { const double e = rec[291];
        acc[30] = fma(e, W[9], acc[30]);
      }
      { const double e = rec[292];
        acc[30] = fma(e, W[1], acc[30]);
      }
      { const double e = rec[293];
        acc[30] = fma(e, W[6], acc[30]);
      }
      { const double e = rec[294];
        acc[30] = fma(e, W[15], acc[30]);
      }
      { const double e = rec[295];
        acc[30] = fma(e, W[4], acc[30]);
      }
      { const double e = rec[296];
        acc[30] = fma(e, W[12], acc[30]);
      }
      { const double e = rec[297];
        acc[30] = fma(e, W[25], acc[30]);
      }
      { const double e = rec[298];
        acc[31] = fma(e, W[0], acc[31]);
      }
      { const double e = rec[299];
        acc[31] = fma(e, W[3], acc[31]);
      }
      { const double e = rec[300];
        acc[31] = fma(e, W[9], acc[31]);
      }
      { const double e = rec[301];
        acc[31] = fma(e, W[2], acc[31]);
      }
      { const double e = rec[302];
        acc[31] = fma(e, W[8], acc[31]);
      }
      { const double e = rec[303];
        acc[31] = fma(e, W[18], acc[31]);
      }
      { const double e = rec[304];
        acc[31] = fma(e, W[1], acc[31]);
      }
      { const double e = rec[305];
        acc[31] = fma(e, W[6], acc[31]);
      }
      { const double e = rec[306];
        acc[31] = fma(e, W[15], acc[31]);
      }
      { const double e = rec[307];
        acc[31] = fma(e, W[5], acc[31]);
      }
      { const double e = rec[308];
        acc[31] = fma(e, W[14], acc[31]);
      }
      { const double e = rec[309];
        acc[31] = fma(e, W[28], acc[31]);
      }
      { const double e = rec[310];
        acc[32] = fma(e, W[0], acc[32]);
      }
      { const double e = rec[311];
        acc[32] = fma(e, W[3], acc[32]);
      }
      { const double e = rec[312];
        acc[32] = fma(e, W[9], acc[32]);
      }
      { const double e = rec[313];
        acc[32] = fma(e, W[19], acc[32]);
      }
      { const double e = rec[314];
        acc[32] = fma(e, W[1], acc[32]);
      }
      { const double e = rec[315];
        acc[32] = fma(e, W[6], acc[32]);
      }
      { const double e = rec[316];
        acc[32] = fma(e, W[15], acc[32]);
      }
      { const double e = rec[317];
        acc[32] = fma(e, W[29], acc[32]);
      }
      { const double e = rec[318];
        acc[33] = fma(e, W[0], acc[33]);
      }
      { const double e = rec[319];
        acc[33] = fma(e, W[3], acc[33]);
      }
      { const double e = rec[320];
        acc[33] = fma(e, W[9], acc[33]);
      }
      { const double e = rec[321];
        acc[33] = fma(e, W[2], acc[33]);
      }
      { const double e = rec[322];
        acc[33] = fma(e, W[8], acc[33]);
      }
      { const double e = rec[323];
        acc[33] = fma(e, W[18], acc[33]);
      }
      { const double e = rec[324];
        acc[33] = fma(e, W[7], acc[33]);
      }
      { const double e = rec[325];
        acc[33] = fma(e, W[17], acc[33]);
      }
      { const double e = rec[326];
        acc[33] = fma(e, W[32], acc[33]);
      }
      { const double e = rec[327];
        acc[34] = fma(e, W[0], acc[34]);
      }
      { const double e = rec[328];
        acc[34] = fma(e, W[3], acc[34]);
      }
      { const double e = rec[329];
        acc[34] = fma(e, W[9], acc[34]);
      }
      { const double e = rec[330];
        acc[34] = fma(e, W[19], acc[34]);
      }
      { const double e = rec[331];
        acc[34] = fma(e, W[2], acc[34]);
      }
      { const double e = rec[332];
        acc[34] = fma(e, W[8], acc[34]);
      }
      { const double e = rec[333];
        acc[34] = fma(e, W[18], acc[34]);
      }
      { const double e = rec[334];
        acc[34] = fma(e, W[33], acc[34]);
      }
      { const double e = rec[335];
        acc[35] = fma(e, W[0], acc[35]);
      }
      { const double e = rec[336];
        acc[35] = fma(e, W[3], acc[35]);
      }
      { const double e = rec[337];
        acc[35] = fma(e, W[9], acc[35]);
      }
      { const double e = rec[338];
        acc[35] = fma(e, W[19], acc[35]);
      }
      { const double e = rec[339];
        acc[35] = fma(e, W[34], acc[35]);
      }
      }
    }
    if (active) {
    out[(row0 + 0) * ldo + col + 6] = 1.0 * acc[7];
    out[(row0 + 1) * ldo + col + 6] = 1.0 * acc[10];
    out[(row0 + 2) * ldo + col + 6] = -0.2886751345948129 * acc[6] + -0.2886751345948129 * acc[9] + 0.5773502691896258 * acc[11];
    out[(row0 + 3) * ldo + col + 6] = 1.0 * acc[8];
    out[(row0 + 4) * ldo + col + 6] = 0.5 * acc[6] + -0.5 * acc[9];
    out[(row0 + 5) * ldo + col + 6] = 1.0 * acc[25];
    out[(row0 + 6) * ldo + col + 6] = 1.0 * acc[28];
    out[(row0 + 7) * ldo + col + 6] = -0.2886751345948129 * acc[24] + -0.2886751345948129 * acc[27] + 0.5773502691896258 * acc[29];
    out[(row0 + 8) * ldo + col + 6] = 1.0 * acc[26];
    out[(row0 + 9) * ldo + col + 6] = 0.5 * acc[24] + -0.5 * acc[27];
    out[(row0 + 10) * ldo + col + 6] = -0.2886751345948129 * acc[1] + -0.2886751345948129 * acc[19] + 0.5773502691896258 * acc[31];
    out[(row0 + 11) * ldo + col + 6] = -0.2886751345948129 * acc[4] + -0.2886751345948129 * acc[22] + 0.5773502691896258 * acc[34];
    out[(row0 + 12) * ldo + col + 6] = 0.08333333333333336 * acc[0] + 0.08333333333333336 * acc[3] + -0.1666666666666667 * acc[5] + 0.08333333333333336 * acc[18] + 0.08333333333333336 * acc[21] + -0.1666666666666667 * acc[23] + -0.1666666666666667 * acc[30] + -0.1666666666666667 * acc[33] + 0.3333333333333334 * acc[35];
    out[(row0 + 13) * ldo + col + 6] = -0.2886751345948129 * acc[2] + -0.2886751345948129 * acc[20] + 0.5773502691896258 * acc[32];
    out[(row0 + 14) * ldo + col + 6] = -0.14433756729740646 * acc[0] + 0.14433756729740646 * acc[3] + -0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + 0.2886751345948129 * acc[30] + -0.2886751345948129 * acc[33];
    out[(row0 + 15) * ldo + col + 6] = 1.0 * acc[13];
    out[(row0 + 16) * ldo + col + 6] = 1.0 * acc[16];
    out[(row0 + 17) * ldo + col + 6] = -0.2886751345948129 * acc[12] + -0.2886751345948129 * acc[15] + 0.5773502691896258 * acc[17];
    out[(row0 + 18) * ldo + col + 6] = 1.0 * acc[14];
    out[(row0 + 19) * ldo + col + 6] = 0.5 * acc[12] + -0.5 * acc[15];
    out[(row0 + 20) * ldo + col + 6] = 0.5 * acc[1] + -0.5 * acc[19];
    out[(row0 + 21) * ldo + col + 6] = 0.5 * acc[4] + -0.5 * acc[22];
    out[(row0 + 22) * ldo + col + 6] = -0.14433756729740646 * acc[0] + -0.14433756729740646 * acc[3] + 0.2886751345948129 * acc[5] + 0.14433756729740646 * acc[18] + 0.14433756729740646 * acc[21] + -0.2886751345948129 * acc[23];
    out[(row0 + 23) * ldo + col + 6] = 0.5 * acc[2] + -0.5 * acc[20];
    out[(row0 + 24) * ldo + col + 6] = 0.25 * acc[0] + -0.25 * acc[3] + -0.25 * acc[18] + 0.25 * acc[21];
    }
    break;
  }
  default: break;
  }
}

void register_gen3_22(Gen3Table& t) {
  t[2][2][0] = Gen3Launch{&int3c2e_gen_220, 1, 340};
  t[2][2][1] = Gen3Launch{&int3c2e_gen_221, 3, 340};
  t[2][2][2] = Gen3Launch{&int3c2e_gen_222, 5, 340};
  t[2][2][3] = Gen3Launch{&int3c2e_gen_223, 7, 340};
}

}  // namespace dfhip
