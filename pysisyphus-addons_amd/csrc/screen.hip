// QVL screening on the device: pair records and the screened-in prefixes.
#include "hermite.h"
#include "linalg.h"

namespace dfhip {

// Pair record [centre x, y, z, extent, (min a + min b)^(1/4), Q_ab, 0, 0] of
// each shell pair (common/screening.f90:135-146): |d_a d_b|-weighted mean of
// the primitive centres (:235-265) and the largest primitive extent
// sqrt(2/p) erfinv(1 - thresh_ws) + |R - P| (:198-232).  One thread per pair.
__global__ void qvl_pair_kernel(DevShells bra, const int* __restrict__ s1v, const int* __restrict__ s2v,
                                const double* __restrict__ qab, int npairs, double erfi, double* __restrict__ rec) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int s1 = s1v[p], s2 = s2v[p];
  const int na = bra.nprim[s1], nb = bra.nprim[s2], a0 = bra.pptr[s1], b0 = bra.pptr[s2];
  const double A[3] = {bra.xyz[3 * s1], bra.xyz[3 * s1 + 1], bra.xyz[3 * s1 + 2]};
  const double B[3] = {bra.xyz[3 * s2], bra.xyz[3 * s2 + 1], bra.xyz[3 * s2 + 2]};
  double Rc[3] = {0.0, 0.0, 0.0}, den = 0.0, amin = bra.exps[a0], bmin = bra.exps[b0];
  {
#pragma clang fp contract(off)
    for (int i = 0; i < na; i++) {
      const double ax = bra.exps[a0 + i];
      amin = fmin(amin, ax);
      for (int j = 0; j < nb; j++) {
        const double bx = bra.exps[b0 + j], px = ax + bx;
        const double w = fabs(bra.coefs[a0 + i] * bra.coefs[b0 + j]);
        // the primitive centre first, then its weight (R = R + abs_dab*P, :256-259)
        for (int k = 0; k < 3; k++) Rc[k] += w * ((ax * A[k] + bx * B[k]) / px);
        den += w;
      }
    }
    for (int j = 0; j < nb; j++) bmin = fmin(bmin, bra.exps[b0 + j]);
    for (int k = 0; k < 3; k++) Rc[k] /= den;
    double ext = 0.0;
    for (int i = 0; i < na; i++)
      for (int j = 0; j < nb; j++) {
        const double ax = bra.exps[a0 + i], bx = bra.exps[b0 + j], px = ax + bx;
        const double tmp = sqrt(2.0 / px) * erfi;
        double dist = 0.0;
        for (int k = 0; k < 3; k++) {
          const double Pk = (ax * A[k] + bx * B[k]) / px;
          dist += (Rc[k] - Pk) * (Rc[k] - Pk);
        }
        dist = sqrt(dist);
        if (tmp + dist > ext) ext = tmp + dist;
      }
    double* r = rec + 8 * (size_t)p;
    r[0] = Rc[0];
    r[1] = Rc[1];
    r[2] = Rc[2];
    r[3] = ext;
    r[4] = pow(amin + bmin, 0.25);  // (..)**0.25d0, :143
    r[5] = qab[p];
    r[6] = 0.0;
    r[7] = 0.0;
  }
}

// Screened-in prefix of the aux list of angular momentum L for every pair
// (the last kept list entry + 1), the number of kept shells and their total
// primitive count.  One thread per pair, entries in list order.
__global__ void qvl_screen_kernel(const double* __restrict__ prec, int npairs, DevShells ket,
                                  const int* __restrict__ alist, int nlist, int L, const double* __restrict__ xrec,
                                  double thresh, int* __restrict__ nsurv, int* __restrict__ nkeep,
                                  int* __restrict__ nprimkeep) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double pr[6];
  for (int k = 0; k < 6; k++) pr[k] = prec[8 * (size_t)p + k];
  int last = 0, cnt = 0, pc = 0;
  // Q_ab = 0 (a pair dropped by its Schwarz bound): every estimate is 0
  const int n = (thresh > 0.0 && !(pr[5] > 0.0)) ? 0 : nlist;
  for (int i = 0; i < n; i++) {
    const int sh = alist[i];
    if (qvl_keep(pr, xrec + 4 * (size_t)sh, ket.xyz[3 * sh], ket.xyz[3 * sh + 1], ket.xyz[3 * sh + 2], L, thresh)) {
      last = i + 1;
      cnt++;
      pc += ket.nprim[sh];
    }
  }
  nsurv[p] = last;
  nkeep[p] = cnt;
  nprimkeep[p] = pc;
}

// The kept entries of every pair's screened-in prefix, in list order:
// kidx[koff[p] ..] for pair p (koff: exclusive prefix sums of the kept
// counts of qvl_screen_kernel).  The same test as there, one thread per pair.
__global__ void qvl_compact_kernel(const double* __restrict__ prec, int npairs, DevShells ket,
                                   const int* __restrict__ alist, int L, const double* __restrict__ xrec,
                                   double thresh, const int* __restrict__ nsurv, const int64_t* __restrict__ koff,
                                   int* __restrict__ kidx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  double pr[6];
  for (int k = 0; k < 6; k++) pr[k] = prec[8 * (size_t)p + k];
  int* out = kidx + koff[p];
  const int n = nsurv[p];
  int j = 0;
  for (int i = 0; i < n; i++) {
    const int sh = alist[i];
    if (qvl_keep(pr, xrec + 4 * (size_t)sh, ket.xyz[3 * sh], ket.xyz[3 * sh + 1], ket.xyz[3 * sh + 2], L, thresh))
      out[j++] = i;
  }
}

// Schwarz screened prefixes (common/screening.f90 test Q_ab Q_C >= thresh,
// the comparison of the host sweep): for pair p and aux list L (Q_C sorted
// non-increasing in qcs[loff[L] .. loff[L] + nl[L])), the number of list
// entries with !(Q_ab Q_C < thresh).  The products are non-increasing along
// the list, so the entries that pass are a prefix: binary search for its
// end.  all: no screening (the whole list).  One thread per (pair, L).
__global__ void schwarz_prefix_kernel(const double* __restrict__ qab, int npairs, const double* __restrict__ qcs,
                                      const int* __restrict__ loff, const int* __restrict__ nl, double thresh,
                                      int all, int* __restrict__ nsflat) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, L = blockIdx.y;
  if (p >= npairs) return;
  const int n = nl[L];
  int lo = 0, hi = n;
  if (!all) {
    const double qa = qab[p];
    const double* q = qcs + loff[L];
    // invariant: entries < lo pass, entries >= hi fail
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (qa * q[mid] < thresh) hi = mid;
      else lo = mid + 1;
    }
  } else {
    lo = n;
  }
  nsflat[(size_t)L * npairs + p] = lo;
}

void schwarz_prefixes(const double* qab, int npairs, const double* qcs, const int* loff, const int* nl, int nL,
                      double thresh, bool all, int* nsflat, hipStream_t st) {
  if (npairs <= 0 || nL <= 0) return;
  hipLaunchKernelGGL(schwarz_prefix_kernel, dim3((npairs + 255) / 256, nL), dim3(256), 0, st, qab, npairs, qcs, loff,
                     nl, thresh, all ? 1 : 0, nsflat);
}

void qvl_pair_records(const DevShells& bra, const int* s1, const int* s2, const double* qab, int npairs, double erfi,
                      double* rec, hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(qvl_pair_kernel, dim3((npairs + 127) / 128), dim3(128), 0, st, bra, s1, s2, qab, npairs, erfi,
                     rec);
}

void qvl_screen(const double* prec, int npairs, const DevShells& ket, const int* alist, int nlist, int L,
                const double* xrec, double thresh, int* nsurv, int* nkeep, int* nprimkeep, hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(qvl_screen_kernel, dim3((npairs + 127) / 128), dim3(128), 0, st, prec, npairs, ket, alist, nlist,
                     L, xrec, thresh, nsurv, nkeep, nprimkeep);
}

void qvl_compact(const double* prec, int npairs, const DevShells& ket, const int* alist, int L, const double* xrec,
                 double thresh, const int* nsurv, const int64_t* koff, int* kidx, hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(qvl_compact_kernel, dim3((npairs + 127) / 128), dim3(128), 0, st, prec, npairs, ket, alist, L,
                     xrec, thresh, nsurv, koff, kidx);
}

// Kept-list window of every pair for one launch of the matrix-core kernels:
// with the list window [cbeg, cend), cbeg = max(lo, nbeg[p]) (lo without
// nbeg) and cend = min(hi, nend[p]), ja[p] / jb[p] = the first kept entries
// kidx[koff[p] ..] at or after cbeg / cend (both 0 when the window is empty),
// so that each workgroup of the pair reads two ints instead of searching its
// kept list.  One thread per pair.
__global__ void qvl_window_kernel(const int64_t* __restrict__ koff, const int* __restrict__ kidx, int npairs, int lo,
                                  int hi, const int* __restrict__ nbeg, const int* __restrict__ nend,
                                  int* __restrict__ ja, int* __restrict__ jb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int cbeg = nbeg ? max(lo, nbeg[p]) : lo, cend = min(hi, nend[p]);
  if (cbeg >= cend) {
    ja[p] = 0;
    jb[p] = 0;
    return;
  }
  const int* kl = kidx + koff[p];
  const int nk = (int)(koff[p + 1] - koff[p]);
  auto lower = [&](int v) {
    int a = 0, b = nk;
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (kl[mid] < v) a = mid + 1;
      else b = mid;
    }
    return a;
  };
  ja[p] = lower(cbeg);
  jb[p] = lower(cend);
}

void qvl_windows(const int64_t* koff, const int* kidx, int npairs, int lo, int hi, const int* nbeg, const int* nend,
                 int* ja, int* jb, hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(qvl_window_kernel, dim3((npairs + 255) / 256), dim3(256), 0, st, koff, kidx, npairs, lo, hi, nbeg,
                     nend, ja, jb);
}

// *count += the kept entries kidx[koff[p] ..] at or after nfrom[p], over the
// pairs (the kept triplets of the windows [nfrom, nsurv)).  One thread per pair.
__global__ void qvl_count_from_kernel(const int64_t* __restrict__ koff, const int* __restrict__ kidx,
                                      const int* __restrict__ nfrom, int npairs,
                                      unsigned long long* __restrict__ count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int* kl = kidx + koff[p];
  const int nk = (int)(koff[p + 1] - koff[p]), v = nfrom[p];
  int lo = 0, hi = nk;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (kl[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  if (nk > lo) atomicAdd(count, (unsigned long long)(nk - lo));
}

void qvl_count_from(const int64_t* koff, const int* kidx, const int* nfrom, int npairs, unsigned long long* count,
                    hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(qvl_count_from_kernel, dim3((npairs + 255) / 256), dim3(256), 0, st, koff, kidx, nfrom, npairs,
                     count);
}

}  // namespace dfhip
