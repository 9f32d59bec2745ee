// Host launchers of the dense kernels (linalg.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "dfhip_common.h"

namespace dfhip {

// out = in^T for a row-major nr x nc matrix
void transpose(const double* in, int64_t nr, int64_t nc, double* out, hipStream_t st);
// Dp[r * ldd + I] for the npf rows r (pair functions; rowpair[r] = pair of r)
void pack_density(const int* rowpair, const int* s1, const int* s2, const int64_t* row, int64_t npf, const int* L,
                  const int* f0, const double* D, int64_t sI, int64_t si, int64_t sj, int ndens, double* Dp, int ldd,
                  hipStream_t st);
// The packing for the C layout (density index slowest) as a tiled transpose;
// pairof[s*nsh + t] = pair index of {s, t} | (1 << 30 if oriented (s, t)),
// fshell[j] = shell of function j.
void pack_density_t(const int* pairof, const int* fshell, const int64_t* row, const int* L, const int* f0, int nsh,
                    int nbf, const double* D, int64_t sI, int ndens, double* Dp, int ldd, hipStream_t st);
// rowpair[r] = p for the rows of pair p (row[p] .. row[p+1]-1, the last ends
// at npf); with dprow: rowperm[r] = dprow[p] + r - row[p]
void fill_rowpair(const int64_t* row, int npairs, int64_t npf, int* rowpair, const int64_t* dprow, int* rowperm,
                  hipStream_t st);
// Screened-prefix mask of the staging matrix T: entry (row k, column P) is
// live iff colia[P] < nsurv[colL[P] * npairs + rowpair[k]] (column arrays
// relative to the column block).
struct ScreenMask {
  const int* rowpair;
  const int* rowperm;  // packed density row of each staging-matrix row
  const int* colL;
  const int* colia;
  const int* nsurv;
  int npairs;
};

// ipt: densities per thread of the output tile (0 or an unsupported value:
// contract_ipt(ndens))
void contract_T(const double* T, int64_t ldt, const double* Dp, int ldd, int ncol, int ndens, int nseg,
                const int64_t* seg, const ScreenMask& mk, double* part, int nsplit, const int* orig, double* d,
                int64_t sP, int64_t sI, hipStream_t st, int ipt = 0);
int contract_tile();
int contract_ipt(int ndens);
bool contract_ipt_ok(int ipt);
// output tiles of one contraction (column tiles x density tiles)
int64_t contract_ntiles(int64_t ncol, int ndens, int ipt = 0);
// Cholesky of [[M], [I]] (n columns, 2n rows, lda): L in the leading block,
// L^-T in the border rows n..2n-1.
void cholesky_inverse_rows(double* A, int64_t lda, int64_t n, int* info, hipStream_t st);
void set_identity_rows(double* A, int64_t lda, int64_t n, hipStream_t st);
// BT (ndens x n, column-major) = DT (ndens x n, ldd) times the upper triangular LT (ldl);
// part holds apply_linv_part_size(n, ndens) doubles
int64_t apply_linv_part_size(int64_t n, int ndens);
void apply_linv(const double* DT, int64_t ldd, const double* LT, int64_t ldl, int64_t n, int ndens, double* part,
                double* BT, hipStream_t st);
// G (ndens x ndens, column-major) = -BT BT^T; part holds gram_chunks(n, ndens) ndens^2 doubles
int gram_chunks(int64_t n, int ndens);
void gram_neg(const double* BT, int64_t n, int ndens, double* part, double* G, hipStream_t st);
void bt_to_b(const double* BT, int64_t n, int ndens, double* B, hipStream_t st);
void assemble_j(const double* A, int64_t lda, int64_t n, int ndens, int nstates, double* J, hipStream_t st);
void schwarz_aux_from_metric(const double* M, int64_t n, const int* L, const int* f0, int nsh, double* q,
                             hipStream_t st, bool compact = false);
void boys_probe(const int* n, const double* T, int64_t npts, const double* tab, double* F, hipStream_t st);

// QVL screening (screen.hip): pair records (8 doubles per pair) and, per aux
// list of angular momentum L, the screened-in prefix, kept shells and kept
// primitives of every pair.
void qvl_pair_records(const DevShells& bra, const int* s1, const int* s2, const double* qab, int npairs, double erfi,
                      double* rec, hipStream_t st);
void qvl_screen(const double* prec, int npairs, const DevShells& ket, const int* alist, int nlist, int L,
                const double* xrec, double thresh, int* nsurv, int* nkeep, int* nprimkeep, hipStream_t st);
// kidx[koff[p] ..] = the kept list entries (ascending) of pair p's screened-in
// prefix nsurv[p] of the aux list of angular momentum L
void qvl_compact(const double* prec, int npairs, const DevShells& ket, const int* alist, int L, const double* xrec,
                 double thresh, const int* nsurv, const int64_t* koff, int* kidx, hipStream_t st);
// *count += the kept entries of every pair at or after list entry nfrom[p]
void qvl_windows(const int64_t* koff, const int* kidx, int npairs, int lo, int hi, const int* nbeg, const int* nend,
                 int* ja, int* jb, hipStream_t st);
void qvl_count_from(const int64_t* koff, const int* kidx, const int* nfrom, int npairs, unsigned long long* count,
                    hipStream_t st);

// Schwarz screened prefixes (screen.hip) of every pair and aux list L into
// nsflat[L * npairs + p]: qab per pair, qcs the Q_C of the lists (sorted
// non-increasing, list L at loff[L], nl[L] entries); all: no screening.
void schwarz_prefixes(const double* qab, int npairs, const double* qcs, const int* loff, const int* nl, int nL,
                      double thresh, bool all, int* nsflat, hipStream_t st);

// The metric from its computed side (linalg.hip): fL[u] = L of the shell of
// aux function u.
void symmetrize_metric(double* M, int64_t ldm, int64_t n, const int* fL, hipStream_t st);

// Shell-pair classes (la >= lb) of a basis, for generating the pair list on
// the device: the shells of each L in one list (ascending shell index per
// L); per class its first pair, its shells' offsets in that list, the number
// of lb shells, the functions per pair and its first row.  Inside a class
// the pairs run over the la shells, then the lb shells (for la == lb only
// the lb shell <= the la shell).
constexpr int PC_MAX = (LMAX_ORB + 1) * (LMAX_ORB + 2) / 2;
struct PairClassTable {
  int ncls;
  int begin[PC_MAX + 1];
  int la[PC_MAX], lb[PC_MAX];
  int off_a[PC_MAX], off_b[PC_MAX], n_b[PC_MAX], w[PC_MAX];
  int64_t row0[PC_MAX];
};
// s1[p], s2[p] and (if prow0) the first row prow0[p] of every pair
void gen_pairs(const PairClassTable& t, const int* shflat, int* s1, int* s2, int64_t* prow0, hipStream_t st);

constexpr int CHOL_BLOCK = 64;

}  // namespace dfhip
