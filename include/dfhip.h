/*
 * dfhip.h -- C ABI of libdfhip.so, the MI355X (gfx950) density-fitting
 * Coulomb library.  Drop-in for the native symbols pysisyphus-addons'
 * Cython layer binds, plus extended entry points.
 *
 * Basis arrays follow the reference ABI (common/shells.f90:97-110):
 *   bas_centers  int32 column-major (nshells, 3): centre index, atomic
 *                number, 1-based pointer to x, y, z (bohr) in bas_data
 *   bas_spec     int32 column-major (nshells, 5): shell index, L, number of
 *                primitives, number of contractions, 1-based pointer to the
 *                contraction coefficients (exponents follow) in bas_data
 *   bas_data     float64
 * Spherical functions are ordered shell by shell, m = -l..l inside a shell.
 *
 * The reference's entry points return nothing and abort on fatal errors;
 * here every call records a status readable with dfhip_last_status() and
 * dfhip_last_error() (0 = success), and never aborts the process.
 */
#ifndef DFHIP_H
#define DFHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference-compatible entry points -------------------------------- */

/* One-time initialization of tables and device state.
 * Replaces f_init, common/integrals/init.f90:13-27. */
void f_init(void);

/* B = L^-1 d with d(P, I) = sum_{mu nu} (mu nu|P) D_I(mu, nu), Schwarz-screened
 * at the single-precision literal 1e-10, L the Cholesky factor of (P|Q).
 * densities: column-major (ndens, nbfs, nbfs), packed lower-triangular state
 * pairs (1,1),(2,1),(2,2),...; df_tensor: column-major (naux, ndens), written.
 * naux must equal the number of spherical auxiliary functions.
 * Replaces f_contract_coulomb_densities_2d, diabatization/intor_wrapper.f90:12-43
 * (bound in diabatization/_intor.pyx:11-31). */
void f_contract_coulomb_densities_2d(int32_t nshells, int32_t ndata, const int32_t *bas_centers,
                                     const int32_t *bas_spec, const double *bas_data,
                                     int32_t nshells_aux, int32_t ndata_aux,
                                     const int32_t *bas_centers_aux, const int32_t *bas_spec_aux,
                                     const double *bas_data_aux, int32_t ndens, int32_t nbfs,
                                     const double *densities, int32_t naux, double *df_tensor);

/* Coulomb tensor J_IJKL = sum_P B(P, IJ) B(P, KL), column-major (nstates)^4,
 * all 8 permutations written.
 * Replaces f_contract_coulomb_densities_4d, diabatization/intor_wrapper.f90:45-76
 * (bound in diabatization/_intor.pyx:32-52). */
void f_contract_coulomb_densities_4d(int32_t nshells, int32_t ndata, const int32_t *bas_centers,
                                     const int32_t *bas_spec, const double *bas_data,
                                     int32_t nshells_aux, int32_t ndata_aux,
                                     const int32_t *bas_centers_aux, const int32_t *bas_spec_aux,
                                     const double *bas_data_aux, int32_t ndens, int32_t nbfs,
                                     const double *densities, int32_t nstates,
                                     double *coulomb_tensor);

/* integrals[c] = sqrt(||(C|C)||_F) per shell.
 * Replaces f_int_schwarz_aux, intor/intor_wrapper.f90:12-24. */
void f_int_schwarz_aux(int32_t nshells, int32_t ndata, const int32_t *bas_centers,
                       const int32_t *bas_spec, const double *bas_data, double *integrals);

/* integrals[a(a+1)/2 + b] = sqrt(||(ab|ab)||_F) for a >= b (0-based).
 * Replaces f_int_schwarz_bra, intor/intor_wrapper.f90:26-38. */
void f_int_schwarz_bra(int32_t nshells, int32_t ndata, const int32_t *bas_centers,
                       const int32_t *bas_spec, const double *bas_data, double *integrals);

/* df1c_tensor (column-major nbfs x nbfs_aux) = D1 L^-T with D1(mu, Q) the
 * diagonal bra elements (mu mu|Q) of the Schwarz-screened (aa|C) blocks and L
 * the Cholesky factor of (P|Q).
 * Replaces f_intor_df1c, std/intor_wrapper.f90:40-62 (std/intor.f90:15-112). */
void f_intor_df1c(int32_t nshells, int32_t ndata, const int32_t *bas_centers, const int32_t *bas_spec,
                  const double *bas_data, int32_t nshells_aux, int32_t ndata_aux,
                  const int32_t *bas_centers_aux, const int32_t *bas_spec_aux, const double *bas_data_aux,
                  int32_t nbfs, int32_t nbfs_aux, double *df1c_tensor);

/* eri2c_tensor (column-major nbfs x nbfs) = df1c df1c^T, the density-fitted
 * (aa|bb).  Replaces f_intor_eri2c, std/intor_wrapper.f90:12-38
 * (std/intor.f90:114-131). */
void f_intor_eri2c(int32_t nshells, int32_t ndata, const int32_t *bas_centers, const int32_t *bas_spec,
                   const double *bas_data, int32_t nshells_aux, int32_t ndata_aux,
                   const int32_t *bas_centers_aux, const int32_t *bas_spec_aux, const double *bas_data_aux,
                   int32_t nbfs, double *eri2c_tensor);

/* df2c_mo_tensor (column-major nmos(nmos+1)/2 x nbfs_aux): row i(i+1)/2 + a
 * (i >= a, 0-based) = sum_Q [sum_{mu nu} C(mu, i) C(nu, a) (mu nu|Q)] L^-T(Q, P),
 * mo_tensor column-major (nbfs, nmos).  The reference's extra drops of
 * single values <= 1e-10 (std/intor.f90:244-266) are not applied.
 * Replaces f_intor_df2c_mo, std/intor_wrapper.f90:64-88 (std/intor.f90:133-292). */
void f_intor_df2c_mo(int32_t nshells, int32_t ndata, const int32_t *bas_centers, const int32_t *bas_spec,
                     const double *bas_data, int32_t nshells_aux, int32_t ndata_aux,
                     const int32_t *bas_centers_aux, const int32_t *bas_spec_aux,
                     const double *bas_data_aux, int32_t nbfs, int32_t nmos, const double *mo_tensor,
                     int32_t nbfs_aux, double *df2c_mo_tensor);

/* Screening-estimate records of all triplets of the pairs a >= b that are
 * neither concentric nor small (Q_ab <= 1e-10), in the reference's loop order,
 * written to int_estimates.bin in the working directory as 4 float64 per
 * triplet: Schwarz estimate, QVL estimate, ||(ab|C)||_F, |contracted centre - C|;
 * the reference's log lines go to stdout.  The metric and its inverse square
 * root are computed (and timed) as in the reference.
 * Replaces f_benchmark_int3c2e, intor/benchmark_wrapper.f90:12-38
 * (intor/benchmark.f90:15-132). */
void f_benchmark_int3c2e(int32_t nshells, int32_t ndata, const int32_t *bas_centers,
                         const int32_t *bas_spec, const double *bas_data, int32_t nshells_aux,
                         int32_t ndata_aux, const int32_t *bas_centers_aux, const int32_t *bas_spec_aux,
                         const double *bas_data_aux);

/* grid_densities (column-major ndens x npoints, overwritten) = the densities
 * of the Cartesian AO density matrices (column-major ndens x nbfs x nbfs,
 * element (j, mu, nu) read for mu >= nu; nbfs = Cartesian functions of the
 * basis, pysisyphus order) at the points grid3d (column-major 3 x npoints).
 * Points go in blocks of blk_size; a pair (mu, nu) enters a block's points
 * when |f D_j(mu, nu) mean_nu mean_mu| >= thresh (f = 2 off the diagonal,
 * means of |chi| over the block).  Shells up to L = 4 (the reference returns
 * NaN above; this library reports DFHIP_ERR_UNSUPPORTED).
 * Replaces f_eval_densities, grid/grid_wrapper.f90:10-39
 * (grid/grid.f90:112-286). */
void f_eval_densities(int32_t nshells, int32_t ndata, const int32_t *bas_centers, const int32_t *bas_spec,
                      const double *bas_data, int32_t npoints, const double *grid3d, int32_t ndens,
                      int32_t nbfs, const double *densities, double *grid_densities, int32_t blk_size,
                      double thresh);

/* ---- status ------------------------------------------------------------ */

enum {
    DFHIP_OK = 0,
    DFHIP_ERR_ARG = 1,       /* invalid argument or basis data */
    DFHIP_ERR_HIP = 2,       /* HIP runtime error / no device */
    DFHIP_ERR_NOT_SPD = 3,   /* metric not positive definite */
    DFHIP_ERR_UNSUPPORTED = 4,
    DFHIP_ERR_MEMORY = 5
};

/* Status of the last call on this thread and its message. */
int dfhip_last_status(void);
const char *dfhip_last_error(void);

/* ---- extended entry points (no reference counterpart) ------------------ */

typedef struct {
    int32_t nshells, ndata;
    const int32_t *bas_centers; /* column-major (nshells, 3) */
    const int32_t *bas_spec;    /* column-major (nshells, 5) */
    const double *bas_data;
} dfhip_basis;

enum { DFHIP_FP64 = 0, DFHIP_MIXED = 1 };
enum { DFHIP_SCREEN_SCHWARZ = 0, DFHIP_SCREEN_NONE = 1, DFHIP_SCREEN_QVL = 2 };
enum { DFHIP_DENS_FORTRAN = 0, DFHIP_DENS_C = 1 };

typedef struct {
    int32_t precision;   /* DFHIP_FP64, or DFHIP_MIXED: 3c2e triplets whose Schwarz
                            product Q_ab Q_C is below fp32_below are evaluated
                            with FP32 arithmetic (FP64 Boys values and
                            prefactors); everything else is FP64 */
    int32_t screening;   /* DFHIP_SCREEN_SCHWARZ (reference), _NONE, or _QVL: the
                            screener's qvl_estimate (common/screening.f90:267-299)
                            with thresh_ws = 1e-4 and aux extents at 1e-8 */
    double thresh;       /* skip when the estimate is < thresh; default (double)1e-10f */
    int32_t dens_layout; /* DFHIP_DENS_FORTRAN: (ndens,nbf,nbf) column-major,
                            DFHIP_DENS_C: (ndens,nbf,nbf) row-major */
    int32_t verbose;     /* 1: print sizes, phase times and screening counts;
                            2: also the wall time of each set-up step before the
                            integrals (synchronizing, for diagnosis); 3: instead
                            of 2, when the host and when the stream reached each
                            step, without waits (the env DFHIP_VERBOSE=n sets
                            the default) */
    int64_t max_tile_bytes; /* cap of the (ab|P) staging matrix, 0 = automatic */
    int32_t dens_on_device; /* densities is a device pointer (HBM-resident input) */
    int32_t reserved;
    /* (a class whose FP32 Et table exceeds a sixth of the free HBM, or the
       bytes in the environment variable DFHIP_ET32_CAP, runs FP64) */
    double fp32_below;   /* DFHIP_MIXED: FP32 below this Schwarz product (default
                            1e-2: J within 1e-9 of the FP64 path on the test
                            systems, 8 to 200 atoms); the FP32 windows of the
                            generated kernels (tuning table or DFHIP_GEN32=1)
                            stay below 1e-3, where they move J by at most
                            6e-11 (4.4e-8 on one batch molecule at 1e-2) */
} dfhip_options;

typedef struct {
    int64_t ntriplets;   /* nshells(nshells+1)/2 * nshells_aux */
    int64_t ncomputed;   /* screened-in triplets */
    int64_t nskipped;
    int64_t nbf, naux, npair_functions; /* rows of the staging matrix: the functions of
                                           the shell pairs with Q_ab > 0 */
    int32_t ncolumn_blocks;
    /* phase times in ms (HIP events): ms_schwarz, ms_pack (the set-up from
       the end of the Schwarz kernels to the first 3c2e kernel),
       ms_int3c2e, ms_contract and ms_assemble (B^T = d^T L^-T, -B^T B and
       the J assembly) on the library's stream; the density packing, the
       metric (ms_metric) and its Cholesky run on a second stream beside the
       Schwarz kernels and the integrals, ms_cholesky is the part of the
       Cholesky still running after the contraction */
    double ms_upload, ms_schwarz, ms_metric, ms_cholesky, ms_int3c2e, ms_pack, ms_contract,
        ms_solve, ms_assemble, ms_download, ms_total;
    double ms_h2d_densities; /* copying host densities (second stream) */
    double ms_host_plan;     /* host time: pair sorting, screening prefixes, tables */
    /* screened-in triplets and primitive triplets per class, index
       (la*(la+1)/2 + lb) * 6 + lc with la >= lb the pair's angular momenta */
    int64_t class_triplets[15 * 6];
    int64_t class_prim_triplets[15 * 6];
    int64_t ncomputed_fp32; /* DFHIP_MIXED: triplets inside the FP32 windows */
    int64_t npairs_zero_q;  /* shell pairs with Q_ab = 0: dropped by their bound before the
                               (ab|ab) block (see dfhip_schwarz_bounds), or underflowing */
} dfhip_stats;

void dfhip_default_options(dfhip_options *opt);

/* Full path with options and statistics.  df_tensor (column-major naux x
 * ndens) and coulomb_tensor (column-major nstates^4) may each be NULL.
 * Returns the status. */
int dfhip_contract_coulomb_densities(const dfhip_basis *bas, const dfhip_basis *aux, int32_t ndens,
                                     int32_t nbfs, const double *densities,
                                     const dfhip_options *opt, double *df_tensor,
                                     int32_t nstates, double *coulomb_tensor, dfhip_stats *stats);

/* The batched form: nmol independent molecules.  Molecule k has the bases
 * bas[k] / aux[k], ndens[k] densities of nbfs[k] functions at densities[k]
 * (layout opt->dens_layout), and receives its Coulomb tensor
 * (nstates[k]^4, column-major) in coulomb_tensors[k] and, if stats is not
 * NULL, its statistics in stats[k].  The molecules run concurrently on
 * nlanes streams (0 = 4, at most 8), each with its own workspace, one host
 * thread per stream; the results equal those of one call per molecule.  The
 * first failing molecule stops the batch; the message names it. */
int dfhip_contract_coulomb_densities_batch(int32_t nmol, const dfhip_basis *bas, const dfhip_basis *aux,
                                           const int32_t *ndens, const int32_t *nbfs,
                                           const double *const *densities, const dfhip_options *opt,
                                           const int32_t *nstates, double *const *coulomb_tensors,
                                           dfhip_stats *stats, int32_t nlanes);

/* Dense screened 3c2e tensor out[(i*nbf + j)*naux + P] (row-major, both
 * (i,j) orders filled, screened-out triplets zero), for parity checks on
 * small systems. */
int dfhip_int3c2e_dense(const dfhip_basis *bas, const dfhip_basis *aux, const dfhip_options *opt,
                        double *out, dfhip_stats *stats);

/* Frees the device memory the library keeps between calls (staging matrix,
 * bordered metric, packed densities and the cached scratch blocks of earlier
 * calls); the next call re-allocates. */
int dfhip_release_workspace(void);

/* The QVL screener's records as the device path computes them
 * (common/screening.f90:101-163 set-up): pair_records[8 k + 0..7] for the
 * packed pair k = a(a+1)/2 + b, a >= b (0-based) = contracted centre x, y, z,
 * contracted extent (thresh_ws 1e-4), (min a + min b)^(1/4), Q_ab, 0, 0;
 * aux_records[4 c + 0..3] = auxiliary extent (prec 1e-8),
 * beta = min exp^(-(2L+3)/4) sqrt((2L-1)!!), Q_C, 0. */
int dfhip_qvl_records(const dfhip_basis *bas, const dfhip_basis *aux, double *pair_records,
                      double *aux_records);

/* Allocation-free form of f_int_schwarz_aux: integrals[c] = sqrt(||(C|C)||_F)
 * per aux shell, host arrays in and out, synchronous.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  It calls no hipMalloc, hipFree or
 * pinned allocation: every device buffer lives in a static 16 MiB device array
 * per GPU (88 bytes per aux function plus about 100 bytes per shell must fit,
 * e.g. 170,000 functions in h shells; a larger basis is rejected with
 * DFHIP_ERR_ARG).  f_init() must have run on the current device first.
 * Calls in one process are serialized by the library; the array keeps the last
 * call's contents. */
int dfhip_int_schwarz_aux_noalloc(const dfhip_basis *aux, double *integrals);

/* (P|Q)^power, naux x naux, by the eigen route of matrix_powerh
 * (common/linalg.f90:36-58): rocSOLVER dsyevd, then V diag(lambda^power) V^T. */
int dfhip_metric_powerh(const dfhip_basis *aux, double power, double *out);

/* The records of f_benchmark_int3c2e in memory: records (4 float64 each) may be
 * NULL to count them; *nrecords receives the count. */
int dfhip_int3c2e_estimates(const dfhip_basis *bas, const dfhip_basis *aux, double *records,
                            int64_t capacity, int64_t *nrecords);

/* Rigorous upper bounds of the Schwarz norms, on the host: pair_bounds[k]
 * >= Q_ab for the packed pair k = a(a+1)/2 + b, a >= b (0-based), and
 * aux_bounds[c] >= Q_C.  No counterpart in the reference: the screened
 * contraction drops, before their (ab|ab) blocks are evaluated, the pairs
 * whose bound times the largest factor of their triplets' estimates is
 * below half the threshold (DFHIP_NO_PAIR_SKIP=1 evaluates every pair). */
int dfhip_schwarz_bounds(const dfhip_basis *bas, const dfhip_basis *aux, double *pair_bounds,
                         double *aux_bounds);

/* Metric (P|Q), naux x naux (symmetric). */
int dfhip_metric(const dfhip_basis *aux, double *out);

/* F[i] = F_{n[i]}(T[i]), 0 <= n[i] <= 17, T[i] >= 0, on the device. */
int dfhip_boys(int64_t npts, const int32_t *n, const double *T, double *F);

/* ---- autotuner (no counterpart in the reference) -------------------------
 * dfhip_autotune measures, on the caller's workload (arguments as for
 * dfhip_contract_coulomb_densities; opt selects the precision and screening
 * that are tuned), the kernel choices that the library otherwise takes from
 * fixed rules:
 *   - per (la, lb, lc) class that has both a generated VALU kernel and a
 *     matrix-core kernel, and per precision window (0 = FP64, 1 = the FP32
 *     window of the mixed path), which of the two runs it;
 *   - for this density count, the contraction's output tile (code 1-5 or 7:
 *     the FP64 VALU kernel with 16 x code densities per tile; 102-107: the
 *     FP64 matrix-core kernel with 32 x (code - 100)) and the scale of its
 *     split-K depth (1/2, 1, 2).
 * One full contraction per candidate, timed with HIP events per class; the
 * rule's choice stays unless another is at least 2 % faster.  The table
 * replaces this process's table and is written to path (NULL: not written).
 * dfhip_load_tuning reads such a file (NULL clears the table); the library
 * also reads the file named by the environment variable DFHIP_TUNING when it
 * sets up its device.  Results do not depend on the table beyond rounding. */
int dfhip_autotune(const dfhip_basis *bas, const dfhip_basis *aux, int32_t ndens, int32_t nbfs,
                   const double *densities, const dfhip_options *opt, const char *path);
int dfhip_load_tuning(const char *path);
/* 1 generated, 2 matrix cores, 0 not tuned, -1 invalid arguments */
int32_t dfhip_tuning_choice(int32_t la, int32_t lb, int32_t lc, int32_t window);
/* 1 and *ipt (tile code), *split_scale when the table has an entry for ndens, else 0 */
int32_t dfhip_tuning_contract(int32_t ndens, int32_t *ipt, double *split_scale);

#ifdef __cplusplus
}
#endif

#endif /* DFHIP_H */
