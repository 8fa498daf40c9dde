/*
 * corr_abi.h — C-ABI of the MI355X (gfx950) post-HF contraction library
 * (libcorr_gpu.so).
 *
 * Every entry point takes plain pointers and sizes, returns an int status and
 * never throws.  Array arguments named d_* are device (HBM) addresses; h_*
 * are host addresses.  Work is enqueued on the context's stream; functions
 * that return host scalars (energies) synchronise that stream before they
 * return.  The library never allocates host memory for the caller; device
 * scratch lives in the context and is reused between calls.
 *
 * Layouts follow the reference (row-major, 64-bit indices here):
 *   MOs        C[p*N + mu]                        (row = MO)
 *   Fock, FMo  F[p*N + q]
 *   prec_ints  (ij|kl) at ((i*N + j)*N + k)*N + l (chemist notation)
 *   BPQ        B[Q*N*N + mu*N + nu]
 *   Bia        B[Q*o*v + i*v + a]
 *   T1         [i*v + a],  T2 [((i*o + j)*v + a)*v + b]   (spin orbitals)
 *
 * Reference interfaces replaced (Masterluke87/corr):
 *   build_FMo             src/POSTHF/motrans.cc:8-32
 *   MOtrans / trans_I /
 *   mo2int_op             src/POSTHF/motrans.cc:35-211, 341-503
 *   run_canonical_mp2     src/POSTHF/can_mp2.cc:11-49
 *   calculate_ri (arith.) src/POSTHF/ri.cc:50-69, 102-103
 *   transform_ri          src/POSTHF/ri.cc:147-172
 *   run_canonical_mp2_ri  src/POSTHF/can_mp2.cc:53-132
 *   ccsd_ur and helpers   src/POSTHF/ccsd.cc:6-580,
 *                         src/POSTHF/ccsd_intermediates.cc:6-287
 *   run_non_canonical_mp2 src/POSTHF/loc_mp2.cc:7-92
 *   build_Fmat            src/SCF/ops_rhf.cc:466-585
 *   build_Pmat            src/SCF/scf.cc:44-61
 *   run_scf (+ calc_S12,  src/SCF/scf.cc:105-320,
 *   diag_Fmat, check_initial_guess)  src/SCF/ops_rhf.cc:15-37, 342-373, 425-430
 *   localize_PM /         syshfwplug/pymodule.py:75-167
 *   localize_boys
 */
#ifndef CORR_ABI_H
#define CORR_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CORR_OK            0
#define CORR_EINVAL        1   /* bad dimension, null pointer, bad enum      */
#define CORR_ENOMEM        2   /* device allocation failed                   */
#define CORR_EHIP          3   /* a HIP runtime call failed                  */
#define CORR_ENOTCONV      4   /* iteration did not converge                 */
#define CORR_ENOTPD        5   /* metric not positive definite (Cholesky)    */
#define CORR_ENODEV        6   /* no usable gfx950 device                    */

/* ---- arithmetic of the bulk contractions ------------------------------- */
#define CORR_PREC_FP64     0   /* FP64 MFMA, FP64 accumulate (plain method)  */
#define CORR_PREC_FP32     1   /* FP32 MFMA, FP32 accumulate                 */
#define CORR_PREC_BF16X3   2   /* hi/lo BF16 split, 3 products, FP32 acc.    */
#define CORR_PREC_INT8X5   3   /* corr_dgemm only: FP64 emulated on int8      */
                               /* MFMA, 5 slices per operand (Ozaki), exact   */
                               /* int32 sums; error below 2^-35 of the row    */
                               /* and column maxima                           */

/* ---- RI metric factorisation ------------------------------------------- */
#define CORR_METRIC_EIGEN     0 /* J^-1/2 = V L^-1/2 V^T, cutoff as ri.cc:56-66 */
#define CORR_METRIC_CHOLESKY  1 /* L^-1 with J = L L^T (same B B^T products)   */

/* ---- CCSD iteration mode ------------------------------------------------ */
#define CORR_CCSD_REFERENCE   0 /* fixed iteration count, no DIIS (ccsd.cc:564) */
#define CORR_CCSD_DIIS        1 /* iterate to |dE| and residual thresholds     */

typedef struct corr_ctx corr_ctx;

typedef struct corr_mp2_energy {
    double e_os;       /* EMP2_OS  (opposite spin)                  */
    double e_ss;       /* EMP2_SS  (same spin)                      */
    double e_mp2;      /* EMP2 = EMP2_OS + EMP2_SS                  */
    double e_scs;      /* EMP2(SCS) = EMP2_SS/3 + 6/5 EMP2_OS       */
} corr_mp2_energy;

/* ---- version / context --------------------------------------------------- */
int         corr_version(void);
int         corr_ctx_create(int device, corr_ctx** out);
int         corr_ctx_destroy(corr_ctx* ctx);
/* Enqueue on an existing hipStream_t (passed as void*); NULL is the HIP
 * default (null) stream.  A new context uses a stream of its own.  The
 * context's device scratch is shared by every call on it, and several entry
 * points return before their kernels finish, so a change of stream orders
 * the new stream after all work already enqueued on the old one (an event
 * recorded on the old stream, waited for by the new one): calls on one
 * context stay one ordered queue whichever streams they name. */
int         corr_ctx_set_stream(corr_ctx* ctx, void* hip_stream);
int         corr_ctx_synchronize(corr_ctx* ctx);
const char* corr_ctx_last_error(const corr_ctx* ctx);
/* Releases the context's cached device scratch. */
int         corr_ctx_trim(corr_ctx* ctx);
/* Caps the device scratch the context may cache, in bytes (0: no cap).  A
 * call whose workspace would exceed it fails with CORR_ENOMEM and leaves the
 * context usable; corr_ctx_trim releases what is held.  No reference
 * counterpart (the reference never checks allocations). */
int         corr_ctx_set_scratch_limit(corr_ctx* ctx, int64_t bytes);
/* Kernel timing: when enabled, the main kernel of each entry point is
 * bracketed by HIP events on the context stream.  corr_ctx_kernel_time
 * returns the summed duration (ms) and launch count of one kernel name
 * ("rimp2_energy", "rimp2_gradient", "gemm", ...) since the last reset;
 * name == NULL resets all counters.  No reference counterpart (the
 * reference only has omp_get_wtime timers, main.cc:42-45, 159-166). */
int         corr_ctx_enable_timing(corr_ctx* ctx, int on);
int         corr_ctx_kernel_time(corr_ctx* ctx, const char* name,
                                 double* h_total_ms, int64_t* h_launches);

/* ---- device memory helpers (for C/C++ hosts without a tensor library) ---- */
int corr_dmalloc(corr_ctx* ctx, void** d_ptr, int64_t bytes);
int corr_dfree(corr_ctx* ctx, void* d_ptr);
int corr_h2d(corr_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes);
int corr_d2h(corr_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes);

/* ---- generic strided-batched MFMA GEMM (row-major) ----------------------
 * C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b],  b < batch.
 * op(A) is M x K: transa = 0 -> A is M x K (lda >= K), 1 -> A is K x M.
 * op(B) is K x N: transb = 0 -> B is K x N (ldb >= N), 1 -> B is N x K.
 * All operands FP64 in memory; prec selects the MFMA arithmetic.
 * CORR_PREC_INT8X5 takes batch = 1 only and K <= 75000;
 * a non-finite entry makes its row of C (in A) or column (in B) NaN.
 * Used for every GEMM-shaped contraction of the path. */
int corr_dgemm(corr_ctx* ctx, int transa, int transb,
               int64_t M, int64_t N, int64_t K, double alpha,
               const double* d_A, int64_t lda, int64_t strideA,
               const double* d_B, int64_t ldb, int64_t strideB,
               double beta, double* d_C, int64_t ldc, int64_t strideC,
               int64_t batch, int prec);

/* ---- a1: FMo = C F C^T  (motrans.cc:8-32) -------------------------------- */
int corr_build_fmo(corr_ctx* ctx, int64_t nroao, const double* d_MOs,
                   const double* d_Fmat, double* d_FMo);

/* ---- a2-a4: 4-index AO->MO transform (motrans.cc:35-211, 341-503) --------
 * Unfold the sorted unique AO list (5 permutation classes, bounds in
 * sortcount[0..3], nrofint entries, indices 4 x uint16 per entry) into the
 * dense N^4 AO tensor, then apply four quarter transforms.  d_work must hold
 * N^4 doubles; d_prec_ints receives (ij|kl).  d_work is overwritten. */
int corr_ao_unfold(corr_ctx* ctx, int64_t nroao, int64_t nrofint,
                   const int64_t* h_sortcount, const double* d_intval,
                   const uint16_t* d_intnums, double* d_dense_ao);
int corr_motrans(corr_ctx* ctx, int64_t nroao, const double* d_MOs,
                 const double* d_dense_ao, double* d_work,
                 double* d_prec_ints, int prec);

/* ---- a6: canonical MP2 from (ia|jb) in prec_ints (can_mp2.cc:11-49) ------ */
int corr_canonical_mp2(corr_ctx* ctx, int64_t nroao, int64_t nocc,
                       const double* d_prec_ints, const double* d_FMo,
                       corr_mp2_energy* h_out);

/* ---- a7: RI metric (ri.cc:50-69) -----------------------------------------
 * Writes the naux x naux matrix M with M^T M = J^-1 (on the kept space):
 * EIGEN -> symmetric J^-1/2 with the reference cutoff (eigenvalues with
 * l/l_max < 1e-10 or l < 0 dropped); CHOLESKY -> lower-triangular L^-1.
 * h_nkept (optional) receives the number of kept eigenvalues / naux. */
int corr_ri_metric(corr_ctx* ctx, int64_t naux, const double* d_J,
                   int mode, double* d_M, int64_t* h_nkept);

/* ---- a7/a8: B^Q_ia formation ---------------------------------------------
 * corr_ri_bpq:      BPQ = M * (P|mu nu)              (ri.cc:102-103)
 * corr_transform_ri: Bia[Q,i,a] = sum C_i,mu C_a,nu BPQ[Q,mu,nu]  (ri.cc:147-172)
 * corr_ri_form_bia: Bia = M * ((P|mu nu) -> (P|ia)), transform first, then
 *                   the metric: 1.97 instead of 7.8 TFLOP at N=768, naux=2432.
 *                   metric_mode tells how M was made (CHOLESKY: M is lower
 *                   triangular and its zero half is skipped).  With EIGEN the
 *                   metric product runs in FP64 whatever prec asks: the dense
 *                   J^-1/2 amplifies FP32 rounding (4 uEh in EMP2 at (H2O)32).
 *                   (P|mu nu) must be symmetric in (mu, nu), as 3-index
 *                   Coulomb integrals are: nu is contracted with the occupied
 *                   and mu with the virtual block, and Bia is written by the
 *                   metric GEMM in place (no transpose).
 * Precision: CORR_PREC_FP64 and CORR_PREC_FP32 run every GEMM in that
 * precision (the metric product of EIGEN always in FP64).  CORR_PREC_BF16X3
 * asks for the mixed path's B formation, with no FP32 rounding (the AO->MO
 * sums cancel: FP32 leaves ~0.3 uEh per GEMM in EMP2 at (H2O)32): in
 * corr_ri_form_bia and corr_ri_bia_from_j the first AO->MO GEMM in FP64, the
 * second and the CHOLESKY metric product on int8 MFMA as CORR_PREC_INT8X5 (B
 * within 3e-7 of FP64 relative to its maximum, 0.003 uEh in EMP2);
 * corr_transform_ri runs its GEMMs in FP64 and corr_ri_bpq in FP32. */
int corr_ri_bpq(corr_ctx* ctx, int64_t naux, int64_t nroao, const double* d_M,
                const double* d_Pmn, double* d_BPQ, int prec);
int corr_transform_ri(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc,
                      const double* d_MOs, const double* d_BPQ, double* d_Bia,
                      int prec);
int corr_ri_form_bia(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc,
                     const double* d_MOs, const double* d_Pmn, const double* d_M,
                     int metric_mode, double* d_Bia, int prec);
/* corr_ri_bia_from_j: the arithmetic of calculate_ri + transform_ri
 * (ri.cc:50-69, 102-103, 147-172) from J and (P|mu nu) in one call: the
 * metric factorisation (corr_ri_metric) runs on a second stream of the
 * context, overlapped with the AO->MO GEMMs of corr_ri_form_bia, and the
 * metric GEMM waits for it.  d_M receives the metric factor (NULL: context
 * scratch); h_nkept as in corr_ri_metric (may be NULL). */
int corr_ri_bia_from_j(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc,
                       const double* d_MOs, const double* d_Pmn, const double* d_J,
                       int metric_mode, double* d_M, double* d_Bia, int prec,
                       int64_t* h_nkept);

/* ---- a9: RI-MP2 energy, fused pair-block GEMM + denominator + reduction ---
 * (can_mp2.cc:53-132).  eps_occ[i] = FMo[i,i], eps_vir[a] = FMo[o+a,o+a].
 * Sizes: nocc, nvir < 2^20, and with CORR_PREC_BF16X3 nocc <= 65532 (the
 * tile list's 14-bit occupied block index); CORR_EINVAL otherwise. */
int corr_rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                      const double* d_Bia, const double* d_eps_occ,
                      const double* d_eps_vir, int prec, corr_mp2_energy* h_out);

/* corr_ri_mp2_energy_from_j: corr_ri_bia_from_j followed by corr_rimp2_energy
 * (calculate_ri + transform_ri + run_canonical_mp2_ri, ri.cc:50-172,
 * can_mp2.cc:53-132) without handing B back.  With the Cholesky metric and
 * CORR_PREC_BF16X3 (the mixed path) B is never written out: the int8 metric
 * product stores it as the BF16 hi/lo panels the energy kernel reads (same
 * energies, bit for bit, as the two calls).  Other metric modes and
 * precisions form B in context scratch and call the energy.  h_nkept as in
 * corr_ri_metric (may be NULL). */
int corr_ri_mp2_energy_from_j(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc,
                              const double* d_MOs, const double* d_Pmn, const double* d_J,
                              int metric_mode, const double* d_eps_occ,
                              const double* d_eps_vir, int prec, corr_mp2_energy* h_out,
                              int64_t* h_nkept);

/* corr_rimp2_energy_noalloc: corr_rimp2_energy without any device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture setup, tight serving loops, harnesses that
 * forbid hipMalloc).  Its workspace is a static device array of
 * CORR_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends; problems needing more return CORR_EINVAL with the size in
 * the error string.  Same kernels and results as corr_rimp2_energy (bit for
 * bit); no kernel timing is recorded.  Calls that overlap on one GPU share the
 * array and race: serialise them. */
#define CORR_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int corr_rimp2_energy_noalloc(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                              const double* d_Bia, const double* d_eps_occ,
                              const double* d_eps_vir, int prec, corr_mp2_energy* h_out);

/* ---- RI-MP2 energy gradient (no reference counterpart: the reference has no
 * gradient).  Writes dE/dBia (same layout as Bia), dE/d eps_occ, dE/d eps_vir
 * for E = EMP2 (c_os = c_ss = 1) or any (c_os, c_ss) spin-component scaling.
 * Pair blocks are recomputed, the o^2 v^2 amplitudes are never stored. */
int corr_rimp2_gradient(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                        const double* d_Bia, const double* d_eps_occ,
                        const double* d_eps_vir, double c_os, double c_ss,
                        int prec, double* d_dBia, double* d_deps_occ,
                        double* d_deps_vir, corr_mp2_energy* h_out);

/* ---- a10-a19: spin-orbital CCSD (ccsd.cc:6-580, ccsd_intermediates.cc) ----
 * Inputs: nroao spatial MOs, nroe electrons, prec_ints (ij|kl) over spatial
 * MOs (N^4), MOs, Hmat (AO core Hamiltonian), ion_rep.
 * Outputs: per-iteration energies (h_energies[maxiter]), iterations done,
 * final T1 (o x v) and T2 (o^2 v^2) in the reference layout (device).
 * mode CORR_CCSD_REFERENCE reproduces ccsd_ur: maxiter Jacobi updates, the
 * energy printed at the top of each iteration; CORR_CCSD_DIIS adds DIIS(8)
 * and stops when |dE| < e_tol and the residual norm < r_tol.
 * h_emp2 (optional) receives the spin-orbital MP2 energy of the guess. */
typedef struct corr_ccsd_opts {
    int    mode;          /* CORR_CCSD_REFERENCE or CORR_CCSD_DIIS          */
    int    maxiter;       /* 20 in the reference                            */
    int    diis_size;     /* subspace size for CORR_CCSD_DIIS (8)           */
    int    prec;          /* arithmetic of the ladder/ring/T1 contractions  */
    double e_tol;         /* energy change threshold (DIIS mode)            */
    double r_tol;         /* amplitude-change norm threshold (DIIS mode)    */
} corr_ccsd_opts;

typedef struct corr_ccsd_result {
    double e_hf_check;    /* El of calc_e_check (ccsd.cc:50-81), + ion_rep  */
    double e_one_check;   /* "one:" of calc_e_check: sum_i h_ii + ion_rep  */
    double e_fsum_check;  /* "fsum-0.5twoe:": sum_i f_ii - twoe/2 + ion_rep */
    double e_mp2;         /* spin-orbital MP2 of the guess (ccsd.cc:131-143) */
    double e_ccsd;        /* last energy evaluated                           */
    double d_e;           /* last energy change (E_old - E)                  */
    double r_norm;        /* last amplitude-change norm                      */
    int    iterations;    /* energy evaluations done                         */
    int    converged;     /* DIIS mode: 1 if both thresholds met             */
} corr_ccsd_result;

int corr_ccsd(corr_ctx* ctx, int64_t nroao, int64_t nroe,
              const double* d_prec_ints, const double* d_MOs,
              const double* d_Hmat, double ion_rep,
              const corr_ccsd_opts* opts, double* h_energies,
              double* d_T1, double* d_T2, corr_ccsd_result* h_out);

/* Building blocks of corr_ccsd, exported for parity testing. */
/* ints_so[pqrs] = <pq||rs> over 2N spin orbitals (ccsd.cc:6-9, 509-535) */
int corr_ccsd_ints_so(corr_ctx* ctx, int64_t nroao, const double* d_prec_ints,
                      double* d_ints_so);
/* One CCSD update from (T1, T2) -> (T1n, T2n) with ints_so and the
 * spin-orbital Fock matrix f (norb x norb), exactly one pass of
 * ccsd_build_taus / ccsd_build_intermediates / ccsd_update_T1 / _T2. */
int corr_ccsd_update(corr_ctx* ctx, int64_t nocc, int64_t nvir,
                     const double* d_ints_so, const double* d_f,
                     const double* d_T1, const double* d_T2,
                     double* d_T1n, double* d_T2n, int prec);
/* E(CCSD) of ccsd_energy (ccsd.cc:147-219) for given amplitudes. */
int corr_ccsd_energy(corr_ctx* ctx, int64_t nocc, int64_t nvir,
                     const double* d_ints_so, const double* d_f,
                     const double* d_T1, const double* d_T2, double* h_e);

/* ---- RHF input side (the step before the post-HF path) --------------------
 * Fock build F = H + J(P) - 1/2 K(P), replacing build_Fmat
 * (src/SCF/ops_rhf.cc:549-585 with the perm_* image rules at :466-547).
 * corr_build_fmat reads the dense AO tensor of corr_ao_unfold (nroao <= 512;
 * it relies on the tensor's (pq)<->(qp) and (rs)<->(sr) symmetry, which the
 * unfold guarantees); corr_build_fmat_list takes the reference's own
 * arguments (sorted unique list + sortcount) and unfolds into context
 * scratch (nroao^4 doubles) on every call.  F must not alias P or H. */
int corr_build_fmat(corr_ctx* ctx, int64_t nroao, const double* d_dense_ao,
                    const double* d_Pmat, const double* d_Hmat,
                    double* d_Fmat);
int corr_build_fmat_list(corr_ctx* ctx, int64_t nroao, int64_t nrofint,
                         const int64_t* h_sortcount, const double* d_intval,
                         const uint16_t* d_intnums, const double* d_Pmat,
                         const double* d_Hmat, double* d_Fmat);
/* P = 2(1-damp) C_occ^T C_occ + damp P (in place), nocc = nroe/2
 * (build_Pmat, src/SCF/scf.cc:44-61). */
int corr_build_pmat(corr_ctx* ctx, int64_t nroao, int64_t nroe,
                    const double* d_MOs, double damp, double* d_Pmat);

/* Damped RHF with DIIS (run_scf, src/SCF/scf.cc:145-320), preceded by
 * calc_S12 (ops_rhf.cc:342-373) and, when core_guess_check is set,
 * check_initial_guess (ops_rhf.cc:15-37).  d_MOs holds the starting orbitals
 * (row = MO) and receives the final ones; d_MOens the orbital energies of the
 * last diagonalisation; d_Pmat and d_Fmat the density and Fock matrix rebuilt
 * after the loop, as the reference leaves them for build_FMo.
 * h_energies and h_errors (optional, maxiter doubles each) receive each
 * iteration's ESCF (electronic + ion_rep) and mean |S P F - F P S|.  Fails with CORR_ENOTPD when S is not positive
 * definite.  Defaults of the reference: maxiter 100 (main.cc:114),
 * diis_size 8, diis_thresh 1e-3, damp 0.5, e_tol 1e-8 (scf.cc:170-190, 207). */
typedef struct corr_scf_opts {
    int    maxiter;
    int    diis_size;         /* 1..15                                        */
    int    core_guess_check;  /* 1: core guess when |trace C| < 0.1           */
    int    reserved;
    double diis_thresh;       /* mean |S P F - F P S| that enables DIIS       */
    double damp;              /* density damping until DIIS starts            */
    double e_tol;             /* stop when |E - E_old| <= e_tol               */
} corr_scf_opts;

typedef struct corr_scf_result {
    double e_scf;         /* last iteration's ESCF (electronic + ion_rep)     */
    double e_one;         /* "ONE E:" line, tr(P H) + ion_rep                 */
    double e_two;         /* "TWO E:" line, (tr(P F) - tr(P H)) / 2           */
    double e_total;       /* "TOTAL E:" line, (tr(P H) + tr(P F))/2 + ion_rep */
    double last_de;       /* last energy change                               */
    double last_error;    /* last mean |S P F - F P S|                        */
    int    iterations;
    int    converged;     /* |last_de| <= e_tol                               */
    int    diis;          /* DIIS extrapolation was used                      */
    int    diis_start;    /* iteration at which it started                    */
    int    diis_singular; /* extrapolations skipped: singular DIIS system     */
    int    core_guess;    /* the core guess replaced the input orbitals       */
} corr_scf_result;

int corr_run_scf(corr_ctx* ctx, int64_t nroao, int64_t nroe,
                 const double* d_Hmat, const double* d_Smat,
                 const double* d_dense_ao, double* d_MOs, double* d_MOens,
                 double* d_Pmat, double* d_Fmat, double ion_rep,
                 const corr_scf_opts* opts, double* h_energies,
                 double* h_errors, corr_scf_result* h_out);

/* RI-JK variants: J and K from B^Q_mu nu (naux x N x N, e.g. corr_ri_bpq of
 * the metric and (P|mu nu)) instead of the full integral list:
 *   J = sum_Q B^Q tr(B^Q P),  K = sum_Q B^Q P B^Q.
 * No counterpart in the reference, whose SCF needs the full integral list (at
 * (H2O)32, N = 768, the dense tensor would take 2.8 TB); they make the RHF
 * in front of the post-HF path possible at that size.  Scratch: naux N^2
 * doubles. */
int corr_build_fmat_ri(corr_ctx* ctx, int64_t nroao, int64_t naux,
                       const double* d_BPQ, const double* d_Pmat,
                       const double* d_Hmat, double* d_Fmat);
int corr_run_scf_ri(corr_ctx* ctx, int64_t nroao, int64_t nroe,
                    const double* d_Hmat, const double* d_Smat, int64_t naux,
                    const double* d_BPQ, double* d_MOs, double* d_MOens,
                    double* d_Pmat, double* d_Fmat, double ion_rep,
                    const corr_scf_opts* opts, double* h_energies,
                    double* h_errors, corr_scf_result* h_out);

/* ---- non-canonical MP2 (run_non_canonical_mp2, loc_mp2.cc:7-92) ----------
 * Semi-canonical MP2 amplitudes with the off-diagonal occupied block of FMo:
 * maxiter Jacobi updates (20 in the reference).  h_energies[maxiter] and
 * h_rsum[maxiter] receive each iteration's "E(sem-loc)" and "Rsum" (computed
 * before that iteration's update, as printed by the reference).  d_T
 * (optional, nocc^2 nvir^2 doubles, [i][j][a][b]) receives the final
 * amplitudes.  Scratch: 4 nocc^2 nvir^2 doubles. */
int corr_non_canonical_mp2(corr_ctx* ctx, int64_t nroao, int64_t nocc,
                           const double* d_prec_ints, const double* d_FMo,
                           int maxiter, double* h_energies, double* h_rsum,
                           double* d_T);

/* ---- Gaussian integrals (replacing the reference's libint2 calls) ---------
 * A basis is a list of contracted spherical shells in plain host arrays:
 * shell s has angular momentum l[s] (0..4), nprim[s] primitives and centre
 * center[3s..3s+2] (bohr); exps and coefs hold the primitives of all shells in
 * order, coefs being the contraction coefficients of normalised primitives as
 * basis-set files (.g94) list them.  Functions are numbered shell by shell,
 * 2l+1 per shell, m = -l..l (real solid harmonics); every contracted function
 * has unit self-overlap, as libint2 normalises them.  Orbital shells up to d.
 *   corr_int1e  S, T, V (N x N)        calculate_libint_oei, ops_rhf.cc:77-103
 *   corr_int2c  (P|Q) (naux x naux)    calculate_ri, ri.cc:25-60
 *   corr_int3c  (P|mu nu) [naux][N][N] calculate_ri, ri.cc:62-100
 *   corr_int4c  dense (mu nu|la si)     calculate_libint_tei, ops_rhf.cc:618-723
 *               (the reference keeps the sorted unique list instead)
 *   corr_int1e_dipole  D_x, D_y, D_z [3][N][N] about h_origin: psi4's
 *               MintsHelper.ao_dipole in the Boys localisation,
 *               syshfwplug/pymodule.py:156-160 */
typedef struct corr_basis {
    int64_t        nshell;
    const int32_t* l;
    const int32_t* nprim;
    const double*  center;
    const double*  exps;
    const double*  coefs;
} corr_basis;

int corr_int1e(corr_ctx* ctx, const corr_basis* obs, int64_t natom,
               const double* h_charges, const double* h_coords,
               double* d_S, double* d_T, double* d_V);
int corr_int2c(corr_ctx* ctx, const corr_basis* aux, double* d_J);
int corr_int3c(corr_ctx* ctx, const corr_basis* aux, const corr_basis* obs,
               double* d_Pmn);
int corr_int4c(corr_ctx* ctx, const corr_basis* obs, double* d_dense);
int corr_int1e_dipole(corr_ctx* ctx, const corr_basis* obs, const double* h_origin,
                      double* d_D);

/* ---- orbital localisation (syshfwplug/pymodule.py:75-167) -----------------
 * localize_PM / localize_boys of the reference: maximise
 * f = sum_K sum_i (Q_K)_ii^2 over rotations of the nocc orbitals in d_C
 * (nocc x nbf, one orbital per row, as run_scf's MOs), with
 *   PIPEK-MEZEY (calculate_PM, :143-153): d_M = S (nbf x nbf) and
 *     h_bf_atom[nbf] the atom of every function (natom atoms); Q_A holds the
 *     Mulliken populations of atom A;
 *   BOYS (calculate_boys, :156-167): d_M = D_x, D_y, D_z (3 x nbf x nbf),
 *     h_bf_atom unused.
 * The reference maximises with BFGS over U = expm(X); here parallel Jacobi
 * sweeps (round-robin pairs) run until a sweep raises f by at most
 * tol * max(f, 1) or max_sweeps is reached.  d_L receives the localised
 * orbitals (same layout as d_C; order and signs are the algorithm's). */
#define CORR_LOC_PM    0
#define CORR_LOC_BOYS  1

typedef struct corr_loc_opts {
    int    max_sweeps;   /* 100 */
    double tol;          /* 1e-14 */
} corr_loc_opts;

typedef struct corr_loc_result {
    double spread_initial;
    double spread;
    int    sweeps;
    int    converged;
} corr_loc_result;

int corr_localize(corr_ctx* ctx, int64_t nbf, int64_t nocc, const double* d_C,
                  int kind, const double* d_M, const int32_t* h_bf_atom,
                  int64_t natom, const corr_loc_opts* opts, double* d_L,
                  corr_loc_result* h_res);

#ifdef __cplusplus
}
#endif

#endif /* CORR_ABI_H */
