/*
 * corr_dropin.h — C++ drop-in replacements for the reference's post-HF entry
 * points (libcorr_dropin.so), with the reference's exact signatures.  They
 * keep the reference's memory ownership and print the reference's result
 * lines; the arithmetic runs on the GPU through the C-ABI of corr_abi.h.
 *
 * The data structs mirror src/UTIL/util.h:7-65 member for member, so objects
 * allocated by the reference driver (src/main.cc) are read with the same
 * layout.  util.h holds std::vector<libint2::Atom>; a vector's layout does not
 * depend on its element type, so the mirror uses a local Atom with the same
 * members as libint2::Atom (atomic_number, x, y, z).  A build that has libint2
 * defines CORR_DROPIN_USE_REFERENCE_UTIL and includes the reference's util.h
 * instead.
 */
#ifndef CORR_DROPIN_H
#define CORR_DROPIN_H

#include <string>
#include <vector>

#ifdef CORR_DROPIN_USE_REFERENCE_UTIL
#include "util.h"
#else
namespace corr_dropin {
struct Atom {
    int atomic_number;
    double x, y, z;
};
}  // namespace corr_dropin

struct pHF {                  /* util.h:7-13 */
    double* FMo;
    double* prec_ints;
    double* ints_so;
    double* BPQ;
    double* Bia;
};

struct systeminfo {           /* util.h:16-45 */
    int nroe;
    int nroao;
    int nroa;
    int naux_1;
    int naux_2;
    long long int nrofint;
    long long int nrofaux;
    long long int nrofaux2;
    std::string basisNameOB;
    std::string basisNameJK;
    std::string basisNameRI;
    double* coord;
    double* charges;
    double* zeff;
    double* mass;
    std::vector<corr_dropin::Atom> atoms;
    double ion_rep;
    int scfiter;
    std::string prefix;
};

struct OEints {               /* util.h:47-59 */
    double* Hmat;
    double* Tmat;
    double* Smat;
    double* Som12;
    double* Vmat;
    double* MOens;
    double* MOs;
    double* Pmat;
    double* Fmat;
};

struct TEints {               /* util.h:61-65 */
    double* intval;
    unsigned short* intnums;
    long long int sortcount[4];
};
#endif

/* src/POSTHF/motrans.h:21-22 — FMo = C F C^T; callee allocates postHF->FMo. */
void build_FMo(systeminfo* sysinfo, OEints* onemats, pHF* postHF);
/* src/POSTHF/motrans.h:21 — (ij|kl) for all N^4; callee allocates postHF->prec_ints. */
void MOtrans(systeminfo* sysinfo, OEints* onemats, TEints* twomats, pHF* postHF);
/* src/POSTHF/motrans.h:12-17 — reads PREFIX.rimp2 (BPQ) and writes Bia. */
void read_transform_ri(std::string prefix, int nroe, int nroao, int naux_2, double* MOs, double* Bia);
/* src/POSTHF/can_mp2.h:4-5 */
void run_canonical_mp2(systeminfo* sysinfo, pHF* postHF);
void run_canonical_mp2_ri(systeminfo* sysinfo, pHF* postHF);
/* src/POSTHF/ri.h:12 — Bia from BPQ; caller allocates postHF->Bia. */
void transform_ri(systeminfo* sysinfo, OEints* onemats, pHF* postHF);
/* Arithmetic of src/POSTHF/ri.cc:50-69,102-103 (calculate_ri without the
 * libint2 integral evaluation): J (naux x naux) and (P|mu nu)
 * (naux x N x N) in, BPQ = J^-1/2 (P|mu nu) out (caller-allocated). */
void calculate_ri_arith(int naux, int nroao, const double* J, const double* Pmn, double* BPQ);
/* src/SCF/ops_rhf.h — F = H + J(P) - 1/2 K(P) from the sorted unique list. */
void build_Fmat(int nroao, double* Fmat, double* Pmat, double* Hmat, double* intval, unsigned short* intnums,
                long long int* sortcount, long long int nrofint);
/* src/SCF/scf.cc:44 — P = 2(1-damp) C_occ^T C_occ + damp P. */
void build_Pmat(double* C, int nroao, int nroe, double* Pmat, double damp);
/* src/SCF/scf.h:5 — damped RHF + DIIS from onemats->MOs (the caller has run
 * check_initial_guess, as main.cc:111-112 does); leaves MOs, MOens, Pmat and
 * Fmat in onemats as the reference does.  S^-1/2 is recomputed on the GPU. */
void run_scf(systeminfo* sysinfo, OEints* onemats, TEints* twomats);
/* src/POSTHF/loc_mp2.h — 20 semi-canonical iterations, "E(sem-loc)" lines. */
void run_non_canonical_mp2(int nroao, int nroe, double* prec_ints, double* FMo);
/* src/POSTHF/ccsd.h:6-7 */
double get_integral(double* ints, long long int& istep, long long int& jstep, long long int& kstep, int& i,
                    int& j, int& k, int& l);
void ccsd_ur(systeminfo* sysinfo, OEints* onemats, pHF* postHF);

#endif /* CORR_DROPIN_H */
