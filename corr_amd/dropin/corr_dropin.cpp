// Drop-in replacements for the reference's post-HF and SCF entry points (the
// C++ signatures of src/POSTHF/{motrans,can_mp2,ri,ccsd}.h and src/SCF).  Host code only: all
// arithmetic is done by libcorr_gpu.so through the C-ABI.  Ownership follows
// the reference (callee new[] for FMo, prec_ints, ints_so; caller-allocated
// BPQ and Bia), and results are printed with the reference's labels and
// widths.  Errors end the process with a message, as the reference's own
// failure paths do (exit() in ops_io.cc / ops_rhf.cc).
//
// Environment overrides (defaults reproduce the reference):
//   CORR_PREC       fp64 | fp32 | bf16x3   arithmetic of the bulk contractions
//   CORR_CCSD_MODE  reference | diis        20 fixed iterations, or DIIS to 1e-8
#include "../../include/corr_dropin.h"
#include "../../include/corr_abi.h"

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#define PWIDTH_L 12
#define PWIDTH_R 16

namespace {

corr_ctx* g_ctx = nullptr;

[[noreturn]] void die(const char* what, int st) {
    std::cerr << "corr_dropin: " << what << " failed with status " << st << ": "
              << (g_ctx ? corr_ctx_last_error(g_ctx) : "") << std::endl;
    std::exit(3);
}

corr_ctx* ctx() {
    if (!g_ctx) {
        const int st = corr_ctx_create(0, &g_ctx);
        if (st != CORR_OK) die("corr_ctx_create", st);
    }
    return g_ctx;
}

void check(const char* what, int st) {
    if (st != CORR_OK) die(what, st);
}

int prec_from_env() {
    const char* p = std::getenv("CORR_PREC");
    if (!p || !std::strcmp(p, "fp64")) return CORR_PREC_FP64;
    if (!std::strcmp(p, "fp32")) return CORR_PREC_FP32;
    if (!std::strcmp(p, "bf16x3")) return CORR_PREC_BF16X3;
    std::cerr << "corr_dropin: unknown CORR_PREC '" << p << "'" << std::endl;
    std::exit(1);
}

// Device buffer owned by the wrapper call.
struct DBuf {
    void* p = nullptr;
    explicit DBuf(int64_t bytes) { check("corr_dmalloc", corr_dmalloc(ctx(), &p, bytes)); }
    DBuf(const void* host, int64_t bytes) : DBuf(bytes) { check("corr_h2d", corr_h2d(ctx(), p, host, bytes)); }
    ~DBuf() { corr_dfree(ctx(), p); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    double* d() const { return static_cast<double*>(p); }
    void to_host(void* host, int64_t bytes) const { check("corr_d2h", corr_d2h(ctx(), host, p, bytes)); }
};

// Device copy of the caller's prec_ints.  It is uploaded on every call: the
// host array may have been edited, or freed and its address reused, since
// MOtrans wrote it, and the N^4 upload is small next to the N^4 transform.
DBuf* device_prec_ints(const double* host, int64_t N) {
    return new DBuf(host, sizeof(double) * N * N * N * N);
}

void print_mp2(const char* title, const char* rule, const char* scs_title, const char* scs_rule,
               const corr_mp2_energy& e) {
    std::cout << title;
    std::cout << rule;
    std::cout << std::setw(PWIDTH_L) << "EMP2_SS:" << std::setw(PWIDTH_R) << e.e_ss << '\n';
    std::cout << std::setw(PWIDTH_L) << "EMP2_OS:" << std::setw(PWIDTH_R) << e.e_os << '\n';
    std::cout << std::setw(PWIDTH_L) << "EMP2:" << std::setw(PWIDTH_R) << e.e_mp2 << '\n';
    std::cout << scs_title;
    std::cout << scs_rule;
    std::cout << std::setw(PWIDTH_L) << "EMP2_SS:" << std::setw(PWIDTH_R) << e.e_ss / 3.0 << '\n';
    std::cout << std::setw(PWIDTH_L) << "EMP2_OS:" << std::setw(PWIDTH_R) << e.e_os * 6.0 / 5.0 << '\n';
    std::cout << std::setw(PWIDTH_L) << "EMP2(SCS):" << std::setw(PWIDTH_R) << e.e_ss / 3. + e.e_os * 6.0 / 5.0
              << '\n';
}

void transform_ri_device(int naux, int N, int nroe, const double* MOs, const double* d_BPQ, double* Bia) {
    const int o = nroe / 2, v = N - o;
    DBuf dC(MOs, sizeof(double) * N * N);
    DBuf dB(sizeof(double) * (int64_t)naux * o * v);
    check("corr_transform_ri", corr_transform_ri(ctx(), naux, N, o, dC.d(), d_BPQ, dB.d(), CORR_PREC_FP64));
    dB.to_host(Bia, sizeof(double) * (int64_t)naux * o * v);
}

}  // namespace

// motrans.cc:8-32
void build_FMo(systeminfo* sysinfo, OEints* onemats, pHF* postHF) {
    const int N = sysinfo->nroao;
    postHF->FMo = new double[(size_t)N * N];
    DBuf dC(onemats->MOs, sizeof(double) * N * N), dF(onemats->Fmat, sizeof(double) * N * N);
    DBuf dFMo(sizeof(double) * N * N);
    check("corr_build_fmo", corr_build_fmo(ctx(), N, dC.d(), dF.d(), dFMo.d()));
    dFMo.to_host(postHF->FMo, sizeof(double) * N * N);
}

// motrans.cc:35-211 (the full transform; the reference's debug transform and
// progress prints are not reproduced)
void MOtrans(systeminfo* sysinfo, OEints* onemats, TEints* twomats, pHF* postHF) {
    const int64_t N = sysinfo->nroao, n4 = N * N * N * N, nint = sysinfo->nrofint;
    std::cout << "Need " << n4 * (int64_t)sizeof(double) << " bytes ("
              << n4 * (int64_t)sizeof(double) / 1024 / 1024 << " MB) for precalculation\n";
    postHF->prec_ints = new double[(size_t)n4];
    DBuf dC(onemats->MOs, sizeof(double) * N * N);
    DBuf dval(twomats->intval, sizeof(double) * (nint > 0 ? nint : 1));
    DBuf dnum(twomats->intnums, sizeof(unsigned short) * 4 * (nint > 0 ? nint : 1));
    DBuf dense(sizeof(double) * n4), work(sizeof(double) * n4);
    DBuf out(sizeof(double) * n4);
    const int64_t sc[4] = {twomats->sortcount[0], twomats->sortcount[1], twomats->sortcount[2],
                           twomats->sortcount[3]};
    check("corr_ao_unfold", corr_ao_unfold(ctx(), N, nint, sc, dval.d(), (const uint16_t*)dnum.p, dense.d()));
    check("corr_motrans", corr_motrans(ctx(), N, dC.d(), dense.d(), work.d(), out.d(),
                                       prec_from_env() == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32));
    out.to_host(postHF->prec_ints, sizeof(double) * n4);
}

// can_mp2.cc:11-49
void run_canonical_mp2(systeminfo* sysinfo, pHF* postHF) {
    const int N = sysinfo->nroao, o = sysinfo->nroe / 2;
    DBuf* dI = device_prec_ints(postHF->prec_ints, N);
    DBuf dF(postHF->FMo, sizeof(double) * N * N);
    corr_mp2_energy e;
    check("corr_canonical_mp2", corr_canonical_mp2(ctx(), N, o, dI->d(), dF.d(), &e));
    delete dI;
    print_mp2("\nMP2-Results:\n", "------------\n", "\nSCS-MP2-Results:\n", "------------\n", e);
}

// can_mp2.cc:53-132
void run_canonical_mp2_ri(systeminfo* sysinfo, pHF* postHF) {
    const int N = sysinfo->nroao, o = sysinfo->nroe / 2, v = N - o, naux = sysinfo->naux_2;
    std::vector<double> eo(o), ev(v);
    for (int i = 0; i < o; ++i) eo[i] = postHF->FMo[(size_t)i * N + i];
    for (int a = 0; a < v; ++a) ev[a] = postHF->FMo[(size_t)(a + o) * N + (a + o)];
    DBuf dB(postHF->Bia, sizeof(double) * (int64_t)naux * o * v);
    DBuf deo(eo.data(), sizeof(double) * o), dev(ev.data(), sizeof(double) * v);
    corr_mp2_energy e;
    check("corr_rimp2_energy", corr_rimp2_energy(ctx(), naux, o, v, dB.d(), deo.d(), dev.d(), prec_from_env(), &e));
    print_mp2("\nRI-MP2-Results:\n", "---------------\n", "\nSCS-RI-MP2-Results:\n", "-------------------\n", e);
}

// ri.cc:50-69, 102-103 without the libint2 integral evaluation
void calculate_ri_arith(int naux, int nroao, const double* J, const double* Pmn, double* BPQ) {
    const int64_t N2 = (int64_t)nroao * nroao;
    DBuf dJ(J, sizeof(double) * naux * naux), dM(sizeof(double) * naux * naux);
    DBuf dP(Pmn, sizeof(double) * naux * N2), dBPQ(sizeof(double) * naux * N2);
    int64_t kept = 0;
    check("corr_ri_metric", corr_ri_metric(ctx(), naux, dJ.d(), CORR_METRIC_EIGEN, dM.d(), &kept));
    check("corr_ri_bpq", corr_ri_bpq(ctx(), naux, nroao, dM.d(), dP.d(), dBPQ.d(), CORR_PREC_FP64));
    dBPQ.to_host(BPQ, sizeof(double) * naux * N2);
}

// ri.cc:147-172
void transform_ri(systeminfo* sysinfo, OEints* onemats, pHF* postHF) {
    const int N = sysinfo->nroao, naux = sysinfo->naux_2;
    DBuf dBPQ(postHF->BPQ, sizeof(double) * (int64_t)naux * N * N);
    transform_ri_device(naux, N, sysinfo->nroe, onemats->MOs, dBPQ.d(), postHF->Bia);
}

// ri.cc:109-145: BPQ read from PREFIX.rimp2
void read_transform_ri(std::string prefix, int nroe, int nroao, int naux_2, double* MOs, double* Bia) {
    const int64_t count = (int64_t)naux_2 * nroao * nroao;
    std::vector<double> BPQ(count);
    std::ifstream datf(prefix + ".rimp2", std::ios::binary);
    datf.read(reinterpret_cast<char*>(BPQ.data()), count * sizeof(double));
    if (!datf) {
        std::cerr << "corr_dropin: cannot read " << count << " doubles from " << prefix << ".rimp2" << std::endl;
        std::exit(3);
    }
    DBuf dBPQ(BPQ.data(), sizeof(double) * count);
    transform_ri_device(naux_2, nroao, nroe, MOs, dBPQ.d(), Bia);
}

// loc_mp2.cc:7-92
void run_non_canonical_mp2(int nroao, int nroe, double* prec_ints, double* FMo) {
    const int64_t N = nroao, o = nroe / 2;
    std::cout << "\nCalculating semi-canonical amplitudes ... ";
    DBuf* dI = device_prec_ints(prec_ints, N);
    DBuf dF(FMo, sizeof(double) * N * N);
    std::vector<double> E(20), R(20);
    check("corr_non_canonical_mp2", corr_non_canonical_mp2(ctx(), N, o, dI->d(), dF.d(), 20, E.data(), R.data(), nullptr));
    delete dI;
    std::cout << "done" << '\n';
    for (int it = 0; it < 20; ++it)
        std::cout << std::fixed << std::setw(10) << std::setprecision(10) << "E(sem-loc):" << std::setw(16) << E[it]
                  << '\t' << "Rsum:" << R[it] << '\n';
}

// ops_rhf.cc:549-585
void build_Fmat(int nroao, double* Fmat, double* Pmat, double* Hmat, double* intval, unsigned short* intnums,
                long long int* sortcount, long long int nrofint) {
    const int64_t N = nroao, nint = nrofint;
    DBuf dval(intval, sizeof(double) * (nint > 0 ? nint : 1));
    DBuf dnum(intnums, sizeof(unsigned short) * 4 * (nint > 0 ? nint : 1));
    DBuf dP(Pmat, sizeof(double) * N * N), dH(Hmat, sizeof(double) * N * N), dF(sizeof(double) * N * N);
    const int64_t sc[4] = {sortcount[0], sortcount[1], sortcount[2], sortcount[3]};
    check("corr_build_fmat_list", corr_build_fmat_list(ctx(), N, nint, sc, dval.d(), (const uint16_t*)dnum.p, dP.d(),
                                                       dH.d(), dF.d()));
    dF.to_host(Fmat, sizeof(double) * N * N);
}

// scf.cc:44-61
void build_Pmat(double* C, int nroao, int nroe, double* Pmat, double damp) {
    const int64_t N = nroao;
    DBuf dC(C, sizeof(double) * N * N), dP(Pmat, sizeof(double) * N * N);
    check("corr_build_pmat", corr_build_pmat(ctx(), N, nroe, dC.d(), damp, dP.d()));
    dP.to_host(Pmat, sizeof(double) * N * N);
}

// scf.cc:145-320.  The iteration lines are printed after the GPU loop ends;
// t[s] is the mean wall time of one iteration.
void run_scf(systeminfo* sysinfo, OEints* onemats, TEints* twomats) {
    std::cout << "\nSCF:\n----" << '\n';
    const int64_t N = sysinfo->nroao, n2 = N * N, n4 = n2 * n2, nint = sysinfo->nrofint;
    DBuf dval(twomats->intval, sizeof(double) * (nint > 0 ? nint : 1));
    DBuf dnum(twomats->intnums, sizeof(unsigned short) * 4 * (nint > 0 ? nint : 1));
    DBuf dense(sizeof(double) * n4);
    const int64_t sc[4] = {twomats->sortcount[0], twomats->sortcount[1], twomats->sortcount[2],
                           twomats->sortcount[3]};
    check("corr_ao_unfold", corr_ao_unfold(ctx(), N, nint, sc, dval.d(), (const uint16_t*)dnum.p, dense.d()));
    DBuf dH(onemats->Hmat, sizeof(double) * n2), dS(onemats->Smat, sizeof(double) * n2);
    DBuf dC(onemats->MOs, sizeof(double) * n2), dE(sizeof(double) * N);
    DBuf dP(sizeof(double) * n2), dF(sizeof(double) * n2);
    corr_scf_opts opts;
    opts.maxiter = sysinfo->scfiter > 0 ? sysinfo->scfiter : 100;
    opts.diis_size = 8;
    opts.core_guess_check = 0;
    opts.reserved = 0;
    opts.diis_thresh = 1e-3;
    opts.damp = 0.5;
    opts.e_tol = 1e-8;
    std::vector<double> energies(opts.maxiter), errors(opts.maxiter);
    corr_scf_result r;
    const auto t0 = std::chrono::steady_clock::now();
    check("corr_run_scf", corr_run_scf(ctx(), N, sysinfo->nroe, dH.d(), dS.d(), dense.d(), dC.d(), dE.d(), dP.d(),
                                       dF.d(), sysinfo->ion_rep, &opts, energies.data(), errors.data(), &r));
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    dC.to_host(onemats->MOs, sizeof(double) * n2);
    dE.to_host(onemats->MOens, sizeof(double) * N);
    dP.to_host(onemats->Pmat, sizeof(double) * n2);
    dF.to_host(onemats->Fmat, sizeof(double) * n2);
    std::cout << std::fixed << std::setprecision(10);
    std::cout << std::setw(-4) << "Iter" << std::setw(16) << "ESCF" << std::setw(16) << "DE" << std::setw(16)
              << "ERROR" << std::setw(16) << "t[s]" << "\n";
    double e_prev = 100.0 + sysinfo->ion_rep;
    for (int it = 0; it < r.iterations; ++it) {
        if (r.diis && it == r.diis_start) std::cout << "Starting DIIS \n";
        std::cout << std::setw(4) << it << ":" << std::setw(16) << energies[it] << std::setw(16)
                  << energies[it] - e_prev << std::setw(16) << errors[it] << std::setw(16)
                  << dt / (r.iterations > 0 ? r.iterations : 1) << '\n';
        e_prev = energies[it];
    }
    std::cout << "\nHF-Results:\n";
    std::cout << "-----------\n";
    std::cout << std::setw(PWIDTH_L) << "  ONE E:" << std::setw(PWIDTH_R) << r.e_one << '\n';
    std::cout << std::setw(PWIDTH_L) << "  TWO E:" << std::setw(PWIDTH_R) << r.e_two << '\n';
    std::cout << std::setw(PWIDTH_L) << "TOTAL E:" << std::setw(PWIDTH_R) << r.e_total << '\n';
}

// ccsd.cc:6-9 (index accessor over the spatial (ij|kl) array)
double get_integral(double* ints, long long int& istep, long long int& jstep, long long int& kstep, int& i, int& j,
                    int& k, int& l) {
    return ints[(i / 2) * istep + (k / 2) * jstep + (j / 2) * kstep + l / 2] * (i % 2 == k % 2) * (j % 2 == l % 2) -
           ints[(i / 2) * istep + (l / 2) * jstep + (j / 2) * kstep + k / 2] * (i % 2 == l % 2) * (j % 2 == k % 2);
}

// ccsd.cc:538-580
void ccsd_ur(systeminfo* sysinfo, OEints* onemats, pHF* postHF) {
    const int N = sysinfo->nroao, nocc = sysinfo->nroe, norb = 2 * N, nvir = norb - nocc;
    const char* mode = std::getenv("CORR_CCSD_MODE");
    const bool diis = mode && !std::strcmp(mode, "diis");
    std::cout << "\nCCSD:\n----" << '\n';
    const long long mem_a = 2LL * nocc * nvir + 2LL * nocc * nocc * nvir * nvir;
    const long long mem_int = 1LL * nvir * nvir + 1LL * nocc * nocc + 1LL * nocc * nvir + 1LL * nocc * nocc * nocc * nocc +
                              1LL * nvir * nvir * nvir * nvir + 1LL * nocc * nvir * nvir * nocc +
                              2LL * nocc * nocc * nvir * nvir;
    std::cout << "\n Allocating " << mem_a / 1024 / 1024 << "MB(" << mem_a / 1024 / 1024 / 1024 << "Gb) for amplitues\n";
    std::cout << "Allocating " << mem_int / 1024 / 1024 << "MB(" << mem_int / 1024 / 1024 / 1024 << "Gb) for intermediates\n";
    std::cout.flush();

    DBuf* dI = device_prec_ints(postHF->prec_ints, N);
    DBuf dC(onemats->MOs, sizeof(double) * N * N), dH(onemats->Hmat, sizeof(double) * N * N);
    const int64_t L1 = (int64_t)nocc * nvir, L2 = L1 * L1;
    DBuf dT1(sizeof(double) * L1), dT2(sizeof(double) * L2);
    corr_ccsd_opts opts;
    opts.mode = diis ? CORR_CCSD_DIIS : CORR_CCSD_REFERENCE;
    opts.maxiter = diis ? 200 : 20;
    opts.diis_size = 8;
    const int p = prec_from_env();
    opts.prec = p == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32;
    opts.e_tol = 1e-8;
    opts.r_tol = 1e-6;
    std::vector<double> energies(opts.maxiter);
    corr_ccsd_result r;
    const int st = corr_ccsd(ctx(), N, nocc, dI->d(), dC.d(), dH.d(), sysinfo->ion_rep, &opts, energies.data(), dT1.d(),
                             dT2.d(), &r);
    if (st != CORR_OK && st != CORR_ENOTCONV) die("corr_ccsd", st);
    // <pq||rs> as the reference leaves it in postHF->ints_so (ccsd.cc:514)
    {
        const int64_t n4 = (int64_t)norb * norb * norb * norb;
        postHF->ints_so = new double[(size_t)n4];
        DBuf dV(sizeof(double) * n4);
        check("corr_ccsd_ints_so", corr_ccsd_ints_so(ctx(), N, dI->d(), dV.d()));
        dV.to_host(postHF->ints_so, sizeof(double) * n4);
    }
    delete dI;
    std::cout << "El:" << r.e_hf_check << std::endl;
    std::cout << "one:" << r.e_one_check << std::endl;
    std::cout << "fsum-0.5twoe:" << r.e_fsum_check;
    std::cout << "\n\nEMP2: " << r.e_mp2;
    double e_old = 0.0;
    for (int it = 0; it < r.iterations; ++it) {
        std::cout << it << "E(CCSD)= " << energies[it] << " dE:" << e_old - energies[it] << "  " << "\n";
        e_old = energies[it];
    }
}
