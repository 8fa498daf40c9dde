// Runs the reference's post-HF sequence (src/main.cc:130-175) through the
// drop-in entry points on inputs read from a binary file:
//   int64 N, nroe, naux, nrofint, sortcount[4]; double ion_rep;
//   double MOs[N*N], Fmat[N*N], Hmat[N*N], intval[nrofint];
//   uint16 intnums[4*nrofint]; double J[naux*naux], Pmn[naux*N*N]
// and prints the reference's result lines.  With --scf the file also holds
// Smat[N*N] after Pmn; run_scf then starts from MOs and produces the MOs and
// Fmat the post-HF steps use (main.cc:111-121), and build_Pmat/build_Fmat
// rebuild the Fock matrix from the final MOs ("Fcheck:" line: its trace and
// sum of squares).  With --ncmp2, run_non_canonical_mp2 (loc_mp2.cc, not
// called by main.cc) runs after MOtrans.
#include "../include/corr_dropin.h"

#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

template <typename T>
static void rd(std::ifstream& f, T* p, int64_t n) {
    f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
    if (!f) {
        std::cerr << "dropin_check: short input file" << std::endl;
        std::exit(2);
    }
}

int main(int argc, char** argv) {
    bool scf = false, ncmp2 = false;
    for (int a = 2; a < argc; ++a) {
        const std::string f(argv[a]);
        if (f == "--scf") scf = true;
        else if (f == "--ncmp2") ncmp2 = true;
        else {
            std::cerr << "usage: corr_dropin_check INPUT [--scf] [--ncmp2]" << std::endl;
            return 1;
        }
    }
    if (argc < 2) {
        std::cerr << "usage: corr_dropin_check INPUT [--scf] [--ncmp2]" << std::endl;
        return 1;
    }
    std::ifstream f(argv[1], std::ios::binary);
    int64_t hdr[8];
    rd(f, hdr, 8);
    double ion_rep;
    rd(f, &ion_rep, 1);
    const int64_t N = hdr[0], nroe = hdr[1], naux = hdr[2], nint = hdr[3];
    std::vector<double> MOs(N * N), Fmat(N * N), Hmat(N * N), intval(nint);
    std::vector<unsigned short> intnums(4 * nint);
    std::vector<double> J(naux * naux), Pmn(naux * N * N);
    rd(f, MOs.data(), N * N);
    rd(f, Fmat.data(), N * N);
    rd(f, Hmat.data(), N * N);
    rd(f, intval.data(), nint);
    rd(f, intnums.data(), 4 * nint);
    rd(f, J.data(), naux * naux);
    rd(f, Pmn.data(), naux * N * N);
    std::vector<double> Smat(scf ? N * N : 0), MOens(N), Pmat(N * N);
    if (scf) rd(f, Smat.data(), N * N);

    systeminfo* sys = new systeminfo;
    sys->nroe = (int)nroe;
    sys->nroao = (int)N;
    sys->naux_2 = (int)naux;
    sys->nrofint = nint;
    sys->ion_rep = ion_rep;
    OEints* one = new OEints;
    one->MOs = MOs.data();
    one->Fmat = Fmat.data();
    one->Hmat = Hmat.data();
    one->Smat = Smat.data();
    one->MOens = MOens.data();
    one->Pmat = Pmat.data();
    TEints* two = new TEints;
    two->intval = intval.data();
    two->intnums = intnums.data();
    for (int k = 0; k < 4; ++k) two->sortcount[k] = hdr[4 + k];
    pHF* post = new pHF;
    sys->scfiter = 100;
    std::cout << std::fixed << std::setprecision(10);   // the reference's stream state after SCF
    if (scf) {
        run_scf(sys, one, two);
        std::vector<double> P(N * N, 0.0), F2(N * N);
        build_Pmat(MOs.data(), (int)N, (int)nroe, P.data(), 0.0);
        long long sc[4] = {two->sortcount[0], two->sortcount[1], two->sortcount[2], two->sortcount[3]};
        build_Fmat((int)N, F2.data(), P.data(), Hmat.data(), intval.data(), intnums.data(), sc, nint);
        double tr = 0.0, ss = 0.0;
        for (int64_t i = 0; i < N; ++i) tr += F2[i * N + i];
        for (double x : F2) ss += x * x;
        std::cout << "Fcheck: " << tr << " " << ss << "\n";
    }
    build_FMo(sys, one, post);
    MOtrans(sys, one, two, post);
    run_canonical_mp2(sys, post);
    if (ncmp2) run_non_canonical_mp2((int)N, (int)nroe, post->prec_ints, post->FMo);
    post->BPQ = new double[naux * N * N];
    calculate_ri_arith((int)naux, (int)N, J.data(), Pmn.data(), post->BPQ);
    post->Bia = new double[naux * (nroe / 2) * (N - nroe / 2)];
    transform_ri(sys, one, post);
    run_canonical_mp2_ri(sys, post);
    ccsd_ur(sys, one, post);
    std::cout << "\n--END--" << std::endl;
    return 0;
}
