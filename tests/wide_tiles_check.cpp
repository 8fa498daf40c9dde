// CPU check of the BF16x3 energy kernel's tile list (corr_amd/csrc/wide_tiles.hpp),
// built with g++ (no device code) by tests/test_wide_tiles.py.  For each
// (o, v) given on the command line it enumerates the list exactly as the
// device builder does (per-super-tile counts, then the tiles, the diagonal
// ones after all the others) and checks:
//   - the host count, the per-super-tile counts and the emitted tiles agree;
//   - every field is in range (I, J0, J1 < nIB; z, w < nAC) and the tiles
//     with z < w come first;
//   - coverage: with the kernel's pair weights, every occupied pair i <= j < o
//     and every virtual block pair A <= B is covered with total weight 2
//     (i < j) or 1 (i = j), and every virtual pair once per occupied unit.
// Prints one line per shape; exits 1 on the first failure.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "wide_tiles.hpp"

using namespace corr;

static int fail(const char* what, int o, int v) {
    std::fprintf(stderr, "FAIL o=%d v=%d: %s\n", o, v, what);
    return 1;
}

static int check(int o, int v) {
    const int nIB = (o + 3) / 4, nAC = (v + 31) / 32;
    int64_t nT = 0;
    const WtGrid g = wt_grid(nIB, nAC, nT);
    // per-super-tile counts, as the device scan uses them
    int64_t nOff = 0, nDiag = 0;
    for (int64_t st = 0; st < g.NS; ++st) {
        int n, nd;
        wt_count(g, st, n, nd);
        nOff += n - nd;
        nDiag += nd;
    }
    if (nOff + nDiag != nT) return fail("counts", o, v);
    std::vector<int> x, y, z, w;
    std::vector<int64_t> off, diag;
    for (int64_t st = 0; st < g.NS; ++st) {
        int n, nd;
        wt_count(g, st, n, nd);
        int64_t seen = 0, seenDiag = 0;
        wt_emit(g, st, [&](int xx, int yy, int zz, int ww, bool d) {
            ++seen;
            if (d) ++seenDiag;
            (d ? diag : off).push_back((int64_t)x.size());
            x.push_back(xx);
            y.push_back(yy);
            z.push_back(zz);
            w.push_back(ww);
        });
        if (seen != n || seenDiag != nd) return fail("super-tile count", o, v);
    }
    if ((int64_t)off.size() != nOff || (int64_t)diag.size() != nDiag) return fail("emitted", o, v);
    // the list as stored: off-diagonal tiles, then diagonal ones
    std::vector<int64_t> order(off);
    order.insert(order.end(), diag.begin(), diag.end());
    // coverage[(i, j) pair index][virtual pair index] in weight units
    const int64_t nPair = (int64_t)o * (o + 1) / 2, nVP = (int64_t)nAC * (nAC + 1) / 2;
    std::vector<float> cov((size_t)(nPair * nVP), 0.0f);
    auto pidx = [&](int i, int j) { return (int64_t)i * o - (int64_t)i * (i - 1) / 2 + (j - i); };
    auto vidx = [&](int a, int b) { return (int64_t)a * nAC - (int64_t)a * (a - 1) / 2 + (b - a); };
    for (int64_t k = 0; k < (int64_t)order.size(); ++k) {
        const int64_t t = order[k];
        const int I = x[t] & WT_IMASK, J0 = y[t];
        const bool j1ok = !(x[t] & WT_NOJ1);
        const int J1 = j1ok ? J0 + ((x[t] & WT_STEP2) ? 2 : 1) : J0;
        if (I >= nIB || J0 >= nIB || J1 >= nIB || z[t] >= nAC || w[t] >= nAC) return fail("field range", o, v);
        const bool isDiag = z[t] >= w[t];
        if (isDiag != (k >= (int64_t)off.size())) return fail("diagonal tiles not last", o, v);
        int vp[2][2];
        int nv = 0;
        if (z[t] < w[t]) { vp[0][0] = z[t]; vp[0][1] = w[t]; nv = 1; }
        else if (z[t] > w[t]) { vp[0][0] = vp[0][1] = w[t]; vp[1][0] = vp[1][1] = z[t]; nv = 2; }
        else { vp[0][0] = vp[0][1] = z[t]; nv = 1; }
        for (int half = 0; half < 2; ++half) {
            const int Jb = half ? J1 : J0;
            const bool present = half == 0 || j1ok;
            for (int i = 4 * I; i < 4 * I + 4; ++i)
                for (int j = 4 * Jb; j < 4 * Jb + 4; ++j) {
                    if (i >= o || j >= o) continue;
                    const double wt = wt_pair_weight(I, Jb, i, j, present);
                    if (wt == 0.0) continue;
                    const int a = i < j ? i : j, b = i < j ? j : i;
                    for (int q = 0; q < nv; ++q) cov[(size_t)(pidx(a, b) * nVP + vidx(vp[q][0], vp[q][1]))] += (float)wt;
                }
        }
    }
    for (int i = 0; i < o; ++i)
        for (int j = i; j < o; ++j)
            for (int a = 0; a < nAC; ++a)
                for (int b = a; b < nAC; ++b)
                    if (cov[(size_t)(pidx(i, j) * nVP + vidx(a, b))] != (i < j ? 2.0f : 1.0f))
                        return fail("coverage", o, v);
    std::printf("o=%d v=%d nIB=%d nAC=%d tiles=%lld (diagonal %lld)\n", o, v, nIB, nAC, (long long)nT,
                (long long)nDiag);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3 || argc % 2 == 0) {
        std::fprintf(stderr, "usage: %s o v [o v ...]\n", argv[0]);
        return 2;
    }
    for (int k = 1; k + 1 < argc; k += 2)
        if (check(std::atoi(argv[k]), std::atoi(argv[k + 1]))) return 1;
    return 0;
}
