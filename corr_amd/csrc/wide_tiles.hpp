// Tile list of the BF16x3 RI-MP2 energy kernel (rimp2_energy_bf16x3_wide in
// rimp2.hip), restating run_canonical_mp2_ri's pair loop (can_mp2.cc:82-112)
// as tiles of the o*v x o*v matrix M^{ij}_ab.  Host and device code: the
// device builds the list, the host counts it, and a CPU test enumerates it
// with the same functions (no HIP types here).
//
// Occupied side: a tile pairs the 4-block I with two 4-blocks J0, J1.  Row I
// takes the blocks J >= I two at a time, (I, I + 1), (I + 2, I + 3), ...;
// when an odd count leaves the last block nIB - 1 alone, the pair
// {I, nIB - 1} goes to a leftover tile of row block nIB - 1, two such pairs
// per tile (the energy is symmetric in i and j).  Every block pair I <= J is
// then in exactly one tile, and only the diagonal blocks carry pairs i > j
// (410 tiles per virtual unit at 40 blocks; 420 when every tile took an
// aligned block pair (2G, 2G + 1) and the odd rows wasted half of their first
// tile).  Virtual side: the pairs A < B, and the diagonal blocks two per tile.
//
// Order: super-tiles of WT_GI rows of one parity (I = 2r + par, WT_GI
// consecutive r) x one block-pair column m (J0 = 2m + par, J1 = J0 + 1) x
// (WT_GA x WT_GA virtual blocks, A-group <= B-group), row-major (row group,
// parity, m, group pair), then one super-tile of leftover tiles per group
// pair; tiles inside row-major (unit, A, B).  Consecutive tiles, which run
// together on one XCD, share most of their panels in its L2: a super-tile
// reads 4 row panels and 2 column panels per virtual block (62.5 ms against
// 66.8 ms for plain order at (H2O)32; rows of both parities in one
// super-tile read 3 column panels and lowered the L2 hit rate from 78% to
// 75%).  The diagonal tiles come after all the others, in the same order:
// placed among them they took the L2 hit rate of a (H2O)32 launch from 81%
// to 50% (2.5 times the HBM traffic).
//
// A tile is (x, y, z, w): x = I | (J1 = J0 + 2 ? WT_STEP2 : 0) | (no J1 ?
// WT_NOJ1 : 0), y = J0, and (z, w) the virtual pair (z < w: blocks z and w;
// z > w: the two diagonal blocks w and z; z = w: one diagonal block alone).
#pragma once

#include <algorithm>
#include <cstdint>

#if defined(__HIPCC__)
#define CORR_HD __host__ __device__
#else
#define CORR_HD
#endif

namespace corr {

constexpr int WT_GI = 4, WT_GA = 4;
constexpr int WT_STEP2 = 1 << 14, WT_NOJ1 = 1 << 15, WT_IMASK = WT_STEP2 - 1;

struct WtGrid {
    int nIB, nAC, nM, nSP;
    int64_t NSreg, NS;
};

// Tiles of one occupied unit in the virtual group pair (gA, gB): every
// (A, B) with A < B, and on a diagonal group the diagonal blocks two at a
// time (one tile holds (A, A) and (A + 1, A + 1); an odd last block alone).
CORR_HD inline int wt_npairs(int gA, int gB, int nAC) {
    const int nA = std::min(nAC, (gA + 1) * WT_GA) - gA * WT_GA;
    const int nB = std::min(nAC, (gB + 1) * WT_GA) - gB * WT_GA;
    return gA < gB ? nA * nB : nA * (nA - 1) / 2 + (nA + 1) / 2;
}
CORR_HD inline int wt_ndiagpairs(int gA, int nAC) {
    return (std::min(nAC, (gA + 1) * WT_GA) - gA * WT_GA + 1) / 2;
}

// Sets the virtual group pair (gA, gB) of super-tile st, then calls f(x, y)
// for its occupied units in order.
template <typename F>
CORR_HD inline void wt_units(const WtGrid& g, int64_t st, int& gA, int& gB, F f) {
    const int sp = (int)(st < g.NSreg ? st % g.nSP : st - g.NSreg);
    // (gA, gB), gA <= gB, row-major over nAG groups
    const int nAG = (g.nAC + WT_GA - 1) / WT_GA;
    int r = 0, rem = sp;
    while (rem >= nAG - r) { rem -= nAG - r; ++r; }
    gA = r;
    gB = r + rem;
    if (st < g.NSreg) {
        const int m = (int)((st / g.nSP) % g.nM);
        const int qp = (int)(st / ((int64_t)g.nSP * g.nM));   // row group q and parity: 2 q + par
        const int par = qp & 1, J0 = 2 * m + par;
        for (int r = (qp >> 1) * WT_GI; r < ((qp >> 1) + 1) * WT_GI; ++r) {
            const int I = 2 * r + par;
            if (I < g.nIB && J0 >= I && J0 + 1 <= g.nIB - 1) f(I, J0);
        }
    } else {
        const int last = g.nIB - 1;
        for (int J0 = last & 1; J0 <= last; J0 += 4) f(last | (J0 + 2 <= last ? WT_STEP2 : WT_NOJ1), J0);
    }
}

// Tiles of super-tile st: all of them (n) and the diagonal ones among them (nd).
CORR_HD inline void wt_count(const WtGrid& g, int64_t st, int& n, int& nd) {
    int gA, gB, nu = 0;
    wt_units(g, st, gA, gB, [&](int, int) { ++nu; });
    n = nu * wt_npairs(gA, gB, g.nAC);
    nd = gA == gB ? nu * wt_ndiagpairs(gA, g.nAC) : 0;
}

// Calls emit(x, y, z, w, diagonal) for the tiles of super-tile st in order.
template <typename E>
CORR_HD inline void wt_emit(const WtGrid& g, int64_t st, E emit) {
    int gA, gB;
    wt_units(g, st, gA, gB, [&](int x, int y) {
        const int aEnd = std::min(g.nAC, (gA + 1) * WT_GA), bEnd = std::min(g.nAC, (gB + 1) * WT_GA);
        for (int A = gA * WT_GA; A < aEnd; ++A) {
            if (gA == gB && (A - gA * WT_GA) % 2 == 0) emit(x, y, A + 1 < aEnd ? A + 1 : A, A, true);
            for (int B = std::max(A + 1, gB * WT_GA); B < bEnd; ++B) emit(x, y, A, B, false);
        }
    });
}

// The grid of an nIB x nAC problem and its tile count.
inline WtGrid wt_grid(int nIB, int nAC, int64_t& nT) {
    WtGrid g;
    const int nQ = ((nIB + 1) / 2 + WT_GI - 1) / WT_GI, nAG = (nAC + WT_GA - 1) / WT_GA;
    g.nIB = nIB;
    g.nAC = nAC;
    g.nM = (nIB + 1) / 2;
    g.nSP = nAG * (nAG + 1) / 2;
    g.NSreg = (int64_t)nQ * 2 * g.nM * g.nSP;
    g.NS = g.NSreg + g.nSP;
    nT = 0;
    for (int64_t st = 0; st < g.NS; ++st) {
        int n, nd;
        wt_count(g, st, n, nd);
        nT += n;
    }
    return g;
}

// Weight of the occupied pair (i, j) of a tile's blocks (I, Jb): every block
// pair I != Jb is in one tile only and stands for (i, j) and (j, i); inside a
// diagonal block the pairs i < j count twice, i = j once, i > j not at all.
CORR_HD inline double wt_pair_weight(int I, int Jb, int i, int j, bool present) {
    if (!present) return 0.0;
    if (I != Jb) return 2.0;
    return i < j ? 2.0 : (i == j ? 1.0 : 0.0);
}

}  // namespace corr
