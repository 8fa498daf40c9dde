// Gaussian integrals over spherical contracted shells on the GPU, replacing
// the reference's libint2 calls (libint2 2.4.2, not vendored):
//
//  calculate_libint_oei   src/SCF/ops_rhf.cc:77-103    S, T, V
//  calculate_libint_tei   src/SCF/ops_rhf.cc:618-723   (mu nu|la si) (dense here)
//  calculate_ri           src/POSTHF/ri.cc:25-100      (P|Q), (P|mu nu)
//
// Scheme: McMurchie-Davidson (J. Comput. Phys. 26, 218 (1978)).  Every
// charge distribution is a set of primitive "Hermite objects": an orbital
// shell pair (a >= b) or one aux shell (paired with a zero-exponent s
// function).  A Hermite object of total order L = la + lb has, per primitive
// pair, an exponent p, a centre P, a weight K (contraction coefficients and
// normalisation) and the table E[m][tuv] of its spherical components m over
// the Hermite Gaussians Lambda_tuv, t + u + v <= L (the spherical transform
// is folded into the table).  A Coulomb integral between two objects is
//   (X_m|Y_n) = sum_prims 2 pi^(5/2) / (p q sqrt(p+q)) K_X K_Y
//               sum_tuv E_X[m][tuv] sum_TUV (-1)^(T+U+V) E_Y[n][TUV] R_{t+T,u+U,v+V}
// with R the Hermite Coulomb integrals of alpha = pq/(p+q), R_PQ (Boys
// function F_n).  Kernels are instantiated per class (LX, LY), so R, the
// partial sums and the accumulators have compile-time indices and stay in
// registers; one thread computes one component m of X against all of Y.
// The same kernel serves (P|Q), (P|mu nu) and (mu nu|la si); only the output
// scatter differs.  Real solid harmonics follow the recursion of Helgaker,
// Joergensen & Olsen, Molecular Electronic-Structure Theory, eq. 6.4.70-72;
// each contracted spherical function has unit self-overlap (libint2's
// convention), with the closed form of the self-overlap folded into K.
#include "common.hpp"
#include <cmath>
#include <vector>

namespace corr {

namespace {

constexpr int LMAX = 4;               // shells up to g
constexpr int NCART_MAX = 15;         // (LMAX+1)(LMAX+2)/2
constexpr int NSPH_MAX = 9;           // 2 LMAX + 1

__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// position of (t,u,v) in the Hermite index set ordered by n = t+u+v, then t
// descending, then u descending
__host__ __device__ constexpr int hidx(int t, int u, int v) {
    return (t + u + v) * (t + u + v + 1) * (t + u + v + 2) / 6 + (u + v) * (u + v + 1) / 2 + v;
}
__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
// component rows of an object of order L (the largest shell pair of that order)
__host__ __device__ constexpr int nrows(int L) { return L == 0 ? 1 : L == 1 ? 3 : L == 2 ? 9 : L == 3 ? 15 : L == 4 ? 25 : L == 5 ? 35 : 49; }

// ---------------------------------------------------------------- host ----

// Real solid harmonics: T[m + l][c] = coefficient of Cartesian c (x-major order
// of (i,j,k), i + j + k = l) in S_lm.
struct SolidHarmonics {
    double T[LMAX + 1][NSPH_MAX][NCART_MAX] = {};
    SolidHarmonics() {
        // polynomial in x^i y^j z^k with i + j + k = l, stored dense [i][j][k]
        typedef std::vector<double> Poly;   // index (i*(LMAX+3) + j)*(LMAX+3) + k
        const int D = LMAX + 3;
        auto at = [&](int i, int j, int k) { return (i * D + j) * D + k; };
        std::vector<std::vector<Poly>> S(LMAX + 1);
        for (int l = 0; l <= LMAX; ++l) S[l].assign(2 * l + 1, Poly(D * D * D, 0.0));
        S[0][0][at(0, 0, 0)] = 1.0;
        auto mulxyz = [&](const Poly& p, int dx, int dy, int dz, double f, Poly& out) {
            for (int i = 0; i + dx < D; ++i)
                for (int j = 0; j + dy < D; ++j)
                    for (int k = 0; k + dz < D; ++k)
                        if (p[at(i, j, k)] != 0.0) out[at(i + dx, j + dy, k + dz)] += f * p[at(i, j, k)];
        };
        for (int l = 0; l < LMAX; ++l) {
            const double d = l == 0 ? 1.0 : 0.0;
            const double f = std::sqrt(std::pow(2.0, d) * (2 * l + 1) / (2.0 * l + 2));
            Poly& top = S[l + 1][2 * (l + 1)];     // m = l + 1
            Poly& bot = S[l + 1][0];               // m = -(l + 1)
            mulxyz(S[l][2 * l], 1, 0, 0, f, top);
            mulxyz(S[l][0], 0, 1, 0, -f * (1 - d), top);
            mulxyz(S[l][2 * l], 0, 1, 0, f, bot);
            mulxyz(S[l][0], 1, 0, 0, f * (1 - d), bot);
            for (int m = -l; m <= l; ++m) {
                const double den = std::sqrt((double)(l + m + 1) * (l - m + 1));
                Poly& o = S[l + 1][m + l + 1];
                mulxyz(S[l][m + l], 0, 0, 1, (2 * l + 1) / den, o);
                if (l >= 1 && std::abs(m) <= l - 1) {
                    const double g = -std::sqrt((double)(l + m) * (l - m)) / den;
                    const Poly& q = S[l - 1][m + l - 1];
                    mulxyz(q, 2, 0, 0, g, o);
                    mulxyz(q, 0, 2, 0, g, o);
                    mulxyz(q, 0, 0, 2, g, o);
                }
            }
        }
        for (int l = 0; l <= LMAX; ++l) {
            int c = 0;
            for (int i = l; i >= 0; --i)
                for (int j = l - i; j >= 0; --j, ++c)
                    for (int m = 0; m < 2 * l + 1; ++m) T[l][m][c] = S[l][m][at(i, j, l - i - j)];
        }
    }
};

const SolidHarmonics& solid() {
    static const SolidHarmonics s;
    return s;
}

double dfact(int n) {
    double r = 1.0;
    for (; n > 1; n -= 2) r *= n;
    return r;
}

struct HShell {
    int l;
    double C[3];
    std::vector<double> e, c;   // c includes primitive and contracted normalisation
    int64_t off;                // first function
};

int load_basis(corr_ctx* ctx, const corr_basis* b, std::vector<HShell>& out, int64_t* nfun) {
    CORR_REQUIRE(ctx, b && b->nshell > 0 && b->l && b->nprim && b->center && b->exps && b->coefs,
                 "integrals: incomplete basis description");
    out.clear();
    int64_t po = 0, fo = 0;
    for (int64_t s = 0; s < b->nshell; ++s) {
        HShell h;
        h.l = b->l[s];
        CORR_REQUIRE(ctx, h.l >= 0 && h.l <= LMAX, "integrals: angular momentum above 4");
        CORR_REQUIRE(ctx, b->nprim[s] > 0, "integrals: shell without primitives");
        for (int k = 0; k < 3; ++k) h.C[k] = b->center[3 * s + k];
        const int l = h.l;
        for (int i = 0; i < b->nprim[s]; ++i) {
            const double a = b->exps[po + i];
            CORR_REQUIRE(ctx, a > 0.0, "integrals: non-positive exponent");
            h.e.push_back(a);
            h.c.push_back(b->coefs[po + i] * std::pow(2 * a / M_PI, 0.75) * std::pow(4 * a, 0.5 * l) /
                          std::sqrt(dfact(2 * l - 1)));
        }
        po += b->nprim[s];
        // self-overlap of S_lm e^{-a r^2} contractions: 4 pi/(2l+1) (2l+1)!!/(2^(l+2) p^(l+1)) sqrt(pi/p)
        double so = 0.0;
        for (size_t i = 0; i < h.e.size(); ++i)
            for (size_t j = 0; j < h.e.size(); ++j) {
                const double p = h.e[i] + h.e[j];
                so += h.c[i] * h.c[j] * 4 * M_PI / (2 * l + 1) * dfact(2 * l + 1) / (std::pow(2.0, l + 2) *
                      std::pow(p, l + 1)) * std::sqrt(M_PI / p);
            }
        CORR_REQUIRE(ctx, so > 0.0, "integrals: shell with zero norm");
        for (auto& c : h.c) c /= std::sqrt(so);
        h.off = fo;
        fo += 2 * l + 1;
        out.push_back(std::move(h));
    }
    *nfun = fo;
    return CORR_OK;
}

// One primitive pair of a Hermite object.
struct Prim {
    double p, P[3], K;
    double A[3], B[3], a, b;   // the two primitives (b = 0 for an aux shell)
    int64_t eoff;              // first double of E[rows][nherm(L)]
};

struct Obj {
    int L, la, lb, ncomp;     // ncomp = (2la+1)(2lb+1)
    int nprim;
    int64_t prim0;
    int64_t fa, fb;           // first functions of the two shells
    int nb;                   // 2lb+1
    int same;                 // a == b (orbital pair on one shell)
};

struct ObjSet {
    std::vector<Obj> obj;
    std::vector<Prim> prim;
    int64_t etab = 0;          // doubles
};

// Orbital shell pairs (one per unordered pair of shells, oriented so that
// la >= lb: the table kernels exist for la >= lb only) with primitive-pair
// screening on the overlap magnitude |K| e^{-mu AB^2} (pi/p)^{3/2}.
void build_pairs(const std::vector<HShell>& sh, ObjSet& S, double thr) {
    for (size_t i0 = 0; i0 < sh.size(); ++i0)
        for (size_t i1 = 0; i1 <= i0; ++i1) {
            const bool keep = sh[i0].l >= sh[i1].l;
            const size_t ia = keep ? i0 : i1, ib = keep ? i1 : i0;
            const HShell &A = sh[ia], &B = sh[ib];
            Obj o{A.l + B.l, A.l, B.l, (2 * A.l + 1) * (2 * B.l + 1), 0, (int64_t)S.prim.size(), A.off, B.off,
                  2 * B.l + 1, ia == ib ? 1 : 0};
            double ab2 = 0;
            for (int k = 0; k < 3; ++k) ab2 += (A.C[k] - B.C[k]) * (A.C[k] - B.C[k]);
            for (size_t i = 0; i < A.e.size(); ++i)
                for (size_t j = 0; j < B.e.size(); ++j) {
                    const double a = A.e[i], b = B.e[j], p = a + b;
                    const double K = A.c[i] * B.c[j];
                    if (std::fabs(K) * std::exp(-a * b / p * ab2) * std::pow(M_PI / p, 1.5) < thr) continue;
                    Prim q;
                    q.p = p;
                    q.K = K;
                    q.a = a;
                    q.b = b;
                    for (int k = 0; k < 3; ++k) {
                        q.P[k] = (a * A.C[k] + b * B.C[k]) / p;
                        q.A[k] = A.C[k];
                        q.B[k] = B.C[k];
                    }
                    q.eoff = S.etab;
                    S.etab += (int64_t)nrows(o.L) * nherm(o.L);
                    S.prim.push_back(q);
                    ++o.nprim;
                }
            if (o.nprim > 0) S.obj.push_back(o);
        }
}

// Aux shells as Hermite objects (partner: zero-exponent s on the same centre).
void build_singles(const std::vector<HShell>& sh, ObjSet& S) {
    for (const HShell& A : sh) {
        Obj o{A.l, A.l, 0, 2 * A.l + 1, 0, (int64_t)S.prim.size(), A.off, 0, 1, 0};
        for (size_t i = 0; i < A.e.size(); ++i) {
            Prim q;
            q.p = A.e[i];
            q.K = A.c[i];
            q.a = A.e[i];
            q.b = 0.0;
            for (int k = 0; k < 3; ++k) q.P[k] = q.A[k] = q.B[k] = A.C[k];
            q.eoff = S.etab;
            S.etab += (int64_t)nrows(o.L) * nherm(o.L);
            S.prim.push_back(q);
            ++o.nprim;
        }
        S.obj.push_back(o);
    }
}

// --------------------------------------------------------------- device ----

struct DevSet {
    const Obj* obj;
    const Prim* prim;
    const double* E;
};

// 1D Hermite coefficients E[i][j][t] of x_A^i x_B^j e^{-a x_A^2 - b x_B^2}
// (the Gaussian overlap factor e^{-mu X_AB^2} is applied once in 3D).
template <int LA, int LB>
__device__ inline void herm1d(double a, double b, double A, double B, double (&E)[LA + 1][LB + 1][LA + LB + 1]) {
    const double p = a + b;
    const double P = (a * A + b * B) / p;
    const double XPA = P - A, XPB = P - B, h = 0.5 / p;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j)
#pragma unroll
            for (int t = 0; t <= LA + LB; ++t) E[i][j][t] = 0.0;
    E[0][0][0] = 1.0;
#pragma unroll
    for (int i = 0; i <= LA; ++i)
#pragma unroll
        for (int j = 0; j <= LB; ++j) {
            if (i == 0 && j == 0) continue;
#pragma unroll
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (i > 0) {
                    v = XPA * E[i - 1][j][t] + (t + 1 <= i - 1 + j ? (t + 1) * E[i - 1][j][t + 1] : 0.0);
                    if (t > 0) v += h * E[i - 1][j][t - 1];
                } else {
                    v = XPB * E[i][j - 1][t] + (t + 1 <= i + j - 1 ? (t + 1) * E[i][j - 1][t + 1] : 0.0);
                    if (t > 0) v += h * E[i][j - 1][t - 1];
                }
                E[i][j][t] = v;
            }
        }
}

// Table rows of one primitive pair: E[ma*nb + mb][hidx(t,u,v)] =
// e^{-mu AB^2} sum_{ca,cb} Ta[ma][ca] Tb[mb][cb] Ex Ey Ez; rows past ncomp zero.
template <int LA, int LB>
__global__ __launch_bounds__(64) void etab_kernel(DevSet s, const int64_t* __restrict__ prims, int64_t n,
                                                  const double* __restrict__ sph, double* __restrict__ E) {
    constexpr int L = LA + LB, NH = nherm(L), NR = nrows(L);
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const Prim q = s.prim[prims[e]];
    double Ex[LA + 1][LB + 1][L + 1], Ey[LA + 1][LB + 1][L + 1], Ez[LA + 1][LB + 1][L + 1];
    herm1d<LA, LB>(q.a, q.b, q.A[0], q.B[0], Ex);
    herm1d<LA, LB>(q.a, q.b, q.A[1], q.B[1], Ey);
    herm1d<LA, LB>(q.a, q.b, q.A[2], q.B[2], Ez);
    double ab2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) ab2 += (q.A[k] - q.B[k]) * (q.A[k] - q.B[k]);
    const double g = q.b > 0.0 ? exp(-q.a * q.b / q.p * ab2) : 1.0;
    double* out = E + q.eoff;
    const double* Ta = sph + LA * NSPH_MAX * NCART_MAX;
    const double* Tb = sph + LB * NSPH_MAX * NCART_MAX;
    for (int r = (2 * LA + 1) * (2 * LB + 1); r < NR; ++r)
        for (int k = 0; k < NH; ++k) out[r * NH + k] = 0.0;
    for (int ma = 0; ma < 2 * LA + 1; ++ma)
        for (int mb = 0; mb < 2 * LB + 1; ++mb) {
            double acc[NH];
#pragma unroll
            for (int k = 0; k < NH; ++k) acc[k] = 0.0;
            int ca = 0;
#pragma unroll
            for (int ia = LA; ia >= 0; --ia)
#pragma unroll
                for (int ja = LA - ia; ja >= 0; --ja, ++ca) {
                    const int ka = LA - ia - ja;
                    const double wa = Ta[ma * NCART_MAX + ca];
                    int cb = 0;
#pragma unroll
                    for (int ib = LB; ib >= 0; --ib)
#pragma unroll
                        for (int jb = LB - ib; jb >= 0; --jb, ++cb) {
                            const int kb = LB - ib - jb;
                            const double w = wa * Tb[mb * NCART_MAX + cb];
#pragma unroll
                            for (int t = 0; t <= ia + ib; ++t)
#pragma unroll
                                for (int u = 0; u <= ja + jb; ++u)
#pragma unroll
                                    for (int v = 0; v <= ka + kb; ++v)
                                        acc[hidx(t, u, v)] += w * Ex[ia][ib][t] * Ey[ja][jb][u] * Ez[ka][kb][v];
                        }
                }
            const int r = ma * (2 * LB + 1) + mb;
#pragma unroll
            for (int k = 0; k < NH; ++k) out[r * NH + k] = g * acc[k];
        }
}

// Boys function F_n(T), n = 0..L: series for F_L and downward recursion for
// T < 12, else F_0 = (1/2) sqrt(pi/T) erf(sqrt T) and upward recursion.
template <int L>
__device__ inline void boys(double T, double (&F)[L + 1]) {
    if (T < 12.0) {
        const double e = exp(-T);
        double term = 1.0 / (2 * L + 1), sum = term;
        for (int k = 1; k < 120; ++k) {
            term *= 2.0 * T / (2 * L + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        F[L] = e * sum;
#pragma unroll
        for (int m = L; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + e) / (2 * m - 1);
    } else {
        const double e = exp(-T), rt = sqrt(T);
        F[0] = 0.886226925452758014 / rt * erf(rt);   // sqrt(pi)/2
        const double h = 0.5 / T;
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) * h;
    }
}

// R[hidx(t,u,v)] = R^0_tuv(alpha, X, Y, Z), t + u + v <= L, computed in place
// from n = L down to 0 (orders high to low, so R^{n+1} entries are read before
// they are overwritten).
template <int L>
__device__ inline void hermite_R(double alpha, double X, double Y, double Z, double (&R)[nherm(L)]) {
    double F[L + 1];
    boys<L>(alpha * (X * X + Y * Y + Z * Z), F);
    double pw[L + 1];
    pw[0] = 1.0;
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
    R[0] = pw[L] * F[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
#pragma unroll
        for (int k = L - n; k >= 1; --k)
#pragma unroll
            for (int t = k; t >= 0; --t)
#pragma unroll
                for (int u = k - t; u >= 0; --u) {
                    const int v = k - t - u;
                    double r;
                    if (t > 0)
                        r = X * R[hidx(t - 1, u, v)] + (t > 1 ? (t - 1) * R[hidx(t - 2, u, v)] : 0.0);
                    else if (u > 0)
                        r = Y * R[hidx(t, u - 1, v)] + (u > 1 ? (u - 1) * R[hidx(t, u - 2, v)] : 0.0);
                    else
                        r = Z * R[hidx(t, u, v - 1)] + (v > 1 ? (v - 1) * R[hidx(t, u, v - 2)] : 0.0);
                    R[hidx(t, u, v)] = r;
                }
        R[0] = pw[n] * F[n];
    }
}

enum { MODE_3C = 0, MODE_2C = 1, MODE_4C = 2 };

struct CoulArgs {
    DevSet X, Y;
    const int32_t* xl;     // object indices of the class
    const int32_t* yl;
    int64_t nx, ny;
    int ncx;               // rows of X handled (max ncomp in the class)
    int triangle;          // only ix >= iy (same list)
    int mode;
    int64_t N, naux;       // output dims
    double* out;
};

template <int LX, int LY>
__global__ __launch_bounds__(128) void coulomb_kernel(CoulArgs a) {
    constexpr int L = LX + LY, NHX = nherm(LX), NHY = nherm(LY), NRY = nrows(LY);
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= a.nx * a.ny * a.ncx) return;
    const int mX = (int)(tid % a.ncx);
    const int64_t rest = tid / a.ncx;
    const int64_t iy = rest % a.ny, ix = rest / a.ny;
    if (a.triangle && ix < iy) return;
    const Obj ox = a.X.obj[a.xl[ix]];
    const Obj oy = a.Y.obj[a.yl[iy]];
    if (mX >= ox.ncomp) return;
    double acc[NRY];
#pragma unroll
    for (int m = 0; m < NRY; ++m) acc[m] = 0.0;
    for (int px = 0; px < ox.nprim; ++px) {
        const Prim qx = a.X.prim[ox.prim0 + px];
        double h[NHX];
        const double* ex = a.X.E + qx.eoff + (int64_t)mX * NHX;
#pragma unroll
        for (int k = 0; k < NHX; ++k) h[k] = ex[k];
        for (int py = 0; py < oy.nprim; ++py) {
            const Prim qy = a.Y.prim[oy.prim0 + py];
            const double p = qx.p, q = qy.p, alpha = p * q / (p + q);
            double R[nherm(L)];
            hermite_R<L>(alpha, qx.P[0] - qy.P[0], qx.P[1] - qy.P[1], qx.P[2] - qy.P[2], R);
            // g[TUV] = (-1)^(T+U+V) sum_tuv h[tuv] R[t+T, u+U, v+V]
            double g[NHY];
#pragma unroll
            for (int nY = 0; nY <= LY; ++nY)
#pragma unroll
                for (int T = nY; T >= 0; --T)
#pragma unroll
                    for (int U = nY - T; U >= 0; --U) {
                        const int V = nY - T - U;
                        double s = 0.0;
#pragma unroll
                        for (int nX = 0; nX <= LX; ++nX)
#pragma unroll
                            for (int t = nX; t >= 0; --t)
#pragma unroll
                                for (int u = nX - t; u >= 0; --u) {
                                    const int v = nX - t - u;
                                    s += h[hidx(t, u, v)] * R[hidx(t + T, u + U, v + V)];
                                }
                        g[hidx(T, U, V)] = (nY & 1) ? -s : s;
                    }
            const double pref = 34.986836655249725693 / (p * q * sqrt(p + q)) * qx.K * qy.K;   // 2 pi^(5/2)
            const double* ey = a.Y.E + qy.eoff;
#pragma unroll
            for (int m = 0; m < NRY; ++m) {
                double s = 0.0;
#pragma unroll
                for (int k = 0; k < NHY; ++k) s += ey[m * NHY + k] * g[k];
                acc[m] += pref * s;
            }
        }
    }
    const int nby = oy.nb;
    if (a.mode == MODE_3C) {
        // X: aux shell (function ox.fa + mX); Y: orbital pair (a, b)
        const int64_t P = ox.fa + mX;
        double* o = a.out + P * a.N * a.N;
#pragma unroll
        for (int m = 0; m < NRY; ++m) {
            if (m >= oy.ncomp) break;
            const int64_t i = oy.fa + m / nby, j = oy.fb + m % nby;
            o[i * a.N + j] = acc[m];
            o[j * a.N + i] = acc[m];
        }
    } else if (a.mode == MODE_2C) {
        const int64_t P = ox.fa + mX;
#pragma unroll
        for (int m = 0; m < NRY; ++m) {
            if (m >= oy.ncomp) break;
            const int64_t Q = oy.fa + m;
            a.out[P * a.naux + Q] = acc[m];
            a.out[Q * a.naux + P] = acc[m];
        }
    } else {
        const int64_t N = a.N;
        const int64_t i = ox.fa + mX / ox.nb, j = ox.fb + mX % ox.nb;
#pragma unroll
        for (int m = 0; m < NRY; ++m) {
            if (m >= oy.ncomp) break;
            const int64_t k = oy.fa + m / nby, l = oy.fb + m % nby;
            const double v = acc[m];
            a.out[((i * N + j) * N + k) * N + l] = v;
            a.out[((j * N + i) * N + k) * N + l] = v;
            a.out[((i * N + j) * N + l) * N + k] = v;
            a.out[((j * N + i) * N + l) * N + k] = v;
            a.out[((k * N + l) * N + i) * N + j] = v;
            a.out[((l * N + k) * N + i) * N + j] = v;
            a.out[((k * N + l) * N + j) * N + i] = v;
            a.out[((l * N + k) * N + j) * N + i] = v;
        }
    }
}

// One-electron integrals of one component (ma, mb) of an orbital pair:
// S and T from 1D overlaps (Obara-Saika recurrence), V from the Hermite table.
template <int L>
__device__ inline double nuclear_one(const Prim& q, const double* e, int64_t ncen, const double* Z,
                                     const double* Cxyz) {
    constexpr int NH = nherm(L);
    double s = 0.0;
    for (int64_t c = 0; c < ncen; ++c) {
        double R[NH];
        hermite_R<L>(q.p, q.P[0] - Cxyz[3 * c], q.P[1] - Cxyz[3 * c + 1], q.P[2] - Cxyz[3 * c + 2], R);
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < NH; ++k) t += e[k] * R[k];
        s -= Z[c] * t;
    }
    return s * 6.283185307179586477 / q.p;
}

// 1D overlap factors o1[d][i][j] = <x_A^i | x_B^j> along d, for i <= la and
// j <= jmax, without the sqrt(pi/p) e^{-mu X_AB^2} factor (Obara-Saika).
__device__ inline void overlap_1d(const Prim& q, int la, int jmax, double (&o1)[3][LMAX + 1][LMAX + 3]) {
    for (int d = 0; d < 3; ++d) {
        const double XPA = q.P[d] - q.A[d], XPB = q.P[d] - q.B[d], h = 0.5 / q.p;
        for (int i = 0; i <= la; ++i)
            for (int j = 0; j <= jmax; ++j) {
                double v;
                if (i == 0 && j == 0) v = 1.0;
                else if (i > 0) v = XPA * o1[d][i - 1][j] + h * ((i - 1) * (i > 1 ? o1[d][i - 2][j] : 0.0) +
                                                             j * (j > 0 ? o1[d][i - 1][j - 1] : 0.0));
                else v = XPB * o1[d][i][j - 1] + h * (j - 1 > 0 ? (j - 1) * o1[d][i][j - 2] : 0.0);
                o1[d][i][j] = v;
            }
    }
}

// <a| r_d - O_d |b> for d = x, y, z: along d, x - O = (x - B) + (B - O), so
// the 1D factor is o1[i][j+1] + (B_d - O_d) o1[i][j].  D is [3][N][N].
__global__ __launch_bounds__(128) void dipole_kernel(DevSet s, const int32_t* __restrict__ objs, int64_t nobj,
                                                     const double* __restrict__ sph, double ox, double oy, double oz,
                                                     int64_t N, double* __restrict__ D) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= nobj * 25) return;
    const int m = (int)(tid % 25);
    const Obj o = s.obj[objs[tid / 25]];
    if (m >= o.ncomp) return;
    const int ma = m / o.nb, mb = m % o.nb, la = o.la, lb = o.lb;
    const double* Ta = sph + la * NSPH_MAX * NCART_MAX + ma * NCART_MAX;
    const double* Tb = sph + lb * NSPH_MAX * NCART_MAX + mb * NCART_MAX;
    const double org[3] = {ox, oy, oz};
    double sD[3] = {0.0, 0.0, 0.0};
    for (int pi = 0; pi < o.nprim; ++pi) {
        const Prim q = s.prim[o.prim0 + pi];
        double o1[3][LMAX + 1][LMAX + 3];
        overlap_1d(q, la, lb + 1, o1);
        double ab2 = 0.0;
        for (int k = 0; k < 3; ++k) ab2 += (q.A[k] - q.B[k]) * (q.A[k] - q.B[k]);
        const double pre = q.K * exp(-q.a * q.b / q.p * ab2) * pow(M_PI / q.p, 1.5);
        int ca = 0;
        for (int ia = la; ia >= 0; --ia)
            for (int ja = la - ia; ja >= 0; --ja, ++ca) {
                const double wa = Ta[ca];
                if (wa == 0.0) continue;
                const int ka = la - ia - ja, ai[3] = {ia, ja, ka};
                int cb = 0;
                for (int ib = lb; ib >= 0; --ib)
                    for (int jb = lb - ib; jb >= 0; --jb, ++cb) {
                        const double w = wa * Tb[cb];
                        if (w == 0.0) continue;
                        const int kb = lb - ib - jb, bi[3] = {ib, jb, kb};
                        double sd[3], md[3];
                        for (int d = 0; d < 3; ++d) {
                            const int i = ai[d], j = bi[d];
                            sd[d] = o1[d][i][j];
                            md[d] = o1[d][i][j + 1] + (q.B[d] - org[d]) * o1[d][i][j];
                        }
                        sD[0] += pre * w * md[0] * sd[1] * sd[2];
                        sD[1] += pre * w * sd[0] * md[1] * sd[2];
                        sD[2] += pre * w * sd[0] * sd[1] * md[2];
                    }
            }
    }
    const int64_t i = o.fa + ma, j = o.fb + mb;
    for (int d = 0; d < 3; ++d) D[(d * N + i) * N + j] = D[(d * N + j) * N + i] = sD[d];
}

__global__ __launch_bounds__(128) void oneel_kernel(DevSet s, const int32_t* __restrict__ objs, int64_t nobj,
                                                    const double* __restrict__ sph, int64_t ncen,
                                                    const double* __restrict__ Z, const double* __restrict__ Cxyz,
                                                    int64_t N, double* __restrict__ Sm, double* __restrict__ Tm,
                                                    double* __restrict__ Vm) {
    const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (tid >= nobj * 25) return;
    const int m = (int)(tid % 25);
    const Obj o = s.obj[objs[tid / 25]];
    if (m >= o.ncomp) return;
    const int ma = m / o.nb, mb = m % o.nb, la = o.la, lb = o.lb;
    const double* Ta = sph + la * NSPH_MAX * NCART_MAX + ma * NCART_MAX;
    const double* Tb = sph + lb * NSPH_MAX * NCART_MAX + mb * NCART_MAX;
    double sS = 0.0, sT = 0.0, sV = 0.0;
    for (int pi = 0; pi < o.nprim; ++pi) {
        const Prim q = s.prim[o.prim0 + pi];
        // 1D overlaps o1[d][i][j], i <= la, j <= lb + 2 (without sqrt(pi/p) e^{-mu X^2})
        double o1[3][LMAX + 1][LMAX + 3];
        overlap_1d(q, la, lb + 2, o1);
        double ab2 = 0.0;
        for (int k = 0; k < 3; ++k) ab2 += (q.A[k] - q.B[k]) * (q.A[k] - q.B[k]);
        const double pre = q.K * exp(-q.a * q.b / q.p * ab2) * pow(M_PI / q.p, 1.5);
        int ca = 0;
        for (int ia = la; ia >= 0; --ia)
            for (int ja = la - ia; ja >= 0; --ja, ++ca) {
                const double wa = Ta[ca];
                if (wa == 0.0) continue;
                const int ka = la - ia - ja, ai[3] = {ia, ja, ka};
                int cb = 0;
                for (int ib = lb; ib >= 0; --ib)
                    for (int jb = lb - ib; jb >= 0; --jb, ++cb) {
                        const double w = wa * Tb[cb];
                        if (w == 0.0) continue;
                        const int kb = lb - ib - jb, bi[3] = {ib, jb, kb};
                        double sd[3], kd[3];
                        for (int d = 0; d < 3; ++d) {
                            const int i = ai[d], j = bi[d];
                            sd[d] = o1[d][i][j];
                            double k = 4 * q.b * q.b * o1[d][i][j + 2] - 2 * q.b * (2 * j + 1) * o1[d][i][j];
                            if (j >= 2) k += j * (j - 1) * o1[d][i][j - 2];
                            kd[d] = -0.5 * k;
                        }
                        sS += pre * w * sd[0] * sd[1] * sd[2];
                        sT += pre * w * (kd[0] * sd[1] * sd[2] + sd[0] * kd[1] * sd[2] + sd[0] * sd[1] * kd[2]);
                    }
            }
        const double* e = s.E + q.eoff + (int64_t)m * nherm(o.L);
        double v;
        switch (o.L) {
            case 0: v = nuclear_one<0>(q, e, ncen, Z, Cxyz); break;
            case 1: v = nuclear_one<1>(q, e, ncen, Z, Cxyz); break;
            case 2: v = nuclear_one<2>(q, e, ncen, Z, Cxyz); break;
            case 3: v = nuclear_one<3>(q, e, ncen, Z, Cxyz); break;
            case 4: v = nuclear_one<4>(q, e, ncen, Z, Cxyz); break;
            case 5: v = nuclear_one<5>(q, e, ncen, Z, Cxyz); break;
            case 6: v = nuclear_one<6>(q, e, ncen, Z, Cxyz); break;
            case 7: v = nuclear_one<7>(q, e, ncen, Z, Cxyz); break;
            default: v = nuclear_one<8>(q, e, ncen, Z, Cxyz); break;
        }
        sV += q.K * v;
    }
    const int64_t i = o.fa + ma, j = o.fb + mb;
    Sm[i * N + j] = Sm[j * N + i] = sS;
    Tm[i * N + j] = Tm[j * N + i] = sT;
    Vm[i * N + j] = Vm[j * N + i] = sV;
}

// ----------------------------------------------------------- host driver ----

struct DevObjSet {
    Obj* obj = nullptr;
    Prim* prim = nullptr;
    double* E = nullptr;
    int64_t nobj = 0;
    std::vector<Obj> host;   // copy for class lists
    DevSet view() const { return DevSet{obj, prim, E}; }
};

template <int LA, int LB>
int launch_etab(corr_ctx* ctx, DevObjSet& d, const std::vector<int64_t>& prims, const double* sph, const char* tag) {
    if (prims.empty()) return CORR_OK;
    int64_t* dp = nullptr;
    const std::string nm = std::string("int_etab_list_") + tag;
    CORR_TRY(scratch(ctx, nm.c_str(), sizeof(int64_t) * prims.size(), (void**)&dp));
    CORR_HIP(ctx, hipMemcpyAsync(dp, prims.data(), sizeof(int64_t) * prims.size(), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL((etab_kernel<LA, LB>), dim3((unsigned)cdiv((int64_t)prims.size(), 64)), dim3(64), 0, ctx->stream,
                       d.view(), dp, (int64_t)prims.size(), sph, d.E);
    CORR_LAUNCHED(ctx, "etab_kernel");
    // the list buffer is reused by the next class: wait before refilling
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

int upload_set(corr_ctx* ctx, const ObjSet& S, const char* tag, const double* sph, DevObjSet& d) {
    const std::string t(tag);
    CORR_TRY(scratch(ctx, ("int_obj_" + t).c_str(), sizeof(Obj) * std::max<size_t>(S.obj.size(), 1),
                     (void**)&d.obj));
    CORR_TRY(scratch(ctx, ("int_prim_" + t).c_str(), sizeof(Prim) * std::max<size_t>(S.prim.size(), 1),
                     (void**)&d.prim));
    CORR_TRY(scratch(ctx, ("int_E_" + t).c_str(), sizeof(double) * std::max<int64_t>(S.etab, 1), (void**)&d.E));
    CORR_HIP(ctx, hipMemcpyAsync(d.obj, S.obj.data(), sizeof(Obj) * S.obj.size(), hipMemcpyHostToDevice, ctx->stream));
    CORR_HIP(ctx, hipMemcpyAsync(d.prim, S.prim.data(), sizeof(Prim) * S.prim.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    d.nobj = (int64_t)S.obj.size();
    d.host = S.obj;
    // primitive lists per (la, lb)
    std::vector<int64_t> lists[LMAX + 1][LMAX + 1];
    for (const Obj& o : S.obj) {
        CORR_REQUIRE(ctx, o.la >= o.lb, "integrals: shell pair with la < lb");
        for (int k = 0; k < o.nprim; ++k) lists[o.la][o.lb].push_back(o.prim0 + k);
    }
#define CORR_ETAB(A, B) CORR_TRY((launch_etab<A, B>(ctx, d, lists[A][B], sph, tag)))
    CORR_ETAB(0, 0); CORR_ETAB(1, 0); CORR_ETAB(1, 1); CORR_ETAB(2, 0); CORR_ETAB(2, 1); CORR_ETAB(2, 2);
    CORR_ETAB(3, 0); CORR_ETAB(3, 1); CORR_ETAB(3, 2); CORR_ETAB(3, 3); CORR_ETAB(4, 0); CORR_ETAB(4, 1);
    CORR_ETAB(4, 2); CORR_ETAB(4, 3); CORR_ETAB(4, 4);
#undef CORR_ETAB
    return CORR_OK;
}

int sph_table(corr_ctx* ctx, double** out) {
    CORR_TRY(scratch(ctx, "int_sph", sizeof(SolidHarmonics::T), (void**)out));
    CORR_HIP(ctx, hipMemcpyAsync(*out, solid().T, sizeof(SolidHarmonics::T), hipMemcpyHostToDevice, ctx->stream));
    return CORR_OK;
}

template <int LX, int LY>
int launch_coul(corr_ctx* ctx, CoulArgs a) {
    const int64_t n = a.nx * a.ny * a.ncx;
    if (n == 0) return CORR_OK;
    const int64_t nb = cdiv(n, 128);
    CORR_REQUIRE(ctx, nb < ((int64_t)1 << 31), "integrals: class too large for one launch");
    hipLaunchKernelGGL((coulomb_kernel<LX, LY>), dim3((unsigned)nb), dim3(128), 0, ctx->stream, a);
    CORR_LAUNCHED(ctx, "coulomb_kernel");
    return CORR_OK;
}

typedef int (*CoulFn)(corr_ctx*, CoulArgs);
template <int LX>
CoulFn coul_row(int LY) {
    switch (LY) {
        case 0: return launch_coul<LX, 0>;
        case 1: return launch_coul<LX, 1>;
        case 2: return launch_coul<LX, 2>;
        case 3: return launch_coul<LX, 3>;
        default: return launch_coul<LX, 4>;
    }
}
CoulFn coul_fn(int LX, int LY) {
    switch (LX) {
        case 0: return coul_row<0>(LY);
        case 1: return coul_row<1>(LY);
        case 2: return coul_row<2>(LY);
        case 3: return coul_row<3>(LY);
        default: return coul_row<4>(LY);
    }
}

// Launches every (LX, LY) class of X objects against Y objects.
int run_coulomb(corr_ctx* ctx, const DevObjSet& X, const DevObjSet& Y, bool same, int mode, int64_t N, int64_t naux,
                double* out) {
    std::vector<int32_t> xl[LMAX + 1], yl[LMAX + 1];
    int ncx[LMAX + 1] = {};
    for (int64_t i = 0; i < (int64_t)X.host.size(); ++i) {
        xl[X.host[i].L].push_back((int32_t)i);
        ncx[X.host[i].L] = std::max(ncx[X.host[i].L], X.host[i].ncomp);
    }
    for (int64_t i = 0; i < (int64_t)Y.host.size(); ++i) yl[Y.host[i].L].push_back((int32_t)i);
    int32_t *dx = nullptr, *dy = nullptr;
    for (int LX = 0; LX <= LMAX; ++LX) {
        if (xl[LX].empty()) continue;
        CORR_TRY(scratch(ctx, "int_xl", sizeof(int32_t) * xl[LX].size(), (void**)&dx));
        CORR_HIP(ctx, hipMemcpyAsync(dx, xl[LX].data(), sizeof(int32_t) * xl[LX].size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        for (int LY = 0; LY <= LMAX; ++LY) {
            if (yl[LY].empty()) continue;
            if (same && LY > LX) continue;
            CORR_TRY(scratch(ctx, "int_yl", sizeof(int32_t) * yl[LY].size(), (void**)&dy));
            CORR_HIP(ctx, hipMemcpyAsync(dy, yl[LY].data(), sizeof(int32_t) * yl[LY].size(), hipMemcpyHostToDevice,
                                         ctx->stream));
            CoulArgs a{X.view(), Y.view(), dx, dy, (int64_t)xl[LX].size(), (int64_t)yl[LY].size(), ncx[LX],
                       (same && LX == LY) ? 1 : 0, mode, N, naux, out};
            CORR_TRY(coul_fn(LX, LY)(ctx, a));
            // the class lists are refilled for the next class
            CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
    }
    return CORR_OK;
}

int check_orbital(corr_ctx* ctx, const std::vector<HShell>& sh) {
    for (const HShell& s : sh) CORR_REQUIRE(ctx, s.l <= 2, "integrals: orbital shells above d need LX+LY > 8");
    return CORR_OK;
}

}  // namespace

int int3c(corr_ctx* ctx, const corr_basis* aux, const corr_basis* obs, double* Pmn) {
    CORR_REQUIRE(ctx, Pmn, "int3c: null output");
    std::vector<HShell> ash, osh;
    int64_t na = 0, n = 0;
    CORR_TRY(load_basis(ctx, aux, ash, &na));
    CORR_TRY(load_basis(ctx, obs, osh, &n));
    CORR_TRY(check_orbital(ctx, osh));
    for (const HShell& s : ash) CORR_REQUIRE(ctx, s.l <= 4, "int3c: aux shells above g");
    double* sph = nullptr;
    CORR_TRY(sph_table(ctx, &sph));
    ObjSet X, Y;
    build_singles(ash, X);
    build_pairs(osh, Y, 1e-15);
    DevObjSet dX, dY;
    CORR_TRY(upload_set(ctx, X, "x", sph, dX));
    CORR_TRY(upload_set(ctx, Y, "y", sph, dY));
    // screened pairs are zero
    CORR_HIP(ctx, hipMemsetAsync(Pmn, 0, sizeof(double) * na * n * n, ctx->stream));
    return run_coulomb(ctx, dX, dY, false, MODE_3C, n, na, Pmn);
}

int int2c(corr_ctx* ctx, const corr_basis* aux, double* J) {
    CORR_REQUIRE(ctx, J, "int2c: null output");
    std::vector<HShell> ash;
    int64_t na = 0;
    CORR_TRY(load_basis(ctx, aux, ash, &na));
    double* sph = nullptr;
    CORR_TRY(sph_table(ctx, &sph));
    ObjSet X;
    build_singles(ash, X);
    DevObjSet dX;
    CORR_TRY(upload_set(ctx, X, "x", sph, dX));
    return run_coulomb(ctx, dX, dX, true, MODE_2C, 0, na, J);
}

int int4c(corr_ctx* ctx, const corr_basis* obs, double* dense) {
    CORR_REQUIRE(ctx, dense, "int4c: null output");
    std::vector<HShell> osh;
    int64_t n = 0;
    CORR_TRY(load_basis(ctx, obs, osh, &n));
    CORR_TRY(check_orbital(ctx, osh));
    double* sph = nullptr;
    CORR_TRY(sph_table(ctx, &sph));
    ObjSet X;
    build_pairs(osh, X, 1e-15);
    DevObjSet dX;
    CORR_TRY(upload_set(ctx, X, "x", sph, dX));
    CORR_HIP(ctx, hipMemsetAsync(dense, 0, sizeof(double) * n * n * n * n, ctx->stream));
    return run_coulomb(ctx, dX, dX, true, MODE_4C, n, 0, dense);
}

int int1e(corr_ctx* ctx, const corr_basis* obs, int64_t natom, const double* charges, const double* coords, double* S,
          double* T, double* V) {
    CORR_REQUIRE(ctx, S && T && V && natom >= 0 && (natom == 0 || (charges && coords)), "int1e: bad arguments");
    std::vector<HShell> osh;
    int64_t n = 0;
    CORR_TRY(load_basis(ctx, obs, osh, &n));
    CORR_TRY(check_orbital(ctx, osh));
    double* sph = nullptr;
    CORR_TRY(sph_table(ctx, &sph));
    ObjSet X;
    build_pairs(osh, X, 0.0);
    DevObjSet dX;
    CORR_TRY(upload_set(ctx, X, "x", sph, dX));
    double* cz = nullptr;
    CORR_TRY(scratch(ctx, "int_nuclei", sizeof(double) * 4 * std::max<int64_t>(natom, 1), (void**)&cz));
    if (natom > 0) {
        CORR_HIP(ctx, hipMemcpyAsync(cz, charges, sizeof(double) * natom, hipMemcpyHostToDevice, ctx->stream));
        CORR_HIP(ctx, hipMemcpyAsync(cz + natom, coords, sizeof(double) * 3 * natom, hipMemcpyHostToDevice,
                                     ctx->stream));
    }
    std::vector<int32_t> all(X.obj.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int32_t)i;
    int32_t* dl = nullptr;
    CORR_TRY(scratch(ctx, "int_xl", sizeof(int32_t) * all.size(), (void**)&dl));
    CORR_HIP(ctx, hipMemcpyAsync(dl, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, ctx->stream));
    const int64_t nt = (int64_t)all.size() * 25;
    hipLaunchKernelGGL(oneel_kernel, dim3((unsigned)cdiv(nt, 128)), dim3(128), 0, ctx->stream, dX.view(), dl,
                       (int64_t)all.size(), sph, natom, cz, cz + natom, n, S, T, V);
    CORR_LAUNCHED(ctx, "oneel_kernel");
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

int int1e_dipole(corr_ctx* ctx, const corr_basis* obs, const double* origin, double* D) {
    CORR_REQUIRE(ctx, D && origin, "int1e_dipole: bad arguments");
    std::vector<HShell> osh;
    int64_t n = 0;
    CORR_TRY(load_basis(ctx, obs, osh, &n));
    CORR_TRY(check_orbital(ctx, osh));
    double* sph = nullptr;
    CORR_TRY(sph_table(ctx, &sph));
    ObjSet X;
    build_pairs(osh, X, 0.0);
    DevObjSet dX;
    CORR_TRY(upload_set(ctx, X, "x", sph, dX));
    std::vector<int32_t> all(X.obj.size());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int32_t)i;
    int32_t* dl = nullptr;
    CORR_TRY(scratch(ctx, "int_xl", sizeof(int32_t) * all.size(), (void**)&dl));
    CORR_HIP(ctx, hipMemcpyAsync(dl, all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice, ctx->stream));
    const int64_t nt = (int64_t)all.size() * 25;
    hipLaunchKernelGGL(dipole_kernel, dim3((unsigned)cdiv(nt, 128)), dim3(128), 0, ctx->stream, dX.view(), dl,
                       (int64_t)all.size(), sph, origin[0], origin[1], origin[2], n, D);
    CORR_LAUNCHED(ctx, "dipole_kernel");
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

}  // namespace corr

extern "C" {

int corr_int1e(corr_ctx* ctx, const corr_basis* obs, int64_t natom, const double* h_charges, const double* h_coords,
               double* d_S, double* d_T, double* d_V) {
    CORR_CHECK_CTX(ctx);
    return corr::int1e(ctx, obs, natom, h_charges, h_coords, d_S, d_T, d_V);
}

int corr_int2c(corr_ctx* ctx, const corr_basis* aux, double* d_J) {
    CORR_CHECK_CTX(ctx);
    return corr::int2c(ctx, aux, d_J);
}

int corr_int3c(corr_ctx* ctx, const corr_basis* aux, const corr_basis* obs, double* d_Pmn) {
    CORR_CHECK_CTX(ctx);
    return corr::int3c(ctx, aux, obs, d_Pmn);
}

int corr_int4c(corr_ctx* ctx, const corr_basis* obs, double* d_dense) {
    CORR_CHECK_CTX(ctx);
    return corr::int4c(ctx, obs, d_dense);
}

int corr_int1e_dipole(corr_ctx* ctx, const corr_basis* obs, const double* h_origin, double* d_D) {
    CORR_CHECK_CTX(ctx);
    return corr::int1e_dipole(ctx, obs, h_origin, d_D);
}

}  // extern "C"
