// Spin-orbital CCSD (Stanton & Gauss) on CDNA4: ccsd_ur and its helpers
// (src/POSTHF/ccsd.cc:6-580, src/POSTHF/ccsd_intermediates.cc:6-287).
//
// Every contraction of an iteration is one MFMA GEMM (FP64 or FP32) on
// contiguous blocks of <pq||rs>; index permutations that a GEMM cannot absorb
// through its leading dimensions and batch strides are done by one generic
// 4-D gather-accumulate kernel.  Amplitude updates, denominators, energies and
// DIIS stay FP64.  Spin orbital p is spatial p/2 with spin p%2; the nocc
// lowest spin orbitals are occupied (ccsd.cc:540-543).
#include "common.hpp"
#include <cmath>
#include <vector>

namespace corr {

namespace {

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 * 4 ? 65536 * 4 : (g < 1 ? 1 : g));
}

#define GRID_STRIDE(e, n) \
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (n); e += (int64_t)gridDim.x * blockDim.x)

// <pq||rs> = (pr|qs) d(sp,sr) d(sq,ss) - (ps|qr) d(sp,ss) d(sq,sr)   (ccsd.cc:6-9)
__global__ void ints_so_kernel(const double* __restrict__ I, int64_t N, double* __restrict__ out) {
    const int64_t n = 2 * N, n4 = n * n * n * n;
    GRID_STRIDE(e, n4) {
        const int64_t s = e % n, r = (e / n) % n, q = (e / (n * n)) % n, p = e / (n * n * n);
        const int64_t P = p >> 1, Q = q >> 1, R = r >> 1, S = s >> 1;
        double x = 0.0;
        if ((p & 1) == (r & 1) && (q & 1) == (s & 1)) x += I[((P * N + R) * N + Q) * N + S];
        if ((p & 1) == (s & 1) && (q & 1) == (r & 1)) x -= I[((P * N + S) * N + Q) * N + R];
        out[e] = x;
    }
}

struct Perm4 {
    int64_t d[4];   // destination dims (contiguous, d[3] fastest)
    int64_t s[4];   // source strides for the 4 destination indices
};

// dst[i0][i1][i2][i3] = beta dst + alpha src[i0 s0 + i1 s1 + i2 s2 + i3 s3]
// Grid (ceil(d2 d3 / 256), d1, d0): i0 and i1 come from the block, (i2, i3)
// from one 32-bit division (d2 d3 < 2^31 is checked by the launcher).
__global__ void permute4_kernel(double* __restrict__ dst, const double* __restrict__ src, Perm4 p,
                                double alpha, double beta) {
    const uint32_t d3 = (uint32_t)p.d[3], n23 = (uint32_t)(p.d[2] * p.d[3]);
    const uint32_t e23 = blockIdx.x * blockDim.x + threadIdx.x;
    if (e23 >= n23) return;
    const int64_t i0 = blockIdx.z, i1 = blockIdx.y;
    const uint32_t i2 = e23 / d3, i3 = e23 - i2 * d3;
    const double x = alpha * src[i0 * p.s[0] + i1 * p.s[1] + (int64_t)i2 * p.s[2] + (int64_t)i3 * p.s[3]];
    const int64_t e = (i0 * p.d[1] + i1) * (int64_t)n23 + e23;
    dst[e] = (beta == 0.0) ? x : beta * dst[e] + x;
}

// The same permutation when the source is contiguous along destination index
// k < 3 (s_k = 1) and not along i3: 32 x 32 tiles of (i_k, i3) are read along
// i_k into LDS and written along i3, so both sides are coalesced.  Grid
// (tiles of (i_k, i3), d_b, d_a) with a < b the two remaining indices.
__global__ __launch_bounds__(256) void permute4_tiled_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                             Perm4 p, int k, double alpha, double beta) {
    __shared__ double t[32][33];
    const int64_t dk = p.d[k], n3 = (p.d[3] + 31) / 32;
    const int64_t tk = blockIdx.x / n3, t3 = blockIdx.x % n3;
    const int64_t ia = blockIdx.z, ib = blockIdx.y;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    auto index = [&](int64_t ik, int64_t& i0, int64_t& i1, int64_t& i2) {
        i0 = k == 0 ? ik : ia;
        i1 = k == 0 ? ia : (k == 1 ? ik : ib);
        i2 = k == 2 ? ik : ib;
    };
    for (int j = ty; j < 32; j += 8) {
        const int64_t ik = tk * 32 + tx, i3 = t3 * 32 + j;
        if (ik < dk && i3 < p.d[3]) {
            int64_t i0, i1, i2;
            index(ik, i0, i1, i2);
            t[j][tx] = src[i0 * p.s[0] + i1 * p.s[1] + i2 * p.s[2] + i3 * p.s[3]];
        }
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t ik = tk * 32 + j, i3 = t3 * 32 + tx;
        if (ik < dk && i3 < p.d[3]) {
            int64_t i0, i1, i2;
            index(ik, i0, i1, i2);
            const int64_t e = ((i0 * p.d[1] + i1) * p.d[2] + i2) * p.d[3] + i3;
            const double x = alpha * t[tx][j];
            dst[e] = (beta == 0.0) ? x : beta * dst[e] + x;
        }
    }
}

// Antisymmetric pair packing of the particle-particle ladder.  Pair p of n
// orbitals is (x_p < y_p) (index arrays px, py); rows p of the 2-D grid
// (blockIdx.y, looping past 65535) and columns q (blockIdx.x, threads).
#define PAIR_LOOP(P, Q)                                                                   \
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;                     \
    if (q >= (Q)) return;                                                                 \
    for (int64_t p = blockIdx.y; p < (P); p += gridDim.y)

// VP[p][q] = <ab||ef>, (a<b) = pair p and (e<f) = pair q of the virtuals, from
// the spin-orbital tensor V (n^4, [p][q][r][s]).
__global__ void pack_vvvv_kernel(const double* __restrict__ V, int64_t n, int64_t o, const int32_t* __restrict__ pa,
                                 const int32_t* __restrict__ pb, int64_t Pv, double* __restrict__ out) {
    PAIR_LOOP(Pv, Pv) {
        const int64_t a = o + pa[p], b = o + pb[p], e = o + pa[q], f = o + pb[q];
        out[p * Pv + q] = V[((a * n + b) * n + e) * n + f];
    }
}

// out[a][m][q] = <am||ef>, (e<f) = pair q; row index a*o + m.
__global__ void pack_vovv_kernel(const double* __restrict__ V, int64_t n, int64_t o, int64_t v,
                                 const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, int64_t Pv,
                                 double* __restrict__ out) {
    PAIR_LOOP(v * o, Pv) {
        const int64_t a = o + p / o, m = p % o, e = o + pa[q], f = o + pb[q];
        out[p * Pv + q] = V[((a * n + m) * n + e) * n + f];
    }
}

// tauP[p][q] = tau[i][j][e][f], (i<j) = pair p of the occupied, (e<f) = pair q.
__global__ void pack_tau_kernel(const double* __restrict__ tau, int64_t o, int64_t v, const int32_t* __restrict__ pi,
                                const int32_t* __restrict__ pj, const int32_t* __restrict__ pa,
                                const int32_t* __restrict__ pb, int64_t Po, int64_t Pv, double* __restrict__ out) {
    PAIR_LOOP(Po, Pv) {
        const int64_t i = pi[p], j = pj[p], e = pa[q], f = pb[q];
        out[p * Pv + q] = tau[((i * o + j) * v + e) * v + f];
    }
}

// WP[p][q] = W[m][n][i][j], (m<n) = pair p and (i<j) = pair q of the occupied.
__global__ void pack_oooo_kernel(const double* __restrict__ W, int64_t o, const int32_t* __restrict__ pi,
                                 const int32_t* __restrict__ pj, int64_t Po, double* __restrict__ out) {
    PAIR_LOOP(Po, Po) {
        const int64_t m = pi[p], n = pj[p], i = pi[q], j = pj[q];
        out[p * Po + q] = W[((m * o + n) * o + i) * o + j];
    }
}

// WP[p][q] = VP[p][q] - XP[a][b][q] + XP[b][a][q], (a<b) = pair p:
// <ab||ef> - P(ab) t_mb <am||ef> with XP[a][b][q] = sum_m t_mb <am||ef>.
__global__ void wabef_packed_kernel(const double* __restrict__ VP, const double* __restrict__ XP, int64_t v,
                                    const int32_t* __restrict__ pa, const int32_t* __restrict__ pb, int64_t Pv,
                                    double* __restrict__ WP) {
    PAIR_LOOP(Pv, Pv) {
        const int64_t a = pa[p], b = pb[p];
        WP[p * Pv + q] = VP[p * Pv + q] - XP[(a * v + b) * Pv + q] + XP[(b * v + a) * Pv + q];
    }
}

// T2n[i][j][a][b] += Y[p][q] with the antisymmetry of T2 in (ij) and (ab):
// (i<j) = pair p, (a<b) = pair q; each T2 element has one source.
__global__ void scatter_pp_kernel(const double* __restrict__ Y, int64_t o, int64_t v, const int32_t* __restrict__ pi,
                                  const int32_t* __restrict__ pj, const int32_t* __restrict__ pa,
                                  const int32_t* __restrict__ pb, int64_t Po, int64_t Pv, double* __restrict__ T2n) {
    PAIR_LOOP(Po, Pv) {
        const int64_t i = pi[p], j = pj[p], a = pa[q], b = pb[q];
        const double y = Y[p * Pv + q];
        T2n[((i * o + j) * v + a) * v + b] += y;
        T2n[((j * o + i) * v + a) * v + b] -= y;
        T2n[((i * o + j) * v + b) * v + a] -= y;
        T2n[((j * o + i) * v + b) * v + a] += y;
    }
}
#undef PAIR_LOOP

// Spin-orbital h from hs = C H C^T and f = h + sum_{j<nocc} <pj||qj>  (ccsd.cc:11-48)
__global__ void form_fock_kernel(const double* __restrict__ hs, const double* __restrict__ V, int64_t N,
                                 int64_t nocc, double* __restrict__ h, double* __restrict__ f) {
    const int64_t n = 2 * N;
    GRID_STRIDE(e, n * n) {
        const int64_t p = e / n, q = e % n;
        const double hv = ((p & 1) == (q & 1)) ? hs[(p >> 1) * N + (q >> 1)] : 0.0;
        double t = 0.0;
        for (int64_t j = 0; j < nocc; ++j) t += V[((p * n + j) * n + q) * n + j];
        h[e] = hv;
        f[e] = hv + t;
    }
}

// T1n /= f_ii - f_aa ; T2n /= f_ii + f_jj - f_aa - f_bb
__global__ void denom_kernel(double* __restrict__ T1, double* __restrict__ T2, const double* __restrict__ f,
                             int64_t o, int64_t v) {
    const int64_t n = o + v;
    GRID_STRIDE(e, o * v + o * o * v * v) {
        if (e < o * v) {
            const int64_t i = e / v, a = e % v;
            if (T1) T1[e] /= (f[i * n + i] - f[(o + a) * n + o + a]);
        } else {
            const int64_t x = e - o * v;
            const int64_t b = x % v, a = (x / v) % v, j = (x / (v * v)) % o, i = x / (v * v * o);
            T2[x] /= (f[i * n + i] + f[j * n + j] - f[(o + a) * n + o + a] - f[(o + b) * n + o + b]);
        }
    }
}

// tau~ = t2 + 1/2 (t_ia t_jb - t_ib t_ja), tau = t2 + t_ia t_jb - t_ib t_ja  (ccsd.cc:221-251)
__global__ void taus_kernel(const double* __restrict__ T1, const double* __restrict__ T2, int64_t o,
                            int64_t v, double* __restrict__ tau_s, double* __restrict__ tau) {
    GRID_STRIDE(e, o * o * v * v) {
        const int64_t b = e % v, a = (e / v) % v, j = (e / (v * v)) % o, i = e / (v * v * o);
        const double x = T1[i * v + a] * T1[j * v + b] - T1[i * v + b] * T1[j * v + a];
        tau_s[e] = T2[e] + 0.5 * x;
        tau[e] = T2[e] + x;
    }
}

// Fae = (1-d_ae) f_ae - 1/2 sum_m f_{m a} t_m^a (the f index of
// ccsd_intermediates.cc:33); Fmi = (1-d_mi) f_mi
__global__ void f_init_kernel(const double* __restrict__ f, const double* __restrict__ T1, int64_t o,
                              int64_t v, double* __restrict__ Fae, double* __restrict__ Fmi) {
    const int64_t n = o + v;
    GRID_STRIDE(e, v * v + o * o) {
        if (e < v * v) {
            const int64_t a = e / v, ee = e % v;
            double x = (a == ee) ? 0.0 : f[(o + a) * n + o + ee];
            double q = 0.0;
            for (int64_t m = 0; m < o; ++m) q += f[m * n + o + a] * T1[m * v + a];
            Fae[e] = x - 0.5 * q;
        } else {
            const int64_t x = e - v * v, m = x / o, i = x % o;
            Fmi[x] = (m == i) ? 0.0 : f[m * n + i];
        }
    }
}

// Zp[n][f][j][b] = 1/2 t2[j][n][f][b] + t_jf t_nb   (Wmbej, ccsd_intermediates.cc:278-283)
__global__ void zp_kernel(const double* __restrict__ T1, const double* __restrict__ T2, int64_t o,
                          int64_t v, double* __restrict__ Zp) {
    GRID_STRIDE(e, o * o * v * v) {
        const int64_t b = e % v, j = (e / v) % o, f = (e / (v * o)) % v, nn = e / (v * o * v);
        Zp[e] = 0.5 * T2[((j * o + nn) * v + f) * v + b] + T1[j * v + f] * T1[nn * v + b];
    }
}

// ---- spin-sorted compound indices ---------------------------------------------
// Spin orbitals alternate alpha / beta (spin = index % 2, as in ints_so_kernel).
// An (occupied p, virtual x) pair falls in class 0 (alpha alpha), 1 (beta beta),
// 2 (alpha beta) or 3 (beta alpha); sorted index = class * (o/2)(v/2) +
// (p/2)(v/2) + x/2.  Pair matrices of the ring terms (rows and columns such
// pairs) are block-sparse in this order: amplitude-like ones (t_imae, <mn||ef>,
// Z_nfjb) vanish outside blocks (0|1, 0|1), (2, 3), (3, 2); W_mbej-like ones
// outside (0|1, 0|1), (2, 2), (3, 3).
struct SpinDims {
    int64_t o, v;   // even
};
__device__ inline int64_t spin_sorted(int64_t p, int64_t x, int64_t vh, int64_t B4) {
    const int64_t sp = p & 1, sx = x & 1;
    const int64_t cls = sp == sx ? sp : 2 + sp;
    return cls * B4 + (p >> 1) * vh + (x >> 1);
}
__device__ inline void spin_unsorted(int64_t r, int64_t vh, int64_t B4, int64_t& p, int64_t& x) {
    const int64_t cls = r / B4, loc = r % B4;
    const int64_t sp = cls < 2 ? cls : cls - 2, sx = cls < 2 ? cls : 3 - cls;
    p = 2 * (loc / vh) + sp;
    x = 2 * (loc % vh) + sx;
}

// dst[r'][c'] (sorted, ov x ov) = src[p s0 + x s1 + q s2 + y s3] for r' = sorted(p, x),
// c' = sorted(q, y).
__global__ void spin_gather_kernel(double* __restrict__ dst, const double* __restrict__ src, SpinDims d,
                                   int64_t s0, int64_t s1, int64_t s2, int64_t s3) {
    const int64_t ov = d.o * d.v, vh = d.v / 2, B4 = ov / 4;
    GRID_STRIDE(e, ov * ov) {
        int64_t p, x, q, y;
        spin_unsorted(e / ov, vh, B4, p, x);
        spin_unsorted(e % ov, vh, B4, q, y);
        dst[e] = src[p * s0 + x * s1 + q * s2 + y * s3];
    }
}

// dst (dims n0..n3, contiguous) = beta dst + alpha S[sorted(p, x)][sorted(q, y)], where
// destination index k is the role[k]-th of (p, x, q, y).
struct SpinScatter {
    int64_t n[4];
    int role[4];
};
__global__ void spin_scatter_kernel(double* __restrict__ dst, const double* __restrict__ S, SpinDims d, SpinScatter m,
                                    double alpha, double beta) {
    const int64_t ov = d.o * d.v, vh = d.v / 2, B4 = ov / 4;
    const int64_t total = m.n[0] * m.n[1] * m.n[2] * m.n[3];
    GRID_STRIDE(e, total) {
        int64_t idx[4], rem = e;
        for (int k = 3; k >= 0; --k) {
            idx[k] = rem % m.n[k];
            rem /= m.n[k];
        }
        int64_t pxqy[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) pxqy[m.role[k]] = idx[k];
        const double val = S[spin_sorted(pxqy[0], pxqy[1], vh, B4) * ov + spin_sorted(pxqy[2], pxqy[3], vh, B4)];
        dst[e] = beta == 0.0 ? alpha * val : beta * dst[e] + alpha * val;
    }
}

// Largest |amplitude| that spin forbids: T1[i][a] with spin(i) != spin(a), T2[i][j][a][b]
// with spin(i) + spin(j) != spin(a) + spin(b).  Block (i, j) scans T2[i][j] (and
// block 0 also T1); one maximum per block (order-free).
__global__ __launch_bounds__(256) void spin_forbidden_kernel(const double* __restrict__ T1,
                                                             const double* __restrict__ T2, int o, int v,
                                                             double* __restrict__ partial) {
    __shared__ double sh[256];
    const int ij = blockIdx.x, i = ij / o, j = ij % o, sij = (i & 1) + (j & 1);
    const double* t = T2 + (int64_t)ij * v * v;
    double m = 0.0;
    for (int e = threadIdx.x; e < v * v; e += 256) {
        const int a = e / v, b = e - a * v;
        if (((a & 1) + (b & 1)) != sij) m = fmax(m, fabs(t[e]));
    }
    if (ij == 0)
        for (int e = threadIdx.x; e < o * v; e += 256) {
            const int ii = e / v, a = e - ii * v;
            if ((ii & 1) != (a & 1)) m = fmax(m, fabs(T1[e]));
        }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[ij] = sh[0];
}

__global__ void dot_kernel(const double* __restrict__ x, const double* __restrict__ y, int64_t n,
                           double* __restrict__ partial) {
    __shared__ double sh[256];
    double acc = 0.0;
    GRID_STRIDE(e, n) acc += x[e] * y[e];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// calc_e_check (ccsd.cc:50-81): partial sums of [sum_i h_ii, 1/2 twoe, sum_i f_ii]
// with twoe = sum_ij [(ii|jj) - (ij|ij) d(si,sj)] over occupied spin orbitals.
__global__ void ehf_kernel(const double* __restrict__ I, const double* __restrict__ h,
                           const double* __restrict__ f, int64_t N, int64_t nocc, double* __restrict__ partial) {
    __shared__ double sh[3][256];
    const int64_t n = 2 * N;
    double acc[3] = {0.0, 0.0, 0.0};
    GRID_STRIDE(e, nocc * nocc) {
        const int64_t i = e / nocc, j = e % nocc, si = i >> 1, sj = j >> 1;
        double t = I[((si * N + si) * N + sj) * N + sj];
        if ((i & 1) == (j & 1)) t -= I[((si * N + sj) * N + si) * N + sj];
        acc[1] += 0.5 * t;
        if (j == 0) {
            acc[0] += h[i * n + i];
            acc[2] += f[i * n + i];
        }
    }
    for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[blockIdx.x * 3 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n, double a,
                             double b) {
    GRID_STRIDE(e, n) y[e] = (b == 0.0 ? 0.0 : b * y[e]) + a * x[e];
}

__global__ void sub_kernel(double* __restrict__ r, const double* __restrict__ x, const double* __restrict__ y,
                           int64_t n) {
    GRID_STRIDE(e, n) r[e] = x[e] - y[e];
}

}  // namespace

// ------------------------------------------------------------------------------
struct CCSD {
    corr_ctx* ctx;
    int64_t o, v, n;
    int prec;
    // constant blocks of <pq||rs>
    double *Voooo, *Vooov, *Voovo, *Voovv, *Vovoo, *Vovvo, *Vovvv, *Vvovv;
    double* VaeK;   // VaeK[(a e)][(m f)] = <am||fe>: the Fae term as one matrix-vector product
    // packed over antisymmetric pairs: <ab||ef> [(a<b)][(e<f)], <am||ef> [a][m][(e<f)]
    double *VvvvvP, *VvovvP;
    int64_t Po = 0, Pv = 0;            // o(o-1)/2, v(v-1)/2
    int64_t po_cls[4] = {0, 0, 0, 0}, pv_cls[4] = {0, 0, 0, 0};   // class starts of the pair lists
    int32_t *po_i = nullptr, *po_j = nullptr, *pv_a = nullptr, *pv_b = nullptr;
    double *PeOOf, *Pmenf, *Pmine, *Pianf, *Pmbje, *Pabje;
    double *fo;   // f (n x n)
    // per-iteration work
    double *tau_s, *tau, *tauS_a, *Fae, *Faep, *Fmi, *Fmip, *Fme, *Yk, *Wmnij, *WabefP, *XabP, *tauP, *Ypp, *Wmbej;
    double *VoovvP, *WmnijP;   // <mn||ef> over (m<n) x (e<f); Wmnij over (m<n) x (i<j)
    double *Zp, *W2, *Wp, *T2p, *T2a, *X1, *X2, *Zr, *U, *Q, *Tmp1;
    double* partial;
    // Spin-blocked ring terms (V from ints_so_kernel: spin = index % 2).  Exact
    // only while the amplitudes have no spin-forbidden part: the blocks that
    // spin forbids are skipped.  The caller checks the new amplitudes after
    // each update (spin_forbidden_max) and leaves this path for good once a
    // forbidden amplitude exceeds SPIN_FORBIDDEN_TOL: for an RHF reference whose
    // spin-flip sector is unstable, rounding noise there grows as in the dense
    // spin-orbital iteration, which then runs on to its spin-broken root.
    bool spin = false;
    static constexpr double SPIN_FORBIDDEN_TOL = 1e-12;
    int spin_forbidden_max(const double* T1, const double* T2, double* out) {
        const int64_t nb = o * o;
        hipLaunchKernelGGL(spin_forbidden_kernel, dim3((unsigned)nb), dim3(256), 0, ctx->stream, T1, T2, (int)o, (int)v,
                           spinmax);
        CORR_LAUNCHED(ctx, "spin_forbidden_kernel");
        std::vector<double> h((size_t)nb);
        CORR_HIP(ctx, hipMemcpyAsync(h.data(), spinmax, sizeof(double) * nb, hipMemcpyDeviceToHost, ctx->stream));
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        double m = 0.0;
        for (double x : h) m = (x > m || x != x) ? x : m;   // a NaN is kept
        *out = m;
        return CORR_OK;
    }
    double *PmenfS = nullptr, *ZpS = nullptr, *W2S = nullptr, *T2S = nullptr, *WpS = nullptr, *ZrS = nullptr;
    double* spinmax = nullptr;   // o^2 block maxima of spin_forbidden_kernel

    int spin_gather(double* dst, const double* src, int64_t s0, int64_t s1, int64_t s2, int64_t s3) {
        const int64_t ov = o * v;
        hipLaunchKernelGGL(spin_gather_kernel, dim3(grid_for(ov * ov)), dim3(256), 0, ctx->stream, dst, src,
                           SpinDims{o, v}, s0, s1, s2, s3);
        CORR_LAUNCHED(ctx, "spin_gather_kernel");
        return CORR_OK;
    }
    int spin_scatter(double* dst, const double* S, std::initializer_list<int64_t> dims,
                     std::initializer_list<int> roles, double alpha, double beta) {
        SpinScatter m;
        int k = 0;
        for (auto d : dims) m.n[k++] = d;
        k = 0;
        for (auto r : roles) m.role[k++] = r;
        hipLaunchKernelGGL(spin_scatter_kernel, dim3(grid_for(m.n[0] * m.n[1] * m.n[2] * m.n[3])), dim3(256), 0,
                           ctx->stream, dst, S, SpinDims{o, v}, m, alpha, beta);
        CORR_LAUNCHED(ctx, "spin_scatter_kernel");
        return CORR_OK;
    }
    // C = A B on sorted pair matrices: the (0|1 x 0|1) super-block in one GEMM,
    // then the two cross blocks: A (ra, ca) B (ca, cb) -> C (ra, cb) for
    // (ra, ca, cb) = x[0..2] and x[3..5].
    int spin_ring(const double* A, const double* B, double* C, const int (&x)[6]) {
        const int64_t ov = o * v, B4 = ov / 4;
        CORR_TRY(mm(0, 0, 2 * B4, 2 * B4, 2 * B4, 1.0, A, ov, 0, B, ov, 0, 0.0, C, ov, 0));
        for (int k = 0; k < 2; ++k) {
            const int64_t ra = x[3 * k] * B4, ca = x[3 * k + 1] * B4, cb = x[3 * k + 2] * B4;
            CORR_TRY(mm(0, 0, B4, B4, B4, 1.0, A + ra * ov + ca, ov, 0, B + ca * ov + cb, ov, 0, 0.0, C + ra * ov + cb,
                        ov, 0));
        }
        return CORR_OK;
    }

    int perm(double* dst, const double* src, std::initializer_list<int64_t> dims,
             std::initializer_list<int64_t> strides, double alpha = 1.0, double beta = 0.0) {
        Perm4 p;
        int k = 0;
        for (auto d : dims) p.d[k++] = d;
        k = 0;
        for (auto s : strides) p.s[k++] = s;
        const int64_t n23 = p.d[2] * p.d[3];
        CORR_REQUIRE(ctx, n23 < ((int64_t)1 << 31) && p.d[1] <= 65535 && p.d[0] <= 65535, "ccsd: permute too large");
        if (n23 == 0 || p.d[0] == 0 || p.d[1] == 0) return CORR_OK;
        // source contiguous along i_k (k < 3) but not along i3: tiled transpose
        int kc = -1;
        for (int c = 0; c < 3; ++c)
            if (p.s[c] == 1 && p.d[c] >= 8) kc = c;
        if (p.s[3] != 1 && kc >= 0 && p.d[3] >= 8) {
            const int64_t da = kc == 0 ? p.d[1] : p.d[0], db = kc == 2 ? p.d[1] : p.d[2];
            const int64_t tiles = cdiv(p.d[kc], 32) * cdiv(p.d[3], 32);
            CORR_REQUIRE(ctx, da <= 65535 && db <= 65535 && tiles < ((int64_t)1 << 31), "ccsd: permute too large");
            hipLaunchKernelGGL(permute4_tiled_kernel, dim3((unsigned)tiles, (unsigned)db, (unsigned)da), dim3(256), 0,
                               ctx->stream, dst, src, p, kc, alpha, beta);
            CORR_LAUNCHED(ctx, "permute4_tiled_kernel");
            return CORR_OK;
        }
        hipLaunchKernelGGL(permute4_kernel, dim3((unsigned)cdiv(n23, 256), (unsigned)p.d[1], (unsigned)p.d[0]),
                           dim3(256), 0, ctx->stream, dst, src, p, alpha, beta);
        CORR_LAUNCHED(ctx, "permute4_kernel");
        return CORR_OK;
    }
    int mm(int ta, int tb, int64_t M, int64_t N_, int64_t K, double al, const double* A, int64_t lda,
           int64_t sA, const double* B, int64_t ldb, int64_t sB, double be, double* C, int64_t ldc,
           int64_t sC, int64_t batch = 1) {
        return gemm(ctx, ta, tb, M, N_, K, al, A, lda, sA, B, ldb, sB, be, C, ldc, sC, batch, prec);
    }
    int dot(const double* x, const double* y, int64_t len, double* out) {
        const int nb = 512;
        hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(256), 0, ctx->stream, x, y, len, partial);
        CORR_LAUNCHED(ctx, "dot_kernel");
        return reduce_sum(ctx, partial, nb, 1, out);
    }

    int alloc(double** p, const char* name, int64_t count) {
        return scratch(ctx, name, sizeof(double) * (size_t)count, (void**)p);
    }

    // Extracts the constant blocks from ints_so (n^4, [p][q][r][s]).
    int setup(const double* V) {
        const int64_t s0 = n * n * n, s1 = n * n, s2 = n;
        const int64_t O4 = o * o * o * o, O3V = o * o * o * v, O2V2 = o * o * v * v, OV3 = o * v * v * v;
        CORR_TRY(alloc(&Voooo, "cc_Voooo", O4));
        CORR_TRY(alloc(&Vooov, "cc_Vooov", O3V));
        CORR_TRY(alloc(&Voovo, "cc_Voovo", O3V));
        CORR_TRY(alloc(&Voovv, "cc_Voovv", O2V2));
        CORR_TRY(alloc(&Vovoo, "cc_Vovoo", O3V));
        CORR_TRY(alloc(&Vovvo, "cc_Vovvo", O2V2));
        CORR_TRY(alloc(&Vovvv, "cc_Vovvv", OV3));
        CORR_TRY(alloc(&Vvovv, "cc_Vvovv", OV3));
        // pair index arrays of the packed pp ladder, x < y, ordered by spin
        // class (both alpha, both beta, mixed; spin = index % 2): with
        // spin-symmetric amplitudes the ladder is block-diagonal in the classes
        Po = o * (o - 1) / 2;
        Pv = v * (v - 1) / 2;
        {
            std::vector<int32_t> h(2 * (Po + Pv) + 4);
            auto fill = [&](int64_t n, int32_t* first, int32_t* second, int64_t* bounds) {
                int64_t k = 0;
                const int cls_order[3] = {0, 2, 1};   // (x&1)+(y&1): alpha alpha, beta beta, mixed
                for (int c = 0; c < 3; ++c) {
                    bounds[c] = k;
                    for (int64_t x = 0; x < n; ++x)
                        for (int64_t y = x + 1; y < n; ++y)
                            if ((x & 1) + (y & 1) == cls_order[c]) {
                                first[k] = (int32_t)x;
                                second[k] = (int32_t)y;
                                ++k;
                            }
                }
                bounds[3] = k;
            };
            fill(o, h.data(), h.data() + Po, po_cls);
            fill(v, h.data() + 2 * Po, h.data() + 2 * Po + Pv, pv_cls);
            int32_t* d = nullptr;
            CORR_TRY(scratch(ctx, "cc_pairs", sizeof(int32_t) * h.size(), (void**)&d));
            CORR_HIP(ctx, hipMemcpyAsync(d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, ctx->stream));
            CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));   // h is read by the copy
            po_i = d;
            po_j = d + Po;
            pv_a = d + 2 * Po;
            pv_b = d + 2 * Po + Pv;
        }
        CORR_TRY(alloc(&VvvvvP, "cc_VvvvvP", std::max<int64_t>(Pv * Pv, 1)));
        CORR_TRY(alloc(&VvovvP, "cc_VvovvP", std::max<int64_t>(v * o * Pv, 1)));
        CORR_TRY(alloc(&PeOOf, "cc_PeOOf", O2V2));
        CORR_TRY(alloc(&Pmenf, "cc_Pmenf", O2V2));
        CORR_TRY(alloc(&Pmine, "cc_Pmine", O3V));
        CORR_TRY(alloc(&Pianf, "cc_Pianf", O2V2));
        CORR_TRY(alloc(&Pmbje, "cc_Pmbje", O2V2));
        CORR_TRY(alloc(&Pabje, "cc_Pabje", OV3));
        const int64_t ob = o;   // offset of the virtual block
        auto at = [&](int64_t p, int64_t q, int64_t r, int64_t s) { return V + p * s0 + q * s1 + r * s2 + s; };
        CORR_TRY(perm(Voooo, at(0, 0, 0, 0), {o, o, o, o}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Vooov, at(0, 0, 0, ob), {o, o, o, v}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Voovo, at(0, 0, ob, 0), {o, o, v, o}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Voovv, at(0, 0, ob, ob), {o, o, v, v}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Vovoo, at(0, ob, 0, 0), {o, v, o, o}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Vovvo, at(0, ob, ob, 0), {o, v, v, o}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Vovvv, at(0, ob, ob, ob), {o, v, v, v}, {s0, s1, s2, 1}));
        CORR_TRY(perm(Vvovv, at(ob, 0, ob, ob), {v, o, v, v}, {s0, s1, s2, 1}));
        CORR_TRY(alloc(&VaeK, "cc_VaeK", OV3));
        CORR_TRY(perm(VaeK, Vvovv, {v, v, o, v}, {o * v * v, 1, v * v, v}));
        if (Pv > 0) {
            hipLaunchKernelGGL(pack_vvvv_kernel, dim3((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(Pv, 65535)),
                               dim3(256), 0, ctx->stream, V, n, o, pv_a, pv_b, Pv, VvvvvP);
            CORR_LAUNCHED(ctx, "pack_vvvv_kernel");
            hipLaunchKernelGGL(pack_vovv_kernel,
                               dim3((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(v * o, 65535)), dim3(256), 0,
                               ctx->stream, V, n, o, v, pv_a, pv_b, Pv, VvovvP);
            CORR_LAUNCHED(ctx, "pack_vovv_kernel");
        }
        // PeOOf[e][m][n][f] = <mn||ef>
        CORR_TRY(perm(PeOOf, at(0, 0, ob, ob), {v, o, o, v}, {s2, s0, s1, 1}));
        // Pmenf[m][e][n][f] = <mn||ef>
        CORR_TRY(perm(Pmenf, at(0, 0, ob, ob), {o, v, o, v}, {s0, s2, s1, 1}));
        // Pmine[m][i][n][e] = <mn||ie>
        CORR_TRY(perm(Pmine, at(0, 0, 0, ob), {o, o, o, v}, {s0, s2, s1, 1}));
        // Pianf[i][a][n][f] = <na||if>
        CORR_TRY(perm(Pianf, at(0, ob, 0, ob), {o, v, o, v}, {s2, s1, s0, 1}));
        // Pmbje[m][b][j][e] = <mb||ej>
        CORR_TRY(perm(Pmbje, at(0, ob, ob, 0), {o, v, o, v}, {s0, s1, 1, s2}));
        // Pabje[a][b][j][e] = <ab||ej>
        CORR_TRY(perm(Pabje, at(ob, ob, ob, 0), {v, v, o, v}, {s0, s1, 1, s2}));

        CORR_TRY(alloc(&tau_s, "cc_tau_s", O2V2));
        CORR_TRY(alloc(&tau, "cc_tau", O2V2));
        CORR_TRY(alloc(&tauS_a, "cc_tauS_a", O2V2));
        CORR_TRY(alloc(&Fae, "cc_Fae", v * v));
        CORR_TRY(alloc(&Faep, "cc_Faep", v * v));
        CORR_TRY(alloc(&Fmi, "cc_Fmi", o * o));
        CORR_TRY(alloc(&Fmip, "cc_Fmip", o * o));
        CORR_TRY(alloc(&Fme, "cc_Fme", o * v));
        CORR_TRY(alloc(&Yk, "cc_Yk", o * v));
        CORR_TRY(alloc(&Wmnij, "cc_Wmnij", O4));
        CORR_TRY(alloc(&WabefP, "cc_WabefP", std::max<int64_t>(Pv * Pv, 1)));
        CORR_TRY(alloc(&XabP, "cc_XabP", std::max<int64_t>(v * v * Pv, 1)));
        CORR_TRY(alloc(&tauP, "cc_tauP", std::max<int64_t>(Po * Pv, 1)));
        CORR_TRY(alloc(&VoovvP, "cc_VoovvP", std::max<int64_t>(Po * Pv, 1)));
        CORR_TRY(alloc(&WmnijP, "cc_WmnijP", std::max<int64_t>(Po * Po, 1)));
        if (Po > 0 && Pv > 0) {
            hipLaunchKernelGGL(pack_tau_kernel, dim3((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(Po, 65535)),
                               dim3(256), 0, ctx->stream, Voovv, o, v, po_i, po_j, pv_a, pv_b, Po, Pv, VoovvP);
            CORR_LAUNCHED(ctx, "pack_tau_kernel");
        }
        CORR_TRY(alloc(&Ypp, "cc_Ypp", std::max<int64_t>(Po * Pv, 1)));
        CORR_TRY(alloc(&Wmbej, "cc_Wmbej", O2V2));
        CORR_TRY(alloc(&Zp, "cc_Zp", O2V2));
        CORR_TRY(alloc(&W2, "cc_W2", O2V2));
        CORR_TRY(alloc(&Wp, "cc_Wp", O2V2));
        CORR_TRY(alloc(&T2p, "cc_T2p", O2V2));
        CORR_TRY(alloc(&T2a, "cc_T2a", O2V2));
        CORR_TRY(alloc(&X1, "cc_X1", O2V2));
        CORR_TRY(alloc(&X2, "cc_X2", O2V2 > O3V ? O2V2 : O3V));
        CORR_TRY(alloc(&Zr, "cc_Zr", O2V2));
        CORR_TRY(alloc(&U, "cc_U", O3V));
        CORR_TRY(alloc(&Q, "cc_Q", O2V2));
        CORR_TRY(alloc(&Tmp1, "cc_Tmp1", o * v));
        CORR_TRY(alloc(&partial, "cc_partial", 512));
        if (spin) {
            const int64_t ov = o * v;
            CORR_REQUIRE(ctx, o % 2 == 0 && v % 2 == 0, "ccsd: spin blocks need even nocc and nvir");
            CORR_TRY(alloc(&PmenfS, "cc_PmenfS", ov * ov));
            CORR_TRY(alloc(&ZpS, "cc_ZpS", ov * ov));
            CORR_TRY(alloc(&W2S, "cc_W2S", ov * ov));
            CORR_TRY(alloc(&T2S, "cc_T2S", ov * ov));
            CORR_TRY(alloc(&WpS, "cc_WpS", ov * ov));
            CORR_TRY(alloc(&ZrS, "cc_ZrS", ov * ov));
            CORR_TRY(alloc(&spinmax, "cc_spinmax", o * o));
            CORR_REQUIRE(ctx, o * o < ((int64_t)1 << 31) && v * v < ((int64_t)1 << 31), "ccsd: too many orbitals");
            // the products never write the blocks that vanish by spin: zero them once
            CORR_HIP(ctx, hipMemsetAsync(W2S, 0, sizeof(double) * ov * ov, ctx->stream));
            CORR_HIP(ctx, hipMemsetAsync(ZrS, 0, sizeof(double) * ov * ov, ctx->stream));
            CORR_HIP(ctx, hipMemsetAsync(Ypp, 0, sizeof(double) * std::max<int64_t>(Po * Pv, 1), ctx->stream));
            // PmenfS[(m e)][(n f)] = <mn||ef>
            CORR_TRY(spin_gather(PmenfS, Pmenf, v * ov, ov, v, 1));
        }
        return CORR_OK;
    }

    // ccsd_energy (ccsd.cc:147-219): E = f_kc t_kc + 1/4 t2.<kl||dc> + 1/2 t_kd Y_kd
    int energy(const double* f, const double* T1, const double* T2, double* E) {
        const int64_t ov = o * v;
        // f_ov block as a contiguous copy in Tmp1
        CORR_TRY(perm(Tmp1, f + o, {1, 1, o, v}, {0, 0, n, 1}));
        // Y[(k d)] = sum_{(l c)} <kl||dc> t_lc  (FP64: energy terms stay FP64)
        CORR_TRY(gemm(ctx, 0, 0, ov, 1, ov, 1.0, Pmenf, ov, 0, T1, 1, 0, 0.0, Yk, 1, 0, 1, CORR_PREC_FP64));
        double e1, e2, e3;
        CORR_TRY(dot(Tmp1, T1, ov, &e1));
        CORR_TRY(dot(T2, Voovv, o * o * v * v, &e2));
        CORR_TRY(dot(T1, Yk, ov, &e3));
        *E = e1 + 0.25 * e2 + 0.5 * e3;
        return CORR_OK;
    }

    // One update (T1, T2) -> (T1n, T2n)  (ccsd.cc:221-467, ccsd_intermediates.cc)
    int update(const double* f, const double* T1, const double* T2, double* T1n, double* T2n) {
        const int64_t ov = o * v, o2 = o * o, v2 = v * v, O2V2 = o2 * v2;
        const unsigned g2 = grid_for(O2V2);
        hipLaunchKernelGGL(taus_kernel, dim3(g2), dim3(256), 0, ctx->stream, T1, T2, o, v, tau_s, tau);
        CORR_LAUNCHED(ctx, "taus_kernel");
        // ---- Fae, Fmi, Fme
        hipLaunchKernelGGL(f_init_kernel, dim3(grid_for(v2 + o2)), dim3(256), 0, ctx->stream, f, T1, o, v, Fae, Fmi);
        CORR_LAUNCHED(ctx, "f_init_kernel");
        // Fae[a][e] += sum_mf t_mf <ma||fe> = - sum_(mf) t_(mf) <am||fe> (one matrix-vector
        // product over the rows (a e) of VaeK)
        CORR_TRY(mm(0, 0, v * v, 1, ov, -1.0, VaeK, ov, 0, T1, 1, 0, 1.0, Fae, 1, 0));
        // Fae[a][e] -= 1/2 sum_(mn f) tau~[mn a f] <mn||ef>
        CORR_TRY(perm(tauS_a, tau_s, {v, o, o, v}, {v, o * v2, v2, 1}));
        CORR_TRY(mm(0, 1, v, v, o2 * v, -0.5, tauS_a, o2 * v, 0, PeOOf, o2 * v, 0, 1.0, Fae, v, 0));
        // Fmi[m][i] += 1/2 sum_e f_me t_ie
        CORR_TRY(perm(Tmp1, f + o, {1, 1, o, v}, {0, 0, n, 1}));
        CORR_TRY(mm(0, 1, o, o, v, 0.5, Tmp1, v, 0, T1, v, 0, 1.0, Fmi, o, 0));
        // Fmi[(m i)] += sum_(n e) <mn||ie> t_ne
        CORR_TRY(mm(0, 0, o2, 1, ov, 1.0, Pmine, ov, 0, T1, 1, 0, 1.0, Fmi, 1, 0));
        // Fmi[m][i] += 1/2 sum_(n e f) <mn||ef> tau~[i n e f]
        CORR_TRY(mm(0, 1, o, o, o * v2, 0.5, Voovv, o * v2, 0, tau_s, o * v2, 0, 1.0, Fmi, o, 0));
        // Fme[(m e)] = f_me + sum_(n f) <mn||ef> t_nf
        CORR_TRY(mm(0, 0, ov, 1, ov, 1.0, Pmenf, ov, 0, T1, 1, 0, 0.0, Fme, 1, 0));
        hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(ov)), dim3(256), 0, ctx->stream, Fme, Tmp1, ov, 1.0, 1.0);
        CORR_LAUNCHED(ctx, "axpby_kernel");
        // ---- Wmnij = <mn||ij> + P(ij) t_je <mn||ie> + 1/2 tau_ijef <mn||ef>
        // The reference's Wmnij has 1/4 here and its Wabef a term 1/4 tau_mnab
        // <mn||ef> (ccsd_intermediates.cc:127-230).  That Wabef term adds
        // 1/2 tau_ijef * 1/4 tau_mnab <mn||ef> = 1/8 tau_mnab sum_ef tau_ijef
        // <mn||ef> to the pp ladder, which is what Wmnij's 1/4 term adds to the
        // hh ladder; so Wmnij carries 1/2 and Wabef has no tau term: the
        // amplitudes are the same and one o^2 v^4 GEMM is saved.
        CORR_TRY(perm(Wmnij, Voooo, {o, o, o, o}, {o * o2, o2, o, 1}));
        CORR_TRY(mm(0, 1, o * o2, o, v, 1.0, Vooov, v, 0, T1, v, 0, 1.0, Wmnij, o, 0));
        CORR_TRY(mm(0, 1, o, o, v, -1.0, T1, v, 0, Vooov, v, o * v, 1.0, Wmnij, o, o2, o2));
        // ... and its tau term on the packed pairs: W_mnij is antisymmetric in
        // (mn) and (ij), so WmnijP[(m<n)][(i<j)] += sum_{e<f} <mn||ef> tau_ijef
        // (= 1/2 sum_ef), 1/8 of the unpacked product; with spin-symmetric
        // amplitudes only the diagonal pair-class blocks are nonzero.
        if (Po > 0 && Pv > 0) {
            const dim3 gpp((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(Po, 65535));
            hipLaunchKernelGGL(pack_tau_kernel, gpp, dim3(256), 0, ctx->stream, tau, o, v, po_i, po_j, pv_a, pv_b, Po,
                               Pv, tauP);
            CORR_LAUNCHED(ctx, "pack_tau_kernel");
            hipLaunchKernelGGL(pack_oooo_kernel, dim3((unsigned)cdiv(Po, 256), (unsigned)std::min<int64_t>(Po, 65535)),
                               dim3(256), 0, ctx->stream, Wmnij, o, po_i, po_j, Po, WmnijP);
            CORR_LAUNCHED(ctx, "pack_oooo_kernel");
            if (spin) {
                for (int c = 0; c < 3; ++c) {
                    const int64_t r0 = po_cls[c], nr = po_cls[c + 1] - r0, k0 = pv_cls[c], nk = pv_cls[c + 1] - k0;
                    if (nr > 0 && nk > 0)
                        CORR_TRY(mm(0, 1, nr, nr, nk, 1.0, VoovvP + r0 * Pv + k0, Pv, 0, tauP + r0 * Pv + k0, Pv, 0,
                                    1.0, WmnijP + r0 * Po + r0, Po, 0));
                }
            } else {
                CORR_TRY(mm(0, 1, Po, Po, Pv, 1.0, VoovvP, Pv, 0, tauP, Pv, 0, 1.0, WmnijP, Po, 0));
            }
        }
        // ---- Wabef = <ab||ef> - P(ab) t_mb <am||ef>  (its tau term is in Wmnij),
        // antisymmetric in (ab) and (ef): kept only for a < b, e < f.
        // XabP[a][b][(e<f)] = sum_m t_mb <am||ef> (one GEMM per a), then one pass
        // WabefP = VvvvvP - XabP[a][b] + XabP[b][a]
        if (Pv > 0) {
            CORR_TRY(mm(1, 0, v, Pv, o, 1.0, T1, v, 0, VvovvP, Pv, o * Pv, 0.0, XabP, Pv, v * Pv, v));
            hipLaunchKernelGGL(wabef_packed_kernel,
                               dim3((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(Pv, 65535)), dim3(256), 0,
                               ctx->stream, VvvvvP, XabP, v, pv_a, pv_b, Pv, WabefP);
            CORR_LAUNCHED(ctx, "wabef_packed_kernel");
        }
        // ---- Wmbej = <mb||ej> + t_jf <mb||ef> - t_nb <mn||ej> - (1/2 t_jnfb + t_jf t_nb) <mn||ef>
        CORR_TRY(perm(Wmbej, Vovvo, {o, v, v, o}, {v2 * o, v * o, o, 1}));
        CORR_TRY(mm(0, 1, ov * v, o, v, 1.0, Vovvv, v, 0, T1, v, 0, 1.0, Wmbej, o, 0));
        CORR_TRY(mm(1, 0, v, v * o, o, -1.0, T1, v, 0, Voovo, v * o, o * v * o, 1.0, Wmbej, v * o, v * v * o, o));
        hipLaunchKernelGGL(zp_kernel, dim3(g2), dim3(256), 0, ctx->stream, T1, T2, o, v, Zp);
        CORR_LAUNCHED(ctx, "zp_kernel");
        if (spin) {
            // W2S = PmenfS ZpS: blocks (2,3)(3,2) and (3,2)(2,3) beside the 0|1 super-block
            CORR_TRY(spin_gather(ZpS, Zp, v * ov, ov, v, 1));
            CORR_TRY(spin_ring(PmenfS, ZpS, W2S, {2, 3, 2, 3, 2, 3}));
            // Wmbej[m][b][e][j] -= W2[(m e)][(j b)]
            CORR_TRY(spin_scatter(Wmbej, W2S, {o, v, v, o}, {0, 3, 1, 2}, -1.0, 1.0));
        } else {
            CORR_TRY(mm(0, 0, ov, ov, ov, 1.0, Pmenf, ov, 0, Zp, ov, 0, 0.0, W2, ov, 0));  // W2[m][e][j][b]
            CORR_TRY(perm(Wmbej, W2, {o, v, v, o}, {v * ov, 1, ov, v}, -1.0, 1.0));
        }

        // ---- T1 update (ccsd.cc:266-333)
        CORR_TRY(perm(T1n, f + o, {1, 1, o, v}, {0, 0, n, 1}));
        CORR_TRY(mm(0, 1, o, v, v, 1.0, T1, v, 0, Fae, v, 0, 1.0, T1n, v, 0));
        CORR_TRY(mm(1, 0, o, v, o, -1.0, Fmi, o, 0, T1, v, 0, 1.0, T1n, v, 0));
        CORR_TRY(perm(T2p, T2, {o, v, o, v}, {o * v2, v, v2, 1}));      // T2p[i][a][m][e] = t_imae
        CORR_TRY(mm(0, 0, ov, 1, ov, 1.0, T2p, ov, 0, Fme, 1, 0, 1.0, T1n, 1, 0));
        CORR_TRY(mm(0, 0, ov, 1, ov, -1.0, Pianf, ov, 0, T1, 1, 0, 1.0, T1n, 1, 0));
        // -1/2 t_imef <ma||ef> = +1/2 t_i(mef) <am||ef>
        CORR_TRY(mm(0, 1, o, v, o * v2, 0.5, T2, o * v2, 0, Vvovv, o * v2, 0, 1.0, T1n, v, 0));
        // -1/2 t_mnae <nm||ei> = +1/2 sum_(mn e) t_mnae <mn||ei>
        CORR_TRY(perm(T2a, T2, {v, o, o, v}, {v, o * v2, v2, 1}));      // T2a[a][m][n][e]
        CORR_TRY(mm(0, 0, v, o, o2 * v, 1.0, T2a, o2 * v, 0, Voovo, o, 0, 0.0, Q, o, 0));  // Q[a][i]
        CORR_TRY(perm(T1n, Q, {1, 1, o, v}, {0, 0, 1, o}, 0.5, 1.0));

        // ---- T2 update (ccsd.cc:336-467)
        CORR_TRY(perm(T2n, Voovv, {o, o, v, v}, {o * v2, v2, v, 1}));
        // Fae'[b][e] = Fae - 1/2 t_mb Fme ; P(ab) t_ijae Fae'_be
        CORR_TRY(perm(Faep, Fae, {1, 1, v, v}, {0, 0, v, 1}));
        CORR_TRY(mm(1, 0, v, v, o, -0.5, T1, v, 0, Fme, v, 0, 1.0, Faep, v, 0));
        CORR_TRY(mm(0, 1, o2 * v, v, v, 1.0, T2, v, 0, Faep, v, 0, 0.0, X1, v, 0));
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {o * v2, v2, v, 1}, 1.0, 1.0));
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {o * v2, v2, 1, v}, -1.0, 1.0));
        // Fmi'[m][j] = Fmi + 1/2 Fme t_je ; -P(ij) t_imab Fmi'_mj
        CORR_TRY(perm(Fmip, Fmi, {1, 1, o, o}, {0, 0, o, 1}));
        CORR_TRY(mm(0, 1, o, o, v, 0.5, Fme, v, 0, T1, v, 0, 1.0, Fmip, o, 0));
        CORR_TRY(mm(1, 0, o, v2, o, 1.0, Fmip, o, 0, T2, v2, o * v2, 0.0, X2, v2, o * v2, o));
        CORR_TRY(perm(T2n, X2, {o, o, v, v}, {o * v2, v2, v, 1}, -1.0, 1.0));
        CORR_TRY(perm(T2n, X2, {o, o, v, v}, {v2, o * v2, v, 1}, 1.0, 1.0));
        // pp ladder: 1/2 sum_ef tau_ijef W_abef = sum_{e<f} tau_ijef W_abef, and hh
        // ladder: 1/2 sum_mn tau_mnab W_mnij = sum_{m<n} W_mnij tau_mnab, computed
        // for i < j and a < b only (the products are symmetric under e <-> f and
        // m <-> n, T2 antisymmetric in (ij) and (ab)): GEMMs over the packed pairs
        // (tauP from the Wmnij step), 1/8 of the unpacked work, then each result
        // goes to its four T2 elements.
        if (Po > 0 && Pv > 0) {
            const dim3 gp((unsigned)cdiv(Pv, 256), (unsigned)std::min<int64_t>(Po, 65535));
            if (spin) {
                // diagonal class blocks only (the others stay zero from setup)
                for (int c = 0; c < 3; ++c) {
                    const int64_t r0 = po_cls[c], nr = po_cls[c + 1] - r0, k0 = pv_cls[c], nk = pv_cls[c + 1] - k0;
                    if (nr > 0 && nk > 0) {
                        CORR_TRY(mm(0, 1, nr, nk, nk, 1.0, tauP + r0 * Pv + k0, Pv, 0, WabefP + k0 * Pv + k0, Pv, 0,
                                    0.0, Ypp + r0 * Pv + k0, Pv, 0));
                        CORR_TRY(mm(1, 0, nr, nk, nr, 1.0, WmnijP + r0 * Po + r0, Po, 0, tauP + r0 * Pv + k0, Pv, 0,
                                    1.0, Ypp + r0 * Pv + k0, Pv, 0));
                    }
                }
            } else {
                CORR_TRY(mm(0, 1, Po, Pv, Pv, 1.0, tauP, Pv, 0, WabefP, Pv, 0, 0.0, Ypp, Pv, 0));
                CORR_TRY(mm(1, 0, Po, Pv, Po, 1.0, WmnijP, Po, 0, tauP, Pv, 0, 1.0, Ypp, Pv, 0));
            }
            hipLaunchKernelGGL(scatter_pp_kernel, gp, dim3(256), 0, ctx->stream, Ypp, o, v, po_i, po_j, pv_a, pv_b, Po,
                               Pv, T2n);
            CORR_LAUNCHED(ctx, "scatter_pp_kernel");
        }
        // ring: Zr[(ia)][(jb)] = t_imae W_mbej - t_ie t_ma <mb||ej>
        if (spin) {
            // ZrS = T2S WpS: blocks (2,3)(3,3) and (3,2)(2,2) beside the 0|1 super-block
            CORR_TRY(spin_gather(T2S, T2, o * v2, v, v2, 1));              // t_imae: (i a) (m e)
            CORR_TRY(spin_gather(WpS, Wmbej, v2 * o, o, 1, v * o));        // W_mbej: (m e) (j b)
            CORR_TRY(spin_ring(T2S, WpS, ZrS, {2, 3, 3, 3, 2, 2}));
            CORR_TRY(spin_scatter(Zr, ZrS, {o, v, o, v}, {0, 1, 2, 3}, 1.0, 0.0));
        } else {
            CORR_TRY(perm(Wp, Wmbej, {o, v, o, v}, {v2 * o, o, 1, v * o}));   // Wp[m][e][j][b]
            CORR_TRY(mm(0, 0, ov, ov, ov, 1.0, T2p, ov, 0, Wp, ov, 0, 0.0, Zr, ov, 0));
        }
        CORR_TRY(mm(0, 1, ov * o, o, v, 1.0, Pmbje, v, 0, T1, v, 0, 0.0, U, o, 0));   // U[m][b][j][i]
        CORR_TRY(mm(1, 0, v, v * o2, o, 1.0, T1, v, 0, U, v * o2, 0, 0.0, Q, v * o2, 0));  // Q[a][b][j][i]
        CORR_TRY(perm(Zr, Q, {o, v, o, v}, {1, v * o2, o, o2}, -1.0, 1.0));
        CORR_TRY(perm(T2n, Zr, {o, o, v, v}, {v * ov, v, ov, 1}, 1.0, 1.0));   // +Zr[i a j b]
        CORR_TRY(perm(T2n, Zr, {o, o, v, v}, {v * ov, v, 1, ov}, -1.0, 1.0));  // -Zr[i b j a]
        CORR_TRY(perm(T2n, Zr, {o, o, v, v}, {v, v * ov, ov, 1}, -1.0, 1.0));  // -Zr[j a i b]
        CORR_TRY(perm(T2n, Zr, {o, o, v, v}, {v, v * ov, 1, ov}, 1.0, 1.0));   // +Zr[j b i a]
        // P(ij) t_ie <ab||ej>
        CORR_TRY(mm(0, 1, v2 * o, o, v, 1.0, Pabje, v, 0, T1, v, 0, 0.0, X1, o, 0));   // X1[a][b][j][i]
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {1, o, v * o2, o2}, 1.0, 1.0));
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {o, 1, v * o2, o2}, -1.0, 1.0));
        // -P(ab) t_ma <mb||ij>
        CORR_TRY(mm(1, 0, v, v * o2, o, 1.0, T1, v, 0, Vovoo, v * o2, 0, 0.0, X1, v * o2, 0));  // X1[a][b][i][j]
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {o, 1, v * o2, o2}, -1.0, 1.0));
        CORR_TRY(perm(T2n, X1, {o, o, v, v}, {o, 1, o2, v * o2}, 1.0, 1.0));

        hipLaunchKernelGGL(denom_kernel, dim3(grid_for(ov + O2V2)), dim3(256), 0, ctx->stream, T1n, T2n, f, o, v);
        CORR_LAUNCHED(ctx, "denom_kernel");
        return CORR_OK;
    }
};

static int check_sizes(corr_ctx* ctx, int64_t o, int64_t v) {
    CORR_REQUIRE(ctx, o > 0 && v > 0, "ccsd: nocc and nvir must be positive");
    CORR_REQUIRE(ctx, v * v * v * v < ((int64_t)1 << 40), "ccsd: v^4 too large");
    return CORR_OK;
}

int ccsd_ints_so(corr_ctx* ctx, int64_t N, const double* I, double* out) {
    CORR_REQUIRE(ctx, N > 0 && I && out, "ccsd_ints_so: bad arguments");
    const int64_t n = 2 * N;
    hipLaunchKernelGGL(ints_so_kernel, dim3(grid_for(n * n * n * n)), dim3(256), 0, ctx->stream, I, N, out);
    CORR_LAUNCHED(ctx, "ints_so_kernel");
    return CORR_OK;
}

int ccsd_update(corr_ctx* ctx, int64_t o, int64_t v, const double* V, const double* f, const double* T1,
                const double* T2, double* T1n, double* T2n, int prec) {
    CORR_TRY(check_sizes(ctx, o, v));
    CORR_REQUIRE(ctx, V && f && T1 && T2 && T1n && T2n, "ccsd_update: null pointer");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32, "ccsd: precision must be FP64 or FP32");
    CCSD cc{ctx, o, v, o + v, prec};
    CORR_TRY(cc.setup(V));
    KernelTimer t(ctx, "ccsd_update");
    return cc.update(f, T1, T2, T1n, T2n);
}

int ccsd_energy(corr_ctx* ctx, int64_t o, int64_t v, const double* V, const double* f, const double* T1,
                const double* T2, double* E) {
    CORR_TRY(check_sizes(ctx, o, v));
    CORR_REQUIRE(ctx, V && f && T1 && T2 && E, "ccsd_energy: null pointer");
    CCSD cc{ctx, o, v, o + v, CORR_PREC_FP64};
    CORR_TRY(cc.setup(V));
    return cc.energy(f, T1, T2, E);
}

// Solves the (m+1) x (m+1) DIIS system by Gaussian elimination with partial pivoting.
bool diis_solve(std::vector<double> B, int m, std::vector<double>& c) {
    const int k = m + 1;
    std::vector<double> rhs(k, 0.0);
    rhs[m] = -1.0;
    for (int col = 0; col < k; ++col) {
        int piv = col;
        for (int r = col + 1; r < k; ++r)
            if (std::fabs(B[r * k + col]) > std::fabs(B[piv * k + col])) piv = r;
        if (std::fabs(B[piv * k + col]) < 1e-300) return false;
        if (piv != col) {
            for (int x = 0; x < k; ++x) std::swap(B[col * k + x], B[piv * k + x]);
            std::swap(rhs[col], rhs[piv]);
        }
        for (int r = col + 1; r < k; ++r) {
            const double fct = B[r * k + col] / B[col * k + col];
            for (int x = col; x < k; ++x) B[r * k + x] -= fct * B[col * k + x];
            rhs[r] -= fct * rhs[col];
        }
    }
    c.assign(k, 0.0);
    for (int r = k - 1; r >= 0; --r) {
        double s = rhs[r];
        for (int x = r + 1; x < k; ++x) s -= B[r * k + x] * c[x];
        c[r] = s / B[r * k + r];
    }
    c.resize(m);
    return true;
}

int ccsd_run(corr_ctx* ctx, int64_t N, int64_t nroe, const double* I, const double* C, const double* H,
             double ion_rep, const corr_ccsd_opts* opts, double* h_energies, double* T1out, double* T2out,
             corr_ccsd_result* res) {
    CORR_REQUIRE(ctx, I && C && H && opts && res, "ccsd: null pointer");
    CORR_REQUIRE(ctx, N > 0 && nroe > 0 && nroe < 2 * N && nroe % 2 == 0, "ccsd: bad nroao/nroe");
    CORR_REQUIRE(ctx, opts->mode == CORR_CCSD_REFERENCE || opts->mode == CORR_CCSD_DIIS, "ccsd: bad mode");
    CORR_REQUIRE(ctx, opts->maxiter >= 1, "ccsd: maxiter must be >= 1");
    CORR_REQUIRE(ctx, opts->prec == CORR_PREC_FP64 || opts->prec == CORR_PREC_FP32,
                 "ccsd: precision must be FP64 or FP32");
    const int64_t n = 2 * N, o = nroe, v = n - o;
    CORR_TRY(check_sizes(ctx, o, v));
    const int64_t L1 = o * v, L2 = o * o * v * v, L = L1 + L2;
    double *V, *hs, *hso, *f, *tmp;
    CORR_TRY(scratch(ctx, "cc_ints_so", sizeof(double) * (size_t)(n * n * n * n), (void**)&V));
    CORR_TRY(scratch(ctx, "cc_hs", sizeof(double) * N * N * 2, (void**)&hs));
    tmp = hs + N * N;
    CORR_TRY(scratch(ctx, "cc_hso", sizeof(double) * n * n, (void**)&hso));
    CORR_TRY(scratch(ctx, "cc_f", sizeof(double) * n * n, (void**)&f));
    CORR_TRY(ccsd_ints_so(ctx, N, I, V));
    // h = C H C^T (spatial), then spin-blocked; f = h + sum_j <pj||qj>
    CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, C, N, 0, H, N, 0, 0.0, tmp, N, 0, 1, CORR_PREC_FP64));
    CORR_TRY(gemm(ctx, 0, 1, N, N, N, 1.0, tmp, N, 0, C, N, 0, 0.0, hs, N, 0, 1, CORR_PREC_FP64));
    hipLaunchKernelGGL(form_fock_kernel, dim3(grid_for(n * n)), dim3(256), 0, ctx->stream, hs, V, N, o, hso, f);
    CORR_LAUNCHED(ctx, "form_fock_kernel");

    CCSD cc{ctx, o, v, n, opts->prec};
    cc.spin = true;   // V from ints_so_kernel: alternating alpha / beta spin orbitals
    CORR_TRY(cc.setup(V));
    // HF check
    {
        const int nb = 64;
        hipLaunchKernelGGL(ehf_kernel, dim3(nb), dim3(256), 0, ctx->stream, I, hso, f, N, o, cc.partial);
        CORR_LAUNCHED(ctx, "ehf_kernel");
        double e[3];
        CORR_TRY(reduce_sum(ctx, cc.partial, nb, 3, e));
        res->e_hf_check = e[0] + e[1] + ion_rep;
        res->e_one_check = e[0] + ion_rep;
        res->e_fsum_check = e[2] - e[1] + ion_rep;
    }
    // amplitude vectors x = [T1 | T2]
    const int nd = opts->mode == CORR_CCSD_DIIS ? std::max(2, opts->diis_size) : 0;
    double *X, *Xn;
    CORR_TRY(scratch(ctx, "cc_X", sizeof(double) * L, (void**)&X));
    CORR_TRY(scratch(ctx, "cc_Xn", sizeof(double) * L, (void**)&Xn));
    double *hist = nullptr, *errs = nullptr;
    if (nd) {
        CORR_TRY(scratch(ctx, "cc_diis_x", sizeof(double) * L * nd, (void**)&hist));
        CORR_TRY(scratch(ctx, "cc_diis_r", sizeof(double) * L * nd, (void**)&errs));
    }
    // guess: T1 = 0, T2 = <ij||ab>/D  (ccsd.cc:84-144)
    CORR_HIP(ctx, hipMemsetAsync(X, 0, sizeof(double) * L1, ctx->stream));
    CORR_HIP(ctx, hipMemcpyAsync(X + L1, cc.Voovv, sizeof(double) * L2, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(denom_kernel, dim3(grid_for(L)), dim3(256), 0, ctx->stream, (double*)nullptr, X + L1, f, o, v);
    CORR_LAUNCHED(ctx, "denom_kernel");
    {
        double e;
        CORR_TRY(cc.dot(X + L1, cc.Voovv, L2, &e));
        res->e_mp2 = 0.25 * e;
    }
    res->iterations = 0;
    res->converged = 0;
    res->r_norm = 0.0;
    double e_old = 0.0, e = 0.0;
    std::vector<double> Bm;   // DIIS overlaps of the stored residuals (nd x nd)
    if (nd) Bm.assign((size_t)nd * nd, 0.0);
    int nstored = 0, head = 0;
    for (int it = 0; it < opts->maxiter; ++it) {
        CORR_TRY(cc.energy(f, X, X + L1, &e));
        if (h_energies) h_energies[it] = e;
        res->iterations = it + 1;
        res->d_e = e_old - e;
        res->e_ccsd = e;
        if (opts->mode == CORR_CCSD_DIIS && it > 0 && std::fabs(e - e_old) < opts->e_tol &&
            res->r_norm < opts->r_tol) {
            res->converged = 1;
            break;
        }
        e_old = e;
        {
            KernelTimer t(ctx, "ccsd_update");
            CORR_TRY(cc.update(f, X, X + L1, Xn, Xn + L1));
        }
        if (cc.spin) {
            double fm;
            CORR_TRY(cc.spin_forbidden_max(Xn, Xn + L1, &fm));
            if (!(fm <= CCSD::SPIN_FORBIDDEN_TOL)) cc.spin = false;   // NaN included
        }
        if (!nd) {
            std::swap(X, Xn);
            // the swapped buffers keep their scratch registration by address only
            continue;
        }
        // residual r = Xn - X, stored with Xn in slot head
        double* xs = hist + (int64_t)head * L;
        double* rs = errs + (int64_t)head * L;
        hipLaunchKernelGGL(sub_kernel, dim3(grid_for(L)), dim3(256), 0, ctx->stream, rs, Xn, X, L);
        CORR_LAUNCHED(ctx, "sub_kernel");
        CORR_HIP(ctx, hipMemcpyAsync(xs, Xn, sizeof(double) * L, hipMemcpyDeviceToDevice, ctx->stream));
        if (nstored < nd) ++nstored;
        for (int k = 0; k < nstored; ++k) {
            double d;
            CORR_TRY(cc.dot(rs, errs + (int64_t)k * L, L, &d));
            Bm[(size_t)head * nd + k] = Bm[(size_t)k * nd + head] = d;
        }
        res->r_norm = std::sqrt(Bm[(size_t)head * nd + head]);
        head = (head + 1) % nd;
        std::vector<double> c;
        bool ok = false;
        if (nstored >= 2) {
            const int m = nstored;
            std::vector<double> Bs((size_t)(m + 1) * (m + 1), -1.0);
            Bs[(size_t)m * (m + 1) + m] = 0.0;
            for (int a = 0; a < m; ++a)
                for (int b = 0; b < m; ++b) Bs[(size_t)a * (m + 1) + b] = Bm[(size_t)a * nd + b];
            ok = diis_solve(Bs, m, c);
        }
        if (ok) {
            for (int k = 0; k < nstored; ++k) {
                hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(L)), dim3(256), 0, ctx->stream, X,
                                   hist + (int64_t)k * L, L, c[k], k == 0 ? 0.0 : 1.0);
                CORR_LAUNCHED(ctx, "axpby_kernel");
            }
        } else {
            std::swap(X, Xn);
        }
    }
    if (T1out) CORR_HIP(ctx, hipMemcpyAsync(T1out, X, sizeof(double) * L1, hipMemcpyDeviceToDevice, ctx->stream));
    if (T2out) CORR_HIP(ctx, hipMemcpyAsync(T2out, X + L1, sizeof(double) * L2, hipMemcpyDeviceToDevice, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (opts->mode == CORR_CCSD_DIIS && !res->converged) return set_error(ctx, CORR_ENOTCONV, "ccsd: not converged");
    return CORR_OK;
}

}  // namespace corr

extern "C" {

int corr_ccsd_ints_so(corr_ctx* ctx, int64_t nroao, const double* d_prec_ints, double* d_ints_so) {
    CORR_CHECK_CTX(ctx);
    return corr::ccsd_ints_so(ctx, nroao, d_prec_ints, d_ints_so);
}

int corr_ccsd_update(corr_ctx* ctx, int64_t nocc, int64_t nvir, const double* d_ints_so, const double* d_f,
                     const double* d_T1, const double* d_T2, double* d_T1n, double* d_T2n, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::ccsd_update(ctx, nocc, nvir, d_ints_so, d_f, d_T1, d_T2, d_T1n, d_T2n, prec);
}

int corr_ccsd_energy(corr_ctx* ctx, int64_t nocc, int64_t nvir, const double* d_ints_so, const double* d_f,
                     const double* d_T1, const double* d_T2, double* h_e) {
    CORR_CHECK_CTX(ctx);
    return corr::ccsd_energy(ctx, nocc, nvir, d_ints_so, d_f, d_T1, d_T2, h_e);
}

int corr_ccsd(corr_ctx* ctx, int64_t nroao, int64_t nroe, const double* d_prec_ints, const double* d_MOs,
              const double* d_Hmat, double ion_rep, const corr_ccsd_opts* opts, double* h_energies,
              double* d_T1, double* d_T2, corr_ccsd_result* h_out) {
    CORR_CHECK_CTX(ctx);
    return corr::ccsd_run(ctx, nroao, nroe, d_prec_ints, d_MOs, d_Hmat, ion_rep, opts, h_energies, d_T1, d_T2,
                          h_out);
}

}  // extern "C"
