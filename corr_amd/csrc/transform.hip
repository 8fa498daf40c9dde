// MO Fock matrix, 4-index AO->MO transform, canonical MP2 and RI B formation.
//
//  build_FMo            src/POSTHF/motrans.cc:8-32      FMo = C F C^T (2 GEMMs)
//  trans_I / MOtrans    src/POSTHF/motrans.cc:341-403   unfold the sorted unique
//                       AO list into the dense N^4 tensor (scatter kernel)
//  mo2int_op            src/POSTHF/motrans.cc:407-503   four quarter transforms;
//                       each is one GEMM out[m,n] = sum_k in[k,m] C[n,k] with
//                       M = N^3, so the contracted index is always the leading
//                       one and the result index rotates to the back:
//                       [a][bcd] -> [bcd][i] -> [cdi][j] -> [dij][k] -> [ijk][l]
//  run_canonical_mp2    src/POSTHF/can_mp2.cc:11-49     fused gather + energy
//  calculate_ri (BPQ)   src/POSTHF/ri.cc:102-103        one GEMM
//  transform_ri         src/POSTHF/ri.cc:147-172        two GEMMs
#include "common.hpp"

#include <vector>

namespace corr {

namespace {

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 * 4 ? 65536 * 4 : (g < 1 ? 1 : g));
}

__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

// Images of one sorted unique integral, by permutation class (motrans.cc:358-402).
// Index order of each image is (p q | r s) of the dense AO tensor.
__global__ void unfold_kernel(int64_t N, int64_t nrofint, int64_t s0, int64_t s1, int64_t s2,
                              int64_t s3, const double* __restrict__ val,
                              const uint16_t* __restrict__ idx, double* __restrict__ dense) {
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < nrofint;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = idx[4 * x + 0], b = idx[4 * x + 1], c = idx[4 * x + 2], d = idx[4 * x + 3];
        if (a >= N || b >= N || c >= N || d >= N) continue;
        const double w = val[x];
        auto put = [&](int64_t p, int64_t q, int64_t r, int64_t s) {
            dense[((p * N + q) * N + r) * N + s] = w;
        };
        put(a, b, c, d);                       // (1) every class
        if (x < s0) continue;                  // PERM_1
        if (x < s1) { put(c, d, a, b); continue; }                       // PERM_15: (1)(5)
        if (x < s2) { put(a, b, d, c); put(b, a, c, d); put(b, a, d, c); continue; }  // PERM_1234
        if (x < s3) { put(a, b, d, c); put(c, d, a, b); put(d, c, a, b); continue; }  // PERM_1256
        put(a, b, d, c); put(b, a, c, d); put(b, a, d, c);               // PERM_ALL
        put(c, d, a, b); put(d, c, a, b); put(c, d, b, a); put(d, c, b, a);
    }
}

// Pair-symmetric form of the same unfold.  Every image that a permutation class
// produces is a permutation of (a b | c d) that keeps the pairs {a,b}, {c,d}
// and swaps or keeps them, so the dense tensor is D[p][q][r][s] =
// V[pair(p,q)][pair(r,s)] with pair(x,y) = max(x,y)(max(x,y)+1)/2 + min(x,y) and
// V the symmetric P x P pair matrix (P = N(N+1)/2).  Step 1 puts each list
// entry at V[max(ab,cd)][min(ab,cd)] (one write; consecutive for a sorted
// list), step 2 mirrors the lower triangle into the upper one through LDS
// tiles, step 3 fills one dense row (p,q) per workgroup with coalesced writes,
// reading row pair(p,q) of V.
__device__ inline int64_t sym_pair(int64_t x, int64_t y) {
    return x >= y ? x * (x + 1) / 2 + y : y * (y + 1) / 2 + x;
}

__global__ void pair_scatter_kernel(int64_t N, int64_t nrofint, const double* __restrict__ val,
                                    const uint16_t* __restrict__ idx, double* __restrict__ V) {
    const int64_t P = N * (N + 1) / 2;
    for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < nrofint;
         x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = idx[4 * x + 0], b = idx[4 * x + 1], c = idx[4 * x + 2], d = idx[4 * x + 3];
        if (a >= N || b >= N || c >= N || d >= N) continue;
        const int64_t u = sym_pair(a, b), w = sym_pair(c, d);
        V[(u >= w ? u * P + w : w * P + u)] = val[x];
    }
}

// V[i][j] = V[j][i] for j > i: block (bi, bj), bj >= bi, reads the 64 x 64 tile
// (bj, bi) of the lower triangle and writes its transpose into tile (bi, bj).
__global__ __launch_bounds__(256) void pair_mirror_kernel(int64_t P, double* __restrict__ V) {
    const int64_t bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;
    __shared__ double t[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int64_t row = bj * 64 + r, col = bi * 64 + tx;
        t[r][tx] = (row < P && col < P) ? V[row * P + col] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int64_t row = bi * 64 + r, col = bj * 64 + tx;
        if (row < P && col < P && col > row) V[row * P + col] = t[tx][r];
    }
}

// Pair u = blockIdx.x = pair(p, q), p >= q: the N x N blocks D[p][q][.][.] and
// D[q][p][.][.] are the same symmetric matrix, whose lower triangle is row u
// of V packed (element (r, s), r >= s, at r(r+1)/2 + s); both are written from
// one read of the row.  Lower and diagonal 32 x 32 tiles read the packed row
// along s; an upper tile (r < s) is the transpose of a lower one, read along r
// into LDS and written along s, so every global access is coalesced.
__global__ __launch_bounds__(256) void pair_expand_kernel(int64_t N, const double* __restrict__ V,
                                                          double* __restrict__ dense) {
    __shared__ double t[32][33];
    const int64_t P = N * (N + 1) / 2;
    const int n = (int)N;
    const int64_t u = blockIdx.x;
    int64_t p = (int64_t)((sqrt(8.0 * (double)u + 1.0) - 1.0) * 0.5);
    while (p * (p + 1) / 2 > u) --p;
    while ((p + 1) * (p + 2) / 2 <= u) ++p;
    const int64_t q = u - p * (p + 1) / 2;
    const double* Vu = V + u * P;
    double* out = dense + (p * N + q) * N * N;
    double* out2 = dense + (q * N + p) * N * N;
    const bool two = p != q;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int nt = (n + 31) / 32;
    for (int R = 0; R < nt; ++R) {
        for (int S = 0; S < nt; ++S) {
            if (R >= S) {
                for (int k = ty; k < 32; k += 8) {
                    const int r = R * 32 + k, c = S * 32 + tx;
                    if (r < n && c < n) {
                        const double x = Vu[r >= c ? (int64_t)r * (r + 1) / 2 + c : (int64_t)c * (c + 1) / 2 + r];
                        out[(int64_t)r * n + c] = x;
                        if (two) out2[(int64_t)r * n + c] = x;
                    }
                }
            } else {
                for (int k = ty; k < 32; k += 8) {
                    const int c = S * 32 + k, r = R * 32 + tx;
                    t[k][tx] = (r < n && c < n) ? Vu[(int64_t)c * (c + 1) / 2 + r] : 0.0;
                }
                __syncthreads();
                for (int k = ty; k < 32; k += 8) {
                    const int r = R * 32 + k, c = S * 32 + tx;
                    if (r < n && c < n) {
                        out[(int64_t)r * n + c] = t[tx][k];
                        if (two) out2[(int64_t)r * n + c] = t[tx][k];
                    }
                }
                __syncthreads();
            }
        }
    }
}

// (ia|jb) = prec_ints[((i*N + o+a)*N + j)*N + o+b].  One workgroup per
// (i, j, 32 x 32 tile of (a, b)): the tile of (ia|jb) is read along b and
// the tile of (ib|ja) along a (its mirror, staged through LDS), so every read
// is coalesced; one partial pair per workgroup.
__global__ __launch_bounds__(256) void canonical_mp2_kernel(const double* __restrict__ I4,
                                                            const double* __restrict__ FMo,
                                                            int64_t N, int64_t o, int nt,
                                                            double* __restrict__ partial) {
    __shared__ double tX[32][33];   // tX[b][a] = (ib|ja)
    __shared__ double sh[2][256];
    const int64_t v = N - o;
    const int64_t blk = blockIdx.x;
    const int bt = (int)(blk % nt), at = (int)((blk / nt) % nt);
    const int64_t j = (blk / ((int64_t)nt * nt)) % o, i = blk / ((int64_t)nt * nt * o);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t a0 = (int64_t)at * 32, b0 = (int64_t)bt * 32;
    for (int r = ty; r < 32; r += 8) {
        const int64_t bb = b0 + r, aa = a0 + tx;
        tX[r][tx] = (bb < v && aa < v) ? I4[((i * N + o + bb) * N + j) * N + o + aa] : 0.0;
    }
    __syncthreads();
    const double fij = FMo[i * N + i] + FMo[j * N + j];
    double sos = 0.0, sss = 0.0;
    for (int r = ty; r < 32; r += 8) {
        const int64_t aa = a0 + r, bb = b0 + tx;
        if (aa < v && bb < v) {
            const double iajb = I4[((i * N + o + aa) * N + j) * N + o + bb];
            const double ibja = tX[tx][r];
            const double D = FMo[(o + aa) * N + o + aa] + FMo[(o + bb) * N + o + bb] - fij;
            sos -= iajb * iajb / D;
            sss -= (iajb - ibja) * iajb / D;
        }
    }
    sh[0][threadIdx.x] = sos;
    sh[1][threadIdx.x] = sss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blk + 0] = sh[0][0];
        partial[2 * blk + 1] = sh[1][0];
    }
}

int gemm_prec(int prec) { return prec == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32; }

// Precision of the AO->MO GEMMs of B formation.  Their sums cancel: FP32
// leaves a 3e-6 rms relative error in (P|ia) at (H2O)32 / def2-SVP, which the
// Cholesky metric amplifies to ~1e-4 in B and ~0.3 uEh in EMP2.  FP32 only
// when asked for; FP64 for FP64 and for the mixed path (BF16X3), whose second
// GEMM in ao2ov_sym runs on int8 MFMA with 5 slices per operand ((P|ia)
// within 1.1e-8 of FP64 relative to its maximum; 4 slices would give 1e-6
// and 0.02 uEh).
static int ao_prec(int prec) { return prec == CORR_PREC_FP32 ? CORR_PREC_FP32 : CORR_PREC_FP64; }

// Precision of the metric product Y = M X.  The dense J^-1/2 of the eigen
// metric has entries of order lambda_min^-1/2 whose products cancel in Y, so
// reduced-precision rounding of that product is amplified (FP32: 4.1 uEh in
// EMP2 on the (H2O)32 / def2-SVP-RI inputs): with the eigen metric it always
// runs in FP64, as the factorisation does.  With the Cholesky factor an FP32
// product still costs 0.28 uEh there; the mixed path (BF16X3) runs it on the
// int8 matrix cores with 5 slices per operand (max relative error of B
// 2.8e-8 from this product, 2.6e-7 with the int8 AO->MO GEMM before it;
// 0.003 uEh), at a quarter of the FP64 MFMA time.
static int metric_gemm_prec(int mode, int prec) {
    if (mode == CORR_METRIC_EIGEN || prec == CORR_PREC_FP64) return CORR_PREC_FP64;
    return prec == CORR_PREC_BF16X3 ? CORR_PREC_INT8X5 : CORR_PREC_FP32;
}

}  // namespace

int build_fmo(corr_ctx* ctx, int64_t N, const double* C, const double* F, double* FMo) {
    CORR_REQUIRE(ctx, N > 0 && C && F && FMo, "build_fmo: bad arguments");
    double* T = nullptr;
    CORR_TRY(scratch(ctx, "fmo_T", sizeof(double) * N * N, (void**)&T));
    CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, C, N, 0, F, N, 0, 0.0, T, N, 0, 1, CORR_PREC_FP64));
    CORR_TRY(gemm(ctx, 0, 1, N, N, N, 1.0, T, N, 0, C, N, 0, 0.0, FMo, N, 0, 1, CORR_PREC_FP64));
    return CORR_OK;
}

int ao_unfold(corr_ctx* ctx, int64_t N, int64_t nrofint, const int64_t* sc, const double* val,
              const uint16_t* idx, double* dense) {
    CORR_REQUIRE(ctx, N > 0 && N <= 65536 && nrofint >= 0 && sc && dense, "ao_unfold: bad arguments");
    CORR_REQUIRE(ctx, nrofint == 0 || (val && idx), "ao_unfold: null list");
    CORR_REQUIRE(ctx, 0 <= sc[0] && sc[0] <= sc[1] && sc[1] <= sc[2] && sc[2] <= sc[3] && sc[3] <= nrofint,
                 "ao_unfold: sortcount must be non-decreasing and <= nrofint");
    const int64_t n4 = N * N * N * N;
    const int64_t P = N * (N + 1) / 2;
    // Pair-symmetric expansion while the P x P pair matrix is at most 32 GiB
    // (N <= ~480), else the per-class image scatter.
    if (P * P * 8 <= ((int64_t)32 << 30)) {
        double* V = nullptr;
        CORR_TRY(scratch(ctx, "unfold_pairs", sizeof(double) * P * P, (void**)&V));
        hipLaunchKernelGGL(zero_kernel, dim3(grid_for(P * P)), dim3(256), 0, ctx->stream, V, P * P);
        CORR_LAUNCHED(ctx, "zero_kernel");
        if (nrofint > 0) {
            hipLaunchKernelGGL(pair_scatter_kernel, dim3(grid_for(nrofint)), dim3(256), 0, ctx->stream, N, nrofint,
                               val, idx, V);
            CORR_LAUNCHED(ctx, "pair_scatter_kernel");
        }
        const unsigned nb = (unsigned)cdiv(P, 64);   // P < 2^17 here
        hipLaunchKernelGGL(pair_mirror_kernel, dim3(nb, nb), dim3(256), 0, ctx->stream, P, V);
        CORR_LAUNCHED(ctx, "pair_mirror_kernel");
        hipLaunchKernelGGL(pair_expand_kernel, dim3((unsigned)P), dim3(256), 0, ctx->stream, N, V, dense);
        CORR_LAUNCHED(ctx, "pair_expand_kernel");
        return CORR_OK;
    }
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n4)), dim3(256), 0, ctx->stream, dense, n4);
    CORR_LAUNCHED(ctx, "zero_kernel");
    if (nrofint > 0) {
        hipLaunchKernelGGL(unfold_kernel, dim3(grid_for(nrofint)), dim3(256), 0, ctx->stream, N, nrofint,
                           sc[0], sc[1], sc[2], sc[3], val, idx, dense);
        CORR_LAUNCHED(ctx, "unfold_kernel");
    }
    return CORR_OK;
}

// (ij|kl) = sum C_i,mu C_j,nu C_k,la C_l,si (mu nu|la si) as four quarter
// transforms, each contracting the leading index of the current tensor and
// appending the new index last (one (N^3 x N) x (N x N) GEMM reading the
// tensor as [first][rest]).  The AO tensor is symmetric in (mu, nu), so the
// half-transformed T2[la si][i j] is symmetric in (i, j): quarter 2 writes
// only the packed pairs i <= j (P = N(N+1)/2 of N^2), quarters 3 and 4 run
// on them (half the work each: 6 N^5 instead of 8 N^5 FLOP in total), and
// quarter 4 writes every result to (ij|kl) and (ji|kl).
//   Q1 ao[mu][nu la si]           -> work[nu la si][i]
//   Q2 work[nu][la si i]          -> out[la si][(ij)]     (pairs)
//   Q3 out[la][si (ij)]           -> work[si (ij)][k]
//   Q4 work[si][(ij) k]           -> out[i j k][l], out[j i k][l]
int motrans(corr_ctx* ctx, int64_t N, const double* C, const double* ao, double* work, double* out,
            int prec) {
    CORR_REQUIRE(ctx, N > 0 && C && ao && work && out, "motrans: bad arguments");
    CORR_REQUIRE(ctx, ao != out && work != out && work != ao, "motrans: buffers must not alias");
    const int64_t N2 = N * N, N3 = N2 * N, P = N * (N + 1) / 2;
    const int p = gemm_prec(prec);
    CORR_TRY(gemm(ctx, 1, 1, N3, N, N, 1.0, ao, N3, 0, C, N, 0, 0.0, work, N, 0, 1, p));
    CORR_TRY(gemm_mapped(ctx, 1, 1, N3, N, N, work, N3, C, N, out, 1, N, p));
    CORR_TRY(gemm(ctx, 1, 1, N * P, N, N, 1.0, out, N * P, 0, C, N, 0, 0.0, work, N, 0, 1, p));
    return gemm_mapped(ctx, 1, 1, P * N, N, N, work, P * N, C, N, out, 2, N, p);
}

int canonical_mp2(corr_ctx* ctx, int64_t N, int64_t o, const double* I4, const double* FMo,
                  corr_mp2_energy* out) {
    CORR_REQUIRE(ctx, N > 0 && o > 0 && o < N && I4 && FMo && out, "canonical_mp2: bad arguments");
    const int64_t v = N - o;
    const int nt = (int)cdiv(v, 32);
    const int64_t nblk = o * o * (int64_t)nt * nt;
    CORR_REQUIRE(ctx, nblk < ((int64_t)1 << 31), "canonical_mp2: too many tiles");
    double* partial = nullptr;
    CORR_TRY(scratch(ctx, "cmp2_partial", sizeof(double) * 2 * nblk, (void**)&partial));
    hipLaunchKernelGGL(canonical_mp2_kernel, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, I4, FMo, N, o, nt,
                       partial);
    CORR_LAUNCHED(ctx, "canonical_mp2_kernel");
    double s[2];
    CORR_TRY(reduce_sum(ctx, partial, nblk, 2, s));
    out->e_os = s[0];
    out->e_ss = s[1];
    out->e_mp2 = s[0] + s[1];
    out->e_scs = s[1] / 3.0 + s[0] * 6.0 / 5.0;
    return CORR_OK;
}

int ri_bpq(corr_ctx* ctx, int64_t naux, int64_t N, const double* M, const double* Pmn, double* BPQ,
           int prec) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && M && Pmn && BPQ, "ri_bpq: bad arguments");
    return gemm(ctx, 0, 0, naux, N * N, naux, 1.0, M, naux, 0, Pmn, N * N, 0, 0.0, BPQ, N * N, 0, 1,
                gemm_prec(prec));
}

// X[Q][a][i] = sum_{mu nu} C[o+a,nu] C[i,mu] Y[Q][mu][nu]: the first AO index
// is contracted with the occupied block, as in transform_ri, into
// Z[Q][nu][i] (one tile column holds all o <= 160), then the second with the
// virtual block.  Output layout [Q][a][i].
static int ao2vo(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Y,
                 double* X, int p) {
    const int64_t v = N - o;
    double* Z = nullptr;
    CORR_TRY(scratch(ctx, "ri_Znui", sizeof(double) * naux * N * o, (void**)&Z));
    // Z[Q] (N x o) = Y[Q]^T (N x N) * C_occ^T (N x o)
    CORR_TRY(gemm(ctx, 1, 1, N, o, N, 1.0, Y, N, N * N, C, N, 0, 0.0, Z, o, N * o, naux, p));
    // X[Q] (v x o) = C_vir (v x N) * Z[Q] (N x o)
    CORR_TRY(gemm(ctx, 0, 0, v, o, N, 1.0, C + o * N, N, 0, Z, o, N * o, 0.0, X, o, v * o, naux, p));
    return CORR_OK;
}

// X[Q][i][a] = sum_{mu nu} C[o+a,mu] C[i,nu] Y[Q][mu][nu] for Y[Q] symmetric in
// (mu, nu) (3-index Coulomb integrals and their metric products are), in the
// [Q][i][a] layout the metric GEMM and the energy read, with no transpose:
//   Zt[mu][Q][i] = sum_nu Y[Q][mu][nu] C[i][nu]    (batched over mu: rows Q
//                  at stride N^2, K-contiguous; naux x o x N per batch)
//   X[(Q i)][a]  = sum_mu Zt[mu][(Q i)] C[o+a][mu]  (one GEMM, naux o x v x N)
// prec is the caller's request: FP64 or FP32 for both GEMMs, and for the mixed
// path (BF16X3) the first in FP64 and the second emulated on int8 MFMA.
static int ao2ov_sym(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Y,
                     double* X, int prec) {
    const int64_t v = N - o;
    const int p1 = ao_prec(prec), p2 = prec == CORR_PREC_BF16X3 ? CORR_PREC_INT8X5 : ao_prec(prec);
    double* Zt = nullptr;
    CORR_TRY(scratch(ctx, "ri_Znui", sizeof(double) * naux * N * o, (void**)&Zt));
    CORR_TRY(gemm(ctx, 0, 1, naux, o, N, 1.0, Y, N * N, N, C, N, 0, 0.0, Zt, o, naux * o, N, p1));
    CORR_TRY(gemm(ctx, 1, 1, naux * o, v, N, 1.0, Zt, naux * o, 0, C + o * N, N, 0, 0.0, X, v, 0, 1, p2));
    return CORR_OK;
}

// dst[q][i][a] = src[q][a][i]  (batched transpose through LDS)
__global__ __launch_bounds__(256) void transpose_vo_kernel(const double* __restrict__ src,
                                                           double* __restrict__ dst, int64_t v, int64_t o,
                                                           int64_t tv, int64_t to) {
    __shared__ double t[32][33];
    const int64_t q = blockIdx.y;
    const int64_t ta = (blockIdx.x / to) * 32, ti = (blockIdx.x % to) * 32;
    const double* s = src + q * v * o;
    double* d = dst + q * v * o;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int r = ty; r < 32; r += 8) {
        const int64_t a = ta + r, i = ti + tx;
        t[r][tx] = (a < v && i < o) ? s[a * o + i] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t i = ti + r, a = ta + tx;
        if (a < v && i < o) d[i * v + a] = t[tx][r];
    }
}

static int transpose_vo(corr_ctx* ctx, int64_t naux, int64_t v, int64_t o, const double* src, double* dst) {
    const int64_t tv = cdiv(v, 32), to = cdiv(o, 32);
    for (int64_t q0 = 0; q0 < naux; q0 += 65535) {
        const int64_t nq = std::min<int64_t>(65535, naux - q0);
        hipLaunchKernelGGL(transpose_vo_kernel, dim3((unsigned)(tv * to), (unsigned)nq), dim3(256), 0, ctx->stream,
                           src + q0 * v * o, dst + q0 * v * o, v, o, tv, to);
        CORR_LAUNCHED(ctx, "transpose_vo_kernel");
    }
    return CORR_OK;
}

int transform_ri(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C,
                 const double* BPQ, double* Bia, int prec) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && C && BPQ && Bia, "transform_ri: bad arguments");
    double* X = nullptr;
    CORR_TRY(scratch(ctx, "ri_Xvo", sizeof(double) * naux * o * (N - o), (void**)&X));
    CORR_TRY(ao2vo(ctx, naux, N, o, C, BPQ, X, ao_prec(prec)));
    return transpose_vo(ctx, naux, N - o, o, X, Bia);
}

int ri_form_bia(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Pmn,
                const double* M, int mode, double* Bia, int prec) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && C && Pmn && M && Bia,
                 "ri_form_bia: bad arguments");
    CORR_REQUIRE(ctx, mode == CORR_METRIC_EIGEN || mode == CORR_METRIC_CHOLESKY, "ri_form_bia: unknown metric mode");
    const int64_t v = N - o;
    double* X = nullptr;
    CORR_TRY(scratch(ctx, "ri_Xov", sizeof(double) * naux * o * v, (void**)&X));
    CORR_TRY(ao2ov_sym(ctx, naux, N, o, C, Pmn, X, prec));
    // Bia[Q][(i a)] = sum_P M[Q][P] X[P][(i a)]   (M lower triangular for CHOLESKY)
    return gemm(ctx, 0, 0, naux, o * v, naux, 1.0, M, naux, 0, X, o * v, 0, 0.0, Bia, o * v, 0, 1,
                metric_gemm_prec(mode, prec), mode == CORR_METRIC_CHOLESKY ? GEMM_TRI_A_LOWER : 0);
}

// The context's two overlap streams: aux_stream for the latency-bound metric
// chain on one eighth of the CUs (the first ncu/8 of the CU mask) and
// part_stream for the large GEMMs beside it on the rest, so that the chain's
// small launches never queue behind GEMM workgroups.  Measured on the
// (H2O)32 B formation (FP64, Cholesky): 26.7 ms against 29.1-29.5 ms with
// both on all CUs; 16 or 24 reserved CUs gave 27.8 and 26.7 ms, 8 gave
// 38.6 ms.  The mixed path (int8 AO->MO GEMM, one-pass slicing): 17.2-17.5 ms
// with 32 reserved CUs, 17.3-17.5 with 24, 18.9-19.2 with 40 or 48, and
// 19.2-19.4 with 32 or 64 CUs taken as every eighth CU (c % 8 == 0, < 2)
// instead of the first ones.  Falls back to one unmasked aux_stream when
// masks are unavailable.
static int overlap_streams(corr_ctx* ctx) {
    if (ctx->aux_stream) return CORR_OK;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && ncu >= 64) {
        const int words = (ncu + 31) / 32;
        std::vector<uint32_t> aux(words, 0u), part(words, 0u);
        for (int c = 0; c < ncu; ++c) (c < ncu / 8 ? aux : part)[c / 32] |= 1u << (c % 32);
        hipStream_t a = nullptr, p = nullptr;
        if (hipExtStreamCreateWithCUMask(&a, (uint32_t)words, aux.data()) == hipSuccess &&
            hipExtStreamCreateWithCUMask(&p, (uint32_t)words, part.data()) == hipSuccess) {
            ctx->aux_stream = a;
            ctx->part_stream = p;
            ctx->aux_masked = true;
            return CORR_OK;
        }
        if (a) (void)hipStreamDestroy(a);
        if (p) (void)hipStreamDestroy(p);
        (void)hipGetLastError();
    }
    CORR_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    return CORR_OK;
}

// B from J and (P|mu nu) with the metric factorisation overlapped: the
// metric runs on the context's auxiliary stream and the AO->MO GEMMs beside
// it (on the CUs the auxiliary stream leaves free, when the CUs are split),
// both ordered after the caller's prior work by an event; the metric GEMM
// runs on the caller's stream once both have finished.  When that GEMM runs
// on int8 MFMA (the mixed path), the AO->MO stream slices (P|ia) after its
// GEMMs, and L^-1 is sliced after the join, on all CUs (the chain, on one
// eighth of them, ends after the AO->MO side).
// pan != nullptr (Cholesky metric, mixed path only): the metric product writes
// the BF16x3 energy's panels instead of Bia (which may then be null).
static int ri_bia_from_j_impl(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C,
                              const double* Pmn, const double* J, int mode, double* M, double* Bia, int prec,
                              int64_t* nkept, const Bf16Panels* pan) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && C && Pmn && J && (Bia || pan),
                 "ri_bia_from_j: bad arguments");
    CORR_REQUIRE(ctx, !pan || metric_gemm_prec(mode, prec) == CORR_PREC_INT8X5,
                 "ri_bia_from_j: panel output needs the int8 metric product");
    CORR_REQUIRE(ctx, mode == CORR_METRIC_EIGEN || mode == CORR_METRIC_CHOLESKY, "ri_bia_from_j: unknown metric mode");
    CORR_REQUIRE(ctx, J != M, "ri_bia_from_j: J and M must not alias");
    const int64_t v = N - o;
    double* X = nullptr;
    CORR_TRY(scratch(ctx, "ri_Xov", sizeof(double) * naux * o * v, (void**)&X));
    if (!M) CORR_TRY(scratch(ctx, "ri_M", sizeof(double) * naux * naux, (void**)&M));
    CORR_TRY(overlap_streams(ctx));
    if (!ctx->ev_fork) CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    if (!ctx->ev_join) CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    if (!ctx->ev_part) CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_part, hipEventDisableTiming));
    hipStream_t main = ctx->stream;
    const bool int8m = metric_gemm_prec(mode, prec) == CORR_PREC_INT8X5;
    OzOperand opA, opB;
    if (mode == CORR_METRIC_EIGEN && ctx->aux_masked) {
        // the eigensolver wants the whole device: no overlap
        CORR_TRY(ao2ov_sym(ctx, naux, N, o, C, Pmn, X, prec));
        CORR_TRY(ri_metric(ctx, naux, J, mode, M, nkept));
        return gemm(ctx, 0, 0, naux, o * v, naux, 1.0, M, naux, 0, X, o * v, 0, 0.0, Bia, o * v, 0, 1,
                    metric_gemm_prec(mode, prec), false);
    }
    CORR_HIP(ctx, hipEventRecord(ctx->ev_fork, main));
    int rc = CORR_OK;
    if (ctx->aux_masked) {
        CORR_HIP(ctx, hipStreamWaitEvent(ctx->part_stream, ctx->ev_fork, 0));
        ctx->stream = ctx->part_stream;
        rc = ao2ov_sym(ctx, naux, N, o, C, Pmn, X, prec);
        if (rc == CORR_OK && int8m) rc = ozaki_prepare(ctx, "oz_metric_X", X, o * v, o * v, naux, false, &opB);
        if (rc == CORR_OK && hipEventRecord(ctx->ev_part, ctx->part_stream) != hipSuccess)
            rc = set_error(ctx, CORR_EHIP, "ri_bia_from_j: event record failed");
        ctx->stream = main;
        if (rc != CORR_OK) {
            (void)hipStreamSynchronize(ctx->part_stream);
            return rc;
        }
    } else {
        CORR_TRY(ao2ov_sym(ctx, naux, N, o, C, Pmn, X, prec));
        if (int8m) CORR_TRY(ozaki_prepare(ctx, "oz_metric_X", X, o * v, o * v, naux, false, &opB));
    }
    CORR_HIP(ctx, hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0));
    ctx->stream = ctx->aux_stream;
    rc = ri_metric(ctx, naux, J, mode, M, nkept);
    if (rc == CORR_OK && int8m && !ctx->aux_masked)
        rc = ozaki_prepare(ctx, "oz_metric_L", M, naux, naux, naux, true, &opA);
    if (rc == CORR_OK && hipEventRecord(ctx->ev_join, ctx->aux_stream) != hipSuccess)
        rc = set_error(ctx, CORR_EHIP, "ri_bia_from_j: event record failed");
    ctx->stream = main;
    if (rc != CORR_OK) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        if (ctx->part_stream) (void)hipStreamSynchronize(ctx->part_stream);
        return rc;
    }
    if (ctx->aux_masked) CORR_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_part, 0));
    CORR_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_join, 0));
    if (int8m && ctx->aux_masked) CORR_TRY(ozaki_prepare(ctx, "oz_metric_L", M, naux, naux, naux, true, &opA));
    if (int8m) return ozaki_gemm_prepared(ctx, opA, opB, 1.0, 0.0, Bia, o * v, true, pan);
    return gemm(ctx, 0, 0, naux, o * v, naux, 1.0, M, naux, 0, X, o * v, 0, 0.0, Bia, o * v, 0, 1,
                metric_gemm_prec(mode, prec), mode == CORR_METRIC_CHOLESKY ? GEMM_TRI_A_LOWER : 0);
}

int ri_bia_from_j(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Pmn,
                  const double* J, int mode, double* M, double* Bia, int prec, int64_t* nkept) {
    return ri_bia_from_j_impl(ctx, naux, N, o, C, Pmn, J, mode, M, Bia, prec, nkept, nullptr);
}

// RI-MP2 from J and (P|mu nu): B formation as ri_bia_from_j, then the energy.
// The mixed path with the Cholesky metric never writes B: its int8 metric
// product fills the BF16x3 energy's panels (the repack and 8 bytes per B
// element of HBM writes saved); every other combination forms B in scratch.
int ri_mp2_from_j(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Pmn,
                  const double* J, int mode, const double* eo, const double* ev, int prec, corr_mp2_energy* out,
                  int64_t* nkept) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && eo && ev && out, "ri_mp2_from_j: bad arguments");
    const int64_t v = N - o;
    if (mode == CORR_METRIC_CHOLESKY && prec == CORR_PREC_BF16X3) {
        Bf16Panels pan;
        CORR_TRY(rimp2_bf16_panels(ctx, naux, o, v, &pan));
        CORR_TRY(ri_bia_from_j_impl(ctx, naux, N, o, C, Pmn, J, mode, nullptr, nullptr, prec, nkept, &pan));
        return rimp2_energy_prepacked(ctx, naux, o, v, eo, ev, out);
    }
    double* B = nullptr;
    CORR_TRY(scratch(ctx, "ri_mp2_B", sizeof(double) * naux * o * v, (void**)&B));
    CORR_TRY(ri_bia_from_j(ctx, naux, N, o, C, Pmn, J, mode, nullptr, B, prec, nkept));
    return rimp2_energy(ctx, naux, o, v, B, eo, ev, prec, out);
}

}  // namespace corr

extern "C" {

int corr_ri_bia_from_j(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc, const double* d_MOs,
                       const double* d_Pmn, const double* d_J, int metric_mode, double* d_M, double* d_Bia,
                       int prec, int64_t* h_nkept) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_bia_from_j(ctx, naux, nroao, nocc, d_MOs, d_Pmn, d_J, metric_mode, d_M, d_Bia, prec, h_nkept);
}

int corr_ri_mp2_energy_from_j(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc, const double* d_MOs,
                              const double* d_Pmn, const double* d_J, int metric_mode, const double* d_eps_occ,
                              const double* d_eps_vir, int prec, corr_mp2_energy* h_out, int64_t* h_nkept) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_mp2_from_j(ctx, naux, nroao, nocc, d_MOs, d_Pmn, d_J, metric_mode, d_eps_occ, d_eps_vir, prec,
                               h_out, h_nkept);
}

int corr_build_fmo(corr_ctx* ctx, int64_t nroao, const double* d_MOs, const double* d_Fmat,
                   double* d_FMo) {
    CORR_CHECK_CTX(ctx);
    return corr::build_fmo(ctx, nroao, d_MOs, d_Fmat, d_FMo);
}

int corr_ao_unfold(corr_ctx* ctx, int64_t nroao, int64_t nrofint, const int64_t* h_sortcount,
                   const double* d_intval, const uint16_t* d_intnums, double* d_dense_ao) {
    CORR_CHECK_CTX(ctx);
    return corr::ao_unfold(ctx, nroao, nrofint, h_sortcount, d_intval, d_intnums, d_dense_ao);
}

int corr_motrans(corr_ctx* ctx, int64_t nroao, const double* d_MOs, const double* d_dense_ao,
                 double* d_work, double* d_prec_ints, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::motrans(ctx, nroao, d_MOs, d_dense_ao, d_work, d_prec_ints, prec);
}

int corr_canonical_mp2(corr_ctx* ctx, int64_t nroao, int64_t nocc, const double* d_prec_ints,
                       const double* d_FMo, corr_mp2_energy* h_out) {
    CORR_CHECK_CTX(ctx);
    return corr::canonical_mp2(ctx, nroao, nocc, d_prec_ints, d_FMo, h_out);
}

int corr_ri_bpq(corr_ctx* ctx, int64_t naux, int64_t nroao, const double* d_M, const double* d_Pmn,
                double* d_BPQ, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_bpq(ctx, naux, nroao, d_M, d_Pmn, d_BPQ, prec);
}

int corr_transform_ri(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc, const double* d_MOs,
                      const double* d_BPQ, double* d_Bia, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::transform_ri(ctx, naux, nroao, nocc, d_MOs, d_BPQ, d_Bia, prec);
}

int corr_ri_form_bia(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc, const double* d_MOs,
                     const double* d_Pmn, const double* d_M, int metric_mode, double* d_Bia, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_form_bia(ctx, naux, nroao, nocc, d_MOs, d_Pmn, d_M, metric_mode, d_Bia, prec);
}

}  // extern "C"
