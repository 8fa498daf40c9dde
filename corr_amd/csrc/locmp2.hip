// Non-canonical ("semi-local") MP2: run_non_canonical_mp2,
// src/POSTHF/loc_mp2.cc:7-92.
//
//   T0_ijab = -(ia|jb) / D_ijab,            D = f_aa + f_bb - f_ii - f_jj
//   per iteration:
//     G_ijab = -sum_{k != i} f_ik T_kjab     one GEMM over the occupied block
//     R_ijab = (ia|jb) + D T_ijab + G_ijab + G_jiba
//     Rsum = sum |R|,  E = sum ((ia|jb) + R)(2 T_ijab - T_ijba)
//     T_ijab -= R_ijab / D_ijab
// The residual, both sums and the update are one pass over 32 x 32 (a, b)
// tiles of each (i, j) block; the transposed operands T_ijba and G_jiba are
// staged through LDS, so every global access is coalesced.  T is
// double-buffered (the update of T_ijab must not reach a thread still reading
// it as T_ijba), and the sums are fixed-order, so iterations are
// bit-reproducible.  Layout of K, T and G: [i][j][a][b].
#include "common.hpp"
#include <vector>

namespace corr {

namespace {

constexpr int TB = 32;   // tile edge

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

// K[i][j][a][b] = (ia|jb) from prec_ints; T = -K / D.
__global__ void ncmp2_setup_kernel(const double* __restrict__ I4, const double* __restrict__ F, int64_t N,
                                   int64_t o, double* __restrict__ K, double* __restrict__ T) {
    const int64_t v = N - o, n = o * o * v * v;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e % v, a = (e / v) % v, j = (e / (v * v)) % o, i = e / (v * v * o);
        const double k = I4[((i * N + a + o) * N + j) * N + b + o];
        const double d = F[(o + a) * N + o + a] + F[(o + b) * N + o + b] - F[i * N + i] - F[j * N + j];
        K[e] = k;
        T[e] = -k / d;
    }
}

// Occupied block of F with a zeroed diagonal (the k != i of loc_mp2.cc:46).
__global__ void ncmp2_foff_kernel(const double* __restrict__ F, int64_t N, int64_t o, double* __restrict__ Fo) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= o * o) return;
    const int64_t i = e / o, k = e % o;
    Fo[e] = i == k ? 0.0 : F[i * N + k];
}

// One workgroup per (i, j, a-tile, b-tile): residual, the two sums (per-block
// partials, fixed tree) and Tn = T - R / D.
__global__ __launch_bounds__(256) void ncmp2_step_kernel(const double* __restrict__ K, const double* __restrict__ T,
                                                         const double* __restrict__ G, const double* __restrict__ F,
                                                         int64_t N, int64_t o, int nt, double* __restrict__ Tn,
                                                         double* __restrict__ partial) {
    __shared__ double tT[TB][TB + 1];   // T[i][j][b][a] for this tile
    __shared__ double tG[TB][TB + 1];   // G[j][i][b][a]
    __shared__ double sh[2][256];
    const int64_t v = N - o;
    const int64_t blk = blockIdx.x;
    const int bt = (int)(blk % nt), at = (int)((blk / nt) % nt);
    const int64_t j = (blk / ((int64_t)nt * nt)) % o, i = blk / ((int64_t)nt * nt * o);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int64_t a0 = (int64_t)at * TB, b0 = (int64_t)bt * TB;
    const double* Tij = T + (i * o + j) * v * v;
    const double* Gji = G + (j * o + i) * v * v;
    // staged transposes: rows b of the (a0, b0) tile's mirror, columns a
    for (int r = ty; r < TB; r += 8) {
        const int64_t bb = b0 + r, aa = a0 + tx;
        const bool ok = bb < v && aa < v;
        tT[r][tx] = ok ? Tij[bb * v + aa] : 0.0;
        tG[r][tx] = ok ? Gji[bb * v + aa] : 0.0;
    }
    __syncthreads();
    const double fi = F[i * N + i], fj = F[j * N + j];
    double sr = 0.0, se = 0.0;
    for (int r = ty; r < TB; r += 8) {
        const int64_t aa = a0 + r, bb = b0 + tx;
        if (aa < v && bb < v) {
            const int64_t e = (i * o + j) * v * v + aa * v + bb;
            const double d = F[(o + aa) * N + o + aa] + F[(o + bb) * N + o + bb] - fi - fj;
            const double t = T[e];
            const double res = K[e] + d * t + G[e] + tG[tx][r];
            sr += fabs(res);
            se += (K[e] + res) * (2.0 * t - tT[tx][r]);
            Tn[e] = t - res / d;
        }
    }
    sh[0][threadIdx.x] = sr;
    sh[1][threadIdx.x] = se;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blk] = sh[0][0];
        partial[2 * blk + 1] = sh[1][0];
    }
}

}  // namespace

int non_canonical_mp2(corr_ctx* ctx, int64_t N, int64_t o, const double* I4, const double* F, int maxiter,
                      double* h_energies, double* h_rsum, double* Tout) {
    CORR_REQUIRE(ctx, I4 && F && N > 0 && o > 0 && o < N && maxiter >= 0, "non_canonical_mp2: bad arguments");
    CORR_REQUIRE(ctx, maxiter == 0 || (h_energies && h_rsum), "non_canonical_mp2: null result arrays");
    const int64_t v = N - o, n = o * o * v * v;
    const int nt = (int)cdiv(v, TB);
    const int64_t nblk = o * o * (int64_t)nt * nt;
    CORR_REQUIRE(ctx, nblk < ((int64_t)1 << 31), "non_canonical_mp2: too many tiles");
    double *K = nullptr, *T = nullptr, *T2 = nullptr, *G = nullptr, *Fo = nullptr, *partial = nullptr;
    CORR_TRY(scratch(ctx, "ncmp2_K", sizeof(double) * n, (void**)&K));
    CORR_TRY(scratch(ctx, "ncmp2_T", sizeof(double) * n, (void**)&T));
    CORR_TRY(scratch(ctx, "ncmp2_T2", sizeof(double) * n, (void**)&T2));
    CORR_TRY(scratch(ctx, "ncmp2_G", sizeof(double) * n, (void**)&G));
    CORR_TRY(scratch(ctx, "ncmp2_Fo", sizeof(double) * o * o, (void**)&Fo));
    CORR_TRY(scratch(ctx, "ncmp2_partial", sizeof(double) * 2 * nblk, (void**)&partial));
    hipLaunchKernelGGL(ncmp2_setup_kernel, dim3(grid_for(n)), dim3(256), 0, ctx->stream, I4, F, N, o, K, T);
    CORR_LAUNCHED(ctx, "ncmp2_setup_kernel");
    hipLaunchKernelGGL(ncmp2_foff_kernel, dim3((unsigned)cdiv(o * o, 256)), dim3(256), 0, ctx->stream, F, N, o, Fo);
    CORR_LAUNCHED(ctx, "ncmp2_foff_kernel");
    for (int it = 0; it < maxiter; ++it) {
        // G[i][(j,a,b)] = -sum_k Fo[i][k] T[k][(j,a,b)]
        CORR_TRY(gemm(ctx, 0, 0, o, o * v * v, o, -1.0, Fo, o, 0, T, o * v * v, 0, 0.0, G, o * v * v, 0, 1,
                      CORR_PREC_FP64));
        {
            KernelTimer tm(ctx, "ncmp2_step_kernel");
            hipLaunchKernelGGL(ncmp2_step_kernel, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, K, T, G, F, N, o,
                               nt, T2, partial);
            CORR_LAUNCHED(ctx, "ncmp2_step_kernel");
        }
        double s[2];
        CORR_TRY(reduce_sum(ctx, partial, nblk, 2, s));
        h_rsum[it] = s[0];
        h_energies[it] = s[1];
        std::swap(T, T2);
    }
    if (Tout) CORR_HIP(ctx, hipMemcpyAsync(Tout, T, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    return CORR_OK;
}

}  // namespace corr

extern "C" int corr_non_canonical_mp2(corr_ctx* ctx, int64_t nroao, int64_t nocc, const double* d_prec_ints,
                                      const double* d_FMo, int maxiter, double* h_energies, double* h_rsum,
                                      double* d_T) {
    CORR_CHECK_CTX(ctx);
    return corr::non_canonical_mp2(ctx, nroao, nocc, d_prec_ints, d_FMo, maxiter, h_energies, h_rsum, d_T);
}
