// RI metric factorisation (the arithmetic of calculate_ri, src/POSTHF/ri.cc:50-69).
//
// CORR_METRIC_EIGEN:    J = V diag(l) V^T (rocSOLVER dsyevd, FP64), then
//                       J^-1/2 = V diag(s) V^T with s = 0 where l/l_max < 1e-10
//                       or l < 0, else l^-1/2  (ri.cc:56-69).
// CORR_METRIC_CHOLESKY: J = L L^T by a blocked right-looking Cholesky on the
//                       GPU (64x64 diagonal blocks factored and inverted in
//                       LDS, panel and trailing updates on the MFMA GEMM),
//                       then L^-1 by recursive halving (two GEMMs per split,
//                       X21 = -X22 L21 X11).  Both give
//                       M with M^T M = J^-1, hence identical B^T B products.
#include "common.hpp"

#include <string>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <vector>
#include <cmath>

namespace corr {

namespace {
constexpr int NB = 64;

// Factor the kb x kb diagonal block at A[k0,k0] (lda) in place (lower) and
// write its inverse (lower, kb x kb, ld NB) to Linv.  Sets *notpd on failure.
//
// The 2080 lower-triangle entries are owned by the threads (9 slots each) and
// live in registers.  Factorisation is right-looking and unscaled: step j
// subtracts S[i][j] S[l][j] / S[j][j] from every owned (i, l) with l > j,
// reading column j from a double-buffered LDS column that the owners of
// column j published at the end of step j-1; the scaling by 1/sqrt(S[l][l])
// is applied at the end.  The inverse is a right-looking forward
// substitution of L X = I in the same ownership: step k finalises row k of X
// (X[k][c] = R[k][c] / L[k][k]) and subtracts L[i][k] X[k][c] from the rows
// below.  One barrier per step in both phases.
constexpr int TRI = NB * (NB + 1) / 2;
constexpr int SLOTS = (TRI + 255) / 256;

__global__ __launch_bounds__(256) void chol_diag_kernel(double* __restrict__ A, int64_t lda,
                                                        int64_t k0, int kb,
                                                        double* __restrict__ Linv,
                                                        int* __restrict__ notpd) {
    __shared__ double col[2][NB];
    __shared__ double Ls[NB][NB + 1];
    __shared__ int bad;
    const int tid = threadIdx.x;
    if (tid == 0) bad = 0;
    int si[SLOTS], sl[SLOTS];
    bool act[SLOTS];
    double val[SLOTS];
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
        const int e = tid + 256 * u;
        int i = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        si[u] = i;
        sl[u] = e - i * (i + 1) / 2;
        act[u] = e < TRI && si[u] < kb;
        val[u] = act[u] ? A[(k0 + si[u]) * lda + (k0 + sl[u])] : 0.0;
        if (act[u] && sl[u] == 0) col[0][si[u]] = val[u];
    }
    // LDS indices of the slots, clamped into the arrays (inactive slots read
    // a valid element whose value they never use)
    int ri[SLOTS], rl[SLOTS];
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
        ri[u] = si[u] < NB ? si[u] : NB - 1;
        rl[u] = sl[u] < NB ? sl[u] : NB - 1;
    }
    // factorisation.  All LDS reads of a step are issued before any update,
    // with no branch between them, so their latencies overlap.
    for (int j = 0; j < kb; ++j) {
        __syncthreads();
        const double* cj = col[j & 1];
        double dj = cj[j];
        double a[SLOTS], b[SLOTS];
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            a[u] = cj[ri[u]];
            b[u] = cj[rl[u]];
        }
        if (!(dj > 0.0)) {
            if (tid == 0) bad = 1;
            dj = 1.0;
        }
        const double inv = 1.0 / dj;
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            const bool upd = act[u] && sl[u] > j;
            const double nv = val[u] - a[u] * b[u] * inv;
            val[u] = upd ? nv : val[u];
            if (upd && sl[u] == j + 1) col[(j + 1) & 1][si[u]] = val[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SLOTS; ++u)
        if (act[u] && si[u] == sl[u]) col[0][si[u]] = val[u] > 0.0 ? sqrt(val[u]) : 1.0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) {
        if (act[u]) {
            const double d = col[0][sl[u]];
            val[u] = (si[u] == sl[u]) ? d : val[u] / d;
            Ls[si[u]][sl[u]] = val[u];
            A[(k0 + si[u]) * lda + (k0 + sl[u])] = val[u];
        }
    }
    // inverse: R = I, then row by row
#pragma unroll
    for (int u = 0; u < SLOTS; ++u) val[u] = (act[u] && si[u] == sl[u]) ? 1.0 : 0.0;
    __syncthreads();
    for (int k = 0; k < kb; ++k) {
        double* rk = col[k & 1];
        const double dkk = Ls[k][k];
#pragma unroll
        for (int u = 0; u < SLOTS; ++u)
            if (act[u] && si[u] == k) {
                val[u] /= dkk;
                rk[sl[u]] = val[u];
            }
        __syncthreads();
        double a[SLOTS], b[SLOTS];
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            a[u] = Ls[ri[u]][k];
            b[u] = rk[rl[u]];
        }
#pragma unroll
        for (int u = 0; u < SLOTS; ++u) {
            const double nv = val[u] - a[u] * b[u];
            val[u] = (act[u] && si[u] > k && sl[u] <= k) ? nv : val[u];
        }
    }
    for (int e = tid; e < NB * NB; e += 256) {
        const int r = e / NB, c = e % NB;
        if (r < kb && c < kb) {
            if (c > r) {
                A[(k0 + r) * lda + (k0 + c)] = 0.0;
                Linv[r * NB + c] = 0.0;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < SLOTS; ++u)
        if (act[u]) Linv[si[u] * NB + sl[u]] = val[u];
    if (tid == 0 && bad) *notpd = 1;
}

__global__ void copy2d_kernel(const double* __restrict__ src, int64_t lds, double* __restrict__ dst,
                              int64_t ldd, int64_t rows, int64_t cols) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < rows * cols;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / cols, c = e % cols;
        dst[r * ldd + c] = src[r * lds + c];
    }
}

__global__ void zero_upper_kernel(double* __restrict__ A, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / n, c = e % n;
        if (c > r) A[e] = 0.0;
    }
}

__global__ void fill_kernel(double* __restrict__ A, int64_t n, double v) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        A[e] = v;
}

// rows of W scaled: W[c][:] *= s(lambda_c) with the reference cutoff;
// nkept[0] counts the kept eigenvalues, nkept[1] the non-finite ones.
__global__ void eig_scale_kernel(double* __restrict__ W, const double* __restrict__ lam, int64_t n,
                                 int64_t* __restrict__ nkept) {
    const double lmax = lam[n - 1];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = e / n;
        const double l = lam[c];
        const double s = (l / lmax < 1.0e-10 || l < 0.0) ? 0.0 : 1.0 / sqrt(l);
        W[e] *= s;
        if (e % n == 0 && s != 0.0) atomicAdd((unsigned long long*)nkept, 1ull);
        if (e % n == 0 && !isfinite(l)) atomicAdd((unsigned long long*)(nkept + 1), 1ull);
    }
}

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}
}  // namespace

// X[r0:r0+nr, r0:r0+nr] = inverse of the lower factor block of L (ld n);
// the block spans nb whole NB-blocks starting at block r0 / NB.
static int trinv_rec(corr_ctx* ctx, const double* L, int64_t n, int64_t r0, int64_t nb, int64_t nr,
                     double* X, const double* Linv, double* T) {
    if (nb == 1) {
        hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for(nr * nr)), dim3(256), 0, ctx->stream,
                           Linv + (r0 / NB) * NB * NB, (int64_t)NB, X + r0 * n + r0, n, nr, nr);
        CORR_LAUNCHED(ctx, "copy2d_kernel");
        return CORR_OK;
    }
    const int64_t nb1 = (nb + 1) / 2, n1 = nb1 * NB, n2 = nr - n1;
    CORR_TRY(trinv_rec(ctx, L, n, r0, nb1, n1, X, Linv, T));
    CORR_TRY(trinv_rec(ctx, L, n, r0 + n1, nb - nb1, n2, X, Linv, T));
    // T (n2 x n1) = L21 X11 ;  X21 = -X22 T  (X22 lower triangular)
    CORR_TRY(gemm(ctx, 0, 0, n2, n1, n1, 1.0, L + (r0 + n1) * n + r0, n, 0, X + r0 * n + r0, n, 0,
                  0.0, T, n1, 0, 1, CORR_PREC_FP64, GEMM_TRI_B_LOWER));
    CORR_TRY(gemm(ctx, 0, 0, n2, n1, n2, -1.0, X + (r0 + n1) * n + (r0 + n1), n, 0, T, n1, 0, 0.0,
                  X + (r0 + n1) * n + r0, n, 0, 1, CORR_PREC_FP64, GEMM_TRI_A_LOWER));
    return CORR_OK;
}

static int cholesky_inverse(corr_ctx* ctx, int64_t n, const double* J, double* M) {
    double *A = nullptr, *P = nullptr, *Linv = nullptr;
    int* flag = nullptr;
    CORR_TRY(scratch(ctx, "chol_A", sizeof(double) * n * n, (void**)&A));
    CORR_TRY(scratch(ctx, "chol_P", sizeof(double) * n * NB, (void**)&P));
    const int64_t nblk = cdiv(n, NB);
    CORR_TRY(scratch(ctx, "chol_Linv", sizeof(double) * NB * NB * nblk, (void**)&Linv));
    CORR_TRY(scratch(ctx, "chol_flag", sizeof(int) * 4, (void**)&flag));
    CORR_HIP(ctx, hipMemcpyAsync(A, J, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    CORR_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
    for (int64_t kb0 = 0; kb0 < nblk; ++kb0) {
        const int64_t k0 = kb0 * NB;
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        double* Li = Linv + kb0 * NB * NB;
        hipLaunchKernelGGL(chol_diag_kernel, dim3(1), dim3(256), 0, ctx->stream, A, n, k0, kb, Li, flag);
        CORR_LAUNCHED(ctx, "chol_diag_kernel");
        const int64_t m = n - k0 - kb;
        if (m <= 0) break;
        // Panel: L[k0+kb:, k0:k0+kb] = A[...] * Linv^T
        double* Ap = A + (k0 + kb) * n + k0;
        CORR_TRY(gemm(ctx, 0, 1, m, kb, kb, 1.0, Ap, n, 0, Li, NB, 0, 0.0, P, NB, 0, 1, CORR_PREC_FP64));
        hipLaunchKernelGGL(copy2d_kernel, dim3(grid_for(m * kb)), dim3(256), 0, ctx->stream, P,
                           (int64_t)NB, Ap, n, m, (int64_t)kb);
        CORR_LAUNCHED(ctx, "copy2d_kernel");
        // Trailing update: A22 -= P P^T on and below the diagonal (the upper
        // triangle is never read and is zeroed at the end)
        double* A22 = A + (k0 + kb) * n + (k0 + kb);
        CORR_TRY(gemm(ctx, 0, 1, m, m, kb, -1.0, P, NB, 0, P, NB, 0, 1.0, A22, n, 0, 1, CORR_PREC_FP64,
                      GEMM_TRI_C_LOWER));
    }
    hipLaunchKernelGGL(zero_upper_kernel, dim3(grid_for(n * n)), dim3(256), 0, ctx->stream, A, n);
    CORR_LAUNCHED(ctx, "zero_upper_kernel");
    // X = L^-1 by recursive halving over whole diagonal blocks:
    //   inv([[L11, 0], [L21, L22]]) = [[X11, 0], [-X22 L21 X11, X22]]
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n * n)), dim3(256), 0, ctx->stream, M, n * n, 0.0);
    CORR_LAUNCHED(ctx, "fill_kernel");
    double* T = nullptr;
    CORR_TRY(scratch(ctx, "chol_T", sizeof(double) * (n / 2 + NB) * (n / 2 + NB), (void**)&T));
    CORR_TRY(trinv_rec(ctx, A, n, 0, nblk, n, M, Linv, T));
    int h_flag = 0;
    CORR_HIP(ctx, hipMemcpyAsync(&h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h_flag) return set_error(ctx, CORR_ENOTPD, "ri metric: J is not positive definite");
    return CORR_OK;
}

int blas_handle(corr_ctx* ctx, rocblas_handle* h) {
    if (!ctx->blas_handle) {
        rocblas_handle hh = nullptr;
        if (rocblas_create_handle(&hh) != rocblas_status_success)
            return set_error(ctx, CORR_EHIP, "rocblas_create_handle failed");
        ctx->blas_handle = hh;
    }
    *h = (rocblas_handle)ctx->blas_handle;
    if (rocblas_set_stream(*h, ctx->stream) != rocblas_status_success)
        return set_error(ctx, CORR_EHIP, "rocblas_set_stream failed");
    return CORR_OK;
}

int sym_eig(corr_ctx* ctx, int64_t n, double* A, double* lam) {
    CORR_REQUIRE(ctx, n > 0 && n < (1 << 30), "sym_eig: bad order");
    double* E = nullptr;
    int* info = nullptr;
    CORR_TRY(scratch(ctx, "eig_E", sizeof(double) * n, (void**)&E));
    CORR_TRY(scratch(ctx, "eig_info", 64, (void**)&info));
    CORR_HIP(ctx, hipMemsetAsync(info, 0, 64, ctx->stream));
    rocblas_handle h;
    CORR_TRY(blas_handle(ctx, &h));
    // A symmetric row-major buffer is also its column-major form; rocSOLVER
    // reads the column-major upper triangle (the reference's dsyev 'U').  On
    // return the buffer holds the eigenvectors as columns (column-major),
    // i.e. row c of the row-major view is eigenvector c, eigenvalues ascending.
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, (rocblas_int)n, A,
                         (rocblas_int)n, lam, E, (rocblas_int*)info) != rocblas_status_success)
        return set_error(ctx, CORR_EHIP, "rocsolver_dsyevd failed");
    int h_info = 0;
    CORR_HIP(ctx, hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h_info != 0) return set_error(ctx, CORR_ENOTCONV, "rocsolver_dsyevd did not converge");
    return CORR_OK;
}

static int eigen_invsqrt(corr_ctx* ctx, int64_t n, const double* J, double* M, int64_t* nkept) {
    CORR_REQUIRE(ctx, n < (1 << 30), "ri metric: naux too large for dsyevd");
    double *W = nullptr, *lam = nullptr;
    int64_t* kept = nullptr;
    CORR_TRY(scratch(ctx, "eig_W", sizeof(double) * n * n, (void**)&W));
    CORR_TRY(scratch(ctx, "eig_lam", sizeof(double) * n, (void**)&lam));
    CORR_TRY(scratch(ctx, "eig_kept", 64, (void**)&kept));
    CORR_HIP(ctx, hipMemcpyAsync(W, J, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    CORR_HIP(ctx, hipMemsetAsync(kept, 0, 64, ctx->stream));
    CORR_TRY(sym_eig(ctx, n, W, lam));
    // Wt = rows scaled copy; M = W^T diag(s) W
    double* Ws = nullptr;
    CORR_TRY(scratch(ctx, "eig_Ws", sizeof(double) * n * n, (void**)&Ws));
    CORR_HIP(ctx, hipMemcpyAsync(Ws, W, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(eig_scale_kernel, dim3(grid_for(n * n)), dim3(256), 0, ctx->stream, Ws, lam, n, kept);
    CORR_LAUNCHED(ctx, "eig_scale_kernel");
    CORR_TRY(gemm(ctx, 1, 0, n, n, n, 1.0, W, n, 0, Ws, n, 0, 0.0, M, n, 0, 1, CORR_PREC_FP64));
    int64_t h_kept[2] = {0, 0};
    CORR_HIP(ctx, hipMemcpyAsync(h_kept, kept, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h_kept[1] != 0)
        return set_error(ctx, CORR_ENOTPD, "ri metric: J has " + std::to_string(h_kept[1]) +
                                               " non-finite eigenvalues (NaN or Inf in J)");
    if (nkept) *nkept = h_kept[0];
    return CORR_OK;
}

int ri_metric(corr_ctx* ctx, int64_t naux, const double* J, int mode, double* M, int64_t* nkept) {
    CORR_REQUIRE(ctx, naux > 0 && J && M, "ri metric: bad arguments");
    CORR_REQUIRE(ctx, J != M, "ri metric: J and M must not alias");
    if (mode == CORR_METRIC_CHOLESKY) {
        CORR_TRY(cholesky_inverse(ctx, naux, J, M));
        if (nkept) *nkept = naux;
        return CORR_OK;
    }
    CORR_REQUIRE(ctx, mode == CORR_METRIC_EIGEN, "ri metric: unknown mode");
    return eigen_invsqrt(ctx, naux, J, M, nkept);
}

}  // namespace corr

extern "C" int corr_blas_release(corr_ctx* ctx) {
    if (ctx && ctx->blas_handle) {
        rocblas_destroy_handle((rocblas_handle)ctx->blas_handle);
        ctx->blas_handle = nullptr;
    }
    return CORR_OK;
}

extern "C" int corr_ri_metric(corr_ctx* ctx, int64_t naux, const double* d_J, int mode,
                              double* d_M, int64_t* h_nkept) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_metric(ctx, naux, d_J, mode, d_M, h_nkept);
}
