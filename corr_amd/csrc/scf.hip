// RHF input side: the Fock build and the damped RHF + DIIS driver.
//
//  build_Fmat, perm_*    src/SCF/ops_rhf.cc:466-585   F = H + J(P) - 1/2 K(P)
//  build_Pmat            src/SCF/scf.cc:44-61         P = 2(1-d) C_o^T C_o + d P_old
//  calc_diis_error       src/SCF/scf.cc:105-141       S P F - F P S
//  run_scf               src/SCF/scf.cc:145-320
//  calc_S12              src/SCF/ops_rhf.cc:342-373   S^-1/2 (no cutoff)
//  diag_Fmat             src/SCF/ops_rhf.cc:425-430   T F T, eigenvectors, C = V T
//  check_initial_guess   src/SCF/ops_rhf.cc:15-37
//  Calc_e_el, oneE, twoE src/SCF/ops_rhf.cc:593-601, src/SCF/scf.cc:83-101
//
// The reference's build_Fmat adds, for every sorted unique integral, the J and
// K terms of each image its permutation class owns.  Those images are the
// entry's distinct (pq|rs) orbit, so the sum equals
//   F[p][r'] = H + sum_rs (pq|rs) P[r][s] (at r' = q) - 1/2 sum_qs (pq|r's) P[q][s]
// over the dense tensor D[p][q][r][s] that the unfold produces.  The dense
// tensor has the pair symmetries D[p][q] = D[q][p] (as N x N slabs) and
// D[p][q][r][s] = D[p][q][s][r], so one pass over the slabs with p >= q reads
// half of it:
//   J[p][q] = J[q][p] = sum_rs D[pq][r][s] P[r][s]
//   Kpart[p][q][t] = sum_r D[pq][r][t] P[q][r]   (row symmetry of the slab)
//   Kpart[q][p][t] = sum_r D[pq][r][t] P[p][r]
// and K[p][t] = sum_q Kpart[p][q][t] in a second, fixed-order pass.  Every sum
// has a fixed order, so the build is bit-reproducible (the reference's OpenMP
// array reductions are not).
#include "common.hpp"
#include <rocblas/rocblas.h>
#include <cmath>
#include <vector>

namespace corr {

int ao_unfold(corr_ctx* ctx, int64_t N, int64_t nrofint, const int64_t* sc, const double* val,
              const uint16_t* idx, double* dense);

namespace {

constexpr int FOCK_THREADS = 256;
constexpr int FOCK_WAVES = FOCK_THREADS / 64;
constexpr int FOCK_MAX_NJ = 8;                 // N <= 512
constexpr int MAX_DIIS = 15;

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// One workgroup per slab (p, q), p >= q.  Wave w streams rows r = w, w + 4, ...
// of the slab (coalesced N-long rows); lane l owns the VEC adjacent columns
// t = VEC l + 64 VEC j + v and keeps the two K column sums and its share of the
// J sum in registers, so no cross-lane reduction happens inside the row loop.
// VEC = 2 (16-byte loads, N even) halves the load instructions against VEC = 1.
// RU rows are loaded before any of them is used.  Out-of-range rows and
// columns load element 0 of the slab (in bounds) and are masked to zero.
template <int NJ, int RU, int VEC>
__global__ __launch_bounds__(FOCK_THREADS) void fock_slab_kernel(const double* __restrict__ D,
                                                                 const double* __restrict__ P, int64_t N,
                                                                 double* __restrict__ J,
                                                                 double* __restrict__ Kpart) {
    typedef double vd __attribute__((ext_vector_type(VEC)));
    __shared__ double colq[FOCK_WAVES][NJ * 64 * VEC];
    __shared__ double colp[FOCK_WAVES][NJ * 64 * VEC];
    __shared__ double jw[FOCK_WAVES];
    const int64_t b = blockIdx.x;
    int64_t p = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while (p * (p + 1) / 2 > b) --p;
    while ((p + 1) * (p + 2) / 2 <= b) ++p;
    const int64_t q = b - p * (p + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double* slab = D + (p * N + q) * N * N;
    const double* Pp = P + p * N;
    const double* Pq = P + q * N;

    double kq[NJ][VEC], kp[NJ][VEC];
    bool on[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        on[j] = VEC * lane + 64 * VEC * j < N;
#pragma unroll
        for (int v = 0; v < VEC; ++v) kq[j][v] = kp[j][v] = 0.0;
    }
    double jacc = 0.0;
    for (int64_t r0 = w; r0 < N; r0 += FOCK_WAVES * RU) {
        vd d[RU][NJ], pr[RU][NJ];
        double wq[RU], wp[RU];
#pragma unroll
        for (int u = 0; u < RU; ++u) {
            const int64_t r = r0 + FOCK_WAVES * u;
            const bool rv = r < N;
            const int64_t rc = rv ? r : 0;
            const double a = Pq[rc], c = Pp[rc];
            wq[u] = rv ? a : 0.0;
            wp[u] = rv ? c : 0.0;
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                const bool ok = rv && on[j];
                const int64_t col = ok ? VEC * lane + 64 * VEC * j : 0;
                const vd dv = *(const vd*)(slab + rc * N + col);
                const vd pv = *(const vd*)(P + rc * N + col);
                d[u][j] = ok ? dv : (vd)(0.0);
                pr[u][j] = ok ? pv : (vd)(0.0);
            }
        }
#pragma unroll
        for (int u = 0; u < RU; ++u) {
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    jacc = fma(d[u][j][v], pr[u][j][v], jacc);
                    kq[j][v] = fma(d[u][j][v], wq[u], kq[j][v]);
                    kp[j][v] = fma(d[u][j][v], wp[u], kp[j][v]);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            colq[w][VEC * lane + 64 * VEC * j + v] = kq[j][v];
            colp[w][VEC * lane + 64 * VEC * j + v] = kp[j][v];
        }
    }
    jacc = wave_sum(jacc);
    if (lane == 0) jw[w] = jacc;
    __syncthreads();
    for (int64_t t = threadIdx.x; t < N; t += FOCK_THREADS) {
        double sq = colq[0][t], sp = colp[0][t];
#pragma unroll
        for (int x = 1; x < FOCK_WAVES; ++x) {
            sq += colq[x][t];
            sp += colp[x][t];
        }
        Kpart[(p * N + q) * N + t] = sq;
        if (p != q) Kpart[(q * N + p) * N + t] = sp;
    }
    if (threadIdx.x == 0) {
        double s = jw[0];
#pragma unroll
        for (int x = 1; x < FOCK_WAVES; ++x) s += jw[x];
        J[p * N + q] = s;
        J[q * N + p] = s;
    }
}

// F[p][t] = H[p][t] + J[p][t] - 1/2 sum_q Kpart[p][q][t], q ascending.
__global__ __launch_bounds__(256) void fock_combine_kernel(const double* __restrict__ H,
                                                           const double* __restrict__ J,
                                                           const double* __restrict__ Kpart, int64_t N,
                                                           double* __restrict__ F) {
    const int64_t p = blockIdx.y;
    const int64_t t = blockIdx.x * 256 + threadIdx.x;
    if (t >= N) return;
    const double* kp = Kpart + p * N * N + t;
    double k = 0.0;
    for (int64_t q = 0; q < N; ++q) k += kp[q * N];
    F[p * N + t] = H[p * N + t] + J[p * N + t] - 0.5 * k;
}

__global__ void fill_kernel(double* __restrict__ x, int64_t n, double v) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = v;
}

// E = X - X^T
__global__ void antisym_kernel(const double* __restrict__ X, int64_t N, double* __restrict__ E) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < N * N;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = e / N, j = e % N;
        E[e] = X[e] - X[j * N + i];
    }
}

// rows of W scaled by 1/sqrt(lambda_row)
__global__ void invsqrt_rows_kernel(double* __restrict__ W, const double* __restrict__ lam, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n * n;
         e += (int64_t)gridDim.x * blockDim.x)
        W[e] /= sqrt(lam[e / n]);
}

struct DotSet {
    const double* x[MAX_DIIS + 1];
    const double* y[MAX_DIIS + 1];
    int nd;        // number of (x, y) pairs
    int abs0;      // series 0 is sum |x0| instead of a dot product
};

// Per-block partial sums of nd series, written to partial[blk * nd + s]; the
// fixed-order reduce_sum completes them.
__global__ __launch_bounds__(256) void dots_kernel(DotSet ds, int64_t n, double* __restrict__ partial) {
    __shared__ double sh[256];
    for (int s = 0; s < ds.nd; ++s) {
        const double* x = ds.x[s];
        const double* y = ds.y[s];
        double acc = 0.0;
        for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
            acc += (s == 0 && ds.abs0) ? fabs(x[e]) : x[e] * y[e];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[blockIdx.x * ds.nd + s] = sh[0];
        __syncthreads();
    }
}

struct Comb {
    const double* x[MAX_DIIS];
    double c[MAX_DIIS];
    int m;
};

// out = sum_j c_j x_j, j ascending
__global__ void lincomb_kernel(Comb cb, int64_t n, double* __restrict__ out) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int j = 0; j < cb.m; ++j) v += cb.c[j] * cb.x[j][e];
        out[e] = v;
    }
}

// sum of the diagonal of the N x N matrix C (one workgroup, fixed order)
__global__ __launch_bounds__(256) void trace_kernel(const double* __restrict__ C, int64_t N,
                                                    double* __restrict__ out) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += 256) acc += C[i * N + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// out[r] = sum_k B[r][k] x[k], one workgroup per row, fixed-order tree.
__global__ __launch_bounds__(256) void rowdot_kernel(const double* __restrict__ B, const double* __restrict__ x,
                                                     int64_t K, double* __restrict__ out) {
    __shared__ double sh[256];
    const double* row = B + (int64_t)blockIdx.x * K;
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < K; k += 256) acc = fma(row[k], x[k], acc);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// out[k] = sum_r B[r][k] d[r], r ascending (coalesced over k).
__global__ __launch_bounds__(256) void coldot_kernel(const double* __restrict__ B, const double* __restrict__ d,
                                                     int64_t rows, int64_t K, double* __restrict__ out) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= K) return;
    double acc = 0.0;
    for (int64_t r = 0; r < rows; ++r) acc = fma(B[r * K + k], d[r], acc);
    out[k] = acc;
}

// F = H + J - 1/2 K
__global__ void jk_combine_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                  const double* __restrict__ K, int64_t n, double* __restrict__ F) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        F[e] = H[e] + J[e] - 0.5 * K[e];
}

template <int NJ, int VEC>
int launch_slab(corr_ctx* ctx, const double* D, const double* P, int64_t N, double* J, double* Kpart) {
    constexpr int RU = NJ * VEC <= 2 ? 8 : (NJ * VEC <= 4 ? 4 : 2);
    const int64_t pairs = N * (N + 1) / 2;
    KernelTimer tm(ctx, "fock_slab_kernel");
    hipLaunchKernelGGL((fock_slab_kernel<NJ, RU, VEC>), dim3((unsigned)pairs), dim3(FOCK_THREADS), 0, ctx->stream,
                       D, P, N, J, Kpart);
    CORR_LAUNCHED(ctx, "fock_slab_kernel");
    return CORR_OK;
}

// Sums of several series over n elements, to the host.
int dots(corr_ctx* ctx, const DotSet& ds, int64_t n, double* h_out) {
    double* partial = nullptr;
    const int64_t nb = std::min<int64_t>(cdiv(n, 256), 512);
    CORR_TRY(scratch(ctx, "scf_partial", sizeof(double) * nb * ds.nd, (void**)&partial));
    hipLaunchKernelGGL(dots_kernel, dim3((unsigned)nb), dim3(256), 0, ctx->stream, ds, n, partial);
    CORR_LAUNCHED(ctx, "dots_kernel");
    return reduce_sum(ctx, partial, nb, ds.nd, h_out);
}

}  // namespace

int build_fmat(corr_ctx* ctx, int64_t N, const double* D, const double* P, const double* H, double* F) {
    CORR_REQUIRE(ctx, N > 0 && D && P && H && F, "build_fmat: bad arguments");
    CORR_REQUIRE(ctx, N <= 64 * FOCK_MAX_NJ, "build_fmat: nroao > 512 (the dense tensor would exceed 512 GB)");
    CORR_REQUIRE(ctx, F != P && F != H, "build_fmat: F must not alias P or H");
    double *J = nullptr, *Kpart = nullptr;
    CORR_TRY(scratch(ctx, "fock_J", sizeof(double) * N * N, (void**)&J));
    CORR_TRY(scratch(ctx, "fock_Kpart", sizeof(double) * N * N * N, (void**)&Kpart));
    // 16-byte loads need every row start 16-byte aligned: N even (the slabs
    // start at multiples of N^2 doubles and allocations are 256-byte aligned).
    if (N % 2 == 0 && ((uintptr_t)D % 16) == 0 && ((uintptr_t)P % 16) == 0) {
        switch ((int)cdiv(N, 128)) {
            case 1: CORR_TRY((launch_slab<1, 2>(ctx, D, P, N, J, Kpart))); break;
            case 2: CORR_TRY((launch_slab<2, 2>(ctx, D, P, N, J, Kpart))); break;
            case 3: CORR_TRY((launch_slab<3, 2>(ctx, D, P, N, J, Kpart))); break;
            default: CORR_TRY((launch_slab<4, 2>(ctx, D, P, N, J, Kpart))); break;
        }
    } else {
        switch ((int)cdiv(N, 64)) {
            case 1: CORR_TRY((launch_slab<1, 1>(ctx, D, P, N, J, Kpart))); break;
            case 2: CORR_TRY((launch_slab<2, 1>(ctx, D, P, N, J, Kpart))); break;
            case 3: CORR_TRY((launch_slab<3, 1>(ctx, D, P, N, J, Kpart))); break;
            case 4: CORR_TRY((launch_slab<4, 1>(ctx, D, P, N, J, Kpart))); break;
            case 5: CORR_TRY((launch_slab<5, 1>(ctx, D, P, N, J, Kpart))); break;
            case 6: CORR_TRY((launch_slab<6, 1>(ctx, D, P, N, J, Kpart))); break;
            case 7: CORR_TRY((launch_slab<7, 1>(ctx, D, P, N, J, Kpart))); break;
            default: CORR_TRY((launch_slab<8, 1>(ctx, D, P, N, J, Kpart))); break;
        }
    }
    hipLaunchKernelGGL(fock_combine_kernel, dim3((unsigned)cdiv(N, 256), (unsigned)N), dim3(256), 0, ctx->stream, H,
                       J, Kpart, N, F);
    CORR_LAUNCHED(ctx, "fock_combine_kernel");
    return CORR_OK;
}

int build_fmat_list(corr_ctx* ctx, int64_t N, int64_t nrofint, const int64_t* sc, const double* val,
                    const uint16_t* idx, const double* P, const double* H, double* F) {
    CORR_REQUIRE(ctx, N > 0 && N <= 64 * FOCK_MAX_NJ, "build_fmat: nroao must be in 1..512");
    double* D = nullptr;
    CORR_TRY(scratch(ctx, "fock_dense", sizeof(double) * N * N * N * N, (void**)&D));
    CORR_TRY(ao_unfold(ctx, N, nrofint, sc, val, idx, D));
    return build_fmat(ctx, N, D, P, H, F);
}

// RI-JK Fock matrix from B^Q_mu nu (naux x N x N, (mu nu|la si) ~ sum_Q B^Q_mu nu B^Q_la si):
//   J = sum_Q B^Q (sum_la si B^Q_la si P_la si)           two fixed-order GEMV passes
//   K = sum_Q B^Q P B^Q = Y B,  Y[mu][(Q,la)] = (B^Q P)[mu][la]
// Y is written in [mu][Q][la] order by a strided-batched GEMM (ldc = naux N),
// so K is one GEMM with inner dimension naux N (split-K, fixed-order reduce).
// No counterpart in the reference, whose SCF needs the full integral list.
int build_fmat_ri(corr_ctx* ctx, int64_t N, int64_t naux, const double* B, const double* P, const double* H,
                  double* F) {
    CORR_REQUIRE(ctx, N > 0 && naux > 0 && B && P && H && F, "build_fmat_ri: bad arguments");
    CORR_REQUIRE(ctx, F != P && F != H, "build_fmat_ri: F must not alias P or H");
    CORR_REQUIRE(ctx, naux < (1 << 30), "build_fmat_ri: naux too large");
    const int64_t nn = N * N;
    double *d = nullptr, *J = nullptr, *K = nullptr, *Y = nullptr;
    CORR_TRY(scratch(ctx, "rijk_d", sizeof(double) * naux, (void**)&d));
    CORR_TRY(scratch(ctx, "rijk_J", sizeof(double) * nn, (void**)&J));
    CORR_TRY(scratch(ctx, "rijk_K", sizeof(double) * nn, (void**)&K));
    CORR_TRY(scratch(ctx, "rijk_Y", sizeof(double) * naux * nn, (void**)&Y));
    hipLaunchKernelGGL(rowdot_kernel, dim3((unsigned)naux), dim3(256), 0, ctx->stream, B, P, nn, d);
    CORR_LAUNCHED(ctx, "rowdot_kernel");
    hipLaunchKernelGGL(coldot_kernel, dim3((unsigned)cdiv(nn, 256)), dim3(256), 0, ctx->stream, B, d, naux, nn, J);
    CORR_LAUNCHED(ctx, "coldot_kernel");
    CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, B, N, nn, P, N, 0, 0.0, Y, naux * N, N, naux, CORR_PREC_FP64));
    CORR_TRY(gemm(ctx, 0, 0, N, N, naux * N, 1.0, Y, naux * N, 0, B, N, 0, 0.0, K, N, 0, 1, CORR_PREC_FP64));
    hipLaunchKernelGGL(jk_combine_kernel, dim3(grid_for(nn)), dim3(256), 0, ctx->stream, H, J, K, nn, F);
    CORR_LAUNCHED(ctx, "jk_combine_kernel");
    return CORR_OK;
}

int build_pmat(corr_ctx* ctx, int64_t N, int64_t nroe, const double* C, double damp, double* P) {
    CORR_REQUIRE(ctx, N > 0 && C && P && nroe >= 2 && nroe / 2 <= N, "build_pmat: need 2 <= nroe <= 2 nroao");
    CORR_REQUIRE(ctx, C != P, "build_pmat: P must not alias the MOs");
    const int64_t nocc = nroe / 2;
    // column-major dgemm("N","T") of scf.cc:56 == row-major C_o^T C_o
    return gemm(ctx, 1, 0, N, N, nocc, 2.0 * (1.0 - damp), C, N, 0, C, N, 0, damp, P, N, 0, 1, CORR_PREC_FP64);
}

namespace {

struct ScfWork {
    int64_t N;
    double *T, *W, *X, *Y, *lam, *ringF, *ringE, *trace;
};

// T = S^-1/2 (calc_S12); fails with CORR_ENOTPD when S has an eigenvalue <= 0.
int calc_s12(corr_ctx* ctx, const ScfWork& w, const double* S) {
    const int64_t N = w.N;
    CORR_HIP(ctx, hipMemcpyAsync(w.W, S, sizeof(double) * N * N, hipMemcpyDeviceToDevice, ctx->stream));
    CORR_TRY(sym_eig(ctx, N, w.W, w.lam));
    double lmin = 0.0;
    CORR_HIP(ctx, hipMemcpyAsync(&lmin, w.lam, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!(lmin > 0.0)) return set_error(ctx, CORR_ENOTPD, "run_scf: overlap matrix is not positive definite");
    CORR_HIP(ctx, hipMemcpyAsync(w.X, w.W, sizeof(double) * N * N, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(invsqrt_rows_kernel, dim3(grid_for(N * N)), dim3(256), 0, ctx->stream, w.X, w.lam, N);
    CORR_LAUNCHED(ctx, "invsqrt_rows_kernel");
    return gemm(ctx, 1, 0, N, N, N, 1.0, w.W, N, 0, w.X, N, 0, 0.0, w.T, N, 0, 1, CORR_PREC_FP64);
}

// diag_Fmat: F' = T F T, F' = V diag(eps) V^T, C = V^T T (rows = MOs).
int diag_fmat(corr_ctx* ctx, const ScfWork& w, const double* F, double* C, double* eps) {
    const int64_t N = w.N;
    CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, w.T, N, 0, F, N, 0, 0.0, w.Y, N, 0, 1, CORR_PREC_FP64));
    CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, w.Y, N, 0, w.T, N, 0, 0.0, w.X, N, 0, 1, CORR_PREC_FP64));
    CORR_TRY(sym_eig(ctx, N, w.X, eps));
    return gemm(ctx, 0, 0, N, N, N, 1.0, w.X, N, 0, w.T, N, 0, 0.0, C, N, 0, 1, CORR_PREC_FP64);
}

}  // namespace

template <class Fock>
int run_scf_impl(corr_ctx* ctx, int64_t N, int64_t nroe, const double* H, const double* S, const Fock& fock,
                 double* C, double* MOens, double* P, double* F, double ion_rep, const corr_scf_opts* opts,
                 double* h_energies, double* h_errors, corr_scf_result* res) {
    CORR_REQUIRE(ctx, H && S && C && MOens && P && F && opts && res, "run_scf: null pointer");
    CORR_REQUIRE(ctx, nroe > 0 && nroe % 2 == 0 && nroe / 2 <= N, "run_scf: nroe must be even and <= 2 nroao");
    CORR_REQUIRE(ctx, opts->maxiter >= 0 && opts->diis_size >= 1 && opts->diis_size <= MAX_DIIS,
                 "run_scf: maxiter >= 0 and diis_size in 1..15");
    CORR_REQUIRE(ctx, opts->damp >= 0.0 && opts->damp < 1.0 && opts->e_tol >= 0.0, "run_scf: bad damp/e_tol");
    const int m = opts->diis_size;
    const int64_t nn = N * N;
    ScfWork w{N};
    CORR_TRY(scratch(ctx, "scf_T", sizeof(double) * nn, (void**)&w.T));
    CORR_TRY(scratch(ctx, "scf_W", sizeof(double) * nn, (void**)&w.W));
    CORR_TRY(scratch(ctx, "scf_X", sizeof(double) * nn, (void**)&w.X));
    CORR_TRY(scratch(ctx, "scf_Y", sizeof(double) * nn, (void**)&w.Y));
    CORR_TRY(scratch(ctx, "scf_lam", sizeof(double) * N, (void**)&w.lam));
    CORR_TRY(scratch(ctx, "scf_ringF", sizeof(double) * nn * m, (void**)&w.ringF));
    CORR_TRY(scratch(ctx, "scf_ringE", sizeof(double) * nn * m, (void**)&w.ringE));
    CORR_TRY(scratch(ctx, "scf_trace", sizeof(double) * 8, (void**)&w.trace));
    *res = corr_scf_result{};

    CORR_TRY(calc_s12(ctx, w, S));
    if (opts->core_guess_check) {
        hipLaunchKernelGGL(trace_kernel, dim3(1), dim3(256), 0, ctx->stream, C, N, w.trace);
        CORR_LAUNCHED(ctx, "trace_kernel");
        double tr = 0.0;
        CORR_HIP(ctx, hipMemcpyAsync(&tr, w.trace, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (std::fabs(tr) < 0.1) {
            CORR_TRY(diag_fmat(ctx, w, H, C, MOens));
            res->core_guess = 1;
        }
    }
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(nn * m)), dim3(256), 0, ctx->stream, w.ringF, nn * m, 0.0);
    CORR_LAUNCHED(ctx, "fill_kernel");
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(nn * m)), dim3(256), 0, ctx->stream, w.ringE, nn * m, 0.0);
    CORR_LAUNCHED(ctx, "fill_kernel");
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(nn)), dim3(256), 0, ctx->stream, P, nn, 0.0);
    CORR_LAUNCHED(ctx, "fill_kernel");
    CORR_TRY(build_pmat(ctx, N, nroe, C, 0.0, P));

    std::vector<int> ord(m);                        // ring slots, oldest first
    for (int i = 0; i < m; ++i) ord[i] = i;
    std::vector<double> dotc((size_t)m * m, 0.0);   // dot(E_slot_i, E_slot_j)
    double damp = opts->damp, DE = 10.0, Eold = 100.0, error = 0.0;
    bool diis = false;
    int it = 0;
    while (it < opts->maxiter && std::fabs(DE) > opts->e_tol) {
        CORR_TRY(build_pmat(ctx, N, nroe, C, damp, P));
        CORR_TRY(fock(P, F));
        const int slot = ord[0];
        for (int i = 0; i + 1 < m; ++i) ord[i] = ord[i + 1];
        ord[m - 1] = slot;
        double* Fs = w.ringF + (int64_t)slot * nn;
        double* Es = w.ringE + (int64_t)slot * nn;
        CORR_HIP(ctx, hipMemcpyAsync(Fs, F, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
        // E = S (P F) - F (P S) = X - X^T with X = S P F (S, P, F symmetric)
        CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, P, N, 0, F, N, 0, 0.0, w.Y, N, 0, 1, CORR_PREC_FP64));
        CORR_TRY(gemm(ctx, 0, 0, N, N, N, 1.0, S, N, 0, w.Y, N, 0, 0.0, w.X, N, 0, 1, CORR_PREC_FP64));
        hipLaunchKernelGGL(antisym_kernel, dim3(grid_for(nn)), dim3(256), 0, ctx->stream, w.X, N, Es);
        CORR_LAUNCHED(ctx, "antisym_kernel");
        DotSet ds{};
        ds.nd = m + 1;
        ds.abs0 = 1;
        ds.x[0] = Es;
        ds.y[0] = Es;
        for (int j = 0; j < m; ++j) {
            ds.x[j + 1] = Es;
            ds.y[j + 1] = w.ringE + (int64_t)j * nn;
        }
        double sums[MAX_DIIS + 1];
        CORR_TRY(dots(ctx, ds, nn, sums));
        error = sums[0] / (double)nn;
        for (int j = 0; j < m; ++j) dotc[(size_t)slot * m + j] = dotc[(size_t)j * m + slot] = sums[j + 1];
        if (error < opts->diis_thresh && it >= m) {
            if (!diis) {
                diis = true;
                damp = 0.0;
                res->diis_start = it;
            }
            const int k = m + 1;
            std::vector<double> B((size_t)k * k, 0.0), c;
            for (int i = 0; i < m; ++i) {
                for (int j = 0; j < m; ++j) B[(size_t)i * k + j] = dotc[(size_t)ord[i] * m + ord[j]];
                B[(size_t)i * k + m] = -1.0;
                B[(size_t)m * k + i] = -1.0;
            }
            if (diis_solve(B, m, c)) {
                Comb cb{};
                cb.m = m;
                for (int j = 0; j < m; ++j) {
                    cb.x[j] = w.ringF + (int64_t)ord[j] * nn;
                    cb.c[j] = c[j];
                }
                hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(nn)), dim3(256), 0, ctx->stream, cb, nn, F);
                CORR_LAUNCHED(ctx, "lincomb_kernel");
            } else {
                res->diis_singular += 1;
            }
        }
        DotSet de{};
        de.nd = 2;
        de.x[0] = P;
        de.y[0] = F;
        de.x[1] = P;
        de.y[1] = H;
        double pe[2];
        CORR_TRY(dots(ctx, de, nn, pe));
        const double Escf = 0.5 * (pe[0] + pe[1]);
        DE = Escf - Eold;
        Eold = Escf;
        if (h_energies) h_energies[it] = Escf + ion_rep;
        if (h_errors) h_errors[it] = error;
        CORR_TRY(diag_fmat(ctx, w, F, C, MOens));
        ++it;
    }
    CORR_TRY(build_pmat(ctx, N, nroe, C, damp, P));
    CORR_TRY(fock(P, F));
    DotSet fin{};
    fin.nd = 2;
    fin.x[0] = P;
    fin.y[0] = H;
    fin.x[1] = P;
    fin.y[1] = F;
    double oe[2];
    CORR_TRY(dots(ctx, fin, nn, oe));
    res->e_scf = Eold + ion_rep;
    res->e_one = oe[0] + ion_rep;
    res->e_two = 0.5 * (oe[1] - oe[0]);
    res->e_total = 0.5 * (oe[0] + oe[1]) + ion_rep;
    res->last_de = DE;
    res->last_error = error;
    res->iterations = it;
    res->converged = std::fabs(DE) <= opts->e_tol ? 1 : 0;
    res->diis = diis ? 1 : 0;
    return CORR_OK;
}

int run_scf(corr_ctx* ctx, int64_t N, int64_t nroe, const double* H, const double* S, const double* D, double* C,
            double* MOens, double* P, double* F, double ion_rep, const corr_scf_opts* opts, double* h_energies,
            double* h_errors, corr_scf_result* res) {
    CORR_REQUIRE(ctx, D, "run_scf: null integral tensor");
    CORR_REQUIRE(ctx, N > 0 && N <= 64 * FOCK_MAX_NJ, "run_scf: nroao must be in 1..512");
    auto fock = [&](const double* Pm, double* Fm) { return build_fmat(ctx, N, D, Pm, H, Fm); };
    return run_scf_impl(ctx, N, nroe, H, S, fock, C, MOens, P, F, ion_rep, opts, h_energies, h_errors, res);
}

int run_scf_ri(corr_ctx* ctx, int64_t N, int64_t nroe, const double* H, const double* S, int64_t naux,
               const double* B, double* C, double* MOens, double* P, double* F, double ion_rep,
               const corr_scf_opts* opts, double* h_energies, double* h_errors, corr_scf_result* res) {
    CORR_REQUIRE(ctx, B && naux > 0 && N > 0, "run_scf_ri: bad arguments");
    auto fock = [&](const double* Pm, double* Fm) { return build_fmat_ri(ctx, N, naux, B, Pm, H, Fm); };
    return run_scf_impl(ctx, N, nroe, H, S, fock, C, MOens, P, F, ion_rep, opts, h_energies, h_errors, res);
}

}  // namespace corr

extern "C" {

int corr_build_fmat(corr_ctx* ctx, int64_t nroao, const double* d_dense_ao, const double* d_Pmat,
                    const double* d_Hmat, double* d_Fmat) {
    CORR_CHECK_CTX(ctx);
    return corr::build_fmat(ctx, nroao, d_dense_ao, d_Pmat, d_Hmat, d_Fmat);
}

int corr_build_fmat_list(corr_ctx* ctx, int64_t nroao, int64_t nrofint, const int64_t* h_sortcount,
                         const double* d_intval, const uint16_t* d_intnums, const double* d_Pmat,
                         const double* d_Hmat, double* d_Fmat) {
    CORR_CHECK_CTX(ctx);
    return corr::build_fmat_list(ctx, nroao, nrofint, h_sortcount, d_intval, d_intnums, d_Pmat, d_Hmat, d_Fmat);
}

int corr_build_pmat(corr_ctx* ctx, int64_t nroao, int64_t nroe, const double* d_MOs, double damp, double* d_Pmat) {
    CORR_CHECK_CTX(ctx);
    return corr::build_pmat(ctx, nroao, nroe, d_MOs, damp, d_Pmat);
}

int corr_build_fmat_ri(corr_ctx* ctx, int64_t nroao, int64_t naux, const double* d_BPQ, const double* d_Pmat,
                       const double* d_Hmat, double* d_Fmat) {
    CORR_CHECK_CTX(ctx);
    return corr::build_fmat_ri(ctx, nroao, naux, d_BPQ, d_Pmat, d_Hmat, d_Fmat);
}

int corr_run_scf_ri(corr_ctx* ctx, int64_t nroao, int64_t nroe, const double* d_Hmat, const double* d_Smat,
                    int64_t naux, const double* d_BPQ, double* d_MOs, double* d_MOens, double* d_Pmat,
                    double* d_Fmat, double ion_rep, const corr_scf_opts* opts, double* h_energies, double* h_errors,
                    corr_scf_result* res) {
    CORR_CHECK_CTX(ctx);
    return corr::run_scf_ri(ctx, nroao, nroe, d_Hmat, d_Smat, naux, d_BPQ, d_MOs, d_MOens, d_Pmat, d_Fmat, ion_rep,
                            opts, h_energies, h_errors, res);
}

int corr_run_scf(corr_ctx* ctx, int64_t nroao, int64_t nroe, const double* d_Hmat, const double* d_Smat,
                 const double* d_dense_ao, double* d_MOs, double* d_MOens, double* d_Pmat, double* d_Fmat,
                 double ion_rep, const corr_scf_opts* opts, double* h_energies, double* h_errors,
                 corr_scf_result* res) {
    CORR_CHECK_CTX(ctx);
    return corr::run_scf(ctx, nroao, nroe, d_Hmat, d_Smat, d_dense_ao, d_MOs, d_MOens, d_Pmat, d_Fmat, ion_rep,
                         opts, h_energies, h_errors, res);
}

}  // extern "C"
