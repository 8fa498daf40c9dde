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
// V the symmetric P x P pair matrix (P = N(N+1)/2).  Step 1 puts each list entry
// at V[ab][cd] and V[cd][ab]; step 2 fills one dense row (p,q) per workgroup
// with coalesced writes, reading row pair(p,q) of V (L2-resident).
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
        V[u * P + w] = val[x];
        V[w * P + u] = val[x];
    }
}

__global__ __launch_bounds__(256) void pair_expand_kernel(int64_t N, const double* __restrict__ V,
                                                          double* __restrict__ dense) {
    const int64_t P = N * (N + 1) / 2;
    const int64_t row = blockIdx.x;   // p * N + q
    const int64_t p = row / N, q = row % N;
    const double* Vu = V + sym_pair(p, q) * P;
    double* out = dense + row * N * N;
    for (int64_t e = threadIdx.x; e < N * N; e += 256) out[e] = Vu[sym_pair(e / N, e % N)];
}

// (ia|jb) = prec_ints[((i*N + o+a)*N + j)*N + o+b]; one partial pair per block.
__global__ __launch_bounds__(256) void canonical_mp2_kernel(const double* __restrict__ I4,
                                                            const double* __restrict__ FMo,
                                                            int64_t N, int64_t o,
                                                            double* __restrict__ partial) {
    const int64_t v = N - o;
    const int64_t total = o * o * v * v;
    double sos = 0.0, sss = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = e % v, a = (e / v) % v, j = (e / (v * v)) % o, i = e / (v * v * o);
        const double iajb = I4[((i * N + o + a) * N + j) * N + o + b];
        const double ibja = I4[((i * N + o + b) * N + j) * N + o + a];
        const double D = FMo[(o + a) * N + o + a] + FMo[(o + b) * N + o + b] - FMo[i * N + i] -
                         FMo[j * N + j];
        sos -= iajb * iajb / D;
        sss -= (iajb - ibja) * iajb / D;
    }
    __shared__ double sh[2][256];
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
        partial[2 * blockIdx.x + 0] = sh[0][0];
        partial[2 * blockIdx.x + 1] = sh[1][0];
    }
}

int gemm_prec(int prec) { return prec == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32; }

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
        CORR_REQUIRE(ctx, N * N < ((int64_t)1 << 31), "ao_unfold: N too large");
        hipLaunchKernelGGL(pair_expand_kernel, dim3((unsigned)(N * N)), dim3(256), 0, ctx->stream, N, V, dense);
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

int motrans(corr_ctx* ctx, int64_t N, const double* C, const double* ao, double* work, double* out,
            int prec) {
    CORR_REQUIRE(ctx, N > 0 && C && ao && work && out, "motrans: bad arguments");
    CORR_REQUIRE(ctx, ao != out && work != out && work != ao, "motrans: buffers must not alias");
    const int64_t N3 = N * N * N;
    const int p = gemm_prec(prec);
    const double* in[4] = {ao, work, out, work};
    double* dst[4] = {work, out, work, out};
    for (int s = 0; s < 4; ++s)
        CORR_TRY(gemm(ctx, 1, 1, N3, N, N, 1.0, in[s], N3, 0, C, N, 0, 0.0, dst[s], N, 0, 1, p));
    return CORR_OK;
}

int canonical_mp2(corr_ctx* ctx, int64_t N, int64_t o, const double* I4, const double* FMo,
                  corr_mp2_energy* out) {
    CORR_REQUIRE(ctx, N > 0 && o > 0 && o < N && I4 && FMo && out, "canonical_mp2: bad arguments");
    const int nblk = 1024;
    double* partial = nullptr;
    CORR_TRY(scratch(ctx, "cmp2_partial", sizeof(double) * 2 * nblk, (void**)&partial));
    hipLaunchKernelGGL(canonical_mp2_kernel, dim3(nblk), dim3(256), 0, ctx->stream, I4, FMo, N, o, partial);
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
    CORR_TRY(ao2vo(ctx, naux, N, o, C, BPQ, X, gemm_prec(prec)));
    return transpose_vo(ctx, naux, N - o, o, X, Bia);
}

int ri_form_bia(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Pmn,
                const double* M, int mode, double* Bia, int prec) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && C && Pmn && M && Bia,
                 "ri_form_bia: bad arguments");
    CORR_REQUIRE(ctx, mode == CORR_METRIC_EIGEN || mode == CORR_METRIC_CHOLESKY, "ri_form_bia: unknown metric mode");
    const int64_t v = N - o;
    const int p = gemm_prec(prec);
    double *X = nullptr, *Y = nullptr;
    CORR_TRY(scratch(ctx, "ri_Xvo", sizeof(double) * naux * o * v, (void**)&X));
    CORR_TRY(scratch(ctx, "ri_Yvo", sizeof(double) * naux * o * v, (void**)&Y));
    CORR_TRY(ao2vo(ctx, naux, N, o, C, Pmn, X, p));
    // Y[Q][(a i)] = sum_P M[Q][P] X[P][(a i)]   (M lower triangular for CHOLESKY)
    CORR_TRY(gemm(ctx, 0, 0, naux, o * v, naux, 1.0, M, naux, 0, X, o * v, 0, 0.0, Y, o * v, 0, 1, p,
                  mode == CORR_METRIC_CHOLESKY));
    return transpose_vo(ctx, naux, v, o, Y, Bia);
}

// B from J and (P|mu nu) with the metric factorisation overlapped: the AO->MO
// GEMMs are queued on the context stream first, the metric runs on the
// context's auxiliary stream (ordered after the caller's prior work by an
// event), and the metric GEMM waits for it by a second event.
int ri_bia_from_j(corr_ctx* ctx, int64_t naux, int64_t N, int64_t o, const double* C, const double* Pmn,
                  const double* J, int mode, double* M, double* Bia, int prec, int64_t* nkept) {
    CORR_REQUIRE(ctx, naux > 0 && N > 0 && o > 0 && o < N && C && Pmn && J && Bia,
                 "ri_bia_from_j: bad arguments");
    CORR_REQUIRE(ctx, mode == CORR_METRIC_EIGEN || mode == CORR_METRIC_CHOLESKY, "ri_bia_from_j: unknown metric mode");
    CORR_REQUIRE(ctx, J != M, "ri_bia_from_j: J and M must not alias");
    const int64_t v = N - o;
    const int p = gemm_prec(prec);
    double *X = nullptr, *Y = nullptr;
    CORR_TRY(scratch(ctx, "ri_Xvo", sizeof(double) * naux * o * v, (void**)&X));
    CORR_TRY(scratch(ctx, "ri_Yvo", sizeof(double) * naux * o * v, (void**)&Y));
    if (!M) CORR_TRY(scratch(ctx, "ri_M", sizeof(double) * naux * naux, (void**)&M));
    if (!ctx->aux_stream) CORR_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    if (!ctx->ev_fork) CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    if (!ctx->ev_join) CORR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    hipStream_t main = ctx->stream;
    CORR_HIP(ctx, hipEventRecord(ctx->ev_fork, main));
    CORR_TRY(ao2vo(ctx, naux, N, o, C, Pmn, X, p));
    CORR_HIP(ctx, hipStreamWaitEvent(ctx->aux_stream, ctx->ev_fork, 0));
    ctx->stream = ctx->aux_stream;
    int rc = ri_metric(ctx, naux, J, mode, M, nkept);
    if (rc == CORR_OK && hipEventRecord(ctx->ev_join, ctx->aux_stream) != hipSuccess)
        rc = set_error(ctx, CORR_EHIP, "ri_bia_from_j: event record failed");
    ctx->stream = main;
    if (rc != CORR_OK) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        return rc;
    }
    CORR_HIP(ctx, hipStreamWaitEvent(main, ctx->ev_join, 0));
    CORR_TRY(gemm(ctx, 0, 0, naux, o * v, naux, 1.0, M, naux, 0, X, o * v, 0, 0.0, Y, o * v, 0, 1, p,
                  mode == CORR_METRIC_CHOLESKY));
    return transpose_vo(ctx, naux, v, o, Y, Bia);
}

}  // namespace corr

extern "C" {

int corr_ri_bia_from_j(corr_ctx* ctx, int64_t naux, int64_t nroao, int64_t nocc, const double* d_MOs,
                       const double* d_Pmn, const double* d_J, int metric_mode, double* d_M, double* d_Bia,
                       int prec, int64_t* h_nkept) {
    CORR_CHECK_CTX(ctx);
    return corr::ri_bia_from_j(ctx, naux, nroao, nocc, d_MOs, d_Pmn, d_J, metric_mode, d_M, d_Bia, prec, h_nkept);
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
