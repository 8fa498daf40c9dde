// Generic strided-batched GEMM on CDNA4 matrix cores (row-major operands).
//
//   C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b]
//
// FP64 path: v_mfma_f64_16x16x4_f64, FP64 accumulate.
// FP32 path: operands rounded to FP32 when staged into LDS,
//            v_mfma_f32_16x16x4_f32, FP32 accumulate, epilogue in FP64.
//
// A workgroup is 8 wave64s (two per SIMD); each wave owns a 64x32 block of C
// built from 4x2 MFMA 16x16 tiles.  Tiles of A and B are staged through LDS as [k][m]
// (m contiguous, padded so that a wave's fragment read touches all 64 banks
// once) with the next K-slab prefetched into registers while the current one
// is consumed.  Global loads are lane-contiguous along whichever dimension
// the operand is contiguous in.
#include "common.hpp"

namespace corr {

struct GemmArgs {
    int64_t M, N, K;
    const double* A; int64_t lda, sA;
    const double* B; int64_t ldb, sB;
    double* C; int64_t ldc, sC;
    double alpha, beta;
    int64_t tilesM;
    int a_lower;
};

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef corr_d4 acc_t;
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef corr_f4 acc_t;
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// One operand slab (BK deep, BDIM wide) of a GEMM tile.
//  CONTIG_ALONG_K == false: element (k, x) at P[k*ld + x] (x contiguous); LDS
//    tile [k][x] with rows padded by 16 elements, so a fragment read (16 x
//    values per lane group) and a store (consecutive x per lane) both touch
//    distinct banks.
//  CONTIG_ALONG_K == true : element (k, x) at P[x*ld + k] (k contiguous); LDS
//    tile [x][k] with rows of BK + KPAD elements (36 floats / 18 doubles:
//    a row stride of 36 dwords), so the 16 rows of a fragment read and the
//    consecutive-k stores of a lane group fall in distinct banks.
template <typename T> struct KPad;
template <> struct KPad<float> { static constexpr int v = 4; };
template <> struct KPad<double> { static constexpr int v = 2; };

template <int BDIM, int BK, bool CONTIG_ALONG_K, int NT, typename T>
struct Slab {
    static constexpr int ROWS = CONTIG_ALONG_K ? BDIM : BK;
    static constexpr int LDW = CONTIG_ALONG_K ? BK + KPad<T>::v : BDIM + 16;
    static constexpr int ELEMS = (BDIM * BK + NT - 1) / NT;

    __device__ static inline void coords(int e, int& kk, int& xx) {
        if (CONTIG_ALONG_K) { xx = e / BK; kk = e % BK; }
        else                { kk = e / BDIM; xx = e % BDIM; }
    }
    __device__ static inline void load(const double* __restrict__ P, int64_t ld, int64_t x0, int64_t xmax,
                                       int64_t k0, int64_t kmax, T* regs) {
#pragma unroll
        for (int u = 0; u < ELEMS; ++u) {
            const int e = u * NT + (int)threadIdx.x;
            int kk, xx;
            coords(e, kk, xx);
            const int64_t k = k0 + kk, x = x0 + xx;
            double v = 0.0;
            if (e < BDIM * BK && k < kmax && x < xmax) v = CONTIG_ALONG_K ? P[x * ld + k] : P[k * ld + x];
            regs[u] = (T)v;
        }
    }
    __device__ static inline void store(T (*S)[LDW], const T* regs) {
#pragma unroll
        for (int u = 0; u < ELEMS; ++u) {
            const int e = u * NT + (int)threadIdx.x;
            if (BDIM * BK % NT != 0 && e >= BDIM * BK) break;
            int kk, xx;
            coords(e, kk, xx);
            if (CONTIG_ALONG_K) S[xx][kk] = regs[u];
            else S[kk][xx] = regs[u];
        }
    }
    __device__ static inline T frag(const T (*S)[LDW], int k, int x) {
        return CONTIG_ALONG_K ? S[x][k] : S[k][x];
    }
};

template <typename T> struct GemmSlab;
template <> struct GemmSlab<double> { static constexpr int BK = 16; };
template <> struct GemmSlab<float> { static constexpr int BK = 32; };

// WM x WN waves (8 or 10, two or more per SIMD); each wave owns 64 x 32 of C
// (4 x 2 MFMA tiles).  BM = 64 WM, BN = 32 WN.
template <int TA, int TB, typename T, int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(GemmArgs g) {
    constexpr int NT = 64 * WM * WN;
    constexpr int BM = 64 * WM, BN = 32 * WN, BK = GemmSlab<T>::BK;
    typedef Slab<BM, BK, TA == 0, NT, T> SA;   // op(A)(m,k): TA==0 -> A[m*lda+k]; TA==1 -> A[k*lda+m]
    typedef Slab<BN, BK, TB == 1, NT, T> SB;   // op(B)(k,n): TB==0 -> B[k*ldb+n]; TB==1 -> B[n*ldb+k]
    __shared__ T As[SA::ROWS][SA::LDW];
    __shared__ T Bs[SB::ROWS][SB::LDW];
    typedef typename Mfma<T>::acc_t acc_t;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int64_t tm = blockIdx.x % g.tilesM, tn = blockIdx.x / g.tilesM;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const int64_t b = blockIdx.y;
    const double* A = g.A + b * g.sA;
    const double* B = g.B + b * g.sB;
    double* C = g.C + b * g.sC;

    int64_t kEnd = g.K;
    if (g.a_lower) kEnd = (m0 + BM < g.K) ? (m0 + BM) : g.K;
    const int64_t kTiles = (kEnd + BK - 1) / BK;

    acc_t acc[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[SA::ELEMS], rb[SB::ELEMS];
    if (kTiles > 0) {
        SA::load(A, g.lda, m0, g.M, 0, kEnd, ra);
        SB::load(B, g.ldb, n0, g.N, 0, kEnd, rb);
    }
    for (int64_t kt = 0; kt < kTiles; ++kt) {
        __syncthreads();
        SA::store(As, ra);
        SB::store(Bs, rb);
        __syncthreads();
        if (kt + 1 < kTiles) {
            SA::load(A, g.lda, m0, g.M, (kt + 1) * BK, kEnd, ra);
            SB::load(B, g.ldb, n0, g.N, (kt + 1) * BK, kEnd, rb);
        }
#pragma unroll 2
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T fa[4], fb[2];
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[i] = SA::frag(As, kr, wm * 64 + i * 16 + (lane & 15));
#pragma unroll
            for (int j = 0; j < 2; ++j) fb[j] = SB::frag(Bs, kr, wn * 32 + j * 16 + (lane & 15));
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(fa[i], fb[j], acc[i][j]);
        }
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t row = m0 + wm * 64 + i * 16 + AccLayout<acc_t>::row(lane, r);
                if (row < g.M && col < g.N) {
                    double v = g.alpha * (double)acc[i][j][r];
                    if (g.beta != 0.0) v += g.beta * C[row * g.ldc + col];
                    C[row * g.ldc + col] = v;
                }
            }
        }
}

template <int TA, int TB, typename T>
static int launch_t(corr_ctx* ctx, GemmArgs& a, int64_t batch) {
    // Tile shapes: 128 x 160 when N is in (128, 160] (o = 160 of (H2O)32 in one
    // tile), 256 x 64 for N <= 64 or a multiple of 64 that 128 does not divide,
    // otherwise 128 x 128.
    int cfg = 0;
    if (a.N > 128 && a.N <= 160) cfg = 2;
    else if ((a.N <= 64) || (a.N % 128 != 0 && a.N % 64 == 0)) cfg = 1;
    const int BM = cfg == 1 ? 256 : 128, BN = cfg == 0 ? 128 : (cfg == 1 ? 64 : 160);
    const int nt = cfg == 2 ? 640 : 512;
    a.tilesM = cdiv(a.M, BM);
    const int64_t tiles = a.tilesM * cdiv(a.N, BN);
    CORR_REQUIRE(ctx, tiles < (int64_t)1 << 31, "gemm: too many tiles");
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = (batch - b0 < 65535) ? (batch - b0) : 65535;
        GemmArgs s = a;
        s.A += b0 * a.sA; s.B += b0 * a.sB; s.C += b0 * a.sC;
        dim3 grid((unsigned)tiles, (unsigned)nb);
        KernelTimer timer(ctx, "gemm");
        if (cfg == 1)
            hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 4, 2>), grid, dim3(nt), 0, ctx->stream, s);
        else if (cfg == 2)
            hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 5>), grid, dim3(nt), 0, ctx->stream, s);
        else
            hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 4>), grid, dim3(nt), 0, ctx->stream, s);
        CORR_LAUNCHED(ctx, "gemm_kernel");
    }
    return CORR_OK;
}

int gemm(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K,
         double alpha, const double* A, int64_t lda, int64_t sA,
         const double* B, int64_t ldb, int64_t sB, double beta, double* C,
         int64_t ldc, int64_t sC, int64_t batch, int prec, bool a_lower) {
    CORR_REQUIRE(ctx, M >= 0 && N >= 0 && K >= 0 && batch >= 0, "gemm: negative size");
    if (M == 0 || N == 0 || batch == 0) return CORR_OK;
    CORR_REQUIRE(ctx, A && B && C, "gemm: null operand");
    CORR_REQUIRE(ctx, transa == 0 || transa == 1, "gemm: transa must be 0 or 1");
    CORR_REQUIRE(ctx, transb == 0 || transb == 1, "gemm: transb must be 0 or 1");
    CORR_REQUIRE(ctx, lda >= (transa ? M : K) && ldb >= (transb ? K : N) && ldc >= N,
                 "gemm: leading dimension too small");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32 || prec == CORR_PREC_BF16X3,
                 "gemm: unknown precision");
    GemmArgs a{M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, alpha, beta, 0, a_lower ? 1 : 0};
    const bool f64 = (prec == CORR_PREC_FP64);
    const int sel = transa * 2 + transb;
    if (f64) {
        switch (sel) {
            case 0: return launch_t<0, 0, double>(ctx, a, batch);
            case 1: return launch_t<0, 1, double>(ctx, a, batch);
            case 2: return launch_t<1, 0, double>(ctx, a, batch);
            default: return launch_t<1, 1, double>(ctx, a, batch);
        }
    }
    switch (sel) {
        case 0: return launch_t<0, 0, float>(ctx, a, batch);
        case 1: return launch_t<0, 1, float>(ctx, a, batch);
        case 2: return launch_t<1, 0, float>(ctx, a, batch);
        default: return launch_t<1, 1, float>(ctx, a, batch);
    }
}

// Deterministic reduction: one workgroup, fixed strided order, fixed tree.
__global__ __launch_bounds__(256) void reduce_kernel(const double* __restrict__ d, int64_t n,
                                                     int nser, double* __restrict__ out) {
    __shared__ double sh[256];
    for (int s = 0; s < nser; ++s) {
        double acc = 0.0;
        for (int64_t i = threadIdx.x; i < n; i += 256) acc += d[i * nser + s];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[s] = sh[0];
        __syncthreads();
    }
}

int reduce_sum(corr_ctx* ctx, const double* d_data, int64_t n, int nser, double* h_out) {
    double* d_out = nullptr;
    CORR_TRY(scratch(ctx, "reduce_out", sizeof(double) * 8, (void**)&d_out));
    CORR_REQUIRE(ctx, nser >= 1 && nser <= 8, "reduce_sum: 1..8 series");
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(256), 0, ctx->stream, d_data, n, nser, d_out);
    CORR_LAUNCHED(ctx, "reduce_kernel");
    CORR_HIP(ctx, hipMemcpyAsync(h_out, d_out, sizeof(double) * nser, hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

}  // namespace corr

extern "C" int corr_dgemm(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N,
                          int64_t K, double alpha, const double* d_A, int64_t lda,
                          int64_t strideA, const double* d_B, int64_t ldb, int64_t strideB,
                          double beta, double* d_C, int64_t ldc, int64_t strideC,
                          int64_t batch, int prec) {
    CORR_CHECK_CTX(ctx);
    return corr::gemm(ctx, transa, transb, M, N, K, alpha, d_A, lda, strideA, d_B, ldb,
                      strideB, beta, d_C, ldc, strideC, batch, prec);
}
