// Generic strided-batched GEMM on CDNA4 matrix cores (row-major operands).
//
//   C[b] = alpha * op(A[b]) op(B[b]) + beta * C[b]
//
// FP64 path: v_mfma_f64_16x16x4_f64, FP64 accumulate.
// FP32 path: operands rounded to FP32 when staged into LDS,
//            v_mfma_f32_16x16x4_f32, FP32 accumulate, epilogue in FP64.
//
// A workgroup is 8, 10 or 12 wave64s; each wave owns a 64x32 block of C built from
// 4x2 MFMA 16x16 tiles.  A K-slab is BK deep (32 for FP32, 16 for FP64); the
// four k-rows of one MFMA step are spread over the slab as k = g * BK/4 + t
// (lane group g = lane / 16, step t), so that a lane's k values within a slab
// are contiguous.  An operand that is contiguous along K is staged as [x][k]
// rows and read as two ds_read_b128 per fragment row; an operand contiguous
// along x is staged as [k][x] and read as scalars.  Global loads move pairs of
// doubles; LDS holds two slabs, so each slab costs one barrier.
//
// FP64 workgroups whose tile lies inside C, on 16-byte aligned operands,
// stage with global_load_lds_dwordx4 instead (no VGPR round trip, no
// ds_write), into the same LDS layouts, three slabs deep with one barrier per
// slab; a K remainder below one slab goes through the register path.
#include "common.hpp"

#include <type_traits>

namespace corr {

struct GemmArgs {
    int64_t M, N, K;
    const double* A; int64_t lda, sA;
    const double* B; int64_t ldb, sB;
    double* C; int64_t ldc, sC;
    double alpha, beta;
    int64_t tilesM;
    int tri;          // GEMM_TRI_* flags
    int64_t ksplit;   // > 0: blockIdx.y is a K chunk of this depth (C = partial slice), not a batch index
    int dma;          // FP64 operands 16-byte aligned (rows, batch and K-chunk offsets): global_load_lds staging
    int epi;          // output mapping (GEMM_EPI_*), beta = 0 unless PLAIN
    int64_t en;       // index extent of the PAIRS / MIRROR mappings
};

// Output mappings of element (row, col) of op(A) op(B), with n = en and
// P = n (n + 1) / 2 (used by the 4-index transform):
//  PLAIN : C[row * ldc + col]
//  PAIRS : row = X n + i, col = j: C[X P + tri(i, j)] for j >= i only, tri(i, j)
//          = i n - i (i - 1) / 2 + (j - i) (the packed upper pairs of (i, j))
//  MIRROR: row = q n + k with q = tri(i, j), col = l: C[((i n + j) n + k) n + l]
//          and, for i != j, C[((j n + i) n + k) n + l]
enum { GEMM_EPI_PLAIN = 0, GEMM_EPI_PAIRS = 1, GEMM_EPI_MIRROR = 2 };

// (i, j), i <= j < n, of the packed pair index q = tri(i, j).
__device__ inline void gemm_pair_decode(int64_t q, int64_t n, int64_t& i, int64_t& j) {
    const double b = 2.0 * (double)n + 1.0;
    int64_t r = (int64_t)((b - sqrt(b * b - 8.0 * (double)q)) * 0.5);
    if (r < 0) r = 0;
    if (r > n - 1) r = n - 1;
    auto start = [n](int64_t x) { return x * n - x * (x - 1) / 2; };
    while (r > 0 && start(r) > q) --r;
    while (r + 1 < n && start(r + 1) <= q) ++r;
    i = r;
    j = r + (q - start(r));
}

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

typedef double corr_d2 __attribute__((ext_vector_type(2)));

// Slab geometry: BK elements of k per slab, KV = BK/4 per lane group, CH
// elements per 16-byte LDS chunk; XPAD pads [k][x] rows to a stride of
// 2 (FP32) or 4 (FP64) modulo 8 elements (4 lane groups in distinct banks).
template <typename T> struct GemmSlab;
template <> struct GemmSlab<double> {
    static constexpr int BK = 16, KV = 4, CH = 2, XPAD = 4;
};
template <> struct GemmSlab<float> {
    static constexpr int BK = 32, KV = 8, CH = 4, XPAD = 2;
};

// Slot of 16-byte chunk ch in a 128-byte [x][k] row x (an involution in ch):
// keeps the ds_read_b128 fragment reads of 16 consecutive rows conflict-free.
__device__ inline int gemm_chunk(int x, int ch) {
    return ch ^ (((x >> 1) & 1) | (((x >> 3) & 1) << 2));
}

// One operand slab (BK deep, BDIM wide) of a GEMM tile, staged as element pairs.
//  KC == true : element (k, x) at P[x*ld + k]; LDS [BDIM][BK], chunk-swizzled rows.
//  KC == false: element (k, x) at P[k*ld + x]; LDS [BK][BDIM + XPAD].
template <int BDIM, bool KC, int NT, typename T>
struct Opnd {
    typedef GemmSlab<T> G;
    static constexpr int BK = G::BK, KV = G::KV;
    static constexpr int LDW = KC ? BK : BDIM + G::XPAD;
    static constexpr int SIZE = KC ? BDIM * BK : BK * LDW;   // elements of one LDS slab
    static constexpr int PAIRS = BDIM * BK / 2;
    static constexpr int PPT = (PAIRS + NT - 1) / NT;

    __device__ static inline void coords(int e, int& kk, int& xx) {
        if (KC) { xx = e / (BK / 2); kk = (e % (BK / 2)) * 2; }
        else    { kk = e / (BDIM / 2); xx = (e % (BDIM / 2)) * 2; }
    }
    __device__ static inline void load(const double* __restrict__ P, int64_t ld, int64_t x0, int64_t xmax,
                                       int64_t k0, int64_t kmax, corr_d2* regs) {
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const int e = u * NT + (int)threadIdx.x;
            int kk, xx;
            coords(e, kk, xx);
            const int64_t k = k0 + kk, x = x0 + xx;
            corr_d2 v = {0.0, 0.0};
            if (PAIRS % NT == 0 || e < PAIRS) {
                // The pair runs along the contiguous dimension: (k, k+1) or (x, x+1).
                const int64_t i0 = KC ? x * ld + k : k * ld + x;
                const bool in0 = k < kmax && x < xmax;
                const bool in1 = KC ? (k + 1 < kmax && x < xmax) : (k < kmax && x + 1 < xmax);
                if (in0 && in1 && ((reinterpret_cast<uintptr_t>(P + i0) & 15) == 0)) {
                    v = *reinterpret_cast<const corr_d2*>(P + i0);
                } else {
                    if (in0) v.x = P[i0];
                    if (in1) v.y = P[i0 + 1];
                }
            }
            regs[u] = v;
        }
    }
    __device__ static inline void store(T* S, const corr_d2* regs) {
#pragma unroll
        for (int u = 0; u < PPT; ++u) {
            const int e = u * NT + (int)threadIdx.x;
            if (PAIRS % NT != 0 && e >= PAIRS) break;
            int kk, xx;
            coords(e, kk, xx);
            const int off = KC ? xx * BK + gemm_chunk(xx, kk / G::CH) * G::CH + kk % G::CH : kk * LDW + xx;
            if (sizeof(T) == 4) {
                float2 f = make_float2((float)regs[u].x, (float)regs[u].y);
                *reinterpret_cast<float2*>(S + off) = f;
            } else {
                *reinterpret_cast<corr_d2*>(S + off) = regs[u];
            }
        }
    }
    // Values h*PER .. h*PER+PER-1 of the lane group's KV k positions (slab
    // positions g*KV + ...) for row/column x; PER = elements per 16 bytes.
    static constexpr int PER = 16 / sizeof(T);
    __device__ static inline void frag(const T* S, int g, int h, int x, T (&f)[PER]) {
        if (KC) {
            typedef T vec_t __attribute__((ext_vector_type(PER)));
            const vec_t q = *reinterpret_cast<const vec_t*>(S + x * BK + gemm_chunk(x, g * (KV / PER) + h) * PER);
#pragma unroll
            for (int e = 0; e < PER; ++e) f[e] = q[e];
        } else {
#pragma unroll
            for (int e = 0; e < PER; ++e) f[e] = S[(g * KV + h * PER + e) * LDW + x];
        }
    }
};

typedef __attribute__((address_space(3))) void* gemm_lds_ptr;
typedef __attribute__((address_space(1))) const void* gemm_gbl_ptr;

// global_load_lds staging of one FP64 operand slab into the register path's
// LDS layout.  Instruction i of the operand's NI moves up to 1 KiB; lane l
// writes bytes [16 l, 16 l + 16) from the instruction's LDS base.
//  KC:  8 rows of [x][16] (128 B each) per instruction; lane l fills row
//       8 i + l / 8, chunk slot l % 8, which holds logical chunk
//       gemm_chunk(row, l % 8) (the swizzle is an involution).
//  !KC: row k of [16][BDIM + XPAD] takes NPR instructions of up to 128
//       doubles; lanes past the row's end are masked off.
template <int BDIM, bool KC>
struct DmaOpnd {
    static constexpr int BK = 16, LDW = KC ? BK : BDIM + 4;
    static constexpr int NPR = (BDIM + 127) / 128;
    static constexpr int NI = KC ? BDIM / 8 : BK * NPR;
    __device__ static inline void copy(double* S, const double* __restrict__ P, int64_t ld, int64_t x0, int64_t k0,
                                       int i, int lane) {
        if (KC) {
            const int r = 8 * i + (lane >> 3);
            const int ch = gemm_chunk(r, lane & 7);
            const double* src = P + (x0 + r) * ld + k0 + 2 * ch;
            __builtin_amdgcn_global_load_lds((gemm_gbl_ptr)src, (gemm_lds_ptr)(S + 8 * i * BK), 16, 0, 0);
        } else {
            const int k = i / NPR, j = i % NPR;
            const int x = 128 * j + 2 * lane;
            if (x < BDIM) {
                const double* src = P + (k0 + k) * ld + x0 + x;
                __builtin_amdgcn_global_load_lds((gemm_gbl_ptr)src, (gemm_lds_ptr)(S + k * LDW + 128 * j), 16, 0, 0);
            }
        }
    }
};

// WM x WN waves (8, 10 or 12; two or more per SIMD); each wave owns BI x BJ
// MFMA 16x16 tiles of C (16 BI x 16 BJ).  BM = 16 BI WM, BN = 16 BJ WN.
template <int TA, int TB, typename T, int WM, int WN, int BI = 4, int BJ = 2>
struct GemmTile {
    static constexpr int NT = 64 * WM * WN;
    static constexpr int BM = 16 * BI * WM, BN = 16 * BJ * WN, BK = GemmSlab<T>::BK, KV = GemmSlab<T>::KV;
    typedef Opnd<BM, TA == 0, NT, T> OA;   // op(A)(m,k): TA==0 -> A[m*lda+k]; TA==1 -> A[k*lda+m]
    typedef Opnd<BN, TB == 1, NT, T> OB;   // op(B)(k,n): TB==0 -> B[k*ldb+n]; TB==1 -> B[n*ldb+k]
    typedef typename Mfma<T>::acc_t acc_t;

    // MFMAs of one staged slab.
    __device__ static inline void compute(const T* As, const T* Bs, int wm, int wn, int lg, int lc,
                                          acc_t (&acc)[BI][BJ]) {
        constexpr int PER = OA::PER;
#pragma unroll
        for (int h = 0; h < KV / PER; ++h) {
            T fa[BI][PER], fb[BJ][PER];
#pragma unroll
            for (int i = 0; i < BI; ++i) OA::frag(As, lg, h, wm * 16 * BI + i * 16 + lc, fa[i]);
#pragma unroll
            for (int j = 0; j < BJ; ++j) OB::frag(Bs, lg, h, wn * 16 * BJ + j * 16 + lc, fb[j]);
#pragma unroll
            for (int t = 0; t < PER; ++t)
#pragma unroll
                for (int i = 0; i < BI; ++i)
#pragma unroll
                    for (int j = 0; j < BJ; ++j) acc[i][j] = Mfma<T>::mma(fa[i][t], fb[j][t], acc[i][j]);
        }
    }

    // Register-path slabs [kt0, kTiles) through the double buffer (As0/Bs0,
    // As1/Bs1): global -> registers -> ds_write, loads two slabs ahead.
    __device__ static inline void regs_loop(const GemmArgs& g, const double* A, const double* B, int64_t m0,
                                            int64_t n0, int64_t kEnd, int64_t kt0, int64_t kTiles, T* As0, T* As1,
                                            T* Bs0, T* Bs1, int wm, int wn, int lg, int lc, acc_t (&acc)[BI][BJ]) {
        corr_d2 ra[OA::PPT], rb[OB::PPT];
        if (kt0 < kTiles) {
            OA::load(A, g.lda, m0, g.M, kt0 * BK, kEnd, ra);
            OB::load(B, g.ldb, n0, g.N, kt0 * BK, kEnd, rb);
            OA::store(As0, ra);
            OB::store(Bs0, rb);
            if (kt0 + 1 < kTiles) {
                OA::load(A, g.lda, m0, g.M, (kt0 + 1) * BK, kEnd, ra);
                OB::load(B, g.ldb, n0, g.N, (kt0 + 1) * BK, kEnd, rb);
            }
        }
        __syncthreads();
        for (int64_t kt = kt0; kt < kTiles; ++kt) {
            const bool odd = ((kt - kt0) & 1) != 0;
            if (kt + 1 < kTiles) {
                OA::store(odd ? As0 : As1, ra);
                OB::store(odd ? Bs0 : Bs1, rb);
                if (kt + 2 < kTiles) {
                    OA::load(A, g.lda, m0, g.M, (kt + 2) * BK, kEnd, ra);
                    OB::load(B, g.ldb, n0, g.N, (kt + 2) * BK, kEnd, rb);
                }
            }
            compute(odd ? As1 : As0, odd ? Bs1 : Bs0, wm, wn, lg, lc, acc);
            __syncthreads();
        }
    }
};

template <int TA, int TB, typename T, int WM, int WN, int EPI, int BI = 4, int BJ = 2>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(GemmArgs g) {
    typedef GemmTile<TA, TB, T, WM, WN, BI, BJ> GT;
    constexpr int BM = GT::BM, BN = GT::BN, BK = GT::BK;
    typedef typename GT::OA OA;
    typedef typename GT::OB OB;
    typedef typename GT::acc_t acc_t;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN, lg = lane >> 4, lc = lane & 15;
    const int64_t tm = blockIdx.x % g.tilesM, tn = blockIdx.x / g.tilesM;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const int64_t b = blockIdx.y;
    const double* A;
    const double* B;
    int64_t kEnd = g.K;
    if (g.ksplit > 0) {
        const int64_t k0 = b * g.ksplit;
        A = g.A + (TA ? k0 * g.lda : k0);
        B = g.B + (TB ? k0 : k0 * g.ldb);
        kEnd = (g.K - k0 < g.ksplit) ? (g.K - k0) : g.ksplit;
    } else {
        A = g.A + b * g.sA;
        B = g.B + b * g.sB;
    }
    double* C = g.C + b * g.sC;
    if (g.tri & GEMM_TRI_C_LOWER) {
        if (n0 >= m0 + BM) return;   // the tile lies above the diagonal
    }
    if (g.tri & GEMM_TRI_A_LOWER) kEnd = (m0 + BM < kEnd) ? (m0 + BM) : kEnd;
    if (g.tri & GEMM_TRI_B_LOWER) {
        // rows k < n0 of op(B) are zero: start the K range at n0
        const int64_t kb = n0 < kEnd ? n0 : kEnd;
        A += TA ? kb * g.lda : kb;
        B += TB ? kb : kb * g.ldb;
        kEnd -= kb;
    }
    const int64_t kTiles = (kEnd + BK - 1) / BK;

    acc_t acc[BI][BJ];
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < BJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    if constexpr (std::is_same<T, double>::value) {
        // Three slab buffers [A slab | B slab] (the register path uses the first two).
        typedef DmaOpnd<BM, TA == 0> DA;
        typedef DmaOpnd<BN, TB == 1> DB;
        constexpr int SLAB = OA::SIZE + OB::SIZE;
        constexpr int NW = WM * WN, NI = DA::NI + DB::NI, CPW = (NI + NW - 1) / NW;
        static_assert(CPW <= 15, "gemm: copies per wave exceed the vmcnt field used");
        static_assert(DA::LDW == OA::LDW && DB::LDW == OB::LDW && DA::BK == BK,
                      "gemm: staged and register layouts must agree");
        __shared__ __attribute__((aligned(16))) double S[3][SLAB];
        const int64_t kFull = kEnd / BK;
        if (g.dma && m0 + BM <= g.M && n0 + BN <= g.N && kFull > 0) {
            // Wave w issues copies w*CPW .. w*CPW+CPW-1 of each slab (the last
            // ones repeat copy NI-1 so every wave issues CPW).  Slab s lives in
            // buffer s % 3; copies run two slabs ahead, and the barrier before
            // slab s+1 waits until at most CPW copies (slab s+2) are in flight.
            auto copy_slab = [&](int64_t st, int buf) {
                double* As = S[buf];
                double* Bs = S[buf] + OA::SIZE;
#pragma unroll
                for (int c = 0; c < CPW; ++c) {
                    int q = wave * CPW + c;
                    q = q < NI ? q : NI - 1;
                    if (q < DA::NI) DA::copy(As, A, g.lda, m0, st * BK, q, lane);
                    else DB::copy(Bs, B, g.ldb, n0, st * BK, q - DA::NI, lane);
                }
            };
            copy_slab(0, 0);
            copy_slab(kFull > 1 ? 1 : 0, 1);
            __builtin_amdgcn_s_waitcnt(CPW | (0x7 << 4));   // slab 0 landed (vmcnt CPW, lgkmcnt 0)
            __builtin_amdgcn_s_barrier();
            __asm__ volatile("" ::: "memory");
            int buf = 0;
            for (int64_t kt = 0; kt < kFull; ++kt) {
                const int nb2 = buf >= 1 ? buf - 1 : 2;                 // buffer of slab kt+2 (= kt-1's)
                copy_slab(kt + 2 < kFull ? kt + 2 : kFull - 1, nb2);
                GT::compute(S[buf], S[buf] + OA::SIZE, wm, wn, lg, lc, acc);
                __builtin_amdgcn_s_waitcnt(CPW | (0x7 << 4));
                __asm__ volatile("" ::: "memory");
                __builtin_amdgcn_s_barrier();
                __asm__ volatile("" ::: "memory");
                buf = buf == 2 ? 0 : buf + 1;
            }
            __builtin_amdgcn_s_waitcnt((0x7 << 4) | (0xF << 8));  // vmcnt 0: no copy outlives the loop
            __syncthreads();
            if (kFull < kTiles)   // the K remainder, through registers
                GT::regs_loop(g, A, B, m0, n0, kEnd, kFull, kTiles, S[0], S[1], S[0] + OA::SIZE, S[1] + OA::SIZE,
                              wm, wn, lg, lc, acc);
        } else {
            GT::regs_loop(g, A, B, m0, n0, kEnd, 0, kTiles, S[0], S[1], S[0] + OA::SIZE, S[1] + OA::SIZE, wm, wn,
                          lg, lc, acc);
        }
    } else {
        __shared__ __attribute__((aligned(16))) T As[2][OA::SIZE];
        __shared__ __attribute__((aligned(16))) T Bs[2][OB::SIZE];
        GT::regs_loop(g, A, B, m0, n0, kEnd, 0, kTiles, As[0], As[1], Bs[0], Bs[1], wm, wn, lg, lc, acc);
    }

    // Mapped outputs: each of this lane's 16 rows (row = X n + ri, ri < n)
    // gets its destination base once, from the tile's first row and a
    // bounded subtraction (no division per element).
    const int64_t x0 = EPI == GEMM_EPI_PLAIN ? 0 : m0 / g.en, r0 = EPI == GEMM_EPI_PLAIN ? 0 : m0 % g.en;
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int off = wm * 16 * BI + i * 16 + AccLayout<acc_t>::row(lane, r);
            const int64_t row = m0 + off;
            if (row >= g.M) continue;
            int64_t d0 = 0, d1 = -1, ri = 0;
            if constexpr (EPI != GEMM_EPI_PLAIN) {
                const int64_t n = g.en;
                int64_t x = x0;
                ri = r0 + off;
                while (ri >= n) { ri -= n; ++x; }
                if constexpr (EPI == GEMM_EPI_PAIRS) {
                    // row (x, i = ri), column j -> x P + i n - i (i - 1) / 2 + j - i
                    d0 = x * (n * (n + 1) / 2) + ri * n - ri * (ri - 1) / 2 - ri;
                } else {
                    // row (q = x, k = ri), column l -> ((pi n + pj) n + k) n + l and (pj pi k l)
                    int64_t pi, pj;
                    gemm_pair_decode(x, n, pi, pj);
                    d0 = ((pi * n + pj) * n + ri) * n;
                    if (pi != pj) d1 = ((pj * n + pi) * n + ri) * n;
                }
            }
#pragma unroll
            for (int j = 0; j < BJ; ++j) {
                const int64_t col = n0 + wn * 16 * BJ + j * 16 + lc;
                if (col >= g.N) continue;
                double v = g.alpha * (double)acc[i][j][r];
                if constexpr (EPI == GEMM_EPI_PLAIN) {
                    if (g.beta != 0.0) v += g.beta * C[row * g.ldc + col];
                    C[row * g.ldc + col] = v;
                } else if constexpr (EPI == GEMM_EPI_PAIRS) {
                    if (col >= ri) C[d0 + col] = v;
                } else {
                    C[d0 + col] = v;
                    if (d1 >= 0) C[d1 + col] = v;
                }
            }
        }
}

#ifndef CORR_GEMM_FORCE
#define CORR_GEMM_FORCE (-1)
#endif

template <int TA, int TB, typename T, int EPI = GEMM_EPI_PLAIN>
static int launch_t(corr_ctx* ctx, GemmArgs& a, int64_t batch) {
    // Tile shapes: 0: 128 x 128 (8 waves of 4 x 2 MFMA tiles), 1: 256 x 64 (8 x 4 x 2),
    // 2: 128 x 160 (10 x 4 x 2), 3: 128 x 192 (12 x 4 x 2), 4: 128 x 160 (8 waves of
    // 2 x 5), 5: 128 x 256 (8 x 4 x 4), 6: 256 x 128 (8 x 4 x 4), 7: 128 x 192 (8 x 2 x 6).
    const int BMs[8] = {128, 256, 128, 128, 128, 128, 256, 128}, BNs[8] = {128, 64, 160, 192, 160, 256, 128, 192};
    int cfg = 0;
    int64_t best = -1;
    // The shape with the least weighted padded area of C wins; ties keep the
    // order below.  Padded areas are weighted by the measured time per useful
    // FLOP on MI355X at equal padding (FP64 and FP32, 6080^3 .. 23104^2 x 1600):
    // 128 x 192 runs 12 waves, 3 per SIMD (1.00); 128 x 128 (1.06) and
    // 256 x 64 (1.11) run 2 per SIMD; 128 x 160 on 10 waves runs (3, 3, 2, 2)
    // waves per SIMD and idles matrix cores (1.20).  FP64 only, from the
    // B-formation GEMMs of (H2O)32: 128 x 160 on 8 waves of 2 x 5 (1.03;
    // 7.6 ms against 9.5 ms for 128 x 192 at N = o = 160) and 256 x 128 on 8
    // waves of 4 x 4 (1.01; 6.2 against 6.6 ms for 128 x 128 at
    // 389120 x 608 x 768); 128 x 256 of 4 x 4 is not faster than 128 x 192
    // and stays out.  For a single column of 192-wide tiles (FP64), 128 x 192
    // on 8 waves of 2 x 6 (0.96; the (H2O)8 4-index quarters, N^3 x 192 x 192:
    // MOtrans 29.3 against 30.4 ms).  Outputs with fewer than 512 tiles of 128 x 128 keep
    // unweighted areas: there the tile count (filling the 256 CUs) matters
    // more than per-tile speed.
    const bool big = cdiv(a.M, 128) * cdiv(a.N, 128) * batch >= 512;
    const bool f64 = std::is_same<T, double>::value;
    const int64_t weight[8] = {big ? 106 : 100, big ? 111 : 100, big ? 120 : 100, 100,
                               f64 ? (big ? 103 : 100) : 9999, 9999, f64 ? (big ? 101 : 100) : 9999,
                               f64 && big && a.N <= 192 ? 96 : 9999};
    for (int c : {0, 3, 2, 1, 4, 5, 6, 7}) {
        const int64_t area = cdiv(a.M, BMs[c]) * BMs[c] * (cdiv(a.N, BNs[c]) * BNs[c]) / 100 * weight[c];
        if (best < 0 || area < best) { best = area; cfg = c; }
    }
    if (CORR_GEMM_FORCE >= 0) cfg = CORR_GEMM_FORCE;
    const int BM = BMs[cfg], BN = BNs[cfg];
    const int nt = cfg == 3 ? 768 : (cfg == 2 ? 640 : 512);
    a.tilesM = cdiv(a.M, BM);
    const int64_t tiles = a.tilesM * cdiv(a.N, BN);
    CORR_REQUIRE(ctx, tiles < (int64_t)1 << 31, "gemm: too many tiles");
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = (batch - b0 < 65535) ? (batch - b0) : 65535;
        GemmArgs s = a;
        if (a.ksplit > 0) {
            s.A = a.A; s.B = a.B;   // chunk offsets are applied in the kernel from blockIdx.y
            CORR_REQUIRE(ctx, batch <= 65535, "gemm: too many K chunks");
        } else {
            s.A += b0 * a.sA; s.B += b0 * a.sB;
        }
        s.C += b0 * a.sC;
        dim3 grid((unsigned)tiles, (unsigned)nb);
        KernelTimer timer(ctx, "gemm");
        switch (cfg) {
            case 1: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 4, 2, EPI>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 2: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 5, EPI>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 3: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 6, EPI>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 4: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 4, 2, EPI, 2, 5>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 5: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 4, EPI, 4, 4>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 6: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 4, 2, EPI, 4, 4>), grid, dim3(nt), 0, ctx->stream, s); break;
            case 7: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 4, 2, EPI, 2, 6>), grid, dim3(nt), 0, ctx->stream, s); break;
            default: hipLaunchKernelGGL((gemm_kernel<TA, TB, T, 2, 4, EPI>), grid, dim3(nt), 0, ctx->stream, s); break;
        }
        CORR_LAUNCHED(ctx, "gemm_kernel");
    }
    return CORR_OK;
}

// C = alpha * sum_s P[s] + beta * C, chunks summed in order s = 0, 1, ...
__global__ void splitk_reduce_kernel(const double* __restrict__ P, int64_t S, int64_t M, int64_t N,
                                     double alpha, double beta, double* __restrict__ C, int64_t ldc) {
    const int64_t MN = M * N;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < MN; e += (int64_t)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int64_t c = 0; c < S; ++c) s += P[c * MN + e];
        const int64_t m = e / N, n = e % N;
        double v = alpha * s;
        if (beta != 0.0) v += beta * C[m * ldc + n];
        C[m * ldc + n] = v;
    }
}

static int dispatch(corr_ctx* ctx, int transa, int transb, int f64, GemmArgs& a, int64_t batch) {
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

// Matrix-vector products (M = 1 or N = 1; memory-bound, FP64 on the vector
// units whatever prec asks).  Output t of batch b: y[t] = alpha sum_k
// W[t st + k sk] x[k sx] + beta y[t], with W, x, y offset by the batch strides.
// sk == 1: one wave per output, lanes stride k, fixed xor-shuffle sum; else
// one thread per output, k in order.  Both deterministic.
struct GemvArgs {
    int64_t T, K;
    const double* W; int64_t st, sk, sW;
    const double* x; int64_t sx, sX;
    double* y; int64_t sy, sY;
    double alpha, beta;
};
__global__ __launch_bounds__(256) void gemv_rows_kernel(GemvArgs g) {
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6), b = blockIdx.y;
    if (t >= g.T) return;
    const double* W = g.W + b * g.sW + t * g.st;
    const double* x = g.x + b * g.sX;
    double s = 0.0;
    for (int64_t k = lane; k < g.K; k += 64) s += W[k] * x[k * g.sx];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        double* y = g.y + b * g.sY + t * g.sy;
        *y = g.alpha * s + (g.beta != 0.0 ? g.beta * *y : 0.0);
    }
}
__global__ __launch_bounds__(256) void gemv_cols_kernel(GemvArgs g) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (t >= g.T) return;
    const double* W = g.W + b * g.sW + t * g.st;
    const double* x = g.x + b * g.sX;
    double s = 0.0;
    for (int64_t k = 0; k < g.K; ++k) s += W[k * g.sk] * x[k * g.sx];
    double* y = g.y + b * g.sY + t * g.sy;
    *y = g.alpha * s + (g.beta != 0.0 ? g.beta * *y : 0.0);
}

static int gemv(corr_ctx* ctx, const GemvArgs& a, int64_t batch) {
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
        const int64_t nb = std::min<int64_t>(65535, batch - b0);
        GemvArgs s = a;
        s.W += b0 * a.sW;
        s.x += b0 * a.sX;
        s.y += b0 * a.sY;
        if (a.sk == 1) {
            CORR_REQUIRE(ctx, cdiv(a.T, 4) < ((int64_t)1 << 31), "gemv: too many rows");
            hipLaunchKernelGGL(gemv_rows_kernel, dim3((unsigned)cdiv(a.T, 4), (unsigned)nb), dim3(256), 0, ctx->stream,
                               s);
            CORR_LAUNCHED(ctx, "gemv_rows_kernel");
        } else {
            CORR_REQUIRE(ctx, cdiv(a.T, 256) < ((int64_t)1 << 31), "gemv: too many columns");
            hipLaunchKernelGGL(gemv_cols_kernel, dim3((unsigned)cdiv(a.T, 256), (unsigned)nb), dim3(256), 0,
                               ctx->stream, s);
            CORR_LAUNCHED(ctx, "gemv_cols_kernel");
        }
    }
    return CORR_OK;
}

// FP64 operands whose slab rows, batch strides and K-chunk offsets all start
// on 16-byte boundaries can be staged by global_load_lds_dwordx4.
static int dma_ok(int f64, const double* A, int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB) {
    const auto al16 = [](const double* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    return f64 && al16(A) && al16(B) && lda % 2 == 0 && ldb % 2 == 0 && sA % 2 == 0 && sB % 2 == 0 ? 1 : 0;
}

int gemm(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K,
         double alpha, const double* A, int64_t lda, int64_t sA,
         const double* B, int64_t ldb, int64_t sB, double beta, double* C,
         int64_t ldc, int64_t sC, int64_t batch, int prec, int tri) {
    CORR_REQUIRE(ctx, M >= 0 && N >= 0 && K >= 0 && batch >= 0, "gemm: negative size");
    if (M == 0 || N == 0 || batch == 0) return CORR_OK;
    CORR_REQUIRE(ctx, A && B && C, "gemm: null operand");
    CORR_REQUIRE(ctx, transa == 0 || transa == 1, "gemm: transa must be 0 or 1");
    CORR_REQUIRE(ctx, transb == 0 || transb == 1, "gemm: transb must be 0 or 1");
    CORR_REQUIRE(ctx, lda >= (transa ? M : K) && ldb >= (transb ? K : N) && ldc >= N,
                 "gemm: leading dimension too small");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32 || prec == CORR_PREC_BF16X3 ||
                          prec == CORR_PREC_INT8X5,
                 "gemm: unknown precision");
    if (prec == CORR_PREC_INT8X5) {
        CORR_REQUIRE(ctx, batch == 1, "gemm int8x5: batch 1 only");
        CORR_REQUIRE(ctx, (tri & ~GEMM_TRI_A_LOWER) == 0, "gemm int8x5: only a lower-triangular A");
        return gemm_ozaki(ctx, transa, transb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri != 0);
    }
    const int f64 = (prec == CORR_PREC_FP64) ? 1 : 0;
    if ((N == 1 || M == 1) && tri == 0 && K > 0) {
        GemvArgs g;
        if (N == 1) {   // y[m] = sum_k op(A)(m, k) op(B)(k, 0), y at C[m ldc]
            g = GemvArgs{M, K, A, transa ? 1 : lda, transa ? lda : 1, sA, B, transb ? 1 : ldb, sB, C, ldc, sC,
                         alpha, beta};
        } else {        // y[n] = sum_k op(A)(0, k) op(B)(k, n), y at C[n]
            g = GemvArgs{N, K, B, transb ? ldb : 1, transb ? 1 : ldb, sB, A, transa ? lda : 1, sA, C, 1, sC,
                         alpha, beta};
        }
        return gemv(ctx, g, batch);
    }
    // Split K when the output has fewer tiles than the 256 CUs and K is deep
    // (matrix-vector products included): chunks of depth kc >= 512 write
    // partial products to a workspace, then a fixed-order sum applies alpha
    // and beta (deterministic).
    const int64_t tiles_est = cdiv(M, 128) * cdiv(N, 128) * batch;
    if (batch == 1 && tri == 0 && tiles_est < 256 && K >= 2048) {
        int64_t S = cdiv(512, tiles_est);
        if (S > K / 512) S = K / 512;
        const int64_t cap = (int64_t)(256 << 20) / (8 * M * N);   // workspace at most 256 MiB
        if (S > cap) S = cap;
        if (S >= 2) {
            const int64_t kc = rup(cdiv(K, S), 32);
            S = cdiv(K, kc);
            double* P = nullptr;
            // one workspace per stream: the auxiliary stream may run GEMMs concurrently
            const char* ws = (ctx->aux_stream && ctx->stream == ctx->aux_stream) ? "gemm_splitk_aux" : "gemm_splitk";
            CORR_TRY(scratch(ctx, ws, sizeof(double) * S * M * N, (void**)&P));
            GemmArgs a{M, N, K, A, lda, 0, B, ldb, 0, P, N, M * N, 1.0, 0.0, 0, 0, kc};
            a.dma = dma_ok(f64, A, lda, 0, B, ldb, 0) && kc % 2 == 0;
            a.epi = GEMM_EPI_PLAIN;
            a.en = 0;
            CORR_TRY(dispatch(ctx, transa, transb, f64, a, S));
            const int64_t grid = std::min<int64_t>(cdiv(M * N, 256), 4096);
            hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)grid), dim3(256), 0, ctx->stream, P, S, M, N,
                               alpha, beta, C, ldc);
            CORR_LAUNCHED(ctx, "splitk_reduce_kernel");
            return CORR_OK;
        }
    }
    GemmArgs a{M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, alpha, beta, 0, tri, 0};
    a.dma = dma_ok(f64, A, lda, sA, B, ldb, sB);
    a.epi = GEMM_EPI_PLAIN;
    a.en = 0;
    return dispatch(ctx, transa, transb, f64, a, batch);
}

int gemm_mapped(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K, const double* A,
                int64_t lda, const double* B, int64_t ldb, double* C, int mapping, int64_t n, int prec) {
    CORR_REQUIRE(ctx, M > 0 && N > 0 && K > 0 && A && B && C && n > 0, "gemm_mapped: bad arguments");
    CORR_REQUIRE(ctx, (mapping == 1 || mapping == 2) && M % n == 0 && (mapping == 2 || N == n),
                 "gemm_mapped: rows must be whole blocks of n (and N = n for PAIRS)");
    CORR_REQUIRE(ctx, lda >= (transa ? M : K) && ldb >= (transb ? K : N), "gemm_mapped: leading dimension too small");
    CORR_REQUIRE(ctx, transa == 1 && transb == 1, "gemm_mapped: only op(A) = A^T, op(B) = B^T");
    const int f64 = (prec == CORR_PREC_FP64) ? 1 : 0;
    GemmArgs a{M, N, K, A, lda, 0, B, ldb, 0, C, N, 0, 1.0, 0.0, 0, 0, 0};
    a.dma = dma_ok(f64, A, lda, 0, B, ldb, 0);
    a.epi = mapping;
    a.en = n;
    if (mapping == GEMM_EPI_PAIRS)
        return f64 ? launch_t<1, 1, double, GEMM_EPI_PAIRS>(ctx, a, 1) : launch_t<1, 1, float, GEMM_EPI_PAIRS>(ctx, a, 1);
    return f64 ? launch_t<1, 1, double, GEMM_EPI_MIRROR>(ctx, a, 1) : launch_t<1, 1, float, GEMM_EPI_MIRROR>(ctx, a, 1);
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

// First pass of a long reduction: block b sums the fixed chunk
// [b*chunk, (b+1)*chunk) of every series in a fixed order.
__global__ __launch_bounds__(256) void reduce_chunks_kernel(const double* __restrict__ d, int64_t n, int nser,
                                                            int64_t chunk, double* __restrict__ out) {
    __shared__ double sh[256];
    const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < n ? lo + chunk : n;
    for (int s = 0; s < nser; ++s) {
        double acc = 0.0;
        for (int64_t i = lo + threadIdx.x; i < hi; i += 256) acc += d[i * nser + s];
        sh[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[blockIdx.x * nser + s] = sh[0];
        __syncthreads();
    }
}

int reduce_sum(corr_ctx* ctx, const double* d_data, int64_t n, int nser, double* h_out) {
    double *d_out = nullptr, *d_mid = nullptr;
    CORR_TRY(scratch(ctx, "reduce_out", sizeof(double) * 16, (void**)&d_out));
    if (n > 64 * REDUCE_NCHUNK)
        CORR_TRY(scratch(ctx, "reduce_mid", sizeof(double) * 16 * REDUCE_NCHUNK, (void**)&d_mid));
    return reduce_sum_ws(ctx, d_data, n, nser, h_out, d_mid, d_out);
}

int reduce_sum_ws(corr_ctx* ctx, const double* d_data, int64_t n, int nser, double* h_out, double* d_mid,
                  double* d_out) {
    CORR_REQUIRE(ctx, nser >= 1 && nser <= 16, "reduce_sum: 1..16 series");
    // Long inputs: 256 fixed chunks first, so the single-workgroup pass is short.
    constexpr int64_t NCHUNK = REDUCE_NCHUNK;
    if (n > 64 * NCHUNK) {
        hipLaunchKernelGGL(reduce_chunks_kernel, dim3((unsigned)NCHUNK), dim3(256), 0, ctx->stream, d_data, n, nser,
                           cdiv(n, NCHUNK), d_mid);
        CORR_LAUNCHED(ctx, "reduce_chunks_kernel");
        d_data = d_mid;
        n = NCHUNK;
    }
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
