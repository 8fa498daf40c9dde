// FP64 GEMM emulated on the int8 matrix cores (Ozaki splitting).
//
// C = alpha op(A) op(B) + beta C for row-major FP64 operands (op = identity or
// transpose).  Every row m of op(A) gets a power-of-two scale sig[m] >=
// max_k |op(A)[m][k]| (127/128 headroom) and every column n of op(B) a scale
// tau[n]; the scaled values x in (-1, 1) are cut
// into S = 5 int8 slices by round-to-nearest,
//     r_0 = x,  q_s = rint(128 r_s),  r_{s+1} = 128 r_s - q_s,
// so x = sum_s q_s 2^-7(s+1) + r_S 2^-7S with |r_S| <= 1/2 (|q_0| <= 127,
// |q_s| <= 64 after).  Every product q_s(A) q_t(B) with s + t <= S - 1 is an
// exact int32 sum on v_mfma_i32_32x32x32_i8 (15 of the 25 slice pairs; the
// dropped ones and the remainders are below 2^-35 of the row and column
// maxima), accumulated per diagonal d = s + t, and the epilogue combines the
// five diagonals in FP64 by Horner's rule in 2^-7 (exact up to the last step)
// and applies sig[m] tau[n] 2^-14.  Results are bit-reproducible: the integer
// sums are exact and the FP64 steps run in a fixed order.
//
// Non-finite entries (and maxima of 2^1023 and above) make their row's (A) or
// column's (B) scale NaN, so the whole row or column of C is NaN.  K is
// limited to 75000 (int32 range of the widest diagonal:
// K (2 127 64 + 3 64 64) < 2^31).
//
// Operands are sliced into the kernel's LDS stage order first (rows of op(A)
// and columns of op(B) whose k run is contiguous in memory: a scale pass and a
// slice pass over (row tile, stage) blocks; the others: one block per 128-row
// tile, maxima then slices), so each stage copy is a contiguous 1 KiB run per
// wave: a stage holds 32 k of the S slices of 128 rows
// of A and of 128 columns of B, [slice][row][32 bytes], the two 16-byte k
// chunks of a row in swizzled slots (conflict-free ds_read_b128 fragments).
#include "common.hpp"

#include <algorithm>
#include <cfloat>
#include <type_traits>

namespace corr {

namespace {
constexpr int OZ_S = 5;                          // slices per operand
constexpr int OZ_T = 128;                        // rows (columns) of a tile
constexpr int OZ_KS = 32;                        // k per stage
constexpr int OZ_SLICE = OZ_T * OZ_KS;           // bytes of one slice of a stage (4 KiB)
constexpr int OZ_PANEL = OZ_S * OZ_SLICE;        // one operand's stage (20 KiB)
constexpr int OZ_NQ = 2;                         // 32-column blocks per wave (8 waves)
// Stage buffers of OZ_KD x 40 KiB (OZ_KD 32-deep sub-stages per barrier).
// Measured on the (H2O)32 B formation against three buffers of one
// sub-stage: four buffers (all 160 KiB of LDS, copies three stages ahead) no
// faster, two of two sub-stages (half the barriers, copies one stage ahead)
// 1% slower; fragments read one slice group ahead of their MFMAs (pinned by
// scheduling barriers) no faster either.
#ifndef OZ_NBUF
#define OZ_NBUF 3
#endif
#ifndef OZ_KD
#define OZ_KD 1
#endif
constexpr int64_t OZ_KMAX = 75000;

typedef int oz_i4 __attribute__((ext_vector_type(4)));
typedef int oz_i16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* oz_lds_ptr;
typedef __attribute__((address_space(1))) const void* oz_gbl_ptr;

// Slot of 16-byte k chunk ch of row r in a 32-byte LDS row (an involution in ch).
__device__ inline int oz_slot(int row, int ch) { return ch ^ ((row >> 3) & 1); }

// Scale 2^e for a row (column) with maximum magnitude mx: |x| / 2^e < 127.5/128
// for every |x| <= mx, so the first slice stays within [-127, 127].  1 for an
// all-zero row, NaN when an entry is not finite.
__device__ inline double oz_scale(double mx, bool bad) {
    if (bad) return __builtin_nan("");
    if (!(mx > 0.0)) return 1.0;
    int e;
    const double f = frexp(mx, &e);   // mx = f 2^e, f in [0.5, 1)
    if (f * 128.0 >= 127.5) ++e;
    if (e > 1023) return __builtin_nan("");   // 2^e would overflow
    return ldexp(1.0, e);
}

// inv_scale = 0 marks a row (column) with a NaN scale: its slices are zero.
__device__ inline void oz_slices(double x, double inv_scale, int8_t (&q)[OZ_S]) {
    double r = inv_scale == 0.0 ? 0.0 : x * inv_scale;   // power-of-two scaling: exact
#pragma unroll
    for (int s = 0; s < OZ_S; ++s) {
        r *= 128.0;
        const double t = rint(r);
        r -= t;
        q[s] = (int8_t)(int)t;
    }
}

// 1 / scale as an exact power of two (0 for a NaN scale: its slices are zero).
__device__ inline double oz_inv(double scale) {
    if (scale != scale) return 0.0;
    int e;
    (void)frexp(scale, &e);
    return ldexp(1.0, 1 - e);
}

// Operand rows r (of op(A), or columns of op(B)) whose k run is contiguous:
// element (r, k) at X[r ld + k].  Scales, one wave per row.
__global__ __launch_bounds__(256) void oz_scale_kmajor_kernel(const double* __restrict__ A, int64_t lda, int64_t M,
                                                              int64_t K, double* __restrict__ sig) {
    const int64_t m = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (m >= M) return;
    double mx = 0.0;
    bool bad = false;
    for (int64_t k = lane; k < K; k += 64) {
        const double a = fabs(A[m * lda + k]);
        bad |= !(a <= DBL_MAX);
        mx = a > mx ? a : mx;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
        bad |= __shfl_xor((int)bad, off, 64) != 0;
    }
    if (lane == 0) sig[m] = oz_scale(mx, bad);
}

// Slices of k-contiguous rows into stage order: one block per (row tile,
// stage); thread t takes row t / 2 and k chunk t % 2 (16 values).  Rows past M
// and k past K are zero.
__global__ __launch_bounds__(256) void oz_split_kmajor_kernel(const double* __restrict__ A, int64_t lda, int64_t M,
                                                              int64_t K, int64_t Ks, const double* __restrict__ sig,
                                                              int8_t* __restrict__ Ap) {
    const int64_t st = blockIdx.x % Ks, mt = blockIdx.x / Ks;
    const int r = threadIdx.x >> 1, ch = threadIdx.x & 1;
    const int64_t m = mt * OZ_T + r, k0 = st * OZ_KS + ch * 16;
    const double inv = m < M ? oz_inv(sig[m]) : 0.0;
    int8_t q[16][OZ_S];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const double x = (m < M && k0 + j < K) ? A[m * lda + k0 + j] : 0.0;
        oz_slices(x, inv, q[j]);
    }
    int8_t* out = Ap + (mt * Ks + st) * OZ_PANEL + r * 32 + oz_slot(r, ch) * 16;
#pragma unroll
    for (int s = 0; s < OZ_S; ++s) {
        oz_i4 w;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            w[u] = (int)(uint8_t)q[4 * u][s] | ((int)(uint8_t)q[4 * u + 1][s] << 8) |
                   ((int)(uint8_t)q[4 * u + 2][s] << 16) | ((int)(uint8_t)q[4 * u + 3][s] << 24);
        *reinterpret_cast<oz_i4*>(out + s * OZ_SLICE) = w;
    }
}

// Scales and slices of operand rows laid out along memory rows: element (r, k)
// at X[k ld + r] (columns of op(B) = B, rows of op(A) = A^T).  One block per
// 128-row tile; thread t takes row t / 2 and, in every stage, k chunk t % 2.
// Pass 1 finds the maxima (the two chunk threads of a row are adjacent
// lanes), pass 2 writes the slices of every stage.
__global__ __launch_bounds__(256) void oz_split_rmajor_kernel(const double* __restrict__ B, int64_t ldb, int64_t N,
                                                              int64_t K, int64_t Ks, double* __restrict__ tau,
                                                              int8_t* __restrict__ Bp) {
    const int64_t nt = blockIdx.x;
    const int c = threadIdx.x >> 1, ch = threadIdx.x & 1;
    const int64_t n = nt * OZ_T + c;
    const bool col_ok = n < N;
    double mx = 0.0;
    bool bad = false;
    for (int64_t st = 0; st < Ks; ++st) {
        const int64_t k0 = st * OZ_KS + ch * 16;
        double x[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = (col_ok && k0 + j < K) ? B[(k0 + j) * ldb + n] : 0.0;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double a = fabs(x[j]);
            bad |= !(a <= DBL_MAX);
            mx = a > mx ? a : mx;
        }
    }
    {
        const double o = __shfl_xor(mx, 1, 64);
        mx = o > mx ? o : mx;
        bad |= __shfl_xor((int)bad, 1, 64) != 0;
    }
    const double scale = oz_scale(mx, bad);
    if (col_ok && ch == 0) tau[n] = scale;
    const double inv = col_ok ? oz_inv(scale) : 0.0;
    for (int64_t st = 0; st < Ks; ++st) {
        const int64_t k0 = st * OZ_KS + ch * 16;
        int8_t q[16][OZ_S];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double x = (col_ok && k0 + j < K) ? B[(k0 + j) * ldb + n] : 0.0;
            oz_slices(x, inv, q[j]);
        }
        int8_t* out = Bp + (nt * Ks + st) * OZ_PANEL + c * 32 + oz_slot(c, ch) * 16;
#pragma unroll
        for (int s = 0; s < OZ_S; ++s) {
            oz_i4 w;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                w[u] = (int)(uint8_t)q[4 * u][s] | ((int)(uint8_t)q[4 * u + 1][s] << 8) |
                       ((int)(uint8_t)q[4 * u + 2][s] << 16) | ((int)(uint8_t)q[4 * u + 3][s] << 24);
            *reinterpret_cast<oz_i4*>(out + s * OZ_SLICE) = w;
        }
    }
}

// The same slices in one pass when K is short (2 Ks 16-deep chunks <= 64 U):
// a workgroup of 1024 threads takes 16 columns n and the whole K range in
// registers (thread t: column t % 16, chunks t / 16 + 64 u), so the operand is
// read once; the column maxima go over the four lanes of a column in a wave,
// then over the 16 waves through LDS.  Eight workgroups cover a 128-column
// panel, the columns past N included (zero slices), as in the two-pass
// kernel.  Each slice is taken in place in the value registers, with the
// arithmetic of oz_slices.
template <int U>
__global__ __launch_bounds__(1024) void oz_split_rmajor1_kernel(const double* __restrict__ B, int64_t ldb,
                                                               int64_t N, int64_t K, int64_t Ks,
                                                               double* __restrict__ tau, int8_t* __restrict__ Bp) {
    __shared__ double smx[16][16];
    __shared__ int sbad[16][16];
    const int t = threadIdx.x, c = t & 15, ck0 = t >> 4, wave = t >> 6, lane = t & 63;
    const int64_t n = (int64_t)blockIdx.x * 16 + c;
    const bool col_ok = n < N;
    const int64_t nck = 2 * Ks;
    double x[U][16];
    double mx = 0.0;
    bool bad = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t ck = ck0 + 64 * u;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int64_t k = ck * 16 + j;
            x[u][j] = (col_ok && ck < nck && k < K) ? B[k * ldb + n] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const double a = fabs(x[u][j]);
            bad |= !(a <= DBL_MAX);
            mx = a > mx ? a : mx;
        }
    }
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
        const double o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
        bad |= __shfl_xor((int)bad, off, 64) != 0;
    }
    if (lane < 16) {
        smx[wave][lane] = mx;
        sbad[wave][lane] = bad;
    }
    __syncthreads();
    double cm = 0.0;
    bool cb = false;
#pragma unroll
    for (int w = 0; w < 16; ++w) {
        const double o = smx[w][c];
        cm = o > cm ? o : cm;
        cb |= sbad[w][c] != 0;
    }
    const double scale = oz_scale(cm, cb);
    if (col_ok && ck0 == 0) tau[n] = scale;
    const double inv = col_ok ? oz_inv(scale) : 0.0;
    const int64_t nt = n / OZ_T;
    const int col = (int)(n % OZ_T);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int64_t ck = ck0 + 64 * u;
        if (ck >= nck) continue;
        const int ch = (int)(ck & 1);
        int8_t* out = Bp + (nt * Ks + (ck >> 1)) * OZ_PANEL + col * 32 + oz_slot(col, ch) * 16;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[u][j] = inv == 0.0 ? 0.0 : x[u][j] * inv;   // power-of-two scaling: exact
#pragma unroll
        for (int sl = 0; sl < OZ_S; ++sl) {
            oz_i4 w;
#pragma unroll
            for (int v4 = 0; v4 < 4; ++v4) {
                int word = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    double& r = x[u][4 * v4 + b];
                    r *= 128.0;
                    const double q = rint(r);
                    r -= q;
                    word |= (int)(uint8_t)(int8_t)(int)q << (8 * b);
                }
                w[v4] = word;
            }
            *reinterpret_cast<oz_i4*>(out + sl * OZ_SLICE) = w;
        }
    }
}

__device__ inline int64_t oz_xcd_logical(int64_t per) {
    const int64_t w = blockIdx.x;
    return (w % 8) * per + (w / 8);
}

// Waits until at most N of this wave's global->LDS copies are outstanding and its
// LDS reads are back, then meets the workgroup barrier.
#define OZ_STAGE_BARRIER(N)                                     \
    {                                                           \
        __builtin_amdgcn_s_waitcnt((N) | (0x7 << 4));           \
        __asm__ volatile("" ::: "memory");                      \
        __builtin_amdgcn_s_barrier();                           \
        __asm__ volatile("" ::: "memory");                      \
    }

// One 128 x 128 tile of C per workgroup, waves as 4 (rows) x 4 / NQ (columns):
// wave w owns 32 rows and NQ 32-column blocks, with five int32 diagonal
// accumulators per block (80 NQ registers).  NQ = 2 (8 waves of up to 256
// registers) is used: with NQ = 1 (16 waves of 128) the accumulators and
// fragments spill.  Stage buffers: OZ_NBUF of 40 KiB (A slices, then B
// slices), copies OZ_NBUF - 1 stages ahead; the 40 1 KiB copies of a stage
// are dealt out to the waves in turn.  With a_lower the K range of
// row tile mt stops at 128 (mt + 1).  Tiles come in groups of GM x GN (about
// the 32 workgroups one XCD runs at once; row tiles fastest within a group,
// groups of one column block consecutive, deepest row tiles first), so that
// concurrent workgroups share both operands' panels in L2.
// BF16 hi/lo of a double for the energy kernel's panels: the conversion of the
// RI-MP2 repack (round-to-nearest-even of the FP32 value; lo from x - hi).
__device__ inline uint16_t oz_f32_to_bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline void oz_bf16_split(double x, uint16_t& hi, uint16_t& lo) {
    hi = oz_f32_to_bf16_rn((float)x);
    lo = oz_f32_to_bf16_rn((float)(x - (double)__uint_as_float(((uint32_t)hi) << 16)));
}

template <int NQ, bool PAN>
__global__ __launch_bounds__(64 * 16 / NQ) void oz_gemm_kernel(const int8_t* __restrict__ Ap,
                                                               const int8_t* __restrict__ Bp,
                                                               const double* __restrict__ sig,
                                                               const double* __restrict__ tau,
                                                               double* __restrict__ C, int64_t ldc, int64_t M,
                                                               int64_t N, int64_t Ks, int nMT, int64_t nNT,
                                                               int GM, int GN, int64_t nT, int a_lower, double alpha,
                                                               double beta, int64_t per, Bf16Panels pan) {
    constexpr int W = 16 / NQ, WC = 4 / NQ;                 // waves, wave columns
    constexpr int NCH = 2 * OZ_PANEL / 1024;                 // 1 KiB copies per stage (40)
    constexpr int NC = NCH / W;                              // copies per wave and stage
    static_assert(NCH % W == 0, "oz_gemm: copies must divide among the waves");
    __shared__ __attribute__((aligned(16))) int8_t L[OZ_NBUF][OZ_KD][2 * OZ_PANEL];
    const int64_t id = oz_xcd_logical(per);
    if (id >= nT) return;
    const int64_t nGM = (nMT + GM - 1) / GM, g = id / (GM * GN), r = id % (GM * GN);
    const int mt = nMT - 1 - (int)((g % nGM) * GM + r % GM);
    const int64_t nt = (g / nGM) * GN + r / GM;
    if (mt < 0 || nt >= nNT) return;   // past the edge of a partial group
    const int64_t Kc = a_lower ? (Ks < 4 * (int64_t)(mt + 1) ? Ks : 4 * (int64_t)(mt + 1)) : Ks;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r32 = lane & 31, kh = lane >> 5;
    const int wr = wave / WC, wc = wave % WC;

    // copies: chunks c = wave + W k of the stage's 40 (20 of A, then 20 of B)
    const int8_t* src[NC];
    int seg[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
        const int c = wave + W * k;
        const int8_t* base = c < OZ_PANEL / 1024 ? Ap + (int64_t)mt * Ks * OZ_PANEL + c * 1024
                                                 : Bp + nt * Ks * OZ_PANEL + (c - OZ_PANEL / 1024) * 1024;
        src[k] = base + lane * 16;
        seg[k] = c * 1024;
    }
    const int rowA = wr * 32 + r32;
    const int offA = rowA * 32 + oz_slot(rowA, kh) * 16;
    int offB[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int col = (NQ * wc + q) * 32 + r32;
        offB[q] = OZ_PANEL + col * 32 + oz_slot(col, kh) * 16;
    }
    oz_i16 acc[OZ_S][NQ];
#pragma unroll
    for (int d = 0; d < OZ_S; ++d)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[d][q][r] = 0;

    // A stage is OZ_KD 32-deep sub-stages; stage j lives in buffer j % OZ_NBUF,
    // copies run OZ_NBUF - 1 stages ahead (sub-stages past the end repeat
    // sub-stage Kc - 1 and are not computed), and the barrier closing stage j
    // waits until this wave's copies of stage j + 1 have landed.
    const int64_t Sc = (Kc + OZ_KD - 1) / OZ_KD;
    auto copy_stage = [&](int b, int64_t j) {
#pragma unroll
        for (int h = 0; h < OZ_KD; ++h) {
            const int64_t st = j * OZ_KD + h < Kc ? j * OZ_KD + h : Kc - 1;
#pragma unroll
            for (int k = 0; k < NC; ++k)
                __builtin_amdgcn_global_load_lds((oz_gbl_ptr)(src[k] + st * OZ_PANEL),
                                                 (oz_lds_ptr)(&L[b][h][0] + seg[k]), 16, 0, 0);
        }
    };
#pragma unroll
    for (int p = 0; p < OZ_NBUF - 1; ++p) copy_stage(p, p < Sc ? p : Sc - 1);
    OZ_STAGE_BARRIER(NC * OZ_KD * (OZ_NBUF - 2))
    int buf = 0;
    for (int64_t j = 0; j < Sc; ++j) {
        copy_stage(buf == 0 ? OZ_NBUF - 1 : buf - 1, j + OZ_NBUF - 1 < Sc ? j + OZ_NBUF - 1 : Sc - 1);
#pragma unroll
        for (int h = 0; h < OZ_KD; ++h) {
            if (OZ_KD > 1 && j * OZ_KD + h >= Kc) break;
            const int8_t* S = &L[buf][h][0];
            oz_i4 a[OZ_S];
#pragma unroll
            for (int sa = 0; sa < OZ_S; ++sa) a[sa] = *reinterpret_cast<const oz_i4*>(S + sa * OZ_SLICE + offA);
#pragma unroll
            for (int t = 0; t < OZ_S; ++t) {
                oz_i4 b[NQ];
#pragma unroll
                for (int q = 0; q < NQ; ++q) b[q] = *reinterpret_cast<const oz_i4*>(S + t * OZ_SLICE + offB[q]);
#pragma unroll
                for (int sa = 0; sa + t < OZ_S; ++sa)
#pragma unroll
                    for (int q = 0; q < NQ; ++q)
                        acc[sa + t][q] =
                            __builtin_amdgcn_mfma_i32_32x32x32_i8(a[sa], b[q], acc[sa + t][q], 0, 0, 0);
            }
        }
        OZ_STAGE_BARRIER(NC * OZ_KD * (OZ_NBUF - 2))
        buf = buf == OZ_NBUF - 1 ? 0 : buf + 1;
    }
    __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xF << 8));

    // element r of lane l: C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] of its 32 x 32 block;
    // the 16 row scales and 2 column scales of a lane are loaded together
    constexpr double R7 = 1.0 / 128.0, R14 = 1.0 / 16384.0;
    const int64_t m0 = (int64_t)mt * OZ_T + wr * 32 + 4 * kh;
    double sg[16], tn[NQ];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int64_t m = m0 + 8 * (r / 4) + (r % 4);
        sg[r] = m < M ? sig[m] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int64_t n = nt * OZ_T + (NQ * wc + q) * 32 + r32;
        tn[q] = n < N ? tau[n] : 0.0;
    }
    auto store = [&](auto with_beta) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int64_t n = nt * OZ_T + (NQ * wc + q) * 32 + r32;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t m = m0 + 8 * (r / 4) + (r % 4);
                double h = (double)acc[4][q][r];
                h = h * R7 + (double)acc[3][q][r];
                h = h * R7 + (double)acc[2][q][r];
                h = h * R7 + (double)acc[1][q][r];
                h = h * R7 + (double)acc[0][q][r];
                const double val = alpha * ((h * R14) * sg[r] * tn[q]);
                if (m < M && n < N) {
                    double* c = C + m * ldc + n;
                    if constexpr (decltype(with_beta)::value) *c = val + beta * *c;
                    else *c = val;
                }
            }
        }
    };
    if constexpr (PAN) {
        // groups of four rows (4 k of one 16-byte chunk) per column: an
        // 8-byte hi store and an 8-byte lo store
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int64_t n = nt * OZ_T + (NQ * wc + q) * 32 + r32;
            if (n >= N) continue;
            const int i = (int)(n / pan.v), a = (int)(n % pan.v);
            const int c = (i % 4) * 32 + a % 32;
            const int64_t blk = (int64_t)(i / 4) * pan.nAC + a / 32;
#pragma unroll
            for (int gq = 0; gq < 4; ++gq) {
                const int64_t Q0 = m0 + 8 * gq;
                if (Q0 >= pan.Kp) continue;
                uint16_t hv[4], lv[4];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int r = 4 * gq + rr;
                    double h = (double)acc[4][q][r];
                    h = h * R7 + (double)acc[3][q][r];
                    h = h * R7 + (double)acc[2][q][r];
                    h = h * R7 + (double)acc[1][q][r];
                    h = h * R7 + (double)acc[0][q][r];
                    const double val = Q0 + rr < M ? (h * R14) * sg[r] * tn[q] : 0.0;
                    oz_bf16_split(val, hv[rr], lv[rr]);
                }
                const int ch = (int)((Q0 % 16) / 8), slot = ch ^ ((c >> 3) & 1);
                uint16_t* dst = pan.p + (blk * pan.Ks + Q0 / 16) * (2 * OZ_T * 16) + c * 16 + slot * 8 + Q0 % 8;
                uint2 wh, wl;
                wh.x = hv[0] | ((uint32_t)hv[1] << 16);
                wh.y = hv[2] | ((uint32_t)hv[3] << 16);
                wl.x = lv[0] | ((uint32_t)lv[1] << 16);
                wl.y = lv[2] | ((uint32_t)lv[3] << 16);
                *reinterpret_cast<uint2*>(dst) = wh;
                *reinterpret_cast<uint2*>(dst + OZ_T * 16) = wl;
            }
        }
    } else {
        if (beta == 0.0) store(std::false_type{});
        else store(std::true_type{});
    }
}

}  // namespace

// Scales and stage-order slices of the R x K operand X (kmajor: (r, k) at
// X[r ld + k], else at X[k ld + r]).
static int oz_split(corr_ctx* ctx, const double* X, int64_t ld, int64_t R, int64_t K, int64_t Ks, bool kmajor,
                    double* scale, int8_t* P) {
    if (kmajor) {
        hipLaunchKernelGGL(oz_scale_kmajor_kernel, dim3((unsigned)cdiv(R, 4)), dim3(256), 0, ctx->stream, X, ld, R,
                           K, scale);
        CORR_LAUNCHED(ctx, "oz_scale_kmajor_kernel");
        hipLaunchKernelGGL(oz_split_kmajor_kernel, dim3((unsigned)(cdiv(R, OZ_T) * Ks)), dim3(256), 0, ctx->stream,
                           X, ld, R, K, Ks, scale, P);
        CORR_LAUNCHED(ctx, "oz_split_kmajor_kernel");
    } else if (2 * Ks <= 64 * 3) {
        // short K: one pass, the operand held in registers (Ks <= 96, K <= 3072)
        const dim3 grid((unsigned)(cdiv(R, OZ_T) * (OZ_T / 16)));
        if (2 * Ks <= 64)
            hipLaunchKernelGGL(oz_split_rmajor1_kernel<1>, grid, dim3(1024), 0, ctx->stream, X, ld, R, K, Ks, scale, P);
        else if (2 * Ks <= 128)
            hipLaunchKernelGGL(oz_split_rmajor1_kernel<2>, grid, dim3(1024), 0, ctx->stream, X, ld, R, K, Ks, scale, P);
        else
            hipLaunchKernelGGL(oz_split_rmajor1_kernel<3>, grid, dim3(1024), 0, ctx->stream, X, ld, R, K, Ks, scale, P);
        CORR_LAUNCHED(ctx, "oz_split_rmajor1_kernel");
    } else {
        hipLaunchKernelGGL(oz_split_rmajor_kernel, dim3((unsigned)cdiv(R, OZ_T)), dim3(256), 0, ctx->stream, X, ld,
                           R, K, Ks, scale, P);
        CORR_LAUNCHED(ctx, "oz_split_rmajor_kernel");
    }
    return CORR_OK;
}

int ozaki_prepare(corr_ctx* ctx, const char* name, const double* X, int64_t ld, int64_t R, int64_t K,
                  bool kmajor, OzOperand* out) {
    CORR_REQUIRE(ctx, R > 0 && K >= 0 && X && out && ld >= (kmajor ? K : R), "int8x5 operand: bad arguments");
    CORR_REQUIRE(ctx, K <= OZ_KMAX, "gemm int8x5: K above 75000 (int32 range of the slice sums)");
    const int64_t nRT = cdiv(R, OZ_T), Ks = cdiv(K > 0 ? K : 1, OZ_KS);
    CORR_REQUIRE(ctx, nRT * Ks < ((int64_t)1 << 31), "gemm int8x5: operand too large for one slicing launch");
    const std::string n(name);
    int8_t* P = nullptr;
    double* sc = nullptr;
    CORR_TRY(scratch(ctx, (n + "_slices").c_str(), (size_t)nRT * Ks * OZ_PANEL, (void**)&P));
    CORR_TRY(scratch(ctx, (n + "_scales").c_str(), sizeof(double) * nRT * OZ_T, (void**)&sc));
    KernelTimer kt(ctx, "oz_split");
    CORR_TRY(oz_split(ctx, X, ld, R, K, Ks, kmajor, sc, P));
    *out = OzOperand{P, sc, R, K, Ks};
    return CORR_OK;
}

int ozaki_gemm_prepared(corr_ctx* ctx, const OzOperand& A, const OzOperand& B, double alpha, double beta, double* C,
                        int64_t ldc, bool a_lower, const Bf16Panels* pan) {
    CORR_REQUIRE(ctx, A.slices && B.slices && A.K == B.K, "gemm int8x5: operands do not match");
    CORR_REQUIRE(ctx, pan ? (pan->p && alpha == 1.0 && beta == 0.0 && (int64_t)pan->o * pan->v == B.R &&
                            pan->Kp % 16 == 0 && pan->Kp >= A.R && pan->Ks == pan->Kp / 16)
                          : (C && ldc >= B.R),
                 "gemm int8x5: bad output");
    const int64_t M = A.R, N = B.R, Ks = A.Ks;
    const int64_t nMT = cdiv(M, OZ_T), nNT = cdiv(N, OZ_T);
    // groups of about 32 tiles: up to 8 column tiles, the rest in rows
    const int GN = (int)std::min<int64_t>(nNT, 8), GM = std::max(1, 32 / GN);
    const int64_t nT = cdiv(nMT, GM) * GM * cdiv(nNT, GN) * GN, per = cdiv(nT, 8);
    CORR_REQUIRE(ctx, nMT < (1 << 30) && 8 * per < ((int64_t)1 << 31), "gemm int8x5: too many tiles for one launch");
    KernelTimer kt(ctx, "oz_gemm");
    if (pan)
        hipLaunchKernelGGL((oz_gemm_kernel<OZ_NQ, true>), dim3((unsigned)(8 * per)), dim3(64 * 16 / OZ_NQ), 0,
                           ctx->stream, A.slices, B.slices, A.scales, B.scales, C, ldc, M, N, Ks, (int)nMT, nNT, GM,
                           GN, nT, a_lower ? 1 : 0, alpha, beta, per, *pan);
    else
        hipLaunchKernelGGL((oz_gemm_kernel<OZ_NQ, false>), dim3((unsigned)(8 * per)), dim3(64 * 16 / OZ_NQ), 0,
                           ctx->stream, A.slices, B.slices, A.scales, B.scales, C, ldc, M, N, Ks, (int)nMT, nNT, GM,
                           GN, nT, a_lower ? 1 : 0, alpha, beta, per, Bf16Panels{});
    CORR_LAUNCHED(ctx, "oz_gemm_kernel");
    return CORR_OK;
}

int gemm_ozaki(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K, double alpha,
               const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
               bool a_lower) {
    CORR_REQUIRE(ctx, M >= 0 && N >= 0 && K >= 0, "gemm int8x5: negative size");
    CORR_REQUIRE(ctx, K <= OZ_KMAX, "gemm int8x5: K above 75000 (int32 range of the slice sums)");
    if (M == 0 || N == 0) return CORR_OK;
    CORR_REQUIRE(ctx, A && B && C && lda >= (transa ? M : K) && ldb >= (transb ? K : N) && ldc >= N,
                 "gemm int8x5: bad operands");
    OzOperand oa, ob;
    CORR_TRY(ozaki_prepare(ctx, "oz_A", A, lda, M, K, !transa, &oa));
    CORR_TRY(ozaki_prepare(ctx, "oz_B", B, ldb, N, K, transb != 0, &ob));
    return ozaki_gemm_prepared(ctx, oa, ob, alpha, beta, C, ldc, a_lower);
}

}  // namespace corr
