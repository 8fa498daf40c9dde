// RI-MP2 energy: fused pair-block GEMM + FP64 denominator + deterministic
// FP64 reduction.  Restates run_canonical_mp2_ri (src/POSTHF/can_mp2.cc:53-132):
//
//   M^{ij}_ab = sum_Q B^Q_ia B^Q_jb
//   EMP2_OS  = - sum_{ij ab} M_ab^2 / D,   EMP2_SS = - sum M_ab (M_ab - M_ba) / D
//   D = e_a + e_b - e_i - e_j
//
// The reference loops over all o^2 ordered pairs with one dgemm each.  Here
// the o*v x o*v matrix G[(i,a),(j,b)] = M^{ij}_ab is tiled with 128x128 tiles
// whose rows are (4 occupied) x (32 virtual) and whose columns are likewise
// (4 occupied) x (32 virtual).  A workgroup owns one tile T1 = G[I x A, J x B]
// and its exchange partner T2[(i,a),(j,b)] = M^{ij}_ba computed as
// B[:, J x A]^T B[:, I x B], so that M_ab and M_ba of every element sit in the
// same lane and register: the exchange energy needs no data movement.
// Only I <= J and A <= B tiles are launched; weights restore the ordered-pair
// sum of the reference exactly:
//   I <  J: x2 (the (j,i) pair gives the same sum with a<->b)
//   A == B: x1/2 (T1 and T2 cover the same elements)
//
// B is repacked once into panels: panel (I, A) is a contiguous
// [Kp][128] slab (FP64/FP32), or a [Kp/16][hi/lo][128][2][8] slab of BF16
// halves (BF16x3) already in the energy kernel's LDS order, so every K-slab a
// workgroup stages into LDS is one contiguous, fully coalesced read, moved by
// direct global->LDS copies in all three kernels.
#include "common.hpp"

#include <type_traits>

namespace corr {

namespace {

constexpr int PI = 4;     // occupied per panel
constexpr int PA = 32;    // virtuals per panel
constexpr int PW = PI * PA;  // 128 panel width

__device__ inline void tri_decode(int64_t p, int64_t n, int& r, int& c) {
    // p enumerates (r, c), r <= c < n, row-major.
    const double b = 2.0 * n + 1.0;
    int64_t rr = (int64_t)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
    if (rr < 0) rr = 0;
    if (rr > n - 1) rr = n - 1;
    auto start = [n](int64_t x) { return x * n - x * (x - 1) / 2; };
    while (rr > 0 && start(rr) > p) --rr;
    while (rr + 1 < n && start(rr + 1) <= p) ++rr;
    r = (int)rr;
    c = (int)(rr + (p - start(rr)));
}

// Map the dispatch order (round-robin over the 8 XCDs) to a logical order in
// which consecutive ids run on the same XCD and share panels through its L2.
__device__ inline int64_t xcd_logical(int64_t per) {
    const int64_t w = blockIdx.x;
    return (w % 8) * per + (w / 8);
}

__device__ inline uint16_t f32_to_bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Slot of 16-byte k-chunk ch (0 or 1) in a 32-byte LDS row r of the BF16 energy
// kernel (an involution in ch): keeps its ds_read_b128 fragment reads, 32 rows
// per instruction, free of bank conflicts on the 64-bank LDS.
__device__ inline int bf16_slot(int row, int ch) { return ch ^ ((row >> 3) & 1); }

}  // namespace

// ---- repack ----------------------------------------------------------------
template <typename T>
__global__ void repack_real(const double* __restrict__ Bia, T* __restrict__ Bp, int64_t naux,
                            int o, int v, int nAC, int64_t Kp, int64_t total) {
    for (int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; d < total;
         d += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(d % PW);
        const int64_t q = (d / PW) % Kp;
        const int64_t blk = d / (PW * Kp);
        const int ib = (int)(blk / nAC), ac = (int)(blk % nAC);
        const int i = ib * PI + c / PA, a = ac * PA + c % PA;
        double x = 0.0;
        if (q < naux && i < o && a < v) x = Bia[(q * o + i) * (int64_t)v + a];
        Bp[d] = (T)x;
    }
}

// One block per (panel, 32-row K slab): read [32 q][128 c] coalesced along a,
// write two 16-deep stages, each [hi/lo][128 c][2 slots][8 q], in the energy
// kernel's LDS order (slot s of row c holds k-chunk bf16_slot(c, s)).
__global__ __launch_bounds__(256) void repack_bf16(const double* __restrict__ Bia,
                                                   uint16_t* __restrict__ Bp, int64_t naux, int o,
                                                   int v, int nAC, int64_t Kc) {
    __shared__ double S[32][PW + 1];
    const int64_t kc = blockIdx.x % Kc;
    const int64_t blk = blockIdx.x / Kc;
    const int ib = (int)(blk / nAC), ac = (int)(blk % nAC);
    for (int e = threadIdx.x; e < 32 * PW; e += 256) {
        const int q = e / PW, c = e % PW;
        const int64_t Q = kc * 32 + q;
        const int i = ib * PI + c / PA, a = ac * PA + c % PA;
        S[q][c] = (Q < naux && i < o && a < v) ? Bia[(Q * o + i) * (int64_t)v + a] : 0.0;
    }
    __syncthreads();
    // the two stages of this slab are consecutive: 2 x [2][PW][2][8] bf16
    uint16_t* out = Bp + (blk * Kc + kc) * (2 * PW * 32);
    for (int e = threadIdx.x; e < 2 * 2 * PW * 2; e += 256) {
        const int h = e / (2 * PW * 2), r0 = e % (2 * PW * 2);
        const int hl = r0 / (PW * 2), c = (r0 % (PW * 2)) / 2, kq = bf16_slot(c, r0 % 2);
        uint16_t hv[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const double x = S[h * 16 + kq * 8 + t][c];
            const uint16_t hi = f32_to_bf16_rn((float)x);
            if (hl == 0) hv[t] = hi;
            else hv[t] = f32_to_bf16_rn((float)(x - (double)bf16_to_f32(hi)));
        }
        uint4 pk;
        pk.x = hv[0] | ((uint32_t)hv[1] << 16);
        pk.y = hv[2] | ((uint32_t)hv[3] << 16);
        pk.z = hv[4] | ((uint32_t)hv[5] << 16);
        pk.w = hv[6] | ((uint32_t)hv[7] << 16);
        *reinterpret_cast<uint4*>(out + (int64_t)e * 8) = pk;
    }
}

// ---- shared pieces -------------------------------------------------------------
// A workgroup is 8 wave64s (two per SIMD).  Wave (wr, wc), wr in 0..1,
// wc in 0..3, owns rows (i = 2 wr + {0,1}) x (32 virtuals a) and the column
// block (j = wc) x (32 virtuals b) of both T1 and T2: 4 x 2 MFMA 16x16 tiles
// each.  Row block rb = (i offset rb/2, a half rb%2); column block cb = b half.
constexpr int WG_THREADS = 512;

struct PairTile {
    int IB, JB, AC, BC;
};

template <typename ACC>
__device__ inline void energy_epilogue(const ACC (&acc1)[4][2], const ACC (&acc2)[4][2],
                                       const PairTile& t, int o, int v,
                                       const double* __restrict__ eo,
                                       const double* __restrict__ ev, double& sos, double& sx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const int j = t.JB * PI + wc;
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const int i = t.IB * PI + 2 * wr + rb / 2;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const int b = t.BC * PA + cb * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = t.AC * PA + (rb % 2) * 16 + AccLayout<ACC>::row(lane, r);
                if (i < o && j < o && a < v && b < v) {
                    const double m1 = (double)acc1[rb][cb][r], m2 = (double)acc2[rb][cb][r];
                    const double dinv = 1.0 / (ev[a] + ev[b] - eo[i] - eo[j]);
                    sos += (m1 * m1 + m2 * m2) * dinv;
                    sx += m1 * m2 * dinv;
                }
            }
        }
    }
}

__device__ inline bool decode_tile(int nIB, int nAC, int64_t nWG, int64_t per, PairTile& t,
                                   int64_t& id) {
    id = xcd_logical(per);
    if (id >= nWG) return false;
    const int64_t nAB = (int64_t)nAC * (nAC + 1) / 2;
    tri_decode(id / nAB, nIB, t.IB, t.JB);
    tri_decode(id % nAB, nAC, t.AC, t.BC);
    return true;
}

__device__ inline double block_sum8(double v, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < WG_THREADS / 64; ++w) s += sh[w];
    return s;
}

__device__ inline void write_partial(double sos, double sx, const PairTile& t, int64_t id,
                                     double* sh, double* __restrict__ partial) {
    const double w = (t.IB < t.JB ? 2.0 : 1.0) * (t.AC < t.BC ? 1.0 : 0.5);
    const double os = block_sum8(sos, sh);
    const double x = block_sum8(sx, sh);
    if (threadIdx.x == 0) {
        partial[id * 2 + 0] = -w * os;
        partial[id * 2 + 1] = 2.0 * w * x;
    }
}

// ---- FP64 / FP32 kernel --------------------------------------------------------
// KS: K depth of one stage (32 KiB of the four panels in LDS).
template <typename T> struct RealMfma;
template <> struct RealMfma<double> {
    typedef corr_d4 acc_t;
    static constexpr int KS = 8;    // 2 MFMA k-steps of 64 cycles per stage
    __device__ static inline acc_t mma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct RealMfma<float> {
    typedef corr_f4 acc_t;
    static constexpr int KS = 16;   // 4 MFMA k-steps of 32 cycles per stage
    __device__ static inline acc_t mma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// ---- BF16x3 kernel -------------------------------------------------------------
// M ~= hi.hi + hi.lo + lo.hi with hi = bf16(x), lo = bf16(x - hi), FP32 accumulate,
// on v_mfma_f32_32x32x16_bf16: lane l holds A[l%32][8(l/32) + t] (t < 8), and
// accumulator element r of lane l is C[8(r/4) + 4(l/32) + r%4][l%32].
typedef float corr_f16v __attribute__((ext_vector_type(16)));

__device__ inline corr_f16v mma_bf16_32(const corr_u4& a, const corr_u4& b, corr_f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(corr_bf8, a),
                                                   __builtin_bit_cast(corr_bf8, b), c, 0, 0, 0);
}

// K advances in 16-deep stages; a stage of one panel is [hi/lo][128 rows][16 k]
// bf16 (8 KiB), stored in the panel exactly as in LDS: a row is two 16-byte
// k-chunks, chunk ch of row r in slot bf16_slot(r, ch).  Stages move global ->
// LDS with global_load_lds_dwordx4 (no VGPR round trip, no ds_write): each wave
// copies one (panel, hi/lo) half, 4 contiguous 1 KiB segments.  Four stage
// buffers (4 x 32 KiB, separate LDS objects) keep the copies three stages ahead
// of the stage the matrix cores consume, with one barrier per stage.
typedef __attribute__((address_space(3))) void* corr_lds_ptr;
typedef __attribute__((address_space(1))) const void* corr_gbl_ptr;

typedef uint16_t Bf16Stage[4][2][PW][16];   // [panel][hi/lo][row][k slot]

__device__ inline corr_u4 lds_frag(const Bf16Stage& S, int p, int hl, int row, int ko) {
    return *reinterpret_cast<const corr_u4*>(&S[p][hl][row][ko]);
}

// Copies this wave's half-panel of stage st into S (4 x 1 KiB, one instruction each).
__device__ inline void bf16_stage_copy(Bf16Stage& S, const uint16_t* Pw, int pw, int hw, int64_t st, int lane) {
    constexpr int SUB = 2 * PW * 16;
    const uint16_t* src = Pw + st * SUB + hw * (PW * 16) + lane * 8;
#pragma unroll
    for (int s = 0; s < 4; ++s)
        __builtin_amdgcn_global_load_lds((corr_gbl_ptr)(src + s * 512), (corr_lds_ptr)&S[pw][hw][32 * s][0],
                                         16, 0, 0);
}

// One stage of this wave's four 32 x 32 tiles: T1_x = rows ra[x] of panel 0 x
// rows rb of panel 1, T2_x = rows rb of panel 2 x rows ra[x] of panel 3.
__device__ inline void bf16_stage_compute(const Bf16Stage& S, const int (&ra)[2], int rb, int ko,
                                          corr_f16v (&acc1)[2], corr_f16v (&acc2)[2]) {
    corr_u4 ah[2], al[2], dh[2], dl[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        ah[x] = lds_frag(S, 0, 0, ra[x], ko);
        al[x] = lds_frag(S, 0, 1, ra[x], ko);
        dh[x] = lds_frag(S, 3, 0, ra[x], ko);
        dl[x] = lds_frag(S, 3, 1, ra[x], ko);
    }
    const corr_u4 bh = lds_frag(S, 1, 0, rb, ko), bl = lds_frag(S, 1, 1, rb, ko);
    const corr_u4 ch = lds_frag(S, 2, 0, rb, ko), cl = lds_frag(S, 2, 1, rb, ko);
    // The three products of one accumulator are dependent: issue them term by
    // term across the four accumulators, so dependent MFMAs are four apart.
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        acc1[x] = mma_bf16_32(al[x], bh, acc1[x]);
        acc2[x] = mma_bf16_32(cl, dh[x], acc2[x]);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        acc1[x] = mma_bf16_32(ah[x], bl, acc1[x]);
        acc2[x] = mma_bf16_32(ch, dl[x], acc2[x]);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        acc1[x] = mma_bf16_32(ah[x], bh, acc1[x]);
        acc2[x] = mma_bf16_32(ch, dh[x], acc2[x]);
    }
}

// Waits until this wave's outstanding global->LDS copies number at most N.
#define CORR_WAIT_VMCNT(N) __builtin_amdgcn_s_waitcnt((N) | (0x7 << 4) | (0xF << 8))
// Stage barrier: this wave's LDS reads are complete and at most N of its copies
// are outstanding (s_waitcnt vmcnt(N) lgkmcnt(0)), then s_barrier.  Unlike
// __syncthreads() it does not wait for the younger copies still in flight; the
// empty asm statements keep the compiler from moving LDS accesses across it.
#define CORR_STAGE_BARRIER(N)                                   \
    {                                                           \
        __builtin_amdgcn_s_waitcnt((N) | (0x7 << 4));           \
        __asm__ volatile("" ::: "memory");                      \
        __builtin_amdgcn_s_barrier();                           \
        __asm__ volatile("" ::: "memory");                      \
    }

// Wave (wr, wc), wr in 0..1, wc in 0..3: T1_x = M^{ij}_ab for i = 2 wr + x,
// j = wc (rows a, columns b, all 32 x 32 of the tile's virtual blocks), and
// T2_x = B[:, (J,A) row j]^T B[:, (I,B) row i] = M^{ji}_ab = M^{ij}_ba, so the
// exchange partner of every element sits in the same lane and register.
// Accumulation is FP32 on two levels: the MFMA accumulators are added into
// FP32 totals and cleared every FLUSH groups of four 16-deep stages (64 K
// values), which shortens each rounding chain from 456 MFMA steps (152
// stages at naux = 2432) to 48 plus 10 total updates.  On the (H2O)32 / def2-SVP inputs the error of the energy
// against FP64 drops from 0.77 to 0.48 uEh (0.47 with FP64 totals, which take
// twice the registers and ran 2.5% slower; one level: 0.77 uEh).  The totals
// take 64 VGPRs (206 in all, no spills, 2 waves/SIMD as the 128 KiB of LDS
// sets).
constexpr int BF16X3_FLUSH = 4;

template <int FLUSH>
__global__ __launch_bounds__(WG_THREADS) void rimp2_energy_bf16x3(
    const uint16_t* __restrict__ Bp, const double* __restrict__ eo, const double* __restrict__ ev,
    int o, int v, int nIB, int nAC, int64_t Kc, int64_t nWG, int64_t per,
    double* __restrict__ partial) {
    constexpr int SUB = 2 * PW * 16;            // bf16 per panel per stage (hi + lo)
    __shared__ __attribute__((aligned(16))) Bf16Stage L0;
    __shared__ __attribute__((aligned(16))) Bf16Stage L1;
    __shared__ __attribute__((aligned(16))) Bf16Stage L2;
    __shared__ __attribute__((aligned(16))) Bf16Stage L3;
    __shared__ double sh[WG_THREADS / 64];

    PairTile t;
    int64_t id;
    if (!decode_tile(nIB, nAC, nWG, per, t, id)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 2, wc = wave & 3, r32 = lane & 31, kh = lane >> 5;

    // Panels: 0 = (I,A), 1 = (J,B), 2 = (J,A), 3 = (I,B).  Wave w copies panel w/2, half w%2.
    const int pw = wave >> 1, hw = wave & 1;
    const int ibw = (pw == 0 || pw == 3) ? t.IB : t.JB;
    const int acw = (pw == 0 || pw == 2) ? t.AC : t.BC;
    const uint16_t* Pw = Bp + ((int64_t)ibw * nAC + acw) * Kc * SUB;

    const int ra[2] = {(2 * wr) * PA + r32, (2 * wr + 1) * PA + r32};
    const int rb = wc * PA + r32;
    // Rows read by this lane all share r % 32 = r32, so they share the chunk slot.
    const int ko = bf16_slot(r32, kh) * 8;

    corr_f16v acc1[2], acc2[2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int r = 0; r < 16; ++r) { acc1[x][r] = 0.0f; acc2[x][r] = 0.0f; }
    // FP32 totals of the accumulators, flushed every FLUSH 4-stage groups
    float d1[FLUSH > 0 ? 2 : 1][FLUSH > 0 ? 16 : 1], d2[FLUSH > 0 ? 2 : 1][FLUSH > 0 ? 16 : 1];
    if constexpr (FLUSH > 0) {
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int r = 0; r < 16; ++r) { d1[x][r] = 0.0f; d2[x][r] = 0.0f; }
    }
    auto flush = [&]() {
        if constexpr (FLUSH > 0) {
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    d1[x][r] += acc1[x][r];
                    d2[x][r] += acc2[x][r];
                    acc1[x][r] = 0.0f;
                    acc2[x][r] = 0.0f;
                }
        }
    };

    // Stage s lives in buffer s % 4.  Before the barrier that opens stage s,
    // every wave has finished reading stage s-1 and has landed its copy of
    // stage s; after it, the buffer of stage s-1 takes the copy of stage s+3.
    // Kc is a multiple of 4 (naux is padded to 64).
    // Copies in flight at each barrier: this wave's 4 instructions per stage;
    // stage s+1 must have landed while s+2 and s+3 may still run (vmcnt 8).
    // The copy of each step is unconditional (no branch for the compiler's
    // wait insertion to merge): the last three steps re-copy stage Kc-1 into
    // the buffer just released, which is never read again.
    bf16_stage_copy(L0, Pw, pw, hw, 0, lane);
    bf16_stage_copy(L1, Pw, pw, hw, Kc > 1 ? 1 : Kc - 1, lane);
    bf16_stage_copy(L2, Pw, pw, hw, Kc > 2 ? 2 : Kc - 1, lane);
    CORR_STAGE_BARRIER(8)
#define CORR_STAGE(CUR, NEXT3, s_)                                                       \
    {                                                                                    \
        bf16_stage_copy(NEXT3, Pw, pw, hw, ((s_) + 3 < Kc) ? (s_) + 3 : Kc - 1, lane);   \
        bf16_stage_compute(CUR, ra, rb, ko, acc1, acc2);                                 \
        CORR_STAGE_BARRIER(8)                                                            \
    }
    for (int64_t s = 0; s < Kc; s += 4) {
        CORR_STAGE(L0, L3, s)
        CORR_STAGE(L1, L0, s + 1)
        CORR_STAGE(L2, L1, s + 2)
        CORR_STAGE(L3, L2, s + 3)
        if constexpr (FLUSH > 0) {
            if ((s / 4 + 1) % FLUSH == 0) flush();
        }
    }
#undef CORR_STAGE
    CORR_WAIT_VMCNT(0);   // no copy outlives the loop
    flush();

    double sos = 0.0, sx = 0.0;
    const int j = t.JB * PI + wc;
    const int b = t.BC * PA + r32;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
        const int i = t.IB * PI + 2 * wr + x;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int a = t.AC * PA + 8 * (r / 4) + 4 * kh + (r % 4);
            if (i < o && j < o && a < v && b < v) {
                const double m1 = FLUSH > 0 ? (double)d1[x][r] : (double)acc1[x][r];
                const double m2 = FLUSH > 0 ? (double)d2[x][r] : (double)acc2[x][r];
                const double dinv = 1.0 / (ev[a] + ev[b] - eo[i] - eo[j]);
                sos += (m1 * m1 + m2 * m2) * dinv;
                sx += m1 * m2 * dinv;
            }
        }
    }
    write_partial(sos, sx, t, id, sh, partial);
}
// ---- FP64 / FP32 kernel: stages -----------------------------------------------
// A stage holds KS k-rows of the four 128-wide panels, [panel][k][128], moved
// global -> LDS by global_load_lds_dwordx4 (32 x 1 KiB per stage, four per
// wave), four stage buffers deep as in the BF16x3 kernel.  Row k is stored
// rotated by 16 (k % 4) elements, which places the four k-rows that one MFMA
// step reads in distinct bank groups (what a 16-element row pad did when the
// stages went through registers).  The rotation is applied on the source
// side: lane l of a copy writes LDS bytes [16 l, 16 l + 16) of its 1 KiB and
// loads the global chunk that belongs there.
template <typename T> using RealStage = T[4][RealMfma<T>::KS][PW];

__device__ inline int real_rot(int k, int c) { return (c + 16 * (k & 3)) & (PW - 1); }

template <typename T>
__device__ inline void real_stage_copy(RealStage<T>& S, const T* const (&P)[4], int64_t st, int wave, int lane) {
    constexpr int KS = RealMfma<T>::KS, EPC = 16 / sizeof(T);   // elements per 16-byte chunk
    constexpr int RPC = 1024 / (PW * (int)sizeof(T));             // k-rows per 1 KiB copy
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int q = wave * 4 + s;                 // copy q of 32: panel q / 8, KiB q % 8
        const int p = q >> 3, k0 = (q & 7) * RPC;
        const int k = k0 + lane / (64 / RPC);
        const int pos = (lane % (64 / RPC)) * EPC;  // stored column of this lane's chunk
        const int c = (pos - 16 * (k & 3)) & (PW - 1);
        const T* src = P[p] + (st * KS + k) * (int64_t)PW + c;
        __builtin_amdgcn_global_load_lds((corr_gbl_ptr)src, (corr_lds_ptr)&S[p][k0][0], 16, 0, 0);
    }
}

// Wave (wr, wc) as in the BF16x3 kernel, on 16 x 16 MFMA tiles: rows
// i = 2 wr + {0,1} x 32 virtuals a (4 row tiles), columns j = wc x 32
// virtuals b (2 column tiles), for T1 and its exchange partner T2.
template <typename T>
__device__ inline void real_stage_compute(const RealStage<T>& S, const int (&c1)[4], int c2, const int (&c4)[2][2],
                                          int lane, typename RealMfma<T>::acc_t (&acc1)[4][2],
                                          typename RealMfma<T>::acc_t (&acc2)[4][2]) {
    constexpr int KS = RealMfma<T>::KS;
#pragma unroll
    for (int kk = 0; kk < KS; kk += 4) {
        const int kr = kk + (lane >> 4);
        T f1[4], f2[2], f3[2], f4[2][2];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) f1[rb] = S[0][kr][real_rot(kr, c1[rb])];
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            f2[y] = S[1][kr][real_rot(kr, c2 + y * 16)];
            f3[y] = S[2][kr][real_rot(kr, c2 + y * 16)];
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) f4[x][y] = S[3][kr][real_rot(kr, c4[x][y])];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                acc1[rb][cb] = RealMfma<T>::mma(f1[rb], f2[cb], acc1[rb][cb]);
                acc2[rb][cb] = RealMfma<T>::mma(f3[rb % 2], f4[rb / 2][cb], acc2[rb][cb]);
            }
    }
}

template <typename T>
__global__ __launch_bounds__(WG_THREADS) void rimp2_energy_real(
    const T* __restrict__ Bp, const double* __restrict__ eo, const double* __restrict__ ev,
    int o, int v, int nIB, int nAC, int64_t Kp, int64_t nWG, int64_t per,
    double* __restrict__ partial) {
    constexpr int KS = RealMfma<T>::KS;
    __shared__ __attribute__((aligned(16))) RealStage<T> L0;
    __shared__ __attribute__((aligned(16))) RealStage<T> L1;
    __shared__ __attribute__((aligned(16))) RealStage<T> L2;
    __shared__ __attribute__((aligned(16))) RealStage<T> L3;
    __shared__ double sh[WG_THREADS / 64];
    typedef typename RealMfma<T>::acc_t acc_t;

    PairTile t;
    int64_t id;
    if (!decode_tile(nIB, nAC, nWG, per, t, id)) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 2, wc = wave & 3, cl = lane & 15;

    const T* const P[4] = {Bp + ((int64_t)t.IB * nAC + t.AC) * Kp * PW,    // (I, A)
                           Bp + ((int64_t)t.JB * nAC + t.BC) * Kp * PW,    // (J, B)
                           Bp + ((int64_t)t.JB * nAC + t.AC) * Kp * PW,    // (J, A)
                           Bp + ((int64_t)t.IB * nAC + t.BC) * Kp * PW};   // (I, B)

    acc_t acc1[4][2], acc2[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) { acc1[x][y] = acc_t{0, 0, 0, 0}; acc2[x][y] = acc_t{0, 0, 0, 0}; }
    // FP32: accumulation on two levels, as in the BF16x3 kernel.  The MFMA
    // accumulators are added into FP32 totals and cleared every 64 K values,
    // so each rounding chain is 16 MFMA steps plus one total update per 64 K
    // values instead of naux / 4 steps.
    constexpr bool TWO_LEVEL = std::is_same<T, float>::value;
    acc_t tot1[TWO_LEVEL ? 4 : 1][TWO_LEVEL ? 2 : 1], tot2[TWO_LEVEL ? 4 : 1][TWO_LEVEL ? 2 : 1];
    if constexpr (TWO_LEVEL) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) { tot1[x][y] = acc_t{0, 0, 0, 0}; tot2[x][y] = acc_t{0, 0, 0, 0}; }
    }
    auto flush = [&]() {
        if constexpr (TWO_LEVEL) {
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    tot1[x][y] += acc1[x][y];
                    tot2[x][y] += acc2[x][y];
                    acc1[x][y] = acc_t{0, 0, 0, 0};
                    acc2[x][y] = acc_t{0, 0, 0, 0};
                }
        }
    };

    // fragment columns inside the 128-wide panels
    int c1[4], c4[2][2];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) c1[rb] = (2 * wr + rb / 2) * PA + (rb % 2) * 16 + cl;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) c4[x][y] = (2 * wr + x) * PA + y * 16 + cl;
    const int c2 = wc * PA + cl;

    // Stage s lives in buffer s % 4; the schedule of copies and barriers is
    // the BF16x3 kernel's (copies three stages ahead, vmcnt 8 at each
    // barrier, the last three steps re-copying the final stage into a buffer
    // that is never read again).  nK is a multiple of 4 (Kp is one of 64).
    const int64_t nK = Kp / KS;
    real_stage_copy<T>(L0, P, 0, wave, lane);
    real_stage_copy<T>(L1, P, nK > 1 ? 1 : nK - 1, wave, lane);
    real_stage_copy<T>(L2, P, nK > 2 ? 2 : nK - 1, wave, lane);
    CORR_STAGE_BARRIER(8)
#define CORR_RSTAGE(CUR, NEXT3, s_)                                                          \
    {                                                                                        \
        real_stage_copy<T>(NEXT3, P, ((s_) + 3 < nK) ? (s_) + 3 : nK - 1, wave, lane);       \
        real_stage_compute<T>(CUR, c1, c2, c4, lane, acc1, acc2);                            \
        CORR_STAGE_BARRIER(8)                                                                \
    }
    constexpr int FLUSH_GROUPS = 64 / (4 * KS);   // groups of four stages per 64 K values (FP32: 1)
    for (int64_t s = 0; s < nK; s += 4) {
        CORR_RSTAGE(L0, L3, s)
        CORR_RSTAGE(L1, L0, s + 1)
        CORR_RSTAGE(L2, L1, s + 2)
        CORR_RSTAGE(L3, L2, s + 3)
        if constexpr (TWO_LEVEL) {
            if ((s / 4 + 1) % FLUSH_GROUPS == 0) flush();
        }
    }
#undef CORR_RSTAGE
    CORR_WAIT_VMCNT(0);   // no copy outlives the loop
    double sos = 0.0, sx = 0.0;
    if constexpr (TWO_LEVEL) {
        flush();
        energy_epilogue(tot1, tot2, t, o, v, eo, ev, sos, sx);
    } else {
        energy_epilogue(acc1, acc2, t, o, v, eo, ev, sos, sx);
    }
    write_partial(sos, sx, t, id, sh, partial);
}

#undef CORR_WAIT_VMCNT
#undef CORR_STAGE_BARRIER

// ---- host side -------------------------------------------------------------------
struct RiLayout {
    int nIB, nAC;
    int64_t Kp, nWG, per;
};

static RiLayout ri_layout(int64_t naux, int64_t o, int64_t v) {
    RiLayout l;
    l.nIB = (int)cdiv(o, PI);
    l.nAC = (int)cdiv(v, PA);
    l.Kp = rup(naux, 64);
    l.nWG = (int64_t)l.nIB * (l.nIB + 1) / 2 * ((int64_t)l.nAC * (l.nAC + 1) / 2);
    l.per = cdiv(l.nWG, 8);
    return l;
}

// Repacks Bia into the panel layout for the given precision; returns the
// device pointer in *out.
int ri_repack(corr_ctx* ctx, const RiLayout& l, int64_t naux, int o, int v, const double* Bia,
              int prec, const void** out) {
    const int64_t elems = (int64_t)l.nIB * l.nAC * l.Kp * PW;
    const size_t esz = prec == CORR_PREC_FP64 ? 8 : 4;   // bf16 hi+lo = 4 B
    void* Bp = nullptr;
    CORR_TRY(scratch(ctx, "rimp2_panels", elems * esz, &Bp));
    if (prec == CORR_PREC_BF16X3) {
        const int64_t Kc = l.Kp / 32;
        const int64_t blocks = (int64_t)l.nIB * l.nAC * Kc;
        hipLaunchKernelGGL(repack_bf16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, Bia,
                           (uint16_t*)Bp, naux, o, v, l.nAC, Kc);
    } else {
        const int64_t grid = std::min<int64_t>(cdiv(elems, 256), 65536 * 4);
        if (prec == CORR_PREC_FP64)
            hipLaunchKernelGGL(repack_real<double>, dim3((unsigned)grid), dim3(256), 0, ctx->stream,
                               Bia, (double*)Bp, naux, o, v, l.nAC, l.Kp, elems);
        else
            hipLaunchKernelGGL(repack_real<float>, dim3((unsigned)grid), dim3(256), 0, ctx->stream,
                               Bia, (float*)Bp, naux, o, v, l.nAC, l.Kp, elems);
    }
    CORR_LAUNCHED(ctx, "rimp2 repack");
    *out = Bp;
    return CORR_OK;
}

int rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia,
                 const double* eo, const double* ev, int prec, corr_mp2_energy* out) {
    CORR_REQUIRE(ctx, naux > 0 && nocc > 0 && nvir > 0, "rimp2: sizes must be positive");
    CORR_REQUIRE(ctx, nocc < (1 << 20) && nvir < (1 << 20), "rimp2: nocc/nvir too large");
    CORR_REQUIRE(ctx, Bia && eo && ev && out, "rimp2: null pointer");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32 || prec == CORR_PREC_BF16X3,
                 "rimp2: unknown precision");
    const RiLayout l = ri_layout(naux, nocc, nvir);
    CORR_REQUIRE(ctx, l.per * 8 < ((int64_t)1 << 31), "rimp2: too many tiles");
    const void* Bp = nullptr;
    CORR_TRY(ri_repack(ctx, l, naux, (int)nocc, (int)nvir, Bia, prec, &Bp));
    double* partial = nullptr;
    CORR_TRY(scratch(ctx, "rimp2_partial", sizeof(double) * 2 * l.nWG, (void**)&partial));
    const dim3 grid((unsigned)(l.per * 8)), block(WG_THREADS);
    KernelTimer timer(ctx, "rimp2_energy");
    if (prec == CORR_PREC_FP64)
        hipLaunchKernelGGL(rimp2_energy_real<double>, grid, block, 0, ctx->stream, (const double*)Bp,
                           eo, ev, (int)nocc, (int)nvir, l.nIB, l.nAC, l.Kp, l.nWG, l.per, partial);
    else if (prec == CORR_PREC_FP32)
        hipLaunchKernelGGL(rimp2_energy_real<float>, grid, block, 0, ctx->stream, (const float*)Bp,
                           eo, ev, (int)nocc, (int)nvir, l.nIB, l.nAC, l.Kp, l.nWG, l.per, partial);
    else
        hipLaunchKernelGGL(rimp2_energy_bf16x3<BF16X3_FLUSH>, grid, block, 0, ctx->stream, (const uint16_t*)Bp,
                           eo, ev, (int)nocc, (int)nvir, l.nIB, l.nAC, l.Kp / 16, l.nWG, l.per, partial);
    CORR_LAUNCHED(ctx, "rimp2_energy");
    timer.stop();
    double s[2];
    CORR_TRY(reduce_sum(ctx, partial, l.nWG, 2, s));
    out->e_os = s[0];
    out->e_ss = s[0] + s[1];
    out->e_mp2 = out->e_os + out->e_ss;
    out->e_scs = out->e_ss / 3.0 + out->e_os * 6.0 / 5.0;
    return CORR_OK;
}

}  // namespace corr

extern "C" int corr_rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                                 const double* d_Bia, const double* d_eps_occ,
                                 const double* d_eps_vir, int prec, corr_mp2_energy* h_out) {
    CORR_CHECK_CTX(ctx);
    return corr::rimp2_energy(ctx, naux, nocc, nvir, d_Bia, d_eps_occ, d_eps_vir, prec, h_out);
}
