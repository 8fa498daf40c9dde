// RI-MP2 energy: fused pair-block GEMM + FP64 denominator + deterministic
// FP64 reduction.  Restates run_canonical_mp2_ri (src/POSTHF/can_mp2.cc:53-132):
//
//   M^{ij}_ab = sum_Q B^Q_ia B^Q_jb
//   EMP2_OS  = - sum_{ij ab} M_ab^2 / D,   EMP2_SS = - sum M_ab (M_ab - M_ba) / D
//   D = e_a + e_b - e_i - e_j
//
// The reference loops over all o^2 ordered pairs with one dgemm each.  Here
// the o*v x o*v matrix G[(i,a),(j,b)] = M^{ij}_ab is tiled.  A workgroup owns
// a tile T1 = G[I x A, J x B] and its exchange partner T2[(i,a),(j,b)] =
// M^{ij}_ba computed as B[:, J x A]^T B[:, I x B], so that M_ab and M_ba of
// every element meet in one workgroup.  Only occupied pairs i <= j and virtual
// blocks A <= B are computed; weights restore the reference's ordered-pair sum
// exactly: x2 for i < j (the (j,i) pair gives the same sum with a<->b), x1/2 on
// A == B tiles (T1 and T2 cover the same elements).
//
// FP64 / FP32 (rimp2_energy_real): 128 x 128 tiles, rows (4 occupied) x (32
// virtual), 8 waves each holding T1 and T2 of its elements in registers.
// BF16x3 (rimp2_energy_bf16x3_wide): 128 x 256 tiles (occupied 4-block x
// 8-block), 16 waves, T1 and T2 on separate waves, swapped through LDS once
// per tile; per-pair weights handle the 4 x 8 blocks that straddle i = j.
//
// B is repacked once into panels of 4 occupied x 32 virtual columns over K:
// a contiguous [Kp][128] slab (FP64/FP32), or [Kp/16][hi/lo][128][2][8] of
// BF16 halves (BF16x3) already in the kernel's LDS order, so every K-slab a
// workgroup stages into LDS is one contiguous, fully coalesced read, moved by
// direct global->LDS copies.
#include "common.hpp"
#include "wide_tiles.hpp"

#include <algorithm>
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

// One block per (panel, 16-deep stage): read [16 q][128 c] coalesced along a,
// write the stage as [hi/lo][128 c][2 slots][8 q] in the energy kernels' LDS
// order (slot s of row c holds k-chunk bf16_slot(c, s)).  Stages of a panel
// are consecutive: panel blk, stage ks at Bp + (blk Ks + ks) 4096.
__global__ __launch_bounds__(256) void repack_bf16(const double* __restrict__ Bia,
                                                   uint16_t* __restrict__ Bp, int64_t naux, int o,
                                                   int v, int nAC, int64_t Ks) {
    __shared__ double S[16][PW + 1];
    const int64_t ks = blockIdx.x % Ks;
    const int64_t blk = blockIdx.x / Ks;
    const int ib = (int)(blk / nAC), ac = (int)(blk % nAC);
    for (int e = threadIdx.x; e < 16 * PW; e += 256) {
        const int q = e / PW, c = e % PW;
        const int64_t Q = ks * 16 + q;
        const int i = ib * PI + c / PA, a = ac * PA + c % PA;
        S[q][c] = (Q < naux && i < o && a < v) ? Bia[(Q * o + i) * (int64_t)v + a] : 0.0;
    }
    __syncthreads();
    uint16_t* out = Bp + (blk * Ks + ks) * (2 * PW * 16);
    {
        const int e = threadIdx.x;                  // 256 = [hi/lo][128 rows][... one 16-byte slot]
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int r0 = e + 256 * half;          // 0..511 = [hi/lo][128][2 slots]
            const int hl = r0 / (PW * 2), c = (r0 % (PW * 2)) / 2, kq = bf16_slot(c, r0 % 2);
            uint16_t hv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const double x = S[kq * 8 + t][c];
                const uint16_t hi = f32_to_bf16_rn((float)x);
                if (hl == 0) hv[t] = hi;
                else hv[t] = f32_to_bf16_rn((float)(x - (double)bf16_to_f32(hi)));
            }
            uint4 pk;
            pk.x = hv[0] | ((uint32_t)hv[1] << 16);
            pk.y = hv[2] | ((uint32_t)hv[3] << 16);
            pk.z = hv[4] | ((uint32_t)hv[5] << 16);
            pk.w = hv[6] | ((uint32_t)hv[7] << 16);
            *reinterpret_cast<uint4*>(out + (int64_t)r0 * 8) = pk;
        }
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

// ---- BF16x3 pieces --------------------------------------------------------------
// M ~= hi.hi + hi.lo + lo.hi with hi = bf16(x), lo = bf16(x - hi), FP32 accumulate,
// on v_mfma_f32_32x32x16_bf16: lane l holds A[l%32][8(l/32) + t] (t < 8), and
// accumulator element r of lane l is C[8(r/4) + 4(l/32) + r%4][l%32].
typedef float corr_f16v __attribute__((ext_vector_type(16)));

__device__ inline corr_f16v mma_bf16_32(const corr_u4& a, const corr_u4& b, corr_f16v c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(corr_bf8, a),
                                                   __builtin_bit_cast(corr_bf8, b), c, 0, 0, 0);
}

typedef __attribute__((address_space(3))) void* corr_lds_ptr;
typedef __attribute__((address_space(1))) const void* corr_gbl_ptr;

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

// ---- BF16x3 kernel, wide tile ------------------------------------------------
// A workgroup owns the occupied 4-block I, the occupied 8-block Jg (two
// 4-blocks J0 = 2 Jg, J1 = 2 Jg + 1) and the virtual blocks A <= B: T1 =
// M^{ij}_ab on (I,A) x (Jg,B), 128 x 256, and its exchange partner T2 of the
// same shape, from six panels (I,A), (J0,B), (J1,B), (J0,A), (J1,A), (I,B):
// 48 KiB per 16-deep stage for 384 MFMAs, where a 128 x 128 tile pair moves
// 32 KiB for 192.  Sixteen waves (four per SIMD, 128 registers each): waves 0-7
// hold T1 and waves 8-15 T2 in 2 x 2 blocks of 32 x 32 MFMA tiles, accumulated
// on one FP32 level.  Pairs (i, j) carry their own weight: 2 for i < j, 1 for
// i = j, 0 for i > j (the 4 x 8 occupied tiles straddle the diagonal).
// Three stage buffers, copies two stages ahead.  The tile of each workgroup
// comes from a list built on the device (ushort4: I, Jg, z, w): z < w is the
// virtual pair (A, B) = (z, w); z > w holds the two diagonal blocks A = w and
// A' = z (the T1 waves compute M^{ij} on A x A, the T2 waves on A' x A', from
// the same six-panel stages, and each wave finds M_ba in its own block);
// z = w is the diagonal block A alone (T1 and T2 then cover the same
// elements, each at half weight).
constexpr int WIDE_THREADS = 1024;
typedef uint16_t Bf16Stage6[6][2][PW][16];   // [panel][hi/lo][row][k slot]

__device__ inline corr_u4 lds_frag6(const Bf16Stage6& S, int p, int hl, int row, int ko) {
    return *reinterpret_cast<const corr_u4*>(&S[p][hl][row][ko]);
}

// NC global -> LDS copies of 1 KiB of a stage: src[k] + st * 8 KiB -> LDS offset seg[k].
template <int NC>
__device__ inline void wide_stage_copy(Bf16Stage6& S, const uint16_t* const (&src)[NC], const int (&seg)[NC],
                                       int64_t st) {
    constexpr int SUB = 2 * PW * 16;
#pragma unroll
    for (int k = 0; k < NC; ++k)
        __builtin_amdgcn_global_load_lds((corr_gbl_ptr)(src[k] + st * SUB),
                                         (corr_lds_ptr)(&S[0][0][0][0] + seg[k]), 16, 0, 0);
}

// NA x NB blocks: acc[p][q] += rows rA[p] of panel pa x rows rB[q] of panel pb.
template <int NA, int NB>
__device__ inline void wide_stage_compute(const Bf16Stage6& S, int pa, int pb, const int (&rA)[NA],
                                          const int (&rB)[NB], int ko, corr_f16v (&acc)[NA][NB]) {
    corr_u4 a[NA][2], b[NB][2];
#pragma unroll
    for (int p = 0; p < NA; ++p) a[p][1] = lds_frag6(S, pa, 1, rA[p], ko);
#pragma unroll
    for (int q = 0; q < NB; ++q) b[q][0] = lds_frag6(S, pb, 0, rB[q], ko);
#pragma unroll
    for (int p = 0; p < NA; ++p) a[p][0] = lds_frag6(S, pa, 0, rA[p], ko);
#pragma unroll
    for (int q = 0; q < NB; ++q) b[q][1] = lds_frag6(S, pb, 1, rB[q], ko);
    // lo.hi, hi.lo, then hi.hi (small terms first)
#pragma unroll
    for (int p = 0; p < NA; ++p)
#pragma unroll
        for (int q = 0; q < NB; ++q) acc[p][q] = mma_bf16_32(a[p][1], b[q][0], acc[p][q]);
#pragma unroll
    for (int p = 0; p < NA; ++p)
#pragma unroll
        for (int q = 0; q < NB; ++q) acc[p][q] = mma_bf16_32(a[p][0], b[q][1], acc[p][q]);
#pragma unroll
    for (int p = 0; p < NA; ++p)
#pragma unroll
        for (int q = 0; q < NB; ++q) acc[p][q] = mma_bf16_32(a[p][0], b[q][0], acc[p][q]);
}

// The K loop: stage s lives in buffer s % 3.  Iteration s copies stage s+2
// into buffer (s+2) % 3, which held stage s-1 (read before the barrier closing
// iteration s-1), computes stage s, then waits until this wave's copies of
// stage s+1 have landed (the NC of s+2 may still run) and its reads are back,
// and meets the barrier.  Copies past the end re-copy stage Kc-1 (into a
// buffer no later stage reads).  Ends with no copy in flight.
template <int NC, int NA, int NB, typename Pro>
__device__ inline void wide_k_loop(Bf16Stage6& L0, Bf16Stage6& L1, Bf16Stage6& L2, const uint16_t* const (&src)[NC],
                                   const int (&seg)[NC], int64_t Kc, int pa, int pb, const int (&rA)[NA],
                                   const int (&rB)[NB], int ko, corr_f16v (&acc)[NA][NB], Pro pro) {
    wide_stage_copy<NC>(L0, src, seg, 0);
    wide_stage_copy<NC>(L1, src, seg, Kc > 1 ? 1 : Kc - 1);
    CORR_STAGE_BARRIER(NC)
    pro();   // the caller's prologue work (older loads have landed: no wait for the copies)
#define CORR_WIDE_STAGE(CUR, COPY, s_)                                                        \
    {                                                                                         \
        wide_stage_copy<NC>(COPY, src, seg, ((s_) + 2 < Kc) ? (s_) + 2 : Kc - 1);             \
        wide_stage_compute<NA, NB>(CUR, pa, pb, rA, rB, ko, acc);                             \
        CORR_STAGE_BARRIER(NC)                                                                \
    }
    // whole groups of three stages in the loop, the one or two left after it: a
    // branch inside the loop would make the compiler wait for every copy in
    // flight before the next LDS read
    const int64_t K3 = Kc - Kc % 3;
    for (int64_t s = 0; s < K3; s += 3) {
        CORR_WIDE_STAGE(L0, L2, s)
        CORR_WIDE_STAGE(L1, L0, s + 1)
        CORR_WIDE_STAGE(L2, L1, s + 2)
    }
    if (K3 < Kc) CORR_WIDE_STAGE(L0, L2, K3)
    if (K3 + 1 < Kc) CORR_WIDE_STAGE(L1, L0, K3 + 1)
#undef CORR_WIDE_STAGE
    CORR_WAIT_VMCNT(0);
}

// 1/D for D > 0: FP32 reciprocal refined by two FP64 Newton steps.  D = +inf
// (an infinite orbital energy) gives 0, as IEEE division does; NaN stays NaN.
__device__ inline double rcp_pos(double d) {
    const double x0 = (double)__builtin_amdgcn_rcpf((float)d);
    double x = fma(x0, fma(-d, x0, 1.0), x0);
    x = fma(x, fma(-d, x, 1.0), x);
    return x0 == 0.0 ? 0.0 : x;
}

// One element's share of the opposite-spin and exchange sums, weight w / D.
__device__ inline void wide_energy_term(double m1, double m2, double ea, double eij, double w, bool valid,
                                        double& sos, double& sx) {
    const double dinv = valid ? w * rcp_pos(ea + eij) : 0.0;
    sos += (m1 * m1 + m2 * m2) * dinv;
    sx += m1 * m2 * dinv;
}

// One tile of the wide kernel; PAIR2: the tile holds two diagonal blocks.
template <bool PAIR2>
__device__ __forceinline__ void wide_tile(Bf16Stage6& L0, Bf16Stage6& L1, Bf16Stage6& L2, double* sh, double* evAB,
                                          const uint16_t* Bp,
                                          ushort4 tl, const double* __restrict__ eo, const double* __restrict__ ev,
                                          int o, int v, int nIB, int nAC, int64_t Kc, int64_t id,
                                          double* __restrict__ partial) {
    constexpr int SUB = 2 * PW * 16;
    const int IB = tl.x & WT_IMASK, J0 = tl.y, AC = PAIR2 ? tl.w : tl.z, BC = PAIR2 ? tl.z : tl.w;
    const bool j1ok = !(tl.x & WT_NOJ1);
    const int J1 = j1ok ? J0 + ((tl.x & WT_STEP2) ? 2 : 1) : J0;   // no J1: J0's panel again, weight 0
    // loaded first, stored into evAB once the first stage copies are issued
    // (the wait then covers this load only) and read after the K loop
    const int t32 = threadIdx.x & 31;
    const int evb = ((threadIdx.x & 63) < PA ? AC : BC) * PA + t32;
    const double ev_own = ev[evb < v ? evb : v - 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r32 = lane & 31, kh = lane >> 5;
    const int ko = bf16_slot(r32, kh) * 8;

    // panel pointers: 0 (I,A), 1 (J0,B), 2 (J1,B), 3 (J0,A), 4 (J1,A), 5 (I,B), or for
    // two diagonal blocks 0 (I,A), 1 (J0,A), 2 (J1,A), 3 (J0,A'), 4 (J1,A'), 5 (I,A');
    auto panel = [=](int p) -> const uint16_t* {
        const int ob = (p == 0 || p == 5) ? IB : (p == 1 || p == 3) ? J0 : J1;
        const bool first = PAIR2 ? p < 3 : (p == 0 || p == 3 || p == 4);
        const int vb = first ? AC : BC;
        return Bp + ((int64_t)ob * nAC + vb) * Kc * SUB;
    };
    const double wab = AC < BC && !PAIR2 ? 1.0 : 0.5;
    double sos = 0.0, sx = 0.0;

    {
        const int half = wave >> 3, t8 = wave & 7, ip = t8 >> 2, jp = t8 & 3;
        const uint16_t* src[3];
        int seg[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int c = 3 * wave + k;                  // copy c of 48: panel c / 8, KiB c % 8
            src[k] = panel(c >> 3) + (c & 7) * 512 + lane * 8;
            seg[k] = (c >> 3) * SUB + (c & 7) * 512;
        }
        // T1 wave: acc[p][q] = (I,A) rows i = 2ip+p  x  (Jg,B) rows j = 2jp+q.
        // T2 wave: acc[p][q] = (Jg,A) rows j = 2jp+p  x  (I,B) rows i = 2ip+q, i.e.
        //          M^{ij}_ba for T1's block (x, y) = (q, p).
        const int pa = half ? 3 + (jp >> 1) : 0;
        const int pb = half ? 5 : 1 + (jp >> 1);
        const int ra0 = half ? (2 * jp) & 3 : 2 * ip;
        const int rb0 = half ? 2 * ip : (2 * jp) & 3;
        const int rA[2] = {ra0 * PA + r32, (ra0 + 1) * PA + r32};
        const int rB[2] = {rb0 * PA + r32, (rb0 + 1) * PA + r32};
        corr_f16v acc[2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[p][q][r] = 0.0f;
        wide_k_loop<3, 2, 2>(L0, L1, L2, src, seg, Kc, pa, pb, rA, rB, ko, acc, [&]() {
            if (threadIdx.x < 2 * PA) evAB[threadIdx.x] = ev_own;
        });
        if constexpr (PAIR2) {
            // Two diagonal blocks: each wave holds whole 32 x 32 blocks, M^{ij} on
            // A x A (T1 waves) or A' x A' (T2 waves), and M_ba of an element is the
            // transposed element of its own block.  Block by block a wave stores
            // the block (element (r, c) at r * 33 + c: 4.1 KiB, six waves per 48 KiB
            // buffer; the odd row stride keeps the transposed reads conflict-free)
            // and reads M_ab and M_ba back.  Each (a, b) and (b, a) is visited, at
            // half weight.
            constexpr int XB = 32 * 33;
            float* X = reinterpret_cast<float*>(wave < 6 ? &L0[0][0][0][0] : wave < 12 ? &L1[0][0][0][0]
                                                                                      : &L2[0][0][0][0]) +
                       (wave % 6) * XB;
            const int vb = (half ? BC : AC) * PA;
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int al = 8 * (r / 4) + 4 * kh + (r % 4);
                        X[al * 33 + r32] = acc[p][q][r];
                    }
                    __syncthreads();
                    // T1 wave: rows i = 2ip+p, columns j = 2jp+q; T2 wave: rows j = 2jp+p, columns i = 2ip+q
                    const int i = IB * PI + 2 * ip + (half ? q : p);
                    const int jb = jp < 2 ? J0 : J1, j = jb * PI + ((2 * jp + (half ? p : q)) & 3);
                    const double w = wt_pair_weight(IB, jb, i, j, jp < 2 || j1ok) * 0.5;
                    if (w > 0.0 && i < o && j < o && vb + r32 < v) {
                        const double eij = ev[vb + r32] - eo[i] - eo[j];
#pragma unroll 4
                        for (int r = 0; r < 16; ++r) {
                            const int al = 8 * (r / 4) + 4 * kh + (r % 4);
                            const float m = X[al * 33 + r32], mt = X[r32 * 33 + al];
                            wide_energy_term((double)m, (double)mt, evAB[half * PA + al], eij, w, vb + al < v,
                                             sos, sx);
                        }
                    }
                }
        } else {
            // The energy of block (x, y) of wave pair t8 (i = 2ip+x, j = 2jp+y) is
            // taken by the T1 wave for x = 0 and by the T2 wave for x = 1, so the two
            // swap one block row through LDS: the T1 wave stores its x = 1 blocks,
            // the T2 wave its x = 0 blocks (M_ba of T1's x = 0), 16 KiB per wave
            // pair, three pairs per 48 KiB buffer.
            __syncthreads();
            float* X = reinterpret_cast<float*>(t8 < 3 ? &L0[0][0][0][0] : t8 < 6 ? &L1[0][0][0][0]
                                                                                 : &L2[0][0][0][0]) +
                       (t8 % 3) * 4096;
#pragma unroll
            for (int y = 0; y < 2; ++y)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if (half) X[2048 + (y * 16 + r) * 64 + lane] = acc[y][0][r];
                    else X[(y * 16 + r) * 64 + lane] = acc[1][y][r];
                }
            __syncthreads();
            const int b = BC * PA + r32;
            const int i = IB * PI + 2 * ip + half;
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int jb = jp < 2 ? J0 : J1, j = jb * PI + ((2 * jp + y) & 3);
                const double w = wt_pair_weight(IB, jb, i, j, jp < 2 || j1ok) * wab;
                if (w > 0.0 && i < o && j < o && b < v) {
                    const double eij = ev[b] - eo[i] - eo[j];
                    float mo[16];   // T1 wave: M_ba from the T2 wave; T2 wave: M_ab from the T1 wave
#pragma unroll
                    for (int r = 0; r < 16; ++r) mo[r] = X[(half ? 0 : 2048) + (y * 16 + r) * 64 + lane];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const double mr = half ? (double)acc[y][1][r] : (double)acc[0][y][r];
                        const int al = 8 * (r / 4) + 4 * kh + (r % 4);
                        wide_energy_term(mr, (double)mo[r], evAB[al], eij, w, AC * PA + al < v, sos, sx);
                    }
                }
            }
        }
    }
    // fixed-order workgroup sums
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sos += __shfl_xor(sos, off, 64);
        sx += __shfl_xor(sx, off, 64);
    }
    __syncthreads();
    if (lane == 0) sh[wave] = sos;
    __syncthreads();
    double tos = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < WIDE_THREADS / 64; ++w) tos += sh[w];
    __syncthreads();
    if (lane == 0) sh[wave] = sx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double tx = 0.0;
        for (int w = 0; w < WIDE_THREADS / 64; ++w) tx += sh[w];
        partial[id * 2 + 0] = -tos;
        partial[id * 2 + 1] = 2.0 * tx;
    }
}

__global__ __launch_bounds__(WIDE_THREADS) void rimp2_energy_bf16x3_wide(
    const uint16_t* __restrict__ Bp, const ushort4* __restrict__ tiles, const double* __restrict__ eo,
    const double* __restrict__ ev, int o, int v, int nIB, int nAC, int64_t Kc, int64_t nT, int64_t per,
    double* __restrict__ partial) {
    // three separate buffers: the compiler can then tell that reads of one do
    // not overlap the copies into another, and needs no wait for them
    __shared__ __attribute__((aligned(16))) Bf16Stage6 L0;
    __shared__ __attribute__((aligned(16))) Bf16Stage6 L1;
    __shared__ __attribute__((aligned(16))) Bf16Stage6 L2;
    __shared__ double sh[WIDE_THREADS / 64];
    __shared__ double evAB[2 * PA];   // eps_vir of the blocks A and B (or A'), read after the K loop

    const int64_t id = xcd_logical(per);
    if (id >= nT) return;
    const ushort4 tl = tiles[id];
    // the two kinds of tile are separate code: nothing is live across this branch
    if (tl.z > tl.w) wide_tile<true>(L0, L1, L2, sh, evAB, Bp, tl, eo, ev, o, v, nIB, nAC, Kc, id, partial);
    else wide_tile<false>(L0, L1, L2, sh, evAB, Bp, tl, eo, ev, o, v, nIB, nAC, Kc, id, partial);
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

// Gradient outputs of the pair-tile kernel (GRAD = true): R = [(c_os + c_ss)
// M_ab - c_ss M_ba] / D into the symmetric ov x ov matrix R (rows (i a)), and
// 72 FP64 partials of dE/deps per tile: [A block 32 | B block 32 | I block 4 |
// J block 4].
struct RGradArgs {
    double* R;
    double* peps;
    double cs, css;
};

// Epilogue of the gradient: dE/deps partials (a 128 x 128 FP64 matrix of
// g = w N / D^2 in LDS, row and column sums in a fixed order), then the four
// symmetric positions of R for every element pair: R1 = R[(i a)][(j b)] and
// R2^T = R[(j a)][(i b)] straight from the registers, R2 = R[(i b)][(j a)] and
// R1^T = R[(j b)][(i a)] through a per-wave 32 x 33 LDS transpose (rows of 32
// contiguous doubles).  A position reached twice (A = B, i = j, and the i > j
// pairs of I = J tiles) is written once.
template <typename ACC>
__device__ inline void grad_epilogue(const ACC (&acc1)[4][2], const ACC (&acc2)[4][2], const PairTile& t, int o,
                                     int v, const double* __restrict__ eo, const double* __restrict__ ev,
                                     const RGradArgs& g, double* b0, double* b1, double* b2, double* b3,
                                     int64_t id, double& sos, double& sx) {
    // the four stage buffers are separate LDS objects: pick one by index
    auto Lb = [&](int k) { return k == 0 ? b0 : k == 1 ? b1 : k == 2 ? b2 : b3; };
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    const int j = t.JB * PI + wc;
    const double wt = (t.IB < t.JB ? 2.0 : 1.0) * (t.AC < t.BC ? 1.0 : 0.5);
    const int64_t ov = (int64_t)o * v;
    // phase 1: energy sums and g into LDS: row (il a) -> buffer il, row a of it,
    // column (wc b) stored at (col + a) & 127 (bank rotation)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
        const int il = 2 * wr + rb / 2, i = t.IB * PI + il;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
            const int bl = cb * 16 + (lane & 15), b = t.BC * PA + bl;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int al = (rb % 2) * 16 + AccLayout<ACC>::row(lane, r), a = t.AC * PA + al;
                double gv = 0.0;
                if (i < o && j < o && a < v && b < v) {
                    const double m1 = (double)acc1[rb][cb][r], m2 = (double)acc2[rb][cb][r];
                    const double dinv = 1.0 / (ev[a] + ev[b] - eo[i] - eo[j]);
                    sos += (m1 * m1 + m2 * m2) * dinv;
                    sx += m1 * m2 * dinv;
                    const double nt = g.cs * (m1 * m1 + m2 * m2) - 2.0 * g.css * m1 * m2;
                    gv = wt * nt * dinv * dinv;
                }
                Lb(il)[al * PW + ((wc * PA + bl + al) & (PW - 1))] = gv;
            }
        }
    }
    __syncthreads();
    // row sums (threads 0..127: row (il a)) and column sums (128..255: column (jl b))
    __shared__ double sums[2 * PW];
    if (threadIdx.x < 2 * PW) {
        double s = 0.0;
        if (threadIdx.x < PW) {
            const int row = threadIdx.x, il = row / PA, al = row % PA;
            for (int c = 0; c < PW; ++c) s += Lb(il)[al * PW + ((c + al) & (PW - 1))];
        } else {
            const int col = threadIdx.x - PW;
            for (int row = 0; row < PW; ++row) {
                const int il = row / PA, al = row % PA;
                s += Lb(il)[al * PW + ((col + al) & (PW - 1))];
            }
        }
        sums[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < 72) {
        const int k = threadIdx.x;
        double s = 0.0;
        if (k < 32) {
            for (int il = 0; il < PI; ++il) s += sums[il * PA + k];                 // eps_a, a in A
        } else if (k < 64) {
            for (int jl = 0; jl < PI; ++jl) s += sums[PW + jl * PA + (k - 32)];     // eps_b, b in B
        } else if (k < 68) {
            for (int al = 0; al < PA; ++al) s -= sums[(k - 64) * PA + al];          // eps_i, i in I
        } else {
            for (int bl = 0; bl < PA; ++bl) s -= sums[PW + (k - 68) * PA + bl];     // eps_j, j in J
        }
        g.peps[id * 72 + k] = s;
    }
    __syncthreads();
    // phase 2: R, one occupied pair (i = 2wr + p, j) at a time
    double* V = Lb(wave / 3) + (wave % 3) * (PA * (PA + 1));
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        const int i = t.IB * PI + 2 * wr + p;
        const bool live = i < o && j < o && !(t.IB == t.JB && i > j);
        const bool w1t = live && !(i == j && t.AC == t.BC);
        const bool w2 = live && t.AC != t.BC && i != j;
        double r1[2][2][4], r2[2][2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int rb = 2 * p + h;
                const int bl = cb * 16 + (lane & 15), b = t.BC * PA + bl;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int al = h * 16 + AccLayout<ACC>::row(lane, r), a = t.AC * PA + al;
                    const double m1 = (double)acc1[rb][cb][r], m2 = (double)acc2[rb][cb][r];
                    const bool ok = live && a < v && b < v;
                    const double dinv = ok ? 1.0 / (ev[a] + ev[b] - eo[i] - eo[j]) : 0.0;
                    r1[h][cb][r] = (g.cs * m1 - g.css * m2) * dinv;
                    r2[h][cb][r] = (g.cs * m2 - g.css * m1) * dinv;
                    if (ok) {
                        g.R[((int64_t)i * v + a) * ov + (int64_t)j * v + b] = r1[h][cb][r];
                        if (w2) g.R[((int64_t)j * v + a) * ov + (int64_t)i * v + b] = r2[h][cb][r];
                    }
                }
            }
        // transposed positions through LDS: V[a][b], read along a
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            __syncthreads();
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int al = h * 16 + AccLayout<ACC>::row(lane, r), bl = cb * 16 + (lane & 15);
                        V[al * (PA + 1) + bl] = pass == 0 ? r2[h][cb][r] : r1[h][cb][r];
                    }
            __syncthreads();
            const bool wr_ok = pass == 0 ? w2 : w1t;
            const int al = lane & 31, a = t.AC * PA + al;
            if (wr_ok && a < v) {
                for (int bb = lane >> 5; bb < PA; bb += 2) {
                    const int b = t.BC * PA + bb;
                    if (b >= v) continue;
                    const double x = V[al * (PA + 1) + bb];
                    if (pass == 0) g.R[((int64_t)i * v + b) * ov + (int64_t)j * v + a] = x;   // R2 at (i b),(j a)
                    else g.R[((int64_t)j * v + b) * ov + (int64_t)i * v + a] = x;            // R1^T at (j b),(i a)
                }
            }
        }
    }
}

template <typename T, bool GRAD>
__global__ __launch_bounds__(WG_THREADS) void rimp2_energy_real(
    const T* __restrict__ Bp, const double* __restrict__ eo, const double* __restrict__ ev,
    int o, int v, int nIB, int nAC, int64_t Kp, int64_t nWG, int64_t per,
    double* __restrict__ partial, RGradArgs gargs) {
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
    if constexpr (GRAD) {
        __syncthreads();   // the stage buffers become the epilogue's LDS
        double* b0 = reinterpret_cast<double*>(&L0[0][0][0]);
        double* b1 = reinterpret_cast<double*>(&L1[0][0][0]);
        double* b2 = reinterpret_cast<double*>(&L2[0][0][0]);
        double* b3 = reinterpret_cast<double*>(&L3[0][0][0]);
        if constexpr (TWO_LEVEL) {
            flush();
            grad_epilogue(tot1, tot2, t, o, v, eo, ev, gargs, b0, b1, b2, b3, id, sos, sx);
        } else {
            grad_epilogue(acc1, acc2, t, o, v, eo, ev, gargs, b0, b1, b2, b3, id, sos, sx);
        }
    } else if constexpr (TWO_LEVEL) {
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

// kstage: K padding of the panels (64 for the 4-buffer kernels, 16 for the
// wide kernel: whole 16-deep stages).
static RiLayout ri_layout(int64_t naux, int64_t o, int64_t v, int64_t kstage = 64) {
    RiLayout l;
    l.nIB = (int)cdiv(o, PI);
    l.nAC = (int)cdiv(v, PA);
    l.Kp = rup(naux, kstage);
    l.nWG = (int64_t)l.nIB * (l.nIB + 1) / 2 * ((int64_t)l.nAC * (l.nAC + 1) / 2);
    l.per = cdiv(l.nWG, 8);
    return l;
}

// One workgroup of 1024 threads builds the wide kernel's tile list
// (wide_tiles.hpp): thread t takes a contiguous chunk of the NS super-tiles,
// counts its tiles, a workgroup scan gives its offsets, and it writes its
// tiles in order, the diagonal ones after all the others.
__global__ __launch_bounds__(1024) void wide_tile_list(ushort4* __restrict__ tiles, WtGrid g) {
    __shared__ int64_t sc[1024], sd[1024];
    const int t = threadIdx.x;
    const int64_t C = (g.NS + 1023) / 1024;
    const int64_t s0 = std::min<int64_t>(g.NS, t * C), s1 = std::min<int64_t>(g.NS, s0 + C);
    int64_t n = 0, nd = 0;   // this thread's tiles with A < B, and its diagonal tiles
    for (int64_t st = s0; st < s1; ++st) {
        int c, d;
        wt_count(g, st, c, d);
        n += c - d;
        nd += d;
    }
    sc[t] = n;
    sd[t] = nd;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t x = t >= off ? sc[t - off] : 0, y = t >= off ? sd[t - off] : 0;
        __syncthreads();
        sc[t] += x;
        sd[t] += y;
        __syncthreads();
    }
    int64_t k = sc[t] - n, kd = sc[1023] + sd[t] - nd;   // exclusive offsets
    for (int64_t st = s0; st < s1; ++st)
        wt_emit(g, st, [&](int x, int y, int z, int w, bool diag) {
            tiles[diag ? kd++ : k++] =
                make_ushort4((unsigned short)x, (unsigned short)y, (unsigned short)z, (unsigned short)w);
        });
}

struct WideTiles {
    WtGrid g;
    int64_t nT;
};

static WideTiles wide_tiles(int nIB, int nAC) {
    WideTiles w;
    w.g = wt_grid(nIB, nAC, w.nT);
    return w;
}

struct RimpPlan {
    RiLayout l;
    WideTiles wt;    // BF16x3 only
    size_t b_panels, b_partial, b_tiles;
};

static RimpPlan rimp2_plan(int64_t naux, int64_t o, int64_t v, int prec) {
    RimpPlan p;
    const bool bf = prec == CORR_PREC_BF16X3;
    p.l = ri_layout(naux, o, v, bf ? 16 : 64);
    const size_t esz = prec == CORR_PREC_FP64 ? 8 : 4;   // bf16 hi+lo = 4 B
    p.b_panels = (size_t)p.l.nIB * p.l.nAC * p.l.Kp * PW * esz;
    if (bf) {
        p.wt = wide_tiles(p.l.nIB, p.l.nAC);
        p.b_partial = sizeof(double) * 2 * p.wt.nT;
        p.b_tiles = sizeof(ushort4) * p.wt.nT;
    } else {
        p.wt = WideTiles{WtGrid{0, 0, 0, 0, 0, 0}, 0};
        p.b_partial = sizeof(double) * 2 * p.l.nWG;
        p.b_tiles = 0;
    }
    return p;
}

// Repacks Bia into the panel layout for the given precision at Bp.
static int ri_repack(corr_ctx* ctx, const RiLayout& l, int64_t naux, int o, int v, const double* Bia, int prec,
                     void* Bp) {
    const int64_t elems = (int64_t)l.nIB * l.nAC * l.Kp * PW;
    if (prec == CORR_PREC_BF16X3) {
        const int64_t Ks = l.Kp / 16;
        const int64_t blocks = (int64_t)l.nIB * l.nAC * Ks;
        hipLaunchKernelGGL(repack_bf16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, Bia,
                           (uint16_t*)Bp, naux, o, v, l.nAC, Ks);
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
    return CORR_OK;
}

// Zeroes uint4 runs: run b (of nrun) covers [b * stride + off, + len) uint4.
__global__ void zero_runs_kernel(uint4* __restrict__ p, int64_t nrun, int64_t stride, int64_t off, int64_t len) {
    const int64_t total = nrun * len;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
        p[(e / len) * stride + off + e % len] = make_uint4(0u, 0u, 0u, 0u);
}

// The energy.  ws == nullptr: workspace from the context's cached scratch;
// otherwise the caller's (sized by rimp2_workspace_bytes).  timed: bracket
// the energy kernel with events when the context's timing is on.  prepacked:
// the BF16x3 panels in the context scratch already hold B (filled through
// rimp2_bf16_panels); Bia is not read.
static int rimp2_energy_impl(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia,
                             const double* eo, const double* ev, int prec, corr_mp2_energy* out, const RimpWs* ws,
                             bool timed, bool prepacked = false) {
    CORR_REQUIRE(ctx, naux > 0 && nocc > 0 && nvir > 0, "rimp2: sizes must be positive");
    CORR_REQUIRE(ctx, nocc < (1 << 20) && nvir < (1 << 20), "rimp2: nocc/nvir too large");
    CORR_REQUIRE(ctx, (Bia || prepacked) && eo && ev && out, "rimp2: null pointer");
    CORR_REQUIRE(ctx, !prepacked || (prec == CORR_PREC_BF16X3 && !ws), "rimp2: prepacked panels are BF16x3 scratch");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32 || prec == CORR_PREC_BF16X3,
                 "rimp2: unknown precision");
    const bool bf = prec == CORR_PREC_BF16X3;
    CORR_REQUIRE(ctx, !bf || cdiv(nocc, PI) <= WT_IMASK, "rimp2: nocc too large for the BF16x3 tile list");
    const RimpPlan p = rimp2_plan(naux, nocc, nvir, prec);
    const RiLayout& l = p.l;
    const int64_t nWG = bf ? p.wt.nT : l.nWG, per = cdiv(nWG, 8);
    CORR_REQUIRE(ctx, per * 8 < ((int64_t)1 << 31), "rimp2: too many tiles");
    RimpWs w;
    if (ws) {
        w = *ws;
    } else {
        CORR_TRY(scratch(ctx, "rimp2_panels", p.b_panels, &w.panels));
        CORR_TRY(scratch(ctx, "rimp2_partial", p.b_partial, (void**)&w.partial));
        if (bf) CORR_TRY(scratch(ctx, "rimp2_tiles", p.b_tiles, (void**)&w.tiles));
        CORR_TRY(scratch(ctx, "reduce_out", sizeof(double) * 16, (void**)&w.red_out));
        CORR_TRY(scratch(ctx, "reduce_mid", sizeof(double) * 16 * REDUCE_NCHUNK, (void**)&w.red_mid));
    }
    if (!prepacked) CORR_TRY(ri_repack(ctx, l, naux, (int)nocc, (int)nvir, Bia, prec, w.panels));
    if (bf) {
        hipLaunchKernelGGL(wide_tile_list, dim3(1), dim3(1024), 0, ctx->stream, w.tiles, p.wt.g);
        CORR_LAUNCHED(ctx, "wide_tile_list");
    }
    {
        KernelTimer timer(ctx, "rimp2_energy", timed);
        if (bf)
            hipLaunchKernelGGL(rimp2_energy_bf16x3_wide, dim3((unsigned)(per * 8)), dim3(WIDE_THREADS), 0,
                               ctx->stream, (const uint16_t*)w.panels, w.tiles, eo, ev, (int)nocc, (int)nvir, l.nIB,
                               l.nAC, l.Kp / 16, nWG, per, w.partial);
        else if (prec == CORR_PREC_FP64)
            hipLaunchKernelGGL((rimp2_energy_real<double, false>), dim3((unsigned)(per * 8)), dim3(WG_THREADS),
                               0, ctx->stream, (const double*)w.panels, eo, ev, (int)nocc, (int)nvir, l.nIB, l.nAC,
                               l.Kp, nWG, per, w.partial, RGradArgs{nullptr, nullptr, 0.0, 0.0});
        else
            hipLaunchKernelGGL((rimp2_energy_real<float, false>), dim3((unsigned)(per * 8)), dim3(WG_THREADS),
                               0, ctx->stream, (const float*)w.panels, eo, ev, (int)nocc, (int)nvir, l.nIB, l.nAC,
                               l.Kp, nWG, per, w.partial, RGradArgs{nullptr, nullptr, 0.0, 0.0});
        CORR_LAUNCHED(ctx, "rimp2_energy");
    }
    double s[2];
    CORR_TRY(reduce_sum_ws(ctx, w.partial, nWG, 2, s, w.red_mid, w.red_out));
    out->e_os = s[0];
    out->e_ss = s[0] + s[1];
    out->e_mp2 = out->e_os + out->e_ss;
    out->e_scs = out->e_ss / 3.0 + out->e_os * 6.0 / 5.0;
    return CORR_OK;
}

int rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia,
                 const double* eo, const double* ev, int prec, corr_mp2_energy* out) {
    return rimp2_energy_impl(ctx, naux, nocc, nvir, Bia, eo, ev, prec, out, nullptr, true);
}

int rimp2_bf16_panels(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, Bf16Panels* out) {
    CORR_REQUIRE(ctx, naux > 0 && nocc > 0 && nvir > 0 && nocc < (1 << 20) && nvir < (1 << 20) && out,
                 "rimp2 panels: bad arguments");
    const RimpPlan p = rimp2_plan(naux, nocc, nvir, CORR_PREC_BF16X3);
    const RiLayout& l = p.l;
    void* panels = nullptr;
    CORR_TRY(scratch(ctx, "rimp2_panels", p.b_panels, &panels));
    const int64_t Ks = l.Kp / 16, nblk = (int64_t)l.nIB * l.nAC, stage = 2 * PW * 16 / 8;   // uint4 per stage
    // a producer writes rows Q < Kp of the o x v columns; zero the rest: the
    // stages from naux / 16 on (rows past naux) and, when o or v leave part of
    // a panel empty, every panel whole
    const bool ragged = nocc % PI != 0 || nvir % PA != 0;
    const int64_t ks0 = ragged ? 0 : naux / 16;
    if (ks0 < Ks) {
        const int64_t len = (Ks - ks0) * stage;
        const int64_t grid = std::min<int64_t>(cdiv(nblk * len, 256), 65536);
        hipLaunchKernelGGL(zero_runs_kernel, dim3((unsigned)grid), dim3(256), 0, ctx->stream, (uint4*)panels, nblk,
                           Ks * stage, ks0 * stage, len);
        CORR_LAUNCHED(ctx, "zero_runs_kernel");
    }
    *out = Bf16Panels{(uint16_t*)panels, (int)nocc, (int)nvir, l.nAC, l.Kp, Ks};
    return CORR_OK;
}

int rimp2_energy_prepacked(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* eo,
                           const double* ev, corr_mp2_energy* out) {
    return rimp2_energy_impl(ctx, naux, nocc, nvir, nullptr, eo, ev, CORR_PREC_BF16X3, out, nullptr, true, true);
}

// dE/deps from the per-tile partials, in a fixed order: block t < o sums,
// for occupied i = t, the I-slots of tiles (I = i/4, J >= I) and the J-slots
// of tiles (I <= J = i/4) over every virtual pair; block o + a likewise the
// A-slots of tiles (A = a/32, B >= A) and the B-slots of (A <= B = a/32).
// Thread k takes the terms k, k + 256, ...; a fixed LDS tree adds the 256.
__device__ inline int64_t tri_index(int64_t r, int64_t c, int64_t n) { return r * n - r * (r - 1) / 2 + (c - r); }

__global__ __launch_bounds__(256) void grad_eps_reduce(const double* __restrict__ peps, int nIB, int nAC, int o,
                                                       int v, double* __restrict__ deo, double* __restrict__ dev) {
    __shared__ double sh[256];
    const int t = blockIdx.x, k = threadIdx.x;
    const int64_t nAB = (int64_t)nAC * (nAC + 1) / 2, nIJ = (int64_t)nIB * (nIB + 1) / 2;
    double s = 0.0;
    if (t < o) {
        const int Y = t / PI, li = t % PI;
        const int64_t n1 = (int64_t)(nIB - Y) * nAB, n = n1 + (int64_t)(Y + 1) * nAB;
        for (int64_t e = k; e < n; e += 256) {
            if (e < n1) {
                const int64_t J = Y + e / nAB;
                s += peps[(tri_index(Y, J, nIB) * nAB + e % nAB) * 72 + 64 + li];
            } else {
                const int64_t I = (e - n1) / nAB;
                s += peps[(tri_index(I, Y, nIB) * nAB + (e - n1) % nAB) * 72 + 68 + li];
            }
        }
    } else {
        const int a = t - o, X = a / PA, la = a % PA;
        const int64_t n = nIJ * (nAC + 1);
        for (int64_t e = k; e < n; e += 256) {
            const int64_t ij = e / (nAC + 1), r = e % (nAC + 1);
            if (r < nAC - X) s += peps[(ij * nAB + tri_index(X, X + r, nAC)) * 72 + la];
            else s += peps[(ij * nAB + tri_index(r - (nAC - X), X, nAC)) * 72 + 32 + la];
        }
    }
    sh[k] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (k < w) sh[k] += sh[k + w];
        __syncthreads();
    }
    if (k == 0) {
        if (t < o) deo[t] = sh[0];
        else dev[t - o] = sh[0];
    }
}

// The pair-tile part of the gradient: R (symmetric ov x ov, rows (i a)),
// dE/deps and the energies, for E = c_os EMP2_OS + c_ss EMP2_SS; FP64 or FP32
// (two-level accumulation) MFMA.  dE/dB = -4 B R is left to the caller.
int rimp2_grad_tiles(corr_ctx* ctx, int64_t naux, int64_t o, int64_t v, const double* Bia, const double* eo,
                     const double* ev, double c_os, double c_ss, int prec, double* R, double* deo, double* dev,
                     corr_mp2_energy* out) {
    const int p = prec == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32;
    const RimpPlan pl = rimp2_plan(naux, o, v, p);
    const RiLayout& l = pl.l;
    const int64_t nWG = l.nWG, per = cdiv(nWG, 8);
    CORR_REQUIRE(ctx, per * 8 < ((int64_t)1 << 31), "rimp2_gradient: too many tiles");
    void* panels = nullptr;
    double *partial = nullptr, *peps = nullptr;
    CORR_TRY(scratch(ctx, "rimp2_panels", pl.b_panels, &panels));
    CORR_TRY(scratch(ctx, "rimp2_partial", pl.b_partial, (void**)&partial));
    CORR_TRY(scratch(ctx, "grad_peps", sizeof(double) * 72 * nWG, (void**)&peps));
    CORR_TRY(ri_repack(ctx, l, naux, (int)o, (int)v, Bia, p, panels));
    const RGradArgs g{R, peps, c_os + c_ss, c_ss};
    if (p == CORR_PREC_FP64)
        hipLaunchKernelGGL((rimp2_energy_real<double, true>), dim3((unsigned)(per * 8)), dim3(WG_THREADS), 0,
                           ctx->stream, (const double*)panels, eo, ev, (int)o, (int)v, l.nIB, l.nAC, l.Kp, nWG, per,
                           partial, g);
    else
        hipLaunchKernelGGL((rimp2_energy_real<float, true>), dim3((unsigned)(per * 8)), dim3(WG_THREADS), 0,
                           ctx->stream, (const float*)panels, eo, ev, (int)o, (int)v, l.nIB, l.nAC, l.Kp, nWG, per,
                           partial, g);
    CORR_LAUNCHED(ctx, "rimp2_energy_real (gradient)");
    hipLaunchKernelGGL(grad_eps_reduce, dim3((unsigned)(o + v)), dim3(256), 0, ctx->stream, peps, l.nIB,
                       l.nAC, (int)o, (int)v, deo, dev);
    CORR_LAUNCHED(ctx, "grad_eps_reduce");
    double s[2];
    CORR_TRY(reduce_sum(ctx, partial, nWG, 2, s));
    if (out) {
        out->e_os = s[0];
        out->e_ss = s[0] + s[1];
        out->e_mp2 = out->e_os + out->e_ss;
        out->e_scs = out->e_ss / 3.0 + out->e_os * 6.0 / 5.0;
    }
    return CORR_OK;
}

int64_t rimp2_workspace(int64_t naux, int64_t nocc, int64_t nvir, int prec, void* base, RimpWs* ws) {
    if (naux <= 0 || nocc <= 0 || nvir <= 0 || nocc >= (1 << 20) || nvir >= (1 << 20)) return -1;
    if (prec != CORR_PREC_FP64 && prec != CORR_PREC_FP32 && prec != CORR_PREC_BF16X3) return -1;
    const RimpPlan p = rimp2_plan(naux, nocc, nvir, prec);
    size_t off = 0;
    auto take = [&](size_t bytes) -> void* {
        void* q = base ? (void*)((char*)base + off) : nullptr;
        off += (bytes + 255) / 256 * 256;
        return q;
    };
    RimpWs w;
    w.panels = take(p.b_panels);
    w.partial = (double*)take(p.b_partial);
    w.tiles = (ushort4*)take(p.b_tiles);
    w.red_mid = (double*)take(sizeof(double) * 16 * REDUCE_NCHUNK);
    w.red_out = (double*)take(sizeof(double) * 16);
    if (ws) *ws = w;
    return (int64_t)off;
}

int rimp2_energy_with(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia, const double* eo,
                      const double* ev, int prec, corr_mp2_energy* out, const RimpWs* ws) {
    return rimp2_energy_impl(ctx, naux, nocc, nvir, Bia, eo, ev, prec, out, ws, false);
}

}  // namespace corr

extern "C" int corr_rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                                 const double* d_Bia, const double* d_eps_occ,
                                 const double* d_eps_vir, int prec, corr_mp2_energy* h_out) {
    CORR_CHECK_CTX(ctx);
    return corr::rimp2_energy(ctx, naux, nocc, nvir, d_Bia, d_eps_occ, d_eps_vir, prec, h_out);
}
