// Gradient of the RI-MP2 energy with respect to B^Q_ia and the orbital
// energies (no counterpart in the reference, which only prints energies).
//
//   E = c_os EMP2_OS + c_ss EMP2_SS = - sum_{ij ab} N^{ij}_ab / D^{ij}_ab
//   M^{ij}_ab = sum_Q B^Q_ia B^Q_jb,  D = e_a + e_b - e_i - e_j
//   N^{ij}_ab = (c_os + c_ss) M_ab^2 - c_ss M_ab M_ba
//   R^{ij}_ab = [(c_os + c_ss) M_ab - c_ss M_ba] / D
//   dE/dB^Q_ia = -4 sum_{jb} R^{ij}_ab B^Q_jb
//   dE/de_a = 2 sum_{ijb} N^{ij}_ab / D^2,   dE/de_i = -2 sum_{jab} N^{ij}_ab / D^2
//
// Up to 48 GB of R (o v <= ~77000) the pair-tile MFMA kernel of the energy
// recomputes M on its symmetric (I <= J, A <= B) tiles and writes all of R
// and per-tile dE/de partials from registers; one GEMM gives dE/dB = -4 B R.
// Larger problems: occupied orbitals are processed in blocks I: the pair blocks
// G_I[(i a)][(j b)] = M^{ij}_ab for i in I are recomputed with one MFMA GEMM
// (never the full o^2 v^2 set), a fused kernel turns them into R_I and the
// FP64 partial sums for dE/de and E, and one more GEMM adds -4 B R_I^T to
// the I columns of dE/dB.  All reductions are in a fixed order.
#include "common.hpp"

namespace corr {

namespace {

constexpr int RT = 32;   // (a, b) tile edge of the pair-block kernel
// Largest R (ov x ov doubles) that the fused path holds at once; larger
// problems (or R above half of a context scratch limit) take the
// occupied-blocked path with slabs of ~6 GB (a quarter of the limit).
constexpr double FUSED_R_BYTES = 48.0e9;

struct GradArgs {
    const double* G;      // [bo*v][o*v] pair blocks of this occupied block
    double* R;            // same layout, output
    const double* eo;
    const double* ev;
    int o, v, bo, i0, nT;
    double cs, css;       // c_os + c_ss, c_ss
    double* pv;           // [bo*o][nT][v]  partials of dE/de_a (summed over b in a tile)
    double* pij;          // [bo*o][nT*nT][3]: dE/de_i, E_OS, E_SS partials
};

// One block per (pair (i,j), tile (ta, tb)).  Stages the (a in ta, b in tb)
// tile of M^{ij} and its transpose (b in tb, a in ta) of G through LDS so that
// both reads and the R write are row-contiguous.
__global__ __launch_bounds__(256) void rimp2_grad_pairs_kernel(GradArgs g) {
    __shared__ double Mab[RT][RT + 1];
    __shared__ double Mba[RT][RT + 1];
    __shared__ double red[3][256];
    const int nt2 = g.nT * g.nT;
    const int64_t blk = blockIdx.x;
    const int tile = (int)(blk % nt2);
    const int64_t pair = blk / nt2;          // ii * o + j
    const int ta = tile / g.nT, tb = tile % g.nT;
    const int ii = (int)(pair / g.o), j = (int)(pair % g.o);
    const int i = g.i0 + ii;
    const int64_t ld = (int64_t)g.o * g.v;
    const int tid = threadIdx.x;
    for (int e = tid; e < RT * RT; e += 256) {
        const int r = e / RT, c = e % RT;
        const int a = ta * RT + r, b = tb * RT + c;     // Mab tile: rows a, cols b
        Mab[r][c] = (a < g.v && b < g.v) ? g.G[((int64_t)ii * g.v + a) * ld + (int64_t)j * g.v + b] : 0.0;
        const int b2 = tb * RT + r, a2 = ta * RT + c;   // transposed tile: rows b, cols a
        Mba[r][c] = (a2 < g.v && b2 < g.v) ? g.G[((int64_t)ii * g.v + b2) * ld + (int64_t)j * g.v + a2] : 0.0;
    }
    __syncthreads();
    // thread -> (row r = tid / 8, 4 columns c = (tid % 8) + 8 q)
    const int r = tid / 8;
    const int a = ta * RT + r;
    double sa = 0.0, si = 0.0, sos = 0.0, sss = 0.0;
    if (a < g.v) {
        const double ea = g.ev[a] - g.eo[i] - g.eo[j];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = (tid % 8) + 8 * q;
            const int b = tb * RT + c;
            if (b < g.v) {
                const double mab = Mab[r][c], mba = Mba[c][r];
                const double D = ea + g.ev[b];
                const double dinv = 1.0 / D;
                g.R[((int64_t)ii * g.v + a) * ld + (int64_t)j * g.v + b] = (g.cs * mab - g.css * mba) * dinv;
                const double nn = g.cs * mab * mab - g.css * mab * mba;
                const double w = nn * dinv * dinv;
                sa += 2.0 * w;
                si -= 2.0 * w;
                sos -= mab * mab * dinv;
                sss -= (mab * mab - mab * mba) * dinv;
            }
        }
    }
    // per-row sum over the 8 threads of row r (fixed butterfly order)
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) sa += __shfl_xor(sa, off, 8);
    if ((tid % 8) == 0 && a < g.v) g.pv[(pair * g.nT + tb) * (int64_t)g.v + a] = sa;
    red[0][tid] = si;
    red[1][tid] = sos;
    red[2][tid] = sss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < 3; ++k) red[k][tid] += red[k][tid + w];
        __syncthreads();
    }
    if (tid < 3) g.pij[(pair * nt2 + tile) * 3 + tid] = red[tid][0];
}

// Fixed-order accumulation of the block's partials:
// deps_vir[a] += sum_{pair, tb} pv ; deps_occ[i] += sum_{j, tiles} pij[.][0]
// energy[0..1] += sum pij[.][1..2]
// Workgroups 0..nvb-1 own 64 consecutive virtuals each (lanes = a, coalesced;
// the 4 waves take interleaved partial rows and are summed in a fixed order).
// The next bo workgroups own one occupied each, the last two one energy each
// (strided partial sums, fixed LDS tree).  Reruns are bit-identical.
__global__ __launch_bounds__(256) void rimp2_grad_accum_kernel(const double* __restrict__ pv,
                                                               const double* __restrict__ pij, int o, int v,
                                                               int bo, int i0, int nT, double* __restrict__ deo,
                                                               double* __restrict__ dev,
                                                               double* __restrict__ energy) {
    __shared__ double sh[256];
    const int nt2 = nT * nT;
    const int64_t npair = (int64_t)bo * o;
    const int nvb = (v + 63) / 64;
    const int tid = threadIdx.x;
    double s = 0.0;
    if ((int)blockIdx.x < nvb) {
        const int lane = tid & 63, w = tid >> 6;
        const int a = blockIdx.x * 64 + lane;
        const int64_t n = npair * nT;
        if (a < v)
            for (int64_t e = w; e < n; e += 4) s += pv[e * v + a];
        sh[tid] = s;
        __syncthreads();
        if (w == 0 && a < v) dev[a] += (sh[lane] + sh[64 + lane]) + (sh[128 + lane] + sh[192 + lane]);
        return;
    }
    const int r = (int)blockIdx.x - nvb;   // < bo: deps_occ of occupied i0 + r; bo, bo+1: E_OS, E_SS
    const double* src;
    int64_t n;
    if (r < bo) {
        src = pij + (int64_t)r * o * nt2 * 3;
        n = (int64_t)o * nt2;
    } else {
        src = pij + 1 + (r - bo);
        n = npair * nt2;
    }
    for (int64_t e = tid; e < n; e += 256) s += src[e * 3];
    sh[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) sh[tid] += sh[tid + w];
        __syncthreads();
    }
    if (tid == 0) {
        if (r < bo) deo[i0 + r] += sh[0];
        else energy[r - bo] += sh[0];
    }
}

__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

}  // namespace

int rimp2_gradient(corr_ctx* ctx, int64_t naux, int64_t o, int64_t v, const double* B, const double* eo,
                   const double* ev, double c_os, double c_ss, int prec, double* dB, double* deo, double* dev,
                   corr_mp2_energy* out) {
    CORR_REQUIRE(ctx, naux > 0 && o > 0 && v > 0 && o < (1 << 20) && v < (1 << 20), "rimp2_gradient: bad sizes");
    CORR_REQUIRE(ctx, B && eo && ev && dB && deo && dev, "rimp2_gradient: null pointer");
    CORR_REQUIRE(ctx, prec == CORR_PREC_FP64 || prec == CORR_PREC_FP32 || prec == CORR_PREC_BF16X3,
                 "rimp2_gradient: unknown precision");
    const int gp = prec == CORR_PREC_FP64 ? CORR_PREC_FP64 : CORR_PREC_FP32;
    const int64_t ov = o * v;
    // a context scratch limit also bounds R (at most half the limit) and the slabs
    const double lim = ctx->scratch_limit ? (double)ctx->scratch_limit : 0.0;
    const double rbytes = 8.0 * (double)ov * (double)ov;
    if (rbytes <= FUSED_R_BYTES && (lim == 0.0 || rbytes <= 0.5 * lim)) {
        // fused: the pair-tile kernel writes the whole symmetric R and the
        // dE/deps partials, one GEMM forms dE/dB = -4 B R
        double* R;
        CORR_TRY(scratch(ctx, "grad_R", sizeof(double) * ov * ov, (void**)&R));
        KernelTimer timer(ctx, "rimp2_gradient");
        corr_mp2_energy e{};
        CORR_TRY(rimp2_grad_tiles(ctx, naux, o, v, B, eo, ev, c_os, c_ss, gp, R, deo, dev, &e));
        CORR_TRY(gemm(ctx, 0, 0, naux, ov, ov, -4.0, B, ov, 0, R, ov, 0, 0.0, dB, ov, 0, 1, gp));
        timer.stop();
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (out) *out = e;
        return CORR_OK;
    }
    // occupied block size: G and R slabs of at most ~6 GB each
    const double slab = lim == 0.0 ? 6.0e9 : std::min(6.0e9, 0.25 * lim);
    int64_t bo = (int64_t)(slab / (8.0 * (double)v * (double)ov));
    if (bo < 1) bo = 1;
    if (bo > o) bo = o;
    const int nT = (int)cdiv(v, RT);
    double *G, *R, *pv, *pij, *acc;
    CORR_TRY(scratch(ctx, "grad_G", sizeof(double) * bo * v * ov, (void**)&G));
    CORR_TRY(scratch(ctx, "grad_R", sizeof(double) * bo * v * ov, (void**)&R));
    CORR_TRY(scratch(ctx, "grad_pv", sizeof(double) * bo * o * nT * v, (void**)&pv));
    CORR_TRY(scratch(ctx, "grad_pij", sizeof(double) * bo * o * nT * nT * 3, (void**)&pij));
    CORR_TRY(scratch(ctx, "grad_acc", sizeof(double) * 8, (void**)&acc));
    auto zero = [&](double* p, int64_t n) {
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<int64_t>(cdiv(n, 256), 65536)), dim3(256), 0,
                           ctx->stream, p, n);
    };
    zero(deo, o);
    zero(dev, v);
    zero(acc, 8);
    CORR_LAUNCHED(ctx, "zero_kernel");
    KernelTimer timer(ctx, "rimp2_gradient");
    for (int64_t i0 = 0; i0 < o; i0 += bo) {
        const int64_t nb = std::min<int64_t>(bo, o - i0);
        // G_I[(i a)][(j b)] = sum_Q B[Q][(i a)] B[Q][(j b)]
        CORR_TRY(gemm(ctx, 1, 0, nb * v, ov, naux, 1.0, B + i0 * v, ov, 0, B, ov, 0, 0.0, G, ov, 0, 1, gp));
        GradArgs a{G, R, eo, ev, (int)o, (int)v, (int)nb, (int)i0, nT, c_os + c_ss, c_ss, pv, pij};
        const int64_t blocks = nb * o * nT * nT;
        CORR_REQUIRE(ctx, blocks < ((int64_t)1 << 31), "rimp2_gradient: too many tiles");
        hipLaunchKernelGGL(rimp2_grad_pairs_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, a);
        CORR_LAUNCHED(ctx, "rimp2_grad_pairs_kernel");
        const int nacc = (int)(cdiv(v, 64) + nb + 2);
        hipLaunchKernelGGL(rimp2_grad_accum_kernel, dim3((unsigned)nacc), dim3(256), 0, ctx->stream, pv, pij,
                           (int)o, (int)v, (int)nb, (int)i0, nT, deo, dev, acc);
        CORR_LAUNCHED(ctx, "rimp2_grad_accum_kernel");
        // dB[:, I, :] = -4 B R_I^T
        CORR_TRY(gemm(ctx, 0, 1, naux, nb * v, ov, -4.0, B, ov, 0, R, ov, 0, 0.0, dB + i0 * v, ov, 0, 1, gp));
    }
    timer.stop();
    double e[2];
    CORR_HIP(ctx, hipMemcpyAsync(e, acc, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (out) {
        out->e_os = e[0];
        out->e_ss = e[1];
        out->e_mp2 = e[0] + e[1];
        out->e_scs = e[1] / 3.0 + e[0] * 6.0 / 5.0;
    }
    return CORR_OK;
}

}  // namespace corr

extern "C" int corr_rimp2_gradient(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* d_Bia,
                                   const double* d_eps_occ, const double* d_eps_vir, double c_os, double c_ss,
                                   int prec, double* d_dBia, double* d_deps_occ, double* d_deps_vir,
                                   corr_mp2_energy* h_out) {
    CORR_CHECK_CTX(ctx);
    return corr::rimp2_gradient(ctx, naux, nocc, nvir, d_Bia, d_eps_occ, d_eps_vir, c_os, c_ss, prec, d_dBia,
                                d_deps_occ, d_deps_vir, h_out);
}
