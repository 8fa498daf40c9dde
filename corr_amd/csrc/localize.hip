// Orbital localisation: Pipek-Mezey and Boys (syshfwplug/pymodule.py:75-167).
//
// Both maximise f = sum_K sum_i (Q_K)_ii^2 over rotations of the occupied
// orbitals, with Q_K = C M_K C^T (rows of C are orbitals):
//   Pipek-Mezey (calculate_PM, pymodule.py:143-153): one Q_A per atom,
//     (Q_A)_ij = (sum_{k on A} C_ik (C S)_jk + (C S)_ik C_jk) / 2, whose
//     diagonal is the Mulliken population of orbital i on A;
//   Boys (calculate_boys, pymodule.py:156-167): Q_d = C D_d C^T for the AO
//     dipole matrices D_x, D_y, D_z.
// The reference maximises with scipy BFGS over U = expm(X); here parallel
// Jacobi sweeps do it on the GPU: each sweep is n_even - 1 rounds of n_even/2
// disjoint orbital pairs (the round-robin schedule), one kernel finds every
// pair's angle (4 g = atan2(B, -A), A = sum_K [Q_st^2 - (Q_ss - Q_tt)^2/4],
// B = sum_K Q_st (Q_ss - Q_tt); Pipek & Mezey, J. Chem. Phys. 90, 4916
// (1989)), a second applies all the rotations at once, out of place:
// Q_K <- J^T Q_K J and C <- J^T C.  Sums are fixed-order.
#include "common.hpp"
#include <cmath>
#include <vector>

namespace corr {

namespace {

unsigned grid_for(int64_t n) {
    int64_t g = cdiv(n, 256);
    return (unsigned)(g > 65536 ? 65536 : (g < 1 ? 1 : g));
}

// Q_A[i][j] for every atom A: functions of A are idx[ptr[A] .. ptr[A+1]).
__global__ void pm_q_kernel(const double* __restrict__ C, const double* __restrict__ CS, int64_t n, int64_t nbf,
                            const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx, int64_t natom,
                            double* __restrict__ Q) {
    const int64_t total = natom * n * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e % n, i = (e / n) % n, A = e / (n * n);
        double s = 0.0;
        for (int32_t q = ptr[A]; q < ptr[A + 1]; ++q) {
            const int64_t k = idx[q];
            s += C[i * nbf + k] * CS[j * nbf + k] + CS[i * nbf + k] * C[j * nbf + k];
        }
        Q[e] = 0.5 * s;
    }
}

// One wave per pair of this round: the angle, the gain A + sqrt(A^2 + B^2),
// and each orbital's (partner, cos, signed sin).  Orbitals without a partner
// keep (self, 1, 0), set by the caller's initialisation of coef.
__global__ __launch_bounds__(256) void loc_angle_kernel(const double* __restrict__ Q, int64_t n, int64_t nK,
                                                        const int32_t* __restrict__ pairs, int npair,
                                                        double* __restrict__ coef, double* __restrict__ gain) {
    const int w = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (w >= npair) return;
    const int64_t s = pairs[2 * w], t = pairs[2 * w + 1];
    double A = 0.0, B = 0.0;
    if (s >= 0) {
        for (int64_t K = lane; K < nK; K += 64) {
            const double* q = Q + K * n * n;
            const double qst = q[s * n + t], d = q[s * n + s] - q[t * n + t];
            A += qst * qst - 0.25 * d * d;
            B += qst * d;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        A += __shfl_xor(A, off, 64);
        B += __shfl_xor(B, off, 64);
    }
    if (lane == 0) {
        if (s < 0) {
            gain[w] = 0.0;
            return;
        }
        const double g = 0.25 * atan2(B, -A);
        const double c = cos(g), sn = sin(g);
        gain[w] = A + hypot(A, B);
        // orbital s' = c s + sn t, t' = -sn s + c t
        coef[3 * s + 0] = (double)t;
        coef[3 * s + 1] = c;
        coef[3 * s + 2] = sn;
        coef[3 * t + 0] = (double)s;
        coef[3 * t + 1] = c;
        coef[3 * t + 2] = -sn;
    }
}

// Q2_K = J^T Q_K J and L2 = J^T L for the rotations of one round:
// Q'_ij = c_i c_j Q_ij + c_i s_j Q_i,p(j) + s_i c_j Q_p(i),j + s_i s_j Q_p(i),p(j),
// L'_i = c_i L_i + s_i L_p(i).
__global__ void loc_apply_kernel(const double* __restrict__ Q, const double* __restrict__ L,
                                 const double* __restrict__ coef, int64_t n, int64_t nK, int64_t nbf,
                                 double* __restrict__ Q2, double* __restrict__ L2) {
    const int64_t nq = nK * n * n, total = nq + n * nbf;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        if (e < nq) {
            const int64_t j = e % n, i = (e / n) % n, K = e / (n * n);
            const double* q = Q + K * n * n;
            const int64_t pi = (int64_t)coef[3 * i], pj = (int64_t)coef[3 * j];
            const double ci = coef[3 * i + 1], si = coef[3 * i + 2], cj = coef[3 * j + 1], sj = coef[3 * j + 2];
            Q2[e] = ci * cj * q[i * n + j] + ci * sj * q[i * n + pj] + si * cj * q[pi * n + j] +
                    si * sj * q[pi * n + pj];
        } else {
            const int64_t r = e - nq, mu = r % nbf, i = r / nbf;
            const int64_t pi = (int64_t)coef[3 * i];
            L2[r] = coef[3 * i + 1] * L[i * nbf + mu] + coef[3 * i + 2] * L[pi * nbf + mu];
        }
    }
}

__global__ void loc_coef_init_kernel(double* __restrict__ coef, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    coef[3 * i] = (double)i;
    coef[3 * i + 1] = 1.0;
    coef[3 * i + 2] = 0.0;
}

// diag[K*n + i] = Q_K[i][i]^2
__global__ void loc_diag2_kernel(const double* __restrict__ Q, int64_t n, int64_t nK, double* __restrict__ out) {
    const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (e >= nK * n) return;
    const int64_t i = e % n, K = e / n;
    const double d = Q[(K * n + i) * n + i];
    out[e] = d * d;
}

}  // namespace

int localize(corr_ctx* ctx, int64_t nbf, int64_t n, const double* C, int kind, const double* M,
             const int32_t* bf_atom, int64_t natom, const corr_loc_opts* opts, double* Lout, corr_loc_result* res) {
    CORR_REQUIRE(ctx, nbf > 0 && n > 0 && n <= nbf && C && M && Lout && opts && res, "localize: bad arguments");
    CORR_REQUIRE(ctx, kind == CORR_LOC_PM || kind == CORR_LOC_BOYS, "localize: kind must be PM or BOYS");
    CORR_REQUIRE(ctx, kind != CORR_LOC_PM || (bf_atom && natom > 0), "localize: PM needs bf_atom and natom");
    CORR_REQUIRE(ctx, opts->max_sweeps >= 0 && opts->tol >= 0.0, "localize: bad options");
    const int64_t nK = kind == CORR_LOC_PM ? natom : 3;
    double *Q, *Q2, *L, *L2, *T, *coef, *gain, *diag;
    CORR_TRY(scratch(ctx, "loc_Q", sizeof(double) * nK * n * n, (void**)&Q));
    CORR_TRY(scratch(ctx, "loc_Q2", sizeof(double) * nK * n * n, (void**)&Q2));
    CORR_TRY(scratch(ctx, "loc_L", sizeof(double) * n * nbf, (void**)&L));
    CORR_TRY(scratch(ctx, "loc_L2", sizeof(double) * n * nbf, (void**)&L2));
    CORR_TRY(scratch(ctx, "loc_T", sizeof(double) * n * nbf, (void**)&T));
    CORR_TRY(scratch(ctx, "loc_coef", sizeof(double) * 3 * n, (void**)&coef));
    CORR_TRY(scratch(ctx, "loc_diag", sizeof(double) * nK * n, (void**)&diag));
    CORR_HIP(ctx, hipMemcpyAsync(L, C, sizeof(double) * n * nbf, hipMemcpyDeviceToDevice, ctx->stream));

    if (kind == CORR_LOC_PM) {
        // functions of each atom (CSR), from the per-function atom index
        std::vector<int32_t> ptr(natom + 1, 0), idx(nbf);
        for (int64_t k = 0; k < nbf; ++k) {
            CORR_REQUIRE(ctx, bf_atom[k] >= 0 && bf_atom[k] < natom, "localize: bf_atom out of range");
            ++ptr[bf_atom[k] + 1];
        }
        for (int64_t A = 0; A < natom; ++A) ptr[A + 1] += ptr[A];
        std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
        for (int64_t k = 0; k < nbf; ++k) idx[fill[bf_atom[k]]++] = (int32_t)k;
        int32_t* dptr = nullptr;
        CORR_TRY(scratch(ctx, "loc_atoms", sizeof(int32_t) * (natom + 1 + nbf), (void**)&dptr));
        CORR_HIP(ctx, hipMemcpyAsync(dptr, ptr.data(), sizeof(int32_t) * (natom + 1), hipMemcpyHostToDevice,
                                     ctx->stream));
        CORR_HIP(ctx, hipMemcpyAsync(dptr + natom + 1, idx.data(), sizeof(int32_t) * nbf, hipMemcpyHostToDevice,
                                     ctx->stream));
        // T = C S, then Q_A
        CORR_TRY(gemm(ctx, 0, 0, n, nbf, nbf, 1.0, C, nbf, 0, M, nbf, 0, 0.0, T, nbf, 0, 1, CORR_PREC_FP64));
        hipLaunchKernelGGL(pm_q_kernel, dim3(grid_for(nK * n * n)), dim3(256), 0, ctx->stream, C, T, n, nbf, dptr,
                           dptr + natom + 1, natom, Q);
        CORR_LAUNCHED(ctx, "pm_q_kernel");
        // the host vectors are read by the copies above: wait before they go
        CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    } else {
        for (int d = 0; d < 3; ++d) {
            CORR_TRY(gemm(ctx, 0, 0, n, nbf, nbf, 1.0, C, nbf, 0, M + d * nbf * nbf, nbf, 0, 0.0, T, nbf, 0, 1,
                          CORR_PREC_FP64));
            CORR_TRY(gemm(ctx, 0, 1, n, n, nbf, 1.0, T, nbf, 0, C, nbf, 0, 0.0, Q + d * n * n, n, 0, 1,
                          CORR_PREC_FP64));
        }
    }

    auto spread = [&](const double* q, double* out) -> int {
        hipLaunchKernelGGL(loc_diag2_kernel, dim3((unsigned)cdiv(nK * n, 256)), dim3(256), 0, ctx->stream, q, n, nK,
                           diag);
        CORR_LAUNCHED(ctx, "loc_diag2_kernel");
        return reduce_sum(ctx, diag, nK * n, 1, out);
    };
    CORR_TRY(spread(Q, &res->spread_initial));
    res->spread = res->spread_initial;
    res->sweeps = 0;
    res->converged = 0;

    // round-robin schedule over m = n rounded up to even: position 0 fixed,
    // positions 1..m-1 rotate; index n (odd n) is a dummy (pair skipped)
    const int64_t m = n + (n & 1);
    const int npair = (int)(m / 2);
    const int64_t rounds = m - 1;
    std::vector<int32_t> sched(2 * npair * std::max<int64_t>(rounds, 1), -1);
    for (int64_t r = 0; r < rounds; ++r) {
        std::vector<int64_t> pos(m);
        pos[0] = 0;
        for (int64_t k = 1; k < m; ++k) pos[k] = 1 + (k - 1 + r) % (m - 1);
        for (int k = 0; k < npair; ++k) {
            int64_t s = pos[k], t = pos[m - 1 - k];
            if (s > t) std::swap(s, t);
            int32_t* pr = &sched[2 * (r * npair + k)];
            if (t >= n) {   // dummy partner
                pr[0] = -1;
                pr[1] = -1;
            } else {
                pr[0] = (int32_t)s;
                pr[1] = (int32_t)t;
            }
        }
    }
    int32_t* dsched = nullptr;
    CORR_TRY(scratch(ctx, "loc_sched", sizeof(int32_t) * sched.size(), (void**)&dsched));
    CORR_TRY(scratch(ctx, "loc_gain", sizeof(double) * npair * std::max<int64_t>(rounds, 1), (void**)&gain));
    CORR_HIP(ctx, hipMemcpyAsync(dsched, sched.data(), sizeof(int32_t) * sched.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));

    if (n > 1) {
        for (int sw = 0; sw < opts->max_sweeps; ++sw) {
            for (int64_t r = 0; r < rounds; ++r) {
                hipLaunchKernelGGL(loc_coef_init_kernel, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream,
                                   coef, n);
                CORR_LAUNCHED(ctx, "loc_coef_init_kernel");
                hipLaunchKernelGGL(loc_angle_kernel, dim3((unsigned)cdiv((int64_t)npair * 64, 256)), dim3(256), 0,
                                   ctx->stream, Q, n, nK, dsched + 2 * r * npair, npair, coef, gain + r * npair);
                CORR_LAUNCHED(ctx, "loc_angle_kernel");
                hipLaunchKernelGGL(loc_apply_kernel, dim3(grid_for(nK * n * n + n * nbf)), dim3(256), 0, ctx->stream,
                                   Q, L, coef, n, nK, nbf, Q2, L2);
                CORR_LAUNCHED(ctx, "loc_apply_kernel");
                std::swap(Q, Q2);
                std::swap(L, L2);
            }
            double g = 0.0;
            CORR_TRY(reduce_sum(ctx, gain, (int64_t)npair * rounds, 1, &g));
            CORR_TRY(spread(Q, &res->spread));
            res->sweeps = sw + 1;
            if (g <= opts->tol * std::max(res->spread, 1.0)) {
                res->converged = 1;
                break;
            }
        }
    } else {
        res->converged = 1;
    }
    CORR_HIP(ctx, hipMemcpyAsync(Lout, L, sizeof(double) * n * nbf, hipMemcpyDeviceToDevice, ctx->stream));
    CORR_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CORR_OK;
}

}  // namespace corr

extern "C" int corr_localize(corr_ctx* ctx, int64_t nbf, int64_t nocc, const double* d_C, int kind,
                             const double* d_M, const int32_t* h_bf_atom, int64_t natom, const corr_loc_opts* opts,
                             double* d_L, corr_loc_result* h_res) {
    CORR_CHECK_CTX(ctx);
    return corr::localize(ctx, nbf, nocc, d_C, kind, d_M, h_bf_atom, natom, opts, d_L, h_res);
}
