// Shared host/device infrastructure of libcorr_gpu.so: the context, status
// handling, cached device scratch and small device helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <map>
#include <vector>
#include "../../include/corr_abi.h"

struct corr_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string last_error;
    struct Buf { void* p = nullptr; size_t bytes = 0; };
    std::map<std::string, Buf> scratch;
    size_t scratch_limit = 0;          // bytes of cached scratch allowed (0: no limit)
    void* blas_handle = nullptr;   // rocBLAS handle used by the eigen metric
    hipStream_t aux_stream = nullptr;   // second stream (default priority) for overlapped work
    hipStream_t part_stream = nullptr;  // stream on the CUs that aux_stream does not use
    bool aux_masked = false;            // aux_stream and part_stream split the CUs
    hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_part = nullptr;
    hipEvent_t ev_switch = nullptr;     // orders a new stream after the previous one
    bool timing = false;
    struct Timed { std::string name; hipEvent_t start, stop; };
    std::vector<Timed> pending;                  // recorded, not yet read
    std::map<std::string, std::pair<double, int64_t>> kernel_ms;
};

namespace corr {

int set_error(corr_ctx* ctx, int code, const std::string& msg);

// Device scratch owned by the context, reused across calls (grown on demand).
int scratch(corr_ctx* ctx, const char* name, size_t bytes, void** out);

// Brackets one kernel launch with events when the context's timing is on.
struct KernelTimer {
    corr_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    KernelTimer(corr_ctx* c, const char* n, bool on = true) : ctx(c), name(n) {
        if (on && ctx->timing && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
            (void)hipEventRecord(a, ctx->stream);
        else
            a = b = nullptr;
    }
    void stop() {
        if (a && b) {
            (void)hipEventRecord(b, ctx->stream);
            ctx->pending.push_back({name, a, b});
            a = b = nullptr;
        }
    }
    ~KernelTimer() { stop(); }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t rup(int64_t a, int64_t b) { return cdiv(a, b) * b; }

// Triangular structure a GEMM may skip (flags of gemm's tri argument):
//   A_LOWER: op(A) is lower triangular (entries with k > m are zero): the K
//            range of a row tile ends after its last row;
//   B_LOWER: op(B) is lower triangular (entries with k < n are zero): the K
//            range of a column tile starts at its first column;
//   C_LOWER: only C on and below the diagonal is wanted: tiles wholly above
//            it are skipped (and left untouched; tiles across it are written
//            whole).
enum { GEMM_TRI_A_LOWER = 1, GEMM_TRI_B_LOWER = 2, GEMM_TRI_C_LOWER = 4 };

// GEMM launcher shared by all modules (row-major, see corr_dgemm); tri takes
// GEMM_TRI_* flags (0 or 1 = A_LOWER for the Cholesky-metric application).
int gemm(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K,
         double alpha, const double* A, int64_t lda, int64_t sA,
         const double* B, int64_t ldb, int64_t sB, double beta, double* C,
         int64_t ldc, int64_t sC, int64_t batch, int prec, int tri = 0);

// C = alpha op(A) op(B) + beta C (row-major, K <= 75000) as an FP64 GEMM
// emulated on the int8 matrix cores with 5 slices per operand (ozaki.hip).
int gemm_ozaki(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K, double alpha,
               const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
               bool a_lower);

// The same in two steps, so that each operand can be sliced on its own
// stream as soon as it exists: ozaki_prepare cuts the R x K operand X (rows
// of op(A), or columns of op(B); kmajor: element (r, k) at X[r ld + k], else
// at X[k ld + r]) into the context scratch buffers named after `name`;
// ozaki_gemm_prepared multiplies a prepared A (M = A.R rows) and B (N = B.R
// columns) on ctx->stream, after both preparations.
struct OzOperand {
    int8_t* slices = nullptr;
    double* scales = nullptr;
    int64_t R = 0, K = 0, Ks = 0;
};
int ozaki_prepare(corr_ctx* ctx, const char* name, const double* X, int64_t ld, int64_t R, int64_t K, bool kmajor,
                  OzOperand* out);

// The BF16x3 RI-MP2 energy kernel's panels (rimp2.hip) as a GEMM output:
// element (Q, n = i v + a) goes, as hi = bf16(x) and lo = bf16(x - hi), to
// panel (i / 4, a / 32), 16-deep stage Q / 16, row (i % 4) 32 + a % 32, in the
// energy kernels' LDS order; Kp rows per panel (Ks = Kp / 16 stages).
struct Bf16Panels {
    uint16_t* p = nullptr;
    int o = 0, v = 0, nAC = 0;
    int64_t Kp = 0, Ks = 0;
};
// pan != nullptr: C is not written; the product goes to the panels (alpha 1,
// beta 0, rows up to Kp, the pad rows Q >= M as zeros).
int ozaki_gemm_prepared(corr_ctx* ctx, const OzOperand& A, const OzOperand& B, double alpha, double beta, double* C,
                        int64_t ldc, bool a_lower, const Bf16Panels* pan = nullptr);

// op(A) op(B) (M x N, alpha 1) written through an output mapping of the
// 4-index transform (gemm.hip): mapping 1 (PAIRS): row X n + i, column j ->
// C[X P + tri(i, j)] for j >= i, P = n (n + 1) / 2; mapping 2 (MIRROR): row
// tri(i, j) n + k, column l -> C[((i n + j) n + k) n + l] and (j i k l).
int gemm_mapped(corr_ctx* ctx, int transa, int transb, int64_t M, int64_t N, int64_t K, const double* A,
                int64_t lda, const double* B, int64_t ldb, double* C, int mapping, int64_t n, int prec);

// J^-1/2 (EIGEN) or L^-1 of J = L L^T (CHOLESKY) into M (metric.hip).
int ri_metric(corr_ctx* ctx, int64_t naux, const double* J, int mode, double* M, int64_t* nkept);

// Eigen-decomposition of the symmetric n x n matrix A (device, in place):
// row c of A becomes eigenvector c, lam (device) the eigenvalues ascending
// (metric.hip).  Synchronises the stream.
int sym_eig(corr_ctx* ctx, int64_t n, double* A, double* lam);

// Solves the (m+1) x (m+1) DIIS system [B -1; -1 0] c = [0; -1] (row-major B
// of order m+1 with the border already set) by Gaussian elimination with
// partial pivoting; false when it is singular (ccsd.hip).
bool diis_solve(std::vector<double> B, int m, std::vector<double>& c);

// Deterministic sum of n doubles (device) into *h_out (host), nser (<= 16)
// interleaved series: data[i*nser + s].
int reduce_sum(corr_ctx* ctx, const double* d_data, int64_t n, int nser, double* h_out);
// The same with caller-provided device workspace: d_out holds 16 doubles,
// d_mid 16 * REDUCE_NCHUNK (used when n > 64 * REDUCE_NCHUNK).
constexpr int64_t REDUCE_NCHUNK = 256;
int reduce_sum_ws(corr_ctx* ctx, const double* d_data, int64_t n, int nser, double* h_out, double* d_mid,
                  double* d_out);

// Device workspace of one RI-MP2 energy call (rimp2.hip): the repacked
// panels, the per-tile partial sums, the BF16x3 tile list and the reduction's
// two buffers.
struct RimpWs {
    void* panels = nullptr;
    double* partial = nullptr;
    ushort4* tiles = nullptr;
    double* red_mid = nullptr;
    double* red_out = nullptr;
};
// Carves the workspace of one energy call out of base (nullptr: sizes only)
// in 256-byte aligned pieces; returns the bytes needed, or -1 for arguments
// the energy rejects.
int64_t rimp2_workspace(int64_t naux, int64_t nocc, int64_t nvir, int prec, void* base, RimpWs* ws);
// The energy on a caller-provided workspace (no allocation, no kernel timing).
int rimp2_energy_with(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia,
                      const double* eo, const double* ev, int prec, corr_mp2_energy* out, const RimpWs* ws);

// The BF16x3 energy's panels in the context scratch as a producer's target
// (pad rows past naux and, for ragged o or v, the empty panel rows zeroed),
// and the energy on them as they stand (rimp2.hip).
int rimp2_bf16_panels(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, Bf16Panels* out);
int rimp2_energy_prepacked(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* eo,
                           const double* ev, corr_mp2_energy* out);
int rimp2_energy(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir, const double* Bia, const double* eo,
                 const double* ev, int prec, corr_mp2_energy* out);

// Pair-tile part of the RI-MP2 gradient (rimp2.hip): the symmetric matrix R
// (ov x ov doubles at R, every element written), dE/deps and the energies.
int rimp2_grad_tiles(corr_ctx* ctx, int64_t naux, int64_t o, int64_t v, const double* Bia, const double* eo,
                     const double* ev, double c_os, double c_ss, int prec, double* R, double* deo, double* dev,
                     corr_mp2_energy* out);

// Device address of the allocation-free entry points' static scratch on this
// device (noalloc.hip), looked up once per device by corr_ctx_create with the
// device current; nullptr if the lookup failed.
int noalloc_register_device(int device);
void* noalloc_scratch(int device, size_t* bytes);

}  // namespace corr

#define CORR_CHECK_CTX(ctx)                                         \
    do {                                                            \
        if (!(ctx)) return CORR_EINVAL;                             \
        hipError_t e_ = hipSetDevice((ctx)->device);                \
        if (e_ != hipSuccess)                                       \
            return corr::set_error(ctx, CORR_EHIP,                  \
                std::string("hipSetDevice: ") + hipGetErrorString(e_)); \
    } while (0)

#define CORR_HIP(ctx, expr)                                         \
    do {                                                            \
        hipError_t e_ = (expr);                                     \
        if (e_ != hipSuccess)                                       \
            return corr::set_error(ctx, CORR_EHIP,                  \
                std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CORR_LAUNCHED(ctx, what)                                    \
    do {                                                            \
        hipError_t e_ = hipGetLastError();                          \
        if (e_ != hipSuccess)                                       \
            return corr::set_error(ctx, CORR_EHIP,                  \
                std::string("launch ") + (what) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define CORR_TRY(expr)                                              \
    do { int s_ = (expr); if (s_ != CORR_OK) return s_; } while (0)

#define CORR_REQUIRE(ctx, cond, msg)                                \
    do { if (!(cond)) return corr::set_error(ctx, CORR_EINVAL, msg); } while (0)

// MFMA operand/accumulator vector types for gfx950.
typedef double corr_d4 __attribute__((ext_vector_type(4)));
typedef float corr_f4 __attribute__((ext_vector_type(4)));
typedef short corr_s8 __attribute__((ext_vector_type(8)));
typedef __bf16 corr_bf8 __attribute__((ext_vector_type(8)));
typedef unsigned int corr_u4 __attribute__((ext_vector_type(4)));

// Row (within the 16x16 output tile) of accumulator element r held by a lane;
// the column is lane % 16.  FP32 accumulators of the 16x16 MFMAs hold four
// consecutive rows per lane group; v_mfma_f64_16x16x4_f64 interleaves them
// (lane group g, element r -> row g + 4 r), as measured on gfx950.
template <typename ACC> struct AccLayout {
    __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct AccLayout<corr_d4> {
    __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
