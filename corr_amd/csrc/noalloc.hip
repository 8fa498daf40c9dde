// Allocation-free entry point of the RI-MP2 energy (corr_rimp2_energy_noalloc).
//
// Its workspace (repacked panels, per-tile partial sums, tile list, reduction
// buffers) lives in one file-scope device array of CORR_NOALLOC_SCRATCH_BYTES,
// which every GPU the library runs on holds until the process ends.  The
// array's name is a host-side stand-in: its device address on each device is
// looked up once, with that device current, when a context is created for it
// (hipGetSymbolAddress), and only that address is handed to kernels.  The
// entry point then calls no hipMalloc / hipFree and records no kernel timing
// events; every workspace byte it reads was written earlier in the same call,
// so nothing needs clearing between calls.  Calls that overlap on one device
// share the array and race.
#include "common.hpp"

#include <mutex>
#include <string>

static __device__ __align__(256) unsigned char scratch[CORR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
bool g_looked_up[MAX_DEVICES] = {};

}  // namespace

namespace corr {

int noalloc_register_device(int device) {
    if (device < 0 || device >= MAX_DEVICES) return CORR_EINVAL;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_looked_up[device]) return g_ptr[device] ? CORR_OK : CORR_EHIP;
    g_looked_up[device] = true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_ptr[device] = nullptr;
        return CORR_EHIP;
    }
    g_ptr[device] = p;
    return CORR_OK;
}

void* noalloc_scratch(int device, size_t* bytes) {
    if (bytes) *bytes = CORR_NOALLOC_SCRATCH_BYTES;
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}

}  // namespace corr

extern "C" int corr_rimp2_energy_noalloc(corr_ctx* ctx, int64_t naux, int64_t nocc, int64_t nvir,
                                         const double* d_Bia, const double* d_eps_occ,
                                         const double* d_eps_vir, int prec, corr_mp2_energy* h_out) {
    CORR_CHECK_CTX(ctx);
    size_t cap = 0;
    void* base = corr::noalloc_scratch(ctx->device, &cap);
    if (!base)
        return corr::set_error(ctx, CORR_EHIP, "rimp2_energy_noalloc: the static scratch of device " +
                                                   std::to_string(ctx->device) + " is not available");
    const int64_t need = corr::rimp2_workspace(naux, nocc, nvir, prec, nullptr, nullptr);
    CORR_REQUIRE(ctx, need >= 0, "rimp2_energy_noalloc: bad sizes or precision");
    if ((size_t)need > cap)
        return corr::set_error(ctx, CORR_EINVAL,
                               "rimp2_energy_noalloc: this problem needs " + std::to_string(need) +
                                   " bytes of workspace; the static scratch holds " + std::to_string(cap) +
                                   " bytes (use corr_rimp2_energy)");
    corr::RimpWs ws;
    corr::rimp2_workspace(naux, nocc, nvir, prec, base, &ws);
    return corr::rimp2_energy_with(ctx, naux, nocc, nvir, d_Bia, d_eps_occ, d_eps_vir, prec, h_out, &ws);
}
