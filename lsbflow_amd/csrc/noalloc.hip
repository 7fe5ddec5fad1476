// Allocation-free variant of lsb_horfft_dev for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The only device memory it needs beyond the caller's arrays, the
// staging buffer of the inverse transform's y pass, is the file-scope array below:
// 256 MiB (2^25 doubles, one 512 x 256 x 193 spectral field) per GPU, held until the
// process ends.  Host code never uses the array's name as an address: lsb_create
// resolves its device address once per device (hipGetSymbolAddress with that device
// current) and the entry point passes that address to the kernels.  Every call on a
// GPU shares that GPU's array, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "lsb_internal.h"


static constexpr size_t kScratchDoubles = (size_t)1 << 25;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

double* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

}  // namespace

namespace lsb {

int noalloc_resolve(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("lsb_create: device ordinal out of range for the scratch table");
        return LSB_E_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return LSB_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("lsb_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return LSB_E_HIP;
    }
    g_scratch[device] = (double*)p;
    return LSB_OK;
}

}  // namespace lsb

using namespace lsb;

// lsb_horfft_dev without device allocation (see the header for the contract).
extern "C" int lsb_horfft_dev_noalloc(lsb_ctx* c, double* ain, double* ainf, int is, int nfields, void* stream) {
    if (!c || !ain || !ainf || nfields < 1 || (is != -1 && is != 1)) {
        set_error("horfft_dev_noalloc: null context or array, nfields < 1, or is not -1/+1");
        return LSB_E_ARG;
    }
    double* scr = scratch_of(c->device);
    if (!scr) {
        set_error("horfft_dev_noalloc: scratch address not resolved for this device (lsb_create did not run)");
        return LSB_E_STATE;
    }
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    // fields per launch: the inverse y pass stages nzp * ny * nx doubles per field
    int batch = kMaxFields;
    if (is == 1 && horfft_has_ops(c->hp)) {
        const size_t per = (size_t)c->nzp * c->ny * c->nx;
        if (per > kScratchDoubles) {
            set_error("horfft_dev_noalloc: one field (" + std::to_string(per * 8) +
                      " bytes of staging) exceeds the scratch maximum of " + std::to_string(kScratchDoubles * 8) +
                      " bytes");
            return LSB_E_ARG;
        }
        batch = (int)std::min<size_t>((size_t)kMaxFields, kScratchDoubles / per);
    }
    for (int f0 = 0; f0 < nfields; f0 += batch) {
        const int nb = std::min(batch, nfields - f0);
        const double* pin[kMaxFields];
        double* pout[kMaxFields];
        for (int q = 0; q < nb; ++q) {
            if (is == -1) {
                pin[q] = ain + (size_t)(f0 + q) * c->nphys;
                pout[q] = ainf + (size_t)(f0 + q) * c->nspec;
            } else {
                pin[q] = ainf + (size_t)(f0 + q) * c->nspec;
                pout[q] = ain + (size_t)(f0 + q) * c->nphys;
            }
        }
        if (is == -1)
            LSB_TRY(horfft_fwd(c, c->hp, pin, pout, nb, c->nzp, s));
        else
            LSB_TRY(horfft_inv_ops(c, c->hp, pin, pout, nb, c->nzp, nullptr, s, scr));
    }
    return LSB_OK;
}
