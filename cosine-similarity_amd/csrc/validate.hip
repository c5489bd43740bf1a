// validate.hip -- device work on a spectrum set before it is compared:
// cs_validate_spectra (the CSR checks of cg_check.h over a set the caller did
// not pack itself) and cs_spectra_norms (the per-spectrum norms).
#include "cg_check.h"
#include "cs_internal.h"

namespace cs {
namespace {

constexpr int kValidateThreads = 256;
constexpr int64_t kValidateGrid = 4096;

__global__ void k_clear_fault(uint32_t *fault) {
    if (threadIdx.x == 0) *fault = 0u;
}

// One thread per spectrum (grid-stride); a spectrum's faults are ORed into
// the flag, one atomic per faulty spectrum.
__global__ void __launch_bounds__(kValidateThreads) k_validate(cs_spectra s, uint32_t *fault) {
    const int64_t stride = (int64_t)gridDim.x * kValidateThreads;
    for (int64_t x = blockIdx.x * (int64_t)kValidateThreads + threadIdx.x; x < s.n_spectra;
         x += stride) {
        const uint32_t f =
            spectrum_faults(s.offsets, x, s.n_peaks, s.max_peaks, s.max_abs_mz, s.mz, s.mz_exact);
        if (f) atomicOr(fault, f);
    }
}

// norm[x] = sqrt(sum over x's peaks, in peak order, of (mz_pow * int_pow)^2)
// (int_pow alone when mz_pow is NULL): one thread per spectrum, a sequential
// float64 sum, so the value is the one a left-to-right host sum gives.
__global__ void __launch_bounds__(kValidateThreads) k_norms(cs_spectra s, double *norm) {
    const int64_t stride = (int64_t)gridDim.x * kValidateThreads;
    for (int64_t x = blockIdx.x * (int64_t)kValidateThreads + threadIdx.x; x < s.n_spectra;
         x += stride) {
        double acc = 0.0;
        for (int64_t p = s.offsets[x]; p < s.offsets[x + 1]; ++p) {
            const double w = s.mz_pow ? s.mz_pow[p] * s.int_pow[p] : s.int_pow[p];
            acc = acc + w * w;
        }
        norm[x] = sqrt(acc);
    }
}

}  // namespace
}  // namespace cs

using namespace cs;

extern "C" cs_status cs_validate_spectra(const cs_spectra *spectra, uint32_t *fault,
                                         void *stream) {
    if (!spectra || !fault) return fail(CS_ERR_INVALID, "spectra/fault is NULL");
    const cs_spectra &s = *spectra;
    if (s.n_spectra < 0 || s.n_peaks < 0) return fail(CS_ERR_INVALID, "negative size");
    if (s.n_spectra > 0 && !s.offsets) return fail(CS_ERR_INVALID, "offsets must not be NULL");
    if (s.n_peaks > 0 && !s.mz) return fail(CS_ERR_INVALID, "mz must not be NULL");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_clear_fault, dim3(1), dim3(64), 0, hs, fault);
    CS_LAUNCH_CHECK("k_clear_fault");
    if (s.n_spectra == 0) return CS_OK;
    int64_t blocks = (s.n_spectra + kValidateThreads - 1) / kValidateThreads;
    if (blocks > kValidateGrid) blocks = kValidateGrid;
    hipLaunchKernelGGL(k_validate, dim3((unsigned)blocks), dim3(kValidateThreads), 0, hs, s, fault);
    CS_LAUNCH_CHECK("k_validate");
    return CS_OK;
}

extern "C" cs_status cs_spectra_norms(const cs_spectra *spectra, double *norm, void *stream) {
    if (!spectra || !norm) return fail(CS_ERR_INVALID, "spectra/norm is NULL");
    const cs_spectra &s = *spectra;
    if (s.n_spectra < 0 || s.n_peaks < 0) return fail(CS_ERR_INVALID, "negative size");
    if (s.n_spectra == 0) return CS_OK;
    if (!s.offsets) return fail(CS_ERR_INVALID, "offsets must not be NULL");
    if (s.n_peaks > 0 && !s.int_pow) return fail(CS_ERR_INVALID, "int_pow must not be NULL");
    hipStream_t hs = static_cast<hipStream_t>(stream);
    int64_t blocks = (s.n_spectra + kValidateThreads - 1) / kValidateThreads;
    if (blocks > kValidateGrid) blocks = kValidateGrid;
    hipLaunchKernelGGL(k_norms, dim3((unsigned)blocks), dim3(kValidateThreads), 0, hs, s, norm);
    CS_LAUNCH_CHECK("k_norms");
    return CS_OK;
}
