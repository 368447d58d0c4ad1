// dmpm_calculate_volumes_noalloc: calculate_point_cloud_volumes for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  The particle planes the kernels work on live in a
// file-scope device array instead of memory allocated per call.
#include <hip/hip_runtime.h>

#include <exception>
#include <mutex>
#include <string>

#include "../../include/dmpm.h"
#include "dmpm_kernels.h"

using namespace dmpm;

// Four planes (x0, x1, x2, m) of at most DMPM_NOALLOC_SCRATCH_BYTES / (4 *
// element size) particles.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` only names the symbol; its
// device address on the current GPU comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[DMPM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int fail(int code, const std::string& msg) {
    dmpm::set_last_error(msg);
    return code;
}

}  // namespace

extern "C" int dmpm_calculate_volumes_noalloc(const dmpm_config* cfg, int n, const void* x, const void* m,
                                              void* vol, void* grid_m, int dtype, void* stream) {
    if (!cfg || !x || !m || !vol || !grid_m) return fail(DMPM_ERR_ARG, "null argument");
    if (n < 0) return fail(DMPM_ERR_ARG, "n must be non-negative");
    if (dtype != DMPM_DTYPE_F32 && dtype != DMPM_DTYPE_F64) return fail(DMPM_ERR_ARG, "unknown dtype");
    for (int i = 0; i < 3; ++i)
        if (cfg->dims[i] <= 0) return fail(DMPM_ERR_ARG, "grid dims must be positive");
    if (!(cfg->dx > 0)) return fail(DMPM_ERR_ARG, "dx must be positive");
    if (cfg->order != 2 && cfg->order != 3) return fail(DMPM_ERR_ARG, "order must be 2 or 3");
    const long long es = dtype == DMPM_DTYPE_F64 ? 8 : 4;
    const long long ld = ((long long)(n > 0 ? n : 1) + 63) / 64 * 64;
    if (4 * ld * es > (long long)DMPM_NOALLOC_SCRATCH_BYTES)
        return fail(DMPM_ERR_ARG, "dmpm_calculate_volumes_noalloc: " + std::to_string(n) +
                                      " particles exceed the scratch of " +
                                      std::to_string(DMPM_NOALLOC_SCRATCH_BYTES) + " bytes (at most " +
                                      std::to_string(DMPM_NOALLOC_SCRATCH_BYTES / (4 * es)) +
                                      " particles for this dtype)");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(DMPM_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(DMPM_ERR_ARG, "device index out of range");
    void* ptr = nullptr;
    {
        // first call on this device (it is current): the array's device address here
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            e = hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                g_scratch_ptr[dev] = nullptr;
                return fail(DMPM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            }
        }
        ptr = g_scratch_ptr[dev];
    }
    if (n == 0) return DMPM_OK;
    const Prec pr = dtype == DMPM_DTYPE_F64 ? PREC_F64 : PREC_F32;
    // the scratch planes are fully written (import) before any kernel reads
    // them, so nothing in them needs clearing
    try {
        mass_volume_pass(pr, *cfg, n, ld, x, m, ptr, grid_m, vol, (hipStream_t)stream);
    } catch (const std::exception& ex) {
        return fail(DMPM_ERR_ARG, ex.what());
    }
    e = hipGetLastError();
    if (e != hipSuccess) return fail(DMPM_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return DMPM_OK;
}
