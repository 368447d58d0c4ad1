// Mesh sampling support: generalised winding numbers of lattice points with
// respect to a triangle mesh, computed on the device.  The caller builds the
// lattice and keeps the points with w >= 1/2, which is the inside test of
// GeometryLoading::GeneratePointCloudFromWatertightTriangleMesh
// (GeometryLoading.cpp:10-83: libigl's signed distance with sign type
// SIGNED_DISTANCE_TYPE_WINDING_NUMBER, S <= 0 <=> w >= 1/2).
#include "../../include/dmpm.h"
#include "dmpm_kernels.h"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct MeshError : std::runtime_error {
    int code;
    MeshError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hipchk(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw MeshError(DMPM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Device buffers freed on every exit path.
struct DevBuf {
    void* p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (bytes == 0) bytes = 16;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            throw MeshError(DMPM_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int dmpm_winding_numbers(const double* V, int nv, const int* F, int nf, const double* P, int np,
                                    double* w) {
    try {
        if (nv < 0 || nf < 0 || np < 0) throw MeshError(DMPM_ERR_ARG, "negative size");
        if (np == 0) return DMPM_OK;
        if (!P || !w || (nf > 0 && (!V || !F))) throw MeshError(DMPM_ERR_ARG, "null array");
        std::vector<double> tri((size_t)nf * 9);
        for (int t = 0; t < nf; ++t)
            for (int c = 0; c < 3; ++c) {
                const int v = F[3 * (size_t)t + c];
                if (v < 0 || v >= nv)
                    throw MeshError(DMPM_ERR_ARG, "triangle " + std::to_string(t) + " has vertex index " +
                                                      std::to_string(v) + " outside [0, " + std::to_string(nv) + ")");
                for (int d = 0; d < 3; ++d) tri[9 * (size_t)t + 3 * c + d] = V[3 * (size_t)v + d];
            }
        // split the triangles over enough blocks to fill the chip, with the
        // partial sums bounded to 512 MiB
        const long long nbx = (np + 255) / 256;
        long long chunks = std::max(1LL, std::min<long long>((nf + 255) / 256, (4096 + nbx - 1) / nbx));
        chunks = std::max(1LL, std::min(chunks, (512LL << 20) / (8LL * np)));
        hipStream_t s = nullptr;
        hipchk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } sg{s};
        DevBuf dtri(tri.size() * sizeof(double)), dp((size_t)np * 3 * sizeof(double)),
            dpart((size_t)chunks * np * sizeof(double)), dw((size_t)np * sizeof(double));
        hipchk(hipMemcpyAsync(dtri.p, tri.data(), tri.size() * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        hipchk(hipMemcpyAsync(dp.p, P, (size_t)np * 3 * sizeof(double), hipMemcpyHostToDevice, s), "copy");
        dmpm::winding_numbers((const double*)dtri.p, nf, (const double*)dp.p, np, (double*)dpart.p, (int)chunks,
                              (double*)dw.p, s);
        hipchk(hipGetLastError(), "winding-number launch");
        hipchk(hipMemcpyAsync(w, dw.p, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        hipchk(hipStreamSynchronize(s), "hipStreamSynchronize");
        return DMPM_OK;
    } catch (const MeshError& e) {
        dmpm::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        dmpm::set_last_error(e.what());
        return DMPM_ERR_ARG;
    }
}
