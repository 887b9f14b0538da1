// H|psi> on device buffers without any device allocation (sed_matvec_device_noalloc).
//
// No reference counterpart.  For callers that cannot allocate (code under HIP graph
// capture, tight serving loops): the one scratch vector a matvec needs (xhat =
// sqrt(stab) D x, the prescaled input in the gauge-free basis) lives in a file-scope
// device array of kScratchBytes, one copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel: its device address is looked up once
// per device (hipGetSymbolAddress, at sed_create) and kept in g_scratch[device].
#include <mutex>

#include "sed_internal.hpp"
#include "sed_mv.hpp"

namespace sed {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};
static std::string g_scratch_err[kMaxDevices];

// Device address of this device's scratch array (the device is current).  Called by
// sed_create for its device.
int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(SED_EINVAL, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return SED_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        return fail(SED_EDEVICE, g_scratch_err[device]);
    }
    g_scratch[device] = p;
    return SED_OK;
}

// xhat[r] = sqrt(stab[r]) * (F[r] x[r]) with the operations of the allocating path: the
// gauge product in FP64, rounded to the vector type, then the prescale in that type.
template <typename T>
__global__ void __launch_bounds__(256) k_gauge_prescale(const T* __restrict__ x, const double2* __restrict__ F,
                                                        const uint8_t* __restrict__ stab, T* __restrict__ xh,
                                                        int64_t n) {
    using R = decltype(T().x);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        T v = x[r];
        if (F) {
            const double2 a = cmul(F[r], make_double2((double)v.x, (double)v.y));
            v.x = (R)a.x;
            v.y = (R)a.y;
        }
        const R s = sqrt((R)stab[r]);
        T o;
        o.x = v.x * s;
        o.y = v.y * s;
        xh[r] = o;
    }
}

// y <- conj(F) y in place (back to the sector basis), F in FP64
template <typename T>
__global__ void __launch_bounds__(256) k_gauge_back(T* __restrict__ y, const double2* __restrict__ F, int64_t n) {
    using R = decltype(T().x);
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
        double2 f = F[r];
        f.y = -f.y;
        const double2 a = cmul(f, make_double2((double)y[r].x, (double)y[r].y));
        T o;
        o.x = (R)a.x;
        o.y = (R)a.y;
        y[r] = o;
    }
}

}  // namespace sed

using namespace sed;

extern "C" int sed_matvec_device_noalloc(sed_sector* sec, const void* x_dev, void* y_dev, int precision) {
    if (!sec || !x_dev || !y_dev) return fail(SED_EINVAL, "NULL argument");
    if (precision != SED_PREC_FP64 && precision != SED_PREC_FP32) return fail(SED_EINVAL, "bad precision");
    if (!sec->has_terms) return fail(SED_ESTATE, "matvec before sed_sector_set_terms");
    const int64_t n = sec->dim;
    const bool fp64 = precision == SED_PREC_FP64;
    const size_t need = (size_t)n * (fp64 ? 16 : 8);
    if (need > kScratchBytes)
        return fail(SED_EINVAL, "sector dimension " + std::to_string(n) + " needs " + std::to_string(need) +
                                    " bytes of scratch; the no-allocation matvec holds " +
                                    std::to_string(kScratchBytes) + " per GPU");
    const int dev = sec->cat->ctx->device;
    void* xh = nullptr;
    std::string why;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (dev >= 0 && dev < kMaxDevices) {
            xh = g_scratch[dev];
            why = g_scratch_err[dev];
        }
    }
    if (!xh) return fail(SED_ESTATE, "scratch array not set up on this device at sed_create: " + why);
    if (n == 0) return SED_OK;
    SED_HIP(hipSetDevice(dev));
    hipStream_t s = sec->cat->ctx->stream;
    const double2* F = sec->has_gauge ? reinterpret_cast<const double2*>(sec->d_gauge.p) : nullptr;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 20);
    if (fp64)
        hipLaunchKernelGGL(k_gauge_prescale<double2>, dim3(g), dim3(256), 0, s,
                           reinterpret_cast<const double2*>(x_dev), F, sec->d_stab.p, reinterpret_cast<double2*>(xh),
                           n);
    else
        hipLaunchKernelGGL(k_gauge_prescale<float2>, dim3(g), dim3(256), 0, s, reinterpret_cast<const float2*>(x_dev),
                           F, sec->d_stab.p, reinterpret_cast<float2*>(xh), n);
    SED_HIP(hipGetLastError());
    SED_TRY(launch_matvec_xhat_fixed(sec, xh, y_dev, fp64, s));
    if (F) {
        if (fp64)
            hipLaunchKernelGGL(k_gauge_back<double2>, dim3(g), dim3(256), 0, s, reinterpret_cast<double2*>(y_dev), F, n);
        else
            hipLaunchKernelGGL(k_gauge_back<float2>, dim3(g), dim3(256), 0, s, reinterpret_cast<float2*>(y_dev), F, n);
        SED_HIP(hipGetLastError());
    }
    return SED_OK;
}
