// cgs_noalloc.hip -- the centred 3D FFT on a caller's device array without
// any device allocation (cgs_fft_device_noalloc in include/cgs.h).
//
// cgs_fft_forward / cgs_fft_backward keep their twiddle tables (and a
// staging copy of the grid) in device memory allocated by cgs_fft_create.
// This variant transforms a device array in place and keeps the twiddle
// tables of the three axes in `scratch`, a file-scope device array of fixed
// size that every GPU holds once for the life of the process: a caller under
// HIP graph capture, or one that must not allocate, can use it.  The tables
// are rewritten by a small kernel at every call, so the array needs no
// clearing; calls that overlap on one GPU (different streams) share the
// array and race.
#include <cmath>
#include <mutex>

#include "cgs.h"
#include "cgs_internal.h"
#include "cgs_kernels.h"

using namespace cgs;

constexpr int kMaxLen = 512;          // longest supported line (fft_len_supported)
constexpr int kMaxDevices = 64;

// Twiddle tables of the x, y and z axes: exp(-2 pi i n / L) for n < L, axis a
// at offset a * kMaxLen.  12 KiB.
static __device__ __align__(256) float2 scratch[3 * kMaxLen];

namespace {

std::mutex g_mu;
float2* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device

// Device address of this GPU's `scratch` (looked up once per device, with
// that device current).
hipError_t scratch_ptr(float2** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (float2*)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

// tw[a * kMaxLen + n] = exp(-2 pi i n / L_a), in FP64 rounded to float (the
// same expression the host uses for the tables of cgs_fft_create).
__global__ void k_twiddles(float2* tw, int l0, int l1, int l2) {
    const int a = blockIdx.y;
    const int L = a == 0 ? l0 : (a == 1 ? l1 : l2);
    const double two_pi = 6.283185307179586476925286766559;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L; n += gridDim.x * blockDim.x) {
        const double ang = -two_pi * (double)n / (double)L;
        tw[a * kMaxLen + n] = make_float2((float)cos(ang), (float)sin(ang));
    }
}

}  // namespace

extern "C" int cgs_fft_device_noalloc(const int dims[3], float* a_dev, int dir, int rescale, void* stream) {
    if (!dims || !a_dev || (dir != -1 && dir != 1)) return CGS_ERR_ARG;
    for (int a = 0; a < 3; ++a)
        if (dims[a] > kMaxLen || !fft_len_supported(dims[a])) return CGS_ERR_ARG;
    float2* tw = nullptr;
    if (scratch_ptr(&tw) != hipSuccess) return CGS_ERR_HIP;
    hipStream_t st = (hipStream_t)stream;
    const long long ng = (long long)dims[0] * dims[1] * dims[2];
    float2* g = (float2*)a_dev;
    k_twiddles<<<dim3(2, 3), 256, 0, st>>>(tw, dims[0], dims[1], dims[2]);
    if (hipGetLastError() != hipSuccess) return CGS_ERR_HIP;
    if (launch_fft_z1(dims[2], dir, g, (long long)dims[0] * dims[1], tw + 2 * kMaxLen, st) != hipSuccess)
        return CGS_ERR_HIP;
    PlainArgs y{};
    y.g = g; y.os = (long long)dims[1] * dims[2]; y.ls = dims[2]; y.ncols = dims[2]; y.tw = tw + kMaxLen;
    y.outer_per_bin = dims[0];
    if (launch_fft_plain(dims[1], dir, y, dims[0], st) != hipSuccess) return CGS_ERR_HIP;
    PlainArgs x{};
    x.g = g; x.os = ng; x.ls = (long long)dims[1] * dims[2]; x.ncols = x.ls; x.tw = tw;
    x.outer_per_bin = 1;
    if (launch_fft_plain(dims[0], dir, x, 1, st) != hipSuccess) return CGS_ERR_HIP;
    if (dir > 0 && rescale && launch_scale_d(g, ng, 1.0 / std::sqrt((double)ng), st) != hipSuccess)
        return CGS_ERR_HIP;
    return CGS_OK;
}
