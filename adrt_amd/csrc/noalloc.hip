// Forward ADRT that allocates nothing: adrt_hip_adrt_noalloc.
//
// The fused two-pass forward needs an intermediate X between its passes
// (4 n (n + 2^k1 + margins) elements per image).  adrt_hip_adrt takes it
// from a caller workspace; this entry point keeps it in a file-scope device
// array instead, so that callers which cannot allocate (code under HIP graph
// capture, tight serving loops) pass only the input and the output.
//
// The array holds ADRT_HIP_NOALLOC_SCRATCH_BYTES on every device the library
// runs on, until the process ends.  Its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) on the first call
// made on that device, and only that address is ever handed to kernels: in
// host code the array's name is a host-side stand-in, not a device pointer.
//
// Every call on a device shares that device's array and leaves its contents
// behind.  Nothing in it has to start at zero: pass 1 writes every element
// of X that pass 2 reads (the rows and their -0.0 / +0.0 margins) before
// pass 2 runs, in stream order.  Calls that overlap on one device (other
// streams, other threads) race on the array and must be serialised by the
// caller.
#include <mutex>

#include "common.hpp"

namespace adrt_hip {
namespace {

static_assert(ADRT_HIP_NOALLOC_SCRATCH_BYTES % 256 == 0, "whole 256-byte lines");
__device__ __align__(256) unsigned char scratch[ADRT_HIP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::once_flag g_once[kMaxDevices];
void *g_scratch[kMaxDevices];
hipError_t g_lookup[kMaxDevices];

// Device address of `scratch` on the current device (nullptr on failure,
// with the runtime's error in *err).
void *scratch_address(hipError_t *err) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        *err = e != hipSuccess ? e : hipErrorInvalidDevice;
        return nullptr;
    }
    std::call_once(g_once[dev], [dev] {
        void *p = nullptr;
        g_lookup[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_lookup[dev] != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
        }
        g_scratch[dev] = p;
    });
    *err = g_lookup[dev];
    return g_scratch[dev];
}

}  // namespace
}  // namespace adrt_hip

using namespace adrt_hip;

extern "C" int adrt_hip_adrt_noalloc(const void *in, void *out, int dtype, size_t batch, size_t n,
                                     void *stream) {
    const size_t shape[3] = {batch, n, n};
    if (!adrt_hip_valid_adrt_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    size_t elems;
    const size_t oshape[4] = {batch, 4, 2 * n - 1, n};
    if (!product_ok(oshape, 4, &elems) || !mul_ok(elems, 8, &elems)) return ADRT_HIP_ERR_TOO_BIG;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // The one-launch LDS path (n <= 64) needs no intermediate at all.
    if (use_small(ADRT_HIP_OP_ADRT, batch, n))
        return launch_adrt(in, out, nullptr, 0, dtype, batch, n, s);
    size_t need;
    if (use_fused(ADRT_HIP_OP_ADRT, n))
        need = fused_workspace_per_image(dtype, n);  // one image per launch group at least
    else
        need = adrt_workspace(dtype, batch, n);  // the streaming path: output-sized
    if (need > ADRT_HIP_NOALLOC_SCRATCH_BYTES) return ADRT_HIP_ERR_SCRATCH;
    hipError_t err = hipSuccess;
    void *ws = scratch_address(&err);
    if (!ws) {
        set_last_error(err);
        return ADRT_HIP_ERR_DEVICE;
    }
    if (use_fused(ADRT_HIP_OP_ADRT, n))
        return run_fused_budget(in, out, ws, dtype, batch, n, ADRT_HIP_NOALLOC_SCRATCH_BYTES, s);
    return launch_adrt(in, out, ws, ADRT_HIP_NOALLOC_SCRATCH_BYTES, dtype, batch, n, s);
}
