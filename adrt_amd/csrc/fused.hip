// Fused two-pass forward ADRT and back-projection for gfx950 (log2 n in
// [6, 12]): one kernel per pass over a launch group of images, the stage
// split and launch-group configuration the autotuner sets, and the C ABI
// glue.  The per-item code, and the account of how the passes reproduce the
// reference's additions exactly, is in fused_items.hpp.

#include <atomic>
#include <mutex>

#include "fused_items.hpp"

namespace adrt_hip {
namespace fused {

#ifndef ADRT_HOST_EMULATION
template <typename T, int KB>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(lds_waves_per_simd<Pass1<T, KB>, T>(
                                   sizeof(T) == 4 ? Pass1<T, KB>::kSimdWaves : 4))))
fwd_pass1(const T *__restrict__ img_base, T *__restrict__ X_base, int n, int first_image,
          int items) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    run_item<Pass1<T, KB>, T>(items, reinterpret_cast<T *>(smem_raw), img_base, X_base, n,
                              first_image);
}

template <typename T, int KD>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(lds_waves_per_simd<Pass2<T, KD>, T>(
                                   sizeof(T) == 4 ? 8 : 4))))
fwd_pass2(const T *__restrict__ X_base, T *__restrict__ out_base, int n, int k1,
          int first_image, int items) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    run_item<Pass2<T, KD>, T>(items, reinterpret_cast<T *>(smem_raw), X_base, out_base, n, k1,
                               first_image);
}
template <typename T, int K2>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(lds_waves_per_simd<BdrtA<T, K2>, T>(
                                   sizeof(T) == 4 ? 8 : 4))))
bdrt_passA(const T *__restrict__ y_base, T *__restrict__ Z_base, int n, int first_image,
           int zrows, int items) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    run_item<BdrtA<T, K2>, T>(items, reinterpret_cast<T *>(smem_raw), y_base, Z_base, n,
                              first_image, zrows);
}

template <typename T, int K4>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(lds_waves_per_simd<BdrtB<T, K4>, T>(
                                   sizeof(T) == 4 ? 8 : 4))))
bdrt_passB(const T *__restrict__ Z_base, T *__restrict__ out_base, int n, int kA,
           int first_image, int rows, int items) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    run_item<BdrtB<T, K4>, T>(items, reinterpret_cast<T *>(smem_raw), Z_base, out_base, n, kA,
                              first_image, rows);
}
#endif  // ADRT_HOST_EMULATION

#ifndef ADRT_HOST_EMULATION
// Raises the dynamic-LDS limit of kernel `Kern` to `lds` bytes, once per
// device and kernel (each kernel address is its own instantiation, so the
// once-flags are per kernel).  A failure is cleared from the runtime's
// error state and returned, so the caller reports it instead of launching.
template <auto Kern>
static hipError_t lds_limit(size_t lds) {
    if (lds == 0) return hipSuccess;
    constexpr int kMaxDevices = 64;
    static std::once_flag once[kMaxDevices];
    static hipError_t result[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return hipErrorInvalidDevice;
    }
    std::call_once(once[dev], [&] {
        result[dev] = hipFuncSetAttribute(reinterpret_cast<const void *>(Kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        if (result[dev] != hipSuccess) (void)hipGetLastError();
    });
    return result[dev];
}

// One work-group per item, rounded up to whole rows of the 8 XCDs.
static unsigned launch_grid(int items) { return unsigned((items + kXcds - 1) / kXcds * kXcds); }

template <typename T, int KB>
static hipError_t launch_pass1(const T *img, T *X, int n, int first, int count, hipStream_t s) {
    using P = Pass1<T, KB>;
    const int items = P::num_items(n, count);
    constexpr size_t lds = P::lds_elems * sizeof(T);
    const hipError_t e = lds_limit<fwd_pass1<T, KB>>(lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fwd_pass1<T, KB>), dim3(launch_grid(items)), dim3(kThreads), lds, s, img,
                       X, n, first, items);
    return hipSuccess;
}

template <typename T, int KD>
static hipError_t launch_pass2(const T *X, T *out, int n, int k1, int first, int count,
                               hipStream_t s) {
    using P = Pass2<T, KD>;
    const int items = P::num_items(n, k1, count);
    constexpr size_t lds = P::lds_elems * sizeof(T);
    const hipError_t e = lds_limit<fwd_pass2<T, KD>>(lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fwd_pass2<T, KD>), dim3(launch_grid(items)), dim3(kThreads), lds, s, X,
                       out, n, k1, first, items);
    return hipSuccess;
}

template <typename T>
static hipError_t run_group(const T *img, T *X, T *out, int n, int L, int k1, int first,
                            int count, hipStream_t s) {
    hipError_t e;
    switch (k1 - 3) {
        case 0: e = launch_pass1<T, 0>(img, X, n, first, count, s); break;
        case 1: e = launch_pass1<T, 1>(img, X, n, first, count, s); break;
        case 2: e = launch_pass1<T, 2>(img, X, n, first, count, s); break;
        default: e = launch_pass1<T, 3>(img, X, n, first, count, s); break;
    }
    if (e != hipSuccess) return e;
    switch (L - k1 - 3) {
        case 0: return launch_pass2<T, 0>(X, out, n, k1, first, count, s);
        case 1: return launch_pass2<T, 1>(X, out, n, k1, first, count, s);
        case 2: return launch_pass2<T, 2>(X, out, n, k1, first, count, s);
        default: return launch_pass2<T, 3>(X, out, n, k1, first, count, s);
    }
}

template <typename T, int K2>
static hipError_t launch_bdrt_a(const T *y, T *Z, int n, int first, int count, int zrows,
                                hipStream_t s) {
    using P = BdrtA<T, K2>;
    const int items = P::num_items(n, count);
    constexpr size_t lds = P::lds_elems * sizeof(T);
    const hipError_t e = lds_limit<bdrt_passA<T, K2>>(lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((bdrt_passA<T, K2>), dim3(launch_grid(items)), dim3(kThreads), lds, s, y,
                       Z, n, first, zrows, items);
    return hipSuccess;
}

template <typename T, int K4>
static hipError_t launch_bdrt_b(const T *Z, T *out, int n, int kA, int first, int count, int rows,
                                hipStream_t s) {
    using P = BdrtB<T, K4>;
    const int items = P::num_items(kA, count, rows);
    constexpr size_t lds = P::lds_elems * sizeof(T);
    const hipError_t e = lds_limit<bdrt_passB<T, K4>>(lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((bdrt_passB<T, K4>), dim3(launch_grid(items)), dim3(kThreads), lds, s, Z,
                       out, n, kA, first, rows, items);
    return hipSuccess;
}

// bdrt split: pass A takes kA stages, pass B kB = L - kA.  Output rows
// [0, rows) of every quadrant are stored (rows = 2n - 1: all of them);
// pass A then skips the items whose Z rows those do not read.
template <typename T>
static hipError_t run_bdrt_group(const T *y, T *Z, T *out, int n, int L, int kA, int first,
                                 int count, int rows, hipStream_t s) {
    hipError_t e;
    const int zrows = bdrt_z_rows(n, rows, L - kA);
    switch (kA - 3) {
        case 0: e = launch_bdrt_a<T, 0>(y, Z, n, first, count, zrows, s); break;
        case 1: e = launch_bdrt_a<T, 1>(y, Z, n, first, count, zrows, s); break;
        case 2: e = launch_bdrt_a<T, 2>(y, Z, n, first, count, zrows, s); break;
        default: e = launch_bdrt_a<T, 3>(y, Z, n, first, count, zrows, s); break;
    }
    if (e != hipSuccess) return e;
    switch (L - kA - 3) {
        case 0: return launch_bdrt_b<T, 0>(Z, out, n, kA, first, count, rows, s);
        case 1: return launch_bdrt_b<T, 1>(Z, out, n, kA, first, count, rows, s);
        case 2: return launch_bdrt_b<T, 2>(Z, out, n, kA, first, count, rows, s);
        default: return launch_bdrt_b<T, 3>(Z, out, n, kA, first, count, rows, s);
    }
}
#endif  // ADRT_HOST_EMULATION

}  // namespace fused

#ifndef ADRT_HOST_EMULATION
bool fused_supported(size_t n) {
    const int L = num_iters(n);
    return is_pow2(n) && L >= 6 && L <= 12;
}

// Per-(op, log2 n) configuration, set by the autotuner through the C ABI.
namespace {
struct Tuned {
    std::atomic<int> variant{ADRT_HIP_VARIANT_DEFAULT};
    std::atomic<size_t> group_bytes{0};
};
Tuned g_tuned[3][64];

Tuned *tuned(int op, size_t n) {
    if ((op != ADRT_HIP_OP_ADRT && op != ADRT_HIP_OP_BDRT && op != ADRT_HIP_OP_IADRT) ||
        !is_pow2(n))
        return nullptr;
    return &g_tuned[op][num_iters(n)];
}

size_t group_budget(int op, size_t n) {
    const Tuned *t = tuned(op, n);
    const size_t b = t ? t->group_bytes.load(std::memory_order_relaxed) : 0;
    return b ? b : fused::kDefaultGroupBytes;
}
}  // namespace

int fused_variant(int op, size_t n) {
    const Tuned *t = tuned(op, n);
    return t ? t->variant.load(std::memory_order_relaxed) : ADRT_HIP_VARIANT_DEFAULT;
}

// The one-launch LDS path: the default up to n = 64 (at 64 it measured as
// fast as the fused kernels or faster: 40 / 37 us against 42 / 45 us for
// forward / bdrt of 256 fp32 images, 55 / 61 against 75 / 77 us in fp64).
static bool small_chosen(int op, size_t n) {
    if (op != ADRT_HIP_OP_ADRT && op != ADRT_HIP_OP_BDRT) return false;
    const int v = fused_variant(op, n);
    return v == ADRT_HIP_VARIANT_SINGLE || (v == ADRT_HIP_VARIANT_DEFAULT && n <= 64);
}

bool use_small(int op, size_t B, size_t n) { return small_chosen(op, n) && small_supported(B, n); }

// The fused kernels: log2 n in [6, 12] unless streaming or the LDS path is
// configured (at n = 64 they are reached through ALT_SPLIT, whose split
// equals the default one for even log2 n).
bool use_fused(int op, size_t n) {
    const int v = fused_variant(op, n);
    return fused_supported(n) && v != ADRT_HIP_VARIANT_STREAMING && !small_chosen(op, n);
}

int set_fused_config(int op, size_t n, int variant, size_t group_bytes) {
    Tuned *t = tuned(op, n);
    if (!t) return ADRT_HIP_ERR_ARG;
    if (variant < ADRT_HIP_VARIANT_DEFAULT || variant > ADRT_HIP_VARIANT_SINGLE)
        return ADRT_HIP_ERR_ARG;
    if (variant == ADRT_HIP_VARIANT_ALT_SPLIT && op != ADRT_HIP_OP_IADRT && !fused_supported(n))
        return ADRT_HIP_ERR_ARG;
    if (variant == ADRT_HIP_VARIANT_SINGLE && (op == ADRT_HIP_OP_IADRT || !small_supported(1, n)))
        return ADRT_HIP_ERR_ARG;
    t->variant.store(variant, std::memory_order_relaxed);
    t->group_bytes.store(group_bytes, std::memory_order_relaxed);
    return ADRT_HIP_OK;
}

int get_fused_config(int op, size_t n, int *variant, size_t *group_bytes) {
    const Tuned *t = tuned(op, n);
    if (!t) return ADRT_HIP_ERR_ARG;
    if (variant) *variant = t->variant.load(std::memory_order_relaxed);
    if (group_bytes) *group_bytes = group_budget(op, n);
    return ADRT_HIP_OK;
}

size_t fused_workspace(int dtype, size_t B, size_t n) {
    const size_t elem = dtype == ADRT_HIP_F64 ? 8 : 4;
    const int L = num_iters(n);
    const size_t per = fused::x_bytes_per_image(int(n), L, elem);
    return fused::group_images(B, per, group_budget(ADRT_HIP_OP_ADRT, n)) * per;
}

size_t fused_workspace_per_image(int dtype, size_t n) {
    return fused::x_bytes_per_image(int(n), num_iters(n), dtype == ADRT_HIP_F64 ? 8 : 4);
}

template <typename T>
static int run_fused_t(const T *in, T *out, T *ws, size_t B, size_t n, size_t budget,
                       hipStream_t s) {
    const int L = num_iters(n);
    const int k1 = fused::fwd_pass1_stages(L, fused_variant(ADRT_HIP_OP_ADRT, n), int(sizeof(T)));
    const size_t per = fused::x_bytes_per_image(int(n), L, sizeof(T));
    const size_t G = fused::group_images(B, per, budget);
    for (size_t first = 0; first < B; first += G) {
        const size_t count = (B - first) < G ? (B - first) : G;
        const hipError_t e =
            fused::run_group<T>(in, ws, out, int(n), L, k1, int(first), int(count), s);
        if (e != hipSuccess) {
            set_last_error(e);
            return ADRT_HIP_ERR_DEVICE;
        }
    }
    return launch_status();
}

int run_fused_budget(const void *in, void *out, void *ws, int dtype, size_t B, size_t n,
                     size_t group_bytes, hipStream_t s) {
    if (dtype == ADRT_HIP_F32)
        return run_fused_t(static_cast<const float *>(in), static_cast<float *>(out),
                           static_cast<float *>(ws), B, n, group_bytes, s);
    return run_fused_t(static_cast<const double *>(in), static_cast<double *>(out),
                       static_cast<double *>(ws), B, n, group_bytes, s);
}

int run_fused(const void *in, void *out, void *ws, int dtype, size_t B, size_t n, hipStream_t s) {
    return run_fused_budget(in, out, ws, dtype, B, n, group_budget(ADRT_HIP_OP_ADRT, n), s);
}

size_t fused_bdrt_workspace(int dtype, size_t B, size_t n) {
    const size_t elem = dtype == ADRT_HIP_F64 ? 8 : 4;
    const size_t per = fused::z_bytes_per_image(int(n), elem);
    return fused::group_images(B, per, group_budget(ADRT_HIP_OP_BDRT, n)) * per;
}

template <typename T>
static int run_fused_bdrt_t(const T *in, T *out, T *ws, size_t B, size_t n, int rows,
                            hipStream_t s) {
    const int L = num_iters(n);
    const int kA = fused::bdrt_passA_stages(L, fused_variant(ADRT_HIP_OP_BDRT, n));
    const size_t per = fused::z_bytes_per_image(int(n), sizeof(T));
    const size_t G = fused::group_images(B, per, group_budget(ADRT_HIP_OP_BDRT, n));
    for (size_t first = 0; first < B; first += G) {
        const size_t count = (B - first) < G ? (B - first) : G;
        const hipError_t e =
            fused::run_bdrt_group<T>(in, ws, out, int(n), L, kA, int(first), int(count), rows, s);
        if (e != hipSuccess) {
            set_last_error(e);
            return ADRT_HIP_ERR_DEVICE;
        }
    }
    return launch_status();
}

int run_fused_bdrt(const void *in, void *out, void *ws, int dtype, size_t B, size_t n, int rows,
                   hipStream_t s) {
    if (dtype == ADRT_HIP_F32)
        return run_fused_bdrt_t(static_cast<const float *>(in), static_cast<float *>(out),
                                static_cast<float *>(ws), B, n, rows, s);
    return run_fused_bdrt_t(static_cast<const double *>(in), static_cast<double *>(out),
                            static_cast<double *>(ws), B, n, rows, s);
}
#endif  // ADRT_HOST_EMULATION

}  // namespace adrt_hip
