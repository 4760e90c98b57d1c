// Shared definitions for the gfx950 ADRT kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/adrt_hip.h"

namespace adrt_hip {

using std::size_t;

constexpr int kBlock = 256;

// Status of the last HIP runtime call that failed on this thread.
void set_last_error(hipError_t err);

// Converts a HIP error (if any) from the last launch into a status code.
int launch_status();

// num_iters(n) = bit_width(n) - has_single_bit(n)  (adrt_cdefs_common.cpp:140-142)
inline int num_iters(size_t n) {
    int width = 0;
    for (size_t v = n; v != 0; v >>= 1) ++width;
    const bool single = n != 0 && (n & (n - 1)) == 0;
    return width - (single ? 1 : 0);
}

inline bool is_pow2(size_t n) { return n != 0 && (n & (n - 1)) == 0; }

// max_size = 2^(digits-1) (adrt_cdefs_common.cpp:62)
constexpr size_t kMaxSize = size_t(1) << (sizeof(size_t) * 8 - 1);

// a * b with overflow detection (adrt_cdefs_common.cpp:64-171 semantics).
inline bool mul_ok(size_t a, size_t b, size_t *out) {
    return !__builtin_mul_overflow(a, b, out);
}

inline bool product_ok(const size_t *shape, int nd, size_t *out) {
    size_t p = 1;
    for (int i = 0; i < nd; ++i) {
        if (!mul_ok(p, shape[i], &p)) return false;
    }
    *out = p;
    return true;
}

inline unsigned grid_for(size_t work, unsigned block = kBlock, size_t cap = 1u << 20) {
    size_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

template <typename T>
struct NegZero;
template <>
struct NegZero<float> {
    __device__ static float value() { return -0.0f; }
};
template <>
struct NegZero<double> {
    __device__ static double value() { return -0.0; }
};

}  // namespace adrt_hip

// Launchers implemented per source file; all return a status code.
namespace adrt_hip {
int launch_adrt(const void *in, void *out, void *ws, size_t ws_bytes, int dtype, size_t B,
                size_t N, hipStream_t s);
size_t adrt_workspace(int dtype, size_t B, size_t N);
bool fused_supported(size_t n);
// Tuned configuration of the fused paths (op = ADRT_HIP_OP_ADRT or _BDRT).
int fused_variant(int op, size_t n);
bool use_fused(int op, size_t n);
int set_fused_config(int op, size_t n, int variant, size_t group_bytes);
int get_fused_config(int op, size_t n, int *variant, size_t *group_bytes);
size_t fused_workspace(int dtype, size_t B, size_t n);
// One-launch LDS path (small.hip) for power-of-two n <= 64: supported sizes,
// whether (op, n) is configured to use it, and one op over one batch.
bool small_supported(size_t B, size_t n);
bool use_small(int op, size_t B, size_t n);
int launch_small(const void *src, void *out, int op, int dtype, size_t B, size_t n,
                 hipStream_t s);
int run_fused(const void *in, void *out, void *ws, int dtype, size_t B, size_t n, hipStream_t s);
// The fused forward with launch groups of at most max(1, group_bytes /
// per-image intermediate) images, and that per-image intermediate size.
int run_fused_budget(const void *in, void *out, void *ws, int dtype, size_t B, size_t n,
                     size_t group_bytes, hipStream_t s);
size_t fused_workspace_per_image(int dtype, size_t n);
size_t fused_bdrt_workspace(int dtype, size_t B, size_t n);
int run_fused_bdrt(const void *in, void *out, void *ws, int dtype, size_t B, size_t n, int rows,
                   hipStream_t s);
int launch_adrt_step(const void *in, void *out, int dtype, size_t B, size_t N, int step,
                     hipStream_t s);
// Stores at least output rows [0, rows) of every quadrant (rows <= 2N - 1);
// the fused path stores exactly those and does only the work they need.
int launch_bdrt(const void *in, void *out, void *ws, size_t ws_bytes, int dtype, size_t B,
                size_t N, int rows, hipStream_t s);
size_t bdrt_workspace(int dtype, size_t B, size_t N);
int launch_bdrt_step(const void *in, void *out, int dtype, size_t B, size_t N, int step,
                     hipStream_t s);
int launch_iadrt(const void *in, void *out, void *ws, size_t ws_bytes, int dtype, size_t B,
                 size_t N, hipStream_t s);
size_t iadrt_workspace(int dtype, size_t B, size_t N);
int launch_fmg_restriction(const void *in, void *out, int dtype, size_t B, size_t N,
                           hipStream_t s);
int launch_fmg_prolongation(const void *in, void *out, int dtype, size_t B, size_t H,
                            size_t W, hipStream_t s);
int launch_fmg_highpass(const void *in, void *out, int dtype, size_t B, size_t H, size_t W,
                        hipStream_t s);
int launch_adrt_init(const void *in, void *out, size_t elem, size_t B, size_t N, hipStream_t s);
int launch_truncate(const void *in, void *out, size_t elem, size_t outer, size_t N,
                    hipStream_t s);
int launch_stitch(const void *in, void *out, size_t elem, size_t outer, size_t n, size_t vc,
                  bool inverse, hipStream_t s);
int launch_truncate_combine(const void *in, void *out, int dtype, size_t B, size_t N,
                            double divisor, int mode, hipStream_t s);
}  // namespace adrt_hip
