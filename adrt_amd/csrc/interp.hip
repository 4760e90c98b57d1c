// Nearest-neighbour interpolation of an ADRT output onto a regular
// Cartesian (theta, t) grid: adrt.utils.interp_to_cart
// (adrt_cdefs_interp_adrtcart.hpp:60-114, binding adrt_cdefs_py.cpp:621-682,
// which computes its index math in float).
//
// out[b][o][a] (o < n offsets, a < 4n angles) = factor(a) *
// in[b][q(a)][hi(o, a)][si(a)], or 0 when hi falls outside [0, 2n-1).  Every
// transcendental quantity depends on the angle alone (q, si, tan, cos, the
// scale factor) or on the offset alone (t), so the host computes those 5n
// values with the C++ standard functions, in float as the reference does,
// and the device evaluates only the per-element arithmetic
//   h0 = (0.5 + tan/2) + (+-t)/cos,  hi = (round(h0 * 2n) - 1) / 2
// with the same IEEE operations in the same order (no contraction), then
// gathers.  The result is therefore the reference's, bit for bit, whatever
// the device's own tan/cos would round to.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <type_traits>

#include "common.hpp"

namespace adrt_hip {
namespace interp {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kSqrt2 = 1.41421356237309504880f;

// Table layout (host buffer and its device copy): for A = 4n angles
//   int32 q[A], int32 si[A], int32 sgn[A], float tan[A], float cos[A],
//   T factor[A] (8-byte aligned), float t[n]
struct Layout {
    size_t q, si, sgn, tan, cos, factor, t, bytes;
};

inline Layout layout(size_t n, size_t elem) {
    const size_t A = 4 * n;
    Layout l;
    l.q = 0;
    l.si = l.q + 4 * A;
    l.sgn = l.si + 4 * A;
    l.tan = l.sgn + 4 * A;
    l.cos = l.tan + 4 * A;
    l.factor = (l.cos + 4 * A + 7) / 8 * 8;
    l.t = l.factor + elem * A;
    l.bytes = l.t + 4 * n;
    return l;
}

// lerp(a, b, f) for endpoints of opposite signs: f * b + (1 - f) * a, the
// form C++'s std::lerp takes for such endpoints (exact at both ends).
inline float lerp_opposite(float a, float b, float f) { return f * b + (1.0f - f) * a; }

template <typename T>
void fill(size_t n, unsigned char *buf) {
    using larger = std::conditional_t<(std::numeric_limits<T>::digits > std::numeric_limits<float>::digits),
                                      T, float>;
    const Layout L = layout(n, sizeof(T));
    const size_t A = 4 * n;
    int *q = reinterpret_cast<int *>(buf + L.q);
    int *si = reinterpret_cast<int *>(buf + L.si);
    int *sgn = reinterpret_cast<int *>(buf + L.sgn);
    float *tn = reinterpret_cast<float *>(buf + L.tan);
    float *cs = reinterpret_cast<float *>(buf + L.cos);
    T *factor = reinterpret_cast<T *>(buf + L.factor);
    float *t = reinterpret_cast<float *>(buf + L.t);
    const float pi_2 = kPi / 2.0f, pi_4 = kPi / 4.0f, pi_8 = kPi / 8.0f;
    const float s22 = kSqrt2 / 2.0f;
    const float t_left = s22 - (s22 / static_cast<float>(n));
    const float th_left = pi_2 - (pi_8 / static_cast<float>(n));
    for (size_t a = 0; a < A; ++a) {
        const float frac = static_cast<float>(a) / static_cast<float>(A - 1);
        const float th = th_left * lerp_opposite(1.0f, -1.0f, frac);
        const int qa = static_cast<int>(std::floor(std::clamp(-th / pi_4, -2.0f, 1.0f))) + 2;
        const float th0 = pi_4 - std::abs(std::abs(th) - pi_4);
        const float tan_theta = std::clamp(std::tan(th0), 0.0f, 1.0f);
        const float s = std::round(tan_theta * static_cast<float>(n - 1));
        const larger sidea = static_cast<larger>(s) / static_cast<larger>(n - 1);
        q[a] = qa;
        si[a] = static_cast<int>(s);
        sgn[a] = qa % 2 == 0 ? 1 : -1;
        tn[a] = tan_theta;
        cs[a] = std::cos(th0);
        factor[a] = static_cast<T>(std::sqrt(sidea * sidea + static_cast<larger>(1)));
    }
    for (size_t o = 0; o < n; ++o) {
        const float frac = static_cast<float>(o) / static_cast<float>(n - 1);
        t[o] = t_left * lerp_opposite(-1.0f, 1.0f, frac);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    interp_kernel(const T *__restrict__ in, T *__restrict__ out, const unsigned char *__restrict__ tab,
                  Layout L, unsigned n) {
    const unsigned A = 4 * n;
    const unsigned chunks = (A + kBlock - 1) / kBlock;
    const size_t row = blockIdx.x / chunks;  // b * n + o
    const unsigned a = (blockIdx.x % chunks) * kBlock + threadIdx.x;
    if (a >= A) return;
    const unsigned o = unsigned(row % n);
    const size_t b = row / n;
    const int q = reinterpret_cast<const int *>(tab + L.q)[a];
    const int si = reinterpret_cast<const int *>(tab + L.si)[a];
    const int sgn = reinterpret_cast<const int *>(tab + L.sgn)[a];
    const float tn = reinterpret_cast<const float *>(tab + L.tan)[a];
    const float cs = reinterpret_cast<const float *>(tab + L.cos)[a];
    const T factor = reinterpret_cast<const T *>(tab + L.factor)[a];
    const float t = reinterpret_cast<const float *>(tab + L.t)[o];
    const float h0 = (0.5f + (tn / 2.0f)) + ((sgn >= 0 ? t : -t) / cs);
    const float hi = (roundf(h0 * static_cast<float>(2 * n)) - 1.0f) / 2.0f;
    T v = T(0);
    if (hi >= 0.0f && hi < static_cast<float>(2 * n - 1)) {
        const size_t h = static_cast<size_t>(hi);
        v = factor * in[((b * 4 + size_t(q)) * (2 * size_t(n) - 1) + h) * n + size_t(si)];
    }
    out[(size_t(row)) * A + a] = v;
}

}  // namespace interp

using namespace interp;

static bool interp_shape_ok(size_t batch, size_t n) {
    // the output shape rule (a valid ADRT output, n > 1) and the float index
    // limit: 4n must be exactly representable in float
    if (batch < 1 || n < 2 || !is_pow2(n)) return false;
    return true;
}

extern "C" {

size_t adrt_hip_interp_table_bytes(int dtype, size_t n) {
    if ((dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) || n < 2 || !is_pow2(n)) return 0;
    return layout(n, dtype == ADRT_HIP_F64 ? 8 : 4).bytes;
}

int adrt_hip_interp_table(int dtype, size_t n, void *host_buf, size_t bytes) {
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (n < 2 || !is_pow2(n)) return ADRT_HIP_ERR_SHAPE;
    if (n > (size_t(1) << std::numeric_limits<float>::digits) / 4) return ADRT_HIP_ERR_TOO_BIG;
    if (!host_buf) return ADRT_HIP_ERR_NULL;
    if (bytes < adrt_hip_interp_table_bytes(dtype, n)) return ADRT_HIP_ERR_WORKSPACE;
    std::memset(host_buf, 0, adrt_hip_interp_table_bytes(dtype, n));
    if (dtype == ADRT_HIP_F32)
        fill<float>(n, static_cast<unsigned char *>(host_buf));
    else
        fill<double>(n, static_cast<unsigned char *>(host_buf));
    return ADRT_HIP_OK;
}

int adrt_hip_interp_to_cart(const void *in, void *out, const void *table, size_t table_bytes,
                            int dtype, size_t batch, size_t n, void *stream) {
    if (!interp_shape_ok(batch, n)) return ADRT_HIP_ERR_SHAPE;
    if (n > (size_t(1) << std::numeric_limits<float>::digits) / 4) return ADRT_HIP_ERR_TOO_BIG;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out || !table) return ADRT_HIP_ERR_NULL;
    const size_t elem = dtype == ADRT_HIP_F64 ? 8 : 4;
    if (table_bytes < layout(n, elem).bytes) return ADRT_HIP_ERR_WORKSPACE;
    const Layout L = layout(n, elem);
    const unsigned A = unsigned(4 * n);
    // one block per (output row b * n + o, chunk of 256 angles)
    const size_t blocks = batch * n * ((A + kBlock - 1) / kBlock);
    if (blocks > 0x7fffffffu) return ADRT_HIP_ERR_TOO_BIG;
    const dim3 grid(static_cast<unsigned>(blocks), 1, 1);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(interp_kernel<float>, grid, dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out),
                           static_cast<const unsigned char *>(table), L, unsigned(n));
    else
        hipLaunchKernelGGL(interp_kernel<double>, grid, dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out),
                           static_cast<const unsigned char *>(table), L, unsigned(n));
    return launch_status();
}

}  // extern "C"

}  // namespace adrt_hip
