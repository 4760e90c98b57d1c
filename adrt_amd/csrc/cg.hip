// Batched conjugate-gradient vector kernels: per-image float64 inner
// products and the fused CG updates, for a batch of B independent systems of
// m unknowns each (the images, C-contiguous).
//
//   dot:       out[b] = sum_i u[b][i] * v[b][i]                 (float64)
//   update:    x += a_b * p,  r -= a_b * q,  out[b] = sum_i r[b][i]^2
//   direction: p = z + c_b * p
//
// with a_b, c_b the float64 scalars rounded to the element type, so each
// update is the same two roundings as the unfused expression x + a * p.
// Sums are deterministic: each block reduces a fixed chunk in a fixed order
// and a second pass adds the chunks of an image in order.
#include "common.hpp"

namespace adrt_hip {
namespace cgk {

constexpr int kThreads = 256;

__device__ __forceinline__ double block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kThreads / 64; ++w) s += red[w];
    return s;
}

// Chunk c of image b: elements [c * chunk, min((c+1) * chunk, m)).
template <typename T>
__global__ void __launch_bounds__(kThreads)
    dot_kernel(const T *__restrict__ u, const T *__restrict__ v, double *__restrict__ partial,
               size_t m, size_t chunk, int chunks) {
    __shared__ double red[kThreads / 64];
    const int c = blockIdx.x;
    const size_t b = blockIdx.y;
    const size_t lo = size_t(c) * chunk;
    const size_t hi = lo + chunk < m ? lo + chunk : m;
    const T *ub = u + b * m, *vb = v + b * m;
    double s = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += kThreads) s += double(ub[i]) * double(vb[i]);
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[b * chunks + c] = s;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    update_kernel(T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
                  const T *__restrict__ q, const double *__restrict__ alpha,
                  double *__restrict__ partial, size_t m, size_t chunk, int chunks) {
    __shared__ double red[kThreads / 64];
    const int c = blockIdx.x;
    const size_t b = blockIdx.y;
    const size_t lo = size_t(c) * chunk;
    const size_t hi = lo + chunk < m ? lo + chunk : m;
    const T a = T(alpha[b]);
    T *xb = x + b * m, *rb = r + b * m;
    const T *pb = p + b * m, *qb = q + b * m;
    double s = 0.0;
    for (size_t i = lo + threadIdx.x; i < hi; i += kThreads) {
        const T ap = a * pb[i];
        const T aq = a * qb[i];
        xb[i] = xb[i] + ap;
        const T rn = rb[i] - aq;
        rb[i] = rn;
        s += double(rn) * double(rn);
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) partial[b * chunks + c] = s;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    direction_kernel(T *__restrict__ p, const T *__restrict__ z, const double *__restrict__ beta,
                     size_t m, size_t chunk) {
    const size_t b = blockIdx.y;
    const size_t lo = size_t(blockIdx.x) * chunk;
    const size_t hi = lo + chunk < m ? lo + chunk : m;
    const T c = T(beta[b]);
    T *pb = p + b * m;
    const T *zb = z + b * m;
    for (size_t i = lo + threadIdx.x; i < hi; i += kThreads) {
        const T cp = c * pb[i];
        pb[i] = zb[i] + cp;
    }
}

// out[b] = sum over c of partial[b][c], in order.
__global__ void __launch_bounds__(kThreads)
    rows_kernel(const double *__restrict__ partial, double *__restrict__ out, size_t B, int chunks) {
    const size_t b = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= B) return;
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += partial[b * chunks + c];
    out[b] = s;
}

struct Split {
    size_t chunk;
    int chunks;
};

// About 2048 blocks in all, each chunk at least 16 K elements.
inline Split split(size_t B, size_t m) {
    size_t want = (2048 + B - 1) / B;
    size_t chunk = (m + want - 1) / want;
    if (chunk < 16384) chunk = 16384;
    chunk = (chunk + kThreads - 1) / kThreads * kThreads;
    Split s;
    s.chunk = chunk;
    s.chunks = int((m + chunk - 1) / chunk);
    return s;
}

}  // namespace cgk

using namespace cgk;

static bool batch_ok(size_t B, size_t m) { return B >= 1 && B <= 65535 && m >= 1; }

extern "C" {

size_t adrt_hip_cg_workspace(size_t batch, size_t m) {
    if (!batch_ok(batch, m)) return 0;
    return size_t(split(batch, m).chunks) * batch * sizeof(double);
}

int adrt_hip_batched_dot(const void *u, const void *v, double *out, void *workspace,
                         size_t workspace_bytes, int dtype, size_t batch, size_t m, void *stream) {
    if (!batch_ok(batch, m)) return ADRT_HIP_ERR_SHAPE;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!u || !v || !out || !workspace) return ADRT_HIP_ERR_NULL;
    if (workspace_bytes < adrt_hip_cg_workspace(batch, m)) return ADRT_HIP_ERR_WORKSPACE;
    const Split sp = split(batch, m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *part = static_cast<double *>(workspace);
    const dim3 grid(unsigned(sp.chunks), unsigned(batch));
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(dot_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<const float *>(u), static_cast<const float *>(v), part, m,
                           sp.chunk, sp.chunks);
    else
        hipLaunchKernelGGL(dot_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<const double *>(u), static_cast<const double *>(v), part, m,
                           sp.chunk, sp.chunks);
    hipLaunchKernelGGL(rows_kernel, dim3(unsigned((batch + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, (const double *)part, out, batch, sp.chunks);
    return launch_status();
}

int adrt_hip_cg_update(void *x, void *r, const void *p, const void *q, const double *alpha,
                       double *rr, void *workspace, size_t workspace_bytes, int dtype,
                       size_t batch, size_t m, void *stream) {
    if (!batch_ok(batch, m)) return ADRT_HIP_ERR_SHAPE;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!x || !r || !p || !q || !alpha || !rr || !workspace) return ADRT_HIP_ERR_NULL;
    if (workspace_bytes < adrt_hip_cg_workspace(batch, m)) return ADRT_HIP_ERR_WORKSPACE;
    const Split sp = split(batch, m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *part = static_cast<double *>(workspace);
    const dim3 grid(unsigned(sp.chunks), unsigned(batch));
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(update_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<float *>(x), static_cast<float *>(r),
                           static_cast<const float *>(p), static_cast<const float *>(q), alpha,
                           part, m, sp.chunk, sp.chunks);
    else
        hipLaunchKernelGGL(update_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<double *>(x), static_cast<double *>(r),
                           static_cast<const double *>(p), static_cast<const double *>(q), alpha,
                           part, m, sp.chunk, sp.chunks);
    hipLaunchKernelGGL(rows_kernel, dim3(unsigned((batch + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, s, (const double *)part, rr, batch, sp.chunks);
    return launch_status();
}

int adrt_hip_cg_direction(void *p, const void *z, const double *beta, int dtype, size_t batch,
                          size_t m, void *stream) {
    if (!batch_ok(batch, m)) return ADRT_HIP_ERR_SHAPE;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!p || !z || !beta) return ADRT_HIP_ERR_NULL;
    const Split sp = split(batch, m);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(unsigned(sp.chunks), unsigned(batch));
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(direction_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<float *>(p), static_cast<const float *>(z), beta, m,
                           sp.chunk);
    else
        hipLaunchKernelGGL(direction_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<double *>(p), static_cast<const double *>(z), beta, m,
                           sp.chunk);
    return launch_status();
}

}  // extern "C"

}  // namespace adrt_hip
