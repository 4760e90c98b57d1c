// Full-multigrid operators (Press FMG) on gfx950; adrt_cdefs_fmg.hpp:53-171.
// Built with -ffp-contract=off: every product is rounded before the sums, so
// the high-pass result is bit-identical to the reference.
#include "common.hpp"

namespace adrt_hip {

// (B,4,2n-1,n) -> (B,4,n-1,n/2): (x[2r][2c] + x[2r+1][2c]) / 4.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    restriction_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t n) {
    const size_t Ri = 2 * n - 1, Ro = n - 1, Co = n / 2;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t c = idx % Co;
        const size_t line = idx / Co;
        const size_t r = line % Ro, bq = line / Ro;
        const T va = in[(bq * Ri + 2 * r) * n + 2 * c];
        const T vb = in[(bq * Ri + 2 * r + 1) * n + 2 * c];
        out[idx] = (va + vb) / T(4);
    }
}

// (B,h,w) -> (B,2h,2w): 2x2 replication.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    prolongation_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t h,
                        size_t w) {
    const size_t W2 = 2 * w;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t oc = idx % W2;
        const size_t line = idx / W2;
        const size_t orow = line % (2 * h), b = line / (2 * h);
        out[idx] = in[(b * h + (orow >> 1)) * w + (oc >> 1)];
    }
}

// 3x3 kernel [[a,b,a],[b,c,b],[a,b,a]], a=-1/16, b=-1/8, c=3/4, reflected
// borders; sums grouped column by column as in the reference.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    highpass_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t h,
                    size_t w) {
    const T ka = T(-0.0625), kb = T(-0.125), kc = T(0.75);
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t c = idx % w;
        const size_t line = idx / w;
        const size_t r = line % h, b = line / h;
        const T *img = in + b * h * w;
        const size_t pr = (r == 0) ? 1 : r - 1;
        const size_t nr = (r == h - 1) ? r - 1 : r + 1;
        const T *P = img + pr * w, *M = img + r * w, *Q = img + nr * w;
        T res;
        if (c == 0) {
            const T v11 = ka * P[1], v12 = kb * P[0];
            const T v21 = kb * M[1], v22 = kc * M[0];
            const T v31 = ka * Q[1], v32 = kb * Q[0];
            res = (v11 + v21 + v31) + (v12 + v22 + v32) + (v11 + v21 + v31);
        } else if (c == w - 1) {
            const T v11 = ka * P[c - 1], v12 = kb * P[c];
            const T v21 = kb * M[c - 1], v22 = kc * M[c];
            const T v31 = ka * Q[c - 1], v32 = kb * Q[c];
            res = (v11 + v21 + v31) + (v12 + v22 + v32) + (v11 + v21 + v31);
        } else {
            const T v11 = ka * P[c - 1], v12 = kb * P[c], v13 = ka * P[c + 1];
            const T v21 = kb * M[c - 1], v22 = kc * M[c], v23 = kb * M[c + 1];
            const T v31 = ka * Q[c - 1], v32 = kb * Q[c], v33 = ka * Q[c + 1];
            res = (v11 + v21 + v31) + (v12 + v22 + v32) + (v13 + v23 + v33);
        }
        out[idx] = res;
    }
}

int launch_fmg_restriction(const void *in, void *out, int dtype, size_t B, size_t n,
                           hipStream_t s) {
    const size_t total = B * 4 * (n - 1) * (n / 2);
    if (total == 0) return ADRT_HIP_OK;
    const unsigned g = grid_for(total);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(restriction_kernel<float>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, n);
    else
        hipLaunchKernelGGL(restriction_kernel<double>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, n);
    return launch_status();
}

int launch_fmg_prolongation(const void *in, void *out, int dtype, size_t B, size_t h, size_t w,
                            hipStream_t s) {
    const size_t total = B * 4 * h * w;
    const unsigned g = grid_for(total);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(prolongation_kernel<float>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, h, w);
    else
        hipLaunchKernelGGL(prolongation_kernel<double>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, h,
                           w);
    return launch_status();
}

int launch_fmg_highpass(const void *in, void *out, int dtype, size_t B, size_t h, size_t w,
                        hipStream_t s) {
    const size_t total = B * h * w;
    const unsigned g = grid_for(total);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(highpass_kernel<float>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, h, w);
    else
        hipLaunchKernelGGL(highpass_kernel<double>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, h,
                           w);
    return launch_status();
}

}  // namespace adrt_hip
