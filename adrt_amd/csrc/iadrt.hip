// Exact inverse ADRT (iadrt) on gfx950.
//
// The stage is a serial reverse-row recurrence per (b, q, l, col) chain
// (adrt_cdefs_iadrt.hpp:52-105).  A parallel scan would reorder the
// additions, so every chain keeps the reference's order: val = 0; val += a;
// val -= b; val += out[row+1].  Chains are independent, one lane each.
#include "common.hpp"
#include "transpose.hpp"

namespace adrt_hip {

// in (B,4,Li,Ci,R) -> out (B,4,2Li,Ci/2,R); chain = (bq, l, col).
template <typename T>
__global__ void __launch_bounds__(kBlock)
    iadrt_stage_kernel(const T *__restrict__ in, T *__restrict__ out, size_t chains, size_t Li,
                       size_t Ci, size_t R) {
    const size_t Lo = 2 * Li, Co = Ci / 2;
    for (size_t ch = size_t(blockIdx.x) * blockDim.x + threadIdx.x; ch < chains;
         ch += size_t(gridDim.x) * blockDim.x) {
        const size_t col = ch % Co;
        const size_t l = (ch / Co) % Lo;
        const size_t bq = ch / (Co * Lo);
        const size_t pl = l >> 1;
        const T *src = in + ((bq * Li + pl) * Ci) * R;
        const T *s_even = src + (2 * col) * R;
        const T *s_odd = src + (2 * col + 1) * R;
        const bool has_odd = 2 * col + 1 < Ci;
        T *dst = out + ch * R;
        T acc = T(0);
        for (size_t rr = 0; rr < R; ++rr) {
            const size_t row = R - rr - 1;
            T val = T(0);
            if ((l & 1) == 0) {
                val += s_even[row];
                if (row + 1 < R && has_odd) val -= s_odd[row + 1];
            } else {
                if (row + 1 + col < R) {
                    if (has_odd) val += s_odd[row + 1 + col];
                    val -= s_even[row + 1 + col];
                }
            }
            if (row + 1 < R) val += acc;
            dst[row] = val;
            acc = val;
        }
    }
}

template <typename T>
static int run_streaming(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const size_t C = 2 * n - 1;
    const int L = num_iters(n);
    T *bufs[2] = {(L % 2 == 0) ? ws : out, (L % 2 == 0) ? out : ws};
    launch_transpose<T>(in, bufs[0], B * 4, C, n, s);
    const size_t chains = B * 4 * n;
    const unsigned grid = grid_for(chains);
    size_t Li = 1, Ci = n;
    for (int k = 0; k < L; ++k) {
        hipLaunchKernelGGL(iadrt_stage_kernel<T>, dim3(grid), dim3(kBlock), 0, s, bufs[k & 1],
                           bufs[(k + 1) & 1], chains, Li, Ci, C);
        Li *= 2;
        Ci /= 2;
    }
    launch_transpose<T>(bufs[L & 1], out, B * 4, n, C, s);
    return launch_status();
}

size_t iadrt_workspace(int dtype, size_t B, size_t n) {
    const size_t s = dtype == ADRT_HIP_F64 ? 8 : 4;
    return B * 4 * n * (2 * n - 1) * s;
}

int launch_iadrt(const void *in, void *out, void *ws, size_t, int dtype, size_t B, size_t n,
                 hipStream_t s) {
    if (dtype == ADRT_HIP_F32)
        return run_streaming(static_cast<const float *>(in), static_cast<float *>(out),
                             static_cast<float *>(ws), B, n, s);
    return run_streaming(static_cast<const double *>(in), static_cast<double *>(out),
                         static_cast<double *>(ws), B, n, s);
}

}  // namespace adrt_hip
