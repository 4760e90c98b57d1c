// Forward ADRT on gfx950.
//
// Streaming path (any n): quadrant load + zero pad, log2(n) butterfly stages
// over the (B, 4, R, S, 2n-1) buffer with offsets fastest, then a tiled
// transposed copy-out.  Each stage element is formed by the same single IEEE
// add of the same two operands as adrt_core (adrt_cdefs_adrt.hpp:55-96), so
// the result is bit-identical to the reference.
#include "common.hpp"
#include "transpose.hpp"

namespace adrt_hip {

// buf[bq][r][c] for c < 2n-1 from the image; quadrant formulas follow the
// loads of adrt_cdefs_adrt.hpp:123-186:
//   q0: img[r][n-1-c]  q1: img[n-1-c][r]  q2: img[c][r]  q3: img[n-1-r][n-1-c]
// and zero for c >= n.  One block per (b, q, r) line.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    fwd_load_kernel(const T *__restrict__ img, T *__restrict__ buf, size_t lines, size_t n) {
    const size_t C = 2 * n - 1;
    for (size_t line = blockIdx.x; line < lines; line += gridDim.x) {
        const size_t r = line % n;
        const size_t bq = line / n;
        const size_t q = bq & 3, b = bq >> 2;
        const T *im = img + b * n * n;
        T *dst = buf + line * C;
        for (size_t c = threadIdx.x; c < C; c += blockDim.x) {
            T v = T(0);
            if (c < n) {
                switch (q) {
                    case 0: v = im[r * n + (n - 1 - c)]; break;
                    case 1: v = im[(n - 1 - c) * n + r]; break;
                    case 2: v = im[c * n + r]; break;
                    default: v = im[(n - 1 - r) * n + (n - 1 - c)]; break;
                }
            }
            dst[c] = v;
        }
    }
}

// One butterfly stage.  Output line = (bq*R + r)*S + a, S = 2^log2s.
// out[c] = x[c] (c < ceil(a/2)) else x[c] + y[c - ceil(a/2)], where x is
// input row 2r, y input row 2r+1, both at angle a/2.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    fwd_stage_kernel(const T *__restrict__ in, T *__restrict__ out, size_t lines, int log2s,
                     size_t C) {
    const size_t S = size_t(1) << log2s;
    const size_t S_in = S >> 1;
    for (size_t line = blockIdx.x; line < lines; line += gridDim.x) {
        const size_t a = line & (S - 1);
        const size_t bqr = line >> log2s;
        const T *x = in + ((2 * bqr) * S_in + (a >> 1)) * C;
        const T *y = in + ((2 * bqr + 1) * S_in + (a >> 1)) * C;
        const size_t sh = (a + 1) >> 1;
        T *o = out + line * C;
        for (size_t c = threadIdx.x; c < C; c += blockDim.x) {
            T v = x[c];
            if (c >= sh) v = v + y[c - sh];
            o[c] = v;
        }
    }
}

// adrt_step on the output layout (adrt_cdefs_adrt.hpp:214-258); one thread
// per output element of a (bq, row) line of n columns.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    fwd_step_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t n,
                    int log2n, int step) {
    const size_t C = 2 * n - 1;
    const size_t e = size_t(1) << step, e2 = e << 1;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t col = idx & (n - 1);
        const size_t line = idx >> log2n;
        const size_t row = line % C;
        const T *plane = in + (line - row) * n;
        const size_t cb = col / e2, ci = col % e2;
        T v = plane[row * n + 2 * cb * e + (ci >> 1)];
        const size_t sh = (ci + 1) >> 1;
        if (sh <= row) v = v + plane[(row - sh) * n + (2 * cb + 1) * e + (ci >> 1)];
        out[idx] = v;
    }
}

template <typename T>
static int run_streaming(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const size_t C = 2 * n - 1;
    const int L = num_iters(n);
    const size_t lines = B * 4 * n;
    // Ping-pong between out and ws so that the last stage lands in ws.
    T *bufs[2] = {(L % 2 == 0) ? ws : out, (L % 2 == 0) ? out : ws};
    const unsigned grid = unsigned(lines < (1u << 20) ? lines : (1u << 20));
    hipLaunchKernelGGL(fwd_load_kernel<T>, dim3(grid), dim3(kBlock), 0, s, in, bufs[0], lines, n);
    for (int i = 0; i < L; ++i) {
        hipLaunchKernelGGL(fwd_stage_kernel<T>, dim3(grid), dim3(kBlock), 0, s, bufs[i & 1],
                           bufs[(i + 1) & 1], lines, i + 1, C);
    }
    // bufs[L & 1] == ws holds (B*4, n, C) with angles as rows; transpose to (B*4, C, n).
    launch_transpose<T>(bufs[L & 1], out, B * 4, n, C, s);
    return launch_status();
}

size_t adrt_workspace(int dtype, size_t B, size_t n) {
    if (use_small(ADRT_HIP_OP_ADRT, B, n)) return 0;
    if (use_fused(ADRT_HIP_OP_ADRT, n)) return fused_workspace(dtype, B, n);
    const size_t s = dtype == ADRT_HIP_F64 ? 8 : 4;
    return B * 4 * n * (2 * n - 1) * s;
}

int launch_adrt(const void *in, void *out, void *ws, size_t, int dtype, size_t B, size_t n,
                hipStream_t s) {
    if (use_small(ADRT_HIP_OP_ADRT, B, n))
        return launch_small(in, out, ADRT_HIP_OP_ADRT, dtype, B, n, s);
    if (use_fused(ADRT_HIP_OP_ADRT, n)) return run_fused(in, out, ws, dtype, B, n, s);
    if (dtype == ADRT_HIP_F32)
        return run_streaming(static_cast<const float *>(in), static_cast<float *>(out),
                             static_cast<float *>(ws), B, n, s);
    return run_streaming(static_cast<const double *>(in), static_cast<double *>(out),
                         static_cast<double *>(ws), B, n, s);
}

int launch_adrt_step(const void *in, void *out, int dtype, size_t B, size_t n, int step,
                     hipStream_t s) {
    const size_t total = B * 4 * (2 * n - 1) * n;
    const int log2n = num_iters(n);
    const unsigned grid = grid_for(total);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(fwd_step_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, n,
                           log2n, step);
    else
        hipLaunchKernelGGL(fwd_step_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, n,
                           log2n, step);
    return launch_status();
}

}  // namespace adrt_hip
