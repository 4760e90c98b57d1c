// Back-projection (bdrt) on gfx950.
//
// Streaming path (any n): transposed load into the (B, 4, R, S, 2n-1) buffer
// with rows fastest, log2(n) stages, transposed store.  Each stage output is
// the same single IEEE add of the same operands as bdrt_core
// (adrt_cdefs_bdrt.hpp:55-116); its literal copy (last row of a right
// section) and zero tail are reproduced exactly.
#include "common.hpp"
#include "transpose.hpp"

namespace adrt_hip {

// Output line = (bq*R + si)*S + sp with R = 2^(L-k-1) output angle pairs and
// S = 2^log2s sections; left sections (sp even) take x[row] + y[row], right
// sections (sp odd) take x[row+si] + y[row+si+1] shifted down by si, a copy
// of x[C-1] at row C-1-si and zeros below.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bdrt_stage_kernel(const T *__restrict__ in, T *__restrict__ out, size_t lines, int log2s,
                      size_t R, size_t C) {
    const size_t S = size_t(1) << log2s;
    const size_t S_in = S >> 1;
    for (size_t line = blockIdx.x; line < lines; line += gridDim.x) {
        const size_t sp = line & (S - 1);
        const size_t bqsi = line >> log2s;
        const size_t si = bqsi & (R - 1);
        const size_t sec = sp >> 1;
        const T *x = in + ((2 * bqsi) * S_in + sec) * C;
        const T *y = in + ((2 * bqsi + 1) * S_in + sec) * C;
        T *o = out + line * C;
        if ((sp & 1) == 0) {
            for (size_t r = threadIdx.x; r < C; r += blockDim.x) o[r] = x[r] + y[r];
        } else {
            for (size_t r = threadIdx.x; r < C; r += blockDim.x) {
                const size_t row = r + si;
                T v = T(0);
                if (row < C - 1)
                    v = x[row] + y[row + 1];
                else if (row == C - 1)
                    v = x[row];
                o[r] = v;
            }
        }
    }
}

// bdrt_step on the output layout (adrt_cdefs_bdrt.hpp:189-244).  Missing
// operands are literal zeros that are added, as in the reference.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    bdrt_step_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t n,
                     int log2n, int step) {
    const size_t C = 2 * n - 1;
    const size_t e = size_t(1) << (log2n - step - 1);
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t col = idx & (n - 1);
        const size_t line = idx >> log2n;
        const size_t row = line % C;
        const T *plane = in + (line - row) * n;
        const size_t cb = col / e, ci = col % e;
        const size_t bc = 2 * (ci + e * (cb >> 1));
        T v;
        if ((cb & 1) == 0) {
            v = plane[row * n + bc] + plane[row * n + bc + 1];
        } else {
            const size_t br = row + ci;
            T xa = T(0), yb = T(0);
            if (br < C) xa = plane[br * n + bc];
            if (br + 1 < C) yb = plane[(br + 1) * n + bc + 1];
            v = xa + yb;
        }
        out[idx] = v;
    }
}

template <typename T>
static int run_streaming(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const size_t C = 2 * n - 1;
    const int L = num_iters(n);
    const size_t lines = B * 4 * n;
    T *bufs[2] = {(L % 2 == 0) ? ws : out, (L % 2 == 0) ? out : ws};
    // (B*4, C, n) -> (B*4, n, C): angle-major, rows fastest.
    launch_transpose<T>(in, bufs[0], B * 4, C, n, s);
    const unsigned grid = unsigned(lines < (1u << 20) ? lines : (1u << 20));
    for (int k = 0; k < L; ++k) {
        const size_t R = size_t(1) << (L - k - 1);
        hipLaunchKernelGGL(bdrt_stage_kernel<T>, dim3(grid), dim3(kBlock), 0, s, bufs[k & 1],
                           bufs[(k + 1) & 1], lines, k + 1, R, C);
    }
    launch_transpose<T>(bufs[L & 1], out, B * 4, n, C, s);
    return launch_status();
}

size_t bdrt_workspace(int dtype, size_t B, size_t n) {
    if (use_small(ADRT_HIP_OP_BDRT, B, n)) return 0;
    if (use_fused(ADRT_HIP_OP_BDRT, n)) return fused_bdrt_workspace(dtype, B, n);
    const size_t s = dtype == ADRT_HIP_F64 ? 8 : 4;
    return B * 4 * n * (2 * n - 1) * s;
}

// log2 n in [6, 12]: the fused two-pass back-projection (fused.hip);
// other sizes stream one stage per kernel.
int launch_bdrt(const void *in, void *out, void *ws, size_t, int dtype, size_t B, size_t n,
                int rows, hipStream_t s) {
    if (use_small(ADRT_HIP_OP_BDRT, B, n))
        return launch_small(in, out, ADRT_HIP_OP_BDRT, dtype, B, n, s);
    if (use_fused(ADRT_HIP_OP_BDRT, n)) return run_fused_bdrt(in, out, ws, dtype, B, n, rows, s);
    if (dtype == ADRT_HIP_F32)
        return run_streaming(static_cast<const float *>(in), static_cast<float *>(out),
                             static_cast<float *>(ws), B, n, s);
    return run_streaming(static_cast<const double *>(in), static_cast<double *>(out),
                         static_cast<double *>(ws), B, n, s);
}

int launch_bdrt_step(const void *in, void *out, int dtype, size_t B, size_t n, int step,
                     hipStream_t s) {
    const size_t total = B * 4 * (2 * n - 1) * n;
    const int log2n = num_iters(n);
    const unsigned grid = grid_for(total);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(bdrt_step_kernel<float>, dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, n,
                           log2n, step);
    else
        hipLaunchKernelGGL(bdrt_step_kernel<double>, dim3(grid), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, n,
                           log2n, step);
    return launch_status();
}

}  // namespace adrt_hip
