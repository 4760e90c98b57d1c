// Batched tiled transpose through LDS: in (planes, rows, cols) -> out (planes, cols, rows).
#pragma once

#include "common.hpp"

namespace adrt_hip {

constexpr int kTile = 32;
constexpr int kTileRows = 8;

// grid: x = ceil(cols/32), y = ceil(rows/32), z = plane chunk; block 32x8.
template <typename T>
__global__ void __launch_bounds__(kTile *kTileRows)
    transpose_kernel(const T *__restrict__ in, T *__restrict__ out, size_t planes, size_t rows,
                     size_t cols) {
    __shared__ T tile[kTile][kTile + 1];
    const size_t c0 = size_t(blockIdx.x) * kTile, r0 = size_t(blockIdx.y) * kTile;
    const size_t plane_elems = rows * cols;
    for (size_t p = blockIdx.z; p < planes; p += gridDim.z) {
        const T *src = in + p * plane_elems;
        T *dst = out + p * plane_elems;
        for (int i = threadIdx.y; i < kTile; i += kTileRows) {
            const size_t r = r0 + i, c = c0 + threadIdx.x;
            if (r < rows && c < cols) tile[i][threadIdx.x] = src[r * cols + c];
        }
        __syncthreads();
        for (int i = threadIdx.y; i < kTile; i += kTileRows) {
            const size_t c = c0 + i, r = r0 + threadIdx.x;
            if (r < rows && c < cols) dst[c * rows + r] = tile[threadIdx.x][i];
        }
        __syncthreads();
    }
}

template <typename T>
inline void launch_transpose(const T *in, T *out, size_t planes, size_t rows, size_t cols,
                             hipStream_t s) {
    const dim3 block(kTile, kTileRows);
    size_t gz = planes < 65535 ? planes : 65535;
    const dim3 grid(unsigned((cols + kTile - 1) / kTile), unsigned((rows + kTile - 1) / kTile),
                    unsigned(gz));
    hipLaunchKernelGGL(transpose_kernel<T>, grid, block, 0, s, in, out, planes, rows, cols);
}

}  // namespace adrt_hip
