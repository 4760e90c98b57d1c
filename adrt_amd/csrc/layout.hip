// Layout kernels: adrt_init (core.py:123-176), truncate (utils.py:191-242)
// and the fused truncate + quadrant combine used by the solvers.
#include <cstdint>

#include "common.hpp"

namespace adrt_hip {

// out (B,4,2n-1,n): rows i < n hold the quadrant flips, rows >= n are zero.
//   q0[i][j] = a[j][n-1-i]   q1[i][j] = a[n-1-i][j]
//   q2[i][j] = a[i][j]       q3[i][j] = a[n-1-j][n-1-i]
template <typename E>
__global__ void __launch_bounds__(kBlock)
    init_kernel(const E *__restrict__ in, E *__restrict__ out, size_t total, size_t n, int log2n) {
    const size_t C = 2 * n - 1;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t j = idx & (n - 1);
        const size_t line = idx >> log2n;
        const size_t i = line % C;
        const size_t bq = line / C;
        const size_t q = bq & 3, b = bq >> 2;
        E v = E(0);
        if (i < n) {
            const E *a = in + b * n * n;
            switch (q) {
                case 0: v = a[j * n + (n - 1 - i)]; break;
                case 1: v = a[(n - 1 - i) * n + j]; break;
                case 2: v = a[i * n + j]; break;
                default: v = a[(n - 1 - j) * n + (n - 1 - i)]; break;
            }
        }
        out[idx] = v;
    }
}

// Inverse of init on the top n x n square of each quadrant:
//   t0[i][j] = x0[n-1-j][i]  t1[i][j] = x1[n-1-i][j]
//   t2[i][j] = x2[i][j]      t3[i][j] = x3[n-1-j][n-1-i]
template <typename E>
__device__ inline E truncate_at(const E *plane4, size_t q, size_t i, size_t j, size_t n) {
    const E *x = plane4 + q * (2 * n - 1) * n;
    switch (q) {
        case 0: return x[(n - 1 - j) * n + i];
        case 1: return x[(n - 1 - i) * n + j];
        case 2: return x[i * n + j];
        default: return x[(n - 1 - j) * n + (n - 1 - i)];
    }
}

template <typename E>
__global__ void __launch_bounds__(kBlock)
    truncate_kernel(const E *__restrict__ in, E *__restrict__ out, size_t total, size_t n) {
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t j = idx % n;
        const size_t i = (idx / n) % n;
        const size_t oq = idx / (n * n);
        const size_t q = oq & 3, o = oq >> 2;
        out[idx] = truncate_at(in + o * 4 * (2 * n - 1) * n, q, i, j, n);
    }
}

// out[b][i][j] = ((t0/d + t1/d) + t2/d) + t3/d, then / 4 for mode 0.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    truncate_combine_kernel(const T *__restrict__ in, T *__restrict__ out, size_t total, size_t n,
                            T divisor, int divide, int mode) {
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t j = idx % n;
        const size_t i = (idx / n) % n;
        const size_t b = idx / (n * n);
        const T *p = in + b * 4 * (2 * n - 1) * n;
        T t0 = truncate_at(p, 0, i, j, n), t1 = truncate_at(p, 1, i, j, n);
        T t2 = truncate_at(p, 2, i, j, n), t3 = truncate_at(p, 3, i, j, n);
        if (divide) {
            t0 = t0 / divisor;
            t1 = t1 / divisor;
            t2 = t2 / divisor;
            t3 = t3 / divisor;
        }
        T acc = ((t0 + t1) + t2) + t3;
        if (mode == 0) acc = acc / T(4);
        out[idx] = acc;
    }
}

// Tiled form for n >= 32 (n a power of two): one 32 x 32 output tile of one
// image per block.  Quadrants 0 and 3 are transposed reads; their source
// blocks are staged through LDS with row-wise (coalesced) loads.  The
// arithmetic per element is the kernel's above, in the same order.
constexpr int kTile = 32;
template <typename T>
__global__ void __launch_bounds__(kBlock)
    truncate_combine_tiled_kernel(const T *__restrict__ in, T *__restrict__ out, int n,
                                  int tiles_per_row, T divisor, int divide, int mode) {
    __shared__ T s0[kTile][kTile + 1];
    __shared__ T s3[kTile][kTile + 1];
    const int tile = blockIdx.x;
    const int i0 = (tile / tiles_per_row) * kTile, j0 = (tile % tiles_per_row) * kTile;
    const size_t b = blockIdx.y;
    const size_t plane = size_t(2 * n - 1) * n;
    const T *x0 = in + b * 4 * plane;
    const T *x1 = x0 + plane, *x2 = x1 + plane, *x3 = x2 + plane;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;  // 32 x 8
    // s0[jj][ii] = x0[n-1-(j0+jj)][i0+ii];  s3[jj][ii] = x3[n-1-(j0+jj)][n-1-(i0+ii)]
    for (int jj = ty; jj < kTile; jj += kBlock / kTile) {
        const size_t row = size_t(n - 1 - (j0 + jj)) * n;
        s0[jj][tx] = x0[row + i0 + tx];
        s3[jj][kTile - 1 - tx] = x3[row + (n - kTile - i0) + tx];
    }
    __syncthreads();
    for (int ii = ty; ii < kTile; ii += kBlock / kTile) {
        const int i = i0 + ii, j = j0 + tx;
        T t0 = s0[tx][ii];
        T t1 = x1[size_t(n - 1 - i) * n + j];
        T t2 = x2[size_t(i) * n + j];
        T t3 = s3[tx][ii];
        if (divide) {
            t0 = t0 / divisor;
            t1 = t1 / divisor;
            t2 = t2 / divisor;
            t3 = t3 / divisor;
        }
        T acc = ((t0 + t1) + t2) + t3;
        if (mode == 0) acc = acc / T(4);
        out[(b * n + i) * n + j] = acc;
    }
}

// stitch_adrt (utils.py:65-130): out (outer, 3n-2, 4 * vc), vc = n or n - 1.
// Quadrant i occupies column block i; quadrants 0, 1 fill rows [0, 2n-1),
// quadrants 2, 3 rows [n-1, 3n-2); odd quadrants are flipped in both axes;
// everything else is zero.
template <typename E>
__global__ void __launch_bounds__(kBlock)
    stitch_kernel(const E *__restrict__ in, E *__restrict__ out, size_t total, size_t n,
                  size_t vc) {
    const size_t R = 2 * n - 1, rows = 3 * n - 2, cols = 4 * vc;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t col = idx % cols;
        const size_t r = (idx / cols) % rows;
        const size_t o = idx / (cols * rows);
        const size_t i = col / vc, c = col % vc;
        const size_t lo = i < 2 ? 0 : n - 1;
        E v = E(0);
        if (r >= lo && r < lo + R) {
            size_t rr = r - lo, cc = c;
            if (i % 2) {
                rr = R - 1 - rr;
                cc = n - 1 - cc;
            }
            v = in[((o * 4 + i) * R + rr) * n + cc];
        }
        out[idx] = v;
    }
}

// unstitch_adrt (utils.py:133-188): the inverse, out (outer, 4, 2n-1, n)
// from (outer, 3n-2, 4 * vc).  With vc = n - 1 the last column of quadrant
// q (before its flip) is column 0 of quadrant q+1 (mod 4) on the same
// stitched row, row-reversed for q = 3.
template <typename E>
__global__ void __launch_bounds__(kBlock)
    unstitch_kernel(const E *__restrict__ in, E *__restrict__ out, size_t total, size_t n,
                    size_t vc) {
    const size_t R = 2 * n - 1, rows = 3 * n - 2, cols = 4 * vc;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const size_t c = idx % n;
        const size_t r = (idx / n) % R;
        const size_t q = (idx / (n * R)) % 4;
        const size_t o = idx / (n * R * 4);
        size_t rr = r, cc = c;
        if (q % 2) {
            rr = R - 1 - rr;
            cc = n - 1 - cc;
        }
        const size_t srow = rr + (q < 2 ? 0 : n - 1);
        const E *img = in + o * rows * cols;
        E v;
        if (cc < vc) {
            v = img[srow * cols + q * vc + cc];
        } else if (q == 3) {
            v = img[(rows - 1 - srow) * cols];
        } else {
            v = img[srow * cols + (q + 1) * vc];
        }
        out[idx] = v;
    }
}

template <typename E>
static void init_typed(const void *in, void *out, size_t B, size_t n, hipStream_t s) {
    const size_t total = B * 4 * (2 * n - 1) * n;
    hipLaunchKernelGGL(init_kernel<E>, dim3(grid_for(total)), dim3(kBlock), 0, s,
                       static_cast<const E *>(in), static_cast<E *>(out), total, n, num_iters(n));
}

template <typename E>
static void truncate_typed(const void *in, void *out, size_t outer, size_t n, hipStream_t s) {
    const size_t total = outer * 4 * n * n;
    hipLaunchKernelGGL(truncate_kernel<E>, dim3(grid_for(total)), dim3(kBlock), 0, s,
                       static_cast<const E *>(in), static_cast<E *>(out), total, n);
}

int launch_adrt_init(const void *in, void *out, size_t elem, size_t B, size_t n, hipStream_t s) {
    switch (elem) {
        case 1: init_typed<uint8_t>(in, out, B, n, s); break;
        case 2: init_typed<uint16_t>(in, out, B, n, s); break;
        case 4: init_typed<uint32_t>(in, out, B, n, s); break;
        default: init_typed<uint64_t>(in, out, B, n, s); break;
    }
    return launch_status();
}

int launch_truncate(const void *in, void *out, size_t elem, size_t outer, size_t n,
                    hipStream_t s) {
    switch (elem) {
        case 1: truncate_typed<uint8_t>(in, out, outer, n, s); break;
        case 2: truncate_typed<uint16_t>(in, out, outer, n, s); break;
        case 4: truncate_typed<uint32_t>(in, out, outer, n, s); break;
        default: truncate_typed<uint64_t>(in, out, outer, n, s); break;
    }
    return launch_status();
}

int launch_truncate_combine(const void *in, void *out, int dtype, size_t B, size_t n,
                            double divisor, int mode, hipStream_t s) {
    const size_t total = B * n * n;
    const unsigned g = grid_for(total);
    const int divide = divisor != 1.0;
    if (n >= size_t(kTile) && B <= 65535) {
        const int tpr = int(n / kTile);
        const dim3 grid(unsigned(tpr * tpr), unsigned(B));
        if (dtype == ADRT_HIP_F32)
            hipLaunchKernelGGL(truncate_combine_tiled_kernel<float>, grid, dim3(kBlock), 0, s,
                               static_cast<const float *>(in), static_cast<float *>(out), int(n),
                               tpr, float(divisor), divide, mode);
        else
            hipLaunchKernelGGL(truncate_combine_tiled_kernel<double>, grid, dim3(kBlock), 0, s,
                               static_cast<const double *>(in), static_cast<double *>(out), int(n),
                               tpr, divisor, divide, mode);
        return launch_status();
    }
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(truncate_combine_kernel<float>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const float *>(in), static_cast<float *>(out), total, n,
                           float(divisor), divide, mode);
    else
        hipLaunchKernelGGL(truncate_combine_kernel<double>, dim3(g), dim3(kBlock), 0, s,
                           static_cast<const double *>(in), static_cast<double *>(out), total, n,
                           divisor, divide, mode);
    return launch_status();
}

template <typename E>
static void stitch_typed(const void *in, void *out, size_t outer, size_t n, size_t vc,
                         bool inverse, hipStream_t s) {
    if (!inverse) {
        const size_t total = outer * (3 * n - 2) * 4 * vc;
        hipLaunchKernelGGL(stitch_kernel<E>, dim3(grid_for(total)), dim3(kBlock), 0, s,
                           static_cast<const E *>(in), static_cast<E *>(out), total, n, vc);
    } else {
        const size_t total = outer * 4 * (2 * n - 1) * n;
        hipLaunchKernelGGL(unstitch_kernel<E>, dim3(grid_for(total)), dim3(kBlock), 0, s,
                           static_cast<const E *>(in), static_cast<E *>(out), total, n, vc);
    }
}

int launch_stitch(const void *in, void *out, size_t elem, size_t outer, size_t n, size_t vc,
                  bool inverse, hipStream_t s) {
    switch (elem) {
        case 1: stitch_typed<uint8_t>(in, out, outer, n, vc, inverse, s); break;
        case 2: stitch_typed<uint16_t>(in, out, outer, n, vc, inverse, s); break;
        case 4: stitch_typed<uint32_t>(in, out, outer, n, vc, inverse, s); break;
        default: stitch_typed<uint64_t>(in, out, outer, n, vc, inverse, s); break;
    }
    return launch_status();
}

}  // namespace adrt_hip
