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
// Sums are deterministic and independent of the batch size: each block
// reduces a fixed 16 K-element chunk of one image in a fixed order and a
// second pass adds the chunks of an image in order.
//
// The pcg_* entry points run a whole (preconditioned) CG iteration's scalar
// logic on the device, over a per-image state buffer, so that one iteration
// is the operator, the preconditioner and four small launch pairs with no
// host round trip:
//   pcg_update:    pq = p.q; alpha = rz / pq (0 for inactive images or
//                  pq <= 0); x += alpha p; r -= alpha q; rr = r.r;
//                  iters += active; next = active && sqrt(rr) > tol
//   pcg_direction: rz' = r.z (preconditioned images) or rr; an image whose
//                  preconditioned step fails (rz' <= 0 or pq <= 0) drops the
//                  preconditioner and restarts from p = r; otherwise
//                  beta = rz' / rz, p = z + beta p; active = next;
//                  count = number of active images
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

// Order of a chunk's sum: element i of [lo, hi) goes to thread
// ((i - lo) / 4) mod kThreads and to that thread's accumulator (i - lo) mod 4;
// the thread returns (s0 + s1) + (s2 + s3), and block_sum adds the threads.
// The order depends on the chunk alone, so an image's sum does not depend on
// the batch, and dot_kernel and update_kernel (r.r) sum alike.  A thread's
// elements are independent, so its loads are all in flight together.
constexpr int kSub = 4;

// Four consecutive elements as one vector (16 bytes of float, 32 of double):
// used when m is a multiple of 4, so that every chunk starts and ends on a
// whole vector and a lane's four elements are one coalesced access.
template <typename T>
using vec4 = T __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ bool vec_ok(const T *p) {
    return (reinterpret_cast<unsigned long long>(p) & (sizeof(vec4<T>) - 1)) == 0;
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
    double sj[kSub] = {0.0, 0.0, 0.0, 0.0};
    if (m % kSub == 0 && vec_ok(ub) && vec_ok(vb)) {
#pragma unroll 4
        for (size_t base = lo + kSub * threadIdx.x; base < hi; base += kSub * kThreads) {
            const vec4<T> a = *reinterpret_cast<const vec4<T> *>(ub + base);
            const vec4<T> c = *reinterpret_cast<const vec4<T> *>(vb + base);
#pragma unroll
            for (int j = 0; j < kSub; ++j) sj[j] += double(a[j]) * double(c[j]);
        }
    } else {
#pragma unroll 4
        for (size_t base = lo + kSub * threadIdx.x; base < hi; base += kSub * kThreads) {
#pragma unroll
            for (int j = 0; j < kSub; ++j)
                if (base + j < hi) sj[j] += double(ub[base + j]) * double(vb[base + j]);
        }
    }
    double s = (sj[0] + sj[1]) + (sj[2] + sj[3]);
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
    double sj[kSub] = {0.0, 0.0, 0.0, 0.0};
    if (m % kSub == 0 && vec_ok(pb) && vec_ok(qb) && vec_ok(xb) && vec_ok(rb)) {
#pragma unroll 2
        for (size_t base = lo + kSub * threadIdx.x; base < hi; base += kSub * kThreads) {
            const vec4<T> p4 = *reinterpret_cast<const vec4<T> *>(pb + base);
            const vec4<T> q4 = *reinterpret_cast<const vec4<T> *>(qb + base);
            vec4<T> x4 = *reinterpret_cast<const vec4<T> *>(xb + base);
            vec4<T> r4 = *reinterpret_cast<const vec4<T> *>(rb + base);
#pragma unroll
            for (int j = 0; j < kSub; ++j) {
                const T ap = a * p4[j];
                const T aq = a * q4[j];
                x4[j] = x4[j] + ap;
                r4[j] = r4[j] - aq;
                sj[j] += double(r4[j]) * double(r4[j]);
            }
            *reinterpret_cast<vec4<T> *>(xb + base) = x4;
            *reinterpret_cast<vec4<T> *>(rb + base) = r4;
        }
    } else {
#pragma unroll 4
        for (size_t base = lo + kSub * threadIdx.x; base < hi; base += kSub * kThreads) {
#pragma unroll
            for (int j = 0; j < kSub; ++j) {
                const size_t i = base + j;
                if (i < hi) {
                    const T ap = a * pb[i];
                    const T aq = a * qb[i];
                    xb[i] = xb[i] + ap;
                    const T rn = rb[i] - aq;
                    rb[i] = rn;
                    sj[j] += double(rn) * double(rn);
                }
            }
        }
    }
    double s = (sj[0] + sj[1]) + (sj[2] + sj[3]);
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
#pragma unroll 4
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

// Chunks of 16 K elements whatever the batch size, so an image's sums (and
// everything computed from them) do not depend on the other images in the
// batch: a solver that drops converged images, or a caller that splits a
// batch, gets the same values bit for bit.
constexpr size_t kChunk = 16384;

inline Split split(size_t B, size_t m) {
    (void)B;
    Split s;
    s.chunk = kChunk;
    s.chunks = int((m + kChunk - 1) / kChunk);
    return s;
}

// Per-image CG state, pcg_state_bytes(B) bytes:
//   double rz[B], pq[B], rr[B], tol[B], alpha[B], beta[B]
//   int32  active[B], next[B], use_p[B], iters[B], count
struct PcgState {
    double *rz, *pq, *rr, *tol, *alpha, *beta;
    int *active, *next, *use_p, *iters, *count;
};

inline size_t pcg_bytes(size_t B) { return 64 * B + 8; }

inline PcgState pcg_state(void *base, size_t B) {
    double *d = static_cast<double *>(base);
    int *i = reinterpret_cast<int *>(d + 6 * B);
    PcgState s;
    s.rz = d;
    s.pq = d + B;
    s.rr = d + 2 * B;
    s.tol = d + 3 * B;
    s.alpha = d + 4 * B;
    s.beta = d + 5 * B;
    s.active = i;
    s.next = i + B;
    s.use_p = i + 2 * B;
    s.iters = i + 3 * B;
    s.count = i + 4 * B;
    return s;
}

__device__ __forceinline__ double row_sum(const double *partial, size_t b, int chunks) {
    double v = 0.0;
    for (int c = 0; c < chunks; ++c) v += partial[b * chunks + c];
    return v;
}

// pq and the step length of each image
__global__ void __launch_bounds__(kThreads)
    pq_rows_kernel(const double *__restrict__ partial, PcgState s, size_t B, int chunks) {
    const size_t b = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= B) return;
    const double pq = row_sum(partial, b, chunks);
    s.pq[b] = pq;
    s.alpha[b] = (s.active[b] && pq > 0.0) ? s.rz[b] / pq : 0.0;
}

// ||r||^2, the iteration count and the convergence test; clears the count
// that pcg_direction accumulates
__global__ void __launch_bounds__(kThreads)
    rr_rows_kernel(const double *__restrict__ partial, PcgState s, size_t B, int chunks) {
    if (blockIdx.x == 0 && threadIdx.x == 0) s.count[0] = 0;
    const size_t b = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= B) return;
    const double rr = row_sum(partial, b, chunks);
    s.rr[b] = rr;
    const int act = s.active[b];
    s.iters[b] += act;
    s.next[b] = (act && sqrt(rr) > s.tol[b]) ? 1 : 0;
}

// beta, the preconditioner's breakdown test and the new active set
// (partial == nullptr: no preconditioner, rz' = rr)
__global__ void __launch_bounds__(kThreads)
    rz_rows_kernel(const double *__restrict__ partial, PcgState s, size_t B, int chunks) {
    const size_t b = size_t(blockIdx.x) * kThreads + threadIdx.x;
    if (b >= B) return;
    const double rr = s.rr[b], rz = s.rz[b];
    const double denom = rz != 0.0 ? rz : 1.0;
    double rzn = rr, beta = rr / denom;
    if (partial != nullptr && s.use_p[b]) {
        const double d = row_sum(partial, b, chunks);
        if (s.next[b] && (d <= 0.0 || s.pq[b] <= 0.0)) {
            s.use_p[b] = 0;  // restart as plain CG from p = r
            rzn = rr;
            beta = 0.0;
        } else {
            rzn = d;
            beta = d / denom;
        }
    }
    s.rz[b] = rzn;
    s.beta[b] = beta;
    const int nxt = s.next[b];
    s.active[b] = nxt;
    if (nxt) atomicAdd(s.count, 1);
}

// p = (use_p ? z : r) + beta p
template <typename T>
__global__ void __launch_bounds__(kThreads)
    pcg_dir_kernel(T *__restrict__ p, const T *__restrict__ r, const T *__restrict__ z,
                   const PcgState s, size_t m, size_t chunk) {
    const size_t b = blockIdx.y;
    const size_t lo = size_t(blockIdx.x) * chunk;
    const size_t hi = lo + chunk < m ? lo + chunk : m;
    const T c = T(s.beta[b]);
    T *pb = p + b * m;
    const T *zb = (z != nullptr && s.use_p[b]) ? z + b * m : r + b * m;
#pragma unroll 4
    for (size_t i = lo + threadIdx.x; i < hi; i += kThreads) {
        const T cp = c * pb[i];
        pb[i] = zb[i] + cp;
    }
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

size_t adrt_hip_pcg_state_bytes(size_t batch) { return batch >= 1 ? pcg_bytes(batch) : 0; }

int adrt_hip_pcg_update(void *x, void *r, const void *p, const void *q, void *state,
                        void *workspace, size_t workspace_bytes, int dtype, size_t batch,
                        size_t m, void *stream) {
    if (!batch_ok(batch, m)) return ADRT_HIP_ERR_SHAPE;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!x || !r || !p || !q || !state || !workspace) return ADRT_HIP_ERR_NULL;
    if (workspace_bytes < adrt_hip_cg_workspace(batch, m)) return ADRT_HIP_ERR_WORKSPACE;
    const Split sp = split(batch, m);
    const PcgState st = pcg_state(state, batch);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *part = static_cast<double *>(workspace);
    const dim3 grid(unsigned(sp.chunks), unsigned(batch));
    const dim3 rows(unsigned((batch + kThreads - 1) / kThreads));
    if (dtype == ADRT_HIP_F32) {
        hipLaunchKernelGGL(dot_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<const float *>(p), static_cast<const float *>(q), part, m,
                           sp.chunk, sp.chunks);
        hipLaunchKernelGGL(pq_rows_kernel, rows, dim3(kThreads), 0, s, (const double *)part, st,
                           batch, sp.chunks);
        hipLaunchKernelGGL(update_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<float *>(x), static_cast<float *>(r),
                           static_cast<const float *>(p), static_cast<const float *>(q),
                           (const double *)st.alpha, part, m, sp.chunk, sp.chunks);
    } else {
        hipLaunchKernelGGL(dot_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<const double *>(p), static_cast<const double *>(q), part,
                           m, sp.chunk, sp.chunks);
        hipLaunchKernelGGL(pq_rows_kernel, rows, dim3(kThreads), 0, s, (const double *)part, st,
                           batch, sp.chunks);
        hipLaunchKernelGGL(update_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<double *>(x), static_cast<double *>(r),
                           static_cast<const double *>(p), static_cast<const double *>(q),
                           (const double *)st.alpha, part, m, sp.chunk, sp.chunks);
    }
    hipLaunchKernelGGL(rr_rows_kernel, rows, dim3(kThreads), 0, s, (const double *)part, st,
                       batch, sp.chunks);
    return launch_status();
}

int adrt_hip_pcg_direction(void *p, const void *r, const void *z, void *state, void *workspace,
                           size_t workspace_bytes, int dtype, size_t batch, size_t m,
                           void *stream) {
    if (!batch_ok(batch, m)) return ADRT_HIP_ERR_SHAPE;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!p || !r || !state || !workspace) return ADRT_HIP_ERR_NULL;
    if (workspace_bytes < adrt_hip_cg_workspace(batch, m)) return ADRT_HIP_ERR_WORKSPACE;
    const Split sp = split(batch, m);
    const PcgState st = pcg_state(state, batch);
    hipStream_t s = static_cast<hipStream_t>(stream);
    double *part = static_cast<double *>(workspace);
    const dim3 grid(unsigned(sp.chunks), unsigned(batch));
    const dim3 rows(unsigned((batch + kThreads - 1) / kThreads));
    if (z != nullptr) {
        if (dtype == ADRT_HIP_F32)
            hipLaunchKernelGGL(dot_kernel<float>, grid, dim3(kThreads), 0, s,
                               static_cast<const float *>(r), static_cast<const float *>(z), part,
                               m, sp.chunk, sp.chunks);
        else
            hipLaunchKernelGGL(dot_kernel<double>, grid, dim3(kThreads), 0, s,
                               static_cast<const double *>(r), static_cast<const double *>(z),
                               part, m, sp.chunk, sp.chunks);
    }
    hipLaunchKernelGGL(rz_rows_kernel, rows, dim3(kThreads), 0, s,
                       z != nullptr ? (const double *)part : nullptr, st, batch, sp.chunks);
    if (dtype == ADRT_HIP_F32)
        hipLaunchKernelGGL(pcg_dir_kernel<float>, grid, dim3(kThreads), 0, s,
                           static_cast<float *>(p), static_cast<const float *>(r),
                           static_cast<const float *>(z), st, m, sp.chunk);
    else
        hipLaunchKernelGGL(pcg_dir_kernel<double>, grid, dim3(kThreads), 0, s,
                           static_cast<double *>(p), static_cast<const double *>(r),
                           static_cast<const double *>(z), st, m, sp.chunk);
    return launch_status();
}

}  // extern "C"

}  // namespace adrt_hip
