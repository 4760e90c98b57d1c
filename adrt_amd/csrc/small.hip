// One-launch LDS path for small sides (n <= 64) on gfx950.
//
// A work-group owns a chunk of planes (a plane is one quadrant of one image)
// of one size class and carries it through every stage in LDS: the quadrant
// load with zero padding, the log2 n forward stages, the transposed store of
// the forward result and, when a back-projection is asked for, the log2 n
// back-projection stages on the forward result still held in LDS (or on ADRT
// data loaded from memory) and their transposed store.  Nothing but the input
// and the outputs touches HBM, and one launch replaces the 2 + log2 n kernels
// of the streaming path.
//
// Each stage element is the same single IEEE add of the same two operands as
// fwd_stage_kernel and bdrt_stage_kernel (adrt_cdefs_adrt.hpp:55-96,
// adrt_cdefs_bdrt.hpp:55-116), including the zero padding, the literal copy
// and the zero tail of the back-projection, so the results are bit-identical
// to every other path.
//
// The launch is persistent: a grid of resident work-groups strides over the
// chunks of up to kMaxClasses size classes (the small sides of an uneven
// batch), largest side first.  Chunk sizes are chosen so that every chunk
// has the same 128 line segments of 64 elements per stage: one plane at
// n = 64, four (one image) at n = 32, eight at n = 16, and so on.  Within
// every 32 consecutive slots the four planes of one n = 64 image are given
// to work-groups of one XCD (work-group ids are dealt round-robin over the 8
// XCDs), so the image is read from HBM once and from that XCD's L2 three
// times.
//
// With ADRT_HOST_EMULATION defined only the phase functions are compiled, so
// tests/host_emulation can run every chunk on the CPU.
#include <atomic>

#include "small_chunk.hpp"

namespace adrt_hip {
namespace small {

#ifndef ADRT_HOST_EMULATION

template <typename T>
__global__ void __launch_bounds__(kThreads) small_kernel(Job job) {
    __shared__ T lds[2 * kMaxElems];
    const int slots = (job.total + 31) & ~31;
    for (int s = blockIdx.x; s < slots; s += gridDim.x) {
        const int c = slot_chunk(s);
        if (c >= job.total) continue;  // uniform over the work-group
        const Class cl = class_of(job, c);
        const Chunk k = chunk_of(cl, c);
        const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
        const int lane = int(threadIdx.x) & 63;
        run_chunk<T>(cl, k, lds, lds + kMaxElems, [&](auto f) {
            f(wave, lane);
            __syncthreads();
        });
    }
}

namespace {
std::atomic<int> g_cus[64];

int resident_groups(int elem) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < 64) cus = g_cus[dev].load();
    if (cus <= 0) {
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        if (dev >= 0 && dev < 64) g_cus[dev].store(cus);
    }
    // LDS: 2 * kMaxElems elements per work-group, 160 KiB per CU.
    const int per_cu = (160 * 1024) / (2 * kMaxElems * elem);
    const int g = (cus * (per_cu > 0 ? per_cu : 1)) & ~31;
    return g > 32 ? g : 32;
}
}  // namespace

// max_groups > 0 caps the grid (rounded down to a multiple of 32, at least
// 32) so that the launch leaves CUs to work running beside it.
template <typename T>
int launch_job(const Job &job, int max_groups, hipStream_t s) {
    if (job.total == 0) return ADRT_HIP_OK;
    const int slots = (job.total + 31) & ~31;
    int resident = resident_groups(int(sizeof(T)));
    if (max_groups > 0 && max_groups < resident) resident = max_groups < 32 ? 32 : max_groups & ~31;
    const int grid = slots < resident ? slots : resident;
    hipLaunchKernelGGL(small_kernel<T>, dim3(grid), dim3(kThreads), 0, s, job);
    return launch_status();
}

#endif  // ADRT_HOST_EMULATION

}  // namespace small

#ifndef ADRT_HOST_EMULATION

bool small_supported(size_t B, size_t n) {
    if (!is_pow2(n) || n > size_t(small::kMaxN)) return false;
    const long long per = small::planes_per_chunk(num_iters(n));
    return B <= size_t(small::kMaxChunks) && (4 * (long long)B + per - 1) / per <= small::kMaxChunks;
}

// One op over one batch: the forward ADRT (src = images, fwd = out) or the
// back-projection (src = ADRT data, back = out).
int launch_small(const void *src, void *out, int op, int dtype, size_t B, size_t n,
                 hipStream_t s) {
    small::Job job{};
    const bool bd = op == ADRT_HIP_OP_BDRT;
    if (!small::add_class(job, src, bd ? nullptr : out, bd ? out : nullptr, B, n, bd))
        return ADRT_HIP_ERR_TOO_BIG;
    return dtype == ADRT_HIP_F32 ? small::launch_job<float>(job, 0, s)
                                 : small::launch_job<double>(job, 0, s);
}

}  // namespace adrt_hip

extern "C" int adrt_hip_adrt_bdrt_small(const adrt_hip_batch *batches, int count, int dtype,
                                        int max_groups, void *stream) {
    using namespace adrt_hip;
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!batches) return ADRT_HIP_ERR_NULL;
    if (count < 1 || count > small::kMaxClasses || max_groups < 0) return ADRT_HIP_ERR_ARG;
    int order[small::kMaxClasses];
    for (int i = 0; i < count; ++i) {
        const adrt_hip_batch &b = batches[i];
        if (!is_pow2(b.n) || b.n > size_t(small::kMaxN)) return ADRT_HIP_ERR_SHAPE;
        if (!b.adrt && !b.bdrt) return ADRT_HIP_ERR_ARG;
        if (b.batch && !b.images) return ADRT_HIP_ERR_NULL;
        order[i] = i;
    }
    // largest side first (stable)
    for (int i = 1; i < count; ++i)
        for (int j = i; j > 0 && batches[order[j]].n > batches[order[j - 1]].n; --j) {
            const int t = order[j];
            order[j] = order[j - 1];
            order[j - 1] = t;
        }
    small::Job job{};
    for (int i = 0; i < count; ++i) {
        const adrt_hip_batch &b = batches[order[i]];
        if (b.batch == 0) continue;
        if (!small::add_class(job, b.images, b.adrt, b.bdrt, b.batch, b.n, false))
            return ADRT_HIP_ERR_TOO_BIG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    return dtype == ADRT_HIP_F32 ? small::launch_job<float>(job, max_groups, s)
                                 : small::launch_job<double>(job, max_groups, s);
}

#else

}  // namespace adrt_hip

#endif  // ADRT_HOST_EMULATION
