// Forward ADRT + back-projection of many images of uneven power-of-two sides
// (n <= 1024) in one persistent launch with a stage-ordered work queue.
//
// The work of a batch is cut into queue entries ("bundles"), laid out in
// stage order:
//
//   1. forward pass 1 of every image of side n >= 64, largest side first
//      (fused_items.hpp: fwd_pass1 items, `bundle` consecutive items per
//      entry, so that the quadrant pairs that read the same image rows run
//      back to back on one work-group and meet in its XCD's L2);
//   2. every chunk of the sides n <= 32 (small_chunk.hpp: the whole forward
//      and back-projection of a few planes in LDS; no dependencies);
//   3. forward pass 2 of every image of side n >= 64;
//   4. back-projection pass A, then 5. pass B, of those images.
//
// Work-groups take entries in queue order from one atomic counter.  An entry
// of steps 3-5 needs the whole previous pass of its own image: it waits until
// that image's completion counter (raised by every finished entry of the
// previous pass) reaches the pass's item count.  Every entry depends only on
// entries before it in the queue, which were taken earlier by work-groups
// that are running, so every wait ends without any assumption on how many
// work-groups are resident.  Waits are bounded anyway: after kMaxPolls polls
// the wait is counted in the state's timeout word and abandoned, so the
// kernel always exits (the results of such a launch are invalid;
// adrt_hip_queue_timeouts reports it).
//
// Memory ordering: a producer entry's stores are published with an
// agent-scope release fence in every thread, a work-group barrier and the
// counter increment; the consumer's single polling thread takes an
// agent-scope acquire fence after the counter is complete and passes it on
// through a work-group barrier.  (The L2 of each XCD is not coherent with
// the others; the fences write back and invalidate it as needed.)
//
// The intermediate of the forward passes (X) and of the back-projection (Z)
// of one image share one workspace slot: pass A of an image starts only
// after its pass 2, the last reader of X, has finished.
//
// The last work-group to leave resets the queue counter and every completion
// counter, so the next launch (a graph replay) starts from zero without a
// memset.  Launches that share one state must not overlap.
//
// Every element is computed by the same item / chunk code as the per-pass
// kernels, so results are bit-identical to them (and to the reference).
//
// Cost on MI355X (1000 float32 images, sides 32..1024 with p ~ 1/n): 10.6 ms
// per pipeline with 8 items per entry (52 ms with 1), against 1.5 ms for
// one forward and one back-projection call per side on 4 streams.  Timing
// variants put most of the difference on the per-entry publication: the
// release fence writes back the producer XCD's L2 and the acquire
// invalidates the consumer's (4.0 ms without them); taking entries from one
// atomic counter costs the next part (2.8 ms with a static schedule and no
// fences), and the kernel that holds the code of every pass spills
// registers.  A kernel boundary pays for the coherence once per launch, so
// the per-side calls stay the faster pipeline and the autotuner keeps them;
// the queue remains the one-launch form for callers that need it.
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

#include "fused_items.hpp"
#include "small_chunk.hpp"

namespace adrt_hip {
namespace queue {

constexpr int kThreads = 512;
constexpr int kMaxBatches = 16;
constexpr int kMaxSegments = 4 * kMaxBatches;
constexpr int kMaxN = 1024;
constexpr int kSmallMaxN = 32;  // sides run as LDS chunks; larger ones as fused passes
// LDS elements of the largest fused pass with at most 5 stages per pass
// (2^5 rows of fused::kRow); small chunks are sized to fit it too.
constexpr int kLdsElems = 32 * fused::kRow;
constexpr unsigned kMaxPolls = 1u << 21;

static_assert(kThreads == fused::kThreads && kThreads == small::kThreads, "one work-group shape");

enum Pass : int { kPass1 = 0, kPass2 = 1, kPassA = 2, kPassB = 3, kSmall = 4 };

// One run of queue entries: pass `pass` (stages split by `k`) over `images`
// images of side n, or the LDS chunks of one small side.
struct Seg {
    int pass;
    int K;          // fused: stages of the pass's second level (0..2)
    int n;
    int k;          // pass 2: k1 (stages of pass 1); pass B: kA; small: log2 n
    int per_image;  // fused: items per image
    int images;
    int first;      // first queue position
    int entries;    // queue positions in the run
    int bundle;     // fused: items per entry
    int done;       // counter of image 0 raised by this pass (-1: none)
    int wait;       // counter of image 0 this pass waits for (-1: none)
    int wait_target;
    // Fused passes read `in` and write `out` (image i at + i * stride elements):
    //   pass 1: images -> X slot; pass 2: X slot -> forward output;
    //   pass A: forward output -> Z slot; pass B: Z slot -> back-projection.
    // Small chunks: images -> forward output (`out`) and back-projection (`aux`).
    const void *in;
    void *out;
    void *aux;
    long long in_stride;
    long long out_stride;
    long long planes;  // small: 4 * batch
    int per_chunk;     // small: planes per chunk
    int pad[3];
};

struct Header {
    int next;      // queue position counter
    int exited;    // work-groups done
    int timeouts;  // waits abandoned (never reset by the kernel)
    int total;     // queue length
    int segments;
    int counters;  // completion counters after the segment table
    int pad[10];
};

struct State {
    Header h;
    Seg seg[kMaxSegments];
    int counter[1];  // `counters` entries
};

inline size_t state_bytes_for(int counters) {
    return offsetof(State, counter) + size_t(counters > 0 ? counters : 1) * sizeof(int);
}

#ifndef ADRT_HOST_EMULATION

__device__ __forceinline__ int atomic_add(int *p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0: wait until *c >= target, then acquire.
__device__ __forceinline__ void wait_for(int *c, int target, int *timeouts) {
    unsigned polls = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if (++polls > kMaxPolls) {
            atomic_add(timeouts, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(4);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// One fused item of one image: in / out are that image's input and output.
template <typename T, int K>
__device__ __forceinline__ void run_fused(int pass, const T *in, T *out, int n, int k, int item,
                                          T *lds) {
    switch (pass) {
        case kPass1: fused::run_item_at<fused::Pass1<T, K>, T>(item, lds, in, out, n, 0); break;
        case kPass2: fused::run_item_at<fused::Pass2<T, K>, T>(item, lds, in, out, n, k, 0); break;
        case kPassA:
            fused::run_item_at<fused::BdrtA<T, K>, T>(item, lds, in, out, n, 0,
                                                      2 * n - 1 + 5);  // every Z row
            break;
        default: fused::run_item_at<fused::BdrtB<T, K>, T>(item, lds, in, out, n, k, 0, 2 * n - 1); break;
    }
}

template <typename T>
__device__ __forceinline__ void run_small(const Seg &s, int chunk, T *lds) {
    small::Class cl{};
    cl.src = s.in;
    cl.fwd = s.out;
    cl.back = s.aux;
    cl.planes = s.planes;
    cl.log2n = s.k;
    cl.per_chunk = s.per_chunk;
    cl.dC = small::make_div(unsigned(2 * s.n - 1));
    cl.first = 0;
    cl.src_is_adrt = 0;
    const small::Chunk k = small::chunk_of(cl, chunk);
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int lane = int(threadIdx.x) & 63;
    small::run_chunk<T>(cl, k, lds, lds + kLdsElems / 2, [&](auto f) {
        f(wave, lane);
        __syncthreads();
    });
}

// The segment table is read-only during the launch; the header and the
// completion counters are the only memory the work-groups share.
//
// Waves: the register budget the kernel is compiled for (waves per SIMD):
// 8 (64 VGPRs, 4 work-groups per CU; the item code of all passes together
// then spills some registers) or 4 (128 VGPRs, 2 work-groups per CU).
template <typename T, int Waves>
__global__ void __attribute__((amdgpu_flat_work_group_size(kThreads, kThreads),
                               amdgpu_waves_per_eu(Waves)))
queue_kernel(State *st) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    int *bcast = reinterpret_cast<int *>(smem_raw);
    const Seg *__restrict__ segs = st->seg;
    int *counter = st->counter;
    const int tid = int(threadIdx.x);

    if (tid == 0) bcast[0] = atomic_add(&st->h.next, 1);
    __syncthreads();
    int pos = __builtin_amdgcn_readfirstlane(bcast[0]);
    __syncthreads();
    int si = 0;
    while (pos < st->h.total) {
        // Take the next entry now; its number is needed only at the end.
        int ticket = 0;
        if (tid == 0) ticket = atomic_add(&st->h.next, 1);
        // Positions only grow, so the segment index only moves forward.
        while (si + 1 < st->h.segments && segs[si + 1].first <= pos) ++si;
        const int local = pos - segs[si].first;
        if (segs[si].pass == kSmall) {
            run_small<T>(segs[si], local, lds);
        } else {
            // Decode the entry into the few values the items need.
            const Seg &s = segs[si];
            const int pass = s.pass, K = s.K, n = s.n, k = s.k;
            const int item0 = local * s.bundle;
            const int image = item0 / s.per_image;
            const int first = item0 - image * s.per_image;
            const int items = min(s.bundle, s.images * s.per_image - item0);
            const T *in = static_cast<const T *>(s.in) + image * s.in_stride;
            T *out = static_cast<T *>(s.out) + image * s.out_stride;
            int *done = s.done >= 0 ? counter + s.done + image : nullptr;
            if (s.wait >= 0) {
                if (tid == 0) wait_for(counter + s.wait + image, s.wait_target, &st->h.timeouts);
                __syncthreads();
            }
            for (int i = 0; i < items; ++i) {
                if (i > 0) __syncthreads();  // the previous item's phase 2 has read the LDS
                switch (K) {
                    case 0: run_fused<T, 0>(pass, in, out, n, k, first + i, lds); break;
                    case 1: run_fused<T, 1>(pass, in, out, n, k, first + i, lds); break;
                    default: run_fused<T, 2>(pass, in, out, n, k, first + i, lds); break;
                }
            }
            if (done) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __syncthreads();
                if (tid == 0)
                    __hip_atomic_fetch_add(done, items, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        if (tid == 0) bcast[0] = ticket;
        __syncthreads();
        pos = __builtin_amdgcn_readfirstlane(bcast[0]);
        __syncthreads();
    }

    // The last work-group out resets the counters for the next launch.
    if (tid == 0)
        bcast[0] = __hip_atomic_fetch_add(&st->h.exited, 1, __ATOMIC_ACQ_REL,
                                          __HIP_MEMORY_SCOPE_AGENT) == int(gridDim.x) - 1;
    __syncthreads();
    if (bcast[0]) {
        const int nc = st->h.counters;
        for (int i = tid; i < nc; i += kThreads)
            __hip_atomic_store(&counter[i], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tid == 0) {
            __hip_atomic_store(&st->h.next, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->h.exited, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <auto Kern>
static hipError_t lds_limit(size_t lds) {
    constexpr int kMaxDevices = 64;
    static std::once_flag once[kMaxDevices];
    static hipError_t result[kMaxDevices];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return hipErrorInvalidDevice;
    }
    std::call_once(once[dev], [&] {
        // the largest the kernel ever asks for: kLdsElems of double
        result[dev] = hipFuncSetAttribute(reinterpret_cast<const void *>(Kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          int(kLdsElems * sizeof(double)));
        if (result[dev] != hipSuccess) (void)hipGetLastError();
    });
    (void)lds;
    return result[dev];
}

static int compute_units() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    return cus;
}

#endif  // ADRT_HOST_EMULATION

// ---------------------------------------------------------------------------
// Host planning.
// ---------------------------------------------------------------------------

struct Plan {
    State host;  // header + segments (counters are zero)
    int counters = 0;
    size_t ws_bytes = 0;
    int lds_bytes = 0;
};

inline size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

inline size_t slot_elems(int n, size_t elem) {
    const int L = num_iters(size_t(n));
    const size_t x = fused::x_bytes_per_image(n, L, elem);
    const size_t z = fused::z_bytes_per_image(n, elem);
    return round_up(x > z ? x : z, 256) / elem;
}

// Fused item counts per image.
inline int items_per_image(int pass, int n, int k1, int kA) {
    // The counts do not depend on T; float instantiations are used for them.
    switch (pass) {
        case kPass1:
            switch (k1 - 3) {
                case 0: return fused::Pass1<float, 0>::num_items(n, 1);
                case 1: return fused::Pass1<float, 1>::num_items(n, 1);
                default: return fused::Pass1<float, 2>::num_items(n, 1);
            }
        case kPass2:
            switch (num_iters(size_t(n)) - k1 - 3) {
                case 0: return fused::Pass2<float, 0>::num_items(n, k1, 1);
                case 1: return fused::Pass2<float, 1>::num_items(n, k1, 1);
                default: return fused::Pass2<float, 2>::num_items(n, k1, 1);
            }
        case kPassA:
            switch (kA - 3) {
                case 0: return fused::BdrtA<float, 0>::num_items(n, 1);
                case 1: return fused::BdrtA<float, 1>::num_items(n, 1);
                default: return fused::BdrtA<float, 2>::num_items(n, 1);
            }
        default:
            switch (num_iters(size_t(n)) - kA - 3) {
                case 0: return fused::BdrtB<float, 0>::num_items(kA, 1, 2 * n - 1);
                case 1: return fused::BdrtB<float, 1>::num_items(kA, 1, 2 * n - 1);
                default: return fused::BdrtB<float, 2>::num_items(kA, 1, 2 * n - 1);
            }
    }
}

// LDS elements of a fused pass with second-level depth K.
inline size_t pass_lds_elems(int pass, int K) {
    switch (pass) {
        case kPass1:
            return K == 0 ? fused::Pass1<float, 0>::lds_elems
                          : K == 1 ? fused::Pass1<float, 1>::lds_elems : fused::Pass1<float, 2>::lds_elems;
        case kPass2:
            return K == 0 ? fused::Pass2<float, 0>::lds_elems
                          : K == 1 ? fused::Pass2<float, 1>::lds_elems : fused::Pass2<float, 2>::lds_elems;
        case kPassA:
            return K == 0 ? fused::BdrtA<float, 0>::lds_elems
                          : K == 1 ? fused::BdrtA<float, 1>::lds_elems : fused::BdrtA<float, 2>::lds_elems;
        default:
            return K == 0 ? fused::BdrtB<float, 0>::lds_elems
                          : K == 1 ? fused::BdrtB<float, 1>::lds_elems : fused::BdrtB<float, 2>::lds_elems;
    }
}

// Planes per small chunk: the default chunk, shrunk until two buffers of it
// fit kLdsElems.
inline int small_per_chunk(int n) {
    const int L = num_iters(size_t(n));
    const int plane = n * (2 * n - 1);
    int per = small::planes_per_chunk(L);
    const int fit = (kLdsElems / 2) / plane;
    if (per > fit) per = fit;
    return per > 0 ? per : 1;
}

// Variant -> stage split, as the per-pass kernels choose it.
int fused_variant_for(int op, size_t n);

// Builds the queue for `count` batches (validated by the caller) into plan;
// workspace == nullptr only sizes it.  Returns an ADRT_HIP_* status.
int build(const adrt_hip_batch *b, int count, int dtype, int bundle, void *workspace, Plan &p);

}  // namespace queue
}  // namespace adrt_hip

#ifndef ADRT_HOST_EMULATION

namespace adrt_hip {
namespace queue {

int build(const adrt_hip_batch *b, int count, int dtype, int bundle, void *workspace, Plan &p) {
    const size_t elem = dtype == ADRT_HIP_F64 ? 8 : 4;
    p = Plan{};
    // batch order: larger sides first (stable)
    int order[kMaxBatches];
    for (int i = 0; i < count; ++i) order[i] = i;
    for (int i = 1; i < count; ++i)
        for (int j = i; j > 0 && b[order[j]].n > b[order[j - 1]].n; --j) {
            const int t = order[j];
            order[j] = order[j - 1];
            order[j - 1] = t;
        }
    struct Big {
        const adrt_hip_batch *b;
        int k1, kA, c0;  // stage splits; first completion counter
        size_t slot0;    // first slot (in slot_elems of its own side)
        char *slots;
    };
    Big big[kMaxBatches];
    int nbig = 0;
    size_t ws = 0;
    int counters = 0;
    for (int i = 0; i < count; ++i) {
        const adrt_hip_batch &e = b[order[i]];
        if (e.batch == 0 || e.n <= size_t(kSmallMaxN)) continue;
        Big &g = big[nbig++];
        g.b = &e;
        const int L = num_iters(e.n);
        g.k1 = fused::fwd_pass1_stages(L, fused_variant_for(ADRT_HIP_OP_ADRT, e.n), int(elem));
        g.kA = fused::bdrt_passA_stages(L, fused_variant_for(ADRT_HIP_OP_BDRT, e.n));
        g.c0 = counters;
        counters += 3 * int(e.batch);
        g.slots = workspace ? static_cast<char *>(workspace) + ws : nullptr;
        ws += e.batch * slot_elems(int(e.n), elem) * elem;
    }
    int nseg = 0;
    long long pos = 0;
    size_t lds = 0;
    auto add_fused = [&](const Big &g, int pass) {
        Seg &s = p.host.seg[nseg++];
        const adrt_hip_batch &e = *g.b;
        const int n = int(e.n), L = num_iters(e.n);
        s.pass = pass;
        s.n = n;
        s.K = pass == kPass1 ? g.k1 - 3 : pass == kPass2 ? L - g.k1 - 3
                                        : pass == kPassA ? g.kA - 3 : L - g.kA - 3;
        s.k = pass == kPass2 ? g.k1 : pass == kPassB ? g.kA : 0;
        s.per_image = items_per_image(pass, n, g.k1, g.kA);
        s.images = int(e.batch);
        s.bundle = bundle;
        while (s.per_image % s.bundle) s.bundle /= 2;  // an entry never spans two images
        s.first = int(pos);
        s.entries = int((size_t(s.images) * s.per_image + s.bundle - 1) / s.bundle);
        pos += s.entries;
        const long long plane = (long long)n * (2 * n - 1);
        const long long slot = (long long)slot_elems(n, elem);
        s.done = pass == kPassB ? -1 : g.c0 + pass * s.images;
        s.wait = pass == kPass1 ? -1 : g.c0 + (pass - 1) * s.images;
        s.wait_target = pass == kPass1 ? 0 : items_per_image(pass - 1, n, g.k1, g.kA);
        switch (pass) {
            case kPass1:
                s.in = e.images, s.in_stride = (long long)n * n;
                s.out = g.slots, s.out_stride = slot;
                break;
            case kPass2:
                s.in = g.slots, s.in_stride = slot;
                s.out = e.adrt, s.out_stride = 4 * plane;
                break;
            case kPassA:
                s.in = e.adrt, s.in_stride = 4 * plane;
                s.out = g.slots, s.out_stride = slot;
                break;
            default:
                s.in = g.slots, s.in_stride = slot;
                s.out = e.bdrt, s.out_stride = 4 * plane;
                break;
        }
        const size_t need = pass_lds_elems(pass, s.K) * elem;
        if (need > lds) lds = need;
    };
    for (int i = 0; i < nbig; ++i) add_fused(big[i], kPass1);
    for (int i = 0; i < count; ++i) {
        const adrt_hip_batch &e = b[order[i]];
        if (e.batch == 0 || e.n > size_t(kSmallMaxN)) continue;
        Seg &s = p.host.seg[nseg++];
        const int n = int(e.n);
        s.pass = kSmall;
        s.n = n;
        s.k = num_iters(e.n);
        s.images = int(e.batch);
        s.planes = 4ll * (long long)e.batch;
        s.per_chunk = small_per_chunk(n);
        s.first = int(pos);
        s.entries = int((s.planes + s.per_chunk - 1) / s.per_chunk);
        pos += s.entries;
        s.done = s.wait = -1;
        s.in = e.images;
        s.out = e.adrt;
        s.aux = e.bdrt;
        // buffer a at 0, buffer b at kLdsElems / 2 (run_small)
        const size_t need =
            (size_t(kLdsElems / 2) + size_t(s.per_chunk) * size_t(n) * size_t(2 * n - 1)) * elem;
        if (need > lds) lds = need;
    }
    for (int i = 0; i < nbig; ++i) add_fused(big[i], kPass2);
    for (int i = 0; i < nbig; ++i)
        if (big[i].b->bdrt) add_fused(big[i], kPassA);
    for (int i = 0; i < nbig; ++i)
        if (big[i].b->bdrt) add_fused(big[i], kPassB);
    if (pos >= (1ll << 30)) return ADRT_HIP_ERR_TOO_BIG;
    p.host.h.total = int(pos);
    p.host.h.segments = nseg;
    p.host.h.counters = counters;
    p.counters = counters;
    p.ws_bytes = ws;
    p.lds_bytes = int(lds < 16 ? 16 : lds);
    return ADRT_HIP_OK;
}

int fused_variant_for(int op, size_t n) { return fused_variant(op, n); }

static int validate(const adrt_hip_batch *b, int count, int dtype, int bundle) {
    if (dtype != ADRT_HIP_F32 && dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (!b) return ADRT_HIP_ERR_NULL;
    if (count < 1 || count > kMaxBatches) return ADRT_HIP_ERR_ARG;
    if (bundle != 1 && bundle != 2 && bundle != 4 && bundle != 8) return ADRT_HIP_ERR_ARG;
    for (int i = 0; i < count; ++i) {
        const adrt_hip_batch &e = b[i];
        if (!is_pow2(e.n) || e.n > size_t(kMaxN)) return ADRT_HIP_ERR_SHAPE;
        if (e.batch == 0) continue;
        if (e.batch > (size_t(1) << 20)) return ADRT_HIP_ERR_TOO_BIG;
        if (!e.images) return ADRT_HIP_ERR_NULL;
        if (!e.adrt && !e.bdrt) return ADRT_HIP_ERR_ARG;
        // the fused passes read the forward output back for the back-projection
        if (e.n > size_t(kSmallMaxN) && !e.adrt) return ADRT_HIP_ERR_NULL;
    }
    return ADRT_HIP_OK;
}

}  // namespace queue
}  // namespace adrt_hip

using namespace adrt_hip;

extern "C" size_t adrt_hip_queue_workspace_bytes(const adrt_hip_batch *batches, int count,
                                                 int dtype) {
    if (queue::validate(batches, count, dtype, 2) != ADRT_HIP_OK) return 0;
    queue::Plan p;
    if (queue::build(batches, count, dtype, 2, nullptr, p) != ADRT_HIP_OK) return 0;
    return p.ws_bytes;
}

extern "C" size_t adrt_hip_queue_state_bytes(const adrt_hip_batch *batches, int count) {
    if (queue::validate(batches, count, ADRT_HIP_F32, 2) != ADRT_HIP_OK) return 0;
    queue::Plan p;
    if (queue::build(batches, count, ADRT_HIP_F32, 2, nullptr, p) != ADRT_HIP_OK) return 0;
    return queue::state_bytes_for(p.counters);
}

extern "C" int adrt_hip_queue_plan(const adrt_hip_batch *batches, int count, int dtype,
                                   int bundle, int waves, void *workspace, size_t workspace_bytes,
                                   void *state, size_t state_bytes, adrt_hip_queue *plan) {
    if (!plan) return ADRT_HIP_ERR_NULL;
    if (bundle == 0) bundle = 8;
    if (waves == 0) waves = dtype == ADRT_HIP_F64 ? 4 : 8;
    if (waves != 4 && (waves != 8 || dtype == ADRT_HIP_F64)) return ADRT_HIP_ERR_ARG;
    int st = queue::validate(batches, count, dtype, bundle);
    if (st != ADRT_HIP_OK) return st;
    if (!state) return ADRT_HIP_ERR_NULL;
    queue::Plan p;
    st = queue::build(batches, count, dtype, bundle, nullptr, p);
    if (st != ADRT_HIP_OK) return st;
    if (p.ws_bytes > 0 && !workspace) return ADRT_HIP_ERR_NULL;
    if (workspace_bytes < p.ws_bytes) return ADRT_HIP_ERR_WORKSPACE;
    const size_t need = queue::state_bytes_for(p.counters);
    if (state_bytes < need) return ADRT_HIP_ERR_WORKSPACE;
    queue::build(batches, count, dtype, bundle, workspace, p);
    // header + segments, then zeroed counters, in one synchronous copy
    std::vector<unsigned char> img(need, 0);
    std::memcpy(img.data(), &p.host, offsetof(queue::State, counter));
    const hipError_t e = hipMemcpy(state, img.data(), need, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_last_error(e);
        return ADRT_HIP_ERR_DEVICE;
    }
    plan->state = state;
    plan->dtype = dtype;
    plan->entries = p.host.h.total;
    plan->segments = p.host.h.segments;
    plan->counters = p.counters;
    plan->lds_bytes = p.lds_bytes;
    plan->waves = waves;
    return ADRT_HIP_OK;
}

extern "C" int adrt_hip_queue_run(const adrt_hip_queue *plan, int max_groups, void *stream) {
    if (!plan || !plan->state) return ADRT_HIP_ERR_NULL;
    if (plan->dtype != ADRT_HIP_F32 && plan->dtype != ADRT_HIP_F64) return ADRT_HIP_ERR_DTYPE;
    if (max_groups < 0 || plan->lds_bytes <= 0 ||
        size_t(plan->lds_bytes) > queue::kLdsElems * sizeof(double) ||
        (plan->waves != 4 && plan->waves != 8) || (plan->waves == 8 && plan->dtype != ADRT_HIP_F32))
        return ADRT_HIP_ERR_ARG;
    if (plan->entries == 0) return ADRT_HIP_OK;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // work-groups per CU: 160 KiB of LDS, and `waves` per SIMD of 8-wave groups
    const int by_lds = (160 * 1024) / plan->lds_bytes;
    const int by_waves = plan->waves / 2;
    const int per_cu = by_lds < by_waves ? by_lds : by_waves;
    int grid = queue::compute_units() * (per_cu > 0 ? per_cu : 1);
    if (max_groups > 0 && max_groups < grid) grid = max_groups;
    if (grid > plan->entries) grid = plan->entries;
    auto *st = static_cast<queue::State *>(plan->state);
    hipError_t e;
    if (plan->dtype == ADRT_HIP_F32 && plan->waves == 8) {
        e = queue::lds_limit<queue::queue_kernel<float, 8>>(size_t(plan->lds_bytes));
        if (e == hipSuccess)
            hipLaunchKernelGGL((queue::queue_kernel<float, 8>), dim3(grid), dim3(queue::kThreads),
                               size_t(plan->lds_bytes), s, st);
    } else if (plan->dtype == ADRT_HIP_F32) {
        e = queue::lds_limit<queue::queue_kernel<float, 4>>(size_t(plan->lds_bytes));
        if (e == hipSuccess)
            hipLaunchKernelGGL((queue::queue_kernel<float, 4>), dim3(grid), dim3(queue::kThreads),
                               size_t(plan->lds_bytes), s, st);
    } else {
        e = queue::lds_limit<queue::queue_kernel<double, 4>>(size_t(plan->lds_bytes));
        if (e == hipSuccess)
            hipLaunchKernelGGL((queue::queue_kernel<double, 4>), dim3(grid), dim3(queue::kThreads),
                               size_t(plan->lds_bytes), s, st);
    }
    if (e != hipSuccess) {
        set_last_error(e);
        return ADRT_HIP_ERR_DEVICE;
    }
    return launch_status();
}

extern "C" int adrt_hip_queue_timeouts(const adrt_hip_queue *plan, int *count) {
    if (!plan || !plan->state || !count) return ADRT_HIP_ERR_NULL;
    auto *st = static_cast<queue::State *>(plan->state);
    int v = 0;
    hipError_t e = hipMemcpy(&v, &st->h.timeouts, sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess && v != 0) e = hipMemset(&st->h.timeouts, 0, sizeof(int));
    if (e != hipSuccess) {
        set_last_error(e);
        return ADRT_HIP_ERR_DEVICE;
    }
    *count = v;
    return ADRT_HIP_OK;
}

#endif  // ADRT_HOST_EMULATION
