// Chunks of the one-launch LDS path for small sides (n <= 64): the chunk
// sizes, the per-phase load / stage / store code and the whole-chunk
// driver run_chunk.  small.hip launches them in its persistent kernel; the
// queue kernel (queue.hip) runs chunks of sides n <= 32 from its work queue.
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
#pragma once

#include "common.hpp"

namespace adrt_hip {
namespace small {

constexpr int kMaxLog2 = 6;                  // n <= 64
constexpr int kMaxN = 1 << kMaxLog2;
constexpr int kMaxElems = kMaxN * (2 * kMaxN - 1);  // LDS elements per buffer
constexpr int kThreads = 512;
constexpr int kMaxClasses = 8;

// Exact e / d for e * d < 2^32 (every use: e < kMaxElems, d < 128):
// with m = ceil(2^32 / d), e * m / 2^32 exceeds e / d by less than e / 2^32,
// which is below 1 / d, so the floor is unchanged.
struct Div {
    unsigned m;
    unsigned d;
};

__host__ __device__ inline Div make_div(unsigned d) {
    return Div{d > 1 ? unsigned(0xFFFFFFFFull / d + 1) : 0u, d};
}

__host__ __device__ __forceinline__ unsigned divide(unsigned e, Div v) {
#ifdef ADRT_HOST_EMULATION
    const unsigned q = unsigned((static_cast<unsigned long long>(e) * v.m) >> 32);
#else
    const unsigned q = __umulhi(e, v.m);
#endif
    return v.d == 1 ? e : q;
}

// A size class: `src` holds images (batch, n, n), or for a back-projection
// alone ADRT data (batch, 4, 2n-1, n); `fwd` (forward result) and `back`
// (back-projection of the forward result, or of `src` when src_is_adrt) are
// (batch, 4, 2n-1, n) or null.
struct Class {
    const void *src;
    void *fwd;
    void *back;
    long long planes;  // 4 * batch
    Div dC;            // division by C = 2n - 1
    int log2n;
    int per_chunk;     // planes per chunk
    int first;         // index of the class's first chunk
    int src_is_adrt;
};

struct Job {
    Class cls[kMaxClasses];
    int count;
    int total;  // chunks over all classes
};

// Work is dealt to the kWaves wavefronts of a work-group in units of one
// wavefront-wide piece (a line segment of 64 elements, an image row, or 64
// consecutive elements of memory); a phase has at most kWaves * kUnits
// units, so a wavefront takes at most kUnits of them (unit wave + t * kWaves).
// Everything that depends on the unit alone (line, angle, section, the
// operand lines) is wavefront-uniform and computed on the scalar unit; a
// lane adds only its column.  Each phase first gathers all of its operands
// into registers, reading clamped in-bounds addresses for units and lanes
// past the end so that the loads do not wait on one another, then writes.
constexpr int kWaves = kThreads / 64;
constexpr int kUnits = 16;
constexpr int kMaxUnits = kWaves * kUnits;

// Lines of C = 2n-1 elements take ceil(C / 64) segments: 2 at n = 64, else 1.
__host__ __device__ inline int segments_log2(int log2n) { return log2n >= 6 ? 1 : 0; }

// Planes per chunk: a chunk has kMaxUnits line segments (and at most
// kMaxUnits image rows): 1 plane at n = 64, 4 at 32, 8 at 16, ..., 128 at 1.
inline int planes_per_chunk(int log2n) {
    const int n = 1 << log2n;
    const int per = kMaxUnits / (n << segments_log2(log2n));
    return per > 0 ? per : 1;
}

// One chunk: planes [p0, p0 + np) of a class of side n = 2^L.
struct Chunk {
    long long p0;
    int np;
    int L;
    int n;
    int C;    // 2n - 1
    int seg;  // log2 of the segments per line
    Div dC;   // division by C
};

// Slot s of the persistent loop -> chunk.  A bijection on every aligned
// group of 32 slots: slots with equal s mod 8 (one XCD when the grid is a
// multiple of 8) get four consecutive chunks.
__host__ __device__ __forceinline__ int slot_chunk(int s) {
    return (s & ~31) | ((s & 7) << 2) | ((s >> 3) & 3);
}

__host__ __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// Quadrant load (adrt_cdefs_adrt.hpp:123-186, as fwd_load_kernel) into
// lines (plane, r) of C elements:
//   q0: img[r][n-1-c]  q1: img[n-1-c][r]  q2: img[c][r]  q3: img[n-1-r][n-1-c]
// A unit is one image row i of one plane, read coalesced (lane j = column)
// and scattered into LDS; the same unit zeroes columns [n, C) of line
// (plane, i).
template <typename T>
__host__ __device__ __forceinline__ void load_image(const T *img, T *buf, const Chunk &k, int wave,
                                                    int lane) {
    const int n = k.n, L = k.L, C = k.C;
    const int units = k.np * n;
    const T *base = img + size_t(k.p0 >> 2) * size_t(n) * size_t(n);
    const int q0 = int(k.p0 & 3);
    const int j = imin(lane, n - 1);
    const int jC = j * C;
    T v[kUnits];
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int u = imin(wave + t * kWaves, units - 1);
        const int pl = u >> L, i = u & (n - 1);
        v[t] = base[(size_t((q0 + pl) >> 2) * size_t(n) + size_t(i)) * size_t(n) + size_t(j)];
    }
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int u = wave + t * kWaves;
        if (u >= units) break;
        const int pl = u >> L, i = u & (n - 1);
        const int plane = pl * n * C;
        int at;
        switch ((q0 + pl) & 3) {
            case 0: at = plane + i * C + (n - 1 - j); break;
            case 1: at = plane + jC + (n - 1 - i); break;
            case 2: at = plane + jC + i; break;
            default: at = plane + (n - 1 - i) * C + (n - 1 - j); break;
        }
        if (lane < n) buf[at] = v[t];
        if (lane < n - 1) buf[plane + i * C + n + lane] = T(0);
    }
}

// ADRT data (planes, C, n) -> lines (plane, angle) of C rows (the transposed
// load of the back-projection).  A unit is 64 consecutive elements of the
// chunk's memory: element f is row R = f / n (R = plane * C + r), angle
// a = f mod n.
template <typename T>
__host__ __device__ __forceinline__ void load_adrt(const T *src, T *buf, const Chunk &k, int wave,
                                                   int lane) {
    const int n = k.n, L = k.L, C = k.C;
    const int total = k.np * C * n;
    const int units = (total + 63) >> 6;
    const T *base = src + size_t(k.p0) * size_t(C) * size_t(n);
    T v[kUnits];
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int f = imin((wave + t * kWaves) * 64 + lane, total - 1);
        v[t] = base[f];
    }
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int u = wave + t * kWaves;
        if (u >= units) break;
        const int f = u * 64 + lane;
        if (f < total) {
            const int R = f >> L, a = f & (n - 1);
            const int pl = int(divide(unsigned(R), k.dC));
            buf[(pl * n + a) * C + (R - pl * C)] = v[t];
        }
    }
}

// Lines (plane, angle) of C elements -> (planes, C, n) in memory, by the
// units of load_adrt.
template <typename T>
__host__ __device__ __forceinline__ void store_planes(const T *buf, T *dst, const Chunk &k, int wave,
                                                      int lane) {
    const int n = k.n, L = k.L, C = k.C;
    const int total = k.np * C * n;
    const int units = (total + 63) >> 6;
    T *base = dst + size_t(k.p0) * size_t(C) * size_t(n);
    T v[kUnits];
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int f = imin((wave + t * kWaves) * 64 + lane, total - 1);
        const int R = f >> L, a = f & (n - 1);
        const int pl = int(divide(unsigned(R), k.dC));
        v[t] = buf[(pl * n + a) * C + (R - pl * C)];
    }
#pragma unroll
    for (int t = 0; t < kUnits; ++t) {
        const int u = wave + t * kWaves;
        if (u >= units) break;
        const int f = u * 64 + lane;
        if (f < total) base[f] = v[t];
    }
}

// Stage units: one 64-element segment of a pair of output lines that read
// the same two input lines, so the pair's index math is done once; a chunk
// has kMaxUnits / 2 of them per stage.
constexpr int kPairUnits = kUnits / 2;

// Forward stage with S = 2^log2s angles per line group (fwd_stage_kernel):
// out line (bqr, a): out[c] = x[c] (c < ceil(a/2)) else x[c] + y[c - ceil(a/2)],
// x = input line (2 bqr, a/2), y = input line (2 bqr + 1, a/2).  Angles
// a = 2k and 2k + 1 share x and y (shifts k and k + 1): a unit reads x[c],
// y[c - k] and y[c - k - 1] and writes both lines.
template <typename T>
__host__ __device__ __forceinline__ void fwd_stage(const T *in, T *out, const Chunk &k, int log2s,
                                                   int wave, int lane) {
    const int C = k.C, seg = k.seg;
    const int S_in = (1 << log2s) >> 1;
    const int units = ((k.np * k.n) >> 1) << seg;
    T ve[kPairUnits], vo[kPairUnits];
#pragma unroll
    for (int t = 0; t < kPairUnits; ++t) {
        const int u = imin(wave + t * kWaves, units - 1);
        const int pair = u >> seg;  // = bqr * S_in + kk
        const int kk = pair & (S_in - 1);
        const int bqr = pair >> (log2s - 1);
        const int xo = ((2 * bqr) * S_in + kk) * C;
        const int yo = xo + S_in * C;
        const int c = ((u & ((1 << seg) - 1)) << 6) + lane;
        const T xv = in[xo + imin(c, C - 1)];
        const T y0 = in[yo + imin(c >= kk ? c - kk : 0, C - 1)];
        const T y1 = in[yo + imin(c > kk ? c - kk - 1 : 0, C - 1)];
        ve[t] = c >= kk ? xv + y0 : xv;
        vo[t] = c > kk ? xv + y1 : xv;
    }
#pragma unroll
    for (int t = 0; t < kPairUnits; ++t) {
        const int u = wave + t * kWaves;
        if (u >= units) break;
        const int c = ((u & ((1 << seg) - 1)) << 6) + lane;
        if (c < C) {
            T *o = out + 2 * (u >> seg) * C + c;
            o[0] = ve[t];
            o[C] = vo[t];
        }
    }
}

// Back-projection stage with R = 2^(L-log2s) angle pairs and S = 2^log2s
// sections (bdrt_stage_kernel): left sections x[r] + y[r]; right sections
// x[r+si] + y[r+si+1], a copy of x[C-1] at r + si == C-1 and zeros below.
// The left and right sections of one section pair read the same x and y
// lines: a unit writes both.
template <typename T>
__host__ __device__ __forceinline__ void bdrt_stage(const T *in, T *out, const Chunk &k, int log2s,
                                                    int wave, int lane) {
    const int C = k.C, seg = k.seg;
    const int S_in = (1 << log2s) >> 1;
    const int Rm = (1 << (k.L - log2s)) - 1;
    const int units = ((k.np * k.n) >> 1) << seg;
    T vl[kPairUnits], vr[kPairUnits];
#pragma unroll
    for (int t = 0; t < kPairUnits; ++t) {
        const int u = imin(wave + t * kWaves, units - 1);
        const int pair = u >> seg;  // = bqsi * S_in + sec
        const int sec = pair & (S_in - 1);
        const int bqsi = pair >> (log2s - 1);
        const int si = bqsi & Rm;
        const int xo = ((2 * bqsi) * S_in + sec) * C;
        const int yo = xo + S_in * C;
        const int r = ((u & ((1 << seg) - 1)) << 6) + lane;
        const int rc = imin(r, C - 1);
        const int row = r + si;
        const T xl = in[xo + rc];
        const T yl = in[yo + rc];
        const T xr = in[xo + imin(row, C - 1)];
        const T yr = in[yo + imin(row + 1, C - 1)];
        vl[t] = xl + yl;
        T w = xr + yr;
        if (row >= C - 1) w = row == C - 1 ? xr : T(0);
        vr[t] = w;
    }
#pragma unroll
    for (int t = 0; t < kPairUnits; ++t) {
        const int u = wave + t * kWaves;
        if (u >= units) break;
        const int r = ((u & ((1 << seg) - 1)) << 6) + lane;
        if (r < C) {
            T *o = out + 2 * (u >> seg) * C + r;
            o[0] = vl[t];
            o[C] = vr[t];
        }
    }
}

// Class and chunk of chunk index c (classes in launch order; unrolled
// selects keep the class table in kernel arguments).
__host__ __device__ __forceinline__ Class class_of(const Job &job, int c) {
    Class cl = job.cls[0];
#pragma unroll
    for (int i = 1; i < kMaxClasses; ++i)
        if (i < job.count && c >= job.cls[i].first) cl = job.cls[i];
    return cl;
}

__host__ __device__ __forceinline__ Chunk chunk_of(const Class &cl, int c) {
    Chunk k;
    k.L = cl.log2n;
    k.n = 1 << cl.log2n;
    k.C = 2 * k.n - 1;
    k.p0 = (long long)(c - cl.first) * cl.per_chunk;
    const long long left = cl.planes - k.p0;
    k.np = int(left < cl.per_chunk ? left : cl.per_chunk);
    k.seg = segments_log2(cl.log2n);
    k.dC = cl.dC;
    return k;
}

// The whole chunk; exec(f) runs f(wave, lane) for every thread of the
// work-group and then synchronises it.  Buffers a and b hold kMaxElems each.
template <typename T, typename Exec>
__host__ __device__ __forceinline__ void run_chunk(const Class &cl, const Chunk &k, T *a, T *b,
                                                   Exec &&exec) {
    T *cur = a, *nxt = b;
    if (cl.src_is_adrt) {
        exec([&](int w, int l) { load_adrt(static_cast<const T *>(cl.src), cur, k, w, l); });
    } else {
        exec([&](int w, int l) { load_image(static_cast<const T *>(cl.src), cur, k, w, l); });
        for (int s = 1; s <= k.L; ++s) {
            exec([&](int w, int l) { fwd_stage(cur, nxt, k, s, w, l); });
            T *t = cur;
            cur = nxt;
            nxt = t;
        }
        // The store reads `cur`, the first back-projection stage reads `cur`
        // and writes `nxt`: no synchronisation is needed between them.
        if (cl.fwd) {
            exec([&](int w, int l) { store_planes(cur, static_cast<T *>(cl.fwd), k, w, l); });
        }
    }
    if (cl.back) {
        for (int s = 1; s <= k.L; ++s) {
            exec([&](int w, int l) { bdrt_stage(cur, nxt, k, s, w, l); });
            T *t = cur;
            cur = nxt;
            nxt = t;
        }
        exec([&](int w, int l) { store_planes(cur, static_cast<T *>(cl.back), k, w, l); });
    }
}

// Appends a class of `batch` images of side n (a power of two <= kMaxN);
// false when the job is full or would exceed kMaxChunks chunks.
constexpr long long kMaxChunks = 1ll << 30;

inline bool add_class(Job &job, const void *src, void *fwd, void *back, size_t batch, size_t n,
                      bool src_is_adrt) {
    if (job.count >= kMaxClasses || batch > size_t(kMaxChunks)) return false;
    Class &cl = job.cls[job.count];
    cl.src = src;
    cl.fwd = fwd;
    cl.back = back;
    cl.log2n = num_iters(n);
    cl.planes = (long long)(4 * batch);
    cl.per_chunk = planes_per_chunk(cl.log2n);
    const unsigned C = unsigned(2 * n - 1);
    cl.dC = make_div(C);
    cl.first = job.total;
    cl.src_is_adrt = src_is_adrt ? 1 : 0;
    const long long chunks = (cl.planes + cl.per_chunk - 1) / cl.per_chunk;
    if (job.total + chunks > kMaxChunks) return false;
    job.total += int(chunks);
    job.count += 1;
    return true;
}

}  // namespace small
}  // namespace adrt_hip
