// Fused two-pass forward ADRT for gfx950 (log2 n in [6, 12]).
//
// Structure.  Stage i of the reference (adrt_core, adrt_cdefs_adrt.hpp:55-96)
// merges row blocks (2r, 2r+1) and maps angle a to a/2 with the odd block
// shifted by ceil(a/2).  Unrolling the recursion splits the L = log2 n stages
// into two passes with one intermediate X in HBM:
//
//   pass 1: for every row block J of 2^k1 rows, the first k1 stages
//           -> X[J][s][c] for the 2^k1 stage-k1 angles s
//   pass 2: for every angle s, the last k2 = L - k1 stages over the 2^k2
//           blocks X[J][s][.], where block J is pre-shifted by s*J; the pass
//           produces the final angles s*2^k2 + t directly in output layout.
//
// Each pass is itself two "levels" of at most 3 stages.  A level holds 2^K
// rows x W offsets per thread in registers and runs its K stages with
// compile-time shifts (ceil(t/2) of the local angle t); the level input angle
// m contributes m * j to the shift of local row j, which is applied as a skew
// when the rows are loaded.  Levels exchange data through LDS.  Every output
// element is produced by the same sequence of single IEEE adds, with the same
// operands, as in the reference, so results are bit-identical.  The
// reference's plain copy for c < ceil(a/2) is reproduced by adding -0.0
// (x + -0.0 == x for every x), so offsets below zero load as -0.0 and
// offsets past the data load as +0.0 (the reference's zero padding).
#include "common.hpp"

namespace adrt_hip {
namespace fused {

template <typename T>
__host__ __device__ __forceinline__ T neg_zero() {
    return -T(0);
}

// K butterfly stages on 2^K rows held in registers.
//   v[j][i]: level-input row j, slot i <-> offset base - H + i (H = 2^K - 1)
//            after row j's skew.
//   o[t][i]: local angle t at offset base + i.
// Slot i of a stage output is only valid for i >= (accumulated shift); the
// final slots H.. are always valid because the total shift of row j is <= j.
template <int K, int W, typename T>
__host__ __device__ __forceinline__ void small_adrt(const T (&v)[1 << K][W + (1 << K) - 1],
                                           T (&o)[1 << K][W]) {
    constexpr int R0 = 1 << K;
    constexpr int H = R0 - 1;
    constexpr int S = W + H;
    T a[R0][S];
#pragma unroll
    for (int j = 0; j < R0; ++j)
#pragma unroll
        for (int i = 0; i < S; ++i) a[j][i] = v[j][i];
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int A = 1 << u;    // angles per block before this stage
        const int R = R0 >> u;   // blocks before this stage
        T b[R0][S];
#pragma unroll
        for (int r = 0; r < R / 2; ++r) {
#pragma unroll
            for (int t = 0; t < 2 * A; ++t) {
                const int sh = (t + 1) >> 1;
                const int xa = (2 * r) * A + (t >> 1);
                const int ya = (2 * r + 1) * A + (t >> 1);
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    if (i >= sh)
                        b[r * 2 * A + t][i] = a[xa][i] + a[ya][i - sh];
                    else
                        b[r * 2 * A + t][i] = a[xa][i];  // invalid slot, never read
                }
            }
        }
#pragma unroll
        for (int j = 0; j < R0; ++j)
#pragma unroll
            for (int i = 0; i < S; ++i) a[j][i] = b[j][i];
    }
#pragma unroll
    for (int t = 0; t < R0; ++t)
#pragma unroll
        for (int i = 0; i < W; ++i) o[t][i] = a[t][H + i];
}

// Quadrant view of the zero-padded init buffer (adrt_cdefs_adrt.hpp:123-186):
//   q0: img[r][n-1-c]  q1: img[n-1-c][r]  q2: img[c][r]  q3: img[n-1-r][n-1-c]
// c < 0 -> -0.0 (reproduces the reference's copies), c >= n -> +0.0.
template <typename T>
__host__ __device__ __forceinline__ T quad_at(const T *__restrict__ img, int q, int n, int r, int c) {
    if (c < 0) return neg_zero<T>();
    if (c >= n) return T(0);
    switch (q) {
        case 0: return img[(size_t)r * n + (n - 1 - c)];
        case 1: return img[(size_t)(n - 1 - c) * n + r];
        case 2: return img[(size_t)c * n + r];
        default: return img[(size_t)(n - 1 - r) * n + (n - 1 - c)];
    }
}

// LDS rows are stored lane-interleaved: column c of a row of length LR (a
// multiple of W) lives at (c % W) * (LR / W) + c / W, so the W-strided
// columns read by consecutive lanes fall in consecutive banks.
template <int W>
__host__ __device__ __forceinline__ int lds_col(int c, int LR) {
    return (c % W) * (LR / W) + c / W;
}

constexpr int kThreads = 256;

// Whole windows of width W needed below a chunk start to cover `halo` offsets.
template <int W>
__host__ __device__ constexpr int pre_windows(int halo) {
    return (halo + W - 1) / W;
}

template <int R, int G, int W>
__host__ __device__ constexpr int pre_windows_total() {
    int t = 0;
    for (int g = 0; g < G; ++g) t += pre_windows<W>(R * g);
    return t;
}

// ---------------------------------------------------------------------------
// Pass 1: rows J*2^k1 .. +2^k1 of quadrant q, offsets [c0, c0+D) of X.
// Level A: sub-blocks g of 2^KA rows -> 2^KA angles; level B: 2^KB sub-blocks
// at angle mA -> angles s = mA*2^KB + tB.  The work of one block is split in
// phases separated by a block barrier; the phases are plain functions of
// (block, thread) so they also run on the host for checking.
// ---------------------------------------------------------------------------
template <typename T, int KA, int KB, int W, int D>
struct Pass1 {
    static constexpr int k1 = KA + KB;
    static constexpr int RA = 1 << KA, RB = 1 << KB;
    static constexpr int HA = RA - 1, HB = RB - 1;
    static constexpr int LR = D + (1 << k1);  // LDS row: offsets [c0 - 2^k1, c0 + D)
    static constexpr int wins_base = D / W;
    static constexpr size_t lds_elems = KB == 0 ? 0 : size_t(1 << k1) * LR;

    struct Item {
        const T *img;
        T *X;
        int q, n, c0, row0, W1, P1;
        bool valid;
    };

    __host__ __device__ static int num_chunks(int n) { return (n + (1 << k1) - 1 + D - 1) / D; }

    __host__ __device__ static Item item(const T *img_base, T *X_base, int n, int first_image,
                                         int nimg, int block) {
        Item it;
        it.n = n;
        it.W1 = n + (1 << k1) - 1;  // valid offsets of X
        it.P1 = n + (1 << k1);      // row pitch of X
        const int nJ = n >> k1;
        const int nchunk = num_chunks(n);
        const int chunk = block % nchunk;
        block /= nchunk;
        const int J = block % nJ;
        block /= nJ;
        it.q = block & 3;
        const int b = first_image + (block >> 2);
        it.valid = b < first_image + nimg;
        it.c0 = chunk * D;
        it.img = img_base + (size_t)b * n * n;
        it.X = X_base + ((size_t)(b - first_image) * 4 + it.q) * nJ * ((size_t)1 << k1) * it.P1 +
               (size_t)J * ((size_t)1 << k1) * it.P1;
        it.row0 = J << k1;
        return it;
    }

    // Level A (or the only level when KB == 0).
    __host__ __device__ static void phase0(const Item &it, T *lds, int tid) {
        if constexpr (KB == 0) {
            for (int task = tid; task < wins_base; task += kThreads) {
                const int w0 = it.c0 + task * W;
                T v[RA][W + HA];
#pragma unroll
                for (int j = 0; j < RA; ++j)
#pragma unroll
                    for (int i = 0; i < W + HA; ++i)
                        v[j][i] = quad_at(it.img, it.q, it.n, it.row0 + j, w0 - HA + i);
                T o[RA][W];
                small_adrt<KA, W, T>(v, o);
#pragma unroll
                for (int t = 0; t < RA; ++t)
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        if (w0 + i < it.W1) it.X[(size_t)t * it.P1 + w0 + i] = o[t][i];
            }
        } else {
            // Sub-block g needs offsets [c0 - RA*g, c0 + D), in whole windows.
            constexpr int tasks_a = RB * wins_base + pre_windows_total<RA, RB, W>();
            for (int task = tid; task < tasks_a; task += kThreads) {
                int g = 0, rem = task;
#pragma unroll
                for (int gg = 0; gg < RB; ++gg) {
                    const int cnt = wins_base + pre_windows<W>(RA * gg);
                    if (g == gg && rem >= cnt) {
                        rem -= cnt;
                        g = gg + 1;
                    }
                }
                const int w0 = it.c0 - pre_windows<W>(RA * g) * W + rem * W;
                T v[RA][W + HA];
#pragma unroll
                for (int j = 0; j < RA; ++j)
#pragma unroll
                    for (int i = 0; i < W + HA; ++i)
                        v[j][i] = quad_at(it.img, it.q, it.n, it.row0 + g * RA + j, w0 - HA + i);
                T o[RA][W];
                small_adrt<KA, W, T>(v, o);
                const int cbase = w0 - (it.c0 - (1 << k1));
#pragma unroll
                for (int t = 0; t < RA; ++t)
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        lds[(g * RA + t) * LR + lds_col<W>(cbase + i, LR)] = o[t][i];
            }
        }
    }

    // Level B: angle mA, windows in [c0, c0 + D); sub-block g skewed by mA*g.
    __host__ __device__ static void phase1(const Item &it, const T *lds, int tid) {
        if constexpr (KB != 0) {
            for (int task = tid; task < RA * wins_base; task += kThreads) {
                const int mA = task / wins_base;
                const int w0 = it.c0 + (task % wins_base) * W;
                T v[RB][W + HB];
#pragma unroll
                for (int g = 0; g < RB; ++g) {
                    const T *row = lds + (g * RA + mA) * LR;
                    const int cb = w0 - HB - mA * g - (it.c0 - (1 << k1));  // >= RA > 0
#pragma unroll
                    for (int i = 0; i < W + HB; ++i) v[g][i] = row[lds_col<W>(cb + i, LR)];
                }
                T o[RB][W];
                small_adrt<KB, W, T>(v, o);
#pragma unroll
                for (int t = 0; t < RB; ++t) {
                    T *xr = it.X + (size_t)(mA * RB + t) * it.P1;
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        if (w0 + i < it.W1) xr[w0 + i] = o[t][i];
                }
            }
        }
    }
};

#ifndef ADRT_HOST_EMULATION
template <typename T, int KA, int KB, int W, int D>
__global__ void __launch_bounds__(kThreads)
    fwd_pass1(const T *__restrict__ img_base, T *__restrict__ X_base, int n, int first_image,
              int nimg) {
    using P = Pass1<T, KA, KB, W, D>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const typename P::Item it = P::item(img_base, X_base, n, first_image, nimg, blockIdx.x);
    if (!it.valid) return;
    P::phase0(it, lds, threadIdx.x);
    if constexpr (KB != 0) {
        __syncthreads();
        P::phase1(it, lds, threadIdx.x);
    }
}
#endif  // ADRT_HOST_EMULATION

// ---------------------------------------------------------------------------
// Pass 2: angle s of quadrant q, output offsets [d0, d0 + D).
// Level C: groups g of 2^KC blocks J = g*2^KC + j at angle s (row j skewed by
// s*j) -> angles tC; level D: 2^KD groups at angle m = s*2^KC + tC (group g
// skewed by m*g) -> final angles s*2^k2 + tC*2^KD + tD.
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T x_at(const T *__restrict__ xrow, int c, int W1) {
    if (c < 0) return neg_zero<T>();
    if (c >= W1) return T(0);
    return xrow[c];
}

template <typename T, int KC, int KD, int W, int D>
struct Pass2 {
    static constexpr int k2 = KC + KD;
    static constexpr int RC = 1 << KC, RD = 1 << KD;
    static constexpr int HC = RC - 1, HD = RD - 1;
    static constexpr int LR = D + (1 << k2);
    static constexpr int LP = LR + 8;  // row pitch: spreads level-D lanes over banks
    static constexpr int wins_base = D / W;
    static constexpr size_t lds_elems = KD == 0 ? 0 : size_t(1 << k2) * LP;

    struct Item {
        const T *X;
        T *out;
        int n, s, d0, C, W1, abase;
        size_t xrow_stride;
        bool valid;
    };

    __host__ __device__ static int num_chunks(int n) { return (2 * n - 1 + D - 1) / D; }

    __host__ __device__ static Item item(const T *X_base, T *out_base, int n, int k1,
                                         int first_image, int nimg, int block) {
        Item it;
        it.n = n;
        it.C = 2 * n - 1;
        it.W1 = n + (1 << k1) - 1;
        const int P1 = n + (1 << k1);
        const int nS = 1 << k1;
        const int nchunk = num_chunks(n);
        const int chunk = block % nchunk;
        block /= nchunk;
        it.s = block % nS;
        block /= nS;
        const int q = block & 3;
        const int b = first_image + (block >> 2);
        it.valid = b < first_image + nimg;
        it.d0 = chunk * D;
        it.xrow_stride = (size_t)nS * P1;  // between blocks J at fixed s
        it.X = X_base + ((size_t)(b - first_image) * 4 + q) * ((size_t)1 << k2) * it.xrow_stride +
               (size_t)it.s * P1;
        it.out = out_base + ((size_t)b * 4 + q) * it.C * (size_t)n;
        it.abase = it.s << k2;
        return it;
    }

    // Level C (or the only level when KD == 0).
    __host__ __device__ static void phase0(const Item &it, T *lds, int tid) {
        const int s = it.s;
        if constexpr (KD == 0) {
            for (int task = tid; task < wins_base; task += kThreads) {
                const int w0 = it.d0 + task * W;
                T v[RC][W + HC];
#pragma unroll
                for (int j = 0; j < RC; ++j) {
                    const T *xr = it.X + (size_t)j * it.xrow_stride;
#pragma unroll
                    for (int i = 0; i < W + HC; ++i) v[j][i] = x_at(xr, w0 - HC + i - s * j, it.W1);
                }
                T o[RC][W];
                small_adrt<KC, W, T>(v, o);
#pragma unroll
                for (int i = 0; i < W; ++i)
                    if (w0 + i < it.C) {
                        T *orow = it.out + (size_t)(w0 + i) * it.n + it.abase;
#pragma unroll
                        for (int t = 0; t < RC; ++t) orow[t] = o[t][i];
                    }
            }
        } else {
            // Group g covers [Sg, d0 + D - s*RC*g), Sg = d0 - s*RC*g - pre(RC*g)*W.
            constexpr int tasks_c = RD * wins_base + pre_windows_total<RC, RD, W>();
            for (int task = tid; task < tasks_c; task += kThreads) {
                int g = 0, rem = task;
#pragma unroll
                for (int gg = 0; gg < RD; ++gg) {
                    const int cnt = wins_base + pre_windows<W>(RC * gg);
                    if (g == gg && rem >= cnt) {
                        rem -= cnt;
                        g = gg + 1;
                    }
                }
                const int Sg = it.d0 - s * RC * g - pre_windows<W>(RC * g) * W;
                const int w0 = Sg + rem * W;
                T v[RC][W + HC];
#pragma unroll
                for (int j = 0; j < RC; ++j) {
                    const T *xr = it.X + (size_t)(g * RC + j) * it.xrow_stride;
#pragma unroll
                    for (int i = 0; i < W + HC; ++i) v[j][i] = x_at(xr, w0 - HC + i - s * j, it.W1);
                }
                T o[RC][W];
                small_adrt<KC, W, T>(v, o);
#pragma unroll
                for (int t = 0; t < RC; ++t)
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        lds[(g * RC + t) * LP + lds_col<W>(rem * W + i, LR)] = o[t][i];
            }
        }
    }

    // Level D: m = s*RC + tC; group g read at offset w0 + i - HD - m*g.
    __host__ __device__ static void phase1(const Item &it, const T *lds, int tid) {
        if constexpr (KD != 0) {
            for (int task = tid; task < RC * wins_base; task += kThreads) {
                const int tC = task % RC;
                const int w0 = it.d0 + (task / RC) * W;
                T v[RD][W + HD];
#pragma unroll
                for (int g = 0; g < RD; ++g) {
                    const T *row = lds + (g * RC + tC) * LP;
                    const int cb = (w0 - it.d0) - HD - tC * g + pre_windows<W>(RC * g) * W;
#pragma unroll
                    for (int i = 0; i < W + HD; ++i) {
                        int col = cb + i;
                        col = col < 0 ? 0 : col;  // clamped slots are never used
                        v[g][i] = row[lds_col<W>(col, LR)];
                    }
                }
                T o[RD][W];
                small_adrt<KD, W, T>(v, o);
                const int a0 = it.abase + tC * RD;
#pragma unroll
                for (int i = 0; i < W; ++i)
                    if (w0 + i < it.C) {
                        T *orow = it.out + (size_t)(w0 + i) * it.n + a0;
#pragma unroll
                        for (int t = 0; t < RD; ++t) orow[t] = o[t][i];
                    }
            }
        }
    }
};

#ifndef ADRT_HOST_EMULATION
template <typename T, int KC, int KD, int W, int D>
__global__ void __launch_bounds__(kThreads)
    fwd_pass2(const T *__restrict__ X_base, T *__restrict__ out_base, int n, int k1,
              int first_image, int nimg) {
    using P = Pass2<T, KC, KD, W, D>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const typename P::Item it = P::item(X_base, out_base, n, k1, first_image, nimg, blockIdx.x);
    if (!it.valid) return;
    P::phase0(it, lds, threadIdx.x);
    if constexpr (KD != 0) {
        __syncthreads();
        P::phase1(it, lds, threadIdx.x);
    }
}
#endif  // ADRT_HOST_EMULATION

}  // namespace fused
}  // namespace adrt_hip

namespace adrt_hip {
namespace fused {

// Launch geometry per element type: offsets per thread (W) and per work
// item (D1 for pass 1, D2 for pass 2).
template <typename T>
struct Geo;
template <>
struct Geo<float> {
    static constexpr int W = 8, D1 = 256, D2 = 256;
};
template <>
struct Geo<double> {
    static constexpr int W = 4, D1 = 128, D2 = 128;
};

template <typename T, int KA, int KB>
using P1Of = Pass1<T, KA, KB, Geo<T>::W, Geo<T>::D1>;
template <typename T, int KC, int KD>
using P2Of = Pass2<T, KC, KD, Geo<T>::W, Geo<T>::D2>;

#ifndef ADRT_HOST_EMULATION
template <typename T, int KA, int KB>
static void launch_pass1(const T *img, T *X, int n, int first, int count, hipStream_t s) {
    using P = P1Of<T, KA, KB>;
    const unsigned grid = unsigned(count) * 4u * unsigned(n >> P::k1) * unsigned(P::num_chunks(n));
    constexpr size_t lds = P::lds_elems * sizeof(T);
    auto kern = fwd_pass1<T, KA, KB, Geo<T>::W, Geo<T>::D1>;
    static bool attr_set = false;
    if (lds > 0 && !attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds, s, img, X, n, first, count);
}

template <typename T, int KC, int KD>
static void launch_pass2(const T *X, T *out, int n, int k1, int first, int count,
                         hipStream_t s) {
    using P = P2Of<T, KC, KD>;
    const unsigned grid = unsigned(count) * 4u * unsigned(1 << k1) * unsigned(P::num_chunks(n));
    constexpr size_t lds = P::lds_elems * sizeof(T);
    auto kern = fwd_pass2<T, KC, KD, Geo<T>::W, Geo<T>::D2>;
    static bool attr_set = false;
    if (lds > 0 && !attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), lds, s, X, out, n, k1, first, count);
}

// Split of L = log2 n stages: pass 1 takes k1 = ceil(L/2), pass 2 the rest;
// each pass is a 3-stage level plus a level of k-3 stages.
template <typename T>
static void run_group(const T *img, T *X, T *out, int n, int L, int first, int count,
                      hipStream_t s) {
    switch (L) {
        case 6:
            launch_pass1<T, 3, 0>(img, X, n, first, count, s);
            launch_pass2<T, 3, 0>(X, out, n, 3, first, count, s);
            break;
        case 7:
            launch_pass1<T, 3, 1>(img, X, n, first, count, s);
            launch_pass2<T, 3, 0>(X, out, n, 4, first, count, s);
            break;
        case 8:
            launch_pass1<T, 3, 1>(img, X, n, first, count, s);
            launch_pass2<T, 3, 1>(X, out, n, 4, first, count, s);
            break;
        case 9:
            launch_pass1<T, 3, 2>(img, X, n, first, count, s);
            launch_pass2<T, 3, 1>(X, out, n, 5, first, count, s);
            break;
        case 10:
            launch_pass1<T, 3, 2>(img, X, n, first, count, s);
            launch_pass2<T, 3, 2>(X, out, n, 5, first, count, s);
            break;
        case 11:
            launch_pass1<T, 3, 3>(img, X, n, first, count, s);
            launch_pass2<T, 3, 2>(X, out, n, 6, first, count, s);
            break;
        default:  // 12
            launch_pass1<T, 3, 3>(img, X, n, first, count, s);
            launch_pass2<T, 3, 3>(X, out, n, 6, first, count, s);
            break;
    }
}

#endif  // ADRT_HOST_EMULATION

inline int pass1_stages(int L) { return (L + 1) / 2; }

// Images per group: keep one group's intermediate at or below ~64 MiB so it
// is written and re-read while resident in the 256 MB Infinity Cache.
inline size_t group_images(size_t B, size_t per_image_bytes) {
    const size_t budget = size_t(64) << 20;
    size_t g = per_image_bytes >= budget ? 1 : budget / per_image_bytes;
    return g < B ? g : B;
}

inline size_t x_bytes_per_image(int n, int L, size_t elem) {
    const int k1 = pass1_stages(L);
    return size_t(4) * n * size_t(n + (1 << k1)) * elem;
}

}  // namespace fused

#ifndef ADRT_HOST_EMULATION

bool fused_supported(size_t n) {
    const int L = num_iters(n);
    return is_pow2(n) && L >= 6 && L <= 12;
}

size_t fused_workspace(int dtype, size_t B, size_t n) {
    const size_t elem = dtype == ADRT_HIP_F64 ? 8 : 4;
    const int L = num_iters(n);
    const size_t per = fused::x_bytes_per_image(int(n), L, elem);
    return fused::group_images(B, per) * per;
}

template <typename T>
static int run_fused_t(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const int L = num_iters(n);
    const size_t per = fused::x_bytes_per_image(int(n), L, sizeof(T));
    const size_t G = fused::group_images(B, per);
    for (size_t first = 0; first < B; first += G) {
        const size_t count = (B - first) < G ? (B - first) : G;
        fused::run_group<T>(in, ws, out, int(n), L, int(first), int(count), s);
    }
    return launch_status();
}

int run_fused(const void *in, void *out, void *ws, int dtype, size_t B, size_t n, hipStream_t s) {
    if (dtype == ADRT_HIP_F32)
        return run_fused_t(static_cast<const float *>(in), static_cast<float *>(out),
                           static_cast<float *>(ws), B, n, s);
    return run_fused_t(static_cast<const double *>(in), static_cast<double *>(out),
                       static_cast<double *>(ws), B, n, s);
}

#endif  // ADRT_HOST_EMULATION

}  // namespace adrt_hip
