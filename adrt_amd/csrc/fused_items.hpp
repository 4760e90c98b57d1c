// Work items of the fused two-pass forward ADRT and back-projection for
// gfx950 (log2 n in [6, 12]): the per-item load / compute / store code of
// fwd_pass1, fwd_pass2, bdrt_passA and bdrt_passB, the stage splits and
// the intermediate sizes.  fused.hip launches them one pass per kernel; the
// persistent queue kernel (queue.hip) runs them from one work queue.
//
// Structure.  Stage i of the reference (adrt_core, adrt_cdefs_adrt.hpp:55-96)
// merges row blocks (2r, 2r+1) and maps angle a to a/2 with the odd block
// shifted by ceil(a/2).  Unrolling the recursion splits the L = log2 n stages
// into two passes with one intermediate X in HBM:
//
//   pass 1: for every row block J of 2^k1 rows, the first k1 stages
//           -> X[J][s][c] for the 2^k1 stage-k1 angles s
//   pass 2: for every angle s, the last k2 = L - k1 stages over the 2^k2
//           blocks X[J][s][.], block J pre-shifted by s*J; the pass writes
//           the final angles s*2^k2 + t directly in the output layout.
//
// Each pass is two "levels": a 3-stage level and a level of k-3 stages.  A
// level runs on a strip: one wavefront, lane l holding 4 consecutive offsets
// 4l..4l+3 of 2^K rows in registers.  Its K stages use compile-time shifts
// ceil(t/2) of the local angle t; operands that lie in the previous lane are
// read with DPP wave_shr:1, which the compiler folds into v_add_f32_dpp, so a
// stage is exactly one add per element with no halo loads.  The level input
// angle m contributes m*j to the shift of local row j and is applied as a
// skew when row j is loaded.  Offsets of a strip below 2^K - 1 have no valid
// halo and are discarded; each strip delivers offsets [64, 256) (lanes 16+),
// which leaves room for the skew spread of the next level's input.  Levels
// exchange through LDS.
//
// Exactness.  Every output element is the same sequence of single IEEE adds
// of the same operands as in the reference, so results are bit-identical.
// The reference's plain copy for c < ceil(a/2) is reproduced by adding -0.0
// (x + -0.0 == x for every x): offsets below zero load as -0.0, offsets past
// the data load as +0.0 (the reference's zero padding).
//
// Host emulation.  The per-lane load/store functions are __host__ __device__;
// with ADRT_HOST_EMULATION defined only those and the whole-strip compute are
// compiled, so tests/host_emulation can run every block on the CPU.
#pragma once

#include "common.hpp"

namespace adrt_hip {
namespace fused {

constexpr int kW = 4;                  // offsets per lane
constexpr int kLanes = 64;             // lanes per strip (one wavefront)
constexpr int kStrip = kW * kLanes;    // offsets per strip
constexpr int kOut = 192;              // offsets a strip delivers: [kLead, kStrip)
constexpr int kLead = kStrip - kOut;   // 64
constexpr int kWaves = 8;              // wavefronts per work-group
constexpr int kThreads = kWaves * kLanes;
constexpr int kPad = 64;               // LDS row front padding (columns)
constexpr int kRow = kStrip + kPad;    // LDS row length (columns)
constexpr int kSlice = kRow / kW;      // columns per interleave slice

template <typename T>
__host__ __device__ __forceinline__ T neg_zero() {
    return -T(0);
}

// An LDS row holds strip columns c in [-kPad, kStrip) at position
// (c' & 3) * kSlice + (c' >> 2), c' = c + kPad: interleaved by residue so
// that column 4*lane + u, read by every lane of a wave for one wave-uniform
// u, is bank-consecutive.  Its address is lane + lds_uniform(u).
__host__ __device__ __forceinline__ int lds_uniform(int u) {
    const int cp = u + kPad;
    return (cp & (kW - 1)) * kSlice + (cp >> 2);
}

// Wave-uniform value in a scalar register (identity on the host).
__host__ __device__ __forceinline__ int wave_uniform(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// Work-group -> item.  Work-groups are dispatched to the 8 XCDs
// round-robin, so work-group w runs on XCD w % 8.  Each XCD owns a
// contiguous eighth of the items, so the work-groups resident on one XCD at
// a time work on neighbouring items, which share input rows in its L2.  The
// mapping only affects speed: with grid = 8 * ceil(items / 8) every item is
// visited exactly once (and a smaller multiple of 8 makes each work-group
// step through several).
constexpr int kXcds = 8;
struct Schedule {
    int idx, hi, step;
};
__host__ __device__ __forceinline__ Schedule schedule(int block, int grid, int items) {
    const int per = (items + kXcds - 1) / kXcds;
    const int lo = (block % kXcds) * per;
    Schedule sc;
    sc.idx = lo + block / kXcds;
    sc.hi = items < lo + per ? items : lo + per;
    sc.step = grid / kXcds;
    return sc;
}

// Quadrant views of the zero-padded init buffer (adrt_cdefs_adrt.hpp:123-186),
// row r and offset c of quadrant q:
//   q0: img[r][n-1-c]  q1: img[n-1-c][r]  q2: img[c][r]  q3: img[n-1-r][n-1-c]
// c < 0 reads -0.0 (reproduces the reference's copies), c >= n reads +0.0.

// 16-byte (fp32) / 32-byte (fp64) vector moves through a native vector type
// so the temporaries stay in registers; p needs only element alignment.
#ifndef ADRT_HOST_EMULATION
template <typename T>
using vec4_t = T __attribute__((ext_vector_type(4)));
#else
template <typename T>
struct vec4_t {
    T x, y, z, w;
};
#endif

template <typename T>
__host__ __device__ __forceinline__ void load_vec4(const T *p, T &a, T &b, T &c, T &d) {
    vec4_t<T> t;
    __builtin_memcpy(&t, p, sizeof(t));
    a = t.x;
    b = t.y;
    c = t.z;
    d = t.w;
}

template <typename T>
__host__ __device__ __forceinline__ void store_vec4(T *p, T a, T b, T c, T d) {
    vec4_t<T> t;
    t.x = a;
    t.y = b;
    t.z = c;
    t.w = d;
    __builtin_memcpy(p, &t, sizeof(t));
}

// Streaming store for data no kernel of this call reads again (the final
// output): it is not kept in L2, which leaves L2 to the intermediate.
template <typename T>
__host__ __device__ __forceinline__ void store_vec4_stream(T *p, T a, T b, T c, T d) {
#ifndef ADRT_HOST_EMULATION
    vec4_t<T> t;
    t.x = a;
    t.y = b;
    t.z = c;
    t.w = d;
    __builtin_nontemporal_store(t, reinterpret_cast<vec4_t<T> *>(p));
#else
    store_vec4(p, a, b, c, d);
#endif
}

// Padded windows.  A strip row needs four consecutive offsets of a row that
// has `lim` (>= 4) valid elements and reads -0.0 below offset 0 and +0.0 from
// `lim` on.  Every window is read with one unconditional 16-byte load from a
// start clamped into the row; windows that cross an edge are then patched in
// a separate step.  Keeping the loads free of divergent branches lets all
// rows of a strip be in flight at once (a branch around a load would make
// the compiler wait for it before the paths join).
__host__ __device__ __forceinline__ int clamp_start(int c, int lim) {
    return c < 0 ? 0 : (c > lim - 4 ? lim - 4 : c);
}

// t_k for k in [0, 3], `below` for k < 0, `above` for k > 3: a chain of
// selects on four scalars (an array here would be indexed through scratch).
template <typename T>
__host__ __device__ __forceinline__ T pick4(T t0, T t1, T t2, T t3, int k, T below, T above) {
    T v = k < 0 ? below : t0;
    v = k == 1 ? t1 : v;
    v = k == 2 ? t2 : v;
    v = k == 3 ? t3 : v;
    return k > 3 ? above : v;
}

// Reversed window: out[i] = row[lim - 1 - (c + i)], i.e. offsets c..c+3 of a
// row read backwards; offsets below 0 read -0.0, offsets from lim on +0.0.
template <typename T>
__host__ __device__ __forceinline__ void load_window_rev(const T *__restrict__ row, int c, int lim,
                                                         T (&out)[4]) {
    load_vec4(row + unsigned(clamp_start(lim - 4 - c, lim)), out[3], out[2], out[1], out[0]);
}

template <typename T>
__host__ __device__ __forceinline__ void patch_window_rev(int c, int lim, T (&v)[4]) {
    if (c < 0 || c > lim - 4) {
        const T t0 = v[3], t1 = v[2], t2 = v[1], t3 = v[0];  // ascending columns
        const int e = (lim - 4 - c) - clamp_start(lim - 4 - c, lim);
        // offset c + i is column lim - 1 - c - i = clamped start + e + 3 - i
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = pick4(t0, t1, t2, t3, e + 3 - i, T(0), neg_zero<T>());
    }
}

// ---------------------------------------------------------------------------
// Strip compute: K stages on R = 2^K rows, a[j][i] = row j at strip offset
// 4*lane + i -> a[t][i] = local angle t.  Valid for strip offsets >= R - 1.
// ---------------------------------------------------------------------------
#ifndef ADRT_HOST_EMULATION
template <typename T>
__device__ __forceinline__ T lane_shr1(T v);

template <>
__device__ __forceinline__ float lane_shr1<float>(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}

template <>
__device__ __forceinline__ double lane_shr1<double>(double v) {
    const long long bits = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(bits & 0xffffffff), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(bits >> 32), 0x138, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) |
                                                  (unsigned long long)(unsigned)lo));
}

template <int K, typename T>
__device__ __forceinline__ void strip_adrt(T (&a)[1 << K][kW]) {
    constexpr int R0 = 1 << K;
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int A = 1 << u;   // angles per block before this stage
        const int R = R0 >> u;  // blocks before this stage
        T b[R0][kW];
#pragma unroll
        for (int r = 0; r < R / 2; ++r) {
#pragma unroll
            for (int t = 0; t < 2 * A; ++t) {
                const int sh = (t + 1) >> 1;  // <= kW for K <= 3
                const int xa = (2 * r) * A + (t >> 1);
                const int ya = (2 * r + 1) * A + (t >> 1);
#pragma unroll
                for (int i = 0; i < kW; ++i) {
                    const T y = (i >= sh) ? a[ya][i - sh] : lane_shr1(a[ya][kW + i - sh]);
                    b[r * 2 * A + t][i] = a[xa][i] + y;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < R0; ++j)
#pragma unroll
            for (int i = 0; i < kW; ++i) a[j][i] = b[j][i];
    }
}
#endif

// Whole-strip restatement for host emulation: a[lane][row][i]; offsets with
// no valid halo are poisoned with NaN.
template <int K, typename T>
inline void strip_adrt_host(T (*a)[1 << K][kW]) {
    constexpr int R0 = 1 << K;
    static T cur[R0][kStrip], nxt[R0][kStrip];
    for (int l = 0; l < kLanes; ++l)
        for (int j = 0; j < R0; ++j)
            for (int i = 0; i < kW; ++i) cur[j][l * kW + i] = a[l][j][i];
    for (int u = 0; u < K; ++u) {
        const int A = 1 << u, R = R0 >> u;
        for (int r = 0; r < R / 2; ++r)
            for (int t = 0; t < 2 * A; ++t) {
                const int sh = (t + 1) >> 1;
                const int xa = (2 * r) * A + (t >> 1), ya = (2 * r + 1) * A + (t >> 1);
                for (int o = 0; o < kStrip; ++o)
                    nxt[r * 2 * A + t][o] = cur[xa][o] + (o >= sh ? cur[ya][o - sh] : T(NAN));
            }
        for (int j = 0; j < R0; ++j)
            for (int o = 0; o < kStrip; ++o) cur[j][o] = nxt[j][o];
    }
    for (int l = 0; l < kLanes; ++l)
        for (int j = 0; j < R0; ++j)
            for (int i = 0; i < kW; ++i) a[l][j][i] = cur[j][l * kW + i];
}

// ---------------------------------------------------------------------------
// Back-projection strip: K bdrt stages (adrt_cdefs_bdrt.hpp:55-116) on
// R = 2^K indices held as a[index * S + section][i] (S sections per index,
// strip offset 4*lane + i).  Stage u pairs indices (2j, 2j+1) into index j
// with sections 2*sec (left: x + y) and 2*sec + 1 (right: x at offset +j plus
// y at offset +j+1); the level's common shift (the pair index of the
// enclosing group) is applied outside, as a skew.  Operands above the lane
// come from the next lane (DPP wave_shl:1).  After K stages a[s] holds
// section s.  Valid for strip offsets below 256 - (2^K - 1).
// ---------------------------------------------------------------------------
#ifndef ADRT_HOST_EMULATION
template <typename T>
__device__ __forceinline__ T lane_shl1(T v);

template <>
__device__ __forceinline__ float lane_shl1<float>(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

template <>
__device__ __forceinline__ double lane_shl1<double>(double v) {
    const long long bits = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(bits & 0xffffffff), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(bits >> 32), 0x130, 0xf, 0xf, false);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) |
                                                  (unsigned long long)(unsigned)lo));
}

template <int K, typename T>
__device__ __forceinline__ void strip_bdrt(T (&a)[1 << K][kW]) {
    constexpr int R0 = 1 << K;
#pragma unroll
    for (int u = 0; u < K; ++u) {
        const int S = 1 << u;   // sections per index before this stage
        const int I = R0 >> u;  // indices before this stage
        T b[R0][kW];
#pragma unroll
        for (int j = 0; j < I / 2; ++j) {
#pragma unroll
            for (int sec = 0; sec < S; ++sec) {
                const int xa = (2 * j) * S + sec, ya = (2 * j + 1) * S + sec;
#pragma unroll
                for (int i = 0; i < kW; ++i) b[j * 2 * S + 2 * sec][i] = a[xa][i] + a[ya][i];
#pragma unroll
                for (int i = 0; i < kW; ++i) {
                    const T x = (i + j < kW) ? a[xa][i + j] : lane_shl1(a[xa][i + j - kW]);
                    const T y = (i + j + 1 < kW) ? a[ya][i + j + 1]
                                                 : lane_shl1(a[ya][i + j + 1 - kW]);
                    b[j * 2 * S + 2 * sec + 1][i] = x + y;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R0; ++r)
#pragma unroll
            for (int i = 0; i < kW; ++i) a[r][i] = b[r][i];
    }
}
#endif

// Whole-strip restatement for host emulation; offsets whose operands fall
// past the strip are poisoned with NaN.
template <int K, typename T>
inline void strip_bdrt_host(T (*a)[1 << K][kW]) {
    constexpr int R0 = 1 << K;
    static T cur[R0][kStrip], nxt[R0][kStrip];
    for (int l = 0; l < kLanes; ++l)
        for (int r = 0; r < R0; ++r)
            for (int i = 0; i < kW; ++i) cur[r][l * kW + i] = a[l][r][i];
    for (int u = 0; u < K; ++u) {
        const int S = 1 << u, I = R0 >> u;
        for (int j = 0; j < I / 2; ++j)
            for (int sec = 0; sec < S; ++sec) {
                const int xa = (2 * j) * S + sec, ya = (2 * j + 1) * S + sec;
                for (int o = 0; o < kStrip; ++o) {
                    nxt[j * 2 * S + 2 * sec][o] = cur[xa][o] + cur[ya][o];
                    nxt[j * 2 * S + 2 * sec + 1][o] =
                        o + j + 1 < kStrip ? cur[xa][o + j] + cur[ya][o + j + 1] : T(NAN);
                }
            }
        for (int r = 0; r < R0; ++r)
            for (int o = 0; o < kStrip; ++o) cur[r][o] = nxt[r][o];
    }
    for (int l = 0; l < kLanes; ++l)
        for (int r = 0; r < R0; ++r)
            for (int i = 0; i < kW; ++i) a[l][r][i] = cur[r][l * kW + i];
}

// Intermediate X rows (element offsets relative to the row's offset 0):
//   [-4, 0)         -0.0
//   [0, W1)         the W1 = n + 2^k1 - 1 valid offsets
//   [W1, W1 + 4)    +0.0 (and +0.0 wherever pass 1's last chunk runs past
//                   W1: those values come from zero padding only)
// A window of 4 offsets whose start is clamped into [-4, W1] therefore reads
// exactly the padded row (-0.0 below 0, +0.0 from W1 on) with one
// unconditional load.  Pass 1 writes the margins.  The left margin (kXL
// elements) and the pitch are whole 128-byte lines (256 bytes for fp64), so
// pass 1's row segments are line-aligned.
constexpr int kXL = 32;
constexpr int kXPad = 4;  // margin elements read on each side
__host__ __device__ __forceinline__ int x_valid(int n, int k1) { return n + (1 << k1) - 1; }
__host__ __device__ __forceinline__ int x_chunks(int n, int k1) {
    return (x_valid(n, k1) + kOut - 1) / kOut;
}
__host__ __device__ __forceinline__ int x_pitch(int n, int k1) {
    const int w = x_valid(n, k1) + kXPad;
    const int span = x_chunks(n, k1) * kOut;
    return kXL + ((w > span ? w : span) + 31) / 32 * 32;
}

// ---------------------------------------------------------------------------
// Pass 1: rows row0 .. row0 + 2^k1 of quadrant q, X offsets [c0, c0 + kOut).
//   phase 0 (level A): strip g (2^KB of them): rows row0 + 8g + j, base
//     c0 - HB - 7g - 7 -> LDS rows (g, tA)
//   phase 1 (level B): strip mA (8 of them): LDS rows (g, mA) skewed by
//     mA*g, base c0 - kLead -> X[J][mA*2^KB + tB][c] for c in [c0, c0+kOut)
// With KB == 0 phase 0 writes X directly and phase 1 is empty.
// With KB == 3 (64 rows) the LDS rows go in two parts of tA in [0, 4) and
// [4, 8): strips mA = 0..3 load the first, then every strip stores its
// second half over it and strips 4..7 load that (40 KB of LDS in fp32
// instead of 80, so three work-groups fit per CU instead of two).
// ---------------------------------------------------------------------------
template <typename T, int KB>
struct Pass1 {
    static constexpr bool kStaged = false;
    static constexpr int KA = 3;
    static constexpr int k1 = KA + KB;
    static constexpr int RA = 1 << KA, RB = 1 << KB;
    static constexpr int HB = RB - 1;
    static constexpr int R0 = RA, R1 = RB;
    static constexpr int tasks0 = RB;
    static constexpr int tasks1 = KB == 0 ? 0 : RA;
    static constexpr int kParts = KB == 3 ? 2 : 1;     // LDS parts of the level A -> B rows
    static constexpr int kPartRows = RA / kParts;      // rows tA of a strip per part
    static constexpr size_t lds_elems = KB == 0 ? 0 : size_t(1 << k1) / kParts * kRow;
    // waves per SIMD for the register budget: three 8-wave work-groups per CU
    // with parts (80 VGPRs; at 64, four work-groups, it spills 9 registers and
    // measured 69 against 50 us per 2-image group at 2048^2), else the budget
    // the LDS allows
    static constexpr int kSimdWaves = kParts > 1 ? 6 : 8;

#ifndef ADRT_HOST_EMULATION
    __device__ __forceinline__ static void strip0(T (&v)[R0][kW]) { strip_adrt<KA, T>(v); }
    __device__ __forceinline__ static void strip1(T (&v)[R1][kW]) { strip_adrt<KB, T>(v); }
#endif
    static void strip0_host(T (*v)[R0][kW]) { strip_adrt_host<KA, T>(v); }
    static void strip1_host(T (*v)[R1][kW]) { strip_adrt_host<KB, T>(v); }

    struct Item {
        const T *img;
        T *X;  // rows of block J: X + s * P1
        int q, n, c0, row0, W1, P1;
    };

    __host__ __device__ __forceinline__ static int num_chunks(int n) { return x_chunks(n, k1); }
    __host__ __device__ __forceinline__ static int num_items(int n, int nimg) {
        return nimg * 4 * (n >> k1) * num_chunks(n);
    }
    __host__ __device__ __forceinline__ static bool skip(const Item &) { return false; }

    // Item `block` in [0, num_items) -> (image, quadrant, row block J, chunk),
    // ordered so that the two quadrants reading the same image data are
    // neighbours (and so run at the same time in one XCD's L2):
    //   block = (((b * 2 + pair) * nJ + J') * nchunk + chunk') * 2 + m
    //   pair 0: m = 0 -> q0, block J', chunk'; m = 1 -> q3, block nJ-1-J'
    //           (the same image rows, reversed), chunk'
    //   pair 1: m = 0 -> q2, block J', chunk'; m = 1 -> q1, block J', chunk
    //           nchunk-1-chunk' (its offsets run over the image rows backwards)
    __host__ __device__ __forceinline__ static Item item(const T *img_base, T *X_base, int n,
                                                         int first_image, int block) {
        Item it;
        it.n = n;
        it.W1 = x_valid(n, k1);  // valid offsets of X
        it.P1 = x_pitch(n, k1);  // row pitch of X (multiple of 4)
        const int nJ = n >> k1;
        const int nchunk = num_chunks(n);
        const int m = block & 1;
        block >>= 1;
        int chunk = block % nchunk;
        block /= nchunk;
        int J = block % nJ;
        block /= nJ;
        const int pair = block & 1;
        const int b = first_image + (block >> 1);
        if (pair == 0) {
            it.q = m == 0 ? 0 : 3;
            if (m == 1) J = nJ - 1 - J;
        } else {
            it.q = m == 0 ? 2 : 1;
            if (m == 1) chunk = nchunk - 1 - chunk;
        }
        it.c0 = chunk * kOut;
        it.img = img_base + (size_t)b * n * n;
        it.X = X_base + ((size_t)(b - first_image) * 4 + it.q) * nJ * ((size_t)1 << k1) * it.P1 +
               (size_t)J * ((size_t)1 << k1) * it.P1 + kXL;
        it.row0 = J << k1;
        return it;
    }

    __host__ __device__ __forceinline__ static int base0(const Item &it, int g) {
        return KB == 0 ? it.c0 - kLead : it.c0 - HB - (RA - 1) * g - (RA - 1);
    }

    // Level A input: 8 quadrant rows at offsets base0 + 4*lane + i, in two
    // steps: every load is issued first (from start offsets clamped into the
    // image), then the windows that cross an image edge are patched, so all
    // rows' loads are in flight together.
    //
    // q1/q2: rows are image columns r0..r0+7; offset c+i is image row
    // n-1-(c+i) (q1) or c+i (q2), clamped into the image for the load.
    __host__ __device__ __forceinline__ static void issue_cols(const Item &it, int c, int r0,
                                                               T (&v)[RA][kW]) {
        const int n = it.n;
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int cc = c + i;
            int irow = it.q == 1 ? n - 1 - cc : cc;
            irow = irow < 0 ? 0 : (irow > n - 1 ? n - 1 : irow);
            const T *src = it.img + unsigned(irow * n + r0);
            load_vec4(src, v[0][i], v[1][i], v[2][i], v[3][i]);
            load_vec4(src + 4, v[4][i], v[5][i], v[6][i], v[7][i]);
        }
    }
    __host__ __device__ __forceinline__ static void patch_cols(int c, int n, T (&v)[RA][kW]) {
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int cc = c + i;
            if (cc < 0 || cc >= n) {
                const T pad = cc < 0 ? neg_zero<T>() : T(0);
#pragma unroll
                for (int j = 0; j < RA; ++j) v[j][i] = pad;
            }
        }
    }
    // q0/q3: rows are image rows (q0: r, q3: n-1-r); offsets c..c+3 are the
    // columns n-1-c .. n-4-c.
    __host__ __device__ __forceinline__ static void issue_rows(const Item &it, int c, int r0,
                                                               T (&v)[RA][kW]) {
        const int n = it.n;
#pragma unroll
        for (int j = 0; j < RA; ++j) {
            const int irow = it.q == 0 ? r0 + j : n - 1 - (r0 + j);
            load_window_rev(it.img + (size_t)irow * n, c, n, v[j]);
        }
    }
    __host__ __device__ __forceinline__ static void patch_rows(int c, int n, T (&v)[RA][kW]) {
#pragma unroll
        for (int j = 0; j < RA; ++j) patch_window_rev(c, n, v[j]);
    }

    __host__ __device__ __forceinline__ static void load0(const Item &it, int g, int lane,
                                                          T (&v)[RA][kW]) {
        const int c = base0(it, g) + kW * lane;
        if (it.q == 1 || it.q == 2) {
            issue_cols(it, c, it.row0 + g * RA, v);
            patch_cols(c, it.n, v);
        } else {
            issue_rows(it, c, it.row0 + g * RA, v);
            patch_rows(c, it.n, v);
        }
    }

    // X rows mA*R + t at offsets c0 - kLead + 4*lane + i; lanes 16..63 deliver
    // [c0, c0 + kOut) (a last chunk running past W1 delivers +0.0 there,
    // computed from zero padding only).  Lane 0, which delivers nothing,
    // writes the row margins: -0.0 before offset 0 (first chunk) and +0.0
    // from the end of the last chunk up to W1 + 4, if it ends before that.
    template <int R>
    __host__ __device__ __forceinline__ static void store_x(const Item &it, int mA, int lane,
                                                            const T (&o)[R][kW]) {
        const int c = it.c0 - kLead + kW * lane;
        if (lane >= kLead / kW) {
#pragma unroll
            for (int t = 0; t < R; ++t)
                store_vec4(it.X + (size_t)(mA * R + t) * it.P1 + unsigned(c), o[t][0], o[t][1],
                           o[t][2], o[t][3]);
        }
        if (lane == 0) {
            if (it.c0 == 0) {
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    T *row = it.X + (size_t)(mA * R + t) * it.P1;
                    store_vec4(row - kXPad, neg_zero<T>(), neg_zero<T>(), neg_zero<T>(),
                               neg_zero<T>());
                }
            }
            if (it.c0 + kOut > it.W1) {  // the last chunk (W1 is odd, c0 + kOut even)
                for (int z = it.c0 + kOut; z < it.W1 + kXPad; z += 4) {
#pragma unroll
                    for (int t = 0; t < R; ++t)
                        store_vec4(it.X + (size_t)(mA * R + t) * it.P1 + unsigned(z), T(0), T(0),
                                   T(0), T(0));
                }
            }
        }
    }

    // Rows tA of part `part` of strip g: LDS row g * kPartRows + tA % kPartRows.
    __host__ __device__ __forceinline__ static void store0_part(const Item &it, int g, int lane,
                                                                const T (&o)[RA][kW], T *lds,
                                                                int part) {
        (void)it;
#pragma unroll
        for (int u = 0; u < kPartRows; ++u)
#pragma unroll
            for (int i = 0; i < kW; ++i)
                lds[(g * kPartRows + u) * kRow + lds_uniform(i) + lane] = o[part * kPartRows + u][i];
    }

    __host__ __device__ __forceinline__ static void store0(const Item &it, int g, int lane,
                                                           const T (&o)[RA][kW], T *lds) {
        if constexpr (KB == 0) {
            (void)lds;
            store_x<RA>(it, g, lane, o);
        } else {
            static_assert(kParts == 1, "parted rows are stored by store0_part");
            store0_part(it, g, lane, o, lds, 0);
        }
    }

    // Level B input: LDS rows (g, mA), skewed by mA*g.  Strip column of
    // offset c0 - kLead + 4*lane + i - mA*g relative to base0(g):
    // 4*lane + i - kLead + HB + 7 + (7 - mA)*g, always in [-kPad, kStrip).
    // With parts, the LDS holds part mA / kPartRows.
    __host__ __device__ __forceinline__ static void load1(const Item &it, int mA, int lane,
                                                          const T *lds, T (&v)[RB][kW]) {
        (void)it;
#pragma unroll
        for (int g = 0; g < RB; ++g) {
            const int ub = -kLead + HB + (RA - 1) + (RA - 1 - mA) * g;
            const T *row = lds + (g * kPartRows + mA % kPartRows) * kRow + lane;
#pragma unroll
            for (int i = 0; i < kW; ++i) v[g][i] = row[lds_uniform(ub + i)];
        }
    }

    __host__ __device__ __forceinline__ static void store1(const Item &it, int mA, int lane,
                                                           const T (&o)[RB][kW], T *lds) {
        (void)lds;
        store_x<RB>(it, mA, lane, o);
    }

    __host__ __device__ __forceinline__ static void phase2(const Item &, const T *, int) {}
};

// ---------------------------------------------------------------------------
// Pass 2: angle s of quadrant q, output offsets [d0, d0 + kOut).
//   phase 0 (level C): strip g (2^KD of them): blocks J = 8g + j of X at
//     angle s, row j skewed by s*j, base d0 - HD - (8s+7)g - 7 -> LDS rows (g, tC)
//   phase 1 (level D): strip tC (8 of them): LDS rows (g, tC) skewed by
//     (8s + tC)*g, base d0 - kLead -> output tile [d][tC*2^KD + tD], written
//     over the level-C rows after every strip has loaded
//   phase 2: the tile (kOut rows of 2^k2 angles, swizzled by 16-byte group)
//     is stored row by row: out[d0 + d][s*2^k2 + a], 2^k2 contiguous values.
// With KD == 0 phase 0 writes the tile directly.
// ---------------------------------------------------------------------------
template <typename T, int KD>
struct Pass2 {
    static constexpr bool kStaged = false;
    static constexpr int kParts = 1;
    static constexpr int KC = 3;
    static constexpr int k2 = KC + KD;
    static constexpr int RC = 1 << KC, RD = 1 << KD;
    static constexpr int HD = RD - 1;
    static constexpr int R0 = RC, R1 = RD;
    static constexpr int A2 = 1 << k2;   // angles per tile row
    static constexpr int G = A2 / 4;     // 4-element groups per tile row
    static constexpr int tasks0 = RD;
    static constexpr int tasks1 = KD == 0 ? 0 : RC;
    static constexpr size_t lc_elems = KD == 0 ? 0 : size_t(1 << k2) * kRow;
    static constexpr size_t tile_elems = size_t(kOut) * A2;
    // The tile reuses the level-C rows once every level-D strip has loaded.
    static constexpr size_t lds_elems = lc_elems > tile_elems ? lc_elems : tile_elems;

#ifndef ADRT_HOST_EMULATION
    __device__ __forceinline__ static void strip0(T (&v)[R0][kW]) { strip_adrt<KC, T>(v); }
    __device__ __forceinline__ static void strip1(T (&v)[R1][kW]) { strip_adrt<KD, T>(v); }
#endif
    static void strip0_host(T (*v)[R0][kW]) { strip_adrt_host<KC, T>(v); }
    static void strip1_host(T (*v)[R1][kW]) { strip_adrt_host<KD, T>(v); }

    struct Item {
        const T *X;  // row (J, s) at X + J * xrow_stride
        T *out;      // plane (b, q)
        int n, s, d0, C, W1, abase;
        size_t xrow_stride;
    };

    __host__ __device__ __forceinline__ static int num_chunks(int n) {
        return (2 * n - 1 + kOut - 1) / kOut;
    }
    __host__ __device__ __forceinline__ static int num_items(int n, int k1, int nimg) {
        return nimg * 4 * (1 << k1) * num_chunks(n);
    }
    __host__ __device__ __forceinline__ static bool skip(const Item &) { return false; }

    // Item `block` in [0, num_items): (image, quadrant, angle s, chunk).
    __host__ __device__ __forceinline__ static Item item(const T *X_base, T *out_base, int n, int k1,
                                                         int first_image, int block) {
        Item it;
        it.n = n;
        it.C = 2 * n - 1;
        it.W1 = x_valid(n, k1);
        const int P1 = x_pitch(n, k1);
        const int nS = 1 << k1;
        const int nchunk = num_chunks(n);
        const int chunk = block % nchunk;
        block /= nchunk;
        it.s = block % nS;
        block /= nS;
        const int q = block & 3;
        const int b = first_image + (block >> 2);
        it.d0 = chunk * kOut;
        it.xrow_stride = (size_t)nS * P1;  // between blocks J at fixed s
        it.X = X_base + ((size_t)(b - first_image) * 4 + q) * ((size_t)1 << k2) * it.xrow_stride +
               (size_t)it.s * P1 + kXL;
        it.out = out_base + ((size_t)b * 4 + q) * it.C * (size_t)n;
        it.abase = it.s << k2;
        return it;
    }

    __host__ __device__ __forceinline__ static int base0(const Item &it, int g) {
        return KD == 0 ? it.d0 - kLead : it.d0 - HD - (RC * it.s + RC - 1) * g - (RC - 1);
    }

    // Swizzled tile position of (row d, angle a): 16-byte group a/4 is XORed
    // with (d/4) mod G so that lanes writing consecutive rows hit distinct banks.
    __host__ __device__ __forceinline__ static int tile_pos(int d, int a) {
        return d * A2 + 4 * ((a >> 2) ^ ((d >> 2) & (G - 1))) + (a & 3);
    }

    // Level C input: blocks 8g + j at angle s, row j skewed by s*j.  The
    // row margins make every window one load with its start clamped into
    // [-4, W1] (a window reaching past either edge reads the margin).
    __host__ __device__ __forceinline__ static void load0(const Item &it, int g, int lane,
                                                          T (&v)[RC][kW]) {
        const int c = base0(it, g) + kW * lane;
#pragma unroll
        for (int j = 0; j < RC; ++j) {
            // start clamped into [-4, W1], taken relative to offset -4
            int cj = c - it.s * j + kXPad;
            cj = cj < 0 ? 0 : (cj > it.W1 + kXPad ? it.W1 + kXPad : cj);
            load_vec4(it.X - kXPad + (size_t)(g * RC + j) * it.xrow_stride + unsigned(cj), v[j][0],
                      v[j][1], v[j][2], v[j][3]);
        }
    }

    // Tile rows d = 4*lane + i - kLead (lanes 16..63), angles a0 + t.
    template <int R>
    __host__ __device__ __forceinline__ static void store_tile(int a0, int lane, const T (&o)[R][kW],
                                                               T *tile) {
        if (lane < kLead / kW) return;
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int d = kW * lane + i - kLead;
            if constexpr (R % 4 == 0) {
#pragma unroll
                for (int t = 0; t < R; t += 4) {
                    T *p = tile + tile_pos(d, a0 + t);  // 4-element group, 16-byte aligned
#pragma unroll
                    for (int k = 0; k < 4; ++k) p[k] = o[t + k][i];
                }
            } else {
#pragma unroll
                for (int t = 0; t < R; ++t) tile[tile_pos(d, a0 + t)] = o[t][i];
            }
        }
    }

    __host__ __device__ __forceinline__ static void store0(const Item &it, int g, int lane,
                                                           const T (&o)[RC][kW], T *lds) {
        (void)it;
        if constexpr (KD == 0) {
            store_tile<RC>(0, lane, o, lds);
        } else {
#pragma unroll
            for (int t = 0; t < RC; ++t)
#pragma unroll
                for (int i = 0; i < kW; ++i) lds[(g * RC + t) * kRow + lds_uniform(i) + lane] = o[t][i];
        }
    }

    // Level D input: LDS rows (g, tC), skewed by (8s + tC)*g.  Strip column of
    // offset d0 - kLead + 4*lane + i - (8s + tC)*g relative to base0(g):
    // 4*lane + i - kLead + HD + 7 + (7 - tC)*g, always in [-kPad, kStrip).
    __host__ __device__ __forceinline__ static void load1(const Item &it, int tC, int lane,
                                                          const T *lds, T (&v)[RD][kW]) {
        (void)it;
#pragma unroll
        for (int g = 0; g < RD; ++g) {
            const int ub = -kLead + HD + (RC - 1) + (RC - 1 - tC) * g;
            const T *row = lds + (g * RC + tC) * kRow + lane;
#pragma unroll
            for (int i = 0; i < kW; ++i) v[g][i] = row[lds_uniform(ub + i)];
        }
    }

    __host__ __device__ __forceinline__ static void store1(const Item &it, int tC, int lane,
                                                           const T (&o)[RD][kW], T *lds) {
        (void)it;
        store_tile<RD>(tC * RD, lane, o, lds);
    }

    // Coalesced store of the tile: thread tid handles (row, group) pairs.
    __host__ __device__ __forceinline__ static void phase2(const Item &it, const T *lds, int tid) {
        const T *tile = lds;
        for (int f = tid; f < kOut * G; f += kThreads) {
            const int d = f / G, grp = f % G;
            const int od = it.d0 + d;
            if (od >= it.C) continue;
            const T *p = tile + tile_pos(d, 4 * grp);
            store_vec4_stream(it.out + (size_t)od * it.n + it.abase + 4 * grp, p[0], p[1], p[2], p[3]);
        }
    }
};

// ===========================================================================
// Fused two-pass back-projection (bdrt, adrt_cdefs_bdrt.hpp:55-187).
//
// Stage i pairs angle groups (2si, 2si+1) of the previous stage into group
// si; section s of the pair becomes sections 2s (x + y, no shift) and 2s+1
// (x at row r+si plus y at row r+si+1).  Unrolled, a group m of R = 2^K
// consecutive groups at one level goes through K stages whose shifts are
// m * (R / 2^(u+1)) + (local pair index) at stage u, so the K-stage result
// for output section s_loc is the local K-stage strip (strip_bdrt) shifted
// by the common m * s_loc: computed at strip offset rho, it belongs to row
// rho - m * s_loc.  The L stages split into
//   pass A (kA = floor(L/2) stages): for every group p of 2^kA input angles
//     and chunk of input rows, a 3-stage level over 8 angles and a
//     (kA-3)-stage level over the level-A groups, stored to Z[p][s][r] with
//     the per-section skew p * s (store side);
//   pass B (kB = ceil(L/2) stages): for every section s of Z and chunk of
//     output rows, a 3-stage level over 8 groups p and a (kB-3)-stage level
//     whose inputs are skewed by g' * s3 (load side), written through an
//     LDS tile as rows of 2^kB consecutive output columns.
// Rows past the last real row C-1 read as -0.0 at row C and +0.0 above, the
// values that make the reference's boundary rule (right[C-1-si] = x[C-1],
// zeros after) fall out of the same adds; the result is bit-identical to
// the reference up to the sign of zero results (canonical -0.0 == +0.0).
// ===========================================================================

// Z rows: kZL elements of scratch, rows [0, C), margins C: -0.0 and
// C+1..C+4: +0.0, then scratch up to the pitch (whole 128-byte lines).
constexpr int kZL = 32;
__host__ __device__ __forceinline__ int z_pitch(int n) { return kZL + (2 * n - 1 + 8 + 31) / 32 * 32; }

// Z rows that the output rows [0, rows) of a kB-stage pass B are made of.
// Output row d adds Z rows d .. d + 2^kB - 1 only: the level-3 strip reaches
// 7 rows ahead (4 + 2 + 1), level 4 reads level-3 offsets up to
// (2^(kB-3) - 1) * 7 ahead and its own strip reaches 2^(kB-3) - 1 more.
// Rows from rows + 2^kB - 1 on, and past C + 4 (the last margin row pass A
// writes), are not needed: a kept chunk's 256-row window may still load
// them, but they reach only its rows from `rows` on, which are not stored.
// With rows = C this is every row pass A writes.
__host__ __device__ __forceinline__ int bdrt_z_rows(int n, int rows, int kB) {
    const int reach = rows + (1 << kB) - 1;
    const int full = 2 * n - 1 + 5;
    return reach < full ? reach : full;
}

// LDS column c = 4*lane + u of a row (u >= 0, c < kRow): interleaved by
// residue so that one wave-uniform u is bank-consecutive across lanes.
__host__ __device__ __forceinline__ int lds_col(int u) { return (u & (kW - 1)) * kSlice + (u >> 2); }

template <typename T, int K2>
struct BdrtA {
    static constexpr int kA = 3 + K2;
    static constexpr int kParts = 1;
    static constexpr int R0 = 8, R1 = 1 << K2;
    static constexpr int tasks0 = R1;                // level-1 strips g
    static constexpr int tasks1 = K2 == 0 ? 0 : 8;  // level-2 strips s1
    static constexpr size_t lds_elems = K2 == 0 ? 0 : size_t(R1) * 8 * kRow;
    // With LDS (K2 > 0) the item's input block - 256 rows x 2^kA angles,
    // 2^kA * sizeof(T) contiguous bytes per row - is first staged through
    // LDS by the whole work-group with row-wise loads, transposed to
    // [angle][row] so that each strip lane reads its 4 rows of an angle as
    // one 16-byte LDS read.  (Read directly, a strip's loads touch one cache
    // line per lane.)  The staging area is reused by the level-1 rows after a
    // barrier.
    static constexpr bool kStaged = K2 > 0;
    static constexpr int kSA = 8 * R1;                 // angles of an item
    static constexpr int kSP = kLanes * kW + 4;        // staged row pitch
    static_assert(!kStaged || size_t(kSA) * kSP <= lds_elems, "staging fits the LDS");

#ifndef ADRT_HOST_EMULATION
    __device__ __forceinline__ static void strip0(T (&v)[R0][kW]) { strip_bdrt<3, T>(v); }
    __device__ __forceinline__ static void strip1(T (&v)[R1][kW]) { strip_bdrt<K2, T>(v); }
#endif
    static void strip0_host(T (*v)[R0][kW]) { strip_bdrt_host<3, T>(v); }
    static void strip1_host(T (*v)[R1][kW]) { strip_bdrt_host<K2, T>(v); }

    struct Item {
        const T *y;  // input plane (C, n)
        T *Z;        // row 0 of section 0 of group p; section s at + s * PZ
        int n, C, p, c0, PZ;
        int zrows;  // Z rows pass B reads (bdrt_z_rows)
    };

    // Chunks of input rows: outputs reach row C + 4 (margins) after the
    // largest skew (groups - 1) * (sections - 1).
    __host__ __device__ __forceinline__ static int num_chunks(int n) {
        const int C = 2 * n - 1;
        const int maxskew = ((n >> kA) - 1) * ((1 << kA) - 1);
        return (C + 5 + maxskew + kOut - 1) / kOut;
    }
    __host__ __device__ __forceinline__ static int num_items(int n, int nimg) {
        return nimg * 4 * (n >> kA) * num_chunks(n);
    }

    // Item `block` -> (image, quadrant, group p, chunk), chunk fastest.
    __host__ __device__ __forceinline__ static Item item(const T *y_base, T *Z_base, int n,
                                                         int first_image, int zrows, int block) {
        Item it;
        it.n = n;
        it.C = 2 * n - 1;
        it.PZ = z_pitch(n);
        it.zrows = zrows;
        const int nP = n >> kA, nchunk = num_chunks(n);
        const int chunk = block % nchunk;
        block /= nchunk;
        it.p = block % nP;
        block /= nP;
        const int q = block & 3;
        const int b = first_image + (block >> 2);
        it.c0 = chunk * kOut;
        it.y = y_base + ((size_t)b * 4 + q) * it.C * (size_t)n;
        it.Z = Z_base + (((size_t)(b - first_image) * 4 + q) * nP + it.p) * ((size_t)1 << kA) * it.PZ +
               kZL;
        return it;
    }

    // Level-1 strip g: input rows c0 + 4*lane + i (row C reads -0.0, rows
    // above +0.0), angles p * 2^kA + 8g + v for v in [0, 8).  All loads are
    // issued before the padding is applied.
    __host__ __device__ __forceinline__ static void load0(const Item &it, int g, int lane,
                                                          T (&v)[R0][kW]) {
        const int a0 = (it.p << kA) + 8 * g;
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int r = it.c0 + kW * lane + i;
            const int rc = r < it.C ? r : it.C - 1;
            const T *src = it.y + unsigned(rc * it.n + a0);
            load_vec4(src, v[0][i], v[1][i], v[2][i], v[3][i]);
            load_vec4(src + 4, v[4][i], v[5][i], v[6][i], v[7][i]);
        }
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int r = it.c0 + kW * lane + i;
            if (r >= it.C) {
                const T pad = r == it.C ? neg_zero<T>() : T(0);
#pragma unroll
                for (int j = 0; j < R0; ++j) v[j][i] = pad;
            }
        }
    }

    // stg[a][rr] = y[min(c0 + rr, C - 1)][p * 2^kA + a], rr < 256, a < 2^kA.
    // A thread's vectors are all loaded before any is stored, so its loads
    // are in flight together.
    __host__ __device__ __forceinline__ static void stage(const Item &it, T *lds, int tid) {
        constexpr int vecs = kSA / 4;  // 4-element vectors per input row
        constexpr int per = kLanes * kW * vecs / kThreads;
        static_assert(per * kThreads == kLanes * kW * vecs, "whole vectors per thread");
        T v[per][4];
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int f = tid + q * kThreads;
            const int rr = f / vecs, vq = f - rr * vecs;
            const int r = it.c0 + rr;
            const int rc = r < it.C ? r : it.C - 1;
            load_vec4(it.y + unsigned(rc * it.n + (it.p << kA) + 4 * vq), v[q][0], v[q][1],
                      v[q][2], v[q][3]);
        }
#pragma unroll
        for (int q = 0; q < per; ++q) {
            const int f = tid + q * kThreads;
            const int rr = f / vecs, vq = f - rr * vecs;
            T *col = lds + (4 * vq) * kSP + rr;
            col[0] = v[q][0];
            col[kSP] = v[q][1];
            col[2 * kSP] = v[q][2];
            col[3 * kSP] = v[q][3];
        }
    }

    // Items that pass B does not need:
    //  - section s of the item goes to Z rows from c0 - p*s on, so with s <
    //    2^kA every row it stores lies at or past c0 - p*(2^kA - 1); when
    //    that is at or past the rows pass B reads, nothing it stores is read;
    //  - an item whose input rows all lie past the last row (c0 > C) sees
    //    only +0 padding: all it would store is +0 (and -0 at row C) into Z
    //    rows r > C - p*s, which pass B makes itself (BdrtB::load0).  About
    //    a third of the items are such (they exist only for the skew).
    __host__ __device__ __forceinline__ static bool skip(const Item &it) {
        return it.c0 > it.C || it.c0 - it.p * ((1 << kA) - 1) >= it.zrows;
    }

    // load0 from the staged block (same values, then the same padding).
    __host__ __device__ __forceinline__ static void load0s(const Item &it, int g, int lane,
                                                           const T *lds, T (&v)[R0][kW]) {
#pragma unroll
        for (int j = 0; j < R0; ++j) {
            const T *src = lds + (R0 * g + j) * kSP + kW * lane;
            load_vec4(src, v[j][0], v[j][1], v[j][2], v[j][3]);
        }
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int r = it.c0 + kW * lane + i;
            if (r >= it.C) {
                const T pad = r == it.C ? neg_zero<T>() : T(0);
#pragma unroll
                for (int j = 0; j < R0; ++j) v[j][i] = pad;
            }
        }
    }

    // Section s at strip offsets 4*lane + i (lanes 0..47, the chunk's kOut
    // offsets) -> Z rows c0 + 4*lane + i - p*s; rows from C on get the
    // margin values.  Windows holding no row of [0, C + 5) are skipped; the
    // others stay inside the row's scratch ends.
    __host__ __device__ __forceinline__ static void store_z(const Item &it, int s, int lane,
                                                            const T (&o)[kW]) {
        if (lane >= kOut / kW) return;
        const int r0 = it.c0 + kW * lane - it.p * s;
        if (r0 < -3 || r0 > it.C + 4) return;
        T v[kW];
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int r = r0 + i;
            v[i] = r < it.C ? o[i] : (r == it.C ? neg_zero<T>() : T(0));
        }
        store_vec4(it.Z + (size_t)s * it.PZ + r0, v[0], v[1], v[2], v[3]);
    }

    __host__ __device__ __forceinline__ static void store0(const Item &it, int g, int lane,
                                                           const T (&o)[R0][kW], T *lds) {
        if constexpr (K2 == 0) {
            (void)lds;
            (void)g;
#pragma unroll
            for (int t = 0; t < R0; ++t) store_z(it, t, lane, o[t]);
        } else {
            (void)it;
#pragma unroll
            for (int t = 0; t < R0; ++t)
#pragma unroll
                for (int i = 0; i < kW; ++i) lds[(g * R0 + t) * kRow + lds_col(i) + lane] = o[t][i];
        }
    }

    // Level-2 strip s1: input g is level-1 section s1 of strip g at strip
    // offset + g*s1 (the relative skew of the groups R1*p + g).
    __host__ __device__ __forceinline__ static void load1(const Item &it, int s1, int lane,
                                                          const T *lds, T (&v)[R1][kW]) {
        (void)it;
#pragma unroll
        for (int g = 0; g < R1; ++g) {
            const T *row = lds + (g * R0 + s1) * kRow + lane;
#pragma unroll
            for (int i = 0; i < kW; ++i) v[g][i] = row[lds_col(i + g * s1)];
        }
    }

    // Level-2 output s2 of strip s1 is section R1*s1 + s2; the strip's
    // offsets are rows c0 + offset - R1*p*s1 before the section skew p*s2.
    __host__ __device__ __forceinline__ static void store1(const Item &it, int s1, int lane,
                                                           const T (&o)[R1][kW], T *lds) {
        (void)lds;
#pragma unroll
        for (int s2 = 0; s2 < R1; ++s2) store_z(it, R1 * s1 + s2, lane, o[s2]);
    }

    __host__ __device__ __forceinline__ static void phase2(const Item &, const T *, int) {}
};

template <typename T, int K4>
struct BdrtB {
    static constexpr int kParts = 1;
    static constexpr bool kStaged = false;
    static constexpr int kB = 3 + K4;
    static constexpr int R0 = 8, R1 = 1 << K4;
    static constexpr int A2 = 1 << kB;  // output columns per tile row
    static constexpr int G = A2 / 4;    // 4-element groups per tile row
    static constexpr int tasks0 = R1;                // level-3 strips g'
    static constexpr int tasks1 = K4 == 0 ? 0 : 8;  // level-4 strips s3
    static constexpr size_t lc_elems = K4 == 0 ? 0 : size_t(R1) * 8 * kRow;
    static constexpr size_t tile_elems = size_t(kOut) * A2;
    static constexpr size_t lds_elems = lc_elems > tile_elems ? lc_elems : tile_elems;

#ifndef ADRT_HOST_EMULATION
    __device__ __forceinline__ static void strip0(T (&v)[R0][kW]) { strip_bdrt<3, T>(v); }
    __device__ __forceinline__ static void strip1(T (&v)[R1][kW]) { strip_bdrt<K4, T>(v); }
#endif
    static void strip0_host(T (*v)[R0][kW]) { strip_bdrt_host<3, T>(v); }
    static void strip1_host(T (*v)[R1][kW]) { strip_bdrt_host<K4, T>(v); }

    struct Item {
        const T *Z;  // row 0 of (group 0, section s); group p at + p * pstride
        T *out;      // output plane (C, n)
        int n, C, d0, abase, s;
        int rows;    // output rows stored: [0, rows), rows <= C
        size_t pstride;
    };

    // Chunks of kOut output rows covering rows [0, rows).
    __host__ __device__ __forceinline__ static int num_chunks(int rows) {
        return (rows + kOut - 1) / kOut;
    }
    __host__ __device__ __forceinline__ static int num_items(int kA, int nimg, int rows) {
        return nimg * 4 * (1 << kA) * num_chunks(rows);
    }

    __host__ __device__ __forceinline__ static bool skip(const Item &) { return false; }

    // Item `block` -> (image, quadrant, section s, chunk), chunk fastest.
    __host__ __device__ __forceinline__ static Item item(const T *Z_base, T *out_base, int n, int kA,
                                                         int first_image, int rows, int block) {
        Item it;
        it.n = n;
        it.C = 2 * n - 1;
        it.rows = rows;
        const int PZ = z_pitch(n);
        const int nS = 1 << kA, nchunk = num_chunks(rows);
        const int chunk = block % nchunk;
        block /= nchunk;
        const int s = block % nS;
        block /= nS;
        const int q = block & 3;
        const int b = first_image + (block >> 2);
        it.d0 = chunk * kOut;
        it.pstride = (size_t)nS * PZ;
        it.s = s;
        it.Z = Z_base + ((size_t)(b - first_image) * 4 + q) * (size_t)n * PZ + (size_t)s * PZ + kZL;
        it.out = out_base + ((size_t)b * 4 + q) * it.C * (size_t)n;
        it.abase = s << kB;
        return it;
    }

    // Level-3 strip g': Z rows d0 + 4*lane + i of groups p = 8g' + j.  Z row
    // r of (p, s) adds input rows from r + p*s on, so past lim = C - p*s it
    // holds +0 from padding only (row C: the -0 margin), and pass A does not
    // write the items that hold nothing else (BdrtA::skip).  Those rows are
    // made here: a window wholly past lim is loaded from lim's window (a row
    // that is written) and replaced, one that crosses lim is patched.  A
    // window starting past C + 1 starts at C + 1.
    __host__ __device__ __forceinline__ static void load0(const Item &it, int g, int lane,
                                                          T (&v)[R0][kW]) {
        int c = it.d0 + kW * lane;
        c = c > it.C + 1 ? it.C + 1 : c;
#pragma unroll
        for (int j = 0; j < R0; ++j) {
            const int lim = it.C - (8 * g + j) * it.s;  // > 0: p*s < C
            const int cl = c > lim ? (lim > 3 ? lim - 3 : 0) : c;
            load_vec4(it.Z + (size_t)(8 * g + j) * it.pstride + unsigned(cl), v[j][0], v[j][1],
                      v[j][2], v[j][3]);
        }
#pragma unroll
        for (int j = 0; j < R0; ++j) {
            const int lim = it.C - (8 * g + j) * it.s;
#pragma unroll
            for (int i = 0; i < kW; ++i) {
                const int r = c + i;
                if (r > lim) v[j][i] = r == it.C ? neg_zero<T>() : T(0);
            }
        }
    }

    // Swizzled tile position of (row d, column a), as in the forward's tile.
    __host__ __device__ __forceinline__ static int tile_pos(int d, int a) {
        return d * A2 + 4 * ((a >> 2) ^ ((d >> 2) & (G - 1))) + (a & 3);
    }

    template <int R>
    __host__ __device__ __forceinline__ static void store_tile(int a0, int lane, const T (&o)[R][kW],
                                                               T *tile) {
        if (lane >= kOut / kW) return;
#pragma unroll
        for (int i = 0; i < kW; ++i) {
            const int d = kW * lane + i;
#pragma unroll
            for (int t = 0; t < R; ++t) tile[tile_pos(d, a0 + t)] = o[t][i];
        }
    }

    __host__ __device__ __forceinline__ static void store0(const Item &it, int g, int lane,
                                                           const T (&o)[R0][kW], T *lds) {
        (void)it;
        if constexpr (K4 == 0) {
            (void)g;
            store_tile<R0>(0, lane, o, lds);
        } else {
#pragma unroll
            for (int t = 0; t < R0; ++t)
#pragma unroll
                for (int i = 0; i < kW; ++i) lds[(g * R0 + t) * kRow + lds_col(i) + lane] = o[t][i];
        }
    }

    // Level-4 strip s3: input g' is level-3 section s3 of strip g' at strip
    // offset + g'*s3.
    __host__ __device__ __forceinline__ static void load1(const Item &it, int s3, int lane,
                                                          const T *lds, T (&v)[R1][kW]) {
        (void)it;
#pragma unroll
        for (int g = 0; g < R1; ++g) {
            const T *row = lds + (g * R0 + s3) * kRow + lane;
#pragma unroll
            for (int i = 0; i < kW; ++i) v[g][i] = row[lds_col(i + g * s3)];
        }
    }

    __host__ __device__ __forceinline__ static void store1(const Item &it, int s3, int lane,
                                                           const T (&o)[R1][kW], T *lds) {
        (void)it;
        store_tile<R1>(s3 * R1, lane, o, lds);
    }

    // Tile rows d -> out[d0 + d][abase .. abase + A2), coalesced.
    __host__ __device__ __forceinline__ static void phase2(const Item &it, const T *lds, int tid) {
        const T *tile = lds;
        for (int f = tid; f < kOut * G; f += kThreads) {
            const int d = f / G, grp = f % G;
            const int od = it.d0 + d;
            if (od >= it.rows) continue;
            const T *p = tile + tile_pos(d, 4 * grp);
            store_vec4_stream(it.out + (size_t)od * it.n + it.abase + 4 * grp, p[0], p[1], p[2], p[3]);
        }
    }
};

inline size_t z_bytes_per_image(int n, size_t elem) {
    return size_t(4) * n * size_t(z_pitch(n)) * elem;
}

#ifndef ADRT_HOST_EMULATION

// Item `idx` on the whole work-group: phase 0 strips (tasks0 <= kWaves,
// one per wave), barrier, phase 1 strips (exactly one per wave) loaded and
// computed in registers, barrier (every read of phase-0 data done), phase 1
// stores, barrier, phase 2 cooperative stores.  The LDS must be free on
// entry; phase 2 still reads it on return.
template <typename P, typename T, typename... A>
__device__ __forceinline__ void run_item_at(int idx, T *lds, A... args) {
    static_assert(P::tasks0 <= kWaves, "at most one phase-0 strip per wave");
    static_assert(P::tasks1 == 0 || P::tasks1 == kWaves, "one phase-1 strip per wave");
    const typename P::Item it = P::item(args..., idx);
    if (P::skip(it)) return;  // work-group uniform
    const int wave = wave_uniform(threadIdx.x / kLanes);
    const int lane = threadIdx.x % kLanes;
    if constexpr (P::kStaged) {
        // cooperative staged load; the staging area is the level-1 area, so
        // every strip reads its block before any strip stores
        P::stage(it, lds, threadIdx.x);
        __syncthreads();
        T v[P::R0][kW];
        if (wave < P::tasks0) P::load0s(it, wave, lane, lds, v);
        __syncthreads();
        if (wave < P::tasks0) {
            P::strip0(v);
            P::store0(it, wave, lane, v, lds);
        }
    } else if constexpr (P::kParts == 2) {
        // level A in every strip; the level B strips of part 0 load their
        // rows, every strip stores its part-1 rows over them, then the level
        // B strips of part 1 load (a strip holds its part-1 rows and its
        // level B rows together for one barrier only)
        static_assert(P::tasks0 == kWaves && P::tasks1 == kWaves, "one strip per wave");
        T v[P::R0][kW];
        P::load0(it, wave, lane, v);
        P::strip0(v);
        P::store0_part(it, wave, lane, v, lds, 0);
        __syncthreads();
        T w[P::R1][kW];
        const bool first = wave < P::kPartRows;  // level B strip of part 0
        if (first) P::load1(it, wave, lane, lds, w);
        __syncthreads();
        P::store0_part(it, wave, lane, v, lds, 1);
        __syncthreads();
        if (!first) P::load1(it, wave, lane, lds, w);
        P::strip1(w);
        P::store1(it, wave, lane, w, lds);
        return;
    } else if (wave < P::tasks0) {
        T v[P::R0][kW];
        P::load0(it, wave, lane, v);
        P::strip0(v);
        P::store0(it, wave, lane, v, lds);
    }
    __syncthreads();
    if constexpr (P::tasks1 > 0) {
        T w[P::R1][kW];
        P::load1(it, wave, lane, lds, w);
        P::strip1(w);
        __syncthreads();
        P::store1(it, wave, lane, w, lds);
        __syncthreads();
    }
    P::phase2(it, lds, threadIdx.x);
}

// One item per work-group, by the XCD-aware schedule.
template <typename P, typename T, typename... A>
__device__ __forceinline__ void run_item(int items, T *lds, A... args) {
    const Schedule sc = schedule(blockIdx.x, gridDim.x, items);
    if (sc.idx >= sc.hi) return;
    run_item_at<P, T>(sc.idx, lds, args...);
}

// Register budget: waves per SIMD the kernel's LDS allows (160 KB per CU,
// at most 4 work-groups per CU), capped by what the strips' registers need
// without spilling (`cap` waves: 64, 80, 128 or 256 VGPRs).
template <typename P, typename T>
constexpr int lds_waves_per_simd(int cap) {
    constexpr size_t bytes = P::lds_elems * sizeof(T);
    constexpr int wgs = bytes == 0 ? 4 : int((size_t(160) << 10) / bytes);
    const int w = (wgs < 4 ? (wgs < 1 ? 1 : wgs) : 4) * kWaves / 4;
    return w < cap ? w : cap;
}

#endif  // ADRT_HOST_EMULATION

// Split of L = log2 n stages.  Forward: pass 1 takes k1 = 3 + KB stages,
// pass 2 the remaining k2 = 3 + KD; variant 0 has k1 = ceil(L/2), variant 1
// k1 = floor(L/2) (the same for even L).  Back-projection: pass A takes
// kA = floor(L/2) (variant 0) or ceil(L/2) (variant 1).  For L in [6, 12]
// every pass keeps 3..6 stages.
// Stage splits (they differ only for odd L).  The default is the split that
// measured faster for the size and element type (MI355X, ms per call):
//   forward pass 1: floor(L/2) stages for fp64 or L <= 9 (fp64 2048^2:
//     2.39 against 2.92; fp32 128^2: 0.081 against 0.132), else ceil(L/2)
//     (fp32 2048^2: equal);
//   bdrt pass A: ceil(L/2) for L <= 9 (fp32 128^2: 0.175 against 0.191;
//     fp64 512^2: 1.01 against 1.05), else floor(L/2) (2048^2: 4.38
//     against 4.83 in fp64, 1.62 against 1.86 in fp32).
// ALT_SPLIT (variant 1) selects the other split.
inline int fwd_pass1_stages(int L, int variant, int elem = 4) {
    const bool low = elem == 8 || L <= 9;
    return (variant == 1) != low ? L / 2 : (L + 1) / 2;
}
inline int bdrt_passA_stages(int L, int variant) {
    const bool high = L <= 9;
    return (variant == 1) != high ? (L + 1) / 2 : L / 2;
}

// Largest k1 of either variant: sizes X for both.
inline int pass1_stages(int L) { return (L + 1) / 2; }

// Images per group: by default keep one group's intermediate at or below
// ~136 MiB (two 2048^2 fp32 images) so that pass 2 re-reads it largely from
// the 256 MB Infinity Cache; the final output bypasses that cache (streaming
// stores).  Fewer, larger groups also mean fewer launches.
constexpr size_t kDefaultGroupBytes = size_t(136) << 20;
inline size_t group_images(size_t B, size_t per_image_bytes,
                           size_t budget = kDefaultGroupBytes) {
    size_t g = per_image_bytes >= budget ? 1 : budget / per_image_bytes;
    return g < B ? g : B;
}

inline size_t x_bytes_per_image(int n, int L, size_t elem) {
    return size_t(4) * n * size_t(x_pitch(n, pass1_stages(L))) * elem;
}

}  // namespace fused
}  // namespace adrt_hip
