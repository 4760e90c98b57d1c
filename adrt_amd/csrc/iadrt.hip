// Exact inverse ADRT (iadrt) on gfx950.
//
// A stage (adrt_cdefs_iadrt.hpp:52-105) maps Li processed column groups of
// Ci columns to 2Li groups of Ci/2 columns.  Output chain (l, col) is a
// serial recurrence down the rows, in the reference's order:
//   val = 0; val += a; val -= b; (if row + 1 < R) val += out[row + 1]
// with, for src chains 2col and 2col+1 of group pl = l/2,
//   l even: a = src[2col][row],         b = src[2col+1][row+1]
//   l odd:  a = src[2col+1][row+1+col], b = src[2col][row+1+col]
// (terms outside the rows drop out).  A parallel scan would reorder these
// additions, so every chain stays serial; chains are independent.
//
// Layout.  Every stage works in the API layout itself, (B, 4, R, n) with
// chain (l, col) = column l*Co + col (R = 2n-1 rows, n columns): the first
// stage reads the input array and the last writes the output, with no
// transposes.  One wavefront owns 64 consecutive output columns and sweeps
// the rows downward, each lane carrying its chain's running value.
//
// Coalescing.  An even chain reads its src pair at rows s and s+1; an odd
// chain at row r+1+col, which differs per lane.  Odd lane col therefore runs
// (col - c0) rows behind (c0 = the wave's first column when the wave holds
// one group, else 0), so at step s every odd lane reads src row s + 1 + c0:
// each step is one 8-byte (fp32) pair load per lane from one or two rows.
// The lagging outputs go through a 64-row LDS ring per wave and row s is
// stored, whole and coalesced, at step s, when every lane has produced it.
// Pair loads are issued a block of kSteps steps ahead of their use.
#include "common.hpp"

namespace adrt_hip {
namespace iadrt_native {

constexpr int kLanes = 64;
constexpr int kRing = 64;   // rows in flight: lag 0..63
constexpr int kSteps = 8;   // steps per prefetch block

// Wave-uniform parameters of one wavefront of a stage.
struct Wave {
    int R, n, Ci, Co;
    int ch0;      // first output column of the wave
    int c0;       // column base of the lag
    int maxlag;   // largest lag in the wave
    int top;      // first step: R - 1 + maxlag
};

__host__ __device__ inline Wave make_wave(int R, int n, int Li, int wave_in_plane) {
    Wave w;
    w.R = R;
    w.n = n;
    w.Ci = n / Li;
    w.Co = w.Ci / 2;
    w.ch0 = wave_in_plane * kLanes;
    const bool one_group = w.Co >= kLanes;
    w.c0 = one_group ? w.ch0 % w.Co : 0;
    const int l0 = w.ch0 / w.Co;
    w.maxlag = one_group ? ((l0 & 1) ? kLanes - 1 : 0) : w.Co - 1;
    w.top = R - 1 + w.maxlag;
    return w;
}

// Per-lane constants.
struct Lane {
    bool live;   // column exists
    bool odd;
    int pc;      // src column of the (2col, 2col+1) pair
    int roff;    // src row = step + roff
    int lag;     // output row = step - lag
};

__host__ __device__ inline Lane make_lane(const Wave &w, int lane) {
    Lane ln;
    const int ch = w.ch0 + lane;
    ln.live = ch < w.n;
    const int chc = ln.live ? ch : w.ch0;
    const int l = chc / w.Co, col = chc % w.Co;
    ln.odd = (l & 1) != 0;
    ln.pc = (l >> 1) * w.Ci + 2 * col;
    ln.roff = ln.odd ? 1 + w.c0 : 0;
    ln.lag = ln.odd ? col - w.c0 : 0;
    return ln;
}

// One step of one lane: `p` is the src pair at row step + roff, `prev` the
// pair this lane loaded at the previous step (row step + 1 for even lanes).
// Returns true and the output in `val` when the lane produces row
// step - lag at this step.
template <typename T>
__host__ __device__ inline bool step(const Wave &w, const Lane &ln, int s, T p0, T p1, T prev1,
                                     T &acc, T &val, int &row) {
    row = s - ln.lag;
    if (!ln.live || row < 0 || row >= w.R) return false;
    T v = T(0);
    if (!ln.odd) {
        v += p0;                          // src[2col][row]
        if (row + 1 < w.R) v -= prev1;    // src[2col+1][row+1]
    } else if (s + ln.roff < w.R) {       // row + 1 + col < R
        v += p1;                          // src[2col+1][row+1+col]
        v -= p0;                          // src[2col][row+1+col]
    }
    if (row + 1 < w.R) v += acc;
    acc = v;
    val = v;
    return true;
}

#ifndef ADRT_HOST_EMULATION
template <typename T>
struct Pair;
template <>
struct Pair<float> {
    using type = float2;
};
template <>
struct Pair<double> {
    using type = double2;
};

// grid: planes * waves_per_plane single-wave work-groups.
template <typename T>
__global__ void __launch_bounds__(kLanes)
    stage_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li,
                 int waves_per_plane) {
    using P2 = typename Pair<T>::type;
    __shared__ T ring[kRing][kLanes + 1];
    const int lane = threadIdx.x;
    const size_t plane = blockIdx.x / waves_per_plane;
    const Wave w = make_wave(R, n, Li, int(blockIdx.x % waves_per_plane));
    const Lane ln = make_lane(w, lane);
    const T *src = in + plane * size_t(R) * n;
    T *dst = out + plane * size_t(R) * n;
    const bool store_lane = w.ch0 + lane < n;

    // src pair at row s + roff, clamped into the rows (unused when clamped)
    auto load = [&](int s) -> P2 {
        int r = s + ln.roff;
        r = r < 0 ? 0 : (r > R - 1 ? R - 1 : r);
        return *reinterpret_cast<const P2 *>(src + size_t(r) * n + (ln.live ? ln.pc : 0));
    };

    // Row s of the wave's columns is flushed at step s.  Every lane stores at
    // every step, so each block issues exactly kSteps stores after its
    // prefetch and the wait for the prefetched pairs leaves them in flight:
    // steps above the last row store junk into row R-1, which its real value
    // overwrites at step R-1 (same lane, same address, later); lanes past n
    // repeat the last live lane's store (same value, same address).
    const int flush_lane = store_lane ? lane : n - 1 - w.ch0;
    T *flush_col = dst + w.ch0 + flush_lane;
    auto run_step = [&](int sk, const P2 &p, T &acc, T &prev1) {
        T val;
        int row;
        if (step<T>(w, ln, sk, p.x, p.y, prev1, acc, val, row)) ring[row % kRing][lane] = val;
        prev1 = p.y;
        const int fr = sk < R ? sk : R - 1;
        flush_col[size_t(fr) * n] = ring[sk % kRing][flush_lane];
    };

    T acc = T(0);
    T prev1 = T(0);
    P2 cur[kSteps];
    int s = w.top;
#pragma unroll
    for (int k = 0; k < kSteps; ++k) cur[k] = load(s - k);
    for (; s >= kSteps - 1; s -= kSteps) {  // whole blocks
        P2 nxt[kSteps];
#pragma unroll
        for (int k = 0; k < kSteps; ++k) nxt[k] = load(s - kSteps - k);
#pragma unroll
        for (int k = 0; k < kSteps; ++k) run_step(s - k, cur[k], acc, prev1);
#pragma unroll
        for (int k = 0; k < kSteps; ++k) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < kSteps; ++k)  // tail: steps s..0
        if (k <= s) run_step(s - k, cur[k], acc, prev1);
}
#endif  // ADRT_HOST_EMULATION

}  // namespace iadrt_native

#ifndef ADRT_HOST_EMULATION
template <typename T>
static int run_native(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const size_t R = 2 * n - 1;
    const int L = num_iters(n);
    if (L == 0) {
        if (hipMemcpyAsync(out, in, B * 4 * R * n * sizeof(T), hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
            return launch_status();
        return launch_status();
    }
    // ping-pong so that the last stage writes `out`
    T *bufs[2] = {(L % 2 == 1) ? out : ws, (L % 2 == 1) ? ws : out};
    const int waves_per_plane = int((n + iadrt_native::kLanes - 1) / iadrt_native::kLanes);
    const unsigned grid = unsigned(B * 4 * size_t(waves_per_plane));
    const T *src = in;
    int Li = 1;
    for (int k = 0; k < L; ++k) {
        T *dst = bufs[k & 1];
        hipLaunchKernelGGL(iadrt_native::stage_kernel<T>, dim3(grid), dim3(iadrt_native::kLanes),
                           0, s, src, dst, int(R), int(n), Li, waves_per_plane);
        src = dst;
        Li *= 2;
    }
    return launch_status();
}

size_t iadrt_workspace(int dtype, size_t B, size_t n) {
    const size_t s = dtype == ADRT_HIP_F64 ? 8 : 4;
    return B * 4 * n * (2 * n - 1) * s;
}

int launch_iadrt(const void *in, void *out, void *ws, size_t, int dtype, size_t B, size_t n,
                 hipStream_t s) {
    if (dtype == ADRT_HIP_F32)
        return run_native(static_cast<const float *>(in), static_cast<float *>(out),
                          static_cast<float *>(ws), B, n, s);
    return run_native(static_cast<const double *>(in), static_cast<double *>(out),
                      static_cast<double *>(ws), B, n, s);
}
#endif  // ADRT_HOST_EMULATION

}  // namespace adrt_hip
