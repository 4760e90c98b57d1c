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
#include <type_traits>

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

// Paired stages (Co >= 32).  Output chain (l odd, col) reads its pair at
// source row r + 1 + col; its rows r >= R - 1 - col have no source term and
// are 0 + 0 = +0.  Indexing the odd chain by t = R - 2 - col - r, it reads
// source row R - 1 - t: the row the even chain (l - 1, col) reads at the
// same t (output row R - 1 - t).  So a wave holds the 32 even chains of
// columns c0 .. c0 + 31 of group 2pl in lanes 0-31 and the 32 odd chains of
// the same columns of group 2pl + 1 in lanes 32-63, and at step t every lane
// loads its pair from source row R - 1 - t: lanes L and L + 32 load the same
// 8 bytes, so each pair leaves memory once per stage.  Even lanes store their
// row directly (one 128-byte row segment per step); an odd lane's row lags
// by col - c0, so odd rows pass through a 32-row LDS ring of each lane's own
// column and row f = R - 2 - c0 - t is stored whole at step t.  The odd
// chains' +0 rows above R - 2 - c0 are filled after the sweep; the ring
// starts at +0 for the rows a lane skips before its first source row.
constexpr int kHalf = 32;
// Steps per prefetch block of the paired kernel: pair loads in flight per
// lane (with one store per step, a block's loads and stores stay within the
// 63 outstanding memory operations the counter tracks).
constexpr int kPairSteps = 16;

struct PairLane {
    bool odd;
    int col;  // column within the group
    int pc;   // source column of the (2col, 2col+1) pair
    int oc;   // output column
};

// wave_in_plane -> the lane's chain; c0 = the wave's first column in its group
__host__ __device__ inline PairLane make_pair_lane(int n, int Li, int wave_in_plane, int lane,
                                                   int &c0) {
    const int Ci = n / Li, Co = Ci / 2;
    const int wpp = Co / kHalf;  // waves per pair of groups
    const int pl = wave_in_plane / wpp;
    c0 = (wave_in_plane % wpp) * kHalf;
    PairLane p;
    p.odd = lane >= kHalf;
    p.col = c0 + (lane & (kHalf - 1));
    p.pc = pl * Ci + 2 * p.col;
    p.oc = (2 * pl + (p.odd ? 1 : 0)) * Co + p.col;
    return p;
}

// Step at source row y = R - 1 - t: even lanes produce row y, odd lanes row
// y - 1 - col (if >= 0), in the reference's order of additions.
template <typename T>
__host__ __device__ inline bool pair_step(const PairLane &pp, int y, int R, T p0, T p1, T prev1,
                                          T &acc, T &val, int &row) {
    T v = T(0);
    if (!pp.odd) {
        row = y;
        v += p0;                          // src[2col][row]
        if (y + 1 < R) v -= prev1;        // src[2col+1][row+1]
        if (y + 1 < R) v += acc;
    } else {
        row = y - 1 - pp.col;
        if (row < 0) return false;
        v += p1;                          // src[2col+1][row+1+col], row+1+col = y < R
        v -= p0;                          // src[2col][row+1+col]
        v += acc;                         // row + 1 < R
    }
    acc = v;
    val = v;
    return true;
}

// Two paired stages in one pass (stage k with Co >= 32 and stage k + 1).
// The wave's 64 stage-k chains (lanes 0-31: even group 2pl, columns
// c0 .. c0 + 31; lanes 32-63: odd group 2pl + 1, the same columns; computed
// exactly as in the pair kernel) are the sources of 64 stage-(k+1) chains of
// columns col' = c0/2 + i (i < 16) in the four output groups 4pl + type:
//   lane 2i       type A (group 4pl, even):      a = S_e[2col'][r],     b = S_e[2col'+1][r+1]
//   lane 2i + 1   type B (group 4pl + 1, odd):   a = S_e[2col'+1][u],   b = S_e[2col'][u]
//   lane 32 + 2i  type C (group 4pl + 2, even):  a = S_o[2col'][r],     b = S_o[2col'+1][r+1]
//   lane 32 + 2i + 1  type D (group 4pl + 3):    a = S_o[2col'+1][u],   b = S_o[2col'][u]
// (u = r + 1 + col'; S_e / S_o the stage-k chains of the even / odd group).
// A stage-k odd chain of column c produces row y - 1 - c while reading
// source row y, so at step y each operand is a value the lane itself or its
// partner lane (lane ^ 1) produced at this step or one or two steps before:
//   A: a = own now,  b = partner's previous,  row r = y
//   B: a = own now,  b = partner's now,       r = y - 1 - col'
//   C: a = own now,  b = partner's 2 steps ago, r = y - 1 - 2col'
//   D: a = own previous, b = partner's now,   r = y - 2 - 3col'
// Each lane keeps its last three stage-k values and sends its partner the
// one the partner needs (one lane swap per step).  Rows of a stage-k odd
// chain above its first source row, and of every stage-(k+1) chain above
// its first row reached in the sweep, are +0 (their sources are +0 rows);
// the histories start at +0 for them.  Output rows lag by type and column;
// each type stores, at step y, its row y - minlag (the row its least-lagged
// lane produces now, which every lane of the type has produced by then)
// through a ring per type in LDS of 16 (B), 32 (C) or 64 (D) rows (lag
// spreads 15, 30, 45); A stores its row y directly.  The rows above the
// first flushed row are filled with +0 after the sweep.
constexpr int kQuadRing = 16 * (16 + 32 + 64);  // ring elements of B, C and D

struct QuadLane {
    PairLane p;  // the lane's stage-k chain
    int type;    // 0 A, 1 B, 2 C, 3 D
    int colq;    // stage-(k+1) column in its group
    int lag;     // stage-(k+1) output row = y - lag
    int minlag;  // smallest lag of the type in the wave (flush row = y - minlag)
    int oc;      // stage-(k+1) output column
    int rbase, rmask, ri;  // ring element of row r: rbase + (r & rmask) * 16 + ri
};

__host__ __device__ inline QuadLane make_quad_lane(int n, int Li, int wave_in_plane, int lane,
                                                   int &c0) {
    QuadLane q;
    q.p = make_pair_lane(n, Li, wave_in_plane, lane, c0);
    const int Ci = n / Li, Co = Ci / 2, Coq = Co / 2;
    const int e = lane & (kHalf - 1);
    const int i = e >> 1;
    const int pl = wave_in_plane / (Co / kHalf);
    q.type = ((lane >> 5) << 1) | (e & 1);
    q.colq = c0 / 2 + i;
    q.oc = (4 * pl + q.type) * Coq + q.colq;
    const int cm = c0 / 2;
    const int k = q.type == 0 ? 0 : (q.type == 1 ? 1 : (q.type == 2 ? 2 : 3));
    const int mult = k == 3 ? 3 : k;        // lag = add + mult * col'
    const int add = k == 0 ? 0 : (k == 3 ? 2 : 1);
    q.lag = add + mult * q.colq;
    q.minlag = add + mult * cm;
    q.ri = i;
    q.rmask = k == 1 ? 15 : (k == 2 ? 31 : 63);
    q.rbase = k == 1 ? 0 : (k == 2 ? 16 * 16 : 16 * (16 + 32));
    return q;
}

// The value this lane sends its partner at a step (h0 = this step's stage-k
// value, h1 the previous, h2 the one before).
template <typename T>
__host__ __device__ __forceinline__ T quad_send(const QuadLane &q, T h0, T h1, T h2) {
    return q.type == 1 ? h1 : (q.type == 3 ? h2 : h0);
}

// Stage k + 1 at step y: a = the lane's own operand (D: h1, else h0), b =
// the partner's; the reference's order of additions.
template <typename T>
__host__ __device__ __forceinline__ bool quad_step(const QuadLane &q, int y, int R, T a, T b,
                                                   T &acc, T &val, int &row) {
    row = y - q.lag;
    if (row < 0) return false;
    T v = T(0);
    v += a;
    if (q.type == 0 || q.type == 2) {  // even chains: b enters while row + 1 < R
        if (row + 1 < R) v -= b;
    } else {  // odd chains: row + 1 + col' < R at every row the sweep reaches
        v -= b;
    }
    if (row + 1 < R) v += acc;
    acc = v;
    val = v;
    return true;
}

// Fused tail.  Once a stage's input groups are at most 64 columns wide
// (Ci <= 64), one wavefront's 64 columns hold whole groups of that stage and
// of every later one, so the last S stages run in one pass: the first of
// them (t = 0) reads memory exactly as the stage kernel does (lagged odd
// chains) and writes its rows into a ring in LDS; each later stage t works on
// row s + t at step s, reading rows its producer wrote at earlier steps
// (rows x .. x + Co) from the producer's ring, and the last stage stores its
// row straight to memory (every lane produces the same row at a step, so the
// store is whole and coalesced).
//
// Ring sizes (rows, a power of two each).  Stage 0 writes rows down to
// s - maxlag (maxlag = Co_0 - 1) while stage 1 still reads rows up to
// s + 1 + Co_1, so its ring spans maxlag + Co_1 + 2 = 3 Co_0 / 2 + 1 rows; a
// later stage t writes row s + t while its consumer reads rows up to
// s + t + 1 + Co_{t+1}: Co_{t+1} + 2 rows.  The first stage of an S-stage
// tail has Ci = 2^S columns per group.
constexpr int kMaxTail = 6;  // stages 64, 32, .., 2 columns wide
constexpr int kTailSteps = 16;  // steps per prefetch block (1 load, <= 1 store a step)

__host__ __device__ constexpr int pow2_at_least(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}
// rows of the ring stage t (0 <= t <= S - 2) writes for stage t + 1
__host__ __device__ constexpr int tail_ring_rows(int S, int t) {
    return t == 0 ? pow2_at_least(3 * (1 << (S - 1)) / 2 + 1)
                  : pow2_at_least((1 << (S - t - 2)) + 2);
}
// first row of that ring in the kernel's ring buffer
__host__ __device__ constexpr int tail_ring_base(int S, int t) {
    int b = 0;
    for (int u = 0; u < t; ++u) b += tail_ring_rows(S, u);
    return b;
}
__host__ __device__ constexpr int tail_rows_total(int S) { return tail_ring_base(S, S - 1); }

// The operands of row x of a later stage's chain in its producer's ring
// (rows base .. base + mask, pitch kLanes + 1):
//   even: a = src[2col][x],          b = src[2col+1][x+1]  (b used if x+1 < R)
//   odd:  a = src[2col+1][x+1+col],  b = src[2col][x+1+col] (both if x+1+col < R)
// As per-lane constants: a sits at ring row x + offa, column ca, and enters
// while x < lima (b likewise), so a step computes each index with one add,
// one mask and one multiply-add and each condition with one compare.  Both
// slots are read unconditionally (a ring slot exists for every row) and the
// conditions are applied by selects, so a step has no branches.
struct TailLane {
    bool live;
    bool odd;
    int col;
    int pc;  // ring column of the src pair (2col, 2col+1) of the group
    int offa, offb, ca, cb, lima, limb;
};

__host__ __device__ inline TailLane make_tail_lane(int n, int Ci, int ch0, int lane, int R) {
    TailLane t;
    const int ch = ch0 + lane;
    t.live = ch < n;
    const int Co = Ci / 2;
    const int chc = t.live ? ch : ch0;
    const int l = chc / Co;
    t.col = chc % Co;
    t.odd = (l & 1) != 0;
    t.pc = (l >> 1) * Ci + 2 * t.col - ch0;
    t.offa = t.odd ? 1 + t.col : 0;
    t.offb = t.odd ? 1 + t.col : 1;
    t.ca = t.pc + (t.odd ? 1 : 0);
    t.cb = t.pc + (t.odd ? 0 : 1);
    t.lima = t.odd ? R - 1 - t.col : R;  // odd: x + 1 + col < R; even: always (x < R)
    t.limb = t.odd ? R - 1 - t.col : R - 1;  // even: x + 1 < R
    return t;
}

struct TailOperands {
    int ia, ib;   // ring element indices of a and b
    bool use_a;   // a enters (even: always)
    bool use_b;   // b enters
};

__host__ __device__ __forceinline__ TailOperands tail_operands(const TailLane &tl, int x,
                                                               int base, int mask) {
    TailOperands o;
    o.ia = (base + ((x + tl.offa) & mask)) * (kLanes + 1) + tl.ca;
    o.ib = (base + ((x + tl.offb) & mask)) * (kLanes + 1) + tl.cb;
    o.use_a = x < tl.lima;
    o.use_b = x < tl.limb;
    return o;
}

// val = 0; val += a; val -= b; (if x + 1 < R) val += acc, the reference's
// order, with each skipped term skipped exactly (selects, not zeros added:
// -0 + 0 would turn a -0 into +0).
template <typename T>
__host__ __device__ __forceinline__ T tail_combine(const TailOperands &o, T a, T b, T acc,
                                                   bool add_acc) {
    T v = o.use_a ? T(0) + a : T(0);
    v = o.use_b ? v - b : v;
    return add_acc ? v + acc : v;
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

// A paired stage (Co >= 32): grid = planes * (n / 64) single-wave
// work-groups.
template <typename T>
__global__ void __launch_bounds__(kLanes)
    pair_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li,
                int waves_per_plane) {
    using P2 = typename Pair<T>::type;
    __shared__ T ring[kHalf][kHalf];  // [row & 31][odd lane]: bank = lane
    const int lane = threadIdx.x;
    const int hl = lane & (kHalf - 1);
    const size_t plane = blockIdx.x / waves_per_plane;
    int c0;
    const PairLane pp = make_pair_lane(n, Li, int(blockIdx.x % waves_per_plane), lane, c0);
    const T *src = in + plane * size_t(R) * n + pp.pc;
    T *dst = out + plane * size_t(R) * n + pp.oc;
    if (pp.odd)
        for (int r = 0; r < kHalf; ++r) ring[r][hl] = T(0);

    auto load = [&](int t) -> P2 {
        const int y = R - 1 - t;
        return *reinterpret_cast<const P2 *>(src + size_t(y < 0 ? 0 : y) * n);
    };
    T acc = T(0), prev1 = T(0);
    auto run_step = [&](int t, const P2 &p) {
        const int y = R - 1 - t;
        T val;
        int row;
        const bool made = pair_step<T>(pp, y, R, p.x, p.y, prev1, acc, val, row);
        if (pp.odd && made) ring[row & (kHalf - 1)][hl] = val;
        prev1 = p.y;
        // one store instruction per step: even lanes their row y, odd lanes
        // row f = y - 1 - c0, which every odd lane has produced
        const int f = y - 1 - c0;
        const int orow = pp.odd ? f : y;
        const T ov = pp.odd ? ring[f & (kHalf - 1)][hl] : val;
        if (orow >= 0) dst[size_t(orow) * n] = ov;
    };

    P2 cur[kPairSteps];
    int t = 0;
#pragma unroll
    for (int k = 0; k < kPairSteps; ++k) cur[k] = load(k);
    for (; t + kPairSteps <= R; t += kPairSteps) {  // whole blocks
        P2 nxt[kPairSteps];
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) nxt[k] = load(t + kPairSteps + k);
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) run_step(t + k, cur[k]);
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < kPairSteps; ++k)  // tail: steps t .. R-1
        if (t + k < R) run_step(t + k, cur[k]);
    if (pp.odd)  // the odd chains' +0 rows
        for (int r = R - 1 - c0; r < R; ++r) dst[size_t(r) * n] = T(0);
}

template <typename T>
__device__ __forceinline__ T lane_swap1(T v);
template <>
__device__ __forceinline__ float lane_swap1<float>(float v) {
    return __builtin_bit_cast(float, __shfl_xor(__builtin_bit_cast(int, v), 1));
}
template <>
__device__ __forceinline__ double lane_swap1<double>(double v) {
    return __shfl_xor(v, 1);
}

// Two paired stages (stage k has Co >= 32): grid = planes * (n / 64)
// single-wave work-groups; `in` holds stage k's input, `out` receives stage
// k + 1's output.
template <typename T>
__global__ void __launch_bounds__(kLanes)
    quad_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li,
                int waves_per_plane) {
    using P2 = typename Pair<T>::type;
    __shared__ T ring[kQuadRing];
    const int lane = threadIdx.x;
    const size_t plane = blockIdx.x / waves_per_plane;
    int c0;
    const QuadLane q = make_quad_lane(n, Li, int(blockIdx.x % waves_per_plane), lane, c0);
    const T *src = in + plane * size_t(R) * n + q.p.pc;
    T *dst = out + plane * size_t(R) * n + q.oc;
    for (int e = lane; e < kQuadRing; e += kLanes) ring[e] = T(0);
    __syncthreads();

    auto load = [&](int t) -> P2 {
        const int y = R - 1 - t;
        return *reinterpret_cast<const P2 *>(src + size_t(y < 0 ? 0 : y) * n);
    };
    T acc = T(0), prev1 = T(0), acc2 = T(0);
    T h0 = T(0), h1 = T(0), h2 = T(0);
    auto run_step = [&](int t, const P2 &p) {
        const int y = R - 1 - t;
        T val;
        int row;
        // stage k (values of rows a stage-k odd chain does not reach stay +0)
        const bool made = pair_step<T>(q.p, y, R, p.x, p.y, prev1, acc, val, row);
        prev1 = p.y;
        h2 = h1;
        h1 = h0;
        h0 = made ? val : T(0);
        // stage k + 1
        const T b = lane_swap1<T>(quad_send<T>(q, h0, h1, h2));
        const T a = q.type == 3 ? h1 : h0;
        T v2;
        int row2;
        const bool made2 = quad_step<T>(q, y, R, a, b, acc2, v2, row2);
        if (q.type != 0 && made2) ring[q.rbase + (row2 & q.rmask) * 16 + q.ri] = v2;
        // one store per step: A its row y, the others their type's row y - minlag
        const int f = y - q.minlag;
        const T ov = q.type == 0 ? v2 : ring[q.rbase + (f & q.rmask) * 16 + q.ri];
        if (f >= 0) dst[size_t(f) * n] = ov;
    };

    P2 cur[kPairSteps];
    int t = 0;
#pragma unroll
    for (int k = 0; k < kPairSteps; ++k) cur[k] = load(k);
    for (; t + kPairSteps <= R; t += kPairSteps) {  // whole blocks
        P2 nxt[kPairSteps];
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) nxt[k] = load(t + kPairSteps + k);
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) run_step(t + k, cur[k]);
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < kPairSteps; ++k)  // tail: steps t .. R-1
        if (t + k < R) run_step(t + k, cur[k]);
    for (int r = R - q.minlag; r < R; ++r) dst[size_t(r) * n] = T(0);  // the +0 rows
}

// The last S stages (the first with Li0 groups, Ci = n / Li0 = 2^S), one
// single-wave work-group per 64 columns of a plane.
template <typename T, int S>
__global__ void __launch_bounds__(kLanes)
    tail_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li0,
                int waves_per_plane) {
    using P2 = typename Pair<T>::type;
    constexpr int kM0 = tail_ring_rows(S, 0);
    __shared__ T rbuf[tail_rows_total(S) * (kLanes + 1)];
    const int lane = threadIdx.x;
    const size_t plane = blockIdx.x / waves_per_plane;
    const Wave w = make_wave(R, n, Li0, int(blockIdx.x % waves_per_plane));
    const Lane ln = make_lane(w, lane);
    const T *src = in + plane * size_t(R) * n;
    T *dst = out + plane * size_t(R) * n + w.ch0 + lane;
    TailLane tl[S - 1];
#pragma unroll
    for (int t = 1; t < S; ++t) tl[t - 1] = make_tail_lane(n, n / (Li0 << t), w.ch0, lane, R);

    auto load = [&](int s) -> P2 {
        int r = s + ln.roff;
        r = r < 0 ? 0 : (r > R - 1 ? R - 1 : r);
        return *reinterpret_cast<const P2 *>(src + size_t(r) * n + (ln.live ? ln.pc : 0));
    };

    T acc0 = T(0), prev1 = T(0);
    T acc[S - 1];
#pragma unroll
    for (int t = 0; t < S - 1; ++t) acc[t] = T(0);
    // kChecked: some later stage's row sk + t may lie outside [0, R) (the
    // first and last steps); a block of steps with every stage inside runs
    // without the per-stage range branches
    auto run_step = [&](int sk, const P2 &p, auto checked) {
        constexpr bool kChecked = decltype(checked)::value;
        // every later stage's operands first (rows written at earlier steps),
        // so the step's LDS reads are issued together
        T a[S - 1], b[S - 1];
        TailOperands o[S - 1];
#pragma unroll
        for (int t = 1; t < S; ++t) {
            o[t - 1] = tail_operands(tl[t - 1], sk + t, tail_ring_base(S, t - 1),
                                     tail_ring_rows(S, t - 1) - 1);
            a[t - 1] = rbuf[o[t - 1].ia];
            b[t - 1] = rbuf[o[t - 1].ib];
        }
        // stage 0 (the step function on the loaded pair)
        T val;
        int row;
        if (step<T>(w, ln, sk, p.x, p.y, prev1, acc0, val, row))
            rbuf[(row & (kM0 - 1)) * (kLanes + 1) + lane] = val;
        prev1 = p.y;
        // later stages: row x = sk + t is wave-uniform
#pragma unroll
        for (int t = 1; t < S; ++t) {
            const int x = sk + t;
            if (kChecked && (x < 0 || x >= R)) continue;
            const T v = tail_combine<T>(o[t - 1], a[t - 1], b[t - 1], acc[t - 1], x + 1 < R);
            acc[t - 1] = v;
            if (t < S - 1)
                rbuf[(tail_ring_base(S, t) + (x & (tail_ring_rows(S, t) - 1))) * (kLanes + 1) +
                     lane] = v;
            else if (tl[t - 1].live)
                dst[size_t(x) * n] = v;
        }
    };

    // steps w.top .. -(S-1): the last stage reaches row 0 at step -(S-1);
    // loads for steps below 0 are clamped into the rows and unused
    const int s_end = -(S - 1);
    P2 cur[kTailSteps];
    int s = w.top;
#pragma unroll
    for (int k = 0; k < kTailSteps; ++k) cur[k] = load(s - k);
    using Checked = std::true_type;
    using Unchecked = std::false_type;
    for (; s - (kTailSteps - 1) >= s_end; s -= kTailSteps) {  // whole blocks
        P2 nxt[kTailSteps];
#pragma unroll
        for (int k = 0; k < kTailSteps; ++k) nxt[k] = load(s - kTailSteps - k);
        if (s <= R - S && s - (kTailSteps - 1) >= 0) {  // rows s+1 .. s+S-1 inside
#pragma unroll
            for (int k = 0; k < kTailSteps; ++k) run_step(s - k, cur[k], Unchecked{});
        } else {
#pragma unroll
            for (int k = 0; k < kTailSteps; ++k) run_step(s - k, cur[k], Checked{});
        }
#pragma unroll
        for (int k = 0; k < kTailSteps; ++k) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < kTailSteps; ++k)  // tail: steps s .. s_end
        if (s - k >= s_end) run_step(s - k, cur[k], Checked{});
}
#endif  // ADRT_HOST_EMULATION

}  // namespace iadrt_native

// Stages in the fused tail for L = log2 n stages and a configuration
// variant: DEFAULT at most 5 (measured at n = 2048, float32 and float64:
// 5 stages beat 6, whose first ring needs 64 rows and halves the resident
// waves, and 4, which leaves one more pass to memory); ALT_SPLIT at most 2;
// STREAMING none (one kernel per stage).
constexpr int kDefaultTail = 5;

template <typename T>
inline int tail_stages(int L, int variant = ADRT_HIP_VARIANT_DEFAULT) {
    if (variant == ADRT_HIP_VARIANT_STREAMING) return 0;
    const int cap = variant == ADRT_HIP_VARIANT_ALT_SPLIT ? kDefaultTail / 2 : kDefaultTail;
    return L < cap ? L : cap;
}

// Passes of the k0 stages before the tail: two paired stages (Co >= 32)
// per pass while both remain (with `quad`), else one.
inline int iadrt_passes(int n, int k0, bool quad) {
    int passes = 0, Li = 1;
    for (int k = 0; k < k0; ++passes) {
        const bool paired = n / Li / 2 >= iadrt_native::kHalf;
        const int step = paired && quad && k + 1 < k0 ? 2 : 1;
        k += step;
        Li <<= step;
    }
    return passes;
}

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
    // the last S stages run fused (S = 1: no tail kernel); before them,
    // pairs of paired stages run as one pass each (not in STREAMING)
    const int variant = fused_variant(ADRT_HIP_OP_IADRT, n);
    const int S = tail_stages<T>(L, variant);
    const bool quad = variant != ADRT_HIP_VARIANT_STREAMING;
    const int k0 = L - (S >= 2 ? S : 0);  // stages before the tail
    const int passes = iadrt_passes(int(n), k0, quad) + (S >= 2 ? 1 : 0);
    // ping-pong so that the last pass writes `out`
    T *bufs[2] = {(passes % 2 == 1) ? out : ws, (passes % 2 == 1) ? ws : out};
    const int waves_per_plane = int((n + iadrt_native::kLanes - 1) / iadrt_native::kLanes);
    const unsigned grid = unsigned(B * 4 * size_t(waves_per_plane));
    const T *src = in;
    int Li = 1;
    int pass = 0;
    for (int k = 0; k < k0; ++pass) {
        T *dst = bufs[pass & 1];
        const bool paired = int(n) / Li / 2 >= iadrt_native::kHalf;  // Co >= 32
        if (paired && quad && k + 1 < k0) {  // stages k and k + 1 in one pass
            hipLaunchKernelGGL(iadrt_native::quad_kernel<T>, dim3(grid),
                               dim3(iadrt_native::kLanes), 0, s, src, dst, int(R), int(n), Li,
                               waves_per_plane);
            k += 2;
            Li *= 4;
        } else {
            if (paired)  // even and odd chains paired
                hipLaunchKernelGGL(iadrt_native::pair_kernel<T>, dim3(grid),
                                   dim3(iadrt_native::kLanes), 0, s, src, dst, int(R), int(n), Li,
                                   waves_per_plane);
            else
                hipLaunchKernelGGL(iadrt_native::stage_kernel<T>, dim3(grid),
                                   dim3(iadrt_native::kLanes), 0, s, src, dst, int(R), int(n),
                                   Li, waves_per_plane);
            k += 1;
            Li *= 2;
        }
        src = dst;
    }
    if (S >= 2) {
        T *dst = bufs[pass & 1];
        const dim3 g(grid), blk(iadrt_native::kLanes);
        const int r = int(R), nn = int(n);
        switch (S) {
            case 2: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 2>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            case 3: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 3>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            case 4: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 4>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            case 5: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 5>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            default: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 6>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
        }
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
