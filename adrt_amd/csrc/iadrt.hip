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
// Kernels, by the stages a pass runs (run_native picks them per stage):
//   stage_kernel  one stage, any group width (below, and where paired
//                 chains do not fit: Co < 32);
//   pair_kernel   one stage with Co >= 32: the even and odd chains of 32
//                 columns in one wave, each source pair loaded once;
//   quad_kernel   two stages (stage k with Co >= 32): float64, and float32
//                 where the wide pass does not apply (Co = 32, or store
//                 offsets past 32 bits);
//   wide_kernel   two stages (float32, Co >= 64), two chains per lane and
//                 whole-line stores;
//   tail_kernel   the last S <= 6 stages (groups of at most 64 columns).
//
// Coalescing (stage_kernel).  An even chain reads its src pair at rows s and
// s+1; an odd chain at row r+1+col, which differs per lane.  Odd lane col therefore runs
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
// Row pitch of the LDS rings, in elements.  Lagged lanes access different
// rows of a ring at one step: at pitch 65 the 16 odd chains of a group (rows
// s - col, columns c0 + col) all fall in one bank.  At 68 the tail's accesses
// of a step conflict for about one cycle in all (64 four-byte banks), and,
// unlike at 64, a row address stays one 24-bit multiply-add.
constexpr int kPitch = kLanes + 4;

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
// lane produces now, which every lane of the type has produced by then), so
// a lane's value of step t is stored at step t + d, d = lag - minlag (0 for A
// and for the first lane of each type, up to 15, 30 and 45 for B, C and D).
// The values wait in a per-lane delay line in LDS, indexed by step: slot
// t & mask of the lane's column in its type's region.  The rows above the
// first flushed row are filled with +0 after the sweep.
//
// Steps run in blocks of 16 that start at multiples of 16, so a block writes
// 16 consecutive slots of each line (one address and 16 immediate offsets);
// its 16 reads start at slot (t0 - d) & mask and may run past the last slot,
// so the first 16 slots of B, C and D are mirrored after the last (written
// twice in the blocks that start at slot 0).  Regions (16 columns each): A 16
// slots, B and C 32 + 16, D 64 + 16; their bases lie 16 elements apart
// modulo 64, so the four types' columns of one slot fall in distinct banks.
constexpr int kQuadAhead = 32;         // pair registers: loads run kQuadAhead - 1 steps ahead
constexpr int kQuadMirror = 16;        // mirrored slots
constexpr int kQuadMirrorPeriod = 32;  // smallest mirrored line: blocks at t0 % 32 == 0 mirror
constexpr int kQuadRing = 1840 + (64 + kQuadMirror) * 16;  // ring elements (D's region ends it)

__host__ __device__ constexpr int quad_slots(int type) {
    return type == 0 ? 16 : (type == 3 ? 64 : 32);
}
__host__ __device__ constexpr int quad_region(int type) {
    return type == 0 ? 0 : (type == 1 ? 272 : (type == 2 ? 1056 : 1840));
}

struct QuadLane {
    PairLane p;  // the lane's stage-k chain
    int type;    // 0 A, 1 B, 2 C, 3 D
    int colq;    // stage-(k+1) column in its group
    int lag;     // stage-(k+1) output row = y - lag
    int minlag;  // smallest lag of the type in the wave (flush row = y - minlag)
    int oc;      // stage-(k+1) output column
    int i;       // column of the lane's delay line in its type's region
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
    const int mult = q.type;  // lag = add + mult * col'
    const int add = q.type == 0 ? 0 : (q.type == 3 ? 2 : 1);
    q.lag = add + mult * q.colq;
    q.minlag = add + mult * cm;
    q.i = i;
    return q;
}

// The wave's largest flush lag (type D's minlag): stores address row
// y - quad_mlmax(c0) plus a per-lane offset of quad_store_off elements.
__host__ __device__ inline int quad_mlmax(int c0) { return 2 + 3 * (c0 / 2); }
__host__ __device__ inline size_t quad_store_off(const QuadLane &q, int n, int mlmax) {
    return size_t(mlmax - q.minlag) * size_t(n) + size_t(q.oc);
}
// Whether every store offset of an n-column plane, in bytes, fits 32 bits
// (mlmax <= 3n/4, oc < n).
template <typename T>
__host__ __device__ inline bool quad_narrow(size_t n) {
    return (3 * (n / 4) + 3) * n * sizeof(T) <= size_t(0xffffffffu);
}

// Last block start t0 at which every lane of the wave produces a stage-k
// value, a stage-(k+1) row and a flush row at all 16 steps t0 .. t0 + 15:
// odd stage-k chains need y >= col + 1 (col <= c0 + 31), stage-(k+1) chains
// y >= lag (lag <= 3 c0 / 2 + 47), flushes y >= minlag.  With t0 >= 16 as
// well (row y + 1 < R for every chain), such a block needs no checks.
__host__ __device__ inline int quad_bulk_last(int R, int c0) { return R - 63 - 3 * (c0 / 2); }

// Per-lane constants of the branch-free step (t = step, y = R - 1 - t).
struct QuadConst {
    bool odd;      // the stage-k chain is odd
    bool send1;    // B sends h1
    bool send2;    // D sends h2 (others h0)
    bool own1;     // D: a = h1 (others h0)
    bool first_a;  // A: at t = 0 (row R - 1) neither b nor acc enters
    int made_lim;  // stage k produces a value while t <= made_lim
    int made2_lim; // stage k + 1 produces a row while t <= made2_lim
    int flush_lim; // the flush row y - minlag exists while t <= flush_lim
    int col0;      // element of slot 0 of the lane's delay line
    int mask;      // slots - 1
    int delay;     // lag - minlag
    int mirror;    // elements from slot s to its mirror (0: A, never read past its end)
};

__host__ __device__ inline QuadConst make_quad_const(const QuadLane &q, int R) {
    QuadConst c;
    c.odd = q.p.odd;
    c.send1 = q.type == 1;
    c.send2 = q.type == 3;
    c.own1 = q.type == 3;
    c.first_a = q.type == 0;
    c.made_lim = q.p.odd ? R - 2 - q.p.col : R;  // odd: row y - 1 - col >= 0
    c.made2_lim = R - 1 - q.lag;
    c.flush_lim = R - 1 - q.minlag;
    c.col0 = quad_region(q.type) + q.i;
    c.mask = quad_slots(q.type) - 1;
    c.delay = q.lag - q.minlag;
    c.mirror = q.type == 0 ? 0 : quad_slots(q.type) * 16;
    return c;
}

// Ring elements of a block's first step (step k adds 16 k): the write, the
// mirror write (= the write where none is due) and the read.
struct QuadBlock {
    int w, m, r;
};

__host__ __device__ inline QuadBlock quad_block(const QuadConst &c, int t0) {
    QuadBlock b;
    const int ws = t0 & c.mask;
    b.w = c.col0 + ws * 16;
    b.m = ws == 0 ? b.w + c.mirror : b.w;
    b.r = c.col0 + ((t0 - c.delay) & c.mask) * 16;
    return b;
}

template <typename T>
struct QuadState {
    T acc, prev1;  // stage k: running value, the previous pair's second element
    T h0, h1, h2;  // stage-k values of this step and the two before (+0 where none)
    T acc2;        // stage k + 1 running value
};

// Stage k at step t from the lane's source pair (p0, p1) at row y, in the
// reference's order: even  v = 0 + p0 - prev1 + acc  (only p0 at t = 0),
// odd  v = 0 + p1 - p0 + acc.  Returns the value the lane sends its partner.
// Without kChecked every condition must hold (a bulk block).
template <bool kChecked, typename T>
__host__ __device__ __forceinline__ T quad_stage_k(const QuadConst &c, int t, T p0, T p1,
                                                   QuadState<T> &s) {
    const T x = c.odd ? p1 : p0;
    const T z = c.odd ? p0 : s.prev1;
    const T u = T(0) + x;
    T v = (u - z) + s.acc;
    s.prev1 = p1;
    if (kChecked) {
        v = (c.odd || t != 0) ? v : u;
        const bool made = t <= c.made_lim;
        s.acc = made ? v : s.acc;
        v = made ? v : T(0);
    } else {
        s.acc = v;
    }
    s.h2 = s.h1;
    s.h1 = s.h0;
    s.h0 = v;
    return c.send2 ? s.h2 : (c.send1 ? s.h1 : s.h0);
}

// Stage k + 1 at step t: a = the lane's own operand, b = the partner's;
// v = 0 + a - b + acc2 (A at t = 0: only a).  Returns v (junk where the row
// is negative: it is never flushed).
template <bool kChecked, typename T>
__host__ __device__ __forceinline__ T quad_stage_k1(const QuadConst &c, int t, T b,
                                                    QuadState<T> &s) {
    const T a = c.own1 ? s.h1 : s.h0;
    const T u = T(0) + a;
    T v = (u - b) + s.acc2;
    if (kChecked) {
        v = (c.first_a && t == 0) ? u : v;
        s.acc2 = t <= c.made2_lim ? v : s.acc2;
    } else {
        s.acc2 = v;
    }
    return v;
}

// Wide two-stage pass (float32, stage k with Co >= 64): grid = planes *
// (n / 128) single-wave work-groups.  A wave covers stage-k columns c0 ..
// c0 + 63 of the group pair pl and lane j (< 32) holds two stage-k chains of
// the even group (P = column 2col', Q = 2col' + 1, col' = c0/2 + j), lane
// 32 + j the same two of the odd group; each lane loads the four source
// values of its two pairs as one 16-byte vector.  The lane then holds both
// sources of its stage-(k+1) chains E and F, so no value crosses lanes:
//   lanes 0-31:  E = type A: a = P now,  b = Q one step ago
//                F = type B: a = Q now,  b = P now
//   lanes 32-63: E = type C: a = P now,  b = Q two steps ago
//                F = type D: a = Q one step ago, b = P now
// (the operand timing of the quad kernel, within one lane).  Each type's 32
// output columns are consecutive, so a flush store writes whole 128-byte
// lines (in the quad kernel a type's 16 columns are half a line, the other
// half written by the neighbouring wave 16-48 steps later).  The delay lines
// (per lane, one per chain E / F) span 0, 62, 31 and 93 steps for A, C, B, D:
// A 16 slots, C 64, B 32, D 128, plus 16 mirrored slots for C, B, D; a
// region is [slot][32 lanes], and C and D start 32 elements (mod 64) after A
// and B, so the two halves of a wave's access fall in distinct banks.
constexpr int kWideRing = 4704 + (128 + 16) * 32;  // ring elements (D's region ends it)
constexpr int kWideAhead = 16;  // source vectors in the register ring (loads 15 steps ahead)

__host__ __device__ constexpr int wide_slots(int type) {  // 0 A, 1 B, 2 C, 3 D
    return type == 0 ? 16 : (type == 1 ? 32 : (type == 2 ? 64 : 128));
}
__host__ __device__ constexpr int wide_region(int type) {
    return type == 0 ? 0 : (type == 2 ? 544 : (type == 1 ? 3136 : 4704));
}

struct WideLane {
    bool odd;          // lanes 32-63: the odd stage-k group (types C, D)
    int j;             // lane & 31
    int colq;          // stage-(k+1) column c0/2 + j
    int src;           // first of the four source columns (2 * 2col')
    int ocE, ocF;      // output columns of E and F
    int lagE, lagF;    // output row = y - lag
    int minE, minF;    // smallest lag of each type in the wave (flush row = y - min)
};

__host__ __device__ inline WideLane make_wide_lane(int n, int Li, int wave_in_plane, int lane,
                                                   int &c0) {
    const int Ci = n / Li, Co = Ci / 2, Coq = Co / 2;
    const int wpg = Co / kLanes;  // waves per group pair
    const int pl = wave_in_plane / wpg;
    c0 = (wave_in_plane % wpg) * kLanes;
    const int cm = c0 / 2;
    WideLane w;
    w.odd = lane >= kHalf;
    w.j = lane & (kHalf - 1);
    w.colq = cm + w.j;
    w.src = pl * Ci + 4 * w.colq;
    w.ocE = (4 * pl + (w.odd ? 2 : 0)) * Coq + w.colq;
    w.ocF = (4 * pl + (w.odd ? 3 : 1)) * Coq + w.colq;
    w.lagE = w.odd ? 1 + 2 * w.colq : 0;
    w.lagF = w.odd ? 2 + 3 * w.colq : 1 + w.colq;
    w.minE = w.odd ? 1 + 2 * cm : 0;
    w.minF = w.odd ? 2 + 3 * cm : 1 + cm;
    return w;
}

// The wave's largest flush lags (types C and D): E stores address row
// y - wide_mlE(c0), F row y - wide_mlF(c0), plus a per-lane offset.
__host__ __device__ inline int wide_mlE(int c0) { return 1 + 2 * (c0 / 2); }
__host__ __device__ inline int wide_mlF(int c0) { return 2 + 3 * (c0 / 2); }

// Last iteration start t0 (two 16-step blocks, t0 .. t0 + 31) at which every
// lane produces both stage-k values, both stage-(k+1) rows and both flush
// rows: odd chains need y >= 2col' + 2 (col' <= c0/2 + 31), D chains y >= lag
// (<= 3 c0/2 + 95), so t0 + 31 <= R - 96 - 3 c0/2; with t0 >= 16 as well
// no check is needed.
__host__ __device__ inline int wide_bulk_last(int R, int c0) { return R - 127 - 3 * (c0 / 2); }

struct WideConst {
    bool odd;
    int limP, limQ;    // stage-k chains P / Q produce a value while t <= lim
    int limE, limF;    // stage-(k+1) rows exist while t <= lim
    int flushE, flushF;  // flush rows exist while t <= flush
    int colE, colF;    // element of slot 0 of each delay line
    int maskE, maskF;  // slots - 1
    int dE, dF;        // delays
    int mirE, mirF;    // elements from slot s to its mirror (0: none)
};

__host__ __device__ inline WideConst make_wide_const(const WideLane &w, int R) {
    WideConst c;
    c.odd = w.odd;
    c.limP = w.odd ? R - 2 - 2 * w.colq : R;      // odd P: row y - 1 - 2col' >= 0
    c.limQ = w.odd ? R - 3 - 2 * w.colq : R;      // odd Q: row y - 2 - 2col' >= 0
    c.limE = R - 1 - w.lagE;
    c.limF = R - 1 - w.lagF;
    c.flushE = R - 1 - w.minE;
    c.flushF = R - 1 - w.minF;
    const int tE = w.odd ? 2 : 0, tF = w.odd ? 3 : 1;
    c.colE = wide_region(tE) + w.j;
    c.colF = wide_region(tF) + w.j;
    c.maskE = wide_slots(tE) - 1;
    c.maskF = wide_slots(tF) - 1;
    c.dE = w.lagE - w.minE;
    c.dF = w.lagF - w.minF;
    c.mirE = w.odd ? wide_slots(tE) * 32 : 0;
    c.mirF = wide_slots(tF) * 32;
    return c;
}

struct WideBlock {
    int wE, mE, rE, wF, mF, rF;  // write, mirror write, read elements of step 0 (step k: + 32 k)
};

__host__ __device__ inline WideBlock wide_block(const WideConst &c, int t0) {
    WideBlock b;
    const int sE = t0 & c.maskE, sF = t0 & c.maskF;
    b.wE = c.colE + sE * 32;
    b.mE = sE == 0 ? b.wE + c.mirE : b.wE;
    b.rE = c.colE + ((t0 - c.dE) & c.maskE) * 32;
    b.wF = c.colF + sF * 32;
    b.mF = sF == 0 ? b.wF + c.mirF : b.wF;
    b.rF = c.colF + ((t0 - c.dF) & c.maskF) * 32;
    return b;
}

template <typename T>
struct WideState {
    T accP, accQ, prevP, prevQ;  // stage k: running values, the previous second elements
    T hP, hQ0, hQ1, hQ2;         // stage-k values (+0 where none): P now, Q now and before
    T accE, accF;                // stage k + 1
};

// One step of a lane from its source values (p0, p1, p2, p3) at row y, both
// stages in the reference's order; returns E's and F's values (junk where
// their row is negative: never flushed).
template <bool kChecked, typename T>
__host__ __device__ __forceinline__ void wide_step(const WideConst &c, int t, T p0, T p1, T p2,
                                                   T p3, WideState<T> &s, T &vE, T &vF) {
    // stage k: even  v = 0 + x0 - prev + acc (x0 only at t = 0), odd  v = 0 + x1 - x0 + acc
    const T xP = c.odd ? p1 : p0, zP = c.odd ? p0 : s.prevP;
    const T xQ = c.odd ? p3 : p2, zQ = c.odd ? p2 : s.prevQ;
    const T uP = T(0) + xP, uQ = T(0) + xQ;
    T vP = (uP - zP) + s.accP;
    T vQ = (uQ - zQ) + s.accQ;
    s.prevP = p1;
    s.prevQ = p3;
    if (kChecked) {
        const bool full = c.odd || t != 0;
        vP = full ? vP : uP;
        vQ = full ? vQ : uQ;
        const bool mP = t <= c.limP, mQ = t <= c.limQ;
        s.accP = mP ? vP : s.accP;
        s.accQ = mQ ? vQ : s.accQ;
        vP = mP ? vP : T(0);
        vQ = mQ ? vQ : T(0);
    } else {
        s.accP = vP;
        s.accQ = vQ;
    }
    s.hP = vP;
    s.hQ2 = s.hQ1;
    s.hQ1 = s.hQ0;
    s.hQ0 = vQ;
    // stage k + 1: v = 0 + a - b + acc (A at t = 0: only a)
    const T bE = c.odd ? s.hQ2 : s.hQ1;
    const T aF = c.odd ? s.hQ1 : s.hQ0;
    const T uE = T(0) + s.hP, uF = T(0) + aF;
    vE = (uE - bE) + s.accE;
    vF = (uF - s.hP) + s.accF;
    if (kChecked) {
        vE = (!c.odd && t == 0) ? uE : vE;
        s.accE = t <= c.limE ? vE : s.accE;
        s.accF = t <= c.limF ? vF : s.accF;
    } else {
        s.accE = vE;
        s.accF = vF;
    }
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
// Highest step s0 at which a block s0 .. s0 - 15 of the tail runs without
// checks: every later stage t's row x = s + t has x + 1 < R, and an odd
// chain's operands (row x + 1 + col, col < Co_t = 2^(S-1-t)) lie inside the
// rows, so every operand and the running value enter each addition:
// s + t + Co_t < R for every t.
__host__ __device__ constexpr int tail_free_top(int R, int S) {
    int m = 0;
    for (int t = 1; t < S; ++t) {
        const int v = t + (1 << (S - 1 - t));
        m = v > m ? v : m;
    }
    return R - 1 - m;
}
// Whether the tail's block of steps s0 .. s0 - 15 runs without checks: the
// later stages as above, and every stage-0 row s - lag in [0, R - 1) with its
// source row s + roff < R (s0 - 15 >= maxlag); all the wave's columns must
// exist as well.
__host__ __device__ constexpr bool tail_block_free(int s0, int R, int S, int maxlag) {
    return s0 <= tail_free_top(R, S) && s0 - (kTailSteps - 1) - maxlag >= 0;
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
// (rows base .. base + mask, pitch kPitch):
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
    o.ia = (base + ((x + tl.offa) & mask)) * kPitch + tl.ca;
    o.ib = (base + ((x + tl.offb) & mask)) * kPitch + tl.cb;
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
    __shared__ T ring[kRing][kPitch];
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

// Global-memory bytes: a pointer that passes through an asm statement keeps
// its address space (and so global rather than flat instructions) only when
// its type names it.
using GByte = __attribute__((address_space(1))) char;

// The value of lane ^ 1 (DPP quad_perm [1, 0, 3, 2]; every lane active).
template <typename T>
__device__ __forceinline__ T partner(T v);
template <>
__device__ __forceinline__ float partner<float>(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
}
template <>
__device__ __forceinline__ double partner<double>(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u)), 0xB1, 0xF, 0xF, true));
    const unsigned hi =
        unsigned(__builtin_amdgcn_mov_dpp(int(unsigned(u >> 32)), 0xB1, 0xF, 0xF, true));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// Two paired stages (stage k has Co >= 32): grid = planes * (n / 64)
// single-wave work-groups; `in` holds stage k's input, `out` receives stage
// k + 1's output.
//
// Addressing.  A row is a wave-uniform (scalar) row address plus a per-lane
// byte offset, so a load or store costs no vector instruction for its
// address: the offsets pass through an empty asm statement once per block,
// which keeps them in the block beside the accesses (hoisted out of the loop,
// their 64-bit extension would cost an add per access), and the row address
// steps by one row through another, which keeps the compiler from folding
// the steps into one per-lane address.  Flush stores wait
// to the end of their block, off the serial chain of the LDS reads they
// store.  Bulk blocks store at row y - mlmax plus the lane's offset of
// mlmax - minlag rows (32 bits: the library runs two-stage passes only
// where quad_narrow holds);
// checked blocks store per lane, where their rows exist.
template <typename T>
__global__ void __launch_bounds__(kLanes)
    quad_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li,
                int waves_per_plane) {
    using P2 = T __attribute__((ext_vector_type(2)));  // a native vector: loads from address space 1
    using False = std::integral_constant<bool, false>;
    using True = std::integral_constant<bool, true>;
    __shared__ T ring[kQuadRing];
    const int lane = threadIdx.x;
    const size_t plane = blockIdx.x / waves_per_plane;
    int c0;
    const QuadLane q = make_quad_lane(n, Li, int(blockIdx.x % waves_per_plane), lane, c0);
    const QuadConst c = make_quad_const(q, R);
    const int mlmax = quad_mlmax(c0);
    const size_t row_bytes = size_t(n) * sizeof(T);
    const GByte *gsrc = (const GByte *)(in + plane * size_t(R) * n);
    GByte *gdst = (GByte *)(out + plane * size_t(R) * n);
    T *dst_lane = out + plane * size_t(R) * n + q.oc;
    const uint32_t src_off = uint32_t(q.p.pc) * uint32_t(sizeof(T));
    const uint32_t dst_off = uint32_t(quad_store_off(q, n, mlmax) * sizeof(T));
    for (int e = lane; e < kQuadRing; e += kLanes) ring[e] = T(0);
    __syncthreads();

    QuadState<T> s;
    s.acc = s.prev1 = s.h0 = s.h1 = s.h2 = s.acc2 = T(0);
    // step t0 + k: both stages, the delay-line write(s), the read of the flush value
    auto step = [&](auto checked, auto mirrored, int t0, int k, const P2 &p,
                    const QuadBlock &b) -> T {
        constexpr bool kChecked = decltype(checked)::value;
        const int t = t0 + k;
        const T send = quad_stage_k<kChecked, T>(c, t, p.x, p.y, s);
        const T v2 = quad_stage_k1<kChecked, T>(c, t, partner<T>(send), s);
        ring[b.w + 16 * k] = v2;
        if (decltype(mirrored)::value) ring[b.m + 16 * k] = v2;
        return ring[b.r + 16 * k];
    };
    // a bulk block's flushes: rows R - 1 - t0 - k - minlag, all present
    auto flush_bulk = [&](int t0, const T *ov) {
        uint32_t off = dst_off;
        asm volatile("" : "+v"(off));
        GByte *row = gdst + size_t(R - 1 - t0 - mlmax) * row_bytes;
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) {
            if (k > 0) {
                row -= row_bytes;
                asm volatile("" : "+s"(row));
            }
            *reinterpret_cast<__attribute__((address_space(1))) T *>(row + off) = ov[k];
        }
    };
    // a checked block's flushes, steps t0 .. t0 + count - 1
    auto flush_checked = [&](int t0, int count, const T *ov) {
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k)
            if (k < count && t0 + k <= c.flush_lim)
                dst_lane[size_t(R - 1 - t0 - k - q.minlag) * n] = ov[k];
    };

    // Loads run kQuadAhead - 1 steps ahead in a ring of registers: slot
    // t % kQuadAhead holds the pair of step t; step t reads it, and slot
    // (t - 1) % kQuadAhead (step t - 1's pair, whose second element step t
    // still reads as prev1) is refilled after step t with the pair of step
    // t + kQuadAhead - 1.  Each refill comes after its slot's last reader in
    // program order, and the refills' row addresses form one chain through
    // asm statements, so every slot keeps one register across iterations (no
    // copies, which would wait for every load in flight).  An iteration is
    // two 16-step blocks from t0 % 32 == 0 (the first may mirror, the second
    // never does).
    using GP2 = const __attribute__((address_space(1))) P2;
    P2 buf[kQuadAhead];
    {
        uint32_t off = src_off;
        asm volatile("" : "+v"(off));
#pragma unroll
        for (int k = 0; k < kQuadAhead - 1; ++k) {  // steps 0 .. kQuadAhead - 2 (clamped)
            const int y = R - 1 - k;
            buf[k] = *reinterpret_cast<GP2 *>(gsrc + size_t(y < 0 ? 0 : y) * row_bytes + off);
        }
    }
    // an iteration's refills: steps t0 + k + kQuadAhead - 1, rows y0 - k
    // (clamped to row 0 in the last iterations; never needed where checks are)
    auto refill_rows = [&](auto clamped, int t0, auto body) {
        const int y0 = R - 1 - t0 - (kQuadAhead - 1);
        uint32_t off = src_off;
        asm volatile("" : "+v"(off));
        const GByte *row = gsrc + size_t(y0 < 0 ? 0 : y0) * row_bytes;
        auto refill = [&](int k) {
            if (k > 0) {
                if (decltype(clamped)::value)
                    row = y0 - k >= 0 ? row - row_bytes : row;
                else
                    row -= row_bytes;
                asm volatile("" : "+s"(row));
            }
            buf[(k + kQuadAhead - 1) % kQuadAhead] = *reinterpret_cast<GP2 *>(row + off);
            // no instruction crosses a step: hoisting later steps' operand
            // selects would wait for their loads early and empty the ring
            __builtin_amdgcn_sched_barrier(0);
        };
        body(refill);
    };
    auto checked_iteration = [&](int t0) {
        refill_rows(True(), t0, [&](auto &refill) {
#pragma unroll
            for (int h = 0; h < kQuadAhead; h += kPairSteps) {
                const QuadBlock b = quad_block(c, t0 + h);
                T ov[kPairSteps];
#pragma unroll
                for (int k = 0; k < kPairSteps; ++k) {
                    ov[k] = step(True(), True(), t0 + h, k, buf[h + k], b);
                    refill(h + k);
                }
                flush_checked(t0 + h, kPairSteps, ov);
            }
        });
    };
    auto bulk_iteration = [&](int t0) {
        refill_rows(False(), t0, [&](auto &refill) {  // rows y0 - 31 >= 0: see quad_bulk_last
            T ov[kPairSteps];
            QuadBlock b = quad_block(c, t0);
#pragma unroll
            for (int k = 0; k < kPairSteps; ++k) {
                ov[k] = step(False(), True(), t0, k, buf[k], b);
                refill(k);
            }
            flush_bulk(t0, ov);
            b = quad_block(c, t0 + kPairSteps);
#pragma unroll
            for (int k = 0; k < kPairSteps; ++k) {
                ov[k] = step(False(), False(), t0 + kPairSteps, k, buf[kPairSteps + k], b);
                refill(kPairSteps + k);
            }
            flush_bulk(t0 + kPairSteps, ov);
        });
    };
    // whole iterations: checked until the first check-free one (t0 >= 16,
    // both blocks within quad_bulk_last), check-free while they last, checked
    // after; three loops, so that no loop body branches
    const int bulk_last = quad_bulk_last(R, c0);
    int t0 = 0;
    for (; t0 + kQuadAhead <= R && !(t0 >= kPairSteps && t0 + kPairSteps <= bulk_last);
         t0 += kQuadAhead)
        checked_iteration(t0);
    for (; t0 + kQuadAhead <= R && t0 + kPairSteps <= bulk_last; t0 += kQuadAhead)
        bulk_iteration(t0);
    for (; t0 + kQuadAhead <= R; t0 += kQuadAhead) checked_iteration(t0);
#pragma unroll
    for (int h = 0; h < kQuadAhead; h += kPairSteps) {  // tail: steps t0 .. R - 1
        const QuadBlock b = quad_block(c, t0 + h);
        T ov[kPairSteps];
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k)
            if (t0 + h + k < R) ov[k] = step(True(), True(), t0 + h, k, buf[h + k], b);
        flush_checked(t0 + h, R - t0 - h, ov);
    }
    for (int r = R - q.minlag; r < R; ++r) dst_lane[size_t(r) * n] = T(0);  // the +0 rows
}

// The wide two-stage pass (float32; see WideLane), addressed and pipelined as
// the quad kernel: scalar row addresses plus per-lane offsets, source loads
// kWideAhead - 1 steps ahead in a register ring, flush stores at the end of
// each 16-step block, three loops of 32-step iterations so that no loop body
// branches.
template <typename T>
__global__ void __launch_bounds__(kLanes)
    wide_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li,
                int waves_per_plane) {
    using P4 = T __attribute__((ext_vector_type(4)));
    using GP4 = const __attribute__((address_space(1))) P4;
    using GT = __attribute__((address_space(1))) T;
    using False = std::integral_constant<bool, false>;
    using True = std::integral_constant<bool, true>;
    __shared__ T ring[kWideRing];
    const int lane = threadIdx.x;
    const size_t plane = blockIdx.x / waves_per_plane;
    int c0;
    const WideLane wl = make_wide_lane(n, Li, int(blockIdx.x % waves_per_plane), lane, c0);
    const WideConst c = make_wide_const(wl, R);
    // wave-uniform (scalar) flush-row lags
    const int mlE = __builtin_amdgcn_readfirstlane(wide_mlE(c0));
    const int mlF = __builtin_amdgcn_readfirstlane(wide_mlF(c0));
    const size_t row_bytes = size_t(n) * sizeof(T);
    const GByte *gsrc = (const GByte *)(in + plane * size_t(R) * n);
    GByte *gdst = (GByte *)(out + plane * size_t(R) * n);
    T *dst_plane = out + plane * size_t(R) * n;
    const uint32_t src_off = uint32_t(wl.src) * uint32_t(sizeof(T));
    const uint32_t offE = uint32_t((size_t(mlE - wl.minE) * n + size_t(wl.ocE)) * sizeof(T));
    const uint32_t offF = uint32_t((size_t(mlF - wl.minF) * n + size_t(wl.ocF)) * sizeof(T));
    for (int e = lane; e < kWideRing; e += kLanes) ring[e] = T(0);
    __syncthreads();

    WideState<T> st;
    st.accP = st.accQ = st.prevP = st.prevQ = st.hP = st.hQ0 = st.hQ1 = st.hQ2 = T(0);
    st.accE = st.accF = T(0);
    // step t0 + k: both stages, the delay-line writes, the reads of the flush values
    auto step = [&](auto checked, auto mirrored, int t0, int k, const P4 &p, const WideBlock &b,
                    T &oE, T &oF) {
        constexpr bool kChecked = decltype(checked)::value;
        T vE, vF;
        wide_step<kChecked, T>(c, t0 + k, p.x, p.y, p.z, p.w, st, vE, vF);
        ring[b.wE + 32 * k] = vE;
        ring[b.wF + 32 * k] = vF;
        if (decltype(mirrored)::value) {
            ring[b.mE + 32 * k] = vE;
            ring[b.mF + 32 * k] = vF;
        }
        oE = ring[b.rE + 32 * k];
        oF = ring[b.rF + 32 * k];
    };
    // a check-free block's flushes: E rows R - 1 - t0 - k - minE (stored at
    // row R - 1 - t0 - k - mlE plus the lane's offset), F likewise
    auto flush_bulk = [&](int t0, const T *oE, const T *oF) {
        uint32_t e = offE, f = offF;
        asm volatile("" : "+v"(e), "+v"(f));
        GByte *rowE = gdst + size_t(R - 1 - t0 - mlE) * row_bytes;
        GByte *rowF = gdst + size_t(R - 1 - t0 - mlF) * row_bytes;
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) {
            if (k > 0) {
                rowE -= row_bytes;
                rowF -= row_bytes;
                asm volatile("" : "+s"(rowE));
                asm volatile("" : "+s"(rowF));
            }
            *reinterpret_cast<GT *>(rowE + e) = oE[k];
            *reinterpret_cast<GT *>(rowF + f) = oF[k];
        }
    };
    // a checked block's flushes, steps t0 .. t0 + count - 1
    auto flush_checked = [&](int t0, int count, const T *oE, const T *oF) {
#pragma unroll
        for (int k = 0; k < kPairSteps; ++k) {
            const int t = t0 + k;
            if (k < count && t <= c.flushE)
                dst_plane[size_t(R - 1 - t - wl.minE) * n + wl.ocE] = oE[k];
            if (k < count && t <= c.flushF)
                dst_plane[size_t(R - 1 - t - wl.minF) * n + wl.ocF] = oF[k];
        }
    };

    // the register ring of source vectors, as in the quad kernel but 16
    // vectors deep (64 registers; at 32 the allocator splits the slots)
    P4 buf[kWideAhead];
    {
        uint32_t off = src_off;
        asm volatile("" : "+v"(off));
#pragma unroll
        for (int k = 0; k < kWideAhead - 1; ++k) {
            const int y = R - 1 - k;
            buf[k] = *reinterpret_cast<GP4 *>(gsrc + size_t(y < 0 ? 0 : y) * row_bytes + off);
        }
    }
    auto refill_rows = [&](auto clamped, int t0, auto body) {
        const int y0 = R - 1 - t0 - (kWideAhead - 1);
        uint32_t off = src_off;
        asm volatile("" : "+v"(off));
        const GByte *row = gsrc + size_t(y0 < 0 ? 0 : y0) * row_bytes;
        auto refill = [&](int k) {
            if (k > 0) {
                if (decltype(clamped)::value)
                    row = y0 - k >= 0 ? row - row_bytes : row;
                else
                    row -= row_bytes;
                asm volatile("" : "+s"(row));
            }
            buf[(k + kWideAhead - 1) % kWideAhead] = *reinterpret_cast<GP4 *>(row + off);
            __builtin_amdgcn_sched_barrier(0);
        };
        body(refill);
    };
    auto checked_iteration = [&](int t0) {
        refill_rows(True(), t0, [&](auto &refill) {
#pragma unroll
            for (int h = 0; h < kQuadAhead; h += kPairSteps) {
                const WideBlock b = wide_block(c, t0 + h);
                T oE[kPairSteps], oF[kPairSteps];
#pragma unroll
                for (int k = 0; k < kPairSteps; ++k) {
                    step(True(), True(), t0 + h, k, buf[(h + k) % kWideAhead], b, oE[k], oF[k]);
                    refill(h + k);
                }
                flush_checked(t0 + h, kPairSteps, oE, oF);
            }
        });
    };
    auto bulk_iteration = [&](int t0) {
        refill_rows(False(), t0, [&](auto &refill) {  // rows y0 - 31 >= 0: see wide_bulk_last
            T oE[kPairSteps], oF[kPairSteps];
            WideBlock b = wide_block(c, t0);
#pragma unroll
            for (int k = 0; k < kPairSteps; ++k) {
                step(False(), True(), t0, k, buf[k % kWideAhead], b, oE[k], oF[k]);
                refill(k);
            }
            flush_bulk(t0, oE, oF);
            b = wide_block(c, t0 + kPairSteps);
#pragma unroll
            for (int k = 0; k < kPairSteps; ++k) {
                step(False(), False(), t0 + kPairSteps, k, buf[(kPairSteps + k) % kWideAhead], b,
                     oE[k], oF[k]);
                refill(kPairSteps + k);
            }
            flush_bulk(t0 + kPairSteps, oE, oF);
        });
    };
    const int bulk_last = wide_bulk_last(R, c0);
    int t0 = 0;
    for (; t0 + kQuadAhead <= R && !(t0 >= kPairSteps && t0 <= bulk_last); t0 += kQuadAhead)
        checked_iteration(t0);
    for (; t0 + kQuadAhead <= R && t0 <= bulk_last; t0 += kQuadAhead) bulk_iteration(t0);
    for (; t0 + kQuadAhead <= R; t0 += kQuadAhead) checked_iteration(t0);
    // tail: steps t0 .. R - 1 (up to 31, more than the ring holds ahead, so
    // the refills go on, clamped into the rows)
    refill_rows(True(), t0, [&](auto &refill) {
#pragma unroll
        for (int h = 0; h < kQuadAhead; h += kPairSteps) {
            const WideBlock b = wide_block(c, t0 + h);
            T oE[kPairSteps], oF[kPairSteps];
#pragma unroll
            for (int k = 0; k < kPairSteps; ++k) {
                if (t0 + h + k < R)
                    step(True(), True(), t0 + h, k, buf[(h + k) % kWideAhead], b, oE[k], oF[k]);
                refill(h + k);
            }
            flush_checked(t0 + h, R - t0 - h, oE, oF);
        }
    });
    // the +0 rows above each type's first flushed row
    for (int r = R - wl.minE; r < R; ++r) dst_plane[size_t(r) * n + wl.ocE] = T(0);
    for (int r = R - wl.minF; r < R; ++r) dst_plane[size_t(r) * n + wl.ocF] = T(0);
}

// The last S stages (the first with Li0 groups, Ci = n / Li0 = 2^S), one
// single-wave work-group per 64 columns of a plane.
template <typename T, int S, bool kAllLive>
__global__ void __launch_bounds__(kLanes)
    tail_kernel(const T *__restrict__ in, T *__restrict__ out, int R, int n, int Li0,
                int waves_per_plane) {
    using P2 = typename Pair<T>::type;
    constexpr int kM0 = tail_ring_rows(S, 0);
    __shared__ T rbuf[tail_rows_total(S) * kPitch];
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
    // kChecked: some stage's row may lie outside the rows (the first and last
    // steps); a block with every stage inside (tail_block_free) runs without
    // the range branches and selects
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
        if constexpr (kChecked) {
            T val;
            int row;
            if (step<T>(w, ln, sk, p.x, p.y, prev1, acc0, val, row))
                rbuf[(row & (kM0 - 1)) * kPitch + lane] = val;
        } else {  // the lane's row sk - lag: even  0 + p0 - prev1 + acc, odd  0 + p1 - p0 + acc
            const T x0 = ln.odd ? p.y : p.x, z0 = ln.odd ? p.x : prev1;
            const T v0 = ((T(0) + x0) - z0) + acc0;
            acc0 = v0;
            rbuf[((sk - ln.lag) & (kM0 - 1)) * kPitch + lane] = v0;
        }
        prev1 = p.y;
        // later stages: row x = sk + t is wave-uniform
#pragma unroll
        for (int t = 1; t < S; ++t) {
            const int x = sk + t;
            if (kChecked && (x < 0 || x >= R)) continue;
            // unchecked: both operands and acc enter (see tail_free_top)
            const T v = kChecked
                            ? tail_combine<T>(o[t - 1], a[t - 1], b[t - 1], acc[t - 1], x + 1 < R)
                            : ((T(0) + a[t - 1]) - b[t - 1]) + acc[t - 1];
            acc[t - 1] = v;
            if (t < S - 1)
                rbuf[(tail_ring_base(S, t) + (x & (tail_ring_rows(S, t) - 1))) * kPitch +
                     lane] = v;
            else if (kAllLive || tl[t - 1].live)
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
        if (kAllLive && tail_block_free(s, R, S, w.maxlag)) {
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
// The wide two-stage pass: grid = planes * (n / 128) waves (float32 only).
template <typename T>
static void launch_wide(const T *src, T *dst, size_t B, size_t n, int R, int Li, hipStream_t s) {
    if constexpr (sizeof(T) == 4) {
        const int wpp = int(n / 128);
        hipLaunchKernelGGL((iadrt_native::wide_kernel<T>), dim3(unsigned(B * 4 * wpp)),
                           dim3(iadrt_native::kLanes), 0, s, src, dst, R, int(n), Li, wpp);
    }
}

template <typename T>
static int run_native(const T *in, T *out, T *ws, size_t B, size_t n, hipStream_t s) {
    const size_t R = 2 * n - 1;
    const int L = num_iters(n);
    if (L == 0) {  // n = 1: the inverse of a single-element transform is a copy
        const hipError_t e =
            hipMemcpyAsync(out, in, B * 4 * R * n * sizeof(T), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            set_last_error(e);
            return ADRT_HIP_ERR_DEVICE;
        }
        return launch_status();
    }
    // the last S stages run fused (S = 1: no tail kernel); before them,
    // pairs of paired stages run as one pass each (not in STREAMING)
    const int variant = fused_variant(ADRT_HIP_OP_IADRT, n);
    const int S = tail_stages<T>(L, variant);
    // two stages per pass unless STREAMING, and only where every store offset
    // of a plane fits 32 bits (else each stage runs as its own paired pass)
    const bool quad = variant != ADRT_HIP_VARIANT_STREAMING && iadrt_native::quad_narrow<T>(n);
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
            // float32 with Co >= 64: the wide pass
            if (sizeof(T) == 4 && int(n) / Li / 2 >= iadrt_native::kLanes)
                launch_wide<T>(src, dst, B, n, int(R), Li, s);
            else
                hipLaunchKernelGGL((iadrt_native::quad_kernel<T>), dim3(grid),
                                   dim3(iadrt_native::kLanes), 0, s, src, dst, int(R), int(n),
                                   Li, waves_per_plane);
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
        // every column of every wave exists when n is a multiple of 64
        if (n % iadrt_native::kLanes == 0) {
            switch (S) {
                case 2: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 2, true>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                case 3: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 3, true>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                case 4: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 4, true>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                case 5: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 5, true>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                default: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 6, true>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            }
        } else {  // n < 64: one wave with n live lanes, L = log2 n <= 5 stages
            switch (S) {
                case 2: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 2, false>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                case 3: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 3, false>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                case 4: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 4, false>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
                default: hipLaunchKernelGGL((iadrt_native::tail_kernel<T, 5, false>), g, blk, 0, s, src, dst, r, nn, Li, waves_per_plane); break;
            }
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
