// Host emulation of the iadrt stage and fused-tail kernels
// (adrt_amd/csrc/iadrt.hip).
//
// Runs every wavefront of every stage on the CPU with the kernel's own
// per-lane step function: lanes advance in lockstep, step by step, through a
// 64-row ring exactly sized as the kernel's LDS ring (poisoned with NaN), and
// row s is flushed at step s with the kernel's unconditional stores.  Pair loads use the kernel's clamped rows.
// Buffers are exactly sized, so an out-of-bounds access aborts under
// AddressSanitizer.  Usage:  iadrt_emu f32|f64 B n in.bin out.bin [tail_stages] [quad]
// (tail_stages: how many final stages the fused tail runs; default: the
// library's choice; 0: none; quad 0: every paired stage in its own pass,
// default 1: two paired stages per pass where both precede the tail).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define ADRT_HOST_EMULATION 1
#include "../../adrt_amd/csrc/iadrt.hip"

using namespace adrt_hip;
using namespace adrt_hip::iadrt_native;

template <typename T>
static void emulate_stage(const std::vector<T> &in, std::vector<T> &out, int planes, int R, int n,
                          int Li) {
    const int waves = (n + kLanes - 1) / kLanes;
    std::vector<T> ring(size_t(kRing) * kPitch);
    for (int plane = 0; plane < planes; ++plane) {
        const T *src = in.data() + size_t(plane) * R * n;
        T *dst = out.data() + size_t(plane) * R * n;
        for (int wv = 0; wv < waves; ++wv) {
            const Wave w = make_wave(R, n, Li, wv);
            Lane ln[kLanes];
            T acc[kLanes], prev1[kLanes];
            for (int l = 0; l < kLanes; ++l) {
                ln[l] = make_lane(w, l);
                acc[l] = T(0);
                prev1[l] = T(0);
            }
            for (auto &x : ring) x = T(NAN);
            for (int s = w.top; s >= 0; --s) {
                for (int l = 0; l < kLanes; ++l) {
                    int r = s + ln[l].roff;
                    r = r < 0 ? 0 : (r > R - 1 ? R - 1 : r);
                    const size_t at = size_t(r) * n + (ln[l].live ? ln[l].pc : 0);
                    const T p0 = src[at], p1 = src[at + 1];
                    T val;
                    int row;
                    if (step<T>(w, ln[l], s, p0, p1, prev1[l], acc[l], val, row))
                        ring[size_t(row % kRing) * kPitch + l] = val;
                    prev1[l] = p1;
                }
                // the kernel's unconditional flush: steps above the last row
                // write row R-1 (overwritten at step R-1), lanes past n repeat
                // the last live lane
                for (int l = 0; l < kLanes; ++l) {
                    const int fl = w.ch0 + l < n ? l : n - 1 - w.ch0;
                    const int fr = s < R ? s : R - 1;
                    dst[size_t(fr) * n + w.ch0 + fl] = ring[size_t(s % kRing) * kPitch + fl];
                }
            }
        }
    }
}

// The paired stage kernel (Co >= 32): lanes in lockstep, the odd lanes'
// 32-row rings of their own columns, tagged with the row each slot holds
// (the +0 start is its own tag); every flush checks that it reads its row,
// or a +0 row the lane skipped.
template <typename T>
static void emulate_pair_stage(const std::vector<T> &in, std::vector<T> &out, int planes, int R,
                               int n, int Li) {
    const int waves = n / kLanes;
    const int kZero = -(1 << 30);
    for (int plane = 0; plane < planes; ++plane) {
        const T *src = in.data() + size_t(plane) * R * n;
        T *dst = out.data() + size_t(plane) * R * n;
        for (int wv = 0; wv < waves; ++wv) {
            PairLane pp[kLanes];
            int c0 = 0;
            for (int l = 0; l < kLanes; ++l) pp[l] = make_pair_lane(n, Li, wv, l, c0);
            std::vector<T> ring(size_t(kHalf) * kHalf, T(0));
            std::vector<int> tag(ring.size(), kZero);
            T acc[kLanes], prev1[kLanes];
            for (int l = 0; l < kLanes; ++l) {
                acc[l] = T(0);
                prev1[l] = T(0);
                if (pp[l].pc < 0 || pp[l].pc + 1 >= n || pp[l].oc < 0 || pp[l].oc >= n)
                    std::abort();
            }
            for (int t = 0; t < R; ++t) {
                const int y = R - 1 - t;
                for (int l = 0; l < kLanes; ++l) {
                    const size_t at = size_t(y) * n + pp[l].pc;
                    const T p0 = src[at], p1 = src[at + 1];
                    T val;
                    int row;
                    if (pair_step<T>(pp[l], y, R, p0, p1, prev1[l], acc[l], val, row)) {
                        if (!pp[l].odd) {
                            dst[size_t(row) * n + pp[l].oc] = val;
                        } else {
                            const size_t slot = size_t(row & (kHalf - 1)) * kHalf + (l & (kHalf - 1));
                            ring[slot] = val;
                            tag[slot] = row;
                        }
                    }
                    prev1[l] = p1;
                }
                const int f = y - 1 - c0;
                if (f < 0) continue;
                for (int l = kHalf; l < kLanes; ++l) {
                    const size_t slot = size_t(f & (kHalf - 1)) * kHalf + (l & (kHalf - 1));
                    const bool skipped = tag[slot] == kZero && f > R - 2 - pp[l].col;
                    if (tag[slot] != f && !skipped) {
                        std::fprintf(stderr, "stale pair flush: row %d lane %d\n", f, l);
                        std::abort();
                    }
                    dst[size_t(f) * n + pp[l].oc] = ring[slot];
                }
            }
            for (int l = kHalf; l < kLanes; ++l)
                for (int r = R - 1 - c0; r < R; ++r) dst[size_t(r) * n + pp[l].oc] = T(0);
        }
    }
}

// The register ring of source rows that the quad and wide kernels keep:
// `depth` slots; the prologue loads steps 0 .. depth - 2 into slots 0 ..
// depth - 2; each step t reads slot t % depth and then, where the kernel
// refills (whole iterations; in the wide kernel the final partial one too),
// refills slot (t - 1) % depth with step t + depth - 1.  `use` checks that
// the slot read at step t holds step t.
struct RowRing {
    std::vector<int> step;
    explicit RowRing(int depth) : step(size_t(depth), -1) {
        for (int k = 0; k < depth - 1; ++k) step[size_t(k)] = k;
    }
    void use(int t, bool refill) {
        const int d = int(step.size());
        if (step[size_t(t % d)] != t) {
            std::fprintf(stderr, "register ring: step %d reads slot %d holding step %d\n", t, t % d,
                         step[size_t(t % d)]);
            std::abort();
        }
        if (refill) step[size_t((t + d - 1) % d)] = t + d - 1;
    }
};

// Two paired stages in one pass (the quad kernel): lanes in lockstep, in the
// kernel's blocks of 16 steps (two per iteration of kQuadAhead steps, the
// first with mirror writes where check-free); a check-free block runs the unchecked step
// only after verifying, for every lane and step, each condition that step
// drops.  The partner exchange is lane ^ 1.  The delay lines are the
// kernel's ring (its element indices, its mirror writes; blocks without
// mirror writes verify that none was due), every element tagged with the
// step that wrote it, and every flush checks that it reads the value of step
// t - delay, or an element no step has written when t - delay < 0 (the +0
// rows above the lane's first row).  Check-free blocks store through the
// kernel's addressing (row y - mlmax plus the lane's offset, which must fit
// 32 bits where the kernel takes it so).
template <typename T>
static void emulate_quad_stage(const std::vector<T> &in, std::vector<T> &out, int planes, int R,
                               int n, int Li) {
    const int waves = n / kLanes;
    const int kZero = -(1 << 30);
    const bool narrow = quad_narrow<T>(size_t(n));
    for (int type = 0; type < 3; ++type)  // regions in order, disjoint, inside the ring
        if (quad_region(type) + 16 * (quad_slots(type) + (type ? kQuadMirror : 0)) >
            quad_region(type + 1))
            std::abort();
    if (quad_region(3) + 16 * (quad_slots(3) + kQuadMirror) != kQuadRing) std::abort();
    for (int plane = 0; plane < planes; ++plane) {
        const T *src = in.data() + size_t(plane) * R * n;
        T *dst = out.data() + size_t(plane) * R * n;
        for (int wv = 0; wv < waves; ++wv) {
            QuadLane q[kLanes];
            QuadConst c[kLanes];
            int c0 = 0;
            for (int l = 0; l < kLanes; ++l) {
                q[l] = make_quad_lane(n, Li, wv, l, c0);
                c[l] = make_quad_const(q[l], R);
                if (q[l].p.pc < 0 || q[l].p.pc + 1 >= n || q[l].oc < 0 || q[l].oc >= n ||
                    q[l].i < 0 || q[l].i > 15)
                    std::abort();
                if (c[l].mirror != (q[l].type ? 16 * (c[l].mask + 1) : 0)) std::abort();
            }
            const int mlmax = quad_mlmax(c0);
            for (int l = 0; l < kLanes; ++l) {
                if (q[l].minlag > mlmax) std::abort();
                if (narrow && quad_store_off(q[l], n, mlmax) * sizeof(T) > 0xffffffffull)
                    std::abort();
            }
            std::vector<T> ring(kQuadRing, T(0));
            std::vector<int> tag(kQuadRing, kZero);
            QuadState<T> st[kLanes];
            for (auto &x : st) x.acc = x.prev1 = x.h0 = x.h1 = x.h2 = x.acc2 = T(0);
            const int bulk_last = quad_bulk_last(R, c0);
            RowRing rows(kQuadAhead);
            // iterations of kQuadAhead steps (two blocks), check-free as a whole
            // (the kernel's three loops take the iterations where this holds,
            // an interval of t0, in the middle one)
            for (int t0 = 0; t0 < R; t0 += kPairSteps) {
                const int it0 = t0 - t0 % kQuadAhead;
                const bool bulk = it0 + kQuadAhead <= R && it0 >= kPairSteps &&
                                  it0 + kPairSteps <= bulk_last;
                const bool mirrored = !bulk || t0 == it0;
                QuadBlock b[kLanes];
                for (int l = 0; l < kLanes; ++l) {
                    b[l] = quad_block(c[l], t0);
                    if (!mirrored && b[l].m != b[l].w) {
                        std::fprintf(stderr, "mirror write skipped: step %d lane %d\n", t0, l);
                        std::abort();
                    }
                    for (int k = 0; bulk && k < kPairSteps; ++k) {
                        const int t = t0 + k;
                        if (t == 0 || t > c[l].made_lim || t > c[l].made2_lim ||
                            t > c[l].flush_lim) {
                            std::fprintf(stderr, "check-free block with a check due: step %d lane %d\n",
                                         t, l);
                            std::abort();
                        }
                    }
                }
                for (int k = 0; k < kPairSteps && t0 + k < R; ++k) {
                    const int t = t0 + k, y = R - 1 - t;
                    rows.use(t, it0 + kQuadAhead <= R);  // no refills in the final partial iteration
                    T send[kLanes];
                    for (int l = 0; l < kLanes; ++l) {
                        const size_t at = size_t(y) * n + q[l].p.pc;
                        const T p0 = src[at], p1 = src[at + 1];
                        send[l] = bulk ? quad_stage_k<false, T>(c[l], t, p0, p1, st[l])
                                       : quad_stage_k<true, T>(c[l], t, p0, p1, st[l]);
                    }
                    for (int l = 0; l < kLanes; ++l) {
                        const T v2 = bulk ? quad_stage_k1<false, T>(c[l], t, send[l ^ 1], st[l])
                                          : quad_stage_k1<true, T>(c[l], t, send[l ^ 1], st[l]);
                        ring.at(size_t(b[l].w + 16 * k)) = v2;
                        tag.at(size_t(b[l].w + 16 * k)) = t;
                        if (mirrored) {
                            ring.at(size_t(b[l].m + 16 * k)) = v2;
                            tag.at(size_t(b[l].m + 16 * k)) = t;
                        }
                        const size_t e = size_t(b[l].r + 16 * k);
                        const T ov = ring.at(e);
                        if (t > c[l].flush_lim) continue;
                        const int want = t - c[l].delay;
                        if (want >= 0 ? tag[e] != want : tag[e] != kZero) {
                            std::fprintf(stderr, "stale quad flush: step %d lane %d\n", t, l);
                            std::abort();
                        }
                        const int f = y - q[l].minlag;
                        size_t at = size_t(f) * n + q[l].oc;
                        if (bulk) {  // the kernel's row y - mlmax plus the lane's offset
                            const size_t at2 = size_t(y - mlmax) * n + quad_store_off(q[l], n, mlmax);
                            if (y < mlmax || at2 != at) std::abort();
                        }
                        dst[at] = ov;
                    }
                }
            }
            for (int l = 0; l < kLanes; ++l)
                for (int r = R - q[l].minlag; r < R; ++r) dst[size_t(r) * n + q[l].oc] = T(0);
        }
    }
}

// The wide two-stage pass (float32, Co >= 64): lanes in lockstep in the
// kernel's 32-step iterations of two 16-step blocks; a check-free iteration
// runs the unchecked step only after verifying every condition it drops.
// Each lane holds both chains of its two stage-(k+1) chains (no exchange).
// The two delay lines per lane are the kernel's ring (its element indices and
// mirror writes; blocks without mirror writes verify none was due), every
// element tagged with the step that wrote it; each flush checks it reads the
// value of step t - delay, or an unwritten +0 element when t - delay < 0.
// Check-free blocks store through the kernel's addressing (row y - ml plus
// the lane's offset, which must fit 32 bits).
template <typename T>
static void emulate_wide_stage(const std::vector<T> &in, std::vector<T> &out, int planes, int R,
                               int n, int Li) {
    const int waves = n / 128;
    const int kZero = -(1 << 30);
    if (!quad_narrow<T>(size_t(n))) std::abort();  // the library takes the wide pass only then
    // regions in order, disjoint, inside the ring, C and D 32 elements (mod 64) after A and B
    const int order[4] = {0, 2, 1, 3};
    for (int i = 0; i < 3; ++i) {
        const int t = order[i], u = order[i + 1];
        if (wide_region(t) + 32 * (wide_slots(t) + (t ? 16 : 0)) > wide_region(u)) std::abort();
    }
    if (wide_region(3) + 32 * (wide_slots(3) + 16) != kWideRing) std::abort();
    if ((wide_region(2) - wide_region(0)) % 64 != 32 || (wide_region(3) - wide_region(1)) % 64 != 32)
        std::abort();
    for (int plane = 0; plane < planes; ++plane) {
        const T *src = in.data() + size_t(plane) * R * n;
        T *dst = out.data() + size_t(plane) * R * n;
        for (int wv = 0; wv < waves; ++wv) {
            WideLane w[kLanes];
            WideConst c[kLanes];
            int c0 = 0;
            for (int l = 0; l < kLanes; ++l) {
                w[l] = make_wide_lane(n, Li, wv, l, c0);
                c[l] = make_wide_const(w[l], R);
                if (w[l].src < 0 || w[l].src + 3 >= n || w[l].src % 4 != 0 || w[l].ocE < 0 ||
                    w[l].ocE >= n || w[l].ocF < 0 || w[l].ocF >= n)
                    std::abort();
            }
            const int mlE = wide_mlE(c0), mlF = wide_mlF(c0);
            for (int l = 0; l < kLanes; ++l) {
                if (w[l].minE > mlE || w[l].minF > mlF) std::abort();
                const size_t oE = (size_t(mlE - w[l].minE) * n + w[l].ocE) * sizeof(T);
                const size_t oF = (size_t(mlF - w[l].minF) * n + w[l].ocF) * sizeof(T);
                if (oE > 0xffffffffull || oF > 0xffffffffull) std::abort();
            }
            std::vector<T> ring(kWideRing, T(0));
            std::vector<int> tag(kWideRing, kZero);
            WideState<T> st[kLanes];
            for (auto &x : st) {
                x.accP = x.accQ = x.prevP = x.prevQ = x.hP = x.hQ0 = x.hQ1 = x.hQ2 = T(0);
                x.accE = x.accF = T(0);
            }
            const int bulk_last = wide_bulk_last(R, c0);
            RowRing rows(kWideAhead);
            auto put = [&](size_t e, T v, int t) {
                ring.at(e) = v;
                tag.at(e) = t;
            };
            for (int t0 = 0; t0 < R; t0 += kPairSteps) {
                const int it0 = t0 - t0 % kQuadAhead;
                const bool bulk = it0 + kQuadAhead <= R && it0 >= kPairSteps && it0 <= bulk_last;
                const bool mirrored = !bulk || t0 == it0;
                WideBlock b[kLanes];
                for (int l = 0; l < kLanes; ++l) {
                    b[l] = wide_block(c[l], t0);
                    if (!mirrored && (b[l].mE != b[l].wE || b[l].mF != b[l].wF)) {
                        std::fprintf(stderr, "wide mirror write skipped: step %d lane %d\n", t0, l);
                        std::abort();
                    }
                    for (int k = 0; bulk && k < kPairSteps; ++k) {
                        const int t = t0 + k;
                        if (t == 0 || t > c[l].limP || t > c[l].limQ || t > c[l].limE ||
                            t > c[l].limF || t > c[l].flushE || t > c[l].flushF) {
                            std::fprintf(stderr, "wide check-free block with a check due: step %d lane %d\n",
                                         t, l);
                            std::abort();
                        }
                    }
                }
                for (int k = 0; k < kPairSteps && t0 + k < R; ++k) {
                    const int t = t0 + k, y = R - 1 - t;
                    rows.use(t, true);
                    for (int l = 0; l < kLanes; ++l) {
                        const T *p = src + size_t(y) * n + w[l].src;
                        T vE, vF;
                        if (bulk)
                            wide_step<false, T>(c[l], t, p[0], p[1], p[2], p[3], st[l], vE, vF);
                        else
                            wide_step<true, T>(c[l], t, p[0], p[1], p[2], p[3], st[l], vE, vF);
                        put(size_t(b[l].wE + 32 * k), vE, t);
                        put(size_t(b[l].wF + 32 * k), vF, t);
                        if (mirrored) {
                            put(size_t(b[l].mE + 32 * k), vE, t);
                            put(size_t(b[l].mF + 32 * k), vF, t);
                        }
                        const int line[2][4] = {{b[l].rE, c[l].flushE, c[l].dE, w[l].minE},
                                                {b[l].rF, c[l].flushF, c[l].dF, w[l].minF}};
                        const int oc[2] = {w[l].ocE, w[l].ocF}, ml[2] = {mlE, mlF};
                        for (int h = 0; h < 2; ++h) {
                            const size_t e = size_t(line[h][0] + 32 * k);
                            const T ov = ring.at(e);
                            if (t > line[h][1]) continue;
                            const int want = t - line[h][2];
                            if (want >= 0 ? tag[e] != want : tag[e] != kZero) {
                                std::fprintf(stderr, "stale wide flush: step %d lane %d line %d\n", t, l, h);
                                std::abort();
                            }
                            const size_t at = size_t(y - line[h][3]) * n + oc[h];
                            if (bulk) {
                                const size_t at2 = size_t(y - ml[h]) * n +
                                                   (size_t(ml[h] - line[h][3]) * n + oc[h]);
                                if (y < ml[h] || at2 != at) std::abort();
                            }
                            dst[at] = ov;
                        }
                    }
                }
            }
            for (int l = 0; l < kLanes; ++l) {
                for (int r = R - w[l].minE; r < R; ++r) dst[size_t(r) * n + w[l].ocE] = T(0);
                for (int r = R - w[l].minF; r < R; ++r) dst[size_t(r) * n + w[l].ocF] = T(0);
            }
        }
    }
}

// The fused tail kernel: the last S stages, the first with Li0 groups.  Lanes
// run in lockstep, in the kernel's order within a step: every later stage's
// ring reads, then stage 0, then the later stages' results.  The ring buffer
// has exactly the kernel's rows, and every slot carries the row it holds:
// every ring read that enters a result checks that it reads the row it wants
// (a stale or unwritten slot aborts).
template <typename T>
static void emulate_tail(const std::vector<T> &in, std::vector<T> &out, int planes, int R, int n,
                         int Li0, int S) {
    const int waves = (n + kLanes - 1) / kLanes;
    const int P = kPitch;
    const int M0 = tail_ring_rows(S, 0);
    std::vector<T> rbuf(size_t(tail_rows_total(S)) * P);
    std::vector<int> tag(rbuf.size());
    auto check = [&](int base, int M, int row, int col) {
        if (tag.at(size_t(base + (row & (M - 1))) * P + col) != row) {
            std::fprintf(stderr, "stale ring read: ring base %d row %d col %d\n", base, row, col);
            std::abort();
        }
    };
    for (int plane = 0; plane < planes; ++plane) {
        const T *src = in.data() + size_t(plane) * R * n;
        T *dst = out.data() + size_t(plane) * R * n;
        for (int wv = 0; wv < waves; ++wv) {
            const Wave w = make_wave(R, n, Li0, wv);
            if (w.c0 != 0 || w.maxlag != w.Ci / 2 - 1 || w.Ci != (1 << S)) std::abort();
            Lane ln[kLanes];
            std::vector<TailLane> tl(size_t(S) * kLanes);
            T acc0[kLanes], prev1[kLanes];
            std::vector<T> acc(size_t(S) * kLanes, T(0));
            for (int l = 0; l < kLanes; ++l) {
                ln[l] = make_lane(w, l);
                acc0[l] = T(0);
                prev1[l] = T(0);
                for (int t = 1; t < S; ++t)
                    tl[size_t(t) * kLanes + l] = make_tail_lane(n, n / (Li0 << t), w.ch0, l, R);
            }
            for (auto &x : rbuf) x = T(NAN);
            for (auto &x : tag) x = -1 << 20;
            // the kernel's blocks of kTailSteps steps: a whole block that it
            // runs without range checks must have every later stage's row
            // inside [0, R), with x + 1 < R and both operands of every live
            // lane entering (the unchecked combine adds them all)
            const bool all_live = n % kLanes == 0;
            for (int sb = w.top; sb - (kTailSteps - 1) >= -(S - 1); sb -= kTailSteps)
                if (all_live && tail_block_free(sb, R, S, w.maxlag))
                    for (int k = 0; k < kTailSteps; ++k) {
                        const int sk = sb - k;
                        for (int l = 0; l < kLanes; ++l) {  // the branch-free stage 0
                            const int row = sk - ln[l].lag;
                            if (!ln[l].live || row < 0 || row + 1 >= R || sk + ln[l].roff >= R) {
                                std::fprintf(stderr, "tail unchecked block with a stage-0 check due: s %d lane %d\n",
                                             sk, l);
                                std::abort();
                            }
                        }
                        for (int t = 1; t < S; ++t) {
                            const int x = sb - k + t;
                            if (x < 0 || x + 1 >= R) std::abort();
                            for (int l = 0; l < kLanes; ++l) {
                                const TailLane &u = tl[size_t(t) * kLanes + l];
                                const TailOperands o = tail_operands(u, x, 0, 0);
                                if (u.live && (!o.use_a || !o.use_b)) {
                                    std::fprintf(stderr, "tail unchecked block drops an operand: x %d lane %d\n",
                                                 x, l);
                                    std::abort();
                                }
                            }
                        }
                    }
            for (int s = w.top; s >= -(S - 1); --s) {
                std::vector<TailOperands> ops(size_t(S) * kLanes);
                std::vector<T> va(size_t(S) * kLanes), vb(size_t(S) * kLanes);
                for (int t = 1; t < S; ++t) {
                    const int x = s + t;
                    const int base = tail_ring_base(S, t - 1), M = tail_ring_rows(S, t - 1);
                    for (int l = 0; l < kLanes; ++l) {
                        const TailLane &u = tl[size_t(t) * kLanes + l];
                        if (u.pc < 0 || u.pc + 1 > kLanes - 1) std::abort();
                        const TailOperands o = tail_operands(u, x, base, M - 1);
                        const size_t k = size_t(t) * kLanes + l;
                        ops[k] = o;
                        va[k] = rbuf.at(size_t(o.ia));
                        vb[k] = rbuf.at(size_t(o.ib));
                        if (!u.live || x < 0 || x >= R) continue;
                        const int ro = x + 1 + u.col;
                        if (o.use_a) check(base, M, u.odd ? ro : x, u.pc + (u.odd ? 1 : 0));
                        if (o.use_b) check(base, M, u.odd ? ro : x + 1, u.pc + (u.odd ? 0 : 1));
                    }
                }
                for (int l = 0; l < kLanes; ++l) {
                    int r = s + ln[l].roff;
                    r = r < 0 ? 0 : (r > R - 1 ? R - 1 : r);
                    const size_t at = size_t(r) * n + (ln[l].live ? ln[l].pc : 0);
                    const T p0 = src[at], p1 = src[at + 1];
                    T val;
                    int row;
                    if (step<T>(w, ln[l], s, p0, p1, prev1[l], acc0[l], val, row)) {
                        rbuf.at(size_t(row & (M0 - 1)) * P + l) = val;
                        tag.at(size_t(row & (M0 - 1)) * P + l) = row;
                    }
                    prev1[l] = p1;
                }
                for (int t = 1; t < S; ++t) {
                    const int x = s + t;
                    if (x < 0 || x >= R) continue;
                    const bool last = t == S - 1;
                    for (int l = 0; l < kLanes; ++l) {
                        const size_t k = size_t(t) * kLanes + l;
                        T &a = acc[k];
                        const T v = tail_combine<T>(ops[k], va[k], vb[k], a, x + 1 < R);
                        a = v;
                        if (!last) {
                            const size_t at =
                                size_t(tail_ring_base(S, t) + (x & (tail_ring_rows(S, t) - 1))) * P + l;
                            rbuf.at(at) = v;
                            tag.at(at) = x;
                        } else if (tl[k].live) {
                            dst[size_t(x) * n + w.ch0 + l] = v;
                        }
                    }
                }
            }
        }
    }
}

template <typename T>
static int run(int B, int n, const char *fin, const char *fout, int tail, bool quad) {
    const int R = 2 * n - 1;
    const size_t elems = size_t(B) * 4 * R * n;
    std::vector<T> a(elems), b(elems, T(NAN));
    FILE *f = std::fopen(fin, "rb");
    if (!f || std::fread(a.data(), sizeof(T), elems, f) != elems) return 2;
    std::fclose(f);
    const int L = num_iters(size_t(n));
    // tail < 0: the library's choice; 0 or 1: no fused tail
    const int S = tail < 0 ? tail_stages<T>(L) : (tail < L ? tail : L);
    const int k0 = L - (S >= 2 ? S : 0);
    int Li = 1;
    for (int k = 0; k < k0;) {
        const bool paired = n / Li / 2 >= kHalf;
        // the library's choice: two stages per pass where store offsets fit 32
        // bits, the wide pass for float32 with Co >= 64
        if (paired && quad && quad_narrow<T>(size_t(n)) && k + 1 < k0) {
            if (sizeof(T) == 4 && n / Li / 2 >= kLanes)
                emulate_wide_stage<T>(a, b, B * 4, R, n, Li);
            else
                emulate_quad_stage<T>(a, b, B * 4, R, n, Li);
            k += 2;
            Li *= 4;
        } else {
            if (paired)
                emulate_pair_stage<T>(a, b, B * 4, R, n, Li);
            else
                emulate_stage<T>(a, b, B * 4, R, n, Li);
            k += 1;
            Li *= 2;
        }
        std::swap(a, b);
        for (auto &x : b) x = T(NAN);
    }
    if (S >= 2) {
        emulate_tail<T>(a, b, B * 4, R, n, Li, S);
        std::swap(a, b);
    }
    f = std::fopen(fout, "wb");
    if (!f || std::fwrite(a.data(), sizeof(T), elems, f) != elems) return 4;
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 6 || argc > 8) {
        std::fprintf(stderr, "usage: %s f32|f64 B n in.bin out.bin [tail_stages] [quad 0|1]\n",
                     argv[0]);
        return 1;
    }
    const int B = std::atoi(argv[2]), n = std::atoi(argv[3]);
    const int tail = argc >= 7 ? std::atoi(argv[6]) : -1;
    const bool quad = argc == 8 ? std::atoi(argv[7]) != 0 : true;
    if (std::strcmp(argv[1], "f32") == 0) return run<float>(B, n, argv[4], argv[5], tail, quad);
    return run<double>(B, n, argv[4], argv[5], tail, quad);
}
