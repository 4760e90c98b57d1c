// Host emulation of the fused forward kernels (adrt_amd/csrc/fused_items.hpp).
//
// Compiles the exact per-lane load/store functions of Pass1/Pass2 for the
// CPU and runs every launch block by block: for each wavefront task all 64
// lanes load, the strip is computed with the whole-strip restatement (offsets
// without a valid halo poisoned with NaN), then all lanes store.  LDS is an
// exactly sized heap buffer filled with NaN before every block and X is sized
// exactly as the library's workspace, so any out-of-bounds access aborts
// under AddressSanitizer and any read of an unwritten or invalid slot shows
// up as NaN in the result.  Usage:
//   fused_emu f32|f64 B n in.bin out.bin [slots] [fwd|bdrt] [variant] [group_bytes] [rows]
// `slots` (a multiple of 8) caps the grid so that work-groups run several
// items each; 0 (default) is the library's one-item-per-work-group grid.
// `rows` > 0 runs the back-projection for output rows [0, rows) only, as
// adrt_hip_bdrt_rows does (the other rows stay NaN in the output).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define ADRT_HOST_EMULATION 1
#include "../../adrt_amd/csrc/fused_items.hpp"

using namespace adrt_hip;
using namespace adrt_hip::fused;


// Runs the persistent schedule of `grid` work-groups over `items` items, one
// item at a time, exactly as the kernels' loop assigns them.
template <typename P, typename T, typename... Args>
static void emulate_grid(int items, unsigned grid, std::vector<int> &seen, Args... args) {
    std::vector<T> lds(P::lds_elems > 0 ? P::lds_elems : 1);
    std::vector<T> buf0(size_t(kLanes) * P::R0 * kW), buf1(size_t(kLanes) * P::R1 * kW);
    auto *v0 = reinterpret_cast<T(*)[P::R0][kW]>(buf0.data());
    auto *v1 = reinterpret_cast<T(*)[P::R1][kW]>(buf1.data());
    seen.assign(size_t(items), 0);
    for (unsigned blk = 0; blk < grid; ++blk) {
        const Schedule sc = schedule(int(blk), int(grid), items);
        for (int idx = sc.idx; idx < sc.hi; idx += sc.step) {
            seen[size_t(idx)]++;
            const auto it = P::item(args..., idx);
            if (P::skip(it)) continue;  // the kernel's early return
            for (auto &x : lds) x = T(NAN);
            T *l = P::lds_elems > 0 ? lds.data() : nullptr;
            if constexpr (P::kStaged) {
                // staged load by the whole work-group, then every strip reads
                // its block before any strip stores (they share the LDS area)
                for (int tid = 0; tid < kThreads; ++tid) P::stage(it, l, tid);
                std::vector<T> keep0(size_t(P::tasks0) * buf0.size());
                for (int task = 0; task < P::tasks0; ++task) {
                    for (int lane = 0; lane < kLanes; ++lane) P::load0s(it, task, lane, l, v0[lane]);
                    std::memcpy(keep0.data() + size_t(task) * buf0.size(), buf0.data(), buf0.size() * sizeof(T));
                }
                for (auto &x : lds) x = T(NAN);
                for (int task = 0; task < P::tasks0; ++task) {
                    std::memcpy(buf0.data(), keep0.data() + size_t(task) * buf0.size(), buf0.size() * sizeof(T));
                    P::strip0_host(v0);
                    for (int lane = 0; lane < kLanes; ++lane) P::store0(it, task, lane, v0[lane], l);
                }
            } else if constexpr (P::kParts == 2) {
                // the kernel's two parts: level A everywhere; part 0 rows, the
                // first strips' level B loads; part 1 rows over them (LDS
                // poisoned in between), the other strips' loads
                std::vector<T> keepA(size_t(P::tasks0) * buf0.size());
                for (int task = 0; task < P::tasks0; ++task) {
                    for (int lane = 0; lane < kLanes; ++lane) P::load0(it, task, lane, v0[lane]);
                    P::strip0_host(v0);
                    std::memcpy(keepA.data() + size_t(task) * buf0.size(), buf0.data(), buf0.size() * sizeof(T));
                }
                for (int part = 0; part < 2; ++part) {
                    for (auto &x : lds) x = T(NAN);
                    for (int task = 0; task < P::tasks0; ++task) {
                        std::memcpy(buf0.data(), keepA.data() + size_t(task) * buf0.size(), buf0.size() * sizeof(T));
                        for (int lane = 0; lane < kLanes; ++lane)
                            P::store0_part(it, task, lane, v0[lane], l, part);
                    }
                    for (int task = part * P::kPartRows; task < (part + 1) * P::kPartRows; ++task) {
                        for (int lane = 0; lane < kLanes; ++lane) P::load1(it, task, lane, l, v1[lane]);
                        P::strip1_host(v1);
                        for (int lane = 0; lane < kLanes; ++lane) P::store1(it, task, lane, v1[lane], l);
                    }
                }
                continue;
            } else {
                for (int task = 0; task < P::tasks0; ++task) {
                    for (int lane = 0; lane < kLanes; ++lane) P::load0(it, task, lane, v0[lane]);
                    P::strip0_host(v0);
                    for (int lane = 0; lane < kLanes; ++lane) P::store0(it, task, lane, v0[lane], l);
                }
            }
            // Phase 1: every strip loads and computes before any stores (the
            // stores may overwrite the LDS rows the loads read).
            std::vector<T> keep(size_t(P::tasks1 > 0 ? P::tasks1 : 1) * buf1.size());
            for (int task = 0; task < P::tasks1; ++task) {
                for (int lane = 0; lane < kLanes; ++lane) P::load1(it, task, lane, l, v1[lane]);
                P::strip1_host(v1);
                std::memcpy(keep.data() + size_t(task) * buf1.size(), buf1.data(), buf1.size() * sizeof(T));
            }
            for (int task = 0; task < P::tasks1; ++task) {
                std::memcpy(buf1.data(), keep.data() + size_t(task) * buf1.size(), buf1.size() * sizeof(T));
                for (int lane = 0; lane < kLanes; ++lane) P::store1(it, task, lane, v1[lane], l);
            }
            for (int tid = 0; tid < kThreads; ++tid) P::phase2(it, l, tid);
        }
    }
}

// Every item of a launch must be visited exactly once.
static bool once_each(const std::vector<int> &seen) {
    for (int c : seen)
        if (c != 1) return false;
    return true;
}

// The library launches one work-group per item (items rounded up to 8).  A
// nonzero `slots` caps the grid instead, so each work-group steps through
// several items: the schedule must still visit every item exactly once.
static unsigned grid_for_slots(int items, unsigned slots) {
    const unsigned need = unsigned((items + kXcds - 1) / kXcds * kXcds);
    return slots == 0 || need < slots ? need : slots;
}

static unsigned g_slots = 0;
static int g_variant = 0;  // stage split (ADRT_HIP_VARIANT_DEFAULT / _ALT_SPLIT)
static size_t g_budget = kDefaultGroupBytes;  // intermediate bytes per launch group
static int g_rows = 0;  // back-projection output rows stored (0: all 2n - 1)

template <typename T, int KB, int KD>
static int emulate_group(const T *img, T *X, T *out, int n, int first, int count) {
    using P1 = Pass1<T, KB>;
    using P2 = Pass2<T, KD>;
    std::vector<int> seen;
    const int items1 = P1::num_items(n, count);
    emulate_grid<P1, T>(items1, grid_for_slots(items1, g_slots), seen, img, X, n, first);
    if (!once_each(seen)) return 5;
    const int items2 = P2::num_items(n, P1::k1, count);
    emulate_grid<P2, T>(items2, grid_for_slots(items2, g_slots), seen, (const T *)X, out, n, P1::k1,
                        first);
    if (!once_each(seen)) return 6;
    return 0;
}

template <typename T, int K2, int K4>
static int emulate_bdrt_group(const T *y, T *Z, T *out, int n, int first, int count) {
    using PA = BdrtA<T, K2>;
    using PB = BdrtB<T, K4>;
    std::vector<int> seen;
    const int rows = g_rows > 0 ? g_rows : 2 * n - 1;
    const int zrows = bdrt_z_rows(n, rows, PB::kB);
    const int itemsA = PA::num_items(n, count);
    emulate_grid<PA, T>(itemsA, grid_for_slots(itemsA, g_slots), seen, y, Z, n, first, zrows);
    if (!once_each(seen)) return 5;
    // every Z row from zrows on back to NaN, written or not: the kept output
    // rows must not depend on any of them
    const int PZ = z_pitch(n), nP = n >> PA::kA, nS = 1 << PA::kA;
    for (int img = 0; img < count; ++img)
        for (int q = 0; q < 4; ++q)
            for (int p = 0; p < nP; ++p)
                for (int sec = 0; sec < nS; ++sec) {
                    T *row = Z + ((size_t(img * 4 + q) * nP + p) * nS + sec) * PZ + kZL;
                    for (int r = zrows; r < PZ - kZL; ++r) row[r] = T(NAN);
                }
    const int itemsB = PB::num_items(PA::kA, count, rows);
    emulate_grid<PB, T>(itemsB, grid_for_slots(itemsB, g_slots), seen, (const T *)Z, out, n, PA::kA,
                        first, rows);
    if (!once_each(seen)) return 6;
    return 0;
}

// Back-projection: the input is (B, 4, 2n-1, n), as the output.
template <typename T>
static int run_bdrt(int B, int n, const char *fin, const char *fout) {
    const size_t elems = size_t(B) * 4 * (2 * n - 1) * n;
    std::vector<T> y(elems), out(elems, T(NAN));
    FILE *f = std::fopen(fin, "rb");
    if (!f || std::fread(y.data(), sizeof(T), elems, f) != elems) return 2;
    std::fclose(f);
    const int L = num_iters(size_t(n));
    const size_t per = z_bytes_per_image(n, sizeof(T));
    const size_t G = group_images(size_t(B), per, g_budget);
    std::vector<T> Z(G * per / sizeof(T), T(NAN));
    const int kA = bdrt_passA_stages(L, g_variant);
    for (int first = 0; first < B; first += int(G)) {
        const int count = (B - first) < int(G) ? (B - first) : int(G);
        const int code = (kA - 3) * 4 + (L - kA - 3);
        int st = 3;
        switch (code) {
            case 0: st = emulate_bdrt_group<T, 0, 0>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 1: st = emulate_bdrt_group<T, 0, 1>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 4: st = emulate_bdrt_group<T, 1, 0>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 5: st = emulate_bdrt_group<T, 1, 1>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 6: st = emulate_bdrt_group<T, 1, 2>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 9: st = emulate_bdrt_group<T, 2, 1>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 10: st = emulate_bdrt_group<T, 2, 2>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 11: st = emulate_bdrt_group<T, 2, 3>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 14: st = emulate_bdrt_group<T, 3, 2>(y.data(), Z.data(), out.data(), n, first, count); break;
            case 15: st = emulate_bdrt_group<T, 3, 3>(y.data(), Z.data(), out.data(), n, first, count); break;
            default: break;
        }
        if (st != 0) return st;
    }
    f = std::fopen(fout, "wb");
    if (!f || std::fwrite(out.data(), sizeof(T), elems, f) != elems) return 4;
    std::fclose(f);
    return 0;
}

template <typename T>
static int run(int B, int n, const char *fin, const char *fout) {
    const size_t in_elems = size_t(B) * n * n;
    const size_t out_elems = size_t(B) * 4 * (2 * n - 1) * n;
    std::vector<T> img(in_elems), out(out_elems, T(NAN));
    FILE *f = std::fopen(fin, "rb");
    if (!f || std::fread(img.data(), sizeof(T), in_elems, f) != in_elems) return 2;
    std::fclose(f);
    const int L = num_iters(size_t(n));
    const size_t per = x_bytes_per_image(n, L, sizeof(T));
    const size_t G = group_images(size_t(B), per, g_budget);
    std::vector<T> X(G * per / sizeof(T), T(NAN));
    const int k1 = fwd_pass1_stages(L, g_variant, int(sizeof(T)));
    for (int first = 0; first < B; first += int(G)) {
        const int count = (B - first) < int(G) ? (B - first) : int(G);
        const int code = (k1 - 3) * 4 + (L - k1 - 3);
        int st = 3;
        switch (code) {
            case 0: st = emulate_group<T, 0, 0>(img.data(), X.data(), out.data(), n, first, count); break;
            case 1: st = emulate_group<T, 0, 1>(img.data(), X.data(), out.data(), n, first, count); break;
            case 4: st = emulate_group<T, 1, 0>(img.data(), X.data(), out.data(), n, first, count); break;
            case 5: st = emulate_group<T, 1, 1>(img.data(), X.data(), out.data(), n, first, count); break;
            case 6: st = emulate_group<T, 1, 2>(img.data(), X.data(), out.data(), n, first, count); break;
            case 9: st = emulate_group<T, 2, 1>(img.data(), X.data(), out.data(), n, first, count); break;
            case 10: st = emulate_group<T, 2, 2>(img.data(), X.data(), out.data(), n, first, count); break;
            case 11: st = emulate_group<T, 2, 3>(img.data(), X.data(), out.data(), n, first, count); break;
            case 14: st = emulate_group<T, 3, 2>(img.data(), X.data(), out.data(), n, first, count); break;
            case 15: st = emulate_group<T, 3, 3>(img.data(), X.data(), out.data(), n, first, count); break;
            default: break;
        }
        if (st != 0) return st;
    }
    f = std::fopen(fout, "wb");
    if (!f || std::fwrite(out.data(), sizeof(T), out_elems, f) != out_elems) return 4;
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 6 || argc > 11) {
        std::fprintf(stderr,
                     "usage: %s f32|f64 B n in.bin out.bin [slots] [fwd|bdrt] [variant] [group_bytes]"
                     " [rows]\n",
                     argv[0]);
        return 1;
    }
    if (argc >= 7) g_slots = unsigned(std::atoi(argv[6]));
    const bool bdrt = argc >= 8 && std::strcmp(argv[7], "bdrt") == 0;
    if (argc >= 9) g_variant = std::atoi(argv[8]);
    if (argc >= 10 && std::atoll(argv[9]) > 0) g_budget = size_t(std::atoll(argv[9]));
    if (argc == 11) g_rows = std::atoi(argv[10]);
    const int B = std::atoi(argv[2]), n = std::atoi(argv[3]);
    if (std::strcmp(argv[1], "f32") == 0)
        return bdrt ? run_bdrt<float>(B, n, argv[4], argv[5]) : run<float>(B, n, argv[4], argv[5]);
    return bdrt ? run_bdrt<double>(B, n, argv[4], argv[5]) : run<double>(B, n, argv[4], argv[5]);
}
