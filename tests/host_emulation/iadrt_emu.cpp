// Host emulation of the iadrt stage kernel (adrt_amd/csrc/iadrt.hip).
//
// Runs every wavefront of every stage on the CPU with the kernel's own
// per-lane step function: lanes advance in lockstep, step by step, through a
// 64-row ring exactly sized as the kernel's LDS ring (poisoned with NaN), and
// row s is flushed at step s with the kernel's unconditional stores.  Pair loads use the kernel's clamped rows.
// Buffers are exactly sized, so an out-of-bounds access aborts under
// AddressSanitizer.  Usage:  iadrt_emu f32|f64 B n in.bin out.bin
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
    std::vector<T> ring(size_t(kRing) * (kLanes + 1));
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
                        ring[size_t(row % kRing) * (kLanes + 1) + l] = val;
                    prev1[l] = p1;
                }
                // the kernel's unconditional flush: steps above the last row
                // write row R-1 (overwritten at step R-1), lanes past n repeat
                // the last live lane
                for (int l = 0; l < kLanes; ++l) {
                    const int fl = w.ch0 + l < n ? l : n - 1 - w.ch0;
                    const int fr = s < R ? s : R - 1;
                    dst[size_t(fr) * n + w.ch0 + fl] = ring[size_t(s % kRing) * (kLanes + 1) + fl];
                }
            }
        }
    }
}

template <typename T>
static int run(int B, int n, const char *fin, const char *fout) {
    const int R = 2 * n - 1;
    const size_t elems = size_t(B) * 4 * R * n;
    std::vector<T> a(elems), b(elems, T(NAN));
    FILE *f = std::fopen(fin, "rb");
    if (!f || std::fread(a.data(), sizeof(T), elems, f) != elems) return 2;
    std::fclose(f);
    const int L = num_iters(size_t(n));
    int Li = 1;
    for (int k = 0; k < L; ++k) {
        emulate_stage<T>(a, b, B * 4, R, n, Li);
        std::swap(a, b);
        for (auto &x : b) x = T(NAN);
        Li *= 2;
    }
    f = std::fopen(fout, "wb");
    if (!f || std::fwrite(a.data(), sizeof(T), elems, f) != elems) return 4;
    std::fclose(f);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s f32|f64 B n in.bin out.bin\n", argv[0]);
        return 1;
    }
    const int B = std::atoi(argv[2]), n = std::atoi(argv[3]);
    if (std::strcmp(argv[1], "f32") == 0) return run<float>(B, n, argv[4], argv[5]);
    return run<double>(B, n, argv[4], argv[5]);
}
