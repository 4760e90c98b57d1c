// Host emulation of the one-launch LDS kernel (adrt_amd/csrc/small_chunk.hpp).
//
// Compiles the exact phase functions, the class/chunk decoding and the slot
// permutation for the CPU and runs the persistent schedule work-group by
// work-group: every phase runs for all kThreads threads before the next one
// (the kernel's barrier), on an LDS buffer of exactly 2 * kMaxElems elements
// filled with NaN before every chunk.  Global arrays are sized exactly, so
// an out-of-bounds access aborts under AddressSanitizer; a chunk run twice or
// never aborts too.  Usage:
//   small_emu f32|f64 fused|fwd|bdrt grid  n1 B1 in1 adrt1 bdrt1  [n2 B2 ...]
// fused: images -> adrt and bdrt(adrt); fwd: images -> adrt; bdrt: ADRT data
// `in` -> bdrt.  Output paths of outputs the mode does not write are ignored.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#define ADRT_HOST_EMULATION 1
#include "../../adrt_amd/csrc/small_chunk.hpp"

using namespace adrt_hip;
using namespace adrt_hip::small;

template <typename T>
static std::vector<T> read_file(const char *path, size_t count) {
    std::vector<T> v(count);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", count, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

template <typename T>
static void write_file(const char *path, const std::vector<T> &v) {
    FILE *f = std::fopen(path, "wb");
    if (!f || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
}

struct Entry {
    size_t n, B;
    const char *in, *adrt, *bdrt;
};

template <typename T>
static int run(const std::string &mode, int grid, const std::vector<Entry> &entries) {
    std::vector<std::vector<T>> ins, fwds, backs;
    Job job{};
    for (const Entry &e : entries) {
        const size_t out = e.B * 4 * (2 * e.n - 1) * e.n;
        const size_t in = mode == "bdrt" ? out : e.B * e.n * e.n;
        ins.push_back(read_file<T>(e.in, in));
        fwds.emplace_back(mode == "bdrt" ? 0 : out, T(NAN));
        backs.emplace_back(mode == "fwd" ? 0 : out, T(NAN));
    }
    for (size_t i = 0; i < entries.size(); ++i) {
        const Entry &e = entries[i];
        void *fwd = fwds[i].empty() ? nullptr : fwds[i].data();
        void *back = backs[i].empty() ? nullptr : backs[i].data();
        if (!add_class(job, ins[i].data(), fwd, back, e.B, e.n, mode == "bdrt")) {
            std::fprintf(stderr, "add_class refused entry %zu\n", i);
            return 3;
        }
    }
    std::vector<T> lds(2 * size_t(kMaxElems));
    std::vector<int> seen(size_t(job.total), 0);
    const int slots = (job.total + 31) & ~31;
    for (int blk = 0; blk < grid; ++blk) {
        for (int s = blk; s < slots; s += grid) {
            const int c = slot_chunk(s);
            if (c >= job.total) continue;
            seen[size_t(c)]++;
            const Class cl = class_of(job, c);
            const Chunk k = chunk_of(cl, c);
            if (k.np < 1 || k.np > cl.per_chunk || k.np * k.n * k.C > kMaxElems ||
                (k.np * k.n << k.seg) > kMaxUnits || k.np * k.C * k.n > 64 * kMaxUnits) {
                std::fprintf(stderr, "chunk %d: %d planes do not fit\n", c, k.np);
                return 4;
            }
            for (auto &x : lds) x = T(NAN);
            run_chunk<T>(cl, k, lds.data(), lds.data() + kMaxElems, [&](auto f) {
                for (int tid = 0; tid < kThreads; ++tid) f(tid >> 6, tid & 63);
            });
        }
    }
    for (int c = 0; c < job.total; ++c)
        if (seen[size_t(c)] != 1) {
            std::fprintf(stderr, "chunk %d ran %d times\n", c, seen[size_t(c)]);
            return 5;
        }
    for (size_t i = 0; i < entries.size(); ++i) {
        if (!fwds[i].empty()) write_file(entries[i].adrt, fwds[i]);
        if (!backs[i].empty()) write_file(entries[i].bdrt, backs[i]);
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 9 || (argc - 4) % 5 != 0) {
        std::fprintf(stderr, "usage: small_emu f32|f64 fused|fwd|bdrt grid n B in adrt bdrt ...\n");
        return 2;
    }
    const std::string dt = argv[1], mode = argv[2];
    const int grid = std::atoi(argv[3]);
    std::vector<Entry> entries;
    for (int a = 4; a + 4 < argc; a += 5)
        entries.push_back(Entry{size_t(std::atoll(argv[a])), size_t(std::atoll(argv[a + 1])),
                                argv[a + 2], argv[a + 3], argv[a + 4]});
    if (grid < 1 || (mode != "fused" && mode != "fwd" && mode != "bdrt")) return 2;
    return dt == "f32" ? run<float>(mode, grid, entries) : run<double>(mode, grid, entries);
}
