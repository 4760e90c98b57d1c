// Host emulation of the fused forward kernels (adrt_amd/csrc/fused.hip).
//
// Compiles the exact Pass1/Pass2 phase functions for the CPU and runs every
// (block, thread) of every launch serially, with a block barrier between the
// phases, LDS modelled as an exactly sized heap buffer and X sized exactly as
// the library's workspace.  Built with AddressSanitizer on the host, so any
// out-of-bounds access of the kernels' index math aborts.  Usage:
//   fused_emu f32|f64 B n in.bin out.bin
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define ADRT_HOST_EMULATION 1
#include "../../adrt_amd/csrc/fused.hip"

using namespace adrt_hip;
using namespace adrt_hip::fused;

template <typename P, typename... Args>
static void emulate_grid(unsigned grid, Args... args) {
    using T = std::remove_cv_t<std::remove_pointer_t<decltype(P::Item::X)>>;
    std::vector<T> lds_buf(P::lds_elems > 0 ? P::lds_elems : 1);
    for (unsigned blk = 0; blk < grid; ++blk) {
        const auto it = P::item(args..., int(blk));
        if (!it.valid) continue;
        // Fill LDS with NaN so reads of never-written slots would poison results.
        for (auto &v : lds_buf) v = T(NAN);
        T *lds = P::lds_elems > 0 ? lds_buf.data() : nullptr;
        for (int tid = 0; tid < kThreads; ++tid) P::phase0(it, lds, tid);
        for (int tid = 0; tid < kThreads; ++tid) P::phase1(it, lds, tid);
    }
}

template <typename T, int KA, int KB, int KC, int KD>
static void emulate_group(const T *img, T *X, T *out, int n, int first, int count) {
    using P1 = P1Of<T, KA, KB>;
    using P2 = P2Of<T, KC, KD>;
    const unsigned g1 = unsigned(count) * 4u * unsigned(n >> P1::k1) * unsigned(P1::num_chunks(n));
    emulate_grid<P1>(g1, img, X, n, first, count);
    const unsigned g2 = unsigned(count) * 4u * unsigned(1 << P1::k1) * unsigned(P2::num_chunks(n));
    emulate_grid<P2>(g2, (const T *)X, out, n, P1::k1, first, count);
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
    const size_t G = group_images(size_t(B), per);
    std::vector<T> X(G * per / sizeof(T), T(NAN));
    for (int first = 0; first < B; first += int(G)) {
        const int count = (B - first) < int(G) ? (B - first) : int(G);
        switch (L) {
            case 6: emulate_group<T, 3, 0, 3, 0>(img.data(), X.data(), out.data(), n, first, count); break;
            case 7: emulate_group<T, 3, 1, 3, 0>(img.data(), X.data(), out.data(), n, first, count); break;
            case 8: emulate_group<T, 3, 1, 3, 1>(img.data(), X.data(), out.data(), n, first, count); break;
            case 9: emulate_group<T, 3, 2, 3, 1>(img.data(), X.data(), out.data(), n, first, count); break;
            case 10: emulate_group<T, 3, 2, 3, 2>(img.data(), X.data(), out.data(), n, first, count); break;
            case 11: emulate_group<T, 3, 3, 3, 2>(img.data(), X.data(), out.data(), n, first, count); break;
            case 12: emulate_group<T, 3, 3, 3, 3>(img.data(), X.data(), out.data(), n, first, count); break;
            default: return 3;
        }
    }
    f = std::fopen(fout, "wb");
    if (!f || std::fwrite(out.data(), sizeof(T), out_elems, f) != out_elems) return 4;
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
