// Kernels of the learned initial guess / preconditioner for CG inversion:
// a 4-level U-Net (3x3 convolutions, BF16 weights and activations, FP32
// accumulation on the matrix cores) with its training kernels, a
// multiscale linear preconditioner (fp32), and a phantom rasteriser that
// generates training data on the device.  None of this has a counterpart in
// the reference package; it builds on its transforms.
//
// Activations are NHWC bf16 with a pixel stride (elements) that may exceed
// the channel count, so a U-Net skip connection and the upsampled coarse
// features can share one buffer (zero-copy concatenation).  Channel counts
// and strides are multiples of 8, so 8 channels are one 16-byte load.
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
//   forward / input gradient: D[pixel][co] = sum_k X[pixel + tap(k)][ci(k)] * W[co][k],
//     k = tap * Cin + ci (weights stored [Cout][9][Cin]);
//   weight gradient: D[co][ci] = sum_pixel dZ[pixel][co] * X[pixel + tap][ci],
//     per tap, reduced over pixel partitions in a second pass.
// Operand fragments of one MFMA: A lane l holds row l%16 with k-group l/16
// (8 consecutive k), B lane l holds column l%16 with the same k-group, and
// the result lane l holds column l%16, rows 4*(l/16) + 0..3.
#include "common.hpp"

namespace adrt_hip {
namespace nn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaveLanes = 64;
constexpr int kWgThreads = 256;
constexpr int kMT = 4;  // 16-pixel row tiles per wave (forward / input gradient)

__device__ __forceinline__ bf16x8 load8(const bf16 *p) {
    return *reinterpret_cast<const bf16x8 *>(p);
}

// v if keep, else all zeros (branch-free: the load above is always issued).
__device__ __forceinline__ bf16x8 keep_or_zero(bf16x8 v, bool keep) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    const unsigned m = keep ? 0xffffffffu : 0u;
    u.x &= m;
    u.y &= m;
    u.z &= m;
    u.w &= m;
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// 3x3 convolution, zero padding, stride 1 (forward and input gradient).
// ---------------------------------------------------------------------------
enum ConvMode : int {
    kReluBias = 0,   // y = relu(conv + bias)                       (bf16)
    kResidual = 1,   // y32[p] = resid[p] + conv[p][0] + bias[0]     (fp32, channel 0 only)
    kMaskInput = 2,  // y = conv * (mask > 0)                       (bf16)
    kLinear = 3,     // y = conv                                    (bf16)
};

struct ConvArgs {
    const bf16 *x;
    const bf16 *w;  // [Cout][9][Cin]
    const float *bias;
    bf16 *y;
    float *y32;
    const float *resid;
    const bf16 *mask;
    long long P;  // pixels = B * H * W
    int x_stride, y_stride, mask_stride;
    int H, W, Cin, Cout;
    // input channels [c_split, Cin) come from x2 at half resolution (nearest
    // upsampling: pixel (i, j) reads x2 pixel (i/2, j/2), channel c - c_split);
    // c_split == Cin: x alone (LDS-tiled kernel only)
    const bf16 *x2;
    int x2_stride, c_split;
};

// Wave = 64 consecutive pixels (4 tiles of 16) x 16*NT output channels.  A
// work-group's 4 waves share the output-channel block (their weight
// fragments hit the same cache lines).
template <int NT, int MODE>
__global__ void __launch_bounds__(kWgThreads) conv3x3_kernel(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int n_co_blocks = a.Cout / (16 * NT);
    const int co0 = int(blockIdx.x % n_co_blocks) * 16 * NT;
    const long long px0 = ((long long)(blockIdx.x / n_co_blocks) * 4 + wave) * 64;
    if (px0 >= a.P) return;  // no barriers below

    // A rows: pixel px0 + 16*mt + r
    int py[kMT], pxx[kMT];
    long long pp[kMT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) {
        const long long p = px0 + 16 * mt + r;
        pp[mt] = p;
        pxx[mt] = int(p % a.W);
        py[mt] = int((p / a.W) % a.H);
    }
    f32x4 acc[kMT][NT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int K = 9 * a.Cin;
    int k = 8 * g;  // this lane group's first k
    int t = k / a.Cin, ci = k - t * a.Cin;
    const bf16 *wrow[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) wrow[nt] = a.w + (size_t)(co0 + 16 * nt + r) * K;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const bool kok = k < K;
        const int tc = kok ? t : 4;
        const int dy = tc / 3 - 1, dx = tc - (tc / 3) * 3 - 1;
        const long long delta = ((long long)dy * a.W + dx) * a.x_stride + (kok ? ci : 0);
        bf16x8 A[kMT], Bf[NT];
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt) {
            const bool ok = kok && unsigned(py[mt] + dy) < unsigned(a.H) &&
                            unsigned(pxx[mt] + dx) < unsigned(a.W);
            const long long off = pp[mt] * a.x_stride + (ok ? delta : 0);
            A[mt] = keep_or_zero(load8(a.x + off), ok);
        }
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) Bf[nt] = keep_or_zero(load8(wrow[nt] + (kok ? k : 0)), kok);
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(A[mt], Bf[nt], acc[mt][nt]);
        k += 32;
        ci += 32;
        while (ci >= a.Cin) {
            ci -= a.Cin;
            ++t;
        }
    }

#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = co0 + 16 * nt + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const long long p = px0 + 16 * mt + 4 * g + j;
                float v = acc[mt][nt][j];
                if (MODE == kReluBias) {
                    v += a.bias[co];
                    v = v > 0.f ? v : 0.f;
                    a.y[p * a.y_stride + co] = bf16(v);
                } else if (MODE == kResidual) {
                    if (co == 0) a.y32[p] = a.resid[p] + (v + a.bias[0]);
                } else if (MODE == kMaskInput) {
                    const float m = float(a.mask[p * a.mask_stride + co]);
                    a.y[p * a.y_stride + co] = bf16(m > 0.f ? v : 0.f);
                } else {
                    a.y[p * a.y_stride + co] = bf16(v);
                }
            }
        }
    }
}

// LDS-tiled form for W a multiple of 16 (>= 16), H a multiple of 4 and
// Cin >= 16: a work-group computes 4 output rows x TW = min(W, 64) columns x
// 16*NT output channels (one row per wave).  Input channels go through LDS
// 32 at a time: the (4+2) x (TW+2) pixel tile with its zero-padded halo is
// loaded once per chunk and read by all 9 taps from LDS.  Inside a chunk
// k = tap * ccv + c runs through the chunk's taps and channels as in the
// kernel above.  Pixel rows in LDS are padded to 40 channels (80 bytes), so
// 16 lanes reading 16 consecutive pixels hit distinct banks.
constexpr int kTH = 4, kCC = 32, kCCP = kCC + 8, kTWmax = 64;
// 16-byte tile vectors per thread at most: (kTH + 2) (kTWmax + 2) kCC / 8
// over kWgThreads threads
constexpr int kFill = ((kTH + 2) * (kTWmax + 2) * (kCC / 8) + kWgThreads - 1) / kWgThreads;

template <int NT, int MODE>
__global__ void __launch_bounds__(kWgThreads) conv3x3_lds_kernel(ConvArgs a, int TW,
                                                                 int tw2_recip) {
    __shared__ __align__(16) bf16 tile[kTH + 2][kTWmax + 2][kCCP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int co0 = int(blockIdx.y) * 16 * NT;
    const int tiles_x = a.W / TW, tiles_y = a.H / kTH;
    int tt = int(blockIdx.x);
    const int tx = tt % tiles_x;
    tt /= tiles_x;
    const int ty = tt % tiles_y;
    const long long b = tt / tiles_y;
    const int x0 = tx * TW, y0 = ty * kTH;
    const int MT = TW / 16;
    const long long img = b * a.H * a.W;  // first pixel of the image
    const long long img2 = b * (a.H >> 1) * (a.W >> 1);  // ... in x2

    f32x4 acc[kMT][NT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int K = 9 * a.Cin;
    for (int c0 = 0; c0 < a.Cin; c0 += kCC) {
        const int ccv = a.Cin - c0 < kCC ? a.Cin - c0 : kCC;
        const int nvec = ccv / 8;
        const int npix = (kTH + 2) * (TW + 2);
        __syncthreads();  // the previous chunk's readers are done
        // nvec is 1, 2 or 4 (shifts); pix / (TW + 2) by the exact reciprocal
        // multiply (TW in {16, 32, 48, 64}, pix < 6 (TW + 2)).  A thread's
        // (at most kFill) vectors are all loaded before any is stored, so
        // their loads are in flight together; vectors past the tile load
        // the image's first pixel and are not stored.
        const int lg = nvec == 4 ? 2 : (nvec == 2 ? 1 : 0);
        const int nfill = npix * nvec;
        bf16x8 fv[kFill];
        int fdst[kFill];
#pragma unroll
        for (int j = 0; j < kFill; ++j) {
            const int i0 = int(threadIdx.x) + j * kWgThreads;
            const int i = i0 < nfill ? i0 : 0;
            const int v = i & (nvec - 1), pix = i >> lg;
            const int py = (pix * tw2_recip) >> 16, px = pix - py * (TW + 2);
            const int gy = y0 + py - 1, gx = x0 + px - 1;
            const bool ok = unsigned(gy) < unsigned(a.H) && unsigned(gx) < unsigned(a.W);
            const int ch = c0 + 8 * v;
            const bf16 *src;
            if (ch < a.c_split) {
                const long long q = ok ? img + (long long)gy * a.W + gx : img;
                src = a.x + q * a.x_stride + ch;
            } else {
                const long long q = ok ? img2 + (long long)(gy >> 1) * (a.W >> 1) + (gx >> 1) : img2;
                src = a.x2 + q * a.x2_stride + (ch - a.c_split);
            }
            fv[j] = keep_or_zero(load8(src), ok);
            fdst[j] = (py * (kTWmax + 2) + px) * kCCP + 8 * v;
        }
#pragma unroll
        for (int j = 0; j < kFill; ++j)
            if (int(threadIdx.x) + j * kWgThreads < nfill)
                *reinterpret_cast<bf16x8 *>(&tile[0][0][0] + fdst[j]) = fv[j];
        __syncthreads();
        const int Kc = 9 * ccv;
        int k = 8 * g;
        int tap = k / ccv, ci = k - tap * ccv;
        // weight fragments of k (zero past Kc, read from tap 4, channel 0)
        auto load_b = [&](int kk, int tp, int cn, bf16x8 (&Bo)[NT]) {
            const bool ok = kk < Kc;
            const int tc = ok ? tp : 4, cc = ok ? cn : 0;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                Bo[nt] = keep_or_zero(
                    load8(a.w + (size_t)(co0 + 16 * nt + r) * K + tc * a.Cin + c0 + cc), ok);
        };
        bf16x8 Bf[NT];
        load_b(k, tap, ci, Bf);
        for (int k0 = 0; k0 < Kc; k0 += 32) {
            const bool kok = k < Kc;
            const int tc = kok ? tap : 4;
            const int dy = tc / 3, dx = tc - (tc / 3) * 3;
            const int cc = kok ? ci : 0;
            bf16x8 A[kMT];
#pragma unroll
            for (int mt = 0; mt < kMT; ++mt)
                if (mt < MT)
                    A[mt] = keep_or_zero(
                        *reinterpret_cast<const bf16x8 *>(&tile[wave + dy][16 * mt + r + dx][cc]),
                        kok);
            // the next step's weights are loaded before this step's products
            const int kn = k + 32;
            int tapn = tap, cin = ci + 32;
            while (cin >= ccv) {
                cin -= ccv;
                ++tapn;
            }
            bf16x8 Bn[NT];
            load_b(kn, tapn, cin, Bn);
#pragma unroll
            for (int mt = 0; mt < kMT; ++mt)
                if (mt < MT)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(A[mt], Bf[nt], acc[mt][nt]);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) Bf[nt] = Bn[nt];
            k = kn;
            tap = tapn;
            ci = cin;
        }
    }

    const long long prow = img + (long long)(y0 + wave) * a.W + x0;
    if (MODE == kResidual) {
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt) {
            if (mt >= MT) break;
            if (r == 0 && co0 == 0) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const long long p = prow + 16 * mt + 4 * g + j;
                    a.y32[p] = a.resid[p] + (acc[mt][0][j] + a.bias[0]);
                }
            }
        }
        return;
    }
    // bf16 outputs go through LDS, one 16-channel slice at a time: each wave
    // writes its row's [16*MT pixels][16 channels] block, then stores it as
    // 16-byte vectors (8 channels of one pixel per lane).
    bf16 *ob = &tile[0][0][0] + wave * (kTWmax * 16);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int co = co0 + 16 * nt + r;
        const float bco = MODE == kReluBias ? a.bias[co] : 0.f;
        __syncthreads();  // the tile's (or the previous slice's) readers are done
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt) {
            if (mt >= MT) break;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int px = 16 * mt + 4 * g + j;
                float v = acc[mt][nt][j];
                if (MODE == kReluBias) {
                    v += bco;
                    v = v > 0.f ? v : 0.f;
                } else if (MODE == kMaskInput) {
                    const float m = float(a.mask[(prow + px) * a.mask_stride + co]);
                    v = m > 0.f ? v : 0.f;
                }
                ob[px * 16 + r] = bf16(v);
            }
        }
        __syncthreads();
        for (int e = lane; e < 32 * MT; e += kWaveLanes) {
            const int px = e >> 1, h = e & 1;
            const bf16x8 v = *reinterpret_cast<const bf16x8 *>(&ob[px * 16 + 8 * h]);
            *reinterpret_cast<bf16x8 *>(&a.y[(prow + px) * a.y_stride + co0 + 16 * nt + 8 * h]) = v;
        }
    }
}

// ---------------------------------------------------------------------------
// The same convolution for the wide layers: W a multiple of 64, H a multiple
// of 4, Cin a multiple of 16, 16 output channels per work-group (Cout 16 or
// 32).  Tiles, channel chunks (32 at a time, then a last one of 16),
// the K order k = tap * ccv + c, the MFMA operand layout and the epilogue
// are those of conv3x3_lds_kernel, so the results are bit-identical to it.
// What changes is where the index work happens.  A chunk of ccv channels is
// 9 (ccv = 32) or 4.5 (ccv = 16) MFMA K steps whose taps and channels are
// known at compile time, so every A fragment is an LDS read at a constant
// offset from one per-lane base (for ccv = 16 a step covers two taps, one
// per half of the lane groups: a select between two constants).  The
// chunk's weight fragments are loaded once, together with the tile fill,
// and the fill's pixel decomposition uses the constant tile width.  The
// narrow layers issued about 50 VALU instructions per MFMA in the general
// kernel and were bound by them.  It serves the 16-channel layers: with 32
// output channels per work-group the chunk's weights need 72 more registers
// and the kernel drops to 2 waves per SIMD, and two 16-channel work-groups
// per tile (the mapping below) measured no faster than the general kernel
// for the 32-channel layers (MI355X, 1000 images at 128^2).
// ---------------------------------------------------------------------------
constexpr int kWideRow = kTWmax + 2;  // LDS tile row (pixels)

__device__ __forceinline__ constexpr int tap_offset(int t) {
    return ((t / 3) * kWideRow + (t % 3)) * kCCP;  // tile element offset of tap t
}

// Weight fragments of a chunk: step s, lane group g -> k = 32 s + 8 g, i.e.
// tap s / channels 8g (ccv = 32) or tap 2s + g/2 / channels 8(g%2) (ccv = 16;
// tap 9 of the last step is past the chunk: zero).  wl: the lane's weight
// row (output channel co0 + r), Cin: the layer's input channels.
template <int NT, int CCV>
__device__ __forceinline__ void wide_load_b(const bf16 *wl, int Cin, int c0, int g,
                                            bf16x8 (&B)[CCV == 32 ? 9 : 5][NT]) {
    constexpr int S = CCV == 32 ? 9 : 5;
    const int K = 9 * Cin;
    const bf16 *wc = wl + c0 + (CCV == 32 ? 8 * g : (g >> 1) * Cin + 8 * (g & 1));
    const int wstep = CCV == 32 ? Cin : 2 * Cin;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const bool ok = CCV == 32 || s < S - 1 || g < 2;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
            B[s][nt] = keep_or_zero(load8(wc + (ok ? s * wstep : 0) + (size_t)nt * 16 * K), ok);
    }
}

// The chunk's MFMAs on the filled tile: lane (r, g) of tile mt reads pixel
// 16 mt + r of row `wave` shifted by its tap, channels of its group, at a
// constant offset from the lane's base.
template <int NT, int CCV>
__device__ __forceinline__ void wide_mfma(const bf16 (&tile)[kTH + 2][kWideRow][kCCP], int wave,
                                          int g, int r, const bf16x8 (&B)[CCV == 32 ? 9 : 5][NT],
                                          f32x4 (&acc)[kMT][NT]) {
    constexpr int S = CCV == 32 ? 9 : 5;
    const bf16 *tl = &tile[0][0][0] + (wave * kWideRow + r) * kCCP + (CCV == 32 ? 8 * g : 8 * (g & 1));
    const bool hi = g >= 2;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int off;
        if (CCV == 32) {
            off = tap_offset(s);
        } else {
            const int t0 = 2 * s, t1 = 2 * s + 1 < 9 ? 2 * s + 1 : 8;
            off = hi ? tap_offset(t1) : tap_offset(t0);
        }
        bf16x8 A[kMT];
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt) {
            A[mt] = *reinterpret_cast<const bf16x8 *>(tl + off + 16 * mt * kCCP);
            if (CCV == 16 && s == S - 1) A[mt] = keep_or_zero(A[mt], !hi);
        }
#pragma unroll
        for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = mfma(A[mt], B[s][nt], acc[mt][nt]);
    }
}

template <int NT, int CCV>
__device__ __forceinline__ void wide_chunk(const ConvArgs &a, bf16 (&tile)[kTH + 2][kWideRow][kCCP],
                                           const bf16 *xb, const bf16 *x2b, const bf16 *wl, int c0,
                                           int y0, int x0, int wave, int g, int r,
                                           f32x4 (&acc)[kMT][NT]) {
    constexpr int NV = CCV / 8;                  // 16-byte vectors per pixel
    constexpr int LG = CCV == 32 ? 2 : 1;
    constexpr int NFILL = (kTH + 2) * kWideRow * NV;
    constexpr int FILL = (NFILL + kWgThreads - 1) / kWgThreads;
    const int tid = int(threadIdx.x);

    bf16x8 B[CCV == 32 ? 9 : 5][NT];
    wide_load_b<NT, CCV>(wl, a.Cin, c0, g, B);

    __syncthreads();  // the previous chunk's readers are done
    // tile fill: (kTH + 2) x kWideRow pixels with a zero halo; a thread's
    // vectors all hold channels c0 + 8v (v = tid % NV) of their pixels
    const int v = tid & (NV - 1);
    const int ch = c0 + 8 * v;
    const bool up = ch >= a.c_split;
    bf16x8 fv[FILL];
#pragma unroll
    for (int j = 0; j < FILL; ++j) {
        const int i = tid + j * kWgThreads;
        const int pix = (i < NFILL ? i : 0) >> LG;
        const int py = pix / kWideRow, px = pix - py * kWideRow;
        const int gy = y0 + py - 1, gx = x0 + px - 1;
        const bool ok = unsigned(gy) < unsigned(a.H) && unsigned(gx) < unsigned(a.W);
        const bf16 *src;
        if (up)
            src = x2b + (ok ? (gy >> 1) * (a.W >> 1) + (gx >> 1) : 0) * a.x2_stride + (ch - a.c_split);
        else
            src = xb + (ok ? gy * a.W + gx : 0) * a.x_stride + ch;
        fv[j] = keep_or_zero(load8(src), ok);
    }
#pragma unroll
    for (int j = 0; j < FILL; ++j) {
        const int i = tid + j * kWgThreads;
        if (i < NFILL)
            *reinterpret_cast<bf16x8 *>(&tile[0][0][0] + (i >> LG) * kCCP + 8 * v) = fv[j];
    }
    __syncthreads();
    wide_mfma<NT, CCV>(tile, wave, g, r, B, acc);
}

// bf16 outputs (16 channels co0 .. co0 + 15 of the work-group's 4 x 64
// pixels, one row per wave) through LDS, as in conv3x3_lds_kernel; then, if
// `pooled`, the 2 x 32 averages of their 2 x 2 blocks as pool_kernel forms
// them.  The tile's readers must be done (a barrier comes first).
template <int MODE>
__device__ __forceinline__ void wide_store(bf16 (&tile)[kTH + 2][kWideRow][kCCP],
                                           const f32x4 (&acc)[kMT][1], const float *bias, int co0,
                                           bf16 *y, int y_stride, long long prow, int wave, int lane,
                                           int g, int r, bf16 *pooled, int pooled_stride,
                                           long long pq0, int Wc) {
    bf16 *ob0 = &tile[0][0][0];
    bf16 *ob = ob0 + wave * (kTWmax * 16);
    const int co = co0 + r;
    const float bco = MODE == kReluBias ? bias[co] : 0.f;
    __syncthreads();  // the tile's readers are done
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int px = 16 * mt + 4 * g + j;
            float val = acc[mt][0][j];
            if (MODE == kReluBias) {
                val += bco;
                val = val > 0.f ? val : 0.f;
            }
            ob[px * 16 + r] = bf16(val);
        }
    }
    __syncthreads();
    for (int e = lane; e < 32 * kMT; e += kWaveLanes) {
        const int px = e >> 1, h = e & 1;
        const bf16x8 val = *reinterpret_cast<const bf16x8 *>(&ob[px * 16 + 8 * h]);
        *reinterpret_cast<bf16x8 *>(&y[(prow + px) * y_stride + co0 + 8 * h]) = val;
    }
    if (pooled) {
        // pooled row pr (of 2), column pc (of 32), channels 8h .. 8h + 7:
        // 0.25 * ((x[2pr][2pc] + x[2pr][2pc+1]) + (x[2pr+1][2pc] + x[2pr+1][2pc+1]))
        const int e = int(threadIdx.x);
        if (e < 2 * (kTWmax / 2) * 2) {
            const int h = e & 1, pc = (e >> 1) & (kTWmax / 2 - 1), pr = e >> 6;
            const bf16 *r0 = ob0 + (2 * pr) * (kTWmax * 16), *r1 = r0 + kTWmax * 16;
            const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(r0 + (2 * pc) * 16 + 8 * h);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(r0 + (2 * pc + 1) * 16 + 8 * h);
            const bf16x8 a2 = *reinterpret_cast<const bf16x8 *>(r1 + (2 * pc) * 16 + 8 * h);
            const bf16x8 a3 = *reinterpret_cast<const bf16x8 *>(r1 + (2 * pc + 1) * 16 + 8 * h);
            bf16x8 o;
#pragma unroll
            for (int j = 0; j < 8; ++j)
                o[j] = bf16(0.25f * ((float(a0[j]) + float(a1[j])) + (float(a2[j]) + float(a3[j]))));
            *reinterpret_cast<bf16x8 *>(&pooled[(pq0 + (long long)pr * Wc + pc) * pooled_stride +
                                                co0 + 8 * h]) = o;
        }
    }
}

// Work-group -> (tile, 16-channel group): the cg groups of one tile go to
// work-groups 8 apart, which run on one XCD (work-groups are dealt to the 8
// XCDs round-robin) at about the same time, so the tile is read from HBM
// once and from that XCD's L2 for the other groups.
template <int MODE>
__global__ void __launch_bounds__(kWgThreads) conv3x3_wide_kernel(ConvArgs a, int tiles) {
    constexpr int NT = 1;
    static_assert(MODE == kReluBias || MODE == kLinear, "bf16 outputs");
    __shared__ __align__(16) bf16 tile[kTH + 2][kWideRow][kCCP];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
    const int g = lane >> 4, r = lane & 15;
    const int cg = a.Cout / 16;
    const int bid = int(blockIdx.x);
    const int full = tiles / 8 * 8;
    int tt, grp;
    if (bid < full * cg) {
        const int rem = bid % (8 * cg);
        tt = bid / (8 * cg) * 8 + rem % 8;
        grp = rem / 8;
    } else {
        tt = full + (bid - full * cg) / cg;
        grp = (bid - full * cg) % cg;
    }
    const int co0 = grp * 16;
    const int tiles_x = a.W / kTWmax, tiles_y = a.H / kTH;
    const int tx = tt % tiles_x;
    tt /= tiles_x;
    const int ty = tt % tiles_y;
    const long long b = tt / tiles_y;
    const int x0 = tx * kTWmax, y0 = ty * kTH;
    const long long img = b * a.H * a.W;  // first pixel of the image
    const bf16 *xb = a.x + img * a.x_stride;
    const bf16 *x2b = a.x2 ? a.x2 + b * (a.H >> 1) * (a.W >> 1) * a.x2_stride : a.x;
    const bf16 *wl = a.w + (size_t)(co0 + r) * (9 * a.Cin);

    f32x4 acc[kMT][NT];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};

    int c0 = 0;
    for (; c0 + 32 <= a.Cin; c0 += 32)
        wide_chunk<NT, 32>(a, tile, xb, x2b, wl, c0, y0, x0, wave, g, r, acc);
    if (c0 < a.Cin) wide_chunk<NT, 16>(a, tile, xb, x2b, wl, c0, y0, x0, wave, g, r, acc);

    const long long prow = img + (long long)(y0 + wave) * a.W + x0;
    wide_store<MODE>(tile, acc, a.bias, co0, a.y, a.y_stride, prow, wave, lane, g, r, nullptr, 0,
                     0, 0);
}

// ---------------------------------------------------------------------------
// The encoder's full-resolution level in one kernel (inference): e0a (the
// first layer, 1 -> 16 channels from the float32 input), e0b (16 -> 16) and
// the 2 x 2 average pooling.  A work-group's e0b tile is 4 x 64 pixels; the
// 6 x 66 e0a values it reads are computed into the LDS tile from an 8 x 68
// tile of bf16(scale * z), with conv_in_kernel's arithmetic (one fp32 fma
// chain per output in tap order, + bias, ReLU, bf16), so e0a never goes
// through HBM.  e0b is conv3x3_wide_kernel's (same MFMA operands), written
// as the skip connection, and its 2 x 2 averages (pool_kernel's arithmetic)
// go to the pooled output.  Results equal conv_in + conv3x3 + pool bit for
// bit.
// ---------------------------------------------------------------------------
struct Enc0Args {
    const float *z;
    float scale;
    const bf16 *w1;  // e0a weights [16][9][cinw1] (channel 0 real)
    int cinw1;
    const float *b1;
    const bf16 *w2;  // e0b weights [16][9][16]
    const float *b2;
    bf16 *skip;
    int skip_stride;
    bf16 *pooled;
    int pooled_stride;
    int H, W;
};

__global__ void __launch_bounds__(kWgThreads) encoder0_kernel(Enc0Args e) {
    __shared__ __align__(16) bf16 tile[kTH + 2][kWideRow][kCCP];
    __shared__ float zt[kTH + 4][kTWmax + 4];
    __shared__ __align__(16) float w1l[9][16];
    const int tid = int(threadIdx.x);
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int g = lane >> 4, r = lane & 15;
    const int tiles_x = e.W / kTWmax, tiles_y = e.H / kTH;
    int tt = int(blockIdx.x);
    const int tx = tt % tiles_x;
    tt /= tiles_x;
    const int ty = tt % tiles_y;
    const long long b = tt / tiles_y;
    const int x0 = tx * kTWmax, y0 = ty * kTH;
    const long long img = b * e.H * e.W;

    bf16x8 B[5][1];
    wide_load_b<1, 16>(e.w2 + (size_t)r * (9 * 16), 16, 0, g, B);
    // bf16(scale * z) around the tile (zero outside the image), e0a weights
    for (int i = tid; i < (kTH + 4) * (kTWmax + 4); i += kWgThreads) {
        const int rr = i / (kTWmax + 4), cc = i - rr * (kTWmax + 4);
        const int gy = y0 + rr - 2, gx = x0 + cc - 2;
        const bool ok = unsigned(gy) < unsigned(e.H) && unsigned(gx) < unsigned(e.W);
        zt[rr][cc] = ok ? float(bf16(e.scale * e.z[img + (long long)gy * e.W + gx])) : 0.f;
    }
    if (tid < 144) {
        const int t = tid >> 4, co = tid & 15;
        w1l[t][co] = float(e.w1[((size_t)co * 9 + t) * e.cinw1]);
    }
    __syncthreads();
    // e0a at the tile's (kTH + 2) x kWideRow pixels: vector i = (pixel, 8
    // channels 8v); pixels outside the image are e0b's zero padding
    {
        constexpr int NFILL = (kTH + 2) * kWideRow * 2;
        constexpr int FILL = (NFILL + kWgThreads - 1) / kWgThreads;
        const int v = tid & 1;
        float bv[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) bv[c] = e.b1[8 * v + c];
#pragma unroll 1
        for (int j = 0; j < FILL; ++j) {
            const int i = tid + j * kWgThreads;
            if (i >= NFILL) break;
            const int pix = i >> 1;
            const int py = pix / kWideRow, px = pix - py * kWideRow;
            const int gy = y0 + py - 1, gx = x0 + px - 1;
            const bool ok = unsigned(gy) < unsigned(e.H) && unsigned(gx) < unsigned(e.W);
            float acc[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = 0.f;
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
                for (int dx = 0; dx < 3; ++dx) {
                    const float xv = zt[py + dy][px + dx];
                    const float4 wa = *reinterpret_cast<const float4 *>(&w1l[dy * 3 + dx][8 * v]);
                    const float4 wb = *reinterpret_cast<const float4 *>(&w1l[dy * 3 + dx][8 * v + 4]);
                    acc[0] = __builtin_fmaf(wa.x, xv, acc[0]);
                    acc[1] = __builtin_fmaf(wa.y, xv, acc[1]);
                    acc[2] = __builtin_fmaf(wa.z, xv, acc[2]);
                    acc[3] = __builtin_fmaf(wa.w, xv, acc[3]);
                    acc[4] = __builtin_fmaf(wb.x, xv, acc[4]);
                    acc[5] = __builtin_fmaf(wb.y, xv, acc[5]);
                    acc[6] = __builtin_fmaf(wb.z, xv, acc[6]);
                    acc[7] = __builtin_fmaf(wb.w, xv, acc[7]);
                }
            }
            bf16x8 o;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const float a = acc[c] + bv[c];
                o[c] = bf16(ok && a > 0.f ? a : 0.f);
            }
            *reinterpret_cast<bf16x8 *>(&tile[0][0][0] + pix * kCCP + 8 * v) = o;
        }
    }
    __syncthreads();
    f32x4 acc[kMT][1];
#pragma unroll
    for (int mt = 0; mt < kMT; ++mt) acc[mt][0] = f32x4{0.f, 0.f, 0.f, 0.f};
    wide_mfma<1, 16>(tile, wave, g, r, B, acc);
    const long long prow = img + (long long)(y0 + wave) * e.W + x0;
    const int Wc = e.W >> 1;
    const long long pq0 = (b * (e.H >> 1) + (y0 >> 1)) * Wc + (x0 >> 1);
    wide_store<kReluBias>(tile, acc, e.b2, 0, e.skip, e.skip_stride, prow, wave, lane, g, r,
                          e.pooled, e.pooled_stride, pq0, Wc);
}

// ---------------------------------------------------------------------------
// Weight gradient: partial[part][co][t][ci] = sum over the part's pixels of
// dZ[p][co] * X[p + tap t][ci]; partial_bias[part][co] = sum dZ[p][co].
// Wave = one (16 co) x (16 ci) tile for all 9 taps over one pixel part.
// ---------------------------------------------------------------------------
struct WgradArgs {
    const bf16 *x;
    const bf16 *dz;
    float *part;       // [parts][Cout][9][Cin]
    float *part_bias;  // [parts][Cout]
    long long P;
    long long px_per_part;
    int parts;
    int x_stride, dz_stride;
    int H, W, Cin, Cout;
};

__device__ __forceinline__ bf16x8 pack8(const bf16 (&v)[8]) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = v[j];
    return r;
}

__global__ void __launch_bounds__(kWgThreads) wgrad_kernel(WgradArgs a) {
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, r = lane & 15;
    const int n_ci = (a.Cin + 15) / 16, n_co = a.Cout / 16;
    const long long wv = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long tiles = (long long)n_ci * n_co;
    if (wv >= tiles * a.parts) return;  // no barriers below
    const int part = int(wv / tiles);
    const int tile = int(wv % tiles);
    const int co0 = (tile / n_ci) * 16, ci0 = (tile % n_ci) * 16;
    const int ci = ci0 + r;
    const bool ci_ok = ci < a.Cin;
    const int cic = ci_ok ? ci : 0;
    const bool do_bias = ci0 == 0;

    f32x4 acc[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    f32x4 accb = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
#pragma unroll
    for (int j = 0; j < 8; ++j) ones[j] = bf16(1.0f);

    const long long pa = (long long)part * a.px_per_part;
    long long pb = pa + a.px_per_part;
    if (pb > a.P) pb = a.P;
    for (long long k0 = pa; k0 < pb; k0 += 32) {
        // this lane group's 8 pixels: k0 + 8g .. +7, one image row (W % 8 == 0)
        const long long p0 = k0 + 8 * g;
        const bool pok = p0 < pb;
        const long long p0c = pok ? p0 : pa;
        const int x0 = int(p0c % a.W);
        const int y0 = int((p0c / a.W) % a.H);
        bf16 av[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bf16 v = a.dz[(p0c + j) * a.dz_stride + co0 + r];
            av[j] = pok ? v : bf16(0.0f);
        }
        const bf16x8 A = pack8(av);
#pragma unroll
        for (int dyi = 0; dyi < 3; ++dyi) {
            const int yy = y0 + dyi - 1;
            const bool yok = pok && ci_ok && unsigned(yy) < unsigned(a.H);
            // window of 10 pixels x0-1 .. x0+8 in row yy
            bf16 win[10];
#pragma unroll
            for (int u = 0; u < 10; ++u) {
                const int xx = x0 + u - 1;
                const bool ok = yok && unsigned(xx) < unsigned(a.W);
                const long long q = ok ? p0c + (long long)(dyi - 1) * a.W + (u - 1) : p0c;
                const bf16 v = a.x[q * a.x_stride + cic];
                win[u] = ok ? v : bf16(0.0f);
            }
#pragma unroll
            for (int dxi = 0; dxi < 3; ++dxi) {
                bf16 bv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) bv[j] = win[j + dxi];
                acc[dyi * 3 + dxi] = mfma(A, pack8(bv), acc[dyi * 3 + dxi]);
            }
        }
        if (do_bias) accb = mfma(A, ones, accb);
    }
    // result lane: column ci0 + r, rows co0 + 4g + j
    float *dst = a.part + (size_t)part * a.Cout * 9 * a.Cin;
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (ci_ok) dst[((size_t)(co0 + 4 * g + j) * 9 + t) * a.Cin + ci] = acc[t][j];
    if (do_bias && r == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) a.part_bias[(size_t)part * a.Cout + co0 + 4 * g + j] = accb[j];
    }
}

// out[i] = sum_p part[p][i], i < count (weights), then the bias rows.
__global__ void __launch_bounds__(kWgThreads)
    reduce_parts_kernel(const float *part, float *out, long long count, int parts) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int p = 0; p < parts; ++p) s += part[(size_t)p * count + i];
    out[i] = s;
}

// W'[ci][t'][co] = W[co][8 - t'][ci]: the input-gradient convolution's weights.
__global__ void __launch_bounds__(kWgThreads)
    transpose_weights_kernel(const bf16 *w, bf16 *wt, int Cin, int Cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = Cout * 9 * Cin;
    if (i >= total) return;
    const int ci = i % Cin, t = (i / Cin) % 9, co = i / (9 * Cin);
    wt[((size_t)ci * 9 + (8 - t)) * Cout + co] = w[i];
}

// ---------------------------------------------------------------------------
// 2x2 average pooling, nearest 2x upsampling and their gradients (NHWC, 8
// channels per thread).
// ---------------------------------------------------------------------------
struct Vec8 {
    float v[8];
};
__device__ __forceinline__ Vec8 to_f32(bf16x8 b) {
    Vec8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r.v[j] = float(b[j]);
    return r;
}
__device__ __forceinline__ bf16x8 to_bf16(const Vec8 &f) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] = bf16(f.v[j]);
    return r;
}
__device__ __forceinline__ void store8(bf16 *p, bf16x8 v) { *reinterpret_cast<bf16x8 *>(p) = v; }

// Grids: blockIdx.y = image, x covers one image's (pixel, 8-channel group)
// items in 32-bit arithmetic.

// y (B, H/2, W/2) = average of 2x2 blocks of x (B, H, W); C channels.
__global__ void __launch_bounds__(kWgThreads)
    pool_kernel(const bf16 *x, int xs, bf16 *y, int ys, int H, int W, int C) {
    const int c8 = C / 8, Wc = W / 2, Hc = H / 2;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Hc * Wc * c8) return;
    const int c = (i % c8) * 8, q = i / c8;
    const int xq = q % Wc, yq = q / Wc;
    const long long b = blockIdx.y;
    const long long p00 = (b * H + 2 * yq) * W + 2 * xq;
    const Vec8 a0 = to_f32(load8(x + p00 * xs + c)), a1 = to_f32(load8(x + (p00 + 1) * xs + c));
    const Vec8 a2 = to_f32(load8(x + (p00 + W) * xs + c)),
               a3 = to_f32(load8(x + (p00 + W + 1) * xs + c));
    Vec8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.v[j] = 0.25f * ((a0.v[j] + a1.v[j]) + (a2.v[j] + a3.v[j]));
    store8(y + ((b * Hc + yq) * Wc + xq) * ys + c, to_bf16(o));
}

// ds (B, H, W) = (dp[parent] / 4 + dskip) * (s > 0)   (H, W: the fine size)
__global__ void __launch_bounds__(kWgThreads)
    pool_backward_kernel(const bf16 *dp, int dps, const bf16 *dskip, int dks, const bf16 *s,
                         int ss, bf16 *ds, int dss, int H, int W, int C) {
    const int c8 = C / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W * c8) return;
    const int c = (i % c8) * 8, pi = i / c8;
    const int xx = pi % W, yy = pi / W;
    const long long b = blockIdx.y;
    const long long p = b * H * W + pi;
    const long long q = (b * (H / 2) + yy / 2) * (W / 2) + xx / 2;
    const Vec8 g = to_f32(load8(dp + q * dps + c));
    const Vec8 k = to_f32(load8(dskip + p * dks + c));
    const Vec8 m = to_f32(load8(s + p * ss + c));
    Vec8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.v[j] = m.v[j] > 0.f ? 0.25f * g.v[j] + k.v[j] : 0.f;
    store8(ds + p * dss + c, to_bf16(o));
}

// y (B, 2H, 2W) = x[parent]   (H, W: the coarse size); a thread writes the
// 2x2 children of one coarse pixel.
__global__ void __launch_bounds__(kWgThreads)
    upsample_kernel(const bf16 *x, int xs, bf16 *y, int ys, int H, int W, int C) {
    const int c8 = C / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W * c8) return;
    const int c = (i % c8) * 8, qi = i / c8;
    const int xq = qi % W, yq = qi / W;
    const long long b = blockIdx.y;
    const bf16x8 v = load8(x + (b * H * W + qi) * xs + c);
    const long long p00 = (b * 2 * H + 2 * yq) * (2 * W) + 2 * xq;
    store8(y + p00 * ys + c, v);
    store8(y + (p00 + 1) * ys + c, v);
    store8(y + (p00 + 2 * W) * ys + c, v);
    store8(y + (p00 + 2 * W + 1) * ys + c, v);
}

// dx (B, H, W) = sum of the 2x2 children of dy (B, 2H, 2W)   (H, W coarse)
__global__ void __launch_bounds__(kWgThreads)
    upsample_backward_kernel(const bf16 *dy, int dys, bf16 *dx, int dxs, int H, int W, int C) {
    const int c8 = C / 8;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H * W * c8) return;
    const int c = (i % c8) * 8, qi = i / c8;
    const int xq = qi % W, yq = qi / W;
    const long long b = blockIdx.y;
    const long long p00 = (b * 2 * H + 2 * yq) * (2 * W) + 2 * xq;
    const Vec8 a0 = to_f32(load8(dy + p00 * dys + c)), a1 = to_f32(load8(dy + (p00 + 1) * dys + c));
    const Vec8 a2 = to_f32(load8(dy + (p00 + 2 * W) * dys + c)),
               a3 = to_f32(load8(dy + (p00 + 2 * W + 1) * dys + c));
    Vec8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o.v[j] = (a0.v[j] + a1.v[j]) + (a2.v[j] + a3.v[j]);
    store8(dx + (b * H * W + qi) * dxs + c, to_bf16(o));
}

// y (P, 8) bf16: channel 0 = scale * x[p], channels 1..7 = 0.
__global__ void __launch_bounds__(kWgThreads)
    pack_input_kernel(const float *x, bf16 *y, long long P, float scale) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = bf16(0.0f);
    v[0] = bf16(scale * x[p]);
    store8(y + p * 8, v);
}

// dz (P, 16) bf16: channel 0 = scale * (pred - target), others 0.
__global__ void __launch_bounds__(kWgThreads)
    loss_grad_kernel(const float *pred, const float *target, bf16 *dz, long long P, float scale) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    bf16x8 v0, v1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        v0[j] = bf16(0.0f);
        v1[j] = bf16(0.0f);
    }
    v0[0] = bf16(scale * (pred[p] - target[p]));
    store8(dz + p * 16, v0);
    store8(dz + p * 16 + 8, v1);
}

// Adam (Kingma & Ba 2015, with bias correction c1 = 1 - beta1^t,
// c2 = 1 - beta2^t) on fp32 master weights; w16 receives the bf16 copy.
__global__ void __launch_bounds__(kWgThreads)
    adam_kernel(float *w, const float *g, float *m, float *v, bf16 *w16, long long count, float lr,
                float b1, float b2, float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = bf16(wi);
}

// ---------------------------------------------------------------------------
// Multiscale preconditioner P r = sum_l D_l^T (h_l * (D_l r)), fp32:
// D = 2x2 averaging, D^T = nearest upsampling / 4, h_l a 5x5 stencil with
// zero padding (symmetric coefficients, so P is symmetric).
// ---------------------------------------------------------------------------
// One block per (output row of the whole batch, chunk of 256 columns):
// row R = b * Hc + yq reads input rows 2R and 2R + 1 (H = 2 Hc).
__global__ void __launch_bounds__(kWgThreads)
    pyr_pool_kernel(const float *__restrict__ x, float *__restrict__ y, int W, int chunks) {
    const unsigned R = blockIdx.x / unsigned(chunks);
    const int xq = int(blockIdx.x - R * unsigned(chunks)) * kWgThreads + int(threadIdx.x);
    const int Wc = W / 2;
    if (xq >= Wc) return;
    const float *top = x + size_t(2 * R) * W + 2 * xq;
    const float2 a = *reinterpret_cast<const float2 *>(top);
    const float2 c = *reinterpret_cast<const float2 *>(top + W);
    y[size_t(R) * Wc + xq] = 0.25f * ((a.x + a.y) + (c.x + c.y));
}

// The same average, two outputs per thread from 16-byte loads, for w / 4 a
// power of two (2^lq pairs per output row) and 16-byte aligned x, 8-byte
// aligned y; thread p -> output row p >> lq, pair p & (2^lq - 1).
__global__ void __launch_bounds__(kWgThreads)
    pyr_pool4_kernel(const float *__restrict__ x, float *__restrict__ y, int W, int lq,
                     long long pairs) {
    const long long p = (long long)blockIdx.x * kWgThreads + threadIdx.x;
    if (p >= pairs) return;
    const long long R = p >> lq;
    const int xp = int(p & ((1LL << lq) - 1));
    const float *top = x + size_t(2 * R) * W + 4 * xp;
    const float4 a = *reinterpret_cast<const float4 *>(top);
    const float4 c = *reinterpret_cast<const float4 *>(top + W);
    float2 o;
    o.x = 0.25f * ((a.x + a.y) + (c.x + c.y));
    o.y = 0.25f * ((a.z + a.w) + (c.z + c.w));
    *reinterpret_cast<float2 *>(y + size_t(R) * (W / 2) + 2 * xp) = o;
}

// y = h * x (5x5, zero padding) + coarse[parent] / 4 (coarse may be null),
// one block per 64 x 32 tile of one image.  The tile and its 2-pixel halo
// are staged in LDS (zeros outside the image); each thread computes a 4 x 2
// block of outputs from a 8 x 6 window read as 128-bit rows.
constexpr int kPyrTW = 64, kPyrTH = 32;
constexpr int kPyrLW = kPyrTW + 4, kPyrLH = kPyrTH + 4;  // row pitch 68: 16-byte rows

__global__ void __launch_bounds__(kWgThreads)
    pyr_stencil_up_kernel(const float *__restrict__ x, const float *__restrict__ coarse,
                          float *__restrict__ y, const float *__restrict__ coef, int H, int W,
                          int tiles_x, int tiles, int vec_store) {
    __shared__ __align__(16) float t[kPyrLH][kPyrLW];
    const unsigned b = blockIdx.x / unsigned(tiles);
    const int tt = int(blockIdx.x - b * unsigned(tiles));
    const int ty = tt / tiles_x, tx = tt - ty * tiles_x;
    const int x0 = tx * kPyrTW, y0 = ty * kPyrTH;
    const float *img = x + size_t(b) * H * W;
    // All of a thread's tile loads are issued before any is stored (points
    // outside the image load the image's first pixel and store 0).
    constexpr int kN = kPyrLH * kPyrLW, kFillPyr = (kN + kWgThreads - 1) / kWgThreads;
    float fv[kFillPyr];
#pragma unroll
    for (int q = 0; q < kFillPyr; ++q) {
        const int i0 = int(threadIdx.x) + q * kWgThreads;
        const int i = i0 < kN ? i0 : 0;
        const int r = i / kPyrLW, c = i - r * kPyrLW;
        const int gy = y0 - 2 + r, gx = x0 - 2 + c;
        const bool ok = unsigned(gy) < unsigned(H) && unsigned(gx) < unsigned(W);
        const float v = img[ok ? size_t(gy) * W + gx : 0];
        fv[q] = ok ? v : 0.f;
    }
    // thread -> columns 4j .. 4j+3, rows 2i, 2i+1 of the tile; rows 2i and
    // 2i + 1 share their coarse parent row, columns 4j .. 4j+3 two parents
    const int j = threadIdx.x & 15, i = threadIdx.x >> 4;
    const int cx = x0 + 4 * j, ry = y0 + 2 * i;
    const bool inside = cx < W && ry < H;
    float par[2] = {0.f, 0.f};
    if (coarse && inside) {
        const float *cr = coarse + (size_t(b) * (H / 2) + ry / 2) * (W / 2) + cx / 2;
        par[0] = cr[0];
        if (cx + 2 < W) par[1] = cr[1];
    }
#pragma unroll
    for (int q = 0; q < kFillPyr; ++q)
        if (int(threadIdx.x) + q * kWgThreads < kN) (&t[0][0])[int(threadIdx.x) + q * kWgThreads] = fv[q];
    float cf[25];
#pragma unroll
    for (int k = 0; k < 25; ++k) cf[k] = coef[k];
    __syncthreads();
    if (!inside) return;
    float win[6][8];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        const float4 a = *reinterpret_cast<const float4 *>(&t[2 * i + r][4 * j]);
        const float4 c = *reinterpret_cast<const float4 *>(&t[2 * i + r][4 * j + 4]);
        win[r][0] = a.x;
        win[r][1] = a.y;
        win[r][2] = a.z;
        win[r][3] = a.w;
        win[r][4] = c.x;
        win[r][5] = c.y;
        win[r][6] = c.z;
        win[r][7] = c.w;
    }
#pragma unroll
    for (int oy = 0; oy < 2; ++oy) {
        const int yy = ry + oy;
        if (yy >= H) break;
        float o[4];
#pragma unroll
        for (int ox = 0; ox < 4; ++ox) {
            float s = 0.f;
#pragma unroll
            for (int dy = 0; dy < 5; ++dy)
#pragma unroll
                for (int dx = 0; dx < 5; ++dx) s += cf[dy * 5 + dx] * win[oy + dy][ox + dx];
            o[ox] = s;
        }
        if (coarse) {
#pragma unroll
            for (int ox = 0; ox < 4; ++ox)
                if (cx + ox < W) o[ox] += 0.25f * par[ox / 2];
        }
        float *dst = y + (size_t(b) * H + yy) * W + cx;
        if (vec_store && cx + 4 <= W) {
            *reinterpret_cast<float4 *>(dst) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int ox = 0; ox < 4; ++ox)
                if (cx + ox < W) dst[ox] = o[ox];
        }
    }
}

// partial[block][o] = sum over the block's pixels p of e[p] * w[p + o],
// o in the 5x5 window (zero outside each image).
__global__ void __launch_bounds__(kWgThreads)
    pyr_correlate_kernel(const float *e, const float *w, float *partial, long long P, int H, int W) {
    __shared__ float red[25][kWgThreads / kWaveLanes];
    float s[25];
#pragma unroll
    for (int o = 0; o < 25; ++o) s[o] = 0.f;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < P;
         p += (long long)gridDim.x * blockDim.x) {
        const int xx = int(p % W), yy = int((p / W) % H);
        const float ev = e[p];
#pragma unroll
        for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
            for (int dx = -2; dx <= 2; ++dx) {
                const bool ok = unsigned(yy + dy) < unsigned(H) && unsigned(xx + dx) < unsigned(W);
                const float wv = ok ? w[p + (long long)dy * W + dx] : 0.f;
                s[(dy + 2) * 5 + dx + 2] += ev * wv;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 0; o < 25; ++o) {
        float v = s[o];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) red[o][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 25) {
        float v = 0.f;
        for (int k = 0; k < kWgThreads / kWaveLanes; ++k) v += red[threadIdx.x][k];
        partial[(size_t)blockIdx.x * 25 + threadIdx.x] = v;
    }
}

// ---------------------------------------------------------------------------
// The network's first and last convolutions at full resolution, on the
// vector ALU: their real channel counts (1 in, 16 out; 16 in, 1 out) would
// fill a 16-wide matrix tile with zeros.  A work-group computes a tile of
// kDTH x kDTW output pixels (4 consecutive pixels per thread) from an LDS
// copy of the input tile with its 1-pixel halo (zero outside the image);
// each output is one chain of fp32 fused multiply-adds over the taps (and
// channels), in tap order, then + bias.
// ---------------------------------------------------------------------------
constexpr int kDTH = 16, kDTW = 64;
constexpr int kDPX = 4;  // pixels per thread (kDTH * kDTW / kWgThreads)

// y[p][co] = relu(bias[co] + sum_t w[co][t][0] * bf16(scale * z[p + tap t]))
// for co < 16 (the first layer: one real input channel; w rows of cinw).
__global__ void __launch_bounds__(kWgThreads) __attribute__((amdgpu_waves_per_eu(4, 8)))
    conv_in_kernel(const float *z, float scale, const bf16 *w, int cinw, const float *bias,
                   bf16 *y, int y_stride, int H, int W) {
    __shared__ float tile[kDTH + 2][kDTW + 3];
    __shared__ float wl[9][16];
    const int tiles_x = (W + kDTW - 1) / kDTW, tiles_y = (H + kDTH - 1) / kDTH;
    int tt = int(blockIdx.x);
    const int tx = tt % tiles_x;
    tt /= tiles_x;
    const int ty = tt % tiles_y;
    const long long b = tt / tiles_y;
    const int x0 = tx * kDTW, y0 = ty * kDTH;
    const float *zi = z + b * H * W;
    for (int i = threadIdx.x; i < (kDTH + 2) * (kDTW + 2); i += kWgThreads) {
        const int r = i / (kDTW + 2), c = i - r * (kDTW + 2);
        const int gy = y0 + r - 1, gx = x0 + c - 1;
        const bool ok = unsigned(gy) < unsigned(H) && unsigned(gx) < unsigned(W);
        tile[r][c] = ok ? float(bf16(scale * zi[(long long)gy * W + gx])) : 0.f;
    }
    if (threadIdx.x < 144) {
        const int t = threadIdx.x >> 4, co = threadIdx.x & 15;
        wl[t][co] = float(w[((size_t)co * 9 + t) * cinw]);
    }
    __syncthreads();
    const int r = threadIdx.x / (kDTW / kDPX), c0 = (threadIdx.x % (kDTW / kDPX)) * kDPX;
    float acc[kDPX][16];
#pragma unroll
    for (int p = 0; p < kDPX; ++p)
#pragma unroll
        for (int co = 0; co < 16; ++co) acc[p][co] = 0.f;
#pragma unroll 1
    for (int dy = 0; dy < 3; ++dy) {
        float xv[kDPX + 2];
#pragma unroll
        for (int u = 0; u < kDPX + 2; ++u) xv[u] = tile[r + dy][c0 + u];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
#pragma unroll
            for (int co = 0; co < 16; ++co) {
                const float wv = wl[dy * 3 + dx][co];
#pragma unroll
                for (int p = 0; p < kDPX; ++p) acc[p][co] = __builtin_fmaf(wv, xv[p + dx], acc[p][co]);
            }
        }
    }
    const int gy = y0 + r;
    if (gy >= H) return;
#pragma unroll
    for (int p = 0; p < kDPX; ++p) {
        const int gx = x0 + c0 + p;
        if (gx >= W) continue;
        bf16x8 v0, v1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float a0 = acc[p][j] + bias[j], a1 = acc[p][8 + j] + bias[8 + j];
            v0[j] = bf16(a0 > 0.f ? a0 : 0.f);
            v1[j] = bf16(a1 > 0.f ? a1 : 0.f);
        }
        bf16 *dst = y + ((b * H + gy) * (long long)W + gx) * y_stride;
        *reinterpret_cast<bf16x8 *>(dst) = v0;
        *reinterpret_cast<bf16x8 *>(dst + 8) = v1;
    }
}

// y32[p] = resid[p] + (sum_{t, ci} w[0][t][ci] * x[p + tap t][ci] + bias[0])
// (the last layer: 16 input channels, one real output channel).
__global__ void __launch_bounds__(kWgThreads) __attribute__((amdgpu_waves_per_eu(4, 8)))
    conv_out_kernel(const bf16 *x, int x_stride, const bf16 *w, const float *bias, float *y32,
                    const float *resid, int H, int W) {
    __shared__ __align__(16) bf16 tile[kDTH + 2][kDTW + 2][16];
    __shared__ float wl[9][16];
    const int tiles_x = (W + kDTW - 1) / kDTW, tiles_y = (H + kDTH - 1) / kDTH;
    int tt = int(blockIdx.x);
    const int tx = tt % tiles_x;
    tt /= tiles_x;
    const int ty = tt % tiles_y;
    const long long b = tt / tiles_y;
    const int x0 = tx * kDTW, y0 = ty * kDTH;
    const long long img = b * H * W;
    for (int i = threadIdx.x; i < (kDTH + 2) * (kDTW + 2) * 2; i += kWgThreads) {
        const int half = i & 1, pix = i >> 1;
        const int r = pix / (kDTW + 2), c = pix - r * (kDTW + 2);
        const int gy = y0 + r - 1, gx = x0 + c - 1;
        const bool ok = unsigned(gy) < unsigned(H) && unsigned(gx) < unsigned(W);
        const long long q = ok ? img + (long long)gy * W + gx : img;
        *reinterpret_cast<bf16x8 *>(&tile[r][c][8 * half]) =
            keep_or_zero(load8(x + q * x_stride + 8 * half), ok);
    }
    if (threadIdx.x < 144) {
        const int t = threadIdx.x >> 4, ci = threadIdx.x & 15;
        wl[t][ci] = float(w[(size_t)t * 16 + ci]);  // output channel 0: w[0][t][ci]
    }
    __syncthreads();
    const int r = threadIdx.x / (kDTW / kDPX), c0 = (threadIdx.x % (kDTW / kDPX)) * kDPX;
    // taps in order; within a tap the 16 channels in order (channels are
    // read 4 at a time, so few values are live at once)
    float acc[kDPX];
#pragma unroll
    for (int p = 0; p < kDPX; ++p) acc[p] = 0.f;
#pragma unroll 1
    for (int dy = 0; dy < 3; ++dy) {
#pragma unroll 1
        for (int dx = 0; dx < 3; ++dx) {
#pragma unroll
            for (int c4 = 0; c4 < 16; c4 += 4) {
                float wv[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) wv[c] = wl[dy * 3 + dx][c4 + c];
#pragma unroll
                for (int p = 0; p < kDPX; ++p) {
                    const bf16 *px = &tile[r + dy][c0 + p + dx][c4];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[p] = __builtin_fmaf(wv[c], float(px[c]), acc[p]);
                }
            }
        }
    }
    const int gy = y0 + r;
    if (gy >= H) return;
#pragma unroll
    for (int p = 0; p < kDPX; ++p) {
        const int gx = x0 + c0 + p;
        if (gx >= W) continue;
        const long long q = img + (long long)gy * W + gx;
        y32[q] = resid[q] + (acc[p] + bias[0]);
    }
}

// ---------------------------------------------------------------------------
// Phantoms: out[b][i][j] = sum of val over the ellipses of image b whose
// closed interior contains the pixel centre (x, y) = (-1 + (2j+1)/n,
// 1 - (2i+1)/n); params[b][e] = (cx, cy, cos, sin, ax, ay, val), float64.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWgThreads)
    phantom_kernel(const double *params, const int *counts, int max_e, float *out, long long P,
                   int n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int j = int(p % n), i = int((p / n) % n);
    const long long b = p / ((long long)n * n);
    const double xs = -1.0 + (2.0 * j + 1.0) / n;
    const double ys = 1.0 - (2.0 * i + 1.0) / n;
    const double *e = params + b * max_e * 7;
    double acc = 0.0;
    for (int k = 0; k < counts[b]; ++k, e += 7) {
        const double dx = xs - e[0], dy = ys - e[1];
        const double u = (dx * e[2] + dy * e[3]) / e[4];
        const double v = (-dx * e[3] + dy * e[2]) / e[5];
        if (u * u + v * v <= 1.0) acc += e[6];
    }
    out[p] = float(acc);
}

inline unsigned blocks_for(long long work, int block = kWgThreads) {
    long long g = (work + block - 1) / block;
    return unsigned(g < 1 ? 1 : g);
}

template <int NT>
static void launch_conv_lds(const ConvArgs &a, int mode, hipStream_t s) {
    if (a.W % kTWmax == 0 && a.Cin % 16 == 0 && a.Cout == 16 &&
        (mode == kReluBias || mode == kLinear)) {
        const long long tiles = (a.P / a.W / kTH) * (a.W / kTWmax);
        const unsigned grid = unsigned(tiles * (a.Cout / 16));
        if (mode == kReluBias)
            hipLaunchKernelGGL((conv3x3_wide_kernel<kReluBias>), dim3(grid), dim3(kWgThreads), 0, s,
                               a, int(tiles));
        else
            hipLaunchKernelGGL((conv3x3_wide_kernel<kLinear>), dim3(grid), dim3(kWgThreads), 0, s,
                               a, int(tiles));
        return;
    }
    const int TW = a.W < kTWmax ? a.W : kTWmax;
    // floor(p / (TW + 2)) == (p * recip) >> 16 for every tile pixel index p
    // (checked for TW in {16, 32, 48, 64})
    const int recip = (65536 + TW + 1) / (TW + 2);
    const long long tiles = (a.P / a.W / kTH) * (a.W / TW);
    const dim3 grid(unsigned(tiles), unsigned(a.Cout / (16 * NT)));
    switch (mode) {
        case kReluBias: hipLaunchKernelGGL((conv3x3_lds_kernel<NT, kReluBias>), grid, dim3(kWgThreads), 0, s, a, TW, recip); break;
        case kResidual: hipLaunchKernelGGL((conv3x3_lds_kernel<NT, kResidual>), grid, dim3(kWgThreads), 0, s, a, TW, recip); break;
        case kMaskInput: hipLaunchKernelGGL((conv3x3_lds_kernel<NT, kMaskInput>), grid, dim3(kWgThreads), 0, s, a, TW, recip); break;
        default: hipLaunchKernelGGL((conv3x3_lds_kernel<NT, kLinear>), grid, dim3(kWgThreads), 0, s, a, TW, recip); break;
    }
}

template <int NT>
static void launch_conv_nt(const ConvArgs &a, int mode, hipStream_t s) {
    // the tile fill takes 8, 16 or 32 channels per chunk (power-of-two vectors)
    if (a.W % 16 == 0 && a.H % kTH == 0 && a.Cin >= 16 && a.Cin % kCC != 24 &&
        (a.W <= kTWmax || a.W % kTWmax == 0)) {
        launch_conv_lds<NT>(a, mode, s);
        return;
    }
    const long long px_blocks = a.P / 64;
    const unsigned grid = unsigned((px_blocks + 3) / 4 * (a.Cout / (16 * NT)));
    switch (mode) {
        case kReluBias: hipLaunchKernelGGL((conv3x3_kernel<NT, kReluBias>), dim3(grid), dim3(kWgThreads), 0, s, a); break;
        case kResidual: hipLaunchKernelGGL((conv3x3_kernel<NT, kResidual>), dim3(grid), dim3(kWgThreads), 0, s, a); break;
        case kMaskInput: hipLaunchKernelGGL((conv3x3_kernel<NT, kMaskInput>), dim3(grid), dim3(kWgThreads), 0, s, a); break;
        default: hipLaunchKernelGGL((conv3x3_kernel<NT, kLinear>), dim3(grid), dim3(kWgThreads), 0, s, a); break;
    }
}

struct WgradPlan {
    int parts;
    long long px_per_part;
};

inline WgradPlan wgrad_plan(long long P, int Cin, int Cout) {
    const long long tiles = (long long)((Cin + 15) / 16) * (Cout / 16);
    long long parts = 8192 / tiles;
    if (parts < 1) parts = 1;
    const long long max_parts = (P + 31) / 32;
    if (parts > max_parts) parts = max_parts;
    long long per = (P + parts - 1) / parts;
    per = (per + 31) / 32 * 32;
    WgradPlan w;
    w.px_per_part = per;
    w.parts = int((P + per - 1) / per);
    return w;
}

}  // namespace nn

using namespace nn;

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

extern "C" {

int adrt_hip_nn_conv3x3(const void *x, int x_stride, const void *w, const float *bias, void *y,
                        int y_stride, float *y32, const float *resid, const void *mask,
                        int mask_stride, int batch, int h, int w_, int cin, int cout, int mode,
                        void *stream) {
    if (batch < 1 || h < 1 || w_ < 1 || cin < 8 || cin % 8 != 0 || cout < 16 ||
        cout % 16 != 0 || x_stride < cin || x_stride % 8 != 0 || mode < 0 || mode > 3)
        return ADRT_HIP_ERR_ARG;
    const long long P = (long long)batch * h * w_;
    if (P % 64 != 0) return ADRT_HIP_ERR_ARG;
    if (!x || !w || !aligned16(x) || !aligned16(w)) return ADRT_HIP_ERR_NULL;
    if (mode == kResidual) {
        if (!y32 || !resid || !bias) return ADRT_HIP_ERR_NULL;
    } else {
        if (!y || y_stride < cout) return ADRT_HIP_ERR_ARG;
        if (!aligned16(y)) return ADRT_HIP_ERR_NULL;
        if (mode == kReluBias && !bias) return ADRT_HIP_ERR_NULL;
        if (mode == kMaskInput && (!mask || mask_stride < cout)) return ADRT_HIP_ERR_NULL;
    }
    ConvArgs a;
    a.x = static_cast<const bf16 *>(x);
    a.w = static_cast<const bf16 *>(w);
    a.bias = bias;
    a.y = static_cast<bf16 *>(y);
    a.y32 = y32;
    a.resid = resid;
    a.mask = static_cast<const bf16 *>(mask);
    a.P = P;
    a.x_stride = x_stride;
    a.y_stride = y_stride;
    a.mask_stride = mask_stride;
    a.H = h;
    a.W = w_;
    a.Cin = cin;
    a.Cout = cout;
    a.x2 = nullptr;
    a.x2_stride = 0;
    a.c_split = cin;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (cout % 64 == 0)
        launch_conv_nt<4>(a, mode, s);
    else if (cout % 32 == 0)
        launch_conv_nt<2>(a, mode, s);
    else
        launch_conv_nt<1>(a, mode, s);
    return launch_status();
}

int adrt_hip_nn_conv3x3_upcat(const void *x, int x_stride, int c_skip, const void *up,
                              int up_stride, const void *w, const float *bias, void *y,
                              int y_stride, int batch, int h, int w_, int cin, int cout, int mode,
                              void *stream) {
    if (batch < 1 || h < 4 || w_ < 16 || h % kTH != 0 || w_ % 16 != 0 ||
        (w_ > kTWmax && w_ % kTWmax != 0) || cin < 16 || cin % 8 != 0 || cin % kCC == 24 ||
        cout < 16 ||
        cout % 16 != 0 || c_skip < 8 || c_skip % 8 != 0 || c_skip >= cin ||
        x_stride < c_skip || x_stride % 8 != 0 || up_stride < cin - c_skip ||
        up_stride % 8 != 0 || y_stride < cout || (mode != kReluBias && mode != kLinear))
        return ADRT_HIP_ERR_ARG;
    if (!x || !up || !w || !y || !aligned16(x) || !aligned16(up) || !aligned16(w) ||
        !aligned16(y))
        return ADRT_HIP_ERR_NULL;
    if (mode == kReluBias && !bias) return ADRT_HIP_ERR_NULL;
    ConvArgs a;
    a.x = static_cast<const bf16 *>(x);
    a.w = static_cast<const bf16 *>(w);
    a.bias = bias;
    a.y = static_cast<bf16 *>(y);
    a.y32 = nullptr;
    a.resid = nullptr;
    a.mask = nullptr;
    a.P = (long long)batch * h * w_;
    a.x_stride = x_stride;
    a.y_stride = y_stride;
    a.mask_stride = 0;
    a.H = h;
    a.W = w_;
    a.Cin = cin;
    a.Cout = cout;
    a.x2 = static_cast<const bf16 *>(up);
    a.x2_stride = up_stride;
    a.c_split = c_skip;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (cout % 64 == 0)
        launch_conv_lds<4>(a, mode, s);
    else if (cout % 32 == 0)
        launch_conv_lds<2>(a, mode, s);
    else
        launch_conv_lds<1>(a, mode, s);
    return launch_status();
}

size_t adrt_hip_nn_wgrad_workspace(int batch, int h, int w, int cin, int cout) {
    if (batch < 1 || h < 1 || w < 8 || cin < 8 || cout < 16) return 0;
    const long long P = (long long)batch * h * w;
    const WgradPlan pl = wgrad_plan(P, cin, cout);
    return size_t(pl.parts) * (size_t(cout) * 9 * cin + cout) * sizeof(float);
}

int adrt_hip_nn_wgrad(const void *x, int x_stride, const void *dz, int dz_stride, float *dw,
                      float *dbias, void *ws, size_t ws_bytes, int batch, int h, int w, int cin,
                      int cout, void *stream) {
    if (batch < 1 || h < 1 || w < 8 || w % 8 != 0 || cin < 8 || cin % 8 != 0 || cout < 16 ||
        cout % 16 != 0 || x_stride < cin || dz_stride < cout)
        return ADRT_HIP_ERR_ARG;
    if (!x || !dz || !dw || !dbias || !ws) return ADRT_HIP_ERR_NULL;
    if (ws_bytes < adrt_hip_nn_wgrad_workspace(batch, h, w, cin, cout)) return ADRT_HIP_ERR_WORKSPACE;
    const long long P = (long long)batch * h * w;
    const WgradPlan pl = wgrad_plan(P, cin, cout);
    WgradArgs a;
    a.x = static_cast<const bf16 *>(x);
    a.dz = static_cast<const bf16 *>(dz);
    a.part = static_cast<float *>(ws);
    a.part_bias = a.part + (size_t)pl.parts * cout * 9 * cin;
    a.P = P;
    a.px_per_part = pl.px_per_part;
    a.parts = pl.parts;
    a.x_stride = x_stride;
    a.dz_stride = dz_stride;
    a.H = h;
    a.W = w;
    a.Cin = cin;
    a.Cout = cout;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long waves = (long long)((cin + 15) / 16) * (cout / 16) * pl.parts;
    hipLaunchKernelGGL(wgrad_kernel, dim3(unsigned((waves + 3) / 4)), dim3(kWgThreads), 0, s, a);
    const long long nw = (long long)cout * 9 * cin;
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(nw)), dim3(kWgThreads), 0, s,
                       (const float *)a.part, dw, nw, pl.parts);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(blocks_for(cout)), dim3(kWgThreads), 0, s,
                       (const float *)a.part_bias, dbias, (long long)cout, pl.parts);
    return launch_status();
}

int adrt_hip_nn_transpose_weights(const void *w, void *wt, int cin, int cout, void *stream) {
    if (cin < 1 || cout < 1) return ADRT_HIP_ERR_ARG;
    if (!w || !wt) return ADRT_HIP_ERR_NULL;
    const int total = cout * 9 * cin;
    hipLaunchKernelGGL(transpose_weights_kernel, dim3(blocks_for(total)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16 *>(w),
                       static_cast<bf16 *>(wt), cin, cout);
    return launch_status();
}

static bool nhwc_ok(int h, int w, int c, int stride) {
    return h >= 2 && w >= 2 && h % 2 == 0 && w % 2 == 0 && c >= 8 && c % 8 == 0 && stride >= c &&
           stride % 8 == 0;
}

int adrt_hip_nn_pool(const void *x, int x_stride, void *y, int y_stride, int batch, int h, int w,
                     int c, void *stream) {
    if (batch < 1 || !nhwc_ok(h, w, c, x_stride) || y_stride < c || y_stride % 8) return ADRT_HIP_ERR_ARG;
    if (!x || !y) return ADRT_HIP_ERR_NULL;
    const long long items = (long long)(h / 2) * (w / 2) * (c / 8);
    if (items >= (1ll << 31) || batch > 65535) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(pool_kernel, dim3(blocks_for(items), unsigned(batch)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16 *>(x), x_stride,
                       static_cast<bf16 *>(y), y_stride, h, w, c);
    return launch_status();
}

int adrt_hip_nn_pool_backward(const void *dp, int dp_stride, const void *dskip, int dskip_stride,
                              const void *s, int s_stride, void *ds, int ds_stride, int batch,
                              int h, int w, int c, void *stream) {
    if (batch < 1 || !nhwc_ok(h, w, c, s_stride) || dp_stride < c || dskip_stride < c ||
        ds_stride < c || dp_stride % 8 || dskip_stride % 8 || ds_stride % 8)
        return ADRT_HIP_ERR_ARG;
    if (!dp || !dskip || !s || !ds) return ADRT_HIP_ERR_NULL;
    const long long items = (long long)h * w * (c / 8);
    if (items >= (1ll << 31) || batch > 65535) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(pool_backward_kernel, dim3(blocks_for(items), unsigned(batch)),
                       dim3(kWgThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16 *>(dp), dp_stride, static_cast<const bf16 *>(dskip),
                       dskip_stride, static_cast<const bf16 *>(s), s_stride,
                       static_cast<bf16 *>(ds), ds_stride, h, w, c);
    return launch_status();
}

int adrt_hip_nn_upsample(const void *x, int x_stride, void *y, int y_stride, int batch, int h,
                         int w, int c, void *stream) {
    if (batch < 1 || h < 1 || w < 1 || c < 8 || c % 8 || x_stride < c || x_stride % 8 ||
        y_stride < c || y_stride % 8)
        return ADRT_HIP_ERR_ARG;
    if (!x || !y) return ADRT_HIP_ERR_NULL;
    const long long items = (long long)h * w * (c / 8);
    if (items >= (1ll << 31) || batch > 65535) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(upsample_kernel, dim3(blocks_for(items), unsigned(batch)), dim3(kWgThreads),
                       0, static_cast<hipStream_t>(stream), static_cast<const bf16 *>(x), x_stride,
                       static_cast<bf16 *>(y), y_stride, h, w, c);
    return launch_status();
}

int adrt_hip_nn_upsample_backward(const void *dy, int dy_stride, void *dx, int dx_stride,
                                  int batch, int h, int w, int c, void *stream) {
    if (batch < 1 || h < 1 || w < 1 || c < 8 || c % 8 || dy_stride < c || dy_stride % 8 ||
        dx_stride < c || dx_stride % 8)
        return ADRT_HIP_ERR_ARG;
    if (!dy || !dx) return ADRT_HIP_ERR_NULL;
    const long long items = (long long)h * w * (c / 8);
    if (items >= (1ll << 31) || batch > 65535) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(upsample_backward_kernel, dim3(blocks_for(items), unsigned(batch)),
                       dim3(kWgThreads), 0, static_cast<hipStream_t>(stream),
                       static_cast<const bf16 *>(dy), dy_stride, static_cast<bf16 *>(dx), dx_stride,
                       h, w, c);
    return launch_status();
}

int adrt_hip_nn_pack_input(const float *x, void *y, int batch, int h, int w, float scale,
                           void *stream) {
    if (batch < 1 || h < 1 || w < 1) return ADRT_HIP_ERR_ARG;
    if (!x || !y) return ADRT_HIP_ERR_NULL;
    const long long P = (long long)batch * h * w;
    hipLaunchKernelGGL(pack_input_kernel, dim3(blocks_for(P)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), x, static_cast<bf16 *>(y), P, scale);
    return launch_status();
}

int adrt_hip_nn_conv_in(const float *z, float scale, const void *w, int cinw, const float *bias,
                        void *y, int y_stride, int batch, int h, int w_, void *stream) {
    if (batch < 1 || h < 1 || w_ < 1 || cinw < 1 || y_stride < 16 || y_stride % 8 != 0)
        return ADRT_HIP_ERR_ARG;
    if (!z || !w || !bias || !y || !aligned16(y)) return ADRT_HIP_ERR_NULL;
    const long long tiles =
        (long long)batch * ((h + kDTH - 1) / kDTH) * ((w_ + kDTW - 1) / kDTW);
    hipLaunchKernelGGL(conv_in_kernel, dim3(unsigned(tiles)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), z, scale, static_cast<const bf16 *>(w),
                       cinw, bias, static_cast<bf16 *>(y), y_stride, h, w_);
    return launch_status();
}

int adrt_hip_nn_encoder0(const float *z, float scale, const void *w1, int cinw1, const float *b1,
                         const void *w2, const float *b2, void *skip, int skip_stride,
                         void *pooled, int pooled_stride, int batch, int h, int w_,
                         void *stream) {
    if (batch < 1 || h < kTH || w_ < kTWmax || h % kTH != 0 || w_ % kTWmax != 0 || cinw1 < 1 ||
        skip_stride < 16 || skip_stride % 8 != 0 || pooled_stride < 16 || pooled_stride % 8 != 0)
        return ADRT_HIP_ERR_ARG;
    if (!z || !w1 || !b1 || !w2 || !b2 || !skip || !pooled || !aligned16(w2) || !aligned16(skip) ||
        !aligned16(pooled))
        return ADRT_HIP_ERR_NULL;
    Enc0Args e;
    e.z = z;
    e.scale = scale;
    e.w1 = static_cast<const bf16 *>(w1);
    e.cinw1 = cinw1;
    e.b1 = b1;
    e.w2 = static_cast<const bf16 *>(w2);
    e.b2 = b2;
    e.skip = static_cast<bf16 *>(skip);
    e.skip_stride = skip_stride;
    e.pooled = static_cast<bf16 *>(pooled);
    e.pooled_stride = pooled_stride;
    e.H = h;
    e.W = w_;
    const long long tiles = (long long)batch * (h / kTH) * (w_ / kTWmax);
    hipLaunchKernelGGL(encoder0_kernel, dim3(unsigned(tiles)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), e);
    return launch_status();
}

int adrt_hip_nn_conv_out(const void *x, int x_stride, const void *w, const float *bias, float *y32,
                         const float *resid, int batch, int h, int w_, void *stream) {
    if (batch < 1 || h < 1 || w_ < 1 || x_stride < 16 || x_stride % 8 != 0)
        return ADRT_HIP_ERR_ARG;
    if (!x || !w || !bias || !y32 || !resid || !aligned16(x)) return ADRT_HIP_ERR_NULL;
    const long long tiles =
        (long long)batch * ((h + kDTH - 1) / kDTH) * ((w_ + kDTW - 1) / kDTW);
    hipLaunchKernelGGL(conv_out_kernel, dim3(unsigned(tiles)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), static_cast<const bf16 *>(x), x_stride,
                       static_cast<const bf16 *>(w), bias, y32, resid, h, w_);
    return launch_status();
}

int adrt_hip_nn_loss_grad(const float *pred, const float *target, void *dz, int batch, int h,
                          int w, float scale, void *stream) {
    if (batch < 1 || h < 1 || w < 1) return ADRT_HIP_ERR_ARG;
    if (!pred || !target || !dz) return ADRT_HIP_ERR_NULL;
    const long long P = (long long)batch * h * w;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(P)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), pred, target, static_cast<bf16 *>(dz), P,
                       scale);
    return launch_status();
}

int adrt_hip_nn_adam(float *w, const float *g, float *m, float *v, void *w16, size_t count,
                     float lr, float beta1, float beta2, float eps, float c1, float c2,
                     void *stream) {
    if (!w || !g || !m || !v) return ADRT_HIP_ERR_NULL;
    if (count == 0) return ADRT_HIP_OK;
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for((long long)count)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), w, g, m, v, static_cast<bf16 *>(w16),
                       (long long)count, lr, beta1, beta2, eps, c1, c2);
    return launch_status();
}

int adrt_hip_pyr_pool(const float *x, float *y, int batch, int h, int w, void *stream) {
    if (batch < 1 || h < 2 || w < 2 || h % 2 || w % 2) return ADRT_HIP_ERR_ARG;
    if (!x || !y) return ADRT_HIP_ERR_NULL;
    if (reinterpret_cast<uintptr_t>(x) % 8) return ADRT_HIP_ERR_ARG;  // float2 loads
    const int q = w / 4;
    if (w % 4 == 0 && (q & (q - 1)) == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(y) % 8 == 0) {
        int lq = 0;
        while ((1 << lq) < q) ++lq;
        const long long pairs = (long long)batch * (h / 2) * q;
        const long long nb = (pairs + kWgThreads - 1) / kWgThreads;
        if (nb > 0x7fffffffLL) return ADRT_HIP_ERR_TOO_BIG;
        hipLaunchKernelGGL(pyr_pool4_kernel, dim3(unsigned(nb), 1, 1), dim3(kWgThreads), 0,
                           static_cast<hipStream_t>(stream), x, y, w, lq, pairs);
        return launch_status();
    }
    const int chunks = (w / 2 + kWgThreads - 1) / kWgThreads;
    const long long blocks = (long long)batch * (h / 2) * chunks;
    if (blocks > 0x7fffffffLL) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(pyr_pool_kernel, dim3(unsigned(blocks), 1, 1), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), x, y, w, chunks);
    return launch_status();
}

int adrt_hip_pyr_stencil_up(const float *x, const float *coarse, float *y, const float *coef25,
                            int batch, int h, int w, void *stream) {
    if (batch < 1 || h < 1 || w < 1 || (coarse && (h % 2 || w % 2))) return ADRT_HIP_ERR_ARG;
    if (!x || !y || !coef25) return ADRT_HIP_ERR_NULL;
    const int tiles_x = (w + kPyrTW - 1) / kPyrTW;
    const int tiles = tiles_x * ((h + kPyrTH - 1) / kPyrTH);
    const long long blocks = (long long)batch * tiles;
    if (blocks > 0x7fffffffLL) return ADRT_HIP_ERR_TOO_BIG;
    hipLaunchKernelGGL(pyr_stencil_up_kernel, dim3(unsigned(blocks), 1, 1), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), x, coarse, y, coef25, h, w, tiles_x,
                       tiles, int(w % 4 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0));
    return launch_status();
}

int adrt_hip_pyr_correlate(const float *e, const float *w, float *partial, int parts, int batch,
                           int h, int wd, void *stream) {
    if (batch < 1 || h < 1 || wd < 1 || parts < 1) return ADRT_HIP_ERR_ARG;
    if (!e || !w || !partial) return ADRT_HIP_ERR_NULL;
    const long long P = (long long)batch * h * wd;
    hipLaunchKernelGGL(pyr_correlate_kernel, dim3(unsigned(parts)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), e, w, partial, P, h, wd);
    return launch_status();
}

int adrt_hip_phantom_raster(const double *params, const int *counts, int max_ellipses, float *out,
                            int batch, int n, void *stream) {
    if (batch < 1 || n < 1 || max_ellipses < 1) return ADRT_HIP_ERR_ARG;
    if (!params || !counts || !out) return ADRT_HIP_ERR_NULL;
    const long long P = (long long)batch * n * n;
    hipLaunchKernelGGL(phantom_kernel, dim3(blocks_for(P)), dim3(kWgThreads), 0,
                       static_cast<hipStream_t>(stream), params, counts, max_ellipses, out, P, n);
    return launch_status();
}

}  // extern "C"

}  // namespace adrt_hip
