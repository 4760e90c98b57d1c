// Radial DCT-domain preconditioner of the learned CG solver (no reference
// counterpart): y = IDCT2(H * DCT2(x)) for float32 images (batch, n, n),
// with the orthonormal DCT-II along both axes, H an (n, n) table of positive
// weights and IDCT2 its transpose (the orthonormal DCT-III).  H radial and
// positive makes the filter symmetric positive definite.
//
// A 1-D orthonormal DCT-II of length n (Makhoul 1980): reorder
// v[k] = x[2k], v[n-1-k] = x[2k+1] (k < n/2); V = DFT_n(v);
// X[m] = 2 Re(e^{-i pi m / 2n} V[m]) * s(m), s(0) = sqrt(1/4n),
// s(m > 0) = sqrt(1/2n).  Its inverse: Y[m] = X[m] / s(m),
// V[m] = (1/2) e^{i pi m / 2n} (Y[m] - i Y[n-m]) (Y[n] = 0), v = IDFT_n(V),
// x[2k] = v[k], x[2k+1] = v[n-1-k].  The DFTs are in-LDS radix-2
// decimation-in-frequency FFTs (natural order in, bit-reversed out), so no
// permutation pass is needed: the post-processing reads V[bitrev(m)].
//
// A work-group holds R = 8192 / n whole rows, two real rows per complex FFT
// row (4096 complex values, 32 KB of LDS), and the twiddles (n <= 1024).
// The 2-D transform is three row passes over the batch:
//   pass 0 (kForward):  rows of x -> DCT-II -> written transposed
//   pass 1 (kFiltered): rows (now columns of x) -> DCT-II -> * H -> DCT-III
//                       -> written transposed (back to x's orientation)
//   pass 2 (kInverse):  rows -> DCT-III -> written in place
// Transposed stores go through LDS with a padded pitch so that each column
// of R rows leaves as one contiguous run.
#include <cmath>

#include "common.hpp"

namespace adrt_hip {
namespace dctp {

constexpr int kThreads = 256;
constexpr int kElems = 4096;                  // complex values per work-group
constexpr int kPer = kElems / kThreads;       // elements per thread
constexpr int kMaxN = 1024;                   // longest row (the twiddles live in LDS)

enum Mode : int { kForward = 0, kFiltered = 1, kInverse = 2 };

struct Args {
    const float *in;
    float *out;
    const float2 *tw;  // [n/2] e^{-2 pi i t / n}, then [n] e^{-i pi m / 2n}
    const float *H;    // (n, n), kFiltered only
    long long rows;    // batch * n
    float s0, s1;      // s(0), s(m > 0)
    float r0, r1;      // 1 / s(0), 1 / s(m > 0)
};

__device__ __forceinline__ int bitrev(int m, int L) {
    return int(__builtin_bitreverse32(unsigned(m)) >> (32 - L));
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// LDS position of complex element e of the work-group's rows: one padding
// element after every 16, so the strided accesses of the FFT steps (stride
// 16 elements in the last radix-16 step) fall on distinct banks.
__device__ __forceinline__ int pad(int e) { return e + (e >> 4); }
constexpr int kPadded = kElems + kElems / 16;
// LDS buffer (complex elements): the padded FFT rows, or as reals the
// transposed staging of 2 kElems / n rows of n (+1 per column)
constexpr int kStaged = (2 * kElems + kMaxN + 1) / 2;
constexpr int kBuf = kPadded > kStaged ? kPadded : kStaged;

__device__ __forceinline__ float2 twid(const float2 *tw, int t, bool inverse) {
    float2 w = tw[t];
    if (inverse) w.y = -w.y;
    return w;
}

// In-place DIF FFT of the R rows of `buf` (row r holds elements r * n ...,
// at LDS positions pad(.)); forward (e^{-2 pi i ...}) or inverse (conjugate
// twiddles, unscaled); `tw` (in LDS) holds W^t = e^{-2 pi i t / n}, t < n/2.
// Radix-2 stage s pairs elements half = n >> (s + 1) apart, with twiddle
// W^(pos << s) on the difference (pos = position within the half).  Step<K>
// does stages s .. s+K-1 in registers on the 2^K elements
// {base + t q : t < 2^K} (q = n >> (s + K)), with exactly the operations of
// those K radix-2 stages: stage s+u pairs t with t + h, h = 2^(K-1-u), at
// position (t mod h) q + p.  Steps of 4 stages (radix 16) go first; one
// shorter step finishes.
//
// log2 n (L), the first stage (s) and the direction are compile-time, so
// the index arithmetic folds into constants and shifts.
template <int K, int L, int s, bool inverse>
__device__ __forceinline__ void fft_step(float2 *buf, const float2 *tw) {
    constexpr int E = 1 << K;
    constexpr int lq = L - s - K;  // log2 q
    constexpr int q = 1 << lq;
    constexpr int gpr = 1 << (L - K);  // groups per row
#pragma unroll
    for (int j = threadIdx.x; j < kElems / E; j += kThreads) {
        const int r = j >> (L - K), jj = j & (gpr - 1);
        const int g = jj >> lq, p = jj & (q - 1);
        const int base = (r << L) + (g << (lq + K)) + p;
        float2 a[E];
#pragma unroll
        for (int t = 0; t < E; ++t) a[t] = buf[pad(base + t * q)];
#pragma unroll
        for (int u = 0; u < K; ++u) {
            const int h = E >> (u + 1);
#pragma unroll
            for (int t = 0; t < E; ++t) {
                if (t & h) continue;
                const float2 w = twid(tw, ((t & (h - 1)) * q + p) << (s + u), inverse);
                const float2 x = a[t], y = a[t + h];
                a[t] = make_float2(x.x + y.x, x.y + y.y);
                a[t + h] = cmul(make_float2(x.x - y.x, x.y - y.y), w);
            }
        }
#pragma unroll
        for (int t = 0; t < E; ++t) buf[pad(base + t * q)] = a[t];
    }
    __syncthreads();
}

template <int L, bool inverse, int s = 0>
__device__ __forceinline__ void fft_dif(float2 *buf, const float2 *tw) {
    if constexpr (s + 4 <= L) {
        fft_step<4, L, s, inverse>(buf, tw);
        fft_dif<L, inverse, s + 4>(buf, tw);
    } else if constexpr (s < L) {
        fft_step<L - s, L, s, inverse>(buf, tw);
    }
}

// Two real rows per complex FFT: real rows 2c and 2c + 1 of the work-group
// (reordered as above: v_a, v_b) travel as one complex row z = v_a + i v_b
// of length n, so a work-group holds 2 * kElems / n real rows.
//   forward: Z = DFT(z); with Zc = conj(Z[(n - m) mod n]),
//            V_a = (Z[m] + Zc) / 2 and V_b = (Z[m] - Zc) / (2i), so
//            X_a[m] = Re(pt[m] (Z[m] + Zc)) s(m), X_b[m] = Im(pt[m] (Z[m] - Zc)) s(m);
//   inverse: V_a, V_b are Hermitian (V[n - m] = conj(V[m]) for real Y), so
//            IDFT(V_a + i V_b) = v_a + i v_b: real parts row 2c, imaginary 2c + 1.
template <int MODE, int L>
__global__ void __launch_bounds__(kThreads) dct_rows_kernel(Args a) {
    __shared__ __align__(16) float2 buf[kBuf];
    extern __shared__ __align__(16) float2 tws[];  // n/2 + n twiddles (dynamic LDS)
    float *stg = reinterpret_cast<float *>(buf);  // real staging view (2 * kBuf floats)
    constexpr int n = 1 << L;
    constexpr int R = 2 * (kElems >> L);  // real rows per work-group
    const long long g0 = (long long)blockIdx.x * R;  // first row of this work-group
    const int tid = threadIdx.x;
    for (int t = tid; t < (n >> 1) + n; t += kThreads) tws[t] = a.tw[t];
    const float2 *tw = tws;
    const float2 *pt = tws + (n >> 1);
    // element q of the thread: complex row c = e >> L, column e & (n - 1),
    // carrying real rows 2c (va) and 2c + 1 (vb)
    float va[kPer], vb[kPer];

    if (MODE != kInverse) {
        // load two rows per complex row, Makhoul reorder into the buffer
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const int c = e >> L, k = e & (n - 1);
            const long long ra = g0 + 2 * c;
            const float xa = ra < a.rows ? a.in[ra * n + k] : 0.f;
            const float xb = ra + 1 < a.rows ? a.in[(ra + 1) * n + k] : 0.f;
            const int vp = (k & 1) ? n - 1 - (k >> 1) : (k >> 1);
            buf[pad((c << L) + vp)] = make_float2(xa, xb);
        }
        __syncthreads();
        fft_dif<L, false>(buf, tw);
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const int c = e >> L, m = e & (n - 1);
            const float2 Zm = buf[pad((c << L) + bitrev(m, L))];
            const float2 Zn = buf[pad((c << L) + bitrev((n - m) & (n - 1), L))];
            const float ax = Zm.x + Zn.x, ay = Zm.y - Zn.y;  // Z[m] + conj(Z[n-m])
            const float bx = Zm.x - Zn.x, by = Zm.y + Zn.y;  // Z[m] - conj(Z[n-m])
            const float2 p = pt[m];
            const float sm = m == 0 ? a.s0 : a.s1;
            va[q] = (p.x * ax - p.y * ay) * sm;
            vb[q] = (p.x * by + p.y * bx) * sm;
        }
        if (MODE == kFiltered) {
            // * H[j][m] with j = the row's index in its image (the frequency
            // of the other axis, after the transposed first pass)
#pragma unroll
            for (int q = 0; q < kPer; ++q) {
                const int e = tid + q * kThreads;
                const int c = e >> L, m = e & (n - 1);
                const int ja = int((g0 + 2 * c) & (n - 1));
                va[q] *= a.H[(size_t)ja * n + m];
                vb[q] *= a.H[(size_t)(ja + 1) * n + m];
            }
        }
        __syncthreads();  // every read of buf is done
    } else {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const long long ra = g0 + 2 * (e >> L);
            const int k = e & (n - 1);
            va[q] = ra < a.rows ? a.in[ra * n + k] : 0.f;
            vb[q] = ra + 1 < a.rows ? a.in[(ra + 1) * n + k] : 0.f;
        }
    }

    if (MODE == kInverse) __syncthreads();  // the twiddles are in LDS
    if (MODE != kForward) {
        // Y = X / s in natural order (real row r at stg[r * n]), then
        // V[m] = (1/2) conj(pt[m]) (Y[m] - i Y[n-m]) per row and W = V_a + i V_b
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const int c = e >> L, m = e & (n - 1);
            const float rm = m == 0 ? a.r0 : a.r1;
            stg[((2 * c) << L) + m] = va[q] * rm;
            stg[((2 * c + 1) << L) + m] = vb[q] * rm;
        }
        __syncthreads();
        float2 W[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const int c = e >> L, m = e & (n - 1);
            const float2 cc = pt[m];  // e^{-i pi m / 2n}; its conjugate is applied
            const int ia = (2 * c) << L, ib = ia + n;
            const float ya = stg[ia + m], yna = m == 0 ? 0.f : stg[ia + n - m];
            const float yb = stg[ib + m], ynb = m == 0 ? 0.f : stg[ib + n - m];
            const float vax = 0.5f * (cc.x * ya - cc.y * yna), vay = 0.5f * (-cc.y * ya - cc.x * yna);
            const float vbx = 0.5f * (cc.x * yb - cc.y * ynb), vby = 0.5f * (-cc.y * yb - cc.x * ynb);
            W[q] = make_float2(vax - vby, vay + vbx);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q) buf[pad(tid + q * kThreads)] = W[q];
        __syncthreads();
        fft_dif<L, true>(buf, tw);
        constexpr float inv_n = 1.f / float(n);
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const int c = e >> L, k = e & (n - 1);
            const int j = (k & 1) ? n - 1 - (k >> 1) : (k >> 1);
            const float2 z = buf[pad((c << L) + bitrev(j, L))];
            va[q] = z.x * inv_n;
            vb[q] = z.y * inv_n;
        }
    }

    if (MODE == kInverse) {
        // rows back in place, coalesced
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads;
            const long long ra = g0 + 2 * (e >> L);
            const int k = e & (n - 1);
            if (ra < a.rows) a.out[ra * n + k] = va[q];
            if (ra + 1 < a.rows) a.out[(ra + 1) * n + k] = vb[q];
        }
        return;
    }
    // transposed store: element (row r, m) -> out[b][m][i], i = row index in
    // its image; staged as stg[m * (R + 1) + r] so each m leaves as a run of rows
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + q * kThreads;
        const int c = e >> L, m = e & (n - 1);
        stg[m * (R + 1) + 2 * c] = va[q];
        stg[m * (R + 1) + 2 * c + 1] = vb[q];
    }
    __syncthreads();
    constexpr int lr = 13 - L;  // log2 R
#pragma unroll
    for (int q = 0; q < 2 * kPer; ++q) {
        const int e = tid + q * kThreads;
        const int m = e >> lr, r = e & (R - 1);
        const long long gr = g0 + r;
        if (gr < a.rows) {
            const long long b = gr >> L, i = gr & (n - 1);
            a.out[(b << (2 * L)) + ((long long)m << L) + i] = stg[m * (R + 1) + r];
        }
    }
}

template <int MODE, int L>
static void launch_l(const Args &a, hipStream_t s) {
    constexpr long long R = 2 * (kElems >> L);  // real rows per work-group
    const unsigned grid = unsigned((a.rows + R - 1) / R);
    const size_t lds = size_t(((1 << L) >> 1) + (1 << L)) * sizeof(float2);
    hipLaunchKernelGGL((dct_rows_kernel<MODE, L>), dim3(grid), dim3(kThreads), lds, s, a);
}

template <int MODE>
static void launch(const Args &a, int L, hipStream_t s) {
    switch (L) {
        case 1: launch_l<MODE, 1>(a, s); break;
        case 2: launch_l<MODE, 2>(a, s); break;
        case 3: launch_l<MODE, 3>(a, s); break;
        case 4: launch_l<MODE, 4>(a, s); break;
        case 5: launch_l<MODE, 5>(a, s); break;
        case 6: launch_l<MODE, 6>(a, s); break;
        case 7: launch_l<MODE, 7>(a, s); break;
        case 8: launch_l<MODE, 8>(a, s); break;
        case 9: launch_l<MODE, 9>(a, s); break;
        default: launch_l<MODE, 10>(a, s); break;
    }
}

}  // namespace dctp
}  // namespace adrt_hip

using namespace adrt_hip;

extern "C" size_t adrt_hip_dct_table_len(size_t n) { return n / 2 + n; }

extern "C" int adrt_hip_dct_table(size_t n, float *host, size_t len) {
    if (n < 2 || n > size_t(dctp::kMaxN) || !is_pow2(n)) return ADRT_HIP_ERR_SHAPE;
    if (!host) return ADRT_HIP_ERR_NULL;
    if (len < 2 * adrt_hip_dct_table_len(n)) return ADRT_HIP_ERR_ARG;
    const double pi = 3.14159265358979323846;
    for (size_t t = 0; t < n / 2; ++t) {
        const double ang = 2.0 * pi * double(t) / double(n);
        host[2 * t] = float(std::cos(ang));
        host[2 * t + 1] = float(-std::sin(ang));
    }
    float *pt = host + n;
    for (size_t m = 0; m < n; ++m) {
        const double ang = pi * double(m) / (2.0 * double(n));
        pt[2 * m] = float(std::cos(ang));
        pt[2 * m + 1] = float(-std::sin(ang));
    }
    return ADRT_HIP_OK;
}

extern "C" int adrt_hip_dct_filter(const float *x, float *y, const float *table, const float *H,
                                   float *tmp, size_t batch, size_t n, int mode, void *stream) {
    if (n < 2 || n > size_t(dctp::kMaxN) || !is_pow2(n) || batch == 0) return ADRT_HIP_ERR_SHAPE;
    if (mode != 0 && mode != 1) return ADRT_HIP_ERR_ARG;
    if (!x || !y || !table || !tmp || (mode == 1 && !H)) return ADRT_HIP_ERR_NULL;
    if (x == y || x == tmp || y == tmp) return ADRT_HIP_ERR_ARG;
    size_t elems;
    if (!mul_ok(batch, n * n, &elems) || elems > (size_t(1) << 62)) return ADRT_HIP_ERR_TOO_BIG;
    dctp::Args a;
    a.tw = reinterpret_cast<const float2 *>(table);
    a.H = H;
    a.rows = (long long)(batch * n);
    const int L = num_iters(n);
    a.s0 = float(std::sqrt(1.0 / (4.0 * double(n))));
    a.s1 = float(std::sqrt(1.0 / (2.0 * double(n))));
    a.r0 = float(std::sqrt(4.0 * double(n)));
    a.r1 = float(std::sqrt(2.0 * double(n)));
    const hipStream_t s = static_cast<hipStream_t>(stream);
    a.in = x;
    a.out = tmp;
    dctp::launch<dctp::kForward>(a, L, s);
    a.in = tmp;
    a.out = y;
    if (mode == 0) {
        dctp::launch<dctp::kForward>(a, L, s);
        return launch_status();
    }
    dctp::launch<dctp::kFiltered>(a, L, s);
    a.in = y;
    a.out = y;
    dctp::launch<dctp::kInverse>(a, L, s);
    return launch_status();
}
