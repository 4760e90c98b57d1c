/*
 * CPU oracle for the ADRT transform core -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is included twice by adrt_oracle.c, once with REAL=float and
 * once with REAL=double.  Every routine restates one routine of the
 * reference's native core (src/adrt/adrt_cdefs_*.hpp) loop for loop, so that
 * each output element is formed by exactly the same sequence of IEEE
 * additions as in the reference.  Only tests/, the smoke check and the CPU
 * baseline leg of bench.py may call it; the product path never does.
 *
 * Buffers follow the reference's virtual 5-d layouts:
 *   forward: (B, 4, R, S, C)   R row-blocks, S angles, C = 2N-1 offsets
 *   bdrt:    (B, 4, R, S, C)   R angle-pairs, S sections, C rows
 *   iadrt:   (B, 4, L, C, R)   L processed columns, C columns, R rows
 * all C-contiguous, last index fastest.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* One forward butterfly stage; follows adrt_cdefs_adrt.hpp:55-96. */
static void FN(fwd_stage_)(const REAL *in, REAL *out, size_t B, size_t R_in,
                           size_t S_in, size_t C) {
    const size_t R = R_in / 2, S = S_in * 2;
    const long long outer = (long long)(B * 4 * R);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / R, r = (size_t)idx % R;
        const REAL *lo = in + ((bq * R_in + 2 * r) * S_in) * C;      /* row 2r   */
        const REAL *hi = in + ((bq * R_in + 2 * r + 1) * S_in) * C;  /* row 2r+1 */
        REAL *dst = out + ((bq * R + r) * S) * C;
        for (size_t a = 0; a < S; ++a) {
            const size_t src_a = a / 2;
            const size_t shift = (a + 1) / 2; /* ceil(a/2) */
            const REAL *x = lo + src_a * C, *y = hi + src_a * C;
            REAL *o = dst + a * C;
            for (size_t c = 0; c < shift && c < C; ++c) o[c] = x[c];
            for (size_t c = shift; c < C; ++c) o[c] = x[c] + y[c - shift];
        }
    }
}

/* Forward ADRT; follows adrt_basic in adrt_cdefs_adrt.hpp:100-212. */
int FN(oracle_adrt_)(const REAL *data, size_t B, size_t N, REAL *out) {
    const size_t C = 2 * N - 1;
    const size_t total = B * 4 * N * C;
    REAL *buf_a = (REAL *)malloc(total * sizeof(REAL));
    REAL *buf_b = (REAL *)malloc(total * sizeof(REAL));
    if (!buf_a || !buf_b) { free(buf_a); free(buf_b); return 1; }
    /* quadrant loads, adrt_cdefs_adrt.hpp:123-175; zero pad :176-186 */
    const long long rows = (long long)(B * N);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < rows; ++idx) {
        const size_t b = (size_t)idx / N, row = (size_t)idx % N;
        const REAL *img = data + b * N * N;
        for (size_t col = 0; col < N; ++col) {
            buf_a[((b * 4 + 0) * N + row) * C + (N - col - 1)] = img[row * N + col];
            buf_a[((b * 4 + 1) * N + (N - row - 1)) * C + (N - col - 1)] = img[col * N + (N - row - 1)];
            buf_a[((b * 4 + 2) * N + (N - row - 1)) * C + (N - col - 1)] = img[(N - col - 1) * N + (N - row - 1)];
            buf_a[((b * 4 + 3) * N + row) * C + (N - col - 1)] = img[(N - row - 1) * N + col];
        }
        for (size_t q = 0; q < 4; ++q)
            for (size_t col = N; col < C; ++col) buf_a[((b * 4 + q) * N + row) * C + col] = (REAL)0;
    }
    size_t R = N, S = 1;
    while (R > 1) {
        FN(fwd_stage_)(buf_a, buf_b, B, R, S, C);
        R /= 2; S *= 2;
        REAL *t = buf_a; buf_a = buf_b; buf_b = t;
    }
    /* transposed copy-out, adrt_cdefs_adrt.hpp:194-210 */
    const long long planes = (long long)(B * 4);
#pragma omp parallel for schedule(static)
    for (long long bq = 0; bq < planes; ++bq)
        for (size_t d = 0; d < C; ++d)
            for (size_t a = 0; a < N; ++a)
                out[((size_t)bq * C + d) * N + a] = buf_a[((size_t)bq * N + a) * C + d];
    free(buf_a); free(buf_b);
    return 0;
}

/* One forward stage on the output layout; follows adrt_step in
 * adrt_cdefs_adrt.hpp:214-258. Layout (B, 4, 2N-1, N). */
int FN(oracle_adrt_step_)(const REAL *data, size_t B, size_t N, int step, REAL *out) {
    const size_t C = 2 * N - 1;
    const size_t e = (size_t)1 << step, e2 = e << 1;
    const size_t nblk = (N + e2 - 1) / e2;
    const long long outer = (long long)(B * 4 * C);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / C, row = (size_t)idx % C;
        const REAL *plane = data + bq * C * N;
        REAL *dst = out + bq * C * N + row * N;
        for (size_t cb = 0; cb < nblk; ++cb) {
            const size_t col0 = cb * e2;
            const size_t ncols = (e2 < N - col0) ? e2 : N - col0;
            const size_t split = (2 * row + 1 < ncols) ? 2 * row + 1 : ncols;
            for (size_t ci = 0; ci < split; ++ci) {
                const REAL x = plane[row * N + 2 * cb * e + ci / 2];
                const REAL y = plane[(row - (ci + 1) / 2) * N + (2 * cb + 1) * e + ci / 2];
                dst[col0 + ci] = x + y;
            }
            for (size_t ci = split; ci < ncols; ++ci)
                dst[col0 + ci] = plane[row * N + 2 * cb * e + ci / 2];
        }
    }
    return 0;
}

/* One back-projection stage; follows bdrt_core in adrt_cdefs_bdrt.hpp:55-116. */
static void FN(bdrt_stage_)(const REAL *in, REAL *out, size_t B, size_t R_in,
                            size_t S_in, size_t C) {
    const size_t R = R_in / 2, S = S_in * 2;
    const long long outer = (long long)(B * 4 * R);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / R, si = (size_t)idx % R;
        for (size_t sec = 0; sec < S_in; ++sec) {
            const REAL *x = in + ((bq * R_in + 2 * si) * S_in + sec) * C;
            const REAL *y = in + ((bq * R_in + 2 * si + 1) * S_in + sec) * C;
            REAL *left = out + ((bq * R + si) * S + 2 * sec) * C;
            REAL *right = out + ((bq * R + si) * S + 2 * sec + 1) * C;
            for (size_t row = 0; row < si; ++row) left[row] = x[row] + y[row];
            for (size_t row = si; row < C - 1; ++row) {
                const REAL la = x[row], lb = y[row], lbb = y[row + 1];
                left[row] = la + lb;
                right[row - si] = la + lbb;
            }
            left[C - 1] = x[C - 1] + y[C - 1];
            right[C - 1 - si] = x[C - 1];
            for (size_t z = C - si; z < C; ++z) right[z] = (REAL)0;
        }
    }
}

/* Back-projection; follows bdrt_basic in adrt_cdefs_bdrt.hpp:120-187. */
int FN(oracle_bdrt_)(const REAL *data, size_t B, size_t N, REAL *out) {
    const size_t C = 2 * N - 1;
    const size_t total = B * 4 * N * C;
    REAL *buf_a = (REAL *)malloc(total * sizeof(REAL));
    REAL *buf_b = (REAL *)malloc(total * sizeof(REAL));
    if (!buf_a || !buf_b) { free(buf_a); free(buf_b); return 1; }
    const long long planes = (long long)(B * 4);
#pragma omp parallel for schedule(static)
    for (long long bq = 0; bq < planes; ++bq)
        for (size_t row = 0; row < C; ++row)
            for (size_t col = 0; col < N; ++col)
                buf_a[((size_t)bq * N + col) * C + row] = data[((size_t)bq * C + row) * N + col];
    size_t R = N, S = 1;
    while (R > 1) {
        FN(bdrt_stage_)(buf_a, buf_b, B, R, S, C);
        R /= 2; S *= 2;
        REAL *t = buf_a; buf_a = buf_b; buf_b = t;
    }
#pragma omp parallel for schedule(static)
    for (long long bq = 0; bq < planes; ++bq)
        for (size_t row = 0; row < C; ++row)
            for (size_t col = 0; col < N; ++col)
                out[((size_t)bq * C + row) * N + col] = buf_a[((size_t)bq * N + col) * C + row];
    free(buf_a); free(buf_b);
    return 0;
}

/* One back-projection stage on the output layout; follows bdrt_step in
 * adrt_cdefs_bdrt.hpp:189-244. Out-of-range operands are literal zeros. */
int FN(oracle_bdrt_step_)(const REAL *data, size_t B, size_t N, int step, REAL *out) {
    const size_t C = 2 * N - 1;
    int L = 0;
    while (((size_t)1 << L) < N) ++L;
    const size_t e = (size_t)1 << (L - step - 1);
    const size_t nblk = (N + e - 1) / e;
    const long long outer = (long long)(B * 4 * C);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / C, row = (size_t)idx % C;
        const REAL *plane = data + bq * C * N;
        REAL *dst = out + bq * C * N + row * N;
        for (size_t cb = 0; cb < nblk; ++cb) {
            const size_t col0 = cb * e;
            const size_t ncols = (e < N - col0) ? e : N - col0;
            const size_t base = e * (cb / 2);
            if (cb % 2 == 0) {
                for (size_t ci = 0; ci < ncols; ++ci) {
                    const size_t bc = 2 * (ci + base);
                    dst[col0 + ci] = plane[row * N + bc] + plane[row * N + bc + 1];
                }
            } else {
                for (size_t ci = 0; ci < ncols; ++ci) {
                    const size_t bc = 2 * (ci + base);
                    const size_t br = row + ci;
                    REAL x = 0, y = 0;
                    if (br < C) x = plane[br * N + bc];
                    if (br + 1 < C) y = plane[(br + 1) * N + bc + 1];
                    dst[col0 + ci] = x + y;
                }
            }
        }
    }
    return 0;
}

/* One inverse stage; follows iadrt_core in adrt_cdefs_iadrt.hpp:52-105.
 * in (B,4,Li,Ci,R) -> out (B,4,2Li,Ci/2,R); the row loop is serial. */
static void FN(iadrt_stage_)(const REAL *in, REAL *out, size_t B, size_t Li,
                             size_t Ci, size_t R) {
    const size_t Lo = 2 * Li, Co = Ci / 2;
    const long long outer = (long long)(B * 4 * Lo * Co);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / (Lo * Co);
        const size_t l = ((size_t)idx / Co) % Lo, col = (size_t)idx % Co;
        const size_t pl = l / 2;
        const REAL *src = in + ((bq * Li + pl) * Ci) * R;
        REAL *dst = out + ((bq * Lo + l) * Co + col) * R;
        for (size_t rr = 0; rr < R; ++rr) {
            const size_t row = R - rr - 1;
            REAL val = 0;
            if (l % 2 == 0) {
                val += src[(2 * col) * R + row];
                if (row + 1 < R && 2 * col + 1 < Ci) val -= src[(2 * col + 1) * R + row + 1];
            } else {
                if (row + 1 + col < R) {
                    if (2 * col + 1 < Ci) val += src[(2 * col + 1) * R + row + 1 + col];
                    val -= src[(2 * col) * R + row + 1 + col];
                }
            }
            if (row + 1 < R) val += dst[row + 1];
            dst[row] = val;
        }
    }
}

/* Exact inverse; follows iadrt_basic in adrt_cdefs_iadrt.hpp:109-176. */
int FN(oracle_iadrt_)(const REAL *data, size_t B, size_t N, REAL *out) {
    const size_t C = 2 * N - 1;
    const size_t total = B * 4 * N * C;
    REAL *buf_a = (REAL *)malloc(total * sizeof(REAL));
    REAL *buf_b = (REAL *)malloc(total * sizeof(REAL));
    if (!buf_a || !buf_b) { free(buf_a); free(buf_b); return 1; }
    const long long planes = (long long)(B * 4);
#pragma omp parallel for schedule(static)
    for (long long bq = 0; bq < planes; ++bq)
        for (size_t c = 0; c < N; ++c)
            for (size_t r = 0; r < C; ++r)
                buf_a[((size_t)bq * N + c) * C + r] = data[((size_t)bq * C + r) * N + c];
    size_t Li = 1, Ci = N;
    while (Ci > 1) {
        FN(iadrt_stage_)(buf_a, buf_b, B, Li, Ci, C);
        Li *= 2; Ci /= 2;
        REAL *t = buf_a; buf_a = buf_b; buf_b = t;
    }
#pragma omp parallel for schedule(static)
    for (long long bq = 0; bq < planes; ++bq)
        for (size_t r = 0; r < C; ++r)
            for (size_t c = 0; c < N; ++c)
                out[((size_t)bq * C + r) * N + c] = buf_a[((size_t)bq * N + c) * C + r];
    free(buf_a); free(buf_b);
    return 0;
}

/* Multigrid restriction; follows fmg_restriction in adrt_cdefs_fmg.hpp:53-73.
 * (B,4,2N-1,N) -> (B,4,N-1,N/2). */
int FN(oracle_fmg_restriction_)(const REAL *data, size_t B, size_t N, REAL *out) {
    const size_t Ri = 2 * N - 1, Ro = N - 1, Co = N / 2;
    const long long outer = (long long)(B * 4 * Ro);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t bq = (size_t)idx / Ro, r = (size_t)idx % Ro;
        for (size_t c = 0; c < Co; ++c) {
            const REAL va = data[(bq * Ri + 2 * r) * N + 2 * c];
            const REAL vb = data[(bq * Ri + 2 * r + 1) * N + 2 * c];
            out[(bq * Ro + r) * Co + c] = (va + vb) / (REAL)4;
        }
    }
    return 0;
}

/* Multigrid prolongation; follows fmg_prolongation in adrt_cdefs_fmg.hpp:75-95. */
int FN(oracle_fmg_prolongation_)(const REAL *data, size_t B, size_t H, size_t W, REAL *out) {
    const long long outer = (long long)(B * H);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t b = (size_t)idx / H, r = (size_t)idx % H;
        for (size_t c = 0; c < W; ++c) {
            const REAL v = data[(b * H + r) * W + c];
            REAL *o = out + (b * 2 * H + 2 * r) * 2 * W;
            o[2 * c] = v; o[2 * c + 1] = v;
            o[2 * W + 2 * c] = v; o[2 * W + 2 * c + 1] = v;
        }
    }
    return 0;
}

/* Multigrid high-pass; follows fmg_highpass in adrt_cdefs_fmg.hpp:97-171.
 * Products are rounded before the sums; sums group as in the reference. */
int FN(oracle_fmg_highpass_)(const REAL *data, size_t B, size_t H, size_t W, REAL *out) {
    const REAL ka = (REAL)-0.0625, kb = (REAL)-0.125, kc = (REAL)0.75;
    const long long outer = (long long)(B * H);
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < outer; ++idx) {
        const size_t b = (size_t)idx / H, r = (size_t)idx % H;
        const REAL *img = data + b * H * W;
        const size_t pr = (r == 0) ? 1 : r - 1;
        const size_t nr = (r == H - 1) ? r - 1 : r + 1;
        const REAL *P = img + pr * W, *M = img + r * W, *Q = img + nr * W;
        REAL *o = out + (b * H + r) * W;
        {
            const REAL v11 = ka * P[1], v12 = kb * P[0];
            const REAL v21 = kb * M[1], v22 = kc * M[0];
            const REAL v31 = ka * Q[1], v32 = kb * Q[0];
            o[0] = (v11 + v21 + v31) + (v12 + v22 + v32) + (v11 + v21 + v31);
        }
        for (size_t c = 1; c + 1 < W; ++c) {
            const REAL v11 = ka * P[c - 1], v12 = kb * P[c], v13 = ka * P[c + 1];
            const REAL v21 = kb * M[c - 1], v22 = kc * M[c], v23 = kb * M[c + 1];
            const REAL v31 = ka * Q[c - 1], v32 = kb * Q[c], v33 = ka * Q[c + 1];
            o[c] = (v11 + v21 + v31) + (v12 + v22 + v32) + (v13 + v23 + v33);
        }
        {
            const size_t c = W - 1;
            const REAL v11 = ka * P[c - 1], v12 = kb * P[c];
            const REAL v21 = kb * M[c - 1], v22 = kc * M[c];
            const REAL v31 = ka * Q[c - 1], v32 = kb * Q[c];
            o[c] = (v11 + v21 + v31) + (v12 + v22 + v32) + (v11 + v21 + v31);
        }
    }
    return 0;
}

#undef FN
#undef CAT
#undef CAT2
