// C ABI entry points: validation (same rules as the reference's
// adrt_cdefs_common.cpp:176-330), then asynchronous launches.
#include <cstdio>
#include <cstring>

#include "common.hpp"

namespace adrt_hip {

static thread_local char g_last_error[256] = "";

void set_last_error(hipError_t err) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", hipGetErrorName(err),
                  hipGetErrorString(err));
}

int launch_status() {
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) {
        set_last_error(err);
        return ADRT_HIP_ERR_DEVICE;
    }
    return ADRT_HIP_OK;
}

static bool dtype_ok(int dtype) { return dtype == ADRT_HIP_F32 || dtype == ADRT_HIP_F64; }

static size_t dtype_size(int dtype) { return dtype == ADRT_HIP_F64 ? 8 : 4; }

// ADRT-output shape (B, 4, 2n-1, n) check (adrt_cdefs_common.cpp:185-192).
static bool valid_output(size_t B, size_t n) {
    return B > 0 && n > 0 && n <= kMaxSize && is_pow2(n);
}

// Element count of a (B, 4, 2n-1, n) array, false on overflow.
static bool output_elems(size_t B, size_t n, size_t *out) {
    const size_t shape[4] = {B, 4, 2 * n - 1, n};
    return product_ok(shape, 4, out);
}

}  // namespace adrt_hip

using namespace adrt_hip;

extern "C" {

const char *adrt_hip_status_string(int status) {
    switch (status) {
        case ADRT_HIP_OK: return "ok";
        case ADRT_HIP_ERR_SHAPE: return "invalid shape";
        case ADRT_HIP_ERR_STEP: return "step out of range";
        case ADRT_HIP_ERR_DTYPE: return "unsupported dtype";
        case ADRT_HIP_ERR_TOO_BIG: return "array is too big";
        case ADRT_HIP_ERR_NULL: return "null pointer";
        case ADRT_HIP_ERR_WORKSPACE: return "workspace too small";
        case ADRT_HIP_ERR_DEVICE: return "HIP device error";
        case ADRT_HIP_ERR_ARG: return "invalid argument";
        case ADRT_HIP_ERR_SCRATCH: return "needs more than the no-allocation scratch buffer";
        default: return "unknown status";
    }
}

const char *adrt_hip_last_error(void) { return g_last_error; }

int adrt_hip_abi_version(void) { return 101; }

int adrt_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int adrt_hip_num_iters(size_t n) { return num_iters(n); }

int adrt_hip_valid_adrt_shape(const size_t shape[3]) {
    if (!shape) return 0;
    return shape[0] > 0 && shape[1] > 0 && shape[2] > 0 && shape[1] == shape[2] &&
           shape[2] <= kMaxSize && is_pow2(shape[2]);
}

int adrt_hip_valid_output_shape(const size_t shape[4]) {
    if (!shape) return 0;
    for (int i = 0; i < 4; ++i)
        if (shape[i] == 0) return 0;
    return shape[1] == 4 && shape[3] <= kMaxSize && shape[2] == shape[3] * 2 - 1 &&
           is_pow2(shape[3]);
}

int adrt_hip_valid_step(const size_t shape[4], int step) {
    if (!shape) return 0;
    return step >= 0 && step < num_iters(shape[3]);
}

int adrt_hip_valid_restriction_shape(const size_t shape[4]) {
    if (!shape) return 0;
    for (int i = 0; i < 4; ++i)
        if (shape[i] == 0) return 0;
    return shape[1] == 4 && shape[3] <= kMaxSize && shape[2] == shape[3] * 2 - 1 &&
           shape[3] >= 2 && shape[3] % 2 == 0;
}

int adrt_hip_valid_prolongation_shape(const size_t shape[3]) {
    if (!shape) return 0;
    const size_t lim = SIZE_MAX / 2;
    return shape[0] > 0 && shape[1] > 0 && shape[2] > 0 && shape[1] <= lim && shape[2] <= lim;
}

int adrt_hip_valid_highpass_shape(const size_t shape[3]) {
    if (!shape) return 0;
    return shape[0] > 0 && shape[1] >= 2 && shape[2] >= 2;
}

size_t adrt_hip_workspace_bytes(int op, int dtype, size_t batch, size_t n) {
    if (!dtype_ok(dtype) || !valid_output(batch, n)) return 0;
    switch (op) {
        case ADRT_HIP_OP_ADRT: return adrt_workspace(dtype, batch, n);
        case ADRT_HIP_OP_BDRT: return bdrt_workspace(dtype, batch, n);
        case ADRT_HIP_OP_IADRT: return iadrt_workspace(dtype, batch, n);
        default: return 0;
    }
}

int adrt_hip_adrt(const void *in, void *out, void *workspace, size_t workspace_bytes, int dtype,
                  size_t batch, size_t n, void *stream) {
    const size_t shape[3] = {batch, n, n};
    if (!adrt_hip_valid_adrt_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    size_t elems;
    if (!output_elems(batch, n, &elems) || !mul_ok(elems, 8, &elems)) return ADRT_HIP_ERR_TOO_BIG;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    const size_t need = adrt_workspace(dtype, batch, n);
    if (workspace_bytes < need || (need && !workspace)) return ADRT_HIP_ERR_WORKSPACE;
    return launch_adrt(in, out, workspace, workspace_bytes, dtype, batch, n,
                       static_cast<hipStream_t>(stream));
}

int adrt_hip_adrt_step(const void *in, void *out, int dtype, size_t batch, size_t n, int step,
                       void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_output_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    if (!adrt_hip_valid_step(shape, step)) return ADRT_HIP_ERR_STEP;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_adrt_step(in, out, dtype, batch, n, step, static_cast<hipStream_t>(stream));
}

static int bdrt_checked(const void *in, void *out, void *workspace, size_t workspace_bytes,
                        int dtype, size_t batch, size_t n, size_t rows, void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_output_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    size_t elems;
    if (!output_elems(batch, n, &elems) || !mul_ok(elems, 8, &elems)) return ADRT_HIP_ERR_TOO_BIG;
    if (rows == 0 || rows > 2 * n - 1) return ADRT_HIP_ERR_ARG;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    const size_t need = bdrt_workspace(dtype, batch, n);
    if (workspace_bytes < need || (need && !workspace)) return ADRT_HIP_ERR_WORKSPACE;
    return launch_bdrt(in, out, workspace, workspace_bytes, dtype, batch, n, int(rows),
                       static_cast<hipStream_t>(stream));
}

int adrt_hip_bdrt(const void *in, void *out, void *workspace, size_t workspace_bytes, int dtype,
                  size_t batch, size_t n, void *stream) {
    return bdrt_checked(in, out, workspace, workspace_bytes, dtype, batch, n, 2 * n - 1, stream);
}

int adrt_hip_bdrt_rows(const void *in, void *out, void *workspace, size_t workspace_bytes,
                       int dtype, size_t batch, size_t n, size_t rows, void *stream) {
    return bdrt_checked(in, out, workspace, workspace_bytes, dtype, batch, n, rows, stream);
}

int adrt_hip_bdrt_step(const void *in, void *out, int dtype, size_t batch, size_t n, int step,
                       void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_output_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    if (!adrt_hip_valid_step(shape, step)) return ADRT_HIP_ERR_STEP;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_bdrt_step(in, out, dtype, batch, n, step, static_cast<hipStream_t>(stream));
}

int adrt_hip_iadrt(const void *in, void *out, void *workspace, size_t workspace_bytes, int dtype,
                   size_t batch, size_t n, void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_output_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    size_t elems;
    if (!output_elems(batch, n, &elems) || !mul_ok(elems, 8, &elems)) return ADRT_HIP_ERR_TOO_BIG;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    const size_t need = iadrt_workspace(dtype, batch, n);
    if (workspace_bytes < need || (need && !workspace)) return ADRT_HIP_ERR_WORKSPACE;
    return launch_iadrt(in, out, workspace, workspace_bytes, dtype, batch, n,
                        static_cast<hipStream_t>(stream));
}

int adrt_hip_fmg_restriction(const void *in, void *out, int dtype, size_t batch, size_t n,
                             void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_restriction_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_fmg_restriction(in, out, dtype, batch, n, static_cast<hipStream_t>(stream));
}

int adrt_hip_fmg_prolongation(const void *in, void *out, int dtype, size_t batch, size_t h,
                              size_t w, void *stream) {
    const size_t shape[3] = {batch, h, w};
    if (!adrt_hip_valid_prolongation_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    const size_t oshape[3] = {batch, 2 * h, 2 * w};
    size_t elems;
    if (!product_ok(oshape, 3, &elems) || !mul_ok(elems, 8, &elems)) return ADRT_HIP_ERR_TOO_BIG;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_fmg_prolongation(in, out, dtype, batch, h, w, static_cast<hipStream_t>(stream));
}

int adrt_hip_fmg_highpass(const void *in, void *out, int dtype, size_t batch, size_t h, size_t w,
                          void *stream) {
    const size_t shape[3] = {batch, h, w};
    if (!adrt_hip_valid_highpass_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_fmg_highpass(in, out, dtype, batch, h, w, static_cast<hipStream_t>(stream));
}

int adrt_hip_adrt_init(const void *in, void *out, size_t elem_bytes, size_t batch, size_t n,
                       void *stream) {
    const size_t shape[3] = {batch, n, n};
    if (!adrt_hip_valid_adrt_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_adrt_init(in, out, elem_bytes, batch, n, static_cast<hipStream_t>(stream));
}

int adrt_hip_truncate(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                      void *stream) {
    if (outer == 0 || n == 0) return ADRT_HIP_ERR_SHAPE;
    if (elem_bytes != 1 && elem_bytes != 2 && elem_bytes != 4 && elem_bytes != 8)
        return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_truncate(in, out, elem_bytes, outer, n, static_cast<hipStream_t>(stream));
}

int adrt_hip_truncate_combine(const void *in, void *out, int dtype, size_t batch, size_t n,
                              double divisor, int mode, void *stream) {
    if (batch == 0 || n == 0) return ADRT_HIP_ERR_SHAPE;
    if (mode != 0 && mode != 1) return ADRT_HIP_ERR_ARG;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return launch_truncate_combine(in, out, dtype, batch, n, divisor, mode,
                                   static_cast<hipStream_t>(stream));
}

size_t adrt_hip_stage_chain_workspace(int dtype, size_t batch, size_t n) {
    if (!dtype_ok(dtype) || !valid_output(batch, n)) return 0;
    return batch * 4 * (2 * n - 1) * n * (dtype == ADRT_HIP_F64 ? 8 : 4);
}

int adrt_hip_stage_chain(const void *in, void *out, void *workspace, size_t workspace_bytes,
                         int dtype, size_t batch, size_t n, int first, int last, void *stream) {
    const size_t shape[4] = {batch, 4, 2 * n - 1, n};
    if (n == 0 || !adrt_hip_valid_output_shape(shape)) return ADRT_HIP_ERR_SHAPE;
    const int L = num_iters(n);
    if (first < 0 || last > L || first >= last) return ADRT_HIP_ERR_STEP;
    if (!dtype_ok(dtype)) return ADRT_HIP_ERR_DTYPE;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    const int steps = last - first;
    const size_t need = steps > 1 ? adrt_hip_stage_chain_workspace(dtype, batch, n) : 0;
    if (workspace_bytes < need || (need && !workspace)) return ADRT_HIP_ERR_WORKSPACE;
    // ping-pong so that the last step writes `out`
    void *bufs[2] = {steps % 2 == 1 ? out : workspace, steps % 2 == 1 ? workspace : out};
    const void *src = in;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int k = 0; k < steps; ++k) {
        const int st = launch_adrt_step(src, bufs[k & 1], dtype, batch, n, first + k, s);
        if (st != ADRT_HIP_OK) return st;
        src = bufs[k & 1];
    }
    return ADRT_HIP_OK;
}

static int stitch_args(size_t elem, size_t outer, size_t n, int flag, const void *in,
                       const void *out) {
    if (outer == 0 || n == 0 || (flag != 0 && flag != 1) || (flag == 1 && n < 2))
        return ADRT_HIP_ERR_SHAPE;
    if (elem != 1 && elem != 2 && elem != 4 && elem != 8) return ADRT_HIP_ERR_DTYPE;
    size_t total;
    const size_t dims[4] = {outer, 3 * n, 4, n};
    if (n > (SIZE_MAX >> 4) || !product_ok(dims, 4, &total)) return ADRT_HIP_ERR_TOO_BIG;
    if (!in || !out) return ADRT_HIP_ERR_NULL;
    return ADRT_HIP_OK;
}

int adrt_hip_stitch(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                    int remove_repeated, void *stream) {
    const int st = stitch_args(elem_bytes, outer, n, remove_repeated, in, out);
    if (st != ADRT_HIP_OK) return st;
    return launch_stitch(in, out, elem_bytes, outer, n, remove_repeated ? n - 1 : n, false,
                         static_cast<hipStream_t>(stream));
}

int adrt_hip_unstitch(const void *in, void *out, size_t elem_bytes, size_t outer, size_t n,
                      int removed_repeated, void *stream) {
    const int st = stitch_args(elem_bytes, outer, n, removed_repeated, in, out);
    if (st != ADRT_HIP_OK) return st;
    return launch_stitch(in, out, elem_bytes, outer, n, removed_repeated ? n - 1 : n, true,
                         static_cast<hipStream_t>(stream));
}

int adrt_hip_set_config(int op, size_t n, int variant, size_t group_bytes) {
    return set_fused_config(op, n, variant, group_bytes);
}

int adrt_hip_get_config(int op, size_t n, int *variant, size_t *group_bytes) {
    return get_fused_config(op, n, variant, group_bytes);
}

}  // extern "C"
