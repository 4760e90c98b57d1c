// C-ABI shim over the reference's own native core -- TEST INFRASTRUCTURE ONLY.
//
// This file holds no algorithm: it includes the reference's headers from the
// reference checkout (src/adrt/adrt_cdefs_*.hpp, compiled where they lie by
// oracle/Makefile) and exposes adrt_basic / adrt_step / bdrt_basic /
// bdrt_step / iadrt_basic / fmg_* / interp_adrtcart (float index math, as
// the reference's binding instantiates it) through plain C entry points so the oracle
// can be checked against the reference itself and bench.py can time the
// reference's OpenMP CPU path.  The built library lands in oracle/_ref/.
#include <array>
#include <cstddef>
#include <cstdlib>
#include <new>
#include "adrt_cdefs_common.hpp"
#include "adrt_cdefs_adrt.hpp"
#include "adrt_cdefs_bdrt.hpp"
#include "adrt_cdefs_iadrt.hpp"
#include "adrt_cdefs_fmg.hpp"
#include "adrt_cdefs_interp_adrtcart.hpp"

namespace {

template <typename T>
T *scratch(std::size_t n) {
    return static_cast<T *>(std::malloc(n * sizeof(T)));
}

template <typename T>
int run_adrt(const T *in, std::size_t B, std::size_t N, T *out) {
    const std::array<std::size_t, 3> shape{B, N, N};
    T *tmp = scratch<T>(B * 4 * N * (2 * N - 1));
    if (!tmp) return 1;
    adrt::adrt_basic<T>(in, shape, tmp, out);
    std::free(tmp);
    return 0;
}

template <typename T>
int run_bdrt(const T *in, std::size_t B, std::size_t N, T *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    T *tmp = scratch<T>(B * 4 * N * (2 * N - 1));
    if (!tmp) return 1;
    adrt::bdrt_basic<T>(in, shape, tmp, out);
    std::free(tmp);
    return 0;
}

template <typename T>
int run_iadrt(const T *in, std::size_t B, std::size_t N, T *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    T *tmp = scratch<T>(B * 4 * N * (2 * N - 1));
    if (!tmp) return 1;
    adrt::iadrt_basic<T>(in, shape, tmp, out);
    std::free(tmp);
    return 0;
}

}  // namespace

extern "C" {

int ref_openmp_enabled(void) { return adrt::_const::openmp_enabled ? 1 : 0; }

int ref_adrt_f32(const float *in, std::size_t B, std::size_t N, float *out) { return run_adrt(in, B, N, out); }
int ref_adrt_f64(const double *in, std::size_t B, std::size_t N, double *out) { return run_adrt(in, B, N, out); }
int ref_bdrt_f32(const float *in, std::size_t B, std::size_t N, float *out) { return run_bdrt(in, B, N, out); }
int ref_bdrt_f64(const double *in, std::size_t B, std::size_t N, double *out) { return run_bdrt(in, B, N, out); }
int ref_iadrt_f32(const float *in, std::size_t B, std::size_t N, float *out) { return run_iadrt(in, B, N, out); }
int ref_iadrt_f64(const double *in, std::size_t B, std::size_t N, double *out) { return run_iadrt(in, B, N, out); }

int ref_adrt_step_f32(const float *in, std::size_t B, std::size_t N, int step, float *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::adrt_step<float>(in, shape, out, step);
    return 0;
}
int ref_adrt_step_f64(const double *in, std::size_t B, std::size_t N, int step, double *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::adrt_step<double>(in, shape, out, step);
    return 0;
}
int ref_bdrt_step_f32(const float *in, std::size_t B, std::size_t N, int step, float *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::bdrt_step<float>(in, shape, out, step);
    return 0;
}
int ref_bdrt_step_f64(const double *in, std::size_t B, std::size_t N, int step, double *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::bdrt_step<double>(in, shape, out, step);
    return 0;
}

int ref_fmg_restriction_f32(const float *in, std::size_t B, std::size_t N, float *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::fmg_restriction<float>(in, shape, out);
    return 0;
}
int ref_fmg_restriction_f64(const double *in, std::size_t B, std::size_t N, double *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::fmg_restriction<double>(in, shape, out);
    return 0;
}
int ref_fmg_prolongation_f32(const float *in, std::size_t B, std::size_t H, std::size_t W, float *out) {
    const std::array<std::size_t, 3> shape{B, H, W};
    adrt::fmg_prolongation<float>(in, shape, out);
    return 0;
}
int ref_fmg_prolongation_f64(const double *in, std::size_t B, std::size_t H, std::size_t W, double *out) {
    const std::array<std::size_t, 3> shape{B, H, W};
    adrt::fmg_prolongation<double>(in, shape, out);
    return 0;
}
int ref_fmg_highpass_f32(const float *in, std::size_t B, std::size_t H, std::size_t W, float *out) {
    const std::array<std::size_t, 3> shape{B, H, W};
    adrt::fmg_highpass<float>(in, shape, out);
    return 0;
}
int ref_fmg_highpass_f64(const double *in, std::size_t B, std::size_t H, std::size_t W, double *out) {
    const std::array<std::size_t, 3> shape{B, H, W};
    adrt::fmg_highpass<double>(in, shape, out);
    return 0;
}

int ref_interp_f32(const float *in, std::size_t B, std::size_t N, float *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::interp_adrtcart<float, float>(in, shape, out);
    return 0;
}
int ref_interp_f64(const double *in, std::size_t B, std::size_t N, double *out) {
    const std::array<std::size_t, 4> shape{B, 4, 2 * N - 1, N};
    adrt::interp_adrtcart<double, float>(in, shape, out);
    return 0;
}

}  // extern "C"
