// Host-side prep helpers of the stage-2 library: the row-parallel pieces of
// prepare_stage2_inputs (crates/pystamps-core/src/native_stage2.rs:449-467,
// 594-606) over the caller's arrays, on host threads.
#include "pystamps_hip.h"
#include "ps2_host_util.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace ps2 {
namespace {

// |v| as glibc hypotf: the f64 sum of exact f64 squares, one f64 square root,
// rounded to f32 (Complex32::norm, num-complex -> f32::hypot)
inline float hypotf_glibc(float x, float y) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    if (std::isinf(ax) || std::isinf(ay)) return INFINITY;
    if (std::isnan(ax) || std::isnan(ay)) return ax * ay;
    if (ax == 0.0f) return ay;
    if (ay == 0.0f) return ax;
    const double dx = (double)x, dy = (double)y;
    return (float)std::sqrt(dx * dx + dy * dy);
}

int thread_count(int32_t n_threads, int64_t rows) {
    int n = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
    if (n <= 0) n = 1;
    n = std::min<int64_t>(n, std::max<int64_t>(1, rows / 4096));
    return std::max(1, n);
}

template <typename F>
void parallel_rows(int64_t rows, int nthreads, F fn) {
    if (nthreads <= 1) {
        fn(0, rows);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nthreads);
    for (int t = 0; t < nthreads; ++t) {
        const int64_t a = rows * t / nthreads, b = rows * (t + 1) / nthreads;
        th.emplace_back([=]() { fn(a, b); });
    }
    for (auto &x : th) x.join();
}

int check_cols(const int64_t *cols, int64_t n_sel, int64_t n_cols) {
    for (int64_t i = 0; i < n_sel; ++i)
        if (cols[i] < 0 || cols[i] >= n_cols) {
            set_error("column index " + std::to_string(cols[i]) + " outside 0.." + std::to_string(n_cols - 1));
            return PS2_E_ARG;
        }
    return PS2_OK;
}

}  // namespace
}  // namespace ps2

using namespace ps2;

extern "C" int ps2_hip_normalize_phase(const float *ph, int64_t n_rows, int64_t n_cols,
                                       const int64_t *cols, int64_t n_sel, int32_t n_threads,
                                       float *ph_nm, float *amp) {
    if (n_rows < 0 || n_cols < 0 || n_sel < 0 ||
        (n_rows > 0 && n_sel > 0 && (!ph || !cols || !ph_nm || !amp))) {
        set_error("normalize_phase: invalid argument");
        return PS2_E_ARG;
    }
    if (int rc = check_cols(cols, n_sel, n_cols)) return rc;
    parallel_rows(n_rows, thread_count(n_threads, n_rows), [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r) {
            const float *src = ph + 2 * r * n_cols;
            float *dn = ph_nm + 2 * r * n_sel;
            float *da = amp + r * n_sel;
            for (int64_t i = 0; i < n_sel; ++i) {
                const float re = src[2 * cols[i]], im = src[2 * cols[i] + 1];
                const float mag = hypotf_glibc(re, im);
                const float safe = mag == 0.0f ? 1.0f : mag;
                da[i] = safe;
                dn[2 * i] = re / safe;
                dn[2 * i + 1] = im / safe;
            }
        }
    });
    return PS2_OK;
}

extern "C" int ps2_hip_select_columns_f64(const double *src, int64_t n_rows, int64_t n_cols,
                                          const int64_t *cols, int64_t n_sel, int32_t n_threads,
                                          double *dst) {
    if (n_rows < 0 || n_cols < 0 || n_sel < 0 || (n_rows > 0 && n_sel > 0 && (!src || !cols || !dst))) {
        set_error("select_columns_f64: invalid argument");
        return PS2_E_ARG;
    }
    if (int rc = check_cols(cols, n_sel, n_cols)) return rc;
    parallel_rows(n_rows, thread_count(n_threads, n_rows), [&](int64_t a, int64_t b) {
        for (int64_t r = a; r < b; ++r)
            for (int64_t i = 0; i < n_sel; ++i) dst[r * n_sel + i] = src[r * n_cols + cols[i]];
    });
    return PS2_OK;
}
