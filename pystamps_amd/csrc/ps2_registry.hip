// Kernels of the pySTAMPS kernel registry on the GPU (gfx950): the `hip`
// provider of stage2_topofit / stage2_topofit_row_invariant /
// stage2_topofit_coh_row_invariant, stage2_grid_accumulate and
// stage2_histogram (pystamps/kernels/accelerated.py:392-426).
//
// The topofit kernels follow the Python pipeline's single-row fit
// (_ps_topofit_single, pystamps/pipeline/ported.py:2355-2430), in FP64:
// zero elements are left out of the fit, every local maximum of the trial
// coherence within 2e-4 of the maximum is refined, and the refined coherence
// picks among several candidates (endpoint pairs keep the coarse argmax).
#include "pystamps_hip.h"
#include "ps2_host_util.hpp"
#include "ps2_staging.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

namespace ps2 {
namespace {

constexpr double kPi4 = 0.78539816339744830962;
constexpr double kNearMaxTol = 2.0e-4;

struct PyFitParams {
    int64_t P;      // rows in this launch
    int64_t row0;   // first row of this launch
    int I, T, tn, rowinv;
    int cohrule;    // _ps_topofit_batch_row_invariant_coh's rule (shared bperp only)
};

// Element (row k, interferogram i) of the fit's arrays, for either layout:
// ifg-major planes [I][n] (c_i = n, c_k = 1) or the caller's row-major
// [n][I] (c_i = 1, c_k = I).  bperp has b_k = 0 when every row shares it.
struct RowView {
    const double2 *cpx;
    int64_t c_i, c_k;
    const double *bp;
    int64_t b_i, b_k;
    float2 *res;   // may be null
    int64_t r_i, r_k;
    __device__ double2 v(int64_t k, int i) const { return cpx[(int64_t)i * c_i + k * c_k]; }
    __device__ double b(int64_t k, int i) const { return bp[(int64_t)i * b_i + k * b_k]; }
};

// exp(-i ph) * (re + i im)
__device__ inline void rot(double re, double im, double ph, double &orr, double &oi) {
    double s, c;
    sincos(ph, &s, &c);
    orr = re * c + im * s;
    oi = im * c - re * s;
}

// _ps_topofit_refine_candidate (ported.py:2480-2513) of row k; writes the
// residual when write_res.
__device__ void py_refine(const RowView &rv, int64_t k, int I, double den_lin, double k0, double &K,
                          double &C, double &coh, bool write_res) {
    double ore = 0.0, oim = 0.0;
    for (int i = 0; i < I; ++i) {
        const double2 v = rv.v(k, i);
        if (v.x == 0.0 && v.y == 0.0) continue;
        double rr, ri;
        rot(v.x, v.y, k0 * rv.b(k, i), rr, ri);
        ore += rr;
        oim += ri;
    }
    double num = 0.0;
    for (int i = 0; i < I; ++i) {
        const double2 v = rv.v(k, i);
        if (v.x == 0.0 && v.y == 0.0) continue;
        const double b = rv.b(k, i);
        double rr, ri;
        rot(v.x, v.y, k0 * b, rr, ri);
        const double w = hypot(v.x, v.y);
        const double pr = rr * ore + ri * oim, pi = ri * ore - rr * oim;
        num += (w * b) * (w * atan2(pi, pr));
    }
    const double kk = k0 + num / den_lin;
    double mr = 0.0, mi = 0.0, d2 = 0.0;
    for (int i = 0; i < I; ++i) {
        const double2 v = rv.v(k, i);
        double rr = 0.0, ri = 0.0;
        if (v.x != 0.0 || v.y != 0.0) {
            rot(v.x, v.y, kk * rv.b(k, i), rr, ri);
            mr += rr;
            mi += ri;
            d2 += hypot(rr, ri);
        }
        if (write_res) rv.res[(int64_t)i * rv.r_i + k * rv.r_k] = make_float2((float)rr, (float)ri);
    }
    if (d2 == 0.0) d2 = 1.0;
    K = kk;
    C = atan2(mi, mr);
    coh = hypot(mr, mi) / d2;
}

// one thread per row k = row0 + local; cohs [T][P] scratch of the launch's
// rows; Kout/Cout/cohout indexed by k
__global__ __launch_bounds__(256) void k_py_topofit(RowView rv, double *__restrict__ cohs,
                                                    double *__restrict__ Kout,
                                                    double *__restrict__ Cout,
                                                    double *__restrict__ cohout, PyFitParams p) {
    const int64_t local = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= p.P) return;
    const int64_t k = p.row0 + local;
    const int I = p.I, T = p.T;
    if (p.cohrule) {
        // ported.py:2561-2608: trial grid over the whole bperp vector's range
        // (zero elements included); a row with exactly one near-maximum local
        // peak on it is refined from that peak, every other row continues
        // below with _ps_topofit_single's fit (:2611-2614)
        double fmn = INFINITY, fmx = -INFINITY, dn = 0.0, dl = 0.0;
        for (int i = 0; i < I; ++i) {
            const double b = rv.b(k, i);
            fmn = fmin(fmn, b);
            fmx = fmax(fmx, b);
            const double2 v = rv.v(k, i);
            if (v.x == 0.0 && v.y == 0.0) continue;
            const double w = hypot(v.x, v.y);
            dn += w;
            dl += (w * b) * (w * b);
        }
        double fr = fmx - fmn;
        if (fr == 0.0) fr = 1.0;
        if (dn == 0.0) dn = 1.0;
        if (dl == 0.0) dl = 1.0;
        double cm = -INFINITY;
        for (int t = 0; t < T; ++t) {
            const double tm = (double)(t - p.tn);
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < I; ++i) {
                const double2 v = rv.v(k, i);
                if (v.x == 0.0 && v.y == 0.0) continue;
                double rr, ri;
                rot(v.x, v.y, (rv.b(k, i) / fr * kPi4) * tm, rr, ri);
                sr += rr;
                si += ri;
            }
            const double c = hypot(sr, si) / dn;
            cohs[(int64_t)t * p.P + local] = c;
            cm = fmax(cm, c);
        }
        int count = 0, only = 0;
        for (int t = 0; t < T; ++t) {
            const double c = cohs[(int64_t)t * p.P + local];
            bool lm;
            if (T == 1)
                lm = true;
            else if (t == 0)
                lm = c >= cohs[(int64_t)1 * p.P + local];
            else if (t == T - 1)
                lm = c >= cohs[(int64_t)(T - 2) * p.P + local];
            else
                lm = c >= cohs[(int64_t)(t - 1) * p.P + local] && c >= cohs[(int64_t)(t + 1) * p.P + local];
            if (lm && c >= cm - kNearMaxTol) {
                if (count == 0) only = t;
                ++count;
            }
        }
        if (count == 1) {
            double K, C, coh;
            py_refine(rv, k, I, dl, kPi4 / fr * (double)(only - p.tn), K, C, coh, false);
            Kout[k] = K;
            Cout[k] = C;
            cohout[k] = coh;
            return;
        }
    }
    double mn = INFINITY, mx = -INFINITY, denom = 0.0, den_lin = 0.0;
    int m = 0;
    for (int i = 0; i < I; ++i) {
        const double2 v = rv.v(k, i);
        if (v.x == 0.0 && v.y == 0.0) continue;
        const double b = rv.b(k, i);
        mn = fmin(mn, b);
        mx = fmax(mx, b);
        const double w = hypot(v.x, v.y);
        denom += w;
        den_lin += (w * b) * (w * b);
        ++m;
    }
    if (m == 0) {
        Kout[k] = Cout[k] = cohout[k] = __builtin_nan("");
        if (rv.res)
            for (int i = 0; i < I; ++i) rv.res[(int64_t)i * rv.r_i + k * rv.r_k] = make_float2(0.0f, 0.0f);
        return;
    }
    double range = mx - mn;
    if (range == 0.0) range = 1.0;
    if (denom == 0.0) denom = 1.0;
    if (den_lin == 0.0) den_lin = 1.0;
    double cmax = -INFINITY;
    int amax = 0;
    for (int t = 0; t < T; ++t) {
        const double tm = (double)(t - p.tn);
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < I; ++i) {
            const double2 v = rv.v(k, i);
            if (v.x == 0.0 && v.y == 0.0) continue;
            double rr, ri;
            rot(v.x, v.y, (rv.b(k, i) / range * kPi4) * tm, rr, ri);
            sr += rr;
            si += ri;
        }
        const double c = hypot(sr, si) / denom;
        cohs[(int64_t)t * p.P + local] = c;
        if (c > cmax) {
            cmax = c;
            amax = t;
        }
    }
    // candidates in ascending trial order (ported.py:2433-2448) and the
    // selection rule (:2451-2477), tracked on the fly
    int n = 0, first = -1, last = -1, coarse_best = -1, refined_best = -1;
    double coarse_v = -INFINITY, refined_v = -INFINITY;
    for (int t = 0; t < T; ++t) {
        const double c = cohs[(int64_t)t * p.P + local];
        bool lm;
        if (T == 1)
            lm = true;
        else if (t == 0)
            lm = c >= cohs[(int64_t)1 * p.P + local];
        else if (t == T - 1)
            lm = c >= cohs[(int64_t)(T - 2) * p.P + local];
        else
            lm = c >= cohs[(int64_t)(t - 1) * p.P + local] && c >= cohs[(int64_t)(t + 1) * p.P + local];
        if (!(lm && c >= cmax - kNearMaxTol)) continue;
        if (n == 0) first = t;
        last = t;
        ++n;
        if (c > coarse_v) {
            coarse_v = c;
            coarse_best = t;
        }
        double Kc, Cc, rc;
        py_refine(rv, k, I, den_lin, kPi4 / range * (double)(t - p.tn), Kc, Cc, rc, false);
        if (rc > refined_v) {
            refined_v = rc;
            refined_best = t;
        }
    }
    int sel;
    if (n == 0)
        sel = amax;
    else if (n == 1 || (n == 2 && first == 0 && last == T - 1))
        sel = coarse_best;
    else
        sel = refined_best;
    double K, C, coh;
    py_refine(rv, k, I, den_lin, kPi4 / range * (double)(sel - p.tn), K, C, coh, rv.res != nullptr);
    Kout[k] = K;
    Cout[k] = C;
    cohout[k] = coh;
}

// np.add.at order: thread per (cell, ifg), ifg fastest; the cell's rows in
// ascending row order, complex64 sums
__global__ __launch_bounds__(256) void k_py_grid_accumulate(const float2 *__restrict__ ph,
                                                            const int64_t *__restrict__ order,
                                                            const int64_t *__restrict__ cell_start,
                                                            float2 *__restrict__ out, int64_t C,
                                                            int I) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= C * I) return;
    const int64_t c = gid / I;
    const int i = (int)(gid - c * I);
    float sr = 0.0f, si = 0.0f;
    for (int64_t q = cell_start[c]; q < cell_start[c + 1]; ++q) {
        const float2 v = ph[order[q] * I + i];
        sr += v.x;
        si += v.y;
    }
    out[gid] = make_float2(sr, si);
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() {
        if (p) {
            AllocGuard g(capture_mutex());
            (void)hipFree(p);
        }
    }
    hipError_t alloc(size_t n) {
        AllocGuard g(capture_mutex());
        return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1));
    }
};

#define REG_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            set_error(std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
            return e_ == hipErrorOutOfMemory ? PS2_E_NOMEM : PS2_E_DEVICE;                \
        }                                                                                 \
    } while (0)

// A private non-blocking stream for one call: no device-wide synchronisation,
// so calls run beside other work on the device (sessions, graph captures).
struct CallStream {
    hipStream_t s = nullptr;
    ~CallStream() {
        if (s) {
            (void)hipStreamSynchronize(s);
            AllocGuard g(capture_mutex());
            (void)hipStreamDestroy(s);
        }
    }
    hipError_t create() {
        AllocGuard g(capture_mutex());
        return hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    }
};

int set_device(int32_t device) {
    int n = 0;
    REG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("device " + std::to_string(device) + " not present (" + std::to_string(n) +
                  " visible)");
        return PS2_E_ARG;
    }
    REG_HIP(hipSetDevice(device));
    return PS2_OK;
}

}  // namespace
}  // namespace ps2

using namespace ps2;

namespace {

// ps2_hip_topofit_rows and ps2_hip_topofit_coh_row_invariant: host arrays in,
// host arrays out, one kernel launch on a stream of the call's own.  k_out and
// c_out may be null only under the coherence rule.
int topofit_rows_host(const char *who, const double *cpx, const double *bperp, int32_t row_invariant,
                      int cohrule, int64_t n_rows, int64_t n_ifg, double n_trial_wraps,
                      int32_t device, double *k_out, double *c_out, double *coh_out,
                      float *res_out) {
    const std::string name(who);
    if (n_rows < 0 || n_ifg <= 0 ||
        (n_rows > 0 && (!cpx || !bperp || !coh_out || (!cohrule && (!k_out || !c_out))))) {
        set_error(name + ": invalid argument");
        return PS2_E_ARG;
    }
    if (!(n_trial_wraps >= 0.0) || !std::isfinite(n_trial_wraps) || n_trial_wraps > 1.0e6) {
        set_error(name + ": n_trial_wraps must be finite and >= 0");
        return PS2_E_ARG;
    }
    if (n_rows == 0) return PS2_OK;
    if (n_rows * n_ifg >= ((int64_t)1 << 31) / 2) {
        set_error(name + ": n_rows * n_ifg too large for one call");
        return PS2_E_ARG;
    }
    int rc = set_device(device);
    if (rc) return rc;
    const int64_t P = n_rows, I = n_ifg;
    const int tn = (int)std::ceil(8.0 * n_trial_wraps);
    const int T = 2 * tn + 1;
    // ifg-major copies for coalesced per-row threads
    std::vector<double> hc(2 * P * I), hb(row_invariant ? I : P * I);
    for (int64_t k = 0; k < P; ++k)
        for (int64_t i = 0; i < I; ++i) {
            hc[2 * (i * P + k)] = cpx[2 * (k * I + i)];
            hc[2 * (i * P + k) + 1] = cpx[2 * (k * I + i) + 1];
            if (!row_invariant) hb[i * P + k] = bperp[k * I + i];
        }
    if (row_invariant)
        for (int64_t i = 0; i < I; ++i) hb[i] = bperp[i];
    DevBuf<double2> dc;
    DevBuf<double> db, dcoh_t, dK, dC, dcoh;
    DevBuf<float2> dres;
    REG_HIP(dc.alloc(P * I));
    REG_HIP(db.alloc(hb.size()));
    REG_HIP(dcoh_t.alloc((size_t)T * P));
    REG_HIP(dK.alloc(P));
    REG_HIP(dC.alloc(P));
    REG_HIP(dcoh.alloc(P));
    if (res_out) REG_HIP(dres.alloc(P * I));
    CallStream cs;
    REG_HIP(cs.create());
    REG_HIP(hipMemcpyAsync(dc.p, hc.data(), sizeof(double) * hc.size(), hipMemcpyHostToDevice, cs.s));
    REG_HIP(hipMemcpyAsync(db.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, cs.s));
    PyFitParams prm{P, 0, (int)I, T, tn, row_invariant ? 1 : 0, cohrule};
    const RowView rv{dc.p, P, 1, db.p, row_invariant ? 1 : P, row_invariant ? 0 : 1,
                     res_out ? dres.p : nullptr, P, 1};
    hipLaunchKernelGGL(k_py_topofit, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, cs.s, rv, dcoh_t.p,
                       dK.p, dC.p, dcoh.p, prm);
    REG_HIP(hipGetLastError());
    if (k_out) REG_HIP(hipMemcpyAsync(k_out, dK.p, sizeof(double) * P, hipMemcpyDeviceToHost, cs.s));
    if (c_out) REG_HIP(hipMemcpyAsync(c_out, dC.p, sizeof(double) * P, hipMemcpyDeviceToHost, cs.s));
    REG_HIP(hipMemcpyAsync(coh_out, dcoh.p, sizeof(double) * P, hipMemcpyDeviceToHost, cs.s));
    REG_HIP(hipStreamSynchronize(cs.s));
    if (res_out) {
        std::vector<float2> hr(P * I);
        REG_HIP(hipMemcpyAsync(hr.data(), dres.p, sizeof(float2) * hr.size(), hipMemcpyDeviceToHost, cs.s));
        REG_HIP(hipStreamSynchronize(cs.s));
        for (int64_t k = 0; k < P; ++k)
            for (int64_t i = 0; i < I; ++i) {
                res_out[2 * (k * I + i)] = hr[i * P + k].x;
                res_out[2 * (k * I + i) + 1] = hr[i * P + k].y;
            }
    }
    return PS2_OK;
}

}  // namespace

extern "C" int ps2_hip_topofit_rows(const double *cpx, const double *bperp, int32_t row_invariant,
                                    int64_t n_rows, int64_t n_ifg, double n_trial_wraps,
                                    int32_t device, double *k_out, double *c_out,
                                    double *coh_out, float *res_out) {
    return topofit_rows_host("topofit_rows", cpx, bperp, row_invariant, 0, n_rows, n_ifg,
                             n_trial_wraps, device, k_out, c_out, coh_out, res_out);
}

extern "C" int ps2_hip_topofit_coh_row_invariant(const double *cpx, const double *bperp_vec,
                                                 int64_t n_rows, int64_t n_ifg,
                                                 double n_trial_wraps, int32_t device,
                                                 double *coh_out) {
    return topofit_rows_host("topofit_coh_row_invariant", cpx, bperp_vec, 1, 1, n_rows, n_ifg,
                             n_trial_wraps, device, nullptr, nullptr, coh_out, nullptr);
}

// ---- allocation-free topofit ----------------------------------------------
// ps2_hip_topofit_rows_noalloc keeps the per-row trial coherences of one chunk
// of rows ([T][chunk] doubles) in this array instead of a hipMalloc'd buffer.
// Each GPU the library runs on holds its own copy until the process ends.
// Host code uses the array only through the per-device address that
// hipGetSymbolAddress returns (noalloc_scratch_address).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kNoAllocRowBlock = 256;   // rows per workgroup of k_py_topofit
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// address of `scratch` on the current device (looked up once per device)
int noalloc_scratch_address(int device, double **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) {
        set_error("topofit_rows_noalloc: device ordinal " + std::to_string(device) + " out of range");
        return PS2_E_ARG;
    }
    if (!g_scratch_addr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("topofit_rows_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return PS2_E_DEVICE;
        }
        g_scratch_addr[device] = ptr;
    }
    *out = static_cast<double *>(g_scratch_addr[device]);
    return PS2_OK;
}

}  // namespace

extern "C" int ps2_hip_topofit_rows_noalloc(const double *d_cpx, const double *d_bperp,
                                            int32_t row_invariant, int64_t n_rows, int64_t n_ifg,
                                            double n_trial_wraps, int32_t device, void *stream,
                                            double *d_k_out, double *d_c_out, double *d_coh_out,
                                            float *d_res_out) {
    if (n_rows < 0 || n_ifg <= 0 || n_ifg > (1 << 20) ||
        (n_rows > 0 && (!d_cpx || !d_bperp || !d_k_out || !d_c_out || !d_coh_out))) {
        set_error("topofit_rows_noalloc: invalid argument");
        return PS2_E_ARG;
    }
    if (!(n_trial_wraps >= 0.0) || !std::isfinite(n_trial_wraps) || n_trial_wraps > 1.0e6) {
        set_error("topofit_rows_noalloc: n_trial_wraps must be finite and >= 0");
        return PS2_E_ARG;
    }
    if (n_rows > (((int64_t)1 << 62) / n_ifg)) {
        set_error("topofit_rows_noalloc: n_rows * n_ifg too large");
        return PS2_E_ARG;
    }
    const int tn = (int)std::ceil(8.0 * n_trial_wraps);
    const int64_t T = 2 * (int64_t)tn + 1;
    // rows per launch: whole workgroups whose [T][rows] coherences fit the scratch
    const int64_t cap_rows = (int64_t)(kNoAllocScratchBytes / sizeof(double)) / T / kNoAllocRowBlock *
                             kNoAllocRowBlock;
    if (cap_rows < kNoAllocRowBlock) {
        set_error("topofit_rows_noalloc: " + std::to_string(T) + " trials need " +
                  std::to_string(T * kNoAllocRowBlock * (int64_t)sizeof(double)) +
                  " bytes of scratch per 256 rows; the scratch holds " +
                  std::to_string(kNoAllocScratchBytes) + " bytes (at most " +
                  std::to_string(kNoAllocScratchBytes / sizeof(double) / kNoAllocRowBlock) + " trials)");
        return PS2_E_ARG;
    }
    if (n_rows == 0) return PS2_OK;
    int rc = set_device(device);
    if (rc) return rc;
    double *cohs = nullptr;
    if ((rc = noalloc_scratch_address(device, &cohs))) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    // the caller's row-major layouts: element (k, i) at k * n_ifg + i
    const RowView rv{reinterpret_cast<const double2 *>(d_cpx), 1, n_ifg, d_bperp, 1,
                     row_invariant ? 0 : n_ifg, reinterpret_cast<float2 *>(d_res_out), 1, n_ifg};
    for (int64_t row0 = 0; row0 < n_rows; row0 += cap_rows) {
        const int64_t rows = std::min(cap_rows, n_rows - row0);
        PyFitParams prm{rows, row0, (int)n_ifg, (int)T, tn, row_invariant ? 1 : 0, 0};
        hipLaunchKernelGGL(k_py_topofit, dim3((unsigned)((rows + kNoAllocRowBlock - 1) / kNoAllocRowBlock)),
                           dim3(kNoAllocRowBlock), 0, st, rv, cohs, d_k_out, d_c_out, d_coh_out, prm);
        REG_HIP(hipGetLastError());
    }
    REG_HIP(hipStreamSynchronize(st));
    return PS2_OK;
}

extern "C" int ps2_hip_grid_accumulate(const float *ph_weight, const int64_t *grid_lin,
                                       int64_t n_rows, int64_t n_ifg, int64_t n_i, int64_t n_j,
                                       int32_t device, float *out) {
    if (n_rows < 0 || n_ifg <= 0 || n_i <= 0 || n_j <= 0 || !out ||
        (n_rows > 0 && (!ph_weight || !grid_lin))) {
        set_error("grid_accumulate: invalid argument");
        return PS2_E_ARG;
    }
    const int64_t C = n_i * n_j, P = n_rows, I = n_ifg;
    std::vector<int64_t> start(C + 1, 0), order(P > 0 ? P : 1);
    for (int64_t k = 0; k < P; ++k) {
        if (grid_lin[k] < 0 || grid_lin[k] >= C) {
            set_error("grid_accumulate: grid index " + std::to_string(grid_lin[k]) + " of row " +
                      std::to_string(k) + " is out of range for " + std::to_string(n_i) + "x" +
                      std::to_string(n_j) + " cells");
            return PS2_E_ARG;
        }
        ++start[grid_lin[k] + 1];
    }
    for (int64_t c = 0; c < C; ++c) start[c + 1] += start[c];
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t k = 0; k < P; ++k) order[fill[grid_lin[k]]++] = k;
    }
    int rc = set_device(device);
    if (rc) return rc;
    DevBuf<float2> dph, dout;
    DevBuf<int64_t> dorder, dstart;
    REG_HIP(dph.alloc(P * I));
    REG_HIP(dout.alloc(C * I));
    REG_HIP(dorder.alloc(P));
    REG_HIP(dstart.alloc(C + 1));
    CallStream cs;
    REG_HIP(cs.create());
    if (P > 0) {
        REG_HIP(hipMemcpyAsync(dph.p, ph_weight, sizeof(float2) * P * I, hipMemcpyHostToDevice, cs.s));
        REG_HIP(hipMemcpyAsync(dorder.p, order.data(), sizeof(int64_t) * P, hipMemcpyHostToDevice, cs.s));
    }
    REG_HIP(hipMemcpyAsync(dstart.p, start.data(), sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, cs.s));
    hipLaunchKernelGGL(k_py_grid_accumulate, dim3((unsigned)((C * I + 255) / 256)), dim3(256), 0, cs.s,
                       dph.p, dorder.p, dstart.p, dout.p, C, (int)I);
    REG_HIP(hipGetLastError());
    REG_HIP(hipMemcpyAsync(out, dout.p, sizeof(float2) * C * I, hipMemcpyDeviceToHost, cs.s));
    REG_HIP(hipStreamSynchronize(cs.s));
    return PS2_OK;
}

extern "C" int ps2_hip_histogram_with_centers(const double *values, int64_t n,
                                              const double *centers, int64_t n_centers,
                                              double *out) {
    if (n < 0 || n_centers < 0 || (n > 0 && !values) || (n_centers > 0 && (!centers || !out))) {
        set_error("histogram_with_centers: invalid argument");
        return PS2_E_ARG;
    }
    if (n_centers == 0) return PS2_OK;
    for (int64_t b = 0; b < n_centers; ++b) out[b] = 0.0;
    if (n_centers == 1) {
        for (int64_t q = 0; q < n; ++q)
            if (std::isfinite(values[q])) out[0] += 1.0;
        return PS2_OK;
    }
    const HistRule hr = make_hist_rule(centers, n_centers);
    for (int64_t q = 0; q < n; ++q) {
        const double v = values[q];
        if (!std::isfinite(v)) continue;
        int64_t b;
        if (hr.equal) {
            double a = std::ceil((v - hr.cutoff0) / hr.d);
            if (!(a > 0.0)) a = 0.0;
            if (a > (double)(n_centers - 1)) a = (double)(n_centers - 1);
            b = (int64_t)a;
        } else {
            b = std::lower_bound(hr.mids.begin(), hr.mids.end(), v) - hr.mids.begin();
            if (b > n_centers - 1) b = n_centers - 1;
        }
        out[b] += 1.0;
    }
    return PS2_OK;
}
