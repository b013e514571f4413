// Random-coherence histogram Nr of the stage-2 prep on the device
// (random_coherence_histogram, native_stage2.rs:1310-1512, single-master
// branch).  The MATLAB v5 uniform draws are serial and are made on the host;
// each random phase row is then fitted in FP64 by one device thread, exactly
// as topofit_phase_row_coh + refine_phase_candidate_coh, and binned with
// integer atomics, so the counts are exact and order-independent.
#include "pystamps_hip.h"
#include "ps2_host_util.hpp"
#include "ps2_staging.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

namespace ps2 {

struct NrParams {
    int64_t n_rows;
    int I, tn;
    double range;   // bperp range (1 if 0)
    double den_lin; // sum bp^2 (1 if 0)
    int equal;
    double cutoff0, d;
    int nbins;
};

constexpr int NR_TC = 8;

__global__ __launch_bounds__(256) void k_nr_rows(const double *__restrict__ ph,
                                                 const double *__restrict__ bp,
                                                 const double *__restrict__ mids,
                                                 uint32_t *__restrict__ hist, NrParams p) {
    __shared__ uint32_t lh[128];
    for (int q = threadIdx.x; q < 128; q += blockDim.x) lh[q] = 0;
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < p.n_rows) {
        const int I = p.I;
        const double pi4 = 0.78539816339744830962;
        double best = -INFINITY;
        int best_t = 0;
        bool first = true;
        // trial scan, chunks of NR_TC trials in ascending order; per trial the
        // ifg sum runs in ascending order (native_stage2.rs:1458-1469)
        for (int t0 = -p.tn; t0 <= p.tn; t0 += NR_TC) {
            double sr[NR_TC], si[NR_TC];
#pragma unroll
            for (int q = 0; q < NR_TC; ++q) sr[q] = si[q] = 0.0;
            for (int i = 0; i < I; ++i) {
                const double phi = ph[r + (int64_t)i * p.n_rows];
                const double a = (bp[i] / p.range) * pi4;
#pragma unroll
                for (int q = 0; q < NR_TC; ++q) {
                    const int t = t0 + q;
                    if (t <= p.tn) {
                        double sn, cs;
                        sincos(phi - a * (double)t, &sn, &cs);
                        sr[q] += cs;
                        si[q] += sn;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NR_TC; ++q) {
                const int t = t0 + q;
                if (t <= p.tn) {
                    const double c = hypot(sr[q], si[q]) / (double)I;
                    if (first) {
                        best = c;
                        best_t = t;
                        first = false;
                    } else if (c > best) {
                        best = c;
                        best_t = t;
                    }
                }
            }
        }
        const double k0 = pi4 / p.range * (double)best_t;
        double ore = 0.0, oim = 0.0;
        for (int i = 0; i < I; ++i) {
            double sn, cs;
            sincos(ph[r + (int64_t)i * p.n_rows] - k0 * bp[i], &sn, &cs);
            ore += cs;
            oim += sn;
        }
        const double cre = ore, cim = -oim;
        double num = 0.0;
        for (int i = 0; i < I; ++i) {
            double sn, cs;
            sincos(ph[r + (int64_t)i * p.n_rows] - k0 * bp[i], &sn, &cs);
            const double pr = cs * cre - sn * cim, pim = cs * cim + sn * cre;
            num += bp[i] * atan2(pim, pr);
        }
        const double kk = k0 + num / p.den_lin;
        double fr = 0.0, fi = 0.0;
        for (int i = 0; i < I; ++i) {
            double sn, cs;
            sincos(ph[r + (int64_t)i * p.n_rows] - kk * bp[i], &sn, &cs);
            fr += cs;
            fi += sn;
        }
        const double coh = hypot(fr, fi) / (double)(I > 1 ? I : 1);
        int bin = -1;
        if (__builtin_isfinite(coh)) {
            if (p.equal) {
                if (p.d == 0.0) {
                    bin = 0;
                } else {
                    double a = ceil((coh - p.cutoff0) / p.d);
                    if (a < 0.0) a = 0.0;
                    if (a > (double)(p.nbins - 1)) a = (double)(p.nbins - 1);
                    bin = (int)a;
                }
            } else {
                bin = 0;
                while (bin < p.nbins - 1 && mids[bin] < coh) ++bin;
            }
        }
        if (bin >= 0) atomicAdd(&lh[bin], 1u);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < p.nbins; q += blockDim.x)
        if (lh[q]) atomicAdd(&hist[q], lh[q]);
}

}  // namespace ps2

using namespace ps2;

#define NR_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            set_error(std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
            rc = PS2_E_DEVICE;                                                             \
            goto done;                                                                     \
        }                                                                                  \
    } while (0)

namespace {

// uniform phases, column-major [n_rows][n_ifg] (x 2 pi): single master draws
// the interferogram columns directly (native_stage2.rs:1336-1349); small
// baseline draws image columns and takes b - a (:1321-1334)
int nr_check(const double *bperp_nm, int64_t n_ifg, const double *trial_values, int64_t n_trial,
             const double *coh_bins, int64_t n_rows, const double *out_nr,
             const double *out_nr_max_nz_ix) {
    if (!bperp_nm || !trial_values || !coh_bins || !out_nr || !out_nr_max_nz_ix || n_ifg <= 0 ||
        n_rows < 0) {
        set_error("random_coherence_histogram: invalid argument");
        return PS2_E_ARG;
    }
    if (n_trial < 1 || n_trial % 2 != 1) {
        set_error("trial_values must hold 2*tn+1 values");
        return PS2_E_ARG;
    }
    const int64_t tn = n_trial / 2;
    for (int64_t t = 0; t < n_trial; ++t)
        if (trial_values[t] != (double)(t - tn)) {
            set_error("trial_values must be the consecutive integers -tn..tn");
            return PS2_E_ARG;
        }
    return PS2_OK;
}

// FP64 fit of every row of ph on the device, histogram, Nr_max_nz_ix
int nr_histogram(const std::vector<double> &ph, const double *bperp_nm, int64_t n_ifg,
                 int64_t n_trial, const double *coh_bins, int64_t n_rows, int32_t device,
                 double *out_nr, double *out_nr_max_nz_ix) {
    int rc = PS2_OK;
    double *d_ph = nullptr, *d_bp = nullptr, *d_mids = nullptr;
    uint32_t *d_hist = nullptr;
    hipStream_t st = nullptr;   // private stream: no device-wide synchronisation
    std::vector<uint32_t> counts(128, 0);
    NrParams p{};
    {
        int ndev = 0;
        NR_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) {
            set_error("device not present");
            rc = PS2_E_ARG;
            goto done;
        }
        NR_HIP(hipSetDevice(device));
        double mn = INFINITY, mx = -INFINITY, den = 0.0;
        for (int64_t i = 0; i < n_ifg; ++i) {
            mn = std::fmin(mn, bperp_nm[i]);
            mx = std::fmax(mx, bperp_nm[i]);
            den += bperp_nm[i] * bperp_nm[i];
        }
        p.n_rows = n_rows;
        p.I = (int)n_ifg;
        p.tn = (int)(n_trial / 2);
        p.range = (mx - mn) == 0.0 ? 1.0 : (mx - mn);
        p.den_lin = den == 0.0 ? 1.0 : den;
        const HistRule hr = make_hist_rule(coh_bins, PS2_COH_BINS);
        p.equal = hr.equal ? 1 : 0;
        p.cutoff0 = hr.cutoff0;
        p.d = hr.d;
        p.nbins = PS2_COH_BINS;
        std::vector<double> mids(128, 0.0);
        for (size_t k = 0; k < hr.mids.size() && k < 128; ++k) mids[k] = hr.mids[k];
        {
            AllocGuard g(capture_mutex());
            NR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            NR_HIP(hipMalloc(&d_ph, sizeof(double) * ph.size()));
            NR_HIP(hipMalloc(&d_bp, sizeof(double) * n_ifg));
            NR_HIP(hipMalloc(&d_mids, sizeof(double) * 128));
            NR_HIP(hipMalloc(&d_hist, sizeof(uint32_t) * 128));
        }
        if ((rc = staging_h2d(d_ph, ph.data(), sizeof(double) * ph.size(), st))) goto done;
        NR_HIP(hipMemcpyAsync(d_bp, bperp_nm, sizeof(double) * n_ifg, hipMemcpyHostToDevice, st));
        NR_HIP(hipMemcpyAsync(d_mids, mids.data(), sizeof(double) * 128, hipMemcpyHostToDevice, st));
        NR_HIP(hipMemsetAsync(d_hist, 0, sizeof(uint32_t) * 128, st));
        if (n_rows > 0) {
            hipLaunchKernelGGL(k_nr_rows, dim3((unsigned)((n_rows + 255) / 256)), dim3(256), 0, st, d_ph,
                               d_bp, d_mids, d_hist, p);
            NR_HIP(hipGetLastError());
        }
        NR_HIP(hipMemcpyAsync(counts.data(), d_hist, sizeof(uint32_t) * 128, hipMemcpyDeviceToHost, st));
        NR_HIP(hipStreamSynchronize(st));
    }
    {
        int last = -1;
        for (int b = 0; b < PS2_COH_BINS; ++b) {
            out_nr[b] = (double)counts[b];
            if (counts[b] > 0) last = b;
        }
        *out_nr_max_nz_ix = last >= 0 ? (double)(last + 1) : 1.0;
    }
done:
    if (st) (void)hipStreamSynchronize(st);
    AllocGuard g(capture_mutex());
    if (d_ph) (void)hipFree(d_ph);
    if (d_bp) (void)hipFree(d_bp);
    if (d_mids) (void)hipFree(d_mids);
    if (d_hist) (void)hipFree(d_hist);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;

// rows x cols uniform draws of the MATLAB v5 stream, x 2 pi, in draw order
std::vector<double> uniform_draws(uint32_t seed, int64_t rows, int64_t cols) {
    const int64_t total = rows * cols;
    std::vector<double> v((size_t)(total > 0 ? total : 1), 0.0);
    MatlabV5Uniform rng(seed);
    for (int64_t q = 0; q < total; ++q) v[q] = rng.next() * kTwoPi;
    return v;
}

}  // namespace

extern "C" int ps2_hip_random_coherence_histogram(const double *bperp_nm, int64_t n_ifg,
                                                  const double *trial_values, int64_t n_trial,
                                                  const double *coh_bins, int64_t n_rows,
                                                  uint32_t seed, int32_t device, double *out_nr,
                                                  double *out_nr_max_nz_ix) {
    int rc = nr_check(bperp_nm, n_ifg, trial_values, n_trial, coh_bins, n_rows, out_nr, out_nr_max_nz_ix);
    if (rc) return rc;
    const std::vector<double> ph = uniform_draws(seed, n_rows, n_ifg);
    return nr_histogram(ph, bperp_nm, n_ifg, n_trial, coh_bins, n_rows, device, out_nr, out_nr_max_nz_ix);
}

extern "C" int ps2_hip_random_coherence_histogram_sb(const double *bperp_nm, int64_t n_ifg,
                                                     const int32_t *image_a, const int32_t *image_b,
                                                     int64_t n_image, const double *trial_values,
                                                     int64_t n_trial, const double *coh_bins,
                                                     int64_t n_rows, uint32_t seed, int32_t device,
                                                     double *out_nr, double *out_nr_max_nz_ix) {
    int rc = nr_check(bperp_nm, n_ifg, trial_values, n_trial, coh_bins, n_rows, out_nr, out_nr_max_nz_ix);
    if (rc) return rc;
    if (!image_a || !image_b || n_image <= 0) {
        set_error("random_coherence_histogram_sb: image indices are required");
        return PS2_E_ARG;
    }
    for (int64_t i = 0; i < n_ifg; ++i)
        if (image_a[i] < 0 || image_a[i] >= n_image || image_b[i] < 0 || image_b[i] >= n_image) {
            set_error("random_coherence_histogram_sb: image index out of range at ifg " +
                      std::to_string(i));
            return PS2_E_ARG;
        }
    const std::vector<double> img = uniform_draws(seed, n_rows, n_image);
    const int64_t total = n_rows * n_ifg;
    std::vector<double> ph((size_t)(total > 0 ? total : 1));
    for (int64_t i = 0; i < n_ifg; ++i)
        for (int64_t r = 0; r < n_rows; ++r)
            ph[r + i * n_rows] = img[r + (int64_t)image_b[i] * n_rows] - img[r + (int64_t)image_a[i] * n_rows];
    return nr_histogram(ph, bperp_nm, n_ifg, n_trial, coh_bins, n_rows, device, out_nr, out_nr_max_nz_ix);
}
