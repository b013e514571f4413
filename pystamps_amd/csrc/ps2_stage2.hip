// pystamps_hip: stage-2 gamma loop on one MI355X (gfx950).
// C ABI declared in include/pystamps_hip.h.
#include "pystamps_hip.h"
#include "ps2_host_util.hpp"
#include "ps2_kernels.hpp"
#include "ps2_staging.hpp"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/file.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <utility>
#include <string>
#include <thread>
#include <vector>

namespace ps2 {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

std::shared_mutex &capture_mutex() {
    static std::shared_mutex *m = new std::shared_mutex();   // process lifetime
    return *m;
}

#define PS2_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)); \
            return PS2_E_DEVICE;                                                        \
        }                                                                               \
    } while (0)

#define PS2_LAUNCHED()                                                                  \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_));    \
            return PS2_E_DEVICE;                                                        \
        }                                                                               \
    } while (0)

static int arg_error(const std::string &msg) {
    set_error(msg);
    return PS2_E_ARG;
}

static inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Device buffer that frees itself.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p) {
            AllocGuard g(capture_mutex());
            (void)hipFree(p);
        }
    }
    int alloc(size_t n) {
        AllocGuard g(capture_mutex());
        if (p) {
            (void)hipFree(p);
            p = nullptr;
        }
        bytes = n;
        if (n == 0) return PS2_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("device allocation of " + std::to_string(n) + " bytes failed: " +
                      hipGetErrorString(e));
            return PS2_E_NOMEM;
        }
        return PS2_OK;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

struct HostPinned {
    void *p = nullptr;
    ~HostPinned() {
        if (p) {
            AllocGuard g(capture_mutex());
            (void)hipHostFree(p);
        }
    }
    int alloc(size_t n, unsigned flags = hipHostMallocDefault) {
        AllocGuard g(capture_mutex());
        hipError_t e = hipHostMalloc(&p, n, flags);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("pinned allocation failed: ") + hipGetErrorString(e));
            return PS2_E_NOMEM;
        }
        return PS2_OK;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Bulk host <-> device copies of caller-owned (pageable) arrays.  A direct
// hipMemcpy from pageable memory runs at ~10-17 GB/s and faults the caller's
// fresh pages in on one thread; here the data moves through two pinned chunks
// of one process-wide pool: the DMA of one chunk overlaps a multi-threaded
// host copy of the other.  Transfers share the pool under a mutex.
class PinnedStaging {
  public:
    static constexpr size_t kChunk = (size_t)64 << 20;
    static constexpr size_t kDirect = (size_t)2 << 20;   // below this: plain hipMemcpy

    int d2h(void *dst, const void *src, size_t n, hipStream_t st) {
        if (n == 0) return PS2_OK;
        if (n < kDirect) {
            PS2_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, st));
            PS2_HIP(hipStreamSynchronize(st));
            return PS2_OK;
        }
        std::lock_guard<std::mutex> lk(mu_);
        int rc;
        if ((rc = ensure())) return rc;
        // on an error, wait for the DMA already queued before the pool is released
        if ((rc = d2h_chunks(dst, src, n, st))) (void)hipStreamSynchronize(st);
        return rc;
    }

    int h2d(void *dst, const void *src, size_t n, hipStream_t st) {
        if (n == 0) return PS2_OK;
        if (n < kDirect) {
            PS2_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, st));
            PS2_HIP(hipStreamSynchronize(st));
            return PS2_OK;
        }
        std::lock_guard<std::mutex> lk(mu_);
        int rc;
        if ((rc = ensure())) return rc;
        if ((rc = h2d_chunks(dst, src, n, st))) (void)hipStreamSynchronize(st);
        return rc;
    }

  private:
    int d2h_chunks(void *dst, const void *src, size_t n, hipStream_t st) {
        int rc;
        Events ev;
        if ((rc = ev.create())) return rc;
        const size_t nk = (n + kChunk - 1) / kChunk;
        auto issue = [&](size_t k) -> int {
            const size_t off = k * kChunk, len = std::min(kChunk, n - off);
            PS2_HIP(hipMemcpyAsync(buf_[k & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, st));
            PS2_HIP(hipEventRecord(ev.e[k & 1], st));
            return PS2_OK;
        };
        if ((rc = issue(0))) return rc;
        for (size_t k = 0; k < nk; ++k) {
            // the other buffer held chunk k-1, already copied out below
            if (k + 1 < nk && (rc = issue(k + 1))) return rc;
            PS2_HIP(hipEventSynchronize(ev.e[k & 1]));
            const size_t off = k * kChunk;
            host_copy((char *)dst + off, buf_[k & 1], std::min(kChunk, n - off));
        }
        return PS2_OK;
    }

    int h2d_chunks(void *dst, const void *src, size_t n, hipStream_t st) {
        int rc;
        Events ev;
        if ((rc = ev.create())) return rc;
        const size_t nk = (n + kChunk - 1) / kChunk;
        for (size_t k = 0; k < nk; ++k) {
            // this buffer's previous DMA (chunk k-2) must have left it
            if (k >= 2) PS2_HIP(hipEventSynchronize(ev.e[k & 1]));
            const size_t off = k * kChunk, len = std::min(kChunk, n - off);
            host_copy(buf_[k & 1], (const char *)src + off, len);
            PS2_HIP(hipMemcpyAsync((char *)dst + off, buf_[k & 1], len, hipMemcpyHostToDevice, st));
            PS2_HIP(hipEventRecord(ev.e[k & 1], st));
        }
        PS2_HIP(hipStreamSynchronize(st));
        return PS2_OK;
    }

    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        int create() {
            for (auto &x : e) PS2_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
            return PS2_OK;
        }
        ~Events() {
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    };

    int ensure() {
        AllocGuard g(capture_mutex());
        for (auto &b : buf_) {
            if (b) continue;
            hipError_t e = hipHostMalloc(&b, kChunk, hipHostMallocPortable);
            if (e != hipSuccess) {
                b = nullptr;
                set_error(std::string("pinned staging allocation failed: ") + hipGetErrorString(e));
                return PS2_E_NOMEM;
            }
        }
        return PS2_OK;
    }

    static void host_copy(void *dst, const void *src, size_t n) {
        constexpr size_t kSlice = (size_t)4 << 20;
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>({8, hw, (n + kSlice - 1) / kSlice});
        if (nt <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        std::vector<std::thread> th;
        th.reserve(nt);
        for (size_t t = 0; t < nt; ++t) {
            const size_t a = n * t / nt, b = n * (t + 1) / nt;
            th.emplace_back([=]() { std::memcpy((char *)dst + a, (const char *)src + a, b - a); });
        }
        for (auto &x : th) x.join();
    }

    std::mutex mu_;
    void *buf_[2] = {nullptr, nullptr};
};

PinnedStaging &staging() {
    static PinnedStaging *pool = new PinnedStaging();   // process lifetime
    return *pool;
}

int staging_h2d(void *dst_device, const void *src_host, size_t bytes, hipStream_t stream) {
    return staging().h2d(dst_device, src_host, bytes, stream);
}

int staging_d2h(void *dst_host, const void *src_device, size_t bytes, hipStream_t stream) {
    return staging().d2h(dst_host, src_device, bytes, stream);
}

}  // namespace ps2

using namespace ps2;

// CLAP strip segmentation of a session (clap_segments)
struct ClapSegs {
    int nseg, wseg, nplanes;
};

constexpr int CLAP_TEAMS = 2;   // teams (interferograms) per CLAP wavefront

// The loop kernels of one gamma iteration in launch order, then the outputs.
// A recorded launch (below) carries its stage, so that the launches of
// several patches can be merged stage by stage.
enum Stage {
    ST_PREPARE, ST_FILL, ST_CLAP_A, ST_CLAP_B, ST_SCAN, ST_RESCORE, ST_REFINE, ST_REDECIDE, ST_POST, ST_CONTROL,
    ST_WEIGHTS, ST_OUTPUTS, ST_COUNT
};

// A loop-kernel launch kept for later: the kernel (its multi-patch
// instantiation), grid and workgroup size, the argument struct's bytes and
// the layout of that kernel's MultiPack (its size and where the ranges and
// the count sit; the structs are at multiples of the struct size).
struct LaunchRec {
    const void *fn;
    uint32_t blocks, threads;
    int stage;
    std::vector<char> args;
    size_t pack_size, range_off, n_off;
};

struct ps2_session;
// drops the merged-launch graphs that include session s (its destructor)
void batch_forget(const ps2_session *s);

struct ps2_session {
    std::mutex mu;
    int form = PS2_FORM_FP32;
    int device = 0;
    int verbose = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    // sizes / parameters
    int64_t P = 0, I = 0, ni = 0, nj = 0, C = 0;
    int tn = 0, T = 0, rowinv = 0;
    int nwi = 0, nwj = 0;
    bool windows = false;
    // CLAP window: n_win x n_win in an n_ex-point transform, step n_inc;
    // gclap: any other window than 24 of 32 (the generic kernels)
    int n_win = CLAP_WIN, n_ex = CLAP_FFT, n_inc = CLAP_INC;
    bool gclap = false;
    DBuf gpart, gweight;
    ClapSegs segs{1, 1, 4};   // set with the device state
    // tiles (model rules, or the autotuner's choice): fill workgroup cells,
    // FP32 scan unit-step source, CLAP segmentation above
    int fill_cb = 128;
    int scan_table = 1;
    int tile_source = PS2_TILES_MODEL;
    bool autotune = false;
    double alpha = 1.0, beta = 0.3, nr_max_nz_ix = 1.0, conv = 1e-4;
    int64_t lct = 31, max_iter = 25;
    int weighting = PS2_WEIGHT_PSQUARE;
    float rescore_bound = 0.0f;
    std::vector<int64_t> order;  // sorted position k -> caller row
    std::vector<double> nr_base;
    HistRule hist_rule;
    // device state
    DBuf phnm, bp, bp64, amp, grid, filt, part, lp, rng, wt, offset;
    // FP32 form: the f64 baselines as bp (f32) + the int32 residuals bplo
    // (exact, Bp64), or, where some value does not split, the f64 plane bp64
    DBuf bplo;
    // outputs in the caller's layouts (row order), written by k_final_outputs
    // at the end of every loop (and by the fit step)
    DBuf o_weight, o_patch, o_res, o_vec, o_tix, o_grid;
    DBuf d_order, cell_start, cell_of, tix, rescore_list, rescore_count, hist, mids;
    DBuf flags;   // FP32 form: rows the fused small-patch fit marks for the FP64 re-decision
    DBuf recheck_list;   // FP32 form: rows k_refine queues for k_redecide (count in rescore_count[1])
    DBuf K, Kprev, W, Wprev, W0, Cc, coh, cohprev, nopt, prand_hi, rms_part;
    int n_rms_blocks = 0;   // row groups of the loop's post step, one RMS partial each
    int n_cu = 256;         // compute units of the device
    int clap_slots = 1;     // CLAP strip wavefronts resident on the device
    // device-resident loop: control block, gate of the loop kernels (set
    // while a loop is being launched), one iteration captured as a graph
    DBuf ctl, nr0;   // control block; the base Nr histogram (its starting Nr)
    HostPinned h_ctl;
    // coherent pinned copy of the control block that a captured loop's last
    // kernel writes (k_publish_ctl), and its device address
    HostPinned h_ctl_pub;
    LoopCtl *d_ctl_pub = nullptr;
    const int32_t *gate = nullptr;
    // while set, the launch functions instantiated for several patches record
    // their launches here instead of launching them
    std::vector<LaunchRec> *rec = nullptr;
    CtlParams ctl_params{};
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool no_graph = false;   // a capture failed: launch step-wise
    // per-kernel timing (ps2_hip_session_set_timing)
    bool timing = false;
    std::vector<hipEvent_t> tev;
    struct TimerRec {
        int slot, start;   // kernel slot, index of the start event in tev
        int64_t iter;      // loop iteration (1-based) the launch belongs to, 0 outside one
        bool launched;     // a kernel of the slot was launched with the pair
    };
    std::vector<TimerRec> tpending;
    int tused = 0;
    int64_t timer_iter = 0;
    int cur_rec = -1;   // tpending entry the next launches are timed by, -1 none
    double slot_ms[PS2_NUM_KERNEL_SLOTS] = {};
    int64_t slot_launches[PS2_NUM_KERNEL_SLOTS] = {};
    // results of the last run
    bool ran = false;
    bool loop_outputs = false;   // o_* hold the last loop's outputs (not a fit step's)
    int64_t i_loop = 0;
    double gamma_change_save = 0.0, loop_seconds = 0.0;
    std::vector<double> nr;
    int64_t n_rescored = 0, n_rechecked = 0;

    ~ps2_session() {
        batch_forget(this);
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
            AllocGuard g(capture_mutex());
            if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
            if (graph) (void)hipGraphDestroy(graph);
            for (auto &e : ev)
                if (e) (void)hipEventDestroy(e);
            for (auto &e : tev)
                if (e) (void)hipEventDestroy(e);
            (void)hipStreamDestroy(stream);
        }
    }
    size_t rsz() const { return form == PS2_FORM_FP64 ? sizeof(double) : sizeof(float); }
};

// Launch of a loop kernel k with argument struct a: for one patch (M false)
// on the session's stream now; for several (M true) recorded in s->rec.
template <bool M, typename A>
void kgo(ps2_session *s, void (*k)(KArg<M, A>), dim3 g, dim3 b, const A &a, int stage) {
    if constexpr (M) {
        static_assert(sizeof(MultiPack<A>) <= 4096, "kernel argument pack too large");
        LaunchRec r{reinterpret_cast<const void *>(k), g.x, b.x, stage, {}, sizeof(MultiPack<A>),
                    offsetof(MultiPack<A>, range), offsetof(MultiPack<A>, n)};
        r.args.assign(reinterpret_cast<const char *>(&a), reinterpret_cast<const char *>(&a) + sizeof(A));
        s->rec->push_back(std::move(r));
    } else if (s->cur_rec >= 0) {
        auto &tr = s->tpending[s->cur_rec];
        hipExtLaunchKernelGGL(k, g, b, 0, s->stream, tr.launched ? nullptr : s->tev[tr.start],
                              s->tev[tr.start + 1], 0, a);
        tr.launched = true;
    } else {
        hipLaunchKernelGGL(k, g, b, 0, s->stream, a);
    }
}

// Patches of at most 65,536 rows (at most one 256-row group per CU) run the
// fused k_fit_tail (re-score + refine + post step) in their loops.
static inline bool small_patch(const ps2_session *s) { return s->P <= 65536; }

// FP32 trial scan: unit steps e^{-i phi} from a per-(pixel, ifg) table, or
// computed in the scan from the pixel's bperp.  Computing them is the better
// choice where the scan is bound by HBM bandwidth: 8-trial chunks (fewer than
// 8 positive trials) over large patches with per-pixel bperp.  Invariant
// bperp rows keep a one-row table.
static inline bool scan_step_table_rule(const ps2_session *s) {
    return s->form == PS2_FORM_FP32 && (s->rowinv || s->tn >= 8 || s->P < 65536);
}
static inline bool scan_step_table(const ps2_session *s) {
    return s->form == PS2_FORM_FP32 && (s->scan_table != 0 || s->rowinv || s->tn >= 8);
}

// Host -> device copy on the session's stream, complete on return (the
// library issues nothing on the legacy default stream, which would wait on
// other threads' streams, including one being captured).
static hipError_t h2d_now(ps2_session *s, void *dst, const void *src, size_t n) {
    hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e;
}

namespace {

std::mutex g_init_mu;

int validate(const ps2_patch_in *in) {
    if (!in) return arg_error("ps2_patch_in is NULL");
    if (in->n_ps <= 0) return arg_error("ps1.mat missing valid n_ps");
    if (in->n_ifg <= 0) return arg_error("n_ifg must be positive");
    if (in->n_i <= 0 || in->n_j <= 0) return arg_error("grid extent must be positive");
    if (!in->ph_nm || !in->grid_lin || !in->d_a || !in->low_pass || !in->trial_values ||
        !in->nr_base || !in->coh_bins || !in->bperp_row)
        return arg_error("a required input array is NULL");
    if (!in->row_invariant && !in->bperp_mat)
        return arg_error("bp1.bperp_mat is required for non-invariant stage-2 baselines");
    if (in->n_win_ex < 1 || in->n_win_ex > GCLAP_MAX || in->clap_window < 1 || in->clap_window > in->n_win_ex)
        return arg_error("CLAP window outside the HIP backend's range: n_win_ex=" + std::to_string(in->n_win_ex) +
                         " clap_window=" + std::to_string(in->clap_window) + " (1 <= clap_window <= n_win_ex <= " +
                         std::to_string(GCLAP_MAX) + ", i.e. clap_win up to 64)");
    if (in->n_trial < 1 || (in->n_trial % 2) != 1)
        return arg_error("trial_values must hold 2*tn+1 values");
    const int64_t tn = in->n_trial / 2;
    for (int64_t t = 0; t < in->n_trial; ++t)
        if (in->trial_values[t] != (double)(t - tn))
            return arg_error("trial_values must be the consecutive integers -tn..tn (trial_values(), native_stage2.rs:1305-1308)");
    if (tn > 4096) return arg_error("trial count too large");
    // size limits first (before any per-pixel read), as divisions so that no
    // product of the caller's extents can overflow; every product stays below
    // 2^31, so pixel and cell counts fit the int32 counting sort
    const int64_t lim = ((int64_t)1 << 31) - 1;
    if (in->n_i > lim / in->n_j || in->n_ps > lim / in->n_ifg || in->n_i * in->n_j > lim / in->n_ifg)
        return arg_error("patch too large for one session (n_ps*n_ifg and n_i*n_j*n_ifg must be below 2^31)");
    const int64_t C = in->n_i * in->n_j;
    for (int64_t p = 0; p < in->n_ps; ++p)
        if (in->grid_lin[p] < 0 || in->grid_lin[p] >= C)
            return arg_error("grid_lin[" + std::to_string(p) + "] outside the n_i x n_j grid");
    if (in->weighting == PS2_WEIGHT_SNR && !in->amp) return arg_error("amp is required for SNR weighting");
    return PS2_OK;
}

template <typename R>
int upload_inputs(ps2_session *s, const ps2_patch_in *in) {
    const int64_t P = s->P, I = s->I;
    // cell-sorted pixel order (stable counting sort of grid_lin)
    std::vector<int32_t> cstart(s->C + 1, 0);
    for (int64_t p = 0; p < P; ++p) cstart[in->grid_lin[p] + 1]++;
    for (int64_t c = 0; c < s->C; ++c) cstart[c + 1] += cstart[c];
    std::vector<int32_t> fillpos(cstart.begin(), cstart.end() - 1);
    s->order.assign(P, 0);
    std::vector<int32_t> cell_of(P);
    for (int64_t p = 0; p < P; ++p) {
        const int64_t c = in->grid_lin[p];
        const int32_t k = fillpos[c]++;
        s->order[k] = p;
        cell_of[k] = (int32_t)c;
    }
    int rc;
    if ((rc = s->cell_start.alloc(sizeof(int32_t) * (s->C + 1)))) return rc;
    if ((rc = s->cell_of.alloc(sizeof(int32_t) * P))) return rc;
    if ((rc = s->d_order.alloc(sizeof(int64_t) * P))) return rc;
    PS2_HIP(h2d_now(s, s->cell_start.p, cstart.data(), sizeof(int32_t) * (s->C + 1)));
    PS2_HIP(h2d_now(s, s->cell_of.p, cell_of.data(), sizeof(int32_t) * P));
    PS2_HIP(h2d_now(s, s->d_order.p, s->order.data(), sizeof(int64_t) * P));

    // the row-major inputs go up as they are; the device permutes them into
    // sorted ifg-major planes (one staging buffer, reused)
    DBuf stage;
    const unsigned nb = blocks_for(P * I, 256);
    if ((rc = stage.alloc(sizeof(double) * P * I))) return rc;
    {   // ph_nm planes [I][P]
        if ((rc = s->phnm.alloc(sizeof(cplx<R>) * P * I))) return rc;
        if ((rc = staging().h2d(stage.p, in->ph_nm, sizeof(float) * 2 * P * I, s->stream))) return rc;
        hipLaunchKernelGGL(k_planes_cplx<R>, dim3(nb), dim3(256), 0, s->stream, stage.as<float2>(),
                           s->d_order.as<int64_t>(), s->phnm.as<cplx<R>>(), P, (int)I);
        PS2_LAUNCHED();
    }
    {   // baselines: R copy for the bulk; the FP32 form's f64 values for the
        // FP64 re-score as an exact split (Bp64: 4 more bytes per element) or,
        // where some value does not split, as an f64 plane (8)
        const int64_t n = s->rowinv ? I : P * I;
        if ((rc = s->bp.alloc(sizeof(R) * n))) return rc;
        if (s->rowinv) {
            std::vector<R> h((size_t)n);
            for (int64_t f = 0; f < I; ++f) h[f] = (R)in->bperp_row[f];
            PS2_HIP(h2d_now(s, s->bp.p, h.data(), sizeof(R) * n));
            if (sizeof(R) == 4) {
                if ((rc = s->bp64.alloc(sizeof(double) * n))) return rc;
                PS2_HIP(h2d_now(s, s->bp64.p, in->bperp_row, sizeof(double) * n));
            }
        } else {
            PS2_HIP(hipStreamSynchronize(s->stream));  // the staging buffer is free again
            if ((rc = staging().h2d(stage.p, in->bperp_mat, sizeof(double) * P * I, s->stream))) return rc;
            DBuf inexact;
            if (sizeof(R) == 4) {
                if ((rc = s->bplo.alloc(sizeof(int32_t) * n))) return rc;
                if ((rc = inexact.alloc(sizeof(uint32_t)))) return rc;
                PS2_HIP(hipMemsetAsync(inexact.p, 0, sizeof(uint32_t), s->stream));
            }
            hipLaunchKernelGGL(k_planes_real<R>, dim3(nb), dim3(256), 0, s->stream, stage.as<double>(),
                               s->d_order.as<int64_t>(), s->bp.as<R>(), (double *)nullptr,
                               sizeof(R) == 4 ? s->bplo.as<int32_t>() : nullptr,
                               sizeof(R) == 4 ? inexact.as<uint32_t>() : nullptr, P, (int)I);
            PS2_LAUNCHED();
            if (sizeof(R) == 4) {
                uint32_t n_inexact = 0;
                PS2_HIP(hipMemcpyAsync(&n_inexact, inexact.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
                PS2_HIP(hipStreamSynchronize(s->stream));
                if (n_inexact) {   // the f64 plane instead
                    if ((rc = s->bplo.alloc(0))) return rc;
                    if ((rc = s->bp64.alloc(sizeof(double) * n))) return rc;
                    hipLaunchKernelGGL(k_planes_real<R>, dim3(nb), dim3(256), 0, s->stream, stage.as<double>(),
                                       s->d_order.as<int64_t>(), s->bp.as<R>(), s->bp64.as<double>(),
                                       (int32_t *)nullptr, (uint32_t *)nullptr, P, (int)I);
                    PS2_LAUNCHED();
                }
            }
        }
    }
    if (in->amp && s->weighting == PS2_WEIGHT_SNR) {
        if ((rc = s->amp.alloc(sizeof(float) * P * I))) return rc;
        PS2_HIP(hipStreamSynchronize(s->stream));
        if ((rc = staging().h2d(stage.p, in->amp, sizeof(float) * P * I, s->stream))) return rc;
        hipLaunchKernelGGL(k_planes_f32, dim3(nb), dim3(256), 0, s->stream, stage.as<float>(),
                           s->d_order.as<int64_t>(), s->amp.as<float>(), P, (int)I);
        PS2_LAUNCHED();
    }
    PS2_HIP(hipStreamSynchronize(s->stream));  // before the staging buffer is freed
    {   // initial weighting 1/D_A (0 where D_A == 0), native_stage2.rs:242-246
        std::vector<double> w0(P);
        for (int64_t k = 0; k < P; ++k) {
            const double da = in->d_a[s->order[k]];
            w0[k] = da != 0.0 ? 1.0 / da : 0.0;
        }
        if ((rc = s->W0.alloc(sizeof(double) * P))) return rc;
        PS2_HIP(h2d_now(s, s->W0.p, w0.data(), sizeof(double) * P));
    }
    if (!s->gclap) {   // CLAP constants: low_pass (the window weight is computed in the kernel)
        // stored transposed: column c's 32 values are contiguous (the CLAP
        // kernel's lane c reads them with vector loads)
        std::vector<R> lp(CLAP_FFT * CLAP_FFT);
        for (int r = 0; r < CLAP_FFT; ++r)
            for (int c = 0; c < CLAP_FFT; ++c) lp[c * CLAP_FFT + r] = (R)in->low_pass[r * CLAP_FFT + c];
        if ((rc = s->lp.alloc(sizeof(R) * lp.size()))) return rc;
        PS2_HIP(h2d_now(s, s->lp.p, lp.data(), sizeof(R) * lp.size()));
    } else {   // any other window: low_pass in its own layout, window weight x 1/n^2
        const int n = s->n_ex, nw = s->n_win;
        std::vector<R> lp((size_t)n * n);
        for (size_t k = 0; k < lp.size(); ++k) lp[k] = (R)in->low_pass[k];
        if ((rc = s->lp.alloc(sizeof(R) * lp.size()))) return rc;
        PS2_HIP(h2d_now(s, s->lp.p, lp.data(), sizeof(R) * lp.size()));
        // clap_window_weight (native_stage2.rs:1014-1037): r + c + 1e-6 on a
        // half x half quadrant, mirrored; an odd window's last row and column
        // stay 0
        const int half = nw / 2;
        std::vector<double> wt((size_t)nw * nw, 0.0), top((size_t)half * nw, 0.0);
        for (int r = 0; r < half; ++r)
            for (int c = 0; c < half; ++c) {
                top[(size_t)r * nw + c] = (double)r + (double)c + 1.0e-6;
                top[(size_t)r * nw + half + c] = (double)r + (double)(half - 1 - c) + 1.0e-6;
            }
        for (int r = 0; r < half; ++r)
            for (int c = 0; c < nw; ++c) {
                wt[(size_t)r * nw + c] = top[(size_t)r * nw + c];
                wt[(size_t)(half + r) * nw + c] = top[(size_t)(half - 1 - r) * nw + c];
            }
        const double inv_scale = 1.0 / (double)(n * n);
        for (double &v : wt) v = v * inv_scale;
        if ((rc = s->gweight.alloc(sizeof(double) * wt.size()))) return rc;
        PS2_HIP(h2d_now(s, s->gweight.p, wt.data(), sizeof(double) * wt.size()));
    }
    return PS2_OK;
}

// Window segments of the CLAP strips: each window row is split into nseg
// segments of wseg windows; one strip (a wavefront) walks one segment of one
// window row for one pair of interferograms.  The strips run in rounds of
// `slots` (CUs x strip wavefronts resident per CU), and a strip's time is
// about wseg windows plus a quarter window of start and flush, so the launch
// takes ~ ceil(strips / slots) x (wseg + 1/4) window times (per XCD: strip u
// runs on XCD u % 8).  Each segment also adds 18 overlap columns per window
// row to the partial planes that the combine reads: 2-7.5e-5 x I x nwi
// window times per segment measured on the combine (500k x 60, 4M x 120,
// 1M x 200); the model takes 3e-5.  nseg minimises the sum: a last round that holds a few strips
// costs as much as a full one.  Segment s covers columns
// [6 wseg s, 6 wseg (s + 1) + 18), so it overlaps the next ceil(3 / wseg)
// segments; segments write their partial sums into nplanes = 1 + ceil(3 / wseg)
// planes in turn (2 from 3 windows per segment on), so no two overlapping
// segments share a plane.
// Segmentations ranked by the model's cost (cheapest first; ties keep the
// longer segments).  The model's choice is the first; the autotuner times the
// first few.
static std::vector<ClapSegs> clap_segment_candidates(int nwi, int nwj, int ngroups, int I, int slots) {
    const int64_t rows = std::max<int64_t>(1, (int64_t)nwi * ngroups);
    const double seg_cost = 3e-5 * (double)I * (double)nwi;
    const int64_t xslots = std::max<int64_t>(1, slots / PS2_NUM_XCD);
    const int nw = std::max(1, nwj);
    std::vector<std::pair<double, ClapSegs>> c;
    for (int nseg = 1; nseg <= nw; ++nseg) {
        const int wseg = (nw + nseg - 1) / nseg;
        if ((nw + wseg - 1) / wseg != nseg) continue;   // same segments as a smaller nseg
        const int64_t per_xcd = (rows * nseg + PS2_NUM_XCD - 1) / PS2_NUM_XCD;
        const double cost = (double)((per_xcd + xslots - 1) / xslots) * (wseg + 0.25) + seg_cost * nseg;
        c.push_back({cost, ClapSegs{nseg, wseg, 1 + (3 + wseg - 1) / wseg}});
    }
    std::stable_sort(c.begin(), c.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    std::vector<ClapSegs> out;
    for (const auto &x : c) out.push_back(x.second);
    if (out.empty()) out.push_back(ClapSegs{1, nw, 1 + (3 + nw - 1) / nw});
    return out;
}

static ClapSegs clap_segments(int nwi, int nwj, int ngroups, int I, int slots) {
    return clap_segment_candidates(nwi, nwj, ngroups, I, slots).front();
}

template <typename R>
int alloc_state(ps2_session *s) {
    const int64_t P = s->P, I = s->I, C = s->C;
    int rc;
    if ((rc = s->grid.alloc(sizeof(cplx<R>) * C * I))) return rc;
    if ((rc = s->filt.alloc(sizeof(cplx<R>) * C * I))) return rc;
    PS2_HIP(hipDeviceGetAttribute(&s->n_cu, hipDeviceAttributeMultiprocessorCount, s->device));
    s->n_cu = std::max(1, s->n_cu);
    if (s->windows && s->gclap) {
        const double bytes = (double)sizeof(cplx<R>) * (double)s->nwi * (double)s->nwj * (double)s->n_win *
                             (double)s->n_win * (double)I;
        if (bytes > 32e9)
            return arg_error("CLAP window of " + std::to_string(s->n_win) + " cells on a " + std::to_string(s->ni) +
                             "x" + std::to_string(s->nj) + " grid needs " + std::to_string(bytes / 1e9) +
                             " GB of window outputs (limit 32 GB per session)");
        if ((rc = s->gpart.alloc((size_t)bytes))) return rc;
    } else if (s->windows) {
        // strip slots: CUs x CLAP wavefronts resident per CU (register bound)
        const int n_cu = s->n_cu;
        int per_cu = 0;
        PS2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&k_clap_strip<false, R, CLAP_TEAMS, true>), 32 * CLAP_TEAMS, 0));
        s->clap_slots = std::max(1, n_cu) * std::max(1, per_cu);
        s->segs = clap_segments(s->nwi, s->nwj, (int)((I + CLAP_TEAMS - 1) / CLAP_TEAMS), (int)I, s->clap_slots);
        // the autotuner tries segmentations with up to 4 planes; one
        // interferogram's planes stay below 2^31 entries (the combine's
        // offsets are 32-bit)
        const int planes = s->autotune ? 4 : s->segs.nplanes;
        if ((double)s->nwi * 4.0 * CLAP_WIN * (double)s->nj >= 2147483647.0)
            return arg_error("grid too large for the CLAP partial planes (n_i x n_j x 16 must stay below 2^31)");
        if ((rc = s->part.alloc(sizeof(cplx<R>) * (size_t)s->nwi * planes * CLAP_WIN * s->nj * I)))
            return rc;
    }
    s->fill_cb = (double)P / (double)std::max<int64_t>(C, 1) < 12.0 ? 128 : 64;
    s->scan_table = scan_step_table_rule(s) ? 1 : 0;
    if ((rc = s->rng.alloc(sizeof(R) * P))) return rc;
    // the FP32 unit-step table (one row with invariant bperp rows), unless
    // the scan computes the steps itself; the chunked scan (48 or more
    // positive trials) always reads it
    const bool table = sizeof(R) == 4 && (scan_step_table(s) || s->tn >= 48 || s->autotune);
    if (table)
        if ((rc = s->wt.alloc(sizeof(cplx<R>) * (s->rowinv ? I : P * I)))) return rc;
    hipLaunchKernelGGL(k_trial_setup<R>, dim3(blocks_for(P, 256)), dim3(256), 0, s->stream,
                       s->bp.as<R>(), s->rowinv, P, (int)I, s->rng.as<R>(),
                       table ? s->wt.as<cplx<R>>() : nullptr);
    PS2_LAUNCHED();
    if ((rc = s->offset.alloc(sizeof(cplx<R>) * P))) return rc;
    if ((rc = s->tix.alloc(sizeof(int32_t) * P))) return rc;
    if ((rc = s->rescore_list.alloc(sizeof(int32_t) * P))) return rc;
    if ((rc = s->rescore_count.alloc(sizeof(uint32_t) * 4))) return rc;
    if (sizeof(R) == 4) {
        if ((rc = s->flags.alloc(std::max<int64_t>(P, 1)))) return rc;
        if ((rc = s->recheck_list.alloc(sizeof(int32_t) * std::max<int64_t>(P, 1)))) return rc;
    }
    if ((rc = s->hist.alloc(sizeof(uint32_t) * 128))) return rc;
    for (DBuf *b : {&s->K, &s->Kprev, &s->W, &s->Wprev, &s->Cc, &s->coh, &s->cohprev, &s->nopt})
        if ((rc = b->alloc(sizeof(double) * P))) return rc;
    if ((rc = s->prand_hi.alloc(sizeof(double) * 2048))) return rc;
    if ((rc = s->ctl.alloc(sizeof(LoopCtl)))) return rc;
    if ((rc = s->h_ctl.alloc(sizeof(LoopCtl)))) return rc;
    if ((rc = s->h_ctl_pub.alloc(sizeof(LoopCtl), hipHostMallocCoherent))) return rc;
    PS2_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&s->d_ctl_pub), s->h_ctl_pub.p, 0));
    if ((rc = s->nr0.alloc(sizeof(double) * CTL_BINS))) return rc;
    PS2_HIP(hipMemcpyAsync(s->nr0.p, s->nr_base.data(), sizeof(double) * CTL_BINS, hipMemcpyHostToDevice,
                           s->stream));
    PS2_HIP(hipStreamSynchronize(s->stream));
    // one partial per row group of the loop's post step: 64 rows in the FP32
    // form's small-patch fit (k_fit_small), else 256 (k_post, k_fit_tail)
    s->n_rms_blocks = (int)blocks_for(P, s->form == PS2_FORM_FP32 && small_patch(s) ? SMALL_GROUP : 256);
    if ((rc = s->rms_part.alloc(sizeof(double) * s->n_rms_blocks))) return rc;
    {   // control-kernel constants: gausswin(7, 2.5) taps and their sum (native_stage2.rs:1685-1699)
        CtlParams &cp = s->ctl_params;
        cp.P = P;
        cp.n_rms_blocks = s->n_rms_blocks;
        cp.max_iter = (int)std::min<int64_t>(std::max<int64_t>(1, s->max_iter), 1 << 30);
        cp.psquare = s->weighting == PS2_WEIGHT_PSQUARE ? 1 : 0;
        cp.lct = (int)std::min<int64_t>(std::max<int64_t>(s->lct, -1), 1 << 30);
        cp.conv = s->conv;
        cp.nr_max_nz_ix = s->nr_max_nz_ix;
        const double sd = 6.0 / 5.0;
        cp.gsum = 0.0;
        for (int k = 0; k < 7; ++k) {
            const double x = (double)k - 3.0;
            cp.g7[k] = std::exp(-0.5 * (x / sd) * (x / sd));
        }
        for (int k = 0; k < 7; ++k) cp.gsum += cp.g7[k];
    }
    // output buffers in the caller's layouts
    if ((rc = s->o_weight.alloc(sizeof(float2) * P * I))) return rc;
    if ((rc = s->o_patch.alloc(sizeof(float2) * P * I))) return rc;
    if ((rc = s->o_res.alloc(sizeof(float) * P * I))) return rc;
    if ((rc = s->o_vec.alloc(sizeof(double) * 5 * P))) return rc;
    if ((rc = s->o_tix.alloc(sizeof(int32_t) * P))) return rc;
    if ((rc = s->o_grid.alloc(sizeof(float2) * C * I))) return rc;
    if ((rc = s->mids.alloc(sizeof(double) * 128))) return rc;
    std::vector<double> mids(128, 0.0);
    for (size_t k = 0; k < s->hist_rule.mids.size() && k < 128; ++k) mids[k] = s->hist_rule.mids[k];
    PS2_HIP(h2d_now(s, s->mids.p, mids.data(), sizeof(double) * 128));
    return PS2_OK;
}

// --- per-kernel timing ---------------------------------------------------------

enum Slot { SLOT_FILL = 0, SLOT_CLAP_STRIP, SLOT_CLAP_COMBINE, SLOT_SCAN, SLOT_RESCORE, SLOT_REFINE,
            SLOT_POST, SLOT_WEIGHTS, SLOT_OUTPUTS };

// The kernels launched in a SlotTimer's scope carry its event pair in their
// dispatches (kgo): the start event takes the first kernel's start time, the
// stop event the last kernel's end time.  No marker packet goes between the
// kernels, so they run back to back as in a graph.  A scope that launched
// nothing leaves no record.
struct SlotTimer {
    ps2_session *s;
    bool on;
    SlotTimer(ps2_session *s_, int slot)
        : s(s_), on(s_->timing && s_->cur_rec < 0 && s_->tused + 2 <= (int)s_->tev.size()) {
        if (!on) return;
        s->tpending.push_back({slot, s->tused, s->timer_iter, false});
        s->cur_rec = (int)s->tpending.size() - 1;
        s->tused += 2;
    }
    ~SlotTimer() {
        if (!on) return;
        if (!s->tpending[s->cur_rec].launched) s->tpending.pop_back();
        s->cur_rec = -1;
    }
};

// Accumulates the recorded pairs of iterations up to last_iter (launches
// outside an iteration always); the stream must be synchronised.
void collect_timers(ps2_session *s, int64_t last_iter = INT64_MAX) {
    for (auto &pr : s->tpending) {
        float ms = 0.0f;
        if (pr.iter > last_iter || !pr.launched) continue;
        if (hipEventElapsedTime(&ms, s->tev[pr.start], s->tev[pr.start + 1]) == hipSuccess) {
            s->slot_ms[pr.slot] += ms;
            s->slot_launches[pr.slot] += 1;
        }
    }
    s->tpending.clear();
    s->tused = 0;
    s->cur_rec = -1;
}

// --- kernel launch helpers ---------------------------------------------------


template <typename R, bool M = false>
int launch_fill_grid(ps2_session *s, const double *K, const double *W, double *Kprev = nullptr,
                     double *Wprev = nullptr, bool psquare_in_fill = false) {
    const int64_t n = s->C * s->I;
    SlotTimer tm(s, SLOT_FILL);
    // inside the loop with P-square weighting, iterations after the first
    // compute the weights from coh and the P-square table (FillWeights)
    const FillWeights fw{W, psquare_in_fill ? s->coh.as<double>() : nullptr, s->prand_hi.as<double>(),
                         psquare_in_fill ? &s->ctl.as<LoopCtl>()->i_loop : nullptr, CTL_HI};
    if (sizeof(R) == 4) {
        // 128 cells x 2 planes per workgroup below 12 pixels per cell, else
        // 64 cells x 4 planes (measured: 128 x 2 fastest of 32 x 8, 64 x 4,
        // 128 x 2, 256 x 1 at 500k x 60 = 5.6 pixels per cell; 64 x 4 6%
        // faster at 4M x 120 = 25 per cell, where reading K and W once per 4
        // planes pays).  The LDS chunk is the smallest of 256 / 512 / 1024
        // pixels covering the mean pixel count of the workgroup's cells (larger
        // ranges take several chunks: LDS beyond 16 KB per workgroup costs more
        // in occupancy than the extra passes).
        const double density = (double)s->P / (double)std::max<int64_t>(s->C, 1);
        // the cells per workgroup come from the session's tiles
        auto launch = [&](auto cb_tag) {
            constexpr int CB = decltype(cb_tag)::value;
            const double mean = density * CB;
            const int64_t ncg = blocks_for(s->C, CB);
            const int npg = (int)blocks_for(s->I, 256 / CB);
            const dim3 grid((unsigned)(blocks_for(ncg, PS2_NUM_XCD) * PS2_NUM_XCD * npg));
            const FillStagedArgs<R> a{s->phnm.as<cplx<R>>(), s->bp.as<R>(), s->rowinv, K, fw,
                                      s->cell_start.as<int32_t>(), s->grid.as<cplx<R>>(), s->C, (int)s->I,
                                      s->P, npg, Kprev, Wprev, s->gate};
            auto go = [&](auto kern) { kgo<M>(s, kern, grid, dim3(256), a, ST_FILL); };
            if (mean <= 256.0)
                go(k_fill_grid_staged<M, R, CB, 256>);
            else if (mean <= 512.0)
                go(k_fill_grid_staged<M, R, CB, 512>);
            else
                go(k_fill_grid_staged<M, R, CB, 1024>);
        };
        if (s->fill_cb == 128)
            launch(std::integral_constant<int, 128>{});
        else
            launch(std::integral_constant<int, 64>{});
    } else {
        const FillArgs<R> a{s->phnm.as<cplx<R>>(), s->bp.as<R>(), s->rowinv, K, fw,
                            s->cell_start.as<int32_t>(), s->grid.as<cplx<R>>(), s->C, (int)s->I, s->P,
                            Kprev, Wprev, s->gate};
        kgo<M>(s, k_fill_grid<M, R>, dim3(blocks_for(n, 256)), dim3(256), a, ST_FILL);
    }
    PS2_LAUNCHED();
    return PS2_OK;
}

template <typename R>
ClapParams<R> clap_params(ps2_session *s, int teams) {
    ClapParams<R> p;
    const double sd = 6.0 / 5.0;  // (7-1)/(2*2.5)
    for (int k = 0; k < 7; ++k) {
        const double x = (double)k - 3.0;
        p.k7[k] = (R)std::exp(-0.5 * (x / sd) * (x / sd));
    }
    p.alpha = (R)s->alpha;
    p.beta = (R)s->beta;
    p.alpha_is_one = std::fabs(s->alpha - 1.0) <= 2.220446049250313e-16;
    p.ni = (int)s->ni;
    p.nj = (int)s->nj;
    p.I = (int)s->I;
    p.nwi = s->nwi;
    p.nwj = s->nwj;
    p.ngroups = (int)((s->I + teams - 1) / teams);
    const ClapSegs cs = s->segs;
    p.wseg = cs.wseg;
    p.nseg = cs.nseg;
    p.nplanes = cs.nplanes;
    p.units = p.ngroups * p.nseg;
    p.upx = (p.units + PS2_NUM_XCD - 1) / PS2_NUM_XCD;
    return p;
}

template <typename R, bool M = false>
int launch_clap(ps2_session *s, bool unit) {
    const int64_t n = s->C * s->I;
    if (!s->windows) {   // empty CLAP: zero filtered grid (a kernel: this may be graph-captured)
        const int64_t words = (int64_t)(sizeof(cplx<R>) / 4) * n;
        kgo<M>(s, k_zero_u32<M>, dim3(blocks_for(words, 256)), dim3(256),
               ZeroArgs{s->filt.as<uint32_t>(), words, s->gate}, ST_CLAP_A);
        PS2_LAUNCHED();
        return PS2_OK;
    }
    if (s->gclap) {   // any other window than 24 of 32: the generic kernels
        GClapArgs<R> g{};
        g.grid = s->grid.as<cplx<R>>();
        g.gpart = s->gpart.as<cplx<R>>();
        g.lp = s->lp.as<R>();
        g.weight = s->gweight.as<double>();
        g.ni = (int)s->ni;
        g.nj = (int)s->nj;
        g.I = (int)s->I;
        g.nwi = s->nwi;
        g.nwj = s->nwj;
        g.n_win = s->n_win;
        g.n = s->n_ex;
        g.n_inc = s->n_inc;
        g.alpha = (R)s->alpha;
        g.beta = (R)s->beta;
        const double sd = 6.0 / 5.0;  // gausswin(7, 2.5): std (7-1)/(2*2.5)
        for (int k = 0; k < 7; ++k) {
            const double x = (double)k - 3.0;
            g.k7[k] = (R)std::exp(-0.5 * (x / sd) * (x / sd));
        }
        g.gate = s->gate;
        const bool one = std::fabs(s->alpha - 1.0) <= 2.220446049250313e-16;
        const dim3 gw((unsigned)((int64_t)s->nwi * s->nwj * s->I));
        {
            SlotTimer tm(s, SLOT_CLAP_STRIP);
            auto win = [&](auto nmax) {
                constexpr int NM = decltype(nmax)::value;
                if (one)
                    kgo<M>(s, k_gclap_window<M, R, NM, true>, gw, dim3(256), g, ST_CLAP_A);
                else
                    kgo<M>(s, k_gclap_window<M, R, NM, false>, gw, dim3(256), g, ST_CLAP_A);
            };
            if (s->n_ex <= 16)
                win(std::integral_constant<int, 16>{});
            else if (s->n_ex <= 32)
                win(std::integral_constant<int, 32>{});
            else
                win(std::integral_constant<int, 64>{});
            PS2_LAUNCHED();
        }
        SlotTimer tm(s, SLOT_CLAP_COMBINE);
        const GCombineArgs<R> ca{g, s->filt.as<cplx<R>>()};
        const dim3 gc(blocks_for(n, 256));
        if (unit)
            kgo<M>(s, k_gclap_combine<M, R, true>, gc, dim3(256), ca, ST_CLAP_B);
        else
            kgo<M>(s, k_gclap_combine<M, R, false>, gc, dim3(256), ca, ST_CLAP_B);
        PS2_LAUNCHED();
        return PS2_OK;
    }
    constexpr int TEAMS = CLAP_TEAMS;
    const ClapParams<R> prm = clap_params<R>(s, TEAMS);
    {
        SlotTimer tm(s, SLOT_CLAP_STRIP);
        const dim3 g((unsigned)s->nwi * prm.upx * PS2_NUM_XCD);
        const StripArgs<R> a{s->grid.as<cplx<R>>(), s->part.as<cplx<R>>(), s->lp.as<R>(), prm, s->gate};
        if (prm.alpha_is_one)
            kgo<M>(s, k_clap_strip<M, R, TEAMS, true>, g, dim3(32 * TEAMS), a, ST_CLAP_A);
        else
            kgo<M>(s, k_clap_strip<M, R, TEAMS, false>, g, dim3(32 * TEAMS), a, ST_CLAP_A);
        PS2_LAUNCHED();
    }
    SlotTimer tm(s, SLOT_CLAP_COMBINE);
    const CombineArgs<R> ca{s->part.as<cplx<R>>(), s->filt.as<cplx<R>>(), (int)s->ni, (int)s->nj, (int)s->I,
                            s->nwi, prm.nseg, prm.wseg, prm.nplanes, s->gate};
    auto combine = [&](auto kern) {
        const dim3 g(blocks_for(s->ni, COMB_TR) * blocks_for(s->nj, COMB_TC) * (unsigned)s->I);
        kgo<M>(s, kern, g, dim3(256), ca, ST_CLAP_B);
    };
    if (prm.nplanes <= 2)
        unit ? combine(k_clap_combine<M, R, true, 2>) : combine(k_clap_combine<M, R, false, 2>);
    else
        unit ? combine(k_clap_combine<M, R, true, 4>) : combine(k_clap_combine<M, R, false, 4>);
    PS2_LAUNCHED();
    return PS2_OK;
}

// The FP32 form's f64 baselines (Bp64); unused by the FP64 form.
static Bp64 bp64_of(const ps2_session *s) {
    if (s->form != PS2_FORM_FP32) return Bp64{nullptr, nullptr, nullptr};
    return Bp64{s->bp64.as<double>(), s->bp.as<float>(), s->bplo.as<int32_t>()};
}

FitParams fit_params(ps2_session *s) {
    FitParams f;
    f.P = s->P;
    f.C = s->C;
    f.I = (int)s->I;
    f.tn = s->tn;
    f.T = s->T;
    f.rowinv = s->rowinv;
    f.step_table = scan_step_table(s) ? 1 : 0;
    f.rescore_bound = s->rescore_bound;
    return f;
}

// positive trials (incl. 0) per scan chunk: the smallest of 8/16/24/32/48
// holding 0..tn in the FP32 form (one pass over the data), 8/16 in the FP64 form
int scan_chunk(const ps2_session *s) {
    if (s->tn < 8) return 8;
    if (s->form != PS2_FORM_FP32 || s->tn < 16) return 16;
    if (s->tn < 24) return 24;
    if (s->tn < 32) return 32;
    return 48;
}

template <typename R, bool M = false>
int launch_fit(ps2_session *s, bool in_loop = false) {
    const FitParams fp = fit_params(s);
    const unsigned nb = blocks_for(s->P, 256);
    if (sizeof(R) == 4 && !in_loop) {   // (inside the loop the control step clears it)
        kgo<M>(s, k_zero_u32<M>, dim3(1), dim3(64), ZeroArgs{s->rescore_count.as<uint32_t>(), 4, s->gate},
               ST_SCAN);
        PS2_LAUNCHED();
    }
    // the FP32 form's small patches in the loop scan inside k_fit_small (8-trial chunks)
    const bool scan_in_fit = sizeof(R) == 4 && in_loop && small_patch(s) && scan_chunk(s) == 8;
    if (!scan_in_fit) {
    SlotTimer tm(s, SLOT_SCAN);
    const ScanArgs<R> sa{s->phnm.as<cplx<R>>(), s->bp.as<R>(), s->wt.as<cplx<R>>(), s->filt.as<cplx<R>>(),
                         s->cell_of.as<int32_t>(), s->tix.as<int32_t>(), s->rng.as<R>(),
                         s->offset.as<cplx<R>>(), s->rescore_list.as<int32_t>(),
                         s->rescore_count.as<uint32_t>(), fp, s->gate};
    auto go = [&](auto kern) { kgo<M>(s, kern, dim3(nb), dim3(256), sa, ST_SCAN); };
    // trials per chunk: the FP32 form scans all trials in one pass up to 48
    // positive trials (larger grids take 48-trial chunks)
    const int tc = scan_chunk(s);
    if constexpr (sizeof(R) == 4) {
        // FP32: every trial fits one chunk (tn < tc) except beyond 47
        // positive trials, which take 48-trial chunks
        if (tc == 8 && small_patch(s)) {
            // small patches: a quad of lanes per row, 64 rows per workgroup
            auto goq = [&](auto kern) {
                kgo<M>(s, kern, dim3(blocks_for(s->P, 64)), dim3(256), sa, ST_SCAN);
            };
            switch (s->tn) {
            case 1: goq(k_scan_quad<M, 1>); break;
            case 2: goq(k_scan_quad<M, 2>); break;
            case 3: goq(k_scan_quad<M, 3>); break;
            case 4: goq(k_scan_quad<M, 4>); break;
            case 5: goq(k_scan_quad<M, 5>); break;
            case 6: goq(k_scan_quad<M, 6>); break;
            case 7: goq(k_scan_quad<M, 7>); break;
            default: goq(k_scan_quad<M, 0>); break;
            }
        } else if (tc == 8) {
            // the trial count as a compile-time constant (no per-trial test)
            switch (s->tn) {
            case 1: go(k_scan<M, R, 8, true, 1>); break;
            case 2: go(k_scan<M, R, 8, true, 2>); break;
            case 3: go(k_scan<M, R, 8, true, 3>); break;
            case 4: go(k_scan<M, R, 8, true, 4>); break;
            case 5: go(k_scan<M, R, 8, true, 5>); break;
            case 6: go(k_scan<M, R, 8, true, 6>); break;
            case 7: go(k_scan<M, R, 8, true, 7>); break;
            default: go(k_scan<M, R, 8, true>); break;
            }
        } else if (tc == 16)
            go(k_scan<M, R, 16, true>);
        else if (tc == 24)
            go(k_scan<M, R, 24, true>);
        else if (tc == 32)
            go(k_scan<M, R, 32, true>);
        else if (s->tn < 48)
            go(k_scan<M, R, 48, true>);
        else
            go(k_scan<M, R, 48>);
    } else {
        if (tc == 8)
            go(k_scan<M, R, 8>);
        else
            go(k_scan<M, R, 16>);
    }
    PS2_LAUNCHED();
    }
    RmsHistOut ro{};
    ro.mids = s->mids.as<double>();
    ro.hp.P = s->P;
    ro.hp.equal = s->hist_rule.equal ? 1 : 0;
    ro.hp.cutoff0 = s->hist_rule.cutoff0;
    ro.hp.d = s->hist_rule.d;
    ro.hp.nbins = PS2_COH_BINS;
    if (in_loop) {
        ro.coh_prev = s->cohprev.as<double>();
        ro.part = s->rms_part.as<double>();
        ro.hist = s->hist.as<uint32_t>();
    }
    RecheckIn rc{};
    rc.bp64 = bp64_of(s);
    rc.enabled = sizeof(R) == 4 ? 1 : 0;
    rc.psquare = s->weighting == PS2_WEIGHT_PSQUARE ? 1 : 0;
    if (in_loop && small_patch(s)) {
        // re-score, refine and post step of each row group in one workgroup:
        // 64-row groups with a quad of lanes per row in the refine (FP32
        // form), 256-row groups with a thread per row (FP64 form)
        SlotTimer tm(s, SLOT_REFINE);
        const FitTailArgs<R> a{s->phnm.as<cplx<R>>(), s->bp.as<R>(), rc.bp64, s->filt.as<cplx<R>>(),
                               s->cell_of.as<int32_t>(), s->tix.as<int32_t>(), s->rng.as<R>(),
                               s->offset.as<cplx<R>>(), s->K.as<double>(), s->Cc.as<double>(),
                               s->coh.as<double>(), s->nopt.as<double>(), s->rescore_list.as<int32_t>(),
                               s->rescore_count.as<uint32_t>(), s->wt.as<cplx<R>>(), fp, ro.mids, ro.hp, rc,
                               s->flags.as<uint8_t>(), ro, sizeof(R) == 4 ? 1 : 0, s->gate};
        if constexpr (sizeof(R) == 4) {
            const dim3 g(blocks_for(s->P, SMALL_GROUP));
            if (scan_in_fit) {
                switch (s->tn) {
                case 1: kgo<M>(s, k_fit_small<M, true, 1>, g, dim3(256), a, ST_REFINE); break;
                case 2: kgo<M>(s, k_fit_small<M, true, 2>, g, dim3(256), a, ST_REFINE); break;
                case 3: kgo<M>(s, k_fit_small<M, true, 3>, g, dim3(256), a, ST_REFINE); break;
                case 4: kgo<M>(s, k_fit_small<M, true, 4>, g, dim3(256), a, ST_REFINE); break;
                case 5: kgo<M>(s, k_fit_small<M, true, 5>, g, dim3(256), a, ST_REFINE); break;
                case 6: kgo<M>(s, k_fit_small<M, true, 6>, g, dim3(256), a, ST_REFINE); break;
                case 7: kgo<M>(s, k_fit_small<M, true, 7>, g, dim3(256), a, ST_REFINE); break;
                default: kgo<M>(s, k_fit_small<M, true, 0>, g, dim3(256), a, ST_REFINE); break;
                }
            } else {
                kgo<M>(s, k_fit_small<M>, g, dim3(256), a, ST_REFINE);
            }
        } else {
            kgo<M>(s, k_fit_tail<M, R>, dim3(nb), dim3(256), a, ST_REFINE);
        }
        PS2_LAUNCHED();
        return PS2_OK;
    }
    if (sizeof(R) == 4) {
        SlotTimer tm(s, SLOT_RESCORE);
        // one wavefront per queued row (grid-stride); the queue holds ~0.2-1% of
        // the rows, and launching many more workgroups than that costs more
        // than it saves (16,384 of them: 12 us at 20,000 rows for ~20 rows)
        const RescoreArgs<R> a{s->phnm.as<cplx<R>>(), rc.bp64, s->filt.as<cplx<R>>(),
                               s->cell_of.as<int32_t>(), s->rescore_list.as<int32_t>(),
                               s->rescore_count.as<uint32_t>(), s->tix.as<int32_t>(), s->offset.as<cplx<R>>(),
                               fp, s->gate};
        kgo<M>(s, k_rescore<M, R>, dim3((unsigned)std::min<int64_t>(s->P, 4096)), dim3(64), a, ST_RESCORE);
        PS2_LAUNCHED();
    }
    // the refine kernel marks the rows whose discrete choices lie within its
    // error bound (FP32 form); k_post refines them again in FP64 and, inside
    // the loop, computes the gamma-change RMS partials, coh_prev <- coh and
    // the coherence histogram (k_control then takes the loop decisions)
    {
        SlotTimer tm(s, SLOT_REFINE);
        // the FP32 form queues its flagged rows for k_redecide (the count in
        // rescore_count[1], the loop's re-decision statistic)
        const RefineArgs<R> a{s->phnm.as<cplx<R>>(), s->bp.as<R>(), s->filt.as<cplx<R>>(),
                              s->cell_of.as<int32_t>(), s->tix.as<int32_t>(), s->rng.as<R>(),
                              s->offset.as<cplx<R>>(), s->K.as<double>(), s->Cc.as<double>(),
                              s->coh.as<double>(), s->nopt.as<double>(), fp, ro.mids, ro.hp, rc,
                              s->flags.as<uint8_t>(), rc.enabled ? s->recheck_list.as<int32_t>() : nullptr,
                              s->rescore_count.as<uint32_t>() + 1, s->gate};
        kgo<M>(s, k_refine<M, R>, dim3(nb), dim3(256), a, ST_REFINE);
        PS2_LAUNCHED();
    }
    if (!in_loop && !rc.enabled) return PS2_OK;
    SlotTimer tm(s, SLOT_POST);
    if (rc.enabled) {
        // the queued rows, one wavefront each over the whole GPU
        const RedecideArgs da{reinterpret_cast<const cplx<float> *>(s->phnm.p), rc.bp64,
                              reinterpret_cast<const cplx<float> *>(s->filt.p), s->cell_of.as<int32_t>(),
                              s->tix.as<int32_t>(), s->K.as<double>(), s->Cc.as<double>(), s->coh.as<double>(),
                              fp, s->recheck_list.as<int32_t>(), s->rescore_count.as<uint32_t>() + 1, s->gate};
        kgo<M>(s, k_redecide<M>, dim3((unsigned)std::min<int64_t>(s->P, 4096)), dim3(64), da, ST_REDECIDE);
        PS2_LAUNCHED();
        if (!in_loop) return PS2_OK;
    }
    // grid-stride over the 256-row groups: 2 workgroups per CU at most; the
    // re-decided rows are already rewritten, so the post step only reads coh
    RecheckIn prc = rc;
    prc.enabled = 0;
    const PostArgs<R> pa{s->phnm.as<cplx<R>>(), s->filt.as<cplx<R>>(), s->cell_of.as<int32_t>(),
                         s->tix.as<int32_t>(), s->K.as<double>(), s->Cc.as<double>(), s->coh.as<double>(), fp,
                         prc, s->flags.as<uint8_t>(), ro, in_loop ? 1 : 0, s->rescore_count.as<uint32_t>(),
                         s->gate};
    kgo<M>(s, k_post<M, R>, dim3(std::min<unsigned>(nb, 2 * (unsigned)s->n_cu)), dim3(256), pa, ST_POST);
    PS2_LAUNCHED();
    return PS2_OK;
}

// One gamma iteration (native_stage2.rs:261-390): every launch is gated by
// the control block, so a whole loop can be launched without the host seeing
// the stop decision.
template <typename R, bool M = false>
int launch_iteration(ps2_session *s) {
    const int64_t P = s->P;
    int rc;
    if (s->verbose) PS2_HIP(hipEventRecord(s->ev[2], s->stream));
    // the fill also keeps the K and weighting it used (for the ph_weight
    // output); a small patch's fill computes the P-square weights of the
    // previous iteration
    const bool psquare = s->weighting == PS2_WEIGHT_PSQUARE;
    const bool psquare_in_fill = psquare && small_patch(s);
    if ((rc = launch_fill_grid<R, M>(s, s->K.as<double>(), s->W.as<double>(), s->Kprev.as<double>(),
                                     s->Wprev.as<double>(), psquare_in_fill)))
        return rc;
    if (s->verbose) PS2_HIP(hipEventRecord(s->ev[3], s->stream));
    if ((rc = launch_clap<R, M>(s, true))) return rc;
    if (s->verbose) PS2_HIP(hipEventRecord(s->ev[4], s->stream));
    if ((rc = launch_fit<R, M>(s, true))) return rc;
    if (s->verbose) PS2_HIP(hipEventRecord(s->ev[5], s->stream));
    // gamma-change RMS, stop test, Nr scaling and the P-square table
    const ControlArgs ca{s->ctl.as<LoopCtl>(), s->rms_part.as<double>(), s->hist.as<uint32_t>(),
                         s->rescore_count.as<uint32_t>(), s->prand_hi.as<double>(), s->ctl_params};
    kgo<M>(s, k_control<M>, dim3(1), dim3(256), ca, ST_CONTROL);
    PS2_LAUNCHED();
    if (psquare_in_fill) return PS2_OK;   // the next fill computes them
    SlotTimer tm(s, SLOT_WEIGHTS);
    if (psquare)
        kgo<M>(s, k_weights_psquare<M>, dim3(blocks_for(P, 256)), dim3(256),
               PsquareArgs{s->coh.as<double>(), s->prand_hi.as<double>(), CTL_HI, s->W.as<double>(), P, s->gate},
               ST_WEIGHTS);
    else
        kgo<M>(s, k_weights_snr<M, R>, dim3(blocks_for(P, 256)), dim3(256),
               SnrArgs<R>{s->phnm.as<cplx<R>>(), s->amp.as<float>(), s->bp.as<R>(), s->filt.as<cplx<R>>(),
                          s->cell_of.as<int32_t>(), s->K.as<double>(), s->nopt.as<double>(), s->W.as<double>(),
                          P, (int)s->I, s->rowinv, s->C, s->gate},
               ST_WEIGHTS);
    PS2_LAUNCHED();
    return PS2_OK;
}

// The outputs in the caller's layouts (k_final_outputs).  loop: every
// write_pm1 argument of the loop; otherwise (the fit step) ph_patch, ph_res,
// K/C/coh/N_opt and the trial index.
template <typename R, bool M = false>
int launch_outputs(ps2_session *s, bool loop) {
    const int64_t P = s->P, I = s->I, C = s->C;
    SlotTimer tm(s, SLOT_OUTPUTS);
    FinalIn fi;
    fi.K = s->K.as<double>();
    fi.Cc = s->Cc.as<double>();
    fi.coh = s->coh.as<double>();
    fi.nopt = s->nopt.as<double>();
    fi.coh_prev = loop ? s->cohprev.as<double>() : nullptr;
    fi.Kw = s->Kprev.as<double>();
    fi.Ww = s->Wprev.as<double>();
    fi.tix = s->tix.as<int32_t>();
    FinalOut fo;
    fo.weight = loop ? s->o_weight.as<float2>() : nullptr;
    fo.patch = s->o_patch.as<float2>();
    fo.res = s->o_res.as<float>();
    fo.vec = s->o_vec.as<double>();
    fo.tix = s->o_tix.as<int32_t>();
    fo.grid = loop ? s->o_grid.as<float2>() : nullptr;
    auto go = [&](auto tp_tag, auto tf_tag, auto nt_tag) {
        constexpr int TP = decltype(tp_tag)::value, TF = decltype(tf_tag)::value;
        constexpr int NT = decltype(nt_tag)::value;
        fo.n_ifg_tiles = (int)blocks_for(I, TF);
        fo.n_pix_blocks = (int64_t)blocks_for(P, TP) * fo.n_ifg_tiles;
        fo.n_grid_ifg_tiles = (int)blocks_for(I, 32);
        const int64_t n_grid_blocks = loop ? (int64_t)blocks_for(C, 32) * fo.n_grid_ifg_tiles : 0;
        const OutputsArgs<R> a{s->phnm.as<cplx<R>>(), s->bp.as<R>(), s->rowinv, s->filt.as<cplx<R>>(),
                               s->grid.as<cplx<R>>(), s->cell_of.as<int32_t>(), s->d_order.as<int64_t>(), fi,
                               fo, P, (int)I, C};
        kgo<M>(s, k_final_outputs<M, R, TP, TF, NT>, dim3((unsigned)(fo.n_pix_blocks + n_grid_blocks)), dim3(NT),
               a, ST_OUTPUTS);
    };
    // the tile spans the whole row where it can (at most 256 interferograms).
    // The stores are row segments at the caller's row positions: for pixels
    // listed in random spatial order these are scattered 0.25-1 KB writes,
    // which bound the kernel (wider pixel tiles or 512/1024 threads measured
    // no faster at 500k x 60, 4M x 120 or 1M x 200).
    using std::integral_constant;
    if (I <= 64)
        go(integral_constant<int, 32>{}, integral_constant<int, 64>{}, integral_constant<int, 256>{});
    else if (I <= 128)
        go(integral_constant<int, 16>{}, integral_constant<int, 128>{}, integral_constant<int, 256>{});
    else
        go(integral_constant<int, 16>{}, integral_constant<int, 256>{}, integral_constant<int, 256>{});
    PS2_LAUNCHED();
    return PS2_OK;
}

// --- autotuner ------------------------------------------------------------------
// Tiles measured on the session's own data, keyed on the patch's shape and
// kept in a text cache (one line per shape), FP32 form.  Candidates: fill
// workgroups of 64 or 128 cells (per-cell sums in the same order: bit-identical
// grids), the scan's unit steps from the table or computed from bperp (the same
// bits), and the CLAP strip segmentations the model ranks first (sums grouped
// differently: equal to f32 rounding).  The tuned tiles apply to every later
// run of the session; a cache hit applies them without measuring.

static std::string tile_key(const ps2_session *s) {
    char b[256];
    const int pb = (int)std::lround(2.0 * std::log2((double)std::max<int64_t>(s->P, 1)));
    std::snprintf(b, sizeof b, "ps2tiles1 cu=%d form=%d I=%lld T=%d rowinv=%d P2=%d win=%dx%d", s->n_cu, s->form,
                  (long long)s->I, s->T, s->rowinv, pb, s->nwi, s->nwj);
    return b;
}

struct TileLine {
    int fill_cb = 0, scan_table = -1, nseg = 0;
};

// cache lookup / append under an advisory lock of the file
static int tile_cache(const char *path, const std::string &key, TileLine *found, const TileLine *add) {
    const int fd = ::open(path, O_RDWR | O_CREAT | O_APPEND, 0644);
    if (fd < 0) {
        set_error(std::string("autotune cache ") + path + ": " + std::strerror(errno));
        return PS2_E_ARG;
    }
    ::flock(fd, add ? LOCK_EX : LOCK_SH);
    int rc = PS2_OK;
    if (found) {
        FILE *f = ::fdopen(::dup(fd), "r");
        if (f) {
            ::rewind(f);
            char line[512];
            while (std::fgets(line, sizeof line, f)) {
                const size_t n = key.size();
                if (std::strncmp(line, key.c_str(), n) != 0 || line[n] != ' ') continue;
                TileLine t;
                if (std::sscanf(line + n, " fill_cells=%d scan_step_table=%d clap_segments=%d", &t.fill_cb,
                                &t.scan_table, &t.nseg) == 3) {
                    *found = t;
                    break;
                }
            }
            std::fclose(f);
        }
    }
    if (add) {
        char b[96];
        const int len = std::snprintf(b, sizeof b, " fill_cells=%d scan_step_table=%d clap_segments=%d\n", add->fill_cb,
                                      add->scan_table, add->nseg);
        const std::string line = key + std::string(b, (size_t)len);
        if (::write(fd, line.data(), line.size()) != (ssize_t)line.size()) {
            set_error(std::string("autotune cache ") + path + ": write failed");
            rc = PS2_E_ARG;
        }
    }
    ::flock(fd, LOCK_UN);
    ::close(fd);
    return rc;
}

template <typename R>
int apply_segments(ps2_session *s, int nseg) {
    const int ngroups = (int)((s->I + CLAP_TEAMS - 1) / CLAP_TEAMS);
    for (const ClapSegs &c : clap_segment_candidates(s->nwi, s->nwj, ngroups, (int)s->I, s->clap_slots))
        if (c.nseg == nseg) {
            s->segs = c;
            return PS2_OK;
        }
    return arg_error("autotune cache: no CLAP segmentation with " + std::to_string(nseg) + " segments");
}

template <typename R>
int autotune_tiles(ps2_session *s, const char *path) {
    const std::string key = tile_key(s);
    TileLine hit;
    int rc;
    if (path && (rc = tile_cache(path, key, &hit, nullptr))) return rc;
    const bool scan_choice = !s->rowinv && s->tn < 8;   // the one scan that can compute its steps
    if (hit.fill_cb) {
        s->fill_cb = hit.fill_cb == 64 ? 64 : 128;
        if (scan_choice) s->scan_table = hit.scan_table ? 1 : 0;
        if (s->windows && !s->gclap && (rc = apply_segments<R>(s, hit.nseg))) return rc;
        s->tile_source = PS2_TILES_CACHE;
    } else {
        // measured on this session's inputs from the initial state (K = 0,
        // weighting 1/D_A): average of 3 launches after one warm-up
        hipEvent_t e0 = nullptr, e1 = nullptr;
        PS2_HIP(hipEventCreate(&e0));
        PS2_HIP(hipEventCreate(&e1));
        auto timed = [&](auto &&launch, float *ms) -> int {
            int r;
            if ((r = launch())) return r;
            PS2_HIP(hipEventRecord(e0, s->stream));
            for (int q = 0; q < 3 && !r; ++q) r = launch();
            if (r) return r;
            PS2_HIP(hipEventRecord(e1, s->stream));
            PS2_HIP(hipEventSynchronize(e1));
            PS2_HIP(hipEventElapsedTime(ms, e0, e1));
            return PS2_OK;
        };
        auto done = [&](int r) {
            (void)hipStreamSynchronize(s->stream);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return r;
        };
        PS2_HIP(hipMemsetAsync(s->K.p, 0, sizeof(double) * s->P, s->stream));
        PS2_HIP(hipMemcpyAsync(s->W.p, s->W0.p, sizeof(double) * s->P, hipMemcpyDeviceToDevice, s->stream));
        float best = -1.0f;
        int best_cb = s->fill_cb;
        for (int cb : {64, 128}) {
            s->fill_cb = cb;
            float ms = 0.0f;
            if ((rc = timed([&] { return launch_fill_grid<R>(s, s->K.as<double>(), s->W.as<double>()); }, &ms)))
                return done(rc);
            if (best < 0.0f || ms < best) best = ms, best_cb = cb;
        }
        s->fill_cb = best_cb;
        if (s->windows && !s->gclap) {
            const int ngroups = (int)((s->I + CLAP_TEAMS - 1) / CLAP_TEAMS);
            std::vector<ClapSegs> cand = clap_segment_candidates(s->nwi, s->nwj, ngroups, (int)s->I, s->clap_slots);
            if (cand.size() > 4) cand.resize(4);
            best = -1.0f;
            ClapSegs best_seg = cand.front();
            for (const ClapSegs &c : cand) {
                s->segs = c;
                float ms = 0.0f;
                if ((rc = timed([&] { return launch_clap<R>(s, true); }, &ms))) return done(rc);
                if (best < 0.0f || ms < best) best = ms, best_seg = c;
            }
            s->segs = best_seg;
            if ((rc = launch_clap<R>(s, true))) return done(rc);   // the filtered grid the scan reads
        } else if (s->windows) {
            if ((rc = launch_clap<R>(s, true))) return done(rc);   // (any other window: no tiles)
        }
        if (scan_choice) {
            best = -1.0f;
            int best_t = s->scan_table;
            for (int tab : {1, 0}) {
                s->scan_table = tab;
                float ms = 0.0f;
                if ((rc = timed([&] { return launch_fit<R>(s, false); }, &ms))) return done(rc);
                if (best < 0.0f || ms < best) best = ms, best_t = tab;
            }
            s->scan_table = best_t;
        }
        if ((rc = done(PS2_OK))) return rc;
        s->tile_source = PS2_TILES_MEASURED;
        if (path) {
            TileLine t;
            t.fill_cb = s->fill_cb;
            t.scan_table = s->scan_table;
            t.nseg = s->windows && !s->gclap ? s->segs.nseg : 0;
            if ((rc = tile_cache(path, key, nullptr, &t))) return rc;
        }
    }
    // the step table stays only where the scan reads it
    if (!(scan_step_table(s) || s->tn >= 48)) (void)s->wt.alloc(0);
    return PS2_OK;
}

// The gamma loop from the initial state (native_stage2.rs:242-258) to the
// stop decision, and the outputs of the last iteration.  Default: the
// starting state, gamma_max_iterations iterations and the outputs are
// captured as one HIP graph once per session (one iteration, launched that
// many times after the starting state is set, when there are more than
// LOOP_GRAPH_MAX) and launched; the control kernel's stop
// flag turns the remaining kernels into early returns, and the host waits
// once.  With per-kernel timing the iterations are launched kernel by kernel,
// the timing events carried in the dispatches, all of them queued before the
// one host wait as in the graph (so the device runs them back to back, as in
// an untimed loop), and the timers of iterations after the stop decision are
// dropped.  With verbose output the host reads the control block after each
// iteration.
constexpr int64_t LOOP_GRAPH_MAX = 64;
// timer events per queued iteration (at most 12 timed launches each)
constexpr int TIMERS_PER_ITERATION = 24;

// The loop's starting state, on stream st (not captured): K = 0,
// coh_prev = 0, the first iteration's histogram and re-score counter (k_control
// clears them for every later one), the initial weighting, the control block.
int loop_prepare(ps2_session *s, hipStream_t st) {
    const int64_t P = s->P;
    LoopCtl *hc = s->h_ctl.as<LoopCtl>();
    PS2_HIP(hipMemsetAsync(s->K.p, 0, sizeof(double) * P, st));
    PS2_HIP(hipMemsetAsync(s->cohprev.p, 0, sizeof(double) * P, st));
    PS2_HIP(hipMemsetAsync(s->hist.p, 0, sizeof(uint32_t) * CTL_HIST_WORDS, st));
    PS2_HIP(hipMemsetAsync(s->rescore_count.p, 0, sizeof(uint32_t) * 4, st));
    PS2_HIP(hipMemcpyAsync(s->W.p, s->W0.p, sizeof(double) * P, hipMemcpyDeviceToDevice, st));
    std::memset(hc, 0, sizeof(LoopCtl));
    hc->i_loop = 1;
    for (int b = 0; b < CTL_BINS; ++b) hc->nr[b] = s->nr_base[b];
    PS2_HIP(hipMemcpyAsync(s->ctl.p, hc, sizeof(LoopCtl), hipMemcpyHostToDevice, st));
    return PS2_OK;
}

// The same starting state as one kernel (recorded for the merged launches).
template <bool M>
int launch_prepare(ps2_session *s) {
    const PrepareArgs a{s->K.as<double>(), s->cohprev.as<double>(), s->W.as<double>(), s->W0.as<double>(),
                        s->hist.as<uint32_t>(), s->rescore_count.as<uint32_t>(), s->ctl.as<LoopCtl>(),
                        s->nr0.as<double>(), s->P};
    kgo<M>(s, k_loop_prepare<M>, dim3(blocks_for(s->P, 256)), dim3(256), a, ST_PREPARE);
    PS2_LAUNCHED();
    return PS2_OK;
}

// The loop's results from the control block read back into h_ctl.
void loop_finish(ps2_session *s, double seconds) {
    const LoopCtl *hc = s->h_ctl.as<LoopCtl>();
    s->loop_seconds = seconds;
    s->i_loop = hc->i_loop;
    s->gamma_change_save = hc->gcs;
    s->nr.assign(hc->nr, hc->nr + CTL_BINS);
    s->n_rescored = (int64_t)hc->rescored;
    s->n_rechecked = (int64_t)hc->rechecked;
    s->ran = true;
    s->loop_outputs = true;
}

template <typename R>
int run_loop(ps2_session *s) {
    int rc;
    LoopCtl *hc = s->h_ctl.as<LoopCtl>();
    s->gate = &s->ctl.as<LoopCtl>()->stopped;
    const int64_t max_iter = s->ctl_params.max_iter;
    bool stepwise = s->timing || s->verbose || s->no_graph;
    // timing without verbose output: every iteration queued, one host wait
    const bool queued_timing = s->timing && !s->verbose;
    if (queued_timing) {
        const size_t need = (size_t)TIMERS_PER_ITERATION * (size_t)max_iter + 64;
        if (max_iter > 4096) return arg_error("per-kernel timing supports at most 4096 iterations per loop");
        if (s->tev.size() < need) {
            AllocGuard g(capture_mutex());
            const size_t have = s->tev.size();
            s->tev.resize(need, nullptr);
            for (size_t e = have; e < need; ++e) PS2_HIP(hipEventCreate(&s->tev[e]));
        }
    }
    // the graph holds every iteration (up to LOOP_GRAPH_MAX of them, else one
    // iteration replayed): each graph launch costs ~30 us more than the
    // ~10 us between two of its kernels
    const int64_t per_graph = max_iter <= LOOP_GRAPH_MAX ? max_iter : 1;
    // a graph that holds every iteration starts with the loop's starting
    // state (one kernel) and ends with the output kernel: one launch per loop
    const bool outputs_in_graph = per_graph == max_iter;
    if (!stepwise && !s->graph_exec) {
        // no allocation anywhere in the library while the capture is open
        std::unique_lock<std::shared_mutex> cap(capture_mutex());
        PS2_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        rc = outputs_in_graph ? launch_prepare<false>(s) : 0;
        for (int64_t it = 0; it < per_graph && !rc; ++it) rc = launch_iteration<R>(s);
        if (!rc && outputs_in_graph) rc = launch_outputs<R>(s, true);
        if (!rc && outputs_in_graph) {
            hipLaunchKernelGGL(k_publish_ctl, dim3(1), dim3(64), 0, s->stream, s->ctl.as<LoopCtl>(), s->d_ctl_pub);
            rc = hipGetLastError() == hipSuccess ? 0 : PS2_E_DEVICE;
        }
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->stream, &g);
        cap.unlock();
        hipError_t ei = hipErrorUnknown;
        if (!rc && e == hipSuccess && g) ei = hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0);
        if (ei == hipSuccess) {
            s->graph = g;
        } else {
            // The capture failed, e.g. invalidated by another host thread's
            // allocation, free or blocking copy on the device meanwhile.
            // Nothing of the captured iteration ran: this session launches its
            // iterations one by one from now on (same kernels, same results).
            if (g) (void)hipGraphDestroy(g);
            s->graph_exec = nullptr;
            (void)hipGetLastError();
            s->no_graph = true;
            stepwise = true;
        }
    }
    PS2_HIP(hipEventRecord(s->ev[0], s->stream));
    if ((stepwise || !outputs_in_graph) && (rc = loop_prepare(s, s->stream))) {
        s->gate = nullptr;
        return rc;
    }
    if (!stepwise) {
        for (int64_t it = 0; it < max_iter; it += per_graph) PS2_HIP(hipGraphLaunch(s->graph_exec, s->stream));
        if (!outputs_in_graph && (rc = launch_outputs<R>(s, true))) {
            s->gate = nullptr;
            return rc;
        }
    } else if (queued_timing) {
        for (int64_t it = 0; it < max_iter; ++it) {
            s->timer_iter = it + 1;
            if ((rc = launch_iteration<R>(s))) {
                s->timer_iter = 0;
                s->gate = nullptr;
                return rc;
            }
        }
        s->timer_iter = 0;
        s->gate = nullptr;
        if ((rc = launch_outputs<R>(s, true))) return rc;
    } else {
        for (int64_t it = 0; it < max_iter; ++it) {
            const auto t_it = std::chrono::steady_clock::now();
            if ((rc = launch_iteration<R>(s))) {
                s->gate = nullptr;
                return rc;
            }
            PS2_HIP(hipMemcpyAsync(hc, s->ctl.p, sizeof(LoopCtl), hipMemcpyDeviceToHost, s->stream));
            PS2_HIP(hipStreamSynchronize(s->stream));
            collect_timers(s);
            if (s->verbose) {
                float t_fill = 0, t_clap = 0, t_fit = 0;
                (void)hipEventElapsedTime(&t_fill, s->ev[2], s->ev[3]);
                (void)hipEventElapsedTime(&t_clap, s->ev[3], s->ev[4]);
                (void)hipEventElapsedTime(&t_fit, s->ev[4], s->ev[5]);
                const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_it).count();
                std::fprintf(stderr,
                             "stage2 hip iteration %lld: fill+grid=%.3fms clap=%.3fms topofit=%.3fms "
                             "wall=%.3fms gamma_change=%.6e stop=%s\n",
                             (long long)(it + 1), t_fill, t_clap, t_fit, wall * 1e3, hc->gcs,
                             hc->stopped ? "true" : "false");
            }
            if (hc->stopped) break;
        }
        s->gate = nullptr;
        if ((rc = launch_outputs<R>(s, true))) return rc;
    }
    PS2_HIP(hipEventRecord(s->ev[1], s->stream));
    // a graph holding the whole loop ends with the control block written to
    // h_ctl_pub; else it is copied down
    const bool published = !stepwise && outputs_in_graph;
    if (!published) PS2_HIP(hipMemcpyAsync(hc, s->ctl.p, sizeof(LoopCtl), hipMemcpyDeviceToHost, s->stream));
    PS2_HIP(hipStreamSynchronize(s->stream));
    if (published) std::memcpy(hc, s->h_ctl_pub.p, sizeof(LoopCtl));
    s->gate = nullptr;
    collect_timers(s, hc->i_loop);
    float ms = 0.0f;
    PS2_HIP(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
    loop_finish(s, (double)ms * 1e-3);
    return PS2_OK;
}

// The loop launched an iteration at a time, the state after each written in
// the caller's layouts and handed to fn (ps2_hip_session_run_iterations).
int download_outputs(ps2_session *s, ps2_patch_out *out, bool loop_outputs);

template <typename R>
int run_loop_iterations(ps2_session *s, ps2_patch_out *state, ps2_iteration_fn fn, void *user) {
    int rc = PS2_OK;
    LoopCtl *hc = s->h_ctl.as<LoopCtl>();
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = loop_prepare(s, s->stream))) return rc;
    const int32_t *gate = &s->ctl.as<LoopCtl>()->stopped;
    for (int64_t it = 0; it < s->ctl_params.max_iter && rc == PS2_OK; ++it) {
        s->gate = gate;
        if ((rc = launch_iteration<R>(s))) break;
        s->gate = nullptr;   // the state's outputs, also after the stop decision
        if ((rc = launch_outputs<R>(s, true))) break;
        PS2_HIP(hipMemcpyAsync(hc, s->ctl.p, sizeof(LoopCtl), hipMemcpyDeviceToHost, s->stream));
        PS2_HIP(hipStreamSynchronize(s->stream));
        collect_timers(s);
        s->i_loop = it + 1;
        s->gamma_change_save = hc->gcs;
        s->nr.assign(hc->nr, hc->nr + CTL_BINS);
        s->n_rescored = (int64_t)hc->rescored;
        s->n_rechecked = (int64_t)hc->rechecked;
        s->ran = true;
        s->loop_outputs = true;
        if ((rc = download_outputs(s, state, true))) break;
        if (fn) {
            const int u = fn(user, it + 1, state);
            if (u) {
                set_error("iteration callback returned " + std::to_string(u));
                rc = u;
                break;
            }
        }
        if (hc->stopped) break;
    }
    s->gate = nullptr;
    if (rc) return rc;
    loop_finish(s, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    return PS2_OK;
}

// download: the outputs are already in the caller's layouts on the device
// (k_final_outputs); each goes down in one bulk copy.

int download_outputs(ps2_session *s, ps2_patch_out *out, bool loop_outputs) {
    const int64_t P = s->P, I = s->I, C = s->C;
    const size_t plane = (size_t)P * I;
    int rc;
    double *const vec_dst[5] = {out->k_ps, out->c_ps, out->coh_ps, out->n_opt,
                                loop_outputs ? out->coh_ps_save : nullptr};
    for (int v = 0; v < 5; ++v)
        if (vec_dst[v] && (rc = staging().d2h(vec_dst[v], s->o_vec.as<double>() + (size_t)v * P,
                                              sizeof(double) * P, s->stream)))
            return rc;
    if (out->trial_ix && (rc = staging().d2h(out->trial_ix, s->o_tix.p, sizeof(int32_t) * P, s->stream)))
        return rc;
    if (out->ph_patch && (rc = staging().d2h(out->ph_patch, s->o_patch.p, sizeof(float2) * plane, s->stream)))
        return rc;
    if (out->ph_res && (rc = staging().d2h(out->ph_res, s->o_res.p, sizeof(float) * plane, s->stream)))
        return rc;
    if (!loop_outputs) return PS2_OK;
    if (out->ph_weight && (rc = staging().d2h(out->ph_weight, s->o_weight.p, sizeof(float2) * plane, s->stream)))
        return rc;
    if (out->ph_grid && (rc = staging().d2h(out->ph_grid, s->o_grid.p, sizeof(float2) * C * I, s->stream)))
        return rc;
    if (out->nr)
        for (int b = 0; b < PS2_COH_BINS; ++b) out->nr[b] = s->nr[b];
    out->gamma_change_save = s->gamma_change_save;
    out->i_loop = s->i_loop;
    out->n_rescored = s->n_rescored;
    out->n_rechecked = s->n_rechecked;
    out->loop_seconds = s->loop_seconds;
    return PS2_OK;
}

// --- several patches per launch ---------------------------------------------
// ps2_hip_sessions_run runs the loops of several sessions of one device as
// merged launches: every loop kernel is launched once per gamma iteration for
// a chain of patches (MultiPack: each patch keeps its own workgroups,
// arguments, gate and convergence; up to PS2_PACK patches per launch).  A
// chain holds patches whose per-iteration data fit together in a part of the
// device's 256 MB last-level cache (kChainBytes), so that each kernel of an
// iteration finds the previous kernel's planes there; a larger patch is a
// chain of its own.  Each chain is a sequence of merged launches on its own
// stream, and the chains overlap (the memory-bound kernels of one with the
// ALU-bound CLAP of another); all chains, all iterations and the outputs are
// captured in one HIP graph (forked from and joined to the first session's
// stream).  The launches are recorded from the same launch
// functions as a single patch's (launch_iteration<R, true>), so every patch
// runs exactly the kernels and arguments of its own loop: results are
// bit-identical to the single runs.  The graph is kept for the next call with
// the same sessions.

// batch_build / sessions_run_merged: the sessions cannot run merged (the
// caller runs them concurrently instead)
constexpr int kNotMerged = 1;
constexpr double kChainBytes = 128e6;

struct MergedLaunch {
    const void *fn;
    uint32_t blocks, threads;
    std::vector<char> pack;   // the kernel's MultiPack argument
};

struct Chain {
    hipStream_t stream;
    // the starting state, one iteration's launches in stage order, the outputs
    std::vector<MergedLaunch> prepare, iter, outputs;
};

struct BatchGraph {
    std::vector<ps2_session *> ss;
    int device = 0;
    std::vector<Chain> chains;
    std::vector<hipEvent_t> events;   // fork, then one join per further chain
    int64_t max_iter = 0, per_graph = 0;
    bool outputs_in_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchGraph() {
        (void)hipSetDevice(device);
        AllocGuard g(capture_mutex());
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
    }
};

std::mutex g_batch_mu;
// most recent last; a run holds its own reference while it uses one
std::vector<std::shared_ptr<BatchGraph>> g_batches;
constexpr size_t kBatchGraphsKept = 4;

void batch_forget_impl(const ps2_session *s) {
    std::vector<std::shared_ptr<BatchGraph>> drop;
    {
        std::lock_guard<std::mutex> lk(g_batch_mu);
        for (auto it = g_batches.begin(); it != g_batches.end();) {
            if (std::find((*it)->ss.begin(), (*it)->ss.end(), s) != (*it)->ss.end()) {
                drop.push_back(std::move(*it));
                it = g_batches.erase(it);
            } else {
                ++it;
            }
        }
    }
    // the graphs are released here, outside the list's lock
}

hipError_t launch_merged(const MergedLaunch &m, hipStream_t st) {
    void *kargs[1] = {const_cast<char *>(m.pack.data())};
    return hipLaunchKernel(m.fn, dim3(m.blocks), dim3(m.threads), kargs, 0, st);
}

// bytes one gamma iteration reads and writes per patch (sorted pixel planes,
// bperp and unit steps; grid, filtered grid and CLAP partials)
double loop_bytes(const ps2_session *s) {
    const double r = (double)s->rsz();
    return (double)s->I * ((double)s->P * (3.0 * r + 8.0) + (double)s->C * 6.0 * r);
}

// Records every session's iteration and outputs launches and merges them, per
// chain, stage by stage: one launch per (stage, kernel) with the chain's
// sessions in order, each session's workgroups starting at a multiple of 8.
int batch_build(const std::vector<ps2_session *> &ss, BatchGraph &bg) {
    const size_t n = ss.size();
    std::vector<std::vector<LaunchRec>> recs(n);
    int rc = PS2_OK;
    for (size_t q = 0; q < n && !rc; ++q) {
        ps2_session *s = ss[q];
        s->gate = &s->ctl.as<LoopCtl>()->stopped;
        s->rec = &recs[q];
        rc = launch_prepare<true>(s);
        if (!rc && s->form == PS2_FORM_FP64) {
            rc = launch_iteration<double, true>(s);
            if (!rc) rc = launch_outputs<double, true>(s, true);
        } else if (!rc) {
            rc = launch_iteration<float, true>(s);
            if (!rc) rc = launch_outputs<float, true>(s, true);
        }
        s->rec = nullptr;
        s->gate = nullptr;
        int last = -1;
        for (const auto &r : recs[q]) {
            if (r.stage <= last) return kNotMerged;
            last = r.stage;
        }
    }
    if (rc) return rc;
    // chains: first fit of the patches, largest first, into chains of at
    // most kChainBytes per iteration (a larger patch alone)
    std::vector<size_t> by_size(n);
    for (size_t q = 0; q < n; ++q) by_size[q] = q;
    std::stable_sort(by_size.begin(), by_size.end(),
                     [&](size_t x, size_t y) { return loop_bytes(ss[x]) > loop_bytes(ss[y]); });
    std::vector<double> load;
    std::vector<int> chain_of(n, 0);
    for (size_t q : by_size) {
        size_t c = 0;
        while (c < load.size() && load[c] + loop_bytes(ss[q]) > kChainBytes) ++c;
        if (c == load.size()) load.push_back(0.0);
        chain_of[q] = (int)c;
        load[c] += loop_bytes(ss[q]);
    }
    const size_t nch = load.size();
    // groups of one (chain, stage, kernel), sessions in order
    struct Group {
        int chain, stage;
        const void *fn;
        uint32_t threads;
        std::vector<const LaunchRec *> r;
    };
    std::vector<Group> groups;
    for (size_t c = 0; c < nch; ++c)
        for (int st = 0; st < ST_COUNT; ++st)
            for (size_t q = 0; q < n; ++q) {
                if (chain_of[q] != (int)c) continue;
                for (const auto &r : recs[q]) {
                    if (r.stage != st || r.blocks == 0) continue;
                    Group *g = nullptr;
                    for (auto &x : groups)
                        if (x.chain == (int)c && x.stage == st && x.fn == r.fn && x.r.size() < PS2_PACK) g = &x;
                    if (!g) {
                        groups.push_back(Group{(int)c, st, r.fn, r.threads, {}});
                        g = &groups.back();
                    }
                    if (g->threads != r.threads) return kNotMerged;
                    if (!g->r.empty() && g->r[0]->args.size() != r.args.size()) return kNotMerged;
                    g->r.push_back(&r);
                }
            }
    // each group's pack: the argument structs, the ranges, the count
    std::vector<std::vector<char>> packs(groups.size());
    std::vector<uint32_t> blocks(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) {
        const LaunchRec &r0 = *groups[g].r[0];
        const size_t asz = r0.args.size();
        std::vector<char> &pk = packs[g];
        pk.assign(r0.pack_size, 0);
        uint32_t *rng = reinterpret_cast<uint32_t *>(pk.data() + r0.range_off);
        uint64_t start = 0;
        for (size_t e = 0; e < groups[g].r.size(); ++e) {
            const LaunchRec &r = *groups[g].r[e];
            std::memcpy(pk.data() + e * asz, r.args.data(), asz);
            rng[2 * e] = (uint32_t)start;
            rng[2 * e + 1] = r.blocks;
            start = (start + r.blocks + PS2_NUM_XCD - 1) / PS2_NUM_XCD * PS2_NUM_XCD;
            if (start >= ((uint64_t)1 << 31)) return kNotMerged;
        }
        const int32_t cnt = (int32_t)groups[g].r.size();
        std::memcpy(pk.data() + r0.n_off, &cnt, sizeof(cnt));
        // the grid ends with the last range (no padding after it)
        blocks[g] = rng[2 * (cnt - 1)] + groups[g].r.back()->blocks;
    }
    // chain c runs on the stream of its first session (chain 0: ss[0]'s)
    bg.chains.assign(nch, Chain{});
    for (size_t c = 0; c < nch; ++c)
        for (size_t q = 0; q < n; ++q)
            if (chain_of[q] == (int)c) {
                bg.chains[c].stream = ss[q]->stream;
                break;
            }
    if (chain_of[0] != 0) {   // ss[0] leads chain 0 (the graph's launch stream)
        const int c0 = chain_of[0];
        std::swap(bg.chains[0].stream, bg.chains[c0].stream);
        for (auto &g : groups) g.chain = g.chain == 0 ? c0 : g.chain == c0 ? 0 : g.chain;
    }
    for (size_t g = 0; g < groups.size(); ++g) {
        MergedLaunch m;
        m.fn = groups[g].fn;
        m.blocks = blocks[g];
        m.threads = groups[g].threads;
        m.pack = std::move(packs[g]);
        Chain &ch = bg.chains[groups[g].chain];
        const int st = groups[g].stage;
        (st == ST_PREPARE ? ch.prepare : st == ST_OUTPUTS ? ch.outputs : ch.iter).push_back(std::move(m));
    }
    bg.ss = ss;
    bg.device = ss[0]->device;
    bg.max_iter = 0;
    for (auto *s : ss) bg.max_iter = std::max<int64_t>(bg.max_iter, s->ctl_params.max_iter);
    bg.per_graph = bg.max_iter <= LOOP_GRAPH_MAX ? bg.max_iter : 1;
    // a graph of every iteration also holds the starting state and the outputs
    bg.outputs_in_graph = bg.per_graph == bg.max_iter;
    bg.events.assign(nch, nullptr);
    for (auto &e : bg.events) PS2_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // capture: fork the chains from chain 0's stream, each chain's iterations
    // (with the starting state and the outputs when every iteration is in),
    // join them back
    hipStream_t st0 = bg.chains[0].stream;
    std::unique_lock<std::shared_mutex> cap(capture_mutex());
    PS2_HIP(hipStreamBeginCapture(st0, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipEventRecord(bg.events[0], st0);
    for (size_t c = 1; c < nch && le == hipSuccess; ++c) le = hipStreamWaitEvent(bg.chains[c].stream, bg.events[0], 0);
    for (size_t c = 0; c < nch && le == hipSuccess; ++c) {
        const Chain &ch = bg.chains[c];
        if (bg.outputs_in_graph)
            for (const auto &m : ch.prepare)
                if (le == hipSuccess) le = launch_merged(m, ch.stream);
        for (int64_t it = 0; it < bg.per_graph && le == hipSuccess; ++it)
            for (const auto &m : ch.iter)
                if (le == hipSuccess) le = launch_merged(m, ch.stream);
        if (bg.outputs_in_graph)
            for (const auto &m : ch.outputs)
                if (le == hipSuccess) le = launch_merged(m, ch.stream);
    }
    for (size_t c = 1; c < nch && le == hipSuccess; ++c) {
        le = hipEventRecord(bg.events[c], bg.chains[c].stream);
        if (le == hipSuccess) le = hipStreamWaitEvent(st0, bg.events[c], 0);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(st0, &graph);
    cap.unlock();
    hipError_t ie = hipErrorUnknown;
    if (le == hipSuccess && ee == hipSuccess && graph) ie = hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        // e.g. invalidated by another host thread's allocation meanwhile;
        // nothing of it ran
        if (graph) (void)hipGraphDestroy(graph);
        bg.exec = nullptr;
        (void)hipGetLastError();
        return kNotMerged;
    }
    bg.graph = graph;
    return PS2_OK;
}

// Runs the sessions' loops as merged launches (see above); the caller holds
// their locks.  Returns kNotMerged without running anything when they cannot
// be merged (timing or verbose sessions, several devices).
int sessions_run_merged(const std::vector<ps2_session *> &ss) {
    const size_t n = ss.size();
    for (auto *s : ss)
        if (s->timing || s->verbose || s->device != ss[0]->device) return kNotMerged;
    PS2_HIP(hipSetDevice(ss[0]->device));
    for (auto *s : ss) PS2_HIP(hipStreamSynchronize(s->stream));
    std::shared_ptr<BatchGraph> bg;
    {
        std::lock_guard<std::mutex> lk(g_batch_mu);
        for (size_t b = 0; b < g_batches.size(); ++b)
            if (g_batches[b]->ss == ss) {
                std::rotate(g_batches.begin() + b, g_batches.begin() + b + 1, g_batches.end());
                bg = g_batches.back();
                break;
            }
    }
    if (!bg) {
        auto nb = std::make_shared<BatchGraph>();
        int rc = batch_build(ss, *nb);
        if (rc) return rc;
        std::vector<std::shared_ptr<BatchGraph>> drop;
        std::lock_guard<std::mutex> lk(g_batch_mu);
        g_batches.push_back(nb);
        bg = nb;
        while (g_batches.size() > kBatchGraphsKept) {
            drop.push_back(std::move(g_batches.front()));
            g_batches.erase(g_batches.begin());
        }
    }
    hipStream_t st = bg->chains[0].stream;
    PS2_HIP(hipEventRecord(ss[0]->ev[0], st));
    if (!bg->outputs_in_graph)
        for (const auto &ch : bg->chains)
            for (const auto &m : ch.prepare) PS2_HIP(launch_merged(m, st));
    for (int64_t it = 0; it < bg->max_iter; it += bg->per_graph) PS2_HIP(hipGraphLaunch(bg->exec, st));
    if (!bg->outputs_in_graph)
        for (const auto &ch : bg->chains)
            for (const auto &m : ch.outputs) PS2_HIP(launch_merged(m, st));
    for (auto *s : ss)
        PS2_HIP(hipMemcpyAsync(s->h_ctl.p, s->ctl.p, sizeof(LoopCtl), hipMemcpyDeviceToHost, st));
    PS2_HIP(hipEventRecord(ss[0]->ev[1], st));
    PS2_HIP(hipStreamSynchronize(st));
    float ms = 0.0f;
    PS2_HIP(hipEventElapsedTime(&ms, ss[0]->ev[0], ss[0]->ev[1]));
    for (size_t q = 0; q < n; ++q) loop_finish(ss[q], (double)ms * 1e-3);
    return PS2_OK;
}

int session_create(const ps2_patch_in *in, const ps2_run_opts *opts, ps2_session **out) {
    if (!out) return arg_error("session out pointer is NULL");
    *out = nullptr;
    int rc = validate(in);
    if (rc) return rc;
    std::unique_ptr<ps2_session> s(new ps2_session());
    s->form = opts ? opts->form : PS2_FORM_FP32;
    if (s->form != PS2_FORM_FP32 && s->form != PS2_FORM_FP64) return arg_error("unknown form");
    s->device = opts ? opts->device : 0;
    s->verbose = opts ? opts->verbose : 0;
    s->autotune = opts && opts->autotune && s->form == PS2_FORM_FP32;
    {
        std::lock_guard<std::mutex> lk(g_init_mu);
        int n = 0;
        PS2_HIP(hipGetDeviceCount(&n));
        if (s->device < 0 || s->device >= n)
            return arg_error("device " + std::to_string(s->device) + " not present (" + std::to_string(n) + " visible)");
        PS2_HIP(hipSetDevice(s->device));
        AllocGuard g(capture_mutex());
        PS2_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        for (auto &e : s->ev) PS2_HIP(hipEventCreate(&e));
    }
    s->P = in->n_ps;
    s->I = in->n_ifg;
    s->ni = in->n_i;
    s->nj = in->n_j;
    s->C = in->n_i * in->n_j;
    s->T = (int)in->n_trial;
    s->tn = (int)(in->n_trial / 2);
    s->rowinv = in->row_invariant ? 1 : 0;
    s->alpha = in->clap_alpha;
    s->beta = in->clap_beta;
    s->nr_max_nz_ix = in->nr_max_nz_ix;
    s->conv = in->gamma_change_convergence;
    s->lct = in->low_coh_thresh;
    s->max_iter = in->gamma_max_iterations;
    s->weighting = in->weighting;
    s->nr_base.assign(in->nr_base, in->nr_base + PS2_COH_BINS);
    s->hist_rule = make_hist_rule(in->coh_bins, PS2_COH_BINS);
    // window rows/columns (native_stage2.rs:754-759)
    s->n_win = (int)in->clap_window;
    s->n_ex = (int)in->n_win_ex;
    s->n_inc = std::max(1, s->n_win / 4);
    s->gclap = s->n_win != CLAP_WIN || s->n_ex != CLAP_FFT;
    s->nwi = (int)std::ceil((double)s->ni / (double)s->n_inc) - 3;
    s->nwj = (int)std::ceil((double)s->nj / (double)s->n_inc) - 3;
    s->windows = s->nwi > 0 && s->nwj > 0;
    if (s->windows && (s->ni < s->n_win || s->nj < s->n_win))
        return arg_error("grid " + std::to_string(s->ni) + "x" + std::to_string(s->nj) +
                         " has CLAP windows but is narrower than the " + std::to_string(s->n_win) +
                         "-cell window (the reference's window placement underflows here)");
    // derived FP32 re-score bound: the scan's worst-case rounding error on one
    // coherence, doubled (recurrence drift over one trial chunk of 8-48 trials
    // + an I-term f32 sum + the sincos anchors), in units of 2^-24
    if (opts && opts->rescore_bound > 0.0)
        s->rescore_bound = (float)opts->rescore_bound;
    else
        s->rescore_bound = (float)(2.0 * ((double)scan_chunk(s.get()) + 8.0 + 2.0 * (double)s->I) *
                                   5.9604644775390625e-08);
    if (s->form == PS2_FORM_FP64) {
        if ((rc = upload_inputs<double>(s.get(), in))) return rc;
        if ((rc = alloc_state<double>(s.get()))) return rc;
    } else {
        if ((rc = upload_inputs<float>(s.get(), in))) return rc;
        if ((rc = alloc_state<float>(s.get()))) return rc;
        if (s->autotune && (rc = autotune_tiles<float>(s.get(), opts->autotune_cache))) return rc;
    }
    *out = s.release();
    return PS2_OK;
}

}  // namespace

void batch_forget(const ps2_session *s) { batch_forget_impl(s); }

// ============================================================================
// C ABI

namespace {

// Runs job(q) for q = 0..n-1 on min(n, kBatchWorkers) host threads (each
// patch's loop is on its own stream, so at most that many loops are in
// flight).  After the first failure no new job starts; the failure with the
// lowest index is reported as "patch <base + q>: <message>".
constexpr size_t kBatchWorkers = 16;

template <typename Job>
int run_jobs(size_t n, Job job, size_t base = 0) {
    std::vector<int> rcs(n, PS2_OK);
    std::vector<std::string> errs(n);
    std::atomic<size_t> next{0};
    std::atomic<bool> failed{false};
    auto worker = [&]() {
        for (;;) {
            const size_t q = next.fetch_add(1);
            if (q >= n || failed.load()) return;
            rcs[q] = job(q);
            if (rcs[q]) {
                errs[q] = g_last_error;
                failed.store(true);
            }
        }
    };
    const size_t nw = std::min(n, kBatchWorkers);
    std::vector<std::thread> th;
    th.reserve(nw);
    for (size_t w = 0; w < nw; ++w) th.emplace_back(worker);
    for (auto &t : th) t.join();
    for (size_t q = 0; q < n; ++q)
        if (rcs[q]) {
            set_error("patch " + std::to_string(base + q) + ": " + errs[q]);
            return rcs[q];
        }
    return PS2_OK;
}

}  // namespace

extern "C" {

int ps2_hip_abi_version(void) { return PS2_HIP_ABI_VERSION; }

const char *ps2_hip_last_error(void) { return g_last_error.c_str(); }

int ps2_hip_device_count(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e == hipErrorNoDevice) return 0;
    if (e != hipSuccess) {
        set_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
        return PS2_E_DEVICE;
    }
    return n;
}

int ps2_hip_session_create(const ps2_patch_in *in, const ps2_run_opts *opts, ps2_session **out) {
    return session_create(in, opts, out);
}

int ps2_hip_session_run(ps2_session *s, int64_t *out_i_loop, double *out_loop_seconds) {
    if (!s) return arg_error("session is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    int rc = s->form == PS2_FORM_FP64 ? run_loop<double>(s) : run_loop<float>(s);
    if (rc) return rc;
    if (out_i_loop) *out_i_loop = s->i_loop;
    if (out_loop_seconds) *out_loop_seconds = s->loop_seconds;
    return PS2_OK;
}

int ps2_hip_session_run_iterations(ps2_session *s, ps2_patch_out *state, ps2_iteration_fn fn, void *user,
                                   int64_t *out_i_loop) {
    if (!s || !state) return arg_error("session and state are required");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    const int rc = s->form == PS2_FORM_FP64 ? run_loop_iterations<double>(s, state, fn, user)
                                             : run_loop_iterations<float>(s, state, fn, user);
    if (rc) return rc;
    if (out_i_loop) *out_i_loop = s->i_loop;
    return PS2_OK;
}

int ps2_hip_session_download(ps2_session *s, ps2_patch_out *out) {
    if (!s || !out) return arg_error("session or output is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->ran || !s->loop_outputs) {
        set_error("session_download before session_run (or after a single step)");
        return PS2_E_STATE;
    }
    PS2_HIP(hipSetDevice(s->device));
    return download_outputs(s, out, true);
}

void ps2_hip_session_destroy(ps2_session *s) { delete s; }

int ps2_hip_session_export_device(ps2_session *s, ps2_device_outputs *out) {
    if (!s || !out) return arg_error("session or output is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->ran) {
        set_error("session_export_device before session_run");
        return PS2_E_STATE;
    }
    if (!s->loop_outputs) {
        set_error("session_export_device after a single step: run the loop again");
        return PS2_E_STATE;
    }
    PS2_HIP(hipSetDevice(s->device));
    const int64_t P = s->P, I = s->I;
    PS2_HIP(hipStreamSynchronize(s->stream));
    out->device = s->device;
    out->n_ps = P;
    out->n_ifg = I;
    out->k_ps = s->o_vec.as<double>();
    out->c_ps = s->o_vec.as<double>() + P;
    out->coh_ps = s->o_vec.as<double>() + 2 * P;
    out->n_opt = s->o_vec.as<double>() + 3 * P;
    out->ph_patch = s->o_patch.as<float>();
    out->ph_res = s->o_res.as<float>();
    for (int b = 0; b < PS2_COH_BINS; ++b) out->nr[b] = s->nr[b];
    out->gamma_change_save = s->gamma_change_save;
    out->i_loop = s->i_loop;
    return PS2_OK;
}

int ps2_hip_session_tiles(ps2_session *s, ps2_tiles *out) {
    if (!s || !out) return arg_error("session or output is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    out->fill_cells = s->form == PS2_FORM_FP32 ? s->fill_cb : 1;
    out->scan_step_table = scan_step_table(s) ? 1 : 0;
    out->clap_segments = s->windows && !s->gclap ? s->segs.nseg : 0;
    out->clap_windows_per_segment = s->windows && !s->gclap ? s->segs.wseg : 0;
    out->source = s->tile_source;
    return PS2_OK;
}

int ps2_hip_session_set_timing(ps2_session *s, int32_t enable) {
    if (!s) return arg_error("session is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    if (enable && s->tev.empty()) {
        s->tev.resize(64, nullptr);
        for (auto &e : s->tev) PS2_HIP(hipEventCreate(&e));
    }
    s->timing = enable != 0;
    for (int k = 0; k < PS2_NUM_KERNEL_SLOTS; ++k) {
        s->slot_ms[k] = 0.0;
        s->slot_launches[k] = 0;
    }
    return PS2_OK;
}

int ps2_hip_session_kernel_times(ps2_session *s, double *ms, int64_t *launches) {
    if (!s || !ms || !launches) return arg_error("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    for (int k = 0; k < PS2_NUM_KERNEL_SLOTS; ++k) {
        ms[k] = s->slot_ms[k];
        launches[k] = s->slot_launches[k];
    }
    return PS2_OK;
}

int ps2_hip_run(const ps2_patch_in *in, ps2_patch_out *out, const ps2_run_opts *opts) {
    if (!out) return arg_error("ps2_patch_out is NULL");
    ps2_session *s = nullptr;
    int rc = session_create(in, opts, &s);
    if (rc) return rc;
    rc = ps2_hip_session_run(s, nullptr, nullptr);
    if (!rc) rc = ps2_hip_session_download(s, out);
    ps2_hip_session_destroy(s);
    return rc;
}

int ps2_hip_sessions_run(size_t n, ps2_session *const *sessions, int64_t *i_loops, double *seconds) {
    if (n && !sessions) return arg_error("sessions is NULL");
    for (size_t q = 0; q < n; ++q)
        if (!sessions[q]) return arg_error("session " + std::to_string(q) + " is NULL");
    std::vector<ps2_session *> ss(sessions, sessions + n);
    std::vector<ps2_session *> sorted(ss);
    std::sort(sorted.begin(), sorted.end());
    if (n >= 2 && std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end()) {
        // every loop in one merged launch sequence (locks taken in address order)
        std::vector<std::unique_lock<std::mutex>> locks;
        locks.reserve(n);
        for (auto *s : sorted) locks.emplace_back(s->mu);
        const int rc = sessions_run_merged(ss);
        if (rc != kNotMerged) {
            if (rc) return rc;
            for (size_t q = 0; q < n; ++q) {
                if (i_loops) i_loops[q] = ss[q]->i_loop;
                if (seconds) seconds[q] = ss[q]->loop_seconds;
            }
            return PS2_OK;
        }
    }
    return run_jobs(n, [&](size_t q) {
        int64_t il = 0;
        double sec = 0.0;
        const int rc = ps2_hip_session_run(sessions[q], &il, &sec);
        if (i_loops) i_loops[q] = il;
        if (seconds) seconds[q] = sec;
        return rc;
    });
}

int ps2_hip_run_batch(size_t n, const ps2_patch_in *in, ps2_patch_out *out, const ps2_run_opts *opts) {
    if (n && (!in || !out)) return arg_error("batch arrays are NULL");
    // waves of kBatchWorkers patches: upload all, run their loops together,
    // download all, free.  Device memory holds one wave, and no allocation,
    // free or blocking copy runs while a loop's graph is being captured.
    for (size_t w0 = 0; w0 < n; w0 += kBatchWorkers) {
        const size_t m = std::min(kBatchWorkers, n - w0);
        std::vector<ps2_session *> ss(m, nullptr);
        int rc = run_jobs(m, [&](size_t q) { return session_create(&in[w0 + q], opts, &ss[q]); }, w0);
        if (!rc) rc = ps2_hip_sessions_run(m, ss.data(), nullptr, nullptr);
        if (!rc)
            rc = run_jobs(m, [&](size_t q) { return ps2_hip_session_download(ss[q], &out[w0 + q]); }, w0);
        for (auto *x : ss) ps2_hip_session_destroy(x);
        if (rc) return rc;
    }
    return PS2_OK;
}

// --- single steps -------------------------------------------------------------

// host copy between the caller's [C][I] complex layout and the device's [I][C]
static int plane_upload(ps2_session *s, void *dev, const void *host) {
    const int64_t C = s->C, I = s->I;
    const size_t rs = s->rsz();
    std::vector<char> t((size_t)(2 * rs * C * I));
    for (int64_t c = 0; c < C; ++c)
        for (int64_t f = 0; f < I; ++f)
            std::memcpy(&t[2 * rs * (f * C + c)], (const char *)host + 2 * rs * (c * I + f), 2 * rs);
    PS2_HIP(h2d_now(s, dev, t.data(), t.size()));
    return PS2_OK;
}

static int plane_download(ps2_session *s, void *host, const void *dev) {
    const int64_t C = s->C, I = s->I;
    const size_t rs = s->rsz();
    std::vector<char> t((size_t)(2 * rs * C * I));
    PS2_HIP(hipMemcpyAsync(t.data(), dev, t.size(), hipMemcpyDeviceToHost, s->stream));
    PS2_HIP(hipStreamSynchronize(s->stream));
    for (int64_t c = 0; c < C; ++c)
        for (int64_t f = 0; f < I; ++f)
            std::memcpy((char *)host + 2 * rs * (c * I + f), &t[2 * rs * (f * C + c)], 2 * rs);
    return PS2_OK;
}

int ps2_hip_step_fill_grid(ps2_session *s, const double *k_ps, const double *weighting, void *ph_grid_out) {
    if (!s || !k_ps || !weighting || !ph_grid_out) return arg_error("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    std::vector<double> hk(s->P), hw(s->P);
    for (int64_t k = 0; k < s->P; ++k) {
        hk[k] = k_ps[s->order[k]];
        hw[k] = weighting[s->order[k]];
    }
    PS2_HIP(h2d_now(s, s->K.p, hk.data(), sizeof(double) * s->P));
    PS2_HIP(h2d_now(s, s->W.p, hw.data(), sizeof(double) * s->P));
    int rc = s->form == PS2_FORM_FP64 ? launch_fill_grid<double>(s, s->K.as<double>(), s->W.as<double>())
                                      : launch_fill_grid<float>(s, s->K.as<double>(), s->W.as<double>());
    if (rc) return rc;
    PS2_HIP(hipStreamSynchronize(s->stream));
    collect_timers(s);
    return plane_download(s, ph_grid_out, s->grid.p);
}

int ps2_hip_step_clap(ps2_session *s, const void *ph_grid, void *ph_filt_out) {
    if (!s || !ph_grid || !ph_filt_out) return arg_error("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    int rc = plane_upload(s, s->grid.p, ph_grid);
    if (rc) return rc;
    rc = s->form == PS2_FORM_FP64 ? launch_clap<double>(s, false) : launch_clap<float>(s, false);
    if (rc) return rc;
    PS2_HIP(hipStreamSynchronize(s->stream));
    collect_timers(s);
    return plane_download(s, ph_filt_out, s->filt.p);
}

int ps2_hip_step_fit(ps2_session *s, const void *ph_filt, ps2_patch_out *out) {
    if (!s || !ph_filt || !out) return arg_error("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    PS2_HIP(hipSetDevice(s->device));
    int rc = plane_upload(s, s->filt.p, ph_filt);
    if (rc) return rc;
    // the loop's filt planes hold unit phases (extract_patch_phase)
    const int64_t n = s->C * s->I;
    if (s->form == PS2_FORM_FP64)
        hipLaunchKernelGGL(k_unit_phase<double>, dim3(blocks_for(n, 256)), dim3(256), 0, s->stream,
                           s->filt.as<cplx<double>>(), n);
    else
        hipLaunchKernelGGL(k_unit_phase<float>, dim3(blocks_for(n, 256)), dim3(256), 0, s->stream,
                           s->filt.as<cplx<float>>(), n);
    PS2_LAUNCHED();
    rc = s->form == PS2_FORM_FP64 ? launch_fit<double>(s) : launch_fit<float>(s);
    if (rc) return rc;
    rc = s->form == PS2_FORM_FP64 ? launch_outputs<double>(s, false) : launch_outputs<float>(s, false);
    if (rc) return rc;
    s->loop_outputs = false;
    uint32_t cnt[2] = {0u, 0u};
    if (s->form == PS2_FORM_FP32)
        PS2_HIP(hipMemcpyAsync(cnt, s->rescore_count.p, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
    PS2_HIP(hipStreamSynchronize(s->stream));
    collect_timers(s);
    out->n_rescored = cnt[0];
    out->n_rechecked = cnt[1];
    return download_outputs(s, out, false);
}

}  // extern "C"
