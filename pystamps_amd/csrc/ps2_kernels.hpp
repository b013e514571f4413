// Stage-2 gamma-loop kernels for gfx950, templated on the real type R
// (float: FP32 form, double: plain FP64 form).
//
// Device layouts (set up once per patch by the host, see ps2_stage2.hip):
//   pixels are stored in cell-sorted order k (stable sort of grid_lin), so the
//   pixels of one grid cell are contiguous and each cell is a CSR range
//   cell_start[c] .. cell_start[c+1];
//   every per-(pixel, ifg) and per-(cell, ifg) array is ifg-major, one plane
//   per interferogram, so a wavefront of consecutive pixels (or cells) reads
//   consecutive addresses:
//   phnm  [I][P] complex R   ph_nm in sorted pixel order
//   bp    [I][P] R           bperp_mat (or [I] when the rows are invariant)
//   grid, filt [I][C] complex R
//   part  [I][nwi][nplanes][4][n_j][6] complex R  per-window-row CLAP
//         partial sums in groups of 6 window rows (one window step), each
//         group column-major; overlapping window segments in different planes
#pragma once

#include "ps2_device.hpp"

#include <type_traits>

namespace ps2 {

#ifndef PS2_CLAP_MEDIAN_HIST
#define PS2_CLAP_MEDIAN_HIST 1   // 1: histogram median (FP32 form), 0: bitwise search
#endif

constexpr int CLAP_FFT = 32;   // n_win_ex
constexpr int CLAP_WIN = 24;   // n_win
// Loop gate: the device-resident gamma loop launches every iteration's
// kernels up front; once its control kernel has taken the stop decision the
// flag is set and each loop kernel returns at entry.  nullptr = ungated
// (single-step use).
__device__ inline bool loop_stopped(const int32_t *gate) { return gate && *gate; }

constexpr int CLAP_INC = 6;    // n_inc = n_win / 4

// gfx950 dispatches workgroup b to XCD b % 8 (8 XCDs, each with its own L2)
#define PS2_NUM_XCD 8

// Kernel launches of the gamma loop take their arguments as one struct A.
// Launched for one patch (template flag M = false) a kernel takes A itself and
// uses its own workgroup index.  Launched once for up to PS2_PACK patches
// (M = true) it takes a MultiPack: workgroups range[2j] .. range[2j] +
// range[2j+1] - 1 run patch j with a[j] and local index b - range[2j].  The
// ranges start at multiples of 8, so a local index keeps the XCD of the global
// one; workgroups of the padding between ranges return at once.  The pack is
// a kernel argument, like A, so the pointers in it are known to address
// global memory (loaded from a table in memory they would be generic pointers
// and every access a flat one).
#define PS2_PACK 8
template <typename A>
struct MultiPack {
    A a[PS2_PACK];
    uint32_t range[2 * PS2_PACK];   // (first workgroup, workgroups), ascending
    int32_t n;
};
template <bool M, typename A>
using KArg = typename std::conditional<M, MultiPack<A>, A>::type;

template <typename A>
__device__ __forceinline__ bool multi_locate(const MultiPack<A> &p, A &a, uint32_t &bid, uint32_t &nblk) {
    const uint32_t b = blockIdx.x;
    int j = p.n - 1;
    while (j > 0 && p.range[2 * j] > b) --j;
    const uint32_t loc = b - p.range[2 * j];
    const uint32_t cnt = p.range[2 * j + 1];
    if (loc >= cnt) return false;
    a = p.a[j];
    bid = loc;
    nblk = cnt;
    return true;
}

// ---------------------------------------------------------------------------
// fill_phase_weight value (native_stage2.rs:648-652):
//   phase = -(bp * K); src * (cos, sin) * w   (all in R)
template <typename R>
__device__ inline cplx<R> weight_value(cplx<R> s, R b, double K, double W) {
    const R kk = (R)K, ww = (R)W;
    const R phase = -(b * kk);
    R sn, cs;
    RT<R>::sincos_(phase, &sn, &cs);
    const R re = s.re * cs - s.im * sn;
    const R im = s.re * sn + s.im * cs;
    return {re * ww, im * ww};
}

// w = (1 - prand_hi[round_half_away(1000 coh)])^2, the index clamped into
// the table (native_stage2.rs:1674-1682)
__device__ inline double psquare_weight(double coh, const double *__restrict__ prand_hi, int nhi) {
    const double v = coh * 1000.0;
    double r = v >= 0.0 ? floor(v + 0.5) : ceil(v - 0.5);
    const double hi = (double)(nhi > 0 ? nhi - 1 : 0);
    if (r < 0.0) r = 0.0;
    if (r > hi) r = hi;
    const int ix = __builtin_isnan(r) ? 0 : (int)r;
    const double p = prand_hi[ix];
    return (1.0 - p) * (1.0 - p);
}

// The weighting a fill applies: W[k], or, with P-square weighting in every
// iteration after the first (the control block's iteration counter), the
// weight of the pixel's coherence in the P-square table that the previous
// iteration's control step built, computed here in place of a weights
// kernel between the two iterations (psquare_weighting, native_stage2.rs:
// 1674-1682; the first iteration's W is the initial weighting).  Used for
// small patches, whose iterations are bound by launch latency: the fill's
// plane-group workgroups each compute a pixel's weight again, which costs a
// large patch more than the separate kernel.
struct FillWeights {
    const double *W;
    const double *coh, *prand_hi;   // null coh: W only
    const int32_t *i_loop;          // the control block's iteration counter (1-based)
    int nhi;
    __device__ bool from_coh() const { return coh && *i_loop > 1; }
    __device__ double at(int64_t k, bool fc) const { return fc ? psquare_weight(coh[k], prand_hi, nhi) : W[k]; }
};

// fill_phase_weight + accumulate_grid (native_stage2.rs:629-657, 722-740):
// one thread per (ifg, cell); the cell's pixels are summed in ascending row
// order, which is the reference's sequential order, so no atomics are needed.
// With Kprev/Wprev non-null, the K and weighting this fill used are kept
// there (for the ph_weight output of the last fill).
template <typename R>
__device__ __forceinline__ void fill_grid_body(
    const cplx<R> *__restrict__ phnm,
    const R *__restrict__ bp,
    int bp_rowinv,
    const double *__restrict__ K,
    FillWeights fw,
    const int32_t *__restrict__ cell_start,
    cplx<R> *__restrict__ grid,
    int64_t C,
    int I,
    int64_t P,
    double *__restrict__ Kprev,
    double *__restrict__ Wprev,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t gid = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (gid >= C * I) return;
    const bool fc = fw.from_coh();
    const int f = (int)(gid / C);
    const int64_t c = gid - (int64_t)f * C;
    const int32_t k0 = cell_start[c], k1 = cell_start[c + 1];
    const bool keep = Kprev && f == 0;
    R sre = (R)0, sim = (R)0;
    for (int32_t k = k0; k < k1; ++k) {
        const int64_t e = (int64_t)f * P + k;
        const R b = bp_rowinv ? bp[f] : bp[e];
        const double kk = K[k], ww = fw.at(k, fc);
        if (keep) {
            Kprev[k] = kk;
            Wprev[k] = ww;
        }
        const cplx<R> v = weight_value<R>(phnm[e], b, kk, ww);
        sre += v.re;
        sim += v.im;
    }
    grid[gid] = {sre, sim};
}

template <typename R>
struct FillArgs {
    const cplx<R> * phnm;
    const R * bp;
    int bp_rowinv;
    const double * K;
    FillWeights fw;
    const int32_t * cell_start;
    cplx<R> * grid;
    int64_t C;
    int I;
    int64_t P;
    double * Kprev;
    double * Wprev;
    const int32_t * gate;
};

template <bool M, typename R>
__global__ __launch_bounds__(256) void k_fill_grid(KArg<M, FillArgs<R>> p) {
    if constexpr (M) {
        FillArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        fill_grid_body<R>(a.phnm, a.bp, a.bp_rowinv, a.K, a.fw, a.cell_start, a.grid, a.C, a.I, a.P, a.Kprev, a.Wprev, a.gate, bid_, nblk_);
    } else {
        fill_grid_body<R>(p.phnm, p.bp, p.bp_rowinv, p.K, p.fw, p.cell_start, p.grid, p.C, p.I, p.P, p.Kprev, p.Wprev, p.gate, blockIdx.x, gridDim.x);
    }
}

// The same step with the phase weights computed cooperatively: a 256-thread
// workgroup covers CB consecutive cells and FG = 256 / CB interferogram
// planes; the weights of the cells' pixels (a contiguous sorted-pixel range)
// are computed with coalesced loads into LDS, CHUNK pixels per plane at a
// time, then each thread adds its cell's weights in ascending row order, so
// the sums are the same as k_fill_grid's.  CHUNK is sized to the mean pixel
// count of CB cells: LDS per workgroup bounds occupancy.  Kprev/Wprev as in
// k_fill_grid.
template <typename R, int CB, int CHUNK>
__device__ __forceinline__ void fill_grid_staged_body(
    const cplx<R> *__restrict__ phnm,
    const R *__restrict__ bp,
    int bp_rowinv,
    const double *__restrict__ K,
    FillWeights fw,
    const int32_t *__restrict__ cell_start,
    cplx<R> *__restrict__ grid,
    int64_t C,
    int I,
    int64_t P,
    int npg,
    double *__restrict__ Kprev,
    double *__restrict__ Wprev,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    constexpr int FG = 256 / CB;
    __shared__ cplx<R> vals[FG][CHUNK];
    const int t = threadIdx.x;
    // XCD-aware order: cell group cg belongs to XCD cg % 8, and its npg
    // plane groups are consecutive workgroups there, so the group's K, W and
    // cell_start reads are shared through that XCD's L2
    const int xcd = bid_ % PS2_NUM_XCD;
    const int loc = bid_ / PS2_NUM_XCD;
    const int cgl = loc / npg;
    const int pg = loc - cgl * npg;
    const int64_t c0 = ((int64_t)cgl * PS2_NUM_XCD + xcd) * CB;
    if (c0 >= C) return;
    const int f0 = pg * FG;
    const int64_t cend = c0 + CB < C ? c0 + CB : C;
    const int32_t ks = cell_start[c0], ke = cell_start[cend];
    const int64_t c = c0 + (t % CB);
    const int fs = t / CB;
    const int f = f0 + fs;
    const bool mine = c < C && f < I;
    const int32_t my0 = mine ? cell_start[c] : 0, my1 = mine ? cell_start[c + 1] : 0;
    const bool fc = fw.from_coh();
    R sre = (R)0, sim = (R)0;
    constexpr int NJ = CHUNK / 256;   // pixels per thread and chunk
    for (int32_t base = ks; base < ke; base += CHUNK) {
        const int n = ke - base < CHUNK ? ke - base : CHUNK;
        // every load of the thread's NJ x FG elements is issued before the
        // first weight is computed
        double kk[NJ], ww[NJ];
        cplx<R> sv[NJ][FG];
        R bv[NJ][FG];
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            const int j = t + u * 256;
            if (j < n) {
                const int64_t k = base + j;
                kk[u] = K[k];
                ww[u] = fw.at(k, fc);
                if (Kprev && pg == 0) {   // each pixel once: its cell group's first plane group
                    Kprev[k] = kk[u];
                    Wprev[k] = ww[u];
                }
#pragma unroll
                for (int q = 0; q < FG; ++q) {
                    const int ff = f0 + q;
                    if (ff < I) {
                        const int64_t e = (int64_t)ff * P + k;
                        bv[u][q] = bp_rowinv ? bp[ff] : bp[e];
                        sv[u][q] = phnm[e];
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NJ; ++u) {
            const int j = t + u * 256;
            if (j < n) {
#pragma unroll
                for (int q = 0; q < FG; ++q)
                    if (f0 + q < I) vals[q][j] = weight_value<R>(sv[u][q], bv[u][q], kk[u], ww[u]);
            }
        }
        __syncthreads();
        const int32_t lo = my0 > base ? my0 : base;
        const int32_t hi = my1 < base + n ? my1 : base + n;
        // ascending order; eight LDS reads in flight per step
        for (int32_t k = lo; k < hi; k += 8) {
            cplx<R> v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < hi) v[u] = vals[fs][k + u - base];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k + u < hi) {
                    sre += v[u].re;
                    sim += v[u].im;
                }
        }
        __syncthreads();
    }
    if (mine) grid[(int64_t)f * C + c] = {sre, sim};
}

template <typename R>
struct FillStagedArgs {
    const cplx<R> * phnm;
    const R * bp;
    int bp_rowinv;
    const double * K;
    FillWeights fw;
    const int32_t * cell_start;
    cplx<R> * grid;
    int64_t C;
    int I;
    int64_t P;
    int npg;
    double * Kprev;
    double * Wprev;
    const int32_t * gate;
};

template <bool M, typename R, int CB, int CHUNK>
__global__ __launch_bounds__(256) void k_fill_grid_staged(KArg<M, FillStagedArgs<R>> p) {
    if constexpr (M) {
        FillStagedArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        fill_grid_staged_body<R, CB, CHUNK>(a.phnm, a.bp, a.bp_rowinv, a.K, a.fw, a.cell_start, a.grid, a.C, a.I, a.P, a.npg, a.Kprev, a.Wprev, a.gate, bid_, nblk_);
    } else {
        fill_grid_staged_body<R, CB, CHUNK>(p.phnm, p.bp, p.bp_rowinv, p.K, p.fw, p.cell_start, p.grid, p.C, p.I, p.P, p.npg, p.Kprev, p.Wprev, p.gate, blockIdx.x, gridDim.x);
    }
}

// Outputs in the caller's row order [P][I] from sorted ifg-major planes [I][P]:
// a workgroup takes TP sorted pixels x TF interferograms (TP x TF = 2048 or
// 4096), reads the planes along the pixel axis, and writes each pixel's run of
// interferograms as one contiguous row segment (order[k] scatters whole rows,
// not single elements).  TF covers the whole row where I <= 256, so no row is
// split between workgroups (a split leaves partial cache lines in the middle
// of rows: 2.7 TB/s at 4M x 120 with 64-wide tiles).

// ---------------------------------------------------------------------------
// CLAP (native_stage2.rs:742-1184).
//
// One workgroup per (window row ix, group of TEAMS interferograms, segment of
// wseg consecutive windows of the row).  A team of 32 lanes (half a wavefront)
// owns one interferogram and walks its segment's windows left to right; each
// window is a 32x32 FFT held one row / one column per lane in registers,
// transposed through LDS.  The weighted window outputs of the segment are
// summed per team in a rolling 24-column register accumulator; every column
// block that no later window of the segment touches is flushed to `part`
// (segment s in plane s % nplanes: a segment's columns reach 18 past its
// windows' start columns, so it overlaps the next ceil(3 / wseg) segments,
// and nplanes = 1 + ceil(3 / wseg) keeps overlapping segments on different
// planes).  A second kernel adds, in a fixed order, the (at most four) window
// rows and (at most nplanes) segments covering each cell.  No atomics, so the
// result is deterministic.

template <typename R>
struct ClapParams {
    R k7[7];        // gausswin(7, 2.5) smoothing taps (native_stage2.rs:1044-1053)
    R alpha, beta;  // clap_alpha, clap_beta
    int alpha_is_one;
    int ni, nj, I, nwi, nwj, ngroups;
    int nseg, wseg;   // window segments per row, windows per segment
    int nplanes;      // partial-sum planes per window row: 1 + ceil(3 / wseg)
    int units;        // strips: ngroups * nseg (interferogram pair x segment)
    int upx;          // strips per XCD: ceil(units / PS2_NUM_XCD)
};

// Wave-level LDS synchronisation: a team is half a wavefront and never shares
// LDS with another wavefront, and LDS operations of one wavefront complete in
// issue order, so ordering the compiler's memory operations at wavefront scope
// is enough between a team's LDS writes and its reads.
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Count over the calling lane's team (32 lanes) of keys[r] <= t, with t uniform
// within the team.  One ballot per register; the popcounts run on the scalar
// unit.
template <typename K>
__device__ inline int team_count_le(const K (&keys)[32], K t, bool hi_team) {
    int lo = 0, hi = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        const unsigned long long m = __ballot(keys[r] <= t);
        lo += __popc((unsigned)m);
        hi += __popc((unsigned)(m >> 32));
    }
    return hi_team ? hi : lo;
}

// Keys of non-negative reals: the IEEE bit pattern with the sign cleared
// orders exactly like the value, and converts back without loss.
template <typename R>
__device__ inline typename RT<R>::key_t median_key(R v) {
    using K = typename RT<R>::key_t;
    constexpr K signmask = ~(K(1) << (sizeof(K) * 8 - 1));
    return RT<R>::key(v) & signmask;
}

// median of the 1024 non-negative values held (as keys) 32 per lane by a team
// (native_stage2.rs:1168-1184: sort, mean of the two middle values).  Bitwise
// selection of the 0-based 511th key, starting below the team's common
// min/max key prefix; the 512th is that key again or the next larger one.
template <typename R>
__device__ inline R team_median_bits(const typename RT<R>::key_t (&keys)[32], bool hi_team) {
    using K = typename RT<R>::key_t;
    constexpr int NB = (int)sizeof(K) * 8;
    K kmin = ~K(0), kmax = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        kmin = keys[r] < kmin ? keys[r] : kmin;
        kmax = keys[r] > kmax ? keys[r] : kmax;
    }
    kmin = team_min_key<K>(kmin);
    kmax = ~team_min_key<K>(~kmax);
    const K diff = kmin ^ kmax;
    int top = -1;
    if (diff) top = NB - 1 - (int)(sizeof(K) == 4 ? __clz((unsigned)diff) : __clzll((long long)diff));
    K ans = top >= 0 ? (kmin & ~((K(2) << top) - 1)) : kmin;
    for (int bit = top; bit >= 0; --bit) {
        const K t = ans | ((K(1) << bit) - 1);
        if (team_count_le<K>(keys, t, hi_team) < 512) ans |= K(1) << bit;
    }
    K nxt = ans;
    if (team_count_le<K>(keys, ans, hi_team) < 513) {
        K mn = ~K(0);
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if (keys[r] > ans && keys[r] < mn) mn = keys[r];
        nxt = team_min_key<K>(mn);
    }
    return (RT<R>::from_key(ans) + RT<R>::from_key(nxt)) / (R)2;
}

// 7-tap Gaussian at position c of a zero-extended 32-vector; the taps are
// symmetric (gausswin), so pairs are added before the multiply.
template <typename R>
__device__ inline R smooth7(const R (&h)[32], int c, const R (&k)[7]) {
    auto at = [&](int q) { return (q >= 0 && q < 32) ? h[q] : (R)0; };
    R s = h[c] * k[3];
    s = fma(at(c - 1) + at(c + 1), k[2], s);
    s = fma(at(c - 2) + at(c + 2), k[1], s);
    s = fma(at(c - 3) + at(c + 3), k[0], s);
    return s;
}

// Exact median for the FP32 form: a 1024-bucket histogram of the keys
// (bucket = (key - kmin) >> shift, monotone in the value; float keys are
// close to logarithmic, so the buckets stay narrow near the median) built with
// LDS atomics in the team's tile, then exact selection among the few keys of
// the bucket holding rank 511 (rank 512 is in the same bucket or is the
// smallest key above it).  A crowded bucket falls back to the bitwise search.
// Bucket b = 32 L + q lives at word 36 L + q: lane L owns buckets
// [32L, 32L + 32), a 16-byte aligned run that it reads with 8 four-word loads
// for its total (the 36-word rows put 16 lanes' runs on 64 distinct banks),
// and the owner's bucket row (lane q reads 36 owner + q) is contiguous.  The
// region is cleared with four-word stores.  hist (>= 1152 words, 16-byte
// aligned) and cand (>= 64 words, distinct from hist) are the team's LDS
// scratch.
template <typename R>
__device__ inline R team_median(const typename RT<R>::key_t (&keys)[32], bool hi_team, R *hist_mem,
                                R *cand_mem) {
    if constexpr (sizeof(R) != 4) {
        return team_median_bits<R>(keys, hi_team);
    } else {
        const int lane = threadIdx.x & 31;
        uint32_t *hist = reinterpret_cast<uint32_t *>(hist_mem);
        uint32_t *cand = reinterpret_cast<uint32_t *>(cand_mem);
        uint32_t kmin = 0xffffffffu, kmax = 0u;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            kmin = keys[r] < kmin ? keys[r] : kmin;
            kmax = keys[r] > kmax ? keys[r] : kmax;
        }
        kmin = team_min_key<uint32_t>(kmin);
        kmax = ~team_min_key<uint32_t>(~kmax);
        if (kmin == kmax) return __uint_as_float(kmin);
        const uint32_t span = kmax - kmin;
        const int bits = 32 - __clz((int)span);
        const int shift = bits > 10 ? bits - 10 : 0;
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        u4 *hist4 = reinterpret_cast<u4 *>(hist);
#pragma unroll
        for (int q = 0; q < 9; ++q) hist4[q * 32 + lane] = u4{0u, 0u, 0u, 0u};
        wave_sync();
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t bk = (keys[r] - kmin) >> shift;
            atomicAdd(&hist[bk + (bk >> 5) * 4u], 1u);
        }
        wave_sync();
        // per-lane totals, team inclusive scan, owner of rank 511
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const u4 h = hist4[lane * 9 + q];
            tot += (h.x + h.y) + (h.z + h.w);
        }
        const uint32_t incl = team_incl_sum_u32(tot);
        const uint32_t excl = incl - tot;
        int ol, oh;
        team_first_lanes(__ballot(excl <= 511u && 511u < incl), ol, oh);
        const int owner = (hi_team ? oh : ol) & 31;
        const uint32_t base = team_pick_u32(excl, ol, oh, hi_team);
        // the owner's 32 buckets, one per lane, scanned the same way
        const uint32_t c = hist[owner * 36 + lane];
        const uint32_t run = team_incl_sum_u32(c) + base;
        int sl, sh;
        team_first_lanes(__ballot(run > 511u), sl, sh);
        const int qs = (hi_team ? sh : sl) & 31;
        const uint32_t n511 = team_pick_u32(c, sl, sh, hi_team);
        const uint32_t below511 = team_pick_u32(run, sl, sh, hi_team) - n511;
        const uint32_t b511 = (uint32_t)(owner * 32 + qs);
        if (n511 > 64u) return team_median_bits<R>(keys, hi_team);
        // bucket b511 holds the keys lo <= k < lo + width
        const uint32_t lo = kmin + (b511 << shift), width = 1u << shift;
        // compact the keys of bucket b511 into cand, in team order
        uint32_t mine = 0;
#pragma unroll
        for (int r = 0; r < 32; ++r) mine += keys[r] - lo < width ? 1u : 0u;
        uint32_t off = team_incl_sum_u32(mine) - mine;
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if (keys[r] - lo < width) cand[off++] = keys[r];
        wave_sync();
        // rank of the candidates q = lane and q = lane + 32 within the bucket
        // (ties by position), eight reads in flight per step
        const uint32_t j0 = 511u - below511;
        const uint32_t qa = (uint32_t)lane, qb = (uint32_t)lane + 32u;
        const uint32_t xa = qa < n511 ? cand[qa] : 0xffffffffu;
        const uint32_t xb = qb < n511 ? cand[qb] : 0xffffffffu;
        uint32_t ra = 0, rb = 0;
        for (uint32_t m0 = 0; m0 < n511; m0 += 8u) {
            uint32_t y[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) y[u] = cand[m0 + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t m = m0 + (uint32_t)u;
                const bool in = m < n511;
                ra += (in && (y[u] < xa || (y[u] == xa && m < qa))) ? 1u : 0u;
                rb += (in && (y[u] < xb || (y[u] == xb && m < qb))) ? 1u : 0u;
            }
        }
        uint32_t v0 = 0xffffffffu, v1 = 0xffffffffu;
        if (qa < n511 && ra == j0) v0 = xa;
        if (qb < n511 && rb == j0) v0 = xb;
        if (qa < n511 && ra == j0 + 1u) v1 = xa;
        if (qb < n511 && rb == j0 + 1u) v1 = xb;
        const uint32_t k511 = team_min_key<uint32_t>(v0);
        uint32_t k512 = team_min_key<uint32_t>(v1);
        if (j0 + 1u >= n511) {
            uint32_t mn = 0xffffffffu;
            const uint32_t hi = lo + width;
#pragma unroll
            for (int r = 0; r < 32; ++r)
                if (keys[r] >= hi && keys[r] < mn) mn = keys[r];
            k512 = team_min_key<uint32_t>(mn);
        }
        return (__uint_as_float(k511) + __uint_as_float(k512)) / 2.0f;
    }
}

template <typename R, int TEAMS, bool ALPHA_ONE>
__device__ __forceinline__ void clap_strip_body(
    const cplx<R> *__restrict__ grid,
    cplx<R> *__restrict__ part,
    const R *__restrict__ lp,
    ClapParams<R> prm,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int team = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const bool hi_team = (threadIdx.x & 63) >= 32;
    // XCD-aware order: workgroup b runs on XCD b % 8.  Strip u (interferogram
    // pair g, segment seg) belongs to XCD u % 8, which walks its strips' window
    // rows top to bottom, so the 4 window rows that read a grid row run close
    // together on the XCD whose L2 holds it.
    const int xcd = bid_ % PS2_NUM_XCD;
    const int loc = bid_ / PS2_NUM_XCD;
    const int ix = loc / prm.upx;
    const int u = (loc - ix * prm.upx) * PS2_NUM_XCD + xcd;
    if (u >= prm.units) return;
    const int g = u / prm.nseg;
    const int seg = u - g * prm.nseg;
    const int f = g * TEAMS + team;
    const int jx0 = seg * prm.wseg;
    const int jx1 = jx0 + prm.wseg < prm.nwj ? jx0 + prm.wseg : prm.nwj;
    const bool active = f < prm.I;
    const int ni = prm.ni, nj = prm.nj;
    int i1 = ix * CLAP_INC, rs = 0;
    if (i1 + CLAP_WIN > ni) {
        rs = i1 + CLAP_WIN - ni;
        i1 = ni - CLAP_WIN;
    }
    // separate re / im planes; a row of 33 keeps row and column walks on
    // distinct banks.  One array per team: re plane, im plane, then 32 words
    // of padding so that the other team's planes start half the banks apart;
    // 16-byte aligned for the median's four-word histogram access.
    constexpr int TILE = 2 * 32 * 33 + 128 / (int)sizeof(R);
    __shared__ __attribute__((aligned(16))) R tile_s[TEAMS][TILE];
    R (&tbr)[32][33] = *reinterpret_cast<R (*)[32][33]>(&tile_s[team][0]);
    R (&tbi)[32][33] = *reinterpret_cast<R (*)[32][33]>(&tile_s[team][32 * 33]);
    const int cS = (lane + 16) & 31;  // fftshift of this lane's column
    // rolling accumulator of output row lr = lane - rs: slot e holds column
    // 6*jx + e of the current window jx (native_stage2.rs:904-915)
    R ar[CLAP_WIN], ai[CLAP_WIN];
#pragma unroll
    for (int e = 0; e < CLAP_WIN; ++e) ar[e] = ai[e] = (R)0;
    const bool out_row = active && lane < CLAP_WIN && lane >= rs;
    const int lr = lane - rs;
    const cplx<R> *gsrc = grid + (int64_t)f * ni * nj + (int64_t)(i1 + lane) * nj;
    const int lrc = lr < 0 ? 0 : (lr > CLAP_WIN - 1 ? CLAP_WIN - 1 : lr);
    // window weight of output row lrc (native_stage2.rs:1014-1037): r + c +
    // 1e-6 on a quadrant, mirrored, so with a(x) = min(x, 23 - x) the weight
    // is a(row) + a(col) + 1e-6, here with the inverse FFT's 1/32^2 folded in.
    // The integer sum is exact; in f32, k + fl32(1e-6) rounds like the f64
    // value k + 1e-6 for every k in [0, 22], so this equals the f64 table
    // rounded to R.
    R wv[CLAP_WIN / 2];
    {
        const int arow = lrc < CLAP_WIN / 2 ? lrc : CLAP_WIN - 1 - lrc;
        const R bias = std::is_same<R, float>::value ? (R)1.0e-6f : (R)1.0e-6;
#pragma unroll
        for (int q = 0; q < CLAP_WIN / 2; ++q)
            wv[q] = ((R)(arow + q) + bias) * (R)(1.0 / (CLAP_FFT * CLAP_FFT));
    }
    // plane of 6-row groups, each column-major: a flush of 6 columns writes one
    // contiguous run of 36 entries per group
    cplx<R> *dst = part + (((int64_t)f * prm.nwi + ix) * prm.nplanes + seg % prm.nplanes) * (int64_t)nj * CLAP_WIN +
                   (int64_t)(lrc / CLAP_INC) * nj * CLAP_INC + lrc % CLAP_INC;

    for (int jx = jx0; jx < jx1; ++jx) {
        // re-read the low-pass table (L1-resident) every window instead of
        // holding 32 loop-invariant values in registers
        const R *lpg = lp;
        asm volatile("" : "+s"(lpg));
        // the asm hides the address space; restore it so these stay global
        // (not flat) loads, which do not count against the LDS wait counter
        const __attribute__((address_space(1))) R *lpw =
            (const __attribute__((address_space(1))) R *)lpg;
        int j1 = jx * CLAP_INC, cs = 0;
        if (j1 + CLAP_WIN > nj) {
            cs = j1 + CLAP_WIN - nj;
            j1 = nj - CLAP_WIN;
        }
        R xr[32], xi[32];
        // window row `lane`, zero padded to 32 (NaN -> 0, native_stage2.rs:852-858)
        if (lane < CLAP_WIN && active) {
            // two cells per 16-byte load (rows are 8-byte aligned; gfx950
            // global loads accept that)
            typedef R pair4 __attribute__((ext_vector_type(4), aligned(sizeof(cplx<R>))));
            const pair4 *src = reinterpret_cast<const pair4 *>(gsrc + j1);
#pragma unroll
            for (int c = 0; c < CLAP_WIN; c += 2) {
                const pair4 v = src[c / 2];
                const bool z0 = RT<R>::isnan_(v.x) || RT<R>::isnan_(v.y);
                const bool z1 = RT<R>::isnan_(v.z) || RT<R>::isnan_(v.w);
                xr[c] = z0 ? (R)0 : v.x;
                xi[c] = z0 ? (R)0 : v.y;
                xr[c + 1] = z1 ? (R)0 : v.z;
                xi[c + 1] = z1 ? (R)0 : v.w;
            }
#pragma unroll
            for (int c = CLAP_WIN; c < 32; ++c) xr[c] = xi[c] = (R)0;
            fft32<R, false, CLAP_WIN>(xr, xi);
        } else {
#pragma unroll
            for (int c = 0; c < 32; ++c) xr[c] = xi[c] = (R)0;
        }
        wave_sync();  // the previous window's readers of the tile are done
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            tbr[lane][c] = xr[c];
            tbi[lane][c] = xi[c];
        }
        wave_sync();
        // rows 24..31 of the row transforms are zero (padding rows)
#pragma unroll
        for (int r = 0; r < CLAP_WIN; ++r) {
            xr[r] = tbr[r][lane];
            xi[r] = tbi[r][lane];
        }
#pragma unroll
        for (int r = CLAP_WIN; r < 32; ++r) xr[r] = xi[r] = (R)0;
        fft32<R, false, CLAP_WIN>(xr, xi);  // lane c now holds column c of FFT2
        wave_sync();
        // |F| at fftshifted position (re plane)
#pragma unroll
        for (int r = 0; r < 32; ++r) tbr[(r + 16) & 31][cS] = RT<R>::abs_(xr[r], xi[r]);
        wave_sync();
        {   // separable 7-tap smoothing, zero fill: rows (lane = shifted row) -> im plane
            R h[32];
#pragma unroll
            for (int c = 0; c < 32; ++c) h[c] = tbr[lane][c];
#pragma unroll
            for (int c = 0; c < 32; ++c) tbi[lane][c] = smooth7<R>(h, c, prm.k7);
        }
        wave_sync();
        // columns: lane c takes its shifted column cS and smooths it in
        // registers; ifftshift back, H[r][c] = Hf[(r+16)%32][cS], kept as
        // median keys (exact: H >= 0)
        typename RT<R>::key_t hk[32];
        {
            R t[32];
#pragma unroll
            for (int r = 0; r < 32; ++r) t[r] = tbi[r][cS];
#pragma unroll
            for (int r = 0; r < 32; ++r) hk[r] = median_key<R>(smooth7<R>(t, (r + 16) & 31, prm.k7));
        }
        wave_sync();  // the tile is scratch for the median from here to the gain
#if PS2_CLAP_MEDIAN_HIST
        const R med = team_median<R>(hk, hi_team, &tile_s[team][0], &tile_s[team][1152]);
#else
        const R med = team_median_bits<R>(hk, hi_team);
#endif
        const R scl = RT<R>::gain_scale(med);
        // low-pass column `lane` (the table is stored transposed): 4 values per load
        typedef R lp4 __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) lp4 *lpc =
            (const __attribute__((address_space(1))) lp4 *)(lpw + lane * 32);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const lp4 l4 = lpc[q];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = 4 * q + u;
                R v = ALPHA_ONE ? RT<R>::gain_rel_minus_one(RT<R>::from_key(hk[r]), med, scl)
                                : RT<R>::pow_(RT<R>::gain_div(RT<R>::from_key(hk[r]), med, scl), prm.alpha) - (R)1;
                if (v < (R)0) v = (R)0;
                const R gsc = fma(v, prm.beta, l4[u]);
                xr[r] *= gsc;
                xi[r] *= gsc;
            }
        }
        // inverse along columns; only the window's 24 output rows go on
        fft32<R, true, 32, CLAP_WIN>(xr, xi);
        wave_sync();
#pragma unroll
        for (int r = 0; r < CLAP_WIN; ++r) {
            tbr[r][lane] = xr[r];
            tbi[r][lane] = xi[r];
        }
        wave_sync();
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            xr[c] = tbr[lane][c];
            xi[c] = tbi[lane][c];
        }
        fft32<R, true, 32, CLAP_WIN>(xr, xi);  // inverse along rows: spatial row `lane`
        const bool last = jx == jx1 - 1;
        if (cs == 0) {
            if (out_row) {
#pragma unroll
                for (int e = 0; e < CLAP_WIN; ++e) {
                    const R fct = wv[e < CLAP_WIN / 2 ? e : CLAP_WIN - 1 - e];
                    ar[e] = fma(xr[e], fct, ar[e]);
                    ai[e] = fma(xi[e], fct, ai[e]);
                }
            }
        } else {
            // clamped last window column: effective column e = lj - cs; the
            // shift is dynamic, so it goes through the team's LDS tile
            wave_sync();
#pragma unroll
            for (int c = 0; c < CLAP_WIN; ++c) {
                tbr[lane][c] = xr[c];
                tbi[lane][c] = xi[c];
            }
            wave_sync();
            if (out_row) {
#pragma unroll
                for (int e = 0; e < CLAP_WIN; ++e) {
                    if (e < CLAP_WIN - cs) {
                        const R fct = wv[e < CLAP_WIN / 2 ? e : CLAP_WIN - 1 - e];
                        ar[e] += tbr[lane][e + cs] * fct;
                        ai[e] += tbi[lane][e + cs] * fct;
                    }
                }
            }
        }
        // flush the columns no later window of this row touches, then shift
        const int cbeg = jx * CLAP_INC;
        if (out_row) {
            if (!last) {
#pragma unroll
                for (int e = 0; e < CLAP_INC; ++e) dst[(int64_t)(cbeg + e) * CLAP_INC] = {ar[e], ai[e]};
            } else {
#pragma unroll
                for (int e = 0; e < CLAP_WIN; ++e)
                    if (cbeg + e < nj) dst[(int64_t)(cbeg + e) * CLAP_INC] = {ar[e], ai[e]};
            }
        }
#pragma unroll
        for (int e = 0; e < CLAP_WIN - CLAP_INC; ++e) {
            ar[e] = ar[e + CLAP_INC];
            ai[e] = ai[e + CLAP_INC];
        }
#pragma unroll
        for (int e = CLAP_WIN - CLAP_INC; e < CLAP_WIN; ++e) ar[e] = ai[e] = (R)0;
    }
}

template <typename R>
struct StripArgs {
    const cplx<R> * grid;
    cplx<R> * part;
    const R * lp;
    ClapParams<R> prm;
    const int32_t * gate;
};

// two wavefronts per SIMD in the FP32 form (254 VGPRs); the FP64 form's
// registers allow one
template <bool M, typename R, int TEAMS, bool ALPHA_ONE>
__global__ __launch_bounds__(32 * TEAMS) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 2 : 1, 8))) void k_clap_strip(KArg<M, StripArgs<R>> p) {
    if constexpr (M) {
        StripArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        clap_strip_body<R, TEAMS, ALPHA_ONE>(a.grid, a.part, a.lp, a.prm, a.gate, bid_, nblk_);
    } else {
        clap_strip_body<R, TEAMS, ALPHA_ONE>(p.grid, p.part, p.lp, p.prm, p.gate, blockIdx.x, gridDim.x);
    }
}

// unit phase of a complex value with the reference's f32/f64 arithmetic
// (normalize_complex_unit_magnitude, native_stage2.rs:1198-1205)
template <typename R>
__device__ inline cplx<R> unit_phase(cplx<R> v) {
    const R mag = RT<R>::hypot_(v.re, v.im);
    if (mag != (R)0) {
        v.re /= mag;
        v.im /= mag;
    }
    return v;
}

// Sum of the window rows covering each cell (ascending row order).  With
// UNIT the cell's unit phase is stored instead: that is ph_patch of every
// pixel in the cell (extract_patch_phase gathers then normalises each copy of
// the same value, native_stage2.rs:1186-1196).  Workgroup = a tile of
// COMB_TR = 24 rows x COMB_TC columns of one interferogram; every 6-row group
// of the tile lies in one 6-row group of each of the (at most 4) window rows
// over it, a contiguous run of 6 x COMB_TC entries (the planes hold 6-row
// groups column-major), and lanes take (group row, column, group) fastest
// first, so each load reads whole runs.  Each cell adds its (at most
// 4 x nplanes) entries in the fixed order window row ascending, segment
// ascending; the sums go through an LDS tile and are stored row-major.
// Grid: tiles x I, the interferogram slowest.
constexpr int COMB_TR = 4 * CLAP_INC, COMB_TC = 32;
template <typename R, bool UNIT, int MAXSEG>
__device__ __forceinline__ void clap_combine_body(
    const cplx<R> *__restrict__ part,
    cplx<R> *__restrict__ filt,
    int ni,
    int nj,
    int I,
    int nwi,
    int nseg,
    int wseg,
    int nplanes,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    static_assert(COMB_TR * COMB_TC % 256 == 0, "whole tile per pass");
    __shared__ cplx<R> tile[COMB_TC][COMB_TR + 1];   // [column][row]
    const int tj_n = (nj + COMB_TC - 1) / COMB_TC;
    const int tiles = ((ni + COMB_TR - 1) / COMB_TR) * tj_n;
    const int f = (int)(bid_ / tiles), tl = (int)(bid_ - (unsigned)f * tiles);
    const int r0 = (tl / tj_n) * COMB_TR, c0 = (tl % tj_n) * COMB_TC;
    const int64_t C = (int64_t)ni * nj;
    const cplx<R> *pf = part + (int64_t)f * nwi * nplanes * nj * CLAP_WIN;
    // segments whose columns [span s, span s + span + 18) contain c: s0..s1,
    // at most nplanes <= 4 of them
    const int span = CLAP_INC * wseg;
#pragma unroll
    for (int h = 0; h < COMB_TR * COMB_TC / 256; ++h) {
        const int e = (int)threadIdx.x + 256 * h;
        const int r6 = e % CLAP_INC, rest = e / CLAP_INC;
        const int cl = rest % COMB_TC, rl = (rest / COMB_TC) * CLAP_INC + r6;
        const int r = r0 + rl, c = c0 + cl;
        cplx<R> out{(R)0, (R)0};
        if (r < ni && c < nj) {
            int ix_hi = r / CLAP_INC;
            if (ix_hi > nwi - 1) ix_hi = nwi - 1;
            const int ix_lo = r >= CLAP_WIN - 1 ? (r - (CLAP_WIN - 1) + CLAP_INC - 1) / CLAP_INC : 0;
            int s1 = c / span;
            if (s1 > nseg - 1) s1 = nseg - 1;
            int s0;
            if constexpr (MAXSEG == 2) {   // wseg >= 3: s1, and s1 - 1 inside s1's first 18 columns
                s0 = s1 > 0 && c < s1 * span + (CLAP_WIN - CLAP_INC) ? s1 - 1 : s1;
            } else {
                const int d = c - span - (CLAP_WIN - CLAP_INC);
                s0 = d < 0 ? 0 : d / span + 1;
            }
            // at most 4 window rows (ix_hi - ix_lo <= 3) x MAXSEG (>= nplanes)
            // segments; the 6 rows of a group share ix_lo, so for each q the
            // group's lanes read one run.  Every load is issued before the
            // first add.
            cplx<R> v[4][MAXSEG];
            bool ok[4][MAXSEG];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ix = ix_lo + q;
                const int lr = r - ix * CLAP_INC;
#pragma unroll
                for (int t = 0; t < MAXSEG; ++t) {
                    const int sg = s0 + t;
                    ok[q][t] = ix <= ix_hi && sg <= s1;
                    const int plane = MAXSEG == 2 ? (sg & 1) : sg % nplanes;
                    // (an address inside the planes also where the entry is not used;
                    // one interferogram's planes hold fewer than 2^31 entries, a
                    // host-checked size, so the offset is 32-bit arithmetic)
                    const int at =
                        ok[q][t] ? (((ix * nplanes + plane) * (CLAP_WIN / CLAP_INC) + lr / CLAP_INC) * nj + c) *
                                           CLAP_INC + r6
                                 : 0;
                    v[q][t] = ok[q][t] ? pf[at] : cplx<R>{(R)0, (R)0};
                }
            }
            R sre = (R)0, sim = (R)0;
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int t = 0; t < MAXSEG; ++t)
                    if (ok[q][t]) {
                        sre += v[q][t].re;
                        sim += v[q][t].im;
                    }
            out = UNIT ? unit_phase<R>({sre, sim}) : cplx<R>{sre, sim};
        }
        tile[cl][rl] = out;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < COMB_TR * COMB_TC / 256; ++h) {
        const int e = (int)threadIdx.x + 256 * h;
        const int rl = e / COMB_TC, cl = e - rl * COMB_TC;
        const int r = r0 + rl, c = c0 + cl;
        if (r < ni && c < nj) filt[(int64_t)f * C + (int64_t)r * nj + c] = tile[cl][rl];
    }
}

template <typename R>
struct CombineArgs {
    const cplx<R> * part;
    cplx<R> * filt;
    int ni;
    int nj;
    int I;
    int nwi;
    int nseg;
    int wseg;
    int nplanes;
    const int32_t * gate;
};

template <bool M, typename R, bool UNIT, int MAXSEG>
__global__ __launch_bounds__(256) void k_clap_combine(KArg<M, CombineArgs<R>> p) {
    if constexpr (M) {
        CombineArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        clap_combine_body<R, UNIT, MAXSEG>(a.part, a.filt, a.ni, a.nj, a.I, a.nwi, a.nseg, a.wseg, a.nplanes, a.gate, bid_, nblk_);
    } else {
        clap_combine_body<R, UNIT, MAXSEG>(p.part, p.filt, p.ni, p.nj, p.I, p.nwi, p.nseg, p.wseg, p.nplanes, p.gate, blockIdx.x, gridDim.x);
    }
}

// --- CLAP for any window size (clap_win other than 32) ------------------------
// The same filter (native_stage2.rs:742-1184) for an n_win x n_win window in
// an n x n transform, n = n_win_ex = round(clap_win) up to GCLAP_MAX and
// n_win = round(0.75 clap_win), window step n_inc = max(n_win / 4, 1):
// one 256-thread workgroup per (window, interferogram) with the window in LDS;
// direct n-point DFTs (twiddles e^{-2 pi i m / n} in R from f64 values), the
// reference's fftshift by floor(n / 2) and ifftshift by ceil(n / 2), the
// separable 7-tap smoothing with zero fill, the exact median (bitonic sort of
// the n^2 gains, mean of the two middle values for an even count), the gain
// and low-pass, the inverse DFT, and the window weight times 1 / n^2; the
// weighted window goes to its own slot of `gpart`.  k_gclap_combine then adds,
// per cell, the windows over it in window order (row of windows ascending,
// then column), in f64 as the reference's accumulator, and rounds once to R.
// NMAX (16, 32 or 64) sizes the LDS; kernels of the default window (24 of 32)
// are the strip kernels above.
constexpr int GCLAP_MAX = 64;

template <typename R>
struct GClapArgs {
    const cplx<R> *grid;     // [I][ni][nj]
    cplx<R> *gpart;          // [I][nwi * nwj][n_win][n_win]
    const R *lp;             // low_pass [n][n], the reference's layout
    const double *weight;    // [n_win][n_win] window weight / n^2
    int ni, nj, I, nwi, nwj, n_win, n, n_inc;
    R alpha, beta;
    R k7[7];
    const int32_t *gate;
};

// one pass of an n-point DFT over an n x n plane held in LDS: along rows
// (dst[r][k] = sum_m src[r][m] w^(km)) or along columns (dst[k][c] = sum_m
// src[m][c] w^(km)), m < mlim (the rest of src is zero), outputs with both
// indices below olim; w = e^{-2 pi i / n}, conjugated for INV.
template <typename R, bool ROWS, bool INV>
__device__ inline void gclap_dft_pass(const R *__restrict__ sr, const R *__restrict__ si, R *__restrict__ dr,
                                      R *__restrict__ di, const R *__restrict__ twr, const R *__restrict__ twi,
                                      int n, int mlim, int olim) {
    for (int e = threadIdx.x; e < olim * olim; e += blockDim.x) {
        const int a = e / olim, k = e - a * olim;   // ROWS: row a, output k; else output a, column k
        const int kk = ROWS ? k : a;
        R accr = (R)0, acci = (R)0;
        int t = 0;
        for (int m = 0; m < mlim; ++m) {
            const int at = ROWS ? a * n + m : m * n + k;
            const R xr = sr[at], xi = si[at];
            const R wr = twr[t], wi = INV ? -twi[t] : twi[t];
            accr += xr * wr - xi * wi;
            acci += xr * wi + xi * wr;
            t += kk;
            if (t >= n) t -= n;
        }
        const int o = ROWS ? a * n + k : a * n + k;
        dr[o] = accr;
        di[o] = acci;
    }
}

template <typename R, int NMAX, bool ALPHA_ONE>
__device__ __forceinline__ void gclap_window_body(const GClapArgs<R> &a, const uint32_t bid_) {
    if (loop_stopped(a.gate)) return;
    constexpr int NN = NMAX * NMAX;
    __shared__ R xr[NN], xi[NN], yr[NN], yi[NN], twr[NMAX], twi[NMAX];
    const int n = a.n, nn = n * n, nw = a.n_win, t = threadIdx.x;
    const int nwin = a.nwi * a.nwj;
    const int w = (int)(bid_ % (uint32_t)nwin), f = (int)(bid_ / (uint32_t)nwin);
    const int wa = w / a.nwj, wb = w - (w / a.nwj) * a.nwj;
    int i1 = wa * a.n_inc, rs = 0;
    if (i1 + nw > a.ni) {
        rs = i1 + nw - a.ni;
        i1 = a.ni - nw;
    }
    int j1 = wb * a.n_inc, cs = 0;
    if (j1 + nw > a.nj) {
        cs = j1 + nw - a.nj;
        j1 = a.nj - nw;
    }
    for (int m = t; m < n; m += blockDim.x) {
        double sn, cn;
        sincospi(2.0 * (double)m / (double)n, &sn, &cn);
        twr[m] = (R)cn;
        twi[m] = (R)-sn;
    }
    // the window, zero padded to n x n (NaN -> 0, native_stage2.rs:852-858)
    const cplx<R> *g = a.grid + (int64_t)f * a.ni * a.nj;
    for (int k = t; k < nn; k += blockDim.x) {
        const int r = k / n, c = k - r * n;
        R vr = (R)0, vi = (R)0;
        if (r < nw && c < nw) {
            const cplx<R> v = g[(int64_t)(i1 + r) * a.nj + j1 + c];
            if (!RT<R>::isnan_(v.re) && !RT<R>::isnan_(v.im)) {
                vr = v.re;
                vi = v.im;
            }
        }
        xr[k] = vr;
        xi[k] = vi;
    }
    __syncthreads();
    gclap_dft_pass<R, true, false>(xr, xi, yr, yi, twr, twi, n, nw, n);   // rows (columns >= nw zero)
    __syncthreads();
    gclap_dft_pass<R, false, false>(yr, yi, xr, xi, twr, twi, n, nw, n);  // columns (rows >= nw zero)
    __syncthreads();
    // |F| at the fftshifted position (shift floor(n / 2)) -> yr
    const int fsh = n / 2, ish = (n + 1) / 2;
    for (int k = t; k < nn; k += blockDim.x) {
        const int r = k / n, c = k - r * n;
        const int src = ((r + fsh) % n) * n + (c + fsh) % n;
        yr[k] = RT<R>::abs_(xr[src], xi[src]);
    }
    __syncthreads();
    // 7-tap smoothing along rows (zero fill) -> yi, along columns -> yr
    for (int k = t; k < nn; k += blockDim.x) {
        const int r = k / n, c = k - r * n;
        R sm = (R)0;
        for (int q = 0; q < 7; ++q) {
            const int sc = c + q - 3;
            if (sc >= 0 && sc < n) sm += yr[r * n + sc] * a.k7[q];
        }
        yi[k] = sm;
    }
    __syncthreads();
    for (int k = t; k < nn; k += blockDim.x) {
        const int r = k / n, c = k - r * n;
        R sm = (R)0;
        for (int q = 0; q < 7; ++q) {
            const int sr2 = r + q - 3;
            if (sr2 >= 0 && sr2 < n) sm += yi[sr2 * n + c] * a.k7[q];
        }
        yr[k] = sm;
    }
    __syncthreads();
    // ifftshift (shift ceil(n / 2)) -> yi: the gains H in transform order
    for (int k = t; k < nn; k += blockDim.x) {
        const int r = k / n, c = k - r * n;
        yi[k] = yr[((r + ish) % n) * n + (c + ish) % n];
    }
    __syncthreads();
    // exact median: bitonic sort of a copy (padded with +inf to a power of two)
    int p2 = 1;
    while (p2 < nn) p2 <<= 1;
    for (int k = t; k < p2; k += blockDim.x) yr[k] = k < nn ? yi[k] : (R)INFINITY;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int e = t; e < p2 / 2; e += blockDim.x) {
                const int lo = (e / stride) * stride * 2 + (e % stride), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const R u = yr[lo], v = yr[hi];
                if ((u > v) == up) {
                    yr[lo] = v;
                    yr[hi] = u;
                }
            }
            __syncthreads();
        }
    }
    const R med = (nn % 2 == 0) ? (yr[nn / 2 - 1] + yr[nn / 2]) / (R)2 : yr[nn / 2];
    // gain and low-pass (native_stage2.rs:876-895)
    for (int k = t; k < nn; k += blockDim.x) {
        R v = yi[k];
        if (med != (R)0) v /= med;
        v = ALPHA_ONE ? v - (R)1 : RT<R>::pow_(v, a.alpha) - (R)1;
        if (v < (R)0) v = (R)0;
        const R gsc = v * a.beta + a.lp[k];
        xr[k] *= gsc;
        xi[k] *= gsc;
    }
    __syncthreads();
    gclap_dft_pass<R, true, true>(xr, xi, yr, yi, twr, twi, n, n, n);     // inverse rows
    __syncthreads();
    gclap_dft_pass<R, false, true>(yr, yi, xr, xi, twr, twi, n, n, nw);   // inverse columns, the window only
    __syncthreads();
    cplx<R> *o = a.gpart + ((int64_t)f * nwin + w) * nw * nw;
    for (int k = t; k < nw * nw; k += blockDim.x) {
        const int li = k / nw, lj = k - li * nw;
        R wt = (R)0;
        if (li >= rs && lj >= cs) wt = (R)a.weight[(li - rs) * nw + (lj - cs)];
        o[k] = {xr[li * n + lj] * wt, xi[li * n + lj] * wt};
    }
}

template <bool M, typename R, int NMAX, bool ALPHA_ONE>
__global__ __launch_bounds__(256) void k_gclap_window(KArg<M, GClapArgs<R>> p) {
    if constexpr (M) {
        GClapArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        gclap_window_body<R, NMAX, ALPHA_ONE>(a, bid_);
    } else {
        gclap_window_body<R, NMAX, ALPHA_ONE>(p, blockIdx.x);
    }
}

// per (interferogram, cell): the windows over the cell in window order, f64 sum
template <typename R, bool UNIT>
__device__ __forceinline__ void gclap_combine_body(const GClapArgs<R> &a, cplx<R> *__restrict__ filt,
                                                   const uint32_t bid_) {
    if (loop_stopped(a.gate)) return;
    const int64_t C = (int64_t)a.ni * a.nj;
    const int64_t gid = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (gid >= C * a.I) return;
    const int f = (int)(gid / C);
    const int64_t cell = gid - (int64_t)f * C;
    const int r = (int)(cell / a.nj), c = (int)(cell - (int64_t)r * a.nj);
    const int nw = a.n_win, nwin = a.nwi * a.nwj;
    const cplx<R> *pf = a.gpart + (int64_t)f * nwin * nw * nw;
    double sr = 0.0, si = 0.0;
    // windows ascending; the start row (column) of window q is
    // min(q n_inc, ni - n_win), non-decreasing in q
    const int wa0 = r - nw + 1 > 0 ? (r - nw + 1 + a.n_inc - 1) / a.n_inc : 0;
    const int wb0 = c - nw + 1 > 0 ? (c - nw + 1 + a.n_inc - 1) / a.n_inc : 0;
    for (int wa = wa0 < a.nwi ? wa0 : a.nwi; wa < a.nwi; ++wa) {
        const int i1 = wa * a.n_inc + nw > a.ni ? a.ni - nw : wa * a.n_inc;
        if (i1 > r) break;
        if (r >= i1 + nw) continue;
        for (int wb = wb0 < a.nwj ? wb0 : a.nwj; wb < a.nwj; ++wb) {
            const int j1 = wb * a.n_inc + nw > a.nj ? a.nj - nw : wb * a.n_inc;
            if (j1 > c) break;
            if (c >= j1 + nw) continue;
            const cplx<R> v = pf[((int64_t)(wa * a.nwj + wb) * nw + (r - i1)) * nw + (c - j1)];
            sr += (double)v.re;
            si += (double)v.im;
        }
    }
    const cplx<R> sum{(R)sr, (R)si};
    filt[gid] = UNIT ? unit_phase<R>(sum) : sum;
}

template <typename R>
struct GCombineArgs {
    GClapArgs<R> g;
    cplx<R> *filt;
};

template <bool M, typename R, bool UNIT>
__global__ __launch_bounds__(256) void k_gclap_combine(KArg<M, GCombineArgs<R>> p) {
    if constexpr (M) {
        GCombineArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        gclap_combine_body<R, UNIT>(a.g, a.filt, bid_);
    } else {
        gclap_combine_body<R, UNIT>(p.g, p.filt, blockIdx.x);
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_unit_phase(cplx<R> *__restrict__ v, int64_t n) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < n) v[gid] = unit_phase<R>(v[gid]);
}

// ---------------------------------------------------------------------------
// psdph of fit_stage2_rows (native_stage2.rs:686-694): conj(ph_patch) * ph_nm,
// where ph_patch is the cell's unit phase (the loop's filt planes hold unit
// phases, see k_clap_combine).
template <typename R>
__device__ inline cplx<R> psdph_value(cplx<R> p, cplx<R> nm) {
    const R ar = p.re, ai = -p.im;
    return {ar * nm.re - ai * nm.im, ar * nm.im + ai * nm.re};
}

// v * e^{-i*phase} with the reference's literal complex product
template <typename R>
__device__ inline cplx<R> rotate_neg(cplx<R> v, R phase) {
    R sn, cs;
    RT<R>::sincos_(phase, &sn, &cs);
    const R nsn = -sn;
    return {v.re * cs - v.im * nsn, v.re * nsn + v.im * cs};
}

struct FitParams {
    int64_t P, C;
    int I, tn, T, rowinv;
    int step_table;       // FP32 scan: unit steps from the session table (else from bperp)
    float rescore_bound;  // FP32 form only
};

// Session upload: the caller's row-major [P][I] arrays, copied as they are,
// become ifg-major planes [I][P] in cell-sorted pixel order.  Thread per
// (sorted pixel k, ifg f), f fastest, so each source row is read contiguously.
template <typename R>
__global__ __launch_bounds__(256) void k_planes_cplx(const float2 *__restrict__ src,
                                                     const int64_t *__restrict__ order,
                                                     cplx<R> *__restrict__ dst, int64_t P, int I) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * I) return;
    const int64_t k = gid / I;
    const int f = (int)(gid - k * I);
    const float2 v = src[order[k] * I + f];
    dst[(int64_t)f * P + k] = {(R)v.x, (R)v.y};
}

// Planes of the caller's f64 baselines: dst in R; with dst64 the f64 values;
// with dst_lo (FP32 form) the int32 residuals of the exact split v = (float)v
// + lo 2^(E - 54) (Bp64), values that do not split (denormal, infinite or NaN
// hi, a residual off the grid) counted in *inexact.
template <typename R>
__global__ __launch_bounds__(256) void k_planes_real(const double *__restrict__ src,
                                                     const int64_t *__restrict__ order,
                                                     R *__restrict__ dst, double *__restrict__ dst64,
                                                     int32_t *__restrict__ dst_lo, uint32_t *__restrict__ inexact,
                                                     int64_t P, int I) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * I) return;
    const int64_t k = gid / I;
    const int f = (int)(gid - k * I);
    const double v = src[order[k] * I + f];
    const int64_t e = (int64_t)f * P + k;
    dst[e] = (R)v;
    if (dst64) dst64[e] = v;
    if (dst_lo) {
        const float h = (float)v;
        const int be = (int)((__float_as_uint(h) >> 23) & 0xffu);
        int32_t lo = 0;
        bool ok;
        if (be == 0 || be == 255) {
            ok = v == 0.0;   // a zero (either sign) is h itself
        } else {
            const double r = (v - (double)h) * ldexp(1.0, 180 - be);
            ok = fabs(r) <= 1073741824.0 && r == rint(r);
            if (ok) {
                lo = (int32_t)r;
                ok = (lo == 0 ? (double)h : (double)h + ldexp((double)lo, be - 180)) == v;
            }
        }
        dst_lo[e] = lo;
        if (!ok) atomicAdd(inexact, 1u);
    }
}

__global__ __launch_bounds__(256) void k_planes_f32(const float *__restrict__ src,
                                                    const int64_t *__restrict__ order,
                                                    float *__restrict__ dst, int64_t P, int I) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= P * I) return;
    const int64_t k = gid / I;
    const int f = (int)(gid - k * I);
    dst[(int64_t)f * P + k] = src[order[k] * I + f];
}

// sorted pixel k -> the caller's row order[k]
template <typename T>
__global__ __launch_bounds__(256) void k_unpermute(const T *__restrict__ src, const int64_t *__restrict__ order,
                                                   T *__restrict__ dst, int64_t P) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < P) dst[order[k]] = src[k];
}

// Loop-invariant inputs of the trial scan (native_stage2.rs:1215-1225),
// computed once per session: the pixel's bperp range max - min (1 when zero)
// and, for the FP32 form, the unit trial step e^{-i phi} with
// phi = (bp / range) * pi/4 per interferogram, with the scan's own f32
// arithmetic.  wt may be NULL (FP64 form, or an FP32 scan that computes the
// steps itself).  With row-invariant bperp every pixel has the same range and
// wt holds one step per interferogram.
template <typename R>
__global__ __launch_bounds__(256) void k_trial_setup(const R *__restrict__ bp, int rowinv,
                                                     int64_t P, int I, R *__restrict__ rng,
                                                     cplx<R> *__restrict__ wt) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P) return;
    const R *b = rowinv ? bp : bp + k;
    const int64_t bs = rowinv ? 1 : P;
    R mn = (R)INFINITY, mx = -(R)INFINITY;
    for (int i = 0; i < I; ++i) {
        const R bi = b[i * bs];
        if (bi < mn) mn = bi;
        if (bi > mx) mx = bi;
    }
    R range = mx - mn;
    if (range == (R)0) range = (R)1;
    rng[k] = range;
    if (!wt) return;
    const R pi4 = RT<R>::pi4();
    if (!rowinv) {
        for (int i = 0; i < I; ++i) {
            R sn, cs;
            RT<R>::sincos_((b[i * bs] / range) * pi4, &sn, &cs);
            wt[i * P + k] = {cs, -sn};
        }
    } else if (k < I) {
        R sn, cs;
        RT<R>::sincos_((bp[k] / range) * pi4, &sn, &cs);
        wt[k] = {cs, -sn};
    }
}

// FP32 scan accumulators: (pre, pim) hold A_q = sum v Re(z^q) and (nre, nim)
// B_q = sum v Im(z^q) for q in [q0, qmax]; turn them into the trial sums
// sum v z^q = A + iB (pre, pim) and sum v conj(z)^q = A - iB (nre, nim).
template <typename R, int TC>
__device__ inline void scan_sums_from_ab(R (&pre)[TC], R (&pim)[TC], R (&nre)[TC], R (&nim)[TC], int q0,
                                         int qmax) {
#pragma unroll
    for (int q = 0; q < TC; ++q) {
        if (q >= q0 && q <= qmax) {
            const R ar = pre[q], ai = pim[q], br = nre[q], bi = nim[q];
            pre[q] = ar - bi;
            pim[q] = ai + br;
            nre[q] = ar + bi;
            nim[q] = ai - br;
        }
    }
}

// e / d with the bits of the IEEE quotient, from rcp = 1 / d (correctly
// rounded, once per pixel): q0 = e rcp is within 1.5 ulp of e / d; one
// correction q + (e - d q) rcp (the residual is exact with an FMA) brings it
// within 1 ulp, and a second is then correctly rounded (Markstein's theorem),
// for every quotient in the normal range.  Five instructions instead of the
// eleven of a division.  The corrections are written negated,
// -((-r) rcp + (-q)), which keeps the sign of a zero quotient.
__device__ inline float div_from_rcp(float e, float d, float rcp) {
    float q = e * rcp;
    q = -fmaf(-fmaf(-q, d, e), rcp, -q);
    q = -fmaf(-fmaf(-q, d, e), rcp, -q);
    return q;
}

// Trial scan of topofit_row (native_stage2.rs:1207-1256) for one pixel per
// thread.  Trials are processed in chunks of TC positive trials t and their
// negatives -t.
//   FP32 form (R=float): per ifg the rotation e^{-i phi t} comes from the
//     complex recurrence z_{t+1} = z_t * e^{-i phi}, re-anchored with an exact
//     sincos at the first trial of each chunk; e^{-i phi (-t)} = conj(z_t).
//     Rows whose best-vs-second margin is below rescore_bound are queued for
//     the FP64 re-score.
//   FP64 form (R=double): every rotation is an exact sincos, as the reference.
// Output: tix (selected trial index, -1 for invalid rows) and offset (the
// trial sum at the selected trial).  rng and wt come from k_trial_setup.
// SINGLE (FP32 form, tn < TC): only the single-chunk path is compiled, so the
// kernel needs fewer registers and more wavefronts fit per SIMD.
// NQ (single-chunk path only): tn known at compile time (NQ = tn > 0), so the
// trial loop has no per-trial test of a runtime count; 0 reads tn at run time.
template <typename R, int TC, bool SINGLE = false, int NQ = 0>
__device__ __forceinline__ void scan_body(
    const cplx<R> *__restrict__ phnm,
    const R *__restrict__ bp,
    const cplx<R> *__restrict__ wt,
    const cplx<R> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    int32_t *__restrict__ tix,
    const R *__restrict__ rng,
    cplx<R> *__restrict__ offset,
    int32_t *__restrict__ rescore_list,
    uint32_t *__restrict__ rescore_count,
    FitParams prm,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t k = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (k >= prm.P) return;
    const int I = prm.I;
    const int64_t P = prm.P;
    const cplx<R> *nm = phnm + k;                      // nm[i * P]
    const cplx<R> *fl = filt + cell_of[k];             // fl[i * C]
    const R *b = prm.rowinv ? bp : bp + k;
    const int64_t bs = prm.rowinv ? 1 : P;            // b[i * bs]
    const cplx<R> *w1 = wt ? (prm.rowinv ? wt : wt + k) : nullptr;  // FP32 table: w1[i * bs]
    const int64_t C = prm.C;
    const R range = rng[k];
    const R rcp = (R)1 / range;   // (the on-the-fly unit step of the 8-trial scan)
    const R pi4 = RT<R>::pi4();
    const int tn = prm.tn;

    // validity (any zero psdph element invalidates the row, :686-697) and the
    // denominator sum |psdph| ride along the first trial chunk
    bool valid = true;
    R denom = (R)0;
    R best = -(R)INFINITY, second = -(R)INFINITY;
    R bre = (R)0, bim = (R)0;   // trial sum at the selected trial: refine's offset
    int best_ix = 0;
    auto offer = [&](R val, int idx, R sre, R sim) {
        if (val > best || (val == best && idx < best_ix)) {
            if (best_ix != idx) second = best > second ? best : second;
            best = val;
            best_ix = idx;
            bre = sre;
            bim = sim;
        } else if (val > second) {
            second = val;
        }
    };

    static_assert(!SINGLE || sizeof(R) == 4, "the single-chunk scan is the FP32 form's");
    static_assert(NQ == 0 || (SINGLE && NQ < TC), "a compile-time trial count needs the single-chunk scan");
    if constexpr (sizeof(R) == 4) {
        if (SINGLE || tn < TC) {
            const int qmax = NQ > 0 ? NQ : tn;   // the host passes NQ = tn
            // one chunk: trials 0..tn and their negatives.  Trial 0 adds v
            // itself and the recurrence starts at z = e^{-i phi}, which is what
            // the general loop computes from z = 1 (exact multiplies by 1, 0).
            // The unit step e^{-i phi} comes from the session's table or, where
            // the host found the scan bound by HBM bandwidth (8-trial chunks
            // over large patches, per-pixel bperp), is computed from the
            // pixel's bperp: 4 bytes read instead of 8, with the table's
            // values (the quotient bperp / range has the IEEE bits, see
            // div_from_rcp, and the rest are the same f32 operations).  The
            // 8-trial loop
            // keeps the loads of the next two interferograms in flight.
            R pre[TC], pim[TC], nre[TC], nim[TC];
#pragma unroll
            for (int q = 0; q < TC; ++q) pre[q] = pim[q] = nre[q] = nim[q] = (R)0;
            auto run = [&](auto table) {
                constexpr bool TABLE = decltype(table)::value;
                using Src = typename std::conditional<TABLE, cplx<R>, R>::type;
                auto load = [&](int i) -> Src {
                    if constexpr (TABLE)
                        return w1[(int64_t)i * bs];
                    else
                        return b[(int64_t)i * bs];
                };
                auto step = [&](Src e) -> cplx<R> {
                    if constexpr (TABLE) {
                        return e;
                    } else {
                        R sn, cs;
                        RT<R>::sincos_(div_from_rcp(e, range, rcp) * pi4, &sn, &cs);
                        return {cs, -sn};
                    }
                };
                const int i1 = I > 1 ? 1 : 0;
                cplx<R> cf = fl[0], cn = nm[0];
                Src ce = load(0);
                cplx<R> mf = fl[(int64_t)i1 * C], mn = nm[(int64_t)i1 * P];
                Src me = load(i1);
                // (8 trials) unrolled by the period of the register rotation,
                // so the copies below become renaming: a copy of a register
                // that a load in flight targets would wait for that load
                constexpr int ROT_UNROLL = TC == 8 ? 3 : 1;
#pragma unroll ROT_UNROLL
                for (int i = 0; i < I; ++i) {
                    const int i2 = i + 2 < I ? i + 2 : I - 1;
                    const cplx<R> nf = fl[(int64_t)i2 * C], nn = nm[(int64_t)i2 * P];
                    const Src ne = load(i2);
                    const cplx<R> v = psdph_value<R>(cf, cn);
                    if (v.re == (R)0 && v.im == (R)0) valid = false;
                    denom += RT<R>::abs_(v.re, v.im);
                    pre[0] += v.re;
                    pim[0] += v.im;
                    const cplx<R> w = step(ce);
                    const R wr = w.re, wi = w.im;
                    R zr = wr, zi = wi;
#pragma unroll
                    for (int q = 1; q < TC; ++q) {
                        if (q <= qmax) {
                            // A_q = sum v Re(z^q) in (pre, pim), B_q = sum v
                            // Im(z^q) in (nre, nim): both trial sums v z^q =
                            // A + iB and v conj(z)^q = A - iB from four FMAs
                            pre[q] = fmaf(v.re, zr, pre[q]);
                            pim[q] = fmaf(v.im, zr, pim[q]);
                            nre[q] = fmaf(v.re, zi, nre[q]);
                            nim[q] = fmaf(v.im, zi, nim[q]);
                            const R nzr = fmaf(zr, wr, -zi * wi);
                            const R nzi = fmaf(zr, wi, zi * wr);
                            zr = nzr;
                            zi = nzi;
                        }
                    }
                    cf = mf;
                    cn = mn;
                    ce = me;
                    mf = nf;
                    mn = nn;
                    me = ne;
                }
            };
            if constexpr (TC == 8) {
                if (prm.step_table)
                    run(std::true_type{});
                else
                    run(std::false_type{});
            } else {
                run(std::true_type{});
            }
            if (!valid) {
                tix[k] = -1;
                return;
            }
            if (denom == (R)0) denom = (R)1;
            scan_sums_from_ab<R, TC>(pre, pim, nre, nim, 1, qmax);
#pragma unroll
            for (int q = 0; q < TC; ++q) {
                if (q <= qmax) {
                    offer(RT<R>::abs_(pre[q], pim[q]) / denom, q + tn, pre[q], pim[q]);
                    if (q != 0) offer(RT<R>::abs_(nre[q], nim[q]) / denom, tn - q, nre[q], nim[q]);
                }
            }
            tix[k] = best_ix;
            offset[k] = {bre, bim};
            if (prm.T > 1 && best - second < (R)prm.rescore_bound) {
                const uint32_t slot = atomicAdd(rescore_count, 1u);
                rescore_list[slot] = (int32_t)k;
            }
            return;
        }
    }
    if constexpr (SINGLE) return;
    for (int t0 = 0; t0 <= tn; t0 += TC) {
        // chunk: positive trials t0 .. t0+TC-1 (t0 = 0 holds trial 0) and their negatives
        R pre[TC], pim[TC], nre[TC], nim[TC];
#pragma unroll
        for (int q = 0; q < TC; ++q) pre[q] = pim[q] = nre[q] = nim[q] = (R)0;
        for (int i = 0; i < I; ++i) {
            const cplx<R> v = psdph_value<R>(fl[i * C], nm[i * P]);
            if (t0 == 0) {
                if (v.re == (R)0 && v.im == (R)0) valid = false;
                denom += RT<R>::abs_(v.re, v.im);
            }
            if constexpr (sizeof(R) == 4) {
                const cplx<R> e1 = w1[i * bs];  // e^{-i phi}
                const R wr = e1.re, wi = e1.im;
                R zr, zi;
                if (t0 == 0) {
                    zr = (R)1;
                    zi = (R)0;
                } else {
                    const R phi = (b[i * bs] / range) * pi4;
                    R sa, ca;
                    RT<R>::sincos_(phi * (R)t0, &sa, &ca);
                    zr = ca;
                    zi = -sa;
                }
#pragma unroll
                for (int q = 0; q < TC; ++q) {
                    if (t0 + q <= tn) {
                        // A and B as in the single-chunk path
                        pre[q] = fmaf(v.re, zr, pre[q]);
                        pim[q] = fmaf(v.im, zr, pim[q]);
                        nre[q] = fmaf(v.re, zi, nre[q]);
                        nim[q] = fmaf(v.im, zi, nim[q]);
                        const R nzr = fmaf(zr, wr, -zi * wi);
                        const R nzi = fmaf(zr, wi, zi * wr);
                        zr = nzr;
                        zi = nzi;
                    }
                }
            } else {
                const R phi = (b[i * bs] / range) * pi4;
#pragma unroll
                for (int q = 0; q < TC; ++q) {
                    const int t = t0 + q;
                    if (t <= tn) {
                        const cplx<R> a = rotate_neg<R>(v, phi * (R)t);
                        pre[q] += a.re;
                        pim[q] += a.im;
                        if (t != 0) {
                            const cplx<R> bb = rotate_neg<R>(v, phi * (R)(-t));
                            nre[q] += bb.re;
                            nim[q] += bb.im;
                        }
                    }
                }
            }
        }
        if constexpr (sizeof(R) == 4) scan_sums_from_ab<R, TC>(pre, pim, nre, nim, 0, tn - t0);
        if (t0 == 0) {
            if (!valid) {
                tix[k] = -1;
                return;
            }
            if (denom == (R)0) denom = (R)1;
        }
#pragma unroll
        for (int q = 0; q < TC; ++q) {
            const int t = t0 + q;
            if (t <= tn) {
                offer(RT<R>::abs_(pre[q], pim[q]) / denom, t + tn, pre[q], pim[q]);
                if (t != 0) offer(RT<R>::abs_(nre[q], nim[q]) / denom, tn - t, nre[q], nim[q]);
            }
        }
    }
    tix[k] = best_ix;
    offset[k] = {bre, bim};
    if constexpr (sizeof(R) == 4) {
        if (prm.T > 1 && best - second < (R)prm.rescore_bound) {
            const uint32_t slot = atomicAdd(rescore_count, 1u);
            rescore_list[slot] = (int32_t)k;
        }
    }
}

// --- one row per quad of lanes (small patches, FP32 form) ---------------------
// Broadcast of lane S of each quad (4 consecutive lanes) to the quad.
template <int S>
__device__ __forceinline__ uint32_t quad_bcast_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, S * 0x55, 0xf, 0xf, false);
}
template <int S>
__device__ __forceinline__ float quad_bcast_f32(float v) {
    return __uint_as_float(quad_bcast_u32<S>(__float_as_uint(v)));
}
template <int S>
__device__ __forceinline__ double quad_bcast_f64(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = quad_bcast_u32<S>((uint32_t)u), hi = quad_bcast_u32<S>((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float quad_max_f32(float v) {
    v = fmaxf(v, __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0xB1, 0xf, 0xf, false)));
    return fmaxf(v, __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0x4E, 0xf, 0xf, false)));
}
__device__ __forceinline__ float quad_min_f32(float v) {
    v = fminf(v, __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0xB1, 0xf, 0xf, false)));
    return fminf(v, __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0x4E, 0xf, 0xf, false)));
}

// Sum over the quad, the same value in all four lanes ((x0 + x1) + (x2 + x3)
// up to the order of each commutative addition, so bit-identical).
__device__ __forceinline__ float quad_sum_f32(float v) {
    v += __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0xB1, 0xf, 0xf, false));
    return v + __uint_as_float((uint32_t)__builtin_amdgcn_update_dpp(0, (int)__float_as_uint(v), 0x4E, 0xf, 0xf, false));
}
__device__ __forceinline__ bool quad_all(bool b) {
    int v = b ? 1 : 0;
    v &= __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);
    v &= __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);
    return v != 0;
}

// Adds the terms x of the quad's lanes 0 .. n-1 (n <= 4, uniform) to acc in
// lane order, in every lane of the quad.
__device__ __forceinline__ void quad_add_f64(double &acc, double x, int n) {
    const double x0 = quad_bcast_f64<0>(x), x1 = quad_bcast_f64<1>(x);
    const double x2 = quad_bcast_f64<2>(x), x3 = quad_bcast_f64<3>(x);
    acc += x0;
    if (n > 1) acc += x1;
    if (n > 2) acc += x2;
    if (n > 3) acc += x3;
}
__device__ __forceinline__ void quad_add_f32_as_f64(double &acc, float x, int n) {
    const float x0 = quad_bcast_f32<0>(x), x1 = quad_bcast_f32<1>(x);
    const float x2 = quad_bcast_f32<2>(x), x3 = quad_bcast_f32<3>(x);
    acc += (double)x0;
    if (n > 1) acc += (double)x1;
    if (n > 2) acc += (double)x2;
    if (n > 3) acc += (double)x3;
}

template <typename R>
struct ScanArgs {
    const cplx<R> * phnm;
    const R * bp;
    const cplx<R> * wt;
    const cplx<R> * filt;
    const int32_t * cell_of;
    int32_t * tix;
    const R * rng;
    cplx<R> * offset;
    int32_t * rescore_list;
    uint32_t * rescore_count;
    FitParams prm;
    const int32_t * gate;
};

template <bool M, typename R, int TC, bool SINGLE = false, int NQ = 0>
__global__ __launch_bounds__(256) void k_scan(KArg<M, ScanArgs<R>> p) {
    if constexpr (M) {
        ScanArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        scan_body<R, TC, SINGLE, NQ>(a.phnm, a.bp, a.wt, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.rescore_list, a.rescore_count, a.prm, a.gate, bid_, nblk_);
    } else {
        scan_body<R, TC, SINGLE, NQ>(p.phnm, p.bp, p.wt, p.filt, p.cell_of, p.tix, p.rng, p.offset, p.rescore_list, p.rescore_count, p.prm, p.gate, blockIdx.x, gridDim.x);
    }
}

// The FP32 form's single-chunk scan (fewer than 8 positive trials) for small
// patches, one row per quad of lanes: lane `sub` of the quad accumulates the
// trial sums of interferograms sub, sub + 4, ... (the same per-element
// arithmetic as scan_body), the quad adds its four partial sums, and the
// selection, the offset and the re-score queueing follow as in scan_body.
// The f32 sums run in another order than the one-thread scan, a difference
// inside the bound the re-score margin is derived from (an I-term f32 sum in
// any order).  64 rows per 256-thread workgroup.
template <int NQ>
__device__ __forceinline__ void scan_quad_body(
    const cplx<float> *__restrict__ phnm,
    const float *__restrict__ bp,
    const cplx<float> *__restrict__ wt,
    const cplx<float> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    int32_t *__restrict__ tix,
    const float *__restrict__ rng,
    cplx<float> *__restrict__ offset,
    int32_t *__restrict__ rescore_list,
    uint32_t *__restrict__ rescore_count,
    FitParams prm,
    const int32_t *__restrict__ gate,
    const uint32_t bid_,
    int32_t *local_queue = nullptr,
    int *local_count = nullptr) {
    using R = float;
    constexpr int TC = 8;
    if (loop_stopped(gate)) return;
    const int64_t k = (int64_t)bid_ * 64 + (threadIdx.x >> 2);
    const int sub = threadIdx.x & 3;
    if (k >= prm.P) return;   // (a quad's lanes share k)
    const int I = prm.I;
    const int64_t P = prm.P, C = prm.C;
    const cplx<R> *nm = phnm + k;
    const cplx<R> *fl = filt + cell_of[k];
    const R *b = prm.rowinv ? bp : bp + k;
    const int64_t bs = prm.rowinv ? 1 : P;
    const cplx<R> *w1 = wt ? (prm.rowinv ? wt : wt + k) : nullptr;
    const R range = rng[k];
    const R rcp = (R)1 / range;
    const R pi4 = RT<R>::pi4();
    const int tn = prm.tn;
    const int qmax = NQ > 0 ? NQ : tn;
    bool valid = true;
    R denom = (R)0;
    R pre[TC], pim[TC], nre[TC], nim[TC];
#pragma unroll
    for (int q = 0; q < TC; ++q) pre[q] = pim[q] = nre[q] = nim[q] = (R)0;
    const bool table = prm.step_table != 0;
    for (int i = sub; i < I; i += 4) {
        const cplx<R> v = psdph_value<R>(fl[(int64_t)i * C], nm[(int64_t)i * P]);
        if (v.re == (R)0 && v.im == (R)0) valid = false;
        denom += RT<R>::abs_(v.re, v.im);
        pre[0] += v.re;
        pim[0] += v.im;
        R wr, wi;
        if (table) {
            const cplx<R> e = w1[(int64_t)i * bs];
            wr = e.re;
            wi = e.im;
        } else {
            R sn, cs;
            RT<R>::sincos_(div_from_rcp(b[(int64_t)i * bs], range, rcp) * pi4, &sn, &cs);
            wr = cs;
            wi = -sn;
        }
        R zr = wr, zi = wi;
#pragma unroll
        for (int q = 1; q < TC; ++q) {
            if (q <= qmax) {
                pre[q] = fmaf(v.re, zr, pre[q]);
                pim[q] = fmaf(v.im, zr, pim[q]);
                nre[q] = fmaf(v.re, zi, nre[q]);
                nim[q] = fmaf(v.im, zi, nim[q]);
                const R nzr = fmaf(zr, wr, -zi * wi);
                const R nzi = fmaf(zr, wi, zi * wr);
                zr = nzr;
                zi = nzi;
            }
        }
    }
    valid = quad_all(valid);
    denom = quad_sum_f32(denom);
#pragma unroll
    for (int q = 0; q < TC; ++q) {
        if (q <= qmax) {
            pre[q] = quad_sum_f32(pre[q]);
            pim[q] = quad_sum_f32(pim[q]);
            nre[q] = quad_sum_f32(nre[q]);
            nim[q] = quad_sum_f32(nim[q]);
        }
    }
    if (sub != 0) return;
    if (!valid) {
        tix[k] = -1;
        return;
    }
    if (denom == (R)0) denom = (R)1;
    scan_sums_from_ab<R, TC>(pre, pim, nre, nim, 1, qmax);
    R best = -(R)INFINITY, second = -(R)INFINITY;
    R bre = (R)0, bim = (R)0;
    int best_ix = 0;
    auto offer = [&](R val, int idx, R sre, R sim) {
        if (val > best || (val == best && idx < best_ix)) {
            if (best_ix != idx) second = best > second ? best : second;
            best = val;
            best_ix = idx;
            bre = sre;
            bim = sim;
        } else if (val > second) {
            second = val;
        }
    };
#pragma unroll
    for (int q = 0; q < TC; ++q) {
        if (q <= qmax) {
            offer(RT<R>::abs_(pre[q], pim[q]) / denom, q + tn, pre[q], pim[q]);
            if (q != 0) offer(RT<R>::abs_(nre[q], nim[q]) / denom, tn - q, nre[q], nim[q]);
        }
    }
    tix[k] = best_ix;
    offset[k] = {bre, bim};
    if (prm.T > 1 && best - second < (R)prm.rescore_bound) {
        if (local_queue) {   // the workgroup's own queue in LDS (k_fit_small<..., true>)
            local_queue[atomicAdd(local_count, 1)] = (int32_t)k;
        } else {
            const uint32_t slot = atomicAdd(rescore_count, 1u);
            rescore_list[slot] = (int32_t)k;
        }
    }
}

template <bool M, int NQ>
__global__ __launch_bounds__(256) void k_scan_quad(KArg<M, ScanArgs<float>> p) {
    if constexpr (M) {
        ScanArgs<float> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        scan_quad_body<NQ>(a.phnm, a.bp, a.wt, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.rescore_list,
                           a.rescore_count, a.prm, a.gate, bid_);
    } else {
        scan_quad_body<NQ>(p.phnm, p.bp, p.wt, p.filt, p.cell_of, p.tix, p.rng, p.offset, p.rescore_list,
                           p.rescore_count, p.prm, p.gate, blockIdx.x);
    }
}

// FP64 re-score of near-tie rows (FP32 form): one 64-lane wavefront per
// queued row, lanes over interferograms, wave sums per trial; first argmax
// with strict > over trials in ascending order, as ref64's topofit_row
// (native_stage2.rs:1243-1253).  bp64 holds the caller's f64 bperp (the f32
// copy used by the bulk is exact only to f32).
__device__ inline double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// four wave sums at once (the exchanges of the four values overlap)
__device__ inline void wave_sum4_f64(double &a, double &b, double &c, double &d) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double ta = __shfl_xor(a, o, 64), tb = __shfl_xor(b, o, 64);
        const double tc = __shfl_xor(c, o, 64), td = __shfl_xor(d, o, 64);
        a += ta;
        b += tb;
        c += tc;
        d += td;
    }
}

// The caller's f64 baselines for the FP32 form's FP64 re-score and
// re-decision: the f64 plane itself (`full`), or, where every value of the
// plane splits exactly, the f32 bulk plane `hi` plus an int32 residual `lo`:
// v = hi + lo 2^(E - 54), E the frexp exponent of hi (the residual v - hi of
// a normal hi has at most 29 significant bits on the grid 2^(E - 54), so the
// split is exact and needs 4 bytes per element instead of 8).  k_planes_real
// writes the split and counts the values it cannot represent.
struct Bp64 {
    const double *full;
    const float *hi;
    const int32_t *lo;
    __device__ __forceinline__ double at(int64_t e) const {
        if (full) return full[e];
        const float h = hi[e];
        const int32_t l = lo[e];
        if (l == 0) return (double)h;   // (keeps the sign of a zero)
        const int be = (int)((__float_as_uint(h) >> 23) & 0xffu);
        return (double)h + ldexp((double)l, be - 180);
    }
};

template <typename RF>
__device__ __forceinline__ void rescore_body(
    const cplx<RF> *__restrict__ phnm,
    const Bp64 bp64,
    const cplx<RF> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    const int32_t *__restrict__ rescore_list,
    const uint32_t *__restrict__ rescore_count,
    int32_t *__restrict__ tix,
    cplx<RF> *__restrict__ offset,
    FitParams prm,
    const int32_t *__restrict__ gate,
    const int64_t row_lo, const int64_t row_hi,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    // one wavefront per queued row (wavefront bid_ of nblk_, rows in
    // [row_lo, row_hi) only); lane l holds interferograms l, l + 64, ... (up
    // to RESCORE_PER_LANE of them) for every trial; larger I re-reads the rest
    // per trial
    constexpr int RESCORE_PER_LANE = 4;
    const uint32_t n = *rescore_count;
    const int lane = threadIdx.x & 63;
    for (uint32_t q = bid_; q < n; q += nblk_) {
        const int64_t k = rescore_list[q];
        if (k < row_lo || k >= row_hi) continue;
        const int I = prm.I;
        const int64_t P = prm.P, C = prm.C;
        const cplx<RF> *nm = phnm + k;
        const cplx<RF> *fl = filt + cell_of[k];
        auto bat = [&](int i) { return bp64.at(prm.rowinv ? (int64_t)i : (int64_t)i * P + k); };
        double vr[RESCORE_PER_LANE], vi[RESCORE_PER_LANE], vb[RESCORE_PER_LANE];
        double denom = 0.0, mn = INFINITY, mx = -INFINITY;
#pragma unroll
        for (int u = 0; u < RESCORE_PER_LANE; ++u) {
            const int i = lane + 64 * u;
            vr[u] = vi[u] = vb[u] = 0.0;
            if (i < I) {
                const cplx<RF> v = psdph_value<RF>(fl[(int64_t)i * C], nm[(int64_t)i * P]);
                vr[u] = (double)v.re;
                vi[u] = (double)v.im;
                vb[u] = bat(i);
            }
        }
        for (int i = lane; i < I; i += 64) {
            double re, im, bi;
            const int u = i / 64;
            if (u < RESCORE_PER_LANE) {
                re = u == 0 ? vr[0] : u == 1 ? vr[1] : u == 2 ? vr[2] : vr[3];
                im = u == 0 ? vi[0] : u == 1 ? vi[1] : u == 2 ? vi[2] : vi[3];
                bi = u == 0 ? vb[0] : u == 1 ? vb[1] : u == 2 ? vb[2] : vb[3];
            } else {
                const cplx<RF> v = psdph_value<RF>(fl[(int64_t)i * C], nm[(int64_t)i * P]);
                re = (double)v.re;
                im = (double)v.im;
                bi = bat(i);
            }
            denom += hypot(re, im);
            mn = bi < mn ? bi : mn;
            mx = bi > mx ? bi : mx;
        }
        denom = wave_sum_f64(denom);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, o, 64));
            mx = fmax(mx, __shfl_xor(mx, o, 64));
        }
        if (denom == 0.0) denom = 1.0;
        double range = mx - mn;
        if (range == 0.0) range = 1.0;
        double bestv = -INFINITY, bsr = 0.0, bsi = 0.0;
        int besti = -1;
        if (I <= 64 * RESCORE_PER_LANE) {
            // e^{-i phi t} by an f64 recurrence from t = 0 (drift ~t ulp, far
            // below the f64 sums' own rounding): per positive trial t the
            // wave sums A = sum v Re(z^t), B = sum v Im(z^t) give both trial
            // sums, v z^t = A + iB (trial t) and v conj(z)^t = A - iB (trial
            // -t); the first argmax over ascending trial index is kept with
            // the smaller index winning exact ties
            // (only the nu = ceil(I / 64) slots that hold interferograms: the
            // others are zero and stay out of the serial chain)
            const int nu = (I + 63) / 64;
            double zr[RESCORE_PER_LANE], zi[RESCORE_PER_LANE], wr[RESCORE_PER_LANE], wi[RESCORE_PER_LANE];
#pragma unroll
            for (int u = 0; u < RESCORE_PER_LANE; ++u) {
                zr[u] = 1.0;
                zi[u] = 0.0;
                wr[u] = 1.0;
                wi[u] = 0.0;
                if (u < nu) {
                    double sn, cs;
                    sincos((vb[u] / range) * 0.78539816339744830962, &sn, &cs);
                    wr[u] = cs;
                    wi[u] = -sn;
                }
            }
            auto offer = [&](double sr, double si, int idx) {
                const double c = hypot(sr, si) / denom;
                if (besti < 0 || c > bestv || (c == bestv && idx < besti)) {
                    bestv = c;
                    besti = idx;
                    bsr = sr;
                    bsi = si;
                }
            };
            for (int t = 0; t <= prm.tn; ++t) {
                double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
#pragma unroll
                for (int u = 0; u < RESCORE_PER_LANE; ++u) {
                    if (u < nu) {
                        ar = fma(vr[u], zr[u], ar);
                        ai = fma(vi[u], zr[u], ai);
                        br = fma(vr[u], zi[u], br);
                        bi = fma(vi[u], zi[u], bi);
                        const double nzr = zr[u] * wr[u] - zi[u] * wi[u];
                        const double nzi = zr[u] * wi[u] + zi[u] * wr[u];
                        zr[u] = nzr;
                        zi[u] = nzi;
                    }
                }
                wave_sum4_f64(ar, ai, br, bi);
                offer(ar - bi, ai + br, prm.tn + t);
                if (t > 0) offer(ar + bi, ai - br, prm.tn - t);
            }
        } else
        for (int t = 0; t < prm.T; ++t) {
            const double tv = (double)(t - prm.tn);
            double sr = 0.0, si = 0.0;
            for (int i = lane; i < I; i += 64) {
                double re, im, bi;
                const int u = i / 64;
                if (u < RESCORE_PER_LANE) {
                    re = u == 0 ? vr[0] : u == 1 ? vr[1] : u == 2 ? vr[2] : vr[3];
                    im = u == 0 ? vi[0] : u == 1 ? vi[1] : u == 2 ? vi[2] : vi[3];
                    bi = u == 0 ? vb[0] : u == 1 ? vb[1] : u == 2 ? vb[2] : vb[3];
                } else {
                    const cplx<RF> v = psdph_value<RF>(fl[(int64_t)i * C], nm[(int64_t)i * P]);
                    re = (double)v.re;
                    im = (double)v.im;
                    bi = bat(i);
                }
                const double ph = ((bi / range) * 0.78539816339744830962) * tv;
                double sn, cs;
                sincos(ph, &sn, &cs);
                const double nsn = -sn;
                sr += re * cs - im * nsn;
                si += re * nsn + im * cs;
            }
            sr = wave_sum_f64(sr);
            si = wave_sum_f64(si);
            const double c = hypot(sr, si) / denom;
            if (besti < 0 || c > bestv) {
                bestv = c;
                besti = t;
                bsr = sr;
                bsi = si;
            }
        }
        if (lane == 0 && besti >= 0 && besti < prm.T && besti != tix[k]) {
            tix[k] = besti;
            offset[k] = {(RF)bsr, (RF)bsi};
        }
    }
}

template <typename RF>
struct RescoreArgs {
    const cplx<RF> * phnm;
    Bp64 bp64;
    const cplx<RF> * filt;
    const int32_t * cell_of;
    const int32_t * rescore_list;
    const uint32_t * rescore_count;
    int32_t * tix;
    cplx<RF> * offset;
    FitParams prm;
    const int32_t * gate;
};

template <bool M, typename RF>
__global__ __launch_bounds__(64) void k_rescore(KArg<M, RescoreArgs<RF>> p) {
    if constexpr (M) {
        RescoreArgs<RF> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        rescore_body<RF>(a.phnm, a.bp64, a.filt, a.cell_of, a.rescore_list, a.rescore_count, a.tix, a.offset, a.prm, a.gate, 0, a.prm.P, bid_, nblk_);
    } else {
        rescore_body<RF>(p.phnm, p.bp64, p.filt, p.cell_of, p.rescore_list, p.rescore_count, p.tix, p.offset, p.prm, p.gate, 0, p.prm.P, blockIdx.x, gridDim.x);
    }
}

// Coherence histogram rule (hist_with_centers, native_stage2.rs:1538-1622):
// the bin of one value, equal-spacing fast path or the bin mid-points.
struct HistParams {
    int64_t P;
    int equal;       // equal-spacing rule
    double cutoff0, d;
    int nbins;
};

__device__ inline int hist_bin(double v, const double *__restrict__ mids, HistParams hp) {
    if (!__builtin_isfinite(v)) return -1;
    if (hp.equal) {
        if (hp.d == 0.0) return 0;
        double a = ceil((v - hp.cutoff0) / hp.d);
        if (a < 0.0) a = 0.0;
        if (a > (double)(hp.nbins - 1)) a = (double)(hp.nbins - 1);
        return (int)a;
    }
    int b = 0;
    while (b < hp.nbins - 1 && mids[b] < v) ++b;
    return b;
}

// refine_candidate (native_stage2.rs:1259-1303) of row k from the selected
// trial: offset sum, f64 least-squares step mopt, final residual sum.  Writes
// K, C, coh and N_opt of the row and returns its coh.  ROWINV: one bperp row
// shared by every pixel.
template <typename R, bool ROWINV>
__device__ inline double refine_row(int64_t k, const cplx<R> *__restrict__ phnm, const R *__restrict__ bp,
                                    const cplx<R> *__restrict__ filt, const int32_t *__restrict__ cell_of,
                                    const int32_t *__restrict__ tix, const R *__restrict__ rng,
                                    const cplx<R> *__restrict__ offset, double *__restrict__ Kout,
                                    double *__restrict__ Cout, double *__restrict__ cohout,
                                    double *__restrict__ nopt, const FitParams &prm,
                                    float *max_angle, float *cut_margin) {
    const int t = tix[k];
    if (t < 0) {
        Kout[k] = __builtin_nan("");
        Cout[k] = 0.0;
        cohout[k] = 0.0;
        nopt[k] = 0.0;
        return 0.0;
    }
    const int I = prm.I;
    const int64_t P = prm.P, C = prm.C;
    const cplx<R> *nm = phnm + k;
    const cplx<R> *fl = filt + cell_of[k];
    auto bat = [&](int i) { return ROWINV ? bp[i] : bp[(int64_t)i * P + k]; };
    const R range = rng[k];
    R k0 = (R)(RT<R>::pi4() / range * (R)(t - prm.tn));
    // offset = sum psdph e^{-i k0 bp} (:1261-1267) is the trial scan's sum at
    // the selected trial (same angles up to rounding), kept by k_scan/k_rescore
    const cplx<R> off = offset[k];
    const R ocr = off.re, oci = -off.im;
    double num = 0.0, den = 0.0;
    // largest |k0 bp| and smallest distance of arg(r conj(offset)) to the
    // +-pi cut of this pass (the f64 re-decision takes rows near the cut)
    R amax1 = (R)0, cut = (R)RT<R>::pi4() * (R)4;
    // both passes keep the loads of the next two interferograms in flight
    // while the current one is processed; the loops are unrolled by the
    // period of that register rotation so the copies become renaming (a copy
    // of a register that a load in flight targets would wait for the load)
    const int i1 = I > 1 ? 1 : 0;
    cplx<R> cf = fl[0], cn = nm[0];
    R cb = bat(0);
    cplx<R> mf = fl[(int64_t)i1 * C], mn = nm[(int64_t)i1 * P];
    R mb = bat(i1);
#pragma unroll 3
    for (int i = 0; i < I; ++i) {
        const int i2 = i + 2 < I ? i + 2 : I - 1;
        const cplx<R> nf = fl[(int64_t)i2 * C], nn = nm[(int64_t)i2 * P];
        const R nb = bat(i2);
        const cplx<R> v = psdph_value<R>(cf, cn);
        const R bi = cb;
        cf = mf;
        cn = mn;
        cb = mb;
        mf = nf;
        mn = nn;
        mb = nb;
        const double wgt = (double)RT<R>::abs_(v.re, v.im);
        const double wb = wgt * (double)bi;
        den += wb * wb;
        const R ang = k0 * bi;
        const cplx<R> r = rotate_neg<R>(v, ang);
        const R pr = r.re * ocr - r.im * oci, pim = r.re * oci + r.im * ocr;
        const R a = RT<R>::atan2_(pim, pr);
        num += wb * (wgt * (double)a);
        if constexpr (sizeof(R) == 4) {   // (the FP64 form has no re-decision)
            amax1 = fmax(amax1, fabs(ang));
            cut = fmin(cut, (R)RT<R>::pi4() * (R)4 - fabs(a));
        }
    }
    if (den == 0.0) den = 1.0;
    k0 = (R)((double)k0 + num / den);
    // residual sums over the interferograms in f64 (the FP32 form's terms are
    // f32; f64 accumulation keeps the I-term sum's rounding out of coh)
    double mre = 0.0, mim = 0.0, dn = 0.0;
    R amax = (R)0;   // largest |k0 bp| of this pass (bounds the f32 angle rounding)
    cf = fl[0];
    cn = nm[0];
    cb = bat(0);
    mf = fl[(int64_t)i1 * C];
    mn = nm[(int64_t)i1 * P];
    mb = bat(i1);
#pragma unroll 3
    for (int i = 0; i < I; ++i) {
        const int i2 = i + 2 < I ? i + 2 : I - 1;
        const cplx<R> nf = fl[(int64_t)i2 * C], nn = nm[(int64_t)i2 * P];
        const R nb = bat(i2);
        const R ang = k0 * cb;
        const cplx<R> r = rotate_neg<R>(psdph_value<R>(cf, cn), ang);
        if constexpr (sizeof(R) == 4) amax = fmax(amax, fabs(ang));
        cf = mf;
        cn = mn;
        cb = mb;
        mf = nf;
        mn = nn;
        mb = nb;
        mre += (double)r.re;
        mim += (double)r.im;
        dn += (double)RT<R>::abs_(r.re, r.im);
    }
    if (dn == 0.0) dn = 1.0;
    Kout[k] = (double)k0;
    Cout[k] = (double)RT<R>::atan2_((R)mim, (R)mre);
    const double coh = hypot(mre, mim) / dn;
    cohout[k] = coh;
    nopt[k] = 1.0;
    *max_angle = (float)fmax(amax, amax1);
    *cut_margin = (float)cut;
    return coh;
}


// Outputs of the loop's gamma-change and histogram step (native_stage2.rs:
// 2045-2057 and hist_with_centers :1538-1622), taken by k_post over the
// coherences just computed: per group of 256 rows the sum of
// (coh - coh_prev)^2 in a fixed tree (one partial per group, summed in a
// fixed order by k_control), coh_prev <- coh, and the bin counts (integer,
// order-free).
struct RmsHistOut {
    double *coh_prev;   // NULL: no RMS/histogram step (single-step use)
    double *part;       // [ceil(P / 256)]
    uint32_t *hist;
    const double *mids;
    HistParams hp;
};


// FP64 re-decision of the P-square inputs (FP32 form).  The loop turns each
// coherence into two discrete choices: the P-square lookup index
// round_half_away(1000 coh) (native_stage2.rs:1674-1682) and the coherence
// histogram bin (hist_with_centers :1593-1606).  A row whose FP32 coherence
// lies within the refine's error bound of a choice boundary is refined again
// in FP64 (one wavefront per row, lanes over interferograms, from the same
// trial index, filtered phase and ph_nm, with the caller's f64 baselines), as
// refine_candidate (:1259-1303) does in the FP64 oracle; its K, C and coh are
// then the FP64 values.
struct RecheckIn {
    Bp64 bp64;            // [I][P] (or [I] with invariant rows): the caller's f64 bperp
    int enabled;          // FP32 form
    int psquare;          // P-square weighting: also rows near a lookup / bin boundary
};

// distance of coh to the nearest lookup or histogram boundary
__device__ inline double decision_distance(double coh, const double *__restrict__ mids, const HistParams &hp) {
    const double x = 1000.0 * coh;
    double dist = fabs(x - floor(x) - 0.5) * 1e-3;   // round-half boundaries n + 1/2
    if (hp.equal) {
        if (hp.d != 0.0) {
            const double y = (coh - hp.cutoff0) / hp.d;   // bin edges cutoff0 + j d
            dist = fmin(dist, fabs(y - rint(y)) * hp.d);
        }
    } else {
        for (int b = 0; b + 1 < hp.nbins; ++b) dist = fmin(dist, fabs(coh - mids[b]));
    }
    return dist;
}

// FP64 refine of row k by one wavefront (every lane returns the row's coh).
// Lane l holds interferograms l, l + 64, ... (up to RW of them, loaded once);
// beyond 64 RW interferograms the rest are read again in each pass.
template <int RW>
__device__ inline double refine_row_f64_wave(int64_t k, const cplx<float> *__restrict__ phnm,
                                             const Bp64 &bp64,
                                             const cplx<float> *__restrict__ filt,
                                             const int32_t *__restrict__ cell_of,
                                             const int32_t *__restrict__ tix, double *__restrict__ Kout,
                                             double *__restrict__ Cout, double *__restrict__ cohout,
                                             const FitParams &prm, int lane) {
    const int I = prm.I;
    const int64_t P = prm.P, C = prm.C;
    const cplx<float> *nm = phnm + k;
    const cplx<float> *fl = filt + cell_of[k];
    auto bat = [&](int i) { return bp64.at(prm.rowinv ? (int64_t)i : (int64_t)i * P + k); };
    // psdph in f64: conj(unit filtered phase) x ph_nm
    auto load = [&](int i, double &re, double &im, double &b) {
        const cplx<float> f = fl[(int64_t)i * C], n = nm[(int64_t)i * P];
        double fr = (double)f.re, fi = (double)f.im;
        const double a = hypot(fr, fi);
        if (a != 0.0) {
            fr /= a;
            fi /= a;
        }
        const double nr = (double)n.re, ni = (double)n.im;
        re = fr * nr + fi * ni;
        im = fr * ni - fi * nr;
        b = bat(i);
    };
    double vr[RW], vi[RW], vb[RW];
#pragma unroll
    for (int u = 0; u < RW; ++u) {
        vr[u] = vi[u] = vb[u] = 0.0;
        if (lane + 64 * u < I) load(lane + 64 * u, vr[u], vi[u], vb[u]);
    }
    // element i of this lane: registers for u < RW, global memory beyond
    auto elem = [&](int u, int i, double &re, double &im, double &b) {
        if (u < RW) {
#pragma unroll
            for (int w = 0; w < RW; ++w)
                if (w == u) {
                    re = vr[w];
                    im = vi[w];
                    b = vb[w];
                }
        } else {
            load(i, re, im, b);
        }
    };
    double mn = INFINITY, mx = -INFINITY;
    for (int i = lane, u = 0; i < I; i += 64, ++u) {
        double re, im, b;
        elem(u, i, re, im, b);
        mn = fmin(mn, b);
        mx = fmax(mx, b);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mn = fmin(mn, __shfl_xor(mn, o, 64));
        mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    double range = mx - mn;
    if (range == 0.0) range = 1.0;
    double k0 = 0.78539816339744830962 / range * (double)(tix[k] - prm.tn);
    double ore = 0.0, oim = 0.0;
    for (int i = lane, u = 0; i < I; i += 64, ++u) {
        double re, im, b, sn, cs;
        elem(u, i, re, im, b);
        sincos(k0 * b, &sn, &cs);
        ore += re * cs + im * sn;
        oim += im * cs - re * sn;
    }
    ore = wave_sum_f64(ore);
    oim = wave_sum_f64(oim);
    double num = 0.0, den = 0.0;
    for (int i = lane, u = 0; i < I; i += 64, ++u) {
        double re, im, b, sn, cs;
        elem(u, i, re, im, b);
        sincos(k0 * b, &sn, &cs);
        const double rr = re * cs + im * sn, ri = im * cs - re * sn;
        const double w = hypot(re, im), wb = w * b;
        den += wb * wb;
        const double pr = rr * ore + ri * oim, pim = ri * ore - rr * oim;   // r conj(offset)
        num += wb * (w * atan2(pim, pr));
    }
    num = wave_sum_f64(num);
    den = wave_sum_f64(den);
    if (den == 0.0) den = 1.0;
    k0 += num / den;
    double mre = 0.0, mim = 0.0, dn = 0.0;
    for (int i = lane, u = 0; i < I; i += 64, ++u) {
        double re, im, b, sn, cs;
        elem(u, i, re, im, b);
        sincos(k0 * b, &sn, &cs);
        const double rr = re * cs + im * sn, ri = im * cs - re * sn;
        mre += rr;
        mim += ri;
        dn += hypot(rr, ri);
    }
    double unused = 0.0;
    wave_sum4_f64(mre, mim, dn, unused);
    if (dn == 0.0) dn = 1.0;
    const double coh = hypot(mre, mim) / dn;
    if (lane == 0) {
        Kout[k] = k0;
        Cout[k] = atan2(mim, mre);
        cohout[k] = coh;
    }
    return coh;
}

// refine_row (FP32 form) of row k by the 4 lanes of a quad: lane `sub` takes
// interferograms sub, sub + 4, ...; each lane computes its own terms (the
// expensive part: psdph, rotation, atan2) and every lane adds the quad's
// terms in ascending interferogram order.  The terms are the same rounded
// f64 values and the sums run in the same order as refine_row's, so every
// output is bit-identical to refine_row<float>; the row's work is spread over
// 4 lanes, for patches too small to give every SIMD more than one wavefront of
// rows.  Every lane returns the row's coh, max |k0 bp| and cut margin; lane 0
// writes the outputs.
template <bool ROWINV>
__device__ inline double refine_row_quad(int64_t k, int sub, const cplx<float> *__restrict__ phnm,
                                         const float *__restrict__ bp, const cplx<float> *__restrict__ filt,
                                         const int32_t *__restrict__ cell_of, const int32_t *__restrict__ tix,
                                         const float *__restrict__ rng, const cplx<float> *__restrict__ offset,
                                         double *__restrict__ Kout, double *__restrict__ Cout,
                                         double *__restrict__ cohout, double *__restrict__ nopt,
                                         const FitParams &prm, float *max_angle, float *cut_margin) {
    using R = float;
    const int t = tix[k];
    if (t < 0) {
        if (sub == 0) {
            Kout[k] = __builtin_nan("");
            Cout[k] = 0.0;
            cohout[k] = 0.0;
            nopt[k] = 0.0;
        }
        *max_angle = 0.0f;
        *cut_margin = 4.0f;
        return 0.0;
    }
    const int I = prm.I;
    const int64_t P = prm.P, C = prm.C;
    const cplx<R> *nm = phnm + k;
    const cplx<R> *fl = filt + cell_of[k];
    auto bat = [&](int i) { return ROWINV ? bp[i] : bp[(int64_t)i * P + k]; };
    const R range = rng[k];
    R k0 = (R)(RT<R>::pi4() / range * (R)(t - prm.tn));
    const cplx<R> off = offset[k];
    const R ocr = off.re, oci = -off.im;
    double num = 0.0, den = 0.0;
    R amax1 = (R)0, cut = (R)RT<R>::pi4() * (R)4;
    for (int i0 = 0; i0 < I; i0 += 4) {
        const int i = i0 + sub;
        double dd = 0.0, nn = 0.0;
        if (i < I) {
            const cplx<R> v = psdph_value<R>(fl[(int64_t)i * C], nm[(int64_t)i * P]);
            const R bi = bat(i);
            const double wgt = (double)RT<R>::abs_(v.re, v.im);
            const double wb = wgt * (double)bi;
            dd = wb * wb;
            const R ang = k0 * bi;
            const cplx<R> r = rotate_neg<R>(v, ang);
            const R pr = r.re * ocr - r.im * oci, pim = r.re * oci + r.im * ocr;
            const R a = RT<R>::atan2_(pim, pr);
            nn = wb * (wgt * (double)a);
            amax1 = fmax(amax1, fabs(ang));
            cut = fmin(cut, (R)RT<R>::pi4() * (R)4 - fabs(a));
        }
        const int n = I - i0 < 4 ? I - i0 : 4;
        quad_add_f64(den, dd, n);
        quad_add_f64(num, nn, n);
    }
    if (den == 0.0) den = 1.0;
    k0 = (R)((double)k0 + num / den);
    double mre = 0.0, mim = 0.0, dn = 0.0;
    R amax = (R)0;
    for (int i0 = 0; i0 < I; i0 += 4) {
        const int i = i0 + sub;
        R rre = (R)0, rim = (R)0, rab = (R)0;
        if (i < I) {
            const R ang = k0 * bat(i);
            const cplx<R> r = rotate_neg<R>(psdph_value<R>(fl[(int64_t)i * C], nm[(int64_t)i * P]), ang);
            amax = fmax(amax, fabs(ang));
            rre = r.re;
            rim = r.im;
            rab = RT<R>::abs_(r.re, r.im);
        }
        const int n = I - i0 < 4 ? I - i0 : 4;
        quad_add_f32_as_f64(mre, rre, n);
        quad_add_f32_as_f64(mim, rim, n);
        quad_add_f32_as_f64(dn, rab, n);
    }
    if (dn == 0.0) dn = 1.0;
    const double coh = hypot(mre, mim) / dn;
    if (sub == 0) {
        Kout[k] = (double)k0;
        Cout[k] = (double)RT<R>::atan2_((R)mim, (R)mre);
        cohout[k] = coh;
        nopt[k] = 1.0;
    }
    *max_angle = quad_max_f32(fmaxf(amax, amax1));
    *cut_margin = quad_min_f32(cut);
    return coh;
}

// The FP32 form's re-decision mark of a refined row (see refine_body).
__device__ inline uint8_t redecide_flag(int t, double coh, float amax, float cut, const double *__restrict__ mids,
                                        HistParams hp, const RecheckIn &rc) {
    const double bound = 2.0 * (8.0 + (double)amax) * 5.9604644775390625e-08;
    return t >= 0 && ((double)cut < bound + 4.0e-7 || (rc.psquare && decision_distance(coh, mids, hp) < bound)) ? 1
                                                                                                                : 0;
}

template <typename R>
__device__ __forceinline__ void refine_body(
    const cplx<R> *__restrict__ phnm,
    const R *__restrict__ bp,
    const cplx<R> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    const int32_t *__restrict__ tix,
    const R *__restrict__ rng,
    const cplx<R> *__restrict__ offset,
    double *__restrict__ Kout,
    double *__restrict__ Cout,
    double *__restrict__ cohout,
    double *__restrict__ nopt,
    FitParams prm,
    const double *__restrict__ mids,
    HistParams hp,
    RecheckIn rc,
    uint8_t *__restrict__ flags,
    int32_t *__restrict__ recheck_list,
    uint32_t *__restrict__ recheck_count,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t k = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (k >= prm.P) return;
    double coh;
    float amax = 0.0f, cut = 4.0f;
    if (prm.rowinv)
        coh = refine_row<R, true>(k, phnm, bp, filt, cell_of, tix, rng, offset, Kout, Cout, cohout, nopt, prm,
                                  &amax, &cut);
    else
        coh = refine_row<R, false>(k, phnm, bp, filt, cell_of, tix, rng, offset, Kout, Cout, cohout, nopt, prm,
                                   &amax, &cut);
    if (rc.enabled) {
        // rows near a lookup / bin boundary: bound on |coh_f32 - coh_f64| of
        // this refine, 2 (8 + max |k0 bp|) 2^-24 (per term: psdph product,
        // rotation, angle rounding <= |k0 bp| ulp; sums in f64).  Rows whose
        // least-squares step has a residual phase within the f32 angle error
        // of the +-pi cut of arg() (the same bound plus the fitted atan2's
        // 3.5e-7 rad) could take the other branch in f64.  Those rows are
        // refined again in FP64: queued for k_redecide, or (no queue) flagged
        // for k_post.
        const uint8_t fl = redecide_flag(tix[k], coh, amax, cut, mids, hp, rc);
        if (recheck_list) {
            if (fl) recheck_list[atomicAdd(recheck_count, 1u)] = (int32_t)k;
        } else {
            flags[k] = fl;
        }
    }
}

template <typename R>
struct RefineArgs {
    const cplx<R> * phnm;
    const R * bp;
    const cplx<R> * filt;
    const int32_t * cell_of;
    const int32_t * tix;
    const R * rng;
    const cplx<R> * offset;
    double * Kout;
    double * Cout;
    double * cohout;
    double * nopt;
    FitParams prm;
    const double * mids;
    HistParams hp;
    RecheckIn rc;
    uint8_t * flags;
    int32_t * recheck_list;     // non-null: flagged rows queued here (k_redecide)
    uint32_t * recheck_count;
    const int32_t * gate;
};

template <bool M, typename R>
__global__ __launch_bounds__(256) void k_refine(KArg<M, RefineArgs<R>> p) {
    if constexpr (M) {
        RefineArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        refine_body<R>(a.phnm, a.bp, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.Kout, a.Cout, a.cohout, a.nopt, a.prm, a.mids, a.hp, a.rc, a.flags, a.recheck_list, a.recheck_count, a.gate, bid_, nblk_);
    } else {
        refine_body<R>(p.phnm, p.bp, p.filt, p.cell_of, p.tix, p.rng, p.offset, p.Kout, p.Cout, p.cohout, p.nopt, p.prm, p.mids, p.hp, p.rc, p.flags, p.recheck_list, p.recheck_count, p.gate, blockIdx.x, gridDim.x);
    }
}

// FP64 re-decision (FP32 form) of the rows k_refine queued: one wavefront per
// queued row, grid-stride over the queue; refine_row_f64_wave rewrites the
// row's K, C and coh.  The rows are independent, so the queue's order (the
// order of the refine's atomics) does not change any result; spreading them
// over the whole GPU keeps a workgroup that owns many of them from setting
// the step's length.
struct RedecideArgs {
    const cplx<float> *phnm;
    Bp64 bp64;
    const cplx<float> *filt;
    const int32_t *cell_of;
    const int32_t *tix;
    double *Kout, *Cout, *cohout;
    FitParams prm;
    const int32_t *list;
    const uint32_t *count;
    const int32_t *gate;
};

__device__ __forceinline__ void redecide_body(const RedecideArgs &a, const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(a.gate)) return;
    const uint32_t n = *a.count;
    const int lane = threadIdx.x & 63;
    for (uint32_t q = bid_; q < n; q += nblk_)
        (void)refine_row_f64_wave<4>(a.list[q], a.phnm, a.bp64, a.filt, a.cell_of, a.tix, a.Kout, a.Cout, a.cohout,
                                     a.prm, lane);
}

template <bool M>
__global__ __launch_bounds__(64) void k_redecide(KArg<M, RedecideArgs> p) {
    if constexpr (M) {
        RedecideArgs a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        redecide_body(a, bid_, nblk_);
    } else {
        redecide_body(p, blockIdx.x, gridDim.x);
    }
}

// The loop's outputs in the caller's layouts (the write_pm1 arguments,
// native_stage2.rs:319-336), one launch after the last iteration:
//  - ph_weight of the last fill (fill_phase_weight :629-657, from the K and
//    weighting that fill used), ph_patch (the unit filtered cell value,
//    extract_patch_phase :1186-1205) and ph_res = arg(residual) of valid rows
//    (:714-716), [P][I] row order.  A pixel tile is TP sorted pixels x TF
//    interferograms: the inputs are read once along the planes
//    (coalesced over pixels), each output goes out through one LDS tile as
//    whole row segments (coalesced over interferograms);
//  - K, C, coh, N_opt, coh_ps_save and the trial index in row order (the
//    first interferogram tile of each pixel tile);
//  - ph_grid [C][I] from the [I][C] planes (32 x 32 transposed tiles, in the
//    blocks after the pixel tiles).
// Any output pointer may be NULL.
struct FinalOut {
    float2 *weight;        // [P][I]
    float2 *patch;         // [P][I]
    float *res;            // [P][I]
    double *vec;           // [5][P]: K, C, coh, N_opt, coh_ps_save
    int32_t *tix;          // [P]
    float2 *grid;          // [C][I]
    int64_t n_pix_blocks;  // pixel tiles: ceil(P / TP) x ceil(I / TF)
    int n_ifg_tiles;       // ceil(I / TF)
    int n_grid_ifg_tiles;  // ceil(I / 32)
};

struct FinalIn {
    const double *K, *Cc, *coh, *nopt, *coh_prev;   // sorted order, [P]
    const double *Kw, *Ww;                          // K and weighting of the last fill
    const int32_t *tix;
};

template <typename R, int TP, int TF, int NT>
__device__ __forceinline__ void final_outputs_body(
    const cplx<R> *__restrict__ phnm,
    const R *__restrict__ bp,
    int rowinv,
    const cplx<R> *__restrict__ filt,
    const cplx<R> *__restrict__ grid,
    const int32_t *__restrict__ cell_of,
    const int64_t *__restrict__ order,
    FinalIn in,
    FinalOut o,
    int64_t P,
    int I,
    int64_t C,
    const uint32_t bid_, const uint32_t nblk_) {
    static_assert(TP * TF % NT == 0 && NT % TP == 0 && NT % TF == 0 && NT % 32 == 0, "tile shape");
    static_assert(TP * (TF + 1) >= 32 * 33, "the grid transpose shares the tile");
    __shared__ float2 tile[TP][TF + 1];
    const int64_t b = bid_;
    if (b >= o.n_pix_blocks) {   // ph_grid transpose tile
        float2(*t)[33] = reinterpret_cast<float2(*)[33]>(&tile[0][0]);
        const int64_t gb = b - o.n_pix_blocks;
        const int64_t c0 = (gb / o.n_grid_ifg_tiles) * 32;
        const int f0 = (int)(gb % o.n_grid_ifg_tiles) * 32;
        const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
        for (int r = ty; r < 32; r += NT / 32) {
            const int f = f0 + r;
            const int64_t c = c0 + tx;
            if (f < I && c < C) {
                const cplx<R> v = grid[(int64_t)f * C + c];
                t[r][tx] = make_float2((float)v.re, (float)v.im);
            }
        }
        __syncthreads();
        for (int r = ty; r < 32; r += NT / 32) {
            const int64_t c = c0 + r;
            const int f = f0 + tx;
            if (f < I && c < C) o.grid[c * I + f] = t[tx][r];
        }
        return;
    }
    const int64_t k0 = (b / o.n_ifg_tiles) * TP;
    const int f0 = (int)(b % o.n_ifg_tiles) * TF;
    const int np = (int)min<int64_t>(TP, P - k0), nf = min(TF, I - f0);
    if (f0 == 0 && (int)threadIdx.x < np) {
        const int64_t k = k0 + threadIdx.x, r = order[k];
        if (o.vec) {
            o.vec[r] = in.K[k];
            o.vec[P + r] = in.Cc[k];
            o.vec[2 * P + r] = in.coh[k];
            o.vec[3 * P + r] = in.nopt[k];
            if (in.coh_prev) o.vec[4 * P + r] = in.coh_prev[k];
        }
        if (o.tix) o.tix[r] = in.tix[k];
    }
    // read side: a thread keeps one pixel of the tile (kl) and takes the
    // interferograms fl0 + 8 q; indices are clamped into the tile so that every
    // load of the thread is issued at once (out-of-tile values are not stored)
    constexpr int PER = TP * TF / NT;
    constexpr int FSTEP = NT / TP;   // interferograms between a thread's reads
    const int kl = threadIdx.x % TP, fl0 = threadIdx.x / TP;
    const int64_t k = k0 + min(kl, np - 1);
    const int32_t cell = cell_of[k];
    const R kres = (R)in.K[k];
    const bool valid = in.nopt[k] == 1.0;
    const double kw = o.weight ? in.Kw[k] : 0.0, ww = o.weight ? in.Ww[k] : 0.0;
    cplx<R> nm[PER], fv[PER];
    R bv[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int f = f0 + min(fl0 + FSTEP * q, nf - 1);
        const int64_t g = (int64_t)f * P + k;
        nm[q] = phnm[g];
        bv[q] = rowinv ? bp[f] : bp[g];
        fv[q] = filt[(int64_t)f * C + cell];
    }
    float2 wv[PER], pv[PER];
    float rv[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        wv[q] = make_float2(0.0f, 0.0f);
        if (o.weight) {
            const cplx<R> v = weight_value<R>(nm[q], bv[q], kw, ww);
            wv[q] = make_float2((float)v.re, (float)v.im);
        }
        pv[q] = make_float2((float)fv[q].re, (float)fv[q].im);
        rv[q] = 0.0f;
        if (o.res && valid) {
            const cplx<R> r = rotate_neg<R>(psdph_value<R>(fv[q], nm[q]), kres * bv[q]);
            rv[q] = (float)RT<R>::atan2_(r.im, r.re);
        }
    }
    // write side: thread t stores column t % TF of rows t / TF + KSTEP q
    constexpr int KSTEP = NT / TF;
    const int fw = threadIdx.x % TF, kw0 = threadIdx.x / TF;
    int64_t orow[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) orow[q] = order[k0 + min(kw0 + KSTEP * q, np - 1)] * I + f0 + fw;
    // one output at a time through the tile, written as row segments
    for (int which = 0; which < 3; ++which) {
        if ((which == 0 && !o.weight) || (which == 1 && !o.patch) || (which == 2 && !o.res)) continue;
#pragma unroll
        for (int q = 0; q < PER; ++q)
            tile[kl][fl0 + FSTEP * q] = which == 0 ? wv[q] : which == 1 ? pv[q] : make_float2(rv[q], 0.0f);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int kr = kw0 + KSTEP * q;
            if (kr < np && fw < nf) {
                const float2 v = tile[kr][fw];
                if (which == 0)
                    o.weight[orow[q]] = v;
                else if (which == 1)
                    o.patch[orow[q]] = v;
                else
                    o.res[orow[q]] = v.x;
            }
        }
        __syncthreads();
    }
}

template <typename R>
struct OutputsArgs {
    const cplx<R> * phnm;
    const R * bp;
    int rowinv;
    const cplx<R> * filt;
    const cplx<R> * grid;
    const int32_t * cell_of;
    const int64_t * order;
    FinalIn in;
    FinalOut o;
    int64_t P;
    int I;
    int64_t C;
};

template <bool M, typename R, int TP, int TF, int NT>
__global__ __launch_bounds__(NT) void k_final_outputs(KArg<M, OutputsArgs<R>> p) {
    if constexpr (M) {
        OutputsArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        final_outputs_body<R, TP, TF, NT>(a.phnm, a.bp, a.rowinv, a.filt, a.grid, a.cell_of, a.order, a.in, a.o, a.P, a.I, a.C, bid_, nblk_);
    } else {
        final_outputs_body<R, TP, TF, NT>(p.phnm, p.bp, p.rowinv, p.filt, p.grid, p.cell_of, p.order, p.in, p.o, p.P, p.I, p.C, blockIdx.x, gridDim.x);
    }
}

__device__ __forceinline__ void zero_u32_body(
    uint32_t * p,
    int64_t n,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t q = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (q < n) p[q] = 0u;
}

struct ZeroArgs {
    uint32_t * p;
    int64_t n;
    const int32_t * gate;
};

template <bool M>
__global__ void k_zero_u32(KArg<M, ZeroArgs> p) {
    if constexpr (M) {
        ZeroArgs a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        zero_u32_body(a.p, a.n, a.gate, bid_, nblk_);
    } else {
        zero_u32_body(p.p, p.n, p.gate, blockIdx.x, gridDim.x);
    }
}


// psquare weighting lookup (native_stage2.rs:1674-1682):
// w = (1 - prand_hi[round_half_away(1000 coh)])^2, once per pixel (large
// patches; a small patch's fill computes it, FillWeights)
__device__ __forceinline__ void weights_psquare_body(
    const double *__restrict__ coh,
    const double *__restrict__ prand_hi,
    int nhi,
    double *__restrict__ W,
    int64_t P,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t k = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (k >= P) return;
    W[k] = psquare_weight(coh[k], prand_hi, nhi);
}

struct PsquareArgs {
    const double * coh;
    const double * prand_hi;
    int nhi;
    double * W;
    int64_t P;
    const int32_t * gate;
};

template <bool M>
__global__ __launch_bounds__(256) void k_weights_psquare(KArg<M, PsquareArgs> p) {
    if constexpr (M) {
        PsquareArgs a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        weights_psquare_body(a.coh, a.prand_hi, a.nhi, a.W, a.P, a.gate, bid_, nblk_);
    } else {
        weights_psquare_body(p.coh, p.prand_hi, p.nhi, p.W, p.P, p.gate, blockIdx.x, gridDim.x);
    }
}

// snr_weighting (native_stage2.rs:1875-1897) from the current residuals.
template <typename R>
__device__ __forceinline__ void weights_snr_body(
    const cplx<R> *__restrict__ phnm,
    const float *__restrict__ amp,
    const R *__restrict__ bp,
    const cplx<R> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    const double *__restrict__ K,
    const double *__restrict__ nopt,
    double *__restrict__ W,
    int64_t P,
    int I,
    int rowinv,
    int64_t C,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    const int64_t k = (int64_t)bid_ * blockDim.x + threadIdx.x;
    if (k >= P) return;
    double g = 0.0, a2 = 0.0;
    const bool valid = nopt[k] == 1.0;
    for (int f = 0; f < I; ++f) {
        const int64_t e = (int64_t)f * P + k;
        const double a = (double)amp[e];
        double pr = 0.0;
        if (valid) {
            const cplx<R> v = psdph_value<R>(filt[(int64_t)f * C + cell_of[k]], phnm[e]);
            const R b = rowinv ? bp[f] : bp[e];
            const cplx<R> r = rotate_neg<R>(v, (R)K[k] * b);
            pr = (double)(R)RT<R>::atan2_(r.im, r.re);
        }
        g += a * cos(pr);
        a2 += a * a;
    }
    const double dd = (double)(I > 1 ? I : 1);
    g /= dd;
    a2 /= dd;
    const double sigma = sqrt(0.5 * (a2 - g * g));
    W[k] = sigma != 0.0 ? g / sigma : 0.0;
}

template <typename R>
struct SnrArgs {
    const cplx<R> * phnm;
    const float * amp;
    const R * bp;
    const cplx<R> * filt;
    const int32_t * cell_of;
    const double * K;
    const double * nopt;
    double * W;
    int64_t P;
    int I;
    int rowinv;
    int64_t C;
    const int32_t * gate;
};

template <bool M, typename R>
__global__ __launch_bounds__(256) void k_weights_snr(KArg<M, SnrArgs<R>> p) {
    if constexpr (M) {
        SnrArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        weights_snr_body<R>(a.phnm, a.amp, a.bp, a.filt, a.cell_of, a.K, a.nopt, a.W, a.P, a.I, a.rowinv, a.C, a.gate, bid_, nblk_);
    } else {
        weights_snr_body<R>(p.phnm, p.amp, p.bp, p.filt, p.cell_of, p.K, p.nopt, p.W, p.P, p.I, p.rowinv, p.C, p.gate, blockIdx.x, gridDim.x);
    }
}


// ---------------------------------------------------------------------------
// Device-resident loop control.  One control kernel per iteration takes the
// stop decision (native_stage2.rs:307-314) and, when the loop goes on with
// P-square weighting, rebuilds the lookup table prand_hi (:364-381 and
// psquare_weighting :1642-1748) with the host restatement's operation order;
// later kernels of the iteration and of further iterations return at entry
// once ctl->stopped is set.
constexpr int CTL_BINS = 100;
constexpr int CTL_HIST_WORDS = 128;   // coherence histogram buffer (CTL_BINS used)
constexpr int CTL_HI = 1001;      // interp([1, prand], 10) keeps 1010 - 9 samples
constexpr int CTL_UP = 1049;      // 101 * 10 + 39-sample delay

struct LoopCtl {
    int32_t stopped;
    int32_t i_loop;       // current iteration (1-based), final count once stopped
    uint64_t rescored;    // FP32 form: rows re-scored in FP64, summed
    uint64_t rechecked;   // FP32 form: rows refined again in FP64 (P-square choices), summed
    double gcs;           // gamma_change_save
    double nr[CTL_BINS];  // Nr, scaled each P-square update
};

// The loop's starting state as a kernel (the captured loop of one patch and
// the merged launches of several start with it; a loop launched kernel by
// kernel sets the same state with memsets and copies): K = 0, coh_prev = 0,
// W = W0, the first iteration's
// histogram and re-score counters zero, and the control block with i_loop 1
// and Nr = the base histogram nr0.  Not gated.
struct PrepareArgs {
    double *K, *cohprev, *W;
    const double *W0;
    uint32_t *hist, *counters;
    LoopCtl *ctl;
    const double *nr0;
    int64_t P;
};

template <bool M>
__global__ __launch_bounds__(256) void k_loop_prepare(KArg<M, PrepareArgs> p) {
    PrepareArgs a;
    uint32_t bid, nblk;
    if constexpr (M) {
        if (!multi_locate(p, a, bid, nblk)) return;
    } else {
        a = p;
        bid = blockIdx.x;
        nblk = gridDim.x;
    }
    (void)nblk;
    const int64_t k = (int64_t)bid * 256 + threadIdx.x;
    if (k < a.P) {
        a.K[k] = 0.0;
        a.cohprev[k] = 0.0;
        a.W[k] = a.W0[k];
    }
    if (bid == 0) {
        for (int w = threadIdx.x; w < CTL_HIST_WORDS; w += 256) a.hist[w] = 0u;
        if (threadIdx.x < 4) a.counters[threadIdx.x] = 0u;
        for (int b = threadIdx.x; b < CTL_BINS; b += 256) a.ctl->nr[b] = a.nr0[b];
        if (threadIdx.x == 0) {
            a.ctl->stopped = 0;
            a.ctl->i_loop = 1;
            a.ctl->rescored = 0;
            a.ctl->rechecked = 0;
            a.ctl->gcs = 0.0;
        }
    }
}

// The control block copied to coherent pinned host memory, the last kernel
// of a captured loop: after its wait on the stream the host reads the loop's
// results there, with no copy queued behind the loop.  Not gated.
__global__ __launch_bounds__(64) void k_publish_ctl(const LoopCtl *__restrict__ src, LoopCtl *__restrict__ dst) {
    static_assert(sizeof(LoopCtl) % 8 == 0, "whole 8-byte words");
    constexpr int NW = (int)(sizeof(LoopCtl) / 8);
    const uint64_t *s = reinterpret_cast<const uint64_t *>(src);
    uint64_t *d = reinterpret_cast<uint64_t *>(dst);
    for (int i = threadIdx.x; i < NW; i += 64) d[i] = s[i];
    __threadfence_system();
}

struct CtlParams {
    int64_t P;
    int n_rms_blocks;
    int max_iter;
    int psquare;          // 1: P-square weights follow (else SNR)
    int lct;              // low_coh_thresh
    double conv;          // gamma_change_convergence
    double nr_max_nz_ix;
    double g7[7];         // gausswin(7, 2.5) taps and their sum
    double gsum;
};

// MATLAB interp(x, 10) FIR: the reference's 79 audited taps (native_stage2.rs:33-113)
__constant__ double kCtlInterpTaps[79] = {
    -7.92925708923623813e-05, -1.73293349766172333e-04, -3.00248431516961407e-04,
    -4.48605550383722462e-04, -5.95835452295412349e-04, -7.08759204842826507e-04,
    -7.46231412934915406e-04, -6.64375214040318490e-04, -4.24121369705182660e-04,
    -3.01053961794528480e-07, 6.08939001185030554e-04,  1.37399755086099230e-03,
    2.22872001918483081e-03,  3.06999955141492101e-03,  3.76333143346069959e-03,
    4.15488007985555935e-03,  4.08970016389962471e-03,  3.43472424182160452e-03,
    2.10413493213892910e-03,  8.39421246326618742e-05,  -2.54786253685349538e-03,
    -5.60918261221970971e-03, -8.81278723412589476e-03, -1.17783753398490003e-02,
    -1.40582499688596801e-02, -1.51753768689126753e-02, -1.46708218626883501e-02,
    -1.21559904304523667e-02, -7.36396963079587880e-03, -1.93792008491816548e-04,
    9.25827682864269544e-03,  2.06852593184593814e-02,  3.35768644427044682e-02,
    4.72485589714582085e-02,  6.08905923193367796e-02,  7.36328154139928304e-02,
    8.46190871379102622e-02,  9.30836882813813049e-02,  9.84216619119140101e-02,
    1.00245459871706619e-01,  9.84216619119134828e-02,  9.30836882813811800e-02,
    8.46190871379124271e-02,  7.36328154139975766e-02,  6.08905923193391457e-02,
    4.72485589714637388e-02,  3.35768644427109145e-02,  2.06852593184574732e-02,
    9.25827682864464527e-03,  -1.93792008489998774e-04, -7.36396963079247527e-03,
    -1.21559904304541604e-02, -1.46708218626876996e-02, -1.51753768689126579e-02,
    -1.40582499688608267e-02, -1.17783753398508096e-02, -8.81278723413165405e-03,
    -5.60918261221661496e-03, -2.54786253685826761e-03, 8.39421246353382272e-05,
    2.10413493213411785e-03,  3.43472424182329154e-03,  4.08970016390024574e-03,
    4.15488007985538154e-03,  3.76333143346146634e-03,  3.06999955141576842e-03,
    2.22872001918783102e-03,  1.37399755086213202e-03,  6.08939001188346478e-04,
    -3.01053965413812172e-07, -4.24121369703597340e-04, -6.64375214039929912e-04,
    -7.46231412935480926e-04, -7.08759204842192465e-04, -5.95835452295058465e-04,
    -4.48605550383040716e-04, -3.00248431517695195e-04, -1.73293349768269614e-04,
    -7.92925708931989381e-05,
};

// run by one workgroup of 256 threads (k_control)
__device__ inline void control_step(LoopCtl *__restrict__ ctl, const double *__restrict__ rms_part,
                                    uint32_t *__restrict__ hist, uint32_t *__restrict__ rescore_count,
                                    double *__restrict__ prand_hi, const CtlParams &cp) {
    constexpr int CTL_THREADS = 256;   // the calling workgroup
    __shared__ double red[CTL_THREADS];
    __shared__ double na[128], nr[128], prand[128], sm[128];
    __shared__ double ext[112];   // [1, smoothed prand], zeros after
    __shared__ int s_stop;
    __shared__ double s_scale;
    if (ctl->stopped) return;
    const int t = threadIdx.x;
    // every global input is read here, before the first barrier, so the
    // phases below wait on one round of loads instead of one per phase
    const double h_t = t < CTL_BINS ? (double)hist[t] : 0.0;
    const double nr_t = t < CTL_BINS ? ctl->nr[t] : 0.0;
    const double gcs0 = ctl->gcs;
    const int i_loop0 = ctl->i_loop;
    const uint32_t nresc = rescore_count ? rescore_count[0] : 0u;
    const uint32_t nrech = rescore_count ? rescore_count[1] : 0u;
    // gamma-change RMS: fixed assignment of block partials to threads, then a
    // fixed tree (deterministic)
    double acc = 0.0;
    for (int q = t; q < cp.n_rms_blocks; q += CTL_THREADS) acc += rms_part[q];
    red[t] = acc;
    __syncthreads();
    for (int o = CTL_THREADS / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    if (t == 0) {
        const double rms = sqrt(red[0] / (double)(cp.P > 1 ? cp.P : 1));
        const double change = rms - gcs0;
        ctl->gcs = rms;
        if (rescore_count) {
            ctl->rescored += nresc;
            ctl->rechecked += nrech;
        }
        s_stop = (fabs(change) < cp.conv || i_loop0 >= cp.max_iter) ? 1 : 0;
    }
    __syncthreads();
    if (s_stop) {
        if (t == 0) ctl->stopped = 1;
        return;
    }
    // the next iteration's histogram and re-score counter start at zero
    // (read above and below before they are cleared)
    auto clear_counters = [&]() {
        if (t < CTL_HIST_WORDS) hist[t] = 0u;
        if (t < 4 && rescore_count) rescore_count[t] = 0u;
    };
    if (!cp.psquare) {
        clear_counters();
        if (t == 0) ctl->i_loop = i_loop0 + 1;
        return;
    }
    // Nr scaling on the low-coherence bins (native_stage2.rs:365-374)
    const int lct = cp.lct < 0 ? 0 : (cp.lct > CTL_BINS ? CTL_BINS : cp.lct);
    if (t < CTL_BINS) {
        na[t] = h_t;
        nr[t] = nr_t;
    }
    clear_counters();
    __syncthreads();
    if (t == 0) {
        // in bin order, ten bins' loads issued at a time (b0 + j <= 99)
        double denom = 0.0, num = 0.0;
        for (int b0 = 0; b0 < lct; b0 += 10) {
            double vr[10], va[10];
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                vr[j] = nr[b0 + j];
                va[j] = na[b0 + j];
            }
#pragma unroll
            for (int j = 0; j < 10; ++j) {
                if (b0 + j < lct) {
                    denom += vr[j];
                    num += va[j];
                }
            }
        }
        s_scale = denom > 0.0 ? num / denom : 1.0;
    }
    __syncthreads();
    // prand = min(Nr / Na, 1), [:lct] = 1, [Nr_max_nz_ix:] = 0 (:1642-1683)
    const int zero_from = cp.nr_max_nz_ix > 0.0 ? (int)cp.nr_max_nz_ix : 0;
    if (t < CTL_BINS) {
        nr[t] *= s_scale;
        ctl->nr[t] = nr[t];
        const double q = nr[t] / (na[t] == 0.0 ? 1.0 : na[t]);
        double pr = (__builtin_isnan(q) || q > 1.0) ? 1.0 : q;
        if (t < lct) pr = 1.0;
        if (t >= zero_from) pr = 0.0;
        prand[t] = pr;
    }
    __syncthreads();
    // gausswin(7) FIR over [1 x 7, prand], normalised (:1685-1713)
    if (t < CTL_BINS) {
        const int n = t + 7;
        double sacc = 0.0;
        for (int k = 0; k < 7; ++k) {
            const int m = n - k;
            sacc += cp.g7[k] * (m < 7 ? 1.0 : prand[m - 7]);
        }
        sm[t] = sacc / cp.gsum;
    }
    __syncthreads();
    // interp([1, prand], 10) (:1715-1748): the zero-stuffed sequence up
    // (up[10 m] = ext[m], zeros between; up[i] = 0 for i >= 1049) through
    // the 79-tap FIR, prand_hi[q] = 10 sum_{k = 0..min(78, n)} tap[k]
    // up[n - k], n = 39 + q.  Only the taps k = n mod 10 + 10 j meet a
    // non-zero input; the others add a zero product, which leaves a sum
    // that starts at +0 unchanged, so the sum of those few terms in tap
    // order is the whole sum, bit for bit.
    if (t < 112) ext[t] = t == 0 ? 1.0 : (t <= CTL_BINS ? sm[t - 1] : 0.0);
    __syncthreads();
    for (int q = t; q < CTL_HI; q += CTL_THREADS) {
        const int n = 39 + q;
        const int r = n % 10;
        double sacc = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = r + 10 * j;
            // n - k = 10 m with m <= 103 (ext has 112 entries)
            if (k <= 78 && k <= n && n - k < CTL_UP) sacc += kCtlInterpTaps[k] * ext[(n - k) / 10];
        }
        prand_hi[q] = sacc * 10.0;
    }
    if (t == 0) ctl->i_loop = i_loop0 + 1;
}

// Post-refine step of one gamma iteration, one launch of ceil(P / 256)
// workgroups of 256 rows:
//  1. (FP32 form) the rows k_refine flagged are refined again in FP64, one
//     wavefront per row (refine_row_f64_wave);
//  2. per workgroup: the gamma-change sum of (coh - coh_prev)^2 in a fixed
//     tree (one partial per workgroup), coh_prev <- coh, and the coherence
//     histogram (integer, order-free);
// and k_control (one workgroup) then takes the RMS over the partials in a
// fixed order, the stop test, the Nr scaling and the P-square table.  (A
// last-workgroup-runs-the-control variant needs a device-scope release fence
// in every workgroup, which writes back the XCD's L2 each time: 134-261 us
// at 500k x 60 instead of a second launch.)  loop = 0 (the single fit step):
// step 1 only.  counters: [0] re-score queue, [1] rows refined again;
// control_step clears them for the next iteration.
template <typename R, int GR = 256>
__device__ __forceinline__ void post_body(
    const cplx<R> *__restrict__ phnm,
    const cplx<R> *__restrict__ filt,
    const int32_t *__restrict__ cell_of,
    const int32_t *__restrict__ tix,
    double *__restrict__ Kout,
    double *__restrict__ Cout,
    double *__restrict__ cohout,
    FitParams prm,
    RecheckIn rc,
    const uint8_t *__restrict__ flags,
    RmsHistOut ro,
    int loop,
    uint32_t *__restrict__ counters,
    const int32_t *__restrict__ gate,
    const uint32_t bid_, const uint32_t nblk_) {
    if (loop_stopped(gate)) return;
    // groups of GR (256 or 64) rows, walked grid-stride: one RMS partial per
    // group (the same fixed tree whatever the grid), the block's histogram
    // counts in LDS and flushed once (one global atomic per non-empty bin and
    // workgroup).  Threads 0 .. GR-1 hold the group's rows.
    static_assert(GR == 256 || GR == 64, "post step groups of 256 or 64 rows");
    __shared__ uint32_t lh[128];
    __shared__ double red[256];
    __shared__ int n_flag;
    __shared__ int flag_tid[256];
    __shared__ double new_coh[256];
    if (threadIdx.x < 128) lh[threadIdx.x] = 0u;
    __syncthreads();   // the bins are zero before any wavefront counts into them
    uint32_t n_rechecked = 0;
    const int64_t n_groups = (prm.P + GR - 1) / GR;
    // the next group's coh and coh_prev are loaded while this one is processed
    auto fetch = [&](int64_t gg, double &c, double &cp) {
        const int64_t kk = gg * GR + threadIdx.x;
        const bool r = (int)threadIdx.x < GR && kk < prm.P;
        c = r ? cohout[kk] : 0.0;
        cp = r && loop ? ro.coh_prev[kk] : 0.0;
    };
    double coh_next = 0.0, prev_next = 0.0;
    if ((int64_t)bid_ < n_groups) fetch(bid_, coh_next, prev_next);
    for (int64_t g = bid_; g < n_groups; g += nblk_) {
        const int64_t k = g * GR + threadIdx.x;
        const bool row = (int)threadIdx.x < GR && k < prm.P;
        double coh = coh_next;
        const double coh_prev = prev_next;
        if (g + nblk_ < n_groups) fetch(g + nblk_, coh_next, prev_next);
        if constexpr (sizeof(R) == 4) {
            if (rc.enabled) {
                if (threadIdx.x == 0) n_flag = 0;
                __syncthreads();
                const bool flag = row && flags[k] != 0;
                if (flag) flag_tid[atomicAdd(&n_flag, 1)] = threadIdx.x;
                __syncthreads();
                const int nf = n_flag;
                const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
                for (int q = wave; q < nf; q += 4) {
                    const int t = flag_tid[q];
                    const double c = refine_row_f64_wave<4>(g * GR + t, reinterpret_cast<const cplx<float> *>(phnm),
                                                         rc.bp64, reinterpret_cast<const cplx<float> *>(filt),
                                                         cell_of, tix, Kout, Cout, cohout, prm, lane);
                    if (lane == 0) new_coh[t] = c;
                }
                __syncthreads();
                if (flag) coh = new_coh[threadIdx.x];
                n_rechecked += (uint32_t)nf;
            }
        }
        if (loop) {
            double d2 = 0.0;
            if (row) {
                const double d = coh - coh_prev;
                d2 = d * d;
                ro.coh_prev[k] = coh;
                const int bin = hist_bin(coh, ro.mids, ro.hp);
                if (bin >= 0) atomicAdd(&lh[bin], 1u);
            }
            // fixed fold tree over the group's rows, s[t] += s[t + o] for
            // o = GR/2 .. 1: the steps across wavefronts through LDS, the
            // steps inside wavefront 0 (o <= 32) by shuffles (the same pairs
            // and order)
            double v = d2;
            if constexpr (GR == 256) {
                red[threadIdx.x] = d2;
                __syncthreads();
                if (threadIdx.x < 128) red[threadIdx.x] += red[threadIdx.x + 128];
                __syncthreads();
                if (threadIdx.x < 64) v = red[threadIdx.x] + red[threadIdx.x + 64];
                __syncthreads();   // red is reused by the next group
            }
            if (threadIdx.x < 64) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
                if (threadIdx.x == 0) ro.part[g] = v;
            }
        }
    }
    if (threadIdx.x == 0 && n_rechecked > 0) atomicAdd(&counters[1], n_rechecked);
    if (!loop) return;
    __syncthreads();
    for (int q = threadIdx.x; q < ro.hp.nbins; q += 256)
        if (lh[q]) atomicAdd(&ro.hist[q], lh[q]);
}

template <typename R>
struct PostArgs {
    const cplx<R> * phnm;
    const cplx<R> * filt;
    const int32_t * cell_of;
    const int32_t * tix;
    double * Kout;
    double * Cout;
    double * cohout;
    FitParams prm;
    RecheckIn rc;
    const uint8_t * flags;
    RmsHistOut ro;
    int loop;
    uint32_t * counters;
    const int32_t * gate;
};

template <bool M, typename R>
__global__ __launch_bounds__(256) void k_post(KArg<M, PostArgs<R>> p) {
    if constexpr (M) {
        PostArgs<R> a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        post_body<R>(a.phnm, a.filt, a.cell_of, a.tix, a.Kout, a.Cout, a.cohout, a.prm, a.rc, a.flags, a.ro, a.loop, a.counters, a.gate, bid_, nblk_);
    } else {
        post_body<R>(p.phnm, p.filt, p.cell_of, p.tix, p.Kout, p.Cout, p.cohout, p.prm, p.rc, p.flags, p.ro, p.loop, p.counters, p.gate, blockIdx.x, gridDim.x);
    }
}

__device__ __forceinline__ void control_body(
    LoopCtl *__restrict__ ctl,
    const double *__restrict__ rms_part,
    uint32_t *__restrict__ hist,
    uint32_t *__restrict__ counters,
    double *__restrict__ prand_hi,
    CtlParams cp,
    const uint32_t bid_, const uint32_t nblk_) {
    control_step(ctl, rms_part, hist, counters, prand_hi, cp);
}

struct ControlArgs {
    LoopCtl * ctl;
    const double * rms_part;
    uint32_t * hist;
    uint32_t * counters;
    double * prand_hi;
    CtlParams cp;
};

template <bool M>
__global__ __launch_bounds__(256) void k_control(KArg<M, ControlArgs> p) {
    if constexpr (M) {
        ControlArgs a;
        uint32_t bid_, nblk_;
        if (!multi_locate(p, a, bid_, nblk_)) return;
        control_body(a.ctl, a.rms_part, a.hist, a.counters, a.prand_hi, a.cp, bid_, nblk_);
    } else {
        control_body(p.ctl, p.rms_part, p.hist, p.counters, p.prand_hi, p.cp, blockIdx.x, gridDim.x);
    }
}


// --- small patches: fewer launches per iteration ------------------------------
// A patch of at most 65,536 rows has at most one 256-row group per CU, so its
// per-row kernels leave most of the device idle and each launch costs its
// fixed start-up.  For such patches the loop runs k_fit_tail in place of
// k_rescore, k_refine and k_post: per 256-row group, the FP32 form's FP64
// re-score of the group's queued rows (one wavefront per row), then the
// refine of the group's rows, then the post step of the group (the FP64
// re-decisions, RMS partial, coh_prev, histogram) -- the same device
// functions and arithmetic, so the results are identical.  (The control step
// and the P-square weights stay two launches: one workgroup looking up 20,000
// weights took 40 us, the separate weights kernel 7.)
template <typename R>
struct FitTailArgs {
    const cplx<R> *phnm;
    const R *bp;
    Bp64 bp64;
    const cplx<R> *filt;
    const int32_t *cell_of;
    int32_t *tix;
    const R *rng;
    cplx<R> *offset;
    double *K, *Cc, *coh, *nopt;
    const int32_t *rescore_list;
    uint32_t *counters;   // [0] re-score queue length, [1] rows refined again
    const cplx<R> *wt;    // unit-step table of the scan (k_fit_small with the scan)
    FitParams prm;
    const double *mids;
    HistParams hp;
    RecheckIn rc;
    uint8_t *flags;
    RmsHistOut ro;
    int rescore;          // 1: FP32 form (re-score the queued rows)
    const int32_t *gate;
};

template <bool M, typename R>
__global__ __launch_bounds__(256) void k_fit_tail(KArg<M, FitTailArgs<R>> p) {
    FitTailArgs<R> a;
    uint32_t bid, nblk;
    if constexpr (M) {
        if (!multi_locate(p, a, bid, nblk)) return;
    } else {
        a = p;
        bid = blockIdx.x;
        nblk = gridDim.x;
    }
    if (a.rescore) {
        const int64_t r0 = (int64_t)bid * 256;
        rescore_body<R>(a.phnm, a.bp64, a.filt, a.cell_of, a.rescore_list, a.counters, a.tix, a.offset, a.prm,
                        a.gate, r0, r0 + 256, threadIdx.x >> 6, 4);
        __syncthreads();
    }
    refine_body<R>(a.phnm, a.bp, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.K, a.Cc, a.coh, a.nopt, a.prm,
                   a.mids, a.hp, a.rc, a.flags, nullptr, nullptr, a.gate, bid, nblk);
    __syncthreads();
    post_body<R>(a.phnm, a.filt, a.cell_of, a.tix, a.K, a.Cc, a.coh, a.prm, a.rc, a.flags, a.ro, 1, a.counters,
                 a.gate, bid, nblk);
}

// The same step for the FP32 form's small patches in groups of 64 rows, one
// row per quad of lanes in the refine (refine_row_quad: bit-identical values,
// 4x the wavefronts): per group the FP64 re-score of its queued rows, the
// refine, then the post step (re-decisions, RMS partial of the 64 rows,
// coh_prev, histogram).  A patch of 20,000 rows gives 79 groups of 256 rows,
// one wavefront on 316 of the 1,024 SIMDs, each thread's dependent chain of
// sincos/atan2 evaluations the bound; in 64-row groups it spreads over every
// CU with a quarter of the chain per lane.
// With SCAN (8-trial chunks, NQ as in k_scan_quad), the group's trial scan
// runs first in the same workgroup (scan_quad_body, the same arithmetic as
// k_scan_quad) and its near-tie rows are queued in LDS for the re-score: one
// launch and no global queue per iteration; the queued-row count still goes to
// counters[0] for the loop's statistics.
constexpr int SMALL_GROUP = 64;
template <bool M, bool SCAN = false, int NQ = 0>
__global__ __launch_bounds__(256) void k_fit_small(KArg<M, FitTailArgs<float>> p) {
    FitTailArgs<float> a;
    uint32_t bid, nblk;
    if constexpr (M) {
        if (!multi_locate(p, a, bid, nblk)) return;
    } else {
        a = p;
        bid = blockIdx.x;
        nblk = gridDim.x;
    }
    if (loop_stopped(a.gate)) return;
    const int64_t r0 = (int64_t)bid * SMALL_GROUP;
    if constexpr (SCAN) {
        __shared__ int32_t q_rows[SMALL_GROUP];
        __shared__ int q_n;
        if (threadIdx.x == 0) q_n = 0;
        __syncthreads();
        scan_quad_body<NQ>(a.phnm, a.bp, a.wt, a.filt, a.cell_of, a.tix, a.rng, a.offset, nullptr, nullptr, a.prm,
                           a.gate, bid, q_rows, &q_n);
        __syncthreads();
        const int nq = q_n;
        if (threadIdx.x == 0 && nq > 0) atomicAdd(&a.counters[0], (uint32_t)nq);
        if (nq > 0) {
            const uint32_t cnt = (uint32_t)nq;
            rescore_body<float>(a.phnm, a.bp64, a.filt, a.cell_of, q_rows, &cnt, a.tix, a.offset, a.prm, a.gate,
                                r0, r0 + SMALL_GROUP, threadIdx.x >> 6, 4);
        }
    } else {
        rescore_body<float>(a.phnm, a.bp64, a.filt, a.cell_of, a.rescore_list, a.counters, a.tix, a.offset, a.prm,
                            a.gate, r0, r0 + SMALL_GROUP, threadIdx.x >> 6, 4);
    }
    __syncthreads();
    {
        const int64_t k = r0 + (threadIdx.x >> 2);
        const int sub = threadIdx.x & 3;
        if (k < a.prm.P) {   // (a quad's lanes share k)
            float amax, cut;
            const double coh =
                a.prm.rowinv
                    ? refine_row_quad<true>(k, sub, a.phnm, a.bp, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.K,
                                            a.Cc, a.coh, a.nopt, a.prm, &amax, &cut)
                    : refine_row_quad<false>(k, sub, a.phnm, a.bp, a.filt, a.cell_of, a.tix, a.rng, a.offset, a.K,
                                             a.Cc, a.coh, a.nopt, a.prm, &amax, &cut);
            if (a.rc.enabled && sub == 0) a.flags[k] = redecide_flag(a.tix[k], coh, amax, cut, a.mids, a.hp, a.rc);
        }
    }
    __syncthreads();
    post_body<float, SMALL_GROUP>(a.phnm, a.filt, a.cell_of, a.tix, a.K, a.Cc, a.coh, a.prm, a.rc, a.flags, a.ro, 1,
                                  a.counters, a.gate, bid, nblk);
}

}  // namespace ps2
