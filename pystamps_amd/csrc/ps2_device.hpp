// Device helpers shared by the stage-2 kernels (gfx950).
//
// Arithmetic conventions: the library is compiled with -ffp-contract=off so a
// product followed by a sum rounds twice, as in the reference (rustc never
// fuses).  Where a kernel wants a fused multiply-add it calls fma()/fmaf()
// explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

namespace ps2 {

template <typename R>
struct cplx {
    R re, im;
};

// FP32 sine and cosine of the per-element rotation angles.  x = q pi/2 + r with
// q = rint(x 2/pi) and pi/2 split into three floats (their sum is pi/2 to
// ~2^-74), subtracted by fused multiply-adds, so r keeps its own relative
// precision for |x| <= 8192.  sin r = r + r s (S1 + s S2 + s^2 S3) and
// cos r = 1 + s (C1 + ... + s^3 C4), s = r^2, with coefficients fitted to sin
// and cos on |r| <= 1.002 pi/4 by reweighted least squares (Lawson's
// minimax iteration); q mod 4 swaps and negates.  Evaluated in float
// arithmetic over 4M arguments in |x| <= 8192, both stay within 1.5 ulp of
// the float64 values.  Wider arguments and non-finite ones go to sincosf.
// sin(-0) comes out +0 (the reduction adds +0); the rotations that use it
// sum the products with other terms, so no result's sign depends on it.
__device__ inline void sincos_fit(float x, float *sp, float *cp) {
    if (__builtin_expect(!(__builtin_fabsf(x) <= 8192.0f), 0)) {
        sincosf(x, sp, cp);
        return;
    }
    const float q = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-q, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-q, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-q, -0x1.ee59dap-50f, r);
    const float s = r * r;
    const float ps = __builtin_fmaf(__builtin_fmaf(-0x1.993ce4p-13f, s, 0x1.110726p-7f), s, -0x1.555544p-3f);
    const float sn = __builtin_fmaf(r * s, ps, r);
    const float pc = __builtin_fmaf(
        __builtin_fmaf(__builtin_fmaf(0x1.990f5ep-16f, s, -0x1.6c07d2p-10f), s, 0x1.55553cp-5f), s, -0.5f);
    const float cn = __builtin_fmaf(s, pc, 1.0f);
    const int qi = (int)q;
    const bool odd = (qi & 1) != 0;
    // sin(q pi/2 + r): sin r, cos r, -sin r, -cos r for q mod 4 = 0..3;
    // cos(q pi/2 + r): cos r, -sin r, -cos r, sin r
    const float S = odd ? cn : sn, Cv = odd ? sn : cn;
    *sp = __int_as_float(__float_as_int(S) ^ ((qi & 2) << 30));
    *cp = __int_as_float(__float_as_int(Cv) ^ (((qi + 1) & 2) << 30));
}

// FP32 atan2 for finite arguments ((0, 0) gives 0 with the sign of y).
// t = min(|x|,|y|) / max(|x|,|y|) in [0, 1] by the hardware reciprocal,
// atan t = t + t s P(s), s = t^2, P of degree 7 fitted like sincos_fit
// (within 8e-8 of atan on [0, 1] in float arithmetic), then pi/2 - a when
// |y| > |x|, pi - a when x < 0, and the sign of y.  Within 3e-7 rad (about
// one ulp of pi) of the float64 atan2 over 4M arguments of magnitude
// e^-20..e^5.
__device__ inline float atan2_fit(float y, float x) {
    const float ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    const float mx = __builtin_fmaxf(__builtin_fmaxf(ax, ay), 1e-30f);
    const float mn = __builtin_fminf(ax, ay);
    const float t = mn * __builtin_amdgcn_rcpf(mx);
    const float s = t * t;
    float p = 0x1.57b160p-9f;
    p = __builtin_fmaf(p, s, -0x1.efda4ep-7f);
    p = __builtin_fmaf(p, s, 0x1.50ddb2p-5f);
    p = __builtin_fmaf(p, s, -0x1.2dbd08p-4f);
    p = __builtin_fmaf(p, s, 0x1.b11b7ap-4f);
    p = __builtin_fmaf(p, s, -0x1.228756p-3f);
    p = __builtin_fmaf(p, s, 0x1.996740p-3f);
    p = __builtin_fmaf(p, s, -0x1.55546cp-2f);
    float a = __builtin_fmaf(t * s, p, t);
    if (ay > ax) a = 0x1.921fb6p+0f - a;
    if (x < 0.0f) a = 0x1.921fb6p+1f - a;
    const float xy = x + y;  // NaN in, NaN out (fmaxf/fminf drop NaNs)
    return xy != xy ? xy : __builtin_copysignf(a, y);
}

// Real-type traits: the FP32 form runs on float, the plain FP64 form on double.
template <typename R>
struct RT;

template <>
struct RT<float> {
    using key_t = uint32_t;
    static constexpr int key_top_bit = 30;  // keys are of non-negative floats
    __device__ static inline float pi4() { return 0.785398185253143310546875f; }  // f32(pi)/4
    __device__ static inline void sincos_(float x, float *s, float *c) { sincos_fit(x, s, c); }
    __device__ static inline float atan2_(float y, float x) { return atan2_fit(y, x); }
    __device__ static inline float sqrt_(float x) { return sqrtf(x); }
    __device__ static inline float pow_(float x, float y) { return powf(x, y); }
    // glibc hypotf: sqrt of the f64 sum of f64 squares, rounded once to f32
    // (both squares are exact in f64).
    __device__ static inline float hypot_(float x, float y) {
        float ax = fabsf(x), ay = fabsf(y);
        if (__builtin_isinf(ax) || __builtin_isinf(ay)) return INFINITY;
        if (__builtin_isnan(ax) || __builtin_isnan(ay)) return ax * ay;
        if (ax == 0.0f) return ay;
        if (ay == 0.0f) return ax;
        double dx = (double)x, dy = (double)y;
        return (float)__builtin_sqrt(dx * dx + dy * dy);
    }
    // |z| where glibc fidelity is not needed (CLAP spectrum magnitude)
    __device__ static inline float abs_(float x, float y) {
        return __builtin_amdgcn_sqrtf(fmaf(x, x, y * y));
    }
    // CLAP gain normalisation h / median: one reciprocal per window
    __device__ static inline float gain_scale(float med) { return med != 0.0f ? 1.0f / med : 1.0f; }
    __device__ static inline float gain_div(float h, float, float scl) { return h * scl; }
    // h / median - 1 in one rounding
    __device__ static inline float gain_rel_minus_one(float h, float, float scl) {
        return fmaf(h, scl, -1.0f);
    }
    __device__ static inline key_t key(float v) { return __float_as_uint(v); }
    __device__ static inline float from_key(key_t k) { return __uint_as_float(k); }
    __device__ static inline bool isnan_(float v) { return __builtin_isnan(v); }
};

template <>
struct RT<double> {
    using key_t = uint64_t;
    static constexpr int key_top_bit = 62;
    __device__ static inline double pi4() { return 0.78539816339744830962; }
    __device__ static inline void sincos_(double x, double *s, double *c) { sincos(x, s, c); }
    __device__ static inline double atan2_(double y, double x) { return atan2(y, x); }
    __device__ static inline double sqrt_(double x) { return sqrt(x); }
    __device__ static inline double pow_(double x, double y) { return pow(x, y); }
    __device__ static inline double hypot_(double x, double y) { return hypot(x, y); }
    __device__ static inline double abs_(double x, double y) { return hypot(x, y); }
    __device__ static inline double gain_scale(double) { return 1.0; }
    __device__ static inline double gain_div(double h, double med, double) {
        return med != 0.0 ? h / med : h;
    }
    __device__ static inline double gain_rel_minus_one(double h, double med, double s) {
        return gain_div(h, med, s) - 1.0;
    }
    __device__ static inline key_t key(double v) { return (uint64_t)__double_as_longlong(v); }
    __device__ static inline double from_key(key_t k) { return __longlong_as_double((long long)k); }
    __device__ static inline bool isnan_(double v) { return __builtin_isnan(v); }
};

// num-complex Complex<T> * Complex<T>: (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re)
template <typename R>
__device__ inline cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

// cos(k*pi/16), k = 0..8
template <typename R>
__device__ constexpr R c16(int k) {
    constexpr double t[9] = {1.0,
                             0.98078528040323044913,
                             0.92387953251128675613,
                             0.83146961230254523708,
                             0.70710678118654752440,
                             0.55557023301960222474,
                             0.38268343236508977173,
                             0.19509032201612826785,
                             0.0};
    return (R)t[k];
}
// e^{-2 pi i k / 32} = cos(k pi/16) - i sin(k pi/16), 0 <= k < 16
template <typename R>
__device__ constexpr R twr(int k) {
    return k <= 8 ? c16<R>(k) : -c16<R>(16 - k);
}
template <typename R>
__device__ constexpr R twi_fwd(int k) {
    return -(k <= 8 ? c16<R>(8 - k) : c16<R>(k - 8));
}

__device__ constexpr int bitrev5(int v) {
    return ((v & 1) << 4) | ((v & 2) << 2) | (v & 4) | ((v & 8) >> 2) | ((v & 16) >> 4);
}

// FP32 butterflies with a general twiddle use the fused form below; the FP64
// form keeps the separately rounded product and sums.
template <typename R>
constexpr bool FMA_BUTTERFLY = sizeof(R) == 4;

// 32-point radix-2 decimation-in-time FFT held in registers, unnormalised:
// forward X[k] = sum x[n] e^{-2 pi i nk/32}, inverse with e^{+...} (the
// rustfft conventions of native_stage2.rs:1055-1095).
// NZ: inputs n >= NZ are zero (the zero-padded window: their first-stage
// butterflies reduce to copies).  NOUT: outputs k >= NOUT are not needed
// (their last-stage halves are left uncomputed; xr/xi[k] are unspecified).
template <typename R, bool INV, int NZ = 32, int NOUT = 32>
__device__ inline void fft32(R (&xr)[32], R (&xi)[32]) {
    R ar[32], ai[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        ar[bitrev5(k)] = xr[k];
        ai[bitrev5(k)] = xi[k];
    }
#pragma unroll
    for (int s = 1; s <= 5; ++s) {
        const int m = 1 << s, half = m >> 1, step = 32 / m;
#pragma unroll
        for (int b = 0; b < 32; b += m) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                const int tw = j * step;
                R br = ar[b + j + half], bi = ai[b + j + half];
                R vr, vi;
                if constexpr (FMA_BUTTERFLY<R>) {
                    if (tw != 0 && tw != 8 && !(s == 1 && bitrev5(b + j + half) >= NZ)) {
                        // x = u + w b in four fused steps, u - w b = 2u - x in
                        // one each: 6 instructions per butterfly instead of 8
                        const R wr = twr<R>(tw);
                        const R wi = INV ? -twi_fwd<R>(tw) : twi_fwd<R>(tw);
                        const R ur = ar[b + j], ui = ai[b + j];
                        const R xr = fma(-wi, bi, fma(wr, br, ur));
                        const R xi = fma(wi, br, fma(wr, bi, ui));
                        ar[b + j] = xr;
                        ai[b + j] = xi;
                        if (s == 5 && b + j + half >= NOUT) continue;
                        ar[b + j + half] = fma((R)2, ur, -xr);
                        ai[b + j + half] = fma((R)2, ui, -xi);
                        continue;
                    }
                }
                if (tw == 0) {
                    vr = br;
                    vi = bi;
                } else if (tw == 8) {
                    // multiply by -i (forward) or +i (inverse)
                    if (INV) {
                        vr = -bi;
                        vi = br;
                    } else {
                        vr = bi;
                        vi = -br;
                    }
                } else {
                    const R wr = twr<R>(tw);
                    const R wi = INV ? -twi_fwd<R>(tw) : twi_fwd<R>(tw);
                    vr = fma(br, wr, -(bi * wi));
                    vi = fma(br, wi, bi * wr);
                }
                R ur = ar[b + j], ui = ai[b + j];
                if (s == 1 && bitrev5(b + j + half) >= NZ) {
                    // the pair's second input is a padding zero
                    ar[b + j + half] = ur;
                    ai[b + j + half] = ui;
                    continue;
                }
                ar[b + j] = ur + vr;
                ai[b + j] = ui + vi;
                if (s == 5 && b + j + half >= NOUT) continue;
                ar[b + j + half] = ur - vr;
                ai[b + j + half] = ui - vi;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) {
        xr[k] = ar[k];
        xi[k] = ai[k];
    }
}

// Sum over the 32 lanes of a half-wave team (lanes 0-31 or 32-63).
__device__ inline int team_sum_i32(int v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) v += __shfl_xor(v, o, 32);
    return v;
}

template <typename K>
__device__ inline K team_min_key(K v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) {
        K w = __shfl_xor(v, o, 32);
        v = w < v ? w : v;
    }
    return v;
}

// Team-wide (32-lane) primitives on 32-bit values.  With PS2_TEAM_DPP they use
// DPP row permutations (rows of 16 lanes) and the gfx950 cross-row swap
// v_permlane16_swap instead of LDS-routed permutes; a team is two DPP rows
// (lanes 0-31 or 32-63 of the wavefront) and every lane of the team must be
// active.
#ifndef PS2_TEAM_DPP
#define PS2_TEAM_DPP 1
#endif

#if PS2_TEAM_DPP
// DPP move within rows of 16 lanes; sources outside the row read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline uint32_t dpp_u32(uint32_t v) {
    return __builtin_amdgcn_update_dpp(0u, v, CTRL, ROW_MASK, 0xf, true);
}
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_ROW_HALF_MIRROR = 0x141,
              DPP_ROW_MIRROR = 0x140, DPP_ROW_SHR = 0x110, DPP_ROW_BCAST15 = 0x142;

template <>
__device__ inline uint32_t team_min_key<uint32_t>(uint32_t v) {
    v = min(v, dpp_u32<DPP_QUAD_XOR1>(v));
    v = min(v, dpp_u32<DPP_QUAD_XOR2>(v));
    v = min(v, dpp_u32<DPP_ROW_HALF_MIRROR>(v));
    v = min(v, dpp_u32<DPP_ROW_MIRROR>(v));
    // rows 0<->1 and 2<->3: each lane sees both rows' minima
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return min((uint32_t)r[0], (uint32_t)r[1]);
}
#endif

// inclusive prefix sum over the team's lanes
__device__ inline uint32_t team_incl_sum_u32(uint32_t v) {
#if PS2_TEAM_DPP
    v += dpp_u32<DPP_ROW_SHR + 1>(v);
    v += dpp_u32<DPP_ROW_SHR + 2>(v);
    v += dpp_u32<DPP_ROW_SHR + 4>(v);
    v += dpp_u32<DPP_ROW_SHR + 8>(v);
    v += dpp_u32<DPP_ROW_BCAST15, 0xa>(v);  // row 0 total into row 1 (row 2's into row 3)
    return v;
#else
    const int lane = threadIdx.x & 31;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const uint32_t w = __shfl_up(v, o, 32);
        if (lane >= o) v += w;
    }
    return v;
#endif
}

// v of team lane ql (team of lanes 0-31) or qh (lanes 32-63); ql and qh are
// the same on every lane of the wavefront
__device__ inline uint32_t team_pick_u32(uint32_t v, int ql, int qh, bool hi_team) {
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, ql & 31);
    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 32 + (qh & 31));
    return hi_team ? b : a;
}

// first set bit of each team's half of a wavefront ballot (-1 if none)
__device__ inline void team_first_lanes(unsigned long long m, int &ql, int &qh) {
    ql = __ffs((unsigned)m) - 1;
    qh = __ffs((unsigned)(m >> 32)) - 1;
}

}  // namespace ps2
