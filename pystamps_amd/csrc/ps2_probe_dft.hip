// Measurement behind the CLAP transform design: the forward 2D transform of
// a CLAP window (24 x 24 cells zero-padded to 32 x 32, complex FP32) as
//   (a) 32-point DFT-matrix products on the FP32 matrix cores
//       (v_mfma_f32_32x32x2f32: one wavefront per window, rows then columns,
//       4 real products per complex product, the zero padding left out of
//       the inner dimension: 96 MFMA instructions per window), and
//   (b) the library's radix-2 FFT in registers (fft32, one 32-lane team per
//       window, rows, an LDS transpose, columns), as in k_clap_strip.
// Both read the same windows, reduce each spectrum to a checksum (so the
// timing is the transform, not a 8 KB store per window), and write the full
// spectrum of the first windows, which the host checks against an f64 DFT.
//
//   ps2_probe_dft            (built by pystamps_amd.build.build_probe)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "ps2_device.hpp"

#define CK(x)                                                                                      \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);             \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

namespace {

constexpr int W = 24, N32 = 32;
constexpr int NCHECK = 8;   // windows whose full spectrum is written and checked

typedef float f32x16 __attribute__((ext_vector_type(16)));

// e^{-2 pi i m / 32}, m = 0..31
__constant__ float2 kTw[32];

// (a) one wavefront per window, 4 windows per 256-thread workgroup.  Operand
// layouts of v_mfma_f32_32x32x2f32: lane l gives A[l % 32][l / 32] and
// B[l / 32][l % 32]; the result D[i][j] is in register r of lane l with
// j = l % 32, i = 8 (r / 4) + 4 (l / 32) + r % 4.  Row pass Y = X F: step m,
// half h takes column c = 2m + h (c < 24).  Column pass Z = F Y: step r, half
// h takes row i = 8 (r / 4) + 4 h + r % 4, which is where register r of the
// row pass holds Y, so Y feeds the column pass from its accumulator registers
// as they are; rows 24..31 of Y are zero (registers 12..15), left out.
__global__ __launch_bounds__(256) void k_dft_mfma(const float2 *__restrict__ win, int n, float *__restrict__ chk,
                                                  float2 *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int w = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (w >= n) return;   // (wavefront-uniform)
    const int h = lane >> 5, q = lane & 31;
    const float2 *x = win + (size_t)w * W * W;
    f32x16 yr = {}, yi = {};
#pragma unroll
    for (int m = 0; m < W / 2; ++m) {
        const int c = 2 * m + h;
        float xr = 0.0f, xi = 0.0f;
        if (q < W) {
            const float2 v = x[q * W + c];
            xr = v.x;
            xi = v.y;
        }
        const float2 f = kTw[(c * q) & 31];
        yr = __builtin_amdgcn_mfma_f32_32x32x2f32(xr, f.x, yr, 0, 0, 0);
        yr = __builtin_amdgcn_mfma_f32_32x32x2f32(-xi, f.y, yr, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_32x32x2f32(xr, f.y, yi, 0, 0, 0);
        yi = __builtin_amdgcn_mfma_f32_32x32x2f32(xi, f.x, yi, 0, 0, 0);
    }
    f32x16 zr = {}, zi = {};
#pragma unroll
    for (int r = 0; r < 12; ++r) {
        const int i = 8 * (r / 4) + 4 * h + (r % 4);
        const float2 f = kTw[(q * i) & 31];
        zr = __builtin_amdgcn_mfma_f32_32x32x2f32(f.x, yr[r], zr, 0, 0, 0);
        zr = __builtin_amdgcn_mfma_f32_32x32x2f32(-f.y, yi[r], zr, 0, 0, 0);
        zi = __builtin_amdgcn_mfma_f32_32x32x2f32(f.x, yi[r], zi, 0, 0, 0);
        zi = __builtin_amdgcn_mfma_f32_32x32x2f32(f.y, yr[r], zi, 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) s += zr[r] * zr[r] + zi[r] * zi[r];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) chk[w] = s;
    if (w < NCHECK) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int k = 8 * (r / 4) + 4 * h + (r % 4);
            out[(size_t)w * N32 * N32 + k * N32 + q] = make_float2(zr[r], zi[r]);
        }
    }
}

__device__ inline void wave_fence_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (b) one 32-lane team per window, 2 windows per 64-thread workgroup: lane =
// row, fft32 over the row's 24 cells, LDS transpose (separate re / im planes,
// rows of 33), lane = column, fft32 over the 24 non-zero rows.
__global__ __launch_bounds__(64) void k_dft_fft(const float2 *__restrict__ win, int n, float *__restrict__ chk,
                                                float2 *__restrict__ out) {
    const int team = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const int w = (int)blockIdx.x * 2 + team;
    const bool valid = w < n;
    __shared__ float tr[2][N32][N32 + 1], ti[2][N32][N32 + 1];
    float xr[32], xi[32];
#pragma unroll
    for (int c = 0; c < 32; ++c) xr[c] = xi[c] = 0.0f;
    if (valid && lane < W) {
        const float2 *x = win + (size_t)w * W * W + lane * W;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            const float2 v = x[c];
            xr[c] = v.x;
            xi[c] = v.y;
        }
    }
    ps2::fft32<float, false, W>(xr, xi);
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        tr[team][lane][c] = xr[c];
        ti[team][lane][c] = xi[c];
    }
    wave_fence_sync();
#pragma unroll
    for (int r = 0; r < W; ++r) {
        xr[r] = tr[team][r][lane];
        xi[r] = ti[team][r][lane];
    }
#pragma unroll
    for (int r = W; r < 32; ++r) xr[r] = xi[r] = 0.0f;
    ps2::fft32<float, false, W>(xr, xi);   // lane = column j, xr[k] = Z[k][j]
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 32; ++k) s += xr[k] * xr[k] + xi[k] * xi[k];
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) s += __shfl_xor(s, o, 32);
    if (valid && lane == 0) chk[w] = s;
    if (valid && w < NCHECK) {
#pragma unroll
        for (int k = 0; k < 32; ++k) out[(size_t)w * N32 * N32 + k * N32 + lane] = make_float2(xr[k], xi[k]);
    }
}

}  // namespace

int main() {
    const int n = 1 << 17;   // windows per launch (C2 transforms 132,540 per iteration)
    std::vector<float2> hw((size_t)n * W * W);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    for (auto &v : hw) {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const double a = (double)(st >> 11) * (6.283185307179586 / 9007199254740992.0);
        v = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2 tw[32];
    for (int m = 0; m < 32; ++m)
        tw[m] = make_float2((float)std::cos(2.0 * M_PI * m / 32.0), (float)-std::sin(2.0 * M_PI * m / 32.0));
    CK(hipMemcpyToSymbol(HIP_SYMBOL(kTw), tw, sizeof tw));
    float2 *dwin, *dout;
    float *dchk;
    CK(hipMalloc(&dwin, sizeof(float2) * hw.size()));
    CK(hipMalloc(&dout, sizeof(float2) * NCHECK * N32 * N32));
    CK(hipMalloc(&dchk, sizeof(float) * n));
    CK(hipMemcpy(dwin, hw.data(), sizeof(float2) * hw.size(), hipMemcpyHostToDevice));
    // f64 reference spectra of the checked windows
    std::vector<double> ref((size_t)NCHECK * N32 * N32 * 2);
    for (int w = 0; w < NCHECK; ++w)
        for (int k = 0; k < N32; ++k)
            for (int j = 0; j < N32; ++j) {
                double sr = 0.0, si = 0.0;
                for (int r = 0; r < W; ++r)
                    for (int c = 0; c < W; ++c) {
                        const float2 v = hw[(size_t)w * W * W + r * W + c];
                        const double a = -2.0 * M_PI * (double)((k * r + j * c) % 32) / 32.0;
                        sr += v.x * std::cos(a) - v.y * std::sin(a);
                        si += v.x * std::sin(a) + v.y * std::cos(a);
                    }
                ref[((size_t)w * N32 * N32 + k * N32 + j) * 2] = sr;
                ref[((size_t)w * N32 * N32 + k * N32 + j) * 2 + 1] = si;
            }
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const char *names[2] = {"mfma DFT-matrix products (fp32 32x32x2)", "radix-2 fft32 in registers + LDS transpose"};
    std::vector<float> chk0;
    for (int v = 0; v < 2; ++v) {
        auto launch = [&]() {
            if (v == 0)
                k_dft_mfma<<<(n + 3) / 4, 256>>>(dwin, n, dchk, dout);
            else
                k_dft_fft<<<(n + 1) / 2, 64>>>(dwin, n, dchk, dout);
        };
        launch();
        CK(hipDeviceSynchronize());
        CK(hipGetLastError());
        std::vector<float2> got((size_t)NCHECK * N32 * N32);
        CK(hipMemcpy(got.data(), dout, sizeof(float2) * got.size(), hipMemcpyDeviceToHost));
        double err = 0.0, mag = 0.0;
        for (size_t e = 0; e < got.size(); ++e) {
            err = std::fmax(err, std::fabs(got[e].x - ref[2 * e]) + std::fabs(got[e].y - ref[2 * e + 1]));
            mag = std::fmax(mag, std::fabs(ref[2 * e]) + std::fabs(ref[2 * e + 1]));
        }
        std::vector<float> chk(n);
        CK(hipMemcpy(chk.data(), dchk, sizeof(float) * n, hipMemcpyDeviceToHost));
        double cdiff = 0.0;
        if (v == 0)
            chk0 = chk;
        else
            for (int w = 0; w < n; ++w) cdiff = std::fmax(cdiff, std::fabs((double)chk[w] - chk0[w]) / chk0[w]);
        const int reps = 20;
        CK(hipEventRecord(e0));
        for (int r = 0; r < reps; ++r) launch();
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        float ms;
        CK(hipEventElapsedTime(&ms, e0, e1));
        const double us = ms * 1e3 / reps;
        std::printf("%-45s: %8.1f us per launch of %d windows, %6.1f windows/us; max |err| %.2e of max |X| %.1f",
                    names[v], us, n, n / us, err, mag);
        if (v == 1) std::printf("; checksums vs mfma: max rel diff %.1e", cdiff);
        std::printf("\n");
        std::fflush(stdout);
    }
    std::printf("done\n");
    return 0;
}
