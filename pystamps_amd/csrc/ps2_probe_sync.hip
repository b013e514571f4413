// Measurement behind the batching design: what one synchronisation point of
// a gamma iteration costs on the device.
//   (1) a kernel boundary inside a captured HIP graph (N tiny kernels in one
//       graph, per kernel);
//   (2) a device-wide barrier inside one persistent launch: every wave
//       releases its writes at agent scope, one arrival per workgroup on an
//       atomic counter, every workgroup polls it (flat), or arrivals counted
//       per XCD first with one arrival per XCD on a top counter and a release
//       flag per XCD that the XCD's workgroups poll (two-level).
// The persistent form of the loop would replace each kernel boundary of an
// iteration by one such barrier.  Every spin is bounded: an exhausted spin
// records an error and its workgroup leaves, so the program always ends.
//
//   ps2_probe_sync            (built by pystamps_amd.build.build_probe)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CK(x)                                                                                      \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);             \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

namespace {

constexpr uint32_t kSpinMax = 1u << 22;   // polls per wait (seconds), then an error
constexpr int kLine = 32;                  // counters 128 bytes apart
constexpr int kXcd = 8;                    // workgroup b runs on XCD b % 8

__device__ bool wait_ge(const uint32_t *p, uint32_t target, uint32_t *err) {
    for (uint32_t s = 0; s < kSpinMax; ++s) {
        if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
        __builtin_amdgcn_s_sleep(1);
    }
    atomicAdd(err, 1u);
    return false;
}

// a little work per phase: each thread updates one float of its workgroup's slice
__device__ void phase_work(float *buf, int per_wg) {
    for (int q = threadIdx.x; q < per_wg; q += blockDim.x) buf[(size_t)blockIdx.x * per_wg + q] += 1.0f;
}

// nb phases separated by flat barriers; cnt[b * kLine] zeroed by the host
__global__ __launch_bounds__(256) void k_flat(uint32_t *cnt, int nb, uint32_t *err, float *buf, int per_wg) {
    __shared__ int ok;
    for (int b = 0; b < nb; ++b) {
        phase_work(buf, per_wg);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            __hip_atomic_fetch_add(&cnt[b * kLine], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = wait_ge(&cnt[b * kLine], gridDim.x, err) ? 1 : 0;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (!ok) return;
    }
}

// two-level: xc[(b * kXcd + x) * kLine] per-XCD arrivals, top[b * kLine] XCD
// arrivals, flag[(b * kXcd + x) * kLine] release of XCD x; gridDim.x a
// multiple of 8 (every XCD has workgroups)
__global__ __launch_bounds__(256) void k_two_level(uint32_t *xc, uint32_t *top, uint32_t *flag, int nb,
                                                   uint32_t *err, float *buf, int per_wg) {
    __shared__ int ok;
    const uint32_t x = blockIdx.x % kXcd;
    const uint32_t nx = gridDim.x / kXcd;
    for (int b = 0; b < nb; ++b) {
        phase_work(buf, per_wg);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t old =
                __hip_atomic_fetch_add(&xc[(b * kXcd + x) * kLine], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old == nx - 1) {   // last of its XCD
                const uint32_t o2 =
                    __hip_atomic_fetch_add(&top[b * kLine], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (o2 == kXcd - 1)   // last XCD: release all
                    for (int y = 0; y < kXcd; ++y)
                        __hip_atomic_store(&flag[(b * kXcd + y) * kLine], 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
            ok = wait_ge(&flag[(b * kXcd + x) * kLine], 1u, err) ? 1 : 0;
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (!ok) return;
    }
}

__global__ __launch_bounds__(256) void k_tiny(float *buf, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) buf[i] += 1.0f;
}

}  // namespace

int main() {
    hipStream_t st;
    CK(hipStreamCreate(&st));
    constexpr int NB = 200;                  // barriers per launch
    constexpr int kMaxG = 512, kMaxPer = 4096;
    uint32_t *cnt, *xc, *top, *flag, *err;
    float *buf;
    CK(hipMalloc(&cnt, sizeof(uint32_t) * NB * kLine));
    CK(hipMalloc(&xc, sizeof(uint32_t) * NB * kXcd * kLine));
    CK(hipMalloc(&top, sizeof(uint32_t) * NB * kLine));
    CK(hipMalloc(&flag, sizeof(uint32_t) * NB * kXcd * kLine));
    CK(hipMalloc(&err, sizeof(uint32_t)));
    CK(hipMalloc(&buf, sizeof(float) * kMaxG * kMaxPer));   // k_tiny uses at most 1024 * 256 of it
    CK(hipMemset(buf, 0, sizeof(float) * kMaxG * kMaxPer));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    int dev;
    CK(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    CK(hipGetDeviceProperties(&pr, dev));
    int occ = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_flat, 256, 0));
    std::printf("device %s, %d CUs, %d resident barrier workgroups per CU\n", pr.gcnArchName, pr.multiProcessorCount,
                occ);
    if (occ * pr.multiProcessorCount < kMaxG) {
        std::printf("fewer than %d workgroups resident: barrier runs skipped\n", kMaxG);
        return 1;
    }
    std::fflush(stdout);
    // (1) kernel boundaries in a graph
    for (int nk : {1, 50}) {
        for (int blocks : {80, 1024}) {
            hipGraph_t g;
            hipGraphExec_t ge;
            CK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < nk; ++i) k_tiny<<<blocks, 256, 0, st>>>(buf, blocks * 256);
            CK(hipStreamEndCapture(st, &g));
            CK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            CK(hipGraphLaunch(ge, st));
            CK(hipStreamSynchronize(st));
            CK(hipEventRecord(e0, st));
            for (int r = 0; r < 20; ++r) CK(hipGraphLaunch(ge, st));
            CK(hipEventRecord(e1, st));
            CK(hipEventSynchronize(e1));
            float ms;
            CK(hipEventElapsedTime(&ms, e0, e1));
            std::printf("graph of %2d tiny kernels x %4d workgroups: %7.2f us per graph, %6.2f us per kernel\n", nk,
                        blocks, ms * 1e3 / 20, ms * 1e3 / 20 / nk);
            std::fflush(stdout);
            CK(hipGraphExecDestroy(ge));
            CK(hipGraphDestroy(g));
        }
    }
    // (2) barriers
    for (int two : {0, 1}) {
        for (int G : {80, 256, 512}) {
            const int per_wg = 256;
            float best = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                CK(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * NB * kLine, st));
                CK(hipMemsetAsync(xc, 0, sizeof(uint32_t) * NB * kXcd * kLine, st));
                CK(hipMemsetAsync(top, 0, sizeof(uint32_t) * NB * kLine, st));
                CK(hipMemsetAsync(flag, 0, sizeof(uint32_t) * NB * kXcd * kLine, st));
                CK(hipMemsetAsync(err, 0, sizeof(uint32_t), st));
                CK(hipEventRecord(e0, st));
                if (two)
                    k_two_level<<<G, 256, 0, st>>>(xc, top, flag, NB, err, buf, per_wg);
                else
                    k_flat<<<G, 256, 0, st>>>(cnt, NB, err, buf, per_wg);
                CK(hipEventRecord(e1, st));
                CK(hipEventSynchronize(e1));
                CK(hipGetLastError());
                float ms;
                CK(hipEventElapsedTime(&ms, e0, e1));
                uint32_t he = 0;
                CK(hipMemcpy(&he, err, sizeof(uint32_t), hipMemcpyDeviceToHost));
                if (he) {
                    std::printf("%s barrier, %d workgroups: %u spins exhausted\n", two ? "two-level" : "flat", G, he);
                    return 2;
                }
                if (ms < best) best = ms;
            }
            std::printf("%-9s barrier, %3d workgroups (256 floats each per phase): %6.2f us per barrier\n",
                        two ? "two-level" : "flat", G, best * 1e3 / NB);
            std::fflush(stdout);
        }
    }
    std::printf("done\n");
    return 0;
}
