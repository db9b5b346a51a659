// Microbenchmark of the hand-off costs the block passes are designed
// around (cholespy_probe_handoff in the C ABI): one-way hand-off of a value
// between two workgroups through agent-scope relaxed stores and loads of a
// sentinel word (and the same with release / acquire fences), a dependent
// chain of agent-scope loads, a 256-thread workgroup barrier, a ticket
// atomic on one address under contention, and back-to-back launches of an
// empty kernel. Every wait is bounded.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "launchers.hpp"

namespace cholespy_amd {

namespace {

constexpr uint32_t kProbeLimit = 1u << 24;
constexpr int kRounds = 2000;                     // ping-pong rounds
constexpr size_t kSideWords = (size_t)kRounds * 16;   // one 128-byte line per round and side
constexpr int kChain = 2000;

__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup 0 publishes round i on side a, workgroup 1 answers on side b.
template <bool FENCED>
__global__ void ping_pong(uint64_t *words, unsigned long long *out, int *gave_up) {
    if (threadIdx.x != 0) return;
    const int me = blockIdx.x;
    uint64_t *a = words, *b = words + kSideWords;
    const unsigned long long t0 = wall_clock64();
    for (int i = 0; i < kRounds; ++i) {
        uint64_t *mine = (me == 0 ? a : b) + (size_t)i * 16;
        const uint64_t *theirs = (me == 0 ? b : a) + (size_t)i * 16;
        if (me == 0) {
            if (FENCED) __atomic_thread_fence(__ATOMIC_RELEASE);
            st_agent(mine, 1 + (uint64_t)i);
        }
        uint32_t spins = 0;
        while (ld_agent(theirs) == 0) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kProbeLimit) {
                *gave_up = 1;
                return;
            }
        }
        if (FENCED) __atomic_thread_fence(__ATOMIC_ACQUIRE);
        if (me == 1) {
            if (FENCED) __atomic_thread_fence(__ATOMIC_RELEASE);
            st_agent(mine, 1 + (uint64_t)i);
        }
    }
    if (me == 0) out[0] = wall_clock64() - t0;
}

__global__ void load_chain(const uint64_t *chain, unsigned long long *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t cur = 0;
    const unsigned long long t0 = wall_clock64();
    for (int i = 0; i < kChain; ++i) cur = ld_agent(chain + cur * 16);
    out[1] = wall_clock64() - t0;
    out[2] = cur;   // keeps the chain live
}

__global__ void barrier_loop(int reps, unsigned long long *out) {
    const unsigned long long t0 = wall_clock64();
    for (int i = 0; i < reps; ++i) __syncthreads();
    if (threadIdx.x == 0) out[3] = wall_clock64() - t0;
}

__global__ void ticket_storm(unsigned int *ctr, int per_wg, unsigned long long *out) {
    if (threadIdx.x != 0) return;
    const unsigned long long t0 = wall_clock64();
    unsigned int sum = 0;
    for (int i = 0; i < per_wg; ++i)
        sum += __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    atomicMax(&out[4], wall_clock64() - t0);
    if (sum == 0xffffffffu) out[5] = sum;
}

__global__ void empty_kernel() {}

}  // namespace

hipError_t probe_handoff(double *res) {
    const int reps = 20000, wgs = 1024, per_wg = 16, launches = 1000;
    uint64_t *words = nullptr, *chain = nullptr;
    unsigned long long *out = nullptr;
    unsigned int *ctr = nullptr;
    int *gave_up = nullptr;
    hipError_t e = hipMalloc(&words, 2 * kSideWords * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&chain, (size_t)kChain * 16 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&out, 8 * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&ctr, sizeof(unsigned int));
    if (e == hipSuccess) e = hipMalloc(&gave_up, sizeof(int));
    std::vector<uint64_t> hc((size_t)kChain * 16, 0);
    for (int i = 0; i < kChain; ++i) hc[(size_t)i * 16] = (uint64_t)(((size_t)i * 7919 + 1) % kChain);
    if (e == hipSuccess) e = hipMemcpy(chain, hc.data(), hc.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    int rate = 0;
    double freq = 100e6;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, 0) == hipSuccess && rate > 0) freq = rate * 1e3;
    unsigned long long h[8] = {0};
    int failed = 0;
    for (int fenced = 0; fenced < 2 && e == hipSuccess; ++fenced) {
        double best = 1e30;
        for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
            e = hipMemset(words, 0, 2 * kSideWords * sizeof(uint64_t));
            if (e == hipSuccess) e = hipMemset(gave_up, 0, sizeof(int));
            if (e != hipSuccess) break;
            if (fenced) ping_pong<true><<<2, 64>>>(words, out, gave_up);
            else ping_pong<false><<<2, 64>>>(words, out, gave_up);
            e = hipDeviceSynchronize();
            if (e == hipSuccess) e = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(&failed, gave_up, sizeof failed, hipMemcpyDeviceToHost);
            if (e == hipSuccess && !failed) best = std::min(best, h[0] / freq * 1e6 / (2.0 * kRounds));
        }
        res[fenced] = failed ? -1.0 : best;
    }
    if (e == hipSuccess) {
        load_chain<<<1, 64>>>(chain, out);
        e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
        res[2] = h[1] / freq * 1e9 / kChain;
    }
    if (e == hipSuccess) {
        barrier_loop<<<1, 256>>>(reps, out);
        e = hipDeviceSynchronize();
        if (e == hipSuccess) e = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
        res[3] = h[3] / freq * 1e9 / reps;
    }
    if (e == hipSuccess) {
        e = hipMemset(ctr, 0, sizeof(unsigned int));
        if (e == hipSuccess) e = hipMemset(out, 0, 8 * sizeof(unsigned long long));
        if (e == hipSuccess) {
            ticket_storm<<<wgs, 64>>>(ctr, per_wg, out);
            e = hipDeviceSynchronize();
        }
        if (e == hipSuccess) e = hipMemcpy(h, out, sizeof h, hipMemcpyDeviceToHost);
        res[4] = h[4] / freq * 1e9 / ((double)wgs * per_wg);
    }
    if (e == hipSuccess) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        for (int i = 0; i < 100 && e == hipSuccess; ++i) empty_kernel<<<1, 64>>>();
        if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
        for (int i = 0; i < launches && e == hipSuccess; ++i) empty_kernel<<<1, 64>>>();
        if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        res[5] = ms * 1e3 / launches;
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
    (void)hipFree(words);
    (void)hipFree(chain);
    (void)hipFree(out);
    (void)hipFree(ctr);
    (void)hipFree(gave_up);
    return e;
}

}  // namespace cholespy_amd
