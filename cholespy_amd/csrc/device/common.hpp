// Shared device helpers for the gfx950 solve-phase kernels.
//
// Cross-workgroup hand-off. The reference publishes a solved row through a
// separate `volatile bool` flag after a __threadfence (kernels/kernels.cu:
// 106-111, 122-130). Here every published value is its own flag: buffers
// that rows wait on are pre-filled with a signalling-NaN sentinel, a
// producer writes each value with one agent-scope relaxed atomic store, and
// a consumer spins with agent-scope relaxed atomic loads until the word is
// no longer the sentinel. Single-copy atomicity of aligned 32/64-bit
// accesses makes a non-sentinel word the final value, so no fences (and no
// L2 write-back / invalidate) are needed on gfx950; only the `sc1` scope bit
// on the accesses themselves. Arithmetic never produces a signalling NaN,
// so a computed result can never be mistaken for the sentinel.
//
// Bounded waits. Every spin loop counts its polls; past a limit it records
// CHOLESPY_ERR_TIMEOUT in the status word, raises the abort word and stops
// waiting, and every other spinning lane sees the abort word and stops too,
// so a broken schedule ends the kernel instead of hanging the device. A wait
// that gives up returns a quiet NaN, so the results it feeds are NaN, never
// finite and wrong. The two status words live in host memory mapped into
// the device (fine-grained, coherent): kernels only store to them, with
// system-scope stores on the give-up path, and the host reads them with no
// device call. Long waits read a device-memory copy of the abort word every
// 256 polls (a host-memory read from thousands of waiting wavefronts would
// load the PCIe link the whole time).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cholespy_amd {

constexpr int kWave = 64;
constexpr int kMaxRhs = 128;
constexpr int kStatusTimeout = 4;   // matches CHOLESPY_ERR_TIMEOUT

template <typename F> struct Bits;
template <> struct Bits<float> {
    using U = uint32_t;
    static constexpr U sentinel = 0x7FA5A5A5u;   // signalling NaN
};
template <> struct Bits<double> {
    using U = uint64_t;
    static constexpr U sentinel = 0x7FF4A5A5A5A5A5A5ull;   // signalling NaN
};

template <typename F>
__device__ __forceinline__ typename Bits<F>::U as_bits(F v) {
    return __builtin_bit_cast(typename Bits<F>::U, v);
}
template <typename F>
__device__ __forceinline__ F from_bits(typename Bits<F>::U u) {
    return __builtin_bit_cast(F, u);
}

// Global-memory view of a pointer: accesses through it compile to global
// (vector-memory) instructions. Where the compiler cannot prove a pointer's
// address space, e.g. one picked from several arrays, it emits a flat access,
// which also counts against the LDS counter; the wait for it then drains
// every other access in flight. Every buffer the hand-off helpers below
// touch is device (or mapped host) memory.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
    return (__attribute__((address_space(1))) T *)p;
}

// Agent-scope relaxed store of one value (its own publication).
template <typename F>
__device__ __forceinline__ void publish(F *p, F v) {
    using U = typename Bits<F>::U;
    __hip_atomic_store(gmem(reinterpret_cast<U *>(p)), as_bits(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

template <typename F>
__device__ __forceinline__ typename Bits<F>::U load_bits(const F *p) {
    using U = typename Bits<F>::U;
    return __hip_atomic_load(gmem(reinterpret_cast<const U *>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}

// First read of a published entry, through the caches (a relaxed atomic at
// workgroup scope is a plain load). Every published location holds the
// sentinel or its final value and changes at most once per launch (launch
// boundaries write back and invalidate the caches), so a stale cached copy
// can only be the sentinel: callers pass it to ready_value, which then
// waits at agent scope. Values read by many workgroups of one XCD (a large
// block's r_B, a separator's z) then come from that XCD's L2.
template <typename F>
__device__ __forceinline__ typename Bits<F>::U peek_bits(const F *p) {
    using U = typename Bits<F>::U;
    return __hip_atomic_load(gmem(reinterpret_cast<const U *>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ int load_int(const int *p) {
    return __hip_atomic_load(gmem(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_int(int *p, int v) {
    __hip_atomic_store(gmem(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Dynamic work queues (below) and the watchdog's launch arguments.
struct QueueArgs {
    unsigned int *ctr;    // nq counters kQueueStride apart, then the done counter
    int nq;
    uint32_t spin_limit;  // poll budget of every bounded wait of the launch
    int *abort;           // device copy of the abort word (read by long waits)
};

// status[0] = error code, status[1] = abort flag, in mapped host memory.
// Plain system-scope stores only (no read-modify-write on host memory):
// every writer stores the same code. The device abort word is raised too.
__device__ __forceinline__ void raise_timeout(int *status, const QueueArgs &qa) {
    __hip_atomic_store(status, kStatusTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(status + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(qa.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Default poll budget of one wait (about 10 s of polling); a handle may set a
// smaller one (cholespy_solver_set_spin_limit) to force the give-up path.
constexpr uint32_t kSpinLimit = 1u << 23;
#ifndef CHOLESPY_SPIN_SLEEP
#define CHOLESPY_SPIN_SLEEP 1
#endif

template <typename F>
__device__ __forceinline__ F give_up_value() {
    return __builtin_nan("");
}

// Waits until *p is no longer the sentinel and returns the value. After
// `limit` polls it records the timeout and returns a quiet NaN; it also
// returns a quiet NaN when it sees the abort word (checked every 256 polls).
template <typename F>
__device__ __forceinline__ F wait_value(const F *p, int *status, const QueueArgs &qa) {
    using U = typename Bits<F>::U;
    U u = load_bits(p);
    if (u != Bits<F>::sentinel) return from_bits<F>(u);
    uint32_t spins = 0;
    while (true) {
        __builtin_amdgcn_s_sleep(CHOLESPY_SPIN_SLEEP);
        u = load_bits(p);
        if (u != Bits<F>::sentinel) return from_bits<F>(u);
        if (++spins >= qa.spin_limit) {
            raise_timeout(status, qa);
            return give_up_value<F>();
        }
        if ((spins & 255u) == 0 && load_int(qa.abort)) return give_up_value<F>();
    }
}

// Waits until *p is non-zero and returns it (same watchdog as wait_value;
// returns 1 when it gives up, and the status word says so).
__device__ __forceinline__ int wait_nonzero(const int *p, int *status, const QueueArgs &qa) {
    int v = load_int(p);
    if (v != 0) return v;
    uint32_t spins = 0;
    while (true) {
        __builtin_amdgcn_s_sleep(1);
        v = load_int(p);
        if (v != 0) return v;
        if (++spins >= qa.spin_limit) {
            raise_timeout(status, qa);
            return 1;
        }
        if ((spins & 255u) == 0 && load_int(qa.abort)) return 1;
    }
}

// Dynamic work queues. Tickets (rows, row chunks or block tasks, numbered in
// dependency order) are taken in increasing order, so a ticket is only ever
// held by a running worker and every wait targets a ticket that is running
// or finished: the schedule needs no co-residency guarantee (the occupancy
// reported for cooperative launches was measured to over-state residency on
// MI355X). To avoid serialising on one atomic address, tickets are spread
// over nq queues: ticket t belongs to queue t mod nq, worker w serves queue
// w mod nq (the first nq workers are dispatched first, so every queue has a
// running worker). Each launch starts from zeroed counters and leaves them
// zeroed: the last workgroup to finish resets them, so launches need no
// host-side state and can be replayed from a captured graph.
constexpr int kMaxQueues = 16;
constexpr int kMaxLeafQueues = 64;   // queues of a fused launch's leaf phase
constexpr int kQueueStride = 16;   // 128 bytes between counters

// Next ticket of queue q, or -1 when the queue is exhausted (n tickets in all).
__device__ __forceinline__ int take_ticket(const QueueArgs &qa, int q, int n) {
    const unsigned int j =
        __hip_atomic_fetch_add(qa.ctr + (size_t)q * kQueueStride, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const long long t = (long long)j * qa.nq + q;
    return t < n ? (int)t : -1;
}

// Called once per workgroup by one thread after its last ticket: the last
// workgroup of the grid zeroes the counters for the next launch. Relaxed is
// enough: a workgroup gets here only after its last ticket's fetch-add has
// returned, so when the last one sees the full count no fetch-add on the
// counters is pending, and the launch boundary orders the zeroing stores
// before the next launch. (An acquire-release here would write back and
// invalidate the workgroup's L2 at every exit, under the tasks still
// running.)
// `extra` (extra_nq > 0): a second set of extra_nq queues the launch used
// (a fused launch's leaf queues, counters kQueueStride apart), zeroed with
// them; every wavefront of the workgroup must be done taking tickets.
__device__ __forceinline__ void finish_queues(const QueueArgs &qa, unsigned int *extra = nullptr,
                                              int extra_nq = 0) {
    unsigned int *done = qa.ctr + (size_t)kMaxQueues * kQueueStride;
    const unsigned int d = __hip_atomic_fetch_add(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (d + 1 == gridDim.x * gridDim.y * gridDim.z) {
        for (int q = 0; q < qa.nq; ++q)
            __hip_atomic_store(qa.ctr + (size_t)q * kQueueStride, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int q = 0; q < extra_nq; ++q)
            __hip_atomic_store(extra + (size_t)q * kQueueStride, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, kWave));
    return v;
}

}  // namespace cholespy_amd
