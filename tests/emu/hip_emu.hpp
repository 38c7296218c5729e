// Host-thread emulation of HIP workgroups, for running the product's kernels
// (dgcom_amd/csrc/swe_kernels.hpp) on a CPU in tests: one std::thread per lane,
// a std::barrier for __syncthreads, function-static storage for __shared__
// (workgroups run one after another), plain memory for __constant__.
#pragma once

#define DGSWE_EMULATE 1

#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

struct double4 { double x, y, z, w; };
struct EmuDim3 {
    unsigned x = 1, y = 1, z = 1;
};
inline thread_local EmuDim3 threadIdx;
inline thread_local EmuDim3 blockIdx;
inline EmuDim3 blockDim;
inline EmuDim3 gridDim;
inline std::barrier<>* emu_barrier = nullptr;

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __constant__
#define __launch_bounds__(x)

inline void __syncthreads() { emu_barrier->arrive_and_wait(); }

inline long long __double_as_longlong(double x) {
    long long r;
    std::memcpy(&r, &x, 8);
    return r;
}
inline double __longlong_as_double(long long x) {
    double r;
    std::memcpy(&r, &x, 8);
    return r;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
    return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned int atomicAdd(unsigned int* p, unsigned int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
    unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline int atomicMax(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
    unsigned long long old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}

// Run `kernel` over `grid` workgroups of `block` lanes, one workgroup at a time.
template <class F>
inline void emu_launch(unsigned grid, unsigned block, F kernel) {
    gridDim.x = grid;
    blockDim.x = block;
    std::barrier<> bar(block);
    emu_barrier = &bar;
    for (unsigned b = 0; b < grid; ++b) {
        std::vector<std::thread> lanes;
        lanes.reserve(block);
        for (unsigned t = 0; t < block; ++t)
            lanes.emplace_back([&, b, t] {
                threadIdx.x = t;
                blockIdx.x = b;
                kernel();
            });
        for (auto& l : lanes) l.join();
    }
    emu_barrier = nullptr;
}
