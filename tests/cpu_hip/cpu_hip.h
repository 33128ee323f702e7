// Host build of the library's film kernels, for CPU tests only: the HIP features that
// film.hip uses, emulated.  Blocks run one after another; the threads of a block run as
// std::threads that meet at a barrier for __syncthreads; __shared__ arrays become static
// (one block at a time uses them); wave shuffles go through a block-wide exchange buffer.
#pragma once
#include <algorithm>
#include <barrier>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>
using std::max;
using std::min;
struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct cpu_uint3 { unsigned x, y, z; };
inline thread_local cpu_uint3 threadIdx, blockIdx;
inline dim3 blockDim, gridDim;
inline std::barrier<> *cpu_bar = nullptr;
inline unsigned char *cpu_dyn_lds = nullptr;
inline std::mutex cpu_atomic_mu;
struct uint4 { unsigned x, y, z, w; };
typedef int hipError_t;
typedef void *hipStream_t;
#define hipSuccess 0
#define hipErrorInvalidValue 1
inline hipError_t hipGetLastError() { return 0; }
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
inline void __syncthreads() { cpu_bar->arrive_and_wait(); }
template <class T> T __shfl_xor(T v, int mask, int)
{
    static T buf[1024];
    const unsigned t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    const T r = buf[t ^ (unsigned)mask];
    __syncthreads();
    return r;
}
inline int atomicMin(int *p, int v) { std::lock_guard<std::mutex> l(cpu_atomic_mu); const int o = *p; *p = min(o, v); return o; }
inline int atomicMax(int *p, int v) { std::lock_guard<std::mutex> l(cpu_atomic_mu); const int o = *p; *p = max(o, v); return o; }
template <class K, class... A> void cpu_launch(K k, dim3 g, dim3 b, size_t lds, A... a)
{
    gridDim = g;
    blockDim = b;
    std::vector<unsigned char> dyn(lds + 16);
    cpu_dyn_lds = dyn.data();
    for (unsigned bz = 0; bz < g.z; bz++)
        for (unsigned by = 0; by < g.y; by++)
            for (unsigned bx = 0; bx < g.x; bx++) {
                std::barrier<> bar(b.x);
                cpu_bar = &bar;
                std::vector<std::thread> th;
                for (unsigned t = 0; t < b.x; t++)
                    th.emplace_back([=]() { threadIdx = {t, 0, 0}; blockIdx = {bx, by, bz}; k(a...); });
                for (auto &x : th) x.join();
            }
}
#define hipLaunchKernelGGL(k, g, b, lds, s, ...) cpu_launch(k, dim3(g), dim3(b), lds, __VA_ARGS__)
