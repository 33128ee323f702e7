// Host build of the library's film kernels and of film_host.cpp, for CPU tests only: the
// HIP features they use, emulated.  Blocks run one after another; the threads of a block
// run as std::threads that meet at a barrier for __syncthreads; __shared__ arrays become
// static (one block at a time uses them); wave shuffles go through a block-wide exchange
// buffer; dynamic LDS is a heap block of exactly the launch's size.  The runtime calls
// act at once on host memory: "device" buffers are malloc'd at their exact size (so a
// sanitizer sees every overrun), copies are memcpy, streams and events are no-ops.
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
struct float4 { float x, y, z, w; };
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
typedef int hipError_t;
typedef void *hipStream_t;
typedef void *hipEvent_t;
#define hipSuccess 0
#define hipErrorInvalidValue 1
#define hipErrorOutOfMemory 2
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
inline hipError_t hipGetLastError() { return 0; }
inline const char *hipGetErrorString(hipError_t e) { return e ? "cpu_hip error" : "no error"; }
inline hipError_t hipSetDevice(int) { return 0; }
inline hipError_t hipMalloc(void **p, size_t n)
{
    *p = malloc(n ? n : 1);
    return *p ? 0 : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p)
{
    free(p);
    return 0;
}
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    memcpy(d, s, n);
    return 0;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t)
{
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
    memset(d, v, n);
    return 0;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipDeviceSynchronize() { return 0; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.0f;
    return 0;
}
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
// One std::thread per lane of a block for the whole launch: every thread runs its lane of
// each block in turn and meets the others at the block barrier between blocks (the
// blocks' __shared__ arrays are reused, as one block at a time runs).
template <class K, class... A> void cpu_launch(K k, dim3 g, dim3 b, size_t lds, A... a)
{
    gridDim = g;
    blockDim = b;
    std::vector<unsigned char> dyn(lds);   // exactly the launch's dynamic LDS
    cpu_dyn_lds = lds ? dyn.data() : nullptr;
    std::barrier<> bar(b.x);
    cpu_bar = &bar;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < b.x; t++)
        th.emplace_back([=, &bar]() {
            threadIdx = {t, 0, 0};
            for (unsigned bz = 0; bz < g.z; bz++)
                for (unsigned by = 0; by < g.y; by++)
                    for (unsigned bx = 0; bx < g.x; bx++) {
                        blockIdx = {bx, by, bz};
                        k(a...);
                        bar.arrive_and_wait();
                    }
        });
    for (auto &x : th) x.join();
}
#define hipLaunchKernelGGL(k, g, b, lds, s, ...) cpu_launch(k, dim3(g), dim3(b), lds, __VA_ARGS__)
