// A CPU stand-in for the parts of the HIP runtime the library uses, so that
// the library's own sources (kernels and host orchestration alike) build
// with g++ under AddressSanitizer and run on the CPU. Test infrastructure:
// only the bounds-checking build in this folder puts it on the include path.
//
// Device memory is host heap memory filled with 0xCD bytes when allocated,
// so that a read of memory no kernel or copy wrote gives large indices and
// NaN-like floats instead of the zeros fresh device memory often holds.
// A launch runs every thread of every block one after another, which is a
// legal schedule for kernels without barriers or wave-level exchange; the
// prefix sum (the one kernel family with both) is replaced by emu_scan.cpp.
#pragma once

#include <chrono>
#include <condition_variable>
#include <mutex>
#include <set>
#include <thread>
#include <vector>
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

// The kernels' bounds assertions, live in this build only.
[[noreturn]] inline void kdn_emu_fail(const char* what, const char* file, int line) {
    std::fprintf(stderr, "KDN_CHECK failed: %s at %s:%d\n", what, file, line);
    std::abort();
}
#define KDN_CHECK(cond) ((cond) ? (void)0 : kdn_emu_fail(#cond, __FILE__, __LINE__))
// A histogram of the cells plus march steps each traced ray took
// (bucket b counts rays with 2^(b-1) < iterations <= 2^b), read through
// kdn_emu_iterations().
inline uint64_t kdn_emu_histogram[16];
inline int kdn_emu_launch_most = 0;  // the most any ray of the running launch took
inline void kdn_emu_count_iterations(int n) {
    ++kdn_emu_histogram[n <= 1 ? 0 : (32 - __builtin_clz((unsigned)n - 1)) & 15];
    kdn_emu_launch_most = n > kdn_emu_launch_most ? n : kdn_emu_launch_most;
}
#define KDN_TRACE_ITERATIONS(n) kdn_emu_count_iterations(n)
extern "C" __attribute__((used, visibility("default"))) inline const uint64_t* kdn_emu_iterations() { return kdn_emu_histogram; }

// this build runs glass scattering, which the device build holds back
// (-DKDN_EMU_AS_DEVICE keeps the device build's refusal, to test it)
#ifndef KDN_EMU_AS_DEVICE
#define KDN_GLASS_ON_DEVICE 1
#endif

#define __device__
#define __constant__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
// a block's shared memory: one array for every thread (threads run one
// after another here, each in its own column)
#define __shared__ static

struct dim3 {
    uint32_t x, y, z;
    constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
inline thread_local dim3 blockIdx{0, 0, 0}, threadIdx{0, 0, 0}, blockDim{1, 1, 1}, gridDim{1, 1, 1};

enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNoDevice = 100 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
struct EmuStream {
    int unused;
};
struct EmuEvent {
    std::chrono::steady_clock::time_point at;
};
typedef EmuStream* hipStream_t;
typedef EmuEvent* hipEvent_t;

inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
// a small device: the launches that size their grid by it run few blocks,
// so their lanes take many queue entries each
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
inline hipError_t hipDeviceGetAttribute(int* value, hipDeviceAttribute_t, int) {
    *value = 2;
    return hipSuccess;
}
inline hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
// a __device__ variable is host memory here: its symbol is its address
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}

template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
    void* m = std::malloc(bytes ? bytes : 1);
    if (!m) return hipErrorOutOfMemory;
    std::memset(m, 0xCD, bytes);
    *p = (T*)m;
    return hipSuccess;
}
inline hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
    if (bytes) std::memcpy(dst, src, bytes);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t = nullptr) {
    return hipMemcpy(dst, src, bytes, k);
}
inline hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t = nullptr) {
    if (bytes) std::memset(dst, v, bytes);
    return hipSuccess;
}
inline hipError_t hipMemset(void* dst, int v, size_t bytes) { return hipMemsetAsync(dst, v, bytes); }

inline hipError_t hipStreamCreate(hipStream_t* s) {
    *s = new EmuStream{0};
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }

inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new EmuEvent{};
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
    e->at = std::chrono::steady_clock::now();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
// every stream of this build runs its work at once, in call order: an
// event and a wait are kept only as calls
constexpr unsigned hipEventDisableTiming = 2;
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->at - a->at).count();
    return hipSuccess;
}

// KDN_EMU_ORDER=reverse runs the threads of a launch last to first: a
// kernel whose output depends on the order of its threads (beyond the
// order of atomic appends) then gives different results.
inline bool kdn_emu_reverse() {
    static const bool reverse = [] {
        const char* v = std::getenv("KDN_EMU_ORDER");
        return v && std::strcmp(v, "reverse") == 0;
    }();
    return reverse;
}

// KDN_EMU_LAUNCHES=1 prints, after every launch that traced rays, its
// thread count and the most cells plus march steps one ray took.
inline bool kdn_emu_print_launches() {
    static const bool on = std::getenv("KDN_EMU_LAUNCHES") != nullptr;
    return on;
}

template <class... P, class... A>
inline void kdn_emu_run(void (*kernel)(P...), dim3 grid, dim3 block, A&&... args);

template <class... P, class... A>
inline void kdn_emu_launch(void (*kernel)(P...), dim3 grid, dim3 block, A&&... args) {
    kdn_emu_launch_most = 0;
    kdn_emu_run(kernel, grid, block, args...);
    if (kdn_emu_print_launches() && kdn_emu_launch_most > 0) {
        std::fprintf(stderr, "[launch] %u threads, most iterations %d\n", grid.x * block.x, kdn_emu_launch_most);
    }
}

// Kernels that synchronise their block (__syncthreads) are marked with
// KDN_SYNCS_BLOCK next to their definition. Their blocks run one after
// another, each block's threads as OS threads at once (their __shared__
// arrays shared as on the device), meeting at a real barrier; every other
// kernel's threads run one after another.
inline std::mutex& kdn_emu_marks_lock() {
    static std::mutex m;
    return m;
}
inline std::set<const void*>& kdn_emu_marks() {
    static std::set<const void*> marks;
    return marks;
}
inline bool kdn_emu_mark(const void* kernel) {
    std::lock_guard<std::mutex> g(kdn_emu_marks_lock());
    kdn_emu_marks().insert(kernel);
    return true;
}
inline bool kdn_emu_marked(const void* kernel) {
    std::lock_guard<std::mutex> g(kdn_emu_marks_lock());
    return kdn_emu_marks().count(kernel) != 0;
}
#define KDN_SYNCS_BLOCK(kernel) static const bool kdn_emu_syncs_##kernel = kdn_emu_mark((const void*)&kernel);

struct KdnEmuBarrier {
    std::mutex m;
    std::condition_variable cv;
    uint64_t count = 0, waiting = 0, generation = 0;
    void wait() {
        std::unique_lock<std::mutex> lock(m);
        const uint64_t g = generation;
        if (++waiting == count) {
            waiting = 0;
            ++generation;
            cv.notify_all();
        } else {
            cv.wait(lock, [&] { return generation != g; });
        }
    }
};
inline thread_local KdnEmuBarrier* kdn_emu_barrier = nullptr;
inline void __syncthreads() {
    if (!kdn_emu_barrier) {
        std::fprintf(stderr, "emulation: a kernel calls __syncthreads but is not marked KDN_SYNCS_BLOCK\n");
        std::abort();
    }
    kdn_emu_barrier->wait();
}

template <class... P, class... A>
inline void kdn_emu_run(void (*kernel)(P...), dim3 grid, dim3 block, A&&... args) {
    gridDim = grid;
    blockDim = block;
    if (kdn_emu_marked((const void*)kernel)) {
        const uint64_t blocks = (uint64_t)grid.x * grid.y * grid.z, threads = (uint64_t)block.x * block.y * block.z;
        for (uint64_t b = 0; b < blocks; ++b) {
            KdnEmuBarrier barrier;
            barrier.count = threads;
            const dim3 at((uint32_t)(b % grid.x), (uint32_t)(b / grid.x % grid.y), (uint32_t)(b / grid.x / grid.y));
            std::vector<std::thread> pool;
            pool.reserve(threads);
            for (uint64_t t = 0; t < threads; ++t) {
                pool.emplace_back([&, t] {
                    gridDim = grid;
                    blockDim = block;
                    blockIdx = at;
                    threadIdx = dim3((uint32_t)(t % block.x), (uint32_t)(t / block.x % block.y), (uint32_t)(t / block.x / block.y));
                    kdn_emu_barrier = &barrier;
                    kernel(args...);
                    kdn_emu_barrier = nullptr;
                });
            }
            for (std::thread& th : pool) th.join();
        }
        return;
    }
    const uint64_t blocks = (uint64_t)grid.x * grid.y * grid.z, threads = (uint64_t)block.x * block.y * block.z;
    const bool reverse = kdn_emu_reverse();
    for (uint64_t bi = 0; bi < blocks; ++bi) {
        const uint64_t b = reverse ? blocks - 1 - bi : bi;
        blockIdx = dim3((uint32_t)(b % grid.x), (uint32_t)(b / grid.x % grid.y), (uint32_t)(b / grid.x / grid.y));
        for (uint64_t ti = 0; ti < threads; ++ti) {
            const uint64_t t = reverse ? threads - 1 - ti : ti;
            threadIdx = dim3((uint32_t)(t % block.x), (uint32_t)(t / block.x % block.y), (uint32_t)(t / block.x / block.y));
            kernel(args...);
        }
    }
}
#define hipLaunchKernelGGL(kernel, grid, block, shared, stream, ...) kdn_emu_launch(kernel, dim3(grid), dim3(block), __VA_ARGS__)

inline uint32_t __float_as_uint(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    return b;
}
inline float __uint_as_float(uint32_t b) {
    float x;
    std::memcpy(&x, &b, 4);
    return x;
}
inline int __float_as_int(float x) { return (int)__float_as_uint(x); }
inline float __int_as_float(int b) { return __uint_as_float((uint32_t)b); }

template <class T>
inline T atomicAdd(T* p, T v) {
    const T old = *p;
    *p = old + v;
    return old;
}
template <class T>
inline T atomicOr(T* p, T v) {
    const T old = *p;
    *p = old | v;
    return old;
}
template <class T>
inline T atomicMax(T* p, T v) {
    const T old = *p;
    if (v > old) *p = v;
    return old;
}
template <class T>
inline T atomicMin(T* p, T v) {
    const T old = *p;
    if (v < old) *p = v;
    return old;
}
