// Host stand-in for the HIP runtime, for the CPU build of the library's kernel and host sources (tests/hostbuild).
//
// Test infrastructure only: the product is built by hipcc for gfx950 and never sees this header.  Here a kernel
// launch runs every workgroup on the calling thread, one fiber per lane.  Wavefront operations (ballot, shuffles,
// readlane, wave barriers) and workgroup barriers are the points where a lane yields; a wavefront's operation is
// resolved once every live lane of the block has reached its next such point, over the lanes that reached the same
// operation (the lanes a divergent branch leaves out do not take part, as with an execution mask).  Device memory is
// host heap memory and every allocation, the dynamic LDS of each launch included, is exactly the size asked for, so
// that AddressSanitizer reports any access outside it.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <tuple>
#include <type_traits>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __shared__ static
#define __constant__
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
    unsigned x, y, z;
    constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotSupported = 801,
    hipErrorUnknown = 999,
} hipError_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipDeviceAttribute_t {
    hipDeviceAttributeMaxSharedMemoryPerBlock = 1,
    hipDeviceAttributeWallClockRate = 2,
} hipDeviceAttribute_t;

struct ihipStream_t;
struct ihipEvent_t;
typedef ihipStream_t *hipStream_t;
typedef ihipEvent_t *hipEvent_t;

#define hipEventDisableSystemFence 0x20000000u
#define hipHostMallocDefault 0u
#define HIP_SYMBOL(s) (&(s))

namespace emu {

struct Lane;
extern Lane *cur;
extern dim3 blockIdxV, blockDimV, gridDimV;

unsigned lane_tid();
void *dyn_lds();

enum WaveKind { W_BALLOT, W_SHFL, W_SHFL_DOWN, W_SHFL_XOR, W_READLANE, W_READFIRST, W_BARRIER, W_SHFL_UP };
uint64_t wave_op(int kind, uint64_t bits, int arg, int line);
int block_op(int orValue, int isOr, int line);

// the body every lane of every workgroup runs (the kernel with its arguments bound)
void run_grid(dim3 grid, dim3 block, std::size_t lds, const std::function<void()> &body);

template <class... P, class... A>
void launch(void (*k)(P...), dim3 grid, dim3 block, std::size_t lds, A &&...a) {
    std::tuple<typename std::decay<P>::type...> args(static_cast<typename std::decay<P>::type>(a)...);
    run_grid(grid, block, lds, [&]() { std::apply(k, args); });
}

template <class T>
inline uint64_t to_bits(T v) {
    static_assert(sizeof(T) <= 8, "wave values are at most 64 bits");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(uint64_t b) {
    T v;
    std::memcpy(&v, &b, sizeof(T));
    return v;
}
template <class T>
inline T wave_value(int kind, T v, int arg, int line) {
    return from_bits<T>(wave_op(kind, to_bits(v), arg, line));
}

inline unsigned lane_id() { return lane_tid() & 63u; }

}  // namespace emu

struct EmuThreadIdx {
    unsigned x, y, z;
};
EmuThreadIdx emu_thread_idx();
#define threadIdx (emu_thread_idx())
#define blockIdx (emu::blockIdxV)
#define blockDim (emu::blockDimV)
#define gridDim (emu::gridDimV)
#define warpSize 64

// ---- wavefront and workgroup operations
#define __syncthreads() ((void)emu::block_op(0, 0, __LINE__))
#define __syncthreads_or(p) emu::block_op((p) ? 1 : 0, 1, __LINE__)
#define __ballot(p) ((unsigned long long)emu::wave_op(emu::W_BALLOT, (p) ? 1u : 0u, 0, __LINE__))
#define __any(p) (emu::wave_op(emu::W_BALLOT, (p) ? 1u : 0u, 0, __LINE__) != 0 ? 1 : 0)
#define __all(p) (emu::wave_op(emu::W_BALLOT, (p) ? 0u : 1u, 0, __LINE__) == 0 ? 1 : 0)
#define EMU_SHFL_PICK(_1, _2, _3, NAME, ...) NAME
#define EMU_SHFL2(v, s) emu::wave_value(emu::W_SHFL, (v), (int)(s), __LINE__)
#define EMU_SHFL3(v, s, w) emu::wave_value(emu::W_SHFL, (v), (int)(s), __LINE__)
#define __shfl(...) EMU_SHFL_PICK(__VA_ARGS__, EMU_SHFL3, EMU_SHFL2, )(__VA_ARGS__)
#define EMU_DOWN2(v, d) emu::wave_value(emu::W_SHFL_DOWN, (v), (int)(d), __LINE__)
#define EMU_DOWN3(v, d, w) emu::wave_value(emu::W_SHFL_DOWN, (v), (int)(d), __LINE__)
#define __shfl_down(...) EMU_SHFL_PICK(__VA_ARGS__, EMU_DOWN3, EMU_DOWN2, )(__VA_ARGS__)
#define EMU_UP2(v, d) emu::wave_value(emu::W_SHFL_UP, (v), (int)(d), __LINE__)
#define EMU_UP3(v, d, w) emu::wave_value(emu::W_SHFL_UP, (v), (int)(d), __LINE__)
#define __shfl_up(...) EMU_SHFL_PICK(__VA_ARGS__, EMU_UP3, EMU_UP2, )(__VA_ARGS__)
#define EMU_XOR2(v, m) emu::wave_value(emu::W_SHFL_XOR, (v), (int)(m), __LINE__)
#define EMU_XOR3(v, m, w) emu::wave_value(emu::W_SHFL_XOR, (v), (int)(m), __LINE__)
#define __shfl_xor(...) EMU_SHFL_PICK(__VA_ARGS__, EMU_XOR3, EMU_XOR2, )(__VA_ARGS__)
#define __builtin_amdgcn_readlane(v, l) emu::wave_value(emu::W_READLANE, (v), (int)(l), __LINE__)
#define __builtin_amdgcn_readfirstlane(v) emu::wave_value(emu::W_READFIRST, (v), 0, __LINE__)
#define __builtin_amdgcn_wave_barrier() ((void)emu::wave_op(emu::W_BARRIER, 0, 0, __LINE__))
#define __builtin_amdgcn_s_setprio(n) ((void)0)

inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned mask, unsigned acc) {
    const unsigned l = emu::lane_id();
    const unsigned below = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
    return acc + (unsigned)__builtin_popcount(mask & below);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned mask, unsigned acc) {
    const unsigned l = emu::lane_id();
    const unsigned below = l < 32 ? 0u : ((1u << (l - 32)) - 1u);
    return acc + (unsigned)__builtin_popcount(mask & below);
}

inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline unsigned __float_as_uint(float f) { return emu::from_bits<unsigned>(emu::to_bits(f)); }
unsigned long long wall_clock64();
inline unsigned __smid() { return 0; }

#ifndef __HIP_MEMORY_SCOPE_AGENT
#define __HIP_MEMORY_SCOPE_AGENT 3
#endif
template <class T>
inline T __hip_atomic_load(const T *p, int, int) {
    return *p;
}
template <class T, class U>
inline T atomicAdd(T *p, U v) {
    T o = *p;
    *p = (T)(o + (T)v);
    return o;
}
template <class T, class U>
inline T atomicOr(T *p, U v) {
    T o = *p;
    *p = (T)(o | (T)v);
    return o;
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
    T o = *p;
    if ((T)v > o) *p = (T)v;
    return o;
}
template <class T, class U>
inline T atomicMin(T *p, U v) {
    T o = *p;
    if ((T)v < o) *p = (T)v;
    return o;
}

template <class A, class B>
inline typename std::common_type<A, B>::type min(A a, B b) {
    return a < b ? a : b;
}
template <class A, class B>
inline typename std::common_type<A, B>::type max(A a, B b) {
    return a > b ? a : b;
}

// ---- launches (run to completion on the calling thread)
#define hipLaunchKernelGGL(k, grid, block, lds, stream, ...) emu::launch(k, dim3(grid), dim3(block), (lds), ##__VA_ARGS__)
#define hipExtLaunchKernelGGL(k, grid, block, lds, stream, start, stop, flags, ...) \
    emu::launch(k, dim3(grid), dim3(block), (lds), ##__VA_ARGS__)

// ---- runtime API
hipError_t hipMalloc(void **p, std::size_t n);
template <class T>
inline hipError_t hipMalloc(T **p, std::size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), n);
}
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, std::size_t n, unsigned flags);
template <class T>
inline hipError_t hipHostMalloc(T **p, std::size_t n, unsigned flags) {
    return hipHostMalloc(reinterpret_cast<void **>(p), n, flags);
}
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, std::size_t n, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, std::size_t n, hipMemcpyKind kind, hipStream_t s);
hipError_t hipMemset(void *p, int v, std::size_t n);
hipError_t hipMemsetAsync(void *p, int v, std::size_t n, hipStream_t s);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
template <class T>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *blocks, T, int, std::size_t) {
    *blocks = 1;
    return hipSuccess;
}
