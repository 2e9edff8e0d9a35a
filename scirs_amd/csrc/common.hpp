// Shared device/host helpers for the gfx950 dense-solve kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#define SCIRS_WAVE 64

// Status codes shared by every entry point (mirrored in include/scirs_gpu.h).
enum {
    SCIRS_OK = 0,
    SCIRS_ERR_ARG = -1,
    SCIRS_ERR_HIP = -2,
    SCIRS_ERR_NOMEM = -3,
    SCIRS_ERR_TIMEOUT = -4,
    SCIRS_ERR_NOT_CONVERGED = -5,
    SCIRS_ERR_UNSUPPORTED = -6,
};

template <typename T> struct num_limits;
template <> struct num_limits<double> { static __host__ __device__ constexpr double eps() { return DBL_EPSILON; } };
template <> struct num_limits<float>  { static __host__ __device__ constexpr float  eps() { return FLT_EPSILON; } };

// Wave-wide arg-max of |v| with the lowest row index winning ties, which is
// the strict '>' scan order of scirs2-linalg/src/lapack.rs:95-104.
// idx < 0 marks "no candidate".
template <typename T>
__device__ __forceinline__ void wave_argmax(T &v, int &idx)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        T ov = __shfl_xor(v, off, SCIRS_WAVE);
        int oi = __shfl_xor(idx, off, SCIRS_WAVE);
        bool take = (oi >= 0) && (idx < 0 || ov > v || (ov == v && oi < idx));
        if (take) { v = ov; idx = oi; }
    }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, SCIRS_WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, SCIRS_WAVE));
    return v;
}

// Grid-wide barrier for cooperative launches.  `bar` is a counter that every
// workgroup increments once per barrier; barrier number t (1-based) completes
// when the counter reaches t * nblocks.  The wait is bounded: after
// `spin_limit` polls the workgroup raises *abort and leaves, so a kernel can
// never wait forever.  All threads fence at agent scope on both sides so that
// global writes made before the barrier by any thread of any workgroup are
// visible to every thread after it.
__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned target, int *abort,
                                             unsigned spin_limit)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ok = 0; break; }
            if (++spins > spin_limit) {
                __hip_atomic_store(abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        s_ok = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return s_ok != 0;
}

// Host-side error helper.
#define SCIRS_HIP_CHECK(expr)                                                   \
    do {                                                                        \
        hipError_t _e = (expr);                                                 \
        if (_e != hipSuccess) {                                                 \
            scirs_set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr,       \
                            hipGetErrorString(_e));                             \
            return SCIRS_ERR_HIP;                                               \
        }                                                                       \
    } while (0)

extern "C" void scirs_set_error(const char *fmt, ...);

// Kernels that take more than 64 KiB of dynamic LDS must opt in once; the
// ceiling is the 160 KiB of a gfx950 CU minus the kernel's static LDS.
// Returns SCIRS_ERR_HIP (with a message) when the request cannot be granted.
int scirs_allow_lds(const void *kernel, size_t dynamic_bytes);
