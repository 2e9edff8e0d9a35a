// Shared device/host helpers for the gfx950 dense-solve kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#include "../../include/scirs_gpu.h"   // status codes SCIRS_OK, SCIRS_ERR_*

#define SCIRS_WAVE 64

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

// Copy an nrows x ncols block of a row-major global matrix into LDS (row pitch
// `pitch`), with every thread issuing UNROLL loads before it stores any, so the
// copy pays the memory latency once per UNROLL elements instead of once per
// element.  Elements outside [0, vrows) x [0, vcols) of the source are
// written as zero.  Call from all threads of the workgroup.
template <typename T, int UNROLL = 8>
__device__ __forceinline__ void g2s_block(T *lds, int pitch, const T *g, long ld, int nrows, int ncols, int vrows,
                                          int vcols)
{
    const int total = nrows * ncols;
    const int nt = blockDim.x;
    for (int base = threadIdx.x; base < total; base += UNROLL * nt) {
        T v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int idx = base + u * nt;
            const int r = idx / ncols, c = idx - r * ncols;
            v[u] = (idx < total && r < vrows && c < vcols) ? g[(long)r * ld + c] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const int idx = base + u * nt;
            const int r = idx / ncols, c = idx - r * ncols;
            if (idx < total) lds[r * pitch + c] = v[u];
        }
    }
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
