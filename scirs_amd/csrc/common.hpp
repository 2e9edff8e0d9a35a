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

// DPP moves (GFX9 family): lane l receives the value of the lane selected by
// CTRL among the rows enabled in ROWMASK; every other lane receives `old`.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i32(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_val(float old, float src)
{
    return __int_as_float(dpp_i32<CTRL, ROWMASK>(__float_as_int(old), __float_as_int(src)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_val(double old, double src)
{
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = dpp_i32<CTRL, ROWMASK>((int)(o & 0xffffffffll), (int)(v & 0xffffffffll));
    const int hi = dpp_i32<CTRL, ROWMASK>((int)(o >> 32), (int)(v >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename T>
__device__ __forceinline__ T readlane_val(T v, int k);
template <>
__device__ __forceinline__ float readlane_val<float>(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
template <>
__device__ __forceinline__ double readlane_val<double>(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// (v, idx) <- the better of itself and the pair DPP brings in: larger value,
// then lower index; idx < 0 is "no candidate".
template <typename T, int CTRL, int ROWMASK>
__device__ __forceinline__ void argmax_dpp_step(T &v, int &idx)
{
    const T ov = dpp_val<CTRL, ROWMASK>((T)0, v);
    const int oi = dpp_i32<CTRL, ROWMASK>(-1, idx);
    if (oi >= 0 && (idx < 0 || ov > v || (ov == v && oi < idx))) { v = ov; idx = oi; }
}

// Wave-wide arg-max of |v| with the lowest row index winning ties, which is
// the strict '>' scan of scirs2-linalg/src/lapack.rs:95-104.  idx < 0 marks
// "no candidate".  Row scans (row_shr 1, 2, 4, 8) and row broadcasts
// (row_bcast 15, 31) run on the VALU through DPP; lane 63 ends with the result,
// which every lane then reads.
template <typename T>
__device__ __forceinline__ void wave_argmax(T &v, int &idx)
{
    argmax_dpp_step<T, 0x111, 0xf>(v, idx);
    argmax_dpp_step<T, 0x112, 0xf>(v, idx);
    argmax_dpp_step<T, 0x114, 0xf>(v, idx);
    argmax_dpp_step<T, 0x118, 0xf>(v, idx);
    argmax_dpp_step<T, 0x142, 0xa>(v, idx);
    argmax_dpp_step<T, 0x143, 0xc>(v, idx);
    v = readlane_val(v, 63);
    idx = __builtin_amdgcn_readlane(idx, 63);
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
