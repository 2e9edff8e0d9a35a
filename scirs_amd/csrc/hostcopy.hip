// Host <-> device copies of row-strided host arrays through pinned staging.
//
// A 2-D copy between pageable host memory and the device (hipMemcpy2DAsync
// on an ordinary host pointer) is carried out one row at a time: a
// 16384 x 16384 FP64 matrix moved at about 5 GB/s that way, and a 512 x 512
// one took 13 ms.  Here host threads pack the rows into one of two pinned
// 32 MiB buffers while the DMA engine moves the other, so the host-side copy
// and the PCIe transfer overlap and each DMA is one large transfer.
//
// The staging buffers and their events are process-wide and used only by the
// host-memory entry points, which all hold the context mutex.
#include "api_internal.hpp"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace scirs {
namespace {

constexpr size_t STAGE_BYTES = (size_t)32 << 20;
constexpr unsigned MAX_THREADS = 8;

struct Stage {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};   // a transfer from / into buf[b] may still run
};
Stage g_stage;

int stage_init()
{
    for (int b = 0; b < 2; ++b) {
        if (!g_stage.buf[b] && hipHostMalloc(&g_stage.buf[b], STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
            g_stage.buf[b] = nullptr;
            scirs_set_error("pinned staging allocation of %zu bytes failed", STAGE_BYTES);
            return SCIRS_ERR_NOMEM;
        }
        if (!g_stage.ev[b]) SCIRS_HIP_CHECK(hipEventCreateWithFlags(&g_stage.ev[b], hipEventDisableTiming));
    }
    return SCIRS_OK;
}

int stage_wait(int b)
{
    if (g_stage.pending[b]) {
        SCIRS_HIP_CHECK(hipEventSynchronize(g_stage.ev[b]));
        g_stage.pending[b] = false;
    }
    return SCIRS_OK;
}

// f(r0, r1) over rows [0, rows), split over up to MAX_THREADS host threads
// with at least 1 MiB each; a thread that cannot be started runs its share
// on the calling thread.
template <class F>
void par_rows(size_t rows, size_t bytes, const F &f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({(size_t)MAX_THREADS, (size_t)hw, rows, bytes / ((size_t)1 << 20) + 1});
    if (nt <= 1) {
        f((size_t)0, rows);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (size_t t = 1; t < nt; ++t) {
        const size_t a = rows * t / nt, e = rows * (t + 1) / nt;
        try {
            th.emplace_back([&f, a, e] { f(a, e); });
        } catch (...) {
            f(a, e);
        }
    }
    f((size_t)0, rows / nt);
    for (std::thread &x : th) x.join();
}

// rows [a, e) of a chunk between a packed buffer and a strided host array
void pack(char *dst, const char *src, size_t pitch, size_t width, size_t a, size_t e)
{
    if (pitch == width)
        memcpy(dst + a * width, src + a * width, (e - a) * width);
    else
        for (size_t r = a; r < e; ++r) memcpy(dst + r * width, src + r * pitch, width);
}

void unpack(char *dst, size_t pitch, const char *src, size_t width, size_t a, size_t e)
{
    if (pitch == width)
        memcpy(dst + a * width, src + a * width, (e - a) * width);
    else
        for (size_t r = a; r < e; ++r) memcpy(dst + r * pitch, src + r * width, width);
}

} // namespace

int copy_h2d(void *d, size_t dpitch, const void *h, size_t hpitch, size_t width, size_t rows, hipStream_t s)
{
    if (rows == 0 || width == 0) return SCIRS_OK;
    if (width > STAGE_BYTES) {   // a row longer than a staging buffer: no matrix of this library has one
        SCIRS_HIP_CHECK(hipMemcpy2DAsync(d, dpitch, h, hpitch, width, rows, hipMemcpyHostToDevice, s));
        return SCIRS_OK;
    }
    if (int rc = stage_init()) return rc;
    const size_t rpc = STAGE_BYTES / width;
    const char *src = (const char *)h;
    char *dst = (char *)d;
    int b = 0;
    for (size_t r0 = 0; r0 < rows; r0 += rpc, b ^= 1) {
        const size_t nr = std::min(rpc, rows - r0);
        if (int rc = stage_wait(b)) return rc;
        char *st = (char *)g_stage.buf[b];
        const char *sc = src + r0 * hpitch;
        par_rows(nr, nr * width, [&](size_t a, size_t e) { pack(st, sc, hpitch, width, a, e); });
        if (dpitch == width)
            SCIRS_HIP_CHECK(hipMemcpyAsync(dst + r0 * dpitch, st, nr * width, hipMemcpyHostToDevice, s));
        else
            SCIRS_HIP_CHECK(
                hipMemcpy2DAsync(dst + r0 * dpitch, dpitch, st, width, width, nr, hipMemcpyHostToDevice, s));
        SCIRS_HIP_CHECK(hipEventRecord(g_stage.ev[b], s));
        g_stage.pending[b] = true;
    }
    return SCIRS_OK;
}

int copy_d2h(void *h, size_t hpitch, const void *d, size_t dpitch, size_t width, size_t rows, hipStream_t s)
{
    if (rows == 0 || width == 0) return SCIRS_OK;
    if (width > STAGE_BYTES) {
        SCIRS_HIP_CHECK(hipMemcpy2DAsync(h, hpitch, d, dpitch, width, rows, hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipStreamSynchronize(s));
        return SCIRS_OK;
    }
    if (int rc = stage_init()) return rc;
    const size_t rpc = STAGE_BYTES / width, nch = (rows + rpc - 1) / rpc;
    const char *src = (const char *)d;
    char *dst = (char *)h;
    auto issue = [&](size_t ch) -> int {
        const int b = (int)(ch & 1);
        const size_t r0 = ch * rpc, nr = std::min(rpc, rows - r0);
        if (int rc = stage_wait(b)) return rc;
        if (dpitch == width)
            SCIRS_HIP_CHECK(hipMemcpyAsync(g_stage.buf[b], src + r0 * dpitch, nr * width, hipMemcpyDeviceToHost, s));
        else
            SCIRS_HIP_CHECK(hipMemcpy2DAsync(g_stage.buf[b], width, src + r0 * dpitch, dpitch, width, nr,
                                             hipMemcpyDeviceToHost, s));
        SCIRS_HIP_CHECK(hipEventRecord(g_stage.ev[b], s));
        g_stage.pending[b] = true;
        return SCIRS_OK;
    };
    for (size_t ch = 0; ch < std::min<size_t>(2, nch); ++ch)
        if (int rc = issue(ch)) return rc;
    for (size_t ch = 0; ch < nch; ++ch) {
        const int b = (int)(ch & 1);
        if (int rc = stage_wait(b)) return rc;
        const size_t r0 = ch * rpc, nr = std::min(rpc, rows - r0);
        const char *st = (const char *)g_stage.buf[b];
        char *dc = dst + r0 * hpitch;
        par_rows(nr, nr * width, [&](size_t a, size_t e) { unpack(dc, hpitch, st, width, a, e); });
        if (ch + 2 < nch)
            if (int rc = issue(ch + 2)) return rc;
    }
    return SCIRS_OK;
}

int copy_h2d_flat(void *d, const void *h, size_t bytes, hipStream_t s)
{
    constexpr size_t W = (size_t)1 << 16;
    const size_t rows = bytes / W, tail = bytes - rows * W;
    if (rows)
        if (int rc = copy_h2d(d, W, h, W, W, rows, s)) return rc;
    if (tail) return copy_h2d((char *)d + rows * W, tail, (const char *)h + rows * W, tail, tail, 1, s);
    return SCIRS_OK;
}

int copy_d2h_flat(void *h, const void *d, size_t bytes, hipStream_t s)
{
    constexpr size_t W = (size_t)1 << 16;
    const size_t rows = bytes / W, tail = bytes - rows * W;
    if (rows)
        if (int rc = copy_d2h(h, W, d, W, W, rows, s)) return rc;
    if (tail) return copy_d2h((char *)h + rows * W, tail, (const char *)d + rows * W, tail, tail, 1, s);
    return SCIRS_OK;
}

} // namespace scirs
