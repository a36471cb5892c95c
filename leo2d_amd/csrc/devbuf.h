// devbuf.h -- owning device buffers and HIP error checks shared by the
// host-side implementation files of libleo2d.  Internal.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../../include/leo2d.h"

namespace leo2d {

// owning device allocation (freed on release() or destruction; movable, not copyable)
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// pinned host staging memory, grown on demand and kept until released
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    HostBuf(HostBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    HostBuf& operator=(HostBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~HostBuf() { release(); }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

inline void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? LEO2D_ERR_ALLOC : LEO2D_ERR_HIP;
        throw HipError(std::string(what) + ": " + hipGetErrorString(e), code);
    }
}

// at least `bytes` of pinned host memory in b (contents not kept when it grows)
inline void halloc(HostBuf& b, size_t bytes, const char* what)
{
    if (bytes <= b.bytes && b.p) return;
    b.release();
    const hipError_t e = hipHostMalloc(&b.p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        check(e, what);
    }
    b.bytes = bytes;
}

inline void dalloc(DevBuf& b, size_t bytes, const char* what)
{
    b.release();
    if (bytes == 0) bytes = 16;
    const hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        // a failed allocation is not sticky, but it stays the runtime's last error:
        // clear it, or the next launch check (hipGetLastError) would report it
        (void)hipGetLastError();
        b.p = nullptr;
        check(e, what);
    }
    b.bytes = bytes;
}

}  // namespace leo2d
