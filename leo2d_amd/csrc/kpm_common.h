// kpm_common.h -- device helpers shared by the KPM step kernels (kpm.hip)
// and the batched recursion (batch.hip): complex/real element arithmetic in
// the recursion's precision, FP64 moment terms, and the counter-based
// random-phase generator.  Internal to libleo2d.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace leo2d {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// ---- element arithmetic ------------------------------------------------
struct cf32 { float x, y; };

template <typename X> struct acc_of;
template <> struct acc_of<double> { using type = double; };
template <> struct acc_of<double2> { using type = double2; };
template <> struct acc_of<float2> { using type = float2; };
template <> struct acc_of<float> { using type = float; };

__device__ inline void zero(double& a) { a = 0.0; }
__device__ inline void zero(double2& a) { a.x = 0.0; a.y = 0.0; }
__device__ inline void zero(float2& a) { a.x = 0.0f; a.y = 0.0f; }
__device__ inline void zero(float& a) { a = 0.0f; }

// acc += h * x
__device__ inline void mac(double& acc, double h, double x) { acc = fma(h, x, acc); }
__device__ inline void mac(double2& acc, double h, double2 x)
{
    acc.x = fma(h, x.x, acc.x);
    acc.y = fma(h, x.y, acc.y);
}
__device__ inline void mac(double2& acc, double2 h, double2 x)
{
    acc.x = fma(h.x, x.x, acc.x);
    acc.x = fma(-h.y, x.y, acc.x);
    acc.y = fma(h.x, x.y, acc.y);
    acc.y = fma(h.y, x.x, acc.y);
}
__device__ inline void mac(float& acc, float h, float x) { acc = fmaf(h, x, acc); }
__device__ inline void mac(float2& acc, float h, float2 x)
{
    acc.x = fmaf(h, x.x, acc.x);
    acc.y = fmaf(h, x.y, acc.y);
}
__device__ inline void mac(float2& acc, float2 h, float2 x)
{
    acc.x = fmaf(h.x, x.x, acc.x);
    acc.x = fmaf(-h.y, x.y, acc.x);
    acc.y = fmaf(h.x, x.y, acc.y);
    acc.y = fmaf(h.y, x.x, acc.y);
}

// out = alpha * acc - prev  (alpha is 1 or 2, exact)
__device__ inline double comb(double alpha, double acc, double prev) { return alpha * acc - prev; }
__device__ inline double2 comb(double alpha, double2 acc, double2 prev)
{
    double2 o;
    o.x = alpha * acc.x - prev.x;
    o.y = alpha * acc.y - prev.y;
    return o;
}
__device__ inline float2 comb(double alpha, float2 acc, float2 prev)
{
    float al = (float)alpha;
    float2 o;
    o.x = al * acc.x - prev.x;
    o.y = al * acc.y - prev.y;
    return o;
}
__device__ inline float comb(double alpha, float acc, float prev) { return (float)alpha * acc - prev; }

// FP64 moment terms
__device__ inline double nrm2(double a) { return a * a; }
__device__ inline double nrm2(double2 a) { return fma(a.x, a.x, a.y * a.y); }
__device__ inline double nrm2(float2 a)
{
    double x = a.x, y = a.y;
    return fma(x, x, y * y);
}
__device__ inline double nrm2(float a)
{
    double x = a;
    return x * x;
}
__device__ inline double redot(double a, double b) { return a * b; }
__device__ inline double redot(double2 a, double2 b) { return fma(a.x, b.x, a.y * b.y); }
__device__ inline double redot(float2 a, float2 b)
{
    double ax = a.x, ay = a.y, bx = b.x, by = b.y;
    return fma(ax, bx, ay * by);
}
__device__ inline double redot(float a, float b) { return (double)a * (double)b; }
__device__ inline double re(double a) { return a; }
__device__ inline double re(double2 a) { return a.x; }
__device__ inline double re(float2 a) { return a.x; }
__device__ inline double re(float a) { return a; }

// ---- 16-lane row broadcasts --------------------------------------------------
// bc16<K>(v): every lane receives v from lane K of its 16-lane row (DPP
// row_newbcast).  A row of 16 lanes loads 16 CSR entries in one access and
// passes them to each other this way.
// Every lane of the row is enabled and row_newbcast always has a valid source,
// so the plain DPP move (bound_ctrl set, no "old" value to materialise) is used.
template <int K>
__device__ inline int bc16(int v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}
template <int K>
__device__ inline double bc16(double v)
{
    // one 64-bit DPP move (gfx950 allows row_newbcast on 64-bit operands)
    return __builtin_amdgcn_mov_dpp(v, 0x150 + K, 0xf, 0xf, true);
}
template <int K>
__device__ inline float bc16(float v) { return __int_as_float(bc16<K>(__float_as_int(v))); }
template <int K>
__device__ inline double2 bc16(double2 v) { return make_double2(bc16<K>(v.x), bc16<K>(v.y)); }
template <int K>
__device__ inline float2 bc16(float2 v) { return make_float2(bc16<K>(v.x), bc16<K>(v.y)); }

// gathers and multiply-adds for the 16 entries of one chunk held by a
// 16-lane row (entry K in lane K); only the first m entries are live
template <typename XT, int R, int K>
__device__ inline void xchunk_gather(XT* xs, int m, int mc, const XT* __restrict__ cur, int voff)
{
    if constexpr (K < 16) {
        const int c = bc16<K>(mc);
        if (K < m) xs[K] = cur[(size_t)c * R + voff];
        xchunk_gather<XT, R, K + 1>(xs, m, mc, cur, voff);
    }
}
template <typename HT, typename XT, typename AT, int K>
__device__ inline void xchunk_mac(AT& acc, int m, HT mh, const XT* xs)
{
    if constexpr (K < 16) {
        const HT h = bc16<K>(mh);
        if (K < m) mac(acc, h, xs[K]);
        xchunk_mac<HT, XT, AT, K + 1>(acc, m, mh, xs);
    }
}

// ---- random-phase start vectors ----------------------------------------------
// phi(seed, r, i) in [0, 2 pi): three splitmix64 finalisations keyed by the
// seed, the global vector index and the orbital, so every vector entry is
// independent of the thread layout that generates it.
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline double random_phase_angle(uint64_t seed, uint64_t r, uint64_t i)
{
    uint64_t h = mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL);
    h = mix64(h ^ (r * 0xD1B54A32D192ED03ULL));
    h = mix64(h ^ (i * 0x8CB92BA72F3D8DD7ULL));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 6.283185307179586;
}

}  // namespace
}  // namespace leo2d
