// Shared helpers of the tropidist_amd HIP library (gfx950): status codes, error capture, byte-lane arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>

#include "../../include/tdp.h"

namespace tdp {

// Last error message of the calling thread (tdp_last_error).
void set_error(const char *fmt, ...);

#define TDP_HIP_CHECK(expr)                                                                         \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) {                                                                     \
            ::tdp::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return TDP_EHIP;                                                                        \
        }                                                                                           \
    } while (0)

#define TDP_LAUNCH_CHECK() TDP_HIP_CHECK(hipGetLastError())

#define TDP_REQUIRE(cond, ...)                                                                      \
    do {                                                                                            \
        if (!(cond)) {                                                                              \
            ::tdp::set_error(__VA_ARGS__);                                                          \
            return TDP_EINVAL;                                                                      \
        }                                                                                           \
    } while (0)

static inline int top_bit(uint64_t x) { return 63 - __builtin_clzll(x); }

static inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------------------------- byte lanes
// Four uint8 costs in one 32-bit word.  Saturating add of a small weight and the byte-wise minimum are done on
// two 16-bit lanes per word (even and odd bytes), where neither can carry into its neighbour.

__device__ __forceinline__ uint32_t sat_add4(uint32_t x, uint32_t w) {
    uint32_t ww = w * 0x00010001u;
    uint32_t e = (x & 0x00FF00FFu) + ww;          // <= 255 + 255 per 16-bit lane
    uint32_t o = ((x >> 8) & 0x00FF00FFu) + ww;
    // clamp each 16-bit lane to 255: lanes above 255 have bit 8 set (w <= 255)
    uint32_t em = ((e >> 8) & 0x00010001u) * 0xFFu;
    uint32_t om = ((o >> 8) & 0x00010001u) * 0xFFu;
    e = (e | em) & 0x00FF00FFu;
    o = (o | om) & 0x00FF00FFu;
    return e | (o << 8);
}

__device__ __forceinline__ uint32_t min4(uint32_t a, uint32_t b) {
    uint32_t ae = a & 0x00FF00FFu, be = b & 0x00FF00FFu;
    uint32_t ao = (a >> 8) & 0x00FF00FFu, bo = (b >> 8) & 0x00FF00FFu;
    // per 16-bit lane: a - b + 0x100 keeps bit 8 set iff a >= b
    uint32_t ge_e = ((ae | 0x01000100u) - be) & 0x01000100u;
    uint32_t ge_o = ((ao | 0x01000100u) - bo) & 0x01000100u;
    uint32_t se = (ge_e >> 8) * 0xFFu;             // 0xFF in lanes where a >= b (take b)
    uint32_t so = (ge_o >> 8) * 0xFFu;
    uint32_t e = (ae & ~se) | (be & se);
    uint32_t o = (ao & ~so) | (bo & so);
    return e | (o << 8);
}

// Byte permutation of one word: out byte i = in byte (i ^ m), m in 0..3.
__device__ __forceinline__ uint32_t xor_bytes4(uint32_t v, unsigned m) {
    if (m & 1) v = ((v & 0x00FF00FFu) << 8) | ((v >> 8) & 0x00FF00FFu);
    if (m & 2) v = (v << 16) | (v >> 16);
    return v;
}

// 16 bytes as four words: out byte i = in byte (i ^ m), m in 0..15.
__device__ __forceinline__ uint4 xor_bytes16(uint4 v, unsigned m) {
    uint32_t d[4] = {v.x, v.y, v.z, v.w};
    unsigned q = m >> 2, r = m & 3;
    uint4 o;
    o.x = xor_bytes4(d[0 ^ q], r);
    o.y = xor_bytes4(d[1 ^ q], r);
    o.z = xor_bytes4(d[2 ^ q], r);
    o.w = xor_bytes4(d[3 ^ q], r);
    return o;
}

__device__ __forceinline__ uint4 sat_add16(uint4 x, uint32_t w) {
    return make_uint4(sat_add4(x.x, w), sat_add4(x.y, w), sat_add4(x.z, w), sat_add4(x.w, w));
}

__device__ __forceinline__ uint4 min16(uint4 a, uint4 b) {
    return make_uint4(min4(a.x, b.x), min4(a.y, b.y), min4(a.z, b.z), min4(a.w, b.w));
}

__device__ __forceinline__ uint8_t sat_add1(uint8_t a, uint32_t w) {
    uint32_t s = (uint32_t)a + w;
    return (uint8_t)(s > 255u ? 255u : s);
}

// Launchers of the per-op dense kernels (tdp_dense.hip), shared by the plan executors.
int launch_update(uint8_t *T, uint64_t n, uint64_t h, uint32_t w, hipStream_t st);
int launch_expand(uint8_t *T, uint64_t n, uint64_t hl, uint32_t w, hipStream_t st);
int launch_fill(uint8_t *T, uint64_t n, uint8_t v, hipStream_t st);
int launch_select(uint8_t *dst, const uint8_t *src, uint64_t n, int pos, int target, uint64_t *n_out,
                  hipStream_t st);
int run_ops_range(const uint64_t *ops, int64_t first, int64_t last, uint8_t *&cur, uint8_t *&other, uint64_t &n,
                  int peak_bits, hipStream_t st);

}  // namespace tdp
