// Dense min-plus trellis DP on gfx950: one streaming kernel per op (update, expand-update, select, fill, hist)
// and a host-side plan executor that launches them on one stream.
//
// Table: uint8 costs, INF = 255, n = 2^b entries in HBM.  Every kernel moves 16-byte chunks (one dwordx4 per lane,
// 1 KiB per wave, fully coalesced).  An update with mask h pairs s with s ^ h; with hb = top bit of h:
//   hb >= 4: chunk x (bit hb clear) pairs with chunk (x ^ (hl >> 4)) | half, and inside the pair byte i of one
//            chunk meets byte i ^ (hl & 15) of the other -- a register byte permutation;
//   hb < 4:  both members of every pair lie in the same chunk.
// Semantics follow cdist/viterbi.c:72-219 (saturating min-plus butterflies on old values).
#include "tdp_common.h"

#include <vector>

namespace tdp {

static __thread char g_err[512];

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

constexpr int kBlock = 256;

static inline unsigned grid_for(uint64_t items) {
    uint64_t g = (items + kBlock - 1) / kBlock;
    return (unsigned)(g ? g : 1);
}

// ------------------------------------------------------------------------------------------------- kernels
// hb >= 4: one thread per chunk pair.  c enumerates the chunks with bit (hb - 4) of the chunk index clear.
__global__ __launch_bounds__(kBlock) void k_update_hi(uint8_t *__restrict__ T, uint64_t npairs, int hb4,
                                                      uint64_t hl_chunk, unsigned m, uint32_t w) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= npairs) return;
    uint64_t low = c & ((1ull << hb4) - 1);
    uint64_t xi = ((c >> hb4) << (hb4 + 1)) | low;
    uint64_t yi = (xi ^ hl_chunk) | (1ull << hb4);
    uint4 *P = reinterpret_cast<uint4 *>(T);
    uint4 x = P[xi], y = P[yi];
    uint4 xp = xor_bytes16(x, m), yp = xor_bytes16(y, m);
    P[xi] = min16(x, sat_add16(yp, w));
    P[yi] = min16(y, sat_add16(xp, w));
}

// hb < 4: one thread per chunk; the pair partner of byte i is byte i ^ h of the same chunk.
__global__ __launch_bounds__(kBlock) void k_update_lo(uint8_t *__restrict__ T, uint64_t nchunks, unsigned h,
                                                      uint32_t w) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nchunks) return;
    uint4 *P = reinterpret_cast<uint4 *>(T);
    uint4 x = P[c];
    P[c] = min16(x, sat_add16(xor_bytes16(x, h), w));
}

// Tables of fewer than 16 entries: one thread, scalar pairs.
__global__ void k_update_tiny(uint8_t *T, uint32_t n, uint32_t h, uint32_t w) {
    uint32_t hb = 1u << (31 - __builtin_clz(h));
    for (uint32_t s = 0; s < n; s++) {
        if (s & hb) continue;
        uint32_t t = s ^ h;
        uint8_t a = T[s], b = T[t];
        uint8_t a1 = sat_add1(a, w), b1 = sat_add1(b, w);
        T[s] = a < b1 ? a : b1;
        T[t] = b < a1 ? b : a1;
    }
}

// U[s ^ hl] = T[s] + w, U = T + n: chunk c of T lands on chunk c ^ (hl >> 4) of U, permuted by hl & 15.
__global__ __launch_bounds__(kBlock) void k_expand(uint8_t *__restrict__ T, uint64_t nchunks, uint64_t hl_chunk,
                                                   unsigned m, uint32_t w) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nchunks) return;
    const uint4 *S = reinterpret_cast<const uint4 *>(T);
    uint4 *U = reinterpret_cast<uint4 *>(T) + nchunks;
    U[c ^ hl_chunk] = sat_add16(xor_bytes16(S[c], m), w);
}

__global__ void k_expand_tiny(uint8_t *T, uint32_t n, uint32_t hl, uint32_t w) {
    for (uint32_t s = 0; s < n; s++) T[n + (s ^ hl)] = sat_add1(T[s], w);
}

// dst chunk c <- src chunk with bit (pos - 4) inserted = target (pos >= 4).
__global__ __launch_bounds__(kBlock) void k_select_hi(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src,
                                                      uint64_t nchunks, int pos4, uint32_t target) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nchunks) return;
    uint64_t low = c & ((1ull << pos4) - 1);
    uint64_t si = ((c >> pos4) << (pos4 + 1)) | ((uint64_t)target << pos4) | low;
    reinterpret_cast<uint4 *>(dst)[c] = reinterpret_cast<const uint4 *>(src)[si];
}

// pos < 4: dst chunk c gathers the 16 bytes with bit pos == target out of src chunks 2c and 2c + 1.
__global__ __launch_bounds__(kBlock) void k_select_lo(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src,
                                                      uint64_t nchunks, int pos, uint32_t target) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= nchunks) return;
    const uint4 *S = reinterpret_cast<const uint4 *>(src);
    uint4 a = S[2 * c], b = S[2 * c + 1];
    uint32_t in[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t out[4] = {0, 0, 0, 0};
    uint32_t lowmask = (1u << pos) - 1;
#pragma unroll
    for (uint32_t k = 0; k < 16; k++) {
        uint32_t s = ((k & ~lowmask) << 1) | (target << pos) | (k & lowmask);   // 0..31
        uint32_t byte = (in[s >> 2] >> (8 * (s & 3))) & 0xFFu;
        out[k >> 2] |= byte << (8 * (k & 3));
    }
    reinterpret_cast<uint4 *>(dst)[c] = make_uint4(out[0], out[1], out[2], out[3]);
}

__global__ void k_select_tiny(uint8_t *dst, const uint8_t *src, uint32_t half, int pos, uint32_t target) {
    uint32_t lowmask = (1u << pos) - 1;
    for (uint32_t k = 0; k < half; k++) dst[k] = src[((k & ~lowmask) << 1) | (target << pos) | (k & lowmask)];
}

__global__ __launch_bounds__(kBlock) void k_fill(uint8_t *__restrict__ T, uint64_t nchunks, uint32_t v4) {
    uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c < nchunks) reinterpret_cast<uint4 *>(T)[c] = make_uint4(v4, v4, v4, v4);
}

__global__ void k_fill_tiny(uint8_t *T, uint32_t n, uint8_t v) {
    for (uint32_t i = 0; i < n; i++) T[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_hist(const uint8_t *__restrict__ T, uint64_t n,
                                                 unsigned long long *__restrict__ hist) {
    __shared__ uint32_t bins[256];
    bins[threadIdx.x] = 0;
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) atomicAdd(&bins[T[i]], 1u);
    __syncthreads();
    if (bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
}

__global__ void k_zero_u64(unsigned long long *p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------------------------------------- small plans
// A whole plan of at most 2^kSmallBits entries in one workgroup, its two tables in LDS, with a barrier between
// steps.  The host groups the op stream into steps: an optional expand, a run of up to kSmallRun updates whose
// masks are linearly independent over GF(2), and an optional select.  A step's updates pair an entry with its
// translates by the masks, so the table falls into cosets of their span, 2^r entries each.  One thread loads a
// coset, applies the run's updates to it in op order, and stores it.
//   - The coset of representative s is {s ^ X_c}, with X_c the XOR of the masks picked by the bits of c.
//   - Update u pairs c with c ^ 2^u.
//   - The representatives are the indices with zeros at the pivot bits of the span's echelon basis (each nonzero
//     element of the span has its top bit at a pivot, so the cosets partition the table).
//   - An expand is folded in by reading the new half as 255.
//   - A select is folded in by storing only the kept entries, compacted, into the other table.
// One launch and about a third as many barriers as ops replace a launch per op (a few hundred for the
// reference's CPU case).  A step is 16 bytes, staged in LDS a chunk at a time, so no step waits on a global load:
//   word 0: expand | select << 1 | position << 2 | target << 6 | r << 7 | pivots (4 bits each, ascending) << 10
//   words 1-2: the run's masks, 16 bits each; word 3: the run's weights, 8 bits each (saturated at 255).
constexpr int kSmallBits = 14;
constexpr int kSmallRun = 4;
constexpr int kSmallBlock = 1024;
constexpr int kSmallSteps = 1024;       // steps per LDS chunk (16 KB)

template <int R, int BLOCK>
__device__ __forceinline__ void small_step(const uint8_t *cur, uint8_t *dst, uint32_t n, uint32_t nold, uint4 st) {
    const bool E = st.x & 1, S = (st.x >> 1) & 1;
    const uint32_t pos = (st.x >> 2) & 15, tb = (st.x >> 6) & 1;
    const uint32_t low = (1u << pos) - 1;
    uint32_t X[1 << R], w[R > 0 ? R : 1];
    X[0] = 0;
#pragma unroll
    for (int c = 1; c < (1 << R); c++) {
        const int q = __builtin_ctz(c);
        const uint32_t a = ((q < 2 ? st.y : st.z) >> (16 * (q & 1))) & 0xFFFF;
        X[c] = X[c & (c - 1)] ^ a;
    }
#pragma unroll
    for (int u = 0; u < R; u++) w[u] = (st.w >> (8 * u)) & 255;
    for (uint32_t g = threadIdx.x; g < (n >> R); g += BLOCK) {
        uint32_t s = g;
#pragma unroll
        for (int q = 0; q < R; q++) {
            const uint32_t p = (st.x >> (10 + 4 * q)) & 15;
            s = ((s >> p) << (p + 1)) | (s & ((1u << p) - 1));
        }
        uint32_t v[1 << R];
#pragma unroll
        for (int c = 0; c < (1 << R); c++) {
            const uint32_t idx = s ^ X[c];
            v[c] = (E && idx >= nold) ? 255u : cur[idx];
        }
#pragma unroll
        for (int u = 0; u < R; u++) {
#pragma unroll
            for (int c = 0; c < (1 << R); c++) {       // no clamp: min(x, y + w) <= x <= 255
                if (c & (1 << u)) continue;
                const uint32_t x = v[c], y = v[c | (1 << u)];
                v[c] = min(x, y + w[u]);
                v[c | (1 << u)] = min(y, x + w[u]);
            }
        }
#pragma unroll
        for (int c = 0; c < (1 << R); c++) {
            const uint32_t idx = s ^ X[c];
            if (!S) dst[idx] = (uint8_t)v[c];
            else if (((idx >> pos) & 1) == tb) dst[(idx & low) | ((idx >> (pos + 1)) << pos)] = (uint8_t)v[c];
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_plan_small(const uint4 *__restrict__ steps, int64_t nsteps,
                                                      uint8_t *value) {
    __shared__ uint8_t tab[2][1 << kSmallBits];
    __shared__ uint4 lsteps[kSmallSteps];
    uint8_t *cur = tab[0], *oth = tab[1];
    uint32_t n = 1;
    if (threadIdx.x == 0) cur[0] = 0;
    for (int64_t base = 0; base < nsteps; base += kSmallSteps) {
        const int m = (int)(nsteps - base < kSmallSteps ? nsteps - base : kSmallSteps);
        __syncthreads();                                    // the previous chunk's steps are all read
        for (int j = threadIdx.x; j < m; j += BLOCK) lsteps[j] = steps[base + j];
        __syncthreads();
        for (int j = 0; j < m; j++) {
            const uint4 raw = lsteps[j];
            const uint4 st = make_uint4(__builtin_amdgcn_readfirstlane(raw.x), __builtin_amdgcn_readfirstlane(raw.y),
                                        __builtin_amdgcn_readfirstlane(raw.z), __builtin_amdgcn_readfirstlane(raw.w));
            const uint32_t nold = n;
            if (st.x & 1) n <<= 1;
            uint8_t *dst = ((st.x >> 1) & 1) ? oth : cur;
            switch ((st.x >> 7) & 7) {
            case 0: small_step<0, BLOCK>(cur, dst, n, nold, st); break;
            case 1: small_step<1, BLOCK>(cur, dst, n, nold, st); break;
            case 2: small_step<2, BLOCK>(cur, dst, n, nold, st); break;
            case 3: small_step<3, BLOCK>(cur, dst, n, nold, st); break;
            default: small_step<4, BLOCK>(cur, dst, n, nold, st); break;
            }
            if ((st.x >> 1) & 1) {
                oth = cur;
                cur = dst;
                n >>= 1;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *value = cur[0];
}

// ------------------------------------------------------------------------------------------------- launchers
int launch_update(uint8_t *T, uint64_t n, uint64_t h, uint32_t w, hipStream_t st) {
    if (h == 0 || n < 2) return TDP_OK;
    TDP_REQUIRE(h < n, "update mask 0x%llx outside a table of %llu entries", (unsigned long long)h,
                (unsigned long long)n);
    if (w > 255) w = 255;
    int hb = top_bit(h);
    if (n < 16) {
        k_update_tiny<<<1, 1, 0, st>>>(T, (uint32_t)n, (uint32_t)h, w);
    } else if (hb >= 4) {
        uint64_t hl = h ^ (1ull << hb);
        uint64_t npairs = n >> 5;
        k_update_hi<<<grid_for(npairs), kBlock, 0, st>>>(T, npairs, hb - 4, hl >> 4, (unsigned)(hl & 15), w);
    } else {
        uint64_t nchunks = n >> 4;
        k_update_lo<<<grid_for(nchunks), kBlock, 0, st>>>(T, nchunks, (unsigned)h, w);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int launch_expand(uint8_t *T, uint64_t n, uint64_t hl, uint32_t w, hipStream_t st) {
    TDP_REQUIRE(hl < n, "expand mask 0x%llx outside a table of %llu entries", (unsigned long long)hl,
                (unsigned long long)n);
    if (w > 255) w = 255;
    if (n < 16) {
        k_expand_tiny<<<1, 1, 0, st>>>(T, (uint32_t)n, (uint32_t)hl, w);
    } else {
        uint64_t nchunks = n >> 4;
        k_expand<<<grid_for(nchunks), kBlock, 0, st>>>(T, nchunks, hl >> 4, (unsigned)(hl & 15), w);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int launch_fill(uint8_t *T, uint64_t n, uint8_t v, hipStream_t st) {
    if (n == 0) return TDP_OK;
    if (n < 16 || (n & 15)) {
        TDP_REQUIRE(n < (1ull << 32), "fill of %llu entries must be a multiple of 16", (unsigned long long)n);
        k_fill_tiny<<<1, 1, 0, st>>>(T, (uint32_t)n, v);
    } else {
        k_fill<<<grid_for(n >> 4), kBlock, 0, st>>>(T, n >> 4, 0x01010101u * v);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int launch_select(uint8_t *dst, const uint8_t *src, uint64_t n, int pos, int target, uint64_t *n_out,
                         hipStream_t st) {
    TDP_REQUIRE(pos >= 0 && pos < 63 && (target == 0 || target == 1), "select pos %d target %d", pos, target);
    if ((n >> (pos + 1)) == 0) {   // the reference returns n unchanged (cdist/viterbi.c:157-158)
        *n_out = n;
        if (dst != src) TDP_HIP_CHECK(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, st));
        return TDP_OK;
    }
    uint64_t half = n >> 1;
    if (half < 16) {
        k_select_tiny<<<1, 1, 0, st>>>(dst, src, (uint32_t)half, pos, (uint32_t)target);
    } else if (pos >= 4) {
        k_select_hi<<<grid_for(half >> 4), kBlock, 0, st>>>(dst, src, half >> 4, pos - 4, (uint32_t)target);
    } else {
        k_select_lo<<<grid_for(half >> 4), kBlock, 0, st>>>(dst, src, half >> 4, pos, (uint32_t)target);
    }
    TDP_LAUNCH_CHECK();
    *n_out = half;
    return TDP_OK;
}

static bool pow2(uint64_t n) { return n && !(n & (n - 1)); }

}  // namespace tdp

using namespace tdp;

// Ops [first, last) of an encoded stream on the table cur (n entries, at most 2^peak_bits), selects writing to
// other and swapping the two; an expand directly followed by an update of its new row is one kernel.
int tdp::run_ops_range(const uint64_t *ops, int64_t first, int64_t last, uint8_t *&cur, uint8_t *&other,
                       uint64_t &n, int peak_bits, hipStream_t st) {
    const uint64_t cap = 1ull << peak_bits;
    int rc = TDP_OK;
    for (int64_t i = first; i < last; i++) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (code == 0) {
            TDP_REQUIRE(2 * n <= cap, "op %lld: expand beyond 2^%d entries", (long long)i, peak_bits);
            const uint64_t *nx = ops + 3 * (i + 1);
            if (i + 1 < last && nx[0] == 1 && (nx[1] >> top_bit(n)) == 1) {
                TDP_REQUIRE(nx[2] >= 1, "op %lld: weight < 1", (long long)(i + 1));
                rc = launch_expand(cur, n, nx[1] ^ n, (uint32_t)(nx[2] > 255 ? 255 : nx[2]), st);
                i++;
            } else {
                rc = launch_fill(cur + n, n, 255, st);
            }
            n <<= 1;
        } else if (code == 1) {
            TDP_REQUIRE(a < n, "op %lld: update mask 0x%llx outside 2^%d entries", (long long)i,
                        (unsigned long long)a, top_bit(n));
            TDP_REQUIRE(b >= 1, "op %lld: weight < 1", (long long)i);
            rc = launch_update(cur, n, a, (uint32_t)(b > 255 ? 255 : b), st);
        } else if (code == 2) {
            TDP_REQUIRE(n >= 2 && a < (uint64_t)top_bit(n) && b <= 1, "op %lld: select pos %llu on 2^%d entries",
                        (long long)i, (unsigned long long)a, top_bit(n));
            uint64_t nn = 0;
            rc = launch_select(other, cur, n, (int)a, (int)b, &nn, st);
            uint8_t *t = cur;
            cur = other;
            other = t;
            n = nn;
        } else {
            TDP_REQUIRE(false, "op %lld: unknown code %llu", (long long)i, (unsigned long long)code);
        }
        if (rc) return rc;
    }
    return TDP_OK;
}

extern "C" {

const char *tdp_last_error(void) { return g_err; }

const char *tdp_version(void) { return "tropidist_amd 0.1 gfx950"; }

int tdp_read_back(void *host, const void *d_src, uint64_t bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(host && d_src, "tdp_read_back: null pointer");
    TDP_HIP_CHECK(hipMemcpyAsync(host, d_src, bytes, hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    return TDP_OK;
}

int tdp_num_cus(int *cus) {
    int dev = 0;
    TDP_HIP_CHECK(hipGetDevice(&dev));
    TDP_HIP_CHECK(hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev));
    return TDP_OK;
}

int tdp_update(uint8_t *d_T, uint64_t n, uint64_t h, void *stream) {
    TDP_REQUIRE(pow2(n), "table size %llu is not a power of two", (unsigned long long)n);
    return launch_update(d_T, n, h, 1, as_stream(stream));
}

int tdp_update_w(uint8_t *d_T, uint64_t n, uint64_t h, int w, void *stream) {
    TDP_REQUIRE(pow2(n), "table size %llu is not a power of two", (unsigned long long)n);
    TDP_REQUIRE(w >= 1, "weight %d < 1", w);
    return launch_update(d_T, n, h, (uint32_t)w, as_stream(stream));
}

int tdp_expand_update(uint8_t *d_T, uint64_t n, uint64_t hl, int w, void *stream) {
    TDP_REQUIRE(pow2(n), "table size %llu is not a power of two", (unsigned long long)n);
    TDP_REQUIRE(w >= 1, "weight %d < 1", w);
    return launch_expand(d_T, n, hl, (uint32_t)w, as_stream(stream));
}

int tdp_select(uint8_t *d_dst, const uint8_t *d_src, uint64_t n, int pos, int target, uint64_t *n_out,
               void *stream) {
    TDP_REQUIRE(pow2(n), "table size %llu is not a power of two", (unsigned long long)n);
    TDP_REQUIRE(d_dst + n / 2 <= d_src || d_src + n <= d_dst, "select source and destination overlap");
    return launch_select(d_dst, d_src, n, pos, target, n_out, as_stream(stream));
}

int tdp_fill(uint8_t *d_T, uint64_t n, uint8_t v, void *stream) { return launch_fill(d_T, n, v, as_stream(stream)); }

int tdp_hist(const uint8_t *d_T, uint64_t n, uint64_t *d_hist, void *stream) {
    hipStream_t st = as_stream(stream);
    k_zero_u64<<<1, 256, 0, st>>>(reinterpret_cast<unsigned long long *>(d_hist), 256);
    TDP_LAUNCH_CHECK();
    if (n == 0) return TDP_OK;
    uint64_t blocks = (n + kBlock * 16 - 1) / (kBlock * 16);
    if (blocks > 4096) blocks = 4096;
    k_hist<<<(unsigned)blocks, kBlock, 0, st>>>(d_T, n, reinterpret_cast<unsigned long long *>(d_hist));
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_plan_small_bits(void) { return kSmallBits; }

int tdp_plan_small_bytes(int64_t nops, uint64_t *bytes) {
    TDP_REQUIRE(nops >= 0 && bytes, "bad op count");
    *bytes = ((uint64_t)nops * 16 + 255) / 256 * 256 + 256;       // at most one step per op, then the result
    return TDP_OK;
}

int tdp_run_plan_small(const uint64_t *ops, int64_t nops, int peak_bits, void *d_work, uint64_t work_bytes,
                       int *value, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(peak_bits >= 0 && peak_bits <= kSmallBits, "a small plan holds at most 2^%d entries (2^%d asked)",
                kSmallBits, peak_bits);
    TDP_REQUIRE(nops >= 0 && (nops == 0 || ops), "bad op stream");
    uint64_t need = 0;
    int rc = tdp_plan_small_bytes(nops, &need);
    if (rc) return rc;
    TDP_REQUIRE(d_work && work_bytes >= need, "small-plan workspace of %llu bytes, %llu needed",
                (unsigned long long)work_bytes, (unsigned long long)need);
    // The checks of the per-op executor, on the host, before anything is launched, while the ops are grouped
    // into steps.  A run of updates ends at the first mask in the span of the run's earlier masks (a mask of 0
    // updates nothing and is dropped).
    const uint64_t cap = 1ull << peak_bits;
    std::vector<uint32_t> steps;
    steps.reserve((size_t)nops * 4);
    uint64_t n = 1;
    int64_t i = 0;
    auto bad_op = [&](int64_t k) -> int {
        const uint64_t code = ops[3 * k], a = ops[3 * k + 1], b = ops[3 * k + 2];
        if (code == 0) {
            TDP_REQUIRE(2 * n <= cap, "op %lld: expand beyond 2^%d entries", (long long)k, peak_bits);
        } else if (code == 1) {
            TDP_REQUIRE(a < n && b >= 1, "op %lld: bad update", (long long)k);
        } else if (code == 2) {
            TDP_REQUIRE(n >= 2 && a < (uint64_t)top_bit(n) && b <= 1, "op %lld: bad select", (long long)k);
        } else {
            TDP_REQUIRE(false, "op %lld: unknown code %llu", (long long)k, (unsigned long long)code);
        }
        return TDP_OK;
    };
    while (i < nops) {
        uint32_t w0 = 0, masks[kSmallRun] = {0, 0, 0, 0}, wts = 0;
        uint32_t basis[kSmallRun];
        int r = 0;
        if (ops[3 * i] == 0) {
            if ((rc = bad_op(i))) return rc;
            w0 |= 1;
            n <<= 1;
            i++;
        }
        // at least 256 cosets a step: a longer run on a small table leaves a few waves with a long serial chain
        const int run = top_bit(n) - 8 < 1 ? 1 : top_bit(n) - 8 > kSmallRun ? kSmallRun : top_bit(n) - 8;
        while (i < nops && ops[3 * i] == 1 && r < run) {
            if ((rc = bad_op(i))) return rc;
            const uint32_t a = (uint32_t)ops[3 * i + 1];
            uint32_t x = a;
            for (int q = 0; q < r; q++)                 // basis kept with distinct top bits, highest first
                if ((x >> top_bit(basis[q])) & 1) x ^= basis[q];
            if (a != 0 && x == 0) break;                // in the span: the next step's
            if (a != 0) {
                int q = r;
                while (q > 0 && top_bit(basis[q - 1]) < top_bit(x)) {
                    basis[q] = basis[q - 1];
                    q--;
                }
                basis[q] = x;
                masks[r] = a;
                wts |= (uint32_t)(ops[3 * i + 2] > 255 ? 255 : ops[3 * i + 2]) << (8 * r);
                r++;
            }
            i++;
        }
        if (i < nops && ops[3 * i] == 2) {
            if ((rc = bad_op(i))) return rc;
            w0 |= 2 | (uint32_t)ops[3 * i + 1] << 2 | (uint32_t)ops[3 * i + 2] << 6;
            i++;
        } else if (i < nops && ops[3 * i] != 0 && ops[3 * i] != 1) {
            return bad_op(i);
        }
        for (int q = 0; q < r; q++) w0 |= (uint32_t)top_bit(basis[r - 1 - q]) << (10 + 4 * q);    // ascending
        w0 |= (uint32_t)r << 7;
        if (w0 & 2) n >>= 1;
        if (w0 == 0) continue;                          // updates of mask 0 only
        steps.push_back(w0);
        steps.push_back(masks[0] | masks[1] << 16);
        steps.push_back(masks[2] | masks[3] << 16);
        steps.push_back(wts);
    }
    TDP_REQUIRE(n == 1, "op stream ends with %llu entries, not 1", (unsigned long long)n);
    const int64_t nsteps = (int64_t)steps.size() / 4;
    uint8_t *w = static_cast<uint8_t *>(d_work);
    const uint64_t off_val = need - 256;
    if (nsteps) TDP_HIP_CHECK(hipMemcpyAsync(w, steps.data(), (size_t)nsteps * 16, hipMemcpyHostToDevice, st));
    const uint4 *d_steps = reinterpret_cast<const uint4 *>(w);
    k_plan_small<kSmallBlock><<<1, kSmallBlock, 0, st>>>(d_steps, nsteps, w + off_val);
    TDP_LAUNCH_CHECK();
    uint8_t v = 0;
    TDP_HIP_CHECK(hipMemcpyAsync(&v, w + off_val, 1, hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    *value = v;
    return TDP_OK;
}

int tdp_run_plan(const uint64_t *ops, int64_t nops, int peak_bits, uint8_t *d_T0, uint8_t *d_T1, int *value,
                 void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(peak_bits >= 0 && peak_bits <= 40, "peak_bits %d out of range", peak_bits);
    TDP_REQUIRE(nops >= 0 && (nops == 0 || ops), "bad op stream");
    uint8_t *cur = d_T0, *other = d_T1;
    uint64_t n = 1;
    int rc = launch_fill(cur, 1, 0, st);
    if (rc) return rc;
    rc = run_ops_range(ops, 0, nops, cur, other, n, peak_bits, st);
    if (rc) return rc;
    TDP_REQUIRE(n == 1, "op stream ends with %llu entries, not 1", (unsigned long long)n);
    uint8_t v = 0;
    TDP_HIP_CHECK(hipMemcpyAsync(&v, cur, 1, hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    *value = v;
    return TDP_OK;
}

}  // extern "C"
