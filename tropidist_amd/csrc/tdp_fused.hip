// Fused-pass executor of the dense trellis DP on gfx950 (the pass layout is compiled on the host by
// tropidist_amd/fused.py; semantics are those of the op stream, cdist/kernel.py:94-138 over cdist/viterbi.c).
//
// One launch = one pass: D update directions (plus updates inside dwords) and the expands / selects of their
// columns, applied to register-resident cubes.  Thread t owns the cube of representative x(t); cube element q
// (q < 2^D) is the dword at input index base_in(t) ^ XOR_{j in q} din[2 + j].
//
// Twisted lanes: register entry k of element q holds the table entry at byte k ^ sigma_q of its dword, where
// sigma_q = XOR_{j in q} lam[j] and lam[j] is the part of update j's mask inside the dword.  The partner of register
// entry k of q along direction j is then register entry k of q ^ (1 << j), so every butterfly is lane-aligned:
// two packed u16 adds and two packed u16 mins per dword pair and 16-bit half (entries stay <= 255, since
// min(x, y + w) <= x, so no saturation is needed).  The twist costs one v_perm per dword on load and on store.
//
// Entries whose new-row coordinate is 1 start at 255 and are not loaded; entries are stored to
// base_out(t) ^ XOR dout[...] when the ending rows' coordinates equal their targets.
#include "tdp_common.h"

namespace tdp {

constexpr int kPassBlock = 256;
constexpr int kMaxRuns = 12;
enum { MAP_IN = 0, MAP_OUT = 1, MAP_NEW = 2, MAP_SEL = 3 };

// A thread's base indices: map m = OR over runs r < nruns[m] of ((tid >> f0) & (2^len - 1)) << sh.
struct PassArgs {
    uint64_t din[8], dout[8];               // per basis vector (e_0, e_1, m'_0 .. m'_{D-1}): index deltas
    uint32_t dnew[8], dsel[8];              // new-row / ending-row coordinate deltas
    uint8_t run_f0[4][kMaxRuns], run_len[4][kMaxRuns], run_sh[4][kMaxRuns];
    int32_t nruns[4];
    uint32_t sel_target;
    int32_t D, nlow, byte_store, b_in, b_out, nfree, pad;
    uint8_t lam[8], w[8], low_lam[8], low_w[8];
};

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 vmin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2 swap16(u16x2 v) { return u16x2{v.y, v.x}; }

// Byte selectors of v_perm_b32 for a twist s (0..3).  Register halves: lo = (r0, r2), hi = (r1, r3) as 16-bit lanes.
// Unpack from a dword d with perm(d, d, sel): lo byte 0 = d byte (0 ^ s), byte 2 = d byte (2 ^ s), zeros between.
__device__ __forceinline__ uint32_t sel_lo(uint32_t s) { return 0x0C000C00u | ((s ^ 2u) << 16) | s; }
__device__ __forceinline__ uint32_t sel_hi(uint32_t s) { return 0x0C000C00u | ((s ^ 3u) << 16) | (s ^ 1u); }
// Pack with perm(hi, lo, sel): pool bytes 0-3 = lo, 4-7 = hi, so r0, r1, r2, r3 sit at pool bytes 0, 4, 2, 6;
// dword byte k takes register entry k ^ s.
__device__ __forceinline__ uint32_t sel_pack(uint32_t s) {
    uint32_t v = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        uint32_t r = k ^ s;
        v |= (((r & 1u) << 2) | (r & 2u)) << (8 * k);
    }
    return v;
}

template <int D, int J>
struct DirLoop {
    static __device__ __forceinline__ void run(u16x2 *lo, u16x2 *hi, const PassArgs &a) {
        const unsigned short w = a.w[J];
        const u16x2 w2 = u16x2{w, w};
#pragma unroll
        for (int q = 0; q < (1 << D); q++) {
            if ((q >> J) & 1) continue;
            const int p = q | (1 << J);
            u16x2 xl = lo[q], xh = hi[q], yl = lo[p], yh = hi[p];
            lo[q] = vmin(xl, yl + w2);
            hi[q] = vmin(xh, yh + w2);
            lo[p] = vmin(yl, xl + w2);
            hi[p] = vmin(yh, xh + w2);
        }
        DirLoop<D, J + 1>::run(lo, hi, a);
    }
};

template <int D>
struct DirLoop<D, D> {
    static __device__ __forceinline__ void run(u16x2 *, u16x2 *, const PassArgs &) {}
};

// An update whose mask lies inside the dword: register entry k pairs with entry k ^ lam in every element.
__device__ __forceinline__ void low_update(u16x2 &l, u16x2 &h, bool swap_halves, bool swap_lanes, u16x2 w2) {
    u16x2 pl = swap_halves ? h : l, ph = swap_halves ? l : h;   // lam bit 0: r0 <-> r1 (lo <-> hi)
    if (swap_lanes) {                                            // lam bit 1: r0 <-> r2 (16-bit lanes)
        pl = swap16(pl);
        ph = swap16(ph);
    }
    l = vmin(l, pl + w2);
    h = vmin(h, ph + w2);
}

// map m of the index bits in `bits` (a subset of the thread index): OR over runs of ((bits >> f0) & mask) << sh.
// The map is linear over disjoint bit sets, so a thread's map is the map of its block's bits (uniform: scalar
// arithmetic) OR the map of its lane bits (only the runs that start below 8 touch them).
template <typename IDX>
__device__ __forceinline__ IDX map_bits(const PassArgs &a, int m, uint64_t bits, bool lane_runs_only) {
    IDX x = 0;
#pragma unroll
    for (int r = 0; r < kMaxRuns; r++) {
        if (r < a.nruns[m] && (!lane_runs_only || a.run_f0[m][r] < 8)) {
            uint64_t mask = (1ull << a.run_len[m][r]) - 1;
            x |= (IDX)(((bits >> a.run_f0[m][r]) & mask) << a.run_sh[m][r]);
        }
    }
    return x;
}

// IDX: uint32_t when both tables hold at most 2^32 entries (byte offsets from a scalar base), else uint64_t.
// FREE_NEW: some new row is a free variable, so whether an element starts at 255 varies across lanes; otherwise it
// depends on the element only and the test is uniform.
template <int D, bool BYTE, bool FREE_NEW, typename IDX>
__device__ __forceinline__ void pass_cube(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                          uint64_t block_bits, uint32_t lane, const PassArgs &a) {
    const IDX bin = map_bits<IDX>(a, MAP_IN, block_bits, false) | map_bits<IDX>(a, MAP_IN, lane, true);
    const IDX bout = map_bits<IDX>(a, MAP_OUT, block_bits, false) | map_bits<IDX>(a, MAP_OUT, lane, true);
    const uint32_t xs = map_bits<uint32_t>(a, MAP_SEL, block_bits, false) | map_bits<uint32_t>(a, MAP_SEL, lane, true);
    uint32_t xn = 0;
    if (FREE_NEW) xn = map_bits<uint32_t>(a, MAP_NEW, block_bits, false) | map_bits<uint32_t>(a, MAP_NEW, lane, true);
    constexpr int Q = 1 << D;
    u16x2 lo[Q], hi[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        IDX ai = bin;
        uint32_t nz = xn, s = 0;
#pragma unroll
        for (int j = 0; j < D; j++)
            if ((q >> j) & 1) {
                ai ^= (IDX)a.din[2 + j];
                nz ^= a.dnew[2 + j];
                s ^= a.lam[j];
            }
        uint32_t d = 0xFFFFFFFFu;
        if (!nz) d = *reinterpret_cast<const uint32_t *>(in + ai);
        lo[q] = as_u16x2(__builtin_amdgcn_perm(d, d, sel_lo(s)));
        hi[q] = as_u16x2(__builtin_amdgcn_perm(d, d, sel_hi(s)));
    }
    DirLoop<D, 0>::run(lo, hi, a);
#pragma unroll
    for (int t = 0; t < 8; t++) {                  // constant indices keep the descriptor in kernel arguments
        if (t >= a.nlow) break;
        const unsigned short w = a.low_w[t];
        const bool sh = a.low_lam[t] & 1, sl = (a.low_lam[t] >> 1) & 1;
#pragma unroll
        for (int q = 0; q < Q; q++) low_update(lo[q], hi[q], sh, sl, u16x2{w, w});
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        IDX ao = bout;
        uint32_t sel = xs, s = 0;
#pragma unroll
        for (int j = 0; j < D; j++)
            if ((q >> j) & 1) {
                ao ^= (IDX)a.dout[2 + j];
                sel ^= a.dsel[2 + j];
                s ^= a.lam[j];
            }
        if (!BYTE) {
            if (sel == a.sel_target)
                *reinterpret_cast<uint32_t *>(out + ao) = __builtin_amdgcn_perm(as_u32(hi[q]), as_u32(lo[q]), sel_pack(s));
        } else {
            const unsigned short r[4] = {lo[q].x, hi[q].x, lo[q].y, hi[q].y};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t m = (uint32_t)k ^ s;           // byte of the dword this register entry came from
                IDX ak = ao ^ (IDX)((m & 1) ? a.dout[0] : 0) ^ (IDX)((m & 2) ? a.dout[1] : 0);
                uint32_t sk = sel ^ ((m & 1) ? a.dsel[0] : 0) ^ ((m & 2) ? a.dsel[1] : 0);
                if (sk == a.sel_target) out[ak] = (uint8_t)r[k];
            }
        }
    }
}

template <int D, bool BYTE, bool FREE_NEW, typename IDX>
__global__ __launch_bounds__(kPassBlock) void k_pass(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                     uint64_t nthreads, const PassArgs a) {
    const uint64_t block = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (block * kPassBlock + threadIdx.x < nthreads)
        pass_cube<D, BYTE, FREE_NEW, IDX>(in, out, block * kPassBlock, threadIdx.x, a);
}

template <int D, bool BYTE, bool FREE_NEW>
static void launch_pass_v(const uint8_t *in, uint8_t *out, uint64_t nthreads, const PassArgs &a, dim3 grid,
                          hipStream_t st) {
    if (a.b_in <= 32 && a.b_out <= 32)
        k_pass<D, BYTE, FREE_NEW, uint32_t><<<grid, kPassBlock, 0, st>>>(in, out, nthreads, a);
    else
        k_pass<D, BYTE, FREE_NEW, uint64_t><<<grid, kPassBlock, 0, st>>>(in, out, nthreads, a);
}

template <int D>
static void launch_pass_d(const uint8_t *in, uint8_t *out, uint64_t nthreads, const PassArgs &a, hipStream_t st) {
    // one thread per cube: blocks beyond 2^20 go to the y dimension (grid.x * 256 stays below 2^32)
    uint64_t blocks = (nthreads + kPassBlock - 1) / kPassBlock;
    dim3 grid((unsigned)(blocks < (1ull << 20) ? blocks : (1ull << 20)), (unsigned)((blocks + (1ull << 20) - 1) >> 20));
    const bool free_new = a.nruns[MAP_NEW] > 0;
    if (a.byte_store) {
        if (free_new) launch_pass_v<D, true, true>(in, out, nthreads, a, grid, st);
        else launch_pass_v<D, true, false>(in, out, nthreads, a, grid, st);
    } else {
        if (free_new) launch_pass_v<D, false, true>(in, out, nthreads, a, grid, st);
        else launch_pass_v<D, false, false>(in, out, nthreads, a, grid, st);
    }
}

// Host-side validation of a descriptor: every address a thread forms is an XOR / OR of the deltas and runs below,
// so their unions bound it.
static int check_pass(const PassArgs &a) {
    TDP_REQUIRE(a.D >= 0 && a.D <= 5 && a.nlow >= 0 && a.nlow <= 8 && a.nfree >= 0 && a.nfree <= 48,
                "pass with D %d, %d low updates, %d free variables", a.D, a.nlow, a.nfree);
    TDP_REQUIRE(a.b_in >= 2 && a.b_in <= 40 && a.b_out >= 0 && a.b_out <= 40, "pass from 2^%d to 2^%d entries",
                a.b_in, a.b_out);
    uint64_t u[4] = {0, 0, 0, 0};
    for (int m = 0; m < 4; m++) {
        TDP_REQUIRE(a.nruns[m] >= 0 && a.nruns[m] <= kMaxRuns, "pass index map %d: %d runs", m, a.nruns[m]);
        for (int r = 0; r < a.nruns[m]; r++) {
            int f0 = a.run_f0[m][r], len = a.run_len[m][r], sh = a.run_sh[m][r];
            TDP_REQUIRE(len >= 1 && f0 + len <= a.nfree && sh + len <= (m < 2 ? 40 : 32),
                        "pass index map %d run %d out of range", m, r);
            u[m] |= ((1ull << len) - 1) << sh;
        }
    }
    uint64_t uin = u[MAP_IN], uout = u[MAP_OUT];
    for (int k = 2; k < a.D + 2; k++) {
        uin |= a.din[k];
        uout |= a.dout[k];
    }
    TDP_REQUIRE((uin | 3) < (1ull << a.b_in) && (uin & 3) == 0 && a.din[0] == 1 && a.din[1] == 2,
                "pass input addresses outside its table or off the dword lanes");
    TDP_REQUIRE((uout | a.dout[0] | a.dout[1]) < (1ull << a.b_out), "pass output addresses outside its table");
    TDP_REQUIRE(a.byte_store || (a.b_out >= 2 && a.dout[0] == 1 && a.dout[1] == 2 && a.dsel[0] == 0 &&
                                 a.dsel[1] == 0 && (uout & 3) == 0),
                "dword stores need the two low output bits on the dword lanes");
    for (int j = 0; j < a.D; j++) TDP_REQUIRE(a.lam[j] < 4 && a.w[j] >= 1, "pass update %d: bad lane mask/weight", j);
    for (int t = 0; t < a.nlow; t++)
        TDP_REQUIRE(a.low_lam[t] >= 1 && a.low_lam[t] < 4 && a.low_w[t] >= 1, "pass low update %d: bad mask", t);
    return TDP_OK;
}

int launch_pass(const uint8_t *in, uint8_t *out, const PassArgs &a, hipStream_t st) {
    int rc = check_pass(a);
    if (rc) return rc;
    uint64_t nthreads = 1ull << a.nfree;
    switch (a.D) {
        case 0: launch_pass_d<0>(in, out, nthreads, a, st); break;
        case 1: launch_pass_d<1>(in, out, nthreads, a, st); break;
        case 2: launch_pass_d<2>(in, out, nthreads, a, st); break;
        case 3: launch_pass_d<3>(in, out, nthreads, a, st); break;
        case 4: launch_pass_d<4>(in, out, nthreads, a, st); break;
        default: launch_pass_d<5>(in, out, nthreads, a, st); break;
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_pass_args_bytes(void) { return (int)sizeof(PassArgs); }

int tdp_run_pass(const uint8_t *d_in, uint8_t *d_out, const void *pass, void *stream) {
    TDP_REQUIRE(d_in && d_out && pass && d_in != d_out, "bad pass arguments");
    return launch_pass(d_in, d_out, *static_cast<const PassArgs *>(pass), as_stream(stream));
}

int tdp_run_program(const uint64_t *ops, int64_t nops, const int64_t *segments, int64_t nseg, const void *passes,
                    int64_t npasses, int peak_bits, uint8_t *d_T0, uint8_t *d_T1, int *value, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(peak_bits >= 0 && peak_bits <= 40 && nseg >= 0 && npasses >= 0, "bad program");
    const PassArgs *P = static_cast<const PassArgs *>(passes);
    const uint64_t cap = 1ull << peak_bits;
    uint8_t *cur = d_T0, *other = d_T1;
    uint64_t n = 1;
    int rc = launch_fill(cur, 1, 0, st);
    if (rc) return rc;
    for (int64_t s = 0; s < nseg; s++) {
        int64_t kind = segments[3 * s], a = segments[3 * s + 1], b = segments[3 * s + 2];
        if (kind == 0) {
            TDP_REQUIRE(a >= 0 && b <= nops && a <= b, "segment %lld: op range", (long long)s);
            rc = run_ops_range(ops, a, b, cur, other, n, peak_bits, st);
        } else {
            TDP_REQUIRE(kind == 1 && a >= 0 && a < npasses, "segment %lld: pass index", (long long)s);
            const PassArgs &p = P[a];
            rc = check_pass(p);
            if (rc) return rc;
            TDP_REQUIRE(n == (1ull << p.b_in) && (1ull << p.b_out) <= cap, "segment %lld: pass on 2^%d entries, "
                        "table holds %llu", (long long)s, p.b_in, (unsigned long long)n);
            rc = launch_pass(cur, other, p, st);
            uint8_t *t = cur;
            cur = other;
            other = t;
            n = 1ull << p.b_out;
        }
        if (rc) return rc;
    }
    TDP_REQUIRE(n == 1, "program ends with %llu entries, not 1", (unsigned long long)n);
    uint8_t v = 0;
    TDP_HIP_CHECK(hipMemcpyAsync(&v, cur, 1, hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    *value = v;
    return TDP_OK;
}

}  // extern "C"
