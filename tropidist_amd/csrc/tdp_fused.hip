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

#include <algorithm>
#include <vector>

namespace tdp {

constexpr int kPassBlock = 256;
constexpr int kPassWaves = 6;       // minimum waves per SIMD asked of the pass kernels
constexpr int kMaxRuns = 12;
enum { MAP_IN = 0, MAP_OUT = 1, MAP_NEW = 2, MAP_SEL = 3 };

// A thread's base indices: map m = OR over runs r < nruns[m] of ((tid >> f0) & (2^len - 1)) << sh, each run packed as
// f0 | len << 8 | sh << 16.  Every field is a naturally aligned 32- or 64-bit word, so the whole descriptor is read
// with scalar loads.
struct PassArgs {
    uint64_t din[8], dout[8];               // per basis vector (e_0, e_1, m'_0 .. m'_{D-1}): index deltas
    uint32_t dnew[8], dsel[8];              // new-row / ending-row coordinate deltas
    uint32_t runs[4][kMaxRuns];
    int32_t nruns[4];
    uint32_t sel_target;
    int32_t D, nlow, byte_store, b_in, b_out, nfree, pad;
    uint32_t lamw[8];                       // update j: lane mask | weight << 8
    uint32_t loww[8];                       // in-dword update t: lane mask | weight << 8
    // per cube element q < 2^D, precomputed on the host for the stores: XOR over the directions in q of dout (low
    // 32 bits, for tables of at most 2^32 entries) and dsel, and the v_perm selector that packs q's twisted lanes
    uint32_t qout[32], qsel[32], qpk[32];
};
static_assert(sizeof(PassArgs) == 880, "pass descriptor layout");

__device__ __forceinline__ uint32_t run_f0(uint32_t r) { return r & 0xFFu; }
__device__ __forceinline__ uint32_t run_len(uint32_t r) { return (r >> 8) & 0xFFu; }
__device__ __forceinline__ uint32_t run_sh(uint32_t r) { return (r >> 16) & 0xFFu; }

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Table accesses through the global address space (tables are device memory; the multi-program kernel reads its
// table pointers from memory, where the compiler could not tell them from generic pointers).
typedef __attribute__((address_space(1))) uint32_t gu32;
typedef __attribute__((address_space(1))) uint8_t gu8;
// NT: a nontemporal load (the input of a pass is read once; on rounds of large tables this leaves more of the
// memory-side cache to the outputs the next round reads).
template <bool NT = false>
__device__ __forceinline__ uint32_t gload32(const uint8_t *base, uint64_t off) {
    const gu32 *p = reinterpret_cast<const gu32 *>(reinterpret_cast<uintptr_t>(base) + off);
    if (NT) return __builtin_nontemporal_load(p);
    return *p;
}
__device__ __forceinline__ void gstore32(uint8_t *base, uint64_t off, uint32_t v) {
    *reinterpret_cast<gu32 *>(reinterpret_cast<uintptr_t>(base) + off) = v;
}
__device__ __forceinline__ void gstore8(uint8_t *base, uint64_t off, uint8_t v) {
    *reinterpret_cast<gu8 *>(reinterpret_cast<uintptr_t>(base) + off) = v;
}

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

// Butterflies of directions J.. of a cube of 2^D elements.
template <int D, int J>
struct DirLoop {
    static __device__ __forceinline__ void run(u16x2 *lo, u16x2 *hi, const PassArgs &a) {
        const unsigned short w = (unsigned short)(a.lamw[J] >> 8);
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
        const uint32_t run = a.runs[m][r];
        if (r < a.nruns[m] && (!lane_runs_only || run_f0(run) < 8)) {
            uint64_t mask = (1ull << run_len(run)) - 1;
            x |= (IDX)(((bits >> run_f0(run)) & mask) << run_sh(run));
        }
    }
    return x;
}

// IDX: uint32_t when both tables hold at most 2^32 entries (byte offsets from a scalar base), else uint64_t.
// FREE_NEW: some new row is a free variable, so whether an element starts at 255 varies across lanes; otherwise it
// depends on the element only and the test is uniform.
// FREE_NEW_ < 0: read from the descriptor at run time (the multi-program kernel, whose jobs differ).
template <int D, int FREE_NEW_, typename IDX, bool NT = false>
__device__ __forceinline__ void cube_load(const uint8_t *__restrict__ in, uint64_t block_bits, uint32_t lane,
                                          const PassArgs &a, uint32_t (&d)[1 << D]) {
    const bool FREE_NEW = FREE_NEW_ < 0 ? a.nruns[MAP_NEW] > 0 : FREE_NEW_ != 0;
    const IDX bin = map_bits<IDX>(a, MAP_IN, block_bits, false) | map_bits<IDX>(a, MAP_IN, lane, true);
    uint32_t xn = 0;
    if (FREE_NEW) xn = map_bits<uint32_t>(a, MAP_NEW, block_bits, false) | map_bits<uint32_t>(a, MAP_NEW, lane, true);
    constexpr int Q = 1 << D;
    // uniform per-element deltas, each the previous one (q with its lowest bit cleared) XOR one direction's:
    // one scalar XOR per element and quantity
    IDX cin[Q];
    uint32_t cnew[Q];
    cin[0] = 0;
    cnew[0] = 0;
#pragma unroll
    for (int q = 1; q < Q; q++) {
        const int j = __builtin_ctz(q), p = q & (q - 1);
        cin[q] = cin[p] ^ (IDX)a.din[2 + j];
        cnew[q] = cnew[p] ^ a.dnew[2 + j];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) d[q] = !(xn ^ cnew[q]) ? gload32<NT>(in, bin ^ cin[q]) : 0xFFFFFFFFu;
}

// The butterflies and stores of a loaded cube.  The twist selectors are affine in sigma:
// sel_lo(s) = sel_lo(0) ^ s * 0x10001, likewise sel_hi.
template <int D, bool BYTE, typename IDX>
__device__ __forceinline__ void cube_finish(uint8_t *__restrict__ out, uint64_t block_bits, uint32_t lane,
                                            const PassArgs &a, const uint32_t (&d)[1 << D]) {
    const IDX bout = map_bits<IDX>(a, MAP_OUT, block_bits, false) | map_bits<IDX>(a, MAP_OUT, lane, true);
    const uint32_t xs = map_bits<uint32_t>(a, MAP_SEL, block_bits, false) | map_bits<uint32_t>(a, MAP_SEL, lane, true);
    constexpr int Q = 1 << D;
    u16x2 lo[Q], hi[Q];
    uint32_t csig[Q];
    csig[0] = 0;
#pragma unroll
    for (int q = 1; q < Q; q++) csig[q] = csig[q & (q - 1)] ^ (a.lamw[__builtin_ctz(q)] & 3u) * 0x10001u;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        lo[q] = as_u16x2(__builtin_amdgcn_perm(d[q], d[q], sel_lo(0) ^ csig[q]));
        hi[q] = as_u16x2(__builtin_amdgcn_perm(d[q], d[q], sel_hi(0) ^ csig[q]));
    }
    DirLoop<D, 0>::run(lo, hi, a);
#pragma unroll
    for (int t = 0; t < 8; t++) {                  // constant indices keep the descriptor in kernel arguments
        if (t >= a.nlow) break;
        const unsigned short w = (unsigned short)(a.loww[t] >> 8);
        const bool sh = a.loww[t] & 1, sl = (a.loww[t] >> 1) & 1;
#pragma unroll
        for (int q = 0; q < Q; q++) low_update(lo[q], hi[q], sh, sl, u16x2{w, w});
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        IDX ao = bout;
        uint32_t sel = xs, s = 0, spk;
        if (sizeof(IDX) == 4 && !BYTE) {
            ao ^= (IDX)a.qout[q];
            sel ^= a.qsel[q];
            spk = a.qpk[q];
        } else {
#pragma unroll
            for (int j = 0; j < D; j++)
                if ((q >> j) & 1) {
                    ao ^= (IDX)a.dout[2 + j];
                    sel ^= a.dsel[2 + j];
                    s ^= a.lamw[j] & 3u;
                }
            spk = sel_pack(s);
        }
        if (!BYTE) {
            if (sel == a.sel_target)
                gstore32(out, ao, __builtin_amdgcn_perm(as_u32(hi[q]), as_u32(lo[q]), spk));
        } else {
            const unsigned short r[4] = {lo[q].x, hi[q].x, lo[q].y, hi[q].y};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t m = (uint32_t)k ^ s;           // byte of the dword this register entry came from
                IDX ak = ao ^ (IDX)((m & 1) ? a.dout[0] : 0) ^ (IDX)((m & 2) ? a.dout[1] : 0);
                uint32_t sk = sel ^ ((m & 1) ? a.dsel[0] : 0) ^ ((m & 2) ? a.dsel[1] : 0);
                if (sk == a.sel_target) gstore8(out, ak, (uint8_t)r[k]);
            }
        }
    }
}

// IDX: uint32_t when both tables hold at most 2^32 entries (byte offsets from a scalar base), else uint64_t.
// FREE_NEW: some new row is a free variable, so whether an element starts at 255 varies across lanes; otherwise it
// depends on the element only and the test is uniform.
// FREE_NEW_ < 0: read from the descriptor at run time (the multi-program kernel, whose jobs differ).
template <int D, bool BYTE, int FREE_NEW_, typename IDX, bool NT = false>
__device__ __forceinline__ void pass_cube(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                          uint64_t block_bits, uint32_t lane, const PassArgs &a) {
    uint32_t d[1 << D];
    cube_load<D, FREE_NEW_, IDX, NT>(in, block_bits, lane, a, d);
    cube_finish<D, BYTE, IDX>(out, block_bits, lane, a, d);
}

// Six waves per SIMD: the cube of a D = 5 pass then fits 80 VGPRs without spilling (the compiler's own choice, 88,
// allows five), and the extra wave hides more of the loads' latency: the widest passes gain 1-4% (kernel trace and
// HIP events).
template <int D, bool BYTE, bool FREE_NEW, typename IDX>
__global__ __launch_bounds__(kPassBlock, kPassWaves) void k_pass(const uint8_t *__restrict__ in, uint8_t *__restrict__ out,
                                                     uint64_t nthreads, const PassArgs a) {
    const uint64_t block = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (block * kPassBlock + threadIdx.x < nthreads)
        pass_cube<D, BYTE, FREE_NEW ? 1 : 0, IDX>(in, out, block * kPassBlock, threadIdx.x, a);
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
            const uint32_t run = a.runs[m][r];
            int f0 = run & 0xFF, len = (run >> 8) & 0xFF, sh = (run >> 16) & 0xFF;
            TDP_REQUIRE((run >> 24) == 0, "pass index map %d run %d: bad packing", m, r);
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
    for (int j = 0; j < a.D; j++)
        TDP_REQUIRE((a.lamw[j] & 0xFF) < 4 && (a.lamw[j] >> 8) >= 1 && (a.lamw[j] >> 8) <= 255,
                    "pass update %d: bad lane mask/weight", j);
    // the per-element tables must be the XORs of the direction deltas (they carry the addresses of 32-bit passes)
    for (int q = 0; q < (1 << a.D); q++) {
        uint64_t xi = 0, xo = 0;
        uint32_t xn = 0, xs = 0, sg = 0;
        for (int j = 0; j < a.D; j++)
            if ((q >> j) & 1) {
                xi ^= a.din[2 + j];
                xo ^= a.dout[2 + j];
                xn ^= a.dnew[2 + j];
                xs ^= a.dsel[2 + j];
                sg ^= a.lamw[j] & 3u;
            }
        uint32_t pk = 0;
        for (uint32_t k = 0; k < 4; k++) {
            uint32_t r = k ^ sg;
            pk |= (((r & 1u) << 2) | (r & 2u)) << (8 * k);
        }
        (void)xi;
        (void)xn;
        TDP_REQUIRE(a.qout[q] == (uint32_t)xo && a.qsel[q] == xs && a.qpk[q] == pk,
                    "pass element table %d inconsistent with its directions", q);
    }
    for (int t = 0; t < a.nlow; t++)
        TDP_REQUIRE((a.loww[t] & 0xFF) >= 1 && (a.loww[t] & 0xFF) < 4 && (a.loww[t] >> 8) >= 1 &&
                    (a.loww[t] >> 8) <= 255, "pass low update %d: bad mask", t);
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

// ------------------------------------------------------------------------------------------ witness mode
// Every step (a pass, or one per-op op; an expand followed by the update of its new top bit is one step) writes a
// fresh slot of the caller's arena, so each step's input table stays in HBM for the traceback.
struct WStep {
    uint64_t in_off, inmask, newmask, fixmask, fixval;   // input slot, its index mask, opened / ended-row variables
    uint64_t m[6];                                      // update masks in the step's variable space, op order
    int64_t op[6];                                      // op index of each update
    uint64_t out_off, n_in;                             // output slot, input entries
    int64_t op0;                                        // first op of a per-op step
    int32_t kind, pass_index, nop, k, b_out, pad;       // kind 0 = pass, 1 = per-op; k updates; output bits
    uint8_t w[6];                                       // update weights
    uint8_t out_var[48];                                // variable held by each output bit
    uint8_t pad2[2];
};
static_assert(sizeof(WStep) == 240, "witness step layout");

constexpr uint64_t kSlotAlign = 256;
static inline uint64_t align_up(uint64_t x) { return (x + kSlotAlign - 1) / kSlotAlign * kSlotAlign; }

__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        uint32_t u = (uint32_t)__shfl_xor((int)v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}

// One wave walks the steps backward.  At a step with k updates, the state x after it and the value V there, lane l
// holds X[x ^ XOR_{j in l} m_j] + sum_{j in l} w_j (X = 255 where an opened row is 1); update i, last first, was
// used iff the least of the lanes that agree with the used set T above i differs from V.
// io = {value, nused, error}: error 1 = an opened row is still 1 at the step's input, 2 = the walk does not end in
// state 0 with value 0.
__device__ void traceback_wave(const uint8_t *__restrict__ arena, const WStep *__restrict__ steps, int64_t nsteps,
                               uint64_t final_off, int64_t *__restrict__ used, int64_t *__restrict__ io) {
    const int lane = threadIdx.x;
    uint32_t v = arena[final_off];
    int64_t nused = 0, err = 0;
    if (v < 255 && nsteps > 0) {
        uint64_t x = steps[nsteps - 1].fixval;
        for (int64_t s = nsteps - 1; s >= 0; s--) {
            const WStep &st = steps[s];
            const int k = st.k;
            const uint32_t V = v;
            uint32_t val = 0xFFFFFFFFu;
            if (lane < (1 << k)) {
                uint64_t y = x;
                uint32_t ws = 0;
                for (int j = 0; j < k; j++)
                    if ((lane >> j) & 1) {
                        y ^= st.m[j];
                        ws += st.w[j];
                    }
                val = ((y & st.newmask) ? 255u : (uint32_t)arena[st.in_off + (y & st.inmask)]) + ws;
            }
            uint32_t T = 0;
            for (int i = k - 1; i >= 0; i--) {
                uint32_t cand = (lane < (1 << k) && ((uint32_t)lane >> i) == (T >> i)) ? val : 0xFFFFFFFFu;
                if (wave_min(cand) != V) {
                    T |= 1u << i;
                    v -= st.w[i];
                    if (lane == 0) used[nused] = st.op[i];
                    nused++;
                }
            }
            for (int j = 0; j < k; j++)
                if ((T >> j) & 1) x ^= st.m[j];
            if (x & st.newmask) {
                err = 1;
                break;
            }
            if (s > 0) {
                const WStep &pv = steps[s - 1];
                uint64_t xp = pv.fixval;
                for (int q = 0; q < pv.b_out; q++)
                    if ((x >> q) & 1) xp |= 1ull << pv.out_var[q];
                x = xp;
            }
        }
        if (!err && (x != 0 || v != 0)) err = 2;
    }
    if (lane == 0) {
        io[0] = (int64_t)arena[final_off];
        io[1] = nused;
        io[2] = err;
    }
}

__global__ __launch_bounds__(64) void k_traceback(const uint8_t *__restrict__ arena, const WStep *__restrict__ steps,
                                                  int64_t nsteps, uint64_t final_off, int64_t *__restrict__ used,
                                                  int64_t *__restrict__ io) {
    traceback_wave(arena, steps, nsteps, final_off, used, io);
}

// ------------------------------------------------------------------------------------------ many programs at once
// One launch runs the next pass of every active program ("round").  The round's jobs own consecutive ranges of
// blocks (first_block ascending); a block finds its job by counting, with all its threads at once, the jobs that
// start at or before it -- one coalesced read of the round's job records, no per-block map.  A job's descriptor,
// tables and thread count are read from device memory.  Jobs of different programs touch disjoint tables, so a
// round is a plain data-parallel launch.
struct Job {
    const uint8_t *in;
    uint8_t *out;
    uint64_t nthreads;
    uint32_t desc, first_block;
};

// Per-op work on a table of at most kFoldEntries entries (the first and last ops of a plan, where the table is
// too small for a pass), run as a job of a round by one thread instead of a launch of its own.  A job whose
// desc is kOpJob has one block, and its `in` points to an OpJob.  Value runs: ops on `cur` (selects write `other`
// and swap).  Witness steps (witness = 1): one select from cur (the step's input slot) to other (its output
// slot), or a copy of cur to other followed by an expand (fill), expand + update, or update on other.
constexpr uint32_t kOpJob = 0xFFFFFFFFu;
constexpr uint64_t kFoldEntries = 64;
struct OpJob {
    uint8_t *cur, *other;
    uint64_t n;
    const uint64_t *ops;
    int64_t nops;
    int32_t witness, pad;
};

// The ops of cdist/viterbi.c on a tiny table, one thread: expand = the new half at 255 (an update that follows
// then fills it, which is what the fused expand-update computes), update = the pairwise saturating min-plus,
// select = the compaction of the kept half.
__device__ void tiny_ops(const OpJob &J) {
    uint8_t *cur = J.cur, *oth = J.other;
    uint64_t n = J.n;
    int64_t i = 0;
    if (J.witness) {
        if (J.ops[0] == 2) {
            const int pos = (int)J.ops[1];
            const uint64_t tgt = J.ops[2], low = (1ull << pos) - 1;
            for (uint64_t k = 0; k < n / 2; k++) oth[k] = cur[((k & ~low) << 1) | (tgt << pos) | (k & low)];
            return;
        }
        for (uint64_t k = 0; k < n; k++) oth[k] = cur[k];
        cur = oth;
    }
    for (; i < J.nops; i++) {
        const uint64_t code = J.ops[3 * i], a = J.ops[3 * i + 1], b = J.ops[3 * i + 2];
        if (code == 0) {
            for (uint64_t k = 0; k < n; k++) cur[n + k] = 255;
            n <<= 1;
        } else if (code == 1) {
            if (a == 0 || n < 2) continue;                  // no pair to update, as the per-op kernels
            const uint32_t w = b > 255 ? 255u : (uint32_t)b;
            const uint64_t hb = 1ull << (63 - __builtin_clzll(a));
            for (uint64_t s2 = 0; s2 < n; s2++) {
                if (s2 & hb) continue;
                const uint64_t t = s2 ^ a;
                const uint32_t x = cur[s2], y = cur[t];
                const uint32_t xw = x + w > 255u ? 255u : x + w, yw = y + w > 255u ? 255u : y + w;
                cur[s2] = (uint8_t)(yw < x ? yw : x);
                cur[t] = (uint8_t)(xw < y ? xw : y);
            }
        } else {
            const int pos = (int)a;
            const uint64_t low = (1ull << pos) - 1;
            for (uint64_t k = 0; k < n / 2; k++) oth[k] = cur[((k & ~low) << 1) | (b << pos) | (k & low)];
            uint8_t *t = cur;
            cur = oth;
            oth = t;
            n >>= 1;
        }
    }
}

// The same bound spills 16 bytes per lane here; the headline step of four programs is still 2% faster with it than
// at five waves.
// NT: nontemporal input loads (rounds larger than the memory-side cache, kNtBytes).
// One block's share of one job of a round: an op job (one thread), or its cube of a pass.
template <bool NT>
__device__ __forceinline__ void pass_job(const PassArgs *__restrict__ descs, const uint8_t *in, uint8_t *out,
                                         uint64_t nthreads, uint32_t desc, uint32_t first_block) {
    if (desc == kOpJob) {
        if (threadIdx.x == 0) tiny_ops(*reinterpret_cast<const OpJob *>(in));
        return;
    }
    const uint64_t local = (uint64_t)(blockIdx.x - first_block);
    if (local * kPassBlock + threadIdx.x >= nthreads) return;
    const PassArgs &a = descs[desc];
    const uint64_t bb = local * kPassBlock;
    // one compiled shape per (D, store mode); the branch is uniform over the block
    switch (a.D * 2 + (a.byte_store ? 1 : 0)) {
        case 10: pass_cube<5, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 11: pass_cube<5, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 8: pass_cube<4, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 9: pass_cube<4, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 6: pass_cube<3, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 7: pass_cube<3, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 4: pass_cube<2, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 5: pass_cube<2, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 2: pass_cube<1, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 3: pass_cube<1, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        case 0: pass_cube<0, false, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
        default: pass_cube<0, true, -1, uint32_t, NT>(in, out, bb, threadIdx.x, a); break;
    }
}

template <bool NT>
__global__ __launch_bounds__(kPassBlock, kPassWaves) void k_pass_multi(const PassArgs *__restrict__ descs,
                                                           const Job *__restrict__ jobs, uint32_t njobs) {
    __shared__ uint32_t s_count;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    uint32_t c = 0;
    for (uint32_t j = threadIdx.x; j < njobs; j += kPassBlock) c += jobs[j].first_block <= blockIdx.x ? 1u : 0u;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(&s_count, c);
    __syncthreads();
    const Job &J = jobs[s_count - 1];
    pass_job<NT>(descs, J.in, J.out, J.nthreads, J.desc, J.first_block);
}

// A round of at most kRoundJobs jobs, its jobs passed by value: a block finds its job by scalar compares of kernel
// arguments instead of loading the job list, the job and then the descriptor one after another.
constexpr int kRoundJobs = 8;
struct RoundJobs {
    Job job[kRoundJobs];
    uint32_t njobs;
};

template <bool NT>
__global__ __launch_bounds__(kPassBlock, kPassWaves) void k_pass_round(const PassArgs *__restrict__ descs,
                                                                       const RoundJobs R) {
    const uint8_t *in = R.job[0].in;
    uint8_t *out = R.job[0].out;
    uint64_t nthreads = R.job[0].nthreads;
    uint32_t desc = R.job[0].desc, first = R.job[0].first_block;
#pragma unroll
    for (int k = 1; k < kRoundJobs; k++)            // uniform: constant indices, scalar selects
        if ((uint32_t)k < R.njobs && blockIdx.x >= R.job[k].first_block) {
            in = R.job[k].in;
            out = R.job[k].out;
            nthreads = R.job[k].nthreads;
            desc = R.job[k].desc;
            first = R.job[k].first_block;
        }
    pass_job<NT>(descs, in, out, nthreads, desc, first);
}

struct TraceJob {
    const uint8_t *arena;
    const WStep *steps;
    int64_t nsteps;
    uint64_t final_off;
    int64_t *used, *io;
};

__global__ __launch_bounds__(64) void k_traceback_multi(const TraceJob *__restrict__ jobs) {
    const TraceJob &J = jobs[blockIdx.x];
    traceback_wave(J.arena, J.steps, J.nsteps, J.final_off, J.used, J.io);
}

// out[i] = *ptrs[i]: the final entries of many value programs.
__global__ void k_gather(const uint8_t *const *__restrict__ ptrs, int n, uint8_t *__restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *ptrs[i];
}

struct WitnessLayout {
    uint64_t tables, steps_off, used_off, io_off, total, nupd;
};

static int witness_layout(const WStep *steps, int64_t nsteps, WitnessLayout *L) {
    TDP_REQUIRE(nsteps >= 1 && steps, "witness program without steps");
    uint64_t nupd = 0, end = 0;
    for (int64_t s = 0; s < nsteps; s++) {
        TDP_REQUIRE(steps[s].k >= 0 && steps[s].k <= 6, "step %lld: %d updates", (long long)s, steps[s].k);
        nupd += (uint64_t)steps[s].k;
        uint64_t e = steps[s].out_off + (steps[s].b_out >= 0 && steps[s].b_out <= 40 ? (1ull << steps[s].b_out) : 0);
        end = e > end ? e : end;
    }
    L->tables = align_up(end);
    L->steps_off = L->tables;
    L->used_off = align_up(L->steps_off + (uint64_t)nsteps * sizeof(WStep));
    L->io_off = align_up(L->used_off + 8 * (nupd ? nupd : 1));
    L->total = L->io_off + 64;
    L->nupd = nupd;
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

int tdp_witness_step_bytes(void) { return (int)sizeof(WStep); }

int tdp_program_witness_bytes(const void *steps, int64_t nsteps, uint64_t *bytes) {
    WitnessLayout L;
    int rc = witness_layout(static_cast<const WStep *>(steps), nsteps, &L);
    if (rc) return rc;
    *bytes = L.total;
    return TDP_OK;
}

}  // extern "C"

namespace tdp {

// Every check of a witness program before anything is launched: the chain of slots inside the arena, and each
// pass or per-op step against the op stream.
static int validate_witness(const uint64_t *ops, int64_t nops, const WStep *S, int64_t nsteps, const PassArgs *P,
                            int64_t npasses, uint64_t arena_bytes, WitnessLayout *Lout) {
    WitnessLayout L;
    int rc = witness_layout(S, nsteps, &L);
    if (rc) return rc;
    TDP_REQUIRE(arena_bytes >= L.total, "witness arena of %llu bytes, %llu needed", (unsigned long long)arena_bytes,
                (unsigned long long)L.total);
    TDP_REQUIRE(S[0].n_in == 1 && S[0].in_off == 0 && S[nsteps - 1].b_out == 0,
                "the steps must lead from the one-entry table to the one-entry table");
    // validate the chain of slots and every step before launching anything
    for (int64_t s = 0; s < nsteps; s++) {
        const WStep &w = S[s];
        TDP_REQUIRE(w.n_in >= 1 && (w.n_in & (w.n_in - 1)) == 0 && w.n_in <= (1ull << 40), "step %lld: bad size",
                    (long long)s);
        TDP_REQUIRE(w.b_out >= 0 && w.b_out <= 40 && w.out_off % kSlotAlign == 0 && w.in_off % kSlotAlign == 0,
                    "step %lld: bad slot", (long long)s);
        TDP_REQUIRE(w.in_off + w.n_in <= w.out_off && w.out_off + (1ull << w.b_out) <= L.tables,
                    "step %lld: slots out of order or outside the arena", (long long)s);
        TDP_REQUIRE(w.inmask == w.n_in - 1, "step %lld: index mask", (long long)s);
        for (int q = 0; q < w.b_out; q++) TDP_REQUIRE(w.out_var[q] < 64, "step %lld: output variable", (long long)s);
        if (s + 1 < nsteps)
            TDP_REQUIRE(S[s + 1].in_off == w.out_off && S[s + 1].n_in == (1ull << w.b_out),
                        "step %lld: the next step does not read this step's output", (long long)s);
        const int bin = top_bit(w.n_in);
        if (w.kind == 0) {
            TDP_REQUIRE(w.pass_index >= 0 && w.pass_index < npasses, "step %lld: pass index", (long long)s);
            const PassArgs &p = P[w.pass_index];
            rc = check_pass(p);
            if (rc) return rc;
            TDP_REQUIRE(p.b_in == bin && p.b_out == w.b_out, "step %lld: pass sizes", (long long)s);
        } else {
            TDP_REQUIRE(w.kind == 1 && w.op0 >= 0 && w.nop >= 1 && w.nop <= 2 && w.op0 + w.nop <= nops,
                        "step %lld: op range", (long long)s);
            const uint64_t *o = ops + 3 * w.op0;
            if (o[0] == 0) {
                TDP_REQUIRE(w.b_out == bin + 1, "step %lld: expand size", (long long)s);
                if (w.nop == 2)
                    TDP_REQUIRE(o[3] == 1 && (o[4] >> bin) == 1 && o[5] >= 1, "step %lld: expand-update",
                                (long long)s);
            } else if (o[0] == 1) {
                TDP_REQUIRE(w.nop == 1 && o[1] < w.n_in && o[1] >= 1 && o[2] >= 1 && w.b_out == bin,
                            "step %lld: update", (long long)s);
            } else {
                TDP_REQUIRE(o[0] == 2 && w.nop == 1 && bin >= 1 && o[1] < (uint64_t)bin && o[2] <= 1 &&
                            w.b_out == bin - 1, "step %lld: select", (long long)s);
            }
        }
    }
    *Lout = L;
    return TDP_OK;
}

// Step s of a validated witness program: a pass into its own slot, or one per-op op on a copy of its input.
static int run_witness_step(const uint64_t *ops, const WStep &w, uint8_t *d_arena, const PassArgs *P, hipStream_t st) {
    uint8_t *in = d_arena + w.in_off, *out = d_arena + w.out_off;
    const uint64_t n = w.n_in;
    if (w.kind == 0) return launch_pass(in, out, P[w.pass_index], st);
    const uint64_t *o = ops + 3 * w.op0;
    if (o[0] == 2) {
        uint64_t nn = 0;
        return launch_select(out, in, n, (int)o[1], (int)o[2], &nn, st);
    }
    TDP_HIP_CHECK(hipMemcpyAsync(out, in, n, hipMemcpyDeviceToDevice, st));
    if (o[0] == 0 && w.nop == 2) return launch_expand(out, n, o[4] ^ n, (uint32_t)(o[5] > 255 ? 255 : o[5]), st);
    if (o[0] == 0) return launch_fill(out + n, n, 255, st);
    return launch_update(out, n, o[1], (uint32_t)(o[2] > 255 ? 255 : o[2]), st);
}

}  // namespace tdp

extern "C" {

int tdp_run_program_witness(const uint64_t *ops, int64_t nops, const void *steps_v, int64_t nsteps,
                            const void *passes, int64_t npasses, uint8_t *d_arena, uint64_t arena_bytes, int *value,
                            int64_t *used, int64_t *nused, void *stream) {
    hipStream_t st = as_stream(stream);
    const WStep *S = static_cast<const WStep *>(steps_v);
    const PassArgs *P = static_cast<const PassArgs *>(passes);
    TDP_REQUIRE(d_arena && S && ops, "bad witness arguments");
    WitnessLayout L;
    int rc = validate_witness(ops, nops, S, nsteps, P, npasses, arena_bytes, &L);
    if (rc) return rc;
    rc = launch_fill(d_arena, 1, 0, st);
    if (rc) return rc;
    for (int64_t s = 0; s < nsteps && !rc; s++) rc = run_witness_step(ops, S[s], d_arena, P, st);
    if (rc) return rc;
    WStep *d_steps = reinterpret_cast<WStep *>(d_arena + L.steps_off);
    int64_t *d_used = reinterpret_cast<int64_t *>(d_arena + L.used_off);
    int64_t *d_io = reinterpret_cast<int64_t *>(d_arena + L.io_off);
    TDP_HIP_CHECK(hipMemcpyAsync(d_steps, S, (size_t)nsteps * sizeof(WStep), hipMemcpyHostToDevice, st));
    k_traceback<<<1, 64, 0, st>>>(d_arena, d_steps, nsteps, S[nsteps - 1].out_off, d_used, d_io);
    TDP_LAUNCH_CHECK();
    int64_t io[3] = {0, 0, 0};
    TDP_HIP_CHECK(hipMemcpyAsync(io, d_io, sizeof(io), hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    TDP_REQUIRE(io[2] == 0, "traceback failed (code %lld)", (long long)io[2]);
    TDP_REQUIRE(io[1] >= 0 && (uint64_t)io[1] <= L.nupd, "traceback used count %lld", (long long)io[1]);
    if (io[1] > 0) TDP_HIP_CHECK(hipMemcpy(used, d_used, (size_t)io[1] * 8, hipMemcpyDeviceToHost));
    *value = (int)io[0];
    *nused = io[1];
    return TDP_OK;
}

}  // extern "C"

namespace tdp {

// Table-size bookkeeping of an op range without launching anything (the checks and effects of run_ops_range):
// n entries, `swapped` flips on every select.
static int dry_ops_range(const uint64_t *ops, int64_t first, int64_t last, uint64_t &n, int peak_bits,
                         bool &swapped) {
    const uint64_t cap = 1ull << peak_bits;
    for (int64_t i = first; i < last; i++) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (code == 0) {
            TDP_REQUIRE(2 * n <= cap, "op %lld: expand beyond 2^%d entries", (long long)i, peak_bits);
            const uint64_t *nx = ops + 3 * (i + 1);
            if (i + 1 < last && nx[0] == 1 && (nx[1] >> top_bit(n)) == 1) {
                TDP_REQUIRE(nx[2] >= 1, "op %lld: weight < 1", (long long)(i + 1));
                i++;
            }
            n <<= 1;
        } else if (code == 1) {
            TDP_REQUIRE(a < n && b >= 1, "op %lld: bad update", (long long)i);
        } else if (code == 2) {
            TDP_REQUIRE(n >= 2 && a < (uint64_t)top_bit(n) && b <= 1, "op %lld: bad select", (long long)i);
            n >>= 1;
            swapped = !swapped;
        } else {
            TDP_REQUIRE(false, "op %lld: unknown code %llu", (long long)i, (unsigned long long)code);
        }
    }
    return TDP_OK;
}

// The whole schedule of a set of programs, computed on the host before anything is launched.
struct Action {
    int kind;             // 0 = per-op range of a value run, 1 = per-op witness step, 2 = one multi-pass round
    int run;
    int64_t a, b;         // kind 0: op range; kind 1: step index; kind 2: first job, number of jobs
    uint8_t *cur, *other; // kind 0: tables at that point
    uint64_t n;           // kind 0: entries at that point
    uint64_t nblocks;     // kind 2: blocks of the round
    bool nt;              // kind 2: nontemporal input loads (the round moves at least kNtBytes)
};
// Rounds whose inputs and outputs together exceed the 256 MB memory-side cache by half or more read their inputs
// nontemporally: on the headline step that raised the rounds of 4 x 2^25 -> 2^26 and 4 x 2^26 entries from 0.52
// to 0.54 of the HBM peak, while rounds that fit the cache (up to 4 x 2^25 -> 2^25) ran slower with it.
constexpr uint64_t kNtBytes = 3ull << 27;

struct Schedule {
    std::vector<uint32_t> desc_base;      // per run: index of its first pass descriptor in the upload
    uint64_t ndescs = 0;                  // descriptors are copied from the callers' arrays at upload
    std::vector<Job> jobs;                // op jobs hold their OpJob index in `in` until the upload
    std::vector<OpJob> opjobs;            // `ops` holds the index of the first op word until the upload
    std::vector<uint64_t> opwords;        // the op triples of the op jobs
    std::vector<Action> actions;
    std::vector<const uint8_t *> finals;  // value runs: the final one-entry table
    std::vector<int> value_runs, witness_runs;
    std::vector<WitnessLayout> layouts;   // per run (witness runs only)
    uint64_t off_jobs, off_opjobs, off_opwords, off_finals, off_results, off_trace, total;
};

// Largest table size along ops [first, last) from n entries (the ops were checked by dry_ops_range).
static uint64_t max_entries(const uint64_t *ops, int64_t first, int64_t last, uint64_t n) {
    uint64_t m = n;
    for (int64_t i = first; i < last; i++) {
        if (ops[3 * i] == 0) n <<= 1;
        else if (ops[3 * i] == 2) n >>= 1;
        m = n > m ? n : m;
    }
    return m;
}

static int build_schedule(const tdp_run_t *runs, int nruns, Schedule &S) {
    TDP_REQUIRE(nruns >= 1 && runs, "no programs");
    struct Ev {
        int kind;          // 0 = op range, 1 = witness per-op step, 2 = pass
        int64_t a, b;      // op range / step index / (pass index, step index)
    };
    std::vector<std::vector<Ev>> evs(nruns);
    std::vector<uint32_t> &desc_base = S.desc_base;
    desc_base.assign(nruns, 0);
    S.layouts.resize(nruns);
    for (int r = 0; r < nruns; r++) {
        const tdp_run_t &R = runs[r];
        const PassArgs *P = static_cast<const PassArgs *>(R.passes);
        TDP_REQUIRE(R.ops && R.nops >= 0 && (R.npasses == 0 || P), "run %d: bad op stream or passes", r);
        TDP_REQUIRE(S.ndescs + (uint64_t)R.npasses < (1ull << 31), "more than 2^31 pass descriptors");
        desc_base[r] = (uint32_t)S.ndescs;
        S.ndescs += (uint64_t)R.npasses;
        for (int64_t k = 0; k < R.npasses; k++) {
            int rc = check_pass(P[k]);
            if (rc) return rc;
            TDP_REQUIRE(P[k].b_in <= 32 && P[k].b_out <= 32, "run %d: multi-program passes are limited to 2^32 "
                        "entries", r);
        }
        if (R.steps) {
            const WStep *W = static_cast<const WStep *>(R.steps);
            TDP_REQUIRE(R.d_arena, "run %d: witness arena missing", r);
            int rc = validate_witness(R.ops, R.nops, W, R.nsteps, P, R.npasses, R.arena_bytes, &S.layouts[r]);
            if (rc) return rc;
            TDP_REQUIRE(R.used_ops, "run %d: used_ops missing", r);
            for (int64_t k = 0; k < R.nsteps; k++)
                evs[r].push_back(W[k].kind == 0 ? Ev{2, W[k].pass_index, k} : Ev{1, k, 0});
            S.witness_runs.push_back(r);
        } else {
            TDP_REQUIRE(R.d_T0 && R.d_T1 && R.peak_bits >= 0 && R.peak_bits <= 32 && R.segments && R.nseg >= 0,
                        "run %d: bad value program", r);
            uint64_t n = 1;
            bool sw = false;
            for (int64_t k = 0; k < R.nseg; k++) {
                const int64_t kind = R.segments[3 * k], a = R.segments[3 * k + 1], b = R.segments[3 * k + 2];
                if (kind == 0) {
                    TDP_REQUIRE(a >= 0 && a <= b && b <= R.nops, "run %d segment %lld: op range", r, (long long)k);
                    int rc = dry_ops_range(R.ops, a, b, n, R.peak_bits, sw);
                    if (rc) return rc;
                    evs[r].push_back(Ev{0, a, b});
                } else {
                    TDP_REQUIRE(kind == 1 && a >= 0 && a < R.npasses, "run %d segment %lld: pass", r, (long long)k);
                    TDP_REQUIRE(n == (1ull << P[a].b_in) && (1ull << P[a].b_out) <= (1ull << R.peak_bits),
                                "run %d segment %lld: pass sizes", r, (long long)k);
                    n = 1ull << P[a].b_out;
                    evs[r].push_back(Ev{2, a, -1});
                }
            }
            TDP_REQUIRE(n == 1, "run %d: the program does not end with one entry", r);
            S.value_runs.push_back(r);
        }
    }
    // rounds: every active program advances through its per-op work to its next pass; the passes form one launch
    std::vector<size_t> cur(nruns, 0);
    std::vector<uint8_t *> tc(nruns), to(nruns);
    std::vector<uint64_t> nn(nruns, 1);
    for (int r = 0; r < nruns; r++) {
        tc[r] = runs[r].steps ? nullptr : runs[r].d_T0;
        to[r] = runs[r].steps ? nullptr : runs[r].d_T1;
    }
    S.finals.assign(nruns, nullptr);
    while (true) {
        const size_t j0 = S.jobs.size();
        uint64_t nblocks = 0;
        uint64_t round_bytes = 0;
        for (int r = 0; r < nruns; r++) {
            const tdp_run_t &R = runs[r];
            // per-op work on a tiny table is this round's job of the program (its next pass waits a round)
            auto foldable = [&](const Ev &e) {
                if (e.kind == 0) return max_entries(R.ops, e.a, e.b, nn[r]) <= kFoldEntries;
                if (e.kind == 1) return static_cast<const WStep *>(R.steps)[e.a].n_in <= kFoldEntries / 2;
                return false;
            };
            if (cur[r] < evs[r].size() && foldable(evs[r][cur[r]])) {
                const Ev &e = evs[r][cur[r]++];
                OpJob oj{};
                int64_t op0, nop;
                if (e.kind == 0) {
                    oj.cur = tc[r];
                    oj.other = to[r];
                    oj.n = nn[r];
                    op0 = e.a;
                    nop = e.b - e.a;
                    bool sw = false;
                    int rc = dry_ops_range(R.ops, e.a, e.b, nn[r], R.peak_bits, sw);
                    if (rc) return rc;
                    if (sw) std::swap(tc[r], to[r]);
                } else {
                    const WStep &w = static_cast<const WStep *>(R.steps)[e.a];
                    oj.cur = R.d_arena + w.in_off;
                    oj.other = R.d_arena + w.out_off;
                    oj.n = w.n_in;
                    oj.witness = 1;
                    op0 = w.op0;
                    nop = w.nop;
                }
                oj.ops = reinterpret_cast<const uint64_t *>((uintptr_t)S.opwords.size());
                oj.nops = nop;
                S.opwords.insert(S.opwords.end(), R.ops + 3 * op0, R.ops + 3 * (op0 + nop));
                Job J;
                J.in = reinterpret_cast<const uint8_t *>((uintptr_t)S.opjobs.size());
                J.out = nullptr;
                J.nthreads = 1;
                J.desc = kOpJob;
                J.first_block = (uint32_t)nblocks;
                nblocks += 1;
                S.opjobs.push_back(oj);
                S.jobs.push_back(J);
                continue;
            }
            while (cur[r] < evs[r].size() && evs[r][cur[r]].kind != 2 && !foldable(evs[r][cur[r]])) {
                const Ev &e = evs[r][cur[r]++];
                Action act{};
                act.run = r;
                if (e.kind == 0) {
                    act.kind = 0;
                    act.a = e.a;
                    act.b = e.b;
                    act.cur = tc[r];
                    act.other = to[r];
                    act.n = nn[r];
                    bool sw = false;
                    int rc = dry_ops_range(R.ops, e.a, e.b, nn[r], R.peak_bits, sw);
                    if (rc) return rc;
                    if (sw) std::swap(tc[r], to[r]);
                } else {
                    act.kind = 1;
                    act.a = e.a;
                }
                S.actions.push_back(act);
            }
            if (cur[r] < evs[r].size() && evs[r][cur[r]].kind != 2) continue;   // folded in a later round
            if (cur[r] >= evs[r].size()) {
                if (!R.steps && !S.finals[r]) S.finals[r] = tc[r];
                continue;
            }
            const Ev &e = evs[r][cur[r]++];
            const PassArgs &p = static_cast<const PassArgs *>(R.passes)[e.a];
            Job J;
            if (R.steps) {
                const WStep &w = static_cast<const WStep *>(R.steps)[e.b];
                J.in = R.d_arena + w.in_off;
                J.out = R.d_arena + w.out_off;
            } else {
                J.in = tc[r];
                J.out = to[r];
                std::swap(tc[r], to[r]);
                nn[r] = 1ull << p.b_out;
            }
            J.nthreads = 1ull << p.nfree;
            J.desc = desc_base[r] + (uint32_t)e.a;
            round_bytes += (1ull << p.b_in) + (1ull << p.b_out);
            const uint64_t nb = (J.nthreads + kPassBlock - 1) / kPassBlock;
            TDP_REQUIRE(nblocks + nb < (1ull << 31), "round of more than 2^31 blocks");
            J.first_block = (uint32_t)nblocks;
            nblocks += nb;
            S.jobs.push_back(J);
        }
        if (S.jobs.size() == j0) break;
        Action act{};
        act.kind = 2;
        act.a = (int64_t)j0;
        act.b = (int64_t)(S.jobs.size() - j0);
        act.nblocks = nblocks;
        act.nt = round_bytes >= kNtBytes;
        S.actions.push_back(act);
    }
    auto up = [](uint64_t x) { return (x + 255) / 256 * 256; };
    S.off_jobs = up(S.ndescs * sizeof(PassArgs));
    S.off_opjobs = S.off_jobs + up(S.jobs.size() * sizeof(Job));
    S.off_opwords = S.off_opjobs + up(S.opjobs.size() * sizeof(OpJob));
    S.off_finals = S.off_opwords + up(S.opwords.size() * 8);
    S.off_results = S.off_finals + up(S.value_runs.size() * 8);
    S.off_trace = S.off_results + up(S.value_runs.size());
    S.total = S.off_trace + up(S.witness_runs.size() * sizeof(TraceJob)) + 256;
    return TDP_OK;
}

}  // namespace tdp

extern "C" {

int tdp_run_record_bytes(void) { return (int)sizeof(tdp_run_t); }

// A bound on the workspace from counts alone (no schedule): every pass a job, every segment or step an op job with
// all of the run's ops, and the layout's alignment slack.  tdp_run_programs checks the exact size.
int tdp_run_programs_workspace(const tdp_run_t *runs, int nruns, uint64_t *bytes) {
    TDP_REQUIRE(runs && nruns >= 1 && bytes, "no programs");
    uint64_t passes = 0, events = 0, ops = 0, values = 0, witnesses = 0;
    for (int r = 0; r < nruns; r++) {
        const tdp_run_t &R = runs[r];
        TDP_REQUIRE(R.npasses >= 0 && R.nops >= 0 && R.nseg >= 0 && R.nsteps >= 0, "run %d: negative counts", r);
        passes += (uint64_t)R.npasses;
        ops += (uint64_t)R.nops;
        if (R.steps) {
            events += (uint64_t)R.nsteps;
            witnesses++;
        } else {
            events += (uint64_t)R.nseg;
            values++;
        }
    }
    *bytes = passes * sizeof(PassArgs) + (passes + events) * sizeof(Job) + events * sizeof(OpJob) + 24 * ops +
             9 * values + witnesses * sizeof(TraceJob) + 8 * 256;
    return TDP_OK;
}

// Per-round timing of tdp_run_programs_timed: HIP events around every multi-program launch, on its stream.
struct RoundTiming {
    double *us, *bytes;
    int64_t cap, *n;
};

// A captured schedule (tdp_programs_graph): the launches of one tdp_run_programs call as a HIP graph, replayed by
// tdp_graph_launch; the value runs' final entries land at results (device) on every replay.
struct ProgramGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const uint8_t *results = nullptr;
    int nvalues = 0;
    uint64_t nodes = 0;
    ~ProgramGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

static int run_programs(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                        uint64_t staging_bytes, void *stream, const RoundTiming *timing, ProgramGraph *capture = nullptr);

int tdp_run_programs(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                     uint64_t staging_bytes, void *stream) {
    return run_programs(runs, nruns, d_work, work_bytes, host_staging, staging_bytes, stream, nullptr);
}

int tdp_run_programs_timed(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                           uint64_t staging_bytes, void *stream, double *round_us, double *round_bytes, int64_t cap,
                           int64_t *nrounds) {
    TDP_REQUIRE(round_us && round_bytes && nrounds && cap >= 0, "tdp_run_programs_timed: bad timing arrays");
    RoundTiming t{round_us, round_bytes, cap, nrounds};
    return run_programs(runs, nruns, d_work, work_bytes, host_staging, staging_bytes, stream, &t);
}

void *tdp_programs_graph(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *stream) {
    // capture on a stream of the library's own (the null stream, which callers often pass, cannot be captured),
    // ordered after the caller's stream so the upload follows the caller's earlier work
    hipStream_t cs = nullptr;
    hipEvent_t after = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&after) != hipSuccess ||
        hipEventRecord(after, as_stream(stream)) != hipSuccess || hipStreamWaitEvent(cs, after, 0) != hipSuccess) {
        set_error("tdp_programs_graph: capture stream setup failed");
        if (after) (void)hipEventDestroy(after);
        if (cs) (void)hipStreamDestroy(cs);
        return nullptr;
    }
    ProgramGraph *g = new ProgramGraph();
    const int rc = run_programs(runs, nruns, d_work, work_bytes, nullptr, 0, cs, nullptr, g);
    (void)hipStreamSynchronize(cs);
    (void)hipEventDestroy(after);
    (void)hipStreamDestroy(cs);
    if (rc != TDP_OK) {
        delete g;
        return nullptr;
    }
    return g;
}

int tdp_graph_launch(void *graph, void *stream) {
    TDP_REQUIRE(graph, "tdp_graph_launch: null graph");
    TDP_HIP_CHECK(hipGraphLaunch(static_cast<ProgramGraph *>(graph)->exec, as_stream(stream)));
    return TDP_OK;
}

int tdp_graph_values(void *graph, int32_t *values, int n, void *stream) {
    TDP_REQUIRE(graph && values, "tdp_graph_values: null argument");
    const ProgramGraph *g = static_cast<ProgramGraph *>(graph);
    TDP_REQUIRE(n == g->nvalues, "tdp_graph_values: %d values asked, the graph has %d", n, g->nvalues);
    std::vector<uint8_t> v(g->nvalues > 0 ? g->nvalues : 1);
    if (g->nvalues) TDP_HIP_CHECK(hipMemcpyAsync(v.data(), g->results, g->nvalues, hipMemcpyDeviceToHost,
                                                 as_stream(stream)));
    TDP_HIP_CHECK(hipStreamSynchronize(as_stream(stream)));
    for (int k = 0; k < n; k++) values[k] = v[k];
    return TDP_OK;
}

int tdp_graph_nodes(void *graph, uint64_t *nodes) {
    TDP_REQUIRE(graph && nodes, "tdp_graph_nodes: null argument");
    *nodes = static_cast<ProgramGraph *>(graph)->nodes;
    return TDP_OK;
}

void tdp_graph_free(void *graph) { delete static_cast<ProgramGraph *>(graph); }

}  // extern "C"

static int run_programs(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                        uint64_t staging_bytes, void *stream, const RoundTiming *timing, ProgramGraph *capture) {
    hipStream_t st = as_stream(stream);
    Schedule S;
    int rc = build_schedule(runs, nruns, S);
    if (rc) return rc;
    TDP_REQUIRE(!capture || S.witness_runs.empty(), "a captured schedule takes value runs only");
    TDP_REQUIRE(d_work && work_bytes >= S.total, "workspace of %llu bytes, %llu needed",
                (unsigned long long)work_bytes, (unsigned long long)S.total);
    // one upload: descriptors, jobs, final-table pointers, traceback records (through the caller's pinned staging
    // buffer when it is large enough)
    std::vector<uint8_t> own;
    uint8_t *hbuf = static_cast<uint8_t *>(host_staging);
    if (!hbuf || staging_bytes < S.total) {
        own.assign(S.total, 0);
        hbuf = own.data();
    }
    struct HostView {
        uint8_t *p;
        uint8_t *data() { return p; }
    } host{hbuf};
    for (int r = 0; r < nruns; r++)       // the validated descriptors, straight from the callers' arrays
        if (runs[r].npasses)
            memcpy(host.data() + (uint64_t)S.desc_base[r] * sizeof(PassArgs), runs[r].passes,
                   (size_t)runs[r].npasses * sizeof(PassArgs));
    for (Job &J : S.jobs)                 // op jobs: their record's device address
        if (J.desc == kOpJob)
            J.in = d_work + S.off_opjobs + (uintptr_t)J.in * sizeof(OpJob);
    for (OpJob &oj : S.opjobs)
        oj.ops = reinterpret_cast<const uint64_t *>(d_work + S.off_opwords + 8 * (uintptr_t)oj.ops);
    memcpy(host.data() + S.off_jobs, S.jobs.data(), S.jobs.size() * sizeof(Job));
    memcpy(host.data() + S.off_opjobs, S.opjobs.data(), S.opjobs.size() * sizeof(OpJob));
    memcpy(host.data() + S.off_opwords, S.opwords.data(), S.opwords.size() * 8);
    for (size_t k = 0; k < S.value_runs.size(); k++) {
        const uint8_t *f = S.finals[S.value_runs[k]];
        memcpy(host.data() + S.off_finals + 8 * k, &f, 8);
    }
    for (size_t k = 0; k < S.witness_runs.size(); k++) {
        const tdp_run_t &R = runs[S.witness_runs[k]];
        const WitnessLayout &L = S.layouts[S.witness_runs[k]];
        const WStep *W = static_cast<const WStep *>(R.steps);
        TraceJob T{R.d_arena, reinterpret_cast<const WStep *>(R.d_arena + L.steps_off), R.nsteps,
                   W[R.nsteps - 1].out_off, reinterpret_cast<int64_t *>(R.d_arena + L.used_off),
                   reinterpret_cast<int64_t *>(R.d_arena + L.io_off)};
        memcpy(host.data() + S.off_trace + k * sizeof(TraceJob), &T, sizeof(T));
    }
    TDP_HIP_CHECK(hipMemcpyAsync(d_work, host.data(), S.total, hipMemcpyHostToDevice, st));
    if (capture) {
        // the upload stays outside the graph (it is the same on every replay); every launch from here on is
        // recorded: kernels and device-to-device copies only
        TDP_HIP_CHECK(hipStreamSynchronize(st));
        TDP_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    }
    struct CaptureGuard {               // ends a capture left open by an error return
        hipStream_t s;
        bool on;
        ~CaptureGuard() {
            if (on) {
                hipGraph_t g = nullptr;
                if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            }
        }
    } cguard{st, capture != nullptr};
    for (int r = 0; r < nruns; r++) {
        rc = launch_fill(runs[r].steps ? runs[r].d_arena : runs[r].d_T0, 1, 0, st);
        if (rc) return rc;
    }
    const PassArgs *d_descs = reinterpret_cast<const PassArgs *>(d_work);
    const Job *d_jobs = reinterpret_cast<const Job *>(d_work + S.off_jobs);
    // timing: the algorithmic bytes of each round (every input entry read once, every output entry written once)
    // and a pair of events around its launch
    std::vector<hipEvent_t> evs;
    std::vector<double> rbytes;
    std::vector<const PassArgs *> desc_of;
    if (timing) {
        desc_of.resize(S.ndescs);
        for (int r = 0; r < nruns; r++)
            for (int64_t k = 0; k < runs[r].npasses; k++)
                desc_of[S.desc_base[r] + k] = static_cast<const PassArgs *>(runs[r].passes) + k;
    }
    struct EventGuard {
        std::vector<hipEvent_t> &e;
        ~EventGuard() {
            for (hipEvent_t x : e) (void)hipEventDestroy(x);
        }
    } guard{evs};
    for (const Action &act : S.actions) {
        const tdp_run_t &R = runs[act.run];
        if (act.kind == 0) {
            uint8_t *c = act.cur, *o = act.other;
            uint64_t n = act.n;
            rc = run_ops_range(R.ops, act.a, act.b, c, o, n, R.peak_bits, st);
        } else if (act.kind == 1) {
            rc = run_witness_step(R.ops, static_cast<const WStep *>(R.steps)[act.a], R.d_arena,
                                  static_cast<const PassArgs *>(R.passes), st);
        } else {
            const bool timed = timing && (int64_t)rbytes.size() < timing->cap;
            if (timed) {
                double b = 0;
                for (int64_t j = act.a; j < act.a + act.b; j++)
                    if (S.jobs[j].desc != kOpJob) {
                        const PassArgs *pa = desc_of[S.jobs[j].desc];
                        b += (double)(1ull << pa->b_in) + (double)(1ull << pa->b_out);
                    }
                rbytes.push_back(b);
                hipEvent_t e0, e1;
                TDP_HIP_CHECK(hipEventCreate(&e0));
                evs.push_back(e0);
                TDP_HIP_CHECK(hipEventCreate(&e1));
                evs.push_back(e1);
                TDP_HIP_CHECK(hipEventRecord(e0, st));
            }
const unsigned nbk = (unsigned)act.nblocks;
            const uint32_t nj = (uint32_t)act.b;
            if (nj <= (uint32_t)kRoundJobs) {
                RoundJobs rj{};
                for (uint32_t k = 0; k < nj; k++) rj.job[k] = S.jobs[act.a + k];
                rj.njobs = nj;
                if (act.nt) k_pass_round<true><<<nbk, kPassBlock, 0, st>>>(d_descs, rj);
                else k_pass_round<false><<<nbk, kPassBlock, 0, st>>>(d_descs, rj);
            } else if (act.nt) {
                k_pass_multi<true><<<nbk, kPassBlock, 0, st>>>(d_descs, d_jobs + act.a, nj);
            } else {
                k_pass_multi<false><<<nbk, kPassBlock, 0, st>>>(d_descs, d_jobs + act.a, nj);
            }
            TDP_LAUNCH_CHECK();
            if (timed) TDP_HIP_CHECK(hipEventRecord(evs.back(), st));
        }
        if (rc) return rc;
    }
    std::vector<uint8_t> vals(S.value_runs.size());
    if (!S.value_runs.empty()) {
        const int nv = (int)S.value_runs.size();
        k_gather<<<(nv + 255) / 256, 256, 0, st>>>(reinterpret_cast<const uint8_t *const *>(d_work + S.off_finals),
                                                   nv, d_work + S.off_results);
        TDP_LAUNCH_CHECK();
        if (!capture)
            TDP_HIP_CHECK(hipMemcpyAsync(vals.data(), d_work + S.off_results, nv, hipMemcpyDeviceToHost, st));
    }
    if (capture) {
        cguard.on = false;
        TDP_HIP_CHECK(hipStreamEndCapture(st, &capture->graph));
        size_t nn = 0;
        TDP_HIP_CHECK(hipGraphGetNodes(capture->graph, nullptr, &nn));
        capture->nodes = nn;
        TDP_HIP_CHECK(hipGraphInstantiate(&capture->exec, capture->graph, nullptr, nullptr, 0));
        capture->results = d_work + S.off_results;
        capture->nvalues = (int)S.value_runs.size();
        return TDP_OK;
    }
    std::vector<int64_t> io(3 * S.witness_runs.size());
    if (!S.witness_runs.empty()) {
        for (size_t k = 0; k < S.witness_runs.size(); k++) {
            const tdp_run_t &R = runs[S.witness_runs[k]];
            const WitnessLayout &L = S.layouts[S.witness_runs[k]];
            TDP_HIP_CHECK(hipMemcpyAsync(R.d_arena + L.steps_off, R.steps, (size_t)R.nsteps * sizeof(WStep),
                                         hipMemcpyHostToDevice, st));
        }
        k_traceback_multi<<<(unsigned)S.witness_runs.size(), 64, 0, st>>>(
            reinterpret_cast<const TraceJob *>(d_work + S.off_trace));
        TDP_LAUNCH_CHECK();
        for (size_t k = 0; k < S.witness_runs.size(); k++) {
            const tdp_run_t &R = runs[S.witness_runs[k]];
            const WitnessLayout &L = S.layouts[S.witness_runs[k]];
            TDP_HIP_CHECK(hipMemcpyAsync(&io[3 * k], R.d_arena + L.io_off, 24, hipMemcpyDeviceToHost, st));
        }
    }
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    if (timing) {
        for (size_t k = 0; k < rbytes.size(); k++) {
            float ms = 0;
            TDP_HIP_CHECK(hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]));
            timing->us[k] = 1e3 * (double)ms;
            timing->bytes[k] = rbytes[k];
        }
        *timing->n = (int64_t)rbytes.size();
    }
    for (size_t k = 0; k < S.value_runs.size(); k++) runs[S.value_runs[k]].value = vals[k];
    for (size_t k = 0; k < S.witness_runs.size(); k++) {
        tdp_run_t &R = runs[S.witness_runs[k]];
        const WitnessLayout &L = S.layouts[S.witness_runs[k]];
        TDP_REQUIRE(io[3 * k + 2] == 0, "run %d: traceback failed (code %lld)", S.witness_runs[k],
                    (long long)io[3 * k + 2]);
        TDP_REQUIRE(io[3 * k + 1] >= 0 && (uint64_t)io[3 * k + 1] <= L.nupd, "run %d: traceback used count",
                    S.witness_runs[k]);
        R.value = (int32_t)io[3 * k];
        R.nused = io[3 * k + 1];
        if (R.nused > 0)
            TDP_HIP_CHECK(hipMemcpy(R.used_ops, R.d_arena + L.used_off, (size_t)R.nused * 8, hipMemcpyDeviceToHost));
    }
    return TDP_OK;
}


