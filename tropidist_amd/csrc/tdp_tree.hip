// Min-plus XOR joins of the tree-realization DP on gfx950 (cdist/tree_dp.c:13-51).
//
// A join enumerates the 2^J compatible (left, right, parent) index triples of one tree node: generators
// 0..pr-1 lift the parent's unit vectors, generators pr..J-1 span the kernel of the child-state sum, and the
// left / right indices are XORs of the generator maps lm / rm.  out[p] = min(out[p], L[l] + R[r]) over the
// triples whose costs are finite and whose sum stays below the infinity.
//
// Full joins: LPP lanes (a power of two <= 64, inside one wave) share a parent p; each lane walks 2^kr / LPP free
// combinations in Gray order (one XOR per step), the lanes reduce their minima with wave shuffles and one lane
// writes out[p] -- no atomics.  Partial Gray-counter ranges [start, stop) (the reference's chunks) use one thread
// per counter and an atomic minimum on the output word.
#include "tdp_common.h"

#include <cstring>
#include <vector>

namespace tdp {

struct JoinMaps {
    uint64_t lm[64];
    uint64_t rm[64];
};

constexpr int kTreeBlock = 256;

template <typename T>
__device__ __forceinline__ void atomic_min_item(T *addr, T v);

template <>
__device__ __forceinline__ void atomic_min_item<uint64_t>(uint64_t *addr, uint64_t v) {
    atomicMin(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)v);
}

template <>
__device__ __forceinline__ void atomic_min_item<uint32_t>(uint32_t *addr, uint32_t v) {
    atomicMin(addr, v);
}

// 8- and 16-bit items: compare-and-swap on the aligned 32-bit word that holds them.
template <typename T>
__device__ __forceinline__ void atomic_min_narrow(T *addr, T v) {
    uintptr_t a = reinterpret_cast<uintptr_t>(addr);
    uint32_t *word = reinterpret_cast<uint32_t *>(a & ~(uintptr_t)3);
    unsigned shift = (unsigned)(a & 3) * 8;
    uint32_t mask = ((1u << (8 * sizeof(T))) - 1) << shift;
    uint32_t old = *word;
    for (;;) {
        uint32_t cur = (old & mask) >> shift;
        if (cur <= v) return;
        uint32_t next = (old & ~mask) | ((uint32_t)v << shift);
        uint32_t seen = atomicCAS(word, old, next);
        if (seen == old) return;
        old = seen;
    }
}

template <>
__device__ __forceinline__ void atomic_min_item<uint16_t>(uint16_t *addr, uint16_t v) { atomic_min_narrow(addr, v); }

template <>
__device__ __forceinline__ void atomic_min_item<uint8_t>(uint8_t *addr, uint8_t v) { atomic_min_narrow(addr, v); }

template <typename T>
__global__ __launch_bounds__(kTreeBlock) void k_tree_join_full(const T *__restrict__ L, const T *__restrict__ R,
                                                               T *__restrict__ out, const JoinMaps maps, int pbits,
                                                               int kbits, int lpp_log, uint64_t inf) {
    uint64_t tid = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    uint64_t p = tid >> lpp_log;
    unsigned sub = (unsigned)(tid & ((1u << lpp_log) - 1));
    bool active = p < (1ull << pbits);
    uint64_t best = inf;
    if (active) {
        uint64_t lp = 0, rp = 0;
        for (int b = 0; b < pbits; b++)
            if ((p >> b) & 1) { lp ^= maps.lm[b]; rp ^= maps.rm[b]; }
        int per_log = kbits - lpp_log;
        uint64_t per = 1ull << per_log, base = (uint64_t)sub << per_log;
        uint64_t li = lp, ri = rp;
        for (int b = 0; b < kbits; b++)
            if ((base >> b) & 1) { li ^= maps.lm[pbits + b]; ri ^= maps.rm[pbits + b]; }
        for (uint64_t t = 0; t < per; t++) {
            if (t) {
                int bit = __builtin_ctzll(t);
                li ^= maps.lm[pbits + bit];
                ri ^= maps.rm[pbits + bit];
            }
            uint64_t a = L[li], b = R[ri];
            if (a < inf && b < inf && a <= inf - b && a + b < best) best = a + b;
        }
    }
    // minimum over the lanes of this parent inside the wave; with more than 64 lanes per parent each wave's
    // minimum goes to out[p] by an atomic minimum
    int red_log = lpp_log < 6 ? lpp_log : 6;
    for (int off = 1; off < (1 << red_log); off <<= 1) {
        uint64_t o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
    }
    if (active && (sub & ((1u << red_log) - 1)) == 0) {
        if (lpp_log <= 6) {
            if (best < (uint64_t)out[p]) out[p] = (T)best;
        } else if (best < inf) {
            atomic_min_item<T>(out + p, (T)best);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kTreeBlock) void k_tree_join_range(const T *__restrict__ L, const T *__restrict__ R,
                                                                T *out, const JoinMaps maps, int pbits,
                                                                uint64_t start, uint64_t stop, uint64_t inf) {
    uint64_t i = start + (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i >= stop) return;
    uint64_t g = i ^ (i >> 1), li = 0, ri = 0, pi = 0;
    for (uint64_t rest = g; rest; rest &= rest - 1) {
        int b = __builtin_ctzll(rest);
        li ^= maps.lm[b];
        ri ^= maps.rm[b];
        if (b < pbits) pi |= 1ull << b;
    }
    uint64_t a = L[li], b = R[ri];
    if (a < inf && b < inf && a <= inf - b) atomic_min_item<T>(out + pi, (T)(a + b));
}

template <typename T>
__global__ void k_tree_leaf(T *out, uint64_t size, uint64_t v0, uint64_t v1, uint64_t fill) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < size) out[i] = (T)(i == 0 ? v0 : (i == 1 ? v1 : fill));
}

template <typename T>
__global__ void k_tree_fill(T *out, uint64_t size, uint64_t v) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < size) out[i] = (T)v;
}

// One workgroup moves a short table toward a lower address in chunks of `gap` bytes (the distance between source and
// destination), in increasing order with a barrier between chunks, so no chunk reads a byte an earlier one wrote.
__global__ __launch_bounds__(1024) void k_move_ordered(uint8_t *dst, const uint8_t *src, uint64_t n, uint64_t gap) {
    for (uint64_t off = 0; off < n; off += gap) {
        const uint64_t len = n - off < gap ? n - off : gap;
        for (uint64_t i = threadIdx.x; i < len; i += blockDim.x) dst[off + i] = src[off + i];
        __syncthreads();
    }
}

__global__ void k_copy_bytes(uint8_t *__restrict__ dst, const uint8_t *__restrict__ src, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * kTreeBlock + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static unsigned tgrid(uint64_t items) { return (unsigned)((items + kTreeBlock - 1) / kTreeBlock); }

// ------------------------------------------------------------------------------------------------ tiled full joins
// A full join (all 2^J generator combinations) of 8- or 16-bit items, blocked through LDS.  Write a combination as
// g = parent part + kernel part u (the kr = J - pr generators that leave the parent alone).  If the kernel maps
// A_k: u -> left index and B_k: u -> right index are injective, every left delta of a parent direction v splits as
// rho_L(v) (a residue modulo im A_k) ^ A_k kappa_L(v), likewise on the right.  Parent directions then fall into
//   S  : rho_L = rho_R = 0            (a pure shift of the kernel index on both sides),
//   A' : rho_R = 0, rho_L != 0        (a new left row, a shift on the right),
//   B' : rho_L = 0, rho_R != 0        (a new right row, a shift on the left),
//   C  : the rest.
// A tile fixes the C directions (and the unused S / A' / B' ones) -- blockIdx -- and holds in LDS the left rows
// Lt[alpha][u] = L[lbase ^ lmap(alpha) ^ A_k u] of its 2^a row directions and the right rows Rt[beta][u] of its 2^b,
// as 16-bit values.  Its 2^(s+a+b) parents (sigma, alpha, beta) are then
//   out[p] = min_u Lt[alpha][u] + Rt[beta][u ^ delta],  delta = kappa_L(sigma) ^ kappa_R(sigma) ^ kappa_L(beta) ^
//                                                                  kappa_R(alpha),
// every pair of the join counted once: 2^(kr+s+a+b) pairs from 2^(kr+a) + 2^(kr+b) gathers.  Sums of two 16-bit
// values up to the infinity (<= 32767) do not overflow, and a sum >= the infinity never beats the table's entry,
// which is the reference's rule (cdist/tree_dp.c:30-36: both finite, sum <= infinity, strictly smaller).
constexpr int kTileMaxKr = 14;
constexpr int kTileMaxRows = 14;
constexpr int kTileMaxOut = 40;
constexpr int kTileMaxT = 40;
constexpr uint32_t kTileLdsEntries = 1u << 15;     // 16-bit entries of a tile's two tables: 64 KiB
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

struct TileArgs {
    uint64_t vl[kTileMaxKr + kTileMaxRows];   // left index deltas: kernel generators, then the A' row directions
    uint64_t vr[kTileMaxKr + kTileMaxRows];   // right: kernel generators, then the B' row directions
    uint64_t tl[kTileMaxT], tr[kTileMaxT], tp[kTileMaxT];   // tile directions: left, right and parent deltas
    uint64_t op[kTileMaxOut];                 // output directions (S, then A', then B'): parent deltas
    uint32_t od[kTileMaxOut];                 // ... and their kernel shifts delta
    int32_t kr, s, a, b, t, lanes_log;
};

// One 64-bit GF(2) echelon basis with coefficient tracking: reduce(v) gives the residue (no pivot bit set) and the
// combination of inserted vectors that was removed.
struct Echelon {
    uint64_t vec[64], coef[64];
    int piv[64], n = 0;
    bool insert(uint64_t v, uint64_t c) {
        for (int i = 0; i < n; i++)
            if ((v >> piv[i]) & 1) {
                v ^= vec[i];
                c ^= coef[i];
            }
        if (!v) return false;
        const int b = 63 - __builtin_clzll(v);
        for (int i = 0; i < n; i++)
            if ((vec[i] >> b) & 1) {
                vec[i] ^= v;
                coef[i] ^= c;
            }
        vec[n] = v;
        coef[n] = c;
        piv[n++] = b;
        return true;
    }
    uint64_t reduce(uint64_t v, uint64_t &c) const {
        c = 0;
        for (int i = 0; i < n; i++)
            if ((v >> piv[i]) & 1) {
                v ^= vec[i];
                c ^= coef[i];
            }
        return v;
    }
};

// Echelon basis over 128-bit images (two words) that records, for each vector of a spanning run, whether it was
// independent; used for kernels and extensions of subspaces of the parent space.
struct Echelon2 {
    uint64_t lo[128], hi[128];
    int piv[128], n = 0;
    bool insert(uint64_t l, uint64_t h) {
        for (int i = 0; i < n; i++) {
            const bool bit = piv[i] >= 64 ? (h >> (piv[i] - 64)) & 1 : (l >> piv[i]) & 1;
            if (bit) {
                l ^= lo[i];
                h ^= hi[i];
            }
        }
        if (!l && !h) return false;
        const int b = h ? 64 + 63 - __builtin_clzll(h) : 63 - __builtin_clzll(l);
        for (int i = 0; i < n; i++) {
            const bool bit = b >= 64 ? (hi[i] >> (b - 64)) & 1 : (lo[i] >> b) & 1;
            if (bit) {
                lo[i] ^= l;
                hi[i] ^= h;
            }
        }
        lo[n] = l;
        hi[n] = h;
        piv[n++] = b;
        return true;
    }
};

// Kernel basis of a linear map from the pr-dimensional parent space (unit vector i -> (img_lo[i], img_hi[i])).
static std::vector<uint64_t> parent_kernel(const uint64_t *img_lo, const uint64_t *img_hi, int pr) {
    std::vector<uint64_t> ker;
    std::vector<uint64_t> rl, rh, rx;
    std::vector<int> rp;
    for (int i = 0; i < pr; i++) {
        uint64_t l = img_lo[i], h = img_hi[i], x = 1ull << i;
        for (size_t k = 0; k < rp.size(); k++) {
            const bool bit = rp[k] >= 64 ? (h >> (rp[k] - 64)) & 1 : (l >> rp[k]) & 1;
            if (bit) {
                l ^= rl[k];
                h ^= rh[k];
                x ^= rx[k];
            }
        }
        if (!l && !h) {
            ker.push_back(x);
            continue;
        }
        const int b = h ? 64 + 63 - __builtin_clzll(h) : 63 - __builtin_clzll(l);
        for (size_t k = 0; k < rp.size(); k++) {
            const bool bit = b >= 64 ? (rh[k] >> (b - 64)) & 1 : (rl[k] >> b) & 1;
            if (bit) {
                rl[k] ^= l;
                rh[k] ^= h;
                rx[k] ^= x;
            }
        }
        rl.push_back(l);
        rh.push_back(h);
        rx.push_back(x);
        rp.push_back(b);
    }
    return ker;
}

// Plans a tiled full join; false when the join does not suit it (kernel maps not injective, too wide a kernel, no
// reuse, values that could overflow 16 bits), and the gather kernel runs it instead.
static bool plan_tiled_join(const uint64_t *lm, const uint64_t *rm, int pr, int J, int itemsize, uint64_t inf,
                            TileArgs &T, uint64_t &tiles, int min_tiles_log = 10) {
    memset(&T, 0, sizeof(T));
    const int kr = J - pr;
    if ((itemsize != 1 && itemsize != 2) || inf > 32767 || kr < 4 || kr > kTileMaxKr || pr < 1 || J > 62)
        return false;
    Echelon EL, ER;
    for (int j = 0; j < kr; j++)
        if (!EL.insert(lm[pr + j], 1ull << j) || !ER.insert(rm[pr + j], 1ull << j)) return false;
    // residues and kernel coordinates of the parent generators
    std::vector<uint64_t> rhoL(pr), kapL(pr), rhoR(pr), kapR(pr);
    for (int i = 0; i < pr; i++) {
        rhoL[i] = EL.reduce(lm[i], kapL[i]);
        rhoR[i] = ER.reduce(rm[i], kapR[i]);
    }
    auto comb = [&](const std::vector<uint64_t> &v, uint64_t x) {
        uint64_t r = 0;
        for (; x; x &= x - 1) r ^= v[__builtin_ctzll(x)];
        return r;
    };
    std::vector<uint64_t> zero(pr, 0), lmv(lm, lm + pr), rmv(rm, rm + pr);
    const std::vector<uint64_t> S = parent_kernel(rhoL.data(), rhoR.data(), pr);
    const std::vector<uint64_t> KR = parent_kernel(rhoR.data(), zero.data(), pr);
    const std::vector<uint64_t> KL = parent_kernel(rhoL.data(), zero.data(), pr);
    // adapted basis: S, then A' extending S inside ker rho_R, then B' inside ker rho_L, then C up to everything
    Echelon2 span;
    std::vector<uint64_t> A, B, C;
    for (uint64_t v : S) span.insert(v, 0);
    for (uint64_t v : KR)
        if (span.insert(v, 0)) A.push_back(v);
    for (uint64_t v : KL)
        if (span.insert(v, 0)) B.push_back(v);
    for (int i = 0; i < pr; i++)
        if (span.insert(1ull << i, 0)) C.push_back(1ull << i);
    if ((int)(S.size() + A.size() + B.size() + C.size()) != pr) return false;
    // tile shape: as many output directions as leave 2^10 tiles (or all of them on small joins), S first (they
    // cost no LDS), then A' and B' alternately while both row sets fit the LDS
    const int out_max = std::min(kTileMaxOut, std::max(0, pr - min_tiles_log));
    int s = std::min<int>((int)S.size(), out_max), a = 0, b = 0;
    for (;;) {
        const bool can_a = a < (int)A.size() && a < kTileMaxRows && s + a + b < out_max &&
                           (1u << (kr + a + 1)) + (1u << (kr + b)) <= kTileLdsEntries;
        const bool can_b = b < (int)B.size() && b < kTileMaxRows && s + a + b < out_max &&
                           (1u << (kr + a)) + (1u << (kr + b + 1)) <= kTileLdsEntries;
        if (can_a && (a <= b || !can_b)) a++;
        else if (can_b) b++;
        else break;
    }
    if ((1u << (kr + a)) + (1u << (kr + b)) > kTileLdsEntries) return false;
    if (s + a + b < 2) return false;                   // too little reuse: the gather kernel does as well
    const int t = pr - s - a - b;
    if (t > kTileMaxT || t > 31) return false;
    T.kr = kr;
    T.s = s;
    T.a = a;
    T.b = b;
    T.t = t;
    for (int j = 0; j < kr; j++) {
        T.vl[j] = lm[pr + j];
        T.vr[j] = rm[pr + j];
    }
    int o = 0;
    for (int i = 0; i < s; i++, o++) {
        T.op[o] = S[i];
        T.od[o] = (uint32_t)(comb(kapL, S[i]) ^ comb(kapR, S[i]));
    }
    for (int i = 0; i < a; i++, o++) {
        T.vl[kr + i] = comb(lmv, A[i]);
        T.op[o] = A[i];
        T.od[o] = (uint32_t)comb(kapR, A[i]);
    }
    for (int i = 0; i < b; i++, o++) {
        T.vr[kr + i] = comb(rmv, B[i]);
        T.op[o] = B[i];
        T.od[o] = (uint32_t)comb(kapL, B[i]);
    }
    std::vector<uint64_t> tdirs(S.begin() + s, S.end());
    tdirs.insert(tdirs.end(), A.begin() + a, A.end());
    tdirs.insert(tdirs.end(), B.begin() + b, B.end());
    tdirs.insert(tdirs.end(), C.begin(), C.end());
    for (int i = 0; i < t; i++) {
        T.tl[i] = comb(lmv, tdirs[i]);
        T.tr[i] = comb(rmv, tdirs[i]);
        T.tp[i] = tdirs[i];
    }
    // lanes per parent: 16 (one 64-byte run of a row per pass) when the rows are long enough
    const int rowdw_log = kr - 1;
    T.lanes_log = std::min(4, rowdw_log);
    tiles = 1ull << t;
    return true;
}

// Loads 2^nbits table entries into LDS as 16-bit values: entry e (bit k -> delta v[k]) from base ^ XOR v.  Each
// thread takes the entries whose low 8 bits are its index and walks the high bits in Gray order.
template <typename T>
__device__ __forceinline__ void tile_gather(const T *__restrict__ tab, uint64_t base, const uint64_t *v, int nbits,
                                            uint16_t *dst) {
    const uint32_t tid = threadIdx.x;
    const int low = nbits < 8 ? nbits : 8;
    if (tid >> low) return;
    uint64_t x = base;
    for (int k = 0; k < low; k++)
        if ((tid >> k) & 1) x ^= v[k];
    const uint32_t nhigh = 1u << (nbits - low);
    for (uint32_t k = 0; k < nhigh; k++) {
        if (k) x ^= v[low + __builtin_ctz(k)];
        const uint32_t e = tid | ((k ^ (k >> 1)) << low);
        dst[e] = (uint16_t)tab[x];
    }
}

template <typename T>
__global__ __launch_bounds__(kTreeBlock) void k_tree_join_tiled(const T *__restrict__ L, const T *__restrict__ R,
                                                                T *__restrict__ out, const TileArgs A) {
    extern __shared__ uint32_t tile_lds[];
    const uint64_t tau = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
    uint64_t lb = 0, rb = 0, pb = 0;
    for (int i = 0; i < A.t; i++)
        if ((tau >> i) & 1) {
            lb ^= A.tl[i];
            rb ^= A.tr[i];
            pb ^= A.tp[i];
        }
    const int kr = A.kr;
    uint16_t *ls = reinterpret_cast<uint16_t *>(tile_lds);
    uint16_t *rs = ls + (1u << (kr + A.a));
    tile_gather<T>(L, lb, A.vl, kr + A.a, ls);
    tile_gather<T>(R, rb, A.vr, kr + A.b, rs);
    __syncthreads();
    const uint32_t *lw = tile_lds;
    const uint32_t *rw = tile_lds + (1u << (kr + A.a - 1));
    const int nod = A.s + A.a + A.b;
    const uint32_t nout = 1u << nod, rowdw = 1u << (kr - 1);
    const int g_log = A.lanes_log;
    const uint32_t G = 1u << g_log, lane = threadIdx.x & (G - 1);
    for (uint32_t o0 = 0; o0 < nout; o0 += kTreeBlock >> g_log) {
        const uint32_t o = o0 + (threadIdx.x >> g_log);
        u16x2_t acc = {0xFFFF, 0xFFFF};     // four independent running minima keep four LDS reads in flight
        u16x2_t acc1 = acc, acc2 = acc, acc3 = acc;
        uint32_t delta = 0;
        uint64_t p = pb;
        if (o < nout) {
            for (uint32_t x = o; x; x &= x - 1) {
                const int i = __builtin_ctz(x);
                delta ^= A.od[i];
                p ^= A.op[i];
            }
            const uint32_t alpha = (o >> A.s) & ((1u << A.a) - 1), beta = o >> (A.s + A.a);
            const uint32_t *lrow = lw + alpha * rowdw, *rrow = rw + beta * rowdw;
            const uint32_t dh = delta >> 1, sh = (delta & 1) * 16;
            // a pair of dwords: the parent's partial minimum over kernel indices 2w and 2w + 1 (the right row read
            // at w ^ (delta >> 1), its halves swapped when delta is odd)
            auto step = [&](u16x2_t &m, uint32_t w) {
                const uint32_t rv = rrow[w ^ dh];
                const uint32_t r = __builtin_amdgcn_alignbit(rv, rv, sh);
                m = __builtin_elementwise_min(m, __builtin_bit_cast(u16x2_t, lrow[w]) + __builtin_bit_cast(u16x2_t, r));
            };
            uint32_t w = lane;
            for (; w + 3 * G < rowdw; w += 4 * G) {
                step(acc, w);
                step(acc1, w + G);
                step(acc2, w + 2 * G);
                step(acc3, w + 3 * G);
            }
            for (; w < rowdw; w += G) step(acc, w);
        }
        acc = __builtin_elementwise_min(__builtin_elementwise_min(acc, acc1), __builtin_elementwise_min(acc2, acc3));
        uint32_t m = acc.x < acc.y ? acc.x : acc.y;
        for (int off = 1; off < (int)G; off <<= 1) {
            const uint32_t q = (uint32_t)__shfl_xor((int)m, off, 64);
            m = q < m ? q : m;
        }
        if (o < nout && lane == 0 && m < (uint32_t)out[p]) out[p] = (T)m;
    }
}

template <typename T>
static int tree_join_t(const void *left, const void *right, void *out, const JoinMaps &maps, unsigned pbits,
                       unsigned gens, uint64_t start, uint64_t stop, uint64_t inf, hipStream_t st) {
    const T *L = static_cast<const T *>(left);
    const T *R = static_cast<const T *>(right);
    T *O = static_cast<T *>(out);
    uint64_t total = 1ull << gens;
    if (start >= stop) return TDP_OK;
    if (start == 0 && stop == total) {
        TileArgs tile;
        uint64_t tiles = 0;
        if (plan_tiled_join(maps.lm, maps.rm, (int)pbits, (int)gens, (int)sizeof(T), inf, tile, tiles)) {
            const size_t lds = 2ull * ((1ull << (tile.kr + tile.a)) + (1ull << (tile.kr + tile.b)));
            dim3 grid((unsigned)(tiles < (1ull << 20) ? tiles : (1ull << 20)), (unsigned)((tiles + (1ull << 20) - 1) >> 20));
            k_tree_join_tiled<T><<<grid, kTreeBlock, lds, st>>>(L, R, O, tile);
            TDP_LAUNCH_CHECK();
            return TDP_OK;
        }
        // lanes per parent: up to one wave, more (with per-wave atomics) when the parents alone cannot fill the
        // chip (about 2^20 threads)
        int kbits = (int)(gens - pbits);
        int lpp_log = kbits < 6 ? kbits : 6;
        if ((int)pbits + lpp_log < 20) lpp_log = kbits < 20 - (int)pbits ? kbits : 20 - (int)pbits;
        uint64_t threads = (1ull << pbits) << lpp_log;
        k_tree_join_full<T><<<tgrid(threads), kTreeBlock, 0, st>>>(L, R, O, maps, (int)pbits, kbits, lpp_log, inf);
    } else {
        k_tree_join_range<T><<<tgrid(stop - start), kTreeBlock, 0, st>>>(L, R, O, maps, (int)pbits, start, stop,
                                                                          inf);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_tree_join(int itemsize, const void *d_left, const void *d_right, void *d_out, const uint64_t *lm,
                  const uint64_t *rm, unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop,
                  uint64_t infinity, void *stream) {
    TDP_REQUIRE(generators <= 62 && parent_bits <= generators, "join with %u generators, %u parent bits",
                generators, parent_bits);
    TDP_REQUIRE(stop <= (1ull << generators), "counter range beyond 2^%u", generators);
    JoinMaps maps;
    memset(&maps, 0, sizeof(maps));
    for (unsigned b = 0; b < generators; b++) {
        maps.lm[b] = lm[b];
        maps.rm[b] = rm[b];
    }
    hipStream_t st = as_stream(stream);
    switch (itemsize) {
        case 1: return tree_join_t<uint8_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 2: return tree_join_t<uint16_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 4: return tree_join_t<uint32_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
        case 8: return tree_join_t<uint64_t>(d_left, d_right, d_out, maps, parent_bits, generators, start, stop, infinity, st);
    }
    TDP_REQUIRE(false, "item size %d", itemsize);
}

int tdp_tree_tile_plan(const uint64_t *lm, const uint64_t *rm, unsigned parent_bits, unsigned generators,
                       int itemsize, uint64_t infinity, int min_tiles_log, void *args, uint64_t *tiles) {
    TDP_REQUIRE(generators <= 62 && parent_bits <= generators, "join with %u generators, %u parent bits",
                generators, parent_bits);
    TDP_REQUIRE(min_tiles_log >= 0 && min_tiles_log <= 40, "min_tiles_log %d", min_tiles_log);
    TileArgs T;
    uint64_t n = 0;
    const bool ok = plan_tiled_join(lm, rm, (int)parent_bits, (int)generators, itemsize, infinity, T, n,
                                    min_tiles_log);
    if (args) memcpy(args, &T, sizeof(T));
    *tiles = ok ? n : 0;
    return TDP_OK;
}

int tdp_tree_tile_args_bytes(void) { return (int)sizeof(TileArgs); }

#define TDP_TREE_JOIN_NAMED(BITS, BYTES)                                                                     \
    int tdp_tree_join_u##BITS(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm,      \
                              const uint64_t *rm, unsigned parent_bits, unsigned generators, uint64_t start, \
                              uint64_t stop, uint64_t infinity, void *stream) {                            \
        return tdp_tree_join(BYTES, d_left, d_right, d_out, lm, rm, parent_bits, generators, start, stop,     \
                             infinity, stream);                                                             \
    }
TDP_TREE_JOIN_NAMED(8, 1)
TDP_TREE_JOIN_NAMED(16, 2)
TDP_TREE_JOIN_NAMED(32, 4)
TDP_TREE_JOIN_NAMED(64, 8)

int tdp_tree_fill(int itemsize, void *d_out, uint64_t size, uint64_t value, void *stream) {
    hipStream_t st = as_stream(stream);
    if (!size) return TDP_OK;
    switch (itemsize) {
        case 1: k_tree_fill<uint8_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint8_t *)d_out, size, value); break;
        case 2: k_tree_fill<uint16_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint16_t *)d_out, size, value); break;
        case 4: k_tree_fill<uint32_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint32_t *)d_out, size, value); break;
        case 8: k_tree_fill<uint64_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint64_t *)d_out, size, value); break;
        default: TDP_REQUIRE(false, "item size %d", itemsize);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

int tdp_tree_leaf(int itemsize, void *d_out, uint64_t size, uint64_t v0, uint64_t v1, uint64_t fill, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(size >= 1, "empty leaf table");
    switch (itemsize) {
        case 1: k_tree_leaf<uint8_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint8_t *)d_out, size, v0, v1, fill); break;
        case 2: k_tree_leaf<uint16_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint16_t *)d_out, size, v0, v1, fill); break;
        case 4: k_tree_leaf<uint32_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint32_t *)d_out, size, v0, v1, fill); break;
        case 8: k_tree_leaf<uint64_t><<<tgrid(size), kTreeBlock, 0, st>>>((uint64_t *)d_out, size, v0, v1, fill); break;
        default: TDP_REQUIRE(false, "item size %d", itemsize);
    }
    TDP_LAUNCH_CHECK();
    return TDP_OK;
}

// Overlap-safe move toward lower addresses (the postorder arena compaction of cdist/tree_dp.c:46-51): copied in
// chunks no longer than the distance between source and destination, in increasing order, so no chunk reads a
// byte an earlier chunk wrote.
int tdp_tree_move(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    hipStream_t st = as_stream(stream);
    uint8_t *dst = static_cast<uint8_t *>(d_dst);
    const uint8_t *src = static_cast<const uint8_t *>(d_src);
    if (!bytes || dst == src) return TDP_OK;
    TDP_REQUIRE(dst < src, "tree_move only moves toward lower addresses");
    uint64_t gap = (uint64_t)(src - dst);
    if (bytes <= (1u << 20) && gap < bytes) {            // short overlapping move: one launch, chunks in order
        k_move_ordered<<<1, 1024, 0, st>>>(dst, src, bytes, gap);
        TDP_LAUNCH_CHECK();
        return TDP_OK;
    }
    for (uint64_t off = 0; off < bytes; off += gap) {
        uint64_t len = bytes - off < gap ? bytes - off : gap;
        if (len >= (1u << 16) && gap >= bytes) {
            TDP_HIP_CHECK(hipMemcpyAsync(dst + off, src + off, len, hipMemcpyDeviceToDevice, st));
        } else {
            k_copy_bytes<<<tgrid(len), kTreeBlock, 0, st>>>(dst + off, src + off, len);
            TDP_LAUNCH_CHECK();
        }
    }
    return TDP_OK;
}

}  // extern "C"
