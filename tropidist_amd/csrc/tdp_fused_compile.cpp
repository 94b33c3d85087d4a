// Host compiler of fused-pass programs in C++: the steps of tropidist_amd/fused.py (compile_program, _build_pass,
// pack, witness_program) on an encoded op stream, producing the segment triples, the pass descriptors of
// tdp_fused.hip and, in witness mode, the step records of its traceback.  The output is identical, field for field,
// to the Python compiler's (tests/test_fused_native.py), which is kept as the readable statement of the rules.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/tdp.h"

namespace tdp {
void set_error(const char *fmt, ...);
}

namespace {

constexpr int kMaxRows = 32, kMaxFree = 48, kMaxLow = 8, kMaxRuns = 12, kMaxVars = 64, kDMax = 5;
constexpr uint64_t kSlotAlign = 256;

// Layouts shared with tdp_fused.hip (checked by size there and in the Python bindings).
struct PassArgs {
    uint64_t din[8], dout[8];
    uint32_t dnew[8], dsel[8];
    uint32_t runs[4][kMaxRuns];             // f0 | len << 8 | sh << 16
    int32_t nruns[4];
    uint32_t sel_target;
    int32_t D, nlow, byte_store, b_in, b_out, nfree, pad;
    uint32_t lamw[8], loww[8];              // lane mask | weight << 8
    uint32_t qout[32], qsel[32], qpk[32];
};
struct WStep {
    uint64_t in_off, inmask, newmask, fixmask, fixval;
    uint64_t m[6];
    int64_t op[6];
    uint64_t out_off, n_in;
    int64_t op0;
    int32_t kind, pass_index, nop, k, b_out, pad;
    uint8_t w[6];
    uint8_t out_var[48];
    uint8_t pad2[2];
};
static_assert(sizeof(PassArgs) == 880, "pass descriptor layout");
static_assert(sizeof(WStep) == 240, "witness step layout");

struct Update {
    uint64_t m;
    uint64_t w;
    int64_t op;
};

struct Pass {
    int64_t first, last;
    int b_in, b_out;
    PassArgs args;
    std::vector<Update> upd_ops;
    uint64_t new_mask, fix_mask, fix_val;
    std::vector<int> out_vars;
    int max_runs;
};

inline int bitlen(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

// GF(2) elimination over 64-bit variable masks (independence tests), pivot = highest set bit.
struct Basis {
    uint64_t rows[64];
    bool has[64];
    Basis() { memset(has, 0, sizeof(has)); }
    uint64_t reduce(uint64_t v) const {
        while (v) {
            int t = bitlen(v) - 1;
            if (!has[t]) return v;
            v ^= rows[t];
        }
        return 0;
    }
    void add(uint64_t v) {
        v = reduce(v);
        int t = bitlen(v) - 1;
        rows[t] = v;
        has[t] = true;
    }
};

// runs [f0, len, shift] of consecutive thread-index bits mapped to consecutive single bits (fused.address_runs)
int address_runs(const uint64_t *vals, int n, uint32_t *runs) {
    int cnt = 0;
    int lf = -1, ll = 0, ls = 0;
    auto emit = [&]() {
        if (cnt < kMaxRuns) runs[cnt] = (uint32_t)lf | (uint32_t)ll << 8 | (uint32_t)ls << 16;
        cnt++;
    };
    for (int f = 0; f < n; f++) {
        uint64_t v = vals[f];
        if (!v) continue;
        int sh = bitlen(v) - 1;
        if (lf >= 0 && lf + ll == f && ls + ll == sh) {
            ll++;
        } else {
            if (lf >= 0) emit();
            lf = f;
            ll = 1;
            ls = sh;
        }
    }
    if (lf >= 0) emit();
    return cnt;
}

// fused._build_pass followed by fused.pack for one pass.  Returns false on a variable-count overflow.
bool build_pass(const uint64_t *ops, int64_t first, int64_t last, const std::vector<int> &in_log, bool logical_out,
                Pass &P, std::vector<int> &out_log) {
    const int b_in = (int)in_log.size();
    // scratch lists reused across calls (one set per thread: serving lanes compile concurrently)
    static thread_local std::vector<int> lay, new_vars, sel_vars, phys, movers, order, free;
    static thread_local std::vector<Update> upd;
    static thread_local std::vector<uint64_t> red, basis;
    lay.assign(in_log.begin(), in_log.end());
    new_vars.clear();
    sel_vars.clear();
    upd.clear();
    int nvar = b_in;
    int sel_t[kMaxVars];
    P.upd_ops.clear();
    for (int64_t i = first; i < last; i++) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (code == 0) {
            if (nvar >= kMaxVars) return false;
            lay.push_back(nvar);
            new_vars.push_back(nvar);
            nvar++;
        } else if (code == 1) {
            uint64_t m = 0;
            for (size_t p = 0; p < lay.size(); p++)
                if ((a >> p) & 1) m |= 1ull << lay[p];
            upd.push_back({m, b, i});
            if (m) P.upd_ops.push_back({m, b, i});
        } else {
            int u = lay[a];
            lay.erase(lay.begin() + (int64_t)a);
            sel_vars.push_back(u);
            sel_t[u] = (int)b;
        }
    }
    const int b_out = (int)lay.size();
    if (logical_out) {
        phys.assign(lay.begin(), lay.end());
    } else {
        char alive[kMaxVars + 1] = {0};
        for (int v : lay) alive[v] = 1;
        phys.assign(b_out, -1);
        movers.clear();
        order.clear();
        for (int v = 0; v < b_in; v++)
            if (alive[v]) {
                if (v < b_out) phys[v] = v;
                else movers.push_back(v);
            }
        for (int k = (int)new_vars.size() - 1; k >= 0; k--)
            if (alive[new_vars[k]]) order.push_back(new_vars[k]);
        order.insert(order.end(), movers.begin(), movers.end());
        size_t o = 0;
        for (int k = 0; k < b_out; k++)
            if (phys[k] < 0) phys[k] = order[o++];
    }
    int out_pos[kMaxVars + 1];
    std::fill(out_pos, out_pos + nvar, -1);
    for (int p = 0; p < b_out; p++) out_pos[phys[p]] = p;
    out_log.resize(lay.size());
    for (size_t p = 0; p < lay.size(); p++) out_log[p] = out_pos[lay[p]];

    struct Dir {
        uint64_t mh;
        uint32_t lam;
        uint64_t w;
    };
    static thread_local std::vector<Dir> dirs;
    static thread_local std::vector<std::pair<uint32_t, uint64_t>> lows;
    dirs.clear();
    lows.clear();
    for (const Update &u : upd) {
        uint64_t mh = u.m & ~3ull;
        uint32_t lam = (uint32_t)(u.m & 3);
        if (mh) dirs.push_back({mh, lam, u.w});
        else if (u.m) lows.push_back({lam, u.w});
    }
    const int D = (int)dirs.size();
    red.clear();
    for (const Dir &d : dirs) red.push_back(d.mh);
    char is_piv[kMaxVars + 1] = {0};
    for (int k = 0; k < D; k++) {
        uint64_t v = red[k];
        // pivot: the highest opened-row variable of the reduced vector, else its highest input bit
        int pv = -1;
        for (int x = nvar - 1; x >= 0; x--)
            if ((v >> x) & 1) {
                if (x >= b_in) {
                    pv = x;
                    break;
                }
                if (pv < 0) pv = x;
            }
        is_piv[pv] = 1;
        for (int k2 = 0; k2 < D; k2++)
            if (k2 != k && ((red[k2] >> pv) & 1)) red[k2] ^= v;
    }
    int new_idx[kMaxVars + 1], sel_idx[kMaxVars + 1];
    std::fill(new_idx, new_idx + nvar, -1);
    std::fill(sel_idx, sel_idx + nvar, -1);
    for (size_t k = 0; k < new_vars.size(); k++) new_idx[new_vars[k]] = (int)k;
    for (size_t k = 0; k < sel_vars.size(); k++) sel_idx[sel_vars[k]] = (int)k;
    auto deltas = [&](uint64_t mask, uint64_t &din, uint64_t &dout, uint32_t &dnew, uint32_t &dsel) {
        din = dout = 0;
        dnew = dsel = 0;
        for (uint64_t x = mask; x; x &= x - 1) {          // the variables of the mask (free ones have one)
            const int v = __builtin_ctzll(x);
            if (v >= nvar) break;
            if (v < b_in) din |= 1ull << v;
            if (out_pos[v] >= 0) dout |= 1ull << out_pos[v];
            if (new_idx[v] >= 0) dnew |= 1u << new_idx[v];
            if (sel_idx[v] >= 0) dsel |= 1u << sel_idx[v];
        }
    };
    PassArgs &A = P.args;
    memset(&A, 0, sizeof(A));
    basis.assign({1, 2});
    for (const Dir &d : dirs) basis.push_back(d.mh);
    for (size_t k = 0; k < basis.size(); k++) deltas(basis[k], A.din[k], A.dout[k], A.dnew[k], A.dsel[k]);
    free.clear();
    for (int v = 2; v < nvar; v++)
        if (!is_piv[v]) free.push_back(v);
    const int nfree = (int)free.size();
    if (nfree > kMaxFree) return false;
    uint64_t fin[64], fout[64], fnew[64], fsel[64];
    for (int f = 0; f < nfree; f++) {
        uint32_t a32, b32;
        deltas(1ull << free[f], fin[f], fout[f], a32, b32);
        fnew[f] = a32;
        fsel[f] = b32;
    }
    const uint64_t *maps[4] = {fin, fout, fnew, fsel};
    P.max_runs = 0;
    for (int m = 0; m < 4; m++) {
        int c = address_runs(maps[m], nfree, A.runs[m]);
        A.nruns[m] = c > kMaxRuns ? kMaxRuns : c;
        P.max_runs = std::max(P.max_runs, c);
    }
    uint32_t target = 0;
    for (size_t k = 0; k < sel_vars.size(); k++) target |= (uint32_t)sel_t[sel_vars[k]] << k;
    A.sel_target = target;
    A.D = D;
    A.nlow = (int)lows.size();
    A.nfree = nfree;
    A.b_in = b_in;
    A.b_out = b_out;
    A.byte_store = !(b_out >= 2 && phys[0] == 0 && phys[1] == 1);
    for (int j = 0; j < D && j < 8; j++) {
        A.lamw[j] = dirs[j].lam | (uint32_t)(dirs[j].w > 255 ? 255 : dirs[j].w) << 8;
    }
    for (size_t t = 0; t < lows.size() && t < 8; t++) {
        A.loww[t] = lows[t].first | (uint32_t)(lows[t].second > 255 ? 255 : lows[t].second) << 8;
    }
    for (int q = 0; q < (1 << std::min(D, 5)); q++) {
        uint64_t xi = 0, xo = 0;
        uint32_t xn = 0, xs = 0, sg = 0;
        for (int j = 0; j < D && j < 5; j++)
            if ((q >> j) & 1) {
                xi ^= A.din[2 + j];
                xo ^= A.dout[2 + j];
                xn ^= A.dnew[2 + j];
                xs ^= A.dsel[2 + j];
                sg ^= dirs[j].lam;
            }
        uint32_t pk = 0;
        for (uint32_t k = 0; k < 4; k++) {
            uint32_t r = k ^ sg;
            pk |= (((r & 1u) << 2) | (r & 2u)) << (8 * k);
        }
        (void)xi;
        (void)xn;
        A.qout[q] = (uint32_t)xo;
        A.qsel[q] = xs;
        A.qpk[q] = pk;
    }
    P.first = first;
    P.last = last;
    P.b_in = b_in;
    P.b_out = b_out;
    P.new_mask = 0;
    for (int v : new_vars) P.new_mask |= 1ull << v;
    P.fix_mask = P.fix_val = 0;
    for (int v : sel_vars) {
        P.fix_mask |= 1ull << v;
        P.fix_val |= (uint64_t)sel_t[v] << v;
    }
    P.out_vars = phys;
    return true;
}

}  // namespace

extern "C" int tdp_fused_compile(const uint64_t *ops, int64_t nops, int dmax, int min_bits, int max_updates,
                                 int64_t *segments, int64_t seg_cap, int64_t *nseg, void *passes_out,
                                 int64_t pass_cap, int64_t *npasses, void *steps_out, int64_t step_cap,
                                 int64_t *nsteps) {
    if (!ops || nops < 0 || dmax < 1 || dmax > kDMax || min_bits < 2 || !segments || !nseg || !passes_out ||
        !npasses) {
        tdp::set_error("tdp_fused_compile: bad arguments");
        return TDP_EINVAL;
    }
    // columns: [expands*, update?, selects*]
    std::vector<std::pair<int64_t, int64_t>> cols;
    for (int64_t i = 0; i < nops;) {
        int64_t j = i;
        while (j < nops && ops[3 * j] == 0) j++;
        if (j < nops && ops[3 * j] == 1) j++;
        while (j < nops && ops[3 * j] == 2) j++;
        if (j == i) j = i + 1;
        cols.push_back({i, j});
        i = j;
    }
    for (int64_t i = 0; i < nops; i++)
        if (ops[3 * i] > 2) {
            tdp::set_error("tdp_fused_compile: op %lld: unknown code", (long long)i);
            return TDP_EINVAL;
        }
    PassArgs *PA = static_cast<PassArgs *>(passes_out);
    std::vector<Pass> passes;
    int64_t ns = 0, np = 0;
    int b = 0;
    std::vector<int> log2phys;
    size_t k = 0;
    const size_t ncol = cols.size();
    while (k < ncol) {
        int64_t i0 = cols[k].first;
        if (b < min_bits) {
            int64_t start = i0;
            while (k < ncol && b < min_bits) {
                for (int64_t i = cols[k].first; i < cols[k].second; i++)
                    b += ops[3 * i] == 0 ? 1 : (ops[3 * i] == 2 ? -1 : 0);
                k++;
            }
            if (ns >= seg_cap) {
                tdp::set_error("tdp_fused_compile: segment capacity");
                return TDP_ECAPACITY;
            }
            segments[3 * ns] = 0;
            segments[3 * ns + 1] = start;
            segments[3 * ns + 2] = cols[k - 1].second;
            ns++;
            log2phys.resize(b);
            for (int q = 0; q < b; q++) log2phys[q] = q;
            continue;
        }
        Basis basis;
        int nlow = 0, D = 0, nnew = 0, nsel = 0;
        const int64_t start = i0;
        size_t kk = k;
        std::vector<int> lay(log2phys);
        int nvar = b;
        while (kk < ncol) {
            const int64_t c0 = cols[kk].first, c1 = cols[kk].second;
            int lay2[kMaxVars];                     // the layout after the column's expands (variables < 64)
            int n2 = (int)std::min<size_t>(lay.size(), kMaxVars);
            std::copy(lay.begin(), lay.begin() + n2, lay2);
            int nv2 = nvar;
            uint64_t m = 0;
            int ne = 0, nsl = 0;
            bool over = (int)lay.size() > kMaxVars;
            for (int64_t i = c0; i < c1 && !over; i++) {
                uint64_t code = ops[3 * i], a = ops[3 * i + 1];
                if (code == 0) {
                    if (nv2 >= kMaxVars || n2 >= kMaxVars) {
                        over = true;                // the column is not taken (checked below)
                        break;
                    }
                    lay2[n2++] = nv2++;
                    ne++;
                } else if (code == 1) {
                    for (int p = 0; p < n2; p++)
                        if ((a >> p) & 1) m |= 1ull << (lay2[p] & 63);
                } else {
                    nsl++;
                }
            }
            const uint64_t mh = m & ~3ull;
            if (over) break;
            if (nnew + ne > kMaxRows || nsel + nsl > kMaxRows) break;
            if (mh) {
                if (basis.reduce(mh) == 0 || D >= dmax) break;
            } else if (m) {
                if (nlow >= kMaxLow) break;
            }
            if (nv2 - 2 - (D + (mh ? 1 : 0)) > kMaxFree) break;
            if (max_updates > 0 && m && D + nlow + 1 > max_updates) break;
            if (nv2 > kMaxVars) break;
            nnew += ne;
            nsel += nsl;
            if (mh) {
                basis.add(mh);
                D++;
            } else if (m) {
                nlow++;
            }
            for (int64_t i = c0; i < c1; i++) {
                uint64_t code = ops[3 * i], a = ops[3 * i + 1];
                if (code == 0) lay.push_back(nvar++);
                else if (code == 2) lay.erase(lay.begin() + (int64_t)a);
            }
            kk++;
        }
        if (kk == k) {
            tdp::set_error("plan column at op %lld exceeds the fused pass limits", (long long)i0);
            return TDP_EINVAL;
        }
        Pass P;
        std::vector<int> out_log;
        while (true) {
            const int64_t last = cols[kk - 1].second;
            int bo = b;
            for (int64_t i = start; i < last; i++) bo += ops[3 * i] == 0 ? 1 : (ops[3 * i] == 2 ? -1 : 0);
            const bool logical = bo < min_bits || kk >= ncol;
            if (!build_pass(ops, start, last, log2phys, logical, P, out_log)) {
                tdp::set_error("plan column at op %lld exceeds the fused pass limits", (long long)i0);
                return TDP_EINVAL;
            }
            if (kk == k + 1 || P.max_runs <= kMaxRuns) break;
            kk--;
        }
        if (P.max_runs > kMaxRuns) {
            tdp::set_error("pass index map with %d runs (at most %d)", P.max_runs, kMaxRuns);
            return TDP_EINVAL;
        }
        if (ns >= seg_cap || np >= pass_cap) {
            tdp::set_error("tdp_fused_compile: segment or pass capacity");
            return TDP_ECAPACITY;
        }
        PA[np] = P.args;
        segments[3 * ns] = 1;
        segments[3 * ns + 1] = np;
        segments[3 * ns + 2] = 0;
        ns++;
        np++;
        log2phys = out_log;
        b = P.b_out;
        k = kk;
        if (steps_out) passes.push_back(std::move(P));
    }
    *nseg = ns;
    *npasses = np;
    if (!steps_out) return TDP_OK;

    // witness steps (fused.witness_program)
    if (!nsteps) {
        tdp::set_error("tdp_fused_compile: nsteps missing");
        return TDP_EINVAL;
    }
    WStep *S = static_cast<WStep *>(steps_out);
    int64_t nst = 0;
    uint64_t off = 0;
    int bits = 0;
    size_t pi = 0;
    auto push = [&](WStep &w, int b_in) -> bool {
        if (nst >= step_cap) return false;
        w.in_off = off;
        w.n_in = 1ull << b_in;
        off += ((1ull << b_in) + kSlotAlign - 1) / kSlotAlign * kSlotAlign;
        w.out_off = off;
        w.inmask = (1ull << b_in) - 1;
        S[nst++] = w;
        return true;
    };
    for (int64_t s = 0; s < ns; s++) {
        if (segments[3 * s] == 1) {
            const Pass &P = passes[pi];
            if (P.upd_ops.size() > 6 || P.b_out > 48) {
                tdp::set_error("witness pass with %zu updates", P.upd_ops.size());
                return TDP_EINVAL;
            }
            WStep w;
            memset(&w, 0, sizeof(w));
            w.kind = 0;
            w.pass_index = (int32_t)pi;
            w.op0 = P.first;
            w.nop = (int32_t)(P.last - P.first);
            w.newmask = P.new_mask;
            w.fixmask = P.fix_mask;
            w.fixval = P.fix_val;
            w.k = (int32_t)P.upd_ops.size();
            for (size_t j = 0; j < P.upd_ops.size(); j++) {
                w.m[j] = P.upd_ops[j].m;
                w.w[j] = (uint8_t)(P.upd_ops[j].w > 255 ? 255 : P.upd_ops[j].w);
                w.op[j] = P.upd_ops[j].op;
            }
            w.b_out = P.b_out;
            for (int q = 0; q < P.b_out; q++) w.out_var[q] = (uint8_t)P.out_vars[q];
            if (!push(w, P.b_in)) goto full;
            bits = P.b_out;
            pi++;
            continue;
        }
        for (int64_t i = segments[3 * s + 1]; i < segments[3 * s + 2];) {
            const uint64_t code = ops[3 * i], a = ops[3 * i + 1], wt = ops[3 * i + 2];
            WStep w;
            memset(&w, 0, sizeof(w));
            w.kind = 1;
            w.pass_index = -1;
            w.op0 = i;
            const int b_in = bits;
            if (code == 0) {
                if (bits >= 47) {
                    tdp::set_error("witness step beyond 2^47 entries");
                    return TDP_EINVAL;
                }
                const bool fuse = i + 1 < segments[3 * s + 2] && ops[3 * (i + 1)] == 1 &&
                                  (ops[3 * (i + 1) + 1] >> bits) == 1;
                w.nop = fuse ? 2 : 1;
                w.newmask = 1ull << bits;
                if (fuse) {
                    w.k = 1;
                    w.m[0] = ops[3 * (i + 1) + 1];
                    w.w[0] = (uint8_t)(ops[3 * (i + 1) + 2] > 255 ? 255 : ops[3 * (i + 1) + 2]);
                    w.op[0] = i + 1;
                }
                w.b_out = bits + 1;
                for (int q = 0; q <= bits; q++) w.out_var[q] = (uint8_t)q;
                i += fuse ? 2 : 1;
                bits++;
            } else if (code == 1) {
                w.nop = 1;
                if (a) {
                    w.k = 1;
                    w.m[0] = a;
                    w.w[0] = (uint8_t)(wt > 255 ? 255 : wt);
                    w.op[0] = i;
                }
                w.b_out = bits;
                for (int q = 0; q < bits; q++) w.out_var[q] = (uint8_t)q;
                i++;
            } else {
                w.nop = 1;
                w.fixmask = 1ull << a;
                w.fixval = wt << a;
                w.b_out = bits - 1;
                for (int q = 0; q < bits - 1; q++) w.out_var[q] = (uint8_t)((uint64_t)q < a ? q : q + 1);
                i++;
                bits--;
            }
            if (!push(w, b_in)) goto full;
        }
    }
    *nsteps = nst;
    return TDP_OK;
full:
    tdp::set_error("tdp_fused_compile: step capacity");
    return TDP_ECAPACITY;
}
