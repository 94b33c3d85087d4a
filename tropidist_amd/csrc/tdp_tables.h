// Device meet-in-the-middle tables shared by the table build/lookup (tdp_tables.hip) and the weight-bounded search
// (tdp_search.hip): the table record behind a tdp_tab handle and the probe rule of cdist/gmitm.c:136-163.
#pragma once
#include "tdp_common.h"

namespace tdp {

constexpr uint64_t kFpMask = 0x7FFFFFFFFFFF0000ull;

struct DevTable {
    uint64_t mask;
    uint64_t *key;
    uint64_t *lgs;              // wide only
    uint8_t *cost;
    unsigned long long *count;
    int k;
    int wide;
    uint64_t log2slots;
};

struct TabView {
    uint64_t mask;
    uint64_t *key;
    uint64_t *lgs;
    uint8_t *cost;
    unsigned long long *count;
};

static inline TabView view(const DevTable *t) { return TabView{t->mask, t->key, t->lgs, t->cost, t->count}; }

// Least cost <= R of an entry with the query's fingerprint and a different label; -1 if none.  The table must be
// complete (no build running): plain loads.
__device__ __forceinline__ int tab_probe(const TabView &t, int wide, uint64_t zh, uint64_t zl, uint64_t label,
                                         int R) {
    uint64_t i = zh & t.mask;
    int best = -1;
    const uint64_t fp = zl & kFpMask, wkey = zl | (1ull << 63);
    for (;;) {
        uint64_t cur = t.key[i];
        if (cur == 0) break;
        bool match = wide ? (cur == wkey && t.lgs[i] != label)
                          : ((cur & kFpMask) == fp && (cur & 0xFFFF) != (label & 0xFFFF));
        if (match) {
            int c = t.cost[i];
            if (c <= R && (best < 0 || c < best)) best = c;
        }
        i = (i + 1) & t.mask;
    }
    return best;
}

}  // namespace tdp
