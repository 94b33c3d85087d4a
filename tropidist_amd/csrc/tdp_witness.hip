// Witness of the dense trellis DP: a minimum-cost column set, by traceback through the tables (the algorithm of
// cdist/auto.py:126-182, which keeps every pre-update table on the host; here the tables stay in HBM and only
// checkpoints are kept).
//
//   forward   the plan runs op by op (no expand-update fusion, so every update's input table exists); the table
//             is copied to a checkpoint slot before every S-th op, S = ceil(sqrt(nops));
//   backward  for each segment, last first: the segment is recomputed from its checkpoint, keeping a copy of the
//             table before each of its updates, and one device thread walks the segment's ops backward:
//             select -> re-insert the target bit; expand -> the new (top) bit must be 0; update -> the column is
//             used iff before[state] != value, then state ^= mask, value -= weight.
// Memory: 2 working tables + ceil(nops / S) checkpoints + S segment tables, 2^peak_bits bytes each.
#include "tdp_common.h"

#include <math.h>

#include <vector>

namespace tdp {

struct WitnessLayout {
    int64_t S, nseg;
    uint64_t cap, off_t1, off_ck, off_seg, off_meta, total;
};

static WitnessLayout witness_layout(int64_t nops, int peak_bits) {
    WitnessLayout L;
    L.S = nops > 0 ? (int64_t)ceil(sqrt((double)nops)) : 1;
    L.nseg = nops > 0 ? (nops + L.S - 1) / L.S : 0;
    L.cap = 1ull << peak_bits;
    uint64_t cap16 = (L.cap + 255) & ~255ull;
    L.off_t1 = cap16;
    L.off_ck = 2 * cap16;
    L.off_seg = L.off_ck + (uint64_t)L.nseg * cap16;
    L.off_meta = L.off_seg + (uint64_t)L.S * cap16;
    // metadata: segment ops (3 S u64), before-table offsets (S), table sizes (S), io (4), used op indices (nops)
    L.total = L.off_meta + 8ull * (5 * (uint64_t)L.S + 4 + (uint64_t)(nops > 0 ? nops : 1));
    return L;
}

// One thread: walks ops [0, nseg) of a segment backward.  io = {state, value, nused, error}.
__global__ void k_traceback(const uint64_t *__restrict__ ops, int64_t first_op, int64_t nseg,
                            const uint8_t *__restrict__ tabs, const uint64_t *__restrict__ toff,
                            const uint64_t *__restrict__ nbefore, int64_t *io, int64_t *used) {
    uint64_t state = (uint64_t)io[0];
    int64_t value = io[1], nused = io[2], err = io[3];
    for (int64_t i = nseg - 1; i >= 0 && !err; i--) {
        uint64_t code = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        uint64_t nb = nbefore[i];
        if (code == 2) {
            uint64_t low = state & ((1ull << a) - 1);
            state = ((state >> a) << (a + 1)) | (b << a) | low;
        } else if (code == 0) {
            if (state >= nb) err = 1;          // entries with the new bit set were INF before the row started
        } else {
            const uint8_t *before = tabs + toff[i];
            if (state >= nb) { err = 2; break; }
            if ((int64_t)before[state] != value) {
                used[nused++] = first_op + i;
                state ^= a;
                value -= (int64_t)b;
                if (state >= nb || (int64_t)before[state] != value) err = 3;
            }
        }
    }
    io[0] = (int64_t)state;
    io[1] = value;
    io[2] = nused;
    io[3] = err;
}

static int run_op(const uint64_t *op, uint8_t *&cur, uint8_t *&other, uint64_t &n, uint64_t cap, hipStream_t st) {
    uint64_t code = op[0], a = op[1], b = op[2];
    if (code == 0) {
        TDP_REQUIRE(2 * n <= cap, "expand beyond the peak table");
        int rc = launch_fill(cur + n, n, 255, st);
        n <<= 1;
        return rc;
    }
    if (code == 1) {
        TDP_REQUIRE(a < n && b >= 1, "update mask 0x%llx / weight %llu invalid on %llu entries",
                    (unsigned long long)a, (unsigned long long)b, (unsigned long long)n);
        return launch_update(cur, n, a, (uint32_t)(b > 255 ? 255 : b), st);
    }
    TDP_REQUIRE(code == 2 && n >= 2 && a < (uint64_t)top_bit(n) && b <= 1, "bad op code %llu",
                (unsigned long long)code);
    uint64_t nn = 0;
    int rc = launch_select(other, cur, n, (int)a, (int)b, &nn, st);
    uint8_t *t = cur;
    cur = other;
    other = t;
    n = nn;
    return rc;
}

}  // namespace tdp

using namespace tdp;

extern "C" {

int tdp_witness_workspace_bytes(int64_t nops, int peak_bits, uint64_t *bytes) {
    TDP_REQUIRE(nops >= 0 && peak_bits >= 0 && peak_bits <= 40, "nops %lld / peak_bits %d", (long long)nops,
                peak_bits);
    *bytes = witness_layout(nops, peak_bits).total;
    return TDP_OK;
}

int tdp_run_plan_witness(const uint64_t *ops, int64_t nops, int peak_bits, uint8_t *d_work, uint64_t work_bytes,
                         int *value, int64_t *used_ops, int64_t *nused, void *stream) {
    hipStream_t st = as_stream(stream);
    TDP_REQUIRE(peak_bits >= 0 && peak_bits <= 40 && nops >= 0, "bad plan");
    WitnessLayout L = witness_layout(nops, peak_bits);
    TDP_REQUIRE(work_bytes >= L.total, "witness workspace of %llu bytes, %llu needed",
                (unsigned long long)work_bytes, (unsigned long long)L.total);
    uint64_t cap16 = L.off_t1;
    uint8_t *T0 = d_work, *T1 = d_work + L.off_t1;
    uint64_t *meta = reinterpret_cast<uint64_t *>(d_work + L.off_meta);
    uint64_t *d_ops = meta, *d_toff = meta + 3 * L.S, *d_nb = meta + 4 * L.S;
    int64_t *d_io = reinterpret_cast<int64_t *>(meta + 5 * L.S);
    int64_t *d_used = d_io + 4;
    std::vector<uint64_t> ck_n(L.nseg > 0 ? L.nseg : 1);

    // forward with checkpoints
    uint8_t *cur = T0, *other = T1;
    uint64_t n = 1;
    int rc = launch_fill(cur, 1, 0, st);
    if (rc) return rc;
    for (int64_t i = 0; i < nops; i++) {
        if (i % L.S == 0) {
            int64_t g = i / L.S;
            ck_n[g] = n;
            TDP_HIP_CHECK(hipMemcpyAsync(d_work + L.off_ck + g * cap16, cur, n, hipMemcpyDeviceToDevice, st));
        }
        if ((rc = run_op(ops + 3 * i, cur, other, n, L.cap, st))) return rc;
    }
    TDP_REQUIRE(n == 1, "op stream ends with %llu entries, not 1", (unsigned long long)n);
    uint8_t v = 0;
    TDP_HIP_CHECK(hipMemcpyAsync(&v, cur, 1, hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    *value = v;
    *nused = 0;
    if (v == 255 || nops == 0) return TDP_OK;       // the caller raises OverflowError / nothing to trace

    int64_t io[4] = {0, (int64_t)v, 0, 0};
    TDP_HIP_CHECK(hipMemcpyAsync(d_io, io, sizeof(io), hipMemcpyHostToDevice, st));
    std::vector<uint64_t> toff(L.S), nb(L.S);
    for (int64_t g = L.nseg - 1; g >= 0; g--) {
        int64_t first = g * L.S, last = first + L.S < nops ? first + L.S : nops;
        n = ck_n[g];
        cur = T0;
        other = T1;
        TDP_HIP_CHECK(hipMemcpyAsync(cur, d_work + L.off_ck + g * cap16, n, hipMemcpyDeviceToDevice, st));
        int64_t k = 0;
        for (int64_t i = first; i < last; i++) {
            nb[i - first] = n;
            toff[i - first] = 0;
            if (ops[3 * i] == 1) {
                toff[i - first] = (uint64_t)k * cap16;
                TDP_HIP_CHECK(hipMemcpyAsync(d_work + L.off_seg + k * cap16, cur, n, hipMemcpyDeviceToDevice, st));
                k++;
            }
            if ((rc = run_op(ops + 3 * i, cur, other, n, L.cap, st))) return rc;
        }
        TDP_HIP_CHECK(hipMemcpyAsync(d_ops, ops + 3 * first, 24 * (last - first), hipMemcpyHostToDevice, st));
        TDP_HIP_CHECK(hipMemcpyAsync(d_toff, toff.data(), 8 * (last - first), hipMemcpyHostToDevice, st));
        TDP_HIP_CHECK(hipMemcpyAsync(d_nb, nb.data(), 8 * (last - first), hipMemcpyHostToDevice, st));
        k_traceback<<<1, 1, 0, st>>>(d_ops, first, last - first, d_work + L.off_seg, d_toff, d_nb, d_io, d_used);
        TDP_LAUNCH_CHECK();
        // the host vectors are reused by the next segment: wait for the copies above
        TDP_HIP_CHECK(hipStreamSynchronize(st));
    }
    TDP_HIP_CHECK(hipMemcpyAsync(io, d_io, sizeof(io), hipMemcpyDeviceToHost, st));
    TDP_HIP_CHECK(hipStreamSynchronize(st));
    TDP_REQUIRE(io[3] == 0, "traceback lost the path (code %lld)", (long long)io[3]);
    TDP_REQUIRE(io[0] == 0 && io[1] == 0, "traceback ended at state %lld value %lld", (long long)io[0],
                (long long)io[1]);
    if (io[2] > 0) TDP_HIP_CHECK(hipMemcpy(used_ops, d_used, 8 * io[2], hipMemcpyDeviceToHost));
    *nused = io[2];
    return TDP_OK;
}

}  // extern "C"
