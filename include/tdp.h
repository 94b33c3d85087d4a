/*
 * tropidist_amd -- C-ABI of the MI355X (gfx950) exact distance evaluators, libtdp.so.
 *
 * Plain pointers and sizes only.  Every pointer named d_* or documented as "device" is HIP device memory owned by
 * the caller (the reference's ownership model: numpy buffers owned by Python, cdist/kernel.py:101-104,
 * cdist/tree_dp.py:435-436, cdist/sparse_dp.py:523-527); `stream` is a hipStream_t (NULL = the null stream).
 * Every entry point returns an int status (TDP_OK = 0); tdp_last_error() describes the last failure of the
 * calling thread.  Nothing here keeps process-global mutable state: calls on different streams and buffers are
 * independent.
 *
 * Each declaration cites the reference interface it replaces (file:line under the reference tree).
 */
#ifndef TDP_H
#define TDP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDP_OK 0
#define TDP_EINVAL 1          /* invalid argument (sizes, masks, malformed op stream) */
#define TDP_EHIP 2            /* a HIP runtime call failed */
#define TDP_ENOMEM 3          /* a device allocation failed (table handles only) */
#define TDP_EDEADLINE 4       /* stopped at the caller's deadline */
#define TDP_ECAPACITY 5       /* a fixed-capacity structure (hash table, output buffer) is full */

/* Message of the calling thread's last failure ("" if none).  No reference counterpart (the reference C returns
 * no error codes, cdist/viterbi.c; Python raises, cdist/kernel.py:18-25). */
const char *tdp_last_error(void);

/* Version string and the number of compute units of the current device (replaces dp_num_threads,
 * cdist/viterbi.c:195-201, as the parallelism report). */
const char *tdp_version(void);
int tdp_num_cus(int *cus);

/* Copy `bytes` of device memory to host memory after the work queued on `stream` and wait for it (the small
 * read-backs of counters and results between launches).  No reference counterpart. */
int tdp_read_back(void *host, const void *d_src, uint64_t bytes, void *stream);

/* ------------------------------------------------------------------ dense trellis, one op per call
 * All act on a uint8 device table d_T of n = 2^b entries, INF = 255, additions saturating at 255. */

/* T[s] <- min(T[s], T[s ^ h] + 1) for all s, in place.  Replaces dp_update (cdist/viterbi.c:72-118). */
int tdp_update(uint8_t *d_T, uint64_t n, uint64_t h, void *stream);

/* Same with + w (w >= 1), saturating.  Replaces dp_update_w (cdist/viterbi.c:203-219). */
int tdp_update_w(uint8_t *d_T, uint64_t n, uint64_t h, int w, void *stream);

/* New top bit flipped by the current column: T[n + (s ^ hl)] = T[s] + w; the buffer holds 2n entries.
 * Replaces dp_expand_update (cdist/viterbi.c:120-138), with a weight. */
int tdp_expand_update(uint8_t *d_T, uint64_t n, uint64_t hl, int w, void *stream);

/* Keep the states whose bit `pos` equals `target`, dropping that bit: d_dst[k] = d_src[insert(k)], k < n/2.
 * d_dst and d_src must not overlap.  *n_out = n/2 (n if n < 2^(pos+1), as the reference).  Out-of-place form of
 * dp_select (cdist/viterbi.c:149-172). */
int tdp_select(uint8_t *d_dst, const uint8_t *d_src, uint64_t n, int pos, int target, uint64_t *n_out,
               void *stream);

/* d_T[0..n) = v.  Replaces dp_fill (cdist/viterbi.c:174-179). */
int tdp_fill(uint8_t *d_T, uint64_t n, uint8_t v, void *stream);

/* d_hist[v] = number of entries equal to v (256 uint64 device counters, overwritten).  Replaces dp_hist
 * (cdist/viterbi.c:181-193). */
int tdp_hist(const uint8_t *d_T, uint64_t n, uint64_t *d_hist, void *stream);

/* ------------------------------------------------------------------ dense trellis, whole plan
 * Op stream: nops triples of uint64 (code, a, b) in host memory: (0, 0, 0) expand, (1, mask, weight) update,
 * (2, pos, target) select -- the op list of cdist/trellis.py:125-182 with an update weight.  The loop of
 * run_plan (cdist/kernel.py:94-138): an expand followed by an update whose top bit is the new bit becomes one
 * expand-update, otherwise the new half is filled with 255.  d_T0 and d_T1 hold 2^peak_bits bytes each (the second
 * receives selects).  *value = the final entry (255 = cost >= 255, the caller raises OverflowError as
 * final_value does, cdist/kernel.py:18-25).  Executed one streaming kernel per op on `stream`. */
int tdp_run_plan(const uint64_t *ops, int64_t nops, int peak_bits, uint8_t *d_T0, uint8_t *d_T1, int *value,
                 void *stream);

/* The same plan in one launch of one workgroup, its tables in LDS, for plans of at most 2^tdp_plan_small_bits()
 * entries.  The host checks the op stream as tdp_run_plan does and groups it into steps: an optional expand, a run
 * of updates with linearly independent masks, an optional select.  Each step runs with one barrier; one thread
 * takes each coset of the run's masks.  The steps are uploaded into d_work (tdp_plan_small_bytes(nops) bytes of
 * device memory) and the result is read back.  The result equals tdp_run_plan's.  On the per-op executor a small
 * plan costs a launch per op, and on the fused one a launch per pass; here it costs one launch.  Counterpart: the
 * loop of run_plan (cdist/kernel.py:94-138) over the ops of cdist/viterbi.c. */
int tdp_plan_small_bits(void);
int tdp_plan_small_bytes(int64_t nops, uint64_t *bytes);
int tdp_run_plan_small(const uint64_t *ops, int64_t nops, int peak_bits, void *d_work, uint64_t work_bytes,
                       int *value, void *stream);

/* The same plan by fused passes: one kernel applies up to five update columns (with the expands and selects of
 * their rows) to register-resident cubes of the table, so a pass reads 2^b_in and writes 2^b_out bytes for all
 * of them.  segments: nseg triples of int64 (0, first, last) = ops [first, last) by the per-op kernels, or (1, k, 0)
 * = pass descriptor k of `passes` (npasses records of tdp_pass_args_bytes() bytes, compiled on the host from the
 * op stream).  The result equals tdp_run_plan's on the same op stream (the loop of run_plan,
 * cdist/kernel.py:94-138, over the ops of cdist/viterbi.c).  TDP_EINVAL if a descriptor addresses outside its
 * tables or does not match the table size the segments before it leave. */
int tdp_pass_args_bytes(void);
/* One pass: d_in holds 2^b_in entries, d_out receives 2^b_out (distinct buffers). */
int tdp_run_pass(const uint8_t *d_in, uint8_t *d_out, const void *pass, void *stream);
int tdp_run_program(const uint64_t *ops, int64_t nops, const int64_t *segments, int64_t nseg, const void *passes,
                    int64_t npasses, int peak_bits, uint8_t *d_T0, uint8_t *d_T1, int *value, void *stream);

/* Witness by fused passes: the program runs as nsteps steps (records of tdp_witness_step_bytes() bytes: a pass of
 * `passes`, or one per-op op, each writing a fresh slot of d_arena so every step's input table is kept), then one
 * wave walks the steps backward and names the updates on one minimum path (the traceback of
 * cdist/auto.py:126-182: a column is used iff the entry before it differs from the value after it).
 * d_arena: tdp_program_witness_bytes(steps) bytes of device memory; used_ops: host array with room for every
 * update of the steps.  *value = the final entry (255 leaves *nused = 0). */
int tdp_witness_step_bytes(void);

/* Many programs at once: each record is a value run (segments + passes on d_T0/d_T1 of 2^peak_bits bytes, as
 * tdp_run_program) or, when steps is not NULL, a witness run (as tdp_run_program_witness, on d_arena with room
 * for used_ops in the host array).  One launch per round runs the next pass of every active program, so programs
 * whose passes are too small to fill the device alone share it; per-op work on tables of at most 64 entries is a
 * one-thread job of a round, larger per-op work is launched between the rounds.  Every run is validated on the host
 * before any device work (TDP_EINVAL).  d_work: at least tdp_run_programs_workspace(runs) bytes (a bound computed
 * from the runs' counts alone); host_staging (optional, NULL): pinned host memory of at
 * least as many bytes, through which the schedule is uploaded.  Outputs: value (the final entry, 255 = cost >= 255),
 * nused / used_ops for witness runs.  Tables of passes must hold at most 2^32 entries.  No reference counterpart
 * (the reference runs one plan at a time, cdist/auto.py:105-123). */
typedef struct {
    const uint64_t *ops;
    int64_t nops;
    const int64_t *segments;
    int64_t nseg;
    const void *passes;
    int64_t npasses;
    const void *steps;
    int64_t nsteps;
    uint8_t *d_T0;
    uint8_t *d_T1;
    uint8_t *d_arena;
    uint64_t arena_bytes;
    int64_t *used_ops;
    int64_t nused;
    int32_t peak_bits;
    int32_t value;
} tdp_run_t;
int tdp_run_record_bytes(void);
int tdp_run_programs_workspace(const tdp_run_t *runs, int nruns, uint64_t *bytes);
int tdp_run_programs(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                     uint64_t staging_bytes, void *stream);
/* tdp_run_programs with HIP events recorded on `stream` around each of the first `cap` multi-program launches:
 * round_us[k] = the launch's duration in microseconds, round_bytes[k] = its algorithmic bytes (sum over its pass
 * jobs of 2^b_in + 2^b_out: every input entry read once, every output entry written once); *nrounds = rounds
 * recorded.  For measurement; the results are those of tdp_run_programs. */
int tdp_run_programs_timed(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *host_staging,
                           uint64_t staging_bytes, void *stream, double *round_us, double *round_bytes, int64_t cap,
                           int64_t *nrounds);

/* The launches of one tdp_run_programs call captured once as a HIP graph and replayed: tdp_programs_graph builds and
 * uploads the schedule (value runs only), records every launch of the call (fills, per-op kernels, rounds, the final
 * gather) with stream capture on `stream` and instantiates it; nothing runs yet.  tdp_graph_launch replays all of
 * it with one launch call; tdp_graph_values waits for the stream and reads the final entries of the value runs, in
 * run order (255 = cost >= 255).  The graph refers to the runs' tables and to d_work, which must outlive it.  For
 * repeated schedules (a benchmark step, a scan certified again): it takes the host's per-launch cost, one launch
 * call per round, off the critical path.  No reference counterpart.  tdp_graph_nodes: the node count. */
void *tdp_programs_graph(tdp_run_t *runs, int nruns, uint8_t *d_work, uint64_t work_bytes, void *stream);
int tdp_graph_launch(void *graph, void *stream);
int tdp_graph_values(void *graph, int32_t *values, int n, void *stream);
int tdp_graph_nodes(void *graph, uint64_t *nodes);
void tdp_graph_free(void *graph);

/* Host compiler of fused programs (no device work): the segments (capacity seg_cap triples), pass descriptors
 * (capacity pass_cap records) and, when steps is not NULL, the witness steps (capacity step_cap records) of an
 * encoded op stream.  dmax <= 5 directions per pass, tables below 2^min_bits (>= 2) run per op, max_updates > 0
 * caps the updates of a pass (the witness mode uses min_bits 2 and max_updates 6).  TDP_EINVAL when a column
 * exceeds the pass limits on its own (callers then use tdp_run_plan).  No reference counterpart. */
int tdp_fused_compile(const uint64_t *ops, int64_t nops, int dmax, int min_bits, int max_updates, int64_t *segments,
                      int64_t seg_cap, int64_t *nseg, void *passes, int64_t pass_cap, int64_t *npasses, void *steps,
                      int64_t step_cap, int64_t *nsteps);
int tdp_program_witness_bytes(const void *steps, int64_t nsteps, uint64_t *bytes);
int tdp_run_program_witness(const uint64_t *ops, int64_t nops, const void *steps, int64_t nsteps,
                            const void *passes, int64_t npasses, uint8_t *d_arena, uint64_t arena_bytes, int *value,
                            int64_t *used_ops, int64_t *nused, void *stream);

/* Witness of a plan: the plan's value and the op indices of the updates on one minimum path, by a checkpointed
 * traceback from the final state 0 (replaces the numpy replay of dp_witness, cdist/auto.py:126-182, which keeps
 * every table).  d_work: caller workspace of tdp_witness_workspace_bytes(nops, peak_bits) bytes; used_ops: host
 * array of nops entries.  *value = 255 leaves *nused = 0 (the caller raises OverflowError). */
int tdp_witness_workspace_bytes(int64_t nops, int peak_bits, uint64_t *bytes);
int tdp_run_plan_witness(const uint64_t *ops, int64_t nops, int peak_bits, uint8_t *d_work, uint64_t work_bytes,
                         int *value, int64_t *used_ops, int64_t *nused, void *stream);

/* ------------------------------------------------------------------ tree-realization DP joins
 * Tables of unsigned integers of itemsize 1, 2, 4 or 8 bytes in device memory; lm / rm: host arrays of the
 * `generators` uint64 generator maps.  For the Gray-code counters i in [start, stop): g = i ^ (i >> 1) selects
 * generators; out[g & (2^parent_bits - 1)] = min(out, L[XOR lm] + R[XOR rm]) when both costs are below `infinity`
 * and the sum does not exceed it.  [0, 2^generators) runs as one atomic-free launch.  Replaces
 * tree_join_u8/u16/u32/u64 (cdist/tree_dp.c:13-44). */
int tdp_tree_join(int itemsize, const void *d_left, const void *d_right, void *d_out, const uint64_t *lm,
                  const uint64_t *rm, unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop,
                  uint64_t infinity, void *stream);
int tdp_tree_join_u8(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm, const uint64_t *rm,
                     unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop, uint64_t infinity,
                     void *stream);
int tdp_tree_join_u16(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm, const uint64_t *rm,
                      unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop, uint64_t infinity,
                      void *stream);
int tdp_tree_join_u32(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm, const uint64_t *rm,
                      unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop, uint64_t infinity,
                      void *stream);
int tdp_tree_join_u64(const void *d_left, const void *d_right, void *d_out, const uint64_t *lm, const uint64_t *rm,
                      unsigned parent_bits, unsigned generators, uint64_t start, uint64_t stop, uint64_t infinity,
                      void *stream);

/* The tiled form tdp_tree_join gives a full join of 1- or 2-byte items (infinity <= 32767, 4..14 kernel generators
 * with injective left and right maps, some reuse): *tiles = the number of tiles (0 when the join runs on the
 * gather kernel instead), args = the tile description (tdp_tree_tile_args_bytes() bytes; layout in
 * csrc/tdp_tree.hip TileArgs).  The launches keep at least 2^10 tiles (min_tiles_log = 10); smaller values give the
 * plan of smaller joins.  Host only, for checking the plan; no counterpart in the reference. */
int tdp_tree_tile_plan(const uint64_t *lm, const uint64_t *rm, unsigned parent_bits, unsigned generators,
                       int itemsize, uint64_t infinity, int min_tiles_log, void *args, uint64_t *tiles);
int tdp_tree_tile_args_bytes(void);

/* Postorder arena compaction: move `bytes` toward a lower address, source and destination may overlap.
 * Replaces tree_move (cdist/tree_dp.c:46-51). */
int tdp_tree_move(void *d_dst, const void *d_src, size_t bytes, void *stream);

/* Leaf and fill initialisation of arena tables (the numpy fills of cdist/tree_dp.py:446-454): out[0] = v0,
 * out[1] = v1 (when size > 1), the rest = fill; tdp_tree_fill sets all `size` items to `value`. */
int tdp_tree_leaf(int itemsize, void *d_out, uint64_t size, uint64_t v0, uint64_t v1, uint64_t fill, void *stream);
int tdp_tree_fill(int itemsize, void *d_out, uint64_t size, uint64_t value, void *stream);

/* Compiler and runtime of a whole tree realisation (the compilation and run of cdist/tree_dp.py:245-505; the reference
 * does both in Python around its C joins).  cols: ncols columns of rw uint64 words; nodes in postorder: node_leaf[i]
 * = column index or -1, node_left[i] / node_right[i] = child node indices (earlier nodes); every column a leaf once.
 * The compiler finds each node's interface im(M_S) /\ im(M_rest), each join's generator maps and the child order
 * of least peak live entries, and checks J = rank(M_A) + rank(M_B) + rank(M_rest) - rank(M) at every join.
 * Returns a handle, NULL on failure with *status TDP_EINVAL (malformed tree, broken identity) or TDP_EDEADLINE
 * (max_seconds >= 0 exceeded between nodes). */
void *tdp_tree_compile(const uint64_t *cols, int ncols, int rw, const int32_t *node_leaf, const int32_t *node_left,
                       const int32_t *node_right, int nnodes, double max_seconds, int *status);
void tdp_tree_free(void *program);
/* out[0] rank(M), [1] peak live table entries of the execution order, [2] host bytes of the program, [3] largest
 * interface rank, [4] largest join dimension J, [5] 1 if every join's maps fit 64-bit words (ranks <= 63, J <= 62). */
int tdp_tree_info(void *program, int64_t *out);
/* Per node: interface rank, the execution order (node ids), the child joined first (-1 at leaves) and four join
 * numbers (left rank, right rank, parent rank, J; -1 at leaves). */
int tdp_tree_nodes(void *program, int32_t *rank, int32_t *order, int32_t *first, int32_t *joins);
/* A join's J generator maps (parent generators first), as tdp_tree_join takes them. */
int tdp_tree_maps(void *program, int node, uint64_t *lm, uint64_t *rm);
/* A node's interface basis: rank vectors of rw words. */
int tdp_tree_interface(void *program, int node, uint64_t *vecs);
/* A particular solution e0 (ceil(ncols / 64) words) of M e0 = target (rw words); *feasible = 0 when the target is
 * outside the column span (the run's None). */
int tdp_tree_solve(void *program, const uint64_t *target, uint64_t *e0, int *feasible);
/* The DP for the target whose particular solution is e0, on the caller's device arena of at least the peak live
 * entries of itemsize bytes (the smallest unsigned width holding ncols + 1; infinity = ncols + 1): leaf tables, the
 * joins in Gray-counter chunks of chunk_pairs, the stack compaction, all on `stream`.  *value = the root entry;
 * stats[0] = join pairs enumerated, stats[1] = the largest live entries seen.  TDP_EDEADLINE when max_seconds >= 0
 * passes before the last launch or before the result is read. */
int tdp_tree_run(void *program, const uint64_t *e0, int itemsize, void *d_arena, uint64_t arena_items,
                 uint64_t chunk_pairs, double max_seconds, uint64_t *value, uint64_t *stats, void *stream);

/* ------------------------------------------------------------------ sparse (weight-truncated) trellis DP
 * A device table is n states of 16 bytes {lo, hv}: key = (hv & (2^56 - 1)) << 64 | lo (up to 120 row slots),
 * value = hv >> 56 (the reference's packed key | value << 56 widened from 56 to 120 key bits,
 * cdist/sparse_dp.c:3-4), with distinct keys in any order, and a hash index d_slots of uint32 entry numbers
 * (capacity a power of two >= 2n).  128-bit masks are passed as host arrays {lo, hi}. */

/* Build the index of a table.  No reference counterpart (the reference keeps tables sorted, cdist/sparse_dp.c:51-114). */
int tdp_sp_index(const void *d_ents, uint64_t n, uint32_t *d_slots, uint64_t capacity, void *stream);

/* One column step T'[s] = min(T[s], T[s ^ upd] + 1) with the closing-row filter, cap, keep_below and lookahead
 * bound (d_cls_mask: ncls {lo, hi} masks, d_cls_w: ncls weights, device).  Writes the new table to d_out (capacity
 * out_cap >= 2n) and, if d_ch is not NULL, the new or improved states to d_ch (capacity ch_cap >= 2n); device
 * counters: [0] table size, [1] changed size, [2] overflow flag.  Replaces sp_candidates + sp_merge
 * (cdist/sparse_dp.c:116-202). */
int tdp_sp_step(const void *d_old, uint64_t n, const uint32_t *d_slots, uint64_t capacity, const uint64_t *upd,
                const uint64_t *sel_mask, const uint64_t *sel_bits, int cap, int keep_below, int ncls,
                const uint64_t *d_cls_mask, const int64_t *d_cls_w, int64_t extra, const uint64_t *tmask, int K,
                int use_h, void *d_out, uint64_t out_cap, void *d_ch, uint64_t ch_cap,
                unsigned long long *d_counters, void *stream);

/* d_vals[i] = value of key {d_keys[2i], d_keys[2i+1]} (255 if absent).  The lookup of NumpyTables/CTables.lookup
 * (cdist/sparse_dp.py:329-336, 557-565), batched. */
int tdp_sp_lookup(const void *d_ents, const uint32_t *d_slots, uint64_t capacity, const uint64_t *d_keys, uint64_t m,
                  uint8_t *d_vals, void *stream);

/* States with value < below into d_out; *d_count = their number.  Replaces CTables.truncate
 * (cdist/sparse_dp.py:533-535). */
int tdp_sp_truncate(const void *d_ents, uint64_t n, int below, void *d_out, unsigned long long *d_count,
                    void *stream);

/* Least value(f) + value(b) over f in d_f and b in the indexed backward table with key(b) = key(f) ^ tmask:
 * d_best = {total, key hi, key lo, forward value} of the least total with the least forward key, total = 2^64-1
 * if none.  Replaces sp_join (cdist/sparse_dp.c:204-236). */
int tdp_sp_join(const void *d_f, uint64_t nf, const void *d_bents, const uint32_t *d_bslots, uint64_t bcapacity,
                const uint64_t *tmask, unsigned long long *d_best, void *stream);

/* d_hist[v] = number of states of value v (256 device counters).  Replaces sp_hist (cdist/sparse_dp.c:238-242). */
int tdp_sp_hist(const void *d_ents, uint64_t n, uint64_t *d_hist, void *stream);

/* ------------------------------------------------------------------ dense tables of the sparse sweeps' fallback
 * A dense table holds every state of a cut's active rows: d_T[i] (uint8, 255 = absent, 2^w entries, 4-byte aligned)
 * for the compact index i of a state, bit k of i = the state's bit at slot slots[k] (w <= 40 slots, ascending,
 * host array).  The Dense class of cdist/sparse_dp.py:221-289. */

/* d_T = 255 everywhere, then d_T[compact(key)] = value for each of the n states (their bits lie in the slots).
 * The sparse-to-dense switch of cdist/sparse_dp.py:698-705. */
int tdp_dn_from_sparse(const void *d_ents, uint64_t n, const uint8_t *slots, int w, uint8_t *d_T, void *stream);

/* The entries of d_T other than 255 as states into d_out (capacity cap, any order); device counters [0] = their
 * number, [1] = 1 if it exceeded cap.  The dense-to-sparse switch of cdist/sparse_dp.py:694-697. */
int tdp_dn_to_sparse(const uint8_t *d_T, int w, const uint8_t *slots, void *d_out, uint64_t cap,
                     unsigned long long *d_counters, void *stream);

/* One column step of a dense table (cdist/sparse_dp.py:680-693): the rows starting at the column take new bits of
 * the "mid" space (w_mid = w_old + nstart bits; their half of it absent), T'[x] = min(T[x], T[x ^ cm] + 1) there,
 * the rows ending at the column are fixed at their target bit and removed, and states above cap, or (use_h) whose
 * value plus ceil((extra + sum_c cls_w[c] |(i ^ tmask) & cls_mask[c]|) / 720720) exceeds K, are dropped.
 * start_pos: mid positions of the starting rows, descending; end_pos / end_bit: mid positions of the ending rows,
 * ascending, and their targets; cls_mask (ncls <= 64) and tmask in the new table's compact space (host arrays).
 * d_U: 2^(w_mid - nend) bytes; *d_count (device) = its entries other than 255. */
int tdp_dn_step(const uint8_t *d_T, int w_old, uint8_t *d_U, int w_mid, const uint8_t *start_pos, int nstart,
                const uint8_t *end_pos, const uint8_t *end_bit, int nend, uint64_t cm, int cap, int K, int ncls,
                const uint64_t *cls_mask, const int64_t *cls_w, int64_t extra, uint64_t tmask, int use_h,
                unsigned long long *d_count, void *stream);

/* tdp_sp_join against a dense backward table: d_best = {total, key hi, key lo, forward value} of the least
 * value(f) + T[compact(key(f) ^ tmask)] with the least forward key.  NumpyTables.join over a Dense table
 * (cdist/sparse_dp.py:420-444). */
int tdp_dn_join(const void *d_f, uint64_t nf, const uint8_t *d_T, int w, const uint8_t *slots, const uint64_t *tmask,
                unsigned long long *d_best, void *stream);

/* ------------------------------------------------------------------ one sparse query, natively
 * A whole query of cdist/sparse_dp.py:985-1095 (mode TDP_SQ_MITM: backward sweep, forward frontier with joins at
 * every cut, traceback) or one forward sweep (TDP_SQ_FORWARD, weights <= K, the value of state 0 at cut N), run by
 * the library on `stream` with its own device tables (stream-ordered allocations, all freed before it returns).
 * The column words are those of tdp_spb_add: prm per position 10 words (column slots, forward closing mask and
 * bits, backward closing mask and bits, each (low 64, high 56 bits)), tm per cut 2 words.  split: the forward
 * share a(c) per cut (N + 1 values, non-increasing, 1..K); the backward sweep keeps weights <= K - a(c).  With
 * use_h, the lookahead bound of each direction ([0] forward, [1] backward) per cut c: bnd_ncls[d][c] classes from
 * index bnd_off[d][c] of bnd_mask[d] (2 words each) and bnd_w[d], and bnd_extra[d][c] (tdp_sp_bounds' layout).
 * Tables are kept at every cut while their bytes stay within keep_bytes, otherwise at checkpoints every `every`
 * cuts, and recomputed from the kept table before them when a join or the traceback needs them.  lower: a join
 * total <= max(lower, 0) ends the forward sweep.  Profile modes sweep one direction keeping weights <= K (no
 * lookahead) and write hist[c * (hist_top + 1) + w] = the states of value w at cut c.  Outputs: per-cut table
 * counts of each direction and changed counts of the frontier (N + 1 each, may be NULL), the witness's positions
 * (room for K + 1), and the result record; a state budget or deadline stops the query with that status.  K <= 254 for
 * one sweep, K <= 508 for the meet in the middle (a(c) <= 254 and K - a(c) <= 254).  The
 * reference's Sweep/query run this in Python around one C call per column (cdist/sparse_dp.py:621-1095).
 * dev_alloc / dev_free (optional, both or neither): the caller's device allocator for every buffer of the query,
 * called on the calling thread with alloc_ctx; dev_alloc returns NULL when it has no memory (the query then fails
 * with TDP_ENOMEM).  A buffer is freed only after the work on `stream` that used it has completed.  Without them
 * the library allocates stream-ordered from the device's default pool. */
#define TDP_SQ_FORWARD 0
#define TDP_SQ_MITM 1
#define TDP_SQ_PROFILE_FW 2
#define TDP_SQ_PROFILE_BW 3
#define TDP_SQ_DONE 0
#define TDP_SQ_STATES 1       /* a table or changed list exceeded max_states at fail_cut (BudgetExceeded "states") */
#define TDP_SQ_DEADLINE 2     /* max_seconds passed at fail_cut (BudgetExceeded "deadline") */
typedef struct {
    int32_t N, K, mode, use_h, witness, lower, every, hist_top;
    uint64_t max_states;      /* 0: no limit */
    double max_seconds;       /* < 0: no limit */
    uint64_t keep_bytes;
    const uint64_t *prm, *tm;
    const int32_t *split;
    const int32_t *bnd_ncls[2];
    const int64_t *bnd_off[2];
    const uint64_t *bnd_mask[2];
    const int64_t *bnd_w[2];
    const int64_t *bnd_extra[2];
    uint64_t *hist;
    void *(*dev_alloc)(uint64_t bytes, void *ctx);
    void (*dev_free)(void *ptr, void *ctx);
    void *alloc_ctx;
} tdp_sq_spec_t;
typedef struct {
    int32_t status, fail_cut, value, join_cut, nwitness, pad;
    uint64_t steps, joined, peak_states, peak_bytes;
    double seconds, backward_seconds, forward_seconds, traceback_seconds;
} tdp_sq_result_t;
int tdp_sq_run(const tdp_sq_spec_t *spec, uint64_t *fw_counts, uint64_t *bw_counts, uint64_t *ch_counts,
               int32_t *witness, tdp_sq_result_t *result, void *stream);

/* The lookahead data of one query's trellis (the bound of cdist/sparse_dp.py:148-192), per cut c: ncls[c] classes
 * at masks + 2 * (c * cap + k) (slot masks, low and high word) with weights + c * cap + k (ascending), and extra[c]
 * (the target row's weight while it lies wholly on the remaining side).  forward selects the direction; targets as
 * tdp_sp_trellis for one query; cap >= the trellis' slot count. */
int tdp_sp_bounds(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                  int forward, int cap, int32_t *ncls, int64_t *extra, uint64_t *masks, int64_t *weights);

/* ------------------------------------------------------------------ many sparse queries at once
 * A handle holding many meet-in-the-middle queries resident on one device (the query of
 * cdist/sparse_dp.py:985-1095 with a uniform split, its column steps and joins from cdist/sparse_dp.c:116-236).
 * tdp_spb_step advances every active query by one column in a few launches and one read-back, so queries admitted
 * between steps share the device (continuous batching).  No reference counterpart: the reference runs one query
 * at a time from Python (cdist/sparse_dp.py:621-820).  Every table stays on the device, with its index, until the
 * query is released; memory comes from the handle's arena of arena_bytes.  Calls on one handle are serialised;
 * all device work goes to `stream`. */
#define TDP_SPB_RUNNING 0
#define TDP_SPB_DONE 1        /* finished: value (or none) known */
#define TDP_SPB_STATES 2      /* a table or changed list exceeded max_states (BudgetExceeded "states") */
#define TDP_SPB_NOMEM 3       /* the arena could not hold the query's tables */
#define TDP_SPB_CANCELLED 4   /* stopped by tdp_spb_cancel */
#define TDP_SPB_FAILED 5      /* a device-side check failed for this job (tdp_last_error); the others go on */

/* NULL on failure (tdp_last_error). */
void *tdp_spb_create(uint64_t arena_bytes, int max_jobs, void *stream);
int tdp_spb_free(void *h);

/* A query of N columns: prm holds per column p ten words -- the column's row-slot mask, the forward closing mask
 * and bits, the backward closing mask and bits, each as (low 64 bits, high 56 bits) -- and tmask the target mask
 * per cut (N + 1 pairs).  The backward sweep keeps weights <= K - split, the forward frontier weights
 * <= split - 1; max_states 0 means no limit.  Writes the query's id. */
int tdp_spb_add(void *h, const uint64_t *prm, const uint64_t *tmask, int32_t N, int32_t K, int32_t split,
                uint64_t max_states, int64_t *id);

/* One column of every active query; *active = queries not finished. */
int tdp_spb_step(void *h, int32_t *active);

/* Ids of the queries that finished (done or failed) since the last call, oldest first, at most cap of them. */
int tdp_spb_poll(void *h, int64_t *ids, int32_t cap, int32_t *n);

/* Arena and progress counters: out[0] arena bytes, [1] bytes in use, [2] most bytes in use so far, [3] queries
 * that found the arena full (TDP_SPB_NOMEM), [4] engine steps so far, [5] queries running, [6] queries that found
 * the arena full at the last step. */
int tdp_spb_stats(void *h, uint64_t *out);

/* Status (TDP_SPB_*), the least weight <= K (-1 if none or not finished) and, if stats is not NULL, four counters:
 * states summed over cuts, the largest table, the join cut, the most arena bytes the query held. */
int tdp_spb_result(void *h, int64_t id, int32_t *status, int32_t *value, uint64_t *stats);

/* Column positions (indices into the query's column order) of a solution of the query's value, traced back through
 * its kept tables on the device. */
int tdp_spb_witness(void *h, int64_t id, int32_t *positions, int32_t cap, int32_t *n);

/* Frees a finished query's tables and id. */
int tdp_spb_release(void *h, int64_t id);

/* Stops a running query between steps (status TDP_SPB_CANCELLED, its tables freed; it is reported by the next poll);
 * a finished query is left as it is.  For callers that give a request up, e.g. when one of its queries gave its
 * memory up. */
int tdp_spb_cancel(void *h, int64_t id);

/* ------------------------------------------------------------------ meet-in-the-middle lookup tables
 * Open addressing over 2^log2slots device slots; all subsets of size <= k of n columns keyed by the 128-bit
 * XOR-linear syndrome hash (zh, zl) with their logical label lg.  The table owns its device memory behind the
 * handle (as the reference's Table, cdist/gmitm.c:36-84). */

/* NULL on failure (the caller raises MemoryError, cdist/search.py:127-129).  Replaces tab_create (gmitm.c:61-76). */
void *tdp_tab_create(int log2slots, int wide);
/* Replaces tab_free (gmitm.c:78-82). */
void tdp_tab_free(void *table);
/* Number of distinct keys (compact: fingerprint and 16-bit label; wide: fingerprint and label), -1 on error.
 * Replaces tab_count (gmitm.c:84). */
int64_t tdp_tab_count(void *table);
/* Insert the empty set and every subset of size 1..k (k <= 8) of the n columns given by host arrays zh, zl, lg
 * (gmitm.c:165-186), on the device. */
int tdp_tab_build(void *table, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k);
/* tdp_tab_build without any device allocation, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same inserts, same table
 * contents, same null-stream synchronous behaviour as tdp_tab_build; the columns and the wide-table retry lists live
 * in a static device array of each GPU (about 4.3 MiB, held until the process ends) instead of memory allocated per
 * call.  Limits: n <= 4096 columns (TDP_EINVAL beyond), at most 65,536 deferred wide inserts per retry round
 * (TDP_ECAPACITY beyond; tdp_tab_build allows 2^20).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (from several host threads) race and must be serialised by the caller. */
int tdp_tab_build_noalloc(void *table, int n, const uint64_t *zh, const uint64_t *zl, const uint64_t *lg, int k);
/* For m device queries: d_out[q] = least cost <= d_R[q] of an entry with the query's fingerprint and a different
 * label, -1 if none.  Batched form of tab_lookup (gmitm.c:136-163, static in the reference). */
int tdp_tab_lookup_batch(void *table, const uint64_t *d_zh, const uint64_t *d_zl, const uint64_t *d_lg,
                         const int *d_R, uint64_t m, int *d_out, void *stream);

/* ------------------------------------------------------------------ weight-bounded search
 * The anchored canonical search of cdist/gmitm.c:188-1023 on the device, with tdp_tab handles as its tables.  The
 * options of the next searches are set as in the reference (its gm_set_* globals, here per host thread); every
 * pointer they keep must stay valid until the search that uses them returns. */

/* Replaces gm_set_cover (gmitm.c:248-250): gate4 > 0 tests the covering bound at nodes with 4|sigma| >= gate4 * R;
 * ascent: one greedy dual-ascent pass. */
void tdp_gm_set_cover(int gate4, int ascent);
/* Replaces gm_set_kernel (gmitm.c:242-244): zero-syndrome zero-logical column sets (CSR), nks = 0 turns it off. */
void tdp_gm_set_kernel(int nks, const int *ks_start, const int *ks_cols, const int *kc_start, const int *kc_ids);
/* Replaces gm_set_kernel_mode (gmitm.c:246): 1 records zero-syndrome sets with zero logical instead of solutions. */
void tdp_gm_set_kernel_mode(int on);
/* Replaces gm_set_full_labels (gmitm.c:240): label-free mode, the logical value of a complete set from lg (N x lw
 * words); NULL turns it off. */
void tdp_gm_set_full_labels(const uint64_t *lg, int lw);
/* Replaces gm_set_pattern (gmitm.c:252-254): class per column (-1: none), on-pattern flag per class, slack. */
void tdp_gm_set_pattern(const int *cls, const uint8_t *onpat, int nclass, int slack);
/* Replaces gm_num_threads (gmitm.c:1025-1031): nodes the device expands at once. */
int tdp_gm_num_threads(void);
/* Replaces gm_search (gmitm.c:857-1023); same arguments, host arrays, tabs[] are tdp_tab handles.  *status gets the
 * reference's return code: 0 exhausted, 1 stopped at the first solution (enum_all = 0), 2 time limit, 4 K > 40.
 * Codes 3 (too many tasks), 5 and 6 (task ledger) never occur: the device search keeps no task list and no ledger.
 * nodes_out[0..40]: nodes visited per depth; stats_out[8]: lookups, hits, false hits, nodes at depth task_depth that
 * reach their expansion (the reference's tasks), cover prunes, ascent prunes, cover tests, solutions found.  Nodes
 * per depth, hits, false hits, tasks and prunes equal the reference's on an exhausted run; lookups and cover tests
 * count every probe and test once (the reference counts its tasks' twice and skips probes its prefix phase does not
 * make).  The returned int is TDP_OK unless the arguments are invalid or the device fails.  The device buffers a
 * search needs are kept for the next one (tdp_gm_release frees them); searches on one device run one at a time. */
/* Nodes expanded per device launch of the next searches (0: the default, 2^19).  No counterpart in the reference:
 * a tuning knob of the device search. */
void tdp_gm_set_chunk(int nodes);
/* Device launches of the calling thread's last tdp_gm_search.  No counterpart in the reference. */
int tdp_gm_last_launches(void);
/* Frees the device buffers the searches keep between calls on the current device (node levels, scratch, uploads).
 * No counterpart in the reference. */
void tdp_gm_release(void);
int tdp_gm_search(int W, int N, int M, int wmax, const uint64_t *col, const uint64_t *lg, const int *row_start,
                  const int *row_cols, const int *col_start, const int *col_rows, int ntab, void **tabs,
                  const uint64_t **tzh, const uint64_t **tzl, int main_tab, const int *rank, int n_anchor,
                  const int *anchors, const int *ex_start, const int *ex_cols, int K, int enum_all, int task_depth,
                  double time_limit, int64_t *nodes_out, int64_t *stats_out, int *best, int sol_cap, int sol_stride,
                  int *sol_buf, int *nsol, int *status);

/* ------------------------------------------------------------------ batched dense plans
 * d_items: device records of tdp_batch_item_bytes() bytes (layout in tropidist_amd/csrc/tdp_batch.hip: op stream,
 * peak bits, checkpoint interval, table / checkpoint / decision-map / result / barrier pointers, group size);
 * d_tickets: ntickets device pairs {item, member}, the `group` members of an item consecutive.  Each workgroup of
 * one persistent launch takes the next ticket from the zeroed counter d_next and runs its share of that plan, the
 * members of a plan meeting at a group barrier between ops; with the witness when asked (checkpointed traceback,
 * one decision bit per state and update).  Result records and barrier words must be zeroed by the caller;
 * `workgroups` (resident at once) must be >= the largest group.  No reference counterpart: the reference runs
 * plans one at a time (cdist/auto.py:105-123); this serves the batched scans and the serving loop. */
int tdp_batch_run(void *d_items, const void *d_tickets, int ntickets, int *d_next, int workgroups, void *stream);
int tdp_batch_item_bytes(void);

/* ------------------------------------------------------------------ host planner
 * Dense DP plans of nq queries sharing one check matrix, in C++ (the steps of cdist/trellis.py:47-193).  h_words:
 * ncols columns of ceil(nrows / 64) uint64 words; order: the column order; targets: nq x ncols bytes, 1 where the
 * column flips the query's logical parity; weights: per column (NULL = unit).  Per query q: ops + 3 * q * ops_cap
 * receives the op stream (code, a, b), columns + q * ops_cap the column position of each update, and nops[q],
 * peak[q], work[q] (sum over updates of the table size), feasible[q].  ops_cap >= ncols + 2 * (nrows + 1). */
int tdp_plan_build(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                   int nq, const int32_t *weights, uint64_t *ops, int64_t ops_cap, int64_t *nops, int32_t *peak,
                   double *work, int32_t *columns, int32_t *feasible);

/* r[i] = GF(2) rank of the first i columns of `order` (i = 0..n; out has n + 1 entries); columns as h_words, ncols
 * bitsets of rw uint64 words.  Replaces the rank profile of cdist/trellis.py:20-36 (cut ranks, best_order). */
int tdp_rank_profile(const uint64_t *h_words, int ncols, int rw, const int32_t *order, int n, int32_t *out);

/* A column order with few open rows at every cut: greedy over the columns sharing a row with those already taken
 * (fewest newly opened rows net of closed ones, then most open rows touched, then lowest index).  No reference
 * counterpart: the reference's orders come from cdist/ordering.py:36-179; any order gives the same exact value. */
int tdp_greedy_order(const uint64_t *h_words, int ncols, int nrows, int start, int32_t *order);

/* The sparse trellis of nq queries in the word layout of tdp_spb_add (prm: nq x ncols x 10 words, tm:
 * nq x (ncols + 1) x 2): rows in minimal span form, the reduced target row, slots by interval colouring, per
 * position the slot mask and closing masks / bits of both directions, per cut the target mask.  Replaces the
 * construction of cdist/sparse_dp.py:63-145 (Trellis.__init__).  feasible[q] = 0 for a target that reduces to
 * zero; nslots[q] > max_slots leaves the query's words unwritten. */
int tdp_sp_trellis(const uint64_t *h_words, int ncols, int nrows, const int32_t *order, const uint8_t *targets,
                   int nq, int max_slots, uint64_t *prm, uint64_t *tm, int32_t *nslots, int32_t *feasible);

#ifdef __cplusplus
}
#endif

#endif /* TDP_H */
