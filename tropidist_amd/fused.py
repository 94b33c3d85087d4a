"""Fused-pass compiler of the dense DP: many ops per pass over the table in HBM.

A pass is a run of whole columns (each column: the expands of the rows starting there, its update, the selects of
the rows ending there).  Updates are min-plus convolutions over GF(2)^b and commute with each other and with the
expands / selects of rows they do not touch, so the whole run can be applied to register-resident "cubes":

  variables   the input table's bits (var i = input bit i) and one new variable per row starting in the pass;
  cube basis  e_0, e_1 (the two low input bits: one dword of four uint8 entries) and the update masks of the pass
              with their bit-0/1 parts removed (m'_j), linearly independent -- D of them, so a cube has 2^(D+2)
              entries held as 2^D dwords;
  threads     one per cube: a representative x with the pivot variables of the basis and bits 0, 1 zero, indexed by
              the remaining ("free") variables, the lowest input bits first so that neighbouring lanes read
              neighbouring dwords;
  expands     a new variable's entries with value 1 start at 255 (not loaded);
  selects     an ending row's variable must equal its target for an entry to be stored, and leaves the output
              layout (the higher output bits move down, the reference's compaction);
  updates     update j pairs cube element q with q ^ (1 << j), entries inside the dword permuted by the mask's
              bit-0/1 part; updates with no bit outside bits 0, 1 act inside each dword.

Every input entry is read once and every output entry written once per pass: a pass moves 2^b_in + 2^b_out bytes
for D updates.  Tables narrower than `min_bits` keep the per-op kernels (their ops are latency-bound).

compile_program(plan) -> Program (segments of per-op ranges and passes); simulate(program, plan) executes the
passes on the CPU with the kernel's exact index arithmetic (used by the tests)."""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .trellis import OP_EXPAND, OP_SELECT, OP_UPDATE

DMAX = 5
MAX_FREE = 48
MAX_LOW = 8
MAX_ROWS = 32                   # rows starting / ending in one pass (32-bit coordinate masks)
MAX_RUNS = 12                   # runs of consecutive free variables per index map (see address_runs)
MAX_VARS = 64                   # input bits plus opened rows of one pass (64-bit variable masks)


@dataclass
class Pass:
    first: int                  # op range [first, last)
    last: int
    b_in: int
    b_out: int
    D: int
    din: list                   # per basis vector (K = D + 2): input index delta
    dout: list                  # output index delta
    dnew: list                  # new-variable coordinate delta (bitmask over new vars)
    dsel: list                  # selected-variable coordinate delta (bitmask over selected vars)
    fin: list                   # per free variable (thread bit order)
    fout: list
    fnew: list
    fsel: list
    sel_target: int
    lam: list                   # per update direction j: bits 0/1 of its mask
    w: list
    low_lam: list = field(default_factory=list)     # updates acting inside each dword
    low_w: list = field(default_factory=list)
    byte_store: bool = False
    # for the traceback: every update of the pass in op order as (variable mask, weight, op index), the new and
    # the ending rows' variables (with the targets), and the variable held by each output bit
    upd_ops: list = field(default_factory=list)
    new_mask: int = 0
    fix_mask: int = 0
    fix_val: int = 0
    out_vars: list = field(default_factory=list)

    @property
    def nfree(self):
        return len(self.fin)

    @property
    def bytes_moved(self):
        return (1 << self.b_in) + (1 << self.b_out)

    def runs(self):
        return [address_runs(v) for v in (self.fin, self.fout, self.fnew, self.fsel)]


def address_runs(vals):
    """A thread's index map as runs: vals[f] is a single bit (or 0) per thread-index bit f, and the map is the OR of
    ((tid >> f0) & (2^len - 1)) << shift over runs [f0, len, shift] of consecutive f with consecutive bits."""
    out = []
    for f, v in enumerate(vals):
        v = int(v)
        if not v:
            continue
        assert v & (v - 1) == 0, "index maps of free variables are single bits"
        sh = v.bit_length() - 1
        if out and out[-1][0] + out[-1][1] == f and out[-1][2] + out[-1][1] == sh:
            out[-1][1] += 1
        else:
            out.append([f, 1, sh])
    return out


@dataclass
class Program:
    segments: list              # ("ops", first, last) | ("pass", Pass)
    peak_bits: int

    @property
    def passes(self):
        return [s[1] for s in self.segments if s[0] == "pass"]


def _columns(enc):
    """Split an encoded op stream into columns: [expands*, update?, selects*] op index ranges."""
    cols, i, n = [], 0, len(enc)
    while i < n:
        j = i
        while j < n and enc[j, 0] == OP_EXPAND:
            j += 1
        if j < n and enc[j, 0] == OP_UPDATE:
            j += 1
        while j < n and enc[j, 0] == OP_SELECT:
            j += 1
        if j == i:              # a select without an update before it (cannot happen in plans): own column
            j = i + 1
        cols.append((i, j))
        i = j
    return cols


class _Basis:
    """GF(2) elimination over variable bitmasks, for independence tests."""

    def __init__(self):
        self.rows = {}          # pivot bit -> reduced vector

    def reduce(self, v):
        while v:
            t = v.bit_length() - 1
            r = self.rows.get(t)
            if r is None:
                return v
            v ^= r
        return 0

    def add(self, v):
        v = self.reduce(v)
        self.rows[v.bit_length() - 1] = v
        return v


def _build_pass(enc, first, last, in_log, logical_out):
    """The descriptor of ops [first, last) on a table of 2^b_in entries whose logical position p (the reference's
    bit order, in which the op stream is written) is held by physical bit in_log[p].  The output's physical layout
    keeps every surviving row on its bit and moves the newest rows into the bits freed by ending rows (so the low
    two bits change rarely), or is the logical order when logical_out.  Returns (pass, out_log)."""
    b_in = len(in_log)
    lay = list(in_log)          # variables (= input physical bits, then new ones) in logical order
    nvar = b_in
    new_vars, sel_vars, sel_t = [], [], {}
    upd = []                    # (var mask, weight)
    upd_ops = []                # (var mask, weight, op index)
    for i in range(first, last):
        code, a, b = (int(x) for x in enc[i])
        if code == OP_EXPAND:
            lay.append(nvar)
            new_vars.append(nvar)
            nvar += 1
        elif code == OP_UPDATE:
            m = 0
            for p in range(len(lay)):
                if (a >> p) & 1:
                    m |= 1 << lay[p]
            upd.append((m, b))
            if m:                   # a zero mask changes nothing and is never on a path
                upd_ops.append((m, b, i))
        else:
            u = lay.pop(a)
            sel_vars.append(u)
            sel_t[u] = b
    if logical_out:
        phys = list(lay)
    else:
        # surviving input rows stay on their bits; freed bits (lowest first) take the newest rows
        b_out = len(lay)
        alive = set(lay)
        phys = [None] * b_out
        movers = []
        for v in range(b_in):
            if v in alive:
                if v < b_out:
                    phys[v] = v
                else:
                    movers.append(v)
        newest_first = [v for v in reversed(new_vars) if v in alive]
        free = [k for k in range(b_out) if phys[k] is None]
        for k, v in zip(free, newest_first + movers):
            phys[k] = v
    assert sorted(phys) == sorted(lay)
    out_pos = {v: p for p, v in enumerate(phys)}
    out_log = [out_pos[v] for v in lay]
    low = 0b11
    dirs, lows = [], []
    for m, w in upd:
        mh = m & ~low
        lam = m & low
        if mh:
            dirs.append((mh, lam, w))
        elif m:                     # a zero mask changes nothing
            lows.append((lam, w))
    D = len(dirs)
    # pivots: new variables first, then the highest input bits
    def pref(v):
        return (1 if v >= b_in else 0, v)
    red = [d[0] for d in dirs]
    pivots = []
    for k in range(D):
        v = red[k]
        assert v, "dependent update directions in one pass"
        bits = [x for x in range(nvar) if (v >> x) & 1]
        pv = max(bits, key=pref)
        pivots.append(pv)
        for k2 in range(D):
            if k2 != k and (red[k2] >> pv) & 1:
                red[k2] ^= v
    piv = set(pivots)
    new_idx = {v: k for k, v in enumerate(new_vars)}
    sel_idx = {v: k for k, v in enumerate(sel_vars)}

    def deltas(mask):
        din = dout = dnew = dsel = 0
        x = mask
        while x:
            low_bit = x & -x
            v = low_bit.bit_length() - 1
            x ^= low_bit
            if v < b_in:
                din |= 1 << v
            if v in out_pos:
                dout |= 1 << out_pos[v]
            if v in new_idx:
                dnew |= 1 << new_idx[v]
            if v in sel_idx:
                dsel |= 1 << sel_idx[v]
        return din, dout, dnew, dsel

    basis = [1, 2] + [d[0] for d in dirs]
    dd = [deltas(b) for b in basis]
    free = [v for v in range(nvar) if v not in piv and v not in (0, 1)]
    free.sort(key=lambda v: (1 if v >= b_in else 0, v))
    fd = [deltas(1 << v) for v in free]
    b_out = len(lay)
    return Pass(first, last, b_in, b_out, D,
                [x[0] for x in dd], [x[1] for x in dd], [x[2] for x in dd], [x[3] for x in dd],
                [x[0] for x in fd], [x[1] for x in fd], [x[2] for x in fd], [x[3] for x in fd],
                sum(sel_t[v] << sel_idx[v] for v in sel_vars),
                [d[1] for d in dirs], [d[2] for d in dirs],
                [x[0] for x in lows], [x[1] for x in lows],
                byte_store=not (b_out >= 2 and phys[0] == 0 and phys[1] == 1),
                upd_ops=upd_ops, new_mask=sum(1 << v for v in new_vars), fix_mask=sum(1 << v for v in sel_vars),
                fix_val=sum(sel_t[v] << v for v in sel_vars), out_vars=list(phys)), out_log


def _bits_after(enc, b, first, last):
    for i in range(first, last):
        c = enc[i, 0]
        b += 1 if c == OP_EXPAND else (-1 if c == OP_SELECT else 0)
    return b


def compile_program(plan, dmax=DMAX, min_bits=12, max_updates=None):
    """Segments of a feasible plan: per-op ranges while the table is narrower than 2^min_bits, passes of up to
    dmax independent update directions otherwise (and at most max_updates updates in all, when given)."""
    enc = plan.encode()
    cols = _columns(enc)
    segs = []
    b = 0                       # table bits before the next column
    log2phys = []               # physical bit of each logical position of the current table
    k = 0
    while k < len(cols):
        i0, i1 = cols[k]
        if b < min_bits:
            # per-op run until the table reaches min_bits at a column boundary
            start = i0
            while k < len(cols) and b < min_bits:
                i0, i1 = cols[k]
                for i in range(i0, i1):
                    c = enc[i, 0]
                    b += 1 if c == OP_EXPAND else (-1 if c == OP_SELECT else 0)
                k += 1
            segs.append(("ops", start, cols[k - 1][1]))
            log2phys = list(range(b))
            continue
        # grow a pass column by column
        basis = _Basis()
        lay_bits = b
        nlow = 0
        start = i0
        bits_now = b
        D = 0
        kk = k
        lay = list(log2phys)
        nvar = b
        nnew = nsel = 0
        while kk < len(cols):
            c0, c1 = cols[kk]
            # the column's update mask over variables, with the layout as it would be after its expands
            lay2 = list(lay)
            nv2 = nvar
            m = 0
            w_ok = True
            for i in range(c0, c1):
                code, a, _ = (int(x) for x in enc[i])
                if code == OP_EXPAND:
                    lay2.append(nv2)
                    nv2 += 1
                elif code == OP_UPDATE:
                    for p in range(len(lay2)):
                        if (a >> p) & 1:
                            m |= 1 << lay2[p]
            mh = m & ~0b11
            ne = sum(1 for i in range(c0, c1) if enc[i, 0] == OP_EXPAND)
            ns = sum(1 for i in range(c0, c1) if enc[i, 0] == OP_SELECT)
            if nnew + ne > MAX_ROWS or nsel + ns > MAX_ROWS:
                break
            if mh:
                r = basis.reduce(mh)
                if r == 0 or D >= dmax:
                    break
            elif m:
                if nlow >= MAX_LOW:
                    break
            if nv2 - 2 - (D + (1 if mh else 0)) > MAX_FREE:
                break
            if max_updates is not None and m and D + nlow + 1 > max_updates:
                break
            if nv2 > MAX_VARS:
                break
            nnew += ne
            nsel += ns
            # accept the column
            if mh:
                basis.add(mh)
                D += 1
            elif m:
                nlow += 1
            for i in range(c0, c1):
                code, a, _ = (int(x) for x in enc[i])
                if code == OP_EXPAND:
                    lay.append(nvar)
                    nvar += 1
                    bits_now += 1
                elif code == OP_SELECT:
                    lay.pop(a)
                    bits_now -= 1
            kk += 1
        if kk == k:             # a column over the limits on its own: no pass can hold it
            raise ValueError("plan column at op %d exceeds the fused pass limits" % i0)
        while True:
            last_op = cols[kk - 1][1]
            b_out = _bits_after(enc, b, start, last_op)
            logical = b_out < min_bits or kk >= len(cols)
            p, out_log = _build_pass(enc, start, last_op, log2phys, logical)
            if kk == k + 1 or max(len(r) for r in p.runs()) <= MAX_RUNS:
                break
            kk -= 1
        log2phys = out_log
        segs.append(("pass", p))
        b = p.b_out
        k = kk
    return Program(segs, plan.peak_bits)


# ------------------------------------------------------------------------------------- CPU model of the kernels
def _apply_ops(T, enc, first, last):
    n = T.size
    for i in range(first, last):
        code, a, b = (int(x) for x in enc[i])
        if code == OP_EXPAND:
            T = np.concatenate([T, np.full(n, 255, np.uint8)])
            n *= 2
        elif code == OP_UPDATE:
            other = T[np.arange(n, dtype=np.int64) ^ a].astype(np.int64) + b
            T = np.minimum(T, np.minimum(other, 255)).astype(np.uint8)
        else:
            T = T[((np.arange(n, dtype=np.int64) >> a) & 1) == b]
            n //= 2
    return T


def _xor_fold(vals, idx):
    """XOR of vals[i] over the set bits i of each entry of idx (int64 array)."""
    out = np.zeros(idx.shape, dtype=np.int64)
    for i, v in enumerate(vals):
        out ^= np.where((idx >> i) & 1, np.int64(v), np.int64(0))
    return out


def _map_index(runs, tid):
    x = np.zeros(tid.shape, dtype=np.int64)
    for f0, ln, sh in runs:
        x |= ((tid >> f0) & ((1 << ln) - 1)) << sh
    return x


def run_pass_cpu(T, p):
    """One pass with the kernel's index arithmetic (run-length index maps, twisted lanes): T (2^b_in uint8) ->
    2^b_out uint8."""
    nthreads = 1 << p.nfree
    tid = np.arange(nthreads, dtype=np.int64)
    rin, rout, rnew, rsel = p.runs()
    base_in, base_out = _map_index(rin, tid), _map_index(rout, tid)
    xnew, xsel = _map_index(rnew, tid), _map_index(rsel, tid)
    Q = 1 << p.D
    sig = [0] * Q
    for q in range(Q):
        for j in range(p.D):
            if (q >> j) & 1:
                sig[q] ^= p.lam[j]
    # reg[t, q, k] = table entry at byte k ^ sig[q] of dword q
    reg = np.full((nthreads, Q, 4), 255, dtype=np.int64)
    for q in range(Q):
        din = base_in.copy()
        nz = xnew.copy()
        for j in range(p.D):
            if (q >> j) & 1:
                din ^= p.din[2 + j]
                nz ^= p.dnew[2 + j]
        ok = nz == 0
        for k in range(4):
            m = k ^ sig[q]
            addr = din ^ ((m & 1) * p.din[0]) ^ ((m >> 1) * p.din[1])
            reg[ok, q, k] = T[addr[ok]]
    for j in range(p.D):                      # lane-aligned butterflies
        w = min(255, int(p.w[j]))
        for q in range(Q):
            if (q >> j) & 1:
                continue
            q2 = q | (1 << j)
            X, Y = reg[:, q, :].copy(), reg[:, q2, :].copy()
            reg[:, q, :] = np.minimum(X, Y + w)
            reg[:, q2, :] = np.minimum(Y, X + w)
    for lam, w in zip(p.low_lam, p.low_w):
        perm = [k ^ lam for k in range(4)]
        reg = np.minimum(reg, reg[:, :, perm] + min(255, int(w)))
    out = np.full(1 << p.b_out, 0, dtype=np.uint8)
    written = np.zeros(1 << p.b_out, dtype=bool)
    for q in range(Q):
        dout = base_out.copy()
        sel = xsel.copy()
        for j in range(p.D):
            if (q >> j) & 1:
                dout ^= p.dout[2 + j]
                sel ^= p.dsel[2 + j]
        for k in range(4):
            m = k ^ sig[q]
            a = dout ^ ((m & 1) * p.dout[0]) ^ ((m >> 1) * p.dout[1])
            s = sel ^ ((m & 1) * p.dsel[0]) ^ ((m >> 1) * p.dsel[1])
            ok = s == p.sel_target
            assert not written[a[ok]].any(), "two entries stored to one output index"
            assert (reg[ok, q, k] <= 255).all()
            out[a[ok]] = reg[ok, q, k]
            written[a[ok]] = True
    assert written.all(), "an output entry was not written"
    return out


def simulate(program, plan):
    """Final value of the plan executed segment by segment on the CPU (per-op ranges by numpy, passes by
    run_pass_cpu)."""
    enc = plan.encode()
    T = np.zeros(1, np.uint8)
    for seg in program.segments:
        if seg[0] == "ops":
            T = _apply_ops(T, enc, seg[1], seg[2])
        else:
            p = seg[1]
            assert T.size == 1 << p.b_in, (T.size, p.b_in)
            T = run_pass_cpu(T, p)
    assert T.size == 1
    return int(T[0])


# ------------------------------------------------------------------------------------------ GPU executor
# runs: f0 | len << 8 | shift << 16; lamw / loww: lane mask | weight << 8 (all 32-bit words for scalar loads)
PASS_DTYPE = np.dtype([("din", "<u8", 8), ("dout", "<u8", 8), ("dnew", "<u4", 8), ("dsel", "<u4", 8),
                       ("runs", "<u4", (4, MAX_RUNS)), ("nruns", "<i4", 4), ("sel_target", "<u4"),
                       ("D", "<i4"), ("nlow", "<i4"), ("byte_store", "<i4"), ("b_in", "<i4"), ("b_out", "<i4"),
                       ("nfree", "<i4"), ("pad", "<i4"), ("lamw", "<u4", 8), ("loww", "<u4", 8),
                       ("qout", "<u4", 32), ("qsel", "<u4", 32), ("qpk", "<u4", 32)], align=True)


def element_tables(p):
    """Per cube element q < 2^D: the XORs over q's directions of dout (low 32 bits) and dsel, and the v_perm
    selector that packs q's twisted lanes -- the tables the kernel's stores read."""
    out = {k: [0] * 32 for k in ("qout", "qsel", "qpk")}
    for q in range(1 << p.D):
        xi = xo = xn = xs = sg = 0
        for j in range(p.D):
            if (q >> j) & 1:
                xi ^= p.din[2 + j]
                xo ^= p.dout[2 + j]
                xn ^= p.dnew[2 + j]
                xs ^= p.dsel[2 + j]
                sg ^= p.lam[j]
        pk = 0
        for k in range(4):
            r = k ^ sg
            pk |= (((r & 1) << 2) | (r & 2)) << (8 * k)
        out["qout"][q], out["qsel"][q], out["qpk"][q] = xo & 0xFFFFFFFF, xs, pk
    return out


def unpack_run(word):
    w = int(word)
    return [w & 0xFF, (w >> 8) & 0xFF, (w >> 16) & 0xFF]
SEG_OPS, SEG_PASS = 0, 1


def pack(program):
    """(segments int64 [nseg, 3], passes PASS_DTYPE [npasses]) in the layout tdp_run_program reads."""
    passes = program.passes
    arr = np.zeros(max(1, len(passes)), dtype=PASS_DTYPE)
    segs = np.zeros((len(program.segments), 3), dtype=np.int64)
    ip = 0
    for s, seg in enumerate(program.segments):
        if seg[0] == "ops":
            segs[s] = (SEG_OPS, seg[1], seg[2])
            continue
        p = seg[1]
        r = arr[ip]
        K = p.D + 2
        r["din"][:K] = p.din
        r["dout"][:K] = p.dout
        r["dnew"][:K] = p.dnew
        r["dsel"][:K] = p.dsel
        for m, runs in enumerate(p.runs()):
            if len(runs) > MAX_RUNS:
                raise ValueError("pass index map with %d runs (at most %d)" % (len(runs), MAX_RUNS))
            r["nruns"][m] = len(runs)
            for i, (f0, ln, sh) in enumerate(runs):
                r["runs"][m, i] = f0 | ln << 8 | sh << 16
        r["sel_target"] = p.sel_target
        r["nfree"], r["nlow"], r["D"] = p.nfree, len(p.low_lam), p.D
        r["byte_store"], r["b_in"], r["b_out"] = int(p.byte_store), p.b_in, p.b_out
        r["lamw"][:p.D] = [int(lam) | min(255, int(w)) << 8 for lam, w in zip(p.lam, p.w)]
        r["loww"][:len(p.low_lam)] = [int(lam) | min(255, int(w)) << 8 for lam, w in zip(p.low_lam, p.low_w)]
        for k, v in element_tables(p).items():
            r[k] = v
        segs[s] = (SEG_PASS, ip, 0)
        ip += 1
    return segs, arr


class NativeProgram:
    """A compiled program as the executor reads it: segments and pass descriptors (and witness steps)."""

    def __init__(self, segs, arr, steps=None):
        self.segs, self.arr, self.steps = segs, arr, steps
        self.npasses = int((segs[:, 0] == SEG_PASS).sum()) if len(segs) else 0

    def pass_bytes(self, i):
        r = self.arr[i]
        return (1 << int(r["b_in"])) + (1 << int(r["b_out"]))

    def pass_kernel(self, i):
        r = self.arr[i]
        return "tdp::k_pass<%d, %s, %s, %s>" % (int(r["D"]), "true" if r["byte_store"] else "false",
                                                "true" if r["nruns"][2] > 0 else "false",
                                                "unsigned int" if max(r["b_in"], r["b_out"]) <= 32
                                                else "unsigned long")


def _cached(plan, key, make):
    """Compiled programs live on the plan object (per compile options), so they last as long as the plan."""
    cache = plan.__dict__.setdefault("_fused_programs", {})
    val = cache.get(key)
    if val is None:
        val = cache[key] = make()
    return val


def program_for(plan, dmax=DMAX, min_bits=12):
    """The native program of a plan, cached per plan object.  ValueError if a column exceeds the pass limits."""
    return _cached(plan, ("run", dmax, min_bits),
                   lambda: NativeProgram(*compile_native(plan, dmax=dmax, min_bits=min_bits)))


def witness_program_for(plan):
    """The native witness-mode program of a plan, cached per plan object."""
    return _cached(plan, ("witness",), lambda: NativeProgram(*compile_native(
        plan, min_bits=2, max_updates=WITNESS_MAX_UPDATES, witness=True)))


def run_program(plan, program=None, T0=None, T1=None, stream=None, raw=False):
    """Min cost of a feasible plan by the fused-pass executor (tdp_run_program); OverflowError at 255 (or the raw
    final entry when raw)."""
    import ctypes
    from . import _lib, kernel
    L = _lib.lib()
    if int(L.tdp_pass_args_bytes()) != PASS_DTYPE.itemsize:
        raise RuntimeError("pass descriptor layout mismatch: library %d bytes, host %d"
                           % (L.tdp_pass_args_bytes(), PASS_DTYPE.itemsize))
    from .trellis import as_plan
    plan = as_plan(plan)
    if program is None:
        prog = program_for(plan)
        segs, arr = prog.segs, prog.arr
    else:
        segs, arr = pack(program)
    ops = plan.encode()
    size = 1 << plan.peak_bits
    T0 = _lib.device_empty(size) if T0 is None else T0
    T1 = _lib.device_empty(size) if T1 is None else T1
    value = ctypes.c_int(-1)
    rc = L.tdp_run_program(ops.ctypes.data_as(ctypes.c_void_p), len(ops), segs.ctypes.data_as(ctypes.c_void_p),
                           len(segs), arr.ctypes.data_as(ctypes.c_void_p), len(arr), plan.peak_bits,
                           _lib.dptr(T0), _lib.dptr(T1), ctypes.byref(value), _lib.stream_handle(stream))
    _lib.check(rc, "tdp_run_program")
    return value.value if raw else kernel.final_value(value.value)


# ------------------------------------------------------------------------------------------ witness (traceback)
# Witness mode runs the program with every step writing a fresh slot of one arena, so the input table of every
# step stays in HBM.  A step is a pass or one per-op op (an expand directly followed by the update of its new top
# bit is one step, as in the per-op executor).  The traceback walks the steps backward in the step's variable
# space (input bits, then the rows it opens): at a step with updates u_0..u_{k-1} (k <= 6), the state x after it
# and its value V there, lane l of one wave loads X[x ^ XOR_{j in l} m_j] + sum_{j in l} w_j (X = 255 where a row
# opened in the step is 1), and update i, last first, was used iff the minimum over the lanes that agree with the
# used set above i differs from V.  Every entry any decision needs is one of the 2^k loaded values.
WITNESS_MAX_UPDATES = 6
STEP_PASS, STEP_OPS = 0, 1
WSTEP_DTYPE = np.dtype([("in_off", "<u8"), ("inmask", "<u8"), ("newmask", "<u8"), ("fixmask", "<u8"),
                        ("fixval", "<u8"), ("m", "<u8", 6), ("op", "<i8", 6), ("out_off", "<u8"), ("n_in", "<u8"),
                        ("op0", "<i8"), ("kind", "<i4"), ("pass_index", "<i4"), ("nop", "<i4"), ("k", "<i4"),
                        ("b_out", "<i4"), ("pad", "<i4"), ("w", "u1", 6), ("out_var", "u1", 48), ("pad2", "u1", 2)],
                       align=True)
SLOT_ALIGN = 256


def _align(x):
    return (x + SLOT_ALIGN - 1) // SLOT_ALIGN * SLOT_ALIGN


def witness_program(plan):
    """(program, steps WSTEP_DTYPE array, table bytes of the arena) of a feasible plan in witness mode: passes of at
    most six updates from 2^2 entries up, single per-op steps below."""
    prog = compile_program(plan, min_bits=2, max_updates=WITNESS_MAX_UPDATES)
    enc = plan.encode()
    rows = []
    off = 0
    b = 0
    pass_index = 0
    for seg in prog.segments:
        if seg[0] == "pass":
            p = seg[1]
            assert len(p.upd_ops) <= WITNESS_MAX_UPDATES and p.b_out <= 48
            rows.append(dict(kind=STEP_PASS, pass_index=pass_index, op0=p.first, nop=p.last - p.first, b_in=p.b_in,
                             b_out=p.b_out, newmask=p.new_mask, fixmask=p.fix_mask, fixval=p.fix_val,
                             upd=p.upd_ops, out_var=p.out_vars))
            pass_index += 1
            b = p.b_out
            continue
        i = seg[1]
        while i < seg[2]:
            code, a, w = (int(x) for x in enc[i])
            if code == OP_EXPAND:
                fuse = i + 1 < seg[2] and int(enc[i + 1, 0]) == OP_UPDATE and (int(enc[i + 1, 1]) >> b) == 1
                upd = [(int(enc[i + 1, 1]), int(enc[i + 1, 2]), i + 1)] if fuse else []
                rows.append(dict(kind=STEP_OPS, pass_index=-1, op0=i, nop=2 if fuse else 1, b_in=b, b_out=b + 1,
                                 newmask=1 << b, fixmask=0, fixval=0, upd=upd, out_var=list(range(b + 1))))
                i += 2 if fuse else 1
                b += 1
            elif code == OP_UPDATE:
                rows.append(dict(kind=STEP_OPS, pass_index=-1, op0=i, nop=1, b_in=b, b_out=b, newmask=0, fixmask=0,
                                 fixval=0, upd=[(a, w, i)] if a else [], out_var=list(range(b))))
                i += 1
            else:
                rows.append(dict(kind=STEP_OPS, pass_index=-1, op0=i, nop=1, b_in=b, b_out=b - 1, newmask=0,
                                 fixmask=1 << a, fixval=w << a,
                                 upd=[], out_var=[q if q < a else q + 1 for q in range(b - 1)]))
                i += 1
                b -= 1
    steps = np.zeros(len(rows), dtype=WSTEP_DTYPE)
    for r, row in zip(steps, rows):
        r["in_off"] = off
        r["n_in"] = 1 << row["b_in"]
        off += _align(1 << row["b_in"])
        r["out_off"] = off
        r["inmask"] = (1 << row["b_in"]) - 1
        r["newmask"], r["fixmask"], r["fixval"] = row["newmask"], row["fixmask"], row["fixval"]
        r["kind"], r["pass_index"], r["op0"], r["nop"] = row["kind"], row["pass_index"], row["op0"], row["nop"]
        r["k"] = len(row["upd"])
        for j, (m, w, i) in enumerate(row["upd"]):
            r["m"][j], r["w"][j], r["op"][j] = m, min(255, w), i
        r["b_out"] = row["b_out"]
        r["out_var"][:len(row["out_var"])] = row["out_var"]
    table_bytes = off + SLOT_ALIGN          # the final one-entry table
    return prog, steps, table_bytes


def witness_cpu(plan):
    """(value, used op indices) by the witness-mode program on the CPU: the forward steps by numpy, the traceback
    with the device kernel's arithmetic.  Used by the tests."""
    prog, steps, _ = witness_program(plan)
    enc = plan.encode()
    passes = prog.passes
    tables = []
    T = np.zeros(1, np.uint8)
    for st in steps:
        assert T.size == int(st["n_in"])
        tables.append(T)
        if st["kind"] == STEP_PASS:
            T = run_pass_cpu(T, passes[int(st["pass_index"])])
        else:
            T = _apply_ops(T, enc, int(st["op0"]), int(st["op0"]) + int(st["nop"]))
    assert T.size == 1
    v = int(T[0])
    if v == 255:
        return v, []
    used = []
    x = int(steps[-1]["fixval"])
    for s in range(len(steps) - 1, -1, -1):
        st = steps[s]
        k = int(st["k"])
        X = tables[s]
        vals = []
        for lane in range(1 << k):
            y, ws = x, 0
            for j in range(k):
                if (lane >> j) & 1:
                    y ^= int(st["m"][j])
                    ws += int(st["w"][j])
            vals.append((255 if y & int(st["newmask"]) else int(X[y & int(st["inmask"])])) + ws)
        # with T the used set above i, the entry before update i at the current state is
        # min(255, min over lanes T | S (S below i) of vals - w(T)), and the current value is V - w(T), V the value
        # after the step: update i was used iff that minimum differs from V (V < 255 on a path)
        V = v
        used_set = 0
        for i in range(k - 1, -1, -1):
            before = min(vals[lane] for lane in range(1 << k) if (lane >> i) == (used_set >> i))
            if before != V:
                used_set |= 1 << i
                v -= int(st["w"][i])
                used.append(int(st["op"][i]))
        for j in range(k):
            if (used_set >> j) & 1:
                x ^= int(st["m"][j])
        assert x & int(st["newmask"]) == 0, "traceback left a row opened in the step set"
        if s > 0:
            prev = steps[s - 1]
            xp = int(prev["fixval"])
            for q in range(int(prev["b_out"])):
                if (x >> q) & 1:
                    xp |= 1 << int(prev["out_var"][q])
            x = xp
    assert x == 0 and v == 0, (x, v)
    return int(T[0]), sorted(used)


def witness(plan, max_bytes=None, stream=None):
    """(value, sorted op indices of the updates on one minimum path) of a feasible plan by the witness-mode program
    on the GPU; None if its arena would exceed max_bytes.  OverflowError at 255."""
    import ctypes
    from . import _lib, kernel
    from .trellis import as_plan
    plan = as_plan(plan)
    L = _lib.lib()
    if int(L.tdp_witness_step_bytes()) != WSTEP_DTYPE.itemsize:
        raise RuntimeError("witness step layout mismatch")
    prog = witness_program_for(plan)
    steps, arr = prog.steps, prog.arr
    nbytes = witness_bytes(plan)
    if max_bytes is not None and nbytes > max_bytes:
        return None
    ops = plan.encode()
    arena = _lib.device_empty(nbytes)
    used = np.zeros(max(1, int(steps["k"].sum())), dtype=np.int64)
    value, nused = ctypes.c_int(-1), ctypes.c_int64(0)
    rc = L.tdp_run_program_witness(ops.ctypes.data_as(ctypes.c_void_p), len(ops),
                                   steps.ctypes.data_as(ctypes.c_void_p), len(steps),
                                   arr.ctypes.data_as(ctypes.c_void_p), len(arr), _lib.dptr(arena), nbytes,
                                   ctypes.byref(value), used.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nused),
                                   _lib.stream_handle(stream))
    _lib.check(rc, "tdp_run_program_witness")
    v = kernel.final_value(value.value)
    return v, sorted(int(i) for i in used[:nused.value])


def witness_bytes(plan):
    """Device bytes of the witness-mode arena of a plan (tables of every step plus the traceback's records).
    ValueError if a column exceeds the pass limits."""
    import ctypes
    from . import _lib
    steps = witness_program_for(plan).steps
    nbytes = ctypes.c_uint64(0)
    _lib.check(_lib.lib().tdp_program_witness_bytes(steps.ctypes.data_as(ctypes.c_void_p), len(steps),
                                                    ctypes.byref(nbytes)), "tdp_program_witness_bytes")
    return nbytes.value


# ------------------------------------------------------------------------------------------ native compiler
def compile_native(plan, dmax=DMAX, min_bits=12, max_updates=None, witness=False):
    """(segments, pass descriptors[, witness steps]) by libtdp's C++ compiler (tdp_fused_compile): the same arrays as
    pack(compile_program(...)) and witness_program(...), about a hundred times faster.  ValueError when a column
    exceeds the pass limits."""
    import ctypes
    from . import _lib
    L = _lib.lib()
    enc = np.ascontiguousarray(plan.encode(), dtype=np.uint64)
    nops = len(enc)
    # capacities by the op count (a pass descriptor is 880 bytes): left uninitialised, since the compiler writes every
    # field of every segment, descriptor and step it reports
    segs = np.empty((nops + 2, 3), dtype=np.int64)
    arr = np.empty(max(1, nops + 1), dtype=PASS_DTYPE)
    steps = np.empty(nops + 2, dtype=WSTEP_DTYPE) if witness else None
    ns, npass, nst = ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    rc = L.tdp_fused_compile(enc.ctypes.data_as(ctypes.c_void_p), nops, dmax, min_bits, max_updates or 0,
                             segs.ctypes.data_as(ctypes.c_void_p), len(segs), ctypes.byref(ns),
                             arr.ctypes.data_as(ctypes.c_void_p), len(arr), ctypes.byref(npass),
                             steps.ctypes.data_as(ctypes.c_void_p) if witness else None,
                             len(steps) if witness else 0, ctypes.byref(nst))
    if rc == 1:
        raise ValueError(L.tdp_last_error().decode(errors="replace"))
    _lib.check(rc, "tdp_fused_compile")
    segs = segs[:ns.value]
    arr = arr[:max(1, npass.value)]
    if witness:
        return segs, arr, steps[:nst.value]
    return segs, arr


# ------------------------------------------------------------------------------------------ many programs at once
def _run_record_type():
    import ctypes
    vp, i64, u64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32

    class TdpRun(ctypes.Structure):
        """tdp_run_t of include/tdp.h."""
        _fields_ = [("ops", vp), ("nops", i64), ("segments", vp), ("nseg", i64), ("passes", vp), ("npasses", i64),
                    ("steps", vp), ("nsteps", i64), ("d_T0", vp), ("d_T1", vp), ("d_arena", vp),
                    ("arena_bytes", u64), ("used_ops", vp), ("nused", i64), ("peak_bits", i32), ("value", i32)]
    return TdpRun


_STAGING = __import__("threading").local()


def _staging(nbytes):
    """The calling thread's pinned host buffer of at least nbytes for schedule uploads (grown on demand, reused by
    the thread's later calls)."""
    from . import _lib
    t = _lib.torch()
    buf = getattr(_STAGING, "buf", None)
    if buf is None or buf.numel() < nbytes:
        buf = _STAGING.buf = t.empty(max(nbytes, 1 << 24), dtype=t.uint8, pin_memory=True)
    return buf


def run_many(plans, witness=None, stream=None, min_bits=2, timing=None):
    """Run many feasible plans together by tdp_run_programs (one launch per round runs the next pass of every
    program).  witness[i] true: a witness run, result (value, sorted used op indices); else the value.  Values of
    255 or more raise OverflowError.  ValueError if a plan's column exceeds the pass limits (run it alone).
    min_bits: value runs use passes from 2^min_bits entries up (program_for); smaller tables run op by op.
    timing (a dict): filled with "round_us" and "round_bytes" (numpy arrays, one entry per multi-program launch:
    its duration by HIP events on the stream and its algorithmic bytes), by tdp_run_programs_timed."""
    import ctypes
    from . import _lib, kernel
    from .trellis import as_plan
    L = _lib.lib()
    TdpRun = _run_record_type()
    if int(L.tdp_run_record_bytes()) != ctypes.sizeof(TdpRun):
        raise RuntimeError("tdp_run_t layout mismatch")
    plans = [as_plan(p) for p in plans]
    witness = [False] * len(plans) if witness is None else list(witness)
    runs = (TdpRun * len(plans))()
    keep, used_of = [], {}
    for r, (plan, w) in enumerate(zip(plans, witness)):
        if not plan.feasible:
            raise ValueError("run_many takes feasible plans only")
        ops = np.ascontiguousarray(plan.encode(), dtype=np.uint64)
        R = runs[r]
        R.ops, R.nops = ops.ctypes.data, len(ops)
        R.peak_bits = plan.peak_bits
        if w:
            prog = witness_program_for(plan)
            nbytes = witness_bytes(plan)
            arena = _lib.device_empty(nbytes)
            used = np.zeros(max(1, int(prog.steps["k"].sum())), dtype=np.int64)
            R.steps, R.nsteps = prog.steps.ctypes.data, len(prog.steps)
            R.d_arena, R.arena_bytes = _lib.dptr(arena), nbytes
            R.used_ops = used.ctypes.data
            used_of[r] = used
            keep.append(arena)
        else:
            # by default passes from 2^2 entries up: narrow heads and tails are rounds too, not one launch per op
            prog = program_for(plan, min_bits=min_bits)
            T0 = _lib.device_empty(1 << plan.peak_bits)
            T1 = _lib.device_empty(1 << plan.peak_bits)
            R.segments, R.nseg = prog.segs.ctypes.data, len(prog.segs)
            R.d_T0, R.d_T1 = _lib.dptr(T0), _lib.dptr(T1)
            keep += [T0, T1]
        R.passes, R.npasses = prog.arr.ctypes.data, prog.npasses
        keep += [ops, prog]
    nb = ctypes.c_uint64(0)
    _lib.check(L.tdp_run_programs_workspace(runs, len(plans), ctypes.byref(nb)), "tdp_run_programs_workspace")
    work = _lib.device_empty(nb.value)
    stage = _staging(nb.value)
    if timing is None:
        _lib.check(L.tdp_run_programs(runs, len(plans), _lib.dptr(work), nb.value, stage.data_ptr(), stage.numel(),
                                      _lib.stream_handle(stream)), "tdp_run_programs")
    else:
        cap = sum(int(getattr(r, "npasses")) for r in runs) + len(plans) + 16
        us, by, nr = np.zeros(cap), np.zeros(cap), ctypes.c_int64(0)
        _lib.check(L.tdp_run_programs_timed(runs, len(plans), _lib.dptr(work), nb.value, stage.data_ptr(),
                                            stage.numel(), _lib.stream_handle(stream), us.ctypes.data,
                                            by.ctypes.data, cap, ctypes.byref(nr)), "tdp_run_programs_timed")
        timing["round_us"], timing["round_bytes"] = us[:nr.value].copy(), by[:nr.value].copy()
    out = []
    for r, w in enumerate(witness):
        v = kernel.final_value(runs[r].value)
        out.append((v, sorted(int(i) for i in used_of[r][:runs[r].nused])) if w else v)
    return out


class ProgramGraph:
    """Value runs of many plans captured once as one HIP graph (tdp_programs_graph) and replayed by launch(): the
    same rounds as run_many, with the host's schedule, upload and per-launch calls paid once at capture.  values()
    waits for the stream and returns the distances in plan order (OverflowError at 255).  The graph owns the plans'
    device tables and workspace while it lives; close() (or garbage collection) frees it."""

    def __init__(self, plans, stream=None, min_bits=2):
        import ctypes
        from . import _lib
        from .trellis import as_plan
        L = self._L = _lib.lib()
        TdpRun = _run_record_type()
        self._plans = [as_plan(p) for p in plans]
        runs = (TdpRun * len(self._plans))()
        self._keep = []
        for r, plan in enumerate(self._plans):
            if not plan.feasible:
                raise ValueError("ProgramGraph takes feasible plans only")
            ops = np.ascontiguousarray(plan.encode(), dtype=np.uint64)
            prog = program_for(plan, min_bits=min_bits)
            T0 = _lib.device_empty(1 << plan.peak_bits)
            T1 = _lib.device_empty(1 << plan.peak_bits)
            R = runs[r]
            R.ops, R.nops, R.peak_bits = ops.ctypes.data, len(ops), plan.peak_bits
            R.segments, R.nseg = prog.segs.ctypes.data, len(prog.segs)
            R.passes, R.npasses = prog.arr.ctypes.data, prog.npasses
            R.d_T0, R.d_T1 = _lib.dptr(T0), _lib.dptr(T1)
            self._keep += [ops, prog, T0, T1]
        nb = ctypes.c_uint64(0)
        _lib.check(L.tdp_run_programs_workspace(runs, len(runs), ctypes.byref(nb)), "tdp_run_programs_workspace")
        self._work = _lib.device_empty(nb.value)
        self._h = L.tdp_programs_graph(runs, len(runs), _lib.dptr(self._work), nb.value, _lib.stream_handle(stream))
        if not self._h:
            raise _lib.TdpError("tdp_programs_graph failed: " + L.tdp_last_error().decode(errors="replace"),
                                _lib.TDP_EHIP)
        n = ctypes.c_uint64(0)
        _lib.check(L.tdp_graph_nodes(self._h, ctypes.byref(n)), "tdp_graph_nodes")
        self.nodes = int(n.value)

    def launch(self, stream=None):
        from . import _lib
        _lib.check(self._L.tdp_graph_launch(self._h, _lib.stream_handle(stream)), "tdp_graph_launch")

    def values(self, stream=None):
        from . import _lib, kernel
        out = np.zeros(len(self._plans), dtype=np.int32)
        _lib.check(self._L.tdp_graph_values(self._h, out.ctypes.data, len(out), _lib.stream_handle(stream)),
                   "tdp_graph_values")
        return [kernel.final_value(v) for v in out]

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._L.tdp_graph_free(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
