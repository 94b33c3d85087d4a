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
        else:
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
                byte_store=not (b_out >= 2 and phys[0] == 0 and phys[1] == 1)), out_log


def _bits_after(enc, b, first, last):
    for i in range(first, last):
        c = enc[i, 0]
        b += 1 if c == OP_EXPAND else (-1 if c == OP_SELECT else 0)
    return b


def compile_program(plan, dmax=DMAX, min_bits=12):
    """Segments of a feasible plan: per-op ranges while the table is narrower than 2^min_bits, passes of up to
    dmax independent update directions otherwise."""
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
        if kk == k:             # a column that cannot start a pass on its own (dependent mask): one-op pass
            kk = k + 1
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
PASS_DTYPE = np.dtype([("din", "<u8", 8), ("dout", "<u8", 8), ("dnew", "<u4", 8), ("dsel", "<u4", 8),
                       ("run_f0", "u1", (4, MAX_RUNS)), ("run_len", "u1", (4, MAX_RUNS)),
                       ("run_sh", "u1", (4, MAX_RUNS)), ("nruns", "<i4", 4), ("sel_target", "<u4"),
                       ("D", "<i4"), ("nlow", "<i4"), ("byte_store", "<i4"), ("b_in", "<i4"), ("b_out", "<i4"),
                       ("nfree", "<i4"), ("pad", "<i4"),
                       ("lam", "u1", 8), ("w", "u1", 8), ("low_lam", "u1", 8), ("low_w", "u1", 8)], align=True)
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
                r["run_f0"][m, i], r["run_len"][m, i], r["run_sh"][m, i] = f0, ln, sh
        r["sel_target"] = p.sel_target
        r["nfree"], r["nlow"], r["D"] = p.nfree, len(p.low_lam), p.D
        r["byte_store"], r["b_in"], r["b_out"] = int(p.byte_store), p.b_in, p.b_out
        r["lam"][:p.D] = p.lam
        r["w"][:p.D] = [min(255, int(x)) for x in p.w]
        r["low_lam"][:len(p.low_lam)] = p.low_lam
        r["low_w"][:len(p.low_w)] = [min(255, int(x)) for x in p.low_w]
        segs[s] = (SEG_PASS, ip, 0)
        ip += 1
    return segs, arr


_compiled = {}


def program_for(plan, dmax=DMAX, min_bits=12):
    """compile_program with a per-plan cache (keyed by the plan object)."""
    key = (id(plan), dmax, min_bits)
    hit = _compiled.get(key)
    if hit is not None and hit[0] is plan:
        return hit[1]
    prog = compile_program(plan, dmax=dmax, min_bits=min_bits)
    packed = pack(prog)
    if len(_compiled) > 256:
        _compiled.clear()
    _compiled[key] = (plan, (prog, packed))
    return prog, packed


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
        _, (segs, arr) = program_for(plan)
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
