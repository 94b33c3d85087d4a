"""GPU parity of the sparse sweeps' dense fallback (tdp_dn_*; cdist/sparse_dp.py:221-289, 680-707): the dense
column step against the oracle's restatement on random tables and layouts, and dense sweeps on the device against
the oracle's sparse sweeps (the fallback changes the representation only, so every cut holds the same states),
brute force and the recorded distances of the fixtures."""
import random

import numpy as np
import pytest

import oracle
from helpers import dense_fixtures
from sparse_helpers import brute, check_witness, random_problem
from tropidist_amd import _lib, problem, sparse_dp
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu


class _Tr:
    nslots = 56

    def bound_data(self, forward):
        return []


def _random_layout(rng, w_max=12):
    """Old slots, starting rows, ending rows (targets) of one column, with room in 56 slots.  One layout in four
    has up to ten starting and ending rows, so that its index maps have more runs of consecutive bits than the
    fast kernel takes (the generic kernel runs it)."""
    many = rng.random() < 0.25
    w = rng.randint(0, 8 if many else w_max)
    pool = rng.sample(range(56), 56)
    old = sorted(pool[:w])
    starts = [(s, 0) for s in sorted(pool[w:w + rng.randint(0, 10 if many else 3)])]
    mid = sorted(old + [s for s, _ in starts])
    nend = rng.randint(0, min(10 if many else 3, len(mid)))
    ends = [(s, rng.randint(0, 1)) for s in sorted(rng.sample(mid, nend))]
    return old, starts, ends, mid


def test_dense_step_equals_oracle_on_random_tables():
    print("case dense step random seed 21", flush=True)
    rng = random.Random(21)
    nrng = np.random.default_rng(21)
    torch = _lib.torch()
    hip, cpu = sparse_dp.HipTables(_Tr()), oracle.OracleTables(_Tr())
    for trial in range(400):
        old, starts, ends, mid = _random_layout(rng, 22 if trial % 50 == 0 else 12)
        T = nrng.integers(0, 12, size=1 << len(old)).astype(np.uint8)
        T[nrng.random(T.size) < nrng.random()] = 255
        upd = sum(1 << s for s in mid if rng.random() < 0.4)
        cap, K = rng.randint(0, 11), rng.randint(0, 14)
        new = [s for s in mid if s not in {e for e, _ in ends}]
        tmask = sum(1 << s for s in new if rng.random() < 0.2)
        hd = None
        if rng.random() < 0.6:
            ncls = rng.randint(0, 4)
            hd = ([sum(1 << s for s in mid if rng.random() < 0.5) for _ in range(ncls)],
                  [rng.randint(1, 720720) for _ in range(ncls)], rng.randint(0, 720720))
        want = cpu.dense_step(oracle.OracleDense(old, T.copy()), starts, ends, upd, cap, K, tmask, hd)
        dev = torch.from_numpy(np.concatenate([T, np.zeros(4, dtype=np.uint8)])).cuda()
        got = hip.dense_step(sparse_dp.DenseTab(dev, old, int(np.count_nonzero(T != 255))), starts, ends, upd, cap,
                             K, tmask, hd)
        assert got.slots == want.slots, trial
        assert np.array_equal(got.T[:1 << got.w].cpu().numpy(), want.T), (trial, old, starts, ends, upd, cap, K)
        assert got.n == want.count(), trial


def test_dense_conversions_and_join_equal_oracle():
    print("case dense conversions seed 5", flush=True)
    rng = random.Random(5)
    nrng = np.random.default_rng(5)
    hip, cpu = sparse_dp.HipTables(_Tr()), oracle.OracleTables(_Tr())
    for trial in range(200):
        slots = sorted(rng.sample(range(56), rng.randint(0, 12)))
        full = 1 << len(slots)
        idx = np.flatnonzero(nrng.random(full) < nrng.random())
        D = oracle.OracleDense(slots, np.full(full, 255, dtype=np.uint8))
        D.T[idx] = nrng.integers(0, 9, size=idx.size).astype(np.uint8)
        keys, vals = D.arrays()
        sp = hip.from_arrays(keys, vals)
        dt = hip.to_dense(sp, slots)
        assert np.array_equal(dt.T[:full].cpu().numpy(), D.T), trial
        back = hip.from_dense(dt)
        got = sorted(zip(*(a.tolist() for a in hip.arrays(back))))
        assert got == sorted(zip(keys.tolist(), vals.tolist())), trial
        # joins: forward states against the dense table, some keys outside its slots
        fk = [int(k) ^ (sum(1 << s for s in slots if rng.random() < 0.3)) for k in keys[:40]]
        fk += [rng.getrandbits(56) for _ in range(5)]
        fk = sorted(set(fk))
        fv = [rng.randint(0, 8) for _ in fk]
        tmask = sum(1 << s for s in slots if rng.random() < 0.2)
        ch_cpu = cpu.from_arrays(np.array(fk, dtype=np.uint64), np.array(fv, dtype=np.uint8))
        ch_hip = hip.from_arrays(fk, fv)
        assert hip.join(ch_hip, dt, tmask) == cpu.join(ch_cpu, D, tmask), trial
        for s in fk[:8]:
            assert hip.lookup(dt, s) == D.lookup(s), trial


def _cut_tables(tr, direction, K, use_h, dense, be, dense_max=1 << 30):
    sw = sparse_dp.Sweep(tr, direction, K, K, use_h, backend=be, dense=dense, dense_max=dense_max)
    out = {}

    def keep(c, tab, changed, extra):
        keys, vals = sw.be.arrays(tab)
        out[c] = sorted(zip((int(k) for k in keys), (int(v) for v in vals)))

    sw.run(on_cut=keep)
    return out, sw


def test_dense_sweeps_hold_the_oracle_sparse_tables():
    print("case dense sweeps seed 11", flush=True)
    rng = random.Random(11)
    dense_cuts = 0
    for _ in range(25):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            for direction in (+1, -1):
                K, use_h = rng.randint(1, p.ncols), rng.random() < 0.5
                want, _ = _cut_tables(tr, direction, K, use_h, None, oracle.OracleTables(tr))
                for dense, dmax in ((1e-9, 1 << 30), (0.25, 1 << 30), (1e-9, 8)):
                    got, sw = _cut_tables(tr, direction, K, use_h, dense, "hip", dmax)
                    assert got == want, (direction, K, use_h, dense, dmax)
                    dense_cuts += sw.dense_cuts
    assert dense_cuts > 100


def test_dense_queries_against_brute_force():
    print("case dense queries seed 3", flush=True)
    rng = random.Random(3)
    checked = 0
    for _ in range(20):
        p = random_problem(rng)
        order = list(range(p.ncols))
        rng.shuffle(order)
        for obs in range(p.nobs):
            tr = sparse_dp.Trellis.for_query(p, obs, order)
            if not tr.feasible:
                continue
            truth = brute(p, obs)
            for K in range(1, p.ncols + 2, 2):
                want = truth if truth <= K else None
                for mode, split, dense in (("forward", None, 1e-9), ("mitm", K // 2 + 1, 0.25)):
                    r = sparse_dp.query(tr, K, mode=mode, split=split, use_h=rng.random() < 0.5, dense=dense,
                                        backend="hip", every=rng.choice([1, 2, 5]))
                    assert r["value"] == want, (mode, split, dense, K, r["value"], want)
                    if want is not None:
                        check_witness(p, r["witness"], want, obs)
                    checked += 1
    assert checked > 100


@pytest.mark.parametrize("name", ["code_bb72", "star_42", "surface_d7", "routing_54"])
def test_dense_fallback_on_recorded_codes(name):
    """tests/test_sparse_dp.py:146-155 on the fixtures: with dense=0.25 the forward sweep and the meet in the
    middle give the recorded distance at K = d (none at d - 1) as the native runner does without the fallback, the
    forward sweep has the oracle's per-cut counts, and the witnesses are valid."""
    r = next(x for x in dense_fixtures() if x["name"] == name)
    p = Problem(r["name"], r["nrows"], r["nobs"], r["h"], r["l"])
    d = min(x for x in r["distances"] if x is not None)
    obs = r["distances"].index(d)
    tr = sparse_dp.Trellis.for_query(p, obs, r["order"])
    print("case dense fixture", name, "d", d, "obs", obs, "slots", tr.nslots, flush=True)
    for K in (d - 1, d):
        plain = sparse_dp.query(tr, K, mode="forward", backend="hip")
        dense = sparse_dp.query(tr, K, mode="forward", backend="hip", dense=0.25)
        assert dense["dense_cuts"] > 0 and plain["dense_cuts"] == 0
        ref = sparse_dp.query(tr, K, mode="forward", backend=oracle.OracleTables(tr), witness=False)
        assert dense["forward_counts"] == ref["forward_counts"]
        assert plain["value"] == dense["value"] == (d if K == d else None)
        mitm = sparse_dp.query(tr, K, mode="mitm", backend="hip", dense=0.25)
        assert mitm["value"] == plain["value"]
        for rec in (dense, mitm):
            if rec["value"] is not None:
                check_witness(p, rec["witness"], d, obs)


def test_dense_fallback_at_the_8bit_boundary_on_gpu():
    print("case dense chain boundary", flush=True)
    for n, mode in ((254, "forward"), (256, "mitm")):
        tr = sparse_dp.Trellis.for_query(problem.chain(n), 0, list(range(n)))
        r = sparse_dp.query(tr, n, mode=mode, backend="hip", use_h=False, dense=0.25)
        assert r["value"] == n and r["witness"] == list(range(n)) and r["dense_cuts"] > 0
