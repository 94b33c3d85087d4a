"""The matrices front end (python -m tropidist_amd --matrices, tropidist_amd.pipeline) on the CPU: MatrixMarket
files as bench/common.py:64-90 writes and reads them, the config syntax, the command line's refusals, and the
decision flow with the dense DP answered by the oracle (its device executors need a GPU), against brute force."""
import os
import random

import pytest

import oracle
from helpers import random_problem
from tropidist_amd import __main__ as cli
from tropidist_amd import auto, pipeline
from tropidist_amd.problem import Problem


def test_mtx_round_trip_and_gf2_cancellation(tmp_path):
    rng = random.Random(3)
    for trial in range(20):
        nrows, ncols = rng.randint(1, 70), rng.randint(0, 40)
        cols = [rng.getrandbits(nrows) for _ in range(ncols)]
        path = str(tmp_path / "m.mtx")
        pipeline.write_mtx(path, cols, nrows)
        assert pipeline.read_mtx(path) == (nrows, ncols, cols)
    with open(path, "w") as f:              # a repeated entry cancels over GF(2); comment lines are skipped
        f.write("%%MatrixMarket matrix coordinate pattern general\n% comment\n3 2 3\n1 1\n3 2\n1 1\n")
    assert pipeline.read_mtx(path) == (3, 2, [0, 4])
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n3 1\n")
    with pytest.raises(ValueError):
        pipeline.read_mtx(path)


def test_load_matrices_drops_zero_and_repeated_columns(tmp_path):
    h, l = [1, 0, 1, 2, 0, 3], [0, 0, 0, 1, 1, 1]
    pipeline.write_mtx(str(tmp_path / "H.mtx"), h, 2)
    pipeline.write_mtx(str(tmp_path / "L.mtx"), l, 1)
    p = pipeline.load_matrices(str(tmp_path / "H.mtx"), str(tmp_path / "L.mtx"))
    assert p.columns == [0, 3, 4, 5] and p.h == [1, 2, 0, 3] and p.l == [0, 1, 1, 1] and p.file_columns == 6
    pipeline.write_mtx(str(tmp_path / "L2.mtx"), l[:5], 1)
    with pytest.raises(ValueError):
        pipeline.load_matrices(str(tmp_path / "H.mtx"), str(tmp_path / "L2.mtx"))


def test_configuration():
    assert pipeline.configuration("") == pipeline.DEFAULTS
    assert pipeline.configuration("dp=1, kmax=12")["kmax"] == 12
    for bad in ("dp=3", "kmax=0", "kmin=5,kmax=4", "nope=1", "dp", "table_gb=0"):
        with pytest.raises(ValueError):
            pipeline.configuration(bad)


def test_command_line_refusals(tmp_path):
    for argv in ([], ["c.stim"], ["c.stim", "--matrices", "H", "L"], ["--matrices", "H", "L", "--config", "dp=7"]):
        with pytest.raises(SystemExit):
            cli.main(argv)


def _oracle_dp(monkeypatch):
    """The dense DP's value and witness from the oracle (values from its executor, witness by brute force)."""
    from itertools import combinations

    def dp_distance(p, order, max_table_bytes=None, deadline=None):
        per = [oracle.run_plan_c(plan, "oracle") if plan.feasible else None for _, plan in auto._plans(p, order)]
        found = [v for v in per if v is not None]
        return (min(found) if found else None), per

    def dp_witness(p, order, obs, max_bytes=None):
        for w in range(1, p.ncols + 1):
            for cols in combinations(range(p.ncols), w):
                s = lg = 0
                for j in cols:
                    s ^= p.h[j]
                    lg ^= p.l[j]
                if s == 0 and (lg >> obs) & 1:
                    return list(cols)
        return None

    monkeypatch.setattr(auto, "dp_distance", dp_distance)
    monkeypatch.setattr(auto, "dp_witness", dp_witness)


def test_dense_route_records_equal_brute_force(monkeypatch):
    _oracle_dp(monkeypatch)
    rng = random.Random(11)
    checked = 0
    for trial in range(80):
        m, k, h, l = random_problem(rng)
        p = Problem("rand", m, k, h, l)
        want = oracle.brute_all(h, l)
        want = want[0] if want else None
        rec = pipeline.certify(p, config="dp=1")
        assert rec["notes"]["decisions"][0].startswith("dense DP")
        if rec["upper_bound"] is None:          # no logical can be flipped
            assert rec["notes"].get("no_logical") and rec["status"] == "bounds"
            continue
        assert rec["status"] == "exact" and rec["witness_ok"] and rec["lower_bound"] == rec["upper_bound"]
        assert rec["trace"][0] == {"t": rec["trace"][0]["t"], "lb": 1}
        if want is not None:
            assert rec["lower_bound"] == want
        checked += 1
    assert checked >= 40


def test_search_route_when_the_dense_table_does_not_fit(monkeypatch):
    """dp=0 (and dp=1 with a table over the budget) go to the search; here the search is answered by brute force,
    and the record takes its lower bound, witness and status."""
    from tropidist_amd import search

    def fake_distance(p, k_table=2, kmax=20, start=1, time_limit=0.0, log=None, **kw):
        from itertools import combinations
        for w in range(start, kmax + 1):
            for cols in combinations(range(p.ncols), w):
                s, lg = p.check(cols)
                if s == 0 and lg:
                    return {"k_table": k_table, "table_entries": 0, "distance": w, "lower_bound": w,
                            "solutions": [tuple(cols)], "runs": [{"K": w, "status": "found", "best": w,
                                                                  "seconds": 0.0}]}
        return {"k_table": k_table, "table_entries": 0, "distance": None, "lower_bound": kmax + 1, "solutions": [],
                "runs": [{"K": kmax, "status": "exhausted", "best": None, "seconds": 0.0}]}

    monkeypatch.setattr(search, "distance", fake_distance)
    p = Problem("rep3", 2, 1, [1, 3, 2], [1, 1, 1])          # repetition code: distance 3
    rec = pipeline.certify(p, config="dp=0")
    assert rec["status"] == "exact" and rec["lower_bound"] == rec["upper_bound"] == 3 and rec["witness_ok"]
    assert rec["notes"]["decisions"] == ["search from threshold 1"]
    rec = pipeline.certify(p, config="dp=0,kmax=2")
    assert rec["status"] == "bounds" and rec["lower_bound"] == 3 and rec["upper_bound"] is None


def test_record_maps_witness_to_file_columns(tmp_path, monkeypatch):
    _oracle_dp(monkeypatch)
    # columns 1 and 3 are duplicates / zero and are dropped; the witness is reported in the files' indices
    h, l = [1, 1, 3, 0, 2], [1, 1, 1, 0, 1]
    pipeline.write_mtx(str(tmp_path / "H.mtx"), h, 2)
    pipeline.write_mtx(str(tmp_path / "L.mtx"), l, 1)
    p = pipeline.load_matrices(str(tmp_path / "H.mtx"), str(tmp_path / "L.mtx"))
    rec = pipeline.certify(p, config="dp=1")
    assert rec["cols"] == 5 and rec["lower_bound"] == rec["upper_bound"] == 3
    s = lg = 0
    for j in rec["witness"]:
        s ^= h[j]
        lg ^= l[j]
    assert s == 0 and lg and len(rec["witness"]) == 3
    out = str(tmp_path / "rec" / "r.json")
    pipeline.save(rec, out)
    assert os.path.exists(out)
