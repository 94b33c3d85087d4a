"""The matrices front end on the GPU: `python -m tropidist_amd --matrices H.mtx L.mtx` writes the record of the
reference's pipeline (bench/common.py:411-466), and each route -- dense DP, sparse meet in the middle, search --
certifies the recorded distances of the reference's dense audit (results/dense_audit_20261021) with a witness
checked against the matrices."""
import json
import os
import subprocess
import sys

import pytest

from helpers import dense_fixtures
from tropidist_amd import pipeline
from tropidist_amd.problem import Problem

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fixture(name):
    for r in dense_fixtures():
        if r["name"] == name:
            return r
    raise AssertionError(name)


def _check(rec, h, l, d):
    assert rec["status"] == "exact", rec["notes"]
    assert rec["lower_bound"] == rec["upper_bound"] == d and rec["witness_ok"]
    s = lg = 0
    for j in rec["witness"]:
        s ^= h[j]
        lg ^= l[j]
    assert s == 0 and lg != 0 and len(rec["witness"]) == d


def test_command_line_dense_route(tmp_path):
    r = _fixture("code_bb72")
    pipeline.write_mtx(str(tmp_path / "bb72.H.mtx"), r["h"], r["nrows"])
    pipeline.write_mtx(str(tmp_path / "bb72.L.mtx"), r["l"], r["nobs"])
    out = str(tmp_path / "bb72.json")
    print("case pipeline cli code_bb72", flush=True)
    run = subprocess.run([sys.executable, "-m", "tropidist_amd", "--matrices", str(tmp_path / "bb72.H.mtx"),
                          str(tmp_path / "bb72.L.mtx"), "--out", out, "--time-limit", "120"], cwd=ROOT,
                         capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stderr[-2000:]
    assert "distance %d, certified" % min(r["distances"]) in run.stdout
    rec = json.load(open(out))
    _check(rec, r["h"], r["l"], min(r["distances"]))
    assert rec["notes"]["decisions"][0].startswith("dense DP")
    assert rec["rows"] == r["nrows"] and rec["cols"] == len(r["h"]) and rec["obs"] == r["nobs"]


@pytest.mark.parametrize("name", ["code_bb72", "star_42"])
def test_sparse_route_when_no_dense_table_fits(name, monkeypatch):
    from tropidist_amd import auto
    r = _fixture(name)
    p = Problem(name, r["nrows"], r["nobs"], r["h"], r["l"])
    estimate = auto.estimate

    def too_wide(*a, **k):                  # the dense table is over the budget
        return dict(estimate(*a, **k), fits=False)

    monkeypatch.setattr(auto, "estimate", too_wide)
    print("case pipeline sparse route", name, flush=True)
    rec = pipeline.certify(p, config={"dp": 2, "kmax": 12}, truth=min(r["distances"]))
    _check(rec, r["h"], r["l"], min(r["distances"]))
    assert rec["notes"]["decisions"][0].startswith("sparse DP") and rec["agrees_with_truth"]
    sweeps = rec["notes"]["sparse"]["sweeps"]
    assert [s["K"] for s in sweeps] == list(range(1, min(r["distances"]) + 1))
    assert all(s["value"] is None for s in sweeps[:-1]) and sweeps[-1]["value"] == min(r["distances"])


@pytest.mark.parametrize("name", ["code_bb72", "surface_d5"])
def test_search_route(name):
    r = _fixture(name)
    p = Problem(name, r["nrows"], r["nobs"], r["h"], r["l"])
    print("case pipeline search route", name, flush=True)
    rec = pipeline.certify(p, config={"dp": 0, "kmax": 12}, time_limit=120.0)
    _check(rec, r["h"], r["l"], min(r["distances"]))
    assert rec["notes"]["decisions"] == ["search from threshold 1"]
