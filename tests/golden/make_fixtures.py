"""Extract golden vectors for the dense DP from the reference's recorded audit (data only, no code):

    results/dense_audit_20261021/inputs/<inst>/<variant>.json.gz     h / l hex columns, truth
    results/dense_audit_20261021/widths/<inst>__<variant>.json.gz    column order, per-observable peak_bits / work
    results/dense_audit_20261021/runs/<inst>__<variant>.json         per-observable distances

into tests/golden/dense_audit.json.gz.  Run in a container that has the reference tree:

    python tests/golden/make_fixtures.py [/root/reference]
"""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

# (instance, variant): chosen to cover table widths 4..28 and 1..12 observables
CASES = [
    ("code_bb72", "full"), ("code_bb90", "full"), ("code_bb108", "full"), ("code_bb144", "full"),
    ("color_d3", "full"), ("color_d5", "full"), ("fx_bb72_code", "full"),
    ("fx_sdbb16_r3", "collapsed"), ("fx_sdbb16_r3", "sector"), ("star_42", "collapsed"),
    ("routing_54", "collapsed"), ("surface_d5", "sector"), ("surface_d7", "sector"),
    ("surface_d11", "collapsed"), ("surface_d25", "collapsed"), ("bb72_r6", "collapsed"),
    ("sdbb_64", "collapsed"), ("walkingcat_54", "collapsed"),
]


def main(ref):
    base = os.path.join(ref, "results", "dense_audit_20261021")
    out = []
    for inst, variant in CASES:
        inp = json.load(gzip.open(os.path.join(base, "inputs", inst, variant + ".json.gz")))
        wid = json.load(gzip.open(os.path.join(base, "widths", "%s__%s.json.gz" % (inst, variant))))
        run_path = os.path.join(base, "runs", "%s__%s.json" % (inst, variant))
        run = json.load(open(run_path)) if os.path.exists(run_path) else None
        rec = {
            "name": inst, "variant": variant, "nrows": inp["nrows"], "nobs": inp["nobs"],
            "h": inp["h"], "l": inp["l"], "truth": inp.get("truth"),
            "order": wid["order"], "profile": wid["profile"], "peak_bits": wid["peak_bits"],
            "total_work": wid["total_work"],
            "plans": [{k: p[k] for k in ("observable", "peak_bits", "work", "feasible")} for p in wid["plans"]],
            "distances": None if run is None else [p.get("distance") for p in run["per_observable"]],
            "source": "results/dense_audit_20261021 (%s, %s)" % (inst, variant),
        }
        out.append(rec)
        print(inst, variant, "peak", rec["peak_bits"], "obs", rec["nobs"], "d", rec["truth"])
    with gzip.open(os.path.join(HERE, "dense_audit.json.gz"), "wt") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
