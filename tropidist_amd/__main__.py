"""Certified distance of a pair of matrices on the GPU, the command line of the reference's `python -m cdist`
(cdist/__main__.py:65-104) for its --matrices form.

    python -m tropidist_amd --matrices H.mtx L.mtx [--out RECORD.json] [--time-limit SECONDS] [--config KEY=VALUE,...]

H and L are MatrixMarket coordinate files over the same columns (for a CSS code, H = H_X and L = L_X give d_Z).  The
record holds the certified lower and upper bounds, the time of each, and the fault set that attains the upper bound
(tropidist_amd.pipeline).  Stim circuits are not read by this build (it has no circuit builder)."""
import argparse
import os
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tropidist_amd", description=__doc__.split("\n\n")[0])
    ap.add_argument("circuit", nargs="?", help="Stim circuit file (not supported by this build)")
    ap.add_argument("--matrices", nargs=2, metavar=("H.mtx", "L.mtx"), help="constraint and logical matrices")
    ap.add_argument("--out", help="record file (default: NAME.distance.json in the current directory)")
    ap.add_argument("--threads", type=int, default=1, help="accepted for the reference's command line; unused")
    ap.add_argument("--time-limit", type=float, default=0.0, help="seconds; 0 means no limit")
    ap.add_argument("--config", default="", help="dp=0|1|2, kmax, kmin, k_table, table_gb (default dp=2)")
    a = ap.parse_args(argv)
    if (a.circuit is None) == (a.matrices is None):
        ap.error("give either a circuit file or --matrices H.mtx L.mtx")
    if a.circuit is not None:
        ap.error("Stim circuits are not supported by this build; export the problem as --matrices H.mtx L.mtx")
    from . import pipeline
    try:
        cfg = pipeline.configuration(a.config)
    except ValueError as exc:
        ap.error(str(exc))
    name = os.path.splitext(os.path.basename(a.matrices[0]))[0]
    out = os.path.abspath(a.out or name + ".distance.json")
    p = pipeline.load_matrices(a.matrices[0], a.matrices[1], name=name)
    rec = pipeline.certify(p, time_limit=a.time_limit, config=cfg, log=lambda m: print(m, flush=True))
    pipeline.save(rec, out)
    lower, upper = rec["lower_bound"], rec["upper_bound"]
    if lower is not None and lower == upper:
        print("distance %d, certified (%s, %.1f s); record in %s" % (lower, rec["status"], rec["seconds"], out))
    else:
        print("distance between %s and %s (%s, %.1f s); record in %s" % (lower, upper, rec["status"], rec["seconds"],
                                                                         out))
    return 0 if rec["status"] != "error" else 1


if __name__ == "__main__":
    sys.exit(main())
