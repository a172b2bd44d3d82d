"""Compare a slice written by `python -m mhdturbulence_amd.tprof --slice FILE` with the
statistics of the reference's t = 15 sample (vissimple_t15_stats.json): prints one row per
statistic, ours against the reference's, and their ratio.

    python tests/golden/compare_vissimple_t15.py FILE
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from vissimple_stats import COLUMNS, QUANTILES, read_slices, slice_stats  # noqa: E402


def main(argv):
    ours = slice_stats(read_slices(argv[1:2]))
    with open(os.path.join(HERE, "vissimple_t15_stats.json")) as f:
        ref = json.load(f)
    print("| statistic | this run | reference sample | ratio |")
    print("|---|---|---|---|")
    for name in COLUMNS:
        for k in ("min", "max", "mean", "std", *[f"q{p}" for p in QUANTILES]):
            a, b = ours[name][k], ref[name][k]
            print(f"| {name} {k} | {a:.4g} | {b:.4g} | {a / b if b else float('nan'):.3f} |")
    for k in ("mix", "vy_rms"):
        print(f"| {k} | {ours[k]:.4g} | {ref[k]:.4g} | {ours[k] / ref[k]:.3f} |")


if __name__ == "__main__":
    main(sys.argv)
