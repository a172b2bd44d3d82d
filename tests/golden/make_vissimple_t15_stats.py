"""Write tests/golden/vissimple_t15_stats.json from the reference's t = 15 sample.

Source: analysis/vissimple/snap00[0-3]-00101.csv of the reference: the four MPI ranks'
75 x 75 parts of the z-slice at the first interior plane of a 150^3 Kelvin-Helmholtz run
(Fortran variant: gamma 5/3, unshifted faces, a random vx kick) at t = 15, columns
"x y d vx vy p phi X1", three significant digits.  The fixture holds the slice's statistics
(vissimple_stats.slice_stats), not the cells.  Run from the repository root:
    python tests/golden/make_vissimple_t15_stats.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from vissimple_stats import read_slices, slice_stats  # noqa: E402

SRC = [f"/root/reference/analysis/vissimple/snap00{r}-00101.csv" for r in range(4)]
DST = os.path.join(HERE, "vissimple_t15_stats.json")


def main():
    a = read_slices(SRC)
    stats = slice_stats(a)
    stats["source"] = "analysis/vissimple/snap00[0-3]-00101.csv (t = 15, 150^3, first interior z-plane)"
    with open(DST, "w") as f:
        json.dump(stats, f, indent=1, sort_keys=True)
    print(f"{a.shape[0]} cells -> {DST}")


if __name__ == "__main__":
    sys.exit(main())
