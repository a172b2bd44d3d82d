"""Write tests/golden/vissimple_t0.csv from the reference's sample snapshot.

Source: analysis/vissimple/snap000-00001.csv of the reference (MPI rank 0 of a 150^3
Kelvin-Helmholtz run, 75x75 slice at k=3, t=0, columns "x y d vx vy p phi X1", three
significant digits).  The fixture keeps every 5th data row and the columns
x, y, d, vy, p, X1 (vx carries a per-(j,k) random kick that no restatement can
reproduce).  Run from the repository root:  python tests/golden/make_vissimple_fixture.py
"""
import os
import sys

SRC = "/root/reference/analysis/vissimple/snap000-00001.csv"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vissimple_t0.csv")


def main():
    rows = []
    with open(SRC) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            v = line.split()
            if len(v) != 8:
                continue
            rows.append(v)
    keep = rows[::5]
    with open(DST, "w") as f:
        f.write("# from analysis/vissimple/snap000-00001.csv (every 5th row): x y d vy p X1\n")
        for v in keep:
            f.write(" ".join([v[0], v[1], v[2], v[4], v[5], v[7]]) + "\n")
    print(f"{len(rows)} rows read, {len(keep)} written to {DST}")


if __name__ == "__main__":
    sys.exit(main())
