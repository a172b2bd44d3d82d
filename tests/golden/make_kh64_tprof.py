"""Write tests/golden/kh64_tprof.csv: the oracle's t-prof of the reference KH problem at
64^3 (src_cpp_kokkos_host/main.cpp:66-184 with its config.hpp constants) from t = 0 to 4.5,
sampled exactly as mhdturbulence_amd.tprof samples it (a chunk of round(0.05 / dt) steps,
dt of the chunk's last step, then one t-prof row; the step that passes t = 4.5 ends it).

Columns: step count, time, mix, sqrt(<vy^2>) (RealTimeAnalysis, src_cpp_omp_device/main.cpp:
23-69), in %.17g.  About 12,700 steps; the oracle needs some 20 minutes on 8 cores.
Run from the repository root:  python tests/golden/make_kh64_tprof.py
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as orc  # noqa: E402

N, T_END, INTERVAL = 64, 4.5, 0.05
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kh64_tprof.csv")


def main():
    orc.lib().orc_set_threads(os.cpu_count() or 1)
    o = orc.Oracle(nx=N, ny=N, nz=N)
    o.generate_grid()
    o.generate_problem_kh()
    d = o.diagnostics()
    rows = [(0, 0.0, d[1], d[2])]
    t, dts = o.run(1, 0.0)
    steps = 1
    per = max(1, round(INTERVAL / dts[0]))
    while t < T_END:
        n, last = 0, None
        while n < per:
            t, lg = o.run(1, t)
            n += 1
            last = lg[0]
            if t > T_END:
                break
        steps += n
        d = o.diagnostics()
        rows.append((steps, t, d[1], d[2]))
        if n < per:
            break
        per = max(1, round(INTERVAL / last))
    with open(DST, "w") as f:
        f.write(f"# oracle t-prof, KH {N}^3 to t={T_END}: steps time mix sqrt(<vy^2>)\n")
        for r in rows:
            f.write(f"{r[0]} {r[1]:.17g} {r[2]:.17g} {r[3]:.17g}\n")


if __name__ == "__main__":
    main()
