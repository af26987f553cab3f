"""Writes the fixture of the dam-break trajectory the GPU time stepper is
compared against: the oracle's own steps (oracle/step.py, the C restatement
with direct momentum solves and its PCG-multigrid to tol) on the 512 x 256
dam break with its rectangular boom, FP64, tol 1e-10, 500 steps.

    python tests/golden/make_dambreak_c1.py [steps] [tol]

A second run at tol 1e-12 (file suffix _tol1e-12) measures how far the
oracle's own trajectory moves when only the solve's rounding changes: the
envelope the GPU's trajectory is held to.

Stores per step: dt, liquid volume, gas centroid (cells), the liquid length
of the bottom row (cells); and the final volume fraction.  Runs in this
container in about half an hour (the oracle is single-threaded outside its
PCG)."""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from oracle import step as S  # noqa: E402
from simcoflow_amd import cases as C  # noqa: E402


def diagnostics(g, vof, vofS):
    area = g.dx[None, :] * g.dy[:, None]
    cx, cy = S.gas_centroid(g, vof, vofS)
    return float((vof * area).sum()), cx, cy, float(vof[0, :].sum())


def main(steps=500, tol=1e-10):
    O.build()
    case = C.case_dambreak(512, 256)
    g = case.grid
    st = S.OracleStepper(O, case, tol=tol)
    rows = []
    t0 = time.time()
    for k in range(steps):
        r = st.step()
        rows.append((r["dt"],) + diagnostics(g, st.vof, case.geo.vofS))
        if k % 25 == 0:
            print(f"step {k}: dt {r['dt']:.3e} its {r['iterations']} ({time.time() - t0:.0f} s)", flush=True)
    rows = np.array(rows)
    suffix = "" if tol == 1e-10 else f"_tol{tol:.0e}"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"dambreak_512x256_{steps}{suffix}.npz")
    np.savez_compressed(out, dt=rows[:, 0], volume=rows[:, 1], centroid_x=rows[:, 2], centroid_y=rows[:, 3],
                        bottom_length=rows[:, 4], vof=st.vof, time=st.time, steps=steps, tol=tol)
    print(f"wrote {out} ({os.path.getsize(out)} bytes)")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 500, float(sys.argv[2]) if len(sys.argv) > 2 else 1e-10)
