"""Writes brio_wu_800x4_mp5_t0.1.npz: the oracle's full Brio-Wu run, the
CPU-runnable benchmark configuration (800 x 4 cells on x in [0,1], y periodic
with square cells, gamma = 2, MP5 + HLLD + CT, SSP-RK3, CFL 0.4, auto dt to
t = 0.1), as the fixture the GPU run is compared with.

The run is the oracle's own auto-dt loop (cfl_limit, then step_unchecked,
the last step clipped to t_end; python/quasar/mhd/cli.py:386-491).  Stored:
the interior of the 8 conserved planes (rho, mx, my, mz, energy, bx_face,
by_face, bz_cell; row-major ny x nx), the step count, the final time and the
dt of every step.  Run from the repository root:

    python tests/golden/make_brio_wu_reference.py
"""
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from oracle_ctypes import Oracle, OracleSolver, ic_spec  # noqa: E402

NX, NY, NG = 800, 4, 3
T_END = 0.1
OUT = HERE / "brio_wu_800x4_mp5_t0.1.npz"


def main():
    oracle = Oracle()
    g = oracle.grid(NX, NY, NG, 1.0, NY / NX)
    spec = ic_spec("brio_wu", 2.0, left=[1, 1, 0, 0, 0, 0.75, 1, 0], right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
    st, planes = oracle.seed(g, spec)
    assert st == 0
    s = OracleSolver(oracle, g, 2.0, 0.4, 5, fluid=(1, 1, 0, 0))
    s.seed(planes)
    t, dts = 0.0, []
    t0 = time.time()
    while t < T_END:
        dt = min(s.cfl_limit(), T_END - t)
        clipped = dt == T_END - t
        s.step_unchecked(dt)
        dts.append(dt)
        t = T_END if clipped else t + dt
        if len(dts) % 200 == 0:
            print(f"step {len(dts)} t={t:.6f} ({time.time() - t0:.0f} s)", flush=True)
    state = s.state()
    pitch = NX + 2 * NG
    interior = np.stack([p.reshape(NY + 2 * NG, pitch)[NG:NG + NY, NG:NG + NX] for p in state])
    np.savez_compressed(OUT, interior=interior, steps=np.array([len(dts)]), final_time=np.array([t]),
                        dts=np.array(dts), divb=np.array([s.divergence_b_max()]))
    print(f"wrote {OUT}: {len(dts)} steps to t={t} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
