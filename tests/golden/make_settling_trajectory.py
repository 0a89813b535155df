"""Writes tests/golden/settling_trajectory.npz: the CPU oracle's run of the
single-sphere settling case (64x64x256, six no-slip walls, one sphere of
D = 16h with 800 markers, rho_p/rho_f = 1.5, Re_g = 40) for 1000 steps.

Per step it records the body's position, velocity and angular velocity, the
fluid's kinetic energy sum(u^2 + v^2 + w^2) over the interior arrays and
the CG iteration count, starting from the seeded initial state of
ibmdem.cases.config_settling(). The GPU test replays the same 1000 steps and
compares against these vectors.

Run from the repository root (about 5 minutes on 8 cores):
    OMP_NUM_THREADS=8 python tests/golden/make_settling_trajectory.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "ibm-dem_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from ibmdem import cases  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "settling_trajectory.npz")


def main(steps: int = 1000) -> None:
    setup = cases.config_settling()
    xs, vs, ws, ke, its = [], [], [], [], []
    t0 = time.time()
    with oracle.simulation(setup.case) as s:
        cases.load(s, setup)
        for k in range(steps):
            info = s.step()
            b = s.download_bodies()
            u, v, w, _ = s.download_fields()
            xs.append(b.x[:, 0].copy())
            vs.append(b.v[:, 0].copy())
            ws.append(b.omega[:, 0].copy())
            ke.append(float((u * u).sum() + (v * v).sum() + (w * w).sum()))
            its.append(info.iterations)
            if k % 100 == 99:
                print(f"step {k + 1}: z = {xs[-1][2]:.6f}, w = {vs[-1][2]:.6e}, "
                      f"{time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(OUT, x=np.array(xs), v=np.array(vs), omega=np.array(ws), ke=np.array(ke),
                        cg_iterations=np.array(its, dtype=np.int32), steps=np.int32(steps))
    print("wrote", OUT)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1000)
