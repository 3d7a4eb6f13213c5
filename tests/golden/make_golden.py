"""Write the golden vectors of the plate workload (8x16 rings, 5 deg,
plunge 0.1c at k = 0.5, dt = 1/8, Euler, steps 0..200) from the oracle.

    python tests/golden/make_golden.py

The reference cannot be built here (Eigen/glog/gflags/protobuf and its
multiple_sheet submodule are absent), so these vectors come from the oracle,
which is pinned separately by the reference's known-answer tests and by
closed forms; the fixture freezes it so that the GPU tests, without the
oracle's run time, still compare against the full 200-step history.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from uvlm_amd import workloads as W  # noqa: E402

SNAP_STEPS = (0, 1, 2, 10, 50, 100, 150, 200)


def main():
    case = W.plate_c1(steps=200)
    sim = oracle.Sim(case)
    loads, gammas = [], {}
    for s in range(case["steps"] + 1):
        sim.step(s)
        loads.append(sim.loads().copy())
        if s in SNAP_STEPS:
            gammas[s] = sim.gamma().copy()
    nodes, gam = sim.wake()
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plate_c1_200.npz")
    np.savez_compressed(out, loads=np.array(loads), snap_steps=np.array(SNAP_STEPS),
                        gammas=np.stack([gammas[s] for s in SNAP_STEPS]),
                        wake_nodes=nodes, wake_gamma=gam, pairs=sim.pairs)
    print("wrote", out, "pairs", sim.pairs)


if __name__ == "__main__":
    main()
