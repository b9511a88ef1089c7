"""Writes nve_drift_c1.json: a 10,000-step FP64 NVE run of the oracle on the 4,800-atom AB
bilayer graphene (30x20 cells per layer, seed 12345, 300 K), C.opt.tersoff C(O) + BNCH.ILP with
Tap, rc 16 A, skin 2 A, rebuild every 10 steps, dt 1 fs.  Records (step, potential energy,
kinetic energy) every 50 steps.  Run from the repository root: python tests/golden/make_nve_drift.py
(about 5 minutes on 8 cores)."""
import json
import os
import platform
import sys
import time

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "lammps-lmff_amd"))

import oracle  # noqa: E402
from lmff import structures as S  # noqa: E402

POT = os.path.join(REPO, "lammps-lmff_amd", "potentials")


def main(nsteps=10000, every=50, threads=8):
    s = S.config_c1()
    o = oracle.Oracle(1, os.path.join(POT, "C.opt.tersoff"), ["C(O)"], os.path.join(POT, "BNCH.ILP"), ["C"],
                      tap_flag=1, threads=threads)
    o.set_system(s)
    t0 = time.time()
    _, _, th = o.run_nve(s.x, s.v, s.mass_per_type, 0.001, nsteps, rebuild_every=10, thermo_every=every)
    out = {"system": "config_c1(seed=12345): 2 x (30x20 cells) AB bilayer graphene, 4800 atoms, 300 K",
           "potentials": "C.opt.tersoff C(O) + BNCH.ILP C, ilp cut 16, tap 1", "neighbor": "skin 2, every 10",
           "dt_ps": 0.001, "nsteps": nsteps, "thermo_every": every, "natoms": int(s.n),
           "libc": " ".join(platform.libc_ver()), "seconds": time.time() - t0,
           "step": th[:, 0].astype(int).tolist(), "pe": th[:, 1].tolist(), "ke": th[:, 2].tolist()}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "nve_drift_c1.json"), "w") as f:
        json.dump(out, f, indent=0)


if __name__ == "__main__":
    main()
