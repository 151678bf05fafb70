"""Writes tests/golden/disk_n1024.npz: a 1,024-disk state (φ = 0.80) and the oracle's outputs on it.

The state: N=1024 bidisperse disks (diameters 1.0/1.4, 50/50, large first), packing
fraction 0.80, harmonic e_c=1, m=1, seeded uniform positions relaxed by exactly 200
FIRE steps (alpha0=0.1, dt=0.01), Gaussian velocities at T=1e-4 (mean removed,
rescaled), cell list with rc=1.7 and rebuild at max displacement > 0.15.

Stored: the inputs, the static order after neighbour set-up, the neighbour pairs
(static ids, in list order), forces/energies, pressure, contacts, and the state after
100 NVE steps at dt=0.01.  The reference ships no golden data; these vectors freeze
the oracle (itself pinned by tests/test_oracle_pins.py) so the GPU path is checked
against fixed numbers as well as the live oracle.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from _helpers import oracle, pkg, workloads  # noqa: E402


def main():
    P, W = pkg(), workloads()
    st = W.disk_config(1024, 0.80, seed=12345)
    spec = W.neighbor_spec(st.side)
    s = W.make_disk_system(P.System, st, spec, backend=oracle())
    s.minimize_fire(P.fire_params(max_steps=200))
    x0, y0 = s.download(P.F_X), s.download(P.F_Y)
    vx0, vy0 = W.gaussian_velocities(st.n, 1e-4, 12345)
    # fresh system from the relaxed state (static order = upload order)
    s = W.make_disk_system(P.System, W.DiskState(st.n, st.side, x0, y0, st.rad, st.mass),
                           spec, backend=oracle())
    s.upload(P.F_VX, vx0)
    s.upload(P.F_VY, vy0)
    s.zero_forces()
    s.compute_forces()
    pi, pj = s.neighbor_pairs()
    out = dict(
        side=st.side, cutoff=spec.cutoff, disp_thresh_sq=spec.disp_thresh_sq,
        n_cells_dim=spec.n_cells_dim, cell_disp_thresh_sq=spec.cell_disp_thresh_sq,
        x=x0, y=y0, vx=vx0, vy=vy0, rad=st.rad, mass=st.mass,
        static_index=s.download(P.F_STATIC_INDEX).astype(np.int32),
        pair_i=pi.astype(np.int32), pair_j=pj.astype(np.int32),
        fx=s.download(P.F_FX), fy=s.download(P.F_FY), pe=s.download(P.F_PE),
        pressure=s.reduce(P.RED_PRESSURE), contacts=s.reduce(P.RED_CONTACTS),
    )
    s.nve_run(0.01, 100)
    out.update(x100=s.download(P.F_X), y100=s.download(P.F_Y), vx100=s.download(P.F_VX),
               vy100=s.download(P.F_VY), pe100=s.reduce(P.RED_PE), ke100=s.reduce(P.RED_KE))
    np.savez_compressed(os.path.join(HERE, "disk_n1024.npz"), **out)
    print("pairs", len(pi), "pe", out["pe"].sum(), "pe100", out["pe100"])


if __name__ == "__main__":
    main()
