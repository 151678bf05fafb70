#!/usr/bin/env python3
"""Headline benchmark: FP64 NVE of N=262,144 bidisperse disks on one MI355X.

Workload: 50/50 diameters 1.0/1.4 at packing fraction 0.85 in a periodic
square box, harmonic contacts (e_c=1, n_c=2, m=1), seeded uniform positions relaxed
by 200 FIRE steps, Gaussian velocities at T=1e-5, dt=0.01, cell list with
rc = 1.4 + 0.3 and rebuild when the max displacement exceeds 0.15.

One step = one velocity-Verlet step of the hot path over all N particles: one
launch of the fused force + kick + kick + drift kernel, which also rebuilds the
neighbour lists in place when a step needs it; a step that needs a cell rebuild
halts its chunk and is finished by the rebuild chain.  `value` is particle-steps/s
over exactly --steps steps with the state resident in HBM, bracketed by device
synchronisation.

The roofline object covers the fused step kernel (the dominant one; one launch is
one NVE step of all N particles).  Algorithmic bytes per launch use the fixed
per-particle-step figure for a fused NVE step, 196 + 4*k_mean (kick+drift 104, force
+ closing kick 92 + 4*k_mean; neighbour gathers not counted again), times N.  The
kernel's own compulsory traffic is lower, 132 + 4*k_mean (it never stores the forces
of intermediate steps), and is reported beside it.  The time is the kernel's total
duration from HIP events recorded on the system's own stream around every chunk of
steps (a hipGraph of fused launches) inside the timed region, divided by the number
of steps: launches that only skip (the rest of a chunk after a step halted for a
cell rebuild) are charged to the steps, not averaged in.

cpu_baseline: the CPU oracle (C++ restatement of the reference, OpenMP) timed on a
bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "FP64 particle-steps/s at N=262k; batched jams/s & p99 latency vs plain loop"
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_pkg():
    P = _load("dpcuda_v2_amd", os.path.join(ROOT, "dpcuda-v2_amd", "__init__.py"))
    W = _load("dpcuda_v2_amd_workloads", os.path.join(ROOT, "dpcuda-v2_amd", "workloads.py"))
    return P, W


def prepare(P, W, n, seed, backend=None, fire_steps=200, temperature=1e-5):
    st = W.disk_config(n, 0.85, seed=seed)
    spec = W.neighbor_spec(st.side)
    s = W.make_disk_system(P.System, st, spec, backend=backend)
    s.minimize_fire(P.fire_params(max_steps=fire_steps))
    vx, vy = W.gaussian_velocities(n, temperature, seed)
    s.upload(P.F_VX, vx)
    s.upload(P.F_VY, vy)
    s.zero_forces()
    s.compute_forces()
    return s


def cpu_baseline(P, W, n, seed, steps):
    """The oracle on this host's cores (bounded sample of the same workload)."""
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        return None
    be = P.Backend(oracle_so, "orc_", gpu=False)
    s = prepare(P, W, n, seed, backend=be)
    t0 = time.perf_counter()
    s.nve_run(0.01, steps)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{steps} NVE steps of the same N={n} workload (after the same 200 "
                      f"FIRE steps), C++ oracle with OpenMP, {threads} threads, "
                      f"{dt:.2f} s"}


# ---------------------------------------------------------------- rigid bumpy FIRE
RIGID_N = 65536


def prepare_rigid(P, W, n, seed, backend=None):
    st = W.bumpy_config(n, 0.80, seed=seed)
    return st, W.make_bumpy_system(P.System, st, backend=backend,
                                   e_c=W.bumpy_contact_scale(st, 1.0))


def rigid_fire(args, P, W):
    """FIRE of N=65,536 rigid bumpy particles to |F|max < 1e-12 (one device-resident
    minimizeFire).  value = particle FIRE steps per second over the whole
    minimisation."""
    n = args.n if args.n != 262144 else RIGID_N
    print(f"[bench] rigid bumpy FIRE N={n}", file=sys.stderr, flush=True)
    st, s = prepare_rigid(P, W, n, args.seed)
    s.set_option(P.OPT_GRAPH, 0 if args.no_graph else 1)
    s.set_option(P.OPT_CHUNK, args.chunk)
    prm = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-12,
                        max_steps=args.fire_max_steps)
    s.synchronize()
    r0 = s.neighbor_stats()
    t0 = time.perf_counter()
    r = s.minimize_fire(prm)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    r1 = s.neighbor_stats()
    s.compute_stress()
    pressure = s.reduce(P.RED_PRESSURE)
    pe = s.reduce(P.RED_PE)
    phi = s.reduce(P.RED_PACKING_FRACTION)
    # per-launch events on a converged state (no rebuilds): the force phase of a step
    # (vertex forces + per-particle sums and closing kick) and the move phase
    s.set_option(P.OPT_TIMING, 1)
    s.reset_kernel_time()
    probe = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=0.0,
                          max_steps=args.steps_probe)
    s.minimize_fire(probe)
    f_ms, f_n = s.kernel_time(P.K_FORCE)
    m_ms, m_n = s.kernel_time(P.K_DRIFT)
    s.set_option(P.OPT_TIMING, 0)
    nv = st.nv
    kv = len(s.vertex_pairs()[0]) / nv
    force_us = f_ms / max(f_n, 1) * 1e3
    move_us = m_ms / max(m_n, 1) * 1e3
    # algorithmic bytes of the force phase of one step (forces, torques, energies, their
    # per-particle sums, the closing kick and the reduction partials): per vertex 16
    # (position) + 4 (owner) + 4 (row length) + 4 k_v (row; the neighbours' positions
    # are the same array, not counted again); per particle 32 (position record: lever
    # arm and mass) + 8 (vertex range) + 32 (force, torque, energy written) + 48
    # (velocity and spin read and written) + 8 (moment of inertia)
    force_bytes = nv * (24.0 + 4.0 * kv) + n * (32 + 8 + 32 + 48 + 8)
    force_gbs = force_bytes / (force_us * 1e-6) / 1e9 if force_us > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
        if os.path.exists(oracle_so):
            be = P.Backend(oracle_so, "orc_", gpu=False)
            _, so = prepare_rigid(P, W, n, args.seed, backend=be)
            k = args.cpu_fire_steps
            c0 = time.perf_counter()
            so.minimize_fire(P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=1e-12, max_steps=k))
            cdt = time.perf_counter() - c0
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            cpu = {"value": n * k / cdt, "unit": "particle-FIRE-steps/s", "cores": threads,
                   "kind": "port",
                   "sample": f"the first {k} FIRE steps of the same N={n} minimisation, C++ "
                             f"oracle with OpenMP, {threads} threads, {cdt:.2f} s"}
    out = {
        "metric": "FP64 rigid bumpy FIRE particle-steps/s to |F|max<1e-12 at N=65,536",
        "value": n * r.steps / elapsed,
        "unit": "particle-FIRE-steps/s",
        "n_gpus": 1,
        "steps": int(r.steps),
        "warmup": 0,
        "ms_per_step": elapsed / max(r.steps, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform positions and orientations)",
        "config": {
            "workload": "FP64 FIRE of N=65536 rigid bumpy particles (diameters "
                        "1.0 / 1.4 with 20 / 28 vertices, 50/50), phi=0.80, vertex radius "
                        "from the touching rule, contact scale (2 r_v / d_min)^2, to "
                        "|F|max < 1e-12, cell list",
            "n_particles": n,
            "n_vertices": nv,
            "converged": bool(r.converged),
            "time_to_converge_s": elapsed,
            "final_pe_per_particle": pe / n,
            "final_pressure": pressure,
            "final_phi": phi,
            "final_fmax": r.fmax_sq ** 0.5,
            "rebuilds": r1["n_rebuilds"] - r0["n_rebuilds"],
            "cell_rebuilds": r1["n_cell_rebuilds"] - r0["n_cell_rebuilds"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_bfire_forces + k_bfire_partials (force phase of one FIRE step)",
            "achieved": force_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": force_gbs / HBM_PEAK_GBS if force_gbs else None,
            # calibrated FETCH_SIZE + WRITE_SIZE of k_bfire_forces (the partials kernel
            # beside it moves under 2 MB)
            "traffic": traffic_from_profiles(n, "k_bfire_forces_rigid_n65536_traffic.json"),
            "algorithmic_bytes_per_launch": force_bytes,
            "force_phase_us": force_us,
            "move_phase_us": move_us,
            "vertex_row_mean": kv,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- NVE with logging
def nve_logged(args, P, W):
    """NVE of N=262,144 disks run by the run_nve driver in the reference's run-directory form: the
    FIRE-relaxed state is written here as a reference-format input directory
    (system/{init,restart}/*.dat in %.17e, particle and log configs), loaded by
    load_configs / loadParticle, and run with the reference's log groups (energy.csv every
    100 steps, positions and velocities every 1000, restart every 10000).  value =
    particle-steps/s with every log written; the same driver run with only the final log
    gives the rate without logging."""
    import shutil
    import subprocess
    import tempfile
    import numpy as np
    n = args.n
    steps = args.steps
    print(f"[bench] preparing N={n} run directory", file=sys.stderr, flush=True)
    s = prepare(P, W, n, args.seed)
    cols = {"positions": (s.download(P.F_X), s.download(P.F_Y)),
            "velocities": (s.download(P.F_VX), s.download(P.F_VY))}
    rad, mass = s.download(P.F_RAD), s.download(P.F_MASS)
    side = s.box()[0]
    del s
    tmp = tempfile.mkdtemp(prefix="nve_logged_", dir=os.environ.get("TMPDIR", "/tmp"))
    drv = os.path.join(ROOT, "dpcuda-v2_amd", "_build", "drivers", "run_nve")

    def logc(names, group, freq):
        return {"log_names": names, "save_style": "lin", "save_freq": freq,
                "reset_save_decade": 10, "min_save_decade": 10, "multiple": 0,
                "group_name": group, "decade": 10}

    def run(tag, energy_every, state_every):
        out = os.path.join(tmp, "out_" + tag)
        os.makedirs(os.path.join(out, "system"))
        for g, c in {"console": logc(["step", "PE/N", "KE/N", "T"], "console", 10**9),
                     "energy": logc(["step", "PE/N", "KE/N", "TE/N", "T", "phi", "P", "Zp"],
                                    "energy", energy_every),
                     "state": logc(["positions", "velocities"], "state", state_every)}.items():
            json.dump(c, open(os.path.join(out, "system", f"{g}_log_config.json"), "w"),
                      indent=4)
        rd = os.path.join(tmp, "run_" + tag)
        os.makedirs(os.path.join(rd, "system"))
        json.dump({"input_dir": src, "output_dir": out, "minimize": False, "num_steps": steps,
                   "dt_dimless": 0.01, "temperature": 0.0},
                  open(os.path.join(rd, "system", "run_config.json"), "w"), indent=4)
        print(f"[bench] run_nve {tag} ({steps} steps)", file=sys.stderr, flush=True)
        p = subprocess.run([drv, rd], capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise SystemExit("run_nve failed: " + p.stderr[-2000:])
        line = [l for l in p.stdout.splitlines() if l.startswith("particle-steps/s")][-1]
        frames = len(os.listdir(os.path.join(out, "trajectories")))
        rows = len(open(os.path.join(out, "system", "energy.csv")).read().splitlines()) - 1
        mb = sum(os.path.getsize(os.path.join(dp, f))
                 for dp, _, fs in os.walk(out) for f in fs) / 1e6
        return (float(line.split()[1]), float(line.split("(logging included, ")[1].split()[0]),
                frames, rows, mb)

    try:
        src = os.path.join(tmp, "src")
        for sub in ("init", "restart"):
            d = os.path.join(src, "system", sub)
            os.makedirs(d)
            for name, (a, b) in cols.items():
                np.savetxt(os.path.join(d, name + ".dat"), np.column_stack([a, b]),
                           fmt="%.17e", delimiter="\t")
            np.savetxt(os.path.join(d, "radii.dat"), rad, fmt="%.17e")
            np.savetxt(os.path.join(d, "masses.dat"), mass, fmt="%.17e")
            np.savetxt(os.path.join(d, "box_size.dat"), [side, side], fmt="%.17e")
            with open(os.path.join(d, "step.dat"), "w") as f:
                f.write("0")
        pc = {"particle_type": "Disk", "n_particles": n, "packing_fraction": 0.85,
              "particle_dim_block": 256, "e_c": 1.0, "n_c": 2.0, "mass": 1.0,
              "size_ratio": 1.4, "count_ratio": 0.5, "seed": args.seed,
              "neighbor_list_config": {"neighbor_cutoff_multiplier": 1.5,
                                       "neighbor_displacement_multiplier": 0.2,
                                       "num_particles_per_cell": 8.0,
                                       "cell_displacement_multiplier": 0.5,
                                       "neighbor_list_update_method": "cell"}}
        json.dump(pc, open(os.path.join(src, "system", "particle_config.json"), "w"), indent=4)
        quiet = run("final_log_only", 10**9, 10**9)
        logged = run("logged", 100, 1000)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    value, secs, frames, rows, mb = logged
    print(json.dumps({
        "metric": "FP64 NVE particle-steps/s at N=262,144 with the reference's logs written",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": secs / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (the N=262144 disk state written as a reference-format input directory)",
        "config": {
            "workload": "disks (N=262144, phi=0.85, FIRE-relaxed) NVE dt_dimless "
                        "0.01 through run_nve <run directory> with the standard neighbour "
                        "config (cutoff 1.5 d_max, cell list): energy.csv every 100 steps "
                        "(PE/N, KE/N, TE/N, T, phi, P, Zp), positions + velocities every "
                        "1000, restart every 10000, final forced log",
            "n_particles": n,
            "final_log_only_particle_steps_per_s": quiet[0],
            "logging_overhead_fraction": 1 - value / quiet[0],
            "trajectory_frames": frames,
            "energy_rows": rows,
            "megabytes_written": mb,
        },
        "roofline": None,
        "cpu_baseline": None,
    }))


# ---------------------------------------------------------------- pair analytics
def pairs(args, P, W):
    """Per-pair analytics (dpm_pair_analytics: energies, forces, separations, ids, overlaps,
    angles, 2x2 Hessian blocks for every neighbour-list slot, in the reference's layout) on
    the N=262,144 disk state (phi 0.85).  value = neighbour-list slots per second; the kernel writes every
    slot of every field, so it is bound by HBM writes."""
    import numpy as np
    n = args.n
    print(f"[bench] preparing N={n} for pair analytics", file=sys.stderr, flush=True)
    s = prepare(P, W, n, args.seed)
    s.nve_run(0.01, 200)
    import ctypes
    r = ctypes.c_int64(0)
    for _ in range(3):
        s._c("pair_analytics", ctypes.byref(r))
    row = r.value
    st = s.neighbor_stats()
    used = st["total_entries"]
    k = args.steps if args.steps < 5000 else 200
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        s._c("pair_analytics", ctypes.byref(r))
    s.synchronize()
    dt = time.perf_counter() - t0
    slots = n * row
    # algorithmic bytes: 18 fields x 8 B written per slot (fill values included, as the
    # reference lays them out); per used slot the neighbour index (4), P0[j] (32) and its
    # static id (4); per row nnum (4), P0[i] (32) and sidx[i] (4)
    alg = slots * 144.0 + used * 40.0 + n * 40.0
    us = dt / k * 1e6
    gbs = alg / (us * 1e-6) / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
        if os.path.exists(oracle_so):
            be = P.Backend(oracle_so, "orc_", gpu=False)
            so = prepare(P, W, n, args.seed, backend=be)
            so.nve_run(0.01, 200)
            kc = 3
            c0 = time.perf_counter()
            for _ in range(kc):
                so._c("pair_analytics", ctypes.byref(r))
            cdt = time.perf_counter() - c0
            cpu = {"value": slots * kc / cdt, "unit": "pair-slots/s", "cores": 1,
                   "kind": "port",
                   "sample": f"{kc} full pair-analytics passes of the same N={n} state, "
                             f"C++ oracle (one thread), {cdt:.2f} s"}
    out = {
        "metric": "FP64 per-pair analytics (force/distance/Hessian pairs) neighbour slots/s at N=262,144",
        "value": slots / (dt / k),
        "unit": "pair-slots/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": 3,
        "ms_per_step": dt / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (N=262144 disk state after 200 FIRE and 200 NVE steps)",
        "config": {
            "workload": "Disk::calculateForceDistancePairs on N=262144 bidisperse "
                        "disks, phi=0.85, cell list rc=1.7; 18 per-slot fields",
            "n_particles": n,
            "row_length": row,
            "used_slots": used,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_pairs_disk",
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": alg,
            "call_us": us,
            "note": "call time includes the launch and one stream synchronisation",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- batched jamming
def _jam_prm(P, args, pressure=None):
    return P.jam_params(pressure_target=args.jam_pressure if pressure is None else pressure,
                        window_rel=args.jam_window, strain_step=args.jam_strain,
                        max_compression_steps=args.jam_csteps,
                        fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=args.jam_fmax,
                                           max_steps=args.jam_fire_steps))


def batch_force_roofline(P, sts, steps):
    """The batched force pass (k_b_force) timed with HIP events over `steps` batched
    steps of a batch holding `sts` from their initial state (every system is inside a
    minimisation then, so every particle slot is worked).  Algorithmic bytes per live slot:
    ownership 4 + position record 32 + row length 4 + 4 k (row) + velocity 16 read, force
    16 + energy 8 + velocity 16 written = 96 + 4 k (neighbour records not counted again)."""
    import numpy as np
    b = P.Batch([st.n for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    b.set_option(P.OPT_TIMING, 1)
    prm = P.jam_params(pressure_target=1e-2, window_rel=0.5, strain_step=1e-2,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-12, max_steps=10 * steps))
    b.jam(prm, steps)
    ms, launches = b.kernel_time()
    entries = b.row_entries()
    b.close()
    n = int(np.sum([st.n for st in sts]))
    k = entries / n
    per = ms / max(launches, 1) * 1e-3
    byts = n * (96.0 + 4.0 * k)
    gbs = byts / per / 1e9
    return {"bound": "hbm", "kernel": "k_b_force (batched force pass)", "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": byts,
            "kernel_us_per_launch": per * 1e6, "timed_launches": launches,
            "row_mean": k, "particles": n}


def batch_rigid_roofline(P, W, sts, steps):
    """The batched rigid force pass (k_rb_force) timed with HIP events over `steps` batched
    steps of a batch holding the rigid systems `sts` from their initial state (every
    system inside a minimisation).  Algorithmic bytes: per vertex 16 (position) + 4
    (owner) + 4 (row length) + 4 k_v (row; the neighbours' positions are the same array,
    not counted again); per particle 32 (record: lever arm and mass) + 8 (vertex range) +
    32 (force, torque, energy written) + 48 (velocity and spin read and written) + 8
    (moment of inertia)."""
    import numpy as np
    b = P.Batch([st.n for st in sts], [st.nv for st in sts])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        b.set_rigid_system(k, st.x, st.y, st.rad, st.mass, st.angle, st.nverts, st.rv,
                           st.side, e_c=W.bumpy_contact_scale(st))
    b.set_option(P.OPT_TIMING, 1)
    prm = P.jam_params(pressure_target=1e-2, window_rel=0.5, strain_step=1e-2,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-12, max_steps=10 * steps))
    b.jam(prm, steps)
    ms, launches = b.kernel_time()
    ventries = b.vertex_row_entries()
    b.close()
    n = int(np.sum([st.n for st in sts]))
    nv = int(np.sum([st.nv for st in sts]))
    kv = ventries / nv
    per = ms / max(launches, 1) * 1e-3
    byts = nv * (24.0 + 4.0 * kv) + n * 128.0
    gbs = byts / per / 1e9
    return {"bound": "hbm", "kernel": "k_rb_force (batched rigid force pass)", "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes_per_launch": byts,
            "kernel_us_per_launch": per * 1e6, "timed_launches": launches,
            "vertex_row_mean": kv, "particles": n, "vertices": nv}


def oracle_jam_sample(P, W, jobs, budget_s):
    """The CPU oracle jamming (system, parameters) jobs of the same workload one after
    another until the time budget is spent (the plain reference algorithm on the host)."""
    oracle_so = os.path.join(ROOT, "oracle", "liboracle.so")
    if not os.path.exists(oracle_so):
        return None
    be = P.Backend(oracle_so, "orc_", gpu=False)
    done, t = 0, 0.0
    for st, prm in jobs:
        if hasattr(st, "angle"):
            s = W.make_bumpy_system(P.System, st, backend=be, e_c=W.bumpy_contact_scale(st),
                                    method=P.NBR_VERLET)
        else:
            s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2),
                                   backend=be)
        c0 = time.perf_counter()
        s.jam(prm)
        t += time.perf_counter() - c0
        done += 1
        if t > budget_s:
            break
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": done / t, "unit": "packings/s", "cores": threads, "kind": "port",
            "sample": f"{done} of the systems jammed one after another by the C++ oracle "
                      f"(OpenMP; systems under 2048 particles run on one thread), {t:.1f} s"}


def jam_batch(args, P, W):
    """Jamming of many independent disk systems of uneven N: the batched engine against
    the plain one-system-at-a-time GPU path (dpm_jam, Verlet rows: all-pairs under 1,024
    particles, from a cell grid above; the same result bits), both on the same systems.
    value = systems jammed (status 0) per second by the batched engine."""
    import numpy as np
    rigid = args.batch_kind == "rigid"
    if rigid:  # rigid bumpy systems (20/28 vertices), the rigid neighbour rules
        ns0 = W.batched_sizes(args.systems, args.seed, args.batch_n_lo, args.batch_n_hi)
        sts = [W.bumpy_config(int(n), args.jam_phi0, seed=args.seed + 1 + k)
               for k, n in enumerate(ns0)]
    else:
        sts = W.batched_systems(args.systems, args.seed, args.batch_n_lo, args.batch_n_hi,
                                args.jam_phi0)
    ns = np.array([st.n for st in sts])
    prm = _jam_prm(P, args)
    rc, thr_sq = 1.7, 3.0 * 0.15 * 0.15
    print(f"[bench] batched jamming: {len(sts)} {args.batch_kind} systems, {int(ns.sum())} "
          f"particles", file=sys.stderr, flush=True)
    b = P.Batch([st.n for st in sts], [st.nv for st in sts] if rigid else None)
    b.config(1.0, 2.0, rc, thr_sq)
    b.set_option(P.OPT_CHUNK, args.batch_chunk)
    b.set_option(P.OPT_GRAPH, 0 if args.no_graph else 1)
    for k, st in enumerate(sts):
        if rigid:
            b.set_rigid_system(k, st.x, st.y, st.rad, st.mass, st.angle, st.nverts, st.rv,
                               st.side, e_c=W.bumpy_contact_scale(st))
        else:
            b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    tuned = None
    if args.autotune:
        cut, chunk, ms = b.autotune(prm, trial_steps=args.autotune_steps)
        blk, pct, ms = b.autotune_shape(prm, trial_steps=args.autotune_steps)
        tuned = {"cutoff": cut, "steps_per_launch": chunk, "force_block": blk,
                 "cell_pct": pct, "trial_ms_per_step": ms}
        print(f"[bench] autotune: cutoff {cut}, {chunk} steps per launch, force block {blk},"
              f" cells {pct}% of the cutoff", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    res, steps = b.jam(prm, args.batch_max_steps)
    elapsed = time.perf_counter() - t0
    fin = b.finish_steps()
    b.close()
    status = np.array([r.status for r in res])
    jammed = int(np.sum(status == 0))
    step_s = elapsed / max(steps, 1)
    lat = np.sort(fin[fin > 0] * step_s)
    p50 = float(np.percentile(lat, 50)) if len(lat) else None
    p99 = float(np.percentile(lat, 99)) if len(lat) else None
    # plain path on a sample of the same systems, one at a time
    rng = np.random.default_rng(args.seed + 1)
    k_sample = min(args.plain_sample, len(sts))
    pick = rng.choice(len(sts), size=k_sample, replace=False)
    print(f"[bench] plain path on {k_sample} of the systems", file=sys.stderr, flush=True)
    t_plain, same = [], True
    for k in pick:
        st = sts[int(k)]
        if rigid:
            s = W.make_bumpy_system(P.System, st, e_c=W.bumpy_contact_scale(st),
                                    method=P.NBR_VERLET)
        else:
            s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2))
        s.set_option(P.OPT_GRAPH, 0 if args.no_graph else 1)
        s.synchronize()
        c0 = time.perf_counter()
        r = s.jam(prm)
        s.synchronize()
        t_plain.append(time.perf_counter() - c0)
        for f in ("status", "compression_steps", "fire_steps_total", "pressure", "pe",
                  "packing_fraction", "contacts", "box_length"):
            same = same and getattr(r, f) == getattr(res[int(k)], f)
        s.close()
    plain_jams_s = 1.0 / (sum(t_plain) / len(t_plain))
    value = jammed / elapsed
    print("[bench] force-pass roofline (event-timed batched steps)", file=sys.stderr,
          flush=True)
    if rigid:
        roof = batch_rigid_roofline(P, W, sts, args.roofline_steps)
    else:
        roof = batch_force_roofline(P, sts, args.roofline_steps)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = oracle_jam_sample(P, W, [(sts[int(k)], prm) for k in pick], args.cpu_budget)
    out = {
        "metric": "batched jammed packings/s vs the plain one-system-at-a-time GPU path",
        "value": value,
        "unit": "packings/s",
        "n_gpus": 1,
        "steps": int(steps),
        "warmup": 0,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded uniform positions; N log-uniform in [{args.batch_n_lo}, "
                f"{args.batch_n_hi}])",
        "config": {
            "workload": f"{len(sts)} bidisperse "
                        f"{'rigid bumpy (20/28 vertices)' if rigid else 'disk'} systems "
                        f"(1.0/1.4, 50/50), N log-uniform "
                        f"in [{args.batch_n_lo}, {args.batch_n_hi}], phi0={args.jam_phi0}, "
                        f"compress by strain {args.jam_strain} with FIRE to |F|max<"
                        f"{args.jam_fmax} (at most {args.jam_fire_steps} steps each), pressure "
                        f"bisection to {args.jam_pressure} (window {args.jam_window}), Verlet "
                        f"rc={'1.5 d_max' if rigid else '1.7'}",
            "systems": len(sts),
            "particles": int(ns.sum()),
            "jammed_status0": jammed,
            "status_counts": {str(int(k)): int(v) for k, v in
                              zip(*np.unique(status, return_counts=True))},
            "elapsed_s": elapsed,
            "latency_p50_s": p50,
            "latency_p99_s": p99,
            "plain_jams_per_s": plain_jams_s,
            "plain_sample": k_sample,
            "plain_mean_s": sum(t_plain) / len(t_plain),
            "speedup_vs_plain": value / plain_jams_s,
            "sample_results_bit_identical": same,
            "autotune": tuned,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


# ---------------------------------------------------------------- serving loop
def serve(args, P, W):
    """Poisson arrivals of jamming requests served by continuous batching (dpm_server)
    against the plain sequential GPU path (one dpm_jam at a time, first come first served)
    on the same arrival sequence.  value = requests jammed (status 0) per second of the
    server over the run."""
    import numpy as np
    reqs = W.serve_requests(args.requests, args.rate, args.seed, args.bumpy_fraction,
                            (args.disk_n_lo, args.disk_n_hi), (args.bumpy_n_lo, args.bumpy_n_hi),
                            (args.p_lo, args.p_hi), args.jam_phi0)
    n_bumpy = sum(1 for r in reqs if r.bumpy)
    print(f"[bench] serving {len(reqs)} requests ({n_bumpy} rigid bumpy) at rate {args.rate}/s",
          file=sys.stderr, flush=True)

    def submit(sv, q):
        st, prm = q.state, _jam_prm(P, args, q.pressure)
        if q.bumpy:
            return sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm,
                             e_c=W.bumpy_contact_scale(st), angle=st.angle,
                             vertices_per_particle=st.nverts, vertex_radius=st.rv)
        return sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm)

    sv = P.Server(max_systems=args.serve_systems, arena_slots=args.serve_slots,
                  budget_bytes=args.serve_budget, vertex_slots=args.serve_vertex_slots)
    ids = [None] * len(reqs)
    t0 = time.perf_counter()
    nxt = 0
    done = 0
    next_note = 10.0
    while done < len(reqs):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            ids[nxt] = submit(sv, reqs[nxt])
            nxt += 1
        done, inflight, queued = sv.progress(args.batch_chunk)
        if now >= next_note:  # a progress line every ten seconds
            print(f"[bench] {now:.0f} s: {done} done, {inflight} running, {queued} queued",
                  file=sys.stderr, flush=True)
            next_note += 10.0
        if inflight == 0 and queued == 0 and nxt < len(reqs):
            time.sleep(max(0.0, min(0.01, reqs[nxt].arrival - (time.perf_counter() - t0))))
    elapsed = time.perf_counter() - t0
    lat = np.array([sv.result(i)["latency_s"] for i in ids])
    served = [sv.result(i)["result"] for i in ids]
    for i in ids:
        sv.release(i)
    sv.close()
    status = np.array([r.status for r in served])
    # the plain path: the same arrivals, one request at a time
    print("[bench] plain sequential path", file=sys.stderr, flush=True)
    k_plain = min(args.plain_sample, len(reqs))
    free_at, lat_plain, busy, same = 0.0, [], 0.0, True
    for k in range(k_plain):
        q = reqs[k]
        st, prm = q.state, _jam_prm(P, args, q.pressure)
        print(f"[bench] plain request {k}: {'rigid' if q.bumpy else 'disk'} N={st.n}",
              file=sys.stderr, flush=True)
        if q.bumpy:
            s = W.make_bumpy_system(P.System, st, e_c=W.bumpy_contact_scale(st),
                                    method=P.NBR_VERLET)
        else:
            s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2))
        s.synchronize()
        c0 = time.perf_counter()
        r = s.jam(prm)
        s.synchronize()
        dt = time.perf_counter() - c0
        s.close()
        start = max(q.arrival, free_at)
        free_at = start + dt
        busy += dt
        lat_plain.append(free_at - q.arrival)
        for f in ("status", "compression_steps", "fire_steps_total", "pressure", "pe",
                  "packing_fraction", "contacts", "box_length"):
            same = same and getattr(r, f) == getattr(served[k], f)
    plain_rate = k_plain / busy
    value = int(np.sum(status == 0)) / elapsed
    lp = np.array(lat_plain)
    disks = [q.state for q in reqs if not q.bumpy]
    roof = batch_force_roofline(P, disks, args.roofline_steps) if disks else None
    cpu = None
    if not args.no_cpu_baseline and disks:
        # disk requests only: one large rigid jam alone can take minutes on the host,
        # beyond the bounded sample
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = oracle_jam_sample(P, W, [(q.state, _jam_prm(P, args, q.pressure))
                                       for q in reqs if not q.bumpy], args.cpu_budget)
        if cpu:
            cpu["sample"] = "the disk requests in arrival order: " + cpu["sample"]
    out = {
        "metric": "served jammed packings/s and latency vs the plain sequential GPU path",
        "value": value,
        "unit": "packings/s",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 0,
        "ms_per_step": elapsed / len(reqs) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Poisson arrivals and configurations)",
        "config": {
            "workload": f"{len(reqs)} requests, Poisson rate {args.rate}/s: "
                        f"{100 * (1 - args.bumpy_fraction):.0f}% disks N log-uniform in "
                        f"[{args.disk_n_lo}, {args.disk_n_hi}], "
                        f"{100 * args.bumpy_fraction:.0f}% rigid bumpy (20/28 vertices) N "
                        f"log-uniform in [{args.bumpy_n_lo}, {args.bumpy_n_hi}]; target "
                        f"pressure log-uniform in [{args.p_lo}, {args.p_hi}]; strain "
                        f"{args.jam_strain}, FIRE to |F|max<{args.jam_fmax} (at most "
                        f"{args.jam_fire_steps} steps each)",
            "requests": len(reqs),
            "rigid_bumpy_requests": n_bumpy,
            "jammed_status0": int(np.sum(status == 0)),
            "status_counts": {str(int(k)): int(v) for k, v in
                              zip(*np.unique(status, return_counts=True))},
            "elapsed_s": elapsed,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "plain_packings_per_s": plain_rate,
            "plain_latency_p50_s": float(np.percentile(lp, 50)),
            "plain_latency_p99_s": float(np.percentile(lp, 99)),
            "plain_sample": k_plain,
            "speedup_vs_plain": value / plain_rate,
            "sample_results_bit_identical": same,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def traffic_from_profiles(n, name="k_nve_fused_n262144_traffic.json"):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        if int(d.get("n_particles", -1)) != n:
            return None
        return float(d["bytes_per_launch"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--cpu-steps", type=int, default=6000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--chunk", type=int, default=32)
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--jam-phi0", type=float, default=0.80)
    ap.add_argument("--jam-pressure", type=float, default=1e-7)
    ap.add_argument("--jam-window", type=float, default=0.05)
    ap.add_argument("--jam-strain", type=float, default=1e-4)
    ap.add_argument("--jam-fmax", type=float, default=1e-12)
    ap.add_argument("--jam-fire-steps", type=int, default=100000)
    ap.add_argument("--jam-csteps", type=int, default=100000)
    ap.add_argument("--batch-chunk", type=int, default=32)
    ap.add_argument("--batch-kind", choices=("disk", "rigid"), default="disk")
    ap.add_argument("--batch-n-lo", type=int, default=64)
    ap.add_argument("--batch-n-hi", type=int, default=2048)
    ap.add_argument("--batch-max-steps", type=int, default=10**9)
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--roofline-steps", type=int, default=100)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--autotune-steps", type=int, default=100)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--bumpy-fraction", type=float, default=0.3)
    ap.add_argument("--disk-n-lo", type=int, default=128)
    ap.add_argument("--disk-n-hi", type=int, default=16384)
    ap.add_argument("--bumpy-n-lo", type=int, default=128)
    ap.add_argument("--bumpy-n-hi", type=int, default=4096)
    ap.add_argument("--p-lo", type=float, default=1e-7)
    ap.add_argument("--p-hi", type=float, default=1e-3)
    ap.add_argument("--serve-systems", type=int, default=1024)
    ap.add_argument("--serve-slots", type=int, default=1 << 18)
    ap.add_argument("--serve-budget", type=int, default=16 << 30)
    ap.add_argument("--serve-vertex-slots", type=int, default=1 << 22)
    ap.add_argument("--workload", choices=("nve-disk", "fire-rigid", "jam-batch", "serve",
                                           "pairs", "nve-logged"),
                    default="nve-disk",
                    help="nve-disk: the headline NVE benchmark (default); fire-rigid: FIRE "
                         "of N=65,536 rigid bumpy particles to |F|max < 1e-12")
    ap.add_argument("--fire-max-steps", type=int, default=200000)
    ap.add_argument("--steps-probe", type=int, default=200)
    ap.add_argument("--cpu-fire-steps", type=int, default=300)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device copy (torch) five times; its "
                         "FETCH_SIZE/WRITE_SIZE counters calibrate the traffic figures")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        a = torch.empty(256 * 2**20 // 8, dtype=torch.float64, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 8, "launches": 5}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: measures one GPU only; runs on several GPUs are not supported "
              "(systems would shard by independent systems or by spatial slabs).",
              file=sys.stderr)
        sys.exit(2)

    P, W = load_pkg()
    if args.workload == "fire-rigid":
        rigid_fire(args, P, W)
        return
    if args.workload == "jam-batch":
        jam_batch(args, P, W)
        return
    if args.workload == "serve":
        serve(args, P, W)
        return
    if args.workload == "pairs":
        pairs(args, P, W)
        return
    if args.workload == "nve-logged":
        nve_logged(args, P, W)
        return
    n = args.n
    print(f"[bench] preparing N={n} (FIRE 200 steps, T=1e-5)", file=sys.stderr, flush=True)
    s = prepare(P, W, n, args.seed)
    s.set_option(P.OPT_GRAPH, 0 if args.no_graph else 1)
    s.set_option(P.OPT_CHUNK, args.chunk)
    # HIP events around every graph-launched chunk (on the system's stream): the fused
    # step kernel's time, chunk by chunk; set before the warm-up, which so also captures
    # the graphs and uses the events in the form the timed region runs
    s.set_option(P.OPT_TIMING, 2)
    print(f"[bench] warm-up {args.warmup} steps, timed region {args.steps} steps",
          file=sys.stderr, flush=True)
    if args.warmup:
        s.nve_run(0.01, args.warmup)
    s.synchronize()
    # nothing slow between the warm-up and the timed region (an idle GPU clocks down)
    r0 = s.neighbor_stats(totals=False)
    s.reset_kernel_time()
    t0 = time.perf_counter()
    s.nve_run(0.01, args.steps)
    s.synchronize()
    t1 = time.perf_counter()
    r1 = s.neighbor_stats()
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    f_ms, f_n = s.kernel_time(P.K_FORCE)
    r_ms, r_n = s.kernel_time(P.K_REBUILD)
    s.set_option(P.OPT_TIMING, 0)
    stats = s.neighbor_stats()
    k_mean = stats["total_entries"] / n
    step_bytes = n * (196.0 + 4.0 * k_mean)
    own_bytes = n * (132.0 + 4.0 * k_mean)
    chunk_steps = max(args.steps - 1, 1)  # nve_run(n): n - 1 steps run in chunks
    f_step_s = f_ms / chunk_steps / 1e3
    f_gbs = step_bytes / f_step_s / 1e9
    r_avg_us = r_ms / max(r_n, 1) * 1e3
    traffic = traffic_from_profiles(n)

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(P, W, n, args.seed, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform positions, FIRE-relaxed, Gaussian velocities)",
        "config": {
            "workload": "FP64 NVE, N=262144 bidisperse disks (1.0/1.4, 50/50), "
                        "phi=0.85, T=1e-5, dt=0.01, cell list rc=1.7, rebuild at max "
                        "displacement > 0.15",
            "n_particles": n,
            "k_mean": k_mean,
            "rebuilds_in_timed_region": r1["n_rebuilds"] - r0["n_rebuilds"],
            "cell_rebuilds_in_timed_region": r1["n_cell_rebuilds"] - r0["n_cell_rebuilds"],
            "graph": not args.no_graph,
            "chunk": args.chunk,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_nve_fused",
            "achieved": f_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": f_gbs / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": step_bytes,
            "kernel_compulsory_bytes_per_launch": own_bytes,
            "kernel_us_per_step": f_step_s * 1e6,
            "timed_chunks": f_n,
            # the same kernel time against the kernel's own compulsory bytes and against
            # the HBM bytes the counters measure per launch
            "frac_on_kernel_compulsory": own_bytes / f_step_s / 1e9 / HBM_PEAK_GBS,
            "frac_on_measured_traffic": (traffic / f_step_s / 1e9 / HBM_PEAK_GBS
                                         if traffic else None),
        },
        "cell_rebuild_recovery": {"mean_us": r_avg_us, "count": r_n},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
