"""The reference's own driver programs, unchanged, over this library's C++ host API.

The seven drivers of the hot path (scripts/create.cpp, runNVE.cpp, runFire.cpp, runAdam.cpp,
runCompress.cpp, runJam.cpp, runArthurJamWithFire.cpp) are copied byte for byte from the
reference checkout into a temporary tree, compiled with g++ against the headers of
dpcuda-v2_amd/host/include (which sit at the reference's include paths) and linked with the
host library built over the CPU oracle (tests/cpp/_build/libdpm_host_oracle.so), then run
on small run directories laid out as the reference's tooling lays them out.  The reference
checkout is only present where this suite is developed; elsewhere the test is skipped.
"""
import csv
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from _helpers import ROOT
from test_host_api import _disk_particle_config, _log_config

REF_SCRIPTS = "/root/reference/scripts"
DRIVERS = ("create", "runNVE", "runFire", "runAdam", "runCompress", "runJam",
           "runArthurJamWithFire")
HOST_INCLUDE = os.path.join(ROOT, "dpcuda-v2_amd", "host", "include")
HOST_ORACLE = os.path.join(ROOT, "tests", "cpp", "_build")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SCRIPTS),
                                reason="the reference checkout is not present")


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    if not os.path.exists(os.path.join(HOST_ORACLE, "libdpm_host_oracle.so")):
        subprocess.check_call(["make", "-C", os.path.join(ROOT, "tests", "cpp"),
                               "_build/libdpm_host_oracle.so"])
    d = tmp_path_factory.mktemp("drivers")
    scripts = d / "scripts"
    scripts.mkdir()
    out = {}
    for name in DRIVERS:
        src = os.path.join(REF_SCRIPTS, name + ".cpp")
        dst = scripts / (name + ".cpp")
        shutil.copyfile(src, dst)
        assert open(src, "rb").read() == open(dst, "rb").read()  # unchanged
        exe = str(scripts / name)
        p = subprocess.run(["g++", "-std=c++17", "-O1", "-I", HOST_INCLUDE, "-o", exe, str(dst),
                            "-L", HOST_ORACLE, "-ldpm_host_oracle", "-Wl,-rpath," + HOST_ORACLE],
                           capture_output=True, text=True)
        assert p.returncode == 0, (name, p.stderr[-3000:])
        out[name] = exe
    # the library's own NVE driver over the same host library, for comparison
    own = str(scripts / "run_nve_own")
    p = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "dpcuda-v2_amd", "host"),
                        "-o", own, os.path.join(ROOT, "dpcuda-v2_amd", "host", "drivers", "run_nve.cpp"),
                        "-L", HOST_ORACLE, "-ldpm_host_oracle", "-Wl,-rpath," + HOST_ORACLE],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    out["run_nve_own"] = own
    return out


def _prepare(out_dir, particle_config=None):
    import json
    sysd = os.path.join(out_dir, "system")
    os.makedirs(sysd, exist_ok=True)
    logs = {"console": _log_config(["step", "PE/N", "KE/N", "T"], "console", 100),
            "energy": _log_config(["step", "PE/N", "KE/N", "TE/N", "T", "phi", "P"], "energy", 50),
            "state": _log_config(["positions", "velocities"], "state", 250)}
    for g, c in logs.items():
        json.dump(c, open(os.path.join(sysd, f"{g}_log_config.json"), "w"), indent=4)
    if particle_config is not None:
        json.dump(particle_config, open(os.path.join(sysd, "particle_config.json"), "w"),
                  indent=4)


def _run(exe, run_dir, run_config, timeout=600):
    import json
    os.makedirs(os.path.join(run_dir, "system"), exist_ok=True)
    json.dump(run_config, open(os.path.join(run_dir, "system", "run_config.json"), "w"), indent=4)
    p = subprocess.run([exe, run_dir], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (exe, p.stdout[-2000:], p.stderr[-2000:])
    assert "Time taken:" in p.stdout or os.path.basename(exe) == "create"
    return p.stdout


def _energy(out_dir):
    rows = list(csv.DictReader(open(os.path.join(out_dir, "system", "energy.csv"))))
    return [{k: float(v) for k, v in r.items()} for r in rows]


def _dat(path):
    return np.loadtxt(path, ndmin=2)


def _bumpy_particle_config(n, phi, seed):
    rv = (1.0 / (1.0 + 1.0 / math.sin(math.pi / 20))) / 2.0
    return {"particle_type": "RigidBumpy", "n_particles": n, "n_vertices": n // 2 * 20 + n // 2 * 28,
            "packing_fraction": phi, "particle_dim_block": 256, "vertex_dim_block": 256,
            "segment_length_per_vertex_diameter": 1.0, "vertex_radius": rv,
            "n_vertices_per_small_particle": 20, "n_vertices_per_large_particle": 28,
            "size_ratio": 1.4, "count_ratio": 0.5, "rotation": True, "e_c": 1.0, "n_c": 2.0,
            "mass": 1.0, "seed": seed,
            "neighbor_list_config": {"neighbor_cutoff_multiplier": 1.5,
                                     "neighbor_displacement_multiplier": 0.2,
                                     "num_particles_per_cell": 8.0,
                                     "cell_displacement_multiplier": 0.5,
                                     "vertex_neighbor_cutoff_multiplier": 1.5,
                                     "vertex_neighbor_displacement_multiplier": 0.2,
                                     "neighbor_list_update_method": "cell"}}


def test_reference_drivers_build_unchanged_and_run(bins, tmp_path):
    base = str(tmp_path)
    # create: a new particle from the particle config, written as init and restart
    src = os.path.join(base, "src")
    _prepare(src, _disk_particle_config(256, 0.80, 7))
    _run(bins["create"], src, {"input_dir": "", "output_dir": src, "minimize": False})
    x0 = _dat(os.path.join(src, "system", "init", "positions.dat"))
    assert x0.shape == (256, 2)

    # runNVE over the reference's unchanged driver and over the library's own driver:
    # the same run directory, byte for byte; total energy conserved
    outs = {}
    for name in ("runNVE", "run_nve_own"):
        out = os.path.join(base, "nve_" + name)
        _prepare(out)
        _run(bins[name], os.path.join(base, "run_" + name),
             {"input_dir": src, "output_dir": out, "minimize": False, "num_steps": 500,
              "dt_dimless": 0.01, "temperature": 1e-4})
        outs[name] = out
    for rel in ("system/energy.csv", "system/restart/positions.dat",
                "system/restart/velocities.dat", "trajectories/t250/positions.dat"):
        a = open(os.path.join(outs["runNVE"], rel)).read()
        assert a == open(os.path.join(outs["run_nve_own"], rel)).read(), rel
    e = _energy(outs["runNVE"])
    assert [r["step"] for r in e] == list(range(0, 500, 50)) + [500]
    te = np.array([r["TE/N"] for r in e])
    assert np.max(np.abs(te - te[0])) < 1e-3 * abs(te[0])

    # runFire: minimised to the PE criteria; the pair arrays of the restart group written
    fire = os.path.join(base, "fire")
    _prepare(fire)
    _run(bins["runFire"], os.path.join(base, "run_fire"),
         {"input_dir": src, "output_dir": fire, "minimize": False, "fire_alpha_init": 0.1,
          "fire_dt": 0.01, "fire_num_steps": 20000, "fire_log_every_n": 1000})
    assert _energy(fire)[-1]["PE/N"] < 1e-12
    assert os.path.exists(os.path.join(fire, "system", "restart", "force_pairs.dat"))

    # runAdam: the energy falls below a tenth of the created state's (the driver logs only
    # its final state)
    pe0 = _energy(src)[0]["PE/N"]
    assert pe0 > 0
    adam = os.path.join(base, "adam")
    _prepare(adam)
    _run(bins["runAdam"], os.path.join(base, "run_adam"),
         {"input_dir": src, "output_dir": adam, "minimize": False, "adam_alpha": 1e-3,
          "adam_beta1": 0.9, "adam_beta2": 0.999, "adam_epsilon": 1e-8,
          "adam_num_steps": 2000, "adam_log_every_n": 1000})
    assert _energy(adam)[-1]["PE/N"] < 0.1 * pe0

    # runCompress: packing fraction raised by delta_phi, then minimised
    comp = os.path.join(base, "compress")
    _prepare(comp)
    _run(bins["runCompress"], os.path.join(base, "run_compress"),
         {"input_dir": src, "output_dir": comp, "minimize": False, "delta_phi": 0.02,
          "avg_pe_tolerance": 1e-16, "avg_pe_diff_tolerance": 1e-16})
    ec = _energy(comp)
    assert ec[-1]["phi"] == pytest.approx(ec[0]["phi"] + 0.02, rel=1e-12)

    # runJam (packing-fraction bisection) and runArthurJamWithFire (pressure bisection) on
    # 64 disks: the latter ends inside its pressure window
    small = os.path.join(base, "small")
    _prepare(small, _disk_particle_config(64, 0.80, 3))
    _run(bins["create"], small, {"input_dir": "", "output_dir": small, "minimize": False})
    jam = os.path.join(base, "jam")
    _prepare(jam)
    _run(bins["runJam"], os.path.join(base, "run_jam"),
         {"input_dir": small, "output_dir": jam, "minimize": False,
          "num_compression_steps": 40, "compression_increment": 2e-3,
          "avg_pe_tolerance": 1e-16, "avg_pe_diff_tolerance": 1e-16, "phi_tolerance": 1e-6,
          "delta_phi_stability": 1e-4})
    assert os.path.exists(os.path.join(jam, "system", "restart", "hessian_ii_yy.dat"))
    jf = os.path.join(base, "jam_fire")
    _prepare(jf)
    _run(bins["runArthurJamWithFire"], os.path.join(base, "run_jam_fire"),
         {"input_dir": small, "output_dir": jf, "minimize": False,
          "num_compression_steps": 300, "compression_rate": 0.99,
          "pressure_tolerance_low": 1e-4, "pressure_tolerance_high": 1.5e-4,
          "length_tolerance": 1e-12})
    last = _energy(jf)[-1]
    assert 1e-4 <= last["P"] <= 1.5e-4
    assert os.path.exists(os.path.join(jf, "system", "last_state", "positions.dat"))

    # a rigid bumpy particle through create and runFire (the RigidBumpy branches)
    rb = os.path.join(base, "rb")
    _prepare(rb, _bumpy_particle_config(16, 0.70, 5))
    _run(bins["create"], rb, {"input_dir": "", "output_dir": rb, "minimize": False})
    rbf = os.path.join(base, "rb_fire")
    _prepare(rbf)
    _run(bins["runFire"], os.path.join(base, "run_rb_fire"),
         {"input_dir": rb, "output_dir": rbf, "minimize": False, "fire_alpha_init": 0.1,
          "fire_dt": 0.01, "fire_num_steps": 5000, "fire_log_every_n": 1000})
    assert os.path.exists(os.path.join(rbf, "system", "restart", "angle_pairs_i.dat"))
