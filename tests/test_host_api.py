"""The C++ host API (dpcuda-v2_amd/host: Particle / Disk / RigidBumpy, NVE / Fire / Adam,
minimizeFire / minimizeAdam, jam_adam / compress_adam / decompress_adam /
compress_to_phi_adam, jamWithFire) run through fixed scenarios by
tests/cpp/host_scenarios.cpp.

The scenario program is built twice: over libdpm.so and over the CPU oracle (the same
host sources with the dpm_ names mapped onto orc_).  CPU tests run the oracle build and
check the host-side logic (seeded fills against std::minstd_rand, ConfigDict semantics,
error behaviour, neighbour/cell set-up and rigid-bumpy geometry values); the GPU test runs
both builds and compares every printed value."""
import math
import os
import subprocess

import pytest

from _helpers import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
BIN_ORACLE = os.path.join(CPP, "_build", "host_scenarios_oracle")
BIN_GPU = os.path.join(CPP, "_build", "host_scenarios_gpu")
DRIVERS = os.path.join(ROOT, "dpcuda-v2_amd", "_build", "drivers")
SCENARIOS = ["rng", "config", "errors", "disk_nve", "disk_fire", "disk_adam", "disk_compress",
             "disk_jam_fire", "bumpy_init"]


def _ensure_built(path):
    if not os.path.exists(path):
        subprocess.check_call(["make", "-C", CPP], stdout=subprocess.DEVNULL)
    assert os.path.exists(path), path


def run(binary, scenario, timeout=300):
    _ensure_built(binary)
    print(f"[host_scenarios] {os.path.basename(binary)} {scenario}", flush=True)
    p = subprocess.run([binary, scenario], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (scenario, p.returncode, p.stderr[-2000:])
    out = {}
    for line in p.stdout.splitlines():
        if line.startswith("@ "):
            key, *vals = line[2:].split()
            out[key] = vals
    assert out.get("done") == ["1"], scenario
    return out


def num(tok):
    return float.fromhex(tok) if ("p" in tok or "x" in tok) else float(int(tok))


_cache = {}


def oracle_run(scenario):
    if scenario not in _cache:
        _cache[scenario] = run(BIN_ORACLE, scenario)
    return _cache[scenario]


# ------------------------------------------------------------------ CPU: host-side logic
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_scenarios_run_over_oracle(scenario):
    out = oracle_run(scenario)
    assert len(out) > 1


def test_seeded_fill_is_thrust_minstd():
    # random_uniform(seed, n) == minstd_rand(seed + n), discard(n), one draw mapped to [a, b)
    assert oracle_run("rng")["rng.mismatches"] == ["0"]


def test_config_dict_semantics():
    out = oracle_run("config")
    assert num(out["config.copy_is_deep"][0]) == 8.0
    assert num(out["config.implicit_phi"][0]) == 0.8
    assert out["config.implicit_n"] == ["1024"] and out["config.kind_is_disk"] == ["1"]
    assert out["config.rb_n_vertices"] == [str(5 * 26 + 5 * 36)]
    assert out["throws.missing_key"] == ["1"] and out["throws.wrong_type"] == ["1"]


def test_error_behaviour():
    out = oracle_run("errors")
    thrown = {k: v for k, v in out.items() if k.startswith("throws.")}
    assert len(thrown) == 9 and all(v == ["1"] for v in thrown.values()), thrown


def test_disk_neighbour_setup_values():
    # N = 256 at phi = 0.85, diameters 1.0 / 1.4 half and half: box side from the disk area,
    # cells of about 8 particles tiling the box (at least 2 d_max = 2.8 wide)
    out = oracle_run("disk_nve")
    side = math.sqrt(math.pi * (128 * 0.25 + 128 * 0.49) / 0.85)
    assert num(out["nve.init.box"][0]) == pytest.approx(side, rel=1e-15)
    n_dim = int(out["nve.n_cells_dim"][0])
    assert n_dim == math.floor(side / math.sqrt(8.0 / (256 / side ** 2))) == 5
    cs = num(out["nve.cell_size"][0])
    assert cs == pytest.approx(side / n_dim, rel=1e-15) and cs >= 2.8
    assert num(out["nve.neighbor_cutoff"][0]) == pytest.approx(2.1, rel=1e-15)
    assert num(out["nve.disp_sq"][0]) == pytest.approx(0.42 ** 2, rel=1e-14)
    assert num(out["nve.cell_disp_sq"][0]) == pytest.approx((0.5 * cs) ** 2, rel=1e-14)
    assert num(out["nve.time_unit"][0]) == 1.0
    # T = 1e-4 after setRandomVelocities, and NVE conserves energy over 200 steps
    assert num(out["nve.temperature"][0]) == pytest.approx(1e-4, rel=1e-12)
    e60 = num(out["nve.step60.pe"][0]) + num(out["nve.step60.ke"][0])
    e200 = num(out["nve.run200.pe"][0]) + num(out["nve.run200.ke"][0])
    assert abs(e200 / e60 - 1) < 1e-4


def test_bumpy_geometry_values():
    # 20 vertices on the small ring, segment = one vertex diameter: r_v solves
    # 2 (R - r_v) sin(pi/20) = 2 r_v with R = 1/2; e_c is scaled by (2 r_v / d_min)^2
    out = oracle_run("bumpy_init")
    rv = num(out["bumpy.vertex_radius"][0])
    assert 2 * (0.5 - rv) * math.sin(math.pi / 20) == pytest.approx(2 * rv, rel=1e-14)
    assert num(out["bumpy.e_c"][0]) == pytest.approx((2 * rv) ** 2, rel=1e-14)
    assert out["bumpy.e_c"] == out["bumpy.config_e_c"]
    assert out["bumpy.n_vertices"] == [str(32 * 20 + 32 * 28)]
    assert out["bumpy.n_dof"] == [str(3 * 64)]
    assert num(out["bumpy.vertex_cutoff"][0]) == pytest.approx(1.5 * 2 * rv, rel=1e-15)
    assert num(out["bumpy.disp_sq"][0]) == pytest.approx((0.2 * rv) ** 2, rel=1e-14)
    # both seeded fills start at value seed + i: angle_i = 2 pi x_i / L_x
    ang = [num(t) for t in out["bumpy.angles"]]
    x = [num(t) for t in out["bumpy.init.x"]]
    side = num(out["bumpy.init.box"][0])
    for a, xi in zip(ang, x):
        assert a == pytest.approx(2 * math.pi * xi / side, rel=1e-12)
    # vertex rings sit at radius r - r_v around each centre
    vx = [num(t) for t in out["bumpy.vx"]]
    vy = [num(t) for t in out["bumpy.vy"]]
    y = [num(t) for t in out["bumpy.init.y"]]
    def rings_at(vx, vy, x, y):
        k = 0
        for i in range(64):
            nv, r = (28, 0.7) if i < 32 else (20, 0.5)
            for j in range(nv):
                d = math.hypot(vx[k] - x[i], vy[k] - y[i])
                assert d == pytest.approx(r - rv, rel=1e-12)
                k += 1

    rings_at(vx, vy, x, y)
    # rigid Adam (Adam over RigidBumpy after a compression): the rings move rigidly with
    # their centres, and the angles change
    rings_at([num(t) for t in out["bumpy.vx_adam"]], [num(t) for t in out["bumpy.vy_adam"]],
             [num(t) for t in out["bumpy.adam200.x"]], [num(t) for t in out["bumpy.adam200.y"]])
    assert out["bumpy.angles_adam"] != out["bumpy.angles100"]


def test_drivers_are_built():
    for d in ("run_nve", "create"):
        assert os.access(os.path.join(DRIVERS, d), os.X_OK), d


# ------------------------------------------------------------------ GPU: same values
EXACT = set(SCENARIOS)


@pytest.mark.gpu
@pytest.mark.parametrize("scenario", SCENARIOS)
def test_host_api_over_hip_matches_oracle(scenario):
    """Every printed value bit-identical (rigid bumpy included: both sides rotate vertices
    with the same sin/cos)."""
    o = oracle_run(scenario)
    g = run(BIN_GPU, scenario)
    assert set(g) == set(o), set(g) ^ set(o)
    for key in o:
        a, b = g[key], o[key]
        assert len(a) == len(b), key
        if scenario in EXACT:
            assert a == b, (key, [(x, y) for x, y in zip(a, b) if x != y][:4])
        else:
            av, bv = [num(t) for t in a], [num(t) for t in b]
            scale = max([abs(v) for v in bv] + [1e-300])
            for x, y in zip(av, bv):
                assert abs(x - y) <= 1e-9 * max(abs(y), 1e-3 * scale), (key, x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("argv", [
    ["run_nve", "n_particles=1024", "num_steps=300", "temperature=1e-4", "loop=device"],
    ["run_nve", "n_particles=512", "num_steps=100", "temperature=1e-4", "loop=host"],
])
def test_drivers_run(argv):
    print("[driver]", " ".join(argv), flush=True)
    p = subprocess.run([os.path.join(DRIVERS, argv[0])] + argv[1:], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "Time taken:" in p.stdout


# ------------------------------------------------------------------ reference-form drivers
def _log_config(names, group, freq=1, style="lin"):
    return {"log_names": names, "save_style": style, "save_freq": freq, "reset_save_decade": 10,
            "min_save_decade": 10, "multiple": 0, "group_name": group, "decade": 10}


def _prepare(out_dir, run_config, particle_config=None):
    """A run directory as the upstream tooling lays it out: system/run_config.json, and in
    the output directory the log configs (and a particle config for a new particle)."""
    import json
    sysd = os.path.join(out_dir, "system")
    os.makedirs(sysd, exist_ok=True)
    logs = {"console": _log_config(["step", "PE/N", "KE/N", "T"], "console", 100),
            "energy": _log_config(["step", "PE/N", "KE/N", "TE/N", "T", "phi", "P"], "energy", 50),
            "state": _log_config(["positions", "velocities"], "state", 250)}
    for g, c in logs.items():
        json.dump(c, open(os.path.join(sysd, f"{g}_log_config.json"), "w"), indent=4)
    if particle_config is not None:
        json.dump(particle_config, open(os.path.join(sysd, "particle_config.json"), "w"), indent=4)
    return sysd


def _disk_particle_config(n, phi, seed):
    return {"particle_type": "Disk", "n_particles": n, "packing_fraction": phi,
            "particle_dim_block": 256, "e_c": 1.0, "n_c": 2.0, "mass": 1.0, "size_ratio": 1.4,
            "count_ratio": 0.5, "seed": seed,
            "neighbor_list_config": {"neighbor_cutoff_multiplier": 1.5,
                                     "neighbor_displacement_multiplier": 0.2,
                                     "num_particles_per_cell": 8.0,
                                     "cell_displacement_multiplier": 0.5,
                                     "neighbor_list_update_method": "cell"}}


def _drive(name, run_dir, run_config, timeout=300):
    import json
    os.makedirs(os.path.join(run_dir, "system"), exist_ok=True)
    json.dump(run_config, open(os.path.join(run_dir, "system", "run_config.json"), "w"), indent=4)
    print("[driver]", name, run_dir, flush=True)
    p = subprocess.run([os.path.join(DRIVERS, name), run_dir], capture_output=True, text=True,
                       timeout=timeout)
    assert p.returncode == 0, (name, p.stdout[-2000:], p.stderr[-2000:])
    return p.stdout


@pytest.mark.gpu
def test_reference_form_drivers(tmp_path):
    """create -> run_nve over the HIP library, each started as `driver <run directory>` and
    writing the reference's directory layout; the NVE steps between two logs run as one
    device loop or one host call per step, with the same files.  (The reference's own
    drivers, unchanged, run over the host library in tests/test_reference_drivers.py.)"""
    base = str(tmp_path)
    src = os.path.join(base, "src")
    _prepare(src, None, _disk_particle_config(256, 0.8, 7))
    _drive("create", src, {"input_dir": "", "output_dir": src, "minimize": False})
    for sub in ("init", "restart"):
        assert os.path.exists(os.path.join(src, "system", sub, "positions.dat"))
    outs = {}
    for loop in ("device", "host"):
        out = os.path.join(base, f"nve_{loop}")
        _prepare(out, None)
        _drive("run_nve", os.path.join(base, f"run_nve_{loop}"),
               {"input_dir": src, "output_dir": out, "minimize": False, "num_steps": 600,
                "dt_dimless": 0.01, "temperature": 1e-4, "loop": loop})
        outs[loop] = out
    for rel in ("system/energy.csv", "system/restart/positions.dat",
                "system/restart/velocities.dat", "trajectories/t500/positions.dat"):
        a = open(os.path.join(outs["device"], rel)).read()
        assert a == open(os.path.join(outs["host"], rel)).read(), rel
    rows = open(os.path.join(outs["device"], "system", "energy.csv")).read().split("\n")[1:-1]
    assert [float(r.split(",")[0]) for r in rows] == list(range(0, 600, 50)) + [600]
    assert open(os.path.join(outs["device"], "system", "restart", "step.dat")).read() == "600"
