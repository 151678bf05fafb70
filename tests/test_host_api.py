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
    k = 0
    for i in range(64):
        nv, r = (28, 0.7) if i < 32 else (20, 0.5)
        for j in range(nv):
            d = math.hypot(vx[k] - x[i], vy[k] - y[i])
            assert d == pytest.approx(r - rv, rel=1e-12)
            k += 1


def test_drivers_are_built():
    for d in ("run_nve", "run_fire", "run_jam", "run_jam_with_fire"):
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
    ["run_fire", "n_particles=512", "fire_num_steps=3000"],
    ["run_jam", "n_particles=64", "num_compression_steps=60", "compression_increment=2e-3"],
    ["run_jam_with_fire", "n_particles=64", "compression_rate=0.99",
     "pressure_tolerance_low=1e-4", "pressure_tolerance_high=1.5e-4"],
])
def test_drivers_run(argv):
    print("[driver]", " ".join(argv), flush=True)
    p = subprocess.run([os.path.join(DRIVERS, argv[0])] + argv[1:], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "Time taken:" in p.stdout
