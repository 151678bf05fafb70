"""On-disk formats and the logging layer (dpcuda-v2_amd/host/io.*, state_io.cpp): ConfigDict
JSON text, the run-directory layout written by IOManager, the .dat / energy.csv formats,
the log schedules, loadParticle from init / restart / frame / a named state, load_configs,
and the per-pair analytics arrays written into restart directories.

The scenario program tests/cpp/host_scenarios.cpp writes a run through the C++ host layer.
CPU tests run its oracle build and check the files against what the formats, the
schedules and the pair law fix independently: Python's json module, "%.17e" round trips,
the closed-form log schedule, and the textbook force / Hessian of the harmonic contact
U(r) = e_c/2 (1 - r/s)^2.  The GPU test runs the libdpm.so build and compares every file
with the oracle build's, byte for byte (separation angles and particle areas, which use
the device's atan2 / acos, to 2 ulp)."""
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from _helpers import ROOT

CPP = os.path.join(ROOT, "tests", "cpp")
BIN_ORACLE = os.path.join(CPP, "_build", "host_scenarios_oracle")
BIN_GPU = os.path.join(CPP, "_build", "host_scenarios_gpu")
DOUBLE = re.compile(r"^-?\d\.\d{17}e[+-]\d{2,3}$")
INT = re.compile(r"^-?\d+$")


def _built(path):
    if not os.path.exists(path):
        subprocess.check_call(["make", "-C", CPP], stdout=subprocess.DEVNULL)
    assert os.path.exists(path), path
    return path


def run(binary, *args, timeout=300):
    print(f"[host_scenarios] {os.path.basename(binary)} {' '.join(args)}", flush=True)
    p = subprocess.run([_built(binary), *args], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, (args, p.returncode, p.stderr[-2000:], p.stdout[-2000:])
    out = {}
    for line in p.stdout.splitlines():
        if line.startswith("@ "):
            key, _, rest = line[2:].partition(" ")
            out[key] = rest
    assert out.get("done") == "1"
    return out, p.stdout


def hexnums(s):
    return [float.fromhex(t) for t in s.split()]


# ------------------------------------------------------------------ JSON text
def test_json_numbers_follow_nlohmann_layout():
    out, _ = run(BIN_ORACLE, "json")
    expect = {"one": "1.0", "tenth": "0.1", "third": "0.3333333333333333", "tiny": "1e-07",
              "small": "0.0001", "smaller": "1e-05", "e14": "100000000000000.0",
              "e15": "1e+15", "e16": "1e+16", "big": "1.5e+300", "neg": "-2.5e-08",
              "zero": "0.0", "negzero": "-0.0", "int": "42", "negint": "-7", "t": "true",
              "f": "false", "vec": '["x","y/N"]', "empty_vec": "[]"}
    for k, v in expect.items():
        assert out["json." + k] == v, (k, out["json." + k])
    # every float reads back to the same double and is no longer than Python's shortest repr
    for k in ("one", "tenth", "third", "tiny", "small", "smaller", "e14", "neg", "big"):
        txt = out["json." + k]
        val = json.loads(txt)
        assert repr(val) == repr(float(txt))
        assert len(txt.replace("e+", "e").replace("e-0", "e-")) <= len(repr(val)) + 2
    assert json.loads(out["json.str"]) == 'a"b\\c\n\t\x01'
    assert out["json.str"] == json.dumps('a"b\\c\n\t\x01')


def test_json_indented_dump_matches_python():
    out, text = run(BIN_ORACLE, "json")
    body = text.split("=== dump4\n", 1)[1].split("\n=== end", 1)[0]
    obj = json.loads(body)
    assert obj["neighbor_list_config"]["num_particles_per_cell"] == 8.0
    assert obj["log_names"] == ["step", "PE/N"] and obj["nested"]["empty"] == {}
    # same text as Python's sorted, 4-space dump (no float here lies in [1e15, 1e16), the
    # one range where Python's repr and nlohmann's layout differ)
    assert body == json.dumps(obj, indent=4, sort_keys=True, ensure_ascii=False)
    assert out["json.roundtrip"] == "1"
    assert json.loads(out["json.parsed"]) == {"a": [1, 2500.0, -0.0, "é\U0001F600"], "b": None}
    for k in ("json_trailing", "json_unterminated", "json_bad_escape"):
        assert out["throws." + k] == "1"


# ------------------------------------------------------------------ run directories
def read_dat(path, cols=1, as_int=False):
    rows = open(path).read().split("\n")
    assert rows[-1] == "", path  # every row ends in a newline
    rows = rows[:-1]
    vals = []
    for r in rows:
        parts = r.split("\t")
        assert len(parts) == cols, (path, r)
        for t in parts:
            assert (INT if as_int else DOUBLE).match(t), (path, t)
        vals.append([int(t) if as_int else float(t) for t in parts])
    return np.array(vals, dtype=np.int64 if as_int else np.float64)


def log_schedule(num_steps, reset, lo, decade=10):
    """Saves of a "log" group: within each reset period of `reset` steps, every `lo` steps
    up to lo*decade, then every lo*decade up to lo*decade^2, ... (and step 0)."""
    out = [0]
    for s in range(1, num_steps):
        r = s - ((s - 1) // reset) * reset  # position in its period, 1..reset
        f = lo
        while r > f * decade:
            f *= decade
        if r % f == 0:
            out.append(s)
    return out


def _io_run(tmp_path_factory, kind, binary=BIN_ORACLE, tag="oracle"):
    d = tmp_path_factory.mktemp(f"io_{kind}_{tag}")
    out, _ = run(binary, f"io_{kind}", str(d))
    return out, str(d / "out"), str(d)


@pytest.fixture(scope="module")
def disk_run(tmp_path_factory):
    return _io_run(tmp_path_factory, "disk")


@pytest.fixture(scope="module")
def bumpy_run(tmp_path_factory):
    return _io_run(tmp_path_factory, "bumpy")


@pytest.mark.parametrize("kind", ["disk", "bumpy"])
def test_run_directory_layout(kind, disk_run, bumpy_run):
    out, root, _ = disk_run if kind == "disk" else bumpy_run
    sysd = os.path.join(root, "system")
    for name in ("energy.csv", "particle_config.json", "integrator_config.json",
                 "console_log_config.json", "energy_log_config.json", "state_log_config.json",
                 "restart_log_config.json", "init", "restart"):
        assert os.path.exists(os.path.join(sysd, name)), name
    # state: "log" schedule over 120 steps with one save decade (reset 120, from 10), plus
    # the forced final log at step 120
    frames = sorted(int(f[1:]) for f in os.listdir(os.path.join(root, "trajectories")))
    assert frames == log_schedule(120, 120, 10) + [120]
    assert frames == [0, 10, 20, 30, 40, 50, 60, 70, 80, 90, 100, 120]
    assert open(os.path.join(sysd, "restart", "step.dat")).read() == "120"
    assert open(os.path.join(sysd, "init", "step.dat")).read() == "0"
    pc = json.load(open(os.path.join(sysd, "particle_config.json")))
    assert pc["particle_type"] == ("Disk" if kind == "disk" else "RigidBumpy")
    assert pc["seed"] == 17
    ic = json.load(open(os.path.join(sysd, "integrator_config.json")))
    assert ic["integrator_type"] == "NVE" and ic["dt"] > 0
    rc = json.load(open(os.path.join(sysd, "restart_log_config.json")))
    assert rc["save_style"] == "lin" and rc["save_freq"] == 50 and "force_pairs" in rc["log_names"]
    for f in os.listdir(sysd):
        if f.endswith(".json"):
            txt = open(os.path.join(sysd, f)).read()
            assert txt == json.dumps(json.loads(txt), indent=4, sort_keys=True, ensure_ascii=False)


def test_log_schedule_closed_form():
    # reset 1000, from 10: 10..100 by 10, 200..1000 by 100, then again from 1010
    s = log_schedule(2101, 1000, 10)
    assert s[:20] == [0] + list(range(10, 101, 10)) + list(range(200, 1001, 100))
    assert s[19:30] == [1000] + list(range(1010, 1101, 10))


@pytest.mark.parametrize("kind", ["disk", "bumpy"])
def test_state_files_format_and_values(kind, disk_run, bumpy_run):
    out, root, _ = disk_run if kind == "disk" else bumpy_run
    n = 96 if kind == "disk" else 48
    t = os.path.join(root, "trajectories", "t120")
    pos = read_dat(os.path.join(t, "positions.dat"), 2)
    assert pos.shape == (n, 2)
    # the files hold exactly the final state the program printed (17 digits round-trip)
    assert np.array_equal(pos[:, 0], hexnums(out["io.final.x"]))
    assert np.array_equal(pos[:, 1], hexnums(out["io.final.y"]))
    vel = read_dat(os.path.join(t, "velocities.dat"), 2)
    assert np.array_equal(vel[:, 0], hexnums(out["io.final.vx"]))
    assert np.array_equal(read_dat(os.path.join(t, "box_size.dat"))[:, 0], hexnums(out["io.final.box"]))
    assert np.array_equal(read_dat(os.path.join(t, "static_particle_index.dat"), as_int=True)[:, 0],
                          np.arange(n))
    ke = read_dat(os.path.join(t, "kinetic_energy.dat"))[:, 0]
    assert ke.shape == (n,) and np.all(ke >= 0)
    assert read_dat(os.path.join(t, "stress_tensor.dat")).shape == (4, 1)
    cc = read_dat(os.path.join(t, "contact_counts.dat"), as_int=True)
    assert cc.shape == (n, 1) and cc.min() >= 0
    # restart equals the final frame for the arrays both hold
    r = os.path.join(root, "system", "restart")
    assert open(os.path.join(r, "positions.dat")).read() == open(os.path.join(t, "positions.dat")).read()
    if kind == "bumpy":
        nvp = read_dat(os.path.join(t, "num_vertices_in_particle.dat"), as_int=True)[:, 0]
        assert sorted(set(nvp)) == [12, 16]
        vp = read_dat(os.path.join(t, "vertex_positions.dat"), 2)
        assert vp.shape == (nvp.sum(), 2)
        vm = read_dat(os.path.join(t, "vertex_masses.dat"))[:, 0]
        m = read_dat(os.path.join(t, "radii.dat"))  # shape check of a per-particle file
        assert m.shape == (n, 1) and np.allclose(vm[:12 if nvp[0] == 12 else 16].sum(), 1.0)


@pytest.mark.parametrize("kind", ["disk", "bumpy"])
def test_energy_csv(kind, disk_run, bumpy_run):
    out, root, _ = disk_run if kind == "disk" else bumpy_run
    n = 96 if kind == "disk" else 48
    lines = open(os.path.join(root, "system", "energy.csv")).read().split("\n")
    assert lines[-1] == ""
    head = lines[0].split(",")
    assert head == ["step", "PE/N", "KE/N", "TE/N", "T", "phi", "P", "Zp", "PE/dof", "KE"]
    rows = [r.split(",") for r in lines[1:-1]]
    for r in rows:
        assert len(r) == len(head) and all(DOUBLE.match(x) for x in r), r
    v = np.array([[float(x) for x in r] for r in rows])
    # "lin" every 120/12 = 10 steps from 0 to 110, then the forced final log at 120
    assert list(v[:, 0]) == list(range(0, 120, 10)) + [120]
    pe_n, ke_n, te_n = v[:, 1], v[:, 2], v[:, 3]
    assert np.allclose(te_n, pe_n + ke_n, rtol=1e-14, atol=0)
    assert np.allclose(v[:, 9], ke_n * n, rtol=1e-14, atol=0)
    dof = 2 * n if kind == "disk" else 3 * n
    assert np.allclose(v[:, 8], pe_n * n / dof, rtol=1e-14, atol=0)
    assert np.all(v[:, 5] == v[0, 5])  # NVE keeps the box
    final_pe = float.fromhex(out["io.final.pe"].split()[0])
    assert v[-1, 1] == pytest.approx(final_pe / n, rel=1e-15)
    # energy drift over 120 steps is small
    assert abs(te_n[-1] / te_n[0] - 1) < 1e-3


@pytest.mark.parametrize("kind", ["disk", "bumpy"])
def test_loaded_particles_equal_written(kind, disk_run, bumpy_run):
    out, _, _ = disk_run if kind == "disk" else bumpy_run
    assert out["io.restart_step"] == "120" and out["io.init_step"] == "0"
    assert out["io.frame_step"] == "120" and out["io.resume_step"] == "120"
    for src in ("restart", "frame", "resume", "last_state"):
        assert out[f"io.{src}.same_state"] == "1", (src, out[f"io.{src}.same_flags"])
    assert out["io.resume_energy_group"] == "1"


def test_disk_pair_arrays(disk_run):
    """Per-slot pair files of the final restart: layout N x row, fill values in unused
    slots, symmetric ids, antisymmetric forces and separations, and the harmonic contact's
    energy, force and Hessian blocks from U(r) = e_c/2 (1 - r/s)^2 (half per slot)."""
    out, root, _ = disk_run
    n, row = 96, int(out["io.row"])
    r = os.path.join(root, "system", "restart")
    ids = read_dat(os.path.join(r, "pair_ids.dat"), 2, as_int=True)
    assert ids.shape == (n * row, 2)
    f = read_dat(os.path.join(r, "force_pairs.dat"), 2)
    d = read_dat(os.path.join(r, "distance_pairs.dat"), 2)
    pe = read_dat(os.path.join(r, "potential_pairs.dat"))[:, 0]
    ov = read_dat(os.path.join(r, "overlap_pairs.dat"))[:, 0]
    rs = read_dat(os.path.join(r, "radsum_pairs.dat"))[:, 0]
    ang = read_dat(os.path.join(r, "pair_separation_angle.dat"))[:, 0]
    H = {k: read_dat(os.path.join(r, f"hessian_{k}.dat"))[:, 0]
         for k in ("pairs_xx", "pairs_xy", "pairs_yx", "pairs_yy", "ii_xx", "ii_xy", "ii_yx", "ii_yy")}
    used = ids[:, 0] >= 0
    assert used.sum() > n
    # fill values of unused slots
    assert np.all(ids[~used] == -1) and np.all(pe[~used] == -1.0) and np.all(f[~used] == 0.0)
    assert np.all(d[~used] == -1.0) and np.all(ov[~used] == -1.0) and np.all(rs[~used] == -1.0)
    assert np.all(ang[~used] == -1.0) and all(np.all(h[~used] == 0.0) for h in H.values())
    # every row holds one particle's neighbours; ids are static and the list is full
    rows = ids[:, 0].reshape(n, row)
    for i in range(n):
        vals = rows[i][rows[i] >= 0]
        assert len(set(vals)) <= 1
    slot = {(a, b): k for k, (a, b) in enumerate(ids) if a >= 0}
    assert all((b, a) in slot for (a, b) in slot)
    radii = read_dat(os.path.join(r, "radii.dat"))[:, 0]
    box = read_dat(os.path.join(r, "box_size.dat"))[:, 0]
    pos = read_dat(os.path.join(r, "positions.dat"), 2)
    e_c = 1.0
    for (a, b), k in slot.items():
        kk = slot[(b, a)]
        assert f[k, 0] == -f[kk, 0] and f[k, 1] == -f[kk, 1]
        assert pe[k] == pe[kk]
        s = radii[a] + radii[b]
        assert rs[k] == s
        dx = pos[a] - pos[b]
        dx -= box * np.round(dx / box)
        assert d[k] == pytest.approx(dx, abs=1e-14)
        rr = math.hypot(*d[k])
        assert ov[k] == pytest.approx(rr - s, abs=1e-14)
        assert ang[k] == pytest.approx(math.atan2(d[k, 1], d[k, 0]), abs=4e-16 * math.pi)
        hp = np.array([[H["pairs_xx"][k], H["pairs_xy"][k]], [H["pairs_yx"][k], H["pairs_yy"][k]]])
        hs = np.array([[H["ii_xx"][k], H["ii_xy"][k]], [H["ii_yx"][k], H["ii_yy"][k]]])
        if rr < s:
            x = 1 - rr / s
            u1, u2 = -e_c / s * x, e_c / s ** 2  # U'(r), U''(r)
            nn = np.outer(d[k], d[k]) / rr ** 2
            # 1 - r/s loses digits for small overlaps: allow a few ulps of r/s relative to x
            tol = max(1e-13, 8e-16 / x)
            assert pe[k] == pytest.approx(e_c / 4 * x * x, rel=2 * tol)
            assert f[k] == pytest.approx(-u1 * d[k] / rr, rel=tol, abs=1e-18)
            assert hs == pytest.approx(u2 * nn + u1 / rr * (np.eye(2) - nn), rel=tol, abs=1e-13)
            assert np.array_equal(hp, -hs)  # d2U/dx_i dx_j = -d2U/dx_i dx_i
        else:
            assert pe[k] == 0.0 and np.all(f[k] == 0) and np.all(hp == 0) and np.all(hs == 0)


def test_bumpy_pair_arrays(bumpy_run):
    """Rigid bumpy pair files: the pair force is the sum of the vertex forces between the
    two rings, its sum over a row is the particle's force (checked against Newton's third
    law here), angles are the two orientations, friction is |F_t|/|F_n|, the vertex
    overlaps are positive where vertices touch, and the 3x3 blocks satisfy H_ij,xy-part =
    -H_ii,xy-part; particle areas equal the vertex polygon plus, at every vertex, the
    sector of its disk outside the polygon (angle pi + 2 pi/m on a regular m-ring whose
    neighbouring vertices only touch): m/2 + 1 vertex disks in all."""
    out, root, _ = bumpy_run
    n, row = 48, int(out["io.row"])
    r = os.path.join(root, "system", "restart")
    ids = read_dat(os.path.join(r, "pair_ids.dat"), 2, as_int=True)
    f = read_dat(os.path.join(r, "force_pairs.dat"), 2)
    ai = read_dat(os.path.join(r, "angle_pairs_i.dat"))[:, 0]
    aj = read_dat(os.path.join(r, "angle_pairs_j.dat"))[:, 0]
    mu = read_dat(os.path.join(r, "pair_friction_coefficient.dat"))[:, 0]
    vo = read_dat(os.path.join(r, "pair_vertex_overlaps.dat"))[:, 0]
    vc = read_dat(os.path.join(r, "this_vertex_contact_counts.dat"), as_int=True)[:, 0]
    d = read_dat(os.path.join(r, "distance_pairs.dat"), 2)
    angles = read_dat(os.path.join(r, "angles.dat"))[:, 0]
    assert ids.shape == (n * row, 2)
    used = ids[:, 0] >= 0
    assert np.all(ai[~used] == -1.0) and np.all(vc[~used] == -1) and np.all(mu[~used] == -1.0)
    slot = {(a, b): k for k, (a, b) in enumerate(ids) if a >= 0}
    touching = 0
    for (a, b), k in slot.items():
        kk = slot[(b, a)]
        assert ai[k] == angles[a] and aj[k] == angles[b]
        assert f[k] == pytest.approx(-f[kk], rel=1e-12, abs=1e-16)
        if np.any(f[k] != 0):
            touching += 1
            assert vc[k] >= 1 and vo[k] > 0
            rr = math.hypot(*d[k])
            nx, ny = f[k, 0] * d[k, 0] / rr, f[k, 1] * d[k, 1] / rr
            ft = math.hypot(f[k, 0] - nx, f[k, 1] - ny)
            assert mu[k] == pytest.approx(ft / math.hypot(nx, ny), rel=1e-12)
        else:
            assert vc[k] == 0 and vo[k] == 0.0 and mu[k] == 0.0
        hp = {c: read_dat(os.path.join(r, f"hessian_pairs_{c}.dat"))[k, 0] for c in ("xx", "xy", "yx", "yy")}
        hs = {c: read_dat(os.path.join(r, f"hessian_ii_{c}.dat"))[k, 0] for c in ("xx", "xy", "yx", "yy")}
        for c in hp:
            assert hp[c] == pytest.approx(-hs[c], rel=1e-12, abs=1e-15)
        if touching > 12:
            break
    assert touching > 0
    # areas: ring polygon + m/2 + 1 vertex disks (segment = one vertex diameter: neighbours touch)
    area = read_dat(os.path.join(r, "area.dat"))[:, 0]
    nvp = read_dat(os.path.join(r, "num_vertices_in_particle.dat"), as_int=True)[:, 0]
    radii = read_dat(os.path.join(r, "radii.dat"))[:, 0]
    rv = (1.0 / (1.0 + 1.0 / math.sin(math.pi / 12))) / 2.0
    for i in range(n):
        R = radii[i] - rv
        m = nvp[i]
        expect = 0.5 * m * R * R * math.sin(2 * math.pi / m) + math.pi * rv * rv * (m / 2 + 1)
        assert area[i] == pytest.approx(expect, rel=1e-12)


# ------------------------------------------------------------------ GPU: same files
def _tree(root):
    files = {}
    for dp, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(dp, f)
            files[os.path.relpath(p, root)] = open(p).read()
    return files


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["disk", "bumpy"])
def test_io_run_over_hip_matches_oracle(kind, tmp_path_factory):
    o_out, o_root, o_dir = _io_run(tmp_path_factory, kind, BIN_ORACLE, "oracle")
    g_out, g_root, g_dir = _io_run(tmp_path_factory, kind, BIN_GPU, "gpu")
    for k in o_out:
        assert g_out[k] == o_out[k], k
    for sub in ("out", "out2"):
        a, b = _tree(os.path.join(g_dir, sub)), _tree(os.path.join(o_dir, sub))
        assert set(a) == set(b), set(a) ^ set(b)
        for name in sorted(b):
            if os.path.basename(name) in ("pair_separation_angle.dat", "area.dat"):
                x = np.array([float(t) for t in a[name].split()])
                y = np.array([float(t) for t in b[name].split()])
                assert np.all(np.abs(x - y) <= 2 * np.spacing(np.abs(y))), name
            else:
                assert a[name] == b[name], name
