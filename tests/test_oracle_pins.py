"""Pins the CPU oracle to what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures, so every oracle function
is checked here against closed forms, invariants or brute force, each chosen so a
plausible slip (a dropped term, a wrong sign or index, a transposed operand) fails:

  pair law (kernels.cuh:154-183)     closed form at one overlap, zero at contact,
                                     F = -dU/dr by finite differences
  pbcDistance (kernels.cuh:65-72)    round-half-away image choice at d = L/2
  cutoff (kernels.cuh:249-260)       strict '<' at exactly rc
  neighbour lists (base/kernels.cu)  cell == verlet == independent brute force
  force sums (disk/kernels.cu:11-57) Newton's third law; PE = sum of pair energies
  NVE (nve.cu:15-25)                 2-body oscillator closed form; time reversal;
                                     O(dt^2) energy error
  FIRE (fire.cu:13-46)               state-machine known answers; two disks relax
  Adam (base/kernels.cu:550-611)     first-step known answer
  rigid Adam (rigid_bumpy/kernels.cu: first-step known answer for centres, angles and
    1004-1093)                       vertex rings; rotation off keeps angles and rings
  stress (disk/kernels.cu:203-251)   2-particle closed form; virial == -dU/dA
  packing fraction (disk.cu:205-207) pi*sum r^2 / A
  cell reorder (particle.cu:1292-)   stable sort by cell id
  rigid geometry (rigid_bumpy)       ring radius, chord, I = m (R - r_v)^2, rotation
                                     round trip, torque balance
  contacts (disk/kernels.cu:177-201, all-pairs brute force (strict <, exact touch not
    rigid_bumpy/kernels.cu:924-997)  counted); rigid: brute force over vertex pairs of
                                     the reference's counting rule
  rigid stress (rigid_bumpy/kernels. two facing vertices: P = d f / 2A; moving spinning
    cu:705-789)                      ring: M V V + I w^2 / 2; virial == -dU/dA with the
                                     vertices scaled affinely
  rigid pair Hessian (rigid_bumpy/   every H_ij and H_ii entry (theta ones included) by
    kernels.cu:482-702)              central differences of the ring-pair energy
  jamming controller (runArthurJam-  hand-derived trace on a square lattice whose
    WithFire.cpp:64-106)             pressure P(L) is closed form
  FIRE's force floor                 a rigid packing translated by 1e4 reaches the same
                                     minimum but not |F|max < 1e-12
"""
import math

import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()


def orc_system(n, side, x, y, rad, mass=None, method=None, rc=None, vx=None, vy=None,
               kind=0, nv=0):
    s = P.System(kind, n, nv, backend=oracle())
    s.set_box(side, side)
    s.set_energy(1.0, 2.0)
    s.set_state(np.asarray(x, float), np.asarray(y, float), np.asarray(rad, float),
                np.ones(n) if mass is None else mass, vx, vy)
    if method is not None:
        if method == P.NBR_ALL:
            s.neighbor_config(P.NBR_ALL)
        else:
            spec = W.neighbor_spec(side, 2 * max(rad), (rc or 1.7) - 2 * max(rad), method)
            W.configure(s, spec)
        s.neighbor_init()
    return s


def forces(s):
    s.zero_forces()
    s.compute_forces()
    return s.download(P.F_FX), s.download(P.F_FY), s.download(P.F_PE)


# ---------------------------------------------------------------- pair law
def test_pair_law_closed_form():
    # sigma = 1.2, d = 1.0: overlap = 1/6, |F| = e_c*ov/sigma, U_pair = ov^2/2, each gets half
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    fx, fy, pe = forces(s)
    ov = 1.0 - 1.0 / 1.2
    assert fx[0] == pytest.approx(-ov / 1.2, rel=1e-15)
    assert fx[1] == pytest.approx(ov / 1.2, rel=1e-15)
    assert fy[0] == 0.0 and fy[1] == 0.0
    assert pe[0] == pytest.approx(ov * ov / 4.0, rel=1e-15)
    assert pe.sum() == pytest.approx(ov * ov / 2.0, rel=1e-15)


def test_pair_law_zero_at_contact_and_beyond():
    for d in (1.2, 1.25, 1.6):
        s = orc_system(2, 10.0, [3.0, 3.0 + d], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
        fx, fy, pe = forces(s)
        assert np.all(fx == 0) and np.all(pe == 0)


def test_pair_force_is_minus_energy_gradient():
    h = 1e-6
    def energy(d):
        s = orc_system(2, 10.0, [3.0, 3.0 + d * 0.6], [5.0, 5.0 + d * 0.8], [0.5, 0.7],
                       method=P.NBR_ALL)
        return forces(s)[2].sum()
    d0 = 1.05
    s = orc_system(2, 10.0, [3.0, 3.0 + d0 * 0.6], [5.0, 5.0 + d0 * 0.8], [0.5, 0.7],
                   method=P.NBR_ALL)
    fx, fy, _ = forces(s)
    dudr = (energy(d0 + h) - energy(d0 - h)) / (2 * h)
    # force on particle 1 along +r equals -dU/dr
    f_along = fx[1] * 0.6 + fy[1] * 0.8
    assert f_along == pytest.approx(-dudr, rel=1e-7)


def test_general_exponent():
    # n_c = 2.5: U_pair = e/n * ov^n, |F| = e ov^(n-1) / sigma
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    s.set_energy(2.0, 2.5)
    fx, _, pe = forces(s)
    ov = 1.0 - 1.0 / 1.2
    assert fx[1] == pytest.approx(2.0 * ov ** 1.5 / 1.2, rel=1e-14)
    assert pe.sum() == pytest.approx(2.0 / 2.5 * ov ** 2.5, rel=1e-14)


def test_minimum_image_across_boundary():
    L = 10.0
    s = orc_system(2, L, [0.2, L - 0.6], [5.0, 5.0], [0.5, 0.5], method=P.NBR_ALL)
    fx, _, _ = forces(s)
    ov = 1.0 - 0.8 / 1.0
    assert fx[0] == pytest.approx(ov, rel=1e-15)   # pushed to +x, away from the image at -0.6
    assert fx[1] == pytest.approx(-ov, rel=1e-15)


def test_half_box_rounds_away_from_zero():
    # d = x1 - x2 = +L/2 exactly: round(0.5) = 1 (half away from zero) gives dx = -L/2,
    # so particle 1 is pushed towards -x; round-half-even would push it to +x.
    L = 2.0
    s = orc_system(2, L, [1.5, 0.5], [1.0, 1.0], [0.6, 0.6], method=P.NBR_ALL)
    fx, _, _ = forces(s)
    assert fx[0] < 0 and fx[1] > 0


# ---------------------------------------------------------------- neighbours
def test_cutoff_is_strict():
    # rc = 1.5 exactly; particles 1.5 apart are not neighbours, 1.4999 apart are
    for gap, expect in ((1.5, 0), (1.4999, 2)):
        s = P.System(0, 2, 0, backend=oracle())
        s.set_box(20.0, 20.0)
        s.set_state(np.array([5.0, 5.0 + gap]), np.array([5.0, 5.0]), np.array([0.5, 0.5]),
                    np.ones(2))
        s.neighbor_config(P.NBR_VERLET, 1.5, 1.0, 0, 0.0)
        s.neighbor_init()
        i, j = s.neighbor_pairs()
        assert len(i) == expect


def brute_pairs(x, y, L, rc):
    n = len(x)
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dx -= L * np.trunc(dx / L + np.copysign(0.5, dx))
    dy -= L * np.trunc(dy / L + np.copysign(0.5, dy))
    d2 = dx * dx + dy * dy
    amb = np.abs(d2 - rc * rc) < 1e-9
    inside = (d2 < rc * rc) & ~np.eye(n, dtype=bool)
    return inside, amb


@pytest.mark.parametrize("method", [2, 3])
def test_pair_sets_match_brute_force(method):
    st = W.disk_config(400, 0.8, seed=3)
    spec = W.neighbor_spec(st.side, method=method)
    s = P.System(0, st.n, 0, backend=oracle())
    s.set_box(st.side, st.side)
    s.set_state(st.x, st.y, st.rad, st.mass)
    W.configure(s, spec)
    s.neighbor_init()
    i, j = s.neighbor_pairs()
    got = np.zeros((st.n, st.n), bool)
    got[i, j] = True
    assert len(i) == got.sum()  # no duplicates
    want, amb = brute_pairs(st.x, st.y, st.side, spec.cutoff)
    assert np.array_equal(got & ~amb, want & ~amb)
    assert np.array_equal(got, got.T)  # full (both-direction) list


def test_cell_reorder_is_stable_sort_by_cell():
    st = W.disk_config(500, 0.7, seed=5)
    spec = W.neighbor_spec(st.side)
    s = W.make_disk_system(P.System, st, spec, backend=oracle())
    order = s.download(P.F_STATIC_INDEX)
    cs = st.side / spec.n_cells_dim
    n = spec.n_cells_dim
    cx = np.mod(np.floor(st.x / cs).astype(np.int64), n)
    cy = np.mod(np.floor(st.y / cs).astype(np.int64), n)
    cell = cx + cy * n
    assert np.array_equal(order, np.argsort(cell, kind="stable"))


def test_neighbor_capacity_grows_as_power_of_two():
    st = W.disk_config(300, 0.9, seed=9)
    s = W.make_disk_system(P.System, st, backend=oracle())
    stats = s.neighbor_stats()
    nn = s.download(P.F_NUM_NEIGHBORS)
    m = int(nn.max())
    assert stats["max_alloc"] == max(4, 2 ** math.ceil(math.log2(m)))


# ---------------------------------------------------------------- force sums
def test_newton_third_law_and_energy_sum():
    st = W.disk_config(600, 0.85, seed=11)
    s = W.make_disk_system(P.System, st, backend=oracle())
    fx, fy, pe = forces(s)
    fb = s.reduce(P.RED_FORCE_BALANCE)
    assert fb / st.n < 1e-14
    # independent brute-force energy
    dx = st.x[:, None] - st.x[None, :]
    dy = st.y[:, None] - st.y[None, :]
    L = st.side
    dx -= L * np.trunc(dx / L + np.copysign(0.5, dx))
    dy -= L * np.trunc(dy / L + np.copysign(0.5, dy))
    d = np.sqrt(dx * dx + dy * dy)
    sig = st.rad[:, None] + st.rad[None, :]
    ov = np.where((d < sig) & ~np.eye(st.n, dtype=bool), 1 - d / sig, 0.0)
    u = 0.5 * (ov ** 2).sum() / 2.0  # each unordered pair counted twice
    assert pe.sum() == pytest.approx(u, rel=1e-12)
    with np.errstate(divide="ignore", invalid="ignore"):
        fm = np.where(ov > 0, ov / (sig * d), 0.0)
    np.testing.assert_allclose(fx, (fm * dx).sum(axis=1), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(fy, (fm * dy).sum(axis=1), rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- NVE
def test_two_body_oscillator():
    sigma, delta0 = 1.2, 0.05
    s = orc_system(2, 10.0, [3.0, 3.0 + sigma - delta0], [5.0, 5.0], [0.6, 0.6],
                   method=P.NBR_ALL)
    forces(s)
    omega = math.sqrt(2.0) / sigma
    dt = 1e-4
    n = int(round((math.pi / 4) / omega / dt))
    s.nve_run(dt, n)
    x = s.download(P.F_X)
    t = n * dt
    expect = sigma - delta0 * math.cos(omega * t)
    assert (x[1] - x[0]) == pytest.approx(expect, abs=1e-8)


def test_time_reversibility():
    st = W.disk_config(256, 0.82, seed=21)
    s = W.make_disk_system(P.System, st, backend=oracle())
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 1)
    s.upload(P.F_VX, vx)
    s.upload(P.F_VY, vy)
    forces(s)
    x0, y0 = s.download(P.F_X), s.download(P.F_Y)
    s.nve_run(0.005, 200)
    s.upload(P.F_VX, -s.download(P.F_VX))
    s.upload(P.F_VY, -s.download(P.F_VY))
    s.nve_run(0.005, 200)
    np.testing.assert_allclose(s.download(P.F_X), x0, atol=1e-10)
    np.testing.assert_allclose(s.download(P.F_Y), y0, atol=1e-10)


def energy_fluct(dt, steps):
    st = W.disk_config(128, 0.84, seed=4)
    s = W.make_disk_system(P.System, st, backend=oracle())
    s.minimize_fire(P.fire_params(max_steps=200))
    vx, vy = W.gaussian_velocities(st.n, 1e-3, 2)
    s.upload(P.F_VX, vx)
    s.upload(P.F_VY, vy)
    forces(s)
    es = []
    for _ in range(steps):
        s.nve_run(dt, int(round(0.05 / dt)))
        es.append(s.reduce(P.RED_PE) + s.reduce(P.RED_KE))
    es = np.array(es)
    return es.max() - es.min()


def test_energy_error_scales_as_dt_squared():
    a = energy_fluct(0.02, 20)
    b = energy_fluct(0.01, 20)
    assert 2.5 < a / b < 6.0


# ---------------------------------------------------------------- FIRE
def test_fire_state_machine_known_answers():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    s.zero_velocities()
    forces(s)
    st = s.fire_state(0.1, 0.01)
    s.fire_step(st)
    # from rest the power is 0 -> uphill branch: dt halves, alpha resets, N_bad = 1
    assert st.power == 0.0 and st.n_bad == 1 and st.n_good == 0
    assert st.dt == pytest.approx(0.005) and st.alpha == pytest.approx(0.1)
    for k in range(5):
        s.fire_step(st)
        assert st.power > 0 and st.n_good == k + 1 and st.dt == pytest.approx(0.005)
    s.fire_step(st)  # N_good = 6 > N_min = 5: dt *= 1.1, alpha *= 0.99
    assert st.n_good == 6
    assert st.dt == pytest.approx(0.0055) and st.alpha == pytest.approx(0.099)


def test_fire_uphill_back_step_known_answer():
    # velocities against the forces -> P < 0: dt is halved first, then x -= v*dt_new/2,
    # v = 0, kick(dt_new/2), mix (v parallel to f: unchanged), drift(dt_new)
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL,
                   vx=np.array([0.3, -0.3]), vy=np.zeros(2))
    fx, _, _ = forces(s)
    x0 = s.download(P.F_X)
    st = s.fire_state(0.1, 0.01)
    s.fire_step(st)
    assert st.power < 0 and st.n_bad == 1
    dt = 0.005
    v0 = np.array([0.3, -0.3])
    expect = x0 - v0 * dt / 2.0 + fx * (dt / 2.0) * dt
    np.testing.assert_allclose(s.download(P.F_X), expect, rtol=0, atol=1e-15)


def test_fire_relaxes_two_disks_to_contact():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    r = s.minimize_fire(P.fire_params(avg_pe_target=1e-20, avg_pe_diff_target=0.0,
                                      max_steps=100000))
    x = s.download(P.F_X)
    assert r.converged == 1
    assert (x[1] - x[0]) >= 1.2 - 1e-9
    assert s.reduce(P.RED_PE) < 2e-20


def test_fire_three_disk_chain():
    s = orc_system(3, 10.0, [3.0, 3.9, 4.8], [5.0, 5.0, 5.0], [0.5, 0.5, 0.5],
                   method=P.NBR_ALL)
    s.minimize_fire(P.fire_params(avg_pe_target=1e-20, avg_pe_diff_target=0.0))
    x = np.sort(s.download(P.F_X))
    assert np.all(np.diff(x) >= 1.0 - 1e-9)
    assert abs(s.download(P.F_Y) - 5.0).max() < 1e-12  # collinear forces keep y


def test_fire_max_force_criterion():
    st = W.disk_config(200, 0.86, seed=8)
    s = W.make_disk_system(P.System, st, backend=oracle())
    r = s.minimize_fire(P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                      fmax_target=1e-10, max_steps=200000))
    assert r.converged == 1
    fx, fy = s.download(P.F_FX), s.download(P.F_FY)
    assert np.max(fx * fx + fy * fy) < 1e-20


# ---------------------------------------------------------------- Adam
def test_adam_first_step_known_answer():
    s = orc_system(2, 10.0, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    x0 = s.download(P.F_X)
    alpha, eps = 1e-3, 1e-8
    s.adam_step(0, alpha, 0.9, 0.999, eps)
    fx = s.download(P.F_FX)
    # m_hat = -f, v_hat = f^2 on the first step: dx = alpha * f / (|f| + eps)
    np.testing.assert_allclose(s.download(P.F_X) - x0, alpha * fx / (np.abs(fx) + eps),
                               rtol=1e-10)
    np.testing.assert_allclose(s.download(P.F_Y), 5.0, rtol=0, atol=0)


# ---------------------------------------------------------------- stress
def test_pressure_two_particles_closed_form():
    L = 10.0
    s = orc_system(2, L, [3.0, 4.0], [5.0, 5.0], [0.6, 0.6], method=P.NBR_ALL)
    p = s.reduce(P.RED_PRESSURE)
    ov = 1.0 - 1.0 / 1.2
    f = ov / 1.2
    assert p == pytest.approx(1.0 * f / (2 * L * L), rel=1e-14)


def test_virial_pressure_equals_minus_dU_dA():
    st = W.disk_config(300, 0.9, seed=13)
    s = W.make_disk_system(P.System, st, backend=oracle())
    s.minimize_fire(P.fire_params(max_steps=300))
    p = s.reduce(P.RED_PRESSURE)
    x, y = s.download(P.F_X), s.download(P.F_Y)
    L = s.box()[0]
    eps = 1e-6

    def u_at(f):
        t = orc_system(st.n, L * f, x * f, y * f, st.rad, method=3)
        return forces(t)[2].sum()

    dA = (L * (1 + eps)) ** 2 - (L * (1 - eps)) ** 2
    p_fd = -(u_at(1 + eps) - u_at(1 - eps)) / dA
    assert p == pytest.approx(p_fd, rel=1e-5)


def test_packing_fraction():
    st = W.disk_config(300, 0.77, seed=1)
    s = W.make_disk_system(P.System, st, backend=oracle())
    assert s.reduce(P.RED_PACKING_FRACTION) == pytest.approx(0.77, rel=1e-13)
    s.scale_to_packing_fraction(0.8)
    assert s.reduce(P.RED_PACKING_FRACTION) == pytest.approx(0.8, rel=1e-13)


def test_canonical_sum_accuracy():
    rng = np.random.default_rng(0)
    lib = oracle().lib
    f = lib.orc_canonical_sum
    import ctypes as C
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_int64]
    for n in (1, 63, 64, 65, 4096, 100001):
        a = rng.integers(-1000, 1000, n).astype(np.float64)
        assert f(a.ctypes.data, n) == float(a.sum())  # exact for integers
        b = rng.standard_normal(n)
        assert f(b.ctypes.data, n) == pytest.approx(math.fsum(b), rel=1e-12, abs=1e-12)


# ---------------------------------------------------------------- rigid bumpy
def bumpy_small(n=2, rotation=True, x=None, y=None, angle=None, side=20.0):
    rad = np.array([0.7, 0.5][:n] if n <= 2 else W.bidisperse_radii(n))
    nverts = np.where(rad == 0.7, 28, 20).astype(np.int64)
    rv = W.vertex_radius()
    s = P.System(1, n, int(nverts.sum()), backend=oracle())
    s.set_box(side, side)
    s.set_energy(1.0, 2.0)
    s.set_state(np.asarray(x, float), np.asarray(y, float), rad, np.ones(n))
    s.upload(P.F_ANGLE, np.asarray(angle, float))
    s.bumpy_setup(nverts, rv, rotation)
    return s, rad, nverts, rv


def test_bumpy_ring_geometry_and_inertia():
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 12.0], y=[5.0, 6.0], angle=[0.3, 1.1])
    vx, vy = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    vm = s.download(P.F_VERTEX_MASS)
    moi = s.download(P.F_MOI)
    start = 0
    for p, (cx, cy) in enumerate([(5.0, 5.0), (12.0, 6.0)]):
        n = nverts[p]
        px, py = vx[start:start + n], vy[start:start + n]
        R = rad[p] - rv
        np.testing.assert_allclose(np.hypot(px - cx, py - cy), R, rtol=1e-14)
        chord = np.hypot(np.diff(px), np.diff(py))
        np.testing.assert_allclose(chord, 2 * R * math.sin(math.pi / n), rtol=1e-12)
        assert vm[start:start + n].sum() == pytest.approx(1.0, rel=1e-14)
        assert moi[p] == pytest.approx(R * R, rel=1e-13)
        start += n
    assert rv == pytest.approx(0.0676365, rel=1e-6)  # value quoted for cfg 3


def test_bumpy_rotation_round_trip():
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 12.0], y=[5.0, 6.0], angle=[0.3, 1.1])
    vx0, vy0 = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    s.upload(P.F_OMEGA, np.array([3.0, -2.0]))
    s.drift(0.1)
    th = s.download(P.F_ANGLE)
    assert th[0] == pytest.approx(0.6) and th[1] == pytest.approx(0.9)
    s.upload(P.F_OMEGA, np.array([-3.0, 2.0]))
    s.drift(0.1)
    np.testing.assert_allclose(s.download(P.F_VERTEX_X), vx0, atol=1e-13)
    np.testing.assert_allclose(s.download(P.F_VERTEX_Y), vy0, atol=1e-13)


def test_bumpy_torque_balance_and_contact():
    # two rings close enough that a few vertices overlap
    s, rad, nverts, rv = bumpy_small(2, x=[5.0, 6.15], y=[5.0, 5.05], angle=[0.0, 0.1])
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    fx, fy, pe = forces(s)
    tq = s.download(P.F_TORQUE)
    assert pe.sum() > 0
    assert abs(fx.sum()) < 1e-15 and abs(fy.sum()) < 1e-15
    # angular momentum balance about the origin: sum(tau_i + x_i x F_i) = 0
    x, y = np.array([5.0, 6.15]), np.array([5.0, 5.05])
    total = (tq + x * fy - y * fx).sum()
    assert abs(total) < 1e-13


@pytest.mark.parametrize("rotation", [True, False])
def test_bumpy_adam_first_step_known_answer(rotation):
    # two overlapping rings: on the first step m_hat = -f, v_hat = f^2 for the force and
    # the torque alike, so the centre moves by alpha f / (|f| + eps), the angle by
    # alpha tau / (|tau| + eps) (rotation on) and every vertex keeps its place in the
    # ring, turned by that angle about the moved centre
    x0, y0, a0 = np.array([5.0, 6.15]), np.array([5.0, 5.05]), np.array([0.0, 0.1])
    s, rad, nverts, rv = bumpy_small(2, rotation=rotation, x=x0, y=y0, angle=a0)
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    vx0, vy0 = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    alpha, eps = 1e-3, 1e-8
    s.adam_reset()
    s.adam_step(0, alpha, 0.9, 0.999, eps)
    fx, fy, tq = s.download(P.F_FX), s.download(P.F_FY), s.download(P.F_TORQUE)
    assert np.all(np.abs(fx) > 1e-6) and np.all(np.abs(tq) > 1e-9)
    x1, y1, a1 = s.download(P.F_X), s.download(P.F_Y), s.download(P.F_ANGLE)
    np.testing.assert_allclose(x1 - x0, alpha * fx / (np.abs(fx) + eps), rtol=1e-10)
    np.testing.assert_allclose(y1 - y0, alpha * fy / (np.abs(fy) + eps), rtol=1e-10)
    da = alpha * tq / (np.abs(tq) + eps) if rotation else np.zeros(2)
    np.testing.assert_allclose(a1 - a0, da, rtol=1e-10, atol=0)
    owner = np.repeat(np.arange(2), nverts)
    rx, ry = vx0 - x0[owner], vy0 - y0[owner]
    c, sn = np.cos(da[owner]), np.sin(da[owner])
    np.testing.assert_allclose(s.download(P.F_VERTEX_X), x1[owner] + rx * c - ry * sn,
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(s.download(P.F_VERTEX_Y), y1[owner] + rx * sn + ry * c,
                               rtol=0, atol=1e-13)


def test_bumpy_adam_relaxes_overlaps():
    # Adam from overlapping rings lowers the energy; the rings stay rigid
    x0, y0 = np.array([5.0, 6.15]), np.array([5.0, 5.05])
    s, rad, nverts, rv = bumpy_small(2, x=x0, y=y0, angle=[0.0, 0.1])
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    s.zero_forces()
    s.compute_forces()
    pe0 = s.reduce(P.RED_PE)
    steps = s.minimize_adam(alpha=1e-3, max_steps=2000)
    assert 0 < steps <= 2000
    assert s.reduce(P.RED_PE) < 1e-3 * pe0
    x, y = s.download(P.F_X), s.download(P.F_Y)
    owner = np.repeat(np.arange(2), nverts)
    d = np.hypot(s.download(P.F_VERTEX_X) - x[owner], s.download(P.F_VERTEX_Y) - y[owner])
    ring = (rad - rv)[owner]
    np.testing.assert_allclose(d, ring, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- shared sin/cos
def _sincos(x):
    import ctypes as C

    from _helpers import ORACLE_SO
    lib = C.CDLL(ORACLE_SO)
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, c = np.empty_like(x), np.empty_like(x)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.orc_sincos.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
    assert lib.orc_sincos(ptr(x), ptr(s), ptr(c), x.size) == 0
    return s, c


@pytest.mark.parametrize("lim,max_ulp", [(math.pi / 4, 1.0), (7.0, 2.0), (100.0, 2.0),
                                         (1e5, 3.0)])
def test_shared_sincos_against_libm(lim, max_ulp):
    # dpm_sincos (rigid-body rotations on both sides) vs the C library's correctly
    # rounded-to-0.5-ulp sin/cos: together at most max_ulp apart
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(-lim, lim, 200000),
                        [0.0, -0.0, 1e-300, 5e-324, math.pi / 4, math.pi / 2, math.pi,
                         -math.pi / 2, 3 * math.pi / 2, 2 * math.pi]])
    x = x[np.abs(x) <= lim]
    s, c = _sincos(x)
    for got, ref in ((s, np.sin(x)), (c, np.cos(x))):
        err = np.abs(got - ref) / np.spacing(np.maximum(np.abs(ref), 2.0 ** -1022))
        assert err.max() <= max_ulp, (lim, err.max())
    assert np.all(s * s + c * c == pytest.approx(1.0, abs=4e-16))
    # exact special values and symmetries of the reduction
    s0, c0 = _sincos(np.array([0.0, -0.0]))
    assert s0[0] == 0.0 and c0[0] == 1.0 and math.copysign(1, s0[1]) == -1.0
    sp, cp = _sincos(x)
    sn, cn = _sincos(-x)
    assert np.array_equal(sn, -sp) and np.array_equal(cn, cp)


# ---------------------------------------------------------------- contacts
def _min_image(d, L):
    # pbcDistance (kernels.cuh:65-72): d - L * round(d / L), round half away from zero
    t = d / L
    return d - L * np.trunc(t + np.copysign(0.5, t))


def test_disk_contacts_brute_force():
    """countContacts (disk/kernels.cu:177-201): per particle, the listed neighbours with
    centre distance strictly below r_i + r_j -- checked against an all-pairs count with
    the minimum image, including a pair placed exactly at contact (not counted)."""
    st = W.disk_config(400, 0.88, seed=21)
    x, y = st.x.copy(), st.y.copy()
    x[1], y[1] = x[0] + (st.rad[0] + st.rad[1]), y[0]  # exactly touching: not a contact
    s = orc_system(st.n, st.side, x, y, st.rad, method=3)
    s.count_contacts()
    got = s.download(P.F_CONTACTS)
    dx = _min_image(x[:, None] - x[None, :], st.side)
    dy = _min_image(y[:, None] - y[None, :], st.side)
    touch = np.sqrt(dx * dx + dy * dy) < st.rad[:, None] + st.rad[None, :]
    np.fill_diagonal(touch, False)
    expect = touch.sum(axis=1)
    assert np.array_equal(got, expect)
    assert expect.sum() > st.n  # a dense packing: many contacts
    assert not touch[0, 1] or np.hypot(dx[0, 1], dy[0, 1]) < st.rad[0] + st.rad[1]
    assert s.reduce(P.RED_CONTACTS) == float(expect.sum())


def _bumpy_random(n, phi, seed, e_c=1.0):
    st = W.bumpy_config(n, phi, seed=seed)
    s = P.System(1, st.n, st.nv, backend=oracle())
    s.set_box(st.side, st.side)
    s.set_energy(e_c, 2.0)
    s.set_state(st.x, st.y, st.rad, st.mass)
    s.upload(P.F_ANGLE, st.angle)
    s.bumpy_setup(st.nverts, st.rv, True)
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * st.rv, 1.4)
    s.neighbor_init()
    return s, st


def test_bumpy_contacts_brute_force():
    """countContacts for rigid bumpy particles (rigid_bumpy/kernels.cu:924-997): particle
    q counts as a contact of p when some vertex of p is closer than 2 r_v to a vertex of
    q; the vertex-contact count adds, for each such q, the number of q's vertices touching
    the first vertex of p (in ring order) that touches q.  Brute force over all vertex
    pairs of the downloaded rings."""
    s, st = _bumpy_random(40, 0.86, seed=5)
    s.count_contacts()
    nc, nvc = s.download(P.F_CONTACTS), s.download(P.F_VERTEX_CONTACTS)
    vx, vy = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    start = np.concatenate([[0], np.cumsum(st.nverts)[:-1]])
    exp_c, exp_v = np.zeros(st.n, int), np.zeros(st.n, int)
    for p in range(st.n):
        vp = slice(start[p], start[p] + st.nverts[p])
        for q in range(st.n):
            if q == p:
                continue
            vq = slice(start[q], start[q] + st.nverts[q])
            dx = _min_image(vx[vp][:, None] - vx[vq][None, :], st.side)
            dy = _min_image(vy[vp][:, None] - vy[vq][None, :], st.side)
            t = np.sqrt(dx * dx + dy * dy) < 2 * st.rv
            rows = np.flatnonzero(t.any(axis=1))
            if rows.size:
                exp_c[p] += 1
                exp_v[p] += int(t[rows[0]].sum())
    assert np.array_equal(nc, exp_c) and np.array_equal(nvc, exp_v)
    assert exp_c.sum() > st.n and np.any(exp_v > exp_c)  # several touching vertices occur


# ---------------------------------------------------------------- rigid stress
def test_bumpy_pressure_two_rings_closed_form():
    """kernelCalcRigidBumpyStressTensor (rigid_bumpy/kernels.cu:705-789) with one touching
    vertex pair on the line of centres: sigma summed over particles is d f / A (each side
    adds d f / 2), so P = d f / (2 A) with f = e_c (1 - d / 2 r_v) / (2 r_v)."""
    rv = W.vertex_radius()
    L = 20.0
    R0, R1 = 0.7 - rv, 0.5 - rv
    dist = 0.13  # < 2 r_v = 0.1353: only the two facing vertices touch
    x0, x1 = 5.0, 5.0 + R0 + dist + R1
    # ring 0 (28 vertices) has a vertex at angle 0; ring 1 (20 vertices) at angle pi
    s, rad, nverts, rv = bumpy_small(2, x=[x0, x1], y=[7.0, 7.0], angle=[0.0, math.pi],
                                     side=L)
    s.set_energy(0.8, 2.0)
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    vx = s.download(P.F_VERTEX_X)
    d_actual = vx[28] - vx[0]  # vertex 0 of each ring sits at its angle: facing each other
    s.compute_stress()
    sig = 2 * rv
    f = 0.8 * (1 - d_actual / sig) / sig
    assert d_actual == pytest.approx(dist, abs=1e-12)
    p = s.reduce(P.RED_PRESSURE)
    assert p == pytest.approx(d_actual * f / (2 * L * L), rel=1e-12)
    assert s.reduce(P.RED_STRESS_YY) == pytest.approx(0.0, abs=1e-25)
    s.count_contacts()
    assert list(s.download(P.F_VERTEX_CONTACTS)) == [1, 1]


def test_bumpy_kinetic_stress_of_a_moving_spinning_ring():
    """The kinetic part of the rigid stress sums m_v v_v v_v over the vertices with
    v_v = V + omega z x r_v: for a regular ring, sum m_v r_v = 0 and sum m_v r_y^2 =
    sum m_v r_x^2 = I/2, so sigma_xx A = M Vx^2 + I omega^2 / 2, sigma_xy A = M Vx Vy."""
    s, rad, nverts, rv = bumpy_small(2, x=[3.0, 12.0], y=[3.0, 12.0], angle=[0.4, 1.3])
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    V = np.array([[0.3, -0.2], [-0.1, 0.25]])
    w = np.array([1.7, -0.6])
    s.upload(P.F_VX, V[:, 0])
    s.upload(P.F_VY, V[:, 1])
    s.upload(P.F_OMEGA, w)
    s.compute_stress()
    A = 20.0 * 20.0
    I = s.download(P.F_MOI)
    sxx, sxy, syy = (s.download(getattr(P, f)) for f in ("F_STRESS_XX", "F_STRESS_XY",
                                                          "F_STRESS_YY"))
    np.testing.assert_allclose(sxx * A, V[:, 0] ** 2 + I * w * w / 2, rtol=1e-12)
    np.testing.assert_allclose(syy * A, V[:, 1] ** 2 + I * w * w / 2, rtol=1e-12)
    np.testing.assert_allclose(sxy * A, V[:, 0] * V[:, 1], rtol=1e-11, atol=1e-15)


def _vertex_energy(vx, vy, owner, L, sig, e_c):
    """Sum over vertex pairs of different particles of the harmonic pair energy
    (e_c / 2) (1 - d / sig)^2 (each pair once)."""
    dx = _min_image(vx[:, None] - vx[None, :], L)
    dy = _min_image(vy[:, None] - vy[None, :], L)
    d = np.sqrt(dx * dx + dy * dy)
    other = owner[:, None] != owner[None, :]
    ov = np.where(other & (d < sig), 1 - d / sig, 0.0)
    return 0.5 * (0.5 * e_c * ov * ov).sum()


def test_bumpy_virial_pressure_equals_minus_dU_dA():
    """The rigid pressure's virial sums d_uv . f_uv over touching vertex pairs, which is
    -dU/dA when every vertex position and the box scale together (affine in the vertices);
    U and the derivative come from the downloaded vertex positions alone."""
    s, st = _bumpy_random(30, 0.86, seed=9, e_c=0.7)
    s.compute_stress()
    p = s.reduce(P.RED_PRESSURE)
    vx, vy = s.download(P.F_VERTEX_X), s.download(P.F_VERTEX_Y)
    owner = np.repeat(np.arange(st.n), st.nverts)
    L, sig = st.side, 2 * st.rv
    s.zero_forces()
    s.compute_forces()
    u0 = _vertex_energy(vx, vy, owner, L, sig, 0.7)
    assert u0 == pytest.approx(s.reduce(P.RED_PE), rel=1e-12) and u0 > 0
    eps = 1e-6
    up = _vertex_energy(vx * (1 + eps), vy * (1 + eps), owner, L * (1 + eps), sig, 0.7)
    um = _vertex_energy(vx * (1 - eps), vy * (1 - eps), owner, L * (1 - eps), sig, 0.7)
    dA = (L * (1 + eps)) ** 2 - (L * (1 - eps)) ** 2
    assert p == pytest.approx(-(up - um) / dA, rel=1e-6)


# ---------------------------------------------------------------- rigid pair Hessian
def test_bumpy_pair_hessian_blocks_by_finite_differences():
    """kernelCalcRigidBumpyForceDistancePairs (rigid_bumpy/kernels.cu:482-702): the 3x3
    blocks H_ij and H_ii of a pair row are the second derivatives of the pair's full
    vertex energy with respect to (x_i, y_i, theta_i) x (x_j, y_j, theta_j) and
    (x_i, y_i, theta_i) x (x_i, y_i, theta_i), where vertex u of ring i sits at
    X_i + R_i (cos(theta_i + 2 pi u / n_i), sin(...)).  Central differences of that energy
    (computed here from the ring geometry) pin every entry, theta ones included."""
    rv = W.vertex_radius()
    x, y, th = np.array([5.0, 6.13]), np.array([5.0, 5.21]), np.array([0.23, 1.01])
    s, rad, nverts, rv = bumpy_small(2, x=x, y=y, angle=th, side=20.0)
    s.set_energy(0.9, 2.0)
    s.neighbor_config(P.NBR_ALL)
    s.bumpy_neighbor_config(1.5 * 2 * rv, 1.4)
    s.neighbor_init()
    pa = s.pair_analytics()
    R = rad - rv
    sig = 2 * rv

    def energy(q):  # q = (x_i, y_i, th_i, x_j, y_j, th_j) of rings 0 and 1
        pts = []
        for k in range(2):
            n = nverts[k]
            a = q[3 * k + 2] + 2 * math.pi * np.arange(n) / n
            pts.append((q[3 * k] + R[k] * np.cos(a), q[3 * k + 1] + R[k] * np.sin(a)))
        dx = pts[0][0][:, None] - pts[1][0][None, :]
        dy = pts[0][1][:, None] - pts[1][1][None, :]
        d = np.sqrt(dx * dx + dy * dy)
        ov = np.where(d < sig, 1 - d / sig, 0.0)
        return (0.5 * 0.9 * ov * ov).sum(), d

    q0 = np.array([x[0], y[0], th[0], x[1], y[1], th[1]])
    e0, d0 = energy(q0)
    assert e0 > 0 and np.sum(d0 < sig) >= 2
    assert np.min(np.abs(d0 - sig)) > 1e-3  # the touching set is stable under the steps
    h = 1e-5

    def d2(a, b):
        def e(sa, sb):
            q = q0.copy()
            q[a] += sa * h
            q[b] += sb * h
            return energy(q)[0]
        return (e(1, 1) - e(1, -1) - e(-1, 1) + e(-1, -1)) / (4 * h * h)

    ids_i, ids_j = pa["pair_ids_i"], pa["pair_ids_j"]
    row, slot = [(r, c) for r in range(2) for c in range(ids_i.shape[1])
                 if ids_i[r, c] == 0 and ids_j[r, c] == 1][0]
    names = "xyt"
    for a in range(3):
        for b in range(3):
            hij = pa[f"hessian_pairs_{names[a]}{names[b]}"][row, slot]
            hii = pa[f"hessian_ii_{names[a]}{names[b]}"][row, slot]
            assert hij == pytest.approx(d2(a, 3 + b), rel=2e-5, abs=1e-6), ("ij", a, b)
            assert hii == pytest.approx(d2(a, b), rel=2e-5, abs=1e-6), ("ii", a, b)
    # the pair potential is half the pair energy (each side stores its half)
    assert pa["potential_pairs"][row, slot] == pytest.approx(e0 / 2, rel=1e-12)


# ---------------------------------------------------------------- jamming controller
def test_jam_controller_known_answer_on_a_square_lattice():
    """The pressure bisection of scripts/runArthurJamWithFire.cpp:64-106 on a 4 x 4 square
    lattice of unit disks (sigma = 1): every force vanishes by symmetry, so each FIRE
    minimisation leaves the lattice as it is, and with spacing a = L / 4 the pressure is
    P(a) = (1 - a) / a (four contacts each; zero for a >= 1).  With strain 0.05 from
    a = 1.1 and the window [0.05, 0.06] the trace is, by hand (L = 4 a):
      1  L 4.4                P = 0        low  -> L * 0.95
      2  L 4.18               P = 0        low  -> L * 0.95
      3  L 3.971              P = 0.0073   low  -> L * 0.95
      4  L 3.77245            P = 0.0603   high -> restore 3.971, L = mid(3.77245, 3.971)
      5  L 3.871725           P = 0.0331   low  -> L = mid(3.77245, 3.871725), drop L_high
      6  L 3.8220875          P = 0.0466   low  -> L * 0.95
      7  L 3.631083125        P = 0.1016   high -> restore 3.8220875, L = mid
      8  L 3.7265853125       P = 0.0734   high -> restore 3.8220875, L = mid
      9  L 3.77433640625      P = 0.0598   inside the window with 64 contacts: done
    so 8 compression steps, status 0, and the final state's closed-form P, PE, phi."""
    m, a0 = 4, 1.1
    n = m * m
    L0 = m * a0
    g = (np.arange(m) + 0.5) * a0
    x, y = np.meshgrid(g, g)
    s = orc_system(n, L0, x.ravel(), y.ravel(), np.full(n, 0.5), method=3, rc=1.5)
    prm = P.jam_params(pressure_target=0.05, window_rel=0.2, strain_step=0.05,
                       max_compression_steps=100, decompress_factor=1.01,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-12, max_steps=100))
    r = s.jam(prm)
    L3, L4 = 3.971, 3.971 * 0.95
    L5 = (L4 + L3) / 2
    L6 = (L4 + L5) / 2
    L7 = L6 * 0.95
    L8 = (L7 + L6) / 2
    L9 = (L8 + L6) / 2
    a = L9 / m
    assert (r.status, r.compression_steps, r.contacts) == (0, 8, 4 * n)
    assert r.box_length == pytest.approx(L9, rel=1e-12)
    assert r.pressure == pytest.approx((1 - a) / a, rel=1e-9)
    assert 0.05 <= r.pressure <= 0.06
    assert r.pe == pytest.approx(2 * n * 0.5 * (1 - a) ** 2, rel=1e-9)
    assert r.packing_fraction == pytest.approx(math.pi * 0.25 / (a * a), rel=1e-12)


def test_fire_force_floor_follows_coordinate_precision():
    """Why FIRE of the 65,536-particle rigid packing stops near |F|max = 2.5e-12: the
    forces cannot resolve below the precision of the coordinates.  The same 256-particle
    rigid packing relaxes to |F|max < 1e-12 in a few thousand steps; translated by
    10,000 (a periodic copy of the same system, coordinates ~200x coarser in the last
    place) it reaches the same minimum -- the energy agrees to 1e-9 relative -- but its
    largest force stays above 1e-11 however long FIRE runs, at the largest step."""
    st = W.bumpy_config(256, 0.80, seed=12345)
    prm = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-12,
                        max_steps=12000)
    out = []
    for shift in (0.0, 1e4):
        st.x, st.y = st.x + shift, st.y + shift
        s = W.make_bumpy_system(P.System, st, backend=oracle(), e_c=W.bumpy_contact_scale(st))
        r = s.minimize_fire(prm)
        out.append((r, s.reduce(P.RED_PE)))
        st.x, st.y = st.x - shift, st.y - shift
    (r0, e0), (r1, e1) = out
    assert r0.converged and math.sqrt(r0.fmax_sq) < 1e-12
    assert not r1.converged and math.sqrt(r1.fmax_sq) > 1e-11 and r1.dt == 0.1
    assert e1 == pytest.approx(e0, rel=1e-9)
