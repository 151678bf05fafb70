"""GPU tests of rigid bumpy systems in the batched engine (dpm_batch_create_mixed): every
rigid system's result -- jam outcome, centres, angles and vertex positions -- is
bit-identical to the same system jammed (or minimised) alone by dpm_jam on a dpm_system with
the Verlet method and the rigid neighbour rules (RigidBumpy::setNeighborSize,
rigid_bumpy.cu:1278-1300), and to the oracle's jam of that system; disk systems in the same
batch keep their own results."""
import numpy as np
import pytest

from _helpers import oracle, pkg, workloads

P = pkg()
W = workloads()
pytestmark = pytest.mark.gpu

FIELDS = ("status", "compression_steps", "fire_steps_total", "pressure", "packing_fraction",
          "pe", "box_length", "contacts")


def say(*a):
    print("[gpu-case]", *a, flush=True)


def params(pressure=1e-3, window=0.2, strain=1e-2, max_steps=800, csteps=200):
    return P.jam_params(pressure_target=pressure, window_rel=window, strain_step=strain,
                        max_compression_steps=csteps,
                        fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=1e-8, max_steps=max_steps))


def plain_rigid(st, run, backend=None, rotation=True):
    """The system alone on a dpm_system (or the oracle): Verlet method, the rigid rules."""
    s = W.make_bumpy_system(P.System, st, backend=backend, e_c=W.bumpy_contact_scale(st),
                            rotation=rotation, method=P.NBR_VERLET)
    r = run(s)
    out = [s.download(f) for f in (P.F_X, P.F_Y, P.F_ANGLE, P.F_VERTEX_X, P.F_VERTEX_Y)]
    return r, out


def plain_disk(st, prm):
    s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2))
    r = s.jam(prm)
    return r, [s.download(P.F_X), s.download(P.F_Y)]


def mixed_batch(sts, rotation=None):
    bumpy = [hasattr(st, "angle") for st in sts]
    b = P.Batch([st.n for st in sts], [st.nv if bp else 0 for st, bp in zip(sts, bumpy)])
    b.config(1.0, 2.0, 1.7, 3.0 * 0.15 * 0.15)
    for k, st in enumerate(sts):
        if bumpy[k]:
            rot = True if rotation is None else rotation[k]
            b.set_rigid_system(k, st.x, st.y, st.rad, st.mass, st.angle, st.nverts, st.rv,
                               st.side, e_c=W.bumpy_contact_scale(st), rotation=rot)
        else:
            b.set_system(k, st.x, st.y, st.rad, st.mass, st.side)
    return b, bumpy


def state_of(b, k, bumpy):
    if bumpy:
        x, y, a, vx, vy, side = b.get_rigid_system(k)
        return [x, y, a, vx, vy], side
    x, y, side = b.get_system(k)
    return [x, y], side


def same(got, ref, k):
    for f in FIELDS:
        assert getattr(got, f) == getattr(ref, f), (k, f, getattr(got, f), getattr(ref, f))


def check_arrays(a, b, k):
    for j, (u, v) in enumerate(zip(a, b)):
        assert np.array_equal(u, v), (k, j, np.max(np.abs(u - v)))


def test_geometry_set_in_the_arena():
    """Vertex rings, from the angles, equal dpm_bumpy_setup's, bit for bit."""
    st = W.bumpy_config(40, 0.72, seed=21)
    b, bumpy = mixed_batch([W.disk_config(64, 0.8, seed=2), st])
    got, side = state_of(b, 1, True)
    s = W.make_bumpy_system(P.System, st, e_c=W.bumpy_contact_scale(st), method=P.NBR_VERLET)
    for g, f in zip(got, (P.F_X, P.F_Y, P.F_ANGLE, P.F_VERTEX_X, P.F_VERTEX_Y)):
        assert np.array_equal(g, s.download(f))
    assert side == st.side
    b.close()


def test_rigid_systems_in_the_batch_equal_each_alone():
    """Three rigid systems (one without rotation) beside two disk systems, one of 200
    particles whose rows come from the cell grid; each equals its own jam alone, and the
    two smallest rigid systems equal the oracle's jam."""
    sts = [W.bumpy_config(24, 0.70, seed=3), W.disk_config(64, 0.80, seed=1),
           W.bumpy_config(40, 0.72, seed=5), W.disk_config(130, 0.80, seed=2),
           W.bumpy_config(200, 0.72, seed=7)]
    rotation = [True, None, False, None, True]
    prm = params()
    b, bumpy = mixed_batch(sts, rotation)
    res, steps = b.jam(prm)
    say("mixed batch steps", steps)
    for k, st in enumerate(sts):
        got, side = state_of(b, k, bumpy[k])
        if bumpy[k]:
            r, ref = plain_rigid(st, lambda s: s.jam(prm), rotation=rotation[k])
        else:
            r, ref = plain_disk(st, prm)
        say("system", k, "rigid" if bumpy[k] else "disk", st.n, r.status, r.compression_steps,
            r.fire_steps_total)
        same(res[k], r, k)
        check_arrays(got, ref, k)
        assert side == r.box_length
        if bumpy[k] and st.n <= 40:
            ro, oref = plain_rigid(st, lambda s: s.jam(prm), backend=oracle(),
                                   rotation=rotation[k])
            same(res[k], ro, k)
            check_arrays(got, oref, k)
    b.close()


def test_rigid_bisection_with_restores():
    """A coarse strain step and a narrow pressure window: the controller overshoots,
    restores its snapshot (a full list rebuild at the restored state) and bisects; every
    system still equals its jam alone."""
    sts = [W.bumpy_config(30, 0.72, seed=11), W.bumpy_config(48, 0.70, seed=12)]
    prm = params(pressure=1e-6, window=0.01, strain=0.2, csteps=120)
    b, bumpy = mixed_batch(sts)
    res, steps = b.jam(prm)
    for k, st in enumerate(sts):
        got, side = state_of(b, k, True)
        r, ref = plain_rigid(st, lambda s: s.jam(prm))
        say("bisection", k, st.n, r.status, r.compression_steps, r.fire_steps_total)
        assert r.compression_steps >= 4  # 20 % steps from below the window overshoot it
        same(res[k], r, k)
        check_arrays(got, ref, k)
        ro, oref = plain_rigid(st, lambda s: s.jam(prm), backend=oracle())
        same(res[k], ro, k)
        check_arrays(got, oref, k)
    b.close()


def test_rigid_minimization_in_the_batch():
    """dpm_batch_minimize on rigid systems: steps, energy, pressure, contacts, packing
    fraction and the final state equal dpm_minimize_fire alone and the reductions."""
    sts = [W.bumpy_config(36, 0.78, seed=31), W.bumpy_config(150, 0.76, seed=32),
           W.disk_config(100, 0.84, seed=33)]
    fp = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-9,
                       max_steps=3000)
    b, bumpy = mixed_batch(sts)
    res, steps = b.minimize(fp)
    for k, st in enumerate(sts):
        got, side = state_of(b, k, bumpy[k])
        if not bumpy[k]:
            continue
        for be in (None, oracle()) if st.n <= 40 else (None,):
            s = W.make_bumpy_system(P.System, st, backend=be, e_c=W.bumpy_contact_scale(st),
                                    method=P.NBR_VERLET)
            fr = s.minimize_fire(fp)
            r = res[k]
            say("rigid minimisation", st.n, fr.steps, fr.converged, "oracle" if be else "gpu")
            assert r.fire_steps_total == fr.steps and r.status == (0 if fr.converged else 1)
            assert r.pe == s.reduce(P.RED_PE) and r.pressure == s.reduce(P.RED_PRESSURE)
            assert r.contacts == s.reduce(P.RED_CONTACTS)
            assert r.packing_fraction == s.reduce(P.RED_PACKING_FRACTION)
            ref = [s.download(f) for f in (P.F_X, P.F_Y, P.F_ANGLE, P.F_VERTEX_X,
                                            P.F_VERTEX_Y)]
            check_arrays(got, ref, k)
    b.close()


def test_bad_rigid_set_up_fails_loudly():
    st = W.bumpy_config(16, 0.7, seed=41)
    b = P.Batch([16, 16], [st.nv, 0])
    with pytest.raises(P.DpmError):  # a disk set on a rigid system
        b.set_system(0, st.x, st.y, st.rad, st.mass, st.side)
    with pytest.raises(P.DpmError):  # a rigid set on a disk system
        b.set_rigid_system(1, st.x, st.y, st.rad, st.mass, st.angle, st.nverts, st.rv, st.side)
    with pytest.raises(P.DpmError):  # vertex counts that do not sum to the system's
        b.set_rigid_system(0, st.x, st.y, st.rad, st.mass, st.angle, st.nverts + 1, st.rv,
                           st.side)
    b.close()


def test_rigid_force_pass_timing_and_vertex_rows():
    """In a batch of rigid systems only, DPM_OPT_TIMING times the rigid force pass (a
    timed launch per launched batched step; launches go by whole chunks, so at least one
    per executed step; the same results), and the stored vertex-row entries equal the
    listed vertex pairs of the same minimisation alone."""
    st = W.bumpy_config(150, 0.76, seed=32)
    fp = P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0, fmax_target=1e-9,
                       max_steps=3000)
    b, _ = mixed_batch([st])
    b.set_option(P.OPT_TIMING, 1)
    res, steps = b.minimize(fp)
    ms, launches = b.kernel_time()
    entries = b.vertex_row_entries()
    got, _ = state_of(b, 0, True)
    b.close()
    say("timed rigid passes", launches, ms, "batched steps", steps, "vertex rows", entries)
    assert launches >= steps > 0 and ms > 0
    r, ref = plain_rigid(st, lambda s: s.minimize_fire(fp))
    check_arrays(got, ref, 0)
    s = W.make_bumpy_system(P.System, st, e_c=W.bumpy_contact_scale(st), method=P.NBR_VERLET)
    s.minimize_fire(fp)
    assert entries == len(s.vertex_pairs()[0])
    s.close()
    d, _ = mixed_batch([W.disk_config(64, 0.8, seed=2)])
    assert d.vertex_row_entries() == 0
    d.close()
