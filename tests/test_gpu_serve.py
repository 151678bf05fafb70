"""GPU tests of the serving loop (dpm_server_*): requests that arrive while others run
are admitted as memory frees up, and every result equals the request's system jammed
(or minimised) alone, bit for bit -- by dpm_jam on the GPU (Verlet method; rigid bumpy
systems with the rigid neighbour rules) and by the oracle's jam (orc_jam)."""
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


def jam(pressure=1e-3, max_steps=1200):
    return P.jam_params(pressure_target=pressure, window_rel=0.2, strain_step=5e-3,
                        max_compression_steps=300,
                        fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=1e-10, max_steps=max_steps))


def alone_disk(st, prm, backend=None):
    s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2), backend=backend)
    r = s.jam(prm)
    return r, s.download(P.F_X), s.download(P.F_Y)


def alone_bumpy(st, prm, e_c, backend=None):
    s = W.make_bumpy_system(P.System, st, e_c=e_c, backend=backend, method=P.NBR_VERLET)
    r = s.jam(prm)
    return r, s.download(P.F_X), s.download(P.F_Y)


def run_until_done(sv, ids, limit=200000):
    for _ in range(limit):
        done, inflight, queued = sv.progress(32)
        if all(sv.result(i)["state"] == 2 for i in ids):
            return
    raise AssertionError("requests did not finish")


def check(got, ref):
    r, x, y = ref
    for f in FIELDS:
        assert getattr(got["result"], f) == getattr(r, f), f
    assert np.array_equal(got["x"], x) and np.array_equal(got["y"], y)


def test_disk_requests_arriving_during_a_batch():
    sts = [W.disk_config(n, phi, seed=s) for n, phi, s in
           ((64, 0.80, 11), (130, 0.81, 12), (90, 0.79, 13), (200, 0.80, 14), (75, 0.82, 15))]
    prms = [jam(1e-3), jam(5e-4), jam(2e-3), jam(1e-3), jam(8e-4)]
    sv = P.Server(max_systems=8, arena_slots=1024, budget_bytes=1 << 30, vertex_slots=0)
    ids = [sv.submit(st.x, st.y, st.rad, st.mass, st.side, p) for st, p in
           zip(sts[:2], prms[:2])]
    for _ in range(20):  # the first two run a while, then the others arrive
        sv.progress(32)
    ids += [sv.submit(st.x, st.y, st.rad, st.mass, st.side, p) for st, p in
            zip(sts[2:], prms[2:])]
    run_until_done(sv, ids)
    for k, (st, p) in enumerate(zip(sts, prms)):
        got = sv.result(ids[k])
        say("request", k, got["result"].status, got["result"].fire_steps_total, got["latency_s"])
        check(got, alone_disk(st, p))
        check(got, alone_disk(st, p, backend=oracle()))
    sv.close()


def test_large_disk_requests_with_cell_grid_rows():
    # requests of 128+ particles get their rows from the cell grid; the 4,000-particle
    # request waits until the first two leave, then takes their arena range (the grid
    # workgroups are listed again at each admission)
    prm = P.jam_params(pressure_target=1e-3, window_rel=0.5, strain_step=1e-2,
                       max_compression_steps=30,
                       fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                          fmax_target=1e-8, max_steps=300))
    sts = [W.disk_config(n, 0.80, seed=s) for n, s in ((2100, 31), (300, 32), (4000, 33))]
    sv = P.Server(max_systems=4, arena_slots=4224, budget_bytes=1 << 30, vertex_slots=0)
    ids = [sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm) for st in sts]
    run_until_done(sv, ids)
    for k, st in enumerate(sts):
        got = sv.result(ids[k])
        say("large request", k, st.n, got["result"].status, got["result"].fire_steps_total)
        check(got, alone_disk(st, prm))
    sv.close()


def test_admission_waits_for_space():
    # an arena of 256 slots holds one 200-particle system at a time: requests queue
    sts = [W.disk_config(n, 0.80, seed=20 + n) for n in (200, 150, 180)]
    prm = jam(max_steps=600)
    sv = P.Server(max_systems=4, arena_slots=256, budget_bytes=1 << 30, vertex_slots=0)
    ids = [sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm) for st in sts]
    done, inflight, queued = sv.progress(1)
    assert inflight == 1 and queued == 2
    run_until_done(sv, ids)
    for k, st in enumerate(sts):
        check(sv.result(ids[k]), alone_disk(st, prm))
    sv.close()


def test_rigid_bumpy_request_next_to_disks():
    b = W.bumpy_config(24, 0.7, seed=3)
    e_c = W.bumpy_contact_scale(b)
    d = W.disk_config(100, 0.8, seed=4)
    prm_b = jam(1e-3, max_steps=400)
    prm_d = jam(1e-3)
    sv = P.Server(max_systems=4, arena_slots=512, budget_bytes=1 << 30, vertex_slots=4096)
    ib = sv.submit(b.x, b.y, b.rad, b.mass, b.side, prm_b, e_c=e_c, angle=b.angle,
                   vertices_per_particle=b.nverts, vertex_radius=b.rv)
    idd = sv.submit(d.x, d.y, d.rad, d.mass, d.side, prm_d)
    run_until_done(sv, [ib, idd])
    check(sv.result(idd), alone_disk(d, prm_d))
    check(sv.result(ib), alone_bumpy(b, prm_b, e_c))
    check(sv.result(ib), alone_bumpy(b, prm_b, e_c, backend=oracle()))
    sv.close()


def test_bad_requests_fail_loudly():
    sv = P.Server(max_systems=2, arena_slots=128, budget_bytes=1 << 20, vertex_slots=0)
    d = W.disk_config(200, 0.8, seed=1)
    with pytest.raises(P.DpmError):  # larger than the arena
        sv.submit(d.x, d.y, d.rad, d.mass, d.side, jam())
    with pytest.raises(P.DpmError):  # this server takes no bumpy requests
        b = W.bumpy_config(8, 0.5, seed=1)
        sv.submit(b.x, b.y, b.rad, b.mass, b.side, jam(), angle=b.angle,
                  vertices_per_particle=b.nverts, vertex_radius=b.rv)
    sv.close()


def relaxed(pressure, max_steps=2000):
    return P.jam_params(pressure_target=pressure, window_rel=0.5, strain_step=1e-2,
                        max_compression_steps=100000,
                        fire=P.fire_params(avg_pe_target=0.0, avg_pe_diff_target=0.0,
                                           fmax_target=1e-6, max_steps=max_steps))


def submit_request(sv, q, prm, mode=P.REQ_JAM):
    st = q.state if hasattr(q, "state") else q
    if hasattr(st, "angle"):
        return sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm,
                         e_c=W.bumpy_contact_scale(st), angle=st.angle,
                         vertices_per_particle=st.nverts, vertex_radius=st.rv, mode=mode)
    return sv.submit(st.x, st.y, st.rad, st.mass, st.side, prm, mode=mode)


def test_poisson_70_30_mix_alone_and_in_the_mix():
    """The serving configuration at a reduced count: ten seeded Poisson arrivals, 70 %
    disks with N log-uniform in [128, 16384] and 30 % rigid bumpy particles with N
    log-uniform in [128, 4096], target pressure log-uniform in [1e-7, 1e-3] (relaxed FIRE
    tolerance, strain 1e-2).  Requests are submitted as they arrive while the server runs;
    every result equals the request jammed alone on the GPU, and the three smallest disk
    requests equal the oracle's jam."""
    import time
    reqs = W.serve_requests(10, 50.0, 8)
    assert sum(q.bumpy for q in reqs) == 3
    sv = P.Server(max_systems=64, arena_slots=1 << 15, budget_bytes=1 << 32,
                  vertex_slots=1 << 19)
    ids = [None] * len(reqs)
    t0, nxt = time.perf_counter(), 0
    while nxt < len(reqs) or not all(sv.result(i)["state"] == 2 for i in ids):
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            ids[nxt] = submit_request(sv, reqs[nxt], relaxed(reqs[nxt].pressure))
            nxt += 1
        sv.progress(32)
        assert time.perf_counter() - t0 < 600, "mix did not finish"
    disks = sorted((q.state.n, k) for k, q in enumerate(reqs) if not q.bumpy)
    for k, q in enumerate(reqs):
        got = sv.result(ids[k])
        prm = relaxed(q.pressure)
        say("mix request", k, "bumpy" if q.bumpy else "disk", q.state.n, q.pressure,
            got["result"].status, got["result"].fire_steps_total, got["latency_s"])
        if q.bumpy:
            check(got, alone_bumpy(q.state, prm, W.bumpy_contact_scale(q.state)))
        else:
            check(got, alone_disk(q.state, prm))
        sv.release(ids[k])
    for n, k in disks[:3]:
        got_ref = alone_disk(reqs[k].state, relaxed(reqs[k].pressure))
        check({"result": got_ref[0], "x": got_ref[1], "y": got_ref[2]},
              alone_disk(reqs[k].state, relaxed(reqs[k].pressure), backend=oracle()))
    with pytest.raises(P.DpmError):  # released requests are gone
        sv.result(ids[0])
    sv.close()


def test_minimization_requests():
    """DPM_REQ_MINIMIZE: minimizeFire alone (disk and rigid bumpy systems in the batch); the
    result equals dpm_minimize_fire on the system alone followed by the reductions, and the
    oracle's; a request cannot be released while it runs."""
    fire_prm = relaxed(1e-3, max_steps=3000)
    d1, d2 = W.disk_config(300, 0.86, seed=61), W.disk_config(1500, 0.84, seed=62)
    b = W.bumpy_config(40, 0.75, seed=63)
    sv = P.Server(max_systems=8, arena_slots=4096, budget_bytes=1 << 30, vertex_slots=4096)
    ids = [submit_request(sv, st, fire_prm, mode=P.REQ_MINIMIZE) for st in (d1, d2, b)]
    with pytest.raises(P.DpmError):
        sv.release(ids[0])  # queued
    run_until_done(sv, ids)
    for st, i in zip((d1, d2, b), ids):
        got = sv.result(i)
        r = got["result"]
        say("minimisation request", st.n, r.status, r.fire_steps_total)
        bumpy = hasattr(st, "angle")
        for be in (None, oracle()):
            if bumpy:
                s = W.make_bumpy_system(P.System, st, backend=be, e_c=W.bumpy_contact_scale(st),
                                        method=P.NBR_VERLET)
            else:
                s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2),
                                       backend=be)
            fr = s.minimize_fire(fire_prm.fire)
            assert r.fire_steps_total == fr.steps and r.status == (0 if fr.converged else 1)
            assert r.pe == s.reduce(P.RED_PE) and r.pressure == s.reduce(P.RED_PRESSURE)
            assert r.contacts == s.reduce(P.RED_CONTACTS) and r.compression_steps == 0
            assert r.packing_fraction == s.reduce(P.RED_PACKING_FRACTION)
            assert np.array_equal(got["x"], s.download(P.F_X))
            assert np.array_equal(got["y"], s.download(P.F_Y))
        sv.release(i)
    sv.close()
