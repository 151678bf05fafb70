"""GPU tests of the serving loop (dpm_server_*): requests that arrive while others run
are admitted as memory frees up, and every result equals the request's system jammed
alone (dpm_jam), bit for bit."""
import numpy as np
import pytest

from _helpers import pkg, workloads

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


def alone_disk(st, prm):
    s = W.make_disk_system(P.System, st, W.neighbor_spec(st.side, method=2))
    r = s.jam(prm)
    return r, s.download(P.F_X), s.download(P.F_Y)


def alone_bumpy(st, prm, e_c):
    s = W.make_bumpy_system(P.System, st, e_c=e_c)
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
    sv = P.Server(max_systems=8, arena_slots=1024, budget_bytes=1 << 30, bumpy_workers=0)
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
    sv = P.Server(max_systems=4, arena_slots=4224, budget_bytes=1 << 30, bumpy_workers=0)
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
    sv = P.Server(max_systems=4, arena_slots=256, budget_bytes=1 << 30, bumpy_workers=0)
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
    sv = P.Server(max_systems=4, arena_slots=512, budget_bytes=1 << 30, bumpy_workers=1)
    ib = sv.submit(b.x, b.y, b.rad, b.mass, b.side, prm_b, e_c=e_c, angle=b.angle,
                   vertices_per_particle=b.nverts, vertex_radius=b.rv)
    idd = sv.submit(d.x, d.y, d.rad, d.mass, d.side, prm_d)
    run_until_done(sv, [ib, idd])
    check(sv.result(idd), alone_disk(d, prm_d))
    check(sv.result(ib), alone_bumpy(b, prm_b, e_c))
    sv.close()


def test_bad_requests_fail_loudly():
    sv = P.Server(max_systems=2, arena_slots=128, budget_bytes=1 << 20, bumpy_workers=0)
    d = W.disk_config(200, 0.8, seed=1)
    with pytest.raises(P.DpmError):  # larger than the arena
        sv.submit(d.x, d.y, d.rad, d.mass, d.side, jam())
    with pytest.raises(P.DpmError):  # this server takes no bumpy requests
        b = W.bumpy_config(8, 0.5, seed=1)
        sv.submit(b.x, b.y, b.rad, b.mass, b.side, jam(), angle=b.angle,
                  vertices_per_particle=b.nverts, vertex_radius=b.rv)
    sv.close()
