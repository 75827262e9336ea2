"""Serving (continuous batching of I-V points) on the GPU, through the C-ABI:
a point's result is the same alone and in a mix that it joins mid-way (bit
for bit: probe counts, particle counts, window currents, steps); admission
respects the device-memory and resident-particle budgets and the admission
order; statuses and step accounting."""
import sys

import pytest

pytestmark = pytest.mark.gpu

cp = pytest.importorskip("cupic_amd")

SAME = ("status", "error", "steps", "windows", "current", "rel_change", "total_probe",
        "particles")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def point(ncx, ncy, ppc, phi_p, run_id):
    return cp.params(ncx=ncx, ncy=ncy, ppc=ppc, phi_p=phi_p, run_id=run_id)


def cfg(**kw):
    base = dict(round_steps=20, window_steps=100, max_steps=400, rel_change=0.005,
                jacobi_max_sweeps=200, jacobi_check_every=50)
    base.update(kw)
    return cp.server_config(**base)


def drain(srv, ids, limit=1000):
    for _ in range(limit):
        active, queued = srv.poll()
        if active == 0 and queued == 0:
            break
    res = [srv.result(i) for i in ids]
    assert all(r["status"] in (cp.POINT_CONVERGED, cp.POINT_CAPPED) for r in res), res
    return res


@pytest.fixture(scope="module")
def dev():
    return cp.init_dev(0)


def test_point_same_alone_and_in_mix(dev):
    a = point(32, 16, 8, -3.0, 7)
    say("alone")
    srv = cp.Server(cfg())
    ia = srv.submit(a)
    alone = drain(srv, [ia])[0]
    srv.close()
    say("  ", {k: alone[k] for k in SAME})
    say("mix")
    srv = cp.Server(cfg())
    others = [srv.submit(point(48, 24, 10, -8.0, 1)), srv.submit(point(40, 12, 6, 1.0, 2))]
    srv.poll()
    srv.poll()
    ia = srv.submit(a)                      # joins a batch already running
    others.append(srv.submit(point(24, 20, 12, -5.0, 3)))
    res = drain(srv, [ia] + others)
    st = srv.stats()
    srv.close()
    say("  rebuilds", st["rebuilds"], "rounds", st["rounds"])
    assert st["rebuilds"] >= 2
    mixed = res[0]
    for k in SAME:
        assert mixed[k] == alone[k], (k, mixed[k], alone[k])
    assert mixed["steps"] % 20 == 0 and mixed["windows"] == mixed["steps"] // 100


def test_admission_respects_budget(dev):
    ps = [point(32, 16, 8, -2.0 - k, 10 + k) for k in range(3)]
    f = cp.footprint(ps[0])
    srv = cp.Server(cfg(hbm_budget=int(2.5 * f), max_steps=200))
    ids = [srv.submit(p) for p in ps]
    active, queued = srv.poll()
    assert (active, queued) == (2, 1)
    assert srv.stats()["hbm_used"] <= 2.5 * f
    res = drain(srv, ids)
    first_done = min(r["t_done"] for r in res[:2])
    assert res[2]["t_admit"] >= first_done
    for r in res:
        assert r["steps"] == 200 or r["status"] == cp.POINT_CONVERGED
    srv.close()


def test_particle_budget_and_admission_order(dev):
    """With a resident-particle budget that holds one point at a time,
    points run in turn: in arrival order (FCFS), or fewest particles first;
    a point larger than the budget is still admitted into an empty server."""
    sizes = [(48, 24, 10), (32, 16, 4), (40, 20, 8)]
    ps = [point(nx, ny, ppc, -3.0, 20 + k) for k, (nx, ny, ppc) in enumerate(sizes)]
    npart = [2 * p.per_row * p.ncy for p in ps]
    for policy, order in ((cp.SERVE_FCFS, [0, 1, 2]),
                          (cp.SERVE_SMALLEST_FIRST, sorted(range(3), key=lambda k: npart[k]))):
        budget = min(npart[i] + npart[j] for i in range(3) for j in range(3) if i != j) - 1
        srv = cp.Server(cfg(max_particles=budget, max_steps=100, policy=policy))
        ids = [srv.submit(p) for p in ps]
        active, queued = srv.poll()
        assert (active, queued) == (1, 2)
        res = drain(srv, ids)
        srv.close()
        admitted = sorted(range(3), key=lambda k: res[k]["t_admit"])
        say("policy", policy, "admission order", admitted, "particles", npart)
        assert admitted == order
        for a, b in zip(order, order[1:]):
            assert res[b]["t_admit"] >= res[a]["t_done"]


def test_request_over_budget_is_rejected(dev):
    p = point(32, 16, 8, -2.0, 1)
    srv = cp.Server(cfg(hbm_budget=cp.footprint(p) // 2))
    with pytest.raises(cp.CupicError):
        srv.submit(p)
    srv.close()
