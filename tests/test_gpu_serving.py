"""Serving loop: continuous batching of independent solves on one GPU.

* a served solve equals the same solve run alone through HpgemAPIDevice
  (same kernels, same arithmetic: bitwise);
* a solve's result is the same alone and in a mix of concurrently resident
  solves of other sizes, orders and operators (bitwise);
* admission respects the HBM budget, and a rejected request does not stop
  the server.
"""
import numpy as np
import pytest

from hpgem_amd import HpgemAPIDevice, Mesh, OP_ACOUSTIC_CENTRAL, HgdgError
from hpgem_amd.serving import Server, run_stream, serving_stream

pytestmark = pytest.mark.gpu


def one_solve(dim, kind, order, cells, seed, steps=6):
    rng = np.random.default_rng(seed)
    amp = rng.uniform(-1, 1, 3)
    k = rng.integers(1, 4, size=(3, dim)).astype(float)
    vel = [0.8, -0.6, 0.5][:dim]
    dt = 0.2 / max(cells) / (2 * order + 1)
    return dict(dim=dim, kind=kind, order=order, cells=cells, amp=amp, k=k, dt=dt, n_steps=steps, velocity=vel,
                wave_speed=1.0)


def submit(srv, r, keep=True):
    return srv.submit(r["dim"], r["kind"], r["order"], r["cells"], r["amp"], r["k"], r["dt"], r["n_steps"],
                      velocity=r["velocity"], wave_speed=r["wave_speed"], keep_state=keep)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("spec", [(2, "advection", 2, [8, 6]), (3, "advection", 3, [5, 4, 6]),
                                  (3, "acoustic", 2, [4, 5, 4]), (2, "acoustic", 4, [6, 6])])
def test_served_solve_equals_direct_solve(precision, spec):
    dim, kind, order, cells = spec
    r = one_solve(dim, kind, order, cells, seed=order)
    srv = Server(precision=precision, hbm_budget_gb=4, max_active=8)
    t = submit(srv, r)
    assert srv.wait(t, 120)
    served = srv.result(t, with_state=True)["state"]
    srv.close()
    mesh = Mesh.structured(cells)
    kw = {}
    if kind == "acoustic":
        kw = dict(kind=OP_ACOUSTIC_CENTRAL, c_inv=np.ones(mesh.n_elements))
    else:
        kw = dict(velocity=r["velocity"])
    dev = HpgemAPIDevice(mesh, order, precision=precision, **kw)
    dev.setInitialSolutionModes(r["amp"], r["k"], field=0 if kind == "acoustic" else -1)
    dev.computeOneTimeStep(0.0, r["dt"], r["n_steps"])
    direct = dev.getTimeIntegrationVector(0)
    dev.close()
    assert served.shape == direct.shape
    np.testing.assert_array_equal(served, direct)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_result_same_alone_and_in_mix(precision):
    target = one_solve(3, "advection", 3, [6, 5, 4], seed=99, steps=10)
    srv = Server(precision=precision, hbm_budget_gb=4, max_active=16)
    t = submit(srv, target)
    assert srv.wait(t, 120)
    alone = srv.result(t, with_state=True)["state"]
    srv.close()
    # the same solve arriving in the middle of a stream of other solves
    others = [one_solve(d, kd, p, c, seed=s, steps=st) for s, (d, kd, p, c, st) in enumerate([
        (2, "advection", 1, [20, 16], 30), (3, "acoustic", 2, [5, 5, 5], 12), (3, "advection", 3, [8, 8, 8], 8),
        (2, "acoustic", 4, [10, 12], 20), (3, "advection", 2, [6, 7, 5], 15), (3, "advection", 3, [4, 4, 4], 25)])]
    srv = Server(precision=precision, hbm_budget_gb=4, max_active=16)
    tickets = [submit(srv, o, keep=False) for o in others[:3]]
    t = submit(srv, target)
    tickets += [submit(srv, o, keep=False) for o in others[3:]]
    assert srv.wait(-1, 300)
    mixed_in = srv.result(t, with_state=True)["state"]
    for tk in tickets:
        assert srv.result(tk)["n_dof"] > 0
    srv.close()
    np.testing.assert_array_equal(mixed_in, alone)


def test_admission_budget_and_stream_statistics():
    reqs = serving_stream(n_requests=24, mean_interarrival_s=0.002, side_min=4, side_max=12, steps=8, seed=3)
    # a budget that holds only a few solves at a time
    srv = Server(precision="fp64", hbm_budget_gb=0.02, max_active=4)
    tickets = run_stream(srv, reqs)
    st = srv.stats()
    print(st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0
    big = max(srv.result(t)["n_dof"] for t in tickets)
    assert st["peak_bytes"] <= max(0.02 * (1 << 30), 200 * big)
    # every request takes 8 steps; the resident set steps together
    assert 8 <= st["ticks"] <= 8 * len(reqs)
    assert 1 <= st["restages"] <= 2 * len(reqs)
    assert 0 < st["restage_s"] <= st["tick_s"]
    srv.close()


def test_bad_request_is_rejected_and_server_keeps_serving():
    srv = Server(precision="fp64", hbm_budget_gb=1, max_active=4)
    with pytest.raises(HgdgError):
        srv.submit(3, "advection", 9, [4, 4, 4], [1.0], [[1, 1, 1]], 0.01, 2)  # order out of range
    r = one_solve(2, "advection", 2, [6, 6], seed=1)
    t = submit(srv, r, keep=False)
    assert srv.wait(t, 60)
    assert srv.result(t)["n_dof"] == 36 * 9
    srv.close()


def test_profiled_serving_times_each_family_and_changes_no_result():
    """With profiling on, every RK step runs as separate launches between HIP
    events: the same launches as the step graph, so every result is bitwise
    the same, and each (kind, dim, order) family of the resident mix reports
    its kernel time (4 stages per step)."""
    specs = [(2, "advection", 2, [8, 6], 6), (3, "advection", 3, [5, 4, 6], 4), (3, "acoustic", 2, [4, 5, 4], 5)]
    reqs = [one_solve(d, kd, p, c, seed=i, steps=st) for i, (d, kd, p, c, st) in enumerate(specs)]
    states = []
    for profile in (False, True):
        srv = Server(precision="fp64", hbm_budget_gb=4, max_active=8)
        if profile:
            srv.set_profiling(True)
        tickets = [submit(srv, r) for r in reqs]
        assert srv.wait(-1, 120)
        states.append([srv.result(t, with_state=True)["state"] for t in tickets])
        if profile:
            fams, upd = srv.family_times()
            print(fams, upd)
            assert set(fams) == {("advection", 2, 2), ("advection", 3, 3), ("acoustic", 3, 2)}
            assert all(ms > 0 and n % 4 == 0 for ms, n in fams.values()) and upd == 0
            # every step of every resident family is timed: at least the longest solve's steps
            assert max(n for _, n in fams.values()) >= 4 * max(r["n_steps"] for r in reqs)
        else:
            assert srv.family_times()[0] == {}
        srv.close()
    for a, b in zip(*states):
        np.testing.assert_array_equal(a, b)
