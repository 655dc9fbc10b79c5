"""Serving loop (include/glst_serve.h) on the GPU.

A request's answer must not depend on what it was batched with: alone, in a
mix, split over several steps by the device budget or by max_batch, and as a
repeat request of a resident system, it equals what a single-system handle
computes for it (FP64 bar 1e-12, mixed bar 1e-6: the batches differ only in
how the back-projection's direction ranges are split).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RCUT = 9.0
TOL = 1e-4


def _system(nw, seed):
    import glst_amd as g
    L = (nw / 0.0334) ** (1.0 / 3.0)
    return L, g.make_water_box(nw, L, seed=seed)


def _dev(h, torch):
    return [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]


def _outs(n, torch):
    return [torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
            for _ in range(4)]


def _single(L, h, ct, torch):
    import glst_amd as g
    f = g.GLSTForce(h[0].size, TOL, L, rcut=RCUT, ct=ct)
    f.calc_ener_force(*_dev(h, torch))
    out = [t.cpu().numpy() for t in f.get_ef()]
    f.close()
    return out


def _check(got, ref, bar):
    got = [t.cpu().numpy() for t in got]
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar


@pytest.mark.parametrize("ct,bar", [("double", 1e-12), ("float", 1e-6)])
def test_requests_alone_and_in_a_mix(cuda, ct, bar):
    import torch
    import glst_amd as g
    systems = [_system(nw, 40 + k)
               for k, nw in enumerate([150, 1000, 400, 2000, 700])]
    refs = [_single(L, h, ct, torch) for L, h in systems]
    # all in one batch
    srv = g.GLSTServer(TOL, 1 << 34, ct=ct)
    ins = [_dev(h, torch) for _, h in systems]
    outs = [_outs(h[0].size, torch) for _, h in systems]
    tickets = [srv.submit(L, RCUT, *i, *o)
               for (L, _), i, o in zip(systems, ins, outs)]
    assert srv.pending() == len(systems)
    assert srv.step() == len(systems)
    assert srv.pending() == 0
    assert srv.completed() == tickets
    for o, r in zip(outs, refs):
        _check(o, r, bar)
    # one request per step
    one = g.GLSTServer(TOL, 1 << 34, max_batch=1, ct=ct)
    outs1 = [_outs(h[0].size, torch) for _, h in systems]
    for (L, _), i, o in zip(systems, ins, outs1):
        one.submit(L, RCUT, *i, *o)
    answered = 0
    while one.pending():
        answered += one.step()
    assert answered == len(systems)
    for o, r in zip(outs1, refs):
        _check(o, r, bar)
    st = one.stats()
    assert st["completed"] == len(systems) and st["batches"] == len(systems)
    srv.close()
    one.close()


def test_budget_admission(cuda):
    import torch
    import glst_amd as g
    small = [_system(600, 60 + k) for k in range(4)]
    big = _system(4000, 70)
    need = max(g.batch_device_bytes([h[0].size], [L], RCUT, TOL)
               for L, h in small)
    # two batches may be in flight, each with half the budget: room for
    # two small requests per batch
    srv = g.GLSTServer(TOL, int(5.0 * need), ct="double")
    ins = [_dev(h, torch) for _, h in small]
    outs = [_outs(h[0].size, torch) for _, h in small]
    for (L, _), i, o in zip(small, ins, outs):
        srv.submit(L, RCUT, *i, *o)
    answered = 0
    while srv.pending():
        answered += srv.step()
    assert answered == 4
    assert srv.stats()["batches"] == 2
    for (L, h), o in zip(small, outs):
        _check(o, _single(L, h, "double", torch), 1e-12)
    # a request larger than a batch's share is refused at submission
    Lb, hb = big
    with pytest.raises(g.GLSTError) as e:
        srv.submit(Lb, RCUT, *_dev(hb, torch), *_outs(hb[0].size, torch))
    assert e.value.status == 4
    assert srv.stats()["rejected"] == 1
    assert srv.pending() == 0
    srv.close()


def test_resident_system_reuses_setup_and_batch(cuda):
    """An MD-style trajectory: the same system with perturbed positions,
    one request per step."""
    import torch
    import glst_amd as g
    L, h = _system(1200, 80)
    srv = g.GLSTServer(TOL, 1 << 34, ct="float")
    rng = np.random.default_rng(0)
    for step in range(3):
        pos = [np.clip(a + rng.normal(0.0, 0.01, a.size), 0.0, None)
               for a in h[:3]]
        hs = (pos[0], pos[1], pos[2], h[3])
        o = _outs(h[0].size, torch)
        i = _dev(hs, torch)  # borrowed until the request completes
        srv.submit(L, RCUT, *i, *o)
        assert srv.step() == 1
        _check(o, _single(L, hs, "float", torch), 1e-6)
    st = srv.stats()
    assert st["setups"] == 1 and st["setup_hits"] == 2
    assert st["reconfigures"] == 1 and st["batches"] == 3
    srv.close()


def test_batches_overlap_in_flight(cuda):
    """poll never blocks: with two slots, a second batch is launched while
    the first is still in flight, and every answer is still exact."""
    import torch
    import glst_amd as g
    systems = [_system(nw, 90 + k) for k, nw in enumerate([3000, 2500, 800])]
    srv = g.GLSTServer(TOL, 1 << 34, max_batch=1, ct="double")
    # inputs are borrowed until the request completes: keep them alive
    ins = [_dev(h, torch) for _, h in systems]
    outs = [_outs(h[0].size, torch) for _, h in systems]
    for (L, h), i, o in zip(systems, ins, outs):
        srv.submit(L, RCUT, *i, *o)
    answered = srv.poll()  # fills both slots, returns without waiting
    st = srv.stats()
    # both slots were filled before the first batch could finish (its
    # completion is queried right after its asynchronous launch); the third
    # request may already have followed if the first batch finished while
    # the second was being packed
    assert st["max_in_flight"] == 2 and st["batches"] >= 2
    while srv.pending():
        answered += srv.step()
    assert answered == 3 and srv.stats()["batches"] == 3
    for (L, h), o in zip(systems, outs):
        _check(o, _single(L, h, "double", torch), 1e-12)
    srv.close()


def test_new_shapes_stream_through_both_slots(cuda):
    """Requests of six grid shapes (2^3 to 7^3 cells), one per batch, both
    slots in flight: the shared shape tables grow with each new shape
    (reallocating while the other slot's batch may still run) and every
    answer stays exact."""
    import torch
    import glst_amd as g
    waters = [150, 900, 2500, 400, 3200, 1200, 6000, 260]
    systems = [_system(nw, 300 + k) for k, nw in enumerate(waters)]
    srv = g.GLSTServer(TOL, 1 << 34, max_batch=1, ct="float")
    ins = [_dev(h, torch) for _, h in systems]
    outs = [_outs(h[0].size, torch) for _, h in systems]
    answered = 0
    for (L, h), i, o in zip(systems, ins, outs):
        srv.submit(L, RCUT, *i, *o)
        answered += srv.poll()
    while srv.pending():
        answered += srv.step()
    assert answered == len(systems)
    st = srv.stats()
    assert st["batches"] == len(systems) and st["max_in_flight"] == 2
    for (L, h), o in zip(systems, outs):
        _check(o, _single(L, h, "float", torch), 1e-6)
    srv.close()
