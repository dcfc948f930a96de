"""Relation-bearing parity at every magnitude, in both device fields.

Each item of tests/golden/planted.json has a relation planted at a known prime (or a composite
that must yield none); see tests/golden/make_planted.py.  Every item runs through the GPU worker
in every setup mode (classes off: plain; classes on under class policies 0, 1 and 2) and, below
2^50, in both fields (FP64 by default, Montgomery forced).  Per run:
  * the relations equal the oracle's over the item's whole P window (sorted p, M, q);
  * the planted relation is among them, with the fixture's M and q (or the planted number is not);
  * the device itself decided the planted prime, M included (the exact records of
    mps_gpu_sieve_relations, not the host's re-check), so the BSGS kernels of both fields are
    shown to find the smallest solution;
  * for a composite, a record was written for it and the device Miller-Rabin test removed it
    (plain mode, where no earlier stage can drop it); in class and order mode it is dropped or
    left to the host.
The relation equation is Gpu.hpp:202-206's; M is defined by run_bsgs, Math.hpp:377-486."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PLANTED = json.load(open(os.path.join(HERE, "golden", "planted.json")))["items"]

MODES = {"plain": (False, 1), "policy0": (True, 0), "policy1": (True, 1), "policy2": (True, 2)}

CASES = [(it, mode, field) for it in PLANTED for mode in MODES for field in ([-1, 0] if 1 in it["fields"] else [-1])]

_want = {}


def oracle_relations(oracle, it):
    key = it["tag"]
    if key not in _want:
        b, c, P0, P1, N0, N1, mod, res = it["item"]
        _want[key], _ = oracle.process_range(oracle.make_item(*it["item"]), P0, P1, 1 << 17, 16)
    return _want[key]


@pytest.mark.parametrize("it,mode,field", CASES,
                         ids=[f"{it['tag']}-{mode}-f{field}" for it, mode, field in CASES])
def test_planted(gpu, mps, oracle, it, mode, field):
    b, c, P0, P1, N0, N1, mod, res = it["item"]
    classes, policy = MODES[mode]
    print(f"planted {it['tag']} {mode} field {field}: {it['item']}", flush=True)
    gpu.set_classes(classes)
    gpu.set_class_policy(policy)
    gpu.configure(field, 0)
    try:
        gpu.reset_stats()
        recs = gpu.sieve_relations(P0, P1, mps.work_item(*it["item"]))
        flagged = gpu.stats()["flagged"]
        rels, st = gpu.run_item(mps.work_item(*it["item"]), chunk_ints=1 << 17)
    finally:
        gpu.set_classes(True)
        gpu.set_class_policy(1)
        gpu.configure(-1, 0)
    assert recs is not None, gpu.last_error()
    assert st.declined == 0
    want = oracle_relations(oracle, it)
    assert rels == want
    p = it["p"]
    dev = {q: M for q, M in recs}
    if it["relation"] is not None:
        rel = tuple(it["relation"])
        assert rel in rels
        assert dev.get(p) == rel[1], f"device record for the planted prime: {dev.get(p)}"
    else:
        assert p not in {r[0] for r in rels}
        if it.get("composite") and mode != "plain":
            # the BSGS base b^qp (class mode) or b^L (order mode) may have a short cycle modulo a
            # composite, which leaves the number to the host's exact check
            assert dev.get(p, mps.HOST_DECIDES) == mps.HOST_DECIDES
        else:
            assert p not in dev
    if it.get("composite") and mode == "plain":
        # base b itself, of order above the N range: the BSGS found the planted solution and
        # wrote a record; the device Miller-Rabin test zeroed it
        assert flagged >= len(recs) + 1
