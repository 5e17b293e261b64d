"""Serving with continuous batching: a stream of equiv and verify requests submitted while others
run. Each result equals the same request run alone (gpu_enum_run), fault-free circuit/rewrite
pairs have no mismatch, every counterexample really separates the two netlists, and small requests
equal the oracle's enumeration on the reference machine."""
import time

import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tesseract_amd import netlists as N

pytestmark = pytest.mark.gpu


def small_request(seed, k, n_in, gates):
    rng = np.random.default_rng([seed, k])
    c = N.circuit(n_in, int(rng.integers(1, 9)), gates, rng)
    r = N.rewrite(c, rng)
    faulty = k % 3 == 0
    if faulty:
        r = N.fault(r, rng)
    return c, r, faulty


def test_server_mix_equals_alone(engine):
    reqs = []
    for k in range(18):
        n_in = [10, 14, 22, 27][k % 4]
        gates = [60, 300, 1500][k % 3]
        c, r, faulty = small_request(7, k, n_in, gates)
        reqs.append((t.net_parse(c.bench()), t.net_parse(r.bench()), c, r, faulty))
    adder = t.gen_adder(8)
    srv = t.Server(engine, streams=4, chunk_words=1 << 14)
    tickets = []
    for k, (a, b, c, r, faulty) in enumerate(reqs):
        tickets.append(srv.submit(t.EN_EQUIV, a, b))
        if k == 5:
            tickets.append(srv.submit(t.EN_VERIFY, adder, spec="c:s = a + b"))
        time.sleep(0.01)
    results = [srv.wait(tk) for tk in tickets]
    times = [srv.times(tk) for tk in tickets]
    srv.close()
    # requests overlapped on the device: one started before another finished
    assert any(times[i]["started_s"] < times[j]["done_s"] and times[j]["started_s"] < times[i]["done_s"]
               for i in range(len(times)) for j in range(i + 1, len(times)))
    adder_res = results.pop(6)
    assert adder_res["hits"] == 65536 and adder_res["hit"] == [2 * h + 1 for h in range(16)]
    for (a, b, c, r, faulty), got in zip(reqs, results):
        alone = engine.enum_run(t.EN_EQUIV, a, b)
        assert got == {**alone, "seconds": got["seconds"]} or (
            (got["hits"], got["hit"], got["tried"], got["ops"]) ==
            (alone["hits"], alone["hit"], alone["tried"], alone["ops"]))
        if not faulty:
            assert got["hits"] == 0
        for x in got["hit"]:
            assert c.evaluate(x) != r.evaluate(x)
        if len(c.inputs) <= 14:
            o = O.enum(t.EN_EQUIV, a, b, b_in=b.inputs, b_out=b.outputs)
            assert (o["hits"], o["hit"], o["tried"]) == (got["hits"], got["hit"], got["tried"])


def test_server_reports_errors(engine):
    srv = t.Server(engine)
    rank = t.gen_rank(9)
    tk = srv.submit(t.EN_VERIFY, rank, spec="p = x")
    with pytest.raises(ValueError, match="id and link cells"):
        srv.wait(tk)
    srv.close()
