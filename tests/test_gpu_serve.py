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
from tests.helpers import golden

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
    # a large request runs first; small ones submitted while it runs must finish before it does
    # (continuous batching, not a queue), and every result must equal the request run alone
    rng = np.random.default_rng(11)
    big_c = N.circuit(36, 8, 20000, rng)
    big_r = N.rewrite(big_c, rng)
    big = (t.net_parse(big_c.bench()), t.net_parse(big_r.bench()))
    reqs = []
    for k in range(18):
        n_in = [10, 14, 18, 22][k % 4]
        gates = [60, 300, 1500][k % 3]
        c, r, faulty = small_request(7, k, n_in, gates)
        reqs.append((t.net_parse(c.bench()), t.net_parse(r.bench()), c, r, faulty))
    adder = golden("adder8")
    srv = t.Server(engine, streams=4, chunk_words=1 << 20)
    big_tk = srv.submit(t.EN_EQUIV, *big)
    while srv.poll(big_tk)[0] < 3:  # running
        time.sleep(0.01)
    tickets = []
    for k, (a, b, c, r, faulty) in enumerate(reqs):
        tickets.append(srv.submit(t.EN_EQUIV, a, b))
        if k == 5:
            tickets.append(srv.submit(t.EN_VERIFY, adder, spec="c:s = a + b"))
    results = [srv.wait(tk) for tk in tickets]
    big_res = srv.wait(big_tk)
    times = [srv.times(tk) for tk in tickets]
    big_times = srv.times(big_tk)
    srv.close()
    print("big request: %.2f s running" % (big_times["done_s"] - big_times["started_s"]), flush=True)
    assert big_res["hits"] == 0 and big_res["tried"] == 1 << 36
    # requests ready (compiled) well before the big one ends are not queued behind it
    inside = [tm for tm in times if tm["compiled_s"] < big_times["done_s"] - 0.05]
    assert len(inside) >= 3
    assert all(tm["done_s"] <= big_times["done_s"] for tm in inside)
    adder_res = results.pop(6)
    assert adder_res["hits"] == 65536 and adder_res["hit"] == [2 * h + 1 for h in range(16)]
    # both tiers served: small requests on the tape interpreter, the small verify too (its spec
    # bit-blasted into the tape), the big one (far more than 2 s of tape) compiled
    assert big_times["tier"] == 0 and times[6]["tier"] == 1
    assert sum(tm["tier"] == 1 for tm in times) >= 10
    key = lambda r: (r["hits"], r["hit"], r["tried"], r["ops"], r["space"])
    for (a, b, c, r, faulty), got in zip(reqs, results):
        alone = engine.enum_run(t.EN_EQUIV, a, b)
        assert key(got) == key(alone)
        if not faulty:
            assert got["hits"] == 0
        for x in got["hit"]:
            assert c.evaluate(x) != r.evaluate(x)
        if len(c.inputs) <= 14:
            o = O.enum(t.EN_EQUIV, a, b, b_in=b.inputs, b_out=b.outputs)
            assert (o["hits"], o["hit"], o["tried"]) == (got["hits"], got["hit"], got["tried"])


def test_server_reports_errors(engine):
    srv = t.Server(engine)
    rank = golden("rank9")
    tk = srv.submit(t.EN_VERIFY, rank, spec="p = x")
    with pytest.raises(ValueError, match="id and link cells"):
        srv.wait(tk)
    srv.close()


def test_server_tiers_agree(engine):
    # the same requests through the tape tier (generous limit) and compiled only (limit < 0)
    reqs = []
    for k in range(8):
        c, r, faulty = small_request(9, k, [12, 20, 26, 16][k % 4], [80, 700, 2500][k % 3])
        reqs.append((t.net_parse(c.bench()), t.net_parse(r.bench())))
    out = {}
    for limit in (1e6, -1.0):
        srv = t.Server(engine, streams=3, chunk_words=1 << 16, tape_max_ms=limit)
        tks = [srv.submit(t.EN_EQUIV, a, b) for a, b in reqs]
        out[limit] = [srv.wait(tk) for tk in tks]
        tiers = [srv.times(tk)["tier"] for tk in tks]
        srv.close()
        assert tiers == [1 if limit > 0 else 0] * len(reqs)
    key = lambda r: (r["hits"], r["hit"], r["tried"], r["ops"], r["space"])
    assert [key(r) for r in out[1e6]] == [key(r) for r in out[-1.0]]


def test_server_tape_then_compiled(engine):
    # every equiv request is worth a compile (tape_max_ms tiny) and races its compile on the tape
    # (tape_race_ms large): each starts on the tape in short chunks and moves to its kernel when the
    # compile is in (2^32-2^33 assignments take the tape longer than a compile); the results equal
    # each request run alone
    reqs = []
    for k in range(6):
        c, r, faulty = small_request(13, k, [32, 33, 32][k % 3], [800, 1500][k % 2])
        reqs.append((t.net_parse(c.bench()), t.net_parse(r.bench()), c, r, faulty))
    srv = t.Server(engine, streams=4, tape_max_ms=1e-6, tape_race_ms=1e6)
    tks = [srv.submit(t.EN_EQUIV, a, b) for a, b, _, _, _ in reqs]
    got = [srv.wait(tk) for tk in tks]
    times = [srv.times(tk) for tk in tks]
    tiers = [tm["tier"] for tm in times]
    srv.close()
    print("tiers", tiers, flush=True)
    assert set(tiers) <= {1, 2} and 2 in tiers
    # the footprint admission counted: a compiled kernel's registers and resident waves, the
    # code and the tape it held
    for tm in times:
        assert tm["waves_per_cu"] > 0 and tm["hbm_bytes"] > 0
        if tm["tier"] == 2:
            assert 0 < tm["vgprs"] <= 512
    key = lambda r: (r["hits"], r["hit"], r["tried"], r["ops"], r["space"])
    for (a, b, c, r, faulty), res in zip(reqs, got):
        assert key(res) == key(engine.enum_run(t.EN_EQUIV, a, b))
        for x in res["hit"]:
            assert c.evaluate(x) != r.evaluate(x)


def test_server_release_frees_finished_requests(engine):
    # a long-running server releases each request once its result has been read: the server then
    # holds nothing for it (its ticket is unknown); unfinished requests cannot be released
    srv = t.Server(engine, streams=2, chunk_words=1 << 18)
    tickets, expect = [], []
    for k in range(8):
        c, r, faulty = small_request(21, k, [10, 16, 22, 26][k % 4], [60, 300][k % 2])
        a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
        tickets.append(srv.submit(t.EN_EQUIV, a, b))
        expect.append(engine.enum_run(t.EN_EQUIV, a, b))
    rng = np.random.default_rng(5)
    big_c = N.circuit(30, 4, 3000, rng)
    big = srv.submit(t.EN_EQUIV, t.net_parse(big_c.bench()), t.net_parse(N.rewrite(big_c, rng).bench()))
    with pytest.raises(ValueError, match="has not finished"):
        srv.release(big)
    assert srv.live() == 9
    for tk, want in zip(tickets, expect):
        got = srv.wait(tk)
        assert (got["hits"], got["hit"], got["tried"]) == (want["hits"], want["hit"], want["tried"])
        srv.release(tk)
    assert srv.wait(big)["hits"] == 0
    srv.release(big)
    assert srv.live() == 0
    with pytest.raises(ValueError, match="no request"):
        srv.release(tickets[0])
    with pytest.raises(ValueError, match="no request"):
        srv.wait(tickets[0])
    srv.close()
