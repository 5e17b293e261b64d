"""The interpreter tier's tape, built by the engine's host code, executed here by a numpy
restatement of the tape kernel's semantics (every word of the enumeration at once) and compared
with the oracle's enumeration: slot reuse, group independence and the hit word are all exercised
without a GPU."""
import numpy as np
import pytest

import tesseract_amd as t
from tesseract_amd import netlists as N
from oracle import pyoracle as O
from tests.helpers import adder, golden

PATTERNS = (0xAAAAAAAA, 0xCCCCCCCC, 0xF0F0F0F0, 0xFF00FF00, 0xFFFF0000)


def run_tape(tp, n):
    """hits and the 16 smallest hit assignments of the tape over all 2^n assignments."""
    space = 1 << n
    nwords = (space + 31) // 32
    gw = np.arange(nwords, dtype=np.uint64)
    f = np.zeros((tp["slots"], nwords), np.uint32)

    def read(ref):
        v = f[(int(ref) & 0x7FFFFFFF) // 256]
        return ~v if int(ref) >> 31 else v

    def entry(off, mask):
        v = f[int(off) // 256]
        return ~v if int(mask) else v

    for lf in tp["leaves"]:
        slot, j = (int(lf) & 0xFFFFFF) // 256, int(lf) >> 24
        f[slot] = PATTERNS[j] if j < 5 else (np.uint64(0) - ((gw >> np.uint64(j - 5)) & np.uint64(1))).astype(np.uint32)
    for g in range(len(tp["count"])):
        a, b, d, ma, mb = tp["ops"][g]  # all of the group's entries, padding included
        size = len(d)
        vals = [entry(a[j], ma[j]) & entry(b[j], mb[j]) for j in range(size)]  # every read first
        for j in range(size):
            f[int(d[j]) // 256] = vals[j]
    hit = read(tp["root"]).copy()
    rem = space - 32 * np.arange(nwords, dtype=np.int64)
    valid = np.where(rem >= 32, 0xFFFFFFFF, (1 << np.minimum(rem, 31)) - 1).astype(np.uint32)
    hit &= valid
    count = int(np.unpackbits(hit.view(np.uint8)).sum())
    first = []
    for w in np.nonzero(hit)[0]:
        m = int(hit[w])
        while m and len(first) < 16:
            first.append(int(w) * 32 + (m & -m).bit_length() - 1)
            m &= m - 1
        if len(first) >= 16:
            break
    return count, first


def test_tape_equiv_matches_oracle():
    checked = faulty_seen = 0
    for k in range(48):
        c, r, n_in, faulty = N.request(3, k)
        if n_in > 16 or len(c.gates) > 4000:
            continue
        a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
        tp = t.enum_tape(t.EN_EQUIV, a, b)
        assert tp["slots"] <= 640
        assert all(1 <= c_ <= tp["ops"].shape[2] for c_ in tp["count"])
        # every offset names a slot of the file; padding writes only the last slot (the sink)
        assert int(tp["ops"][:, :3].max(initial=0)) < 256 * tp["slots"]
        for g, c_ in enumerate(tp["count"]):
            assert (tp["ops"][g, 2, c_:] == 256 * (tp["slots"] - 1)).all()
            assert (tp["ops"][g, 2, :c_] < 256 * (tp["slots"] - 1)).all()
        got = run_tape(tp, n_in)
        o = O.enum(t.EN_EQUIV, a, b, b_in=b.inputs, b_out=b.outputs)
        assert got == (o["hits"], o["hit"]), (k, n_in)
        checked += 1
        faulty_seen += o["hits"] > 0
    assert checked >= 15 and faulty_seen >= 1


def _rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def test_tape_search_matches_toyhash():
    # assignment x sets input k from bit n-1-k, so the hash input is x bit-reversed
    bits = 12
    prog = golden("hash%d_4" % bits)
    h = [O.toyhash(_rev(x, bits), bits, 4) for x in range(1 << bits)]
    for target in (h[0], h[1234], 0xFFF):
        want = [x for x in range(1 << bits) if h[x] == target]
        got = run_tape(t.enum_tape(t.EN_SEARCH, prog, spec="h = %d" % target), bits)
        assert got == (len(want), want[:16])


def test_tape_constant_hit_words():
    # identical circuits: the hit word folds to the constant 0 (slot 0)
    rng = np.random.default_rng(4)
    c = N.circuit(9, 4, 200, rng)
    a = t.net_parse(c.bench())
    tp = t.enum_tape(t.EN_EQUIV, a, t.net_parse(c.bench()))
    assert tp["and_nodes"] == 0 and (tp["root"] & 0x7FFFFFFF) == 0
    assert run_tape(tp, 9) == (0, [])
    # an output and its complement: every assignment differs
    inv = N.Circuit(c.inputs, c.gates + [("y", c.outputs[0], c.outputs[0])], ["y"])
    same = N.Circuit(c.inputs, c.gates + [("y_t", c.outputs[0], c.outputs[0]), ("y", "y_t", "y_t")], ["y"])
    tp = t.enum_tape(t.EN_EQUIV, t.net_parse(inv.bench()), t.net_parse(same.bench()))
    assert run_tape(tp, 9) == (512, list(range(16)))


@pytest.mark.parametrize("name,spec", [
    ("adder8", "c:s = a + b + c"), ("adder8", "c:s = a + b"), ("adder8", "c:s = b - ~a - 1 + c"),
    ("mul6", "p = a * b"), ("mul6", "p = a * b + (a & b & 1)"), ("mul6", "p = (a << 2) * b"),
    ("mul6", "p = -(a * b) ^ 5"), ("mul8", "p = a * b"), ("mul8", "p = a * b | 0x100")])
def test_tape_verify_matches_oracle(name, spec):
    # a verify's spec bit-blasted into the tape (adders, shift-and-add products, logic, shifts):
    # the same hits as the oracle's enumeration against the spec's values
    prog = golden(name)
    tp = t.enum_tape(t.EN_VERIFY, prog, spec=spec)
    sp = t.Spec(spec, prog)
    n = len(prog.inputs)
    expect = np.array([sp.value(x) for x in range(1 << n)], np.uint64)
    cells = [sp.out_cell(i) for i in range(sp.width)]
    o = O.enum(t.EN_VERIFY, prog, out_cells=cells, expect=expect)
    assert run_tape(tp, n) == (o["hits"], o["hit"])


def test_tape_refuses_what_it_cannot_run():
    prog = adder(4)
    with pytest.raises(ValueError, match="tape form"):
        t.enum_tape(t.EN_VERIFY, prog, spec="c:s = (a + b) >> 1")
    rng = np.random.default_rng(5)
    c = N.circuit(12, 8, 3000, rng)
    with pytest.raises(ValueError, match="slots"):
        t.enum_tape(t.EN_EQUIV, t.net_parse(c.bench()), t.net_parse(N.rewrite(c, rng).bench()),
                    max_slots=32)
