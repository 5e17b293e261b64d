"""The generated enumeration kernel, run on the host (tests/kernel_cpu.py): the grouped form (the
lane-free part of the network once per group of 32 words) against the one-word form and against
the oracle's and the reference's enumerations, on searches, verifies and equivalences of at least
24 inputs, whole spaces, shares of 2^32 spaces and ranges that start and end inside a group."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tesseract_amd import netlists as N
from tesseract_amd.programs import golden
from tests import kernel_cpu as K


def _both(mode, a, b=None, spec=None, **kw):
    plain = t.enum_source(mode, a, b, spec=spec, tune={"group": 1}, **kw)
    grouped = t.enum_source(mode, a, b, spec=spec, tune={"group": 2}, **kw)
    assert "gwb" not in plain and "gwb" in grouped
    return plain, grouped


def test_search_hash24(ref):
    # every hit maps back to the target through the hash's closed form (input bits reversed
    # against x), and the count and 16 smallest equal the reference's enum_run
    prog = golden("hash24_5")
    for x in (777, 12345678):
        target = O.toyhash(x, 24, 5)
        plain, grouped = _both(t.EN_SEARCH, prog, spec="h = %d" % target)
        hp, xp = K.run(plain, 24)
        hg, xg = K.run(grouped, 24)
        assert (hp, xp) == (hg, xg)
        assert int(format(x, "024b")[::-1], 2) in xg
        assert all(O.toyhash(int(format(h, "024b")[::-1], 2), 24, 5) == target for h in xg)
        w = ref.enum_run(t.EN_SEARCH, prog.ptr, spec_text="h = %d" % target, spec_inputs=False,
                         threads=8)
        assert hg == w["hits"] and xg[:16] == w["hit"]


@pytest.mark.parametrize("part,part_end", [(0, 0), (37, 38)])
def test_mul16_shares_against_the_reference(ref, part, part_end):
    # shares of 2^24 / 2^25 assignments of the 2^32 space: the correct spec (no hits) and a
    # planted difference (a and b both odd: a quarter of the share), with the reference's own
    # enum_run on the same shares
    prog = golden("mul16")
    for spec in ("p = a * b", "p = a * b + (a & b & 1)"):
        plain, grouped = _both(t.EN_VERIFY, prog, spec=spec, part=part, part_end=part_end, parts=256)
        hp, xp = K.run(plain, 32, part=part, part_end=part_end, parts=256)
        hg, xg = K.run(grouped, 32, part=part, part_end=part_end, parts=256)
        assert (hp, xp) == (hg, xg)
        w = ref.enum_run(t.EN_VERIFY, prog.ptr, spec_text=spec, part=part,
                         part_end=part_end, parts=256, threads=8)
        assert hg == w["hits"] and xg[:16] == w["hit"]
        # share r takes the chunks c = x >> 15 with c % 256 == r: b0 = x bit 15 = r & 1, and a0
        # (x bit 31) is odd in half of each share's 2^24 assignments
        odd_b = sum(r & 1 for r in range(part, part_end + 1))
        assert hg == (0 if spec == "p = a * b" else odd_b << 23)


@pytest.mark.parametrize("spec", ["p = b * a", "p = ~(0 - a * b) + 1",
                                  "p = a * b + (a << 3) - a * 8", "p = (a + 1) * b - b",
                                  "p = a * b + 12345", "p = a * b ^ 1", "p = a * b + (a & 1)",
                                  "p = a * b + (b & 64)"])
def test_mul16_spec_forms(spec):
    # specs equal to a * b in several forms, some kept linear in the lane bits (the grouped
    # kernel's bit-sliced add) and some not (its per-lane values); some differ: a * b ^ 1 and
    # a * b + 12345 in every assignment (bit 0, which no lane and no word of a group reaches, is
    # compared once per group), + (a & 1) where a is odd (a group's whole 32 words, from bit 0
    # up), + (b & 64) where b bit 6 (x bit 9, a word bit of the group) is set
    prog = golden("mul16")
    plain, grouped = _both(t.EN_VERIFY, prog, spec=spec, part=5, part_end=5, parts=1024)
    hp, xp = K.run(plain, 32, part=5, part_end=5, parts=1024)
    hg, xg = K.run(grouped, 32, part=5, part_end=5, parts=1024)
    assert (hp, xp) == (hg, xg)
    want = {"p = a * b + 12345": 1 << 22, "p = a * b ^ 1": 1 << 22,
            "p = a * b + (a & 1)": 1 << 21, "p = a * b + (b & 64)": 1 << 21}
    assert hg == want.get(spec, 0)


def test_ranges_inside_groups():
    # launches over word ranges that begin and end inside a group (the list-overflow halving
    # launches such ranges): the grouped kernel lists exactly the one-word kernel's hits
    prog = golden("mul16")
    plain, grouped = _both(t.EN_VERIFY, prog, spec="p = a * b + (a & b & 1)", part=3, part_end=3,
                           parts=256)
    for i0, i1 in ((37, 70001), (1024 * 64 + 5, 1024 * 64 + 6), (0, 31), (33, 64)):
        assert K.run(plain, 32, part=3, part_end=3, parts=256, i0=i0, i1=i1) == \
            K.run(grouped, 32, part=3, part_end=3, parts=256, i0=i0, i1=i1)


def test_equiv_netlists(ref):
    # circuit against its NAND-level rewrite, with and without an injected fault, 24 and 26
    # inputs: grouped and one-word kernels against the reference's enum_run
    rng = np.random.default_rng(22)
    differ = 0
    for n_in, gates in ((24, 300), (26, 200)):
        c = N.circuit(n_in, 4, gates, rng)
        for faulty in (False, True):
            r = N.rewrite(c, rng)
            if faulty:
                r = N.fault(r, rng)
            a, b = t.net_parse(c.bench()), t.net_parse(r.bench())
            plain, grouped = _both(t.EN_EQUIV, a, b)
            hp, xp = K.run(plain, n_in)
            hg, xg = K.run(grouped, n_in)
            w = ref.enum_run(t.EN_EQUIV, a.ptr, b.ptr, threads=8)
            assert (hp, xp) == (hg, xg)
            assert hg == w["hits"] and xg[:16] == w["hit"]
            if not faulty:
                assert hg == 0
            differ += hg > 0
    assert differ  # the 24-input fault changes 3,159,680 of the 2^24 assignments
