"""gpu_enum_exec_noalloc (the hit list in the library's static device scratch) against
gpu_enum_exec on the same compiled enumerations: two calls in a row, the second on a smaller
input, many-hit enumerations whose list overflows the static list (re-run by halves), the tape
tier, and the size limit."""
import pytest

import tesseract_amd as t
from tests.helpers import golden

pytestmark = pytest.mark.gpu


def same(a, b):
    keys = ("space", "tried", "hits", "hit", "nhit", "ops")
    return {k: a[k] for k in keys} == {k: b[k] for k in keys}


def test_noalloc_equals_exec(engine):
    # first: mul10 against a spec it meets only on odd a and b (2^18 of 2^20 hits, 8,192 hit
    # words: more than the static list holds); second: mul6 with a few hits, a smaller input
    big = t.EnumCode(engine, t.EN_VERIFY, golden("mul10"), spec="p = a * b + (a & b & 1)")
    small = t.EnumCode(engine, t.EN_VERIFY, golden("mul6"), spec="p = a * b + 1")
    print("noalloc: mul10 planted difference, then mul6", flush=True)
    want_big, want_small = big.run(), small.run()
    assert want_big["hits"] == 1 << 18 and want_small["hits"] == 1 << 12
    got_big = big.run_noalloc()
    got_small = small.run_noalloc()
    assert same(got_big, want_big) and same(got_small, want_small)
    # the last call's contents stay in the scratch; the next call clears what it counts
    assert same(big.run_noalloc(), want_big)
    for words in (1, 7, 4096):  # lists too short for one launch: the halving path
        assert same(small.run_noalloc(words), want_small)
        assert same(big.run_noalloc(words), want_big)


def test_noalloc_tape_tier(engine):
    prog = golden("hash16_5")
    for spec in ("h = 12345", "h = 7"):
        code = t.EnumCode(engine, t.EN_SEARCH, prog, spec=spec, tune={"tape": 1})
        print("noalloc: tape search", spec, flush=True)
        assert same(code.run_noalloc(), code.run())
        assert same(code.run_noalloc(2), code.run())


def test_noalloc_rejects_a_larger_list(engine):
    code = t.EnumCode(engine, t.EN_VERIFY, golden("mul4"), spec="p = a * b")
    with pytest.raises(ValueError, match="exceeds the static scratch"):
        code.run_noalloc(4097)
    assert code.run_noalloc(4096)["hits"] == 0
