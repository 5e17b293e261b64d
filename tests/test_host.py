"""Host side (CPU): the reference's programs stored as data (against the reference's own generators
and assembler, and against closed-form instruction counts), specs against the reference's known
values, error messages and parser, the program writer's gates, adders and registers against truth
tables, .bench netlists against the reference's loader, Beneš routing, random programs, the C
adapter of INTEGRATION.md against the reference's headers, and that the library exports every
declared C entry point."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O
from tesseract_amd import programs as P
from tests.helpers import adder, golden, life_cube

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"


def test_exports_every_declared_symbol():
    lib = C.CDLL(t.LIB_PATH)
    names = []
    for h in ("tesseract_host.h", "tesseract_gpu.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b((?:ts|gpu)_[a-z0-9_]+)\s*\(", text)
    missing = [n for n in set(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(set(names)) > 50


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="the reference's headers are not present")
def test_integration_adapter_compiles_against_reference_headers():
    # the binding INTEGRATION.md gives a maintainer, with its _Static_asserts on the layouts of
    # Program, Instr, CellInfo, Line and EnumResult, compiled against the reference's own headers
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, flags=re.S)
    adapter = [b for b in blocks if "_Static_assert" in b]
    assert adapter, "INTEGRATION.md has no adapter block"
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "gpu.c")
        with open(src, "w") as f:
            f.write("#include <stdio.h>\n#include <stdlib.h>\n" + adapter[0])
        r = subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-fsyntax-only", "-Werror=implicit",
                            "-I" + REF_SRC, "-I" + os.path.join(ROOT, "include"), src],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


# ---- the reference's programs as data ------------------------------------------------------

GOLDEN = sorted(f[:-4] for f in os.listdir(P.GOLDEN) if f.endswith(".npz"))


def test_golden_instruction_counts():
    # closed forms: mul16 = 16*AND6 + 16*BUF4 + 15*(16*AND6 + HA11 + 15*FA19); hash32 =
    # 8*32*17 + 3*popcount(round constants) + 32*4; Life 32x32 tile on 8+8 dims = 1024*123 +
    # 132*57 + 2*296; adder8 = 8 full adders of 19; PERM8 / SORT8 as ROADMAP.md:729,738 states
    assert golden("mul16").len == 16 * 6 + 16 * 4 + 15 * (16 * 6 + 11 + 15 * 19) == 6040
    assert golden("hash32_8").len == 4861
    life, board = life_cube(32, 32, 8, 8)
    assert life.len == 1024 * 123 + 132 * 57 + 2 * 296 == 134068
    assert life.dims == 16 and life.groups == 128 and len(board) == 1024
    assert golden("adder8").len == 152 == adder(8).len
    assert golden("life_8_8").len == 64 * 123
    assert golden("perm8").len == 1224 and golden("sort8").len == 7425
    assert golden("polymul").len == 220758
    assert golden("xor").len == 8 and golden("full_adder").len == 19 and golden("c17").len == 30


def _ref_program(ref, made_by):
    """The reference's program for a fixture's made_by, rebuilt with the reference's code."""
    L = ref.lib()
    m = re.match(r"(\w+)\(([\d, ]*)\)", made_by)
    if made_by.startswith("examples/"):
        p = ref.Prog()
        err = C.create_string_buffer(512)
        L.load_program.argtypes = [C.c_char_p, C.c_void_p, C.c_char_p, C.c_size_t]
        assert L.load_program(("/root/reference/" + made_by).encode(), p.ptr, err, 512), err.value
        return p, {}
    if made_by.startswith("lib/cube.inc"):
        what = made_by.split()[1]
        v = " ".join("v%d" % i for i in range(8))
        c = " ".join("c%d" % i for i in range(17))
        inc = '.include "/root/reference/lib/std.inc"\n.include "/root/reference/lib/cube.inc"\n'
        body = (".in %s %s\n.out %s\nPERM8 v c\n" % (v, c, v) if what == "PERM8" else
                ".in %s\n.out %s\nSORT8 v\n" % (v, v))
        p = ref.Prog()
        err = C.create_string_buffer(512)
        L.asm_string.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_char_p,
                                 C.c_size_t]
        assert L.asm_string((inc + body).encode(), b"x", b".", p.ptr, err, 512), err.value
        return p, {}
    fn, args = m.group(1), [int(a) for a in m.group(2).replace(" ", "").split(",") if a]
    p = ref.Prog()
    extra = {}
    if fn == "gen_life_cube":
        bd = np.zeros(args[0] * args[1], np.uint32)
        L.gen_life_cube(p.ptr, *args, bd.ctypes.data)
        extra["board"] = bd
    elif fn == "gen_life":
        bd = np.zeros(args[0] * args[1], np.uint32)
        L.gen_life(p.ptr, *args, bd.ctypes.data)
        extra["board"] = bd
    elif fn == "gen_beta":
        bc = np.zeros(2048, np.uint32)
        L.gen_beta(p.ptr, bc.ctypes.data)
        extra["cells"] = bc
    elif fn == "gen_gamma":
        gc = np.zeros(8192, np.uint8)
        assert L.gen_gamma(p.ptr, *args, gc.ctypes.data)
        extra["cells"] = gc
    elif fn == "gen_polymul":
        L.gen_polymul(p.ptr)
        zf = np.zeros((9, 512), np.uint16)
        zi = np.zeros((9, 512), np.uint16)
        L.ntt_twiddles(zf.ctypes.data, zi.ctypes.data)
        extra.update(zf=zf, zi=zi)
    else:
        getattr(L, fn)(p.ptr, *args)
    return p, extra


@pytest.mark.parametrize("name", GOLDEN)
def test_golden_program_is_the_references(ref, name):
    # every stored program is instruction for instruction, name for name, the program the
    # reference's own generator, assembler or netlist loader makes today
    ours = golden(name)
    theirs, extra = _ref_program(ref, ours.made_by)
    oa, aa = ours.code()
    ob, ab = theirs.code()
    assert ours.len == theirs.s.len and (oa == ob).all() and (aa == ab).all()
    assert [ours.name(i) for i in range(ours.ncells)] == theirs.names()
    assert ours.inputs == [theirs.s.inputs[i] for i in range(theirs.s.ninputs)]
    assert ours.outputs == [theirs.s.outputs[i] for i in range(theirs.s.noutputs)]
    assert ours.has_special == bool(theirs.s.special)
    if ours.has_special:
        info = C.cast(theirs.s.special, C.POINTER(C.c_uint32 * 2))
        for c in range(ours.ncells):
            kind, dim = info[c][0] & 0xFF, (info[c][0] >> 8) & 0xFF
            assert (ours.kind(c), ours.cell_dim(c), ours._p.special[c].src) == (kind, dim, info[c][1])
    for k, v in extra.items():
        assert np.array_equal(ours.extra[k], v), k


def test_program_files_refuse_malformed_programs(tmp_path):
    p = golden("mul4")
    P.save(str(tmp_path / "ok.npz"), p, "copy")
    q = P.load(str(tmp_path / "ok.npz"))
    assert (q.code()[1] == p.code()[1]).all() and q.ncells == p.ncells
    with np.load(str(tmp_path / "ok.npz")) as z:
        arrays = {k: z[k] for k in z.files}
    bad = dict(arrays, code=np.concatenate([arrays["code"], [np.uint32(p.ncells + 3)]]))
    np.savez(str(tmp_path / "bad.npz"), **bad)
    with pytest.raises(ValueError, match="names cell"):
        P.load(str(tmp_path / "bad.npz"))
    short = dict(arrays, names=arrays["names"][:10])
    np.savez(str(tmp_path / "short.npz"), **short)
    with pytest.raises(ValueError, match="names end"):
        P.load(str(tmp_path / "short.npz"))


def test_toyhash_oracle_is_the_references(ref):
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 1 << 32, 200, dtype=np.uint64):
        assert O.toyhash(int(x), 32, 8) == ref.lib().toyhash(int(x), 32, 8)


# ---- specs ------------------------------------------------------------------------------------

def test_spec_values():
    # the reference's known answers (tests/test_work.c:191-210) on its adder8; x: a7..a0 b7..b0 c
    add = golden("adder8")
    cases = [("c:s = a + b + c", (200 << 9) | (100 << 1) | 1, 301), ("s = a * 2 + b", (3 << 9) | (4 << 1), 10),
             ("s = (a + b) * 2", (3 << 9) | (4 << 1), 14), ("s = a - b", (3 << 9) | (4 << 1), 255),
             ("s = ~a & 0xf0 | b ^ 1", (0x0F << 9) | (0x02 << 1), 0xF3), ("s = 1 << 3 >> 1", 0, 4),
             ("c = -c", 1, 1)]
    for text, x, want in cases:
        assert t.Spec(text, add).value(x) == want, text


def test_spec_errors():
    add = golden("adder8")
    bad = [("q = a", "no output called q", True), ("s = z", "no input called z", True),
           ("s = a +", "ends too early", True), ("s = (a", "expected )", True),
           ("s a", "expected =", True), ("s = a", "must be a number", False),
           ("s = a + " + " + ".join(["a"] * 200), "too long", True),
           ("s = " + "a + (" * 70 + "a" + ")" * 70, "nested too deeply", True),
           ("s = a < b", "unexpected '<'", True), ("= a", "expected an output name", True)]
    for text, msg, inputs in bad:
        with pytest.raises(ValueError, match=re.escape(msg)):
            t.Spec(text, add, inputs)


def _ref_spec(ref, text, prog, inputs):
    """(value function or None, message) of the reference's spec_parse on our program (the
    program structs are layout-identical)."""
    err = C.create_string_buffer(512)
    s = ref.lib().spec_parse(text.encode(), C.cast(prog.ptr, C.c_void_p), int(inputs), err, 512)
    return s, err.value.decode()


def _random_spec(rng):
    atoms = ["a", "b", "c", "A", "B0", "x", "s", "7", "0x1f", "017", "0", "64", "99999999999999999999",
             "0x", "a1", "_q", "aa"]
    ops = ["+", "-", "*", "&", "|", "^", "<<", ">>", "<", ">", "~", "(", ")", " ", "", "=", ":", "$"]
    left = rng.choice(["s", "c:s", "s:c", "c", "q", "s0", "s:c:s", "", "S", "s:"])
    parts = []
    for _ in range(int(rng.integers(1, 14))):
        parts.append(str(rng.choice(atoms if rng.random() < 0.5 else ops)))
    return "%s = %s" % (left, " ".join(parts) if rng.random() < 0.5 else "".join(parts))


def _random_expr(rng, depth=0):
    r = rng.random()
    if depth > 5 or r < 0.3:
        return str(rng.choice(["a", "b", "c", "A", "7", "0x1f", "017", "0", "64", "3", "255"]))
    if r < 0.45:
        return str(rng.choice(["~", "-"])) + _random_expr(rng, depth + 1)
    if r < 0.55:
        return "(" + _random_expr(rng, depth + 1) + ")"
    op = str(rng.choice(["+", "-", "*", "&", "|", "^", "<<", ">>"]))
    sp = " " if rng.random() < 0.5 else ""
    return _random_expr(rng, depth + 1) + sp + op + sp + _random_expr(rng, depth + 1)


def test_spec_parser_against_reference(ref):
    # the reference's own parser decides acceptance, error messages and values: random texts over
    # the grammar's tokens, most of them malformed, and well-formed ones evaluated at random x
    add = golden("adder8")
    L = ref.lib()
    rng = np.random.default_rng(7)
    accepted = 0
    texts = [_random_spec(rng) for _ in range(3000)]
    texts += ["%s = %s" % (rng.choice(["s", "c:s", "s:c"]), _random_expr(rng)) for _ in range(1500)]
    texts += ["c:s = a + b + c", "s = a * b - ~c << 2", "s = -(a ^ b) >> 1 | c & 3", "c:s = a << 63",
              "s = a >> 64", "s = (((a)))", "s = ~-~a", "c: s = 0x10 * b"]
    for text in texts:
        for inputs in (True, False):
            theirs, msg = _ref_spec(ref, text, add, inputs)
            try:
                ours = t.Spec(text, add, inputs)
            except ValueError as e:
                assert not theirs, (text, str(e))
                assert str(e) == msg, text
                continue
            assert theirs, (text, msg)
            accepted += 1
            L.spec_width.argtypes = [C.c_void_p]
            assert ours.width == L.spec_width(theirs), text
            for x in rng.integers(0, 1 << 17, 16):
                assert ours.value(int(x)) == L.spec_value(theirs, int(x)), text
            L.spec_free(theirs)
    assert accepted > 1500


# ---- the program writer -------------------------------------------------------------------

def _eval(prog, inputs, passes=1, start=None):
    """Outputs of prog for assignments `inputs` (array of ints, input k = bit n-1-k), one machine
    per assignment (at most 512), after `passes` passes from power-on (or `start` lines)."""
    n = len(prog.inputs)
    mem, rr = O.reset_lines(prog, 1) if start is None else start
    mem = mem.copy()
    lanes = np.zeros(512, np.uint64)
    lanes[:len(inputs)] = inputs
    for k, cell in enumerate(prog.inputs):
        bits = ((lanes >> np.uint64(n - 1 - k)) & np.uint64(1)).astype(np.uint8)
        mem[0, cell] = np.packbits(bits, bitorder="little").view(np.uint64)
    mem, rr = O.run_lines(prog, mem, rr, passes)
    out = []
    for cell in prog.outputs:
        out.append(np.unpackbits(mem[0, cell].view(np.uint8), bitorder="little")[:len(inputs)])
    return np.array(out), mem, rr


GATES = {t.G_BUF: lambda v: v[0], t.G_NOT: lambda v: 1 - v[0],
         t.G_AND: lambda v: np.bitwise_and.reduce(v), t.G_NAND: lambda v: 1 - np.bitwise_and.reduce(v),
         t.G_OR: lambda v: np.bitwise_or.reduce(v), t.G_NOR: lambda v: 1 - np.bitwise_or.reduce(v),
         t.G_XOR: lambda v: np.bitwise_xor.reduce(v), t.G_XNOR: lambda v: 1 - np.bitwise_xor.reduce(v)}


@pytest.mark.parametrize("gate", sorted(GATES))
def test_gates_truth_tables(gate):
    # every gate at fan-in 1..5 (BUF, NOT: 1), exhaustively, and written over one of its inputs
    for n in ([1] if gate in (t.G_BUF, t.G_NOT) else range(1, 6)):
        for over_input in (False, True):
            b = t.Builder()
            ins = [b.cell("x%d" % i) for i in range(n)]
            out = ins[n - 1] if over_input else b.cell("o")
            for c in ins:
                b.input(c)
            b.gate(gate, out, ins)
            b.output(out)
            prog = b.finish()
            x = np.arange(1 << n, dtype=np.uint64)
            got, _, _ = _eval(prog, x)
            v = np.array([(x >> np.uint64(n - 1 - k)) & np.uint64(1) for k in range(n)]).astype(np.uint8)
            assert (got[0] == GATES[gate](v)).all(), (t.lib().ts_gate_name(gate), n, over_input)
    # the published instruction counts (with the store): AND/OR 6, XOR 8, BUF 4, NOT 3, NAND 5
    want = {t.G_AND: 6, t.G_OR: 6, t.G_XOR: 8, t.G_NAND: 5, t.G_NOR: 7, t.G_XNOR: 9}
    if gate in want:
        b = t.Builder()
        x, y, o = b.cell("x"), b.cell("y"), b.cell("o")
        b.gate(gate, o, [x, y])
        assert len(b) == want[gate]
        b.finish()


def test_adders_exhaustive():
    for n in (1, 3, 4):
        prog = adder(n)
        assert prog.len == 19 * n
        x = np.arange(1 << (2 * n + 1), dtype=np.uint64)
        got, _, _ = _eval(prog, x[:512])
        a, bb, c = (x[:512] >> np.uint64(n + 1)), (x[:512] >> np.uint64(1)) & np.uint64((1 << n) - 1), x[:512] & np.uint64(1)
        total = a + bb + c
        # outputs s<n-1>..s0 then the carry
        value = sum(got[k].astype(np.uint64) << np.uint64(n - 1 - k) for k in range(n))
        value = value + (got[n].astype(np.uint64) << np.uint64(n))
        assert (value == total).all()
    b = t.Builder()
    x, y, s, k = b.cell("x"), b.cell("y"), b.cell("s"), b.cell("k")
    for c in (x, y):
        b.input(c)
    b.half_adder(s, k, x, y)
    b.output(s)
    b.output(k)
    prog = b.finish()
    assert prog.len == 11
    got, _, _ = _eval(prog, np.arange(4, dtype=np.uint64))
    assert list(got[0]) == [0, 1, 1, 0] and list(got[1]) == [0, 0, 0, 1]


def test_registers_take_their_values_at_once():
    # a shift register moves one place per pass; a swap and a three-way rotation (cycles, which
    # go through staging cells) exchange their old values; a register reading itself keeps it
    b = t.Builder()
    x = b.cell("x")
    q = [b.cell("q%d" % i) for i in range(4)]
    r = [b.cell("r%d" % i) for i in range(3)]
    u, v, w = b.cell("u"), b.cell("v"), b.cell("w")
    b.input(x)
    b.reg(q[0], x)
    for i in range(1, 4):
        b.reg(q[i], q[i - 1])
    b.reg(r[0], r[1])
    b.reg(r[1], r[2])
    b.reg(r[2], r[0])
    b.reg(u, v)
    b.reg(v, u)
    b.reg(w, w)
    for c in q + r + [u, v, w]:
        b.output(c)
    prog = b.finish()
    rng = np.random.default_rng(3)
    start = O.random_lines(prog, 1, 9)
    old = {c: start[0][0, c].copy() for c in q + r + [u, v, w, x]}
    _, mem, _ = _eval(prog, rng.integers(0, 2, 512).astype(np.uint64), 1, start)
    assert (mem[0, q[0]] != 0).any()
    for i in range(1, 4):
        assert (mem[0, q[i]] == old[q[i - 1]]).all()
    assert (mem[0, r[0]] == old[r[1]]).all() and (mem[0, r[1]] == old[r[2]]).all()
    assert (mem[0, r[2]] == old[r[0]]).all()
    assert (mem[0, u] == old[v]).all() and (mem[0, v] == old[u]).all() and (mem[0, w] == old[w]).all()


def test_writer_reports_misuse_instead_of_aborting():
    for make, msg in [
        (lambda b: b.gate(99, b.cell("o"), [b.cell("x")]), "unknown gate"),
        (lambda b: b.gate(t.G_AND, 1, [b.cell("x"), b.cell("y")]), "read-only"),
        (lambda b: b.gate(t.G_AND, b.cell("o"), [b.cell("x"), 4000]), "only"),
        (lambda b: b.gate(t.G_NOT, b.cell("o"), [b.cell("x"), b.cell("y")]), "inputs"),
        (lambda b: b.nand(77), "only"),
        (lambda b: b.link("l", 500, 3), "link"),
        (lambda b: b.reg(0, b.cell("x")), "register"),
    ]:
        b = t.Builder()
        make(b)
        with pytest.raises(ValueError, match=msg):
            b.finish()


def test_writer_id_and_link_cells():
    b = t.Builder()
    x = b.cell("x")
    i3 = b.cell("id3")
    ln = b.link("n", x, 10)
    b.nand(ln)
    b.nand(i3)
    b.sto(x)
    p = b.finish()
    assert p.kind(i3) == t.CK_ID and p.cell_dim(i3) == 3
    assert p.kind(ln) == t.CK_LINK and p.cell_dim(ln) == 10 and p.dims == 11


# ---- Beneš, random programs ------------------------------------------------------------------

def test_benes_routes_every_permutation():
    rng = np.random.default_rng(3)
    for _ in range(20):
        dest = rng.permutation(512).astype(np.uint16)
        ctrl = t.benes_route(dest)
        for s in range(17):  # both machines of a pair carry the same setting
            d = t.benes_dim(s)
            k = np.arange(512)
            assert (ctrl[s][k] == ctrl[s][k ^ (1 << d)]).all()
        v = rng.integers(0, 1 << 40, 512, dtype=np.uint64)
        out = O.benes_apply(ctrl, v)
        assert (out[dest] == v).all()
    assert [t.benes_dim(s) for s in range(17)] == list(range(9)) + list(range(7, -1, -1))


def test_random_program_properties():
    a = t.random_program(20, 500, 42)
    b = t.random_program(20, 500, 42)
    c = t.random_program(20, 500, 43)
    ops, addrs = a.code()
    assert a.ncells == 22
    assert (addrs[ops == 1] >= 2).all() and (addrs < 22).all()
    assert 200 < (ops == 0).sum() < 300
    assert (addrs[ops == 0] < 2).any()
    assert (a.code()[1] == b.code()[1]).all() and not (a.code()[1] == c.code()[1]).all()


def test_random_program_is_the_references(ref):
    for args in ((64, 2000, 5), (16, 300, 9), (4096, 1000, 1)):
        ours, theirs = t.random_program(*args), ref.make("random_program", *args)
        assert (ours.code()[0] == theirs.code()[0]).all() and (ours.code()[1] == theirs.code()[1]).all()


def test_t1_start_matches_reference_stream():
    # group 0 is the reference's t1_start (src/tester.c:83-91): splitmix64 from seed ^ 0x5DEECE66D
    p = t.random_program(5, 10, 1)
    mem, rr = t.t1_start(p, 99, groups=2)
    s = C.c_uint64(99 ^ 0x5DEECE66D)
    words = [t.lib().ts_splitmix64(C.byref(s)) for _ in range(2 * 8 * 6)]
    assert list(mem[0, 2:, :].reshape(-1)) == words[:40]
    assert list(rr[0]) == words[40:48]
    assert list(mem[1, 2:, :].reshape(-1)) == words[48:88]


# ---- .bench netlists ---------------------------------------------------------------------------

def test_netlist_errors_and_dff():
    with pytest.raises(ValueError, match="driven twice"):
        t.net_parse("INPUT(a)\nb = NOT(a)\nb = BUF(a)\n")
    with pytest.raises(ValueError, match="loop without a DFF"):
        t.net_parse("INPUT(a)\nb = AND(a, c)\nc = NOT(b)\n")
    p = t.net_parse("INPUT(a)\nOUTPUT(q)\nq = DFF(d)\nd = XOR(a, q)\n")
    assert p.len > 0


def test_c17_program():
    # the reference's examples/c17.bench loaded here from its text: 6 NANDs of 5 instructions
    # (ROADMAP.md:299: 30 instructions); its function on all 32 assignments
    text = C17
    p = t.net_parse(text)
    assert p.len == 30 and len(p.inputs) == 5 and len(p.outputs) == 2
    got, _, _ = _eval(p, np.arange(32, dtype=np.uint64))
    theirs, _, _ = _eval(golden("c17"), np.arange(32, dtype=np.uint64))
    for x in range(32):
        a = [(x >> (4 - k)) & 1 for k in range(5)]  # inputs 1, 2, 3, 6, 7
        n10, n11 = 1 - (a[0] & a[2]), 1 - (a[2] & a[3])
        n16, n19 = 1 - (a[1] & n11), 1 - (n11 & a[4])
        assert [int(got[0][x]), int(got[1][x])] == [1 - (n10 & n16), 1 - (n16 & n19)]
    assert (got == theirs).all()


# ISCAS-85 c17: 5 inputs, 2 outputs, 6 NAND gates (the public benchmark circuit)
C17 = """INPUT(1)
INPUT(2)
INPUT(3)
INPUT(6)
INPUT(7)
OUTPUT(22)
OUTPUT(23)
10 = NAND(1, 3)
11 = NAND(3, 6)
16 = NAND(2, 11)
19 = NAND(11, 7)
22 = NAND(10, 16)
23 = NAND(16, 19)
"""


def _ref_net(ref, text):
    p = ref.Prog()
    e = C.create_string_buffer(512)
    ok = ref.lib().net_parse(text.encode(), b"net", p.ptr, e, len(e))
    return (p if ok else None), e.value.decode()


def _named(names):
    return {n: i for i, n in enumerate(names) if not n.startswith("__")}


def _random_netlist(rng, n_in, gates, dffs):
    ins = ["i%d" % k for k in range(n_in)]
    sig = list(ins)
    lines = ["INPUT(%s)" % s for s in ins]
    qs = ["q%d" % k for k in range(dffs)]
    body = []
    for k in range(gates):
        kind = str(rng.choice(["AND", "NAND", "OR", "NOR", "XOR", "XNOR", "NOT", "BUFF"]))
        fan = 1 if kind in ("NOT", "BUFF") else int(rng.integers(1, 5))
        src = sig + qs
        args = [src[int(rng.integers(0, len(src)))] for _ in range(fan)]
        name = "g%d" % k
        body.append("%s = %s(%s)" % (name, kind, ", ".join(args)))
        sig.append(name)
    for k in range(dffs):
        body.append("%s = DFF(%s)" % (qs[k], sig[int(rng.integers(0, len(sig)))]))
    rng.shuffle(body)  # gates in any order
    outs = [sig[-1 - k] for k in range(min(4, gates))] + qs[:2]
    return "\n".join(lines + ["OUTPUT(%s)" % o for o in outs] + body) + "\n"


def test_netlists_same_function_as_reference(ref):
    # the reference's loader and ours on the same texts: the same instruction count where there
    # is no flip-flop (gates cost the same), and the same values in every named cell after three
    # passes from the same random state (flip-flops included)
    from tesseract_amd import netlists as N
    rng = np.random.default_rng(3)
    texts = [C17, "INPUT(a)\nOUTPUT(q)\nq = DFF(d)\nd = XOR(a, q)\n",
             "INPUT(x)\nINPUT(y)\nOUTPUT(o)\nOUTPUT(p)\no = NOR(x, y)\np = XNOR(o, y)\n"]
    for k in range(4):
        c = N.circuit(int(rng.integers(4, 20)), int(rng.integers(1, 9)), int(rng.integers(20, 600)), rng)
        texts += [c.bench(), N.rewrite(c, rng).bench()]
    for k in range(12):
        texts.append(_random_netlist(rng, int(rng.integers(2, 9)), int(rng.integers(5, 60)),
                                     int(rng.integers(0, 6))))
    for text in texts:
        ours = t.net_parse(text)
        theirs, msg = _ref_net(ref, text)
        assert theirs is not None, msg
        if "DFF" not in text:
            assert ours.len == theirs.s.len
        on = _named([ours.name(i) for i in range(ours.ncells)])
        tn = _named(theirs.names())
        assert set(on) == set(tn)
        mem, rr = O.random_lines(ours, 1, 17)
        tmem = np.zeros((1, theirs.s.ncells, 8), np.uint64)
        tmem[0, 1] = ~np.uint64(0)
        for name, c in on.items():
            tmem[0, tn[name]] = mem[0, c]
        got, _ = O.run_lines(ours, mem, rr, 3)
        want, _ = O.run_lines(theirs, tmem, rr.copy(), 3)
        for name, c in on.items():
            assert (got[0, c] == want[0, tn[name]]).all(), (name, text[:200])


def test_netlist_messages_as_reference(ref):
    # malformed netlists: the reference's loader and ours refuse the same texts with the same
    # "file:line: message", and accept the same ones
    rng = np.random.default_rng(8)
    base = ["INPUT(a)", "INPUT(b)", "OUTPUT(z)", "z = AND(a, b)", "y = NOT(z)", "OUTPUT(y)"]
    mutants = ["INPUT a", "INPUT(a", "OUTPUT()", "z AND(a, b)", "z = FOO(a, b)", "z = AND(a b)",
               "z = AND(a, )", "z = NOT(a, b)", "q = DFF(a, b)", "(x)", "w = OR(v, a)", "a = BUF(b)",
               "INPUT(z)", "z = XOR(a, b)", "v = AND(v, a)", "u = NAND(t, a)\nt = OR(u, b)",
               "n.1 = AND(a, b)\nn_1 = OR(a, b)", "ZERO = AND(a, b)", "one = OR(a, b)", "  # only a comment",
               "\t", "9x = AND(a, b)", "x$y = dff(a)", "input(c)", "Input(c)", "m = buff(a)",
               "k = XNOR(a, b, a, b)", "p = NOR(a)", "r = AND(a, b) trailing"]
    for _ in range(300):
        lines = list(base)
        for _ in range(int(rng.integers(1, 3))):
            lines.insert(int(rng.integers(0, len(lines) + 1)), str(rng.choice(mutants)))
        text = "\n".join(lines) + ("\n" if rng.random() < 0.5 else "")
        theirs, msg = _ref_net(ref, text)
        if theirs is None:
            with pytest.raises(ValueError) as e:
                t.net_parse(text, "net")
            assert str(e.value) == msg, text
        else:
            t.net_parse(text, "net")


def test_generated_kernels_compile_for_gfx950():
    # hipRTC needs no GPU: every kernel form the engine emits must compile for gfx950
    progs = [golden("adder8"), golden("rank12"), golden("perm8"), t.random_program(30, 300, 1),
             life_cube(4, 4, 6, 6)[0], golden("cube_any"), golden("rule110")]
    # cube programs whose links read cells no pass writes (loads inside cross-lane moves) and
    # cells written before any read (stored as the last pass makes them)
    rng = np.random.default_rng(11)
    for dims in (9, 12):
        b = t.Builder()
        plain = [b.cell("c%d" % i) for i in range(10)]
        ids = [b.cell("ID%d" % d) for d in range(dims)]
        links = [b.link("l%d" % k, (plain + ids)[int(rng.integers(0, 10 + dims))], k % dims)
                 for k in range(3 * dims)]
        for k in range(3 * dims):  # links of links: a move inside the word and across lanes
            links.append(b.link("m%d" % k, links[k], (k + 5) % dims))
        for _ in range(200):
            if rng.random() < 0.55:
                b.nand(([0, 1] + plain + ids + links)[int(rng.integers(0, 12 + 7 * dims))])
            else:
                b.sto(plain[int(rng.integers(0, 6))])
        progs.append(b.finish())
    for p in progs:
        src, info = t.source(p)
        t.rtc_check(src)
    t.rtc_check(t.enum_source(t.EN_VERIFY, golden("mul8"), spec="p = a * b"))
    t.rtc_check(t.enum_source(t.EN_VERIFY, golden("mul8"), spec="p = a * b", ticks=2))
    t.rtc_check(t.enum_source(t.EN_SEARCH, golden("hash16_5"), spec="h = 77"))
    x = t.net_parse("INPUT(a)\nINPUT(b)\nOUTPUT(o)\no = XOR(a, b)\n")
    t.rtc_check(t.enum_source(t.EN_EQUIV, x, x))


def test_cube_in_one_workgroup_source():
    # a cube of 128-1,024 words compiles to one resident kernel whose workgroup is the cube (its
    # reads across waves through LDS, a barrier pair per chunk of each level); larger cubes and
    # lds_links = 2 keep the segmented form; every form compiles for gfx950
    from tests.test_gpu_run import random_cube_program
    for dims in (12, 14, 15):
        p = random_cube_program(16, 300, dims, seed=dims, links=30)
        src, info = t.source(p)
        assert info["resident"] == 1 and info["segments"] == 1
        assert "__launch_bounds__(%d)" % (1 << (dims - 5)) in src
        assert "ts_xs[" in src and src.count("__syncthreads();") % 2 == 0
        t.rtc_check(src)
        src2, info2 = t.source(p, {"lds_links": 2})
        assert info2["resident"] == 0 and "ts_xs[" not in src2
    p = random_cube_program(16, 300, 16, seed=16, links=30)
    src, info = t.source(p)
    assert info["resident"] == 0 and "ts_xs[" not in src
    # more values crossing the waves at one level than 64 KB of rows holds (48 rows of 512
    # words, 32 at a time): the level is written and read in two chunks
    from tests.test_gpu_run import wide_cross_program
    src, info = t.source(wide_cross_program(), {"lds_links": 1})
    assert "__shared__ u32 ts_xs[%d];" % (32 * 512) in src
    t.rtc_check(src)
    # the default weighs one cube's pass in both forms: gamma on 12 dims (6,311 LUTs one thread
    # evaluates in one workgroup, 5,945 in the segments' largest slices) stays segmented
    gamma = golden("gamma_4_4_4")
    assert t.source(gamma)[1]["segments"] == 2
    one = t.source(gamma, {"lds_links": 1})[1]
    assert one["resident"] == 1 and one["chain"] == one["bitop3"]


def test_pass_kernels_refuse_iterative_ilp():
    # the compiler faults (the process dies) on some kernels scheduled iterative-ilp, the
    # reference's PERM8 among them: the engine refuses that strategy with a message (pass kernels
    # here, before any source is made; enumeration kernels when they are compiled, a GPU test)
    with pytest.raises(RuntimeError, match="iterative-ilp"):
        t.source(golden("perm8"), {"sched": 2})
    t.source(golden("perm8"), {"sched": 3})


def test_enum_validation_messages():
    big = t.net_parse("\n".join("INPUT(i%d)" % k for k in range(47)) + "\nOUTPUT(o)\no = AND(i0, i1)\n")
    with pytest.raises(ValueError, match="47 inputs is too many"):
        t.enum_source(t.EN_VERIFY, big, spec="o = 1")
    with pytest.raises(ValueError, match="id and link cells connect the lanes"):
        t.enum_source(t.EN_VERIFY, golden("rank9"), spec="p = x")
    with pytest.raises(ValueError, match="bad --part"):
        t.enum_source(t.EN_VERIFY, golden("mul2"), spec="p = a * b", part=2, part_end=1, parts=3)


def test_t1_start_group_continues_the_stream():
    # the start state of one group, computed at its offset in the splitmix64 stream, equals the
    # library's t1_start of all groups up to it (group 0: the reference's t1_start)
    from tesseract_amd.workloads import t1_start_group
    p = t.random_program(9, 40, 5)
    mem, rr = t.t1_start(p, 1234, groups=5)
    for g in range(5):
        m, r = t1_start_group(p, 1234, g)
        assert np.array_equal(m, mem[g]) and np.array_equal(r, rr[g]), g


def test_batch_lds_classes():
    # sizes in LDS share at most three classes; neighbours whose wave counts per CU differ least
    # share one (a class sizes its LDS for its largest item)
    from tesseract_amd.workloads import lds_classes
    assert lds_classes({18: 19 * 256}) == {18: 0}
    assert lds_classes({18: 4864, 66: 8576, 258: 16576, 1026: 65728}) == {18: 0, 66: 0, 258: 1,
                                                                         1026: 2}
    # 16, 9, 4 and 2 waves per CU: the first two are closest
    assert lds_classes({18: 4864, 66: 17152, 258: 33152, 1026: 65728}) == {18: 0, 66: 0, 258: 1,
                                                                          1026: 2}
    got = lds_classes({c: 1000 * c for c in range(1, 9)})
    assert sorted(set(got.values())) == [0, 1, 2]
    assert all(got[a] <= got[b] for a in got for b in got if a < b)
