"""The batch interpreter's window records (gpu_batch_records, no device): run as the interpreter
kernel runs them — every load of a window before any of its slots, then each slot's
rr = NAND ? !(rr & v) : rr and its store of rr — they must move every machine exactly as the
reference machine does (the oracle's run_lines), pass after pass. This pins the encoder's rules
on the CPU: a window ends before a NAND that reads a cell stored earlier in it, a STO shares the
slot of the NAND before it, a STO to a read-only cell or one overwritten later in the pass before
any read takes no slot, padding touches only the sink cell, and a slot flagged as storing nothing
may skip its store (as may a slot without a NAND its load)."""
import numpy as np
import pytest

import tesseract_amd as t
from oracle import pyoracle as O


def run_records(prog, rec, item_words, mem, rr, passes, skip=False):
    # a store word with bit 31 set stores nothing: its slot's rr goes to the sink, or nowhere
    # with skip (as the kernel's branch-around form runs it), where loads for slots without a
    # NAND are skipped too
    row = 4 * item_words
    groups = mem.shape[0]
    cells = np.zeros((prog.ncells + 1, groups, 8), np.uint64)  # the last one is the sink
    cells[:prog.ncells] = np.transpose(mem, (1, 0, 2))
    cells[0] = 0
    cells[1] = ~np.uint64(0)
    rr = rr.copy()
    win = rec.shape[2]
    for _ in range(passes):
        for w in rec:
            loads, stores = w[0], w[1]
            v = [cells[(int(x) & 0x7FFFFFFF) // row].copy() if not skip or int(x) >> 31 else None
                 for x in loads]
            for k in range(win):
                if int(loads[k]) >> 31:
                    rr = ~(rr & v[k])
                if not skip or not int(stores[k]) >> 31:
                    cells[(int(stores[k]) & 0x7FFFFFFF) // row] = rr
    return np.transpose(cells[:prog.ncells], (1, 0, 2)), rr


def readonly_program():
    b = t.Builder()
    x, y, z = b.cell("x"), b.cell("y"), b.cell("z")
    for op, c in [("n", x), ("s", 0), ("s", y), ("n", 1), ("s", 1), ("s", x), ("n", 0), ("n", y),
                  ("s", y), ("s", z), ("s", z), ("n", z), ("n", x), ("s", 0), ("s", x)]:
        (b.nand if op == "n" else b.sto)(c)
    return b.finish()


PROGRAMS = [("readonly", readonly_program)] + [
    ("rand%d_%d" % (c, n), (lambda c=c, n=n, s=s: t.random_program(c, n, s)))
    for c, n, s in [(1, 50, 1), (2, 80, 2), (5, 200, 3), (16, 300, 4), (100, 700, 5), (600, 2500, 6)]]


@pytest.mark.parametrize("win", [8, 16])
@pytest.mark.parametrize("item_words", [16, 64])
@pytest.mark.parametrize("name,make", PROGRAMS, ids=[n for n, _ in PROGRAMS])
def test_records_run_like_the_machine(name, make, win, item_words):
    prog = make()
    rec = t.batch_records(prog, item_words, win)
    assert rec.shape[1:] == (2, win)
    row = 4 * item_words
    sink = prog.ncells * row
    # offsets are whole rows inside the cells and the sink; read-only cells are never stored
    offs = np.concatenate([rec[:, 0].ravel() & 0x7FFFFFFF, rec[:, 1].ravel() & 0x7FFFFFFF])
    assert (offs % row == 0).all() and offs.max(initial=0) <= sink
    assert not np.isin(rec[:, 1] & 0x7FFFFFFF, [0, row]).any()
    # bit 31 of a store word marks a slot that stores nothing, and names the sink; bit 31 of a
    # load word marks a NAND, any other load names the sink
    assert ((rec[:, 1] >> 31 == 0) | (rec[:, 1] & 0x7FFFFFFF == sink)).all()
    assert ((rec[:, 0] >> 31 == 1) | (rec[:, 0] == sink)).all()
    mem, rr = O.random_lines(prog, 2, 11)
    want_mem, want_rr = O.run_lines(prog, mem, rr, 3)
    for skip in (False, True):
        got_mem, got_rr = run_records(prog, rec, item_words, mem, rr, 3, skip)
        assert np.array_equal(got_rr, want_rr), skip
        assert np.array_equal(got_mem[:, 2:], want_mem[:, 2:]), skip


def live_stores(prog):
    """STOs to writable cells whose value is read, or is the cell's value at the end of the pass:
    the next access to the cell later in the pass is not another STO."""
    ops, addr = prog.code()
    live, stored_next = 0, set()
    for o, a in zip(ops[::-1], addr[::-1]):
        a = int(a)
        if o == t.OP_NAND:
            stored_next.discard(a)
        else:
            if a >= 2 and a not in stored_next:
                live += 1
            stored_next.add(a)
    return live


def test_records_slots_per_instruction():
    # a STO shares the slot of the NAND before it, and a STO overwritten later in the pass before
    # any read takes no slot: a long random program takes about 0.61 slots per instruction (one
    # per NAND, one per live STO that follows a STO), padding aside, and stores only its live STOs
    prog = t.random_program(4000, 20000, 9)
    rec = t.batch_records(prog, 64, 16)
    sink = prog.ncells * 256
    nand, sto = (rec[:, 0] >> 31) != 0, (rec[:, 1] >> 31) == 0
    used = int(nand.sum() + sto.sum() - (nand & sto).sum())
    assert 0.56 * prog.len < used < 0.66 * prog.len
    assert rec.shape[0] * 16 < 0.66 * prog.len
    assert int(sto.sum()) == live_stores(prog) < 0.35 * prog.len


def test_records_refuse_other_windows():
    with pytest.raises(ValueError):
        t.batch_records(t.random_program(5, 20, 1), 64, 12)


def test_records_refuse_a_cell_outside_the_program():
    # a caller's program whose instruction names a cell >= ncells gets an error, not an encoding
    # whose byte offsets run past the item's cells
    prog = t.random_program(16, 40, 3)
    t.batch_records(prog)
    prog._p.code[7].addr = prog.ncells + 5
    with pytest.raises(ValueError, match="reads cell"):
        t.batch_records(prog)
    prog._p.code[7].addr = 2
    prog._p.code[9].op = 7
    with pytest.raises(ValueError, match="opcode"):
        t.batch_records(prog)

