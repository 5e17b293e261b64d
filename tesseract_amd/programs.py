"""Program files: NAND/STO programs stored as data, loaded into the library's ts_program.

The engine runs whatever program its caller hands it (the reference host passes its own Program
by pointer). Programs made elsewhere, e.g. by the reference's generators and assembler, travel
as .npz files (numpy, no pickled objects):

  code     u32 per instruction: cell address | opcode << 31 (0 = NAND, 1 = STO)
  ncells   the number of cells, ZERO and ONE included
  names    the cell names, each ended by a NUL byte
  inputs, outputs   u32 cells in declaration order
  kind, dim, src    per cell (only for programs with id or link cells): 0 plain, 1 id, 2 link
  x_*      extra arrays that belong with the program (e.g. a Life board's cell map)
  made_by  a description of where the program came from

tests/golden/programs holds the reference's workload programs in this form (written by
tests/golden/make_programs.py); golden(name) loads one of them.
"""
import ctypes as C
import json
import os

import numpy as np

from . import Program, _err, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "programs")


def load(path):
    """The program in `path`; its extra arrays as .extra (name without x_), its origin as
    .made_by."""
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    code = np.ascontiguousarray(arrays["code"], np.uint32)
    names = np.ascontiguousarray(arrays["names"], np.uint8)
    ins = np.ascontiguousarray(arrays["inputs"], np.uint32)
    outs = np.ascontiguousarray(arrays["outputs"], np.uint32)
    special = "kind" in arrays
    kind = np.ascontiguousarray(arrays["kind"], np.uint8) if special else None
    dim = np.ascontiguousarray(arrays["dim"], np.uint8) if special else None
    src = np.ascontiguousarray(arrays["src"], np.uint32) if special else None
    ptr = lambda a: None if a is None else a.ctypes.data
    p = Program()
    e = _err()
    if not lib().ts_program_from_arrays(p.ptr, ptr(code), len(code), int(arrays["ncells"]),
                                        ptr(names), len(names), ptr(ins), len(ins), ptr(outs),
                                        len(outs), ptr(kind), ptr(dim), ptr(src), e, len(e)):
        p._owned = False
        raise ValueError("%s: %s" % (path, e.value.decode()))
    p.extra = {k[2:]: v for k, v in arrays.items() if k.startswith("x_")}
    p.made_by = json.loads(bytes(arrays["made_by"]).decode()) if "made_by" in arrays else None
    return p


def golden(name):
    """One of the reference's programs from tests/golden/programs (e.g. "mul16", "perm8",
    "life_cube_32_32_8_8", "adder8")."""
    return load(os.path.join(GOLDEN, name + ".npz"))


def save(path, prog, made_by="", **extra):
    """Writes prog in the program-file form (extra arrays stored as x_<name>)."""
    ops, addrs = prog.code()
    arrays = {
        "code": (addrs.astype(np.uint32) | (ops.astype(np.uint32) << np.uint32(31))),
        "ncells": np.array(prog.ncells, np.uint32),
        "names": np.frombuffer(b"".join(prog.name(c).encode() + b"\0"
                                        for c in range(prog.ncells)), np.uint8),
        "inputs": np.array(prog.inputs, np.uint32),
        "outputs": np.array(prog.outputs, np.uint32),
        "made_by": np.frombuffer(json.dumps(made_by).encode(), np.uint8),
    }
    if prog.has_special:
        arrays["kind"] = np.array([prog.kind(c) for c in range(prog.ncells)], np.uint8)
        arrays["dim"] = np.array([prog.cell_dim(c) for c in range(prog.ncells)], np.uint8)
        arrays["src"] = np.array([prog._p.special[c].src for c in range(prog.ncells)], np.uint32)
    for k, v in extra.items():
        arrays["x_" + k] = v
    np.savez_compressed(path, **arrays)
