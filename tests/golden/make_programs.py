"""Writes the program fixtures under tests/golden/programs/ from the reference's own code.

The engine takes programs as data (a caller's Program, reference src/isa.h:36-46); the workloads,
tests and benchmark run the reference's programs, made by the reference's generators and assembler
as compiled into oracle/_ref (`make -C oracle ref`, from the sources under /root/reference):

  mul<b>              gen_mul(b)                          src/gen.c:39-75
  hash<b>_<r>         gen_hash(b, r)                      src/gen.c:98-162
  rank<d>             gen_rank(d)                         src/gen.c:327-...
  wht12, xorconv      gen_wht(12), gen_xorconv()          src/gen.c
  polymul             gen_polymul() + ntt_twiddles        src/gen.c:456-669
  life_cube_<w>_<h>_<r>_<c>  gen_life_cube + its board map (x_board: the cell of each tile cell,
                      row-major) and x_geometry (w, h, rbits, cbits)   src/gen.c:285-319
  life_8_8            gen_life(8, 8) + its board map      src/gen.c:169-219
  perm8, sort8        PERM8 v c / SORT8 v of lib/cube.inc assembled as the reference's cube demos do
                      (src/cubedemo.c:360-366)
  adder8, xor, full_adder, counter4, mux, cube_any, cube_rank, cube_max, rule110
                      examples/*.asm through the reference's assembler (asm_file)
  c17                 examples/c17.bench through the reference's load_program
  beta                gen_beta() + its BetaCells (x_cells, the struct's words)     src/beta.c:678
  gamma_<x>_<y>_<z>   gen_gamma(x, y, z) + its GammaCells (x_cells, the struct's bytes) and
                      x_geometry (x, y, z bits)                                   src/gamma.c:703

Each file is a numpy .npz (load with allow_pickle=False): code (u32, addr | op << 31), ncells,
names (the cell names, each NUL-terminated), inputs, outputs, kind/dim/src (only for programs with
id or link cells), x_* extra arrays, made_by (a description). Run in the build container:
  python tests/golden/make_programs.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"
OUT = os.path.join(HERE, "programs")

from oracle import pyref  # noqa: E402


class RefCellInfo(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("dim", C.c_uint8), ("src", C.c_uint32)]


def bind():
    L = pyref.lib()
    L.asm_string.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_void_p, C.c_char_p, C.c_size_t]
    L.asm_string.restype = C.c_int
    L.asm_file.argtypes = [C.c_char_p, C.c_void_p, C.c_char_p, C.c_size_t]
    L.asm_file.restype = C.c_int
    L.load_program.argtypes = [C.c_char_p, C.c_void_p, C.c_char_p, C.c_size_t]
    L.load_program.restype = C.c_int
    return L


def save(name, prog, made_by, **extra):
    s = prog.s
    n = s.len
    raw = np.ctypeslib.as_array(C.cast(s.code, C.POINTER(C.c_uint32)), shape=(2 * n,)).reshape(n, 2) \
        if n else np.zeros((0, 2), np.uint32)
    code = (raw[:, 1].astype(np.uint32) | ((raw[:, 0] & 0xFF).astype(np.uint32) << np.uint32(31)))
    names = b"".join(s.cell_names[i] + b"\0" for i in range(s.ncells))
    arrays = {
        "code": code.astype(np.uint32),
        "ncells": np.array(s.ncells, np.uint32),
        "names": np.frombuffer(names, np.uint8),
        "inputs": np.array([s.inputs[i] for i in range(s.ninputs)], np.uint32),
        "outputs": np.array([s.outputs[i] for i in range(s.noutputs)], np.uint32),
        "made_by": np.frombuffer(json.dumps(made_by).encode(), np.uint8),
    }
    if s.special:
        info = C.cast(s.special, C.POINTER(RefCellInfo))
        arrays["kind"] = np.array([info[c].kind for c in range(s.ncells)], np.uint8)
        arrays["dim"] = np.array([info[c].dim for c in range(s.ncells)], np.uint8)
        arrays["src"] = np.array([info[c].src for c in range(s.ncells)], np.uint32)
    for k, v in extra.items():
        arrays["x_" + k] = v
    path = os.path.join(OUT, name + ".npz")
    np.savez_compressed(path, **arrays)
    return path


def assemble(L, text, what):
    p = pyref.Prog()
    err = C.create_string_buffer(512)
    if not L.asm_string(text.encode(), what.encode(), b".", p.ptr, err, len(err)):
        raise RuntimeError(err.value.decode())
    return p


def load(L, path):
    p = pyref.Prog()
    err = C.create_string_buffer(512)
    if not L.load_program(path.encode(), p.ptr, err, len(err)):
        raise RuntimeError(err.value.decode())
    return p


def main():
    if not pyref.available():
        sys.exit("oracle/_ref is not built: make -C oracle ref (needs the reference sources)")
    L = bind()
    os.makedirs(OUT, exist_ok=True)
    made = []
    for b in list(range(1, 11)) + [16]:
        made.append(save("mul%d" % b, pyref.make("gen_mul", b), "gen_mul(%d)" % b))
    for b, r in ((3, 5), (12, 4), (16, 5), (24, 5), (32, 8)):
        p = pyref.Prog()
        if not L.gen_hash(p.ptr, b, r):
            raise RuntimeError("gen_hash(%d, %d)" % (b, r))
        made.append(save("hash%d_%d" % (b, r), p, "gen_hash(%d, %d)" % (b, r)))
    for d in (9, 12, 16):
        made.append(save("rank%d" % d, pyref.make("gen_rank", d), "gen_rank(%d)" % d))
    made.append(save("wht12", pyref.make("gen_wht", 12), "gen_wht(12)"))
    made.append(save("xorconv", pyref.make("gen_xorconv"), "gen_xorconv()"))
    zf = np.zeros((9, 512), np.uint16)
    zi = np.zeros((9, 512), np.uint16)
    L.ntt_twiddles(zf.ctypes.data, zi.ctypes.data)
    made.append(save("polymul", pyref.make("gen_polymul"), "gen_polymul() and ntt_twiddles()",
                     zf=zf, zi=zi))
    for args in ((32, 32, 8, 8), (4, 4, 3, 3), (4, 4, 6, 6), (8, 6, 5, 6), (8, 8, 4, 5)):
        bd = np.zeros(args[0] * args[1], np.uint32)
        p = pyref.Prog()
        L.gen_life_cube(p.ptr, *args, bd.ctypes.data)
        made.append(save("life_cube_%d_%d_%d_%d" % args, p, "gen_life_cube(%d, %d, %d, %d)" % args,
                         board=bd, geometry=np.array(args, np.uint32)))
    bd = np.zeros(64, np.uint32)
    p = pyref.Prog()
    L.gen_life(p.ptr, 8, 8, bd.ctypes.data)
    made.append(save("life_8_8", p, "gen_life(8, 8)", board=bd, geometry=np.array((8, 8), np.uint32)))
    lib_inc = '.include "%s/lib/std.inc"\n.include "%s/lib/cube.inc"\n' % (REF, REF)
    v = " ".join("v%d" % i for i in range(8))
    c = " ".join("c%d" % i for i in range(17))
    made.append(save("perm8", assemble(L, lib_inc + ".in %s %s\n.out %s\nPERM8 v c\n" % (v, c, v),
                                       "perm8"), "lib/cube.inc PERM8 v c (src/cubedemo.c:363-366)"))
    made.append(save("sort8", assemble(L, lib_inc + ".in %s\n.out %s\nSORT8 v\n" % (v, v), "sort8"),
                     "lib/cube.inc SORT8 v (src/cubedemo.c:359-360)"))
    for ex in ("adder8", "xor", "full_adder", "counter4", "mux", "cube_any", "cube_rank",
               "cube_max", "rule110"):
        made.append(save(ex, load(L, os.path.join(REF, "examples", ex + ".asm")),
                         "examples/%s.asm" % ex))
    made.append(save("c17", load(L, os.path.join(REF, "examples", "c17.bench")),
                     "examples/c17.bench"))
    bc = np.zeros(2048, np.uint32)  # BetaCells: 1,325 words, the rest zero
    p = pyref.Prog()
    L.gen_beta(p.ptr, bc.ctypes.data)
    made.append(save("beta", p, "gen_beta() and its BetaCells", cells=bc))
    for geo in ((3, 3, 3), (4, 4, 4)):
        gc = np.zeros(8192, np.uint8)  # GammaCells: 2,224 bytes, the rest zero
        p = pyref.Prog()
        if not L.gen_gamma(p.ptr, *geo, gc.ctypes.data):
            raise RuntimeError("gen_gamma%r" % (geo,))
        made.append(save("gamma_%d_%d_%d" % geo, p, "gen_gamma(%d, %d, %d) and its GammaCells" % geo,
                         cells=gc, geometry=np.array(geo, np.uint32)))
    total = sum(os.path.getsize(p) for p in made)
    print("%d programs, %.1f KiB" % (len(made), total / 1024))


if __name__ == "__main__":
    main()
