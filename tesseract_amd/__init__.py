"""tesseract_amd: an MI355X engine for the bit-sliced NAND/STO machine.

Python mirror of the C interfaces in include/: programs (written with the program writer, loaded
from .bench netlists or from program files, see tesseract_amd.programs), verify/search specs, and
the GPU engine (gpu_compile, gpu_run, gpu_enum_run, batches, the server). Everything runs in
libtesseract.so; this module only binds it with ctypes. There is no CPU fallback: without the
built library, or without a GPU, the engine calls raise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtesseract.so")

OP_NAND, OP_STO = 0, 1
CELL_ZERO, CELL_ONE, CELL_FIRST_FREE = 0, 1, 2
CK_PLAIN, CK_ID, CK_LINK = 0, 1, 2
EN_VERIFY, EN_EQUIV, EN_SEARCH = 0, 1, 2
G_BUF, G_NOT, G_AND, G_NAND, G_OR, G_NOR, G_XOR, G_XNOR = range(8)
CUBE_DIMS, CUBE_SIZE = 9, 512
BENES_ROUNDS = 17


class Instr(C.Structure):
    _fields_ = [("op", C.c_uint8), ("addr", C.c_uint32)]


class CellInfo(C.Structure):
    _fields_ = [("kind", C.c_uint8), ("dim", C.c_uint8), ("src", C.c_uint32)]


class ProgramStruct(C.Structure):
    """Layout-identical to the reference's Program (src/isa.h:36-46)."""
    _fields_ = [("code", C.POINTER(Instr)), ("len", C.c_size_t), ("ncells", C.c_uint32),
                ("cell_names", C.POINTER(C.c_char_p)), ("inputs", C.POINTER(C.c_uint32)),
                ("ninputs", C.c_size_t), ("outputs", C.POINTER(C.c_uint32)),
                ("noutputs", C.c_size_t), ("special", C.POINTER(CellInfo))]


class EnumResultStruct(C.Structure):
    """Layout-identical to the reference's EnumResult (src/enum.h:26-35)."""
    _fields_ = [("space", C.c_uint64), ("tried", C.c_uint64), ("hits", C.c_uint64),
                ("hit", C.c_uint64 * 16), ("nhit", C.c_int), ("seconds", C.c_double),
                ("ops", C.c_uint64), ("threads", C.c_int)]


class GpuTune(C.Structure):
    _fields_ = [("block", C.c_uint), ("slice_nodes", C.c_uint), ("lut3", C.c_uint), ("grid", C.c_uint),
                ("graph_passes", C.c_uint), ("tape", C.c_uint), ("waves", C.c_uint),
                ("max_waves", C.c_uint), ("sched", C.c_uint), ("group", C.c_uint),
                ("lds_links", C.c_uint)]


class GpuT1Choice(C.Structure):
    _fields_ = [("cells", C.c_uint32), ("residency", C.c_uint), ("rate", C.c_double * 2),
                ("sample_programs", C.c_uint64)]


class GpuEnumInfo(C.Structure):
    _fields_ = [("compile_seconds", C.c_double), ("tape", C.c_uint), ("sched", C.c_uint),
                ("sched_ms", C.c_double * 5), ("sample_words", C.c_uint64)]


class GpuCodeInfo(C.Structure):
    _fields_ = [("segments", C.c_uint32), ("resident", C.c_uint32), ("slices", C.c_uint32),
                ("bitop3", C.c_uint64), ("values", C.c_uint64), ("compile_seconds", C.c_double),
                ("dims", C.c_uint32), ("chain", C.c_uint64)]


class GpuBatchItem(C.Structure):
    _fields_ = [("prog", C.POINTER(ProgramStruct)), ("seed", C.c_uint64),
                ("lam", C.POINTER(C.c_uint64)), ("mem", C.c_void_p), ("rr", C.c_void_p),
                ("lam_group", C.c_uint), ("lam_groups", C.c_uint), ("residency", C.c_uint),
                ("win", C.c_uint), ("words", C.c_uint)]


class GpuBatchStats(C.Structure):
    _fields_ = [("hist", C.c_uint64 * 64), ("lam_sum", C.c_uint64), ("longest", C.c_uint64),
                ("passes", C.c_uint64), ("unfinished", C.c_uint64)]


class GpuBatchTune(C.Structure):
    _fields_ = [("groups", C.c_uint), ("max_passes", C.c_uint), ("plain", C.c_uint),
                ("lds_cells", C.c_uint), ("waves_per_cu", C.c_uint), ("lds_large_cells", C.c_uint),
                ("seconds", C.c_double), ("sink_branch", C.c_uint)]


class GpuBatchInfo(C.Structure):
    _fields_ = [("ms", C.c_double), ("ops", C.c_uint64), ("launches", C.c_uint64),
                ("lds_cells", C.c_uint), ("class_ms", C.c_double * 4), ("class_ops", C.c_uint64 * 4)]


class GpuServeTune(C.Structure):
    _fields_ = [("streams", C.c_uint), ("compile_threads", C.c_uint), ("memory_budget", C.c_uint64),
                ("chunk_words", C.c_uint64), ("max_active", C.c_uint), ("tape_max_ms", C.c_double),
                ("tape_race_ms", C.c_double)]


class GpuRequestTimes(C.Structure):
    _fields_ = [("submitted_s", C.c_double), ("compiled_s", C.c_double), ("admitted_s", C.c_double),
                ("started_s", C.c_double), ("done_s", C.c_double), ("launches", C.c_uint64),
                ("tier", C.c_uint64), ("vgprs", C.c_uint32), ("lds_bytes", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("hbm_bytes", C.c_uint64)]


class GpuTapeStruct(C.Structure):
    _fields_ = [("ops", C.POINTER(C.c_uint32)), ("count", C.POINTER(C.c_uint32)),
                ("leaves", C.POINTER(C.c_uint32)), ("ngroups", C.c_uint32), ("nleaves", C.c_uint32),
                ("root", C.c_uint32), ("slots", C.c_uint32), ("and_nodes", C.c_uint64)]


class T1SizeStruct(C.Structure):
    """Layout-identical to the reference's T1Size (src/tester.h:30-37)."""
    _fields_ = [("cells", C.c_uint32), ("len", C.c_size_t), ("programs", C.c_uint64),
                ("starts", C.c_uint64), ("passes", C.c_uint64), ("unfinished", C.c_uint64),
                ("hist", C.c_uint64 * 64), ("longest", C.c_uint64), ("longest_seed", C.c_uint64)]


class T1ConfigStruct(C.Structure):
    """Layout-identical to the reference's T1Config (src/tester.h:39-50)."""
    _fields_ = [("cells", C.c_uint32 * 32), ("nsizes", C.c_int), ("ratio", C.c_double),
                ("programs", C.c_uint64), ("seconds", C.c_double), ("max_passes", C.c_uint64),
                ("seed", C.c_uint64), ("threads", C.c_int), ("idle", C.c_int),
                ("excluded", C.c_void_p)]


class GpuEnumStruct(C.Structure):
    _fields_ = [("mode", C.c_int), ("a", C.POINTER(ProgramStruct)), ("b", C.POINTER(ProgramStruct)),
                ("spec", C.c_char_p), ("ticks", C.c_uint), ("part", C.c_int), ("part_end", C.c_int),
                ("parts", C.c_int), ("tune", C.POINTER(GpuTune))]


_lib = None


def lib():
    """The native library; raises if it has not been built (python -m tesseract_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtesseract.so is not built: run python -m tesseract_amd.build")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(ProgramStruct)
        vp, sz, u32, u64, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int
        sig = {
            "ts_program_free": (None, [P]),
            "ts_program_dims": (C.c_uint, [P]),
            "ts_program_copy": (i32, [P, P]),
            "ts_random_program": (None, [P, u32, sz, u64]),
            "ts_t1_program_seed": (u64, [u64, u32, u64]),
            "ts_t1_start": (None, [P, u64, C.c_uint, vp, vp]),
            "ts_t1_bucket": (i32, [u64]),
            "ts_splitmix64": (u64, [C.POINTER(u64)]),
            "ts_net_parse": (i32, [C.c_char_p, C.c_char_p, P, C.c_char_p, sz]),
            "ts_spec_parse": (vp, [C.c_char_p, P, i32, C.c_char_p, sz]),
            "ts_spec_free": (None, [vp]),
            "ts_spec_width": (i32, [vp]),
            "ts_spec_value": (u64, [vp, u64]),
            "ts_spec_out_cell": (u32, [vp, i32]),
            "ts_benes_route": (None, [C.POINTER(C.c_uint16), vp]),
            "ts_t1_program_len": (sz, [C.POINTER(T1ConfigStruct), u32]),
            "ts_t1_raw": (vp, [C.POINTER(T1SizeStruct), i32, C.c_char_p, C.c_double, i32, u64]),
            "ts_t1_merge": (i32, [C.POINTER(C.c_char_p), i32, C.POINTER(T1SizeStruct),
                                  C.POINTER(C.c_double), C.POINTER(i32), C.POINTER(u64)]),
            "ts_enum_raw": (vp, [i32, C.POINTER(EnumResultStruct)]),
            "ts_enum_merge": (i32, [C.POINTER(C.c_char_p), i32, C.POINTER(EnumResultStruct),
                                    C.c_char_p, sz]),
            "gpu_t1_run": (i32, [vp, C.POINTER(T1ConfigStruct), C.POINTER(T1SizeStruct),
                                 C.POINTER(C.c_double), C.c_char_p, sz]),
            "gpu_t1_choices": (i32, [vp, C.POINTER(GpuT1Choice), i32]),
            "ts_benes_dim": (i32, [i32]),
            "ts_b_new": (vp, []),
            "ts_b_cell": (u32, [vp, C.c_char_p]),
            "ts_b_has": (i32, [vp, C.c_char_p]),
            "ts_b_link": (u32, [vp, C.c_char_p, u32, i32]),
            "ts_b_input": (None, [vp, u32]),
            "ts_b_output": (None, [vp, u32]),
            "ts_b_nand": (None, [vp, u32]),
            "ts_b_sto": (None, [vp, u32]),
            "ts_b_gate": (None, [vp, i32, u32, C.POINTER(u32), i32]),
            "ts_b_reg": (None, [vp, u32, u32]),
            "ts_b_len": (sz, [vp]),
            "ts_b_finish": (i32, [vp, P, C.c_char_p, sz]),
            "ts_b_half_adder": (None, [vp, u32, u32, u32, u32]),
            "ts_b_full_adder": (None, [vp, u32, u32, u32, u32, u32]),
            "ts_program_from_arrays": (i32, [P, vp, sz, u32, vp, sz, vp, sz, vp, sz, vp, vp, vp,
                                              C.c_char_p, sz]),
            "ts_enum_add_hit": (None, [C.POINTER(EnumResultStruct), u64]),
            "gpu_engine_open": (i32, [i32, C.POINTER(vp), C.c_char_p, sz]),
            "gpu_engine_close": (None, [vp]),
            "gpu_engine_describe": (C.c_char_p, [vp]),
            "gpu_last_ms": (C.c_double, [vp]),
            "gpu_valu_rate": (i32, [vp, C.POINTER(C.c_double), C.c_char_p, sz]),
            "gpu_hbm_rate": (i32, [vp, sz, C.POINTER(C.c_double), C.c_char_p, sz]),
            "gpu_compile": (i32, [vp, P, C.POINTER(GpuTune), C.POINTER(vp), C.POINTER(C.c_char_p),
                                  C.c_char_p, sz]),
            "gpu_code_free": (None, [vp]),
            "gpu_code_stats": (i32, [vp, C.POINTER(GpuCodeInfo)]),
            "gpu_state_new": (i32, [vp, C.c_uint, C.POINTER(vp), C.c_char_p, sz]),
            "gpu_state_free": (None, [vp]),
            "gpu_state_reset": (i32, [vp, C.c_char_p, sz]),
            "gpu_state_put": (i32, [vp, C.c_uint, C.c_uint, vp, vp, C.c_char_p, sz]),
            "gpu_run_steps": (i32, [vp, vp, u64, C.c_char_p, sz]),
            "gpu_state_peek": (i32, [vp, u64, u32, C.POINTER(C.c_uint8), C.c_char_p, sz]),
            "gpu_state_poke": (i32, [vp, u64, u32, C.c_uint8, C.c_char_p, sz]),
            "gpu_state_pc": (u64, [vp]),
            "gpu_state_get": (i32, [vp, C.c_uint, C.c_uint, vp, vp, C.c_char_p, sz]),
            "gpu_run": (i32, [vp, vp, u64, C.c_char_p, sz]),
            "gpu_enum_run": (i32, [vp, C.POINTER(GpuEnumStruct), C.POINTER(EnumResultStruct),
                                   C.c_char_p, sz]),
            "gpu_source": (i32, [P, C.POINTER(GpuTune), C.POINTER(vp), C.POINTER(GpuCodeInfo),
                                 C.c_char_p, sz]),
            "gpu_enum_source": (i32, [C.POINTER(GpuEnumStruct), C.POINTER(vp), C.c_char_p, sz]),
            "gpu_enum_tape": (i32, [C.POINTER(GpuEnumStruct), C.c_uint, C.POINTER(GpuTapeStruct),
                                    C.c_char_p, sz]),
            "gpu_tape_free": (None, [C.POINTER(GpuTapeStruct)]),
            "gpu_rtc_check": (i32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_double),
                                    C.POINTER(sz), C.c_char_p, sz]),
            "gpu_enum_compile": (i32, [vp, C.POINTER(GpuEnumStruct), C.POINTER(vp), C.c_char_p, sz]),
            "gpu_enum_exec": (i32, [vp, C.POINTER(EnumResultStruct), C.c_char_p, sz]),
            "gpu_enum_exec_noalloc": (i32, [vp, C.c_uint, C.POINTER(EnumResultStruct), C.c_char_p,
                                            sz]),
            "gpu_enum_stats": (i32, [vp, C.POINTER(GpuEnumInfo)]),
            "gpu_enum_free": (None, [vp]),
            "gpu_batch_records": (i32, [P, C.c_uint, C.c_uint, C.POINTER(C.c_uint32), sz,
                                        C.POINTER(C.c_size_t), C.c_char_p, sz]),
            "gpu_batch_run": (i32, [vp, C.POINTER(GpuBatchItem), sz, C.POINTER(GpuBatchTune),
                                    C.POINTER(GpuBatchStats), C.POINTER(GpuBatchInfo), C.c_char_p, sz]),
            "gpu_server_open": (i32, [vp, C.POINTER(GpuServeTune), C.POINTER(vp), C.c_char_p, sz]),
            "gpu_server_submit": (i32, [vp, C.POINTER(GpuEnumStruct), C.POINTER(u64), C.c_char_p, sz]),
            "gpu_server_poll": (i32, [vp, u64, C.POINTER(EnumResultStruct), C.POINTER(i32),
                                      C.c_char_p, sz]),
            "gpu_server_wait": (i32, [vp, u64, C.POINTER(EnumResultStruct), C.c_char_p, sz]),
            "gpu_server_times": (i32, [vp, u64, C.POINTER(GpuRequestTimes)]),
            "gpu_server_close": (None, [vp]),
            "gpu_server_release": (i32, [vp, u64, C.c_char_p, sz]),
            "gpu_server_live": (u64, [vp]),
            "free": (None, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _err():
    return C.create_string_buffer(1024)


class Program:
    """A NAND/STO program owned by the native library (a ts_program). Programs loaded from files
    carry their extra arrays (extra) and origin (made_by)."""

    extra = {}
    made_by = None

    def __init__(self):
        self._p = ProgramStruct()
        self._owned = True

    def __del__(self):
        if getattr(self, "_owned", False) and _lib is not None:
            _lib.ts_program_free(C.byref(self._p))
            self._owned = False

    @property
    def ptr(self):
        return C.pointer(self._p)

    @property
    def len(self):
        return int(self._p.len)

    @property
    def ncells(self):
        return int(self._p.ncells)

    def name(self, cell):
        return self._p.cell_names[cell].decode()

    @property
    def inputs(self):
        return [self._p.inputs[i] for i in range(self._p.ninputs)]

    @property
    def outputs(self):
        return [self._p.outputs[i] for i in range(self._p.noutputs)]

    @property
    def has_special(self):
        return bool(self._p.special)

    def kind(self, cell):
        return self._p.special[cell].kind if self._p.special else CK_PLAIN

    def cell_dim(self, cell):
        """The dimension of an id or link cell."""
        return self._p.special[cell].dim if self._p.special else 0

    def code(self):
        """(ops, addrs) as numpy arrays."""
        n = self.len
        if n == 0:
            return np.zeros(0, np.uint8), np.zeros(0, np.uint32)
        raw = np.ctypeslib.as_array(C.cast(self._p.code, C.POINTER(C.c_uint32)), shape=(n * 2,))
        raw = raw.reshape(n, 2)
        return (raw[:, 0] & 0xFF).astype(np.uint8), raw[:, 1].copy()

    def cell(self, name):
        for i in range(self.ncells):
            if self.name(i).lower() == name.lower():
                return i
        raise KeyError(name)

    @property
    def dims(self):
        return int(lib().ts_program_dims(self.ptr))

    @property
    def groups(self):
        """Groups of 512 machines in one cube of this program."""
        return 1 << (self.dims - CUBE_DIMS)


def _make(fn, *args):
    p = Program()
    r = fn(p.ptr, *args)
    return p, r


def random_program(cells, length, seed):
    return _make(lib().ts_random_program, cells, length, seed)[0]


def t1_program_seed(seed, cells, k):
    return int(lib().ts_t1_program_seed(seed, cells, k))


def t1_start(prog, seed, groups=1):
    """Start states (mem [groups, ncells, 8] u64, rr [groups, 8] u64): group 0 = reference t1_start."""
    mem = np.zeros((groups, prog.ncells, 8), np.uint64)
    rr = np.zeros((groups, 8), np.uint64)
    mem[:, CELL_ONE, :] = np.uint64(0xFFFFFFFFFFFFFFFF)
    lib().ts_t1_start(prog.ptr, seed, groups, mem.ctypes.data, rr.ctypes.data)
    return mem, rr


def t1_bucket(lam):
    return int(lib().ts_t1_bucket(lam))


def net_parse(text, name="<netlist>"):
    p = Program()
    e = _err()
    if not lib().ts_net_parse(text.encode(), name.encode(), p.ptr, e, len(e)):
        p._owned = False
        raise ValueError(e.value.decode())
    return p


class Builder:
    """The gate builder (reference src/build.h), for hand-made programs."""

    def __init__(self):
        self._b = lib().ts_b_new()

    def cell(self, name):
        return int(lib().ts_b_cell(self._b, name.encode()))

    def has(self, name):
        return bool(lib().ts_b_has(self._b, name.encode()))

    def link(self, name, src, dim):
        return int(lib().ts_b_link(self._b, name.encode(), src, dim))

    def input(self, c):
        lib().ts_b_input(self._b, c)

    def output(self, c):
        lib().ts_b_output(self._b, c)

    def nand(self, c):
        lib().ts_b_nand(self._b, c)

    def sto(self, c):
        lib().ts_b_sto(self._b, c)

    def gate(self, kind, out, ins):
        arr = (C.c_uint32 * len(ins))(*ins)
        lib().ts_b_gate(self._b, kind, out, arr, len(ins))

    def reg(self, q, d):
        lib().ts_b_reg(self._b, q, d)

    def __len__(self):
        return int(lib().ts_b_len(self._b))

    def half_adder(self, s, k, x, y):
        lib().ts_b_half_adder(self._b, s, k, x, y)

    def full_adder(self, s, k, x, y, c):
        lib().ts_b_full_adder(self._b, s, k, x, y, c)

    def finish(self):
        p = Program()
        e = _err()
        ok = lib().ts_b_finish(self._b, p.ptr, e, len(e))
        self._b = None
        if not ok:
            p._owned = False
            raise ValueError(e.value.decode())
        return p


class Spec:
    def __init__(self, text, prog, inputs_allowed=True):
        e = _err()
        self._s = lib().ts_spec_parse(text.encode(), prog.ptr, int(inputs_allowed), e, len(e))
        if not self._s:
            raise ValueError(e.value.decode())
        self.prog = prog

    def __del__(self):
        if getattr(self, "_s", None) and _lib is not None:
            _lib.ts_spec_free(self._s)

    @property
    def width(self):
        return int(lib().ts_spec_width(self._s))

    def value(self, x):
        return int(lib().ts_spec_value(self._s, x))

    def out_cell(self, bit):
        return int(lib().ts_spec_out_cell(self._s, bit))


def benes_route(dest):
    dest = np.ascontiguousarray(dest, dtype=np.uint16)
    ctrl = np.zeros((BENES_ROUNDS, CUBE_SIZE), np.uint8)
    lib().ts_benes_route(dest.ctypes.data_as(C.POINTER(C.c_uint16)), ctrl.ctypes.data)
    return ctrl


def benes_dim(r):
    return int(lib().ts_benes_dim(r))


def _take_string(ptr):
    s = C.cast(ptr, C.c_char_p).value.decode()
    lib().free(ptr)
    return s


# ---- tester 1 summaries and raw results ------------------------------------------------------

def t1_config(cells=(8, 16, 32, 64, 128, 256), ratio=4.0, programs=64, max_passes=1 << 20, seed=1,
              seconds=0.0):
    """A T1Config with the reference's tester1 defaults (src/main.c:1071-1081)."""
    c = T1ConfigStruct()
    for i, n in enumerate(cells):
        c.cells[i] = n
    c.nsizes = len(cells)
    c.ratio = ratio
    c.programs = programs
    c.seconds = seconds
    c.max_passes = max_passes
    c.seed = seed
    return c


def t1_program_len(cfg, cells):
    return int(lib().ts_t1_program_len(C.byref(cfg), cells))


def t1_size_dict(x):
    return {"cells": x.cells, "len": x.len, "programs": x.programs, "starts": x.starts,
            "passes": x.passes, "unfinished": x.unfinished, "hist": list(x.hist),
            "longest": x.longest, "longest_seed": x.longest_seed}


def _t1_sizes(sizes):
    arr = (T1SizeStruct * max(1, len(sizes)))()
    for i, d in enumerate(sizes):
        for k, v in d.items():
            if k == "hist":
                for b, h in enumerate(v):
                    arr[i].hist[b] = h
            else:
                setattr(arr[i], k, v)
    return arr


def t1_raw(sizes, host, seconds, threads, max_passes):
    """ts_t1_raw: the reference's tester1 --raw text for per-size summaries (dicts)."""
    arr = _t1_sizes(sizes)
    ptr = lib().ts_t1_raw(arr, len(sizes), host.encode(), seconds, threads, max_passes)
    if not ptr:
        raise MemoryError
    return _take_string(ptr)


def t1_merge(paths):
    """ts_t1_merge over raw files: (sizes, seconds, threads, max_passes)."""
    files = (C.c_char_p * len(paths))(*[os.fsencode(p) for p in paths])
    out = (T1SizeStruct * 32)()
    sec, th, mp = C.c_double(), C.c_int(), C.c_uint64()
    n = lib().ts_t1_merge(files, len(paths), out, C.byref(sec), C.byref(th), C.byref(mp))
    if n < 0:
        raise OSError("cannot read a file to merge")
    return [t1_size_dict(out[i]) for i in range(n)], sec.value, th.value, mp.value


def enum_raw(mode, result):
    """ts_enum_raw: the reference's verify/equiv/search --raw text for a result dict."""
    r = EnumResultStruct()
    r.space, r.tried, r.hits = result["space"], result["tried"], result["hits"]
    for i, x in enumerate(result["hit"]):
        r.hit[i] = x
    r.nhit = len(result["hit"])
    r.seconds, r.ops, r.threads = result["seconds"], result["ops"], result.get("threads", 0)
    ptr = lib().ts_enum_raw(mode, C.byref(r))
    if not ptr:
        raise MemoryError
    return _take_string(ptr)


def enum_merge(paths):
    """ts_enum_merge over raw files: the summed result (with its threads)."""
    files = (C.c_char_p * len(paths))(*[os.fsencode(p) for p in paths])
    r = EnumResultStruct()
    e = _err()
    if not lib().ts_enum_merge(files, len(paths), C.byref(r), e, len(e)):
        raise OSError(e.value.decode())
    out = _result(r)
    out["threads"] = r.threads
    return out


def source(prog, tune=None):
    """The HIP source gpu_compile would build for prog, and its code statistics (no device)."""
    out = C.c_void_p()
    info = GpuCodeInfo()
    e = _err()
    if not lib().gpu_source(prog.ptr, _tune(tune), C.byref(out), C.byref(info), e, len(e)):
        raise RuntimeError(e.value.decode())
    return _take_string(out), {f: getattr(info, f) for f, _ in GpuCodeInfo._fields_}


def _enum_struct(mode, a, b, spec, ticks, part, part_end, parts, tune):
    en = GpuEnumStruct()
    en.mode = mode
    en.a = a.ptr
    en.b = b.ptr if b is not None else None
    en.spec = spec.encode() if spec is not None else None
    en.ticks = ticks
    en.part, en.part_end, en.parts = part, part_end, parts
    en.tune = _tune(tune)
    return en


def enum_source(mode, a, b=None, spec=None, ticks=1, part=0, part_end=0, parts=1, tune=None):
    """The HIP source gpu_enum_run would build (no device)."""
    en = _enum_struct(mode, a, b, spec, ticks, part, part_end, parts, tune)
    out = C.c_void_p()
    e = _err()
    if not lib().gpu_enum_source(C.byref(en), C.byref(out), e, len(e)):
        raise ValueError(e.value.decode())
    return _take_string(out)


TAPE_GROUP = 16  # entries per tape group (gpu_tape)


def enum_tape(mode, a, b=None, spec=None, max_slots=0):
    """The interpreter tier's tape of an equiv/search enumeration (no device): a dict of numpy
    arrays ops (groups x 5 x TAPE_GROUP: fields a, b, d, ma, mb), count, leaves and root, slots,
    and_nodes."""
    import numpy as np
    en = _enum_struct(mode, a, b, spec, 1, 0, 0, 1, None)
    t = GpuTapeStruct()
    e = _err()
    if not lib().gpu_enum_tape(C.byref(en), max_slots, C.byref(t), e, len(e)):
        raise ValueError(e.value.decode())
    try:
        g = t.ngroups
        out = {"ops": np.ctypeslib.as_array(t.ops, (g * 5 * TAPE_GROUP,)).reshape(
                   g, 5, TAPE_GROUP).copy() if g else np.zeros((0, 5, TAPE_GROUP), np.uint32),
               "count": np.ctypeslib.as_array(t.count, (g,)).copy() if g else np.zeros(0, np.uint32),
               "leaves": np.ctypeslib.as_array(t.leaves, (t.nleaves,)).copy() if t.nleaves else
               np.zeros(0, np.uint32),
               "root": t.root, "slots": t.slots, "and_nodes": t.and_nodes}
    finally:
        lib().gpu_tape_free(C.byref(t))
    return out


def batch_records(prog, item_words=64, win=16):
    """The batch interpreter's window records of a program (no device): an array of windows x 2 x
    win (load offsets with bit 31 = NAND, then store offsets with bit 31 = nothing stored; bytes of
    cell x 4 x item_words)."""
    import numpy as np
    n = C.c_size_t()
    e = _err()
    if not lib().gpu_batch_records(prog.ptr, item_words, win, None, 0, C.byref(n), e, len(e)):
        raise ValueError(e.value.decode())
    out = np.zeros(n.value, np.uint32)
    if not lib().gpu_batch_records(prog.ptr, item_words, win,
                                   out.ctypes.data_as(C.POINTER(C.c_uint32)), n.value, C.byref(n),
                                   e, len(e)):
        raise ValueError(e.value.decode())
    return out.reshape(-1, 2, win)


def rtc_check(text, arch="gfx950"):
    """hipRTC-compiles HIP source for arch (no device needed); returns (seconds, code bytes)."""
    secs = C.c_double()
    size = C.c_size_t()
    e = _err()
    if not lib().gpu_rtc_check(text.encode(), arch.encode(), C.byref(secs), C.byref(size), e, len(e)):
        raise RuntimeError(e.value.decode())
    return secs.value, size.value


def _result(r):
    return {"space": r.space, "tried": r.tried, "hits": r.hits, "hit": list(r.hit[:r.nhit]),
            "nhit": r.nhit, "seconds": r.seconds, "ops": r.ops}


# ---- the GPU engine --------------------------------------------------------------------------

def _tune(tune):
    if tune is None:
        return None
    t = GpuTune()
    for k, v in tune.items():
        setattr(t, k, v)
    return C.pointer(t)


class Engine:
    """gpu_engine: one per device."""

    def __init__(self, device=0):
        e = _err()
        h = C.c_void_p()
        if not lib().gpu_engine_open(device, C.byref(h), e, len(e)):
            raise RuntimeError("gpu_engine_open: " + e.value.decode())
        self._e = h

    def close(self):
        if self._e:
            lib().gpu_engine_close(self._e)
            self._e = None

    def __del__(self):
        if getattr(self, "_e", None) and _lib is not None:
            _lib.gpu_engine_close(self._e)
            self._e = None

    def describe(self):
        return lib().gpu_engine_describe(self._e).decode()

    def valu_rate(self):
        """Measured 32-bit v_bitop3_b32 lane-operations per second."""
        v = C.c_double()
        e = _err()
        if not lib().gpu_valu_rate(self._e, C.byref(v), e, len(e)):
            raise RuntimeError(e.value.decode())
        return v.value

    def hbm_rate(self, nbytes=1 << 30):
        """Measured bytes/s read + written by a plain copy of nbytes between device buffers."""
        v = C.c_double()
        e = _err()
        if not lib().gpu_hbm_rate(self._e, nbytes, C.byref(v), e, len(e)):
            raise RuntimeError(e.value.decode())
        return v.value

    @property
    def last_ms(self):
        return float(lib().gpu_last_ms(self._e))

    def batch_run(self, progs, seeds, groups=2048, max_passes=64, plain=False, lds_cells=0,
                  waves_per_cu=0, lds_large_cells=0, lams=False, starts=None, lam_groups=None,
                  residency=None, win=None, seconds=0.0, words=None, sink_branch=0):
        """gpu_batch_run: per program a dict(hist, lam_sum, longest, passes, unfinished), with
        lams also "lam" (machine 512 g + k's cycle length, 0 = none found); plus the run info.
        starts: per program None (t1_start of its seed) or (mem, rr) as uint64 arrays
        (groups, ncells, 8), (groups, 8). lam_groups: per program a group g, whose 512 cycle
        lengths alone come back (lams=True). residency, win: per program (0 = automatic; 1 = LDS,
        2 = HBM; 8 or 16 slots per window). seconds > 0 stops the run that long after launch.
        words: per program the words (of 32 machines) of its work items, 16, 32 or 64 (0 = by
        groups). sink_branch: bit c set = class c branches around slots' missing loads and
        stores."""
        import numpy as np
        n = len(progs)
        items = (GpuBatchItem * n)()
        per = 512 if lam_groups is not None else groups * 512
        lam = [np.zeros(per, np.uint64) for _ in range(n)] if lams else None
        keep = []
        for i, (p, sd) in enumerate(zip(progs, seeds)):
            items[i].prog = p.ptr
            items[i].seed = sd
            if starts is not None and starts[i] is not None:
                m = np.ascontiguousarray(starts[i][0], dtype=np.uint64)
                r = np.ascontiguousarray(starts[i][1], dtype=np.uint64)
                if m.shape != (groups, p.ncells, 8) or r.shape != (groups, 8):
                    raise ValueError("a start state is (groups, ncells, 8) and (groups, 8) words")
                keep += [m, r]
                items[i].mem, items[i].rr = m.ctypes.data, r.ctypes.data
            if lams:
                items[i].lam = lam[i].ctypes.data_as(C.POINTER(C.c_uint64))
                if lam_groups is not None:
                    items[i].lam_group, items[i].lam_groups = int(lam_groups[i]), 1
            if residency is not None:
                items[i].residency = int(residency[i])
            if win is not None:
                items[i].win = int(win[i])
            if words is not None:
                items[i].words = int(words[i])
        tune = GpuBatchTune(groups, max_passes, int(plain), lds_cells, waves_per_cu, lds_large_cells,
                            float(seconds), sink_branch)
        stats = (GpuBatchStats * n)()
        info = GpuBatchInfo()
        e = _err()
        if not lib().gpu_batch_run(self._e, items, n, C.byref(tune), stats, C.byref(info), e, len(e)):
            raise RuntimeError("gpu_batch_run: " + e.value.decode())
        out = [{"hist": list(st.hist), "lam_sum": st.lam_sum, "longest": st.longest,
                "passes": st.passes, "unfinished": st.unfinished} for st in stats]
        if lams:
            for o, a in zip(out, lam):
                o["lam"] = a
        return out, {"ms": info.ms, "ops": info.ops, "launches": info.launches,
                     "class_ms": list(info.class_ms), "class_ops": list(info.class_ops)}

    def t1_run(self, cfg):
        """gpu_t1_run: tester 1 on the device; (per-size summaries, seconds)."""
        out = (T1SizeStruct * max(1, cfg.nsizes))()
        sec = C.c_double()
        e = _err()
        if not lib().gpu_t1_run(self._e, C.byref(cfg), out, C.byref(sec), e, len(e)):
            raise ValueError("gpu_t1_run: " + e.value.decode())
        return [t1_size_dict(out[i]) for i in range(cfg.nsizes)], sec.value

    def t1_choices(self):
        """gpu_t1_choices: the residency the last t1_run chose per size, with the sample rates."""
        out = (GpuT1Choice * 32)()
        n = lib().gpu_t1_choices(self._e, out, 32)
        return [{"cells": c.cells, "residency": "LDS" if c.residency == 1 else "HBM",
                 "lds_ops_per_s": c.rate[0], "hbm_ops_per_s": c.rate[1],
                 "sample_programs": c.sample_programs} for c in out[:n]]

    def compile(self, prog, tune=None, listing=False):
        return Code(self, prog, tune, listing)

    def enum_run(self, mode, a, b=None, spec=None, ticks=1, part=0, part_end=0, parts=1, tune=None):
        en = _enum_struct(mode, a, b, spec, ticks, part, part_end, parts, tune)
        r = EnumResultStruct()
        e = _err()
        if not lib().gpu_enum_run(self._e, C.byref(en), C.byref(r), e, len(e)):
            raise ValueError(e.value.decode())
        return _result(r)


class EnumCode:
    """gpu_enum_code: a compiled enumeration that can run many times."""

    def __init__(self, engine, mode, a, b=None, spec=None, ticks=1, part=0, part_end=0, parts=1,
                 tune=None):
        self.engine = engine
        self._keep = (a, b)
        en = _enum_struct(mode, a, b, spec, ticks, part, part_end, parts, tune)
        h = C.c_void_p()
        e = _err()
        if not lib().gpu_enum_compile(engine._e, C.byref(en), C.byref(h), e, len(e)):
            raise ValueError(e.value.decode())
        self._c = h

    def __del__(self):
        if getattr(self, "_c", None) and _lib is not None:
            _lib.gpu_enum_free(self._c)
            self._c = None

    def run(self):
        r = EnumResultStruct()
        e = _err()
        if not lib().gpu_enum_exec(self._c, C.byref(r), e, len(e)):
            raise RuntimeError(e.value.decode())
        return _result(r)

    def stats(self):
        """gpu_enum_stats: compile_seconds, tape, sched (the scheduling strategy kept), sched_ms
        (each timed candidate's sample time), sample_words."""
        i = GpuEnumInfo()
        lib().gpu_enum_stats(self._c, C.byref(i))
        return {"compile_seconds": i.compile_seconds, "tape": i.tape, "sched": i.sched,
                "sched_ms": list(i.sched_ms), "sample_words": i.sample_words}

    def run_noalloc(self, list_words=0):
        """gpu_enum_exec_noalloc: the same enumeration with the hit list in the library's static
        device scratch (at most 4,096 hit words per launch; 0 = all of it)."""
        r = EnumResultStruct()
        e = _err()
        if not lib().gpu_enum_exec_noalloc(self._c, list_words, C.byref(r), e, len(e)):
            raise ValueError(e.value.decode())
        return _result(r)


class Server:
    """gpu_server: continuous-batching request loop on one engine."""

    def __init__(self, engine, streams=0, compile_threads=0, memory_budget=0, chunk_words=0,
                 max_active=0, tape_max_ms=0.0, tape_race_ms=0.0):
        self.engine = engine
        tune = GpuServeTune(streams, compile_threads, memory_budget, chunk_words, max_active,
                            tape_max_ms, tape_race_ms)
        h = C.c_void_p()
        e = _err()
        if not lib().gpu_server_open(engine._e, C.byref(tune), C.byref(h), e, len(e)):
            raise RuntimeError("gpu_server_open: " + e.value.decode())
        self._s = h

    def submit(self, mode, a, b=None, spec=None, ticks=1, part=0, part_end=0, parts=1):
        en = _enum_struct(mode, a, b, spec, ticks, part, part_end, parts, None)
        tk = C.c_uint64()
        e = _err()
        if not lib().gpu_server_submit(self._s, C.byref(en), C.byref(tk), e, len(e)):
            raise RuntimeError("gpu_server_submit: " + e.value.decode())
        return tk.value

    def poll(self, ticket):
        r = EnumResultStruct()
        st = C.c_int()
        e = _err()
        if not lib().gpu_server_poll(self._s, ticket, C.byref(r), C.byref(st), e, len(e)):
            raise RuntimeError(e.value.decode())
        return st.value, (_result(r) if st.value == 4 else None), e.value.decode()

    def wait(self, ticket):
        r = EnumResultStruct()
        e = _err()
        if not lib().gpu_server_wait(self._s, ticket, C.byref(r), e, len(e)):
            raise ValueError(e.value.decode())
        return _result(r)

    def times(self, ticket):
        t = GpuRequestTimes()
        lib().gpu_server_times(self._s, ticket, C.byref(t))
        return {f: getattr(t, f) for f, _ in GpuRequestTimes._fields_}

    def release(self, ticket):
        """Frees a finished request; its ticket is unknown afterwards."""
        e = _err()
        if not lib().gpu_server_release(self._s, ticket, e, len(e)):
            raise ValueError(e.value.decode())

    def live(self):
        """Requests the server holds (submitted, not released)."""
        return int(lib().gpu_server_live(self._s))

    def close(self):
        if self._s:
            lib().gpu_server_close(self._s)
            self._s = None

    def __del__(self):
        if getattr(self, "_s", None) and _lib is not None:
            _lib.gpu_server_close(self._s)
            self._s = None


class Code:
    """gpu_code: one compiled program."""

    def __init__(self, engine, prog, tune=None, listing=False):
        self.engine = engine
        self.prog = prog
        e = _err()
        h = C.c_void_p()
        lst = C.c_char_p()
        if not lib().gpu_compile(engine._e, prog.ptr, _tune(tune), C.byref(h),
                                 C.byref(lst) if listing else None, e, len(e)):
            raise RuntimeError("gpu_compile: " + e.value.decode())
        self._c = h
        self.listing = lst.value.decode() if listing and lst.value else None

    def __del__(self):
        if getattr(self, "_c", None) and _lib is not None:
            _lib.gpu_code_free(self._c)
            self._c = None

    def stats(self):
        i = GpuCodeInfo()
        lib().gpu_code_stats(self._c, C.byref(i))
        return {f: getattr(i, f) for f, _ in GpuCodeInfo._fields_}

    def state(self, groups=None):
        return State(self, groups if groups is not None else self.prog.groups)


class State:
    """gpu_state: device-resident machines of one program."""

    def __init__(self, code, groups):
        self.code = code
        self.groups = groups
        e = _err()
        h = C.c_void_p()
        if not lib().gpu_state_new(code._c, groups, C.byref(h), e, len(e)):
            raise RuntimeError("gpu_state_new: " + e.value.decode())
        self._s = h

    def __del__(self):
        if getattr(self, "_s", None) and _lib is not None:
            _lib.gpu_state_free(self._s)
            self._s = None

    def put(self, mem, rr, group0=0):
        mem = np.ascontiguousarray(mem, dtype=np.uint64)
        rr = np.ascontiguousarray(rr, dtype=np.uint64)
        n = rr.shape[0]
        assert mem.shape == (n, self.code.prog.ncells, 8) and rr.shape == (n, 8)
        e = _err()
        if not lib().gpu_state_put(self._s, group0, n, mem.ctypes.data, rr.ctypes.data, e, len(e)):
            raise RuntimeError("gpu_state_put: " + e.value.decode())

    def get(self, group0=0, ngroups=None):
        n = self.groups - group0 if ngroups is None else ngroups
        mem = np.zeros((n, self.code.prog.ncells, 8), np.uint64)
        rr = np.zeros((n, 8), np.uint64)
        e = _err()
        if not lib().gpu_state_get(self._s, group0, n, mem.ctypes.data, rr.ctypes.data, e, len(e)):
            raise RuntimeError("gpu_state_get: " + e.value.decode())
        return mem, rr

    def reset(self):
        e = _err()
        if not lib().gpu_state_reset(self._s, e, len(e)):
            raise RuntimeError(e.value.decode())

    def run(self, passes):
        e = _err()
        if not lib().gpu_run(self.code._c, self._s, passes, e, len(e)):
            raise RuntimeError("gpu_run: " + e.value.decode())
        return self.code.engine.last_ms

    def run_steps(self, steps):
        """gpu_run_steps: `steps` instructions in every machine from the current pc."""
        e = _err()
        if not lib().gpu_run_steps(self.code._c, self._s, steps, e, len(e)):
            raise RuntimeError("gpu_run_steps: " + e.value.decode())

    @property
    def pc(self):
        return int(lib().gpu_state_pc(self._s))

    def peek(self, machine, cell):
        """gpu_state_peek: one machine's cell (cell == ncells: RR), links resolved."""
        v = C.c_uint8()
        e = _err()
        if not lib().gpu_state_peek(self._s, machine, cell, C.byref(v), e, len(e)):
            raise RuntimeError(e.value.decode())
        return int(v.value)

    def poke(self, machine, cell, value):
        """gpu_state_poke: set one machine's cell (ignored for read-only cells)."""
        e = _err()
        if not lib().gpu_state_poke(self._s, machine, cell, int(value), e, len(e)):
            raise RuntimeError(e.value.decode())
