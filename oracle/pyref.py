"""Bindings to the reference's own engine, compiled from its sources by `make -C oracle ref` into
oracle/_ref/libalpha_ref.so. Test infrastructure only: it pins the oracle and gives the CPU
baseline its "reference" kind. Absent (and every caller skips) where the build could not run.

Structs mirrored here follow the reference headers: Program (src/isa.h:36-46), Enum
(src/enum.h:17-24), EnumResult (src/enum.h:26-35), RefMachine (src/ref.h:11-20). t1_run, t1_merge
and t1_print_raw take T1Config / T1Size pointers (src/tester.h:30-50); callers pass any
layout-identical structure.
"""
import ctypes as C
import os

import numpy as np

from oracle.pyoracle import REF_LIB

_r = None


def available():
    return os.path.exists(REF_LIB)


class RefProgram(C.Structure):
    _fields_ = [("code", C.c_void_p), ("len", C.c_size_t), ("ncells", C.c_uint32),
                ("cell_names", C.POINTER(C.c_char_p)), ("inputs", C.POINTER(C.c_uint32)),
                ("ninputs", C.c_size_t), ("outputs", C.POINTER(C.c_uint32)),
                ("noutputs", C.c_size_t), ("special", C.c_void_p)]


class RefEnum(C.Structure):
    _fields_ = [("mode", C.c_int), ("a", C.c_void_p), ("b", C.c_void_p), ("spec", C.c_void_p),
                ("ticks", C.c_uint), ("threads", C.c_int), ("idle", C.c_int),
                ("excluded", C.c_void_p), ("part", C.c_int), ("part_end", C.c_int),
                ("parts", C.c_int)]


class RefEnumResult(C.Structure):
    _fields_ = [("space", C.c_uint64), ("tried", C.c_uint64), ("hits", C.c_uint64),
                ("hit", C.c_uint64 * 16), ("nhit", C.c_int), ("seconds", C.c_double),
                ("ops", C.c_uint64), ("threads", C.c_int)]


class RefMachine(C.Structure):
    _fields_ = [("prog", C.c_void_p), ("mem", C.POINTER(C.c_uint8)), ("rr", C.c_uint8),
                ("pc", C.c_size_t), ("steps", C.c_uint64), ("ticks", C.c_uint64),
                ("lane", C.c_uint), ("cube", C.c_void_p)]


def lib():
    global _r
    if _r is None:
        L = C.CDLL(REF_LIB)
        vp = C.c_void_p
        for name in ("gen_mul", "gen_rank", "gen_wht"):
            getattr(L, name).argtypes = [vp, C.c_int]
        L.gen_hash.argtypes = [vp, C.c_int, C.c_int]
        L.gen_hash.restype = C.c_int
        L.gen_life.argtypes = [vp, C.c_int, C.c_int, vp]
        L.gen_life_cube.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
        L.gen_xorconv.argtypes = [vp]
        L.random_program.argtypes = [vp, C.c_uint32, C.c_size_t, C.c_uint64]
        L.t1_program_seed.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64]
        L.t1_program_seed.restype = C.c_uint64
        L.program_free.argtypes = [vp]
        L.toyhash.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.toyhash.restype = C.c_uint64
        L.spec_parse.argtypes = [C.c_char_p, vp, C.c_int, C.c_char_p, C.c_size_t]
        L.spec_parse.restype = vp
        L.spec_free.argtypes = [vp]
        L.spec_value.argtypes = [vp, C.c_uint64]
        L.spec_value.restype = C.c_uint64
        L.enum_run.argtypes = [C.POINTER(RefEnum), C.POINTER(RefEnumResult), C.c_char_p, C.c_size_t]
        L.enum_run.restype = C.c_int
        L.net_parse.argtypes = [C.c_char_p, C.c_char_p, vp, C.c_char_p, C.c_size_t]
        L.net_parse.restype = C.c_int
        L.ref_cube_init.argtypes = [vp, vp]
        L.ref_cube_init.restype = C.c_int
        L.ref_cube_free.argtypes = [vp]
        L.ref_cube_tick.argtypes = [vp]
        L.ref_cube_step.argtypes = [vp]
        L.ref_cube_machines.argtypes = [vp]
        L.ref_cube_machines.restype = C.c_uint
        L.ref_set.argtypes = [vp, C.c_uint32, C.c_uint8]
        L.ref_get.argtypes = [vp, C.c_uint32]
        L.ref_get.restype = C.c_uint8
        L.benes_route.argtypes = [vp, vp]
        L.t1_cycles.argtypes = [vp, C.c_uint64, C.c_uint64, vp, vp, vp]
        L.t1_cycles.restype = C.c_uint64
        L.cpu_has_avx2.restype = C.c_int
        L.gen_polymul.argtypes = [vp]
        L.ntt_twiddles.argtypes = [vp, vp]
        L.polymul_scalar.argtypes = [vp, vp, vp]
        L.lane_init.argtypes = [vp, vp]
        L.lane_init.restype = C.c_int
        L.lane_free.argtypes = [vp]
        L.jit_compile.argtypes = [vp, C.c_int, vp, vp, C.c_char_p, C.c_size_t]
        L.jit_compile.restype = C.c_int
        L.jit_free.argtypes = [vp]
        L.jit_run.argtypes = [vp, vp, C.c_uint64]
        L.t1_run.argtypes = [vp, vp, C.POINTER(C.c_int)]
        L.t1_run.restype = C.c_double
        L.t1_print_raw.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_uint64]
        L.t1_merge.argtypes = [C.POINTER(C.c_char_p), C.c_int, vp, C.POINTER(C.c_double),
                               C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
        L.t1_merge.restype = C.c_int
        # beta (src/beta.h) and gamma (src/gamma.h), the large cube clients
        L.gen_beta.argtypes = [vp, vp]
        L.beta_reset.argtypes = [vp]
        L.beta_step.argtypes = [vp]
        L.beta_to_lane.argtypes = [vp, vp, C.c_uint, vp]
        L.beta_from_lane.argtypes = [vp, vp, C.c_uint, vp]
        L.gen_gamma.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp]
        L.gen_gamma.restype = C.c_int
        L.gamma_world_init.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.gamma_world_init.restype = C.c_int
        L.gamma_world_free.argtypes = [vp]
        L.gamma_soup.argtypes = [vp, C.c_double, C.c_uint, C.c_uint64]
        L.gamma_soup.restype = C.c_size_t
        L.gamma_predators.argtypes = [vp, C.c_size_t, C.c_uint64]
        L.gamma_predators.restype = C.c_size_t
        L.gamma_step.argtypes = [vp]
        L.gamma_to_cube.argtypes = [vp, vp, vp]
        L.gamma_from_cube.argtypes = [vp, vp, vp]
        L.gamma_compare.argtypes = [vp, vp, C.POINTER(C.c_char_p)]
        L.gamma_compare.restype = C.c_long
        L.lane_cube_init.argtypes = [vp, vp]
        L.lane_cube_init.restype = C.c_int
        L.lane_cube_free.argtypes = [vp]
        _r = L
    return _r


class Prog:
    """A program made by the reference's generators (owned by the reference library)."""

    def __init__(self):
        self.s = RefProgram()

    def __del__(self):
        if _r is not None:
            _r.program_free(C.byref(self.s))

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), C.c_void_p)

    def code(self):
        n = self.s.len
        raw = np.ctypeslib.as_array(C.cast(self.s.code, C.POINTER(C.c_uint32)), shape=(2 * n,)).reshape(n, 2)
        return (raw[:, 0] & 0xFF).astype(np.uint8), raw[:, 1].copy()

    def names(self):
        return [self.s.cell_names[i].decode() for i in range(self.s.ncells)]


def make(name, *args):
    p = Prog()
    getattr(lib(), name)(p.ptr, *args)
    return p


def enum_run(mode, a_ptr, b_ptr=None, spec_text=None, spec_inputs=True, ticks=1, part=0, part_end=0,
             parts=1, threads=0):
    L = lib()
    err = C.create_string_buffer(512)
    # any pointer to a layout-identical program (a reference Prog's or the engine's ts_program)
    a_ptr = C.cast(a_ptr, C.c_void_p)
    b_ptr = None if b_ptr is None else C.cast(b_ptr, C.c_void_p)
    spec = None
    if spec_text is not None:
        spec = L.spec_parse(spec_text.encode(), a_ptr, int(spec_inputs), err, len(err))
        if not spec:
            raise ValueError(err.value.decode())
    e = RefEnum(mode, a_ptr, b_ptr, spec, ticks, threads, 0, None, part, part_end, parts)
    r = RefEnumResult()
    ok = L.enum_run(C.byref(e), C.byref(r), err, len(err))
    if spec:
        L.spec_free(spec)
    if not ok:
        raise ValueError(err.value.decode())
    return {"space": r.space, "tried": r.tried, "hits": r.hits, "hit": list(r.hit[:r.nhit]),
            "nhit": r.nhit, "seconds": r.seconds, "ops": r.ops, "threads": r.threads}


def t1_cycles(prog, seed, max_passes):
    """The reference's cycle search on one group of 512 machines (its jit-compiled code):
    (lambda per lane, passes run)."""
    lam = np.zeros(512, np.uint64)
    aborted = C.c_int()
    passes = lib().t1_cycles(prog.ptr, seed, max_passes, lam.ctypes.data, None, C.byref(aborted))
    return lam, int(passes)


def t1_sample(progs, seeds, max_passes, threads, seconds):
    """CPU baseline of the batch workload: t1_cycles on group 0 of each program, programs taken in
    order by `threads` workers (ctypes releases the GIL) until `seconds` have passed. Returns
    (programs done, 1-bit ops as the reference counts them, seconds)."""
    import threading
    import time
    lock = threading.Lock()
    state = {"next": 0, "done": 0, "ops": 0}
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                k = state["next"]
                if k >= len(progs) or time.perf_counter() - t0 > seconds:
                    return
                state["next"] += 1
            _, ran = t1_cycles(progs[k], seeds[k], max_passes)
            with lock:
                state["done"] += 1
                state["ops"] += ran * progs[k].s.len * 512  # src/tester.c:305
    pool = [threading.Thread(target=worker) for _ in range(threads)]
    for th in pool:
        th.start()
    for th in pool:
        th.join()
    return state["done"], state["ops"], time.perf_counter() - t0


def t1_run(cfg_ptr, out_ptr):
    """The reference's tester 1 on host threads (cfg.threads of them): (wall seconds, threads)."""
    th = C.c_int()
    wall = lib().t1_run(cfg_ptr, out_ptr, C.byref(th))
    return wall, th.value


def t1_print_raw(sizes_ptr, n, seconds, threads, max_passes):
    """The text t1_print_raw writes to stdout (captured through file descriptor 1)."""
    import tempfile
    libc = C.CDLL(None)
    with tempfile.TemporaryFile() as tmp:
        saved = os.dup(1)
        libc.fflush(None)
        os.dup2(tmp.fileno(), 1)
        try:
            lib().t1_print_raw(sizes_ptr, n, seconds, threads, max_passes)
            libc.fflush(None)
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def t1_merge(paths, out_ptr):
    """The reference's merge of raw files: (sizes, seconds, threads, max_passes)."""
    files = (C.c_char_p * len(paths))(*[os.fsencode(p) for p in paths])
    sec, th, mp = C.c_double(), C.c_int(), C.c_uint64()
    n = lib().t1_merge(files, len(paths), out_ptr, C.byref(sec), C.byref(th), C.byref(mp))
    return n, sec.value, th.value, mp.value


def _aligned(nbytes, align=64):
    raw = C.create_string_buffer(nbytes + align)
    off = (-C.addressof(raw)) % align
    return raw, C.addressof(raw) + off


def jit_run_cells(prog_ptr, cells, passes, read):
    """The reference's jit on one 512-machine group: a LaneMachine at power-on (lane_init), the
    lines of `cells` (cell -> 8 u64) written into its memory, `passes` passes by jit_run, timed.
    Returns ({cell: 8 u64 line} for the cells in `read`, seconds of jit_run). LaneMachine is
    src/lanes.h:38-50 (64-byte aligned, its memory pointer at offset 8); JitCode src/jit.h:25-36."""
    L = lib()
    keep_m, m = _aligned(1024)
    keep_j, j = _aligned(512)
    if not L.lane_init(m, prog_ptr):
        raise MemoryError
    err = C.create_string_buffer(512)
    try:
        if not L.jit_compile(prog_ptr, 0, j, None, err, len(err)):
            raise ValueError(err.value.decode())
        mem = C.c_void_p.from_address(m + 8).value
        for cell, line in cells.items():
            line = np.ascontiguousarray(line, dtype=np.uint64)
            C.memmove(mem + 64 * cell, line.ctypes.data, 64)
        import time
        t0 = time.perf_counter()
        L.jit_run(j, m, passes)
        secs = time.perf_counter() - t0
        out = {}
        for cell in read:
            line = np.zeros(8, np.uint64)
            C.memmove(line.ctypes.data, mem + 64 * cell, 64)
            out[cell] = line
        L.jit_free(j)
        return out, secs
    finally:
        L.lane_free(m)
        del keep_m, keep_j


def cube_run(prog_ptr, ncells, mem, rr, passes, steps=0):
    """The reference machine cube (ref_cube_*) from line state, as the oracle's run_lines: whole
    passes (ref_cube_tick), then `steps` single instructions (ref_cube_step)."""
    L = lib()
    n = L.ref_cube_machines(prog_ptr)
    cube = (RefMachine * n)()
    if not L.ref_cube_init(C.cast(cube, C.c_void_p), prog_ptr):
        raise MemoryError
    mem = np.array(mem, dtype=np.uint64)
    rr = np.array(rr, dtype=np.uint64)
    bits = np.unpackbits(mem.view(np.uint8), bitorder="little").reshape(mem.shape[0], ncells, 512)
    rbits = np.unpackbits(rr.view(np.uint8), bitorder="little").reshape(rr.shape[0], 512)
    for k in range(n):
        m = C.cast(C.byref(cube, k * C.sizeof(RefMachine)), C.c_void_p)
        for c in range(2, ncells):
            L.ref_set(m, c, int(bits[k // 512, c, k % 512]))
        cube[k].rr = int(rbits[k // 512, k % 512])
    for _ in range(passes):
        L.ref_cube_tick(C.cast(cube, C.c_void_p))
    for _ in range(steps):
        L.ref_cube_step(C.cast(cube, C.c_void_p))
    out = bits.copy()
    outr = rbits.copy()
    for k in range(n):
        m = C.cast(C.byref(cube, k * C.sizeof(RefMachine)), C.c_void_p)
        for c in range(2, ncells):
            out[k // 512, c, k % 512] = cube[k].mem[c]
        outr[k // 512, k % 512] = cube[k].rr
    L.ref_cube_free(C.cast(cube, C.c_void_p))
    mem2 = np.packbits(out, axis=-1, bitorder="little").view(np.uint64).reshape(mem.shape)
    rr2 = np.packbits(outr, axis=-1, bitorder="little").view(np.uint64).reshape(rr.shape)
    return mem2, rr2


# ---- beta and gamma state through the reference's own packers ---------------------------------

class Beta(C.Structure):
    """src/beta.h Beta: one beta computer's state."""
    _fields_ = [("mem", C.c_uint8 * 256), ("a", C.c_uint8), ("x", C.c_uint8), ("pc", C.c_uint8),
                ("c", C.c_uint8), ("inp", C.c_uint8), ("in_valid", C.c_uint8), ("out", C.c_uint8),
                ("out_strobe", C.c_uint8), ("halted", C.c_uint8), ("waiting", C.c_uint8)]


# src/lanes.h: LaneMachine (prog, mem, then the 64-byte-aligned rr line at offset 64; 192 bytes in
# all) and LaneCube (prog, dims, groups, g)
_LM_SIZE, _LM_MEM, _LM_RR = 192, 8, 64


def _machine_lines(base, ncells):
    """(mem[ncells, 8], rr[8]) of the LaneMachine at address base, copied."""
    mem_ptr = C.c_void_p.from_address(base + _LM_MEM).value
    mem = np.ctypeslib.as_array(C.cast(mem_ptr, C.POINTER(C.c_uint64)), (ncells * 8,))
    rr = np.ctypeslib.as_array(C.cast(base + _LM_RR, C.POINTER(C.c_uint64)), (8,))
    return mem.reshape(ncells, 8).copy(), rr.copy()


def _set_machine_lines(base, ncells, mem, rr):
    mem_ptr = C.c_void_p.from_address(base + _LM_MEM).value
    dst = np.ctypeslib.as_array(C.cast(mem_ptr, C.POINTER(C.c_uint64)), (ncells * 8,))
    dst[:] = np.asarray(mem, np.uint64).reshape(-1)
    np.ctypeslib.as_array(C.cast(base + _LM_RR, C.POINTER(C.c_uint64)), (8,))[:] = rr


def beta_lines(prog_ptr, ncells, cells_ptr, betas):
    """Lines (groups, ncells, 8), (groups, 8) holding the Beta states `betas` (512 per group) in a
    LaneMachine of the program, through the reference's beta_to_lane."""
    L = lib()
    groups = (len(betas) + 511) // 512
    mem = np.zeros((groups, ncells, 8), np.uint64)
    rr = np.zeros((groups, 8), np.uint64)
    for g in range(groups):  # a fresh machine per group: power-on state, then the betas
        m = (C.c_uint8 * _LM_SIZE)()
        if not L.lane_init(m, prog_ptr):
            raise MemoryError
        try:
            for lane in range(512):
                k = 512 * g + lane
                if k < len(betas):
                    L.beta_to_lane(m, cells_ptr, lane, C.byref(betas[k]))
            mem[g], rr[g] = _machine_lines(C.addressof(m), ncells)
        finally:
            L.lane_free(m)
    return mem, rr


def beta_states(prog_ptr, ncells, cells_ptr, mem, rr, count):
    """The Beta states of lines (groups, ncells, 8) through the reference's beta_from_lane."""
    L = lib()
    m = (C.c_uint8 * _LM_SIZE)()
    if not L.lane_init(m, prog_ptr):
        raise MemoryError
    out = []
    try:
        for g in range(rr.shape[0]):
            _set_machine_lines(C.addressof(m), ncells, mem[g], rr[g])
            for lane in range(512):
                if len(out) < count:
                    b = Beta()
                    L.beta_from_lane(m, cells_ptr, lane, C.byref(b))
                    out.append(b)
    finally:
        L.lane_free(m)
    return out


class GammaWorld:
    """A reference GammaWorld (src/gamma.h), held in a buffer the reference fills."""

    def __init__(self, xbits, ybits, zbits):
        self.buf = (C.c_uint8 * 4096)()
        if not lib().gamma_world_init(self.buf, xbits, ybits, zbits):
            raise MemoryError

    def __del__(self):
        if _r is not None:
            _r.gamma_world_free(self.buf)


def gamma_lines(prog_ptr, ncells, cells_ptr, world):
    """The world's cube state as lines (groups, ncells, 8), (groups, 8): gamma_to_cube into the
    reference's LaneCube of the program."""
    L = lib()
    lc = (C.c_uint8 * 64)()
    if not L.lane_cube_init(lc, prog_ptr):
        raise MemoryError
    try:
        L.gamma_to_cube(lc, cells_ptr, world.buf)
        groups = C.c_uint.from_address(C.addressof(lc) + 12).value
        g = C.c_void_p.from_address(C.addressof(lc) + 16).value
        mem = np.zeros((groups, ncells, 8), np.uint64)
        rr = np.zeros((groups, 8), np.uint64)
        for k in range(groups):
            mem[k], rr[k] = _machine_lines(g + k * _LM_SIZE, ncells)
    finally:
        L.lane_cube_free(lc)
    return mem, rr


def gamma_world_of(prog_ptr, ncells, cells_ptr, mem, rr, world):
    """gamma_from_cube of lines (groups, ncells, 8) into `world`."""
    L = lib()
    lc = (C.c_uint8 * 64)()
    if not L.lane_cube_init(lc, prog_ptr):
        raise MemoryError
    try:
        g = C.c_void_p.from_address(C.addressof(lc) + 16).value
        for k in range(rr.shape[0]):
            _set_machine_lines(g + k * _LM_SIZE, ncells, mem[k], rr[k])
        L.gamma_from_cube(lc, cells_ptr, world.buf)
    finally:
        L.lane_cube_free(lc)


def gamma_compare(a, b):
    """(first differing voxel or -1, the part's name) of two worlds."""
    what = C.c_char_p()
    v = lib().gamma_compare(a.buf, b.buf, C.byref(what))
    return v, (what.value.decode() if what.value else "")
