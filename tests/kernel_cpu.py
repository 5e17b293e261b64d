"""The generated ts_enum kernel run on the host: its source (as gpu_enum_source gives it) compiled
as C++ for the CPU, with plain-C definitions of the three AMDGPU builtins it uses (v_bitop3_b32,
v_bfe_i32, v_perm_b32) and of the hit report, and its threads run one after another. The kernel's
loops are per lane except for the hit report's wave steps, which the host version replaces by a
plain record of every hit word, so the sequential run computes what the device's lanes compute.

This checks code generation (the network split, the spec epilogue, the word and group loops,
shares and ranges) on any machine; the device runs the same source in tests marked gpu."""
import ctypes as C
import hashlib
import os
import subprocess
import tempfile

import numpy as np

CLANG = "/opt/rocm/llvm/bin/clang++"

HOST_PRELUDE = r"""
#include <stdint.h>
#include <vector>
typedef unsigned int u32;
typedef unsigned long long u64;
#define TS_DEV static inline
#define __global__
#define __launch_bounds__(x)
#define __restrict__
#define __attribute__(x)
struct Dim3 { u32 x; };
static Dim3 threadIdx, blockIdx, gridDim;
// v_bitop3_b32: bit k of imm is the result where (S0, S1, S2) = (k >> 2 & 1, k >> 1 & 1, k & 1)
static inline u32 host_bitop3(u32 a, u32 b, u32 c, u32 imm) {
  u32 r = 0;
  for (int k = 0; k < 8; k++)
    if (imm >> k & 1u)
      r |= ((k & 4) ? a : ~a) & ((k & 2) ? b : ~b) & ((k & 1) ? c : ~c);
  return r;
}
// v_bfe_i32: the sign-extended field of w bits at offset o
static inline int host_sbfe(int x, u32 o, u32 w) {
  o &= 31u; w &= 31u;
  if (!w) return 0;
  u32 v = ((u32)x >> o) & ((w == 32u) ? 0xffffffffu : ((1u << w) - 1u));
  if (v >> (w - 1) & 1u) v |= ~((w == 32u) ? 0xffffffffu : ((1u << w) - 1u));
  return (int)v;
}
// v_perm_b32: selector bytes 0-3 take bytes of S1, 4-7 bytes of S0
static inline u32 host_perm(u32 s0, u32 s1, u32 sel) {
  const u64 v = ((u64)s0 << 32) | s1;
  u32 r = 0;
  for (int b = 0; b < 4; b++) {
    u32 s = sel >> (8 * b) & 0xffu, byte;
    if (s < 8) byte = (u32)(v >> (8 * s)) & 0xffu;
    else if (s == 12) byte = 0;
    else byte = 0xffu;
    r |= byte << (8 * b);
  }
  return r;
}
#define __builtin_amdgcn_bitop3_b32(a, b, c, i) host_bitop3((a), (b), (c), (i))
#define __builtin_amdgcn_sbfe(x, o, w) host_sbfe((x), (o), (w))
#define __builtin_amdgcn_perm(a, b, s) host_perm((a), (b), (s))
struct HostHit { u64 gw; u32 mask; };
static std::vector<HostHit> *host_hits;
TS_DEV void ts_report_hits(u64 *, u64 *, u64, u64 gw, u32 hit, u32 *acc, u32 *) {
  if (hit) { *acc += __builtin_popcount(hit); host_hits->push_back({gw, hit}); }
}
TS_DEV void ts_count_hits(u64 *out, u32 acc) { out[0] += acc; }
TS_DEV u32 ts_shl32(u32 a, u32 s) { return s < 32u ? a << s : 0u; }
TS_DEV u64 ts_shl64(u64 a, u64 s) { return s < 64ull ? a << s : 0ull; }
TS_DEV u64 ts_shr64(u64 a, u64 s) { return s < 64ull ? a >> s : 0ull; }
"""

HOST_MAIN = r"""
// every thread of a one-workgroup grid of `block` threads, one after another; the hit words in
// (gw, mask) pairs into hw (at most cap of them), their number returned, the hit count in *hits
extern "C" long long host_enum(u32 block, u64 i0, u64 i1, u64 space, u32 parts, u32 part,
                               u32 nshare, u64 *hits, u64 *hw, u64 cap) {
  std::vector<HostHit> all;
  host_hits = &all;
  u64 out[2] = {0, 0};
  gridDim.x = 1; blockIdx.x = 0;
  for (u32 t = 0; t < block; t++) {
    threadIdx.x = t;
    ts_enum(out, nullptr, 0, i0, i1, space, parts, part, nshare);
  }
  *hits = out[0];
  for (u64 k = 0; k < (u64)all.size() && k < cap; k++) { hw[2 * k] = all[k].gw; hw[2 * k + 1] = all[k].mask; }
  return (long long)all.size();
}
"""

_cache = {}


def build(source):
    """The kernel source compiled for the host: a ctypes library with host_enum."""
    body = source[source.index('extern "C" __global__'):]
    text = HOST_PRELUDE + body + HOST_MAIN
    key = hashlib.sha256(text.encode()).hexdigest()[:16]
    if key in _cache:
        return _cache[key]
    d = tempfile.mkdtemp(prefix="ts_kernel_cpu_")
    src = os.path.join(d, "k.cpp")
    so = os.path.join(d, "k.so")
    with open(src, "w") as f:
        f.write(text)
    subprocess.run([CLANG, "-O2", "-std=c++17", "-w", "-shared", "-fPIC", "-o", so, src],
                   check=True, capture_output=True)
    lib = C.CDLL(so)
    lib.host_enum.restype = C.c_longlong
    lib.host_enum.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32,
                              C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.c_void_p,
                              C.c_uint64]
    _cache[key] = lib
    return lib


def shares(n, part=0, part_end=0, parts=1):
    """space, nshare and active words of a share, as the engine computes them."""
    space = 1 << n
    passes = (space + 511) // 512
    nchunks = (passes + 63) // 64
    nshare = part_end - part + 1
    if parts == 1:
        words = (space + 31) // 32
    else:
        words = ((nchunks + parts - 1) // parts) * nshare * 1024
    return space, nshare, words


def run(source, n, part=0, part_end=0, parts=1, i0=0, i1=None, block=256, cap=1 << 22):
    """(hits, sorted x of every listed hit) of the kernel over active words [i0, i1)."""
    lib = build(source)
    space, nshare, words = shares(n, part, part_end, parts)
    if i1 is None:
        i1 = words
    hits = C.c_uint64()
    hw = np.zeros(2 * cap, np.uint64)
    nw = lib.host_enum(block, i0, i1, space, parts, part, nshare, C.byref(hits),
                       hw.ctypes.data, cap)
    assert nw <= cap, "more hit words than the buffer holds"
    gw, m = hw[0:2 * nw:2], hw[1:2 * nw:2]
    lane = np.arange(32, dtype=np.uint64)
    on = ((m[:, None] >> lane) & np.uint64(1)).astype(bool)
    xs = np.sort((gw[:, None] * np.uint64(32) + lane)[on])
    return hits.value, [int(x) for x in xs]
