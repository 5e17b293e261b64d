// HIP source generation for lowered programs. See lower.hpp.
//
// Mapping: every AND node of a cone gets its 3-feasible cuts enumerated; an area-flow pass picks
// one cut per node and the cover is taken from the roots down, so each chosen cut becomes one
// v_bitop3_b32 whose 8-bit truth table absorbs every complemented edge inside the cut. Values that
// leave the LUT network (stores, cross-lane moves) are kept in the polarity most of their uses
// want, so a NAND result stored to a cell costs no separate NOT.
#include <algorithm>
#include <array>
#include <functional>
#include <map>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <sstream>

#include "lower.hpp"
#include "tesseract_host.h"

namespace ts {

namespace {

struct Cut {
    uint8_t  n = 0;
    uint32_t leaf[3];
    float    flow = 0;
};

// One cone mapped to 3-input LUTs.
struct Mapping {
    std::vector<uint8_t>  in_cone;   // node in the cone of the roots
    std::vector<uint8_t>  is_lut;    // AND node materialised as a LUT
    std::vector<Cut>      best;      // chosen cut per LUT node
    std::vector<uint8_t>  tt;        // truth table per LUT node (already in stored polarity)
    std::vector<uint8_t>  inv;       // the variable holds the complement of the node
};

static bool merge(const Cut &a, const Cut &b, Cut &out)
{
    uint32_t tmp[6];
    int n = 0;
    for (int i = 0; i < a.n; i++)
        tmp[n++] = a.leaf[i];
    for (int i = 0; i < b.n; i++) {
        bool dup = false;
        for (int k = 0; k < a.n; k++)
            dup |= tmp[k] == b.leaf[i];
        if (!dup)
            tmp[n++] = b.leaf[i];
    }
    if (n > 3)
        return false;
    std::sort(tmp, tmp + n);
    out.n = (uint8_t)n;
    for (int i = 0; i < n; i++)
        out.leaf[i] = tmp[i];
    return true;
}

static bool same_cut(const Cut &a, const Cut &b)
{
    if (a.n != b.n)
        return false;
    for (int i = 0; i < a.n; i++)
        if (a.leaf[i] != b.leaf[i])
            return false;
    return true;
}

// Truth table (over leaf patterns 0xF0, 0xCC, 0xAA) of node n above the given cut.
static uint8_t simulate(const Aig &g, uint32_t n, const Cut &c, const std::vector<uint8_t> &inv,
                        std::vector<int16_t> &memo, std::vector<uint32_t> &touched)
{
    static const uint8_t pattern[3] = {0xF0, 0xCC, 0xAA};
    for (int i = 0; i < c.n; i++)
        if (c.leaf[i] == n)
            return (uint8_t)(inv[n] ? ~pattern[i] : pattern[i]);
    if (memo[n] >= 0)
        return (uint8_t)memo[n];
    const Node &nd = g.nodes[n];
    uint8_t v;
    if (nd.kind == N_CONST) {
        v = 0;
    } else {
        uint8_t x = simulate(g, nd.a >> 1, c, inv, memo, touched);
        uint8_t y = simulate(g, nd.b >> 1, c, inv, memo, touched);
        if (nd.a & 1)
            x = (uint8_t)~x;
        if (nd.b & 1)
            y = (uint8_t)~y;
        v = x & y;
    }
    memo[n] = v;
    touched.push_back(n);
    return v;
}

// roots: literals whose values leave the network (stores, hit logic, cross-lane operands).
static void map_luts(const Aig &g, const std::vector<Lit> &roots, bool lut3, Mapping &m)
{
    const size_t N = g.nodes.size();
    m.in_cone.assign(N, 0);
    m.is_lut.assign(N, 0);
    m.best.assign(N, Cut());
    m.tt.assign(N, 0);
    m.inv.assign(N, 0);

    // the cone, through PERM operands too
    std::vector<uint32_t> stack;
    for (Lit r : roots)
        stack.push_back(r >> 1);
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        if (m.in_cone[n])
            continue;
        m.in_cone[n] = 1;
        const Node &nd = g.nodes[n];
        if (nd.kind == N_AND) {
            stack.push_back(nd.a >> 1);
            stack.push_back(nd.b >> 1);
        } else if (nd.kind == N_PERM) {
            stack.push_back(nd.a >> 1);
        }
    }

    // fanout estimates for area flow
    std::vector<uint32_t> fanout(N, 0);
    for (size_t n = 0; n < N; n++) {
        if (!m.in_cone[n])
            continue;
        const Node &nd = g.nodes[n];
        if (nd.kind == N_AND) {
            fanout[nd.a >> 1]++;
            fanout[nd.b >> 1]++;
        } else if (nd.kind == N_PERM) {
            fanout[nd.a >> 1]++;
        }
    }
    // nodes that must be materialised whatever the cut choice
    std::vector<uint8_t> forced(N, 0);
    for (Lit r : roots)
        forced[r >> 1] = 1;
    for (size_t n = 0; n < N; n++)
        if (m.in_cone[n] && g.nodes[n].kind == N_PERM)
            forced[g.nodes[n].a >> 1] = 1;

    // cut enumeration in topological (index) order
    const int KEEP = lut3 ? 8 : 1;
    std::vector<std::vector<Cut>> cuts(N);
    std::vector<float> flow(N, 0.f);
    for (size_t n = 0; n < N; n++) {
        if (!m.in_cone[n])
            continue;
        const Node &nd = g.nodes[n];
        Cut trivial;
        trivial.n = 1;
        trivial.leaf[0] = (uint32_t)n;
        if (nd.kind != N_AND) {
            flow[n] = 0;
            cuts[n].push_back(trivial);
            continue;
        }
        uint32_t a = nd.a >> 1, b = nd.b >> 1;
        std::vector<Cut> cand;
        if (!lut3) {
            Cut c;
            c.n = 2;
            c.leaf[0] = std::min(a, b);
            c.leaf[1] = std::max(a, b);
            if (a == b)
                c.n = 1;
            cand.push_back(c);
        } else {
            for (const Cut &x : cuts[a])
                for (const Cut &y : cuts[b]) {
                    Cut c;
                    if (!merge(x, y, c))
                        continue;
                    bool seen = false;
                    for (const Cut &z : cand)
                        seen |= same_cut(z, c);
                    if (!seen)
                        cand.push_back(c);
                }
        }
        for (Cut &c : cand) {
            float f = 1.f;
            for (int i = 0; i < c.n; i++) {
                uint32_t l = c.leaf[i];
                if (!forced[l])  // a materialised node is a free leaf
                    f += flow[l] / (float)std::max<uint32_t>(1, fanout[l]);
            }
            c.flow = f;
        }
        std::sort(cand.begin(), cand.end(), [](const Cut &x, const Cut &y) {
            return x.flow < y.flow || (x.flow == y.flow && x.n < y.n);
        });
        if ((int)cand.size() > KEEP)
            cand.resize(KEEP);
        m.best[n] = cand.front();
        flow[n] = cand.front().flow;
        cuts[n] = cand;
        cuts[n].push_back(trivial);  // the node itself as a leaf of a bigger cut
    }

    // Area recovery with exact local area (reference-counted cover): refs[n] = LUTs of the cover
    // that read n, plus one for a forced node. Dereferencing a LUT's cut frees every LUT only it
    // kept alive; the exact area of a cut is the number of LUTs its selection adds. Each cover node
    // in topological order takes the cut of least exact area among its kept cuts; two passes.
    std::vector<uint32_t> refs(N, 0);
    std::function<uint32_t(uint32_t, const Cut &)> ref_cut, deref_cut;
    auto is_and = [&](uint32_t l) { return g.nodes[l].kind == N_AND && !forced[l]; };
    ref_cut = [&](uint32_t n, const Cut &c) -> uint32_t {
        uint32_t area = 1;
        for (int i = 0; i < c.n; i++)
            if (c.leaf[i] != n && is_and(c.leaf[i]) && refs[c.leaf[i]]++ == 0)
                area += ref_cut(c.leaf[i], m.best[c.leaf[i]]);
        return area;
    };
    deref_cut = [&](uint32_t n, const Cut &c) -> uint32_t {
        uint32_t area = 1;
        for (int i = 0; i < c.n; i++)
            if (c.leaf[i] != n && is_and(c.leaf[i]) && --refs[c.leaf[i]] == 0)
                area += deref_cut(c.leaf[i], m.best[c.leaf[i]]);
        return area;
    };
    for (size_t n = 0; n < N; n++)
        if (m.in_cone[n] && forced[n] && g.nodes[n].kind == N_AND) {
            refs[n] = 1;
            ref_cut((uint32_t)n, m.best[n]);
        }
    for (int pass = 0; lut3 && pass < 2; pass++)
        for (size_t n = 0; n < N; n++) {
            if (!m.in_cone[n] || g.nodes[n].kind != N_AND || !refs[n])
                continue;
            deref_cut((uint32_t)n, m.best[n]);
            uint32_t best_area = UINT32_MAX;
            size_t best_k = 0;
            for (size_t k = 0; k < cuts[n].size(); k++) {
                const Cut &c = cuts[n][k];
                if (c.n == 1 && c.leaf[0] == n)
                    continue;  // the trivial cut
                uint32_t a = ref_cut((uint32_t)n, c);
                deref_cut((uint32_t)n, c);
                if (a < best_area) {
                    best_area = a;
                    best_k = k;
                }
            }
            m.best[n] = cuts[n][best_k];
            ref_cut((uint32_t)n, m.best[n]);
        }

    // Joint moves for pairs of cover nodes that read the same leaves (a full adder's sum and
    // carry over the two halves of an inner XOR): moving one alone never pays, since the other
    // keeps the shared leaves alive, but moving both onto another cut they both have can free
    // those leaves.
    if (lut3) {
        std::map<std::array<uint32_t, 4>, std::vector<uint32_t>> by_leaves;
        for (size_t n = 0; n < N; n++)
            if (m.in_cone[n] && g.nodes[n].kind == N_AND && refs[n] && m.best[n].n == 3)
                by_leaves[{m.best[n].leaf[0], m.best[n].leaf[1], m.best[n].leaf[2], 3}].push_back(
                    (uint32_t)n);
        for (auto &kv : by_leaves) {
            const std::vector<uint32_t> &grp = kv.second;
            for (size_t i = 0; i + 1 < grp.size(); i += 2) {
                const uint32_t u = grp[i], v = grp[i + 1];
                if (!refs[u] || !refs[v] || !same_cut(m.best[u], m.best[v]))
                    continue;
                deref_cut(u, m.best[u]);
                deref_cut(v, m.best[v]);
                uint32_t cur = ref_cut(u, m.best[u]) + ref_cut(v, m.best[v]);
                deref_cut(v, m.best[v]);
                deref_cut(u, m.best[u]);
                const Cut *pick = nullptr;
                for (const Cut &c : cuts[u]) {
                    if (c.n == 1 && c.leaf[0] == u)
                        continue;
                    bool in_v = false;
                    for (const Cut &d : cuts[v])
                        in_v |= same_cut(c, d);
                    if (!in_v || (c.n == 1 && c.leaf[0] == v))
                        continue;
                    uint32_t a = ref_cut(u, c) + ref_cut(v, c);
                    deref_cut(v, c);
                    deref_cut(u, c);
                    if (a < cur) {
                        cur = a;
                        pick = &c;
                    }
                }
                if (pick) {
                    m.best[u] = *pick;
                    m.best[v] = *pick;
                }
                ref_cut(u, m.best[u]);
                ref_cut(v, m.best[v]);
            }
        }
    }

    // the cover: forced nodes and every AND node a chosen cut reads
    for (size_t n = 0; n < N; n++)
        if (m.in_cone[n] && g.nodes[n].kind == N_AND && (forced[n] || refs[n]))
            m.is_lut[n] = 1;

    // polarity: uses outside the LUT network (roots and PERM operands) decide
    std::vector<int> pos(N, 0), neg(N, 0);
    for (Lit r : roots)
        (r & 1 ? neg : pos)[r >> 1]++;
    for (size_t n = 0; n < N; n++)
        if (m.in_cone[n] && g.nodes[n].kind == N_PERM)
            (g.nodes[n].a & 1 ? neg : pos)[g.nodes[n].a >> 1]++;
    for (size_t n = 0; n < N; n++)
        if (m.is_lut[n] && neg[n] > pos[n])
            m.inv[n] = 1;

    std::vector<int16_t> memo(N, -1);
    std::vector<uint32_t> touched;
    for (size_t n = 0; n < N; n++) {
        if (!m.is_lut[n])
            continue;
        uint8_t t = simulate(g, (uint32_t)n, m.best[n], m.inv, memo, touched);
        for (uint32_t k : touched)
            memo[k] = -1;
        touched.clear();
        m.tt[n] = m.inv[n] ? (uint8_t)~t : t;
    }
}

std::string hex32(uint32_t v)
{
    char b[16];
    snprintf(b, sizeof b, "0x%08xu", v);
    return b;
}

// The word held by wave lane (lane ^ x), 0 < x < 64: DPP quad permutations for x < 4,
// ds_swizzle's xor mode inside 32 lanes, ds_bpermute across the halves.
std::string xl_expr(const std::string &v, uint32_t x)
{
    static const unsigned quad[4] = {0, 0xB1, 0x4E, 0x1B};
    char k[32];
    if (x < 4) {
        snprintf(k, sizeof k, "0x%X", quad[x]);
        return "(u32)__builtin_amdgcn_mov_dpp((int)(" + v + "), " + k + ", 0xF, 0xF, false)";
    }
    if (x < 32) {
        snprintf(k, sizeof k, "0x%X", (x << 10) | 0x1Fu);
        return "(u32)__builtin_amdgcn_ds_swizzle((int)(" + v + "), " + k + ")";
    }
    return "ts_xl_far(" + v + ", " + std::to_string(x) + "u)";
}

// Common device helpers, emitted once per module.
// Hits, reported a wave at a time. Each lane counts its hits in a register (*acc) and the wave
// adds its total to out[0] with one atomic at the end (ts_count_hits); a word with hits takes a
// list entry (word index, 32-bit hit mask), one atomic on out[1] reserving the wave's entries.
// A wave that finds the list full (out[1] at or past the capacity: the caller lists the range
// again in parts) makes sure out[1] is past it and stops reserving (*full, the same in every
// lane). Every lane of the wave takes part; a wave without hits only ballots.
const char *kReportHits = R"(TS_DEV void ts_report_hits(u64 *out, u64 *list, u64 cap, u64 gw, u32 hit, u32 *acc, u32 *full) {
  const u64 who = __ballot(hit != 0u);
  if (!who) return;
  *acc += __builtin_popcount(hit);
  if (*full) return;
  const u64 n = (u64)__builtin_popcountll(who);
  const int leader = __builtin_ctzll(who);
  u64 base = 0;
  if ((int)(threadIdx.x & 63u) == leader) {
    const u64 seen = __atomic_load_n(&out[1], __ATOMIC_RELAXED);
    if (seen < cap) base = atomicAdd(&out[1], n);
    else { base = cap; if (seen == cap) atomicAdd(&out[1], 1ull); }
  }
  base = __shfl(base, leader);
  if (base + n > cap) *full = 1u;
  if (hit) {
    const u64 k = base + __builtin_amdgcn_mbcnt_hi((u32)(who >> 32), __builtin_amdgcn_mbcnt_lo((u32)who, 0u));
    if (k < cap) { list[2 * k] = gw; list[2 * k + 1] = hit; }
  }
}
TS_DEV void ts_count_hits(u64 *out, u32 acc) {
  u64 sum = acc;
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if ((threadIdx.x & 63u) == 0u && sum) atomicAdd(&out[0], sum);
}
)";

const char *kPrelude = R"(typedef unsigned int u32;
typedef unsigned long long u64;
#define TS_DEV __device__ static inline __attribute__((always_inline))
// lane k of a word takes lane k ^ m (m < 32): bit permutations inside the word
TS_DEV u32 ts_pw(u32 v, u32 m) {
  if (m & 1u)  v = ((v & 0x55555555u) << 1) | ((v >> 1) & 0x55555555u);
  if (m & 2u)  v = ((v & 0x33333333u) << 2) | ((v >> 2) & 0x33333333u);
  if (m & 4u)  v = ((v & 0x0F0F0F0Fu) << 4) | ((v >> 4) & 0x0F0F0F0Fu);
  if (m & 8u)  v = ((v & 0x00FF00FFu) << 8) | ((v >> 8) & 0x00FF00FFu);
  if (m & 16u) v = (v << 16) | (v >> 16);
  return v;
}
// the word held by wave lane (lane ^ x), x >= 32: any lane through LDS permute hardware
TS_DEV u32 ts_xl_far(u32 v, u32 x) {
  return (u32)__builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 63u) ^ x) << 2), (int)v);
}
)";

// Emits the mapped cone as statements "u32 vN = ...;". leaf(n) gives the expression of a non-AND
// node (loads, inputs, ids); PERM nodes are emitted here from their operand.
struct Emitter {
    const Aig &g;
    const Mapping &m;
    std::ostringstream &o;
    std::string prefix;  // variable name prefix
    size_t luts = 0, values = 0;
    // lines written right after a node's definition (stores of outputs as soon as they exist)
    const std::unordered_map<uint32_t, std::vector<std::string>> *after = nullptr;
    // when set, only the nodes marked here are defined (the rest are defined elsewhere)
    const std::vector<uint8_t> *only = nullptr;

    Emitter(const Aig &g_, const Mapping &m_, std::ostringstream &o_, std::string p)
        : g(g_), m(m_), o(o_), prefix(std::move(p)) {}

    std::string var(uint32_t n) const { return prefix + std::to_string(n); }

    // expression of literal l as a plain value
    std::string val(Lit l) const
    {
        uint32_t n = l >> 1;
        if (g.nodes[n].kind == N_CONST)
            return (l & 1) ? "0xffffffffu" : "0u";
        bool c = (l & 1) != 0;
        if (g.nodes[n].kind == N_AND && m.inv[n])
            c = !c;
        return c ? "(~" + var(n) + ")" : var(n);
    }

    template <class LeafFn>
    void emit(LeafFn leaf)
    {
        for (size_t n = 1; n < g.nodes.size(); n++) {
            if (!m.in_cone[n] || (only && !(*only)[n]))
                continue;
            const Node &nd = g.nodes[n];
            if (nd.kind == N_AND) {
                if (!m.is_lut[n])
                    continue;
                const Cut &c = m.best[n];
                uint8_t tt = m.tt[n];
                o << "  u32 " << var((uint32_t)n) << " = ";
                if (c.n == 3) {
                    o << "__builtin_amdgcn_bitop3_b32(" << var(c.leaf[0]) << ", " << var(c.leaf[1])
                      << ", " << var(c.leaf[2]) << ", " << (unsigned)tt << "u);\n";
                } else if (c.n == 2) {
                    o << "__builtin_amdgcn_bitop3_b32(" << var(c.leaf[0]) << ", " << var(c.leaf[1])
                      << ", " << var(c.leaf[0]) << ", " << (unsigned)tt << "u);\n";
                } else {
                    // one leaf: the table is 0xF0 (copy) or 0x0F (complement), or constant
                    if (tt == 0xF0)
                        o << var(c.leaf[0]) << ";\n";
                    else if (tt == 0x0F)
                        o << "~" << var(c.leaf[0]) << ";\n";
                    else
                        o << (tt ? "0xffffffffu" : "0u") << ";\n";
                }
                luts++;
            } else if (nd.kind == N_PERM) {
                std::string x = val(nd.a);
                uint32_t inword = nd.b & 31u, lanes = nd.b >> 5;
                std::string e = x;
                if (inword)
                    e = "ts_pw(" + e + ", " + std::to_string(inword) + "u)";
                if (lanes)
                    e = xl_expr(e, lanes);
                o << "  u32 " << var((uint32_t)n) << " = " << e << ";\n";
            } else {
                o << "  u32 " << var((uint32_t)n) << " = " << leaf((uint32_t)n, nd) << ";\n";
            }
            values++;
            if (after) {
                auto it = after->find((uint32_t)n);
                if (it != after->end())
                    for (const std::string &line : it->second)
                        o << line;
            }
        }
    }
};

// The resident pass of a cube that runs in one workgroup of B threads (thread = word of the cube):
// a value read across the cube's waves goes through LDS. A node's level counts the reads across
// waves on its longest path from the start of the pass; level 0 is emitted first, then for each
// level k the values its cross-wave reads take are written to LDS (rows of B words, at most
// `rows` at a time), a barrier, each read of the partner word's row (thread ^ the read's word
// mask), a barrier before the rows are written again, and then the level's other nodes. Every
// thread of the workgroup runs the same straight-line code, so every barrier is met by all.
// `state` names the register holding a cell's value at the start of the pass (cross-wave reads of
// those are level 1). Returns the LDS rows the code uses.
template <class LeafFn>
unsigned emit_levels(const Aig &g, const Mapping &m, Emitter &e, std::ostringstream &o, unsigned B,
                     unsigned rows_max, const std::vector<Version> &versions, LeafFn leaf)
{
    const size_t N = g.nodes.size();
    std::vector<int> lev(N, 0);
    std::vector<uint8_t> cross(N, 0);
    int top = 0;
    for (size_t n = 1; n < N; n++) {
        if (!m.in_cone[n])
            continue;
        const Node &nd = g.nodes[n];
        if (nd.kind == N_AND) {
            if (!m.is_lut[n])
                continue;
            const Cut &c = m.best[n];
            int l = 0;
            for (int i = 0; i < c.n; i++)
                l = std::max(l, lev[c.leaf[i]]);
            lev[n] = l;
        } else if (nd.kind == N_PERM) {
            cross[n] = (nd.b >> 5) >= 64u;
            lev[n] = lev[nd.a >> 1] + cross[n];
        } else if (nd.kind == N_LOAD) {
            cross[n] = (nd.b >> 5) >= 64u;
            lev[n] = cross[n];
        }
        top = std::max(top, lev[n]);
    }
    unsigned rows_used = 0;
    std::vector<uint8_t> only(N, 0);
    const std::vector<uint8_t> *saved = e.only;
    for (int k = 0; k <= top; k++) {
        if (k > 0) {
            // the values this level's cross-wave reads take, each written once
            std::vector<std::string> src;
            std::unordered_map<std::string, unsigned> row_of;
            std::vector<std::pair<uint32_t, unsigned>> reads;  // node, source index
            for (size_t n = 1; n < N; n++) {
                if (!m.in_cone[n] || !cross[n] || lev[n] != k)
                    continue;
                const Node &nd = g.nodes[n];
                const std::string x = nd.kind == N_PERM ? e.var(nd.a >> 1)
                                                        : "s" + std::to_string(versions[nd.a].cell);
                auto it = row_of.find(x);
                if (it == row_of.end()) {
                    it = row_of.emplace(x, (unsigned)src.size()).first;
                    src.push_back(x);
                }
                reads.push_back({(uint32_t)n, it->second});
            }
            for (size_t c0 = 0; c0 < src.size(); c0 += rows_max) {
                const size_t c1 = std::min(src.size(), c0 + rows_max);
                rows_used = std::max<unsigned>(rows_used, (unsigned)(c1 - c0));
                for (size_t j = c0; j < c1; j++)
                    o << "  ts_xs[" << (j - c0) * B << "u + tid] = " << src[j] << ";\n";
                o << "  __syncthreads();\n";
                for (const auto &r : reads) {
                    if (r.second < c0 || r.second >= c1)
                        continue;
                    const Node &nd = g.nodes[r.first];
                    const uint32_t inword = nd.b & 31u, words = nd.b >> 5;
                    std::string x = "ts_xs[" + std::to_string((r.second - c0) * B) + "u + (tid ^ " +
                                    std::to_string(words) + "u)]";
                    if (inword)
                        x = "ts_pw(" + x + ", " + std::to_string(inword) + "u)";
                    // a PERM's operand row holds the operand's variable, which may hold its
                    // complement (and the PERM may read the complemented literal)
                    if (nd.kind == N_PERM && (((nd.a & 1u) != 0) != (m.inv[nd.a >> 1] != 0)))
                        x = "(~" + x + ")";
                    o << "  u32 " << e.var(r.first) << " = " << x << ";\n";
                    e.values++;
                    if (e.after) {
                        auto it = e.after->find(r.first);
                        if (it != e.after->end())
                            for (const std::string &line : it->second)
                                o << line;
                    }
                }
                o << "  __syncthreads();\n";
            }
        }
        for (size_t n = 0; n < N; n++)
            only[n] = m.in_cone[n] && lev[n] == k && !cross[n];
        e.only = &only;
        e.emit(leaf);
    }
    e.only = saved;
    return rows_used;
}

}  // namespace

// ---- run kernels (streamed and resident) ------------------------------------------------------

void codegen_run(const Lowered &lw, const CodegenOptions &opt, KernelSource &ks)
{
    std::ostringstream o;
    o << kPrelude;
    const unsigned cube_wl = lw.dims - 5;  // words per cube = 2^cube_wl
    const char *ptr_of[3] = {"cur", "prev", "init"};
    // the launch bounds, and the waves per SIMD the register allocation must leave room for. A
    // resident program's pass is one long block of straight-line code: left alone the compiler
    // spends up to all 512 registers on it (one wave per SIMD); held to 256 it spills a little and
    // runs two waves per SIMD, 1.5x the throughput on the NTT polynomial product.
    const unsigned waves = opt.waves ? opt.waves : lw.resident ? 2u : 0u;
    const std::string bounds =
        "__launch_bounds__(" + std::to_string(opt.block) + ")" +
        (waves > 1 ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(waves) + ")))"
                   : std::string());

    auto id_expr = [&](uint32_t d) -> std::string {
        static const uint32_t pat[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u,
                                        0xFFFF0000u};
        if (d < 5)
            return hex32(pat[d]);
        return "(0u - ((cw >> " + std::to_string(d - 5) + ") & 1u))";
    };

    if (lw.resident) {
        const Segment &s = lw.segs[0];
        std::vector<Lit> roots;
        for (const SegOut &out : s.outs)
            roots.push_back(out.v);
        Mapping m;
        map_luts(s.g, roots, opt.lut3, m);
        // a whole cube in one workgroup (thread = word of the cube): reads across its waves go
        // through LDS, level by level (emit_levels)
        const bool in_wg = lw.wg_dims > lw.wave_dims;
        const unsigned B = in_wg ? 1u << (lw.wg_dims - 5) : opt.block;
        const unsigned rows_max = in_wg ? (64u * 1024u) / (4u * B) : 0u;
        // a workgroup's waves share the 4 SIMDs of its CU: at least B / 256 of them per SIMD
        const unsigned wg_waves = in_wg ? std::max(waves, B / 256u) : waves;
        const std::string wg_bounds =
            "__launch_bounds__(" + std::to_string(B) + ")" +
            (wg_waves > 1 ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(wg_waves) + ")))"
                          : std::string());
        // state variables carried from pass to pass: the cells whose value at the start of a pass
        // is read (version -1). A cell the pass writes before any read of it is no loop-carried
        // variable: its value of the last pass is stored as the last pass makes it, so no register
        // holds it across the loop.
        std::vector<uint8_t> is_state(lw.ncells + 1, 0);
        for (const Version &v : lw.versions)
            is_state[v.cell] = 1;
        o << "extern \"C\" __global__ void " << (in_wg ? wg_bounds : bounds)
          << " ts_pass(const u32* __restrict__ prev, u32* __restrict__ cur, const u32* "
             "__restrict__ init, u32 nw, u32 passes) {\n"
             "  const u32 w = blockIdx.x * "
          << B << "u + threadIdx.x;\n";
        // (a workgroup is a whole cube and the state whole cubes: no thread is past nw, and all
        // of them reach every barrier)
        if (!in_wg)
            o << "  if (w >= nw) return;\n";
        o << "  const u32 cw = w & " << ((1u << cube_wl) - 1) << "u;\n  (void)cw;\n";
        const size_t xs_at = o.tellp();  // the LDS rows are declared here once their count is known
        if (in_wg)
            o << "  const u32 tid = threadIdx.x;\n";
        o << "  if (passes == 0u) {  // the state as it was\n";
        for (const SegOut &out : s.outs)
            o << "    cur[(size_t)" << lw.final_slot[out.cell] << "u * nw + w] = prev[(size_t)"
              << lw.final_slot[out.cell] << "u * nw + w];\n";
        o << "    return;\n  }\n";
        // every cell read at the start of a pass is loaded before the loop, so that all the
        // loads are in flight at once (issued where they are read, each round of a short program
        // waited for its own)
        for (uint32_t c = 0; c <= lw.ncells; c++) {
            if (!is_state[c])
                continue;
            int32_t fs = lw.final_slot[c];
            if (fs >= 0)
                o << "  u32 s" << c << " = prev[(size_t)" << fs << "u * nw + w];\n";
            else
                o << "  u32 s" << c << " = init[(size_t)" << lw.init_slot[c] << "u * nw + w];\n";
        }
        auto leaf = [&](uint32_t, const Node &nd) -> std::string {
            if (nd.kind == N_LOAD) {
                const Version &ver = lw.versions[nd.a];
                std::string x = "s" + std::to_string(ver.cell);
                uint32_t inword = nd.b & 31u, lanes = nd.b >> 5;
                if (inword)
                    x = "ts_pw(" + x + ", " + std::to_string(inword) + "u)";
                if (lanes)
                    x = xl_expr(x, lanes);
                return x;
            }
            if (nd.kind == N_IDBIT)
                return id_expr(nd.a);
            return "0u";
        };
        // Every pass but the last updates only the carried cells: what else the pass computes is
        // overwritten by the next pass before anything reads it, and the compiler drops it. The
        // last pass runs apart, storing each cell's value the moment it exists, so that no value
        // waits in a register for the end of the pass.
        unsigned rows = 0;
        o << "  for (u32 pass = 1; pass < passes; pass++) {\n";
        {
            Emitter e(s.g, m, o, "v");
            if (in_wg)
                rows = std::max(rows, emit_levels(s.g, m, e, o, B, rows_max, lw.versions, leaf));
            else
                e.emit(leaf);
            // all new values first, then the state update (every read above saw the old state)
            for (size_t i = 0; i < s.outs.size(); i++)
                if (is_state[s.outs[i].cell])
                    o << "    u32 n" << i << " = " << e.val(s.outs[i].v) << ";\n";
            for (size_t i = 0; i < s.outs.size(); i++)
                if (is_state[s.outs[i].cell])
                    o << "    s" << s.outs[i].cell << " = n" << i << ";\n";
        }
        o << "  }\n  {  // the last pass\n";
        Emitter e(s.g, m, o, "v");
        std::unordered_map<uint32_t, std::vector<std::string>> after;
        std::vector<std::string> constant_outs;
        for (const SegOut &out : s.outs) {
            std::string line = "    cur[(size_t)" + std::to_string(lw.final_slot[out.cell]) +
                               "u * nw + w] = " + e.val(out.v) + ";\n";
            if (s.g.nodes[out.v >> 1].kind == N_CONST)
                constant_outs.push_back(line);
            else
                after[out.v >> 1].push_back(line);
        }
        e.after = &after;
        if (in_wg)
            rows = std::max(rows, emit_levels(s.g, m, e, o, B, rows_max, lw.versions, leaf));
        else
            e.emit(leaf);
        for (const std::string &line : constant_outs)
            o << line;
        o << "  }\n}\n";
        ks.names.push_back("ts_pass");
        if (in_wg) {
            ks.block = B;
            if (rows) {
                std::string text = o.str();
                text.insert(xs_at, "  __shared__ u32 ts_xs[" + std::to_string(rows * B) + "];\n");
                o.str(text);
                o.seekp(0, std::ios_base::end);
            }
        }
        ks.bitop3 += e.luts;
        ks.values += e.values;
        ks.chain_luts = e.luts;
        ks.text = o.str();
        return;
    }

    // streamed: one kernel per segment, slices switch on the thread's slice index
    for (size_t si = 0; si < lw.segs.size(); si++) {
        const Segment &s = lw.segs[si];
        std::string name = "ts_seg" + std::to_string(si);
        ks.names.push_back(name);
        o << "extern \"C\" __global__ void " << bounds << " " << name
          << "(const u32* __restrict__ prev, u32* __restrict__ cur, const u32* __restrict__ "
             "init, u32 nw) {\n"
             "  const u32 t = blockIdx.x * "
          << opt.block << "u + threadIdx.x;\n"
          << "  const u32 slice = t / nw, w = t - slice * nw;\n"
          << "  if (slice >= " << s.slices.size() << "u) return;\n"
          << "  const u32 cw = w & " << ((1u << cube_wl) - 1) << "u;\n  (void)cw;\n"
          << "  switch (slice) {\n";
        size_t widest = 0;
        for (size_t k = 0; k < s.slices.size(); k++) {
            std::vector<Lit> roots;
            for (uint32_t oi : s.slices[k])
                roots.push_back(s.outs[oi].v);
            Mapping m;
            map_luts(s.g, roots, opt.lut3, m);
            o << "  case " << k << "u: {\n";
            Emitter e(s.g, m, o, "v");
            e.emit([&](uint32_t, const Node &nd) -> std::string {
                if (nd.kind == N_LOAD) {
                    uint32_t where = lw.ver_where[nd.a], slot = (uint32_t)lw.ver_slot[nd.a];
                    uint32_t inword = nd.b & 31u, wx = nd.b >> 5;
                    std::string idx = wx ? "(w ^ " + std::to_string(wx) + "u)" : "w";
                    std::string x = std::string(ptr_of[where]) + "[(size_t)" +
                                    std::to_string(slot) + "u * nw + " + idx + "]";
                    if (inword)
                        x = "ts_pw(" + x + ", " + std::to_string(inword) + "u)";
                    return x;
                }
                if (nd.kind == N_IDBIT)
                    return id_expr(nd.a);
                return "0u";
            });
            for (uint32_t oi : s.slices[k]) {
                const SegOut &out = s.outs[oi];
                o << "    cur[(size_t)" << out.slot << "u * nw + w] = " << e.val(out.v) << ";\n";
            }
            o << "  } break;\n";
            ks.bitop3 += e.luts;
            ks.values += e.values;
            widest = std::max(widest, e.luts);
        }
        ks.chain_luts += widest;
        o << "  default: break;\n  }\n}\n";
    }
    ks.text = o.str();
}

// ---- exhaustive enumeration ---------------------------------------------------------------

namespace {

// Symbolic execution of a link-free pass. cur[c] holds the literal of every cell (RR = ncells) on
// entry and on exit; written[c] is set for cells the pass stores to.
void execute_linkfree(const ts_program *p, Aig &g, std::vector<Lit> &cur, std::vector<uint8_t> *written)
{
    const uint32_t RR = p->ncells;
    for (size_t i = 0; i < p->len; i++) {
        const ts_instr in = p->code[i];
        if (in.op == TS_OP_STO) {
            if (in.addr >= TS_CELL_FIRST_FREE) {
                cur[in.addr] = cur[RR];
                if (written)
                    (*written)[in.addr] = 1;
            }
            continue;
        }
        Lit v = in.addr == TS_CELL_ZERO ? 0 : in.addr == TS_CELL_ONE ? 1 : cur[in.addr];
        cur[RR] = lit_not(g.And(cur[RR], v));
        if (written)
            (*written)[RR] = 1;
    }
}

// The per-lane expected value for verify, as C source. x bits 0..4 are the lane (compile-time
// constant per unrolled lane), bits >= 5 come from the word index.
struct SpecCode {
    std::string prelude;                 // per word: the var bases
    std::vector<std::string> lane_expr;  // per lane
    bool u32 = false;
    // The result's low bits that no lane bit reaches: with only +, -, *, &, |, ^, ~, negation and
    // left shifts by constants, bit i of the result depends on bits <= i of the operands, so bits
    // below the lowest vector bit fed by a lane (x bits 0..4) are the same in all 32 lanes.
    int lane_free = 0;
    // By the same rule, bits below the lowest vector bit fed by x bits 5..9 (the word within a
    // group of 32 words) are the same in every word of the group.
    int group_free = 0;
};

bool lower_spec(const ts_spec *sp, SpecCode &sc, std::string &err)
{
    const int nv = ts_spec_nvars(sp), nr = ts_spec_nrpn(sp), width = ts_spec_width(sp);
    const ts_spec_item *rpn = ts_spec_rpn(sp);
    // u32 arithmetic gives the same low 32 bits when nothing shifts right and every left shift
    // is by a constant
    bool u32 = width <= 32;
    for (int i = 0; i < nr && u32; i++) {
        if (rpn[i].op == TS_T_SHR)
            u32 = false;
        if (rpn[i].op == TS_T_SHL && (i == 0 || rpn[i - 1].op != TS_T_NUM))
            u32 = false;
    }
    sc.u32 = u32;
    bool low_monotone = true;
    for (int i = 0; i < nr; i++)
        if (rpn[i].op == TS_T_SHR || (rpn[i].op == TS_T_SHL && (i == 0 || rpn[i - 1].op != TS_T_NUM)))
            low_monotone = false;
    if (low_monotone) {
        int m = 64, mg = 64;
        for (int k = 0; k < nv; k++)
            for (int b = 0; b < ts_spec_var_bits(sp, k); b++) {
                const int pos = ts_spec_var_pos(sp, k, b);
                if (pos < 5)
                    m = std::min(m, b);
                else if (pos < 10)
                    mg = std::min(mg, b);
            }
        sc.lane_free = std::min(m, width);
        sc.group_free = std::min(mg, width);
    }
    const char *T = u32 ? "u32" : "u64";
    std::ostringstream pre;
    for (int k = 0; k < nv; k++) {
        int bits = ts_spec_var_bits(sp, k);
        std::string acc = "0ull";
        int b = 0;
        while (b < bits) {
            int p0 = ts_spec_var_pos(sp, k, b);
            if (p0 < 5) {
                b++;
                continue;
            }
            int len = 1, dir = 0;
            if (b + 1 < bits) {
                int p1 = ts_spec_var_pos(sp, k, b + 1);
                if (p1 == p0 + 1 && p1 >= 5)
                    dir = 1;
                else if (p1 == p0 - 1 && p1 >= 5)
                    dir = -1;
            }
            if (dir)
                while (b + len < bits && ts_spec_var_pos(sp, k, b + len) == p0 + dir * len &&
                       ts_spec_var_pos(sp, k, b + len) >= 5)
                    len++;
            unsigned long long mask = len == 64 ? ~0ull : (1ull << len) - 1;
            char term[160];
            if (dir >= 0)
                snprintf(term, sizeof term, "(((xb >> %d) & 0x%llxull) << %d)", p0, mask, b);
            else
                snprintf(term, sizeof term, "(((xr >> %d) & 0x%llxull) << %d)", 63 - p0, mask, b);
            acc += std::string(" | ") + term;
            b += len;
        }
        pre << "    const u64 vb" << k << " = " << acc << ";\n";
    }
    sc.prelude = pre.str();
    sc.lane_expr.resize(32);
    for (int l = 0; l < 32; l++) {
        std::vector<std::string> st;
        for (int i = 0; i < nr; i++) {
            const ts_spec_item &it = rpn[i];
            char b[96];
            switch (it.op) {
            case TS_T_NUM:
                snprintf(b, sizeof b, "((%s)0x%llxull)", T, (unsigned long long)it.v);
                st.push_back(b);
                break;
            case TS_T_VAR: {
                int k = (int)it.v;
                unsigned long long c = 0;
                for (int bit = 0; bit < ts_spec_var_bits(sp, k); bit++) {
                    int p = ts_spec_var_pos(sp, k, bit);
                    if (p < 5 && (l >> p & 1))
                        c |= 1ull << bit;
                }
                snprintf(b, sizeof b, "((%s)(vb%d | 0x%llxull))", T, k, c);
                st.push_back(b);
                break;
            }
            case TS_T_NOT:
                st.back() = "(~" + st.back() + ")";
                break;
            case TS_T_NEG:
                st.back() = "((" + std::string(T) + ")0 - " + st.back() + ")";
                break;
            default: {
                if (st.size() < 2) {
                    err = "internal: bad spec program";
                    return false;
                }
                std::string r = st.back();
                st.pop_back();
                std::string &x = st.back();
                const char *op = it.op == TS_T_ADD   ? "+"
                                 : it.op == TS_T_SUB ? "-"
                                 : it.op == TS_T_MUL ? "*"
                                 : it.op == TS_T_AND ? "&"
                                 : it.op == TS_T_OR  ? "|"
                                 : it.op == TS_T_XOR ? "^"
                                                     : nullptr;
                if (op)
                    x = "(" + x + " " + op + " " + r + ")";
                else if (it.op == TS_T_SHL)
                    x = std::string(u32 ? "ts_shl32(" : "ts_shl64(") + x + ", " + r + ")";
                else
                    x = "ts_shr64(" + x + ", " + r + ")";
            }
            }
        }
        if (st.size() != 1) {
            err = "internal: bad spec program";
            return false;
        }
        sc.lane_expr[l] = st[0];
    }
    return true;
}

// The expected value as B + sum over the lane bits t of x of (bit t) * C_t, when the spec is
// linear in them: sums, differences, negation, complement (~v = -v - 1), products with a factor
// no lane bit reaches and left shifts by such amounts keep the form; &, |, ^ and right shifts
// keep it only between lane-free operands. B and C_t are C expressions over the per-word var
// bases vb<k> (lane bits cleared). dep_c: the x bits the C_t read; tz: bits below it are zero in
// every C_t, so the lanes agree there.
struct SpecLin {
    std::string b, c[5];
    bool has_c[5] = {false, false, false, false, false};
    uint64_t dep_c = 0;
    int tz = 64;
};

bool linear_spec(const ts_spec *sp, bool u32, SpecLin &out)
{
    struct Sym {
        std::string b, c[5];
        bool has[5] = {false, false, false, false, false};
        uint64_t dep_b = 0, dep_c = 0;
        int tz[5] = {64, 64, 64, 64, 64};
        bool lanes() const { return has[0] || has[1] || has[2] || has[3] || has[4]; }
    };
    const int nr = ts_spec_nrpn(sp);
    const ts_spec_item *rpn = ts_spec_rpn(sp);
    const std::string T = u32 ? "u32" : "u64";
    const std::string zero = "((" + T + ")0)";
    std::vector<Sym> st;
    for (int i = 0; i < nr; i++) {
        const ts_spec_item &it = rpn[i];
        char buf[96];
        if (it.op == TS_T_NUM) {
            Sym s;
            snprintf(buf, sizeof buf, "((%s)0x%llxull)", T.c_str(), (unsigned long long)it.v);
            s.b = buf;
            st.push_back(s);
            continue;
        }
        if (it.op == TS_T_VAR) {
            const int k = (int)it.v;
            Sym s;
            s.b = "((" + T + ")vb" + std::to_string(k) + ")";
            for (int bit = 0; bit < ts_spec_var_bits(sp, k); bit++) {
                const int p = ts_spec_var_pos(sp, k, bit);
                if (p < 5) {
                    snprintf(buf, sizeof buf, "((%s)0x%llxull)", T.c_str(), 1ull << bit);
                    s.c[p] = buf;
                    s.has[p] = true;
                    s.tz[p] = bit;
                } else {
                    s.dep_b |= 1ull << p;
                }
            }
            st.push_back(s);
            continue;
        }
        if (it.op == TS_T_NOT || it.op == TS_T_NEG) {
            if (st.empty())
                return false;
            Sym &s = st.back();
            s.b = it.op == TS_T_NOT ? "(~" + s.b + ")" : "(" + zero + " - " + s.b + ")";
            for (int t = 0; t < 5; t++)
                if (s.has[t])
                    s.c[t] = "(" + zero + " - " + s.c[t] + ")";
            continue;
        }
        if (st.size() < 2)
            return false;
        Sym r = st.back();
        st.pop_back();
        Sym &x = st.back();
        switch (it.op) {
        case TS_T_ADD:
        case TS_T_SUB: {
            const char *op = it.op == TS_T_ADD ? " + " : " - ";
            x.b = "(" + x.b + op + r.b + ")";
            for (int t = 0; t < 5; t++) {
                if (r.has[t])
                    x.c[t] = x.has[t] ? "(" + x.c[t] + op + r.c[t] + ")"
                                      : it.op == TS_T_ADD ? r.c[t] : "(" + zero + " - " + r.c[t] + ")";
                x.has[t] = x.has[t] || r.has[t];
                x.tz[t] = std::min(x.tz[t], r.tz[t]);
            }
            x.dep_b |= r.dep_b;
            x.dep_c |= r.dep_c;
            break;
        }
        case TS_T_MUL: {
            if (x.lanes() && r.lanes())
                return false;
            const Sym &lin = x.lanes() ? x : r, &f = x.lanes() ? r : x;
            Sym o;
            o.b = "(" + x.b + " * " + r.b + ")";
            o.dep_b = x.dep_b | r.dep_b;
            for (int t = 0; t < 5; t++)
                if (lin.has[t]) {
                    o.c[t] = "(" + lin.c[t] + " * " + f.b + ")";
                    o.has[t] = true;
                    o.tz[t] = lin.tz[t];
                }
            o.dep_c = lin.dep_c | f.dep_b;
            x = o;
            break;
        }
        case TS_T_SHL: {
            if (r.lanes())
                return false;
            const char *fn = u32 ? "ts_shl32(" : "ts_shl64(";
            const bool num = i > 0 && rpn[i - 1].op == TS_T_NUM;
            x.b = std::string(fn) + x.b + ", " + r.b + ")";
            for (int t = 0; t < 5; t++)
                if (x.has[t]) {
                    x.c[t] = std::string(fn) + x.c[t] + ", " + r.b + ")";
                    x.tz[t] = num ? (int)std::min<uint64_t>(64, x.tz[t] + rpn[i - 1].v) : x.tz[t];
                }
            x.dep_b |= r.dep_b;
            x.dep_c |= r.dep_b;
            break;
        }
        default: {  // &, |, ^, >>: lane-free operands only
            if (x.lanes() || r.lanes())
                return false;
            if (it.op == TS_T_SHR)
                x.b = "ts_shr64(" + x.b + ", " + r.b + ")";
            else
                x.b = "(" + x.b + (it.op == TS_T_AND ? " & " : it.op == TS_T_OR ? " | " : " ^ ") +
                      r.b + ")";
            x.dep_b |= r.dep_b;
        }
        }
    }
    if (st.size() != 1)
        return false;
    const Sym &s = st[0];
    out.b = s.b;
    out.dep_c = s.dep_c;
    out.tz = 64;
    for (int t = 0; t < 5; t++) {
        out.c[t] = s.has[t] ? s.c[t] : zero;
        out.has_c[t] = s.has[t];
        if (s.has[t])
            out.tz = std::min(out.tz, s.tz[t]);
    }
    return true;
}

// 32 x 32 bit transpose of e0..e31 in place: stage j swaps the j-bit blocks of word pairs
// (k, k + j). The 16- and 8-bit stages are byte shuffles, one v_perm_b32 per word (a byte
// selector picks from {S0 = high word, S1 = low word}: 0-3 = bytes of S1, 4-7 = bytes of S0);
// the 4-, 2- and 1-bit stages a shift and a bitwise select (v_bitop3_b32) per word.
void emit_transpose(std::ostringstream &o, const char *e)
{
    static const uint32_t masks[5] = {0x0000FFFFu, 0x00FF00FFu, 0x0F0F0F0Fu, 0x33333333u, 0x55555555u};
    // j = 16: a' = [a0 a1 b0 b1], b' = [a2 a3 b2 b3]; j = 8: a' = [a0 b0 a2 b2], b' = [a1 b1 a3 b3]
    static const uint32_t sel_lo[2] = {0x05040100u, 0x06020400u}, sel_hi[2] = {0x07060302u, 0x07030501u};
    int j = 16;
    for (int st = 0; st < 5; st++, j >>= 1)
        for (int k = 0; k < 32; k++) {
            if (k & j)
                continue;
            const std::string a = e + std::to_string(k), b = e + std::to_string(k + j);
            if (st < 2)
                o << "    { const u32 a = __builtin_amdgcn_perm(" << b << ", " << a << ", "
                  << hex32(sel_lo[st]) << "), b = __builtin_amdgcn_perm(" << b << ", " << a << ", "
                  << hex32(sel_hi[st]) << "); " << a << " = a; " << b << " = b; }\n";
            else  // bitop3 0xD8 = S2 ? S1 : S0 (a bitwise select)
                o << "    { const u32 a = __builtin_amdgcn_bitop3_b32(" << a << ", " << b << " << " << j
                  << ", " << hex32(masks[st] << j) << ", 0xD8), b = __builtin_amdgcn_bitop3_b32(" << b
                  << ", " << a << " >> " << j << ", " << hex32(masks[st]) << ", 0xD8); " << a
                  << " = a; " << b << " = b; }\n";
        }
}

}  // namespace

void enum_outputs(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                  const EnumLowering &el, Aig &g, std::vector<Lit> &oa, std::vector<Lit> &ob,
                  uint32_t *a_end, uint32_t *b_end)
{
    const unsigned n = el.n;
    auto run = [&](const ts_program *p, const std::vector<uint32_t> &in_cells) {
        std::vector<Lit> cur(p->ncells + 1, 0);
        for (unsigned k = 0; k < n; k++)
            cur[in_cells[k]] = g.InBit(n - 1 - k);  // input k = bit n-1-k of x (src/enum.c:113-133)
        cur[TS_CELL_ZERO] = 0;
        cur[TS_CELL_ONE] = 1;
        execute_linkfree(p, g, cur, nullptr);
        return cur;
    };
    oa.clear();
    ob.clear();
    std::vector<Lit> ra = run(a, std::vector<uint32_t>(a->inputs, a->inputs + n));
    if (a_end)
        *a_end = (uint32_t)g.nodes.size();
    if (el.mode == TS_EN_EQUIV) {
        std::vector<uint32_t> b_in;
        for (unsigned k = 0; k < n; k++)
            b_in.push_back(b->inputs[in_map[k]]);
        std::vector<Lit> rb = run(b, b_in);
        for (size_t o = 0; o < a->noutputs; o++) {
            oa.push_back(ra[a->outputs[o]]);
            ob.push_back(rb[b->outputs[out_map[o]]]);
        }
    } else {
        for (uint32_t c : el.out_cells_a)
            oa.push_back(ra[c]);
    }
    if (b_end)
        *b_end = (uint32_t)g.nodes.size();
}

bool codegen_enum(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                  const EnumLowering &el, const CodegenOptions &opt, KernelSource &ks,
                  std::string &err, const ts_spec *spec)
{
    const unsigned n = el.n;
    Aig g;
    const bool loop = el.ticks > 1;
    // outputs to compare: o_k words
    std::vector<uint32_t> cells_a, cells_b;
    if (el.mode == TS_EN_EQUIV) {
        for (size_t o = 0; o < a->noutputs; o++) {
            cells_a.push_back(a->outputs[o]);
            cells_b.push_back(b->outputs[out_map[o]]);
        }
    } else {
        cells_a = el.out_cells_a;
    }

    // ---- the network
    std::vector<Lit> ra, rb;
    std::vector<uint8_t> wa(a->ncells + 1, 0), wb;
    std::vector<uint32_t> b_in;
    if (el.mode == TS_EN_EQUIV)
        for (unsigned k = 0; k < n; k++)
            b_in.push_back(b->inputs[in_map[k]]);
    std::vector<Lit> out_a, out_b;  // ticks == 1: the compared outputs
    if (!loop) {
        enum_outputs(a, b, in_map, out_map, el, g, out_a, out_b, nullptr, nullptr);
    } else {
        auto init_cells = [&](const ts_program *p, std::vector<Lit> &cur, uint32_t base) {
            cur.assign(p->ncells + 1, 0);
            for (uint32_t c = 0; c <= p->ncells; c++)
                cur[c] = g.Load(base + c, 0);
            cur[TS_CELL_ZERO] = 0;
            cur[TS_CELL_ONE] = 1;
        };
        init_cells(a, ra, 0);
        execute_linkfree(a, g, ra, &wa);
        if (el.mode == TS_EN_EQUIV) {
            wb.assign(b->ncells + 1, 0);
            init_cells(b, rb, a->ncells + 1);
            execute_linkfree(b, g, rb, &wb);
        }
    }

    // roots: ticks == 1 the compared outputs; ticks > 1 every written cell (state update)
    std::vector<Lit> roots;
    if (!loop) {
        roots = out_a;
        roots.insert(roots.end(), out_b.begin(), out_b.end());
    } else {
        for (uint32_t c = 0; c <= a->ncells; c++)
            if (wa[c])
                roots.push_back(ra[c]);
        for (uint32_t c = 0; el.mode == TS_EN_EQUIV && c <= b->ncells; c++)
            if (wb[c])
                roots.push_back(rb[c]);
    }
    Mapping m;
    map_luts(g, roots, opt.lut3, m);

    // ---- groups of 32 words. x bits 0..4 are the lanes of a word; a thread that takes the 32
    // consecutive words of a group sees x bits 5..9 run through all 32 values and the bits above
    // stay fixed. A LUT whose inputs reach no lane bit is evaluated once per group on group planes
    // (bit w = its value in word w, the same in all 32 lanes of that word), and a word takes bit w
    // of each such plane its lane-dependent LUTs read as a sign-extended one-bit field. cls: 0 =
    // the same in every word of the group (x bits >= 10 only), 1 = varies with x bits 5..9,
    // 2 = reaches a lane bit.
    std::vector<uint8_t> cls;
    size_t group_luts = 0;
    bool grouped = false;
    if (!loop && n >= 24 && opt.group != 1) {
        cls.assign(g.nodes.size(), 0);
        for (size_t k = 1; k < g.nodes.size(); k++) {
            if (!m.in_cone[k])
                continue;
            const Node &nd = g.nodes[k];
            if (nd.kind == N_INBIT) {
                cls[k] = nd.a < 5 ? 2 : nd.a < 10 ? 1 : 0;
            } else if (nd.kind == N_AND && m.is_lut[k]) {
                uint8_t c = 0;
                for (int l = 0; l < m.best[k].n; l++)
                    c = std::max(c, cls[m.best[k].leaf[l]]);
                cls[k] = c;
                if (c < 2)
                    group_luts++;
            }
        }
        grouped = opt.group == 2 || group_luts >= 32;
    }
    ks.group = grouped ? 32 : 1;

    // verify: the per-lane expected values, and in groups the linear form when its lane terms
    // read only x bits >= 10 (the same in every word of a group)
    SpecCode sc;
    SpecLin sl;
    bool lin = false;
    if (el.mode == TS_EN_VERIFY) {
        if (!lower_spec(spec, sc, err))
            return false;
        lin = grouped && linear_spec(spec, sc.u32, sl) && (sl.dep_c & 0x3E0ull) == 0;
    }
    const int lin_free = lin ? std::min(sl.tz, el.width) : 0;  // bits every lane of a word shares

    std::ostringstream o;
    o << kPrelude << kReportHits;
    o << "TS_DEV u32 ts_shl32(u32 a, u32 s) { return s < 32u ? a << s : 0u; }\n"
         "TS_DEV u64 ts_shl64(u64 a, u64 s) { return s < 64ull ? a << s : 0ull; }\n"
         "TS_DEV u64 ts_shr64(u64 a, u64 s) { return s < 64ull ? a >> s : 0ull; }\n";
    // Word indices: 32 bits while the space has at most 2^36 assignments (i1 <= 2^31, so a
    // round's base plus the grid's stride stays below 2^32), else 64. The loop is wave-uniform:
    // every lane of a wave takes the same rounds (a lane past i1 has no valid assignment), so the
    // hit report's cross-lane steps see every lane.
    const bool narrow = n <= 36;
    const char *W = narrow ? "u32" : "u64";
    o << "extern \"C\" __global__ void __launch_bounds__(" << opt.block << ")"
      << (opt.waves > 1 ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(opt.waves) + ")))"
                        : std::string())
      << " ts_enum(u64* __restrict__ out, u64* __restrict__ list, u64 cap, u64 i0_, u64 i1_, "
         "u64 space, u32 parts, u32 part, u32 nshare) {\n"
      << "  const " << W << " i0 = (" << W << ")i0_, i1 = (" << W << ")i1_;\n"
      << "  u32 hits = 0u, full = 0u;  // this lane's hits; the hit list is full\n";
    // input bit j of x: j < 5 lane pattern, else a bit of the word index (a sign-extended
    // one-bit field: 0 or all ones)
    static const uint32_t pat[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u, 0xFF00FF00u, 0xFFFF0000u};
    auto wordbit = [&](const char *word, uint32_t j) -> std::string {
        if (narrow || j - 5 < 32)
            return "(u32)__builtin_amdgcn_sbfe((int)(u32)" + std::string(word) + ", " +
                   std::to_string(j - 5) + "u, 1u)";
        return "(u32)__builtin_amdgcn_sbfe((int)(u32)((u64)" + std::string(word) + " >> 32), " +
               std::to_string(j - 37) + "u, 1u)";
    };
    auto inbit = [&](uint32_t j) -> std::string {
        return j < 5 ? hex32(pat[j]) : wordbit("gw", j);
    };
    // active word index I -> word G of the whole space: a share takes every parts-th chunk of
    // 64 passes of 512 machines (1024 words)
    auto emit_chunk_map = [&](const std::string &I, const std::string &G) {
        o << "    if (parts > 1u) {  // active chunk -> chunk of the space\n"
          << "      const " << W << " ci = " << I << " >> 10, q = ci / nshare, r = part + ci % nshare;\n"
          << "      " << G << " = ((q * parts + r) << 10) | (" << I << " & 1023u);\n"
          << "    }\n";
    };

    std::vector<std::string> oa, ob;  // expressions of the compared output words
    std::vector<uint8_t> hoisted(el.width, 0);  // verify bits compared once per group (gmiss)
    if (grouped) {
        o << "  const " << W << " g0 = i0 >> 5, g1 = (i1 + 31u) >> 5;\n"
          << "  for (" << W << " gbase = g0 + (" << W << ")blockIdx.x * " << opt.block
          << "u + (threadIdx.x & ~63u); gbase < g1; gbase += (" << W << ")gridDim.x * "
          << opt.block << "u) {\n"
          << "    const " << W << " ib = (gbase + (threadIdx.x & 63u)) << 5;  // the group's first word\n"
          << "    " << W << " gwb = ib;\n";
        emit_chunk_map("ib", "gwb");
        // the group's part of the network on group planes
        std::vector<uint8_t> in_group(g.nodes.size(), 0), in_word(g.nodes.size(), 0);
        for (size_t k = 1; k < g.nodes.size(); k++)
            if (m.in_cone[k]) {
                const bool def = g.nodes[k].kind == N_INBIT || m.is_lut[k];
                (cls[k] < 2 ? in_group : in_word)[k] = def;
            }
        auto gleaf = [&](uint32_t, const Node &nd) -> std::string {
            if (nd.kind == N_INBIT)
                return nd.a < 10 ? hex32(pat[nd.a - 5]) : wordbit("gwb", nd.a);
            return "0u";
        };
        Emitter eg(g, m, o, "g");
        eg.only = &in_group;
        eg.emit(gleaf);
        ks.bitop3 += eg.luts;
        if (lin && lin_free < el.width) {
            // T(k) = sum of C_t over the set bits t of lane k, the same in every word of the
            // group; tp<i> = bit i of T(k) in lane k
            const char *T = sc.u32 ? "u32" : "u64";
            o << "    u32 tp" << lin_free;
            for (int bit = lin_free + 1; bit < el.width; bit++)
                o << ", tp" << bit;
            o << ";\n    {\n"
              << "    const u64 xb = (u64)gwb << 5;\n"
              << "    const u64 xr = __builtin_bitreverse64(xb);\n    (void)xr;\n"
              << sc.prelude;
            for (int t = 0; t < 5; t++)
                o << "    const " << T << " c" << t << " = " << sl.c[t] << ";\n";
            o << "    const " << T << " k0 = 0;\n";
            for (int k = 1; k < 32; k++)
                o << "    const " << T << " k" << k << " = k" << (k & (k - 1)) << " + c"
                  << __builtin_ctz(k) << ";\n";
            for (int half = 0; half < 2; half++) {
                const int lo = std::max(lin_free, 32 * half), hi = std::min(el.width, 32 * half + 32);
                if (lo >= hi)
                    continue;
                o << "    {\n";
                for (int k = 0; k < 32; k++)
                    o << "    u32 t" << k << " = (u32)(k" << k << (half ? " >> 32" : "") << ");\n";
                emit_transpose(o, "t");
                for (int bit = lo; bit < hi; bit++)
                    o << "    tp" << bit << " = t" << bit - 32 * half << ";\n";
                o << "    }\n";
            }
            o << "    }\n";
        }
        // verify bits that no lane and no word of the group reaches (below lin_free and
        // group_free), against output planes that are the same in every word of the group: one
        // compare per group instead of one per word
        if (lin) {
            const char *T = sc.u32 ? "u32" : "u64";
            const int gf = std::min(lin_free, sc.group_free);
            std::vector<int> bits;
            for (int bit = 0; bit < gf; bit++)
                if (cls[out_a[bit] >> 1] == 0)
                    bits.push_back(bit);
            if (!bits.empty()) {
                o << "    u32 gmiss = 0u;\n    {\n"
                  << "    const u64 xb = (u64)gwb << 5;\n"
                  << "    const u64 xr = __builtin_bitreverse64(xb);\n    (void)xr;\n"
                  << sc.prelude << "    const " << T << " bg = " << sl.b << ";\n";
                for (int bit : bits) {
                    o << "    gmiss = __builtin_amdgcn_bitop3_b32(gmiss, (u32)__builtin_amdgcn_sbfe("
                         "(int)(u32)(bg"
                      << (bit >= 32 ? " >> 32" : "") << "), " << bit % 32 << "u, 1u), "
                      << eg.val(out_a[bit]) << ", 0xF6u);\n";
                    hoisted[bit] = 1;
                }
                o << "    }\n";
            }
        }
        // the words of the group inside [i0, i1) and inside the space: bit w for word ib + w
        o << "    u32 gvalid;\n    {\n"
          << "    const u64 sw = (space + 31ull) >> 5;  // words of the space\n"
          << "    const u64 lo = (u64)i0 > (u64)ib ? (u64)i0 - (u64)ib : 0ull;\n"
          << "    u64 hi = (u64)i1 > (u64)ib ? (u64)i1 - (u64)ib : 0ull;\n"
          << "    const u64 hs = sw > (u64)gwb ? sw - (u64)gwb : 0ull;\n"
          << "    hi = hi < hs ? hi : hs;\n"
          << "    hi = hi < 32ull ? hi : 32ull;\n"
          << "    gvalid = hi > lo ? (u32)(((1ull << hi) - 1ull) & ~((1ull << lo) - 1ull)) : 0u;\n"
          << "    }\n";
        // the group values a word reads: LUT leaves of its own part and compared outputs
        std::vector<uint8_t> import(g.nodes.size(), 0);
        for (size_t k = 1; k < g.nodes.size(); k++)
            if (in_word[k] && g.nodes[k].kind == N_AND)
                for (int l = 0; l < m.best[k].n; l++)
                    if (cls[m.best[k].leaf[l]] < 2)
                        import[m.best[k].leaf[l]] = 1;
        for (Lit v : out_a)
            if (g.nodes[v >> 1].kind != N_CONST && cls[v >> 1] < 2)
                import[v >> 1] = 1;
        for (Lit v : out_b)
            if (g.nodes[v >> 1].kind != N_CONST && cls[v >> 1] < 2)
                import[v >> 1] = 1;
        o << "    #pragma unroll 1\n"
          << "    for (u32 w = 0; w < 32u; w++) {\n"
          << "    const " << W << " gw = gwb + w;\n"
          << "    const u64 xb = (u64)gw << 5;\n"
          << "    const u32 valid = (u32)__builtin_amdgcn_sbfe((int)gvalid, w, 1u);\n";
        for (size_t k = 1; k < g.nodes.size(); k++)
            if (import[k])
                o << "    const u32 v" << k << " = "
                  << (cls[k] == 1 ? "(u32)__builtin_amdgcn_sbfe((int)g" + std::to_string(k) + ", w, 1u)"
                                  : "g" + std::to_string(k))
                  << ";\n";
        Emitter ew(g, m, o, "v");
        ew.only = &in_word;
        ew.emit([&](uint32_t, const Node &nd) -> std::string {
            return nd.kind == N_INBIT ? hex32(pat[nd.a]) : std::string("0u");
        });
        ks.bitop3 += ew.luts;
        for (Lit v : out_a)
            oa.push_back(ew.val(v));
        for (Lit v : out_b)
            ob.push_back(ew.val(v));
    } else {
        o << "  for (" << W << " base = i0 + (" << W << ")blockIdx.x * " << opt.block
          << "u + (threadIdx.x & ~63u); base < i1; base += (" << W << ")gridDim.x * " << opt.block
          << "u) {\n"
          << "    const " << W << " i = base + (threadIdx.x & 63u);\n"
          << "    " << W << " gw = i;\n";
        emit_chunk_map("i", "gw");
        o << "    const u64 xb = (u64)gw << 5;\n";
        if (n >= 5)  // the space is a multiple of 32: a word is wholly valid or wholly not
            o << "    const u32 valid = i < i1 && xb < space ? 0xffffffffu : 0u;\n";
        else
            o << "    const u32 valid = i >= i1 || xb >= space ? 0u\n"
                 "                      : space - xb >= 32ull ? 0xffffffffu : ((1u << (u32)(space - xb)) - 1u);\n";
        auto leaf = [&](uint32_t, const Node &nd) -> std::string {
            if (nd.kind == N_INBIT)
                return inbit(nd.a);
            if (nd.kind == N_LOAD)
                return "s" + std::to_string(nd.a);
            return "0u";
        };
        if (!loop) {
            Emitter e(g, m, o, "v");
            e.emit(leaf);
            ks.bitop3 += e.luts;
            for (Lit v : out_a)
                oa.push_back(e.val(v));
            for (Lit v : out_b)
                ob.push_back(e.val(v));
        } else {
            // state s<c> (a) and s<ncells_a + 1 + c> (b): power-on with the inputs set
            auto init_state = [&](const ts_program *p, const uint32_t *in_cells, uint32_t base) {
                for (uint32_t c = 0; c <= p->ncells; c++) {
                    std::string v = "0u";
                    if (c == TS_CELL_ONE)
                        v = "0xffffffffu";
                    for (unsigned k = 0; k < n; k++)
                        if (in_cells[k] == c && c >= TS_CELL_FIRST_FREE)
                            v = inbit(n - 1 - k);
                    o << "    u32 s" << base + c << " = " << v << ";\n";
                }
            };
            init_state(a, a->inputs, 0);
            if (el.mode == TS_EN_EQUIV)
                init_state(b, b_in.data(), a->ncells + 1);
            o << "    for (u32 tick = 0; tick < " << el.ticks << "u; tick++) {\n";
            Emitter e(g, m, o, "v");
            e.emit(leaf);
            ks.bitop3 += e.luts;
            std::vector<std::pair<uint32_t, std::string>> upd;
            for (uint32_t c = 0; c <= a->ncells; c++)
                if (wa[c])
                    upd.push_back({c, e.val(ra[c])});
            for (uint32_t c = 0; el.mode == TS_EN_EQUIV && c <= b->ncells; c++)
                if (wb[c])
                    upd.push_back({a->ncells + 1 + c, e.val(rb[c])});
            for (size_t i = 0; i < upd.size(); i++)
                o << "      u32 n" << i << " = " << upd[i].second << ";\n";
            for (size_t i = 0; i < upd.size(); i++)
                o << "      s" << upd[i].first << " = n" << i << ";\n";
            o << "    }\n";
            for (uint32_t c : cells_a)
                oa.push_back("s" + std::to_string(c));
            for (uint32_t c : cells_b)
                ob.push_back("s" + std::to_string(a->ncells + 1 + c));
        }
    }

    // hit word
    o << "    u32 hit;\n";
    if (el.mode == TS_EN_SEARCH) {
        o << "    hit = valid";
        for (int bit = 0; bit < el.width; bit++)
            o << " & " << ((el.target >> bit & 1) ? "" : "~") << oa[bit];
        o << ";\n";
    } else if (el.mode == TS_EN_EQUIV) {
        o << "    hit = 0u";
        for (size_t k = 0; k < oa.size(); k++)
            o << " | (" << oa[k] << " ^ " << ob[k] << ")";
        o << ";\n    hit &= valid;\n";
    } else {
        o << "    const u64 xr = __builtin_bitreverse64(xb);\n    (void)xr;\n" << sc.prelude;
        const char *T = sc.u32 ? "u32" : "u64";
        // mismatches gathered in four accumulators, one v_bitop3_b32 per bit: m | (t ^ o), truth
        // table 0xF6 (bit 4 S0 + 2 S1 + S2 of the table is the result for inputs S0, S1, S2)
        o << "    u32 miss0 = 0u, miss1 = 0u, miss2 = 0u, miss3 = 0u;\n";
        int acc = 0;
        auto gather = [&](const std::string &t, const std::string &out) {
            o << "    miss" << acc << " = __builtin_amdgcn_bitop3_b32(miss" << acc << ", " << t
              << ", " << out << ", 0xF6u);\n";
            acc = (acc + 1) % 4;
        };
        auto bit_of = [](const std::string &v, int bit) {
            return "(u32)__builtin_amdgcn_sbfe((int)(u32)(" + v + (bit >= 32 ? " >> 32" : "") +
                   "), " + std::to_string(bit % 32) + "u, 1u)";
        };
        if (lin) {
            // expected value of lane k = B + T(k): B the same in all lanes of the word, T's bit
            // planes from the group; a bit-sliced ripple add (xor3 0x96, majority 0xE8) gives
            // each expected bit plane, compared as it is made
            o << "    const " << T << " bw = " << sl.b << ";\n";
            for (int bit = 0; bit < lin_free; bit++)
                if (!hoisted[bit])
                    gather(bit_of("bw", bit), oa[bit]);
            for (int bit = lin_free; bit < el.width; bit++) {
                const std::string tp = "tp" + std::to_string(bit), rb = "r" + std::to_string(bit);
                o << "    const u32 " << rb << " = " << bit_of("bw", bit) << ";\n";
                if (bit == lin_free) {
                    o << "    u32 cy = " << tp << " & " << rb << ";\n";
                    gather("(" + tp + " ^ " + rb + ")", oa[bit]);
                } else {
                    o << "    const u32 s" << bit << " = __builtin_amdgcn_bitop3_b32(" << tp << ", "
                      << rb << ", cy, 0x96u);\n";
                    if (bit + 1 < el.width)
                        o << "    cy = __builtin_amdgcn_bitop3_b32(" << tp << ", " << rb
                          << ", cy, 0xE8u);\n";
                    gather("s" + std::to_string(bit), oa[bit]);
                }
            }
            o << "    hit = (miss0 | miss1 | miss2 | miss3"
              << (std::count(hoisted.begin(), hoisted.end(), 1) ? " | gmiss" : "") << ") & valid;\n";
        } else {
            for (int l = 0; l < 32; l++)
                o << "    const " << T << " e" << l << " = " << sc.lane_expr[l] << ";\n";
            // lane-free low bits: every lane expects bit `bit` of e0
            for (int bit = 0; bit < sc.lane_free; bit++)
                gather("(u32)__builtin_amdgcn_sbfe((int)(u32)(e0" + std::string(bit >= 32 ? " >> 32" : "") +
                           "), " + std::to_string(bit % 32) + "u, 1u)",
                       oa[bit]);
            // the rest from the 32 x 32 transpose of the expected values (the compiler drops the
            // transpose work of rows no compare reads)
            for (int half = 0; half < (el.width > 32 ? 2 : 1); half++) {
                if (std::max(sc.lane_free, 32 * half) >= std::min(el.width, 32 * half + 32))
                    continue;
                o << "    {\n";
                for (int l = 0; l < 32; l++)
                    o << "    u32 t" << l << " = (u32)(e" << l << (half ? " >> 32" : "") << ");\n";
                emit_transpose(o, "t");
                for (int bit = std::max(sc.lane_free, 32 * half); bit < el.width && bit < 32 * half + 32;
                     bit++)
                    gather("t" + std::to_string(bit - 32 * half), oa[bit]);
                o << "    }\n";
            }
            o << "    hit = (miss0 | miss1 | miss2 | miss3) & valid;\n";
        }
    }
    o << "    ts_report_hits(out, list, cap, (u64)gw, hit, &hits, &full);\n";
    if (grouped)
        o << "    }\n";
    o << "  }\n"
         "  ts_count_hits(out, hits);\n}\n";
    ks.names.push_back("ts_enum");
    ks.text = o.str();
    return true;
}

}  // namespace ts
