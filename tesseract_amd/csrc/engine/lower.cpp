// Program -> AIG -> segments and slices. See lower.hpp.
#include "lower.hpp"

#include <algorithm>
#include <unordered_map>
#include <cstdlib>
#include <cstdio>

namespace ts {

static inline uint64_t pair_key(uint32_t a, uint32_t b) { return (uint64_t)a << 32 | b; }

Aig::Aig()
{
    nodes.push_back(Node{N_CONST, 0, 0});
}

Lit Aig::And(Lit x, Lit y)
{
    if (x > y)
        std::swap(x, y);
    if (x == 0)
        return 0;          // false & y
    if (x == 1)
        return y;          // true & y
    if (x == y)
        return x;
    if (x == lit_not(y))
        return 0;
    uint64_t k = pair_key(x, y);
    auto it = and_table.find(k);
    if (it != and_table.end())
        return it->second << 1;
    uint32_t id = (uint32_t)nodes.size();
    nodes.push_back(Node{N_AND, x, y});
    and_table.emplace(k, id);
    return id << 1;
}

Lit Aig::Leaf(uint8_t kind, uint32_t a, uint32_t b)
{
    uint64_t k = (uint64_t)kind << 56 ^ (uint64_t)a << 24 ^ b;  // a < 2^32, b < 2^16 here
    auto it = leaf_table.find(k);
    if (it != leaf_table.end()) {
        const Node &n = nodes[it->second];
        if (n.kind == kind && n.a == a && n.b == b)
            return it->second << 1;
    }
    uint32_t id = (uint32_t)nodes.size();
    nodes.push_back(Node{kind, a, b});
    leaf_table[k] = id;
    return id << 1;
}

// The value of x in machine k ^ mask.
Lit Aig::Perm(Lit x, uint32_t mask)
{
    if (mask == 0)
        return x;
    const Node n = nodes[x >> 1];
    Lit c = x & 1u;
    switch (n.kind) {
    case N_CONST:
        return x;
    case N_LOAD:
        return Load(n.a, n.b ^ mask) ^ c;
    case N_IDBIT:
        return IdBit(n.a) ^ c ^ ((mask >> n.a) & 1u);
    case N_PERM:
        return Perm(n.a, n.b ^ mask) ^ c;
    default:
        return Leaf(N_PERM, x & ~1u, mask) ^ c;
    }
}

// ---- lowering ------------------------------------------------------------------------------

namespace {

struct PassBuilder {
    const ts_program *p;
    const LowerOptions &opt;
    Lowered &lw;
    uint32_t RR;
    std::vector<Lit> cur;
    std::vector<uint8_t> valid;          // written in the current segment
    std::vector<int> last_write;         // segment of the last write so far in this pass
    std::vector<std::vector<uint32_t>> version_ids;  // per cell: (seg + 1) -> version id
    std::vector<uint32_t> written_order; // cells written in the current segment, first-write order

    PassBuilder(const ts_program *p_, const LowerOptions &o, Lowered &l)
        : p(p_), opt(o), lw(l), RR(p_->ncells)
    {
        cur.assign(RR + 1, 0);
        valid.assign(RR + 1, 0);
        last_write.assign(RR + 1, -1);
        version_ids.resize(RR + 1);
        lw.segs.emplace_back();
    }

    Aig &g() { return lw.segs.back().g; }
    int seg() const { return (int)lw.segs.size() - 1; }

    uint32_t version(uint32_t cell, int s)
    {
        auto &v = version_ids[cell];
        size_t at = (size_t)(s + 1);
        if (v.size() <= at)
            v.resize(at + 1, UINT32_MAX);
        if (v[at] == UINT32_MAX) {
            v[at] = (uint32_t)lw.versions.size();
            lw.versions.push_back(Version{cell, s});
        }
        return v[at];
    }

    Lit value(uint32_t cell)
    {
        if (valid[cell])
            return cur[cell];
        return g().Load(version(cell, last_write[cell]), 0);
    }

    void write(uint32_t cell, Lit v)
    {
        if (!valid[cell]) {
            valid[cell] = 1;
            written_order.push_back(cell);
        }
        cur[cell] = v;
    }

    void close_segment()
    {
        Segment &s = lw.segs.back();
        for (uint32_t c : written_order) {
            s.outs.push_back(SegOut{c, cur[c], -1});
            last_write[c] = seg();
            valid[c] = 0;
        }
        written_order.clear();
    }

    void cut()
    {
        close_segment();
        lw.segs.emplace_back();
    }

    bool run(std::string &err)
    {
        for (size_t i = 0; i < p->len; i++) {
            const ts_instr in = p->code[i];
            if (in.addr >= p->ncells) {
                char b[96];
                snprintf(b, sizeof b, "instruction %zu reads cell %u of %u", i, in.addr, p->ncells);
                err = b;
                return false;
            }
            if (in.op == TS_OP_STO) {
                if (!ts_readonly(in.addr))
                    write(in.addr, value(RR));
                continue;
            }
            unsigned mask = 0;
            uint32_t c = resolve(in.addr, &mask);
            Lit v;
            if (c == TS_CELL_ZERO)
                v = 0;
            else if (c == TS_CELL_ONE)
                v = 1;
            else if (p->special && p->special[c].kind == TS_CK_ID)
                v = g().IdBit(p->special[c].dim) ^ ((mask >> p->special[c].dim) & 1u);
            else if (mask == 0)
                v = value(c);
            else {
                Lit base = value(c);
                if ((mask >> lw.wg_dims) && g().derived(base)) {
                    // a value computed in this segment, read from a machine in another workgroup:
                    // computed again here from that machine's start-of-segment values if its cone
                    // is small, else the segment ends (the value crosses as a slot)
                    if (!remat(base, mask, v)) {
                        cut();
                        v = g().Perm(value(c), mask);
                    }
                } else {
                    v = g().Perm(base, mask);
                }
            }
            Lit r = value(RR);
            write(RR, lit_not(g().And(r, v)));
        }
        close_segment();
        return true;
    }

    // Cones of at most this many computed nodes are recomputed rather than cut for.
    static constexpr size_t kRematNodes = 128;

    // x as machine k ^ mask computes it, rebuilt from loads and ids of that machine: for a cone
    // of at most kRematNodes AND/PERM nodes. Reads across dims the wave sees stay PERM nodes.
    std::vector<uint32_t> seen_at;  // node -> the remat call that last visited it
    uint32_t remat_calls = 0;

    bool remat(Lit x, uint32_t mask, Lit &out)
    {
        if (seen_at.size() < g().nodes.size())
            seen_at.resize(g().nodes.size() * 2, 0);
        const uint32_t stamp = ++remat_calls;
        std::vector<uint32_t> stack{x >> 1};
        size_t count = 0;
        while (!stack.empty()) {
            uint32_t n = stack.back();
            stack.pop_back();
            if (seen_at[n] == stamp)
                continue;
            seen_at[n] = stamp;
            const Node nd = g().nodes[n];
            if (nd.kind == N_AND) {
                stack.push_back(nd.a >> 1);
                stack.push_back(nd.b >> 1);
            } else if (nd.kind == N_PERM) {
                stack.push_back(nd.a >> 1);
            } else {
                continue;
            }
            if (++count > kRematNodes)
                return false;
        }
        std::unordered_map<uint64_t, Lit> memo;
        out = moved(x, mask, memo);
        return true;
    }

    Lit moved(Lit x, uint32_t mask, std::unordered_map<uint64_t, Lit> &memo)
    {
        if (!(mask >> lw.wg_dims))
            return g().Perm(x, mask);  // inside the workgroup: a move of the value
        const uint32_t n = x >> 1;
        const Lit c = x & 1u;
        const uint64_t key = (uint64_t)n << 32 | mask;
        auto it = memo.find(key);
        if (it != memo.end())
            return it->second ^ c;
        const Node nd = g().nodes[n];
        Lit r;
        if (nd.kind == N_AND)
            r = g().And(moved(nd.a, mask, memo), moved(nd.b, mask, memo));
        else if (nd.kind == N_PERM)
            r = moved(nd.a, nd.b ^ mask, memo);
        else
            r = g().Perm(n << 1, mask);  // loads, ids, constants move directly
        memo.emplace(key, r);
        return r ^ c;
    }

    bool ts_readonly(uint32_t c) const
    {
        return c < TS_CELL_FIRST_FREE || (p->special && p->special[c].kind != TS_CK_PLAIN);
    }

    uint32_t resolve(uint32_t c, unsigned *mask) const
    {
        unsigned m = 0;
        while (p->special && p->special[c].kind == TS_CK_LINK) {
            m ^= 1u << p->special[c].dim;
            c = p->special[c].src;
        }
        *mask = m;
        return c;
    }
};

// Nodes reachable from the given literals, in index order.
static void cone(const Aig &g, const std::vector<Lit> &roots, std::vector<uint8_t> &mark)
{
    mark.assign(g.nodes.size(), 0);
    std::vector<uint32_t> stack;
    for (Lit r : roots)
        stack.push_back(r >> 1);
    while (!stack.empty()) {
        uint32_t n = stack.back();
        stack.pop_back();
        if (mark[n])
            continue;
        mark[n] = 1;
        const Node &nd = g.nodes[n];
        if (nd.kind == N_AND) {
            stack.push_back(nd.a >> 1);
            stack.push_back(nd.b >> 1);
        } else if (nd.kind == N_PERM) {
            stack.push_back(nd.a >> 1);
        }
    }
}

// Splits a segment's stored outputs into slices of about `target` AIG nodes each, in output order
// (program order keeps neighbouring cells, which share inputs, together). Returns the nodes the
// slices hold together (a node in the cones of several slices counts in each).
static uint64_t slice_outputs(Segment &s, std::vector<uint32_t> stored, unsigned target);

// Slices of the outputs every pass stores, then slices of the final_only outputs; the nodes they
// hold together.
static uint64_t make_slices(Segment &s, unsigned target)
{
    std::vector<uint32_t> always, last;
    for (uint32_t i = 0; i < s.outs.size(); i++)
        if (s.outs[i].slot >= 0)
            (s.outs[i].final_only ? last : always).push_back(i);
    s.slices.clear();
    uint64_t held = slice_outputs(s, always, target);
    s.always = s.slices.size();
    return held + slice_outputs(s, last, target);
}

static uint64_t slice_outputs(Segment &s, std::vector<uint32_t> stored, unsigned target)
{
    if (stored.empty())
        return 0;
    if (!target) {
        s.slices.push_back(stored);
        return 0;
    }
    // visit order: by the output's node index, so cells computed together stay together
    std::stable_sort(stored.begin(), stored.end(), [&](uint32_t x, uint32_t y) {
        return (s.outs[x].v >> 1) < (s.outs[y].v >> 1);
    });
    std::vector<uint8_t> mark(s.g.nodes.size(), 0);
    std::vector<uint32_t> touched, stack;
    std::vector<uint32_t> slice;
    size_t size = 0;
    uint64_t held = 0;
    for (uint32_t oi : stored) {
        stack.push_back(s.outs[oi].v >> 1);
        size_t added = 0;
        while (!stack.empty()) {
            uint32_t n = stack.back();
            stack.pop_back();
            if (mark[n])
                continue;
            mark[n] = 1;
            touched.push_back(n);
            const Node &nd = s.g.nodes[n];
            if (nd.kind == N_AND) {
                added++;
                stack.push_back(nd.a >> 1);
                stack.push_back(nd.b >> 1);
            } else if (nd.kind == N_PERM) {
                added++;
                stack.push_back(nd.a >> 1);
            }
        }
        slice.push_back(oi);
        size += added;
        held += added;
        if (size >= target) {
            s.slices.push_back(slice);
            slice.clear();
            size = 0;
            for (uint32_t n : touched)
                mark[n] = 0;
            touched.clear();
        }
    }
    if (!slice.empty())
        s.slices.push_back(slice);
    return held;
}

// Slice size when the caller leaves it open. A big cube (2^14 machines and more) is usually one
// state of a few cubes, so its parallelism must come from slices: aim at kSliceThreads threads per
// segment launch (8 waves per SIMD of a 256-CU device), at least 50 nodes per slice so the
// re-computation of shared logic stays small. Smaller cubes run many cubes per state and keep
// large slices. (Life on 2^16 machines: 1,500-node slices 27.8 us/generation, 200-node 12.7 us;
// with last-pass-only outputs, a 200-node floor 8.3 us and a 50-node floor 8.0 us.)
static constexpr unsigned kSliceLarge = 1500, kSliceSmall = 50;

static uint64_t slices_wanted(unsigned dims)
{
    const uint64_t kSliceThreads = 256ull * 4 * 8 * 64;
    const uint64_t words = 1ull << (dims - 5);
    return std::max<uint64_t>(1, kSliceThreads / words);
}

static unsigned auto_slice_nodes(const Segment &s, unsigned dims)
{
    if (dims < 14)
        return kSliceLarge;
    uint64_t nodes = 0;
    for (const Node &n : s.g.nodes)
        nodes += n.kind == N_AND || n.kind == N_PERM;
    return (unsigned)std::min<uint64_t>(kSliceLarge,
                                        std::max<uint64_t>(kSliceSmall, nodes / slices_wanted(dims)));
}

// Slices share cones, so cutting a segment by its node count makes more slices than wanted;
// once, the size is scaled so that the count comes near the wanted one. Life 8192^2 (256
// wanted): 516 slices ran 6.9 us per generation, 272 6.6 us (the time is not smooth in the
// slice size: a caller that knows its program passes slice_nodes).
// Slices recompute the logic their cones share; where outputs' cones overlap so much that the
// slices would hold more than kMostHeld times the segment's nodes (a program whose every output
// reads most of the network, like the gamma world's tick: 127 slices, 37 x the nodes at 1,500),
// the slices grow until they hold at most that.
static constexpr uint64_t kMostHeld = 3;

static void auto_slices(Segment &s, unsigned dims)
{
    unsigned target = auto_slice_nodes(s, dims);
    uint64_t held = make_slices(s, target);
    if (dims >= 14) {
        const uint64_t want = slices_wanted(dims);
        if (s.slices.size() > want + want / 2) {
            target = (unsigned)std::min<uint64_t>(kSliceLarge, target * s.slices.size() / want);
            held = make_slices(s, target);
        }
    }
    uint64_t nodes = 0;
    for (const Node &n : s.g.nodes)
        nodes += n.kind == N_AND || n.kind == N_PERM;
    while (held > kMostHeld * nodes && target < nodes) {
        target = (unsigned)std::min<uint64_t>(nodes, 2ull * target);
        held = make_slices(s, target);
    }
}

}  // namespace

bool lower_program(const ts_program *p, const LowerOptions &opt, Lowered &lw, std::string &err)
{
    lw = Lowered();
    lw.ncells = p->ncells;
    unsigned dims = TS_CUBE_DIMS;
    for (uint32_t c = 0; p->special && c < p->ncells; c++)
        if (p->special[c].kind != TS_CK_PLAIN && p->special[c].dim + 1u > dims)
            dims = p->special[c].dim + 1u;
    lw.dims = dims;
    unsigned wl = opt.cube_words_log2 < 6 ? opt.cube_words_log2 : 6;
    lw.wave_dims = 5 + wl;
    lw.wg_dims = opt.cube_words_log2 > 6 && opt.cube_words_log2 <= opt.wg_words_log2
                     ? 5 + opt.cube_words_log2
                     : lw.wave_dims;

    PassBuilder pb(p, opt, lw);
    if (!pb.run(err))
        return false;

    const uint32_t total = p->ncells + 1;
    lw.last_seg = pb.last_write;

    // where every version lives
    lw.ver_where.resize(lw.versions.size());
    std::vector<std::vector<int32_t>> slot_of(total);  // per cell: (seg) -> slot
    auto need_slot = [&](uint32_t cell, int s) {
        auto &v = slot_of[cell];
        if (v.size() <= (size_t)s)
            v.resize((size_t)s + 1, -1);
        if (v[(size_t)s] < 0)
            v[(size_t)s] = (int32_t)lw.nslots++;
        return v[(size_t)s];
    };
    lw.final_slot.assign(total, -1);
    lw.init_slot.assign(total, -1);
    for (uint32_t c = 0; c < total; c++) {
        bool plain = c == p->ncells ||
                     (c >= TS_CELL_FIRST_FREE && !(p->special && p->special[c].kind != TS_CK_PLAIN));
        if (!plain)
            continue;
        if (lw.last_seg[c] >= 0)
            lw.final_slot[c] = need_slot(c, lw.last_seg[c]);
        else
            lw.init_slot[c] = (int32_t)lw.ninit++;
    }
    lw.ver_slot.resize(lw.versions.size());
    for (size_t v = 0; v < lw.versions.size(); v++) {
        const Version &ver = lw.versions[v];
        if (ver.seg >= 0) {
            lw.ver_where[v] = WHERE_CUR;
            lw.ver_slot[v] = need_slot(ver.cell, ver.seg);
        } else if (lw.last_seg[ver.cell] >= 0) {
            lw.ver_where[v] = WHERE_PREV;
            lw.ver_slot[v] = lw.final_slot[ver.cell];
        } else {
            lw.ver_where[v] = WHERE_INIT;
            lw.ver_slot[v] = lw.init_slot[ver.cell];
            if (lw.ver_slot[v] < 0) {
                err = "internal: a read of a cell without storage";
                return false;
            }
        }
    }
    // a cell is upward exposed when a pass reads its start-of-pass value; a version is read later
    // in the pass when some segment loads it
    std::vector<uint8_t> upward(total, 0);
    std::vector<std::vector<uint8_t>> read_later(total);
    for (const Version &ver : lw.versions) {
        if (ver.seg < 0) {
            upward[ver.cell] = 1;
        } else {
            auto &r = read_later[ver.cell];
            if (r.size() <= (size_t)ver.seg)
                r.resize((size_t)ver.seg + 1, 0);
            r[(size_t)ver.seg] = 1;
        }
    }
    for (size_t s = 0; s < lw.segs.size(); s++)
        for (SegOut &o : lw.segs[s].outs) {
            auto &v = slot_of[o.cell];
            o.slot = s < v.size() ? v[s] : -1;
            const auto &r = read_later[o.cell];
            const bool later = s < r.size() && r[s];
            o.final_only = o.slot >= 0 && (int)s == lw.last_seg[o.cell] && !upward[o.cell] && !later;
        }

    // resident: one segment and every read of a start-of-pass value inside the workgroup
    lw.resident = lw.segs.size() == 1;
    for (const Node &n : lw.segs[0].g.nodes)
        if (n.kind == N_LOAD && (n.b >> lw.wg_dims))
            lw.resident = false;
    for (Segment &s : lw.segs) {
        if (lw.resident || opt.slice_nodes)
            make_slices(s, lw.resident ? 0 : opt.slice_nodes);
        else
            auto_slices(s, lw.dims);
    }
    return true;
}

}  // namespace ts
