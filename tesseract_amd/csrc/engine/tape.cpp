// Tape building: an equiv/search enumeration pass as data for the precompiled interpreter kernel
// (tape_kernel.hip). No hipRTC compile is needed, so a request whose input space is small starts within
// milliseconds instead of the seconds LLVM spends on a large straight-line kernel.
#include <algorithm>
#include <queue>

#include "lower.hpp"

namespace ts {

bool spec_bits(Aig &g, const ts_spec *sp, int width, std::vector<Lit> &bits, std::string &err)
{
    typedef std::vector<Lit> Word;
    const int W = width;
    auto add = [&](const Word &x, const Word &y, Lit carry) {
        Word r(W);
        for (int i = 0; i < W; i++) {
            const Lit t = g.Xor(x[i], y[i]);
            r[i] = g.Xor(t, carry);
            if (i + 1 < W)
                carry = g.Or(g.And(x[i], y[i]), g.And(carry, t));
        }
        return r;
    };
    auto inv = [&](Word x) {
        for (Lit &l : x)
            l = lit_not(l);
        return x;
    };
    const int nr = ts_spec_nrpn(sp);
    const ts_spec_item *rpn = ts_spec_rpn(sp);
    std::vector<Word> st;
    for (int i = 0; i < nr; i++) {
        const ts_spec_item &it = rpn[i];
        if (it.op == TS_T_NUM) {
            Word w(W);
            for (int b = 0; b < W; b++)
                w[b] = (Lit)((it.v >> b) & 1);
            st.push_back(w);
            continue;
        }
        if (it.op == TS_T_VAR) {
            const int k = (int)it.v;
            Word w(W, 0);
            for (int b = 0; b < W && b < ts_spec_var_bits(sp, k); b++)
                w[b] = g.InBit((uint32_t)ts_spec_var_pos(sp, k, b));
            st.push_back(w);
            continue;
        }
        if (st.empty()) {
            err = "internal: bad spec program";
            return false;
        }
        if (it.op == TS_T_NOT) {
            st.back() = inv(st.back());
            continue;
        }
        if (it.op == TS_T_NEG) {  // -x = ~x + 1
            st.back() = add(inv(st.back()), Word(W, 0), 1);
            continue;
        }
        if (st.size() < 2) {
            err = "internal: bad spec program";
            return false;
        }
        const Word y = st.back();
        st.pop_back();
        Word &x = st.back();
        switch (it.op) {
        case TS_T_ADD:
            x = add(x, y, 0);
            break;
        case TS_T_SUB:  // x + ~y + 1
            x = add(x, inv(y), 1);
            break;
        case TS_T_MUL: {  // the rows y_j & (x << j), added
            Word acc(W, 0);
            for (int j = 0; j < W; j++) {
                if (y[j] == 0)
                    continue;
                Word row(W, 0);
                for (int b = j; b < W; b++)
                    row[b] = g.And(x[b - j], y[j]);
                acc = add(acc, row, 0);
            }
            x = acc;
            break;
        }
        case TS_T_AND:
        case TS_T_OR:
        case TS_T_XOR:
            for (int b = 0; b < W; b++)
                x[b] = it.op == TS_T_AND ? g.And(x[b], y[b])
                       : it.op == TS_T_OR ? g.Or(x[b], y[b])
                                          : g.Xor(x[b], y[b]);
            break;
        case TS_T_SHL: {
            if (i == 0 || rpn[i - 1].op != TS_T_NUM) {
                err = "a shift by a non-constant has no tape form";
                return false;
            }
            const uint64_t s = rpn[i - 1].v;
            Word r(W, 0);
            for (int b = 0; b < W; b++)
                if ((uint64_t)b >= s)
                    r[b] = x[b - s];
            x = r;
            break;
        }
        default:
            err = "a right shift has no tape form";
            return false;
        }
    }
    if (st.size() != 1) {
        err = "internal: bad spec program";
        return false;
    }
    bits = st[0];
    return true;
}

bool build_tape(const ts_program *a, const ts_program *b, const int *in_map, const int *out_map,
                const EnumLowering &el, unsigned max_slots, Tape &t, std::string &err,
                const ts_spec *spec)
{
    if (el.ticks != 1 || (el.mode == TS_EN_VERIFY && !spec)) {
        err = "only enumerations with one tick have a tape form";
        return false;
    }
    Aig g;
    std::vector<Lit> oa, ob;
    uint32_t a_end = 0, b_end = 0;
    enum_outputs(a, b, in_map, out_map, el, g, oa, ob, &a_end, &b_end);
    std::vector<Lit> want;  // verify: the spec's bits, after the program's nodes
    if (el.mode == TS_EN_VERIFY && !spec_bits(g, spec, el.width, want, err))
        return false;
    if (el.mode == TS_EN_VERIFY)
        b_end = (uint32_t)g.nodes.size();

    // the hit word: equiv = OR_k (a_k ^ b_k), verify = OR_bit (out ^ spec bit), search =
    // AND_bit (bit == target bit) (src/enum.c:150-213)
    Lit hit;
    if (el.mode == TS_EN_EQUIV) {
        hit = 0;
        for (size_t k = 0; k < oa.size(); k++)
            hit = g.Or(hit, g.Xor(oa[k], ob[k]));
    } else if (el.mode == TS_EN_VERIFY) {
        hit = 0;
        for (int bit = 0; bit < el.width; bit++)
            hit = g.Or(hit, g.Xor(oa[bit], want[bit]));
    } else {
        hit = 1;
        for (int bit = 0; bit < el.width; bit++)
            hit = g.And(hit, (el.target >> bit & 1) ? oa[bit] : lit_not(oa[bit]));
    }
    const uint32_t nn = (uint32_t)g.nodes.size();

    // the cone of the hit word, with each node's readers inside it
    std::vector<uint8_t> in_cone(nn, 0);
    std::vector<uint32_t> stack{hit >> 1}, uses(nn, 0);
    while (!stack.empty()) {
        uint32_t v = stack.back();
        stack.pop_back();
        if (in_cone[v])
            continue;
        in_cone[v] = 1;
        const Node &nd = g.nodes[v];
        if (nd.kind == N_AND) {
            stack.push_back(nd.a >> 1);
            stack.push_back(nd.b >> 1);
        }
    }
    std::vector<std::vector<uint32_t>> readers(nn);
    size_t and_nodes = 0;
    for (uint32_t v = 0; v < nn; v++) {
        if (!in_cone[v] || g.nodes[v].kind != N_AND)
            continue;
        and_nodes++;
        const Node &nd = g.nodes[v];
        uses[nd.a >> 1]++;
        uses[nd.b >> 1]++;
        readers[nd.a >> 1].push_back(v);
        if ((nd.b >> 1) != (nd.a >> 1))
            readers[nd.b >> 1].push_back(v);
    }
    uses[hit >> 1]++;  // the final read

    // relative position of a node in the execution that created it: interleaves a and b
    auto when = [&](uint32_t v) -> double {
        if (v < a_end)
            return (double)v / std::max<uint32_t>(1, a_end);
        if (v < b_end)
            return (double)(v - a_end) / std::max<uint32_t>(1, b_end - a_end);
        return 1.0 + (double)(v - b_end) / std::max<uint32_t>(1, nn - b_end);
    };

    t = Tape();
    t.and_nodes = and_nodes;
    std::vector<int32_t> slot(nn, -1);
    std::vector<uint32_t> free_slots;
    uint32_t next_slot = 1;  // slot 0 = constant 0
    auto alloc = [&]() -> uint32_t {
        if (!free_slots.empty()) {
            uint32_t s = free_slots.back();
            free_slots.pop_back();
            return s;
        }
        return next_slot++;
    };
    slot[0] = 0;
    std::vector<uint8_t> done(nn, 0);
    done[0] = 1;
    for (uint32_t v = 1; v < nn; v++)
        if (in_cone[v] && g.nodes[v].kind == N_INBIT) {
            slot[v] = (int32_t)alloc();
            done[v] = 1;
            t.leaves.push_back((uint32_t)slot[v] * Tape::kSlotBytes | g.nodes[v].a << 24);
        }

    typedef std::pair<double, uint32_t> Key;
    std::priority_queue<Key, std::vector<Key>, std::greater<Key>> ready;
    std::vector<uint8_t> waiting(nn, 0);  // operands not yet computed
    for (uint32_t v = 0; v < nn; v++) {
        if (!in_cone[v] || g.nodes[v].kind != N_AND)
            continue;
        const Node &nd = g.nodes[v];
        waiting[v] = (uint8_t)(!done[nd.a >> 1] + ((nd.b >> 1) != (nd.a >> 1) && !done[nd.b >> 1]));
        if (!waiting[v])
            ready.push(Key(when(v), v));
    }
    std::vector<uint32_t> group;
    std::vector<size_t> padding;  // ops words that name the sink slot
    size_t scheduled = 0;
    while (!ready.empty()) {
        group.clear();
        while (!ready.empty() && group.size() < Tape::kGroup) {
            group.push_back(ready.top().second);
            ready.pop();
        }
        // operands read for the last time free their slots for the next group's writes: the
        // kernel's waves split a group's entries, so a write may land while another wave still
        // reads this group's operands
        std::vector<uint32_t> freed;
        for (uint32_t v : group) {
            const Node &nd = g.nodes[v];
            for (uint32_t u : {nd.a >> 1, nd.b >> 1})
                if (--uses[u] == 0 && u != 0)
                    freed.push_back((uint32_t)slot[u]);
        }
        const size_t base = t.ops.size();
        t.ops.resize(base + Tape::kWords, 0);
        for (size_t j = 0; j < Tape::kGroup; j++) {
            uint32_t *e = &t.ops[base + j];
            if (j < group.size()) {
                const uint32_t v = group[j];
                const Node &nd = g.nodes[v];
                slot[v] = (int32_t)alloc();
                e[0] = (uint32_t)slot[nd.a >> 1] * Tape::kSlotBytes;
                e[Tape::kGroup] = (uint32_t)slot[nd.b >> 1] * Tape::kSlotBytes;
                e[2 * Tape::kGroup] = (uint32_t)slot[v] * Tape::kSlotBytes;
                e[3 * Tape::kGroup] = 0u - (nd.a & 1u);
                e[4 * Tape::kGroup] = 0u - (nd.b & 1u);
            } else {
                // padding: reads slot 0, writes the sink slot (numbered once the file size is known)
                padding.push_back(base + 2 * Tape::kGroup + j);
            }
        }
        free_slots.insert(free_slots.end(), freed.begin(), freed.end());
        t.count.push_back((uint32_t)group.size());
        scheduled += group.size();
        for (uint32_t v : group) {
            done[v] = 1;
            for (uint32_t r : readers[v])
                if (--waiting[r] == 0)
                    ready.push(Key(when(r), r));
        }
        if (next_slot + 1 > max_slots) {
            err = "the tape needs more than " + std::to_string(max_slots) + " slots";
            return false;
        }
    }
    if (scheduled != and_nodes) {
        err = "tape scheduling left nodes behind";
        return false;
    }
    t.root = (uint32_t)slot[hit >> 1] * Tape::kSlotBytes | (hit & 1u) << 31;
    const uint32_t sink = next_slot;
    for (size_t w : padding)
        t.ops[w] = sink * Tape::kSlotBytes;
    t.slots = sink + 1;
    return true;
}

}  // namespace ts
