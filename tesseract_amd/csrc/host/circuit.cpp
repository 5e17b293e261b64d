// Circuits into NAND/STO programs: the program writer behind the ts_b_* calls (named cells,
// gates, adders, flip-flops, cube links) and the ISCAS .bench loader (ts_net_parse).
//
// Interfaces kept from the reference: the gate names (src/build.h), the shortest instruction
// sequence of each gate (src/build.c:177-307, the reference's superoptimizer results, so programs
// have the reference's instruction counts), registers that all take their new value at the end of a
// pass, and the .bench grammar, messages and cell-name rule (src/net.c). Everything else is this
// file's own: the sequences are data run by one small interpreter, registers are ordered so that only
// registers on a cycle go through staging cells, and the netlist is ordered by logic level.
#include <algorithm>
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <string_view>
#include <unordered_map>
#include <utility>
#include <vector>

#include "tesseract_host.h"

namespace {

// ---- gate sequences ----------------------------------------------------------------------------

// Operands of a sequence step: the gate's inputs x, y, c, its results s, k, the writer's
// temporaries and the constant cells.
enum Arg : uint8_t { X, Y, CIN, S, K, T0, T1, T2, T3, ZERO, ONE, kArgs };
enum : uint8_t { N = TS_OP_NAND, W = TS_OP_STO };

struct Step {
    uint8_t op;
    Arg     arg;
};

// RR = x ^ y in 7 instructions
constexpr Step kXor2[] = {{N, ZERO}, {N, X}, {N, Y}, {W, T0}, {N, Y}, {N, X}, {N, T0}};
// s = x ^ y, k = x & y in 11 (s may be x or y)
constexpr Step kHalfAdd[] = {{N, ZERO}, {N, X}, {N, Y}, {W, T0}, {N, Y}, {N, ONE},
                             {W, K},    {N, ONE}, {N, X}, {N, T0}, {W, S}};
// s = x ^ y ^ c, k = carry in 19 (s may be x or y, k may be c)
constexpr Step kFullAdd[] = {{N, ZERO}, {N, X}, {W, T0}, {N, ONE}, {N, Y},  {W, T1}, {N, Y},
                             {N, T0},   {W, T2}, {N, T1}, {N, CIN}, {W, T3}, {N, CIN}, {N, T1},
                             {W, K},    {N, ONE}, {N, T2}, {N, T3}, {W, S}};
// RR = v: "load" and "load complemented"
constexpr Step kLoad[] = {{N, ZERO}, {N, X}, {N, ONE}};
constexpr Step kLoadNot[] = {{N, ZERO}, {N, X}};

constexpr const char *kGateNames[TS_G_COUNT] = {"BUF", "NOT", "AND", "NAND",
                                                "OR",  "NOR", "XOR", "XNOR"};

std::string fold(std::string_view s)
{
    std::string k(s);
    for (char &c : k)
        c = (char)tolower((unsigned char)c);
    return k;
}

bool same_fold(std::string_view a, std::string_view b) { return fold(a) == fold(b); }

}  // namespace

// ---- the program writer ------------------------------------------------------------------------

struct ts_builder {
    std::vector<ts_instr>    code;
    std::vector<std::string> names;
    std::vector<ts_cellinfo> info;
    std::unordered_map<std::string, uint32_t> by_name;  // folded name -> cell
    std::vector<uint32_t>    ins, outs;
    std::vector<std::pair<uint32_t, uint32_t>> regs;   // (q, d): q <- d at the end of the pass
    uint32_t                 temp[5] = {0, 0, 0, 0, 0};  // created on first use
    bool                     special = false;
    std::string              error;                     // the first misuse, reported by finish

    ts_builder()
    {
        cell("ZERO");
        cell("ONE");
    }

    void misuse(const char *fmt, ...)
    {
        if (!error.empty())
            return;
        char b[256];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof b, fmt, ap);
        va_end(ap);
        error = b;
    }

    uint32_t ncells() const { return (uint32_t)names.size(); }

    uint32_t cell(std::string_view name)
    {
        std::string k = fold(name);
        auto it = by_name.find(k);
        if (it != by_name.end())
            return it->second;
        const uint32_t c = ncells();
        names.emplace_back(name);
        ts_cellinfo ci{};
        const int d = ts_program_id_name(names.back().c_str());
        if (d >= 0) {  // ID0 .. ID15 read a bit of the machine number
            ci.kind = TS_CK_ID;
            ci.dim = (uint8_t)d;
            special = true;
        }
        info.push_back(ci);
        by_name.emplace(std::move(k), c);
        return c;
    }

    bool has(std::string_view name) const { return by_name.count(fold(name)) != 0; }

    bool writable(uint32_t c) const
    {
        return c >= TS_CELL_FIRST_FREE && c < ncells() && info[c].kind == TS_CK_PLAIN;
    }

    // temporaries: names a sanitised netlist signal never takes (they would start with "n")
    uint32_t tmp(int k)
    {
        if (!temp[k]) {
            char nm[8];
            snprintf(nm, sizeof nm, "__w%d", k);
            temp[k] = cell(nm);
        }
        return temp[k];
    }

    void emit(uint8_t op, uint32_t c)
    {
        if (c >= ncells()) {
            misuse("instruction on cell %u, but only %u cells exist", c, ncells());
            return;
        }
        code.push_back(ts_instr{op, c});
    }

    // runs a sequence with its operands bound
    template <size_t n>
    void run(const Step (&seq)[n], const uint32_t (&bind)[kArgs])
    {
        for (const Step &st : seq)
            emit(st.op, bind[st.arg]);
    }

    void bindings(uint32_t (&b)[kArgs], uint32_t x, uint32_t y = 0, uint32_t c = 0, uint32_t s = 0,
                  uint32_t k = 0, int temps = 0)
    {
        b[X] = x;
        b[Y] = y;
        b[CIN] = c;
        b[S] = s;
        b[K] = k;
        for (int i = 0; i < 4; i++)
            b[T0 + i] = i < temps ? tmp(i) : 0;
        b[ZERO] = TS_CELL_ZERO;
        b[ONE] = TS_CELL_ONE;
    }

    void load(uint32_t v, bool complemented)
    {
        uint32_t b[kArgs];
        bindings(b, v);
        if (complemented)
            run(kLoadNot, b);
        else
            run(kLoad, b);
    }

    void xor2(uint32_t x, uint32_t y)
    {
        uint32_t b[kArgs];
        bindings(b, x, y, 0, 0, 0, 1);
        run(kXor2, b);
    }

    void gate(int type, uint32_t out, const uint32_t *in, int n)
    {
        if (type < 0 || type >= TS_G_COUNT) {
            misuse("unknown gate type %d", type);
            return;
        }
        if (n < 1 || ((type == TS_G_BUF || type == TS_G_NOT) && n != 1)) {
            misuse("%s with %d inputs", kGateNames[type], n);
            return;
        }
        if (!writable(out)) {
            misuse("%s writes cell %u, which is read-only or missing", kGateNames[type], out);
            return;
        }
        for (int i = 0; i < n; i++)
            if (in[i] >= ncells()) {
                misuse("%s reads cell %u, but only %u cells exist", kGateNames[type], in[i],
                       ncells());
                return;
            }
        const bool inverted = type == TS_G_NAND || type == TS_G_NOR || type == TS_G_XNOR;
        switch (type) {
        case TS_G_BUF:
        case TS_G_NOT:
            load(in[0], type == TS_G_NOT);
            break;
        case TS_G_AND:
        case TS_G_NAND:
            // RR = in0, then RR = !(acc & in i), re-inverted between inputs
            load(in[0], false);
            for (int i = 1; i < n; i++) {
                emit(N, in[i]);
                if (i + 1 < n)
                    emit(N, TS_CELL_ONE);
            }
            if ((n == 1) == inverted)
                emit(N, TS_CELL_ONE);
            break;
        case TS_G_OR:
        case TS_G_NOR: {
            // RR = !in0; each further input: t = !acc, RR = !(!(acc | !x) & t) = acc | x
            const uint32_t t = tmp(0);
            load(in[0], true);
            for (int i = 1; i < n; i++) {
                if (i > 1)
                    emit(N, TS_CELL_ONE);
                emit(W, t);
                emit(N, in[i]);
                emit(N, t);
            }
            if ((n == 1) != inverted)
                emit(N, TS_CELL_ONE);
            break;
        }
        default: {  // XOR, XNOR: pairwise through an accumulator
            if (n == 1) {
                load(in[0], false);
            } else {
                xor2(in[0], in[1]);
                for (int i = 2; i < n; i++) {
                    const uint32_t acc = tmp(4);
                    emit(W, acc);
                    xor2(acc, in[i]);
                }
            }
            if (inverted)
                emit(N, TS_CELL_ONE);
            break;
        }
        }
        emit(W, out);
    }

    void half_adder(uint32_t s, uint32_t k, uint32_t x, uint32_t y)
    {
        if (!writable(s) || !writable(k) || s == k || x >= ncells() || y >= ncells()) {
            misuse("half adder on unusable cells");
            return;
        }
        uint32_t b[kArgs];
        bindings(b, x, y, 0, s, k, 1);
        run(kHalfAdd, b);
    }

    void full_adder(uint32_t s, uint32_t k, uint32_t x, uint32_t y, uint32_t c)
    {
        if (!writable(s) || !writable(k) || s == k || x >= ncells() || y >= ncells() ||
            c >= ncells()) {
            misuse("full adder on unusable cells");
            return;
        }
        uint32_t b[kArgs];
        bindings(b, x, y, c, s, k, 4);
        run(kFullAdd, b);
    }

    uint32_t link(std::string_view name, uint32_t src, int dim)
    {
        if (has(name) || src >= ncells() || dim < 0 || dim >= TS_CUBE_MAX_DIMS) {
            misuse("link %.*s: the name exists, or no cell %u, or no dimension %d",
                   (int)name.size(), name.data(), src, dim);
            return TS_CELL_ZERO;
        }
        const uint32_t c = cell(name);
        info[c].kind = TS_CK_LINK;
        info[c].dim = (uint8_t)dim;
        info[c].src = src;
        special = true;
        return c;
    }

    // Registers take their new value all at once: a register is written only after every register
    // reading its old value, so a chain (a shift register) needs no copies; registers left on a
    // cycle (a swap) are staged, all of them read before any is written.
    void registers()
    {
        const size_t n = regs.size();
        std::unordered_map<uint32_t, size_t> reg_of_q;
        for (size_t r = 0; r < n; r++)
            reg_of_q[regs[r].first] = r;
        std::vector<size_t> readers(n, 0);  // unwritten registers reading register r's q
        std::vector<long> feeds(n, -1);     // the register whose q register r reads
        for (size_t r = 0; r < n; r++) {
            auto it = reg_of_q.find(regs[r].second);
            if (it != reg_of_q.end() && it->second != r) {
                feeds[r] = (long)it->second;
                readers[it->second]++;
            }
        }
        std::vector<char> done(n, 0);
        std::vector<size_t> ready;
        for (size_t r = 0; r < n; r++)
            if (!readers[r])
                ready.push_back(r);
        for (size_t at = 0; at < ready.size(); at++) {
            const size_t r = ready[at];
            load(regs[r].second, false);
            emit(W, regs[r].first);
            done[r] = 1;
            if (feeds[r] >= 0 && --readers[(size_t)feeds[r]] == 0)
                ready.push_back((size_t)feeds[r]);
        }
        std::vector<std::pair<size_t, uint32_t>> staged;
        for (size_t r = 0; r < n; r++)
            if (!done[r]) {
                char nm[32];
                snprintf(nm, sizeof nm, "__q%zu", r);
                const uint32_t st = cell(nm);
                load(regs[r].second, false);
                emit(W, st);
                staged.emplace_back(r, st);
            }
        for (auto &rs : staged) {
            load(rs.second, false);
            emit(W, regs[rs.first].first);
        }
    }

    // the program, in memory ts_program_free releases
    bool finish(ts_program *p)
    {
        registers();
        memset(p, 0, sizeof *p);
        if (!error.empty())
            return false;
        const uint32_t nc = ncells();
        p->code = (ts_instr *)malloc(std::max<size_t>(1, code.size()) * sizeof(ts_instr));
        p->cell_names = (char **)calloc(std::max<uint32_t>(1, nc), sizeof(char *));
        p->inputs = (uint32_t *)malloc(std::max<size_t>(1, ins.size()) * sizeof(uint32_t));
        p->outputs = (uint32_t *)malloc(std::max<size_t>(1, outs.size()) * sizeof(uint32_t));
        if (special)
            p->special = (ts_cellinfo *)malloc(nc * sizeof(ts_cellinfo));
        bool ok = p->code && p->cell_names && p->inputs && p->outputs && (!special || p->special);
        for (uint32_t c = 0; ok && c < nc; c++)
            ok = (p->cell_names[c] = strdup(names[c].c_str())) != nullptr;
        p->ncells = nc;
        if (!ok) {
            ts_program_free(p);
            error = "out of memory";
            return false;
        }
        std::copy(code.begin(), code.end(), p->code);
        std::copy(ins.begin(), ins.end(), p->inputs);
        std::copy(outs.begin(), outs.end(), p->outputs);
        if (special)
            std::copy(info.begin(), info.end(), p->special);
        p->len = code.size();
        p->ninputs = ins.size();
        p->noutputs = outs.size();
        return true;
    }
};

// Every entry below catches allocation failures: the writer then records "out of memory" and
// ts_b_finish reports it.
#define GUARD(b, body)                                                                         \
    do {                                                                                       \
        try {                                                                                  \
            body;                                                                              \
        } catch (const std::bad_alloc &) {                                                     \
            (b)->misuse("out of memory");                                                      \
        }                                                                                      \
    } while (0)

extern "C" {

ts_builder *ts_b_new(void)
{
    try {
        return new ts_builder();
    } catch (const std::bad_alloc &) {
        return nullptr;
    }
}

uint32_t ts_b_cell(ts_builder *b, const char *name)
{
    uint32_t c = TS_CELL_ZERO;
    GUARD(b, c = b->cell(name));
    return c;
}

int ts_b_has(const ts_builder *b, const char *name) { return b->has(name); }

void ts_b_input(ts_builder *b, uint32_t cell)
{
    if (cell >= b->ncells())
        b->misuse("input cell %u does not exist", cell);
    else
        GUARD(b, b->ins.push_back(cell));
}

void ts_b_output(ts_builder *b, uint32_t cell)
{
    if (cell >= b->ncells())
        b->misuse("output cell %u does not exist", cell);
    else
        GUARD(b, b->outs.push_back(cell));
}

void ts_b_nand(ts_builder *b, uint32_t cell) { GUARD(b, b->emit(N, cell)); }
void ts_b_sto(ts_builder *b, uint32_t cell) { GUARD(b, b->emit(W, cell)); }

void ts_b_gate(ts_builder *b, int type, uint32_t out, const uint32_t *in, int n)
{
    GUARD(b, b->gate(type, out, in, n));
}

void ts_b_gate2(ts_builder *b, int type, uint32_t out, uint32_t x, uint32_t y)
{
    const uint32_t in[2] = {x, y};
    GUARD(b, b->gate(type, out, in, 2));
}

void ts_b_reg(ts_builder *b, uint32_t q, uint32_t d)
{
    if (!b->writable(q) || d >= b->ncells())
        b->misuse("register %u <- %u on unusable cells", q, d);
    else
        GUARD(b, b->regs.emplace_back(q, d));
}

void ts_b_half_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y)
{
    GUARD(b, b->half_adder(s, k, x, y));
}

void ts_b_full_adder(ts_builder *b, uint32_t s, uint32_t k, uint32_t x, uint32_t y, uint32_t c)
{
    GUARD(b, b->full_adder(s, k, x, y, c));
}

uint32_t ts_b_link(ts_builder *b, const char *name, uint32_t src, int dim)
{
    uint32_t c = TS_CELL_ZERO;
    GUARD(b, c = b->link(name, src, dim));
    return c;
}

size_t ts_b_len(const ts_builder *b) { return b->code.size(); }

int ts_b_finish(ts_builder *b, ts_program *out, char *err, size_t errlen)
{
    bool ok = false;
    try {
        ok = b->finish(out);
    } catch (const std::bad_alloc &) {
        memset(out, 0, sizeof *out);
        b->error = "out of memory";
    }
    if (!ok && err && errlen)
        snprintf(err, errlen, "%s", b->error.c_str());
    delete b;
    return ok;
}

const char *ts_gate_name(int type) { return type >= 0 && type < TS_G_COUNT ? kGateNames[type] : "?"; }

int ts_gate_type(const char *name)
{
    if (same_fold(name, "BUFF"))
        return TS_G_BUF;
    for (int t = 0; t < TS_G_COUNT; t++)
        if (same_fold(name, kGateNames[t]))
            return t;
    return -1;
}

}  // extern "C"

// ---- .bench netlists ---------------------------------------------------------------------------

namespace {

constexpr int kDff = -2;  // gate type of a flip-flop

struct BenchNet {
    struct Signal {
        std::string name;
        int         driver = -1;  // gate index; -1 = not driven, -2 = primary input
    };
    struct Gate {
        int              type;
        int              out;
        std::vector<int> in;
        int              line;
    };
    std::string                          file;
    std::vector<Signal>                  sig;
    std::unordered_map<std::string, int> sig_of;
    std::vector<Gate>                    gates;
    std::vector<int>                     inputs, outputs;
    std::string                          err;

    bool fail(int line, const char *fmt, ...)
    {
        char msg[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(msg, sizeof msg, fmt, ap);
        va_end(ap);
        char b[1024];
        snprintf(b, sizeof b, "%s:%d: %s", file.c_str(), line, msg);
        err = b;
        return false;
    }

    int signal(std::string_view name)
    {
        auto it = sig_of.find(std::string(name));
        if (it != sig_of.end())
            return it->second;
        sig.push_back({std::string(name), -1});
        sig_of.emplace(std::string(name), (int)sig.size() - 1);
        return (int)sig.size() - 1;
    }

    // ---- one line ----

    static void blank(std::string_view &s)
    {
        while (!s.empty() && (s[0] == ' ' || s[0] == '\t'))
            s.remove_prefix(1);
    }

    static bool name_char(char c) { return isalnum((unsigned char)c) || (c && strchr("_.[]$", c)); }

    // a signal name after blanks: letters, digits and _ . [ ] $; empty if there is none
    static std::string_view name(std::string_view &s)
    {
        blank(s);
        size_t n = 0;
        while (n < s.size() && name_char(s[n]))
            n++;
        std::string_view nm = s.substr(0, n);
        s.remove_prefix(n);
        return nm;
    }

    static bool punct(std::string_view &s, char c)
    {
        blank(s);
        if (s.empty() || s[0] != c)
            return false;
        s.remove_prefix(1);
        return true;
    }

    bool line(std::string_view s, int ln)
    {
        const size_t hash = s.find('#');
        if (hash != std::string_view::npos)
            s = s.substr(0, hash);
        const std::string word(name(s));
        if (word.empty()) {
            while (!s.empty() && (s[0] == ' ' || s[0] == '\t' || s[0] == '\r'))
                s.remove_prefix(1);
            return s.empty() || fail(ln, "cannot read this line");
        }
        const bool in = word == "INPUT" || word == "input", out = word == "OUTPUT" || word == "output";
        if (in || out) {
            std::string_view arg;
            if (!punct(s, '(') || (arg = name(s)).empty() || !punct(s, ')'))
                return fail(ln, "expected %s(name)", word.c_str());
            const int k = signal(arg);
            if (out) {
                outputs.push_back(k);
                return true;
            }
            if (sig[(size_t)k].driver != -1)
                return fail(ln, "%s is driven twice", std::string(arg).c_str());
            sig[(size_t)k].driver = -2;
            inputs.push_back(k);
            return true;
        }
        std::string_view type_sv;
        if (!punct(s, '=') || (type_sv = name(s)).empty() || !punct(s, '('))
            return fail(ln, "expected name = GATE(inputs)");
        const std::string type(type_sv);
        int t = ts_gate_type(type.c_str());
        if (t < 0 && (type == "DFF" || type == "dff"))
            t = kDff;
        if (t == -1)
            return fail(ln, "unknown gate %s", type.c_str());
        Gate g{t, signal(word), {}, ln};
        if (sig[(size_t)g.out].driver != -1)
            return fail(ln, "%s is driven twice", word.c_str());
        do {
            const std::string_view arg = name(s);
            if (arg.empty())
                return fail(ln, "expected an input name");
            g.in.push_back(signal(arg));
        } while (punct(s, ','));
        if (!punct(s, ')'))
            return fail(ln, "expected )");
        if ((t == TS_G_NOT || t == TS_G_BUF || t == kDff) && g.in.size() != 1)
            return fail(ln, "%s takes one input", type.c_str());
        sig[(size_t)g.out].driver = (int)gates.size();
        gates.push_back(std::move(g));
        return true;
    }

    // ---- the program ----

    // A gate's logic level: 1 + the highest level of the combinational gates it reads (inputs and
    // flip-flop outputs are level 0). -1: the gate is on, or behind, a loop without a flip-flop.
    std::vector<int> levels() const
    {
        enum : int { kUnseen = -3, kOpen = -2, kLoop = -1 };
        std::vector<int> level(gates.size(), kUnseen);
        std::vector<std::pair<int, size_t>> stack;  // (gate, next input to look at)
        for (size_t g0 = 0; g0 < gates.size(); g0++) {
            if (gates[g0].type == kDff || level[g0] != kUnseen)
                continue;
            stack.push_back({(int)g0, 0});
            level[g0] = kOpen;
            while (!stack.empty()) {
                auto &top = stack.back();
                const Gate &g = gates[(size_t)top.first];
                if (top.second < g.in.size()) {
                    const int d = sig[(size_t)g.in[top.second++]].driver;
                    if (d < 0 || gates[(size_t)d].type == kDff)
                        continue;
                    if (level[(size_t)d] == kUnseen) {
                        level[(size_t)d] = kOpen;
                        stack.push_back({d, 0});
                    } else if (level[(size_t)d] == kOpen) {
                        level[(size_t)d] = kLoop;  // a cycle: everything on the stack is behind it
                    }
                    continue;
                }
                int lv = 0;
                for (int s : g.in) {
                    const int d = sig[(size_t)s].driver;
                    if (d < 0 || gates[(size_t)d].type == kDff)
                        continue;
                    if (level[(size_t)d] < 0) {
                        lv = kLoop;
                        break;
                    }
                    lv = std::max(lv, level[(size_t)d] + 1);
                }
                // a gate found open again by its own cone stays on the loop
                level[(size_t)top.first] =
                    level[(size_t)top.first] == kLoop ? kLoop : (lv < 0 ? kLoop : std::max(lv, 1));
                stack.pop_back();
            }
        }
        return level;
    }

    // "x.y" -> "x_y"; names starting with a digit, "__" or named ZERO or ONE get an "n" in front
    static std::string cell_name(const std::string &s)
    {
        std::string out;
        const bool clash = (!s.empty() && isdigit((unsigned char)s[0])) || same_fold(s, "ZERO") ||
                           same_fold(s, "ONE") || s.rfind("__", 0) == 0;
        if (clash)
            out.push_back('n');
        for (char c : s)
            out.push_back(isalnum((unsigned char)c) || c == '_' ? c : '_');
        return out;
    }

    bool build(ts_program *out)
    {
        for (const Signal &s : sig)
            if (s.driver == -1)
                return fail(0, "signal %s is used but never driven", s.name.c_str());
        const std::vector<int> level = levels();
        std::vector<int> order;
        for (size_t g = 0; g < gates.size(); g++) {
            if (gates[g].type == kDff)
                continue;
            if (level[g] < 0)
                return fail(gates[g].line, "%s is part of a loop without a DFF",
                            sig[(size_t)gates[g].out].name.c_str());
            order.push_back((int)g);
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return level[(size_t)a] < level[(size_t)b]; });
        std::unique_ptr<ts_builder> b(new ts_builder());
        std::vector<uint32_t> cell(sig.size());
        for (size_t s = 0; s < sig.size(); s++) {
            const std::string nm = cell_name(sig[s].name);
            if (b->has(nm))
                return fail(0, "two signals both become cell %s", nm.c_str());
            cell[s] = b->cell(nm);
        }
        for (int s : inputs)
            b->ins.push_back(cell[(size_t)s]);
        std::vector<uint32_t> in;
        for (int g : order) {
            const Gate &gt = gates[(size_t)g];
            in.clear();
            for (int s : gt.in)
                in.push_back(cell[(size_t)s]);
            b->gate(gt.type, cell[(size_t)gt.out], in.data(), (int)in.size());
        }
        for (const Gate &gt : gates)
            if (gt.type == kDff)
                b->regs.emplace_back(cell[(size_t)gt.out], cell[(size_t)gt.in[0]]);
        for (int s : outputs)
            b->outs.push_back(cell[(size_t)s]);
        char msg[256];
        if (!ts_b_finish(b.release(), out, msg, sizeof msg))
            return fail(0, "%s", msg);
        return true;
    }
};

}  // namespace

// Reference net_parse (src/net.h:17): text in the ISCAS .bench format, name for messages.
extern "C" int ts_net_parse(const char *text, const char *name, ts_program *out, char *err,
                            size_t errlen)
{
    memset(out, 0, sizeof *out);
    try {
        BenchNet net;
        net.file = name ? name : "<netlist>";
        std::string_view all(text);
        int ln = 1;
        bool ok = true;
        while (ok) {
            const size_t nl = all.find('\n');
            ok = net.line(all.substr(0, nl), ln++);
            if (nl == std::string_view::npos)
                break;
            all.remove_prefix(nl + 1);
        }
        if (ok)
            ok = net.build(out);
        if (!ok)
            snprintf(err, errlen, "%s", net.err.c_str());
        return ok;
    } catch (const std::bad_alloc &) {
        snprintf(err, errlen, "out of memory");
        return 0;
    }
}
