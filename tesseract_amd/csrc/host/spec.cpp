// Specifications of verify and search: "c:s = a + b + c" against a program's outputs and inputs.
//
// The interface is the reference's (src/spec.h): the grammar, operator precedence, 64-bit
// arithmetic with shifts of 64 or more giving 0, the vector naming rule, the limits and every error
// message, so that a command line that the reference accepts or refuses is accepted or refused here
// with the same words. The implementation is a binding-power (Pratt) parser that builds an
// expression tree; the tree is evaluated by recursion and also flattened into postfix order for
// engines that lower it into their own code.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "tesseract_host.h"

namespace {

// the reference's limits (src/spec.c): postfix items, evaluation stack depth, input vectors,
// output vectors on the left side, characters of a name
constexpr int kMaxItems = 256;
constexpr int kMaxDepth = 64;
constexpr int kMaxVars = 32;
constexpr int kMaxOutVectors = 64;
constexpr size_t kNameChars = 63;

bool same_fold(const char *a, size_t na, const char *b, size_t nb)
{
    if (na != nb)
        return false;
    for (size_t i = 0; i < na; i++)
        if (tolower((unsigned char)a[i]) != tolower((unsigned char)b[i]))
            return false;
    return true;
}

}  // namespace

struct ts_spec {
    struct Node {
        int      op;         // TS_T_*
        uint64_t v;          // number, or variable index
        int      lhs, rhs;   // children (-1: none); a unary operator uses lhs
    };
    struct Var {
        std::string      name;
        std::vector<int> pos;  // pos[b]: the bit of the assignment x that is bit b of the vector
    };
    std::vector<uint32_t>     out;      // output cell of left-side bit i (0 = lowest)
    std::vector<Var>          vars;
    std::vector<Node>         nodes;
    int                       root = -1;
    std::vector<ts_spec_item> postfix;  // the tree in evaluation order
};

// vector `name` among cells list[0..n): the cell called exactly that (case-insensitive), else the
// cells name0, name1, ... with decimal suffixes without leading zeros (reference src/spec.c:49-82)
extern "C" int ts_vec_resolve(const ts_program *p, const uint32_t *list, size_t n, const char *name,
                              int *idx)
{
    const size_t len = strlen(name);
    for (size_t i = 0; i < n; i++) {
        const char *c = p->cell_names[list[i]];
        if (same_fold(c, strlen(c), name, len)) {
            idx[0] = (int)i;
            return 1;
        }
    }
    uint64_t have = 0;
    int width = 0;
    for (size_t i = 0; i < n; i++) {
        const char *c = p->cell_names[list[i]];
        const size_t cl = strlen(c);
        if (cl <= len || !same_fold(c, len, name, len))
            continue;
        const char *d = c + len;
        bool digits = true;
        for (const char *q = d; *q; q++)
            digits = digits && isdigit((unsigned char)*q);
        if (!digits || (d[0] == '0' && d[1]))
            continue;
        // the suffix as a number, saturating: anything past 63 is refused
        uint64_t bit = 0;
        for (const char *q = d; *q && bit < 64; q++)
            bit = bit * 10 + (uint64_t)(*q - '0');
        if (bit >= 64)
            return -1;
        idx[bit] = (int)i;
        have |= 1ull << bit;
        width = std::max(width, (int)bit + 1);
    }
    if (width && have != (width == 64 ? ~0ull : (1ull << width) - 1))
        return -1;
    return width;
}

namespace {

class SpecParser {
  public:
    SpecParser(const char *text, const ts_program *p, bool inputs_allowed, ts_spec &out)
        : s_(text), p_(p), inputs_allowed_(inputs_allowed), sp_(out)
    {
    }

    bool parse(std::string &err)
    {
        outputs();
        if (ok()) {
            blank();
            if (*s_ == '=')
                s_++;
            else
                fail("expected = after the outputs");
        }
        if (ok())
            sp_.root = expr(0);
        if (ok()) {
            blank();
            if (*s_)
                fail("unexpected '%c'", *s_);
        }
        if (!ok()) {
            err = err_;
            return false;
        }
        flatten(sp_.root);
        return true;
    }

  private:
    const char        *s_;
    const ts_program  *p_;
    bool               inputs_allowed_;
    ts_spec           &sp_;
    std::string        err_;
    bool               failed_ = false;
    int                waiting_ = 0;  // left operands on the evaluation stack while a right one parses

    bool ok() const { return !failed_; }

    void fail(const char *fmt, ...)
    {
        if (failed_)
            return;
        failed_ = true;
        char b[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(b, sizeof b, fmt, ap);
        va_end(ap);
        err_ = b;
    }

    void blank()
    {
        while (*s_ == ' ' || *s_ == '\t')
            s_++;
    }

    // [A-Za-z_][A-Za-z0-9_]*, kept to its first 63 characters
    bool name(std::string &out)
    {
        if (!isalpha((unsigned char)*s_) && *s_ != '_')
            return false;
        out.clear();
        for (; isalnum((unsigned char)*s_) || *s_ == '_'; s_++)
            if (out.size() < kNameChars)
                out.push_back(*s_);
        return true;
    }

    // Left side: output vectors joined by ':', the first one the highest bits.
    void outputs()
    {
        std::vector<std::vector<int>> parts;
        int total = 0;
        for (;;) {
            std::string nm;
            blank();
            if (!name(nm)) {
                fail("expected an output name before =");
                return;
            }
            std::vector<int> idx(64);
            const int w = ts_vec_resolve(p_, p_->outputs, p_->noutputs, nm.c_str(), idx.data());
            if (w <= 0) {
                fail(w < 0 ? "output vector %s has missing bits or more than 64"
                           : "no output called %s",
                     nm.c_str());
                return;
            }
            idx.resize((size_t)w);
            parts.push_back(std::move(idx));
            total += w;
            blank();
            if (*s_ != ':' || (int)parts.size() == kMaxOutVectors)
                break;
            s_++;
        }
        if (total > 64) {
            fail("the left side has %d bits, at most 64 work", total);
            return;
        }
        for (auto it = parts.rbegin(); it != parts.rend(); ++it)
            for (int i : *it)
                sp_.out.push_back(p_->outputs[i]);
    }

    int add(int op, uint64_t v, int lhs, int rhs)
    {
        if ((int)sp_.nodes.size() == kMaxItems) {
            fail("the expression is too long");
            return -1;
        }
        if ((op == TS_T_NUM || op == TS_T_VAR) && waiting_ + 1 > kMaxDepth) {
            fail("the expression is nested too deeply");
            return -1;
        }
        sp_.nodes.push_back({op, v, lhs, rhs});
        return (int)sp_.nodes.size() - 1;
    }

    // an input vector's index, added on first use
    int variable(const std::string &nm)
    {
        for (size_t k = 0; k < sp_.vars.size(); k++)
            if (same_fold(sp_.vars[k].name.data(), sp_.vars[k].name.size(), nm.data(), nm.size()))
                return (int)k;
        if (!inputs_allowed_) {
            fail("'%s': the right side must be a number here", nm.c_str());
            return -1;
        }
        if ((int)sp_.vars.size() == kMaxVars) {
            fail("too many input vectors");
            return -1;
        }
        std::vector<int> idx(64);
        const int w = ts_vec_resolve(p_, p_->inputs, p_->ninputs, nm.c_str(), idx.data());
        if (w <= 0) {
            fail(w < 0 ? "input vector %s has missing bits or more than 64" : "no input called %s",
                 nm.c_str());
            return -1;
        }
        ts_spec::Var v;
        v.name = nm;
        for (int b = 0; b < w; b++)
            v.pos.push_back((int)p_->ninputs - 1 - idx[(size_t)b]);
        sp_.vars.push_back(std::move(v));
        return (int)sp_.vars.size() - 1;
    }

    // prefix forms: ~x, -x (binding tighter than any binary operator), (e), numbers, names
    int operand()
    {
        blank();
        const char c = *s_;
        if (c == '~' || c == '-') {
            s_++;
            const int x = operand();
            return ok() ? add(c == '~' ? TS_T_NOT : TS_T_NEG, 0, x, -1) : -1;
        }
        if (c == '(') {
            s_++;
            const int x = expr(0);
            if (!ok())
                return -1;
            blank();
            if (*s_ != ')') {
                fail("expected )");
                return -1;
            }
            s_++;
            return x;
        }
        if (isdigit((unsigned char)c)) {
            char *end = nullptr;
            const unsigned long long v = strtoull(s_, &end, 0);  // 0x.. hex, 0.. octal, decimal
            s_ = end;
            return add(TS_T_NUM, v, -1, -1);
        }
        std::string nm;
        if (name(nm)) {
            const int k = variable(nm);
            return k < 0 ? -1 : add(TS_T_VAR, (uint64_t)k, -1, -1);
        }
        if (c)
            fail("unexpected '%c'", c);
        else
            fail("the expression ends too early");
        return -1;
    }

    // the binary operator at the cursor: its node kind, binding power (higher binds tighter) and
    // length; false if there is none (a lone '<' or '>' is not an operator)
    bool infix(int *op, int *bp, int *len) const
    {
        *len = 1;
        switch (*s_) {
        case '|': *op = TS_T_OR; *bp = 1; return true;
        case '^': *op = TS_T_XOR; *bp = 2; return true;
        case '&': *op = TS_T_AND; *bp = 3; return true;
        case '<': *op = TS_T_SHL; *bp = 4; *len = 2; return s_[1] == '<';
        case '>': *op = TS_T_SHR; *bp = 4; *len = 2; return s_[1] == '>';
        case '+': *op = TS_T_ADD; *bp = 5; return true;
        case '-': *op = TS_T_SUB; *bp = 5; return true;
        case '*': *op = TS_T_MUL; *bp = 6; return true;
        default: return false;
        }
    }

    // operators of binding power >= min_bp, left-associative
    int expr(int min_bp)
    {
        int lhs = operand();
        while (ok()) {
            blank();
            int op, bp, len;
            if (!infix(&op, &bp, &len) || bp < min_bp)
                break;
            s_ += len;
            waiting_++;
            const int rhs = expr(bp + 1);
            waiting_--;
            if (!ok())
                return -1;
            lhs = add(op, 0, lhs, rhs);
        }
        return ok() ? lhs : -1;
    }

    void flatten(int n)
    {
        const ts_spec::Node &nd = sp_.nodes[(size_t)n];
        if (nd.lhs >= 0)
            flatten(nd.lhs);
        if (nd.rhs >= 0)
            flatten(nd.rhs);
        sp_.postfix.push_back({nd.op, nd.v});
    }
};

uint64_t eval(const ts_spec &s, int n, const uint64_t *var)
{
    const ts_spec::Node &nd = s.nodes[(size_t)n];
    switch (nd.op) {
    case TS_T_NUM: return nd.v;
    case TS_T_VAR: return var[nd.v];
    case TS_T_NOT: return ~eval(s, nd.lhs, var);
    case TS_T_NEG: return 0 - eval(s, nd.lhs, var);
    default: break;
    }
    const uint64_t a = eval(s, nd.lhs, var), b = eval(s, nd.rhs, var);
    switch (nd.op) {
    case TS_T_ADD: return a + b;
    case TS_T_SUB: return a - b;
    case TS_T_MUL: return a * b;
    case TS_T_AND: return a & b;
    case TS_T_OR: return a | b;
    case TS_T_XOR: return a ^ b;
    case TS_T_SHL: return b < 64 ? a << b : 0;
    default: return b < 64 ? a >> b : 0;  // TS_T_SHR
    }
}

}  // namespace

extern "C" ts_spec *ts_spec_parse(const char *text, const ts_program *p, int inputs_allowed,
                                  char *err, size_t errlen)
{
    try {
        ts_spec *s = new ts_spec();
        std::string msg;
        SpecParser ps(text, p, inputs_allowed != 0, *s);
        if (!ps.parse(msg)) {
            delete s;
            snprintf(err, errlen, "%s", msg.c_str());
            return nullptr;
        }
        return s;
    } catch (const std::bad_alloc &) {
        snprintf(err, errlen, "out of memory");
        return nullptr;
    }
}

extern "C" void ts_spec_free(ts_spec *s) { delete s; }
extern "C" int ts_spec_width(const ts_spec *s) { return (int)s->out.size(); }
extern "C" uint32_t ts_spec_out_cell(const ts_spec *s, int bit) { return s->out[(size_t)bit]; }
extern "C" int ts_spec_nrpn(const ts_spec *s) { return (int)s->postfix.size(); }
extern "C" const ts_spec_item *ts_spec_rpn(const ts_spec *s) { return s->postfix.data(); }
extern "C" int ts_spec_nvars(const ts_spec *s) { return (int)s->vars.size(); }
extern "C" int ts_spec_var_bits(const ts_spec *s, int k) { return (int)s->vars[(size_t)k].pos.size(); }
extern "C" int ts_spec_var_pos(const ts_spec *s, int k, int bit)
{
    return s->vars[(size_t)k].pos[(size_t)bit];
}

extern "C" uint64_t ts_spec_value(const ts_spec *s, uint64_t x)
{
    uint64_t var[kMaxVars];
    for (size_t k = 0; k < s->vars.size(); k++) {
        uint64_t v = 0;
        const std::vector<int> &pos = s->vars[k].pos;
        for (size_t b = 0; b < pos.size(); b++)
            v |= (x >> pos[b] & 1) << b;
        var[k] = v;
    }
    const uint64_t r = eval(*s, s->root, var);
    const size_t w = s->out.size();
    return w >= 64 ? r : r & ((1ull << w) - 1);
}
