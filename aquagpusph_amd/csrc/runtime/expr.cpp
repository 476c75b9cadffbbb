#include "expr.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>

#include "../aq_internal.hpp"

namespace aq {
namespace rt {

namespace {

const std::map<std::string, int32_t>& fn1_table()
{
    static const std::map<std::string, int32_t> t = {
        { "sin", F_SIN },     { "cos", F_COS },     { "tan", F_TAN },
        { "asin", F_ASIN },   { "acos", F_ACOS },   { "atan", F_ATAN },
        { "sinh", F_SINH },   { "cosh", F_COSH },   { "tanh", F_TANH },
        { "asinh", F_ASINH }, { "acosh", F_ACOSH }, { "atanh", F_ATANH },
        { "log2", F_LOG2 },   { "log10", F_LOG10 }, { "log", F_LN },
        { "ln", F_LN },       { "exp", F_EXP },     { "sqrt", F_SQRT },
        { "sign", F_SIGN },   { "rint", F_RINT },   { "abs", F_ABS },
        { "floor", F_FLOOR }, { "ceil", F_CEIL },
    };
    return t;
}

const std::map<std::string, int32_t>& fn2_table()
{
    static const std::map<std::string, int32_t> t = {
        { "min", F_MIN }, { "max", F_MAX }, { "pow", F_POW2 },
        { "atan2", F_ATAN2 }, { "mod", F_MOD },
    };
    return t;
}

struct P {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& m)
    {
        throw Error(AQ_EPARSE, "expression \"" + s + "\": " + m +
                                   " at position " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && isspace((unsigned char)s[i])) i++;
    }
    bool eat(const char* t)
    {
        ws();
        const size_t n = strlen(t);
        if (s.compare(i, n, t) == 0) {
            i += n;
            return true;
        }
        return false;
    }
    ExprP mk(Expr::Kind k)
    {
        auto e = std::make_shared<Expr>();
        e->kind = k;
        return e;
    }
    ExprP bin(const std::string& op, ExprP a, ExprP b)
    {
        auto e = mk(Expr::BINARY);
        e->op = op;
        e->args = { a, b };
        return e;
    }
    ExprP ternary()
    {
        ExprP c = lor();
        if (eat("?")) {
            ExprP a = ternary();
            if (!eat(":")) fail("expected ':'");
            ExprP b = ternary();
            auto e = mk(Expr::TERNARY);
            e->args = { c, a, b };
            return e;
        }
        return c;
    }
    ExprP lor()
    {
        ExprP a = land();
        while (eat("||")) a = bin("||", a, land());
        return a;
    }
    ExprP land()
    {
        ExprP a = cmp();
        while (eat("&&")) a = bin("&&", a, cmp());
        return a;
    }
    ExprP cmp()
    {
        ExprP a = add();
        for (;;) {
            if (eat("==")) a = bin("==", a, add());
            else if (eat("!=")) a = bin("!=", a, add());
            else if (eat("<=")) a = bin("<=", a, add());
            else if (eat(">=")) a = bin(">=", a, add());
            else if (eat("<")) a = bin("<", a, add());
            else if (eat(">")) a = bin(">", a, add());
            else return a;
        }
    }
    ExprP add()
    {
        ExprP a = mul();
        for (;;) {
            if (eat("+")) a = bin("+", a, mul());
            else if (eat("-")) a = bin("-", a, mul());
            else return a;
        }
    }
    ExprP mul()
    {
        ExprP a = unary();
        for (;;) {
            if (eat("*")) a = bin("*", a, unary());
            else if (eat("/")) a = bin("/", a, unary());
            else return a;
        }
    }
    // unary sign binds looser than '^' (muparser: -2^2 = -4)
    ExprP unary()
    {
        if (eat("-")) {
            auto e = mk(Expr::UNARY);
            e->op = "-";
            e->args = { unary() };
            return e;
        }
        if (eat("+")) return unary();
        if (eat("!")) {
            auto e = mk(Expr::UNARY);
            e->op = "!";
            e->args = { unary() };
            return e;
        }
        return power();
    }
    ExprP power()
    {
        ExprP a = primary();
        if (eat("^")) return bin("^", a, unary());  // right associative
        return a;
    }
    ExprP primary()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        const char c = s[i];
        if (c == '(') {
            i++;
            ExprP e = ternary();
            if (!eat(")")) fail("expected ')'");
            return e;
        }
        if (isdigit((unsigned char)c) || c == '.') {
            const char* b = s.c_str() + i;
            char* end = nullptr;
            const double v = strtod(b, &end);
            if (end == b) fail("bad number");
            i += end - b;
            if (i < s.size() && (s[i] == 'f' || s[i] == 'F')) i++;
            auto e = mk(Expr::NUM);
            e->value = v;
            return e;
        }
        if (isalpha((unsigned char)c) || c == '_') {
            const size_t b = i;
            while (i < s.size() &&
                   (isalnum((unsigned char)s[i]) || s[i] == '_'))
                i++;
            const std::string id = s.substr(b, i - b);
            if (eat("(")) {
                auto e = mk(Expr::CALL);
                e->name = id;
                if (!eat(")")) {
                    for (;;) {
                        e->args.push_back(ternary());
                        if (eat(")")) break;
                        if (!eat(",")) fail("expected ',' or ')'");
                    }
                }
                return e;
            }
            auto e = mk(Expr::SYM);
            e->name = id;
            return e;
        }
        fail(std::string("unexpected character '") + c + "'");
    }
};

bool constant(const std::string& n, double& v)
{
    if (n == "_pi" || n == "pi" || n == "M_PI") {
        v = 3.141592653589793238462643383279502884;
        return true;
    }
    if (n == "_e") {
        v = 2.718281828459045235360287471352662498;
        return true;
    }
    if (n == "INFINITY") {
        v = INFINITY;
        return true;
    }
    return false;
}

}  // namespace

ExprP parse_expr(const std::string& text)
{
    P p{ text };
    ExprP e = p.ternary();
    p.ws();
    if (p.i != text.size()) p.fail("unexpected trailing input");
    return e;
}

std::vector<std::string> split_components(const std::string& text)
{
    std::vector<std::string> out;
    int depth = 0;
    std::string cur;
    for (char c : text) {
        if (c == '(') depth++;
        if (c == ')') depth--;
        if (c == ',' && depth == 0) {
            out.push_back(cur);
            cur.clear();
        } else
            cur += c;
    }
    out.push_back(cur);
    for (auto& s : out) {
        size_t b = s.find_first_not_of(" \t\r\n");
        size_t e = s.find_last_not_of(" \t\r\n");
        s = (b == std::string::npos) ? "" : s.substr(b, e - b + 1);
    }
    return out;
}

double fn1(int32_t f, double x)
{
    switch (f) {
    case F_SIN: return sin(x);
    case F_COS: return cos(x);
    case F_TAN: return tan(x);
    case F_ASIN: return asin(x);
    case F_ACOS: return acos(x);
    case F_ATAN: return atan(x);
    case F_SINH: return sinh(x);
    case F_COSH: return cosh(x);
    case F_TANH: return tanh(x);
    case F_ASINH: return asinh(x);
    case F_ACOSH: return acosh(x);
    case F_ATANH: return atanh(x);
    case F_LOG2: return log2(x);
    case F_LOG10: return log10(x);
    case F_LN: return log(x);
    case F_EXP: return exp(x);
    case F_SQRT: return sqrt(x);
    case F_SIGN: return (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0);
    case F_RINT: return rint(x);
    case F_ABS: return fabs(x);
    case F_FLOOR: return floor(x);
    case F_CEIL: return ceil(x);
    }
    return NAN;
}

double fn2(int32_t f, double a, double b)
{
    switch (f) {
    case F_MIN: return (b < a) ? b : a;
    case F_MAX: return (a < b) ? b : a;
    case F_POW2: return pow(a, b);
    case F_ATAN2: return atan2(a, b);
    case F_MOD: {
        const int w = (int)b < 1 ? 1 : (int)b;
        return (double)((int)a % w);
    }
    }
    return NAN;
}

static double binop(const std::string& op, double a, double b)
{
    if (op == "+") return a + b;
    if (op == "-") return a - b;
    if (op == "*") return a * b;
    if (op == "/") return a / b;
    if (op == "^") return pow(a, b);
    if (op == "<") return a < b;
    if (op == ">") return a > b;
    if (op == "<=") return a <= b;
    if (op == ">=") return a >= b;
    if (op == "==") return a == b;
    if (op == "!=") return a != b;
    if (op == "&&") return (a != 0) && (b != 0);
    if (op == "||") return (a != 0) || (b != 0);
    throw Error(AQ_EPARSE, "unknown operator " + op);
}

double eval_expr(const ExprP& e, const SymbolFn& sym)
{
    switch (e->kind) {
    case Expr::NUM:
        return e->value;
    case Expr::SYM: {
        double v;
        if (sym && sym(e->name, v)) return v;
        if (constant(e->name, v)) return v;
        throw Error(AQ_ENOTFOUND, "unknown symbol '" + e->name +
                                      "' in expression");
    }
    case Expr::UNARY: {
        const double a = eval_expr(e->args[0], sym);
        return e->op == "-" ? -a : (a == 0 ? 1.0 : 0.0);
    }
    case Expr::BINARY:
        return binop(e->op, eval_expr(e->args[0], sym),
                     eval_expr(e->args[1], sym));
    case Expr::TERNARY:
        return eval_expr(e->args[0], sym) != 0 ? eval_expr(e->args[1], sym)
                                                : eval_expr(e->args[2], sym);
    case Expr::CALL: {
        auto f1 = fn1_table().find(e->name);
        if (f1 != fn1_table().end()) {
            if (e->args.size() != 1)
                throw Error(AQ_EPARSE, e->name + "() takes one argument");
            return fn1(f1->second, eval_expr(e->args[0], sym));
        }
        auto f2 = fn2_table().find(e->name);
        if (f2 != fn2_table().end()) {
            if (e->args.empty())
                throw Error(AQ_EPARSE, e->name + "() needs arguments");
            if (f2->second == F_MIN || f2->second == F_MAX) {
                double v = eval_expr(e->args[0], sym);
                for (size_t k = 1; k < e->args.size(); k++)
                    v = fn2(f2->second, v, eval_expr(e->args[k], sym));
                return v;
            }
            if (e->args.size() != 2)
                throw Error(AQ_EPARSE, e->name + "() takes two arguments");
            return fn2(f2->second, eval_expr(e->args[0], sym),
                       eval_expr(e->args[1], sym));
        }
        if (e->name == "sum" || e->name == "avg") {
            double v = 0;
            for (auto& a : e->args) v += eval_expr(a, sym);
            return e->name == "sum" ? v : v / (double)e->args.size();
        }
        throw Error(AQ_EPARSE, "unknown function " + e->name + "()");
    }
    }
    return NAN;
}

void expr_symbols(const ExprP& e, std::set<std::string>& out)
{
    if (e->kind == Expr::SYM) {
        double v;
        if (!constant(e->name, v)) out.insert(e->name);
    }
    for (auto& a : e->args) expr_symbols(a, out);
}

// ---- bytecode -------------------------------------------------------------
void VmProgram::compile(const ExprP& e, const SlotFn& slots)
{
    switch (e->kind) {
    case Expr::NUM:
        ops.push_back({ VM_PUSH, 0, e->value });
        return;
    case Expr::SYM: {
        SlotRef r;
        if (slots(e->name, r)) {
            const int32_t code = r.base == 'f'   ? VM_LOADF
                                 : r.base == 'd' ? VM_LOADD
                                 : r.base == 'i' ? VM_LOADI
                                                 : VM_LOADU;
            ops.push_back({ code, (int32_t)r.offset, 0.0 });
            return;
        }
        double v;
        if (constant(e->name, v)) {
            ops.push_back({ VM_PUSH, 0, v });
            return;
        }
        throw Error(AQ_ENOTFOUND, "unknown scalar '" + e->name +
                                      "' in a device expression");
    }
    case Expr::UNARY:
        compile(e->args[0], slots);
        ops.push_back({ e->op == "-" ? VM_NEG : VM_NOT, 0, 0.0 });
        return;
    case Expr::BINARY: {
        compile(e->args[0], slots);
        compile(e->args[1], slots);
        static const std::map<std::string, int32_t> m = {
            { "+", VM_ADD }, { "-", VM_SUB }, { "*", VM_MUL },
            { "/", VM_DIV }, { "^", VM_POW }, { "<", VM_LT },
            { ">", VM_GT },  { "<=", VM_LE }, { ">=", VM_GE },
            { "==", VM_EQ }, { "!=", VM_NE }, { "&&", VM_AND },
            { "||", VM_OR },
        };
        ops.push_back({ m.at(e->op), 0, 0.0 });
        return;
    }
    case Expr::TERNARY:
        compile(e->args[0], slots);
        compile(e->args[1], slots);
        compile(e->args[2], slots);
        ops.push_back({ VM_SEL, 0, 0.0 });
        return;
    case Expr::CALL: {
        auto f1 = fn1_table().find(e->name);
        if (f1 != fn1_table().end()) {
            if (e->args.size() != 1)
                throw Error(AQ_EPARSE, e->name + "() takes one argument");
            compile(e->args[0], slots);
            ops.push_back({ VM_F1, f1->second, 0.0 });
            return;
        }
        auto f2 = fn2_table().find(e->name);
        if (f2 != fn2_table().end()) {
            if (e->args.empty())
                throw Error(AQ_EPARSE, e->name + "() needs arguments");
            if (f2->second == F_MIN || f2->second == F_MAX) {
                compile(e->args[0], slots);
                for (size_t k = 1; k < e->args.size(); k++) {
                    compile(e->args[k], slots);
                    ops.push_back({ VM_F2, f2->second, 0.0 });
                }
                return;
            }
            if (e->args.size() != 2)
                throw Error(AQ_EPARSE, e->name + "() takes two arguments");
            compile(e->args[0], slots);
            compile(e->args[1], slots);
            ops.push_back({ VM_F2, f2->second, 0.0 });
            return;
        }
        if (e->name == "sum" || e->name == "avg") {
            ops.push_back({ VM_PUSH, 0, 0.0 });
            for (auto& a : e->args) {
                compile(a, slots);
                ops.push_back({ VM_ADD, 0, 0.0 });
            }
            if (e->name == "avg") {
                ops.push_back({ VM_PUSH, 0, (double)e->args.size() });
                ops.push_back({ VM_DIV, 0, 0.0 });
            }
            return;
        }
        throw Error(AQ_EPARSE, "unknown function " + e->name + "()");
    }
    }
}

void VmProgram::store(const SlotRef& s)
{
    const int32_t code = s.base == 'f'   ? VM_STOREF
                         : s.base == 'd' ? VM_STORED
                         : s.base == 'i' ? VM_STOREI
                                         : VM_STOREU;
    ops.push_back({ code, (int32_t)s.offset, 0.0 });
}

void VmProgram::assert_true(int32_t id)
{
    ops.push_back({ VM_ASSERT, id, 0.0 });
}

void VmProgram::end()
{
    ops.push_back({ VM_END, 0, 0.0 });
    int sp = 0, top = 0;
    for (const VmOp& op : ops) {
        switch (op.code) {
        case VM_END: break;
        case VM_PUSH: case VM_LOADF: case VM_LOADI: case VM_LOADU:
        case VM_LOADD: sp++; break;
        case VM_STOREF: case VM_STOREI: case VM_STOREU: case VM_STORED:
        case VM_ASSERT: sp--; break;
        case VM_NEG: case VM_NOT: case VM_F1: break;
        case VM_SEL: sp -= 2; break;
        default: sp--; break;  // binary operators and two-argument calls
        }
        if (sp < 0)
            throw Error(AQ_EINVAL, "scalar expression: malformed program");
        top = std::max(top, sp);
    }
    if (top > VM_STACK)
        throw Error(AQ_EUNSUP, "scalar expression too deep for the device "
                               "interpreter (" + std::to_string(top) +
                               " values at once, at most " +
                               std::to_string(VM_STACK) + ")");
}

}  // namespace rt
}  // namespace aq
