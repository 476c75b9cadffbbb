// Scalar expressions of the case files (variable values, set_scalar, assert,
// definitions), evaluated in double and narrowed to the target type like the
// reference tokenizer (aquagpusph/Tokenizer/Tokenizer_muparser.hpp:86-134:
// muparser in double, then an implicit C++ conversion).
//
// The same syntax tree is either evaluated on the host or compiled into a
// small stack bytecode that a one-thread kernel runs on the device, so that
// set_scalar / assert tools read and write the device-resident scalars and a
// whole step can be replayed from a HIP graph without host round trips.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <set>
#include <string>
#include <vector>

namespace aq {
namespace rt {

struct Expr {
    enum Kind { NUM, SYM, UNARY, BINARY, TERNARY, CALL } kind = NUM;
    double value = 0.0;
    std::string name;  // SYM / CALL
    std::string op;    // UNARY / BINARY operator
    std::vector<std::shared_ptr<Expr>> args;
};
typedef std::shared_ptr<Expr> ExprP;

// Parse one expression (no top-level commas).  Throws aq::Error(AQ_EPARSE).
ExprP parse_expr(const std::string& text);
// Split "a, b, c" at top-level commas.
std::vector<std::string> split_components(const std::string& text);

typedef std::function<bool(const std::string&, double&)> SymbolFn;
double eval_expr(const ExprP& e, const SymbolFn& sym);
void expr_symbols(const ExprP& e, std::set<std::string>& out);

// ---- device bytecode ------------------------------------------------------
enum VmCode : int32_t {
    VM_END = 0,
    VM_PUSH,
    VM_LOADF,
    VM_LOADI,
    VM_LOADU,
    VM_STOREF,
    VM_STOREI,
    VM_STOREU,
    VM_NEG,
    VM_NOT,
    VM_ADD,
    VM_SUB,
    VM_MUL,
    VM_DIV,
    VM_POW,
    VM_LT,
    VM_GT,
    VM_LE,
    VM_GE,
    VM_EQ,
    VM_NE,
    VM_AND,
    VM_OR,
    VM_SEL,
    VM_F1,
    VM_F2,
    VM_ASSERT,
    VM_LOADD,
    VM_STORED,
};

enum VmFn1 : int32_t {
    F_SIN, F_COS, F_TAN, F_ASIN, F_ACOS, F_ATAN, F_SINH, F_COSH, F_TANH,
    F_ASINH, F_ACOSH, F_ATANH, F_LOG2, F_LOG10, F_LN, F_EXP, F_SQRT, F_SIGN,
    F_RINT, F_ABS, F_FLOOR, F_CEIL
};
enum VmFn2 : int32_t { F_MIN, F_MAX, F_POW2, F_ATAN2, F_MOD };

struct VmOp {
    int32_t code;
    int32_t arg;
    double imm;
};

// Where a scalar component lives in the device scalar arena.
struct SlotRef {
    uint32_t offset;  // byte offset of the component
    char base;        // 'f' float, 'i' int32, 'u' uint32
};
typedef std::function<bool(const std::string&, SlotRef&)> SlotFn;

// Depth of the device interpreter's value stack.
constexpr int VM_STACK = 32;

class VmProgram {
  public:
    // push code that leaves the value of e on the stack
    void compile(const ExprP& e, const SlotFn& slots);
    void store(const SlotRef& s);
    void assert_true(int32_t id);
    // close the program; raises AQ_EUNSUP when an expression needs more
    // than VM_STACK values at once (the device stack is a fixed array)
    void end();
    std::vector<VmOp> ops;
};

// Host-side reference interpreter of the bytecode (used by tests and by the
// eager executor when no device is wanted).
double fn1(int32_t f, double x);
double fn2(int32_t f, double a, double b);

}  // namespace rt
}  // namespace aq
