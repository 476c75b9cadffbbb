// One-thread interpreter of the scalar bytecode (set_scalar / assert tools).
// Runs in double like the reference's host tokenizer and narrows on store
// with the C++ conversions the reference applies (Tokenizer_muparser.hpp).
#include "sim.hpp"

namespace aq {
namespace rt {

__device__ double dfn1(int32_t f, double x)
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

__device__ double dfn2(int32_t f, double a, double b)
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

__global__ void vm_kernel(const VmOp* __restrict__ ops, uint8_t* arena,
                          uint32_t* status, int32_t* fail_id)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double st[VM_STACK];
    int sp = 0;
    for (int pc = 0;; pc++) {
        const VmOp op = ops[pc];
        switch (op.code) {
        case VM_END:
            return;
        case VM_PUSH:
            st[sp++] = op.imm;
            break;
        case VM_LOADF:
            st[sp++] = (double)*(const float*)(arena + op.arg);
            break;
        case VM_LOADI:
            st[sp++] = (double)*(const int32_t*)(arena + op.arg);
            break;
        case VM_LOADU:
            st[sp++] = (double)*(const uint32_t*)(arena + op.arg);
            break;
        case VM_LOADD:
            st[sp++] = *(const double*)(arena + op.arg);
            break;
        case VM_STORED:
            *(double*)(arena + op.arg) = st[--sp];
            break;
        case VM_STOREF:
            *(float*)(arena + op.arg) = (float)st[--sp];
            break;
        case VM_STOREI:
            *(int32_t*)(arena + op.arg) = (int32_t)st[--sp];
            break;
        case VM_STOREU:
            *(uint32_t*)(arena + op.arg) = (uint32_t)(int64_t)st[--sp];
            break;
        case VM_NEG:
            st[sp - 1] = -st[sp - 1];
            break;
        case VM_NOT:
            st[sp - 1] = (st[sp - 1] == 0.0) ? 1.0 : 0.0;
            break;
        case VM_F1:
            st[sp - 1] = dfn1(op.arg, st[sp - 1]);
            break;
        case VM_SEL: {
            const double b = st[--sp], a = st[--sp], c = st[--sp];
            st[sp++] = (c != 0.0) ? a : b;
            break;
        }
        case VM_ASSERT: {
            const double c = st[--sp];
            if (!(c != 0.0)) {
                atomicOr(status, 2u);
                *fail_id = op.arg;
            }
            break;
        }
        default: {
            const double b = st[--sp], a = st[--sp];
            double r = 0.0;
            switch (op.code) {
            case VM_ADD: r = a + b; break;
            case VM_SUB: r = a - b; break;
            case VM_MUL: r = a * b; break;
            case VM_DIV: r = a / b; break;
            case VM_POW: r = pow(a, b); break;
            case VM_LT: r = a < b; break;
            case VM_GT: r = a > b; break;
            case VM_LE: r = a <= b; break;
            case VM_GE: r = a >= b; break;
            case VM_EQ: r = a == b; break;
            case VM_NE: r = a != b; break;
            case VM_AND: r = (a != 0.0) && (b != 0.0); break;
            case VM_OR: r = (a != 0.0) || (b != 0.0); break;
            case VM_F2: r = dfn2(op.arg, a, b); break;
            default: r = NAN;
            }
            st[sp++] = r;
        }
        }
    }
}

void vm_launch(const VmOp* ops, uint8_t* arena, uint32_t* status,
               int32_t* fail_id, hipStream_t s)
{
    hipLaunchKernelGGL(vm_kernel, dim3(1), dim3(64), 0, s, ops, arena, status,
                       fail_id);
}

}  // namespace rt
}  // namespace aq
