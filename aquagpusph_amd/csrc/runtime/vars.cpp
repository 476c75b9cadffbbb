// Variable registry with a device-resident scalar arena.
#include <cmath>
#include <cstring>

#include "sim.hpp"

namespace aq {
namespace rt {

static bool starts(const std::string& s, const std::string& p)
{
    return s.compare(0, p.size(), p) == 0;
}

static std::string trim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    size_t e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

Type parse_type(const std::string& t0, int dims)
{
    Type T;
    std::string t = trim(t0);
    if (!t.empty() && t.back() == '*') {
        T.array = true;
        t = trim(t.substr(0, t.size() - 1));
    }
    // aliases (Variable.cpp:1547-1625, addr_bits = 32)
    if (t == "int32") t = "int";
    else if (t == "uint32" || t == "uint" || t == "size_t") t = "unsigned int";
    else if (t == "ssize_t") t = "int";
    else if (starts(t, "svec")) t = "uivec" + t.substr(4);
    else if (starts(t, "ssvec")) t = "ivec" + t.substr(5);
    else if (starts(t, "int") && t != "int") t = "ivec" + t.substr(3);
    else if (starts(t, "uint") && t != "uint") t = "uivec" + t.substr(4);
    else if (starts(t, "float") && t != "float") t = "vec" + t.substr(5);
    else if (starts(t, "fvec")) t = "vec" + t.substr(4);
    else if (starts(t, "double") && t != "double") t = "dvec" + t.substr(6);
    const std::string dflt = dims == 3 ? "4" : "2";
    if (t == "vec" || t == "ivec" || t == "uivec" || t == "dvec") t += dflt;

    if (t == "matrix") {
        // float4 (2x2) in 2D, float16 (4x4) in 3D, types/{2D,3D}.h
        if (!T.array)
            throw Error(AQ_EUNSUP, "scalar matrix variables are not supported");
        T.base = 'f';
        T.ncomp = dims == 3 ? 16 : 4;
        T.stored = T.ncomp;
        T.canon = "matrix*";
        return T;
    } else if (t == "float") {
        T.base = 'f';
    } else if (t == "double") {
        T.base = 'd';
    } else if (t == "int") {
        T.base = 'i';
    } else if (t == "unsigned int") {
        T.base = 'u';
    } else {
        std::string pre;
        if (starts(t, "uivec")) {
            T.base = 'u';
            pre = "uivec";
        } else if (starts(t, "ivec")) {
            T.base = 'i';
            pre = "ivec";
        } else if (starts(t, "vec")) {
            T.base = 'f';
            pre = "vec";
        } else if (starts(t, "dvec")) {
            T.base = 'd';
            pre = "dvec";
        } else {
            throw Error(AQ_EUNSUP, "unsupported variable type '" + t0 + "'");
        }
        const std::string n = t.substr(pre.size());
        if (n != "2" && n != "3" && n != "4")
            throw Error(AQ_EUNSUP, "unsupported variable type '" + t0 + "'");
        T.ncomp = std::stoi(n);
    }
    T.stored = T.ncomp == 3 ? 4 : T.ncomp;
    T.canon = t + (T.array ? "*" : "");
    if (!T.array && T.elem_bytes() > 16)
        throw Error(AQ_EUNSUP, "scalar variables of type '" + t0 +
                                   "' do not fit a 16-byte slot");
    return T;
}

Variables::Variables(int dims) : dims_(dims)
{
    host_arena_.assign(ARENA_SLOTS * SLOT_BYTES, 0);
    AQ_HIP(hipMalloc(&dev_arena_, ARENA_SLOTS * SLOT_BYTES));
    AQ_HIP(hipMemset(dev_arena_, 0, ARENA_SLOTS * SLOT_BYTES));
}

Variables::~Variables()
{
    for (auto& v : vars_)
        if (v->dev) (void)hipFree(v->dev);
    if (dev_arena_) (void)hipFree(dev_arena_);
}

Var* Variables::get(const std::string& name)
{
    auto it = by_name_.find(name);
    return it == by_name_.end() ? nullptr : it->second;
}

Var& Variables::at(const std::string& name)
{
    Var* v = get(name);
    if (!v) throw Error(AQ_ENOTFOUND, "undeclared variable '" + name + "'");
    return *v;
}

static const char* SUFFIX[4] = { "_x", "_y", "_z", "_w" };

static double load_comp(const uint8_t* p, char base)
{
    if (base == 'f') return *(const float*)p;
    if (base == 'd') return *(const double*)p;
    if (base == 'i') return *(const int32_t*)p;
    return *(const uint32_t*)p;
}

bool Variables::symbol(const std::string& name, double& out) const
{
    auto it = by_name_.find(name);
    if (it != by_name_.end()) {
        const Var& v = *it->second;
        if (v.is_array()) return false;
        out = load_comp(host_arena_.data() + v.slot, v.type.base);
        return true;
    }
    if (name.size() > 2 && name[name.size() - 2] == '_') {
        for (int k = 0; k < 4; k++) {
            if (name.compare(name.size() - 2, 2, SUFFIX[k]) != 0) continue;
            auto jt = by_name_.find(name.substr(0, name.size() - 2));
            if (jt == by_name_.end() || jt->second->is_array()) return false;
            const Var& v = *jt->second;
            if (k >= v.type.ncomp) return false;
            out = load_comp(host_arena_.data() + v.slot +
                                v.type.comp_bytes() * k,
                            v.type.base);
            return true;
        }
    }
    return false;
}

bool Variables::slot(const std::string& name, SlotRef& r) const
{
    auto it = by_name_.find(name);
    if (it != by_name_.end()) {
        const Var& v = *it->second;
        if (v.is_array() || v.type.ncomp != 1) return false;
        r.offset = v.slot;
        r.base = v.type.base;
        return true;
    }
    if (name.size() > 2 && name[name.size() - 2] == '_') {
        for (int k = 0; k < 4; k++) {
            if (name.compare(name.size() - 2, 2, SUFFIX[k]) != 0) continue;
            auto jt = by_name_.find(name.substr(0, name.size() - 2));
            if (jt == by_name_.end() || jt->second->is_array()) return false;
            const Var& v = *jt->second;
            if (k >= v.type.ncomp) return false;
            r.offset = v.slot + (uint32_t)v.type.comp_bytes() * k;
            r.base = v.type.base;
            return true;
        }
    }
    return false;
}

double Variables::eval(const std::string& expr) const
{
    const std::string t = trim(expr);
    // straight numbers first, like the reference tokenizer
    if (!t.empty()) {
        char* end = nullptr;
        const double v = strtod(t.c_str(), &end);
        if (end && *end == '\0') return v;
    }
    return eval_expr(parse_expr(t), [this](const std::string& n, double& v) {
        return symbol(n, v);
    });
}

void Variables::set_host(Var& v, const std::vector<double>& vals)
{
    uint8_t* p = host_arena_.data() + v.slot;
    for (int k = 0; k < v.type.ncomp && k < (int)vals.size(); k++) {
        const double x = vals[k];
        if (v.type.base == 'f') ((float*)p)[k] = (float)x;
        else if (v.type.base == 'd') ((double*)p)[k] = x;
        else if (v.type.base == 'i') ((int32_t*)p)[k] = (int32_t)x;
        else ((uint32_t*)p)[k] = (uint32_t)(int64_t)x;
    }
    dirty_ = true;
}

std::vector<double> Variables::get_host(const Var& v) const
{
    std::vector<double> out(v.type.ncomp);
    const uint8_t* p = host_arena_.data() + v.slot;
    for (int k = 0; k < v.type.ncomp; k++)
        out[k] = load_comp(p + v.type.comp_bytes() * k, v.type.base);
    return out;
}

void Variables::push(hipStream_t s)
{
    AQ_HIP(hipMemcpyAsync(dev_arena_, host_arena_.data(), host_arena_.size(),
                          hipMemcpyHostToDevice, s));
    AQ_HIP(hipStreamSynchronize(s));
    dirty_ = false;
}

void Variables::pull(hipStream_t s)
{
    AQ_HIP(hipMemcpyAsync(host_arena_.data(), dev_arena_, host_arena_.size(),
                          hipMemcpyDeviceToHost, s));
    AQ_HIP(hipStreamSynchronize(s));
}

void Variables::allocate(Var& v, size_t length)
{
    if (v.dev) {
        AQ_HIP(hipFree(v.dev));
        v.dev = nullptr;
    }
    v.length = length;
    v.bytes = length * v.type.elem_bytes();
    if (v.bytes) {
        AQ_HIP(hipMalloc(&v.dev, v.bytes));
        AQ_HIP(hipMemset(v.dev, 0, v.bytes));
    }
}

Var* Variables::reg(const std::string& name, const std::string& type,
                    const std::string& length, const std::string& value)
{
    if (name.empty()) throw Error(AQ_EINVAL, "empty variable name");
    auto old = by_name_.find(name);
    if (old != by_name_.end()) {
        Var* o = old->second;
        if (o->dev) AQ_HIP(hipFree(o->dev));
        if (!o->is_array()) free_slots_.push_back(o->slot);
        by_name_.erase(old);
        for (size_t k = 0; k < vars_.size(); k++)
            if (vars_[k].get() == o) {
                vars_.erase(vars_.begin() + k);
                break;
            }
    }
    auto v = std::make_unique<Var>();
    v->name = name;
    v->type = parse_type(type, dims_);
    Var* p = v.get();
    if (p->is_array()) {
        p->length_expr = length;
        double n = 0;
        if (!trim(length).empty()) n = eval(length);
        if (!(n >= 0)) throw Error(AQ_EINVAL, "negative length for " + name);
        vars_.push_back(std::move(v));
        by_name_[name] = p;
        allocate(*p, (size_t)n);
    } else {
        if (!free_slots_.empty()) {
            p->slot = free_slots_.back();
            free_slots_.pop_back();
        } else {
            if (next_slot_ >= ARENA_SLOTS)
                throw Error(AQ_ENOMEM, "too many scalar variables");
            p->slot = (next_slot_++) * SLOT_BYTES;
        }
        memset(host_arena_.data() + p->slot, 0, SLOT_BYTES);
        vars_.push_back(std::move(v));
        by_name_[name] = p;
        if (!trim(value).empty()) {
            std::vector<double> vals;
            for (auto& c : split_components(value)) vals.push_back(eval(c));
            if ((int)vals.size() != 1 && (int)vals.size() < p->type.ncomp &&
                p->type.ncomp > 1) {
                // fewer components than the type: the rest stay 0
            }
            set_host(*p, vals);
        }
        dirty_ = true;
    }
    return p;
}

}  // namespace rt
}  // namespace aq
