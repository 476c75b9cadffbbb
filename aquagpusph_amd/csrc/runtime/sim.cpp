// The runtime: variables + definitions + tool pipeline (CalcServer.cpp).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <dlfcn.h>
#include <filesystem>
#include <limits>

#include "sim.hpp"

namespace fs = std::filesystem;

namespace aq {
namespace rt {

static std::string trim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    size_t e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

// folder holding resources/Presets (the package folder of this library)
static std::string package_dir()
{
    Dl_info info;
    if (dladdr((void*)&package_dir, &info) && info.dli_fname)
        return fs::path(info.dli_fname).parent_path().string();
    return fs::current_path().string();
}

Sim::Sim(int d) : dims(d), vars(d)
{
    if (d != 2 && d != 3) throw Error(AQ_EINVAL, "dims must be 2 or 3");
    AQ_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // [0] failure bits, [1] iter at the end of the first failed step
    AQ_HIP(hipMalloc(&d_status, 2 * sizeof(uint32_t)));
    AQ_HIP(hipMalloc(&d_fail_id, sizeof(int32_t)));
    AQ_HIP(hipMemset(d_status, 0, sizeof(uint32_t)));
    AQ_HIP(hipMemset(d_status + 1, 0xff, sizeof(uint32_t)));
    AQ_HIP(hipMemset(d_fail_id, 0xff, sizeof(int32_t)));
}

Sim::~Sim()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& kv : segments) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    tools.clear();
    for (void* b : { (void*)fusion.near_buf, (void*)clist.cell_end,
                     (void*)clist.heads, (void*)clist.ocells,
                     (void*)clist.wcells, (void*)clist.nlist,
                     (void*)clist.bcells, (void*)clist.bcount,
                     (void*)clist.bbase, (void*)clist.bsums,
                     (void*)clist.n_heads, clist.ws,
                     (void*)sel_list, (void*)sel_count })
        if (b) (void)hipFree(b);
    if (fusion.ev_fork) (void)hipEventDestroy(fusion.ev_fork);
    if (fusion.ev_join) (void)hipEventDestroy(fusion.ev_join);
    if (fusion.side_stream) (void)hipStreamDestroy(fusion.side_stream);
    if (fusion.ev0) (void)hipEventDestroy(fusion.ev0);
    if (fusion.ev1) (void)hipEventDestroy(fusion.ev1);
    if (d_status) (void)hipFree(d_status);
    if (d_fail_id) (void)hipFree(d_fail_id);
    if (stream) (void)hipStreamDestroy(stream);
}

void* Sim::ptr(Var* v)
{
    return v->is_array() ? v->dev : vars.scalar_dev(*v);
}

void* Sim::ptr(const std::string& name) { return ptr(&vars.at(name)); }

int Sim::assert_id(const std::string& name)
{
    assert_names.push_back(name);
    return (int)assert_names.size() - 1;
}

std::string Sim::def(const std::string& n) const
{
    auto it = def_values.find(n);
    if (it == def_values.end())
        throw Error(AQ_ENOTFOUND, "undefined preset definition '" + n + "'");
    return it->second;
}

float Sim::def_float(const std::string& n) const
{
    std::string v = def(n);
    if (!v.empty() && (v.back() == 'f' || v.back() == 'F')) v.pop_back();
    return strtof(v.c_str(), nullptr);
}

Problem load_problem(const std::string& xml_path, int dims)
{
    Problem P;
    P.dims = dims;
    P.base_path = package_dir();
    parse_problem(xml_path, P, "");
    // the <Reports> run after every tool (CalcServer.cpp:472-520)
    for (auto& r : P.reports) P.tools.push_back(r);
    P.reports.clear();
    return P;
}

void Sim::load(const std::string& xml_path)
{
    Problem P = load_problem(xml_path, dims);
    build(P);
}

static unsigned long next_pow2(unsigned long n)
{
    unsigned long p = 1;
    while (p < n) p <<= 1;
    return p;
}

void Sim::build(Problem& P)
{
    problem = P;
    size_t n = 0;
    for (auto& s : P.sets) n += s.n;
    if (n == 0) throw Error(AQ_EINVAL, "the case defines no particles");
    if (n > 0xFFFFFFF0ull) throw Error(AQ_EINVAL, "too many particles");
    N = (uint32_t)n;
    // built-in variables, CalcServer.cpp:178-230
    const std::string sN = std::to_string(N);
    vars.reg("mpi_rank", "unsigned int", "", "0");
    vars.reg("mpi_size", "unsigned int", "", "1");
    vars.reg("dims", "unsigned int", "", std::to_string(dims));
    vars.reg("t", "float", "", "0");
    vars.reg("dt", "float", "", "0");
    vars.reg("iter", "unsigned int", "", "0");
    vars.reg("frame", "unsigned int", "", "0");
    {
        char b[64];
        snprintf(b, sizeof(b), "%.9g", (double)(FLT_MAX / 2.f));
        vars.reg("end_t", "float", "", b);
    }
    vars.reg("end_iter", "unsigned int", "",
             std::to_string(std::numeric_limits<int32_t>::max()));
    vars.reg("end_frame", "unsigned int", "",
             std::to_string(std::numeric_limits<int32_t>::max()));
    vars.reg("N", "size_t", "", sN);
    vars.reg("n_sets", "unsigned int", "", std::to_string(P.sets.size()));
    vars.reg("n_radix", "size_t", "",
             std::to_string(next_pow2((N + 1023) / 1024 * 1024)));
    vars.reg("n_cells", "svec4", "", "1, 1, 1, 1");
    vars.reg("support", "float", "", "2");
    vars.reg("id", "size_t*", sN, "");
    vars.reg("r", "vec*", sN, "");
    vars.reg("iset", "unsigned int*", sN, "");
    vars.reg("id_sorted", "size_t*", sN, "");
    vars.reg("id_unsorted", "size_t*", sN, "");
    vars.reg("icell", "size_t*", sN, "");
    vars.reg("ihoc", "size_t*", "n_cells_w", "");
    // user variables, in registration order
    for (auto& v : P.variables)
        vars.reg(std::get<0>(v), std::get<1>(v), std::get<2>(v),
                 std::get<3>(v));
    // end criteria
    if (P.timing.end_t >= 0) vars.set_host(vars.at("end_t"), { P.timing.end_t });
    if (P.timing.end_iter >= 0)
        vars.set_host(vars.at("end_iter"), { (double)P.timing.end_iter });
    if (P.timing.end_frame >= 0)
        vars.set_host(vars.at("end_frame"), { (double)P.timing.end_frame });
    definitions = P.definitions;
    make_defs();
    // tools: one instance per description (a description inserted at
    // several places runs at each of them)
    std::map<const ToolDesc*, std::shared_ptr<Tool>> made;
    for (auto& d : P.tools) {
        auto it = made.find(d.get());
        if (it == made.end()) it = made.emplace(d.get(), make_tool(*this, *d)).first;
        tools.push_back(it->second);
    }
    load_sets();
}

// CalcServer.cpp:241-260: evaluated definitions become "%#G" + "f"
void Sim::make_defs()
{
    def_values.clear();
    for (auto& d : definitions) {
        std::string v = d.value;
        if (d.evaluate && !d.value.empty()) {
            const float f = (float)vars.eval(d.value);
            char b[128];
            snprintf(b, sizeof(b), "%#G", (double)f);
            v = std::string(b) + "f";
        }
        def_values[d.name] = v;
    }
    auto num = [&](const std::string& n, float dflt) {
        return defined(n) && !def(n).empty() ? def_float(n) : dflt;
    };
    // the Wendland kernel is the only kernel family built
    if (defined("KERNEL_NAME") && trim(def("KERNEL_NAME")) != "Wendland")
        throw Error(AQ_EUNSUP, "KERNEL_NAME=" + def("KERNEL_NAME") +
                                   " is not supported (Wendland only)");
    const float h = (float)vars.get_host(vars.at(vars.get("h") ? "h" : "support"))[0];
    kdefs.H = num("H", h);
    kdefs.CONW = num("CONW", 1.f / std::pow(kdefs.H, (float)dims));
    kdefs.CONF = num("CONF", 1.f / std::pow(kdefs.H, (float)dims + 2.f));
    kdefs.DIMS = num("DIMS", (float)dims);
    kdefs.lap_formulation = AQ_LAP_MONAGHAN;
    if (defined("__LAP_FORMULATION__")) {
        std::string v = trim(def("__LAP_FORMULATION__"));
        for (int k = 0; k < 4 && defined(v); k++) v = trim(def(v));
        kdefs.lap_formulation = atoi(v.c_str());
        if (kdefs.lap_formulation != AQ_LAP_MONAGHAN &&
            kdefs.lap_formulation != AQ_LAP_MORRIS)
            throw Error(AQ_EUNSUP, "unknown __LAP_FORMULATION__ " + v);
    }
}

void Sim::load_sets()
{
    // iset / id of every particle, per-set scalars, then the particle files
    std::vector<uint32_t> iset(N), id(N);
    size_t off = 0;
    for (size_t k = 0; k < problem.sets.size(); k++) {
        for (size_t i = 0; i < problem.sets[k].n; i++) {
            iset[off + i] = (uint32_t)k;
            id[off + i] = (uint32_t)(off + i);
        }
        off += problem.sets[k].n;
    }
    AQ_HIP(hipMemcpy(vars.at("iset").dev, iset.data(), N * 4,
                     hipMemcpyHostToDevice));
    AQ_HIP(hipMemcpy(vars.at("id").dev, id.data(), N * 4,
                     hipMemcpyHostToDevice));
    for (size_t k = 0; k < problem.sets.size(); k++)
        for (auto& sc : problem.sets[k].scalars) {
            Var& v = vars.at(sc.first);
            if (!v.is_array() || v.length <= k)
                throw Error(AQ_EINVAL, "set scalar '" + sc.first +
                                           "' must be an array of length "
                                           "n_sets");
            std::vector<double> vals;
            for (auto& c : split_components(sc.second))
                vals.push_back(vars.eval(c));
            if (v.type.base == 'd')
                throw Error(AQ_EUNSUP, "set scalar '" + sc.first +
                                           "': double arrays are not "
                                           "supported here");
            std::vector<uint32_t> w(v.type.stored, 0);
            for (int c = 0; c < v.type.ncomp && c < (int)vals.size(); c++) {
                if (v.type.base == 'f') {
                    float f = (float)vals[c];
                    memcpy(&w[c], &f, 4);
                } else if (v.type.base == 'i') {
                    int32_t x = (int32_t)vals[c];
                    memcpy(&w[c], &x, 4);
                } else
                    w[c] = (uint32_t)(int64_t)vals[c];
            }
            AQ_HIP(hipMemcpy((char*)v.dev + k * v.type.elem_bytes(), w.data(),
                             v.type.elem_bytes(), hipMemcpyHostToDevice));
        }
    off = 0;
    for (auto& s : problem.sets) {
        for (auto& l : s.loads) {
            std::string f = l.format;
            for (auto& c : f) c = (char)tolower((unsigned char)c);
            if (f != "fastascii" && f != "ascii")
                throw Error(AQ_EUNSUP, "particles format '" + l.format +
                                           "' is not supported");
            load_fastascii(*this, l.file, off, s.n, l.fields);
        }
        off += s.n;
    }
}

void Sim::setup()
{
    for (auto& t : tools) t->setup(*this);
    match_scopes();
    // cell-parallel walks unless AQ_WALK=thread asks for the per-particle
    // ones (which also serve more than 2^28 particles)
    const char* walk = getenv("AQ_WALK");
    clist = CellList();
    clist.enabled = !(walk && strcmp(walk, "thread") == 0) && N < (1u << 28);
    clist.icell = vars.get("icell");
    clist.ihoc = vars.get("ihoc");
    clist.n_cells = vars.get("n_cells");
    // boundary-integral element kernels: cell-parallel unless
    // AQ_BI_WALK=thread asks for the per-particle forms (over compacted
    // particle lists while the cell list is on); the kernels whose owners
    // are wall elements or other non-fluid particles walk only the cells
    // holding such particles (the owner-cell list) unless AQ_BI_WALK=cells
    // asks them to visit every cell
    const char* bw = getenv("AQ_BI_WALK");
    bi_cells = clist.enabled && !(bw && strcmp(bw, "thread") == 0);
    bi_group = bi_cells && !(bw && strcmp(bw, "cells") == 0);
    // the fused walk hands its cells out in blocked order unless
    // AQ_WALK_ORDER=rows asks for the row-major cell list
    const char* wo = getenv("AQ_WALK_ORDER");
    clist.blocked = clist.enabled && !(wo && strcmp(wo, "rows") == 0);
    if (clist.enabled && N && !sel_list) {
        AQ_HIP(hipMalloc(&sel_list, (size_t)N * sizeof(uint32_t)));
        AQ_HIP(hipMalloc(&sel_count, sizeof(uint32_t)));
    }
    plan_fusion();
    plan_tail();
    plan_batches();
    // the compensated sums live in the fused cell walk only
    if (problem.is_defined("__KAHAN_SUMS__") &&
        !(fusion.active && clist.enabled))
        throw Error(AQ_EUNSUP,
                    "__KAHAN_SUMS__ needs the fused cell-parallel walk (cfd "
                    "Shepard + cfd interactions + cfd lap p in a row, no "
                    "AQ_WALK=thread or AQ_NO_FUSION)");
    check_periodic();
    vars.push(stream);
    AQ_HIP(hipStreamSynchronize(stream));
    is_setup = true;
}

static bool getenv_true(const char* name);

// Runs of consecutive set tools, of copy tools, each in one launch
// (Tool::batch_kind), and of set_scalar tools, unless AQ_NO_BATCH is set.
// A run ends at a tool of another kind (control tools included, so the
// pipeline never jumps into one), at a `once` tool or at a tool whose
// arrays meet what the run writes (or that writes what it reads), and holds
// at most 8 tools; dummy tools in between are passed over.
void Sim::plan_batches()
{
    for (auto& t : tools) {
        t->batch_next.clear();
        t->batch_member = false;
    }
    no_batch = getenv_true("AQ_NO_BATCH");
    if (no_batch) return;
    size_t a = 0;
    while (a < tools.size()) {
        const int k = tools[a]->batch_kind();
        if (!k || tools[a]->once()) {
            a++;
            continue;
        }
        std::vector<const void*> rd, wr;
        tools[a]->batch_io(rd, wr);
        size_t b = a + 1, n = 1;
        while (b < tools.size() && n < 8) {
            if (tools[b]->batch_transparent()) {
                b++;
                continue;
            }
            if (tools[b]->batch_kind() != k || tools[b]->once()) break;
            std::vector<const void*> r2, w2;
            tools[b]->batch_io(r2, w2);
            auto meets = [](const std::vector<const void*>& x,
                            const std::vector<const void*>& y) {
                for (const void* p : x)
                    if (std::find(y.begin(), y.end(), p) != y.end())
                        return true;
                return false;
            };
            if (meets(r2, wr) || meets(w2, wr) || meets(w2, rd)) break;
            rd.insert(rd.end(), r2.begin(), r2.end());
            wr.insert(wr.end(), w2.begin(), w2.end());
            tools[a]->batch_next.push_back(tools[b].get());
            tools[b]->batch_member = true;
            b++;
            n++;
        }
        if (!tools[a]->batch_next.empty()) tools[a]->batch_setup(*this);
        a = b;
    }
}

// A periodic box (periodic_L declared): every neighbour walk of the
// pipeline must have a periodic form.  Those are the fused cell walk (cfd
// Shepard + cfd interactions + cfd lap p, with the cell list) and the
// cell-parallel LES gradient and viscous term, and the sensors; any other
// kernel that walks the link-list is refused.  Periodic boxes hold fluid
// particles (and sensors) only.
void Sim::check_periodic()
{
    if (!vars.get("periodic_L")) return;
    static const std::set<std::string> fused = {
        "cfd/Shepard.cl:entry", "cfd/Interactions.cl:entry",
        "cfd/deltaSPH.cl:lapp"
    };
    static const std::set<std::string> cells = {
        "cfd/LES/GradU.cl:entry", "cfd/LES/Interactions.cl:entry"
    };
    // per-particle walks with a periodic form
    static const std::set<std::string> part = { "cfd/Sensors.cl:entry" };
    for (auto& t : tools) {
        const std::string k = t->key();
        if (k.empty() || t->type() != "kernel") continue;
        const KernelSpec* spec = find_kernel(k);
        if (!spec) continue;
        bool walks = false;
        for (auto& a : spec->args) walks = walks || a == "ihoc";
        if (!walks) continue;
        if (fused.count(k)) {
            if (!(fusion.active && clist.enabled &&
                  (fusion.anchor == t.get() || fusion.absorbed.count(t.get()))))
                throw Error(AQ_EUNSUP,
                            "periodic box: tool \"" + t->name() +
                                "\" walks the neighbours outside the fused "
                                "cell-parallel walk (AQ_NO_FUSION, "
                                "AQ_WALK=thread or a tool in between)");
            continue;
        }
        if (cells.count(k) && clist.enabled) continue;
        if (part.count(k)) continue;
        throw Error(AQ_EUNSUP, "periodic box: tool \"" + t->name() + "\" (" +
                                   k + ") has no periodic neighbour walk");
    }
}

// Kernel tools that neither move particles nor change imove, r, u, rho, m, p
// of fluid particles: they may sit between fused tools.
static bool getenv_true(const char* name)
{
    const char* e = getenv(name);
    return e && *e && strcmp(e, "0") != 0;
}

static bool fusion_transparent(const Tool& t)
{
    if (t.type() == "dummy" || t.type().compare(0, 7, "report_") == 0)
        return !t.once();
    static const std::set<std::string> ok = {
        "cfd/Sensors.cl:entry",
        "cfd/SensorsRenormalization.cl:entry",
        "cfd/deltaSPH.cl:lapp",
        "cfd/Boundary/BI/LapU.cl:freeslip",
        "cfd/Boundary/BI/GradP.cl:freeslip",
        "cfd/Boundary/BI/Interpolation.cl:entry",
        "cfd/Boundary/BI/InterpolationShepard.cl:entry",
        "cfd/Interactions.cl:entry",
    };
    return t.type() == "kernel" && !t.once() && ok.count(t.key());
}

// The pair kernels compute with IEEE-754 correctly rounded division and
// square root and no contraction (results bit-identical to the oracle)
// unless the case's Device compile_flags ask OpenCL for relaxed math
// (-cl-fast-relaxed-math or -cl-unsafe-math-optimizations): then the fused
// walk uses reciprocal-based division (within the OpenCL C default 2.5
// ulp), the hardware square root (within 3 ulp), contracted multiply-adds
// and products regrouped around 1/rho, as those flags allow.
static bool relaxed_math(const std::string& flags)
{
    for (const char* f : { "-cl-fast-relaxed-math",
                           "-cl-unsafe-math-optimizations" })
        if (flags.find(f) != std::string::npos) return true;
    return false;
}

void Sim::plan_fusion()
{
    fusion = FusionPlan();
    const char* env = getenv("AQ_NO_FUSION");
    if (env && *env && strcmp(env, "0") != 0) return;
    auto find = [&](const std::string& k) -> int {
        int at = -1;
        for (size_t i = 0; i < tools.size(); i++)
            if (tools[i]->key() == k) {
                if (at >= 0) return -2;  // several: do not fuse
                at = (int)i;
            }
        return at;
    };
    const int iI = find("cfd/Interactions.cl:entry");
    if (iI < 0 || tools[iI]->once()) return;
    auto between = [&](int a, int b) {
        for (int k = a + 1; k < b; k++)
            if (!fusion_transparent(*tools[k])) return false;
        return true;
    };
    const int iS2 = find("cfd/Boundary/BI/Shepard.cl:compute");
    const int iS1 = find("cfd/Shepard.cl:entry");
    const int iL = find("cfd/deltaSPH.cl:lapp");
    const int iB = find("cfd/Boundary/BI/Interactions.cl:entry");
    int anchor = iI;
    if (iS2 >= 0 && iS1 < 0 && iS2 < iI && !tools[iS2]->once() &&
        between(iS2, iI)) {
        fusion.shep = 2;
        anchor = iS2;
    } else if (iS1 >= 0 && iS2 < 0 && iS1 < iI && !tools[iS1]->once() &&
               between(iS1, iI)) {
        fusion.shep = 1;
        anchor = iS1;
    }
    if (iL > iI && !tools[iL]->once() && between(iI, iL) && vars.get("lap_p"))
        fusion.lapp = true;
    if (iB > anchor && !tools[iB]->once() && between(anchor, iB))
        fusion.near = true;
    auto V = [&](const char* n) { return vars.get(n); };
    fusion.imove = V("imove");
    fusion.r = V("r");
    fusion.u = V("u");
    fusion.rho = V("rho");
    fusion.m = V("m");
    fusion.p = V("p");
    fusion.normal = V("normal");
    fusion.tangent = V("tangent");
    fusion.binormal = V("binormal");
    fusion.shepard = V("shepard");
    fusion.grad_p = V("grad_p");
    fusion.lap_u = V("lap_u");
    fusion.div_u = V("div_u");
    fusion.lap_p = V("lap_p");
    fusion.icell = V("icell");
    fusion.ihoc = V("ihoc");
    fusion.n_cells = V("n_cells");
    for (Var* v : { fusion.imove, fusion.r, fusion.u, fusion.rho, fusion.m,
                    fusion.p, fusion.grad_p, fusion.lap_u, fusion.div_u,
                    fusion.icell, fusion.ihoc, fusion.n_cells })
        if (!v) return;
    if (fusion.shep && !fusion.shepard) return;
    if (fusion.shep == 2 &&
        !(fusion.normal && fusion.tangent && fusion.binormal))
        return;
    fusion.anchor = tools[anchor].get();
    if (anchor != iI) fusion.absorbed.insert(tools[iI].get());
    if (fusion.lapp) fusion.absorbed.insert(tools[iL].get());
    if (fusion.near) {
        AQ_HIP(hipMalloc(&fusion.near_buf, N));
        AQ_HIP(hipMemset(fusion.near_buf, 1, N));
    }
    // the boundary-integral element chain beside the walk (opt-in,
    // AQ_BI_OVERLAP=1; BI walls with the near flags): BI LapU, GradP,
    // Interpolation and InterpolationShepard between the anchor and BI
    // Interactions, and between the anchor and BI Interactions only the
    // fused tools, dummies and the sensors (which read fluid fields and
    // write sensor fields).  On the 3D dam break it does not pay: the walk,
    // bound by VALU issue at four wavefronts per SIMD, loses more to the
    // slots it leaves the chain than the chain's 1 ms saves.
    if (fusion.shep == 2 && fusion.near && getenv_true("AQ_BI_OVERLAP")) {
        static const char* keys[] = {
            "cfd/Boundary/BI/LapU.cl:freeslip",
            "cfd/Boundary/BI/GradP.cl:freeslip",
            "cfd/Boundary/BI/Interpolation.cl:entry",
            "cfd/Boundary/BI/InterpolationShepard.cl:entry" };
        std::vector<int> at;
        bool ok = true;
        for (const char* k : keys) {
            const int i = find(k);
            ok = ok && i > anchor && i < iB && !tools[i]->once() &&
                 (at.empty() || i > at.back());
            at.push_back(i);
        }
        for (int k = anchor + 1; ok && k < iB; k++) {
            const Tool& t = *tools[k];
            if (std::find(at.begin(), at.end(), k) != at.end()) continue;
            if (k == iI || (fusion.lapp && k == iL)) continue;
            if ((t.type() == "dummy" || t.type().compare(0, 7, "report_") == 0)
                && !t.once())
                continue;
            if (t.type() == "kernel" && !t.once() &&
                (t.key() == "cfd/Sensors.cl:entry" ||
                 t.key() == "cfd/SensorsRenormalization.cl:entry"))
                continue;
            ok = false;
        }
        if (ok) {
            for (int i : at) {
                fusion.side.push_back(tools[i].get());
                fusion.side_set.insert(tools[i].get());
            }
            fusion.join = tools[iB].get();
            AQ_HIP(hipStreamCreateWithFlags(&fusion.side_stream,
                                            hipStreamNonBlocking));
            AQ_HIP(hipEventCreateWithFlags(&fusion.ev_fork,
                                           hipEventDisableTiming));
            AQ_HIP(hipEventCreateWithFlags(&fusion.ev_join,
                                           hipEventDisableTiming));
        }
    }
    if (const char* wb = getenv("AQ_WALK_BLOCKS"))
        fusion.walk_blocks = (unsigned)std::max(0, atoi(wb));
    if (const char* ws = getenv("AQ_WALK_SEGW"))
        fusion.walk_segw = atoi(ws) == 16 ? 16 : 8;
    fusion.math = relaxed_math(problem.compile_flags) ? 1 : 0;
    if (const char* fm = getenv("AQ_MATH"))
        fusion.math = strcmp(fm, "relaxed") == 0 ? 1 : 0;
    if (problem.is_defined("__KAHAN_SUMS__")) fusion.math = 2;
    if (const char* dv = getenv("AQ_DIV"))
        fusion.fast_div = strcmp(dv, "ieee") != 0;
    fusion.active = true;
}

}  // namespace rt

void launch_step_tail(int dims, bool apply, bool delta_sph,
                      const uint32_t* iset, const int32_t* imove,
                      const float* shepard, float* grad_p, float* lap_u,
                      float* div_u, float* dudt, float* drhodt,
                      const float* visc_dyn, const float* g,
                      const float* lap_p, const float* refd,
                      const float* delta, float* r, float* u, float* rho,
                      const float* dudt_in, const float* drhodt_in,
                      uint32_t N, const float* dt, const uint32_t* status,
                      hipStream_t s);

void launch_sort_chain(int dims, const uint32_t* perm, uint32_t N,
                       const uint32_t* id_in, uint32_t* id,
                       const uint32_t* iset_in, uint32_t* iset,
                       const int* imove_in, int* imove, const float* r_in,
                       float* r, const float* normal_in, float* normal,
                       const float* tangent_in, float* tangent,
                       const float* rho_in, float* rho, const float* m_in,
                       float* m, const float* u_in, float* u,
                       const float* dudt, float* dudt_in,
                       const float* drhodt, float* drhodt_in, float* p,
                       const float* refd, const float* cs, const float* p0,
                       float* binormal, hipStream_t s);
void launch_step_head(int dims, int32_t* imove, const float* r,
                      const float* u, const float* dudt, const float* rho,
                      const float* drhodt, float* r_in, float* u_in,
                      float* dudt_in, float* rho_in, float* drhodt_in,
                      float* m, uint32_t N, const float* dt,
                      const uint32_t* status, const float* dmin,
                      const float* dmax, hipStream_t s);
void launch_fluid_pass(int dims, int shep, bool lapp, const int32_t* imove,
                       const float* r, const float* u, const float* rho,
                       const float* m, const float* p, const float* normal,
                       const float* tangent, const float* binormal,
                       float* shepard, float* grad_p, float* lap_u,
                       float* div_u, float* lap_p, uint8_t* near, uint32_t N,
                       const uint32_t* icell, const uint32_t* ihoc,
                       const uint32_t* n_cells, const uint32_t* status,
                       aq_defs d, hipStream_t s);
void launch_bi_shepard_compute_cells(int, const int32_t*, const float*,
                                     const float*, const float*,
                                     const float*, const float*, float*,
                                     const CellArgs&, aq_defs, float, int,
                                     hipStream_t);
void launch_fluid_pass_cells(int dims, int shep, bool lapp,
                             const int32_t* imove, const float* r,
                             const float* u, const float* rho, const float* m,
                             const float* p, const float* normal,
                             const float* tangent, const float* binormal,
                             float* shepard, float* grad_p, float* lap_u,
                             float* div_u, float* lap_p, uint8_t* near,
                             uint32_t N, const uint32_t* icell,
                             const uint32_t* ihoc, const uint32_t* n_cells,
                             const uint32_t* cell_end, const uint32_t* heads,
                             const uint32_t* n_heads, const uint32_t* status,
                             aq_defs d, float smax, unsigned nblocks,
                             int segw, int flagged, int math,
                             unsigned ranges, int fast_div, hipStream_t s,
                             const float* per, unsigned reserve_per_cu,
                             uint32_t* nlist, const uint32_t* blist);
unsigned side_wgs_per_cu();
void launch_shepard(int, const int32_t*, const float*, const float*,
                    const float*, float*, uint32_t, const uint32_t*,
                    const uint32_t*, const uint32_t*, const uint32_t*,
                    aq_defs, int, hipStream_t, const PartWalk*);
void launch_bi_shepard_compute(int, const int32_t*, const float*,
                               const float*, const float*, const float*,
                               const float*, float*, uint32_t,
                               const uint32_t*, const uint32_t*,
                               const uint32_t*, const uint32_t*, aq_defs,
                               hipStream_t, int, const PartWalk*);

namespace rt {

CellArgs Sim::cell_args()
{
    return CellArgs{ (const uint32_t*)clist.icell->dev,
                     (const uint32_t*)clist.ihoc->dev,
                     (const uint32_t*)ptr(clist.n_cells),
                     clist.cell_end,
                     clist.heads,
                     clist.n_heads,
                     d_status,
                     N,
                     clist.kinds ? 1u : 0u,
                     (bi_group && clist.kinds) ? clist.ocells : nullptr,
                     (bi_group && clist.kinds) ? clist.wcells : nullptr,
                     // filled by the fused walk of the same step
                     (bi_group && clist.kinds && fusion.active && fusion.near)
                         ? clist.nlist
                         : nullptr,
                     (clist.blocked && clist.kinds) ? clist.bcells : nullptr,
                     in_side ? AQ_CL_SIDE : 0u };
}

void Sim::join_side(hipStream_t s)
{
    if (!fusion.side_launched) return;
    AQ_HIP(hipStreamWaitEvent(s, fusion.ev_join, 0));
    fusion.side_launched = false;
}

PartWalk Sim::part_walk(const Var* ic, const Var* ih, const Var* nc)
{
    PartWalk pw;
    if (!clist.enabled || !sel_list) return pw;
    pw.list = sel_list;
    pw.count = sel_count;
    if (clist.usable(ic, ih) && nc == clist.n_cells) {
        pw.cell_end = clist.cell_end;
        pw.flagged = clist.kinds ? 1u : 0u;
    }
    return pw;
}

// The step's per-particle tail in one kernel (launch_step_tail): cfd Rates
// and the improved-Euler corrector, with BI Shepard apply before them and
// the delta-SPH rate between them when present, each tool once, in that
// order, with only dummy tools between (so no other tool reads or writes
// their fields in between).  AQ_NO_FUSION turns it off.
void Sim::plan_tail()
{
    tail = TailPlan();
    head = TailPlan();
    sortc = TailPlan();
    const char* env = getenv("AQ_NO_FUSION");
    if (env && *env && strcmp(env, "0") != 0) return;
    {
        // the reorder chain: Sort stage 1, stage 2, EOS, Binormal, each
        // once, in that order, with only dummies between
        const char* keys[4] = { "basic/Sort.cl:stage1", "basic/Sort.cl:stage2",
                                "basic/EOS.cl:entry",
                                "basic/Binormal.cl:entry" };
        int at[4] = { -1, -1, -1, -1 }, cnt[4] = { 0, 0, 0, 0 };
        for (size_t i = 0; i < tools.size(); i++)
            for (int k = 0; k < 4; k++)
                if (tools[i]->key() == keys[k]) {
                    at[k] = (int)i;
                    cnt[k]++;
                }
        bool ok = true;
        for (int k = 0; k < 4 && ok; k++) {
            ok = cnt[k] == 1 && !tools[at[k]]->once() &&
                 (k == 0 || at[k] > at[k - 1]);
            for (int t = k ? at[k - 1] + 1 : at[k]; ok && k && t < at[k]; t++)
                ok = tools[t]->type() == "dummy";
        }
        const size_t vb = (dims == 2 ? 2 : 4) * 4;
        const std::pair<const char*, size_t> need[] = {
            { "id_unsorted", 4 }, { "id_in", 4 }, { "id", 4 },
            { "iset_in", 4 }, { "iset", 4 }, { "imove_in", 4 },
            { "imove", 4 }, { "r_in", vb }, { "r", vb },
            { "normal_in", vb }, { "normal", vb }, { "tangent_in", vb },
            { "tangent", vb }, { "rho_in", 4 }, { "rho", 4 }, { "m_in", 4 },
            { "m", 4 }, { "u_in", vb }, { "u", vb }, { "dudt", vb },
            { "dudt_in", vb }, { "drhodt", 4 }, { "drhodt_in", 4 },
            { "p", 4 }, { "binormal", vb } };
        for (auto& nv : need) {
            Var* v = vars.get(nv.first);
            ok = ok && v && v->is_array() && v->length >= N &&
                 v->type.elem_bytes() == nv.second;
        }
        for (const char* n : { "refd", "cs", "p0" }) ok = ok && vars.get(n);
        if (ok) {
            sortc.anchor = tools[at[0]].get();
            for (int k = 1; k < 4; k++) sortc.absorbed.insert(tools[at[k]].get());
        }
    }
    {
        // the head: predictor, then Domain, with only dummies between
        int iP = -1, iM = -1, nP = 0, nM = 0;
        for (size_t i = 0; i < tools.size(); i++) {
            const std::string k = tools[i]->key();
            if (k == "basic/time_scheme/improved_euler.cl:predictor") {
                iP = (int)i;
                nP++;
            }
            if (k == "basic/Domain.cl:entry") {
                iM = (int)i;
                nM++;
            }
        }
        bool ok = nP == 1 && nM == 1 && iP < iM && !tools[iP]->once() &&
                  !tools[iM]->once();
        for (int t = iP + 1; ok && t < iM; t++)
            ok = tools[t]->type() == "dummy";
        for (const char* n : { "imove", "r", "u", "dudt", "rho", "drhodt",
                               "r_in", "u_in", "dudt_in", "rho_in",
                               "drhodt_in", "dt", "m", "domain_min",
                               "domain_max" })
            ok = ok && vars.get(n);
        if (ok) {
            head.anchor = tools[iP].get();
            head.absorbed.insert(tools[iM].get());
        }
    }
    auto find = [&](const std::string& k) -> int {
        int at = -1;
        for (size_t i = 0; i < tools.size(); i++)
            if (tools[i]->key() == k) {
                if (at >= 0) return -2;
                at = (int)i;
            }
        return at;
    };
    const int iA = find("cfd/Boundary/BI/Shepard.cl:apply");
    const int iR = find("cfd/Rates.cl:entry");
    const int iD = find("cfd/deltaSPH.cl:deltaSPH");
    const int iC = find("basic/time_scheme/improved_euler.cl:corrector");
    if (iR < 0 || iC < 0 || iA == -2 || iD == -2) return;
    std::vector<int> chain;
    if (iA >= 0) chain.push_back(iA);
    chain.push_back(iR);
    if (iD >= 0) chain.push_back(iD);
    chain.push_back(iC);
    for (size_t k = 0; k < chain.size(); k++) {
        if (tools[chain[k]]->once()) return;
        if (k == 0) continue;
        if (chain[k] <= chain[k - 1]) return;
        for (int t = chain[k - 1] + 1; t < chain[k]; t++)
            if (tools[t]->type() != "dummy") return;
    }
    for (const char* n : { "iset", "imove", "grad_p", "lap_u", "div_u",
                           "dudt", "drhodt", "visc_dyn", "g", "r", "u", "rho",
                           "dudt_in", "drhodt_in", "dt" })
        if (!vars.get(n)) return;
    if (iA >= 0 && !vars.get("shepard")) return;
    if (iD >= 0 && !(vars.get("lap_p") && vars.get("refd") &&
                     vars.get("delta")))
        return;
    tail.apply = iA >= 0;
    tail.delta = iD >= 0;
    tail.anchor = tools[chain[0]].get();
    for (size_t k = 1; k < chain.size(); k++)
        tail.absorbed.insert(tools[chain[k]].get());
}

void Sim::run_sort_chain(hipStream_t s)
{
    auto F = [&](const char* n) { return (float*)ptr(vars.get(n)); };
    auto U = [&](const char* n) { return (uint32_t*)ptr(vars.get(n)); };
    launch_sort_chain(dims, U("id_unsorted"), N, U("id_in"), U("id"),
                      U("iset_in"), U("iset"), (const int*)U("imove_in"),
                      (int*)U("imove"), F("r_in"), F("r"), F("normal_in"),
                      F("normal"), F("tangent_in"), F("tangent"), F("rho_in"),
                      F("rho"), F("m_in"), F("m"), F("u_in"), F("u"),
                      F("dudt"), F("dudt_in"), F("drhodt"), F("drhodt_in"),
                      F("p"), F("refd"), F("cs"), F("p0"), F("binormal"), s);
}

void Sim::run_head(hipStream_t s)
{
    auto F = [&](const char* n) { return (float*)ptr(vars.get(n)); };
    launch_step_head(dims, (int32_t*)ptr(vars.get("imove")), F("r"), F("u"),
                     F("dudt"), F("rho"), F("drhodt"), F("r_in"), F("u_in"),
                     F("dudt_in"), F("rho_in"), F("drhodt_in"), F("m"), N,
                     F("dt"), d_status, F("domain_min"), F("domain_max"), s);
}

void Sim::run_tail(hipStream_t s)
{
    auto F = [&](const char* n) { return (float*)ptr(vars.get(n)); };
    launch_step_tail(dims, tail.apply, tail.delta,
                     (const uint32_t*)ptr(vars.get("iset")),
                     (const int32_t*)ptr(vars.get("imove")),
                     tail.apply ? F("shepard") : nullptr, F("grad_p"),
                     F("lap_u"), F("div_u"), F("dudt"), F("drhodt"),
                     F("visc_dyn"), F("g"), tail.delta ? F("lap_p") : nullptr,
                     tail.delta ? F("refd") : nullptr,
                     tail.delta ? F("delta") : nullptr, F("r"), F("u"),
                     F("rho"), F("dudt_in"), F("drhodt_in"), N, F("dt"),
                     d_status, s);
}

void Sim::run_fused(hipStream_t s)
{
    FusionPlan& f = fusion;
    auto P = [&](Var* v) { return v ? ptr(v) : nullptr; };
    const float* fr = (const float*)P(f.r);
    const PartWalk pw = part_walk(f.icell, f.ihoc, f.n_cells);
    // the boundary-integral element chain on the side stream, beside the
    // walk: the non-fluid part of the Shepard tool, then the side tools
    if (f.shep == 2 && !f.side.empty() && !side_off && bi_cells &&
        clist.usable(f.icell, f.ihoc)) {
        hipStream_t ss = f.side_stream;
        AQ_HIP(hipEventRecord(f.ev_fork, s));
        AQ_HIP(hipStreamWaitEvent(ss, f.ev_fork, 0));
        struct InSide {
            bool& b;
            InSide(bool& x) : b(x) { b = true; }
            ~InSide() { b = false; }
        } guard_(in_side);
        launch_bi_shepard_compute_cells(
            dims, (const int32_t*)P(f.imove), fr, (const float*)P(f.normal),
            (const float*)P(f.tangent), (const float*)P(f.binormal),
            (const float*)P(f.m), (float*)P(f.shepard), cell_args(), kdefs,
            support_threshold(kdefs.H), 1, ss);
        for (Tool* t : f.side) t->launch_kernel(*this, ss);
        AQ_HIP(hipEventRecord(f.ev_join, ss));
        f.side_launched = true;
    }
    // non-fluid part of the Shepard tool first (sensors, elements)
    else if (f.shep == 2 && bi_cells && clist.usable(f.icell, f.ihoc))
        launch_bi_shepard_compute_cells(
            dims, (const int32_t*)P(f.imove), fr, (const float*)P(f.normal),
            (const float*)P(f.tangent), (const float*)P(f.binormal),
            (const float*)P(f.m), (float*)P(f.shepard), cell_args(), kdefs,
            support_threshold(kdefs.H), 1, s);
    else if (f.shep == 2)
        launch_bi_shepard_compute(dims, (const int32_t*)P(f.imove), fr,
                                  (const float*)P(f.normal),
                                  (const float*)P(f.tangent),
                                  (const float*)P(f.binormal),
                                  (const float*)P(f.m), (float*)P(f.shepard),
                                  N, (const uint32_t*)P(f.icell),
                                  (const uint32_t*)P(f.ihoc),
                                  (const uint32_t*)P(f.n_cells), d_status,
                                  kdefs, s, 1, &pw);
    else if (f.shep == 1)
        launch_shepard(dims, (const int32_t*)P(f.imove), fr,
                       (const float*)P(f.rho), (const float*)P(f.m),
                       (float*)P(f.shepard), N, (const uint32_t*)P(f.icell),
                       (const uint32_t*)P(f.ihoc),
                       (const uint32_t*)P(f.n_cells), d_status, kdefs, 1 | 2,
                       s, &pw);
    if (f.timing) AQ_HIP(hipEventRecord(f.ev0, s));
    if (clist.usable(f.icell, f.ihoc)) {
        launch_fluid_pass_cells(
            dims, f.shep, f.lapp, (const int32_t*)P(f.imove), fr,
            (const float*)P(f.u), (const float*)P(f.rho),
            (const float*)P(f.m), (const float*)P(f.p),
            (const float*)P(f.normal), (const float*)P(f.tangent),
            (const float*)P(f.binormal), (float*)P(f.shepard),
            (float*)P(f.grad_p), (float*)P(f.lap_u), (float*)P(f.div_u),
            (float*)P(f.lap_p), f.near ? f.near_buf : nullptr, N,
            (const uint32_t*)P(f.icell), (const uint32_t*)P(f.ihoc),
            (const uint32_t*)P(f.n_cells), clist.cell_end, clist.heads,
            clist.n_heads,
            d_status, kdefs, support_threshold(kdefs.H), f.walk_blocks,
            f.walk_segw, clist.kinds ? 1 : 0, f.math, f.walk_ranges,
            f.fast_div ? 1 : 0, s, periodic_L(),
            f.side_launched ? side_wgs_per_cu() : 0u,
            (bi_group && clist.kinds) ? clist.nlist : nullptr,
            clist.blocked ? clist.bcells : nullptr);
        if (f.timing) AQ_HIP(hipEventRecord(f.ev1, s));
        return;
    }
    if (f.math == 2)
        throw Error(AQ_EUNSUP, "__KAHAN_SUMS__: the cell list is not usable "
                               "for this walk");
    if (periodic_L())
        throw Error(AQ_EUNSUP, "periodic box: the cell list is not usable "
                               "for this walk");
    launch_fluid_pass(dims, f.shep, f.lapp, (const int32_t*)P(f.imove), fr,
                      (const float*)P(f.u), (const float*)P(f.rho),
                      (const float*)P(f.m), (const float*)P(f.p),
                      (const float*)P(f.normal), (const float*)P(f.tangent),
                      (const float*)P(f.binormal), (float*)P(f.shepard),
                      (float*)P(f.grad_p), (float*)P(f.lap_u),
                      (float*)P(f.div_u), (float*)P(f.lap_p),
                      f.near ? f.near_buf : nullptr, N,
                      (const uint32_t*)P(f.icell), (const uint32_t*)P(f.ihoc),
                      (const uint32_t*)P(f.n_cells), d_status, kdefs, s);
    if (f.timing) AQ_HIP(hipEventRecord(f.ev1, s));
}

// Scopes of while/if/end (Conditional.cpp setup: an opener's scope closes
// at the first end that brings the count back to zero; an end's opener is
// found the same way backwards).
void Sim::match_scopes()
{
    const size_t n = tools.size();
    jump.assign(n, 0);
    if_taken.assign(n, 0);
    has_control = false;
    std::vector<char> matched(n, 0);
    for (size_t i = 0; i < n; i++) {
        if (tools[i]->scope() <= 0) continue;
        has_control = true;
        int scope = 1;
        size_t j = i;
        while (j + 1 < n) {
            j++;
            scope += tools[j]->scope();
            if (!scope) break;
        }
        if (scope)
            throw Error(AQ_EPARSE, "unclosed scope of tool \"" +
                                       tools[i]->name() + "\": add an end tool");
        jump[i] = j;
        jump[j] = i;
        matched[j] = 1;
    }
    for (size_t i = 0; i < n; i++)
        if (tools[i]->scope() < 0 && !matched[i])
            throw Error(AQ_EPARSE, "tool \"" + tools[i]->name() +
                                       "\" closes no while/if scope");
}

// The tool after a control tool.  while: the body while the condition
// holds, else past its end.  if: the body once when the condition holds
// (its end hands control back, and the if then jumps out).  end: back to
// its opener.
size_t Sim::control(size_t pc, hipStream_t s)
{
    Tool* t = tools[pc].get();
    if (t->scope() < 0) return jump[pc];
    auto* c = static_cast<ControlTool*>(t);
    if (t->type() == "while") return c->eval(*this, s) ? pc + 1 : jump[pc] + 1;
    if (if_taken[pc]) {
        if_taken[pc] = 0;
        return jump[pc] + 1;
    }
    if (c->eval(*this, s)) {
        if_taken[pc] = 1;
        return pc + 1;
    }
    return jump[pc] + 1;
}

// Tool::execute: a `once` tool runs the first time it is reached only.
void Sim::run_tools(hipStream_t s,
                    const std::function<void(size_t, bool)>& hook,
                    bool stop_on_assert)
{
    size_t pc = 0;
    while (pc < tools.size()) {
        Tool* t = tools[pc].get();
        last_tool = t->name();
        if (t->scope()) {
            pc = control(pc, s);
            continue;
        }
        if (!(t->once() && t->done)) {
            if (hook) hook(pc, false);
            t->execute(*this, s);
            if (hook) hook(pc, true);
            t->done = true;
            if (stop_on_assert && t->type() == "assert") {
                stamp_status(s);
                check_status();
            }
        }
        pc++;
    }
    join_side(s);  // a chain not joined by its tool (never left in flight)
    stamp_status(s);
}

void Sim::stamp_status(hipStream_t s)
{
    Var* it = vars.get("iter");
    const uint32_t* iter =
        (it && !it->is_array() && it->type.base == 'u')
            ? (const uint32_t*)vars.scalar_dev(*it)
            : nullptr;
    status_stamp(d_status, iter, s);
}

// Graph replay of a pipeline with control tools: each run of tools between
// control tools is captured once into its own graph; the host walks the
// control flow between them (a condition is read back when evaluated, as
// the reference does).
void Sim::step_segments(int nsteps)
{
    for (int k = 0; k < nsteps; k++) {
        size_t pc = 0;
        while (pc < tools.size()) {
            if (tools[pc]->scope()) {
                last_tool = tools[pc]->name();
                pc = control(pc, stream);
                continue;
            }
            auto it = segments.find(pc);
            if (it == segments.end()) {
                Segment g;
                g.end = pc;
                while (g.end < tools.size() && !tools[g.end]->scope()) g.end++;
                AQ_HIP(hipStreamBeginCapture(stream,
                                             hipStreamCaptureModeThreadLocal));
                try {
                    for (size_t i = pc; i < g.end; i++) {
                        Tool* t = tools[i].get();
                        if (t->once() && t->done) continue;
                        last_tool = t->name();
                        t->execute(*this, stream);
                    }
                } catch (...) {
                    hipGraph_t tmp;
                    (void)hipStreamEndCapture(stream, &tmp);
                    throw;
                }
                AQ_HIP(hipStreamEndCapture(stream, &g.graph));
                AQ_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr,
                                           0));
                it = segments.emplace(pc, g).first;
            }
            AQ_HIP(hipGraphLaunch(it->second.exec, stream));
            pc = it->second.end;
        }
        stamp_status(stream);
    }
}

void Sim::sync() { AQ_HIP(hipStreamSynchronize(stream)); }

void Sim::check_status()
{
    uint32_t st[2] = { 0, 0xFFFFFFFFu };
    int32_t id = -1;
    AQ_HIP(hipMemcpyAsync(st, d_status, 8, hipMemcpyDeviceToHost, stream));
    AQ_HIP(hipMemcpyAsync(&id, d_fail_id, 4, hipMemcpyDeviceToHost, stream));
    AQ_HIP(hipStreamSynchronize(stream));
    // the step that first flagged it, by the iteration counter at its end;
    // the corrector has left the particles as they were since
    const std::string when =
        st[1] == 0xFFFFFFFFu
            ? std::string()
            : " in the step ending at iter = " + std::to_string(st[1]) +
                  " (the particles are kept as they were before it)";
    if (st[0] & 4u)
        throw Error(AQ_EINVAL,
                    "link-list: a periodic axis of the box (periodic_L) is "
                    "shorter than 3 cells of support * h" + when);
    if (st[0] & 1u)
        throw Error(AQ_ENOMEM,
                    "link-list: the cell grid outgrew the ihoc capacity "
                    "(particles left the domain box)" + when);
    if (st[0] & 2u)
        throw Error(AQ_EASSERT,
                    "assertion failed in tool " +
                        (id >= 0 && id < (int)assert_names.size()
                             ? assert_names[id]
                             : std::string("?")) + when);
}

void Sim::step(int nsteps)
{
    if (!is_setup) throw Error(AQ_EINVAL, "setup() has not been called");
    if (vars.host_dirty()) vars.push(stream);
    for (int k = 0; k < nsteps; k++) run_tools(stream, {}, true);
    check_status();
}

void Sim::step_graph(int nsteps)
{
    if (!is_setup) throw Error(AQ_EINVAL, "setup() has not been called");
    if (nsteps <= 0) return;
    if (vars.host_dirty()) vars.push(stream);
    if (has_control) {
        if (!graph_ready) {
            for (auto& t : tools)
                if (!t->graph_safe() && !t->scope())
                    throw Error(AQ_EUNSUP, "tool \"" + t->name() +
                                               "\" cannot be captured in a "
                                               "graph");
            // the first step runs eagerly so that `once` tools are done
            run_tools(stream);
            nsteps--;
            graph_ready = true;
        }
        step_segments(nsteps);
        check_status();
        return;
    }
    if (!graph_ready) {
        for (auto& t : tools)
            if (!t->graph_safe())
                throw Error(AQ_EUNSUP, "tool \"" + t->name() +
                                           "\" cannot be captured in a "
                                           "graph");
        // the first step runs eagerly so that `once` tools are done
        run_tools(stream);
        nsteps--;
        AQ_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            run_tools(stream);
        } catch (...) {
            hipGraph_t g;
            (void)hipStreamEndCapture(stream, &g);
            throw;
        }
        AQ_HIP(hipStreamEndCapture(stream, &graph));
        AQ_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
        graph_ready = true;
    }
    for (int k = 0; k < nsteps; k++)
        AQ_HIP(hipGraphLaunch(graph_exec, stream));
    check_status();
}

}  // namespace rt
}  // namespace aq
