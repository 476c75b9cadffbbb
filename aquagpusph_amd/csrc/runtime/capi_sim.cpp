// C ABI of the runtime (include/aquagpusph_amd.h, "simulation" section).
#include <cstring>

#include "sim.hpp"

using namespace aq;
using namespace aq::rt;

struct aq_sim {
    std::unique_ptr<Sim> sim;
    std::string scratch;
};

static Sim& SIM(aq_sim* s)
{
    if (!s || !s->sim) throw Error(AQ_EINVAL, "null simulation handle");
    return *s->sim;
}

extern "C" {

AQ_API int aq_sim_create(const char* xml_path, int dims, aq_sim** out)
{
    return guard([&] {
        AQ_REQUIRE(xml_path && out, "aq_sim_create: null argument");
        *out = nullptr;
        auto h = std::make_unique<aq_sim>();
        h->sim = std::make_unique<Sim>(dims);
        h->sim->load(xml_path);
        h->sim->setup();
        *out = h.release();
    });
}

AQ_API void aq_sim_destroy(aq_sim* s)
{
    if (!s) return;
    try {
        if (s->sim) s->sim->sync();
    } catch (...) {
    }
    delete s;
}

AQ_API int aq_sim_step(aq_sim* s, int nsteps)
{
    return guard([&] { SIM(s).step(nsteps); });
}

AQ_API int aq_sim_step_graph(aq_sim* s, int nsteps)
{
    return guard([&] { SIM(s).step_graph(nsteps); });
}

AQ_API int aq_sim_sync(aq_sim* s)
{
    return guard([&] { SIM(s).sync(); });
}

AQ_API void* aq_sim_stream(aq_sim* s)
{
    return (s && s->sim) ? (void*)s->sim->stream : nullptr;
}

AQ_API uint32_t aq_sim_n_particles(aq_sim* s)
{
    return (s && s->sim) ? s->sim->N : 0;
}

AQ_API int aq_sim_var_info(aq_sim* s, const char* name, char* type_buf,
                           size_t type_len, size_t* length, size_t* elem_bytes)
{
    return guard([&] {
        Var& v = SIM(s).vars.at(name ? name : "");
        if (type_buf && type_len) {
            strncpy(type_buf, v.type.canon.c_str(), type_len - 1);
            type_buf[type_len - 1] = '\0';
        }
        if (length) *length = v.is_array() ? v.length : 1;
        if (elem_bytes) *elem_bytes = v.type.elem_bytes();
    });
}

AQ_API int aq_sim_get(aq_sim* s, const char* name, void* host, size_t bytes)
{
    return guard([&] {
        Sim& S = SIM(s);
        Var& v = S.vars.at(name ? name : "");
        AQ_REQUIRE(host, "aq_sim_get: null buffer");
        if (v.is_array()) {
            AQ_REQUIRE(bytes <= v.bytes, "aq_sim_get: too many bytes");
            AQ_HIP(hipStreamSynchronize(S.stream));
            AQ_HIP(hipMemcpy(host, v.dev, bytes, hipMemcpyDeviceToHost));
        } else {
            AQ_REQUIRE(bytes <= v.type.elem_bytes(),
                       "aq_sim_get: too many bytes");
            if (!S.vars.host_dirty()) S.vars.pull(S.stream);
            memcpy(host, S.vars.scalar_host(v), bytes);
        }
    });
}

AQ_API int aq_sim_set(aq_sim* s, const char* name, const void* host,
                      size_t bytes)
{
    return guard([&] {
        Sim& S = SIM(s);
        Var& v = S.vars.at(name ? name : "");
        AQ_REQUIRE(host, "aq_sim_set: null buffer");
        if (v.is_array()) {
            AQ_REQUIRE(bytes <= v.bytes, "aq_sim_set: too many bytes");
            AQ_HIP(hipStreamSynchronize(S.stream));
            AQ_HIP(hipMemcpy(v.dev, host, bytes, hipMemcpyHostToDevice));
        } else {
            AQ_REQUIRE(bytes <= v.type.elem_bytes(),
                       "aq_sim_set: too many bytes");
            if (!S.vars.host_dirty()) S.vars.pull(S.stream);
            memcpy(S.vars.scalar_host(v), host, bytes);
            S.vars.push(S.stream);
        }
    });
}

AQ_API void* aq_sim_device_ptr(aq_sim* s, const char* name)
{
    try {
        Sim& S = SIM(s);
        Var* v = S.vars.get(name ? name : "");
        return v ? S.ptr(v) : nullptr;
    } catch (...) {
        return nullptr;
    }
}

AQ_API int aq_sim_n_tools(aq_sim* s)
{
    return (s && s->sim) ? (int)s->sim->tools.size() : 0;
}

AQ_API const char* aq_sim_tool_name(aq_sim* s, int i)
{
    if (!s || !s->sim || i < 0 || i >= (int)s->sim->tools.size()) return "";
    return s->sim->tools[i]->name().c_str();
}

AQ_API const char* aq_sim_tool_type(aq_sim* s, int i)
{
    if (!s || !s->sim || i < 0 || i >= (int)s->sim->tools.size()) return "";
    return s->sim->tools[i]->type().c_str();
}

AQ_API int aq_sim_definition(aq_sim* s, const char* name, char* buf,
                             size_t len)
{
    return guard([&] {
        const std::string v = SIM(s).def(name ? name : "");
        AQ_REQUIRE(buf && len, "aq_sim_definition: null buffer");
        strncpy(buf, v.c_str(), len - 1);
        buf[len - 1] = '\0';
    });
}

// The AQUAgpusph main loop (main.cpp:160-180): steps until the Timing end
// criteria, writing the particle sets and the state file at each output
// point (at most max_steps steps when max_steps > 0).
AQ_API int aq_sim_run(aq_sim* s, long max_steps, int graph, long* steps)
{
    return guard([&] {
        const long n = SIM(s).run(max_steps, graph != 0);
        if (steps) *steps = n;
    });
}

// Write every set's Save files (original particle order) and the state
// file now (FileManager::save, FileManager.cpp:146-155).
AQ_API int aq_sim_save(aq_sim* s)
{
    return guard([&] { SIM(s).save_now(); });
}

// Per-tool GPU time: runs `nsteps` eager steps with HIP events around every
// tool and writes the mean milliseconds of tool i to ms[i] (the
// report_performance counterpart, aquagpusph/CalcServer/Reports/
// Performance.cpp).
AQ_API int aq_sim_profile_tools(aq_sim* s, int nsteps, double* ms)
{
    return guard([&] {
        Sim& S = SIM(s);
        AQ_REQUIRE(ms, "aq_sim_profile_tools: null output");
        const size_t nt = S.tools.size();
        // a pair of events around every execution of a tool (tools inside
        // while loops run several times per step), from a growing pool
        std::vector<hipEvent_t> ev;
        std::vector<size_t> which;
        std::vector<double> acc(nt, 0.0);
        FusionPlan& F = S.fusion;
        double walk_acc = 0.0;
        int walk_n = 0;
        if (F.active) {
            if (!F.ev0) AQ_HIP(hipEventCreate(&F.ev0));
            if (!F.ev1) AQ_HIP(hipEventCreate(&F.ev1));
        }
        if (S.vars.host_dirty()) S.vars.push(S.stream);
        // every tool timed on its own: no element chain beside the walk
        struct SideOff {
            bool& b;
            SideOff(bool& x) : b(x) { b = true; }
            ~SideOff() { b = false; }
        } side_off_(S.side_off);
        for (int k = 0; k < nsteps; k++) {
            F.timing = F.active;
            size_t used = 0;
            which.clear();
            S.run_tools(S.stream, [&](size_t t, bool after) {
                if (used == ev.size()) {
                    hipEvent_t e;
                    AQ_HIP(hipEventCreate(&e));
                    ev.push_back(e);
                }
                AQ_HIP(hipEventRecord(ev[used++], S.stream));
                if (!after) which.push_back(t);
            });
            F.timing = false;
            if (used) AQ_HIP(hipEventSynchronize(ev[used - 1]));
            if (F.active) {
                float x = 0.f;
                if (hipEventElapsedTime(&x, F.ev0, F.ev1) == hipSuccess) {
                    walk_acc += x;
                    walk_n++;
                }
            }
            for (size_t e = 0; e < which.size(); e++) {
                float x = 0.f;
                AQ_HIP(hipEventElapsedTime(&x, ev[2 * e], ev[2 * e + 1]));
                acc[which[e]] += x;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (size_t t = 0; t < nt; t++) ms[t] = nsteps ? acc[t] / nsteps : 0.0;
        F.walk_ms = walk_n ? walk_acc / walk_n : -1.0;
        S.check_status();
    });
}

AQ_API int aq_sim_autotune(aq_sim* s, int nsteps, char* report, size_t len)
{
    return guard([&] {
        const std::string r = SIM(s).autotune(nsteps);
        if (report && len) {
            strncpy(report, r.c_str(), len - 1);
            report[len - 1] = '\0';
        }
    });
}

AQ_API int aq_sim_walk_ms(aq_sim* s, double* ms)
{
    return guard(
        [&] {
            AQ_REQUIRE(ms, "aq_sim_walk_ms: null output");
            *ms = SIM(s).fusion.walk_ms;
        },
        false);
}

// Parse a case without touching the GPU and describe what the runtime would
// build: one line per tool "name\ttype\tkey" (key = kernel registry key or
// the tool's main variable), then "#var name\ttype\tlength\tvalue" lines.
AQ_API int aq_xml_resolve(const char* xml_path, int dims, char* out,
                          size_t len, size_t* needed)
{
    return guard([&] {
        AQ_REQUIRE(xml_path, "aq_xml_resolve: null path");
        AQ_REQUIRE(dims == 2 || dims == 3, "dims must be 2 or 3");
        Problem P = load_problem(xml_path, dims);
        std::string txt;
        for (auto& t : P.tools) {
            std::string key;
            if (t->get("type") == "kernel")
                key = kernel_key(t->get("path"), t->get("entry_point"));
            else
                key = t->get("in");
            txt += t->get("name") + "\t" + t->get("type") + "\t" + key + "\n";
        }
        for (auto& v : P.variables)
            txt += "#var " + std::get<0>(v) + "\t" + std::get<1>(v) + "\t" +
                   std::get<2>(v) + "\t" + std::get<3>(v) + "\n";
        for (auto& d : P.definitions)
            txt += "#def " + d.name + "\t" + d.value + "\t" +
                   (d.evaluate ? "1" : "0") + "\n";
        size_t n = 0;
        for (auto& s : P.sets) n += s.n;
        txt += "#N " + std::to_string(n) + "\n";
        if (needed) *needed = txt.size() + 1;
        if (out && len) {
            strncpy(out, txt.c_str(), len - 1);
            out[len - 1] = '\0';
        }
    }, false);
}

AQ_API int aq_sim_n_registered_kernels(void)
{
    return (int)registered_kernels().size();
}

AQ_API const char* aq_sim_registered_kernel(int i)
{
    static std::vector<std::string> k = registered_kernels();
    return (i >= 0 && i < (int)k.size()) ? k[i].c_str() : "";
}

}  // extern "C"
