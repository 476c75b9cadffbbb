// Output of a run, as the AQUAgpusph executable does it: the time manager
// that decides when to stop and when to write (aquagpusph/main.cpp:160-180,
// InputOutput/TimeManager.cpp:103-150), the particle savers that write each
// set in the original particle order (InputOutput/Particles.cpp:250-310,
// ASCII.cpp:240-306, CalcServer::getUnsortedMem), and the state file
// AQUAgpusph.save.N.xml from which a run resumes (State.cpp:195-207,
// 1517-1916).
#include <cfloat>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "sim.hpp"

namespace aq {

void launch_unsort(const uint32_t* id, const void* in, void* out, uint32_t N,
                   uint32_t words, hipStream_t s);

namespace rt {

namespace fs = std::filesystem;

static std::string lower(std::string s)
{
    for (auto& c : s) c = (char)tolower((unsigned char)c);
    return s;
}

static void replace_all(std::string& s, const std::string& a,
                        const std::string& b)
{
    for (size_t p = s.find(a); p != std::string::npos;
         p = s.find(a, p + b.size()))
        s.replace(p, a.size(), b);
}

// newFilePath (AuxiliarMethods.cpp:215-256): "%d" or "{index}" take the
// first index from `i` on whose file does not exist yet, zero-padded to
// `digits`; a name without an index must not exist yet.
static std::string new_file_path(const std::string& pattern, unsigned& i,
                                 unsigned digits)
{
    std::string p = pattern;
    replace_all(p, "%d", "{index}");
    if (p.find("{index}") == std::string::npos) {
        if (fs::exists(p))
            throw Error(AQ_EINVAL, "the output file '" + p +
                                       "' exists already");
        return p;
    }
    while (true) {
        char b[32];
        snprintf(b, sizeof(b), "%0*u", (int)digits, i);
        std::string f = p;
        replace_all(f, "{index}", b);
        if (!fs::exists(f)) return f;
        i++;
    }
}

static std::string xml_escape(const std::string& s)
{
    std::string o;
    for (char c : s) {
        switch (c) {
        case '&': o += "&amp;"; break;
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        case '"': o += "&quot;"; break;
        default: o += c;
        }
    }
    return o;
}

// Rows of particles [off, off + n) in the original order: the fields are
// unsorted through `id` on the device first.  Fields are separated by ',',
// components by ' ', floats written with 7 significant digits (the
// reference's std::numeric_limits<float>::digits10 + 1), integers whole.
static void write_ascii(Sim& S, const std::string& path, size_t off, size_t n,
                        const std::vector<std::string>& fields, double t)
{
    Var& id = S.vars.at("id");
    std::vector<Var*> vs;
    size_t big = 0;
    for (auto& name : fields) {
        Var& v = S.vars.at(name);
        if (!v.is_array() || v.length < S.N)
            throw Error(AQ_EINVAL, "cannot save '" + name + "': not an "
                                                            "array of N");
        if (v.type.base == 'd')
            throw Error(AQ_EUNSUP, "cannot save the double field '" + name +
                                       "' (ASCII saves 32-bit fields)");
        vs.push_back(&v);
        big = std::max(big, (size_t)S.N * v.type.elem_bytes());
    }
    void* tmp = nullptr;
    AQ_HIP(hipMalloc(&tmp, big));
    std::vector<std::vector<uint32_t>> data(vs.size());
    try {
        for (size_t k = 0; k < vs.size(); k++) {
            const Var& v = *vs[k];
            launch_unsort((const uint32_t*)id.dev, v.dev, tmp, S.N,
                          (uint32_t)v.type.stored, S.stream);
            AQ_HIP(hipGetLastError());
            data[k].resize(n * v.type.stored);
            AQ_HIP(hipMemcpyAsync(data[k].data(),
                                  (char*)tmp + off * v.type.elem_bytes(),
                                  n * v.type.elem_bytes(),
                                  hipMemcpyDeviceToHost, S.stream));
            AQ_HIP(hipStreamSynchronize(S.stream));
        }
    } catch (...) {
        (void)hipFree(tmp);
        throw;
    }
    AQ_HIP(hipFree(tmp));
    std::ofstream f(path);
    if (!f) throw Error(AQ_ENOTFOUND, "cannot write '" + path + "'");
    char b[64];
    snprintf(b, sizeof(b), "%g", t);
    // The header spans 18 lines, 17 comment lines and a blank one, as the
    // reference's (ASCII::print_header): readers such as the reference's
    // check scripts skip exactly that many lines before the data.
    std::string names;
    for (size_t k = 0; k < fields.size(); k++)
        names += fields[k] + (k + 1 < fields.size() ? ", " : "");
    const std::string rule(57, '#');
    f << rule << "\n#\n#    aquagpusph_amd particle file (ASCII)\n#\n"
      << "#    t = " << b << " s\n#    particles: " << n
      << "\n#    fields: " << names
      << "\n#\n#    one particle per line, in the original order;\n"
      << "#    fields separated by ',', vector components by ' '\n"
      << "#    (the layout of the FastASCII loader)\n#\n#\n#\n#\n#\n"
      << rule << "\n\n";
    std::string line;
    for (size_t i = 0; i < n; i++) {
        line.clear();
        for (size_t k = 0; k < vs.size(); k++) {
            const Type& ty = vs[k]->type;
            for (int c = 0; c < ty.ncomp; c++) {
                const uint32_t w = data[k][i * ty.stored + c];
                if (ty.base == 'f') {
                    float x;
                    memcpy(&x, &w, 4);
                    snprintf(b, sizeof(b), "%.9g", (double)x);  // round trip
                } else if (ty.base == 'i')
                    snprintf(b, sizeof(b), "%d", (int32_t)w);
                else
                    snprintf(b, sizeof(b), "%u", w);
                line += b;
                if (c + 1 < ty.ncomp) line += ' ';
            }
            if (k + 1 < vs.size()) line += ',';
        }
        line += '\n';
        f << line;
    }
    if (!f) throw Error(AQ_ENOTFOUND, "failure writing '" + path + "'");
}

// Scalar and array-element values as the XML value text: components
// separated by ", ", floats to 9 significant digits (exact round trip).
static std::string value_text(const Type& ty, const uint8_t* p)
{
    std::string s;
    char b[64];
    for (int c = 0; c < ty.ncomp; c++) {
        if (ty.base == 'd') {
            double x;
            memcpy(&x, p + 8 * c, 8);
            snprintf(b, sizeof(b), "%.17g", x);
        } else if (ty.base == 'f') {
            float x;
            memcpy(&x, p + 4 * c, 4);
            snprintf(b, sizeof(b), "%.9g", (double)x);
        } else if (ty.base == 'i') {
            int32_t x;
            memcpy(&x, p + 4 * c, 4);
            snprintf(b, sizeof(b), "%d", x);
        } else {
            uint32_t x;
            memcpy(&x, p + 4 * c, 4);
            snprintf(b, sizeof(b), "%u", x);
        }
        s += b;
        if (c + 1 < ty.ncomp) s += ", ";
    }
    return s;
}

// The state file: the resolved problem (settings, variables with their
// current values, definitions, the tools in pipeline order, timing) and
// every set loading the particle file just written, so that a run resumes
// from it (State::write, State.cpp:1517-1916).
static void write_state(Sim& S, const std::string& path)
{
    const Problem& P = S.problem;
    std::ostringstream o;
    o << "<?xml version=\"1.0\" ?>\n<sphInput>\n";
    o << "    <Settings>\n";
    if (!P.base_path.empty())
        o << "        <RootPath path=\"" << xml_escape(P.base_path) << "\" />\n";
    if (!P.compile_flags.empty())
        o << "        <Device platform=\"0\" device=\"0\" compile_flags=\""
          << xml_escape(P.compile_flags) << "\" />\n";
    o << "    </Settings>\n";
    // variables (skipping the "__" internals, as the reference)
    o << "    <Variables>\n";
    for (auto& vp : S.vars.all()) {
        const Var& v = *vp;
        if (v.name.compare(0, 2, "__") == 0) continue;
        o << "        <Variable name=\"" << xml_escape(v.name) << "\" type=\""
          << xml_escape(v.type.canon) << "\"";
        if (v.is_array())
            o << " length=\"" << v.length << "\" />\n";
        else
            o << " value=\""
              << value_text(v.type, (const uint8_t*)S.vars.scalar_host(
                                        const_cast<Var&>(v)))
              << "\" />\n";
    }
    o << "    </Variables>\n";
    o << "    <Definitions>\n";
    for (auto& d : P.definitions)
        o << "        <Define name=\"" << xml_escape(d.name) << "\" value=\""
          << xml_escape(d.value) << "\" evaluate=\""
          << (d.evaluate ? "true" : "false") << "\" />\n";
    o << "    </Definitions>\n";
    // tools in pipeline order; placement attributes are dropped (every
    // tool is added in turn) and script paths made absolute
    static const std::set<std::string> placement = {
        "action", "at",    "before", "after", "before_prefix", "after_prefix",
        "ifdef",  "ifndef"
    };
    o << "    <Tools>\n";
    for (auto& td : P.tools) {
        o << "        <Tool";
        for (auto& kv : td->attrs) {
            if (placement.count(kv.first)) continue;
            std::string val = kv.second;
            if (kv.first == "path" && !val.empty() && fs::path(val).is_relative())
                val = (fs::path(td->xml_dir) / val).lexically_normal().string();
            o << " " << kv.first << "=\"" << xml_escape(val) << "\"";
        }
        if (!td->text.empty())
            o << ">" << xml_escape(td->text) << "</Tool>\n";
        else
            o << " />\n";
    }
    o << "    </Tools>\n";
    const TimingDesc& T = P.timing;
    o << "    <Timing>\n";
    if (T.end_t >= 0)
        o << "        <Option name=\"End\" type=\"Time\" value=\"" << T.end_t
          << "\" />\n";
    if (T.end_iter >= 0)
        o << "        <Option name=\"End\" type=\"Steps\" value=\"" << T.end_iter
          << "\" />\n";
    if (T.end_frame >= 0)
        o << "        <Option name=\"End\" type=\"Frames\" value=\""
          << T.end_frame << "\" />\n";
    if (T.output_fps >= 0)
        o << "        <Option name=\"Output\" type=\"FPS\" value=\""
          << T.output_fps << "\" />\n";
    if (T.output_ipf >= 0)
        o << "        <Option name=\"Output\" type=\"IPF\" value=\""
          << T.output_ipf << "\" />\n";
    o << "    </Timing>\n";
    for (size_t k = 0; k < P.sets.size(); k++) {
        const SetDesc& sd = P.sets[k];
        o << "    <ParticlesSet n=\"" << sd.n << "\">\n";
        for (auto& sc : sd.scalars) {
            Var& v = S.vars.at(sc.first);
            std::vector<uint8_t> w(v.type.elem_bytes());
            AQ_HIP(hipMemcpy(w.data(), (char*)v.dev + k * w.size(), w.size(),
                             hipMemcpyDeviceToHost));
            o << "        <Scalar name=\"" << xml_escape(sc.first)
              << "\" value=\"" << value_text(v.type, w.data()) << "\" />\n";
        }
        for (size_t j = 0; j < sd.saves.size(); j++) {
            const LoadDesc& s = sd.saves[j];
            std::string flist;
            for (size_t f = 0; f < s.fields.size(); f++)
                flist += s.fields[f] + (f + 1 < s.fields.size() ? "," : "");
            if (j == 0 && k < S.out.last_file.size() &&
                !S.out.last_file[k].empty())
                o << "        <Load format=\"" << xml_escape(s.format)
                  << "\" file=\""
                  << xml_escape(fs::absolute(S.out.last_file[k]).string())
                  << "\" fields=\"" << xml_escape(flist) << "\" />\n";
            o << "        <Save format=\"" << xml_escape(s.format)
              << "\" file=\"" << xml_escape(s.file) << "\" fields=\""
              << xml_escape(flist) << "\" />\n";
        }
        o << "    </ParticlesSet>\n";
    }
    o << "</sphInput>\n";
    std::ofstream f(path);
    if (!f) throw Error(AQ_ENOTFOUND, "cannot write '" + path + "'");
    f << o.str();
}

static void sync_scalars(Sim& S)
{
    if (S.vars.host_dirty()) S.vars.push(S.stream);
    S.vars.pull(S.stream);
}

static void start_output(Sim& S)
{
    if (S.out.started) return;
    S.out.next_index.assign(S.problem.sets.size(), 0u);
    S.out.last_file.assign(S.problem.sets.size(), std::string());
    // the first state file name not taken yet (State.cpp:195-207)
    for (unsigned i = 0;; i++) {
        const std::string f = "AQUAgpusph.save." + std::to_string(i) + ".xml";
        if (!fs::exists(f)) {
            S.out.state_file = f;
            break;
        }
    }
    sync_scalars(S);
    S.out.time = S.vars.get_host(S.vars.at("t"))[0];
    S.out.step = (unsigned long)S.vars.get_host(S.vars.at("iter"))[0];
    S.out.started = true;
}

void Sim::save_now()
{
    if (!is_setup) throw Error(AQ_EINVAL, "setup() has not been called");
    start_output(*this);
    sync_scalars(*this);
    const double t = vars.get_host(vars.at("t"))[0];
    size_t off = 0;
    for (size_t k = 0; k < problem.sets.size(); k++) {
        const SetDesc& sd = problem.sets[k];
        for (size_t j = 0; j < sd.saves.size(); j++) {
            const LoadDesc& s = sd.saves[j];
            const std::string fmt = lower(s.format);
            if (fmt != "ascii" && fmt != "fastascii")
                throw Error(AQ_EUNSUP, "particles output format '" + s.format +
                                           "' is not supported (ASCII)");
            // a name without an index takes ".{index}.dat" (ASCII::create)
            std::string pattern = s.file;
            if (pattern.find("{index}") == std::string::npos &&
                pattern.find("%d") == std::string::npos)
                pattern += ".{index}.dat";
            const std::string path =
                new_file_path(pattern, out.next_index[k], 5);
            write_ascii(*this, path, off, sd.n, s.fields, t);
            out.next_index[k]++;
            if (j == 0) out.last_file[k] = path;
        }
        off += sd.n;
    }
    write_state(*this, out.state_file);
}

long Sim::run(long max_steps, bool graph)
{
    if (!is_setup) throw Error(AQ_EINVAL, "setup() has not been called");
    start_output(*this);
    const TimingDesc& T = problem.timing;
    auto get = [&](const char* n) { return vars.get_host(vars.at(n))[0]; };
    auto must_stop = [&]() {
        return get("t") >= get("end_t") || get("iter") >= get("end_iter") ||
               get("frame") >= get("end_frame");
    };
    auto next_frame = [&]() {
        vars.set_host(vars.at("frame"), { get("frame") + 1.0 });
    };
    // TimeManager::mustPrintOutput
    auto must_print = [&]() {
        const double t = get("t");
        const unsigned long it = (unsigned long)get("iter");
        if (t < 0) return false;
        if ((T.output_fps >= 0 || T.output_ipf >= 0) && get("frame") == 0 &&
            it == 1) {
            out.time = t;
            out.step = it;
            next_frame();
            return true;
        }
        if (T.output_fps > 0 && t - out.time >= 1.0 / T.output_fps) {
            out.time += 1.0 / T.output_fps;
            out.step = it;
            next_frame();
            return true;
        }
        if (T.output_ipf > 0 && (long)(it - out.step) >= T.output_ipf) {
            out.time = t;
            out.step = it;
            next_frame();
            return true;
        }
        return must_stop();
    };
    long ran = 0;
    sync_scalars(*this);
    while (!must_stop()) {
        bool print = false;
        while (!print) {
            if (graph)
                step_graph(1);
            else
                step(1);
            ran++;
            sync_scalars(*this);
            print = must_print();
            if (max_steps > 0 && ran >= max_steps) break;
        }
        if (print) save_now();
        if (max_steps > 0 && ran >= max_steps) break;
        sync_scalars(*this);
    }
    return ran;
}

}  // namespace rt
}  // namespace aq
