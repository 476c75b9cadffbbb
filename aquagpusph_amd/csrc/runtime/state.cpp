// XML problem setup: the subset of the reference's State::parse
// (aquagpusph/InputOutput/State.cpp:276-1400) that the presets and cases use.
//
// Order and semantics kept from the reference:
//   * <Include> without `when` (or when="begin") is parsed before the file's
//     own content, when="end" after it; `prefix` prefixes tool names;
//   * each file contributes Settings, Variables, Definitions, Tools,
//     Timing and ParticlesSet, in that order;
//   * tools are added, inserted (at / before / after / before_prefix /
//     after_prefix, comma lists take the first/last match), replaced or
//     removed by glob-matched name (State.cpp:84-100, 636-738, 739-1033);
//   * later variables and definitions override earlier ones.
#include <filesystem>
#include <sstream>

#include "sim.hpp"

namespace fs = std::filesystem;

namespace aq {
namespace rt {

static std::string lower(std::string s)
{
    for (auto& c : s) c = (char)tolower((unsigned char)c);
    return s;
}

static std::string trim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    size_t e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

bool Problem::is_defined(const std::string& n) const
{
    for (auto& d : definitions)
        if (d.name == n) return true;
    return false;
}

// glob with '*' and '?' (State.cpp:52-89)
static bool glob(const char* p, const char* s)
{
    if (*p == '\0') return *s == '\0';
    if (*p == '*') return glob(p + 1, s) || (*s && glob(p, s + 1));
    if (*s && (*p == '?' || *p == *s)) return glob(p + 1, s + 1);
    return false;
}
static bool name_match(const std::string& pat, const std::string& name)
{
    return pat == name || glob(pat.c_str(), name.c_str());
}

// State::findPath: as given, then under the root path, then next to the XML
static std::string find_path(const std::string& p, const Problem& P)
{
    fs::path fp(p);
    if (fs::exists(fp)) return fs::canonical(fp).string();
    std::vector<std::string> tried{ (fs::current_path() / fp).string() };
    if (fp.is_relative()) {
        if (!P.base_path.empty()) {
            fs::path f = fs::path(P.base_path) / fp;
            if (fs::exists(f)) return fs::canonical(f).string();
            tried.push_back(f.string());
        }
        if (!P.xml_stack.empty()) {
            fs::path f = fs::path(P.xml_stack.back()) / fp;
            if (fs::exists(f)) return fs::canonical(f).string();
            tried.push_back(f.string());
        }
    }
    std::string m = "no such file '" + p + "'; tried:";
    for (auto& t : tried) m += " '" + t + "'";
    throw Error(AQ_ENOTFOUND, m);
}

static std::vector<size_t> places_exact(const std::string& list,
                                        const Problem& P,
                                        const std::string& prefix)
{
    std::vector<size_t> out;
    std::stringstream ss(list);
    std::string s;
    while (std::getline(ss, s, ',')) {
        const std::string n = prefix + s;
        for (size_t k = 0; k < P.tools.size(); k++)
            if (P.tools[k]->get("name") == n) out.push_back(k);
    }
    return out;
}

static std::vector<size_t> places_glob(const std::string& name,
                                       const Problem& P,
                                       const std::string& prefix)
{
    std::vector<size_t> out;
    const std::string n = prefix + name;
    for (size_t k = 0; k < P.tools.size(); k++)
        if (name_match(n, P.tools[k]->get("name"))) out.push_back(k);
    return out;
}

// Settings: RootPath, and the compile_flags of the Device elements
// (State.cpp:464-515), which select the arithmetic of the pair kernels.
static void parse_settings(const XmlNode& root, Problem& P)
{
    for (auto* s : root.all("Settings")) {
        for (auto* r : s->all("RootPath"))
            if (r->has("path")) P.base_path = find_path(r->attr("path"), P);
        for (auto* d : s->all("Device"))
            if (d->has("compile_flags"))
                P.compile_flags += " " + d->attr("compile_flags");
    }
}

static void parse_variables(const XmlNode& root, Problem& P)
{
    for (auto* vs : root.all("Variables"))
        for (auto* v : vs->all("Variable")) {
            const std::string name = v->attr("name");
            if (name.empty())
                throw Error(AQ_EPARSE, "Variable without a name");
            if (name.compare(0, 2, "__") == 0)
                throw Error(AQ_EPARSE, "variable prefix '__' is reserved");
            for (auto sfx : { "_x", "_y", "_z", "_w" })
                if (name.size() > 2 &&
                    name.compare(name.size() - 2, 2, sfx) == 0)
                    throw Error(AQ_EPARSE, std::string("variable suffix '") +
                                               sfx + "' is reserved");
            const std::string type = v->attr("type");
            if (type.find('*') == std::string::npos)
                P.variables.emplace_back(name, type, "1", v->attr("value"));
            else
                P.variables.emplace_back(name, type, v->attr("length"), "");
        }
}

static void parse_definitions(const XmlNode& root, Problem& P)
{
    for (auto* ds : root.all("Definitions"))
        for (auto* d : ds->all("Define")) {
            Definition def;
            def.name = d->attr("name");
            if (def.name.empty())
                throw Error(AQ_EPARSE, "Define without a name");
            if (d->has("value")) {
                def.value = d->attr("value");
                const std::string ev = lower(d->attr("evaluate"));
                def.evaluate = ev == "true" || ev == "yes";
            }
            // later definitions override earlier ones
            for (auto it = P.definitions.begin(); it != P.definitions.end();)
                it = (it->name == def.name) ? P.definitions.erase(it) : it + 1;
            P.definitions.push_back(def);
        }
}

static void tool_attr(ToolDesc& t, const XmlNode& e, const std::string& a,
                      const char* def)
{
    if (e.has(a)) t.attrs[a] = e.attr(a);
    else if (def) t.attrs[a] = def;
    else
        throw Error(AQ_EPARSE, "tool \"" + t.get("name") +
                                   "\" requires the missing attribute \"" +
                                   a + "\"");
}

static void parse_tools(const XmlNode& root, Problem& P,
                        const std::string& prefix)
{
    for (auto* ts : root.all("Tools"))
        for (auto* e : ts->all("Tool")) {
            if (!e->has("name")) throw Error(AQ_EPARSE, "Tool without a name");
            if (!e->has("type")) throw Error(AQ_EPARSE, "Tool without a type");
            auto t = std::make_shared<ToolDesc>();
            t->attrs["name"] = prefix + e->attr("name");
            t->attrs["type"] = e->attr("type");
            t->attrs["once"] = e->has("once") ? lower(e->attr("once")) : "false";
            t->xml_dir = P.xml_stack.empty() ? "" : P.xml_stack.back();
            if (e->has("ifdef") && !P.is_defined(e->attr("ifdef"))) continue;
            if (!e->has("ifdef") && e->has("ifndef") &&
                P.is_defined(e->attr("ifndef")))
                continue;
            const std::string action = e->attr("action", "add");
            if (action == "add") {
                P.tools.push_back(t);
            } else if (action == "insert" || action == "try_insert") {
                const bool soft = action == "try_insert";
                std::vector<size_t> places;
                if (e->has("at")) {
                    places.push_back((size_t)std::stoul(e->attr("at")));
                } else if (e->has("before") || e->has("before_prefix")) {
                    const bool pre = !e->has("before");
                    const std::string s =
                        pre ? e->attr("before_prefix") : e->attr("before");
                    const std::string px = pre ? prefix : "";
                    if (s.find(',') != std::string::npos) {
                        auto all = places_exact(s, P, px);
                        if (!all.empty())
                            places.push_back(
                                *std::min_element(all.begin(), all.end()));
                    } else {
                        places = places_glob(s, P, px);
                    }
                } else if (e->has("after") || e->has("after_prefix")) {
                    const bool pre = !e->has("after");
                    const std::string s =
                        pre ? e->attr("after_prefix") : e->attr("after");
                    const std::string px = pre ? prefix : "";
                    if (s.find(',') != std::string::npos) {
                        auto all = places_exact(s, P, px);
                        if (!all.empty())
                            places.push_back(
                                *std::max_element(all.begin(), all.end()) + 1);
                    } else {
                        for (auto k : places_glob(s, P, px))
                            places.push_back(k + 1);
                    }
                } else {
                    throw Error(AQ_EPARSE, "tool \"" + t->get("name") +
                                               "\": insert needs at, before "
                                               "or after");
                }
                if (places.empty()) {
                    if (soft) continue;
                    throw Error(AQ_EPARSE, "tool \"" + t->get("name") +
                                               "\": the referred tool cannot "
                                               "be found");
                }
                for (size_t k = places.size(); k > 0; k--) {
                    const size_t at = std::min(places[k - 1], P.tools.size());
                    P.tools.insert(P.tools.begin() + at, t);
                }
            } else if (action == "remove" || action == "try_remove") {
                auto places = places_glob(e->attr("name"), P, prefix);
                if (places.empty()) {
                    if (action == "try_remove") continue;
                    throw Error(AQ_EPARSE, "cannot remove tool \"" +
                                               t->get("name") + "\": no such "
                                               "tool");
                }
                for (size_t k = places.size(); k > 0; k--)
                    P.tools.erase(P.tools.begin() + places[k - 1]);
                continue;
            } else if (action == "replace" || action == "try_replace") {
                auto places = places_glob(e->attr("name"), P, prefix);
                if (places.empty()) {
                    if (action == "try_replace") continue;
                    throw Error(AQ_EPARSE, "cannot replace tool \"" +
                                               t->get("name") + "\": no such "
                                               "tool");
                }
                for (auto k : places) P.tools[k] = t;
            } else {
                throw Error(AQ_EPARSE, "invalid tool action '" + action + "'");
            }

            // per-type attributes (State.cpp:1034-1200)
            const std::string type = t->get("type");
            if (type == "kernel") {
                tool_attr(*t, *e, "path", nullptr);
                tool_attr(*t, *e, "entry_point", "entry");
                tool_attr(*t, *e, "n", "");
            } else if (type == "copy") {
                tool_attr(*t, *e, "in", nullptr);
                tool_attr(*t, *e, "out", nullptr);
            } else if (type == "set" || type == "set_scalar") {
                tool_attr(*t, *e, "in", nullptr);
                tool_attr(*t, *e, "value", nullptr);
            } else if (type == "reduction") {
                tool_attr(*t, *e, "in", nullptr);
                tool_attr(*t, *e, "out", nullptr);
                tool_attr(*t, *e, "null", nullptr);
                t->text = trim(e->text_content());
                if (t->text.empty())
                    throw Error(AQ_EPARSE, "reduction \"" + t->get("name") +
                                               "\" without an operation");
            } else if (type == "link-list") {
                tool_attr(*t, *e, "in", "r");
                tool_attr(*t, *e, "min", "r_min");
                tool_attr(*t, *e, "max", "r_max");
                tool_attr(*t, *e, "ihoc", "ihoc");
                tool_attr(*t, *e, "icell", "icell");
                tool_attr(*t, *e, "n_cells", "n_cells");
                tool_attr(*t, *e, "perm", "id_unsorted");
                tool_attr(*t, *e, "inv_perm", "id_sorted");
                tool_attr(*t, *e, "recompute_grid", "true");
                tool_attr(*t, *e, "sorter", "radix-sort");
            } else if (type == "radix-sort" || type == "sort") {
                tool_attr(*t, *e, "in", nullptr);
                tool_attr(*t, *e, "perm", nullptr);
                tool_attr(*t, *e, "inv_perm", nullptr);
            } else if (type == "assert" || type == "if" || type == "while") {
                tool_attr(*t, *e, "condition", nullptr);
            } else if (type == "installable") {
                tool_attr(*t, *e, "path", nullptr);
                t->attrs["path"] = find_path(t->get("path"), P);
                // every other attribute reaches aq_tool_create as key/value
                for (auto& kv : e->attrs)
                    if (!t->attrs.count(kv.first)) t->attrs[kv.first] = kv.second;
            } else {
                for (auto& kv : e->attrs)
                    if (!t->attrs.count(kv.first)) t->attrs[kv.first] = kv.second;
            }
            if (type == "kernel") {
                // keep the path lexically resolved against the declaring
                // file; the registry keys on the part under "Scripts/"
                fs::path kp(t->get("path"));
                if (kp.is_relative() && !t->xml_dir.empty())
                    kp = fs::path(t->xml_dir) / kp;
                t->attrs["path"] = kp.lexically_normal().string();
            }
        }
}

// <Reports><Report type="screen|file|performance" .../>: the report tools
// of the same names, run after every tool (State.cpp:1395-1500)
static void parse_reports(const XmlNode& root, Problem& P,
                          const std::string& prefix)
{
    for (auto* rs : root.all("Reports"))
        for (auto* e : rs->all("Report")) {
            if (!e->has("type")) throw Error(AQ_EPARSE, "Report without type");
            auto t = std::make_shared<ToolDesc>();
            for (auto& kv : e->attrs) t->attrs[kv.first] = kv.second;
            t->attrs["name"] = prefix + e->attr("name");
            t->attrs["type"] = "report_" + e->attr("type");
            t->attrs["once"] = "false";
            t->xml_dir = P.xml_stack.empty() ? "" : P.xml_stack.back();
            const std::string ty = e->attr("type");
            if ((ty == "screen" || ty == "file" || ty == "particles") &&
                !e->has("fields"))
                throw Error(AQ_EPARSE, "report \"" + t->get("name") +
                                           "\" without fields");
            if ((ty == "file" || ty == "particles") && !e->has("path"))
                throw Error(AQ_EPARSE, "report \"" + t->get("name") +
                                           "\" without an output path");
            P.reports.push_back(t);
        }
}

static void parse_timing(const XmlNode& root, Problem& P)
{
    for (auto* ts : root.all("Timing"))
        for (auto* o : ts->all("Option")) {
            const std::string name = o->attr("name");
            const std::string type = lower(o->attr("type"));
            const std::string val = o->attr("value");
            if (name == "End") {
                if (type == "time") P.timing.end_t = std::stod(val);
                else if (type == "steps" || type == "iterations")
                    P.timing.end_iter = std::stol(val);
                else if (type == "frames") P.timing.end_frame = std::stol(val);
            } else if (name == "Output") {
                if (type == "fps") P.timing.output_fps = std::stod(val);
                else if (type == "ipf") P.timing.output_ipf = std::stol(val);
            }
        }
}

static void parse_sets(const XmlNode& root, Problem& P)
{
    for (auto* s : root.all("ParticlesSet")) {
        SetDesc d;
        const std::string n = trim(s->attr("n"));
        if (n.empty()) throw Error(AQ_EPARSE, "ParticlesSet without n");
        d.n = (size_t)std::stoull(n);
        for (auto* sc : s->all("Scalar"))
            d.scalars.emplace_back(sc->attr("name"), sc->attr("value"));
        auto fields = [](const std::string& f) {
            std::vector<std::string> v;
            for (auto& c : split_components(f))
                if (!c.empty()) v.push_back(c);
            return v;
        };
        for (auto* l : s->all("Load")) {
            LoadDesc ld{ l->attr("format"), l->attr("file"),
                         fields(l->attr("fields")) };
            if (!ld.file.empty()) {
                fs::path fp(ld.file);
                if (fp.is_relative() && !fs::exists(fp) && !P.xml_stack.empty())
                    fp = fs::path(P.xml_stack.back()) / fp;
                ld.file = fp.string();
            }
            d.loads.push_back(ld);
        }
        for (auto* l : s->all("Save"))
            d.saves.push_back({ l->attr("format"), l->attr("file"),
                                fields(l->attr("fields")) });
        P.sets.push_back(d);
    }
}

void parse_problem(const std::string& path, Problem& P,
                   const std::string& prefix)
{
    const std::string full = find_path(path, P);
    auto root = xml_parse_file(full);
    P.xml_stack.push_back(fs::path(full).parent_path().string());
    auto includes = [&](bool at_end) {
        for (auto* inc : root->all("Include")) {
            const std::string when = inc->attr("when", "begin");
            if ((when == "end") != at_end) continue;
            const std::string file = trim(inc->attr("file"));
            const std::string px = inc->has("prefix") ? inc->attr("prefix")
                                                      : prefix;
            parse_problem(find_path(file, P), P, px);
        }
    };
    includes(false);
    parse_settings(*root, P);
    parse_variables(*root, P);
    parse_definitions(*root, P);
    parse_tools(*root, P, prefix);
    parse_reports(*root, P, prefix);
    parse_timing(*root, P);
    parse_sets(*root, P);
    includes(true);
    P.xml_stack.pop_back();
}

}  // namespace rt
}  // namespace aq
