// Built-in tool types (the reference's CalcServer/*.cpp tools).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dlfcn.h>
#include <sstream>
#include <filesystem>
#include <limits>

#include "aquagpusph_amd_tool.h"
#include "sim.hpp"

namespace aq {

void linklist_all(int dims, const float* r, uint32_t N, float* r_min,
                  float* r_max, const float* support, const float* h,
                  uint32_t* n_cells, uint32_t* icell, uint32_t* id_unsorted,
                  uint32_t* id_sorted, uint32_t* ihoc, uint32_t cap,
                  uint32_t* status, void* ws, hipStream_t s,
                  const float* per_min, const float* per_L);
size_t linklist_workspace(uint32_t N);
size_t cell_list_workspace(uint32_t N);
void cell_list(const uint32_t* icell, uint32_t N, uint32_t* cell_end,
               uint32_t* heads, uint32_t* n_heads, const uint32_t* status,
               void* ws, const uint32_t* perm, const int* imove,
               hipStream_t s, uint32_t* ocells, uint32_t* wcells, int dims,
               const uint32_t* ihoc, const uint32_t* n_cells, size_t ocap,
               uint32_t* bcells, uint32_t* bcount, uint32_t* bbase,
               uint32_t* bsums, uint32_t nbmax);
size_t cell_blocks_bound(int dims, size_t cap);
size_t scan_workspace(uint32_t m);
void radix_sort(uint32_t* keys, uint32_t* perm, uint32_t* inv, uint32_t n,
                unsigned key_bits, void* workspace, hipStream_t s);
size_t radix_sort_workspace(uint32_t n);
unsigned radix_key_bits(uint64_t max_val);
size_t reduce_workspace(uint32_t n, int ncomp);
void reduce(int op, int type, int nc, const void* in, uint32_t n, void* out,
            const void* ident_dev, void* ws, hipStream_t s);

namespace rt {

void fill_launch(void* dst, const uint32_t* pattern_dev, int stored, size_t n,
                 hipStream_t s);

static std::string trim(const std::string& s)
{
    size_t b = s.find_first_not_of(" \t\r\n");
    size_t e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

static std::string squeeze(const std::string& s)
{
    std::string o;
    for (char c : s)
        if (!isspace((unsigned char)c)) o += c;
    return o;
}

// device buffer owned by a tool (freed with it)
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    void alloc(size_t bytes)
    {
        if (bytes <= n) return;
        if (p) AQ_HIP(hipFree(p));
        AQ_HIP(hipMalloc(&p, bytes));
        n = bytes;
    }
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// "VEC_ZERO", "-VEC_INFINITY", "INFINITY", "0.f", constant expressions
static std::vector<double> macro_value(const std::string& v0, const Type& t,
                                       int dims, const Variables& vars)
{
    const std::string v = squeeze(v0);
    const double inf = std::numeric_limits<double>::infinity();
    const int n = t.ncomp;
    std::vector<double> out(n, 0.0);
    auto vec = [&](double x, bool all) {
        for (int k = 0; k < n; k++) out[k] = (k < dims || all) ? x : 0.0;
    };
    if (v == "VEC_ZERO" || v == "MAT_ZERO") vec(0.0, true);
    else if (v == "VEC_ONE") vec(1.0, false);
    else if (v == "VEC_ALL_ONE") vec(1.0, true);
    else if (v == "VEC_INFINITY") vec(inf, false);
    else if (v == "VEC_ALL_INFINITY") vec(inf, true);
    else if (v == "VEC_NEG_INFINITY" || v == "-VEC_INFINITY") {
        vec(-inf, false);
        if (n == 4 && dims == 3) out[3] = -0.0;
    } else if (v == "VEC_ALL_NEG_INFINITY" || v == "-VEC_ALL_INFINITY")
        vec(-inf, true);
    else if (v == "INFINITY") vec(inf, true);
    else if (v == "-INFINITY") vec(-inf, true);
    else {
        // one value fills every component; a list fills them in order
        auto comps = split_components(v0);
        if (comps.size() == 1) {
            const double x = vars.eval(comps[0]);
            for (int k = 0; k < n; k++) out[k] = x;
        } else {
            for (int k = 0; k < n && k < (int)comps.size(); k++)
                out[k] = vars.eval(comps[k]);
        }
    }
    return out;
}

static std::vector<uint32_t> words(const std::vector<double>& v,
                                   const Type& t)
{
    std::vector<uint32_t> w(t.elem_bytes() / 4, 0u);
    for (int k = 0; k < t.ncomp && k < (int)v.size(); k++) {
        if (t.base == 'd') {
            std::memcpy(&w[2 * k], &v[k], 8);
        } else if (t.base == 'f') {
            float f = (float)v[k];
            std::memcpy(&w[k], &f, 4);
        } else if (t.base == 'i') {
            int32_t x = (int32_t)v[k];
            std::memcpy(&w[k], &x, 4);
        } else {
            w[k] = (uint32_t)(int64_t)v[k];
        }
    }
    return w;
}

// ---- dummy ---------------------------------------------------------------------
class DummyTool : public Tool {
  public:
    using Tool::Tool;
    void execute(Sim&, hipStream_t) override {}
    bool batch_transparent() const override { return true; }
};

// ---- reports: Reports/Screen.cpp, TabFile.cpp, Performance.cpp -----------------
// The reported variables are copied into pinned memory on the stream and a
// host function, enqueued behind the copy, formats them; both are stream
// operations, so a report inside a captured step becomes a copy node and a
// host node of the graph and never stalls the stream.
struct ReportField {
    Var* v;
    size_t off, bytes;
};

// Variable::asString: scalars as numbers, vectors as "(x,y,...)", arrays as
// their elements in brackets.
static std::string report_value(const Type& ty, const uint8_t* p, size_t n)
{
    auto one = [&](const uint8_t* q) {
        std::string s;
        char b[64];
        for (int c = 0; c < ty.ncomp; c++) {
            if (ty.base == 'd') {
                double x;
                memcpy(&x, q + 8 * c, 8);
                snprintf(b, sizeof(b), "%.17g", x);
            } else if (ty.base == 'f') {
                float x;
                memcpy(&x, q + 4 * c, 4);
                snprintf(b, sizeof(b), "%.9g", (double)x);
            } else if (ty.base == 'i') {
                int32_t x;
                memcpy(&x, q + 4 * c, 4);
                snprintf(b, sizeof(b), "%d", x);
            } else {
                uint32_t x;
                memcpy(&x, q + 4 * c, 4);
                snprintf(b, sizeof(b), "%u", x);
            }
            s += b;
            if (c + 1 < ty.ncomp) s += ",";
        }
        return ty.ncomp > 1 ? "(" + s + ")" : s;
    };
    if (!ty.array) return one(p);
    std::string s = "[";
    for (size_t k = 0; k < n; k++) {
        s += one(p + k * ty.elem_bytes());
        if (k + 1 < n) s += ",";
    }
    return s + "]";
}

class ReportTool : public Tool {
  public:
    // kind 0: screen, 1: file, 2: performance
    ReportTool(const ToolDesc& d, int kind)
        : Tool(d.get("name"), d.get("type"), d.get("once") == "true"),
          d_(d), kind_(kind)
    {
    }
    ~ReportTool() override
    {
        if (host_) (void)hipHostFree(host_);
    }
    void setup(Sim& S) override
    {
        if (host_) return;  // the same tool placed twice
        S_ = &S;
        std::string fields = d_.get("fields");
        if (kind_ == 2) fields = "t,end_t,iter,end_iter,frame,end_frame";
        if (kind_ != 2 && trim(fields).empty())
            throw Error(AQ_EPARSE, "report \"" + name_ + "\" without fields");
        // Report::processFields: ';' breaks lines, ',' or ' ' separates
        size_t off = 0;
        std::stringstream lines(fields);
        std::string line;
        while (std::getline(lines, line, ';')) {
            for (auto& c : line)
                if (c == ' ') c = ',';
            std::stringstream ls(line);
            std::string n;
            size_t in_line = 0;
            while (std::getline(ls, n, ',')) {
                n = trim(n);
                if (n.empty()) continue;
                Var* v = S.vars.get(n);
                if (!v)
                    throw Error(AQ_ENOTFOUND, "report \"" + name_ +
                                                  "\" asks for the undeclared "
                                                  "variable \"" + n + "\"");
                const size_t b = v->is_array() ? v->bytes
                                               : v->type.elem_bytes();
                if (b > (1u << 16))
                    throw Error(AQ_EUNSUP, "report \"" + name_ + "\": \"" +
                                               n + "\" is too large to "
                                               "report (64 KiB at most)");
                fields_.push_back({ v, off, b });
                off += (b + 15) & ~(size_t)15;
                in_line++;
            }
            per_line_.push_back(in_line);
        }
        bytes_ = off ? off : 16;
        AQ_HIP(hipHostMalloc((void**)&host_, bytes_, hipHostMallocDefault));
        if (kind_ == 1) {
            path_ = d_.get("path");
            if (path_.empty())
                throw Error(AQ_EPARSE, "report \"" + name_ +
                                           "\" of type file without a path");
            file_.reset(fopen(path_.c_str(), "w"));
            if (!file_)
                throw Error(AQ_ENOTFOUND, "cannot write '" + path_ + "'");
            std::string h = "# ";
            for (auto& f : fields_) h += f.v->name + " ";
            h += "\n";
            fputs(h.c_str(), file_.get());
            fflush(file_.get());
        } else if (kind_ == 2) {
            path_ = d_.get("path");
            for (auto& v : S.vars.all())
                if (v->is_array()) mem_ += v->bytes;
        }
    }
    void execute(Sim& S, hipStream_t s) override
    {
        for (auto& f : fields_)
            AQ_HIP(hipMemcpyAsync(host_ + f.off, S.ptr(f.v), f.bytes,
                                  hipMemcpyDeviceToHost, s));
        AQ_HIP(hipLaunchHostFunc(s, &ReportTool::cb, this));
    }

  private:
    static void cb(void* p) { ((ReportTool*)p)->print(); }
    std::string text(const char* sep_in_line) const
    {
        std::string out;
        size_t k = 0;
        for (size_t l = 0; l < per_line_.size(); l++) {
            for (size_t j = 0; j < per_line_[l]; j++, k++) {
                const auto& f = fields_[k];
                const size_t n = f.v->is_array() ? f.v->length : 1;
                out += report_value(f.v->type, host_ + f.off, n);
                if (j + 1 < per_line_[l]) out += sep_in_line;
            }
            if (l + 1 < per_line_.size()) out += kind_ == 1 ? " " : "\n";
        }
        return out;
    }
    double scalar(size_t k) const
    {
        const auto& f = fields_[k];
        const uint8_t* q = host_ + f.off;
        if (f.v->type.base == 'f') {
            float x;
            memcpy(&x, q, 4);
            return x;
        }
        if (f.v->type.base == 'd') {
            double x;
            memcpy(&x, q, 8);
            return x;
        }
        uint32_t x;
        memcpy(&x, q, 4);
        return x;
    }
    void print()
    {
        if (kind_ == 0) {
            // Screen: the tool name, then each line of name=value pairs
            std::string out = name_ + ":\n";
            size_t k = 0;
            for (size_t l = 0; l < per_line_.size(); l++) {
                for (size_t j = 0; j < per_line_[l]; j++, k++) {
                    const auto& f = fields_[k];
                    const size_t n = f.v->is_array() ? f.v->length : 1;
                    out += f.v->name + "=" +
                           report_value(f.v->type, host_ + f.off, n);
                    out += (j + 1 < per_line_[l]) ? " " : "\n";
                }
            }
            fputs(out.c_str(), stdout);
            fflush(stdout);
        } else if (kind_ == 1) {
            // TabFile: one line per execution, space separated
            const std::string out = text(" ") + "\n";
            fputs(out.c_str(), file_.get());
            fflush(file_.get());
        } else {
            // Performance: memory, wall time since the last execution,
            // progress by time / steps / frames and the time to arrival
            const auto now = std::chrono::steady_clock::now();
            const double el = last_set_ ? std::chrono::duration<double>(
                                              now - last_).count()
                                        : 0.0;
            last_ = now;
            last_set_ = true;
            double prog = 0.0;
            for (int c = 0; c < 3; c++) {
                const double a = scalar(2 * c), b = scalar(2 * c + 1);
                if (b > 0) prog = std::max(prog, a / b);
            }
            total_ += el;
            const double eta = prog > 0 ? total_ * (1.0 / prog - 1.0) : 0.0;
            char b[256];
            snprintf(b, sizeof(b),
                     "%s:\nMemory=%zuMB\nElapsed=%.6gs\nPercentage=%.2f"
                     "   ETA=%.6gs\n",
                     name_.c_str(), mem_ >> 20, el, 100.0 * prog, eta);
            fputs(b, stdout);
            fflush(stdout);
            if (!path_.empty()) {
                FILE* f = fopen(path_.c_str(), "w");
                if (f) {
                    fprintf(f,
                            "{\n    \"progress\": %.9g,\n    \"elapsed\": "
                            "%.9g,\n    \"snapshot\": {\"step\": %.0f, "
                            "\"samples\": []}\n}\n",
                            prog, el * 1e9, scalar(2));
                    fclose(f);
                }
            }
        }
    }

    struct FileClose {
        void operator()(FILE* f) const
        {
            if (f) fclose(f);
        }
    };
    ToolDesc d_;
    int kind_;
    Sim* S_ = nullptr;
    std::vector<ReportField> fields_;
    std::vector<size_t> per_line_;
    uint8_t* host_ = nullptr;
    size_t bytes_ = 0;
    std::string path_;
    std::unique_ptr<FILE, FileClose> file_;
    size_t mem_ = 0;
    std::chrono::steady_clock::time_point last_;
    bool last_set_ = false;
    double total_ = 0.0;
};

// ---- kernel: registry of HIP implementations --------------------------------
class KernelTool : public Tool {
  public:
    KernelTool(const ToolDesc& d)
        : Tool(d.get("name"), "kernel", d.get("once") == "true"),
          path_(d.get("path")), entry_(d.get("entry_point", "entry"))
    {
    }
    void setup(Sim& S) override
    {
        key_ = kernel_key(path_, entry_);
        spec_ = find_kernel(key_);
        if (!spec_)
            throw Error(AQ_EUNSUP,
                        "tool \"" + name_ + "\": no HIP implementation for "
                        "script \"" + path_ + "\" entry point \"" + entry_ +
                        "\" (registry key \"" + key_ + "\"); "
                        "provide one through an installable tool");
        bound_.clear();
        for (auto& a : spec_->args) {
            Var* v = S.vars.get(a);
            if (!v)
                throw Error(AQ_ENOTFOUND,
                            "tool \"" + name_ + "\" (" + key_ +
                                ") requires the undeclared variable \"" + a +
                                "\"");
            bound_.push_back(v);
        }
    }
    void execute(Sim& S, hipStream_t s) override
    {
        if (S.tail.anchor == this) {
            S.run_tail(s);
            return;
        }
        if (S.head.anchor == this) {
            S.run_head(s);
            return;
        }
        if (S.sortc.anchor == this) {
            S.run_sort_chain(s);
            return;
        }
        if (S.tail.absorbed.count(this) || S.head.absorbed.count(this) ||
            S.sortc.absorbed.count(this))
            return;
        if (S.fusion.active) {
            if (S.fusion.anchor == this) {
                S.run_fused(s);
                return;
            }
            if (S.fusion.absorbed.count(this)) return;
            if (S.fusion.side_launched) {
                // launched beside the walk by the anchor
                if (S.fusion.side_set.count(this)) return;
                if (S.fusion.join == this) S.join_side(s);
            }
        }
        spec_->launch(S, bound_.data(), s);
    }
    void launch_kernel(Sim& S, hipStream_t s) override
    {
        spec_->launch(S, bound_.data(), s);
    }
    std::string key() const override { return key_; }

  private:
    std::string path_, entry_, key_;
    const KernelSpec* spec_ = nullptr;
    std::vector<Var*> bound_;
};

// ---- link-list: LinkList.cpp ------------------------------------------------
class LinkListTool : public Tool {
  public:
    LinkListTool(const ToolDesc& d)
        : Tool(d.get("name"), "link-list", d.get("once") == "true"), d_(d)
    {
        const std::string so = d.get("sorter", "radix-sort");
        if (so != "radix-sort")
            throw Error(AQ_EUNSUP, "link-list \"" + name_ + "\": sorter '" +
                                       so + "' is not supported (only the "
                                       "stable radix-sort)");
        if (d.get("recompute_grid", "true") != "true")
            throw Error(AQ_EUNSUP, "link-list \"" + name_ +
                                       "\": recompute_grid=false is not "
                                       "supported");
    }
    void setup(Sim& S) override
    {
        in_ = &S.vars.at(d_.get("in"));
        mn_ = &S.vars.at(d_.get("min"));
        mx_ = &S.vars.at(d_.get("max"));
        ihoc_ = &S.vars.at(d_.get("ihoc"));
        icell_ = &S.vars.at(d_.get("icell"));
        nc_ = &S.vars.at(d_.get("n_cells"));
        perm_ = &S.vars.at(d_.get("perm"));
        inv_ = &S.vars.at(d_.get("inv_perm"));
        sup_ = &S.vars.at("support");
        h_ = &S.vars.at("h");
        if (in_->type.base != 'f' || !in_->is_array())
            throw Error(AQ_EINVAL, "link-list input must be a vec array");
        // a periodic box (Presets/cfd/periodic.xml): periodic_min and
        // periodic_L, the axes with L > 0 are periodic
        per_min_ = S.vars.get("periodic_min");
        per_L_ = S.vars.get("periodic_L");
        if (per_min_ || per_L_) {
            for (Var* v : { per_min_, per_L_ })
                if (!v || v->is_array() || v->type.base != 'f' ||
                    v->type.ncomp < S.dims)
                    throw Error(AQ_EINVAL,
                                "link-list \"" + name_ +
                                    "\": periodic_min and periodic_L must "
                                    "both be vec variables");
        }
        ws_.alloc(linklist_workspace(S.N));
    }
    // ihoc capacity: the grid of the domain box when domain_min/domain_max
    // exist (Domain.cl keeps every particle inside it), else twice the
    // extent of the current positions in every direction
    size_t capacity(Sim& S)
    {
        if (cap_) return cap_;
        const float cl = (float)S.vars.get_host(*sup_)[0] *
                         (float)S.vars.get_host(*h_)[0];
        if (!(cl > 0.f))
            throw Error(AQ_EINVAL, "zero cell length (support * h)");
        std::vector<double> lo(3, 0.0), hi(3, 0.0);
        Var* dmn = S.vars.get("domain_min");
        Var* dmx = S.vars.get("domain_max");
        bool dom = false;
        if (dmn && dmx && !dmn->is_array() && !dmx->is_array()) {
            auto a = S.vars.get_host(*dmn), b = S.vars.get_host(*dmx);
            dom = true;
            for (int k = 0; k < S.dims; k++) {
                lo[k] = a[k];
                hi[k] = b[k];
                if (!(hi[k] > lo[k])) dom = false;
            }
        }
        Var& r = S.vars.at("r");
        if (!dom && r.length) {
            std::vector<float> h(r.length * r.type.stored);
            AQ_HIP(hipMemcpy(h.data(), r.dev, r.bytes,
                             hipMemcpyDeviceToHost));
            for (int k = 0; k < S.dims; k++) {
                lo[k] = INFINITY;
                hi[k] = -INFINITY;
            }
            for (size_t i = 0; i < r.length; i++)
                for (int k = 0; k < S.dims; k++) {
                    const double x = h[i * r.type.stored + k];
                    if (std::isfinite(x)) {
                        lo[k] = std::min(lo[k], x);
                        hi[k] = std::max(hi[k], x);
                    }
                }
            for (int k = 0; k < S.dims; k++) {
                const double e = std::max(hi[k] - lo[k], (double)cl);
                lo[k] -= 0.5 * e;
                hi[k] += 0.5 * e;
            }
        }
        double cells = 1.0;
        for (int k = 0; k < S.dims; k++)
            cells *= std::floor((hi[k] - lo[k]) / cl) + 8.0;
        if (cells > 4.0e9)
            throw Error(AQ_ENOMEM, "link-list grid too large");
        cap_ = (size_t)cells;
        return cap_;
    }
    void execute(Sim& S, hipStream_t s) override
    {
        // the capacity is sized on the first (always eager) execution, from
        // the particles as loaded or set by the caller
        if (!cap_) {
            AQ_HIP(hipStreamSynchronize(s));
            const size_t c = capacity(S);
            if (ihoc_->length < c) S.vars.allocate(*ihoc_, c);
        }
        linklist_all(S.dims, (const float*)in_->dev, S.N,
                     (float*)S.ptr(mn_), (float*)S.ptr(mx_),
                     (const float*)S.ptr(sup_), (const float*)S.ptr(h_),
                     (uint32_t*)S.ptr(nc_), (uint32_t*)icell_->dev,
                     (uint32_t*)perm_->dev, (uint32_t*)inv_->dev,
                     (uint32_t*)ihoc_->dev, (uint32_t)ihoc_->length,
                     S.d_status, ws_.p, s,
                     per_min_ ? (const float*)S.ptr(per_min_) : nullptr,
                     per_L_ ? (const float*)S.ptr(per_L_) : nullptr);
        // the non-empty cell list of the cell-parallel walks
        CellList& c = S.clist;
        if (c.enabled && c.icell == icell_ && c.ihoc == ihoc_) {
            if (c.cap < ihoc_->length) {
                // first step (eager) or a grown ihoc: never under capture
                if (c.cell_end) AQ_HIP(hipFree(c.cell_end));
                c.cell_end = nullptr;
                AQ_HIP(hipMalloc(&c.cell_end, ihoc_->length * 4));
                c.cap = ihoc_->length;
                // cell triplets: at most one per non-empty cell, and the
                // owner lists one more per AQ_OWNER_CHUNK non-fluid
                // particles of the cells split into runs
                const size_t ncap = std::min<size_t>(S.N, c.cap);
                for (uint32_t** l : { &c.ocells, &c.wcells, &c.nlist,
                                      &c.bcells }) {
                    if (*l) AQ_HIP(hipFree(*l));
                    *l = nullptr;
                    const bool own = l == &c.ocells || l == &c.wcells;
                    AQ_HIP(hipMalloc(
                        l, (own ? aq_owner_list_cap(ncap, S.N) : ncap) * 12));
                }
                // the blocked order's per-block counters (zero between
                // steps) and bases
                for (uint32_t** l : { &c.bcount, &c.bbase, &c.bsums }) {
                    if (*l) AQ_HIP(hipFree(*l));
                    *l = nullptr;
                }
                c.nbmax = (uint32_t)cell_blocks_bound(S.dims, c.cap);
                AQ_HIP(hipMalloc(&c.bcount, (size_t)c.nbmax * 4));
                AQ_HIP(hipMemset(c.bcount, 0, (size_t)c.nbmax * 4));
                AQ_HIP(hipMalloc(&c.bbase, (size_t)c.nbmax * 4));
                AQ_HIP(hipMalloc(&c.bsums, scan_workspace(c.nbmax) + 256));
            }
            if (!c.heads) {
                AQ_HIP(hipMalloc(&c.heads, (size_t)S.N * 4));
                // [0] count, then the walks' hand-out counters (one
                // 128-byte line each, from word 32), the range record and
                // the owner-cell count (word AQ_CL_NOCELLS)
                AQ_HIP(hipMalloc(&c.n_heads, AQ_CL_WORDS * 4));
                AQ_HIP(hipMalloc(&c.ws, cell_list_workspace(S.N)));
                AQ_HIP(hipMemset(c.n_heads, 0, AQ_CL_WORDS * 4));
            }
            Var* im = S.vars.get("imove");
            const bool kinds = im && im->is_array() && im->length >= S.N &&
                               im->type.base == 'i';
            cell_list((const uint32_t*)icell_->dev, S.N, c.cell_end, c.heads,
                      c.n_heads, S.d_status, c.ws,
                      kinds ? (const uint32_t*)perm_->dev : nullptr,
                      kinds ? (const int*)im->dev : nullptr, s,
                      kinds ? c.ocells : nullptr, c.wcells, S.dims,
                      (const uint32_t*)ihoc_->dev,
                      (const uint32_t*)S.ptr(nc_),
                      std::min<size_t>(S.N, c.cap),
                      c.blocked ? c.bcells : nullptr, c.bcount, c.bbase,
                      c.bsums, c.nbmax);
            c.kinds = kinds;
            c.filled = true;
        }
    }

  private:
    ToolDesc d_;
    Var *in_, *mn_, *mx_, *ihoc_, *icell_, *nc_, *perm_, *inv_, *sup_, *h_;
    Var *per_min_ = nullptr, *per_L_ = nullptr;
    DevBuf ws_;
    size_t cap_ = 0;
};

// ---- radix-sort: RadixSort.cpp ----------------------------------------------
// type="sort" (Sort.cpp, a bitonic sort) runs the same stable radix sort:
// its contract is sorted keys with consistent permutations (perm[i] is the
// original index of the i-th sorted key, inv_perm its inverse), which the
// stable order fulfils; among equal keys the bitonic order is not specified.
class RadixSortTool : public Tool {
  public:
    RadixSortTool(const ToolDesc& d)
        : Tool(d.get("name"), d.get("type"), d.get("once") == "true"), d_(d)
    {
    }
    void setup(Sim& S) override
    {
        in_ = &S.vars.at(d_.get("in"));
        perm_ = &S.vars.at(d_.get("perm"));
        inv_ = &S.vars.at(d_.get("inv_perm"));
        for (Var* v : { in_, perm_, inv_ })
            if (!v->is_array() || v->type.base != 'u' || v->type.ncomp != 1)
                throw Error(AQ_EINVAL, type_ + " \"" + name_ +
                                           "\": variables must be "
                                           "unsigned int arrays");
        if (perm_->length != in_->length || inv_->length != in_->length)
            throw Error(AQ_EINVAL, type_ + " \"" + name_ +
                                       "\": lengths mismatch");
        ws_.alloc(radix_sort_workspace((uint32_t)in_->length));
    }
    void execute(Sim&, hipStream_t s) override
    {
        radix_sort((uint32_t*)in_->dev, (uint32_t*)perm_->dev,
                   (uint32_t*)inv_->dev, (uint32_t)in_->length,
                   radix_key_bits(0xFFFFFFFFull >> 1), ws_.p, s);
    }

  private:
    ToolDesc d_;
    Var *in_, *perm_, *inv_;
    DevBuf ws_;
};

// ---- reduction: Reduction.cpp -----------------------------------------------
class ReductionTool : public Tool {
  public:
    ReductionTool(const ToolDesc& d)
        : Tool(d.get("name"), "reduction", d.get("once") == "true"), d_(d)
    {
        // the operation is a C statement; its closing ';' may be left out
        std::string op = squeeze(d.text);
        if (!op.empty() && op.back() != ';') op += ';';
        if (op == "c=min(a,b);" || op == "c=(a<b)?a:b;" ||
            op == "c=(a<=b)?a:b;" || op == "c=(b<a)?b:a;")
            op_ = AQ_RED_MIN;
        else if (op == "c=max(a,b);" || op == "c=(a>b)?a:b;" ||
                 op == "c=(a>=b)?a:b;" || op == "c=(b>a)?b:a;")
            op_ = AQ_RED_MAX;
        else if (op == "c=a+b;" || op == "c=b+a;")
            op_ = AQ_RED_SUM;
        else
            throw Error(AQ_EUNSUP, "reduction \"" + name_ +
                                       "\": unsupported operation \"" +
                                       d.text + "\" (min, max or sum)");
    }
    void setup(Sim& S) override
    {
        in_ = &S.vars.at(d_.get("in"));
        out_ = &S.vars.at(d_.get("out"));
        if (!in_->is_array() || out_->is_array())
            throw Error(AQ_EINVAL, "reduction \"" + name_ +
                                       "\": in must be an array, out a "
                                       "scalar");
        // a float array may be reduced into a double of the same shape
        // (summed in double)
        const bool widen = in_->type.base == 'f' && out_->type.base == 'd';
        if ((in_->type.base != out_->type.base && !widen) ||
            in_->type.stored != out_->type.stored)
            throw Error(AQ_EINVAL, "reduction \"" + name_ +
                                       "\": in and out types differ");
        nc_ = in_->type.stored;
        auto id = words(macro_value(d_.get("null"), out_->type, S.dims, S.vars),
                        out_->type);
        ident_.alloc(id.size() * 4);
        AQ_HIP(hipMemcpy(ident_.p, id.data(), id.size() * 4,
                         hipMemcpyHostToDevice));
        ws_.alloc(reduce_workspace((uint32_t)in_->length, nc_));
    }
    void execute(Sim& S, hipStream_t s) override
    {
        const char ib = in_->type.base, ob = out_->type.base;
        const int type = (ib == 'f' && ob == 'd') ? AQ_T_FLOAT_F64
                         : ib == 'd'              ? AQ_T_DOUBLE
                         : ib == 'f'              ? AQ_T_FLOAT
                         : ib == 'i'              ? AQ_T_INT
                                                  : AQ_T_UINT;
        reduce(op_, type, nc_, in_->dev, (uint32_t)in_->length, S.ptr(out_),
               ident_.p, ws_.p, s);
    }

  private:
    ToolDesc d_;
    int op_ = 0, nc_ = 1;
    Var *in_, *out_;
    DevBuf ident_, ws_;
};

// ---- set: Set.cpp -----------------------------------------------------------
class SetTool : public Tool {
  public:
    SetTool(const ToolDesc& d)
        : Tool(d.get("name"), "set", d.get("once") == "true"), d_(d)
    {
    }
    void setup(Sim& S) override
    {
        v_ = &S.vars.at(d_.get("in"));
        if (!v_->is_array())
            throw Error(AQ_EINVAL, "set \"" + name_ + "\": \"" + v_->name +
                                       "\" is not an array");
        auto w = words(macro_value(d_.get("value"), v_->type, S.dims, S.vars),
                       v_->type);
        pat_.alloc(w.size() * 4);
        AQ_HIP(hipMemcpy(pat_.p, w.data(), w.size() * 4,
                         hipMemcpyHostToDevice));
    }
    void execute(Sim&, hipStream_t s) override
    {
        if (batch_member) return;  // filled by the run's first set tool
        if (batch_next.empty()) {
            fill_launch(v_->dev, (const uint32_t*)pat_.p,
                        (int)(v_->type.elem_bytes() / 4), v_->length, s);
            return;
        }
        FillBatch b;
        add_to(b);
        for (Tool* t : batch_next) static_cast<SetTool*>(t)->add_to(b);
        fill_batch(b, s);
    }
    int batch_kind() const override { return 1; }
    void batch_io(std::vector<const void*>&,
                  std::vector<const void*>& wr) const override
    {
        wr.push_back(v_->dev);
    }

  private:
    void add_to(FillBatch& b) const
    {
        b.dst[b.cnt] = v_->dev;
        b.pat[b.cnt] = (const uint32_t*)pat_.p;
        b.stored[b.cnt] = (int)(v_->type.elem_bytes() / 4);
        b.n[b.cnt] = v_->length;
        b.cnt++;
    }
    ToolDesc d_;
    Var* v_;
    DevBuf pat_;
};

// ---- set_scalar / assert: SetScalar.cpp, Assert.cpp -------------------------
static int vm_depth(const std::vector<VmOp>& ops)
{
    int d = 0, mx = 0;
    for (auto& o : ops) {
        switch (o.code) {
        case VM_PUSH: case VM_LOADF: case VM_LOADI: case VM_LOADU:
        case VM_LOADD: d++; break;
        case VM_STOREF: case VM_STOREI: case VM_STOREU: case VM_STORED:
        case VM_ASSERT: d--; break;
        case VM_NEG: case VM_NOT: case VM_F1: case VM_END: break;
        case VM_SEL: d -= 2; break;
        default: d--; break;
        }
        mx = std::max(mx, d);
    }
    return mx;
}

class VmTool : public Tool {
  public:
    using Tool::Tool;
    void upload(const VmProgram& p)
    {
        if (vm_depth(p.ops) > 30)
            throw Error(AQ_EUNSUP, "expression of tool \"" + name_ +
                                       "\" is too deep");
        prog_.alloc(p.ops.size() * sizeof(VmOp));
        AQ_HIP(hipMemcpy(prog_.p, p.ops.data(), p.ops.size() * sizeof(VmOp),
                         hipMemcpyHostToDevice));
        ops_ = p.ops;
    }
    void execute(Sim& S, hipStream_t s) override
    {
        if (batch_member) return;  // run by the first tool of its run
        const DevBuf& prog = batch_next.empty() ? prog_ : batch_prog_;
        vm_launch((const VmOp*)prog.p, S.vars.dev_arena(), S.d_status,
                  S.d_fail_id, s);
    }
    // the run's programs one after the other (each leaves the stack empty,
    // so the next starts as it would alone)
    void batch_setup(Sim&) override
    {
        std::vector<VmOp> all;
        auto append = [&](const std::vector<VmOp>& o) {
            for (const VmOp& op : o)
                if (op.code != VM_END) all.push_back(op);
        };
        append(ops_);
        for (Tool* t : batch_next) append(static_cast<VmTool*>(t)->ops_);
        all.push_back(VmOp{ VM_END, 0, 0.0 });
        batch_prog_.alloc(all.size() * sizeof(VmOp));
        AQ_HIP(hipMemcpy(batch_prog_.p, all.data(), all.size() * sizeof(VmOp),
                         hipMemcpyHostToDevice));
    }

  protected:
    DevBuf prog_, batch_prog_;
    std::vector<VmOp> ops_;
};

class SetScalarTool : public VmTool {
  public:
    SetScalarTool(const ToolDesc& d)
        : VmTool(d.get("name"), "set_scalar", d.get("once") == "true"), d_(d)
    {
    }
    void setup(Sim& S) override
    {
        Var& v = S.vars.at(d_.get("in"));
        if (v.is_array())
            throw Error(AQ_EINVAL, "set_scalar \"" + name_ + "\": \"" +
                                       v.name + "\" is an array");
        auto comps = split_components(d_.get("value"));
        auto slots = [&](const std::string& n, SlotRef& r) {
            return S.vars.slot(n, r);
        };
        // every component is evaluated from the values before the tool
        // runs, then stored (a swap "v_y,v_x" exchanges them)
        VmProgram p;
        const int nc = std::min(v.type.ncomp, (int)comps.size());
        for (int k = 0; k < nc; k++) p.compile(parse_expr(comps[k]), slots);
        for (int k = nc - 1; k >= 0; k--)
            p.store(SlotRef{ v.slot + (uint32_t)v.type.comp_bytes() * k,
                             v.type.base });
        p.end();
        upload(p);
    }
    int batch_kind() const override { return 3; }

  private:
    ToolDesc d_;
};

class AssertTool : public VmTool {
  public:
    AssertTool(const ToolDesc& d)
        : VmTool(d.get("name"), "assert", d.get("once") == "true"), d_(d)
    {
    }
    void setup(Sim& S) override
    {
        VmProgram p;
        p.compile(parse_expr(d_.get("condition")),
                  [&](const std::string& n, SlotRef& r) {
                      return S.vars.slot(n, r);
                  });
        p.assert_true(S.assert_id(name_ + " (" + d_.get("condition") + ")"));
        p.end();
        upload(p);
    }

  private:
    ToolDesc d_;
};

// ---- while / if / end: Conditional.cpp ------------------------------------------
// The condition runs on the device VM into a hidden int scalar, which is
// read back into pinned memory (a stream synchronisation, the counterpart of
// clWaitForEvents in Conditional::next_tool) for the host to pick the next
// tool.
class CondTool : public ControlTool {
  public:
    CondTool(const ToolDesc& d)
        : ControlTool(d.get("name"), d.get("type"), d.get("once") == "true"),
          d_(d)
    {
    }
    ~CondTool() override
    {
        if (host_) (void)hipHostFree(host_);
    }
    int scope() const override { return 1; }
    void setup(Sim& S) override
    {
        if (prog_.p) return;  // the same tool placed twice
        static int count = 0;
        Var* res = S.vars.reg("__condition " + std::to_string(count++) + " " +
                                  name_,
                              "int", "", "0");
        slot_ = res->slot;
        VmProgram p;
        p.compile(parse_expr(d_.get("condition")),
                  [&](const std::string& n, SlotRef& r) {
                      return S.vars.slot(n, r);
                  });
        p.store(SlotRef{ slot_, 'i' });
        p.end();
        if (vm_depth(p.ops) > 30)
            throw Error(AQ_EUNSUP, "condition of tool \"" + name_ +
                                       "\" is too deep");
        prog_.alloc(p.ops.size() * sizeof(VmOp));
        AQ_HIP(hipMemcpy(prog_.p, p.ops.data(), p.ops.size() * sizeof(VmOp),
                         hipMemcpyHostToDevice));
        AQ_HIP(hipHostMalloc((void**)&host_, sizeof(int32_t),
                             hipHostMallocDefault));
    }
    bool eval(Sim& S, hipStream_t s) override
    {
        vm_launch((const VmOp*)prog_.p, S.vars.dev_arena(), S.d_status,
                  S.d_fail_id, s);
        AQ_HIP(hipMemcpyAsync(host_, S.vars.dev_arena() + slot_,
                              sizeof(int32_t), hipMemcpyDeviceToHost, s));
        AQ_HIP(hipStreamSynchronize(s));
        return *host_ != 0;
    }

  private:
    ToolDesc d_;
    DevBuf prog_;
    uint32_t slot_ = 0;
    int32_t* host_ = nullptr;
};

class EndTool : public ControlTool {
  public:
    EndTool(const ToolDesc& d)
        : ControlTool(d.get("name"), "end", d.get("once") == "true")
    {
    }
    int scope() const override { return -1; }
    bool eval(Sim&, hipStream_t) override { return true; }
};

// ---- copy: Copy.cpp -----------------------------------------------------------
class CopyTool : public Tool {
  public:
    CopyTool(const ToolDesc& d)
        : Tool(d.get("name"), "copy", d.get("once") == "true"), d_(d)
    {
    }
    void setup(Sim& S) override
    {
        in_ = &S.vars.at(d_.get("in"));
        out_ = &S.vars.at(d_.get("out"));
        if (!in_->is_array() || !out_->is_array())
            throw Error(AQ_EINVAL, "copy \"" + name_ + "\": arrays expected");
        if (in_->type.elem_bytes() != out_->type.elem_bytes())
            throw Error(AQ_EINVAL, "copy \"" + name_ + "\": types differ");
    }
    void execute(Sim&, hipStream_t s) override
    {
        if (batch_member) return;  // copied by the run's first copy tool
        if (batch_next.empty()) {
            const size_t b = std::min(in_->bytes, out_->bytes);
            if (b)
                AQ_HIP(hipMemcpyAsync(out_->dev, in_->dev, b,
                                      hipMemcpyDeviceToDevice, s));
            return;
        }
        CopyBatch b;
        add_to(b);
        for (Tool* t : batch_next) static_cast<CopyTool*>(t)->add_to(b);
        copy_batch(b, s);
    }
    // whole words between distinct arrays (the batched copy copies words,
    // and a run's copies must not read what another writes)
    int batch_kind() const override
    {
        return (std::min(in_->bytes, out_->bytes) % 4 == 0 &&
                in_->dev != out_->dev)
                   ? 2
                   : 0;
    }
    void batch_io(std::vector<const void*>& rd,
                  std::vector<const void*>& wr) const override
    {
        rd.push_back(in_->dev);
        wr.push_back(out_->dev);
    }

  private:
    void add_to(CopyBatch& b) const
    {
        b.src[b.cnt] = in_->dev;
        b.dst[b.cnt] = out_->dev;
        b.words[b.cnt] = std::min(in_->bytes, out_->bytes) / 4;
        b.cnt++;
    }
    ToolDesc d_;
    Var *in_, *out_;
};

// ---- installable: out-of-tree tools through a pure C ABI --------------------
// The plugin exports the aq_tool_* functions of aquagpusph_amd_tool.h
// (replacing the C++ create_object() of tests/ExternalTool/tool.cpp:27-32).
typedef aq_tool_ctx aq_tool_ctx_t;

static void* plugin_var(void* sim, const char* name, size_t* bytes)
{
    Sim* S = (Sim*)sim;
    Var* v = S->vars.get(name);
    if (!v) return nullptr;
    if (bytes) *bytes = v->is_array() ? v->bytes : v->type.elem_bytes();
    return S->ptr(v);
}

class InstallableTool : public Tool {
  public:
    InstallableTool(const ToolDesc& d)
        : Tool(d.get("name"), "installable", d.get("once") == "true"), d_(d)
    {
    }
    ~InstallableTool() override
    {
        if (obj_ && destroy_) destroy_(obj_);
        if (h_) dlclose(h_);
    }
    void setup(Sim& S) override
    {
        h_ = dlopen(d_.get("path").c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h_)
            throw Error(AQ_ENOTFOUND, "installable \"" + name_ + "\": " +
                                          dlerror());
        auto create = (int (*)(const char*, const char* const*, void**))dlsym(
            h_, "aq_tool_create");
        setup_ = (int (*)(void*, const aq_tool_ctx_t*))dlsym(h_,
                                                            "aq_tool_setup");
        enqueue_ = (int (*)(void*, void*))dlsym(h_, "aq_tool_enqueue");
        destroy_ = (void (*)(void*))dlsym(h_, "aq_tool_destroy");
        if (!create || !setup_ || !enqueue_ || !destroy_)
            throw Error(AQ_ENOTFOUND, "installable \"" + name_ +
                                          "\": missing aq_tool_* symbols");
        std::vector<std::string> kv;
        for (auto& a : d_.attrs) {
            kv.push_back(a.first);
            kv.push_back(a.second);
        }
        std::vector<const char*> ckv;
        for (auto& s : kv) ckv.push_back(s.c_str());
        ckv.push_back(nullptr);
        if (create(name_.c_str(), ckv.data(), &obj_) != 0 || !obj_)
            throw Error(AQ_EINVAL, "installable \"" + name_ +
                                       "\": aq_tool_create failed");
        ctx_.sim = &S;
        ctx_.var = plugin_var;
        if (setup_(obj_, &ctx_) != 0)
            throw Error(AQ_EINVAL, "installable \"" + name_ +
                                       "\": aq_tool_setup failed");
    }
    void execute(Sim&, hipStream_t s) override
    {
        if (enqueue_(obj_, (void*)s) != 0)
            throw Error(AQ_EINVAL, "installable \"" + name_ +
                                       "\": aq_tool_enqueue failed");
    }

  private:
    ToolDesc d_;
    void* h_ = nullptr;
    void* obj_ = nullptr;
    int (*setup_)(void*, const aq_tool_ctx_t*) = nullptr;
    int (*enqueue_)(void*, void*) = nullptr;
    void (*destroy_)(void*) = nullptr;
    aq_tool_ctx_t ctx_{};
};

std::shared_ptr<Tool> make_tool(Sim& S, const ToolDesc& d)
{
    (void)S;
    const std::string t = d.get("type");
    const bool once = d.get("once") == "true";
    if (t == "kernel") return std::make_shared<KernelTool>(d);
    if (t == "link-list") return std::make_shared<LinkListTool>(d);
    if (t == "radix-sort" || t == "sort")
        return std::make_shared<RadixSortTool>(d);
    if (t == "reduction") return std::make_shared<ReductionTool>(d);
    if (t == "set") return std::make_shared<SetTool>(d);
    if (t == "set_scalar") return std::make_shared<SetScalarTool>(d);
    if (t == "assert") return std::make_shared<AssertTool>(d);
    if (t == "while" || t == "if") return std::make_shared<CondTool>(d);
    if (t == "end") return std::make_shared<EndTool>(d);
    if (t == "copy") return std::make_shared<CopyTool>(d);
    if (t == "installable") return std::make_shared<InstallableTool>(d);
    if (t == "dummy") return std::make_shared<DummyTool>(d.get("name"), t, once);
    if (t == "report_screen") return std::make_shared<ReportTool>(d, 0);
    if (t == "report_file") return std::make_shared<ReportTool>(d, 1);
    if (t == "report_performance") return std::make_shared<ReportTool>(d, 2);
    throw Error(AQ_EUNSUP, "tool \"" + d.get("name") + "\": type '" + t +
                               "' is not supported by this build");
}

}  // namespace rt
}  // namespace aq
