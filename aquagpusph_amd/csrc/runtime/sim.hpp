// Host runtime: the counterpart of the reference's CalcServer
// (aquagpusph/CalcServer/CalcServer.{hpp,cpp}), its variable registry
// (aquagpusph/Variable.{hpp,cpp}), tool plugin API (CalcServer/Tool.hpp)
// and XML problem setup (InputOutput/State.cpp, ProblemSetup.hpp).
//
// MI355X-first differences from the reference (same XML surface):
//   * every scalar variable lives in one device "scalar arena" with a host
//     mirror; kernels read scalars through device pointers and set_scalar /
//     assert / reduction / link-list tools write them on the device, so a
//     step never waits for the host and can be replayed from a HIP graph;
//   * array variables are device buffers in the reference layout (one
//     buffer per field, vec = float2/float4);
//   * kernel tools resolve (script path, entry point) to hand-written HIP
//     kernels through a registry instead of compiling OpenCL at run time.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../aq_internal.hpp"
#include "expr.hpp"
#include "xml.hpp"

namespace aq {
namespace rt {

// ---- types ------------------------------------------------------------------
struct Type {
    char base = 'f';    // 'f' float, 'i' int32, 'u' uint32, 'd' double
    int ncomp = 1;      // logical components
    int stored = 1;     // stored components (vec3 is padded to 4)
    bool array = false;
    std::string canon;  // canonical reference type name
    size_t comp_bytes() const { return base == 'd' ? 8 : 4; }
    size_t elem_bytes() const { return comp_bytes() * (size_t)stored; }
};
// Resolve a reference type string ("vec*", "size_t", "svec4", "float", ...)
// with the aliasing of aquagpusph/Variable.cpp:1547-1625 (addr_bits = 32).
Type parse_type(const std::string& t, int dims);

struct Var {
    std::string name;
    Type type;
    uint32_t slot = 0;  // scalars: byte offset in the scalar arena
    std::string length_expr;
    size_t length = 0;  // arrays: number of elements
    void* dev = nullptr;
    size_t bytes = 0;
    bool is_array() const { return type.array; }
};

class Variables {
  public:
    static constexpr size_t ARENA_SLOTS = 4096;
    static constexpr size_t SLOT_BYTES = 16;

    explicit Variables(int dims);
    ~Variables();
    Variables(const Variables&) = delete;

    // Variables::registerVariable: replaces an existing variable of the same
    // name, evaluates the value/length with the variables registered so far.
    Var* reg(const std::string& name, const std::string& type,
             const std::string& length, const std::string& value);
    Var* get(const std::string& name);
    Var& at(const std::string& name);
    const std::vector<std::unique_ptr<Var>>& all() const { return vars_; }

    // expressions over the host mirror of the scalars ("g_x", "n_cells_w")
    bool symbol(const std::string& name, double& v) const;
    bool slot(const std::string& name, SlotRef& r) const;
    double eval(const std::string& expr) const;

    // host mirror access (component-wise, as double)
    void set_host(Var& v, const std::vector<double>& vals);
    std::vector<double> get_host(const Var& v) const;
    void* scalar_dev(const Var& v) const { return dev_arena_ + v.slot; }
    void* scalar_host(const Var& v) { return host_arena_.data() + v.slot; }
    std::vector<uint8_t>& host_arena() { return host_arena_; }

    // whole-arena transfers (synchronous w.r.t. the stream)
    void push(hipStream_t s);
    void pull(hipStream_t s);
    bool host_dirty() const { return dirty_; }

    // (re)allocate an array variable's device buffer, zero-filled
    void allocate(Var& v, size_t length);
    uint8_t* dev_arena() const { return dev_arena_; }

  private:
    int dims_;
    std::vector<std::unique_ptr<Var>> vars_;
    std::map<std::string, Var*> by_name_;
    std::vector<uint8_t> host_arena_;
    uint8_t* dev_arena_ = nullptr;
    uint32_t next_slot_ = 0;
    std::vector<uint32_t> free_slots_;
    bool dirty_ = true;
};

class Sim;

// ---- tools (Tool.hpp plugin API) -------------------------------------------
class Tool {
  public:
    Tool(const std::string& name, const std::string& type, bool once)
        : name_(name), type_(type), once_(once)
    {
    }
    virtual ~Tool() {}
    virtual void setup(Sim& S) { (void)S; }
    // enqueue the tool's work on the stream; must not synchronise
    virtual void execute(Sim& S, hipStream_t s) = 0;
    // false when execute() needs the host between launches (eager only)
    virtual bool graph_safe() const { return true; }
    // registry key of kernel tools ("" for other tools)
    virtual std::string key() const { return ""; }
    // enqueue the tool's own work on `s`, outside the fusion plan (the
    // kernel tools launched on the side stream)
    virtual void launch_kernel(Sim& S, hipStream_t s) { execute(S, s); }
    // +1 for the scope openers (while, if), -1 for end (Tool.hpp
    // scope_modifier); such tools are run by the pipeline, not execute()
    virtual int scope() const { return 0; }
    // Runs of consecutive tools of one batch kind (1 set, 2 copy, 3 set_scalar;
// 0 none)
    // run as one launch: the first tool's execute() does the whole run
    // (batch_next: the tools after it), the others' execute() nothing.
    // Runs never cross a control tool, where the pipeline may jump, and
    // hold no `once` tool.
    virtual int batch_kind() const { return 0; }
    // tools that do nothing (dummy markers) neither end a run nor join it
    virtual bool batch_transparent() const { return false; }
    // called on a run's first tool once the run is formed
    virtual void batch_setup(Sim& S) { (void)S; }
    // the arrays a batchable tool reads and writes: a run holds no tool
    // that reads or writes what an earlier tool of the run writes, or
    // writes what an earlier one reads
    virtual void batch_io(std::vector<const void*>& rd,
                          std::vector<const void*>& wr) const
    {
        (void)rd;
        (void)wr;
    }
    std::vector<Tool*> batch_next;
    bool batch_member = false;
    const std::string& name() const { return name_; }
    const std::string& type() const { return type_; }
    bool once() const { return once_; }
    bool done = false;

  protected:
    std::string name_, type_;
    bool once_;
};

// While / If condition tools (aquagpusph/CalcServer/Conditional.cpp)
class ControlTool : public Tool {
  public:
    using Tool::Tool;
    void execute(Sim&, hipStream_t) override {}
    bool graph_safe() const override { return false; }
    // evaluate the condition (reads it back: synchronises the stream)
    virtual bool eval(Sim& S, hipStream_t s) = 0;
};

// Tool description collected by the XML parser (ProblemSetup::sphTool)
struct ToolDesc {
    std::map<std::string, std::string> attrs;  // name, type, path, ...
    std::string text;                          // reduction operation body
    std::string xml_dir;                       // folder of the declaring file
    std::string get(const std::string& k, const std::string& d = "") const
    {
        auto it = attrs.find(k);
        return it == attrs.end() ? d : it->second;
    }
};

struct Definition {
    std::string name, value;
    bool evaluate = false;
};

struct LoadDesc {
    std::string format, file;
    std::vector<std::string> fields;
};

struct SetDesc {
    size_t n = 0;
    std::vector<std::pair<std::string, std::string>> scalars;
    std::vector<LoadDesc> loads, saves;
};

struct TimingDesc {
    double end_t = -1;       // Time
    long end_iter = -1;      // Steps
    long end_frame = -1;     // Frames
    double output_fps = -1;  // FPS
    long output_ipf = -1;    // IPF
};

// Everything parsed from the XML files before the runtime is built.
struct Problem {
    int dims = 3;
    std::vector<std::tuple<std::string, std::string, std::string, std::string>>
        variables;  // name, type, length, value
    std::vector<Definition> definitions;
    std::vector<std::shared_ptr<ToolDesc>> tools;
    // <Reports> (State.cpp:1395-1500), appended to the tools once parsed
    std::vector<std::shared_ptr<ToolDesc>> reports;
    std::vector<SetDesc> sets;
    TimingDesc timing;
    std::string base_path;
    std::string compile_flags;  // of the <Device> settings
    std::vector<std::string> xml_stack;
    bool is_defined(const std::string& n) const;
};

void parse_problem(const std::string& path, Problem& P,
                   const std::string& prefix);
// The problem of a case file, with the package folder as root path (no GPU
// work).
Problem load_problem(const std::string& xml_path, int dims);

// ---- kernel registry ----------------------------------------------------------
struct KernelBinding;
struct KernelSpec {
    std::vector<std::string> args;  // variable names bound by name
    // launch with the bound variables (same order as args)
    void (*launch)(Sim& S, Var* const* a, hipStream_t s);
};
const KernelSpec* find_kernel(const std::string& key);
std::vector<std::string> registered_kernels();
// "…/Scripts/cfd/Interactions.cl" + "entry" -> "cfd/Interactions.cl:entry"
std::string kernel_key(const std::string& path, const std::string& entry);

// ---- fusion of neighbour walks ------------------------------------------------
// When the tool list holds the Shepard, cfd-interactions and lapp tools with
// nothing in between that changes what they read, one walk per fluid
// particle evaluates all of them (each sum in the reference order, so the
// results are bit-identical) at the position of the first; the others
// become no-ops.  The walk also flags fluid particles with a wall element
// inside the support so the boundary-integral interactions skip the rest.
struct FusionPlan {
    bool active = false;
    int shep = 0;  // 0 none, 1 cfd/Shepard.cl, 2 BI compute
    bool lapp = false;
    bool near = false;
    const Tool* anchor = nullptr;
    std::set<const Tool*> absorbed;
    Var *imove = nullptr, *r = nullptr, *u = nullptr, *rho = nullptr,
        *m = nullptr, *p = nullptr, *normal = nullptr, *tangent = nullptr,
        *binormal = nullptr, *shepard = nullptr, *grad_p = nullptr,
        *lap_u = nullptr, *div_u = nullptr, *lap_p = nullptr,
        *icell = nullptr, *ihoc = nullptr, *n_cells = nullptr;
    uint8_t* near_buf = nullptr;
    unsigned walk_blocks = 0;  // 0: as many workgroups as fit at once
    int walk_segw = 8;  // accept-bit words per lane and segment (8 or 16)
    // cell hand-out ranges (1 or 8, one per XCD); 0: AQ_WALK_RANGES or 8
    unsigned walk_ranges = 0;
    // fused walk arithmetic: 0 IEEE-754 (correctly rounded division and
    // square root, no contraction); 1 relaxed (reciprocal division,
    // hardware square root, contracted multiply-adds); 2 IEEE terms with
    // compensated momentum and density sums (the __KAHAN_SUMS__ definition)
    int math = 0;
    // IEEE terms (math 0 and 2) through the division and square root
    // sequences without range scaling where the fields' ranges allow them
    // (bit-identical); AQ_DIV=ieee keeps the compiler's sequences
    bool fast_div = true;
    // profiling: events around the walk launch while `timing` is set
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double walk_ms = -1.0;  // mean of the last profile
    // Boundary-integral element chain beside the walk: the non-fluid part
    // of BI Shepard, then the `side` tools (BI LapU, GradP, Interpolation,
    // InterpolationShepard) run on side_stream while the fused walk runs on
    // the main stream; the main stream waits for them before `join` (BI
    // Interactions).  They read the fluid fields the walk only reads and
    // write element fields the walk and the tools in between never read.
    std::set<const Tool*> side_set;
    std::vector<Tool*> side;
    const Tool* join = nullptr;
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool side_launched = false;  // this step's chain is in flight
};

// ---- non-empty cell list of the cell-parallel walks ---------------------------
// Filled by the link-list tool that writes the "icell"/"ihoc" variables,
// right after its sort; pair kernels bound to those variables then run the
// cell-parallel walk instead of the per-particle one.
struct CellList {
    bool enabled = false;  // AQ_WALK != thread and N < 2^28
    bool filled = false;   // the link-list tool has run once
    bool kinds = false;    // cell_end carries the particle-kind flags
    Var *icell = nullptr, *ihoc = nullptr, *n_cells = nullptr;
    uint32_t* cell_end = nullptr;
    size_t cap = 0;  // cells cell_end holds
    uint32_t* heads = nullptr;
    uint32_t* ocells = nullptr;   // owner-cell triplets (cell_kinds_k)
    uint32_t* wcells = nullptr;   // wet owner cells (owner_wet_k)
    uint32_t* nlist = nullptr;    // near cells (the fused walk)
    // the non-empty cells in blocked order (the fused walk's hand-out),
    // with the per-block counters, bases and scan sums
    bool blocked = false;
    uint32_t* bcells = nullptr;
    uint32_t *bcount = nullptr, *bbase = nullptr, *bsums = nullptr;
    uint32_t nbmax = 0;
    uint32_t* n_heads = nullptr;  // [0] count, [32..] hand-out counters
    void* ws = nullptr;
    bool usable(const Var* ic, const Var* ih) const
    {
        return enabled && filled && ic == icell && ih == ihoc;
    }
};

// ---- the runtime ------------------------------------------------------------
class Sim {
  public:
    Sim(int dims);
    ~Sim();

    void load(const std::string& xml_path);  // parse + build
    void build(Problem& P);                  // from a parsed problem
    void setup();                            // tool setup + initial upload
    void step(int nsteps);                   // eager execution
    void step_graph(int nsteps);             // HIP-graph replay
    // time the fused-walk launch configurations on this case (state saved
    // and restored) and keep the fastest; returns a report
    std::string autotune(int nsteps);
    void drop_graphs();
    void check_status();                     // raise on device failures
    void sync();

    int dims;
    Variables vars;
    std::vector<std::shared_ptr<Tool>> tools;
    std::vector<Definition> definitions;
    std::map<std::string, std::string> def_values;  // formatted (-D values)
    aq_defs kdefs{};
    uint32_t N = 0;
    hipStream_t stream = nullptr;
    // [0]: bit0 ihoc overflow, bit1 assert failed; [1]: iter at the end of
    // the first step that flagged one (0xFFFFFFFF: none)
    uint32_t* d_status = nullptr;
    int32_t* d_fail_id = nullptr;
    std::vector<std::string> assert_names;
    Problem problem;
    std::string last_tool;
    bool graph_ready = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // control flow: jump[i] = matching end of an opener, opener of an end
    bool has_control = false;
    std::vector<size_t> jump;
    std::vector<char> if_taken;
    // with control flow, one graph per run of tools between control tools,
    // keyed by the run's first tool index
    struct Segment {
        size_t end = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::map<size_t, Segment> segments;
    // output (aquagpusph/TimeManager.cpp, FileManager.cpp): next file index
    // and last file of every set, the state file, the last output point
    struct Output {
        bool started = false;
        std::vector<unsigned> next_index;
        std::vector<std::string> last_file;
        std::string state_file;
        double time = 0.0;
        unsigned long step = 0;
    } out;
    // run until the Timing end criteria, saving at the output points;
    // returns the steps run (at most max_steps when > 0)
    long run(long max_steps, bool graph);
    // write every set's Save files and the state file now
    void save_now();
    bool is_setup = false;
    FusionPlan fusion;
    // The per-particle tail of the step (BI Shepard apply, cfd Rates, the
    // delta-SPH rate, the improved-Euler corrector) in one kernel when
    // those tools follow one another with only dummy tools between
    struct TailPlan {
        const Tool* anchor = nullptr;
        std::set<const Tool*> absorbed;
        bool apply = false, delta = false;
    } tail;
    // The head of the step (improved-Euler predictor, then Domain) in one
    // kernel, on the same terms
    TailPlan head;
    // The reorder (Sort stage 1 and 2), EOS and Binormal in one kernel
    TailPlan sortc;
    CellList clist;
    bool bi_cells = false;  // boundary-integral element kernels cell-parallel
    bool bi_group = false;  // ... over the owner-cell list
    // no side-stream chain (per-tool profiling), and the side-stream
    // launches in progress (cell_args then hands out the side counters)
    bool side_off = false;
    bool in_side = false;
    // AQ_NO_BATCH: set / copy tools and the sort stages' gathers one launch
    // each
    bool no_batch = false;
    // scratch of the per-particle walk kernels' compacted particle lists
    // (with the cell-parallel walks on): N indices and one counter
    uint32_t* sel_list = nullptr;
    uint32_t* sel_count = nullptr;
    void run_fused(hipStream_t s);
    void plan_tail();
    void run_tail(hipStream_t s);
    void run_head(hipStream_t s);
    void run_sort_chain(hipStream_t s);
    // the main stream waits for the side chain (no-op if none in flight)
    void join_side(hipStream_t s);
    CellArgs cell_args();
    // aids of a per-particle walk kernel bound to these link-list variables
    PartWalk part_walk(const Var* icell, const Var* ihoc, const Var* n_cells);

    bool defined(const std::string& n) const
    {
        return def_values.count(n) > 0;
    }
    std::string def(const std::string& n) const;
    float def_float(const std::string& n) const;
    // device pointer of a variable (array buffer or scalar slot)
    void* ptr(const std::string& name);
    void* ptr(Var* v);
    int assert_id(const std::string& name);
    // device address of periodic_L when the case declares a periodic box
    // (Presets/cfd/periodic.xml), nullptr otherwise
    const float* periodic_L()
    {
        Var* v = vars.get("periodic_L");
        return v ? (const float*)ptr(v) : nullptr;
    }
    // one step of the pipeline in order, following while/if/end; `hook` is
    // called around each executed tool (index, false before / true after).
    // With stop_on_assert (eager steps only: it synchronises the stream),
    // a failed assert raises right away and the tools after it in the step
    // are not run (the reference skips them: Tool.cpp:112-118,
    // tests/3D/Assert).
    void run_tools(hipStream_t s,
                   const std::function<void(size_t, bool)>& hook = {},
                   bool stop_on_assert = false);
    // record the iteration of the first failed step (end of each step)
    void stamp_status(hipStream_t s);

  private:
    void check_periodic();
    size_t control(size_t pc, hipStream_t s);
    void match_scopes();
    void step_segments(int nsteps);
    void plan_fusion();
    void plan_batches();
    void make_defs();
    void load_sets();
};

std::shared_ptr<Tool> make_tool(Sim& S, const ToolDesc& d);

// particle file I/O (FastASCII / ASCII rows, comma or blank separated)
void load_fastascii(Sim& S, const std::string& path, size_t offset, size_t n,
                    const std::vector<std::string>& fields);

// scalar bytecode on the device
void vm_launch(const VmOp* ops, uint8_t* arena, uint32_t* status,
               int32_t* fail_id, hipStream_t s);
void status_stamp(uint32_t* status, const uint32_t* iter, hipStream_t s);
// Batched fills (dst[f][i * stored[f] + k] = pat[f][k], i < n[f]) and
// whole-word copies (words[f] words from src[f] to dst[f]), up to 8 each
struct FillBatch {
    void* dst[8];
    const uint32_t* pat[8];
    int stored[8];
    size_t n[8];
    int cnt = 0;
};
struct CopyBatch {
    const void* src[8];
    void* dst[8];
    size_t words[8];
    int cnt = 0;
};
void fill_batch(const FillBatch& b, hipStream_t s);
void copy_batch(const CopyBatch& b, hipStream_t s);

}  // namespace rt
}  // namespace aq
