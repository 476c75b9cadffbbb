// Autotuner of the fused cell walk: the segment width (8 or 16 accept-bit
// words per lane), the cell hand-out ranges (one per XCD, or one) and the
// grid size (all resident workgroups, or half of them) are timed on the
// case itself, with HIP events around the walk over a few eager steps, and
// the fastest is kept.  The state is saved before and restored after, so
// tuning does not advance the simulation; captured graphs are dropped so
// that the next graph step records the chosen launch.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "sim.hpp"

namespace aq {
namespace rt {

void Sim::drop_graphs()
{
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    for (auto& kv : segments) {
        if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
        if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
    }
    segments.clear();
    graph_ready = false;
}

std::string Sim::autotune(int nsteps)
{
    if (!is_setup) throw Error(AQ_EINVAL, "setup() has not been called");
    FusionPlan& F = fusion;
    if (!F.active || !clist.enabled || nsteps <= 0)
        return "nothing to tune (no fused cell walk)";
    if (vars.host_dirty()) vars.push(stream);
    AQ_HIP(hipStreamSynchronize(stream));
    // save the state: every array and the scalar arena.  (An array that a
    // tool grows during the trial steps, the lazily sized ihoc, is one
    // that tool rebuilds every step: its saved prefix is restored.)
    struct Saved {
        Var* v;
        void* b;
        size_t bytes;
    };
    std::vector<Saved> saved;
    for (auto& vp : vars.all()) {
        Var* v = vp.get();
        if (!v->is_array() || !v->dev || !v->bytes) continue;
        void* b = nullptr;
        AQ_HIP(hipMalloc(&b, v->bytes));
        AQ_HIP(hipMemcpyAsync(b, v->dev, v->bytes, hipMemcpyDeviceToDevice,
                              stream));
        saved.push_back({ v, b, v->bytes });
    }
    vars.pull(stream);
    const std::vector<uint8_t> arena = vars.host_arena();
    if (!F.ev0) AQ_HIP(hipEventCreate(&F.ev0));
    if (!F.ev1) AQ_HIP(hipEventCreate(&F.ev1));

    struct Cand {
        int segw;
        unsigned ranges;
        unsigned blocks;
        double ms;
    };
    // grid: every resident workgroup (0), or one per compute unit
    int dev = 0, ncu = 0;
    AQ_HIP(hipGetDevice(&dev));
    AQ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                 dev));
    std::vector<Cand> cands;
    for (int segw : { 8, 16 })
        for (unsigned ranges : { 8u, 1u })
            for (unsigned blocks : { 0u, (unsigned)ncu })
                cands.push_back({ segw, ranges, blocks, 0.0 });
    const Cand keep{ F.walk_segw, F.walk_ranges, F.walk_blocks, 0.0 };
    std::string report;
    Cand best{ 0, 0, 0, 1e300 };
    for (auto& c : cands) {
        F.walk_segw = c.segw;
        F.walk_ranges = c.ranges;
        F.walk_blocks = c.blocks;
        double acc = 0.0;
        int n = 0;
        for (int k = 0; k < nsteps + 1; k++) {
            F.timing = true;
            run_tools(stream);
            F.timing = false;
            AQ_HIP(hipEventSynchronize(F.ev1));
            float x = 0.f;
            // the first step of a candidate is a warm-up
            if (k && hipEventElapsedTime(&x, F.ev0, F.ev1) == hipSuccess) {
                acc += x;
                n++;
            }
        }
        c.ms = n ? acc / n : 1e300;
        char line[160];
        snprintf(line, sizeof(line), "segw=%d ranges=%u blocks=%u: %.4f ms\n",
                 c.segw, c.ranges, c.blocks, c.ms);
        report += line;
        if (c.ms < best.ms) best = c;
        // back to the saved state before the next candidate
        for (auto& sv : saved)
            AQ_HIP(hipMemcpyAsync(sv.v->dev, sv.b,
                                  std::min(sv.bytes, sv.v->bytes),
                                  hipMemcpyDeviceToDevice, stream));
        vars.host_arena() = arena;
        vars.push(stream);
    }
    AQ_HIP(hipStreamSynchronize(stream));
    for (auto& sv : saved) (void)hipFree(sv.b);
    check_status();
    if (best.ms >= 1e300) {
        F.walk_segw = keep.segw;
        F.walk_ranges = keep.ranges;
        F.walk_blocks = keep.blocks;
        return report + "no timing: kept the configuration\n";
    }
    F.walk_segw = best.segw;
    F.walk_ranges = best.ranges;
    F.walk_blocks = best.blocks;
    drop_graphs();
    char line[160];
    snprintf(line, sizeof(line), "chosen: segw=%d ranges=%u blocks=%u\n",
             best.segw, best.ranges, best.blocks);
    return report + line;
}

}  // namespace rt
}  // namespace aq
