/*
 * A demonstration out-of-tree tool for type="installable" (the counterpart
 * of the reference's tests/ExternalTool/tool.cpp, which logs a line each
 * time it runs).  This one announces itself on stdout when it is first
 * enqueued and, when its <Tool> element names `in` and `out` variables of
 * the same size, copies `in` into `out` on the runtime's stream (a plain
 * device-to-device copy, so the tool also works inside a captured HIP
 * graph).  Built by __graft_entry__.build() with gcc into
 * tests/plugin/libaq_demo_tool.so.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "aquagpusph_amd_tool.h"

typedef struct {
    char name[128];
    char in[64], out[64];
    void *src, *dst;
    size_t bytes;
    int announced;
} demo_tool;

int aq_tool_create(const char* name, const char* const* kv, void** tool)
{
    demo_tool* t = (demo_tool*)calloc(1, sizeof(demo_tool));
    if (!t) return -1;
    snprintf(t->name, sizeof(t->name), "%s", name ? name : "");
    for (const char* const* p = kv; p && p[0] && p[1]; p += 2) {
        if (!strcmp(p[0], "in")) snprintf(t->in, sizeof(t->in), "%s", p[1]);
        if (!strcmp(p[0], "out"))
            snprintf(t->out, sizeof(t->out), "%s", p[1]);
    }
    *tool = t;
    return 0;
}

int aq_tool_setup(void* tool, const aq_tool_ctx* ctx)
{
    demo_tool* t = (demo_tool*)tool;
    if (!t->in[0] && !t->out[0]) return 0;
    size_t bi = 0, bo = 0;
    t->src = ctx->var(ctx->sim, t->in, &bi);
    t->dst = ctx->var(ctx->sim, t->out, &bo);
    if (!t->src || !t->dst || bi != bo) return -2;
    t->bytes = bi;
    return 0;
}

int aq_tool_enqueue(void* tool, void* stream)
{
    demo_tool* t = (demo_tool*)tool;
    if (!t->announced) {
        printf("[INFO] (%s): Executing Installed demo tool!\n", t->name);
        fflush(stdout);
        t->announced = 1;
    }
    if (t->bytes &&
        hipMemcpyAsync(t->dst, t->src, t->bytes, hipMemcpyDeviceToDevice,
                       (hipStream_t)stream) != hipSuccess)
        return -3;
    return 0;
}

void aq_tool_destroy(void* tool) { free(tool); }
