/*
 * Out-of-tree tools for the aquagpusph_amd runtime (type="installable").
 *
 * The counterpart of the reference's installable tools: a shared library
 * named by the tool's `path` attribute, loaded with dlopen when the tool is
 * set up (aquagpusph/CalcServer/CalcServer.cpp:375-400).  The reference
 * resolves one C++ symbol, create_object(std::string, bool), returning an
 * Aqua::CalcServer::Tool* (tests/ExternalTool/tool.hpp:28-40,
 * tool.cpp:27-32).  Here the plugin exports four plain C functions instead,
 * so no C++ type crosses the boundary and the runtime can enqueue the tool
 * inside a captured HIP graph:
 *
 *   aq_tool_create   construct the tool from its name and the XML attributes
 *                    of its <Tool> element (key, value, key, value, ...,
 *                    NULL), like create_object(name, once);
 *   aq_tool_setup    look up the variables it needs (Tool::setup);
 *   aq_tool_enqueue  enqueue its work on `stream` (a hipStream_t) without
 *                    synchronising, allocating or calling the host
 *                    (Tool::_execute);
 *   aq_tool_destroy  release it (the CalcServer deletes its tools).
 *
 * Functions return 0 on success; any other value makes the runtime raise
 * AQ_EINVAL naming the tool.  The plugin implements these symbols; the
 * runtime library does not export them.
 */
#ifndef AQUAGPUSPH_AMD_TOOL_H
#define AQUAGPUSPH_AMD_TOOL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Passed to aq_tool_setup.  var(sim, name, &bytes) returns the device
 * address of a variable (array buffer or scalar slot) and its size in
 * bytes, or NULL when no variable has that name. */
typedef struct aq_tool_ctx {
    void* sim;
    void* (*var)(void* sim, const char* name, size_t* bytes);
} aq_tool_ctx;

int aq_tool_create(const char* name, const char* const* kv, void** tool);
int aq_tool_setup(void* tool, const aq_tool_ctx* ctx);
int aq_tool_enqueue(void* tool, void* stream);
void aq_tool_destroy(void* tool);

#ifdef __cplusplus
}
#endif

#endif /* AQUAGPUSPH_AMD_TOOL_H */
