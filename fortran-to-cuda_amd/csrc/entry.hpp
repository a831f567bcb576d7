// Glue shared by the hand-written HIP entries: layout checks, exact access sections and submission of a
// kernel as one task of the context's graph queue. An entry follows the order of the reference's
// generated entries (compiler/emission/cuda/scoped.py:237-330): ABI and mode checks, layout and type
// checks per array, alias checks of written arrays, definition kill of INTENT(OUT) arrays, device
// reservation, then one queued task per kernel with its exact read and (over)write sections.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../include/fort_scope.h"

namespace ftc {

#define FTC_CHECK(expr)                  \
    do {                                 \
        int ftc_status_ = (expr);        \
        if (ftc_status_) return ftc_status_; \
    } while (false)

inline int entry_preamble(int mode)
{
    if (fort_scope_abi_version() != FORT_SCOPE_ABI_VERSION)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "incompatible shared runtime ABI");
    if (mode < 0 || mode > 1) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "unknown shared execution mode");
    // This library holds device implementations only; the caller's original statements are the native
    // form of every entry.
    if (mode == FORT_SCOPE_NATIVE) return FORT_SCOPE_DECLINED;
    return FORT_SCOPE_OK;
}

// Reads the layout of a rank-3 array of `type` and returns its extents.
inline int layout3(fort_scope_t ctx, fort_buffer_t h, uint32_t type, std::size_t elem, std::size_t ext[3])
{
    fort_scope_layout l{};
    FTC_CHECK(fort_scope_layout_get(ctx, h, &l));
    if (l.rank != 3 || l.type != type || l.element_bytes != elem)
        return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "entry layout/type mismatch");
    for (int d = 0; d < 3; ++d) ext[d] = l.extents[d];
    return FORT_SCOPE_OK;
}

// A zero-based half-open box of a rank-3 array, in storage that outlives the submission call.
struct Sec3 {
    std::size_t lo[3], hi[3];
};

struct Binding3 {
    fort_buffer_t handle;
    std::vector<Sec3> reads, writes, overwrites;
};

// Submits one kernel with its bindings to the context queue. `args` is one struct passed by value.
inline int enqueue_kernel(fort_scope_t ctx, uint64_t code_key, const void *function, dim3 grid, dim3 block,
                          std::size_t shared, const void *args, std::size_t arg_bytes,
                          const std::vector<Binding3> &bindings)
{
    std::vector<fort_scope_plan_binding> plan(bindings.size());
    std::vector<std::vector<fort_scope_section>> rs(bindings.size()), ws(bindings.size()), os(bindings.size());
    for (std::size_t i = 0; i < bindings.size(); ++i) {
        for (const Sec3 &s : bindings[i].reads) rs[i].push_back({s.lo, s.hi});
        for (const Sec3 &s : bindings[i].writes) ws[i].push_back({s.lo, s.hi});
        // An overwrite is also a write: the runtime takes must-overwrite sections within may-write ones.
        for (const Sec3 &s : bindings[i].overwrites) {
            ws[i].push_back({s.lo, s.hi});
            os[i].push_back({s.lo, s.hi});
        }
        plan[i].buffer = bindings[i].handle;
        plan[i].access.flags = 0;
        plan[i].access.read_count = rs[i].size();
        plan[i].access.reads = rs[i].data();
        plan[i].access.write_count = ws[i].size();
        plan[i].access.writes = ws[i].data();
        plan[i].access.overwrite_count = os[i].size();
        plan[i].access.overwrites = os[i].data();
    }
    void *argv[1] = {const_cast<void *>(args)};
    std::size_t sizes[1] = {arg_bytes};
    fort_scope_graph_task_v1 task{};
    task.kernel.version = 1;
    task.kernel.function = function;
    task.kernel.grid[0] = grid.x;
    task.kernel.grid[1] = grid.y;
    task.kernel.grid[2] = grid.z;
    task.kernel.block[0] = block.x;
    task.kernel.block[1] = block.y;
    task.kernel.block[2] = block.z;
    task.kernel.shared_bytes = shared;
    task.kernel.arguments = argv;
    task.kernel.argument_count = 1;
    task.bindings = plan.data();
    task.binding_count = plan.size();
    return fort_scope_graph_enqueue_v1(ctx, code_key, &task, sizes);
}

inline int reserve(fort_scope_t ctx, const std::vector<fort_buffer_t> &handles, std::vector<void *> &devices)
{
    devices.assign(handles.size(), nullptr);
    return fort_scope_device_reserve_many_v1(ctx, handles.data(), handles.size(), devices.data());
}

inline int distinct_outputs(const std::vector<fort_buffer_t> &outs, const std::vector<fort_buffer_t> &ins)
{
    for (std::size_t a = 0; a < outs.size(); ++a) {
        for (std::size_t b = 0; b < a; ++b)
            if (outs[a] == outs[b])
                return fort_scope_report_error(FORT_SCOPE_ALIAS, "writable entry arguments alias");
        for (fort_buffer_t in : ins)
            if (outs[a] == in) return fort_scope_report_error(FORT_SCOPE_ALIAS, "writable entry arguments alias");
    }
    return FORT_SCOPE_OK;
}

}  // namespace ftc
