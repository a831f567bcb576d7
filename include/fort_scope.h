/* fort_scope: the residency runtime's C ABI, MI355X (gfx950) implementation.
 *
 * Each declaration restates one entry point of the reference runtime ABI
 * (compiler/runtime/scoped_runtime.h, FORT_SCOPE_ABI_VERSION 1) with the same name, argument meaning,
 * status codes and struct layouts, so that the reference's Fortran interface module
 * (compiler/runtime/scoped_memory.f90) and generated entries bind to it unchanged. The citation on each
 * line is the reference declaration it replaces. Functions of the reference ABI not listed here are not
 * provided by this library.
 *
 * Semantics kept from the reference: handles are opaque 64-bit values; host arrays are borrowed and must
 * stay valid until unregister or close; coordinates are zero-based physical offsets with exclusive upper
 * bounds; one caller at a time per context; every call returns an int status (0 = OK) with a thread-local
 * message behind fort_scope_error().
 */
#ifndef FTC_FORT_SCOPE_H
#define FTC_FORT_SCOPE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FORT_SCOPE_ABI_VERSION 1               /* scoped_runtime.h:16 */
typedef uint64_t fort_scope_t;                 /* scoped_runtime.h:17 */
typedef uint64_t fort_buffer_t;                /* scoped_runtime.h:18 */
typedef uint64_t fort_scope_domain_t;          /* scoped_runtime.h:19 */

enum fort_scope_status {                       /* scoped_runtime.h:42-49 */
    FORT_SCOPE_OK = 0, FORT_SCOPE_ARGUMENT = 1, FORT_SCOPE_STALE = 2, FORT_SCOPE_ALIAS = 3,
    FORT_SCOPE_RESOURCE = 4, FORT_SCOPE_BOUNDARY = 5, FORT_SCOPE_EXECUTION = 6,
    FORT_SCOPE_UNINITIALIZED = 7, FORT_SCOPE_STATE = 8, FORT_SCOPE_DECLINED = 9
};
enum fort_scope_type {                         /* scoped_runtime.h:50-53 */
    FORT_SCOPE_BYTES = 0, FORT_SCOPE_REAL32 = 1, FORT_SCOPE_REAL64 = 2, FORT_SCOPE_INTEGER32 = 3,
    FORT_SCOPE_LOGICAL = 4
};
enum fort_scope_access_flags {                 /* scoped_runtime.h:54-56 */
    FORT_SCOPE_READ_ALL = 1, FORT_SCOPE_WRITE_ALL = 2, FORT_SCOPE_OVERWRITE_ALL = 4
};
enum fort_scope_execution_mode { FORT_SCOPE_NATIVE = 0, FORT_SCOPE_GPU = 1 }; /* scoped_runtime.h:57 */

typedef struct fort_scope_section {            /* scoped_runtime.h:64-67 */
    const size_t *lower;
    const size_t *upper;
} fort_scope_section;
typedef struct fort_scope_access {             /* scoped_runtime.h:68-76 */
    uint32_t flags;
    size_t read_count;
    const fort_scope_section *reads;
    size_t write_count;
    const fort_scope_section *writes;
    size_t overwrite_count;
    const fort_scope_section *overwrites;
} fort_scope_access;
typedef struct fort_scope_layout {             /* scoped_runtime.h:77-84 */
    uint32_t rank, type;
    size_t element_bytes;
    void *host;
    const size_t *extents;
    const int64_t *lower_bounds;
    uint64_t generation;
} fort_scope_layout;
typedef struct fort_scope_stats {              /* scoped_runtime.h:93-97 */
    uint64_t uploads, downloads, upload_bytes, download_bytes;
    uint64_t allocations, allocated_bytes, peak_device_bytes;
    uint64_t launches, waits, reconciliations;
} fort_scope_stats;
typedef struct fort_scope_scratch_lease_v1 {   /* scoped_runtime.h:105-110 */
    uint32_t version, reserved;
    uint64_t token;
    void *device;
    size_t bytes, capacity;
} fort_scope_scratch_lease_v1;
/* Scratch and completion counters (scoped_runtime.h:58-63, 111-117). */
typedef struct fort_scope_scratch_stats_v1 {
    uint32_t version, active;
    uint64_t acquisitions, allocations, reuses, releases, grows;
    uint64_t capacity_bytes, active_bytes, peak_scratch_bytes, peak_total_device_bytes;
} fort_scope_scratch_stats_v1;
typedef struct fort_scope_completion_stats_v1 {
    uint32_t version, reserved;
    uint64_t event_records, event_queries, event_waits, stream_waits;
    uint64_t created_events, reused_events, independent_host_tasks;
} fort_scope_completion_stats_v1;
/* Submission provenance for diagnostics (scoped_runtime.h:20-41). */
enum fort_scope_trace_field {
    FORT_SCOPE_TRACE_OWNER = 1, FORT_SCOPE_TRACE_PROCEDURE = 2, FORT_SCOPE_TRACE_SEGMENT = 4,
    FORT_SCOPE_TRACE_OPERATION = 8, FORT_SCOPE_TRACE_IMPLEMENTATION = 16, FORT_SCOPE_TRACE_BOUNDARY = 32,
    FORT_SCOPE_TRACE_ALL = 63
};
typedef struct fort_scope_trace_state_v1 {
    uint32_t version, reserved;
    char owner[65], procedure[65], segment[65], operation[65], implementation[65], boundary[65];
} fort_scope_trace_state_v1;
typedef struct fort_scope_plan_binding {       /* scoped_runtime.h:120-123 */
    fort_buffer_t buffer;
    fort_scope_access access;
} fort_scope_plan_binding;
typedef struct fort_scope_residency_stats_v1 { /* scoped_runtime.h:336-341 */
    uint32_t version, reserved;
    uint64_t reserve_batches, requested_resources, reuse_hits;
    uint64_t evictions, evicted_bytes, dirty_evictions, dirty_download_bytes, pressure_flushes;
} fort_scope_residency_stats_v1;
typedef struct fort_scope_kernel_v1 {          /* scoped_runtime.h:362-369 */
    uint32_t version, reserved;
    const void *function;
    uint32_t grid[3], block[3];
    size_t shared_bytes;
    void *const *arguments;
    size_t argument_count;
} fort_scope_kernel_v1;
typedef struct fort_scope_graph_stats_v1 {     /* scoped_runtime.h:370-374 */
    uint32_t version, reserved;
    uint64_t creations, instantiations, rebuilds, submissions, replays;
    uint64_t parameter_updates, logical_kernels, live_graphs, destructions;
} fort_scope_graph_stats_v1;
typedef struct fort_scope_graph_task_v1 {      /* scoped_runtime.h:375-379 */
    fort_scope_kernel_v1 kernel;
    const fort_scope_plan_binding *bindings;
    size_t binding_count;
} fort_scope_graph_task_v1;
typedef struct fort_scope_graph_queue_stats_v1 { /* scoped_runtime.h:380-385 */
    uint32_t version, queue_enabled, graphs_enabled, reserved;
    uint64_t accepted_tasks, submissions, conflict_flushes, explicit_flushes;
    uint64_t pending_tasks, peak_pending_tasks, copied_argument_bytes;
    uint64_t direct_submissions, direct_kernels, graph_resource_fallbacks;
} fort_scope_graph_queue_stats_v1;
typedef struct fort_scope_external_binding_v1 { /* scoped_runtime.h:265-270 */
    fort_buffer_t buffer;
    uint32_t type, rank;
    size_t element_bytes;
    fort_scope_access access;
} fort_scope_external_binding_v1;
typedef struct fort_scope_external_view_v1 {   /* scoped_runtime.h:271-279 */
    fort_buffer_t buffer;
    uint64_t generation;
    uint32_t type, rank;
    size_t element_bytes, bytes;
    const size_t *extents, *strides; /* strides in bytes */
    const int64_t *lower_bounds;
    void *device;
} fort_scope_external_view_v1;
typedef struct fort_scope_external_call_v1 {   /* scoped_runtime.h:280-286 */
    uint32_t version, reserved;
    const fort_scope_external_binding_v1 *bindings;
    size_t binding_count;
    uint64_t workspace_lease;
    size_t workspace_bytes;
} fort_scope_external_call_v1;
typedef struct fort_scope_external_report_v1 { /* scoped_runtime.h:287-291 */
    uint32_t version;
    int32_t status;
    uint32_t started, committed;
} fort_scope_external_report_v1;
typedef int (*fort_scope_external_worker_v1)(void *user, const fort_scope_external_view_v1 *views, size_t count,
                                             void *stream, int *started); /* scoped_runtime.h:292-293 */

uint32_t fort_scope_abi_version(void);                                                /* :124 */
const char *fort_scope_error(void);                                                   /* :125 */
int fort_scope_create(int device, fort_scope_t *context);                            /* :126 */
int fort_scope_domain_create_v1(int device, fort_scope_domain_t *domain);            /* :138 */
int fort_scope_create_in_domain_v1(fort_scope_domain_t domain, fort_scope_t *context); /* :139 */
int fort_scope_domain_get_v1(fort_scope_t context, fort_scope_domain_t *domain);     /* :142 */
int fort_scope_domain_device_get_v1(fort_scope_domain_t domain, int *device);        /* :145 */
int fort_scope_domain_stats_get_v1(fort_scope_domain_t domain, fort_scope_stats *stats); /* :149 */
int fort_scope_domain_close_v1(fort_scope_domain_t domain);                          /* :150 */
int fort_scope_domain_abandon_v1(fort_scope_domain_t domain);                        /* :155 */
int fort_scope_domain_workspace_reserve_v1(fort_scope_domain_t domain, size_t bytes, uint64_t *lease); /* :163 */
int fort_scope_domain_workspace_resize_v1(fort_scope_domain_t domain, uint64_t lease, size_t bytes);   /* :164 */
int fort_scope_domain_workspace_release_v1(fort_scope_domain_t domain, uint64_t lease);              /* :165 */
int fort_scope_set_device_budget(fort_scope_t context, size_t bytes);                /* :168 */
int fort_scope_device_budget_default_v1(int device, size_t *bytes);                  /* :174 */
int fort_scope_domain_budget_get_v1(fort_scope_domain_t domain, size_t *budget, size_t *unclaimed); /* :178 */
int fort_scope_scratch_acquire_v1(fort_scope_t context, size_t bytes, fort_scope_scratch_lease_v1 *lease); /* :186-187 */
int fort_scope_scratch_release_v1(fort_scope_t context, uint64_t token);             /* :188 */
int fort_scope_scratch_stats_get_v1(fort_scope_t context, fort_scope_scratch_stats_v1 *stats); /* :189 */
/* This runtime orders all work of a context on its one stream and records no events: the event counters
 * stay zero, stream_waits counts stream synchronisations, independent_host_tasks counts host accesses
 * that needed no transfer. (:190) */
int fort_scope_completion_stats_get_v1(fort_scope_t context, fort_scope_completion_stats_v1 *stats);
/* Attribution of later submissions, kept per context while FORT_RUNTIME_TRACE is set (otherwise both
 * calls return at once): masked fields are replaced (null or empty clears them), *previous receives the
 * state before the call; restore reinstates a saved state. Never changes status, coherence or order.
 * (:38-41) */
void fort_scope_trace_set_v1(fort_scope_t context, uint32_t update_mask, const char *owner, const char *procedure,
                             const char *segment, const char *operation, const char *implementation,
                             const char *boundary, fort_scope_trace_state_v1 *previous);
void fort_scope_trace_restore_v1(fort_scope_t context, const fort_scope_trace_state_v1 *previous);
int fort_scope_register(fort_scope_t context, uint64_t identity, uint64_t generation,
                        const fort_scope_layout *layout, int host_initialized, fort_buffer_t *buffer); /* :191-192 */
int fort_scope_register_sections(fort_scope_t context, uint64_t identity, uint64_t generation,
                                 const fort_scope_layout *layout, const fort_scope_section *initialized,
                                 size_t initialized_count, fort_buffer_t *buffer); /* :195-197 */
int fort_scope_forget_definition(fort_scope_t context, fort_buffer_t buffer);       /* :217 */
int fort_scope_layout_get(fort_scope_t context, fort_buffer_t buffer, fort_scope_layout *layout); /* :218 */
int fort_scope_access_begin_v1(fort_scope_t context, const fort_scope_plan_binding *bindings, size_t count,
                               int device, void **devices);                          /* :232-233 */
/* Original CPU work (the unchanged source body) with may-access envelopes: the tracked values of the read
 * and write regions reach the host, holes that were never defined are neither copied nor defined, device
 * replicas of reads stay; host_end invalidates the device copies of possible writes and only overwrites
 * add definitions. All bindings are checked before any transfer. (scoped_runtime.h:234-246) */
int fort_scope_original_host_begin_v1(fort_scope_t context, const fort_scope_plan_binding *bindings,
                                      size_t count);
/* Commit a descriptor-only exchange of two registered identities that the caller has already executed
 * (a MOVE_ALLOC swap): first_after must describe the second physical allocation exactly (host address,
 * extents, lower bounds, generation) and second_after the first. Only the identity map changes; no data
 * moves. Any failure poisons the context and its domain and returns EXECUTION, with the outputs
 * untouched. (scoped_runtime.h:198-211) */
int fort_scope_identity_swap_v1(fort_scope_t context, fort_buffer_t first, fort_buffer_t second,
                                const fort_scope_layout *first_after, const fort_scope_layout *second_after,
                                fort_buffer_t *first_current, fort_buffer_t *second_current);
int fort_scope_external_execute_v1(fort_scope_t context, const fort_scope_external_call_v1 *call,
                                   fort_scope_external_worker_v1 worker, void *user,
                                   fort_scope_external_report_v1 *report);          /* :294-296 */
int fort_scope_observe_v1(fort_scope_t context, const fort_buffer_t *buffers, size_t count); /* :303 */
int fort_scope_observe_all_v1(fort_scope_t context);                                /* :309 */
int fort_scope_discard_v1(fort_scope_t context, fort_buffer_t buffer);              /* :310 */
/* Dead values its owner is about to define again: as discard when no queued task uses the buffer, else
 * the values stay until the new definition replaces them and the queue is not submitted. (:311-314) */
int fort_scope_recycle_v1(fort_scope_t context, fort_buffer_t buffer);
/* Destroys a dead resource without publication, after its queued and running uses complete; the handle
 * becomes stale. (:298-300, 315) */
int fort_scope_retire_v1(fort_scope_t context, fort_buffer_t buffer);
/* A reached source-read precondition: values of the sections outside the defined coverage become defined
 * with a current host copy; defined values keep their freshness; no transfer or wait. Exceeding the region
 * limit commits nothing (BOUNDARY). (:85-92) */
int fort_scope_admit_source_reads_v1(fort_scope_t context, fort_buffer_t buffer, const fort_scope_section *sections,
                                     size_t count);
int fort_scope_host_begin(fort_scope_t context, fort_buffer_t buffer, const fort_scope_access *access); /* :316 */
int fort_scope_host_end(fort_scope_t context, fort_buffer_t buffer);                /* :317 */
int fort_scope_device_begin(fort_scope_t context, fort_buffer_t buffer, const fort_scope_access *access,
                            void **device);                                         /* :318 */
int fort_scope_device_end(fort_scope_t context, fort_buffer_t buffer);              /* :319 */
int fort_scope_device_reserve_v1(fort_scope_t context, fort_buffer_t buffer, void **device); /* :323 */
int fort_scope_device_reserve_many_v1(fort_scope_t context, const fort_buffer_t *buffers, size_t count,
                                      void **devices);                              /* :330-331 */
int fort_scope_buffer_valid_v1(fort_scope_t context, fort_buffer_t buffer, int *valid); /* :334 */
int fort_scope_residency_stats_get_v1(fort_scope_t context, fort_scope_residency_stats_v1 *stats); /* :341 */
int fort_scope_cancel_access(fort_scope_t context, fort_buffer_t buffer);           /* :343 */
int fort_scope_gpu_enter(fort_scope_t context, int *previous_device, void **stream); /* :346 */
int fort_scope_gpu_leave(fort_scope_t context, int previous_device);                /* :347 */
int fort_scope_note_launch(fort_scope_t context);                                   /* :348 */
int fort_scope_graph_enqueue_v1(fort_scope_t context, uint64_t code_key, const fort_scope_graph_task_v1 *task,
                                const size_t *argument_sizes);                      /* :399-400 */
/* One launch whose kernel runs several independent members, each with its own effects. Accepted only when
 * no member reads or writes a section that another member writes (judged on the actual sections);
 * otherwise returns BOUNDARY, queues nothing, and the caller runs the members one after another. The
 * accepted launch carries the union of the members' effects. The CPU backend returns RESOURCE for an
 * accepted nonempty launch, after the overlap check. */
typedef struct fort_scope_graph_member_v1 {    /* scoped_runtime.h:407-410 */
    const fort_scope_plan_binding *bindings;
    size_t binding_count;
} fort_scope_graph_member_v1;
int fort_scope_graph_enqueue_members_v1(fort_scope_t context, uint64_t code_key, const fort_scope_kernel_v1 *kernel,
                                        const size_t *argument_sizes, const fort_scope_graph_member_v1 *members,
                                        size_t member_count);                       /* :411-413 */
/* An ordered kernel group on the context stream (each node after its predecessor), for kernels whose
 * resources the caller already acquired (device_end follows). A nonzero group key reuses its executable
 * graph, rebuilt when functions or argument counts change, every replay updating all node parameters;
 * key 0 launches directly. Argument values are copied during the call. (scoped_runtime.h:349-360, 417) */
int fort_scope_graph_submit_v1(fort_scope_t context, uint64_t group_key, const fort_scope_kernel_v1 *kernels,
                               size_t count);
/* Checked group execution: the tasks' effects are simulated in order, so a producer may define what a
 * later task reads; only the requirements defined before the group are uploaded, and definitions and
 * coherence are committed after a successful submission. No binding may be prepared; device storage is
 * reserved (device_reserve) before the arguments are built. The CPU backend returns RESOURCE after
 * validation. (scoped_runtime.h:418-426) */
int fort_scope_graph_execute_v1(fort_scope_t context, uint64_t group_key, const fort_scope_graph_task_v1 *tasks,
                                size_t count);
int fort_scope_graph_flush_v1(fort_scope_t context);                                /* :414 */
int fort_scope_graph_queue_cancel_v1(fort_scope_t context);                         /* :415 */
int fort_scope_graph_queue_stats_get_v1(fort_scope_t context, fort_scope_graph_queue_stats_v1 *stats); /* :416 */
int fort_scope_graph_stats_get_v1(fort_scope_t context, fort_scope_graph_stats_v1 *stats); /* :427 */
int fort_scope_execution_error(fort_scope_t context, const char *message);         /* :430 */
int fort_scope_report_error(int status, const char *message);                       /* :431 */
int fort_scope_wait(fort_scope_t context);                                          /* :432 */
int fort_scope_stats_get(fort_scope_t context, fort_scope_stats *stats);            /* :433 */
int fort_scope_unregister(fort_scope_t context, fort_buffer_t buffer);              /* :434 */
int fort_scope_close(fort_scope_t context);                                         /* :435 */
int fort_scope_abandon(fort_scope_t context);                                       /* :438 */

#ifdef __cplusplus
}
#endif
#endif
