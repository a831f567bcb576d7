// fort_scope runtime for MI355X: residency domains, HBM mirrors of borrowed host arrays with exact section
// coherence, a task queue that submits each batch as one hipGraph, scratch leases and the external-provider
// transaction. The ABI it implements is declared in include/fort_scope.h.
//
// Reference behaviour followed (compiler/runtime/scoped_runtime.cu): registration and layouts (:1834-1900),
// prepared device/host accesses with uploads of only the missing sections (:899-954, :1010-1036), downloads
// on observation or host access (:734-777, :873-885), LRU eviction under a byte budget (:1434-1529), the
// command queue with graph replay (:1038-1149, :1227-1401), poisoning after execution errors (:363-384).
//
// Built twice: with hipcc for the device backend, and with g++ -DFTC_CPU_BACKEND as a reference backend
// whose "device" memory is host memory and whose graph submission returns RESOURCE, so that coherence,
// eviction and queue bookkeeping are testable without a GPU.
#include "../../include/fort_scope.h"
#include "regions.hpp"

#include <algorithm>
#include <atomic>
#include <cfenv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#ifndef FTC_CPU_BACKEND
#include <hip/hip_runtime_api.h>
#endif

namespace ftc {

// ------------------------------------------------------------------------------------------------ errors
struct Error {
    int status;
    std::string message;
};
static thread_local std::string g_error;

[[noreturn]] static void fail(int status, const std::string &message) { throw Error{status, message}; }
static void require(bool ok, int status, const char *message)
{
    if (!ok) fail(status, message);
}

// ------------------------------------------------------------------------------------------------ device
namespace dev {
#ifdef FTC_CPU_BACKEND
using Stream = void *;
static int set(int) { return 0; }
static int current() { return 0; }
static void *alloc(std::size_t n) { return std::malloc(n ? n : 1); }
static void release(void *p) { std::free(p); }
static void copy2d(void *dst, std::size_t dpitch, const void *src, std::size_t spitch, std::size_t width,
                   std::size_t height, bool, Stream)
{
    for (std::size_t r = 0; r < height; ++r)
        std::memcpy(static_cast<char *>(dst) + r * dpitch, static_cast<const char *>(src) + r * spitch, width);
}
static void sync(Stream) {}
constexpr bool kStaged = false;  // "device" memory is host memory: strided copies are plain memcpy loops
static void *host_alloc(std::size_t n) { return std::malloc(n ? n : 1); }
static void host_release(void *p) { std::free(p); }
static Stream make_stream() { return nullptr; }
static void prime(Stream, void *, std::size_t) {}
static void destroy_stream(Stream) {}
static std::size_t free_bytes() { return std::size_t(64) << 30; }
static std::size_t total_bytes() { return std::size_t(64) << 30; }
#else
using Stream = hipStream_t;
static void check(hipError_t e, int status, const char *what)
{
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(status, std::string(what) + ": " + hipGetErrorString(e));
    }
}
static int set(int d)
{
    check(hipSetDevice(d), FORT_SCOPE_RESOURCE, "hipSetDevice");
    return 0;
}
static int current()
{
    int d = 0;
    check(hipGetDevice(&d), FORT_SCOPE_RESOURCE, "hipGetDevice");
    return d;
}
static void *alloc(std::size_t n)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, n ? n : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return p;
}
static void release(void *p) { (void)hipFree(p); }
static void copy2d(void *dst, std::size_t dpitch, const void *src, std::size_t spitch, std::size_t width,
                   std::size_t height, bool h2d, Stream s)
{
    const hipMemcpyKind kind = h2d ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    // A contiguous block takes the linear copy: the pitched one spent milliseconds per megabyte in the
    // runtime for it, pinned or not.
    if (height == 1 || (dpitch == width && spitch == width))
        check(hipMemcpyAsync(dst, src, width * height, kind, s), FORT_SCOPE_EXECUTION, "hipMemcpyAsync");
    else
        check(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, s), FORT_SCOPE_EXECUTION,
              "hipMemcpy2DAsync");
}
static void sync(Stream s) { check(hipStreamSynchronize(s), FORT_SCOPE_EXECUTION, "hipStreamSynchronize"); }
// Strided copies between the device and pageable host memory (the application's arrays) move row by row
// inside the HIP runtime, at a few MB/s for rows of a few hundred bytes; between the device and pinned
// memory they are pitched DMA. So strided boxes travel through a pinned staging buffer.
constexpr bool kStaged = true;
static void *host_alloc(std::size_t n)
{
    void *p = nullptr;
    check(hipHostMalloc(&p, n ? n : 1, hipHostMallocDefault), FORT_SCOPE_RESOURCE, "hipHostMalloc");
    return p;
}
static void host_release(void *p)
{
    if (p) (void)hipHostFree(p);
}
static Stream make_stream()
{
    hipStream_t s = nullptr;
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), FORT_SCOPE_RESOURCE, "hipStreamCreate");
    return s;
}
// The first large linear copy of a process and the first pitched one each set up their copy path,
// which took 6-7 ms apiece on an MI355X (small copies take another path and do not). Copies of `bytes`
// each way between the staging buffer `pinned` and a device block, linear from aligned and unaligned
// addresses and pitched, move that cost to the context's creation, so that the first upload and the
// first observation of an epoch do not pay it.
static void prime(Stream s, void *pinned, std::size_t bytes)
{
    void *d = nullptr;
    check(hipMalloc(&d, bytes + 64), FORT_SCOPE_RESOURCE, "hipMalloc");
    for (std::size_t skew : {std::size_t(0), std::size_t(8)}) {
        char *dd = static_cast<char *>(d) + skew, *hh = static_cast<char *>(pinned) + skew;
        check(hipMemcpyAsync(dd, hh, bytes, hipMemcpyHostToDevice, s), FORT_SCOPE_RESOURCE, "hipMemcpyAsync");
        check(hipMemcpyAsync(hh, dd, bytes, hipMemcpyDeviceToHost, s), FORT_SCOPE_RESOURCE, "hipMemcpyAsync");
    }
    const std::size_t width = 512, pitch = 1024, rows = bytes / pitch;
    check(hipMemcpy2DAsync(d, pitch, pinned, width, width, rows, hipMemcpyHostToDevice, s), FORT_SCOPE_RESOURCE,
          "hipMemcpy2DAsync");
    check(hipMemcpy2DAsync(pinned, width, d, pitch, width, rows, hipMemcpyDeviceToHost, s), FORT_SCOPE_RESOURCE,
          "hipMemcpy2DAsync");
    check(hipStreamSynchronize(s), FORT_SCOPE_RESOURCE, "hipStreamSynchronize");
    (void)hipFree(d);
}
static void destroy_stream(Stream s)
{
    if (s) (void)hipStreamDestroy(s);
}
static std::size_t free_bytes()
{
    std::size_t f = 0, t = 0;
    check(hipMemGetInfo(&f, &t), FORT_SCOPE_RESOURCE, "hipMemGetInfo");
    return f;
}
static std::size_t total_bytes()
{
    std::size_t f = 0, t = 0;
    check(hipMemGetInfo(&f, &t), FORT_SCOPE_RESOURCE, "hipMemGetInfo");
    return t;
}
#endif
}  // namespace dev

// ------------------------------------------------------------------------------------------------ state
struct Domain;
struct Context;

struct Buffer {
    fort_buffer_t handle = 0;
    Domain *domain = nullptr;
    uint64_t identity = 0, generation = 0;
    uint32_t rank = 0, type = 0;
    std::size_t elem = 0;
    void *host = nullptr;
    std::size_t extents[kMaxRank] = {};
    int64_t lower_bounds[kMaxRank] = {};
    std::size_t count = 0, bytes = 0;
    void *device = nullptr;
    Region init, host_cur, dev_cur;
    // A prepared access: 0 none, 1 device, 2 host; with the sections it may change.
    int prepared = 0;
    bool prep_original = false;  // the prepared host access is original host work
    std::vector<Box> prep_writes, prep_overwrites;
    uint64_t last_use = 0;
    int queued = 0;  // accepted tasks not yet submitted that use this buffer

    Box full() const
    {
        Box b;
        b.rank = static_cast<int>(rank);
        for (uint32_t d = 0; d < rank; ++d) b.hi[d] = extents[d];
        return b;
    }
};

struct Domain {
    fort_scope_domain_t handle = 0;
    int device = 0;
    std::size_t budget = 0;
    bool budget_known = false;
    bool budget_warned = false;  // the notice of a task that does not fit was printed
    std::map<fort_buffer_t, std::unique_ptr<Buffer>> buffers;
    std::map<uint64_t, fort_buffer_t> by_identity;
    Context *borrower = nullptr;
    bool is_private = false;
    bool failed = false;
    std::size_t device_bytes = 0, workspace_bytes = 0;
    std::map<uint64_t, std::size_t> workspace;
    uint64_t next_lease = 1;
    uint64_t clock = 0;
    fort_scope_stats stats{};
    fort_scope_residency_stats_v1 residency{};
    std::mutex mu;
};

#ifndef FTC_CPU_BACKEND
// External work captured from the stream its worker was given: one child-graph node of the chain.
struct CapturedWork {
    hipGraph_t graph = nullptr;
    CapturedWork() = default;
    CapturedWork(const CapturedWork &) = delete;
    CapturedWork &operator=(const CapturedWork &) = delete;
    ~CapturedWork()
    {
        if (graph) (void)hipGraphDestroy(graph);
    }
};
#endif

struct Task {
    uint64_t code_key = 0;
    const void *function = nullptr;
    uint32_t grid[3] = {}, block[3] = {};
    std::size_t shared = 0;
    std::vector<unsigned char> args;
    std::vector<std::size_t> offsets, sizes;
    std::vector<Buffer *> buffers;
    uint64_t threads = 0;
#ifndef FTC_CPU_BACKEND
    std::shared_ptr<CapturedWork> captured;  // set: the task is this captured graph, not a kernel
#endif
};

#ifndef FTC_CPU_BACKEND
// One instantiated graph of a topology, with the argument values and launch shapes it holds.
struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
    std::vector<std::vector<unsigned char>> args;
    std::vector<std::vector<uint32_t>> dims;
    uint64_t last_use = 0;
};
// A topology (the ordered kernels and their argument sizes) and up to kGraphVariants instantiations of
// it, so that a step alternating between buffer pairs replays without parameter updates.
constexpr std::size_t kGraphVariants = 4;
struct GraphFamily {
    uint64_t seen = 0, clock = 0;
    std::vector<GraphEntry> variants;
};
#endif

struct Context {
    fort_scope_t handle = 0;
    Domain *domain = nullptr;
    bool owns_domain = false;
    dev::Stream stream = nullptr;
    bool stream_ready = false;
    bool poisoned = false;
    std::vector<Task> queue;
    uint64_t queued_threads = 0;
    bool queue_enabled = true, graphs_enabled = true;
    uint64_t queue_work = uint64_t(1) << 18;
    fort_scope_graph_stats_v1 gstats{};
    fort_scope_graph_queue_stats_v1 qstats{};
#ifndef FTC_CPU_BACKEND
    std::unordered_map<uint64_t, GraphFamily> graphs;
    std::unordered_map<uint64_t, GraphEntry> groups;  // graph_submit_v1 / graph_execute_v1, by group key
    // External workers run under stream capture on this stream when work is held for whole-step graphs.
    hipStream_t capture_stream = nullptr;
    // Executables of captured work launched on their own, destroyed once the stream has drained.
    std::vector<hipGraphExec_t> retired;
#endif
    void *scratch = nullptr;
    std::size_t scratch_capacity = 0;
    // Pinned staging of strided section copies; `staging_queued` while an upload from it may still run.
    void *staging = nullptr;
    std::size_t staging_capacity = 0;
    bool staging_queued = false;
    uint64_t scratch_token = 0, next_token = 1;
    bool scratch_active = false;
    fort_scope_scratch_stats_v1 sstats{};
    fort_scope_completion_stats_v1 cstats{};
    fort_scope_trace_state_v1 trace{};
};

static std::mutex g_registry;
static std::map<fort_scope_t, std::unique_ptr<Context>> g_contexts;
static std::map<fort_scope_domain_t, std::unique_ptr<Domain>> g_domains;
static std::atomic<uint64_t> g_next_handle{0x5c0e000000000001ull};

static uint64_t fresh_handle() { return g_next_handle.fetch_add(0x9e37ull); }

static uint64_t env_u64(const char *name, uint64_t fallback)
{
    const char *v = std::getenv(name);
    if (!v || !*v) return fallback;
    char *end = nullptr;
    unsigned long long x = std::strtoull(v, &end, 10);
    return end && *end == 0 ? x : fallback;
}

// Parses "123", "64M", "2G" (binary multiples).
static bool parse_bytes(const char *v, std::size_t *out)
{
    if (!v || !*v) return false;
    char *end = nullptr;
    unsigned long long x = std::strtoull(v, &end, 10);
    if (end == v) return false;
    unsigned shift = 0;
    if (*end == 'K' || *end == 'k') shift = 10, ++end;
    else if (*end == 'M' || *end == 'm') shift = 20, ++end;
    else if (*end == 'G' || *end == 'g') shift = 30, ++end;
    if (*end) return false;
    *out = static_cast<std::size_t>(x) << shift;
    return true;
}

// ------------------------------------------------------------------------------------------------ lookup
static Context &context(fort_scope_t h)
{
    auto it = g_contexts.find(h);
    require(it != g_contexts.end(), FORT_SCOPE_STALE, "unknown or closed scoped context");
    return *it->second;
}

static Domain &domain_of(fort_scope_domain_t h)
{
    auto it = g_domains.find(h);
    require(it != g_domains.end(), FORT_SCOPE_STALE, "unknown or closed residency domain");
    return *it->second;
}

static Buffer &buffer(Context &c, fort_buffer_t h)
{
    auto it = c.domain->buffers.find(h);
    require(it != c.domain->buffers.end(), FORT_SCOPE_STALE, "unknown or retired buffer handle");
    return *it->second;
}

// ------------------------------------------------------------------------------------------------ copies
// Column-major strides of a buffer in elements.
static void strides(const Buffer &b, std::size_t *s)
{
    s[0] = 1;
    for (uint32_t d = 1; d < b.rank; ++d) s[d] = s[d - 1] * b.extents[d - 1];
}

// Copies one box between the host array and its mirror: the leading dimensions the box spans completely
// form one contiguous run, the next dimension is the row count of a pitched 2D copy, and the remaining
// dimensions are iterated (compiler/runtime/section_copy.hpp:14-80 plans the same decomposition).
static void ensure_stream(Context &c);

static void copy_box(Context &c, Buffer &b, const Box &box, bool h2d)
{
    if (box.empty()) return;
    std::size_t s[kMaxRank];
    strides(b, s);
    int r = static_cast<int>(b.rank);
    int m = 0;  // dims [0, m) are spanned completely
    while (m < r - 1 && box.lo[m] == 0 && box.hi[m] == b.extents[m]) ++m;
    std::size_t run = (box.hi[m] - box.lo[m]) * s[m] * b.elem;
    int row_dim = m + 1;
    std::size_t rows = row_dim < r ? box.hi[row_dim] - box.lo[row_dim] : 1;
    std::size_t pitch = row_dim < r ? s[row_dim] * b.elem : run;
    std::size_t planes = 1;
    for (int d = row_dim + 1; d < r; ++d) planes *= box.hi[d] - box.lo[d];
    const std::size_t packed = run * rows;  // one plane of the box, rows back to back
    const std::size_t bytes = box.volume() * b.elem;
    auto count = [&]() {
        if (h2d) {
            c.domain->stats.uploads++;
            c.domain->stats.upload_bytes += bytes;
        } else {
            c.domain->stats.downloads++;
            c.domain->stats.download_bytes += bytes;
        }
    };
    auto offset_of = [&](const std::size_t *at) {
        std::size_t off = 0;
        for (int d = 0; d < r; ++d) off += at[d] * s[d];
        return off * b.elem;
    };
    auto ensure_staging = [&](std::size_t need) {
        if (c.staging_capacity >= need) return;
        ensure_stream(c);
        if (c.staging_queued) dev::sync(c.stream);
        c.staging_queued = false;
        dev::host_release(c.staging);
        c.staging = nullptr;
        c.staging_capacity = 0;
        c.staging = dev::host_alloc(need);
        c.staging_capacity = need;
    };
    // One contiguous run goes directly; anything strided is staged (HIP backend).
    const bool staged = dev::kStaged && (rows > 1 || planes > 1);
    std::size_t idx[kMaxRank];
    for (int d = 0; d < r; ++d) idx[d] = box.lo[d];
    auto next_plane = [&]() {
        int d = row_dim + 1;
        while (d < r) {
            if (++idx[d] < box.hi[d]) return true;
            idx[d] = box.lo[d];
            ++d;
        }
        return false;
    };
    if (staged && !h2d) {
        // A download of a dense box copies its whole span in one transfer and unpacks its rows on the
        // host; the cells between the rows come along unused.
        std::size_t last[kMaxRank];
        for (int d = 0; d < r; ++d) last[d] = box.hi[d] - 1;
        // The span starts and ends on 256-byte boundaries of the allocation where it can (the copy
        // engine's fast path), clipped to the allocation.
        const std::size_t first = offset_of(box.lo) & ~std::size_t(255);
        const std::size_t end = std::min(b.bytes, (offset_of(last) + b.elem + 255) & ~std::size_t(255));
        const std::size_t span = end - first;
        if (span <= 2 * bytes + 512 && span <= c.staging_capacity) {
            ensure_staging(span);
            char *stage = static_cast<char *>(c.staging);
            dev::copy2d(stage, span, static_cast<char *>(b.device) + first, span, span, 1, false, c.stream);
            dev::sync(c.stream);
            c.staging_queued = false;
            do {
                const std::size_t off = offset_of(idx);
                for (std::size_t q = 0; q < rows; ++q)
                    std::memcpy(static_cast<char *>(b.host) + off + q * pitch, stage + off - first + q * pitch, run);
            } while (next_plane());
            count();
            return;
        }
    }
    // Otherwise the box moves in batches of planes, each plane one pitched copy.
    std::size_t batch = planes;
    if (staged) {
        batch = std::max<std::size_t>(packed, std::min<std::size_t>(packed * planes, c.staging_capacity)) / packed;
        ensure_staging(batch * packed);
    }
    std::vector<std::size_t> offsets;  // byte offsets of the planes of one batch
    auto flush_batch = [&]() {
        if (offsets.empty()) return;
        char *stage = static_cast<char *>(c.staging);
        if (h2d) {
            if (c.staging_queued) dev::sync(c.stream);  // an earlier upload may still read the buffer
            for (std::size_t p = 0; p < offsets.size(); ++p)
                for (std::size_t q = 0; q < rows; ++q)
                    std::memcpy(stage + p * packed + q * run, static_cast<char *>(b.host) + offsets[p] + q * pitch, run);
            for (std::size_t p = 0; p < offsets.size(); ++p)
                dev::copy2d(static_cast<char *>(b.device) + offsets[p], pitch, stage + p * packed, run, run, rows, true,
                            c.stream);
            c.staging_queued = true;
        } else {
            for (std::size_t p = 0; p < offsets.size(); ++p)
                dev::copy2d(stage + p * packed, run, static_cast<char *>(b.device) + offsets[p], pitch, run, rows,
                            false, c.stream);
            dev::sync(c.stream);
            c.staging_queued = false;
            for (std::size_t p = 0; p < offsets.size(); ++p)
                for (std::size_t q = 0; q < rows; ++q)
                    std::memcpy(static_cast<char *>(b.host) + offsets[p] + q * pitch, stage + p * packed + q * run, run);
        }
        offsets.clear();
    };
    do {
        const std::size_t off = offset_of(idx);
        if (staged) {
            offsets.push_back(off);
            if (offsets.size() == batch) flush_batch();
        } else {
            char *h = static_cast<char *>(b.host) + off;
            char *g = static_cast<char *>(b.device) + off;
            if (h2d) dev::copy2d(g, pitch, h, pitch, run, rows, true, c.stream);
            else dev::copy2d(h, pitch, g, pitch, run, rows, false, c.stream);
        }
    } while (next_plane());
    flush_batch();
    count();
}

// Pinned staging every context starts with: strided boxes up to this size move in one batch.
constexpr std::size_t kStagingBytes = std::size_t(16) << 20;

static void ensure_stream(Context &c)
{
    if (c.stream_ready) return;
    dev::set(c.domain->device);
    c.stream = dev::make_stream();
    c.stream_ready = true;
    if (dev::kStaged && !c.staging) {
        c.staging = dev::host_alloc(kStagingBytes);
        c.staging_capacity = kStagingBytes;
        dev::prime(c.stream, c.staging, std::size_t(4) << 20);
    }
}

static void wait(Context &c)
{
    if (!c.stream_ready) return;
    dev::sync(c.stream);
    c.domain->stats.waits++;
    c.cstats.stream_waits++;
#ifndef FTC_CPU_BACKEND
    for (hipGraphExec_t e : c.retired) (void)hipGraphExecDestroy(e);
    c.retired.clear();
#endif
}

// ------------------------------------------------------------------------------------------------ queue
static void flush(Context &c);

static void drain_buffer(Context &c, const Buffer &b)
{
    if (b.queued > 0) {
        c.qstats.conflict_flushes++;
        flush(c);
    }
}

// ------------------------------------------------------------------------------------------------ memory
static void evict_one(Context &c, Buffer &victim)
{
    // Publish the values only the device holds, then drop the mirror.
    std::vector<Box> dirty;
    for (const Box &box : victim.dev_cur.boxes)
        for (const Box &m : victim.host_cur.missing(box)) dirty.push_back(m);
    for (const Box &box : dirty) {
        copy_box(c, victim, box, false);
        c.domain->residency.dirty_download_bytes += box.volume() * victim.elem;
    }
    if (!dirty.empty()) {
        wait(c);
        c.domain->residency.dirty_evictions++;
        for (const Box &box : dirty) require(victim.host_cur.add(box), FORT_SCOPE_RESOURCE, "region limit");
    }
    dev::release(victim.device);
    victim.device = nullptr;
    victim.dev_cur.clear();
    c.domain->device_bytes -= victim.bytes;
    c.domain->stats.allocated_bytes -= victim.bytes;
    c.domain->residency.evictions++;
    c.domain->residency.evicted_bytes += victim.bytes;
}

// FORT_RESIDENT_DEVICE_BUDGET if set, else the free device memory less a reserve of one sixteenth of
// the device, at least 512 MiB (compiler/runtime/scoped_runtime.h:169-174). Called with the registry
// lock held or from the public wrapper; it takes no lock itself.
static std::size_t default_budget(int device)
{
    std::size_t parsed = 0;
    if (parse_bytes(std::getenv("FORT_RESIDENT_DEVICE_BUDGET"), &parsed)) return parsed;
    int previous = dev::current();
    dev::set(device);
    std::size_t free = dev::free_bytes(), total = dev::total_bytes();
    dev::set(previous);
    std::size_t reserve = std::max(total / 16, std::size_t(512) << 20);
    return free > reserve ? free - reserve : 0;
}

static std::size_t budget_of(Domain &d)
{
    if (!d.budget_known) {
        d.budget = default_budget(d.device);
        d.budget_known = true;
    }
    return d.budget;
}

// A device task whose arrays cannot be on the device together under the budget is refused with RESOURCE;
// an application's spliced block then runs the procedure's own loops on the host. The run is told once
// per domain that the budget costs it time (scoped_runtime.cu:534-543 says the same).
[[noreturn]] static void refuse_over_budget(Domain &d, std::size_t budget, const char *message)
{
    if (!d.budget_warned) {
        d.budget_warned = true;
        std::fprintf(stderr,
                     "Resident domain: the arrays in use exceed the device budget of %zu bytes; loop nests whose "
                     "arrays do not fit together run on the host.\n",
                     budget);
    }
    fail(FORT_SCOPE_RESOURCE, message);
}

// Gives every requested buffer a device allocation, evicting least-recently-used unrequested mirrors
// that no prepared access or queued task holds.
static void reserve(Context &c, Buffer *const *list, std::size_t n)
{
    Domain &d = *c.domain;
    d.residency.reserve_batches++;
    d.residency.requested_resources += n;
    std::size_t need = 0;
    for (std::size_t i = 0; i < n; ++i) {
        bool dup = false;
        for (std::size_t j = 0; j < i; ++j) dup |= list[j] == list[i];
        if (!dup && !list[i]->device) need += list[i]->bytes;
        if (list[i]->device) d.residency.reuse_hits++;
    }
    std::size_t budget = budget_of(d);
    std::size_t pinned = d.workspace_bytes + c.scratch_capacity;
    if (need > budget) refuse_over_budget(d, budget, "task footprint exceeds the device budget");
    if (d.device_bytes + pinned + need > budget && !c.queue.empty()) {
        d.residency.pressure_flushes++;
        flush(c);
    }
    while (d.device_bytes + pinned + need > budget) {
        Buffer *victim = nullptr;
        for (auto &kv : d.buffers) {
            Buffer *b = kv.second.get();
            if (!b->device || b->prepared || b->queued) continue;
            bool requested = false;
            for (std::size_t i = 0; i < n; ++i) requested |= list[i] == b;
            if (requested) continue;
            if (!victim || b->last_use < victim->last_use) victim = b;
        }
        if (!victim) refuse_over_budget(d, budget, "device budget exhausted by pinned buffers");
        evict_one(c, *victim);
    }
    for (std::size_t i = 0; i < n; ++i) {
        Buffer &b = *list[i];
        b.last_use = ++d.clock;
        if (b.device) continue;
        dev::set(d.device);
        b.device = dev::alloc(b.bytes);
        require(b.device != nullptr, FORT_SCOPE_RESOURCE, "device allocation failed");
        d.device_bytes += b.bytes;
        d.stats.allocations++;
        d.stats.allocated_bytes += b.bytes;
        std::size_t total = d.device_bytes + d.workspace_bytes + c.scratch_capacity;
        if (total > d.stats.peak_device_bytes) d.stats.peak_device_bytes = total;
    }
}

// ------------------------------------------------------------------------------------------------ access
struct Effects {
    std::vector<Box> reads, writes, overwrites;
};

static Box section_box(const Buffer &b, const fort_scope_section &s)
{
    require(s.lower && s.upper, FORT_SCOPE_ARGUMENT, "section without bounds");
    Box box;
    box.rank = static_cast<int>(b.rank);
    for (uint32_t d = 0; d < b.rank; ++d) {
        require(s.lower[d] <= s.upper[d] && s.upper[d] <= b.extents[d], FORT_SCOPE_ARGUMENT,
                "section outside the registered layout");
        box.lo[d] = s.lower[d];
        box.hi[d] = s.upper[d];
    }
    return box;
}

static Effects effects(const Buffer &b, const fort_scope_access *a)
{
    Effects e;
    require(a != nullptr, FORT_SCOPE_ARGUMENT, "missing access description");
    require((a->flags & ~7u) == 0, FORT_SCOPE_ARGUMENT, "unknown access flags");
    if (a->flags & FORT_SCOPE_READ_ALL) e.reads.push_back(b.full());
    if (a->flags & FORT_SCOPE_WRITE_ALL) e.writes.push_back(b.full());
    if (a->flags & FORT_SCOPE_OVERWRITE_ALL) e.overwrites.push_back(b.full());
    require(!a->read_count || a->reads, FORT_SCOPE_ARGUMENT, "missing read sections");
    require(!a->write_count || a->writes, FORT_SCOPE_ARGUMENT, "missing write sections");
    require(!a->overwrite_count || a->overwrites, FORT_SCOPE_ARGUMENT, "missing overwrite sections");
    // Empty sections name nothing and are dropped (compiler/runtime/scoped_runtime.cu:804).
    auto add = [&](std::vector<Box> &list, const fort_scope_section &s) {
        const Box box = section_box(b, s);
        if (!box.empty()) list.push_back(box);
    };
    for (std::size_t i = 0; i < a->read_count; ++i) add(e.reads, a->reads[i]);
    for (std::size_t i = 0; i < a->write_count; ++i) add(e.writes, a->writes[i]);
    for (std::size_t i = 0; i < a->overwrite_count; ++i) add(e.overwrites, a->overwrites[i]);
    // What an access must overwrite it may write (compiler/runtime/scoped_runtime.cu:816-817).
    Region may;
    for (const Box &w : e.writes) require(may.add(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
    for (const Box &o : e.overwrites)
        require(may.covers(o), FORT_SCOPE_ARGUMENT, "must-overwrite exceeds may-write sections");
    return e;
}

// The values a buffer physically holds: what is defined, and whatever either copy holds current. A
// definition kill (fort_scope_forget_definition) leaves the currency, so values written before it are
// still published (compiler/runtime/scoped_regions.hpp:189-196).
static Region physical(const Buffer &b)
{
    Region p = b.init;
    for (const Region *r : {&b.host_cur, &b.dev_cur})
        for (const Box &x : r->boxes) require(p.add(x), FORT_SCOPE_RESOURCE, "region limit exceeded");
    return p;
}

// The sections of b to copy to the device (device) or to the host before an access with effects e: its
// reads and the parts of its writes it does not overwrite (a partial write leaves the rest of its region
// as it was), where they are not current there. A strict access must find all of them defined
// (compiler/runtime/scoped_runtime.cu:906-929); the may-access of original host work (strict false) may
// cover holes, and copies only the values the buffer physically holds.
static std::vector<Box> plan_transfer(const Buffer &b, const Effects &e, bool device, bool strict = true)
{
    std::vector<Box> wanted = e.reads;
    Region over;
    for (const Box &o : e.overwrites) require(over.add(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
    for (const Box &w : e.writes)
        for (const Box &kept : over.missing(w)) wanted.push_back(kept);
    if (strict)
        for (const Box &r : wanted)
            require(b.init.covers(r), FORT_SCOPE_UNINITIALIZED,
                    "task reads values that were never defined");
    const Region known = strict ? b.init : physical(b);
    const Region &have = device ? b.dev_cur : b.host_cur;
    std::vector<Box> need;
    Region acc;
    for (const Box &box : wanted)
        for (const Box &held : known.within(box))
            for (const Box &m : have.missing(held))
                for (const Box &fresh : acc.missing(m)) {
                    need.push_back(fresh);
                    require(acc.add(fresh), FORT_SCOPE_RESOURCE, "region limit exceeded");
                }
    return need;
}

// Commits the effects of a finished (or accepted) access to the coherence state: what it may write is
// defined (its partial writes were defined before) and current on its side only
// (compiler/runtime/scoped_regions.hpp:219-225). Original host work defines and makes current only what
// it proves to overwrite, and takes the device copy of everything it may write as outdated (:230-236).
static void commit(Buffer &b, const std::vector<Box> &writes, const std::vector<Box> &overwrites, bool device,
                   bool original = false)
{
    Region &mine = device ? b.dev_cur : b.host_cur;
    Region &other = device ? b.host_cur : b.dev_cur;
    for (const Box &o : overwrites) require(b.init.add(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
    for (const Box &w : writes) {
        if (!original) {
            require(b.init.add(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
            require(mine.add(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
        }
        require(other.remove(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
    }
    for (const Box &o : overwrites) {
        require(mine.add(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
        require(other.remove(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
    }
}

// Commit with the reference's reconciliation (compiler/runtime/scoped_runtime.cu:930-950) for a buffer
// whose regions would need more rectangles than the limit: the committing side first takes every value
// only the other side holds (outside what the access wrote), then holds all physical values alone, and the
// other copy, now redundant, is dropped. The regions are restored first if even that cannot be written.
static void commit_or_reconcile(Context &c, Buffer &b, const std::vector<Box> &writes,
                                const std::vector<Box> &overwrites, bool device, bool original = false)
{
    const Region init = b.init, hc = b.host_cur, dc = b.dev_cur;
    auto restore = [&] {
        b.init = init;
        b.host_cur = hc;
        b.dev_cur = dc;
    };
    try {
        commit(b, writes, overwrites, device, original);
        return;
    } catch (const Error &e) {
        restore();
        if (e.status != FORT_SCOPE_RESOURCE || original) throw;
    }
    Region written;
    for (const std::vector<Box> *list : {&writes, &overwrites})
        for (const Box &w : *list) require(written.add(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
    const Region &mine = device ? dc : hc, &other = device ? hc : dc;
    std::vector<Box> pieces;
    for (const Box &o : other.boxes)
        for (const Box &m : mine.missing(o))
            for (const Box &x : written.missing(m)) pieces.push_back(x);
    Region all = init, defined = init;
    try {
        for (const Region *r : {&hc, &dc, static_cast<const Region *>(&written)})
            for (const Box &x : r->boxes) require(all.add(x), FORT_SCOPE_RESOURCE, "region limit exceeded");
        for (const Box &x : written.boxes) require(defined.add(x), FORT_SCOPE_RESOURCE, "region limit exceeded");
    } catch (const Error &) {
        fail(FORT_SCOPE_RESOURCE, "the buffer's regions exceed the rectangle limit even after reconciliation");
    }
    if (!pieces.empty()) {
        ensure_stream(c);
        for (const Box &x : pieces) copy_box(c, b, x, device);
        if (!device) wait(c);
    }
    b.init = defined;
    (device ? b.dev_cur : b.host_cur) = all;
    (device ? b.host_cur : b.dev_cur).clear();
    c.domain->stats.reconciliations++;
}

// Before an access begins: its commit must be representable, directly or after a reconciliation, or the
// access is refused with BOUNDARY while nothing has changed, so the caller can still run its original
// statements (compiler/runtime/scoped_runtime.cu:899-954, the preparation simulated on private state;
// test_scoped_runtime.py:341-357).
static void check_commit(const Buffer &b, const Effects &e, bool device, bool original = false)
{
    Buffer t = b;
    try {
        commit(t, e.writes, e.overwrites, device, original);
        return;
    } catch (const Error &err) {
        if (err.status != FORT_SCOPE_RESOURCE) throw;
    }
    bool fits = !original;
    Region all = b.init, defined = b.init;
    for (const Region *r : {&b.host_cur, &b.dev_cur})
        for (const Box &x : r->boxes) fits = fits && all.add(x);
    for (const std::vector<Box> *list : {&e.writes, &e.overwrites})
        for (const Box &x : *list) fits = fits && all.add(x) && defined.add(x);
    require(fits, FORT_SCOPE_BOUNDARY, "section fragmentation requires a scope boundary");
}

// An access's whole acquisition, planned on private state before anything moves
// (compiler/runtime/scoped_runtime.cu:899-954): the boxes to copy to the acquiring side and that side's
// currency once they are there, checked so that the access's commit then fits, directly or after a
// reconciliation. When the transfers themselves would split the side's currency past the rectangle limit,
// the side takes every value only the other side holds instead and then holds all physical values. What
// cannot be represented even so refuses the access with BOUNDARY while nothing has changed.
struct Acquisition {
    std::vector<Box> copies;
    Region side;
    bool reconciled = false;
};

static Acquisition plan_acquisition(const Buffer &b, const Effects &e, bool device, bool original)
{
    const Region &mine = device ? b.dev_cur : b.host_cur;
    const Region &other = device ? b.host_cur : b.dev_cur;
    Acquisition a;
    bool fragmented = false;
    try {
        a.copies = plan_transfer(b, e, device, !original);
        a.side = mine;
        for (const Box &x : a.copies)
            if (!a.side.add(x)) {
                fragmented = true;
                break;
            }
    } catch (const Error &err) {
        if (err.status != FORT_SCOPE_RESOURCE) throw;
        fragmented = true;
    }
    if (fragmented) {
        if (!original)
            for (const Box &r : e.reads)
                require(b.init.covers(r), FORT_SCOPE_UNINITIALIZED, "task reads values that were never defined");
        a.copies.clear();
        for (const Box &o : other.boxes)
            for (const Box &m : mine.missing(o)) a.copies.push_back(m);
        try {
            a.side = physical(b);
        } catch (const Error &err) {
            if (err.status != FORT_SCOPE_RESOURCE) throw;
            fail(FORT_SCOPE_BOUNDARY, "section fragmentation requires a scope boundary");
        }
        a.reconciled = true;
    }
    Buffer t = b;
    (device ? t.dev_cur : t.host_cur) = a.side;
    check_commit(t, e, device, original);
    return a;
}

// Carries out a planned acquisition: the copies, then the side's currency as planned. Host copies are
// complete on return.
static void acquire(Context &c, Buffer &b, const Acquisition &a, bool device)
{
    if (!a.copies.empty()) {
        require(b.device != nullptr || device, FORT_SCOPE_STATE, "defined values are on no copy");
        ensure_stream(c);
        for (const Box &box : a.copies) copy_box(c, b, box, device);
        if (!device) wait(c);
    }
    (device ? b.dev_cur : b.host_cur) = a.side;
    if (a.reconciled) c.domain->stats.reconciliations++;
}

static void prepare_device(Context &c, const fort_scope_plan_binding *bindings, std::size_t count, void **devices,
                           bool commit_now)
{
    require(!count || bindings, FORT_SCOPE_ARGUMENT, "missing task access bindings");
    require(!count || devices, FORT_SCOPE_ARGUMENT, "missing device pointer outputs");
    std::vector<Buffer *> list;
    std::vector<Effects> eff;
    for (std::size_t i = 0; i < count; ++i) {
        Buffer &b = buffer(c, bindings[i].buffer);
        for (Buffer *o : list) require(o != &b, FORT_SCOPE_ALIAS, "task access bindings repeat a canonical buffer");
        require(!b.prepared, FORT_SCOPE_STATE, "buffer already has a prepared access");
        list.push_back(&b);
        eff.push_back(effects(b, &bindings[i].access));
    }
    // Every binding is planned before any buffer changes, so a later refusal leaves the earlier ones as
    // they were.
    std::vector<Acquisition> plans;
    for (std::size_t i = 0; i < count; ++i) plans.push_back(plan_acquisition(*list[i], eff[i], true, false));
    if (!count) return;
    ensure_stream(c);
    reserve(c, list.data(), list.size());
    for (std::size_t i = 0; i < count; ++i) {
        Buffer &b = *list[i];
        acquire(c, b, plans[i], true);
        devices[i] = b.device;
        if (commit_now) {
            commit_or_reconcile(c, b, eff[i].writes, eff[i].overwrites, true);
        } else {
            b.prepared = 1;
            b.prep_writes = eff[i].writes;
            b.prep_overwrites = eff[i].overwrites;
        }
    }
}

static void end_access(Context &c, Buffer &b, bool device, bool cancel)
{
    require(b.prepared == (device ? 1 : 2), FORT_SCOPE_STATE, "no prepared access of this kind on the buffer");
    // The access is over whatever the commit does: a failure leaves no prepared access behind.
    struct Clear {
        Buffer &b;
        ~Clear()
        {
            b.prepared = 0;
            b.prep_original = false;
            b.prep_writes.clear();
            b.prep_overwrites.clear();
        }
    } clear{b};
    if (!cancel) commit_or_reconcile(c, b, b.prep_writes, b.prep_overwrites, device, !device && b.prep_original);
    if (!cancel && device) c.domain->stats.launches++;
}

static void download(Context &c, Buffer &b, const std::vector<Box> &boxes)
{
    if (boxes.empty()) return;
    require(b.device != nullptr, FORT_SCOPE_STATE, "defined values are on no copy");
    for (const Box &box : boxes) copy_box(c, b, box, false);
    wait(c);
    for (const Box &box : boxes) require(b.host_cur.add(box), FORT_SCOPE_RESOURCE, "region limit exceeded");
}

// The values the buffer physically holds and the host does not: its defined values less what the host
// holds, then the device's current values outside both (compiler/runtime/scoped_runtime.cu:1410-1413).
// The pieces are disjoint and come from subtractions only, so no union has to fit the region limit.
static std::vector<Box> unpublished(const Buffer &b)
{
    std::vector<Box> need;
    for (const Box &i : b.init.boxes)
        for (const Box &m : b.host_cur.missing(i)) need.push_back(m);
    for (const Box &d : b.dev_cur.boxes)
        for (const Box &m : b.host_cur.missing(d))
            for (const Box &x : b.init.missing(m)) need.push_back(x);
    return need;
}

// Publication: every value the buffer physically holds reaches the host.
static void observe(Context &c, Buffer &b)
{
    drain_buffer(c, b);
    download(c, b, unpublished(b));
}

// ------------------------------------------------------------------------------------------------ graphs
static uint64_t fnv(uint64_t h, const void *p, std::size_t n)
{
    const unsigned char *s = static_cast<const unsigned char *>(p);
    for (std::size_t i = 0; i < n; ++i) {
        h ^= s[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

static void launch_direct(Context &c, Task &t)
{
#ifdef FTC_CPU_BACKEND
    (void)c;
    (void)t;
    fail(FORT_SCOPE_RESOURCE, "the CPU reference backend executes no device kernels");
#else
    if (t.captured) {
        hipGraphExec_t exec = nullptr;
        hipError_t e = hipGraphInstantiate(&exec, t.captured->graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, c.stream);
        if (exec) c.retired.push_back(exec);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            c.poisoned = true;
            fail(FORT_SCOPE_EXECUTION, std::string("captured work launch failed: ") + hipGetErrorString(e));
        }
        return;
    }
    std::vector<void *> params(t.offsets.size());
    for (std::size_t i = 0; i < t.offsets.size(); ++i) params[i] = t.args.data() + t.offsets[i];
    hipError_t e = hipLaunchKernel(t.function, dim3(t.grid[0], t.grid[1], t.grid[2]),
                                   dim3(t.block[0], t.block[1], t.block[2]), params.data(), t.shared, c.stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c.poisoned = true;
        fail(FORT_SCOPE_EXECUTION, std::string("kernel launch failed: ") + hipGetErrorString(e));
    }
#endif
}

#ifndef FTC_CPU_BACKEND
static hipKernelNodeParams node_params(Task &t, std::vector<void *> &params)
{
    params.resize(t.offsets.size());
    for (std::size_t i = 0; i < t.offsets.size(); ++i) params[i] = t.args.data() + t.offsets[i];
    hipKernelNodeParams p{};
    p.blockDim = dim3(t.block[0], t.block[1], t.block[2]);
    p.gridDim = dim3(t.grid[0], t.grid[1], t.grid[2]);
    p.func = const_cast<void *>(t.function);
    p.kernelParams = params.data();
    p.sharedMemBytes = static_cast<unsigned>(t.shared);
    p.extra = nullptr;
    return p;
}

static void destroy_graph(Context &c, GraphEntry &g)
{
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    if (g.exec) c.gstats.destructions++;
    g.exec = nullptr;
    g.graph = nullptr;
    g.nodes.clear();
}

// Builds a chain graph (each node depends on its predecessor, so kernel boundaries stay ordered) and
// instantiates it. Returns false if a graph resource could not be created; nothing was launched then.
static bool build_graph(Context &c, GraphEntry &g, std::vector<Task> &tasks)
{
    destroy_graph(c, g);
    if (hipGraphCreate(&g.graph, 0) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    hipGraphNode_t prev = nullptr;
    for (Task &t : tasks) {
        std::vector<void *> params;
        hipGraphNode_t node = nullptr;
        hipError_t e;
        if (t.captured) {
            e = hipGraphAddChildGraphNode(&node, g.graph, prev ? &prev : nullptr, prev ? 1 : 0, t.captured->graph);
        } else {
            hipKernelNodeParams p = node_params(t, params);
            e = hipGraphAddKernelNode(&node, g.graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            destroy_graph(c, g);
            return false;
        }
        g.nodes.push_back(node);
        prev = node;
    }
    if (hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        destroy_graph(c, g);
        return false;
    }
    c.gstats.creations++;
    c.gstats.instantiations++;
    g.args.clear();
    g.dims.clear();
    for (Task &t : tasks) {
        g.args.push_back(t.args);
        g.dims.push_back({t.grid[0], t.grid[1], t.grid[2], t.block[0], t.block[1], t.block[2],
                          static_cast<uint32_t>(t.shared)});
    }
    return true;
}
#endif

// Submits the queued tasks in order: the first time a topology is met its kernels launch directly; from
// the second time on it runs as one instantiated chain graph. A topology keeps up to kGraphVariants
// instantiations with different argument values (a step alternating between two buffer sets replays two
// graphs with no update); past that the least recently used one has its changed node parameters updated
// in place. (compiler/runtime/scoped_graphs.hpp:134-247 warms up, builds and replays in the same order.)
static void flush(Context &c)
{
    if (c.queue.empty()) return;
    std::vector<Task> tasks;
    tasks.swap(c.queue);
    c.queued_threads = 0;
    for (Task &t : tasks)
        for (Buffer *b : t.buffers) b->queued--;
    c.qstats.pending_tasks = 0;
    c.qstats.submissions++;
    c.gstats.submissions++;
    c.gstats.logical_kernels += tasks.size();
#ifdef FTC_CPU_BACKEND
    launch_direct(c, tasks.front());
#else
    dev::set(c.domain->device);
    if (!c.graphs_enabled) {
        for (Task &t : tasks) launch_direct(c, t);
        c.qstats.direct_submissions++;
        c.qstats.direct_kernels += tasks.size();
        c.domain->stats.launches += tasks.size();
        return;
    }
    uint64_t key = 0xcbf29ce484222325ull;
    for (Task &t : tasks) {
        key = fnv(key, &t.code_key, sizeof t.code_key);
        key = fnv(key, &t.function, sizeof t.function);
        key = fnv(key, t.sizes.data(), t.sizes.size() * sizeof(std::size_t));
    }
    GraphFamily &family = c.graphs[key];
    family.seen++;
    if (family.seen == 1) {
        for (Task &t : tasks) launch_direct(c, t);
        c.qstats.direct_submissions++;
        c.qstats.direct_kernels += tasks.size();
        c.domain->stats.launches += tasks.size();
        return;
    }
    auto dims_of = [](const Task &t) {
        return std::vector<uint32_t>{t.grid[0], t.grid[1], t.grid[2], t.block[0], t.block[1], t.block[2],
                                     static_cast<uint32_t>(t.shared)};
    };
    GraphEntry *g = nullptr;
    bool fresh = false;
    for (GraphEntry &v : family.variants) {
        bool same = true;
        for (std::size_t i = 0; same && i < tasks.size(); ++i)
            same = v.args[i] == tasks[i].args && v.dims[i] == dims_of(tasks[i]);
        if (same) {
            g = &v;
            break;
        }
    }
    if (g) {
        c.gstats.replays++;
    } else if (family.variants.size() < kGraphVariants) {
        family.variants.emplace_back();
        g = &family.variants.back();
        if (!build_graph(c, *g, tasks)) {
            family.variants.pop_back();
            c.qstats.graph_resource_fallbacks++;
            for (Task &t : tasks) launch_direct(c, t);
            c.domain->stats.launches += tasks.size();
            return;
        }
        c.gstats.live_graphs++;
        fresh = true;
    } else {
        // Every variant slot is taken: update the least recently used one in place.
        g = &family.variants.front();
        for (GraphEntry &v : family.variants)
            if (v.last_use < g->last_use) g = &v;
        for (std::size_t i = 0; i < tasks.size(); ++i) {
            Task &t = tasks[i];
            std::vector<uint32_t> dims = dims_of(t);
            if (t.captured || (g->args[i] == t.args && g->dims[i] == dims)) continue;
            std::vector<void *> params;
            hipKernelNodeParams p = node_params(t, params);
            hipError_t e = hipGraphExecKernelNodeSetParams(g->exec, g->nodes[i], &p);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                c.poisoned = true;
                fail(FORT_SCOPE_EXECUTION, std::string("graph parameter update failed: ") + hipGetErrorString(e));
            }
            g->args[i] = t.args;
            g->dims[i] = dims;
            c.gstats.parameter_updates++;
        }
        c.gstats.replays++;
    }
    if (!fresh) {
        // Captured work is captured anew each time; its node takes the new capture's parameters.
        for (std::size_t i = 0; i < tasks.size(); ++i) {
            if (!tasks[i].captured) continue;
            hipError_t e = hipGraphExecChildGraphNodeSetParams(g->exec, g->nodes[i], tasks[i].captured->graph);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                c.poisoned = true;
                fail(FORT_SCOPE_EXECUTION, std::string("captured work update failed: ") + hipGetErrorString(e));
            }
            c.gstats.parameter_updates++;
        }
    }
    g->last_use = ++family.clock;
    hipError_t e = hipGraphLaunch(g->exec, c.stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c.poisoned = true;
        fail(FORT_SCOPE_EXECUTION, std::string("graph launch failed: ") + hipGetErrorString(e));
    }
    c.domain->stats.launches += tasks.size();
#endif
}

// ------------------------------------------------------------------------------------------------ guard
enum class Mode { Change, Query };

// Runs one ABI call on a live context: saves and restores the caller's floating-point environment,
// converts errors to status codes and refuses work on a poisoned context.
template <class F>
static int with(fort_scope_t h, F &&body, Mode mode = Mode::Change)
{
    std::fenv_t env;
    std::fegetenv(&env);
    int status = FORT_SCOPE_OK;
    try {
        std::lock_guard<std::mutex> lock(g_registry);
        Context &c = context(h);
        if (mode == Mode::Change && c.poisoned) fail(FORT_SCOPE_EXECUTION, "context failed; only abandon is allowed");
        std::lock_guard<std::mutex> dl(c.domain->mu);
        try {
            body(c);
        } catch (const Error &e) {
            if (e.status == FORT_SCOPE_EXECUTION) c.poisoned = true;
            throw;
        }
    } catch (const Error &e) {
        g_error = e.message;
        status = e.status;
    } catch (const std::exception &e) {
        g_error = e.what();
        status = FORT_SCOPE_RESOURCE;
    }
    std::fesetenv(&env);
    return status;
}

template <class F>
static int global(F &&body)
{
    std::fenv_t env;
    std::fegetenv(&env);
    int status = FORT_SCOPE_OK;
    try {
        std::lock_guard<std::mutex> lock(g_registry);
        body();
    } catch (const Error &e) {
        g_error = e.message;
        status = e.status;
    } catch (const std::exception &e) {
        g_error = e.what();
        status = FORT_SCOPE_RESOURCE;
    }
    std::fesetenv(&env);
    return status;
}

static void close_domain_storage(Domain &d)
{
    for (auto &kv : d.buffers)
        if (kv.second->device) {
            dev::release(kv.second->device);
            d.stats.allocated_bytes -= kv.second->bytes;
        }
    d.buffers.clear();
    d.by_identity.clear();
    d.device_bytes = 0;
}

static void release_context(Context &c)
{
#ifndef FTC_CPU_BACKEND
    for (auto &kv : c.graphs)
        for (GraphEntry &g : kv.second.variants) destroy_graph(c, g);
    c.graphs.clear();
    for (auto &kv : c.groups) destroy_graph(c, kv.second);
    c.groups.clear();
    if (!c.retired.empty() && c.stream_ready) dev::sync(c.stream);
    for (hipGraphExec_t e : c.retired) (void)hipGraphExecDestroy(e);
    c.retired.clear();
    if (c.capture_stream) (void)hipStreamDestroy(c.capture_stream);
    c.capture_stream = nullptr;
#endif
    if (c.scratch) dev::release(c.scratch);
    c.scratch = nullptr;
    if (c.staging) {
        if (c.staging_queued && c.stream_ready) dev::sync(c.stream);
        dev::host_release(c.staging);
    }
    c.staging = nullptr;
    c.staging_capacity = 0;
    c.staging_queued = false;
    if (c.stream_ready) dev::destroy_stream(c.stream);
    c.stream_ready = false;
}

static Context &new_context(Domain &d, bool owns)
{
    require(d.borrower == nullptr, FORT_SCOPE_STATE, "the domain is already borrowed by a context");
    auto c = std::make_unique<Context>();
    c->handle = fresh_handle();
    c->domain = &d;
    c->owns_domain = owns;
    c->queue_enabled = env_u64("FORT_SCOPE_GRAPH_QUEUE", 1) != 0;
    c->graphs_enabled = env_u64("FORT_SCOPE_HIP_GRAPHS", env_u64("FORT_SCOPE_CUDA_GRAPHS", 1)) != 0;
    c->queue_work = env_u64("FORT_SCOPE_GRAPH_QUEUE_WORK", uint64_t(1) << 18);
    c->qstats.version = 1;
    c->qstats.queue_enabled = c->queue_enabled;
#ifdef FTC_CPU_BACKEND
    c->qstats.graphs_enabled = 0;
#else
    c->qstats.graphs_enabled = c->graphs_enabled;
#endif
    c->gstats.version = 1;
    d.borrower = c.get();
    Context &ref = *c;
    g_contexts[c->handle] = std::move(c);
    return ref;
}

static Domain &new_domain(int device)
{
    require(device >= 0, FORT_SCOPE_ARGUMENT, "negative device ordinal");
    auto d = std::make_unique<Domain>();
    d->handle = fresh_handle();
    d->device = device;
    d->residency.version = 1;
    std::size_t bytes = 0;
    if (parse_bytes(std::getenv("FORT_RESIDENT_DEVICE_BUDGET"), &bytes)) {
        d->budget = bytes;
        d->budget_known = true;
    }
    Domain &ref = *d;
    g_domains[d->handle] = std::move(d);
    return ref;
}

}  // namespace ftc

using namespace ftc;

// ================================================================================================ C ABI
extern "C" {

uint32_t fort_scope_abi_version(void) { return FORT_SCOPE_ABI_VERSION; }

const char *fort_scope_error(void) { return g_error.c_str(); }

int fort_scope_report_error(int status, const char *message)
{
    g_error = message ? message : "";
    return status;
}

int fort_scope_device_budget_default_v1(int device, std::size_t *bytes)
{
    return global([&] {
        require(bytes != nullptr, FORT_SCOPE_ARGUMENT, "missing budget output");
        *bytes = default_budget(device);
    });
}

int fort_scope_create(int device, fort_scope_t *out)
{
    return global([&] {
        require(out != nullptr, FORT_SCOPE_ARGUMENT, "missing context output");
        Domain &d = new_domain(device);
        d.is_private = true;
        Context &c = new_context(d, true);
        *out = c.handle;
    });
}

int fort_scope_domain_create_v1(int device, fort_scope_domain_t *out)
{
    return global([&] {
        require(out != nullptr, FORT_SCOPE_ARGUMENT, "missing domain output");
        *out = new_domain(device).handle;
    });
}

int fort_scope_create_in_domain_v1(fort_scope_domain_t domain, fort_scope_t *out)
{
    return global([&] {
        require(out != nullptr, FORT_SCOPE_ARGUMENT, "missing context output");
        Domain &d = domain_of(domain);
        require(!d.failed, FORT_SCOPE_STATE, "the domain failed");
        *out = new_context(d, false).handle;
    });
}

int fort_scope_domain_get_v1(fort_scope_t h, fort_scope_domain_t *out)
{
    return with(h, [&](Context &c) {
        require(out != nullptr, FORT_SCOPE_ARGUMENT, "missing domain output");
        *out = c.owns_domain ? 0 : c.domain->handle;
    }, Mode::Query);
}

int fort_scope_domain_device_get_v1(fort_scope_domain_t domain, int *device)
{
    return global([&] {
        require(device != nullptr, FORT_SCOPE_ARGUMENT, "missing device output");
        *device = domain_of(domain).device;
    });
}

int fort_scope_domain_stats_get_v1(fort_scope_domain_t domain, fort_scope_stats *stats)
{
    return global([&] {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing statistics output");
        *stats = domain_of(domain).stats;
    });
}

int fort_scope_domain_close_v1(fort_scope_domain_t domain)
{
    return global([&] {
        Domain &d = domain_of(domain);
        require(d.borrower == nullptr, FORT_SCOPE_STATE, "a context still borrows the domain");
        require(d.workspace.empty(), FORT_SCOPE_STATE, "workspace leases are outstanding");
        // Publication frontier: every value a buffer physically holds reaches the host before storage is
        // released.
        if (!d.failed && !d.buffers.empty()) {
            Context tmp;
            tmp.domain = &d;
            for (auto &kv : d.buffers) {
                Buffer &b = *kv.second;
                const std::vector<Box> need = unpublished(b);
                if (!need.empty()) {
                    ensure_stream(tmp);
                    download(tmp, b, need);
                }
            }
            release_context(tmp);
        }
        close_domain_storage(d);
        g_domains.erase(domain);
    });
}

int fort_scope_domain_abandon_v1(fort_scope_domain_t domain)
{
    return global([&] {
        Domain &d = domain_of(domain);
        require(d.borrower == nullptr, FORT_SCOPE_STATE, "a context still borrows the domain");
        require(d.workspace.empty(), FORT_SCOPE_STATE, "workspace leases are outstanding");
        close_domain_storage(d);
        g_domains.erase(domain);
    });
}

int fort_scope_domain_workspace_reserve_v1(fort_scope_domain_t domain, std::size_t bytes, uint64_t *lease)
{
    return global([&] {
        require(lease != nullptr, FORT_SCOPE_ARGUMENT, "missing lease output");
        Domain &d = domain_of(domain);
        require(d.workspace_bytes + d.device_bytes + bytes <= budget_of(d), FORT_SCOPE_RESOURCE,
                "workspace exceeds the unclaimed budget");
        uint64_t id = d.next_lease++;
        d.workspace[id] = bytes;
        d.workspace_bytes += bytes;
        *lease = id;
    });
}

int fort_scope_domain_workspace_resize_v1(fort_scope_domain_t domain, uint64_t lease, std::size_t bytes)
{
    return global([&] {
        Domain &d = domain_of(domain);
        auto it = d.workspace.find(lease);
        require(it != d.workspace.end(), FORT_SCOPE_STALE, "unknown workspace lease");
        std::size_t others = d.workspace_bytes - it->second;
        require(others + d.device_bytes + bytes <= budget_of(d), FORT_SCOPE_RESOURCE,
                "workspace exceeds the unclaimed budget");
        d.workspace_bytes = others + bytes;
        it->second = bytes;
    });
}

int fort_scope_domain_workspace_release_v1(fort_scope_domain_t domain, uint64_t lease)
{
    return global([&] {
        Domain &d = domain_of(domain);
        auto it = d.workspace.find(lease);
        require(it != d.workspace.end(), FORT_SCOPE_STALE, "unknown workspace lease");
        d.workspace_bytes -= it->second;
        d.workspace.erase(it);
    });
}

int fort_scope_domain_budget_get_v1(fort_scope_domain_t domain, std::size_t *budget, std::size_t *unclaimed)
{
    return global([&] {
        require(budget && unclaimed, FORT_SCOPE_ARGUMENT, "missing budget outputs");
        Domain &d = domain_of(domain);
        std::size_t b = budget_of(d);
        std::size_t claimed = d.device_bytes + d.workspace_bytes;
        if (d.borrower) claimed += d.borrower->scratch_capacity;
        *budget = b;
        *unclaimed = b > claimed ? b - claimed : 0;
    });
}

int fort_scope_set_device_budget(fort_scope_t h, std::size_t bytes)
{
    return with(h, [&](Context &c) {
        require(c.domain->device_bytes <= bytes, FORT_SCOPE_STATE, "budget below the current allocations");
        c.domain->budget = bytes;
        c.domain->budget_known = true;
    });
}

int fort_scope_scratch_acquire_v1(fort_scope_t h, std::size_t bytes, fort_scope_scratch_lease_v1 *lease)
{
    return with(h, [&](Context &c) {
        require(lease != nullptr, FORT_SCOPE_ARGUMENT, "missing lease output");
        require(!c.scratch_active, FORT_SCOPE_STATE, "a scratch lease is already active");
        if (bytes > c.scratch_capacity) {
            ensure_stream(c);
            flush(c);
            wait(c);
            Domain &d = *c.domain;
            require(d.device_bytes + d.workspace_bytes + bytes <= budget_of(d), FORT_SCOPE_RESOURCE,
                    "scratch exceeds the device budget");
            const bool grow = c.scratch != nullptr;
            if (c.scratch) dev::release(c.scratch);
            c.scratch = nullptr;
            c.scratch_capacity = 0;
            dev::set(d.device);
            c.scratch = dev::alloc(bytes);
            require(c.scratch != nullptr, FORT_SCOPE_RESOURCE, "scratch allocation failed");
            c.scratch_capacity = bytes;
            c.sstats.allocations++;
            if (grow) c.sstats.grows++;
        } else if (bytes) {
            c.sstats.reuses++;
        }
        c.sstats.acquisitions++;
        c.sstats.active_bytes = bytes;
        if (bytes > c.sstats.peak_scratch_bytes) c.sstats.peak_scratch_bytes = bytes;
        {
            const Domain &d = *c.domain;
            const uint64_t total = d.device_bytes + d.workspace_bytes + c.scratch_capacity;
            if (total > c.sstats.peak_total_device_bytes) c.sstats.peak_total_device_bytes = total;
        }
        c.scratch_active = true;
        c.scratch_token = c.next_token++;
        lease->version = 1;
        lease->reserved = 0;
        lease->token = c.scratch_token;
        lease->device = bytes ? c.scratch : nullptr;
        lease->bytes = bytes;
        lease->capacity = c.scratch_capacity;
    });
}

int fort_scope_scratch_release_v1(fort_scope_t h, uint64_t token)
{
    return with(h, [&](Context &c) {
        require(c.scratch_active && token == c.scratch_token, FORT_SCOPE_STATE, "no such active scratch lease");
        c.scratch_active = false;
        c.sstats.releases++;
        c.sstats.active_bytes = 0;
    });
}

int fort_scope_scratch_stats_get_v1(fort_scope_t h, fort_scope_scratch_stats_v1 *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing stats output");
        *stats = c.sstats;
        stats->version = 1;
        stats->active = c.scratch_active ? 1 : 0;
        stats->capacity_bytes = c.scratch_capacity;
    }, Mode::Query);
}

int fort_scope_completion_stats_get_v1(fort_scope_t h, fort_scope_completion_stats_v1 *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing stats output");
        *stats = c.cstats;
        stats->version = 1;
        stats->reserved = 0;
    }, Mode::Query);
}

static bool tracing() { return std::getenv("FORT_RUNTIME_TRACE") != nullptr; }

static void trace_field(char (&dst)[65], const char *src)
{
    std::memset(dst, 0, sizeof dst);
    if (!src) return;
    const std::size_t n = strnlen(src, sizeof dst);
    if (n < sizeof dst) std::memcpy(dst, src, n);  // an ID without its NUL within 65 bytes stays unknown
}

void fort_scope_trace_set_v1(fort_scope_t h, uint32_t mask, const char *owner, const char *procedure,
                             const char *segment, const char *operation, const char *implementation,
                             const char *boundary, fort_scope_trace_state_v1 *previous)
{
    if (!tracing()) return;
    (void)with(h, [&](Context &c) {
        if (previous) *previous = c.trace;
        c.trace.version = 1;
        if (mask & FORT_SCOPE_TRACE_OWNER) trace_field(c.trace.owner, owner);
        if (mask & FORT_SCOPE_TRACE_PROCEDURE) trace_field(c.trace.procedure, procedure);
        if (mask & FORT_SCOPE_TRACE_SEGMENT) trace_field(c.trace.segment, segment);
        if (mask & FORT_SCOPE_TRACE_OPERATION) trace_field(c.trace.operation, operation);
        if (mask & FORT_SCOPE_TRACE_IMPLEMENTATION) trace_field(c.trace.implementation, implementation);
        if (mask & FORT_SCOPE_TRACE_BOUNDARY) trace_field(c.trace.boundary, boundary);
    }, Mode::Query);
}

void fort_scope_trace_restore_v1(fort_scope_t h, const fort_scope_trace_state_v1 *previous)
{
    if (!tracing() || !previous) return;
    (void)with(h, [&](Context &c) { c.trace = *previous; }, Mode::Query);
}

static int register_impl(fort_scope_t h, uint64_t identity, uint64_t generation, const fort_scope_layout *layout,
                         const fort_scope_section *init, std::size_t init_count, bool all, fort_buffer_t *out)
{
    return with(h, [&](Context &c) {
        require(layout && out, FORT_SCOPE_ARGUMENT, "missing layout or handle output");
        require(layout->rank >= 1 && layout->rank <= static_cast<uint32_t>(kMaxRank), FORT_SCOPE_ARGUMENT,
                "unsupported rank");
        require(layout->element_bytes > 0 && layout->extents && layout->host, FORT_SCOPE_ARGUMENT,
                "incomplete layout");
        Domain &d = *c.domain;
        auto found = d.by_identity.find(identity);
        if (found != d.by_identity.end()) {
            Buffer &b = *d.buffers.at(found->second);
            bool same = b.rank == layout->rank && b.elem == layout->element_bytes && b.host == layout->host &&
                        b.generation == generation;
            for (uint32_t k = 0; same && k < b.rank; ++k) same = b.extents[k] == layout->extents[k];
            require(same, FORT_SCOPE_STALE, "identity re-registered with another allocation");
            *out = b.handle;
            return;
        }
        auto b = std::make_unique<Buffer>();
        b->handle = fresh_handle();
        b->domain = &d;
        b->identity = identity;
        b->generation = generation;
        b->rank = layout->rank;
        b->type = layout->type;
        b->elem = layout->element_bytes;
        b->host = layout->host;
        b->count = 1;
        for (uint32_t k = 0; k < b->rank; ++k) {
            b->extents[k] = layout->extents[k];
            b->lower_bounds[k] = layout->lower_bounds ? layout->lower_bounds[k] : 1;
            b->count *= b->extents[k];
        }
        b->bytes = b->count * b->elem;
        if (all && b->count) {
            b->init.add(b->full());
            b->host_cur.add(b->full());
        }
        for (std::size_t i = 0; i < init_count; ++i) {
            Box box = section_box(*b, init[i]);
            require(b->init.add(box) && b->host_cur.add(box), FORT_SCOPE_RESOURCE, "region limit exceeded");
        }
        *out = b->handle;
        d.by_identity[identity] = b->handle;
        d.buffers[b->handle] = std::move(b);
    });
}

int fort_scope_register(fort_scope_t h, uint64_t identity, uint64_t generation, const fort_scope_layout *layout,
                        int host_initialized, fort_buffer_t *out)
{
    return register_impl(h, identity, generation, layout, nullptr, 0, host_initialized != 0, out);
}

int fort_scope_register_sections(fort_scope_t h, uint64_t identity, uint64_t generation,
                                 const fort_scope_layout *layout, const fort_scope_section *initialized,
                                 std::size_t count, fort_buffer_t *out)
{
    if (count && !initialized) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "missing sections");
    return register_impl(h, identity, generation, layout, initialized, count, false, out);
}

int fort_scope_forget_definition(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        // Semantic definitions go; the physical currency stays, so values already written are still
        // published (compiler/runtime/scoped_runtime.cu:1901-1909). Queued tasks that write the buffer
        // keep their effects because the coherence state already reflects them at acceptance.
        b.init.clear();
    });
}

int fort_scope_layout_get(fort_scope_t h, fort_buffer_t handle, fort_scope_layout *layout)
{
    return with(h, [&](Context &c) {
        require(layout != nullptr, FORT_SCOPE_ARGUMENT, "missing layout output");
        Buffer &b = buffer(c, handle);
        layout->rank = b.rank;
        layout->type = b.type;
        layout->element_bytes = b.elem;
        layout->host = b.host;
        layout->extents = b.extents;
        layout->lower_bounds = b.lower_bounds;
        layout->generation = b.generation;
    }, Mode::Query);
}

int fort_scope_access_begin_v1(fort_scope_t h, const fort_scope_plan_binding *bindings, std::size_t count,
                               int device, void **devices)
{
    return with(h, [&](Context &c) {
        require(device == 0 || device == 1, FORT_SCOPE_ARGUMENT, "invalid task access execution mode");
        for (std::size_t i = 0; i < count; ++i) drain_buffer(c, buffer(c, bindings[i].buffer));
        if (device) {
            prepare_device(c, bindings, count, devices, false);
            return;
        }
        std::vector<Buffer *> list;
        std::vector<Effects> eff;
        for (std::size_t i = 0; i < count; ++i) {
            Buffer &b = buffer(c, bindings[i].buffer);
            for (Buffer *o : list) require(o != &b, FORT_SCOPE_ALIAS, "task access bindings repeat a buffer");
            require(!b.prepared, FORT_SCOPE_STATE, "buffer already has a prepared access");
            list.push_back(&b);
            eff.push_back(effects(b, &bindings[i].access));
        }
        std::vector<Acquisition> plans;
        bool moves = false;
        for (std::size_t i = 0; i < count; ++i) {
            plans.push_back(plan_acquisition(*list[i], eff[i], false, false));
            moves = moves || !plans.back().copies.empty();
        }
        if (!moves) c.cstats.independent_host_tasks++;
        for (std::size_t i = 0; i < count; ++i) {
            acquire(c, *list[i], plans[i], false);
            list[i]->prepared = 2;
            list[i]->prep_writes = eff[i].writes;
            list[i]->prep_overwrites = eff[i].overwrites;
        }
    });
}

int fort_scope_original_host_begin_v1(fort_scope_t h, const fort_scope_plan_binding *bindings, std::size_t count)
{
    return with(h, [&](Context &c) {
        require(count == 0 || bindings != nullptr, FORT_SCOPE_ARGUMENT, "missing bindings");
        std::vector<Buffer *> list;
        std::vector<Effects> eff;
        // Every binding is checked before anything moves, so a later bad binding publishes nothing.
        for (std::size_t i = 0; i < count; ++i) {
            Buffer &b = buffer(c, bindings[i].buffer);
            for (Buffer *o : list) require(o != &b, FORT_SCOPE_ALIAS, "task access bindings repeat a buffer");
            require(!b.prepared, FORT_SCOPE_STATE, "buffer already has a prepared access");
            list.push_back(&b);
            eff.push_back(effects(b, &bindings[i].access));
        }
        for (Buffer *b : list) drain_buffer(c, *b);
        std::vector<Acquisition> plans;
        bool moves = false;
        for (std::size_t i = 0; i < count; ++i) {
            plans.push_back(plan_acquisition(*list[i], eff[i], false, true));
            moves = moves || !plans.back().copies.empty();
        }
        if (!moves) c.cstats.independent_host_tasks++;
        for (std::size_t i = 0; i < count; ++i) {
            acquire(c, *list[i], plans[i], false);
            list[i]->prepared = 2;
            list[i]->prep_original = true;
            list[i]->prep_writes = eff[i].writes;
            list[i]->prep_overwrites = eff[i].overwrites;
        }
    });
}

// True when layout l describes buffer b's physical allocation exactly.
static bool describes(const fort_scope_layout &l, const Buffer &b)
{
    if (l.rank != b.rank || l.type != b.type || l.element_bytes != b.elem || l.host != b.host ||
        l.generation != b.generation || !l.extents)
        return false;
    for (uint32_t k = 0; k < b.rank; ++k) {
        const int64_t lower = l.lower_bounds ? l.lower_bounds[k] : 1;
        if (l.extents[k] != b.extents[k] || lower != b.lower_bounds[k]) return false;
    }
    return true;
}

int fort_scope_identity_swap_v1(fort_scope_t h, fort_buffer_t first, fort_buffer_t second,
                                const fort_scope_layout *first_after, const fort_scope_layout *second_after,
                                fort_buffer_t *first_current, fort_buffer_t *second_current)
{
    return with(h, [&](Context &c) {
        Domain &d = *c.domain;
        // The exchange already happened in the caller, so a rejection cannot be answered by a replay: it
        // fails the context and the domain.
        auto reject = [&](const char *why) {
            d.failed = true;
            fail(FORT_SCOPE_EXECUTION, std::string("identity swap: ") + why);
        };
        if (!first_after || !second_after || !first_current || !second_current) reject("missing argument");
        auto fa = d.buffers.find(first), fb = d.buffers.find(second);
        if (fa == d.buffers.end() || fb == d.buffers.end()) reject("unknown buffer");
        if (first == second) reject("the two buffers are one");
        Buffer &a = *fa->second, &b = *fb->second;
        if (a.identity == b.identity) reject("the two identities are one");
        if (a.prepared || b.prepared) reject("a buffer has a prepared access");
        if (!describes(*first_after, b) || !describes(*second_after, a))
            reject("a post-swap layout does not describe the other allocation");
        d.by_identity[a.identity] = b.handle;
        d.by_identity[b.identity] = a.handle;
        std::swap(a.identity, b.identity);
        *first_current = b.handle;
        *second_current = a.handle;
    });
}

int fort_scope_device_begin(fort_scope_t h, fort_buffer_t handle, const fort_scope_access *access, void **device)
{
    fort_scope_plan_binding b{};
    b.buffer = handle;
    if (!access) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "missing access description");
    b.access = *access;
    return fort_scope_access_begin_v1(h, &b, 1, 1, device);
}

int fort_scope_device_end(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) { end_access(c, buffer(c, handle), true, false); });
}

int fort_scope_host_begin(fort_scope_t h, fort_buffer_t handle, const fort_scope_access *access)
{
    fort_scope_plan_binding b{};
    b.buffer = handle;
    if (!access) return fort_scope_report_error(FORT_SCOPE_ARGUMENT, "missing access description");
    b.access = *access;
    return fort_scope_access_begin_v1(h, &b, 1, 0, nullptr);
}

int fort_scope_host_end(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) { end_access(c, buffer(c, handle), false, false); });
}

int fort_scope_cancel_access(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        end_access(c, b, b.prepared == 1, true);
    });
}

int fort_scope_device_reserve_v1(fort_scope_t h, fort_buffer_t handle, void **device)
{
    return fort_scope_device_reserve_many_v1(h, &handle, 1, device);
}

int fort_scope_device_reserve_many_v1(fort_scope_t h, const fort_buffer_t *handles, std::size_t count,
                                      void **devices)
{
    return with(h, [&](Context &c) {
        require(!count || (handles && devices), FORT_SCOPE_ARGUMENT, "missing reserve arguments");
        std::vector<Buffer *> list;
        for (std::size_t i = 0; i < count; ++i) list.push_back(&buffer(c, handles[i]));
        if (!count) return;
        ensure_stream(c);
        reserve(c, list.data(), list.size());
        for (std::size_t i = 0; i < count; ++i) devices[i] = list[i]->device;
    });
}

int fort_scope_buffer_valid_v1(fort_scope_t h, fort_buffer_t handle, int *valid)
{
    return with(h, [&](Context &c) {
        require(valid != nullptr, FORT_SCOPE_ARGUMENT, "missing output");
        *valid = c.domain->buffers.count(handle) ? 1 : 0;
    }, Mode::Query);
}

int fort_scope_residency_stats_get_v1(fort_scope_t h, fort_scope_residency_stats_v1 *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing statistics output");
        *stats = c.domain->residency;
    }, Mode::Query);
}

int fort_scope_observe_v1(fort_scope_t h, const fort_buffer_t *handles, std::size_t count)
{
    return with(h, [&](Context &c) {
        require(!count || handles, FORT_SCOPE_ARGUMENT, "missing buffer list");
        for (std::size_t i = 0; i < count; ++i) {
            Buffer &b = buffer(c, handles[i]);
            require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
            observe(c, b);
        }
    });
}

int fort_scope_observe_all_v1(fort_scope_t h)
{
    return with(h, [&](Context &c) {
        flush(c);
        for (auto &kv : c.domain->buffers) {
            require(!kv.second->prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
            observe(c, *kv.second);
        }
    });
}

int fort_scope_discard_v1(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        drain_buffer(c, b);
        b.init.clear();
        b.host_cur.clear();
        b.dev_cur.clear();
    });
}

int fort_scope_unregister(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        observe(c, b);
        wait(c);
        if (b.device) {
            dev::release(b.device);
            c.domain->device_bytes -= b.bytes;
            c.domain->stats.allocated_bytes -= b.bytes;
        }
        c.domain->by_identity.erase(b.identity);
        c.domain->buffers.erase(handle);
    });
}

int fort_scope_recycle_v1(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        // With queued users the dead values stay until the owner's next definition replaces them, and the
        // queue is not submitted for it; without, the values are forgotten as by discard.
        if (b.queued > 0) return;
        b.init.clear();
        b.host_cur.clear();
        b.dev_cur.clear();
    });
}

int fort_scope_retire_v1(fort_scope_t h, fort_buffer_t handle)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        // Dead values: no publication. Queued and running uses of the storage (device memory and the
        // borrowed host array) complete first; then the handle is gone.
        drain_buffer(c, b);
        wait(c);
        if (b.device) {
            dev::release(b.device);
            c.domain->device_bytes -= b.bytes;
            c.domain->stats.allocated_bytes -= b.bytes;
        }
        c.domain->by_identity.erase(b.identity);
        c.domain->buffers.erase(handle);
    });
}

int fort_scope_admit_source_reads_v1(fort_scope_t h, fort_buffer_t handle, const fort_scope_section *sections,
                                     std::size_t count)
{
    return with(h, [&](Context &c) {
        Buffer &b = buffer(c, handle);
        require(count == 0 || sections != nullptr, FORT_SCOPE_ARGUMENT, "missing sections");
        require(!b.prepared, FORT_SCOPE_STATE, "buffer has a prepared access");
        // Only values outside the defined coverage gain a definition and a current host copy; whatever is
        // defined keeps its freshness (a dirty device copy stays the current one). Both regions are built
        // before either is committed.
        Region init = b.init, host = b.host_cur;
        for (std::size_t i = 0; i < count; ++i) {
            const Box box = section_box(b, sections[i]);
            for (const Box &fresh : b.init.missing(box)) {
                require(init.add(fresh), FORT_SCOPE_BOUNDARY, "admitted reads exceed the region limit");
                // A definition kill does not erase a newer physical device value: only values with no
                // tracked copy take the host's (compiler/runtime/scoped_regions.hpp:205-216).
                for (const Box &untracked : b.dev_cur.missing(fresh))
                    require(host.add(untracked), FORT_SCOPE_BOUNDARY, "admitted reads exceed the region limit");
            }
        }
        b.init = init;
        b.host_cur = host;
    });
}

int fort_scope_gpu_enter(fort_scope_t h, int *previous, void **stream)
{
    return with(h, [&](Context &c) {
        require(previous && stream, FORT_SCOPE_ARGUMENT, "missing outputs");
        *previous = dev::current();
        ensure_stream(c);
        dev::set(c.domain->device);
        *stream = reinterpret_cast<void *>(c.stream);
    });
}

int fort_scope_gpu_leave(fort_scope_t h, int previous)
{
    return with(h, [&](Context &) { dev::set(previous); }, Mode::Query);
}

int fort_scope_note_launch(fort_scope_t h)
{
    return with(h, [&](Context &c) { c.domain->stats.launches++; });
}

// Accepts one launch with its bindings into the context's queue (the effects are committed now; the
// launch runs in acceptance order on the context stream).
static void accept_task(Context &c, uint64_t code_key, const fort_scope_kernel_v1 &k, const std::size_t *argument_sizes,
                        const fort_scope_plan_binding *bindings, std::size_t binding_count)
{
    {
        require(k.function != nullptr, FORT_SCOPE_ARGUMENT, "missing kernel function");
        require(!k.argument_count || (k.arguments && argument_sizes), FORT_SCOPE_ARGUMENT, "missing arguments");
        uint64_t threads = uint64_t(k.grid[0]) * k.grid[1] * k.grid[2] * k.block[0] * k.block[1] * k.block[2];
        if (!threads) return;  // empty launches initialise nothing
#ifdef FTC_CPU_BACKEND
        fail(FORT_SCOPE_RESOURCE, "the CPU reference backend executes no device kernels");
#endif
        // No drain: tasks on one context run in acceptance order on its stream, and uploads only touch
        // cells no accepted task holds on the device.
        std::vector<void *> devices(binding_count);
        prepare_device(c, bindings, binding_count, devices.data(), true);
        Task t;
        t.code_key = code_key;
        t.function = k.function;
        for (int d = 0; d < 3; ++d) {
            t.grid[d] = k.grid[d];
            t.block[d] = k.block[d];
        }
        t.shared = k.shared_bytes;
        t.threads = threads;
        std::size_t off = 0;
        for (std::size_t i = 0; i < k.argument_count; ++i) {
            std::size_t sz = argument_sizes[i];
            std::size_t align = sz >= 8 ? 8 : (sz >= 4 ? 4 : (sz >= 2 ? 2 : 1));
            off = (off + align - 1) / align * align;
            t.offsets.push_back(off);
            t.sizes.push_back(sz);
            off += sz;
        }
        t.args.resize(off);
        for (std::size_t i = 0; i < k.argument_count; ++i)
            std::memcpy(t.args.data() + t.offsets[i], k.arguments[i], t.sizes[i]);
        for (std::size_t i = 0; i < binding_count; ++i) {
            Buffer &b = buffer(c, bindings[i].buffer);
            b.queued++;
            t.buffers.push_back(&b);
        }
        c.qstats.copied_argument_bytes += off;
        c.qstats.accepted_tasks++;
        c.queue.push_back(std::move(t));
        c.queued_threads += threads;
        c.qstats.pending_tasks = c.queue.size();
        if (c.queue.size() > c.qstats.peak_pending_tasks) c.qstats.peak_pending_tasks = c.queue.size();
        if (!c.queue_enabled || (c.queue_work && c.queued_threads >= c.queue_work)) flush(c);
    }
}

int fort_scope_graph_enqueue_v1(fort_scope_t h, uint64_t code_key, const fort_scope_graph_task_v1 *task,
                                const std::size_t *argument_sizes)
{
    return with(h, [&](Context &c) {
        require(task != nullptr, FORT_SCOPE_ARGUMENT, "missing task");
        accept_task(c, code_key, task->kernel, argument_sizes, task->bindings, task->binding_count);
    });
}

// scoped_runtime.h:401-413: members are checked pairwise on their actual sections, then merged into one
// binding per buffer whose sections are the union of the members' sections.
int fort_scope_graph_enqueue_members_v1(fort_scope_t h, uint64_t code_key, const fort_scope_kernel_v1 *kernel,
                                        const std::size_t *argument_sizes, const fort_scope_graph_member_v1 *members,
                                        std::size_t member_count)
{
    return with(h, [&](Context &c) {
        require(kernel != nullptr, FORT_SCOPE_ARGUMENT, "missing kernel");
        require(!member_count || members, FORT_SCOPE_ARGUMENT, "missing members");
        struct Use {
            Buffer *buf;
            fort_buffer_t handle;
            Effects e;
            std::size_t member;
        };
        std::vector<Use> uses;
        for (std::size_t m = 0; m < member_count; ++m) {
            require(!members[m].binding_count || members[m].bindings, FORT_SCOPE_ARGUMENT, "missing member bindings");
            for (std::size_t i = 0; i < members[m].binding_count; ++i) {
                const fort_scope_plan_binding &pb = members[m].bindings[i];
                Buffer &b = buffer(c, pb.buffer);
                for (const Use &u : uses)
                    require(u.member != m || u.buf != &b, FORT_SCOPE_ALIAS, "a member's bindings repeat a buffer");
                uses.push_back({&b, pb.buffer, effects(b, &pb.access), m});
            }
        }
        auto touches = [](const std::vector<Box> &writes, const Effects &o) {
            for (const Box &w : writes)
                for (const std::vector<Box> *list : {&o.reads, &o.writes, &o.overwrites})
                    for (const Box &x : *list)
                        if (w.intersects(x)) return true;
            return false;
        };
        for (const Use &a : uses)
            for (const Use &b : uses)
                if (a.member != b.member && a.buf == b.buf && (touches(a.e.writes, b.e) || touches(a.e.overwrites, b.e)))
                    fail(FORT_SCOPE_BOUNDARY, "launch members overlap: one writes a section another member uses");
        // One binding per distinct buffer, listing every member's boxes as explicit sections.
        std::vector<fort_buffer_t> order;
        for (const Use &u : uses)
            if (std::find(order.begin(), order.end(), u.handle) == order.end()) order.push_back(u.handle);
        std::size_t nbox = 0;
        for (const Use &u : uses) nbox += u.e.reads.size() + u.e.writes.size() + u.e.overwrites.size();
        std::vector<std::size_t> coords(2 * kMaxRank * (nbox ? nbox : 1));
        std::vector<fort_scope_section> sections(nbox ? nbox : 1);
        std::size_t used = 0;
        auto section_of = [&](const Box &x) {
            std::size_t *lo = coords.data() + 2 * kMaxRank * used, *hi = lo + kMaxRank;
            for (int d = 0; d < x.rank; ++d) {
                lo[d] = x.lo[d];
                hi[d] = x.hi[d];
            }
            sections[used] = fort_scope_section{lo, hi};
            return &sections[used++];
        };
        std::vector<fort_scope_plan_binding> merged;
        for (fort_buffer_t handle : order) {
            fort_scope_plan_binding pb{};
            pb.buffer = handle;
            const fort_scope_section *firsts[3] = {nullptr, nullptr, nullptr};
            std::size_t counts[3] = {0, 0, 0};
            for (int kind = 0; kind < 3; ++kind)
                for (const Use &u : uses) {
                    if (u.handle != handle) continue;
                    const std::vector<Box> &list = kind == 0 ? u.e.reads : (kind == 1 ? u.e.writes : u.e.overwrites);
                    for (const Box &x : list) {
                        const fort_scope_section *sp = section_of(x);
                        if (!firsts[kind]) firsts[kind] = sp;
                        counts[kind]++;
                    }
                }
            pb.access.flags = 0;
            pb.access.read_count = counts[0];
            pb.access.reads = firsts[0];
            pb.access.write_count = counts[1];
            pb.access.writes = firsts[1];
            pb.access.overwrite_count = counts[2];
            pb.access.overwrites = firsts[2];
            merged.push_back(pb);
        }
        accept_task(c, code_key, *kernel, argument_sizes, merged.data(), merged.size());
    });
}

// ------------------------------------------------------------------------------------------------ groups
static void check_kernel(const fort_scope_kernel_v1 &k)
{
    require(k.function != nullptr, FORT_SCOPE_ARGUMENT, "missing kernel function");
    require(!k.argument_count || k.arguments, FORT_SCOPE_ARGUMENT, "missing kernel arguments");
    for (int d = 0; d < 3; ++d)
        require(k.grid[d] && k.block[d], FORT_SCOPE_ARGUMENT, "empty kernel launch dimensions");
}

#ifndef FTC_CPU_BACKEND
static hipKernelNodeParams group_params(const fort_scope_kernel_v1 &k)
{
    hipKernelNodeParams p{};
    p.blockDim = dim3(k.block[0], k.block[1], k.block[2]);
    p.gridDim = dim3(k.grid[0], k.grid[1], k.grid[2]);
    p.func = const_cast<void *>(k.function);
    p.kernelParams = const_cast<void **>(k.arguments);  // HIP copies the values during this call
    p.sharedMemBytes = static_cast<unsigned>(k.shared_bytes);
    p.extra = nullptr;
    return p;
}
#endif

// Runs an ordered kernel group on the context stream: one chain graph per nonzero group key, rebuilt
// when the functions or argument counts change and otherwise replayed with every node's parameters,
// dimensions and shared memory updated; group key 0 launches the kernels directly.
static void submit_group(Context &c, uint64_t key, const std::vector<const fort_scope_kernel_v1 *> &ks)
{
#ifdef FTC_CPU_BACKEND
    (void)c;
    (void)key;
    (void)ks;
    fail(FORT_SCOPE_RESOURCE, "the CPU reference backend executes no device kernels");
#else
    ensure_stream(c);
    dev::set(c.domain->device);
    c.gstats.submissions++;
    c.gstats.logical_kernels += ks.size();
    c.domain->stats.launches += ks.size();
    auto poison = [&](const char *what, hipError_t e) {
        (void)hipGetLastError();
        c.poisoned = true;
        fail(FORT_SCOPE_EXECUTION, std::string(what) + hipGetErrorString(e));
    };
    if (key == 0 || !c.graphs_enabled) {
        for (const fort_scope_kernel_v1 *k : ks) {
            hipError_t e = hipLaunchKernel(k->function, dim3(k->grid[0], k->grid[1], k->grid[2]),
                                           dim3(k->block[0], k->block[1], k->block[2]),
                                           const_cast<void **>(k->arguments), k->shared_bytes, c.stream);
            if (e != hipSuccess) poison("kernel launch failed: ", e);
        }
        return;
    }
    GraphEntry &g = c.groups[key];
    // The topology of a group: its functions and argument counts (kept in dims for the comparison).
    bool same = g.exec && g.nodes.size() == ks.size();
    for (std::size_t i = 0; same && i < ks.size(); ++i)
        same = g.dims[i].size() == 2 && g.dims[i][0] == (uint32_t)ks[i]->argument_count &&
               g.args[i].size() == sizeof(void *) &&
               std::memcmp(g.args[i].data(), &ks[i]->function, sizeof(void *)) == 0;
    if (same) {
        for (std::size_t i = 0; i < ks.size(); ++i) {
            hipKernelNodeParams p = group_params(*ks[i]);
            hipError_t e = hipGraphExecKernelNodeSetParams(g.exec, g.nodes[i], &p);
            if (e != hipSuccess) poison("graph parameter update failed: ", e);
            c.gstats.parameter_updates++;
        }
        c.gstats.replays++;
    } else {
        if (g.exec) c.gstats.rebuilds++;
        destroy_graph(c, g);
        hipError_t e = hipGraphCreate(&g.graph, 0);
        if (e != hipSuccess) poison("graph creation failed: ", e);
        hipGraphNode_t prev = nullptr;
        g.args.clear();
        g.dims.clear();
        for (const fort_scope_kernel_v1 *k : ks) {
            hipKernelNodeParams p = group_params(*k);
            hipGraphNode_t node = nullptr;
            e = hipGraphAddKernelNode(&node, g.graph, prev ? &prev : nullptr, prev ? 1 : 0, &p);
            if (e != hipSuccess) poison("graph node creation failed: ", e);
            g.nodes.push_back(node);
            prev = node;
            std::vector<unsigned char> fn(sizeof(void *));
            std::memcpy(fn.data(), &k->function, sizeof(void *));
            g.args.push_back(fn);
            g.dims.push_back({(uint32_t)k->argument_count, 0});
        }
        e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0);
        if (e != hipSuccess) poison("graph instantiation failed: ", e);
        c.gstats.creations++;
        c.gstats.instantiations++;
    }
    hipError_t e = hipGraphLaunch(g.exec, c.stream);
    if (e != hipSuccess) poison("graph launch failed: ", e);
#endif
}

int fort_scope_graph_submit_v1(fort_scope_t h, uint64_t group_key, const fort_scope_kernel_v1 *kernels,
                               std::size_t count)
{
    return with(h, [&](Context &c) {
        if (!count) return;  // a no-op group initializes nothing
        require(kernels != nullptr, FORT_SCOPE_ARGUMENT, "missing kernels");
        std::vector<const fort_scope_kernel_v1 *> ks;
        for (std::size_t i = 0; i < count; ++i) {
            check_kernel(kernels[i]);
            ks.push_back(&kernels[i]);
        }
        // Work queued before the group runs before it.
        flush(c);
        submit_group(c, group_key, ks);
    });
}

int fort_scope_graph_execute_v1(fort_scope_t h, uint64_t group_key, const fort_scope_graph_task_v1 *tasks,
                                std::size_t count)
{
    return with(h, [&](Context &c) {
        if (!count) return;
        require(tasks != nullptr, FORT_SCOPE_ARGUMENT, "missing tasks");
        struct Use {
            Buffer *b;
            Effects e;
        };
        std::vector<std::vector<Use>> uses(count);
        std::vector<Buffer *> order;                   // the group's buffers, first use first
        std::map<Buffer *, Region> defined, produced;  // simulated definitions; written by the group so far
        std::map<Buffer *, Region> need;               // external requirements to upload
        for (std::size_t i = 0; i < count; ++i) {
            const fort_scope_graph_task_v1 &t = tasks[i];
            check_kernel(t.kernel);
            require(!t.binding_count || t.bindings, FORT_SCOPE_ARGUMENT, "missing task bindings");
            for (std::size_t j = 0; j < t.binding_count; ++j) {
                Buffer &b = buffer(c, t.bindings[j].buffer);
                for (const Use &u : uses[i]) require(u.b != &b, FORT_SCOPE_ALIAS, "task bindings repeat a buffer");
                require(!b.prepared, FORT_SCOPE_STATE, "buffer already has a prepared access");
                if (!defined.count(&b)) {
                    defined[&b] = b.init;
                    produced[&b] = Region{};
                    need[&b] = Region{};
                    order.push_back(&b);
                }
                Effects e = effects(b, &t.bindings[j].access);
                // Reads, and what partial writes preserve, must be defined by the state before the group
                // or by an earlier task of it.
                std::vector<Box> touched = e.reads;
                Region over;
                for (const Box &o : e.overwrites) require(over.add(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
                for (const Box &w : e.writes)
                    for (const Box &kept : over.missing(w)) touched.push_back(kept);
                for (const Box &r : touched)
                    require(defined[&b].covers(r), FORT_SCOPE_UNINITIALIZED,
                            "a task of the group reads values that are not defined before it");
                // Values defined before the group that the device lacks and no earlier task produced go up.
                for (const Box &box : touched)
                    for (const Box &d0 : b.init.within(box))
                        for (const Box &m : b.dev_cur.missing(d0))
                            for (const Box &x : produced[&b].missing(m))
                                for (const Box &fresh : need[&b].missing(x))
                                    require(need[&b].add(fresh), FORT_SCOPE_RESOURCE, "region limit exceeded");
                uses[i].push_back({&b, e});
            }
            for (const Use &u : uses[i]) {
                for (const Box &o : u.e.overwrites)
                    require(defined[u.b].add(o) && produced[u.b].add(o), FORT_SCOPE_RESOURCE, "region limit exceeded");
                for (const Box &w : u.e.writes)
                    require(produced[u.b].add(w), FORT_SCOPE_RESOURCE, "region limit exceeded");
            }
        }
#ifdef FTC_CPU_BACKEND
        fail(FORT_SCOPE_RESOURCE, "the CPU reference backend executes no device kernels");
#else
        for (Buffer *b : order) drain_buffer(c, *b);
        ensure_stream(c);
        reserve(c, order.data(), order.size());
        for (Buffer *b : order)
            for (const Box &box : need[b].boxes) {
                copy_box(c, *b, box, true);
                require(b->dev_cur.add(box), FORT_SCOPE_RESOURCE, "region limit exceeded");
            }
        std::vector<const fort_scope_kernel_v1 *> ks;
        for (std::size_t i = 0; i < count; ++i) ks.push_back(&tasks[i].kernel);
        submit_group(c, group_key, ks);
        // Definitions and coherence only after a successful submission, task by task in order.
        for (std::size_t i = 0; i < count; ++i)
            for (const Use &u : uses[i]) commit(*u.b, u.e.writes, u.e.overwrites, true);
#endif
    });
}

int fort_scope_graph_flush_v1(fort_scope_t h)
{
    return with(h, [&](Context &c) {
        if (!c.queue.empty()) c.qstats.explicit_flushes++;
        flush(c);
    });
}

int fort_scope_graph_queue_cancel_v1(fort_scope_t h)
{
    return with(h, [&](Context &c) {
        // Accepted effects were committed at acceptance; cancelling unsubmitted work forgets what it
        // would have written.
        for (Task &t : c.queue)
            for (Buffer *b : t.buffers) {
                b->queued--;
                b->init.clear();
                b->host_cur.clear();
                b->dev_cur.clear();
            }
        c.queue.clear();
        c.queued_threads = 0;
        c.qstats.pending_tasks = 0;
    });
}

int fort_scope_graph_queue_stats_get_v1(fort_scope_t h, fort_scope_graph_queue_stats_v1 *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing statistics output");
        *stats = c.qstats;
    }, Mode::Query);
}

int fort_scope_graph_stats_get_v1(fort_scope_t h, fort_scope_graph_stats_v1 *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing statistics output");
        *stats = c.gstats;
    }, Mode::Query);
}

#ifndef FTC_CPU_BACKEND
// The views an external worker receives of its bound buffers (scoped_runtime.h:247-263).
static void external_views(Context &c, const std::vector<fort_scope_plan_binding> &plan,
                           std::vector<fort_scope_external_view_v1> &views, std::vector<std::vector<std::size_t>> &held)
{
    views.resize(plan.size());
    held.assign(plan.size(), {});
    for (std::size_t i = 0; i < plan.size(); ++i) {
        Buffer &b = buffer(c, plan[i].buffer);
        std::size_t s[kMaxRank];
        strides(b, s);
        for (uint32_t d = 0; d < b.rank; ++d) held[i].push_back(s[d] * b.elem);
        views[i] = fort_scope_external_view_v1{b.handle, b.generation, b.type, b.rank, b.elem, b.bytes,
                                               b.extents, held[i].data(), b.lower_bounds, b.device};
    }
}

// When the context holds its work for whole-step graphs (FORT_SCOPE_GRAPH_QUEUE_WORK=0), an external
// worker runs under stream capture instead of on the context stream: nothing executes, and what it
// enqueued becomes one task of the queue (a child-graph node of the step's chain graph), committed like a
// kernel task. The key of the task is the captured topology (node types, kernels and launch shapes), so a
// step that repeats replays its graph and only refreshes the captured node's parameters. Returns false,
// with nothing changed, where the work cannot be captured (a worker that synchronises, reads back or
// declines, an empty capture): the caller then runs the worker directly on the context stream.
static bool capture_external(Context &c, const std::vector<fort_scope_plan_binding> &plan,
                             fort_scope_external_worker_v1 worker, void *user, fort_scope_external_report_v1 *report)
{
    if (!c.graphs_enabled || !c.queue_enabled || c.queue_work != 0 || c.poisoned) return false;
    ensure_stream(c);
    dev::set(c.domain->device);
    if (!c.capture_stream &&
        hipStreamCreateWithFlags(&c.capture_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        c.capture_stream = nullptr;
        return false;
    }
    std::vector<void *> devices(plan.size());
    prepare_device(c, plan.data(), plan.size(), devices.data(), false);
    auto cancel = [&] {
        for (auto &p : plan) end_access(c, buffer(c, p.buffer), true, true);
    };
    std::vector<fort_scope_external_view_v1> views;
    std::vector<std::vector<std::size_t>> held;
    external_views(c, plan, views, held);
    auto work = std::make_shared<CapturedWork>();
    if (hipStreamBeginCapture(c.capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        cancel();
        return false;
    }
    int started = -1, rc = FORT_SCOPE_EXECUTION;
    try {
        rc = worker(user, views.data(), views.size(), reinterpret_cast<void *>(c.capture_stream), &started);
    } catch (...) {
        rc = FORT_SCOPE_EXECUTION;
    }
    const hipError_t ended = hipStreamEndCapture(c.capture_stream, &work->graph);
    (void)hipGetLastError();
    std::size_t count = 0;
    bool usable = rc == FORT_SCOPE_OK && started == 1 && ended == hipSuccess && work->graph &&
                  hipGraphGetNodes(work->graph, nullptr, &count) == hipSuccess && count > 0;
    Task task;
    if (usable) {
        std::vector<hipGraphNode_t> nodes(count);
        usable = hipGraphGetNodes(work->graph, nodes.data(), &count) == hipSuccess;
        uint64_t key = 0x6a09e667f3bcc909ull;
        for (std::size_t i = 0; usable && i < count; ++i) {
            hipGraphNodeType type;
            usable = hipGraphNodeGetType(nodes[i], &type) == hipSuccess;
            key = fnv(key, &type, sizeof type);
            if (usable && type == hipGraphNodeTypeKernel) {
                hipKernelNodeParams kp{};
                usable = hipGraphKernelNodeGetParams(nodes[i], &kp) == hipSuccess;
                const uint32_t shape[7] = {kp.gridDim.x, kp.gridDim.y, kp.gridDim.z, kp.blockDim.x, kp.blockDim.y,
                                           kp.blockDim.z, kp.sharedMemBytes};
                key = fnv(key, &kp.func, sizeof kp.func);
                key = fnv(key, shape, sizeof shape);
            }
        }
        (void)hipGetLastError();
        task.code_key = key;
    }
    if (!usable) {
        // Nothing ran under capture: the accesses are withdrawn and the worker runs directly.
        cancel();
        return false;
    }
    task.captured = work;
    task.threads = 1;
    for (const auto &p : plan) {
        Buffer &b = buffer(c, p.buffer);
        b.queued++;
        task.buffers.push_back(&b);
    }
    for (const auto &p : plan) end_access(c, buffer(c, p.buffer), true, false);
    c.qstats.accepted_tasks++;
    c.queue.push_back(std::move(task));
    c.queued_threads += 1;
    c.qstats.pending_tasks = c.queue.size();
    if (c.queue.size() > c.qstats.peak_pending_tasks) c.qstats.peak_pending_tasks = c.queue.size();
    if (report) {
        report->started = 1;
        report->committed = 1;
    }
    return true;
}
#endif

int fort_scope_external_execute_v1(fort_scope_t h, const fort_scope_external_call_v1 *call,
                                   fort_scope_external_worker_v1 worker, void *user,
                                   fort_scope_external_report_v1 *report)
{
    if (report) {
        report->version = 1;
        report->status = FORT_SCOPE_OK;
        report->started = 0;
        report->committed = 0;
    }
    int status = with(h, [&](Context &c) {
        require(call && worker, FORT_SCOPE_ARGUMENT, "missing external call or worker");
        require(call->version == 1, FORT_SCOPE_ARGUMENT, "unsupported external call version");
        require(!call->binding_count || call->bindings, FORT_SCOPE_ARGUMENT, "missing external bindings");
        if (!call->binding_count) return;
        std::vector<fort_scope_plan_binding> plan(call->binding_count);
        for (std::size_t i = 0; i < call->binding_count; ++i) {
            const fort_scope_external_binding_v1 &eb = call->bindings[i];
            Buffer &b = buffer(c, eb.buffer);
            require(b.type == eb.type && b.rank == eb.rank && b.elem == eb.element_bytes, FORT_SCOPE_ARGUMENT,
                    "external binding type, rank or element size mismatch");
            plan[i].buffer = eb.buffer;
            plan[i].access = eb.access;
        }
#ifndef FTC_CPU_BACKEND
        if (capture_external(c, plan, worker, user, report)) return;
#endif
        for (auto &pb : plan) drain_buffer(c, buffer(c, pb.buffer));
        std::vector<void *> devices(plan.size());
        prepare_device(c, plan.data(), plan.size(), devices.data(), false);
        std::vector<fort_scope_external_view_v1> views(plan.size());
        std::vector<std::vector<std::size_t>> strides_b(plan.size());
        for (std::size_t i = 0; i < plan.size(); ++i) {
            Buffer &b = buffer(c, plan[i].buffer);
            std::size_t s[kMaxRank];
            strides(b, s);
            for (uint32_t d = 0; d < b.rank; ++d) strides_b[i].push_back(s[d] * b.elem);
            views[i] = fort_scope_external_view_v1{b.handle, b.generation, b.type, b.rank, b.elem, b.bytes,
                                                   b.extents, strides_b[i].data(), b.lower_bounds, b.device};
        }
        int started = -1;
        int rc = FORT_SCOPE_EXECUTION;
        try {
            rc = worker(user, views.data(), views.size(), reinterpret_cast<void *>(c.stream), &started);
        } catch (...) {
            rc = FORT_SCOPE_EXECUTION;
        }
        if (report) report->started = started == 1;
        if (rc == FORT_SCOPE_OK && started == 1) {
            for (auto &p : plan) end_access(c, buffer(c, p.buffer), true, false);
            if (report) report->committed = 1;
            return;
        }
        if (started == 0 && rc != FORT_SCOPE_OK) {
            for (auto &p : plan) end_access(c, buffer(c, p.buffer), true, true);
            fail(rc == FORT_SCOPE_EXECUTION ? FORT_SCOPE_BOUNDARY : rc, "external provider declined before enqueue");
        }
        for (auto &p : plan) end_access(c, buffer(c, p.buffer), true, true);
        c.poisoned = true;
        fail(FORT_SCOPE_EXECUTION, "external provider failed after it started work");
    });
    if (report) report->status = status;
    return status;
}

int fort_scope_execution_error(fort_scope_t h, const char *message)
{
    std::string m = message ? message : "execution error";
    int status = with(h, [&](Context &c) {
        c.poisoned = true;
        fail(FORT_SCOPE_EXECUTION, m);
    }, Mode::Query);
    return status == FORT_SCOPE_STALE ? status : FORT_SCOPE_EXECUTION;
}

int fort_scope_wait(fort_scope_t h)
{
    return with(h, [&](Context &c) {
        flush(c);
        wait(c);
    });
}

int fort_scope_stats_get(fort_scope_t h, fort_scope_stats *stats)
{
    return with(h, [&](Context &c) {
        require(stats != nullptr, FORT_SCOPE_ARGUMENT, "missing statistics output");
        *stats = c.domain->stats;
    }, Mode::Query);
}

int fort_scope_close(fort_scope_t h)
{
    fort_scope_domain_t owned = 0;
    int status = with(h, [&](Context &c) {
        require(!c.scratch_active, FORT_SCOPE_STATE, "a scratch lease is still active");
        for (auto &kv : c.domain->buffers)
            require(!kv.second->prepared, FORT_SCOPE_STATE, "an access is still prepared");
        flush(c);
        wait(c);
        if (tracing()) {
            // What the domain moved and ran, for a run traced with FORT_RUNTIME_TRACE.
            const fort_scope_stats &s = c.domain->stats;
            std::fprintf(stderr,
                         "FORT_SCOPED close launches=%llu uploads=%llu downloads=%llu upload_bytes=%llu "
                         "download_bytes=%llu\n",
                         (unsigned long long)s.launches, (unsigned long long)s.uploads,
                         (unsigned long long)s.downloads, (unsigned long long)s.upload_bytes,
                         (unsigned long long)s.download_bytes);
        }
        release_context(c);
        c.domain->borrower = nullptr;
        if (c.owns_domain) owned = c.domain->handle;
    });
    if (status != FORT_SCOPE_OK) return status;
    {
        std::lock_guard<std::mutex> lock(g_registry);
        g_contexts.erase(h);
    }
    if (owned) return fort_scope_domain_close_v1(owned);
    return FORT_SCOPE_OK;
}

int fort_scope_abandon(fort_scope_t h)
{
    return global([&] {
        Context &c = context(h);
        require(c.poisoned, FORT_SCOPE_STATE, "abandon is only valid after an execution failure");
        c.queue.clear();
        for (auto &kv : c.domain->buffers) {
            kv.second->prepared = 0;
            kv.second->queued = 0;
        }
        if (c.stream_ready) {
#ifndef FTC_CPU_BACKEND
            (void)hipStreamSynchronize(c.stream);
            (void)hipGetLastError();
#endif
        }
        release_context(c);
        Domain *d = c.domain;
        d->borrower = nullptr;
        d->failed = true;
        bool owns = c.owns_domain;
        g_contexts.erase(h);
        if (owns) {
            close_domain_storage(*d);
            g_domains.erase(d->handle);
        }
    });
}

}  // extern "C"
