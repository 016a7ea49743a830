// runtime.hip -- C-ABI runtime: devices, queues, memory, events, graphs,
// profiling and the tuning cache.
#include "runtime.hpp"
#include "device.hpp"
#include "fault.hpp"

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <atomic>
#include <mutex>
#include <sstream>

namespace bcmi {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return BCMI_MEM_ALLOCATION_FAILURE;
    if (e == hipErrorInvalidDevice || e == hipErrorNoDevice) return BCMI_DEVICE_NOT_FOUND;
    if (e == hipErrorInvalidValue) return BCMI_INVALID_VALUE;
    return BCMI_RUNTIME_ERROR;
}

int dtype_size(int dt)
{
    switch (dt) {
    case BCMI_I8: case BCMI_U8: return 1;
    case BCMI_I16: case BCMI_U16: return 2;
    case BCMI_I32: case BCMI_U32: case BCMI_F32: return 4;
    case BCMI_I64: case BCMI_U64: case BCMI_F64: return 8;
    default: return 0;
    }
}

bool dtype_valid(int dt) { return dtype_size(dt) != 0; }

int use_queue(bcmi_queue q)
{
    BCMI_CHECK(q != nullptr, BCMI_INVALID_COMMAND_QUEUE, "null queue");
    BCMI_HIP(hipSetDevice(q->device));
    return 0;
}

static size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int scratch(bcmi_queue q, size_t bytes, void **ptr)
{
    bytes = round_up(bytes < 256 ? 256 : bytes, 256);
    if (bytes > q->scratch_bytes) {
        BCMI_CHECK(!q->capturing, BCMI_OUT_OF_RESOURCES,
                   "scratch arena must grow to %zu bytes during graph capture; "
                   "call bcmi_queue_reserve before capturing", bytes);
        BCMI_HIP(hipStreamSynchronize(q->stream));
        if (q->scratch) BCMI_HIP(hipFree(q->scratch));
        q->scratch = nullptr;
        q->scratch_bytes = 0;
        size_t grow = round_up(bytes + bytes / 8, 1 << 20);
        BCMI_HIP(hipMalloc(&q->scratch, grow));
        q->scratch_bytes = grow;
        q->arena_gen++;
    }
    *ptr = q->scratch;
    return 0;
}

__global__ void k_zero_u64(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

static int zero_status(bcmi_queue q)
{
    if (!q->status_words) return 0;
    size_t blocks = (q->status_words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)blocks), dim3(256), 0, q->stream, q->status, q->status_words);
    BCMI_HIP(hipGetLastError());
    q->lookback_launches = 0;
    return 0;
}

int status_array(bcmi_queue q, size_t words, uint64_t **ptr)
{
    if (words < 1024) words = 1024;
    if (words > q->status_words) {
        BCMI_CHECK(!q->capturing, BCMI_OUT_OF_RESOURCES,
                   "look-back state must grow during graph capture; "
                   "call bcmi_queue_reserve before capturing");
        BCMI_HIP(hipStreamSynchronize(q->stream));
        if (q->status) BCMI_HIP(hipFree(q->status));
        q->status = nullptr;
        q->status_words = 0;
        size_t grow = round_up(words + words / 8, 4096);
        BCMI_HIP(hipMalloc(&q->status, grow * sizeof(uint64_t)));
        q->status_words = grow;
        q->arena_gen++;
        BCMI_TRY(zero_status(q));
    }
    *ptr = q->status;
    return 0;
}

__global__ void k_epoch_advance(Ctrl *ctrl, uint32_t count)
{
    if (threadIdx.x == 0) {
        uint32_t e = __hip_atomic_load(&ctrl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&ctrl->epoch, e + count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

int begin_lookback(bcmi_queue q, int launches)
{
    // Tags carry a 30-bit epoch; clear the status words well before an
    // epoch value can repeat.
    if (q->capturing) {
        q->capture_lookback_launches += (uint64_t)launches;
    } else {
        q->lookback_launches += (uint64_t)launches;
        if (q->lookback_launches > (1ull << 29)) BCMI_TRY(zero_status(q));
    }
    hipLaunchKernelGGL(k_epoch_advance, dim3(1), dim3(64), 0, q->stream, q->ctrl, (uint32_t)launches);
    BCMI_HIP(hipGetLastError());
    return 0;
}

static hipEvent_t pool_event(bcmi_queue q)
{
    if (!q->event_pool.empty()) {
        hipEvent_t e = q->event_pool.back();
        q->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_begin(bcmi_queue q, const char *name, ProfRec *rec)
{
    rec->name = name;
    rec->start = rec->stop = nullptr;
    if (!q->profile || q->capturing) return;
    rec->start = pool_event(q);
    rec->stop = pool_event(q);
    if (rec->start) (void)hipEventRecord(rec->start, q->stream);
}

int prof_end(bcmi_queue q, ProfRec *rec)
{
    if (!rec->start) return 0;
    BCMI_HIP(hipEventRecord(rec->stop, q->stream));
    q->prof_pending.push_back(*rec);
    return 0;
}

int check_launch(const char *name)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, name, __FILE__, __LINE__);
    return 0;
}

// ------------------------------------------------------------ tuning cache
static std::mutex g_tune_mu;
static std::map<std::string, unsigned> g_tune;

unsigned tune_get(const std::string &key, const char *param, unsigned def)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key + "|" + param);
    return it == g_tune.end() ? def : it->second;
}

static int tune_autoload();

template <typename T>
__global__ void k_fill(T *dst, T v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

struct alignas(16) b16 { uint64_t a, b; };

}  // namespace bcmi

using namespace bcmi;

extern "C" {

const char *bcmi_last_error(void) { return g_err; }

int bcmi_get_version(int *major, int *minor)
{
    if (major) *major = 0;
    if (minor) *minor = 5;
    return 0;
}

int bcmi_device_count(int *count)
{
    BCMI_CHECK(count, BCMI_INVALID_VALUE, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) c = 0;
    else if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount", __FILE__, __LINE__);
    *count = c;
    return 0;
}

int bcmi_device_get_info(int device, bcmi_device_info *info)
{
    BCMI_CHECK(info, BCMI_INVALID_VALUE, "null info");
    hipDeviceProp_t p;
    BCMI_HIP(hipGetDeviceProperties(&p, device));
    memset(info, 0, sizeof(*info));
    snprintf(info->name, sizeof(info->name), "%s", p.name);
    if (!info->name[0])  // some drivers report no marketing name
        snprintf(info->name, sizeof(info->name), "%.16s, %d CUs", p.gcnArchName, p.multiProcessorCount);
    snprintf(info->arch, sizeof(info->arch), "%s", p.gcnArchName);
    info->compute_units = p.multiProcessorCount;
    info->wavefront_size = p.warpSize;
    info->clock_khz = p.clockRate;
    info->global_mem_bytes = p.totalGlobalMem;
    info->lds_bytes_per_block = p.sharedMemPerBlock;
    info->pci_bus_id = p.pciBusID;
    return 0;
}

int bcmi_queue_create(int device, bcmi_queue *queue)
{
    BCMI_CHECK(queue, BCMI_INVALID_VALUE, "null queue out-pointer");
    int n = 0;
    BCMI_TRY(bcmi_device_count(&n));
    BCMI_CHECK(device >= 0 && device < n, BCMI_DEVICE_NOT_FOUND, "device %d not found (%d devices)", device, n);
    BCMI_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    BCMI_HIP(hipGetDeviceProperties(&p, device));
    BCMI_CHECK(std::strncmp(p.gcnArchName, "gfx950", 6) == 0, BCMI_DEVICE_NOT_FOUND,
               "device %d is %s; this library is built for gfx950 only", device, p.gcnArchName);
    BCMI_TRY(tune_autoload());
    BCMI_TRY(noalloc_scratch_setup(device));
    static std::atomic<uint64_t> next_id{1};
    bcmi_queue q = new bcmi_queue_s();
    q->id = next_id.fetch_add(1);
    q->device = device;
    q->cu_count = p.multiProcessorCount;
    q->lds_per_block = p.sharedMemPerBlock;
    hipError_t e = hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc((void **)&q->ctrl, sizeof(Ctrl));
    if (e == hipSuccess) e = hipMemset(q->ctrl, 0, sizeof(Ctrl));
    if (e == hipSuccess) e = hipHostMalloc(&q->pinned, 4096, hipHostMallocDefault);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        int s = hip_fail(e, "queue create", __FILE__, __LINE__);
        bcmi_queue_destroy(q);
        return s;
    }
    *queue = q;
    return 0;
}

int bcmi_queue_destroy(bcmi_queue q)
{
    if (!q) return 0;
    (void)hipSetDevice(q->device);
    if (q->stream) (void)hipStreamSynchronize(q->stream);
    for (auto &r : q->prof_pending) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
    }
    for (auto e : q->event_pool) (void)hipEventDestroy(e);
    if (q->scratch) (void)hipFree(q->scratch);
    if (q->status) (void)hipFree(q->status);
    if (q->ctrl) (void)hipFree(q->ctrl);
    if (q->pinned) (void)hipHostFree(q->pinned);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
    return 0;
}

int bcmi_queue_finish(bcmi_queue q)
{
    BCMI_TRY(use_queue(q));
    BCMI_HIP(hipStreamSynchronize(q->stream));
    // a kernel that gave up a look-back wait or saw a stale tile counter
    uint32_t *fault = (uint32_t *)q->pinned;
    BCMI_HIP(hipMemcpyAsync(fault, &q->ctrl->fault, sizeof(uint32_t), hipMemcpyDeviceToHost, q->stream));
    BCMI_HIP(hipStreamSynchronize(q->stream));
    const char *msg;
    const int status = fault_report(*fault, &msg);
    if (status) {
        *fault = 0;
        BCMI_HIP(hipMemcpyAsync(&q->ctrl->fault, fault, sizeof(uint32_t), hipMemcpyHostToDevice, q->stream));
        BCMI_HIP(hipStreamSynchronize(q->stream));
        set_error("%s", msg);
    }
    return status;
}

// The bounded wait every look-back uses, on a word nothing ever sets, with
// the queue's spin limit lowered for the launch: the wait gives up and
// records kFaultSpin, and every kernel exits normally.
__global__ void k_fault_probe(Ctrl *ctrl)
{
    uint32_t spins = 0;
    while (__hip_atomic_load(&ctrl->never_set, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        if (spin_expired(ctrl, spins)) break;
}

__global__ void k_set_spin_limit(Ctrl *ctrl, uint32_t limit) { ctrl->spin_limit = limit; }

int bcmi_debug_fault_probe(bcmi_queue q, unsigned spin_limit)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(spin_limit >= 1 && spin_limit <= (1u << 16), BCMI_INVALID_VALUE,
               "fault probe: spin limit %u outside [1, 65536]", spin_limit);
    BCMI_LAUNCH(q, "set_spin_limit",
                hipLaunchKernelGGL(k_set_spin_limit, dim3(1), dim3(1), 0, q->stream, q->ctrl, (uint32_t)spin_limit));
    BCMI_LAUNCH(q, "fault_probe", hipLaunchKernelGGL(k_fault_probe, dim3(1), dim3(64), 0, q->stream, q->ctrl));
    BCMI_LAUNCH(q, "set_spin_limit",
                hipLaunchKernelGGL(k_set_spin_limit, dim3(1), dim3(1), 0, q->stream, q->ctrl, 0u));
    return 0;
}

int bcmi_queue_device(bcmi_queue q, int *device)
{
    BCMI_CHECK(q && device, BCMI_INVALID_VALUE, "null argument");
    *device = q->device;
    return 0;
}

void *bcmi_queue_native(bcmi_queue q) { return q ? (void *)q->stream : nullptr; }

int bcmi_queue_reserve(bcmi_queue q, size_t n, size_t elem_bytes)
{
    BCMI_TRY(use_queue(q));
    void *p;
    uint64_t *s;
    // temporaries: a ping-pong copy of keys+values plus per-block partials
    BCMI_TRY(scratch(q, n * elem_bytes + (n / 256 + 4096) * 16 + (1 << 20), &p));
    // look-back: one status word per (tile, digit) for the smallest tile
    BCMI_TRY(status_array(q, (n / 1024 + 2) * 256 + 4096, &s));
    return 0;
}

int bcmi_malloc(size_t bytes, void **ptr)
{
    BCMI_CHECK(ptr, BCMI_INVALID_VALUE, "null pointer");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    BCMI_HIP(hipMalloc(ptr, bytes));
    return 0;
}

int bcmi_free(void *ptr)
{
    if (ptr) BCMI_HIP(hipFree(ptr));
    return 0;
}

int bcmi_host_malloc(size_t bytes, void **ptr)
{
    BCMI_CHECK(ptr, BCMI_INVALID_VALUE, "null pointer");
    *ptr = nullptr;
    if (bytes == 0) return 0;
    BCMI_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return 0;
}

int bcmi_host_free(void *ptr)
{
    if (ptr) BCMI_HIP(hipHostFree(ptr));
    return 0;
}

int bcmi_memcpy(bcmi_queue q, void *dst, const void *src, size_t bytes, int kind, int blocking)
{
    BCMI_TRY(use_queue(q));
    if (bytes == 0) return 0;
    BCMI_CHECK(dst && src, BCMI_INVALID_VALUE, "null pointer in copy");
    hipMemcpyKind k = kind == BCMI_H2D ? hipMemcpyHostToDevice
                    : kind == BCMI_D2H ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    BCMI_CHECK(kind >= BCMI_H2D && kind <= BCMI_D2D, BCMI_INVALID_VALUE, "bad copy kind %d", kind);
    BCMI_HIP(hipMemcpyAsync(dst, src, bytes, k, q->stream));
    if (blocking) BCMI_HIP(hipStreamSynchronize(q->stream));
    return 0;
}

int bcmi_fill(bcmi_queue q, void *dst, const void *pattern, size_t pattern_bytes, size_t count)
{
    BCMI_TRY(use_queue(q));
    if (count == 0) return 0;
    BCMI_CHECK(dst && pattern, BCMI_INVALID_VALUE, "null pointer in fill");
    size_t blocks = (count + 255) / 256;
    if (blocks > (size_t)q->cu_count * 16) blocks = (size_t)q->cu_count * 16;
    dim3 g((unsigned)blocks), b(256);
    switch (pattern_bytes) {
    case 1: { uint8_t v; memcpy(&v, pattern, 1); BCMI_LAUNCH(q, "fill", hipLaunchKernelGGL(k_fill<uint8_t>, g, b, 0, q->stream, (uint8_t *)dst, v, count)); break; }
    case 2: { uint16_t v; memcpy(&v, pattern, 2); BCMI_LAUNCH(q, "fill", hipLaunchKernelGGL(k_fill<uint16_t>, g, b, 0, q->stream, (uint16_t *)dst, v, count)); break; }
    case 4: { uint32_t v; memcpy(&v, pattern, 4); BCMI_LAUNCH(q, "fill", hipLaunchKernelGGL(k_fill<uint32_t>, g, b, 0, q->stream, (uint32_t *)dst, v, count)); break; }
    case 8: { uint64_t v; memcpy(&v, pattern, 8); BCMI_LAUNCH(q, "fill", hipLaunchKernelGGL(k_fill<uint64_t>, g, b, 0, q->stream, (uint64_t *)dst, v, count)); break; }
    case 16: { b16 v; memcpy(&v, pattern, 16); BCMI_LAUNCH(q, "fill", hipLaunchKernelGGL(k_fill<b16>, g, b, 0, q->stream, (b16 *)dst, v, count)); break; }
    default: set_error("fill pattern of %zu bytes unsupported", pattern_bytes); return BCMI_INVALID_VALUE;
    }
    return 0;
}

int bcmi_event_create(bcmi_event *event)
{
    BCMI_CHECK(event, BCMI_INVALID_VALUE, "null event out-pointer");
    hipEvent_t e;
    BCMI_HIP(hipEventCreate(&e));
    *event = (bcmi_event)e;
    return 0;
}

int bcmi_event_destroy(bcmi_event event)
{
    if (event) BCMI_HIP(hipEventDestroy((hipEvent_t)event));
    return 0;
}

int bcmi_event_record(bcmi_event event, bcmi_queue q)
{
    BCMI_TRY(use_queue(q));
    BCMI_HIP(hipEventRecord((hipEvent_t)event, q->stream));
    return 0;
}

int bcmi_event_synchronize(bcmi_event event)
{
    BCMI_HIP(hipEventSynchronize((hipEvent_t)event));
    return 0;
}

int bcmi_event_elapsed_ms(bcmi_event start, bcmi_event end, float *ms)
{
    BCMI_CHECK(ms, BCMI_INVALID_VALUE, "null output");
    BCMI_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
    return 0;
}

int bcmi_graph_begin(bcmi_queue q)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(!q->capturing, BCMI_INVALID_OPERATION, "queue is already capturing");
    BCMI_HIP(hipStreamBeginCapture(q->stream, hipStreamCaptureModeThreadLocal));
    q->capturing = true;
    q->capture_lookback_launches = 0;
    return 0;
}

int bcmi_graph_end(bcmi_queue q, bcmi_graph *graph)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(q->capturing, BCMI_INVALID_OPERATION, "queue is not capturing");
    BCMI_CHECK(graph, BCMI_INVALID_VALUE, "null graph out-pointer");
    q->capturing = false;
    bcmi_graph g = new bcmi_graph_s();
    hipError_t e = hipStreamEndCapture(q->stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        int s = hip_fail(e, "graph end", __FILE__, __LINE__);
        bcmi_graph_destroy(g);
        return s;
    }
    g->lookback_launches = q->capture_lookback_launches;
    g->queue_id = q->id;
    g->arena_gen = q->arena_gen;
    *graph = g;
    return 0;
}

int bcmi_graph_launch(bcmi_graph g, bcmi_queue q)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(g && g->exec, BCMI_INVALID_VALUE, "null graph");
    // the graph's kernels hold the capture queue's control block, scratch
    // arena and status array by address
    BCMI_CHECK(g->queue_id == q->id, BCMI_INVALID_COMMAND_QUEUE,
               "graph launched on a queue other than the one it was captured on");
    BCMI_CHECK(g->arena_gen == q->arena_gen, BCMI_INVALID_OPERATION,
               "the queue's scratch arena was reallocated after this graph was captured (a later call "
               "needed more); reserve enough with bcmi_queue_reserve before capturing, then capture again");
    BCMI_HIP(hipGraphLaunch(g->exec, q->stream));
    q->lookback_launches += g->lookback_launches;
    if (q->lookback_launches > (1ull << 29)) BCMI_TRY(zero_status(q));
    return 0;
}

int bcmi_graph_destroy(bcmi_graph g)
{
    if (!g) return 0;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return 0;
}

int bcmi_tune_get(const char *key, const char *param, unsigned def, unsigned *value)
{
    BCMI_CHECK(key && param && value, BCMI_INVALID_VALUE, "null argument");
    *value = tune_get(key, param, def);
    return 0;
}

int bcmi_tune_set(const char *key, const char *param, unsigned value)
{
    BCMI_CHECK(key && param, BCMI_INVALID_VALUE, "null argument");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[std::string(key) + "|" + param] = value;
    return 0;
}

// The cache file is a flat JSON object {"key|param": value, ...}, the
// analogue of ~/.boost_compute/tune/<device>.json
// (detail/parameter_cache.hpp:106-210).
int bcmi_tuning_cache_load(const char *path)
{
    return tune_load(path, true);
}

}  // extern "C"

namespace bcmi {

// Parses the cache file; `override` false keeps entries already set (the
// automatic load at the first queue must not undo an explicit bcmi_tune_set).
int tune_load(const char *path, bool override)
{
    BCMI_CHECK(path, BCMI_INVALID_VALUE, "null path");
    std::ifstream f(path);
    BCMI_CHECK(f.good(), BCMI_INVALID_VALUE, "cannot open tuning cache %s", path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    std::map<std::string, unsigned> parsed;
    size_t i = 0;
    while (true) {
        size_t a = s.find('"', i);
        if (a == std::string::npos) break;
        size_t b = s.find('"', a + 1);
        BCMI_CHECK(b != std::string::npos, BCMI_INVALID_VALUE, "malformed tuning cache");
        size_t c = s.find(':', b);
        BCMI_CHECK(c != std::string::npos, BCMI_INVALID_VALUE, "malformed tuning cache");
        char *end = nullptr;
        unsigned long v = strtoul(s.c_str() + c + 1, &end, 10);
        BCMI_CHECK(end != s.c_str() + c + 1, BCMI_INVALID_VALUE, "malformed tuning cache value");
        parsed[s.substr(a + 1, b - a - 1)] = (unsigned)v;
        i = (size_t)(end - s.c_str());
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &kv : parsed)
        if (override || !g_tune.count(kv.first)) g_tune[kv.first] = kv.second;
    return 0;
}

// Once per process, at the first queue: the file BCMI_TUNING_CACHE names
// (an error if it cannot be read), else tuning/gfx950.json beside this
// library when present -- the autotuner's output (compute_amd/autotune.py),
// in the role of the reference's per-device cache file loaded by
// parameter_cache (detail/parameter_cache.hpp:106-138).
static int tune_autoload()
{
    static std::once_flag once;
    static int status = 0;
    std::call_once(once, [] {
        const char *env = std::getenv("BCMI_TUNING_CACHE");
        if (env && *env) {
            status = tune_load(env, false);
            return;
        }
        Dl_info info;
        if (!dladdr((void *)&tune_autoload, &info) || !info.dli_fname) return;
        std::string dir(info.dli_fname);
        const size_t slash = dir.rfind('/');
        dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
        const std::string path = dir + "/tuning/gfx950.json";
        if (std::ifstream(path).good()) status = tune_load(path.c_str(), false);
    });
    return status;
}

}  // namespace bcmi

extern "C" {

int bcmi_tuning_cache_save(const char *path)
{
    BCMI_CHECK(path, BCMI_INVALID_VALUE, "null path");
    std::ofstream f(path);
    BCMI_CHECK(f.good(), BCMI_INVALID_VALUE, "cannot write tuning cache %s", path);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    f << "{\n";
    size_t k = 0;
    for (auto &kv : g_tune) f << "  \"" << kv.first << "\": " << kv.second << (++k < g_tune.size() ? ",\n" : "\n");
    f << "}\n";
    return 0;
}

int bcmi_profile_enable(bcmi_queue q, int enable)
{
    BCMI_CHECK(q, BCMI_INVALID_COMMAND_QUEUE, "null queue");
    q->profile = enable != 0;
    return 0;
}

static int drain_profile(bcmi_queue q)
{
    if (q->prof_pending.empty()) return 0;
    BCMI_HIP(hipStreamSynchronize(q->stream));
    for (auto &r : q->prof_pending) {
        float ms = 0.f;
        BCMI_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
        auto &acc = q->prof_acc[r.name];
        acc.first += 1;
        acc.second += ms;
        q->event_pool.push_back(r.start);
        q->event_pool.push_back(r.stop);
    }
    q->prof_pending.clear();
    return 0;
}

int bcmi_profile_query(bcmi_queue q, const char *kernel, int *launches, double *total_ms)
{
    BCMI_TRY(use_queue(q));
    BCMI_CHECK(kernel, BCMI_INVALID_VALUE, "null kernel name");
    BCMI_TRY(drain_profile(q));
    auto it = q->prof_acc.find(kernel);
    if (launches) *launches = it == q->prof_acc.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == q->prof_acc.end() ? 0.0 : it->second.second;
    return 0;
}

int bcmi_profile_reset(bcmi_queue q)
{
    BCMI_TRY(use_queue(q));
    BCMI_TRY(drain_profile(q));
    q->prof_acc.clear();
    return 0;
}

}  // extern "C"
