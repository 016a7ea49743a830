// runtime.hpp -- host-side internals shared by the C-ABI entry points.
//
// A bcmi_queue is the MI355X counterpart of the reference's command_queue
// (command_queue.hpp): one in-order HIP stream plus everything the
// algorithms need between calls on that stream:
//   * a grow-only scratch arena for temporaries.  The reference allocates a
//     vector per call and relies on OpenCL's deferred release
//     (radix_sort.hpp:333-336); HIP frees synchronously, so the arena is
//     reused across calls and only grows: a call that needs more syncs the
//     stream, frees the old arena and allocates a larger one.  Each such
//     reallocation (of the arena or the status array) bumps arena_gen; a
//     graph records the generation it was captured under and refuses to
//     replay once it changed, since its kernels hold the freed addresses;
//   * a look-back status array (decoupled look-back scans and onesweep
//     digit offsets).  Status words are tagged with an epoch that the last
//     block of every look-back kernel advances on the device, so the array
//     is never cleared between calls and graph replays stay correct;
//   * a small device control block: the look-back epoch, a fault flag and
//     the radix digit offsets of the current sort.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/bcmi.h"

namespace bcmi {

// Device control block, zero-initialised at queue creation.
struct Ctrl {
    uint32_t epoch;          // look-back epoch counter (advanced by k_epoch_advance)
    uint32_t fault;          // set by a kernel whose look-back wait expired or whose output index was out of range
    uint32_t spin_limit;     // polls before a look-back wait gives up; 0: kSpinLimit (tests lower it)
    uint32_t never_set;      // written by no kernel and no host code: the fault probe's wait target
    uint32_t offsets[8][256];// exclusive digit offsets per radix pass
};

constexpr uint32_t kEpochMask = (1u << 30) - 1u;

struct ProfRec {
    const char *name;
    hipEvent_t start, stop;
};

}  // namespace bcmi

struct bcmi_queue_s {
    uint64_t id = 0;  // unique per queue created in this process
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    size_t lds_per_block = 0;
    // scratch arena
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // look-back status words (uint64)
    uint64_t *status = nullptr;
    size_t status_words = 0;
    uint64_t lookback_launches = 0;
    // bumped whenever scratch or status is reallocated
    uint64_t arena_gen = 0;
    // control block
    bcmi::Ctrl *ctrl = nullptr;
    // pinned host staging for small blocking readbacks
    void *pinned = nullptr;
    bool capturing = false;
    uint64_t capture_lookback_launches = 0;
    // profiling
    bool profile = false;
    std::vector<bcmi::ProfRec> prof_pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<int, double>> prof_acc;
};

struct bcmi_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t lookback_launches = 0;
    uint64_t queue_id = 0;   // the queue it was captured on
    uint64_t arena_gen = 0;  // that queue's arena generation at capture
};

namespace bcmi {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define BCMI_HIP(call)                                                       \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::bcmi::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define BCMI_CHECK(cond, code, ...)                                          \
    do {                                                                     \
        if (!(cond)) { ::bcmi::set_error(__VA_ARGS__); return (code); }     \
    } while (0)

#define BCMI_TRY(expr)                                                       \
    do {                                                                     \
        int _s = (expr);                                                     \
        if (_s != 0) return _s;                                              \
    } while (0)

int dtype_size(int dtype);
bool dtype_valid(int dtype);

// Make the queue's device current for the calling thread.
int use_queue(bcmi_queue q);
// Scratch arena of at least `bytes` (256-byte aligned). Grows only outside
// graph capture.
int scratch(bcmi_queue q, size_t bytes, void **ptr);
// Look-back status array of at least `words` uint64.
int status_array(bcmi_queue q, size_t words, uint64_t **ptr);
// Reserve epochs for the next `launches` look-back launches on the queue:
// enqueues the epoch advance and does the wrap bookkeeping.
int begin_lookback(bcmi_queue q, int launches);

// Profiling brackets around a kernel launch on q->stream.
void prof_begin(bcmi_queue q, const char *name, ProfRec *rec);
int prof_end(bcmi_queue q, ProfRec *rec);
int check_launch(const char *name);

// Queue scratch a radix sort of n keys takes (radix_sort.hip); temporaries
// kept beside a sort go after it.
size_t radix_scratch_bytes(size_t n, size_t key_bytes, size_t value_bytes, int cu_count);

// Looks up the calling device's copy of the allocation-free sort's scratch
// (radix_sort.hip); bcmi_queue_create calls it once per queue.
int noalloc_scratch_setup(int device);

// Persistent grid sizing: blocks of `threads` that fill the device.
inline int cu_blocks(bcmi_queue q, int per_cu) { return q->cu_count * per_cu; }

// Tuning cache (parameter_cache analogue).
unsigned tune_get(const std::string &key, const char *param, unsigned def);
int tune_load(const char *path, bool override);

}  // namespace bcmi

// Launch with profiling brackets and error check.
#define BCMI_LAUNCH(q, name, ...)                                            \
    do {                                                                     \
        ::bcmi::ProfRec _rec;                                                \
        ::bcmi::prof_begin((q), (name), &_rec);                              \
        __VA_ARGS__;                                                         \
        BCMI_TRY(::bcmi::check_launch(name));                                \
        BCMI_TRY(::bcmi::prof_end((q), &_rec));                              \
    } while (0)
