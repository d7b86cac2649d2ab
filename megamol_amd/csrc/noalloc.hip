// noalloc.hip — the render entry point that allocates nothing (mmhrt_render_noalloc).
//
// For callers that cannot allocate device memory (work under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc): the frame goes to device buffers the caller owns, and
// the trace's scratch (the tile counter, overflow flag and instrumentation words, and the
// traversal stacks that do not fit in LDS) lives in `scratch`, a file-scope device array of fixed
// size that every GPU holds for the life of the process. Host code never passes the array's name
// to a kernel (on the host it is only a stand-in symbol): its device address is looked up per
// device with hipGetSymbolAddress when a context is created on that device, and that address is
// what the kernels receive. All calls on one GPU share its array: the counters are cleared by a
// fill kernel at the start of every call, and two calls that overlap on one GPU race.
#include <mutex>

#include "ctx.hpp"

namespace mmhrt {
namespace {

// Persistent lanes whose traversal stacks the scratch holds: 512 workgroups of 256 threads (two
// per CU on 256 CUs). The grid of a call is capped to them.
constexpr uint32_t NOALLOC_LANES = 512u * 256u;
constexpr size_t NOALLOC_WORDS = TRACE_CTR_WORDS + (size_t)NOALLOC_LANES * TRACE_SPILL_PER_LANE;  // 62.9 MB

__device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint32_t* g_scratch[MAX_DEVICES] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return NOALLOC_WORDS * sizeof(uint32_t); }

// The device address of this device's scratch array (the current device must be `device`).
int noalloc_init(mmhrt_ctx_s* c) {
    if (c->device < 0 || c->device >= MAX_DEVICES) return fail(c, MMHRT_E_UNSUPPORTED, "render_noalloc: device index");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[c->device]) return 0;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch symbol address");
    g_scratch[c->device] = static_cast<uint32_t*>(p);
    return 0;
}

int render_noalloc(mmhrt_ctx_s* c, const FrameConsts& fc, const FrameOutputs& out) {
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = c->device >= 0 && c->device < MAX_DEVICES ? g_scratch[c->device] : nullptr;
    }
    if (!base) return fail(c, MMHRT_E_STATE, "render_noalloc: the scratch of this device was not set up");
    int32_t* spill = reinterpret_cast<int32_t*>(base + TRACE_CTR_WORDS);
    return render_frame_into(c, fc, out, base, spill, NOALLOC_LANES, false);
}

}  // namespace mmhrt
