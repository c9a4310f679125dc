// Fiber scheduler and runtime stand-ins of the CPU build (tests/hostbuild): see include/hip/hip_runtime.h.
//
// Every lane of a workgroup is a fiber with its own stack.  A lane runs until it reaches a wavefront operation or a
// workgroup barrier (or returns); once no lane can run, each wavefront's waiting lanes are resolved per operation
// site, and a workgroup barrier is released once every live lane waits at one.  A state in which nothing can be
// resolved is a deadlock (a barrier some lanes never reach) and aborts with the lanes' positions.
#include <hip/hip_runtime.h>

#include <sanitizer/common_interface_defs.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" void emu_switch(void **save, void *to);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {

enum State : uint8_t { RUNNABLE, AT_WAVE, AT_BLOCK, DONE };

struct Lane {
    void *sp = nullptr;
    char *stack = nullptr;
    unsigned tid = 0;
    State state = RUNNABLE;
    int kind = 0, arg = 0, line = 0, orValue = 0, isOr = 0;
    uint64_t bits = 0, result = 0;
    void *fake = nullptr;
};

Lane *cur = nullptr;
dim3 blockIdxV, blockDimV, gridDimV;

namespace {

constexpr std::size_t STACK = 512 * 1024;
std::vector<Lane> lanes;  // grown to the largest workgroup; the first nLanes are in use
std::size_t nLanes = 0;
void *schedSp = nullptr;
void *schedFake = nullptr;
const void *schedBottom = nullptr;
std::size_t schedSize = 0;
const std::function<void()> *body = nullptr;
unsigned char *dynLds = nullptr;
std::size_t dynLdsBytes = 0;
unsigned long long clockTicks = 0;
hipError_t lastError = hipSuccess;

[[noreturn]] void lane_main() {
    __sanitizer_finish_switch_fiber(nullptr, &schedBottom, &schedSize);
    (*body)();
    cur->state = DONE;
    __sanitizer_start_switch_fiber(nullptr, schedBottom, schedSize);
    void *dummy;
    emu_switch(&dummy, schedSp);
    std::abort();
}

void start_lane(Lane &l) {
    if (!l.stack) {
        l.stack = static_cast<char *>(std::malloc(STACK));
        if (!l.stack) {
            std::fprintf(stderr, "hostbuild: no memory for lane stacks\n");
            std::abort();
        }
    }
    // initial frame: six callee-saved registers, then lane_main as the return address of emu_switch
    uintptr_t top = (reinterpret_cast<uintptr_t>(l.stack) + STACK) & ~uintptr_t(15);
    void **sp = reinterpret_cast<void **>(top);
    *--sp = nullptr;                                  // lane_main's own return slot (never used)
    *--sp = reinterpret_cast<void *>(&lane_main);     // popped by ret
    for (int i = 0; i < 6; i++) *--sp = nullptr;      // r15 r14 r13 r12 rbx rbp
    l.sp = sp;
    l.state = RUNNABLE;
}

void resume(Lane &l) {
    cur = &l;
    __sanitizer_start_switch_fiber(&schedFake, l.stack, STACK);
    emu_switch(&schedSp, l.sp);
    __sanitizer_finish_switch_fiber(schedFake, nullptr, nullptr);
    cur = nullptr;
}

// from a lane: back to the scheduler until resolved
void yield() {
    Lane &l = *cur;
    __sanitizer_start_switch_fiber(&l.fake, schedBottom, schedSize);
    emu_switch(&l.sp, schedSp);
    __sanitizer_finish_switch_fiber(l.fake, &schedBottom, &schedSize);
}

[[noreturn]] void deadlock(const char *why) {
    std::fprintf(stderr, "hostbuild: %s in block (%u,%u,%u)\n", why, blockIdxV.x, blockIdxV.y, blockIdxV.z);
    for (std::size_t i = 0; i < nLanes; i++) {
        const Lane &l = lanes[i];
        if (l.state == AT_WAVE || l.state == AT_BLOCK)
            std::fprintf(stderr, "  lane %zu: %s kind %d line %d\n", i, l.state == AT_WAVE ? "wave op" : "barrier",
                         l.kind, l.line);
    }
    std::abort();
}

void resolve_wave_group(Lane *const *g, int n) {
    const int kind = g[0]->kind;
    Lane *byLane[64] = {};
    for (int i = 0; i < n; i++) byLane[g[i]->tid & 63] = g[i];
    for (int i = 1; i < n; i++)
        if (g[i]->kind == W_READLANE && g[i]->arg != g[0]->arg) deadlock("non-uniform readlane index");
    uint64_t mask = 0;
    for (int i = 0; i < n; i++)
        if (g[i]->bits) mask |= 1ull << (g[i]->tid & 63);
    for (int i = 0; i < n; i++) {
        Lane &l = *g[i];
        const unsigned me = l.tid & 63;
        int src = -1;
        switch (kind) {
            case W_BALLOT: l.result = mask; break;
            case W_SHFL: src = l.arg & 63; break;
            case W_SHFL_DOWN: src = (int)me + l.arg < 64 ? (int)me + l.arg : (int)me; break;
            case W_SHFL_UP: src = (int)me - l.arg >= 0 ? (int)me - l.arg : (int)me; break;
            case W_SHFL_XOR: src = (int)(me ^ (unsigned)l.arg) & 63; break;
            case W_READLANE: src = l.arg & 63; break;
            case W_READFIRST: src = (int)(g[0]->tid & 63); break;
            default: l.result = 0; break;
        }
        if (src >= 0) {
            if (!byLane[src]) deadlock("a wavefront operation reads a lane that does not take part");
            l.result = byLane[src]->bits;
        }
        l.state = RUNNABLE;
    }
}

// Resolves the wavefront operations of wave w (lanes [w, w + 64)) once none of its lanes can run: its waiting lanes
// grouped by (operation, source line).  Returns true if some lane became runnable.
bool resolve_wave(std::size_t w) {
    const std::size_t e = std::min(nLanes, w + 64);
    Lane *waiting[64];
    int nw = 0;
    for (std::size_t i = w; i < e; i++) {
        if (lanes[i].state == RUNNABLE) return false;
        if (lanes[i].state == AT_WAVE) waiting[nw++] = &lanes[i];
    }
    bool used[64] = {};
    for (int a = 0; a < nw; a++) {
        if (used[a]) continue;
        Lane *grp[64];
        int ng = 0;
        for (int b = a; b < nw; b++)
            if (!used[b] && waiting[b]->kind == waiting[a]->kind && waiting[b]->line == waiting[a]->line) {
                grp[ng++] = waiting[b];
                used[b] = true;
            }
        resolve_wave_group(grp, ng);
    }
    return nw > 0;
}

// Releases a workgroup barrier once every live lane waits at one; finished once no lane is live.
bool resolve_block(bool &finished) {
    const std::size_t n = nLanes;
    bool anyLive = false;
    int isOr = -1, orAcc = 0;
    for (std::size_t i = 0; i < n; i++) {
        const Lane &l = lanes[i];
        if (l.state == DONE) continue;
        anyLive = true;
        if (l.state != AT_BLOCK) return false;
        if (isOr < 0) isOr = l.isOr;
        if (l.isOr != isOr) deadlock("lanes wait at different kinds of workgroup barrier");
        orAcc |= l.orValue;
    }
    if (!anyLive) {
        finished = true;
        return true;
    }
    for (std::size_t i = 0; i < n; i++) {
        Lane &l = lanes[i];
        if (l.state != AT_BLOCK) continue;
        l.result = (uint64_t)orAcc;
        l.state = RUNNABLE;
    }
    return true;
}

// Order in which wavefronts advance: all runnable lanes in lane order, then every wavefront's operations (the
// default); or, with EMU_SCHED_SEED set, one wavefront at a time drawn at random among those that can advance, so
// that wavefronts run ahead of each other between workgroup barriers by random amounts.
uint64_t schedState = 0;
bool schedRandom = false;

uint64_t next_random() {
    schedState += 0x9E3779B97F4A7C15ull;
    uint64_t z = schedState;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void run_block() {
    const std::size_t nt = nLanes;
    bool finished = false;
    while (!finished) {
        if (!schedRandom) {
            bool ran = false;
            for (std::size_t i = 0; i < nt; i++)
                if (lanes[i].state == RUNNABLE) {
                    resume(lanes[i]);
                    ran = true;
                }
            if (ran) continue;
            bool any = false;
            for (std::size_t w = 0; w < nt; w += 64) any = resolve_wave(w) || any;
            if (any) continue;
            if (!resolve_block(finished)) deadlock("deadlock (no lane can proceed)");
            continue;
        }
        std::size_t cand[16];
        int nc = 0;
        for (std::size_t w = 0; w < nt; w += 64) {
            const std::size_t e = std::min(nt, w + 64);
            bool runnable = false, waiting = false;
            for (std::size_t i = w; i < e; i++) {
                runnable = runnable || lanes[i].state == RUNNABLE;
                waiting = waiting || lanes[i].state == AT_WAVE;
            }
            if (runnable || waiting) cand[nc++] = w;
        }
        if (nc == 0) {
            if (!resolve_block(finished)) deadlock("deadlock (no lane can proceed)");
            continue;
        }
        const std::size_t w = cand[next_random() % (uint64_t)nc];
        bool ran = false;
        for (std::size_t i = w; i < std::min(nt, w + 64); i++)
            if (lanes[i].state == RUNNABLE) {
                resume(lanes[i]);
                ran = true;
            }
        if (!ran) resolve_wave(w);
    }
}

}  // namespace

unsigned lane_tid() { return cur->tid; }
void *dyn_lds() { return dynLds; }

uint64_t wave_op(int kind, uint64_t bits, int arg, int line) {
    Lane &l = *cur;
    l.state = AT_WAVE;
    l.kind = kind;
    l.bits = bits;
    l.arg = arg;
    l.line = line;
    yield();
    return l.result;
}

int block_op(int orValue, int isOr, int line) {
    Lane &l = *cur;
    l.state = AT_BLOCK;
    l.orValue = orValue;
    l.isOr = isOr;
    l.line = line;
    yield();
    return (int)l.result;
}

void run_grid(dim3 grid, dim3 block, std::size_t lds, const std::function<void()> &b) {
    const std::size_t nt = (std::size_t)block.x * block.y * block.z;
    if (nt == 0 || nt > 1024) {
        lastError = hipErrorInvalidValue;
        return;
    }
    if (lds > 160 * 1024) {
        std::fprintf(stderr, "hostbuild: a launch asks for %zu bytes of dynamic LDS (more than 160 KiB)\n", lds);
        std::abort();
    }
    if (lanes.size() < nt) lanes.resize(nt);
    nLanes = nt;
    static bool seeded = false;
    if (!seeded) {
        seeded = true;
        if (const char *sd = std::getenv("EMU_SCHED_SEED")) {
            schedRandom = true;
            schedState = std::strtoull(sd, nullptr, 10);
        }
    }
    body = &b;
    blockDimV = block;
    gridDimV = grid;
    dynLdsBytes = lds;
    dynLds = static_cast<unsigned char *>(std::malloc(lds ? lds : 1));
    for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
            for (unsigned bx = 0; bx < grid.x; bx++) {
                blockIdxV = dim3(bx, by, bz);
                // LDS starts with no defined contents: fill it with a pattern that makes stale reads visible
                std::memset(dynLds, 0xA5, lds ? lds : 1);
                for (std::size_t i = 0; i < nt; i++) {
                    lanes[i].tid = (unsigned)i;
                    start_lane(lanes[i]);
                }
                run_block();
            }
    std::free(dynLds);
    dynLds = nullptr;
    body = nullptr;
}

}  // namespace emu

EmuThreadIdx emu_thread_idx() {
    const unsigned t = emu::cur->tid;
    EmuThreadIdx r;
    r.x = t % emu::blockDimV.x;
    r.y = (t / emu::blockDimV.x) % emu::blockDimV.y;
    r.z = t / (emu::blockDimV.x * emu::blockDimV.y);
    return r;
}

unsigned long long wall_clock64() { return ++emu::clockTicks; }

hipError_t hipMalloc(void **p, std::size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, std::size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *dst, const void *src, std::size_t n, hipMemcpyKind) {
    if (n) std::memmove(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, std::size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(dst, src, n, k);
}
hipError_t hipMemset(void *p, int v, std::size_t n) {
    if (n) std::memset(p, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *p, int v, std::size_t n, hipStream_t) { return hipMemset(p, v, n); }

struct ihipStream_t { int unused; };
struct ihipEvent_t { int unused; };
hipError_t hipStreamCreate(hipStream_t *s) {
    *s = new ihipStream_t{};
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
    *e = new ihipEvent_t{};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 0.f;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int) {
    *v = a == hipDeviceAttributeMaxSharedMemoryPerBlock ? 160 * 1024 : 100000;
    return hipSuccess;
}
hipError_t hipGetLastError() {
    hipError_t e = emu::lastError;
    emu::lastError = hipSuccess;
    return e;
}
hipError_t hipPeekAtLastError() { return emu::lastError; }
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "host build error"; }
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}
