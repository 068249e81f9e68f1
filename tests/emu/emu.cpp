// Fiber executor for the HIP emulation header: the lanes of a workgroup are
// fibers on the calling thread, switched at barriers, wave exchanges and
// waits (s_sleep). Up to EMU_WG_WINDOW workgroups of a grid (default 1) run
// interleaved, round robin over their runnable lanes, so kernels whose
// workgroups wait for one another can run, and orderings that in-order
// execution would hide can show up.
#include <hip/hip_runtime.h>
#include <sanitizer/common_interface_defs.h>

#include <chrono>
#include <map>

// Minimal x86-64 context switch: saves the callee-saved registers on the
// current stack, stores rsp, loads the other stack and returns into it.
extern "C" void emu_switch(void** save_rsp, void* new_rsp);
asm(R"(
.text
.globl emu_switch
.type emu_switch,@function
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

EmuIdx threadIdx, blockIdx, blockDim, gridDim;

namespace emu {
Block* cur_block = nullptr;
int cur_lane_global = 0;

namespace {
constexpr int kMaxThreads = 1024;
constexpr size_t kStack = 512 * 1024;

enum State { kRun, kWaitBlock, kWaitWave, kDone };
struct Ctx;
struct Fiber {
    void* rsp = nullptr;
    char* stack = nullptr;
    State st = kRun;
    EmuIdx tid;
    int local = 0;       // thread index within the workgroup
    Ctx* ctx = nullptr;
};

struct Ctx {                       // one workgroup in flight
    Block blk;
    EmuIdx bidx;
    int n = 0, live = 0, block_arrived = 0;
    std::vector<int> wave_live, wave_arrived;
    std::vector<Fiber*> f;
    unsigned char* heap = nullptr;
    int done = 0;
};

struct Sched {
    std::vector<Fiber> pool;
    Fiber* cur = nullptr;
    void* main_rsp = nullptr;
    const void* main_bottom = nullptr;
    size_t main_size = 0;
    const std::function<void()>* fn = nullptr;
};
Sched S;

void to_main()
{
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, S.main_bottom, S.main_size);
    emu_switch(&S.cur->rsp, S.main_rsp);
    __sanitizer_finish_switch_fiber(fake, &S.main_bottom, &S.main_size);
}

void release_block_if_complete(Ctx& c)
{
    if (c.block_arrived > 0 && c.block_arrived == c.live) {
        for (Fiber* f : c.f)
            if (f->st == kWaitBlock) f->st = kRun;
        c.block_arrived = 0;
    }
}
void release_wave_if_complete(Ctx& c, int w)
{
    if (c.wave_arrived[w] > 0 && c.wave_arrived[w] == c.wave_live[w]) {
        for (int i = 64 * w; i < std::min(c.n, 64 * w + 64); ++i)
            if (c.f[i]->st == kWaitWave) c.f[i]->st = kRun;
        c.wave_arrived[w] = 0;
    }
}

void entry()
{
    __sanitizer_finish_switch_fiber(nullptr, &S.main_bottom, &S.main_size);
    (*S.fn)();
    Fiber* me = S.cur;
    Ctx& c = *me->ctx;
    const int w = me->local / 64;
    c.blk.waves[w].alive[me->local % 64] = false;
    me->st = kDone;
    --c.live;
    --c.wave_live[w];
    release_block_if_complete(c);
    release_wave_if_complete(c, w);
    __sanitizer_start_switch_fiber(nullptr, S.main_bottom, S.main_size);
    void* dead;
    emu_switch(&dead, S.main_rsp);
    __builtin_unreachable();
}

int window_size()
{
    const char* e = std::getenv("EMU_WG_WINDOW");
    const int v = e ? std::atoi(e) : 1;
    return v < 1 ? 1 : v;
}

void start_block(Ctx& c, unsigned bx, unsigned by, unsigned bz, dim3 block, size_t shm, Fiber* fibers)
{
    const int n = block.x * block.y * block.z;
    const int nw = (n + 63) / 64;
    c.blk.waves = std::vector<Wave>(nw);
    c.blk.shvars.clear();
    c.wave_live.assign(nw, 0);
    c.wave_arrived.assign(nw, 0);
    for (int w = 0; w < nw; ++w) {
        const int cnt = std::min(64, n - 64 * w);
        c.wave_live[w] = cnt;
        for (int l = 0; l < 64; ++l) c.blk.waves[w].alive[l] = l < cnt;
    }
    // uninitialised LDS holds garbage on the GPU
    c.heap = shm ? new unsigned char[shm] : nullptr;
    if (c.heap) std::memset(c.heap, 0x7f, shm);
    c.blk.dyn = c.heap;
    c.bidx.x = bx; c.bidx.y = by; c.bidx.z = bz;
    c.n = n;
    c.live = n;
    c.block_arrived = 0;
    c.done = 0;
    c.f.assign(n, nullptr);
    for (int t = 0; t < n; ++t) {
        Fiber& f = fibers[t];
        c.f[t] = &f;
        f.ctx = &c;
        f.local = t;
        f.st = kRun;
        f.tid.x = t % block.x;
        f.tid.y = (t / block.x) % block.y;
        f.tid.z = t / (block.x * block.y);
        // initial frame: six saved registers, then the return address;
        // entry() starts with rsp % 16 == 8 as after a call
        uintptr_t top = ((uintptr_t)(f.stack + kStack)) & ~(uintptr_t)15;
        void** sp = (void**)(top - 8);
        *--sp = (void*)&entry;
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        f.rsp = sp;
    }
}
}  // namespace

void block_sync()
{
    Ctx& c = *S.cur->ctx;
    ++c.block_arrived;
    if (c.block_arrived == c.live) {
        release_block_if_complete(c);
        return;
    }
    S.cur->st = kWaitBlock;
    to_main();
}

void wave_sync()
{
    Ctx& c = *S.cur->ctx;
    const int w = S.cur->local / 64;
    ++c.wave_arrived[w];
    if (c.wave_arrived[w] == c.wave_live[w]) {
        release_wave_if_complete(c, w);
        return;
    }
    S.cur->st = kWaitWave;
    to_main();
}

void yield()
{
    if (S.cur) to_main();
}

unsigned long long realtime()
{
    using clk = std::chrono::steady_clock;
    return (unsigned long long)(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                    clk::now().time_since_epoch()).count() / 10);
}

void* shared_var(int id, size_t bytes)
{
    auto& m = cur_block->shvars;
    auto it = m.find(id);
    if (it == m.end()) {
        std::vector<unsigned char> v(bytes + 16, 0x7f);
        it = m.emplace(id, std::move(v)).first;
    }
    uintptr_t p = (uintptr_t)it->second.data();
    return (void*)((p + 15) & ~(uintptr_t)15);
}

void launch_impl(dim3 grid, dim3 block, size_t shm, const std::function<void()>& fn)
{
    const int n = block.x * block.y * block.z;
    if (n <= 0 || n > kMaxThreads) {
        std::fprintf(stderr, "emu: bad block size %d\n", n);
        std::abort();
    }
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    const long total = (long)grid.x * grid.y * grid.z;
    const int win = (int)std::min<long>(window_size(), total);
    const size_t need = (size_t)n * win;
    if (S.pool.size() < need) {
        const size_t old = S.pool.size();
        // fibers hold pointers into the pool: grow it only while idle
        S.pool.resize(need);
        for (size_t i = old; i < need; ++i) S.pool[i].stack = (char*)std::malloc(kStack);
    }
    S.fn = &fn;
    blockDim.x = block.x; blockDim.y = block.y; blockDim.z = block.z;
    gridDim.x = grid.x; gridDim.y = grid.y; gridDim.z = grid.z;
    std::vector<Ctx> ctx(win);
    std::vector<bool> busy(win, false);
    long next = 0, finished = 0;
    auto block_coords = [&](long id, unsigned& bx, unsigned& by, unsigned& bz) {
        bx = (unsigned)(id % grid.x);
        by = (unsigned)((id / grid.x) % grid.y);
        bz = (unsigned)(id / ((long)grid.x * grid.y));
    };
    while (finished < total) {
        // fill free slots with the next workgroups, in order
        for (int s = 0; s < win && next < total; ++s) {
            if (busy[s]) continue;
            unsigned bx, by, bz;
            block_coords(next++, bx, by, bz);
            start_block(ctx[s], bx, by, bz, block, shm, &S.pool[(size_t)s * n]);
            busy[s] = true;
        }
        bool progressed = false;
        for (int s = 0; s < win; ++s) {
            if (!busy[s]) continue;
            Ctx& c = ctx[s];
            for (int t = 0; t < n; ++t) {
                Fiber& f = *c.f[t];
                if (f.st != kRun) continue;
                S.cur = &f;
                cur_block = &c.blk;
                cur_lane_global = t;
                threadIdx = f.tid;
                blockIdx = c.bidx;
                void* fake = nullptr;
                __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
                emu_switch(&S.main_rsp, f.rsp);
                __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
                progressed = true;
            }
            int done = 0;
            for (int t = 0; t < n; ++t) done += c.f[t]->st == kDone;
            if (done == n) {
                delete[] c.heap;
                c.heap = nullptr;
                c.blk.shvars.clear();
                busy[s] = false;
                ++finished;
            }
        }
        if (!progressed && finished < total) {
            std::fprintf(stderr, "emu: deadlock: no runnable lane in the window\n");
            std::abort();
        }
    }
    S.cur = nullptr;
    cur_block = nullptr;
}
}  // namespace emu

hipError_t hipMalloc(void** p, size_t bytes)
{
    *p = bytes ? std::malloc(bytes) : nullptr;
    if (*p) std::memset(*p, 0x7f, bytes);
    return (*p || !bytes) ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p)
{
    std::free(p);
    return 0;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
