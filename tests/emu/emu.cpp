// Fiber executor for the HIP emulation header: the lanes of one workgroup are
// ucontext fibers on the calling thread, switched at barriers and wave
// exchanges; workgroups run one after another.
#include <hip/hip_runtime.h>
#include <sanitizer/common_interface_defs.h>

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
struct Fiber {
    void* rsp = nullptr;
    char* stack = nullptr;
    State st = kRun;
    EmuIdx tid;
};

struct Sched {
    std::vector<Fiber> f;
    int n = 0;
    int cur = -1;
    void* main_rsp = nullptr;
    const void* main_bottom = nullptr;
    size_t main_size = 0;
    int live = 0;
    int block_arrived = 0;
    std::vector<int> wave_live, wave_arrived;
    const std::function<void()>* fn = nullptr;
};
Sched S;

void to_main()
{
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, S.main_bottom, S.main_size);
    emu_switch(&S.f[S.cur].rsp, S.main_rsp);
    __sanitizer_finish_switch_fiber(fake, &S.main_bottom, &S.main_size);
}

void release_block_if_complete()
{
    if (S.block_arrived > 0 && S.block_arrived == S.live) {
        for (int i = 0; i < S.n; ++i)
            if (S.f[i].st == kWaitBlock) S.f[i].st = kRun;
        S.block_arrived = 0;
    }
}
void release_wave_if_complete(int w)
{
    if (S.wave_arrived[w] > 0 && S.wave_arrived[w] == S.wave_live[w]) {
        for (int i = 64 * w; i < std::min(S.n, 64 * w + 64); ++i)
            if (S.f[i].st == kWaitWave) S.f[i].st = kRun;
        S.wave_arrived[w] = 0;
    }
}

void entry()
{
    __sanitizer_finish_switch_fiber(nullptr, &S.main_bottom, &S.main_size);
    (*S.fn)();
    const int me = S.cur;
    const int w = me / 64;
    cur_block->waves[w].alive[me % 64] = false;
    S.f[me].st = kDone;
    --S.live;
    --S.wave_live[w];
    release_block_if_complete();
    release_wave_if_complete(w);
    __sanitizer_start_switch_fiber(nullptr, S.main_bottom, S.main_size);
    void* dead;
    emu_switch(&dead, S.main_rsp);
    __builtin_unreachable();
}
}  // namespace

void block_sync()
{
    ++S.block_arrived;
    if (S.block_arrived == S.live) {
        release_block_if_complete();
        return;
    }
    S.f[S.cur].st = kWaitBlock;
    to_main();
}

void wave_sync()
{
    const int w = S.cur / 64;
    ++S.wave_arrived[w];
    if (S.wave_arrived[w] == S.wave_live[w]) {
        release_wave_if_complete(w);
        return;
    }
    S.f[S.cur].st = kWaitWave;
    to_main();
}

void launch_impl(dim3 grid, dim3 block, size_t shm, const std::function<void()>& fn)
{
    const int n = block.x * block.y * block.z;
    if (n <= 0 || n > kMaxThreads) {
        std::fprintf(stderr, "emu: bad block size %d\n", n);
        std::abort();
    }
    if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
    if ((int)S.f.size() < n) {
        const int old = (int)S.f.size();
        S.f.resize(n);
        for (int i = old; i < n; ++i) S.f[i].stack = (char*)std::malloc(kStack);
    }
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                Block blk;
                const int nw = (n + 63) / 64;
                blk.waves = std::vector<Wave>(nw);
                S.wave_live.assign(nw, 0);
                S.wave_arrived.assign(nw, 0);
                for (int w = 0; w < nw; ++w) {
                    const int cnt = std::min(64, n - 64 * w);
                    S.wave_live[w] = cnt;
                    for (int l = 0; l < 64; ++l) blk.waves[w].alive[l] = l < cnt;
                }
                // uninitialised LDS holds garbage on the GPU
                unsigned char* heap = shm ? new unsigned char[shm] : nullptr;
                if (heap) std::memset(heap, 0x7f, shm);
                blk.dyn = heap;
                cur_block = &blk;
                blockIdx.x = bx; blockIdx.y = by; blockIdx.z = bz;
                blockDim.x = block.x; blockDim.y = block.y; blockDim.z = block.z;
                gridDim.x = grid.x; gridDim.y = grid.y; gridDim.z = grid.z;
                S.n = n;
                S.live = n;
                S.block_arrived = 0;
                S.fn = &fn;
                for (int t = 0; t < n; ++t) {
                    Fiber& f = S.f[t];
                    f.st = kRun;
                    f.tid.x = t % block.x;
                    f.tid.y = (t / block.x) % block.y;
                    f.tid.z = t / (block.x * block.y);
                    // initial frame: six saved registers, then the return
                    // address; entry() starts with rsp % 16 == 8 as after a call
                    uintptr_t top = ((uintptr_t)(f.stack + kStack)) & ~(uintptr_t)15;
                    void** sp = (void**)(top - 8);          // slot above the return address (alignment pad)
                    *--sp = (void*)&entry;                  // return address
                    for (int r = 0; r < 6; ++r) *--sp = nullptr;
                    f.rsp = sp;
                }
                // round-robin over runnable fibers until all are done
                int done = 0;
                while (done < n) {
                    bool progressed = false;
                    done = 0;
                    for (int t = 0; t < n; ++t) {
                        Fiber& f = S.f[t];
                        if (f.st == kDone) { ++done; continue; }
                        if (f.st != kRun) continue;
                        S.cur = t;
                        cur_lane_global = t;
                        threadIdx = f.tid;
                        void* fake = nullptr;
                        __sanitizer_start_switch_fiber(&fake, f.stack, kStack);
                        emu_switch(&S.main_rsp, f.rsp);
                        __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
                        progressed = true;
                        if (f.st == kDone) ++done;
                    }
                    if (!progressed && done < n) {
                        std::fprintf(stderr, "emu: deadlock in block (%u,%u,%u): live %d, block arrived %d\n",
                                     bx, by, bz, S.live, S.block_arrived);
                        std::abort();
                    }
                }
                delete[] heap;
                cur_block = nullptr;
            }
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
