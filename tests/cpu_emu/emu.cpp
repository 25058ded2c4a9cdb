// Host runtime of the CPU emulation declared in include/hip/hip_runtime.h.
#include <hip/hip_runtime.h>

#include <atomic>
#include <barrier>
#include <chrono>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace emu {
thread_local ThreadCtx tctx;

struct WaveSync {
    std::barrier<> bar;
    uint64_t buf[64];
    explicit WaveSync(int n) : bar(n) {}
};

struct Block {
    std::barrier<> bar;
    std::vector<std::unique_ptr<WaveSync>> waves;
    std::vector<char> shmem;
    explicit Block(int n) : bar(n) {}
};

static std::mutex g_atomic;
static std::atomic<long long> g_launch_id{0};
static std::atomic<long long> g_watch_id{-1};
static std::atomic<double> g_watch_t0{0};
static const char *g_watch_name = "";

static double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static void watchdog()
{
    static std::once_flag once;
    std::call_once(once, [] {
        std::thread([] {
            const char *env = getenv("EMU_TIMEOUT");
            const double lim = env ? atof(env) : 120.0;
            for (;;) {
                std::this_thread::sleep_for(std::chrono::milliseconds(500));
                if (g_watch_id.load() >= 0 && now_s() - g_watch_t0.load() > lim) {
                    fprintf(stderr, "EMU: kernel %s did not finish within %.0f s (deadlock or endless loop)\n",
                            g_watch_name, lim);
                    fflush(stderr);
                    std::abort();
                }
            }
        }).detach();
    });
}

void *dyn_shared() { return ((Block *)tctx.block)->shmem.data(); }

void syncthreads() { ((Block *)tctx.block)->bar.arrive_and_wait(); }

void syncwarp() { ((Block *)tctx.block)->waves[tctx.wave]->bar.arrive_and_wait(); }

void wave_exchange(uint64_t v, uint64_t *all)
{
    Block *b = (Block *)tctx.block;
    WaveSync &w = *b->waves[tctx.wave];
    w.buf[tctx.lane] = v;
    w.bar.arrive_and_wait();
    memcpy(all, w.buf, sizeof(w.buf));
    w.bar.arrive_and_wait();
}

int atomic_add_i(int *p, int v) { std::lock_guard<std::mutex> g(g_atomic); int o = *p; *p = o + v; return o; }
double atomic_add_d(double *p, double v) { std::lock_guard<std::mutex> g(g_atomic); double o = *p; *p = o + v; return o; }
int atomic_or_i(int *p, int v) { std::lock_guard<std::mutex> g(g_atomic); int o = *p; *p = o | v; return o; }

void run_grid(const char *name, dim3 grid, dim3 block, size_t shmem, void (*thunk)(void *), void *ctx)
{
    watchdog();
    const int nt = block.x * block.y * block.z;
    const long long id = g_launch_id++;
    g_watch_name = name;
    static const bool trace = getenv("EMU_TRACE") != nullptr;
    if (trace) fprintf(stderr, "EMU launch %s grid %u block %d lds %zu\n", name, grid.x, nt, shmem);
    g_watch_t0 = now_s();
    g_watch_id = id;
    for (unsigned bz = 0; bz < grid.z; bz++)
        for (unsigned by = 0; by < grid.y; by++)
            for (unsigned bx = 0; bx < grid.x; bx++) {
                Block B(nt);
                B.shmem.assign(shmem ? shmem : 8, (char)0xFF);  // poisoned like uninitialised LDS
                const int nw = (nt + 63) / 64;
                for (int wv = 0; wv < nw; wv++) B.waves.emplace_back(new WaveSync(std::min(64, nt - 64 * wv)));
                std::vector<std::thread> th;
                th.reserve(nt);
                for (int t = 0; t < nt; t++) {
                    th.emplace_back([&, t] {
                        tctx.tid = dim3(t % block.x, (t / block.x) % block.y, t / (block.x * block.y));
                        tctx.bid = dim3(bx, by, bz);
                        tctx.bdim = block;
                        tctx.gdim = grid;
                        tctx.lane = t % 64;
                        tctx.wave = t / 64;
                        tctx.block = &B;
                        thunk(ctx);
                        // a finished thread leaves every barrier it could have joined
                        B.waves[t / 64]->bar.arrive_and_drop();
                        B.bar.arrive_and_drop();
                    });
                }
                for (auto &x : th) x.join();
            }
    g_watch_id = -1;
}
}  // namespace emu

hipError_t hipMalloc(void **p, size_t n)
{
    *p = malloc(n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    memset(*p, 0xFF, n);  // NaN pattern for doubles: use-before-write shows up
    return hipSuccess;
}
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) { memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = (hipStream_t)malloc(1); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 2; return hipSuccess; }
hipError_t hipFuncSetAttribute(const void *, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emulated error"; }
struct emu_event { double t; };
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = new emu_event{0}; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = emu::now_s(); return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { *ms = (float)(1e3 * (b->t - a->t)); return hipSuccess; }


// ---- stream capture and graphs
struct emu_graph {
    std::vector<std::function<void()>> ops;
};
static thread_local emu_graph *g_capture = nullptr;
namespace emu {
bool capturing() { return g_capture != nullptr; }
void record(std::function<void()> f) { g_capture->ops.push_back(std::move(f)); }
}  // namespace emu
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode)
{
    if (g_capture) return hipErrorInvalidValue;
    g_capture = new emu_graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    if (!g_capture) return hipErrorInvalidValue;
    *g = g_capture;
    g_capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t)
{
    *e = new emu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    for (auto &f : e->ops) f();
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }

hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
