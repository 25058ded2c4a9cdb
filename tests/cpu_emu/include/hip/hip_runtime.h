// CPU emulation of the subset of HIP the engine uses, for checking the exact
// kernel source on the host (bounds checks with AddressSanitizer, poisoned
// allocations, barrier/shuffle discipline).  Each block runs its threads as
// OS threads with real barriers; blocks run one after another.  Test
// infrastructure only: never used for the product build.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <cstddef>
#include <functional>
#include <tuple>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu::dyn_shared();

using std::isfinite;

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef enum hipError_t { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 } hipError_t;
typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 } hipDeviceAttribute_t;
typedef enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 0 } hipFuncAttribute;
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;

namespace emu {
struct ThreadCtx {
    dim3 tid, bid, bdim, gdim;
    int lane = 0, wave = 0;
    void *block = nullptr;
};
extern thread_local ThreadCtx tctx;
void *dyn_shared();
void syncthreads();
void wave_exchange(uint64_t v, uint64_t *all);  // every lane of the wave deposits, all read
void syncwarp();                                 // barrier of the lanes of one wave
int atomic_add_i(int *p, int v);
double atomic_add_d(double *p, double v);
int atomic_or_i(int *p, int v);
void run_grid(const char *name, dim3 grid, dim3 block, size_t shmem, void (*thunk)(void *), void *ctx);

template <typename T>
inline T shfl_generic(T v, int src, int width)
{
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    uint64_t u = 0;
    memcpy(&u, &v, sizeof(T));
    uint64_t all[64];
    wave_exchange(u, all);
    const int base = (tctx.lane / width) * width;
    T r;
    memcpy(&r, &all[base + (src % width + width) % width], sizeof(T));
    return r;
}

template <typename F, typename Tup>
struct Thunk {
    F f;
    Tup args;
    static void call(void *p)
    {
        auto *t = (Thunk *)p;
        std::apply(t->f, t->args);
    }
};
}  // namespace emu

#define threadIdx (emu::tctx.tid)
#define blockIdx (emu::tctx.bid)
#define blockDim (emu::tctx.bdim)
#define gridDim (emu::tctx.gdim)

inline void __syncthreads() { emu::syncthreads(); }
inline void __syncwarp() { emu::syncwarp(); }
template <typename T>
inline T __shfl(T v, int src, int width = 64) { return emu::shfl_generic(v, src, width); }
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) { return emu::shfl_generic(v, (emu::tctx.lane % width) ^ mask, width); }
// HIP's global integer min/max
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

// Lane 0's value; the engine uses it only on values it holds wave-uniform, so
// the emulation aborts if any lane of the wave disagrees.
inline int __builtin_amdgcn_readfirstlane(int v)
{
    uint64_t all[64];
    emu::wave_exchange((uint64_t)(uint32_t)v, all);
    const int base = (emu::tctx.lane / 64) * 64;
    for (int l = 0; l < 64; l++)
        if (all[base + l] != all[base]) {
            fprintf(stderr, "emu: readfirstlane of a non-uniform value\n");
            abort();
        }
    return (int)(uint32_t)all[base];
}
// FP64 MFMA 16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// and holds D[4 r + l / 16][l % 16] in accumulator element r = 0..3 (the
// layout measured on gfx950 with a one-hot probe of A and B)
#define CHAMP_DBL4_DEFINED
// HIP's float2 vector type and its constructor
struct float2 {
    float x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }

struct dbl4 {
    double v[4];
    double &operator[](int i) { return v[i]; }
    const double &operator[](int i) const { return v[i]; }
};
inline dbl4 __builtin_amdgcn_mfma_f64_16x16x4f64(double a, double b, dbl4 c, int, int, int)
{
    uint64_t ua, ub, all_a[64], all_b[64];
    memcpy(&ua, &a, 8);
    memcpy(&ub, &b, 8);
    emu::wave_exchange(ua, all_a);
    emu::wave_exchange(ub, all_b);
    const int l = emu::tctx.lane % 64, j = l % 16;
    for (int r = 0; r < 4; r++) {
        const int i = 4 * r + l / 16;
        double s = c[r];
        for (int k = 0; k < 4; k++) {
            double x, y;
            memcpy(&x, &all_a[k * 16 + i], 8);
            memcpy(&y, &all_b[k * 16 + j], 8);
            s += x * y;
        }
        c[r] = s;
    }
    return c;
}
// FP32 MFMA 16x16x4: A and B as above; accumulator element r of lane l holds
// D[4 (l / 16) + r][l % 16] (measured on gfx950 likewise)
struct flt4 {
    float v[4];
    float &operator[](int i) { return v[i]; }
    const float &operator[](int i) const { return v[i]; }
};
inline flt4 __builtin_amdgcn_mfma_f32_16x16x4f32(float a, float b, flt4 c, int, int, int)
{
    uint64_t ua = 0, ub = 0, all_a[64], all_b[64];
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    emu::wave_exchange(ua, all_a);
    emu::wave_exchange(ub, all_b);
    const int l = emu::tctx.lane % 64, j = l % 16;
    for (int r = 0; r < 4; r++) {
        const int i = 4 * (l / 16) + r;
        float s = c[r];
        for (int k = 0; k < 4; k++) {
            float x, y;
            memcpy(&x, &all_a[k * 16 + i], 4);
            memcpy(&y, &all_b[k * 16 + j], 4);
            s += x * y;
        }
        c[r] = s;
    }
    return c;
}
inline int atomicAdd(int *p, int v) { return emu::atomic_add_i(p, v); }
inline double atomicAdd(double *p, double v) { return emu::atomic_add_d(p, v); }
inline int atomicOr(int *p, int v) { return emu::atomic_or_i(p, v); }

hipError_t hipMalloc(void **p, size_t n);
template <typename T>
inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipGetDevice(int *d);
// a __device__ variable is host memory here: its "device address" is its address
#define HIP_SYMBOL(x) ((const void *)&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
hipError_t hipFuncSetAttribute(const void *f, hipFuncAttribute a, int v);
template <typename F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t) { *n = 4; return hipSuccess; }
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipEventCreate(hipEvent_t *e);
#define hipEventDisableTiming 0x2
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
// launches run in issue order on the host, so a cross-stream wait has nothing to wait for
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventSynchronize(hipEvent_t e);

// Stream capture: launches between hipStreamBeginCapture and
// hipStreamEndCapture are recorded (closures over copies of their arguments)
// and run in order by hipGraphLaunch.
typedef enum hipStreamCaptureMode {
    hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2
} hipStreamCaptureMode;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_graph *hipGraphExec_t;
typedef struct emu_graph_node *hipGraphNode_t;
namespace emu {
bool capturing();
void record(std::function<void()> f);
}  // namespace emu
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                                   \
    do {                                                                                              \
        auto emu_args_ = std::make_tuple(__VA_ARGS__);                                                \
        typedef emu::Thunk<decltype(&kernel), decltype(emu_args_)> EmuTT_;                             \
        EmuTT_ emu_t_{&kernel, emu_args_};                                                            \
        const dim3 emu_g_ = dim3(grid), emu_b_ = dim3(block);                                         \
        const size_t emu_sh_ = (size_t)(shmem);                                                       \
        if (emu::capturing())                                                                         \
            emu::record([emu_t_, emu_g_, emu_b_, emu_sh_]() mutable {                                  \
                emu::run_grid(#kernel, emu_g_, emu_b_, emu_sh_, &EmuTT_::call, &emu_t_);              \
            });                                                                                       \
        else                                                                                          \
            emu::run_grid(#kernel, emu_g_, emu_b_, emu_sh_, &EmuTT_::call, &emu_t_);                  \
    } while (0)
