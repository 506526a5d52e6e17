// Shared definitions of the medimage_amd device code and host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mi {

// Error carried from anywhere in the runtime to the C ABI boundary, where
// it becomes a negative status and the thread-local last-error message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what, const char *file, int line);

#define MI_HIP(call) ::mi::hip_check((call), #call, __FILE__, __LINE__)
#define MI_REQUIRE(cond, code, msg) \
    do { if (!(cond)) ::mi::fail((code), (msg)); } while (0)

// Dispatch-timestamp events for the kernel launches of a timed scope: the
// first launch through mi::launch records `a` at its start and `b` at its end
// (the dispatch's own timestamps, no separate event packets); later launches
// of the scope are counted and the scope records `b` again after them.
struct DispatchEvents {
    hipEvent_t a = nullptr, b = nullptr;
    int launches = 0;
};
extern thread_local DispatchEvents *tl_dispatch;

template <typename Kernel, typename... Args>
inline void launch(Kernel kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t st, Args... args) {
    DispatchEvents *d = tl_dispatch;
    if (d && d->launches++ == 0)
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, d->a, d->b, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, shmem, st, args...);
}

constexpr int kWave = 64;        // CDNA wavefront width
constexpr int kBlock = 256;      // 4 waves per workgroup

// ---------------------------------------------------------------------------
// Compensated accumulation.  FP32 runs use Kahan's algorithm per lane; FP64
// runs add plainly.  value() hands the sum on as a double, so the cross-lane
// combination is done in double in both cases.
template <typename T> struct Acc;

template <> struct Acc<float> {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float y = x - c;
        float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    // add(x) when `take`, else nothing (a select, no branch).
    __device__ __forceinline__ void add_if(float x, bool take) {
        const float y = x - c;
        const float t = s + y;
        const float cn = (t - s) - y;
        s = take ? t : s;
        c = take ? cn : c;
    }
    // add(x * l) with the product folded into the compensation step
    // (y = x l - c rounded once).
    __device__ __forceinline__ void add_prod(float x, float l) {
        const float y = __builtin_fmaf(x, l, -c);
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ void add_prod_if(float x, float l, bool take) {
        const float y = __builtin_fmaf(x, l, -c);
        const float t = s + y;
        const float cn = (t - s) - y;
        s = take ? t : s;
        c = take ? cn : c;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <> struct Acc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ void add_if(double x, bool take) { s = take ? s + x : s; }
    __device__ __forceinline__ void add_prod(double x, double l) { s += x * l; }
    __device__ __forceinline__ void add_prod_if(double x, double l, bool take) { s = take ? s + x * l : s; }
    __device__ __forceinline__ double value() const { return s; }
};

// Butterfly sum over `width` consecutive lanes (power of two <= 64).  The
// pairing is fixed, so the result does not depend on timing.
template <int width>
__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
    for (int o = width / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, width);
    return x;
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, kWave));
    return x;
}

// Block-wide sum of one double per thread (blockDim.x == kBlock), fixed order.
__device__ __forceinline__ double block_sum(double x, double *lds /* kBlock/kWave */) {
    x = group_sum<kWave>(x);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) lds[w] = x;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / kWave; ++i) r += lds[i];
    __syncthreads();
    return r;  // valid in thread 0
}

__device__ __forceinline__ double block_max(double x, double *lds) {
    x = wave_max(x);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) lds[w] = x;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kBlock / kWave; ++i) r = fmax(r, lds[i]);
    __syncthreads();
    return r;
}

// Workgroups are dispatched to the 8 XCDs round-robin.  Renumbering them so
// that each XCD owns a contiguous range of the work keeps neighbouring rays
// (or pixels), which gather the same image (or sinogram) lines, in one L2.
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned g = gridDim.x, b = blockIdx.x;
    if (g < 8 || (g & 7)) return b;
    return (b & 7) * (g >> 3) + (b >> 3);
}

// Four consecutive entries of a padded matrix stream (16-byte aligned start).
template <typename T> struct Vec4;
template <> struct Vec4<float> {
    float v[4];
    __device__ __forceinline__ void load(const float *p) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    }
};
template <> struct Vec4<double> {
    double v[4];
    __device__ __forceinline__ void load(const double *p) {
        const double2 a = reinterpret_cast<const double2 *>(p)[0];
        const double2 b = reinterpret_cast<const double2 *>(p)[1];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
};
__device__ __forceinline__ void load_idx4(const uint32_t *p, uint32_t (&o)[4]) {
    const uint4 a = *reinterpret_cast<const uint4 *>(p);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}

// Device-resident scalars of one job's solve.  Kernels read step sizes from
// here and write reduced sums here, so a pass needs no host round trip.
struct Scalars {
    double df_pass;     // sum over subsets of ||g_s - A_s u||^2, this pass
    double df_prev;     // previous pass
    double rel;         // last relative change
    double dd;          // <d, d>
    double ad2;         // ||A_s d||^2
    double nv;          // nV(d)
    double c;           // line-search step of the image update
    double e_nd;        // edge: ||d||^2
    double e_nd2;       // edge: sum n d^2
    double e_nnd;       // edge: sum |central grad d|^2
    double e_c;         // edge step
    double e_maxcd;     // edge: max |c d|
    double gen[8];      // general-purpose slots (GoldenC CG scalars)
    int32_t edge_done;  // edge inner loop stopped early
    int32_t converged;  // rel_tol met
    int32_t iters;      // passes completed
    int32_t done;       // converged or out of passes: kernels leave the job alone
};

}  // namespace mi
