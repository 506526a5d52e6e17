// Shared definitions of the medimage_amd device code and host runtime.
#pragma once

#include <hip/hip_runtime.h>

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
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <> struct Acc<double> {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
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
    int32_t pad;
};

}  // namespace mi
