// common.hpp -- shared device/host helpers of the MI355X backend.
//
// gfx950 only: wave64, column-major 3x3 blocks (the layout of the reference's
// Mat3x3f, crates/ppf-cts-solver/src/cpp/linalg/smat.hpp), fp32 arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#define PPF_HD __host__ __device__ inline

// Global (HBM) address space. Pointers read out of a device-side argument
// struct are otherwise generic, which turns every access into a flat
// instruction that also waits on the LDS/scalar counter.
#define PPF_GLOBAL __attribute__((address_space(1)))

namespace ppf {

template <class T> using gptr = PPF_GLOBAL T *;
template <class T> __device__ inline gptr<T> as_global(T *p) {
    return (gptr<T>)p;
}

constexpr int kWave = 64;
constexpr int kXcd = 8; // MI355X: 8 XCDs, workgroups dispatched round-robin

struct f3 {
    float x, y, z;
};

PPF_HD f3 mk3(float a, float b, float c) { return f3{a, b, c}; }
PPF_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
PPF_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
PPF_HD f3 operator*(float s, f3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
PPF_HD float dot3(f3 a, f3 b) {
    float acc = 0.0f;
    acc = acc + a.x * b.x;
    acc = acc + a.y * b.y;
    acc = acc + a.z * b.z;
    return acc;
}
PPF_HD f3 cross3(f3 a, f3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x);
}
__device__ inline f3 load3(gptr<const float> p) { return mk3(p[0], p[1], p[2]); }
__device__ inline void store3(gptr<float> p, f3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
__device__ inline f3 mv(gptr<const float> a, f3 x) {
    return mk3(a[0] * x.x + a[3] * x.y + a[6] * x.z,
               a[1] * x.x + a[4] * x.y + a[7] * x.z,
               a[2] * x.x + a[5] * x.y + a[8] * x.z);
}
__device__ inline f3 mvt(gptr<const float> a, f3 x) {
    return mk3(a[0] * x.x + a[1] * x.y + a[2] * x.z,
               a[3] * x.x + a[4] * x.y + a[5] * x.z,
               a[6] * x.x + a[7] * x.y + a[8] * x.z);
}
PPF_HD f3 load3(const float *p) { return mk3(p[0], p[1], p[2]); }
PPF_HD void store3(float *p, f3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
// y = A x, column-major A, term order of the reference's UnrolledMat3x3f.
PPF_HD f3 mv(const float *a, f3 x) {
    return mk3(a[0] * x.x + a[3] * x.y + a[6] * x.z,
               a[1] * x.x + a[4] * x.y + a[7] * x.z,
               a[2] * x.x + a[5] * x.y + a[8] * x.z);
}
// y = A^T x.
PPF_HD f3 mvt(const float *a, f3 x) {
    return mk3(a[0] * x.x + a[1] * x.y + a[2] * x.z,
               a[3] * x.x + a[4] * x.y + a[5] * x.z,
               a[6] * x.x + a[7] * x.y + a[8] * x.z);
}

// Sum over a wave64 (butterfly; every lane ends with the total). The order of
// operations is fixed, so the result is deterministic.
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_xor(v, off, kWave);
    }
    return v;
}

// Deterministic block sum for blockDim.x == NT (multiple of 64). Every thread
// returns the total. `lds` needs NT/64 floats.
template <int NT> __device__ inline float block_sum(float v, float *lds) {
    constexpr int NW = NT / kWave;
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        lds[w] = v;
    }
    __syncthreads();
    float t = 0.0f;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        t += lds[k];
    }
    __syncthreads();
    return t;
}

// Workgroup id remapped so that each XCD owns one contiguous range of logical
// ids (the hardware deals workgroups to the 8 XCDs round-robin). Consecutive
// rows then share one XCD's L2. Requires gridDim.x % 8 == 0.
__device__ inline unsigned xcd_logical_block() {
    const unsigned g = gridDim.x, b = blockIdx.x;
    return (b % kXcd) * (g / kXcd) + b / kXcd;
}

// Grid of at least `blocks` workgroups that xcd_logical_block can remap (the
// kernel treats logical ids past its work as empty).
inline uint32_t xcd_grid(uint32_t blocks) {
    return (blocks + kXcd - 1) / kXcd * kXcd;
}

} // namespace ppf

// Error plumbing for the C ABI: every entry point returns 0 on success and a
// nonzero code otherwise; ppf_last_error() holds the message.
namespace ppf {
void set_error(const char *fmt, ...);

// Every device and pinned-host allocation of the library goes through these,
// which count them (the reference's g_device_alloc_count /
// g_device_free_count, main/main.cu:280-286, 1635-1646): after warm-up a
// time step allocates and frees nothing, buffers grow to a high-water mark.
void count_alloc();
void count_free();
void alloc_totals(uint64_t *allocs, uint64_t *frees);

// Walk statistics of the BVH walks (a measurement mode, off by default):
// device counters the walk kernels add to (null when off), and the time of
// each walk kernel launch. A WalkClock brackets one launch of a kind; while
// statistics are on it records HIP events around it and waits for the end.
unsigned long long *walk_stats_device();
class WalkClock {
  public:
    WalkClock(int kind, hipStream_t s);
    void stop();
  private:
    int kind_;
    hipStream_t s_;
    bool on_;
};
// Capacity of a buffer that grows with the contact set: the next power of
// two at or above twice the need (at least `floor`), as the reference sizes
// its pair caches (contact/contact.cu:1389-1396). Growth events are then
// rare, and a step whose sizes stay below the high-water mark allocates
// nothing.
inline uint64_t grow_cap(uint64_t need, uint64_t floor = 1024) {
    uint64_t c = floor;
    while (c < 2 * need) {
        c *= 2;
    }
    return c;
}
template <class T> hipError_t dev_alloc(T **p, size_t bytes) {
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    if (e == hipSuccess) {
        count_alloc();
    }
    return e;
}
inline hipError_t dev_free(void *p) {
    if (p) {
        count_free();
    }
    return hipFree(p);
}
template <class T> hipError_t host_alloc(T **p, size_t bytes) {
    const hipError_t e =
        hipHostMalloc(reinterpret_cast<void **>(p), bytes, hipHostMallocDefault);
    if (e == hipSuccess) {
        count_alloc();
    }
    return e;
}
inline hipError_t host_free(void *p) {
    if (p) {
        count_free();
    }
    return hipHostFree(p);
}
} // namespace ppf

#define PPF_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::ppf::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,        \
                             hipGetErrorString(e_));                           \
            return PPF_ERR_HIP;                                                \
        }                                                                      \
    } while (0)
