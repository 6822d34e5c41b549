// sort.hip -- stable LSD radix sort of (uint32 key, uint32 value) pairs.
//
// 8-bit digits, one histogram + one scatter kernel per digit. Stability comes
// from ranking inside a tile in input order: each 256-item round is ranked
// with wave64 ballots (8 ballots give every lane the mask of lanes holding the
// same digit), per-wave digit counts go through LDS, and a running per-digit
// count carries the order across rounds. Replaces the reference's 4-bit,
// 8-pass sort (kernels/radix_sort.cu:22-178): 30-bit Morton keys need 4 passes.
#include "scan.hpp"
#include "sort.hpp"

namespace ppf {

namespace {

constexpr int kThreads = 256;
constexpr int kRounds = 4;
constexpr int kTile = kThreads * kRounds;
constexpr int kWaves = kThreads / kWave;

__global__ __launch_bounds__(kThreads) void k_digit_hist(
    const uint32_t *__restrict__ keys, size_t n, uint32_t shift,
    uint32_t ntiles, uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kTile;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const size_t idx = base + (size_t)r * kThreads + threadIdx.x;
        if (idx < n) {
            atomicAdd(&h[(keys[idx] >> shift) & 0xffu], 1u);
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_digit_scatter(
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
    size_t n, uint32_t shift, uint32_t ntiles,
    const uint32_t *__restrict__ hist_scan, uint32_t *__restrict__ okeys,
    uint32_t *__restrict__ ovals) {
    __shared__ uint32_t wcount[kWaves][256];
    __shared__ uint32_t running[256];
    running[threadIdx.x] = hist_scan[(size_t)threadIdx.x * ntiles + blockIdx.x];
    const size_t base = (size_t)blockIdx.x * kTile;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    for (int r = 0; r < kRounds; ++r) {
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            wcount[k][threadIdx.x] = 0;
        }
        __syncthreads();
        const size_t idx = base + (size_t)r * kThreads + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys[idx] : 0u;
        const uint32_t d = (key >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank_w = __popcll(peers & below);
        if (valid && rank_w == 0) {
            wcount[w][d] = (uint32_t)__popcll(peers);
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank_w;
            for (int k = 0; k < w; ++k) {
                pos += wcount[k][d];
            }
            okeys[pos] = key;
            ovals[pos] = vals[idx];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            add += wcount[k][threadIdx.x];
        }
        running[threadIdx.x] += add;
        __syncthreads();
    }
}

size_t tiles(size_t n) { return (n + kTile - 1) / kTile; }

} // namespace

size_t radix_sort_scratch_words(size_t n) {
    const size_t h = 256 * tiles(n);
    return 2 * n + 2 * h + scan_scratch_words(h);
}

hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, size_t n,
                            uint32_t key_bits, uint32_t *scratch,
                            hipStream_t s) {
    if (n <= 1) {
        return hipSuccess;
    }
    const uint32_t nt = (uint32_t)tiles(n);
    const size_t h = 256 * (size_t)nt;
    uint32_t *k2 = scratch, *v2 = scratch + n;
    uint32_t *hist = scratch + 2 * n, *hscan = hist + h, *sscr = hscan + h;
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    const uint32_t passes = (key_bits + 7) / 8;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = 8 * p;
        k_digit_hist<<<nt, kThreads, 0, s>>>(ka, n, shift, nt, hist);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            return e;
        }
        e = exclusive_scan_u32(hist, hscan, h, sscr, s);
        if (e != hipSuccess) {
            return e;
        }
        k_digit_scatter<<<nt, kThreads, 0, s>>>(ka, va, n, shift, nt, hscan, kb,
                                                vb);
        e = hipGetLastError();
        if (e != hipSuccess) {
            return e;
        }
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        hipError_t e = hipMemcpyAsync(keys, ka, n * sizeof(uint32_t),
                                      hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) {
            return e;
        }
        return hipMemcpyAsync(vals, va, n * sizeof(uint32_t),
                              hipMemcpyDeviceToDevice, s);
    }
    return hipSuccess;
}

} // namespace ppf
