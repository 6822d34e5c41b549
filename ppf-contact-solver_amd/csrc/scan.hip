// scan.hip -- exclusive prefix sum of uint32 arrays (device-wide).
//
// Three-phase scan: per-block LDS scan of 1024 items, recursive scan of the
// block totals, offset add. Deterministic, no atomics. Used for CSR offsets
// (transpose tables, compaction of pair lists).
#include "scan.hpp"

namespace ppf {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_tile(
    const uint32_t *__restrict__ in, uint32_t *__restrict__ out, size_t n,
    uint32_t *__restrict__ tile_sums) {
    __shared__ uint32_t wsum[kScanThreads / kWave];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        size_t idx = base + (size_t)threadIdx.x * kScanItems + q;
        v[q] = idx < n ? in[idx] : 0u;
        local += v[q];
    }
    // Inclusive scan of `local` across the wave.
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t incl = local;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t t = __shfl_up(incl, off, kWave);
        if (lane >= off) {
            incl += t;
        }
    }
    if (lane == kWave - 1) {
        wsum[w] = incl;
    }
    __syncthreads();
    uint32_t wofs = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kScanThreads / kWave; ++k) {
        if (k < w) {
            wofs += wsum[k];
        }
        total += wsum[k];
    }
    uint32_t run = wofs + incl - local;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        size_t idx = base + (size_t)threadIdx.x * kScanItems + q;
        if (idx < n) {
            out[idx] = run;
        }
        run += v[q];
    }
    if (threadIdx.x == 0 && tile_sums) {
        tile_sums[blockIdx.x] = total;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_add(
    uint32_t *__restrict__ out, size_t n, const uint32_t *__restrict__ ofs) {
    const size_t base = (size_t)blockIdx.x * kScanTile;
    const uint32_t o = ofs[blockIdx.x];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        size_t idx = base + (size_t)threadIdx.x * kScanItems + q;
        if (idx < n) {
            out[idx] += o;
        }
    }
}

size_t tiles_of(size_t n) { return (n + kScanTile - 1) / kScanTile; }

} // namespace

size_t scan_scratch_words(size_t n) {
    size_t words = 0;
    for (size_t m = tiles_of(n); m > 1; m = tiles_of(m)) {
        words += 2 * m;
    }
    return words + 2;
}

hipError_t exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n,
                              uint32_t *scratch, hipStream_t s) {
    if (n == 0) {
        return hipSuccess;
    }
    const size_t nt = tiles_of(n);
    uint32_t *sums = nt > 1 ? scratch : nullptr;
    k_scan_tile<<<(unsigned)nt, kScanThreads, 0, s>>>(in, out, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nt == 1) {
        return e;
    }
    uint32_t *sums_scan = scratch + nt;
    e = exclusive_scan_u32(sums, sums_scan, nt, scratch + 2 * nt, s);
    if (e != hipSuccess) {
        return e;
    }
    k_scan_add<<<(unsigned)nt, kScanThreads, 0, s>>>(out, n, sums_scan);
    return hipGetLastError();
}

} // namespace ppf
