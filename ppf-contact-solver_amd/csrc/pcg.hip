// pcg.hip -- block-Jacobi preconditioned CG on MI355X.
//
// Replaces solver::solve / solver::cg_device
// (crates/ppf-cts-solver/src/cpp/solver/solver.cu:921-1383, 2180-2321) for the
// block-Jacobi preconditioner. Same recurrence, same curvature guard
// (solver.cu:57-114), same L1 relative-residual stopping rule measured against
// the seeded residual |b - A x0|_1, and the same residual-check schedule
// (every 4 iterations, every iteration once within 8*tol), emulated on the
// device so that no host round trip sits inside the loop.
//
// MI355X layout and kernels:
//   * the operator is the reference's split form, an upper-triangle block CSR
//     applied both ways plus a per-row diagonal. The diagonal blocks are merged
//     once per solve into D_i = B_ii + C_i, and the transpose of the strict
//     upper triangle is walked through a per-column (row, slot) table. Rows
//     are dealt to XCDs in contiguous ranges, and a plan may solve in a
//     locality row order, so that a mirrored block's second read mostly hits
//     the L2 of the XCD that read it as a stored block.
//   * one iteration = 3 kernels: SpMV + p.Ap + sum|terms|; the fused x/r/z
//     update with the r.z and |r|_1 partials; p = z + beta p (5 with the
//     factor preconditioner: G r and G^T (G r) between the last two). Each
//     kernel writes one partial per workgroup tile; every workgroup of the
//     next kernel reduces its scene's partials itself in a fixed order, so
//     the scalars are bit-identical everywhere with no atomics and no grid
//     barrier, and every run is deterministic.
//   * a plan may hold many independent systems (scenes) stacked
//     block-diagonally: tiles never straddle scenes and each scene keeps its
//     own scalars and stopping rule.
//   * iterations are captured 8 at a time into a hipGraph; once the device
//     bookkeeping marks the solve done, the remaining replays exit at entry.
#include "common.hpp"
#include "eig3.hpp"
#include "pcg.hpp"
#include "scan.hpp"
#include "sort.hpp"
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace ppf {

namespace {

constexpr int kNT = 256;           // threads per workgroup
constexpr int kGraphIters = 8;     // iterations per captured graph (even)
constexpr unsigned kNone = 0xffffffffu;
constexpr unsigned kMirror = 0x80000000u; // entry slot flag: apply transposed
constexpr float kNoiseSlack = 256.0f;       // solver.cu:69
constexpr float kFloatEps = 1.19209290e-7f; // solver.cu:70
// In-tile mirrors the symmetric SpMV form parks in LDS per workgroup (3
// floats each: 24 KB); a tile's mirrors beyond it are read as in the split
// form. Its tables pack an LDS place (kPlaceBits, kPlaceNone = none) with a
// column into one word per stored slot, and a place (flagged kInTile) or the
// mirrored slot into one word per mirror entry, so columns stay below
// kSymMaxRows.
constexpr uint32_t kPlaceBits = 11;
constexpr uint32_t kPlaceNone = (1u << kPlaceBits) - 1u;
constexpr uint32_t kSymCap = kPlaceNone;
constexpr uint32_t kSymMaxRows = 1u << (32 - kPlaceBits);
constexpr uint32_t kInTile = 0x80000000u;

// Device-resident scalars of one solve. Written by exactly one thread of one
// kernel and read only by later kernels (stream order), except `done`, whose
// late observation inside the kernel that sets it is harmless.
struct Scalars {
    float rz[2];     // r.z of the latest two iterations, by parity
    float err0;
    float tol;
    uint32_t max_iter;
    uint32_t iter;   // completed iterations
    uint32_t next_check;
    uint32_t stride;
    int32_t done;
    int32_t status;
    int32_t breakdown;
    int32_t cause;
    int32_t break_iter;
    uint32_t rep_iter;
    float rep_resid;
    uint32_t bad;    // non-invertible diagonal blocks
};

// Work of one workgroup: rows [lo, hi) of one scene (a plain solve is one
// scene); `first` marks the scene's first tile, whose thread 0 keeps the
// scene's bookkeeping.
struct Tile {
    uint32_t lo, hi, scene, first;
};

struct Args {
    uint32_t n;
    uint32_t G1, G2;     // tiles (workgroups) of the SpMV / of the vector kernels
    uint32_t nscene;
    const Tile *tile1, *tile2; // [G1], [G2]; scene ranges contiguous
    const uint32_t *sc1, *sc2; // per scene its first tile1 / tile2 ([nscene + 1])
    const uint32_t *rp, *col, *tp, *tcol, *tslot, *dslot;
    const uint2 *ent;  // fixed-width entry table (ELL form), or null
    uint32_t ell_m;    // entries per lane of the ELL kernel (row width 4*m)
    // symmetric form: per stored slot (column << kPlaceBits) | its mirror's
    // place among the tile's in-tile mirrors (LDS) or kPlaceNone; per mirror
    // entry kInTile | that place, or the stored slot it mirrors
    const uint32_t *mq, *tq;
    const float *val, *cdiag, *b;
    float *x, *P, *r, *z, *p, *Ap;
    float *part; // [G1 pAp][G1 |terms|][G2 rz][G2 err], one per tile
    float *err_hist; // (one-scene solves)
    Scalars *s;      // [nscene]
    // Factor preconditioner M^-1 = G^T G (null: block Jacobi): Gd the
    // diagonal blocks of G (n), Gl[s] = G_ji for every stored slot s = (i, j),
    // i < j (the lower block mirrored by the slot), both bf16 column-major in
    // 10-entry (20 B) records; gy holds G r.
    const uint16_t *Gd, *Gl, *Gr; // Gr: the rows of G in order (row j at j + tp[j])
    float *gy;
};

__device__ inline Tile load_tile(const Tile *t, uint32_t L) {
    const uint4 v = reinterpret_cast<const uint4 *>(t)[L];
    return Tile{v.x, v.y, v.z, v.w};
}

// ---- pattern preparation (once per sparsity pattern) ----------------------

__global__ void k_count_transpose(uint32_t n, const uint32_t *__restrict__ rp,
                                  const uint32_t *__restrict__ col,
                                  uint32_t *__restrict__ tcount,
                                  uint32_t *__restrict__ dslot,
                                  uint32_t *__restrict__ bad_pattern) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t ds = kNone;
    uint32_t prev = 0;
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const uint32_t j = col[k];
        if (j >= n || j < i || (k > rp[i] && j <= prev)) {
            atomicAdd(bad_pattern, 1u); // not upper, unsorted or out of range
            continue;
        }
        prev = j;
        if (j == i) {
            ds = k;
        } else {
            atomicAdd(tcount + j, 1u);
        }
    }
    dslot[i] = ds;
}

// Radix-sort input for the transpose table: the key of a stored slot is its
// column when the entry is an accepted off-diagonal one (the acceptance rule
// of k_count_transpose), n otherwise; the value is the slot and row[] its row.
__global__ void k_transpose_keys(uint32_t n, const uint32_t *__restrict__ rp,
                                 const uint32_t *__restrict__ col,
                                 uint32_t *__restrict__ key, uint32_t *__restrict__ val,
                                 uint32_t *__restrict__ row) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t prev = 0;
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const uint32_t j = col[k];
        const bool ok = !(j >= n || j < i || (k > rp[i] && j <= prev));
        if (ok) {
            prev = j;
        }
        key[k] = ok && j > i ? j : n;
        val[k] = k;
        row[k] = i;
    }
}

// After a stable sort by column of the slots (which are in row order), the
// first tp[n] entries are every column's mirror list ordered by source row,
// as the reference's transpose table (builder.rs:1336-1345).
__global__ void k_transpose_place(uint32_t nnz, const uint32_t *__restrict__ tp_end,
                                  const uint32_t *__restrict__ val,
                                  const uint32_t *__restrict__ row,
                                  uint32_t *__restrict__ tcol, uint32_t *__restrict__ tslot) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nnz || u >= *tp_end) {
        return;
    }
    const uint32_t k = val[u];
    tcol[u] = row[k];
    tslot[u] = k;
}

// row_offset must start at 0, be non-decreasing and end at nnz.
__global__ void k_check_offsets(uint32_t n, uint32_t nnz,
                                const uint32_t *__restrict__ rp,
                                uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) {
        return;
    }
    const uint32_t v = rp[i];
    const bool ok = (i == 0 ? v == 0 : v >= rp[i - 1]) && (i < n || v == nnz) &&
                    v <= nnz;
    if (!ok) {
        atomicAdd(bad, 1u);
    }
}

// ---- per-solve setup -------------------------------------------------------

// D_i = B_ii + C_i and P_i = D_i^-1 (solver.cu:2191-2195).
__global__ __launch_bounds__(kNT) void k_diag_invert(const Args *__restrict__ A) {
    const Tile T = load_tile(A->tile2, blockIdx.x);
    gptr<const float> val = as_global(A->val);
    gptr<const float> cd = as_global(A->cdiag);
    gptr<float> P = as_global(A->P);
    for (uint32_t i = T.lo + threadIdx.x; i < T.hi; i += kNT) {
        const uint32_t ds = as_global(A->dslot)[i];
        float d[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const float bii = ds != kNone ? val[9 * (size_t)ds + e] : 0.0f;
            d[e] = bii + cd[9 * (size_t)i + e];
        }
        float inv[9];
        const bool ok = bj_invert(d, inv);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            P[9 * (size_t)i + e] = inv[e];
        }
        if (!ok) {
            atomicAdd(&A->s[T.scene].bad, 1u);
        }
    }
}

// Longest row of the split operator: stored blocks plus mirrored blocks.
__global__ void k_row_len_max(uint32_t n, const uint32_t *__restrict__ rp,
                              const uint32_t *__restrict__ tp,
                              uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    atomicMax(out, (rp[i + 1] - rp[i]) + (tp[i + 1] - tp[i]));
}

// Per-row entry table of width W: the row's stored blocks (column, slot),
// then its mirrored blocks (source row, slot | kMirror), padded with kNone.
__global__ void k_build_ell(uint32_t n, uint32_t W, const uint32_t *__restrict__ rp,
                            const uint32_t *__restrict__ col,
                            const uint32_t *__restrict__ tp,
                            const uint32_t *__restrict__ tcol,
                            const uint32_t *__restrict__ tslot,
                            uint2 *__restrict__ ent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint2 *row = ent + (size_t)i * W;
    uint32_t e = 0;
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        row[e++] = make_uint2(col[k], k);
    }
    for (uint32_t k = tp[i]; k < tp[i + 1]; ++k) {
        row[e++] = make_uint2(tcol[k], tslot[k] | kMirror);
    }
    for (; e < W; ++e) {
        row[e] = make_uint2(0u, kNone);
    }
}

// ---- row order (locality) ----------------------------------------------------
//
// A plan may solve in another row order (new row k = original row order[k]),
// e.g. the Morton order of the vertices, so that the rows a mirrored block
// comes from were read shortly before and its gather hits L2. The upper
// pattern is rebuilt in the new order (a block whose rows swap order moves to
// the other row, transposed), and values, right-hand side and iterate are
// permuted on the way in and out of each solve.

// inv[order[k]] = k, with inv preset to kNone: an entry >= n is flagged
// before any write, and a slot claimed twice (a duplicate, which leaves
// another slot unset) is flagged by the exchange.
__global__ void k_inverse_order(uint32_t n, const uint32_t *__restrict__ order,
                                uint32_t *__restrict__ inv, uint32_t *__restrict__ bad) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const uint32_t v = order[k];
    if (v >= n || atomicExch(inv + v, k) != kNone) {
        atomicAdd(bad, 1u);
    }
}

// Entries per new row; flags entries that are not strictly ascending upper
// columns < n in the original pattern (they are skipped).
__global__ void k_order_count(uint32_t n, const uint32_t *__restrict__ rp,
                              const uint32_t *__restrict__ col,
                              const uint32_t *__restrict__ inv, uint32_t *__restrict__ cnt,
                              uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t a = inv[i];
    uint32_t prev = 0;
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const uint32_t j = col[k];
        if (j >= n || j < i || (k > rp[i] && j <= prev)) {
            atomicAdd(bad, 1u);
            continue;
        }
        prev = j;
        atomicAdd(cnt + min(a, inv[j]), 1u);
    }
}

__global__ void k_order_fill(uint32_t n, const uint32_t *__restrict__ rp,
                             const uint32_t *__restrict__ col,
                             const uint32_t *__restrict__ inv,
                             const uint32_t *__restrict__ prp, uint32_t *__restrict__ cursor,
                             uint32_t *__restrict__ pcol, uint32_t *__restrict__ smap) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    const uint32_t a = inv[i];
    uint32_t prev = 0;
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const uint32_t j = col[k];
        if (j >= n || j < i || (k > rp[i] && j <= prev)) {
            continue;
        }
        prev = j;
        const uint32_t b = inv[j];
        const uint32_t r = min(a, b);
        const uint32_t e = prp[r] + atomicAdd(cursor + r, 1u);
        pcol[e] = max(a, b);
        smap[e] = k | (b < a ? kMirror : 0u);
    }
}

// Columns of each new row ascending (the fill order is arbitrary; the sort
// makes the pattern deterministic). Rows are short: insertion sort.
__global__ void k_order_sort(uint32_t n, const uint32_t *__restrict__ prp,
                             uint32_t *__restrict__ pcol, uint32_t *__restrict__ smap) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) {
        return;
    }
    const uint32_t lo = prp[r], hi = prp[r + 1];
    for (uint32_t a = lo + 1; a < hi; ++a) {
        const uint32_t c = pcol[a], m = smap[a];
        uint32_t b = a;
        while (b > lo && pcol[b - 1] > c) {
            pcol[b] = pcol[b - 1];
            smap[b] = smap[b - 1];
            --b;
        }
        pcol[b] = c;
        smap[b] = m;
    }
}

// Blocks of the reordered pattern: entry e takes original slot smap[e],
// transposed when the slot's rows swapped order.
__global__ void k_order_values(uint32_t nnz, const uint32_t *__restrict__ smap,
                               const float *__restrict__ val, float *__restrict__ pval) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) {
        return;
    }
    const uint32_t m = smap[e];
    const float *a = val + 9 * (size_t)(m & ~kMirror);
    float b[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        b[c] = a[c];
    }
    float *o = pval + 9 * (size_t)e;
    if (m & kMirror) {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            o[c] = b[3 * (c % 3) + c / 3];
        }
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            o[c] = b[c];
        }
    }
}

// dst[k] = src[order[k]] (w floats per row), or the inverse (scatter).
template <bool SCATTER>
__global__ void k_order_rows(uint32_t n, uint32_t w, const uint32_t *__restrict__ order,
                             const float *__restrict__ src, float *__restrict__ dst) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * w) {
        return;
    }
    const uint32_t k = t / w, c = t % w;
    const size_t o = (size_t)order[k] * w + c;
    if (SCATTER) {
        dst[o] = src[t];
    } else {
        dst[t] = src[o];
    }
}

// Split form of k_spmv_ell: the same entry list per row (the row's stored
// slots, then its mirrored slots) and the same lane assignment (lane q sums
// entries q, q+4, q+8, ... in that order), so the sums are bit-identical to
// the ELL form, but the entries come from the CSR pattern and the transpose
// table themselves instead of a padded copy (4 B per stored entry, 8 B per
// mirrored one) and rows of any length are taken in passes of 4*M entries,
// all loads of a pass issued before any is used.
template <bool DOT, int M>
__global__ __launch_bounds__(kNT) void k_spmv_split(const Args *__restrict__ A,
                                                    const float *__restrict__ vin,
                                                    float *__restrict__ yout) {
    constexpr uint32_t kRows = kNT / 4;
    __shared__ float lds[kNT / kWave];
    const uint32_t L = xcd_logical_block();
    const Tile T = load_tile(A->tile1, L);
    if (DOT && as_global(A->s + T.scene)->done) {
        return;
    }
    const uint32_t q = threadIdx.x & 3u, sub = threadIdx.x >> 2;
    gptr<const uint32_t> rp = as_global(A->rp);
    gptr<const uint32_t> col = as_global(A->col);
    gptr<const uint32_t> tp = as_global(A->tp);
    gptr<const uint32_t> tcol = as_global(A->tcol);
    gptr<const uint32_t> tslot = as_global(A->tslot);
    gptr<const float> val = as_global(A->val);
    gptr<const float> cd = as_global(A->cdiag);
    gptr<const float> v = as_global(vin);
    gptr<float> y = as_global(yout);
    const uint32_t lo = T.lo, hi = T.hi;
    float dot = 0.0f, adot = 0.0f;
    // row bounds of the next chunk are loaded while this chunk is in flight
    uint32_t nb[4];
    {
        const uint32_t i0 = lo + sub < hi ? lo + sub : lo;
        nb[0] = rp[i0], nb[1] = rp[i0 + 1], nb[2] = tp[i0], nb[3] = tp[i0 + 1];
    }
    for (uint32_t base = lo; base < hi; base += kRows) {
        const uint32_t i = base + sub;
        const bool valid = i < hi;
        const uint32_t ii = valid ? i : lo;
        const uint32_t r0 = nb[0], ns = nb[1] - nb[0], t0 = nb[2], nm = nb[3] - nb[2];
        if (base + kRows < hi) {
            const uint32_t in = base + kRows + sub;
            const uint32_t iin = in < hi ? in : lo;
            nb[0] = rp[iin], nb[1] = rp[iin + 1], nb[2] = tp[iin], nb[3] = tp[iin + 1];
        }
        const f3 vi = load3(v + 3 * (size_t)ii);
        f3 sum = mk3(0, 0, 0), as = mk3(0, 0, 0);
        // the first pass is unconditional (straight-line loads for the common
        // row); rows longer than 4 M entries take further passes
        for (uint32_t k0 = 0; k0 == 0 || k0 < ns + nm; k0 += 4 * M) {
            uint32_t cj[M], sl[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const uint32_t k = k0 + q + 4 * m;
                if (k < ns) {
                    cj[m] = col[r0 + k];
                    sl[m] = r0 + k;
                } else if (k < ns + nm) {
                    cj[m] = tcol[t0 + k - ns];
                    sl[m] = tslot[t0 + k - ns] | kMirror;
                } else {
                    cj[m] = 0;
                    sl[m] = kNone;
                }
            }
            float blk[M][9];
            f3 vj[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const uint32_t slot = sl[m] == kNone ? 0u : (sl[m] & ~kMirror);
                gptr<const float> a = val + 9 * (size_t)slot;
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    blk[m][c] = a[c];
                }
                vj[m] = load3(v + 3 * (size_t)cj[m]);
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                float *b = blk[m];
                if (sl[m] != kNone && (sl[m] & kMirror)) {
                    float t;
                    t = b[1], b[1] = b[3], b[3] = t;
                    t = b[2], b[2] = b[6], b[6] = t;
                    t = b[5], b[5] = b[7], b[7] = t;
                }
                f3 t = mv(b, vj[m]);
                if (sl[m] == kNone) {
                    t = mk3(0, 0, 0);
                }
                sum = sum + t;
                as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
            }
        }
        if (q == 3) {
            const f3 t = mv(cd + 9 * (size_t)ii, vi);
            sum = sum + t;
            as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
        }
#pragma unroll
        for (int off = 2; off > 0; off >>= 1) {
            sum.x += __shfl_xor(sum.x, off, 4);
            sum.y += __shfl_xor(sum.y, off, 4);
            sum.z += __shfl_xor(sum.z, off, 4);
            if (DOT) {
                as.x += __shfl_xor(as.x, off, 4);
                as.y += __shfl_xor(as.y, off, 4);
                as.z += __shfl_xor(as.z, off, 4);
            }
        }
        if (valid && q == 0) {
            store3(y + 3 * (size_t)i, sum);
            if (DOT) {
                dot = dot + vi.x * sum.x;
                dot = dot + vi.y * sum.y;
                dot = dot + vi.z * sum.z;
                adot = adot + fabsf(vi.x) * as.x;
                adot = adot + fabsf(vi.y) * as.y;
                adot = adot + fabsf(vi.z) * as.z;
            }
        }
    }
    if (DOT) {
        dot = block_sum<kNT>(dot, lds);
        adot = block_sum<kNT>(adot, lds);
        if (threadIdx.x == 0) {
            gptr<float> part = as_global(A->part);
            part[L] = dot;
            part[A->G1 + L] = adot;
        }
    }
}

// Tables of the symmetric form for the tiles of tile1 (one workgroup per
// tile). A tile's mirror entries are contiguous in the transpose table (it is
// ordered by destination row); an entry whose source row lies in the tile too
// gets the next place among the tile's in-tile mirrors, while places last
// (kSymCap), and both the entry (tq, flagged kInTile) and the stored slot it
// mirrors (mq, beside the slot's column) hold it. Every other entry holds the
// slot it mirrors, and every stored slot not mirrored in its tile holds
// kPlaceNone (k_sym_init presets mq). *max_inner receives the largest count.
__global__ __launch_bounds__(kNT) void k_sym_init(uint32_t nnz, const uint32_t *__restrict__ col,
                                                  uint32_t *__restrict__ mq) {
    const uint32_t s = blockIdx.x * kNT + threadIdx.x;
    if (s < nnz) {
        mq[s] = (col[s] << kPlaceBits) | kPlaceNone;
    }
}

__global__ __launch_bounds__(kNT) void k_sym_prep(const Args *__restrict__ A,
                                                  uint32_t *__restrict__ mq,
                                                  uint32_t *__restrict__ tq,
                                                  uint32_t *__restrict__ max_inner) {
    __shared__ uint32_t wsum[kNT / kWave];
    const Tile T = load_tile(A->tile1, blockIdx.x);
    if (T.lo >= T.hi) {
        return;
    }
    gptr<const uint32_t> tp = as_global(A->tp);
    gptr<const uint32_t> tcol = as_global(A->tcol);
    gptr<const uint32_t> tslot = as_global(A->tslot);
    const uint32_t tb = tp[T.lo], te = tp[T.hi];
    const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t running = 0;
    for (uint32_t c = tb; c < te; c += kNT) {
        const uint32_t t = c + threadIdx.x;
        const bool in = t < te && tcol[t] >= T.lo;
        const uint64_t ball = __ballot(in);
        const uint32_t below = (uint32_t)__popcll(ball & ((1ull << lane) - 1ull));
        if (lane == 0) {
            wsum[w] = (uint32_t)__popcll(ball);
        }
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int k = 0; k < kNT / kWave; ++k) {
            before += k < (int)w ? wsum[k] : 0u;
            total += wsum[k];
        }
        __syncthreads();
        if (t < te) {
            const uint32_t pos = running + before + below;
            const uint32_t slot = tslot[t];
            if (in && pos < kSymCap) {
                tq[t] = kInTile | pos;
                mq[slot] = (as_global(A->col)[slot] << kPlaceBits) | pos;
            } else {
                tq[t] = slot;
            }
        }
        running += total;
    }
    if (threadIdx.x == 0) {
        atomicMax(max_inner, running);
    }
}

// Symmetric form of k_spmv_split for one-scene plans. A stored block (i, j)
// whose mirror row j lies in the same tile is read once: row i's lanes apply
// it to v_j for row i and, transposed, to v_i for row j, and park the second
// product in LDS (mq), where row j's lanes take it (tq) instead of reading the
// block and v_i again. Rows go in chunks of 64; a chunk's stored entries come
// before a barrier and its mirrored entries after it (mirrors come from rows
// above, so from this chunk or an earlier one). Every product is the split
// form's float expression and every lane sums its entries in the split form's
// order (lane q takes entries q, q+4, ... of the row's stored-then-mirrored
// list), so y is bit-identical to k_spmv_split.
template <bool DOT, int M>
__global__ __launch_bounds__(kNT) void k_spmv_sym(const Args *__restrict__ A,
                                                  const float *__restrict__ vin,
                                                  float *__restrict__ yout) {
    extern __shared__ float mir[]; // 3 floats per in-tile mirror (kSymCap)
    constexpr uint32_t kRows = kNT / 4;
    __shared__ float lds[kNT / kWave];
    const uint32_t L = xcd_logical_block();
    const Tile T = load_tile(A->tile1, L);
    if (DOT && as_global(A->s + T.scene)->done) {
        return;
    }
    const uint32_t q = threadIdx.x & 3u, sub = threadIdx.x >> 2;
    gptr<const uint32_t> rp = as_global(A->rp);
    gptr<const uint32_t> tp = as_global(A->tp);
    gptr<const uint32_t> tcol = as_global(A->tcol);
    gptr<const uint32_t> mq = as_global(A->mq);
    gptr<const uint32_t> tq = as_global(A->tq);
    gptr<const float> val = as_global(A->val);
    gptr<const float> cd = as_global(A->cdiag);
    gptr<const float> v = as_global(vin);
    gptr<float> y = as_global(yout);
    const uint32_t lo = T.lo, hi = T.hi;
    float dot = 0.0f, adot = 0.0f;
    for (uint32_t base = lo; base < hi; base += kRows) {
        const uint32_t i = base + sub;
        const bool valid = i < hi;
        const uint32_t ii = valid ? i : lo;
        const uint32_t r0 = rp[ii], t0 = tp[ii];
        const uint32_t ns = valid ? rp[ii + 1] - r0 : 0u;
        const uint32_t nm = valid ? tp[ii + 1] - t0 : 0u;
        const f3 vi = load3(v + 3 * (size_t)ii);
        f3 sum = mk3(0, 0, 0), as = mk3(0, 0, 0);
        // stored entries: the row's own products, and the transposed ones of
        // in-tile mirrors into LDS
        for (uint32_t k0 = 0; k0 < ns; k0 += 4 * M) {
            uint32_t cj[M], sl[M], mp[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const uint32_t k = k0 + q + 4 * m;
                const bool ok = k < ns;
                const uint32_t w = ok ? mq[r0 + k] : kPlaceNone;
                sl[m] = ok ? r0 + k : kNone;
                cj[m] = w >> kPlaceBits;
                mp[m] = w & kPlaceNone;
            }
            float blk[M][9];
            f3 vj[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                gptr<const float> a = val + 9 * (size_t)(sl[m] == kNone ? 0u : sl[m]);
#pragma unroll
                for (int c = 0; c < 9; ++c) {
                    blk[m][c] = a[c];
                }
                vj[m] = load3(v + 3 * (size_t)cj[m]);
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if (sl[m] != kNone) {
                    const f3 t = mv(blk[m], vj[m]);
                    sum = sum + t;
                    as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
                    if (mp[m] != kPlaceNone) {
                        const f3 u = mvt(blk[m], vi);
                        mir[3 * mp[m]] = u.x;
                        mir[3 * mp[m] + 1] = u.y;
                        mir[3 * mp[m] + 2] = u.z;
                    }
                }
            }
        }
        __syncthreads();
        // mirrored entries: entry ns + u belongs to lane (ns + u) % 4
        for (uint32_t u0 = (q + 4u - (ns & 3u)) & 3u; u0 < nm; u0 += 4 * M) {
            uint32_t tv[M], src[M], sl[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                const uint32_t u = u0 + 4 * m;
                const bool ok = u < nm;
                // the two loads are independent (no extra dependent round
                // trip for the entries read from memory)
                const uint32_t w = ok ? tq[t0 + u] : 0u;
                src[m] = ok ? tcol[t0 + u] : 0u;
                tv[m] = w & kPlaceNone;
                sl[m] = !ok ? kNone : (w & kInTile) ? kMirror /* taken from LDS */ : w;
            }
            float blk[M][9];
            f3 vj[M];
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if (sl[m] != kNone && sl[m] != kMirror) {
                    gptr<const float> a = val + 9 * (size_t)sl[m];
#pragma unroll
                    for (int c = 0; c < 9; ++c) {
                        blk[m][c] = a[c];
                    }
                    vj[m] = load3(v + 3 * (size_t)src[m]);
                } else {
#pragma unroll
                    for (int c = 0; c < 9; ++c) {
                        blk[m][c] = 0.0f;
                    }
                    vj[m] = mk3(0, 0, 0);
                }
            }
#pragma unroll
            for (int m = 0; m < M; ++m) {
                if (sl[m] == kNone) {
                    continue;
                }
                const f3 t = sl[m] == kMirror
                                 ? mk3(mir[3 * tv[m]], mir[3 * tv[m] + 1], mir[3 * tv[m] + 2])
                                 : mvt(blk[m], vj[m]);
                sum = sum + t;
                as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
            }
        }
        if (q == 3) {
            const f3 t = mv(cd + 9 * (size_t)ii, vi);
            sum = sum + t;
            as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
        }
#pragma unroll
        for (int off = 2; off > 0; off >>= 1) {
            sum.x += __shfl_xor(sum.x, off, 4);
            sum.y += __shfl_xor(sum.y, off, 4);
            sum.z += __shfl_xor(sum.z, off, 4);
            if (DOT) {
                as.x += __shfl_xor(as.x, off, 4);
                as.y += __shfl_xor(as.y, off, 4);
                as.z += __shfl_xor(as.z, off, 4);
            }
        }
        if (valid && q == 0) {
            store3(y + 3 * (size_t)i, sum);
            if (DOT) {
                dot = dot + vi.x * sum.x;
                dot = dot + vi.y * sum.y;
                dot = dot + vi.z * sum.z;
                adot = adot + fabsf(vi.x) * as.x;
                adot = adot + fabsf(vi.y) * as.y;
                adot = adot + fabsf(vi.z) * as.z;
            }
        }
    }
    if (DOT) {
        dot = block_sum<kNT>(dot, lds);
        adot = block_sum<kNT>(adot, lds);
        if (threadIdx.x == 0) {
            gptr<float> part = as_global(A->part);
            part[L] = dot;
            part[A->G1 + L] = adot;
        }
    }
}

// ELL form of k_spmv for rows of at most 4*M entries: 4 lanes per row, lane q
// takes entries q, q+4, ... of the row's fixed-width table, so no row-offset
// load sits in the dependent chain and all M block loads of a lane are issued
// before any is consumed. Padding entries read slot 0 and are discarded.
template <bool DOT, int M>
__global__ __launch_bounds__(kNT) void k_spmv_ell(const Args *__restrict__ A,
                                                  const float *__restrict__ vin,
                                                  float *__restrict__ yout) {
    constexpr uint32_t W = 4 * M;
    constexpr uint32_t kRows = kNT / 4;
    __shared__ float lds[kNT / kWave];
    const uint32_t L = xcd_logical_block();
    const Tile T = load_tile(A->tile1, L);
    if (DOT && as_global(A->s + T.scene)->done) {
        return;
    }
    const uint32_t q = threadIdx.x & 3u, sub = threadIdx.x >> 2;
    gptr<const uint64_t> ent = as_global(reinterpret_cast<const uint64_t *>(A->ent));
    gptr<const float> val = as_global(A->val);
    gptr<const float> cd = as_global(A->cdiag);
    gptr<const float> v = as_global(vin);
    gptr<float> y = as_global(yout);
    const uint32_t lo = T.lo, hi = T.hi;
    float dot = 0.0f, adot = 0.0f;
    // The entries of the next chunk of rows are loaded while this chunk's
    // blocks are in flight: one dependent round trip per chunk instead of two.
    uint64_t wnext[M];
    {
        const uint32_t i0 = lo + sub < hi ? lo + sub : lo;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            wnext[m] = lo < hi ? ent[(size_t)i0 * W + q + 4 * m] : 0;
        }
    }
    for (uint32_t base = lo; base < hi; base += kRows) {
        const uint32_t i = base + sub;
        const bool valid = i < hi;
        const uint32_t ii = valid ? i : lo;
        uint2 e[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            e[m] = make_uint2((uint32_t)wnext[m], (uint32_t)(wnext[m] >> 32));
        }
        if (base + kRows < hi) {
            const uint32_t in = base + kRows + sub;
            const uint32_t iin = in < hi ? in : lo;
#pragma unroll
            for (int m = 0; m < M; ++m) {
                wnext[m] = ent[(size_t)iin * W + q + 4 * m];
            }
        }
        const f3 vi = load3(v + 3 * (size_t)ii);
        f3 sum = mk3(0, 0, 0), as = mk3(0, 0, 0);
        float blk[M][9];
        f3 vj[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const uint32_t slot = e[m].y == kNone ? 0u : (e[m].y & ~kMirror);
            gptr<const float> a = val + 9 * (size_t)slot;
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                blk[m][c] = a[c];
            }
            vj[m] = load3(v + 3 * (size_t)e[m].x);
        }
#pragma unroll
        for (int m = 0; m < M; ++m) {
            float *b = blk[m];
            if (e[m].y != kNone && (e[m].y & kMirror)) {
                float t;
                t = b[1], b[1] = b[3], b[3] = t;
                t = b[2], b[2] = b[6], b[6] = t;
                t = b[5], b[5] = b[7], b[7] = t;
            }
            f3 t = mv(b, vj[m]);
            if (e[m].y == kNone) {
                t = mk3(0, 0, 0);
            }
            sum = sum + t;
            as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
        }
        if (q == 3) {
            const f3 t = mv(cd + 9 * (size_t)ii, vi);
            sum = sum + t;
            as = as + mk3(fabsf(t.x), fabsf(t.y), fabsf(t.z));
        }
#pragma unroll
        for (int off = 2; off > 0; off >>= 1) {
            sum.x += __shfl_xor(sum.x, off, 4);
            sum.y += __shfl_xor(sum.y, off, 4);
            sum.z += __shfl_xor(sum.z, off, 4);
            if (DOT) {
                as.x += __shfl_xor(as.x, off, 4);
                as.y += __shfl_xor(as.y, off, 4);
                as.z += __shfl_xor(as.z, off, 4);
            }
        }
        if (valid && q == 0) {
            store3(y + 3 * (size_t)i, sum);
            if (DOT) {
                dot = dot + vi.x * sum.x;
                dot = dot + vi.y * sum.y;
                dot = dot + vi.z * sum.z;
                adot = adot + fabsf(vi.x) * as.x;
                adot = adot + fabsf(vi.y) * as.y;
                adot = adot + fabsf(vi.z) * as.z;
            }
        }
    }
    if (DOT) {
        dot = block_sum<kNT>(dot, lds);
        adot = block_sum<kNT>(adot, lds);
        if (threadIdx.x == 0) {
            gptr<float> part = as_global(A->part);
            part[L] = dot;
            part[A->G1 + L] = adot;
        }
    }
}

// r = b - y, z = P r, p = z, partials of r.z and |r|_1 (solver.cu:1015-1093).
template <bool FAC>
__global__ __launch_bounds__(kNT) void k_init_residual(const Args *__restrict__ A) {
    __shared__ float lds[kNT / kWave];
    const uint32_t L = blockIdx.x;
    const Tile T = load_tile(A->tile2, L);
    const uint32_t lo = T.lo, hi = T.hi;
    gptr<const float> b = as_global(A->b);
    gptr<const float> yv = as_global(A->Ap);
    gptr<const float> P = as_global(A->P);
    gptr<float> r = as_global(A->r);
    gptr<float> z = as_global(A->z);
    gptr<float> p = as_global(A->p);
    float rz = 0.0f, err = 0.0f;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kNT) {
        const f3 bi = load3(b + 3 * (size_t)i);
        const f3 yi = load3(yv + 3 * (size_t)i);
        const f3 ri = mk3(bi.x + (-1.0f) * yi.x, bi.y + (-1.0f) * yi.y,
                          bi.z + (-1.0f) * yi.z);
        store3(r + 3 * (size_t)i, ri);
        if (!FAC) {
            const f3 zi = mv(P + 9 * (size_t)i, ri);
            store3(z + 3 * (size_t)i, zi);
            store3(p + 3 * (size_t)i, zi);
            rz = rz + ri.x * zi.x;
            rz = rz + ri.y * zi.y;
            rz = rz + ri.z * zi.z;
        }
        err = err + fabsf(ri.x);
        err = err + fabsf(ri.y);
        err = err + fabsf(ri.z);
    }
    if (!FAC) {
        rz = block_sum<kNT>(rz, lds);
    }
    err = block_sum<kNT>(err, lds);
    if (threadIdx.x == 0) {
        gptr<float> part = as_global(A->part);
        if (!FAC) {
            part[2 * A->G1 + L] = rz;
        }
        part[2 * A->G1 + A->G2 + L] = err;
    }
}

__device__ inline float reduce_parts(gptr<const float> part, uint32_t cnt,
                                     float *lds) {
    float acc = 0.0f;
    for (uint32_t k = threadIdx.x; k < cnt; k += kNT) {
        acc = acc + part[k];
    }
    return block_sum<kNT>(acc, lds);
}

// Initial scalars and the bookkeeping state (solver.cu:1020-1031, 1170-1175).
// One workgroup per scene.
__global__ __launch_bounds__(kNT) void k_init_scalars(const Args *__restrict__ A) {
    __shared__ float lds[kNT / kWave];
    const uint32_t sc = blockIdx.x;
    const uint32_t t0 = A->sc2[sc], nt = A->sc2[sc + 1] - t0;
    gptr<const float> part = as_global(A->part);
    const float rz0 = reduce_parts(part + 2 * A->G1 + t0, nt, lds);
    const float err0 = reduce_parts(part + 2 * A->G1 + A->G2 + t0, nt, lds);
    if (threadIdx.x == 0) {
        gptr<Scalars> s = as_global(A->s + sc);
        s->rz[0] = rz0;
        s->rz[1] = rz0;
        s->err0 = err0;
        s->iter = 0;
        s->next_check = 1;
        s->stride = 4;
        s->breakdown = 0;
        s->cause = 0;
        s->break_iter = 0;
        s->done = 0;
        s->status = -1;
        if (s->bad) {
            s->done = 1;
            s->status = 5;
        } else if (err0 == 0.0f) {
            s->done = 1;
            s->status = 0;
            s->rep_iter = 1;
            s->rep_resid = 0.0f;
        }
    }
}

// ---- one CG iteration ------------------------------------------------------

// alpha from the SpMV partials with the curvature guard, then x += alpha p,
// r -= alpha Ap, z = P r and the r.z / |r|_1 partials (solver.cu:431-514).
// k = position of this iteration inside the captured graph.
template <bool FAC>
__global__ __launch_bounds__(kNT) void k_update(const Args *__restrict__ A,
                                                uint32_t k) {
    __shared__ float lds[kNT / kWave];
    const uint32_t L = blockIdx.x;
    const Tile T = load_tile(A->tile2, L);
    gptr<Scalars> s = as_global(A->s + T.scene);
    if (s->done) {
        return;
    }
    gptr<const float> part = as_global(A->part);
    const uint32_t t0 = A->sc1[T.scene], nt = A->sc1[T.scene + 1] - t0;
    const float pAp = reduce_parts(part + t0, nt, lds);
    const float absdot = reduce_parts(part + A->G1 + t0, nt, lds);
    const float rz0 = s->rz[k & 1];
    const float bound = kNoiseSlack * kFloatEps * absdot;
    float alpha = 0.0f;
    if (pAp > bound) {
        alpha = rz0 / pAp;
    } else if (T.first && threadIdx.x == 0) {
        s->breakdown = 1;
        s->cause = (pAp < -bound || pAp != pAp) ? 1 : 3;
        s->break_iter = (int32_t)s->iter;
    }
    const uint32_t lo = T.lo, hi = T.hi;
    gptr<float> x = as_global(A->x);
    gptr<float> r = as_global(A->r);
    gptr<float> z = as_global(A->z);
    gptr<const float> p = as_global(A->p);
    gptr<const float> Ap = as_global(A->Ap);
    gptr<const float> P = as_global(A->P);
    float rz = 0.0f, err = 0.0f;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kNT) {
        const size_t o = 3 * (size_t)i;
        const f3 pi = load3(p + o), api = load3(Ap + o);
        f3 xi = load3(x + o), ri = load3(r + o);
        xi = mk3(fmaf(alpha, pi.x, xi.x), fmaf(alpha, pi.y, xi.y),
                 fmaf(alpha, pi.z, xi.z));
        ri = mk3(fmaf(-alpha, api.x, ri.x), fmaf(-alpha, api.y, ri.y),
                 fmaf(-alpha, api.z, ri.z));
        store3(x + o, xi);
        store3(r + o, ri);
        if (!FAC) { // the factor form applies z = G^T G r in two kernels after
            const f3 zi = mv(P + 9 * (size_t)i, ri);
            store3(z + o, zi);
            rz = rz + ri.x * zi.x;
            rz = rz + ri.y * zi.y;
            rz = rz + ri.z * zi.z;
        }
        err = err + fabsf(ri.x);
        err = err + fabsf(ri.y);
        err = err + fabsf(ri.z);
    }
    if (!FAC) {
        rz = block_sum<kNT>(rz, lds);
    }
    err = block_sum<kNT>(err, lds);
    if (threadIdx.x == 0) {
        gptr<float> pw = as_global(A->part);
        if (!FAC) {
            pw[2 * A->G1 + L] = rz;
        }
        pw[2 * A->G1 + A->G2 + L] = err;
    }
}

// A bf16 factor block (10-entry record, column-major) times a vector, or
// its transpose times the vector.
__device__ inline void load_g16(gptr<const uint16_t> g, size_t blk, float *m) {
    gptr<const uint32_t> w = reinterpret_cast<gptr<const uint32_t>>(g + 10 * blk);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        const uint32_t v = w[k];
        m[2 * k] = __uint_as_float(v << 16);
        if (2 * k + 1 < 9) {
            m[2 * k + 1] = __uint_as_float(v & 0xffff0000u);
        }
    }
}

// The factor passes share the SpMV's shape: P lanes per row, lane q taking
// entries q, q + P, ... of the row, M of them per pass with every load of the
// pass issued before any is used (one dependent round trip per pass), the
// lanes' sums combined by shuffles. 4 x 4 measured fastest for both passes on
// the C3 systems: 188-193 us per factor-PCG iteration, against 211 us at
// 2 x 4, 211 at 4 x 2, 218 at 8 x 2, 254 at 2 x 2, and 197 us with G^T y
// alone at 2 x 4.
constexpr uint32_t kFacGP = 4, kFacGM = 4, kFacTP = 4, kFacTM = 4;

// gy = G r: row j's entries are G_jj, then G_ji for its mirrored entries
// (slot s = (i, j), i < j), read in order from the row-ordered copy Gr.
template <uint32_t P, uint32_t M>
__global__ __launch_bounds__(kNT) void k_fac_g(const Args *__restrict__ A, int init) {
    if (!init && as_global(A->s)->done) {
        return;
    }
    const uint32_t n = A->n;
    const uint32_t q = threadIdx.x % P;
    const uint32_t j = blockIdx.x * (kNT / P) + threadIdx.x / P;
    const bool valid = j < n;
    const uint32_t jj = valid ? j : 0;
    gptr<const uint16_t> Gr = as_global(A->Gr);
    gptr<const float> r = as_global(A->r);
    gptr<const uint32_t> tp = as_global(A->tp), tcol = as_global(A->tcol);
    const uint32_t t0 = tp[jj], cnt = 1 + tp[jj + 1] - t0;
    const size_t g0 = (size_t)jj + t0;
    f3 y = mk3(0, 0, 0);
    for (uint32_t k0 = 0; k0 < cnt; k0 += P * M) {
        uint32_t src[M];
        bool ok[M];
#pragma unroll
        for (uint32_t m = 0; m < M; ++m) {
            const uint32_t k = k0 + q + P * m;
            ok[m] = k < cnt;
            src[m] = (k == 0 || !ok[m]) ? jj : tcol[t0 + k - 1];
        }
        float g[M][9];
        f3 v[M];
#pragma unroll
        for (uint32_t m = 0; m < M; ++m) {
            const uint32_t k = k0 + q + P * m;
            load_g16(Gr, ok[m] ? g0 + k : g0, g[m]);
            v[m] = load3(r + 3 * (size_t)src[m]);
        }
#pragma unroll
        for (uint32_t m = 0; m < M; ++m) {
            if (ok[m]) {
                const float *b = g[m];
                const f3 w = v[m];
                y = y + mk3(b[0] * w.x + b[3] * w.y + b[6] * w.z,
                            b[1] * w.x + b[4] * w.y + b[7] * w.z,
                            b[2] * w.x + b[5] * w.y + b[8] * w.z);
            }
        }
    }
#pragma unroll
    for (uint32_t off = P / 2; off > 0; off >>= 1) {
        y.x += __shfl_xor(y.x, off, P);
        y.y += __shfl_xor(y.y, off, P);
        y.z += __shfl_xor(y.z, off, P);
    }
    if (valid && q == 0) {
        store3(as_global(A->gy) + 3 * (size_t)j, y);
    }
}

// z = G^T gy: row i's entries are G_ii (transposed) and, for its stored
// off-diagonal slots s = (i, j), G_ji^T gy_j (the stored diagonal slot
// carries no factor block). Then the r.z partials (and p = z when init).
template <uint32_t P, uint32_t M>
__global__ __launch_bounds__(kNT) void k_fac_gt(const Args *__restrict__ A, int init) {
    constexpr uint32_t kRows = kNT / P;
    __shared__ float lds[kNT / kWave];
    const uint32_t L = blockIdx.x;
    const Tile T = load_tile(A->tile2, L);
    if (!init && as_global(A->s + T.scene)->done) {
        return;
    }
    const uint32_t lo = T.lo, hi = T.hi;
    const uint32_t q = threadIdx.x % P, sub = threadIdx.x / P;
    gptr<const uint16_t> Gd = as_global(A->Gd), Gl = as_global(A->Gl);
    gptr<const float> gy = as_global(A->gy), r = as_global(A->r);
    gptr<const uint32_t> rp = as_global(A->rp), col = as_global(A->col);
    gptr<float> z = as_global(A->z), p = as_global(A->p);
    float rz = 0.0f;
    // row bounds of the next chunk are loaded while this chunk is in flight
    uint32_t nb[2] = {0u, 0u};
    if (lo < hi) {
        const uint32_t i0 = lo + sub < hi ? lo + sub : lo;
        nb[0] = rp[i0], nb[1] = rp[i0 + 1];
    }
    for (uint32_t base = lo; base < hi; base += kRows) {
        const uint32_t i = base + sub;
        const bool valid = i < hi;
        const uint32_t ii = valid ? i : lo;
        const uint32_t r0 = nb[0], cnt = 1 + nb[1] - r0;
        if (base + kRows < hi) {
            const uint32_t in = base + kRows + sub;
            const uint32_t iin = in < hi ? in : lo;
            nb[0] = rp[iin], nb[1] = rp[iin + 1];
        }
        f3 zi = mk3(0, 0, 0);
        for (uint32_t k0 = 0; k0 < cnt; k0 += P * M) {
            uint32_t src[M];
            bool ok[M];
#pragma unroll
            for (uint32_t m = 0; m < M; ++m) {
                const uint32_t k = k0 + q + P * m;
                ok[m] = k < cnt;
                src[m] = (k == 0 || !ok[m]) ? ii : col[r0 + k - 1];
                ok[m] = ok[m] && (k == 0 || src[m] != ii);
            }
            float g[M][9];
            f3 v[M];
#pragma unroll
            for (uint32_t m = 0; m < M; ++m) {
                const uint32_t k = k0 + q + P * m;
                if (k == 0 || !ok[m]) {
                    load_g16(Gd, ii, g[m]);
                } else {
                    load_g16(Gl, r0 + k - 1, g[m]);
                }
                v[m] = load3(gy + 3 * (size_t)src[m]);
            }
#pragma unroll
            for (uint32_t m = 0; m < M; ++m) {
                if (ok[m]) {
                    const float *b = g[m];
                    const f3 w = v[m];
                    zi = zi + mk3(b[0] * w.x + b[1] * w.y + b[2] * w.z,
                                  b[3] * w.x + b[4] * w.y + b[5] * w.z,
                                  b[6] * w.x + b[7] * w.y + b[8] * w.z);
                }
            }
        }
#pragma unroll
        for (uint32_t off = P / 2; off > 0; off >>= 1) {
            zi.x += __shfl_xor(zi.x, off, P);
            zi.y += __shfl_xor(zi.y, off, P);
            zi.z += __shfl_xor(zi.z, off, P);
        }
        if (valid && q == 0) {
            store3(z + 3 * (size_t)i, zi);
            if (init) {
                store3(p + 3 * (size_t)i, zi);
            }
            const f3 ri = load3(r + 3 * (size_t)i);
            rz = rz + ri.x * zi.x;
            rz = rz + ri.y * zi.y;
            rz = rz + ri.z * zi.z;
        }
    }
    rz = block_sum<kNT>(rz, lds);
    if (threadIdx.x == 0) {
        as_global(A->part)[2 * A->G1 + L] = rz;
    }
}

// beta = rz1 / rz0 (breakdown if rz0 <= 0), p = z + beta p, and -- on one
// thread -- the reference's residual-check schedule (solver.cu:1249-1380).
__global__ __launch_bounds__(kNT) void k_pupdate(const Args *__restrict__ A,
                                                 uint32_t k) {
    __shared__ float lds[kNT / kWave];
    const Tile T = load_tile(A->tile2, blockIdx.x);
    gptr<Scalars> s = as_global(A->s + T.scene);
    if (s->done) {
        return;
    }
    gptr<const float> part = as_global(A->part);
    const uint32_t t0 = A->sc2[T.scene], nt = A->sc2[T.scene + 1] - t0;
    const float rz1 = reduce_parts(part + 2 * A->G1 + t0, nt, lds);
    const float rz0 = s->rz[k & 1];
    const bool rz_ok = rz0 > 0.0f;
    const float beta = rz_ok ? rz1 / rz0 : 0.0f;
    const uint32_t lo = 3 * T.lo, hi = 3 * T.hi;
    gptr<float> p = as_global(A->p);
    gptr<const float> z = as_global(A->z);
    for (uint32_t e = lo + threadIdx.x; e < hi; e += kNT) {
        p[e] = fmaf(beta, p[e], 1.0f * z[e]);
    }
    if (!T.first) {
        return;
    }
    const float err = reduce_parts(part + 2 * A->G1 + A->G2 + t0, nt, lds);
    if (threadIdx.x != 0) {
        return;
    }
    if (!rz_ok && s->cause == 0) {
        s->breakdown = 1;
        s->cause = 2;
        s->break_iter = (int32_t)s->iter;
    }
    s->rz[(k + 1) & 1] = rz1;
    const uint32_t it = s->iter + 1;
    s->iter = it;
    if (A->err_hist) {
        as_global(A->err_hist)[it - 1] = err;
    }
    const bool at_cap = it >= s->max_iter;
    if (it == s->next_check || at_cap) {
        const float reresid = err / s->err0;
        s->rep_iter = it;
        s->rep_resid = reresid;
        int32_t st = -1;
        if (reresid < s->tol) {
            st = 0;
        } else if (s->breakdown) {
            st = s->cause == 3 ? 1 : (s->cause == 2 ? 5 : 4);
        } else if (!isfinite(reresid)) {
            st = 3;
        } else if (at_cap) {
            st = 2;
        }
        if (st >= 0) {
            s->status = st;
            s->done = 1;
        } else {
            if (reresid < 8.0f * s->tol) {
                s->stride = 1;
            }
            s->next_check = it + s->stride;
        }
    }
}

uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

// Workgroups of kNT threads of `kernel` the current device holds at once
// (CUs x resident workgroups per CU), or `fallback` if the runtime cannot say.
template <class K> uint32_t resident_blocks(K kernel, uint32_t fallback, size_t lds = 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kNT, lds) != hipSuccess ||
        cus <= 0 || per_cu <= 0) {
        return fallback;
    }
    return (uint32_t)cus * (uint32_t)per_cu;
}
uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

} // namespace

// ---- plan ------------------------------------------------------------------

struct PcgPlanImpl {
    uint32_t n = 0, nnz = 0;
    Args h{};             // host copy of the kernel arguments
    Args *d_args = nullptr;
    Scalars *h_scal = nullptr; // pinned; a one-scene plan holds two (polled in turn)
    hipEvent_t poll_ev[2] = {nullptr, nullptr}; // status copies of a one-scene plan
    void *mem = nullptr;       // one device allocation for the n-sized scratch
    uint32_t *tp = nullptr, *dslot = nullptr;
    uint32_t *scan_tmp = nullptr, *flags = nullptr;
    uint32_t *pat = nullptr;   // transpose columns | slots, grown with nnz
    size_t pat_cap = 0;
    uint32_t *tsort = nullptr; // sort keys | values | rows | radix scratch, grown with nnz
    size_t tsort_cap = 0;
    void *ell = nullptr;       // ELL entry table (when rows are short)
    size_t ell_cap = 0;        // entries
    uint32_t graph_key = 0xffffffffu; // SpMV form the graph was captured for
    bool split = false;        // short rows read the pattern itself (no ELL table)
    uint32_t split_m = 4;      // entries per lane per pass of the split form
    // symmetric form (one-scene plans): in-tile mirrors through LDS
    bool sym = false;
    int form = PPF_SPMV_AUTO;  // the caller's choice of form
    uint32_t *symmem = nullptr; // mq (nnz) | tq (nnz), grown with nnz
    size_t sym_cap = 0;
    const uint32_t *user_rp = nullptr, *user_col = nullptr; // the caller's pattern
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    // scenes (a plain plan is one scene) and their tiles
    uint32_t nscene = 1;
    std::vector<uint32_t> scene_row; // [nscene + 1] first row of each scene
    uint32_t tile_r1 = 0, tile_r2 = 0; // rows per tile; 0: one resident round
    bool compacted = false;    // tiles cover only the scenes still iterating
    // row order: new row k is original row order[k] (caller's device array);
    // null: the original order
    const uint32_t *order = nullptr;
    bool order_built = false;  // the pattern was built in the current order
    void *omem = nullptr;      // inv | prp | cursor (n-sized) | pcd | pb | px
    uint32_t *inv = nullptr, *prp = nullptr, *cursor = nullptr;
    float *pcd = nullptr, *pb = nullptr, *px = nullptr;
    void *emem = nullptr;      // pcol | smap | pval, grown with nnz
    size_t ecap = 0;
    uint32_t *pcol = nullptr, *smap = nullptr;
    float *pval = nullptr;
    void *tmem = nullptr;      // scalars | scene tile starts | tiles | partials
    size_t tmem_cap = 0;
    // A plan on a fixed device region (the allocation-free solve): every
    // buffer is carved from it, nothing is freed, no graph is captured and
    // the status words are read into host memory the caller provides.
    char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};

namespace {

template <class T> T *carve(char *&cur, size_t count) {
    T *p = reinterpret_cast<T *>(cur);
    size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
    cur += bytes;
    return p;
}

size_t carve_bytes(size_t count, size_t elem) {
    return (count * elem + 255) & ~size_t(255);
}

// Device buffers of a plan: from the plan's fixed region when it has one,
// else from the device allocator.
template <class T> hipError_t plan_alloc(PcgPlanImpl *pl, T **p, size_t bytes) {
    if (!pl->arena) {
        return dev_alloc(p, bytes);
    }
    const size_t b = carve_bytes(bytes, 1);
    if (b > pl->arena_cap - pl->arena_used) {
        return hipErrorOutOfMemory;
    }
    *p = reinterpret_cast<T *>(pl->arena + pl->arena_used);
    pl->arena_used += b;
    return hipSuccess;
}
hipError_t plan_free(PcgPlanImpl *pl, void *p) {
    return pl->arena ? hipSuccess : dev_free(p);
}

// The SpMV kernel a plan launches (its graph is captured for one).
uint32_t spmv_key(const PcgPlanImpl &pl) {
    return (pl.h.Gd ? 1000u : 0u) +
           (pl.sym ? 200u + pl.split_m : pl.split ? 100u + pl.split_m : pl.h.ell_m);
}

constexpr size_t kSymLds = kSymCap * 3 * sizeof(float);

// y = M v with the plan's SpMV form (ELL width 4*m, or the CSR walk).
template <bool DOT>
hipError_t launch_spmv(const PcgPlanImpl &pl, const float *v, float *y,
                       hipStream_t st) {
    const Args &h = pl.h;
    if (pl.sym) {
        switch (pl.split_m) {
        case 2: k_spmv_sym<DOT, 2><<<h.G1, kNT, kSymLds, st>>>(pl.d_args, v, y); break;
        case 3: k_spmv_sym<DOT, 3><<<h.G1, kNT, kSymLds, st>>>(pl.d_args, v, y); break;
        default: k_spmv_sym<DOT, 4><<<h.G1, kNT, kSymLds, st>>>(pl.d_args, v, y); break;
        }
        return hipGetLastError();
    }
    if (pl.split) {
        switch (pl.split_m) {
        case 2: k_spmv_split<DOT, 2><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
        case 3: k_spmv_split<DOT, 3><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
        default: k_spmv_split<DOT, 4><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
        }
        return hipGetLastError();
    }
    switch (h.ell_m) {
    case 1: k_spmv_ell<DOT, 1><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    case 2: k_spmv_ell<DOT, 2><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    case 3: k_spmv_ell<DOT, 3><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    case 4: k_spmv_ell<DOT, 4><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    case 6: k_spmv_ell<DOT, 6><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    default: k_spmv_ell<DOT, 8><<<h.G1, kNT, 0, st>>>(pl.d_args, v, y); break;
    }
    return hipGetLastError();
}

// The two factor passes (G r, then G^T (G r) with the r.z partials).
void launch_fac(const PcgPlanImpl &pl, int init, hipStream_t st) {
    k_fac_g<kFacGP, kFacGM><<<cdiv(pl.h.n, kNT / kFacGP), kNT, 0, st>>>(pl.d_args, init);
    k_fac_gt<kFacTP, kFacTM><<<pl.h.G2, kNT, 0, st>>>(pl.d_args, init);
}

hipError_t launch_iteration(const PcgPlanImpl &pl, uint32_t k, hipStream_t st) {
    const Args &h = pl.h;
    hipError_t e = launch_spmv<true>(pl, h.p, h.Ap, st);
    if (e != hipSuccess) {
        return e;
    }
    if (h.Gd) {
        k_update<true><<<h.G2, kNT, 0, st>>>(pl.d_args, k);
        launch_fac(pl, 0, st);
    } else {
        k_update<false><<<h.G2, kNT, 0, st>>>(pl.d_args, k);
    }
    k_pupdate<<<h.G2, kNT, 0, st>>>(pl.d_args, k);
    return hipGetLastError();
}

} // namespace

namespace {

// Transpose table, row-length choice of the SpMV form and the ELL table for
// a pattern. Buffers sized by nnz (or n * row width) grow when needed.
int set_pattern_arrays(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                       hipStream_t st) {
    const uint32_t n = pl->n;
    Args &h = pl->h;
    if (nnz >= kMirror) {
        set_error("pcg: %u stored blocks exceed the entry index range", nnz);
        return PPF_ERR_ARG;
    }
    if (2 * (size_t)nnz > pl->pat_cap) {
        if (pl->pat) {
            PPF_HIP(plan_free(pl, pl->pat));
            pl->pat = nullptr;
        }
        const size_t cap = pl->arena ? 2 * ((size_t)nnz + nnz / 4 + 256) : 2 * grow_cap(nnz);
        PPF_HIP(plan_alloc(pl, &pl->pat, cap * sizeof(uint32_t)));
        pl->pat_cap = cap;
    }
    uint32_t *tcol = pl->pat, *tslot = pl->pat + pl->pat_cap / 2;
    const size_t sort_words = 3 * (size_t)nnz + radix_sort_scratch_words(nnz);
    if (sort_words > pl->tsort_cap) {
        if (pl->tsort) {
            PPF_HIP(plan_free(pl, pl->tsort));
            pl->tsort = nullptr;
        }
        const size_t cap = pl->arena ? sort_words + sort_words / 4 + 256
                                     : 3 * grow_cap(nnz) + radix_sort_scratch_words(grow_cap(nnz));
        PPF_HIP(plan_alloc(pl, &pl->tsort, cap * sizeof(uint32_t)));
        pl->tsort_cap = cap;
    }
    pl->nnz = nnz;
    h.rp = rp;
    h.col = col;
    h.tp = pl->tp;
    h.tcol = tcol;
    h.tslot = tslot;
    h.dslot = pl->dslot;
    uint32_t *flags = pl->flags;
    PPF_HIP(hipMemsetAsync(pl->tp, 0, (n + 1) * sizeof(uint32_t), st));
    PPF_HIP(hipMemsetAsync(flags, 0, 4 * sizeof(uint32_t), st));
    const unsigned gb = cdiv(n, kNT);
    uint32_t bad = 0;
    k_check_offsets<<<cdiv(n + 1, kNT), kNT, 0, st>>>(n, nnz, rp, flags + 1);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(&bad, flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    if (bad) {
        set_error("pcg: row_offset is not a valid CSR offset array (0 .. nnz, non-decreasing)");
        return PPF_ERR_ARG;
    }
    k_count_transpose<<<gb, kNT, 0, st>>>(n, rp, col, pl->tp, pl->dslot, flags);
    PPF_HIP(hipGetLastError());
    PPF_HIP(exclusive_scan_u32(pl->tp, pl->tp, n + 1, pl->scan_tmp, st));
    if (nnz) {
        uint32_t *key = pl->tsort, *val = key + nnz, *row = val + nnz, *scratch = row + nnz;
        k_transpose_keys<<<gb, kNT, 0, st>>>(n, rp, col, key, val, row);
        PPF_HIP(hipGetLastError());
        uint32_t bits = 1;
        while (bits < 32 && (n >> bits) != 0) {
            ++bits;
        }
        PPF_HIP(radix_sort_pairs(key, val, nnz, bits, scratch, st));
        k_transpose_place<<<cdiv(nnz, kNT), kNT, 0, st>>>(nnz, pl->tp + n, val, row, tcol, tslot);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipMemcpyAsync(&bad, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    if (bad) {
        set_error("pcg: %u pattern entries are not strictly ascending upper-triangle "
                  "columns < n",
                  bad);
        return PPF_ERR_ARG;
    }
    // Choose the SpMV form from the longest row (stored + mirrored blocks).
    PPF_HIP(hipMemsetAsync(flags + 2, 0, sizeof(uint32_t), st));
    k_row_len_max<<<gb, kNT, 0, st>>>(n, rp, pl->tp, flags + 2);
    PPF_HIP(hipGetLastError());
    uint32_t maxlen = 0;
    PPF_HIP(hipMemcpyAsync(&maxlen, flags + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    h.ell_m = 0;
    for (uint32_t m : {1u, 2u, 3u, 4u, 6u, 8u}) {
        // Padding entries read block 0, so the ELL form needs nnz >= 1.
        if (nnz > 0 && maxlen <= 4 * m) {
            h.ell_m = m;
            break;
        }
    }
    // SpMV form: the padded ELL table when every row fits it with at most
    // twice the average row (measured fastest on mesh-like patterns), else
    // the split form,
    // whose pass covers the average row (stored + mirrored entries, 2 nnz - n
    // with every diagonal stored) and takes long rows in further passes.
    // A one-scene plan on the device allocator takes the split form's
    // symmetric variant (in-tile mirrors through LDS) wherever it would take
    // the split form. The caller may fix the form instead.
    {
        const double avg = n ? (2.0 * nnz - n) / n : 0.0;
        const uint32_t m = (uint32_t)((avg + 1.0 + 3.999) / 4.0);
        pl->split_m = m < 2 ? 2 : (m > 4 ? 4 : m);
        const bool sym_ok = pl->nscene == 1 && !pl->arena && n <= kSymMaxRows;
        switch (pl->form) {
        case PPF_SPMV_ELL:
            if (!h.ell_m) {
                set_error("pcg: rows of up to %u entries do not fit the ELL form (32 at most)",
                          maxlen);
                return PPF_ERR_ARG;
            }
            pl->split = false;
            break;
        case PPF_SPMV_SPLIT: pl->split = true; break;
        case PPF_SPMV_SYM:
            if (!sym_ok) {
                set_error("pcg: the symmetric SpMV form needs a one-scene plan of at most %u rows",
                          kSymMaxRows);
                return PPF_ERR_ARG;
            }
            pl->split = true;
            break;
        default: pl->split = !(h.ell_m && 4.0 * h.ell_m <= 2.0 * avg); break;
        }
        pl->sym = pl->split && sym_ok && pl->form != PPF_SPMV_SPLIT;
    }
    h.ent = nullptr;
    if (h.ell_m && !pl->split) {
        const uint32_t W = 4 * h.ell_m;
        if ((size_t)n * W > pl->ell_cap) {
            if (pl->ell) {
                PPF_HIP(plan_free(pl, pl->ell));
                pl->ell = nullptr;
            }
            PPF_HIP(plan_alloc(pl, &pl->ell, (size_t)n * W * sizeof(uint2)));
            pl->ell_cap = (size_t)n * W;
        }
        h.ent = static_cast<const uint2 *>(pl->ell);
        k_build_ell<<<gb, kNT, 0, st>>>(n, W, rp, col, pl->tp, tcol, tslot,
                                        static_cast<uint2 *>(pl->ell));
        PPF_HIP(hipGetLastError());
    }
    return PPF_OK;
}

// Tables of the symmetric form for the current pattern and tiles.
int build_sym(PcgPlanImpl *pl, hipStream_t st) {
    Args &h = pl->h;
    h.mq = h.tq = nullptr;
    if (!pl->sym || !h.tile1) {
        return PPF_OK;
    }
    const size_t need = 2 * (size_t)std::max(pl->nnz, 1u);
    if (need > pl->sym_cap) {
        if (pl->symmem) {
            PPF_HIP(dev_free(pl->symmem));
            pl->symmem = nullptr;
            pl->sym_cap = 0;
        }
        const size_t cap = 2 * grow_cap(std::max(pl->nnz, 1u));
        PPF_HIP(dev_alloc(&pl->symmem, cap * sizeof(uint32_t)));
        pl->sym_cap = cap;
    }
    uint32_t *mq = pl->symmem, *tq = pl->symmem + pl->sym_cap / 2;
    if (pl->nnz) {
        k_sym_init<<<cdiv(pl->nnz, kNT), kNT, 0, st>>>(pl->nnz, h.col, mq);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipMemsetAsync(pl->flags + 3, 0, sizeof(uint32_t), st));
    h.mq = mq;
    h.tq = tq;
    PPF_HIP(hipMemcpyAsync(pl->d_args, &h, sizeof(Args), hipMemcpyHostToDevice, st));
    k_sym_prep<<<h.G1, kNT, 0, st>>>(pl->d_args, mq, tq, pl->flags + 3);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

// The plan's pattern: the caller's, or the caller's rebuilt in the plan's row
// order (the original entries are checked the same way either way).
int set_pattern_ordered(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                        hipStream_t st);

int set_pattern(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                hipStream_t st) {
    pl->user_rp = rp;
    pl->user_col = col;
    int rc = pl->order ? set_pattern_ordered(pl, nnz, rp, col, st)
                       : set_pattern_arrays(pl, nnz, rp, col, st);
    return rc == PPF_OK ? build_sym(pl, st) : rc;
}

int set_pattern_ordered(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                        hipStream_t st) {
    const uint32_t n = pl->n;
    if (nnz >= kMirror) {
        set_error("pcg: %u stored blocks exceed the entry index range", nnz);
        return PPF_ERR_ARG;
    }
    if ((size_t)nnz + 1 > pl->ecap) {
        if (pl->emem) {
            PPF_HIP(dev_free(pl->emem));
            pl->emem = nullptr;
            pl->ecap = 0;
        }
        const size_t cap = grow_cap(nnz + 1);
        PPF_HIP(dev_alloc(&pl->emem, cap * (4 + 4 + 36)));
        pl->ecap = cap;
        pl->pcol = static_cast<uint32_t *>(pl->emem);
        pl->smap = pl->pcol + cap;
        pl->pval = reinterpret_cast<float *>(pl->smap + cap);
    }
    uint32_t *flags = pl->flags;
    uint32_t bad = 0;
    PPF_HIP(hipMemsetAsync(flags + 1, 0, sizeof(uint32_t), st));
    k_check_offsets<<<cdiv(n + 1, kNT), kNT, 0, st>>>(n, nnz, rp, flags + 1);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(&bad, flags + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    if (bad) {
        set_error("pcg: row_offset is not a valid CSR offset array (0 .. nnz, non-decreasing)");
        return PPF_ERR_ARG;
    }
    const unsigned gb = cdiv(n, kNT);
    PPF_HIP(hipMemsetAsync(pl->cursor, 0, (n + 1) * sizeof(uint32_t), st));
    PPF_HIP(hipMemsetAsync(flags, 0, sizeof(uint32_t), st));
    k_order_count<<<gb, kNT, 0, st>>>(n, rp, col, pl->inv, pl->cursor, flags);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipMemcpyAsync(&bad, flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    if (bad) {
        set_error("pcg: %u pattern entries are not strictly ascending upper-triangle "
                  "columns < n",
                  bad);
        return PPF_ERR_ARG;
    }
    PPF_HIP(exclusive_scan_u32(pl->cursor, pl->prp, n + 1, pl->scan_tmp, st));
    PPF_HIP(hipMemsetAsync(pl->cursor, 0, (n + 1) * sizeof(uint32_t), st));
    k_order_fill<<<gb, kNT, 0, st>>>(n, rp, col, pl->inv, pl->prp, pl->cursor, pl->pcol,
                                     pl->smap);
    PPF_HIP(hipGetLastError());
    k_order_sort<<<gb, kNT, 0, st>>>(n, pl->prp, pl->pcol, pl->smap);
    PPF_HIP(hipGetLastError());
    const int rc = set_pattern_arrays(pl, nnz, pl->prp, pl->pcol, st);
    pl->order_built = rc == PPF_OK;
    return rc;
}

// Tiles of the SpMV (tile1) and of the vector kernels (tile2). Rows per tile
// r1 / r2 == 0: one resident round over the whole system (one scene), rows
// in multiples of 64. Otherwise every scene is cut into tiles of at most
// r1 / r2 rows (an empty scene gets one empty tile); tile1 is padded to a
// multiple of 8 (the XCD mapping) with empty tiles of the last scene. Each
// scene's partial sums are then those of its own contiguous tile range.
// `active` (per scene, or null for all): scenes left out get no tiles (their
// solves are over). The scalars keep their place across calls.
int set_tiles(PcgPlanImpl *pl, uint32_t r1, uint32_t r2, hipStream_t st,
              const std::vector<uint8_t> *active = nullptr) {
    const uint32_t n = pl->n, ns = pl->nscene;
    Args &h = pl->h;
    std::vector<Tile> t1, t2;
    std::vector<uint32_t> sc1(ns + 1), sc2(ns + 1);
    auto cut = [&](std::vector<Tile> &t, std::vector<uint32_t> &sc, uint32_t rows) {
        for (uint32_t c = 0; c < ns; ++c) {
            const uint32_t a = pl->scene_row[c], b = pl->scene_row[c + 1];
            sc[c] = (uint32_t)t.size();
            if (active && !(*active)[c]) {
                continue;
            }
            if (a == b) {
                t.push_back(Tile{a, a, c, 1u});
            }
            for (uint32_t lo = a; lo < b; lo += rows) {
                t.push_back(Tile{lo, std::min(b, lo + rows), c, lo == a ? 1u : 0u});
            }
        }
    };
    if (r1 == 0 || r2 == 0) {
        if (ns != 1) {
            set_error("pcg: a plan of %u scenes needs explicit tile sizes", ns);
            return PPF_ERR_ARG;
        }
        // One round of resident workgroups per kernel (a second, partly
        // filled round costs as much as the first): the SpMV's register use
        // allows 5 workgroups per CU, the vector kernels 8.
        uint32_t res1 = resident_blocks(k_spmv_split<true, 4>, 1280u);
        if (!pl->arena) {
            // the symmetric form's LDS may hold fewer workgroups per CU
            res1 = std::min(res1, resident_blocks(k_spmv_sym<true, 4>, 1280u, kSymLds));
        }
        const uint32_t g1 = round8(std::min(res1, cdiv(n, 64)));
        const uint32_t g2 = std::min(resident_blocks(k_update<false>, 2048u), cdiv(n, 64));
        r1 = cdiv(cdiv(n, g1), 64) * 64;
        r2 = cdiv(cdiv(n, g2), 64) * 64;
        cut(t1, sc1, r1);
        cut(t2, sc2, r2);
        while (t1.size() < g1) {
            t1.push_back(Tile{n, n, 0u, 0u});
        }
        pl->tile_r1 = pl->tile_r2 = 0;
    } else {
        cut(t1, sc1, r1);
        cut(t2, sc2, r2);
        pl->tile_r1 = r1, pl->tile_r2 = r2;
    }
    pl->compacted = active != nullptr;
    while (t1.size() % 8) {
        t1.push_back(Tile{n, n, ns - 1, 0u});
    }
    sc1[ns] = (uint32_t)t1.size();
    sc2[ns] = (uint32_t)t2.size();
    h.G1 = (uint32_t)t1.size();
    h.G2 = (uint32_t)t2.size();
    h.nscene = ns;
    const size_t bytes = carve_bytes(ns, sizeof(Scalars)) + 2 * carve_bytes(ns + 1, 4) +
                         carve_bytes(t1.size(), sizeof(Tile)) +
                         carve_bytes(t2.size(), sizeof(Tile)) +
                         carve_bytes(2 * t1.size() + 2 * t2.size(), 4);
    if (bytes > pl->tmem_cap) {
        if (active) {
            set_error("pcg: compacted tiles outgrow the full tiling");
            return PPF_ERR_ARG;
        }
        if (pl->tmem) {
            PPF_HIP(plan_free(pl, pl->tmem));
            pl->tmem = nullptr;
            pl->tmem_cap = 0;
        }
        PPF_HIP(plan_alloc(pl, &pl->tmem, bytes));
        pl->tmem_cap = bytes;
    }
    char *cur = static_cast<char *>(pl->tmem);
    h.s = carve<Scalars>(cur, ns);
    uint32_t *dsc1 = carve<uint32_t>(cur, ns + 1);
    uint32_t *dsc2 = carve<uint32_t>(cur, ns + 1);
    Tile *d1 = carve<Tile>(cur, t1.size());
    Tile *d2 = carve<Tile>(cur, t2.size());
    h.tile1 = d1, h.tile2 = d2, h.sc1 = dsc1, h.sc2 = dsc2;
    h.part = carve<float>(cur, 2 * t1.size() + 2 * t2.size());
    PPF_HIP(hipMemcpyAsync(d1, t1.data(), t1.size() * sizeof(Tile), hipMemcpyHostToDevice, st));
    PPF_HIP(hipMemcpyAsync(d2, t2.data(), t2.size() * sizeof(Tile), hipMemcpyHostToDevice, st));
    PPF_HIP(hipMemcpyAsync(dsc1, sc1.data(), sc1.size() * 4, hipMemcpyHostToDevice, st));
    PPF_HIP(hipMemcpyAsync(dsc2, sc2.data(), sc2.size() * 4, hipMemcpyHostToDevice, st));
    PPF_HIP(hipStreamSynchronize(st)); // the host vectors end here
    pl->graph_key = 0xffffffffu;      // grids changed: capture again
    return build_sym(pl, st);
}

// Capture kGraphIters iterations; they read every pointer through d_args, so
// the graph serves every solve and pattern of the plan with the same SpMV form.
int capture_graph(PcgPlanImpl *pl) {
    if (pl->exec) {
        (void)hipGraphExecDestroy(pl->exec);
        pl->exec = nullptr;
    }
    if (!pl->cap) {
        PPF_HIP(hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking));
    }
    hipGraph_t g = nullptr;
    PPF_HIP(hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (uint32_t k = 0; k < (uint32_t)kGraphIters; ++k) {
        hipError_t e2 = launch_iteration(*pl, k, pl->cap);
        if (le == hipSuccess) {
            le = e2;
        }
    }
    hipError_t ce = hipStreamEndCapture(pl->cap, &g);
    if (ce == hipSuccess && le != hipSuccess) {
        ce = le;
    }
    if (ce != hipSuccess) {
        set_error("pcg: graph capture: %s", hipGetErrorString(ce));
        return PPF_ERR_HIP;
    }
    ce = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ce != hipSuccess) {
        pl->exec = nullptr;
        set_error("pcg: hipGraphInstantiate: %s", hipGetErrorString(ce));
        return PPF_ERR_HIP;
    }
    pl->graph_key = spmv_key(*pl);
    return PPF_OK;
}

} // namespace

namespace {

// Every column of a row of scene c must lie in scene c's rows.
__global__ void k_check_scenes(uint32_t n, uint32_t nscene, const uint32_t *__restrict__ srow,
                               const uint32_t *__restrict__ rp,
                               const uint32_t *__restrict__ col, uint32_t *__restrict__ bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t lo = 0, hi = nscene; // last scene whose first row <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (srow[mid] <= i) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const uint32_t a = srow[lo], b = srow[lo + 1];
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
        if (col[k] < a || col[k] >= b) {
            atomicAdd(bad, 1u);
        }
    }
}

} // namespace

namespace {

// Everything a plan holds for its pattern and scenes, on a fresh PcgPlanImpl
// (n-sized scratch, status words, pattern arrays, tiles, the iteration graph).
// A plan with a fixed region takes its buffers from it and captures no graph;
// its status words live in the host memory already in pl->h_scal.
int plan_init(PcgPlanImpl *pl, uint32_t n, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
              uint32_t nscene, const uint32_t *scene_row, hipStream_t st) {
    pl->n = n;
    pl->nscene = nscene;
    pl->scene_row.assign(nscene + 1, 0u);
    if (nscene > 1) {
        std::copy(scene_row, scene_row + nscene + 1, pl->scene_row.begin());
    } else {
        pl->scene_row[1] = n;
    }
    Args &h = pl->h;
    h.n = n;

    const size_t n3 = 3 * (size_t)n;
    const size_t scan_words = scan_scratch_words(n + 1);
    size_t bytes = carve_bytes(n + 1, 4) + carve_bytes(n, 4) +
                   carve_bytes(9 * (size_t)n, 4) + carve_bytes(n3, 4) * 5 +
                   carve_bytes(1, sizeof(Args)) + carve_bytes(scan_words, 4) +
                   carve_bytes(4, 4) + carve_bytes(nscene + 1, 4);
    hipError_t e = plan_alloc(pl, &pl->mem, bytes);
    if (e != hipSuccess) {
        set_error("pcg_plan_create: %zu bytes of scratch: %s", bytes, hipGetErrorString(e));
        return PPF_ERR_HIP;
    }
    char *cur = static_cast<char *>(pl->mem);
    pl->tp = carve<uint32_t>(cur, n + 1);
    pl->dslot = carve<uint32_t>(cur, n);
    h.P = carve<float>(cur, 9 * (size_t)n);
    h.r = carve<float>(cur, n3);
    h.z = carve<float>(cur, n3);
    h.p = carve<float>(cur, n3);
    h.Ap = carve<float>(cur, n3);
    h.gy = carve<float>(cur, n3);
    pl->d_args = carve<Args>(cur, 1);
    pl->scan_tmp = carve<uint32_t>(cur, scan_words);
    pl->flags = carve<uint32_t>(cur, 4);
    uint32_t *d_srow = carve<uint32_t>(cur, nscene + 1);
    if (!pl->arena) {
        const uint32_t nbuf = nscene == 1 ? 2u : nscene;
        e = host_alloc(reinterpret_cast<void **>(&pl->h_scal), nbuf * sizeof(Scalars));
        if (e != hipSuccess) {
            set_error("pcg_plan_create: hipHostMalloc: %s", hipGetErrorString(e));
            return PPF_ERR_HIP;
        }
        if (nscene == 1) {
            for (hipEvent_t &ev : pl->poll_ev) {
                e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (e != hipSuccess) {
                    set_error("pcg_plan_create: hipEventCreate: %s", hipGetErrorString(e));
                    return PPF_ERR_HIP;
                }
            }
        }
    }
    int rc = set_pattern(pl, nnz, rp, col, st);
    if (rc == PPF_OK && nscene > 1) {
        uint32_t bad = 0;
        e = hipMemcpyAsync(d_srow, pl->scene_row.data(), (nscene + 1) * 4,
                           hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemsetAsync(pl->flags + 3, 0, 4, st);
        if (e == hipSuccess) {
            k_check_scenes<<<cdiv(n, kNT), kNT, 0, st>>>(n, nscene, d_srow, rp, col,
                                                         pl->flags + 3);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&bad, pl->flags + 3, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            set_error("pcg_plan_create: %s", hipGetErrorString(e));
            rc = PPF_ERR_HIP;
        } else if (bad) {
            set_error("pcg_plan_create: %u entries couple two scenes", bad);
            rc = PPF_ERR_ARG;
        }
    }
    if (rc == PPF_OK) {
        // a batch starts from mid-size tiles; the autotuner refines them
        rc = nscene > 1 ? set_tiles(pl, 512u, 512u, st) : set_tiles(pl, 0u, 0u, st);
    }
    if (rc == PPF_OK && !pl->arena) {
        rc = capture_graph(pl);
    }
    return rc;
}

} // namespace

int pcg_plan_create_scenes(uint32_t n, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                           uint32_t nscene, const uint32_t *scene_row, hipStream_t st,
                           PcgPlanImpl **out) {
    if (!out || n == 0 || !rp || !col || nnz >= kMirror || nscene == 0 ||
        (nscene > 1 && !scene_row)) {
        set_error("pcg_plan_create: invalid argument");
        return PPF_ERR_ARG;
    }
    if (nscene > 1) {
        bool ok = scene_row[0] == 0 && scene_row[nscene] == n;
        for (uint32_t c = 0; ok && c < nscene; ++c) {
            ok = scene_row[c] <= scene_row[c + 1];
        }
        if (!ok) {
            set_error("pcg_plan_create: scene rows must rise from 0 to n");
            return PPF_ERR_ARG;
        }
    }
    auto *pl = new (std::nothrow) PcgPlanImpl();
    if (!pl) {
        return PPF_ERR_ARG;
    }
    const int rc = plan_init(pl, n, nnz, rp, col, nscene, scene_row, st);
    if (rc != PPF_OK) {
        ppf::pcg_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return PPF_OK;
}

int pcg_plan_create(uint32_t n, uint32_t nnz, const uint32_t *rp,
                    const uint32_t *col, hipStream_t st, PcgPlanImpl **out) {
    return pcg_plan_create_scenes(n, nnz, rp, col, 1u, nullptr, st, out);
}

uint32_t pcg_plan_scenes(const PcgPlanImpl *pl) { return pl->nscene; }

PcgPattern pcg_pattern(const PcgPlanImpl *pl) {
    const Args &h = pl->h;
    return PcgPattern{pl->n, pl->nnz, h.rp, h.col, h.tp, h.tcol, h.tslot, h.dslot};
}

int pcg_plan_repattern(PcgPlanImpl *pl, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                       hipStream_t st) {
    if (!pl || !rp || !col) {
        set_error("pcg_plan_repattern: invalid argument");
        return PPF_ERR_ARG;
    }
    int rc = set_pattern(pl, nnz, rp, col, st);
    if (rc == PPF_OK && spmv_key(*pl) != pl->graph_key) {
        rc = capture_graph(pl);
    }
    return rc;
}

void pcg_plan_destroy(PcgPlanImpl *pl) {
    if (!pl) {
        return;
    }
    if (pl->exec) {
        (void)hipGraphExecDestroy(pl->exec);
    }
    if (pl->cap) {
        (void)hipStreamDestroy(pl->cap);
    }
    for (hipEvent_t ev : pl->poll_ev) {
        if (ev) {
            (void)hipEventDestroy(ev);
        }
    }
    if (pl->h_scal) {
        (void)host_free(pl->h_scal);
    }
    if (pl->mem) {
        (void)dev_free(pl->mem);
    }
    if (pl->pat) {
        (void)dev_free(pl->pat);
    }
    if (pl->tsort) {
        (void)dev_free(pl->tsort);
    }
    if (pl->ell) {
        (void)dev_free(pl->ell);
    }
    if (pl->tmem) {
        (void)dev_free(pl->tmem);
    }
    if (pl->omem) {
        (void)dev_free(pl->omem);
    }
    if (pl->emem) {
        (void)dev_free(pl->emem);
    }
    if (pl->symmem) {
        (void)dev_free(pl->symmem);
    }
    delete pl;
}

int pcg_spmv(PcgPlanImpl *pl, const float *val, const float *cdiag,
             const float *x, float *y, hipStream_t st) {
    Args &h = pl->h;
    if (pl->compacted) { // the last solve ended on a subset of the scenes
        int rc = set_tiles(pl, pl->tile_r1, pl->tile_r2, st);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    const float *ov, *oc, *ob;
    int rc = pcg_order_system(pl, val, cdiag, nullptr, &ov, &oc, &ob, st);
    if (rc != PPF_OK) {
        return rc;
    }
    const float *xin = x;
    float *yout = y;
    if (pl->order) {
        k_order_rows<false><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, x, pl->px);
        PPF_HIP(hipGetLastError());
        xin = pl->px, yout = pl->pb;
    }
    h.val = ov;
    h.cdiag = oc;
    h.b = nullptr;
    h.x = nullptr;
    h.err_hist = nullptr;
    PPF_HIP(hipMemcpyAsync(pl->d_args, &h, sizeof(Args), hipMemcpyHostToDevice,
                           st));
    PPF_HIP(hipMemsetAsync(h.s, 0, pl->nscene * sizeof(Scalars), st));
    k_diag_invert<<<h.G2, kNT, 0, st>>>(pl->d_args);
    PPF_HIP(hipGetLastError());
    PPF_HIP(launch_spmv<false>(*pl, xin, yout, st));
    if (pl->order) {
        k_order_rows<true><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, pl->pb, y);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipStreamSynchronize(st));
    return PPF_OK;
}

int pcg_time_spmv(PcgPlanImpl *pl, const float *val, const float *cdiag,
                  const float *x, float *y, uint32_t reps, float *avg_ms,
                  hipStream_t st) {
    if (!pl || !avg_ms || reps == 0) {
        set_error("pcg_time_spmv: invalid argument");
        return PPF_ERR_ARG;
    }
    Args &h = pl->h;
    if (pl->compacted) { // the last solve ended on a subset of the scenes
        int rc = set_tiles(pl, pl->tile_r1, pl->tile_r2, st);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    const float *ov, *oc, *ob;
    int rc = pcg_order_system(pl, val, cdiag, nullptr, &ov, &oc, &ob, st);
    if (rc != PPF_OK) {
        return rc;
    }
    if (pl->order) { // time the plan's own kernel on the reordered vectors
        k_order_rows<false><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, x, pl->px);
        PPF_HIP(hipGetLastError());
        x = pl->px, y = pl->pb;
    }
    h.val = ov;
    h.cdiag = oc;
    h.b = nullptr;
    h.x = nullptr;
    h.err_hist = nullptr;
    PPF_HIP(hipMemcpyAsync(pl->d_args, &h, sizeof(Args), hipMemcpyHostToDevice,
                           st));
    PPF_HIP(hipMemsetAsync(h.s, 0, pl->nscene * sizeof(Scalars), st));
    k_diag_invert<<<h.G2, kNT, 0, st>>>(pl->d_args);
    PPF_HIP(hipGetLastError());
    hipEvent_t e0, e1;
    PPF_HIP(hipEventCreate(&e0));
    PPF_HIP(hipEventCreate(&e1));
    PPF_HIP(launch_spmv<true>(*pl, x, y, st)); // warm
    PPF_HIP(hipEventRecord(e0, st));
    for (uint32_t r = 0; r < reps; ++r) {
        PPF_HIP(launch_spmv<true>(*pl, x, y, st));
    }
    PPF_HIP(hipEventRecord(e1, st));
    PPF_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    PPF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / (float)reps;
    return PPF_OK;
}

int pcg_solve(PcgPlanImpl *pl, const float *val, const float *cdiag,
              const float *b, float *x, float tol, uint32_t max_iter,
              float *err_hist, ppf_pcg_result *res, hipStream_t st, const PcgFactor *fac) {
    if (!pl || !val || !cdiag || !b || !x || !res || max_iter == 0 ||
        (fac && (!fac->Gd || !fac->Gl || !fac->Gr))) {
        set_error("pcg_solve: invalid argument");
        return PPF_ERR_ARG;
    }
    const uint32_t ns = pl->nscene;
    if (ns > 1 && (fac || err_hist)) {
        set_error("pcg_solve: a plan of %u scenes solves with block Jacobi and no history", ns);
        return PPF_ERR_ARG;
    }
    Args &h = pl->h;
    if (pl->compacted) { // the last solve ended on a subset of the scenes
        int rc = set_tiles(pl, pl->tile_r1, pl->tile_r2, st);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    if (pl->order && ns > 1) {
        set_error("pcg_solve: a reordered plan holds one scene");
        return PPF_ERR_ARG;
    }
    {
        const float *ov, *oc, *ob;
        int rc = pcg_order_system(pl, val, cdiag, b, &ov, &oc, &ob, st);
        if (rc != PPF_OK) {
            return rc;
        }
        val = ov, cdiag = oc, b = ob;
    }
    float *const x_caller = x;
    if (pl->order) {
        k_order_rows<false><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, x, pl->px);
        PPF_HIP(hipGetLastError());
        x = pl->px;
    }
    h.val = val;
    h.cdiag = cdiag;
    h.b = b;
    h.x = x;
    h.err_hist = err_hist;
    h.Gd = fac ? fac->Gd : nullptr;
    h.Gl = fac ? fac->Gl : nullptr;
    h.Gr = fac ? fac->Gr : nullptr;
    const bool Gd = fac != nullptr;
    if (!pl->arena && spmv_key(*pl) != pl->graph_key) {
        // the iteration graph holds the preconditioner's kernels
        int rc = capture_graph(pl);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    PPF_HIP(hipMemcpyAsync(pl->d_args, &h, sizeof(Args), hipMemcpyHostToDevice,
                           st));
    // the pinned status buffer stages the initial scalars of every scene
    PPF_HIP(hipStreamSynchronize(st));
    for (uint32_t c = 0; c < ns; ++c) {
        Scalars init{};
        init.tol = tol;
        init.max_iter = max_iter;
        pl->h_scal[c] = init;
    }
    PPF_HIP(hipMemcpyAsync(h.s, pl->h_scal, ns * sizeof(Scalars), hipMemcpyHostToDevice, st));
    if (!Gd) {
        k_diag_invert<<<h.G2, kNT, 0, st>>>(pl->d_args);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(launch_spmv<false>(*pl, x, h.Ap, st));
    if (Gd) {
        k_init_residual<true><<<h.G2, kNT, 0, st>>>(pl->d_args);
        launch_fac(*pl, 1, st);
    } else {
        k_init_residual<false><<<h.G2, kNT, 0, st>>>(pl->d_args);
    }
    PPF_HIP(hipGetLastError());
    k_init_scalars<<<ns, kNT, 0, st>>>(pl->d_args);
    PPF_HIP(hipGetLastError());
    const uint32_t max_launch = max_iter / kGraphIters + 2;
    if (pl->exec && ns == 1) {
        // One scene: keep one graph of iterations in flight while the host
        // reads the status copied after the graph before it (two pinned
        // status words, an event each), so the device never idles on the
        // host's poll. Graphs launched after the solve is done exit at entry.
        Scalars *buf[2] = {pl->h_scal, pl->h_scal + 1};
        for (uint32_t k = 0; k < 2; ++k) {
            PPF_HIP(hipGraphLaunch(pl->exec, st));
            PPF_HIP(hipMemcpyAsync(buf[k], h.s, sizeof(Scalars), hipMemcpyDeviceToHost, st));
            PPF_HIP(hipEventRecord(pl->poll_ev[k], st));
        }
        uint32_t k = 0;
        for (;; ++k) {
            PPF_HIP(hipEventSynchronize(pl->poll_ev[k & 1]));
            if (buf[k & 1]->done != 0) {
                break;
            }
            if (k + 2 >= max_launch) {
                (void)hipStreamSynchronize(st);
                set_error("pcg_solve: bookkeeping did not terminate");
                return PPF_ERR_SOLVER;
            }
            PPF_HIP(hipGraphLaunch(pl->exec, st));
            PPF_HIP(hipMemcpyAsync(buf[k & 1], h.s, sizeof(Scalars), hipMemcpyDeviceToHost, st));
            PPF_HIP(hipEventRecord(pl->poll_ev[k & 1], st));
        }
        const Scalars done = *buf[k & 1];
        if (pl->order) {
            k_order_rows<true><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, x,
                                                                     x_caller);
            PPF_HIP(hipGetLastError());
        }
        PPF_HIP(hipStreamSynchronize(st)); // the graph still in flight ends here
        res[0].iter = done.rep_iter;
        res[0].resid = done.rep_resid;
        res[0].status = done.status;
        res[0].break_cause = done.cause;
        res[0].break_iter = done.break_iter;
        res[0].err0 = done.err0;
        return PPF_OK;
    }
    // Replay graphs until the device bookkeeping reports every scene done.
    // The first status read follows one graph; the iteration cap bounds the
    // loop. With many scenes, once the scenes still iterating hold under 3/4
    // of the tiled rows, the tiles are cut again over those scenes only (the
    // partial sums are written and read within one iteration, so the switch
    // between graphs is safe).
    uint32_t tiled_rows = pl->n;
    std::vector<uint8_t> active(ns, 1u);
    for (uint32_t g = 0;; ++g) {
        PPF_HIP(hipMemcpyAsync(pl->h_scal, h.s, ns * sizeof(Scalars),
                               hipMemcpyDeviceToHost, st));
        PPF_HIP(hipStreamSynchronize(st));
        uint32_t live_rows = 0, live = 0;
        for (uint32_t c = 0; c < ns; ++c) {
            active[c] = pl->h_scal[c].done == 0;
            if (active[c]) {
                ++live;
                live_rows += pl->scene_row[c + 1] - pl->scene_row[c];
            }
        }
        if (live == 0) {
            break;
        }
        if (g >= max_launch) {
            set_error("pcg_solve: bookkeeping did not terminate");
            return PPF_ERR_SOLVER;
        }
        if (ns > 1 && 4ull * live_rows < 3ull * tiled_rows) {
            int rc = set_tiles(pl, pl->tile_r1, pl->tile_r2, st, &active);
            if (rc == PPF_OK) {
                PPF_HIP(hipMemcpyAsync(pl->d_args, &h, sizeof(Args), hipMemcpyHostToDevice, st));
                rc = capture_graph(pl);
            }
            if (rc != PPF_OK) {
                return rc;
            }
            tiled_rows = live_rows;
        }
        if (pl->exec) {
            PPF_HIP(hipGraphLaunch(pl->exec, st));
        } else { // a plan on a fixed region: the graph's iterations, launched one by one
            for (uint32_t k = 0; k < (uint32_t)kGraphIters; ++k) {
                PPF_HIP(launch_iteration(*pl, k, st));
            }
        }
    }
    if (pl->order) {
        k_order_rows<true><<<cdiv(3 * pl->n, kNT), kNT, 0, st>>>(pl->n, 3, pl->order, x,
                                                                 x_caller);
        PPF_HIP(hipGetLastError());
        PPF_HIP(hipStreamSynchronize(st));
    }
    for (uint32_t c = 0; c < ns; ++c) {
        const Scalars &sc = pl->h_scal[c];
        res[c].iter = sc.rep_iter;
        res[c].resid = sc.rep_resid;
        res[c].status = sc.status;
        res[c].break_cause = sc.cause;
        res[c].break_iter = sc.break_iter;
        res[c].err0 = sc.err0;
    }
    return PPF_OK;
}

// Tile sizes of a many-scene plan by measurement: each candidate pair (SpMV
// rows, vector-kernel rows per tile) runs the same solve (capped at 3 graphs
// of iterations, x in the plan's scratch), the fastest is kept.
int pcg_plan_autotune(PcgPlanImpl *pl, const float *val, const float *cdiag, const float *b,
                      float tol, uint32_t *best_r1, uint32_t *best_r2, float *best_ms,
                      hipStream_t st) {
    if (!pl || !val || !cdiag || !b) {
        set_error("pcg_plan_autotune: invalid argument");
        return PPF_ERR_ARG;
    }
    static const uint32_t k1[] = {128u, 256u, 512u, 1024u, 2048u};
    static const uint32_t k2[] = {256u, 512u, 1024u, 2048u};
    std::vector<ppf_pcg_result> res(pl->nscene);
    hipEvent_t e0, e1;
    PPF_HIP(hipEventCreate(&e0));
    PPF_HIP(hipEventCreate(&e1));
    float best = 3.4e38f;
    uint32_t b1 = 512u, b2 = 512u;
    int rc = PPF_OK;
    for (uint32_t r1 : k1) {
        for (uint32_t r2 : k2) {
            rc = set_tiles(pl, r1, r2, st);
            if (rc == PPF_OK) rc = capture_graph(pl);
            float ms = 0.0f;
            for (int rep = 0; rc == PPF_OK && rep < 2; ++rep) { // the first run warms up
                hipError_t e = hipMemsetAsync(pl->h.gy, 0, 12 * (size_t)pl->n, st);
                if (e == hipSuccess) e = hipEventRecord(e0, st);
                if (e != hipSuccess) {
                    set_error("pcg_plan_autotune: %s", hipGetErrorString(e));
                    rc = PPF_ERR_HIP;
                    break;
                }
                rc = pcg_solve(pl, val, cdiag, b, pl->h.gy, tol, 3 * kGraphIters, nullptr,
                               res.data(), st, nullptr);
                if (rc == PPF_OK) {
                    e = hipEventRecord(e1, st);
                    if (e == hipSuccess) e = hipEventSynchronize(e1);
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                    if (e != hipSuccess) {
                        set_error("pcg_plan_autotune: %s", hipGetErrorString(e));
                        rc = PPF_ERR_HIP;
                    }
                }
            }
            if (rc != PPF_OK) {
                break;
            }
            if (ms < best) {
                best = ms, b1 = r1, b2 = r2;
            }
        }
        if (rc != PPF_OK) {
            break;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == PPF_OK) rc = set_tiles(pl, b1, b2, st);
    if (rc == PPF_OK) rc = capture_graph(pl);
    if (rc == PPF_OK) {
        if (best_r1) *best_r1 = b1;
        if (best_r2) *best_r2 = b2;
        if (best_ms) *best_ms = best;
    }
    return rc;
}

int pcg_plan_set_order(PcgPlanImpl *pl, const uint32_t *order, hipStream_t st) {
    if (!pl || pl->nscene != 1) {
        set_error("pcg_plan_set_order: one-scene plans only");
        return PPF_ERR_ARG;
    }
    pl->order = order;
    pl->order_built = false;
    if (!order) {
        return PPF_OK;
    }
    const uint32_t n = pl->n;
    if (!pl->omem) {
        const size_t bytes = carve_bytes(n, 4) + 2 * carve_bytes(n + 1, 4) +
                             carve_bytes(9 * (size_t)n, 4) + 2 * carve_bytes(3 * (size_t)n, 4);
        PPF_HIP(dev_alloc(&pl->omem, bytes));
        char *cur = static_cast<char *>(pl->omem);
        pl->inv = carve<uint32_t>(cur, n);
        pl->prp = carve<uint32_t>(cur, n + 1);
        pl->cursor = carve<uint32_t>(cur, n + 1);
        pl->pcd = carve<float>(cur, 9 * (size_t)n);
        pl->pb = carve<float>(cur, 3 * (size_t)n);
        pl->px = carve<float>(cur, 3 * (size_t)n);
    }
    PPF_HIP(hipMemsetAsync(pl->inv, 0xff, (size_t)n * sizeof(uint32_t), st));
    PPF_HIP(hipMemsetAsync(pl->flags, 0, sizeof(uint32_t), st));
    k_inverse_order<<<cdiv(n, kNT), kNT, 0, st>>>(n, order, pl->inv, pl->flags);
    PPF_HIP(hipGetLastError());
    uint32_t bad = 0;
    PPF_HIP(hipMemcpyAsync(&bad, pl->flags, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    PPF_HIP(hipStreamSynchronize(st));
    if (bad) {
        pl->order = nullptr; // the plan keeps solving in the original order
        set_error("pcg_plan_set_order: order is not a permutation of 0..n-1 (%u entries out "
                  "of range or repeated)",
                  bad);
        return PPF_ERR_ARG;
    }
    return PPF_OK;
}

bool pcg_plan_ordered(const PcgPlanImpl *pl) { return pl->order != nullptr; }

int pcg_order_system(PcgPlanImpl *pl, const float *val, const float *cdiag, const float *b,
                     const float **oval, const float **ocdiag, const float **ob,
                     hipStream_t st) {
    if (!pl->order) {
        *oval = val, *ocdiag = cdiag, *ob = b;
        return PPF_OK;
    }
    if (!pl->order_built) {
        set_error("pcg: a new row order takes effect at the next pattern (repattern first)");
        return PPF_ERR_ARG;
    }
    const uint32_t n = pl->n;
    k_order_values<<<cdiv(pl->nnz, kNT), kNT, 0, st>>>(pl->nnz, pl->smap, val, pl->pval);
    k_order_rows<false><<<cdiv(9 * n, kNT), kNT, 0, st>>>(n, 9, pl->order, cdiag, pl->pcd);
    if (b) {
        k_order_rows<false><<<cdiv(3 * n, kNT), kNT, 0, st>>>(n, 3, pl->order, b, pl->pb);
    }
    PPF_HIP(hipGetLastError());
    *oval = pl->pval, *ocdiag = pl->pcd, *ob = b ? pl->pb : nullptr;
    return PPF_OK;
}

int pcg_plan_set_spmv_form(PcgPlanImpl *pl, int form, hipStream_t st) {
    if (!pl || form < PPF_SPMV_AUTO || form > PPF_SPMV_SYM) {
        set_error("pcg_plan_set_spmv_form: invalid argument");
        return PPF_ERR_ARG;
    }
    const int old = pl->form;
    pl->form = form;
    int rc = set_pattern(pl, pl->nnz, pl->user_rp, pl->user_col, st);
    if (rc != PPF_OK) {
        pl->form = old;
        (void)set_pattern(pl, pl->nnz, pl->user_rp, pl->user_col, st);
        return rc;
    }
    if (!pl->arena && spmv_key(*pl) != pl->graph_key) {
        rc = capture_graph(pl);
    }
    return rc;
}

int pcg_plan_spmv_form(const PcgPlanImpl *pl) {
    return pl->sym ? PPF_SPMV_SYM : pl->split ? PPF_SPMV_SPLIT : PPF_SPMV_ELL;
}

int pcg_plan_set_tiles(PcgPlanImpl *pl, uint32_t r1, uint32_t r2, hipStream_t st) {
    if (!pl || (pl->nscene > 1 && (r1 == 0 || r2 == 0)) || (r1 == 0) != (r2 == 0)) {
        set_error("pcg_plan_set_tiles: invalid argument");
        return PPF_ERR_ARG;
    }
    int rc = set_tiles(pl, r1, r2, st);
    return rc == PPF_OK ? capture_graph(pl) : rc;
}

} // namespace ppf

// ---- the allocation-free solve ---------------------------------------------

// Fixed device region of ppf_pcg_solve_noalloc: each GPU holds its own copy
// from module load until the process ends. Host code never passes this name
// to a kernel or a copy (on the host it is only a stand-in); it uses the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[ppf::kNoAllocScratchBytes];

namespace ppf {

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device with
// that device current.
int scratch_region(void **out) {
    int dev = 0;
    PPF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("pcg_solve_noalloc: device %d beyond the %d the library tracks", dev,
                  kMaxDevices);
        return PPF_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("pcg_solve_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return PPF_ERR_HIP;
        }
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return PPF_OK;
}

// Upper bound of what plan_init carves for one scene of n rows and nnz
// stored blocks: n-sized scratch, transpose table, sort keys, the widest
// ELL table (32 entries a row) and tiles for the largest possible grid.
size_t noalloc_bytes(uint32_t n, uint32_t nnz) {
    constexpr size_t kMaxTiles = 8192; // 256 CUs x 32 workgroups
    const size_t n3 = 3 * (size_t)n;
    const size_t mem = carve_bytes(n + 1, 4) + carve_bytes(n, 4) +
                       carve_bytes(9 * (size_t)n, 4) + carve_bytes(n3, 4) * 5 +
                       carve_bytes(1, sizeof(Args)) + carve_bytes(scan_scratch_words(n + 1), 4) +
                       carve_bytes(4, 4) + carve_bytes(2, 4);
    const size_t pat = carve_bytes(2 * ((size_t)nnz + nnz / 4 + 256), 4);
    const size_t sort_words = 3 * (size_t)nnz + radix_sort_scratch_words(nnz);
    const size_t tsort = carve_bytes(sort_words + sort_words / 4 + 256, 4);
    const size_t ell = carve_bytes((size_t)n * 32, sizeof(uint2));
    const size_t tiles = carve_bytes(1, sizeof(Scalars)) + 2 * carve_bytes(2, 4) +
                         2 * carve_bytes(kMaxTiles, sizeof(Tile)) +
                         carve_bytes(4 * kMaxTiles, 4);
    return mem + pat + tsort + ell + tiles;
}

} // namespace

int pcg_solve_noalloc(uint32_t n, uint32_t nnz, const uint32_t *rp, const uint32_t *col,
                      const float *val, const float *cdiag, const float *b, float *x,
                      float tol, uint32_t max_iter, ppf_pcg_result *res, hipStream_t st) {
    if (n == 0 || !rp || !col || !val || !cdiag || !b || !x || !res || max_iter == 0 ||
        nnz >= kMirror) {
        set_error("ppf_pcg_solve_noalloc: invalid argument");
        return PPF_ERR_ARG;
    }
    const size_t need = noalloc_bytes(n, nnz);
    if (need > kNoAllocScratchBytes) {
        set_error("ppf_pcg_solve_noalloc: %u rows with %u stored blocks need %zu bytes of "
                  "scratch, beyond the fixed %zu",
                  n, nnz, need, (size_t)kNoAllocScratchBytes);
        return PPF_ERR_ARG;
    }
    void *region = nullptr;
    int rc = scratch_region(&region);
    if (rc != PPF_OK) {
        return rc;
    }
    Scalars status{};
    PcgPlanImpl pl;
    pl.arena = static_cast<char *>(region);
    pl.arena_cap = kNoAllocScratchBytes;
    pl.h_scal = &status;
    rc = plan_init(&pl, n, nnz, rp, col, 1u, nullptr, st);
    if (rc == PPF_OK) {
        rc = pcg_solve(&pl, val, cdiag, b, x, tol, max_iter, nullptr, res, st, nullptr);
    }
    // The solve ends on a synchronous status read, so nothing queued still
    // reads the region or `status`; error paths drain the stream before the
    // next call may reuse them.
    if (rc != PPF_OK) {
        (void)hipStreamSynchronize(st);
    }
    return rc;
}

} // namespace ppf
