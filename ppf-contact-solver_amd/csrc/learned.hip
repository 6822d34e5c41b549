// learned.hip -- the learned block-FSAI preconditioner on MI355X: graph and
// features of a Newton system, the message-passing network on bf16 MFMA
// (fp32 accumulate), the factor G, the Kaporin loss, backpropagation and
// Adam. The reference has no learned preconditioner; its block Jacobi
// (solver.cu:357-400) is what the network starts from (zero output heads)
// and what the solver falls back to.
#include <cmath>
#include <cstring>
#include <vector>

#include "learned.hpp"
#include "scan.hpp"

namespace ppf {

namespace {

constexpr unsigned kNT = 256;
constexpr uint32_t kH = kLpcH, kFN = kLpcFN, kFE = kLpcFE;
constexpr uint32_t kMir = 0x80000000u;
constexpr uint32_t kChunk = 4096; // rows per weight-gradient partial

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

uint32_t cdiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// bf16 with round to nearest even (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline bool pos_bf16(uint16_t h) { return !(h & 0x8000u) && (h & 0x7fffu); }

// ---- 3x3 helpers (column-major b[r + 3c]) ----------------------------------

__device__ inline void mm3(const float *a, const float *b, float *c) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            c[i + 3 * j] = a[i] * b[3 * j] + a[i + 3] * b[1 + 3 * j] + a[i + 6] * b[2 + 3 * j];
}
__device__ inline void mm3t(const float *a, const float *b, float *c) { // a b^T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            c[i + 3 * j] = a[i] * b[j] + a[i + 3] * b[j + 3] + a[i + 6] * b[j + 6];
}

// ---- graph ------------------------------------------------------------------

__global__ void k_edge_count(PcgPattern pt, uint32_t *deg) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    if (j >= pt.n) {
        return;
    }
    const uint32_t ns = pt.rp[j + 1] - pt.rp[j] - (pt.dslot[j] != 0xffffffffu ? 1u : 0u);
    deg[j] = ns + (pt.tp[j + 1] - pt.tp[j]);
}

// Edges into row j: its stored off-diagonal slots (source = column), then its
// mirrored entries (source = the storing row). srow: row of every slot;
// sstore / semir: the store-side and mirror-side edge of every slot.
__global__ void k_edge_fill(PcgPattern pt, const uint32_t *eoff, uint32_t *esrc, uint32_t *eslot,
                            uint32_t *srow, uint32_t *sstore, uint32_t *semir) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    if (j >= pt.n) {
        return;
    }
    uint32_t e = eoff[j];
    for (uint32_t s = pt.rp[j]; s < pt.rp[j + 1]; ++s) {
        srow[s] = j;
        if (pt.col[s] != j) {
            esrc[e] = pt.col[s];
            eslot[e] = s;
            sstore[s] = e;
            ++e;
        } else {
            sstore[s] = 0xffffffffu;
            semir[s] = 0xffffffffu;
        }
    }
    for (uint32_t k = pt.tp[j]; k < pt.tp[j + 1]; ++k) {
        esrc[e] = pt.tcol[k];
        eslot[e] = pt.tslot[k] | kMir;
        semir[pt.tslot[k]] = e;
        ++e;
    }
}

__global__ void k_edge_rev(uint32_t E, const uint32_t *eslot, const uint32_t *sstore,
                           const uint32_t *semir, uint32_t *erev) {
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    if (e >= E) {
        return;
    }
    const uint32_t s = eslot[e] & ~kMir;
    erev[e] = (eslot[e] & kMir) ? sstore[s] : semir[s];
}

// Cholesky of every diagonal block A_jj (stored diagonal + cdiag) and its
// inverse L^-1 (lower, column-major). A non-positive pivot sets bad.
__global__ void k_diag_chol(PcgPattern pt, const float *val, const float *cdiag, float *linv,
                            uint32_t *bad) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    if (j >= pt.n) {
        return;
    }
    float a[9];
    for (int k = 0; k < 9; ++k) {
        a[k] = cdiag[9 * (size_t)j + k];
    }
    if (pt.dslot[j] != 0xffffffffu) {
        for (int k = 0; k < 9; ++k) {
            a[k] += val[9 * (size_t)pt.dslot[j] + k];
        }
    }
    // symmetric part (the stored blocks are symmetric up to rounding)
    const float a00 = a[0], a10 = 0.5f * (a[1] + a[3]), a20 = 0.5f * (a[2] + a[6]);
    const float a11 = a[4], a21 = 0.5f * (a[5] + a[7]), a22 = a[8];
    float l00 = a00 > 0.0f ? sqrtf(a00) : 0.0f;
    const float l10 = l00 > 0.0f ? a10 / l00 : 0.0f, l20 = l00 > 0.0f ? a20 / l00 : 0.0f;
    const float d1 = a11 - l10 * l10;
    const float l11 = d1 > 0.0f ? sqrtf(d1) : 0.0f;
    const float l21 = l11 > 0.0f ? (a21 - l20 * l10) / l11 : 0.0f;
    const float d2 = a22 - l20 * l20 - l21 * l21;
    const float l22 = d2 > 0.0f ? sqrtf(d2) : 0.0f;
    if (!(l00 > 0.0f && l11 > 0.0f && l22 > 0.0f)) {
        atomicAdd(bad, 1u);
        l00 = 1.0f;
    }
    // inverse of the lower factor
    const float i00 = 1.0f / l00, i11 = l11 > 0.0f ? 1.0f / l11 : 1.0f,
                i22 = l22 > 0.0f ? 1.0f / l22 : 1.0f;
    const float i10 = -l10 * i00 * i11;
    const float i21 = -l21 * i11 * i22;
    const float i20 = -(l20 * i00 + l21 * i10) * i22;
    float *o = linv + 9 * (size_t)j;
    o[0] = i00, o[1] = i10, o[2] = i20;
    o[3] = 0.0f, o[4] = i11, o[5] = i21;
    o[6] = 0.0f, o[7] = 0.0f, o[8] = i22;
}

// Edge features: Ahat_e = Linv_dst A_e Linv_src^T row-major (9), its
// Frobenius norm, 1 for a lower entry (source < destination).
__global__ void k_edge_features(PcgPattern pt, uint32_t E, const uint32_t *eoff,
                                const uint32_t *esrc, const uint32_t *eslot, const float *val,
                                const float *linv, uint16_t *f, float *fnorm) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    if (j >= pt.n) {
        return;
    }
    for (uint32_t e = eoff[j]; e < eoff[j + 1]; ++e) {
        const uint32_t i = esrc[e], s = eslot[e] & ~kMir;
        float blk[9], t[9], ah[9];
        const float *v = val + 9 * (size_t)s;
        if (eslot[e] & kMir) { // the stored block (i, j) applied transposed
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    blk[r + 3 * c] = v[c + 3 * r];
        } else {
            for (int k = 0; k < 9; ++k) blk[k] = v[k];
        }
        mm3(linv + 9 * (size_t)j, blk, t);
        mm3t(t, linv + 9 * (size_t)i, ah);
        uint16_t *fe = f + (size_t)e * kFE;
        float nrm = 0.0f;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                const float a = ah[r + 3 * c];
                fe[3 * r + c] = to_bf16(a);
                nrm += a * a;
            }
        nrm = sqrtf(nrm);
        fe[9] = to_bf16(nrm);
        fe[10] = to_bf16(i < j ? 1.0f : 0.0f);
        for (uint32_t k = 11; k < kFE; ++k) fe[k] = 0;
        fnorm[e] = nrm;
    }
}

// Sums over many workgroups: each workgroup writes its sum to parts[blockIdx.x]
// and one workgroup adds those in a fixed order, so the total (and what
// depends on it: features, factor, loss, weights) is the same on every run.
__global__ void k_sum_parts(const float *parts, uint32_t nb, float *out) {
    float v = 0.0f;
    for (uint32_t i = threadIdx.x; i < nb; i += kNT) {
        v += parts[i];
    }
    __shared__ float lds[kNT / kWave];
    v = block_sum<kNT>(v, lds);
    if (threadIdx.x == 0) {
        *out = v;
    }
}

// per-workgroup sums of |Linv_j b_j|^2 (for the RMS of the normalised
// right-hand side)
__global__ void k_bhat_norm(uint32_t n, const float *linv, const float *b, float *parts) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    float v = 0.0f;
    if (j < n) {
        const f3 y = mv(linv + 9 * (size_t)j, load3(b + 3 * (size_t)j));
        v = y.x * y.x + y.y * y.y + y.z * y.z;
    }
    __shared__ float lds[kNT / kWave];
    v = block_sum<kNT>(v, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = v;
    }
}

__global__ void k_node_features(uint32_t n, const uint32_t *eoff, const float *fnorm,
                                const float *linv, const float *b, const float *acc,
                                uint16_t *x) {
    const uint32_t j = blockIdx.x * kNT + threadIdx.x;
    if (j >= n) {
        return;
    }
    float sum = 0.0f, mx = 0.0f;
    for (uint32_t e = eoff[j]; e < eoff[j + 1]; ++e) {
        sum += fnorm[e];
        mx = fmaxf(mx, fnorm[e]);
    }
    const float rms = sqrtf(acc[0] / (float)n);
    const float inv = rms > 0.0f ? 1.0f / rms : 1.0f;
    const f3 y = mv(linv + 9 * (size_t)j, load3(b + 3 * (size_t)j));
    uint16_t *xj = x + (size_t)j * kFN;
    xj[0] = to_bf16(sum);
    xj[1] = to_bf16(mx);
    xj[2] = to_bf16((float)(eoff[j + 1] - eoff[j]) / 16.0f);
    xj[3] = to_bf16(y.x * inv);
    xj[4] = to_bf16(y.y * inv);
    xj[5] = to_bf16(y.z * inv);
    for (uint32_t k = 6; k < kFN; ++k) xj[k] = 0;
}

// ---- dense layers on MFMA ---------------------------------------------------

// One input segment of a dense layer: rows of x (leading dimension ld,
// width k, a multiple of 16), gathered through idx when set; entries whose
// mask (same layout as x) is not positive read as zero.
struct Seg {
    const uint16_t *x = nullptr;
    const uint32_t *idx = nullptr;
    const uint16_t *mask = nullptr;
    uint32_t k = 0, ld = 0;
};

__device__ inline v4s load_frag(const Seg &s, uint32_t row, uint32_t col) {
    const size_t rr = s.idx ? s.idx[row] : row;
    const uint2 w = *reinterpret_cast<const uint2 *>(s.x + rr * s.ld + col);
    uint16_t h[4] = {(uint16_t)(w.x & 0xffffu), (uint16_t)(w.x >> 16), (uint16_t)(w.y & 0xffffu),
                     (uint16_t)(w.y >> 16)};
    if (s.mask) {
        const uint2 mw = *reinterpret_cast<const uint2 *>(s.mask + rr * s.ld + col);
        const uint16_t m[4] = {(uint16_t)(mw.x & 0xffffu), (uint16_t)(mw.x >> 16),
                               (uint16_t)(mw.y & 0xffffu), (uint16_t)(mw.y >> 16)};
        for (int q = 0; q < 4; ++q) {
            if (!pos_bf16(m[q])) h[q] = 0;
        }
    }
    v4s a;
    a[0] = (short)h[0], a[1] = (short)h[1], a[2] = (short)h[2], a[3] = (short)h[3];
    return a;
}

// Y[m][0:16 NT] = act(sum over segments X_seg[m] B_seg + bias) on
// v_mfma_f32_16x16x16_bf16, K = 16 KS (the segments' widths summed). Each
// wave loads its lanes' weight fragments once and keeps them in registers,
// then walks 16-row tiles (grid-stride), so the weights are not re-read per
// tile. Bt is the weight transposed, [16 NT][K] bf16 (row = output column).
// A fragment: lane l holds X[l % 16][4 (l / 16) + 0..3]; B fragment: lane l
// holds B[4 (l / 16) + 0..3][l % 16]; D: lane l holds Y[4 (l / 16) + 0..3][l % 16].
// The k order (segment 0, then segment 1, 16 at a time) is fixed.
template <int NT, int KS>
__global__ __launch_bounds__(256) void k_dense(uint32_t M, Seg s0, Seg s1, const uint16_t *Bt,
                                               const float *bias, int relu, uint16_t *yb,
                                               float *yf) {
    constexpr uint32_t K = 16 * KS;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t ar = lane & 15u, aq = lane >> 4;
    v4s bw[KS][NT];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const uint2 w = *reinterpret_cast<const uint2 *>(Bt + (size_t)(16 * t + ar) * K +
                                                             16 * kk + 4 * aq);
            bw[kk][t][0] = (short)(w.x & 0xffffu), bw[kk][t][1] = (short)(w.x >> 16);
            bw[kk][t][2] = (short)(w.y & 0xffffu), bw[kk][t][3] = (short)(w.y >> 16);
        }
    float bc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        bc[t] = bias ? bias[16 * t + ar] : 0.0f;
    }
    const uint32_t ntile = (M + 15u) / 16u;
    const uint32_t nwave = gridDim.x * 4u;
    for (uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6); tile < ntile; tile += nwave) {
        const uint32_t m0 = tile * 16u;
        const bool arow_ok = m0 + ar < M;
        v4f acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const uint32_t kg = 16u * kk;
            const bool first = kg < s0.k;
            v4s a = v4s{0, 0, 0, 0};
            if (arow_ok) {
                a = first ? load_frag(s0, m0 + ar, kg + 4 * aq)
                          : load_frag(s1, m0 + ar, kg - s0.k + 4 * aq);
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw[kk][t], acc[t], 0, 0, 0);
            }
        }
        constexpr uint32_t ld = 16 * NT;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = m0 + 4 * aq + r, col = 16 * t + ar;
                if (row < M) {
                    float v = acc[t][r] + bc[t];
                    if (relu) {
                        v = fmaxf(v, 0.0f);
                    }
                    if (yb) yb[(size_t)row * ld + col] = to_bf16(v);
                    if (yf) yf[(size_t)row * ld + col] = v;
                }
            }
        }
    }
}

// Workgroups of k_dense for M rows: enough 4-wave groups to cover the
// tiles, at most kDenseGrid (a few rounds of resident waves on 256 CUs), so
// each wave's register-held weights serve many tiles.
constexpr uint32_t kDenseGrid = 2048;
inline uint32_t dense_grid(uint32_t M) { return std::max(1u, std::min(kDenseGrid, (M + 63u) / 64u)); }

// Weight-gradient partials: part[chunk][k][c] = sum over the chunk's rows m
// of X[m][k] D[m][c] (X = the layer's input segments, D = [M][64] bf16), on
// MFMA with the rows as the reduction dimension; cpart[chunk][c] = column
// sums of D (the bias gradient).
__global__ __launch_bounds__(256) void k_wgrad(uint32_t M, Seg s0, Seg s1, const uint16_t *D,
                                               float *part, float *cpart) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t c0 = blockIdx.x * kChunk, c1 = min(M, c0 + kChunk);
    const uint32_t ktot = s0.k + s1.k, nkt = ktot / 16;
    const uint32_t ar = lane & 15u, aq = lane >> 4;
    float *pc = part + (size_t)blockIdx.x * ktot * kH;
    for (uint32_t tile = wave; tile < nkt * 4; tile += 4) {
        const uint32_t kt = tile / 4, nt = tile % 4;
        const uint32_t k = kt * 16 + ar;
        const Seg &s = k < s0.k ? s0 : s1;
        const uint32_t kc = k < s0.k ? k : k - s0.k;
        v4f acc = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (uint32_t m0 = c0; m0 < c1; m0 += 16) {
            v4s a, b;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t m = m0 + 4 * aq + r;
                uint16_t xa = 0, db = 0;
                if (m < c1) {
                    const size_t rr = s.idx ? s.idx[m] : m;
                    xa = s.x[rr * s.ld + kc];
                    db = D[(size_t)m * kH + nt * 16 + ar];
                }
                a[r] = (short)xa;
                b[r] = (short)db;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            pc[(size_t)(kt * 16 + 4 * aq + r) * kH + nt * 16 + ar] = acc[r];
        }
    }
    if (threadIdx.x < kH) {
        float sum = 0.0f;
        for (uint32_t m = c0; m < c1; ++m) {
            sum += from_bf16(D[(size_t)m * kH + threadIdx.x]);
        }
        cpart[(size_t)blockIdx.x * kH + threadIdx.x] = sum;
    }
}

// gW[k][c] += sum over chunks of part (fixed order); gb[c] likewise.
__global__ void k_wgrad_reduce(uint32_t nchunk, uint32_t ktot, const float *part,
                               const float *cpart, float *gW, float *gb) {
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    if (e < ktot * kH) {
        float s = 0.0f;
        for (uint32_t c = 0; c < nchunk; ++c) {
            s += part[(size_t)c * ktot * kH + e];
        }
        gW[e] += s;
    } else if (e < ktot * kH + kH) {
        const uint32_t c = e - ktot * kH;
        float s = 0.0f;
        for (uint32_t k = 0; k < nchunk; ++k) {
            s += cpart[(size_t)k * kH + c];
        }
        gb[c] += s;
    }
}

// Mean of the messages into each row (fp32 sum in edge order).
__global__ void k_aggregate(uint32_t n, const uint32_t *eoff, const uint16_t *m, uint16_t *M) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    const uint32_t j = t / kH, c = t % kH;
    if (j >= n) {
        return;
    }
    float s = 0.0f;
    for (uint32_t e = eoff[j]; e < eoff[j + 1]; ++e) {
        s += from_bf16(m[(size_t)e * kH + c]);
    }
    const uint32_t deg = eoff[j + 1] - eoff[j];
    M[(size_t)j * kH + c] = to_bf16(s / (float)(deg ? deg : 1u));
}

// ---- heads and the factor -----------------------------------------------

// t_j = h3_j Wo + co; u_s = [h3_col(s), h3_row(s), f_semir(s)] Wg + cg for
// every stored off-diagonal slot s (its lower edge), 0 on diagonal slots.
__global__ void k_heads(PcgPattern pt, const uint16_t *h3, const uint16_t *f, const uint32_t *srow,
                        const uint32_t *semir, const float *w, float *t, float *u) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k < pt.n) {
        const uint16_t *h = h3 + (size_t)k * kH;
        float acc[6];
        for (int o = 0; o < 6; ++o) acc[o] = w[LpcLayout::co + o];
        for (uint32_t c = 0; c < kH; ++c) {
            const float hv = from_bf16(h[c]);
            for (int o = 0; o < 6; ++o) acc[o] += hv * w[LpcLayout::Wo + c * 6 + o];
        }
        for (int o = 0; o < 6; ++o) t[6 * (size_t)k + o] = acc[o];
    }
    const uint32_t s = k;
    if (s < pt.nnz) {
        float acc[9];
        const uint32_t i = srow[s], j = pt.col[s];
        if (i == j) {
            for (int o = 0; o < 9; ++o) u[9 * (size_t)s + o] = 0.0f;
            return;
        }
        for (int o = 0; o < 9; ++o) acc[o] = w[LpcLayout::cg + o];
        const uint16_t *hd = h3 + (size_t)j * kH, *hs = h3 + (size_t)i * kH;
        const uint16_t *fe = f + (size_t)semir[s] * kFE;
        for (uint32_t c = 0; c < 2 * kH + kFE; ++c) {
            const float xv = from_bf16(c < kH ? hd[c] : (c < 2 * kH ? hs[c - kH] : fe[c - 2 * kH]));
            for (int o = 0; o < 9; ++o) acc[o] += xv * w[LpcLayout::Wg + c * 9 + o];
        }
        for (int o = 0; o < 9; ++o) u[9 * (size_t)s + o] = acc[o];
    }
}

// Ghat_jj = I + lower(0.5 tanh t), Ghat_s = 0.5 tanh(u) (column-major), and
// G = Ghat L^-1: Gd_j = Ghat_jj Linv_j, Gl_s = Ghat_s Linv_row(s).
__constant__ int kLowR[6] = {0, 1, 1, 2, 2, 2};
__constant__ int kLowC[6] = {0, 0, 1, 0, 1, 2};

__device__ inline void pack_g16(const float *g, uint16_t *o) {
    for (int e = 0; e < 9; ++e) {
        o[e] = to_bf16(g[e]);
    }
    o[9] = 0;
}

__global__ void k_factor(PcgPattern pt, const float *t, const float *u, const float *linv,
                         const uint32_t *srow, float *ghd, float *ghl, float *Gd, float *Gl,
                         uint16_t *Gd16, uint16_t *Gl16) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k < pt.n) {
        float g[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int o = 0; o < 6; ++o) {
            g[kLowR[o] + 3 * kLowC[o]] += 0.5f * tanhf(t[6 * (size_t)k + o]);
        }
        for (int o = 0; o < 9; ++o) ghd[9 * (size_t)k + o] = g[o];
        mm3(g, linv + 9 * (size_t)k, Gd + 9 * (size_t)k);
        pack_g16(Gd + 9 * (size_t)k, Gd16 + 10 * (size_t)k);
    }
    if (k < pt.nnz) {
        float g[9];
        for (int o = 0; o < 9; ++o) g[o] = 0.5f * tanhf(u[9 * (size_t)k + o]);
        for (int o = 0; o < 9; ++o) ghl[9 * (size_t)k + o] = g[o];
        mm3(g, linv + 9 * (size_t)srow[k], Gl + 9 * (size_t)k);
        pack_g16(Gl + 9 * (size_t)k, Gl16 + 10 * (size_t)k);
    }
}

// Rows of G in order for the PCG's G r pass: row j at j + tp[j] holds G_jj,
// then G_ji for its mirrored entries in transpose-table order (4 lanes/row).
__global__ void k_row_records(PcgPattern pt, const uint16_t *Gd16, const uint16_t *Gl16,
                              uint16_t *Gr16) {
    const uint32_t j = blockIdx.x * (kNT / 4) + (threadIdx.x >> 2);
    const uint32_t q = threadIdx.x & 3u;
    if (j >= pt.n) {
        return;
    }
    const uint32_t t0 = pt.tp[j], cnt = 1 + pt.tp[j + 1] - t0;
    const size_t g0 = (size_t)j + t0;
    for (uint32_t k = q; k < cnt; k += 4) {
        const uint16_t *src = k == 0 ? Gd16 + 10 * (size_t)j
                                     : Gl16 + 10 * (size_t)pt.tslot[t0 + k - 1];
        const uint32_t *w = reinterpret_cast<const uint32_t *>(src);
        uint32_t *o = reinterpret_cast<uint32_t *>(Gr16 + 10 * (g0 + k));
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            o[e] = w[e];
        }
    }
}

// ---- loss and its gradient with respect to Ghat -------------------------

// y = Ghat^T z for every probe: y_a = Ghat_aa^T z_a + sum over row a's stored
// off-diagonal slots s = (a, j) of Ghat_s^T z_j; then v = Linv^T y.
__global__ void k_probe_y(PcgPattern pt, uint32_t K, const float *ghd, const float *ghl,
                          const float *linv, const float *z, float *y, float *v) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    const uint32_t a = t % pt.n, k = t / pt.n;
    if (k >= K) {
        return;
    }
    const float *zk = z + (size_t)k * 3 * pt.n;
    f3 ya = mvt(ghd + 9 * (size_t)a, load3(zk + 3 * (size_t)a));
    for (uint32_t s = pt.rp[a]; s < pt.rp[a + 1]; ++s) {
        const uint32_t j = pt.col[s];
        if (j != a) {
            ya = ya + mvt(ghl + 9 * (size_t)s, load3(zk + 3 * (size_t)j));
        }
    }
    store3(y + (size_t)k * 3 * pt.n + 3 * (size_t)a, ya);
    store3(v + (size_t)k * 3 * pt.n + 3 * (size_t)a, mvt(linv + 9 * (size_t)a, ya));
}

// w = 2 Linv (A v) in place over Av, and the partial sums of y . (w / 2).
__global__ void k_probe_w(uint32_t n, uint32_t K, const float *linv, const float *y, float *av,
                          float *parts) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    float q = 0.0f;
    if (t < n * K) {
        const uint32_t a = t % n, k = t / n;
        const size_t o = (size_t)k * 3 * n + 3 * (size_t)a;
        const f3 w = mv(linv + 9 * (size_t)a, load3(av + o));
        const f3 ya = load3(y + o);
        q = ya.x * w.x + ya.y * w.y + ya.z * w.z;
        store3(av + o, mk3(2.0f * w.x, 2.0f * w.y, 2.0f * w.z));
    }
    __shared__ float lds[kNT / kWave];
    q = block_sum<kNT>(q, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = q;
    }
}

// dL/dGhat and the log-determinant part of the loss. acc[0] = K q.
__global__ void k_dghat(PcgPattern pt, uint32_t K, const float *z, const float *w,
                        const float *ghd, const float *acc, const uint32_t *srow, float *dgd,
                        float *dgl, float *parts) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    const float inv = 1.0f / acc[0]; // 1 / (K q)
    const float dcoef = 2.0f / (3.0f * (float)pt.n);
    float ld = 0.0f;
    if (k < pt.n) {
        float g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t p = 0; p < K; ++p) {
            const f3 zj = load3(z + (size_t)p * 3 * pt.n + 3 * (size_t)k);
            const f3 wj = load3(w + (size_t)p * 3 * pt.n + 3 * (size_t)k);
            const float za[3] = {zj.x, zj.y, zj.z}, wb[3] = {wj.x, wj.y, wj.z};
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) g[r + 3 * c] += za[r] * wb[c];
        }
        for (int o = 0; o < 9; ++o) g[o] *= inv;
        for (int d = 0; d < 3; ++d) {
            const float gdd = ghd[9 * (size_t)k + 4 * d];
            g[4 * d] -= dcoef / gdd;
            ld += logf(gdd);
        }
        for (int o = 0; o < 9; ++o) dgd[9 * (size_t)k + o] = g[o];
    }
    if (k < pt.nnz) {
        const uint32_t i = srow[k], j = pt.col[k];
        float g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (i != j) { // Ghat_s = Ghat_ji: z on row j, w on row i
            for (uint32_t p = 0; p < K; ++p) {
                const f3 zj = load3(z + (size_t)p * 3 * pt.n + 3 * (size_t)j);
                const f3 wi = load3(w + (size_t)p * 3 * pt.n + 3 * (size_t)i);
                const float za[3] = {zj.x, zj.y, zj.z}, wb[3] = {wi.x, wi.y, wi.z};
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) g[r + 3 * c] += za[r] * wb[c];
            }
            for (int o = 0; o < 9; ++o) g[o] *= inv;
        }
        for (int o = 0; o < 9; ++o) dgl[9 * (size_t)k + o] = g[o];
    }
    __shared__ float lds[kNT / kWave];
    ld = block_sum<kNT>(ld, lds);
    if (threadIdx.x == 0) {
        parts[blockIdx.x] = ld;
    }
}

// Through the tanh of the heads: dt (n x 6) over dgd's lower entries, du
// (nnz x 9) over dgl; written over dgd / dgl's first entries.
__global__ void k_dhead_out(PcgPattern pt, const float *t, const float *u, float *dgd,
                            float *dgl) {
    const uint32_t k = blockIdx.x * kNT + threadIdx.x;
    if (k < pt.n) {
        float d[6];
        for (int o = 0; o < 6; ++o) {
            const float th = tanhf(t[6 * (size_t)k + o]);
            d[o] = dgd[9 * (size_t)k + kLowR[o] + 3 * kLowC[o]] * 0.5f * (1.0f - th * th);
        }
        for (int o = 0; o < 6; ++o) dgd[9 * (size_t)k + o] = d[o];
    }
    if (k < pt.nnz) {
        for (int o = 0; o < 9; ++o) {
            const float th = tanhf(u[9 * (size_t)k + o]);
            dgl[9 * (size_t)k + o] *= 0.5f * (1.0f - th * th);
        }
    }
}

// Head weight gradients: gWo[c][o] = sum_j h3[j][c] dt[j][o] (fp32, per
// chunk then reduced by k_wgrad_reduce-like sums in fixed order).
__global__ void k_head_wgrad(PcgPattern pt, const uint16_t *h3, const uint16_t *f,
                             const uint32_t *srow, const uint32_t *semir, const float *dt9,
                             const float *du, float *part) {
    // part layout per chunk: [64*6 Wo][6 co][144*9 Wg][9 cg]
    constexpr uint32_t kP = kH * 6 + 6 + (2 * kH + kFE) * 9 + 9;
    const uint32_t e = threadIdx.x + blockIdx.y * kNT;
    if (e >= kP) {
        return;
    }
    const uint32_t c0 = blockIdx.x * kChunk;
    float s = 0.0f;
    if (e < kH * 6 + 6) {
        const uint32_t c = e / 6, o = e % 6;
        const uint32_t c1 = min(pt.n, c0 + kChunk);
        for (uint32_t j = c0; j < c1; ++j) {
            const float x = c < kH ? from_bf16(h3[(size_t)j * kH + c]) : 1.0f;
            s += x * dt9[9 * (size_t)j + o];
        }
    } else {
        const uint32_t q = e - (kH * 6 + 6);
        const uint32_t c = q / 9, o = q % 9;
        const uint32_t c1 = min(pt.nnz, c0 + kChunk);
        for (uint32_t sl = c0; sl < c1; ++sl) {
            const float d = du[9 * (size_t)sl + o];
            if (d == 0.0f) {
                continue;
            }
            float x = 1.0f;
            if (c < kH) {
                x = from_bf16(h3[(size_t)pt.col[sl] * kH + c]);
            } else if (c < 2 * kH) {
                x = from_bf16(h3[(size_t)srow[sl] * kH + c - kH]);
            } else if (c < 2 * kH + kFE) {
                x = from_bf16(f[(size_t)semir[sl] * kFE + c - 2 * kH]);
            }
            s += x * d;
        }
    }
    part[(size_t)blockIdx.x * kP + e] = s;
}

__global__ void k_head_wgrad_reduce(uint32_t nchunk, const float *part, float *grad) {
    constexpr uint32_t kP = kH * 6 + 6 + (2 * kH + kFE) * 9 + 9;
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    if (e >= kP) {
        return;
    }
    float s = 0.0f;
    for (uint32_t c = 0; c < nchunk; ++c) {
        s += part[(size_t)c * kP + e];
    }
    grad[LpcLayout::Wo + e] += s; // Wo, co, Wg, cg are contiguous in this order
}

// dh3[j][c] = sum_o dt[j][o] Wo[c][o] + the lower-edge head's input
// gradients: as destination (mirrored entries of row j), as source (row j's
// stored off-diagonal slots).
__global__ void k_dh3(PcgPattern pt, const float *dt9, const float *du, const float *w,
                      float *dh) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    const uint32_t j = t / kH, c = t % kH;
    if (j >= pt.n) {
        return;
    }
    float s = 0.0f;
    for (int o = 0; o < 6; ++o) s += dt9[9 * (size_t)j + o] * w[LpcLayout::Wo + c * 6 + o];
    for (uint32_t k = pt.tp[j]; k < pt.tp[j + 1]; ++k) {
        const float *d = du + 9 * (size_t)pt.tslot[k];
        for (int o = 0; o < 9; ++o) s += d[o] * w[LpcLayout::Wg + c * 9 + o];
    }
    for (uint32_t sl = pt.rp[j]; sl < pt.rp[j + 1]; ++sl) {
        if (pt.col[sl] == j) {
            continue;
        }
        const float *d = du + 9 * (size_t)sl;
        for (int o = 0; o < 9; ++o) s += d[o] * w[LpcLayout::Wg + (kH + c) * 9 + o];
    }
    dh[(size_t)j * kH + c] = s;
}

// dZ = dh * (h > 0) as bf16.
__global__ void k_relu_grad(uint32_t count, const float *dh, const uint16_t *h, uint16_t *dz) {
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    if (e < count) {
        dz[e] = pos_bf16(h[e]) ? to_bf16(dh[e]) : (uint16_t)0;
    }
}

// Message gradients: dY_e = dM_dst / deg_dst where the message was positive.
__global__ void k_msg_grad(uint32_t n, const uint32_t *eoff, const float *dXn, const uint16_t *m,
                           uint16_t *dY) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    const uint32_t j = t / kH, c = t % kH;
    if (j >= n) {
        return;
    }
    const uint32_t deg = eoff[j + 1] - eoff[j];
    const float g = dXn[(size_t)j * 2 * kH + kH + c] / (float)(deg ? deg : 1u);
    for (uint32_t e = eoff[j]; e < eoff[j + 1]; ++e) {
        dY[(size_t)e * kH + c] = pos_bf16(m[(size_t)e * kH + c]) ? to_bf16(g) : (uint16_t)0;
    }
}

// dh_prev[j] = dXn[j][0:64] + sum over row j's edges e of dXe[rev(e)].
__global__ void k_scatter_src(uint32_t n, const uint32_t *eoff, const uint32_t *erev,
                              const float *dXn, const float *dXe, float *dh) {
    const uint32_t t = blockIdx.x * kNT + threadIdx.x;
    const uint32_t j = t / kH, c = t % kH;
    if (j >= n) {
        return;
    }
    float s = dXn[(size_t)j * 2 * kH + c];
    for (uint32_t e = eoff[j]; e < eoff[j + 1]; ++e) {
        s += dXe[(size_t)erev[e] * kH + c];
    }
    dh[(size_t)j * kH + c] = s;
}

__global__ void k_adam(uint32_t count, float *w, const float *g, float *m, float *v, float lr,
                       float c1, float c2) {
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    if (e >= count) {
        return;
    }
    const float b1 = 0.9f, b2 = 0.999f;
    m[e] = b1 * m[e] + (1.0f - b1) * g[e];
    v[e] = b2 * v[e] + (1.0f - b2) * g[e] * g[e];
    w[e] -= lr * (m[e] / c1) / (sqrtf(v[e] / c2) + 1e-8f);
}

// bf16 operands of the dense layers: wt (transposed [out][in]) for the
// forward pass, wr (row-major [in][out]) for the input gradients.
__global__ void k_weights_bf16(const float *w, uint16_t *wt, uint16_t *wr) {
    const uint32_t e = blockIdx.x * kNT + threadIdx.x;
    // dense layers: W0 (16 in), then per layer We (80 in), Wn (128 in)
    uint32_t base = 0, off = 0, in = 0;
    uint32_t rem = e;
    const uint32_t ins[1 + 2 * kLpcLayers] = {kFN, kH + kFE, 2 * kH, kH + kFE, 2 * kH, kH + kFE,
                                              2 * kH};
    const uint32_t offs[1 + 2 * kLpcLayers] = {LpcLayout::W0, LpcLayout::We(0), LpcLayout::Wn(0),
                                               LpcLayout::We(1), LpcLayout::Wn(1),
                                               LpcLayout::We(2), LpcLayout::Wn(2)};
    int layer = -1;
    for (int l = 0; l < 1 + 2 * (int)kLpcLayers; ++l) {
        const uint32_t sz = ins[l] * kH;
        if (rem < sz) {
            layer = l;
            in = ins[l];
            off = offs[l];
            break;
        }
        rem -= sz;
        base += sz;
    }
    if (layer < 0) {
        return;
    }
    const uint32_t k = rem / kH, c = rem % kH;
    const uint16_t h = to_bf16(w[off + rem]);
    wr[base + rem] = h;
    wt[base + (size_t)c * in + k] = h;
}

constexpr uint32_t kDenseTotal = (kLpcFN + kLpcLayers * (3 * kLpcH + kLpcFE)) * kLpcH;

uint32_t dense_base(int l) { // offset of dense layer l in wt / wr (0: W0, 2k+1: We_k, 2k+2: Wn_k)
    const uint32_t ins[1 + 2 * kLpcLayers] = {kFN, kH + kFE, 2 * kH, kH + kFE, 2 * kH, kH + kFE,
                                              2 * kH};
    uint32_t b = 0;
    for (int k = 0; k < l; ++k) b += ins[k] * kH;
    return b;
}

template <class T> T *carve(char *&cur, size_t count) {
    T *p = reinterpret_cast<T *>(cur);
    cur += (count * sizeof(T) + 255) & ~size_t(255);
    return p;
}

} // namespace

Lpc::~Lpc() { release(); }

void Lpc::release() {
    if (mem_) (void)dev_free(mem_);
    for (void *p : {(void *)w_, (void *)grad_, (void *)am_, (void *)av_, (void *)wt_, (void *)wr_,
                    (void *)pv_, (void *)pav_, (void *)py_}) {
        if (p) (void)dev_free(p);
    }
    mem_ = nullptr;
    w_ = grad_ = am_ = av_ = pv_ = pav_ = py_ = nullptr;
    wt_ = wr_ = nullptr;
}

int Lpc::refresh_bf16(hipStream_t s) {
    k_weights_bf16<<<cdiv(kDenseTotal, kNT), kNT, 0, s>>>(w_, wt_, wr_);
    PPF_HIP(hipGetLastError());
    return PPF_OK;
}

int Lpc::set_weights(const float *host, uint32_t count, hipStream_t s) {
    if (count != LpcLayout::count) {
        set_error("lpc: %u weights given, %u expected", count, LpcLayout::count);
        return PPF_ERR_ARG;
    }
    if (!w_) {
        const size_t b = LpcLayout::count * sizeof(float);
        PPF_HIP(dev_alloc(&w_, b));
        PPF_HIP(dev_alloc(&grad_, b));
        PPF_HIP(dev_alloc(&am_, b));
        PPF_HIP(dev_alloc(&av_, b));
        PPF_HIP(dev_alloc(&wt_, kDenseTotal * sizeof(uint16_t)));
        PPF_HIP(dev_alloc(&wr_, kDenseTotal * sizeof(uint16_t)));
    }
    PPF_HIP(hipMemcpyAsync(w_, host, LpcLayout::count * sizeof(float), hipMemcpyHostToDevice, s));
    PPF_HIP(hipMemsetAsync(am_, 0, LpcLayout::count * sizeof(float), s));
    PPF_HIP(hipMemsetAsync(av_, 0, LpcLayout::count * sizeof(float), s));
    weights_set_ = true;
    return refresh_bf16(s);
}

int Lpc::get_weights(float *host, uint32_t count, hipStream_t s) {
    if (!weights_set_ || count != LpcLayout::count) {
        set_error("lpc: no weights, or %u requested of %u", count, LpcLayout::count);
        return PPF_ERR_ARG;
    }
    PPF_HIP(hipMemcpyAsync(host, w_, count * sizeof(float), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    return PPF_OK;
}

int Lpc::reserve(uint32_t n, uint32_t nnz, uint32_t E) {
    if (n <= ncap_ && nnz <= nzcap_ && E <= ecap_ && mem_) {
        return PPF_OK;
    }
    if (mem_) {
        (void)dev_free(mem_);
        mem_ = nullptr;
    }
    ncap_ = std::max(n, ncap_), nzcap_ = std::max(nnz, nzcap_), ecap_ = std::max(E, ecap_);
    const size_t N = ncap_, Z = nzcap_, EE = std::max<uint32_t>(ecap_, 1);
    const size_t rows = std::max(N, EE);
    const size_t nchunk = (rows + kChunk - 1) / kChunk;
    const size_t part_words =
        std::max(nchunk * (2 * kH + kFE + 16) * kH + nchunk * kH,
                 nchunk * (kH * 6 + 6 + (2 * kH + kFE) * 9 + 9));
    auto b = [](size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); };
    size_t bytes = b(N + 1, 4) + b(EE, 4) * 3 + b(N, 4) + b(Z, 4) * 3 + b(9 * N, 4) +
                   b(kFN * N, 2) + b(kFE * EE, 2) + b(EE, 4) + b(kH * N, 2) * (kLpcLayers + 1) +
                   b(kH * EE, 2) * kLpcLayers + b(kH * N, 2) * kLpcLayers + b(6 * N, 4) +
                   b(9 * Z, 4) * 3 + b(9 * N, 4) * 3 + b(kH * N, 4) + b(9 * Z, 4) + b(16, 4) +
                   b(kH * rows, 2) * 2 + b(2 * kH * N, 4) + b(kH * EE, 4) + b(part_words, 4) +
                   b(scan_scratch_words(N + 1), 4) + b(10 * N, 2) + b(10 * Z, 2) +
                   b(10 * ((size_t)N + Z), 2);
    PPF_HIP(dev_alloc(&mem_, bytes));
    mem_bytes_ = bytes;
    char *cur = static_cast<char *>(mem_);
    eoff_ = carve<uint32_t>(cur, N + 1);
    esrc_ = carve<uint32_t>(cur, EE);
    eslot_ = carve<uint32_t>(cur, EE);
    erev_ = carve<uint32_t>(cur, EE);
    deg_ = carve<uint32_t>(cur, N);
    srow_ = carve<uint32_t>(cur, Z);
    semir_ = carve<uint32_t>(cur, Z);
    sstore_ = carve<uint32_t>(cur, Z);
    linv_ = carve<float>(cur, 9 * N);
    x_ = carve<uint16_t>(cur, kFN * N);
    f_ = carve<uint16_t>(cur, kFE * EE);
    fnorm_ = carve<float>(cur, EE);
    for (uint32_t l = 0; l <= kLpcLayers; ++l) h_[l] = carve<uint16_t>(cur, kH * N);
    for (uint32_t l = 0; l < kLpcLayers; ++l) m_[l] = carve<uint16_t>(cur, kH * EE);
    for (uint32_t l = 0; l < kLpcLayers; ++l) M_[l] = carve<uint16_t>(cur, kH * N);
    t_ = carve<float>(cur, 6 * N);
    u_ = carve<float>(cur, 9 * Z);
    ghl_ = carve<float>(cur, 9 * Z);
    Gl_ = carve<float>(cur, 9 * Z);
    ghd_ = carve<float>(cur, 9 * N);
    Gd_ = carve<float>(cur, 9 * N);
    Gd16_ = carve<uint16_t>(cur, 10 * N);
    Gl16_ = carve<uint16_t>(cur, 10 * Z);
    Gr16_ = carve<uint16_t>(cur, 10 * ((size_t)N + Z));
    dgd_ = carve<float>(cur, 9 * N);
    dh_ = carve<float>(cur, kH * N);
    dgl_ = carve<float>(cur, 9 * Z);
    scal_ = carve<float>(cur, 16);
    dZ_ = carve<uint16_t>(cur, kH * rows);
    dY_ = carve<uint16_t>(cur, kH * rows);
    dXn_ = reinterpret_cast<uint16_t *>(carve<float>(cur, 2 * kH * N));
    dXe_ = reinterpret_cast<uint16_t *>(carve<float>(cur, kH * EE));
    part_ = carve<float>(cur, part_words);
    part_words_ = part_words;
    scan_ = carve<uint32_t>(cur, scan_scratch_words(N + 1));
    if ((size_t)(cur - static_cast<char *>(mem_)) > bytes) {
        set_error("lpc: workspace layout exceeds its allocation");
        return PPF_ERR_ARG;
    }
    return PPF_OK;
}

int Lpc::prepare(const PcgPattern &pt, const float *val, const float *cdiag, const float *b,
                 hipStream_t s) {
    if (!weights_set_) {
        set_error("lpc: weights are not set");
        return PPF_ERR_ARG;
    }
    // edges: 2 per stored off-diagonal slot (counted on the device)
    int rc = reserve(pt.n, pt.nnz, 2 * pt.nnz);
    if (rc != PPF_OK) {
        return rc;
    }
    pt_ = pt;
    n_ = pt.n, nnz_ = pt.nnz;
    factored_ = false;
    k_edge_count<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt, deg_);
    PPF_HIP(hipGetLastError());
    // exclusive scan of deg into eoff (n + 1 entries: deg_[n] is scratch 0)
    PPF_HIP(hipMemcpyAsync(eoff_, deg_, pt.n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    PPF_HIP(hipMemsetAsync(eoff_ + pt.n, 0, sizeof(uint32_t), s));
    PPF_HIP(exclusive_scan_u32(eoff_, eoff_, pt.n + 1, scan_, s));
    uint32_t E = 0;
    PPF_HIP(hipMemcpyAsync(&E, eoff_ + pt.n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    E_ = E;
    k_edge_fill<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt, eoff_, esrc_, eslot_, srow_, sstore_,
                                                 semir_);
    PPF_HIP(hipGetLastError());
    if (E) {
        k_edge_rev<<<cdiv(E, kNT), kNT, 0, s>>>(E, eslot_, sstore_, semir_, erev_);
        PPF_HIP(hipGetLastError());
    }
    PPF_HIP(hipMemsetAsync(scal_, 0, 16 * sizeof(float), s));
    k_diag_chol<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt, val, cdiag, linv_,
                                                 reinterpret_cast<uint32_t *>(scal_ + 8));
    PPF_HIP(hipGetLastError());
    k_edge_features<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt, E, eoff_, esrc_, eslot_, val, linv_, f_,
                                                     fnorm_);
    PPF_HIP(hipGetLastError());
    k_bhat_norm<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt.n, linv_, b, part_);
    PPF_HIP(hipGetLastError());
    k_sum_parts<<<1, kNT, 0, s>>>(part_, cdiv(pt.n, kNT), scal_);
    PPF_HIP(hipGetLastError());
    k_node_features<<<cdiv(pt.n, kNT), kNT, 0, s>>>(pt.n, eoff_, fnorm_, linv_, b, scal_, x_);
    PPF_HIP(hipGetLastError());
    uint32_t bad = 0;
    PPF_HIP(hipMemcpyAsync(&bad, scal_ + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    if (bad) {
        set_error("lpc: %u diagonal blocks are not positive definite", bad);
        return PPF_ERR_SOLVER;
    }
    return PPF_OK;
}

int Lpc::forward(hipStream_t s) {
    if (!mem_ || !n_) {
        set_error("lpc: forward before prepare");
        return PPF_ERR_ARG;
    }
    const uint32_t n = n_, E = E_;
    Seg a, none;
    a.x = x_, a.k = kFN, a.ld = kFN;
    k_dense<4, kFN / 16><<<dense_grid(n), 256, 0, s>>>(n, a, none, wt_ + dense_base(0),
                                                       w_ + LpcLayout::c0, 1, h_[0], nullptr);
    PPF_HIP(hipGetLastError());
    for (uint32_t l = 0; l < kLpcLayers; ++l) {
        if (E) {
            Seg hs, fe;
            hs.x = h_[l], hs.idx = esrc_, hs.k = kH, hs.ld = kH;
            fe.x = f_, fe.k = kFE, fe.ld = kFE;
            k_dense<4, (kH + kFE) / 16><<<dense_grid(E), 256, 0, s>>>(
                E, hs, fe, wt_ + dense_base(1 + 2 * l), w_ + LpcLayout::ce(l), 1, m_[l], nullptr);
            PPF_HIP(hipGetLastError());
        }
        k_aggregate<<<cdiv(n * kH, kNT), kNT, 0, s>>>(n, eoff_, m_[l], M_[l]);
        PPF_HIP(hipGetLastError());
        Seg hn, mn;
        hn.x = h_[l], hn.k = kH, hn.ld = kH;
        mn.x = M_[l], mn.k = kH, mn.ld = kH;
        k_dense<4, 2 * kH / 16><<<dense_grid(n), 256, 0, s>>>(
            n, hn, mn, wt_ + dense_base(2 + 2 * l), w_ + LpcLayout::cn(l), 1, h_[l + 1], nullptr);
        PPF_HIP(hipGetLastError());
    }
    const uint32_t m = std::max(n, nnz_);
    k_heads<<<cdiv(m, kNT), kNT, 0, s>>>(pt_, h_[kLpcLayers], f_, srow_, semir_, w_, t_, u_);
    PPF_HIP(hipGetLastError());
    k_factor<<<cdiv(m, kNT), kNT, 0, s>>>(pt_, t_, u_, linv_, srow_, ghd_, ghl_, Gd_, Gl_,
                                          Gd16_, Gl16_);
    PPF_HIP(hipGetLastError());
    k_row_records<<<cdiv(n, kNT / 4), kNT, 0, s>>>(pt_, Gd16_, Gl16_, Gr16_);
    PPF_HIP(hipGetLastError());
    factored_ = true;
    return PPF_OK;
}

int Lpc::gradient(PcgPlanImpl *plan, const float *val, const float *cdiag, const float *probes,
                  uint32_t K, float *loss, hipStream_t s) {
    const uint32_t n = n_, E = E_;
    if (K == 0 || !probes) {
        set_error("lpc: no probes");
        return PPF_ERR_ARG;
    }
    if (K > nprobe_cap_ || !pv_) {
        for (float *p : {pv_, pav_, py_}) {
            if (p) (void)dev_free(p);
        }
        PPF_HIP(dev_alloc(&pv_, (size_t)K * 3 * n * sizeof(float)));
        PPF_HIP(dev_alloc(&pav_, (size_t)K * 3 * n * sizeof(float)));
        PPF_HIP(dev_alloc(&py_, (size_t)K * 3 * n * sizeof(float)));
        nprobe_cap_ = K;
    }
    PPF_HIP(hipMemsetAsync(grad_, 0, LpcLayout::count * sizeof(float), s));
    PPF_HIP(hipMemsetAsync(scal_, 0, 4 * sizeof(float), s));
    k_probe_y<<<cdiv(K * n, kNT), kNT, 0, s>>>(pt_, K, ghd_, ghl_, linv_, probes, py_, pv_);
    PPF_HIP(hipGetLastError());
    for (uint32_t k = 0; k < K; ++k) {
        int rc = pcg_spmv(plan, val, cdiag, pv_ + (size_t)k * 3 * n, pav_ + (size_t)k * 3 * n, s);
        if (rc != PPF_OK) {
            return rc;
        }
    }
    const uint32_t m = std::max(n, nnz_);
    if (cdiv(K * n, kNT) > part_words_ || cdiv(m, kNT) > part_words_) {
        set_error("lpc: %u probes exceed the partial-sum scratch", K);
        return PPF_ERR_ARG;
    }
    k_probe_w<<<cdiv(K * n, kNT), kNT, 0, s>>>(n, K, linv_, py_, pav_, part_);
    PPF_HIP(hipGetLastError());
    k_sum_parts<<<1, kNT, 0, s>>>(part_, cdiv(K * n, kNT), scal_);
    PPF_HIP(hipGetLastError());
    k_dghat<<<cdiv(m, kNT), kNT, 0, s>>>(pt_, K, probes, pav_, ghd_, scal_, srow_, dgd_, dgl_,
                                         part_);
    PPF_HIP(hipGetLastError());
    k_sum_parts<<<1, kNT, 0, s>>>(part_, cdiv(m, kNT), scal_ + 1);
    PPF_HIP(hipGetLastError());
    k_dhead_out<<<cdiv(m, kNT), kNT, 0, s>>>(pt_, t_, u_, dgd_, dgl_);
    PPF_HIP(hipGetLastError());
    // head weights and dh3
    {
        constexpr uint32_t kP = kH * 6 + 6 + (2 * kH + kFE) * 9 + 9;
        const uint32_t nchunk = cdiv(m, kChunk);
        dim3 grid(nchunk, cdiv(kP, kNT));
        k_head_wgrad<<<grid, kNT, 0, s>>>(pt_, h_[kLpcLayers], f_, srow_, semir_, dgd_, dgl_,
                                         part_);
        PPF_HIP(hipGetLastError());
        k_head_wgrad_reduce<<<cdiv(kP, kNT), kNT, 0, s>>>(nchunk, part_, grad_);
        PPF_HIP(hipGetLastError());
        k_dh3<<<cdiv(n * kH, kNT), kNT, 0, s>>>(pt_, dgd_, dgl_, w_, dh_);
        PPF_HIP(hipGetLastError());
    }
    float *dXn = reinterpret_cast<float *>(dXn_);
    float *dXe = reinterpret_cast<float *>(dXe_);
    float *cpart = part_ + (size_t)cdiv(std::max(n, E), kChunk) * (2 * kH + kFE + 16) * kH;
    for (int l = (int)kLpcLayers - 1; l >= 0; --l) {
        // node layer: h_{l+1} = relu([h_l, M_l] Wn + cn)
        k_relu_grad<<<cdiv(n * kH, kNT), kNT, 0, s>>>(n * kH, dh_, h_[l + 1], dZ_);
        PPF_HIP(hipGetLastError());
        Seg hn, mn, none;
        hn.x = h_[l], hn.k = kH, hn.ld = kH;
        mn.x = M_[l], mn.k = kH, mn.ld = kH;
        uint32_t nch = cdiv(n, kChunk);
        k_wgrad<<<nch, 256, 0, s>>>(n, hn, mn, dZ_, part_, cpart);
        PPF_HIP(hipGetLastError());
        k_wgrad_reduce<<<cdiv(2 * kH * kH + kH, kNT), kNT, 0, s>>>(
            nch, 2 * kH, part_, cpart, grad_ + LpcLayout::Wn(l), grad_ + LpcLayout::cn(l));
        PPF_HIP(hipGetLastError());
        Seg dz;
        dz.x = dZ_, dz.k = kH, dz.ld = kH;
        // dX = dZ Wn^T: Wn row-major [128][64] is the transposed operand of
        // this product (row = output column)
        k_dense<8, kH / 16><<<dense_grid(n), 256, 0, s>>>(n, dz, none, wr_ + dense_base(2 + 2 * l),
                                                          nullptr, 0, nullptr, dXn);
        PPF_HIP(hipGetLastError());
        // edge layer: m = relu([h_src, f] We + ce)
        if (E) {
            k_msg_grad<<<cdiv(n * kH, kNT), kNT, 0, s>>>(n, eoff_, dXn, m_[l], dY_);
            PPF_HIP(hipGetLastError());
            Seg hs, fe;
            hs.x = h_[l], hs.idx = esrc_, hs.k = kH, hs.ld = kH;
            fe.x = f_, fe.k = kFE, fe.ld = kFE;
            nch = cdiv(E, kChunk);
            k_wgrad<<<nch, 256, 0, s>>>(E, hs, fe, dY_, part_, cpart);
            PPF_HIP(hipGetLastError());
            k_wgrad_reduce<<<cdiv((kH + kFE) * kH + kH, kNT), kNT, 0, s>>>(
                nch, kH + kFE, part_, cpart, grad_ + LpcLayout::We(l), grad_ + LpcLayout::ce(l));
            PPF_HIP(hipGetLastError());
            Seg dy;
            dy.x = dY_, dy.k = kH, dy.ld = kH;
            // input gradient of h_src only: the first 64 rows of We
            k_dense<4, kH / 16><<<dense_grid(E), 256, 0, s>>>(
                E, dy, none, wr_ + dense_base(1 + 2 * l), nullptr, 0, nullptr, dXe);
            PPF_HIP(hipGetLastError());
        }
        k_scatter_src<<<cdiv(n * kH, kNT), kNT, 0, s>>>(n, eoff_, erev_, dXn, E ? dXe : dXn,
                                                        dh_);
        PPF_HIP(hipGetLastError());
        if (!E) {
            // no edges: dh_prev is dXn[:, :64] alone (k_scatter_src read no dXe)
        }
    }
    k_relu_grad<<<cdiv(n * kH, kNT), kNT, 0, s>>>(n * kH, dh_, h_[0], dZ_);
    PPF_HIP(hipGetLastError());
    {
        Seg xs, none;
        xs.x = x_, xs.k = kFN, xs.ld = kFN;
        const uint32_t nch = cdiv(n, kChunk);
        k_wgrad<<<nch, 256, 0, s>>>(n, xs, none, dZ_, part_, cpart);
        PPF_HIP(hipGetLastError());
        k_wgrad_reduce<<<cdiv(kFN * kH + kH, kNT), kNT, 0, s>>>(nch, kFN, part_, cpart,
                                                               grad_ + LpcLayout::W0,
                                                               grad_ + LpcLayout::c0);
        PPF_HIP(hipGetLastError());
    }
    float h[2];
    PPF_HIP(hipMemcpyAsync(h, scal_, 2 * sizeof(float), hipMemcpyDeviceToHost, s));
    PPF_HIP(hipStreamSynchronize(s));
    const float q = h[0] / (float)K;
    *loss = logf(q / (3.0f * (float)n)) - 2.0f / (3.0f * (float)n) * h[1];
    return PPF_OK;
}

int Lpc::adam(float lr, uint32_t step, hipStream_t s) {
    const float c1 = 1.0f - powf(0.9f, (float)step), c2 = 1.0f - powf(0.999f, (float)step);
    k_adam<<<cdiv(LpcLayout::count, kNT), kNT, 0, s>>>(LpcLayout::count, w_, grad_, am_, av_, lr,
                                                      c1, c2);
    PPF_HIP(hipGetLastError());
    return refresh_bf16(s);
}

} // namespace ppf
