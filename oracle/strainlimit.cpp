// strainlimit.cpp -- CPU oracle: the strain-limiting barrier and its line
// search.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path, never shipped).
//
// Restates strainlimiting::embed_strainlimiting_force_hessian and
// strainlimiting::line_search (strainlimiting/strainlimiting.cu:17-100,
// 243-305) for shell faces, with the one-rotation Jacobi SVD of 3x2 matrices
// (utility/utility.cu:43-148), the cubic barrier (barrier/cubic.hpp:13-40,
// barrier/barrier.cu:116-131) and the 3x2 eigen-analysis of the projected
// Hessian (eigenanalysis/eigenanalysis.cu:18-78). fp32 throughout, every
// expression in the order the device code evaluates it. Pinned in
// tests/test_oracle_strainlimit.py by numpy's SVD, fp64 finite differences of
// the barrier energy (force and, where the energy is convex, Hessian), a hand
// computed stiffness and the closed-form limit time of a uniform stretch.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Svd32 {
    float U[6]; // 3x2, column-major
    float S[2]; // ascending
    float V[4]; // 2x2, column-major
};

Svd32 svd32(const float F[6]) {
    float A[3][2] = {{F[0], F[3]}, {F[1], F[4]}, {F[2], F[5]}};
    float W[2][2] = {{1.0f, 0.0f}, {0.0f, 1.0f}};
    float a = 0.0f, b = 0.0f, g = 0.0f;
    for (int r = 0; r < 3; ++r) {
        a += A[r][0] * A[r][0];
        b += A[r][1] * A[r][1];
        g += A[r][0] * A[r][1];
    }
    const float zeta = (b - a) / (2.0f * g);
    const float t = (g == 0.0f) ? 0.0f
                                : (zeta >= 0.0f ? 1.0f : -1.0f) /
                                      (std::fabs(zeta) + std::sqrt(1.0f + zeta * zeta));
    const float c = 1.0f / std::sqrt(1.0f + t * t), sn = c * t;
    for (int r = 0; r < 3; ++r) {
        const float ap = A[r][0], aq = A[r][1];
        A[r][0] = c * ap - sn * aq;
        A[r][1] = sn * ap + c * aq;
    }
    for (int r = 0; r < 2; ++r) {
        const float vp = W[r][0], vq = W[r][1];
        W[r][0] = c * vp - sn * vq;
        W[r][1] = sn * vp + c * vq;
    }
    float sv[2], u[3][2];
    for (int k = 0; k < 2; ++k) {
        float n2 = 0.0f;
        for (int r = 0; r < 3; ++r) n2 += A[r][k] * A[r][k];
        const float nrm = std::sqrt(n2);
        sv[k] = nrm;
        const float inv = nrm > 0.0f ? 1.0f / nrm : 0.0f;
        for (int r = 0; r < 3; ++r) u[r][k] = A[r][k] * inv;
    }
    const int i0 = sv[1] < sv[0] ? 1 : 0, i1 = 1 - i0;
    Svd32 out;
    out.S[0] = sv[i0], out.S[1] = sv[i1];
    for (int r = 0; r < 3; ++r) {
        out.U[r] = u[r][i0];
        out.U[3 + r] = u[r][i1];
    }
    for (int r = 0; r < 2; ++r) {
        out.V[r] = W[r][i0];
        out.V[2 + r] = W[r][i1];
    }
    return out;
}

void face_F(const float *x, const uint32_t v[3], const float *inv, float F[6]) {
    float x0[3], x1[3], x2[3], d0[3], d1[3];
    for (int d = 0; d < 3; ++d) {
        x0[d] = x[3 * (size_t)v[0] + d];
        x1[d] = x[3 * (size_t)v[1] + d];
        x2[d] = x[3 * (size_t)v[2] + d];
        d0[d] = x1[d] - x0[d];
        d1[d] = x2[d] - x0[d];
    }
    const float i0 = inv[0], i1 = inv[1], i2 = inv[2], i3 = inv[3];
    for (int d = 0; d < 3; ++d) {
        F[d] = (0.0f + d0[d] * i0) + d1[d] * i1;
        F[3 + d] = (0.0f + d0[d] * i2) + d1[d] * i3;
    }
}

float limit_of(const float *sl, const float *shrink, uint32_t k) {
    const float s = shrink ? shrink[k] : 1.0f;
    return (sl[k] + 1.0f) / s - 1.0f;
}

// Block (i, j) of the upper-stored pattern, transposed when i > j.
void pattern_block(const uint32_t *rp, const uint32_t *ci, const float *val, uint32_t i,
                   uint32_t j, float blk[9]) {
    const uint32_t r = std::min(i, j), c = std::max(i, j);
    std::fill(blk, blk + 9, 0.0f);
    for (uint32_t k = rp[r]; k < rp[r + 1]; ++k) {
        if (ci[k] == c) {
            const float *b = val + 9 * (size_t)k;
            for (int rr = 0; rr < 3; ++rr)
                for (int cc = 0; cc < 3; ++cc)
                    blk[rr + 3 * cc] = i <= j ? b[rr + 3 * cc] : b[cc + 3 * rr];
            return;
        }
    }
}

void cross3(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

// M = U3 Q V^T for a 3x2 mode Q (column-major).
void mode(const float U3[9], const float Q[6], const float V[4], float M[6]) {
    float T[6];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) {
            float acc = 0.0f;
            for (int k = 0; k < 3; ++k) acc = acc + U3[r + 3 * k] * Q[k + 3 * c];
            T[r + 3 * c] = acc;
        }
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r) {
            float acc = 0.0f;
            for (int j = 0; j < 2; ++j) acc = acc + T[r + 3 * j] * V[c + 2 * j];
            M[r + 3 * c] = acc;
        }
}

const uint8_t kFP[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};

// The strain-limit terms of face k: forces of its three vertices (9) and its
// six upper blocks (oriented for the lower vertex index row), already scaled
// by the stiffness. False when the face takes no terms.
bool face_terms(uint32_t k, const uint32_t *face, const float *inv_rest, const uint32_t *fflag,
                const float *fmass, const float *sl_arr, const float *shrink, const float *x,
                const float *x0, const uint32_t *rp, const uint32_t *ci, const float *el,
                float eig_eps, float fo[9], float bo[54], float *stiff_out, float S_out[2]) {
    const float sl = sl_arr[k];
    if ((fflag[k] & 1u) || !(sl > 0.0f)) return false;
    const uint32_t v[3] = {face[3 * (size_t)k], face[3 * (size_t)k + 1], face[3 * (size_t)k + 2]};
    const float *inv = inv_rest + 4 * (size_t)k;
    float F[6];
    face_F(x, v, inv, F);
    const Svd32 svd = svd32(F);
    float gs[2] = {svd.S[0] - 1.0f, svd.S[1] - 1.0f};
    if (S_out) {
        S_out[0] = svd.S[0], S_out[1] = svd.S[1];
    }
    if (!(std::max(gs[0], gs[1]) > 0.0f)) return false;
    float deda[2] = {0.0f, 0.0f}, h[2] = {0.0f, 0.0f};
    for (int i = 0; i < 2; ++i) {
        if (gs[i] > 0.0f) {
            const float y = (sl - gs[i]) - 0.0f;
            const float yy = y - sl;
            deda[i] = -(yy < 0.0f ? -2.0f * yy * yy / sl : 0.0f);
            h[i] = y - sl < 0.0f ? 4.0f * (1.0f - y / sl) : 0.0f;
        }
    }
    float stiff = 0.0f;
    {
        float y[3][3];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) y[a][d] = x0[3 * (size_t)v[a] + d];
        const float s3 = 1.0f / 3.0f;
        float cen[3], wv[3][3];
        for (int d = 0; d < 3; ++d) cen[d] = (s3 * y[0][d] + s3 * y[1][d]) + s3 * y[2][d];
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) wv[a][d] = y[a][d] - cen[d];
        for (int ii = 0; ii < 3; ++ii) {
            float hv[3] = {0.0f, 0.0f, 0.0f};
            for (int jj = 0; jj < 3; ++jj) {
                float blk[9];
                pattern_block(rp, ci, el, v[ii], v[jj], blk);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) hv[r] = hv[r] + blk[r + 3 * c] * wv[jj][c];
            }
            stiff = stiff + ((wv[ii][0] * hv[0] + wv[ii][1] * hv[1]) + wv[ii][2] * hv[2]);
        }
        const float d = limit_of(sl_arr, shrink, k) - std::max(gs[0], gs[1]);
        stiff = stiff + fmass[k] / (d * d);
    }
    if (stiff_out) *stiff_out = stiff;
    const float a0 = gs[0] + 1.0f, a1 = gs[1] + 1.0f;
    float dF[6];
    for (int c = 0; c < 2; ++c)
        for (int r = 0; r < 3; ++r)
            dF[r + 3 * c] = (0.0f + (svd.U[r] * deda[0]) * svd.V[c]) +
                            (svd.U[3 + r] * deda[1]) * svd.V[c + 2];
    float U3[9];
    {
        const float u0[3] = {svd.U[0], svd.U[1], svd.U[2]};
        const float u1[3] = {svd.U[3], svd.U[4], svd.U[5]};
        float u2[3];
        cross3(u0, u1, u2);
        for (int d = 0; d < 3; ++d) {
            U3[d] = u0[d];
            U3[3 + d] = u1[d];
            U3[6 + d] = u2[d];
        }
    }
    const float is2 = 0.7071067811865475f;
    const float denom = a0 - a1;
    float lam[6];
    lam[0] = std::max(0.0f, (deda[0] + deda[1]) / (a0 + a1));
    lam[1] = std::max(0.0f, std::fabs(denom) > eig_eps
                                ? (deda[0] - deda[1]) / denom
                                : 0.5f * (h[0] + h[1]) - 0.5f * (0.0f + 0.0f));
    lam[2] = std::max(0.0f, deda[0] / a0);
    lam[3] = std::max(0.0f, deda[1] / a1);
    const int lo = h[1] < h[0] ? 1 : 0;
    lam[4] = std::max(0.0f, h[lo]);
    lam[5] = std::max(0.0f, h[1 - lo]);
    float H6[36] = {};
    for (int i = 0; i < 6; ++i) {
        if (lam[i] == 0.0f) continue;
        float Q[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (i == 0) {
            Q[3] = is2, Q[1] = -is2;
        } else if (i == 1) {
            Q[3] = is2, Q[1] = is2;
        } else if (i == 2) {
            Q[2] = 1.0f;
        } else if (i == 3) {
            Q[5] = 1.0f;
        } else {
            const int ax = i == 4 ? lo : 1 - lo;
            Q[ax + 3 * ax] = 1.0f;
        }
        float M[6];
        mode(U3, Q, svd.V, M);
        for (int n = 0; n < 6; ++n)
            for (int m = 0; m < 6; ++m) H6[m + 6 * n] = H6[m + 6 * n] + (lam[i] * M[m]) * M[n];
    }
    const float i0 = inv[0], i1 = inv[1], i2 = inv[2], i3 = inv[3];
    const float gg[3][2] = {{-i0 - i1, -i2 - i3}, {i0, i2}, {i1, i3}};
    for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) {
            const float fe = (0.0f + gg[a][0] * dF[d]) + gg[a][1] * dF[3 + d];
            fo[3 * a + d] = 0.0f + stiff * fe;
        }
    for (int p = 0; p < 6; ++p) {
        int a = kFP[p][0], b = kFP[p][1];
        if (v[a] > v[b]) std::swap(a, b);
        float blk[9] = {};
        for (int d = 0; d < 2; ++d)
            for (int e = 0; e < 2; ++e) {
                const float sc = gg[a][d] * gg[b][e];
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < 3; ++r)
                        blk[r + 3 * c] = blk[r + 3 * c] + sc * H6[(3 * d + r) + 6 * (3 * e + c)];
            }
        for (int e = 0; e < 9; ++e) bo[9 * p + e] = 0.0f + stiff * blk[e];
    }
    return true;
}

float max_stretch(const float F[6]) {
    const Svd32 s = svd32(F);
    return std::max(s.S[0] - 1.0f, s.S[1] - 1.0f);
}

} // namespace

extern "C" {

// One-rotation Jacobi SVD of n 3x2 matrices (column-major F, 6 floats each):
// U (6), S ascending (2), V (4) per matrix.
void oracle_svd32(uint32_t n, const float *F, float *U, float *S, float *V) {
    for (uint32_t i = 0; i < n; ++i) {
        const Svd32 s = svd32(F + 6 * (size_t)i);
        std::memcpy(U + 6 * (size_t)i, s.U, sizeof s.U);
        std::memcpy(S + 2 * (size_t)i, s.S, sizeof s.S);
        std::memcpy(V + 4 * (size_t)i, s.V, sizeof s.V);
    }
}

// Per-face strain-limit terms (for the tests): forces (9 per face), the six
// upper blocks (54 per face), the stiffness and the singular values; active
// flags per face.
void oracle_strainlimit_faces(uint32_t nf, const uint32_t *face, const float *inv_rest,
                              const uint32_t *fflag, const float *fmass, const float *sl,
                              const float *shrink, const float *x, const float *x0,
                              const uint32_t *rp, const uint32_t *ci, const float *el,
                              float eig_eps, float *fo, float *bo, float *stiff, float *S,
                              uint8_t *active) {
    for (uint32_t k = 0; k < nf; ++k) {
        float s[2] = {0.0f, 0.0f}, st = 0.0f;
        active[k] = face_terms(k, face, inv_rest, fflag, fmass, sl, shrink, x, x0, rp, ci, el,
                               eig_eps, fo + 9 * (size_t)k, bo + 54 * (size_t)k, &st, s)
                        ? 1
                        : 0;
        stiff[k] = st;
        S[2 * (size_t)k] = s[0], S[2 * (size_t)k + 1] = s[1];
    }
}

// Adds the strain-limit terms of every limited face to force (3 nv) and to
// the stored upper blocks (rp / ci / value): per vertex and per block in face
// order after what they hold, zero block entries skipped (the device gathers'
// order).
void oracle_strainlimit_assemble(uint32_t nv, uint32_t nf, const uint32_t *face,
                                 const float *inv_rest, const uint32_t *fflag,
                                 const float *fmass, const float *sl, const float *shrink,
                                 const float *x, const float *x0, const uint32_t *rp,
                                 const uint32_t *ci, float eig_eps, float *force, float *value) {
    if (!sl) return;
    const size_t nnz = rp[nv];
    std::vector<float> el(value, value + 9 * nnz); // the elastic blocks: stiffness reference
    std::vector<float> fo(9), bo(54);
    std::vector<std::vector<float>> add_f(nv), add_b(nnz);
    struct Term { uint32_t k; float f[9]; float b[54]; };
    std::vector<Term> terms;
    for (uint32_t k = 0; k < nf; ++k) {
        Term t;
        t.k = k;
        if (face_terms(k, face, inv_rest, fflag, fmass, sl, shrink, x, x0, rp, ci, el.data(),
                       eig_eps, t.f, t.b, nullptr, nullptr))
            terms.push_back(t);
    }
    // per vertex and per block, faces in index order
    for (const Term &t : terms) {
        const uint32_t v[3] = {face[3 * (size_t)t.k], face[3 * (size_t)t.k + 1],
                               face[3 * (size_t)t.k + 2]};
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 3; ++d) force[3 * (size_t)v[a] + d] += t.f[3 * a + d];
        for (int p = 0; p < 6; ++p) {
            uint32_t r = v[kFP[p][0]], c = v[kFP[p][1]];
            if (r > c) std::swap(r, c);
            for (uint32_t s = rp[r]; s < rp[r + 1]; ++s) {
                if (ci[s] != c) continue;
                for (int e = 0; e < 9; ++e) {
                    const float y = t.b[9 * p + e];
                    if (y != 0.0f) value[9 * (size_t)s + e] += y;
                }
                break;
            }
        }
    }
}

// Largest t <= max_t (unnormalized) that keeps every limited face's stretch
// below its limit over x0 + t (x1 - x0), by the reference's bisection.
float oracle_strainlimit_toi(uint32_t nf, const uint32_t *face, const float *inv_rest,
                             const uint32_t *fflag, const float *sl, const float *shrink,
                             const float *x0, const float *x1, float max_t) {
    float best = max_t;
    if (!sl) return best;
    for (uint32_t k = 0; k < nf; ++k) {
        if ((fflag[k] & 1u) || !(sl[k] > 0.0f)) continue;
        const uint32_t v[3] = {face[3 * (size_t)k], face[3 * (size_t)k + 1],
                               face[3 * (size_t)k + 2]};
        const float *inv = inv_rest + 4 * (size_t)k;
        const float limit = limit_of(sl, shrink, k);
        float F0[6], F1[6], dF[6], Ft[6];
        face_F(x0, v, inv, F0);
        face_F(x1, v, inv, F1);
        for (int e = 0; e < 6; ++e) dF[e] = F1[e] - F0[e];
        float t = max_t;
        for (int e = 0; e < 6; ++e) Ft[e] = F0[e] + t * dF[e];
        if (max_stretch(Ft) >= limit) {
            float upper = t, lower = 0.0f, window = upper - lower;
            while (true) {
                t = 0.5f * (upper + lower);
                for (int e = 0; e < 6; ++e) Ft[e] = F0[e] + t * dF[e];
                if (max_stretch(Ft) - limit < 0.0f) {
                    lower = t;
                } else {
                    upper = t;
                }
                const float nw = upper - lower;
                if (nw == window) break;
                window = nw;
            }
            t = lower;
        }
        best = std::min(best, t);
    }
    return best;
}

} // extern "C"
