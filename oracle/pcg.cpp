// pcg.cpp -- CPU oracle: block-Jacobi preconditioned CG over the reference's
// split operator M = B_fixed(upper, applied both ways) + C_diag.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP path and the CPU baseline;
// never shipped in the product).
//
// Restates in fp32:
//   * solver::apply (walk form)      crates/ppf-cts-solver/src/cpp/solver/solver.cu:213-259, 812-886
//   * solver::solve (inv_diag build) solver.cu:2180-2195
//   * solver::cg_device recurrence   solver.cu:921-1383
//       r = b - A x0, err0 = |r|_1                      1015-1031
//       z = P r, rz0, p = z                              1088-1093
//       SpMV + pAp + sum|terms|  (walk kernel)           812-886
//       curvature verdict, alpha, x/r/z update           431-514, 57-114
//       rz1, beta = rz1/rz0, roll, iteration counter     516-571
//       p = z + beta p                                   vec_ops.cu:59-67
//       residual check every 4 iterations, every
//       iteration once within 8*tol, return order        1249-1380
//
// FMA policy (the oracle's own, matched by the HIP kernels): the x, r and p
// vector updates are explicit fmaf (the reference's add_scaled_indirect is an
// explicit fmaf, solver.cu:55); every other expression rounds as written.
//
// Reductions are blocked: fixed 1024-element chunks summed left to right, then
// the chunk partials summed left to right. The result is independent of the
// OpenMP thread count, so the oracle is deterministic at any thread count.
#include "oracle_math.hpp"
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {
bool bj_invert(const float *in, float *out);
}

namespace {

const size_t kChunk = 1024;

template <class F> float blocked_sum(size_t n, F term) {
    size_t nc = (n + kChunk - 1) / kChunk;
    std::vector<float> part(nc, 0.0f);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)nc; ++c) {
        size_t lo = (size_t)c * kChunk, hi = lo + kChunk < n ? lo + kChunk : n;
        float acc = 0.0f;
        for (size_t k = lo; k < hi; ++k) {
            acc = acc + term(k);
        }
        part[c] = acc;
    }
    float acc = 0.0f;
    for (size_t c = 0; c < nc; ++c) {
        acc = acc + part[c];
    }
    return acc;
}

struct Operator {
    uint32_t n;
    const uint32_t *rp, *col;
    const float *val, *cd;
    std::vector<uint32_t> tp, tcol, tslot; // transpose table (builder.rs:1336-1345)

    void build_transpose() {
        tp.assign(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) {
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                if (col[k] != i) {
                    tp[col[k] + 1]++;
                }
            }
        }
        for (uint32_t i = 0; i < n; ++i) {
            tp[i + 1] += tp[i];
        }
        tcol.resize(tp[n]);
        tslot.resize(tp[n]);
        std::vector<uint32_t> fill(tp.begin(), tp.end() - 1);
        for (uint32_t i = 0; i < n; ++i) { // ascending row order per column
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                uint32_t j = col[k];
                if (j != i) {
                    tcol[fill[j]] = i;
                    tslot[fill[j]] = k;
                    fill[j]++;
                }
            }
        }
    }

    // y = M p; also per-row dot and |term| partials of the walk kernel.
    void apply(const float *p, float *y, float *rowdot, float *rowabs) const {
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il) {
            uint32_t i = (uint32_t)il;
            orc::V3 sum = orc::v3(0, 0, 0), as = orc::v3(0, 0, 0);
            auto acc = [&](const orc::V3 &v) {
                sum = orc::add(sum, v);
                for (int c = 0; c < 3; ++c) {
                    as[c] = as[c] + std::fabs(v[c]);
                }
            };
            for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
                acc(orc::matvec(val + 9 * (size_t)k, p + 3 * (size_t)col[k]));
            }
            for (uint32_t k = tp[i]; k < tp[i + 1]; ++k) {
                acc(orc::matvec_t(val + 9 * (size_t)tslot[k],
                                  p + 3 * (size_t)tcol[k]));
            }
            acc(orc::matvec(cd + 9 * (size_t)i, p + 3 * (size_t)i));
            float d = 0.0f, a = 0.0f;
            for (int c = 0; c < 3; ++c) {
                y[3 * i + c] = sum[c];
                d = d + p[3 * i + c] * sum[c];
                a = a + std::fabs(p[3 * i + c]) * as[c];
            }
            if (rowdot) {
                rowdot[i] = d;
                rowabs[i] = a;
            }
        }
    }
};

const float kNoiseSlack = 256.0f;              // solver.cu:69
const float kFloatEps = 1.19209290e-7f;        // solver.cu:70

} // namespace

extern "C" {

// Outcome of one solve. status: 0 converged, 1 truncated (curvature inside
// its own round-off; reported as success, solver.cu:1287-1303), 2 iteration
// cap, 3 non-finite residual, 4 p^T A p negative beyond the bound (reference
// aborts), 5 r^T M^-1 r non-positive (reference aborts).
struct oracle_pcg_result {
    uint32_t iter;
    float resid;
    int32_t status;
    int32_t break_cause;
    int32_t break_iter;
    float err0;
};

void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// y = M x for the split operator (solver::apply, walk form).
void oracle_spmv(uint32_t n, const uint32_t *rp, const uint32_t *col,
                 const float *val, const float *cdiag, const float *x,
                 float *y) {
    Operator op{n, rp, col, val, cdiag, {}, {}, {}};
    op.build_transpose();
    op.apply(x, y, nullptr, nullptr);
}

// Merged diagonal blocks D_i = B(i,i) + C_i (solver.cu:2191-2195).
void oracle_merged_diag(uint32_t n, const uint32_t *rp, const uint32_t *col,
                        const float *val, const float *cdiag, float *out9) {
    for (uint32_t i = 0; i < n; ++i) {
        float d[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) {
            if (col[k] == i) {
                for (int e = 0; e < 9; ++e) {
                    d[e] = val[9 * (size_t)k + e];
                }
            }
        }
        for (int e = 0; e < 9; ++e) {
            out9[9 * (size_t)i + e] = d[e] + cdiag[9 * (size_t)i + e];
        }
    }
}

int oracle_pcg(uint32_t n, const uint32_t *rp, const uint32_t *col,
               const float *val, const float *cdiag, const float *b, float *x,
               float tol, uint32_t max_iter, float *err_hist,
               oracle_pcg_result *out) {
    Operator op{n, rp, col, val, cdiag, {}, {}, {}};
    op.build_transpose();
    const size_t n3 = 3 * (size_t)n;
    std::vector<float> P(9 * (size_t)n), D(9 * (size_t)n);
    oracle_merged_diag(n, rp, col, val, cdiag, D.data());
    uint32_t bad = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!orc::bj_invert(&D[9 * (size_t)i], &P[9 * (size_t)i])) {
            ++bad;
        }
    }
    std::memset(out, 0, sizeof(*out));
    if (bad) {
        out->status = 5;
        return out->status;
    }
    std::vector<float> r(n3), z(n3), p(n3), Ap(n3), rowdot(n), rowabs(n);
    op.apply(x, Ap.data(), nullptr, nullptr);
    for (size_t k = 0; k < n3; ++k) {
        r[k] = b[k] + (-1.0f) * Ap[k];
    }
    const float err0f = blocked_sum(n3, [&](size_t k) { return std::fabs(r[k]); });
    const double err0 = err0f;
    out->err0 = err0f;
    if (err0 == 0.0) {
        out->iter = 1;
        out->resid = 0.0f;
        out->status = 0;
        return 0;
    }
    auto precond = [&](uint32_t i) {
        orc::V3 zi = orc::matvec(&P[9 * (size_t)i], &r[3 * (size_t)i]);
        for (int c = 0; c < 3; ++c) {
            z[3 * (size_t)i + c] = zi[c];
        }
    };
    for (uint32_t i = 0; i < n; ++i) {
        precond(i);
    }
    float rz0 = blocked_sum(n3, [&](size_t k) { return r[k] * z[k]; });
    p = z;

    int breakdown = 0, cause = 0, break_iter = 0;
    uint32_t iter = 1, stride = 4, next_check = 1;
    while (true) {
        // SpMV + pAp + |terms| (walk kernel).
        op.apply(p.data(), Ap.data(), rowdot.data(), rowabs.data());
        float pAp = blocked_sum(n, [&](size_t i) { return rowdot[i]; });
        float absdot = blocked_sum(n, [&](size_t i) { return rowabs[i]; });
        // Curvature verdict (solver.cu:86-114).
        float bound = kNoiseSlack * kFloatEps * absdot;
        float alpha = 0.0f;
        if (pAp > bound) {
            alpha = rz0 / pAp;
        } else {
            breakdown = 1;
            cause = (pAp < -bound || pAp != pAp) ? 1 : 3;
            break_iter = (int)iter - 1;
        }
        // x += alpha p, r -= alpha Ap, z = P r.
#pragma omp parallel for schedule(static)
        for (long il = 0; il < (long)n; ++il) {
            for (int c = 0; c < 3; ++c) {
                size_t j = 3 * (size_t)il + c;
                x[j] = std::fmaf(alpha, p[j], x[j]);
                r[j] = std::fmaf(-alpha, Ap[j], r[j]);
            }
            precond((uint32_t)il);
        }
        float rz1 = blocked_sum(n3, [&](size_t k) { return r[k] * z[k]; });
        float err = blocked_sum(n3, [&](size_t k) { return std::fabs(r[k]); });
        float beta;
        if (rz0 > 0.0f) {
            beta = rz1 / rz0;
        } else {
            beta = 0.0f;
            if (cause == 0) {
                breakdown = 1;
                cause = 2;
                break_iter = (int)iter - 1;
            }
        }
        rz0 = rz1;
        if (err_hist) {
            err_hist[iter - 1] = err;
        }
        // p = 1*z + beta*p.
#pragma omp parallel for schedule(static)
        for (long k = 0; k < (long)n3; ++k) {
            p[k] = std::fmaf(beta, p[k], 1.0f * z[k]);
        }

        const bool at_cap = iter >= max_iter;
        if (iter == next_check || at_cap) {
            const double reresid = (double)err / err0;
            out->iter = iter;
            out->resid = (float)reresid;
            out->break_cause = cause;
            out->break_iter = break_iter;
            if (reresid < (double)tol) {
                out->status = 0;
                return 0;
            }
            if (breakdown) {
                out->status = cause == 3 ? 1 : (cause == 2 ? 5 : 4);
                return out->status;
            }
            if (!std::isfinite(reresid)) {
                out->status = 3;
                return 3;
            }
            if (at_cap) {
                out->status = 2;
                return 2;
            }
            if (reresid < 8.0 * (double)tol) {
                stride = 1;
            }
            next_check = iter + stride;
        }
        iter++;
    }
}
}
