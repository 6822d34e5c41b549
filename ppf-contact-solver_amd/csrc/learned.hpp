// learned.hpp -- learned preconditioner: a message-passing network over the
// 3x3-block graph of a Newton system that predicts a block FSAI factor.
//
// For A (stored upper blocks + per-row diagonal), A_ii = L_i L_i^T and
// Ahat = L^-1 A L^-T, the network gives Ghat (block lower triangular on A's
// pattern): Ghat_jj = I + lower(0.5 tanh t_j), Ghat_ji = 0.5 tanh(u_s) for the
// stored slot s = (i, j). The preconditioner is M^-1 = G^T G, G = Ghat L^-1:
// symmetric positive definite for any weights, block Jacobi when the output
// heads are zero. Training minimises the log Kaporin number of
// Ghat Ahat Ghat^T with Rademacher probes. Dense layers run on bf16 MFMA
// (fp32 accumulate); everything else is fp32.
#pragma once
#include "pcg.hpp"

namespace ppf {

constexpr uint32_t kLpcH = 64, kLpcFN = 16, kLpcFE = 16, kLpcLayers = 3;

// Offsets of the weights in the flat fp32 array, row-major [in][out]:
// W0 [16][64] c0 [64]; per layer We [80][64] ce [64] Wn [128][64] cn [64];
// Wo [64][6] co [6]; Wg [144][9] cg [9].
struct LpcLayout {
    static constexpr uint32_t W0 = 0, c0 = W0 + kLpcFN * kLpcH;
    static constexpr uint32_t L0 = c0 + kLpcH;
    static constexpr uint32_t kWe = (kLpcH + kLpcFE) * kLpcH, kWn = 2 * kLpcH * kLpcH;
    static constexpr uint32_t kLayer = kWe + kLpcH + kWn + kLpcH;
    static constexpr uint32_t We(uint32_t l) { return L0 + l * kLayer; }
    static constexpr uint32_t ce(uint32_t l) { return We(l) + kWe; }
    static constexpr uint32_t Wn(uint32_t l) { return ce(l) + kLpcH; }
    static constexpr uint32_t cn(uint32_t l) { return Wn(l) + kWn; }
    static constexpr uint32_t Wo = L0 + kLpcLayers * kLayer, co = Wo + kLpcH * 6;
    static constexpr uint32_t Wg = co + 6, cg = Wg + (2 * kLpcH + kLpcFE) * 9;
    static constexpr uint32_t count = cg + 9;
};

class Lpc {
  public:
    ~Lpc();
    int set_weights(const float *host, uint32_t count, hipStream_t s);
    int get_weights(float *host, uint32_t count, hipStream_t s);
    // Graph, diagonal Cholesky factors and features of the system.
    int prepare(const PcgPattern &pt, const float *val, const float *cdiag, const float *b,
                hipStream_t s);
    // Network and factor (after prepare): Gd (n blocks), Gl (nnz blocks).
    int forward(hipStream_t s);
    // Loss and weight gradients for the probes (nprobe x n x 3) after forward;
    // the plan provides A's SpMV.
    int gradient(PcgPlanImpl *plan, const float *val, const float *cdiag, const float *probes,
                 uint32_t nprobe, float *loss, hipStream_t s);
    int adam(float lr, uint32_t step, hipStream_t s);
    // null until forward has run on the current system
    const float *Gd() const { return factored_ ? Gd_ : nullptr; }
    const float *Gl() const { return factored_ ? Gl_ : nullptr; }
    // the same blocks in bf16 10-entry records (what the PCG applies), plus
    // the rows of G in order; null members until forward has run
    PcgFactor factor() const {
        return factored_ ? PcgFactor{Gd16_, Gl16_, Gr16_} : PcgFactor{nullptr, nullptr, nullptr};
    }
    const float *t() const { return t_; }
    const float *u() const { return u_; }
    const float *grad() const { return grad_; }
    uint32_t n() const { return n_; }
    uint32_t nnz() const { return nnz_; }

  private:
    int reserve(uint32_t n, uint32_t nnz, uint32_t E);
    int refresh_bf16(hipStream_t s);
    void release();
    void *mem_ = nullptr;
    size_t mem_bytes_ = 0;
    uint32_t n_ = 0, nnz_ = 0, E_ = 0, ncap_ = 0, nzcap_ = 0, ecap_ = 0;
    PcgPattern pt_{};
    // weights and optimiser state (fp32) and the bf16 GEMM operands
    float *w_ = nullptr, *grad_ = nullptr, *am_ = nullptr, *av_ = nullptr;
    uint16_t *wt_ = nullptr; // transposed [out][in] per dense layer (forward)
    uint16_t *wr_ = nullptr; // row-major [in][out] per dense layer (backward data)
    bool weights_set_ = false, factored_ = false;
    // graph (all inside mem_)
    uint32_t *eoff_ = nullptr, *esrc_ = nullptr, *eslot_ = nullptr, *erev_ = nullptr;
    uint32_t *deg_ = nullptr, *srow_ = nullptr, *semir_ = nullptr, *sstore_ = nullptr;
    uint32_t *scan_ = nullptr;
    // system data
    float *linv_ = nullptr;
    uint16_t *x_ = nullptr, *f_ = nullptr;
    float *fnorm_ = nullptr;
    uint16_t *h_[kLpcLayers + 1] = {}, *m_[kLpcLayers] = {}, *M_[kLpcLayers] = {};
    float *t_ = nullptr, *u_ = nullptr, *ghd_ = nullptr, *ghl_ = nullptr;
    float *Gd_ = nullptr, *Gl_ = nullptr;
    uint16_t *Gd16_ = nullptr, *Gl16_ = nullptr, *Gr16_ = nullptr;
    // backward scratch
    float *dh_ = nullptr, *dgd_ = nullptr, *dgl_ = nullptr, *scal_ = nullptr;
    uint16_t *dZ_ = nullptr, *dY_ = nullptr, *dXn_ = nullptr, *dXe_ = nullptr;
    float *part_ = nullptr;
    size_t part_words_ = 0;
    float *pv_ = nullptr, *pav_ = nullptr, *py_ = nullptr;
    uint32_t nprobe_cap_ = 0;
};

} // namespace ppf
