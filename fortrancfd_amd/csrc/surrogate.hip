// surrogate.hip -- host side of the chemistry surrogate: parameter and workspace management,
// the forward / backward / Adam sequence of one training step, batched inference with FP64
// fallback, and the surrogate branch of solve_source.  All work is stream-ordered and
// allocation-free once the workspace is sized, so a solve_source with the surrogate can be
// captured in a HIP graph.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fcfd.h"
#include "chem.hpp"
#include "fcfd_common.hpp"
#include "surrogate_impl.hpp"

namespace fcfd {

void set_error(const std::string &msg);

namespace {

constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1.0e-8f;

long wsize(int l) { return long(MLP_K[l]) * MLP_N[l]; }

// Split-K factor for a weight-gradient GEMM with `tiles` output tiles and K = rows: about 256
// workgroups, each K slice a multiple of 16 rows.
int choose_split(int rows, int tiles) {
    int s = std::max(1, 256 / tiles);
    while (s > 1 && rows % (s * 16) != 0) --s;
    return s;
}

template <class T>
hipError_t grow(T *&p, size_t &cap, size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
}

}  // namespace

hipError_t Surrogate::reserve(int rows, bool training) {
    hipError_t e;
    const size_t r = size_t(rows);
    if ((e = grow(X, capX, r * MLP_INP)) != hipSuccess) return e;
    if ((e = grow(Yh, capYh, r * MLP_OUTP)) != hipSuccess) return e;
    if ((e = grow(in_range, capIR, r)) != hipSuccess) return e;
    if ((e = grow(fb_list, capFB, r)) != hipSuccess) return e;
    if ((e = grow(stiff, capST, r)) != hipSuccess) return e;
    if ((e = grow(cand_list, capCL, r)) != hipSuccess) return e;
    if ((e = grow(cand, capCD, r * size_t(mech->host.kk + 1))) != hipSuccess) return e;
    const size_t hr = std::min<size_t>(r, size_t(INFER_CHUNK));
    for (int i = 0; i < 2; ++i)
        if ((e = grow(Hping[i], capH[i], hr * MLP_HID)) != hipSuccess) return e;
    if (!training) return hipSuccess;
    for (int l = 0; l < 4; ++l) {
        if ((e = grow(H[l], capHl[l], r * MLP_HID)) != hipSuccess) return e;
        if ((e = grow(Pre[l], capPre[l], r * MLP_HID)) != hipSuccess) return e;
    }
    if ((e = grow(Yt, capYt, r * MLP_OUTP)) != hipSuccess) return e;
    if ((e = grow(dY, capdY, r * MLP_OUTP)) != hipSuccess) return e;
    for (int i = 0; i < 2; ++i)
        if ((e = grow(dZ[i], capdZ[i], r * MLP_HID)) != hipSuccess) return e;
    if ((e = grow(cs_part, capCS, ((r + 63) / 64) * MLP_HID)) != hipSuccess) return e;
    if ((e = grow(loss_part, capLP, (r * MLP_OUTP + 255) / 256 + 1)) != hipSuccess) return e;
    return hipSuccess;
}

hipError_t Surrogate::gemm(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, int M,
                           int N, int K, int ksplit, int ep, float *C32, uint16_t *C16, float *P32,
                           const float *bias, long ldc, hipStream_t st, const unsigned *m_dev, int m_off,
                           const int *row_live) const {
    GemmArgs g{};
    g.m_dev = m_dev;
    g.m_off = m_off;
    g.row_live = row_live;
    g.A = A; g.B = B; g.sam = sam; g.sak = sak; g.sbk = sbk; g.sbn = sbn;
    g.M = M; g.N = N; g.K = K; g.ksplit = ksplit;
    g.C32 = C32; g.C16 = C16; g.P32 = P32; g.bias = bias; g.ldc = ldc; g.epilogue = ep;
    return launch_gemm(g, st);
}

// Forward of `rows` rows of X (rows multiple of 64).  Training keeps every layer's activations
// and pre-activations; inference ping-pongs two activation buffers chunk by chunk.
hipError_t Surrogate::forward(const uint16_t *x, int rows, float *out, bool training, hipStream_t st,
                              const unsigned *m_dev, const int *row_live) {
    hipError_t e;
    if (training) {
        const uint16_t *in = x;
        for (int l = 0; l < 4; ++l) {
            e = gemm(in, MLP_K[l], 1, W16[l], MLP_N[l], 1, rows, MLP_N[l], MLP_K[l], 1, EP_BIAS_GELU, nullptr, H[l],
                     Pre[l], b32[l], MLP_N[l], st);
            if (e != hipSuccess) return e;
            in = H[l];
        }
        return gemm(in, MLP_HID, 1, W16[4], MLP_OUTP, 1, rows, MLP_OUTP, MLP_HID, 1, EP_BIAS, out, nullptr, nullptr,
                    b32[4], MLP_OUTP, st);
    }
    for (int r0 = 0; r0 < rows; r0 += INFER_CHUNK) {
        const int cr = std::min(INFER_CHUNK, rows - r0);
        const uint16_t *in = x + long(r0) * MLP_INP;
        for (int l = 0; l < 4; ++l) {
            uint16_t *o = Hping[l & 1];
            e = gemm(in, MLP_K[l], 1, W16[l], MLP_N[l], 1, cr, MLP_N[l], MLP_K[l], 1, EP_BIAS_GELU, nullptr, o,
                     nullptr, b32[l], MLP_N[l], st, m_dev, r0, row_live);
            if (e != hipSuccess) return e;
            in = o;
        }
        e = gemm(in, MLP_HID, 1, W16[4], MLP_OUTP, 1, cr, MLP_OUTP, MLP_HID, 1, EP_BIAS, out + long(r0) * MLP_OUTP,
                 nullptr, nullptr, b32[4], MLP_OUTP, st, m_dev, r0, row_live);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Gradients of the mean squared error of the current batch (X, Yt already in place) into gW / gb.
hipError_t Surrogate::backward(int rows, hipStream_t st) {
    hipError_t e;
    if ((e = forward(X, rows, Yh, true, st)) != hipSuccess) return e;
    if ((e = mlp_mse(Yh, Yt, dY, rows, loss_part, d_loss, st)) != hipSuccess) return e;
    const uint16_t *dz = dY;
    for (int l = MLP_LAYERS - 1; l >= 0; --l) {
        const int K = MLP_K[l], N = MLP_N[l];
        const uint16_t *hin = l == 0 ? X : H[l - 1];
        const int tiles = (K / 64) * (N / 64);
        const int ks = choose_split(rows, tiles);
        if ((e = gemm(hin, 1, K, dz, N, 1, K, N, rows, ks, EP_STORE, gpart, nullptr, nullptr, nullptr, N, st)) !=
            hipSuccess)
            return e;
        if ((e = mlp_reduce_splits(gpart, gW[l], ks, wsize(l), st)) != hipSuccess) return e;
        if ((e = mlp_colsum(dz, cs_part, gb[l], rows, N, st)) != hipSuccess) return e;
        if (l > 0) {
            uint16_t *dn = dZ[l & 1];
            if ((e = gemm(dz, N, 1, W16[l], 1, N, rows, K, N, 1, EP_GELU_BWD, nullptr, dn, Pre[l - 1], nullptr, K,
                          st)) != hipSuccess)
                return e;
            dz = dn;
        }
    }
    return hipSuccess;
}

hipError_t Surrogate::adam(float lr, hipStream_t st) {
    ++step;
    const float c1 = 1.0f - std::pow(ADAM_B1, float(step));
    const float c2 = 1.0f - std::pow(ADAM_B2, float(step));
    hipError_t e;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        if ((e = mlp_adam(W32[l], gW[l], mW[l], vW[l], W16[l], wsize(l), lr, ADAM_B1, ADAM_B2, ADAM_EPS, c1, c2, st)) !=
            hipSuccess)
            return e;
        if ((e = mlp_adam(b32[l], gb[l], mb[l], vb[l], nullptr, MLP_N[l], lr, ADAM_B1, ADAM_B2, ADAM_EPS, c1, c2, st)) !=
            hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t Surrogate::network(FeatArgs fa, const int *list, const unsigned *list_n, int rows_max, hipStream_t st) {
    fa.norm = d_norm;
    fa.thermo = mech->d_thermo;
    fa.kk = mech->host.kk;
    fa.list = list;
    fa.list_n = list_n;
    fa.X = X;
    fa.in_range = in_range;
    const int rows = int((long(rows_max) + 63) / 64 * 64);
    hipError_t e = mlp_features(fa, rows, st);
    if (e != hipSuccess) return e;
    // rows outside the training range go to the FP64 integrator whatever the network says: tiles of
    // 64 such rows are not evaluated
    return forward(X, rows, Yh, false, st, list_n, in_range);
}

hipError_t Surrogate::accept(ApplyArgs ap, const ChemArgs &ca, const int *list, const unsigned *list_n, int rows_max,
                             hipStream_t st) {
    ap.thermo = mech->d_thermo;
    ap.norm = d_norm;
    ap.ncf = d_ncf;
    ap.mm = mech->host.mm;
    ap.list = list;
    ap.list_n = list_n;
    ap.Yh = Yh;
    ap.in_range = in_range;
    ap.nfallback = d_nfb;
    ap.fallback = fb_list;
    ap.resid_tol = resid_tol;
    ap.t_tol = t_tol;
    ap.n = rows_max;
    ap.ncand = d_ncand;
    ap.cand_list = cand_list;
    ap.cand = cand;
    hipError_t e = launch_zero_u32(d_nfb, st);
    if (e != hipSuccess) return e;
    if ((e = launch_zero_u32(d_ncand, st)) != hipSuccess) return e;
    if ((e = mlp_apply(ap, mech->host.kk, st)) != hipSuccess) return e;
    return launch_chem_verify(ca, mech->host.kk, cand_list, d_ncand, cand, rows_max, fb_list, d_nfb, verify_tol, st);
}

hipError_t Surrogate::fallback(ChemArgs ca, int chem_wg, hipStream_t st) {
    ca.list = fb_list;
    ca.list_n = d_nfb;
    ca.skip_explicit = 1;
    return launch_chem(ca, mech->host.kk, chem_wg, st);
}

Surrogate::~Surrogate() {
    auto f = [](void *p) { if (p) (void)hipFree(p); };
    for (int l = 0; l < MLP_LAYERS; ++l) {
        f(W32[l]); f(b32[l]); f(W16[l]); f(mW[l]); f(vW[l]); f(mb[l]); f(vb[l]); f(gW[l]); f(gb[l]);
    }
    for (int l = 0; l < 4; ++l) { f(H[l]); f(Pre[l]); }
    for (int i = 0; i < 2; ++i) { f(Hping[i]); f(dZ[i]); }
    f(X); f(Yh); f(Yt); f(dY); f(gpart); f(cs_part); f(loss_part); f(d_loss); f(in_range); f(fb_list);
    f(stiff); f(d_stiff_n); f(d_nfb); f(d_norm); f(d_ncf); f(cand_list); f(cand); f(d_ncand);
    if (stream) (void)hipStreamDestroy(stream);
}

}  // namespace fcfd

using namespace fcfd;

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                      \
            return FCFD_E_HIP;                                                                 \
        }                                                                                      \
    } while (0)

extern "C" {

int fcfd_sur_create(const fcfd_mech *m, fcfd_surrogate **out) {
    if (!m || !out) { set_error("fcfd_sur_create: invalid argument"); return FCFD_E_ARG; }
    *out = nullptr;
    if (m->host.kk + 3 != MLP_IN || m->host.kk + 1 != MLP_OUT) {
        set_error("the surrogate network is sized for 53 species");
        return FCFD_E_UNSUPPORTED;
    }
    if (m->host.mm > 8) { set_error("at most 8 elements"); return FCFD_E_UNSUPPORTED; }
    if (int rc = ensure_dev_tables(m)) return rc;
    auto s = std::make_unique<fcfd_surrogate>();
    s->mech = m;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const size_t w = size_t(wsize(l)), b = size_t(MLP_N[l]);
        HIP_TRY(hipMalloc(&s->W32[l], w * 4)); HIP_TRY(hipMalloc(&s->W16[l], w * 2));
        HIP_TRY(hipMalloc(&s->mW[l], w * 4)); HIP_TRY(hipMalloc(&s->vW[l], w * 4)); HIP_TRY(hipMalloc(&s->gW[l], w * 4));
        HIP_TRY(hipMalloc(&s->b32[l], b * 4)); HIP_TRY(hipMalloc(&s->mb[l], b * 4)); HIP_TRY(hipMalloc(&s->vb[l], b * 4));
        HIP_TRY(hipMalloc(&s->gb[l], b * 4));
        HIP_TRY(hipMemset(s->W32[l], 0, w * 4)); HIP_TRY(hipMemset(s->W16[l], 0, w * 2));
        HIP_TRY(hipMemset(s->mW[l], 0, w * 4)); HIP_TRY(hipMemset(s->vW[l], 0, w * 4));
        HIP_TRY(hipMemset(s->b32[l], 0, b * 4)); HIP_TRY(hipMemset(s->mb[l], 0, b * 4)); HIP_TRY(hipMemset(s->vb[l], 0, b * 4));
    }
    size_t gp = 0;                       // largest split-K partial set (see choose_split)
    for (int l = 0; l < MLP_LAYERS; ++l) {
        const int tiles = (MLP_K[l] / 64) * (MLP_N[l] / 64);
        gp = std::max(gp, size_t(std::max(1, 256 / tiles)) * size_t(wsize(l)));
    }
    HIP_TRY(hipMalloc(&s->gpart, gp * 4));
    HIP_TRY(hipMalloc(&s->d_loss, 4));
    HIP_TRY(hipMalloc(&s->d_nfb, sizeof(unsigned)));
    HIP_TRY(hipMalloc(&s->d_stiff_n, sizeof(unsigned)));
    HIP_TRY(hipMalloc(&s->d_ncand, sizeof(unsigned)));
    HIP_TRY(hipMalloc(&s->d_norm, sizeof(SurrogateNorm)));
    SurrogateNorm nm{};
    for (int c = 0; c < MLP_INP; ++c) { nm.in_std[c] = 1.0; nm.lo[c] = -1e300; nm.hi[c] = 1e300; }
    for (int c = 0; c < MLP_OUTP; ++c) nm.out_std[c] = 1.0;
    s->h_norm = nm;
    HIP_TRY(hipMemcpy(s->d_norm, &nm, sizeof nm, hipMemcpyHostToDevice));
    const int kk = m->host.kk, mm = m->host.mm;
    std::vector<double> ncf(size_t(kk) * mm);
    for (size_t i = 0; i < ncf.size(); ++i) ncf[i] = double(m->host.ncf[i]);
    HIP_TRY(hipMalloc(&s->d_ncf, ncf.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(s->d_ncf, ncf.data(), ncf.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = s.release();
    return FCFD_OK;
}

void fcfd_sur_free(fcfd_surrogate *s) { delete s; }

int fcfd_sur_nparams(void) {
    int n = 0;
    for (int l = 0; l < MLP_LAYERS; ++l) n += int(wsize(l)) + MLP_N[l];
    return n;
}

int fcfd_sur_set_params(fcfd_surrogate *s, const float *params) {
    if (!s || !params) { set_error("fcfd_sur_set_params: invalid argument"); return FCFD_E_ARG; }
    const float *p = params;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        HIP_TRY(hipMemcpyAsync(s->W32[l], p, wsize(l) * 4, hipMemcpyHostToDevice, s->stream));
        p += wsize(l);
        HIP_TRY(hipMemcpyAsync(s->b32[l], p, MLP_N[l] * 4, hipMemcpyHostToDevice, s->stream));
        p += MLP_N[l];
        HIP_TRY(mlp_to_bf16(s->W32[l], s->W16[l], wsize(l), s->stream));
        HIP_TRY(hipMemsetAsync(s->mW[l], 0, wsize(l) * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->vW[l], 0, wsize(l) * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->mb[l], 0, MLP_N[l] * 4, s->stream));
        HIP_TRY(hipMemsetAsync(s->vb[l], 0, MLP_N[l] * 4, s->stream));
    }
    s->step = 0;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FCFD_OK;
}

int fcfd_sur_get_params(fcfd_surrogate *s, float *params) {
    if (!s || !params) { set_error("fcfd_sur_get_params: invalid argument"); return FCFD_E_ARG; }
    float *p = params;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        HIP_TRY(hipMemcpyAsync(p, s->W32[l], wsize(l) * 4, hipMemcpyDeviceToHost, s->stream));
        p += wsize(l);
        HIP_TRY(hipMemcpyAsync(p, s->b32[l], MLP_N[l] * 4, hipMemcpyDeviceToHost, s->stream));
        p += MLP_N[l];
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FCFD_OK;
}

int fcfd_sur_get_grads(fcfd_surrogate *s, float *grads) {
    if (!s || !grads) { set_error("fcfd_sur_get_grads: invalid argument"); return FCFD_E_ARG; }
    float *p = grads;
    for (int l = 0; l < MLP_LAYERS; ++l) {
        HIP_TRY(hipMemcpyAsync(p, s->gW[l], wsize(l) * 4, hipMemcpyDeviceToHost, s->stream));
        p += wsize(l);
        HIP_TRY(hipMemcpyAsync(p, s->gb[l], MLP_N[l] * 4, hipMemcpyDeviceToHost, s->stream));
        p += MLP_N[l];
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return FCFD_OK;
}

int fcfd_sur_set_norm(fcfd_surrogate *s, const double *norm) {
    if (!s || !norm) { set_error("fcfd_sur_set_norm: invalid argument"); return FCFD_E_ARG; }
    SurrogateNorm nm{};
    std::memcpy(&nm, norm, sizeof nm);
    for (int c = 0; c < MLP_INP; ++c)
        if (!(nm.in_std[c] > 0.0)) { set_error("fcfd_sur_set_norm: input scales must be positive"); return FCFD_E_ARG; }
    for (int c = 0; c < MLP_OUTP; ++c)
        if (!(nm.out_std[c] > 0.0)) { set_error("fcfd_sur_set_norm: output scales must be positive"); return FCFD_E_ARG; }
    s->h_norm = nm;
    HIP_TRY(hipMemcpy(s->d_norm, &nm, sizeof nm, hipMemcpyHostToDevice));
    return FCFD_OK;
}

int fcfd_sur_set_thresholds(fcfd_surrogate *s, double resid_tol, double t_tol, double verify_tol) {
    if (!s || !(resid_tol > 0.0) || !(t_tol > 0.0) || !(verify_tol > 0.0)) { set_error("fcfd_sur_set_thresholds: invalid argument"); return FCFD_E_ARG; }
    s->resid_tol = resid_tol;
    s->t_tol = t_tol;
    s->verify_tol = verify_tol;
    return FCFD_OK;
}

int fcfd_sur_features(fcfd_surrogate *s, int n, const double *rho, const double *dt, const double *T,
                      const double *Y, void *X, int *in_range, void *stream) {
    if (!s || n < 1 || !rho || !dt || !T || !Y || !X) { set_error("fcfd_sur_features: invalid argument"); return FCFD_E_ARG; }
    hipStream_t st = static_cast<hipStream_t>(stream);
    FeatArgs f{};
    f.norm = s->d_norm; f.thermo = s->mech->d_thermo; f.kk = s->mech->host.kk; f.layout = 1;
    f.n = n; f.npad = int((long(n) + 63) / 64 * 64);
    f.rho = rho; f.dtv = dt; f.T = T; f.Y = Y;
    f.X = static_cast<uint16_t *>(X); f.in_range = in_range;
    HIP_TRY(mlp_features(f, f.npad, st));
    HIP_TRY(hipStreamSynchronize(st));
    return FCFD_OK;
}

int fcfd_sur_targets(fcfd_surrogate *s, int n, const double *T0, const double *Y0, const double *T1,
                     const double *Y1, float *Tg, void *stream) {
    if (!s || n < 1 || !T0 || !Y0 || !T1 || !Y1 || !Tg) { set_error("fcfd_sur_targets: invalid argument"); return FCFD_E_ARG; }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int npad = int((long(n) + 63) / 64 * 64);
    HIP_TRY(mlp_targets(s->d_norm, s->mech->host.kk, n, npad, T0, Y0, T1, Y1, Tg, st));
    HIP_TRY(hipStreamSynchronize(st));
    return FCFD_OK;
}

int fcfd_sur_forward(fcfd_surrogate *s, int rows, const void *X, float *Yh, void *stream) {
    if (!s || rows < 64 || rows % 64 || !X || !Yh) { set_error("fcfd_sur_forward: rows must be a positive multiple of 64"); return FCFD_E_ARG; }
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(s->reserve(std::min(rows, INFER_CHUNK), false));
    HIP_TRY(s->forward(static_cast<const uint16_t *>(X), rows, Yh, false, st));
    HIP_TRY(hipStreamSynchronize(st));
    return FCFD_OK;
}

int fcfd_sur_train_step(fcfd_surrogate *s, const void *X_all, const float *Tg_all, const int *idx, int batch,
                        float lr, float *loss, void *stream) {
    if (!s || !X_all || !Tg_all || !idx || batch < 64 || batch % 64 || !(lr >= 0.0f)) {
        set_error("fcfd_sur_train_step: invalid argument (batch must be a positive multiple of 64)");
        return FCFD_E_ARG;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(s->reserve(batch, true));
    HIP_TRY(mlp_gather_bf16(static_cast<const uint16_t *>(X_all), s->X, idx, batch, MLP_INP, st));
    HIP_TRY(mlp_gather_f32(Tg_all, s->Yt, idx, batch, MLP_OUTP, st));
    HIP_TRY(s->backward(batch, st));
    if (lr > 0.0f) HIP_TRY(s->adam(lr, st));
    if (loss) {
        HIP_TRY(hipMemcpyAsync(loss, s->d_loss, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *loss /= float(batch) * float(MLP_OUT);
    }
    return FCFD_OK;
}

int fcfd_sur_infer_batch(fcfd_surrogate *s, const fcfd_opts *o, int n, const double *rho, const double *dt,
                         double *T, double *Y, int *accepted, int *n_fallback, int *stats, void *stream) {
    if (!s || !o || n < 1 || !rho || !dt || !T || !Y) { set_error("fcfd_sur_infer_batch: invalid argument"); return FCFD_E_ARG; }
    if (!(o->rtol > 0.0 && o->atol > 0.0)) { set_error("fcfd_sur_infer_batch: atol and rtol must be positive"); return FCFD_E_ARG; }
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int npad = int((long(n) + 63) / 64 * 64);
    HIP_TRY(s->reserve(npad, false));
    ChemArgs ca{};
    ca.thermo = s->mech->d_thermo; ca.kin = static_cast<const DevKin *>(s->mech->d_kin); ca.counter = s->mech->d_counter;
    ca.ncells = n; ca.rtol = o->rtol; ca.atol = o->atol;
    ca.b_rho = rho; ca.b_dt = dt; ca.b_T = T; ca.b_y = Y;
    ca.stats = reinterpret_cast<ChemCellStats *>(stats);
    ca.layout = 1;
    ca.route = accepted;
    HIP_TRY(launch_zero_u32(s->d_stiff_n, st));
    ca.stiff = s->stiff; ca.stiff_n = s->d_stiff_n;
    HIP_TRY(launch_chem_screen(ca, s->mech->host.kk, st));
    FeatArgs fa{};
    fa.layout = 1; fa.n = n; fa.rho = rho; fa.dtv = dt; fa.T = T; fa.Y = Y;
    HIP_TRY(s->network(fa, s->stiff, s->d_stiff_n, n, st));
    ApplyArgs ap{};
    ap.layout = 1; ap.rho = rho; ap.T = T; ap.Y = Y; ap.route = accepted;
    HIP_TRY(s->accept(ap, ca, s->stiff, s->d_stiff_n, n, st));
    HIP_TRY(s->fallback(ca, chem_workgroups(s->mech), st));
    if (n_fallback) {
        unsigned nf = 0;
        HIP_TRY(hipMemcpyAsync(&nf, s->d_nfb, sizeof nf, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *n_fallback = int(nf);
    }
    HIP_TRY(hipStreamSynchronize(st));
    return FCFD_OK;
}

int fcfd_bf16_gemm(int M, int N, int K, const void *A, long sam, long sak, const void *B, long sbk, long sbn,
                   float *C, long ldc, int ksplit, void *stream) {
    if (!A || !B || !C || M < 64 || N < 64 || K < 16 || ksplit < 1) { set_error("fcfd_bf16_gemm: invalid argument"); return FCFD_E_ARG; }
    GemmArgs g{};
    g.A = static_cast<const uint16_t *>(A); g.B = static_cast<const uint16_t *>(B);
    g.sam = sam; g.sak = sak; g.sbk = sbk; g.sbn = sbn; g.M = M; g.N = N; g.K = K; g.ksplit = ksplit;
    g.C32 = C; g.ldc = ldc; g.epilogue = EP_STORE;
    hipError_t e = launch_gemm(g, static_cast<hipStream_t>(stream));
    if (e == hipErrorInvalidValue) { set_error("fcfd_bf16_gemm: M, N must be multiples of 64 and K / ksplit a multiple of 16"); return FCFD_E_ARG; }
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    return FCFD_OK;
}

}  // extern "C"
