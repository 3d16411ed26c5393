// surrogate_impl.hpp -- the surrogate object behind the fcfd_surrogate handle (surrogate.hip).
#pragma once

#include <cstddef>
#include <cstdint>

#include "chem.hpp"
#include "surrogate.hpp"

namespace fcfd {

constexpr int INFER_CHUNK = 65536;     // rows per inference pass through the activation buffers

struct Surrogate {
    const fcfd_mech *mech = nullptr;
    hipStream_t stream = nullptr;        // for the parameter copies of the set/get calls
    // parameters: W[l] is K[l] x N[l] row-major (FP32 master, BF16 copy), biases FP32
    float *W32[MLP_LAYERS] = {}, *b32[MLP_LAYERS] = {};
    uint16_t *W16[MLP_LAYERS] = {};
    float *mW[MLP_LAYERS] = {}, *vW[MLP_LAYERS] = {}, *mb[MLP_LAYERS] = {}, *vb[MLP_LAYERS] = {};
    float *gW[MLP_LAYERS] = {}, *gb[MLP_LAYERS] = {};
    long step = 0;
    // workspace (grown on demand, never inside a captured call)
    uint16_t *X = nullptr;               // [rows][64] features
    float *Yh = nullptr, *Yt = nullptr;  // [rows][64]
    uint16_t *H[4] = {}, *Hping[2] = {}, *dY = nullptr, *dZ[2] = {};
    float *Pre[4] = {};
    float *gpart = nullptr, *cs_part = nullptr, *loss_part = nullptr, *d_loss = nullptr;
    int *in_range = nullptr, *fb_list = nullptr, *stiff = nullptr, *cand_list = nullptr;
    unsigned *d_nfb = nullptr, *d_stiff_n = nullptr, *d_ncand = nullptr;
    double *cand = nullptr;                // [rows][KK + 1] surrogate candidates awaiting verification
    size_t capX = 0, capYh = 0, capYt = 0, capdY = 0, capIR = 0, capFB = 0, capCS = 0, capLP = 0, capST = 0;
    size_t capCL = 0, capCD = 0;
    size_t capH[2] = {}, capHl[4] = {}, capPre[4] = {}, capdZ[2] = {};
    SurrogateNorm h_norm{};
    SurrogateNorm *d_norm = nullptr;
    double *d_ncf = nullptr;
    double resid_tol = 1.0e-5;           // element-conservation residual bound
    double t_tol = 2.0e-3;               // network T vs energy-consistent T (relative)
    double verify_tol = 1.0;             // FP64 trapezoidal residual, RMS in integrator tolerances

    hipError_t reserve(int rows, bool training);
    hipError_t gemm(const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, int M, int N, int K,
                    int ksplit, int ep, float *C32, uint16_t *C16, float *P32, const float *bias, long ldc,
                    hipStream_t st, const unsigned *m_dev = nullptr, int m_off = 0,
                    const int *row_live = nullptr) const;
    // rows: rows allocated in x / out (multiple of 64); m_dev: optional device count of live rows;
    // row_live (inference with m_dev): per-row flags, 64-row tiles with none set are not evaluated
    hipError_t forward(const uint16_t *x, int rows, float *out, bool training, hipStream_t st,
                       const unsigned *m_dev = nullptr, const int *row_live = nullptr);
    hipError_t backward(int rows, hipStream_t st);
    hipError_t adam(float lr, hipStream_t st);
    // Surrogate branch of the chemistry source for the cells list[0 .. *list_n) that the explicit
    // screen rejected: features + network (rows <= rows_max), acceptance test, FP64 verification of
    // the candidates, FP64 SDIRK for the rest.
    hipError_t network(FeatArgs fa, const int *list, const unsigned *list_n, int rows_max, hipStream_t st);
    hipError_t accept(ApplyArgs ap, const ChemArgs &ca, const int *list, const unsigned *list_n, int rows_max,
                      hipStream_t st);
    hipError_t fallback(ChemArgs ca, int chem_wg, hipStream_t st);
    ~Surrogate();
};

int ensure_dev_tables(const fcfd_mech *m);
int chem_workgroups(const fcfd_mech *m);

}  // namespace fcfd

struct fcfd_surrogate : fcfd::Surrogate {};
