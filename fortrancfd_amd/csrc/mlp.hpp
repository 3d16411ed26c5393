// mlp.hpp -- the chemistry surrogate: a 56 -> 4x256 GELU -> 54 multilayer perceptron with BF16
// weights and activations and FP32 accumulation on MFMA (mlp.hip).
#pragma once

#include <cstdint>

#include "fcfd_common.hpp"

namespace fcfd {

constexpr int MLP_IN = 56;       // Box-Cox Y (53), T, log p, log dt
constexpr int MLP_OUT = 54;      // dY (53), dT
constexpr int MLP_HID = 256;
constexpr int MLP_LAYERS = 5;
constexpr int MLP_INP = 64;      // padded widths (multiples of the 16-wide MFMA tiles)
constexpr int MLP_OUTP = 64;

// Layer l maps K[l] -> N[l] columns.
constexpr int MLP_K[MLP_LAYERS] = {MLP_INP, MLP_HID, MLP_HID, MLP_HID, MLP_HID};
constexpr int MLP_N[MLP_LAYERS] = {MLP_HID, MLP_HID, MLP_HID, MLP_HID, MLP_OUTP};

enum GemmEpilogue {
    EP_STORE = 0,      // C32 = acc
    EP_BIAS = 1,       // C32 = acc + bias
    EP_BIAS_GELU = 2,  // P32 = acc + bias; C16 = bf16(gelu(P32))
    EP_GELU_BWD = 3,   // C16 = bf16(acc * gelu'(P32))
};

struct GemmArgs {
    const uint16_t *A;   // bf16, element (m, k) at A[m*sam + k*sak]
    const uint16_t *B;   // bf16, element (k, n) at B[k*sbk + n*sbn]
    long sam, sak, sbk, sbn;
    int M, N, K;         // M, N multiples of 64; K multiple of 16
    int ksplit;          // split-K factor (EP_STORE only): partial s goes to C32 + s*M*ldc
    float *C32;
    uint16_t *C16;
    float *P32;          // pre-activations (EP_BIAS_GELU writes, EP_GELU_BWD reads)
    const float *bias;
    long ldc;
    int epilogue;
    const unsigned *m_dev;   // optional device row count: 64-row tiles with m_off + m0 >= *m_dev are skipped
    int m_off;
    // optional (with m_dev): per-row flags indexed like m_off + row; a 64-row tile none of whose rows
    // below *m_dev is flagged is skipped, its output rows left as they were
    const int *row_live;
};

hipError_t launch_gemm(const GemmArgs &a, hipStream_t st);

}  // namespace fcfd
