// surrogate.hpp -- device-side records of the chemistry surrogate and the launchers of mlp.hip.
#pragma once

#include <cstdint>

#include "fcfd_common.hpp"
#include "mlp.hpp"

namespace fcfd {

// Feature / target scaling and the training range (raw feature units), columns as the network's.
struct SurrogateNorm {
    double in_mean[MLP_INP], in_std[MLP_INP];
    double lo[MLP_INP], hi[MLP_INP];
    double out_mean[MLP_OUTP], out_std[MLP_OUTP];
};

// Feature rows (k_features).
struct FeatArgs {
    const SurrogateNorm *norm;
    const DevThermo *thermo;
    int kk, layout;
    int n, npad;                 // cells; rows written when list is null
    const int *list;             // optional: row r <- cell list[r], r < *list_n
    const unsigned *list_n;
    const double *qf;            // layout 0
    long P;
    int ncell, lv, S;
    const double *dt_dev;
    double dt;
    const double *rho, *dtv, *T, *Y;   // layout 1 (Y [KK][n])
    uint16_t *X;
    int *in_range;
};

// Inference epilogue (k_apply): network output -> accepted state or a fallback-list entry.
struct ApplyArgs {
    const DevThermo *thermo;
    const SurrogateNorm *norm;
    const double *ncf;          // [KK][MM] element counts
    int mm, n, layout;          // layout 0: padded SoA domain (qf), 1: batch arrays
    double *qf;
    long P;
    int ncell, lv, S;
    const double *rho;          // layout 1
    double *T, *Y;              // layout 1: T[n], Y[KK][n]
    const int *list;            // optional: row r is cell list[r], r < *list_n (else row = cell < n)
    const unsigned *list_n;
    const float *Yh;            // [rows][MLP_OUTP] normalised network output
    const int *in_range;        // [rows]
    int *route;                 // [n] or null: 1 surrogate accepted, 0 sent to the FP64 integrator
    unsigned *nfallback;        // fallback-list length (zeroed before the launch)
    int *fallback;              // fallback list (cell indices)
    unsigned *ncand;            // candidate-list length (zeroed before the launch)
    int *cand_list;             // candidate cells
    double *cand;               // [candidate][KK + 1]: Y1, T1
    double resid_tol, t_tol;
};

hipError_t mlp_features(const FeatArgs &f, int rows, hipStream_t st);
hipError_t mlp_targets(const SurrogateNorm *nm, int kk, int n, int npad, const double *T0, const double *Y0,
                       const double *T1, const double *Y1, float *Tg, hipStream_t st);
hipError_t mlp_gather_bf16(const uint16_t *src, uint16_t *dst, const int *idx, int rows, int w, hipStream_t st);
hipError_t mlp_gather_f32(const float *src, float *dst, const int *idx, int rows, int w, hipStream_t st);
hipError_t mlp_mse(const float *Yh, const float *Yt, uint16_t *dY, int rows, float *partial, float *loss,
                   hipStream_t st);
hipError_t mlp_reduce_splits(const float *parts, float *out, int nsplit, long len, hipStream_t st);
hipError_t mlp_colsum(const uint16_t *D, float *partial, float *out, int rows, int cols, hipStream_t st);
hipError_t mlp_adam(float *w, const float *g, float *m, float *v, uint16_t *w16, long n, float lr, float b1, float b2,
                    float eps, float c1, float c2, hipStream_t st);
hipError_t mlp_to_bf16(const float *w, uint16_t *w16, long n, hipStream_t st);
hipError_t mlp_zero_u32(unsigned *p, hipStream_t st);
hipError_t mlp_apply(const ApplyArgs &a, int nsp, hipStream_t st);

}  // namespace fcfd
