// mlp.hip -- kernels of the chemistry surrogate (forward, backward, Adam, feature / target
// construction, and the inference epilogue that decides between the surrogate and the FP64
// integrator for each cell).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] in BF16 with FP32 accumulation on v_mfma_f32_16x16x16bf16_1k.
// A workgroup of 4 waves owns a 64 x 64 tile of C (each wave 32 x 32 = 2 x 2 MFMA blocks); the K loop
// stages a 64 x 16 slice of A and a 16 x 64 slice of B in LDS (B stored n-major so each lane's
// 4-element fragment is one 8-byte read).  Operands are addressed through (row, col) strides, so
// the same kernel serves X*W (forward), dPre*W^T (backward data) and H^T*dPre (backward weights,
// split over K into partial buffers reduced in a fixed order).
#include <cmath>

#include "fcfd_device.hpp"
#include "mlp.hpp"
#include "surrogate.hpp"

namespace fcfd {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {        // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t(u >> 16) | ((u & 0xffffu) ? 0x40 : 0);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// Box-Cox transform of a mass fraction, lambda = 1/10: b(Y) = 10 (Y^0.1 - 1), b(0) = -10, and its
// inverse Y = (b / 10 + 1)^10 (0 where the base is not positive).
__device__ __forceinline__ double boxcox(double y) { return 10.0 * (pow(fmax(y, 0.0), 0.1) - 1.0); }
__device__ __forceinline__ double boxcox_inv(double b) {
    const double v = 0.1 * b + 1.0;
    if (!(v > 0.0)) return 0.0;
    const double v2 = v * v, v4 = v2 * v2, v8 = v4 * v4;
    return v8 * v2;
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gemm(GemmArgs a) {
    __shared__ uint16_t As[64][16 + 4];
    __shared__ uint16_t Bs[64][16 + 4];     // n-major: Bs[n][k]
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int tiles_n = a.N / 64;
    const int tile = blockIdx.x;
    const int m0 = (tile / tiles_n) * 64, n0 = (tile % tiles_n) * 64;
    if (a.m_dev && a.m_off + m0 >= int(*a.m_dev)) return;   // uniform per workgroup, before any barrier
    const int split = blockIdx.y;
    const int kchunk = a.K / a.ksplit;
    const int kbeg = split * kchunk, kend = kbeg + kchunk;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
    const bool a_kfast = a.sak == 1;
    const bool b_nfast = a.sbn == 1;
    for (int k0 = kbeg; k0 < kend; k0 += 16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = tid + r * 256;
            int mm, kk;
            if (a_kfast) { mm = e >> 4; kk = e & 15; } else { mm = e & 63; kk = e >> 6; }
            As[mm][kk] = a.A[long(m0 + mm) * a.sam + long(k0 + kk) * a.sak];
            int nn, kb;
            if (b_nfast) { nn = e & 63; kb = e >> 6; } else { nn = e >> 4; kb = e & 15; }
            Bs[nn][kb] = a.B[long(k0 + kb) * a.sbk + long(n0 + nn) * a.sbn];
        }
        __syncthreads();
        const int r = lane & 15, kq = (lane >> 4) * 4;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const v4s av = *reinterpret_cast<const v4s *>(&As[wm + i * 16 + r][kq]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const v4s bv = *reinterpret_cast<const v4s *>(&Bs[wn + j * 16 + r][kq]);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout: lane holds rows 4*(lane/16)+q, column lane%16 of each 16 x 16 block
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
                const int n = n0 + wn + j * 16 + (lane & 15);
                const long o = long(m) * a.ldc + n;
                const float v = acc[i][j][q];
                switch (a.epilogue) {
                    case EP_STORE: a.C32[long(split) * a.M * a.ldc + o] = v; break;
                    case EP_BIAS: a.C32[o] = v + a.bias[n]; break;
                    case EP_BIAS_GELU: {
                        const float p = v + a.bias[n];
                        if (a.P32) a.P32[o] = p;
                        a.C16[o] = f2bf(gelu(p));
                        break;
                    }
                    default: a.C16[o] = f2bf(v * gelu_d(a.P32[o])); break;
                }
            }
}

hipError_t launch_gemm(const GemmArgs &a, hipStream_t st) {
    if (a.M % 64 || a.N % 64 || a.K % 16 || a.ksplit < 1 || (a.K / a.ksplit) % 16) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm, dim3(unsigned((a.M / 64) * (a.N / 64)), unsigned(a.ksplit)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// Features: x = (b(Y_k), T, log(p/1 atm), log dt) z-scored (b: Box-Cox above); pad to 64 columns.
// in_range: every raw feature inside the training range [lo, hi].
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void features_of(const SurrogateNorm *__restrict__ nm, const DevThermo *__restrict__ th,
                                            int kk, double rho, double T, double dt, const double *y, long stride,
                                            uint16_t *xrow, int *in_range) {
    double swy = 0.0;
    for (int k = 0; k < kk; ++k) swy += y[k * stride] / th->wt[k];
    const double p = rho * 8.31451 * T * swy * 1.0e3;          // Pa (ideal gas, SI)
    bool ok = rho > 0.0 && T > 0.0 && dt > 0.0;
    for (int c = 0; c < MLP_INP; ++c) {
        double raw = 0.0;
        if (c < kk) raw = boxcox(y[c * stride]);
        else if (c == kk) raw = T;
        else if (c == kk + 1) raw = log(p / 101325.0);
        else if (c == kk + 2) raw = log(dt);
        if (c < MLP_IN) {
            if (!(raw >= nm->lo[c] && raw <= nm->hi[c])) ok = false;
            xrow[c] = f2bf(float((raw - nm->in_mean[c]) / nm->in_std[c]));
        } else {
            xrow[c] = 0;
        }
    }
    if (in_range) *in_range = ok ? 1 : 0;
}

// One row of features per cell: rows 0 .. npad of all cells (list == null; rows >= n zero-filled)
// or rows 0 .. *list_n for the cells list[row].
__global__ void k_features(FeatArgs f) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int cell = t;
    if (f.list) {
        if (t >= int(*f.list_n)) return;
        cell = f.list[t];
    } else {
        if (t >= f.npad) return;
        if (t >= f.n) {
            for (int j = 0; j < MLP_INP; ++j) f.X[long(t) * MLP_INP + j] = 0;
            if (f.in_range) f.in_range[t] = 0;
            return;
        }
    }
    const int kk = f.kk;
    double y[MAXK];
    double rho, T, dt;
    if (f.layout == 0) {
        const long col = long(cell / f.ncell) * f.S + f.lv + (cell % f.ncell);
        double s = 0.0;
        for (int k = 0; k < kk; ++k) { y[k] = fmax(f.qf[(3 + k) * f.P + col], 0.0); s += y[k]; }
        for (int k = 0; k < kk; ++k) y[k] /= s;
        rho = f.qf[col];
        T = f.qf[2 * f.P + col];
        dt = f.dt_dev ? *f.dt_dev : f.dt;
    } else {
        for (int k = 0; k < kk; ++k) y[k] = f.Y[long(k) * f.n + cell];
        rho = f.rho[cell];
        T = f.T[cell];
        dt = f.dtv[cell];
    }
    features_of(f.norm, f.thermo, kk, rho, T, dt, y, 1, f.X + long(t) * MLP_INP, f.in_range ? f.in_range + t : nullptr);
}

// Normalised targets: (b(Y1) - b(Y0), T1 - T0) z-scored, columns >= 54 zero.
__global__ void k_targets(const SurrogateNorm *nm, int kk, int n, int npad, const double *T0, const double *Y0,
                          const double *T1, const double *Y1, float *Tg) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= npad) return;
    for (int j = 0; j < MLP_OUTP; ++j) {
        double v = 0.0;
        if (c < n && j < kk)
            v = (boxcox(Y1[long(j) * n + c]) - boxcox(Y0[long(j) * n + c]) - nm->out_mean[j]) / nm->out_std[j];
        else if (c < n && j == kk) v = (T1[c] - T0[c] - nm->out_mean[j]) / nm->out_std[j];
        Tg[long(c) * MLP_OUTP + j] = float(v);
    }
}

// Rows idx[r] of src (width w) -> rows r of dst.
template <class T>
__global__ void k_gather_rows(const T *src, T *dst, const int *idx, int rows, int w) {
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= long(rows) * w) return;
    const int r = int(e / w), j = int(e % w);
    dst[e] = src[long(idx[r]) * w + j];
}

// Mean squared error over the first MLP_OUT columns: dY = 2 (Yh - Yt) / (rows * MLP_OUT) in BF16;
// per-block partial sums of the squared error.
__global__ void __launch_bounds__(256) k_mse(const float *Yh, const float *Yt, uint16_t *dY, int rows, float *partial) {
    __shared__ float red[256];
    const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
    float se = 0.f;
    if (e < long(rows) * MLP_OUTP) {
        const int j = int(e % MLP_OUTP);
        const float d = j < MLP_OUT ? Yh[e] - Yt[e] : 0.f;
        se = d * d;
        dY[e] = f2bf(2.f * d / float(rows * MLP_OUT));
    }
    red[threadIdx.x] = se;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out[j] = sum_s parts[s*len + j] (fixed order).
__global__ void k_reduce_splits(const float *parts, float *out, int nsplit, long len) {
    const long j = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = 0.f;
    for (int q = 0; q < nsplit; ++q) s += parts[q * len + j];
    out[j] = s;
}

// Column sums of a BF16 matrix (bias gradients), first stage: partial[chunk][j] over 64-row chunks.
__global__ void k_colsum(const uint16_t *D, float *partial, int rows, int cols) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int chunk = blockIdx.y;
    if (j >= cols) return;
    float s = 0.f;
    const int r1 = min(rows, (chunk + 1) * 64);
    for (int r = chunk * 64; r < r1; ++r) s += bf2f(D[long(r) * cols + j]);
    partial[long(chunk) * cols + j] = s;
}

__global__ void k_zero_u32(unsigned *p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0u;
}

__global__ void k_sum_partials(const float *p, int n, float *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (int i = 0; i < n; ++i) s += p[i];
        *out = s;
    }
}

// Adam on FP32 master parameters; refreshes the BF16 copies of weights (w16 may be null for biases).
__global__ void k_adam(float *w, const float *g, float *m, float *v, uint16_t *w16, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float *w, uint16_t *w16, long n) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) w16[i] = f2bf(w[i]);
}

// ---------------------------------------------------------------------------------------------
// Inference epilogue: one thread per cell.  Candidate Y1 = Y0 + dY (network), negative entries
// clipped, element balance restored by a weighted projection onto sum_k a_ek dY_k = 0
// (a_ek = atoms of element e per unit mass of species k, weights Y1_k), then T1 from conservation of
// sum_k Y_k u_k(T) / W_k (Newton from the network's T).  Accepted when the inputs were inside the
// training range, the element residual after projection and clipping is <= resid_tol and the
// network temperature agrees with the energy temperature to t_tol; otherwise the cell is queued for
// the FP64 integrator.
// ---------------------------------------------------------------------------------------------
template <int NSP>
__global__ void k_apply(ApplyArgs a) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;   // network row
    int t = row;                                               // cell
    if (a.list) {
        if (row >= int(*a.list_n)) return;
        t = a.list[row];
    } else if (row >= a.n) {
        return;
    }
    const DevThermo *__restrict__ th = a.thermo;
    const SurrogateNorm *__restrict__ nm = a.norm;
    long col = 0;
    double T0, y0[NSP];
    if (a.layout == 0) {
        const int b = t / a.ncell, i = t % a.ncell;
        col = long(b) * a.S + a.lv + i;
        T0 = a.qf[2 * a.P + col];
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) { y0[k] = fmax(a.qf[(3 + k) * a.P + col], 0.0); s = s + y0[k]; }
#pragma unroll
        for (int k = 0; k < NSP; ++k) y0[k] = y0[k] / s;
    } else {
        T0 = a.T[t];
#pragma unroll
        for (int k = 0; k < NSP; ++k) y0[k] = a.Y[long(k) * a.n + t];
    }
    const float *out = a.Yh + long(row) * MLP_OUTP;
    double y1[NSP];
#pragma unroll
    for (int k = 0; k < NSP; ++k) y1[k] = boxcox_inv(boxcox(y0[k]) + (double(out[k]) * nm->out_std[k] + nm->out_mean[k]));
    const double Tnet = T0 + (double(out[NSP]) * nm->out_std[NSP] + nm->out_mean[NSP]);
    bool ok = a.in_range[row] != 0;
    // weighted projection onto element conservation (5 x 5 normal equations)
    constexpr int ME = 8;
    const int mm = a.mm;
    double G[ME][ME], r[ME], el0[ME];
    for (int e = 0; e < mm; ++e) {
        r[e] = 0.0;
        el0[e] = 0.0;
        for (int f = 0; f < mm; ++f) G[e][f] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const double w = fmax(y1[k], 0.0) + 1.0e-12;
        for (int e = 0; e < mm; ++e) {
            const double ae = a.ncf[k * mm + e] / th->wt[k];
            r[e] += ae * (y1[k] - y0[k]);
            el0[e] += ae * y0[k];
            for (int f = 0; f < mm; ++f) G[e][f] += w * ae * (a.ncf[k * mm + f] / th->wt[k]);
        }
    }
    // Gaussian elimination for lambda (elements absent from the mixture are skipped)
    double lam[ME];
    int live[ME];
    for (int e = 0; e < mm; ++e) { live[e] = G[e][e] > 1.0e-300 ? 1 : 0; lam[e] = 0.0; }
    for (int e = 0; e < mm; ++e) {
        if (!live[e]) continue;
        for (int f = e + 1; f < mm; ++f) {
            if (!live[f]) continue;
            const double l = G[f][e] / G[e][e];
            for (int g = e; g < mm; ++g) G[f][g] -= l * G[e][g];
            r[f] -= l * r[e];
        }
    }
    for (int e = mm - 1; e >= 0; --e) {
        if (!live[e]) continue;
        double s = r[e];
        for (int f = e + 1; f < mm; ++f)
            if (live[f]) s -= G[e][f] * lam[f];
        lam[e] = s / G[e][e];
    }
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < NSP; ++k) {
        const double w = fmax(y1[k], 0.0) + 1.0e-12;
        double c = 0.0;
        for (int e = 0; e < mm; ++e) c += lam[e] * (a.ncf[k * mm + e] / th->wt[k]);
        y1[k] = fmax(y1[k] - w * c, 0.0);
        ssum = ssum + y1[k];
    }
    double resid = 0.0;
    for (int e = 0; e < mm; ++e) {
        if (!(el0[e] > 1.0e-300)) continue;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) s += (a.ncf[k * mm + e] / th->wt[k]) * (y1[k] - y0[k]);
        resid = fmax(resid, fabs(s) / el0[e]);
    }
    if (!(resid <= a.resid_tol) || !(fabs(ssum - 1.0) <= 1.0e-3)) ok = false;
    // energy-consistent temperature
    double e0 = 0.0, T1 = Tnet;
    {
        const TPow p = tpow(T0, th->ru);
#pragma unroll
        for (int k = 0; k < NSP; ++k) e0 = e0 + y0[k] * (h_mol(nasa(th, k, T0), p) - p.rut) / th->wt[k];
    }
#pragma unroll
    for (int k = 0; k < NSP; ++k) y1[k] = y1[k] / ssum;
    if (!(T1 > 200.0 && T1 < 6000.0)) ok = false;
    for (int it = 0; it < 20 && ok; ++it) {
        const TPow p = tpow(T1, th->ru);
        double ek = 0.0, cv = 0.0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
            const Nasa c = nasa(th, k, T1);
            ek = ek + y1[k] * (h_mol(c, p) - p.rut) / th->wt[k];
            cv = cv + y1[k] * (cp_mol(c, p) - th->ru) / th->wt[k];
        }
        const double d = (e0 - ek) / cv;
        T1 = T1 + d;
        if (fabs(d) <= 1.0e-12 * T1) break;
    }
    if (!(fabs(T1 - Tnet) <= a.t_tol * T1)) ok = false;
    if (a.route) a.route[t] = ok ? 1 : 0;
    if (!ok) {
        const unsigned q = atomicAdd(a.nfallback, 1u);
        a.fallback[q] = t;
        return;
    }
    if (a.layout == 0) {
#pragma unroll
        for (int k = 0; k < NSP; ++k) a.qf[(3 + k) * a.P + col] = y1[k];
        a.qf[2 * a.P + col] = T1;
    } else {
#pragma unroll
        for (int k = 0; k < NSP; ++k) a.Y[long(k) * a.n + t] = y1[k];
        a.T[t] = T1;
    }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static unsigned blocks(long n, int b) { return unsigned((n + b - 1) / b); }

hipError_t mlp_features(const FeatArgs &f, int rows, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_features, dim3(blocks(rows, 128)), dim3(128), 0, st, f);
    return hipGetLastError();
}

hipError_t mlp_targets(const SurrogateNorm *nm, int kk, int n, int npad, const double *T0, const double *Y0,
                       const double *T1, const double *Y1, float *Tg, hipStream_t st) {
    hipLaunchKernelGGL(k_targets, dim3(blocks(npad, 128)), dim3(128), 0, st, nm, kk, n, npad, T0, Y0, T1, Y1, Tg);
    return hipGetLastError();
}

hipError_t mlp_gather_bf16(const uint16_t *src, uint16_t *dst, const int *idx, int rows, int w, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_rows<uint16_t>, dim3(blocks(long(rows) * w, 256)), dim3(256), 0, st, src, dst, idx, rows, w);
    return hipGetLastError();
}

hipError_t mlp_gather_f32(const float *src, float *dst, const int *idx, int rows, int w, hipStream_t st) {
    hipLaunchKernelGGL(k_gather_rows<float>, dim3(blocks(long(rows) * w, 256)), dim3(256), 0, st, src, dst, idx, rows, w);
    return hipGetLastError();
}

hipError_t mlp_mse(const float *Yh, const float *Yt, uint16_t *dY, int rows, float *partial, float *loss,
                   hipStream_t st) {
    const unsigned nb = blocks(long(rows) * MLP_OUTP, 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, Yh, Yt, dY, rows, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, partial, int(nb), loss);
    return hipGetLastError();
}

hipError_t mlp_reduce_splits(const float *parts, float *out, int nsplit, long len, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_splits, dim3(blocks(len, 256)), dim3(256), 0, st, parts, out, nsplit, len);
    return hipGetLastError();
}

hipError_t mlp_colsum(const uint16_t *D, float *partial, float *out, int rows, int cols, hipStream_t st) {
    const int nchunk = (rows + 63) / 64;
    hipLaunchKernelGGL(k_colsum, dim3(blocks(cols, 64), unsigned(nchunk)), dim3(64), 0, st, D, partial, rows, cols);
    hipLaunchKernelGGL(k_reduce_splits, dim3(blocks(cols, 256)), dim3(256), 0, st, partial, out, nchunk, long(cols));
    return hipGetLastError();
}

hipError_t mlp_zero_u32(unsigned *p, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}

hipError_t mlp_adam(float *w, const float *g, float *m, float *v, uint16_t *w16, long n, float lr, float b1, float b2,
                    float eps, float c1, float c2, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, st, w, g, m, v, w16, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t mlp_to_bf16(const float *w, uint16_t *w16, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, st, w, w16, n);
    return hipGetLastError();
}

hipError_t mlp_apply(const ApplyArgs &a, int nsp, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_apply<N>), dim3(blocks(a.n, 64)), dim3(64), 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
