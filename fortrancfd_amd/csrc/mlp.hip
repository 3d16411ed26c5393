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
    if (a.m_dev && a.row_live) {
        // every wave reads the same 64 flags, so all four leave together, before any barrier
        const int r = a.m_off + m0 + lane;
        if (!__any(r < int(*a.m_dev) && a.row_live[r] != 0)) return;
    }
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
// Wave helpers (one cell per wave, lanes over species / feature columns).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double wsum64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Column c of the mass fractions of a cell: the primitive state's Y renormalised by adjust_sum
// (layout 0, as solve_source sees it) or the batch arrays as given (layout 1).
template <class Args>
__device__ __forceinline__ void cell_state(const Args &f, int cell, int lane, int kk, double &y, double &rho,
                                           double &T, long &col) {
    y = 0.0;
    if (f.layout == 0) {
        col = long(cell / f.ncell) * f.S + f.lv + (cell % f.ncell);
        if (lane < kk) y = fmax(f.qf[(3 + lane) * f.P + col], 0.0);
        y = y / wsum64(y);
        rho = f.qf[col];
        T = f.qf[2 * f.P + col];
    } else {
        col = 0;
        if (lane < kk) y = f.Y[long(lane) * f.n + cell];
        rho = f.rho[cell];
        T = f.T[cell];
    }
}

// ---------------------------------------------------------------------------------------------
// Features: x = (b(Y_k), T, log(p/1 atm), log dt) z-scored (b: Box-Cox above), padded to 64
// columns; in_range = every raw feature inside the training range [lo, hi].  ONE ROW PER WAVE,
// lane = feature column: rows 0 .. npad of all cells (list == null; rows >= n zero-filled) or rows
// 0 .. *list_n for the cells list[row].
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_features(FeatArgs f) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    int cell = t;
    if (f.list) {
        if (t >= int(*f.list_n)) return;
        cell = f.list[t];
    } else {
        if (t >= f.npad) return;
        if (t >= f.n) {
            f.X[long(t) * MLP_INP + lane] = 0;
            if (f.in_range && lane == 0) f.in_range[t] = 0;
            return;
        }
    }
    const int kk = f.kk;
    const SurrogateNorm *__restrict__ nm = f.norm;
    double y, rho, T;
    long col;
    cell_state(f, cell, lane, kk, y, rho, T, col);
    const double dt = f.layout == 0 ? (f.dt_dev ? *f.dt_dev : f.dt) : f.dtv[cell];
    const double swy = wsum64(lane < kk ? y / f.thermo->wt[lane] : 0.0);
    const double p = rho * 8.31451 * T * swy * 1.0e3;          // Pa (ideal gas, SI)
    double raw = 0.0;
    if (lane < kk) raw = boxcox(y);
    else if (lane == kk) raw = T;
    else if (lane == kk + 1) raw = log(p / 101325.0);
    else if (lane == kk + 2) raw = log(dt);
    const bool feat = lane < kk + 3;
    const bool ok_l = !feat || (raw >= nm->lo[lane] && raw <= nm->hi[lane]);
    f.X[long(t) * MLP_INP + lane] = feat ? f2bf(float((raw - nm->in_mean[lane]) / nm->in_std[lane])) : uint16_t(0);
    const bool ok = __all(ok_l) && rho > 0.0 && T > 0.0 && dt > 0.0;
    if (f.in_range && lane == 0) f.in_range[t] = ok ? 1 : 0;
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
// Inference epilogue, ONE CELL PER WAVE (lane = species).  Candidate Y1 = b^-1(b(Y0) + db)
// (network), then the element balance is restored by the weighted projection minimising
// sum_k dY_k^2 / w_k subject to sum_k a_ek (Y1_k - Y0_k) = 0 (a_ek = atoms of element e per unit
// mass of species k, w_k = max(Y1_k, 0) + 1e-12; MM x MM normal equations, elements absent from
// the mixture left out), negative entries are clipped, and T1 follows from conservation of
// sum_k Y_k u_k(T) / W_k (Newton from the network's T).  A candidate when the inputs were inside
// the training range, the element residual after projection and clipping is <= resid_tol, and the
// network temperature agrees with the energy temperature to t_tol: it goes to a candidate buffer
// and list for the FP64 verification (chem_lane.hip); otherwise the cell is appended to the
// fallback list for the FP64 integrator.
// ---------------------------------------------------------------------------------------------
template <int NSP, int MM>
__global__ void __launch_bounds__(256) k_apply(ApplyArgs a) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);        // network row
    int t = row;                                                 // cell
    if (a.list) {
        if (row >= int(*a.list_n)) return;
        t = a.list[row];
    } else if (row >= a.n) {
        return;
    }
    if (a.in_range[row] == 0) {     // outside the training range: the network skipped this row
        if (lane == 0) {
            if (a.route) a.route[t] = 0;
            const unsigned q = atomicAdd(a.nfallback, 1u);
            a.fallback[q] = t;
        }
        return;
    }
    const DevThermo *__restrict__ th = a.thermo;
    const SurrogateNorm *__restrict__ nm = a.norm;
    const bool sp = lane < NSP;
    double y0, rho, T0;
    long col;
    cell_state(a, t, lane, NSP, y0, rho, T0, col);
    (void)rho;
    const float *out = a.Yh + long(row) * MLP_OUTP;
    double y1 = sp ? boxcox_inv(boxcox(y0) + (double(out[lane]) * nm->out_std[lane] + nm->out_mean[lane])) : 0.0;
    const double Tnet = T0 + (double(out[NSP]) * nm->out_std[NSP] + nm->out_mean[NSP]);
    bool ok = true;
    const double iw = sp ? 1.0 / th->wt[lane] : 0.0;
    double ae[MM];
#pragma unroll
    for (int e = 0; e < MM; ++e) ae[e] = sp ? a.ncf[lane * MM + e] * iw : 0.0;
    const double w = sp ? fmax(y1, 0.0) + 1.0e-12 : 0.0;
    double G[MM][MM], r[MM], el0[MM];
#pragma unroll
    for (int e = 0; e < MM; ++e) {
        r[e] = wsum64(ae[e] * (y1 - y0));
        el0[e] = wsum64(ae[e] * y0);
#pragma unroll
        for (int f = e; f < MM; ++f) G[e][f] = wsum64(w * ae[e] * ae[f]);
#pragma unroll
        for (int f = 0; f < e; ++f) G[e][f] = G[f][e];
    }
    double lam[MM];
    bool live[MM];
#pragma unroll
    for (int e = 0; e < MM; ++e) { live[e] = G[e][e] > 1.0e-300; lam[e] = 0.0; }
#pragma unroll
    for (int e = 0; e < MM; ++e) {
#pragma unroll
        for (int f = e + 1; f < MM; ++f) {
            if (!live[e] || !live[f]) continue;
            const double l = G[f][e] / G[e][e];
#pragma unroll
            for (int g = e; g < MM; ++g) G[f][g] -= l * G[e][g];
            r[f] -= l * r[e];
        }
    }
#pragma unroll
    for (int e = MM - 1; e >= 0; --e) {
        if (!live[e]) continue;
        double acc = r[e];
#pragma unroll
        for (int f = e + 1; f < MM; ++f)
            if (live[f]) acc -= G[e][f] * lam[f];
        lam[e] = acc / G[e][e];
    }
    double c = 0.0;
#pragma unroll
    for (int e = 0; e < MM; ++e) c += lam[e] * ae[e];
    y1 = sp ? fmax(y1 - w * c, 0.0) : 0.0;
    const double ssum = wsum64(y1);
    double resid = 0.0;
#pragma unroll
    for (int e = 0; e < MM; ++e) {
        const double s = wsum64(ae[e] * (y1 - y0));
        if (el0[e] > 1.0e-300) resid = fmax(resid, fabs(s) / el0[e]);
    }
    if (!(resid <= a.resid_tol) || !(fabs(ssum - 1.0) <= 1.0e-3)) ok = false;
    // energy-consistent temperature
    double e0;
    {
        const TPow p = tpow(T0, th->ru);
        e0 = wsum64(sp ? y0 * (h_mol(nasa(th, lane, T0), p) - p.rut) * iw : 0.0);
    }
    y1 = y1 / ssum;
    double T1 = Tnet;
    if (!(T1 > 200.0 && T1 < 6000.0)) ok = false;
    for (int it = 0; it < 20 && ok; ++it) {
        const TPow p = tpow(T1, th->ru);
        double ek = 0.0, cv = 0.0;
        if (sp) {
            const Nasa cf = nasa(th, lane, T1);
            ek = y1 * (h_mol(cf, p) - p.rut) * iw;
            cv = y1 * (cp_mol(cf, p) - th->ru) * iw;
        }
        ek = wsum64(ek);
        cv = wsum64(cv);
        const double d = (e0 - ek) / cv;
        T1 = T1 + d;
        if (fabs(d) <= 1.0e-12 * T1) break;
    }
    if (!(fabs(T1 - Tnet) <= a.t_tol * T1)) ok = false;
    if (lane == 0 && a.route) a.route[t] = 0;
    if (!ok) {
        if (lane == 0) {
            const unsigned q = atomicAdd(a.nfallback, 1u);
            a.fallback[q] = t;
        }
        return;
    }
    // candidate: verified against the FP64 right-hand side before it may enter the state
    unsigned q = 0;
    if (lane == 0) q = atomicAdd(a.ncand, 1u);
    q = __shfl(q, 0, 64);
    if (sp) a.cand[long(q) * (NSP + 1) + lane] = y1;
    if (lane == 0) {
        a.cand[long(q) * (NSP + 1) + NSP] = T1;
        a.cand_list[q] = t;
    }
}

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
static unsigned blocks(long n, int b) { return unsigned((n + b - 1) / b); }

hipError_t mlp_features(const FeatArgs &f, int rows, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    if (f.kk + 3 > MLP_INP) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_features, dim3(blocks(rows, 4)), dim3(256), 0, st, f);
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

template <int NSP>
static hipError_t apply_mm(const ApplyArgs &a, hipStream_t st) {
    const dim3 g(blocks(a.n, 4)), b(256);
    switch (a.mm) {
        case 4: hipLaunchKernelGGL((k_apply<NSP, 4>), g, b, 0, st, a); break;
        case 5: hipLaunchKernelGGL((k_apply<NSP, 5>), g, b, 0, st, a); break;
        case 6: hipLaunchKernelGGL((k_apply<NSP, 6>), g, b, 0, st, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t mlp_apply(const ApplyArgs &a, int nsp, hipStream_t st) {
    if (a.n <= 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) return apply_mm<N>(a, st);
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
