// chem.hip -- chemistry source term: the reaction_macks call site (SRC/source.f90:17-84) and the
// stiff integrator behind macks(dt, nspe+1, yiT, rho) (SRC/source.f90:58).  The reference does not
// ship its integrator (module ODEs), so the method is this project's own:
//   state y = (Y_1..Y_KK, T) at fixed density, f from the CKWYR rates (cklib.f:9107-9204),
//   3-stage L-stable SDIRK of order 3 (Alexander 1977, gamma = 0.43586652150845900) with an
//   embedded order-2 solution, simplified Newton per stage with M = I - h*gamma*J (LU with partial
//   pivoting), filtered error M^{-1} h sum (b - b^)_i k_i, analytic mass-action Jacobian (third-body
//   and fall-off factors held fixed) with a finite-difference temperature column, and a final
//   temperature correction that keeps sum_k Y_k u_k(T) / W_k at its initial value.
//
// GPU mapping: ONE CELL PER WAVEFRONT.  A persistent grid pulls cells from a device work queue, so
// cells of very different stiffness (1 step vs. hundreds) balance themselves.  Inside a wave,
// lanes run over reactions (rates, Jacobian scatter) and species (thermo, production rates); the
// 54x54 Newton matrix lives one row per lane in VGPRs and is factored by pivoted elimination
// across lanes; vectors of the state are distributed one component per lane.
#include "chem.hpp"
#include "chem_device.hpp"
#include "fcfd_device.hpp"

namespace fcfd {

namespace {

constexpr double CH_G = 0.435866521508459;
constexpr double CH_A21 = 0.2820667392457705;
constexpr double CH_B1 = 1.20849664917601;
constexpr double CH_B2 = -0.6443631706844692;
constexpr double CH_E1 = 0.43586652150845895;
constexpr double CH_E2 = -0.8717330430169181;
constexpr double CH_E3 = 0.435866521508459;
constexpr double CH_KAPPA = 0.03;
constexpr int CH_MAXIT = 7;
constexpr int CH_MAXSTEPS = 5000;
constexpr int LDJ = 65;   // LDS row stride of the Jacobian (odd in 64-bit words)

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// Lane holding the largest value; ties go to the lower lane (the CPU restatement's order).
__device__ __forceinline__ int wave_argmax(double v) {
    int bi = lane_id();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > v || (ov == v && oi < bi)) { v = ov; bi = oi; }
    }
    return __builtin_amdgcn_readfirstlane(bi);
}

template <int NSP>
struct Shared {
    double J[(NSP + 1) * LDJ];   // Jacobian, overwritten by the LU factors of M
    double c[64], smh[64], u[64], prow[64], wdot[64], cdot[64], ddot[64];
    int perm[64];
    int cell;
};

struct RhsOut {
    double f;      // this lane's component of dy/dt
    double cvm;    // mixture cv, erg/(g K)
    double uk;     // this lane's species internal energy, erg/mol (0 on other lanes)
    double cvk;    // this lane's species cv, erg/(mol K)
};

// f = dy/dt at (y_l, T); y_l is this lane's component (lane < NSP: Y_lane).  Leaves the
// concentrations in sh.c, S/R-H/RT in sh.smh, u_k in sh.u and the production rates in sh.wdot, or,
// with split, CKCDYR's creation / destruction rates in sh.cdot / sh.ddot (cklib.f:1497-1517).
// Rates of progress are scattered into the species sums with LDS atomics; one wave owns the
// workgroup, so the order of those additions is fixed and the result reproducible.
template <int NSP>
__device__ __forceinline__ RhsOut rhs(const DevThermo *__restrict__ th, const DevKin *__restrict__ K, Shared<NSP> &sh,
                      double rho, double y_l, double T, bool split) {
    const int lane = lane_id();
    double uk = 0.0, cpk = 0.0, ck = 0.0;
    if (lane < NSP) {
        ck = rho * y_l / th->wt[lane];
        sh.c[lane] = ck;
        const Nasa a = nasa(th, lane, T);
        const TPow p = tpow(T, th->ru);
        uk = h_mol(a, p) - p.rut;
        cpk = cp_mol(a, p);
        sh.u[lane] = uk;
        sh.smh[lane] = smh_species(a, T);
    }
    sh.wdot[lane] = 0.0;
    sh.cdot[lane] = 0.0;
    sh.ddot[lane] = 0.0;
    const double ctot = wsum(ck);
    __syncthreads();
    const double alogt = log(T), pfac = th->pa / (th->ru * T);
    for (int i = lane; i < K->ii; i += 64) {
        const Rop r = reaction<1>(K, i, T, alogt, pfac, ctot, sh.c, sh.smh);
        const int *nu = K->nu[i], *nk = K->nk[i];
        if (!split) {
            const double q = r.qf - r.qr;
            for (int n = 0; n < 6; ++n)
                if (nk[n] >= 0) atomicAdd(&sh.wdot[nk[n]], q * (double)nu[n]);
        } else {
            for (int n = 0; n < 3; ++n) {
                if (nk[n] >= 0) {
                    const double a = (double)abs(nu[n]);
                    atomicAdd(&sh.cdot[nk[n]], r.qr * a);
                    atomicAdd(&sh.ddot[nk[n]], r.qf * a);
                }
                if (nk[n + 3] >= 0) {
                    const double a = (double)nu[n + 3];
                    atomicAdd(&sh.cdot[nk[n + 3]], r.qf * a);
                    atomicAdd(&sh.ddot[nk[n + 3]], r.qr * a);
                }
            }
        }
    }
    __syncthreads();
    const double wdot = lane < NSP ? sh.wdot[lane] : 0.0;
    const double cvk = cpk - th->ru;
    const double cvm = wsum(lane < NSP ? y_l * cvk / th->wt[lane] : 0.0);
    const double su = wsum(lane < NSP ? uk * wdot : 0.0);
    RhsOut o;
    o.cvm = cvm;
    o.uk = uk;
    o.cvk = cvk;
    o.f = lane < NSP ? wdot * th->wt[lane] / rho : (lane == NSP ? -su / (rho * cvm) : 0.0);
    return o;
}

// Jacobian df/dy into sh.J (rows/cols 0..NSP, NSP = temperature).
template <int NSP>
__device__ __forceinline__ void jacobian(const DevThermo *__restrict__ th, const DevKin *__restrict__ K, Shared<NSP> &sh,
                         double rho, double y_l, double T) {
    const int lane = lane_id();
    constexpr int N = NSP + 1;
    const RhsOut f0 = rhs<NSP>(th, K, sh, rho, y_l, T, false);
    for (int e = lane; e < N * LDJ; e += 64) sh.J[e] = 0.0;
    const double ctot = wsum(lane < NSP ? sh.c[lane] : 0.0);
    __syncthreads();
    const double alogt = log(T), pfac = th->pa / (th->ru * T);
    for (int i = lane; i < K->ii; i += 64) {
        const int *nu = K->nu[i], *nk = K->nk[i];
        if (nu[0] == 0) continue;
        const Rop r = reaction<1>(K, i, T, alogt, pfac, ctot, sh.c, sh.smh);
        double dq[6];
        for (int s = 0; s < 6; ++s) {
            dq[s] = 0.0;
            if (nk[s] < 0) continue;
            const int lo = s < 3 ? 0 : 3;
            const int a = abs(nu[s]);
            double d = (double)a * ipow(sh.c[nk[s]], a - 1);
            for (int t = lo; t < lo + 3; ++t)
                if (t != s && nk[t] >= 0) d = d * ipow(sh.c[nk[t]], abs(nu[t]));
            dq[s] = s < 3 ? r.kf * d : -(r.kr * d);
        }
        for (int s = 0; s < 6; ++s) {
            if (nk[s] < 0) continue;
            for (int t = 0; t < 6; ++t)
                if (nk[t] >= 0) atomicAdd(&sh.J[nk[t] * LDJ + nk[s]], (double)nu[t] * dq[s]);
        }
    }
    __syncthreads();
    // temperature row from the concentration Jacobian, before rows are rescaled to Y
    double trow = 0.0;
    if (lane < NSP) {
        double s = 0.0;
        for (int j = 0; j < NSP; ++j) s = s + sh.u[j] * sh.J[j * LDJ + lane];
        s = s * rho / th->wt[lane];
        const double fT = bcast(f0.f, NSP);
        trow = -s / (rho * f0.cvm) - fT * (f0.cvk / th->wt[lane]) / f0.cvm;
    }
    __syncthreads();
    if (lane < NSP) {
        const double wr = th->wt[lane];
        for (int k = 0; k < NSP; ++k) sh.J[lane * LDJ + k] = sh.J[lane * LDJ + k] * wr / th->wt[k];
        sh.J[NSP * LDJ + lane] = trow;
    }
    // temperature column by a forward difference
    const double dT = T * 1.0e-6;
    const RhsOut f1 = rhs<NSP>(th, K, sh, rho, y_l, T + dT, false);
    if (lane < N) sh.J[lane * LDJ + NSP] = (f1.f - f0.f) / dT;
    __syncthreads();
}

// LU with partial pivoting of M = I - hg*J: rows are factored one per lane in VGPRs and the factors
// written back over sh.J; rank = elimination step of this lane's row (N: not pivoted yet, -1: lane
// beyond the system).  Returns false if a pivot is zero.
template <int NSP>
__device__ __forceinline__ bool lu(Shared<NSP> &sh, double hg, int &rank) {
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    rank = lane < N ? N : -1;
    double m[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        m[j] = 0.0;
        if (lane < N) m[j] = -hg * sh.J[lane * LDJ + j];
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        if (lane == j) m[j] = m[j] + 1.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        const int p = wave_argmax(rank == N ? fabs(m[k]) : -1.0);
        if (lane == p) {
            rank = k;
#pragma unroll
            for (int j = k; j < N; ++j) sh.prow[j] = m[j];
        }
        if (lane == 0) sh.perm[k] = p;
        __syncthreads();
        const double piv = sh.prow[k];
        if (!(piv != 0.0)) ok = false;
        if (rank == N) {
            const double l = m[k] / piv;
            m[k] = l;
#pragma unroll
            for (int j = k + 1; j < N; ++j) m[j] = m[j] - l * sh.prow[j];
        }
        __syncthreads();
    }
    if (lane < N) {
#pragma unroll
        for (int j = 0; j < N; ++j) sh.J[lane * LDJ + j] = m[j];
    }
    __syncthreads();
    return ok;
}

// Solves M x = b with the factors in sh.J; b and x distributed one component per lane.
template <int NSP>
__device__ __forceinline__ double solve(const Shared<NSP> &sh, int rank, double b) {
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    const double *row = sh.J + (lane < N ? lane : 0) * LDJ;
    for (int k = 0; k < N; ++k) {
        const double bp = bcast(b, sh.perm[k]);
        if (rank > k) b = b - row[k] * bp;
    }
    double x = 0.0;
    for (int k = N - 1; k >= 0; --k) {
        const int p = sh.perm[k];
        const double xk = bcast(b / row[k], p);
        if (lane == k) x = xk;
        if (rank >= 0 && rank < k) b = b - row[k] * xk;
    }
    return x;
}

__device__ __forceinline__ double wnorm(double v, double w, int n) {
    const double t = v * w;
    return sqrt(wsum(t * t) / n);
}

// Mixture e = sum Y u / W and cv = sum Y cv / W at T (wave sums); y_l this lane's Y.
template <int NSP>
__device__ __forceinline__ void energy(const DevThermo *__restrict__ th, double y_l, double T, double &e, double &cv) {
    const int lane = lane_id();
    double eu = 0.0, ec = 0.0;
    if (lane < NSP) {
        const Nasa a = nasa(th, lane, T);
        const TPow p = tpow(T, th->ru);
        eu = y_l * (h_mol(a, p) - p.rut) / th->wt[lane];
        ec = y_l * (cp_mol(a, p) - th->ru) / th->wt[lane];
    }
    e = wsum(eu);
    cv = wsum(ec);
}

__device__ __forceinline__ double adjusted(double y_l, int lane, int nsp) {
    const double v = lane < nsp ? fmax(y_l, 0.0) : 0.0;
    return v / wsum(v);
}

template <int NSP>
__global__ void __launch_bounds__(64) k_chem(ChemArgs a) {
    __shared__ Shared<NSP> sh;
    constexpr int N = NSP + 1;
    const int lane = lane_id();
    const DevThermo *__restrict__ th = a.thermo;
    const DevKin *__restrict__ K = a.kin;
    for (;;) {
        if (lane == 0) sh.cell = int(atomicAdd(a.counter, 1u));
        __syncthreads();
        const int item = sh.cell;
        __syncthreads();
        const int nitems = a.list_n ? int(*a.list_n) : a.ncells;
        if (item >= nitems) break;
        const int cell = a.list ? a.list[item] : item;

        // ---- load (reaction_macks, SRC/source.f90:44-56) ----
        long col = 0;
        double rho_si, dt, T0, y_raw = 0.0;
        if (a.layout == 0) {
            const int b = cell / a.ncell, i = cell % a.ncell;
            col = long(b) * a.S + a.lv + i;
            rho_si = a.qf[col];
            T0 = a.qf[2 * a.P + col];
            dt = a.dt_dev ? *a.dt_dev : a.dt;
            if (lane < NSP) y_raw = a.qf[(3 + lane) * a.P + col];
        } else {
            rho_si = a.b_rho[cell];
            dt = a.b_dt[cell];
            T0 = a.b_T[cell];
            if (lane < NSP) y_raw = a.b_y[long(lane) * a.ncells + cell];
        }
        const double rho = rho_si * 1.e-3;
        double y = a.layout == 0 ? adjusted(y_raw, lane, NSP) : y_raw;   // adjust_sum(y_sp)
        double e0, cv0;
        energy<NSP>(th, adjusted(y, lane, NSP), T0, e0, cv0);
        if (lane == NSP) y = T0;

        double t = 0.0, h = dt;
        int steps = 0, rejects = 0, newton = 0;
        int rank = -1;
        // ---- explicit pre-step: Heun over the whole dt, accepted when the embedded Euler error
        //      estimate is within tolerance (quiescent cells never build a Jacobian) ----
        if (!a.skip_explicit) {
            const double w = lane < NSP ? 1.0 / (a.atol + a.rtol * fabs(y))
                                        : (lane == NSP ? 1.0 / (a.rtol * fabs(y)) : 0.0);
            const RhsOut f0 = rhs<NSP>(th, K, sh, rho, y, bcast(y, NSP), false);
            const double y1 = lane < N ? y + dt * f0.f : 0.0;
            const RhsOut f1 = rhs<NSP>(th, K, sh, rho, y1, bcast(y1, NSP), false);
            const double err = wnorm(lane < N ? (dt * 0.5) * (f1.f - f0.f) : 0.0, w, N);
            if (err <= 1.0) {
                if (lane < N) y = y + (dt * 0.5) * (f0.f + f1.f);
                t = dt;
                steps = 1;
            }
        }
        // ---- SDIRK3 over [0, dt] ----
        while (t < dt && steps + rejects < CH_MAXSTEPS) {
            if (h > dt - t) h = dt - t;
            const double Tn = bcast(y, NSP);
            const double w = lane < NSP ? 1.0 / (a.atol + a.rtol * fabs(y))
                                        : (lane == NSP ? 1.0 / (a.rtol * fabs(y)) : 0.0);
            bool have_j = false;
            for (;;) {
                if (!have_j) jacobian<NSP>(th, K, sh, rho, y, Tn);   // lu() overwrites it
                have_j = false;
                const double hg = h * CH_G;
                bool ok = lu<NSP>(sh, hg, rank);
                double k0 = 0.0, k1 = 0.0, k2 = 0.0, z = 0.0;
                for (int st = 0; st < 3 && ok; ++st) {
                    double base = y;
                    if (st == 1) base = base + h * CH_A21 * k0;
                    if (st == 2) base = base + h * CH_B1 * k0 + h * CH_B2 * k1;
                    z = st == 0 ? base : base + hg * (st == 1 ? k0 : k1);
                    double prev = 0.0;
                    bool conv = false;
                    for (int it = 0; it < CH_MAXIT; ++it) {
                        const double Tz = bcast(z, NSP);
                        const RhsOut f = rhs<NSP>(th, K, sh, rho, z, Tz, false);
                        double r = lane < N ? -(z - base - hg * f.f) : 0.0;
                        r = solve<NSP>(sh, rank, r);
                        if (lane < N) z = z + r;
                        ++newton;
                        const double nrm = wnorm(lane < N ? r : 0.0, w, N);
                        if (!(nrm == nrm) || (it > 0 && nrm > 0.9 * prev)) break;
                        if (nrm <= CH_KAPPA) { conv = true; break; }
                        prev = nrm;
                    }
                    if (!conv) { ok = false; break; }
                    const double kv = (z - base) / hg;
                    if (st == 0) k0 = kv; else if (st == 1) k1 = kv; else k2 = kv;
                }
                if (!ok) {
                    ++rejects;
                    h = h * 0.25;
                    if (steps + rejects >= CH_MAXSTEPS) break;
                    continue;
                }
                double ev = lane < N ? h * (CH_E1 * k0 + CH_E2 * k1 + CH_E3 * k2) : 0.0;
                ev = solve<NSP>(sh, rank, ev);
                const double err = wnorm(lane < N ? ev : 0.0, w, N);
                const double fac = 0.9 * pow(fmax(err, 1.0e-10), -1.0 / 3.0);
                if (err <= 1.0) {
                    if (lane < N) y = z;
                    t = t + h;
                    ++steps;
                    h = h * fmin(4.0, fmax(0.2, fac));
                    break;
                }
                ++rejects;
                h = h * fmax(0.2, fac);
                if (steps + rejects >= CH_MAXSTEPS) break;
            }
        }
        const int failed = t < dt ? 1 : 0;

        // ---- energy-consistent temperature ----
        const double ya = adjusted(lane < NSP ? y : 0.0, lane, NSP);
        double T = bcast(y, NSP);
        for (int it = 0; it < 20; ++it) {
            double ek, cvk;
            energy<NSP>(th, ya, T, ek, cvk);
            const double d = (e0 - ek) / cvk;
            T = T + d;
            if (fabs(d) <= 1.0e-12 * T) break;
        }

        // ---- store (SRC/source.f90:60-73) ----
        if (a.layout == 0) {
            if (lane < NSP) a.qf[(3 + lane) * a.P + col] = ya;
            if (lane == 0) a.qf[2 * a.P + col] = T;
            if (a.hrr) {
                // heat_release_rate (SRC/thermo.f90:98-101)
                const double term = lane < NSP ? rho * th->h298[lane] * (ya - y_raw) / dt : 0.0;
                const double s1 = wsum(term);
                // heat_release_rate_omegadot via CKCDYR at the new state (SRC/thermo.f90:128-132)
                rhs<NSP>(th, K, sh, rho, ya, T, true);
                const double cd = lane < NSP ? sh.cdot[lane] : 0.0;
                const double dd = lane < NSP ? sh.ddot[lane] : 0.0;
                const double s2 = wsum(lane < NSP ? th->wt[lane] * th->h298[lane] * (cd - dd) : 0.0);
                if (lane == 0) {
                    a.hrr[cell] = -s1 * 1.e-1;
                    a.hrr_dot[cell] = -s2 * 1.e-1;
                }
            }
        } else {
            if (lane < NSP) a.b_y[long(lane) * a.ncells + cell] = y;
            if (lane == 0) a.b_T[cell] = T;
        }
        if (a.stats && lane == 0) a.stats[cell] = ChemCellStats{steps, rejects, newton, failed};
    }
}

__global__ void k_zero_counter(unsigned int *c) {
    if (threadIdx.x == 0) *c = 0u;
}

}  // namespace

hipError_t launch_zero_u32(unsigned *p, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, p);
    return hipGetLastError();
}

hipError_t launch_chem_full(ChemArgs a, int nsp, int n_workgroups, int *list, unsigned *list_n, hipStream_t st) {
    hipError_t e = launch_zero_u32(list_n, st);
    if (e != hipSuccess) return e;
    a.stiff = list;
    a.stiff_n = list_n;
    a.list = nullptr;
    a.list_n = nullptr;
    if ((e = launch_chem_screen(a, nsp, st)) != hipSuccess) return e;
    a.list = list;
    a.list_n = list_n;
    a.skip_explicit = 1;
    return launch_chem(a, nsp, n_workgroups, st);
}

hipError_t launch_chem(const ChemArgs &a, int nsp, int n_workgroups, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, a.counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nwg = n_workgroups < a.ncells ? n_workgroups : a.ncells;
    if (nwg <= 0) return hipSuccess;
#define FCFD_CASE(N) if (nsp == N) { hipLaunchKernelGGL((k_chem<N>), dim3(nwg), dim3(64), 0, st, a); return hipGetLastError(); }
    FCFD_FOR_EACH_NSP(FCFD_CASE)
#undef FCFD_CASE
    return hipErrorInvalidValue;
}

}  // namespace fcfd
