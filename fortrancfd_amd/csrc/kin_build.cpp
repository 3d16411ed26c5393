// kin_build.cpp -- device kinetics tables from the CHEMKIN-II link-file data (host side).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "chem.hpp"

namespace fcfd {

// Device rows are the mechanism's reactions re-ordered by kind -- fall-off first, then plain
// third-body, then the remaining reactions whose coefficients exceed 3 in magnitude (or whose
// product coefficients are negative), then the rest, by falling count of exponentials -- so a wave,
// which evaluates 64 consecutive rows at once, runs the fall-off and third-body code paths in as
// few row chunks as possible, and the last group (from row nplain0) can take the branch-free
// evaluation.  Every per-reaction result
// is independent of the row order; only the order of the species-sum additions changes.
void build_dev_kin(const Mech &m, DevKin &k) {
    std::memset(&k, 0, sizeof(DevKin));
    k.ii = m.ii;
    k.kk = m.kk;
    k.nthb = int(m.itb.size());
    k.nfal = int(m.ifl.size());
    k.small = std::pow(10.0, -300);        // CKINIT SMALL / EXPARG (cklib.f:4054-4057)
    k.exparg = std::log(std::pow(10.0, 300));
    std::vector<int> tb_of(m.ii, -1), fo_of(m.ii, -1), order;
    for (int t = 0; t < k.nthb; ++t) tb_of[m.itb[t] - 1] = t;
    for (int f = 0; f < k.nfal; ++f) fo_of[m.ifl[f] - 1] = f;
    // the branch-free rows also store their reactants in slots 0 .. nr-1 and products in 3 .. 3+np-1
    auto small = [&](int i) {
        int sum = 0;
        for (int n = 0; n < 6; ++n) {
            const int v = m.nu[size_t(i) * 6 + n];
            if (v > 3 || v < -3 || (n >= 3 && v < 0)) return false;
            sum += v;
        }
        for (int h = 0; h < 2; ++h)
            for (int n = 1; n < 3; ++n)
                if (m.nk[size_t(i) * 6 + 3 * h + n] > 0 && m.nk[size_t(i) * 6 + 3 * h + n - 1] <= 0) return false;
        return sum >= -3 && sum <= 3 && m.nk[size_t(i) * 6] > 0;
    };
    // exponentials a branch-free row evaluates: the Arrhenius factor unless B = E = 0, the
    // equilibrium constant if reversible
    const int np1 = m.npar + 1;
    auto n_exp = [&](int i) {
        const bool arr0 = m.par[size_t(i) * np1 + 1] == 0.0 && m.par[size_t(i) * np1 + 2] == 0.0;
        return (arr0 ? 0 : 1) + (m.ns[i] > 0 ? 1 : 0);
    };
    for (int pass = 0; pass < 6; ++pass) {
        if (pass == 3) k.nplain0 = int(order.size());
        for (int i = 0; i < m.ii; ++i) {
            const int kind = fo_of[i] >= 0 ? 0 : (tb_of[i] >= 0 ? 1 : (small(i) ? 3 : 2));
            // the branch-free rows by falling cost, so that the waves that share a chunk of rows
            // (and its barriers) get rows of equal cost
            if (kind == 3 ? pass == 5 - n_exp(i) : kind == pass) order.push_back(i);
        }
    }
    for (int r = 0; r < m.ii; ++r) {
        const int i = order[r];
        k.orig[r] = i;
        k.A[r] = m.par[size_t(i) * np1 + 0];
        k.B[r] = m.par[size_t(i) * np1 + 1];
        k.E[r] = m.par[size_t(i) * np1 + 2];
        k.rev[r] = m.ns[i] > 0 ? 1 : 0;
        k.arr0[r] = (k.B[r] == 0.0 && k.E[r] == 0.0) ? 1 : 0;
        int s = 0;
        for (int n = 0; n < 6; ++n) {
            k.nu[r][n] = m.nu[size_t(i) * 6 + n];
            k.nk[r][n] = m.nk[size_t(i) * 6 + n] - 1;
            s += k.nu[r][n];
        }
        k.nus[r] = s;
        k.tb[r] = tb_of[i];
        k.fo[r] = fo_of[i];
    }
    // scatter lists of the lane-per-cell kernels, over the device rows
    int ne = 0;
    const int nchunk = (m.ii + LANE_CHUNK - 1) / LANE_CHUNK;
    static_assert(sizeof(k.lane_ent0) / sizeof(int) >= (MAXR + LANE_CHUNK - 1) / LANE_CHUNK * LANE_NW + 1,
                  "lane_ent0 holds one start per chunk and wave");
    for (int c = 0; c < nchunk; ++c)
        for (int w = 0; w < LANE_NW; ++w) {
            k.lane_ent0[c * LANE_NW + w] = ne;
            for (int r = 0; r < LANE_CHUNK; ++r) {
                const int row = c * LANE_CHUNK + r;
                if (row >= m.ii) break;
                for (int n = 0; n < 6; ++n) {
                    const int sp = k.nk[row][n];
                    if (sp >= 0 && sp % LANE_NW == w)
                        k.lane_ent[ne++] = LaneEnt{sp * LANE_NC * 8, r * LANE_NC * 8, double(k.nu[row][n])};
                }
            }
        }
    k.lane_ent0[nchunk * LANE_NW] = ne;
    for (int r = k.nplain0; r < m.ii; ++r) {
        LanePlain &p = k.lane_plain[r];
        p.A = k.A[r];
        p.B = k.B[r];
        p.E = k.E[r];
        int nr = 0, np = 0, pw = 0, big = 0;
        for (int n = 0; n < 6; ++n) {
            const int sp = k.nk[r][n];
            p.nud[n] = double(k.nu[r][n]);
            p.off[n] = (unsigned short)((sp >= 0 ? sp : 0) * LANE_NC * 8);
            if (sp >= 0) (n < 3 ? nr : np) += 1;
            const int a = k.nu[r][n] < 0 ? -k.nu[r][n] : k.nu[r][n];
            pw |= (sp >= 0 ? a : 0) << (2 * n);
            if (sp >= 0 && a > 1) big = 1;
        }
        p.bits = (nr << LP_NR) | (np << LP_NP) | (k.rev[r] ? LP_REV : 0) | (k.arr0[r] ? LP_ARR0 : 0) |
                 ((k.nus[r] + 3) << LP_NUS) | (big ? LP_POW : 0) | (pw << LP_PW) | (k.nu[r][0] == 0 ? LP_DEAD : 0);
    }
    for (int t = 0; t < k.nthb; ++t) {
        k.ntbs[t] = m.ntbs[t];
        for (int l = 0; l < m.mxtb && l < MAXTB; ++l) {
            k.tbk[t][l] = m.nktb[size_t(t) * m.mxtb + l] - 1;
            k.tba[t][l] = m.aik[size_t(t) * m.mxtb + l];
        }
    }
    for (int f = 0; f < k.nfal; ++f) {
        k.ifop[f] = m.ifo[f];
        k.kfal[f] = m.kfl[f] - 1;
        for (int c = 0; c < 8 && c < m.nfar; ++c) k.fpar[f][c] = m.fpar[size_t(f) * m.nfar + c];
    }
}

// CKHMS(298.15) (cklib.f:3598-3618): formation enthalpies for the heat-release diagnostics.
void thermo_h298(DevThermo &t) {
    const double T = 298.15, rut = T * t.ru;
    const double t2 = T * T, t3 = T * t2, t4 = t2 * t2;
    const double tn1 = T / 2.0, tn2 = t2 / 3.0, tn3 = t3 / 4.0, tn4 = t4 / 5.0;
    for (int k = 0; k < t.kk; ++k) {
        const double *a = T > t.tmid[k] ? t.ahi[k] : t.alo[k];
        double s = a[0];
        s = s + tn1 * a[1];
        s = s + tn2 * a[2];
        s = s + tn3 * a[3];
        s = s + tn4 * a[4];
        t.h298[k] = rut * (s + a[5] / T) / t.wt[k];
    }
}

}  // namespace fcfd
