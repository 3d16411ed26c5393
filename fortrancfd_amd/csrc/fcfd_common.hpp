// fcfd_common.hpp -- shared host/device definitions of libfcfd_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace fcfd {

// Species capacity of the device tables.  Kernels are instantiated per species count
// (FCFD_FOR_EACH_NSP: the 53 of GRI-3.0 and the 11 of the Li-Dryer H2/O2 mechanism) so that per-lane
// species arrays stay in registers.
constexpr int MAXK = 64;
constexpr int MAXR = 512;   // reactions
constexpr int MAXTB = 16;   // enhanced third bodies per reaction

#define FCFD_FOR_EACH_NSP(X) X(53) X(11)

// NASA-7 thermo tables in the order CKHML / CKCPML / CKSMH consume them (cklib.f:3537-3556,
// 1849-1865, 7540-7560): a[0..4] polynomial, a[5] enthalpy constant, a[6] entropy constant.
// Range selection is "T > tmid -> high" (cklib.f:3544-3548 with three temperatures).
struct DevThermo {
    int kk;
    double ru;              // erg/(mol K), CKINIT parameter RU (cklib.f:4050)
    double pa;              // dyn/cm^2
    double wt[MAXK];        // g/mol
    double tmid[MAXK];
    double alo[MAXK][7];
    double ahi[MAXK][7];
    double h298[MAXK];      // CKHMS(298.15) erg/g, for heat-release diagnostics
};

// Host-side mechanism read from the CHEMKIN-II link files.
struct Mech {
    int kk = 0, ii = 0, mm = 0, mxsp = 6, mxtb = 0, mxtp = 3, ncp = 5, npar = 3, nfar = 8;
    double ru = 8.314510e7, pa = 1.01325e6;
    std::vector<std::string> names;
    std::vector<int> ncf;                       // [kk*mm] element counts
    std::vector<double> wt, tt, aa;             // [kk], [kk*mxtp], [kk*(mxtp-1)*(ncp+2)]
    std::vector<int> nt;
    std::vector<int> ns, nr, nu, nk;            // [ii], [ii], [ii*mxsp], [ii*mxsp]
    std::vector<double> par;                    // [ii*(npar+1)]
    std::vector<int> ifl, ifo, kfl;             // [nfal]
    std::vector<double> fpar;                   // [nfal*nfar]
    std::vector<int> itb, ntbs, nktb;           // [nthb], [nthb], [nthb*mxtb]
    std::vector<double> aik;                    // [nthb*mxtb]
    // transport (tran.bin)
    bool has_tran = false;
    int no = 0;
    double patmos = 1.01325e6;
    std::vector<double> twt, coflam, cofeta, cofd;   // [kk], [kk*no], [kk*no], [kk*kk*no]
    DevThermo thermo{};                              // filled by build_device_tables
};

int load_mech(const char *chem_bin, const char *tran_bin, Mech &m, std::string &err);
// The same tables from the mechanism text (chem.inp, optionally therm.dat), mech_text.cpp.
int load_mech_text(const char *chem_inp, const char *therm_dat, Mech &m, std::string &err);
// TRANLIB fits of a tran.bin link file for the species of m (mech_io.cpp).
int load_tran_bin(const char *tran_bin, Mech &m, std::string &err);
void thermo_h298(DevThermo &t);

void set_error(const std::string &msg);

}  // namespace fcfd

struct fcfd_mech {
    fcfd::Mech host;
    // device copies of the tables, created on first use on the current device
    int device = -1;
    int ncu = 0;
    fcfd::DevThermo *d_thermo = nullptr;
    void *d_kin = nullptr;               // fcfd::DevKin
    unsigned int *d_counter = nullptr;   // work queue of batched chemistry calls
    int *d_list = nullptr;               // stiff-cell list of batched chemistry calls (grown on demand)
    unsigned *d_list_n = nullptr;
    size_t list_cap = 0;
};
