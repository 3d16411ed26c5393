// mech_io.cpp -- reader of the CHEMKIN-II link files (chem.bin / tran.bin).
//
// The files are Fortran sequential unformatted records (int32 byte count, payload, int32 byte
// count).  Record contents follow the READ lists of CKINIT (cklib.f:4148-4203) and MCINIT
// (tranlib.f:214-312) in the reference's vendored CHEMKIN-II sources.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

#include "fcfd_common.hpp"

namespace fcfd {
namespace {

class RecordReader {
  public:
    bool open(const char *path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) return false;
        buf_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
        pos_ = 0;
        return true;
    }
    // Positions the cursor at the next record's payload; false at end / on a malformed marker.
    bool next() {
        if (pos_ + 4 > buf_.size()) return false;
        int32_t head, tail;
        std::memcpy(&head, buf_.data() + pos_, 4);
        if (head < 0 || pos_ + 8 + size_t(head) > buf_.size()) return false;
        std::memcpy(&tail, buf_.data() + pos_ + 4 + head, 4);
        if (head != tail) return false;
        cur_ = pos_ + 4;
        end_ = cur_ + head;
        pos_ = end_ + 4;
        return true;
    }
    bool ok(size_t bytes) const { return cur_ + bytes <= end_; }
    int32_t i32() { int32_t v = 0; if (ok(4)) std::memcpy(&v, buf_.data() + cur_, 4); cur_ += 4; return v; }
    double f64() { double v = 0; if (ok(8)) std::memcpy(&v, buf_.data() + cur_, 8); cur_ += 8; return v; }
    std::string chars(size_t n) {
        std::string s;
        if (ok(n)) s.assign(buf_.data() + cur_, n);
        cur_ += n;
        while (!s.empty() && s.back() == ' ') s.pop_back();
        return s;
    }
    bool overrun() const { return cur_ > end_; }

  private:
    std::vector<char> buf_;
    size_t pos_ = 0, cur_ = 0, end_ = 0;
};

}  // namespace

int load_mech(const char *chem_bin, const char *tran_bin, Mech &m, std::string &err) {
    RecordReader r;
    if (!chem_bin || !r.open(chem_bin)) { err = std::string("cannot read ") + (chem_bin ? chem_bin : "(null)"); return -2; }
    if (!r.next()) { err = "chem.bin: missing header record"; return -2; }
    std::string vers = r.chars(16), prec = r.chars(16);
    if (prec.find("DOUB") == std::string::npos) { err = "chem.bin is not a DOUBLE PRECISION link file"; return -2; }
    if (!r.next()) { err = "chem.bin: missing size record"; return -2; }
    int h[29];
    for (int &v : h) v = r.i32();
    // LENI LENR LENC MM KK II MAXSP MAXTB MAXTP NTHCF NIPAR NITAR NIFAR NRV NFL NTB NLT NRL NW
    // NCHRG NEI NJA NIJAN NF1 NIF1 NEX NSTO NOR MAXORD (CKMN follows as a real)
    m.mm = h[3]; m.kk = h[4]; m.ii = h[5]; m.mxsp = h[6]; m.mxtb = h[7]; m.mxtp = h[8];
    m.ncp = h[9]; m.npar = h[10]; m.nfar = h[12];
    const int nrv = h[13], nfl = h[14], ntb = h[15];
    const int others = h[16] + h[17] + h[18] + h[20] + h[22] + h[24] + h[25] + h[26] + h[27];
    if (nrv || others) { err = "mechanism uses REV/LT/RLT/HV/EIM/JAN/FIT1/EXCI/RORD/FORD, not supported"; return -4; }
    if (m.kk > MAXK || m.ii > MAXR || m.mxtb > MAXTB || m.mxsp != 6 || m.ncp != 5 || m.mxtp != 3 || m.npar != 3) {
        err = "mechanism dimensions outside the supported range (KK<=64, II<=512, MAXTB<=16, MAXSP=6, NCP=5, MAXTP=3, NPAR=3)";
        return -4;
    }
    if (!r.next()) { err = "chem.bin: missing element record"; return -2; }
    if (!r.next()) { err = "chem.bin: missing species record"; return -2; }
    const int kk = m.kk, ncp2 = m.ncp + 2;
    m.names.resize(kk); m.wt.resize(kk); m.nt.resize(kk); m.ncf.resize(size_t(kk) * m.mm);
    m.tt.resize(size_t(kk) * m.mxtp); m.aa.resize(size_t(kk) * (m.mxtp - 1) * ncp2);
    for (int k = 0; k < kk; ++k) {
        m.names[k] = r.chars(16);
        for (int e = 0; e < m.mm; ++e) m.ncf[size_t(k) * m.mm + e] = r.i32();
        r.i32(); r.i32();                           // phase, charge
        m.wt[k] = r.f64();
        m.nt[k] = r.i32();
        for (int l = 0; l < m.mxtp; ++l) m.tt[size_t(k) * m.mxtp + l] = r.f64();
        for (int l = 0; l < m.mxtp - 1; ++l)
            for (int c = 0; c < ncp2; ++c) m.aa[(size_t(k) * (m.mxtp - 1) + l) * ncp2 + c] = r.f64();
        if (m.nt[k] != 3) { err = "species with other than 3 NASA temperatures"; return -4; }
    }
    if (r.overrun()) { err = "chem.bin: species record too short"; return -2; }
    if (m.ii > 0) {
        if (!r.next()) { err = "chem.bin: missing reaction record"; return -2; }
        const int ii = m.ii, np1 = m.npar + 1;
        m.ns.resize(ii); m.nr.resize(ii); m.par.resize(size_t(ii) * np1);
        m.nu.resize(size_t(ii) * m.mxsp); m.nk.resize(size_t(ii) * m.mxsp);
        for (int i = 0; i < ii; ++i) {
            m.ns[i] = r.i32(); m.nr[i] = r.i32();
            for (int c = 0; c < m.npar; ++c) m.par[size_t(i) * np1 + c] = r.f64();
            m.par[size_t(i) * np1 + m.npar] = 1.0;  // perturbation factor (cklib.f:4162-4164)
            for (int s = 0; s < m.mxsp; ++s) { m.nu[size_t(i) * m.mxsp + s] = r.i32(); m.nk[size_t(i) * m.mxsp + s] = r.i32(); }
        }
        if (r.overrun()) { err = "chem.bin: reaction record too short"; return -2; }
        if (nfl > 0) {
            if (!r.next()) { err = "chem.bin: missing fall-off record"; return -2; }
            m.ifl.resize(nfl); m.ifo.resize(nfl); m.kfl.resize(nfl); m.fpar.resize(size_t(nfl) * m.nfar);
            for (int f = 0; f < nfl; ++f) {
                m.ifl[f] = r.i32(); m.ifo[f] = r.i32(); m.kfl[f] = r.i32();
                for (int c = 0; c < m.nfar; ++c) m.fpar[size_t(f) * m.nfar + c] = r.f64();
            }
        }
        if (ntb > 0) {
            if (!r.next()) { err = "chem.bin: missing third-body record"; return -2; }
            m.itb.resize(ntb); m.ntbs.resize(ntb); m.nktb.resize(size_t(ntb) * m.mxtb); m.aik.resize(size_t(ntb) * m.mxtb);
            for (int t = 0; t < ntb; ++t) {
                m.itb[t] = r.i32(); m.ntbs[t] = r.i32();
                for (int l = 0; l < m.mxtb; ++l) { m.nktb[size_t(t) * m.mxtb + l] = r.i32(); m.aik[size_t(t) * m.mxtb + l] = r.f64(); }
            }
        }
        if (r.overrun()) { err = "chem.bin: fall-off / third-body record too short"; return -2; }
    }

    // Device thermo table.
    DevThermo &d = m.thermo;
    d = DevThermo{};
    d.kk = kk; d.ru = m.ru; d.pa = m.pa;
    for (int k = 0; k < kk; ++k) {
        d.wt[k] = m.wt[k];
        d.tmid[k] = m.tt[size_t(k) * m.mxtp + 1];
        for (int c = 0; c < 7; ++c) {
            d.alo[k][c] = m.aa[(size_t(k) * 2 + 0) * 7 + c];
            d.ahi[k][c] = m.aa[(size_t(k) * 2 + 1) * 7 + c];
        }
    }

    thermo_h298(d);
    return tran_bin ? load_tran_bin(tran_bin, m, err) : 0;
}

int load_tran_bin(const char *tran_bin, Mech &m, std::string &err) {
    const int kk = m.kk;
    {
        RecordReader t;
        if (!t.open(tran_bin)) { err = std::string("cannot read ") + tran_bin; return -2; }
        if (!t.next()) { err = "tran.bin: missing header"; return -2; }
        t.chars(16);
        if (t.chars(16).find("DOUB") == std::string::npos) { err = "tran.bin is not DOUBLE PRECISION"; return -2; }
        if (!t.next()) { err = "tran.bin: missing size record"; return -2; }
        t.i32(); t.i32();
        int no = t.i32(), tk = t.i32();
        t.i32();
        if (tk != kk) { err = "tran.bin species count differs from chem.bin"; return -2; }
        if (!t.next()) { err = "tran.bin: missing data record"; return -2; }
        m.no = no;
        m.patmos = t.f64();
        m.twt.resize(kk);
        for (int k = 0; k < kk; ++k) {
            m.twt[k] = t.f64();
            for (int c = 0; c < 5; ++c) t.f64();   // eps/k, sigma, dipole, polarizability, Zrot
            t.i32();                               // linearity
        }
        m.coflam.resize(size_t(kk) * no); m.cofeta.resize(size_t(kk) * no); m.cofd.resize(size_t(kk) * kk * no);
        for (double &v : m.coflam) v = t.f64();
        for (double &v : m.cofeta) v = t.f64();
        for (double &v : m.cofd) v = t.f64();
        if (t.overrun()) { err = "tran.bin: data record too short"; return -2; }
        m.has_tran = true;
    }
    return 0;
}

}  // namespace fcfd
