// mech_text.cpp -- reader of CHEMKIN-II mechanism TEXT (chem.inp + therm.dat) into the tables the
// link-file reader fills, without going through the interpreter and its binary link file.
//
// The reference turns its mechanism into chem.bin with the CHEMKIN-II interpreter (ckinterp.f, run
// by WORK/KineticModels/Chem_bin/*/comp.sh) and reads that with CKINIT (cklib.f:3972).  This file
// reads the same text and produces the same arrays -- element counts, molar masses, NASA-7
// coefficients and temperature ranges, Arrhenius parameters with the activation energy converted
// to E/R, stoichiometric coefficients and species slots, third-body efficiency lists and fall-off
// parameters -- with the interpreter's conventions where they decide the numbers:
//   * atomic weights are the interpreter's table values, which it stores from single-precision
//     literals (so they are single-precision values widened to double);
//   * a species' molar mass adds atomic weight x count over the element fields of its thermo line,
//     in field order;
//   * E/R = E x EFAC with EFAC = 4.184 / 8.314510 for cal/mole (4.184 a single-precision literal),
//     4184 / 8.314510 for kcal/mole, 1 / 8.314510 for J/mole, 4000 / 8.314510 for kJ/mole, 1 for
//     Kelvins; the LOW parameters of a fall-off reaction are converted alike;
//   * reactant slots 1-3 hold negative coefficients, product slots 4-6 positive ones, species in
//     order of first appearance, repeated species on one side merged into one slot;
//   * every "(+M)" fall-off reaction is also a third-body reaction (its [M] comes from the
//     third-body sum), a "(+species)" one is not; fall-off forms 1 Lindemann (LOW only), 2 SRI (a 3-
//     parameter SRI gets d = 1, e = 0), 3 and 4 Troe with 3 or 4 parameters.
// Unsupported (refused with a message, as the link-file reader refuses them): REV, LT / RLT, HV,
// FORD / RORD, JAN / FIT1, EXCI, PLOG, real stoichiometric coefficients, molecule units.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "fcfd_common.hpp"

namespace fcfd {
namespace {

std::string upper(std::string s) {
    for (char &c : s) c = char(std::toupper((unsigned char)c));
    return s;
}

std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::vector<std::string> tokens(const std::string &s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

// A Fortran-style real: D or d exponents allowed, blanks ignored.
bool parse_real(std::string s, double &v) {
    std::string t;
    for (char c : s)
        if (c != ' ') t += (c == 'D' || c == 'd') ? 'E' : c;
    if (t.empty()) return false;
    char *end = nullptr;
    v = std::strtod(t.c_str(), &end);
    return end && *end == '\0';
}

// Atomic weights of the interpreter's element table (single-precision literals there).
double atomic_weight(const std::string &e) {
    static const std::map<std::string, float> w = {
        {"H", 1.00797f}, {"HE", 4.00260f}, {"LI", 6.93900f}, {"BE", 9.01220f}, {"B", 10.81100f},
        {"C", 12.01115f}, {"N", 14.00670f}, {"O", 15.99940f}, {"F", 18.99840f}, {"NE", 20.18300f},
        {"NA", 22.98980f}, {"MG", 24.31200f}, {"AL", 26.98150f}, {"SI", 28.08600f}, {"P", 30.97380f},
        {"S", 32.06400f}, {"CL", 35.45300f}, {"AR", 39.94800f}, {"K", 39.10200f}, {"CA", 40.08000f},
        {"FE", 55.84700f}, {"KR", 83.80000f}, {"XE", 131.30000f}, {"D", 2.01410f}, {"E", 5.45e-4f}};
    auto it = w.find(upper(e));
    return it == w.end() ? -1.0 : double(it->second);
}

struct Text {
    std::vector<std::string> elems;
    std::vector<double> awt;
    std::vector<std::string> species;
    double tlo = 300.0, tmid = 1000.0, thi = 5000.0;    // THERMO ALL defaults
};

// Strips a '!' comment.
std::string strip_comment(const std::string &l) {
    const size_t p = l.find('!');
    return p == std::string::npos ? l : l.substr(0, p);
}

// NASA-7 records (four 80-column lines, '1'..'4' in column 80) from lines[i0 ..); stops at END or
// REACTIONS.  The first record of a species wins.
int read_thermo(const std::vector<std::string> &lines, size_t i0, const Text &tx, Mech &m,
                std::vector<char> &have, std::string &err) {
    const int kk = int(tx.species.size());
    for (size_t i = i0; i < lines.size(); ++i) {
        const std::string &l = lines[i];
        const std::string t = upper(trim(strip_comment(l)));
        if (t.rfind("END", 0) == 0 || t.rfind("REAC", 0) == 0) break;
        if (l.size() < 80 || l[79] != '1') continue;
        if (i + 3 >= lines.size()) { err = "thermo: truncated species record"; return -1; }
        const auto tk = tokens(l.substr(0, 24));
        if (tk.empty()) continue;
        int k = -1;
        for (int s = 0; s < kk; ++s)
            if (upper(tx.species[s]) == upper(tk[0])) k = s;
        if (k < 0 || have[k]) { i += 3; continue; }
        have[k] = 1;
        // element fields: columns 25-44 (four of 5 characters) and 74-78
        double w = 0.0;
        for (int f = 0; f < 5; ++f) {
            const size_t col = f < 4 ? 24 + 5 * f : 73;
            if (l.size() < col + 5) continue;
            const std::string el = trim(l.substr(col, 2));
            const std::string cnt = trim(l.substr(col + 2, 3));
            if (el.empty() || cnt.empty()) continue;
            double cv;
            if (!parse_real(cnt, cv)) continue;
            const int n = int(cv);
            if (n == 0) continue;
            int e = -1;
            for (size_t q = 0; q < tx.elems.size(); ++q)
                if (upper(tx.elems[q]) == upper(el)) e = int(q);
            if (e < 0) { err = "thermo: element " + el + " of " + tx.species[k] + " is not declared"; return -1; }
            m.ncf[size_t(k) * m.mm + e] = n;
            w = w + tx.awt[e] * double(float(n));
        }
        m.wt[k] = w;
        double tr[3] = {tx.tlo, tx.tmid, tx.thi};
        double v;
        if (parse_real(l.substr(45, 10), v)) tr[0] = v;
        if (parse_real(l.substr(65, 8), v)) tr[1] = v;
        if (parse_real(l.substr(55, 10), v)) tr[2] = v;
        for (int c = 0; c < 3; ++c) m.tt[size_t(k) * 3 + c] = tr[c];
        double a[15];
        for (int r = 0; r < 3; ++r) {
            const std::string &cl = lines[i + 1 + r];
            const int nf = r < 2 ? 5 : 4;
            for (int f = 0; f < nf; ++f) {
                if (cl.size() < size_t(15 * f + 15) || !parse_real(cl.substr(15 * f, 15), a[5 * r + f])) {
                    err = "thermo: bad coefficient field of " + tx.species[k];
                    return -1;
                }
            }
        }
        // line 2: a1..a5 high; line 3: a6, a7 high, a1..a3 low; line 4: a4..a7 low
        double *lo = &m.aa[(size_t(k) * 2 + 0) * 7], *hi = &m.aa[(size_t(k) * 2 + 1) * 7];
        for (int c = 0; c < 7; ++c) hi[c] = a[c];
        for (int c = 0; c < 7; ++c) lo[c] = a[7 + c];
        i += 3;
    }
    return 0;
}

}  // namespace

int load_mech_text(const char *chem_inp, const char *therm_dat, Mech &m, std::string &err) {
    std::vector<std::string> lines;
    {
        std::ifstream f(chem_inp ? chem_inp : "");
        if (!f) { err = std::string("cannot read ") + (chem_inp ? chem_inp : "(null)"); return -2; }
        std::string l;
        while (std::getline(f, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            lines.push_back(l);
        }
    }
    m = Mech{};
    Text tx;
    size_t i = 0;
    // ---- ELEMENTS, SPECIES ----
    auto block = [&](const char *kw, std::vector<std::string> &out) -> bool {
        for (; i < lines.size(); ++i) {
            auto tk = tokens(strip_comment(lines[i]));
            if (tk.empty()) continue;
            if (upper(tk[0]).rfind(kw, 0) != 0) return false;
            bool end = false;
            for (size_t t = 1; t < tk.size(); ++t) {
                if (upper(tk[t]) == "END") { end = true; break; }
                out.push_back(tk[t]);
            }
            ++i;
            for (; !end && i < lines.size(); ++i) {
                for (auto &t : tokens(strip_comment(lines[i]))) {
                    if (upper(t) == "END") { end = true; break; }
                    out.push_back(t);
                }
                if (end) { ++i; break; }
            }
            return true;
        }
        return false;
    };
    std::vector<std::string> el;
    if (!block("ELEM", el)) { err = "chem.inp: no ELEMENTS block"; return -2; }
    for (auto &e : el) {
        std::string name = e;
        double w = -1.0;
        const size_t s1 = e.find('/');
        if (s1 != std::string::npos) {
            name = e.substr(0, s1);
            const size_t s2 = e.find('/', s1 + 1);
            if (s2 == std::string::npos || !parse_real(e.substr(s1 + 1, s2 - s1 - 1), w)) {
                err = "chem.inp: bad element weight " + e;
                return -2;
            }
        } else {
            w = atomic_weight(name);
        }
        if (w <= 0.0) { err = "chem.inp: no atomic weight for element " + name; return -4; }
        tx.elems.push_back(name);
        tx.awt.push_back(w);
    }
    if (!block("SPEC", tx.species)) { err = "chem.inp: no SPECIES block"; return -2; }
    const int kk = int(tx.species.size()), mm = int(tx.elems.size());
    if (kk < 1 || kk > MAXK) { err = "mechanism with " + std::to_string(kk) + " species (supported: 1.." + std::to_string(MAXK) + ")"; return -4; }
    m.kk = kk; m.mm = mm; m.mxsp = 6; m.mxtb = 10; m.mxtp = 3; m.ncp = 5; m.npar = 3; m.nfar = 8;
    m.names = tx.species;
    m.ncf.assign(size_t(kk) * mm, 0);
    m.wt.assign(kk, 0.0);
    m.nt.assign(kk, 3);
    m.tt.assign(size_t(kk) * 3, 0.0);
    m.aa.assign(size_t(kk) * 2 * 7, 0.0);
    std::vector<char> have(kk, 0);
    // ---- THERMO in chem.inp (first), then therm.dat ----
    size_t ir = lines.size();
    for (size_t j = i; j < lines.size(); ++j) {
        const auto tk = tokens(strip_comment(lines[j]));
        if (tk.empty()) continue;
        const std::string k0 = upper(tk[0]);
        if (k0.rfind("THER", 0) == 0) {
            size_t j0 = j + 1;
            if (tk.size() > 1 && upper(tk[1]) == "ALL") {
                const auto tt = tokens(lines[j + 1]);
                if (tt.size() < 3 || !parse_real(tt[0], tx.tlo) || !parse_real(tt[1], tx.tmid) || !parse_real(tt[2], tx.thi)) {
                    err = "chem.inp: THERMO ALL needs the three default temperatures";
                    return -2;
                }
                j0 = j + 2;
            }
            if (read_thermo(lines, j0, tx, m, have, err)) return -2;
        }
        if (k0.rfind("REAC", 0) == 0) { ir = j; break; }
    }
    if (therm_dat) {
        std::ifstream f(therm_dat);
        if (!f) { err = std::string("cannot read ") + therm_dat; return -2; }
        std::vector<std::string> tl;
        std::string l;
        while (std::getline(f, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            tl.push_back(l);
        }
        size_t j0 = 0;
        while (j0 < tl.size() && upper(trim(strip_comment(tl[j0]))).rfind("THER", 0) != 0) ++j0;
        if (j0 < tl.size()) {
            ++j0;
            while (j0 < tl.size() && trim(strip_comment(tl[j0])).empty()) ++j0;
            const auto tt = tokens(strip_comment(tl[j0]));
            Text t2 = tx;
            if (tt.size() >= 3 && parse_real(tt[0], t2.tlo) && parse_real(tt[1], t2.tmid) && parse_real(tt[2], t2.thi)) ++j0;
            if (read_thermo(tl, j0, t2, m, have, err)) return -2;
        }
    }
    for (int k = 0; k < kk; ++k)
        if (!have[k]) { err = "no thermodynamic data for species " + tx.species[k]; return -4; }

    // ---- REACTIONS ----
    if (ir >= lines.size()) { m.ii = 0; goto thermo_tables; }
    {
        double efac = double(4.184f) / 8.314510;     // cal/mole (the default)
        for (auto &t : tokens(strip_comment(lines[ir]))) {
            const std::string u = upper(t);
            if (u.rfind("KELV", 0) == 0) efac = 1.0;
            else if (u.rfind("KCAL", 0) == 0) efac = 4184.0 / 8.314510;
            else if (u.rfind("KJOU", 0) == 0) efac = 4000.0 / 8.314510;
            else if (u.rfind("JOUL", 0) == 0) efac = 1.0 / 8.314510;
            else if (u.rfind("CAL/", 0) == 0) efac = double(4.184f) / 8.314510;
            else if (u.rfind("MOLEC", 0) == 0 || u.rfind("EVOL", 0) == 0) {
                err = "reaction units " + t + " are not supported";
                return -4;
            }
        }
        auto species_index = [&](const std::string &s) -> int {
            for (int k = 0; k < kk; ++k)
                if (upper(tx.species[k]) == upper(s)) return k;
            return -1;
        };
        int ii = 0;
        for (size_t j = ir + 1; j < lines.size(); ++j) {
            const std::string l = trim(strip_comment(lines[j]));
            if (l.empty()) continue;
            if (upper(l).rfind("END", 0) == 0 && l.find('=') == std::string::npos) break;
            if (l.find('=') != std::string::npos) {
                // ---- a reaction line: equation, A, b, E ----
                auto tk = tokens(l);
                if (tk.size() < 4) { err = "reaction line without three Arrhenius parameters: " + l; return -2; }
                double par[3];
                for (int c = 0; c < 3; ++c)
                    if (!parse_real(tk[tk.size() - 3 + c], par[c])) { err = "bad Arrhenius parameter in: " + l; return -2; }
                std::string eq;
                for (size_t t = 0; t + 3 < tk.size(); ++t) eq += tk[t];
                if (ii >= MAXR) { err = "more than " + std::to_string(MAXR) + " reactions"; return -4; }
                int ir_dir = 1;
                size_t p1, p2;
                if ((p1 = eq.find("<=>")) != std::string::npos) p2 = p1 + 3;
                else if ((p1 = eq.find("=>")) != std::string::npos) { p2 = p1 + 2; ir_dir = -1; }
                else { p1 = eq.find('='); p2 = p1 + 1; }
                std::string side[2] = {eq.substr(0, p1), eq.substr(p2)};
                if (side[0].find('.') != std::string::npos || side[1].find('.') != std::string::npos) {
                    err = "real stoichiometric coefficients are not supported: " + eq;
                    return -4;
                }
                // fall-off "(+M)" / "(+species)"
                int ktb[2] = {0, 0};   // 0 none, -1 M, k + 1 species
                for (int s = 0; s < 2; ++s) {
                    const size_t q = side[s].find("(+");
                    if (q == std::string::npos) continue;
                    const size_t r = side[s].find(')', q);
                    if (r == std::string::npos) { err = "bad fall-off declaration in " + eq; return -2; }
                    const std::string who = side[s].substr(q + 2, r - q - 2);
                    if (upper(who) == "M") ktb[s] = -1;
                    else {
                        const int k = species_index(who);
                        if (k < 0) continue;
                        ktb[s] = k + 1;
                    }
                    side[s] = side[s].substr(0, q) + side[s].substr(r + 1);
                }
                bool fall = false, thb = false;
                int kfal = 0;
                if (ktb[0] != 0 || ktb[1] != 0) {
                    if (ktb[0] == -1 && ktb[1] == -1) { fall = true; thb = true; kfal = 0; }
                    else if (ktb[0] == ktb[1]) { fall = true; kfal = ktb[0]; }
                    else { err = "inconsistent fall-off declaration in " + eq; return -2; }
                }
                // species of each side, longest names first across '+' separators
                int nu[6] = {0, 0, 0, 0, 0, 0}, nk[6] = {0, 0, 0, 0, 0, 0};
                int ns = 0, nr = 0;
                for (int s = 0; s < 2; ++s) {
                    const std::string &str = side[s];
                    std::vector<size_t> plus = {size_t(-1)};
                    for (size_t q = 1; q + 1 < str.size(); ++q)
                        if (str[q] == '+') plus.push_back(q);
                    plus.push_back(str.size());
                    size_t a = 0;
                    int slot0 = s == 0 ? 0 : 3, nslot = 0;
                    while (a + 1 < plus.size()) {
                        bool found = false;
                        for (size_t b = plus.size() - 1; b > a; --b) {
                            std::string sp = str.substr(plus[a] + 1, plus[b] - plus[a] - 1);
                            if (upper(sp) == "M") {
                                thb = true;
                                a = b;
                                found = true;
                                break;
                            }
                            size_t nd = 0;
                            while (nd < sp.size() && std::isdigit((unsigned char)sp[nd])) ++nd;
                            int coef = nd ? std::atoi(sp.substr(0, nd).c_str()) : 1;
                            const int k = species_index(sp.substr(nd));
                            if (k < 0) {
                                if (b - a > 1) continue;
                                err = "undeclared species " + sp + " in " + eq;
                                return -2;
                            }
                            if (s == 0) coef = -coef;
                            int slot = -1;
                            for (int q = 0; q < nslot; ++q)
                                if (nk[slot0 + q] == k + 1) slot = slot0 + q;
                            if (slot >= 0) nu[slot] += coef;
                            else {
                                if (nslot == 3) { err = "more than three species on one side of " + eq; return -4; }
                                nu[slot0 + nslot] = coef;
                                nk[slot0 + nslot] = k + 1;
                                ++nslot;
                                ++ns;
                            }
                            a = b;
                            found = true;
                            break;
                        }
                        if (!found) { err = "cannot read " + eq; return -2; }
                    }
                    if (s == 0) nr = nslot;
                }
                m.ns.push_back(ir_dir * ns);
                m.nr.push_back(nr);
                for (int c = 0; c < 3; ++c) m.par.push_back(c == 2 ? par[c] * efac : par[c]);
                m.par.push_back(1.0);
                for (int q = 0; q < 6; ++q) { m.nu.push_back(nu[q]); m.nk.push_back(nk[q]); }
                ++ii;
                if (thb) {
                    m.itb.push_back(ii);
                    m.ntbs.push_back(0);
                    for (int q = 0; q < m.mxtb; ++q) { m.nktb.push_back(0); m.aik.push_back(0.0); }
                }
                if (fall) {
                    m.ifl.push_back(ii);
                    m.ifo.push_back(-1);
                    m.kfl.push_back(kfal);
                    for (int q = 0; q < m.nfar; ++q) m.fpar.push_back(0.0);
                }
                continue;
            }
            // ---- auxiliary line of the last reaction: KEY/values/ ... or DUPLICATE ----
            if (ii == 0) { err = "auxiliary data before the first reaction: " + l; return -2; }
            const bool is_thb = !m.itb.empty() && m.itb.back() == ii;
            const bool is_fal = !m.ifl.empty() && m.ifl.back() == ii;
            std::string rest = l;
            while (!trim(rest).empty()) {
                rest = trim(rest);
                if (upper(rest).rfind("DUP", 0) == 0) {
                    const size_t sp = rest.find_first_of(" \t");
                    rest = sp == std::string::npos ? "" : rest.substr(sp);
                    continue;
                }
                const size_t s1 = rest.find('/');
                const size_t s2 = s1 == std::string::npos ? s1 : rest.find('/', s1 + 1);
                if (s2 == std::string::npos) { err = "cannot read auxiliary data: " + l; return -2; }
                const std::string key = trim(rest.substr(0, s1));
                std::vector<double> vals;
                for (auto &t : tokens(rest.substr(s1 + 1, s2 - s1 - 1))) {
                    double v;
                    if (!parse_real(t, v)) { err = "bad number in auxiliary data: " + l; return -2; }
                    vals.push_back(v);
                }
                rest = rest.substr(s2 + 1);
                const std::string K = upper(key);
                if (K == "LOW" || K == "TROE" || K == "SRI") {
                    if (!is_fal) { err = key + " for a reaction without fall-off: " + l; return -2; }
                    double *fp = &m.fpar[(m.ifl.size() - 1) * m.nfar];
                    int &ifo = m.ifo.back();
                    if (K == "LOW") {
                        if (vals.size() != 3 || ifo > 0) { err = "bad LOW parameters: " + l; return -2; }
                        ifo = std::abs(ifo);
                        fp[0] = vals[0]; fp[1] = vals[1]; fp[2] = vals[2] * efac;
                    } else if (K == "TROE") {
                        if (vals.size() != 3 && vals.size() != 4) { err = "TROE needs 3 or 4 parameters: " + l; return -2; }
                        for (size_t q = 0; q < vals.size(); ++q) fp[3 + q] = vals[q];
                        ifo = int(vals.size()) * ifo;
                    } else {
                        if (vals.size() != 3 && vals.size() != 5) { err = "SRI needs 3 or 5 parameters: " + l; return -2; }
                        for (size_t q = 0; q < vals.size(); ++q) fp[3 + q] = vals[q];
                        if (vals.size() == 3) { fp[6] = 1.0; fp[7] = 0.0; }
                        ifo = 2 * ifo;
                    }
                    continue;
                }
                const int k = species_index(key);
                if (k >= 0) {
                    if (!is_thb || vals.size() != 1) { err = "third-body efficiency for a reaction without +M: " + l; return -2; }
                    const size_t t = m.itb.size() - 1;
                    if (m.ntbs[t] >= m.mxtb) { err = "more than 10 third-body efficiencies: " + l; return -4; }
                    m.nktb[t * m.mxtb + m.ntbs[t]] = k + 1;
                    m.aik[t * m.mxtb + m.ntbs[t]] = vals[0];
                    ++m.ntbs[t];
                    continue;
                }
                err = "auxiliary keyword " + key + " is not supported (" + l + ")";
                return -4;
            }
        }
        m.ii = ii;
        for (size_t f = 0; f < m.ifo.size(); ++f)
            if (m.ifo[f] <= 0) { err = "fall-off reaction " + std::to_string(m.ifl[f]) + " without LOW parameters"; return -2; }
    }
thermo_tables:
    {
        DevThermo &d = m.thermo;
        d = DevThermo{};
        d.kk = kk; d.ru = m.ru; d.pa = m.pa;
        for (int k = 0; k < kk; ++k) {
            d.wt[k] = m.wt[k];
            d.tmid[k] = m.tt[size_t(k) * 3 + 1];
            for (int c = 0; c < 7; ++c) {
                d.alo[k][c] = m.aa[(size_t(k) * 2 + 0) * 7 + c];
                d.ahi[k][c] = m.aa[(size_t(k) * 2 + 1) * 7 + c];
            }
        }
        thermo_h298(d);
    }
    return 0;
}

}  // namespace fcfd
