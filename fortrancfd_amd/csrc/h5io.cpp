// h5io.cpp -- the reference's HDF5 snapshot layout (write_hdf5_b / read_hdf5_restart) on the HDF5 C
// library.  Serial: the single-GPU run owns the whole domain, so the reference's collective MPI-IO
// writes of per-rank hyperslabs become one write of each full dataset (same file content).
#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fcfd_h5.h"

namespace {

thread_local std::string g_err;

int fail(const std::string &m) {
    g_err = m;
    return -1;
}

// SRC/write_hdf5.f90:162-166: 3-character names, leading blanks kept by TRIM
const char *const kNames[15] = {"rho", "u_x", "  T", "  p", " et", " ek", " ei", "HRR",
                                "HRD", "Qrd", " Tw", "QTw", "x_m", " dx", "idx"};

std::string species_name(const char *species, int k) {
    char buf[17];
    std::memcpy(buf, species + 16 * k, 16);
    buf[16] = '\0';
    std::string s(buf);
    while (!s.empty() && s.back() == ' ') s.pop_back();      // trim(species_name(k))
    return s;
}

// Fortran E11.5 of a non-negative value: 0.dddddE+ee
std::string fortran_e11_5(double v) {
    if (v == 0.0) return "0.00000E+00";
    char b[64];
    std::snprintf(b, sizeof b, "%.4E", v);                    // d.ddddE+ee, correctly rounded
    const std::string s(b);
    const size_t epos = s.find('E');
    const std::string digits = s.substr(0, 1) + s.substr(2, epos - 2);
    const int e = std::atoi(s.c_str() + epos + 1) + 1;
    char out[32];
    std::snprintf(out, sizeof out, "0.%sE%c%02d", digits.c_str(), e < 0 ? '-' : '+', e < 0 ? -e : e);
    return out;
}

struct Handle {                        // closes an HDF5 identifier on scope exit
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Handle() { if (id >= 0) close(id); }
    Handle(const Handle &) = delete;
    Handle &operator=(const Handle &) = delete;
};

int write_attr_double(hid_t g, const char *name, double v) {
    const hsize_t one = 1;
    Handle sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
    Handle a(H5Acreate2(g, name, H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Awrite(a.id, H5T_NATIVE_DOUBLE, &v) < 0) return fail(std::string("attribute ") + name);
    return 0;
}

int write_attr_int(hid_t g, const char *name, int v) {
    const hsize_t one = 1;
    Handle sp(H5Screate_simple(1, &one, nullptr), H5Sclose);
    Handle a(H5Acreate2(g, name, H5T_NATIVE_INT, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Awrite(a.id, H5T_NATIVE_INT, &v) < 0) return fail(std::string("attribute ") + name);
    return 0;
}

herr_t collect_name(hid_t, const char *name, const H5L_info_t *, void *op) {
    static_cast<std::vector<std::string> *>(op)->push_back(name);
    return 0;
}

// the last root group in link-name order (read_hdf5_restart: index nlinks - 1)
int last_group(hid_t f, std::string &name) {
    std::vector<std::string> names;
    hsize_t idx = 0;
    if (H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, &idx, collect_name, &names) < 0 || names.empty())
        return fail("no output group in the file");
    name = names.back();
    return 0;
}

template <class T>
int read_attr(hid_t g, const char *name, hid_t type, T *out) {
    if (!out) return 0;
    Handle a(H5Aopen(g, name, H5P_DEFAULT), H5Aclose);
    if (a.id < 0 || H5Aread(a.id, type, out) < 0) return fail(std::string("attribute ") + name);
    return 0;
}

}  // namespace

extern "C" {

const char *fcfd_h5_last_error(void) { return g_err.c_str(); }

int fcfd_h5_write_step(const char *path, int create, int nout, int nt, double dt, double time_total,
                       double time_restart, double cpu_total, double cpu_restart, int ncells, int kk,
                       const char *species, const double *fields) {
    if (!path || ncells < 1 || kk < 1 || !species || !fields) return fail("fcfd_h5_write_step: invalid argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Handle f(create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT) : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT),
             H5Fclose);
    if (f.id < 0) return fail(std::string("cannot open ") + path);
    char step[64];
    std::snprintf(step, sizeof step, "nout%06d_time%s", nout, fortran_e11_5(time_total).c_str());
    Handle g(H5Gcreate2(f.id, step, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (g.id < 0) return fail(std::string("cannot create group ") + step);
    const hsize_t dims = hsize_t(ncells);
    for (int r = 0; r < 15 + kk; ++r) {
        const std::string name = r < 15 ? std::string(kNames[r]) : species_name(species, r - 15);
        Handle sp(H5Screate_simple(1, &dims, nullptr), H5Sclose);
        Handle d(H5Dcreate2(g.id, name.c_str(), H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                 H5Dclose);
        if (d.id < 0) return fail("cannot create dataset '" + name + "'");
        if (H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, fields + size_t(r) * ncells) < 0)
            return fail("cannot write dataset '" + name + "'");
    }
    // SRC/write_hdf5.f90:254-270 (the loop writes names 1..7)
    if (write_attr_double(g.id, "dt", dt) || write_attr_double(g.id, "time from 0s", time_total) ||
        write_attr_double(g.id, "time from restart", time_restart) ||
        write_attr_double(g.id, "cpu time total", cpu_total) ||
        write_attr_double(g.id, "cpu time restart", cpu_restart) ||
        write_attr_int(g.id, "number of output", nout) || write_attr_int(g.id, "number of time step", nt))
        return -1;
    return 0;
}

int fcfd_h5_last_size(const char *path, int *ncells) {
    if (!path || !ncells) return fail("fcfd_h5_last_size: invalid argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return fail(std::string("cannot open ") + path);
    std::string gname;
    if (last_group(f.id, gname)) return -1;
    Handle g(H5Gopen2(f.id, gname.c_str(), H5P_DEFAULT), H5Gclose);
    Handle d(H5Dopen2(g.id, "rho", H5P_DEFAULT), H5Dclose);
    if (d.id < 0) return fail("no rho dataset in " + gname);
    Handle sp(H5Dget_space(d.id), H5Sclose);
    hsize_t dims = 0;
    if (H5Sget_simple_extent_ndims(sp.id) != 1 || H5Sget_simple_extent_dims(sp.id, &dims, nullptr) < 0)
        return fail("rho is not one-dimensional");
    *ncells = int(dims);
    return 0;
}

int fcfd_h5_read_last(const char *path, int ncells, int kk, const char *species, double *fields, int *nout,
                      int *nt, double *time_total, double *cpu_total) {
    if (!path || ncells < 1 || kk < 1 || !species || !fields) return fail("fcfd_h5_read_last: invalid argument");
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    Handle f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (f.id < 0) return fail(std::string("cannot open ") + path);
    std::string gname;
    if (last_group(f.id, gname)) return -1;
    Handle g(H5Gopen2(f.id, gname.c_str(), H5P_DEFAULT), H5Gclose);
    if (g.id < 0) return fail("cannot open group " + gname);
    for (int r = 0; r < 15 + kk; ++r) {
        const std::string name = r < 15 ? std::string(kNames[r]) : species_name(species, r - 15);
        double *dst = fields + size_t(r) * ncells;
        if (H5Lexists(g.id, name.c_str(), H5P_DEFAULT) <= 0) {
            std::memset(dst, 0, sizeof(double) * ncells);
            continue;
        }
        Handle d(H5Dopen2(g.id, name.c_str(), H5P_DEFAULT), H5Dclose);
        Handle sp(H5Dget_space(d.id), H5Sclose);
        hsize_t dims = 0;
        if (d.id < 0 || H5Sget_simple_extent_dims(sp.id, &dims, nullptr) != 1 || dims != hsize_t(ncells))
            return fail("dataset '" + name + "' does not hold " + std::to_string(ncells) + " values");
        if (H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, dst) < 0)
            return fail("cannot read dataset '" + name + "'");
    }
    if (read_attr(g.id, "time from 0s", H5T_NATIVE_DOUBLE, time_total) ||
        read_attr(g.id, "cpu time total", H5T_NATIVE_DOUBLE, cpu_total) ||
        read_attr(g.id, "number of output", H5T_NATIVE_INT, nout) ||
        read_attr(g.id, "number of time step", H5T_NATIVE_INT, nt))
        return -1;
    return 0;
}

}  // extern "C"
