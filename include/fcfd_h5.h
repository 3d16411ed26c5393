/* fcfd_h5.h -- C ABI of libfcfd_h5.so: the reference's HDF5 snapshot format, so that its
 * post-processing scripts read this library's output and its restarts read back.
 *
 * Host-only code over the HDF5 C library (1.10).  Field rows are what fcfd_eos_fields returns:
 * 15 + KK rows of ncells float64 values (rho, u_x, T, p, e_total, e_kinetic, e_internal, HRR,
 * HRR_dot, Qrad, Tw, qTw, x_m, dx, idx, X_1..X_KK).  species: KK names, each in a 16-byte field
 * (NUL-padded).  Every entry point returns 0 or a negative code; fcfd_h5_last_error() holds a
 * message. */
#ifndef FCFD_H5_H
#define FCFD_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* write_hdf5_b (SRC/write_hdf5.f90:48-330) for one output step, appended to path (create = 1:
 * truncate or create it first, init_output :15-33).  Group "noutNNNNNN_time" + time_total in
 * Fortran's E11.5 (:188); one contiguous float64 dataset per row, named rho, u_x, "  T", "  p",
 * " et", " ek", " ei", HRR, HRD, Qrd, " Tw", QTw, x_m, " dx", idx (the reference's 3-character
 * names, leading blanks kept, :162-166) and the species names; group attributes dt,
 * "time from 0s", "time from restart", "cpu time total", "cpu time restart" (float64) and
 * "number of output", "number of time step" (native int) -- the reference's attribute loop stops
 * before its eighth name, "chem directory name" (:276), and so does this one. */
int fcfd_h5_write_step(const char *path, int create, int nout, int nt, double dt, double time_total,
                       double time_restart, double cpu_total, double cpu_restart, int ncells, int kk,
                       const char *species, const double *fields);
/* Size (cells) of the last group of path, in link-name order -- the group read_hdf5_restart
 * opens (SRC/read_hdf5.f90:118-139). */
int fcfd_h5_last_size(const char *path, int *ncells);
/* read_hdf5_restart (SRC/read_hdf5.f90:13-423) on that group: the field rows by dataset name
 * (a species without a dataset reads as 0, as the reference's x_sp stays unset), "time from 0s",
 * "cpu time total", "number of output" and "number of time step".  fields: 15 + KK rows of ncells
 * values; ncells must match the group.  Any output pointer but fields may be NULL. */
int fcfd_h5_read_last(const char *path, int ncells, int kk, const char *species, double *fields,
                      int *nout, int *nt, double *time_total, double *cpu_total);
const char *fcfd_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
