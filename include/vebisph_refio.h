/* vebisph_refio.h — C ABI of the readers and writers of the reference's run files (host code only).
 *
 * The reference (jrcking/vebisph) reads a run from INDAT, IBOUND and IPART, or TEMP_REC.dat on
 * restart, with Fortran list-directed input (source/base/input.F90:12-180), and writes
 * data_directory/PART<n>, TIME, LS, TIME_OUT and TEMP_REC.dat with list-directed output
 * (source/base/output.F90:1-89, source/base/sph2D_incom.F90:52-61).  These routines read and write the
 * same files, so a run directory prepared for the reference drives vb_drive (the GPU driver) and its
 * output files are read by the same post-processing.  Shipped as vebisph_amd/libvebisph_refio.so.
 *
 * Reading follows list-directed input: a READ starts a new record, takes its values from that record
 * and the following ones as needed, and ignores the rest of its last record (so INDAT's trailing
 * comments are skipped); values are separated by blanks or commas; reals accept E, D or no exponent
 * letter (1.0D-3, 1.0-3); logicals take their first letter after an optional '.' (T/F).
 * Writing follows the reference's compiler (flang) list-directed output: each item preceded by one
 * blank; a record continues on a new line (again led by a blank) when an item would pass column 79;
 * integers as digits; reals with the shortest digits that read back to the same double, in fixed
 * form when the value rounded to one significant digit lies in [0.1, 1e15) (".1", "99.", "1.5E-02"
 * otherwise) and in 1P exponent form with at least two exponent digits otherwise.  Values read back
 * bit-exactly; the text differs from flang's only in the last digit of some 16- and 17-digit values,
 * where flang picks another of the strings that read back to the same double (checked on 160 values
 * from 1e-8 to 1e20: 150 identical, 10 of that kind).
 *
 * Every routine returns 0 on success and 1 on failure, with a message in vbio_last_error() (per
 * thread).  Arrays are per-component columns; particle indices are 1-based like the reference's.
 */
#ifndef VEBISPH_REFIO_H
#define VEBISPH_REFIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VBIO_MAX_PATCHES = 16, VBIO_MAX_CIRCLES = 8 };

/* INDAT, input.F90:18-37, in the order it is read */
typedef struct {
  int32_t i_restart;             /* logical */
  int32_t i_open_domain;
  double tmax;
  int32_t i_shift_coeff;
  int32_t np, npfb, nbS;
  double dt_out, h, dx, dt;
  double body_force[2];
  double Ra, Pr, El, beta;
  double dt_coef_cfl, dt_coef_visc;
  double ve_nonlinearity, alpha_evss;
} vbio_indat;

/* IBOUND, input.F90:51-81: patches (node, edge, type, velocity, temperature, periodic parent for
 * type 2, U_inflow for type 3), then circles */
typedef struct {
  int32_t nb_patches;
  double b_node[VBIO_MAX_PATCHES][2], b_edge[VBIO_MAX_PATCHES][2];
  int32_t b_type[VBIO_MAX_PATCHES];
  double b_vel[VBIO_MAX_PATCHES], b_thermal[VBIO_MAX_PATCHES];
  int32_t b_periodic_parent[VBIO_MAX_PATCHES];   /* 0 where the patch is not periodic */
  double u_inflow;                               /* last type-3 patch's value, else 0 */
  int32_t nb_circles;
  double c_centre[VBIO_MAX_CIRCLES][2], c_radius[VBIO_MAX_CIRCLES], c_omega[VBIO_MAX_CIRCLES];
  double c_vel[VBIO_MAX_CIRCLES][2], c_thermal[VBIO_MAX_CIRCLES];
} vbio_ibound;

/* TEMP_REC.dat header lines, output.F90:61-63 (np and npfb already minus the binned particles) */
typedef struct {
  int32_t itime;
  double time;
  int32_t n_dump;
  int32_t np, npfb, nbS, np_inflow;
  double xb_min, xb_max, yb_min, yb_max, dx, av_conc;
} vbio_restart;

const char* vbio_last_error(void);

int32_t vbio_read_indat(const char* path, vbio_indat* out);
int32_t vbio_write_indat(const char* path, const vbio_indat* in);
int32_t vbio_read_ibound(const char* path, vbio_ibound* out);
int32_t vbio_write_ibound(const char* path, const vbio_ibound* in);
/* check_boundaries, input.F90:184-240: nodes joined by edges, edges summing to zero, periodic
 * partners of equal length and opposite direction (tolerance 1e-10); 1 and a message on failure */
int32_t vbio_check_ibound(const vbio_ibound* b);

/* IPART, input.F90:139-144: npfb records x y u v theta dv; *dv is the last record's value (the
 * reference reads it into one scalar) */
int32_t vbio_read_ipart(const char* path, int32_t npfb, double* x, double* y, double* u, double* v,
                        double* theta, double* dv);
int32_t vbio_write_ipart(const char* path, int32_t npfb, const double* x, const double* y, const double* u,
                         const double* v, const double* theta, double dv);

/* PART<n>, output.F90:44-51 (viscoelastic layout): for each of rows 1..n with inbin == 0 one list of
 * x y u v theta P conc a_out n_surf vort tau11 tau12 tau22; with tau11, tau12, tau22 all NULL the
 * Newtonian build's ten columns (output.F90:46-47) */
int32_t vbio_write_part(const char* path, int32_t n, const double* x, const double* y, const double* u,
                        const double* v, const double* theta, const double* p, const double* conc,
                        const double* a_out, const int32_t* n_surf, const double* vort, const double* tau11,
                        const double* tau12, const double* tau22, const int32_t* inbin);
/* the 13 columns of a PART file back, for up to cap particles: column c of particle i at
 * cols13[c*cap + i]; *n gets the count */
int32_t vbio_read_part(const char* path, int32_t cap, int32_t* n, double* cols13);
/* the same for ncols = 13 (viscoelastic build) or 10 (Newtonian build, output.F90:46-47: no stress
 * columns; vbio_write_part writes that layout when tau11, tau12, tau22 are all NULL) */
int32_t vbio_read_part_ncols(const char* path, int32_t ncols, int32_t cap, int32_t* n, double* cols);

/* TEMP_REC.dat, output.F90:60-81: the header, then for rows 1..n with inbin == 0 the lists
 * x y u v theta p dv n_surf and tau11 tau12 tau21 tau22 (with all four stress arrays NULL: the
 * Newtonian build's file, without the stress lines, output.F90:69-71; likewise for the reader) */
int32_t vbio_write_temp_rec(const char* path, const vbio_restart* hd, int32_t n, const double* x,
                            const double* y, const double* u, const double* v, const double* theta,
                            const double* p, double dv, const int32_t* n_surf, const double* tau11,
                            const double* tau12, const double* tau21, const double* tau22,
                            const int32_t* inbin);
/* Reads that file back: the header, then hd->npfb particles (call first with cap = 0 for the
 * header alone).  The reference's restart reader (input.F90:126-128) takes seven values per
 * particle line, x y u v p dv n_surf, against the eight its writer puts there (theta after v), and
 * so reads the temperature as the pressure; this reader takes the eight that are written. */
int32_t vbio_read_temp_rec(const char* path, vbio_restart* hd, int32_t cap, double* x, double* y, double* u,
                           double* v, double* theta, double* p, double* dv, int32_t* n_surf, double* tau11,
                           double* tau12, double* tau21, double* tau22);

/* One list-directed record of the given items appended to the text file at path (created if
 * absent): kinds[k] is 'd' (real, vals[k]) or 'i' (integer, (int64) vals[k]).  The driver's
 * per-step lines: TIME (output.F90:86), LS (:87), TIME_OUT (:55), unit 192 (div_free.F90:182). */
int32_t vbio_append_record(const char* path, const char* kinds, const double* vals);
/* A real as the reference's list-directed output writes it (no leading blank); returns its length,
 * or -1 when buf is too small */
int32_t vbio_format_real(double x, char* buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
