/*
 * CPU oracle for the compact-difference hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C++ restatement of the
 * reference algorithm (emathew1/MPI_3DCompact), used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * It is never linked into, loaded by, or called from the product library
 * (mpi_3dcompact_amd/), and the product path has no CPU fallback.
 *
 * Layout: every 3-D field is a contiguous array with x fastest,
 * index = (k*Ny + j)*Nx + i  (the reference's X-pencil order, src/Macros.hpp:16).
 * The reference computes on pencils in three layouts; per-line arithmetic does
 * not depend on the layout, so the oracle keeps one.
 *
 * Compiled with -ffp-contract=off so that its rounding matches the reference's
 * x86-64 baseline build without FMA (src/Makefile.in:1-4).
 */
#ifndef MC3D_ORACLE_H
#define MC3D_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scheme ids: derivatives */
enum { ORC_CD2 = 0, ORC_PADE6 = 1, ORC_PENTA10 = 2 };
/* filter ids */
enum { ORC_COMPACT8 = 0, ORC_COMPACT10 = 1, ORC_COMMN5PADE = 2, ORC_COMMN5EXP = 3 };
/* boundary types */
enum { ORC_PERIODIC = 0, ORC_DIRICHLET = 1 };

/* ---- banded solvers, src/Utils.cpp:4-224 ---- */
void orc_solve_tri(const double *a, const double *b, const double *c, const double *d,
                   double *x, double *work, int n);
void orc_cyclic_tri(const double *a, const double *b, const double *c, double alpha,
                    double beta, const double *r, int n, double *x);
void orc_solve_penta(const double *e, const double *a, const double *d, const double *c,
                     const double *f, const double *b, double *x, int n);
void orc_cyclic_penta(const double *a, const double *b, const double *c, const double *d,
                      const double *e, const double *f, const double cp[6], double *x, int n);
void orc_invert4(const double m[16], double out[16]);

/* ---- single-line operators ---- */
/* derivative of one line; tpb != NULL selects the transformed-periodic variant
 * with wrap values tpb[0..5] = {Nm3, Nm2, Nm1, Np1, Np2, Np3}. */
int orc_deriv_line(int scheme, int bc, int n, double dd, const double *phi,
                   double *dphi, const double *tpb);
int orc_filter_line(int kind, int bc, int n, double alphaF, const double *phi, double *phiF);
double orc_neumann(int scheme, const double *f);

/* ---- whole-field operators (x-fastest layout) ---- */
/* dir: 0 = xi (x), 1 = eta (y), 2 = zeta (z).  bc3 = BC type of x,y,z. */
int orc_deriv_field(int scheme, int dir, int bc, int Nx, int Ny, int Nz, double dd,
                    const double *in, double *out);
/* TPB: wrap arrays have one value per line, indexed by the line's position in
 * the x-fastest ordering of the two other indices (same order as the field). */
int orc_deriv_field_tpb(int scheme, int dir, int Nx, int Ny, int Nz, double dd,
                        const double *in, double *out, const double *const wrap[6]);
int orc_filter_field(int kind, double alphaF, int dir, const int bc3[3], int Nx, int Ny,
                     int Nz, const double *in, double *out);

/* ---- metrics from mesh coordinates (solveForJacobians) ----
 * fd/bc per direction, N = {Nx, Ny, Nz}, trans[d] = translation vector of
 * periodic direction d; x, y, z and out[10] = {J, J11, ..., J33} are x-fastest
 * fields.  Returns 0, or nonzero for a periodic direction with a scheme the
 * reference's metric path rejects (CD2). */
int orc_metrics(const int fd[3], const int bc[3], const int N[3], const double trans[3][3],
                const double *x, const double *y, const double *z, double *const out[10]);
/* handlePeriodicStuff: face kinds 0 PERIODIC, 1 INTERNALLY_PERIODIC */
void orc_periodic_translations(const int bc[3], const int face_lo[3], const int face_hi[3],
                               const double disp[3][3], double trans[3][3], int set[3]);

/* ---- OpenMP threads of the oracle's parallel loops (CPU baseline timing) ---- */
int orc_max_threads(void);            /* omp_get_max_threads() */
void orc_set_threads(int n);          /* omp_set_num_threads(n) */

#ifdef __cplusplus
}
#endif
#endif
