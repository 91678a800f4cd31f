/*
 * mc3d: MI355X-native (gfx950) per-timestep hot path of a compact-difference
 * curvilinear compressible Navier-Stokes solver, exposed as a plain C ABI.
 *
 * Every entry point returns 0 on success or a negative MC3D_E* code; the text
 * of the last failure on the calling thread is available from
 * mc3d_last_error().  The library never aborts the process.
 *
 * Conventions
 *  - Fields are device-resident FP64 arrays of Nx*Ny*Nz values.  The solver
 *    keeps one layout, x fastest: index = (k*Ny + j)*Nx + i, and solves the
 *    strided lines in place.  The operators and SGS models also take the
 *    reference's other two pencil layouts (src/Macros.hpp:16-18), selected by
 *    a layout argument:
 *      MC3D_XPENCIL  index = (k*Ny + j)*Nx + i   (x fastest, then y, z)
 *      MC3D_YPENCIL  index = (i*Nz + k)*Ny + j   (y fastest, then z, x)
 *      MC3D_ZPENCIL  index = (j*Nx + i)*Nz + k   (z fastest, then x, y)
 *    The reference calls calc1stDerivField / filterField of direction x, y, z
 *    with fields in the X, Y, Z pencil respectively (src/Pade6.cpp:6-43), and
 *    its SGS models work on Y pencils (src/AbstractSGS.cpp, FOR_XYZ_YPEN).
 *  - "Lines" of direction d are enumerated by the two other indices in the
 *    order of the layout, the faster one fastest.  x-fastest: x-lines
 *    q = k*Ny + j, y-lines q = k*Nx + i, z-lines q = j*Nx + i.  In each
 *    direction's own pencil this is the reference's order of its per-line
 *    wrap arrays (x: k*Ny + j, y: i*Nz + k, z: j*Nx + i;
 *    src/Pade6.cpp:56,76,95, src/AbstractSingleBlockMesh.cpp:151,61,242).
 *    Per-line side arrays (transformed-periodic wrap values) use this order.
 *  - Calls are asynchronous on the given stream (a hipStream_t passed as
 *    void*; NULL is the default stream).  The caller owns all field memory;
 *    handles own their precomputed factor tables.
 *  - Direction ids: 0 = xi (x), 1 = eta (y), 2 = zeta (z).
 *    Boundary types: MC3D_PERIODIC, MC3D_DIRICHLET.
 */
#ifndef MC3D_H
#define MC3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MC3D_OK 0
#define MC3D_EINVAL (-1)      /* bad argument or unsupported size */
#define MC3D_EHIP (-2)        /* HIP runtime failure */
#define MC3D_ENOMEM (-3)      /* allocation failure */
#define MC3D_ENUMERIC (-4)    /* factorisation failure (zero pivot) */
#define MC3D_ECONFIG (-5)     /* options file / configuration error */

enum { MC3D_PERIODIC = 0, MC3D_DIRICHLET = 1 };
/* field layouts: the reference's X, Y and Z pencils (src/Macros.hpp:16-18) */
enum { MC3D_XPENCIL = 0, MC3D_YPENCIL = 1, MC3D_ZPENCIL = 2 };
/* derivative schemes: src/Options.hpp:22 (FDType) */
enum { MC3D_CD2 = 0, MC3D_PADE6 = 1, MC3D_PENTA10 = 2 };
/* filter kinds: src/Options.hpp:23 (FilterType) plus the two DSM test filters */
enum { MC3D_COMPACT8 = 0, MC3D_COMPACT10 = 1, MC3D_COMMN5PADE = 2, MC3D_COMMN5EXP = 3 };

typedef struct mc3d_lineop mc3d_lineop; /* opaque derivative or filter operator */

const char *mc3d_last_error(void);
int mc3d_device_count(int *count);

/* ---------------------------------------------------------------------------
 * First-derivative operator along one direction.
 * Replaces: AbstractDerivatives (src/AbstractDerivatives.hpp:13-53) and its
 * constructors Pade6(Domain*, BCType, Direct) src/Pade6.hpp:39,
 * Penta10(...) src/Penta10.hpp:38, CD2(...) src/CD2.hpp:25.
 * dims = {Nx, Ny, Nz}; dd = grid spacing along dir in computational space;
 * layout = MC3D_XPENCIL / YPENCIL / ZPENCIL, the layout of the fields the
 * operator is applied to (the reference's: the pencil of dir).
 */
int mc3d_deriv_create(int scheme, int dir, int bc, const int dims[3], double dd, int layout,
                      mc3d_lineop **out);

/* out[f] = d(in[f])/d(xi_dir) for f < nfields (nfields <= 16; in and out
 * must not alias).  Replaces calc1stDerivField(in, out)
 * src/AbstractDerivatives.hpp:41 / src/Pade6.cpp:3-45. */
int mc3d_deriv_apply(const mc3d_lineop *op, const double *const *in, double *const *out,
                     int nfields, void *stream);

/* Transformed-periodic variant: the wrapped stencil values are taken from
 * wrap[0..5] = {Nm3, Nm2, Nm1, Np1, Np2, Np3} (one value per line, in the line
 * order above; entries beyond the scheme's half-width may be NULL).
 * Replaces calc1stDerivField_TPB(...) src/AbstractDerivatives.hpp:45-47. */
int mc3d_deriv_apply_tpb(const mc3d_lineop *op, const double *in, double *out,
                         const double *const wrap[6], void *stream);

/* ---------------------------------------------------------------------------
 * Compact / test filter along one direction.  bc3 = boundary types of the
 * x, y, z directions: lines lying on a Dirichlet face of another direction are
 * copied unfiltered (src/Compact8Filter.cpp:142-230).
 * Replaces: AbstractFilter (src/AbstractFilter.hpp:13-35) and
 * Compact8Filter(alphaF, Domain*, BC*, BCType, Direct) src/Compact8Filter.hpp:28,
 * Compact10Filter src/Compact10Filter.hpp:30, CommN5PadeTestFilter
 * src/CommN5PadeTestFilter.hpp:28, CommN5ExpTestFilter src/CommN5ExpTestFilter.hpp:24.
 */
int mc3d_filter_create(int kind, double alphaF, int dir, const int bc3[3], const int dims[3], int layout,
                       mc3d_lineop **out);

/* Replaces filterField(in, out) src/AbstractFilter.hpp:34.  in may equal out. */
int mc3d_filter_apply(const mc3d_lineop *op, const double *const *in, double *const *out,
                      int nfields, void *stream);

void mc3d_lineop_destroy(mc3d_lineop *op);

/* Diagnostics of an operator's partitioned solve: chunks per line, rows per
 * chunk, lines per workgroup. */
int mc3d_lineop_info(const mc3d_lineop *op, int *chunks, int *chunk_rows, int *lines_per_block);

/* Diagnostic: name of the kernel instantiation a plain (non-TPB) application
 * of the operator launches, as profilers show it, e.g.
 * "lineop_strided<2, 3, 32, true>" (half bandwidth, stencil half width, rows
 * per chunk, periodic).  NULL for a NULL operator. */
const char *mc3d_lineop_kernel(const mc3d_lineop *op);

/* ---------------------------------------------------------------------------
 * CFL time step over caller-owned fields.
 * Replaces CurvilinearCSolver::calcDtFromCFL (src/CurvilinearCSolver_Core.cpp:189-237):
 *   maxspeed = max over points of J11 (|U| + c)/dx + J22 (|V| + c)/dy + J33 (|W| + c)/dz,
 *   dt = cfl / maxspeed (TIMESTEPPING CONST_CFL; for CONST_DT the reference
 *   sets CFL = dt * maxspeed, which the caller forms from out[1]).
 * fields[7] = {U, V, W, sos, J11, J22, J33}, device, npts * members doubles
 * each, in any layout (the maximum does not depend on it); members boxes of
 * npts points stacked one after the other each get their own maximum and dt:
 * out[2 b] = dt, out[2 b + 1] = maxspeed of box b (device doubles).
 * spacing = {dx, dy, dz} of the computational grid.  The per-workgroup
 * partial maxima (up to 2048 doubles per box) are allocated for the call,
 * and the call waits for the stream before it returns.
 */
int mc3d_cfl_dt(const double *const fields[7], long npts, int members, const double spacing[3], double cfl,
                double *out, void *stream);

/* mc3d_cfl_dt without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as mc3d_cfl_dt; the partial maxima go to a scratch
 * array the library keeps on each GPU for the life of the process (1 MiB,
 * 2048 doubles per box), so members <= 64, else MC3D_EINVAL.  The call only
 * queues its two kernels on `stream` and returns without waiting; out is
 * ready when the stream reaches it.  Calls that overlap on one GPU (two
 * streams, two threads) share that scratch array and race. */
int mc3d_cfl_dt_noalloc(const double *const fields[7], long npts, int members, const double spacing[3],
                        double cfl, double *out, void *stream);

/* ---------------------------------------------------------------------------
 * SGS models over caller-owned device fields.
 * Replaces: AbstractSGS (src/AbstractSGS.hpp:9-42) and its subclasses
 * VremanSGS(cs) src/VremanSGS.hpp:15-41, DSMSGS(cs) src/DSMSGS.hpp:34-111, and
 * the learned closure beside them; doAveraging (src/AbstractSGS.cpp:13-164)
 * runs inside each call as the reference's models call it.
 * options_text: the [DOMAIN] NX NY NZ, [BC] BCX/BCY/BCZ types and [LES] keys
 * (LESMODEL VREMAN | DSM | LEARNED, LESAVERAGING, DSMTESTFILTER, DSMUSETAUKK,
 * DSMCOEFFRANGE, CLOSUREFILE, CLOSUREFALLBACK ...) of the reference's options
 * file (other sections may be present and are ignored).
 * layout: of every field passed (the reference's SGS fields are Y pencils,
 * MC3D_YPENCIL).  J: device Jacobian field (the mesh's J,
 * src/AbstractSingleBlockMesh.cpp:1297-1308), kept by pointer for the
 * filter width (dx dy dz / J)^(1/3).  Work is queued on `stream` (a
 * hipStream_t, NULL: the default stream); the model's work fields (up to about
 * 40 fields for the DSM) are allocated at creation.
 * gradU[3*i + j] = d u_i / d x_j (Cartesian, the reference's gradU[i][j],
 * src/CurvilinearCSolver_Core.cpp:548-561); gradT[j] = dT/dx_j.
 * Outputs are caller-owned device fields.
 */
typedef struct mc3d_sgs mc3d_sgs;
typedef struct mc3d_closure mc3d_closure;

int mc3d_sgs_create(const char *options_text, int layout, const double *J, void *stream, mc3d_sgs **out);
void mc3d_sgs_destroy(mc3d_sgs *s);
/* VremanSGS::getSGSViscosity (src/VremanSGS.hpp:43-96): mu_sgs, then the
 * LESAVERAGING smoothing of mu_sgs (Compact10 alpha_f -0.49 filters for LOCAL). */
int mc3d_sgs_vreman(mc3d_sgs *s, const double *const gradU[9], const double *rho, double *mu_sgs);
/* DSMSGS::calcMuSGSTaukk (src/DSMSGS.hpp:124-343): mu_sgs and tau_kk (zero
 * unless DSMUSETAUKK).  The caller applies tau_kk to p as the solver does
 * (src/CurvilinearCSolver_Core.cpp:571-584) before forming gradT. */
int mc3d_sgs_dsm_mu_taukk(mc3d_sgs *s, const double *const gradU[9], const double *rho, const double *rhoU,
                          const double *rhoV, const double *rhoW, double *mu_sgs, double *taukk);
/* DSMSGS::calcKSGS (src/DSMSGS.hpp:345-444): k_sgs; uses the test-filtered
 * fields of the preceding mc3d_sgs_dsm_mu_taukk call (same state). */
int mc3d_sgs_dsm_ksgs(mc3d_sgs *s, const double *const gradT[3], const double *rho, const double *rhoU,
                      const double *rhoV, const double *rhoW, const double *T, double *k_sgs);
/* DSMCOEFFRANGE: C, CI and Prt of the last DSM calls
 * (src/DSMSGS.hpp:104-109,330-341,435-442) copied into caller fields (stream
 * ordered); NULL outputs are skipped. */
int mc3d_sgs_coefficients(mc3d_sgs *s, double *C, double *CI, double *Prt);
/* Learned closure (LESMODEL LEARNED; network from CLOSUREFILE or
 * mc3d_sgs_set_closure): mu_sgs and k_sgs of the network, with the DSM
 * fallback of CLOSUREFALLBACK, which needs gradT, rhoU, rhoV, rhoW and T
 * (they may be NULL with CLOSUREFALLBACK FALSE).  Waits for the stream
 * (the fallback decision reads the out-of-range count). */
int mc3d_sgs_set_closure(mc3d_sgs *s, const mc3d_closure *c);
int mc3d_sgs_learned(mc3d_sgs *s, const double *const gradU[9], const double *const gradT[3], const double *rho,
                     const double *rhoU, const double *rhoV, const double *rhoW, const double *T, double *mu_sgs,
                     double *k_sgs);
int mc3d_sgs_closure_stats(mc3d_sgs *s, long *evaluations, long *outside_points, long *fallback_stages,
                           long *fallback_points);

/* ---------------------------------------------------------------------------
 * Whole solver step (CurvilinearCSolver + RK + SGS + wall and sponge
 * boundaries, src/CurvilinearCSolver_BC.cpp), configured by the
 * reference's options file text (MPI_3DCompact.in: sections [TIMESTEPPING],
 * [BC], [DOMAIN], [SOLVER], [RESTART], [STATS], [LES], ...).
 * Replaces the construction sequence of src/MPI_3DCompact.cpp:34-316
 * (Options, Domain, CurvilinearCSolver, AbstractRK selection) and
 * AbstractRK::executeSolverLoop (src/AbstractRK.hpp:14).
 * The solver owns its device fields.  Host arrays are x-fastest.
 */
typedef struct mc3d_solver mc3d_solver;

int mc3d_solver_create(const char *options_text, mc3d_solver **out);
/* Parse and validate an options text as mc3d_solver_create does, without a
 * device (host only): the reference's Options(string) constructor
 * (src/Options.hpp:131-426) and bcValidation (src/Options.cpp:178-268),
 * with MC3D_ECONFIG and the message in mc3d_last_error() where the reference
 * prints and calls MPI_Abort.  On success the parsed values are written to
 * buf as "SECTION.KEY = value" lines (enums by name, booleans TRUE / FALSE,
 * numbers with 17 significant digits), truncated to buflen - 1 characters;
 * *needed (if not NULL) receives the full length plus one.  With a restart
 * (FROMRESTART / ONLYGRIDFROMRESTART) the grid is read from the dump only by
 * mc3d_solver_create. */
int mc3d_options_check(const char *options_text, char *buf, int buflen, int *needed);
/* metrics[10] = {J, J11, J12, J13, J21, J22, J23, J31, J32, J33} (host arrays),
 * the fields produced by src/AbstractSingleBlockMesh.cpp:23-1439. */
int mc3d_solver_set_metrics(mc3d_solver *s, const double *const metrics[10]);
/* Metric terms of a single-block mesh on the GPU, without a solver:
 * AbstractSingleBlockMesh::solveForJacobians (src/AbstractSingleBlockMesh.cpp:23-1309).
 * fd[d] and bc[d] select the derivative scheme and boundary of direction d;
 * trans[3*d+e] is component e of the coordinate jump across periodic
 * direction d (PERIODICDISP, :2019-2085; NULL: none).  x, y, z (inputs) and
 * out[10] = {J, J11 .. J33} are device fields of dims[0]*dims[1]*dims[2]
 * doubles, x fastest.  Computational coordinates run over [0, 1].  Work is
 * queued on stream (a hipStream_t, NULL: default) and finished on return;
 * scratch of 39 fields is allocated for the call. */
int mc3d_mesh_metrics(const int fd[3], const int bc[3], const int dims[3], const double trans[9],
                      const double *x, const double *y, const double *z, double *const out[10],
                      void *stream);
/* Sponge strength sigma (host, x fastest) for a solver with SPONGE faces, as
 * SpongeBC builds it (src/SpongeBC.cpp:60-260); mc3d_solver_set_mesh computes
 * it from the mesh itself.  avg (host rho, rhoU, rhoV, rhoW, rhoE; NULL: the
 * initial conserved fields, src/CurvilinearCSolver_BC.cpp:324-345) are the
 * states the sponge relaxes towards.  Required before stepping when a face is
 * a SPONGE and no mesh was given. */
int mc3d_solver_set_sponge(mc3d_solver *s, const double *sigma, const double *const avg[5]);
/* Form the ten metric fields on the GPU from host mesh coordinates x, y, z
 * (x fastest), as AbstractSingleBlockMesh::solveForJacobians does
 * (src/AbstractSingleBlockMesh.cpp:23-1309): conservative form, periodic
 * directions through the transformed-periodic derivative with the wrap values
 * shifted by [BC] PERIODICDISP<d>_<e> (handlePeriodicStuff, :2019-2085).  The
 * computational coordinates run over [0, 1] in each direction.  Used instead of
 * mc3d_solver_set_metrics; CD2 on a periodic direction is MC3D_ECONFIG (the
 * reference has no such path either, :94-97).  With SPONGE faces the sponge
 * strength is formed from the same mesh (see mc3d_solver_set_sponge). */
int mc3d_solver_set_mesh(mc3d_solver *s, const double *x, const double *y, const double *z);
/* primitive initial condition; performs setInitialConditions
 * (src/CurvilinearCSolver_BC.cpp:3-63). */
int mc3d_solver_set_initial(mc3d_solver *s, const double *rho0, const double *U0,
                            const double *V0, const double *W0, const double *p0);
/* advance nsteps full time steps (one executeSolverLoop iteration each) */
int mc3d_solver_step(mc3d_solver *s, int nsteps);
/* run until MAXTIMESTEP / MAXTIME, like executeSolverLoop */
int mc3d_solver_run(mc3d_solver *s);
/* copy a named field to host memory / get its device address.  Names:
 * rho rhoU rhoV rhoW rhoE (the *1 registers), rhok.. rhoEk, rhok2.. rhoEk2,
 * rho2.. rhoE2, U V W p T mu sos Ucurv Vcurv Wcurv mu_sgs k_sgs taukk,
 * C CI Prt (DSMCOEFFRANGE), J J11 .. J33, and dU0 dU1 dU2 dV0 .. dW2 dT0 ..
 * dT2, the computational derivatives (d/dxi, d/deta, d/dzeta) of the last
 * evaluated stage. */
int mc3d_solver_get_field(mc3d_solver *s, const char *name, double *host_out);
/* FROMRESTART ([RESTART] FILENAME): the conserved fields of that SolutionDump
 * file become the initial state through primitive variables, as
 * src/MPI_3DCompact.cpp:214-271 does (call after set_metrics, instead of
 * set_initial).  With FROMRESTART or ONLYGRIDFROMRESTART, mc3d_solver_create
 * takes the grid (and, for a restart, step and time) from the file header
 * (src/MPI_3DCompact.cpp:80-113).  The momentum k2 registers are not in the
 * file, so a restarted run starts them at zero. */
int mc3d_solver_load_restart(mc3d_solver *s);
/* Write a SolutionDump of the current state (dumpSolution,
 * src/CurvilinearCSolver_Core.cpp:1333-1409) with host mesh coordinates
 * x, y, z (NULL entries are written as zeros).  path NULL: "SolutionDump.<step>"
 * in the output directory (or the working directory). */
int mc3d_solver_write_dump(mc3d_solver *s, const char *path, const double *x, const double *y,
                           const double *z);
/* Enable the reference's automatic dumps: every DUMPSTEP steps and at the end
 * of the run (src/CurvilinearCSolver.hpp:367-369, src/CurvilinearCSolver_Core.cpp:1822)
 * into dir (NULL: working directory), with the given host mesh (copied). */
int mc3d_solver_set_output(mc3d_solver *s, const char *dir, const double *x, const double *y,
                           const double *z);
/* The metric identities I_k = dJ_1k/dxi + dJ_2k/deta + dJ_3k/dzeta (k = 1..3)
 * of the current metrics with the solver's first derivatives, as
 * range = {min I_1, max I_1, min I_2, max I_2, min I_3, max I_3}: the check
 * AbstractSingleBlockMesh::solveForJacobians prints
 * (src/AbstractSingleBlockMesh.cpp:1336-1393; set_mesh prints it too). */
int mc3d_solver_metric_identity(mc3d_solver *s, double range[6]);
/* PNG image output of a field plane (PngWriter, src/PngWriter.hpp:83-190;
 * CurvilinearCSolver::addImageOutput, src/CurvilinearCSolver_Core.cpp:1438):
 * every `interval` steps (0: the first step only) the index plane nearest to
 * `fraction` of the grid along `plane` (0, 1, 2) is gathered on the device
 * and drawn as width x height RGB pixels, colour range [vmin, vmax] when
 * has_range is nonzero, else the plane's extremes; colormap 0 GREYSCALE,
 * 1 RAINBOW, 2 BWR.  Written as <var_name>.<step, 6 digits>.png into the
 * output directory (mc3d_solver_set_output; default the working directory).
 * The reference interpolates the curvilinear mesh onto a Cartesian pixel
 * grid; here pixels sample the structured index plane (the same picture for
 * a Cartesian mesh). */
int mc3d_solver_add_image(mc3d_solver *s, int interval, int width, int height, const char *field,
                          const char *var_name, int plane, double fraction, int has_range, double vmin,
                          double vmax, int colormap);
/* The index plane `index` normal to `dir` of a field, into host memory of
 * n1 * n2 doubles, a1 fastest: dir 0 -> (a1, a2) = (j, k), 1 -> (k, i),
 * 2 -> (i, j).  Synchronous. */
int mc3d_solver_slice(mc3d_solver *s, const char *field, int dir, int index, double *host_out);
/* Render an (n2, n1) plane (a1 fastest) as the image outputs do and write it
 * as a PNG file; host only (no device needed). */
int mc3d_image_write(const char *path, const double *plane, int n1, int n2, int width, int height,
                     int has_range, double vmin, double vmax, int colormap);
/* Decaying-HIT statistics of the current state, the turbdata.out hook of
 * tests/hit/Samtaney_D4/MPI_3DCompact.cpp:380-556 (device reductions):
 * stats = {time, tau, Re_lambda, M_t, epsilon = <2 mu S*_ij du_i/dx_j> / <rho>,
 * u', Kolmogorov length, rho', dilatation', vorticity', <rho u.u>}.
 * Overwrites the velocity-derivative work fields.  Synchronous. */
int mc3d_solver_hit_stats(mc3d_solver *s, double stats[11]);
/* The HIT driver's fullStepTemporalHook
 * (tests/hit/Samtaney_D4/MPI_3DCompact.cpp:372-557): after every full step
 * whose number is a multiple of `every` (the driver's 25), append one line of
 * the 11 statistics above to `path` (the driver's turbdata.out), each written
 * with 17 significant digits and followed by a space, as the driver's
 * ofstream with precision(17) writes them.  NULL or "" switches it off. */
int mc3d_solver_set_turbdata(mc3d_solver *s, const char *path, int every);
int mc3d_solver_field_device(mc3d_solver *s, const char *name, double **dev_ptr);
/* Running statistics ([STATS] VELOCITYSTATS / THERMOSTATS, src/Stats.hpp):
 * their accumulated weights (the time they cover); the fields are read with
 * mc3d_solver_get_field as UAVG VAVG WAVG URMS VRMS WRMS UVREY UWREY VWREY and
 * RHOAVG RHORMS PAVG PRMS TAVG TRMS.  Every SolutionDump is accompanied by
 * VelocityStats.<step> / ThermoStats.<step> (weight, then those fields) and,
 * with SPONGE faces, SpongeAvgDump.<step> (the five sponge averages). */
int mc3d_solver_stats_weights(mc3d_solver *s, double *velocity, double *thermo);
int mc3d_solver_scalars(mc3d_solver *s, double *time, double *dt, int *timestep);
/* the solver's HIP stream (as void*), for timing with events */
int mc3d_solver_stream(mc3d_solver *s, void **stream);
int mc3d_solver_sync(mc3d_solver *s);
void mc3d_solver_destroy(mc3d_solver *s);

/* Per-kernel timing of the step with HIP events on the solver stream (the
 * counterpart of SOLVER.USETIMING, src/CurvilinearCSolver_Core.cpp:6,182-185).
 * enable != 0 starts recording (and clears previous totals); the report is a
 * JSON array of {"kernel", "count", "ms", "bytes"} objects, bytes being the
 * algorithmic bytes (distinct full fields read + written) of those launches.
 * Returns MC3D_EINVAL if buf is too small (the required size goes to *needed). */
int mc3d_solver_timing(mc3d_solver *s, int enable);
int mc3d_solver_timing_report(mc3d_solver *s, char *buf, int buflen, int *needed);

/* ---------------------------------------------------------------------------
 * Learned SGS closure (no counterpart in the reference: it replaces the
 * dynamic procedure of DSMSGS, src/DSMSGS.hpp:124-444, behind the same
 * AbstractSGS interface, src/AbstractSGS.hpp:9-42, selected by
 * [LES] LESMODEL = LEARNED with CLOSUREFILE = <network file> and
 * CLOSUREFALLBACK = TRUE|FALSE).
 *
 * Per point, with the Cartesian velocity gradient A = grad u:
 *   features f = (theta/a, (s2 - w2)/a^2, s3/a^3, sw2/a^3, s2w2/a^4, log a,
 *                 log Delta), theta = tr A, S and W the symmetric and
 *                 antisymmetric parts, s2 = S:S, w2 = W:W, s3 = tr S^3,
 *                 sw2 = tr S W^2, s2w2 = tr S^2 W^2, a^2 = s2 + w2,
 *                 Delta = (dxi deta dzeta / J)^(1/3) (src/DSMSGS.hpp:332-333);
 *   network  7 -> 64 -> 64 -> 64 -> 2, tanh-form GELU, on standardised inputs
 *            (x - in_mean) / in_std, BF16 matrix products with FP32 accumulation;
 *   outputs  C = y0 out_std0 + out_mean0, D = y1 out_std1 + out_mean1 (D = C / Pr_t);
 *            mu_sgs = rho C Delta^2 |S|, k_sgs = cp D rho Delta^2 |S|
 *            (the DSM's own forms, src/DSMSGS.hpp:319,406-408,440).
 * A point whose raw features leave [lo, hi] is evaluated on inputs clamped to
 * the range; when more than CLOSUREFALLBACKFRACTION of the points of a stage
 * are out and the fallback is on, the DSM runs on that stage and its mu_sgs
 * and k_sgs replace the network's at those points.
 *
 * Parameters: 8962 FP32 values, W1[64][7] b1[64] W2[64][64] b2[64] W3[64][64]
 * b3[64] W4[2][64] b4[2].  Training data: X[n][7] standardised inputs and
 * T[n][2] standardised targets, FP32 device arrays.
 * ------------------------------------------------------------------------- */
/* weights U(-s, s), s = sqrt(3 / fan_in), zero biases; identity scaling */
int mc3d_closure_create(unsigned long long seed, mc3d_closure **out);
void mc3d_closure_destroy(mc3d_closure *c);
int mc3d_closure_set_params(mc3d_closure *c, const float *params);   /* host[8962]; resets Adam */
int mc3d_closure_get_params(const mc3d_closure *c, float *params);   /* host[8962] */
int mc3d_closure_set_scaling(mc3d_closure *c, const float in_mean[7], const float in_std[7],
                             const float out_mean[2], const float out_std[2], const float lo[7],
                             const float hi[7]);
int mc3d_closure_get_scaling(const mc3d_closure *c, float in_mean[7], float in_std[7],
                             float out_mean[2], float out_std[2], float lo[7], float hi[7]);
/* gradient of L = sum (y - T)^2 / (2 n) over all n samples into grad
 * (device[8962]); *loss = L.  Synchronous. */
int mc3d_closure_gradient(mc3d_closure *c, const float *X, const float *T, long n, float *grad,
                          double *loss, void *stream);
/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the given gradient */
int mc3d_closure_adam_step(mc3d_closure *c, const float *grad, float lr, void *stream);
/* shuffled minibatch Adam; epoch_losses[e] (host, may be NULL) = mean loss of epoch e */
int mc3d_closure_train(mc3d_closure *c, const float *X, const float *T, long n, int epochs, int batch,
                       float lr, unsigned long long seed, double *epoch_losses, void *stream);
/* Y[n][2] = network(X[n][7]) on standardised values (asynchronous) */
int mc3d_closure_infer(const mc3d_closure *c, const float *X, float *Y, long n, void *stream);
int mc3d_closure_save(const mc3d_closure *c, const char *path);
int mc3d_closure_load(mc3d_closure *c, const char *path);
/* copy the network and its scaling into a LESMODEL LEARNED solver */
int mc3d_solver_set_closure(mc3d_solver *s, const mc3d_closure *c);
/* training pairs from a LESMODEL DSM solver with DSMCOEFFRANGE TRUE: the DSM
 * procedure runs on the current state (which is left unchanged) and, for
 * i < count and p = offset + i * stride, X[i][7] = raw features and
 * T[i] = (C, C / Pr_t) of the DSM at p (device FP32).  Synchronous. */
int mc3d_solver_sample_closure(mc3d_solver *s, float *X, float *T, long count, long offset,
                               long stride);
/* derivatives and SGS fields (mu_sgs, k_sgs) of the current state without
 * advancing it (NOMODEL, VREMAN, LEARNED; the DSM modifies the pressure) */
int mc3d_solver_eval_sgs(mc3d_solver *s);
/* closure evaluations (stages), points whose inputs left the trusted range
 * (clamped to it), stages that fell back to the DSM (more than
 * [LES] CLOSUREFALLBACKFRACTION of the points out, default 1e-3), and the
 * points of those stages */
int mc3d_solver_closure_stats(mc3d_solver *s, long *evaluations, long *outside_points,
                              long *fallback_stages, long *fallback_points);

/* ---------------------------------------------------------------------------
 * SolutionDump files (host only, no GPU): five doubles {timeStep, time, Nx,
 * Ny, Nz} then x, y, z, rho, rhoU, rhoV, rhoW, rhoE as Nx*Ny*Nz doubles each,
 * x fastest, native byte order (src/CurvilinearCSolver_Core.cpp:1333-1409;
 * read back by src/MPI_3DCompact.cpp:80-126,214-271).  fields[8] entries may
 * be NULL (skipped on read, zeros on write). */
int mc3d_dump_read_header(const char *path, int *timestep, double *time, int dims[3]);
int mc3d_dump_read(const char *path, double *const fields[8]);
int mc3d_dump_write(const char *path, int timestep, double time, const int dims[3],
                    const double *const fields[8]);

/* Diagnostic (no counterpart in the reference): builds, on the host only,
 * the partition tables an operator would upload, for verification of the
 * factorisation without a GPU.  cls 0 = derivative (kind = scheme), 1 = filter
 * (kind = filter kind); mm = rows per chunk (16 or 32, 0 = the choice the
 * operators make).  sizes[11] receives {N, W, H, P, NI, nbl, nbr, int_off, MM,
 * dense_ok, gb} (gb: the interface band in chunks the fast kernels use, G
 * blocks farther from a chunk being below 1e-17 of its largest); each
 * non-NULL table pointer must have room for the sizes
 * documented in mpi_3dcompact_amd/csrc/lineop.h (lu: P*MM*5, sp: P*MM*2W,
 * G: P*2W*NI, rows: (nbl+nbr)*19, drows: (nbl+nbr)*17, interior: 11,
 * cstart/clen: P). */
int mc3d_plan_tables(int cls, int kind, int bc, int n, double dd, double alphaF, int mm,
                     int sizes[11], double *lu, double *sp, double *G, double *rows,
                     double *drows, double *interior, int *cstart, int *clen);

/* ---------------------------------------------------------------------------
 * Ensembles (no counterpart in the reference, which runs one box per job):
 * many independent periodic boxes of uneven sizes, each a run of the solver
 * configured by the options text (schemes, filter, RK, CFL, LES model NOMODEL or
 * VREMAN with LESAVERAGING NONE / LOCAL), advanced together.  Members of equal
 * size share every launch (stacked along z in device memory); each keeps its
 * own dt from its own CFL reduction, its own time and its own MU_REF.  From
 * the second step on, one step of the whole ensemble is a replayed HIP graph.
 * A member's results equal those of mc3d_solver on the same inputs with the
 * same line-operator launch shapes (bit for bit), and differ by rounding when
 * tuning chose other shapes.  Host arrays are x fastest, one member's box.
 */
typedef struct mc3d_ensemble mc3d_ensemble;

/* dims: 3 per member (nx, ny, nz; each at least 8); mu_ref: per member (NULL:
 * SOLVER.MU_REF of the options).  The options' NX/NY/NZ are not used. */
int mc3d_ensemble_create(const char *options_text, int members, const int *dims, const double *mu_ref,
                         mc3d_ensemble **out);
int mc3d_ensemble_set_metrics(mc3d_ensemble *e, int member, const double *const metrics[10]);
/* before the first step only */
int mc3d_ensemble_set_initial(mc3d_ensemble *e, int member, const double *rho0, const double *U0,
                              const double *V0, const double *W0, const double *p0);
/* Measure and select, per box size and direction, the launch shape of the
 * derivative and filter operators (rows per chunk, kernel family, lines per
 * workgroup).  cache_path (NULL: none) is a text file of earlier results, read
 * first and written back with the new ones.  Synchronous. */
int mc3d_ensemble_tune(mc3d_ensemble *e, const char *cache_path);
int mc3d_ensemble_step(mc3d_ensemble *e, int nsteps);
/* graphs on (default) or off: off issues every step's launches directly */
int mc3d_ensemble_use_graphs(mc3d_ensemble *e, int on);
/* same names as mc3d_solver_get_field */
int mc3d_ensemble_get_field(mc3d_ensemble *e, int member, const char *name, double *host_out);
int mc3d_ensemble_scalars(mc3d_ensemble *e, int member, double *time, double *dt, int *timestep);
/* the ensemble's HIP stream (as void*): every step's work is ordered on it */
int mc3d_ensemble_stream(mc3d_ensemble *e, void **stream);
int mc3d_ensemble_sync(mc3d_ensemble *e);
/* JSON: buckets (box size, members, kernels of the derivative and filter
 * operators per direction, tuning results), graphs captured, graph launches.
 * MC3D_EINVAL if buf is too small (the required size goes to *needed). */
int mc3d_ensemble_info(mc3d_ensemble *e, char *buf, int buflen, int *needed);
void mc3d_ensemble_destroy(mc3d_ensemble *e);

#ifdef __cplusplus
}
#endif
#endif
