/*
 * CPU oracle of one CurvilinearCSolver time step (TEST INFRASTRUCTURE ONLY,
 * see oracle.h).  Restates src/CurvilinearCSolver_Core.cpp:189-1198, the RK
 * schemes (src/TVDRK3.hpp, src/RK4.hpp, src/KenRK4.hpp, src/LSLDDRK4.hpp),
 * src/IdealGas.hpp, src/VremanSGS.hpp, src/DSMSGS.hpp and
 * src/AbstractSGS.cpp, with the boundary handling of
 * src/CurvilinearCSolver_BC.cpp:3-842 (walls, sponge).
 */
#ifndef MC3D_ORC_SOLVER_H
#define MC3D_ORC_SOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TVDRK3 = 0, ORC_RK4 = 1, ORC_KENRK4 = 2, ORC_LSLDDRK4 = 3 };
enum { ORC_NOMODEL = 0, ORC_VREMAN = 1, ORC_DSM = 2 };
/* src/Options.hpp:25 StatsAvgType */
enum { ORC_AVG_NONE = 0, ORC_AVG_XI1 = 1, ORC_AVG_XI2 = 2, ORC_AVG_XI3 = 3, ORC_AVG_LOCAL = 4,
       ORC_AVG_GLOBAL = 5 };
enum { ORC_CONST_DT = 0, ORC_CONST_CFL = 1 };

typedef struct {
    int Nx, Ny, Nz;
    int bc[3];              /* ORC_PERIODIC / ORC_DIRICHLET per direction */
    int fd[3];              /* derivative scheme per direction */
    int filter;             /* ORC_COMPACT8 / ORC_COMPACT10 */
    double alphaF;
    int rk;                 /* ORC_TVDRK3 .. ORC_LSLDDRK4 */
    int les;                /* ORC_NOMODEL / ORC_VREMAN / ORC_DSM */
    int les_avg;            /* ORC_AVG_* */
    int dsm_test_filter;    /* 1: commutative Pade, 2: explicit */
    int dsm_use_taukk;
    int dsm_coeff_range;
    double mu_ref;
    int ts_type;            /* ORC_CONST_DT / ORC_CONST_CFL */
    double cfl, dt;
    int filter_step;
    int substep_filtering;
    double dx, dy, dz;      /* computational spacings (src/Domain.hpp:48-81) */
    int timestep0;
    double time0;
    int face[6];            /* face kinds X0 X1 Y0 Y1 Z0 Z1, ORC_FACE_* (0: periodic) */
    /* running statistics (src/Stats.hpp, src/Stats.cpp): STATS.VELOCITYSTATS,
     * THERMOSTATS, STATSAVGTYPE (ORC_AVG_*), INTERVAL (0: none) */
    int velocity_stats, thermo_stats, stats_avg, stats_interval;
} orc_config;

/* face kinds, in the order of src/Options.hpp BCKind */
enum { ORC_FACE_PERIODIC = 0, ORC_FACE_INTERNALLY_PERIODIC = 1, ORC_FACE_SPONGE = 2,
       ORC_FACE_ADIABATIC_WALL = 3, ORC_FACE_CONST_T_WALL = 4, ORC_FACE_MOVING_ADIABATIC_WALL = 5,
       ORC_FACE_INLET = 6 };

/* field ids for orc_solver_field */
enum {
    ORC_F_RHO = 0, ORC_F_RHOU, ORC_F_RHOV, ORC_F_RHOW, ORC_F_RHOE,   /* *1 registers */
    ORC_F_U, ORC_F_V, ORC_F_W, ORC_F_P, ORC_F_T, ORC_F_MU, ORC_F_SOS,
    ORC_F_MUSGS, ORC_F_TAUKK, ORC_F_KSGS, ORC_F_C, ORC_F_CI, ORC_F_PRT,
    ORC_F_RHOK, ORC_F_RHOUK, ORC_F_RHOVK, ORC_F_RHOWK, ORC_F_RHOEK,
    ORC_F_RHOK2, ORC_F_RHOUK2, ORC_F_RHOVK2, ORC_F_RHOWK2, ORC_F_RHOEK2,
    ORC_F_UAVG, ORC_F_VAVG, ORC_F_WAVG, ORC_F_URMS, ORC_F_VRMS, ORC_F_WRMS,
    ORC_F_UVREY, ORC_F_UWREY, ORC_F_VWREY,
    ORC_F_RHOAVG, ORC_F_RHORMS, ORC_F_PAVG, ORC_F_PRMS, ORC_F_TAVG, ORC_F_TRMS,
    ORC_F_RHO2,
    ORC_F_COUNT
};

typedef struct orc_solver orc_solver;

/* metrics[10] = {J, J11, J12, J13, J21, J22, J23, J31, J32, J33};
 * initial primitive fields rho0, U0, V0, W0, p0.  Arrays are x-fastest. */
orc_solver *orc_solver_create(const orc_config *cfg, const double *const metrics[10],
                              const double *rho0, const double *U0, const double *V0,
                              const double *W0, const double *p0);
/* advance n full time steps (executeSolverLoop iterations) */
int orc_solver_steps(orc_solver *s, int n);
const double *orc_solver_field(const orc_solver *s, int id);
double orc_solver_time(const orc_solver *s);
/* accumulated statistics weights (sum of the time increments): which 0 velocity, 1 thermo */
double orc_solver_stats_weight(const orc_solver *s, int which);
double orc_solver_dt(const orc_solver *s);
int orc_solver_timestep(const orc_solver *s);
void orc_solver_destroy(orc_solver *s);

/* Probes for tests: (1) the RHS J*(-div F) of the five equations at the
 * current *1 state (no time advance, k2 registers untouched); (2) the
 * amplification factor R(z) of one time step of scheme rk on y' = lambda y,
 * z = lambda dt, run through the same register-update code (dual_step selects
 * the 5- or 6-stage half of LSLDDRK4). */
int orc_solver_rhs(orc_solver *s, double *const out[5]);

/* Sponge (src/SpongeBC.cpp, src/CurvilinearCSolver_BC.cpp:324-345): strength
 * field sigma (x fastest) for a solver with SPONGE faces; the sponge averages
 * are avg[5] (rho, rhoU, rhoV, rhoW, rhoE), or the current conserved fields
 * when avg is NULL (the reference's start).  Call after create, before steps. */
int orc_solver_set_sponge(orc_solver *s, const double *sigma, const double *const avg[5]);

/* sigma of SpongeBC::initRectSpongeBC / initCylSpongeBC (src/SpongeBC.cpp:60-260)
 * from mesh coordinates: kind 0 rectilinear with perc[6] = RECT{X0,X1,Y0,Y1,Z0,Z1}PERC,
 * kind 1 cylindrical (orientation 2 only) with axis[3] and rmin; faces[6] as
 * in orc_config.  Returns 0, or -1 for an unsupported cylinder orientation. */
int orc_sponge_sigma(int Nx, int Ny, int Nz, const double *x, const double *y, const double *z,
                     const int faces[6], int kind, double strength, const double perc[6],
                     int cyl_orient, const double axis[3], double rmin, double *sigma);
double orc_rk_amplification(int rk, int dual_step, double z);

#ifdef __cplusplus
}
#endif
#endif
