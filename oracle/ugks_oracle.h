/*
 * ugks_oracle.h -- FP64 CPU oracle for the pyUGKS time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (pyugks_amd)
 * includes, links or calls this code; only tests/, the smoke check and the
 * cpu_baseline leg of bench.py load it, as the checker.
 *
 * Every function restates one reference OpenCL kernel (or one host-side
 * sequencing routine) in plain sequential C, in the reference's own data
 * layout (AoS double2 (g,h) per [i][j][v], double4 per face/cell) and in the
 * reference's evaluation order (velocity sums run v = 0..Nv-1 in order, as
 * the reference does with its default CL_local_size = 1).  The reference
 * file:line each function follows is cited beside it in ugks_oracle.c.
 *
 * Pins: tests/test_oracle_pins.py, tests/test_oracle_flux_limits.py,
 * tests/test_oracle_boundaries.py and tests/test_oracle_joins.py check every
 * function against what the mathematics fixes (closed forms, library
 * quadrature, invariants, limits, brute force); the reference ships no
 * golden vectors to check against.  Every function on the path has such a
 * check; none is left parity-unpinned.
 */
#ifndef UGKS_ORACLE_H
#define UGKS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } od2;
typedef struct { double s0, s1, s2, s3; } od4;

/* Boundary-condition codes, geom/geom_bc_defs.py:27-37 */
enum { OR_ADJACENT = 0, OR_EXTRAPOLATE_OUT = 1, OR_CONSTANT = 2, OR_REFLECT = 3,
       OR_DIFFUSE = 4, OR_PERIODIC = 5, OR_INFLOW = 6, OR_OUTFLOW = 7,
       OR_ADSORBING = 8, OR_BOUNCE_BACK = 9 };

/* Per-block constants: what genHeader bakes into #defines (ugks_CL.py:58-395). */
typedef struct or_params {
    int32_t ni, nj, NI, NJ, NV, ghost, flux_method, relax_type;
    double Pr, Kn_eff, gam, B, K, omega, CFL, umax, vmax;
    /* WALL_{N,E,S,W}_{D,U,V,T,P}; T is already 1/T (ugks_CL.py:298) */
    double wall_D[4], wall_U[4], wall_V[4], wall_T[4], wall_P[4];
    /* adsorbing-wall constants (ugks_CL.py:102-156, :160-185) */
    double beta_n[4], beta_t[4], gamma_f[4], alpha_p[4];
    int32_t dynamic_vartheta[4], constant_adsorb[4];
    double adsorb[4];
    /* velocity tables (ugks_CL.py:318-395) */
    const double *quad;      /* [NV][2] */
    const double *weight;    /* [NV] */
    const int64_t *mirror_NS, *mirror_EW, *mirror_D;  /* [NV] */
    /* Cercignani-Lampis accommodation ALPHA_N / ALPHA_T (ugks_CL.py:134-148) */
    double alpha_n[4], alpha_t[4];
    /* tabulated isotherms ISO / TRI / NBR / N_TRI (ugks_CL.py:160-248): [n_iso][3] points
     * (P, T, theta), [n_tri][3] vertices and neighbours; n_tri 0 where a wall has none */
    const double *iso[4];
    const int32_t *iso_tri[4], *iso_nbr[4];
    int32_t n_iso[4], n_tri[4];
} or_params;

/* Device buffers of one block (ugks_block.py:212-255, :338-407). */
typedef struct or_state {
    od2 *f, *flux_f_S, *flux_f_W, *sigma;           /* [NI][NJ][NV]      */
    od4 *macro;                                     /* [ni][nj]          */
    od2 *Q;                                         /* [ni][nj]          */
    od4 *flux_macro_S, *flux_macro_W;               /* [NI][NJ]          */
    od4 *prim, *aL, *aR, *aT;                       /* [NI][NJ]          */
    od2 *faceQ;                                     /* [NI][NJ]          */
    double *Mxi;                                    /* [NI][NJ][3]       */
    od2 *xy;                                        /* [NI+1][NJ+1]      */
    double *area;                                   /* [NI][NJ]          */
    od2 *centre;                                    /* [NI][NJ]          */
    od2 *mid_side, *normal;                         /* [NI][NJ][2]       */
    double *side_length;                            /* [NI][NJ][2]       */
    od4 *wall_prop, *wall_cover;                    /* [wall_len][4]     */
    double *wall_fluxes;                            /* [wall_len][4][8]  */
    od2 *wall_dist;                                 /* [wall_len][NV]    */
    double *time_step;                              /* [ni][nj]          */
    od4 *residual;                                  /* [ni][nj]          */
    double *bc_mdot;                                /* [wall_len][4]     */
    double *flag;                                   /* [2]               */
} or_state;

/* One face's boundary condition (geom_bc_defs.py:82-170). */
typedef struct or_bc {
    int32_t type, other_block, other_face, orientation, flip, reflect_type; /* reflect: 0=S 1=D 2=CL */
} or_bc;

typedef struct or_block {
    or_params p;
    or_state s;
    or_bc bc[4];          /* N, E, S, W */
} or_block;

/* Global run state (the parts of gdata / UGKSim the step uses). */
typedef struct or_sim {
    int32_t nblocks;
    or_block *blocks;
    double dt, time, CFL, delta_dt, suggest_dt, Kn_eff;
    int32_t step, force_time, dt_update_count;
    double *max_dt;       /* [nblocks] */
} or_sim;

/* ---- reference kernels, one function each ------------------------------ */
void or_cellGeom(const or_params *P, const od2 *xy, double *area, od2 *centre,
                 od2 *mid_side, od2 *normal, double *side_length);
void or_paraBC(const or_params *P, od4 *wall_prop);
void or_clFindDT(const or_params *P, const od2 *xy, const double *area,
                 const od4 *macro, double *time_step, double Kn);
void or_initFunctions(const or_params *P, od2 *Fin, const od4 *macro, const od2 *gQ);
void or_calcMacro(const or_params *P, const od2 *Fin, od4 *macro);
void or_calcQ(const or_params *P, const od2 *Fin, const od4 *macro, od2 *gQ);
void or_getInternal(const or_params *P, const od2 *Fin, od4 *Txyz, od4 *Stress);
void or_iFace(const or_params *P, const od2 *Fin, od2 *iface_f, od2 *fsigma,
              const od2 *centre, const od2 *mid_side, const od2 *normal, int face);
void or_getFaceCons(const or_params *P, const od2 *iface_f, const od2 *normal, int face,
                    od4 *primary, int ob, int ot);
void or_getAL(const or_params *P, const od2 *Fin, const od2 *normal, int face, od4 *gaL, int ob, int ot);
void or_getAR(const or_params *P, const od2 *Fin, const od2 *normal, int face, od4 *gaR, int ob, int ot);
void or_getPLR(const or_params *P, const od2 *centre, const od2 *mid_side, int face,
               od4 *primary, od4 *gaL, od4 *gaR, int ob, int ot);
void or_initMacroFlux(const or_params *P, od4 *flux_macro, int face, double dt,
                      const od4 *primary, const od4 *gaL, const od4 *gaR, od4 *gaT,
                      double *gMxi, int ob, int ot);
void or_calcFaceQ(const or_params *P, const od2 *iface_f, const od4 *primary,
                  const od2 *normal, od2 *faceQ, int face, int ob, int ot);
void or_macroFlux(const or_params *P, const od2 *flux_f, const od2 *fsigma, od4 *flux_macro,
                  const od2 *normal, const double *side_length, int face, double dt,
                  const od4 *primary, const od2 *faceQ, int ob, int ot);
void or_distFlux(const or_params *P, od2 *flux_f, const od2 *fsigma, const od2 *normal,
                 const double *side_length, int face, double dt, const od4 *primary,
                 const od4 *gaL, const od4 *gaR, const od4 *gaT, const double *gMxi,
                 const od2 *faceQ, int ob, int ot);
void or_accommodatingWallDist(const or_params *P, const od2 *normal, int face,
                              const od4 *wall_prop, od2 *flux_f, double dt);
void or_wallFlux(const or_params *P, const od2 *normal, const double *side_length, int face,
                 od2 *flux_f, od4 *flux_macro, double dt);
void or_wallMassEnergyFluxes(const or_params *P, const od2 *normal, const double *side_length,
                             const double *area, int face, const od2 *flux_f,
                             double *wall_flux, double dt);
void or_adsorbingWall_P1(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                         od4 *wall_cover, od2 *wall_dist, const od2 *flux_f, const od4 *macro,
                         double dt, double *flag);
void or_adsorbingWallS_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                          const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                          const od4 *macro, double dt);
void or_adsorbingWallCL_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                           const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                           const od4 *macro, double dt);
double or_bessel_i0(double x);
double or_langmuir_table(const or_params *P, int face, double Pr, double T, int *outside);
double or_cl_kernel(od2 uin, od2 uout, double alpha_n, double alpha_t, double Tw);
void or_adsorbingWallD_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                          const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                          const od4 *macro, double dt);
void or_updateMacro(const or_params *P, const od4 *flux_macro_S, const od4 *flux_macro_W,
                    const double *area, od4 *macro, od4 *residual);
void or_UGKS_update(const or_params *P, od2 *Fin, const od2 *flux_f_S, const od2 *flux_f_W,
                    const double *area, const od4 *macro, const od2 *gQ, const od4 *residual,
                    double dt, double *flag);
void or_getResidual(const or_params *P, const od4 *macro, od4 *residual);
void or_edgeMassFlow(const or_params *P, const od4 *macro, const od2 *normal,
                     const double *side_length, double *mdot, int this_face);
void or_edgeExchange(const or_params *PA, od2 *fA, const od2 *xyA, int this_face,
                     const od2 *fB, const od2 *xyB, int NIB, int NJB, int that_face,
                     int ori, int flip);
void or_edgeExtrapolate(const or_params *P, od2 *Fin, int this_face);
void or_edgeConstant(const or_params *P, od2 *Fin, int this_face, const od4 *wall_prop);
void or_edgeInflow(const or_params *P, od2 *Fin, int this_face, const od2 *normal,
                   const od4 *macro, const od2 *gQ, const od4 *wall_prop);
void or_edgeOutflow(const or_params *P, od2 *Fin, int this_face, const od2 *normal,
                    const od4 *macro, const od2 *gQ, const od4 *wall_prop);
void or_edgeMirror(const or_params *P, od2 *Fin, int this_face);
void or_edgeBounceBack(const or_params *P, od2 *Fin, int this_face);
void or_edgeConstGrad(const or_params *P, od2 *Fin, int this_face);

/* ---- host sequencing (ugks_block.py / ugks_sim.py) --------------------- */
void or_block_init(or_block *b);                         /* initFunctions + paraBC      */
int  or_block_flux(or_block *b, double dt);              /* UGKSBlock.UGKS_flux         */
int  or_block_update(or_block *b, double dt);            /* UGKSBlock.UGKS_update       */
void or_block_updateBC(or_sim *sim, int bi);             /* UGKSBlock.updateBC          */
double or_block_getDT(or_sim *sim, int bi);              /* UGKSBlock.getDT             */
void or_sim_updateAllBC(or_sim *sim);
void or_sim_calcTimeStep(or_sim *sim);
void or_sim_updateTimeStep(or_sim *sim, int limit);
int  or_sim_init(or_sim *sim);                           /* UGKSim.__init__ step part   */
int  or_sim_one_step(or_sim *sim, int get_dt);           /* UGKSim.one_step             */
int  or_sim_run(or_sim *sim, int nsteps);                /* the step loop of UGKSim.run */

/* helpers exposed for direct checks against closed forms */
double or_relaxTime(const or_params *P, od4 prim);
od2 or_fM(const or_params *P, od4 prim, od2 uv, int gv);
od2 or_fS(const or_params *P, od4 prim, od2 Q, od2 uv, od2 M);
void or_momentU(const or_params *P, od4 prim, double *Mu, double *Mv, double *Mxi,
                double *Mu_L, double *Mu_R);
void or_timeIntegrals(double tau, double dt, double *Mt);
od2 or_vanLeer(od2 s1, od2 s2);
od2 or_WENO5(const od2 *S, double dx);
od2 or_NND(const od2 *S, double dx);
od4 or_microSlope(const or_params *P, od4 prim, od4 sw);
od4 or_getPrimary(const or_params *P, od4 w);
od4 or_getConserved(const or_params *P, od4 prim);
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
