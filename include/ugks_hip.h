/*
 * ugks_hip.h -- C ABI of libugks_hip.so, the MI355X (gfx950) UGKS time step.
 *
 * Two layers share one block descriptor:
 *
 *  1. Reference-shaped entry points, one per OpenCL kernel of pyUGKS's time
 *     step.  Each keeps the kernel's name (prefixed ugks_) and argument
 *     order; the block handle and a HIP stream replace the pyopencl queue and
 *     NDRange (the launch shape is derived from the block and the face/offset
 *     arguments, as ugks_block.py derives it).  Device buffers are plain
 *     device pointers in the reference's layout: f-like arrays are
 *     [NI][NJ][NV] pairs (g,h) of double, face/cell arrays [NI][NJ] double4,
 *     macro [ni][nj] double4 etc. (ugks_block.py:212-255, :338-407).
 *
 *  2. Fused entry points (ugks_step_*): one block step in a few launches with
 *     the same buffers, plus the mixed FP32/FP64 and batched variants.
 *
 * Every function returns 0 on success, a negative value on failure
 * (-1: bad argument, -2: unsupported configuration, otherwise -(hipError_t));
 * ugks_last_error() describes the last failure of the calling thread.
 * Errors detected by kernels travel through the reference's flag buffer
 * (flag[0]: 1 = NaN in f, 4 = adsorption isotherm clamp; clError.c:4-13).
 * Calls are stream-ordered; one block must not be used from two streams at
 * once.
 */
#ifndef UGKS_HIP_H
#define UGKS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-block constants: the #define set genHeader bakes per block
 * (ugks_CL.py:58-395), held at run time instead of compiled in. */
typedef struct ugks_block_desc {
    int32_t ni, nj, NI, NJ, NV, ghost;
    int32_t flux_method;            /* 0 vanLeer, 1 WENO5, 3 minmod (ugks_CL.py:33-44) */
    int32_t relax_type;             /* clFunctions.c:8-32 */
    double Pr, Kn_eff, gam, B, K, omega, CFL, umax, vmax;
    double wall_D[4], wall_U[4], wall_V[4], wall_T[4], wall_P[4];  /* N,E,S,W; T is 1/T */
    double beta_n[4], beta_t[4], gamma_f[4], alpha_p[4];           /* ugks_CL.py:102-132 */
    int32_t dynamic_vartheta[4], constant_adsorb[4];
    double adsorb[4];                                              /* ugks_CL.py:163-167 */
    const double *quad;             /* host [NV][2]  (QUAD,   ugks_CL.py:318-333) */
    const double *weight;           /* host [NV]     (WEIGHT, ugks_CL.py:335-346) */
    const int64_t *mirror_NS, *mirror_EW, *mirror_D;  /* host [NV] (ugks_CL.py:348-395) */
    double alpha_n[4], alpha_t[4];  /* Cercignani-Lampis accommodation, -1 off adsorbing walls (ugks_CL.py:134-148) */
    /* Tabulated Langmuir isotherms (clAdsorb.c:171-247, ugks_CL.py:160-248), for adsorbing walls
     * with constant_adsorb 0: n_iso points (P, T, theta) as host [n_iso][3] doubles, and the n_tri
     * triangles of the Delaunay triangulation of (P, T) as host [n_tri][3] vertex indices and
     * [n_tri][3] neighbours (neighbour k opposite vertex k, -1 on the hull).  n_tri 0: no table. */
    const double *iso[4];
    const int32_t *iso_tri[4], *iso_nbr[4];
    int32_t n_iso[4], n_tri[4];
    /* Wall values given as expressions (UDF_D/U/V/T, geom_bc_defs.py:115-146, which
     * ugks_CL.py:278-298 pastes into the kernel source): stack programs in s, x, y, t
     * compiled by the host (opcodes in pyugks_amd/udf.py).  Slot 4*w + q (wall w = N,E,S,W;
     * q = 0 D, 1 U, 2 V, 3 T, the temperature itself) runs udf_len[slot] opcodes from
     * udf_code + udf_start[slot], with constants from udf_const; udf_len 0 keeps the
     * constant wall value.  ugks_block_create checks every program before it is copied. */
    const int32_t *udf_code;
    const double *udf_const;
    int32_t udf_ncode, udf_nconst;
    int32_t udf_start[16], udf_len[16];
} ugks_block_desc;

typedef struct ugks_block ugks_block;

/* ---- runtime ----------------------------------------------------------- */
/* Replaces UGKSBlock.block_CL/update_CL (ugks_block.py:151-189): copies the
 * velocity tables to the device once; no per-block compilation. */
int ugks_block_create(const ugks_block_desc *desc, ugks_block **out);
int ugks_block_destroy(ugks_block *blk);
const char *ugks_last_error(void);
int ugks_get_device_count(int *count);
int ugks_set_device(int device);
int ugks_malloc(void **ptr, size_t bytes);
int ugks_free(void *ptr);
int ugks_memset(void *ptr, int value, size_t bytes, void *stream);
int ugks_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int ugks_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int ugks_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
int ugks_stream_create(void **stream);
int ugks_stream_destroy(void *stream);
int ugks_stream_synchronize(void *stream);
int ugks_event_create(void **event);
int ugks_event_destroy(void *event);
int ugks_event_record(void *event, void *stream);
int ugks_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- reference-shaped kernels (layer 1) -------------------------------- */
/* clKernel.c:31-37, launched (Ni,Nj) at ugks_block.py:748-751 */
int ugks_cellGeom(ugks_block *blk, void *stream, const void *xy, void *area, void *centre,
                  void *mid_side, void *normal, void *side_length);
/* clKernel.c:141-144 (constant WALL_* values), ugks_block.py:603-619 */
int ugks_paraBC(ugks_block *blk, void *stream, const void *para_def, void *wall_prop, double t);
/* clKernel.c:224-227, ugks_block.py:1420-1421 */
int ugks_clFindDT(ugks_block *blk, void *stream, const void *xy, const void *area,
                  const void *macro, void *time_step, double Kn);
/* device max of time_step into out[0] (replaces cl_array.max, ugks_block.py:1426) */
int ugks_time_step_max(ugks_block *blk, void *stream, const void *time_step, void *out);
/* clKernel.c:382-384, ugks_block.py:413-415 */
int ugks_initFunctions(ugks_block *blk, void *stream, void *Fin, const void *macro, const void *gQ);
/* clKernel.c:414-416, ugks_block.py:419-420 */
int ugks_calcMacro(ugks_block *blk, void *stream, const void *Fin, void *macro);
/* clKernel.c:480-483, ugks_block.py:1271-1272 */
int ugks_calcQ(ugks_block *blk, void *stream, const void *Fin, const void *macro, void *gQ);
/* getInternal (clKernel.c:315-372): Txyz [ni][nj] double4 (Tx, Ty, Tz, T) and
 * Stress [ni][nj] double4 (tau_xx, tau_xy, tau_yy, 0) from f, the diagnostics
 * of save_options.internal_data (ugks_block.py:1335-1348) */
int ugks_getInternal(ugks_block *blk, void *stream, const void *Fin, void *Txyz, void *Stress);
/* clUGKS.c:110-117, ugks_block.py:1066-1068 */
int ugks_iFace(ugks_block *blk, void *stream, const void *Fin, void *iface_f, void *fsigma,
               const void *centre, const void *mid_side, const void *normal, int32_t face);
/* clUGKS.c:209-214, ugks_block.py:1107-1109 */
int ugks_getFaceCons(ugks_block *blk, void *stream, const void *iface_f, const void *normal,
                     int32_t face, void *primary, int32_t offset_bottom, int32_t offset_top);
/* clUGKS.c:273-278, ugks_block.py:1111-1113 */
int ugks_getAL(ugks_block *blk, void *stream, const void *Fin, const void *normal, int32_t face,
               void *gaL, int32_t offset_bottom, int32_t offset_top);
/* clUGKS.c:337-342, ugks_block.py:1115-1117 */
int ugks_getAR(ugks_block *blk, void *stream, const void *Fin, const void *normal, int32_t face,
               void *gaR, int32_t offset_bottom, int32_t offset_top);
/* clUGKS.c:401-408, ugks_block.py:1123-1125 */
int ugks_getPLR(ugks_block *blk, void *stream, const void *centre, const void *mid_side,
                int32_t face, void *primary, void *gaL, void *gaR, int32_t offset_bottom,
                int32_t offset_top);
/* clUGKS.c:446-454, ugks_block.py:1129-1132 */
int ugks_initMacroFlux(ugks_block *blk, void *stream, void *flux_macro, int32_t face, double dt,
                       const void *primary, const void *gaL, const void *gaR, void *gaT,
                       void *gMxi, int32_t offset_bottom, int32_t offset_top);
/* clUGKS.c:522-527, ugks_block.py:1138-1140 */
int ugks_calcFaceQ(ugks_block *blk, void *stream, const void *iface_f, const void *primary,
                   const void *normal, void *faceQ, int32_t face, int32_t offset_bottom,
                   int32_t offset_top);
/* clUGKS.c:587-596, ugks_block.py:1145-1148 */
int ugks_macroFlux(ugks_block *blk, void *stream, const void *flux_f, const void *fsigma,
                   void *flux_macro, const void *normal, const void *side_length, int32_t face,
                   double dt, const void *primary, const void *faceQ, int32_t offset_bottom,
                   int32_t offset_top);
/* clUGKS.c:683-695, ugks_block.py:1150-1154 */
int ugks_distFlux(ugks_block *blk, void *stream, void *flux_f, const void *fsigma,
                  const void *normal, const void *side_length, int32_t face, double dt,
                  const void *primary, const void *gaL, const void *gaR, const void *gaT,
                  const void *gMxi, const void *faceQ, int32_t offset_bottom, int32_t offset_top);
/* clUGKS.c:776-781, ugks_block.py:936-938 */
int ugks_accommodatingWallDist(ugks_block *blk, void *stream, const void *normal, int32_t face,
                               const void *wall_prop, void *flux_f, double dt);
/* clUGKS.c:911-915, ugks_block.py:973-975 */
int ugks_wallFlux(ugks_block *blk, void *stream, const void *normal, const void *side_length,
                  int32_t face, void *flux_f, void *flux_macro, double dt);
/* clUGKS.c:1019-1026, ugks_block.py:979-981 */
int ugks_wallMassEnergyFluxes(ugks_block *blk, void *stream, const void *normal,
                              const void *side_length, const void *area, int32_t face,
                              const void *flux_f, void *wall_flux, double dt);
/* clAdsorb.c:253-262, ugks_block.py:942-945 */
int ugks_adsorbingWall_P1(ugks_block *blk, void *stream, const void *normal, int32_t face,
                          const void *wall_prop, void *wall_cover, void *wall_dist,
                          const void *flux_f, const void *macro, double dt, void *flag);
/* clAdsorb.c:656-664, ugks_block.py:966-969 */
int ugks_adsorbingWallS_P2(ugks_block *blk, void *stream, const void *normal, int32_t face,
                           const void *wall_prop, const void *wall_cover, const void *wall_dist,
                           void *flux_f, const void *macro, double dt);
/* clAdsorb.c:821-829, ugks_block.py:960-963 */
int ugks_adsorbingWallD_P2(ugks_block *blk, void *stream, const void *normal, int32_t face,
                           const void *wall_prop, const void *wall_cover, const void *wall_dist,
                           void *flux_f, const void *macro, double dt);
/* clAdsorb.c:488-495, ugks_block.py:954-957: Cercignani-Lampis re-emission,
 * accommodation alpha_n / alpha_t from ugks_block_desc */
int ugks_adsorbingWallCL_P2(ugks_block *blk, void *stream, const void *normal, int32_t face,
                            const void *wall_prop, const void *wall_cover, const void *wall_dist,
                            void *flux_f, const void *macro, double dt);
/* clUGKS.c:1158-1163, ugks_block.py:1279-1281 */
int ugks_updateMacro(ugks_block *blk, void *stream, const void *flux_macro_S,
                     const void *flux_macro_W, const void *area, void *macro, void *residual);
/* clUGKS.c:1189-1197, ugks_block.py:1292-1295 */
int ugks_UGKS_update(ugks_block *blk, void *stream, void *Fin, const void *flux_f_S,
                     const void *flux_f_W, const void *area, const void *macro, const void *gQ,
                     const void *residual, double dt, void *flag);
/* clUGKS.c:1259-1261, ugks_block.py:1311-1312 */
int ugks_getResidual(ugks_block *blk, void *stream, const void *macro, void *residual);
/* clUGKS.c:1282-1287, ugks_block.py:910-911 */
int ugks_edgeMassFlow(ugks_block *blk, void *stream, const void *macro, const void *normal,
                      const void *side_length, void *mdot, int32_t this_face);
/* clGhostRoutines.c:407-414, ugks_block.py:468-470 (blkB supplies NV and the mirror tables
 * are taken from blk; NIB/NJB as in the reference) */
int ugks_edgeExchange(ugks_block *blk, void *stream, void *fA, const void *xyA, int32_t this_face,
                      const void *fB, const void *xyB, int32_t NIB, int32_t NJB,
                      int32_t that_face, int32_t ori, int32_t flip);
/* clGhostRoutines.c:544-546, ugks_block.py:491-492 */
int ugks_edgeExtrapolate(ugks_block *blk, void *stream, void *Fin, int32_t this_face);
/* clGhostRoutines.c:603-606, ugks_block.py:560-561 */
int ugks_edgeConstant(ugks_block *blk, void *stream, void *Fin, int32_t this_face,
                      const void *wall_prop);
/* clGhostRoutines.c:664-670, ugks_block.py:512-514 */
int ugks_edgeInflow(ugks_block *blk, void *stream, void *Fin, int32_t this_face,
                    const void *normal, const void *macro, const void *gQ, const void *wall_prop);
/* clGhostRoutines.c:755-761, ugks_block.py:537-539 */
int ugks_edgeOutflow(ugks_block *blk, void *stream, void *Fin, int32_t this_face,
                     const void *normal, const void *macro, const void *gQ, const void *wall_prop);
/* clGhostRoutines.c:845-847, ugks_block.py:582 */
int ugks_edgeMirror(ugks_block *blk, void *stream, void *Fin, int32_t this_face);
/* clGhostRoutines.c:911-913, ugks_block.py:580 */
int ugks_edgeBounceBack(ugks_block *blk, void *stream, void *Fin, int32_t this_face);
/* clGhostRoutines.c:967-968, ugks_block.py:600-601 */
int ugks_edgeConstGrad(ugks_block *blk, void *stream, void *Fin, int32_t this_face);

/* ---- fused, batched step (layer 2) -------------------------------------- */
/* One plan steps any number of blocks (of one or many independent cases) in
 * eight launches per step with the time step kept on the device.  It replaces
 * the sequence UGKSim.one_step drives per block (ugks_sim.py:407-501:
 * UGKS_flux, UGKS_update, calcTimeStep/updateTimeStep, updateAllBC) and
 * produces the same f, macro, Q, flux_macro, wall_cover and wall_fluxes.
 * Blocks given to one plan share one time step (ugks_sim.py:296-308);
 * independent cases that need their own time steps get their own plans. */
enum { UGKS_FP64 = 0, UGKS_MIXED = 1 };
/* launches per step, in order: face moments, face state, face flux, wall
 * faces, cell update, cell time step, time-step finalise, ghost cells */
#define UGKS_STEP_NKERNELS 8

/* Device buffers of one block, in the layouts of layer 1.  f, flux_f_S,
 * flux_f_W and wall_dist hold (g,h) pairs of double for UGKS_FP64 and of
 * float for UGKS_MIXED; everything else is double.  macro_new is a second
 * [ni][nj] double4 buffer (the step writes the new state there and copies it
 * back to macro). */
typedef struct ugks_block_buffers {
    void *f, *flux_f_S, *flux_f_W, *flux_macro_S, *flux_macro_W;
    void *macro, *macro_new, *residual, *Q;
    const void *xy, *centre, *mid_side, *normal, *area, *side_length;
    void *wall_prop, *wall_cover, *wall_fluxes, *wall_dist, *time_step, *flag;
    const void *para;   /* [wall_len][4] double4 (s, x, y, 0) of each wall cell (ugks_block.py:776-836);
                           read when the block's wall values depend on time */
} ugks_block_buffers;

/* One face's BC (geom/geom_bc_defs.py:27-51): type code, partner block and
 * face for ADJACENT, orientation flag, distribution flip, reflect type of an
 * adsorbing wall (0 specular, 1 diffuse). */
typedef struct ugks_bc_desc {
    int32_t type, other_block, other_face, orientation, flip, reflect_type;
} ugks_bc_desc;

/* Time-step state (the gdata fields of ugks_sim.py:285-311, :491); overshoot
 * reports a first step whose re-evaluated dt differed (ugks_sim.py:443-445);
 * force_time holds dt fixed (UGKSim.force_time, ugks_sim.py:292, :673-675).
 * max_dt is the smallest block maximum time step of the case's last
 * calcTimeStep (UGKSBlock.max_dt, ugks_block.py:1415-1433), against which the
 * first step re-evaluates dt as UGKSim.one_step does (ugks_sim.py:430-446);
 * 0 takes dt as it is.  A plan reports it back as the smallest of its blocks. */
typedef struct ugks_dt_desc {
    double dt, time, CFL, delta_dt, suggest_dt, Kn_eff;
    int32_t step, overshoot, force_time, pad;
    double max_dt;
} ugks_dt_desc;

typedef struct ugks_step_plan ugks_step_plan;

int ugks_step_plan_create(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                          int32_t nblocks, int32_t precision, const ugks_dt_desc *dt, ugks_step_plan **out);
/* Many independent cases in one plan: block b belongs to case case_of_block[b]
 * (blocks of a case contiguous, cases numbered 0..ncases-1 in block order);
 * dt[ncases] holds each case's time-step state, which evolves separately as in
 * one UGKSim per case (ugks_sim.py:285-311).  bcs[].other_block indexes the
 * plan's block list and must stay within the block's case. */
int ugks_step_plan_create_batch(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                                int32_t nblocks, const int32_t *case_of_block, int32_t ncases, int32_t precision,
                                const ugks_dt_desc *dt, ugks_step_plan **out);
int ugks_step_plan_destroy(ugks_step_plan *plan);
/* nsteps full steps: flux, update, time step, BC refresh (ugks_sim.py:407-501).
 * Each call first rederives the per-cell moments the plan keeps from f as it
 * stands, so f may be changed by other calls between runs. */
int ugks_step_run(ugks_step_plan *plan, void *stream, int32_t nsteps);
/* the same steps as replays of a HIP graph of one step, captured on first use
 * (the step has no host round trip, so the graph is valid for the plan's life) */
int ugks_step_run_graph(ugks_step_plan *plan, void *stream, int32_t nsteps);
/* the same steps with HIP events around each launch; kernel_ms[UGKS_STEP_NKERNELS]
 * receives the summed time of each launch of the step, in the order above */
int ugks_step_run_timed(ugks_step_plan *plan, void *stream, int32_t nsteps, float *kernel_ms);
/* Time the face passes with each workgroup size (1, 2 or 4 faces per
 * workgroup, reps launches each) and keep the fastest per pass; choice[2]
 * (may be NULL) receives the picks.  It leaves the simulation state as it was:
 * the face passes only rewrite per-step scratch.  No reference counterpart:
 * the reference takes CL_local_size from its own timing loop
 * (ugks_sim.py:352-405). */
int ugks_step_autotune(ugks_step_plan *plan, void *stream, int32_t reps, int32_t *choice);
/* read back dt, time, step of case 0 (and the per-block maximum time steps if max_dt != NULL) */
int ugks_step_state(ugks_step_plan *plan, void *stream, ugks_dt_desc *out, double *max_dt);
/* the same for every case: out[ncases], max_dt[nblocks] (may be NULL) */
int ugks_step_states(ugks_step_plan *plan, void *stream, ugks_dt_desc *out, double *max_dt);
int ugks_step_set_dt(ugks_step_plan *plan, void *stream, double dt);
int ugks_step_set_dt_case(ugks_step_plan *plan, void *stream, int32_t case_index, double dt);
/* set dt and hold it for the remaining steps (the last step that lands on max_time) */
int ugks_step_force_dt_case(ugks_step_plan *plan, void *stream, int32_t case_index, double dt);
/* 0 if a block of NJ cells per column (ghosts included) and NV velocities can
 * join a plan of this precision, -2 otherwise: the face passes address the
 * stencil of a face with 32-bit byte offsets, so (2R-1)*NJ*NV + NV pairs
 * (R = 3 for WENO5, 2 otherwise) must stay below 2^32 bytes.  Plan creation
 * applies the same check.  Host only, no device call.  No reference
 * counterpart (the reference's kernels index f with size_t). */
int ugks_step_block_fits(int32_t NJ, int32_t NV, int32_t flux_method, int32_t precision);
/* plain device-to-device copy kernel of `bytes` (a multiple of 16): the
 * reference stream for bandwidth and byte-counter calibration */
int ugks_copy_kernel(void *stream, const void *src, void *dst, int64_t bytes);
/* element-wise precision conversion of n values (f of the mixed-precision step) */
int ugks_f64_to_f32(void *stream, const void *src, void *dst, int64_t n);
int ugks_f32_to_f64(void *stream, const void *src, void *dst, int64_t n);

/* ---- allocation-free step ------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nsteps fused steps of ONE case of up to 16 blocks, the
 * same launches as ugks_step_plan_create + ugks_step_run, with every piece of
 * working memory (face sums, face states and records, cell moments, the block
 * records and work tables, the time-step state) carved from a library-owned
 * static device array of ugks_step_noalloc_capacity() = 64 MiB per GPU, which
 * each GPU holds until the process ends.  dt is a ugks_dt_desc in DEVICE
 * memory, read at the start and written back at the end (no host transfer or
 * synchronisation).  Blocks and buffers are the caller's, as for a plan.
 * Returns -2, launching nothing, when the case needs more working memory
 * (about 452 bytes per interior face plus 32 per cell, ghosts included).
 * Calls on one GPU share its array and keep the last call's contents; calls
 * that overlap on one GPU race.  The per-call set-up re-derives everything
 * it reads, so consecutive calls may advance different cases. */
size_t ugks_step_noalloc_capacity(void);
int ugks_step_run_noalloc(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                          int32_t nblocks, int32_t precision, ugks_dt_desc *dt, int32_t nsteps, void *stream);

#ifdef __cplusplus
}
#endif
#endif
