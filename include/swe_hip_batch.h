/*
 * swe_hip_batch.h -- C ABI of the batched engine of libswe_hip.so: many
 * flavour-U blocks of uneven size advanced together on one GPU.
 *
 * Replaces the per-block calls of the reference's in-rank drivers:
 *   - the overdecomposed driver loop (swe_mpi_overdecomp.cpp:263-343):
 *     setGhostLayer / receiveGhostLayer / computeNumericalFluxes / global
 *     dt minimum or getRoundTimestep / updateUnknowns / hasMaxLocalTimestep /
 *     resetStepSizeCounter over every block of the rank;
 *   - the in-rank ghost copy of CONNECT_WITHIN_RANK sides
 *     (SWE_DimensionalSplittingMPIOverdecomp.cpp:273-304) and the bathymetry
 *     copy (:208-235), done on the device;
 *   - the LTS state machine of every block (SWE_Block.hh:311-562).
 * and adds an ensemble mode (independent blocks, each with its own dt and
 * time: as many swe_simple runs, swe_simple.cpp:215-252, side by side).
 *
 * The host arrays of all blocks are packed one after another in the
 * reference's Float2D layout: block k occupies (nx[k]+2)*(ny[k]+2) floats,
 * a[x][y] = raw[off_k + x*(ny[k]+2) + y].
 *
 * Conventions (return codes, swe_hip_last_error, one host thread per handle)
 * are those of swe_hip.h.
 */
#ifndef SWE_HIP_BATCH_H
#define SWE_HIP_BATCH_H

#include <stdint.h>

#include "swe_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* modes */
#define SWE_BATCH_ENSEMBLE 0 /* independent blocks, per-block CFL dt; a unit = one step of every block */
#define SWE_BATCH_GLOBAL 1   /* connected blocks, dt = minimum over blocks (swe_mpi_overdecomp.cpp:297-307);
                                a unit = one step */
#define SWE_BATCH_LTS 2      /* connected blocks, local time stepping with macro step max_timestep_local
                                (SWE_Block.hh:311-562); a unit = one macro step */

typedef struct swe_hip_batch swe_hip_batch;

/* Create a batch of nblocks blocks.  Per block k: nx[k] x ny[k] cells, mesh
 * sizes dx[k], dy[k], boundary types bnd[4k..4k+3] and neighbour indices
 * nbr[4k..4k+3] (order LEFT, RIGHT, BOTTOM, TOP; -1 = none).  A
 * CONNECT_WITHIN_RANK side needs a neighbour whose opposite side is
 * CONNECT_WITHIN_RANK back to k with the same edge length.  solver: AugRie,
 * f-wave, HLLE or Hybrid (flavour U).  max_timestep_local and duration are
 * the LTS macro step and simulation end (the reference's maxTimestepLocal and
 * simulation duration; ignored by the other modes).  At most 2047 blocks. */
int swe_hip_batch_create(int nblocks, const int* nx, const int* ny, const float* dx, const float* dy, const int* bnd,
                         const int* nbr, int solver, int mode, float max_timestep_local, float duration,
                         swe_hip_batch** out);
int swe_hip_batch_destroy(swe_hip_batch* bt);

/* Upload the packed state of every block (ghost cells included) and start
 * from time 0: LTS counters reset, ghost layers of OUTFLOW / WALL sides and
 * corners set as applyBoundaryConditions does (SWE_Block.hh:757-911), and
 * the bathymetry ghost lines of CONNECT_WITHIN_RANK sides copied from the
 * neighbours (SWE_DimensionalSplittingMPIOverdecomp.cpp:208-235).  The
 * bathymetry ghosts of the other sides are taken as uploaded. */
int swe_hip_batch_upload(swe_hip_batch* bt, const float* h, const float* hu, const float* hv, const float* b);
int swe_hip_batch_download(swe_hip_batch* bt, float* h, float* hu, float* hv, float* b);

/* Advance by `units` units (see the modes).  dt_log (may be NULL) receives
 * one float per unit: the global dt of each step (GLOBAL), 0 otherwise.
 * iter_log (may be NULL) receives one int per unit: the driver iterations of
 * each macro step (LTS), 1 otherwise.  elapsed_ms (may be NULL) receives the
 * device time of the whole advance (HIP events on the batch's stream).  An
 * LTS macro step that has not completed after 4 * 1025 iterations ends the
 * call with SWE_HIP_ETIMEOUT. */
int swe_hip_batch_run(swe_hip_batch* bt, int units, float* dt_log, int* iter_log, float* elapsed_ms);

/* Per-block counters: updates performed and (ENSEMBLE) simulated time, both
 * since the last upload; either pointer may be NULL. */
int swe_hip_batch_stats(swe_hip_batch* bt, int64_t* updates, float* t);

/* Tuning: xs = strip columns per work item (4, 8 or 16), groups = workgroups
 * of the persistent sweep / apply grids (1..65535), or 0 for the direct form
 * of the fused sweep (one wave per possible work item, no queue loop; the
 * default; the store strategy's grids then take their default size), chunk = driver
 * iterations per captured hipGraph (1..256), graph = 1 replays captured
 * graphs, 0 launches the kernels one by one, fused = 1 applies each block's
 * update on its next load (one sweep per iteration, the default), 0 runs a
 * separate apply pass; -1 keeps a value.  Results do not depend on the
 * tuning. */
int swe_hip_batch_set_tuning(swe_hip_batch* bt, int xs, int groups, int chunk, int graph, int fused);
int swe_hip_batch_get_tuning(swe_hip_batch* bt, int* xs, int* groups, int* chunk, int* graph, int* fused);

/* Autotuner: times every (xs, groups) candidate on `iterations` driver
 * iterations of the current state, restores the state afterwards and keeps
 * the fastest.  cand_ms (may be NULL, else room for 16 floats) receives the
 * time per iteration of each candidate in the order xs in {4, 8, 16} x
 * groups in {2 x CUs, 4 x CUs, 8 x CUs, 0 (direct)}; *ncand their number. */
int swe_hip_batch_autotune(swe_hip_batch* bt, int iterations, float* cand_ms, int* ncand);

#ifdef __cplusplus
}
#endif

#endif /* SWE_HIP_BATCH_H */
