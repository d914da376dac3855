/*
 * swe_hip.h -- C ABI of the MI355X SWE block library (libswe_hip.so).
 *
 * This is the drop-in boundary for SWE-PPM's dimensional-splitting step.  The
 * reference has no FFI: its blocks are C++ subclasses of SWE_Block
 * (src/blocks/SWE_Block.hh:110-251) whose step is driven through
 *     setGhostLayer()            e.g. SWE_DimensionalSplittingMpi.cpp:188-376
 *     computeNumericalFluxes()   SWE_Block.hh:195 (pure virtual)
 *     getMaxTimestep()           SWE_Block.hh:594-597
 *     updateUnknowns(float dt)   SWE_Block.hh:197 (pure virtual)
 * The C++ subclass SWE_DimensionalSplittingHip (include/swe/) implements
 * those methods on top of the entry points below; anything else (Python
 * ctypes, other languages) binds these symbols directly.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a negative SWE_HIP_E* code;
 *    swe_hip_last_error() returns a description of the last failure on the
 *    calling thread.  The reference methods return void and assert
 *    (SWE_Block.hh:785), so the C++ subclass turns codes into exceptions.
 *  - Host arrays use the reference's Float2D layout: (nx+2) x (ny+2) floats,
 *    a[x][y] = raw[x*(ny+2) + y], ghost cells at 0 and n+1
 *    (src/tools/Float2D.hh:74-80).
 *  - Calls on one handle are ordered on the handle's HIP stream and
 *    serialised by the library (a layer copy or transfer takes the lock of
 *    every block it touches, since reading a block's layer applies its
 *    pending update).  Different handles may be driven from different
 *    threads (the overdecomposed drivers do this,
 *    swe_mpi_overdecomp.cpp:269-323), neighbour exchanges included.
 *  - Enumerations match the reference: BoundaryType and Boundary
 *    (src/types/Boundary.hh:5-28), solver ids of WAVE_PROPAGATION_SOLVER
 *    (SWE_DimensionalSplittingMpi.hh:42-50) plus 3 = Hybrid.
 */
#ifndef SWE_HIP_H
#define SWE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_HIP_OK 0
#define SWE_HIP_EINVAL (-1)  /* bad argument (size, enum, null pointer)        */
#define SWE_HIP_EDEVICE (-2) /* HIP runtime error (allocation, launch, copy)   */
#define SWE_HIP_ESTATE (-3)  /* call not valid in the handle's current state   */
#define SWE_HIP_ETIMEOUT (-4) /* a bounded device-side wait gave up             */

/* solvers */
#define SWE_SOLVER_HLLE 0
#define SWE_SOLVER_FWAVE 1
#define SWE_SOLVER_AUGRIE 2
#define SWE_SOLVER_HYBRID 3

/* step flavours: U = the step of every parallel block of the reference
 * (SWE_DimensionalSplittingMpi.cpp:383-518), S = swe_simple's split sweep
 * (SWE_DimensionalSplitting.cpp:105-210) */
#define SWE_FLAVOUR_U 0
#define SWE_FLAVOUR_S 1

typedef struct swe_hip_block swe_hip_block;

/* Thread-local description of the last error ("" if none). */
const char* swe_hip_last_error(void);

/* Library version string. */
const char* swe_hip_version(void);

/* Create a block of nx x ny interior cells with mesh sizes dx, dy and origin
 * (ox, oy).  Replaces the constructors
 *   SWE_DimensionalSplitting(nx, ny, dx, dy, ox, oy)        SWE_DimensionalSplitting.hh:44-45
 *   SWE_DimensionalSplittingMpi(..., bool localTimestepping) SWE_DimensionalSplittingMpi.hh:55-56
 * Device memory for the state (h, hu, hv, b), the net-update sums and the
 * control block is allocated here.  Boundaries start PASSIVE
 * (SWE_Block.hh:297-300). */
int swe_hip_create(int nx, int ny, float dx, float dy, float ox, float oy, int solver, int flavour,
                   swe_hip_block** out);
int swe_hip_destroy(swe_hip_block* blk);

/* Copy the state from / to host Float2D arrays (b may be NULL on download).
 * Replaces the host-memory accessors getWaterHeight()/getMomentum*()/
 * getBathymetry() (SWE_Block.hh:600-637): the device owns the state, host
 * mirrors are refreshed on demand (the lazy synchronisation SWE_Block.hh:57-82
 * describes). */
int swe_hip_upload(swe_hip_block* blk, const float* h, const float* hu, const float* hv, const float* b);
int swe_hip_download(swe_hip_block* blk, float* h, float* hu, float* hv, float* b);

/* setBoundaryType (SWE_Block.hh:640-642). */
int swe_hip_set_boundary(swe_hip_block* blk, int side, int type);

/* applyBoundaryConditions (SWE_Block.hh:757-911): the body of setGhostLayer
 * for OUTFLOW / WALL sides and the corners. */
int swe_hip_apply_bc(swe_hip_block* blk);

/* applyBoundaryBathymetry (SWE_Block.hh:724-749). */
int swe_hip_apply_bc_bathymetry(swe_hip_block* blk);

/* A gridded field with cell-centred values: value (i, j) sits at
 * (x0 + i dx, y0 + j dy) and covers the cell [x0 + (i - 1/2) dx,
 * x0 + (i + 1/2) dx) x [...]; `values` holds ny rows of nx floats (x fastest,
 * the (y, x) order of a netCDF variable z(y, x)), in host memory. */
typedef struct swe_hip_grid {
    int nx, ny;
    double x0, y0;
    double dx, dy;
    const float* values;
} swe_hip_grid;

/* initScenario (SWE_Block.hh:650-682) of the file-driven scenario
 * (SWE_AsagiScenario.hh:42-151), sampled on the device: every cell, ghosts
 * included, at x = (float)(ox + (i - 1/2) dx), y = (float)(oy + (j - 1/2) dy)
 * in double, gets
 *   bathy = the bathymetry grid's value of the cell containing (x, y)
 *           (indices from floor((x - (x0 - dx/2)) / dx) in double, clamped to
 *           the grid),
 *   b     = bathy + (displacement value at (x, y) if (x, y) lies strictly
 *           inside the displacement grid's float extent, else 0),
 *   h     = bathy > 0 ? 0 : -bathy,  hu = 0 * h,  hv = 0 * h;
 * then applyBoundaryConditions and applyBoundaryBathymetry with the block's
 * boundary types.  displacement may be NULL.  No counterpart entry in the
 * reference: it replaces the host loop of initScenario for this scenario. */
int swe_hip_init_grid_scenario(swe_hip_block* blk, const swe_hip_grid* bathymetry, const swe_hip_grid* displacement);

/* The analytic scenarios of include/swe/scenarios/SWE_simple_scenarios.hh
 * (the reference's SWE_simple_scenarios.hh:37-285 and the centred flat-bed
 * dam break), for swe_hip_init_scenario. */
#define SWE_SCENARIO_HALF_DOMAIN_DRY 0
#define SWE_SCENARIO_RADIAL_DAM_BREAK 1
#define SWE_SCENARIO_RADIAL_BATHYMETRY_DAM_BREAK 2
#define SWE_SCENARIO_BATHYMETRY_DAM_BREAK 3
#define SWE_SCENARIO_SEA_AT_REST 4
#define SWE_SCENARIO_SPLASHING_POOL 5
#define SWE_SCENARIO_SPLASHING_CONE 6
#define SWE_SCENARIO_CENTRED_DAM_BREAK 7

/* initScenario (SWE_Block.hh:650-682) of one of the analytic scenarios above,
 * sampled on the device: every cell, ghosts included, at the same sample
 * point as swe_hip_init_grid_scenario, gets the scenario class's
 * getBathymetry / getWaterHeight evaluated in single precision with the same
 * operations (hu = hv = 0 * h), then applyBoundaryConditions and
 * applyBoundaryBathymetry with the block's boundary types (set them first).
 * The result equals the host loop over the scenario object bit for bit.  No
 * counterpart entry in the reference: it replaces the host loop of
 * initScenario for these scenarios. */
int swe_hip_init_scenario(swe_hip_block* blk, int scenario);

/* The seeded benchmark workloads (swe_ppm_amd/scenarios.py) as parameter
 * sets, for swe_hip_init_seeded.  With (x, y) a cell centre as above and
 * every operation in single precision, each exponential evaluated in double
 * and rounded once to single:
 *   SWE_SEEDED_DAM:    b = sum over k < nbumps of
 *                        bump_a[k] * exp(-((x - bump_x[k])^2 + (y - bump_y[k])^2) / bump_s2[k])
 *                      (summed in k order from 0), eta = eta_in where the
 *                      distance to (dam_x, dam_y) is below dam_radius, eta_out
 *                      elsewhere, h = max(0, eta - b);
 *   SWE_SEEDED_TRENCH: depth = 10 + 3990 * exp(-(x - trench_x)^2 / trench_w2),
 *                      b = -depth, h = depth + 2 * exp(-((x - hump_x)^2 +
 *                      (y - hump_y)^2) / hump_r2);
 * hu = hv = 0. */
#define SWE_SEEDED_DAM 0
#define SWE_SEEDED_TRENCH 1
#define SWE_SEEDED_MAX_BUMPS 32
typedef struct swe_hip_seeded {
    int kind;
    int nbumps;
    float dam_x, dam_y, dam_radius, eta_in, eta_out;
    float trench_x, trench_w2, hump_x, hump_y, hump_r2;
    float bump_x[SWE_SEEDED_MAX_BUMPS];
    float bump_y[SWE_SEEDED_MAX_BUMPS];
    float bump_a[SWE_SEEDED_MAX_BUMPS];
    float bump_s2[SWE_SEEDED_MAX_BUMPS];
} swe_hip_seeded;

/* initScenario (SWE_Block.hh:650-682) of a seeded workload, sampled on the
 * device at every cell, ghosts included, then applyBoundaryConditions and
 * applyBoundaryBathymetry with the block's boundary types (set them first).
 * The result equals swe_ppm_amd.scenarios' host sampling of the same
 * parameters bit for bit.  SWE_HIP_EINVAL for an unknown kind or more than
 * SWE_SEEDED_MAX_BUMPS bumps.  No counterpart entry in the reference: it
 * replaces the host loop of initScenario for these workloads, whose host
 * sampling takes seconds at the largest sizes. */
int swe_hip_init_seeded(swe_hip_block* blk, const swe_hip_seeded* workload);

/* computeNumericalFluxes (flavour U: SWE_DimensionalSplittingMpi.cpp:383-456;
 * flavour S: SWE_DimensionalSplitting.cpp:105-179).  Solves every edge,
 * keeps the net updates on the device and returns the block's CFL time step
 * (the value the reference stores in maxTimestep) in *max_timestep. */
int swe_hip_fluxes(swe_hip_block* blk, float* max_timestep);

/* updateUnknowns (flavour U: SWE_DimensionalSplittingMpi.cpp:484-518, dry
 * clamp included; flavour S: SWE_DimensionalSplitting.cpp:187-210) with time
 * step dt (the reference's U blocks use their maxTimestep member, which the
 * driver may have lowered to the global minimum).  It needs a flux pass of
 * the current state (SWE_HIP_ESTATE otherwise, e.g. a second update after
 * one swe_hip_fluxes).
 *
 * Flavour U defers the work: the update is applied by the next call that
 * needs the state; when that is swe_hip_fluxes after swe_hip_apply_bc (the
 * drivers' order), one fused launch applies the update and the boundary
 * conditions as it loads the cells and solves the next edges.  Once a caller
 * has called update(dt returned) -> apply_bc -> fluxes, each swe_hip_fluxes
 * also queues that next launch at once, with the dt it returns, so the GPU
 * works while the host goes round its loop; a different next call (another
 * dt, another order, any other function) drops it, at the cost of the wasted
 * launch (swe_hip_set_lookahead turns this off).  Results are bit-identical
 * to applying every update at once; quiet tiles are skipped as in the device
 * loop (swe_hip_set_quiet_skip). */
int swe_hip_update(swe_hip_block* blk, float dt);

/* nsteps of the single-block global-time-stepping loop of swe_simple.cpp:
 * 215-252 (setGhostLayer, computeNumericalFluxes, getMaxTimestep,
 * updateUnknowns, t += dt) without leaving the device: dt is reduced and
 * applied on the GPU.  *t_inout is the simulated time (advanced in float as
 * the driver does); dt_log (may be NULL) receives nsteps step sizes. */
int swe_hip_run(swe_hip_block* blk, int nsteps, float* t_inout, float* dt_log);

/* swe_hip_run with HIP-event timing on the block's stream.  *total_ms is
 * the elapsed time of the whole loop.  If kernel_ms is not NULL, events are
 * also recorded between launches and kernel_ms[0..2] receive the summed
 * durations of the boundary and quiet-tile plan, first-pass (flux / max or
 * fused) and second-pass (update / apply) launches; kernel_launches[0..2]
 * their launch counts.  The events themselves slow the loop a little.
 * No reference counterpart: measurement hook for the bench. */
int swe_hip_run_timed(swe_hip_block* blk, int nsteps, float* t_inout, float* total_ms, float* kernel_ms,
                      int* kernel_launches);

/* Wait for all work queued on the block's stream. */
int swe_hip_synchronize(swe_hip_block* blk);

/* Device pointers of the current state (for zero-copy interop) and the pitch
 * (floats between consecutive x columns; equal to ny+2). */
int swe_hip_device_pointers(swe_hip_block* blk, float** h, float** hu, float** hv, float** b, int* pitch);

/* Asynchronous snapshots of the state for output writers, double-buffered
 * (the reference writes synchronously from host mirrors at each checkpoint,
 * swe_simple.cpp:187, 257; no counterpart entry).  swe_hip_snapshot queues a
 * copy of the current h, hu, hv (ghosts included, Float2D layout with pitch
 * ny+2) into the next of SWE_HIP_SNAPSHOT_SLOTS slots and returns the slot at
 * once: a device-side copy in the block's stream order, so that later steps
 * may overwrite the state, then the transfer to pinned host arrays on a
 * stream of its own while the steps go on.  If that slot's previous snapshot
 * has not been released yet, it waits for the release first.
 * swe_hip_snapshot_wait waits for the slot's transfer and returns its host
 * arrays (valid until the slot is released); swe_hip_snapshot_release hands
 * the slot back.  Wait and release may come from another thread than the one
 * stepping the block (a writer thread); they do not take the block's lock. */
#define SWE_HIP_SNAPSHOT_SLOTS 2
int swe_hip_snapshot(swe_hip_block* blk, int* slot);
int swe_hip_snapshot_wait(swe_hip_block* blk, int slot, const float** h, const float** hu, const float** hv);
int swe_hip_snapshot_release(swe_hip_block* blk, int slot);

/* Field masks for the layer functions below. */
#define SWE_FIELD_H 1
#define SWE_FIELD_HU 2
#define SWE_FIELD_HV 4
#define SWE_FIELD_B 8

/* In-rank neighbour exchange: dst's ghost line on `side` <- src's copy layer
 * (the interior line next to the shared edge), for the fields in `fields`.
 * side RIGHT: dst[nx+1][1..ny] <- src[1][1..ny]; LEFT: dst[0][j] <- src[src.nx][j];
 * TOP: dst[i][ny+1] <- src[i][1]; BOTTOM: dst[i][0] <- src[i][src.ny].
 * Replaces the CONNECT_WITHIN_RANK copies of
 * SWE_DimensionalSplittingMPIOverdecomp.cpp:273-304 (h, hu, hv) and :208-235
 * (bathymetry).  Ordered after src's queued work and before src's later
 * work; the shared edge lengths must match.  If src's last update is still
 * pending (swe_hip_update), h, hu, hv are copied with that update applied on
 * the fly and src keeps it pending, so its next flux pass stays one fused
 * launch. */
int swe_hip_copy_layer(swe_hip_block* dst, int side, swe_hip_block* src, int fields);

/* Host transfer of one line of length ny (LEFT/RIGHT) or nx (BOTTOM/TOP),
 * indices 1..n: ghost = 1 selects the ghost line on `side`, ghost = 0 the
 * copy layer next to it.  Packed per field in the order h, hu, hv, b for the
 * bits set in `fields`.  Used by the host-side LTS state machine
 * (SWE_Block.hh:375-524), which interpolates ghost lines on the host. */
int swe_hip_read_layer(swe_hip_block* blk, int side, int ghost, int fields, float* out);
int swe_hip_write_layer(swe_hip_block* blk, int side, int ghost, int fields, const float* in);

/* Batch form of the per-edge solver interface
 *   solver::X<float>::computeNetUpdates(hL, hR, huL, huR, bL, bR,
 *                                       hNetL, hNetR, huNetL, huNetR, maxEdgeSpeed)
 * (call sites SWE_DimensionalSplittingMpi.cpp:403-410, 432-439; the solver
 * bodies live in the un-vendored TUM-I5/swe_solvers submodule).  Evaluates n
 * independent edges on the GPU; all arrays are host arrays of n floats.
 * math selects the arithmetic variant (see swe_hip_set_math). */
int swe_hip_solve_edges(int solver, int math, int64_t n, const float* hL, const float* hR, const float* huL, const float* huR,
                        const float* bL, const float* bR, float* hNetL, float* hNetR, float* huNetL, float* huNetR,
                        float* maxSpeed);

/* Tuning of the device loop (swe_hip_run): strategy 0 = recompute (max pass
 * + update pass), 1 = store (flux pass writing 4 net-update sums + apply
 * pass), 2 = fused (one launch per step that applies the previous update on
 * load; the default); xseg = x columns per wavefront strip (4, 8, 10, 12,
 * 16 or 32; the sweeps of strategies 0 and 1 run 4 as 16); -1 keeps the
 * current value.  All strategies and widths give
 * identical results.  Until the caller sets xseg, a block uses the widest of
 * 16 and 12 for which nyt * ceil(nx/xseg) / (CUs * 4) reaches the solver's
 * threshold (nyt = ceil(ny/62); AugRie 12, 4; with a learned middle state
 * attached 5, 3; HLLE, f-wave, Hybrid 2, 1), else 8 (swe_hip_set_surrogate
 * re-picks it).  The fused loop skips quiet tiles (swe_hip_set_quiet_skip)
 * at widths that are multiples of 4. */
int swe_hip_set_tuning(swe_hip_block* blk, int strategy, int xseg);
int swe_hip_get_tuning(swe_hip_block* blk, int* strategy, int* xseg);

/* Strip-width autotuner of a flavour-U block (no counterpart in the
 * reference, whose CPU loops have no such parameter).  Times `steps` steps of
 * the device loop (swe_hip_run, current strategy) from the block's current
 * state at each of the SWE_HIP_TUNE_WIDTHS widths 4, 8, 10, 12, 16, 32, restores
 * the state (h, hu, hv, the simulated time and step count), and keeps the
 * fastest width as if set with swe_hip_set_tuning.  ms_per_step (may be
 * null) receives the per-step times in that order.  With the fused strategy
 * and the HLLE or f-wave solver it also times the recompute strategy (0) at
 * 16 and 32 columns and keeps it if faster: on blocks that do not fit the
 * GPU's caches recompute moves fewer bytes per cell (swe_hip_get_tuning
 * reports the strategy kept).  Results of later runs do not depend on the
 * choice.  Allocates and frees 3 fields of scratch. */
#define SWE_HIP_TUNE_WIDTHS 6
int swe_hip_autotune(swe_hip_block* blk, int steps, float* ms_per_step);

/* Quiet-tile skipping of the fused device loop (strategy 2), on by default,
 * for the augmented solver (AugRie, with or without a learned middle state)
 * at strip widths that are multiples of 4.  A tile of 62 rows x 4 columns
 * whose net-update sums were all +0 on the last two steps, with the adjacent
 * rows and columns of its four neighbours quiet on the last step, would
 * compute the same zero sums and the same CFL maximum again from an
 * unchanged state; the loop skips it and carries its maximum forward.  Results are bit-identical with
 * and without skipping (enable = 0 computes every tile, e.g. for timing the
 * full sweep).  No reference counterpart: the reference's loops visit every
 * edge every step. */
int swe_hip_set_quiet_skip(swe_hip_block* blk, int enable);

/* Look-ahead of the per-call API (see swe_hip_update), on by default;
 * enable = 0 queues each step only when its calls arrive.  Results are the
 * same either way.  No reference counterpart (the reference's calls run
 * synchronously on the host). */
int swe_hip_set_lookahead(swe_hip_block* blk, int enable);

/* nsteps of swe_hip_run (the fused device loop of flavour U with global time
 * stepping, quiet tiles skipped) on caller-owned device arrays, allocating
 * nothing.  No counterpart in the reference: this entry point is for callers
 * that cannot allocate -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or other
 * allocating routine and no host synchronisation; it only queues kernels on
 * `stream` (a hipStream_t; NULL = the null stream).
 *   h, hu, hv     the state in Float2D layout ((nx+2) x (ny+2), device);
 *                 it holds the state after nsteps on return of the work
 *   b             bathymetry (device), ghosts set by the caller
 *                 (swe_hip_apply_bc_bathymetry's rule)
 *   h2, hu2, hv2  a second state buffer of the same size (device), written
 *   boundary      4 BoundaryType values (left, right, bottom, top): OUTFLOW,
 *                 WALL, CONNECT, CONNECT_WITHIN_RANK or PASSIVE (the last
 *                 three keep their ghost cells as given)
 *   dt_log        device array of nsteps floats for the step sizes, or NULL
 * The working memory (control block, net-update sums, quiet-tile record) is
 * one static device array of SWE_HIP_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends -- a square block up to about 1100 x 1100 cells; a
 * block that needs more is rejected with SWE_HIP_EINVAL.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams,
 * other threads) race and give wrong results.  Results are bit-identical to
 * swe_hip_run on a block with the same state, boundaries and solver. */
#define SWE_HIP_NOALLOC_SCRATCH_BYTES (40u << 20)
int swe_hip_run_noalloc(int nx, int ny, float dx, float dy, int solver, const int* boundary, float* h, float* hu,
                        float* hv, const float* b, float* h2, float* hu2, float* hv2, int nsteps, float* dt_log,
                        void* stream);

/* Arithmetic variant of the flavour-U kernels:
 *   0 = IEEE: no contraction, correctly rounded division and square root;
 *       results are bit-identical to an IEEE single-precision CPU evaluation
 *       of the same formulas (the default);
 *   1 = fast: FMA contraction, hardware reciprocal and square root; results
 *       differ from the IEEE variant by rounding only (about 1e-2 relative
 *       L-infinity after 500 steps of a wet/dry dam break, the same order as
 *       the difference between FP32 and FP64 runs of the scheme). */
int swe_hip_set_math(swe_hip_block* blk, int math);

#ifdef __cplusplus
}
#endif

#endif /* SWE_HIP_H */
