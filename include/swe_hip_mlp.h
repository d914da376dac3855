/*
 * swe_hip_mlp.h -- C ABI of the learned middle state of libswe_hip.so.
 *
 * No reference counterpart: the reference's augmented Riemann solver computes
 * its middle state by Newton iteration inside the un-vendored solver
 * submodule (called from SWE_DimensionalSplittingMpi.cpp:403-410, 432-439).
 * This interface trains a small network on middle states that the library's
 * own exact solver logs, and lets a block's AugRie replace its Newton
 * iteration by a guess (linearised middle state, else the network's) with one
 * FP32 Newton evaluation as residual check and the exact iteration as
 * fallback.
 *
 * Network: 6 -> 64 -> 64 -> 2, ReLU.  Inputs of an edge with both sides wet
 * (after the solver's wet/dry handling), with hs = (hL + hR)/2 and
 * cs = sqrt(9.81 hs):
 *   x = (hL/hs, hR/hs, uL/cs, uR/cs, clamp((bR - bL)/hs, -1, 1), (uR - uL)/cs)
 * outputs y = (h* / hs - (1 - x5 / 2), u* / cs - ((x2 + x3) / 2 - (x1 - x0) / 2)):
 * the corrections to the linearised (acoustic) Riemann solution, with
 * u* = (s1* + s2*) / 2 of the middle-state speeds.
 * FP32 parameter vector (SWE_MLP_PARAMS floats), in order:
 *   W1[6][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
 * (W[in][out]).  A sample is SWE_MLP_SAMPLE_FLOATS floats: x[6], y[2].
 * Training: mean squared error over both outputs, Adam, FP32 forward and
 * backward kernels.  Inference inside the solver: BF16 operands, FP32
 * accumulation on the matrix cores.
 *
 * Conventions (return codes, swe_hip_last_error) are those of swe_hip.h.
 */
#ifndef SWE_HIP_MLP_H
#define SWE_HIP_MLP_H

#include <stdint.h>

#include "swe_hip.h"
#include "swe_hip_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SWE_MLP_PARAMS 4738
#define SWE_MLP_SAMPLE_FLOATS 8

typedef struct swe_hip_mlp swe_hip_mlp;

/* A network with seeded He-initialised weights and an empty device sample
 * store of `capacity` samples. */
int swe_hip_mlp_create(int64_t capacity, uint64_t seed, swe_hip_mlp** out);
/* Frees the network; refused with SWE_HIP_ESTATE while a block or batch has
 * it attached (the blocks read its weights on the device).  Setting new
 * weights (swe_hip_mlp_set_params, swe_hip_mlp_train) first waits for the
 * work queued on the attached blocks' and batches' streams. */
int swe_hip_mlp_destroy(swe_hip_mlp* m);

/* FP32 master parameters (SWE_MLP_PARAMS floats, layout above). */
int swe_hip_mlp_get_params(swe_hip_mlp* m, float* params);
int swe_hip_mlp_set_params(swe_hip_mlp* m, const float* params);

/* Log samples from the current state of a flavour-U block: every edge of the
 * block whose AugRie middle state needs the Newton iteration (both sides wet,
 * not two rarefactions) -- with hard_only != 0 only those whose linearised
 * guess fails the residual check -- is kept with probability `keep` (seeded
 * hash of the edge index) and its inputs and exact FP32 middle state are
 * appended to the store (up to its capacity).  *added (may be NULL) receives
 * the number appended. */
int swe_hip_mlp_log_block(swe_hip_mlp* m, swe_hip_block* blk, float keep, uint64_t seed, int hard_only,
                          int64_t* added);

/* Append host samples / read samples back / count / empty the store. */
int swe_hip_mlp_add_samples(swe_hip_mlp* m, int64_t n, const float* samples);
int swe_hip_mlp_read_samples(swe_hip_mlp* m, int64_t first, int64_t n, float* samples);
int swe_hip_mlp_sample_count(swe_hip_mlp* m, int64_t* count);
int swe_hip_mlp_clear_samples(swe_hip_mlp* m);

/* `steps` Adam steps (learning rate lr, cosine decay to lr/20, beta 0.9 /
 * 0.999) on mini-batches of `batch` samples drawn from the store with a
 * seeded generator.  loss_log (may be NULL) receives the mean loss of every
 * step. */
int swe_hip_mlp_train(swe_hip_mlp* m, int steps, int batch, float lr, uint64_t seed, float* loss_log);

/* Gradient of the mean loss over the n given host samples with respect to
 * the parameters (SWE_MLP_PARAMS floats), computed by the training kernels
 * without updating anything; *loss (may be NULL) receives the mean loss. */
int swe_hip_mlp_gradient(swe_hip_mlp* m, int64_t n, const float* samples, float* grad, float* loss);

/* Network outputs for n host feature vectors (6 floats each) by the solver's
 * BF16 matrix-core inference: out receives 2 floats per sample. */
int swe_hip_mlp_infer(swe_hip_mlp* m, int64_t n, const float* features, float* out);

/* The device feature function for n edges (host arrays): 6 floats per edge. */
int swe_hip_mlp_features(int64_t n, const float* hL, const float* hR, const float* uL, const float* uR,
                         const float* bL, const float* bR, float* features);

/* Middle states of n edges (hL, hR, uL, uR, bL, bR host arrays) as the
 * surrogate solver computes them: linearised guess, network guess for the
 * waves (64 consecutive edges) where it fails somewhere, residual check,
 * exact fallback.  accepted[k] = 1 where the corrected linearised guess was
 * kept, 2 where the corrected network guess was kept, 0 where the exact
 * middle state was computed (no iteration needed, or fallback). */
int swe_hip_mlp_middle(swe_hip_mlp* m, int64_t n, const float* hL, const float* hR, const float* uL, const float* uR,
                       const float* bL, const float* bR, float* hm, float* s1m, float* s2m, int32_t* accepted);

/* Per-edge AugRie with the learned middle state (same arrays as
 * swe_hip_solve_edges). */
int swe_hip_mlp_solve_edges(swe_hip_mlp* m, int64_t n, const float* hL, const float* hR, const float* huL,
                            const float* huR, const float* bL, const float* bR, float* hNetL, float* hNetR,
                            float* huNetL, float* huNetR, float* maxSpeed);

/* Attach the network to an AugRie block or batch (NULL detaches): its steps
 * then use the learned middle state.  The network must outlive the
 * attachment; later parameter changes take effect at once.  Needs the IEEE
 * arithmetic variant.  The fused device loops (the default strategy of
 * swe_hip_run and of the batch) defer the edges whose linearised guess fails
 * its residual check to a second launch per step that runs the network there,
 * with the same results as solving them inside the sweep; for that a queue of
 * 8 bytes per edge of the block (batch: of all blocks) is allocated here (the
 * batch) or at the first fused run (a block) and kept until destruction. */
int swe_hip_set_surrogate(swe_hip_block* blk, swe_hip_mlp* m);
int swe_hip_batch_set_surrogate(swe_hip_batch* bt, swe_hip_mlp* m);

#ifdef __cplusplus
}
#endif

#endif /* SWE_HIP_MLP_H */
