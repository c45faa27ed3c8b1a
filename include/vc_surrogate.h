/*
 * vc_surrogate.h -- learned surrogate of the medial measure (libvoxelcore.so,
 * gfx950).  No counterpart in the reference: the reference evaluates the
 * measure exactly on its Voronoi complex (include/measureforMA_imp.h:1-4,
 * src/voroinfo.cpp:1432-1574); this proposes the per-voxel measure with a
 * small network and keeps the exact measure as the fallback.
 *
 * Network: 4-level 3-D U-Net, 16/32/48/64 channels, two 3^3 convolutions per
 * level (ReLU), 2^3 average pooling, nearest up-sampling with skip
 * concatenation, 1^3 head with two outputs: lambda / 32 and the log of the
 * predicted relative error.  Inputs per voxel: occupancy and
 * sqrt(EDT) / 32.  BF16 weights and activations on MFMA (16x16x32), FP32
 * accumulation, FP32 master weights, gradients and Adam moments.  678,786
 * parameters, flat FP32 order: per convolution W[27][Cin][Cout] (tap
 * (dx+1)*9 + (dy+1)*3 + (dz+1)) then b[Cout], layers L0a L0b L1a L1b L2a L2b
 * L3a L3b L4a L4b L5a L5b L6a L6b, then the head W[16][2], b[2].
 *
 * The thinning stays exact: the surrogate only supplies the float measure
 * the eligibility test reads (as vc_set_measure_f32), so the (26,6)
 * simple-point test preserves topology whatever the measure.
 */
#ifndef VC_SURROGATE_H
#define VC_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vc_ctx vc_ctx;
typedef struct vc_vol vc_vol;
typedef struct vc_surrogate vc_surrogate;

#define VC_SURROGATE_PARAMS 678786
#define VC_SURROGATE_HALO 56 /* tile halo (voxels) >= receptive radius 51 */

typedef struct vc_surrogate_stats {
    int64_t inside;   /* inside voxels                                 */
    int64_t accepted; /* voxels that keep the surrogate measure        */
    int64_t fallback; /* voxels given the exact measure                */
    float ms_infer;   /* network inference (GPU time)                  */
    float ms_exact;   /* exact measure for the fallback (0 if none)    */
    float ms_mix;     /* gate + mix                                    */
} vc_surrogate_stats;

/* He-uniform initial weights from seed (splitmix64, see oracle/unet_ref.py
 * init_params for the exact rule), zero biases. */
int vc_surrogate_create(vc_ctx* ctx, uint64_t seed, vc_surrogate** out);
int vc_surrogate_destroy(vc_ctx* ctx, vc_surrogate* s);
int vc_surrogate_get_params(vc_ctx* ctx, vc_surrogate* s, float* host,
                            int64_t n);
int vc_surrogate_set_params(vc_ctx* ctx, vc_surrogate* s, const float* host,
                            int64_t n);
/* weights file: 8-byte magic "VCUNET01", int64 count, count float32 */
int vc_surrogate_save(vc_ctx* ctx, vc_surrogate* s, const char* path);
int vc_surrogate_load(vc_ctx* ctx, vc_surrogate* s, const char* path);

/* Loss and gradient (host, VC_SURROGATE_PARAMS floats, may be NULL) of the
 * network on one volume whose vc_edt and vc_measure have run; the target is
 * the exact measure sqrt(lam2) / 32 on inside voxels.  No update.  Volume
 * axes are padded with empty voxels to multiples of 8. */
int vc_surrogate_gradients(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                           float* grads, float* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the mean gradient over
 * n volumes; loss (may be NULL) receives the mean loss. */
int vc_surrogate_train_step(vc_ctx* ctx, vc_surrogate* s, vc_vol* const* vols,
                            int n, float lr, float* loss);
/* Network outputs for every voxel of a volume whose vc_edt has run: lambda
 * (voxels) and predicted relative error, host arrays of N floats (either may
 * be NULL).  tile_core = 0: the whole volume in one pass; else cubic tiles
 * of that core edge (multiple of 8) with a VC_SURROGATE_HALO halo, which
 * give the same values as the whole-volume pass. */
int vc_surrogate_infer(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                       int tile_core, float* lambda_out, float* relerr_out);
/* The measure the thinning then uses (installed as by vc_set_measure_f32):
 * on inside voxels the surrogate lambda where the predicted relative error is
 * <= max_rel_err, else the exact sqrt(lam2) (vc_measure is run on the volume
 * when any voxel falls back and it has no exact measure yet).  stats may be
 * NULL. */
int vc_surrogate_measure(vc_ctx* ctx, vc_surrogate* s, vc_vol* vol,
                         float max_rel_err, int tile_core,
                         vc_surrogate_stats* stats);
/* Copy of an activation of the last forward pass as float (which = layer
 * index 0..13 for the ReLU outputs of L0a..L6b, 14 for the two head
 * outputs), channels last, over the padded volume.  After an inference
 * (vc_surrogate_infer / vc_surrogate_measure) layers 0 and 13 are not
 * available (VC_ERR_INVALID): inference computes layers 0 and 1 in one kernel
 * and evaluates the head inside layer 13's convolution, storing neither;
 * after a training pass (vc_surrogate_gradients / train) every layer is. */
int vc_surrogate_activation(vc_ctx* ctx, vc_surrogate* s, int which,
                            float* host, int64_t n);

#ifdef __cplusplus
}
#endif
#endif /* VC_SURROGATE_H */
