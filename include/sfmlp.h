/*
 * sfmlp: the learned surrogate of the hydrodynamics path on MI355X (gfx950).
 *
 * No counterpart in the reference (Stonefish computes hydrodynamics only with
 * the per-face loops).  A small MLP per hull shape maps body-frame features of
 * a body's state (relative linear and angular velocity, direction of gravity,
 * depth of the centre below the local surface, normal of the local surface)
 * to the body-frame force and torque that SolidEntity::ComputeHydrodynamicForces
 * produces (sum of buoyancy, quadratic drag and skin friction after
 * CorrectHydrodynamicForces).  The exact kernels of sfhydro.h generate its
 * training data, audit it and replace it outside its training range.
 *
 * Network: 13 inputs (zero-padded to 32) -> 4 x 256 GELU (exact erf) -> 6
 * outputs (padded to 32).  BF16 weights and activations, FP32 accumulation on
 * the MFMA units, FP32 master weights and Adam moments.  All entry points take
 * device pointers and a hipStream_t (NULL = default stream) and are
 * asynchronous; they return SFM_OK or a negative code, the message of the
 * last error on this thread is sfm_last_error().
 */
#ifndef SFMLP_H
#define SFMLP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFM_OK 0
#define SFM_ERR_INVALID (-1)
#define SFM_ERR_SIZE (-2)
#define SFM_ERR_HIP (-3)

#define SFM_IN 13        /* real inputs */
#define SFM_IN_PAD 32
#define SFM_HIDDEN 256
#define SFM_LAYERS 5     /* weight layers */
#define SFM_OUT 6        /* real outputs */
#define SFM_OUT_PAD 32

typedef struct sfm_net_t* sfm_net;

const char* sfm_last_error(void);

/* Creates a network for batches of up to max_batch samples (a multiple of
 * 256, at most 4,194,304) with weights initialised from `seed`
 * (Kaiming-uniform on the real fan-in, zero biases, zero padding). */
int sfm_create(uint32_t max_batch, uint64_t seed, sfm_net* out);
int sfm_destroy(sfm_net net);

/* Parameter count (202,502 real parameters; storage is padded). */
int sfm_param_count(sfm_net net, uint64_t* real_params);

/* Copies the FP32 master weights and biases of layer l (0..4) in or out
 * (host memory, row-major [out x in] over the padded shape
 * {32,256,256,256,256,32}); `set` 1 copies in (and refreshes the BF16
 * copies), 0 copies out. */
int sfm_layer_io(sfm_net net, int layer, float* weights, float* bias, int set);

#define SFM_MAX_GROUP 16   /* networks per sfm_forward_group call */

/* Inference: x (batch x 32, BF16 bits, row-major, 16-byte aligned, inputs
 * 13..31 zero), y (batch x 6, FP32); batch a multiple of 32 (any size: the
 * fused kernel keeps no per-batch state).  One launch: each workgroup runs a
 * tile of rows through all five layers with the activations in LDS. */
int sfm_forward(sfm_net net, const uint16_t* x, uint32_t batch, float* y, void* stream);

/* Inference of several networks in one launch: x holds counts[0] rows for
 * nets[0], then counts[1] rows for nets[1], ... (each count a multiple of 32,
 * nnets <= SFM_MAX_GROUP); y likewise. */
int sfm_forward_group(const sfm_net* nets, uint32_t nnets, const uint32_t* counts, const uint16_t* x, float* y,
                      void* stream);

/* Inference through the per-layer kernels of the training path (batch <=
 * max_batch); same result as sfm_forward up to FP32 summation order. */
int sfm_forward_layers(sfm_net net, const uint16_t* x, uint32_t batch, float* y, void* stream);

/* One training step on a batch (a multiple of 256): forward, mean-squared-
 * error loss against target (batch x 6, FP32), backward, Adam (lr, beta1,
 * beta2, eps, step counter kept by the network).  loss (device, 1 float)
 * receives the batch loss (sum of squared errors / (batch * 6)). */
int sfm_train_step(sfm_net net, const uint16_t* x, const float* target, uint32_t batch, float lr, float beta1,
                   float beta2, float eps, float* loss, void* stream);

/* The gradients of the last training step (FP32, device): layer l weights
 * [out x in] padded and biases [out]; for checking against a reference. */
int sfm_gradients(sfm_net net, int layer, float* dweights, float* dbias, void* stream);

/* Device pointers of the network's BF16 weights and FP32 biases (b[l]),
 * l = 0..4, for a fused caller (the surrogate step of sfhydro.h).  The
 * weights w[l] ([out x in] padded) are in fragment order: element (r, c) sits
 * at ((((2 (r / 64) + (r / 32) % 2) (in / 16) + c / 16) 64 + r % 32
 * + 32 ((c / 8) % 2)) 8 + c % 8, the order in which the fused kernels'
 * lanes consume them.  They stay valid while the network lives;
 * sfm_train_step updates them in place. */
int sfm_weights(sfm_net net, const uint16_t** w, const float** b);

/* Converts n FP32 values to BF16 bits (round to nearest even), device to device. */
int sfm_to_bf16(const float* x, uint16_t* y, uint64_t n, void* stream);

/* Checks the MFMA operand layout the kernels assume: one 32x32x16 BF16
 * product of a (32x16, row-major) and b (16x32, row-major) into c (32x32,
 * row-major, FP32). */
int sfm_probe_mfma(const uint16_t* a, const uint16_t* b, float* c, void* stream);

/* Average duration (ms) of the MFMA kernels of the last `sfm_train_step` or
 * `sfm_forward` calls made with timing on, and their FLOP count. */
int sfm_timing(sfm_net net, int enable);
int sfm_kernel_ms(sfm_net net, double* ms, double* flops, uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif
