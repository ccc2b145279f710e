// mlp.h — learned warm start for the contact solver: a 16-64-64-64-3 MLP with
// bf16 operands and fp32 accumulation on MFMA (16x16x16 and, for inference, 16x16x32 bf16), used
// between the solver's row build and its warm-start application.
//
// Input features per contact (16, fp32): normal(3), relative linear velocity at
// the contact (3), relative angular velocity (3), distance, invMass A, invMass B,
// previous normal / lateral-1 / lateral-2 impulses, combined friction.  Output:
// initial normal, lateral-1 and lateral-2 impulses.  Both are normalised with
// per-feature mean and scale held beside the weights.
#pragma once
#include "internal.h"

namespace jbg {

constexpr int kMlpIn = 16, kMlpHid = 64, kMlpOut = 3;
// parameter layout (floats): W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[3][64] b4[3]
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHid;
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW3 = kMlpB2 + kMlpHid;
constexpr int kMlpB3 = kMlpW3 + kMlpHid * kMlpHid;
constexpr int kMlpW4 = kMlpB3 + kMlpHid;
constexpr int kMlpB4 = kMlpW4 + kMlpOut * kMlpHid;
constexpr int kMlpParams = kMlpB4 + kMlpOut;
// normalisation: x' = (x - mean_in) * inv_std_in ; y = y' * std_out + mean_out
// words: mean_in[16] inv_std_in[16] std_out[3] mean_out[3]
constexpr int kMlpNormWords = 2 * kMlpIn + 2 * kMlpOut;
// bf16 MFMA operand fragments (64 lanes x 4 values): 40 forward (W), 36 backward (W^T); then the inference
// fragments of layers 2-4 for the K = 32 product (64 lanes x 8 values): 8 W2, 8 W3, 2 W4
constexpr int kMlpFwdFrags = 40, kMlpFrags = 76, kMlpInferFrags = 18;
constexpr int kMlpPackedWords = kMlpFrags * 256 + kMlpInferFrags * 512;

void mlpSetParams(Ctx& c, const float* params, const float* norm);
void mlpGetParams(Ctx& c, float* params, float* norm);
// predictions for n feature rows (host arrays)
void mlpInfer(Ctx& c, const float* x, int64_t n, float* y);
// the same through a fixed per-device staging buffer (no device allocation); n <= kStaticInferMax.
// mlpStaticInit looks the buffer's device address up for the current device (once per device).
constexpr int64_t kStaticInferMax = 65536;
void mlpStaticInit();
void mlpInferStatic(Ctx& c, const float* x, int64_t n, float* y);
// gradient of the mean squared error (normalised outputs) at the current parameters
void mlpGrad(Ctx& c, const float* x, const float* y, int64_t n, float* grad, float* loss);
// Adam training on the device from host data; normalisation from the data; loss per epoch
void mlpTrain(Ctx& c, const float* x, const float* y, int64_t n, int epochs, int batch, float lr, uint64_t seed,
              float* lossPerEpoch);
// predictions for every valid contact of the current colour slots, written into the
// applied words of the normal and friction rows (replacing the scaled cached impulses)
void mlpWarmStart(Ctx& c, int nslotsManifolds);

}  // namespace jbg
