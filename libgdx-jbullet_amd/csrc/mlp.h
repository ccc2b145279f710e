// mlp.h — learned warm start for the contact solver: a 16-64-64-64-3 MLP with
// bf16 weights and fp32 accumulation on MFMA, used between the solver's row
// build and its warm-start application.
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
// normalisation: x' = (x - mean_in) * inv_std_in, y = y' * std_out + mean_out
constexpr int kMlpNormWords = 2 * kMlpIn + 2 * kMlpOut;

// predictions for every valid contact of the current colour slots, written into the
// applied words of the normal and friction rows (replacing the 0.85 * cached impulses)
void mlpWarmStart(Ctx& c, int nslotsManifolds);

}  // namespace jbg
