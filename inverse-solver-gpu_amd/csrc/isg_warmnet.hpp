// isg_warmnet.hpp -- learned warm start for the DBIM total-field solves.
//
// The reference restarts every total-field solve from zero (mom.cpp:324).
// This network maps, per transmitter, the new contrast, the incident field,
// the first-order start tot_prev - alpha dtot and its first-order step
// -alpha dtot (8 real channels; dtot is the Frechet-derivative field the
// step-length phase solved for, mom.cpp:284-290) to a correction of the
// first-order start (2 real channels); the solver then starts from
// first-order start + lambda * correction with lambda chosen per transmitter
// to minimise the initial residual (so the start is never worse than the
// first-order one).  The field channels arrive divided by the transmitter's
// incident-field rms; the correction comes out in units of the rms of the
// transmitter's first-order step alpha dtot (so that small late steps and the
// large first one weigh alike in training).
//
// The image enters space-to-depth by 2 (each 2x2 pixel block becomes one
// position with 4 x 8 = 32 channels) and leaves depth-to-space (16 channels,
// 4 x 2 used): the network runs at half the pixel resolution, 5 positions per
// wavelength at lambda/10 pixels, with nothing discarded.
//
// Network: 4-level convolutional encoder-decoder, 32 base channels.
//   encoder  L0 conv3 32->32, avg-pool,
//            L1 conv3 32->64, avg-pool, L2 conv3 64->128, avg-pool,
//            L3 conv3 128->256 (bottleneck, 1/8 of the network's resolution)
//   decoder  L4 conv1 256->128, L5 conv3 (e2 + up(u3)) 128->128,
//            L6 conv1 128->64,  L7 conv3 (e1 + up(u2)) 64->64,
//            L8 conv1 64->32,   L9 conv3 (e0 + up(u1)) 32->32,
//            L10 conv1 32->16 (channels 0..7 are the output)
// ReLU after every 3x3 convolution; up() is nearest-neighbour 2x.
// Activations are FP32 in HBM, NHWC ([image][row][col][channel]); weights are
// kept as FP32 masters and used as BF16; every convolution rounds its two
// operands to BF16 and accumulates in FP32 on gfx950's
// v_mfma_f32_16x16x32_bf16 (32 input channels per instruction;
// v_mfma_f32_16x16x16_bf16 for the 16-channel input of L10's backward pass).
// Images must have H and W divisible by 16 (three 2x2 poolings after the
// space-to-depth step); the convolutions run on 8x8-position tiles, the last
// ones partial where a level's side is not a multiple of 8.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace isg {

struct NetLayer {
  int ks;      // 1 or 3
  int cin;     // input channels (multiple of 16)
  int cout;    // output channels (multiple of 16)
  int level;   // resolution level, 0 = full
  int relu;    // ReLU on the output
  size_t woff, boff;  // offsets of W [ks][ks][cout][cin] and bias [cout] in the parameter array
};

class WarmNet {
 public:
  static constexpr int kInCh = 8, kOutCh = 2, kLayers = 11;
  static constexpr int kS2D = 2;        // space-to-depth factor at the input
  static constexpr int kTile = 8 * kS2D;  // image sides must be multiples of this (3 pools after s2d)
  static constexpr int kScales = kInCh + 1;  // input channel scales + 1 output scale

  // Workspace for up to max_batch images of max_pixels pixels each.
  WarmNet(int max_batch, long max_pixels, uint64_t seed);
  ~WarmNet();
  bool ok() const { return err_.empty(); }
  const std::string& error() const { return err_; }

  static const std::vector<NetLayer>& layers();
  static size_t num_params();
  int max_batch() const { return max_batch_; }
  long max_pixels() const { return max_pix_; }

  // FP32 master parameters (host copies); set refreshes the BF16 copies and
  // resets the optimiser state.
  void get_params(std::vector<float>& out);
  void set_params(const std::vector<float>& in);
  void get_grads(std::vector<float>& out);
  // Channel scales: in[c] multiplies input channel c, out multiplies the output.
  void set_scales(const float* s) { for (int i = 0; i < kScales; i++) scales_[i] = s[i]; }
  const float* scales() const { return scales_; }

  // in: [batch][H][W][8] FP32, out: [batch][H][W][2] FP32 (device pointers).
  void forward(const float* in, float* out, int batch, int H, int W, hipStream_t s);
  // The same for a caller that writes the network's input itself: packed_input
  // is the space-to-depth input [batch][H/2][W/2][32] (channel 8 (2 dy + dx) + c,
  // already multiplied by the input scales), forward_packed runs the layers and
  // packed_output holds [batch][H/2][W/2][16] (channel 2 (2 dy + dx) + c, c < 2,
  // before the output scale).  packed_input returns null on a bad shape.
  float* packed_input(int batch, int H, int W);
  void forward_packed(int batch, int H, int W, hipStream_t s);
  const float* packed_output() const { return act_[A_OUT]; }
  // Forward + backward of the mean squared error against target
  // [batch][H][W][2]; gradients go to the internal gradient array.  Returns
  // the loss (synchronises s).
  double loss_grad(const float* in, const float* target, int batch, int H, int W, hipStream_t s);
  // Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 masters.
  void adam(float lr, hipStream_t s);

  bool save(const char* path);
  bool load(const char* path);

 private:
  void forward_all(const float* in, int batch, int H, int W, hipStream_t s);
  void forward_layers(int batch, int H, int W, hipStream_t s);
  void conv(int li, const float* x, const float* u, const float* m, float* y, int cstore, bool transpose,
            int batch, int H, int W, hipStream_t s);
  void wgrad(int li, const float* dy, const float* m, const float* x, const float* u, int batch, int H, int W,
             hipStream_t s);
  void refresh_bf16(hipStream_t s);
  void* dalloc(size_t bytes);

  std::string err_;
  int max_batch_ = 0;
  long max_pix_ = 0;
  float scales_[kScales] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
  float* w_ = nullptr;       // FP32 masters
  float* g_ = nullptr;       // gradients
  float* m1_ = nullptr;      // Adam moments
  float* m2_ = nullptr;
  uint16_t* wb_ = nullptr;   // BF16 weights, forward layout
  uint16_t* wt_ = nullptr;   // BF16 weights, input-gradient layout (flipped, transposed)
  float* part_ = nullptr;    // weight-gradient partial sums
  size_t part_len_ = 0;
  double* lpart_ = nullptr;  // loss partials
  float* dscale_ = nullptr;  // device copy of the channel scales
  double* lpart_h_ = nullptr;
  int nlpart_ = 0;
  long step_ = 0;
  // activations (forward) and gradients (backward)
  enum { A_X, A_E0, A_P0, A_E1, A_P1, A_E2, A_P2, A_E3, A_U3, A_D2, A_U2, A_D1, A_U1, A_D0, A_OUT, A_N };
  enum { G_OUT, G_D0, G_IN0, G_U1, G_D1, G_IN1, G_U2, G_D2, G_IN2, G_U3, G_E3, G_P2, G_P1, G_P0, G_N };
  float* act_[A_N] = {};
  float* grd_[G_N] = {};
  std::vector<void*> allocs_;
};

}  // namespace isg
