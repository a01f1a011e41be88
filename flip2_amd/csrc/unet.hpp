// Learned preconditioner: a 3-level 3D U-Net (16/32/64 channels, 3^3
// convolutions, BF16 weights and activations, FP32 accumulation) applied to
// the pressure residual on the dense n^3 grid, trained on GPU with the
// project's own forward/backward kernels and Adam.
//
// Network (x = input channels: r/s and the fluid mask, padded to 16):
//   a0 = relu(conv(x, 16->16))   b0 = relu(conv(a0, 16->16))      level 0 (n)
//   a1 = relu(conv(pool(b0), 16->32))  b1 = relu(conv(a1, 32->32)) level 1 (n/2)
//   a2 = relu(conv(pool(b1), 32->64))  b2 = relu(conv(a2, 64->64)) level 2 (n/4)
//   c1 = relu(conv([up(b2), b1], 96->32))  d1 = relu(conv(c1, 32->32))
//   c0 = relu(conv([up(d1), b0], 48->16))  d0 = relu(conv(c0, 16->16))
//   z  = s * mask * (w . d0 + b)            (s = RMS of r over fluid cells)
#pragma once

#include "common.hpp"
#include "device.hpp"
#include "unet_kernels.hpp"

#include <string>
#include <vector>

namespace flip2 {

struct UNetLayer {
    int cin, cout, level;
    size_t w, b;  // offsets (floats) into the parameter buffer
};

class UNet {
public:
    explicit UNet(int n, hipStream_t stream = nullptr);
    ~UNet();
    UNet(const UNet&) = delete;
    UNet& operator=(const UNet&) = delete;

    int gridSize() const { return n_; }
    size_t numParams() const { return numParams_; }
    void initParams(uint64_t seed);
    void setParams(const float* host, size_t count);
    void getParams(float* host, size_t count);

    // Input assembly on the dense grid: ch0 = r / s, ch1 = mask; out = z.
    uint16_t* input() { return in0_.devPtr(); }          // [n^3][16] (read by the first layer's backward)
    uint16_t* inputCompact() { return in0c_.devPtr(); }  // [n^3][2]: the channels the forward reads
    uint8_t* mask() { return mask_.devPtr(); }
    float* scale() { return scale_.devPtr(); }  // device scalar s
    float* output() { return out_.devPtr(); }   // FP32 [n^3]

    // keepActivations: store every activation the backward pass reads; without
    // it the decoder reads its concatenated inputs in place and the output
    // head runs in the last convolution (same values, less traffic).
    void forward(hipStream_t s, bool keepActivations = true);
    // Level masks (and, if sparse, tile flags) from mask(); call after the
    // input assembly of a new system, before forward().  A forward that
    // stores activations needs sparse = false.
    void setInputMasks(hipStream_t s, bool sparse);
    // Accumulates parameter gradients for dL/dout (FP32 [n^3]).
    void backward(const float* dOut, hipStream_t s);
    void zeroGrad(hipStream_t s);
    void adamStep(float lr, hipStream_t s);
    const float* gradDev() const { return grad_.devPtr(); }

    // Preconditioner apply on the solver's active-brick vectors (brickSlot:
    // brick -> slot, -1 inactive).  newSystem: the fluid cells may differ
    // from the previous apply's (masks and tile flags are rebuilt, z is
    // zeroed); otherwise only r changed.  z is written at fluid cells only.
    // normPairs: null, or the per-slot (sum of r^2, fluid count) pairs of r.
    // stop: null, or a flag that makes the apply's kernels (other than those
    // of a newSystem apply's mask rebuild) return at once while it is set.
    void applyActive(uint32_t numActive, const uint32_t* activeBricks, const int32_t* brickSlot,
                     const uint8_t* cellType, const float* r, float* z, int gridN, bool newSystem,
                     const double* normPairs, hipStream_t s, const uint32_t* stop = nullptr);

    const std::vector<UNetLayer>& layers() const { return layers_; }

private:
    void conv(int layer, const uint16_t* X, uint16_t* Y, const nn::ConvMask& cm, hipStream_t s);
    void convBackward(int layer, float* dY, const uint16_t* Y, const uint16_t* X, float* dX, hipStream_t s);

    int n_, n1_, n2_;
    hipStream_t stream_;
    std::vector<UNetLayer> layers_;
    size_t headW_ = 0, headB_ = 0, numParams_ = 0;
    int step_ = 0;
    DeviceVec<float> params_, grad_, adamM_, adamV_;
    DeviceVec<uint16_t> paramsBf_;
    // activations (BF16, channels-last)
    DeviceVec<uint16_t> in0_, in0c_, a0_, b0_, p0_, a1_, b1_, p1_, a2_, b2_, u1_, c1_, d1_, u0_, c0_, d0_;
    DeviceVec<uint8_t> mask_;
    // masks of levels 1 and 2; per level, tile flags of the current and the
    // previous sparse forward (two sets, swapped each time) and all-ones flags
    DeviceVec<uint8_t> mask1_, mask2_;
    DeviceVec<uint8_t> tileFlags_[2][3], tileOnes_[3];
    int flagSet_ = 0;
    bool tilesDirty_ = true;  // a dense forward (or allocation) left values everywhere
    nn::ConvMask cm_[3] = {};
    DeviceVec<uint32_t> listCompute_[3], listClear_[3], listCounts_[3];
    nn::TileList tl_[3] = {};
    DeviceVec<float> scale_, out_;
    // gradients (FP32)
    DeviceVec<float> gd0_, gc0_, gu0_, gd1_, gc1_, gu1_, gb2_, ga2_, gp1_, gb1_, ga1_, gp0_, gb0_, ga0_;
    DeviceVec<double> red_;
    DeviceVec<uint16_t> dPb_, WT_;  // BF16 gradient of the current layer, transposed weights
    // set by applyActive for its forward: the first layer reads the solver's
    // slots, the head writes them
    nn::SlotInput slotIn_{};
    float* slotOut_ = nullptr;
};

// Dense-grid helpers shared by the trainer and the preconditioner apply.
namespace nn {
// Active-brick vector -> dense grid: in0c ([n^3][2]: r/s, fluid mask),
// mask; s = RMS of r over fluid cells is computed on the device into *scale.
// clear: zero in0c and mask everywhere first (the fluid cells changed).
void launch_scatter_input(uint32_t numActive, const uint32_t* activeBricks, const uint8_t* cellType, const float* r,
                          int n, uint16_t* in0c, uint8_t* mask, float* scale, double* red, bool clear, hipStream_t s);
// *scale = s (as launch_scatter_input) without the scatter.
void launch_input_scale(uint32_t numActive, const uint8_t* cellType, const float* r, double* red, float* scale,
                        hipStream_t s);
// The same from `count` precomputed (sum of r^2, count) pairs.
void launch_pairs_scale(const double* pairs, uint32_t count, double* red, float* scale, hipStream_t s,
                        const uint32_t* stop = nullptr);

// Dense copies of a solver system: b, solution x, operator coefficients
// (u16 diag | mask << 3, brick-tiled order in, dense [z][y][x] out).
void launch_dense_system(uint32_t numActive, const uint32_t* activeBricks, const uint16_t* coef, const double* b,
                         const double* x, int n, float* bD, float* xD, uint16_t* coefD, hipStream_t s);
// Dense copy of an FP32 active-brick vector and the operator coefficients
// (fluid cells only; the caller zeroes vD and coefD).
void launch_dense_vector(uint32_t numActive, const uint32_t* activeBricks, const uint16_t* coef, const float* v, int n,
                         float* vD, uint16_t* coefD, hipStream_t s);
// y = A x on the dense grid (coefficient mask bits: -x, +x, -y, +y, -z, +z).
void launch_dense_apply_A(const uint16_t* coefD, const float* x, float* y, int n, hipStream_t s);
// Training loss pieces: res = A z - b; returns via red[0] = ||res||^2,
// red[1] = ||b||^2; then dOut = 2 A res / ||b||^2.
void launch_residual(const uint16_t* coefD, const float* z, const float* bD, float* res, int n, double* red,
                     hipStream_t s);
void launch_loss_grad(const uint16_t* coefD, const float* res, const double* red, float* dOut, int n, hipStream_t s);
// Damped-Jacobi smoothing around the network on the dense grid:
// r1 = r - A (w D^-1 r);  z = x + w D^-1 (r - A x) with x = w D^-1 r + y.
// The adjoint of y -> z is g -> g - A (w D^-1 g), i.e. the presmooth map.
void launch_dense_presmooth(const uint16_t* coefD, const float* r, float omega, float* r1, int n, hipStream_t s);
void launch_dense_postsmooth(const uint16_t* coefD, const float* r, const float* y, float omega, float* z, int n,
                             hipStream_t s);
// Network input from a dense right-hand side: in0c ([n^3][2]) always, in0
// ([n^3][16], for the first layer's weight gradient) unless null; mask; scale.
void launch_dense_input(const float* bD, const uint16_t* coefD, int n, uint16_t* in0, uint16_t* in0c, uint8_t* mask,
                        float* scale, double* red, hipStream_t s);
}  // namespace nn

}  // namespace flip2
