// Kernels of the learned preconditioner's U-Net (unet_kernels.hip).
// Activations are BF16 channels-last ([z][y][x][C]); gradients FP32.
#pragma once

#include "common.hpp"

namespace flip2 {
namespace nn {

// Masking of a convolution's output (all fields may be null):
//   voxel: the level's mask; Y = mask ? act(conv + bias) : 0;
//   tiles: per 16x4x4 output tile (x fastest), non-zero where the tile holds
//          a masked-in voxel; other tiles are not computed ...
//   prevTiles: ... and, where this flag is set (the tile may still hold
//          values from an earlier launch), are written with zeros.
struct ConvMask {
    const uint8_t* voxel;
    const uint8_t* tiles;
    const uint8_t* prevTiles;
};
// Tiles of a level to compute and to clear (device lists in x-fastest tile
// order, with device counts), for the list-driven convolution.
struct TileList {
    const uint32_t* compute;
    const uint32_t* numCompute;
    const uint32_t* clear;
    const uint32_t* numClear;
    uint32_t tilesX, tilesY;
    // null, or a flag that makes the list convolutions return at once
    // (the solver's converged flag: work queued after convergence)
    const uint32_t* stop = nullptr;
};
// Y = act(conv3x3x3(X, W) + bias), MFMA BF16 with FP32 accumulation.
// W: [27][cout][cin] BF16 (tap = (dz*3 + dy)*3 + dx, offsets d-1).
void launch_conv3_fwd(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* Y, int n, int cin, int cout,
                      bool relu, const ConvMask& mask, hipStream_t s);
// Arguments of the 1x1 output head fused into the last convolution.
struct HeadArgs {
    const float* w;
    const float* b;
    const uint8_t* mask;
    const float* scale;
    float* out;                // dense [n^3] output, or
    float* outSlots;           // the solver's brick-slot layout (slot * 512 + cell), masked voxels only
    const int32_t* brickSlot;  // brick -> slot (bricks of 8^3 voxels, n/8 per axis) for outSlots
};
// Input of the first convolution read from the solver's brick-slot layout:
// voxel v of a brick with slot s >= 0 is r[s * 512 + local]; the residual
// channel is bf16(r / scale[0]) and the mask channel 1 where cellType == 1
// (fluid), both 0 elsewhere.
struct SlotInput {
    const float* r;
    const uint8_t* cellType;
    const int32_t* brickSlot;
    const float* scale;
};
// Decoder convolution reading concat(up2(Xcoarse) [ca channels], Xskip [cb
// channels]) without materialising it; Y = ReLU(conv + bias).
void launch_conv3_fwd_upcat(const uint16_t* Xskip, const uint16_t* Xcoarse, const uint16_t* W, const float* bias,
                            uint16_t* Y, int n, int ca, int cb, int cout, bool relu, const ConvMask& mask,
                            hipStream_t s);
// First layer: 16 output channels from an input whose only non-zero
// channels are the first two, given compactly as X2 [voxel][2] BF16; W is the
// layer's [27][16][16] weights (input channels 2..15 unused).
void launch_conv3_in2(const uint16_t* X2, const uint16_t* W, const float* bias, uint16_t* Y, int n, bool relu,
                      const ConvMask& mask, hipStream_t s);
// The same over the list of tiles with fluid (tl), zeroing the tiles in tl's
// clear list; vmask: the level's voxel mask (outputs outside it are zero).
// slots.r != null: the input comes from slots instead of X2.
void launch_conv3_in2_list(const uint16_t* X2, const SlotInput& slots, const uint16_t* W, const float* bias,
                           uint16_t* Y, int n, bool relu, const uint8_t* vmask, const TileList& tl, hipStream_t s);
// 16 -> 16 convolution with the output head in its epilogue (Y may be null).
void launch_conv3_fwd_head(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* Y, int n, int cin,
                           bool relu, const HeadArgs& head, const ConvMask& mask, hipStream_t s);
// List-driven convolution (the preconditioner apply): computes the tiles in
// tl.compute and zeroes those in tl.clear; vmask as ConvMask::voxel.  ca > 0:
// input = concat(up2(Xcoarse) [ca], X [cin - ca]).  split: workgroups per
// tile (output channels divided among them).  head as launch_conv3_fwd_head.
// pooled != null: Y is also 2x2x2 average-pooled into pooled ([voxel][cout]
// at n / 2), the same values as launch_avgpool on the whole of Y.
void launch_conv3_list(const uint16_t* X, const uint16_t* Xcoarse, const uint16_t* W, const float* bias, uint16_t* Y,
                       uint16_t* pooled, int n, int cin, int ca, int cout, int split, bool relu, const HeadArgs& head,
                       const uint8_t* vmask, const TileList& tl, hipStream_t s);
// Lists of the tiles with cur set (compute) and with cur clear but prev set
// (clear); prev may be null.  counts[0], counts[1] receive their lengths.
void launch_tile_lists(const uint8_t* cur, const uint8_t* prev, uint32_t tiles, uint32_t* compute, uint32_t* clear,
                       uint32_t* counts, hipStream_t s);
void conv_tile_dims(int n, uint32_t& tilesX, uint32_t& tilesY);
// Coarse mask: out[v] = any of the 8 children of v in m (n^3 -> (n/2)^3).
void launch_mask_pool(const uint8_t* m, int n, uint8_t* out, hipStream_t s);
// Per-tile flags of a level mask for the convolution tiles (see ConvMask).
void launch_tile_flags(const uint8_t* m, int n, uint8_t* flags, hipStream_t s);
uint32_t conv_tiles(int n);
// dX (FP32) = conv3x3x3^T(dP) from the BF16 gradient dPb; WT (27*cin*cout
// BF16 scratch) receives the transposed, mirrored weights.
void launch_conv3_dgrad(const uint16_t* dPb, const float* W32, uint16_t* WT, float* dX, int n, int cin, int cout,
                        hipStream_t s);
// dW += sum_v dP (x) X(v + tap) (MFMA on dPb and X), dB += sum_v dP (FP32).
void launch_conv3_wgrad(const uint16_t* dPb, const float* dP, const uint16_t* X, float* dW, float* dB, int n, int cin,
                        int cout, hipStream_t s);
// g <- g * (act > 0) in place; gb = BF16 copy.
void launch_relu_mask_bf16(float* g, const uint16_t* act, uint16_t* gb, long long count, hipStream_t s);
void launch_avgpool(const uint16_t* X, uint16_t* P, int n, int C, hipStream_t s);
void launch_avgpool_bwd(const float* dP, float* dX, int n, int C, hipStream_t s);  // dX += ...
void launch_upcat(const uint16_t* A, const uint16_t* B, uint16_t* U, int n, int Ca, int Cb, hipStream_t s);
void launch_upcat_bwd(const float* dU, float* dA, float* dB, int n, int Ca, int Cb, hipStream_t s);  // +=
void launch_head_fwd(const uint16_t* X, const float* w, const float* b, const uint8_t* mask, const float* scale,
                     float* out, long long nvox, int C, hipStream_t s);
void launch_head_bwd(const float* dOut, const uint16_t* X, const float* w, const uint8_t* mask, const float* scale,
                     float* dX, float* dw, float* db, long long nvox, int C, hipStream_t s);
void launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, int count, float lr, float b1, float b2,
                 float eps, int t, hipStream_t s);
void launch_to_bf16(const float* a, uint16_t* b, long long count, hipStream_t s);
void launch_bf16_to_f32(const uint16_t* a, float* b, long long count, hipStream_t s);
void launch_zero_f32(float* a, long long count, hipStream_t s);

}  // namespace nn
}  // namespace flip2
