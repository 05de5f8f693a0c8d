// Learned preconditioner network (see pcnet.hip).
#pragma once

#include "lt_internal.h"

typedef struct lt_pcnet_s *lt_pcnet;

namespace lt {

struct PcNet {
    static constexpr int kLayers = 8;
    static const int kCin[kLayers], kCout[kLayers], kLevel[kLayers];
    enum Tensor { X0, E1, P1, E2, P2, E3, P3, E4, U4, A5, D3, U3, A6, D2, U2, A7, D1, Y, kTensors };

    int H = 0, W = 0, nbcap = 0;
    long step = 0;
    long woff[kLayers] = {}, boff[kLayers] = {};
    DevBuf<float> params, grads, adam_m, adam_v;  // FP32 master weights, [co][ci][3][3] then bias, per layer
    DevBuf<__bf16> wbf[kLayers];                   // forward weights, BF16 [32*ceil(co/32)][16*ceil(9ci/16)]
    DevBuf<__bf16> wflip[kLayers];                 // data-gradient weights (channels swapped, taps flipped)
    DevBuf<float> act[kTensors], gact[kTensors];
    long tsize[kTensors] = {};
    DevBuf<float> g0, g1, g2, part;
    DevBuf<double> stats, gnorm;

    static long param_count();
    lt_status create(int H, int W, int batch_capacity);
    lt_status init_random(unsigned seed);
    lt_status pack();  // FP32 master weights -> BF16 operand layouts
    lt_status forward(int nb);   // act[X0] -> act[Y]
    lt_status backward(int nb);  // gact[Y] -> grads (after forward)
    lt_status loss(int nb, const float *target, double *loss_out);  // sets gact[Y]
    lt_status adam_step(float lr, float max_norm);

    lt_status conv(const float *x, int ci, int co, int lev, int nb, const __bf16 *w, const float *bias, bool relu,
                   const float *skip, float *actout, float *y);
    lt_status conv_dw(const float *dz, const float *x, int layer, int nb);
};

PcNet *pcnet_of(lt_pcnet h);

}  // namespace lt
