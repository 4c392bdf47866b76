// combine.h -- mPb and combined gradient (reference include/combine.h:15).
#ifndef DMSC_COMBINE_H
#define DMSC_COMBINE_H

// Allocates *p_devMpb and *p_devCombinedg, each [8][cuePitchInFloats].
void combine(int width, int height, int cuePitchInFloats, float* devBg, float* devCga, float* devCgb,
             float* devTg, float** p_devMpb, float** p_devCombinedg, int textonChoice);

#endif
