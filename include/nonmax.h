// nonmax.h -- oriented non-maximum suppression (reference include/nonmax.h:16).
#ifndef DMSC_NONMAX_H
#define DMSC_NONMAX_H

// devPB: [8][pitchFloats]; devNMax: caller-allocated float[width*height].
void nonMaxSuppression(int width, int height, float* devPB, int pitchFloats, float* devNMax);

#endif
