// localcues.h -- Bg/Cga/Cgb/Tg local cues (reference include/localcues.h:20).
#ifndef DMSC_LOCALCUES_H
#define DMSC_LOCALCUES_H

// Allocates the four cue stacks, each [24][*cuePitchInFloats] (scale-major,
// then orientation), on the DEVICE.
void localCues(int width, int height, float* devL, float* devA, float* devB, int* devTextons, float** devBg,
               float** devCga, float** devCgb, float** devTg, int* cuePitchInFloats, int textonChoice);

#endif
