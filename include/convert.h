// convert.h -- colour conversion, mirroring and quantisation (reference
// include/convert.h:13-140).  Functions taking a `**` output allocate it on
// the DEVICE (hipMalloc); the caller frees it.
#ifndef DMSC_CONVERT_H
#define DMSC_CONVERT_H

typedef unsigned int uint;

// Binary PPM (P6, maxval 255) -> DEVICE image, one uint per pixel with R in
// bits 0-7, G in 8-15, B in 16-23.
void loadPPM_rgbU(char* filename, uint* p_width, uint* p_height, uint** p_devRgbU);
void rgbUtoLab3F(uint width, uint height, float gamma, uint* devRgbU, float** p_devL, float** p_devA, float** p_devB);
void normalizeLab(uint width, uint height, float* devL, float* devA, float* devB);
void rgbUtoGreyF(uint width, uint height, uint* devRgbU, float** p_devGrey);
void mirrorImage(uint width, uint height, uint border, float* devInput, float** p_devOutput);
void mirrorImage(uint width, uint height, uint border, float* devInput, float* p_devOutput);
void mirrorImage(uint width, uint height, uint border, int* devInput, int** p_devOutput);
void mirrorImage(uint width, uint height, uint border, int* devInput, int* p_devOutput);
void unMirrorImage(uint width, uint height, uint border, float* devInput, float** p_devOutput);
void unMirrorImage(uint width, uint height, uint border, float* devInput, float* p_devOutput);
void quantizeImage(uint width, uint height, uint nbins, float* devInput, int** p_devOutput);
void quantizeImage(uint width, uint height, uint nbins, float* devInput, int* devOutput);

#endif
