// stencilMVM.h -- host helpers of the stencil matrix (reference
// include/stencilMVM.h:18-36).
#ifndef DMSC_STENCILMVM_H
#define DMSC_STENCILMVM_H

#include <hip/hip_runtime.h>

#include "Stencil.h"

// Pitch (floats) of a row of `width` floats: a multiple of 64 floats.
int findPitchInFloats(int width);
// W -> L_sym = I - D^-1/2 W D^-1/2 in place; returns a new DEVICE array
// D^-1/2 (float[width*height]).  gridDim, blockDim and nDimension are
// accepted for signature compatibility; the launch shape is the library's.
float* convertMatrix(Stencil* theStencil, dim3 gridDim, dim3 blockDim, int nDimension, float* devMatrix);
// y = A x on the DEVICE (the reference exposes this as a kernel over a bound
// texture; here it is a host call).
void stencilMatrixVector(Stencil* theStencil, float* devMatrix, float* devX, float* devY);
// Selects the device with the most memory (reference: stencilMVM.cu).
void chooseLargestGPU(bool verbose);

#endif
