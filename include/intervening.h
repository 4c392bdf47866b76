// intervening.h -- intervening-contour affinity matrix (reference
// include/intervening.h:15).
#ifndef DMSC_INTERVENING_H
#define DMSC_INTERVENING_H

#include "Stencil.h"

// Allocates *p_devMatrix: DIA [stencil area][theStencil.getMatrixPitchInFloats()].
void intervene(Stencil& theStencil, float* devMPb, float** p_devMatrix, float sigma = 0.1f);

#endif
