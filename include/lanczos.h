// lanczos.h -- generalized eigensolve (D - W) v = lambda D v (reference
// include/lanczos.h:17).
#ifndef DMSC_LANCZOS_H
#define DMSC_LANCZOS_H

#include "Stencil.h"

// devMatrix (W, overwritten by L_sym) on the DEVICE; allocates the HOST
// eigenvalues (malloc, float[nEigNum]) and the DEVICE eigenvectors
// (hipMalloc, [nEigNum][width*height]), as the reference does.  Plain FP32
// reference algorithm.
void generalizedEigensolve(Stencil& myStencil, float* devMatrix, int matrixPitchInFloats, int nEigNum,
                           float** p_eigenvalues, float** p_eigenvectors, float tolerance);

#endif
