// spectralPb.h -- oriented spectral Pb (reference include/spectralPb.h:19).
#ifndef DMSC_SPECTRALPB_H
#define DMSC_SPECTRALPB_H

// eigenvalues on the HOST, eigenvectors [nvec][xdim*ydim] on the DEVICE (as
// generalizedEigensolve returns them); results: caller-allocated DEVICE
// [8][res_pitch].
void spectralPb(float* eigenvalues, float* eigenvectors, int xdim, int ydim, int nvec, float* results, int res_pitch);

#endif
