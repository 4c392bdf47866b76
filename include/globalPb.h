// globalPb.h -- gPb from the combined gradient and spectral Pb (reference
// include/globalPb.h:24).
#ifndef DMSC_GLOBALPB_H
#define DMSC_GLOBALPB_H

void StartCalcGPb(int p_nPixels, int p_nMatrixPitch, int p_nOrient, float* devCombinedGradient, float* devspb,
                  float* devmpb, float* devGpball, float* devResult);

#endif
