// skeleton.h -- thinned gPb and sigmoid normalisation (reference
// postprocess/skeleton.h:6-14).  Implemented by libdmsc over
// dmsc_postprocess() / dmsc_normalize_gpb_all() on the per-thread default
// context; errors throw std::runtime_error.
#ifndef DMSC_SKELETON_H
#define DMSC_SKELETON_H

// devGPb [width*height] (replaced by its sigmoid normalisation), devMPb the
// non-max-suppressed mPb [width*height], devGPb_thin [width*height] output;
// matrixPitchInFloats must equal width (the reference passes width).
void PostProcess(int width, int height, unsigned int matrixPitchInFloats, float* devGPb, float* devMPb,
                 float* devGPb_thin);
void NormalizeGpbAll(int p_nPixels, int p_nOrient, int matrixPitchInFloats, float* devGpball);

#endif
