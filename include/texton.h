// texton.h -- texton labels (reference include/texton.h:16).  Implemented by
// libdmsc over dmsc_find_textons() on a per-thread default context
// (device 0, the device's default stream).  Errors throw std::runtime_error.
#ifndef DMSC_TEXTON_H
#define DMSC_TEXTON_H

// 34-filter bank on a greyscale DEVICE image, then k-means (k = 32 for
// textonChoice 1); allocates *p_devTextons (int[width*height], hipMalloc;
// the caller frees it with hipFree).
void findTextons(int width, int height, float* devImage, int** p_devTextons, int textonChoice);

#endif
