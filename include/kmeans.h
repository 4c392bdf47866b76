// kmeans.h -- k-means over filter responses (reference include/kmeans.h:18).
#ifndef DMSC_KMEANS_H
#define DMSC_KMEANS_H

// devResponses: DEVICE [filterCount][nPixels]; allocates *p_devClusters
// (int[nPixels]); returns the label changes of the last iteration.
int kmeans(int textonChoice, int nPixels, int width, int height, int clusterCount, int filterCount,
           float* devResponses, int** p_devClusters, int maxIter = 15, int convThresh = 0);

#endif
