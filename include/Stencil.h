// Stencil.h -- the radius-r disc stencil of the DIA affinity matrix.
// Same class interface as the reference's stencilMatrixMultiply/Stencil.h:12-36
// (constructor, area, offsets, diagonal map, pitch accessors), implemented by
// libdmsc over dmsc_stencil().
#ifndef DMSC_STENCIL_H
#define DMSC_STENCIL_H

typedef unsigned int uint;

class Stencil {
 public:
  Stencil(int radiusIn, int widthIn, int heightIn, int matrixPitchInFloatsIn);
  uint getStencilArea();
  // index of matrix entry (row, col) in the DIA storage (pitch * diagonal +
  // row), -1 outside the stencil (Stencil.cpp:41-59)
  int matrixIndex(int row, int col);
  // delta-coded linear offsets, one per diagonal (Stencil.cpp:16)
  void copyOffsets(int* destOffsets);
  int getMatrixPitch();
  int getMatrixPitchInFloats();
  int getHeight();
  int getWidth();
  int getRadius();
  int getDiameter();
  // (dy + r) * diameter + (dx + r) -> diagonal index, -1 outside the disc
  void copyDiagonalOffsets(int* destOffsets);

 private:
  uint stencilArea;
  int hostOffsets[441];      // radius <= 10
  int diagonalOffsets[441];
  int radius;
  int diameter;
  int width;
  int height;
  int matrixPitchInFloats;
};

#endif
