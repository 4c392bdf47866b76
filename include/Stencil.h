// Stencil.h -- the radius-r disc stencil of the DIA affinity matrix.
// Same class interface and object layout as the reference's
// include/Stencil.h:12-36 (constructor, area, offsets, diagonal map, pitch
// accessors, the .sma reader, and the two 16384-entry tables), implemented
// by libdmsc over dmsc_stencil().
#ifndef DMSC_STENCIL_H
#define DMSC_STENCIL_H

#define CONSTSPACE 16384
#define STENCILAREAMAX CONSTSPACE

typedef unsigned int uint;

class Stencil {
 public:
  Stencil(int radiusIn, int widthIn, int heightIn, int matrixPitchInFloatsIn);
  uint getStencilArea();
  // index of matrix entry (row, col) in the DIA storage (pitch * diagonal +
  // row), -1 outside the stencil (Stencil.cpp:41-59)
  int matrixIndex(int row, int col);
  // delta-coded linear offsets, one per diagonal, then -2e9 padding up to
  // STENCILAREAMAX entries, all of which are copied (Stencil.cpp:16, 29-31, 62)
  void copyOffsets(int* destOffsets);
  int getMatrixPitch();
  int getMatrixPitchInFloats();
  int getHeight();
  int getWidth();
  int getRadius();
  int getDiameter();
  // Stencil.cpp:92-145: a matrix in the .sma format (int n = width * height,
  // int nnz, then per row: int count, double values[count], int
  // columns[count]) as DIA floats [stencilArea][matrixPitchInFloats], zero
  // where the file has no entry; malloc'ed, the caller frees it.  Throws
  // std::runtime_error on a file that cannot be read, a row count other than
  // width * height, or an entry outside the stencil.
  float* readStencilMatrix(char* filename);
  // (dy + r) * diameter + (dx + r) -> diagonal index, -1 outside the disc
  void copyDiagonalOffsets(int* destOffsets);

 private:
  uint stencilArea;
  int hostOffsets[STENCILAREAMAX];
  int diagonalOffsets[CONSTSPACE];
  int radius;
  int diameter;
  int width;
  int height;
  int matrixPitchInFloats;
};

#endif
