// ref_shim.cpp -- C entry points around the reference's host-only C++
// translation units (textons/filters.cpp, stencilMatrixMultiply/Stencil.cpp),
// compiled from /root/reference by `make -C oracle ref` into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle's table builders.
#include "Stencil.h"

void gaussian_2D(float* filter, double sigma_x, double sigma_y, double ori, unsigned int deriv,
                 bool hlbrt, unsigned long support_x, unsigned long support_y);
void createTextonFilters(float** filters, int* nFilterCoefficients, int** radii, float* scale,
                         int nscales, float elongation);

extern "C" {

int ref_texton_filters(float* out, int* radii_out) {
  float* coeffs = 0;
  int* radii = 0;
  int n = 0;
  float scales[2] = {2.0f, (float)(2.0 * 1.41421356237309504880)};
  createTextonFilters(&coeffs, &n, &radii, scales, 2, 3.0f);
  for (int i = 0; i < 9010; i++) out[i] = coeffs[i];
  for (int i = 0; i < 34; i++) radii_out[i] = radii[i];
  return n;
}

void ref_gaussian_2d(float* filter, double sx, double sy, double ori, unsigned deriv, int hil,
                     unsigned long supx, unsigned long supy) {
  gaussian_2D(filter, sx, sy, ori, deriv, hil != 0, supx, supy);
}

int ref_stencil(int radius, int width, int height, int* offsets, int* diag) {
  Stencil* s = new Stencil(radius, width, height, width * height);
  int area = (int)s->getStencilArea();
  int* all = new int[CONSTSPACE];
  s->copyOffsets(all);
  for (int i = 0; i < area; i++) offsets[i] = all[i];
  int d = s->getDiameter();
  s->copyDiagonalOffsets(all);
  for (int i = 0; i < d * d; i++) diag[i] = all[i];
  delete[] all;
  delete s;
  return area;
}

}  // extern "C"
