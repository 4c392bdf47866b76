// tables.hpp -- host-only declarations (no HIP headers): error helpers,
// pipeline constants and the table builders.  tables.cpp and generator.cpp
// include only this header so they can be compiled by the system C++
// compiler with the same floating-point settings as the CPU oracle.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "dmsc.h"

namespace dmsc {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);

#define DMSC_TRY(call)            \
  do {                            \
    int rc_ = (call);             \
    if (rc_ != DMSC_OK) return rc_; \
  } while (0)

#define DMSC_ARG(cond, ...)                                        \
  do {                                                             \
    if (!(cond)) return ::dmsc::fail(DMSC_ERR_ARG, __VA_ARGS__);   \
  } while (0)

// ---------------------------------------------------------------- sizes
constexpr int kBorder = 30;       // local-cue mirror border (localcues.cu:261)
constexpr int kOrient = 8;
constexpr int kScales = 3;
constexpr int kStencilRadius = 5; // 81-point disc (damascene.cu:414)
constexpr int kStencilArea = 81;
constexpr int kNFilters = 34;     // texton bank (17 per scale x 2)
constexpr int kTextons = 32;
constexpr int kMaxIter = 6000;    // Lanczos cap (lanczos.cu:4)
constexpr int kCheckIter = 500;   // Lanczos check period (lanczos.cu:5)
constexpr int kMinSide = 32;      // mirror border 30 needs >= 30 px per side

// Eigenvalues (unsorted) and eigenvectors (column j of row-major Z) of a
// symmetric n x n FP64 matrix: Householder tridiagonalisation + implicit
// symmetric QR (symeig.cpp).  False if the iteration did not converge.
bool sym_eigen(const double* A, int n, double* ev, double* Z);
// The m x m generalized symmetric eigenproblem H s = theta G s of the block
// eigensolver's Rayleigh-Ritz step (symeig.cpp); false if G is not
// positive definite or the QR iteration does not converge.
bool rayleigh_ritz(const double* G, const double* H, int m, std::vector<double>& theta, std::vector<double>& S);
inline int pitch_floats(long n) { return (int)(((n + 63) / 64) * 64); }
inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------- tables
struct TextonBank {
  std::vector<float> coeffs;   // 9010, [filter][dx][dy]
  std::vector<int> radii;      // 34
  int nbound = 0;              // coefficients visible to the convolution
};
const TextonBank& texton_bank();
void gaussian_2d(float* filter, double sx, double sy, double ori, unsigned deriv, bool hilbert,
                 unsigned long supx, unsigned long supy);
void savgol_filter(float* filt, int d, float ra, float rb, float theta);
void parabola_filters(int radius, float* out);          // 8*(2r+1)^2
void spb_filters(float* out);                            // 8*49
int rotation_tables(int width, int height, float thetaPi, int* nw, int* nh, int* bX, int* bY,
                    int* gX, int* gY);
void halfdisc_offsets(int w, int h, float thetaPi, int radius, const int* bX, const int* bY,
                      const int* gX, const int* gY, int rw, int* rect);
int blur_kernel(float sigma, float* k, int* radius);
int stencil_geometry(int radius, int width, int* offsets, int* diagMap);
void gamma_lut(float gamma, float* lut);

}  // namespace dmsc
