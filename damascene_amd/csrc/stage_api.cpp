// stage_api.cpp -- the reference's stage functions (include/texton.h,
// kmeans.h, localcues.h, combine.h, nonmax.h, intervening.h, lanczos.h,
// stencilMVM.h, spectralPb.h, globalPb.h, convert.h, Stencil.h) with their
// names, signatures and ownership rules, implemented over the C ABI.
//
// Each host thread gets a default context on the current device's default
// stream, so calls order with the caller's own default-stream work, as the
// reference's synchronous calls did.  Outputs passed as `**` are allocated
// here with hipMalloc and belong to the caller; host eigenvalues use malloc.
// A failing call throws std::runtime_error with the library's message
// instead of exiting the process.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Stencil.h"
#include "combine.h"
#include "convert.h"
#include "dmsc.h"
#include "globalPb.h"
#include "intervening.h"
#include "kmeans.h"
#include "lanczos.h"
#include "localcues.h"
#include "nonmax.h"
#include "skeleton.h"
#include "spectralPb.h"
#include "stencilMVM.h"
#include "texton.h"

namespace {

void check(int rc, const char* what) {
  if (rc != DMSC_OK) throw std::runtime_error(std::string(what) + ": " + dmsc_last_error_string());
}

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct CtxHolder {
  dmsc_ctx* ctx = nullptr;
  int device = -1;
  ~CtxHolder() {
    if (ctx) dmsc_ctx_destroy(ctx);
  }
};

// Per-thread default context, re-created when the current device changes.
dmsc_ctx* default_ctx() {
  thread_local CtxHolder h;
  int dev = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (!h.ctx || h.device != dev) {
    if (h.ctx) dmsc_ctx_destroy(h.ctx);
    h.ctx = nullptr;
    check(dmsc_ctx_create(&h.ctx, dev, nullptr), "dmsc_ctx_create");
    h.device = dev;
  }
  return h.ctx;
}

template <typename T>
T* dev_alloc(size_t count, const char* what) {
  void* p = nullptr;
  check_hip(hipMalloc(&p, sizeof(T) * (count ? count : 1)), what);
  return (T*)p;
}

}  // namespace

// ------------------------------------------------------------------ Stencil
Stencil::Stencil(int radiusIn, int widthIn, int heightIn, int matrixPitchInFloatsIn)
    : radius(radiusIn), diameter(2 * radiusIn + 1), width(widthIn), height(heightIn),
      matrixPitchInFloats(matrixPitchInFloatsIn) {
  if (radius < 1 || radius > 10) throw std::runtime_error("Stencil: radius must be in [1, 10]");
  const int area = dmsc_stencil(radius, width, hostOffsets, diagonalOffsets);
  if (area <= 0) throw std::runtime_error(std::string("Stencil: ") + dmsc_last_error_string());
  stencilArea = (uint)area;
  for (int i = area; i < STENCILAREAMAX; i++) hostOffsets[i] = -2000000000;  // Stencil.cpp:29-31
  for (int i = diameter * diameter; i < CONSTSPACE; i++) diagonalOffsets[i] = -1;
}
uint Stencil::getStencilArea() { return stencilArea; }
// Index of matrix entry (row, col) in the DIA storage: pitch * diagonal +
// row, with the column differential split into (dy, dx) across row wrap as
// Stencil.cpp:41-59 does; -1 for an entry outside the stencil.
int Stencil::matrixIndex(int row, int col) {
  const int diff = col - row;
  int dx = diff % width, dy = diff / width;
  if (dx > radius) {
    dx -= width;
    dy++;
  }
  if (dx < -radius) {
    dx += width;
    dy--;
  }
  if (dy < -radius || dy > radius || dx < -radius || dx > radius) return -1;
  const int d = diagonalOffsets[(dy + radius) * diameter + dx + radius];
  return d < 0 ? -1 : matrixPitchInFloats * d + row;
}
void Stencil::copyOffsets(int* destOffsets) { std::memcpy(destOffsets, hostOffsets, sizeof(int) * STENCILAREAMAX); }
int Stencil::getMatrixPitch() { return matrixPitchInFloats * (int)sizeof(float); }
int Stencil::getMatrixPitchInFloats() { return matrixPitchInFloats; }
int Stencil::getHeight() { return height; }
int Stencil::getWidth() { return width; }
int Stencil::getRadius() { return radius; }
int Stencil::getDiameter() { return diameter; }
void Stencil::copyDiagonalOffsets(int* destOffsets) {
  std::memcpy(destOffsets, diagonalOffsets, sizeof(int) * diameter * diameter);
}

// Stencil.cpp:92-145 readStencilMatrix (.sma: n, nnz, then per row the
// count, the FP64 values and the int columns), each value rounded to float
// at its DIA slot.
float* Stencil::readStencilMatrix(char* filename) {
  std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(filename, "rb"), std::fclose);
  if (!fp) throw std::runtime_error(std::string("readStencilMatrix: cannot open ") + filename);
  const size_t total = (size_t)stencilArea * (size_t)matrixPitchInFloats;
  std::unique_ptr<float, void (*)(void*)> out((float*)std::calloc(total ? total : 1, sizeof(float)), std::free);
  if (!out) throw std::runtime_error("readStencilMatrix: out of host memory");
  int n = 0, nnz = 0;
  if (std::fread(&n, sizeof(int), 1, fp.get()) != 1 || std::fread(&nnz, sizeof(int), 1, fp.get()) != 1)
    throw std::runtime_error("readStencilMatrix: truncated header");
  if (n != width * height)
    throw std::runtime_error("readStencilMatrix: " + std::to_string(n) + " rows for a " + std::to_string(width) +
                             "x" + std::to_string(height) + " stencil");
  if (nnz < 0 || (long long)nnz >= (long long)stencilArea * n)
    throw std::runtime_error("readStencilMatrix: bad entry count " + std::to_string(nnz));
  std::vector<double> z(stencilArea);
  std::vector<int> x(stencilArea);
  for (int row = 0; row < n; row++) {
    int nz = 0;
    if (std::fread(&nz, sizeof(int), 1, fp.get()) != 1 || nz < 0 || nz > (int)stencilArea)
      throw std::runtime_error("readStencilMatrix: bad entry count in row " + std::to_string(row));
    if (std::fread(z.data(), sizeof(double), nz, fp.get()) != (size_t)nz ||
        std::fread(x.data(), sizeof(int), nz, fp.get()) != (size_t)nz)
      throw std::runtime_error("readStencilMatrix: truncated row " + std::to_string(row));
    for (int k = 0; k < nz; k++) {
      const int idx = (x[k] >= 0 && x[k] < n) ? matrixIndex(row, x[k]) : -1;
      if (idx < 0 || (size_t)idx >= total)
        throw std::runtime_error("readStencilMatrix: entry (" + std::to_string(row) + ", " + std::to_string(x[k]) +
                                 ") is outside the stencil");
      out.get()[idx] = (float)z[k];
    }
  }
  return out.release();
}

// ------------------------------------------------------------------ convert
void loadPPM_rgbU(char* filename, uint* p_width, uint* p_height, uint** p_devRgbU) {
  int w = 0, h = 0;
  check(dmsc_load_ppm(filename, &w, &h, nullptr), "loadPPM_rgbU");
  std::vector<uint32_t> img((size_t)w * h);
  check(dmsc_load_ppm(filename, &w, &h, img.data()), "loadPPM_rgbU");
  uint* d = dev_alloc<uint>(img.size(), "loadPPM_rgbU");
  check_hip(hipMemcpy(d, img.data(), sizeof(uint) * img.size(), hipMemcpyHostToDevice), "loadPPM_rgbU");
  *p_width = (uint)w;
  *p_height = (uint)h;
  *p_devRgbU = d;
}

void rgbUtoLab3F(uint width, uint height, float gamma, uint* devRgbU, float** p_devL, float** p_devA,
                 float** p_devB) {
  const size_t n = (size_t)width * height;
  *p_devL = dev_alloc<float>(n, "rgbUtoLab3F");
  *p_devA = dev_alloc<float>(n, "rgbUtoLab3F");
  *p_devB = dev_alloc<float>(n, "rgbUtoLab3F");
  check(dmsc_rgb_to_lab(default_ctx(), width, height, gamma, devRgbU, *p_devL, *p_devA, *p_devB), "rgbUtoLab3F");
}

void normalizeLab(uint width, uint height, float* devL, float* devA, float* devB) {
  check(dmsc_normalize_lab(default_ctx(), width, height, devL, devA, devB), "normalizeLab");
}

void rgbUtoGreyF(uint width, uint height, uint* devRgbU, float** p_devGrey) {
  *p_devGrey = dev_alloc<float>((size_t)width * height, "rgbUtoGreyF");
  check(dmsc_rgb_to_grey(default_ctx(), width, height, devRgbU, *p_devGrey), "rgbUtoGreyF");
}

void mirrorImage(uint width, uint height, uint border, float* devInput, float** p_devOutput) {
  *p_devOutput = dev_alloc<float>((size_t)(width + 2 * border) * (height + 2 * border), "mirrorImage");
  mirrorImage(width, height, border, devInput, *p_devOutput);
}
void mirrorImage(uint width, uint height, uint border, float* devInput, float* p_devOutput) {
  check(dmsc_mirror_image_f(default_ctx(), width, height, border, devInput, p_devOutput), "mirrorImage");
}
void mirrorImage(uint width, uint height, uint border, int* devInput, int** p_devOutput) {
  *p_devOutput = dev_alloc<int>((size_t)(width + 2 * border) * (height + 2 * border), "mirrorImage");
  mirrorImage(width, height, border, devInput, *p_devOutput);
}
void mirrorImage(uint width, uint height, uint border, int* devInput, int* p_devOutput) {
  check(dmsc_mirror_image_i(default_ctx(), width, height, border, devInput, p_devOutput), "mirrorImage");
}
// width/height are the bordered input's (convert.cu:214-224)
void unMirrorImage(uint width, uint height, uint border, float* devInput, float** p_devOutput) {
  *p_devOutput = dev_alloc<float>((size_t)(width - 2 * border) * (height - 2 * border), "unMirrorImage");
  unMirrorImage(width, height, border, devInput, *p_devOutput);
}
void unMirrorImage(uint width, uint height, uint border, float* devInput, float* p_devOutput) {
  check(dmsc_unmirror_image_f(default_ctx(), width, height, border, devInput, p_devOutput), "unMirrorImage");
}
void quantizeImage(uint width, uint height, uint nbins, float* devInput, int** p_devOutput) {
  *p_devOutput = dev_alloc<int>((size_t)width * height, "quantizeImage");
  quantizeImage(width, height, nbins, devInput, *p_devOutput);
}
void quantizeImage(uint width, uint height, uint nbins, float* devInput, int* devOutput) {
  check(dmsc_quantize_image(default_ctx(), width, height, nbins, devInput, devOutput), "quantizeImage");
}

// ------------------------------------------------------------------ textons
void findTextons(int width, int height, float* devImage, int** p_devTextons, int textonChoice) {
  *p_devTextons = dev_alloc<int>((size_t)width * height, "findTextons");
  check(dmsc_find_textons(default_ctx(), width, height, devImage, *p_devTextons, textonChoice, DMSC_PREC_FP32,
                          nullptr),
        "findTextons");
}

int kmeans(int textonChoice, int nPixels, int width, int height, int clusterCount, int filterCount,
           float* devResponses, int** p_devClusters, int maxIter, int convThresh) {
  (void)textonChoice;
  if (nPixels != width * height) throw std::runtime_error("kmeans: nPixels must equal width * height");
  *p_devClusters = dev_alloc<int>((size_t)nPixels, "kmeans");
  int changes = 0;
  check(dmsc_kmeans(default_ctx(), width, height, clusterCount, filterCount, devResponses, *p_devClusters, maxIter,
                    convThresh, DMSC_PREC_FP32, &changes, nullptr),
        "kmeans");
  return changes;
}

// ------------------------------------------------------------------ local cues, combine, nonmax
void localCues(int width, int height, float* devL, float* devA, float* devB, int* devTextons, float** devBg,
               float** devCga, float** devCgb, float** devTg, int* cuePitchInFloats, int textonChoice) {
  const int pitch = dmsc_pitch_in_floats(width * height);
  float** outs[4] = {devBg, devCga, devCgb, devTg};
  for (float** o : outs) *o = dev_alloc<float>((size_t)24 * pitch, "localCues");
  *cuePitchInFloats = pitch;
  check(dmsc_local_cues(default_ctx(), width, height, devL, devA, devB, devTextons, *devBg, *devCga, *devCgb, *devTg,
                        pitch, textonChoice),
        "localCues");
}

void combine(int width, int height, int cuePitchInFloats, float* devBg, float* devCga, float* devCgb, float* devTg,
             float** p_devMpb, float** p_devCombinedg, int textonChoice) {
  *p_devMpb = dev_alloc<float>((size_t)8 * cuePitchInFloats, "combine");
  *p_devCombinedg = dev_alloc<float>((size_t)8 * cuePitchInFloats, "combine");
  check(dmsc_combine(default_ctx(), width, height, cuePitchInFloats, devBg, devCga, devCgb, devTg, *p_devMpb,
                     *p_devCombinedg, textonChoice),
        "combine");
}

void nonMaxSuppression(int width, int height, float* devPB, int pitchFloats, float* devNMax) {
  check(dmsc_nonmax(default_ctx(), width, height, devPB, pitchFloats, devNMax), "nonMaxSuppression");
}

// ------------------------------------------------------------------ affinities, eigensolve
void intervene(Stencil& theStencil, float* devMPb, float** p_devMatrix, float sigma) {
  const int pitch = theStencil.getMatrixPitchInFloats();
  *p_devMatrix = dev_alloc<float>((size_t)theStencil.getStencilArea() * pitch, "intervene");
  check(dmsc_intervene(default_ctx(), theStencil.getWidth(), theStencil.getHeight(), devMPb, *p_devMatrix, pitch,
                       sigma),
        "intervene");
}

int findPitchInFloats(int width) { return dmsc_pitch_in_floats(width); }

float* convertMatrix(Stencil* theStencil, dim3 gridDim, dim3 blockDim, int nDimension, float* devMatrix) {
  (void)gridDim;
  (void)blockDim;
  (void)nDimension;
  const size_t n = (size_t)theStencil->getWidth() * theStencil->getHeight();
  float* rs = dev_alloc<float>(n, "convertMatrix");
  check(dmsc_convert_matrix(default_ctx(), theStencil->getWidth(), theStencil->getHeight(), devMatrix,
                            theStencil->getMatrixPitchInFloats(), nullptr, rs),
        "convertMatrix");
  return rs;
}

void stencilMatrixVector(Stencil* theStencil, float* devMatrix, float* devX, float* devY) {
  check(dmsc_stencil_spmv(default_ctx(), theStencil->getWidth(), theStencil->getHeight(), devMatrix,
                          theStencil->getMatrixPitchInFloats(), devX, devY),
        "stencilMatrixVector");
}

void chooseLargestGPU(bool verbose) {
  int count = 0;
  check_hip(hipGetDeviceCount(&count), "chooseLargestGPU");
  int best = 0;
  size_t best_mem = 0;
  for (int d = 0; d < count; d++) {
    hipDeviceProp_t p;
    check_hip(hipGetDeviceProperties(&p, d), "chooseLargestGPU");
    if (p.totalGlobalMem > best_mem) {
      best_mem = p.totalGlobalMem;
      best = d;
    }
  }
  check_hip(hipSetDevice(best), "chooseLargestGPU");
  if (verbose) std::printf("Using device %d (%zu MiB)\n", best, best_mem >> 20);
}

void generalizedEigensolve(Stencil& myStencil, float* devMatrix, int matrixPitchInFloats, int nEigNum,
                           float** p_eigenvalues, float** p_eigenvectors, float tolerance) {
  const size_t n = (size_t)myStencil.getWidth() * myStencil.getHeight();
  float* vals = (float*)std::malloc(sizeof(float) * nEigNum);
  if (!vals) throw std::bad_alloc();
  float* vecs = dev_alloc<float>(n * nEigNum, "generalizedEigensolve");
  int it = 0;
  const int rc = dmsc_generalized_eigensolve(default_ctx(), myStencil.getWidth(), myStencil.getHeight(), devMatrix,
                                             matrixPitchInFloats, nEigNum, tolerance, DMSC_PREC_FP32, vals, vecs,
                                             &it);
  if (rc != DMSC_OK) {
    std::free(vals);
    (void)hipFree(vecs);
    check(rc, "generalizedEigensolve");
  }
  *p_eigenvalues = vals;
  *p_eigenvectors = vecs;
}

// ------------------------------------------------------------------ spectral Pb, gPb
void spectralPb(float* eigenvalues, float* eigenvectors, int xdim, int ydim, int nvec, float* results,
                int res_pitch) {
  check(dmsc_spectral_pb(default_ctx(), eigenvalues, eigenvectors, xdim, ydim, nvec, results, res_pitch),
        "spectralPb");
}

void StartCalcGPb(int p_nPixels, int p_nMatrixPitch, int p_nOrient, float* devCombinedGradient, float* devspb,
                  float* devmpb, float* devGpball, float* devResult) {
  check(dmsc_calc_gpb(default_ctx(), p_nPixels, p_nMatrixPitch, p_nOrient, devCombinedGradient, devspb, devmpb,
                      devGpball, devResult),
        "StartCalcGPb");
}

// ------------------------------------------------------------------ post-processing
void PostProcess(int width, int height, unsigned int matrixPitchInFloats, float* devGPb, float* devMPb,
                 float* devGPb_thin) {
  if ((int)matrixPitchInFloats != width) throw std::runtime_error("PostProcess: pitch must equal width");
  check(dmsc_postprocess(default_ctx(), width, height, devGPb, devMPb, devGPb_thin, nullptr), "PostProcess");
}

void NormalizeGpbAll(int p_nPixels, int p_nOrient, int matrixPitchInFloats, float* devGpball) {
  check(dmsc_normalize_gpb_all(default_ctx(), p_nPixels, p_nOrient, matrixPitchInFloats, devGpball),
        "NormalizeGpbAll");
}
