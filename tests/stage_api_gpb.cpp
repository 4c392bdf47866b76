// stage_api_gpb.cpp -- the reference's gPb call sequence (damascene.cu:269-470)
// written against the reference-named C++ stage API of libdmsc (include/*.h),
// as a caller of the reference would write it: PPM in, results out.
//
//   stage_api_gpb image.ppm out_prefix
//
// writes out_prefix.{textons (int32), nmax, gpb (float32), gpball (float32,
// 8 planes of w*h), eigvals (float32, 9)} as raw host arrays, row-major.
// Built by damascene_amd/build.py; run by tests/test_gpu_stage_api.py, which
// compares every output bit for bit with dmsc_gpb on the same image.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "Stencil.h"
#include "combine.h"
#include "convert.h"
#include "globalPb.h"
#include "intervening.h"
#include "lanczos.h"
#include "localcues.h"
#include "nonmax.h"
#include "spectralPb.h"
#include "stencilMVM.h"
#include "texton.h"

namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
void save(const std::string& path, const T* dev, size_t count, size_t pitch = 0, int planes = 1) {
  std::vector<T> host(count * planes);
  for (int p = 0; p < planes; p++)
    hip_ok(hipMemcpy(host.data() + (size_t)p * count, dev + (size_t)p * pitch, sizeof(T) * count,
                     hipMemcpyDeviceToHost),
           "copy back");
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  const size_t put = std::fwrite(host.data(), sizeof(T), host.size(), f);
  std::fclose(f);
  if (put != host.size()) throw std::runtime_error("short write to " + path);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s image.ppm out_prefix\n", argv[0]);
    return 2;
  }
  const std::string out = argv[2];
  try {
    const int nEigNum = 9;
    const float fTolerance = 1e-3f;
    const int textonChoice = 1;  // TEXTON32
    uint width, height;
    uint* devRgbU;
    loadPPM_rgbU(argv[1], &width, &height, &devRgbU);
    const int w = (int)width, h = (int)height, n = w * h;

    float* devGreyscale;
    rgbUtoGreyF(width, height, devRgbU, &devGreyscale);
    int* devTextons;
    findTextons(w, h, devGreyscale, &devTextons, textonChoice);

    float *devL, *devA, *devB;
    rgbUtoLab3F(width, height, 2.5f, devRgbU, &devL, &devA, &devB);
    normalizeLab(width, height, devL, devA, devB);

    float *devBg, *devCga, *devCgb, *devTg;
    int cuePitchInFloats;
    localCues(w, h, devL, devA, devB, devTextons, &devBg, &devCga, &devCgb, &devTg, &cuePitchInFloats,
              textonChoice);

    float *devMPb, *devCombinedGradient;
    combine(w, h, cuePitchInFloats, devBg, devCga, devCgb, devTg, &devMPb, &devCombinedGradient, textonChoice);

    float* devMPbNMax;
    hip_ok(hipMalloc((void**)&devMPbNMax, sizeof(float) * n), "hipMalloc");
    nonMaxSuppression(w, h, devMPb, cuePitchInFloats, devMPbNMax);

    Stencil theStencil(5, w, h, cuePitchInFloats);
    float* devMatrix;
    intervene(theStencil, devMPbNMax, &devMatrix);

    float* eigenvalues;
    float* devEigenvectors;
    generalizedEigensolve(theStencil, devMatrix, cuePitchInFloats, nEigNum, &eigenvalues, &devEigenvectors,
                          fTolerance);

    float* devSpb;
    hip_ok(hipMalloc((void**)&devSpb, sizeof(float) * cuePitchInFloats * 8), "hipMalloc");
    spectralPb(eigenvalues, devEigenvectors, w, h, nEigNum, devSpb, cuePitchInFloats);

    float *devGpbAll, *devGpb;
    hip_ok(hipMalloc((void**)&devGpbAll, sizeof(float) * cuePitchInFloats * 8), "hipMalloc");
    hip_ok(hipMalloc((void**)&devGpb, sizeof(float) * n), "hipMalloc");
    StartCalcGPb(n, cuePitchInFloats, 8, devCombinedGradient, devSpb, devMPbNMax, devGpbAll, devGpb);
    hip_ok(hipDeviceSynchronize(), "synchronise");

    save(out + ".textons", devTextons, n);
    save(out + ".nmax", devMPbNMax, n);
    save(out + ".gpb", devGpb, n);
    save(out + ".gpball", devGpbAll, n, cuePitchInFloats, 8);
    {
      FILE* f = std::fopen((out + ".eigvals").c_str(), "wb");
      if (!f || std::fwrite(eigenvalues, sizeof(float), nEigNum, f) != (size_t)nEigNum)
        throw std::runtime_error("cannot write eigenvalues");
      std::fclose(f);
    }
    std::printf("stage API gPb %dx%d done, cue pitch %d\n", w, h, cuePitchInFloats);

    free(eigenvalues);
    for (void* p : {(void*)devRgbU, (void*)devGreyscale, (void*)devTextons, (void*)devL, (void*)devA, (void*)devB,
                    (void*)devBg, (void*)devCga, (void*)devCgb, (void*)devTg, (void*)devMPb,
                    (void*)devCombinedGradient, (void*)devMPbNMax, (void*)devMatrix, (void*)devEigenvectors,
                    (void*)devSpb, (void*)devGpbAll, (void*)devGpb})
      hip_ok(hipFree(p), "hipFree");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "stage_api_gpb: %s\n", e.what());
    return 1;
  }
  return 0;
}
