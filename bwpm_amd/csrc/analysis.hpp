// Device analysis reductions over field planes (flux, Fresnel reference reflectance, axial
// focus scan, Strehl profile); see analysis.hip.
#pragma once

#include <memory>
#include <vector>

#include "../../include/bwpm_hip.h"

namespace bwpm {

// One grid (ny, nx, dy, dx) at vacuum wavenumber k0 on one device: FFT plans, k tables, work
// buffers and cached mode lists. Not safe for concurrent calls on one object.
class Analysis {
 public:
  Analysis(int device, long long ny, long long nx, double dy, double dx, double k0);
  ~Analysis();
  Analysis(const Analysis&) = delete;
  Analysis& operator=(const Analysis&) = delete;

  // planes: `count` planes (ny, nx, 2); where: BWPM_PLANE_* (host complex128, device complex128,
  // device complex64).
  std::vector<double> plane_power(const void* planes, int where, long long count, double index);
  double reference_reflectance(const void* plane, int where, double n1, double n2);
  void axial_profile(const void* plane, int where, double index, const double* offsets, long long noff, double na,
                     double* intensity);
  void strehl_profile(const void* spectrum, int where, double k, double na, const double* offsets, long long noff,
                      double* values);
  void strehl_modes(const double* ky, const void* spectrum, long long m, double k, const double* offsets,
                    long long noff, double* values);

 private:
  struct Impl;
  std::unique_ptr<Impl> p_;
};

}  // namespace bwpm
