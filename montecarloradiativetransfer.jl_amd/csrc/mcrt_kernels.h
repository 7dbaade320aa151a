// mcrt_kernels.h — parameter blocks shared by the host API and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mcrt.h"
#include "mcrt_device.hpp"

namespace mcrt {

constexpr int kMaxLayers = 16;  // plane layers handled on the device
constexpr int kMaxSca = 16;     // distinct scatterers per solve

// Geometry in FP64 (plane.jl:6-38, sphere.jl:6-21), host-validated.
struct GeoDev {
  int kind;  // 0 plane, 1 sphere
  int nl;
  double tau[kMaxLayers];
  double omega[kMaxLayers];
  int isinf[kMaxLayers];
  int sca[kMaxLayers];  // scatterer slot of each layer
  double kl1;           // mean_free_path(g, 1): the CB phase scale
};

// Bin/quadrature tables (config.jl:77-123) in FP64, device-resident.
struct BinsDev {
  int Nt, Np, Ntb, Npb;
  const double* ct;
  const double* st;
  const double* inorm;
  const double* cp;
  const double* sp;
  const double* ctb;
  const double* stb;
  const double* cpb;
  const double* spb;
};

// Counters, in order: events, rt_evals, cb_evals, rejections, nonfinite, rays,
// single-scattering CB evals.
enum { kCntEvents = 0, kCntRt, kCntCb, kCntRej, kCntNonfinite, kCntRays, kCntCb1, kCntN };

// The drop-in FP64 kernel (structure of src/gpu.jl:177-296).
struct PlainParams {
  int64_t N;
  int64_t maxscat;
  int64_t threads;  // size of the path-history columns
  double min_int;
  double tau0;
  int compute_cb;
  int nsca;
  GeoDev geo;
  BinsDev bins;
  ScaView<double> sca[kMaxSca];
  double I0[4];
  double k0[3];
  double f0[4];  // (cosθ₀, sinθ₀, cosϕ₀, sinϕ₀) of incidence_angle(g)
  uint64_t seed;
  uint64_t ray0;  // stream number of the first ray
  uint32_t instance;
  double* kpath;      // [threads][maxscat][3]  (column per thread, as CuMatrix(maxscat, threads))
  double* xpath;
  int8_t* idxpath;    // [threads][maxscat]
  double* A;          // [5]
  double* Irt;        // [4*Nt*Np]
  double* Icb;        // [4*Ntb*Npb]
  unsigned long long* counters;  // [kCntN]
};

// Host helpers (mcrt_api.cpp): the incident frame and k₀ of simulate
// (MonteCarloRadiativeTransfer.jl:27-34), and the device geometry block.
void incident(const struct mcrt_geometry* g, double k0[3], double f0[4]);
GeoDev make_geo(const struct mcrt_geometry* g, int* layer_slot);

// p_dev: the parameter block in device memory.
hipError_t launch_plain(const PlainParams* p_dev, int blocks, int threads, hipStream_t stream);
// mcrt_plan_eval for the plain form: din/dout are device buffers of n records
// (16 doubles in, 4 out).
hipError_t launch_plain_eval(const PlainParams* p_dev, int what, int64_t n, const double* din,
                             double* dout, hipStream_t stream);

}  // namespace mcrt
