// The rdf table maker on the GPU: many sweeps (rdfmaketable columns,
// rxnparam.cpp:477-564) side by side, one workgroup per sweep.  Each launch
// advances every sweep by one reduced step to its steady flux; the rdf stays
// in device memory between launches (each steady state starts from the
// previous one, as on the host).  The arithmetic per radial point follows the
// host restatement (rdf.cpp) term for term, in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "smd_setup.h"

namespace {

constexpr int kRdfMax = 1024;  // radial points per sweep
constexpr int kRdfThreads = 256;
constexpr double kPiD = 3.14159265358979323846;
constexpr double kSqrt2PiD = 2.50662827462;

__device__ __forceinline__ double shell_d(double r0, double r1, double f0, double f1) {
  return kPiD * (f1 - f0) * (r1 + r0) * (r1 * r1 + r0 * r0) +
         4.0 * kPiD / 3.0 * (r1 * f0 - r0 * f1) * (r1 * r1 + r1 * r0 + r0 * r0);
}

__device__ double erfc_dev(double x) {
  // Numerical Recipes erfcc, as the host (rxnparam.cpp:316-322)
  const double z = fabs(x);
  const double t = 1.0 / (1.0 + 0.5 * z);
  const double ans =
      t * exp(-z * z - 1.26551223 +
              t * (1.00002368 +
                   t * (0.37409196 +
                        t * (0.09678418 +
                             t * (-0.18628806 +
                                  t * (0.27886807 +
                                       t * (-1.13520398 + t * (1.48851587 + t * (-0.82215223 + t * 0.17087277)))))))));
  return x >= 0.0 ? ans : 2.0 - ans;
}

struct RdfSweep {
  double* r;     // [n]
  double* a;     // [n] rdf after absorption (carried between steps)
  double* flux;  // [31]
  double b;      // < 0 irreversible
  int live;      // 0 once a steady state failed
};

// d = diffuse(a) for this thread's points; a2 is the fitted tail amplitude.
__device__ void diffuse_dev(const double* r, const double* a, double* d, int n, double step, double a2) {
  const double norm = 2.0 * kPiD * kSqrt2PiD;
  for (int i = threadIdx.x; i < n; i += kRdfThreads) {
    if (i == 0 && r[0] == 0) {
      double r0 = 0, r1 = 0, f0 = 0, f1 = 0, g = 0, acc = 0;
      for (int j = 1; j < n; ++j) {
        r0 = r1;
        f0 = f1;
        r1 = r[j] / step;
        g = exp(-r1 * r1 / 2.0) / norm;
        f1 = (a[j] - a[0]) * g;
        acc += shell_d(r0, r1, f0, f1);
      }
      acc += (1.0 - a[0]) * (4.0 * kPiD * r1 * g + erfc_dev(r1 / sqrt(2.0)));
      d[0] = acc + a[0];
      continue;
    }
    const double x = r[i] / step;
    double r0 = 0, r1 = 0, f0 = 0, g = exp(-x * x / 2.0) / norm;
    double f1 = (a[0] - a[i]) * g, acc = 0;
    for (int j = r[0] == 0 ? 1 : 0; j < n; ++j) {
      r0 = r1;
      f0 = f1;
      r1 = r[j] / step;
      g = 1.0 / x / r1 * (exp(-(x - r1) * (x - r1) / 2.0) - exp(-(x + r1) * (x + r1) / 2.0)) / (2.0 * norm);
      f1 = (a[j] - a[i]) * g;
      acc += shell_d(r0, r1, f0, f1);
    }
    const double ed = erfc_dev((r1 - x) / sqrt(2.0));
    const double es = erfc_dev((r1 + x) / sqrt(2.0));
    acc += (1.0 - a[i]) * (4.0 * kPiD * r1 * g + 0.5 * (ed + es)) + a2 / 2.0 / x * (ed - es);
    d[i] = acc + a[i];
  }
}

// Thread 0: tail fit of the outer 10% (same order as the host).
__device__ double tail_fit(const double* r, const double* a, int n, double step) {
  double alpha = 0, beta = 0;
  for (int i = (int)(0.9 * n); i < n; ++i) {
    alpha += 1.0 / (r[i] * r[i]);
    beta += (a[i] - 1.0) / r[i];
  }
  return beta / alpha / step;
}

// Thread 0: mass inside r = 1, removed; returns it.
__device__ double absorb_dev(const double* r, double* f, int n) {
  double r0 = 0, r1 = 0, f0 = 0, f1 = f[0], mass = 0;
  int j = r[0] == 0 ? 1 : 0;
  for (; r1 < 1 && j < n; ++j) {
    r0 = r1;
    f0 = f1;
    r1 = r[j];
    f1 = f[j];
    mass += shell_d(r0, r1, f0, f1);
  }
  mass -= shell_d(r0, r1, 0.0, f1);
  for (int k = j - 2; k >= 0; --k) f[k] = 0;
  return mass;
}

__device__ void reverse_dev(const double* r, double* f, int n, double step, double b, double flux) {
  const double k = 1.0 / (4.0 * kPiD * kSqrt2PiD * step * step * step);
  const double bs = b / step;
  for (int i = threadIdx.x; i < n; i += kRdfThreads) {
    const double x = r[i] / step;
    if (b == 0) {
      f[i] += flux * k * 2.0 * exp(-x * x / 2.0);
    } else if (i == 0 && r[0] == 0) {
      f[0] += flux * k * 2.0 * exp(-bs * bs / 2.0);
    } else {
      f[i] += flux * k / x / bs * (exp(-(x - bs) * (x - bs) / 2.0) - exp(-(x + bs) * (x + bs) / 2.0));
    }
  }
}

// Steady state at reduced step steps[blockIdx.x] for every sweep; result in
// flux[k].  Iterations are bounded by maxit (the host's limit).
__global__ __launch_bounds__(kRdfThreads) void k_rdf_step(RdfSweep* sweeps, const double* steps, int n, double eps,
                                                          int k) {
  __shared__ double s_r[kRdfMax], s_a[kRdfMax], s_d[kRdfMax];
  __shared__ double s_a2, s_flux, s_prev;
  __shared__ int s_go;
  RdfSweep& sw = sweeps[blockIdx.x];
  if (!sw.live) {
    if (threadIdx.x == 0) sw.flux[k] = -1.0;
    return;
  }
  const double step = steps[blockIdx.x];
  for (int i = threadIdx.x; i < n; i += kRdfThreads) {
    s_r[i] = sw.r[i];
    s_a[i] = sw.a[i];
  }
  __syncthreads();
  const int maxit = 100000;
  const double maxflux = 1e7;
  if (threadIdx.x == 0) s_a2 = tail_fit(s_r, s_a, n, step);
  __syncthreads();
  diffuse_dev(s_r, s_a, s_d, n, step, s_a2);
  __syncthreads();
  if (threadIdx.x == 0) {
    s_flux = absorb_dev(s_r, s_d, n);
    s_prev = s_flux;
    s_go = 1;
  }
  __syncthreads();
  int it = 0;
  for (; it < maxit; ++it) {
    if (threadIdx.x == 0) {
      const double flux = s_flux, prev = s_prev;
      s_go = it < 30 || (flux < maxflux && fabs((flux - prev) / (prev + 1e-20)) > eps);
      s_prev = flux;
      if (s_go) s_a2 = tail_fit(s_r, s_a, n, step);
    }
    __syncthreads();
    if (!s_go) break;
    diffuse_dev(s_r, s_a, s_d, n, step, s_a2);
    __syncthreads();
    if (sw.b >= 0) reverse_dev(s_r, s_d, n, step, sw.b, s_flux);
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kRdfThreads) s_a[i] = s_d[i];
    __syncthreads();
    if (threadIdx.x == 0) s_flux = absorb_dev(s_r, s_a, n);
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += kRdfThreads) sw.a[i] = s_a[i];
  if (threadIdx.x == 0) {
    const bool bad = it >= maxit || s_flux >= maxflux;
    sw.flux[k] = bad ? -1.0 : s_flux;
    if (bad) sw.live = 0;
  }
}

}  // namespace

// Batched table maker: sweep q has mode modes[q] ('i', 'r' or 'b'), unbinding
// radius bs[q] (ignored for 'i') and direction downs[q]; flux[q * 31 + k]
// belongs to reduced step exp(-3) exp(0.2)^k.
extern "C" smd_status smd_rdf_sweeps_gpu(int device, int n, double eps, int nSweeps, const char* modes,
                                         const double* bs, const int* downs, double* flux) {
  if (n < 10 || n > kRdfMax || !(eps > 0) || nSweeps <= 0 || !modes || !bs || !downs || !flux) {
    smd::set_error("invalid rdf sweep arguments (n must be in [10, 1024])");
    return SMD_ERR_INVALID;
  }
  for (int q = 0; q < nSweeps; ++q) {
    const char m = modes[q];
    if ((m != 'i' && m != 'r' && m != 'b') || (m == 'r' && !(bs[q] >= 1.0)) ||
        (m == 'b' && !(bs[q] >= 0.0 && bs[q] <= 1.0))) {
      smd::set_error("invalid rdf sweep mode or unbinding radius");
      return SMD_ERR_INVALID;
    }
  }
  // host-side grids, initial rdfs and step sequences (as the table maker)
  std::vector<double> r((size_t)nSweeps * n), a((size_t)nSweeps * n), stepSeq((size_t)nSweeps * 31);
  std::vector<double> tmp(n + 2);
  std::vector<RdfSweep> sw(nSweeps);
  for (int q = 0; q < nSweeps; ++q) {
    const char m = modes[q];
    const double b = m == 'i' ? -1.0 : bs[q];
    const double rmax = m == 'i' ? 10.0 : m == 'r' ? b + 3.0 : 5.0;
    smd::rdf_grid(rmax, n, tmp.data());
    double* rq = &r[(size_t)q * n];
    double* aq = &a[(size_t)q * n];
    for (int i = 0; i < n; ++i) rq[i] = tmp[i];
    int i = 0;
    for (; i < n && rq[i] < 1; ++i) aq[i] = 0;
    if (!downs[q] && m == 'i')
      for (; i < n; ++i) aq[i] = 1.0 - 1.0 / rq[i];
    else if (!downs[q] && m == 'r')
      for (; i < n && rq[i] < b; ++i) aq[i] = 1.0 - (b - rq[i]) / rq[i] / (b - 1.0);
    for (; i < n; ++i) aq[i] = 1.0;
    const double inc = downs[q] ? 1.0 / std::exp(0.2) : std::exp(0.2);
    double s = downs[q] ? std::exp(3.0) : std::exp(-3.0);
    for (int k = 0; k < 31; ++k, s *= inc) stepSeq[(size_t)q * 31 + k] = s;
    sw[q].b = b;
    sw[q].live = 1;
  }
  smd_status rc = SMD_OK;
  double *dR = nullptr, *dA = nullptr, *dFlux = nullptr, *dSteps = nullptr;
  RdfSweep* dSw = nullptr;
  std::vector<double> col(nSweeps), out((size_t)nSweeps * 31);
  auto fail = [&](hipError_t e, const char* what) {
    smd::set_error(std::string("HIP error '") + hipGetErrorString(e) + "' in " + what);
    rc = SMD_ERR_HIP;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) fail(e, "hipSetDevice");
  if (rc == SMD_OK && (e = hipMalloc((void**)&dR, r.size() * 8)) != hipSuccess) fail(e, "hipMalloc");
  if (rc == SMD_OK && (e = hipMalloc((void**)&dA, a.size() * 8)) != hipSuccess) fail(e, "hipMalloc");
  if (rc == SMD_OK && (e = hipMalloc((void**)&dFlux, out.size() * 8)) != hipSuccess) fail(e, "hipMalloc");
  if (rc == SMD_OK && (e = hipMalloc((void**)&dSteps, (size_t)nSweeps * 8)) != hipSuccess) fail(e, "hipMalloc");
  if (rc == SMD_OK && (e = hipMalloc((void**)&dSw, (size_t)nSweeps * sizeof(RdfSweep))) != hipSuccess)
    fail(e, "hipMalloc");
  if (rc == SMD_OK) {
    for (int q = 0; q < nSweeps; ++q) {
      sw[q].r = dR + (size_t)q * n;
      sw[q].a = dA + (size_t)q * n;
      sw[q].flux = dFlux + (size_t)q * 31;
    }
    if ((e = hipMemcpy(dR, r.data(), r.size() * 8, hipMemcpyHostToDevice)) != hipSuccess) fail(e, "hipMemcpy");
  }
  if (rc == SMD_OK && (e = hipMemcpy(dA, a.data(), a.size() * 8, hipMemcpyHostToDevice)) != hipSuccess)
    fail(e, "hipMemcpy");
  if (rc == SMD_OK &&
      (e = hipMemcpy(dSw, sw.data(), (size_t)nSweeps * sizeof(RdfSweep), hipMemcpyHostToDevice)) != hipSuccess)
    fail(e, "hipMemcpy");
  for (int k = 0; rc == SMD_OK && k < 31; ++k) {
    for (int q = 0; q < nSweeps; ++q) col[q] = stepSeq[(size_t)q * 31 + k];
    if ((e = hipMemcpy(dSteps, col.data(), (size_t)nSweeps * 8, hipMemcpyHostToDevice)) != hipSuccess) {
      fail(e, "hipMemcpy");
      break;
    }
    hipLaunchKernelGGL(k_rdf_step, dim3(nSweeps), dim3(kRdfThreads), 0, nullptr, dSw, dSteps, n, eps, k);
    if ((e = hipGetLastError()) != hipSuccess) fail(e, "k_rdf_step");
    else if ((e = hipDeviceSynchronize()) != hipSuccess) fail(e, "hipDeviceSynchronize");
  }
  if (rc == SMD_OK && (e = hipMemcpy(out.data(), dFlux, out.size() * 8, hipMemcpyDeviceToHost)) != hipSuccess)
    fail(e, "hipMemcpy");
  if (rc == SMD_OK) {
    // flux[k] is stored at the index of its step in increasing order
    for (int q = 0; q < nSweeps; ++q)
      for (int k = 0; k < 31; ++k) flux[(size_t)q * 31 + (downs[q] ? 30 - k : k)] = out[(size_t)q * 31 + k];
  }
  if (dR) (void)hipFree(dR);
  if (dA) (void)hipFree(dA);
  if (dFlux) (void)hipFree(dFlux);
  if (dSteps) (void)hipFree(dSteps);
  if (dSw) (void)hipFree(dSw);
  return rc;
}
