// An entry point that allocates nothing: the p-norm stress functional of the
// current state with its reduction scratch in a file-scope device array
// instead of the problem's buffers.  For callers that may not allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc).
//
// The array lives once per GPU for the life of the process; every call on a
// GPU shares it and it keeps the last call's contents (each call writes all
// the entries it reads, so nothing needs clearing).  Calls that overlap on
// one GPU race on it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "problem.h"

namespace tmrx {

StressArgs Problem_stress_args(const Material &m, double P, const double *CB, const double *vol);

// partial sums of 256-element blocks: up to NOALLOC_MAX_BLOCKS * 256 =
// 16,777,216 elements; the PCG scalar slots in front (SC_N doubles)
constexpr int NOALLOC_MAX_BLOCKS = 65536;
constexpr int NOALLOC_WORDS = SC_N + NOALLOC_MAX_BLOCKS;

static __device__ __align__(256) double scratch[NOALLOC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// Device address of `scratch` on device `dev` (current), looked up on the
// first call per device.  Host code uses this address, never the array's
// name, which on the host is only a stand-in for the symbol.
static int scratch_address(int dev, double **out, std::string &err) {
  if (dev < 0 || dev >= MAX_DEVICES) {
    err = "stress_pnorm_noalloc: device index out of range";
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("stress_pnorm_noalloc: hipGetSymbolAddress failed: ") +
            hipGetErrorString(e);
      return -1;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

int Problem::stress_pnorm_noalloc(double P, double *f) {
  if (!have_solution_) {
    last_error = "stress_pnorm_noalloc: solve first";
    return -1;
  }
  if (!(P >= 2.0)) {
    last_error = "stress_pnorm_noalloc: the exponent must be >= 2";
    return -1;
  }
  Level &L = levels_[0];
  const int nb = (L.n_elem + 255) / 256;
  if (nb > NOALLOC_MAX_BLOCKS) {
    last_error = "stress_pnorm_noalloc: " + std::to_string(L.n_elem) +
                 " elements exceed the fixed scratch (at most " +
                 std::to_string((long long)NOALLOC_MAX_BLOCKS * 256) + ")";
    return -1;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    last_error = "stress_pnorm_noalloc: hipGetDevice failed";
    return -1;
  }
  double *scr = nullptr;
  if (scratch_address(dev, &scr, last_error)) return -1;
  double *sc = scr, *parts = scr + SC_N;
  const StressArgs sa = Problem_stress_args(mat_, P, CB_.p, cvol_.p);
  launch_stress_sum(L.n_elem, L.d_conn.p, L.dep_ptr.p, L.dep_conn.p, L.dep_w.p, state_mask(),
                    L.d_cls.p, L.d_lev.p, rho_.p, design_mode_ == 1 ? rho_gp_.p : nullptr, u_.p,
                    sa, parts, st_);
  launch_finalize(nb, parts, 4, sc, st_);
  double h = 0.0;
  hipError_t e = take_launch_error();
  if (e == hipSuccess)
    e = hipMemcpyAsync(&h, sc + SC_BB, sizeof(double), hipMemcpyDeviceToHost, st_);
  if (e == hipSuccess) e = hipStreamSynchronize(st_);
  if (e != hipSuccess) {
    last_error = std::string("stress_pnorm_noalloc: ") + hipGetErrorString(e);
    return -1;
  }
  *f = std::pow(h, 1.0 / P);
  last_stress_sum_ = h;
  return 0;
}

}  // namespace tmrx
