// Design-vector transfer between forests (C-ABI, declared in include/tmrx.h):
// TopOptUtils.interpolateDesignVec (tmr/TopOptUtils.py:270-307) builds
// new_filter.createInterpolation(orig_filter) and multiplies the original
// nodal vector by it; here the interpolation CSR comes from the forest
// (the bit-exact createInterpolation restatement) and the product runs on the
// GPU.
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/tmrx.h"
#include "forest.h"
#include "kernels.h"

namespace tmrx {
void set_error(const std::string &s);
}
using tmrx::set_error;
tmrx::Forest *tmrx_forest_impl(tmrx_forest *f);

namespace {
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  bool alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1) == hipSuccess; }
};
}  // namespace

extern "C" int tmrx_forest_interpolate_vec(tmrx_forest *new_forest, tmrx_forest *orig_forest,
                                           int vars_per_node, const double *x_orig,
                                           double *x_new, int on_device) {
  tmrx::Forest *fn = tmrx_forest_impl(new_forest), *fo = tmrx_forest_impl(orig_forest);
  if (!fn || !fo) {
    set_error("tmrx_forest_interpolate_vec: null forest handle");
    return -1;
  }
  if (vars_per_node < 1 || !x_orig || !x_new) {
    set_error("tmrx_forest_interpolate_vec: invalid arguments");
    return -1;
  }
  if (fn->create_nodes() || fo->create_nodes()) {
    set_error("tmrx_forest_interpolate_vec: " + fn->last_error + fo->last_error);
    return -1;
  }
  tmrx::InterpCSR P;
  if (fn->create_interpolation(fo, P)) {
    set_error("tmrx_forest_interpolate_vec: " + fn->last_error);
    return -1;
  }
  const size_t nv = (size_t)vars_per_node;
  const size_t nin = (size_t)P.ncols * nv, nout = (size_t)P.nrows * nv;
  DevBuf dptr, dcol, dw, dx, dy;
  if (!dptr.alloc(P.rowp.size() * sizeof(int)) || !dcol.alloc(P.cols.size() * sizeof(int)) ||
      !dw.alloc(P.w.size() * sizeof(double)) ||
      (!on_device && (!dx.alloc(nin * sizeof(double)) || !dy.alloc(nout * sizeof(double))))) {
    set_error("tmrx_forest_interpolate_vec: device allocation failed");
    return -1;
  }
  const double *x = on_device ? x_orig : (const double *)dx.p;
  double *y = on_device ? x_new : (double *)dy.p;
  hipError_t e = hipMemcpy(dptr.p, P.rowp.data(), P.rowp.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dcol.p, P.cols.data(), P.cols.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(dw.p, P.w.data(), P.w.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && !on_device)
    e = hipMemcpy(dx.p, x_orig, nin * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    tmrx::launch_interp_nodal(P.nrows, (const int *)dptr.p, (const int *)dcol.p,
                              (const double *)dw.p, vars_per_node, x, y, nullptr);
    e = tmrx::take_launch_error();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && !on_device)
    e = hipMemcpy(x_new, dy.p, nout * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    set_error(std::string("tmrx_forest_interpolate_vec: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}
