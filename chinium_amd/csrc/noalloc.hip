// Allocation-free functional evaluation at points.
//
// chx_xc_eval_points stages its host arrays and the functional's code list in
// device buffers it allocates per call. chx_xc_eval_points_noalloc does the
// same evaluation with that staging held in one file-scope device array,
// `scratch`, of fixed size: it calls no hipMalloc / hipFree and creates no
// stream, so it suits callers that must not allocate (serving loops, code
// under allocation-free policies). The array's device address is taken once
// per device with hipGetSymbolAddress (the symbol's name is a host-side
// stand-in, not a device address) and only that address is handed to copies,
// fills and kernels. Every call on a device shares that device's array, which
// keeps the last call's contents: each call rewrites every region it reads
// (inputs copied in, absent sigma / tau cleared by a fill kernel), and calls
// that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace chx {
std::string& last_error();

namespace {

constexpr int kNoAllocMaxCodes = 64;                      // functional codes staged (ints)
constexpr size_t kNoAllocCodeDoubles = kNoAllocMaxCodes / 2;  // their room, in doubles (256 bytes)
// closed shell: rho, sigma, tau, exc, vrho, vsigma, vtau = 7 n doubles;
// polarised: rho 2n, sigma 3n, exc n, vrho 2n, vsigma 3n = 11 n doubles
constexpr size_t kNoAllocDoubles = kNoAllocCodeDoubles + 11 * (size_t)CHX_NOALLOC_MAX_POINTS;

__device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call on that device.
double* scratch_on_current_device() {
  int dev = 0;
  CHX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw Error(CHX_ERR_DEVICE, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw Error(CHX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

void up(double* d, const double* h, size_t n) {
  if (n) CHX_HIP(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, nullptr));
}
void down(double* h, const double* d, size_t n) {
  if (h && n) CHX_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, nullptr));
}

}  // namespace
}  // namespace chx

int chx_xc_eval_points_noalloc(const chx_xc* xc, long n, const double* rho, const double* sigma, const double* tau,
                               double* exc, double* vrho, double* vsigma, double* vtau) {
  using namespace chx;
  try {
    if (!xc || n < 0 || !rho || !exc || !vrho) throw Error(CHX_ERR_ARG, "bad arguments");
    if (n > CHX_NOALLOC_MAX_POINTS)
      throw Error(CHX_ERR_ARG, "chx_xc_eval_points_noalloc takes at most " + std::to_string(CHX_NOALLOC_MAX_POINTS) +
                                   " points per call (its scratch array is fixed); got " + std::to_string(n));
    const int ncodes = (int)xc->codes.size();
    if (ncodes > kNoAllocMaxCodes) throw Error(CHX_ERR_ARG, "too many functional codes for the scratch array");
    if (n == 0) return CHX_OK;
    if (xc->spin == 2 && xc->family >= 2)
      throw Error(CHX_ERR_UNSUPPORTED, "chx_xc_eval_points_noalloc takes spin-polarised LDA and GGA functionals");
    double* base = scratch_on_current_device();
    int* codes = reinterpret_cast<int*>(base);
    double* a = base + kNoAllocCodeDoubles;
    const size_t m = (size_t)n;
    CHX_HIP(hipMemcpyAsync(codes, xc->codes.data(), sizeof(int) * ncodes, hipMemcpyHostToDevice, nullptr));
    if (xc->spin == 2) {  // libxc XC_POLARIZED layout, as chx_xc_eval_points
      double *r = a, *sg = a + 2 * m, *e = a + 5 * m, *vr = a + 6 * m, *vs = a + 8 * m;
      up(r, rho, 2 * m);
      if (sigma) up(sg, sigma, 3 * m);
      launch_eval_functional_pol(nullptr, n, codes, ncodes, r, sigma ? sg : nullptr, nullptr, e, vr,
                                 vsigma ? vs : nullptr, nullptr);
      down(exc, e, m);
      down(vrho, vr, 2 * m);
      down(vsigma, vs, 3 * m);
    } else {
      double *r = a, *sg = a + m, *t = a + 2 * m, *e = a + 3 * m, *vr = a + 4 * m, *vs = a + 5 * m, *vt = a + 6 * m;
      up(r, rho, m);
      if (sigma) up(sg, sigma, m); else launch_fill(nullptr, sg, m, 0.0);
      if (tau) up(t, tau, m); else launch_fill(nullptr, t, m, 0.0);
      launch_eval_functional(nullptr, n, codes, ncodes, r, sg, t, e, vr, vs, vt);
      down(exc, e, m);
      down(vrho, vr, m);
      down(vsigma, vs, m);
      down(vtau, vt, m);
    }
    CHX_HIP(hipStreamSynchronize(nullptr));
    return CHX_OK;
  } catch (const Error& e) {
    last_error() = e.what();
    return e.code;
  } catch (const std::exception& e) {
    last_error() = e.what();
    return CHX_ERR_ARG;
  }
}
