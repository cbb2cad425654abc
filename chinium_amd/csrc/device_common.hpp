// Device helpers shared by the XC kernels: the FP64 MFMA wrapper and AO
// evaluation of one shell at one point.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace chx {

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ dbl4 mfma_f64(double a, double b, dbl4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// AO evaluation of one shell at one point (relative coordinates x, y, z).
// phi_m = R(r) Q_m(x,y,z), R = sum_p c_p exp(-a_p r^2); components in the
// reference order (src/Grid/AO/{S,CartP,PureP,PureD,PureF}.hpp).
template <bool GRAD>
__device__ __forceinline__ int shell_eval(const BasisView& bv, int sh, double px, double py, double pz,
                                          double* v, double* gx, double* gy, double* gz) {
  const double x = px - bv.sh_xyz[3 * sh], y = py - bv.sh_xyz[3 * sh + 1], z = pz - bv.sh_xyz[3 * sh + 2];
  const double r2 = x * x + y * y + z * z;
  const int np = bv.sh_nprim[sh], po = bv.sh_poff[sh];
  double A0 = 0, A1 = 0;
  for (int p = 0; p < np; ++p) {
    const double a = bv.p_exp[po + p];
    const double t = bv.p_coef[po + p] * ::exp(-a * r2);
    A0 += t;
    if (GRAD) A1 += a * t;
  }
  const double dAx = -2.0 * x * A1, dAy = -2.0 * y * A1, dAz = -2.0 * z * A1;
  const int type = bv.sh_type[sh];
  // q and grad q per component
  double q[7], qx[7], qy[7], qz[7];
  int n;
  const double r3 = 1.7320508075688772, r6 = 2.449489742783178, r10 = 3.1622776601683795, r15 = 3.872983346207417;
  switch (type) {
    case 0:
      n = 1; q[0] = 1; qx[0] = qy[0] = qz[0] = 0;
      break;
    case 1:
      n = 3;
      q[0] = x; qx[0] = 1; qy[0] = 0; qz[0] = 0;
      q[1] = y; qx[1] = 0; qy[1] = 1; qz[1] = 0;
      q[2] = z; qx[2] = 0; qy[2] = 0; qz[2] = 1;
      break;
    case -1:
      n = 3;
      q[0] = y; qx[0] = 0; qy[0] = 1; qz[0] = 0;
      q[1] = z; qx[1] = 0; qy[1] = 0; qz[1] = 1;
      q[2] = x; qx[2] = 1; qy[2] = 0; qz[2] = 0;
      break;
    case -2:
      n = 5;
      q[0] = r3 * x * y;                  qx[0] = r3 * y; qy[0] = r3 * x; qz[0] = 0;
      q[1] = r3 * y * z;                  qx[1] = 0;      qy[1] = r3 * z; qz[1] = r3 * y;
      q[2] = z * z - 0.5 * (x * x + y * y); qx[2] = -x;   qy[2] = -y;     qz[2] = 2 * z;
      q[3] = r3 * x * z;                  qx[3] = r3 * z; qy[3] = 0;      qz[3] = r3 * x;
      q[4] = 0.5 * r3 * (x * x - y * y);  qx[4] = r3 * x; qy[4] = -r3 * y; qz[4] = 0;
      break;
    default: {  // -3, pure f
      n = 7;
      const double xx = x * x, yy = y * y, zz = z * z;
      const double c10 = 0.25 * r10, c6 = 0.25 * r6;
      q[0] = c10 * y * (3 * xx - yy);     qx[0] = c10 * 6 * x * y;       qy[0] = c10 * 3 * (xx - yy);        qz[0] = 0;
      q[1] = r15 * x * y * z;             qx[1] = r15 * y * z;           qy[1] = r15 * x * z;                qz[1] = r15 * x * y;
      q[2] = c6 * y * (4 * zz - xx - yy); qx[2] = -c6 * 2 * x * y;       qy[2] = c6 * (4 * zz - xx - 3 * yy); qz[2] = c6 * 8 * y * z;
      q[3] = z * (zz - 1.5 * (xx + yy));  qx[3] = -3 * x * z;            qy[3] = -3 * y * z;                 qz[3] = 3 * zz - 1.5 * (xx + yy);
      q[4] = c6 * x * (4 * zz - xx - yy); qx[4] = c6 * (4 * zz - 3 * xx - yy); qy[4] = -c6 * 2 * x * y;     qz[4] = c6 * 8 * x * z;
      q[5] = 0.5 * r15 * (xx - yy) * z;   qx[5] = r15 * x * z;           qy[5] = -r15 * y * z;               qz[5] = 0.5 * r15 * (xx - yy);
      q[6] = c10 * x * (xx - 3 * yy);     qx[6] = c10 * 3 * (xx - yy);   qy[6] = -c10 * 6 * x * y;           qz[6] = 0;
      break;
    }
  }
#pragma unroll
  for (int m = 0; m < 7; ++m) {
    if (m < n) {
      v[m] = A0 * q[m];
      if (GRAD) {
        gx[m] = dAx * q[m] + A0 * qx[m];
        gy[m] = dAy * q[m] + A0 * qy[m];
        gz[m] = dAz * q[m] + A0 * qz[m];
      }
    }
  }
  return n;
}

// Block index -> work index such that each XCD (blocks are dealt to the 8
// XCDs round-robin) runs one contiguous range of the work list; neighbouring
// work items, which share operand tiles, then meet in the same L2.
__device__ __forceinline__ int xcd_remap(int i, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, r = n % kXcd, x = i % kXcd;
  return x * q + (x < r ? x : r) + i / kXcd;
}

}  // namespace chx
