// Per-element P1 geometry shared by every kernel, so that the COO ops, the
// CSR assembly and the adjoint gradient all use bit-identical factors.
#pragma once
#include <hip/hip_runtime.h>

#include "internal.h"

namespace admfem {

struct ElemGeom {
  double gx[3], gy[3];  // physical gradients of the 3 barycentric shape functions
  double area;          // Heron's formula (deps/MFEM/Common.cpp:9-15, 83)
};

// Heron's formula with Eigen's norm() (sqrt of the squared norm), as
// deps/MFEM/Common.cpp:9-15.
__device__ __forceinline__ double heron_area(double x0, double y0, double x1,
                                             double y1, double x2, double y2) {
#pragma clang fp contract(off)
  double a = sqrt((x0 - x1) * (x0 - x1) + (y0 - y1) * (y0 - y1));
  double b = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  double c = sqrt((x0 - x2) * (x0 - x2) + (y0 - y2) * (y0 - y2));
  double s = (a + b + c) / 2.0;
  return sqrt(s * (s - a) * (s - b) * (s - c));
}

// Gradients of (1-xi-eta, xi, eta) mapped to the physical triangle:
// grad phi_k = (y_{k+1} - y_{k+2}, x_{k+2} - x_{k+1}) / det, the closed form
// of CalcPhysDShape (dshape_ref * J^{-1}) for the affine P1 map
// (deps/MFEM/Common.cpp:119-126).
__device__ __forceinline__ void elem_grads(double x0, double y0, double x1, double y1,
                                           double x2, double y2, double (&gx)[3],
                                           double (&gy)[3]) {
  double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
  double id = 1.0 / det;
  gx[0] = (y1 - y2) * id;
  gy[0] = (x2 - x1) * id;
  gx[1] = (y2 - y0) * id;
  gy[1] = (x0 - x2) * id;
  gx[2] = (y0 - y1) * id;
  gy[2] = (x1 - x0) * id;
}
__device__ __forceinline__ ElemGeom elem_geom(double x0, double y0, double x1,
                                              double y1, double x2, double y2) {
  ElemGeom g;
  elem_grads(x0, y0, x1, y1, x2, y2, g.gx, g.gy);
  g.area = heron_area(x0, y0, x1, y1, x2, y2);
  return g;
}

// Shape value h_r at quadrature point (xi, eta): (1-xi-eta, xi, eta)
// (deps/MFEM/Common.cpp:106-108).
__device__ __forceinline__ double shape(int r, double xi, double eta) {
  return r == 0 ? 1.0 - xi - eta : (r == 1 ? xi : eta);
}

// Gauss weight w = ip.weight * area / 0.5 (deps/MFEM/Common.cpp:116).
__device__ __forceinline__ double gauss_weight(double ipw, double area) {
  return ipw * area / 0.5;
}

// Streamed COO output: triplet t's index pair (ind[2t], ind[2t + 1]) and
// value, written with non-temporal stores (nothing reads them back soon),
// the pair as one 16-byte store when `ind` is 16-byte aligned.
typedef long long ll2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_triplet(long long *ind, double *vv, int64_t t,
                                              long long a, long long b, double v,
                                              bool ind16) {
  if (ind16) {
    const ll2_t ab = {a, b};
    __builtin_nontemporal_store(ab, reinterpret_cast<ll2_t *>(ind) + t);
  } else {
    __builtin_nontemporal_store(a, ind + 2 * t);
    __builtin_nontemporal_store(b, ind + 2 * t + 1);
  }
  __builtin_nontemporal_store(v, vv + t);
}

// Blocks are dispatched round-robin over the 8 XCDs (block b to XCD b % 8).
// This bijection gives each XCD one contiguous range of logical blocks, so
// data shared by neighbouring blocks is reused in that XCD's L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7, x = b & 7;
  return x * per + min(x, rem) + (b >> 3);
}

// Wavefront / block reductions (deterministic: fixed butterfly order).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a block of `NT` threads; result valid in every thread.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < NT / 64; k++) t += sh[k];
  return t;
}

}  // namespace admfem
