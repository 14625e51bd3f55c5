// The P2 (degree 2) basis of MFEM's order-2 H1 triangle, shared by the P2
// operator kernels and the fused P2 system: nodal at the vertices and edge
// midpoints, with barycentric l0 = 1 - xi - eta, l1 = xi, l2 = eta,
//   phi_k = l_k (2 l_k - 1) (k < 3), phi_3 = 4 l0 l1, phi_4 = 4 l1 l2,
//   phi_5 = 4 l2 l0,
// and physical gradients from the constant grad l_k of the affine map
// (CalcPhysDShape, deps/MFEM/Common.cpp:119-126).
#pragma once
#include "geom.h"

namespace admfem {

__device__ __forceinline__ ElemGeom p2_geom(const int *__restrict__ conn, int nelem, int e,
                                            const double *__restrict__ x,
                                            const double *__restrict__ y) {
  double ex[3], ey[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int v = conn[(size_t)k * nelem + e];
    ex[k] = x[v];
    ey[k] = y[v];
  }
  ElemGeom G;
  elem_grads(ex[0], ey[0], ex[1], ey[1], ex[2], ey[2], G.gx, G.gy);
  G.area = 0.0;  // unused: Gauss weights come from the mesh's array (same bits)
  return G;
}

// P2 basis values at (xi, eta).
__device__ __forceinline__ void p2_shape(double xi, double eta, double h[6]) {
#pragma clang fp contract(off)
  const double l0 = 1.0 - xi - eta;
  h[0] = l0 * (2.0 * l0 - 1.0);
  h[1] = xi * (2.0 * xi - 1.0);
  h[2] = eta * (2.0 * eta - 1.0);
  h[3] = 4.0 * l0 * xi;
  h[4] = 4.0 * xi * eta;
  h[5] = 4.0 * eta * l0;
}

// Physical gradients of the P2 basis at (xi, eta).
__device__ __forceinline__ void p2_grad(double xi, double eta, const ElemGeom &G, double gx[6],
                                        double gy[6]) {
#pragma clang fp contract(off)
  const double l[3] = {1.0 - xi - eta, xi, eta};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    gx[k] = (4.0 * l[k] - 1.0) * G.gx[k];
    gy[k] = (4.0 * l[k] - 1.0) * G.gy[k];
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {  // edge j joins vertices a = j, b = (j + 1) % 3
    const int a = j, b = (j + 1) % 3;
    gx[3 + j] = 4.0 * (l[b] * G.gx[a] + l[a] * G.gx[b]);
    gy[3 + j] = 4.0 * (l[b] * G.gy[a] + l[a] * G.gy[b]);
  }
}

}  // namespace admfem
