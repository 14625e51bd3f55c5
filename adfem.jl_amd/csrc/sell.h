// SELL-64 row products shared by the solvers' step kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace admfem {

// A pointer known to address global memory: loads through it compile to
// global (not flat) memory instructions.  For pointers read from memory
// (descriptor tables), where the compiler cannot infer the address space.
template <class T>
using gptr = const __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> global_ptr(const T *p) {
  return (gptr<T>)p;
}
template <class T>
using gmptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gmptr<T> global_mut(T *p) {
  return (gmptr<T>)p;
}

// Row i's product with its SELL row: sum_k sv_k comb(load(c_k)) in k order.
// Callers run one row per lane with every lane of the wave active and the
// wave's rows 64 consecutive rows from a multiple of 64 (the solvers' blocks
// are 4 such waves and npad is a multiple of 64).  A wave's 64 rows are then
// one slice, so the row width is wave-uniform and each
// width up to 8 gets straight-line code: every column and value load, then
// every neighbour's gathered values (load), is issued before the sums.
// (With the loads inside per-k branches the compiler waited for each
// neighbour's values before requesting the next one's.)
template <int W, typename VT, class C, class L, class F>
__device__ __forceinline__ double sell_dot_w(int off, C col, const VT *__restrict__ sv, L load,
                                             F comb) {
  int c[W];
  VT v[W];
  const gptr<VT> gv = global_ptr(sv);
  // FP64 values: the matrix is streamed once per product, so its loads are
  // non-temporal (measured on the FP64 SpMV: 99 -> 82 us at 2048^2; FP32
  // values lose with them, 49 -> 56 us, and keep ordinary loads)
#pragma unroll
  for (int k = 0; k < W; k++) {
    c[k] = col(off + 64 * k);
    if constexpr (sizeof(VT) == 8) v[k] = __builtin_nontemporal_load(gv + off + 64 * k);
    else v[k] = gv[off + 64 * k];
  }
  decltype(load(0)) g[W];
#pragma unroll
  for (int k = 0; k < W; k++) g[k] = load(c[k]);
  __builtin_amdgcn_sched_barrier(0);  // the gathers stay ahead of the sums
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < W; k++) s += (double)v[k] * comb(g[k]);
  return s;
}

// The product given the column of each SELL position (col(e) for entry e).
template <typename VT, class C, class L, class F>
__device__ __forceinline__ double sell_dot_cols(int i, const int *__restrict__ sp, C col,
                                                const VT *__restrict__ sv, L load, F comb) {
  const int sl = i >> 6;
  const int s0 = global_ptr(sp)[sl];
  const int off = s0 + (i & 63);
  const int wfull = (global_ptr(sp)[sl + 1] - s0) >> 6;
  double s = 0.0;
  switch (__builtin_amdgcn_readfirstlane(min(wfull, 8))) {  // wave-uniform
    case 1: s = sell_dot_w<1>(off, col, sv, load, comb); break;
    case 2: s = sell_dot_w<2>(off, col, sv, load, comb); break;
    case 3: s = sell_dot_w<3>(off, col, sv, load, comb); break;
    case 4: s = sell_dot_w<4>(off, col, sv, load, comb); break;
    case 5: s = sell_dot_w<5>(off, col, sv, load, comb); break;
    case 6: s = sell_dot_w<6>(off, col, sv, load, comb); break;
    case 7: s = sell_dot_w<7>(off, col, sv, load, comb); break;
    case 8: s = sell_dot_w<8>(off, col, sv, load, comb); break;
    default: break;
  }
  for (int k = 8; k < wfull; k++)
    s += (double)sv[off + 64 * k] * comb(load(col(off + 64 * k)));
  return s;
}

// with 32-bit column ids (non-temporal loads next to FP64 values)
template <typename VT, class L, class F>
__device__ __forceinline__ double sell_dot(int i, const int *__restrict__ sp,
                                           const int *__restrict__ sc,
                                           const VT *__restrict__ sv, L load, F comb) {
  const gptr<int> gc = global_ptr(sc);
  if constexpr (sizeof(VT) == 8)
    return sell_dot_cols<VT>(
        i, sp, [&](int e) { return __builtin_nontemporal_load(gc + e); }, sv, load, comb);
  else
    return sell_dot_cols<VT>(i, sp, [&](int e) { return gc[e]; }, sv, load, comb);
}

// with 16-bit column offsets from the row
template <typename VT, class L, class F>
__device__ __forceinline__ double sell_dot16(int i, const int *__restrict__ sp,
                                             const int16_t *__restrict__ sd,
                                             const VT *__restrict__ sv, L load, F comb) {
  const gptr<int16_t> gd = global_ptr(sd);
  return sell_dot_cols<VT>(i, sp, [&](int e) { return i + (int)gd[e]; }, sv, load, comb);
}

// the same with f(c) both gathering and combining
template <typename VT, class F>
__device__ __forceinline__ double sell_dot(int i, const int *__restrict__ sp,
                                           const int *__restrict__ sc,
                                           const VT *__restrict__ sv, F f) {
  return sell_dot<VT>(i, sp, sc, sv, f, [](double x) { return x; });
}

// Row i's products with the FP64 and the FP32 matrix of the same pattern in
// one pass over its columns (each sum in the order sell_dot takes):
// s64 = sum_k sv64_k cf(lf(c_k)), s32 = sum_k sv32_k cg(lg(c_k)), every
// gather issued before the sums as in sell_dot.
template <int W, class LF, class CF, class LG, class CG>
__device__ __forceinline__ void sell_dot2_w(int off, const int *__restrict__ sc,
                                            const double *__restrict__ sv64,
                                            const float *__restrict__ sv32, LF lf, CF cf, LG lg,
                                            CG cg, double &s64, double &s32) {
  int c[W];
  double v[W];
  float u[W];
  const gptr<int> gc = global_ptr(sc);
  const gptr<double> g64 = global_ptr(sv64);
  const gptr<float> g32 = global_ptr(sv32);
#pragma unroll
  for (int k = 0; k < W; k++) {
    c[k] = gc[off + 64 * k];
    v[k] = __builtin_nontemporal_load(g64 + off + 64 * k);
    u[k] = g32[off + 64 * k];
  }
  decltype(lf(0)) a[W];
  decltype(lg(0)) b[W];
#pragma unroll
  for (int k = 0; k < W; k++) {
    a[k] = lf(c[k]);
    b[k] = lg(c[k]);
  }
  __builtin_amdgcn_sched_barrier(0);  // the gathers stay ahead of the sums
#pragma unroll
  for (int k = 0; k < W; k++) {
    s64 += v[k] * cf(a[k]);
    s32 += (double)u[k] * cg(b[k]);
  }
}

template <class LF, class CF, class LG, class CG>
__device__ __forceinline__ void sell_dot2(int i, const int *__restrict__ sp,
                                          const int *__restrict__ sc,
                                          const double *__restrict__ sv64,
                                          const float *__restrict__ sv32, LF lf, CF cf, LG lg,
                                          CG cg, double &s64, double &s32) {
  const int sl = i >> 6;
  const int s0 = global_ptr(sp)[sl];
  const int off = s0 + (i & 63);
  const int wfull = (global_ptr(sp)[sl + 1] - s0) >> 6;
  s64 = 0.0;
  s32 = 0.0;
  switch (__builtin_amdgcn_readfirstlane(min(wfull, 8))) {  // wave-uniform
    case 1: sell_dot2_w<1>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 2: sell_dot2_w<2>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 3: sell_dot2_w<3>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 4: sell_dot2_w<4>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 5: sell_dot2_w<5>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 6: sell_dot2_w<6>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 7: sell_dot2_w<7>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    case 8: sell_dot2_w<8>(off, sc, sv64, sv32, lf, cf, lg, cg, s64, s32); break;
    default: break;
  }
  for (int k = 8; k < wfull; k++) {
    const int ck = sc[off + 64 * k];
    s64 += sv64[off + 64 * k] * cf(lf(ck));
    s32 += (double)sv32[off + 64 * k] * cg(lg(ck));
  }
}

}  // namespace admfem
