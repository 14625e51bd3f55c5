// The general sparse solve u = A \ rhs and its gradient, for a matrix given
// as COO triplets -- ADCME's SparseSolver op (the `\` of a SparseTensor,
// README.md:54,75) and SparseSolverGrad: lambda = A^-T dL/du,
// dL/dvv_k = -lambda[i_k] * u[j_k], dL/drhs = lambda.
//
// Duplicates are summed in triplet order (Eigen's setFromTriplets order);
// the matrix must be symmetric positive definite, since the solve is the
// library's Jacobi-PCG (then A^-T = A^-1).
#include <algorithm>
#include <numeric>

#include "internal.h"

namespace admfem {

System *system_from_csr(int n, std::vector<int> rp, std::vector<int> col);

namespace {

// val[slot] = sum of vv over the slot's triplets, in triplet order
__global__ void coo_sum(int nslot, const int *__restrict__ start, const int64_t *__restrict__ trip,
                        const double *__restrict__ vv, double *__restrict__ val) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslot) return;
  double a = 0.0;
  for (int t = start[s]; t < start[s + 1]; t++) a += vv[trip[t]];
  val[s] = a;
}

__global__ void coo_grad(int64_t nnz, const long long *__restrict__ ind,
                         const double *__restrict__ lam, const double *__restrict__ u,
                         double *__restrict__ gvv) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) gvv[k] = -lam[ind[2 * k]] * u[ind[2 * k + 1]];
}

inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

struct CooSystem {
  System *s = nullptr;
  ~CooSystem() { system_free(s); }
};

// Pattern + summed values of the COO matrix, ready to solve.
void build(CooSystem &cs, const long long *ind_dev, const double *vv, int64_t nnz, int n,
           hipStream_t st) {
  REQUIRE(n > 0 && nnz >= 0 && nnz < INT32_MAX, ADMFEM_ERR_ARG, "bad sizes");
  std::vector<long long> ind(2 * (size_t)nnz);
  if (nnz)
    HIPCHECK(hipMemcpyAsync(ind.data(), ind_dev, sizeof(long long) * 2 * nnz,
                            hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  for (int64_t k = 0; k < 2 * nnz; k++)
    REQUIRE(ind[k] >= 0 && ind[k] < n, ADMFEM_ERR_ARG, "triplet index out of range [0, n)");
  std::vector<int64_t> ord(nnz);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    return ind[2 * a] < ind[2 * b] || (ind[2 * a] == ind[2 * b] && ind[2 * a + 1] < ind[2 * b + 1]);
  });
  std::vector<int> rp(n + 1, 0), col, start;
  col.reserve(nnz);
  start.reserve(nnz + 1);
  for (int64_t t = 0; t < nnz; t++) {
    const int64_t k = ord[t];
    const bool fresh = t == 0 || ind[2 * k] != ind[2 * ord[t - 1]] ||
                       ind[2 * k + 1] != ind[2 * ord[t - 1] + 1];
    if (fresh) {
      col.push_back((int)ind[2 * k + 1]);
      start.push_back((int)t);
      rp[ind[2 * k] + 1]++;
    }
  }
  start.push_back((int)nnz);
  for (int i = 0; i < n; i++) rp[i + 1] += rp[i];
  const int nslot = (int)col.size();
  cs.s = system_from_csr(n, std::move(rp), std::move(col));
  DevBuf<int> d_start;
  DevBuf<int64_t> d_trip;
  d_start.upload(start);
  d_trip.upload(ord);
  if (nslot) coo_sum<<<nb(nslot, 256), 256, 0, st>>>(nslot, d_start.p, d_trip.p, vv, cs.s->val.p);
  launch_to_sell(*cs.s, st);
  HIPCHECK(hipStreamSynchronize(st));  // d_start / d_trip go out of scope
}

}  // namespace

void sparse_solve(const long long *ind, const double *vv, int64_t nnz, int n, const double *rhs,
                  double *u, const admfem_solve_opts &o, admfem_solve_stats &stats,
                  hipStream_t st) {
  CooSystem cs;
  build(cs, ind, vv, nnz, n, st);
  cg_solve(*cs.s, rhs, u, o, stats, st);
}

void sparse_solve_grad(const long long *ind, const double *vv, int64_t nnz, int n,
                       const double *u, const double *grad_u, double *grad_vv,
                       double *grad_rhs, const admfem_solve_opts &o, admfem_solve_stats &stats,
                       hipStream_t st) {
  CooSystem cs;
  build(cs, ind, vv, nnz, n, st);
  cg_solve(*cs.s, grad_u, grad_rhs, o, stats, st);  // lambda = A^-1 dL/du = dL/drhs
  if (nnz) coo_grad<<<nb(nnz, 256), 256, 0, st>>>(nnz, ind, grad_rhs, u, grad_vv);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(st));
}

}  // namespace admfem
