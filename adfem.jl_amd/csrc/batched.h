// Device-side descriptors of the batched solver and the serving engine.
#pragma once
#include "internal.h"

namespace admfem {

constexpr int BATCH_NT = 256;  // rows per block of the batched kernels
constexpr int REPLACE_EVERY = 8;  // iterations between replacement points

// One system of the batch: its matrix (topology arrays shared with every
// system of the same mesh) and its vectors.  Node vectors are in the
// topology's internal order, padded to npad.  The FP32 inner iteration works
// on the symmetrically scaled system (see batched.hip).
struct BSys {
  const int *sp, *sc;       // SELL slice offsets / columns (topology)
  const int16_t *sd;        // the columns as 16-bit offsets from the row, or null
  const int *row_bd;        // boundary index per row (topology)
  const double *sv64, *diag64;  // FP64 matrix (replacement residuals)
  const float *sv32;        // FP32 scaled off-diagonals a_ij dsq_i dsq_j
  const double *dsq;        // 1 / sqrt(a_ii)
  const double *b;          // right-hand side
  double *x[2];             // FP64 solution, two buffers (swapped on replacement)
  float *ys[2];             // FP32 scaled correction since the last replacement
  float *r[2], *p[2], *q[2];  // scaled residual, direction, A p: the step reads
                              // one buffer of each and writes the other
  double *part;             // 4 * nblk block partials
  int n, npad, nblk;
};

// CG state of one system: rz (= scaled r.r of the current residual) drives
// the iteration, rr is the true FP64 ||b - A x||^2 of the last replacement;
// alpha and beta are the pending step's.
struct BScal {
  double rtol2, delta2, bb, tol2, rz, alpha, beta, rr, rr_ref, rz_ref, rz_min;
  int status, it, maxit, replace, nrepl, sel, stall;
};

// A Gaussian random field of random Fourier features:
// g(x) = scale * sum_j cos(wx_j x + wy_j y + phi_j).
struct SynField {
  int nmodes;
  double scale;
  double wx[32], wy[32], phi[32];
};

void batch_init(const BSys *D, BScal *S, int sys, int nblk, double rtol2, double delta2,
                int maxit, hipStream_t st);
// the block map of nsys systems (slots slot[k], first logical block
// pref[k], nmap blocks in all) over gmap >= nmap grid blocks, XCD-aware
void batch_map(int nsys, const int *slot, const int *pref, int nmap, int gmap, int2 *map,
               hipStream_t st);
void batch_iterations(const int2 *map, int nmap, const int *list, int nlist, const BSys *D,
                      BScal *S, int iters, hipStream_t st);
void batch_spmv_repeat(const int2 *map, int nmap, const BSys *D, const BScal *S, int repeat,
                       hipStream_t st);
void batch_scale(const System &s, const double *sv64, const double *diag64, double *dsq,
                 float *sv32, hipStream_t st);
void batch_final(const BSys *D, const BScal *S, int sys, int npad, double *out,
                 hipStream_t st);
void synth_inputs(const System &s, const double *base_x, const double *base_y, uint64_t seed,
                  double amp, const SynField &kf, const SynField *uf, double *x, double *y,
                  double2 *xy, double *kappa, double *f, double *u_prev_int, hipStream_t st);

}  // namespace admfem
