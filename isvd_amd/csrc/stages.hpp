// Stage implementations on DEVICE pointers (P = 1: the whole matrix lives on this GPU).
// Layouts are the reference's: yst = Y^T (Nl x m col-major), qt = Qbar^T (l x m),
// ut = U^T (l x m, first k rows), vt = V^T (l x n, first k rows); A col-major (ordera 'C')
// or row-major (ordera 'R').
#pragma once

#include "common.hpp"

namespace isvd {

struct Dims {
  i64 m, n, k, l, N, Nl;
};

struct KnArgs {
  i64 maxit = 256;
  double tol = 1e-4;
};

struct WyArgs {
  i64 maxit = 256;
  double tol = 1e-3, tau0 = 1.0, taumax = 1e14, taumin = 1e-14;
  i64 taumaxit = 20;
  double beta = 0.5, sigma = 1e-4, eta = 0.85;
};

template <typename T>
void gaussian_omega(i64 rows, i64 cols, T *omegat, i64 ld, unsigned long long seed);

template <typename T>
void stage_sketch(const Dims &d, char ordera, const T *a, i64 lda, const T *omegat, i64 ldo,
                  T *yst, i64 ldyst);

// The sketch and the projection Z^T = Qbar^T op(A) for a host-resident A staged through HBM
// by blocks of `cols` columns (the memory-limited path of the reference's GPU stages)
template <typename T>
void stage_sketch_host_blocks(const Dims &d, char ordera, const T *a_host, i64 lda,
                              const T *omegat, i64 ldo, T *yst, i64 ldyst, i64 cols);
template <typename T>
void project_host_blocks(const Dims &d, char ordera, const T *a_host, i64 lda, const T *qt,
                         i64 ldqt, T *zt, i64 cols);

template <typename T>
void stage_orth_gramian(const Dims &d, T *yst, i64 ldyst, bool cholqr = false);

// cholqr: the caller's next stage depends on the bases only through their projectors (KN), so
// CholeskyQR2 may replace the Householder tree where the device certifies it (same span and
// orthogonality to O(u), different rounding); otherwise the Householder tree, the reference's
// factorization.
template <typename T>
void stage_orth_tsqr(const Dims &d, T *yst, i64 ldyst, bool cholqr = false);

// returns iterations in *iter and the final error in *err (host values)
template <typename T>
void stage_integrate_kn(const Dims &d, const KnArgs &args, const T *yst, i64 ldyst, T *qt,
                        i64 ldqt, i64 *iter, double *err);

template <typename T>
void stage_integrate_wy(const Dims &d, const WyArgs &args, const T *yst, i64 ldyst, T *qt,
                        i64 ldqt, i64 *iter, double *mu);

// want_u / want_v mirror ut_root >= -1 / vt_root >= -1
template <typename T>
void stage_post_gramian(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                        T *s, T *ut, i64 ldut, T *vt, i64 ldvt, bool want_u, bool want_v);

template <typename T>
void stage_post_tsqr(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                     T *s, T *ut, i64 ldut, T *vt, i64 ldvt, bool want_u, bool want_v);

// Hierarchical-reduction integration; yst is overwritten
template <typename T>
void stage_integrate_hr(const Dims &d, T *yst, i64 ldyst, T *qt, i64 ldqt);

// Symmetric postprocessing (m == n): s ascending (l values), U^T = first k rows of ut
template <typename T>
void stage_post_symmetric(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                          T *s, T *ut, i64 ldut, bool want_u);

// The integration loops in the coordinates of the sketch basis, from M = Qs^T Qs (Nl x Nl):
// x_out (Nl x l) receives the coefficients of Qbar = Qs x_out.  The FP32 KN loop may take M as
// its two BF16 terms instead (ms != nullptr, M unused): it then iterates with M'' = Mh + Ml and
// forms its products with M on the BF16 matrix cores (split_sym_apply).
struct SplitSym;
template <typename T>
void integrate_kn_coords(const Dims &d, const KnArgs &args, const T *M, T *x_out, i64 *iter,
                         double *err, const SplitSym *ms = nullptr);

template <typename T>
void integrate_wy_coords(const Dims &d, const WyArgs &args, const T *M, T *x_out, i64 *iter,
                         double *mu);

// Postprocessing from Z^T = Qbar^T op(A) (l x n, ld l), which the TSQR variant overwrites.
template <typename T>
void post_gramian_z(const Dims &d, const T *qt, i64 ldqt, T *zt, T *s, T *ut, i64 ldut, T *vt,
                    i64 ldvt, bool want_u, bool want_v);

template <typename T>
void post_tsqr_z(const Dims &d, const T *qt, i64 ldqt, T *zt, T *s, T *ut, i64 ldut, T *vt,
                 i64 ldvt, bool want_u, bool want_v);

// dL/dA of the rank-k truncated SVD from dL/dS (k), dL/dU^T (k x m, ldgut) and dL/dV^T
// (k x n, ldgvt); U, S, V are recomputed from Qbar^T by the Gramian (tsqr = false) or TSQR
// postprocessing.  dA has the storage order of A (ordera).
template <typename T>
void stage_isvd_backward(const Dims &d, bool tsqr, char ordera, const T *a, i64 lda,
                         const T *qt, i64 ldqt, const T *gs, const T *gut, i64 ldgut,
                         const T *gvt, i64 ldgvt, T *da, i64 ldda);

}  // namespace isvd
