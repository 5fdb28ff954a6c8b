// Blocked Householder QR (internal interface shared by the QR and SVD entry points).
#pragma once
#include "common.h"

namespace ptn {

constexpr int QR_NB = 32;       // panel width (columns per compact-WY block)
constexpr int QR_SPLITK = 8;    // K chunks of the V^H C products

// Factorisation B = Q R of an m x n complex128 matrix, held in device workspace:
//   W     column-major m x n; its strict upper triangle is that of R
//   Vf    column-major m x k Householder vectors v_j (v_j[j] = 1, zero above)
//   tau   k scalar factors, H_j = I - tau_j v_j v_j^H
//   rdiag k diagonal entries of R
//   rsup  k superdiagonal entries R[j][j+1] produced by two-reflector launches
//   T     per panel the QR_NB x QR_NB upper-triangular factor of the compact
//         WY form H_j0 ... H_j0+nb-1 = I - V T V^H (row-major, zero padded)
//   scl   per matrix the exact power of two the input was multiplied by
//         before the factorisation (1 unless its largest entry lies outside
//         [2^-256, 2^256], where sums of squares would underflow or overflow);
//         W, rdiag and rsup hold R * scl, the reflectors are scale-free
// plus scratch for the blocked updates (Vf^H C partials, the reduced product).
// Per-matrix strides (complex elements) of a batch of factorisations of
// identical shape; matrix b of a batch uses the arrays offset by b * stride.
struct QrStrides {
  int64_t W, V, k, T, G, P, Z;
};

struct QrFact {
  int64_t m, n, k, npan, ncmax, batch;
  double2 *W, *Vf, *tau, *rdiag, *rsup, *T, *G, *P, *Z;
  double* scl;
  unsigned long long* mxb;   // largest |Re|, |Im| per matrix (bits of a non-negative double)
  QrStrides bs;
};

size_t qr_fact_bytes(int64_t m, int64_t n, int64_t ncmax, int64_t batch = 1);
// carve the buffers of a factorisation (or of `batch` of them) out of `ws`
QrFact qr_fact_layout(int64_t m, int64_t n, int64_t ncmax, void* ws, int64_t batch = 1);
// factor the row-major m x n matrix `a` (leading dimension lda; matrix b of a
// batch at a + b * sa)
int qr_factor(QrFact& f, const double2* a, int64_t lda, hipStream_t s, int64_t sa = 0);
// C <- Q C for the column-major m x nc matrix C (leading dimension ldc, nc <= ncmax;
// matrix b of a batch at C + b * sc), Q = H_0 H_1 ... H_{k-1}. With
// `identity_tail` the caller guarantees that C starts as [I_k; 0] (nc == k), so
// panel j0 only needs columns j0 .. k-1.
int qr_apply_q(const QrFact& f, double2* C, int64_t ldc, int64_t nc, bool identity_tail,
               hipStream_t s, int64_t sc = 0);

}  // namespace ptn
