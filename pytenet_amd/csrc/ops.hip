// Local contractions of the sweep as two large complex products around an
// MPO "mode mix":
//   H_eff / right environment:  T1 = A.R   ->  T2 = W (x) T1  ->  L^T.T2  /  T2.B^H
//   left environment:           T1 = L.B*  ->  T2 = W^T (x) T1 ->  A^T.T2
//   bond operator:              T  = C.R   ->  L^T.T
// The mode mix contracts the (physical, MPO-bond) index pair of T1 with the
// MPO tensor for every (left bond, right bond) pair; it is HBM-bound and
// keeps the two products dense and MFMA-shaped.
#include "common.h"
#include "gemm_impl.h"
#include "ops.h"

namespace ptn {

static inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// precision code with the block-sparse flag removed, and the flag itself
static inline int base_prec(int prec) { return prec & ~PTN_BLOCK_SPARSE; }
static inline bool is_sparse(int prec) { return (prec & PTN_BLOCK_SPARSE) != 0; }
static inline size_t mask_bytes2(int64_t m1, int64_t n1, int64_t k1, int64_t m3, int64_t n3,
                                 int64_t k3) {
  size_t a = gemm_mask_bytes(m1, n1, k1), b = gemm_mask_bytes(m3, n3, k3);
  return a > b ? a : b;
}
static inline size_t csize(int prec) { return prec == PTN_FP64 ? 16 : 8; }
static inline int ctype(int prec) { return prec == PTN_FP64 ? PTN_T_C128 : PTN_T_C64; }
static inline int gemm_compute(int prec) { return prec; }

// ---------------------------------------------------------------------------
// W (Ml, d, d, Mr) viewed as the matrix Wmat[(m, sb), (s, m')]; the left
// environment needs its transpose Wt[(s, m'), (m, sb)].
__global__ void transpose_w_kernel(int64_t rows, int64_t cols, const double2* __restrict__ w,
                                   double2* __restrict__ wt) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * cols) return;
  w += blockIdx.y * rows * cols;     // problem of a batch
  wt += blockIdx.y * rows * cols;
  int64_t i = t / cols, j = t % cols;
  wt[j * rows + i] = w[t];
}

__global__ void convert_c128_c64_kernel(int64_t n, const double2* __restrict__ x,
                                        float2* __restrict__ y) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    y[t] = make_float2((float)x[t].x, (float)x[t].y);
}

int convert_c128_c64(int64_t n, const void* x, void* y, hipStream_t s) {
  if (n == 0) return PTN_OK;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  convert_c128_c64_kernel<<<(unsigned)blocks, 256, 0, s>>>(n, (const double2*)x, (float2*)y);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---------------------------------------------------------------------------
// scratch of the pre-split split-bf16 products of a contraction (the larger
// of its two products; 0 in the other modes)
static size_t split2(int prec, int64_t batch, int64_t m1, int64_t n1, int64_t k1, int64_t m3,
                     int64_t n3, int64_t k3) {
  size_t a, b;
  if (base_prec(prec) == PTN_BF16X3) {
    a = bf16x3_split_bytes(m1, n1, k1, batch);
    b = bf16x3_split_bytes(m3, n3, k3, batch);
  } else if (base_prec(prec) == PTN_BF16X6) {
    a = bf16x6_split_bytes(m1, n1, k1, batch);
    b = bf16x6_split_bytes(m3, n3, k3, batch);
  } else {
    return 0;
  }
  return align256(a > b ? a : b);
}

// FP32 mode: the complex128 input vector of H_eff converted to complex64 once,
// so that the first product streams half the A-operand bytes
static size_t a64_bytes(int prec, const Dims& D, int64_t batch) {
  return base_prec(prec) == PTN_FP32 ? align256((size_t)batch * (size_t)(D.Dl * D.d * D.Dr) * 8)
                                     : 0;
}

size_t heff_ws_bytes(int prec, const Dims& D, int64_t batch) {
  const size_t sp = split2(prec, batch, D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl, D.d * D.Dr,
                           D.Dl * D.Ml);
  const size_t a64 = a64_bytes(prec, D, batch);
  prec = base_prec(prec);
  size_t t1 = (size_t)batch * (size_t)(D.Dl * D.d) * (size_t)(D.Mr * D.Dr) * csize(prec);
  size_t t2 = (size_t)batch * (size_t)(D.Dl * D.Ml) * (size_t)(D.d * D.Dr) * csize(prec);
  return align256(t1) + align256(t2) +
         mask_bytes2(D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl, D.d * D.Dr, D.Dl * D.Ml) + 256 + a64 +
         sp;
}

size_t env_ws_bytes(int prec, const Dims& D, int64_t batch) {
  prec = base_prec(prec);
  // right step: T1 (Dl*d, Mr*Dr), T2 (Dl*Ml, d*Dr); left step: T1 (Dl*Ml, d*Dr),
  // T2 (Dl*d, Mr*Dr); plus the transposed MPO tensor
  size_t a = (size_t)batch * (size_t)(D.Dl * D.d) * (size_t)(D.Mr * D.Dr) * csize(prec);
  size_t b = (size_t)batch * (size_t)(D.Dl * D.Ml) * (size_t)(D.d * D.Dr) * csize(prec);
  size_t wt = (size_t)batch * (size_t)(D.Ml * D.d * D.d * D.Mr) * 16;
  const size_t mr = mask_bytes2(D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl * D.Ml, D.Dl, D.d * D.Dr);
  const size_t ml = mask_bytes2(D.Dl * D.Ml, D.d * D.Dr, D.Dl, D.Dr, D.Mr * D.Dr, D.Dl * D.d);
  const size_t sr = split2(prec, batch, D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl * D.Ml, D.Dl, D.d * D.Dr);
  const size_t sl = split2(prec, batch, D.Dl * D.Ml, D.d * D.Dr, D.Dl, D.Dr, D.Mr * D.Dr, D.Dl * D.d);
  return 2 * align256(a > b ? a : b) + align256(wt) + (mr > ml ? mr : ml) + 256 +
         (sr > sl ? sr : sl);
}

size_t bond_ws_bytes(int prec, int64_t Dl, int64_t Dr, int64_t M, int64_t batch) {
  const size_t sp = split2(prec, batch, Dl, M * Dr, Dr, Dl, Dr, Dl * M);
  prec = base_prec(prec);
  return align256((size_t)batch * (size_t)Dl * (size_t)(M * Dr) * csize(prec)) +
         mask_bytes2(Dl, M * Dr, Dr, Dl, Dr, Dl * M) + 256 + sp;
}

// zero-width virtual bonds (Dl or Dr = 0, as a split that keeps no singular
// value leaves them, pytenet/bond_ops.py:28-29) are valid: the products then
// have an empty output or an empty sum (zeros), as NumPy's tensordot gives
static bool dims_ok(const Dims& D) {
  return D.Dl >= 0 && D.d > 0 && D.Dr >= 0 && D.Ml > 0 && D.Mr > 0 && D.d * D.Mr <= 32 &&
         D.Ml * D.d <= 32;
}

// H_eff: out = L^T . (W (x) (A . R)). `a` and `out` are complex128 or
// complex64 as given by ta / tout; L and R have types tl / tr.
int heff_apply(int prec, const Dims& D, const void* l, int tl, const void* w, const void* r, int tr,
               const void* a, int ta, void* out, int tout, void* ws, size_t ws_bytes,
               hipStream_t s, int64_t batch) {
  if (!dims_ok(D) || batch < 1) return PTN_ERR_ARG;
  const size_t wtot = heff_ws_bytes(prec, D, batch);
  if (ws_bytes < wtot) return PTN_ERR_WORKSPACE;
  const size_t spb = split2(prec, batch, D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl, D.d * D.Dr,
                            D.Dl * D.Ml);
  void* sws = spb ? (char*)ws + (wtot - spb) : nullptr;
  const size_t a64b = a64_bytes(prec, D, batch);
  const bool sp = is_sparse(prec) && batch == 1;
  prec = base_prec(prec);
  const int ct = ctype(prec), cp = gemm_compute(prec);
  const int64_t n1 = (D.Dl * D.d) * (D.Mr * D.Dr), n2 = (D.Dl * D.Ml) * (D.d * D.Dr);
  char* p = (char*)ws;
  void* t1 = p;
  p += align256((size_t)batch * n1 * csize(prec));
  void* t2 = p;
  p += align256((size_t)batch * n2 * csize(prec));
  const size_t mb = mask_bytes2(D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl, D.d * D.Dr, D.Dl * D.Ml);
  void* mws = sp ? p : nullptr;
  const int64_t nA = D.Dl * D.d * D.Dr, nL = D.Dl * D.Ml * D.Dl, nR = D.Dr * D.Mr * D.Dr;
  int st;
  const void* a1 = a;
  int ta1 = ta;
  if (a64b && ta == PTN_T_C128) {
    void* a64 = (char*)ws + (wtot - spb - a64b);
    if ((st = convert_c128_c64(batch * nA, a, a64, s))) return st;
    a1 = a64;
    ta1 = PTN_T_C64;
  }
  ZGemm g1 = {D.Dl * D.d, D.Mr * D.Dr, D.Dr,
              a1, D.Dr, 1, 0,
              r, D.Mr * D.Dr, 1, 0,
              t1, D.Mr * D.Dr, 1, 0, batch, nA, nR, n1, sws, spb};
  if ((st = zgemm(g1, cp, ta1, tr, ct, mws, mb, s))) return st;
  if ((st = mode_mix(D.Dl, D.d * D.Mr, D.Ml * D.d, D.Dr, w, t1, ct, t2, ct, s, batch,
                     D.Ml * D.d * D.d * D.Mr)))
    return st;
  ZGemm g3 = {D.Dl, D.d * D.Dr, D.Dl * D.Ml,
              l, 1, D.Dl, 0,
              t2, D.d * D.Dr, 1, 0,
              out, D.d * D.Dr, 1, 0, batch, nL, n2, nA, sws, spb};
  return zgemm(g3, cp, tl, ct, tout, mws, mb, s);
}

int env_right(int prec, const Dims& D, const void* a, const void* b, const void* w, const void* r,
              int tr, void* out, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch) {
  if (!dims_ok(D) || batch < 1) return PTN_ERR_ARG;
  const size_t wtot = env_ws_bytes(prec, D, batch);
  if (ws_bytes < wtot) return PTN_ERR_WORKSPACE;
  const size_t spb = split2(prec, batch, D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl * D.Ml, D.Dl,
                            D.d * D.Dr);
  void* sws = spb ? (char*)ws + (wtot - spb) : nullptr;
  const bool sp = is_sparse(prec) && batch == 1;
  prec = base_prec(prec);
  const int ct = ctype(prec), cp = gemm_compute(prec);
  const int64_t n1 = (D.Dl * D.d) * (D.Mr * D.Dr), n2 = (D.Dl * D.Ml) * (D.d * D.Dr);
  size_t ta = (size_t)batch * n1 * csize(prec);
  size_t tb = (size_t)batch * n2 * csize(prec);
  char* p = (char*)ws;
  void* t1 = p;
  p += align256(ta > tb ? ta : tb);
  void* t2 = p;
  const size_t wtb = align256((size_t)batch * (size_t)(D.Ml * D.d * D.d * D.Mr) * 16);
  const size_t mb = mask_bytes2(D.Dl * D.d, D.Mr * D.Dr, D.Dr, D.Dl * D.Ml, D.Dl, D.d * D.Dr);
  void* mws = sp ? (char*)ws + 2 * align256(ta > tb ? ta : tb) + wtb : nullptr;
  const int64_t nA = D.Dl * D.d * D.Dr, nR = D.Dr * D.Mr * D.Dr, nO = D.Dl * D.Ml * D.Dl;
  int st;
  ZGemm g1 = {D.Dl * D.d, D.Mr * D.Dr, D.Dr,
              a, D.Dr, 1, 0,
              r, D.Mr * D.Dr, 1, 0,
              t1, D.Mr * D.Dr, 1, 0, batch, nA, nR, n1, sws, spb};
  if ((st = zgemm(g1, cp, PTN_T_C128, tr, ct, mws, mb, s))) return st;
  if ((st = mode_mix(D.Dl, D.d * D.Mr, D.Ml * D.d, D.Dr, w, t1, ct, t2, ct, s, batch,
                     D.Ml * D.d * D.d * D.Mr)))
    return st;
  // out[(l, m), lb] = sum_{(sb, rb)} T2[(l, m), (sb, rb)] conj(B[lb, (sb, rb)])
  ZGemm g3 = {D.Dl * D.Ml, D.Dl, D.d * D.Dr,
              t2, D.d * D.Dr, 1, 0,
              b, 1, D.d * D.Dr, 1,
              out, D.Dl, 1, 0, batch, n2, nA, nO, sws, spb};
  return zgemm(g3, cp, ct, PTN_T_C128, PTN_T_C128, mws, mb, s);
}

int env_left(int prec, const Dims& D, const void* a, const void* b, const void* w, const void* l,
             int tl, void* out, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch) {
  if (!dims_ok(D) || batch < 1) return PTN_ERR_ARG;
  const size_t wtot = env_ws_bytes(prec, D, batch);
  if (ws_bytes < wtot) return PTN_ERR_WORKSPACE;
  const size_t spb = split2(prec, batch, D.Dl * D.Ml, D.d * D.Dr, D.Dl, D.Dr, D.Mr * D.Dr,
                            D.Dl * D.d);
  void* sws = spb ? (char*)ws + (wtot - spb) : nullptr;
  const bool sp = is_sparse(prec) && batch == 1;
  prec = base_prec(prec);
  const int ct = ctype(prec), cp = gemm_compute(prec);
  const int64_t n1 = (D.Dl * D.Ml) * (D.d * D.Dr), n2 = (D.Dl * D.d) * (D.Mr * D.Dr);
  size_t ta = (size_t)batch * n2 * csize(prec);
  size_t tb = (size_t)batch * n1 * csize(prec);
  size_t big = align256(ta > tb ? ta : tb);
  char* p = (char*)ws;
  void* t1 = p;
  void* t2 = p + big;
  void* wt = p + 2 * big;
  const int64_t nW = D.Ml * D.d * D.d * D.Mr;
  const size_t wtb = align256((size_t)batch * (size_t)nW * 16);
  const size_t mb = mask_bytes2(D.Dl * D.Ml, D.d * D.Dr, D.Dl, D.Dr, D.Mr * D.Dr, D.Dl * D.d);
  void* mws = sp ? p + 2 * big + wtb : nullptr;
  int64_t rows = D.Ml * D.d, cols = D.d * D.Mr;
  transpose_w_kernel<<<dim3((unsigned)((rows * cols + 255) / 256), (unsigned)batch), 256, 0, s>>>(
      rows, cols, (const double2*)w, (double2*)wt);
  PTN_LAUNCH_CHECK();
  const int64_t nA = D.Dl * D.d * D.Dr, nL = D.Dl * D.Ml * D.Dl, nO = D.Dr * D.Mr * D.Dr;
  int st;
  // T1[(l, m), (sb, rb)] = sum_lb L[(l, m), lb] conj(B[lb, (sb, rb)])
  ZGemm g1 = {D.Dl * D.Ml, D.d * D.Dr, D.Dl,
              l, D.Dl, 1, 0,
              b, D.d * D.Dr, 1, 1,
              t1, D.d * D.Dr, 1, 0, batch, nL, nA, n1, sws, spb};
  if ((st = zgemm(g1, cp, tl, PTN_T_C128, ct, mws, mb, s))) return st;
  // T2[l, (s, m'), rb] = sum_{(m, sb)} W[m, sb, s, m'] T1[l, (m, sb), rb]
  if ((st = mode_mix(D.Dl, D.Ml * D.d, D.d * D.Mr, D.Dr, wt, t1, ct, t2, ct, s, batch, nW)))
    return st;
  // out[r, (m', rb)] = sum_{(l, s)} A[(l, s), r] T2[(l, s), (m', rb)]
  ZGemm g3 = {D.Dr, D.Mr * D.Dr, D.Dl * D.d,
              a, 1, D.Dr, 0,
              t2, D.Mr * D.Dr, 1, 0,
              out, D.Mr * D.Dr, 1, 0, batch, nA, n2, nO, sws, spb};
  return zgemm(g3, cp, PTN_T_C128, ct, PTN_T_C128, mws, mb, s);
}

int bond_apply(int prec, int64_t Dl, int64_t Dr, int64_t M, const void* l, int tl, const void* r,
               int tr, const void* c, int tcin, void* out, int tout, void* ws, size_t ws_bytes,
               hipStream_t s, int64_t batch) {
  if (Dl < 0 || Dr < 0 || M <= 0 || batch < 1) return PTN_ERR_ARG;
  const size_t wtot = bond_ws_bytes(prec, Dl, Dr, M, batch);
  if (ws_bytes < wtot) return PTN_ERR_WORKSPACE;
  const size_t spb = split2(prec, batch, Dl, M * Dr, Dr, Dl, Dr, Dl * M);
  void* sws = spb ? (char*)ws + (wtot - spb) : nullptr;
  const bool sp = is_sparse(prec) && batch == 1;
  prec = base_prec(prec);
  const int ct = ctype(prec), cp = gemm_compute(prec);
  void* t = ws;
  const size_t mb = mask_bytes2(Dl, M * Dr, Dr, Dl, Dr, Dl * M);
  void* mws = sp ? (char*)ws + align256((size_t)Dl * (size_t)(M * Dr) * csize(prec)) : nullptr;
  const int64_t nT = Dl * M * Dr, nC = Dl * Dr, nL = Dl * M * Dl, nR = Dr * M * Dr;
  int st;
  // T[l, (m, rb)] = sum_r C[l, r] R[r, (m, rb)]
  ZGemm g1 = {Dl, M * Dr, Dr,
              c, Dr, 1, 0,
              r, M * Dr, 1, 0,
              t, M * Dr, 1, 0, batch, nC, nR, nT, sws, spb};
  if ((st = zgemm(g1, cp, tcin, tr, ct, mws, mb, s))) return st;
  // out[lb, rb] = sum_{(l, m)} L[(l, m), lb] T[(l, m), rb]
  ZGemm g3 = {Dl, Dr, Dl * M,
              l, 1, Dl, 0,
              t, Dr, 1, 0,
              out, Dr, 1, 0, batch, nL, nT, nC, sws, spb};
  return zgemm(g3, cp, tl, ct, tout, mws, mb, s);
}

}  // namespace ptn
