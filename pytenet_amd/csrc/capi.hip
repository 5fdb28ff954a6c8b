// extern "C" entry points (declared in include/ptn.h) and the small
// elementwise kernels that do not belong to a larger file.
#include <stdio.h>
#include "common.h"
#include "ops.h"

namespace ptn {

size_t krylov_ws_bytes(const ptn_local_op* op, int numiter, int64_t batch);
int krylov_run(const ptn_local_op* op, int mode, const void* v, double dt_re, double dt_im,
               int numiter, void* out, double* energy, int32_t* info, double* alpha_beta,
               void* ws, size_t ws_bytes, hipStream_t s, int64_t batch);
size_t svd_split_ws_bytes(int64_t m, int64_t n, int distr, bool precond32);
int svd_split_run(int64_t m, int64_t n, const void* theta, int64_t lda, int distr, double tol,
                  int64_t max_bond, bool precond32, void* a0, void* a1, double* s_host,
                  int64_t* k_out, int32_t* info_host, void* ws, size_t ws_bytes, hipStream_t st);
int krylov_expm_noalloc(const ptn_local_op* op, const void* v, double dt_re, double dt_im,
                        int numiter, void* out, int32_t* info, double* alpha_beta, hipStream_t s);
bool krylov_graphs(int on);
int64_t krylov_graph_launches();
size_t qr_ws_bytes(int64_t m, int64_t n, int64_t batch);
int qr_run(int64_t m, int64_t n, const void* a, int64_t lda, void* q, int64_t ldq, void* r,
           int64_t ldr, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch);
size_t svd_ws_bytes(int64_t m, int64_t n, int64_t batch, bool precond32);
int svd_run(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu, double* s,
            void* vh, int64_t ldvh, int32_t* info, void* ws, size_t ws_bytes, hipStream_t st,
            int64_t batch, int64_t sa, int64_t su, int64_t ss, int64_t svh, bool precond32);

static thread_local hipError_t g_last_hip = hipSuccess;
static bool g_autotune = false;
bool autotune_enabled() { return g_autotune; }
void ptn_set_last_hip_error(hipError_t e) { g_last_hip = e; }
bool& noalloc_mode() {
  static thread_local bool f = false;
  return f;
}
SkipFlag& skip_flag() {
  static thread_local SkipFlag f;
  return f;
}

// a (n0, n1, n2) *= s along axis 0 (s[i]) or axis 2 (s[k])
__global__ void k_scale_axis(int64_t n0, int64_t n1, int64_t n2, double2* __restrict__ a,
                             const double* __restrict__ s, int axis) {
  const int64_t total = n0 * n1 * n2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double f = axis == 0 ? s[t / (n1 * n2)] : s[t % n2];
    double2 v = a[t];
    a[t] = make_double2(v.x * f, v.y * f);
  }
}

// max |a| over entries whose three quantum numbers do not sum to zero
// (optionally zeroing them); per-block maxima combined by an atomic max on the
// bit pattern, which orders like the value for non-negative doubles
__global__ void k_qsparse3(int64_t n0, int64_t n1, int64_t n2, const int64_t* __restrict__ q0,
                           const int64_t* __restrict__ q1, const int64_t* __restrict__ q2,
                           double2* __restrict__ a, int zero_them,
                           unsigned long long* __restrict__ out) {
  __shared__ double red[256];
  double m = 0;
  const int64_t total = n0 * n1 * n2;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / (n1 * n2), j = (t / n2) % n1, k = t % n2;
    if (q0[i] + q1[j] + q2[k] != 0) {
      double2 v = a[t];
      m = fmax(m, fmax(fabs(v.x), fabs(v.y)));
      if (zero_them) a[t] = make_double2(0.0, 0.0);
    }
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

}  // namespace ptn

using namespace ptn;

static inline bool precision_ok(int p) {
  p &= ~PTN_BLOCK_SPARSE;
  return p == PTN_FP64 || p == PTN_FP32 || p == PTN_BF16X3 || p == PTN_BF16X6;
}

extern "C" {

int ptn_version(void) { return 100; }

const char* ptn_status_string(int status) {
  switch (status) {
    case PTN_OK: return "ok";
    case PTN_ERR_ARG: return "invalid shape or argument";
    case PTN_ERR_WORKSPACE: return "workspace too small";
    case PTN_ERR_HIP: return "HIP runtime error";
    case PTN_ERR_UNSUPPORTED: return "unsupported precision or shape";
    case PTN_ERR_NONFINITE: return "non-finite singular value";
  }
  return "unknown status";
}

const char* ptn_last_hip_error(void) { return hipGetErrorString(g_last_hip); }

// In the FP32 and split-bf16 modes the environment blocks are converted to
// complex64 once per call (one pass), as the Krylov solver does once per
// local solve: the products then stream half the operand bytes.
static inline bool lowp(int p) { return (p & ~PTN_BLOCK_SPARSE) != PTN_FP64; }
static inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
static inline size_t env64_bytes(int64_t D, int64_t M) { return al256((size_t)(D * M * D) * 8); }

size_t ptn_heff_workspace(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml,
                          int64_t Mr) {
  Dims D = {Dl, d, Dr, Ml, Mr};
  size_t b = al256(heff_ws_bytes(precision, D));
  if (lowp(precision)) b += env64_bytes(Dl, Ml) + env64_bytes(Dr, Mr);
  return b;
}

int ptn_heff_apply(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                   const void* l, const void* w, const void* r, const void* a, void* out, void* ws,
                   size_t ws_bytes, void* stream) {
  if (!precision_ok(precision)) return PTN_ERR_UNSUPPORTED;
  Dims D = {Dl, d, Dr, Ml, Mr};
  if (Dl < 0 || d <= 0 || Dr < 0 || Ml <= 0 || Mr <= 0) return PTN_ERR_ARG;
  if (ws_bytes < ptn_heff_workspace(precision, Dl, d, Dr, Ml, Mr)) return PTN_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const size_t base = al256(heff_ws_bytes(precision, D));
  if (!lowp(precision))
    return heff_apply(precision, D, l, PTN_T_C128, w, r, PTN_T_C128, a, PTN_T_C128, out,
                      PTN_T_C128, ws, base, s);
  char* l64 = (char*)ws + base;
  char* r64 = l64 + env64_bytes(Dl, Ml);
  int st;
  if ((st = convert_c128_c64(Dl * Ml * Dl, l, l64, s))) return st;
  if ((st = convert_c128_c64(Dr * Mr * Dr, r, r64, s))) return st;
  return heff_apply(precision, D, l64, PTN_T_C64, w, r64, PTN_T_C64, a, PTN_T_C128, out,
                    PTN_T_C128, ws, base, s);
}

size_t ptn_bond_workspace(int precision, int64_t Dl, int64_t Dr, int64_t M) {
  size_t b = al256(bond_ws_bytes(precision, Dl, Dr, M));
  if (lowp(precision)) b += env64_bytes(Dl, M) + env64_bytes(Dr, M);
  return b;
}

int ptn_bond_apply(int precision, int64_t Dl, int64_t Dr, int64_t M, const void* l, const void* r,
                   const void* c, void* out, void* ws, size_t ws_bytes, void* stream) {
  if (!precision_ok(precision)) return PTN_ERR_UNSUPPORTED;
  if (Dl < 0 || Dr < 0 || M <= 0) return PTN_ERR_ARG;
  if (ws_bytes < ptn_bond_workspace(precision, Dl, Dr, M)) return PTN_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const size_t base = al256(bond_ws_bytes(precision, Dl, Dr, M));
  if (!lowp(precision))
    return bond_apply(precision, Dl, Dr, M, l, PTN_T_C128, r, PTN_T_C128, c, PTN_T_C128, out,
                      PTN_T_C128, ws, base, s);
  char* l64 = (char*)ws + base;
  char* r64 = l64 + env64_bytes(Dl, M);
  int st;
  if ((st = convert_c128_c64(Dl * M * Dl, l, l64, s))) return st;
  if ((st = convert_c128_c64(Dr * M * Dr, r, r64, s))) return st;
  return bond_apply(precision, Dl, Dr, M, l64, PTN_T_C64, r64, PTN_T_C64, c, PTN_T_C128, out,
                    PTN_T_C128, ws, base, s);
}

int ptn_local_apply(const ptn_local_op* op, const void* a, void* out, int32_t* grouped, void* ws,
                    size_t ws_bytes, void* stream) {
  if (grouped) *grouped = 0;
  if (!op) return PTN_ERR_ARG;
  if (op->kind == PTN_OP_BOND) {
    if (op->Ml != op->Mr) return PTN_ERR_ARG;
    return ptn_bond_apply(op->precision, op->Dl, op->Dr, op->Ml, op->l, op->r, a, out, ws,
                          ws_bytes, stream);
  }
  if (op->kind != PTN_OP_HEFF) return PTN_ERR_ARG;
  SectorPlanRef spr;
  int st;
  if ((st = sector_plan_get(op, SECTORS_HEFF, &spr))) return st;
  const SectorPlan* sp = spr.get();
  if (!sp)
    return ptn_heff_apply(op->precision, op->Dl, op->d, op->Dr, op->Ml, op->Mr, op->l, op->w,
                          op->r, a, out, ws, ws_bytes, stream);
  // a plan exists only for valid shapes with Dl, Dr >= 1
  const Dims D = {op->Dl, op->d, op->Dr, op->Ml, op->Mr};
  if (op->d <= 0 || op->d * op->Mr > 32 || op->Ml * op->d > 32) return PTN_ERR_ARG;
  if (ws_bytes < ptn_heff_workspace(op->precision, op->Dl, op->d, op->Dr, op->Ml, op->Mr))
    return PTN_ERR_WORKSPACE;
  hipStream_t s = (hipStream_t)stream;
  const size_t base = al256(heff_ws_bytes(op->precision, D));
  const void* l = op->l;
  const void* r = op->r;
  int tl = PTN_T_C128;
  if (lowp(op->precision)) {
    char* l64 = (char*)ws + base;
    char* r64 = l64 + env64_bytes(op->Dl, op->Ml);
    if ((st = convert_c128_c64(op->Dl * op->Ml * op->Dl, op->l, l64, s))) return st;
    if ((st = convert_c128_c64(op->Dr * op->Mr * op->Dr, op->r, r64, s))) return st;
    l = l64;
    r = r64;
    tl = PTN_T_C64;
  }
  if ((st = heff_apply_sectors(op->precision, D, sp, l, tl, op->w, r, tl, a, out, ws, base, s)))
    return st;
  if (grouped) *grouped = 1;
  return PTN_OK;
}

static size_t env_workspace_b(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                              int64_t Ml, int64_t Mr) {
  Dims D = {Dl, d, Dr, Ml, Mr};
  size_t b = al256(env_ws_bytes(precision, D, batch));
  if (lowp(precision)) {
    const size_t el = env64_bytes(Dl, Ml), er = env64_bytes(Dr, Mr);
    b += (size_t)batch * (el > er ? el : er);
  }
  return b;
}

static int env_step_b(int left, int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                      int64_t Ml, int64_t Mr, const void* a, const void* b, const void* w,
                      const void* blk, void* out, void* ws, size_t ws_bytes, hipStream_t s) {
  if (!precision_ok(precision)) return PTN_ERR_UNSUPPORTED;
  Dims D = {Dl, d, Dr, Ml, Mr};
  if (Dl < 0 || d <= 0 || Dr < 0 || Ml <= 0 || Mr <= 0 || batch < 1) return PTN_ERR_ARG;
  if (ws_bytes < env_workspace_b(precision, batch, Dl, d, Dr, Ml, Mr)) return PTN_ERR_WORKSPACE;
  const size_t base = al256(env_ws_bytes(precision, D, batch));
  const void* e = blk;
  int te = PTN_T_C128, st;
  if (lowp(precision)) {
    // the environment block stack in complex64, converted once (contiguous)
    char* e64 = (char*)ws + base;
    const int64_t ne = left ? Dl * Ml * Dl : Dr * Mr * Dr;
    if ((st = convert_c128_c64(batch * ne, blk, e64, s))) return st;
    e = e64;
    te = PTN_T_C64;
  }
  if (left) return env_left(precision, D, a, b, w, e, te, out, ws, base, s, batch);
  return env_right(precision, D, a, b, w, e, te, out, ws, base, s, batch);
}

int ptn_env_apply(const ptn_local_op* op, int side, const void* a, const void* b, void* out,
                  int32_t* grouped, void* ws, size_t ws_bytes, void* stream) {
  if (grouped) *grouped = 0;
  if (!op || op->kind != PTN_OP_HEFF || (side != PTN_ENV_LEFT && side != PTN_ENV_RIGHT))
    return PTN_ERR_ARG;
  const bool left = side == PTN_ENV_LEFT;
  SectorPlanRef spr;
  int st;
  if ((st = sector_plan_get(op, left ? SECTORS_ENV_LEFT : SECTORS_ENV_RIGHT, &spr))) return st;
  const SectorPlan* sp = spr.get();
  hipStream_t s = (hipStream_t)stream;
  if (!sp)
    return env_step_b(left ? 1 : 0, op->precision, 1, op->Dl, op->d, op->Dr, op->Ml, op->Mr, a,
                      b, op->w, left ? op->l : op->r, out, ws, ws_bytes, s);
  const Dims D = {op->Dl, op->d, op->Dr, op->Ml, op->Mr};
  if (ws_bytes < env_workspace_b(op->precision, 1, op->Dl, op->d, op->Dr, op->Ml, op->Mr))
    return PTN_ERR_WORKSPACE;
  const size_t base = al256(env_ws_bytes(op->precision, D));
  const void* e = left ? op->l : op->r;
  int te = PTN_T_C128;
  if (lowp(op->precision)) {
    char* e64 = (char*)ws + base;
    const int64_t ne = left ? op->Dl * op->Ml * op->Dl : op->Dr * op->Mr * op->Dr;
    if ((st = convert_c128_c64(ne, e, e64, s))) return st;
    e = e64;
    te = PTN_T_C64;
  }
  st = left ? env_left_sectors(op->precision, D, sp, a, b, op->w, e, te, out, ws, base, s)
            : env_right_sectors(op->precision, D, sp, a, b, op->w, e, te, out, ws, base, s);
  if (st) return st;
  if (grouped) *grouped = 1;
  return PTN_OK;
}

size_t ptn_env_workspace(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml,
                         int64_t Mr) {
  return env_workspace_b(precision, 1, Dl, d, Dr, Ml, Mr);
}

int ptn_env_left(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                 const void* a, const void* b, const void* w, const void* l, void* out, void* ws,
                 size_t ws_bytes, void* stream) {
  return env_step_b(1, precision, 1, Dl, d, Dr, Ml, Mr, a, b, w, l, out, ws, ws_bytes,
                    (hipStream_t)stream);
}

int ptn_env_right(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                  const void* a, const void* b, const void* w, const void* r, void* out, void* ws,
                  size_t ws_bytes, void* stream) {
  return env_step_b(0, precision, 1, Dl, d, Dr, Ml, Mr, a, b, w, r, out, ws, ws_bytes,
                    (hipStream_t)stream);
}

size_t ptn_env_workspace_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                                 int64_t Ml, int64_t Mr) {
  return env_workspace_b(precision, batch, Dl, d, Dr, Ml, Mr);
}

int ptn_env_left_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                         int64_t Ml, int64_t Mr, const void* a, const void* b, const void* w,
                         const void* l, void* out, void* ws, size_t ws_bytes, void* stream) {
  return env_step_b(1, precision, batch, Dl, d, Dr, Ml, Mr, a, b, w, l, out, ws, ws_bytes,
                    (hipStream_t)stream);
}

int ptn_env_right_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                          int64_t Ml, int64_t Mr, const void* a, const void* b, const void* w,
                          const void* r, void* out, void* ws, size_t ws_bytes, void* stream) {
  return env_step_b(0, precision, batch, Dl, d, Dr, Ml, Mr, a, b, w, r, out, ws, ws_bytes,
                    (hipStream_t)stream);
}

int ptn_set_graphs(int on) { return krylov_graphs(on) ? 1 : 0; }

int64_t ptn_graph_launches(void) { return krylov_graph_launches(); }

int ptn_set_autotune(int on) {
  const int prev = g_autotune ? 1 : 0;
  g_autotune = on != 0;
  return prev;
}

size_t ptn_zgemm_workspace(int64_t M, int64_t N, int64_t K) {
  (void)M; (void)N; (void)K;
  return 0;
}

int ptn_zgemm(int64_t M, int64_t N, int64_t K, const void* A, int64_t a_m, int64_t a_k, int conjA,
              const void* B, int64_t b_k, int64_t b_n, int conjB, void* C, int64_t c_m,
              int64_t c_n, int accumulate, void* ws, size_t ws_bytes, void* stream) {
  ZGemm g = {M, N, K, A, a_m, a_k, conjA, B, b_k, b_n, conjB, C, c_m, c_n, accumulate, 1, 0, 0, 0};
  (void)ws;   // the dense product needs no workspace (ptn_zgemm_workspace is 0)
  (void)ws_bytes;
  return zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, (hipStream_t)stream);
}

size_t ptn_krylov_workspace(const ptn_local_op* op, int numiter) {
  return krylov_ws_bytes(op, numiter, 1);
}

int ptn_krylov_expm(const ptn_local_op* op, const void* v, double dt_re, double dt_im, int numiter,
                    void* out, int32_t* info, double* alpha_beta, void* ws, size_t ws_bytes,
                    void* stream) {
  return krylov_run(op, 0, v, dt_re, dt_im, numiter, out, nullptr, info, alpha_beta, ws, ws_bytes,
                    (hipStream_t)stream, 1);
}

int ptn_krylov_expm_noalloc(const ptn_local_op* op, const void* v, double dt_re, double dt_im,
                            int numiter, void* out, int32_t* info, double* alpha_beta,
                            void* stream) {
  return krylov_expm_noalloc(op, v, dt_re, dt_im, numiter, out, info, alpha_beta,
                             (hipStream_t)stream);
}

int ptn_krylov_eigh(const ptn_local_op* op, const void* v, int numiter, void* out, double* energy,
                    int32_t* info, double* alpha_beta, void* ws, size_t ws_bytes, void* stream) {
  return krylov_run(op, 1, v, 0.0, 0.0, numiter, out, energy, info, alpha_beta, ws, ws_bytes,
                    (hipStream_t)stream, 1);
}

size_t ptn_qr_workspace(int64_t m, int64_t n) { return qr_ws_bytes(m, n, 1); }

int ptn_qr(int64_t m, int64_t n, const void* a, int64_t lda, void* q, int64_t ldq, void* r,
           int64_t ldr, void* ws, size_t ws_bytes, void* stream) {
  return qr_run(m, n, a, lda, q, ldq, r, ldr, ws, ws_bytes, (hipStream_t)stream, 1);
}

size_t ptn_svd_workspace(int64_t m, int64_t n) { return svd_ws_bytes(m, n, 1, false); }

int ptn_svd(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu, double* s,
            void* vh, int64_t ldvh, int32_t* info, void* ws, size_t ws_bytes, void* stream) {
  return svd_run(m, n, a, lda, u, ldu, s, vh, ldvh, info, ws, ws_bytes, (hipStream_t)stream, 1,
                 0, 0, 0, 0, false);
}

size_t ptn_svd_split_workspace(int64_t m, int64_t n, int distr, int precond32) {
  return svd_split_ws_bytes(m, n, distr, precond32 != 0);
}

int ptn_svd_split(int64_t m, int64_t n, const void* theta, int64_t lda, int distr, double tol,
                  int64_t max_bond, int precond32, void* a0, void* a1, double* s_host,
                  int64_t* k_out, int32_t* info_host, void* ws, size_t ws_bytes, void* stream) {
  return svd_split_run(m, n, theta, lda, distr, tol, max_bond, precond32 != 0, a0, a1, s_host,
                       k_out, info_host, ws, ws_bytes, (hipStream_t)stream);
}

size_t ptn_svd_workspace_batched(int64_t batch, int64_t m, int64_t n) {
  return svd_ws_bytes(m, n, batch, false);
}

int ptn_svd_batched(int64_t batch, int64_t m, int64_t n, const void* a, void* u, double* s,
                    void* vh, int32_t* info, void* ws, size_t ws_bytes, void* stream) {
  const int64_t k = m < n ? m : n;
  return svd_run(m, n, a, n, u, k, s, vh, n, info, ws, ws_bytes, (hipStream_t)stream, batch,
                 m * n, m * k, k, k * n, false);
}

size_t ptn_svd_workspace_precond32(int64_t batch, int64_t m, int64_t n) {
  return svd_ws_bytes(m, n, batch, true);
}

int ptn_svd_precond32(int64_t batch, int64_t m, int64_t n, const void* a, void* u, double* s,
                      void* vh, int32_t* info, void* ws, size_t ws_bytes, void* stream) {
  const int64_t k = m < n ? m : n;
  return svd_run(m, n, a, n, u, k, s, vh, n, info, ws, ws_bytes, (hipStream_t)stream, batch,
                 m * n, m * k, k, k * n, true);
}

int ptn_zgemm_batched(int64_t batch, int64_t M, int64_t N, int64_t K, const void* A, int64_t a_m,
                      int64_t a_k, int64_t bz_a, int conjA, const void* B, int64_t b_k,
                      int64_t b_n, int64_t bz_b, int conjB, void* C, int64_t c_m, int64_t c_n,
                      int64_t bz_c, int accumulate, void* stream) {
  if (batch < 1 || batch > 65535) return PTN_ERR_ARG;
  ZGemm g = {M, N, K, A, a_m, a_k, conjA, B, b_k, b_n, conjB, C, c_m, c_n, accumulate,
             batch, bz_a, bz_b, bz_c};
  return zgemm(g, PTN_FP64, PTN_T_C128, PTN_T_C128, PTN_T_C128, nullptr, 0, (hipStream_t)stream);
}

size_t ptn_krylov_workspace_batched(const ptn_local_op* op, int numiter, int64_t batch) {
  return krylov_ws_bytes(op, numiter, batch);
}

int ptn_krylov_expm_batched(const ptn_local_op* op, int64_t batch, const void* v, double dt_re,
                            double dt_im, int numiter, void* out, int32_t* info, void* ws,
                            size_t ws_bytes, void* stream) {
  return krylov_run(op, 0, v, dt_re, dt_im, numiter, out, nullptr, info, nullptr, ws, ws_bytes,
                    (hipStream_t)stream, batch);
}

int ptn_krylov_eigh_batched(const ptn_local_op* op, int64_t batch, const void* v, int numiter,
                            void* out, double* energy, int32_t* info, void* ws, size_t ws_bytes,
                            void* stream) {
  return krylov_run(op, 1, v, 0.0, 0.0, numiter, out, energy, info, nullptr, ws, ws_bytes,
                    (hipStream_t)stream, batch);
}

size_t ptn_qr_workspace_batched(int64_t batch, int64_t m, int64_t n) {
  return qr_ws_bytes(m, n, batch);
}

int ptn_qr_batched(int64_t batch, int64_t m, int64_t n, const void* a, void* q, void* r, void* ws,
                   size_t ws_bytes, void* stream) {
  const int64_t k = m < n ? m : n;
  return qr_run(m, n, a, n, q, k, r, n, ws, ws_bytes, (hipStream_t)stream, batch);
}

int ptn_scale_axis(int64_t n0, int64_t n1, int64_t n2, void* a, const double* s, int axis,
                   void* stream) {
  if (n0 < 0 || n1 < 0 || n2 < 0 || (axis != 0 && axis != 2)) return PTN_ERR_ARG;
  const int64_t total = n0 * n1 * n2;
  if (total == 0) return PTN_OK;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_scale_axis<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(n0, n1, n2, (double2*)a, s,
                                                                   axis);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

int ptn_qsparse3(int64_t n0, int64_t n1, int64_t n2, const int64_t* q0, const int64_t* q1,
                 const int64_t* q2, void* a, int zero_them, double* out, void* stream) {
  if (n0 < 0 || n1 < 0 || n2 < 0) return PTN_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  k_zero_u64<<<1, 1, 0, s>>>((unsigned long long*)out);
  const int64_t total = n0 * n1 * n2;
  if (total > 0) {
    int64_t blocks = (total + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    k_qsparse3<<<(unsigned)blocks, 256, 0, s>>>(n0, n1, n2, q0, q1, q2, (double2*)a, zero_them,
                                                (unsigned long long*)out);
  }
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

}  // extern "C"
