// Device-resident Lanczos for the local TDVP/DMRG problems.
//
// Follows pytenet/krylov.py:12-57 (plain three-term recurrence, no
// reorthogonalisation, breakdown when beta_j < 100*n*eps), krylov.py:142-150
// (eigendecomposition of the tridiagonal matrix) and krylov.py:110-136
// (Ritz vector / Krylov exponential). All scalars (alpha, beta, norms, the
// tridiagonal eigensystem and the combination coefficients) live in device
// memory and are FP64, so a whole local solve is enqueued without a host
// round trip. A breakdown or a non-finite value stops the recurrence through a
// flag that later kernels read; the host reads the status words afterwards.
#include <float.h>
#include <map>
#include <mutex>
#include <unordered_map>
#include "common.h"
#include "ops.h"

namespace ptn {

constexpr int LZ_MAX = 64;        // largest supported numiter
constexpr int RED_BLOCKS = 512;   // largest reduction width (partial sums per reduction)

struct LzState {
  double alpha[LZ_MAX];
  double beta[LZ_MAX];
  double coef_re[LZ_MAX];
  double coef_im[LZ_MAX];
  double nrm0;
  double energy;
  int nvalid;      // Lanczos vectors in use
  int breakdown;   // j + 1 of a beta breakdown, 0 if none
  int nonfinite;   // a norm / inner product was not finite
  int zero_start;  // the start vector had norm zero
  int stopped;     // recurrence stopped (breakdown, non-finite or zero start)
  int noconv;      // tridiagonal QL iteration did not converge
};

// Batched form: blockIdx.y selects the problem of a batch; vector arrays are
// stacked with a per-problem stride, LzState and partial-sum arrays are
// indexed by problem. `rb` partial sums per reduction (rb = min(512,
// ceil(n / 256)), a function of n only, so every sum has a fixed order).

// sum of the rb partials in a fixed order; every block of a kernel computes
// the identical value (256 threads)
__device__ double block_sum_partials(const double* __restrict__ part, int rb) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double v = 0;
  for (int i = t; i < rb; i += 256) v += part[i];
  red[t] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  double s = red[0];
  __syncthreads();
  return s;
}

__device__ void block_store_partial(double v, double* out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_state_init(LzState* st, int numiter) {
  st += blockIdx.x;
  int t = threadIdx.x;
  if (t < LZ_MAX) {
    st->alpha[t] = 0; st->beta[t] = 0; st->coef_re[t] = 0; st->coef_im[t] = 0;
  }
  if (t == 0) {
    st->nrm0 = 0; st->energy = 0; st->nvalid = numiter; st->breakdown = 0;
    st->nonfinite = 0; st->zero_start = 0; st->stopped = 0; st->noconv = 0;
  }
}

__global__ __launch_bounds__(256) void k_sumsq_partial(int64_t n, const double2* __restrict__ x,
                                                       int64_t xs, double* __restrict__ part,
                                                       int rb) {
  x += blockIdx.y * xs;
  part += blockIdx.y * rb;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)rb * 256) {
    double2 v = x[i];
    s += v.x * v.x + v.y * v.y;
  }
  block_store_partial(s, part);
}

// Re <w, v> partials
__global__ __launch_bounds__(256) void k_dot_partial(int64_t n, const double2* __restrict__ w,
                                                     int64_t ws_, const double2* __restrict__ v,
                                                     int64_t vs, double* __restrict__ part,
                                                     int rb) {
  w += blockIdx.y * ws_;
  v += blockIdx.y * vs;
  part += blockIdx.y * rb;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)rb * 256) {
    double2 a = w[i], b = v[i];
    s += a.x * b.x + a.y * b.y;
  }
  block_store_partial(s, part);
}

// v0 = x / |x|  (krylov.py:29-31)
__global__ __launch_bounds__(256) void k_start(int64_t n, const double2* __restrict__ x,
                                               int64_t xs, double2* __restrict__ v0, int64_t vs,
                                               const double* __restrict__ part, int rb,
                                               LzState* st) {
  x += blockIdx.y * xs;
  v0 += blockIdx.y * vs;
  part += blockIdx.y * rb;
  st += blockIdx.y;
  double nrm = sqrt(block_sum_partials(part, rb));
  bool bad = !(nrm > 0.0) || !isfinite(nrm);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->nrm0 = nrm;
    if (bad) {
      if (nrm == 0.0) st->zero_start = 1; else st->nonfinite = 1;
      st->stopped = 1;
      st->nvalid = 0;
    }
  }
  if (bad) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double2 a = x[i];
    v0[i] = make_double2(a.x / nrm, a.y / nrm);
  }
}

// alpha_j = Re <w, v_j>;  w -= alpha_j v_j + beta_{j-1} v_{j-1};  partials of |w|^2
__global__ __launch_bounds__(256) void k_alpha_update(int64_t n, double2* __restrict__ w,
                                                      int64_t ws_,
                                                      const double2* __restrict__ vj,
                                                      const double2* __restrict__ vjm1,
                                                      int64_t vs, int j,
                                                      const double* __restrict__ part_in,
                                                      double* __restrict__ part_out, int rb,
                                                      LzState* st) {
  st += blockIdx.y;
  if (st->stopped) return;
  w += blockIdx.y * ws_;
  vj += blockIdx.y * vs;
  vjm1 += blockIdx.y * vs;
  part_in += blockIdx.y * rb;
  part_out += blockIdx.y * rb;
  const double a = block_sum_partials(part_in, rb);
  if (!isfinite(a)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->nonfinite = 1; st->stopped = 1; st->nvalid = j + 1; st->alpha[j] = a;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha[j] = a;
  const double b = (j > 0) ? st->beta[j - 1] : 0.0;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)rb * 256) {
    double2 x = w[i], v = vj[i];
    double tr = a * v.x, ti = a * v.y;
    if (j > 0) {
      double2 u = vjm1[i];
      tr += b * u.x; ti += b * u.y;
    }
    x.x -= tr; x.y -= ti;
    w[i] = x;
    s += x.x * x.x + x.y * x.y;
  }
  block_store_partial(s, part_out);
}

// beta_j = |w|; breakdown test; v_{j+1} = w / beta_j. The test is the
// reference's absolute one (beta_j < 100 n eps) and, for matvecs in reduced
// precision (rel > 0), also beta_j < rel |A v_j| with |A v_j| =
// sqrt(alpha_j^2 + beta_{j-1}^2 + beta_j^2): a residual below the rounding
// noise of a complex64 / split-bf16 matvec is that noise, not a new Krylov
// direction (a start vector in an invariant subspace otherwise continues with
// normalised noise).
__global__ __launch_bounds__(256) void k_beta_next(int64_t n, const double2* __restrict__ w,
                                                   int64_t ws_, double2* __restrict__ vnext,
                                                   int64_t vs, int j,
                                                   const double* __restrict__ part_in, int rb,
                                                   double thresh, double rel, LzState* st) {
  st += blockIdx.y;
  if (st->stopped) return;
  w += blockIdx.y * ws_;
  vnext += blockIdx.y * vs;
  part_in += blockIdx.y * rb;
  const double b = sqrt(block_sum_partials(part_in, rb));
  const bool nf = !isfinite(b);
  const double a = st->alpha[j], bm = j > 0 ? st->beta[j - 1] : 0.0;
  const double thr = fmax(thresh, rel * sqrt(a * a + bm * bm + b * b));
  const bool brk = !nf && (b < thr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (nf) { st->nonfinite = 1; st->stopped = 1; st->nvalid = j + 1; }
    else if (brk) { st->breakdown = j + 1; st->stopped = 1; st->nvalid = j + 1; }
    else st->beta[j] = b;
  }
  if (nf || brk) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double2 x = w[i];
    vnext[i] = make_double2(x.x / b, x.y / b);
  }
}

// last alpha (krylov.py:54-56); one block per problem
__global__ __launch_bounds__(256) void k_alpha_last(int j, const double* __restrict__ part_in,
                                                    int rb, LzState* st) {
  st += blockIdx.x;
  if (st->stopped) return;
  const double a = block_sum_partials(part_in + blockIdx.x * rb, rb);
  if (threadIdx.x == 0) {
    st->alpha[j] = a;
    if (!isfinite(a)) st->nonfinite = 1;
  }
}

// Eigendecomposition of the nv x nv tridiagonal matrix by the implicit QL
// method with Wilkinson shifts (one wave: every lane runs the same scalar
// recurrence on the diagonal/off-diagonal, lane k applies the plane
// rotations to row k of the eigenvector matrix), eigenvalues sorted
// ascending as numpy.linalg.eigh returns them; then the coefficients:
//   mode 0 (expm):  c = U (|v| exp(dt lambda) * U[0, :])    (krylov.py:132-136)
//   mode 1 (eigh):  c = U[:, 0], energy = lambda_0          (krylov.py:116-119)
__global__ __launch_bounds__(64) void k_tridiag(LzState* st, int mode, double dt_re,
                                                double dt_im) {
  st += blockIdx.x;
  __shared__ double d[LZ_MAX], e[LZ_MAX];
  __shared__ double Z[LZ_MAX][LZ_MAX + 1];
  __shared__ double tre[LZ_MAX], tim[LZ_MAX];
  const int lane = threadIdx.x;
  const int n = st->nvalid;
  if (n <= 0 || st->nonfinite) return;
  d[lane] = (lane < n) ? st->alpha[lane] : 0.0;
  e[lane] = (lane < n - 1) ? st->beta[lane] : 0.0;
  for (int k = 0; k < LZ_MAX; ++k) Z[lane][k] = (lane == k) ? 1.0 : 0.0;
  __syncthreads();
  int noconv = 0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      while (m < n - 1) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
        ++m;
      }
      if (m == l) break;
      if (++iter > 100) { noconv = 1; break; }
      // shift from the leading 2 x 2 block of the unreduced part
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i];
        const double b = c * e[i];
        r = hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        const double z1 = Z[lane][i + 1], z0 = Z[lane][i];
        Z[lane][i + 1] = s * z0 + c * z1;
        Z[lane][i] = c * z0 - s * z1;
      }
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // ascending order (selection sort; the permutation is identical in every lane)
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    double dm = d[i];
    for (int j = i + 1; j < n; ++j)
      if (d[j] < dm) { dm = d[j]; k = j; }
    if (k != i) {
      const double t = d[i];
      d[k] = t;
      d[i] = dm;
      const double zt = Z[lane][i];
      Z[lane][i] = Z[lane][k];
      Z[lane][k] = zt;
    }
  }
  __syncthreads();
  if (mode == 0) {
    if (lane < n) {
      // |v| * exp(dt * lambda_j) * U[0, j]
      const double lam = d[lane];
      const double mag = exp(dt_re * lam);
      const double ph = dt_im * lam;
      const double er = st->nrm0 * (mag * cos(ph));
      const double ei = st->nrm0 * (mag * sin(ph));
      tre[lane] = er * Z[0][lane];
      tim[lane] = ei * Z[0][lane];
    }
    __syncthreads();
    if (lane < n) {
      double cr = 0, ci = 0;
      for (int j = 0; j < n; ++j) {
        cr += Z[lane][j] * tre[j];
        ci += Z[lane][j] * tim[j];
      }
      st->coef_re[lane] = cr;
      st->coef_im[lane] = ci;
    }
  } else {
    if (lane < n) {
      st->coef_re[lane] = Z[lane][0];
      st->coef_im[lane] = 0.0;
    }
    if (lane == 0) st->energy = d[0];
  }
  if (lane == 0 && noconv) st->noconv = 1;
}

// out = sum_{j < nvalid} c_j V_j  (krylov.py:118, :136)
// V_j of problem b at V + j * js + b * Vs
__global__ __launch_bounds__(256) void k_combine(int64_t n, const double2* __restrict__ V,
                                                 int64_t js, int64_t Vs,
                                                 const LzState* __restrict__ st,
                                                 double2* __restrict__ out, int64_t os) {
  __shared__ double cr[LZ_MAX], ci[LZ_MAX];
  V += blockIdx.y * Vs;
  out += blockIdx.y * os;
  st += blockIdx.y;
  const int nv = st->nonfinite ? 0 : st->nvalid;
  if (threadIdx.x < LZ_MAX) {
    cr[threadIdx.x] = st->coef_re[threadIdx.x];
    ci[threadIdx.x] = st->coef_im[threadIdx.x];
  }
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double sr = 0, si = 0;
    for (int j = 0; j < nv; ++j) {
      const double2 v = V[(int64_t)j * js + i];
      sr += cr[j] * v.x - ci[j] * v.y;
      si += cr[j] * v.y + ci[j] * v.x;
    }
    out[i] = make_double2(sr, si);
  }
}

__global__ void k_export(const LzState* __restrict__ st, int32_t* info, double* alpha_beta,
                         double* energy, int numiter) {
  const int b = blockIdx.x;
  st += b;
  if (info) info += 4 * b;
  if (energy) energy += b;
  if (alpha_beta) alpha_beta += 2 * numiter * b;
  const int t = threadIdx.x;
  if (t == 0) {
    if (info) {
      info[0] = st->nvalid;
      info[1] = st->breakdown;
      info[2] = st->nonfinite | (st->noconv << 1);
      info[3] = st->zero_start;
    }
    if (energy) *energy = st->energy;
  }
  if (alpha_beta && t < numiter) {
    alpha_beta[t] = st->alpha[t];
    alpha_beta[numiter + t] = st->beta[t];
  }
}

// ---------------------------------------------------------------------------
static inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

struct KrylovLayout {
  size_t v, w, part_a, part_b, state, l64, r64, mv, total;
};

static int64_t op_vec_len(const ptn_local_op* op) {
  return op->kind == PTN_OP_HEFF ? op->Dl * op->d * op->Dr : op->Dl * op->Dr;
}

static int red_width(int64_t n) {
  int64_t rb = (n + 255) / 256;
  if (rb > RED_BLOCKS) rb = RED_BLOCKS;
  return rb < 1 ? 1 : (int)rb;
}

// Krylov vectors are stored as [j][problem][n], so that V_j of all problems is
// one contiguous batch for the batched matvec
static KrylovLayout krylov_layout(const ptn_local_op* op, int numiter, int64_t batch) {
  KrylovLayout k;
  const int64_t n = op_vec_len(op);
  const int rb = red_width(n);
  size_t off = 0;
  k.v = off; off += al((size_t)batch * numiter * n * 16);
  k.w = off; off += al((size_t)batch * n * 16);
  k.part_a = off; off += al((size_t)batch * rb * 8);
  k.part_b = off; off += al((size_t)batch * rb * 8);
  k.state = off; off += al((size_t)batch * sizeof(LzState));
  const bool lowp = (op->precision & ~PTN_BLOCK_SPARSE) != PTN_FP64;
  const int64_t nl = op->Dl * op->Ml * op->Dl, nr = op->Dr * op->Mr * op->Dr;
  k.l64 = off; off += lowp ? al((size_t)batch * nl * 8) : 0;
  k.r64 = off; off += lowp ? al((size_t)batch * nr * 8) : 0;
  k.mv = off;
  if (op->kind == PTN_OP_HEFF) {
    Dims D = {op->Dl, op->d, op->Dr, op->Ml, op->Mr};
    off += al(heff_ws_bytes(op->precision, D, batch));
  } else {
    off += al(bond_ws_bytes(op->precision, op->Dl, op->Dr, op->Ml, batch));
  }
  k.total = off;
  return k;
}

size_t krylov_ws_bytes(const ptn_local_op* op, int numiter, int64_t batch) {
  return krylov_layout(op, numiter, batch).total;
}

// `batch` independent local problems of identical shape: l, w, r, v and out
// stacked contiguously along a leading batch axis; info / energy /
// alpha_beta per problem (4, 1, 2 numiter entries)
static int krylov_enqueue(const ptn_local_op* op, int mode, const void* v, double dt_re,
                          double dt_im, int numiter, void* out, double* energy, int32_t* info,
                          double* alpha_beta, void* ws, size_t ws_bytes, hipStream_t s,
                          int64_t batch, const SectorPlan* sp);

// ---------------------------------------------------------------------------
// hipGraph capture of the local solves. A solve whose vectors are short is
// bound by kernel-launch latency (numiter x ~7 launches); its whole launch
// sequence (state init, every Lanczos iteration with its matvec, the
// tridiagonal eigensolver, the combination) is captured into a graph on a
// private stream and launched as one graph on the caller's stream. The
// executable graphs are cached per shape bucket (operator kind and shapes,
// precision, numiter, mode, batch, device) and updated in place with the
// new call's pointers (hipGraphExecUpdate; the topology of one bucket never
// changes). Iterations after a breakdown are predicated off on the device
// (the "stopped" word), so one graph serves every outcome. Each bucket keeps a
// ring of GRAPH_RING executables; an executable is updated only after its
// previous launch has completed (an event per slot). Long vectors (MFMA-bound
// matvecs of milliseconds) are launched directly.
namespace {
bool g_graphs_on = true;
constexpr int64_t GRAPH_MAX_N = 1 << 20;   // vector length up to which solves are captured
constexpr int GRAPH_RING = 3;
constexpr size_t GRAPH_MAX_BUCKETS = 512;
constexpr int GRAPH_GROUPED = 1 << 20;   // key bit: matvecs through the sector-grouped products
struct GKey {
  int dev, kind, prec, mode, numiter;
  int64_t Dl, d, Dr, Ml, Mr, batch;
  bool operator==(const GKey& o) const {
    return dev == o.dev && kind == o.kind && prec == o.prec && mode == o.mode &&
           numiter == o.numiter && Dl == o.Dl && d == o.d && Dr == o.Dr && Ml == o.Ml &&
           Mr == o.Mr && batch == o.batch;
  }
};
struct GHash {
  size_t operator()(const GKey& k) const {
    size_t h = (size_t)k.dev * 1000003u ^ (size_t)k.kind * 10007u ^ (size_t)k.prec * 101u ^
               (size_t)k.mode * 7u ^ (size_t)k.numiter;
    for (int64_t x : {k.Dl, k.d, k.Dr, k.Ml, k.Mr, k.batch}) h = h * 1315423911u ^ (size_t)x;
    return h;
  }
};
struct GEntry {
  hipGraphExec_t exec[GRAPH_RING] = {};
  hipEvent_t done[GRAPH_RING] = {};
  int next = 0;
};
std::mutex g_gmu;
std::unordered_map<GKey, GEntry, GHash> g_graph_cache;
std::map<int, hipStream_t> g_capture_streams;
int64_t g_graph_launches = 0;

void drop_entry(GEntry& e) {
  for (int i = 0; i < GRAPH_RING; ++i) {
    if (e.done[i]) { (void)hipEventSynchronize(e.done[i]); (void)hipEventDestroy(e.done[i]); }
    if (e.exec[i]) (void)hipGraphExecDestroy(e.exec[i]);
    e.exec[i] = nullptr;
    e.done[i] = nullptr;
  }
}
}  // namespace

bool krylov_graphs(int on) {
  std::lock_guard<std::mutex> lk(g_gmu);
  const bool prev = g_graphs_on;
  if (on >= 0) g_graphs_on = on != 0;
  return prev;
}

int64_t krylov_graph_launches() {
  std::lock_guard<std::mutex> lk(g_gmu);
  return g_graph_launches;
}

int krylov_run(const ptn_local_op* op, int mode, const void* v, double dt_re, double dt_im,
               int numiter, void* out, double* energy, int32_t* info, double* alpha_beta,
               void* ws, size_t ws_bytes, hipStream_t s, int64_t batch) {
  if (!op || numiter < 1 || numiter > LZ_MAX || batch < 1 || batch > 65535) return PTN_ERR_ARG;
  const int64_t n = op->kind == PTN_OP_HEFF ? op->Dl * op->d * op->Dr : op->Dl * op->Dr;
  bool use_graph;
  {
    std::lock_guard<std::mutex> lk(g_gmu);
    use_graph = g_graphs_on && n * batch <= GRAPH_MAX_N;
  }
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cst) != hipSuccess) cst = hipStreamCaptureStatusActive;
  if (cst != hipStreamCaptureStatusNone)
    use_graph = false;   // the caller captures (or the query failed): launch into its capture
  // the sector plan (quantum-number blocks) is looked up or built here, outside
  // any capture (building copies synchronously); under the caller's own capture
  // the tile-mask products run. The reference keeps the plan alive until the
  // solve is enqueued.
  SectorPlanRef spr;
  if (batch == 1 && cst == hipStreamCaptureStatusNone) {
    const int e = sector_plan_get(op, SECTORS_HEFF, &spr);
    if (e) return e;
  }
  const SectorPlan* sp = spr.get();
  if (!use_graph)
    return krylov_enqueue(op, mode, v, dt_re, dt_im, numiter, out, energy, info, alpha_beta, ws,
                          ws_bytes, s, batch, sp);
  int dev = 0;
  PTN_HIP_CHECK(hipGetDevice(&dev));
  hipStream_t cap;
  {
    std::lock_guard<std::mutex> lk(g_gmu);
    auto it = g_capture_streams.find(dev);
    if (it == g_capture_streams.end()) {
      hipStream_t ns;
      PTN_HIP_CHECK(hipStreamCreateWithFlags(&ns, hipStreamNonBlocking));
      it = g_capture_streams.emplace(dev, ns).first;
    }
    cap = it->second;
  }
  PTN_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  const int st = krylov_enqueue(op, mode, v, dt_re, dt_im, numiter, out, energy, info, alpha_beta,
                                ws, ws_bytes, cap, batch, sp);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(cap, &g);
  if (st) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (ec != hipSuccess) { ptn_set_last_hip_error(ec); return PTN_ERR_HIP; }
  const GKey key{dev, op->kind, op->precision | (sp ? GRAPH_GROUPED : 0), mode, numiter,
                 op->Dl, op->d, op->Dr, op->Ml, op->Mr, batch};
  std::lock_guard<std::mutex> lk(g_gmu);
  if (g_graph_cache.size() >= GRAPH_MAX_BUCKETS && !g_graph_cache.count(key)) {
    for (auto& kv : g_graph_cache) drop_entry(kv.second);
    g_graph_cache.clear();
  }
  GEntry& E = g_graph_cache[key];
  const int slot = E.next;
  E.next = (slot + 1) % GRAPH_RING;
  hipError_t e = hipSuccess;
  if (E.exec[slot]) {
    // the slot's previous launch must be complete before its parameters change
    if ((e = hipEventSynchronize(E.done[slot])) != hipSuccess) {
      (void)hipGraphDestroy(g);
      ptn_set_last_hip_error(e);
      return PTN_ERR_HIP;
    }
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    if (hipGraphExecUpdate(E.exec[slot], g, &err_node, &res) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(E.exec[slot]);
      E.exec[slot] = nullptr;
    }
  }
  if (!E.exec[slot] && (e = hipGraphInstantiate(&E.exec[slot], g, nullptr, nullptr, 0)) != hipSuccess) {
    E.exec[slot] = nullptr;
    (void)hipGraphDestroy(g);
    ptn_set_last_hip_error(e);
    return PTN_ERR_HIP;
  }
  (void)hipGraphDestroy(g);
  if (!E.done[slot] && (e = hipEventCreateWithFlags(&E.done[slot], hipEventDisableTiming)) != hipSuccess) {
    ptn_set_last_hip_error(e);
    return PTN_ERR_HIP;
  }
  if ((e = hipGraphLaunch(E.exec[slot], s)) != hipSuccess || (e = hipEventRecord(E.done[slot], s)) != hipSuccess) {
    ptn_set_last_hip_error(e);
    return PTN_ERR_HIP;
  }
  ++g_graph_launches;
  return PTN_OK;
}

static int krylov_enqueue(const ptn_local_op* op, int mode, const void* v, double dt_re,
                          double dt_im, int numiter, void* out, double* energy, int32_t* info,
                          double* alpha_beta, void* ws, size_t ws_bytes, hipStream_t s,
                          int64_t batch, const SectorPlan* sp) {
  if (!op || numiter < 1 || numiter > LZ_MAX || batch < 1 || batch > 65535) return PTN_ERR_ARG;
  if (op->kind != PTN_OP_HEFF && op->kind != PTN_OP_BOND) return PTN_ERR_ARG;
  const int bprec = op->precision & ~PTN_BLOCK_SPARSE;
  if (bprec != PTN_FP64 && bprec != PTN_FP32 && bprec != PTN_BF16X3 && bprec != PTN_BF16X6)
    return PTN_ERR_UNSUPPORTED;
  if (op->Dl <= 0 || op->Dr <= 0 || op->Ml <= 0 || op->Mr <= 0) return PTN_ERR_ARG;
  if (op->kind == PTN_OP_HEFF && (op->d <= 0 || !op->w)) return PTN_ERR_ARG;
  if (op->kind == PTN_OP_BOND && op->Ml != op->Mr) return PTN_ERR_ARG;
  const KrylovLayout K = krylov_layout(op, numiter, batch);
  if (ws_bytes < K.total) return PTN_ERR_WORKSPACE;
  char* base = (char*)ws;
  double2* V = (double2*)(base + K.v);
  double2* w = (double2*)(base + K.w);
  double* pa = (double*)(base + K.part_a);
  double* pb = (double*)(base + K.part_b);
  LzState* st = (LzState*)(base + K.state);
  void* mvws = base + K.mv;
  const size_t mvbytes = K.total - K.mv;
  const int64_t n = op_vec_len(op);
  const int rb = red_width(n);
  const int64_t js = batch * n;   // stride between V_j and V_{j+1}
  const double thresh = 100.0 * (double)n * DBL_EPSILON;
  // relative breakdown level of reduced-precision matvecs (0: FP64, the reference's rule;
  // the three-term split has the accuracy of an FP32 product)
  const double rel = (bprec == PTN_FP32 || bprec == PTN_BF16X6)
                         ? 128.0 * FLT_EPSILON
                         : (bprec == PTN_BF16X3 ? 1.0 / 4096 : 0.0);

  // environment blocks in the matvec's operand precision, converted once per solve
  const void* L = op->l;
  const void* R = op->r;
  int tl = PTN_T_C128, tr = PTN_T_C128;
  int st_;
  if (bprec != PTN_FP64) {
    const int64_t nl = op->Dl * op->Ml * op->Dl, nr = op->Dr * op->Mr * op->Dr;
    if ((st_ = convert_c128_c64(batch * nl, op->l, base + K.l64, s))) return st_;
    if ((st_ = convert_c128_c64(batch * nr, op->r, base + K.r64, s))) return st_;
    L = base + K.l64; R = base + K.r64;
    tl = tr = PTN_T_C64;
  }
  Dims D = {op->Dl, op->d, op->Dr, op->Ml, op->Mr};
  auto matvec = [&](const double2* x, double2* y) -> int {
    if (op->kind == PTN_OP_HEFF && sp && batch == 1)
      return heff_apply_sectors(op->precision, D, sp, L, tl, op->w, R, tr, x, y, mvws, mvbytes, s);
    if (op->kind == PTN_OP_HEFF)
      return heff_apply(op->precision, D, L, tl, op->w, R, tr, x, PTN_T_C128, y, PTN_T_C128,
                        mvws, mvbytes, s, batch);
    return bond_apply(op->precision, op->Dl, op->Dr, op->Ml, L, tl, R, tr, x, PTN_T_C128, y,
                      PTN_T_C128, mvws, mvbytes, s, batch);
  };
  int64_t gblocks = (n + 255) / 256;
  if (gblocks > 2048) gblocks = 2048;
  if (gblocks < 1) gblocks = 1;
  const unsigned B = (unsigned)batch;

  // after a stop (breakdown, non-finite value, zero start) the products of the
  // remaining matvecs of that problem do no work
  static_assert(sizeof(LzState) % sizeof(int) == 0, "LzState stride in ints");
  SkipGuard skip(&st->stopped, (int64_t)(sizeof(LzState) / sizeof(int)));
  k_state_init<<<B, 256, 0, s>>>(st, numiter);
  k_sumsq_partial<<<dim3(rb, B), 256, 0, s>>>(n, (const double2*)v, n, pa, rb);
  k_start<<<dim3((unsigned)gblocks, B), 256, 0, s>>>(n, (const double2*)v, n, V, n, pa, rb, st);
  PTN_LAUNCH_CHECK();
  for (int j = 0; j < numiter - 1; ++j) {
    double2* vj = V + (int64_t)j * js;
    if ((st_ = matvec(vj, w))) return st_;
    k_dot_partial<<<dim3(rb, B), 256, 0, s>>>(n, w, n, vj, n, pa, rb);
    k_alpha_update<<<dim3(rb, B), 256, 0, s>>>(n, w, n, vj, j > 0 ? vj - js : vj, n, j, pa, pb,
                                                rb, st);
    k_beta_next<<<dim3((unsigned)gblocks, B), 256, 0, s>>>(n, w, n, vj + js, n, j, pb, rb,
                                                            thresh, rel, st);
    PTN_LAUNCH_CHECK();
  }
  {
    const int j = numiter - 1;
    double2* vj = V + (int64_t)j * js;
    if ((st_ = matvec(vj, w))) return st_;
    k_dot_partial<<<dim3(rb, B), 256, 0, s>>>(n, w, n, vj, n, pa, rb);
    k_alpha_last<<<B, 256, 0, s>>>(j, pa, rb, st);
  }
  k_tridiag<<<B, 64, 0, s>>>(st, mode, dt_re, dt_im);
  k_combine<<<dim3((unsigned)gblocks, B), 256, 0, s>>>(n, V, js, n, st, (double2*)out, n);
  k_export<<<B, 64, 0, s>>>(st, info, alpha_beta, energy, numiter);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free local solve (ptn_krylov_expm_noalloc): the same kernels as
// ptn_krylov_expm, launched directly (no graph), with the workspace in a
// file-scope device array instead of caller or library allocations. Each GPU
// holds its own copy of the array (for the life of the process); its address
// on a device is taken once with hipGetSymbolAddress while that device is
// current and kept per device: the array's name in host code is only a
// host-side stand-in and is never passed to a kernel. Nothing in the
// workspace needs to start at zero (the state block is initialised by the
// first kernel of every solve, every partial sum is written before it is
// read). Calls that overlap on one GPU share the array and race.
__device__ __align__(256) unsigned char scratch[PTN_NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;
}  // namespace

static int scratch_address(void** out) {
  int dev = 0;
  PTN_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      ptn_set_last_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol);
      return PTN_ERR_HIP;
    }
    it = g_scratch_ptr.emplace(dev, p).first;
  }
  *out = it->second;
  return PTN_OK;
}

int krylov_expm_noalloc(const ptn_local_op* op, const void* v, double dt_re, double dt_im,
                        int numiter, void* out, int32_t* info, double* alpha_beta, hipStream_t s) {
  if (!op || numiter < 1 || numiter > LZ_MAX) return PTN_ERR_ARG;
  if (krylov_ws_bytes(op, numiter, 1) > (size_t)PTN_NOALLOC_SCRATCH_BYTES) return PTN_ERR_WORKSPACE;
  void* ws = nullptr;
  int st = scratch_address(&ws);
  if (st) return st;
  NoAllocGuard guard;
  // no sector plan: building one allocates (the tile-mask products run instead)
  return krylov_enqueue(op, 0, v, dt_re, dt_im, numiter, out, nullptr, info, alpha_beta, ws,
                        (size_t)PTN_NOALLOC_SCRATCH_BYTES, s, 1, nullptr);
}

}  // namespace ptn
