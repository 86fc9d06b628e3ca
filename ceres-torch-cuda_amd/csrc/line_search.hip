// The LINE_SEARCH minimizer on bundle adjustment: _line_search_solve
// (solver.py:360-607) with its search directions (solver.py:1078-1162) and the
// step-size interpolation of line_search.py:9-81.
//
// Device work per trial step: Plus from the accepted state, the evaluation
// (per-observation loss-corrected residual and Jacobian, reduced to the cost
// and J^T r: per camera over its observations in camera_obs order, per point
// over its observations, both fixed), and the dot products the tests need.
// Vectors live in tangent coordinates [cameras 9C | points 3P]; a constant
// block's (or coordinate's) column is zero, so it adds nothing to a dot product
// or a norm, exactly as the reference's vector over the active blocks.
// Dot products are fixed-order two-stage sums; the LBFGS two-loop recursion and
// the BFGS update read their scalars from device memory, so a direction costs
// launches but no host round trip.  The host runs the loop's branches on a few
// scalars read back per trial.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "ba_camera.h"
#include "ba_kernels.h"

namespace ct {
namespace {

constexpr int kLsThreads = 256;
constexpr int kObsTerms = 12;    // per observation: Jc^T r (9), Jp^T r (3)
constexpr int kDotGroups = 256;  // partial sums of a reduction (a fixed grid: a fixed order)
constexpr int kFixedSlots = 16;  // device scalars before the LBFGS rho / alpha slots
constexpr double kEpsLs = 2.220446049250313e-16;

enum Slot { kCost = 0, kA, kB, kC, kD, kE, kF, kT, kU };

// A coefficient: c, or c times device scalar s[idx].
struct Coef {
  double c;
  int idx;
};
__device__ __forceinline__ double coef(const Coef& k, const double* s) { return k.idx < 0 ? k.c : k.c * s[k.idx]; }

// ---------------------------------------------------------------------------
// evaluation: cost and gradient J^T r in tangent coordinates
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kLsThreads) void k_ls_eval(BAView v, const double* __restrict__ cams,
                                                        const double* __restrict__ pts, double* __restrict__ gobs,
                                                        double* __restrict__ part) {
  __shared__ double red[kLsThreads / kWave];
  const int o = blockIdx.x * kLsThreads + threadIdx.x;
  double cost = 0.0;
  if (o < v.N) {
    const int ci = v.obs_camera[o], p = v.obs_point[o];
    const double X[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
    const double obs[2] = {v.obs_xy[2 * (size_t)o], v.obs_xy[2 * (size_t)o + 1]};
    double r[2], Jc[2][9], Jp[2][3];
    camera_jacobian(v, cams, ci, X, obs, r, Jc, Jp);
    if (v.camera_mask) {
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        const double m = v.camera_mask[9 * (size_t)ci + j] ? 1.0 : 0.0;
        Jc[0][j] *= m;
        Jc[1][j] *= m;
      }
    }
    if (v.point_mask) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double m = v.point_mask[3 * (size_t)p + j] ? 1.0 : 0.0;
        Jp[0][j] *= m;
        Jp[1][j] *= m;
      }
    }
    cost = robustify2(v.loss, r, Jc, Jp);
    double* out = gobs + (size_t)kObsTerms * o;
#pragma unroll
    for (int j = 0; j < 9; ++j) out[j] = Jc[0][j] * r[0] + Jc[1][j] * r[1];
#pragma unroll
    for (int j = 0; j < 3; ++j) out[9 + j] = Jp[0][j] * r[0] + Jp[1][j] * r[1];
  }
  const double t = block_sum<kLsThreads>(cost, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// g_c for camera blockIdx.x: its observations in camera_obs order, per-thread
// strided sums, then the wave tree and the waves in order.
__global__ __launch_bounds__(kLsThreads) void k_ls_camera_gradient(BAView v, const double* __restrict__ gobs,
                                                                   double* __restrict__ g) {
  __shared__ double red[9][kLsThreads / kWave];
  const int ci = blockIdx.x;
  const int s0 = v.camera_offsets[ci], s1 = v.camera_offsets[ci + 1];
  double acc[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = s0 + (int)threadIdx.x; s < s1; s += kLsThreads) {
    const double* t = gobs + (size_t)kObsTerms * v.camera_obs[s];
#pragma unroll
    for (int j = 0; j < 9; ++j) acc[j] += t[j];
  }
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < 9; ++j) {
    const double w = wave_sum(acc[j]);
    if (lane == 0) red[j][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double t = 0.0;
    for (int w = 0; w < kLsThreads / kWave; ++w) t += red[threadIdx.x][w];
    g[9 * (size_t)ci + threadIdx.x] = t;
  }
}

// g_p: one thread per point, its observations in order.
__global__ __launch_bounds__(kLsThreads) void k_ls_point_gradient(BAView v, const double* __restrict__ gobs,
                                                                  double* __restrict__ g) {
  const int p = blockIdx.x * kLsThreads + threadIdx.x;
  if (p >= v.P) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int o = v.point_offsets[p]; o < v.point_offsets[p + 1]; ++o) {
    const double* t = gobs + (size_t)kObsTerms * o + 9;
    a0 += t[0];
    a1 += t[1];
    a2 += t[2];
  }
  double* out = g + 9 * (size_t)v.C + 3 * (size_t)p;
  out[0] = a0;
  out[1] = a1;
  out[2] = a2;
}

// ---------------------------------------------------------------------------
// reductions (fixed grids, fixed orders)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kLsThreads) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                         int64_t n, double* __restrict__ part) {
  __shared__ double red[kLsThreads / kWave];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; i < n; i += (int64_t)kDotGroups * kLsThreads)
    acc += a[i] * b[i];
  const double t = block_sum<kLsThreads>(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// max that keeps NaN (torch.max does; fmax would drop it and report a NaN gradient as zero)
__device__ __forceinline__ double nanmax(double a, double b) {
  return (a != a || b != b) ? __builtin_nan("") : fmax(a, b);
}

__device__ __forceinline__ double wave_nanmax(double v) {
  for (int off = 32; off > 0; off >>= 1) v = nanmax(v, __shfl_xor(v, off, kWave));
  return v;
}

__global__ __launch_bounds__(kLsThreads) void k_absmax_part(const double* __restrict__ a, int64_t n,
                                                            double* __restrict__ part) {
  __shared__ double red[kLsThreads / kWave];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; i < n; i += (int64_t)kDotGroups * kLsThreads)
    m = nanmax(m, fabs(a[i]));
  m = wave_nanmax(m);
  if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kLsThreads / kWave; ++w) t = nanmax(t, red[w]);
    part[blockIdx.x] = t;
  }
}

// s[out] = mul * (sum of count partials, in order); max instead of sum when take_max.
__global__ __launch_bounds__(kLsThreads) void k_finish(const double* __restrict__ part, int count, double* s, int out,
                                                       Coef mul, int take_max) {
  __shared__ double red[kLsThreads / kWave];
  double acc = 0.0;
  for (int i = threadIdx.x; i < count; i += kLsThreads) acc = take_max ? nanmax(acc, part[i]) : acc + part[i];
  double t;
  if (take_max) {
    acc = wave_nanmax(acc);
    if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    t = 0.0;
    for (int w = 0; w < kLsThreads / kWave; ++w) t = nanmax(t, red[w]);
  } else {
    t = block_sum<kLsThreads>(acc, red);
  }
  if (threadIdx.x == 0) s[out] = coef(mul, s) * t;
}

// ---------------------------------------------------------------------------
// vector and scalar updates
// ---------------------------------------------------------------------------
// out = ca a + cb b (b may be null)
__global__ __launch_bounds__(kLsThreads) void k_axpby(double* out, const double* a, Coef ca, const double* b, Coef cb,
                                                      int64_t n, const double* __restrict__ s) {
  const double xa = coef(ca, s), xb = b ? coef(cb, s) : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLsThreads)
    out[i] = b ? xa * a[i] + xb * b[i] : xa * a[i];
}

enum ScalarOp { kRecipClamp = 0, kDivClamp = 1, kSub = 2, kMax0 = 3 };

__global__ void k_scalar(double* s, int op, int a, int b, int out) {
  if (threadIdx.x != 0) return;
  switch (op) {
    case kRecipClamp:  // 1 / max(s[a], eps)
      s[out] = 1.0 / fmax(s[a], kEpsLs);
      break;
    case kDivClamp:  // s[a] / max(s[b], eps)
      s[out] = s[a] / fmax(s[b], kEpsLs);
      break;
    case kSub:
      s[out] = s[a] - s[b];
      break;
    default:  // max(s[a], 0)
      s[out] = fmax(s[a], 0.0);
      break;
  }
}

// flag <- 1 unless |d_i - f_i| <= 1e-8 + 1e-5 |f_i| everywhere (torch.allclose defaults)
__global__ __launch_bounds__(kLsThreads) void k_allclose(const double* __restrict__ d, const double* __restrict__ f,
                                                         int64_t n, int* flag) {
  bool bad = false;
  for (int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLsThreads)
    if (!(fabs(d[i] - f[i]) <= 1e-8 + 1e-5 * fabs(f[i]))) bad = true;
  if (__any(bad) && threadIdx.x % kWave == 0) atomicOr(flag, 1);
}

__global__ __launch_bounds__(kLsThreads) void k_fill(double* a, double value, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLsThreads)
    a[i] = value;
}

__global__ void k_clear_flag(int* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// xn = Plus(x, t d): Euclidean on angle-axis cameras and points; QuaternionManifold
// on quaternion rotations (constant rotations untouched, problem.py:462-474).
__global__ __launch_bounds__(kLsThreads) void k_ls_plus(BAView v, const double* __restrict__ x,
                                                        const double* __restrict__ d, Coef tk,
                                                        const double* __restrict__ s, double* __restrict__ xn) {
  const double t = coef(tk, s);
  const size_t cam_values = (size_t)v.cam_stride * v.C;
  const int64_t i = (int64_t)blockIdx.x * kLsThreads + threadIdx.x;
  if (v.cam_model == CT_CAMERA_ANGLE_AXIS) {
    if (i < 9 * (int64_t)v.C) xn[i] = x[i] + t * d[i];
  } else if (i < v.C) {
    double dd[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) dd[j] = t * d[9 * i + j];
    double cam[10];
    load_quat_camera(x + 10 * i, v.cam_model, cam);
    quat_camera_plus(cam, dd, rotation_free(v, (int)i));
    store_quat_camera(cam, v.cam_model, xn + 10 * i);
  }
  if (i < 3 * (int64_t)v.P) xn[cam_values + i] = x[cam_values + i] + t * d[9 * (size_t)v.C + i];
}

// ---------------------------------------------------------------------------
// BFGS: the dense inverse Hessian (n <= CT_LS_MAX_BFGS_N)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kLsThreads) void k_identity(double* H, int n) {
  for (int64_t e = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * kLsThreads)
    H[e] = (e / n == e % n) ? 1.0 : 0.0;
}

// out = H x, one wavefront per row in a fixed order
__global__ __launch_bounds__(kLsThreads) void k_matvec(const double* __restrict__ H, const double* __restrict__ x,
                                                       int n, double* __restrict__ out) {
  const int row = blockIdx.x * (kLsThreads / kWave) + threadIdx.x / kWave, lane = threadIdx.x % kWave;
  if (row >= n) return;
  double acc = 0.0;
  for (int c = lane; c < n; c += kWave) acc += H[(size_t)row * n + c] * x[c];
  acc = wave_sum(acc);
  if (lane == 0) out[row] = acc;
}

// H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T (solver.py:1149-1157),
// expanded with the symmetric H: H - rho (s (Hy)^T + (Hy) s^T) + (rho^2 y^T H y + rho) s s^T.
// s[kU] holds rho, s[kT] holds y^T H y.
__global__ __launch_bounds__(kLsThreads) void k_bfgs_update(double* H, const double* __restrict__ sv,
                                                            const double* __restrict__ hy, int n,
                                                            const double* __restrict__ sc) {
  const double rho = sc[kU], yhy = sc[kT];
  const double c2 = rho * rho * yhy + rho;
  for (int64_t e = (int64_t)blockIdx.x * kLsThreads + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * kLsThreads) {
    const int64_t r = e / n, c = e % n;
    H[e] = H[e] - rho * (sv[r] * hy[c] + hy[r] * sv[c]) + c2 * sv[r] * sv[c];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct LsLayout {
  double *xa, *xt, *g, *gt, *d, *f, *dprev, *gprev, *q, *sv, *yv, *hy, *H, *S, *Y, *gobs, *part, *sc;
  int* flag;
  size_t bytes;
};

LsLayout ls_layout(const ct_ba_problem* p, const ct_line_search_options* o, void* base) {
  Carver cv(base);
  LsLayout L{};
  const size_t C = p->num_cameras, P = p->num_points, N = p->num_observations;
  const size_t stride = p->camera_model == CT_CAMERA_ANGLE_AXIS ? 9 : 10;
  const size_t nx = stride * C + 3 * P, n = 9 * C + 3 * P;
  const size_t m = o->direction == CT_LS_LBFGS ? (size_t)o->max_lbfgs_rank : 0;
  const bool bfgs = o->direction == CT_LS_BFGS;
  L.xa = cv.take<double>(nx);
  L.xt = cv.take<double>(nx);
  L.g = cv.take<double>(n);
  L.gt = cv.take<double>(n);
  L.d = cv.take<double>(n);
  L.f = cv.take<double>(n);
  L.dprev = cv.take<double>(n);
  L.gprev = cv.take<double>(n);
  L.q = cv.take<double>(n);
  L.sv = cv.take<double>(n);
  L.yv = cv.take<double>(n);
  L.hy = cv.take<double>(bfgs ? n : 0);
  L.H = cv.take<double>(bfgs ? n * n : 0);
  L.S = cv.take<double>(m * n);
  L.Y = cv.take<double>(m * n);
  L.gobs = cv.take<double>((size_t)kObsTerms * N);
  const size_t eval_blocks = (N + kLsThreads - 1) / kLsThreads;
  L.part = cv.take<double>(eval_blocks > (size_t)kDotGroups ? eval_blocks : (size_t)kDotGroups);
  L.sc = cv.take<double>(kFixedSlots + 2 * m);
  L.flag = cv.take<int>(1);
  L.bytes = cv.offset + 256;
  return L;
}

class LineSearchRun {
 public:
  LineSearchRun(const BAView& v, const ct_line_search_options& o, const LsLayout& L, hipStream_t s)
      : v_(v), o_(o), L_(L), s_(s) {
    n_ = 9 * (int64_t)v.C + 3 * (int64_t)v.P;
    nx_ = (int64_t)v.cam_stride * v.C + 3 * (int64_t)v.P;
    m_ = o.direction == CT_LS_LBFGS ? o.max_lbfgs_rank : 0;
    host_.assign(kFixedSlots + 2 * (size_t)m_, 0.0);
  }

  int eval(const double* x, double* g) {
    const int blocks = (v_.N + kLsThreads - 1) / kLsThreads;
    if (v_.N > 0) {
      hipLaunchKernelGGL(k_ls_eval, dim3(blocks), dim3(kLsThreads), 0, s_, v_, x, x + (size_t)v_.cam_stride * v_.C,
                         L_.gobs, L_.part);
      CT_LAUNCH_CHECK();
    }
    if (v_.C > 0 && v_.N > 0) {
      hipLaunchKernelGGL(k_ls_camera_gradient, dim3(v_.C), dim3(kLsThreads), 0, s_, v_, L_.gobs, g);
      CT_LAUNCH_CHECK();
    } else if (v_.C > 0) {
      hipLaunchKernelGGL(k_fill, dim3(grid(9 * (int64_t)v_.C)), dim3(kLsThreads), 0, s_, g, 0.0, 9 * (int64_t)v_.C);
      CT_LAUNCH_CHECK();
    }
    if (v_.P > 0) {
      hipLaunchKernelGGL(k_ls_point_gradient, dim3((v_.P + kLsThreads - 1) / kLsThreads), dim3(kLsThreads), 0, s_, v_,
                         L_.gobs, g);
      CT_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kLsThreads), 0, s_, L_.part, blocks, L_.sc, (int)kCost, Coef{1.0, -1},
                       0);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int dot(const double* a, const double* b, int out, Coef mul = Coef{1.0, -1}, int64_t n = -1) {
    hipLaunchKernelGGL(k_dot_part, dim3(kDotGroups), dim3(kLsThreads), 0, s_, a, b, n < 0 ? n_ : n, L_.part);
    CT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kLsThreads), 0, s_, L_.part, kDotGroups, L_.sc, out, mul, 0);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int absmax(const double* a, int out) {
    hipLaunchKernelGGL(k_absmax_part, dim3(kDotGroups), dim3(kLsThreads), 0, s_, a, n_, L_.part);
    CT_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kLsThreads), 0, s_, L_.part, kDotGroups, L_.sc, out, Coef{1.0, -1}, 1);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int axpby(double* out, const double* a, Coef ca, const double* b = nullptr, Coef cb = Coef{0.0, -1}) {
    if (n_ == 0) return CT_OK;
    hipLaunchKernelGGL(k_axpby, dim3(grid(n_)), dim3(kLsThreads), 0, s_, out, a, ca, b, cb, n_, L_.sc);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int scalar(int op, int a, int b, int out) {
    hipLaunchKernelGGL(k_scalar, dim3(1), dim3(kWave), 0, s_, L_.sc, op, a, b, out);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int plus(const double* x, const double* d, double t, double* xn) {
    const int64_t work = std::max<int64_t>(9 * (int64_t)v_.C, 3 * (int64_t)v_.P);
    if (work == 0) return CT_OK;
    hipLaunchKernelGGL(k_ls_plus, dim3((work + kLsThreads - 1) / kLsThreads), dim3(kLsThreads), 0, s_, v_, x, d,
                       Coef{t, -1}, L_.sc, xn);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int read() {
    CT_HIP(hipMemcpyAsync(host_.data(), L_.sc, sizeof(double) * host_.size(), hipMemcpyDeviceToHost, s_));
    CT_HIP(hipStreamSynchronize(s_));
    return CT_OK;
  }

  int allclose(const double* a, const double* b, bool* close) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(kWave), 0, s_, L_.flag);
    CT_LAUNCH_CHECK();
    if (n_ > 0) {
      hipLaunchKernelGGL(k_allclose, dim3(grid(n_)), dim3(kLsThreads), 0, s_, a, b, n_, L_.flag);
      CT_LAUNCH_CHECK();
    }
    int h = 0;
    CT_HIP(hipMemcpyAsync(&h, L_.flag, sizeof(int), hipMemcpyDeviceToHost, s_));
    CT_HIP(hipStreamSynchronize(s_));
    *close = h == 0;
    return CT_OK;
  }

  // the search direction into L.d from the gradient g at x (solver.py:1078-1146)
  int direction(const double* g, bool have_prev, const std::vector<int>& hist) {
    const int dir = o_.direction;
    if (dir == CT_LS_NONLINEAR_CONJUGATE_GRADIENT && have_prev) {
      if (o_.nonlinear_cg_type == CT_NCG_FLETCHER_REEVES) {
        CT_TRY(dot(g, g, kA));
        CT_TRY(dot(L_.gprev, L_.gprev, kB));
      } else {
        CT_TRY(axpby(L_.yv, g, Coef{1.0, -1}, L_.gprev, Coef{-1.0, -1}));  // g - g_prev
        CT_TRY(dot(g, L_.yv, kA));
        if (o_.nonlinear_cg_type == CT_NCG_POLAK_RIBIERE) {
          CT_TRY(dot(L_.gprev, L_.gprev, kB));
        } else {
          CT_TRY(dot(L_.dprev, L_.yv, kB));
        }
      }
      CT_TRY(scalar(kDivClamp, kA, kB, kC));
      CT_TRY(scalar(kMax0, kC, 0, kC));
      return axpby(L_.d, g, Coef{-1.0, -1}, L_.dprev, Coef{1.0, kC});
    }
    if (dir == CT_LS_LBFGS && !hist.empty()) {
      // q = g; newest to oldest: alpha_i = rho_i s_i.q, q -= alpha_i y_i
      CT_TRY(axpby(L_.q, g, Coef{1.0, -1}));
      for (int k = (int)hist.size() - 1; k >= 0; --k) {
        const int i = hist[k];
        CT_TRY(dot(L_.S + (size_t)i * n_, L_.q, alpha_slot(i), Coef{1.0, rho_slot(i)}));
        CT_TRY(axpby(L_.q, L_.q, Coef{1.0, -1}, L_.Y + (size_t)i * n_, Coef{-1.0, alpha_slot(i)}));
      }
      if (o_.use_approximate_eigenvalue_bfgs_scaling) {
        const int last = hist.back();
        CT_TRY(dot(L_.S + (size_t)last * n_, L_.Y + (size_t)last * n_, kA));
        CT_TRY(dot(L_.Y + (size_t)last * n_, L_.Y + (size_t)last * n_, kB));
        CT_TRY(scalar(kDivClamp, kA, kB, kC));
        CT_TRY(axpby(L_.q, L_.q, Coef{1.0, kC}));
      }
      // oldest to newest: beta = rho_i y_i.r, r += s_i (alpha_i - beta)
      for (size_t k = 0; k < hist.size(); ++k) {
        const int i = hist[k];
        CT_TRY(dot(L_.Y + (size_t)i * n_, L_.q, kT, Coef{1.0, rho_slot(i)}));
        CT_TRY(scalar(kSub, alpha_slot(i), kT, kU));
        CT_TRY(axpby(L_.q, L_.q, Coef{1.0, -1}, L_.S + (size_t)i * n_, Coef{1.0, kU}));
      }
      return axpby(L_.d, L_.q, Coef{-1.0, -1});
    }
    if (dir == CT_LS_BFGS) {
      CT_TRY(matvec(L_.H, g, L_.hy));
      return axpby(L_.d, L_.hy, Coef{-1.0, -1});
    }
    return axpby(L_.d, g, Coef{-1.0, -1});
  }

  int matvec(const double* H, const double* x, double* out) {
    if (n_ == 0) return CT_OK;
    const int rows = kLsThreads / kWave;
    hipLaunchKernelGGL(k_matvec, dim3((unsigned)((n_ + rows - 1) / rows)), dim3(kLsThreads), 0, s_, H, x, (int)n_,
                       out);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  int identity() {
    if (o_.direction != CT_LS_BFGS || n_ == 0) return CT_OK;
    hipLaunchKernelGGL(k_identity, dim3(grid(n_ * n_)), dim3(kLsThreads), 0, s_, L_.H, (int)n_);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  // BFGS update with s = L.sv, y = L.yv (the caller checked the curvature)
  int bfgs_update() {
    CT_TRY(dot(L_.yv, L_.sv, kA));
    CT_TRY(scalar(kRecipClamp, kA, 0, kU));  // rho = 1 / (y.s); y.s > 0 here
    CT_TRY(matvec(L_.H, L_.yv, L_.hy));
    CT_TRY(dot(L_.yv, L_.hy, kT));
    hipLaunchKernelGGL(k_bfgs_update, dim3(grid(n_ * n_)), dim3(kLsThreads), 0, s_, L_.H, L_.sv, L_.hy, (int)n_,
                       L_.sc);
    CT_LAUNCH_CHECK();
    return CT_OK;
  }

  double host(int slot) const { return host_[slot]; }
  int rho_slot(int i) const { return kFixedSlots + i; }
  int alpha_slot(int i) const { return kFixedSlots + m_ + i; }
  int64_t n() const { return n_; }
  int64_t nx() const { return nx_; }
  int m() const { return m_; }

 private:
  static unsigned grid(int64_t n) {
    int64_t b = (n + kLsThreads - 1) / kLsThreads;
    if (b > 4096) b = 4096;
    return (unsigned)(b < 1 ? 1 : b);
  }

  BAView v_;
  ct_line_search_options o_;
  LsLayout L_;
  hipStream_t s_;
  int64_t n_ = 0, nx_ = 0;
  int m_ = 0;
  std::vector<double> host_;
};

// line_search.py:37-81
double quadratic_step(double step, double cost, double cand, double der) {
  const double denom = 2.0 * (cand - cost - der * step);
  if (denom <= 0.0 || !std::isfinite(denom)) return 0.5 * step;
  const double t = -(der * step * step) / denom;
  return std::isfinite(t) && t > 0.0 ? t : 0.5 * step;
}

double cubic_step(double step, double cost, double cand, double der, bool has_prev, double prev_step,
                  double prev_cand) {
  if (!has_prev || prev_step == step) return quadratic_step(step, cost, cand, der);
  const double t1 = prev_step, t2 = step;
  const double y1 = prev_cand - cost - der * t1;
  const double y2 = cand - cost - der * t2;
  const double det = t1 * t1 * t2 * t2 * (t2 - t1);
  if (det == 0.0 || !std::isfinite(det)) return quadratic_step(step, cost, cand, der);
  const double c2 = (y1 * t2 * t2 * t2 - y2 * t1 * t1 * t1) / det;
  const double c3 = (y2 * t1 * t1 - y1 * t2 * t2) / det;
  if (!std::isfinite(c2) || !std::isfinite(c3) || std::fabs(c3) < 1e-30) return quadratic_step(step, cost, cand, der);
  const double disc = c2 * c2 - 3.0 * c3 * der;
  if (disc < 0.0) return quadratic_step(step, cost, cand, der);
  const double roots[2] = {(-c2 + std::sqrt(disc)) / (3.0 * c3), (-c2 - std::sqrt(disc)) / (3.0 * c3)};
  double best = 0.0;
  bool found = false;
  for (double r : roots)
    if (std::isfinite(r) && r > 0.0 && r < step && (!found || r < best)) {
      best = r;
      found = true;
    }
  return found ? best : quadratic_step(step, cost, cand, der);
}

// line_search.py:9-35
double next_step_size(const ct_line_search_options& o, double step, double cost, double cand, double der,
                      bool has_prev, double prev_step, double prev_cand) {
  double t;
  if (o.interpolation == CT_LS_BISECTION) {
    t = 0.5 * step;
  } else if (o.interpolation == CT_LS_QUADRATIC) {
    t = quadratic_step(step, cost, cand, der);
  } else {
    t = cubic_step(step, cost, cand, der, has_prev, prev_step, prev_cand);
  }
  const double lower = step * std::max(0.5, std::min(o.max_step_contraction, o.min_step_contraction));
  const double upper = step * std::max(o.max_step_contraction, o.min_step_contraction);
  return std::min(std::max(t, lower), upper);
}

int validate_ls_options(const ct_line_search_options* o, int64_t n) {
  CT_CHECK_ARG(o != nullptr, "options is null");
  CT_CHECK_ARG(o->max_num_iterations >= 0, "max_num_iterations must be >= 0");
  CT_CHECK_ARG(o->direction >= CT_LS_STEEPEST_DESCENT && o->direction <= CT_LS_BFGS, "unknown line search direction");
  CT_CHECK_ARG(o->line_search_type == CT_LS_ARMIJO || o->line_search_type == CT_LS_WOLFE, "unknown line search type");
  CT_CHECK_ARG(o->nonlinear_cg_type >= CT_NCG_FLETCHER_REEVES && o->nonlinear_cg_type <= CT_NCG_HESTENES_STIEFEL,
               "unknown nonlinear conjugate gradient type");
  CT_CHECK_ARG(o->interpolation >= CT_LS_BISECTION && o->interpolation <= CT_LS_CUBIC, "unknown interpolation");
  CT_CHECK_ARG(o->max_lbfgs_rank > 0, "max_lbfgs_rank must be > 0");
  CT_CHECK_ARG(o->max_num_step_size_iterations >= 1, "max_num_line_search_step_size_iterations must be >= 1");
  CT_CHECK_ARG(o->max_num_direction_restarts >= 0, "max_num_line_search_direction_restarts must be >= 0");
  CT_CHECK_ARG(o->min_line_search_step_size > 0.0, "min_line_search_step_size must be > 0");
  if (o->direction == CT_LS_BFGS && n > CT_LS_MAX_BFGS_N) {
    set_error("BFGS keeps a dense inverse Hessian: at most " + std::to_string(CT_LS_MAX_BFGS_N) +
              " tangent coordinates on the device path (LBFGS has no such limit)");
    return CT_ERR_UNSUPPORTED;
  }
  return CT_OK;
}

void record(double* records, int k, const double* vals) {
  if (records) std::memcpy(records + (size_t)k * CT_LS_REC_FIELDS, vals, sizeof(double) * CT_LS_REC_FIELDS);
}

}  // namespace
}  // namespace ct

using namespace ct;

extern "C" {

int ct_ba_line_search_workspace_bytes(const ct_ba_problem* problem, const ct_line_search_options* options,
                                      size_t* bytes) {
  clear_error();
  CT_TRY(ba_validate_problem(problem));
  CT_CHECK_ARG(bytes != nullptr, "bytes is null");
  CT_TRY(validate_ls_options(options, 9 * (int64_t)problem->num_cameras + 3 * (int64_t)problem->num_points));
  *bytes = ls_layout(problem, options, nullptr).bytes;
  return CT_OK;
}

int ct_ba_line_search_solve(const ct_ba_problem* problem, const ct_line_search_options* options, double* cameras,
                            double* points, void* workspace, size_t workspace_bytes, ct_stream_t stream,
                            ct_line_search_report* report, double* records) {
  clear_error();
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  CT_TRY(ba_validate_problem(problem));
  CT_CHECK_ARG(report != nullptr, "report is null");
  const int64_t n_total = 9 * (int64_t)problem->num_cameras + 3 * (int64_t)problem->num_points;
  CT_TRY(validate_ls_options(options, n_total));
  CT_CHECK_ARG(problem->num_cameras == 0 || cameras != nullptr, "cameras is null");
  CT_CHECK_ARG(problem->num_points == 0 || points != nullptr, "points is null");
  const LsLayout need = ls_layout(problem, options, nullptr);
  CT_CHECK_ARG(workspace != nullptr && workspace_bytes >= need.bytes,
               "workspace too small (ct_ba_line_search_workspace_bytes)");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  CT_HIP(hipStreamSynchronize(s));  // the caller's uploads precede the host checks of the tables
  CT_TRY(ba_check_tables(*problem));
  const BAView v = ba_make_view(problem);
  const LsLayout L = ls_layout(problem, options, workspace);
  const ct_line_search_options& o = *options;
  LineSearchRun run(v, o, L, s);
  const size_t cam_values = (size_t)v.cam_stride * v.C;
  // the state x = [cameras | points]
  if (cam_values) CT_HIP(hipMemcpyAsync(L.xa, cameras, sizeof(double) * cam_values, hipMemcpyDeviceToDevice, s));
  if (v.P) CT_HIP(hipMemcpyAsync(L.xa + cam_values, points, sizeof(double) * 3 * v.P, hipMemcpyDeviceToDevice, s));
  CT_TRY(run.identity());

  ct_line_search_report rep{};
  rep.termination = CT_TERM_NO_CONVERGENCE;
  double* xa = L.xa;
  double* xt = L.xt;
  double* g = L.g;
  double* gt = L.gt;
  std::vector<int> hist;  // LBFGS history slots, oldest first
  bool have_prev = false;
  bool have_eval = false;  // g and cost at xa come from the accepted trial
  double cost = 0.0;
  int k = 0;
  for (k = 0; k <= o.max_num_iterations; ++k) {
    // evaluation at x (solver.py:387-395); after an accepted step it is that trial's evaluation,
    // computed from the same state by the same kernels, so its values are the same
    if (!have_eval) {
      CT_TRY(run.eval(xa, g));
    }
    CT_TRY(run.dot(g, g, kA));
    CT_TRY(run.absmax(g, kB));
    CT_TRY(run.read());
    if (!have_eval) cost = run.host(kCost);
    if (k == 0) rep.initial_cost = rep.final_cost = cost;
    rep.num_jacobian_evaluations += 1;
    const double gnorm = std::sqrt(run.host(kA)), gmax = run.host(kB);
    double rec[CT_LS_REC_FIELDS] = {0.0};
    rec[CT_LS_REC_ITERATION] = k;
    rec[CT_LS_REC_COST] = cost;
    rec[CT_LS_REC_GRADIENT_NORM] = gnorm;
    rec[CT_LS_REC_GRADIENT_MAX_NORM] = gmax;
    rep.num_records = k + 1;
    auto finish = [&](int term, int reason) {
      rep.termination = term;
      rep.reason = reason;
      rec[CT_LS_REC_TIMESTAMP] = now();
      record(records, k, rec);
    };
    if (gmax <= o.gradient_tolerance) {
      rep.final_cost = cost;
      finish(CT_TERM_CONVERGENCE, CT_REASON_GRADIENT);
      break;
    }
    if (k == o.max_num_iterations) {
      finish(CT_TERM_NO_CONVERGENCE, CT_REASON_MAX_ITERATIONS);
      break;
    }
    CT_TRY(run.direction(g, have_prev, hist));
    int restarts = 0;
    CT_TRY(run.dot(g, L.d, kC));
    CT_TRY(run.dot(L.d, L.d, kD));
    CT_TRY(run.read());
    double dd = run.host(kC);
    if (std::sqrt(run.host(kD)) <= kEpsLs || dd >= 0.0) {
      restarts += 1;
      CT_TRY(run.axpby(L.d, g, Coef{-1.0, -1}));
      hist.clear();
      CT_TRY(run.identity());
      CT_TRY(run.dot(g, L.d, kC));
      CT_TRY(run.dot(L.d, L.d, kD));
      CT_TRY(run.read());
      dd = run.host(kC);
    }
    if (std::sqrt(run.host(kD)) <= kEpsLs || dd >= 0.0) {
      rep.final_cost = cost;
      rec[CT_LS_REC_DIRECTION_RESTARTS] = restarts;
      finish(CT_TERM_CONVERGENCE, CT_REASON_PROJECTED_GRADIENT);
      break;
    }
    // the fallback -g is tried second unless it is the direction already (solver.py:465-467)
    CT_TRY(run.axpby(L.f, g, Coef{-1.0, -1}));
    bool same = false;
    CT_TRY(run.allclose(L.d, L.f, &same));
    const int num_dirs = (gnorm > kEpsLs && !same) ? 2 : 1;
    bool accepted = false;
    const double* acc_dir = L.d;
    double acc_step = 0.0;
    int trials = 0;
    for (int ti = 0; ti < num_dirs && !accepted; ++ti) {
      if (ti > 0) {
        restarts += 1;
        if (restarts > o.max_num_direction_restarts) break;
      }
      const double* dt = ti == 0 ? L.d : L.f;
      double der = dd;
      if (ti > 0) {
        CT_TRY(run.dot(g, dt, kC));
        CT_TRY(run.read());
        der = run.host(kC);
      }
      double t = 1.0, prev_t = 0.0, prev_c = 0.0;
      bool has_prev_trial = false;
      for (int ls = 0; ls < o.max_num_step_size_iterations; ++ls) {
        CT_TRY(run.plus(xa, dt, t, xt));
        CT_TRY(run.eval(xt, gt));
        if (o.line_search_type == CT_LS_WOLFE) CT_TRY(run.dot(gt, dt, kE));
        CT_TRY(run.read());
        rep.num_jacobian_evaluations += 1;
        trials += 1;
        const double cc = run.host(kCost);
        const bool armijo = cc <= cost + o.sufficient_function_decrease * t * der;
        const bool curvature = o.line_search_type == CT_LS_WOLFE
                                   ? std::fabs(run.host(kE)) <= o.sufficient_curvature_decrease * std::fabs(der)
                                   : true;
        if (armijo && curvature) {
          accepted = true;
          acc_dir = dt;
          acc_step = t;
          rep.final_cost = cc;
          rec[CT_LS_REC_STEP_SIZE] = t;
          rec[CT_LS_REC_LINE_SEARCH_ITERATIONS] = ls + 1;
          rec[CT_LS_REC_FUNCTION_EVALUATIONS] = trials;
          rec[CT_LS_REC_GRADIENT_EVALUATIONS] = trials;
          rec[CT_LS_REC_DIRECTION_RESTARTS] = restarts;
          rec[CT_LS_REC_COST_CHANGE] = cost - cc;
          rec[CT_LS_REC_STEP_IS_SUCCESSFUL] = 1.0;
          rep.num_line_search_steps += ls + 1;
          rep.num_line_search_function_evaluations += trials;
          rep.num_line_search_gradient_evaluations += trials;
          rep.num_line_search_direction_restarts += restarts;
          rep.num_successful_steps += 1;
          cost = cc;  // the cost at the new x (this trial's evaluation)
          break;
        }
        const double nt = next_step_size(o, t, cost, cc, der, has_prev_trial, prev_t, prev_c);
        prev_t = t;
        prev_c = cc;
        has_prev_trial = true;
        t = nt;
        if (t < o.min_line_search_step_size) break;
      }
    }
    if (!accepted) {
      rep.num_unsuccessful_steps += 1;
      rep.num_line_search_steps += trials;
      rep.num_line_search_function_evaluations += trials;
      rep.num_line_search_gradient_evaluations += trials;
      rep.num_line_search_direction_restarts += restarts;
      rec[CT_LS_REC_DIRECTION_RESTARTS] = restarts;
      finish(CT_TERM_NO_CONVERGENCE, CT_REASON_LINE_SEARCH_FAILED);
      break;
    }
    const double cost_before = rec[CT_LS_REC_COST];
    // s = t d, y = g_new - g (solver.py:566-581); the accepted state is xt
    CT_TRY(run.axpby(L.sv, acc_dir, Coef{acc_step, -1}));
    CT_TRY(run.axpby(L.yv, gt, Coef{1.0, -1}, g, Coef{-1.0, -1}));
    CT_TRY(run.dot(L.sv, L.sv, kA));
    CT_TRY(run.dot(L.yv, L.yv, kB));
    CT_TRY(run.dot(L.sv, L.yv, kC));
    CT_TRY(run.absmax(gt, kD));
    CT_TRY(run.dot(xt, xt, kF, Coef{1.0, -1}, run.nx()));
    CT_TRY(run.read());
    const double snorm = std::sqrt(run.host(kA)), ynorm = std::sqrt(run.host(kB));
    rec[CT_LS_REC_STEP_NORM] = snorm;
    const bool positive = run.host(kC) > 10.0 * kEpsLs * snorm * ynorm;
    if (positive) {
      if (o.direction == CT_LS_LBFGS) {
        int slot;
        if ((int)hist.size() < run.m()) {
          slot = (int)hist.size();
        } else {
          slot = hist.front();
          hist.erase(hist.begin());
        }
        CT_TRY(run.axpby(L.S + (size_t)slot * run.n(), L.sv, Coef{1.0, -1}));
        CT_TRY(run.axpby(L.Y + (size_t)slot * run.n(), L.yv, Coef{1.0, -1}));
        CT_TRY(run.dot(L.Y + (size_t)slot * run.n(), L.S + (size_t)slot * run.n(), kT));
        CT_TRY(run.scalar(kRecipClamp, kT, 0, run.rho_slot(slot)));
        hist.push_back(slot);
      } else if (o.direction == CT_LS_BFGS) {
        CT_TRY(run.bfgs_update());
      }
    } else {
      hist.clear();
      CT_TRY(run.identity());
    }
    if (o.direction == CT_LS_NONLINEAR_CONJUGATE_GRADIENT) {
      CT_TRY(run.axpby(L.gprev, g, Coef{1.0, -1}));
      CT_TRY(run.axpby(L.dprev, acc_dir, Coef{1.0, -1}));
    }
    have_prev = true;
    std::swap(xa, xt);
    std::swap(g, gt);
    have_eval = true;
    const double gmax_new = run.host(kD), state_norm = std::sqrt(run.host(kF));
    rec[CT_LS_REC_TIMESTAMP] = now();
    record(records, k, rec);
    if (gmax_new <= o.gradient_tolerance) {
      rep.termination = CT_TERM_CONVERGENCE;
      rep.reason = CT_REASON_GRADIENT;
      break;
    }
    if (snorm <= o.parameter_tolerance * (state_norm + o.parameter_tolerance)) {
      rep.termination = CT_TERM_CONVERGENCE;
      rep.reason = CT_REASON_PARAMETER;
      break;
    }
    if (std::fabs(cost_before - rep.final_cost) <= o.function_tolerance * std::max(cost_before, 1.0)) {
      rep.termination = CT_TERM_CONVERGENCE;
      rep.reason = CT_REASON_FUNCTION;
      break;
    }
    if (now() >= o.max_solver_time_in_seconds) {
      rep.termination = CT_TERM_NO_CONVERGENCE;
      rep.reason = CT_REASON_MAX_TIME;
      break;
    }
  }
  // the last accepted state back to the caller's parameters
  if (cam_values) CT_HIP(hipMemcpyAsync(cameras, xa, sizeof(double) * cam_values, hipMemcpyDeviceToDevice, s));
  if (v.P) CT_HIP(hipMemcpyAsync(points, xa + cam_values, sizeof(double) * 3 * v.P, hipMemcpyDeviceToDevice, s));
  CT_HIP(hipStreamSynchronize(s));
  rep.seconds = now();
  *report = rep;
  return CT_OK;
}

}  // extern "C"
