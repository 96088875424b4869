// C ABI of the single-direction modal time path (glw_modal_execute,
// glw_eta20_execute): workspace layout, twiddle tables, the kernel sequence
// and the synthesis of the output records.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "glw_internal.hpp"
#include "glw_modal.hpp"
#include "glw_modal_launch.hpp"

using glw::modal::Args;
using glw::modal::C;
using glw::modal::Dep;

namespace {

constexpr int kBlock = 256;

int mfail(int code, const std::string& msg) { return glw_report_error(code, msg); }

bool t_length_ok(int n) {
  switch (n) {
#define GLW_CASE(L) case L:
    GLW_MODAL_T_LENGTHS(GLW_CASE)
    return true;
#undef GLW_CASE
    default: return false;
  }
}
// Modal lengths one workgroup transforms in LDS.
bool q_length_direct(int n) {
  switch (n) {
#define GLW_CASE(L) case L:
    GLW_MODAL_Q_LENGTHS(GLW_CASE)
    return true;
#undef GLW_CASE
    default: return false;
  }
}
bool q_length_ok(int n) {
  switch (n) {
#define GLW_CASE(L) case L:
    GLW_MODAL_Q_SPLIT_LENGTHS(GLW_CASE)
    return true;
#undef GLW_CASE
    default: return q_length_direct(n);
  }
}
// Sub-transforms of the modal pass at length n: 1 for an LDS length, else
// R0 = 2, 4 or 8 with n / R0 = 8192 or 6144 (0: none).
int q_split(int n) {
  if (q_length_direct(n)) return 1;
  for (int r0 = 2; r0 <= 8; r0 *= 2)
    if (n == 8192 * r0 || n == 6144 * r0) return r0;
  return 0;
}

// tw[m] = exp(-2 pi i m / n), m < n
__global__ void k_modal_twiddles(C* tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[m] = {c, -s};
}

// Output records (modal_time.py:89-94, 253, 328): value at sample n of
// sum_b coef_b exp(-2 pi i b n / N) over the kept bins, with the phase read
// from the N-point table at (b n) mod N.
struct SynthArgs {
  int count, base, span, order;
  const Dep* dep;
  const C* tw_n;
  const C* coef2;  // [2 span + 1], bins 2 base + w
  const C* coef3;  // [3 span + 1], bins 3 base + w
  const double* energy2;
  const double* energy3;
  double* eta11;
  C* eta22;
  C* eta33;
  double* audit;
};

// Lane part of sum_w coef_w exp(-2 pi i (b0 + w) n / N) over w = lane, lane + 64, ...
// (phase index advanced by 64 n mod N per step).
__device__ __forceinline__ C lane_sum(const C* coef, int nw, int b0, int n, int count, const C* tw, int lane) {
  double re = 0.0, im = 0.0;
  if (lane >= nw) return {0.0, 0.0};
  long long idx = ((long long)(b0 + lane) * n) % count;
  const long long step = (64LL * n) % count;
  for (int w = lane; w < nw; w += 64) {
    const C e = tw[idx];
    const C c = coef[w];
    re += c.x * e.x - c.y * e.y;
    im += c.x * e.y + c.y * e.x;
    idx += step;
    if (idx >= count) idx -= count;
  }
  return {re, im};
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One wave per output sample n (4 per workgroup); wave 0 also sums the audit energies.
__global__ __launch_bounds__(256) void k_modal_synth(SynthArgs s) {
  const int n = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (n == 0 && s.audit) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int w = lane; w < 2 * s.span + 1; w += 64) {
      v[0] += s.energy2[2 * w + 1];
      v[1] += s.energy2[2 * w];
    }
    if (s.order == 3) {
      for (int w = lane; w < 3 * s.span + 1; w += 64) {
        v[2] += s.energy3[2 * w + 1];
        v[3] += s.energy3[2 * w];
      }
    }
    for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
      for (int k = 0; k < 4; ++k) s.audit[k] = v[k];
  }
  if (n >= s.count) return;
  if (s.eta11) {  // Re sum_w A_w exp(-2 pi i (base + w) n / N)
    double re = 0.0;
    if (lane <= s.span) {
      long long idx = ((long long)(s.base + lane) * n) % s.count;
      const long long step = (64LL * n) % s.count;
      for (int w = lane; w <= s.span; w += 64) {
        const C e = s.tw_n[idx];
        re += s.dep[w].re * e.x - s.dep[w].im * e.y;
        idx += step;
        if (idx >= s.count) idx -= s.count;
      }
    }
    re = wave_sum(re);
    if (lane == 0) s.eta11[n] = re;
  }
  C v = lane_sum(s.coef2, 2 * s.span + 1, 2 * s.base, n, s.count, s.tw_n, lane);
  v = {wave_sum(v.x), wave_sum(v.y)};
  if (lane == 0) s.eta22[n] = v;
  if (s.order == 3 && s.eta33) {
    C u = lane_sum(s.coef3, 3 * s.span + 1, 3 * s.base, n, s.count, s.tw_n, lane);
    u = {wave_sum(u.x), wave_sum(u.y)};
    if (lane == 0) s.eta33[n] = u;
  }
}

// eta20 record: Hermitian projection of the signed-bin coefficients
// (_eta20.py:175-185) and the real part of their synthesis (_eta20.py:125):
// c'_0 = Re c_0, c'_s = (c_s + conj c_-s) / 2, eta20 = c'_0 + 2 Re sum_s>0 c'_s e^{-2 pi i s n / N}.
// One wave per output sample, lanes over the bins.
__global__ __launch_bounds__(256) void k_eta20_synth(const C* coef, int span, int count, const C* tw_n, double* out) {
  const int n = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (n >= count) return;
  double acc = 0.0;
  if (lane < span) {
    long long idx = ((long long)(lane + 1) * n) % count;
    const long long step = (64LL * n) % count;
    for (int sb = lane + 1; sb <= span; sb += 64) {
      const C cp = coef[span + sb], cm = coef[span - sb];
      const double re = 0.5 * (cp.x + cm.x), im = 0.5 * (cp.y - cm.y);
      const C e = tw_n[idx];
      acc += re * e.x - im * e.y;
      idx += step;
      if (idx >= count) idx -= count;
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) out[n] = coef[span].x + 2.0 * acc;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

struct Layout {
  size_t tw_t, tw_q, tw_n, src, pair, forcing, coef2, coef3, energy2, energy3, tw_qs, partial, total;
};

Layout modal_layout(const glw_modal_desc& d) {
  Layout L{};
  const size_t cs = sizeof(C);
  const size_t nw2 = 2 * (size_t)d.span + 1, nw3 = 3 * (size_t)d.span + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align_up(bytes);
    return at;
  };
  L.tw_t = take(d.t_len * cs);
  L.tw_q = take(d.q_len * cs);
  L.tw_n = take((size_t)d.count * cs);
  L.src = take(2 * (size_t)glw::modal::out_ld((int)nw2) * d.q_len * cs);
  L.pair = d.order == 3 ? take(glw::modal::kPairFields * nw2 * d.q_len * cs) : 0;
  L.forcing = d.order == 3 ? take(2 * (size_t)glw::modal::out_ld((int)nw3) * d.q_len * cs) : 0;
  L.coef2 = take(nw2 * cs);
  L.coef3 = d.order == 3 ? take(nw3 * cs) : 0;
  L.energy2 = take(2 * nw2 * sizeof(double));
  L.energy3 = d.order == 3 ? take(2 * nw3 * sizeof(double)) : 0;
  const int r0 = q_split(d.q_len);
  if (r0 > 1) {
    L.tw_qs = take((size_t)(d.q_len / r0) * cs);
    L.partial = take((d.order == 3 ? nw3 : nw2) * (size_t)r0 * 4 * sizeof(double));
  }
  L.total = off;
  return L;
}

int check_modal(const glw_modal_desc* d) {
  if (!d) return mfail(GLW_EINVAL, "null modal descriptor");
  if (d->order != 2 && d->order != 3) return mfail(GLW_EINVAL, "modal order must be 2 or 3");
  if (d->count < 9 || d->base_bin < 1 || d->span < 0) return mfail(GLW_EINVAL, "modal record or band is empty");
  if ((long long)d->order * (d->base_bin + d->span) >= d->count)
    return mfail(GLW_EINVAL, "the nonlinear output band does not fit in the time record");
  if (!t_length_ok(d->t_len)) return mfail(GLW_EUNSUPPORTED, "unsupported modal t_len " + std::to_string(d->t_len));
  if (!q_length_ok(d->q_len)) return mfail(GLW_EUNSUPPORTED, "unsupported modal q_len " + std::to_string(d->q_len));
  if (d->t_len < d->order * d->span + 1) return mfail(GLW_EINVAL, "t_len below order * span + 1");
  if (d->kspan < 1 || d->k0 < 0 || d->q_len < d->order * d->kspan + 1)
    return mfail(GLW_EINVAL, "q_len below order * kspan + 1 (or invalid deposit band)");
  if (d->rank < 1 || d->rank > glw::modal::kMaxRank) return mfail(GLW_EINVAL, "GL rank must be 1..16");
  if (!(d->dq > 0) || !(d->fstep > 0) || !(d->depth > 0) || !(d->lambda2 > 0) ||
      (d->order == 3 && (!(d->lambda_inner > 0) || !(d->lambda3 > 0))))
    return mfail(GLW_EINVAL, "modal spacing, frequency step, depth and GL scales must be positive");
  return GLW_OK;
}

int twiddles(C* tw, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_modal_twiddles, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, tw, n);
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

// Sums of the R0 sub-transforms of each column, in a fixed order.
__global__ __launch_bounds__(256) void k_modal_qsum(const double* partial, int r0, int ncol, C* coef, double* energy) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= ncol) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int r = 0; r < r0; ++r)
    for (int k = 0; k < 4; ++k) s[k] += partial[((size_t)w * r0 + r) * 4 + k];
  coef[w] = {s[0], s[1]};
  energy[2 * w] = s[2];
  energy[2 * w + 1] = s[3];
}

// Forward half of a split modal transform on the two columns of `x` ([2][Q][ld]).
int split_pre(C* x, int q_len, int ld, int ncol, const C* tw_q, hipStream_t st) {
  const int r0 = q_split(q_len);
  const dim3 grid((ncol + 63) / 64, q_len / r0, 2);
  if (r0 == 2) hipLaunchKernelGGL(glw::modal::k_modal_pre<2>, grid, dim3(64), 0, st, x, q_len, ld, ncol, tw_q);
  else if (r0 == 4) hipLaunchKernelGGL(glw::modal::k_modal_pre<4>, grid, dim3(64), 0, st, x, q_len, ld, ncol, tw_q);
  else if (r0 == 8) hipLaunchKernelGGL(glw::modal::k_modal_pre<8>, grid, dim3(64), 0, st, x, q_len, ld, ncol, tw_q);
  else return mfail(GLW_EUNSUPPORTED, "split modal length " + std::to_string(q_len));
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

// Inverse half on the five pair fields ([5][ncol][Q]).
int split_post(C* pair, int q_len, int ncol, const C* tw_q, hipStream_t st) {
  const int r0 = q_split(q_len);
  const dim3 grid((q_len / r0 + 255) / 256, ncol, glw::modal::kPairFields);
  if (r0 == 2) hipLaunchKernelGGL(glw::modal::k_modal_post<2>, grid, dim3(256), 0, st, pair, q_len, tw_q);
  else if (r0 == 4) hipLaunchKernelGGL(glw::modal::k_modal_post<4>, grid, dim3(256), 0, st, pair, q_len, tw_q);
  else if (r0 == 8) hipLaunchKernelGGL(glw::modal::k_modal_post<8>, grid, dim3(256), 0, st, pair, q_len, tw_q);
  else return mfail(GLW_EUNSUPPORTED, "split modal length " + std::to_string(q_len));
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

int split_sum(const Args& a, int ncol, hipStream_t st) {
  hipLaunchKernelGGL(k_modal_qsum, dim3((ncol + 255) / 256), dim3(256), 0, st, a.partial,
                     q_split(a.q_len), ncol, a.coef, a.energy);
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

int launched(int rc, const char* what) {
  if (rc == 1) return mfail(GLW_EUNSUPPORTED, std::string(what) + ": unsupported length or launch shape");
  if (rc != 0) return mfail(GLW_EDEVICE, std::string(what) + ": launch failed");
  return GLW_OK;
}

}  // namespace

extern "C" {

int glw_modal_length_supported(int32_t n) {
  return (t_length_ok(n) ? 1 : 0) | (q_length_ok(n) ? 2 : 0) | (q_length_direct(n) ? 4 : 0);
}

int glw_modal_workspace_size(const glw_modal_desc* d, size_t* bytes) {
  if (!bytes) return mfail(GLW_EINVAL, "null argument");
  if (int rc = check_modal(d)) return rc;
  *bytes = modal_layout(*d).total;
  return GLW_OK;
}

}  // extern "C"

// Stages bracketed by `ev` (when given): twiddles, order-2 time pass, eta22
// modal pass, order-3 time pass, eta33 modal pass, output synthesis.
constexpr int kModalStages = 6;

static int modal_run(const glw_modal_desc* d, const double* deposit, void* workspace, size_t workspace_bytes,
                     double* eta11, double* eta22, double* eta33, double* audit, hipStream_t st, hipEvent_t* ev) {
  if (int rc = check_modal(d)) return rc;
  if (!deposit || !workspace || !eta22 || (d->order == 3 && !eta33)) return mfail(GLW_EINVAL, "null device pointer");
  const Layout L = modal_layout(*d);
  if (workspace_bytes < L.total) return mfail(GLW_EINVAL, "modal workspace too small");
  auto mark = [&](int i) -> int {
    if (ev) GLW_HIP_CHECK(hipEventRecord(ev[i], st));
    return GLW_OK;
  };
  char* ws = (char*)workspace;
  Args a{};
  a.q_len = d->q_len;
  a.t_len = d->t_len;
  a.count = d->count;
  a.base = d->base_bin;
  a.span = d->span;
  a.k0 = d->k0;
  a.kspan = d->kspan;
  a.order = d->order;
  a.dq = d->dq;
  a.fstep = d->fstep;
  a.depth = d->depth;
  a.dep = (const Dep*)deposit;
  a.tw_t = (const C*)(ws + L.tw_t);
  a.tw_q = (const C*)(ws + L.tw_q);
  const int r0 = q_split(d->q_len);
  a.tw_qs = r0 > 1 ? (const C*)(ws + L.tw_qs) : a.tw_q;
  a.partial = r0 > 1 ? (double*)(ws + L.partial) : nullptr;
  a.src = (C*)(ws + L.src);
  a.pair = d->order == 3 ? (C*)(ws + L.pair) : nullptr;
  a.forcing = d->order == 3 ? (C*)(ws + L.forcing) : nullptr;
  // e2 and p2 spectra ([2][2 span + 1][q_len]) live in the order-3 forcing
  // array, which the order-3 time pass only writes after they are consumed
  a.pair_scratch = a.forcing;
  a.rank = d->rank;
  a.lam2 = d->lambda2;
  a.lam_in = d->lambda_inner;
  a.lam3 = d->lambda3;
  for (int i = 0; i < glw::modal::kMaxRank; ++i) {
    a.nodes[i] = i < d->rank ? d->nodes[i] : 0.0;
    a.weights[i] = i < d->rank ? d->weights[i] : 0.0;
  }
  if (int rc = mark(0)) return rc;
  if (int rc = twiddles((C*)(ws + L.tw_t), d->t_len, st)) return rc;
  if (int rc = twiddles((C*)(ws + L.tw_q), d->q_len, st)) return rc;
  if (int rc = twiddles((C*)(ws + L.tw_n), d->count, st)) return rc;
  if (r0 > 1)
    if (int rc = twiddles((C*)(ws + L.tw_qs), d->q_len / r0, st)) return rc;
  if (int rc = mark(1)) return rc;
  const int nw2 = 2 * d->span + 1, nw3 = 3 * d->span + 1;
  // order-2 sources, eta22 coefficients (+ pair fields)
  if (int rc = launched(glw::modal::launch_t_order2(a, st), "modal time pass (order 2)")) return rc;
  if (int rc = mark(2)) return rc;
  a.coef = (C*)(ws + L.coef2);
  a.energy = (double*)(ws + L.energy2);
  if (r0 > 1)
    if (int rc = split_pre(a.src, d->q_len, glw::modal::out_ld(nw2), nw2, a.tw_q, st)) return rc;
  if (int rc = launched(glw::modal::launch_q_order2(a, nw2, r0, st), "modal pass (eta22)")) return rc;
  if (d->order == 3)
    if (int rc = launched(glw::modal::launch_q_pairs(a, nw2, r0, st), "modal pass (pair fields)")) return rc;
  if (r0 > 1) {
    if (int rc = split_sum(a, nw2, st)) return rc;
    if (d->order == 3)
      if (int rc = split_post(a.pair, d->q_len, nw2, a.tw_q, st)) return rc;
  }
  if (int rc = mark(3)) return rc;
  if (d->order == 3) {
    if (int rc = launched(glw::modal::launch_t_order3(a, st), "modal time pass (order 3)")) return rc;
    if (int rc = mark(4)) return rc;
    a.coef = (C*)(ws + L.coef3);
    a.energy = (double*)(ws + L.energy3);
    if (r0 > 1)
      if (int rc = split_pre(a.forcing, d->q_len, glw::modal::out_ld(nw3), nw3, a.tw_q, st)) return rc;
    if (int rc = launched(glw::modal::launch_q_order3(a, nw3, r0, st), "modal pass (eta33)")) return rc;
    if (r0 > 1)
      if (int rc = split_sum(a, nw3, st)) return rc;
  } else {
    if (int rc = mark(4)) return rc;
  }
  if (int rc = mark(5)) return rc;
  SynthArgs s{};
  s.count = d->count;
  s.base = d->base_bin;
  s.span = d->span;
  s.order = d->order;
  s.dep = a.dep;
  s.tw_n = (const C*)(ws + L.tw_n);
  s.coef2 = (const C*)(ws + L.coef2);
  s.coef3 = d->order == 3 ? (const C*)(ws + L.coef3) : nullptr;
  s.energy2 = (const double*)(ws + L.energy2);
  s.energy3 = d->order == 3 ? (const double*)(ws + L.energy3) : nullptr;
  s.eta11 = eta11;
  s.eta22 = (C*)eta22;
  s.eta33 = (C*)eta33;
  s.audit = audit;
  hipLaunchKernelGGL(k_modal_synth, dim3((d->count + 3) / 4), dim3(256), 0, st, s);
  GLW_HIP_CHECK(hipGetLastError());
  return mark(6);
}

extern "C" {

int glw_modal_execute(const glw_modal_desc* d, const double* deposit, void* workspace, size_t workspace_bytes,
                      double* eta11, double* eta22, double* eta33, double* audit, void* stream) {
  return modal_run(d, deposit, workspace, workspace_bytes, eta11, eta22, eta33, audit, (hipStream_t)stream, nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free variant of glw_modal_execute: its workspace is `scratch`, a
// static device array of kNoAllocScratchBytes that every GPU the library runs
// on holds from module load to process end.  In host code `scratch` names a
// host-side stand-in, not device memory: its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) and only that
// address is used.  Calls on one GPU share the array, so calls that overlap
// on one GPU (two streams, two threads) race; nothing needs it cleared (every
// stage writes what the next one reads).
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(void** ptr) {
  int dev = -1;
  GLW_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return mfail(GLW_EDEVICE, "device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mfail(GLW_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return GLW_OK;
}
}  // namespace

extern "C" {

int glw_modal_execute_noalloc(const glw_modal_desc* d, const double* deposit, double* eta11, double* eta22,
                              double* eta33, double* audit, void* stream) {
  if (int rc = check_modal(d)) return rc;
  const size_t need = modal_layout(*d).total;
  if (need > kNoAllocScratchBytes)
    return mfail(GLW_EINVAL, "modal level needs " + std::to_string(need) + " bytes of workspace, above the " +
                                 std::to_string(kNoAllocScratchBytes) + "-byte static scratch of " +
                                 "glw_modal_execute_noalloc (use glw_modal_execute with a workspace)");
  void* ws = nullptr;
  if (int rc = scratch_on_current_device(&ws)) return rc;
  return modal_run(d, deposit, ws, kNoAllocScratchBytes, eta11, eta22, eta33, audit, (hipStream_t)stream, nullptr);
}

int glw_modal_stage_times(const glw_modal_desc* d, const double* deposit, void* workspace, size_t workspace_bytes,
                          double* eta11, double* eta22, double* eta33, double* audit, void* stream, float* ms) {
  if (!ms) return mfail(GLW_EINVAL, "null stage-time array");
  hipEvent_t ev[kModalStages + 1] = {};
  int rc = GLW_OK;
  for (int i = 0; i <= kModalStages && rc == GLW_OK; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) rc = mfail(GLW_EDEVICE, "hipEventCreate failed");
  if (rc == GLW_OK)
    rc = modal_run(d, deposit, workspace, workspace_bytes, eta11, eta22, eta33, audit, (hipStream_t)stream, ev);
  if (rc == GLW_OK && hipEventSynchronize(ev[kModalStages]) != hipSuccess)
    rc = mfail(GLW_EDEVICE, "hipEventSynchronize failed");
  for (int i = 0; i < kModalStages && rc == GLW_OK; ++i)
    if (hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]) != hipSuccess) rc = mfail(GLW_EDEVICE, "hipEventElapsedTime failed");
  for (int i = 0; i <= kModalStages; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  return rc;
}

static int check_eta20(const glw_eta20_desc* d) {
  if (!d) return mfail(GLW_EINVAL, "null eta20 descriptor");
  if (d->count < 9 || d->base_bin < 1 || d->span < 0 || 2 * d->span + 1 > d->count)
    return mfail(GLW_EINVAL, "eta20 record or band is invalid");
  if (!t_length_ok(d->t_len) || d->t_len < 2 * d->span + 1)
    return mfail(GLW_EUNSUPPORTED, "eta20 t_len unsupported or below 2 span + 1");
  if (!q_length_direct(d->q_len) || d->kspan < 1 || d->q_len < 2 * d->kspan + 1)
    return mfail(GLW_EUNSUPPORTED, "eta20 q_len unsupported or below 2 kspan + 1");
  if (d->rank < 1 || d->rank > glw::modal::kMaxRank) return mfail(GLW_EINVAL, "GL rank must be 1..16");
  if (!(d->dq > 0) || !(d->fstep > 0) || !(d->depth > 0) || !(d->delta_q > 0))
    return mfail(GLW_EINVAL, "eta20 spacing, frequency step, depth and delta_q must be positive");
  return GLW_OK;
}

static Layout eta20_layout(const glw_eta20_desc& d) {
  Layout L{};
  const size_t cs = sizeof(C), nw = 2 * (size_t)d.span + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += align_up(bytes);
    return at;
  };
  L.tw_t = take(d.t_len * cs);
  L.tw_q = take(d.q_len * cs);
  L.tw_n = take((size_t)d.count * cs);
  L.src = take(2 * (size_t)glw::modal::out_ld((int)nw) * d.q_len * cs);
  L.coef2 = take(nw * cs);
  L.energy2 = take(2 * nw * sizeof(double));
  L.total = off;
  return L;
}

int glw_eta20_workspace_size(const glw_eta20_desc* d, size_t* bytes) {
  if (!bytes) return mfail(GLW_EINVAL, "null argument");
  if (int rc = check_eta20(d)) return rc;
  *bytes = eta20_layout(*d).total;
  return GLW_OK;
}

int glw_eta20_execute(const glw_eta20_desc* d, const double* deposit, void* workspace, size_t workspace_bytes,
                      double* eta20, void* stream) {
  if (int rc = check_eta20(d)) return rc;
  if (!deposit || !workspace || !eta20) return mfail(GLW_EINVAL, "null device pointer");
  const Layout L = eta20_layout(*d);
  if (workspace_bytes < L.total) return mfail(GLW_EINVAL, "eta20 workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  Args a{};
  a.q_len = d->q_len;
  a.t_len = d->t_len;
  a.count = d->count;
  a.base = d->base_bin;
  a.span = d->span;
  a.k0 = d->k0;
  a.kspan = d->kspan;
  a.order = 2;
  a.dq = d->dq;
  a.fstep = d->fstep;
  a.depth = d->depth;
  a.dep = (const Dep*)deposit;
  a.tw_t = (const C*)(ws + L.tw_t);
  a.tw_q = (const C*)(ws + L.tw_q);
  a.tw_qs = a.tw_q;
  a.src = (C*)(ws + L.src);
  a.coef = (C*)(ws + L.coef2);
  a.energy = (double*)(ws + L.energy2);
  a.rank = d->rank;
  a.lam3 = d->delta_q;
  for (int i = 0; i < glw::modal::kMaxRank; ++i) {
    a.nodes[i] = i < d->rank ? d->nodes[i] : 0.0;
    a.weights[i] = i < d->rank ? d->weights[i] : 0.0;
  }
  if (int rc = twiddles((C*)(ws + L.tw_t), d->t_len, st)) return rc;
  if (int rc = twiddles((C*)(ws + L.tw_q), d->q_len, st)) return rc;
  if (int rc = twiddles((C*)(ws + L.tw_n), d->count, st)) return rc;
  if (int rc = launched(glw::modal::launch_t_eta20(a, st), "eta20 time pass")) return rc;
  if (int rc = launched(glw::modal::launch_q_eta20(a, 2 * d->span + 1, st), "eta20 modal pass")) return rc;
  hipLaunchKernelGGL(k_eta20_synth, dim3((d->count + 3) / 4), dim3(256), 0, st,
                     (const C*)(ws + L.coef2), d->span, d->count, (const C*)(ws + L.tw_n), eta20);
  GLW_HIP_CHECK(hipGetLastError());
  return GLW_OK;
}

}  // extern "C"
