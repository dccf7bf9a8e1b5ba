// A simulated device for the sanitizer build of libqxhip's host code.
//
// The host code (dsl.cpp, lower.cpp, runtime.cpp) is compiled unchanged against the
// stand-in headers of this folder and linked with this file instead of the HIP
// runtime, rocBLAS and kernels.hip.  Device memory is host memory from calloc,
// registered by address; streams execute in order at once; graph capture records the
// stream's work and a graph launch re-runs it.  A kernel launch runs no arithmetic:
// it proves, from its arguments alone, that every element the kernel may read or
// write lies inside one live device allocation (the reach of each argument block is
// the sum of its bit or leg weights), and aborts with a message otherwise.  Built
// with AddressSanitizer and UndefinedBehaviorSanitizer, the host code's own memory
// errors surface as well.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../csrc/qx_internal.hpp"

struct SimStream {
  bool capturing = false;
  std::vector<std::function<void()>> work;   // captured work (while capturing)
};
struct SimEvent {
  double t = 0;
};
struct SimGraph {
  std::vector<std::function<void()>> work;
};
struct SimGraphExec {
  std::vector<std::function<void()>> work;
};
struct SimBlas {
  hipStream_t stream = nullptr;
  void *ws = nullptr;
  size_t ws_bytes = 0;
};

namespace {

std::recursive_mutex g_mu;
std::map<uintptr_t, size_t> g_allocs;   // base -> bytes of every live device allocation
double g_clock = 0;                     // simulated time of recorded events
int64_t g_launches = 0;                 // kernel launches checked (eager or replayed)

[[noreturn]] void violation(const std::string &msg) {
  std::fprintf(stderr, "SIM DEVICE VIOLATION: %s\n", msg.c_str());
  std::fflush(stderr);
  std::abort();
}

// [p, p + bytes) must lie inside one live device allocation
void need(const void *p, long double bytes, const char *what) {
  if (bytes <= 0) return;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const uintptr_t a = (uintptr_t)p;
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin()) violation(std::string(what) + ": pointer is not device memory");
  --it;
  if (a < it->first || (long double)(a - it->first) + bytes > (long double)it->second) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %.0Lf bytes at offset %llu of a %zu-byte allocation", what, bytes,
                  (unsigned long long)(a - it->first), it->second);
    violation(buf);
  }
}

bool is_device(const void *p) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const uintptr_t a = (uintptr_t)p;
  auto it = g_allocs.upper_bound(a);
  if (it == g_allocs.begin()) return false;
  --it;
  return a < it->first + it->second;
}

// run `f` now, or record it when the stream is being captured
void submit(hipStream_t s, std::function<void()> f) {
  if (s && s->capturing) {
    s->work.push_back(std::move(f));
    return;
  }
  f();
}

void launch(void *stream, std::function<void()> check) {
  submit((hipStream_t)stream, [check]() {
    check();
    ++g_launches;
  });
}

size_t esize(int prec) { return prec == QX_C128 ? 16 : 8; }

template <typename T>
long double sum_n(const T *w, int n) {
  long double s = 0;
  for (int i = 0; i < n; ++i) s += (long double)w[i];
  return s;
}

template <typename T>
long double max_n(const T *w, int n) {
  long double m = 0;
  for (int i = 0; i < n; ++i) m = std::max(m, (long double)w[i]);
  return m;
}

void *alloc_device(size_t bytes) {
  void *p = std::calloc(bytes ? bytes : 1, 1);
  if (!p) return nullptr;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  g_allocs[(uintptr_t)p] = bytes ? bytes : 1;
  return p;
}

}  // namespace

// ------------------------------------------------------------------ HIP runtime
hipError_t hipSetDevice(int device) { return device == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDevice(int *device) {
  *device = 0;
  return hipSuccess;
}
hipError_t hipGetDeviceCount(int *count) {
  *count = 1;
  return hipSuccess;
}
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "simulated device error"; }

hipError_t hipMalloc(void **p, size_t bytes) {
  *p = alloc_device(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}

hipError_t hipFree(void *p) {
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    auto it = g_allocs.find((uintptr_t)p);
    if (it == g_allocs.end()) violation("hipFree of a pointer that is not a live allocation");
    g_allocs.erase(it);
  }
  std::free(p);
  return hipSuccess;
}

hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes) {
  *total_bytes = size_t(288) << 30;
  *free_bytes = size_t(280) << 30;
  return hipSuccess;
}

hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s) {
  submit(s, [=]() {
    if (kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice) need(dst, bytes, "copy destination");
    if (kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice) need(src, bytes, "copy source");
    if (kind == hipMemcpyHostToDevice && is_device(src)) violation("host-to-device copy from device memory");
    if (kind == hipMemcpyDeviceToHost && is_device(dst)) violation("device-to-host copy into device memory");
    if (bytes) std::memmove(dst, src, bytes);
  });
  return hipSuccess;
}

hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  return hipMemcpyAsync(dst, src, bytes, kind, nullptr);
}

hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  *s = new SimStream;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  if (s && s->capturing) violation("stream destroyed while capturing");
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  if (s && s->capturing) violation("stream synchronized while capturing");
  return hipSuccess;
}

hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new SimEvent;
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  if (!e) violation("record of a null event");
  submit(s, [e]() { e->t = (g_clock += 1e-3); });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  if (!a || !b) violation("elapsed time of a null event");
  *ms = (float)(b->t - a->t);
  return hipSuccess;
}

hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st) {
  *st = (s && s->capturing) ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s || s->capturing) return hipErrorInvalidValue;
  s->capturing = true;
  s->work.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->capturing) return hipErrorInvalidValue;
  s->capturing = false;
  *g = new SimGraph;
  (*g)->work.swap(s->work);
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *ex, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  if (!g) return hipErrorInvalidValue;
  *ex = new SimGraphExec;
  (*ex)->work = g->work;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t ex) {
  delete ex;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t ex, hipStream_t s) {
  if (!ex) return hipErrorInvalidValue;
  if (s && s->capturing) violation("graph launched into a capturing stream");
  for (auto &f : ex->work) f();   // replays check against the allocations live now
  return hipSuccess;
}

// ------------------------------------------------------------------ rocBLAS
rocblas_status rocblas_create_handle(rocblas_handle *h) {
  *h = new SimBlas;
  return rocblas_status_success;
}
rocblas_status rocblas_destroy_handle(rocblas_handle h) {
  delete h;
  return rocblas_status_success;
}
rocblas_status rocblas_set_workspace(rocblas_handle h, void *ws, size_t bytes) {
  need(ws, bytes, "BLAS workspace");
  h->ws = ws;
  h->ws_bytes = bytes;
  return rocblas_status_success;
}
rocblas_status rocblas_set_stream(rocblas_handle h, hipStream_t s) {
  h->stream = s;
  return rocblas_status_success;
}
const char *rocblas_status_to_string(rocblas_status s) {
  return s == rocblas_status_success ? "rocblas_status_success" : "rocblas_status_invalid_size";
}

namespace {
// reach in elements of a column-major (rows x cols) operand, op applied, over the batch
long double mat_reach(bool trans, int64_t rows, int64_t cols, int64_t ld, int64_t stride, int64_t batch) {
  const int64_t r = trans ? cols : rows, c = trans ? rows : cols;   // stored shape
  if (ld < r) violation("leading dimension below the stored row count");
  return (long double)(r - 1) + (long double)(c - 1) * ld + (long double)(batch - 1) * stride + 1;
}

rocblas_status gemm_check(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int64_t m, int64_t n,
                          int64_t k, const void *A, size_t ea, int64_t lda, int64_t sa, const void *B, size_t eb,
                          int64_t ldb, int64_t sb, const void *C, size_t ec, int64_t ldc, int64_t sc, int64_t batch) {
  if (!h) violation("BLAS call without a handle");
  if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return rocblas_status_invalid_size;
  const bool xa = ta == rocblas_operation_transpose, xb = tb == rocblas_operation_transpose;
  const long double ra = mat_reach(xa, m, k, lda, sa, batch) * ea;
  const long double rb = mat_reach(xb, k, n, ldb, sb, batch) * eb;
  const long double rc = mat_reach(false, m, n, ldc, sc, batch) * ec;
  launch(h->stream, [=]() {
    need(A, ra, "BLAS A");
    need(B, rb, "BLAS B");
    need(C, rc, "BLAS C");
  });
  return rocblas_status_success;
}
}  // namespace

rocblas_status rocblas_zgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                             rocblas_int m, rocblas_int n, rocblas_int k,
                                             const rocblas_double_complex *, const rocblas_double_complex *A,
                                             rocblas_int lda, rocblas_stride sa, const rocblas_double_complex *B,
                                             rocblas_int ldb, rocblas_stride sb, const rocblas_double_complex *,
                                             rocblas_double_complex *C, rocblas_int ldc, rocblas_stride sc,
                                             rocblas_int batch) {
  return gemm_check(h, ta, tb, m, n, k, A, 16, lda, sa, B, 16, ldb, sb, C, 16, ldc, sc, batch);
}

rocblas_status rocblas_cgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                             rocblas_int m, rocblas_int n, rocblas_int k,
                                             const rocblas_float_complex *, const rocblas_float_complex *A,
                                             rocblas_int lda, rocblas_stride sa, const rocblas_float_complex *B,
                                             rocblas_int ldb, rocblas_stride sb, const rocblas_float_complex *,
                                             rocblas_float_complex *C, rocblas_int ldc, rocblas_stride sc,
                                             rocblas_int batch) {
  return gemm_check(h, ta, tb, m, n, k, A, 8, lda, sa, B, 8, ldb, sb, C, 8, ldc, sc, batch);
}

rocblas_status rocblas_gemm_strided_batched_ex(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                               rocblas_int m, rocblas_int n, rocblas_int k, const void *,
                                               const void *A, rocblas_datatype, rocblas_int lda, rocblas_stride sa,
                                               const void *B, rocblas_datatype, rocblas_int ldb, rocblas_stride sb,
                                               const void *, const void *C, rocblas_datatype, rocblas_int ldc,
                                               rocblas_stride sc, void *D, rocblas_datatype, rocblas_int ldd,
                                               rocblas_stride sd, rocblas_int batch, rocblas_datatype,
                                               rocblas_gemm_algo, int32_t, uint32_t) {
  if (C != D || ldc != ldd || sc != sd) violation("bf16 GEMM with C differing from D");
  return gemm_check(h, ta, tb, m, n, k, A, 2, lda, sa, B, 2, ldb, sb, C, 4, ldc, sc, batch);
}

// ------------------------------------------------------------------ kernels (checked, not run)
namespace qx {

int scratch_address(int device, void **ptr) {
  static std::map<int, void *> addr;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  auto it = addr.find(device);
  if (it == addr.end()) it = addr.emplace(device, alloc_device(kScratchBytes)).first;
  *ptr = it->second;
  return 0;
}

void launch_contract_bits(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                          const BitArgs &a, void *stream) {
  const size_t es = esize(prec);
  const long double ra = (sum_n(a.wa, a.nbc) + sum_n(a.ka, a.nbk) + 1) * es;
  const long double rb = (sum_n(a.wb, a.nbc) + sum_n(a.kb, a.nbk) + 1) * es;
  const long double rw = a.ksplit > 1 ? (long double)c_elems * a.ksplit * es : 0;
  launch(stream, [=]() {
    need(A, ra, "bits A");
    need(B, rb, "bits B");
    need(C, (long double)c_elems * es, "bits C");
    need(ws, rw, "bits K-split partials");
  });
}

void launch_contract_reduce(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                            const BitArgs &a, void *stream) {
  launch_contract_bits(prec, A, B, C, ws, c_elems, a, stream);
}

void launch_contract_gemm(int prec, const void *A, const void *B, void *C, void *ws, const GemmArgs &g,
                          void *stream) {
  const size_t es = esize(prec);
  const long double ra = (sum_n(g.ha, g.nbh) + sum_n(g.ma, g.nbm) + sum_n(g.ka, g.nbk) + 1) * es;
  const long double rb = (sum_n(g.hb, g.nbh) + sum_n(g.nb, g.nbn) + sum_n(g.kb, g.nbk) + 1) * es;
  const long double rc = (sum_n(g.hc, g.nbh) + sum_n(g.mc, g.nbm) + sum_n(g.nc, g.nbn) + 1) * es;
  const long double rw = g.ksplit > 1 ? (long double)g.c_elems * g.ksplit * es : 0;
  launch(stream, [=]() {
    need(A, ra, "gemm A");
    need(B, rb, "gemm B");
    need(C, rc, "gemm C");
    need(ws, rw, "gemm K-split partials");
  });
}

void launch_contract_rows(int prec, const void *A, const void *B, void *C, const RowArgs &r, void *stream) {
  const size_t es = esize(prec);
  const long double ra = (sum_n(r.ra, r.nbr) + max_n(r.ka, r.nk) + 1) * es;
  const long double rb = (sum_n(r.rb, r.nbr) + max_n(r.kb, r.nk) + max_n(r.nb, r.nn) + 1) * es;
  const long double rc = (sum_n(r.rc, r.nbr) + max_n(r.nc, r.nn) + 1) * es;
  launch(stream, [=]() {
    need(A, ra, "rows A");
    need(B, rb, "rows B");
    need(C, rc, "rows C");
  });
}

void launch_outer(int prec, void *C, int64_t c_elems, const OuterArgs &a, void *stream) {
  const size_t es = esize(prec);
  if (a.nf < 2 || a.nf > kOuterMaxFactors) violation("outer run: factor count");
  std::vector<std::pair<const void *, long double>> reach;
  for (int f = 0; f < a.nf; ++f) {
    long double r = 1;
    for (int k = a.first[f]; k < a.first[f + 1]; ++k) {
      if (a.bit[k] < 0 || a.bit[k] >= kOuterBits || (int64_t(1) << a.bit[k]) >= c_elems)
        violation("outer run: term bit outside C");
      for (int u = a.first[f]; u < k; ++u)
        if (a.bit[u] == a.bit[k]) violation("outer run: a factor names a bit twice (its LDS table slot)");
      r += (long double)a.w[k];
    }
    reach.push_back({a.F[f], r * es});
  }
  launch(stream, [=]() {
    for (auto &x : reach) need(x.first, x.second, "outer run factor");
    need(C, (long double)c_elems * es, "outer run C");
  });
}

void launch_contract_radix(int prec, const void *A, const void *B, void *C, int64_t c_elems, const RadixArgs &r,
                           void *stream) {
  const size_t es = esize(prec);
  long double ra = 1, rb = 1;
  for (int i = 0; i < r.ncl; ++i) ra += (long double)(r.cd[i] - 1) * r.ca[i], rb += (long double)(r.cd[i] - 1) * r.cb[i];
  for (int i = 0; i < r.nkl; ++i) ra += (long double)(r.kd[i] - 1) * r.ka[i], rb += (long double)(r.kd[i] - 1) * r.kb[i];
  launch(stream, [=]() {
    need(A, ra * es, "radix A");
    need(B, rb * es, "radix B");
    need(C, (long double)c_elems * es, "radix C");
  });
}

void launch_fill_outputs(int prec, void *arena, const int64_t *d_out_off, int n_out, const int *d_out_q,
                         const char *d_bits, int64_t nvalid, int64_t batch, int64_t num_chars, void *stream) {
  const size_t es = esize(prec);
  launch(stream, [=]() {
    need(d_out_off, (long double)n_out * 8, "output offsets");
    need(d_out_q, (long double)n_out * 4, "output qubits");
    need(d_bits, (long double)nvalid * num_chars, "bitstrings");
    for (int o = 0; o < n_out; ++o) {
      if (d_out_q[o] < 1 || d_out_q[o] > num_chars) violation("output qubit outside the bitstring");
      need((const char *)arena + d_out_off[o] * (int64_t)es, (long double)2 * batch * es, "output tensor");
    }
  });
}

namespace {
long double root_reach(const int64_t *d_relem_off, int64_t relems, int64_t bstride, int64_t nvalid) {
  need(d_relem_off, (long double)relems * 8, "result element offsets");
  return (max_n(d_relem_off, (int)relems) + (long double)(nvalid - 1) * bstride + 1);
}
}  // namespace

void launch_accumulate(int prec, const void *root, int64_t root_bstride, int64_t batch, int64_t nvalid,
                       int64_t relems, const int64_t *d_relem_off, double *acc, int64_t row0, void *stream) {
  const size_t es = esize(prec);
  if (nvalid > batch) violation("accumulate: more valid rows than the batch");
  launch(stream, [=]() {
    need(root, root_reach(d_relem_off, relems, root_bstride, nvalid) * es, "saved tensor");
    need(acc, (long double)(row0 + nvalid) * relems * 32, "accumulators");
  });
}

void launch_finalize(const double *acc, double *out, int64_t n, void *stream) {
  launch(stream, [=]() {
    need(acc, (long double)n * 32, "accumulators");
    need(out, (long double)n * 16, "results");
  });
}

void launch_add_mapped(int prec, int dst_double, void *dst, const void *src, int64_t n,
                       const std::vector<int64_t> &dims, const std::vector<int64_t> &strides, void *stream) {
  const size_t es = esize(prec);
  long double rd = 1, cnt = 1;
  for (size_t j = 0; j < dims.size(); ++j) rd += (long double)(dims[j] - 1) * strides[j], cnt *= dims[j];
  if (cnt != n) violation("add_mapped: dims do not multiply to n");
  const long double de = dst_double ? 16 : es;
  launch(stream, [=]() {
    need(dst, rd * de, "add destination");
    need(src, (long double)n * es, "add source");
  });
}

void launch_root_cot(int prec, void *dst, int64_t dep, int64_t batch, int64_t nvalid, int64_t relems,
                     const int64_t *d_relem_off, const double *amp, int64_t row0, void *stream) {
  const size_t es = esize(prec);
  if (nvalid > batch) violation("root_cot: more valid rows than the batch");
  launch(stream, [=]() {
    need(dst, root_reach(d_relem_off, relems, dep, dep ? batch : 1) * es, "root cotangent");
    need(amp, (long double)(row0 + nvalid) * relems * 16, "amplitude cotangents");
  });
}

void launch_root_cot_norm(int prec, void *dst, const void *val, int64_t dep, int64_t batch, int64_t nvalid,
                          int64_t relems, const int64_t *d_relem_off, double *amp, int64_t row0, void *stream) {
  const size_t es = esize(prec);
  if (nvalid > batch) violation("root_cot_norm: more valid rows than the batch");
  launch(stream, [=]() {
    const long double r = root_reach(d_relem_off, relems, dep, batch) * es;
    need(dst, r, "root cotangent");
    need(val, r, "root value");
    need(amp, (long double)(row0 + nvalid) * relems * 16, "amplitudes");
  });
}

void launch_small_batch(int prec, const SmallDesc *d_desc, int n, void *stream) {
  const size_t es = esize(prec);
  launch(stream, [=]() {
    need(d_desc, (long double)n * sizeof(SmallDesc), "small-step descriptors");
    for (int i = 0; i < n; ++i) {
      const SmallDesc &d = d_desc[i];
      if (d.nbc > kSmallMaxBits || d.nbk > kSmallMaxBits || d.c0 < 0 || d.c1 <= d.c0 || d.c1 > (1 << d.nbc))
        violation("small-step descriptor out of range");
      need(d.A, (sum_n(d.wa, d.nbc) + sum_n(d.ka, d.nbk) + 1) * es, "small-step A");
      need(d.B, (sum_n(d.wb, d.nbc) + sum_n(d.kb, d.nbk) + 1) * es, "small-step B");
      need(d.C, (long double)d.c1 * es, "small-step C");
    }
  });
}

void launch_small_walk(int prec, const SmallDesc *d_desc, const int *d_level_off, int nlevels, void *stream) {
  launch(stream, [=]() {
    need(d_level_off, (long double)(nlevels + 1) * 4, "level table");
    for (int L = 0; L < nlevels; ++L)
      if (d_level_off[L + 1] < d_level_off[L]) violation("level table not increasing");
  });
  if (nlevels > 0)   // the chunks themselves: as a batch of every level's descriptors
    submit((hipStream_t)stream, [=]() {
      const int n = d_level_off[nlevels] - d_level_off[0];
      launch_small_batch(prec, d_desc + d_level_off[0], n, nullptr);
    });
}

void launch_add_batch(int prec, const AddDesc *d_desc, int n, void *stream) {
  const size_t es = esize(prec);
  launch(stream, [=]() {
    need(d_desc, (long double)n * sizeof(AddDesc), "add descriptors");
    for (int i = 0; i < n; ++i) {
      const AddDesc &d = d_desc[i];
      if (d.nd < 0 || d.nd > kSmallMaxBits) violation("add descriptor out of range");
      long double rd = 1;
      for (int j = 0; j < d.nd; ++j) rd += (long double)(d.dim[j] - 1) * d.st[j];
      need(d.dst, rd * (d.dst_f64 ? 16 : es), "batched add destination");
      need(d.src, (long double)d.n * es, "batched add source");
    }
  });
}

void launch_bf16_split(const void *src, int64_t n, void *out, void *stream) {
  launch(stream, [=]() {
    need(src, (long double)n * 8, "split source");
    need(out, (long double)n * 8, "split planes");
  });
}

void launch_gather(int prec, const void *src, void *dst, int nb, const int64_t *w, void *stream) {
  const size_t es = esize(prec);
  const long double rs = (sum_n(w, nb) + 1) * es;
  launch(stream, [=]() {
    need(src, rs, "gather source");
    need(dst, std::ldexp((long double)es, nb), "gather destination");
  });
}

void launch_planar_merge(const float *re, const float *im, int64_t n, void *dst, void *stream) {
  launch(stream, [=]() {
    need(re, (long double)n * 4, "merge re");
    need(im, (long double)n * 4, "merge im");
    need(dst, (long double)n * 8, "merge destination");
  });
}

void launch_zero_batch(void *base, const int64_t *d_ranges, int n, void *stream) {
  launch(stream, [=]() {
    need(d_ranges, (long double)n * 16, "zero ranges");
    for (int i = 0; i < n; ++i)
      need((char *)base + d_ranges[2 * i] * 16, (long double)d_ranges[2 * i + 1] * 16, "zero range");
  });
}

void launch_contract_chain(int prec, const void *X0, void *Xm, void *ws, const ChainArgs &a,
                           const ChainMergeArgs &merge, void *stream) {
  const size_t es = esize(prec);
  if (a.nsteps < 1 || a.nsteps > kChainMaxSteps || a.nouter < 0 || a.nouter > kMaxBits)
    violation("chain arguments out of range");
  const long double rin = (sum_n(a.win, kChainTileBits) + sum_n(a.oin, a.nouter) + 1) * es;
  const long double rout = (sum_n(a.wout, kChainTileBits) + sum_n(a.oout, a.nouter) + 1) * es;
  std::vector<std::pair<const void *, long double>> s;
  long double packed = 0;
  for (int j = 0; j < a.nsteps; ++j) {
    const ChainStepArgs &t = a.st[j];
    if (t.k < 0 || t.k > 5 || t.h < 0 || t.h > 3 || t.k + t.h < 1)
      violation("chain step " + std::to_string(j) + " out of range: k " + std::to_string(t.k) + ", h " +
                std::to_string(t.h));
    if (t.merged && (t.h != 0 || t.k > 4)) violation("chain merged step with batch legs or over 4 bits");
    long double r = 1;
    for (int i = 0; i < t.k; ++i) r += (long double)t.sk[i] + t.sn[i];
    for (int i = 0; i < t.h; ++i) r += t.sh[i];
    if (!t.merged) s.emplace_back(a.S[j], r * es);
    packed = std::max(packed, (long double)t.soff + std::ldexp(1.0L, t.h + 2 * t.k));
  }
  if (packed > a.s_elems) violation("chain: packed gate blocks beyond s_elems");
  // merged steps: every group names a merged step, covers its members, and writes its
  // K x K block inside the packed blocks; every member's reach inside its gate block
  if (merge.ngroups < 0 || merge.ngroups > a.nsteps) violation("chain merge: group count");
  int merged_steps = 0;
  for (int j = 0; j < a.nsteps; ++j) merged_steps += a.st[j].merged;
  if (merge.ngroups != merged_steps) violation("chain merge: groups do not match the merged steps");
  for (int gi = 0; gi < merge.ngroups; ++gi) {
    const auto &g = merge.g[gi];
    if (g.step < 0 || g.step >= a.nsteps || !a.st[g.step].merged || g.k != a.st[g.step].k || g.soff != a.st[g.step].soff)
      violation("chain merge: group " + std::to_string(gi) + " does not name its merged step");
    if (g.first < 0 || g.count < 2 || g.first + g.count > kChainMaxMerged) violation("chain merge: member range");
    for (int c = g.first; c < g.first + g.count; ++c) {
      const auto &mm = merge.m[c];
      if (mm.kc < 1 || mm.kc > 4) violation("chain merge: member legs");
      long double r = 1;
      for (int i = 0; i < mm.kc; ++i) {
        if (mm.pos[i] < 0 || mm.pos[i] >= g.k) violation("chain merge: member bit outside its step");
        r += (long double)mm.sk[i] + mm.sn[i];
      }
      s.emplace_back(mm.S, r * es);
    }
  }
  if (a.tab_off < a.s_elems) violation("chain: step tables overlap the packed gate blocks");
  if ((int64_t)kChainTabInts * 4 > (int64_t)kChainTabElems * (int64_t)es)
    violation("chain: step table larger than its workspace share");
  launch(stream, [=]() {
    need(X0, rin, "chain X0");
    need(Xm, rout, "chain Xm");
    need(ws, packed * es, "chain packed gate blocks");
    need((const char *)ws + (size_t)a.tab_off * es, (long double)kChainTabInts * 4 * a.nsteps, "chain step tables");
    for (auto &p : s) need(p.first, p.second, "chain gate block");
  });
}

int64_t plain_gemm_units(int prec, const BlasArgs &b) {
  const int tm = prec == QX_C128 || b.split ? 64 : 128, tn = prec == QX_C128 || prec == QX_C64_3M || b.split ? 64 : 128;
  return b.batch * ((b.r + tm - 1) / tm) * ((b.s + tn - 1) / tn);
}

void launch_plain_gemm(int prec, const void *X, const void *Y, void *C, void *ws, const BlasArgs &b, void *stream) {
  const size_t es = esize(prec);
  if (b.r <= 0 || b.s <= 0 || b.k <= 0 || b.batch <= 0 || b.ksplit < 1 || b.k % ((b.split ? 32 : 16) * b.ksplit) ||
      (b.split && prec == QX_C128))
    violation("plain GEMM sizes out of range");
  const long double rx = mat_reach(b.opx, b.r, b.k, b.ldx, b.sx, b.batch) * es;
  const long double ry = mat_reach(b.opy, b.k, b.s, b.ldy, b.sy, b.batch) * es;
  const long double rc = mat_reach(false, b.r, b.s, b.ldc, b.sc, b.batch) * es;
  const long double rw = b.ksplit > 1 ? (long double)b.ksplit * b.batch * b.r * b.s * es : 0;
  launch(stream, [=]() {
    need(X, rx, "plain GEMM X");
    need(Y, ry, "plain GEMM Y");
    need(C, rc, "plain GEMM C");
    need(ws, rw, "plain GEMM K-split partials");
  });
}

void launch_gen_bits(char *bits, int64_t nb, int nch, uint64_t, int64_t, void *stream) {
  launch(stream, [=]() { need(bits, (long double)nb * nch, "generated bitstrings"); });
}

void launch_mfma_probe(int prec, void *out, int blocks, int, void *stream) {
  launch(stream, [=]() { need(out, (long double)blocks * 256 * (prec == QX_C128 ? 8 : 4), "probe output"); });
}

void launch_reject(const double *amps, const char *cand, int64_t nb, int nch, uint64_t, int64_t, double, int,
                   int64_t num_samples, double *state, int64_t *cnt, char *out_bits, double *out_amps,
                   void *stream) {
  launch(stream, [=]() {
    need(amps, (long double)nb * 16, "candidate amplitudes");
    need(cand, (long double)nb * nch, "candidate bitstrings");
    need(state, 8, "rejection state");
    need(cnt, 16, "rejection counters");
    need(out_bits, (long double)num_samples * nch, "accepted bitstrings");
    need(out_amps, (long double)num_samples * 16, "accepted amplitudes");
  });
}

void launch_zero(void *p, int64_t bytes, void *stream) {
  if (bytes % 16) throw Error(QX_ERR_ARGUMENT, "zero fill of a size not a multiple of 16 bytes");
  launch(stream, [=]() { need(p, bytes, "zero fill"); });
}

}  // namespace qx

// number of kernel launches checked so far (eager and graph replays)
extern "C" int64_t qx_sim_launches(void) { return g_launches; }
