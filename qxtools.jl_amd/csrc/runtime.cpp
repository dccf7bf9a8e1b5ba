// C ABI of libqxhip: contexts, plans, the slice/batch walk, statistics.
//
// The walk replaces QXContexts' executor loop (for each bitstring, for each slice
// assignment: load / view / ncon, then sum; bin/qxrun.jl:83-87,
// docs/src/users_guide.md:11-20) and single_amplitude (src/simulation.jl:86-91):
//   once per call   : steps independent of bitstrings and slices
//   per slice       : bitstring-independent steps that read a view
//   per batch       : fill the output tensors of `batch` bitstrings, run the
//                     bitstring-dependent steps, add the saved tensor into FP64
//                     Neumaier accumulators on the device.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <memory>
#include <set>
#include <sstream>
#include <tuple>

#include "qx_internal.hpp"

struct qx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;   // created by the first plan that runs a vendor BLAS step
  void *blas_ws = nullptr;         // its workspace (the handle allocates nothing itself)
  void *scratch = nullptr;         // device address of the library's fixed scratch area
};

namespace {

thread_local std::string g_last_error;

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw qx::Error(QX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
qx_status guard(F &&f) {
  try {
    f();
    g_last_error.clear();
    return QX_OK;
  } catch (const qx::Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return QX_ERR_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return QX_ERR_ARGUMENT;
  }
}

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  std::vector<char> held;   // the bytes of the last upload(), while the buffer stays
  // returns true when the buffer moved (captured graphs that use it are stale)
  bool reserve(size_t n) {
    if (n <= bytes) return false;
    release();
    if (n == 0) return true;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      throw qx::Error(QX_ERR_MEMORY, "device allocation of " + std::to_string(n) + " bytes failed");
    }
    bytes = n;
    return true;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    held.clear();
  }
  // Copy n bytes from host memory in stream order, unless the buffer already holds
  // exactly them from the previous upload (small per-call tables).
  void upload(const void *src, size_t n, hipStream_t s) {
    if (held.size() == n && n > 0 && std::memcmp(held.data(), src, n) == 0) return;
    hip_check(hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s), "copy");
    held.assign((const char *)src, (const char *)src + n);
  }
};

}  // namespace

struct qx_plan {
  qx_ctx *ctx = nullptr;
  int prec = QX_C128;
  qx::Program prog;
  std::map<std::string, std::vector<double>> host_data;
  int64_t opt_batch = 256;
  int64_t opt_arena_mb = -1;
  int64_t opt_graph = 0;
  int64_t last_launches = 0;   // launches of the last amplitudes / vjp call (estimate)
  int64_t opt_profile = 0;
  int64_t opt_autotune = 0;
  int64_t opt_kernel = -1;                  // >= 0: force that kernel wherever it can run
  int64_t opt_chains = 1;                   // fold runs of gate-block steps into chain passes
  int64_t opt_batch_last = 0;               // batched leg slowest in bitstring-dependent values
  int64_t opt_small_walk = 0;               // small sections' levels in one launch (k_small_walk)
  int64_t opt_chain_variant = -1;           // >= 0: chain passes' 16-value path (bits 0-1),
                                            // swizzled tiles (bit 2), barriers (bit 3); tests
  bool finalized = false;

  std::vector<DevBuf> data;                 // per key
  std::map<std::pair<int64_t, uint64_t>, qx::Lowered> lowered;   // per (batch size, fused slice vars)
  std::map<std::pair<int64_t, uint64_t>, qx::VjpLowered> vjp_lowered;
  std::map<std::pair<int64_t, int>, uint64_t> fuse_choice;          // (batch, vjp) -> fused vars
  int64_t opt_fuse = 1;
  const qx::Lowered *last_lowered = nullptr;   // lowering used by the last qx_amplitudes
  DevBuf varena, cot, temp, grad, ampcot, ws;
  DevBuf arena, out_off, out_q, bits, acc, outbuf, relem_off;
  int64_t arena_cap_bytes = 0;
  std::vector<double> step_ms;
  std::vector<int64_t> step_calls;
  std::vector<hipEvent_t> events;           // pool: 2 per profiled launch
  std::vector<hipEvent_t> tune_events;      // pool: 2 per timed candidate launch
  // descriptors of the batched small steps of one lowering (device table)
  struct SmallTable {
    DevBuf desc;
    void *arena = nullptr;                  // arena base the descriptors point into
    std::vector<int> index;                 // per step: descriptor index, or -1
    std::vector<int> count;                 // per step that starts a level: steps in the level
    // one-workgroup walk of a section's levels (k_small_walk): per section, the offset
    // of its level table in `walk` (device ints) and its level count (0: not walked)
    DevBuf walk;
    int walk_at[3] = {0, 0, 0}, walk_levels[3] = {0, 0, 0};
  };
  std::map<const qx::Lowered *, SmallTable> small;   // per lowering (forward or pullback)
  // batched slice-free backward levels of one pullback lowering
  struct BwdTable {
    DevBuf desc, adds, zeros;
    int zero0[3] = {0, 0, 0}, nzero[3] = {0, 0, 0};   // per section: its zero-fill ranges
    std::vector<char *> bases;              // arena, cot, temp, grad the tables point into
    std::vector<int> first[3];              // per section and backward step: its level's group, or -1
    struct Group { int desc0, ndesc, add0, nadd; };
    std::vector<Group> groups;
  };
  std::map<const qx::VjpLowered *, BwdTable> bwd;
  // captured device work of qx_amplitudes per (batch, nb, slice_begin, slice_end)
  std::map<std::tuple<int64_t, int64_t, int64_t, int64_t>, hipGraphExec_t> graphs;
  // captured device work of qx_amplitudes_vjp per (batch, nb, -, fused slice vars)
  std::map<std::tuple<int64_t, int64_t, int64_t, int64_t>, hipGraphExec_t> vjp_graphs;
  void drop_graphs() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
    for (auto &kv : vjp_graphs) (void)hipGraphExecDestroy(kv.second);
    vjp_graphs.clear();
  }
  std::vector<int> event_step;              // step of each recorded pair this call

  size_t esize() const { return prec == QX_C128 ? 16 : 8; }

  ~qx_plan() {
    for (auto &d : data) d.release();
    arena.release();
    out_off.release();
    out_q.release();
    bits.release();
    acc.release();
    outbuf.release();
    relem_off.release();
    varena.release();
    cot.release();
    temp.release();
    grad.release();
    ampcot.release();
    ws.release();
    for (auto &kv : small) kv.second.desc.release(), kv.second.walk.release();
    for (auto &kv : bwd) kv.second.desc.release(), kv.second.adds.release();
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto e : tune_events) (void)hipEventDestroy(e);
    drop_graphs();
  }
};

namespace {

// Contraction steps a lowered step performs (a chain pass: its members, merged or not).
int fused_steps(const qx::Step &st) {
  if (st.kern == qx::Kern::Outer) return st.outer.members;
  if (st.kern != qx::Kern::Chain) return 1;
  int n = st.chain.nsteps;
  for (int j = 0; j < st.chain.nsteps; ++j) n -= st.chain.st[j].merged;
  return n + (int)st.chain_merge.size();
}

// The chain-pass variant the `chain_variant` option names: the path of steps with 16
// contracted values (bits 0-1: vector FMAs, matrix cores with four or with three real
// products), the swizzled tile addressing (bit 2, where the pass has a non-trivial
// swizzle) and a workgroup barrier after every step (bit 3, else wave-local steps
// where the pass allows them).
void force_chain_variant(qx::Step &st, int64_t v) {
  if (st.kern != qx::Kern::Chain) return;
  st.chain.mfma16 = (int)(v & 3);
  bool swz = false;
  for (int b = 0; b < qx::kChainTileBits; ++b) swz |= st.chain.swz[b] != (1 << b);
  st.chain.swizzled = (v & 4) && swz ? 1 : 0;
  if (v & 8) st.chain.wave_local = 0;
}

const qx::Lowered &ensure_lowered(qx_plan *P, int64_t batch, uint64_t fused = 0) {
  auto key = std::make_pair(batch, fused);
  auto it = P->lowered.find(key);
  if (it != P->lowered.end()) return it->second;
  qx::Lowered L = qx::lower(P->prog, batch, false, P->prec, fused, P->opt_chains != 0, P->opt_batch_last != 0);
  if (P->opt_kernel >= 0)
    for (auto &st : L.steps)
      if (qx::force_kernel(st, P->prec, (int)P->opt_kernel)) L.ws_elems = std::max(L.ws_elems, st.ws_elems);
  if (P->opt_chain_variant >= 0)
    for (auto &st : L.steps) force_chain_variant(st, P->opt_chain_variant);
  size_t need = (size_t)L.arena_elems * P->esize();
  if (P->arena_cap_bytes > 0 && (int64_t)need > P->arena_cap_bytes)
    throw qx::Error(QX_ERR_MEMORY, "arena of " + std::to_string(need >> 20) + " MiB for batch " +
                                       std::to_string(batch) + " exceeds the cap of " +
                                       std::to_string(P->arena_cap_bytes >> 20) + " MiB");
  P->lowered.emplace(key, std::move(L));
  return P->lowered.at(key);
}

struct Bases {
  char *arena = nullptr, *cot = nullptr, *temp = nullptr;
};

int64_t value_offset(const qx::Value &v, const std::vector<int64_t> &varval) {
  int64_t off = v.static_off;
  for (auto &t : v.var_terms) off += varval[t.first] * t.second;
  return off;
}

// device pointer of a value for the given slice-variable assignment
void *value_ptr(qx_plan *P, const qx::Value &v, const std::vector<int64_t> &varval, const Bases &b) {
  const int64_t es = (int64_t)P->esize();
  char *base = nullptr;
  switch (v.buf) {
    case qx::Buf::Data: base = (char *)P->data[v.data_key].p; break;
    case qx::Buf::Arena: base = b.arena + v.slot * es; break;
    case qx::Buf::Cot: base = b.cot + v.slot * es; break;
    case qx::Buf::Temp: base = b.temp + v.slot * es; break;
    default: throw qx::Error(QX_ERR_STATE, "value without storage");
  }
  return base + value_offset(v, varval) * es;
}

void *value_ptr(qx_plan *P, const qx::Value &v, const std::vector<int64_t> &varval) {
  Bases b;
  b.arena = (char *)P->arena.p;
  return value_ptr(P, v, varval, b);
}

constexpr size_t kBlasWorkspace = size_t(64) << 20;

// The context's rocBLAS handle, bound to its stream with a workspace of its own, so
// that its calls allocate nothing and can be captured into a graph.  Must be first
// called outside stream capture.
rocblas_handle blas_handle(qx_ctx *c) {
  if (c->blas) return c->blas;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hip_check(hipStreamIsCapturing(c->stream, &cs), "hipStreamIsCapturing");
  if (cs != hipStreamCaptureStatusNone) throw qx::Error(QX_ERR_STATE, "internal: BLAS handle created under capture");
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw qx::Error(QX_ERR_DEVICE, "rocblas_create_handle failed");
  void *ws = nullptr;
  hipError_t e = hipMalloc(&ws, kBlasWorkspace);
  if (e != hipSuccess) {
    rocblas_destroy_handle(h);
    hip_check(e, "hipMalloc (BLAS workspace)");
  }
  if (rocblas_set_workspace(h, ws, kBlasWorkspace) != rocblas_status_success ||
      rocblas_set_stream(h, c->stream) != rocblas_status_success) {
    rocblas_destroy_handle(h);
    (void)hipFree(ws);
    throw qx::Error(QX_ERR_DEVICE, "rocblas handle set-up failed");
  }
  c->blas = h;
  c->blas_ws = ws;
  return h;
}

// Create the BLAS handle before a capture when any of `steps` needs it.
void blas_before_capture(qx_ctx *c, const std::vector<qx::Step> &steps) {
  for (auto &st : steps)
    if (st.kern == qx::Kern::Blas && !st.blas.own) {
      blas_handle(c);
      return;
    }
}

// C = op(X) op(Y) by rocBLAS (strided batched, no copies): the step's operands are
// plain column-major matrices (see qx::BlasArgs).
void launch_blas(qx_plan *P, const qx::BlasArgs &b, const void *A, const void *B, void *C, void *part = nullptr) {
  if (b.ttgt) {
    // plain copies after the split planes: gather A and B, multiply, gather C (the
    // hand-written GEMM's K-split partials after the copies)
    auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
    const int64_t es = (int64_t)P->esize();
    char *w = (char *)P->ws.p + (b.split ? up(8 * b.ex) + up(8 * b.ey) + 2 * up(4 * b.ec) : 0);
    char *pa = w, *pb = pa + up(b.ta * es), *pc = pb + up(b.tb * es);
    char *pp = pc + up(b.r * b.s * b.batch * es);
    hipStream_t stream = P->ctx->stream;
    qx::launch_gather(P->prec, A, pa, b.nga, b.ga, stream);
    qx::launch_gather(P->prec, B, pb, b.ngb, b.gb, stream);
    qx::BlasArgs inner = b;
    inner.ttgt = 0;
    launch_blas(P, inner, pa, pb, pc, pp);
    qx::launch_gather(P->prec, pc, C, b.ngc, b.gc, stream);
    return;
  }
  const void *X = b.swap ? B : A, *Y = b.swap ? A : B;
  if (b.own) {
    qx::launch_plain_gemm(P->prec, X, Y, C, part ? part : P->ws.p, b, P->ctx->stream);
    return;
  }
  rocblas_handle h = blas_handle(P->ctx);
  const rocblas_operation ox = b.opx ? rocblas_operation_transpose : rocblas_operation_none;
  const rocblas_operation oy = b.opy ? rocblas_operation_transpose : rocblas_operation_none;
  rocblas_status r;
  if (b.split) {
    // complex64 operands split into bf16 hi / lo planes (in their own layouts), three
    // bf16 products per real product accumulated in FP32 planes, then interleaved
    hipStream_t stream = P->ctx->stream;
    auto up = [](int64_t bytes) { return (bytes + 255) / 256 * 256; };
    char *w = (char *)P->ws.p;
    uint16_t *xs = (uint16_t *)w;
    uint16_t *ys = (uint16_t *)(w + up(8 * b.ex));
    float *cre = (float *)(w + up(8 * b.ex) + up(8 * b.ey));
    float *cim = (float *)((char *)cre + up(4 * b.ec));
    qx::launch_bf16_split(X, b.ex, xs, stream);
    qx::launch_bf16_split(Y, b.ey, ys, stream);
    auto xp = [&](int part) { return xs + part * b.ex; };   // 0 re hi, 1 re lo, 2 im hi, 3 im lo
    auto yp = [&](int part) { return ys + part * b.ey; };
    const float one = 1.0f, minus = -1.0f, zero = 0.0f;
    auto gemm = [&](const uint16_t *x, const uint16_t *y, float *c, bool neg, bool first) {
      rocblas_status st = rocblas_gemm_strided_batched_ex(
          h, ox, oy, (rocblas_int)b.r, (rocblas_int)b.s, (rocblas_int)b.k, neg ? &minus : &one, x,
          rocblas_datatype_bf16_r, (rocblas_int)b.ldx, b.sx, y, rocblas_datatype_bf16_r, (rocblas_int)b.ldy, b.sy,
          first ? &zero : &one, c, rocblas_datatype_f32_r, (rocblas_int)b.ldc, b.sc, c, rocblas_datatype_f32_r,
          (rocblas_int)b.ldc, b.sc, (rocblas_int)b.batch, rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0);
      if (st != rocblas_status_success)
        throw qx::Error(QX_ERR_DEVICE, std::string("rocBLAS bf16 gemm failed: ") + rocblas_status_to_string(st));
    };
    // one real product u . v as hi.hi + hi.lo + lo.hi (parts: hi = base, lo = base + 1)
    auto product = [&](int u, int v, float *c, bool neg, bool first) {
      gemm(xp(u), yp(v), c, neg, first);
      gemm(xp(u), yp(v + 1), c, neg, false);
      gemm(xp(u + 1), yp(v), c, neg, false);
    };
    product(0, 0, cre, false, true);    // re = Xr Yr - Xi Yi
    product(2, 2, cre, true, false);
    product(0, 2, cim, false, true);    // im = Xr Yi + Xi Yr
    product(2, 0, cim, false, false);
    qx::launch_planar_merge(cre, cim, b.ec, C, stream);
    return;
  }
  if (P->prec == QX_C128) {
    const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
    r = rocblas_zgemm_strided_batched(h, ox, oy, b.r, b.s, b.k, &one, (const rocblas_double_complex *)X, b.ldx,
                                      b.sx, (const rocblas_double_complex *)Y, b.ldy, b.sy, &zero,
                                      (rocblas_double_complex *)C, b.ldc, b.sc, b.batch);
  } else {
    const rocblas_float_complex one{1.0f, 0.0f}, zero{0.0f, 0.0f};
    r = rocblas_cgemm_strided_batched(h, ox, oy, b.r, b.s, b.k, &one, (const rocblas_float_complex *)X, b.ldx,
                                      b.sx, (const rocblas_float_complex *)Y, b.ldy, b.sy, &zero,
                                      (rocblas_float_complex *)C, b.ldc, b.sc, b.batch);
  }
  if (r != rocblas_status_success)
    throw qx::Error(QX_ERR_DEVICE, std::string("rocBLAS gemm failed: ") + rocblas_status_to_string(r));
}

void launch_step(qx_plan *P, const qx::Lowered &L, const qx::Step &st, void *A, void *B, void *C,
                 const std::vector<int64_t> &varval, const Bases *bs = nullptr) {
  hipStream_t stream = P->ctx->stream;
  switch (st.kern) {
    case qx::Kern::Blas: launch_blas(P, st.blas, A, B, C); break;
    case qx::Kern::Chain: {
      qx::ChainArgs a = st.chain;
      for (int j = 0; j < a.nsteps; ++j) {
        const qx::Value &v = L.values[st.chain_small[j]];
        a.S[j] = bs ? value_ptr(P, v, varval, *bs) : value_ptr(P, v, varval);
      }
      qx::ChainMergeArgs m;
      for (size_t i = 0; i < st.chain_merge.size(); ++i) {
        const auto &cm = st.chain_merge[i];
        if (m.ngroups == 0 || m.g[m.ngroups - 1].step != cm.step) {
          auto &g = m.g[m.ngroups++];
          g.step = cm.step;
          g.k = st.chain.st[cm.step].k;
          g.soff = st.chain.st[cm.step].soff;
          g.first = (int32_t)i;
        }
        m.g[m.ngroups - 1].count++;
        auto &mm = m.m[i];
        const qx::Value &v = L.values[cm.small];
        mm.S = bs ? value_ptr(P, v, varval, *bs) : value_ptr(P, v, varval);
        mm.kc = cm.kc;
        for (int q = 0; q < 4; ++q) mm.pos[q] = cm.pos[q], mm.sk[q] = cm.sk[q], mm.sn[q] = cm.sn[q];
      }
      qx::launch_contract_chain(P->prec, A, C, P->ws.p, a, m, stream);
      break;
    }
    case qx::Kern::Outer: {
      qx::OuterArgs a = st.outer;
      for (int f = 0; f < a.nf; ++f) {
        const qx::Value &v = L.values[st.outer_in[f]];
        a.F[f] = bs ? value_ptr(P, v, varval, *bs) : value_ptr(P, v, varval);
      }
      qx::launch_outer(P->prec, C, st.c_elems, a, stream);
      break;
    }
    case qx::Kern::Radix: qx::launch_contract_radix(P->prec, A, B, C, st.c_elems, st.radix, stream); break;
    case qx::Kern::Gemm: qx::launch_contract_gemm(P->prec, A, B, C, P->ws.p, st.gemm, stream); break;
    case qx::Kern::Reduce: qx::launch_contract_reduce(P->prec, A, B, C, P->ws.p, st.c_elems, st.bits, stream); break;
    case qx::Kern::Rows: qx::launch_contract_rows(P->prec, A, B, C, st.rows, stream); break;
    default: qx::launch_contract_bits(P->prec, A, B, C, P->ws.p, st.c_elems, st.bits, stream); break;
  }
}

// Append x to a small-step descriptor table, its C split into workgroup ranges of
// kSmallChunk elements; returns the number of descriptors added.
int push_small(std::vector<qx::SmallDesc> &d, qx::SmallDesc x) {
  // chunks of C (powers of two) sized for about 32 multiply-adds per lane of the
  // workgroup: a chunk of fewer elements than lanes splits K over its lanes
  const int nc = 1 << x.nbc;
  const int64_t nk = int64_t(1) << x.nbk;
  int len = qx::kSmallChunk;
  while (len > 1 && len * nk > 32 * qx::kSmallThreads) len >>= 1;
  len = std::min(len, nc);
  int n = 0;
  for (int c0 = 0; c0 < nc; c0 += len, ++n) {
    x.c0 = c0;
    x.c1 = c0 + len;
    d.push_back(x);
  }
  return n;
}

// Build and upload the descriptor table of L's batched small steps (slice-free, so
// their operand addresses depend only on the arena and data buffers).
void prepare_small(qx_plan *P, const qx::Lowered &L, const Bases &bs) {
  qx_plan::SmallTable &T = P->small[&L];
  if (T.arena == bs.arena && !T.index.empty()) return;
  std::vector<qx::SmallDesc> d;
  T.index.assign(L.steps.size(), -1);
  T.count.assign(L.steps.size(), 0);
  const std::vector<int64_t> zero(P->prog.vars.size(), 0);
  int first = -1;
  for (size_t s = 0; s < L.steps.size(); ++s) {
    const qx::Step &st = L.steps[s];
    if (st.level < 0) continue;
    if (first < 0 || L.steps[first].level != st.level) first = (int)s;
    T.index[s] = (int)d.size();
    qx::SmallDesc x;
    x.A = value_ptr(P, L.values[st.in_a], zero, bs);
    x.B = value_ptr(P, L.values[st.in_b], zero, bs);
    x.C = value_ptr(P, L.values[st.node], zero, bs);
    x.conj_a = st.bits.conj_a;
    x.conj_b = st.bits.conj_b;
    for (auto &g : st.cl)
      for (int64_t m = 1; m < g.dim; m <<= 1) {
        if (x.nbc >= qx::kSmallMaxBits) throw qx::Error(QX_ERR_STATE, "internal: small step bits");
        x.wa[x.nbc] = (int32_t)(g.sa * m);
        x.wb[x.nbc] = (int32_t)(g.sb * m);
        ++x.nbc;
      }
    for (auto &g : st.kl)
      for (int64_t m = 1; m < g.dim; m <<= 1) {
        if (x.nbk >= qx::kSmallMaxBits) throw qx::Error(QX_ERR_STATE, "internal: small step bits");
        x.ka[x.nbk] = (int32_t)(g.sa * m);
        x.kb[x.nbk] = (int32_t)(g.sb * m);
        ++x.nbk;
      }
    T.count[first] += push_small(d, x);
  }
  if (!d.empty()) {
    T.desc.reserve(d.size() * sizeof(qx::SmallDesc));
    hip_check(hipMemcpy(T.desc.p, d.data(), d.size() * sizeof(qx::SmallDesc), hipMemcpyHostToDevice),
              "copy descriptors");
  }
  // level tables of the sections whose small steps are few enough for one workgroup:
  // at most 2^22 multiply-adds and 4096 chunks in all (the levels' descriptors are
  // contiguous, in level order)
  std::vector<int> walk;
  for (int sec = 0; sec < 3; ++sec) {
    T.walk_at[sec] = (int)walk.size();
    T.walk_levels[sec] = 0;
    std::vector<int> offs;
    double macs = 0;
    int chunks = 0;
    bool other_seen = false;   // a step of the section outside the levels, before a level
    for (size_t s = 0; s < L.steps.size(); ++s) {
      if (L.steps[s].section != sec) continue;
      if (L.steps[s].level < 0) {
        other_seen = true;
        continue;
      }
      if (T.count[s] == 0) continue;
      if (other_seen || (!offs.empty() && offs.back() != T.index[s])) {
        offs.clear();   // levels of a section are contiguous; anything else keeps one launch per level
        break;
      }
      if (offs.empty()) offs.push_back(T.index[s]);
      for (int i = T.index[s]; i < T.index[s] + T.count[s]; ++i) {
        macs += double(d[i].c1 - d[i].c0) * double(int64_t(1) << d[i].nbk);
        ++chunks;
      }
      offs.push_back(T.index[s] + T.count[s]);
    }
    if (offs.size() >= 2 && macs <= double(int64_t(1) << 22) && chunks <= 4096) {
      walk.insert(walk.end(), offs.begin(), offs.end());
      T.walk_levels[sec] = (int)offs.size() - 1;
    }
  }
  if (!walk.empty()) {
    T.walk.reserve(walk.size() * sizeof(int));
    hip_check(hipMemcpy(T.walk.p, walk.data(), walk.size() * sizeof(int), hipMemcpyHostToDevice), "copy level table");
  }
  T.arena = bs.arena;
  P->drop_graphs();
}

// Whether section `sec` of L runs its small levels as one walk, and that launch.
bool walks(const qx_plan *P, const qx_plan::SmallTable &T, int sec) {
  return P->opt_small_walk && T.walk_levels[sec] > 0;
}

void run_small_walk(qx_plan *P, const qx::Lowered &L, size_t s, int sec, bool timed = true) {
  hipStream_t stream = P->ctx->stream;
  const qx_plan::SmallTable &T = P->small.at(&L);
  size_t ev = 0;
  timed = timed && P->opt_profile;
  if (timed) {
    ev = P->event_step.size();
    while (P->events.size() < 2 * (ev + 1)) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      P->events.push_back(e);
    }
    P->event_step.push_back((int)s);
    hip_check(hipEventRecord(P->events[2 * ev], stream), "hipEventRecord");
  }
  qx::launch_small_walk(P->prec, (const qx::SmallDesc *)T.desc.p, (const int *)T.walk.p + T.walk_at[sec],
                        T.walk_levels[sec], stream);
  if (timed) hip_check(hipEventRecord(P->events[2 * ev + 1], stream), "hipEventRecord");
}

// Launch the level of batched small steps that starts at step s.
void run_small_level(qx_plan *P, const qx::Lowered &L, size_t s, bool timed = true) {
  hipStream_t stream = P->ctx->stream;
  const qx_plan::SmallTable &T = P->small.at(&L);
  size_t ev = 0;
  timed = timed && P->opt_profile;
  if (timed) {
    ev = P->event_step.size();
    while (P->events.size() < 2 * (ev + 1)) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      P->events.push_back(e);
    }
    P->event_step.push_back((int)s);
    hip_check(hipEventRecord(P->events[2 * ev], stream), "hipEventRecord");
  }
  const qx::SmallDesc *d = (const qx::SmallDesc *)T.desc.p + T.index[s];
  qx::launch_small_batch(P->prec, d, T.count[s], stream);
  if (timed) hip_check(hipEventRecord(P->events[2 * ev + 1], stream), "hipEventRecord");
}

void run_step(qx_plan *P, const qx::Lowered &L, size_t s, const std::vector<int64_t> &varval, bool timed = true) {
  const qx::Step &st = L.steps[s];
  void *A = value_ptr(P, L.values[st.in_a], varval);
  void *B = value_ptr(P, L.values[st.in_b], varval);
  void *C = value_ptr(P, L.values[st.node], varval);
  hipStream_t stream = P->ctx->stream;
  size_t ev = 0;
  timed = timed && P->opt_profile;
  if (timed) {
    ev = P->event_step.size();
    while (P->events.size() < 2 * (ev + 1)) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      P->events.push_back(e);
    }
    P->event_step.push_back((int)s);
    hip_check(hipEventRecord(P->events[2 * ev], stream), "hipEventRecord");
  }
  launch_step(P, L, st, A, B, C, varval);
  if (timed) hip_check(hipEventRecord(P->events[2 * ev + 1], stream), "hipEventRecord");
}

std::string kern_variant(const qx::Step &st);

// Distinct name of a candidate configuration (kernel plus launch choices).
std::string cand_label(const qx::Step &st) {
  std::string v = kern_variant(st);
  if (st.swapped) v += "/swap";
  if (st.kern == qx::Kern::Chain)
    v += "/g" + std::to_string(st.chain.grid) + (st.chain.mfma16 == 2 ? "/mfma3" : st.chain.mfma16 ? "/mfma" : "/fma") + (st.chain.swizzled ? "/swz" : "") +
         (st.chain.wave_local ? "/wl" : "");
  return v;
}

// The autotuner's state for one call: every step's candidate configurations and
// the launches timed so far.  Candidates of a step all write the same C, so
// running them one after another leaves the results unchanged.
struct Tuner {
  static constexpr int kTimedLaunches = 4;   // timed launches per candidate per step
  std::vector<std::vector<qx::Step>> cands;
  std::vector<std::vector<double>> ms;
  std::vector<int> seen;
  std::vector<std::array<int, 3>> rec;        // (step, candidate, event pair)
};

void tune_step(qx_plan *P, const qx::Lowered &L, size_t s, const std::vector<int64_t> &varval, Tuner &T) {
  const qx::Step &st = L.steps[s];
  void *A = value_ptr(P, L.values[st.in_a], varval);
  void *B = value_ptr(P, L.values[st.in_b], varval);
  void *C = value_ptr(P, L.values[st.node], varval);
  const auto &cs = T.cands[s];
  if (cs.size() < 2 || T.seen[s] >= Tuner::kTimedLaunches) {
    launch_step(P, L, st, A, B, C, varval);
    return;
  }
  hipStream_t stream = P->ctx->stream;
  if (T.seen[s] == 0)   // first launch of every candidate untimed: code loading, cold caches
    for (auto &c : cs) launch_step(P, L, c, A, B, C, varval);
  ++T.seen[s];
  for (size_t c = 0; c < cs.size(); ++c) {
    size_t ev = T.rec.size();
    while (P->tune_events.size() < 2 * (ev + 1)) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      P->tune_events.push_back(e);
    }
    T.rec.push_back({(int)s, (int)c, (int)ev});
    hip_check(hipEventRecord(P->tune_events[2 * ev], stream), "hipEventRecord");
    launch_step(P, L, cs[c], A, B, C, varval);
    hip_check(hipEventRecord(P->tune_events[2 * ev + 1], stream), "hipEventRecord");
  }
}

// After the tuning call has drained: keep each step's fastest candidate (the
// current choice on ties) and record what was measured.
void finish_tuning(qx_plan *P, qx::Lowered &L, Tuner &T) {
  for (auto &r : T.rec) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, P->tune_events[2 * r[2]], P->tune_events[2 * r[2] + 1]),
              "hipEventElapsedTime");
    T.ms[r[0]][r[1]] += ms;
  }
  L.tune_log.assign(L.steps.size(), std::string());
  for (size_t s = 0; s < L.steps.size(); ++s) {
    if (T.cands[s].size() < 2 || T.seen[s] == 0) continue;
    size_t best = 0;
    std::ostringstream os;
    os.precision(4);
    for (size_t c = 0; c < T.cands[s].size(); ++c) {
      if (T.ms[s][c] < T.ms[s][best]) best = c;
      os << (c ? "," : "") << "\"" << cand_label(T.cands[s][c]) << "\":" << T.ms[s][c] / T.seen[s];
    }
    L.tune_log[s] = os.str();
    L.steps[s] = T.cands[s][best];
    L.ws_elems = std::max(L.ws_elems, L.steps[s].ws_elems);
  }
  L.tuned = true;
  P->drop_graphs();
}

// After the stream has drained: add every recorded launch's time to its step.
void collect_profile(qx_plan *P, const qx::Lowered &L) {
  if (!P->opt_profile) return;
  if (P->step_ms.size() != L.steps.size()) {
    P->step_ms.assign(L.steps.size(), 0.0);
    P->step_calls.assign(L.steps.size(), 0);
  }
  for (size_t i = 0; i < P->event_step.size(); ++i) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, P->events[2 * i], P->events[2 * i + 1]), "hipEventElapsedTime");
    P->step_ms[P->event_step[i]] += ms;
    P->step_calls[P->event_step[i]] += 1;
  }
  P->event_step.clear();
}

// Slice variables to contract inline rather than loop over: the first j variables
// for the largest j whose lowered program (arena + workspace, or for the vjp the
// kept forward values + cotangents) fits half the memory cap.  Slicing bounds
// memory; when memory does not bind, fewer and larger launches are faster and the
// sum over the fused variables happens inside the contractions (by linearity).
uint64_t choose_fused(qx_plan *P, int64_t batch, bool for_vjp) {
  const int nv = (int)P->prog.vars.size();
  if (!P->opt_fuse || nv == 0) return 0;
  auto key = std::make_pair(batch, (int)for_vjp);
  auto it = P->fuse_choice.find(key);
  if (it != P->fuse_choice.end()) return it->second;
  const int64_t budget = P->arena_cap_bytes / 2;
  const int64_t es = (int64_t)P->esize();
  uint64_t pick = 0;
  for (int j = std::min(nv, 63); j > 0; --j) {
    uint64_t mask = (uint64_t(1) << j) - 1;
    try {
      int64_t need;
      if (for_vjp) {
        qx::VjpLowered V = qx::lower_vjp(P->prog, batch, P->prec, mask, P->opt_batch_last != 0);
        need = (V.fwd.arena_elems + V.fwd.ws_elems + V.cot_elems + V.temp_elems) * es;
      } else {
        qx::Lowered L = qx::lower(P->prog, batch, false, P->prec, mask, P->opt_chains != 0, P->opt_batch_last != 0);
        need = (L.arena_elems + L.ws_elems) * es;
      }
      if (need <= budget) {
        pick = mask;
        break;
      }
    } catch (const qx::Error &e) {
      if (e.code != QX_ERR_UNSUPPORTED) throw;
    }
  }
  P->fuse_choice[key] = pick;
  return pick;
}

// number of looped slices and the slice variables' values of looped slice `sl`
int64_t looped_slices(const qx_plan *P, uint64_t fused) {
  int64_t n = 1;
  for (size_t j = 0; j < P->prog.vars.size(); ++j)
    if (!((fused >> j) & 1)) n *= P->prog.var_dims[j];
  return n;
}

void slice_values(const qx_plan *P, uint64_t fused, int64_t sl, std::vector<int64_t> &varval) {
  for (size_t j = 0; j < P->prog.vars.size(); ++j) {
    if ((fused >> j) & 1) {
      varval[j] = 0;
      continue;
    }
    varval[j] = sl % P->prog.var_dims[j];
    sl /= P->prog.var_dims[j];
  }
}

// Launches of one section of a lowered program per pass: unbatched steps, plus one
// per level of batched small steps.
int64_t section_launches(const qx::Lowered &L, int sec) {
  int64_t n = 0;
  std::set<int> levels;
  for (auto &st : L.steps)
    if (st.section == sec) {
      if (st.level >= 0) levels.insert(st.level);
      else ++n;
    }
  return n + (int64_t)levels.size();
}

// A call of more launches than this runs without graph capture: capturing and
// instantiating a graph of that many nodes costs host time and memory that its replay
// cannot win back, since such a call runs for a long time anyway.
constexpr int64_t kMaxGraphLaunches = int64_t(1) << 15;

void check_bits(qx_plan *P, const char *bits, int64_t nb) {
  const int64_t nch = P->prog.num_chars;
  for (int64_t i = 0; i < nb * nch; ++i) {
    char c = bits[i];
    if (c != '0' && c != '1' && c != '+' && c != '-')
      throw qx::Error(QX_ERR_ARGUMENT, "bitstring characters must be 0, 1, + or -");
  }
}

// Device buffers of one amplitudes call (besides the plan's arena, workspace and data).
struct CallBufs {
  void *acc = nullptr;        // nb * R FP64 Neumaier (sum, compensation) pairs, complex
  void *outbuf = nullptr;     // nb * R complex128 results
  void *out_off = nullptr;    // per output tensor: its arena slot (int64)
  void *out_q = nullptr;      // per output tensor: its qubit (int32)
  void *relem_off = nullptr;  // per result element: offset in the saved tensor (int64)
  void *bits = nullptr;       // nb * num_chars bitstring characters
};

// The device work of one amplitudes call, from zeroing the accumulators to the final
// sum: steps of section 0 once, per looped slice those of section 1, per batch of
// bitstrings the output fill, section 2 and the accumulation into FP64 pairs.
// `timed`: record per-launch events when the plan's "profile" option is on.
void enqueue_amplitudes(qx_plan *P, const qx::Lowered &L, uint64_t fused, int64_t s0, int64_t s1, int64_t nb,
                        int64_t batch, const CallBufs &cb, Tuner *tuner, bool timed) {
  hipStream_t stream = P->ctx->stream;
  const int64_t nch = P->prog.num_chars;
  const int64_t R = L.result_elems;
  const int nout = (int)P->prog.outputs.size();
  const qx::Value &rv = L.values[P->prog.save];
  qx::launch_zero(cb.acc, nb * R * 32, stream);
  const size_t nv = P->prog.vars.size();
  std::vector<int64_t> varval(nv, 0);
  const qx_plan::SmallTable &ST = P->small.at(&L);
  auto run_section = [&](int sec) {
    bool walked = false;
    for (size_t s = 0; s < L.steps.size(); ++s) {
      if (L.steps[s].section != sec) continue;
      if (L.steps[s].level >= 0) {   // batched small steps: one launch per level, or one walk
        if (ST.count[s] == 0) continue;
        if (!walks(P, ST, sec)) {
          run_small_level(P, L, s, timed);
        } else if (!walked) {
          run_small_walk(P, L, s, sec, timed);
          walked = true;
        }
        continue;
      }
      if (tuner)
        tune_step(P, L, s, varval, *tuner);
      else
        run_step(P, L, s, varval, timed);
    }
  };
  run_section(0);
  for (int64_t sl = s0; sl < s1; ++sl) {
    slice_values(P, fused, sl, varval);
    run_section(1);
    for (int64_t row0 = 0; row0 < nb; row0 += batch) {
      int64_t nvalid = std::min<int64_t>(batch, nb - row0);
      qx::launch_fill_outputs(P->prec, P->arena.p, (const int64_t *)cb.out_off, nout, (const int *)cb.out_q,
                              (const char *)cb.bits + row0 * nch, nvalid, batch, nch, stream);
      run_section(2);
      qx::launch_accumulate(P->prec, value_ptr(P, rv, varval), rv.dep ? rv.beta_stride : 0, batch, nvalid, R,
                            (const int64_t *)cb.relem_off, (double *)cb.acc, row0, stream);
    }
  }
  qx::launch_finalize((const double *)cb.acc, (double *)cb.outbuf, nb * R, stream);
}

// Amplitudes of nb bitstrings summed over slices [s0, s1).  `bits` null: the
// bitstrings are already in P->bits on the device (sampling); `out` null: the
// results stay in P->outbuf.
void amplitudes(qx_plan *P, const char *bits, int64_t nb, int64_t s0, int64_t s1, double *out,
                bool device_bits = false) {
  if (!P->finalized) throw qx::Error(QX_ERR_STATE, "plan is not finalized");
  if (nb <= 0) throw qx::Error(QX_ERR_ARGUMENT, "need at least one bitstring");
  if ((!bits && !device_bits) || (!out && !device_bits))
    throw qx::Error(QX_ERR_ARGUMENT, "null bitstring or output pointer");
  int64_t nslices = 1;
  for (int64_t d : P->prog.var_dims) nslices *= d;
  if (s0 < 0 || s1 > nslices || s0 >= s1) throw qx::Error(QX_ERR_ARGUMENT, "slice range out of bounds");
  const int64_t nch = P->prog.num_chars;
  if (bits) check_bits(P, bits, nb);
  int64_t batch = 1;
  while (batch < nb && batch < P->opt_batch) batch <<= 1;
  const uint64_t fused = (s0 == 0 && s1 == nslices) ? choose_fused(P, batch, false) : 0;
  if (fused) {
    s0 = 0;
    s1 = looped_slices(P, fused);
  }
  qx::Lowered &L = const_cast<qx::Lowered &>(ensure_lowered(P, batch, fused));
  if (P->last_lowered != &L) {
    P->step_ms.clear();
    P->step_calls.clear();
  }
  P->last_lowered = &L;
  hipStream_t stream = P->ctx->stream;
  const size_t es = P->esize();
  std::unique_ptr<Tuner> tuner;
  int64_t ws_elems = L.ws_elems;
  if (P->opt_autotune && !L.tuned) {
    tuner = std::make_unique<Tuner>();
    for (auto &st : L.steps) {
      tuner->cands.push_back(qx::kernel_candidates(st, P->prec));
      tuner->ms.emplace_back(tuner->cands.back().size(), 0.0);
      for (auto &c : tuner->cands.back()) ws_elems = std::max(ws_elems, c.ws_elems);
    }
    tuner->seen.assign(L.steps.size(), 0);
  }
  bool moved = P->arena.reserve(std::max<size_t>((size_t)L.arena_elems * es, 256));
  moved |= P->ws.reserve(std::max<size_t>((size_t)ws_elems * es, 256));

  // output tables
  std::vector<int64_t> ooff;
  std::vector<int> oq;
  for (int o : P->prog.outputs) {
    ooff.push_back(L.values[o].slot);
    oq.push_back(P->prog.nodes[o].qubit);
  }
  const int nout = (int)ooff.size();
  moved |= P->out_off.reserve(std::max<size_t>(ooff.size() * 8, 16));
  moved |= P->out_q.reserve(std::max<size_t>(oq.size() * 4, 16));
  if (nout) {
    P->out_off.upload(ooff.data(), ooff.size() * 8, stream);
    P->out_q.upload(oq.data(), oq.size() * 4, stream);
  }
  // root element offsets (logical column-major order of the saved tensor)
  const qx::Value &rv = L.values[P->prog.save];
  const int64_t R = L.result_elems;
  std::vector<int64_t> roff(R);
  for (int64_t e = 0; e < R; ++e) {
    int64_t r = e, o = 0;
    for (size_t j = 0; j < rv.dims.size(); ++j) {
      int64_t d = r % rv.dims[j];
      r /= rv.dims[j];
      o += d * rv.strides[j];
    }
    roff[e] = o;
  }
  moved |= P->relem_off.reserve(std::max<size_t>(R * 8, 16));
  P->relem_off.upload(roff.data(), R * 8, stream);
  size_t nbits = (size_t)nb * nch;
  if (bits) {
    moved |= P->bits.reserve(std::max<size_t>(nbits, 16));
    if (nbits) hip_check(hipMemcpyAsync(P->bits.p, bits, nbits, hipMemcpyHostToDevice, stream), "copy");
  } else if (P->bits.bytes < nbits) {
    throw qx::Error(QX_ERR_STATE, "internal: device bitstrings not staged");
  }
  moved |= P->acc.reserve((size_t)nb * R * 32);
  moved |= P->outbuf.reserve((size_t)nb * R * 16);
  if (moved) P->drop_graphs();
  {
    Bases bs;
    bs.arena = (char *)P->arena.p;
    prepare_small(P, L, bs);
  }

  P->event_step.clear();

  CallBufs cb;
  cb.acc = P->acc.p, cb.outbuf = P->outbuf.p, cb.out_off = P->out_off.p, cb.out_q = P->out_q.p;
  cb.relem_off = P->relem_off.p, cb.bits = P->bits.p;
  auto enqueue = [&]() { enqueue_amplitudes(P, L, fused, s0, s1, nb, batch, cb, tuner.get(), true); };
  const int64_t nrows = (nb + batch - 1) / batch;
  const int64_t launches = 2 + section_launches(L, 0) +
                           (s1 - s0) * (section_launches(L, 1) + nrows * (section_launches(L, 2) + 2));
  P->last_launches = launches;
  if (P->opt_graph && !P->opt_profile && !tuner && launches <= kMaxGraphLaunches) {
    auto key = std::make_tuple(batch, nb, s0, s1 + (fused ? (int64_t(1) << 40) : 0));
    auto it = P->graphs.find(key);
    if (it == P->graphs.end()) {
      hipGraph_t g = nullptr;
      blas_before_capture(P->ctx, L.steps);
      hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue();
      } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_check(e, "hipGraphInstantiate");
      it = P->graphs.emplace(key, ex).first;
    }
    hip_check(hipGraphLaunch(it->second, stream), "hipGraphLaunch");
  } else {
    enqueue();
  }
  if (out) hip_check(hipMemcpyAsync(out, P->outbuf.p, (size_t)nb * R * 16, hipMemcpyDeviceToHost, stream), "copy out");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (tuner) finish_tuning(P, L, *tuner);
  collect_profile(P, L);
}

// qx_amplitudes without device allocations: the plan's batch option as the batch
// size, its arena, workspace, small-step tables (and BLAS handle) as an earlier
// qx_amplitudes call left them, and the per-call buffers in the library's fixed
// scratch area.  No graph capture, tuning or event timing.
void amplitudes_noalloc(qx_plan *P, const char *bits, int64_t nb, int64_t s0, int64_t s1, double *out) {
  if (!P->finalized) throw qx::Error(QX_ERR_STATE, "plan is not finalized");
  if (nb <= 0) throw qx::Error(QX_ERR_ARGUMENT, "need at least one bitstring");
  if (!bits || !out) throw qx::Error(QX_ERR_ARGUMENT, "null bitstring or output pointer");
  int64_t nslices = 1;
  for (int64_t d : P->prog.var_dims) nslices *= d;
  if (s0 < 0 || s1 > nslices || s0 >= s1) throw qx::Error(QX_ERR_ARGUMENT, "slice range out of bounds");
  check_bits(P, bits, nb);
  const int64_t batch = P->opt_batch;
  const uint64_t fused = (s0 == 0 && s1 == nslices) ? choose_fused(P, batch, false) : 0;
  if (fused) {
    s0 = 0;
    s1 = looped_slices(P, fused);
  }
  const qx::Lowered &L = ensure_lowered(P, batch, fused);
  const size_t es = P->esize();
  const char *warm = ": run one qx_amplitudes call of at least the plan's batch option in bitstrings first";
  if ((size_t)L.arena_elems * es > P->arena.bytes || (size_t)L.ws_elems * es > P->ws.bytes || !P->arena.p)
    throw qx::Error(QX_ERR_STATE, std::string("plan memory not sized for this call") + warm);
  auto sm = P->small.find(&L);
  if (sm == P->small.end() || sm->second.arena != P->arena.p || sm->second.index.size() != L.steps.size())
    throw qx::Error(QX_ERR_STATE, std::string("small-step tables not built for this call") + warm);
  for (auto &st : L.steps)
    if (st.kern == qx::Kern::Blas && !st.blas.own && !P->ctx->blas)
      throw qx::Error(QX_ERR_STATE, std::string("library GEMM handle not created") + warm);
  if (!P->ctx->scratch) throw qx::Error(QX_ERR_STATE, "context has no scratch area");
  // per-call buffers, 256-byte aligned, in the scratch area
  const int64_t nch = P->prog.num_chars;
  const int64_t R = L.result_elems;
  std::vector<int64_t> ooff;
  std::vector<int> oq;
  for (int o : P->prog.outputs) {
    ooff.push_back(L.values[o].slot);
    oq.push_back(P->prog.nodes[o].qubit);
  }
  const qx::Value &rv = L.values[P->prog.save];
  std::vector<int64_t> roff(R);
  for (int64_t e = 0; e < R; ++e) {
    int64_t r = e, o = 0;
    for (size_t j = 0; j < rv.dims.size(); ++j) {
      o += (r % rv.dims[j]) * rv.strides[j];
      r /= rv.dims[j];
    }
    roff[e] = o;
  }
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_acc = up((size_t)nb * R * 32), b_out = up((size_t)nb * R * 16), b_roff = up((size_t)R * 8),
               b_ooff = up(ooff.size() * 8), b_oq = up(oq.size() * 4), b_bits = up((size_t)nb * nch);
  const size_t total = b_acc + b_out + b_roff + b_ooff + b_oq + b_bits;
  if (total > qx::kScratchBytes)
    throw qx::Error(QX_ERR_ARGUMENT, "call needs " + std::to_string(total) + " bytes of scratch (nb * result_elems = " +
                                         std::to_string(nb * R) + "), over the fixed " +
                                         std::to_string(qx::kScratchBytes) + " bytes");
  char *base = (char *)P->ctx->scratch;
  CallBufs cb;
  cb.acc = base;
  cb.outbuf = base + b_acc;
  cb.relem_off = (char *)cb.outbuf + b_out;
  cb.out_off = (char *)cb.relem_off + b_roff;
  cb.out_q = (char *)cb.out_off + b_ooff;
  cb.bits = (char *)cb.out_q + b_oq;
  hipStream_t stream = P->ctx->stream;
  hip_check(hipMemcpyAsync(cb.relem_off, roff.data(), (size_t)R * 8, hipMemcpyHostToDevice, stream), "copy");
  if (!ooff.empty()) {
    hip_check(hipMemcpyAsync(cb.out_off, ooff.data(), ooff.size() * 8, hipMemcpyHostToDevice, stream), "copy");
    hip_check(hipMemcpyAsync(cb.out_q, oq.data(), oq.size() * 4, hipMemcpyHostToDevice, stream), "copy");
  }
  if (nb * nch > 0) hip_check(hipMemcpyAsync(cb.bits, bits, (size_t)nb * nch, hipMemcpyHostToDevice, stream), "copy");
  enqueue_amplitudes(P, L, fused, s0, s1, nb, batch, cb, nullptr, false);
  hip_check(hipMemcpyAsync(out, cb.outbuf, (size_t)nb * R * 16, hipMemcpyDeviceToHost, stream), "copy out");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// Stage nb Philox bitstrings (candidates j0 .. j0+nb-1 of stream `seed`) in P->bits.
void stage_bits(qx_plan *P, int64_t nb, uint64_t seed, int64_t j0) {
  const int nch = (int)P->prog.num_chars;
  if (P->bits.reserve(std::max<size_t>((size_t)nb * nch, 16))) P->drop_graphs();
  qx::launch_gen_bits((char *)P->bits.p, nb, nch, seed, j0, P->ctx->stream);
}

void check_sampling(const qx_plan *P) {
  if (!P->finalized) throw qx::Error(QX_ERR_STATE, "plan is not finalized");
  if (P->prog.num_chars > 256) throw qx::Error(QX_ERR_UNSUPPORTED, "sampling supports at most 256 qubits");
}

void amplitudes_uniform(qx_plan *P, int64_t num, uint64_t seed, char *bits_out, double *out) {
  check_sampling(P);
  if (num <= 0) throw qx::Error(QX_ERR_ARGUMENT, "need at least one bitstring");
  if (!bits_out || !out) throw qx::Error(QX_ERR_ARGUMENT, "null output pointer");
  int64_t nslices = 1;
  for (int64_t d : P->prog.var_dims) nslices *= d;
  stage_bits(P, num, seed, 0);
  amplitudes(P, nullptr, num, 0, nslices, out, true);
  const size_t nbits = (size_t)num * P->prog.num_chars;
  if (nbits) hip_check(hipMemcpy(bits_out, P->bits.p, nbits, hipMemcpyDeviceToHost), "copy out");
}

// Rejection sampling in rounds of one batch of candidates: stage the round's
// bitstrings, contract, run the sequential acceptance test on the device; only the
// counters come back to the host between rounds.
void sample_rejection(qx_plan *P, int64_t num_samples, uint64_t seed, double M, int fix_m, int64_t max_candidates,
                      char *bits_out, double *amps_out, double *m_out, int64_t *cand_out) {
  check_sampling(P);
  if (num_samples < 0 || max_candidates < 0) throw qx::Error(QX_ERR_ARGUMENT, "negative count");
  if (!(M > 0) || !std::isfinite(M)) throw qx::Error(QX_ERR_ARGUMENT, "M must be positive and finite");
  if ((!bits_out || !amps_out) && num_samples > 0) throw qx::Error(QX_ERR_ARGUMENT, "null output pointer");
  if (P->lowered.empty() || P->lowered.begin()->second.result_elems != 1)
    throw qx::Error(QX_ERR_UNSUPPORTED, "rejection sampling needs every output leg fixed (one amplitude per bitstring)");
  const int nch = (int)P->prog.num_chars;
  hipStream_t stream = P->ctx->stream;
  int64_t nslices = 1;
  for (int64_t d : P->prog.var_dims) nslices *= d;
  DevBuf obits, oamps, st;
  obits.reserve(std::max<size_t>((size_t)num_samples * nch, 16));
  oamps.reserve(std::max<size_t>((size_t)num_samples * 16, 16));
  st.reserve(32);
  int64_t cnt[2] = {0, 0};
  double state[2] = {M, 0.0};
  hip_check(hipMemcpyAsync(st.p, state, 8, hipMemcpyHostToDevice, stream), "copy");
  hip_check(hipMemcpyAsync((char *)st.p + 16, cnt, 16, hipMemcpyHostToDevice, stream), "copy");
  const double scale = std::ldexp(1.0, nch);
  try {
    while (cnt[0] < num_samples && cnt[1] < max_candidates) {
      const int64_t nb = std::min<int64_t>(P->opt_batch, max_candidates - cnt[1]);
      stage_bits(P, nb, seed, cnt[1]);
      amplitudes(P, nullptr, nb, 0, nslices, nullptr, true);
      qx::launch_reject((const double *)P->outbuf.p, (const char *)P->bits.p, nb, nch, seed, cnt[1], scale, fix_m,
                        num_samples, (double *)st.p, (int64_t *)((char *)st.p + 16), (char *)obits.p,
                        (double *)oamps.p, stream);
      hip_check(hipMemcpyAsync(cnt, (char *)st.p + 16, 16, hipMemcpyDeviceToHost, stream), "copy");
      hip_check(hipMemcpyAsync(state, st.p, 8, hipMemcpyDeviceToHost, stream), "copy");
      hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    if (cnt[0] > 0) {
      hip_check(hipMemcpy(bits_out, obits.p, (size_t)cnt[0] * nch, hipMemcpyDeviceToHost), "copy out");
      hip_check(hipMemcpy(amps_out, oamps.p, (size_t)cnt[0] * 16, hipMemcpyDeviceToHost), "copy out");
    }
  } catch (...) {
    obits.release();
    oamps.release();
    st.release();
    throw;
  }
  obits.release();
  oamps.release();
  st.release();
  if (m_out) *m_out = state[0];
  if (cand_out) *cand_out = cnt[1];
  if (cnt[0] < num_samples)
    throw qx::Error(QX_ERR_STATE, "rejection sampling stopped at max_candidates with " + std::to_string(cnt[0]) +
                                      " of " + std::to_string(num_samples) + " samples");
}

const char *kern_name(const qx::Step &st) {
  if (st.level >= 0) return "small";
  switch (st.kern) {
    case qx::Kern::Radix: return "radix";
    case qx::Kern::Gemm: return "gemm";
    case qx::Kern::Reduce: return "reduce";
    case qx::Kern::Rows: return "rows";
    case qx::Kern::Chain: return "chain";
    case qx::Kern::Outer: return "outer";
    case qx::Kern::Blas: return st.blas.own ? "plain_gemm" : "blas";
    default: return "bits";
  }
}

// Name of the kernel instantiation a step launches, as rocprofv3 reports it
// (template arguments after the precision).
std::string kern_variant(const qx::Step &st) {
  if (st.level >= 0) return "k_small_batch";
  switch (st.kern) {
    case qx::Kern::Rows: {
      int nc = st.rows.nn >= 8 ? 8 : st.rows.nn;
      return "k_contract_rows<" + std::to_string(st.rows.nk) + "," + std::to_string(nc) + ">";
    }
    case qx::Kern::Gemm:
      if (st.gemm.tile == 3) return "k_contract_gemm_mfma";
      if (st.gemm.tile == 4) return "k_contract_gemm_mfma<32>";
      return st.gemm.tile == 0 ? "k_contract_gemm<128,16,16>" : st.gemm.tile == 1 ? "k_contract_gemm<128,32,16>"
                                                                                   : "k_contract_gemm<64,64,16>";
    case qx::Kern::Reduce:
      if (st.bits.skinny)
        return "k_contract_skinny<" + std::to_string(1 << st.bits.nbm) + "," + std::to_string(1 << st.bits.nbn) + ">";
      return "k_contract_reduce";
    case qx::Kern::Radix: return "k_contract_radix";
    case qx::Kern::Chain: return "k_contract_chain";
    case qx::Kern::Outer: return "k_outer_run";
    case qx::Kern::Blas:
      if (st.blas.own)
        return std::string(st.blas.split ? "k_plain_gemm_bf16x3<" : "k_plain_gemm<") + (st.blas.opx ? "T" : "N") +
               (st.blas.opy ? "T" : "N") +
               (st.blas.swap ? ",BA" : ",AB") + (st.blas.gauss ? ",3m" : "") +
               (st.blas.ksplit > 1 ? ",ks" + std::to_string(st.blas.ksplit) : "") + ">" + (st.blas.ttgt ? "+gather" : "");
      return std::string("rocblas_gemm_strided_batched<") + (st.blas.opx ? "T" : "N") + (st.blas.opy ? "T" : "N") +
             (st.blas.swap ? ",BA>" : ",AB>") + (st.blas.ttgt ? "+gather" : "");
    default: return "k_contract_bits";
  }
}

void lowered_json(std::ostringstream &os, const qx::Program &prog, const qx::Lowered &L, int prec,
                  bool with_slots) {
  const size_t es = prec == QX_C128 ? 16 : 8;
  os << "{\"batch\":" << L.batch << ",\"arena_bytes\":" << L.arena_elems * (int64_t)es
     << ",\"result_elems\":" << L.result_elems;
  if (with_slots) {
    os << ",\"slots\":[";
    for (size_t i = 0; i < L.slots.size(); ++i) {
      const auto &s = L.slots[i];
      if (i) os << ",";
      os << "{\"name\":\"" << prog.nodes[s.node].name << "\",\"offset\":" << s.offset << ",\"elems\":" << s.elems
         << ",\"start\":" << s.start << ",\"end\":" << s.end << "}";
    }
    os << "]";
  }
  os << ",\"steps\":[";
  for (size_t s = 0; s < L.steps.size(); ++s) {
    const qx::Step &st = L.steps[s];
    if (s) os << ",";
    double flops = 8.0 * st.H * st.M * st.N * st.K;
    double bytes = (double)es * st.H * (st.M * st.K + st.K * st.N + st.M * st.N);
    if (st.kern == qx::Kern::Chain || st.kern == qx::Kern::Outer) flops = st.chain_flops, bytes = (double)es * st.chain_elems;
    os << "{\"name\":\"" << prog.nodes[st.node].name << "\",\"section\":" << st.section
       << ",\"fused_steps\":" << fused_steps(st) << ",\"B\":" << st.H
       << ",\"M\":" << st.M << ",\"N\":" << st.N << ",\"K\":" << st.K << ",\"flops\":" << flops
       << ",\"bytes\":" << bytes << ",\"kernel\":\"" << kern_name(st) << "\",\"variant\":\""
       << kern_variant(st) << "\",\"config\":\"" << cand_label(st) << "\",\"ws_elems\":" << st.ws_elems;
    if (with_slots) {   // arena values read (views resolved), and the small-step level
      auto root_name = [&](int n) {
        while (prog.nodes[n].op == qx::Op::View || prog.nodes[n].op == qx::Op::Save) n = prog.nodes[n].ia;
        return prog.nodes[n].op == qx::Op::Ncon ? prog.nodes[n].name : std::string();
      };
      os << ",\"level\":" << st.level << ",\"reads\":[\"" << root_name(st.in_a) << "\",\"" << root_name(st.in_b) << "\"";
      for (int c : st.chain_small) os << ",\"" << root_name(c) << "\"";
      for (size_t f = 2; f < st.outer_in.size(); ++f) os << ",\"" << root_name(st.outer_in[f]) << "\"";
      os << "]";
    }
    if (with_slots) {   // leg strides (elements) in A, B and C: free legs in C order, then contracted legs
      os << ",\"legs\":[";
      for (size_t j = 0; j < st.cl.size(); ++j)
        os << (j ? "," : "") << "[" << st.cl[j].dim << "," << st.cl[j].sa << "," << st.cl[j].sb << ","
           << st.cl[j].sc << "]";
      os << "],\"klegs\":[";
      for (size_t j = 0; j < st.kl.size(); ++j)
        os << (j ? "," : "") << "[" << st.kl[j].dim << "," << st.kl[j].sa << "," << st.kl[j].sb << "]";
      os << "]";
      if (st.kern == qx::Kern::Chain) {   // tile layout: store order, LDS swizzle, each step's tile bits
        const qx::ChainArgs &a = st.chain;
        os << ",\"chain_tile\":{\"swizzled\":" << a.swizzled << ",\"wave_local\":" << a.wave_local << ",\"wbits\":["
           << a.wbits[0] << "," << a.wbits[1] << "," << a.wbits[2] << "],\"pout\":[";
        for (int i = 0; i < qx::kChainTileBits; ++i) os << (i ? "," : "") << a.pout[i];
        os << "],\"swz\":[";
        for (int i = 0; i < qx::kChainTileBits; ++i) os << (i ? "," : "") << a.swz[i];
        os << "],\"win\":[";
        for (int i = 0; i < qx::kChainTileBits; ++i) os << (i ? "," : "") << a.win[i];
        os << "],\"wout\":[";
        for (int i = 0; i < qx::kChainTileBits; ++i) os << (i ? "," : "") << a.wout[i];
        os << "],\"q\":[";
        for (int j = 0; j < a.nsteps; ++j) {
          os << (j ? "," : "") << "[";
          for (int i = 0; i < a.st[j].k; ++i) os << (i ? "," : "") << a.st[j].q[i];
          os << "]";
        }
        os << "],\"hq\":[";
        for (int j = 0; j < a.nsteps; ++j) {
          os << (j ? "," : "") << "[";
          for (int i = 0; i < a.st[j].h; ++i) os << (i ? "," : "") << a.st[j].hq[i];
          os << "]";
        }
        os << "]}";
      }
    }
    os << ",\"candidates\":[";
    auto cands = st.level >= 0 ? std::vector<qx::Step>{st} : qx::kernel_candidates(st, prec);
    for (size_t c = 0; c < cands.size(); ++c)
      os << (c ? "," : "") << "\"" << cand_label(cands[c]) << "\"";
    os << "]}";
  }
  os << "]}";
}

const qx::VjpLowered &ensure_vjp(qx_plan *P, int64_t batch, uint64_t fused) {
  auto key = std::make_pair(batch, fused);
  auto it = P->vjp_lowered.find(key);
  if (it != P->vjp_lowered.end()) return it->second;
  qx::VjpLowered V = qx::lower_vjp(P->prog, batch, P->prec, fused, P->opt_batch_last != 0);
  if (P->opt_kernel >= 0) {
    auto force = [&](qx::Step &st) {
      if (qx::force_kernel(st, P->prec, (int)P->opt_kernel)) V.fwd.ws_elems = std::max(V.fwd.ws_elems, st.ws_elems);
    };
    for (auto &st : V.fwd.steps) force(st);
    for (auto &sec : V.bsteps)
      for (auto &b : sec) force(b.step);
  }
  if (P->opt_chain_variant >= 0)
    for (auto &st : V.fwd.steps) force_chain_variant(st, P->opt_chain_variant);
  const size_t es = P->esize();
  size_t need = (size_t)(V.fwd.arena_elems + V.cot_elems + V.temp_elems) * es;
  if (P->arena_cap_bytes > 0 && (int64_t)need > P->arena_cap_bytes)
    throw qx::Error(QX_ERR_MEMORY, "vjp needs " + std::to_string(need >> 20) + " MiB of device memory for batch " +
                                       std::to_string(batch) + ", over the cap of " +
                                       std::to_string(P->arena_cap_bytes >> 20) + " MiB");
  P->vjp_lowered.emplace(key, std::move(V));
  return P->vjp_lowered.at(key);
}

// The pullback of the amplitudes of nb bitstrings for the cotangent `amp_cot`, or, with
// `amp_cot` null, of the loss sum |amp|^2 (cotangent 2 amp formed on the device from
// the same pass's forward; its amplitudes go to `amps_out`; needs every slice variable
// contracted inline, else QX_ERR_UNSUPPORTED).
void vjp(qx_plan *P, const char *bits, int64_t nb, const double *amp_cot, double *grad, double *amps_out = nullptr) {
  if (!P->finalized) throw qx::Error(QX_ERR_STATE, "plan is not finalized");
  if (nb <= 0) throw qx::Error(QX_ERR_ARGUMENT, "need at least one bitstring");
  if (!bits || !grad || (!amp_cot && !amps_out)) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
  const bool norm = amp_cot == nullptr;
  check_bits(P, bits, nb);
  int64_t batch = 1;
  while (batch < nb && batch < P->opt_batch) batch <<= 1;
  const uint64_t fused = choose_fused(P, batch, true);
  const qx::VjpLowered &V = ensure_vjp(P, batch, fused);
  const qx::Lowered &L = V.fwd;
  hipStream_t stream = P->ctx->stream;
  const int64_t es = (int64_t)P->esize();
  const int64_t nch = P->prog.num_chars;
  const int64_t R = L.result_elems;
  bool moved = P->varena.reserve(std::max<size_t>((size_t)L.arena_elems * es, 256));
  moved |= P->ws.reserve(std::max<size_t>((size_t)L.ws_elems * es, 256));
  moved |= P->cot.reserve(std::max<size_t>((size_t)V.cot_elems * es, 256));
  moved |= P->temp.reserve(std::max<size_t>((size_t)V.temp_elems * es, 256));
  std::vector<int64_t> key_off;
  int64_t ke = 0;
  for (auto &k : P->prog.keys) {
    key_off.push_back(ke);
    ke += P->prog.key_elems.at(k);
  }
  moved |= P->grad.reserve(std::max<size_t>((size_t)ke * 16, 16));
  moved |= P->ampcot.reserve((size_t)nb * R * 16);
  if (!norm)
    hip_check(hipMemcpyAsync(P->ampcot.p, amp_cot, (size_t)nb * R * 16, hipMemcpyHostToDevice, stream), "copy");
  size_t nbits = (size_t)nb * nch;
  moved |= P->bits.reserve(std::max<size_t>(nbits, 16));
  if (nbits) hip_check(hipMemcpyAsync(P->bits.p, bits, nbits, hipMemcpyHostToDevice, stream), "copy");
  std::vector<int64_t> ooff;
  std::vector<int> oq;
  for (int o : P->prog.outputs) {
    ooff.push_back(L.values[o].slot);
    oq.push_back(P->prog.nodes[o].qubit);
  }
  const int nout = (int)ooff.size();
  moved |= P->out_off.reserve(std::max<size_t>(ooff.size() * 8, 16));
  moved |= P->out_q.reserve(std::max<size_t>(oq.size() * 4, 16));
  if (nout) {
    P->out_off.upload(ooff.data(), ooff.size() * 8, stream);
    P->out_q.upload(oq.data(), oq.size() * 4, stream);
  }
  const qx::Value &rc = V.cot[V.root];
  const qx::Value &rv = L.values[P->prog.save];
  if (norm && (looped_slices(P, fused) != 1 || !rc.dep || rc.beta_stride != rv.beta_stride))
    throw qx::Error(QX_ERR_UNSUPPORTED, "the norm pullback needs every slice variable contracted inline and a "
                                        "bitstring-dependent root");
  std::vector<int64_t> roff(R);
  {
    for (int64_t e = 0; e < R; ++e) {
      int64_t r = e, o = 0;
      for (size_t j = 0; j < rv.dims.size(); ++j) {
        int64_t d = r % rv.dims[j];
        r /= rv.dims[j];
        o += d * rv.strides[j];
      }
      roff[e] = o;
    }
  }
  moved |= P->relem_off.reserve(std::max<size_t>(R * 8, 16));
  P->relem_off.upload(roff.data(), R * 8, stream);
  if (moved) P->drop_graphs();

  Bases bs;
  bs.arena = (char *)P->varena.p;
  bs.cot = (char *)P->cot.p;
  bs.temp = (char *)P->temp.p;
  const size_t nv = P->prog.vars.size();
  std::vector<int64_t> varval(nv, 0);
  prepare_small(P, L, bs);
  auto run_fwd = [&](int sec) {
    const qx_plan::SmallTable &ST = P->small.at(&L);
    bool walked = false;
    for (size_t s = 0; s < L.steps.size(); ++s) {
      const qx::Step &st = L.steps[s];
      if (st.section != sec) continue;
      if (st.level >= 0) {   // batched small steps: one launch per level, or one walk
        if (ST.count[s] == 0) continue;
        if (!walks(P, ST, sec)) {
          run_small_level(P, L, s);
        } else if (!walked) {
          run_small_walk(P, L, s, sec);
          walked = true;
        }
        continue;
      }
      launch_step(P, L, st, value_ptr(P, L.values[st.in_a], varval, bs),
                  value_ptr(P, L.values[st.in_b], varval, bs), value_ptr(P, L.values[st.node], varval, bs), varval,
                  &bs);
    }
  };
  // operands of a backward step: (A, B, C) of its contraction
  auto bwd_ptrs = [&](const qx::BwdStep &b, void *&A, void *&B, void *&C) {
    const qx::Node &cn = P->prog.nodes[b.step.node];
    const qx::Value &cv = V.cot[b.step.node];
    const bool xa = b.x == cn.ia;
    A = xa ? value_ptr(P, cv, varval, bs) : value_ptr(P, L.values[cn.ia], varval, bs);
    B = xa ? value_ptr(P, L.values[cn.ib], varval, bs) : value_ptr(P, cv, varval, bs);
    C = b.direct ? value_ptr(P, V.cot[b.x], varval, bs) : (void *)(bs.temp + b.temp_off * es);
  };
  // descriptor tables of the batched backward levels (sections 0 and 2; their
  // operands have no slice-dependent offsets)
  qx_plan::BwdTable &BT = P->bwd[&V];
  {
    std::vector<char *> bases = {bs.arena, bs.cot, bs.temp, (char *)P->grad.p};
    if (BT.bases != bases) {
      std::vector<qx::SmallDesc> d;
      std::vector<qx::AddDesc> ad;
      BT.groups.clear();
      for (int sec = 0; sec < 3; ++sec) {
        const auto &bv = V.bsteps[sec];
        BT.first[sec].assign(bv.size(), -1);
        int cur_level = -1;
        for (size_t i = 0; i < bv.size(); ++i) {
          const qx::BwdStep &b = bv[i];
          if (b.level < 0) continue;
          if (b.level != cur_level) {   // the level's first batched step launches the group
            cur_level = b.level;
            BT.groups.push_back({(int)d.size(), 0, (int)ad.size(), 0});
            BT.first[sec][i] = (int)BT.groups.size() - 1;
          }
          auto &g = BT.groups.back();
          qx::SmallDesc x;
          void *A, *B, *C;
          bwd_ptrs(b, A, B, C);
          x.A = A, x.B = B, x.C = C;
          x.conj_a = b.step.bits.conj_a;
          x.conj_b = b.step.bits.conj_b;
          for (auto &l : b.step.cl)
            for (int64_t m = 1; m < l.dim; m <<= 1)
              x.wa[x.nbc] = (int32_t)(l.sa * m), x.wb[x.nbc] = (int32_t)(l.sb * m), ++x.nbc;
          for (auto &l : b.step.kl)
            for (int64_t m = 1; m < l.dim; m <<= 1)
              x.ka[x.nbk] = (int32_t)(l.sa * m), x.kb[x.nbk] = (int32_t)(l.sb * m), ++x.nbk;
          g.ndesc += push_small(d, x);
          if (!b.direct) {
            qx::AddDesc a;
            a.src = C;
            const int64_t xoff = value_offset(L.values[b.x], varval);
            if (b.dest == qx::Buf::Grad) {
              a.dst = (char *)P->grad.p + (key_off[b.dest_key] + xoff) * 16;
            } else {
              a.dst = bs.cot + (V.cot[b.dest_root].slot + xoff) * es;
              a.dst_f64 = P->prec == QX_C128;
            }
            a.n = b.add_elems;
            a.nd = (int32_t)b.add_dims.size();
            for (int j = 0; j < a.nd; ++j) a.dim[j] = b.add_dims[j], a.st[j] = b.add_strides[j];
            ad.push_back(a);
            g.nadd++;
          }
        }
      }
      if (!d.empty()) {
        BT.desc.reserve(d.size() * sizeof(qx::SmallDesc));
        hip_check(hipMemcpy(BT.desc.p, d.data(), d.size() * sizeof(qx::SmallDesc), hipMemcpyHostToDevice), "copy");
      }
      if (!ad.empty()) {
        BT.adds.reserve(ad.size() * sizeof(qx::AddDesc));
        hip_check(hipMemcpy(BT.adds.p, ad.data(), ad.size() * sizeof(qx::AddDesc), hipMemcpyHostToDevice), "copy");
      }
      // zero fills of the accumulating cotangents: one launch per section, ranges of
      // at most 64 KiB per workgroup
      std::vector<int64_t> zr;
      for (int sec = 0; sec < 3; ++sec) {
        BT.zero0[sec] = (int)(zr.size() / 2);
        for (int node : V.zero_at[sec]) {
          const qx::Value &c = V.cot[node];
          const int64_t n16 = (c.elems + 31) / 32 * 32 * es / 16, o16 = c.slot * es / 16;
          for (int64_t o = 0; o < n16; o += 4096) zr.push_back(o16 + o), zr.push_back(std::min<int64_t>(4096, n16 - o));
        }
        BT.nzero[sec] = (int)(zr.size() / 2) - BT.zero0[sec];
      }
      if (!zr.empty()) {
        BT.zeros.reserve(zr.size() * 8);
        hip_check(hipMemcpy(BT.zeros.p, zr.data(), zr.size() * 8, hipMemcpyHostToDevice), "copy");
      }
      BT.bases = bases;
      P->drop_graphs();
    }
  }
  auto zero_cots = [&](int sec) {
    qx::launch_zero_batch(bs.cot, (const int64_t *)BT.zeros.p + 2 * BT.zero0[sec], BT.nzero[sec], stream);
  };
  auto run_bwd = [&](int sec) {
    for (size_t bi = 0; bi < V.bsteps[sec].size(); ++bi) {
      const qx::BwdStep &b = V.bsteps[sec][bi];
      if (b.level >= 0) {   // batched level: contractions, then adds into gradients
        if (BT.first[sec][bi] >= 0) {
          const auto &g = BT.groups[BT.first[sec][bi]];
          qx::launch_small_batch(P->prec, (const qx::SmallDesc *)BT.desc.p + g.desc0, g.ndesc, stream);
          qx::launch_add_batch(P->prec, (const qx::AddDesc *)BT.adds.p + g.add0, g.nadd, stream);
        }
        continue;
      }
      const qx::Node &cn = P->prog.nodes[b.step.node];
      void *A, *B, *C;
      bwd_ptrs(b, A, B, C);
      (void)cn;
      launch_step(P, L, b.step, A, B, C, varval, &bs);
      if (b.direct) continue;
      const int64_t xoff = value_offset(L.values[b.x], varval);
      if (b.dest == qx::Buf::Grad) {
        char *dst = (char *)P->grad.p + (key_off[b.dest_key] + xoff) * 16;
        qx::launch_add_mapped(P->prec, 1, dst, bs.temp, b.add_elems, b.add_dims, b.add_strides, stream);
      } else {
        char *dst = bs.cot + (V.cot[b.dest_root].slot + xoff) * es;
        qx::launch_add_mapped(P->prec, P->prec == QX_C128, dst, bs.temp, b.add_elems, b.add_dims, b.add_strides,
                              stream);
      }
    }
  };
  const int64_t nslices = looped_slices(P, fused);
  // the device work of one call, from zeroing the gradient to the last backward step
  auto enqueue = [&]() {
  qx::launch_zero(P->grad.p, ke * 16, stream);
  run_fwd(0);
  zero_cots(0);
  for (int64_t sl = 0; sl < nslices; ++sl) {
    slice_values(P, fused, sl, varval);
    run_fwd(1);
    zero_cots(1);
    for (int64_t row0 = 0; row0 < nb; row0 += batch) {
      int64_t nvalid = std::min<int64_t>(batch, nb - row0);
      zero_cots(2);
      qx::launch_fill_outputs(P->prec, bs.arena, (const int64_t *)P->out_off.p, nout, (const int *)P->out_q.p,
                              (const char *)P->bits.p + row0 * nch, nvalid, batch, nch, stream);
      run_fwd(2);
      if (norm)
        qx::launch_root_cot_norm(P->prec, value_ptr(P, rc, varval, bs), value_ptr(P, rv, varval, bs), rc.beta_stride,
                                 batch, nvalid, R, (const int64_t *)P->relem_off.p, (double *)P->ampcot.p, row0,
                                 stream);
      else
        qx::launch_root_cot(P->prec, value_ptr(P, rc, varval, bs), rc.dep ? rc.beta_stride : 0, batch, nvalid, R,
                            (const int64_t *)P->relem_off.p, (const double *)P->ampcot.p, row0, stream);
      run_bwd(2);
    }
    run_bwd(1);
  }
  run_bwd(0);
  };
  // forward launches plus at most two per backward step (a contraction and its add)
  int64_t bwd[3] = {0, 0, 0};
  for (int sec = 0; sec < 3; ++sec) bwd[sec] = 2 * (int64_t)V.bsteps[sec].size();
  const int64_t nrows = (nb + batch - 1) / batch;
  const int64_t launches = 1 + section_launches(L, 0) + bwd[0] + 1 +
                           nslices * (section_launches(L, 1) + bwd[1] + 1 +
                                      nrows * (section_launches(L, 2) + bwd[2] + 3));
  P->last_launches = launches;
  if (P->opt_graph && !P->opt_profile && launches <= kMaxGraphLaunches) {
    auto key = std::make_tuple(batch, nb, int64_t(norm ? -2 : -1), (int64_t)fused);
    auto it = P->vjp_graphs.find(key);
    if (it == P->vjp_graphs.end()) {
      hipGraph_t g = nullptr;
      blas_before_capture(P->ctx, L.steps);
      hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue();
      } catch (...) {
        (void)hipStreamEndCapture(stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
      hipGraphExec_t ex = nullptr;
      hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      hip_check(e, "hipGraphInstantiate");
      it = P->vjp_graphs.emplace(key, ex).first;
    }
    hip_check(hipGraphLaunch(it->second, stream), "hipGraphLaunch");
  } else {
    enqueue();
  }
  hip_check(hipMemcpyAsync(grad, P->grad.p, (size_t)ke * 16, hipMemcpyDeviceToHost, stream), "copy out");
  if (norm)
    hip_check(hipMemcpyAsync(amps_out, P->ampcot.p, (size_t)nb * R * 16, hipMemcpyDeviceToHost, stream), "copy out");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

std::string stats_json(const qx_plan *P) {
  std::ostringstream os;
  os.precision(17);
  os << "{\"precision\":" << P->prec << ",\"last_call_launches\":" << P->last_launches
     << ",\"max_graph_launches\":" << kMaxGraphLaunches << ",\"batches\":[";
  bool first = true;
  for (auto &kv : P->lowered) {
    const qx::Lowered &L = kv.second;
    if (!first) os << ",";
    first = false;
    os << "{\"batch\":" << L.batch << ",\"fused\":" << L.fused << ",\"active\":" << (&L == P->last_lowered)
       << ",\"tuned\":" << L.tuned
       << ",\"arena_bytes\":" << L.arena_elems * (int64_t)P->esize() << ",\"result_elems\":" << L.result_elems
       << ",\"steps\":[";
    for (size_t s = 0; s < L.steps.size(); ++s) {
      const qx::Step &st = L.steps[s];
      if (s) os << ",";
      double flops = 8.0 * st.H * st.M * st.N * st.K;
      double bytes = (double)P->esize() * st.H * (st.M * st.K + st.K * st.N + st.M * st.N);
      if (st.kern == qx::Kern::Chain || st.kern == qx::Kern::Outer) flops = st.chain_flops, bytes = (double)P->esize() * st.chain_elems;
      std::string variant = kern_variant(st);
      if (st.level >= 0 && P->opt_small_walk) {   // levels walked by one workgroup
        auto it = P->small.find(&L);
        if (it != P->small.end() && it->second.walk_levels[st.section] > 0) variant = "k_small_walk";
      }
      os << "{\"name\":\"" << P->prog.nodes[st.node].name << "\",\"section\":" << st.section
         << ",\"fused_steps\":" << fused_steps(st)
         << ",\"B\":" << st.H << ",\"M\":" << st.M << ",\"N\":" << st.N << ",\"K\":" << st.K
         << ",\"flops\":" << flops << ",\"bytes\":" << bytes << ",\"kernel\":\""
         << kern_name(st) << "\",\"variant\":\"" << variant << "\",\"config\":\"" << cand_label(st)
         << "\"";
      if (P->opt_profile && &L == P->last_lowered && s < P->step_ms.size())
        os << ",\"ms\":" << P->step_ms[s] << ",\"calls\":" << P->step_calls[s];
      if (s < L.tune_log.size() && !L.tune_log[s].empty()) os << ",\"tuning_ms\":{" << L.tune_log[s] << "}";
      os << "}";
    }
    os << "]}";
  }
  os << "]}";
  return os.str();
}

}  // namespace

extern "C" {

qx_status qx_ctx_create(int device, qx_ctx **out) {
  return guard([&] {
    if (!out) throw qx::Error(QX_ERR_ARGUMENT, "null output pointer");
    *out = nullptr;
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw qx::Error(QX_ERR_ARGUMENT, "no HIP device " + std::to_string(device));
    hip_check(hipSetDevice(device), "hipSetDevice");
    auto *c = new qx_ctx;
    c->device = device;
    // the device address of the fixed scratch area (looked up once per device)
    hipError_t e = (hipError_t)qx::scratch_address(device, &c->scratch);
    if (e != hipSuccess) {
      delete c;
      hip_check(e, "hipGetSymbolAddress (scratch)");
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      hip_check(e, "hipStreamCreate");
    }
    *out = c;
  });
}

qx_status qx_measure_mfma(qx_ctx *ctx, int precision, double *tflops) {
  return guard([&] {
    if (!ctx || !tflops) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    if (precision < QX_C128 || precision > QX_C64_BF16X3) throw qx::Error(QX_ERR_ARGUMENT, "unknown precision");
    hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    const int blocks = 8192, iters = 1024;   // 32 waves per CU on 256 CUs
    const size_t bytes = size_t(blocks) * 256 * (precision == QX_C128 ? 8 : 4);
    void *out = nullptr;
    hip_check(hipMalloc(&out, bytes), "hipMalloc (probe)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    try {
      hip_check(err, "hipEventCreate");
      qx::launch_mfma_probe(precision, out, blocks, iters, ctx->stream);   // untimed: code load
      hip_check(hipEventRecord(e0, ctx->stream), "hipEventRecord");
      qx::launch_mfma_probe(precision, out, blocks, iters, ctx->stream);
      hip_check(hipEventRecord(e1, ctx->stream), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    } catch (...) {
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      (void)hipFree(out);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    const double flops = double(blocks) * 4 /* waves */ * iters * 4 /* products */ * 2.0 * 16 * 16 * 4;
    *tflops = flops / (double(ms) * 1e-3) / 1e12;
  });
}

void qx_ctx_destroy(qx_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
  if (ctx->blas_ws) (void)hipFree(ctx->blas_ws);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

qx_status qx_plan_from_dsl(qx_ctx *ctx, const char *dsl, size_t len, int precision, qx_plan **out) {
  return guard([&] {
    if (!ctx || !out) throw qx::Error(QX_ERR_ARGUMENT, "null context or output pointer");
    *out = nullptr;
    if (precision < QX_C128 || precision > QX_C64_BF16X3) throw qx::Error(QX_ERR_ARGUMENT, "unknown precision");
    auto P = std::make_unique<qx_plan>();
    P->ctx = ctx;
    P->prec = precision;
    P->prog = qx::parse_program(dsl, len);
    *out = P.release();
  });
}

qx_status qx_plan_set_tensor(qx_plan *plan, const char *key, const double *z, int64_t nelem) {
  return guard([&] {
    if (!plan || !key || (!z && nelem > 0)) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    if (plan->finalized) throw qx::Error(QX_ERR_STATE, "plan already finalized");
    auto it = plan->prog.key_elems.find(key);
    if (it == plan->prog.key_elems.end())
      throw qx::Error(QX_ERR_DATA, std::string("data key '") + key + "' is not loaded by the program");
    if (it->second != nelem)
      throw qx::Error(QX_ERR_DATA, std::string("data key '") + key + "' needs " + std::to_string(it->second) +
                                       " elements, got " + std::to_string(nelem));
    plan->host_data[key].assign(z, z + 2 * nelem);
  });
}

qx_status qx_plan_update_tensor(qx_plan *plan, const char *key, const double *z, int64_t nelem) {
  return guard([&] {
    if (!plan || !key || (!z && nelem > 0)) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    if (!plan->finalized) throw qx::Error(QX_ERR_STATE, "plan is not finalized (use qx_plan_set_tensor)");
    auto it = plan->prog.key_elems.find(key);
    if (it == plan->prog.key_elems.end())
      throw qx::Error(QX_ERR_DATA, std::string("data key '") + key + "' is not loaded by the program");
    if (it->second != nelem)
      throw qx::Error(QX_ERR_DATA, std::string("data key '") + key + "' needs " + std::to_string(it->second) +
                                       " elements, got " + std::to_string(nelem));
    const auto &keys = plan->prog.keys;
    const size_t i = (size_t)(std::find(keys.begin(), keys.end(), std::string(key)) - keys.begin());
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    // same buffer, new contents: lowered steps, descriptor tables and captured graphs
    // address the buffer, not its values, so all of them stay valid.  The copy is
    // stream-ordered after earlier calls' work and synchronous towards the caller.
    hipStream_t s = plan->ctx->stream;
    if (plan->prec == QX_C128) {
      hip_check(hipMemcpyAsync(plan->data[i].p, z, (size_t)nelem * 16, hipMemcpyHostToDevice, s), "upload");
    } else {
      std::vector<float> f(z, z + 2 * nelem);
      hip_check(hipMemcpyAsync(plan->data[i].p, f.data(), (size_t)nelem * 8, hipMemcpyHostToDevice, s), "upload");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    plan->host_data[key].assign(z, z + 2 * nelem);
  });
}

qx_status qx_plan_set_option(qx_plan *plan, const char *name, int64_t value) {
  return guard([&] {
    if (!plan || !name) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    std::string n(name);
    bool runtime_opt = (n == "profile" || n == "graph" || n == "fuse_slices" || n == "autotune" || n == "small_walk");
    if (plan->finalized && !runtime_opt) throw qx::Error(QX_ERR_STATE, "option '" + n + "' must be set before finalize");
    if (n == "batch") {
      if (value <= 0 || (value & (value - 1)) || value > (1 << 20))
        throw qx::Error(QX_ERR_ARGUMENT, "batch must be a power of two in [1, 2^20]");
      plan->opt_batch = value;
    } else if (n == "arena_mb") {
      plan->opt_arena_mb = value;
      plan->arena_cap_bytes = value > 0 ? (value << 20) : 0;
    } else if (n == "graph") {
      plan->opt_graph = value;
    } else if (n == "fuse_slices") {
      plan->opt_fuse = value;
    } else if (n == "autotune") {
      plan->opt_autotune = value;
    } else if (n == "chains") {
      plan->opt_chains = value;
    } else if (n == "batch_last") {
      plan->opt_batch_last = value;
    } else if (n == "small_walk") {
      plan->opt_small_walk = value;
      plan->drop_graphs();
    } else if (n == "chain_variant") {
      if (value < -1 || value > 15 || (value & 3) == 3)
        throw qx::Error(QX_ERR_ARGUMENT, "chain_variant must be -1 (heuristic) or 0..15 with bits 0-1 below 3");
      plan->opt_chain_variant = value;
    } else if (n == "kernel") {
      if (value < -1 || value > 12) throw qx::Error(QX_ERR_ARGUMENT, "kernel must be -1 (heuristic) or 0..12");
      plan->opt_kernel = value;
    } else if (n == "profile") {
      // may be toggled at any time; every set clears the accumulated step times
      plan->opt_profile = value;
      plan->step_ms.assign(plan->step_ms.size(), 0.0);
      plan->step_calls.assign(plan->step_calls.size(), 0);
    } else {
      throw qx::Error(QX_ERR_ARGUMENT, "unknown option '" + n + "'");
    }
  });
}

qx_status qx_plan_finalize(qx_plan *plan) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    if (plan->finalized) throw qx::Error(QX_ERR_STATE, "plan already finalized");
    for (auto &k : plan->prog.keys)
      if (!plan->host_data.count(k)) throw qx::Error(QX_ERR_DATA, "data key '" + k + "' was not provided");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    if (plan->arena_cap_bytes == 0) {
      size_t fr = 0, tot = 0;
      hip_check(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
      plan->arena_cap_bytes = (int64_t)(fr * 0.9);
    }
    // lower once up front so label or memory errors surface here
    ensure_lowered(plan, plan->opt_batch);
    plan->data.resize(plan->prog.keys.size());
    const size_t es = plan->esize();
    for (size_t i = 0; i < plan->prog.keys.size(); ++i) {
      const auto &h = plan->host_data[plan->prog.keys[i]];
      size_t n = h.size() / 2;
      plan->data[i].reserve(std::max<size_t>(n * es, 16));
      if (plan->prec == QX_C128) {
        hip_check(hipMemcpy(plan->data[i].p, h.data(), n * 16, hipMemcpyHostToDevice), "upload");
      } else {
        std::vector<float> f(h.begin(), h.end());
        hip_check(hipMemcpy(plan->data[i].p, f.data(), n * 8, hipMemcpyHostToDevice), "upload");
      }
    }
    plan->finalized = true;
  });
}

qx_status qx_plan_info(const qx_plan *plan, int64_t *num_chars, int64_t *result_elems, int64_t *num_slices,
                       int64_t *num_keys, int64_t *key_elems) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    int64_t R = 1;
    for (int64_t d : plan->prog.nodes[plan->prog.save].shape) R *= d;
    int64_t ns = 1;
    for (int64_t d : plan->prog.var_dims) ns *= d;
    int64_t ke = 0;
    for (auto &k : plan->prog.keys) ke += plan->prog.key_elems.at(k);
    if (num_chars) *num_chars = plan->prog.num_chars;
    if (result_elems) *result_elems = R;
    if (num_slices) *num_slices = ns;
    if (num_keys) *num_keys = (int64_t)plan->prog.keys.size();
    if (key_elems) *key_elems = ke;
  });
}

qx_status qx_plan_key(const qx_plan *plan, int64_t i, char *name, size_t cap, int64_t *nelem) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    if (i < 0 || i >= (int64_t)plan->prog.keys.size()) throw qx::Error(QX_ERR_ARGUMENT, "key index out of range");
    const std::string &k = plan->prog.keys[i];
    if (name && cap) {
      size_t n = std::min(cap - 1, k.size());
      std::memcpy(name, k.data(), n);
      name[n] = '\0';
    }
    if (nelem) *nelem = plan->prog.key_elems.at(k);
  });
}

qx_status qx_amplitudes(qx_plan *plan, const char *bits, int64_t nb, int64_t slice_begin, int64_t slice_end,
                        double *out) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    amplitudes(plan, bits, nb, slice_begin, slice_end, out);
  });
}

qx_status qx_amplitudes_noalloc(qx_plan *plan, const char *bits, int64_t nb, int64_t slice_begin, int64_t slice_end,
                                double *out) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    amplitudes_noalloc(plan, bits, nb, slice_begin, slice_end, out);
  });
}

qx_status qx_amplitudes_uniform(qx_plan *plan, int64_t num, uint64_t seed, char *bits_out, double *out) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    amplitudes_uniform(plan, num, seed, bits_out, out);
  });
}

qx_status qx_sample_rejection(qx_plan *plan, int64_t num_samples, uint64_t seed, double M, int fix_M,
                              int64_t max_candidates, char *bits_out, double *amps_out, double *M_out,
                              int64_t *candidates_out) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    sample_rejection(plan, num_samples, seed, M, fix_M, max_candidates, bits_out, amps_out, M_out, candidates_out);
  });
}

qx_status qx_amplitudes_vjp(qx_plan *plan, const char *bits, int64_t nb, const double *amp_cot, double *grad) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    vjp(plan, bits, nb, amp_cot, grad);
  });
}

qx_status qx_norm_vjp(qx_plan *plan, const char *bits, int64_t nb, double *amps, double *grad) {
  return guard([&] {
    if (!plan) throw qx::Error(QX_ERR_ARGUMENT, "null plan");
    if (!amps) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    hip_check(hipSetDevice(plan->ctx->device), "hipSetDevice");
    vjp(plan, bits, nb, nullptr, grad, amps);
  });
}

qx_status qx_plan_stats(const qx_plan *plan, char *json, size_t cap) {
  return guard([&] {
    if (!plan || !json || cap == 0) throw qx::Error(QX_ERR_ARGUMENT, "null argument");
    std::string s = stats_json(plan);
    if (s.size() + 1 > cap)
      throw qx::Error(QX_ERR_ARGUMENT, "stats need " + std::to_string(s.size() + 1) + " bytes");
    std::memcpy(json, s.c_str(), s.size() + 1);
  });
}

void qx_plan_destroy(qx_plan *plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  delete plan;
}

const char *qx_last_error(void) { return g_last_error.c_str(); }

qx_status qx_dsl_inspect(const char *dsl, size_t len, int precision, int64_t batch, uint64_t fused, char *json,
                         size_t cap) {
  return qx_dsl_inspect_ex(dsl, len, precision, batch, fused, 0, json, cap);
}

qx_status qx_dsl_inspect_ex(const char *dsl, size_t len, int precision, int64_t batch, uint64_t fused,
                            int64_t flags, char *json, size_t cap) {
  return guard([&] {
    if (flags & ~int64_t(1)) throw qx::Error(QX_ERR_ARGUMENT, "unknown inspect flags");
    const bool batch_last = flags & 1;
    if (!json || cap == 0) throw qx::Error(QX_ERR_ARGUMENT, "null output buffer");
    if (precision < QX_C128 || precision > QX_C64_BF16X3) throw qx::Error(QX_ERR_ARGUMENT, "unknown precision");
    qx::Program prog = qx::parse_program(dsl, len);
    qx::Lowered L = qx::lower(prog, batch, false, precision, fused, true, batch_last);
    std::ostringstream os;
    os.precision(17);
    lowered_json(os, prog, L, precision, true);
    std::string s = os.str();
    // reverse-mode lowering (its own host-side bounds proof runs here too)
    std::string v;
    try {
      qx::VjpLowered V = qx::lower_vjp(prog, batch, precision, fused, batch_last);
      size_t nb = V.bsteps[0].size() + V.bsteps[1].size() + V.bsteps[2].size();
      std::ostringstream bs;   // backward steps: section, level, kernel, H, M, N, K, destination
      for (int sec = 0; sec < 3; ++sec)
        for (const qx::BwdStep &b : V.bsteps[sec])
          bs << (bs.tellp() > 0 ? "," : "") << "[" << sec << "," << b.level << ",\"" << kern_name(b.step) << "\","
             << b.step.H << "," << b.step.M << "," << b.step.N << "," << b.step.K << "," << (b.direct ? 0 : 1) << "]";
      v = ",\"vjp\":{\"backward_steps\":" + std::to_string(nb) + ",\"forward_elems\":" +
          std::to_string(V.fwd.arena_elems) + ",\"cot_elems\":" + std::to_string(V.cot_elems) +
          ",\"temp_elems\":" + std::to_string(V.temp_elems) + ",\"steps\":[" + bs.str() + "]}";
    } catch (const qx::Error &e) {
      if (e.code != QX_ERR_UNSUPPORTED) throw;
      v = ",\"vjp_error\":\"" + std::string(e.what()) + "\"";
    }
    s.insert(s.size() - 1, v);
    if (s.size() + 1 > cap) throw qx::Error(QX_ERR_ARGUMENT, "inspect needs " + std::to_string(s.size() + 1) + " bytes");
    std::memcpy(json, s.c_str(), s.size() + 1);
  });
}

}  // extern "C"
