// Lowering of a parsed program for one bitstring batch size:
//   * every bitstring-dependent tensor (downstream of an `output` line) gets an
//     extra leg of size `batch`, laid out fastest, so one launch serves a batch of
//     bitstrings and the batch leg makes every such step coalesced;
//   * views are strided views: a base offset plus (slice variable * stride), the
//     viewed leg becoming size 1 (users_guide.md:107-122) -- no copies;
//   * an ncon result is laid out in its DSL label order (leg 1 fastest), size-1
//     legs taking no space;
//   * steps are split into sections run once per call (bitstring- and
//     slice-independent), once per slice, and once per bitstring batch;
//   * arena offsets come from a first-fit over liveness intervals.
#include <algorithm>
#include <climits>
#include <map>
#include <numeric>
#include <set>

#include "qx_internal.hpp"

namespace qx {

static bool pow2(int64_t d) { return d > 0 && (d & (d - 1)) == 0; }
static int ilog2(int64_t d) {
  int r = 0;
  while ((int64_t(1) << r) < d) ++r;
  return r;
}

int64_t first_fit(const std::vector<int64_t> &start, const std::vector<int64_t> &end,
                  const std::vector<int64_t> &size, int esize, std::vector<int64_t> &off) {
  const int64_t align = std::max<int64_t>(1, 256 / esize);
  size_t n = start.size();
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (start[a] != start[b]) return start[a] < start[b];
    return size[a] > size[b];
  });
  off.assign(n, -1);
  std::vector<size_t> placed;
  int64_t total = 0;
  for (size_t i : order) {
    // live intervals overlapping [start, end) in time, sorted by offset
    std::vector<std::pair<int64_t, int64_t>> busy;
    for (size_t j : placed)
      if (start[j] < end[i] && start[i] < end[j]) busy.push_back({off[j], off[j] + size[j]});
    std::sort(busy.begin(), busy.end());
    int64_t cand = 0;
    for (auto &b : busy) {
      if (cand + size[i] <= b.first) break;
      cand = std::max(cand, (b.second + align - 1) / align * align);
    }
    off[i] = cand;
    placed.push_back(i);
    total = std::max(total, cand + size[i]);
  }
  return (total + align - 1) / align * align;
}

// Host-side proof that no kernel can leave its buffers: the largest offset every
// step can form in A, B and C must lie inside that operand's storage.
static void check_bounds(const Program &p, const Lowered &L, const std::vector<int> &root) {
  std::vector<int64_t> key_elems;
  for (auto &k : p.keys) key_elems.push_back(p.key_elems.at(k));
  auto extent = [&](int node) -> std::pair<int64_t, int64_t> {  // (max offset, storage elems)
    const Value &v = L.values[node];
    int64_t mx = v.static_off;
    for (auto &t : v.var_terms) mx += (p.var_dims[t.first] - 1) * t.second;
    int64_t cap = v.buf == Buf::Data ? key_elems[v.data_key] : L.values[root[node]].elems;
    return {mx, cap};
  };
  for (const Step &st : L.steps) {
    if (st.kern == Kern::Chain) {
      const ChainArgs &a = st.chain;
      int64_t mi = 0, mo = 0;
      for (int i = 0; i < kChainTileBits; ++i) mi += a.win[i], mo += a.wout[i];
      for (int j = 0; j < a.nouter; ++j) mi += a.oin[j], mo += a.oout[j];
      auto ex = extent(st.in_a), ec = extent(st.node);
      if (ex.first + mi >= ex.second || ec.first + mo >= ec.second || mo + 1 != st.c_elems ||
          (int64_t(1) << (kChainTileBits + a.nouter)) != st.c_elems)
        throw Error(QX_ERR_STATE, "internal: chain '" + p.nodes[st.node].name + "' addresses outside its operands");
      // LDS addresses are XORs of swz entries: each below the tile size (in bounds) and
      // together a basis over GF(2) (every element has its own address); pout a permutation
      uint32_t basis[kChainTileBits] = {0}, seen = 0;
      for (int b = 0; b < kChainTileBits; ++b) {
        if (a.swz[b] <= 0 || a.swz[b] >= (1 << kChainTileBits) || a.pout[b] < 0 || a.pout[b] >= kChainTileBits)
          throw Error(QX_ERR_STATE, "internal: chain tile layout");
        seen |= 1u << a.pout[b];
        uint32_t x = (uint32_t)a.swz[b];
        for (int k = kChainTileBits - 1; k >= 0 && x; --k)
          if ((x >> k) & 1) {
            if (basis[k]) x ^= basis[k];
            else basis[k] = x, x = 0;
          }
        if (x) throw Error(QX_ERR_STATE, "internal: chain tile layout is not a bijection");
      }
      if (seen != (1u << kChainTileBits) - 1) throw Error(QX_ERR_STATE, "internal: chain store order");
      for (const auto &cm : st.chain_merge) {
        int64_t ms = 0;
        for (int i = 0; i < cm.kc; ++i) {
          ms += cm.sk[i] + cm.sn[i];
          if (cm.pos[i] < 0 || cm.pos[i] >= a.st[cm.step].k || !a.st[cm.step].merged)
            throw Error(QX_ERR_STATE, "internal: chain merge member bits");
        }
        auto es = extent(cm.small);
        if (es.first + ms >= es.second)
          throw Error(QX_ERR_STATE, "internal: chain '" + p.nodes[st.node].name + "' merge member out of range");
      }
      for (int j = 0; j < a.nsteps; ++j) {
        const ChainStepArgs &c = a.st[j];
        if (c.merged) {
          if (c.h != 0 || c.k > 4) throw Error(QX_ERR_STATE, "internal: chain merged step");
          continue;
        }
        int64_t ms = 0;
        for (int i = 0; i < c.k; ++i) ms += c.sk[i] + c.sn[i];
        for (int i = 0; i < c.h; ++i) ms += c.sh[i];
        auto es = extent(st.chain_small[j]);
        if (es.first + ms >= es.second || (int64_t(1) << (c.h + 2 * c.k)) > kChainMaxS)
          throw Error(QX_ERR_STATE, "internal: chain '" + p.nodes[st.node].name + "' operand out of range");
        for (int i = 0; i < c.k; ++i)
          if (c.q[i] < 0 || c.q[i] >= kChainTileBits) throw Error(QX_ERR_STATE, "internal: chain tile bit");
        for (int i = 0; i < c.h; ++i)
          if (c.hq[i] < 0 || c.hq[i] >= kChainTileBits) throw Error(QX_ERR_STATE, "internal: chain tile bit");
      }
      continue;
    }
    if (st.kern == Kern::Outer) {
      const OuterArgs &a = st.outer;
      if (a.nf != (int)st.outer_in.size() || a.nf < 2 || a.nf > kOuterMaxFactors)
        throw Error(QX_ERR_STATE, "internal: outer-product run factors");
      for (int f = 0; f < a.nf; ++f) {
        int64_t mf = 0;
        for (int t = a.first[f]; t < a.first[f + 1]; ++t) {
          if (a.bit[t] < 0 || a.bit[t] >= kOuterBits || (int64_t(1) << a.bit[t]) >= st.c_elems || a.w[t] <= 0)
            throw Error(QX_ERR_STATE, "internal: outer-product run term");
          for (int u = a.first[f]; u < t; ++u)
            if (a.bit[u] == a.bit[t]) throw Error(QX_ERR_STATE, "internal: outer-product run bit named twice");
          mf += a.w[t];
        }
        auto ef = extent(st.outer_in[f]);
        if (ef.first + mf >= ef.second)
          throw Error(QX_ERR_STATE, "internal: outer-product run '" + p.nodes[st.node].name + "' factor out of range");
      }
      auto ec = extent(st.node);
      if (ec.first + st.c_elems - 1 >= ec.second || !pow2(st.c_elems))
        throw Error(QX_ERR_STATE, "internal: outer-product run '" + p.nodes[st.node].name + "' output out of range");
      continue;
    }
    int64_t ma = 0, mb = 0, mc = 0;
    for (auto &g : st.cl) {
      ma += (g.dim - 1) * g.sa;
      mb += (g.dim - 1) * g.sb;
      mc += (g.dim - 1) * g.sc;
    }
    for (auto &g : st.kl) {
      ma += (g.dim - 1) * g.sa;
      mb += (g.dim - 1) * g.sb;
    }
    auto ea = extent(st.in_a), eb = extent(st.in_b), ec = extent(st.node);
    if (ea.first + ma >= ea.second || eb.first + mb >= eb.second || ec.first + mc >= ec.second ||
        mc + 1 != st.c_elems)
      throw Error(QX_ERR_STATE, "internal: step '" + p.nodes[st.node].name + "' addresses outside its operands");
    for (auto &g : st.cl)
      if (g.sa < 0 || g.sb < 0 || g.sc < 0) throw Error(QX_ERR_STATE, "internal: negative stride");
  }
}

// Kernel configurations a power-of-two step can run with.
enum Choice { kBits = 0, kRows = 1, kGemm0 = 2, kGemm1 = 3, kGemm2 = 4, kMfma = 5, kReduce = 6, kMfma32 = 7,
              kBlas = 8, kSkinny = 9, kTtgt = 10, kPlain = 11, kPlain3 = 12, kNumChoices = 13 };

// K split of the hand-written plain GEMM: enough work units to give every CU two
// workgroups, each K range at least 8 stages (128 values) long.
void plain_split(BlasArgs &b, int precision) {
  b.own = 1;
  b.ksplit = 1;
  const int64_t units = plain_gemm_units(precision, b);
  const int64_t stages = b.k / (b.split ? 32 : 16);
  while (units * b.ksplit < 512 && stages / (b.ksplit * 2) >= 8) b.ksplit *= 2;
}

namespace {
struct CB {
  int64_t wa, wb, wc;
};

struct StepBits {
  std::vector<CB> cbits, hb, mb, nbits;
  std::vector<std::pair<int64_t, int64_t>> kbits;
  int nm = 0, nn = 0, nk = 0, nr = 0, hlow = 0;
};

StepBits step_bits(const Step &st) {
  StepBits sb;
  for (auto &g : st.cl)
    for (int i = 0; i < ilog2(g.dim); ++i) sb.cbits.push_back({g.sa << i, g.sb << i, g.sc << i});
  for (auto &g : st.kl)
    for (int i = 0; i < ilog2(g.dim); ++i) sb.kbits.push_back({g.sa << i, g.sb << i});
  for (auto &c : sb.cbits) {
    if (c.wa && c.wb) sb.hb.push_back(c);
    else if (c.wa) sb.mb.push_back(c);
    else sb.nbits.push_back(c);
  }
  sb.nm = (int)sb.mb.size();
  sb.nn = (int)sb.nbits.size();
  sb.nk = (int)sb.kbits.size();
  sb.nr = (int)(sb.hb.size() + sb.mb.size());
  int j = 0;
  for (auto &c : sb.cbits)
    if (c.wa && j < 8) sb.hlow += (c.wb != 0), ++j;
  return sb;
}

bool rows_ok(const StepBits &b) {
  return b.nk <= 5 && b.nn <= 6 && (int64_t(1) << (b.nk + b.nn + b.hlow)) <= 4096;
}
}  // namespace

namespace {
// Bits of one class, weights in two operands, ordered by the first; true when the
// weights in both form dense runs w, 2w, 4w, ... (one matrix dimension each).
bool dense_pair(std::vector<std::pair<int64_t, int64_t>> w, int64_t &lo1, int64_t &lo2) {
  std::sort(w.begin(), w.end());
  lo1 = lo2 = 0;
  if (w.empty()) return true;
  for (size_t i = 1; i < w.size(); ++i)
    if (w[i].first != 2 * w[i - 1].first || w[i].second != 2 * w[i - 1].second) return false;
  lo1 = w[0].first;
  lo2 = w[0].second;
  return true;
}

// Logical matrix Z (rows, cols) stored with dense runs starting at lo_r / lo_c:
// column-major as stored (op 0, ld = lo_c) or transposed (op 1, ld = lo_r).
bool matrix_op(int64_t rows, int64_t cols, int64_t lo_r, int64_t lo_c, int &op, int64_t &ld) {
  if (lo_r == 1 || rows == 1) {
    op = 0;
    ld = cols == 1 ? std::max<int64_t>(rows, 1) : lo_c;
    return ld >= rows;
  }
  if (lo_c == 1 || cols == 1) {
    op = 1;
    ld = rows == 1 ? std::max<int64_t>(cols, 1) : lo_r;
    return ld >= cols;
  }
  return false;
}

// A step that the vendor BLAS can run as one strided-batched GEMM without copies.
bool plain_gemm(const Step &st, BlasArgs &b) {
  if (st.radix.conj_a || st.radix.conj_b) return false;   // conj flags survive reconfiguration here
  std::vector<std::pair<int64_t, int64_t>> am_cm, bn_cn, ak_bk, ha_hb, ha_hc;
  for (auto &g : st.cl) {
    if (!pow2(g.dim)) return false;
    for (int i = 0; i < ilog2(g.dim); ++i) {
      const int64_t wa = g.sa << i, wb = g.sb << i, wc = g.sc << i;
      if (wa && wb) ha_hb.push_back({wa, wb}), ha_hc.push_back({wa, wc});
      else if (wa) am_cm.push_back({wa, wc});
      else bn_cn.push_back({wb, wc});
    }
  }
  for (auto &g : st.kl) {
    if (!pow2(g.dim)) return false;
    for (int i = 0; i < ilog2(g.dim); ++i) ak_bk.push_back({g.sa << i, g.sb << i});
  }
  const int nm = (int)am_cm.size(), nn = (int)bn_cn.size(), nk = (int)ak_bk.size();
  if (nm < 4 || nn < 4 || nk < 4) return false;
  int64_t am, cm, bn, cn, ak, bk, ha, hb, hc, ha2;
  if (!dense_pair(am_cm, am, cm) || !dense_pair(bn_cn, bn, cn) || !dense_pair(ak_bk, ak, bk) ||
      !dense_pair(ha_hb, ha, hb) || !dense_pair(ha_hc, ha2, hc))
    return false;
  const int64_t M = int64_t(1) << nm, N = int64_t(1) << nn, K = int64_t(1) << nk;
  b.batch = int64_t(1) << ha_hb.size();
  int opx, opy;
  int64_t ldx, ldy;
  if (cm == 1) {   // C = M x N column-major: X = A (M x K), Y = B (K x N)
    b.swap = 0;
    b.r = M, b.s = N, b.ldc = cn;
    if (!matrix_op(M, K, am, ak, opx, ldx) || !matrix_op(K, N, bk, bn, opy, ldy)) return false;
    b.sx = ha, b.sy = hb;
  } else if (cn == 1) {   // C^T = N x M: X = B^T (N x K), Y = A^T (K x M)
    b.swap = 1;
    b.r = N, b.s = M, b.ldc = cm;
    if (!matrix_op(N, K, bn, bk, opx, ldx) || !matrix_op(K, M, ak, am, opy, ldy)) return false;
    b.sx = hb, b.sy = ha;
  } else {
    return false;
  }
  if (b.ldc < b.r) return false;
  b.opx = opx, b.opy = opy, b.ldx = ldx, b.ldy = ldy, b.k = K, b.sc = hc;
  const int64_t lim = int64_t(1) << 31;   // rocblas_int sizes and leading dimensions
  return b.r < lim && b.s < lim && b.k < lim && b.ldx < lim && b.ldy < lim && b.ldc < lim && b.batch < lim;
}
}  // namespace

// Configure `st` for kernel configuration `choice`; false if that kernel cannot
// run this step.  Every configuration computes the same C.
bool configure_kernel(Step &st, int precision, int choice) {
  const StepBits sb = step_bits(st);
  const int nm = sb.nm, nn = sb.nn, nk = sb.nk;
  BitArgs &b = st.bits;
  b = BitArgs{};
  b.nbc = (int)sb.cbits.size();
  b.nbk = nk;
  for (int i = 0; i < b.nbc; ++i) b.wa[i] = sb.cbits[i].wa, b.wb[i] = sb.cbits[i].wb;
  for (int i = 0; i < nk; ++i) b.ka[i] = sb.kbits[i].first, b.kb[i] = sb.kbits[i].second;
  st.ws_elems = 0;
  switch (choice) {
    case kBlas: {
      BlasArgs ba;
      if (!plain_gemm(st, ba)) return false;
      if (precision == QX_C64_BF16X3 && ba.r >= 1024 && ba.s >= 1024 && ba.k >= 1024) {
        // operand extents (elements) in their own layouts; workspace: 4 bf16 planes per
        // operand and 2 FP32 planes for C, each 256-byte aligned (complex64 units)
        auto ext = [&](int op, int64_t rows, int64_t cols, int64_t ld, int64_t bs) {
          return (op ? (cols - 1) + (rows - 1) * ld : (rows - 1) + (cols - 1) * ld) + 1 + (ba.batch - 1) * bs;
        };
        ba.split = 1;
        ba.ex = ext(ba.opx, ba.r, ba.k, ba.ldx, ba.sx);
        ba.ey = ext(ba.opy, ba.k, ba.s, ba.ldy, ba.sy);
        ba.ec = ba.r + (ba.s - 1) * ba.ldc + (ba.batch - 1) * ba.sc;
        st.ws_elems = ba.ex + ba.ey + ba.ec + 128;
      }
      st.kern = Kern::Blas;
      st.blas = ba;
      return true;
    }
    case kBits: {
      st.kern = Kern::Bits;
      // small C, long K: split K over workgroups (lanes walk C)
      if (st.c_elems <= (int64_t(1) << 14) && nk >= 8) {
        const int64_t tiles = std::max<int64_t>(1, st.c_elems >> 10);
        const int nk_hi = nk > 9 ? nk - 9 : 0;
        while (tiles * b.ksplit < 1024 && b.ksplit < (int64_t(1) << nk_hi)) b.ksplit *= 2;
      }
      break;
    }
    case kReduce: {
      if (nk < 6) return false;
      st.kern = Kern::Reduce;
      while (st.c_elems * b.ksplit < 8192 && (int64_t(1) << nk) / b.ksplit >= 128) b.ksplit *= 2;
      break;
    }
    case kPlain3:   // complex128 as three real products per complex product
      if (precision != QX_C128) return false;
      [[fallthrough]];
    case kPlain: {
      BlasArgs ba;
      if (!plain_gemm(st, ba)) return false;
      ba.gauss = choice == kPlain3;
      // complex64 steps with M, N, K >= 1024 under QX_C64_BF16X3: the split bf16 product
      ba.split = precision == QX_C64_BF16X3 && ba.r >= 1024 && ba.s >= 1024 && ba.k >= 1024;
      plain_split(ba, precision);
      st.ws_elems = ba.ksplit > 1 ? ba.ksplit * ba.batch * ba.r * ba.s : 0;
      st.kern = Kern::Blas;
      st.blas = ba;
      return true;
    }
    case kTtgt: {
      // a very large step whose operands are not plain matrices: A and B gathered into
      // plain copies, the library GEMM, the result gathered into C (the copies cost a
      // pass over each operand; the autotuner keeps it only where that pays)
      if (nm < 4 || nn < 4 || nk < 4 || 8.0 * st.H * st.M * st.N * st.K < double(int64_t(1) << 34)) return false;
      BlasArgs pa;
      if (plain_gemm(st, pa) || st.radix.conj_a || st.radix.conj_b) return false;
      const int nh = (int)sb.hb.size();
      if (nm + nk + nh > kMaxBits || nk + nn + nh > kMaxBits) return false;
      BlasArgs ba;
      const int64_t M = int64_t(1) << nm, N = int64_t(1) << nn, K = int64_t(1) << nk, H = int64_t(1) << nh;
      if (M >= (int64_t(1) << 31) || N >= (int64_t(1) << 31) || K >= (int64_t(1) << 31) || H >= (int64_t(1) << 31))
        return false;
      for (auto &c : sb.mb) ba.ga[ba.nga++] = c.wa;
      for (auto &k : sb.kbits) ba.ga[ba.nga++] = k.first;
      for (auto &c : sb.hb) ba.ga[ba.nga++] = c.wa;
      for (auto &k : sb.kbits) ba.gb[ba.ngb++] = k.second;
      for (auto &c : sb.nbits) ba.gb[ba.ngb++] = c.wb;
      for (auto &c : sb.hb) ba.gb[ba.ngb++] = c.wb;
      int jm = 0, jn = 0, jh = 0;   // C bits (C order) -> their weight in the plain result
      for (auto &c : sb.cbits) {
        if (c.wa && c.wb) ba.gc[ba.ngc++] = (M * N) << jh++;
        else if (c.wa) ba.gc[ba.ngc++] = int64_t(1) << jm++;
        else ba.gc[ba.ngc++] = M << jn++;
      }
      ba.ttgt = 1;
      ba.r = M, ba.s = N, ba.k = K, ba.ldx = M, ba.ldy = K, ba.ldc = M;
      ba.sx = M * K, ba.sy = K * N, ba.sc = M * N, ba.batch = H;
      ba.ta = M * K * H, ba.tb = K * N * H;
      const int64_t tc = M * N * H;
      st.ws_elems = ba.ta + ba.tb + tc + 512;
      // the copies' product on the hand-written GEMM (its K-split partials after the
      // copies), as the split bf16 product for QX_C64_BF16X3 with M, N, K >= 1024
      ba.split = precision == QX_C64_BF16X3 && M >= 1024 && N >= 1024 && K >= 1024;
      plain_split(ba, precision);
      if (ba.ksplit > 1) st.ws_elems += ba.ksplit * tc + 128;
      st.kern = Kern::Blas;
      st.blas = ba;
      return true;
    }
    case kSkinny: {
      // C = H x M x N with few M N values against a long K: chunks of K so that the
      // H elements times the chunks give about 1024 workgroups, each at least 2^8 long
      if (nk < 12 || nm + nn > 5 || nn > 1) return false;
      st.kern = Kern::Reduce;
      b.skinny = 1;
      b.nbm = nm;
      b.nbn = nn;
      const int nh = (int)sb.hb.size();
      int split_bits = 0;
      while (nh + split_bits < 10 && nk - split_bits > 8) ++split_bits;
      b.kcb = nk - split_bits;
      b.ksplit = 1 << split_bits;
      // lanes walk the low K bits: give them A's smallest strides
      std::vector<std::pair<int64_t, int64_t>> kb(sb.kbits);
      std::sort(kb.begin(), kb.end());
      for (int i = 0; i < nk; ++i) b.ka[i] = kb[i].first, b.kb[i] = kb[i].second;
      break;
    }
    case kRows: {
      if (!rows_ok(sb)) return false;
      st.kern = Kern::Rows;
      RowArgs &r = st.rows;
      r = RowArgs{};
      std::vector<CB> rb;
      for (auto &c : sb.cbits)
        if (c.wa) rb.push_back(c);  // H and M bits, already in C order
      r.nbr = (int)rb.size();
      for (int i = 0; i < std::min(8, r.nbr); ++i)
        if (rb[i].wb) r.hlo |= 1 << i;
      for (int i = 0; i < r.nbr; ++i) r.ra[i] = rb[i].wa, r.rb[i] = rb[i].wb, r.rc[i] = rb[i].wc;
      r.nk = 1 << nk;
      r.nn = 1 << nn;
      for (int k = 0; k < r.nk; ++k) {
        int64_t oa = 0, ob = 0;
        for (int q = 0; q < nk; ++q)
          if ((k >> q) & 1) oa += sb.kbits[q].first, ob += sb.kbits[q].second;
        r.ka[k] = oa;
        r.kb[k] = ob;
      }
      for (int n = 0; n < r.nn; ++n) {
        int64_t ob = 0, oc = 0;
        for (int q = 0; q < nn; ++q)
          if ((n >> q) & 1) ob += sb.nbits[q].wb, oc += sb.nbits[q].wc;
        r.nb[n] = ob;
        r.nc[n] = oc;
      }
      break;
    }
    case kGemm0:
    case kGemm1:
    case kGemm2:
    case kMfma:
    case kMfma32: {
      st.kern = Kern::Gemm;
      GemmArgs &g = st.gemm;
      g = GemmArgs{};
      std::vector<CB> mb = sb.mb, nbits = sb.nbits;
      std::vector<std::pair<int64_t, int64_t>> kbits = sb.kbits;
      auto by_c = [](const CB &x, const CB &y) { return x.wc < y.wc; };
      std::sort(mb.begin(), mb.end(), by_c);
      std::sort(nbits.begin(), nbits.end(), by_c);
      std::sort(kbits.begin(), kbits.end());
      g.nbh = (int)sb.hb.size();
      g.nbm = nm;
      g.nbn = nn;
      g.nbk = nk;
      for (int i = 0; i < g.nbh; ++i) g.ha[i] = sb.hb[i].wa, g.hb[i] = sb.hb[i].wb, g.hc[i] = sb.hb[i].wc;
      for (int i = 0; i < nm; ++i) g.ma[i] = mb[i].wa, g.mc[i] = mb[i].wc;
      for (int i = 0; i < nn; ++i) g.nb[i] = nbits[i].wb, g.nc[i] = nbits[i].wc;
      for (int i = 0; i < nk; ++i) g.ka[i] = kbits[i].first, g.kb[i] = kbits[i].second;
      g.tile = choice == kMfma32 ? 4 : choice - kGemm0;
      const int lm = g.tile >= 2 ? 6 : 7, ln = g.tile == 0 ? 4 : g.tile == 1 ? 5 : 6, lk = g.tile == 4 ? 5 : 4;
      auto min_w = [](const int64_t *w, int n) {
        int64_t m = INT64_MAX;
        for (int i = 0; i < n; ++i) m = std::min(m, w[i]);
        return m;
      };
      int64_t kb_min_b = INT64_MAX, ka_min = min_w(g.ka, std::min(nk, lk));
      for (int i = 0; i < std::min(nk, lk); ++i) kb_min_b = std::min(kb_min_b, g.kb[i]);
      g.a_kfast = ka_min < min_w(g.ma, std::min(nm, lm));
      g.b_kfast = kb_min_b < min_w(g.nb, std::min(nn, ln));
      g.c_nfast = min_w(g.nc, std::min(nn, ln)) < min_w(g.mc, std::min(nm, lm));
      // split K over workgroups when the tiles alone cannot fill the chip
      const int64_t ntiles = (int64_t(1) << g.nbh) * (nm > lm ? int64_t(1) << (nm - lm) : 1) *
                             (nn > ln ? int64_t(1) << (nn - ln) : 1);
      const int64_t chunks = nk > lk ? int64_t(1) << (nk - lk) : 1;
      g.c_elems = st.c_elems;
      while (ntiles * g.ksplit < 1024 && chunks / g.ksplit >= 8) g.ksplit *= 2;
      st.ws_elems = g.ksplit > 1 ? g.ksplit * st.c_elems : 0;
      return true;
    }
    default: return false;
  }
  st.ws_elems = b.ksplit > 1 ? b.ksplit * st.c_elems : 0;
  return true;
}

// The heuristic choice:
//  * Rows   : >= 2^12 rows (H and M), K <= 16, N <= 64, C's fastest bit a row bit
//  * Gemm   : K >= 8, N >= 8, M >= 32 (bits of each class counted without H); complex64
//             with N, M >= 64 on FP32 matrix cores
//  * small C (<= 2^14) with K >= 256: lanes along K (Reduce) when a K bit is A's
//    unit stride, else the gather kernel with K split over workgroups
//  * Bits   : everything else with power-of-two dims
//  * Radix  : any other dims
//  * Plain  : steps of >= 2^30 flops whose operands are plain matrices: the
//             hand-written matrix-core GEMM (k_plain_gemm); with QX_C64_BF16X3, plain
//             steps with M, N, K >= 1024 as its split bf16 product
//  * Blas   : the vendor BLAS, a tuned candidate only
int default_choice(const Step &st, int precision) {
  const StepBits sb = step_bits(st);
  BlasArgs ba;
  if (precision == QX_C64_BF16X3 && plain_gemm(st, ba) && ba.r >= 1024 && ba.s >= 1024 && ba.k >= 1024)
    return kPlain;   // the split bf16 product
  if (8.0 * st.H * st.M * st.N * st.K >= double(int64_t(1) << 30) && plain_gemm(st, ba)) return kPlain;
  if (sb.nr >= 12 && sb.nk <= 4 && sb.nk + sb.nn >= 1 && !sb.cbits.empty() && sb.cbits[0].wa != 0 && rows_ok(sb))
    return kRows;
  if (sb.nk >= 3 && sb.nn >= 3 && sb.nm >= 5) {
    if (precision != QX_C128 && sb.nn >= 6 && sb.nm >= 6) return kMfma;
    return sb.nn <= 4 ? kGemm0 : sb.nn == 5 ? kGemm1 : kGemm2;
  }
  if (st.c_elems <= (int64_t(1) << 14) && sb.nk >= 8) {
    int64_t kmin = INT64_MAX, cmin = INT64_MAX;
    for (auto &kb : sb.kbits) kmin = std::min(kmin, kb.first);
    for (auto &c : sb.cbits)
      if (c.wa) cmin = std::min(cmin, c.wa);
    return kmin < cmin ? kReduce : kBits;
  }
  return kBits;
}

void set_kernel_args(Step &st, int precision) {
  if (st.kern == Kern::Radix) {
    RadixArgs &a = st.radix;
    a.ncl = (int)st.cl.size();
    a.nkl = (int)st.kl.size();
    for (int j = 0; j < a.ncl; ++j) a.cd[j] = st.cl[j].dim, a.ca[j] = st.cl[j].sa, a.cb[j] = st.cl[j].sb;
    for (int j = 0; j < a.nkl; ++j) a.kd[j] = st.kl[j].dim, a.ka[j] = st.kl[j].sa, a.kb[j] = st.kl[j].sb;
    return;
  }
  if (!configure_kernel(st, precision, default_choice(st, precision)))
    throw Error(QX_ERR_STATE, "internal: default kernel cannot run its step");
}

// The same contraction with the operands exchanged (C = B.A): M and N trade places,
// so the row kernels can walk B's free legs (a skinny A with a long B).
static Step swap_operands(const Step &st) {
  Step v = st;
  std::swap(v.in_a, v.in_b);
  for (auto &g : v.cl) std::swap(g.sa, g.sb);
  for (auto &g : v.kl) std::swap(g.sa, g.sb);
  std::swap(v.M, v.N);
  std::swap(v.radix.conj_a, v.radix.conj_b);
  std::swap(v.bits.conj_a, v.bits.conj_b);
  v.swapped = !st.swapped;
  return v;
}

bool force_kernel(Step &st, int precision, int choice) {
  if (st.kern == Kern::Radix || st.kern == Kern::Chain || st.kern == Kern::Outer) return false;
  Step src = st;
  Step v = src;
  if (!configure_kernel(v, precision, choice)) {
    // the row kernel over B's side when A's side cannot hold it
    if (choice != kRows || st.swapped || st.level >= 0) return false;
    src = swap_operands(st);
    const StepBits sw = step_bits(src);
    if (!(sw.nr >= 8 && rows_ok(sw))) return false;
    v = src;
    if (!configure_kernel(v, precision, choice)) return false;
  }
  v.bits.conj_a = v.gemm.conj_a = v.rows.conj_a = src.bits.conj_a;
  v.bits.conj_b = v.gemm.conj_b = v.rows.conj_b = src.bits.conj_b;
  st = std::move(v);
  return true;
}

std::vector<Step> kernel_candidates(const Step &st, int precision) {
  std::vector<Step> out;
  out.push_back(st);
  if (st.kern == Kern::Chain) {
    bool k16 = false;
    for (int j = 0; j < st.chain.nsteps; ++j) k16 |= st.chain.st[j].k == 4;
    bool swz = false;   // a non-trivial swizzle exists
    for (int b = 0; b < kChainTileBits; ++b) swz |= st.chain.swz[b] != (1 << b);
    // m: steps with 16 contracted values on vector FMAs (0) or matrix cores with four
    // (1) or three (2, gate blocks staged in LDS) real products per complex product;
    // wl: wave-local steps or a workgroup barrier after every step
    const bool staged = st.chain.s_elems * (precision == QX_C128 ? 16 : 8) <= kChainSLdsBytes;
    const int wl0 = st.chain.wave_local;
    for (int grid : {256, 512, 1024})
      for (int m : {0, 1, 2})
        for (int z : {0, 1})
          for (int wl = wl0; wl >= 0; --wl) {
            if ((m && !k16) || (m == 2 && !staged) || (z && !swz)) continue;
            if (grid == st.chain.grid && m == st.chain.mfma16 && z == st.chain.swizzled && wl == wl0) continue;
            Step v = st;
            v.chain.grid = grid;
            v.chain.mfma16 = m;
            v.chain.swizzled = z;
            v.chain.wave_local = wl;
            out.push_back(std::move(v));
          }
    return out;
  }
  if (st.kern == Kern::Radix || st.kern == Kern::Outer) return out;
  const StepBits sb = step_bits(st);
  std::vector<int> alts;
  if (sb.nk >= 2 && sb.nn >= 2 && sb.nm >= 4) alts.insert(alts.end(), {kGemm0, kGemm1, kGemm2, kMfma});
  if (sb.nk >= 5 && sb.nn >= 4 && sb.nm >= 4) alts.push_back(kMfma32);
  if (sb.nr >= 8 && rows_ok(sb)) alts.push_back(kRows);
  if (st.c_elems <= (int64_t(1) << 16) && sb.nk >= 6) alts.push_back(kReduce);
  if (sb.nk >= 12 && sb.nm + sb.nn <= 5 && sb.nn <= 1) alts.push_back(kSkinny);
  if (sb.nk >= 4 && sb.nn >= 4 && sb.nm >= 4 && 8.0 * st.H * st.M * st.N * st.K >= double(int64_t(1) << 34))
    alts.push_back(kTtgt);
  if (sb.nk >= 4 && sb.nn >= 4 && sb.nm >= 4) alts.insert(alts.end(), {kPlain, kPlain3, kBlas});
  if (sb.nk <= 6 || st.c_elems <= (int64_t(1) << 16)) alts.push_back(kBits);
  const int def = default_choice(st, precision);
  for (int c : alts) {
    if (c == def) continue;
    Step v = st;
    if (force_kernel(v, precision, c)) out.push_back(std::move(v));
  }
  // a skinny A against a long B: the row and gather kernels over B's side
  if (sb.nn >= 12 && sb.nm <= 6 && !st.swapped && st.level < 0) {
    const Step w = swap_operands(st);
    const StepBits sw = step_bits(w);
    for (int c : {kRows, kBits}) {
      if (c == kRows && !(sw.nr >= 8 && rows_ok(sw))) continue;
      Step v = w;
      if (force_kernel(v, precision, c)) out.push_back(std::move(v));
    }
  }
  return out;
}

Step make_step(const Value &A, const std::vector<int> &alab, bool conj_a, const Value &B,
               const std::vector<int> &blab, bool conj_b, const Value &C, const std::vector<int> &clab,
               int64_t batch, int precision) {
  Step st;
  std::map<int, std::pair<int64_t, int64_t>> la, lb;  // label -> (dim, stride)
  for (size_t j = 0; j < alab.size(); ++j) la[alab[j]] = {A.dims[j], A.strides[j]};
  for (size_t j = 0; j < blab.size(); ++j) lb[blab[j]] = {B.dims[j], B.strides[j]};
  // the batch leg: free in C when C carries it, else contracted between A and B
  if (C.dep) {
    if (!A.dep && !B.dep) throw Error(QX_ERR_STATE, "internal: batch leg in C only");
    st.cl.push_back({batch, A.dep ? A.beta_stride : 0, B.dep ? B.beta_stride : 0, C.beta_stride});
    if (A.dep && B.dep) st.H *= batch;
    else if (A.dep) st.M *= batch;
    else st.N *= batch;
  } else if (A.dep && B.dep) {
    st.kl.push_back({batch, A.beta_stride, B.beta_stride, 0});
    st.K *= batch;
  } else if (A.dep || B.dep) {
    throw Error(QX_ERR_STATE, "internal: batch leg summed out of one operand");
  }
  for (size_t j = 0; j < clab.size(); ++j) {
    int l = clab[j];
    int64_t d = C.dims[j];
    bool ina = la.count(l), inb = lb.count(l);
    if (ina && inb) st.H *= d;
    else if (ina) st.M *= d;
    else st.N *= d;
    if (d == 1) continue;
    st.cl.push_back({d, ina ? la[l].second : 0, inb ? lb[l].second : 0, C.strides[j]});
  }
  // C legs in memory order; C must be dense
  std::stable_sort(st.cl.begin(), st.cl.end(), [](const Leg &x, const Leg &y) { return x.sc < y.sc; });
  int64_t run = 1;
  for (auto &g : st.cl) {
    if (g.sc != run) throw Error(QX_ERR_STATE, "internal: step output is not dense");
    run *= g.dim;
  }
  st.c_elems = run;
  for (size_t j = 0; j < alab.size(); ++j) {
    int l = alab[j];
    if (!lb.count(l) || std::find(clab.begin(), clab.end(), l) != clab.end()) continue;
    int64_t d = A.dims[j];
    st.K *= d;
    if (d == 1) continue;
    st.kl.push_back({d, la[l].second, lb[l].second, 0});
  }
  bool all_pow2 = true;
  int nbc = 0, nbk = 0;
  for (auto &g : st.cl) {
    all_pow2 &= pow2(g.dim);
    nbc += ilog2(g.dim);
  }
  for (auto &g : st.kl) {
    all_pow2 &= pow2(g.dim);
    nbk += ilog2(g.dim);
  }
  st.kern = (all_pow2 && nbc <= kMaxBits && nbk <= kMaxBits) ? Kern::Bits : Kern::Radix;
  if (st.kern == Kern::Radix && ((int)st.cl.size() > kMaxLegs || (int)st.kl.size() > kMaxLegs))
    throw Error(QX_ERR_UNSUPPORTED, "step with more than 48 legs of non power-of-two dims");
  st.radix.conj_a = conj_a;   // read by the kernel choice (plain GEMMs take no conjugate)
  st.radix.conj_b = conj_b;
  set_kernel_args(st, precision);
  st.bits.conj_a = st.gemm.conj_a = st.rows.conj_a = st.radix.conj_a = conj_a;
  st.bits.conj_b = st.gemm.conj_b = st.rows.conj_b = st.radix.conj_b = conj_b;
  return st;
}

// Shapes and slice dependence of every node when the slice variables in `fused`
// (bit j = variable j) are contracted inline instead of looped over: their views
// pass the viewed tensor through unchanged, so the leg keeps its dimension and the
// contractions sum over it (by linearity the same as summing the slices).
static void fused_shapes(const Program &p, uint64_t fused, std::vector<std::vector<int64_t>> &shape,
                         std::vector<char> &sdep) {
  const size_t nn = p.nodes.size();
  shape.assign(nn, {});
  sdep.assign(nn, 0);
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    switch (n.op) {
      case Op::Load:
      case Op::Output: shape[i] = n.shape; break;
      case Op::View:
        shape[i] = shape[n.ia];
        if ((fused >> n.var_id) & 1) {
          sdep[i] = sdep[n.ia];
        } else {
          shape[i][n.pos] = 1;
          sdep[i] = 1;
        }
        break;
      case Op::Ncon: {
        std::map<int, int64_t> dim;
        for (size_t j = 0; j < n.alab.size(); ++j) dim[n.alab[j]] = shape[n.ia][j];
        for (size_t j = 0; j < n.blab.size(); ++j) dim[n.blab[j]] = shape[n.ib][j];
        for (int l : n.clab) shape[i].push_back(dim[l]);
        sdep[i] = sdep[n.ia] || sdep[n.ib];
        break;
      }
      case Op::Save:
        shape[i] = shape[n.ia];
        sdep[i] = sdep[n.ia];
        break;
    }
  }
}

namespace {

// One step seen as X' = S . X with X the large operand: its legs by X stride.
struct ChainView {
  bool x_is_a = true;
  int small = -1, big = -1;                         // nodes
  std::vector<std::pair<int64_t, int64_t>> h, m;    // (X stride, C stride); h also S stride below
  std::vector<int64_t> h_s;                         // S stride of each h leg
  std::vector<std::pair<int64_t, int64_t>> k;       // (X stride, S stride)
  std::vector<std::pair<int64_t, int64_t>> n;       // (S stride, C stride)
};

int64_t value_extent(const Program &p, const Value &v) {   // logical elements of a value
  int64_t e = 1;
  for (int64_t d : v.dims) e *= d;
  (void)p;
  return e;
}

// The step as a chain element, or false: sections 0/1, every leg of dim 2, a small
// operand of at most kChainMaxS elements, as many new legs as contracted ones.
bool chain_view(const Program &p, const Lowered &L, const Step &st, ChainView &cv) {
  if (st.section == 2 || st.bits.conj_a || st.bits.conj_b) return false;
  const Value &A = L.values[st.in_a], &B = L.values[st.in_b];
  if (A.dep || B.dep) return false;
  const int64_t ea = value_extent(p, A), eb = value_extent(p, B);
  cv.x_is_a = ea >= eb;
  cv.big = cv.x_is_a ? st.in_a : st.in_b;
  cv.small = cv.x_is_a ? st.in_b : st.in_a;
  if (std::min(ea, eb) > kChainMaxS || std::max(ea, eb) < (int64_t(1) << 14)) return false;
  // the chain kernel keeps a lane's offsets inside the tile in 32 bits: they are sums
  // of distinct power-of-two strides of the large tensor, below its extent
  if (std::max(ea, eb) >= (int64_t(1) << 32)) return false;
  for (auto &g : st.cl) {
    if (g.dim != 2) return false;
    const int64_t sx = cv.x_is_a ? g.sa : g.sb, ss = cv.x_is_a ? g.sb : g.sa;
    if (sx && ss) {
      cv.h.push_back({sx, g.sc});
      cv.h_s.push_back(ss);
    } else if (sx) {
      cv.m.push_back({sx, g.sc});
    } else {
      cv.n.push_back({ss, g.sc});
    }
  }
  for (auto &g : st.kl) {
    if (g.dim != 2) return false;
    cv.k.push_back({cv.x_is_a ? g.sa : g.sb, cv.x_is_a ? g.sb : g.sa});
  }
  if (cv.k.size() != cv.n.size() || cv.k.size() > 5 || cv.h.size() > 3) return false;
  if ((int64_t(1) << (cv.h.size() + 2 * cv.k.size())) > kChainMaxS) return false;
  return true;
}

struct ChainBuild {
  std::vector<size_t> members;              // step indices
  int x0 = -1, xnode = -1, section = 0;
  std::vector<int64_t> x0_stride;           // per slot (X_0 leg)
  std::map<int64_t, int> slot_of;           // current X stride -> slot
  std::set<int> touched;
  struct S {
    int small;
    std::vector<int> kslot, hslot;
    std::vector<int64_t> sk, sn, sh;
  };
  std::vector<S> steps;
};

// Apply one step to the slot map; false if it does not fit the chain.
bool chain_extend(ChainBuild &cb, const ChainView &cv) {
  const int tile_touch = kChainTileBits - kChainMinPassive;
  std::map<int64_t, int> next;
  ChainBuild::S s;
  s.small = cv.small;
  std::set<int> touched = cb.touched;
  for (auto &hm : cv.h) {
    auto it = cb.slot_of.find(hm.first);
    if (it == cb.slot_of.end()) return false;
    next[hm.second] = it->second;
    s.hslot.push_back(it->second);
    touched.insert(it->second);
  }
  s.sh = cv.h_s;
  for (auto &mm : cv.m) {
    auto it = cb.slot_of.find(mm.first);
    if (it == cb.slot_of.end()) return false;
    next[mm.second] = it->second;
  }
  for (size_t i = 0; i < cv.k.size(); ++i) {
    auto it = cb.slot_of.find(cv.k[i].first);
    if (it == cb.slot_of.end()) return false;
    s.kslot.push_back(it->second);
    s.sk.push_back(cv.k[i].second);
    s.sn.push_back(cv.n[i].first);
    next[cv.n[i].second] = it->second;
    touched.insert(it->second);
  }
  for (auto v : {&s.sk, &s.sn, &s.sh})
    for (int64_t x : *v)
      if (x <= 0 || x > 32767) return false;
  if ((int)touched.size() > tile_touch || next.size() != cb.x0_stride.size()) return false;
  if (cb.steps.size() >= (size_t)kChainMaxSteps) return false;
  cb.slot_of = std::move(next);
  cb.touched = std::move(touched);
  cb.steps.push_back(std::move(s));
  return true;
}

// Turn a chain of at least two steps into one Kern::Chain step.
Step chain_step(const Program &p, const Lowered &L, const ChainBuild &cb) {
  const Step &last = L.steps[cb.members.back()];
  const int nslot = (int)cb.x0_stride.size();
  std::vector<int64_t> xm_stride(nslot, -1);
  for (auto &kv : cb.slot_of) xm_stride[kv.second] = kv.first;
  // tile: touched slots plus the fastest untouched ones, ordered by X_0 stride
  std::vector<int> untouched, tile(cb.touched.begin(), cb.touched.end());
  for (int sl = 0; sl < nslot; ++sl)
    if (!cb.touched.count(sl)) untouched.push_back(sl);
  auto by_in = [&](int a, int b) { return cb.x0_stride[a] < cb.x0_stride[b]; };
  std::sort(untouched.begin(), untouched.end(), by_in);
  size_t u = 0;
  while ((int)tile.size() < kChainTileBits) tile.push_back(untouched[u++]);
  std::sort(tile.begin(), tile.end(), by_in);
  std::vector<int> outer(untouched.begin() + u, untouched.end());
  std::vector<int> pos_of(nslot, -1);
  Step st;
  st.kern = Kern::Chain;
  st.node = last.node;
  st.section = last.section;
  st.in_a = cb.x0;
  st.in_b = cb.steps[0].small;
  st.c_elems = last.c_elems;
  ChainArgs &a = st.chain;
  for (int i = 0; i < kChainTileBits; ++i) {
    pos_of[tile[i]] = i;
    a.win[i] = cb.x0_stride[tile[i]];
    a.wout[i] = xm_stride[tile[i]];
    a.swz[i] = 1 << i;
  }
  std::vector<int> order(kChainTileBits);
  for (int i = 0; i < kChainTileBits; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return a.wout[x] < a.wout[y]; });
  for (int i = 0; i < kChainTileBits; ++i) a.pout[i] = order[i];
  a.nouter = (int)outer.size();
  if (a.nouter > kMaxBits) throw Error(QX_ERR_UNSUPPORTED, "chain: too many legs");
  for (int j = 0; j < a.nouter; ++j) a.oin[j] = cb.x0_stride[outer[j]], a.oout[j] = xm_stride[outer[j]];
  // Merge gate blocks: a step without batch legs joins an earlier group without batch
  // legs when their contracted slots together are at most four (one LDS pass instead
  // of two), moving back past groups whose slots (contracted or batch) it does not
  // touch.
  struct Group {
    std::vector<int> members, slots;   // slots: the union, sorted
    bool batch = false;
    std::set<int> touch;               // contracted and batch slots
  };
  std::vector<Group> groups;
  for (size_t j = 0; j < cb.steps.size(); ++j) {
    const auto &s = cb.steps[j];
    std::set<int> ks(s.kslot.begin(), s.kslot.end());
    bool placed = false;
    if (s.hslot.empty() && s.kslot.size() <= 4) {
      // the reachable groups it could join: best one that it adds no slot to, else one
      // it completes to four slots, else the smallest union
      int best = -1, best_rank = 1 << 30;
      for (int g = (int)groups.size() - 1; g >= 0; --g) {
        const Group &G = groups[g];
        if (!G.batch && (int)G.members.size() < kChainMaxMerged) {
          std::set<int> u(G.slots.begin(), G.slots.end());
          u.insert(ks.begin(), ks.end());
          if (u.size() <= 4) {
            const int rank = u.size() == G.slots.size() ? 0 : u.size() == 4 ? 1 : 1 + (int)u.size();
            if (rank < best_rank) best_rank = rank, best = g;
          }
        }
        bool disjoint = true;
        for (int x : ks) disjoint &= !G.touch.count(x);
        if (!disjoint) break;
      }
      if (best >= 0) {
        Group &G = groups[best];
        std::set<int> u(G.slots.begin(), G.slots.end());
        u.insert(ks.begin(), ks.end());
        G.members.push_back((int)j);
        G.slots.assign(u.begin(), u.end());
        G.touch.insert(ks.begin(), ks.end());
        placed = true;
      }
    }
    if (!placed) {
      Group G;
      G.members = {(int)j};
      G.slots = s.kslot;
      std::sort(G.slots.begin(), G.slots.end());
      G.batch = !s.hslot.empty();
      G.touch.insert(s.kslot.begin(), s.kslot.end());
      G.touch.insert(s.hslot.begin(), s.hslot.end());
      groups.push_back(std::move(G));
    }
  }
  int nmerged = 0;
  for (auto &G : groups) nmerged += G.members.size() > 1 ? (int)G.members.size() : 0;
  if (nmerged > kChainMaxMerged) {   // more members than the merge kernel takes: no merging
    groups.clear();
    for (size_t j = 0; j < cb.steps.size(); ++j) {
      Group G;
      G.members = {(int)j};
      G.slots = cb.steps[j].kslot;
      groups.push_back(std::move(G));
    }
  }
  a.nsteps = (int)groups.size();
  double flops = 0, elems = 0;
  for (size_t j = 0; j < cb.steps.size(); ++j) {
    const Step &m = L.steps[cb.members[j]];
    flops += 8.0 * m.H * m.M * m.N * m.K;
    elems += (double)value_extent(p, L.values[cb.steps[j].small]);
  }
  for (size_t g = 0; g < groups.size(); ++g) {
    const Group &G = groups[g];
    ChainStepArgs &c = a.st[g];
    const auto &s0 = cb.steps[G.members[0]];
    if (G.members.size() == 1) {
      c.k = (int32_t)s0.kslot.size();
      c.h = (int32_t)s0.hslot.size();
      for (int i = 0; i < c.k; ++i) c.q[i] = pos_of[s0.kslot[i]], c.sk[i] = (int16_t)s0.sk[i], c.sn[i] = (int16_t)s0.sn[i];
      for (int i = 0; i < c.h; ++i) c.hq[i] = pos_of[s0.hslot[i]], c.sh[i] = (int16_t)s0.sh[i];
    } else {
      c.k = (int32_t)G.slots.size();
      c.h = 0;
      c.merged = 1;
      for (int i = 0; i < c.k; ++i) c.q[i] = pos_of[G.slots[i]];
      for (int mj : G.members) {
        const auto &s = cb.steps[mj];
        Step::ChainMember cm;
        cm.step = (int)g;
        cm.small = s.small;
        cm.kc = (int)s.kslot.size();
        for (int i = 0; i < cm.kc; ++i) {
          cm.pos[i] = (int)(std::find(G.slots.begin(), G.slots.end(), s.kslot[i]) - G.slots.begin());
          cm.sk[i] = (int)s.sk[i];
          cm.sn[i] = (int)s.sn[i];
        }
        st.chain_merge.push_back(cm);
      }
    }
    c.soff = (int32_t)st.ws_elems;
    st.ws_elems += (int64_t(1) << (c.h + 2 * c.k));
    st.chain_small.push_back(s0.small);
  }
  a.s_elems = (int)st.ws_elems;
  a.tab_off = (int)((st.ws_elems + kChainTabElems - 1) / kChainTabElems * kChainTabElems);
  st.ws_elems = a.tab_off + (int64_t)kChainTabElems * a.nsteps;
  elems += (double)value_extent(p, L.values[cb.x0]) + (double)last.c_elems;
  st.chain_flops = flops;
  st.chain_elems = elems;
  // shape summary for reports: one pass of X with the chain's contracted sizes
  st.H = 1;
  st.M = last.c_elems;
  st.N = 1;
  st.K = 1;
  return st;
}

// Fold runs of chain steps (X_j = S_j . X_{j-1}, each intermediate used once) into
// Kern::Chain steps placed where the run's last step was.  Returns the nodes whose
// values are no longer materialised.
std::vector<int> form_chains(const Program &p, Lowered &L) {
  const size_t nn = p.nodes.size();
  std::vector<int> uses(nn, 0);
  for (const Node &n : p.nodes) {
    if (n.op == Op::Ncon) uses[n.ia]++, uses[n.ib]++;
    if (n.op == Op::View || n.op == Op::Save) uses[n.ia]++;
  }
  std::vector<Step> out;
  std::vector<int> absorbed;
  std::vector<char> drop(L.steps.size(), 0);
  std::vector<std::pair<size_t, Step>> place;   // step index of a run's last member -> chain step
  ChainBuild cb;
  auto close = [&]() {
    if (cb.members.size() >= 2) {
      for (size_t j = 0; j + 1 < cb.members.size(); ++j) {
        drop[cb.members[j]] = 1;
        absorbed.push_back(L.steps[cb.members[j]].node);
      }
      place.push_back({cb.members.back(), chain_step(p, L, cb)});
    }
    cb = ChainBuild();
  };
  auto start = [&](size_t s, const ChainView &cv) {
    cb = ChainBuild();
    const Value &X = L.values[cv.big];
    std::vector<int64_t> strides;
    for (size_t i = 0; i < X.dims.size(); ++i)
      if (X.dims[i] == 2) strides.push_back(X.strides[i]);
      else if (X.dims[i] != 1) return;
    for (size_t i = 0; i < strides.size(); ++i) {
      if (cb.slot_of.count(strides[i])) return;   // aliased legs
      cb.slot_of[strides[i]] = (int)i;
    }
    cb.x0_stride = strides;
    cb.x0 = cv.big;
    cb.section = L.steps[s].section;
    if (!chain_extend(cb, cv)) {
      cb = ChainBuild();
      return;
    }
    cb.members.push_back(s);
    cb.xnode = L.steps[s].node;
  };
  for (size_t s = 0; s < L.steps.size(); ++s) {
    const Step &st = L.steps[s];
    ChainView cv;
    const bool ok = chain_view(p, L, st, cv);
    const bool touches = cb.xnode >= 0 && (st.in_a == cb.xnode || st.in_b == cb.xnode);
    if (touches) {
      ChainBuild keep = cb;
      if (ok && cv.big == cb.xnode && uses[cb.xnode] == 1 && st.section == cb.section && cb.xnode != p.save &&
          chain_extend(cb, cv)) {
        cb.members.push_back(s);
        cb.xnode = st.node;
        continue;
      }
      cb = keep;
      close();
    }
    if (ok && cb.xnode < 0) start(s, cv);
    else if (ok && !touches) {
      // another chain could start here; keep the open one (its steps are interleaved)
    }
  }
  close();
  for (size_t s = 0; s < L.steps.size(); ++s) {
    if (drop[s]) continue;
    bool replaced = false;
    for (auto &pl : place)
      if (pl.first == s) {
        out.push_back(pl.second);
        replaced = true;
      }
    if (!replaced) out.push_back(std::move(L.steps[s]));
  }
  L.steps = std::move(out);
  return absorbed;
}

bool small_step(const Step &st, bool wide);

// A step that only multiplies: no contracted leg, power-of-two legs.
bool outer_product(const Step &st) {
  if (st.kern == Kern::Chain || st.kern == Kern::Outer || st.kern == Kern::Radix || !st.kl.empty() || st.K != 1)
    return false;
  for (const Leg &g : st.cl)
    if (!pow2(g.dim)) return false;
  return true;
}

// The Kern::Outer step of a run of outer products (members in order, each reading the
// previous one's output), or false when the run does not fit OuterArgs.  Walks back
// from the last member: every leg of the final C is followed through the running
// value's operand to the step that brought it in; each factor gets, for every bit of
// every leg it carries, that leg's stride in it.
bool outer_step(const Lowered &L, const std::vector<size_t> &run, Step &out) {
  const int n = (int)run.size();
  const Step &last = L.steps[run.back()];
  struct Pend {
    int64_t dim, c_bit;   // the final leg: dimension and its lowest bit in C
    int64_t s;            // its stride in the current member's output
  };
  std::vector<Pend> pend;
  for (const Leg &g : last.cl) pend.push_back({g.dim, ilog2(g.sc), g.sc});
  // factor terms per member: member i >= 1 adds one factor, member 0 two (A then B)
  std::vector<std::vector<std::pair<int, int64_t>>> terms(n + 1);
  std::vector<int> fnode(n + 1, -1);
  std::vector<int8_t> prev_a(n + 1, 0);
  for (int i = n - 1; i >= 0; --i) {
    const Step &st = L.steps[run[i]];
    if ((int)pend.size() != (int)st.cl.size()) return false;
    std::map<int64_t, const Leg *> by_sc;
    for (const Leg &g : st.cl) by_sc[g.sc] = &g;
    const bool run_a = i > 0 && st.in_a == L.steps[run[i - 1]].node;
    const bool run_b = i > 0 && !run_a && st.in_b == L.steps[run[i - 1]].node;
    if (i > 0 && !run_a && !run_b) return false;
    std::vector<Pend> next;
    auto add_terms = [&](int f, const Pend &pd, int64_t stride) {
      for (int64_t j = 0; (int64_t(1) << j) < pd.dim; ++j) terms[f].push_back({(int)(pd.c_bit + j), stride << j});
    };
    for (const Pend &pd : pend) {
      auto it = by_sc.find(pd.s);
      if (it == by_sc.end() || it->second->dim != pd.dim) return false;
      const Leg &g = *it->second;
      if (i == 0) {
        if (g.sa) add_terms(0, pd, g.sa);
        if (g.sb) add_terms(1, pd, g.sb);
        continue;
      }
      const int64_t s_run = run_a ? g.sa : g.sb, s_fac = run_a ? g.sb : g.sa;
      if (s_run) next.push_back({pd.dim, pd.c_bit, s_run});
      if (s_fac) add_terms(i + 1, pd, s_fac);
    }
    if (i == 0) {
      fnode[0] = st.in_a;
      fnode[1] = st.in_b;
      prev_a[1] = 1;
    } else {
      fnode[i + 1] = run_a ? st.in_b : st.in_a;
      prev_a[i + 1] = run_a ? 1 : 0;
      pend = std::move(next);
    }
  }
  int nterms = 0;
  for (auto &t : terms) nterms += (int)t.size();
  if (n + 1 > kOuterMaxFactors || nterms > kOuterMaxTerms) return false;
  out = Step();
  out.kern = Kern::Outer;
  out.node = last.node;
  out.in_a = fnode[0];
  out.in_b = fnode[1];
  out.outer_in = fnode;
  out.section = last.section;
  out.cl = last.cl;
  out.c_elems = last.c_elems;
  out.M = last.c_elems;
  OuterArgs &a = out.outer;
  a.nf = n + 1;
  a.members = n;
  int k = 0;
  for (int f = 0; f <= n; ++f) {
    a.first[f] = k;
    a.prev_a[f] = prev_a[f];
    for (auto &t : terms[f]) {
      a.bit[k] = (int8_t)t.first;
      a.w[k] = t.second;
      ++k;
    }
  }
  a.first[n + 1] = k;
  double flops = 0, elems = (double)last.c_elems;
  for (size_t m : run) flops += 6.0 * L.steps[m].c_elems;   // one complex product per output element
  for (int f = 0; f <= n; ++f) {
    int64_t fe = 1;   // elements the factor's terms reach
    for (int t = a.first[f]; t < a.first[f + 1]; ++t) fe += a.w[t];
    elems += (double)fe;
  }
  out.chain_flops = flops;
  out.chain_elems = elems;
  return true;
}

// Runs of at least two outer products, each member's output read only by the next
// member (same section, not the saved result), become one Kern::Outer step placed
// where the run's last step was (one launch and one write of the final product
// instead of a launch and an intermediate per member).  Returns the nodes whose
// values are no longer materialised.
std::vector<int> form_outer_runs(const Program &p, Lowered &L) {
  const size_t nn = p.nodes.size();
  std::vector<int> uses(nn, 0);
  for (const Node &n : p.nodes) {
    if (n.op == Op::Ncon) uses[n.ia]++, uses[n.ib]++;
    if (n.op == Op::View || n.op == Op::Save) uses[n.ia]++;
  }
  std::vector<std::vector<size_t>> runs;
  std::map<int, size_t> open;   // output node of a run's last member -> run
  for (size_t s = 0; s < L.steps.size(); ++s) {
    const Step &st = L.steps[s];
    if (!outer_product(st)) continue;
    int from = -1;
    for (int in : {st.in_a, st.in_b}) {
      auto it = open.find(in);
      if (it != open.end() && uses[in] == 1 && in != p.save && L.steps[runs[it->second].back()].section == st.section) {
        from = (int)it->second;
        open.erase(it);
        break;
      }
    }
    if (from < 0) {
      from = (int)runs.size();
      runs.push_back({});
    }
    runs[from].push_back(s);
    open[st.node] = (size_t)from;
  }
  std::vector<int> absorbed;
  std::vector<char> drop(L.steps.size(), 0);
  std::map<size_t, Step> place;
  for (auto &run : runs) {
    // runs that end in a small step stay: the small steps of a section run batched by
    // dependency level, many per launch
    if (run.size() < 2 || small_step(L.steps[run.back()], false)) continue;
    Step f;
    if (!outer_step(L, run, f)) continue;
    for (size_t j = 0; j + 1 < run.size(); ++j) {
      drop[run[j]] = 1;
      absorbed.push_back(L.steps[run[j]].node);
    }
    place[run.back()] = std::move(f);
  }
  std::vector<Step> out;
  for (size_t s = 0; s < L.steps.size(); ++s) {
    if (drop[s]) continue;
    auto it = place.find(s);
    out.push_back(it != place.end() ? std::move(it->second) : std::move(L.steps[s]));
  }
  L.steps = std::move(out);
  return absorbed;
}

}  // namespace

// Wave-local chains: when every step contracts and batches at most four tile bits
// together, the wave index (group bits 6-8 of every step) can sit on the three lowest
// tile bits that no step touches, so that each wave reads and writes only its own
// tile elements from step to step.
void chain_wave_local(ChainArgs &a) {
  uint32_t used = 0;
  bool ok = true;
  for (int j = 0; j < a.nsteps; ++j) {
    const ChainStepArgs &c = a.st[j];
    ok &= c.k + c.h <= 4;
    for (int i = 0; i < c.k; ++i) used |= 1u << c.q[i];
    for (int i = 0; i < c.h; ++i) used |= 1u << c.hq[i];
  }
  int n = 0;
  for (int b = 0; b < kChainTileBits && n < 3; ++b)
    if (!((used >> b) & 1)) a.wbits[n++] = b;
  a.wave_local = ok && n == 3 ? 1 : 0;
}

// A step's free tile bits other than the wave bits of a wave-local chain, in increasing
// order: its group bits 0-3 (the lanes of one access cycle) sit on the first four.
std::vector<int> step_lane_choices(const ChainArgs &a, const ChainStepArgs &c) {
  uint32_t skip = 0;
  for (int i = 0; i < c.k; ++i) skip |= 1u << c.q[i];
  for (int i = 0; i < c.h; ++i) skip |= 1u << c.hq[i];
  if (a.wave_local)
    for (int i = 0; i < 3; ++i) skip |= 1u << a.wbits[i];
  std::vector<int> v;
  for (int b = 0; b < kChainTileBits; ++b)
    if (!((skip >> b) & 1)) v.push_back(b);
  return v;
}

// LDS bank groups: one access cycle serves kLdsCycleBytes, i.e. 2^nb lanes of
// elem_bytes each, and those lanes collide unless their addresses differ in the low nb
// element bits.  Element i sits at P(i) = i ^ sum_{b >= nb} G[b] i_b (G[b] < 2^nb): the
// bank group of i is the linear map L(i) = i_low ^ sum G[b] i_b over GF(2), and a
// cycle's lanes, which differ in nb tile bits, are conflict-free when L maps those bits
// to independent vectors.  A cycle's lanes differ in: the low tile bits (loads), the
// fastest X_m tile bits (stores) and each step's group bits 0-3, its lowest free tile
// bits other than the wave bits of a wave-local chain (in both the vector and the
// matrix-core step, whose result columns are groups).  Every tile bit from nb up may
// carry a pattern, the bits of a lane's element index r included (tile bits 9-12 for
// loads, the last four X_m tile bits for stores: their images then combine by XOR, one
// vector op per element of a refill or a store; kept plain, they would leave eight of
// the thirteen bits without a pattern and many steps' lanes on a few bank groups).
// G minimises the summed conflict degree (weights: 2 per step, 1 per load / store
// pass) by hill climbing from fixed seeds.
void chain_swizzle(ChainArgs &a, int elem_bytes) {
  int nb = 0;
  while ((elem_bytes << (nb + 1)) <= kLdsCycleBytes) ++nb;
  std::vector<std::pair<std::vector<int>, int>> sets;
  {
    std::vector<int> v, w;
    for (int i = 0; i < nb; ++i) v.push_back(i), w.push_back(a.pout[i]);
    sets.push_back({v, 1});
    sets.push_back({w, 1});
  }
  for (int j = 0; j < a.nsteps; ++j) {
    std::vector<int> v = step_lane_choices(a, a.st[j]);
    if ((int)v.size() > nb) v.resize(nb);
    sets.push_back({v, 2});
  }
  std::vector<char> fixed(kChainTileBits, 0);   // G must stay 0
  for (int b = 0; b < nb; ++b) fixed[b] = 1;
  std::vector<uint32_t> G(kChainTileBits, 0);
  auto image = [&](int b) { return b < nb ? 1u << b : G[b]; };
  auto cost = [&]() {
    int total = 0;
    for (auto &st : sets) {
      uint32_t basis[8] = {0};   // GF(2) elimination, basis[k] has leading bit k
      int rank = 0;
      for (int b : st.first) {
        uint32_t x = image(b);
        for (int k = nb - 1; k >= 0 && x; --k)
          if ((x >> k) & 1) {
            if (basis[k]) x ^= basis[k];
            else basis[k] = x, x = 0, ++rank;
          }
      }
      total += st.second << ((int)st.first.size() - rank);
    }
    return total;
  };
  std::vector<uint32_t> best = G;
  int best_cost = cost();
  uint64_t rng = 0x9E3779B97F4A7C15ull;
  for (int restart = 0; restart < 32; ++restart) {
    for (int b = 0; b < kChainTileBits; ++b) {
      rng ^= rng << 13, rng ^= rng >> 7, rng ^= rng << 17;
      G[b] = (restart == 0 || fixed[b]) ? 0u : (uint32_t)(rng % (1u << nb));
    }
    int cur = cost();
    for (bool improved = true; improved;) {
      improved = false;
      for (int b = 0; b < kChainTileBits; ++b) {
        if (fixed[b]) continue;
        uint32_t pick = G[b];
        int pick_cost = cur;
        for (uint32_t v = 0; v < (1u << nb); ++v) {
          G[b] = v;
          const int c = cost();
          if (c < pick_cost) pick_cost = c, pick = v;
        }
        G[b] = pick;
        if (pick_cost < cur) cur = pick_cost, improved = true;
      }
    }
    if (cur < best_cost) best_cost = cur, best = G;
  }
  for (int b = 0; b < kChainTileBits; ++b) a.swz[b] = (int32_t)((1u << b) | (fixed[b] ? 0u : best[b]));
  G.assign(kChainTileBits, 0);
  a.swizzled = best_cost < cost() ? 1 : 0;   // by default only where it removes conflicts
}


namespace {

// Whether a step may run inside a small-step batch (one workgroup per step):
// slice-free, power-of-two legs with at most kSmallMaxBits C and K bits each, at most
// 2^18 MACs, and operand offsets that fit 32 bits.  C and K are each at most 4096,
// except with `wide` (the pullback, whose gradient steps sum long bitstring-batch
// ranges into a few elements): then up to 2^16 MACs with any split.  (The forward
// keeps the limit: its larger steps run faster as multi-workgroup launches.)
bool small_step(const Step &st, bool wide = false) {
  if (st.section == 1 || st.kern == Kern::Chain || st.kern == Kern::Radix || st.kern == Kern::Outer) return false;
  const int64_t macs = st.c_elems * st.K;
  if (macs > (int64_t(1) << 18)) return false;
  if ((st.c_elems > 4096 || st.K > 4096) && (!wide || macs > (int64_t(1) << 16))) return false;
  int nbc = 0, nbk = 0;
  int64_t wmax = 0;
  for (auto &g : st.cl) {
    if (!pow2(g.dim)) return false;
    nbc += ilog2(g.dim);
    wmax = std::max({wmax, g.sa * (g.dim - 1), g.sb * (g.dim - 1)});
  }
  for (auto &g : st.kl) {
    if (!pow2(g.dim)) return false;
    nbk += ilog2(g.dim);
    wmax = std::max({wmax, g.sa * (g.dim - 1), g.sb * (g.dim - 1)});
  }
  return nbc <= kSmallMaxBits && nbk <= kSmallMaxBits && wmax < (int64_t(1) << 30);
}

// Hoist the small steps of section 0 (once per call) and of section 2 (once per
// bitstring batch) whose inputs are available at the section's start (data, values
// of earlier sections, bitstring tensors, or other such steps) to the front of
// their section, grouped by dependency level (a level's steps are independent of
// each other).  Steps reading slice-dependent views are left alone (their
// addresses change per slice).  The order stays topological; liveness is computed
// on the new order afterwards.
void level_small_steps(const Program &p, Lowered &L) {
  const size_t nn = p.nodes.size();
  std::vector<int> root(nn, -1), prod(nn, -1);
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    root[i] = (n.op == Op::View || n.op == Op::Save) ? root[n.ia] : (int)i;
  }
  for (size_t s = 0; s < L.steps.size(); ++s) prod[L.steps[s].node] = (int)s;
  std::vector<int> level(L.steps.size(), -1);
  int maxlev = -1;
  for (size_t s = 0; s < L.steps.size(); ++s) {
    const Step &st = L.steps[s];
    if (!small_step(st)) continue;
    if (!L.values[st.in_a].var_terms.empty() || !L.values[st.in_b].var_terms.empty()) continue;
    int lv = 0;
    bool ok = true;
    for (int in : {st.in_a, st.in_b}) {
      const int r = root[in];
      if (r < 0) continue;
      const int ps = prod[r];
      if (ps < 0) continue;                               // data or bitstring output
      if (L.steps[ps].section != st.section) continue;   // an earlier section's value
      if (level[ps] < 0) {
        ok = false;
        break;
      }
      lv = std::max(lv, level[ps] + 1);
    }
    if (!ok) continue;
    level[s] = lv;
    maxlev = std::max(maxlev, lv);
  }
  if (maxlev < 0) return;
  // levels are numbered per section; section 2's are offset past section 0's
  const int off2 = maxlev + 1;
  std::vector<int> count(2 * (maxlev + 1), 0);
  for (size_t s = 0; s < L.steps.size(); ++s)
    if (level[s] >= 0) count[level[s] + (L.steps[s].section == 2 ? off2 : 0)]++;
  std::vector<Step> out;
  for (int sec : {0, 1, 2}) {
    for (int lv = 0; lv <= maxlev; ++lv)
      for (size_t s = 0; s < L.steps.size(); ++s)
        if (L.steps[s].section == sec && level[s] == lv) {
          const int key = lv + (sec == 2 ? off2 : 0);
          out.push_back(L.steps[s]);
          out.back().level = count[key] > 1 ? key : -1;   // a lone step keeps its own kernel
        }
    for (size_t s = 0; s < L.steps.size(); ++s)
      if (L.steps[s].section == sec && level[s] < 0) out.push_back(L.steps[s]);
  }
  L.steps = std::move(out);
}

}  // namespace

// Leg orders (positions in each node's own dims, fastest first) that turn large
// contractions into plain (strided-batched) column-major matrix products, which the
// vendor BLAS runs without copies.  For C = A.B with legs M (A only), N (B only), K
// (contracted) and H (all three), every operand must hold two runs of its matrix legs
// (A: M and K, B: K and N, C: M and N, each run in one order shared by the operands
// that hold it) followed by the H run (one order everywhere).  An operand whose layout
// is fixed (a data tensor, a value with other consumers, a saved value, or one
// regrouped for a larger contraction) must already have that form and sets the run
// orders; the others (intermediates this contraction alone reads, and its own result)
// are regrouped to match, keeping their own relative orders where nothing is set and
// putting first the run that holds an operand's fastest leg (a result: its M run).  The batched bitstring leg
// stays fastest (slowest with batch_last): its class's run is placed beside it, and
// as an H leg it must be slowest.  Largest contractions (by flops) go first.  Views
// and contractions below 2^30 flops or with M, N or K under 64 are left alone.
// Empty order = DSL order.
static std::vector<std::vector<int>> plain_layouts(const Program &p, const std::vector<std::vector<int64_t>> &shape,
                                                   int64_t batch, bool batch_last) {
  const size_t nn = p.nodes.size();
  std::vector<std::vector<int>> order(nn);
  std::vector<int> uses(nn, 0);
  std::vector<char> pinned(nn, 0);   // saved or viewed: layout fixed
  for (const Node &n : p.nodes) {
    if (n.op == Op::Ncon) uses[n.ia]++, uses[n.ib]++;
    if (n.op == Op::View || n.op == Op::Save) uses[n.ia]++, pinned[n.ia] = 1;
  }
  auto has = [](const std::vector<int> &v, int l) { return std::find(v.begin(), v.end(), l) != v.end(); };
  auto pos = [](const std::vector<int> &v, int l) { return (int)(std::find(v.begin(), v.end(), l) - v.begin()); };
  // class of a label of contraction n: M (A only), N (B only), K (A and B), H (all)
  auto cls = [&](const Node &n, int l) {
    const bool a = has(n.alab, l), b = has(n.blab, l), c = has(n.clab, l);
    return a && b ? (c ? 'H' : 'K') : a ? 'M' : 'N';
  };
  struct Cand {
    size_t node;
    double flops;
    char beta;
  };
  std::vector<Cand> cands;
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    if (n.op != Op::Ncon || n.ia == n.ib) continue;
    if (p.nodes[n.ia].op == Op::View || p.nodes[n.ib].op == Op::View) continue;
    const bool da = p.nodes[n.ia].dep, db = p.nodes[n.ib].dep, dc = n.dep;
    char beta = 0;   // class of the batched bitstring leg
    if (!da && db && dc) beta = 'N';
    else if (da && !db && dc) beta = 'M';
    else if (da && db && dc) beta = 'H';
    else if (da || db || dc) continue;
    if (beta == 'H' && !batch_last) continue;   // a fastest batch leg breaks the matrices
    double prod[4] = {1, 1, 1, 1};               // M N K H
    auto slot = [](char c) { return c == 'M' ? 0 : c == 'N' ? 1 : c == 'K' ? 2 : 3; };
    if (beta) prod[slot(beta)] *= (double)batch;
    bool ok = true, has_h = beta == 'H';
    for (size_t j = 0; j < n.alab.size(); ++j) {
      const int64_t d = shape[n.ia][j];
      ok &= pow2(d);
      const char c = cls(n, n.alab[j]);
      prod[slot(c)] *= (double)d;
      has_h |= c == 'H' && d > 1;
    }
    for (size_t j = 0; j < n.blab.size(); ++j) {
      ok &= pow2(shape[n.ib][j]);
      if (cls(n, n.blab[j]) == 'N') prod[1] *= (double)shape[n.ib][j];
    }
    if (!ok || prod[0] < 64 || prod[1] < 64 || prod[2] < 64) continue;
    if (has_h && (beta == 'M' || beta == 'N')) continue;   // the batch leg would split a run from H
    const double fl = 8.0 * prod[0] * prod[1] * prod[2] * prod[3];
    if (fl >= 1e9) cands.push_back({i, fl, beta});
  }
  std::stable_sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.flops > y.flops; });

  for (const Cand &cd : cands) {
    const size_t i = cd.node;
    const Node &n = p.nodes[i];
    const char beta = cd.beta;
    struct Op3 {
      int node;
      const std::vector<int> *lab;
      bool fixed;
      char r1, r2;            // the two matrix runs of this operand (A: M K, B: N K, C: M N)
      std::vector<int> store;  // labels of legs with dim > 1, fastest first
    };
    Op3 ops[3] = {{n.ia, &n.alab, false, 'M', 'K', {}}, {n.ib, &n.blab, false, 'N', 'K', {}},
                  {(int)i, &n.clab, false, 'M', 'N', {}}};
    bool ok = true;
    for (int o = 0; o < 3; ++o) {
      Op3 &x = ops[o];
      const Node &xn = p.nodes[x.node];
      x.fixed = o == 2 ? (pinned[i] || !order[i].empty())
                       : !(xn.op == Op::Ncon && uses[x.node] == 1 && order[x.node].empty() && !pinned[x.node]);
      const std::vector<int64_t> &dims = shape[x.node];
      std::vector<int> st = order[x.node];
      if (st.empty())
        for (size_t j = 0; j < dims.size(); ++j) st.push_back((int)j);
      for (int j : st)
        if (dims[j] > 1) x.store.push_back((*x.lab)[j]);
      if (!x.fixed) continue;
      // fixed: matrix runs (at most one change of class) then the H run
      size_t e = x.store.size();
      while (e > 0 && cls(n, x.store[e - 1]) == 'H') --e;
      int changes = 0;
      for (size_t k = 0; k < e; ++k) {
        if (cls(n, x.store[k]) == 'H') ok = false;
        if (k && cls(n, x.store[k]) != cls(n, x.store[k - 1])) ++changes;
      }
      if (changes > 1) ok = false;
      // the batch leg sits beside its class's run
      const bool beta_here = (beta == x.r1 || beta == x.r2) && beta != 'H';
      if (beta_here && e > 0 && changes == 1) {
        const char edge = batch_last ? cls(n, x.store[e - 1]) : cls(n, x.store[0]);
        if (edge != beta) ok = false;
      }
    }
    if (!ok) continue;
    // run orders: from fixed operands (which must agree), else the first holder's order
    std::map<char, std::vector<int>> run;
    for (char c : {'K', 'M', 'N', 'H'}) {
      bool set = false;
      for (int o = 0; o < 3 && ok; ++o) {
        Op3 &x = ops[o];
        if (!x.fixed || (c != x.r1 && c != x.r2 && c != 'H')) continue;
        std::vector<int> v;
        for (int l : x.store)
          if (cls(n, l) == c) v.push_back(l);
        if (set && v != run[c]) ok = false;
        run[c] = v, set = true;
      }
      if (set) continue;
      for (int o = 0; o < 3 && !set; ++o) {
        Op3 &x = ops[o];
        if (c != x.r1 && c != x.r2 && c != 'H') continue;
        std::vector<int> v;
        for (int l : x.store)
          if (cls(n, l) == c) v.push_back(l);
        run[c] = v, set = true;
      }
    }
    if (!ok) continue;
    for (int o = 0; o < 3; ++o) {
      Op3 &x = ops[o];
      if (x.fixed) continue;
      // the run next to the batch leg goes on its side, else the run of an operand's
      // fastest leg; results put M first (measured better for their consumers)
      bool r1_first = o == 2 || x.store.empty() || cls(n, x.store[0]) == x.r1;
      if (beta == x.r1) r1_first = !batch_last;
      else if (beta == x.r2) r1_first = batch_last;
      std::vector<int> labs = r1_first ? run[x.r1] : run[x.r2];
      const std::vector<int> &second = r1_first ? run[x.r2] : run[x.r1];
      labs.insert(labs.end(), second.begin(), second.end());
      labs.insert(labs.end(), run['H'].begin(), run['H'].end());
      std::vector<int> ord;
      for (int l : labs) ord.push_back(pos(*x.lab, l));
      if (ord.size() != x.store.size()) {
        ok = false;
        break;
      }
      order[x.node] = ord;
    }
    if (!ok) throw Error(QX_ERR_STATE, "internal: layout regrouping lost a leg");
  }
  return order;
}

Lowered lower(const Program &p, int64_t batch, bool keep_all, int precision, uint64_t fused, bool chains,
              bool batch_last) {
  if (!pow2(batch)) throw Error(QX_ERR_ARGUMENT, "batch must be a power of two");
  Lowered L;
  L.batch = batch;
  L.fused = fused;
  const size_t nn = p.nodes.size();
  L.values.resize(nn);
  std::vector<std::vector<int64_t>> shape;
  std::vector<char> sdep;
  fused_shapes(p, fused, shape, sdep);
  const std::vector<std::vector<int>> leg_order = plain_layouts(p, shape, batch, batch_last);
  std::map<std::string, int> key_index;
  for (size_t i = 0; i < p.keys.size(); ++i) key_index[p.keys[i]] = (int)i;

  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    Value &v = L.values[i];
    switch (n.op) {
      case Op::Load: {
        v.buf = Buf::Data;
        v.data_key = key_index.at(n.key);
        v.dims = shape[i];
        int64_t s = 1;
        for (int64_t d : shape[i]) {
          v.strides.push_back(s);
          s *= d;
        }
        v.elems = s;
        break;
      }
      case Op::Output:
        v.buf = Buf::Arena;
        v.dims = n.shape;
        v.strides = {batch};
        v.dep = true;
        v.beta_stride = 1;
        v.elems = 2 * batch;
        L.out_elems = v.elems;
        break;
      case Op::View: {
        v = L.values[n.ia];
        v.slot = -1;
        if (!((fused >> n.var_id) & 1)) {
          v.var_terms.push_back({n.var_id, v.strides[n.pos]});
          v.dims[n.pos] = 1;
        }
        break;
      }
      case Op::Ncon: {
        v.buf = Buf::Arena;
        v.dims = shape[i];
        v.dep = n.dep;
        int64_t s = 1;
        if (n.dep && !batch_last) {
          v.beta_stride = 1;
          s = batch;
        }
        v.strides.assign(shape[i].size(), 0);   // legs of dim 1 (possibly left out of an order): 0
        const std::vector<int> &ord = leg_order[i];
        for (size_t k = 0; k < (ord.empty() ? shape[i].size() : ord.size()); ++k) {
          const size_t j = ord.empty() ? k : (size_t)ord[k];
          if (shape[i][j] > 1) v.strides[j] = s, s *= shape[i][j];
        }
        if (n.dep && batch_last) {
          v.beta_stride = s;
          s *= batch;
        }
        v.elems = s;
        break;
      }
      case Op::Save:
        v = L.values[n.ia];
        break;
    }
  }

  // steps
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    if (n.op != Op::Ncon) continue;
    Step st = make_step(L.values[n.ia], n.alab, false, L.values[n.ib], n.blab, false, L.values[i], n.clab, batch,
                        precision);
    st.node = (int)i;
    st.in_a = n.ia;
    st.in_b = n.ib;
    st.section = L.values[i].dep ? 2 : (sdep[i] ? 1 : 0);
    L.ws_elems = std::max(L.ws_elems, st.ws_elems);
    L.steps.push_back(std::move(st));
  }

  std::vector<char> absorbed(nn, 0);
  if (chains && !keep_all) {
    for (int a : form_chains(p, L)) absorbed[a] = 1;
    for (int a : form_outer_runs(p, L)) absorbed[a] = 1;
    for (Step &st : L.steps) {
      if (st.kern == Kern::Chain) {
        chain_wave_local(st.chain);
        chain_swizzle(st.chain, precision == QX_C128 ? 16 : 8);
        // steps with 16 contracted values on matrix cores, three real products per
        // complex product where the gate blocks are staged in LDS (the fastest form in
        // measurement; the autotuner also times the others)
        bool k16 = false;
        for (int j = 0; j < st.chain.nsteps; ++j) k16 |= st.chain.st[j].k == 4;
        const bool staged = st.chain.s_elems * (precision == QX_C128 ? 16 : 8) <= kChainSLdsBytes;
        if (k16) st.chain.mfma16 = staged ? 2 : 1;
      }
      L.ws_elems = std::max(L.ws_elems, st.ws_elems);
    }
  }
  if (chains || keep_all) level_small_steps(p, L);

  // result elements per bitstring
  L.result_elems = 1;
  for (int64_t d : shape[p.save]) L.result_elems *= d;

  // liveness: schedule positions = sections 0, 1, then output fill, then section 2
  std::vector<int> order;
  for (int sec = 0; sec < 3; ++sec)
    for (size_t s = 0; s < L.steps.size(); ++s)
      if (L.steps[s].section == sec) order.push_back((int)s);
  std::vector<int64_t> pos_of_step(L.steps.size());
  int64_t fill_pos = 0;
  {
    int64_t pos = 0;
    bool filled = false;
    int prev_level = -1;
    for (int s : order) {
      if (!filled && L.steps[s].section == 2) {
        fill_pos = pos++;
        filled = true;
      }
      // the steps of one batched level run concurrently: one schedule position, so
      // none of them reuses memory another one reads or writes
      const int lv = L.steps[s].level;
      if (lv >= 0 && lv == prev_level) {
        pos_of_step[s] = pos - 1;
        continue;
      }
      prev_level = lv;
      pos_of_step[s] = pos++;
    }
    if (!filled) fill_pos = pos++;
  }
  const int64_t end_pos = (int64_t)order.size() + 2;
  // arena root of each node's value (views resolve to what they view)
  std::vector<int> root(nn, -1);
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    if (n.op == Op::View || n.op == Op::Save) root[i] = root[n.ia];
    else if (n.op == Op::Ncon || n.op == Op::Output) root[i] = (int)i;
  }
  std::vector<int64_t> st_pos(nn, -1), en_pos(nn, -1);
  std::vector<int> sec_of(nn, 2);
  for (size_t s = 0; s < L.steps.size(); ++s) {
    int ni = L.steps[s].node;
    st_pos[ni] = pos_of_step[s];
    en_pos[ni] = pos_of_step[s] + 1;
    sec_of[ni] = L.steps[s].section;
  }
  for (int o : p.outputs) {
    st_pos[o] = fill_pos;
    en_pos[o] = fill_pos + 1;
    sec_of[o] = 2;
  }
  for (size_t s = 0; s < L.steps.size(); ++s) {
    std::vector<int> ins = {L.steps[s].in_a, L.steps[s].in_b};
    ins.insert(ins.end(), L.steps[s].chain_small.begin(), L.steps[s].chain_small.end());
    for (const auto &cm : L.steps[s].chain_merge) ins.push_back(cm.small);
    ins.insert(ins.end(), L.steps[s].outer_in.begin(), L.steps[s].outer_in.end());
    for (int in : ins) {
      int r = root[in];
      if (r < 0) continue;
      int64_t e = pos_of_step[s] + 1;
      if (L.steps[s].section > sec_of[r]) e = end_pos;
      en_pos[r] = std::max(en_pos[r], e);
    }
  }
  int rs = root[p.save];
  if (rs >= 0) en_pos[rs] = end_pos;
  if (keep_all)
    for (size_t i = 0; i < nn; ++i)
      if (root[i] == (int)i) en_pos[i] = end_pos;
  std::vector<int> arena_nodes;
  std::vector<int64_t> S, E, Z, off;
  for (size_t i = 0; i < nn; ++i)
    if (root[i] == (int)i && !absorbed[i]) {
      arena_nodes.push_back((int)i);
      S.push_back(st_pos[i]);
      E.push_back(std::max(en_pos[i], st_pos[i] + 1));
      Z.push_back(L.values[i].elems);
    }
  L.arena_elems = first_fit(S, E, Z, 8, off);
  for (size_t j = 0; j < arena_nodes.size(); ++j) {
    L.values[arena_nodes[j]].slot = off[j];
    L.slots.push_back({arena_nodes[j], off[j], Z[j], S[j], E[j]});
  }
  // propagate slots to views and save
  for (size_t i = 0; i < nn; ++i)
    if (root[i] >= 0 && root[i] != (int)i) L.values[i].slot = L.values[root[i]].slot;
  check_bounds(p, L, root);
  return L;
}

VjpLowered lower_vjp(const Program &p, int64_t batch, int precision, uint64_t fused, bool batch_last) {
  VjpLowered V;
  V.fwd = lower(p, batch, true, precision, fused, false, batch_last);
  const Lowered &L = V.fwd;
  const size_t nn = p.nodes.size();
  // every tensor must feed exactly one consumer (a tree), so each cotangent has
  // one producer per run of its consumer's section
  std::vector<int> uses(nn, 0);
  for (const Node &n : p.nodes) {
    if (n.op == Op::Ncon) uses[n.ia]++, uses[n.ib]++;
    if (n.op == Op::View || n.op == Op::Save) uses[n.ia]++;
  }
  for (size_t i = 0; i < nn; ++i)
    if (uses[i] > 1)
      throw Error(QX_ERR_UNSUPPORTED, "vjp: tensor '" + p.nodes[i].name + "' is used more than once");
  std::vector<int> root(nn, -1), sec(nn, 0);
  for (size_t i = 0; i < nn; ++i) {
    const Node &n = p.nodes[i];
    if (n.op == Op::View || n.op == Op::Save) root[i] = root[n.ia];
    else root[i] = (int)i;
  }
  for (const Step &st : L.steps) sec[st.node] = st.section;
  for (int o : p.outputs) sec[o] = 2;
  V.root = root[p.save];
  if (p.nodes[V.root].op != Op::Ncon) throw Error(QX_ERR_UNSUPPORTED, "vjp: the saved tensor is not a contraction");
  // cotangent buffers: same dense layout as the forward values, no reuse
  V.cot.resize(nn);
  int64_t off = 0;
  for (size_t i = 0; i < nn; ++i) {
    if (p.nodes[i].op != Op::Ncon) continue;
    Value c = L.values[i];
    c.buf = Buf::Cot;
    c.slot = off;
    c.static_off = 0;
    c.var_terms.clear();
    V.cot[i] = c;
    off += (c.elems + 31) / 32 * 32;
  }
  V.cot_elems = off;
  V.bsteps.assign(3, {});
  V.zero_at.assign(3, {});
  auto mark_zero = [&](int node) {
    auto &z = V.zero_at[sec[node]];
    if (std::find(z.begin(), z.end(), node) == z.end()) z.push_back(node);
  };
  if (!L.values[V.root].dep) mark_zero(V.root);   // root cotangent summed over bitstring rows
  std::map<std::string, int> key_index;
  for (size_t i = 0; i < p.keys.size(); ++i) key_index[p.keys[i]] = (int)i;

  for (auto it = L.steps.rbegin(); it != L.steps.rend(); ++it) {
    const Step &fs = *it;
    const Node &cn = p.nodes[fs.node];
    for (int side = 0; side < 2; ++side) {
      const int x = side == 0 ? cn.ia : cn.ib;
      const int rx = root[x];
      const Node &rn = p.nodes[rx];
      if (rn.op == Op::Output) continue;        // bitstring tensors need no gradient
      const Value &XV = L.values[x];
      BwdStep b;
      b.x = x;
      b.direct = (p.nodes[x].op == Op::Ncon) && (sec[x] == fs.section);
      Value out;
      if (b.direct) {
        out = V.cot[x];
      } else {
        out.buf = Buf::Temp;
        out.slot = 0;
        out.dims = XV.dims;
        out.dep = XV.dep;
        int64_t st = 1;
        if (XV.dep) {
          out.beta_stride = 1;
          st = batch;
          b.add_dims.push_back(batch);
          b.add_strides.push_back(XV.beta_stride);
        }
        for (size_t j = 0; j < XV.dims.size(); ++j) {
          out.strides.push_back(XV.dims[j] > 1 ? st : 0);
          if (XV.dims[j] > 1) {
            st *= XV.dims[j];
            b.add_dims.push_back(XV.dims[j]);
            b.add_strides.push_back(XV.strides[j]);
          }
        }
        out.elems = st;
        b.add_elems = st;
        V.temp_elems = std::max(V.temp_elems, st);
        if (rn.op == Op::Load) {
          b.dest = Buf::Grad;
          b.dest_key = key_index.at(rn.key);
        } else {
          b.dest = Buf::Cot;
          b.dest_root = rx;
          mark_zero(rx);
        }
      }
      if (side == 0)
        b.step = make_step(V.cot[fs.node], cn.clab, false, L.values[cn.ib], cn.blab, true, out, cn.alab, batch,
                           precision);
      else
        b.step = make_step(L.values[cn.ia], cn.alab, true, V.cot[fs.node], cn.clab, false, out, cn.blab, batch,
                           precision);
      b.step.node = fs.node;
      b.step.section = fs.section;
      // host-side proof that the step and the add stay inside their buffers
      auto ext = [&](const Value &v, int node) -> std::pair<int64_t, int64_t> {
        int64_t mx = v.static_off;
        for (auto &t : v.var_terms) mx += (p.var_dims[t.first] - 1) * t.second;
        int64_t cap;
        if (v.buf == Buf::Data) cap = p.key_elems.at(p.keys[v.data_key]);
        else if (v.buf == Buf::Cot) cap = V.cot[node].elems;
        else if (v.buf == Buf::Temp) cap = v.elems;
        else cap = L.values[root[node]].elems;
        return {mx, cap};
      };
      const Value &cv = V.cot[fs.node];
      std::pair<int64_t, int64_t> ea = side == 0 ? ext(cv, fs.node) : ext(L.values[cn.ia], cn.ia);
      std::pair<int64_t, int64_t> eb = side == 0 ? ext(L.values[cn.ib], cn.ib) : ext(cv, fs.node);
      std::pair<int64_t, int64_t> ec = ext(out, x);
      int64_t ma = 0, mb = 0, mc = 0;
      for (auto &g : b.step.cl) ma += (g.dim - 1) * g.sa, mb += (g.dim - 1) * g.sb, mc += (g.dim - 1) * g.sc;
      for (auto &g : b.step.kl) ma += (g.dim - 1) * g.sa, mb += (g.dim - 1) * g.sb;
      if (ea.first + ma >= ea.second || eb.first + mb >= eb.second || ec.first + mc >= ec.second)
        throw Error(QX_ERR_STATE, "internal: backward step of '" + cn.name + "' addresses outside its operands");
      if (!b.direct) {
        int64_t md = 0;
        for (size_t j = 0; j < b.add_dims.size(); ++j) md += (b.add_dims[j] - 1) * b.add_strides[j];
        std::pair<int64_t, int64_t> ex = ext(XV, x);
        int64_t cap = b.dest == Buf::Grad ? p.key_elems.at(p.keys[b.dest_key]) : V.cot[b.dest_root].elems;
        if (ex.first + md >= cap)
          throw Error(QX_ERR_STATE, "internal: cotangent add of '" + p.nodes[x].name + "' outside its buffer");
      }
      V.fwd.ws_elems = std::max(V.fwd.ws_elems, b.step.ws_elems);
      V.bsteps[fs.section].push_back(std::move(b));
    }
  }

  // Backward steps of sections 0 and 2 in dependency levels: a step reads the
  // cotangent written one level earlier.  Small ones whose operands have no
  // slice-dependent offsets and that write a cotangent directly, or a temp added
  // into a data key's gradient that no other step adds into, run batched per level
  // (each with its own temp); the order is stable per level.
  std::vector<int> key_uses(p.keys.size(), 0);
  for (auto &sec : V.bsteps)
    for (auto &b : sec)
      if (!b.direct && b.dest == Buf::Grad) key_uses[b.dest_key]++;
  int level_base = 0;
  for (int sec : {0, 2}) {
    auto &bv = V.bsteps[sec];
    std::map<int, size_t> prod;   // node -> backward step writing its cotangent
    for (size_t i = 0; i < bv.size(); ++i)
      if (bv[i].direct) prod[bv[i].x] = i;
    std::vector<int> lev(bv.size(), 0);
    int maxlev = 0;
    for (size_t i = 0; i < bv.size(); ++i) {
      auto it = prod.find(bv[i].step.node);
      lev[i] = it == prod.end() ? 0 : lev[it->second] + 1;
      maxlev = std::max(maxlev, lev[i]);
    }
    auto no_slice_terms = [&](const BwdStep &b) {
      const Node &cn = p.nodes[b.step.node];
      return L.values[cn.ia].var_terms.empty() && L.values[cn.ib].var_terms.empty() &&
             L.values[b.x].var_terms.empty();
    };
    std::vector<size_t> order(bv.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return lev[a] < lev[b]; });
    std::vector<BwdStep> sorted;
    std::map<int, int64_t> level_temp;
    std::map<int, int> level_count;
    std::set<std::pair<int, int>> level_cot;   // (level, cotangent buffer) added into by the level
    for (size_t i : order) {
      BwdStep b = std::move(bv[i]);
      // adds of one level go out in one launch: each destination at most once (a key's
      // gradient has one adding step; a cotangent buffer one per level)
      bool ok = small_step(b.step, true) && no_slice_terms(b) &&
                (b.direct || (b.add_dims.size() <= (size_t)kSmallMaxBits &&
                              ((b.dest == Buf::Grad && key_uses[b.dest_key] == 1) ||
                               (b.dest == Buf::Cot && !level_cot.count({level_base + lev[i], b.dest_root})))));
      if (ok && !b.direct && b.dest == Buf::Cot) level_cot.insert({level_base + lev[i], b.dest_root});
      if (ok) {
        b.level = level_base + lev[i];
        level_count[b.level]++;
        if (!b.direct) {
          b.temp_off = level_temp[b.level];
          level_temp[b.level] += (b.add_elems + 31) / 32 * 32;
          V.temp_elems = std::max(V.temp_elems, level_temp[b.level]);
        }
      }
      sorted.push_back(std::move(b));
    }
    for (auto &b : sorted)   // a lone step keeps its own launch
      if (b.level >= 0 && level_count[b.level] < 2) b.level = -1, b.temp_off = 0;
    bv = std::move(sorted);
    level_base += maxlev + 1;
  }
  return V;
}

}  // namespace qx
