// gfx950 kernels of libqxhip.
//
// Pairwise contraction C[c] = sum_K A[a] B[b] with every index permutation folded
// into the address arithmetic: with all leg dims powers of two, the offset of a C
// element in A (or B) is a weighted sum of the bits of its linear index, so a
// workgroup that owns an aligned run of 2^t C elements gets the high part of every
// offset once and each lane only adds the weights of its t low bits.  Lanes walk
// C's fastest legs, so C stores are always coalesced and A/B loads are coalesced
// whenever C's fastest legs are theirs too (the batch leg, laid out fastest in
// every bitstring-dependent tensor, guarantees it for batched steps).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include <type_traits>

#include <vector>

#include "qx_internal.hpp"

namespace qx {
namespace {

template <typename T> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

template <typename T>
__device__ __forceinline__ void cmac(typename Vec<T>::type &acc, typename Vec<T>::type a,
                                     typename Vec<T>::type b, int conj_a, int conj_b) {
  if (conj_a) a.y = -a.y;
  if (conj_b) b.y = -b.y;
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

constexpr int kThreads = 256;
constexpr int kElems = 4;            // C elements per lane
constexpr int kTileBits = 10;        // log2(kThreads * kElems)
constexpr int kKTabBits = 9;         // low K bits tabulated in LDS

template <typename T>
__global__ __launch_bounds__(kThreads) void k_contract_bits(const typename Vec<T>::type *__restrict__ A,
                                                            const typename Vec<T>::type *__restrict__ B,
                                                            typename Vec<T>::type *__restrict__ C,
                                                            int64_t c_elems, BitArgs a) {
  using V = typename Vec<T>::type;
  __shared__ int64_t tka[1 << kKTabBits], tkb[1 << kKTabBits];
  const int nk_lo = a.nbk < kKTabBits ? a.nbk : kKTabBits;
  const int nkt = 1 << nk_lo;
  for (int k = threadIdx.x; k < nkt; k += kThreads) {
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < nk_lo; ++b)
      if ((k >> b) & 1) {
        oa += a.ka[b];
        ob += a.kb[b];
      }
    tka[k] = oa;
    tkb[k] = ob;
  }
  __syncthreads();
  const int64_t nk_hi_all = int64_t(1) << (a.nbk - nk_lo);
  const int64_t nk_hi = nk_hi_all / a.ksplit;   // high K iterations of one split
  const int64_t tile = int64_t(1) << kTileBits;
  const int64_t ctiles = (c_elems + tile - 1) / tile;
  // per-lane low offsets for the kElems elements of this lane
  int64_t loa[kElems], lob[kElems];
#pragma unroll
  for (int e = 0; e < kElems; ++e) {
    int j = e * kThreads + threadIdx.x;
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < kTileBits && b < a.nbc; ++b)
      if ((j >> b) & 1) {
        oa += a.wa[b];
        ob += a.wb[b];
      }
    loa[e] = oa;
    lob[e] = ob;
  }
  // work unit = (C tile, K split); split s writes its partial at C + s * c_elems
  for (int64_t unit = blockIdx.x; unit < ctiles * a.ksplit; unit += gridDim.x) {
    const int64_t base = (unit % ctiles) * tile;
    const int64_t split = unit / ctiles;
    int64_t hia = 0, hib = 0;
    for (int b = kTileBits; b < a.nbc; ++b)
      if ((base >> b) & 1) {
        hia += a.wa[b];
        hib += a.wb[b];
      }
    V acc[kElems];
#pragma unroll
    for (int e = 0; e < kElems; ++e) acc[e] = V{0, 0};
    for (int64_t kh = split * nk_hi; kh < (split + 1) * nk_hi; ++kh) {
      int64_t oka = 0, okb = 0;
      for (int b = nk_lo; b < a.nbk; ++b)
        if ((kh >> (b - nk_lo)) & 1) {
          oka += a.ka[b];
          okb += a.kb[b];
        }
      for (int k = 0; k < nkt; ++k) {
        const int64_t ka = hia + oka + tka[k], kb = hib + okb + tkb[k];
#pragma unroll
        for (int e = 0; e < kElems; ++e) {
          if (base + e * kThreads + threadIdx.x < c_elems)
            cmac<T>(acc[e], A[ka + loa[e]], B[kb + lob[e]], a.conj_a, a.conj_b);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < kElems; ++e) {
      int64_t idx = base + e * kThreads + threadIdx.x;
      if (idx < c_elems) C[split * c_elems + idx] = acc[e];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_contract_radix(const typename Vec<T>::type *__restrict__ A,
                                                             const typename Vec<T>::type *__restrict__ B,
                                                             typename Vec<T>::type *__restrict__ C,
                                                             int64_t c_elems, RadixArgs a) {
  using V = typename Vec<T>::type;
  int64_t kelems = 1;
  for (int j = 0; j < a.nkl; ++j) kelems *= a.kd[j];
  for (int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x; idx < c_elems;
       idx += (int64_t)gridDim.x * kThreads) {
    int64_t r = idx, oa = 0, ob = 0;
    for (int j = 0; j < a.ncl; ++j) {
      int64_t d = r % a.cd[j];
      r /= a.cd[j];
      oa += d * a.ca[j];
      ob += d * a.cb[j];
    }
    V acc{0, 0};
    for (int64_t k = 0; k < kelems; ++k) {
      int64_t rk = k, ka = 0, kb = 0;
      for (int j = 0; j < a.nkl; ++j) {
        int64_t d = rk % a.kd[j];
        rk /= a.kd[j];
        ka += d * a.ka[j];
        kb += d * a.kb[j];
      }
      cmac<T>(acc, A[oa + ka], B[ob + kb], a.conj_a, a.conj_b);
    }
    C[idx] = acc;
  }
}


// ---------------------------------------------------------------------------
// Tiled complex GEMM over bit-weight index maps: C[h, m, n] = sum_k A[h,m,k] B[h,k,n].
// A TM x TK and a TK x TN tile are gathered into LDS per K chunk, lanes walking
// whichever of the tile's index bits has the smaller stride in that operand (so the
// loads are coalesced whatever the permutation), each lane accumulates an RM x RN
// register micro-tile, and the C tile is staged through LDS so the stores walk C's
// fastest tile bits.
template <int TM, int TN, int TK>
struct GemmTile {
  static constexpr int NT = 256;
  static constexpr int RM = 4;
  static constexpr int RN = (TM * TN) / (NT * RM);
  static constexpr int TX = TM / RM;
  static constexpr int TY = TN / RN;
  static_assert(TX * TY == NT, "tile shape");
  static constexpr int LDS_AB = TK * (TM + TN);
  static constexpr int LDS_C = TM * TN;
  static constexpr int LDS = LDS_AB > LDS_C ? LDS_AB : LDS_C;
};

__device__ __forceinline__ int ilog2_dev(int x) { return 31 - __clz(x); }

template <typename T, int TM, int TN, int TK>
__global__ __launch_bounds__(256) void k_contract_gemm(const typename Vec<T>::type *__restrict__ A,
                                                       const typename Vec<T>::type *__restrict__ B,
                                                       typename Vec<T>::type *__restrict__ C, GemmArgs a) {
  using V = typename Vec<T>::type;
  using G = GemmTile<TM, TN, TK>;
  constexpr int PA = TM + 1, PB = TN + 1, PC = TN + 1;          // padded LDS rows
  constexpr int LDS_AB = TK * (PA + PB), LDS_C = TM * PC;
  constexpr int LDS = LDS_AB > LDS_C ? LDS_AB : LDS_C;
  constexpr int EA = TM * TK / 256, EB = TN * TK / 256;         // tile elements per lane
  static_assert(EA * 256 == TM * TK && EB * 256 == TN * TK, "tile/threads");
  __shared__ V lds[LDS];
  __shared__ int64_t t_am[TM], t_cm[TM], t_bn[TN], t_cn[TN], t_ak[TK], t_bk[TK];
  V *sA = lds;             // [TK][PA]
  V *sB = lds + TK * PA;   // [TK][PB]
  const int tid = threadIdx.x;
  const int lm = ilog2_dev(TM), ln = ilog2_dev(TN), lk = ilog2_dev(TK);
  const int rows = a.nbm < lm ? (1 << a.nbm) : TM;
  const int cols = a.nbn < ln ? (1 << a.nbn) : TN;
  const int kk = a.nbk < lk ? (1 << a.nbk) : TK;
  const int64_t tiles_m = a.nbm > lm ? (int64_t(1) << (a.nbm - lm)) : 1;
  const int64_t tiles_n = a.nbn > ln ? (int64_t(1) << (a.nbn - ln)) : 1;
  const int64_t chunks_all = a.nbk > lk ? (int64_t(1) << (a.nbk - lk)) : 1;
  const int64_t chunks = chunks_all / a.ksplit;                    // K chunks of one unit
  const int64_t ntiles = tiles_m * tiles_n * (int64_t(1) << a.nbh) * a.ksplit;   // units

  for (int i = tid; i < TM; i += 256) {
    int64_t oa = 0, oc = 0;
    for (int b = 0; b < lm && b < a.nbm; ++b)
      if ((i >> b) & 1) oa += a.ma[b], oc += a.mc[b];
    t_am[i] = oa;
    t_cm[i] = oc;
  }
  for (int j = tid; j < TN; j += 256) {
    int64_t ob = 0, oc = 0;
    for (int b = 0; b < ln && b < a.nbn; ++b)
      if ((j >> b) & 1) ob += a.nb[b], oc += a.nc[b];
    t_bn[j] = ob;
    t_cn[j] = oc;
  }
  for (int l = tid; l < TK; l += 256) {
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < lk && b < a.nbk; ++b)
      if ((l >> b) & 1) oa += a.ka[b], ob += a.kb[b];
    t_ak[l] = oa;
    t_bk[l] = ob;
  }
  __syncthreads();

  // fixed per-lane gather offsets inside a tile, and LDS slots they fill
  int64_t offa[EA], offb[EB];
  int slota[EA], slotb[EB];
  bool oka[EA], okb[EB];
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    int e = tid + 256 * j, i, l;
    if (a.a_kfast) i = e / TK, l = e % TK;
    else l = e / TM, i = e % TM;
    oka[j] = i < rows && l < kk;
    offa[j] = oka[j] ? t_am[i] + t_ak[l] : 0;
    slota[j] = l * PA + i;
  }
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    int e = tid + 256 * j, jj, l;
    if (a.b_kfast) jj = e / TK, l = e % TK;
    else l = e / TN, jj = e % TN;
    okb[j] = jj < cols && l < kk;
    offb[j] = okb[j] ? t_bn[jj] + t_bk[l] : 0;
    slotb[j] = l * PB + jj;
  }

  const int64_t nblk = gridDim.x;
  int64_t bid = blockIdx.x;
  if (nblk % 8 == 0) bid = (bid % 8) * (nblk / 8) + bid / 8;   // XCD-contiguous tile ranges
  const int64_t per = (ntiles + nblk - 1) / nblk;
  const int64_t t_begin = bid * per;
  const int64_t t_end = t_begin + per < ntiles ? t_begin + per : ntiles;
  const int64_t items = t_begin < t_end ? (t_end - t_begin) * chunks : 0;

  // a unit is (tile, K split); the split's partial goes to C + split * c_elems
  auto tile_offsets = [&](int64_t unit, int64_t &oa, int64_t &ob, int64_t &oc) {
    const int64_t split = unit % a.ksplit;
    const int64_t tile = unit / a.ksplit;
    int64_t nt = tile % tiles_n;
    int64_t rest = tile / tiles_n;
    int64_t mt = rest % tiles_m;
    int64_t h = rest / tiles_m;
    oa = 0, ob = 0, oc = 0;
    for (int b = 0; b < a.nbh; ++b)
      if ((h >> b) & 1) oa += a.ha[b], ob += a.hb[b], oc += a.hc[b];
    for (int b = lm; b < a.nbm; ++b)
      if (((mt << lm) >> b) & 1) oa += a.ma[b], oc += a.mc[b];
    for (int b = ln; b < a.nbn; ++b)
      if (((nt << ln) >> b) & 1) ob += a.nb[b], oc += a.nc[b];
    oc += split * a.c_elems;
  };
  V ra[EA], rb[EB];
  auto fetch = [&](int64_t item) {
    int64_t unit = t_begin + item / chunks, oa, ob, oc;
    int64_t kc = (unit % a.ksplit) * chunks + item % chunks;
    tile_offsets(unit, oa, ob, oc);
    for (int b = lk; b < a.nbk; ++b)
      if (((kc << lk) >> b) & 1) oa += a.ka[b], ob += a.kb[b];
#pragma unroll
    for (int j = 0; j < EA; ++j) ra[j] = oka[j] ? A[oa + offa[j]] : V{0, 0};
#pragma unroll
    for (int j = 0; j < EB; ++j) rb[j] = okb[j] ? B[ob + offb[j]] : V{0, 0};
  };

  const int tx = tid % G::TX, ty = tid / G::TX;
  V acc[G::RM][G::RN];
  if (items > 0) fetch(0);
  for (int64_t it = 0; it < items; ++it) {
    const int64_t kc = it % chunks;
    if (kc == 0) {
#pragma unroll
      for (int r = 0; r < G::RM; ++r)
#pragma unroll
        for (int c = 0; c < G::RN; ++c) acc[r][c] = V{0, 0};
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      V v = ra[j];
      if (a.conj_a) v.y = -v.y;
      sA[slota[j]] = v;
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      V v = rb[j];
      if (a.conj_b) v.y = -v.y;
      sB[slotb[j]] = v;
    }
    __syncthreads();
    if (it + 1 < items) fetch(it + 1);   // next tile/chunk loads overlap this chunk's FMAs
#pragma unroll 4
    for (int l = 0; l < TK; ++l) {
      V av[G::RM], bv[G::RN];
#pragma unroll
      for (int r = 0; r < G::RM; ++r) av[r] = sA[l * PA + tx + r * G::TX];
#pragma unroll
      for (int c = 0; c < G::RN; ++c) bv[c] = sB[l * PB + ty + c * G::TY];
#pragma unroll
      for (int r = 0; r < G::RM; ++r)
#pragma unroll
        for (int c = 0; c < G::RN; ++c) cmac<T>(acc[r][c], av[r], bv[c], 0, 0);
    }
    if (kc == chunks - 1) {
      int64_t oa, ob, oc;
      tile_offsets(t_begin + it / chunks, oa, ob, oc);
      __syncthreads();
      V *sC = lds;  // [TM][PC]
#pragma unroll
      for (int r = 0; r < G::RM; ++r)
#pragma unroll
        for (int c = 0; c < G::RN; ++c) sC[(tx + r * G::TX) * PC + ty + c * G::TY] = acc[r][c];
      __syncthreads();
      V *Cb = C + oc;
      for (int e = tid; e < TM * TN; e += 256) {
        int i, j;
        if (a.c_nfast) i = e / TN, j = e % TN;
        else j = e / TM, i = e % TM;
        if (i < rows && j < cols) Cb[t_cm[i] + t_cn[j]] = sC[i * PC + j];
      }
    }
  }
}


// 16x16x4 matrix-core products.  Both take A[i][k] from lane i + 16 k and B[k][j]
// from lane j + 16 k; lane l holds D[row(l / 16, r)][l % 16] in element r, with
// row = 4 (l / 16) + r for FP32 and 4 r + l / 16 for FP64.
template <typename T> struct Mfma16;
template <> struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * r + lk; }
};
template <> struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

// Matrix-core rate probe: each wave runs `iters` rounds (a multiple of 16) of four
// independent 16x16x4 products on register operands (no memory traffic in the loop;
// 2 * 16 * 16 * 4 flops each), and each lane stores one sum at the end so that the
// work stays live.  The loop body holds 16 rounds, so that the loop's own instructions
// are few beside the products.
template <typename T>
__global__ __launch_bounds__(256) void k_mfma_probe(T *out, int iters, T seed) {
  typedef typename Mfma16<T>::v4 v4;
  v4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = v4{T(0), T(0), T(0), T(0)};
  const T a = seed + T(threadIdx.x) * T(1e-7), b = seed - T(threadIdx.x) * T(1e-7);
  for (int i = 0; i < iters; i += 16) {
#pragma unroll
    for (int u = 0; u < 16; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] = Mfma16<T>::mma(a, b, acc[j]);
  }
  T s = T(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) s += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  out[(int64_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Complex GEMM on 16x16x4 matrix cores (FP32 for complex64, FP64 for complex128),
// same bit-weight tile gathers as k_contract_gemm.  A 64x64 complex C tile per
// workgroup, 4 waves of 32x32 (2x2 blocks of 16x16).  4M: Re += Ar.Br - Ai.Bi and
// Im += Ar.Bi + Ai.Br as four real products per block; 3M (Gauss, complex64 only):
// RR = Ar.Br, II = Ai.Bi, SS = (Ar+Ai).(Br+Bi), Re = RR - II, Im = SS - RR - II.
// Accumulation in the operand precision.  Operand and result layouts as in Mfma16.
template <typename T, bool GAUSS, int TK>
__global__ __launch_bounds__(256) void k_contract_gemm_mfma(const typename Vec<T>::type *__restrict__ A,
                                                            const typename Vec<T>::type *__restrict__ B,
                                                            typename Vec<T>::type *__restrict__ C, GemmArgs a) {
  using V = typename Vec<T>::type;
  using M = Mfma16<T>;
  using f32x4 = typename M::v4;
  constexpr int TM = 64, TN = 64;
  constexpr int PA = TM + 1, PB = TN + 1, PC = TN + 1;
  constexpr int LDS_AB = TK * (PA + PB), LDS_C = TM * PC;
  constexpr int LDS = LDS_AB > LDS_C ? LDS_AB : LDS_C;
  constexpr int EA = TM * TK / 256, EB = TN * TK / 256;
  __shared__ V lds[LDS];
  __shared__ int64_t t_am[TM], t_cm[TM], t_bn[TN], t_cn[TN], t_ak[TK], t_bk[TK];
  V *sA = lds;
  V *sB = lds + TK * PA;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lm = 6, ln = 6, lk = TK == 16 ? 4 : 5;
  const int rows = a.nbm < lm ? (1 << a.nbm) : TM;
  const int cols = a.nbn < ln ? (1 << a.nbn) : TN;
  const int kk = a.nbk < lk ? (1 << a.nbk) : TK;
  const int64_t tiles_m = a.nbm > lm ? (int64_t(1) << (a.nbm - lm)) : 1;
  const int64_t tiles_n = a.nbn > ln ? (int64_t(1) << (a.nbn - ln)) : 1;
  const int64_t chunks_all = a.nbk > lk ? (int64_t(1) << (a.nbk - lk)) : 1;
  const int64_t chunks = chunks_all / a.ksplit;
  const int64_t ntiles = tiles_m * tiles_n * (int64_t(1) << a.nbh) * a.ksplit;
  for (int i = tid; i < TM; i += 256) {
    int64_t oa = 0, oc = 0;
    for (int b = 0; b < lm && b < a.nbm; ++b)
      if ((i >> b) & 1) oa += a.ma[b], oc += a.mc[b];
    t_am[i] = oa;
    t_cm[i] = oc;
  }
  for (int j = tid; j < TN; j += 256) {
    int64_t ob = 0, oc = 0;
    for (int b = 0; b < ln && b < a.nbn; ++b)
      if ((j >> b) & 1) ob += a.nb[b], oc += a.nc[b];
    t_bn[j] = ob;
    t_cn[j] = oc;
  }
  for (int l = tid; l < TK; l += 256) {
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < lk && b < a.nbk; ++b)
      if ((l >> b) & 1) oa += a.ka[b], ob += a.kb[b];
    t_ak[l] = oa;
    t_bk[l] = ob;
  }
  __syncthreads();
  int64_t offa[EA], offb[EB];
  int slota[EA], slotb[EB];
  bool oka[EA], okb[EB];
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    int e = tid + 256 * j, i, l;
    if (a.a_kfast) i = e / TK, l = e % TK;
    else l = e / TM, i = e % TM;
    oka[j] = i < rows && l < kk;
    offa[j] = oka[j] ? t_am[i] + t_ak[l] : 0;
    slota[j] = l * PA + i;
  }
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    int e = tid + 256 * j, jj, l;
    if (a.b_kfast) jj = e / TK, l = e % TK;
    else l = e / TN, jj = e % TN;
    okb[j] = jj < cols && l < kk;
    offb[j] = okb[j] ? t_bn[jj] + t_bk[l] : 0;
    slotb[j] = l * PB + jj;
  }
  const int64_t nblk = gridDim.x;
  int64_t bid = blockIdx.x;
  if (nblk % 8 == 0) bid = (bid % 8) * (nblk / 8) + bid / 8;
  const int64_t per = (ntiles + nblk - 1) / nblk;
  const int64_t t_begin = bid * per;
  const int64_t t_end = t_begin + per < ntiles ? t_begin + per : ntiles;
  const int64_t items = t_begin < t_end ? (t_end - t_begin) * chunks : 0;
  auto tile_offsets = [&](int64_t unit, int64_t &oa, int64_t &ob, int64_t &oc) {
    const int64_t split = unit % a.ksplit;
    const int64_t tile = unit / a.ksplit;
    int64_t nt = tile % tiles_n;
    int64_t rest = tile / tiles_n;
    int64_t mt = rest % tiles_m;
    int64_t h = rest / tiles_m;
    oa = 0, ob = 0, oc = 0;
    for (int b = 0; b < a.nbh; ++b)
      if ((h >> b) & 1) oa += a.ha[b], ob += a.hb[b], oc += a.hc[b];
    for (int b = lm; b < a.nbm; ++b)
      if (((mt << lm) >> b) & 1) oa += a.ma[b], oc += a.mc[b];
    for (int b = ln; b < a.nbn; ++b)
      if (((nt << ln) >> b) & 1) ob += a.nb[b], oc += a.nc[b];
    oc += split * a.c_elems;
  };
  V ra[EA], rb[EB];
  auto fetch = [&](int64_t item) {
    int64_t unit = t_begin + item / chunks, oa, ob, oc;
    int64_t kc = (unit % a.ksplit) * chunks + item % chunks;
    tile_offsets(unit, oa, ob, oc);
    for (int b = lk; b < a.nbk; ++b)
      if (((kc << lk) >> b) & 1) oa += a.ka[b], ob += a.kb[b];
#pragma unroll
    for (int j = 0; j < EA; ++j) ra[j] = oka[j] ? A[oa + offa[j]] : V{0, 0};
#pragma unroll
    for (int j = 0; j < EB; ++j) rb[j] = okb[j] ? B[ob + offb[j]] : V{0, 0};
  };
  // wave (wm, wn) owns rows 32*wm.., cols 32*wn..; block (bi, bj) of 16x16
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lq = lane >> 4;
  f32x4 acc_rr[2][2], acc_ii[2][2], acc_x[2][2];
  if (items > 0) fetch(0);
  for (int64_t it = 0; it < items; ++it) {
    const int64_t kc = it % chunks;
    if (kc == 0) {
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          acc_rr[bi][bj] = f32x4{0, 0, 0, 0};
          acc_ii[bi][bj] = f32x4{0, 0, 0, 0};
          acc_x[bi][bj] = f32x4{0, 0, 0, 0};
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < EA; ++j) {
      V v = ra[j];
      if (a.conj_a) v.y = -v.y;
      sA[slota[j]] = v;
    }
#pragma unroll
    for (int j = 0; j < EB; ++j) {
      V v = rb[j];
      if (a.conj_b) v.y = -v.y;
      sB[slotb[j]] = v;
    }
    __syncthreads();
    if (it + 1 < items) fetch(it + 1);
#pragma unroll
    for (int k0 = 0; k0 < TK; k0 += 4) {
      V fa[2], fb[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) fa[bi] = sA[(k0 + lq) * PA + 32 * wm + 16 * bi + li];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) fb[bj] = sB[(k0 + lq) * PB + 32 * wn + 16 * bj + li];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          acc_rr[bi][bj] = M::mma(fa[bi].x, fb[bj].x, acc_rr[bi][bj]);
          acc_ii[bi][bj] = M::mma(fa[bi].y, fb[bj].y, acc_ii[bi][bj]);
          if (GAUSS) {
            acc_x[bi][bj] = M::mma(fa[bi].x + fa[bi].y, fb[bj].x + fb[bj].y, acc_x[bi][bj]);
          } else {
            acc_x[bi][bj] = M::mma(fa[bi].x, fb[bj].y, acc_x[bi][bj]);
            acc_x[bi][bj] = M::mma(fa[bi].y, fb[bj].x, acc_x[bi][bj]);
          }
        }
    }
    if (kc == chunks - 1) {
      int64_t oa, ob, oc;
      tile_offsets(t_begin + it / chunks, oa, ob, oc);
      __syncthreads();
      V *sC = lds;
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 32 * wm + 16 * bi + M::row(lq, r), j = 32 * wn + 16 * bj + li;
            const T rr = acc_rr[bi][bj][r], ii = acc_ii[bi][bj][r], x = acc_x[bi][bj][r];
            sC[i * PC + j] = GAUSS ? V{rr - ii, x - rr - ii} : V{rr - ii, x};
          }
      __syncthreads();
      V *Cb = C + oc;
      for (int e = tid; e < TM * TN; e += 256) {
        int i, j;
        if (a.c_nfast) i = e / TN, j = e % TN;
        else j = e / TM, i = e % TM;
        if (i < rows && j < cols) Cb[t_cm[i] + t_cn[j]] = sC[i * PC + j];
      }
    }
  }
}

// C[i] = sum over K-split partials ws[s * n + i].
template <typename T>
__global__ void k_sum_splits(typename Vec<T>::type *__restrict__ C, const typename Vec<T>::type *__restrict__ ws,
                             int64_t n, int ksplit) {
  using V = typename Vec<T>::type;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    V acc = ws[i];
    for (int s = 1; s < ksplit; ++s) {
      V v = ws[(int64_t)s * n + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    C[i] = acc;
  }
}

// One wave per C element, the K range split over lanes, shuffle reduction.
template <typename T>
__global__ __launch_bounds__(256) void k_contract_reduce(const typename Vec<T>::type *__restrict__ A,
                                                         const typename Vec<T>::type *__restrict__ B,
                                                         typename Vec<T>::type *__restrict__ C,
                                                         int64_t c_elems, BitArgs a) {
  using V = typename Vec<T>::type;
  const int lane = threadIdx.x & 63;
  const int64_t nk = int64_t(1) << a.nbk;
  const int64_t kper = nk / a.ksplit;         // K range of one wave
  // per-lane offsets of the low 6 k bits
  int64_t la = 0, lb = 0;
  for (int b = 0; b < 6 && b < a.nbk; ++b)
    if ((lane >> b) & 1) la += a.ka[b], lb += a.kb[b];
  const int64_t waves = (int64_t)gridDim.x * 4;
  const int64_t units = c_elems * a.ksplit;   // (C element, K range); partials land at C[split*c_elems + idx]
  for (int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); u < units; u += waves) {
    const int64_t idx = u % c_elems, split = u / c_elems;
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < a.nbc; ++b)
      if ((idx >> b) & 1) oa += a.wa[b], ob += a.wb[b];
    V acc{0, 0};
    for (int64_t k0 = split * kper; k0 < (split + 1) * kper; k0 += 64) {
      if (k0 + lane >= nk) break;
      int64_t ha = 0, hb = 0;
      for (int b = 6; b < a.nbk; ++b)
        if ((k0 >> b) & 1) ha += a.ka[b], hb += a.kb[b];
      cmac<T>(acc, A[oa + ha + la], B[ob + hb + lb], a.conj_a, a.conj_b);
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc.x += __shfl_xor(acc.x, off, 64);
      acc.y += __shfl_xor(acc.y, off, 64);
    }
    if (lane == 0) C[split * c_elems + idx] = acc;
  }
}


// Row kernel for K <= 32, N <= 64 on many rows: each lane owns one row r (H and
// M bits, lanes along C's fastest bits) and issues all NK loads of its A row at
// once, then produces the row's outputs NC at a time.  B is staged in LDS once per
// workgroup -- one K x N slab per value of the H bits that vary across the
// workgroup's lanes -- so its reads are LDS broadcasts; the LDS size is set at
// launch to what the slabs need.
constexpr int kRowLdsElems = 4096;  // complex elements of staged B at most

template <typename T, int NK, int NC, bool UB>
__global__ __launch_bounds__(256) void k_contract_rows(const typename Vec<T>::type *__restrict__ A,
                                                       const typename Vec<T>::type *__restrict__ B,
                                                       typename Vec<T>::type *__restrict__ C, RowArgs a) {
  using V = typename Vec<T>::type;
  constexpr int LB = 8;  // row bits resolved per lane
  extern __shared__ __align__(16) unsigned char rows_smem[];
  V *sB = reinterpret_cast<V *>(rows_smem);
  const int nn = a.nn, kn = NK * nn;
  const int64_t nrows = int64_t(1) << a.nbr;
  const int lowbits = a.nbr < LB ? a.nbr : LB;
  int64_t la = 0, lc = 0;
  int slab = 0, nslab_bits = 0;
  for (int b = 0; b < lowbits; ++b) {
    const bool on = (threadIdx.x >> b) & 1;
    if (on) la += a.ra[b], lc += a.rc[b];
    if ((a.hlo >> b) & 1) {
      if (on) slab |= 1 << nslab_bits;
      ++nslab_bits;
    }
  }
  const int nslab = 1 << nslab_bits;
  for (int64_t base = (int64_t)blockIdx.x * 256; base < nrows; base += (int64_t)gridDim.x * 256) {
    int64_t oa = la, obh = 0, oc = lc;
    for (int b = LB; b < a.nbr; ++b)
      if ((base >> b) & 1) oa += a.ra[b], obh += a.rb[b], oc += a.rc[b];
    const bool live = base + threadIdx.x < nrows;
    V av[NK];
    if (live) {
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        av[k] = A[oa + a.ka[k]];
        if (a.conj_a) av[k].y = -av[k].y;
      }
    }
    if (UB) {
      // no H bit varies inside the workgroup: B addresses are wave-uniform
      const V *Bu = B + obh;
      if (live) {
        V *Cr = C + oc;
        for (int n0 = 0; n0 < nn; n0 += NC) {
          V acc[NC];
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] = V{0, 0};
#pragma unroll
          for (int k = 0; k < NK; ++k) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if (n0 + c < nn) {
                V bv = Bu[a.kb[k] + a.nb[n0 + c]];
                if (a.conj_b) bv.y = -bv.y;
                cmac<T>(acc[c], av[k], bv, 0, 0);
              }
          }
#pragma unroll
          for (int c = 0; c < NC; ++c)
            if (n0 + c < nn) Cr[a.nc[n0 + c]] = acc[c];
        }
      }
      continue;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nslab * kn; e += 256) {
      const int sl = e / kn, r = e % kn, k = r / nn, n = r % nn;
      int64_t ob = obh;
      for (int b = 0, j = 0; b < lowbits; ++b)
        if ((a.hlo >> b) & 1) {
          if ((sl >> j) & 1) ob += a.rb[b];
          ++j;
        }
      V v = B[ob + a.kb[k] + a.nb[n]];
      if (a.conj_b) v.y = -v.y;
      sB[e] = v;
    }
    __syncthreads();
    if (live) {
      const V *Bs = sB + slab * kn;
      V *Cr = C + oc;
      for (int n0 = 0; n0 < nn; n0 += NC) {
        V acc[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] = V{0, 0};
#pragma unroll
        for (int k = 0; k < NK; ++k) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
            if (n0 + c < nn) cmac<T>(acc[c], av[k], Bs[k * nn + n0 + c], 0, 0);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (n0 + c < nn) Cr[a.nc[n0 + c]] = acc[c];
      }
    }
  }
}

// dst[map(e)] += src[e]: src dense over dims (first fastest), map(e) = sum digit*stride.
struct AddArgs {
  int64_t dim[kMaxLegs], st[kMaxLegs];
  int n = 0;
};

template <typename TS, typename TD>
__global__ void k_add_mapped(typename Vec<TD>::type *__restrict__ dst, const typename Vec<TS>::type *__restrict__ src,
                             int64_t n, AddArgs a) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = e, o = 0;
    for (int j = 0; j < a.n; ++j) {
      int64_t d = r % a.dim[j];
      r /= a.dim[j];
      o += d * a.st[j];
    }
    auto v = src[e];
    auto w = dst[o];
    w.x += (TD)v.x;
    w.y += (TD)v.y;
    dst[o] = w;
  }
}

// Cotangent of the loss sum_b sum_e |amp_{b,e}|^2 at the root: 2 amp, read from the
// root's forward value of the batch (root value and cotangent share one layout); the
// batch's amplitudes are also written out in FP64 (rows past nvalid get a zero
// cotangent).
template <typename T>
__global__ void k_root_cot_norm(typename Vec<T>::type *__restrict__ dst, const typename Vec<T>::type *__restrict__ val,
                                int64_t dep, int64_t batch, int64_t nvalid, int64_t relems,
                                const int64_t *__restrict__ relem_off, double *__restrict__ amp, int64_t row0) {
  using V = typename Vec<T>::type;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < batch * relems;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t beta = t % batch, e = t / batch;
    const int64_t o = relem_off[e] + beta * dep;
    V c{0, 0};
    if (beta < nvalid) {
      const V v = val[o];
      c = V{v.x + v.x, v.y + v.y};
      double *p = amp + 2 * ((row0 + beta) * relems + e);
      p[0] = (double)v.x;
      p[1] = (double)v.y;
    }
    dst[o] = c;
  }
}

template <typename T>
__global__ void k_root_cot(typename Vec<T>::type *__restrict__ dst, int64_t dep, int64_t batch, int64_t nvalid,
                           int64_t relems, const int64_t *__restrict__ relem_off, const double *__restrict__ amp,
                           int64_t row0) {
  using V = typename Vec<T>::type;
  if (dep) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < batch * relems;
         t += (int64_t)gridDim.x * blockDim.x) {
      int64_t beta = t % batch, e = t / batch;
      V v{0, 0};
      if (beta < nvalid) {
        const double *p = amp + 2 * ((row0 + beta) * relems + e);
        v = V{(T)p[0], (T)p[1]};
      }
      dst[relem_off[e] + beta * dep] = v;
    }
  } else {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < relems;
         e += (int64_t)gridDim.x * blockDim.x) {
      double re = 0, im = 0;
      for (int64_t beta = 0; beta < nvalid; ++beta) {
        const double *p = amp + 2 * ((row0 + beta) * relems + e);
        re += p[0];
        im += p[1];
      }
      V w = dst[relem_off[e]];
      w.x += (T)re;
      w.y += (T)im;
      dst[relem_off[e]] = w;
    }
  }
}

// Output tensors <b_q| for a batch: O_i[beta + batch*x], beta fastest.
template <typename T>
__global__ void k_fill_outputs(typename Vec<T>::type *arena, const int64_t *__restrict__ out_off, int n_out,
                               const int *__restrict__ out_q, const char *__restrict__ bits, int64_t nvalid,
                               int64_t batch, int64_t num_chars) {
  using V = typename Vec<T>::type;
  const T s2 = (T)0.70710678118654752440;
  int64_t total = (int64_t)n_out * batch;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(t / batch);
    int64_t beta = t % batch;
    char ch = beta < nvalid ? bits[beta * num_chars + out_q[i] - 1] : '0';
    T v0 = 1, v1 = 0;
    if (ch == '1') v0 = 0, v1 = 1;
    else if (ch == '+') v0 = s2, v1 = s2;
    else if (ch == '-') v0 = s2, v1 = -s2;
    V *o = arena + out_off[i];
    o[beta] = V{v0, 0};
    o[beta + batch] = V{v1, 0};
  }
}

// acc[(row0+beta)*relems + e] holds (sum.re, sum.im, comp.re, comp.im); Neumaier.
__device__ __forceinline__ void neumaier(double &s, double &c, double x) {
  double t = s + x;
  c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
  s = t;
}

template <typename T>
__global__ void k_accumulate(const typename Vec<T>::type *__restrict__ root, int64_t root_bstride, int64_t batch,
                             int64_t nvalid, int64_t relems, const int64_t *__restrict__ relem_off,
                             double *__restrict__ acc, int64_t row0) {
  int64_t total = nvalid * relems;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t beta = t % nvalid;
    int64_t e = t / nvalid;
    auto v = root[relem_off[e] + beta * root_bstride];
    double *a = acc + ((row0 + beta) * relems + e) * 4;
    double s0 = a[0], s1 = a[1], c0 = a[2], c1 = a[3];
    neumaier(s0, c0, (double)v.x);
    neumaier(s1, c1, (double)v.y);
    a[0] = s0;
    a[1] = s1;
    a[2] = c0;
    a[3] = c1;
  }
}

__global__ void k_finalize(const double *__restrict__ acc, double *__restrict__ out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    out[2 * i] = acc[4 * i] + acc[4 * i + 2];
    out[2 * i + 1] = acc[4 * i + 1] + acc[4 * i + 3];
  }
}

__global__ void k_zero(uint4 *p, int64_t n16) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = uint4{0, 0, 0, 0};
}

// Batched zero fill: workgroup i clears range i of `base` (offset and length in
// 16-byte units, pairs in `ranges`).
__global__ void k_zero_batch(uint4 *base, const int64_t *__restrict__ ranges) {
  const int64_t off = ranges[2 * blockIdx.x], n = ranges[2 * blockIdx.x + 1];
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) base[off + i] = uint4{0, 0, 0, 0};
}

// float -> bf16 bits, rounded to nearest with ties to even (finite inputs).
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__global__ void k_bf16_split(const float2 *__restrict__ src, int64_t n, uint16_t *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = src[i];
    const uint16_t rh = bf16_bits(v.x), ih = bf16_bits(v.y);
    out[i] = rh;
    out[n + i] = bf16_bits(v.x - bf16_value(rh));
    out[2 * n + i] = ih;
    out[3 * n + i] = bf16_bits(v.y - bf16_value(ih));
  }
}

struct GatherArgs {
  int nb;
  int64_t w[kMaxBits];
};

template <typename T>
__global__ void k_gather(const typename Vec<T>::type *__restrict__ src, typename Vec<T>::type *__restrict__ dst,
                         int64_t n, GatherArgs g) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t o = 0;
    for (int b = 0; b < g.nb; ++b)
      if ((i >> b) & 1) o += g.w[b];
    dst[i] = src[o];
  }
}

// Bit-permutation copy through LDS: tile bits are the destination's lowest bits and
// the bits that are the source's lowest (source weights 1, 2, 4, ...).  A tile is
// read with lanes along the source's low bits and written with lanes along the
// destination's low bits; the other bits number the tiles.
struct GatherTileArgs {
  int nt, nr;                  // tile bits, other bits
  int tdst[12];                // destination bit of tile bit q (ascending)
  int lbit[12];                // load-order bit j -> tile bit
  int64_t tw[12];              // source weight of tile bit q
  int rdst[kMaxBits];          // destination bit of other bit r
  int64_t rw[kMaxBits];        // source weight of other bit r
};

template <typename T, int NT>
__global__ __launch_bounds__(256) void k_gather_tiled(const typename Vec<T>::type *__restrict__ src,
                                                      typename Vec<T>::type *__restrict__ dst, int64_t ntiles,
                                                      GatherTileArgs g) {
  using V2 = typename Vec<T>::type;
  constexpr int E = (1 << NT) / 256;   // elements per lane per tile
  __shared__ V2 tile[(1 << NT) + (1 << NT) / 32];
  // tile element k = threadIdx.x + 256 e: the offsets of the lane's bits (low 8) and of
  // e's bits (high NT - 8), the same for every tile
  const int tid = threadIdx.x;
  int64_t so_lo = 0, d_lo = 0;
  int loc_lo = 0;
  for (int j = 0; j < 8; ++j)
    if ((tid >> j) & 1) {
      const int q = g.lbit[j];
      loc_lo |= 1 << q;
      so_lo += g.tw[q];
    }
  for (int q = 0; q < 8; ++q)
    if ((tid >> q) & 1) d_lo += int64_t(1) << g.tdst[q];
  int64_t so_hi[E], d_hi[E];
  int loc_hi[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    so_hi[e] = 0, d_hi[e] = 0, loc_hi[e] = 0;
#pragma unroll
    for (int j = 8; j < NT; ++j)
      if ((e >> (j - 8)) & 1) {
        const int q = g.lbit[j];
        loc_hi[e] |= 1 << q;
        so_hi[e] += g.tw[q];
        d_hi[e] += int64_t(1) << g.tdst[j];
      }
  }
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int64_t sb = 0, db = 0;
    for (int r = 0; r < g.nr; ++r)
      if ((t >> r) & 1) sb += g.rw[r], db += int64_t(1) << g.rdst[r];
    V2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = src[sb + so_lo + so_hi[e]];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int loc = loc_lo | loc_hi[e];
      tile[loc + (loc >> 5)] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = tid + 256 * e;
      dst[db + d_lo + d_hi[e]] = tile[k + (k >> 5)];
    }
    __syncthreads();
  }
}

__global__ void k_planar_merge(const float *__restrict__ re, const float *__restrict__ im, int64_t n,
                               float2 *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = float2{re[i], im[i]};
}

int grid_for(int64_t work, int per_block) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > (int64_t(1) << 20)) g = int64_t(1) << 20;
  return (int)g;
}

void check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw Error(QX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}


template <int TM, int TN, int TK>
void gemm_launch(int prec, const void *A, const void *B, void *C, const GemmArgs &a, hipStream_t s) {
  int lm = 0, ln = 0;
  while ((1 << lm) < TM) ++lm;
  while ((1 << ln) < TN) ++ln;
  int64_t tiles = (a.nbm > lm ? (int64_t(1) << (a.nbm - lm)) : 1) * (a.nbn > ln ? (int64_t(1) << (a.nbn - ln)) : 1) *
                  (int64_t(1) << a.nbh) * a.ksplit;
  // persistent grid: a few workgroups per CU, each walking a contiguous tile range
  const int64_t cap = 256 * 4;
  int g = (int)(tiles < cap ? tiles : cap);
  if (prec == QX_C128)
    hipLaunchKernelGGL((k_contract_gemm<double, TM, TN, TK>), dim3(g), dim3(256), 0, s, (const double2 *)A,
                       (const double2 *)B, (double2 *)C, a);
  else
    hipLaunchKernelGGL((k_contract_gemm<float, TM, TN, TK>), dim3(g), dim3(256), 0, s, (const float2 *)A,
                       (const float2 *)B, (float2 *)C, a);
}


template <int NK, int NC>
void rows_launch(int prec, const void *A, const void *B, void *C, const RowArgs &a, hipStream_t s) {
  int g = grid_for(int64_t(1) << a.nbr, 256);
  int hl = 0;
  for (int b = 0; b < 8; ++b) hl += (a.hlo >> b) & 1;
  if (a.hlo == 0) {
    if (prec == QX_C128)
      hipLaunchKernelGGL((k_contract_rows<double, NK, NC, true>), dim3(g), dim3(256), 0, s, (const double2 *)A,
                         (const double2 *)B, (double2 *)C, a);
    else
      hipLaunchKernelGGL((k_contract_rows<float, NK, NC, true>), dim3(g), dim3(256), 0, s, (const float2 *)A,
                         (const float2 *)B, (float2 *)C, a);
    return;
  }
  size_t smem = (size_t)(NK * a.nn) << hl;
  if (prec == QX_C128)
    hipLaunchKernelGGL((k_contract_rows<double, NK, NC, false>), dim3(g), dim3(256), smem * sizeof(double2), s,
                       (const double2 *)A, (const double2 *)B, (double2 *)C, a);
  else
    hipLaunchKernelGGL((k_contract_rows<float, NK, NC, false>), dim3(g), dim3(256), smem * sizeof(float2), s,
                       (const float2 *)A, (const float2 *)B, (float2 *)C, a);
}

template <int NK>
void rows_launch_k(int prec, const void *A, const void *B, void *C, const RowArgs &a, hipStream_t s) {
  switch (a.nn) {
    case 1: rows_launch<NK, 1>(prec, A, B, C, a, s); break;
    case 2: rows_launch<NK, 2>(prec, A, B, C, a, s); break;
    case 4: rows_launch<NK, 4>(prec, A, B, C, a, s); break;
    default: rows_launch<NK, 8>(prec, A, B, C, a, s); break;
  }
}

}  // namespace

// Runs of outer products (OuterArgs): lane per output element, the product of the
// factors in the run's order, each factor's offset the sum of its strides over the
// element's set bits.  The terms first go into a dense LDS table (factor, output bit)
// -> stride (a factor has at most one term per bit), read without loop-carried waits.
// A workgroup covers 256 consecutive elements per trip: the lowest 8 bits are the
// lane's own (their part of every factor's offset is summed once per lane, kept in
// LDS), the bits above are the same for all lanes (summed once per trip, 16 lanes per
// factor).
template <typename T>
__global__ __launch_bounds__(256) void k_outer_run(typename Vec<T>::type *__restrict__ C, int64_t n, OuterArgs a) {
  using V = typename Vec<T>::type;
  __shared__ int64_t tw[kOuterMaxFactors * kOuterBits];
  __shared__ int64_t shi[kOuterMaxFactors];
  __shared__ int64_t slo[kOuterMaxFactors * 256];
  const int t = threadIdx.x;
  const int nf = a.nf, nt = a.first[nf];
  for (int i = t; i < kOuterMaxFactors * kOuterBits; i += 256) tw[i] = 0;
  __syncthreads();
  for (int k = t; k < nt; k += 256) {
    int f = 0;
#pragma unroll
    for (int g = 1; g < kOuterMaxFactors; ++g)
      if (g < nf && k >= a.first[g]) f = g;
    tw[f * kOuterBits + a.bit[k]] = a.w[k];
  }
  __syncthreads();
  // the lane's own low-bit offsets, one per factor (each lane reads only its own)
  for (int f = 0; f < nf; ++f) {
    int64_t o = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if ((t >> b) & 1) o += tw[f * kOuterBits + b];
    slo[f * 256 + t] = o;
  }
  for (int64_t base = (int64_t)blockIdx.x * 256; base < n; base += (int64_t)gridDim.x * 256) {
    {
      // 16 lanes per factor (f < kOuterMaxFactors = 256 / 16), each summing every
      // 16th bit, then across the 16
      static_assert(kOuterMaxFactors * 16 == 256, "16 lanes per factor");
      const int f = t >> 4, j = t & 15;
      long long h = 0;
      if (f < nf) {
#pragma unroll
        for (int b = 8 + j; b < kOuterBits; b += 16) h += ((base >> b) & 1) ? tw[f * kOuterBits + b] : 0;
      }
#pragma unroll
      for (int m = 8; m >= 1; m >>= 1) h += __shfl_xor(h, m, 16);
      if (j == 0 && f < nf) shi[f] = h;
    }
    __syncthreads();
    const int64_t e = base + t;
    if (e < n) {
      // every factor's value loaded before the products (one memory latency per trip)
      V xs[kOuterMaxFactors];
#pragma unroll
      for (int f = 0; f < kOuterMaxFactors; ++f)
        if (f < nf) xs[f] = reinterpret_cast<const V *>(a.F[f])[slo[f * 256 + t] + shi[f]];
      V v = xs[0];
#pragma unroll
      for (int f = 1; f < kOuterMaxFactors; ++f)
        if (f < nf) {
          V r{0, 0};
          if (a.prev_a[f]) cmac<T>(r, v, xs[f], 0, 0);
          else cmac<T>(r, xs[f], v, 0, 0);
          v = r;
        }
      C[e] = v;
    }
    __syncthreads();
  }
}

void launch_outer(int prec, void *C, int64_t c_elems, const OuterArgs &args, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const int g = grid_for(c_elems, 256) < 1024 ? grid_for(c_elems, 256) : 1024;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_outer_run<double>, dim3(g), dim3(256), 0, s, (double2 *)C, c_elems, args);
  else
    hipLaunchKernelGGL(k_outer_run<float>, dim3(g), dim3(256), 0, s, (float2 *)C, c_elems, args);
  check_launch("outer_run");
}

void launch_contract_rows(int prec, const void *A, const void *B, void *C, const RowArgs &args, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  switch (args.nk) {
    case 1: rows_launch_k<1>(prec, A, B, C, args, s); break;
    case 2: rows_launch_k<2>(prec, A, B, C, args, s); break;
    case 4: rows_launch_k<4>(prec, A, B, C, args, s); break;
    case 8: rows_launch_k<8>(prec, A, B, C, args, s); break;
    case 16: rows_launch_k<16>(prec, A, B, C, args, s); break;
    default: rows_launch_k<32>(prec, A, B, C, args, s); break;
  }
  check_launch("contract_rows");
}

namespace {

}  // namespace

void launch_contract_gemm(int prec, const void *A, const void *B, void *C, void *ws, const GemmArgs &args,
                          void *stream) {
  hipStream_t s = (hipStream_t)stream;
  void *dst = args.ksplit > 1 ? ws : C;
  if (args.tile >= 3) {   // matrix cores, 64x64 tiles, K chunks of 16 (tile 3) or 32 (tile 4)
    int64_t tiles = (args.nbm > 6 ? (int64_t(1) << (args.nbm - 6)) : 1) *
                    (args.nbn > 6 ? (int64_t(1) << (args.nbn - 6)) : 1) * (int64_t(1) << args.nbh) * args.ksplit;
    int g = (int)(tiles < 1024 ? tiles : 1024);
#define QX_MFMA_GEMM(TT, V2, GA, TKK)                                                                   \
  hipLaunchKernelGGL((k_contract_gemm_mfma<TT, GA, TKK>), dim3(g), dim3(256), 0, s, (const V2 *)A, \
                     (const V2 *)B, (V2 *)dst, args)
    const bool k32 = args.tile == 4;
    if (prec == QX_C128) {
      if (k32) QX_MFMA_GEMM(double, double2, false, 32); else QX_MFMA_GEMM(double, double2, false, 16);
    } else if (prec == QX_C64_3M) {
      if (k32) QX_MFMA_GEMM(float, float2, true, 32); else QX_MFMA_GEMM(float, float2, true, 16);
    } else {
      if (k32) QX_MFMA_GEMM(float, float2, false, 32); else QX_MFMA_GEMM(float, float2, false, 16);
    }
#undef QX_MFMA_GEMM
  } else switch (args.tile) {
    case 0: gemm_launch<128, 16, 16>(prec, A, B, dst, args, s); break;
    case 1: gemm_launch<128, 32, 16>(prec, A, B, dst, args, s); break;
    default: gemm_launch<64, 64, 16>(prec, A, B, dst, args, s); break;
  }
  check_launch("contract_gemm");
  if (args.ksplit > 1) {
    int g = grid_for(args.c_elems, 256);
    if (prec == QX_C128)
      hipLaunchKernelGGL(k_sum_splits<double>, dim3(g), dim3(256), 0, s, (double2 *)C, (const double2 *)ws,
                         args.c_elems, args.ksplit);
    else
      hipLaunchKernelGGL(k_sum_splits<float>, dim3(g), dim3(256), 0, s, (float2 *)C, (const float2 *)ws,
                         args.c_elems, args.ksplit);
    check_launch("sum_splits");
  }
}

// ---------------------------------------------------------------------------
// Plain complex GEMM on matrix cores: C (r x s, ldc) = op(X) (r x k) . op(Y) (k x s)
// for column-major X, Y, C with batch strides -- the contractions whose operands the
// lowering laid out as plain matrices (op 1 = stored transposed).  A workgroup of four
// waves (2 x 2) owns a TM x TN tile of C for one batch element and one K range:
// K stages of TK values are loaded with lanes along the operand's contiguous index
// (16 bytes per lane for complex128), staged in LDS as [k][m] / [k][n] (one element of
// padding per row), and the next stage is fetched into registers while the matrix
// cores work on the current one.  Each wave holds WM x WN of C as 16 x 16 blocks of
// 16x16x4 products: 4M (Re += Xr.Yr - Xi.Yi, Im += Xr.Yi + Xi.Yr, four products per
// block) or 3M (RR = Xr.Yr, II = Xi.Yi, SS = (Xr+Xi).(Yr+Yi); Re = RR - II,
// Im = SS - RR - II, three products).  Accumulation in the operand precision.  With
// ksplit > 1 each K range writes a dense r x s partial per batch element to `ws`
// (k_plain_sum adds them into C).  Work units are ordered with the C column tile
// fastest and dealt to the eight XCDs in contiguous runs, so the units of one XCD share
// their X panels in its L2.
template <typename T, bool GAUSS> struct PlainTile;
template <> struct PlainTile<double, false> { static constexpr int WM = 32, WN = 32, TK = 16; };
template <> struct PlainTile<double, true> { static constexpr int WM = 32, WN = 32, TK = 16; };
template <> struct PlainTile<float, false> { static constexpr int WM = 64, WN = 64, TK = 16; };
template <> struct PlainTile<float, true> { static constexpr int WM = 64, WN = 32, TK = 16; };

template <typename T, int OPX, int OPY, bool GAUSS>
__global__ __launch_bounds__(256) void k_plain_gemm(const typename Vec<T>::type *__restrict__ X,
                                                    const typename Vec<T>::type *__restrict__ Y,
                                                    typename Vec<T>::type *__restrict__ C,
                                                    typename Vec<T>::type *__restrict__ ws, BlasArgs b) {
  using V = typename Vec<T>::type;
  using M = Mfma16<T>;
  using v4 = typename M::v4;
  constexpr int WM = PlainTile<T, GAUSS>::WM, WN = PlainTile<T, GAUSS>::WN, TK = PlainTile<T, GAUSS>::TK;
  constexpr int TM = 2 * WM, TN = 2 * WN, PA = TM + 1, PB = TN + 1;
  constexpr int BI = WM / 16, BJ = WN / 16;
  constexpr int EA = TM * TK / 256, EB = TN * TK / 256;
  __shared__ V sA[TK * PA];
  __shared__ V sB[TK * PB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lq = lane >> 4;
  const int64_t tr = (b.r + TM - 1) / TM, ts = (b.s + TN - 1) / TN;
  const int64_t units = b.batch * tr * ts * b.ksplit;
  int64_t u = blockIdx.x;
  {
    const int64_t nb = gridDim.x;   // one unit per workgroup; contiguous runs per XCD
    if (nb % 8 == 0) u = (u % 8) * (nb / 8) + u / 8;
  }
  if (u >= units) return;
  const int64_t st = u % ts;
  int64_t rest = u / ts;
  const int64_t ks = rest % b.ksplit;
  rest /= b.ksplit;
  const int64_t rt = rest % tr;
  const int64_t bat = rest / tr;
  const int64_t m0 = rt * TM, n0 = st * TN;
  const int64_t stages = b.k / TK / b.ksplit;
  const int64_t k0 = ks * stages * TK;
  // this thread's elements of the X and Y stages: offsets within a stage and masks
  int64_t offa[EA], offb[EB];
  int slota[EA], slotb[EB];
  bool oka[EA], okb[EB];
#pragma unroll
  for (int j = 0; j < EA; ++j) {
    const int e = tid + 256 * j;
    const int m = OPX ? e / TK : e % TM, kk = OPX ? e % TK : e / TM;
    oka[j] = m0 + m < b.r;
    offa[j] = OPX ? kk + (m0 + m) * b.ldx : (m0 + m) + kk * b.ldx;
    slota[j] = kk * PA + m;
  }
#pragma unroll
  for (int j = 0; j < EB; ++j) {
    const int e = tid + 256 * j;
    const int n = OPY ? e % TN : e / TK, kk = OPY ? e / TN : e % TK;
    okb[j] = n0 + n < b.s;
    offb[j] = OPY ? (n0 + n) + kk * b.ldy : kk + (n0 + n) * b.ldy;
    slotb[j] = kk * PB + n;
  }
  const V *Xb = X + bat * b.sx + (OPX ? k0 : k0 * b.ldx);
  const V *Yb = Y + bat * b.sy + (OPY ? k0 * b.ldy : k0);
  const int64_t xstep = OPX ? TK : TK * b.ldx, ystep = OPY ? TK * b.ldy : TK;
  V ra[EA], rb[EB];
  auto fetch = [&](int64_t sg) {
    const V *xa = Xb + sg * xstep;
    const V *yb = Yb + sg * ystep;
#pragma unroll
    for (int j = 0; j < EA; ++j) ra[j] = oka[j] ? xa[offa[j]] : V{0, 0};
#pragma unroll
    for (int j = 0; j < EB; ++j) rb[j] = okb[j] ? yb[offb[j]] : V{0, 0};
  };
  v4 acc0[BI][BJ], acc1[BI][BJ], acc2[BI][BJ];   // 4M: Re, Im; 3M: RR, II, SS
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) acc0[i][j] = v4{0, 0, 0, 0}, acc1[i][j] = v4{0, 0, 0, 0}, acc2[i][j] = v4{0, 0, 0, 0};
  fetch(0);
  for (int64_t sg = 0; sg < stages; ++sg) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < EA; ++j) sA[slota[j]] = ra[j];
#pragma unroll
    for (int j = 0; j < EB; ++j) sB[slotb[j]] = rb[j];
    __syncthreads();
    if (sg + 1 < stages) fetch(sg + 1);
#pragma unroll
    for (int kc = 0; kc < TK; kc += 4) {
      V fa[BI], fb[BJ];
#pragma unroll
      for (int i = 0; i < BI; ++i) fa[i] = sA[(kc + lq) * PA + WM * wm + 16 * i + li];
#pragma unroll
      for (int j = 0; j < BJ; ++j) fb[j] = sB[(kc + lq) * PB + WN * wn + 16 * j + li];
      if (GAUSS) {
        T fs[BI], gs[BJ];
#pragma unroll
        for (int i = 0; i < BI; ++i) fs[i] = fa[i].x + fa[i].y;
#pragma unroll
        for (int j = 0; j < BJ; ++j) gs[j] = fb[j].x + fb[j].y;
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < BJ; ++j) {
            acc0[i][j] = M::mma(fa[i].x, fb[j].x, acc0[i][j]);
            acc1[i][j] = M::mma(fa[i].y, fb[j].y, acc1[i][j]);
            acc2[i][j] = M::mma(fs[i], gs[j], acc2[i][j]);
          }
      } else {
#pragma unroll
        for (int i = 0; i < BI; ++i) {
          const T nai = -fa[i].y;
#pragma unroll
          for (int j = 0; j < BJ; ++j) {
            acc0[i][j] = M::mma(fa[i].x, fb[j].x, acc0[i][j]);
            acc1[i][j] = M::mma(fa[i].x, fb[j].y, acc1[i][j]);
          }
#pragma unroll
          for (int j = 0; j < BJ; ++j) {
            acc0[i][j] = M::mma(nai, fb[j].y, acc0[i][j]);
            acc1[i][j] = M::mma(fa[i].y, fb[j].x, acc1[i][j]);
          }
        }
      }
    }
  }
  // results: to C, or as this K range's dense partial
  const bool part = b.ksplit > 1;
  V *dst = part ? ws + (ks * b.batch + bat) * b.r * b.s : C + bat * b.sc;
  const int64_t ld = part ? b.r : b.ldc;
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int64_t n = n0 + WN * wn + 16 * j + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t m = m0 + WM * wm + 16 * i + M::row(lq, e);
        V o;
        if (GAUSS) {
          o.x = acc0[i][j][e] - acc1[i][j][e];
          o.y = acc2[i][j][e] - acc0[i][j][e] - acc1[i][j][e];
        } else {
          o.x = acc0[i][j][e];
          o.y = acc1[i][j][e];
        }
        if (m < b.r && n < b.s) dst[m + n * ld] = o;
      }
    }
}

// C (r x s per batch element, ldc, batch stride sc) = sum of the ksplit dense partials.
template <typename T>
__global__ void k_plain_sum(typename Vec<T>::type *__restrict__ C, const typename Vec<T>::type *__restrict__ ws,
                            BlasArgs b) {
  using V = typename Vec<T>::type;
  const int64_t per = b.r * b.s, n = per * b.batch;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    V acc = ws[i];
    for (int s = 1; s < b.ksplit; ++s) {
      const V v = ws[(int64_t)s * n + i];
      acc.x += v.x;
      acc.y += v.y;
    }
    const int64_t bat = i / per, e = i % per;
    C[bat * b.sc + (e % b.r) + (e / b.r) * b.ldc] = acc;
  }
}

// Split bf16 product of a plain complex64 GEMM (QX_C64_BF16X3): every real product
// P.Q is Ph.Qh + Ph.Ql + Pl.Qh with h = bf16(x) and l = bf16(x - h) (round to nearest
// even), on 16x16x32 bf16 matrix cores accumulating in FP32 -- twelve bf16 products per
// complex product (Re = Xr.Yr - Xi.Yi with the Xi parts negated, Im = Xr.Yi + Xi.Yr).
// The split happens while a K stage is staged in LDS: each lane loads two complex64
// values adjacent in k, rounds them, and stores the packed pairs into four planes
// (re hi, re lo, im hi, im lo) laid out [m][k] / [n][k], a row's 16-byte chunks
// XOR-permuted by the row, so a lane's fragment (8 consecutive k values of one row) is
// one 16-byte LDS read.
// Tiles of 64 x 64 of C per workgroup (2 x 2 waves of 2 x 2 blocks), K stages of 32,
// the next stage fetched into registers during the products; K split and unit order as
// in k_plain_gemm.  A lane's k slot is the same in both operands, so the products pair
// equal k whatever order the matrix core sums them in; the result layout is that of
// the 16x16x4 FP32 product (Mfma16<float>).
template <int OPX, int OPY>
__global__ __launch_bounds__(256) void k_plain_gemm_bf16x3(const float2 *__restrict__ X, const float2 *__restrict__ Y,
                                                           float2 *__restrict__ C, float2 *__restrict__ ws,
                                                           BlasArgs b) {
  typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
  typedef float f4 __attribute__((ext_vector_type(4)));
  // a row of a plane: 16 dwords (32 bf16 values of k) in four 16-byte chunks, chunk c
  // of row r stored at chunk c ^ ((r >> 2) & 3): the 16 rows a fragment read spans
  // (one chunk each) and the 4 rows x 16 dwords a wave stores with lanes along k both
  // fall on distinct banks
  constexpr int TM = 64, TN = 64, TK = 32, RW = TK / 2;
  auto slot = [](int row, int dw) { return dw ^ (((row >> 2) & 3) << 2); };
  constexpr int BI = 2, BJ = 2, NP = 4;                           // planes: rh, rl, ih, il
  __shared__ uint32_t sA[NP][TM][RW];
  __shared__ uint32_t sB[NP][TN][RW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, lq = lane >> 4;
  const int64_t tr = (b.r + TM - 1) / TM, ts = (b.s + TN - 1) / TN;
  const int64_t units = b.batch * tr * ts * b.ksplit;
  int64_t u = blockIdx.x;
  {
    const int64_t nb = gridDim.x;
    if (nb % 8 == 0) u = (u % 8) * (nb / 8) + u / 8;
  }
  if (u >= units) return;
  const int64_t st = u % ts;
  int64_t rest = u / ts;
  const int64_t ks = rest % b.ksplit;
  rest /= b.ksplit;
  const int64_t rt = rest % tr;
  const int64_t bat = rest / tr;
  const int64_t m0 = rt * TM, n0 = st * TN;
  const int64_t stages = b.k / TK / b.ksplit;
  const int64_t k0 = ks * stages * TK;
  // four k pairs per thread and operand: (row, pair) of each, and the two elements'
  // offsets inside a stage
  int ra_row[4], ra_pair[4], rb_row[4], rb_pair[4];
  int64_t oa0[4], oa1[4], ob0[4], ob1[4];
  bool oka[4], okb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int m, p;
    if (OPX) p = tid % 16, m = tid / 16 + 16 * j;   // k contiguous in X: lanes along k
    else m = tid % 64, p = tid / 64 + 4 * j;        // m contiguous in X: lanes along m
    ra_row[j] = m, ra_pair[j] = p;
    oka[j] = m0 + m < b.r;
    const int64_t kk = 2 * p;
    oa0[j] = OPX ? kk + (m0 + m) * b.ldx : (m0 + m) + kk * b.ldx;
    oa1[j] = OPX ? oa0[j] + 1 : oa0[j] + b.ldx;
    int n, q;
    if (OPY) n = tid % 64, q = tid / 64 + 4 * j;    // Y stored n x k: lanes along n
    else q = tid % 16, n = tid / 16 + 16 * j;       // Y stored k x n: lanes along k
    rb_row[j] = n, rb_pair[j] = q;
    okb[j] = n0 + n < b.s;
    const int64_t kq = 2 * q;
    ob0[j] = OPY ? (n0 + n) + kq * b.ldy : kq + (n0 + n) * b.ldy;
    ob1[j] = OPY ? ob0[j] + b.ldy : ob0[j] + 1;
  }
  const float2 *Xb = X + bat * b.sx + (OPX ? k0 : k0 * b.ldx);
  const float2 *Yb = Y + bat * b.sy + (OPY ? k0 * b.ldy : k0);
  const int64_t xstep = OPX ? TK : TK * b.ldx, ystep = OPY ? TK * b.ldy : TK;
  float2 xa0[4], xa1[4], yb0[4], yb1[4];
  auto fetch = [&](int64_t sg) {
    const float2 *xa = Xb + sg * xstep;
    const float2 *yb = Yb + sg * ystep;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      xa0[j] = oka[j] ? xa[oa0[j]] : float2{0, 0};
      xa1[j] = oka[j] ? xa[oa1[j]] : float2{0, 0};
      yb0[j] = okb[j] ? yb[ob0[j]] : float2{0, 0};
      yb1[j] = okb[j] ? yb[ob1[j]] : float2{0, 0};
    }
  };
  // (hi, lo) planes of a pair of values adjacent in k, packed two per dword
  auto split_store = [&](uint32_t (*pl)[TM][RW], int row, int pair, float2 v0, float2 v1) {
    const uint16_t rh0 = bf16_bits(v0.x), rh1 = bf16_bits(v1.x), ih0 = bf16_bits(v0.y), ih1 = bf16_bits(v1.y);
    const int d = slot(row, pair);
    pl[0][row][d] = (uint32_t)rh0 | ((uint32_t)rh1 << 16);
    pl[1][row][d] = (uint32_t)bf16_bits(v0.x - bf16_value(rh0)) | ((uint32_t)bf16_bits(v1.x - bf16_value(rh1)) << 16);
    pl[2][row][d] = (uint32_t)ih0 | ((uint32_t)ih1 << 16);
    pl[3][row][d] = (uint32_t)bf16_bits(v0.y - bf16_value(ih0)) | ((uint32_t)bf16_bits(v1.y - bf16_value(ih1)) << 16);
  };
  f4 accr[BI][BJ], acci[BI][BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) accr[i][j] = f4{0, 0, 0, 0}, acci[i][j] = f4{0, 0, 0, 0};
  auto frag = [&](const uint32_t (*pl)[TM][RW], int plane, int row) {
    const uint4 v = *reinterpret_cast<const uint4 *>(&pl[plane][row][slot(row, 4 * lq)]);
    return v;
  };
  auto as_bf8 = [](uint4 v) { return __builtin_bit_cast(bf8, v); };
  auto neg = [](uint4 v) { return uint4{v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u, v.w ^ 0x80008000u}; };
  fetch(0);
  for (int64_t sg = 0; sg < stages; ++sg) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      split_store(sA, ra_row[j], ra_pair[j], xa0[j], xa1[j]);
      split_store(sB, rb_row[j], rb_pair[j], yb0[j], yb1[j]);
    }
    __syncthreads();
    if (sg + 1 < stages) fetch(sg + 1);
    bf8 yrh[BJ], yrl[BJ], yih[BJ], yil[BJ];
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int row = 32 * wn + 16 * j + li;
      yrh[j] = as_bf8(frag(sB, 0, row));
      yrl[j] = as_bf8(frag(sB, 1, row));
      yih[j] = as_bf8(frag(sB, 2, row));
      yil[j] = as_bf8(frag(sB, 3, row));
    }
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      const int row = 32 * wm + 16 * i + li;
      const uint4 ih = frag(sA, 2, row), il = frag(sA, 3, row);
      const bf8 xrh = as_bf8(frag(sA, 0, row)), xrl = as_bf8(frag(sA, 1, row));
      const bf8 xih = as_bf8(ih), xil = as_bf8(il), nxih = as_bf8(neg(ih)), nxil = as_bf8(neg(il));
#pragma unroll
      for (int j = 0; j < BJ; ++j) {
        f4 r = accr[i][j], m = acci[i][j];
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrh, yrh[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrh, yih[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrh, yrl[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrh, yil[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrl, yrh[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xrl, yih[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nxih, yih[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xih, yrh[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nxih, yil[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xih, yrl[j], m, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nxil, yih[j], r, 0, 0, 0);
        m = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xil, yrh[j], m, 0, 0, 0);
        accr[i][j] = r, acci[i][j] = m;
      }
    }
  }
  const bool part = b.ksplit > 1;
  float2 *dst = part ? ws + (ks * b.batch + bat) * b.r * b.s : C + bat * b.sc;
  const int64_t ld = part ? b.r : b.ldc;
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const int64_t n = n0 + 32 * wn + 16 * j + li;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t m = m0 + 32 * wm + 16 * i + Mfma16<float>::row(lq, e);
        if (m < b.r && n < b.s) dst[m + n * ld] = float2{accr[i][j][e], acci[i][j][e]};
      }
    }
}

// Skinny contraction: C[h, m, n] = sum_k A[h, m, k] B[h, k, n] for at most 32 (m, n)
// pairs (N <= 2) and a long K.  Workgroup (h, chunk): each lane walks K indices of the
// chunk (low K bits, A's smallest strides), loads its A value for every m and B value
// for every n once, and keeps the MB x NB sums; the workgroup reduces them (wave
// shuffles, then LDS) into the chunk's partial copy of C.  Every pass reads A and B
// once (the one-wave-per-element reduce reads A N times and B M times).
template <typename T, int MB, int NB>
__global__ __launch_bounds__(256) void k_contract_skinny(const typename Vec<T>::type *__restrict__ A,
                                                         const typename Vec<T>::type *__restrict__ B,
                                                         typename Vec<T>::type *__restrict__ P, int64_t c_elems,
                                                         BitArgs a) {
  using V = typename Vec<T>::type;
  const int chunk = (int)(blockIdx.x % a.ksplit);
  const int64_t h = blockIdx.x / a.ksplit;
  int64_t oa = 0, ob = 0, ch = 0;
  {
    int hi = 0;
    for (int i = 0; i < a.nbc; ++i)
      if (a.wa[i] && a.wb[i]) {
        if ((h >> hi) & 1) oa += a.wa[i], ob += a.wb[i], ch += int64_t(1) << i;
        ++hi;
      }
  }
  int64_t am[MB], cm[MB], bn[NB], cn[NB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    int64_t o = 0, c = 0;
    int j = 0;
    for (int i = 0; i < a.nbc; ++i)
      if (a.wa[i] && !a.wb[i]) {
        if ((m >> j) & 1) o += a.wa[i], c += int64_t(1) << i;
        ++j;
      }
    am[m] = o, cm[m] = c;
  }
#pragma unroll
  for (int n = 0; n < NB; ++n) {
    int64_t o = 0, c = 0;
    int j = 0;
    for (int i = 0; i < a.nbc; ++i)
      if (!a.wa[i] && a.wb[i]) {
        if ((n >> j) & 1) o += a.wb[i], c += int64_t(1) << i;
        ++j;
      }
    bn[n] = o, cn[n] = c;
  }
  for (int b = a.kcb; b < a.nbk; ++b)
    if ((chunk >> (b - a.kcb)) & 1) oa += a.ka[b], ob += a.kb[b];
  V acc[MB][NB];
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < NB; ++n) acc[m][n] = V{0, 0};
  const int64_t kn = int64_t(1) << a.kcb;
  for (int64_t k = threadIdx.x; k < kn; k += 256) {
    int64_t xa = oa, xb = ob;
    for (int b = 0; b < a.kcb; ++b)
      if ((k >> b) & 1) xa += a.ka[b], xb += a.kb[b];
    V av[MB], bv[NB];
#pragma unroll
    for (int m = 0; m < MB; ++m) av[m] = A[xa + am[m]];
#pragma unroll
    for (int n = 0; n < NB; ++n) bv[n] = B[xb + bn[n]];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
      for (int n = 0; n < NB; ++n) cmac<T>(acc[m][n], av[m], bv[n], a.conj_a, a.conj_b);
  }
  __shared__ V red[4][MB * NB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < NB; ++n) {
      V v = acc[m][n];
      for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
      }
      if (lane == 0) red[wave][m * NB + n] = v;
    }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MB; ++m)
#pragma unroll
    for (int n = 0; n < NB; ++n)
      if (threadIdx.x == m * NB + n) {
        const int j = m * NB + n;
        V sum = red[0][j];
        for (int w = 1; w < 4; ++w) sum.x += red[w][j].x, sum.y += red[w][j].y;
        P[(int64_t)chunk * c_elems + ch + cm[m] + cn[n]] = sum;
      }
}

template <typename T>
void launch_skinny(const void *A, const void *B, void *P, int64_t c_elems, const BitArgs &args, hipStream_t s) {
  using V = typename Vec<T>::type;
  const int64_t groups = (c_elems >> (args.nbm + args.nbn)) * args.ksplit;
  const dim3 g((unsigned)groups), t(256);
#define QX_SKINNY(MBITS, NBITS)                                                                             \
  if (args.nbm == MBITS && args.nbn == NBITS) {                                                             \
    hipLaunchKernelGGL((k_contract_skinny<T, 1 << MBITS, 1 << NBITS>), g, t, 0, s, (const V *)A, (const V *)B, \
                       (V *)P, c_elems, args);                                                              \
    return;                                                                                                 \
  }
  QX_SKINNY(0, 0) QX_SKINNY(1, 0) QX_SKINNY(2, 0) QX_SKINNY(3, 0) QX_SKINNY(4, 0) QX_SKINNY(5, 0)
  QX_SKINNY(0, 1) QX_SKINNY(1, 1) QX_SKINNY(2, 1) QX_SKINNY(3, 1) QX_SKINNY(4, 1)
#undef QX_SKINNY
  throw Error(QX_ERR_STATE, "internal: skinny contraction shape");
}

void launch_contract_reduce(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                            const BitArgs &args, void *stream) {
  int g = grid_for(c_elems * args.ksplit, 4);
  hipStream_t s = (hipStream_t)stream;
  void *dst = args.ksplit > 1 ? ws : C;
  if (args.skinny) {
    if (prec == QX_C128)
      launch_skinny<double>(A, B, dst, c_elems, args, s);
    else
      launch_skinny<float>(A, B, dst, c_elems, args, s);
  } else if (prec == QX_C128) {
    hipLaunchKernelGGL(k_contract_reduce<double>, dim3(g), dim3(256), 0, s, (const double2 *)A, (const double2 *)B,
                       (double2 *)dst, c_elems, args);
  } else {
    hipLaunchKernelGGL(k_contract_reduce<float>, dim3(g), dim3(256), 0, s, (const float2 *)A, (const float2 *)B,
                       (float2 *)dst, c_elems, args);
  }
  check_launch("contract_reduce");
  if (args.ksplit > 1) {
    int gs = grid_for(c_elems, 256);
    if (prec == QX_C128)
      hipLaunchKernelGGL(k_sum_splits<double>, dim3(gs), dim3(256), 0, s, (double2 *)C, (const double2 *)ws, c_elems,
                         args.ksplit);
    else
      hipLaunchKernelGGL(k_sum_splits<float>, dim3(gs), dim3(256), 0, s, (float2 *)C, (const float2 *)ws, c_elems,
                         args.ksplit);
    check_launch("sum_splits");
  }
}

void launch_contract_bits(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                          const BitArgs &args, void *stream) {
  int g = grid_for(c_elems * args.ksplit, kThreads * kElems);
  if (args.ksplit > 1) g = grid_for(((c_elems + kThreads * kElems - 1) / (kThreads * kElems)) * args.ksplit, 1);
  hipStream_t s = (hipStream_t)stream;
  void *dst = args.ksplit > 1 ? ws : C;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_contract_bits<double>, dim3(g), dim3(kThreads), 0, s, (const double2 *)A,
                       (const double2 *)B, (double2 *)dst, c_elems, args);
  else
    hipLaunchKernelGGL(k_contract_bits<float>, dim3(g), dim3(kThreads), 0, s, (const float2 *)A, (const float2 *)B,
                       (float2 *)dst, c_elems, args);
  check_launch("contract_bits");
  if (args.ksplit > 1) {
    int gs = grid_for(c_elems, 256);
    if (prec == QX_C128)
      hipLaunchKernelGGL(k_sum_splits<double>, dim3(gs), dim3(256), 0, s, (double2 *)C, (const double2 *)ws, c_elems,
                         args.ksplit);
    else
      hipLaunchKernelGGL(k_sum_splits<float>, dim3(gs), dim3(256), 0, s, (float2 *)C, (const float2 *)ws, c_elems,
                         args.ksplit);
    check_launch("sum_splits");
  }
}

void launch_contract_radix(int prec, const void *A, const void *B, void *C, int64_t c_elems,
                           const RadixArgs &args, void *stream) {
  int g = grid_for(c_elems, kThreads);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_contract_radix<double>, dim3(g), dim3(kThreads), 0, s, (const double2 *)A,
                       (const double2 *)B, (double2 *)C, c_elems, args);
  else
    hipLaunchKernelGGL(k_contract_radix<float>, dim3(g), dim3(kThreads), 0, s, (const float2 *)A,
                       (const float2 *)B, (float2 *)C, c_elems, args);
  check_launch("contract_radix");
}

// ---------------------------------------------------------------------------
// Chain kernel: X_m = S_m ... S_1 X_0 in one pass.  A workgroup owns tiles of
// 2^kChainTileBits elements (the tile bits hold the legs the chain touches and the
// fastest untouched ones; the other bits index tiles).  Per tile: load from X_0
// (lanes along the fastest X_0 legs), apply every S_j in LDS, store to X_m (lanes
// along the fastest X_m legs).  Step j works on groups of 2^k elements that differ
// only in its contracted tile bits q: out[nu] = sum_kappa S[h, kappa, nu] in[kappa],
// the new legs taking the contracted legs' tile bits.  S_j is staged in LDS as
// [h][nu][kappa].
// Calls f(std::integral_constant<int, i>) for i = 0 .. N-1 (indices stay compile-time
// constants so per-lane arrays stay in registers).
template <int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (N > 0) {
    static_for<N - 1>(f);
    f(std::integral_constant<int, N - 1>{});
  }
}

constexpr int kChainTile = 1 << kChainTileBits;
// LDS beside the tile for the packed gate blocks: 32 KiB (160 KiB per workgroup)
template <typename T>
constexpr int kChainSLds = kChainSLdsBytes / (int)sizeof(typename Vec<T>::type);
constexpr int kChainThreads = 512;
constexpr int kChainPer = kChainTile / kChainThreads;   // tile elements per lane (16)

// Workgroup barrier that orders LDS only: outstanding global loads (the next
// tile's prefetch) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// End of a chain step: a workgroup barrier, or, when every wave only touches tile
// elements of its own (wave-local chains: the wave index sits on three tile bits no
// step contracts or batches), an ordering point for the wave alone (one wave's LDS
// operations complete in issue order; this keeps the compiler from moving LDS
// accesses across it).
__device__ __forceinline__ void step_end(bool wave_local) {
  if (wave_local) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  } else {
    lds_barrier();
  }
}

// Slot of packed gate-block element e in the staged copy in LDS: the low four index
// bits XOR the next four, so that the 16 rows of a 16- or 32-value block, which the
// matrix-core steps read one element per row across lanes (rows 256 or 512 bytes
// apart in complex128), land in distinct banks instead of one.  A bijection on every
// aligned run of 16 elements.
__device__ __forceinline__ int ss_slot(int e) { return e ^ ((e >> 4) & 15); }

// Step j's S, packed densely as [h][nu][kappa] at Sp + a.st[j].soff, is read with
// wave-uniform indices (the batch bits h are placed on wave-uniform group bits), so
// its values come through the scalar cache rather than LDS.
//
// LDS addresses: tile element i lives at P(i), the XOR of swz[b] over the set bits b
// of i (swz[b] = 2^b plus, for some bits, a pattern in the low bits that spreads a
// step's lanes over the bank groups; P is a bijection of the tile).  An element's
// address is therefore the XOR of the images of its bit groups.
// LDS image of tile bit b and the combination of images of disjoint bit sets: with
// SWZ the host's swz table and XOR, else the plain bit and OR (cheaper addressing for
// tiles whose accesses have no bank conflicts anyway).
template <bool SWZ>
__device__ __forceinline__ uint32_t tile_img(const int32_t *swz, int b) {
  return SWZ ? (uint32_t)swz[b] : 1u << b;
}
template <bool SWZ>
__device__ __forceinline__ uint32_t tile_cmb(uint32_t x, uint32_t y) {
  return SWZ ? x ^ y : x | y;
}

// Per-step image table (kChainTabInts int32 per step in the workspace, written by
// k_chain_pack for the launch's layout): entry e < 13 - k is the LDS image of group
// bit e (the tile bits outside the step's contracted and batch bits in increasing
// order, then its batch bits), entry 16 + i the image of contracted bit i.  Images of
// distinct bits combine by XOR in either layout (plain images are distinct powers of
// two), so the step code needs no per-step bit scans.
//
// Group g's element kappa sits at D(g) ^ koff(kappa), D(g) the XOR of the images of
// g's set bits.  The FMA path: lane t owns groups t + kChainThreads i, whose D splits
// into the lane part D(t) and the uniform part of i.
template <typename T, int K>
__device__ __forceinline__ void chain_apply(typename Vec<T>::type *tile, const typename Vec<T>::type *__restrict__ Sp,
                                            const ChainStepArgs &s, const int32_t *__restrict__ tb, bool wl) {
  using V = typename Vec<T>::type;
  constexpr int LK = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : K == 16 ? 4 : 5;
  constexpr int TB = kChainTileBits - 4;   // group bits from the lane (kChainThreads = 2^TB)
  static_assert(kChainThreads == 1 << TB, "lane bits");
  const int t = threadIdx.x;
  const int nfree = kChainTileBits - LK - s.h;
  uint32_t kimg[LK > 0 ? LK : 1];
#pragma unroll
  for (int i = 0; i < LK; ++i) kimg[i] = (uint32_t)tb[16 + i];
  auto koff = [&](int kap) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < LK; ++i)
      if ((kap >> i) & 1) o ^= kimg[i];
    return o;
  };
  uint32_t dt = 0;
#pragma unroll
  for (int b = 0; b < TB; ++b) dt ^= ((t >> b) & 1) ? (uint32_t)tb[b] : 0u;
  auto srow = [&](uint32_t g) {
    int h = (int)(g >> nfree) & ((1 << s.h) - 1);
    h = __builtin_amdgcn_readfirstlane(h);
    return Sp + s.soff + h * K * K;
  };
  static_assert(K <= kChainPer, "larger steps run on matrix cores");
  constexpr int NGL = kChainPer / K;   // groups per lane
  if (K >= 4 && NGL > 1 && s.h == 0) {
    // one S for every group: each S value is loaded once and feeds the lane's NGL
    // groups (4 NGL independent FMA chains; with 1 or 2 contracted values the 8 or 16
    // groups' accumulators would not fit beside the prefetch registers)
    uint32_t base[NGL];
    V in[NGL][K];
#pragma unroll
    for (int i = 0; i < NGL; ++i) {
      base[i] = dt;
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((i >> b) & 1) base[i] ^= (uint32_t)tb[TB + b];
#pragma unroll
      for (int kap = 0; kap < K; ++kap) in[i][kap] = tile[base[i] ^ koff(kap)];
    }
    const V *Sr = Sp + s.soff;
#pragma unroll
    for (int nu = 0; nu < K; ++nu) {
      T xa[NGL], xb[NGL], ya[NGL], yb[NGL];
#pragma unroll
      for (int i = 0; i < NGL; ++i) xa[i] = xb[i] = ya[i] = yb[i] = T(0);
#pragma unroll
      for (int kap = 0; kap < K; ++kap) {
        const V sv = Sr[nu * K + kap];
#pragma unroll
        for (int i = 0; i < NGL; ++i) {
          xa[i] = fma(sv.x, in[i][kap].x, xa[i]);
          xb[i] = fma(sv.y, in[i][kap].y, xb[i]);
          ya[i] = fma(sv.x, in[i][kap].y, ya[i]);
          yb[i] = fma(sv.y, in[i][kap].x, yb[i]);
        }
      }
#pragma unroll
      for (int i = 0; i < NGL; ++i) {
        V o;
        o.x = xa[i] - xb[i];
        o.y = ya[i] + yb[i];
        tile[base[i] ^ koff(nu)] = o;
      }
    }
    step_end(wl);
    return;
  }
  // whole groups per lane, updated in place
#pragma unroll
  for (int i = 0; i < kChainPer / K; ++i) {
    const uint32_t g = (uint32_t)(t + kChainThreads * i);
    uint32_t base = dt;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if ((i >> b) & 1) base ^= (uint32_t)tb[TB + b];
    const V *Sr = srow(g);
    V in[K];
#pragma unroll
    for (int kap = 0; kap < K; ++kap) in[kap] = tile[base ^ koff(kap)];
    // the real and imaginary parts accumulate as two partial sums each: four
    // independent FMA chains per output value
#pragma unroll
    for (int nu = 0; nu < K; ++nu) {
      T xa = 0, xb = 0, ya = 0, yb = 0;
#pragma unroll
      for (int kap = 0; kap < K; ++kap) {
        const V sv = Sr[nu * K + kap];
        xa = fma(sv.x, in[kap].x, xa);
        xb = fma(sv.y, in[kap].y, xb);
        ya = fma(sv.x, in[kap].y, ya);
        yb = fma(sv.y, in[kap].x, yb);
      }
      V o;
      o.x = xa - xb;
      o.y = ya + yb;
      tile[base ^ koff(nu)] = o;
    }
  }
  step_end(wl);
}

// Steps with 16 or 32 contracted values on matrix cores: out[g][nu] = sum_kappa
// S[kappa -> nu] in[g][kappa], as 16x16x4 products of S^T (A operand: nu = lane % 16,
// kappa = 4 kb + lane / 16) and in^T (B operand: kappa = 4 kb + lane / 16, group =
// lane % 16), so that the result's columns are groups: 16 lanes with consecutive
// lane % 16 read and write 16 groups that differ in the low group bits, the pattern
// the tile swizzle spreads over the LDS banks (with nu on the columns, the writes
// would spread over the contracted bits, which it cannot always spread).  Wave w owns
// four blocks of 16 groups for one block of 16 nu; its 64 groups share the batch
// bits, so its S fragments stay in registers for the whole step.  Operand and result
// layouts as in Mfma16.  Addresses from the step's image table (as in chain_apply):
// a group block's base from the uniform group bits 4 and up, a group inside it from
// group bits 0-3.
// G3: the complex product as three real products (Gauss: t1 = (ar + ai) br,
// t2 = ar (bi - br), t3 = ai (br + bi); re = t1 - t3, im = t1 + t2), a quarter fewer
// matrix-core ops for one vector add per A element.
template <typename T, int K, bool SL, bool G3 = false>
__device__ __forceinline__ void chain_apply_mfma(typename Vec<T>::type *tile,
                                                 const typename Vec<T>::type *__restrict__ Sp,
                                                 const typename Vec<T>::type *sS, const ChainStepArgs &s,
                                                 const int32_t *__restrict__ tb, bool wl) {
  using V = typename Vec<T>::type;
  using M = Mfma16<T>;
  using v4 = typename M::v4;
  constexpr int LK = K == 16 ? 4 : 5;
  constexpr int NG = kChainTile / K, NB = K / 16, KB = K / 4;
  constexpr int WAVES = kChainThreads / 64, RT = (NG / 16) * NB / WAVES;
  static_assert(RT == 4, "row blocks per wave");
  // the lane from the exec mask (v_mbcnt) rather than the work-item id: the compiler can
  // recompute it where needed instead of keeping a lane value live (and spilling it)
  // across the whole tile loop
  const int lane = (int)__lane_id(), wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x) >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nb = wave % NB;
  const int g0 = (wave / NB) * RT * 16;
  const int nfree = kChainTileBits - LK - s.h;
  uint32_t kimg[LK];
#pragma unroll
  for (int i = 0; i < LK; ++i) kimg[i] = (uint32_t)tb[16 + i];
  auto koff = [&](int kap) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < LK; ++i)
      if ((kap >> i) & 1) o ^= kimg[i];
    return o;
  };
  auto scatter4 = [&](int row) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if ((row >> b) & 1) o ^= (uint32_t)tb[b];
    return o;
  };
  // uniform part of the wave's row blocks: group bits 6 and up from g0 (a multiple of
  // 64), bits 4-5 from the row block
  uint32_t dw = 0;
#pragma unroll
  for (int b = 6; b < kChainTileBits - LK; ++b)
    if ((g0 >> b) & 1) dw ^= (uint32_t)tb[b];
  uint32_t rbase[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
    rbase[rt] = dw ^ ((rt & 1) ? (uint32_t)tb[4] : 0u) ^ ((rt & 2) ? (uint32_t)tb[5] : 0u);
  const int h = (g0 >> nfree) & ((1 << s.h) - 1);   // the wave's groups share their batch bits
  const int so = s.soff + h * K * K + (16 * nb + li) * K + lk;
  T bre[KB], bim[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    // from LDS when staged: a global load here would wait for the prefetch too
    const V v = SL ? sS[ss_slot(so + 4 * kb)] : Sp[so + 4 * kb];
    bre[kb] = v.x;
    bim[kb] = v.y;
  }
  const uint32_t arow = scatter4(li);
  // waves of other nu blocks read these groups: all reads before any write (with
  // one nu block every wave owns its groups alone)
  auto store = [&](int rt, const v4 &re, const v4 &im) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      V o;
      o.x = re[r];
      o.y = im[r];
      tile[rbase[rt] ^ arow ^ koff(16 * nb + M::row(lk, r))] = o;
    }
  };
  if constexpr (!G3) {
    v4 cre[RT], cim[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) cre[rt] = v4{0, 0, 0, 0}, cim[rt] = v4{0, 0, 0, 0};
    // k blocks outer, row blocks inner: the 2 RT accumulators are independent, so
    // consecutive matrix-core ops never wait for each other's results
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      V a[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) a[rt] = tile[rbase[rt] ^ arow ^ koff(4 * kb + lk)];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cre[rt] = M::mma(bre[kb], a[rt].x, cre[rt]);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cim[rt] = M::mma(bim[kb], a[rt].x, cim[rt]);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cre[rt] = M::mma(bim[kb], -a[rt].y, cre[rt]);
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) cim[rt] = M::mma(bre[kb], a[rt].y, cim[rt]);
    }
    if (NB > 1) lds_barrier();
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) store(rt, cre[rt], cim[rt]);
  } else {
    static_assert(NB == 1, "three-product form for one nu block");
    // two row blocks at a time (three accumulators each keep the registers of the
    // four-product form): t1, t2, t3 as above, 6 independent accumulators per k block
#pragma unroll
    for (int half = 0; half < RT / 2; ++half) {
      v4 c1[2], c2[2], c3[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) c1[u] = c2[u] = c3[u] = v4{0, 0, 0, 0};
      // the half's A values are read up front, so that the reads of later k blocks
      // are in flight during the matrix-core ops of earlier ones
      V ah[KB][2];
#pragma unroll
      for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int u = 0; u < 2; ++u) ah[kb][u] = tile[rbase[2 * half + u] ^ arow ^ koff(4 * kb + lk)];
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const T bd = bim[kb] - bre[kb], bs = bre[kb] + bim[kb];
        const V *a = ah[kb];
#pragma unroll
        for (int u = 0; u < 2; ++u) c1[u] = M::mma(bre[kb], a[u].x + a[u].y, c1[u]);
#pragma unroll
        for (int u = 0; u < 2; ++u) c2[u] = M::mma(bd, a[u].x, c2[u]);
#pragma unroll
        for (int u = 0; u < 2; ++u) c3[u] = M::mma(bs, a[u].y, c3[u]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) store(2 * half + u, c1[u] - c3[u], c1[u] + c2[u]);
      // scheduling: the half's 2 KB LDS reads first, then its 6 KB matrix-core ops
      __builtin_amdgcn_sched_group_barrier(0x100, 2 * KB, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 6 * KB, 0);
    }
  }
  step_end(NB == 1 && wl);
}

// Gather every S_j of a chain into the dense [h][nu][kappa] blocks the chain
// kernel reads: step j = blockIdx.y, elements over blockIdx.x (at most kChainMaxS).
// The first workgroup of each step also writes the step's image table (at
// a.tab_off elements, kChainTabInts int32 per step) for the launch's layout.
template <typename T>
__global__ void k_chain_pack(typename Vec<T>::type *__restrict__ Sp, ChainArgs a) {
  using V = typename Vec<T>::type;
  {
    const int j = blockIdx.y;
    const ChainStepArgs &s = a.st[j];
    if (blockIdx.x == 0 && threadIdx.x < kChainTabInts) {
      const int e = threadIdx.x;
      uint32_t used = 0;
      for (int i = 0; i < s.k; ++i) used |= 1u << s.q[i];
      for (int i = 0; i < s.h; ++i) used |= 1u << s.hq[i];
      int bit = -1;
      const int nfree = kChainTileBits - s.k - s.h;
      if (e < nfree) {
        // group bits 6-8 on the wave bits of a wave-local chain, the others on the
        // remaining free tile bits in increasing order
        int order[kChainTileBits];
        int n = 0;
        uint32_t taken = used;
        if (a.wave_local)
          for (int i = 0; i < 3; ++i) taken |= 1u << a.wbits[i];
        for (int b = 0; b < kChainTileBits; ++b) {
          if (a.wave_local && n == 6)
            for (int i = 0; i < 3; ++i) order[n++] = a.wbits[i];
          if (!((taken >> b) & 1)) order[n++] = b;
        }
        if (a.wave_local && n == 6)
          for (int i = 0; i < 3; ++i) order[n++] = a.wbits[i];
        bit = order[e];
      } else if (e < kChainTileBits - s.k) {
        bit = s.hq[e - (kChainTileBits - s.k - s.h)];
      } else if (e >= 16 && e < 16 + s.k) {
        bit = s.q[e - 16];
      }
      const int32_t img = bit < 0 ? 0 : a.swizzled ? a.swz[bit] : (int32_t)(1u << bit);
      reinterpret_cast<int32_t *>(Sp + a.tab_off)[kChainTabInts * j + e] = img;
    }
    const V *S = (const V *)a.S[j];
    const int kk = 1 << s.k, tot = s.merged ? 0 : (1 << s.h) * kk * kk;   // merged: k_chain_merge
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
      const int kap = e % kk, nu = (e / kk) % kk, h = e / (kk * kk);
      int64_t o = 0;
      for (int i = 0; i < s.k; ++i) {
        if ((kap >> i) & 1) o += s.sk[i];
        if ((nu >> i) & 1) o += s.sn[i];
      }
      for (int i = 0; i < s.h; ++i)
        if ((h >> i) & 1) o += s.sh[i];
      Sp[s.soff + e] = S[o];
    }
  }
}

// Merged steps: the product of the members' gate blocks, each acting on some of the
// step's bits (identity on the others), first member applied first.  Workgroup g
// holds the running product [nu][kappa] in LDS (lane nu K + kappa), stages each
// member's block in LDS and applies it: new[nu][kappa] = sum over the member's input
// bits s of S[nu's member bits][s] cur[nu with its member bits set to s][kappa].
// Written packed as [nu][kappa] like the single blocks.
template <typename T>
__global__ __launch_bounds__(256) void k_chain_merge(typename Vec<T>::type *__restrict__ Sp, ChainMergeArgs m) {
  using V = typename Vec<T>::type;
  __shared__ V cur[256], nxt[256], blk[256];
  const ChainMergeArgs::Group &g = m.g[blockIdx.x];
  const int K = 1 << g.k, t = threadIdx.x, nu = t / K, kap = t % K;
  const bool on = t < K * K;
  if (on) cur[t] = V{nu == kap ? T(1) : T(0), T(0)};
  for (int c = 0; c < g.count; ++c) {
    const ChainMergeArgs::Member &mm = m.m[g.first + c];
    const int kk = 1 << mm.kc;
    __syncthreads();
    if (t < kk * kk) {   // blk[so kk + si] = S[si -> sk strides, so -> sn strides]
      const int so = t / kk, si = t % kk;
      int64_t o = 0;
      for (int i = 0; i < mm.kc; ++i) {
        if ((si >> i) & 1) o += mm.sk[i];
        if ((so >> i) & 1) o += mm.sn[i];
      }
      blk[t] = ((const V *)mm.S)[o];
    }
    __syncthreads();
    if (on) {
      int sub_out = 0, rest = nu;
      for (int i = 0; i < mm.kc; ++i) {
        sub_out |= ((nu >> mm.pos[i]) & 1) << i;
        rest &= ~(1 << mm.pos[i]);
      }
      V acc{0, 0};
      for (int si = 0; si < kk; ++si) {
        int x = rest;
        for (int i = 0; i < mm.kc; ++i)
          if ((si >> i) & 1) x |= 1 << mm.pos[i];
        cmac<T>(acc, blk[sub_out * kk + si], cur[x * K + kap], 0, 0);
      }
      nxt[t] = acc;
    }
    __syncthreads();
    if (on) cur[t] = nxt[t];
  }
  if (on) Sp[g.soff + t] = cur[t];
}

// SWZ: tile loads and stores address LDS through the host's swizzle table (a separate
// instance: the plain one keeps the cheaper addressing); the steps take their
// addresses from the image tables, which k_chain_pack writes for the launch's layout.
template <typename T, bool SWZ>
__global__ __launch_bounds__(kChainThreads) void k_contract_chain(const typename Vec<T>::type *__restrict__ X0,
                                                                  typename Vec<T>::type *__restrict__ Xm,
                                                                  const typename Vec<T>::type *__restrict__ Sp,
                                                                  ChainArgs a) {
  using V = typename Vec<T>::type;
  __shared__ V tile[kChainTile];
  __shared__ V sS[kChainSLds<T>];   // packed S_j of every step (at ss_slot), when they fit
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t) >> 6;
  const bool s_lds = a.s_elems <= kChainSLds<T>;
  const int32_t *__restrict__ tab = reinterpret_cast<const int32_t *>(Sp + a.tab_off);
  const bool wl = a.wave_local != 0;
  const int64_t ntiles = int64_t(1) << a.nouter;
  // X_0 / X_m offsets of tile w: the outer bits' strides summed over w's set bits
  auto tile_base = [&](int64_t w, const int64_t *o) {
    int64_t b = 0;
#pragma unroll 4
    for (int j = 0; j < a.nouter; ++j)
      if ((w >> j) & 1) b += o[j];
    return b;
  };
  // the same for tile w + gridDim.x from tile w's offset b, when w + gridDim.x is a
  // tile: with a power-of-two grid (2^gb workgroups) the low gb bits stay and the high
  // part counts up by one (its lowest zero bit c sets, the c ones below clear), so a
  // workgroup's next offset costs about two stride loads instead of one per outer bit
  const int gsz = (int)gridDim.x;
  const bool grid_pow2 = (gsz & (gsz - 1)) == 0;
  const int gb = __builtin_ctz((unsigned)gsz);
  auto next_base = [&](int64_t w, int64_t b, const int64_t *o) {
    if (!grid_pow2) return tile_base(w + gsz, o);
    const int c = __builtin_ctzll(~((uint64_t)w >> gb));
    b += o[gb + c];
    for (int j = 0; j < c; ++j) b -= o[gb + j];
    return b;
  };
  // lane t holds tile elements t + kChainThreads * r: the low tile bits come from t,
  // the top four from r
  constexpr int RB = 4;
  static_assert(kChainPer == 1 << RB && kChainThreads == 1 << (kChainTileBits - RB), "tile split");
  // offsets of the lane's element r = 0 in X_0 and X_m (sums of distinct power-of-two
  // strides of a tensor of fewer than 2^32 elements, which the host guarantees: 32 bits
  // keep these registers, live across every step, few) and its LDS addresses (load /
  // store order)
  uint32_t lin = 0, lout = 0;
  uint32_t sin = 0, sout = 0;
  for (int b = 0; b < kChainTileBits - RB; ++b)
    if ((t >> b) & 1) {
      lin += (uint32_t)a.win[b];
      lout += (uint32_t)a.wout[a.pout[b]];
      sin = tile_cmb<SWZ>(sin, tile_img<SWZ>(a.swz, b));
      sout = tile_cmb<SWZ>(sout, tile_img<SWZ>(a.swz, a.pout[b]));
    }
  // offsets of element r of this lane: the top RB tile bits come from r (a
  // compile-time constant, so these fold to sums of kernel arguments)
  auto rin = [&](auto r) {
    int64_t o = (int64_t)lin;
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if ((r >> i) & 1) o += a.win[kChainTileBits - RB + i];
    return o;
  };
  auto rout = [&](auto r) {
    int64_t o = (int64_t)lout;
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if ((r >> i) & 1) o += a.wout[a.pout[kChainTileBits - RB + i]];
    return o;
  };
  // element r's LDS address: the lane's part (`li`: sin, `lo`: sout) combined with the
  // images of the r bits (in load order the top tile bits; in store order the last four
  // X_m tile bits).  Plain, the load-order part is a constant above every bit of sin,
  // which folds into the LDS instructions; swizzled, one XOR with a uniform value per
  // element.  The lane parts pass through `fresh` at each use, so that the sixteen
  // addresses are formed next to their accesses rather than hoisted out of the tile
  // loop, where they would stay live across the steps (and spill).
  auto fresh = [](uint32_t x) {
    asm volatile("" : "+v"(x));
    return x;
  };
  auto rsi = [&](uint32_t li, auto r) {
    if constexpr (!SWZ) {
      return li + (uint32_t)(kChainThreads * r);
    } else {
      uint32_t o = li;
#pragma unroll
      for (int i = 0; i < RB; ++i)
        if ((r >> i) & 1) o ^= (uint32_t)a.swz[kChainTileBits - RB + i];
      return o;
    }
  };
  auto rso = [&](uint32_t lo, auto r) {
    uint32_t o = lo;
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if ((r >> i) & 1) o = tile_cmb<SWZ>(o, tile_img<SWZ>(a.swz, a.pout[kChainTileBits - RB + i]));
    return o;
  };
  // prefetch registers as native two-element vectors (a struct copy would keep them
  // in private memory)
  typedef T E __attribute__((ext_vector_type(2)));
  const E *X0e = reinterpret_cast<const E *>(X0);
  E *tile_e = reinterpret_cast<E *>(tile);
  E pf[kChainPer];
  // The next tile's loads are spread over this tile's steps (a quarter of them before
  // the steps at 0, n/4, n/2 and 3n/4): issued in one burst they stall the waves at
  // issue until most have returned, and no step work overlaps them.  They are waited
  // for after this tile's stores, which stay in flight across that wait (one in-order
  // counter covers loads and stores).
  auto load_quarter = [&](int q, int64_t bn) {
    auto ld = [&](auto r) { pf[r] = X0e[bn + rin(r)]; };
    switch (q) {
      case 0: static_for<4>([&](auto r) { ld(std::integral_constant<int, decltype(r)::value>{}); }); break;
      case 1: static_for<4>([&](auto r) { ld(std::integral_constant<int, 4 + decltype(r)::value>{}); }); break;
      case 2: static_for<4>([&](auto r) { ld(std::integral_constant<int, 8 + decltype(r)::value>{}); }); break;
      default: static_for<4>([&](auto r) { ld(std::integral_constant<int, 12 + decltype(r)::value>{}); }); break;
    }
  };
  static_assert(kChainPer == 16, "four quarters of four elements");
  int64_t w = blockIdx.x;
  int64_t bi = 0, bo = 0;   // the current tile's X_0 and X_m offsets
  // the first tile's loads and the staged gate blocks' loads in flight together
  if (w < ntiles) {
    bi = tile_base(w, a.oin);
    bo = tile_base(w, a.oout);
    static_for<kChainPer>([&](auto r) { pf[r] = X0e[bi + rin(r)]; });
  }
  if (s_lds) {
    // every lane loads (an index past the blocks reads the last element again), only
    // indices inside are stored
    constexpr int NS = kChainSLds<T> / kChainThreads;
    V sv[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) sv[i] = Sp[min(t + kChainThreads * i, a.s_elems - 1)];
#pragma unroll
    for (int i = 0; i < NS; ++i)
      if (t + kChainThreads * i < a.s_elems) sS[ss_slot(t + kChainThreads * i)] = sv[i];
  }
  // offset of the next tile; its first quarter of loads is issued right after the
  // current tile's refill, so that those loads run during the barrier and not during
  // the tile's first step
  int64_t bn = 0;
  if (w < ntiles) {
    static_for<kChainPer>([&](auto r) { tile_e[rsi(sin, r)] = pf[r]; });
    if (w + gridDim.x < ntiles) {
      bn = next_base(w, bi, a.oin);
      load_quarter(0, bn);
    }
    lds_barrier();
  }
  for (; w < ntiles; w += gridDim.x) {
    const bool more = w + gridDim.x < ntiles;
    const int n = a.nsteps;
    for (int j = 0; j < n; ++j) {
      if (more)
        for (int q = 1; q < 4; ++q)
          if (q * n / 4 == j) load_quarter(q, bn);
      // wave-local steps: the two waves of a SIMD (waves w and w + 4) take turns at the
      // higher issue priority, step by step, so that neither runs far ahead and the
      // other is left to finish its steps alone before the barrier of the stores
      if (wl) {
        if (((wave >> 2) ^ j) & 1)
          __builtin_amdgcn_s_setprio(1);
        else
          __builtin_amdgcn_s_setprio(0);
      }
      const ChainStepArgs &s = a.st[j];
      const int32_t *tb = tab + kChainTabInts * j;
      switch (s.k) {
        case 0: chain_apply<T, 1>(tile, Sp, s, tb, wl); break;
        case 1: chain_apply<T, 2>(tile, Sp, s, tb, wl); break;
        case 2: chain_apply<T, 4>(tile, Sp, s, tb, wl); break;
        case 3: chain_apply<T, 8>(tile, Sp, s, tb, wl); break;
        case 4:
          // mfma16: 0 vector FMAs, 1 matrix cores (four real products), 2 matrix cores
          // (three real products, when the gate blocks are staged in LDS)
          if (!a.mfma16)
            chain_apply<T, 16>(tile, Sp, s, tb, wl);
          else if (a.mfma16 == 2 && s_lds)
            chain_apply_mfma<T, 16, true, true>(tile, Sp, sS, s, tb, wl);
          else if (s_lds)
            chain_apply_mfma<T, 16, true>(tile, Sp, sS, s, tb, wl);
          else
            chain_apply_mfma<T, 16, false>(tile, Sp, sS, s, tb, wl);
          break;
        default:
          if (s_lds)
            chain_apply_mfma<T, 32, true>(tile, Sp, sS, s, tb, wl);
          else
            chain_apply_mfma<T, 32, false>(tile, Sp, sS, s, tb, wl);
          break;
      }
    }
    // the stores read the tile in store order, across waves
    if (wl) {
      __builtin_amdgcn_s_setprio(0);
      lds_barrier();
    }
    E *Xme = reinterpret_cast<E *>(Xm);
    {
      const uint32_t lo = fresh(sout);
      static_for<kChainPer>([&](auto r) { Xme[bo + rout(r)] = tile_e[rso(lo, r)]; });
    }
    if (more) bo = next_base(w, bo, a.oout);
    bi = bn;
    lds_barrier();
    if (more) {
      const uint32_t li = fresh(sin);
      static_for<kChainPer>([&](auto r) { tile_e[rsi(li, r)] = pf[r]; });
      if (w + 2 * (int64_t)gridDim.x < ntiles) {
        bn = next_base(w + gridDim.x, bi, a.oin);
        load_quarter(0, bn);
      }
      lds_barrier();
    }
  }
}

// ---------------------------------------------------------------------------
// Small-step batch: workgroup i contracts C elements [c0, c1) of the step of
// descriptor i (lanes along C, the K loop in registers; a range shorter than the
// workgroup splits each element's K range over 256 / (c1 - c0) lanes, summed in
// LDS).  The steps of one launch are independent.

// Sum over k = k0, k0 + dk, ... < nk of A[oa + offset_a(k)] * B[ob + offset_b(k)] for a
// small-step descriptor, in that order: the operands of 8 terms are loaded before
// they are used, so that their loads are in flight together.
template <typename T>
__device__ __forceinline__ typename Vec<T>::type small_dot(const SmallDesc &s, const typename Vec<T>::type *A,
                                                           const typename Vec<T>::type *B, int64_t oa, int64_t ob,
                                                           int k0, int dk, int nk) {
  using V = typename Vec<T>::type;
  constexpr int U = 8;
  V acc{0, 0};
  int k = k0;
  for (; k + (U - 1) * dk < nk; k += U * dk) {
    V av[U], bv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int kk = k + u * dk;
      int64_t xa = oa, xb = ob;
      for (int b = 0; b < s.nbk; ++b)
        if ((kk >> b) & 1) xa += s.ka[b], xb += s.kb[b];
      av[u] = A[xa];
      bv[u] = B[xb];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) cmac<T>(acc, av[u], bv[u], s.conj_a, s.conj_b);
  }
  for (; k < nk; k += dk) {
    int64_t xa = oa, xb = ob;
    for (int b = 0; b < s.nbk; ++b)
      if ((k >> b) & 1) xa += s.ka[b], xb += s.kb[b];
    cmac<T>(acc, A[xa], B[xb], s.conj_a, s.conj_b);
  }
  return acc;
}

// One chunk of a small step (descriptor `s`) by the workgroup's kSmallThreads lanes.
template <typename T>
__device__ __forceinline__ void small_item(const SmallDesc &s, typename Vec<T>::type *part) {
  using V = typename Vec<T>::type;
  const V *A = (const V *)s.A;
  const V *B = (const V *)s.B;
  V *C = (V *)s.C;
  const int nk = 1 << s.nbk;
  const int len = s.c1 - s.c0;   // a power of two
  if (len < kSmallThreads && nk > 1) {   // lanes split K: 2^lt lanes per element of the chunk
    const int lt = 8 - (__ffs(len) - 1);   // kSmallThreads = 2^8
    const int c = s.c0 + (threadIdx.x >> lt), sub = threadIdx.x & ((1 << lt) - 1);
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < s.nbc; ++b)
      if ((c >> b) & 1) oa += s.wa[b], ob += s.wb[b];
    V acc = small_dot<T>(s, A, B, oa, ob, sub, 1 << lt, nk);
    part[threadIdx.x] = acc;
    __syncthreads();
    if (sub == 0) {
      V sum = part[threadIdx.x];
      for (int j = 1; j < (1 << lt); ++j) {
        sum.x += part[threadIdx.x + j].x;
        sum.y += part[threadIdx.x + j].y;
      }
      C[c] = sum;
    }
    __syncthreads();   // `part` is free again
    return;
  }
  for (int c = s.c0 + threadIdx.x; c < s.c1; c += blockDim.x) {
    int64_t oa = 0, ob = 0;
    for (int b = 0; b < s.nbc; ++b)
      if ((c >> b) & 1) oa += s.wa[b], ob += s.wb[b];
    C[c] = small_dot<T>(s, A, B, oa, ob, 0, 1, nk);
  }
}

template <typename T>
__global__ __launch_bounds__(kSmallThreads) void k_small_batch(const SmallDesc *__restrict__ d) {
  __shared__ typename Vec<T>::type part[kSmallThreads];
  small_item<T>(d[blockIdx.x], part);
}

// Every dependency level of a section's small steps in one launch by one workgroup:
// the levels' chunks in order, a workgroup barrier between levels (so a level reads
// what the previous ones wrote).  For sections whose small steps are tiny in total:
// the levels' launch gaps cost more than their arithmetic.  One workgroup keeps every
// producer and consumer on one CU; workgroups on several XCDs would need an L2
// write-back and invalidate at each level to see each other's results, which costs
// more than a launch.
template <typename T>
__global__ __launch_bounds__(kSmallThreads) void k_small_walk(const SmallDesc *__restrict__ d,
                                                              const int *__restrict__ level_off, int nlevels) {
  __shared__ typename Vec<T>::type part[kSmallThreads];
  for (int L = 0; L < nlevels; ++L) {
    const int i1 = level_off[L + 1];
    for (int i = level_off[L]; i < i1; ++i) small_item<T>(d[i], part);
    __syncthreads();
  }
}

// Batched scatter-add: workgroup i adds descriptor i's dense temp into its FP64
// destination through the destination's strides.  Destinations of one launch are
// disjoint.
template <typename TS>
__global__ __launch_bounds__(256) void k_add_batch(const AddDesc *__restrict__ d) {
  using VS = typename Vec<TS>::type;
  const AddDesc &a = d[blockIdx.x];
  const VS *src = (const VS *)a.src;
  for (int64_t e = threadIdx.x; e < a.n; e += blockDim.x) {
    int64_t r = e, o = 0;
    for (int j = 0; j < a.nd; ++j) {
      const int64_t q = r % a.dim[j];
      r /= a.dim[j];
      o += q * a.st[j];
    }
    const VS v = src[e];
    if (a.dst_f64) {
      double2 *dst = (double2 *)a.dst;
      dst[o].x += (double)v.x;
      dst[o].y += (double)v.y;
    } else {
      VS *dst = (VS *)a.dst;
      dst[o].x += v.x;
      dst[o].y += v.y;
    }
  }
}

// ---------------------------------------------------------------------------
// Output sampling on the device.  Counter-based Philox4x64-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011): candidate j
// of a stream with seed s draws its bits from the block at counter (j, 0, 0, 0),
// key (s, 0) -- character q is bit q of the 256-bit block -- and its acceptance
// uniform from the first word of the block at key (s, 1), top 53 bits.
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

__global__ void k_gen_bits(char *__restrict__ bits, int64_t nb, int nch, uint64_t seed, int64_t j0) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t c[4] = {(uint64_t)(j0 + i), 0, 0, 0};
    philox4x64(c, seed, 0);
    char *o = bits + i * nch;
    for (int q = 0; q < nch; ++q) o[q] = '0' + (char)((c[q >> 6] >> (q & 63)) & 1);
  }
}

// Empirical-supremum rejection over candidates j0 .. j0+nb-1 (amplitudes `amps`,
// bitstrings `cand`), in candidate order: r_j = |a_j|^2 * scale, M_j = M_{j-1} or
// max(M_{j-1}, r_j) when M is not fixed, accept when u_j * M_j < r_j; stop at the
// num_samples-th acceptance.  One workgroup of 256 lanes walks the candidates 256
// at a time with max and sum scans in LDS.  state[0] = M; cnt[0] = accepted so
// far, cnt[1] = candidates consumed so far.
__global__ __launch_bounds__(256) void k_reject(const double2 *__restrict__ amps, const char *__restrict__ cand,
                                                int64_t nb, int nch, uint64_t seed, int64_t j0, double scale,
                                                int fix_m, int64_t num_samples, double *state, int64_t *cnt,
                                                char *__restrict__ out_bits, double2 *__restrict__ out_amps) {
  __shared__ double smax[256];
  __shared__ int64_t ssum[256];
  __shared__ int64_t s_stop;
  const int t = threadIdx.x;
  double carry_m = state[0];
  int64_t carry_n = cnt[0];
  int64_t consumed = nb;
  for (int64_t base = 0; base < nb; base += 256) {
    const int64_t j = base + t;
    const bool live = j < nb;
    double r = 0.0;
    double2 a = {0.0, 0.0};
    if (live) {
      a = amps[j];
      r = __dmul_rn(__dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)), scale);
    }
    // inclusive max scan of r (M_j)
    smax[t] = live ? r : 0.0;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      double v = t >= off ? smax[t - off] : 0.0;
      __syncthreads();
      smax[t] = fmax(smax[t], v);
      __syncthreads();
    }
    const double mj = fix_m ? carry_m : fmax(carry_m, smax[t]);
    bool acc = false;
    if (live) {
      uint64_t c[4] = {(uint64_t)(j0 + j), 0, 0, 0};
      philox4x64(c, seed, 1);
      const double u = (double)(c[0] >> 11) * 0x1.0p-53;
      acc = __dmul_rn(u, mj) < r;
    }
    // inclusive sum scan of acceptances
    ssum[t] = acc ? 1 : 0;
    if (t == 0) s_stop = -1;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int64_t v = t >= off ? ssum[t - off] : 0;
      __syncthreads();
      ssum[t] += v;
      __syncthreads();
    }
    const int64_t rank = carry_n + ssum[t] - 1;   // output slot of an accepted candidate
    if (acc && rank < num_samples) {
      for (int q = 0; q < nch; ++q) out_bits[rank * nch + q] = cand[j * nch + q];
      out_amps[rank] = a;
      if (rank == num_samples - 1) s_stop = j;
    }
    __syncthreads();
    const int64_t stop = s_stop;
    const int last = stop >= 0 ? (int)(stop - base) : (int)((nb - base < 256 ? nb - base : 256) - 1);
    const double m_last = fix_m ? carry_m : fmax(carry_m, smax[last]);
    const int64_t n_last = carry_n + ssum[last];
    __syncthreads();
    carry_m = m_last;
    carry_n = n_last;
    if (stop >= 0) {
      consumed = stop + 1;
      break;
    }
  }
  if (t == 0) {
    state[0] = carry_m;
    cnt[0] = carry_n;
    cnt[1] += consumed;
  }
}

void launch_fill_outputs(int prec, void *arena, const int64_t *d_out_off, int n_out, const int *d_out_q,
                         const char *d_bits, int64_t nvalid, int64_t batch, int64_t num_chars, void *stream) {
  if (n_out == 0) return;
  int g = grid_for((int64_t)n_out * batch, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_fill_outputs<double>, dim3(g), dim3(256), 0, s, (double2 *)arena, d_out_off, n_out,
                       d_out_q, d_bits, nvalid, batch, num_chars);
  else
    hipLaunchKernelGGL(k_fill_outputs<float>, dim3(g), dim3(256), 0, s, (float2 *)arena, d_out_off, n_out,
                       d_out_q, d_bits, nvalid, batch, num_chars);
  check_launch("fill_outputs");
}

void launch_accumulate(int prec, const void *root, int64_t root_bstride, int64_t batch, int64_t nvalid, int64_t relems,
                       const int64_t *d_relem_off, double *acc, int64_t row0, void *stream) {
  int g = grid_for(nvalid * relems, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_accumulate<double>, dim3(g), dim3(256), 0, s, (const double2 *)root, root_bstride, batch,
                       nvalid, relems, d_relem_off, acc, row0);
  else
    hipLaunchKernelGGL(k_accumulate<float>, dim3(g), dim3(256), 0, s, (const float2 *)root, root_bstride, batch,
                       nvalid, relems, d_relem_off, acc, row0);
  check_launch("accumulate");
}


void launch_add_mapped(int prec, int dst_double, void *dst, const void *src, int64_t n, const std::vector<int64_t> &dims,
                       const std::vector<int64_t> &strides, void *stream) {
  if (n <= 0) return;
  if ((int)dims.size() > kMaxLegs) throw Error(QX_ERR_UNSUPPORTED, "add of a tensor with too many legs");
  AddArgs a;
  a.n = (int)dims.size();
  for (int j = 0; j < a.n; ++j) a.dim[j] = dims[j], a.st[j] = strides[j];
  int g = grid_for(n, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL((k_add_mapped<double, double>), dim3(g), dim3(256), 0, s, (double2 *)dst,
                       (const double2 *)src, n, a);
  else if (dst_double)
    hipLaunchKernelGGL((k_add_mapped<float, double>), dim3(g), dim3(256), 0, s, (double2 *)dst,
                       (const float2 *)src, n, a);
  else
    hipLaunchKernelGGL((k_add_mapped<float, float>), dim3(g), dim3(256), 0, s, (float2 *)dst, (const float2 *)src,
                       n, a);
  check_launch("add_mapped");
}

void launch_root_cot_norm(int prec, void *dst, const void *val, int64_t dep, int64_t batch, int64_t nvalid,
                          int64_t relems, const int64_t *d_relem_off, double *amp, int64_t row0, void *stream) {
  const int g = grid_for(batch * relems, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_root_cot_norm<double>, dim3(g), dim3(256), 0, s, (double2 *)dst, (const double2 *)val, dep,
                       batch, nvalid, relems, d_relem_off, amp, row0);
  else
    hipLaunchKernelGGL(k_root_cot_norm<float>, dim3(g), dim3(256), 0, s, (float2 *)dst, (const float2 *)val, dep,
                       batch, nvalid, relems, d_relem_off, amp, row0);
  check_launch("root_cot_norm");
}

void launch_root_cot(int prec, void *dst, int64_t dep, int64_t batch, int64_t nvalid, int64_t relems,
                     const int64_t *d_relem_off, const double *amp, int64_t row0, void *stream) {
  int g = grid_for(dep ? batch * relems : relems, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_root_cot<double>, dim3(g), dim3(256), 0, s, (double2 *)dst, dep, batch, nvalid, relems,
                       d_relem_off, amp, row0);
  else
    hipLaunchKernelGGL(k_root_cot<float>, dim3(g), dim3(256), 0, s, (float2 *)dst, dep, batch, nvalid, relems,
                       d_relem_off, amp, row0);
  check_launch("root_cot");
}

void launch_finalize(const double *acc, double *out, int64_t n, void *stream) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, acc, out, n);
  check_launch("finalize");
}

void launch_bf16_split(const void *src, int64_t n, void *out, void *stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_bf16_split, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, (const float2 *)src, n,
                     (uint16_t *)out);
  check_launch("bf16_split");
}

// Tile bits of the LDS gather: alternately the next destination bit and the bit of
// the next source weight (1, 2, 4, ...) while it exists, until `cap` bits.
bool gather_tiles(int nb, const int64_t *w, int cap, GatherTileArgs &g) {
  bool in_t[kMaxBits] = {};
  std::vector<int> src_order;   // tile bits that are the source's low bits, in source order
  int nt = 0, a = 0, b = 0;
  bool more_b = true;
  while (nt < cap && (a < nb || more_b)) {
    if (a < nb) {
      if (!in_t[a]) in_t[a] = true, ++nt;
      ++a;
    }
    if (nt < cap && more_b) {
      int p = -1;
      for (int q = 0; q < nb; ++q)
        if (w[q] == (int64_t(1) << b)) p = q;
      if (p < 0) {
        more_b = false;
      } else {
        if (!in_t[p]) in_t[p] = true, ++nt;
        src_order.push_back(p);
        ++b;
      }
    }
  }
  g = GatherTileArgs{};
  int pos_of[kMaxBits];
  for (int q = 0; q < nb; ++q) {
    if (in_t[q]) {
      pos_of[q] = g.nt;
      g.tdst[g.nt] = q;
      g.tw[g.nt] = w[q];
      ++g.nt;
    } else {
      g.rdst[g.nr] = q;
      g.rw[g.nr] = w[q];
      ++g.nr;
    }
  }
  int j = 0;
  bool used[12] = {};
  for (int p : src_order) g.lbit[j++] = pos_of[p], used[pos_of[p]] = true;
  for (int q = 0; q < g.nt; ++q)
    if (!used[q]) g.lbit[j++] = q;
  return g.nt == nt && j == g.nt && g.nt >= 6;
}

void launch_gather(int prec, const void *src, void *dst, int nb, const int64_t *w, void *stream) {
  if (nb < 0 || nb > kMaxBits) throw Error(QX_ERR_STATE, "internal: gather bits");
  hipStream_t st = (hipStream_t)stream;
  if (nb >= 10) {
    // tiles of 2^12 complex64 or 2^11 complex128 elements: 33 KiB of LDS with padding
    const int cap = prec == QX_C128 ? 11 : 12;
    GatherTileArgs t;
    if (gather_tiles(nb, w, cap, t) && t.nt == cap) {
      const int64_t ntiles = int64_t(1) << t.nr;
      const int grid = (int)std::min<int64_t>(ntiles, 1 << 16);
      if (prec == QX_C128)
        hipLaunchKernelGGL((k_gather_tiled<double, 11>), dim3(grid), dim3(256), 0, st, (const double2 *)src,
                           (double2 *)dst, ntiles, t);
      else
        hipLaunchKernelGGL((k_gather_tiled<float, 12>), dim3(grid), dim3(256), 0, st, (const float2 *)src,
                           (float2 *)dst, ntiles, t);
      check_launch("gather_tiled");
      return;
    }
  }
  GatherArgs g;
  g.nb = nb;
  for (int b = 0; b < nb; ++b) g.w[b] = w[b];
  const int64_t n = int64_t(1) << nb;
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const double2 *)src,
                       (double2 *)dst, n, g);
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, (const float2 *)src,
                       (float2 *)dst, n, g);
  check_launch("gather");
}

void launch_mfma_probe(int prec, void *out, int blocks, int iters, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_mfma_probe<double>, dim3(blocks), dim3(256), 0, s, (double *)out, iters, 0.5);
  else
    hipLaunchKernelGGL(k_mfma_probe<float>, dim3(blocks), dim3(256), 0, s, (float *)out, iters, 0.5f);
  check_launch("mfma_probe");
}

void launch_planar_merge(const float *re, const float *im, int64_t n, void *dst, void *stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_planar_merge, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, re, im, n,
                     (float2 *)dst);
  check_launch("planar_merge");
}

void launch_zero_batch(void *base, const int64_t *d_ranges, int n, void *stream) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_batch, dim3(n), dim3(256), 0, (hipStream_t)stream, (uint4 *)base, d_ranges);
  check_launch("zero_batch");
}

void launch_zero(void *p, int64_t bytes, void *stream) {
  if (bytes <= 0) return;
  if (bytes % 16) throw Error(QX_ERR_ARGUMENT, "zero fill of a size not a multiple of 16 bytes");
  int64_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n16, 256)), dim3(256), 0, (hipStream_t)stream, (uint4 *)p, n16);
  check_launch("zero");
}

void launch_gen_bits(char *bits, int64_t nb, int nch, uint64_t seed, int64_t j0, void *stream) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_gen_bits, dim3(grid_for(nb, 256)), dim3(256), 0, (hipStream_t)stream, bits, nb, nch, seed, j0);
  check_launch("gen_bits");
}

void launch_reject(const double *amps, const char *cand, int64_t nb, int nch, uint64_t seed, int64_t j0, double scale,
                   int fix_m, int64_t num_samples, double *state, int64_t *cnt, char *out_bits, double *out_amps,
                   void *stream) {
  hipLaunchKernelGGL(k_reject, dim3(1), dim3(256), 0, (hipStream_t)stream, (const double2 *)amps, cand, nb, nch, seed,
                     j0, scale, fix_m, num_samples, state, cnt, out_bits, (double2 *)out_amps);
  check_launch("reject");
}

void launch_contract_chain(int prec, const void *X0, void *Xm, void *ws, const ChainArgs &args,
                           const ChainMergeArgs &merge, void *stream) {
  const int64_t ntiles = int64_t(1) << args.nouter;
  const int g = (int)(ntiles < args.grid ? ntiles : args.grid);
  hipStream_t s = (hipStream_t)stream;
  const dim3 pack_grid(kChainMaxS / 256, args.nsteps);
  if (prec == QX_C128) {
    hipLaunchKernelGGL(k_chain_pack<double>, pack_grid, dim3(256), 0, s, (double2 *)ws, args);
    if (merge.ngroups > 0)
      hipLaunchKernelGGL(k_chain_merge<double>, dim3(merge.ngroups), dim3(256), 0, s, (double2 *)ws, merge);
    if (args.swizzled)
      hipLaunchKernelGGL((k_contract_chain<double, true>), dim3(g), dim3(kChainThreads), 0, s, (const double2 *)X0,
                         (double2 *)Xm, (const double2 *)ws, args);
    else
      hipLaunchKernelGGL((k_contract_chain<double, false>), dim3(g), dim3(kChainThreads), 0, s, (const double2 *)X0,
                         (double2 *)Xm, (const double2 *)ws, args);
  } else {
    hipLaunchKernelGGL(k_chain_pack<float>, pack_grid, dim3(256), 0, s, (float2 *)ws, args);
    if (merge.ngroups > 0)
      hipLaunchKernelGGL(k_chain_merge<float>, dim3(merge.ngroups), dim3(256), 0, s, (float2 *)ws, merge);
    if (args.swizzled)
      hipLaunchKernelGGL((k_contract_chain<float, true>), dim3(g), dim3(kChainThreads), 0, s, (const float2 *)X0,
                         (float2 *)Xm, (const float2 *)ws, args);
    else
      hipLaunchKernelGGL((k_contract_chain<float, false>), dim3(g), dim3(kChainThreads), 0, s, (const float2 *)X0,
                         (float2 *)Xm, (const float2 *)ws, args);
  }
  check_launch("contract_chain");
}

int64_t plain_gemm_units(int prec, const BlasArgs &b) {
  // C tile of a workgroup: 2 x 2 waves of PlainTile<T, GAUSS>; 64 x 64 for the split bf16 product
  const int tm = prec == QX_C128 || b.split ? 64 : 128, tn = prec == QX_C128 || prec == QX_C64_3M || b.split ? 64 : 128;
  return b.batch * ((b.r + tm - 1) / tm) * ((b.s + tn - 1) / tn);
}

void launch_plain_gemm(int prec, const void *X, const void *Y, void *C, void *ws, const BlasArgs &b, void *stream) {
  hipStream_t s = (hipStream_t)stream;
  const int64_t units = plain_gemm_units(prec, b) * b.ksplit;
  if (units > (int64_t(1) << 31) - 1) throw Error(QX_ERR_UNSUPPORTED, "plain GEMM: too many tiles");
  const dim3 g((unsigned)units);
  const int sel = b.opx * 2 + b.opy;
#define QX_PLAIN(TT, V2, OX, OY, GA) \
  hipLaunchKernelGGL((k_plain_gemm<TT, OX, OY, GA>), g, dim3(256), 0, s, (const V2 *)X, (const V2 *)Y, (V2 *)C, \
                     (V2 *)ws, b)
#define QX_PLAIN_OPS(TT, V2, GA)                   \
  switch (sel) {                                   \
    case 0: QX_PLAIN(TT, V2, 0, 0, GA); break;     \
    case 1: QX_PLAIN(TT, V2, 0, 1, GA); break;     \
    case 2: QX_PLAIN(TT, V2, 1, 0, GA); break;     \
    default: QX_PLAIN(TT, V2, 1, 1, GA); break;    \
  }
  if (b.split) {
    if (prec == QX_C128) throw Error(QX_ERR_STATE, "internal: split bf16 product of complex128 operands");
    switch (sel) {
      case 0: hipLaunchKernelGGL((k_plain_gemm_bf16x3<0, 0>), g, dim3(256), 0, s, (const float2 *)X, (const float2 *)Y, (float2 *)C, (float2 *)ws, b); break;
      case 1: hipLaunchKernelGGL((k_plain_gemm_bf16x3<0, 1>), g, dim3(256), 0, s, (const float2 *)X, (const float2 *)Y, (float2 *)C, (float2 *)ws, b); break;
      case 2: hipLaunchKernelGGL((k_plain_gemm_bf16x3<1, 0>), g, dim3(256), 0, s, (const float2 *)X, (const float2 *)Y, (float2 *)C, (float2 *)ws, b); break;
      default: hipLaunchKernelGGL((k_plain_gemm_bf16x3<1, 1>), g, dim3(256), 0, s, (const float2 *)X, (const float2 *)Y, (float2 *)C, (float2 *)ws, b); break;
    }
  } else if (prec == QX_C128) {
    if (b.gauss) {
      QX_PLAIN_OPS(double, double2, true)
    } else {
      QX_PLAIN_OPS(double, double2, false)
    }
  } else if (prec == QX_C64_3M) {
    QX_PLAIN_OPS(float, float2, true)
  } else {
    QX_PLAIN_OPS(float, float2, false)
  }
#undef QX_PLAIN_OPS
#undef QX_PLAIN
  check_launch("plain_gemm");
  if (b.ksplit > 1) {
    const int gs = grid_for(b.r * b.s * b.batch, 256);
    if (prec == QX_C128)
      hipLaunchKernelGGL(k_plain_sum<double>, dim3(gs), dim3(256), 0, s, (double2 *)C, (const double2 *)ws, b);
    else
      hipLaunchKernelGGL(k_plain_sum<float>, dim3(gs), dim3(256), 0, s, (float2 *)C, (const float2 *)ws, b);
    check_launch("plain_sum");
  }
}

void launch_small_batch(int prec, const SmallDesc *d_desc, int n, void *stream) {
  if (n <= 0) return;
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_small_batch<double>, dim3(n), dim3(kSmallThreads), 0, s, d_desc);
  else
    hipLaunchKernelGGL(k_small_batch<float>, dim3(n), dim3(kSmallThreads), 0, s, d_desc);
  check_launch("small_batch");
}

void launch_small_walk(int prec, const SmallDesc *d_desc, const int *d_level_off, int nlevels, void *stream) {
  if (nlevels <= 0) return;
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_small_walk<double>, dim3(1), dim3(kSmallThreads), 0, s, d_desc, d_level_off, nlevels);
  else
    hipLaunchKernelGGL(k_small_walk<float>, dim3(1), dim3(kSmallThreads), 0, s, d_desc, d_level_off, nlevels);
  check_launch("small_walk");
}

void launch_add_batch(int prec, const AddDesc *d_desc, int n, void *stream) {
  if (n <= 0) return;
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_add_batch<double>, dim3(n), dim3(256), 0, s, d_desc);
  else
    hipLaunchKernelGGL(k_add_batch<float>, dim3(n), dim3(256), 0, s, d_desc);
  check_launch("add_batch");
}

}  // namespace qx

// Per-call device memory of the allocation-free entry point (accumulators, staged
// bitstrings, output and index tables).  Host code never passes `scratch` itself to a
// kernel or copy: its host-side name is not a device address; the address comes from
// hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[qx::kScratchBytes];

namespace qx {

int scratch_address(int device, void **ptr) {
  static std::mutex mu;
  static std::map<int, void *> addr;   // device -> address of its copy of `scratch`
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *ptr = it->second;
    return (int)hipSuccess;
  }
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return (int)e;
  if (cur != device) return (int)hipErrorInvalidDevice;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return (int)e;
  addr[device] = p;
  *ptr = p;
  return (int)hipSuccess;
}

}  // namespace qx
