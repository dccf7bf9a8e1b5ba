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
  const int64_t nk_hi = int64_t(1) << (a.nbk - nk_lo);
  const int64_t tile = int64_t(1) << kTileBits;
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
  for (int64_t base = (int64_t)blockIdx.x * tile; base < c_elems; base += (int64_t)gridDim.x * tile) {
    int64_t hia = 0, hib = 0;
    for (int b = kTileBits; b < a.nbc; ++b)
      if ((base >> b) & 1) {
        hia += a.wa[b];
        hib += a.wb[b];
      }
    V acc[kElems];
#pragma unroll
    for (int e = 0; e < kElems; ++e) acc[e] = V{0, 0};
    for (int64_t kh = 0; kh < nk_hi; ++kh) {
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
      if (idx < c_elems) C[idx] = acc[e];
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
__global__ void k_accumulate(const typename Vec<T>::type *__restrict__ root, int root_dep, int64_t batch,
                             int64_t nvalid, int64_t relems, const int64_t *__restrict__ relem_off,
                             double *__restrict__ acc, int64_t row0) {
  int64_t total = nvalid * relems;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t beta = t % nvalid;
    int64_t e = t / nvalid;
    auto v = root[relem_off[e] + (root_dep ? beta : 0)];
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

}  // namespace

void launch_contract_bits(int prec, const void *A, const void *B, void *C, int64_t c_elems, const BitArgs &args,
                          void *stream) {
  int g = grid_for(c_elems, kThreads * kElems);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_contract_bits<double>, dim3(g), dim3(kThreads), 0, s, (const double2 *)A,
                       (const double2 *)B, (double2 *)C, c_elems, args);
  else
    hipLaunchKernelGGL(k_contract_bits<float>, dim3(g), dim3(kThreads), 0, s, (const float2 *)A, (const float2 *)B,
                       (float2 *)C, c_elems, args);
  check_launch("contract_bits");
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

void launch_accumulate(int prec, const void *root, int root_dep, int64_t batch, int64_t nvalid, int64_t relems,
                       const int64_t *d_relem_off, double *acc, int64_t row0, void *stream) {
  int g = grid_for(nvalid * relems, 256);
  hipStream_t s = (hipStream_t)stream;
  if (prec == QX_C128)
    hipLaunchKernelGGL(k_accumulate<double>, dim3(g), dim3(256), 0, s, (const double2 *)root, root_dep, batch,
                       nvalid, relems, d_relem_off, acc, row0);
  else
    hipLaunchKernelGGL(k_accumulate<float>, dim3(g), dim3(256), 0, s, (const float2 *)root, root_dep, batch,
                       nvalid, relems, d_relem_off, acc, row0);
  check_launch("accumulate");
}

void launch_finalize(const double *acc, double *out, int64_t n, void *stream) {
  hipLaunchKernelGGL(k_finalize, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, acc, out, n);
  check_launch("finalize");
}

void launch_zero(void *p, int64_t bytes, void *stream) {
  if (bytes <= 0) return;
  if (bytes % 16) throw Error(QX_ERR_ARGUMENT, "zero fill of a size not a multiple of 16 bytes");
  int64_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n16, 256)), dim3(256), 0, (hipStream_t)stream, (uint4 *)p, n16);
  check_launch("zero");
}

}  // namespace qx
