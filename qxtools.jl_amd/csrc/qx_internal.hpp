// Internal structures of libqxhip: parsed program, lowered steps, kernel launch
// interface.  Nothing here crosses the C ABI.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qxhip.h"

namespace qx {

// Internal error carrying an ABI status code; caught at the C boundary.
struct Error : std::runtime_error {
  qx_status code;
  Error(qx_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// ----------------------------------------------------------------------------
// Parsed .qx program (docs/src/users_guide.md:95-165)
enum class Op { Load, Output, View, Ncon, Save };

struct Node {
  Op op = Op::Load;
  std::string name;
  int line = 0;
  // load
  std::string key;
  std::vector<int64_t> dims;  // logical dims (load), output: {dim}
  // output
  int qubit = 0;
  // view
  std::string src, var;
  int pos = 0;  // 0-based leg
  int64_t vdim = 0;
  // ncon
  std::string a, b;
  std::vector<int> clab, alab, blab;
  // resolved inputs (node indices)
  int ia = -1, ib = -1;
  // derived
  std::vector<int64_t> shape;  // logical shape of this node's tensor
  bool dep = false;            // depends on output (bitstring) tensors
  bool sdep = false;           // depends on a slice variable
  int var_id = -1;             // view: slice variable index
};

struct Program {
  std::vector<Node> nodes;                // topological order as written
  std::map<std::string, int> by_name;     // tensor name -> node
  int save = -1;                          // node index of the saved tensor
  std::vector<std::string> vars;          // slice variables, sorted v1..vk
  std::vector<int64_t> var_dims;
  std::vector<std::string> keys;          // data keys, order of first load
  std::map<std::string, int64_t> key_elems;
  int num_chars = 0;                      // largest output qubit
  std::vector<int> outputs;               // output node indices
};

Program parse_program(const char *text, size_t len);

// ----------------------------------------------------------------------------
// Lowered form for one batch size.
enum class Buf : int8_t { None, Data, Arena, Cot, Temp, Grad };

// Where a tensor value lives and how its logical legs map to memory.
struct Value {
  Buf buf = Buf::None;
  int data_key = -1;               // Buf::Data
  int64_t slot = -1;               // Buf::Arena: element offset in the arena
  int64_t static_off = 0;          // element offset added to the base
  std::vector<std::pair<int, int64_t>> var_terms;  // (slice var, stride): offset += v*stride
  std::vector<int64_t> dims;       // logical dims (views: 1 at the sliced leg)
  std::vector<int64_t> strides;    // element stride of each logical leg
  bool dep = false;                // carries the batch leg (stride beta_stride)
  int64_t beta_stride = 0;
  int64_t elems = 0;               // elements of storage (arena values)
};

constexpr int kMaxBits = 48;   // bits of a C index or a K index
constexpr int kMaxLegs = 48;

struct BitArgs {
  int64_t wa[kMaxBits], wb[kMaxBits];   // weight of each C-index bit in A and B
  int64_t ka[kMaxBits], kb[kMaxBits];   // weight of each K-index bit in A and B
  int nbc = 0, nbk = 0;
  int conj_a = 0, conj_b = 0;
  int ksplit = 1;          // K ranges whose partials are summed after (Reduce and Bits)
  // Reduce only: skinny form (C = H x M x N with M N <= 32, N <= 2, K long): one
  // workgroup per (H element, K chunk of 2^kcb) keeps all M N sums per lane
  int skinny = 0, kcb = 0, nbm = 0, nbn = 0;
};

// Rows kernel: C[r, n] = sum_k A[r, k] B[h(r), k, n] with r over the H and M bits
// (sorted by C weight), k over K bits, n over N bits; nk = 2^nbk, nn = 2^nbn small.
struct RowArgs {
  int64_t ra[kMaxBits], rb[kMaxBits], rc[kMaxBits];   // row bits: weights in A, B (H only), C
  int64_t ka[64], kb[64];                             // per k value (not per bit)
  int64_t nb[64], nc[64];                             // per n value
  int nbr = 0, nk = 1, nn = 1;
  int conj_a = 0, conj_b = 0;
  int hlo = 0;                 // row bits among the low 8 that are H bits (bit mask)
};

struct RadixArgs {
  int64_t cd[kMaxLegs], ca[kMaxLegs], cb[kMaxLegs];
  int64_t kd[kMaxLegs], ka[kMaxLegs], kb[kMaxLegs];
  int ncl = 0, nkl = 0;
  int conj_a = 0, conj_b = 0;
};

// Small slice-free steps (gate-block building) run many per launch: one workgroup
// per step, its bit weights read from a device table of these descriptors.
constexpr int kSmallMaxBits = 18;
constexpr int kSmallChunk = 1024;   // most C elements per workgroup of a small-step batch
constexpr int kSmallThreads = 256;  // lanes per workgroup of a small-step batch
struct SmallDesc {
  const void *A = nullptr, *B = nullptr;
  void *C = nullptr;
  int32_t nbc = 0, nbk = 0, conj_a = 0, conj_b = 0;
  int32_t c0 = 0, c1 = 0;      // C elements [c0, c1) of this workgroup (a large C is split)
  int32_t wa[kSmallMaxBits] = {0}, wb[kSmallMaxBits] = {0};   // C bits
  int32_t ka[kSmallMaxBits] = {0}, kb[kSmallMaxBits] = {0};   // K bits
};

// Batched scatter-add of small temps into FP64 gradients or cotangent buffers:
// dst[map(e)] += src[e], map(e) = sum of digit_j(e) * st[j] over the dims (first
// fastest).
struct AddDesc {
  void *dst = nullptr;
  const void *src = nullptr;
  int64_t n = 0;
  int32_t nd = 0;
  int32_t dst_f64 = 1;   // destination FP64 (a gradient); else the plan's element type
  int64_t dim[kSmallMaxBits] = {0}, st[kSmallMaxBits] = {0};
};

// A chain of steps that each contract a small operand (a gate block: at most 2^10
// elements) into one large tensor X, X_j = S_j . X_{j-1}, run as one pass over X:
// every workgroup loads a tile of 2^kChainTileBits elements of X_0 (the bits of
// the legs the chain touches plus the fastest untouched legs), applies every S_j
// to it in LDS and stores the tile of X_m.  The intermediates never reach memory.
constexpr int kChainTileBits = 13;
constexpr int kChainMaxSteps = 32;
constexpr int kChainMinPassive = 3;   // untouched fast legs kept in the tile
constexpr int kChainMaxS = 1024;      // elements of one staged small operand
constexpr int kChainSLdsBytes = 32 << 10;   // LDS for the packed gate blocks (beside the tile)
// Per-step image tables in the workspace after the packed gate blocks: kChainTabInts
// int32 per step (LDS images of the step's group and contracted bits), reserved as
// kChainTabElems workspace elements per step (128 bytes: 16 complex64 or 8 complex128).
constexpr int kChainTabInts = 32;
constexpr int kChainTabElems = 16;

// The fields the chain kernel reads are 32-bit: it indexes steps with a uniform loop
// counter, and dword-aligned fields load through the scalar unit (narrower ones would
// take vector loads whose waits also drain the next tile's prefetch).
struct ChainStepArgs {
  int32_t k = 0, h = 0;         // contracted legs (each replaced by a new leg), batch legs
  int32_t q[5] = {0}, hq[3] = {0};   // tile bit of each contracted / batch leg
  int32_t soff = 0;             // offset of S packed as [h][nu][kappa] in the workspace
  int16_t sk[5] = {0}, sn[5] = {0}, sh[3] = {0};   // strides in S of contracted, new, batch legs
  int32_t merged = 0;           // 1: S is the product of several gate blocks (k_chain_merge)
};

// Gate blocks of a chain merged into one step (several consecutive steps, or steps
// moved past steps on other tile bits, whose contracted bits together are at most
// four): the merged step's S is the product of its members' blocks, each acting on
// some of the merged step's bits (identity on the others), first member rightmost.
constexpr int kChainMaxMerged = 32;   // members of merged steps in one chain
struct ChainMergeArgs {
  int ngroups = 0;
  struct Group {
    int32_t step = 0, k = 0, first = 0, count = 0;   // merged step index, its bits, members
    int32_t soff = 0;                                 // its packed block in the workspace
  } g[kChainMaxSteps];
  struct Member {
    const void *S = nullptr;   // the member's gate block, set at launch
    int32_t kc = 0;            // its contracted legs
    int32_t pos[4] = {0};      // index of each among the merged step's bits
    int32_t sk[4] = {0}, sn[4] = {0};   // strides in S of its contracted and new legs
  } m[kChainMaxMerged];
};
static_assert(sizeof(ChainMergeArgs) <= 4096, "merge arguments must fit the kernel argument segment");

struct ChainArgs {
  int nsteps = 0, nouter = 0;
  int grid = 256;          // workgroups (each walks tiles); tuned
  int mfma16 = 0;          // steps with 16 contracted values: 0 vector FMAs, 1 matrix cores,
                           // 2 matrix cores with three real products per complex product
  int s_elems = 0;         // elements of all packed S_j (staged in LDS when they fit)
  int tab_off = 0;         // workspace element offset of the per-step image tables
  // wave-local chains (every step has k + h <= 4): the wave index (group bits 6-8 of
  // every step) sits on three tile bits no step touches, so each wave keeps to its own
  // tile elements and the steps need no workgroup barriers
  int wave_local = 0;
  int32_t wbits[3] = {0, 0, 0};
  int64_t win[kChainTileBits] = {0}, wout[kChainTileBits] = {0};   // tile bit -> X_0 / X_m stride
  int32_t pout[kChainTileBits] = {0};  // store order: i-th fastest X_m tile bit -> tile bit
  // LDS address of tile bit b: element i sits at the XOR of swz[b] over its set bits.
  // swz[b] = 2^b plus, for some b above the bank-group bits, a pattern in those bits.
  int32_t swz[kChainTileBits] = {0};
  int swizzled = 0;        // 1: LDS addresses through swz (else plain; swz unused); tuned
  int64_t oin[kMaxBits] = {0}, oout[kMaxBits] = {0};   // outer bits (workgroup index) -> strides
  ChainStepArgs st[kChainMaxSteps];
  const void *S[kChainMaxSteps] = {nullptr};   // small operands, set at launch
};
static_assert(sizeof(ChainArgs) <= 4096, "chain arguments must fit the kernel argument segment");

// C[h, m, n] = sum_k A[h, m, k] B[h, k, n], every index mapped through bit
// weights (all dims powers of two).  m bits are listed in increasing C weight,
// likewise n bits; k bits in increasing A weight.
struct GemmArgs {
  int64_t ha[kMaxBits], hb[kMaxBits], hc[kMaxBits];
  int64_t ma[kMaxBits], mc[kMaxBits];
  int64_t nb[kMaxBits], nc[kMaxBits];
  int64_t ka[kMaxBits], kb[kMaxBits];
  int nbh = 0, nbm = 0, nbn = 0, nbk = 0;
  int conj_a = 0, conj_b = 0;
  int a_kfast = 0, b_kfast = 0, c_nfast = 0;  // coalescing direction of tile loads/stores
  int tile = 0;                               // tile configuration index
  int ksplit = 1;                             // K split over workgroups (partials summed after)
  int64_t c_elems = 0;                        // dense C size: stride between K-split partials
};

// One leg of a step in canonical form: dims and element strides in A, B and C.
struct Leg {
  int64_t dim;
  int64_t sa, sb, sc;
};

// A step whose operands are plain strided-batched column-major matrices, run by the
// vendor BLAS: C (R x S, ldc) = op(X) (R x K) . op(Y) (K x S), batch strides per
// operand.  op: 0 = as stored, 1 = transposed.  swap: X/Y are B/A (C stored
// N-fastest).
struct BlasArgs {
  int opx = 0, opy = 0, swap = 0;
  int64_t r = 0, s = 0, k = 0, ldx = 0, ldy = 0, ldc = 0;
  int64_t sx = 0, sy = 0, sc = 0, batch = 1;
  // QX_C64_BF16X3 on steps with r, s, k >= 1024: each complex64 operand is split into
  // bf16 hi + lo parts (planar re / im, in the operand's own layout, over its extent
  // ex / ey) and C = Xh Yh + Xh Yl + Xl Yh by bf16 GEMMs accumulating in FP32 (planar,
  // extent ec), then interleaved into C.  0: one complex64 GEMM.
  int split = 0;
  int64_t ex = 0, ey = 0, ec = 0;
  // ttgt: the step's operands are not plain; A and B are first gathered into plain
  // [M][K][H] / [K][N][H] copies and C from a plain [M][N][H] result (weights of the
  // copies' bits in the source, dense order), after the split planes in the workspace
  int ttgt = 0, nga = 0, ngb = 0, ngc = 0;
  int64_t ga[kMaxBits] = {0}, gb[kMaxBits] = {0}, gc[kMaxBits] = {0};
  int64_t ta = 0, tb = 0;   // elements of the A and B copies (the result copy: r s batch)
  // own = 1: the hand-written matrix-core GEMM (k_plain_gemm) instead of the vendor
  // BLAS; its K split over ksplit workgroups with dense partials in the workspace;
  // gauss = 1: complex128 as three real products per complex product (complex64 takes
  // the form from its precision)
  int own = 0, ksplit = 1, gauss = 0;
};

// Runs of outer products (steps with no contracted leg, each output feeding only the
// next step of the run): C = F_0 x F_1 x ... over the factors the run reads, legs two
// factors share (the bitstring batch leg) multiplied elementwise.  Output element e
// (C dense, e its offset) is the product of F_f[sum over the set bits b of e of the
// factor's stride for bit b], taken in the run's order: v = F_0, then for each later
// factor v = v F_f (the running value was that step's A operand) or F_f v.
constexpr int kOuterMaxFactors = 16, kOuterMaxTerms = 192;
constexpr int kOuterBits = 48;   // output bits an outer-product run's terms may name
struct OuterArgs {
  int nf = 0;                            // factors
  int members = 0;                       // steps of the run
  int first[kOuterMaxFactors + 1] = {0}; // factor f's terms: [first[f], first[f + 1])
  int8_t prev_a[kOuterMaxFactors] = {0}; // factor f >= 1: the running value was its step's A
  int8_t bit[kOuterMaxTerms] = {0};      // output bit of each term
  int64_t w[kOuterMaxTerms] = {0};       // the factor's stride for that bit
  const void *F[kOuterMaxFactors] = {nullptr};   // set at launch
};
static_assert(sizeof(OuterArgs) <= 4096, "outer-product arguments must fit the kernel argument segment");

enum class Kern : int8_t { Bits, Radix, Gemm, Reduce, Rows, Chain, Blas, Outer };

struct Step {
  int node = -1;           // ncon node
  int in_a = -1, in_b = -1;
  std::vector<Leg> cl;     // C legs in C's memory order (fastest first)
  std::vector<Leg> kl;     // contracted legs
  int64_t H = 1, M = 1, N = 1, K = 1;  // class products (beta counted in its class)
  int64_t c_elems = 1;
  int64_t ws_elems = 0;    // workspace for K-split partials
  Kern kern = Kern::Bits;
  int section = 0;         // 0 shared slice-free, 1 shared per slice, 2 per bitstring batch
  int level = -1;          // >= 0: small slice-free step run with the others of its level in one launch
  bool swapped = false;    // operands exchanged (in_a is the DSL's B): C = B.A, M and N traded
  BitArgs bits;            // Kern::Bits / Kern::Reduce
  GemmArgs gemm;           // Kern::Gemm
  RowArgs rows;            // Kern::Rows
  RadixArgs radix;         // Kern::Radix
  BlasArgs blas;           // Kern::Blas
  ChainArgs chain;         // Kern::Chain: in_a = X_0, node = X_m
  std::vector<int> chain_small;   // Kern::Chain: node of each S_j (a merged step: its first member's)
  struct ChainMember {
    int step = 0, small = -1, kc = 0;   // merged step, gate block node, its contracted legs
    int pos[4] = {0}, sk[4] = {0}, sn[4] = {0};
  };
  std::vector<ChainMember> chain_merge;   // Kern::Chain: members of merged steps, in order
  double chain_flops = 0, chain_elems = 0;   // Kern::Chain / Outer: work and compulsory elements moved
  OuterArgs outer;                 // Kern::Outer: in_a, in_b = its first two factors
  std::vector<int> outer_in;       // Kern::Outer: node of each factor
};

struct ArenaSlot {
  int node;
  int64_t offset, elems, start, end;   // element offset and size, live [start, end)
};

struct Lowered {
  int64_t batch = 0;
  uint64_t fused = 0;               // slice variables contracted inline (bit j = variable j)
  bool tuned = false;               // kernel choices measured by the autotuner
  std::vector<std::string> tune_log;  // per step: mean ms of each candidate (JSON members)
  std::vector<ArenaSlot> slots;     // arena reuse schedule
  std::vector<Value> values;        // per node
  std::vector<Step> steps;          // ncon nodes in order
  int64_t arena_elems = 0;          // forward arena size (elements)
  int64_t ws_elems = 0;             // largest K-split workspace of any step
  int64_t out_elems = 0;            // elements of one output tensor incl. batch
  int64_t result_elems = 1;         // complex values per bitstring
};

// batch_last: lay the batched bitstring leg out slowest in bitstring-dependent
// intermediates (default fastest); outputs keep it fastest.
Lowered lower(const Program &p, int64_t batch, bool keep_all = false, int precision = QX_C128,
              uint64_t fused = 0, bool chains = true, bool batch_last = false);

// C[clab] = sum A[alab] B[blab] over the given value layouts (C dense), kernel
// chosen and arguments filled; conj flags conjugate an operand on load.
Step make_step(const Value &A, const std::vector<int> &alab, bool conj_a, const Value &B,
               const std::vector<int> &blab, bool conj_b, const Value &C, const std::vector<int> &clab,
               int64_t batch, int precision = QX_C128);

// Reverse-mode plan of one batch size (qx_amplitudes_vjp).  The forward keeps
// every value of a slice; each forward step C = A.B contributes Abar = Cbar.conj(B)
// and Bbar = conj(A).Cbar as ordinary steps.  A cotangent goes straight into its
// value's cotangent buffer when its consumer runs in the same loop section,
// otherwise through a temporary that is added (through the value's view offsets
// and strides) into the cotangent of what it views, or into the FP64 gradient of
// a data key.
struct BwdStep {
  Step step;                  // contraction; its C is the cotangent buffer or the temp
  int x = -1;                 // forward node whose cotangent this computes
  bool direct = false;
  Buf dest = Buf::None;       // Cot (cotangent of an ncon value) or Grad (data key)
  int dest_root = -1;         // node (Cot) whose buffer receives the add
  int dest_key = -1;          // data key (Grad)
  std::vector<int64_t> add_dims, add_strides;   // temp (dense) -> destination
  int64_t add_elems = 0;
  int level = -1;             // >= 0: runs batched with the other small steps of its level
  int64_t temp_off = 0;       // offset (elements) of its temp in the temp buffer
};

struct VjpLowered {
  Lowered fwd;                          // forward lowering, every value kept
  std::vector<Value> cot;               // per node: cotangent buffer of ncon values
  std::vector<std::vector<BwdStep>> bsteps;   // per section, execution order
  std::vector<std::vector<int>> zero_at;      // per section: accumulating cotangents
  int64_t cot_elems = 0, temp_elems = 0;
  int root = -1;
};

VjpLowered lower_vjp(const Program &p, int64_t batch, int precision = QX_C128, uint64_t fused = 0,
                     bool batch_last = false);

// First-fit offsets for [start, end) intervals of given sizes (elements), 256-byte
// aligned for element size `esize`.  Returns the total size; offsets written out.
int64_t first_fit(const std::vector<int64_t> &start, const std::vector<int64_t> &end,
                  const std::vector<int64_t> &size, int esize, std::vector<int64_t> &off);

// ----------------------------------------------------------------------------
// Kernel launch interface (kernels.hip).
void set_kernel_args(Step &st, int precision = QX_C128);
// Bytes that one LDS access cycle serves for 16-byte accesses: lanes whose addresses
// differ only above this many bytes collide on the same banks (measured on gfx950 with
// SQ_LDS_BANK_CONFLICT: a swizzle built for 256 bytes leaves no conflicts where one
// for 128 bytes leaves some).
constexpr int kLdsCycleBytes = 256;

// Choose the chain tile's LDS swizzle (ChainArgs::swz) for elements of `elem_bytes`.
void chain_swizzle(ChainArgs &a, int elem_bytes);

// Run `st` with kernel configuration `choice` (0 gather, 1 rows, 2-4 FP tile GEMMs,
// 5 / 7 matrix-core GEMM, 6 reduce, 8 vendor BLAS on plain matrix layouts); false
// (and `st` unchanged) if that kernel cannot.
bool force_kernel(Step &st, int precision, int choice);
// Alternative kernel configurations of a step (the current one first) for the autotuner.
std::vector<Step> kernel_candidates(const Step &st, int precision);
void launch_contract_bits(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                          const BitArgs &args, void *stream);
// With args.ksplit > 1, `ws` receives ksplit partials of C which are then summed
// into C.
void launch_contract_gemm(int prec, const void *A, const void *B, void *C, void *ws, const GemmArgs &args,
                          void *stream);
void launch_contract_rows(int prec, const void *A, const void *B, void *C, const RowArgs &args,
                          void *stream);
// few C elements, long K: one wave per C element, K split over lanes, shuffle reduce
void launch_contract_reduce(int prec, const void *A, const void *B, void *C, void *ws, int64_t c_elems,
                            const BitArgs &args, void *stream);
void launch_contract_radix(int prec, const void *A, const void *B, void *C, int64_t c_elems,
                           const RadixArgs &args, void *stream);
void launch_fill_outputs(int prec, void *arena, const int64_t *d_out_off, int n_out,
                         const int *d_out_q, const char *d_bits, int64_t nvalid, int64_t batch,
                         int64_t num_chars, void *stream);
void launch_accumulate(int prec, const void *root, int64_t root_bstride, int64_t batch, int64_t nvalid,
                       int64_t relems, const int64_t *d_relem_off, double *acc, int64_t row0,
                       void *stream);
void launch_finalize(const double *acc, double *out, int64_t n, void *stream);
// dst[map(e)] += src[e] for src dense over dims (first fastest), map = sum digit*stride
void launch_add_mapped(int prec, int dst_double, void *dst, const void *src, int64_t n,
                       const std::vector<int64_t> &dims, const std::vector<int64_t> &strides,
                       void *stream);
// cotangent of the saved tensor from amp_cot rows [row0, row0 + nvalid)
void launch_root_cot(int prec, void *dst, int64_t dep, int64_t batch, int64_t nvalid, int64_t relems,
                     const int64_t *d_relem_off, const double *amp, int64_t row0, void *stream);
// root cotangent 2 * root value (the loss sum |amp|^2); amplitudes of the batch into `amp`
void launch_root_cot_norm(int prec, void *dst, const void *val, int64_t dep, int64_t batch, int64_t nvalid,
                          int64_t relems, const int64_t *d_relem_off, double *amp, int64_t row0, void *stream);
void launch_small_batch(int prec, const SmallDesc *d_desc, int n, void *stream);
// Levels [0, nlevels) of small-step chunks in one workgroup: level L is the descriptors
// [d_level_off[L], d_level_off[L + 1]), run after every earlier level.
void launch_small_walk(int prec, const SmallDesc *d_desc, const int *d_level_off, int nlevels, void *stream);
void launch_add_batch(int prec, const AddDesc *d_desc, int n, void *stream);
// bf16 split of n complex64 values: out holds four planar n-element bf16 arrays
// (re hi, re lo, im hi, im lo); hi = x rounded to bf16 (nearest, ties to even),
// lo = (x - hi) rounded likewise.
void launch_bf16_split(const void *src, int64_t n, void *out, void *stream);
// dst[i] = src[sum of bit_j(i) * w[j]] for i < 2^nb (a permuted dense copy).
void launch_gather(int prec, const void *src, void *dst, int nb, const int64_t *w, void *stream);
// dst[i] = (re[i], im[i]) as complex64, i < n.
void launch_planar_merge(const float *re, const float *im, int64_t n, void *dst, void *stream);
// d_ranges: n (offset, length) pairs in 16-byte units from base; one workgroup each.
void launch_zero_batch(void *base, const int64_t *d_ranges, int n, void *stream);
// One chain pass: k_chain_pack (gate blocks and step tables into the workspace), for
// merged steps k_chain_merge (their members' product), then k_contract_chain.
void launch_outer(int prec, void *C, int64_t c_elems, const OuterArgs &args, void *stream);
void launch_contract_chain(int prec, const void *X0, void *Xm, void *ws, const ChainArgs &args,
                           const ChainMergeArgs &merge, void *stream);
// C = op(X) op(Y) by the hand-written matrix-core GEMM (b.own): plain column-major
// operands as for the vendor BLAS; with b.ksplit > 1, `ws` holds ksplit * batch * r * s
// elements of partials.  plain_gemm_units: C tiles of one K range.
void launch_plain_gemm(int prec, const void *X, const void *Y, void *C, void *ws, const BlasArgs &b, void *stream);
int64_t plain_gemm_units(int prec, const BlasArgs &b);
// Output sampling: candidate bitstrings from Philox4x64-10 and the sequential
// empirical-supremum rejection test over one round of candidates.
void launch_gen_bits(char *bits, int64_t nb, int nch, uint64_t seed, int64_t j0, void *stream);
// `blocks` workgroups of 256 lanes, each lane `iters` x 4 matrix-core products; writes
// blocks * 256 values of the precision's real type to `out`
void launch_mfma_probe(int prec, void *out, int blocks, int iters, void *stream);
void launch_reject(const double *amps, const char *cand, int64_t nb, int nch, uint64_t seed, int64_t j0, double scale,
                   int fix_m, int64_t num_samples, double *state, int64_t *cnt, char *out_bits, double *out_amps,
                   void *stream);
void launch_zero(void *p, int64_t bytes, void *stream);

// The library's fixed device scratch area (a file-scope device array of kScratchBytes,
// one copy per device): its device address on `device`, looked up once per device
// with that device current and kept.  Returns the HIP status of the lookup.
constexpr size_t kScratchBytes = size_t(8) << 20;
int scratch_address(int device, void **ptr);

}  // namespace qx
