// tiles.hpp — 16x16 complex FP64 register blocks and FP64-MFMA products on gfx950.
//
// A d x d complex matrix (d <= 16*B) is held by a workgroup of B*B wavefronts;
// wavefront w = I*B + J owns block (I, J). Inside a wavefront, lane l = c + 16*q
// holds in register r the element (row q + 4r, column c) of its block: this is
// the result layout of v_mfma_f64_16x16x4_f64 on gfx950 (measured by an
// in-bounds lane-layout check), so products chain without reshuffling: the
// k-step kb of a product sums over inner index k = q + 4*kb, which makes the
// B operand of step kb exactly result register kb. Operands that must be read
// in another layout go through LDS as plain 16 x 17 (padded) complex tiles.
//
// Per output block, a product C = op(A) * op(B) runs, for each k-block K and
// k-step kb, three real MFMAs (Gauss's complex product).
#pragma once
#include <hip/hip_runtime.h>

#ifndef PR_MFMA_4M
#define PR_MFMA_4M 0
#endif

namespace pr {

typedef double d4 __attribute__((ext_vector_type(4)));

struct Blk {
  d4 re, im;
};

// none, transpose, conjugate transpose, conjugate
enum Op { N = 0, T = 1, C = 2, J = 3 };

struct Ln {
  int lane, q, c, w, I, Jb;
};

template <int B>
__device__ inline Ln lane_info() {
  Ln x;
  x.lane = threadIdx.x & 63;
  x.q = x.lane >> 4;
  x.c = x.lane & 15;
  x.w = threadIdx.x >> 6;
  x.I = x.w / B;
  x.Jb = x.w % B;
  return x;
}

// row of the block held by lane-group q in register (or k-step) r
__device__ constexpr int rowof(int q, int r) { return q + 4 * r; }

__device__ inline d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// (-a) b + c (the instruction's negate modifier on the A input)
__device__ inline d4 mfma_nega(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
}

__device__ inline Blk zero_blk() {
  Blk b;
  b.re = d4{0, 0, 0, 0};
  b.im = d4{0, 0, 0, 0};
  return b;
}

// identity restricted to the wavefront's block
__device__ inline Blk eye_blk(const Ln& x) {
  Blk b = zero_blk();
  if (x.I == x.Jb) {
#pragma unroll
    for (int r = 0; r < 4; r++) b.re[r] = (rowof(x.q, r) == x.c) ? 1.0 : 0.0;
  }
  return b;
}

// Pointers held in structs are generic; these accessors let the compiler emit
// global (not flat) memory instructions for the device buffers.
#define PR_GLOBAL __attribute__((address_space(1)))
typedef double dv2 __attribute__((ext_vector_type(2)));

template <class T>
__device__ inline const PR_GLOBAL T* gptr(const T* p) {
  return (const PR_GLOBAL T*)p;
}
template <class T>
__device__ inline PR_GLOBAL T* gptr_w(T* p) {
  return (PR_GLOBAL T*)p;
}
__device__ inline double2 gload2(const double2* p, size_t i) {
  dv2 v = ((const PR_GLOBAL dv2*)p)[i];
  return make_double2(v.x, v.y);
}
__device__ inline void gstore2(double2* p, size_t i, double a, double b) {
  dv2 v;
  v.x = a;
  v.y = b;
  ((PR_GLOBAL dv2*)p)[i] = v;
}

// Global tile image: 256 complex per block, register-major: [r][lane].
__device__ inline Blk load_blk(const double2* __restrict__ img, const Ln& x) {
  Blk b;
  const size_t base = (size_t)x.w * 256;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    double2 v = gload2(img, base + r * 64 + x.lane);
    b.re[r] = v.x;
    b.im[r] = v.y;
  }
  return b;
}

// One 16 x 16 image (row-major, element [i][j] at 16 i + j) loaded transposed:
// lane (c, q) takes [c][q + 4r] into register r, so the block holds the image's
// transpose (an image of U^T gives U, an A operand of op C then reads U^dag).
// Each load is sixteen 64-byte runs instead of one 1 KiB run.
__device__ inline Blk load_blk_t(const double2* __restrict__ img, const Ln& x) {
  Blk b;
  const int base = x.c * 16 + x.q;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    double2 v = gload2(img, base + 4 * r);
    b.re[r] = v.x;
    b.im[r] = v.y;
  }
  return b;
}

__device__ inline void store_blk(double2* __restrict__ img, const Blk& b, const Ln& x) {
  const size_t base = (size_t)x.w * 256;
#pragma unroll
  for (int r = 0; r < 4; r++) gstore2(img, base + r * 64 + x.lane, b.re[r], b.im[r]);
}

// The per-slice propagator images (written once by the segment kernel, read
// once by the gradient walk, gigabytes per evaluation) go through the caches
// non-temporally, so that they do not evict the checkpoints, controls and
// gradient rows the kernels reuse.
__device__ inline Blk load_blk_nt(const double2* __restrict__ img, const Ln& x) {
  Blk b;
  const size_t base = (size_t)x.w * 256;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const dv2 v = __builtin_nontemporal_load((const PR_GLOBAL dv2*)img + base + r * 64 + x.lane);
    b.re[r] = v.x;
    b.im[r] = v.y;
  }
  return b;
}
__device__ inline void store_blk_nt(double2* __restrict__ img, const Blk& b, const Ln& x) {
  const size_t base = (size_t)x.w * 256;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    dv2 v;
    v.x = b.re[r];
    v.y = b.im[r];
    __builtin_nontemporal_store(v, (PR_GLOBAL dv2*)img + base + r * 64 + x.lane);
  }
}

// LDS tile: block blk, row, col -> padded index
__device__ inline int lidx(int blk, int row, int col) { return blk * 272 + row * 17 + col; }

template <int B>
__device__ inline void put(double2* L, const Blk& b, const Ln& x) {
#pragma unroll
  for (int r = 0; r < 4; r++) L[lidx(x.w, rowof(x.q, r), x.c)] = make_double2(b.re[r], b.im[r]);
}

// A-operand element op(A)[16I + c][16K + q + 4kb] for the MFMA k-step kb.
// RAW: the element before conjugation (the caller applies the sign of the
// imaginary part, e.g. as an MFMA negate modifier).
template <int B, int OP, bool RAW = false>
__device__ inline double2 fetch_a(const double2* L, const Blk& own, int K, int kb, const Ln& x) {
  if constexpr (B == 1 && (OP == T || OP == C)) {
    double2 v = make_double2(own.re[kb], own.im[kb]);
    if constexpr (OP == C && !RAW) v.y = -v.y;
    return v;
  } else {
    double2 v;
    if constexpr (OP == N || OP == J)
      v = L[lidx(x.I * B + K, x.c, rowof(x.q, kb))];
    else
      v = L[lidx(K * B + x.I, rowof(x.q, kb), x.c)];
    if constexpr ((OP == C || OP == J) && !RAW) v.y = -v.y;
    return v;
  }
}

// B-operand element op(B)[16K + q + 4kb][16J + c].
template <int B, int OP, bool RAW = false>
__device__ inline double2 fetch_b(const double2* L, const Blk& own, int K, int kb, const Ln& x) {
  if constexpr (B == 1 && (OP == N || OP == J)) {
    double2 v = make_double2(own.re[kb], own.im[kb]);
    if constexpr (OP == J && !RAW) v.y = -v.y;
    return v;
  } else {
    double2 v;
    if constexpr (OP == N || OP == J)
      v = L[lidx(K * B + x.Jb, rowof(x.q, kb), x.c)];
    else
      v = L[lidx(x.Jb * B + K, x.c, rowof(x.q, kb))];
    if constexpr ((OP == C || OP == J) && !RAW) v.y = -v.y;
    return v;
  }
}

template <int B, int OP>
__device__ constexpr bool a_in_lds() {
  return !(B == 1 && (OP == T || OP == C));
}
template <int B, int OP>
__device__ constexpr bool b_in_lds() {
  return !(B == 1 && (OP == N || OP == J));
}

// Workgroup synchronisation of the LDS tiles. A single-wavefront workgroup
// (B == 1) only needs the wave's own LDS accesses ordered, which LDS does in
// issue order; the fences keep the compiler from moving them.
template <int B>
__device__ inline void bsync() {
  if constexpr (B == 1) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Block product op(A) op(B) for this wavefront's output block with three real
// products per k-step (Gauss): t1 = Ar Br, t2 = Ai Bi, t3 = (Ar + Ai)(Br + Bi);
// Re = t1 - t2, Im = t3 - t1 - t2. 12 MFMAs per k-block instead of 16.
template <int B, int OA, int OB>
__device__ inline Blk gemm_prod(const double2* LA, const Blk& Aown, const double2* LB, const Blk& Bown,
                                const Ln& x) {
#if PR_MFMA_4M
  // four real products per k-step: Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br
  Blk r = zero_blk();
#pragma unroll
  for (int K = 0; K < B; K++) {
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      double2 a = fetch_a<B, OA>(LA, Aown, K, kb, x);
      double2 b = fetch_b<B, OB>(LB, Bown, K, kb, x);
      r.re = mfma(a.x, b.x, r.re);
      r.im = mfma(a.x, b.y, r.im);
      r.re = mfma(-a.y, b.y, r.re);
      r.im = mfma(a.y, b.x, r.im);
    }
  }
  return r;
#else
  // conjugated operands: the imaginary sign goes into the MFMA negate
  // modifier and the (free) source negation of the sums, not into VALU
  constexpr bool na = OA == C || OA == J, nb = OB == C || OB == J;
  d4 t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
  for (int K = 0; K < B; K++) {
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const double2 a = fetch_a<B, OA, true>(LA, Aown, K, kb, x);
      const double2 b = fetch_b<B, OB, true>(LB, Bown, K, kb, x);
      t1 = mfma(a.x, b.x, t1);
      t2 = (na != nb) ? mfma_nega(a.y, b.y, t2) : mfma(a.y, b.y, t2);
      t3 = mfma(na ? a.x - a.y : a.x + a.y, nb ? b.x - b.y : b.x + b.y, t3);
    }
  }
  Blk r;
  r.re = t1 - t2;
  r.im = t3 - t1 - t2;
  return r;
#endif
}

// C = op(A) op(B). LA/LB are this workgroup's LDS scratch tiles.
template <int B, int OA, int OB>
__device__ inline Blk mm(const Blk& A, const Blk& Bm, double2* LA, double2* LB, const Ln& x) {
  constexpr bool la = a_in_lds<B, OA>(), lb = b_in_lds<B, OB>();
  if constexpr (la) put<B>(LA, A, x);
  if constexpr (lb) put<B>(LB, Bm, x);
  if constexpr (la || lb) bsync<B>();
  Blk r = gemm_prod<B, OA, OB>(LA, A, LB, Bm, x);
  if constexpr (la || lb) bsync<B>();
  return r;
}

__device__ inline Blk axpby(double a, const Blk& X, double b, const Blk& Y) {
  Blk r;
  r.re = a * X.re + b * Y.re;
  r.im = a * X.im + b * Y.im;
  return r;
}
__device__ inline Blk conj_blk(const Blk& X) {
  Blk r;
  r.re = X.re;
  r.im = -X.im;
  return r;
}
// (ar + i ai) * X
__device__ inline Blk cscale(double ar, double ai, const Blk& X) {
  Blk r;
  r.re = ar * X.re - ai * X.im;
  r.im = ar * X.im + ai * X.re;
  return r;
}

// Cross-lane steps on VALU (no LDS round trip). sum_rows_pairwise: every lane
// gets (x_0 + x_1) + (x_2 + x_3) over the four 16-lane rows at its row offset,
// the same sums as xor-16 then xor-32 exchanges. v_permlane16/32_swap on two
// copies of v leave the lane's pair in the two outputs, whichever is which.
__device__ inline double pair_sum16(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
__device__ inline double pair_sum32(double v) {
  const unsigned lo = __double2loint(v), hi = __double2hiint(v);
  const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
__device__ inline double sum_rows_pairwise(double v) { return pair_sum32(pair_sum16(v)); }
template <int CTRL>
__device__ inline double dpp_d(double v) {
  // every pattern used reads a lane of the same row, so no lane keeps "old"
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
// maximum over the 16 lanes of each row: quad swaps, then half-row and row mirrors
__device__ inline double max_in_row(double v) {
  v = fmax(v, dpp_d<0xB1>(v));   // quad_perm [1,0,3,2]
  v = fmax(v, dpp_d<0x4E>(v));   // quad_perm [2,3,0,1]
  v = fmax(v, dpp_d<0x141>(v));  // row_half_mirror
  return fmax(v, dpp_d<0x140>(v));  // row_mirror
}

// Sum over aligned groups of G lanes (G a power of two <= 64). Each step's
// partner lane holds the same value as the xor-(1, 2, ..., G/2) partner, so
// the sums equal those of xor exchanges bit for bit.
template <int G>
__device__ inline double group_sum(double v) {
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "G must be a power of two <= 64");
  if constexpr (G >= 2) v += dpp_d<0xB1>(v);   // lane ^ 1
  if constexpr (G >= 4) v += dpp_d<0x4E>(v);   // lane ^ 2
  if constexpr (G >= 8) v += dpp_d<0x141>(v);  // other quad of the 8-lane group
  if constexpr (G >= 16) v += dpp_d<0x140>(v); // other half of the row
  if constexpr (G >= 32) v = pair_sum16(v);
  if constexpr (G >= 64) v = pair_sum32(v);
  return v;
}

// Sum of V = 2^H values over the 64 lanes of a wavefront. On return lane l
// holds the total of value index l >> (6 - H) (every lane of its group).
template <int V>
__device__ inline double wave_reduce_spread(double (&v)[V], int lane) {
  constexpr int H = (V >= 64) ? 6 : (V >= 32) ? 5 : (V >= 16) ? 4 : (V >= 8) ? 3 : (V >= 4) ? 2 : (V >= 2) ? 1 : 0;
  static_assert((1 << H) == V, "V must be a power of two <= 64");
#pragma unroll
  for (int t = 0; t < H; t++) {
    const int m = 32 >> t;
    const int half = V >> (t + 1);
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < half; i++) {
      double send = up ? v[i] : v[i + half];
      double keep = up ? v[i + half] : v[i];
      v[i] = keep + __shfl_xor(send, m);
    }
  }
  double s = v[0];
#pragma unroll
  for (int m = (32 >> H); m >= 1; m >>= 1) s += __shfl_xor(s, m);
  return s;
}

}  // namespace pr

// ------------------------------------------------------------ FP32 products
// Mixed-precision form for single-wavefront matrices (d <= 16). The result
// layout of v_mfma_f32_16x16x4f32 is lane c + 16q, register r = element
// (row 4q + r, column c) — the "own" layout of Blk32. Its k-step kb sums over
// k = 4q + kb, so a Blk32 in registers is directly the B operand; the A
// operand comes from an LDS tile of float2 with the same padded indexing as
// the FP64 tiles (lidx), or from the registers of a matrix whose conjugate
// transpose is wanted.
namespace pr {

typedef float f4 __attribute__((ext_vector_type(4)));

struct Blk32 {
  f4 re, im;
};

__device__ inline f4 mfma32(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// row of the own FP32 layout held by lane-group q in register r
__device__ constexpr int rowof32(int q, int r) { return 4 * q + r; }

typedef float fv2 __attribute__((ext_vector_type(2)));
// (FP32 images of D_k: written once, read once, non-temporal like the FP64 ones)
__device__ inline Blk32 load_blk32(const float2* __restrict__ img, const Ln& x) {
  Blk32 b;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const fv2 v = __builtin_nontemporal_load((const PR_GLOBAL fv2*)img + r * 64 + x.lane);
    b.re[r] = v.x;
    b.im[r] = v.y;
  }
  return b;
}
__device__ inline void store_blk32(float2* __restrict__ img, const Blk32& b, const Ln& x) {
#pragma unroll
  for (int r = 0; r < 4; r++) {
    fv2 v;
    v.x = b.re[r];
    v.y = b.im[r];
    __builtin_nontemporal_store(v, (PR_GLOBAL fv2*)img + r * 64 + x.lane);
  }
}

// own-layout FP32 block into an FP32 tile
__device__ inline void put32(float2* L, const Blk32& b, const Ln& x) {
#pragma unroll
  for (int r = 0; r < 4; r++) L[lidx(0, rowof32(x.q, r), x.c)] = make_float2(b.re[r], b.im[r]);
}
// A * Bown with A from the FP32 tile L (Gauss's three real products per k-step)
__device__ inline Blk32 prod32_lds(const float2* L, const Blk32& Bo, const Ln& x) {
  f4 t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    const float2 a = L[lidx(0, x.c, rowof32(x.q, kb))];  // A[c][4q + kb]
    const float br = Bo.re[kb], bi = Bo.im[kb];
    t1 = mfma32(a.x, br, t1);
    t2 = mfma32(a.y, bi, t2);
    t3 = mfma32(a.x + a.y, br + bi, t3);
  }
  Blk32 r;
  r.re = t1 - t2;
  r.im = t3 - t1 - t2;
  return r;
}

// D^dag * Bown, D in the own layout (lane c + 16q, register kb holds
// D[4q + kb][c] = conj(D^dag[c][4q + kb]): the A operand without LDS)
__device__ inline Blk32 prod32_adj(const Blk32& D, const Blk32& Bo) {
  f4 t1 = {0, 0, 0, 0}, t2 = {0, 0, 0, 0}, t3 = {0, 0, 0, 0};
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    const float ar = D.re[kb], ai = -D.im[kb];
    const float br = Bo.re[kb], bi = Bo.im[kb];
    t1 = mfma32(ar, br, t1);
    t2 = mfma32(ai, bi, t2);
    t3 = mfma32(ar + ai, br + bi, t3);
  }
  Blk32 r;
  r.re = t1 - t2;
  r.im = t3 - t1 - t2;
  return r;
}

__device__ inline void add_diag32(Blk32& X, float v, const Ln& x) {
#pragma unroll
  for (int r = 0; r < 4; r++)
    if (rowof32(x.q, r) == x.c) X.re[r] += v;
}

}  // namespace pr
