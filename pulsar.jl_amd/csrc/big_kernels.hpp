// big_kernels.hpp — the GRAPE inner loop for 64 < d <= 128 (5..8 blocks per side).
//
// A 16B x 16B complex FP64 matrix of this size (up to 256 KiB) does not fit a
// CU's LDS, so each workgroup of 16 wavefronts keeps its matrices as tile
// images in global scratch (the same [block][r][lane] layout as the U images)
// and streams products through LDS one k-panel at a time: for each inner
// block K the column panel op(A)(:, K) and the row panel op(B)(K, :) (2B
// padded tiles, at most 68 KiB) are staged, then every wavefront accumulates
// its output blocks b = w, w + 16, ... with FP64 MFMA (four real products
// per k-step). Elementwise steps (polynomial combinations, identity,
// scaling) run on the owned blocks directly from and to global memory.
//
// The kernels cover closed-dynamics gate problems under the reference
// gradient formula (the plan rejects anything else at this size):
//   k_segment_big   (member, segment)  exps and segment products; stores U_k
//   k_combine_big   (member)           prefix products, F_m, Omega
//   k_gradient_big  (member, segment)  C_s = E_s Omega E_s^dag, then per slice
//                                      M = C U^dag, traces, C <- U M
// The algorithm is the one of the d <= 64 kernels (grape_kernels.hpp).
#pragma once
#include "grape_kernels.hpp"

namespace pr {
namespace big {

#ifndef PR_BIG_WAVES
#define PR_BIG_WAVES 16
#endif
constexpr int W = PR_BIG_WAVES;  // wavefronts per workgroup
// third dword of a raw buffer resource on gfx9-family chips (gfx950)
constexpr int BUF_DW3 = 0x00020000;

template <int B>
struct Cfg {
  static constexpr int NB = B * B;
  static constexpr int P = (NB + W - 1) / W;  // output blocks per wavefront
  static constexpr size_t MB = (size_t)NB * 256;
  static constexpr int TILE = 272;  // padded 16 x 17 complex tile
  // LDS: 2B panel tiles, then column sums (B x 16B doubles) and 32 doubles
  static constexpr size_t panel_bytes = (size_t)2 * B * TILE * sizeof(double2);
  static constexpr size_t smem = panel_bytes + (size_t)B * 16 * B * sizeof(double) + 64 * sizeof(double);
};

// per-block lane info: block b of a B x B grid
template <int B>
__device__ inline Ln blk_info(int b) {
  Ln x;
  x.lane = threadIdx.x & 63;
  x.q = x.lane >> 4;
  x.c = x.lane & 15;
  x.w = b;
  x.I = b / B;
  x.Jb = b % B;
  return x;
}

__device__ inline int wave_id() { return threadIdx.x >> 6; }

// Out = op(A) op(B) for OA, OB in {N, T, C}; A, Bm, Out are tile images (Out
// must not alias A or Bm). LP: LDS panel space (2B tiles).
//
// For each inner block K the column panel op(A)(:, K) and the row panel
// op(B)(K, :) are staged in LDS; every wavefront accumulates its output blocks
// b = w, w + 16, ... with FP64 MFMA (four real products per k-step). The next
// K's panels are loaded into registers while the current K's products run
// (their global latency is hidden behind the matrix cores), then written to
// LDS after the barrier. The K loop is not unrolled and the function is
// inlined at every call: a called function spilled its accumulators and
// saved registers around every product.
template <int B, int OA, int OB>
__device__ __attribute__((always_inline)) inline void mm(double2* __restrict__ Out, const double2* A,
                                                         const double2* Bm, double2* LP) {
  using K_ = Cfg<B>;
  // per thread: PH elements of each panel (op(A)(:, K): tiles 0..B-1, op(B)(K, :):
  // tiles B..2B-1), loaded with the panel's own (wave-uniform) base pointer
  constexpr int NH = B * 256;                      // elements per panel
  constexpr int PH = (NH + W * 64 - 1) / (W * 64);
  const int w = wave_id();
  Blk acc[K_::P];
#pragma unroll
  for (int p = 0; p < K_::P; p++) acc[p] = zero_blk();
  double2 pa[PH], pb[PH];
  // buffer loads: the image base stays in scalar registers and each lane keeps
  // a 32-bit offset (64-bit per-lane addresses of both panels spilled); the
  // range check of the resource returns 0 beyond the image
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc((void*)A, 0, (int)(K_::MB * 16), BUF_DW3);
  const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc((void*)Bm, 0, (int)(K_::MB * 16), BUF_DW3);
  auto fetch = [&](int K) {
    // element (t, r, l) of panel K: block src = t B + K or K B + t of the image
    const int ka = ((OA == N) ? K : K * B) * 256 * 16;
    const int kb = ((OB == N) ? K * B : K) * 256 * 16;
#pragma unroll
    for (int i = 0; i < PH; i++) {
      const int e = (int)threadIdx.x + i * W * 64;
      if (NH % (W * 64) == 0 || e < NH) {
        const int t = e >> 8, rl = e & 255;  // rl = r * 64 + l
        const int va = (t * ((OA == N) ? B : 1) * 256 + rl) * 16;
        const int vb = (t * ((OB == N) ? 1 : B) * 256 + rl) * 16;
        pa[i] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(ra, va, ka, 0));
        pb[i] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rb, vb, kb, 0));
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < PH; i++) {
      const int e = (int)threadIdx.x + i * W * 64;
      if (NH % (W * 64) == 0 || e < NH) {
        const int t = e >> 8, r = (e >> 6) & 3, l = e & 63;
        const int o = rowof(l >> 4, r) * 17 + (l & 15);
        LP[t * 272 + o] = pa[i];
        LP[(B + t) * 272 + o] = pb[i];
      }
    }
  };
  fetch(0);
#pragma unroll 1
  for (int K = 0; K < B; K++) {
    stage();
    __syncthreads();
    if (K + 1 < B) fetch(K + 1);
    // (k-steps outermost and not unrolled: holding every operand of the
    // unrolled K step at once spilled the accumulators)
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
#pragma unroll 1
    for (int kb = 0; kb < 4; kb++) {
#pragma unroll
      for (int p = 0; p < K_::P; p++) {
        const int b = w + W * p;
        if (b < K_::NB) {
          const int I = b / B, J = b % B;
          // op(A)(I, K)[c][q + 4kb] and op(B)(K, J)[q + 4kb][c] before
          // conjugation: the signs of the imaginary parts go into the MFMA
          // negate modifiers (Re += ar br - ai bi, Im += ar bi + ai br)
          const double2 a = (OA == N) ? LP[I * 272 + c * 17 + rowof(q, kb)] : LP[I * 272 + rowof(q, kb) * 17 + c];
          const double2 bb = (OB == N) ? LP[(B + J) * 272 + rowof(q, kb) * 17 + c]
                                       : LP[(B + J) * 272 + c * 17 + rowof(q, kb)];
          constexpr bool ca = OA == C, cb = OB == C;
          acc[p].re = mfma(a.x, bb.x, acc[p].re);
          acc[p].im = cb ? mfma_nega(a.x, bb.y, acc[p].im) : mfma(a.x, bb.y, acc[p].im);
          acc[p].re = (ca != cb) ? mfma(a.y, bb.y, acc[p].re) : mfma_nega(a.y, bb.y, acc[p].re);
          acc[p].im = ca ? mfma_nega(a.y, bb.x, acc[p].im) : mfma(a.y, bb.x, acc[p].im);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < K_::P; p++) {
    const int b = w + W * p;
    if (b < K_::NB) store_blk(Out, acc[p], blk_info<B>(b));
  }
  __syncthreads();
}

// D = alpha I + a1 X1 + a2 X2 + a3 X3 (elementwise on owned blocks; X3 may be null)
template <int B>
__device__ void combo(double2* D, double alpha, double a1, const double2* X1, double a2, const double2* X2,
                      double a3, const double2* X3) {
  const int w = wave_id();
  for (int b = w; b < Cfg<B>::NB; b += W) {
    const Ln x = blk_info<B>(b);
    Blk r = load_blk(X1, x);
    r.re *= a1;
    r.im *= a1;
    const Blk y = load_blk(X2, x);
    r.re += a2 * y.re;
    r.im += a2 * y.im;
    if (X3) {
      const Blk z = load_blk(X3, x);
      r.re += a3 * z.re;
      r.im += a3 * z.im;
    }
    add_diag(r, alpha, x);
    store_blk(D, r, x);
  }
  __syncthreads();
}

__device__ inline void store_eye_blocks(double2* D, int NB_, int B_) {
  for (int b = wave_id(); b < NB_; b += W) {
    Ln x;
    x.lane = threadIdx.x & 63;
    x.q = x.lane >> 4;
    x.c = x.lane & 15;
    x.w = b;
    x.I = b / B_;
    x.Jb = b % B_;
    store_blk(D, eye_blk(x), x);
  }
}

// max over the workgroup of v (every thread contributes)
__device__ inline double wg_max(double v, double* red) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  if ((threadIdx.x & 63) == 0) red[wave_id()] = v;
  __syncthreads();
  double out = 0;
  for (int i = 0; i < W; i++) out = fmax(out, red[i]);
  __syncthreads();
  return out;
}

// workgroup sum of a complex value
__device__ inline double2 wg_sum2(double2 v, double* red) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    v.x += __shfl_xor(v.x, m);
    v.y += __shfl_xor(v.y, m);
  }
  if ((threadIdx.x & 63) == 0) {
    red[2 * wave_id()] = v.x;
    red[2 * wave_id() + 1] = v.y;
  }
  __syncthreads();
  double2 out = make_double2(0, 0);
  for (int i = 0; i < W; i++) {
    out.x += red[2 * i];
    out.y += red[2 * i + 1];
  }
  __syncthreads();
  return out;
}

// X = generator of slice k into the image X (owned blocks); returns the 1-norm bound
template <int B>
__device__ double assemble_norm(const ProbDev& P, int m, int k, const double* u, double2* X, double* colsum,
                                double* red) {
  constexpr int NB = Cfg<B>::NB;
  const double2* H0 = P.H0 + (P.h0pm ? (size_t)m * NB * 256 : 0);
  for (int i = threadIdx.x; i < 16 * B * B; i += W * 64) colsum[i] = 0;
  __syncthreads();
  for (int b = wave_id(); b < NB; b += W) {
    const Ln x = blk_info<B>(b);
    const Blk Xb = assemble<B>(P, m, k, u, load_blk(H0, x), x);
    store_blk(X, Xb, x);
    double cs = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) cs += abs_bound(Xb.re[r], Xb.im[r]);
    cs += __shfl_xor(cs, 16);
    cs += __shfl_xor(cs, 32);
    if (x.q == 0) colsum[x.I * 16 * B + 16 * x.Jb + x.c] = cs;  // [row block][column]
  }
  __syncthreads();
  double best = 0;
  if ((int)threadIdx.x < 16 * B) {
    double t = 0;
    for (int I = 0; I < B; I++) t += colsum[I * 16 * B + threadIdx.x];
    best = t;
  }
  return wg_max(best, red);
}

template <int B>
__device__ void scale_blocks(double2* X, double f) {
  for (int b = wave_id(); b < Cfg<B>::NB; b += W) {
    const Ln x = blk_info<B>(b);
    Blk r = load_blk(X, x);
    r.re *= f;
    r.im *= f;
    store_blk(X, r, x);
  }
  __syncthreads();
}

}  // namespace big

// scratch matrices per work item of the big kernels: the segment walk's X,
// X^2, X^3, R, Q, T, T' and, without stored images, U_k; the gradient walk's
// C, M, Y and, recomputing U_k, X, X^2, X^3, R, Q and U_k
constexpr int BIG_SCRATCH = 9;

namespace big {

// U = exp(X_k) for slice k of member m into the image U: the generator and
// its 1-norm bound, the scaling exponent and degree (choose_scaling), the
// Taylor polynomial of degree 6 or 9 in Paterson-Stockmeyer form and the
// squarings. The segment walk and, for plans that keep no U_k images, the
// gradient walk both call it, so a recomputed U_k is the stored one bit for
// bit. Returns the scaling exponent; deg6 tells the degree.
template <int B>
__device__ int exp_slice(const ProbDev& P, int m, int k, const double* __restrict__ u, double2* X, double2* X2,
                         double2* X3, double2* R, double2* Q, double2* U, double2* LP, double* colsum,
                         double* red, bool& deg6) {
  const double c2 = 1.0 / 2, c3 = 1.0 / 6, c4 = 1.0 / 24, c5 = 1.0 / 120, c6 = 1.0 / 720, c7 = 1.0 / 5040,
               c8 = 1.0 / 40320, c9 = 1.0 / 362880;
  const int s = choose_scaling(assemble_norm<B>(P, m, k, u, X, colsum, red), deg6);
  if (s) scale_blocks<B>(X, ldexp(1.0, -s));
  mm<B, N, N>(X2, X, X, LP);
  mm<B, N, N>(X3, X, X2, LP);
  if (deg6) {
    combo<B>(R, c3, c4, X, c5, X2, c6, X3);
  } else {
    combo<B>(Q, c6, c7, X, c8, X2, c9, X3);
    mm<B, N, N>(R, X3, Q, LP);
    combo<B>(R, c3, 1.0, R, c5, X2, c4, X);
  }
  mm<B, N, N>(Q, X3, R, LP);
  combo<B>(U, 1.0, 1.0, Q, c2, X2, 1.0, X);
  for (int i = 0; i < s; i++) {
    mm<B, N, N>(Q, U, U, LP);
    combo<B>(U, 0.0, 1.0, Q, 0.0, Q, 0.0, nullptr);
  }
  return s;
}

}  // namespace big

template <int B>
__global__ __launch_bounds__(64 * big::W) void k_segment_big(const ProbDev* __restrict__ probs,
                                                      const WorkItem* __restrict__ items,
                                                      const double* __restrict__ u, double2* __restrict__ scratch) {
  using K_ = big::Cfg<B>;
  constexpr size_t MB = K_::MB;
  const WorkItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  extern __shared__ double2 smem_raw[];
  double2* LP = smem_raw;
  double* colsum = reinterpret_cast<double*>(reinterpret_cast<char*>(smem_raw) + K_::panel_bytes);
  double* red = colsum + 16 * B * B;
  double2* S = scratch + (size_t)blockIdx.x * BIG_SCRATCH * MB;
  double2 *X = S, *X2 = S + MB, *X3 = S + 2 * MB, *R = S + 3 * MB, *Q = S + 4 * MB, *T = S + 5 * MB,
          *T2 = S + 6 * MB, *Uw = S + 7 * MB;
  const int m = it.member;
  big::store_eye_blocks(T, K_::NB, B);
  __syncthreads();
  int smax = 0, n9 = 0, nsq = 0;
  for (int k = it.k0; k < it.k1; k++) {
    // the image of U_k itself, or (no stored images) a scratch matrix
    double2* U = P.recompute ? Uw : P.Ust + ((size_t)m * P.nt + k) * MB;
    bool deg6;
    const int s = big::exp_slice<B>(P, m, k, u, X, X2, X3, R, Q, U, LP, colsum, red, deg6);
    smax = max(smax, s);
    n9 += deg6 ? 0 : 1;
    nsq += s;
    big::mm<B, N, N>(T2, U, T, LP);
    double2* tmp = T;
    T = T2;
    T2 = tmp;
  }
  for (int b = big::wave_id(); b < K_::NB; b += big::W) {
    const Ln x = big::blk_info<B>(b);
    store_blk(P.Tseg + ((size_t)m * P.nseg + it.seg) * MB, load_blk(T, x), x);
  }
  store_exp_counts(P, m, it.seg, smax, n9, nsq, 0);
}

template <int B>
__global__ __launch_bounds__(64 * big::W) void k_combine_big(const ProbDev* __restrict__ probs,
                                                      const int2* __restrict__ members,
                                                      double2* __restrict__ scratch) {
  using K_ = big::Cfg<B>;
  constexpr size_t MB = K_::MB;
  const int2 pm = members[blockIdx.x];
  const ProbDev& P = probs[pm.x];
  const int m = pm.y;
  extern __shared__ double2 smem_raw[];
  double2* LP = smem_raw;
  double* colsum = reinterpret_cast<double*>(reinterpret_cast<char*>(smem_raw) + K_::panel_bytes);
  double* red = colsum + 16 * B * B;
  double2* S = scratch + (size_t)blockIdx.x * BIG_SCRATCH * MB;
  double2* E0 = P.Eseg + (size_t)m * (P.nseg + 1) * MB;
  big::store_eye_blocks(E0, K_::NB, B);
  __syncthreads();
  for (int s = 0; s < P.nseg; s++)
    big::mm<B, N, N>(E0 + (size_t)(s + 1) * MB, P.Tseg + ((size_t)m * P.nseg + s) * MB, E0 + (size_t)s * MB, LP);
  const double2* U = E0 + (size_t)P.nseg * MB;
  // unitarity defect
  big::mm<B, C, N>(S, U, U, LP);
  double v = 0;
  for (int b = big::wave_id(); b < K_::NB; b += big::W) {
    const Ln x = big::blk_info<B>(b);
    Blk D = load_blk(S, x);
    add_diag(D, -1.0, x);
#pragma unroll
    for (int r = 0; r < 4; r++) v = fmax(v, fmax(fabs(D.re[r]), fabs(D.im[r])));
  }
  v = big::wg_max(v, red);
  if (threadIdx.x == 0) P.udef[m] = v;
  // gate: Phi = Tr(U_t^dag U)/d, F = |Phi|^2, Omega = conj(w) U_t^dag U with w = 2 Phi / d
  double2 part = make_double2(0, 0);
  for (int b = big::wave_id(); b < K_::NB; b += big::W) {
    const Ln x = big::blk_info<B>(b);
    const Blk Tg = load_blk(P.Ut, x), Ub = load_blk(U, x);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      part.x += Tg.re[r] * Ub.re[r] + Tg.im[r] * Ub.im[r];
      part.y += Tg.re[r] * Ub.im[r] - Tg.im[r] * Ub.re[r];
    }
  }
  const double2 tr = big::wg_sum2(part, red);
  const double phr = tr.x / P.d, phi = tr.y / P.d;
  if (threadIdx.x == 0) P.Fm[m] = phr * phr + phi * phi;
  const double wr = 2.0 * phr / P.d, wi = 2.0 * phi / P.d;
  big::mm<B, C, N>(S, P.Ut, U, LP);
  for (int b = big::wave_id(); b < K_::NB; b += big::W) {
    const Ln x = big::blk_info<B>(b);
    store_blk(P.Om + (size_t)m * MB, cscale(wr, -wi, load_blk(S, x)), x);
  }
}

// Without stored U_k images (P.recompute) the walk starts from the segment's
// checkpoint E_s and rebuilds each U_k with exp_slice, as k_segment_big made it.
template <int B>
__global__ __launch_bounds__(64 * big::W) void k_gradient_big(const ProbDev* __restrict__ probs,
                                                       const WorkItem* __restrict__ items,
                                                       const double* __restrict__ u, double2* __restrict__ scratch) {
  using K_ = big::Cfg<B>;
  constexpr size_t MB = K_::MB;
  const WorkItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int m = it.member;
  extern __shared__ double2 smem_raw[];
  double2* LP = smem_raw;
  double* colsum = reinterpret_cast<double*>(reinterpret_cast<char*>(smem_raw) + K_::panel_bytes);
  double* red = colsum + 16 * B * B;
  double2* S = scratch + (size_t)blockIdx.x * BIG_SCRATCH * MB;
  double2 *Cm = S, *Mk = S + MB, *Y = S + 2 * MB;
  double2 *X = S + 3 * MB, *X2 = S + 4 * MB, *X3 = S + 5 * MB, *R = S + 6 * MB, *Q = S + 7 * MB, *Uw = S + 8 * MB;
  const double2* Es = P.Eseg + ((size_t)m * (P.nseg + 1) + it.seg) * MB;
  big::mm<B, N, N>(Y, Es, P.Om + (size_t)m * MB, LP);
  big::mm<B, N, C>(Cm, Y, Es, LP);
  const int nc = P.nc, w = big::wave_id(), lane = threadIdx.x & 63;
  for (int k = it.k0; k < it.k1; k++) {
    const double2* U = P.Ust + ((size_t)m * P.nt + k) * MB;
    if (P.recompute) {
      bool deg6;
      (void)big::exp_slice<B>(P, m, k, u, X, X2, X3, R, Q, Uw, LP, colsum, red, deg6);
      U = Uw;
    }
    big::mm<B, N, C>(Mk, Cm, U, LP);  // M_k = C_k U_k^dag
    // G_m[j,k] = Re(-i dt s_mj Tr(A_j M)): wavefront w takes controls j = w, w + 16, ...
    // (M = C_k when the first-order term sits after U_k, P.after)
    const double2* Mt = P.after ? Cm : Mk;
    for (int j = w; j < nc; j += big::W) {
      double sv = 0;
      const int e1 = gptr(P.tr_ptr)[j + 1];
      for (int e = gptr(P.tr_ptr)[j] + lane; e < e1; e += 64) {
        // LDS-tile index of the element (b, a) of M -> image offset of that element
        const int idx = gptr(P.tr_idx)[e];
        const int blk = idx / 272, rem = idx % 272, row = rem / 17, col = rem % 17;
        const double2 mv = gload2(Mt, (size_t)blk * 256 + (row >> 2) * 64 + col + 16 * (row & 3));
        const double2 a = gload2(P.tr_v, e);
        sv += a.x * mv.x + a.y * mv.y;  // Im(a * M[b][a]), a stored as (Im, Re)
      }
#pragma unroll
      for (int msk = 32; msk >= 1; msk >>= 1) sv += __shfl_xor(sv, msk);
      if (lane == 0) {
        const double f = slice_dt(P, k) * P.scale[(size_t)m * nc + j];
        gptr_w(P.Gm)[(size_t)m * nc * P.nt + (size_t)k * nc + j] = f * sv;
      }
    }
    big::mm<B, N, N>(Cm, U, Mk, LP);  // C_{k+1} = U_k M_k
  }
}

}  // namespace pr
