// split_kernels.hpp — the FP16x2 split-product form (PR_SPLIT16) for d <= 16.
//
// A product Z Y of two 16x16 complex matrices runs on v_mfma_f32_16x16x32_f16
// with each operand split into two FP16 parts, v = 2^-e (hi + lo), hi the
// nearest FP16 of the scaled value and lo the nearest FP16 of the remainder:
// the three products hi hi + hi lo + lo hi accumulate in FP32 and carry about
// 22 bits, FP32-class. Where both operands are one matrix (the square X X of
// the exp polynomial) the dropped lo lo term has one sign for every slice (the
// square of the remainder of an anti-Hermitian matrix is negative
// semidefinite), so it is kept there: without it the bias added up over 2000
// slices to |dF| ~ 2.7e-9 near F = 1 in a CPU emulation of this arithmetic, with
// it 7.5e-11 (FP32 products: 6.4e-11).
//
// Complex operands use the real 2x2 block form with K = 32: for Z Y,
//   Re = [Re Z | -Im Z] [Re Y ; Im Y],  Im = [Im Z | Re Z] [Re Y ; Im Y],
// and the k index of lane c + 16q, slot i (i < 4: real part of row 4q + i,
// i >= 4: imaginary part of row 4q + i - 4) matches the result layout of the
// instruction (lane c + 16q, register i = element (4q + i, c), the "own"
// layout of Blk32; checked by an in-bounds lane-layout test). So the B operand
// of a product is the split of own(Y) as it stands, and the A operand is the
// split of own(Z^T). For Hermitian or anti-Hermitian Z, own(Z^T) is own(Z)
// with signs changed, so those A operands come from Z's own registers.
//
// The exp polynomial, the gradient walk's increments and the stored images
// D_k = U_k^dag - I use these products; generators, the forward chain
// T <- U_k T, accumulations of the walk, traces and reductions stay FP64.
#pragma once
#include "grape_kernels.hpp"

namespace pr {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

// four FP16 values (elements 0..3 of the own layout) in two dwords
struct H4 {
  unsigned w0, w1;
};
// the split of 4 complex own-layout elements: hi and lo of the real and imaginary parts
struct Sp {
  H4 rh, ih, rl, il;
};

__device__ inline unsigned pk16(float a, float b) {
  fv2 v;
  v.x = a;
  v.y = b;
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, h2v));  // nearest (v_cvt_pk_f16_f32)
}
// (fp16(a - h.lo), fp16(b - h.hi)) for the packed hi pair h, with the FP16
// halves read in place (the compiler would convert each half out first)
__device__ inline unsigned lo16(unsigned h, float a, float b) {
  unsigned l;
  asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l) : "v"(h), "v"(a));
  asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(l) : "v"(h), "v"(b));
  return l;
}
// hi = fp16(s a), lo = fp16(s a - hi) for a pair, without forming s a in FP32
__device__ inline void split_pair_scaled(float a, float b, float s, unsigned& h, unsigned& l) {
  asm("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(h) : "v"(a), "v"(s));
  asm("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(h) : "v"(b), "v"(s));
  asm("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(a), "v"(s), "v"(h));
  asm("v_fma_mixhi_f16 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l) : "v"(b), "v"(s), "v"(h));
}

__device__ inline void split4(const f4& v, H4& h, H4& l) {
  h.w0 = pk16(v[0], v[1]);
  h.w1 = pk16(v[2], v[3]);
  l.w0 = lo16(h.w0, v[0], v[1]);
  l.w1 = lo16(h.w1, v[2], v[3]);
}
__device__ inline void split4s(const f4& v, float s, H4& h, H4& l) {
  split_pair_scaled(v[0], v[1], s, h.w0, l.w0);
  split_pair_scaled(v[2], v[3], s, h.w1, l.w1);
}
__device__ inline Sp split(const Blk32& b) {
  Sp s;
  split4(b.re, s.rh, s.rl);
  split4(b.im, s.ih, s.il);
  return s;
}
__device__ inline Sp split_scaled(const Blk32& b, float sc) {
  Sp s;
  split4s(b.re, sc, s.rh, s.rl);
  split4s(b.im, sc, s.ih, s.il);
  return s;
}

__device__ inline H4 negh(const H4& a) { return H4{a.w0 ^ 0x80008000u, a.w1 ^ 0x80008000u}; }
__device__ inline h8 cat(const H4& a, const H4& b) { return __builtin_bit_cast(h8, u4v{a.w0, a.w1, b.w0, b.w1}); }

__device__ inline f4 mfma16(const h8& a, const h8& b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// The two A operands (hi and lo parts each) and one B operand of a product.
struct AOp {
  h8 a1h, a1l, a2h, a2l;
};
struct BOp {
  h8 h, l;
};
__device__ inline BOp bop(const Sp& s) { return BOp{cat(s.rh, s.ih), cat(s.rl, s.il)}; }
// A operand of Z from the split of own(Z) for Z Hermitian (the product gives Z Y)
// or anti-Hermitian (the product gives -Z Y): own(Z^T) = own(Z) up to signs
__device__ inline AOp aop_sym(const Sp& s) {
  return AOp{cat(s.rh, s.ih), cat(s.rl, s.il), cat(negh(s.ih), s.rh), cat(negh(s.il), s.rl)};
}
// A operand of Z from the split of own(Z^T) (general Z)
__device__ inline AOp aop_t(const Sp& t) {
  return AOp{cat(t.rh, negh(t.ih)), cat(t.rl, negh(t.il)), cat(t.ih, t.rh), cat(t.il, t.rl)};
}

// Z Y from the split operands, small terms first. LL: also lo lo.
template <bool LL>
__device__ inline Blk32 mm16(const AOp& a, const BOp& b) {
  const f4 z = {0, 0, 0, 0};
  f4 cr, ci;
  if constexpr (LL) {
    cr = mfma16(a.a1l, b.l, z);
    ci = mfma16(a.a2l, b.l, z);
    cr = mfma16(a.a1h, b.l, cr);
    ci = mfma16(a.a2h, b.l, ci);
  } else {
    cr = mfma16(a.a1h, b.l, z);
    ci = mfma16(a.a2h, b.l, z);
  }
  cr = mfma16(a.a1l, b.h, cr);
  ci = mfma16(a.a2l, b.h, ci);
  cr = mfma16(a.a1h, b.h, cr);
  ci = mfma16(a.a2h, b.h, ci);
  Blk32 r;
  r.re = cr;
  r.im = ci;
  return r;
}

// Stored image of D_k = U_k^dag - I: the split of 2^10 D_k, hi dwords then lo
// dwords of every lane ([2][64 lanes][4 dwords], 2 KiB per member-slice)
constexpr float SPLIT_SD = 1024.f;
// the split images pass the caches non-temporally (written once, read once)
#ifndef PR_NT_IMAGES
#define PR_NT_IMAGES 2
#endif
__device__ inline void store_split(unsigned* __restrict__ img, const Sp& s, const Ln& x) {
  u4v hi = {s.rh.w0, s.rh.w1, s.ih.w0, s.ih.w1};
  u4v lo = {s.rl.w0, s.rl.w1, s.il.w0, s.il.w1};
#if PR_NT_IMAGES
  __builtin_nontemporal_store(hi, (PR_GLOBAL u4v*)img + x.lane);
  __builtin_nontemporal_store(lo, (PR_GLOBAL u4v*)img + 64 + x.lane);
#else
  ((PR_GLOBAL u4v*)img)[x.lane] = hi;
  ((PR_GLOBAL u4v*)img)[64 + x.lane] = lo;
#endif
}
// one lane's hi (dwords 0..3) or lo (64 + lane) part of a stored split image
__device__ inline u4v load_split_part(const unsigned* __restrict__ img, int i) {
#if PR_NT_IMAGES > 1
  return __builtin_nontemporal_load((const PR_GLOBAL u4v*)img + i);
#else
  return ((const PR_GLOBAL u4v*)img)[i];
#endif
}
__device__ inline BOp load_split(const unsigned* __restrict__ img, const Ln& x, Sp& s) {
  const u4v hi = ((const PR_GLOBAL u4v*)img)[x.lane];
  const u4v lo = ((const PR_GLOBAL u4v*)img)[64 + x.lane];
  s.rh = H4{hi.x, hi.y};
  s.ih = H4{hi.z, hi.w};
  s.rl = H4{lo.x, lo.y};
  s.il = H4{lo.z, lo.w};
  return BOp{__builtin_bit_cast(h8, hi), __builtin_bit_cast(h8, lo)};
}

// Largest generator 1-norm the split form takes: the degree-6 remainder
// ||X||^7 / 7! is then below 5e-15 per slice (2000 slices stay far below the
// 1e-9 gate of the form). The bound is the triangle inequality over the
// operators, ||X||_1 <= dt (||A0||_1 + sum_j |s_mj u_jk| ||A_j||_1).
constexpr double THETA_SPLIT = 0.03;
// scale of Y = X_k^T in its split: |entries| <= 0.03 * 512 and (512 Y)^3 stays in FP16 range
constexpr float SPLIT_SY = 512.f;

#ifndef PR_SPLIT_SEGMENT_WAVES
#define PR_SPLIT_SEGMENT_WAVES 4
#endif
#ifndef PR_SPLIT_GRADIENT_WAVES
#define PR_SPLIT_GRADIENT_WAVES 4
#endif

// Set-up of the split segment walk held in registers for a whole work item:
// the member's drift (own layout) and, for NP > 0, the NP slot passes of the
// assembly (the operator entries of each register position and the lane of
// their control's weight; the same every slice).
template <int NP>
struct SplitAsm {
  int4 id[NP > 0 ? NP : 1];
  double2 v[NP > 0 ? NP : 1][4];
};

template <int NP>
__device__ inline void split_asm_load(const ProbDev& P, SplitAsm<NP>& a, const Ln& x) {
  if constexpr (NP > 0) {
    typedef int iv4 __attribute__((ext_vector_type(4)));
    const int p0 = P.so_pass[0];
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const iv4 t = ((const PR_GLOBAL iv4*)P.so_idx)[(unsigned)((p0 + i) * 64 + x.lane)];
      a.id[i] = make_int4(t.x, t.y, t.z, t.w);
#pragma unroll
      for (int r = 0; r < 4; r++) a.v[i][r] = gload2(P.so_v, (unsigned)(((p0 + i) * 4 + r) * 64 + x.lane));
    }
  }
}

// X_k in the own layout, FP64: with NP > 0 from the registers of SplitAsm and
// the lane weight wl = s_mj u_jk of lane j (fetched by ds_bpermute), else by
// the general assembly (tables read from memory).
template <int NP, bool PRE>
__device__ inline Blk split_asm(const ProbDev& P, int m, int k, const double* __restrict__ u, double wl,
                                const SplitAsm<NP>& a, const Blk& H0, const Ln& x) {
  if constexpr (NP == 0) {
    return assemble<1, true>(P, m, k, u, H0, x);
  } else {
    Blk H = H0;
    const int lo = __double2loint(wl), hi = __double2hiint(wl);
#pragma unroll
    for (int i = 0; i < NP; i++) {
      const int ids[4] = {a.id[i].x, a.id[i].y, a.id[i].z, a.id[i].w};
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const double wj = __hiloint2double(__builtin_amdgcn_ds_bpermute(ids[r], hi),
                                           __builtin_amdgcn_ds_bpermute(ids[r], lo));
        H.re[r] = fma(wj, a.v[i][r].x, H.re[r]);
        H.im[r] = fma(wj, a.v[i][r].y, H.im[r]);
      }
    }
    if constexpr (PRE) {
      return H;
    } else {
      // (a prescaled problem in a group that is not all prescaled: no factor)
      if (P.prescaled) return H;
      const double dt = slice_dt(P, k);
      return cscale(P.cre * dt, P.cim * dt, H);
    }
  }
}

// split_asm with the slot tables read from memory, for any pass count np
// (the recomputing walk, which cannot hold the tables in registers): the same
// FMAs in the same order as split_asm<np, PRE>, so the same X.
__device__ inline Blk split_asm_mem(const ProbDev& P, int m, int k, const double* __restrict__ u, double wl, int np,
                                    bool pre, const Blk& H0, const Ln& x) {
  if (np == 0) return assemble<1, true>(P, m, k, u, H0, x);
  typedef int iv4 __attribute__((ext_vector_type(4)));
  Blk H = H0;
  const int lo = __double2loint(wl), hi = __double2hiint(wl);
  const int p0 = P.so_pass[0];
  for (int i = 0; i < np; i++) {
    const iv4 t = ((const PR_GLOBAL iv4*)P.so_idx)[(unsigned)((p0 + i) * 64 + x.lane)];
    const int ids[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double2 v = gload2(P.so_v, (unsigned)(((p0 + i) * 4 + r) * 64 + x.lane));
      const double wj = __hiloint2double(__builtin_amdgcn_ds_bpermute(ids[r], hi),
                                         __builtin_amdgcn_ds_bpermute(ids[r], lo));
      H.re[r] = fma(wj, v.x, H.re[r]);
      H.im[r] = fma(wj, v.y, H.im[r]);
    }
  }
  if (pre || P.prescaled) return H;
  const double dt = slice_dt(P, k);
  return cscale(P.cre * dt, P.cim * dt, H);
}

// coefficients of the split exp polynomial (scaled products, Q4 = 4 Q)
struct SplitCoef {
  static constexpr float q1 = 4.0f / 6, qy = 4.0f / 24, qn2 = 4.0f / (120.0f * 512 * 512),
                         qs3 = 4.0f / (720.0f * 512 * 512 * 512);
  static constexpr float rn2 = 1.0f / (2.0f * 512 * 512), rp3 = 1.0f / (4.0f * 512 * 512 * 512);
};

// One slice of the split exp from its FP64 generator X (own layout): Y = X^T =
// -conj(X) in FP32 and R^T = exp(Y) - I - Y by the degree-6 Paterson-Stockmeyer
// polynomial on split products,
//   N2 = -(s Y)^2, S3 = (s Y)^3, Q = I/6 + Y/24 + Y^2/120 + Y^3/720,
//   R^T = Y^2/2 + Y^3 Q.
// Returns the split of 2^10 D_k, D_k = U_k^dag - I = conj(Y + R^T) (the stored
// image), and R^T in r. The segment kernel and the recomputing walk both call
// it, so they hold the same D_k.
__device__ inline Sp split_exp_slice(const Blk& X, const f4& diag, Blk32& y, Blk32& r) {
  y.re = -__builtin_convertvector(X.re, f4);
  y.im = __builtin_convertvector(X.im, f4);
  const Sp sY = split_scaled(y, SPLIT_SY);
  const AOp aY = aop_sym(sY);
  const Blk32 n2 = mm16<true>(aY, bop(sY));          // -(s Y)^2
  const Blk32 s3 = mm16<false>(aY, bop(split(n2)));  // -(s Y)(-(s Y)^2) = (s Y)^3
  Blk32 q;
  q.re = SplitCoef::qy * y.re - SplitCoef::qn2 * n2.re + SplitCoef::qs3 * s3.re + SplitCoef::q1 * diag;
  q.im = SplitCoef::qy * y.im - SplitCoef::qn2 * n2.im + SplitCoef::qs3 * s3.im;
  const Blk32 p3 = mm16<false>(aop_sym(split(s3)), bop(split(q)));  // -(s Y)^3 (4 Q)
  r.re = -SplitCoef::rn2 * n2.re - SplitCoef::rp3 * p3.re;
  r.im = -SplitCoef::rn2 * n2.im - SplitCoef::rp3 * p3.im;
  Blk32 dd;
  dd.re = y.re + r.re;
  dd.im = -(y.im + r.im);
  return split_scaled(dd, SPLIT_SD);
}

// Slices of a segment whose product is accumulated in the Delta-form (below)
// before an FP64 product folds it into the running segment product.
constexpr int SPLIT_GROUP = 128;
// doubles of controls the split segment kernel stages in LDS at a time
#ifndef PR_SPLIT_CTRL_CHUNK
#define PR_SPLIT_CTRL_CHUNK 256
#endif
constexpr int SPLIT_CTRL_CHUNK = PR_SPLIT_CTRL_CHUNK;
// its LDS: two tiles and the control chunk
constexpr size_t SPLIT_SEGMENT_SMEM = Smem<1>::segment + SPLIT_CTRL_CHUNK * sizeof(double);

#ifndef PR_SPLIT_SEGMENT_WAVES_REG
#define PR_SPLIT_SEGMENT_WAVES_REG 3
#endif

// k_segment for d <= 16 closed dynamics in the split form. First, for the
// whole work item, the bound ||X_k||_1 <= dt (||A0||_1 + sum_j |s_mj u_jk|
// ||A_j||_1) of every slice against THETA_SPLIT, one slice per lane: an item
// with a slice above it flags itself in P.seg_big and stops before any work;
// k_segment<1, true, true>, launched next, redoes it in FP64 (FP64 images,
// P.big set for its slices). Then per slice: the FP64 generator X (own
// layout, NP > 0: the assembly tables held in registers, the next slice's
// controls fetched one slice ahead) and Y = X^T = -conj(X); R^T = exp(Y) - I - Y
// by the degree-6 Paterson-Stockmeyer polynomial on split products,
//   N2 = -(s Y)^2, S3 = (s Y)^3, Q = I/6 + Y/24 + Y^2/120 + Y^3/720,
//   R^T = Y^2/2 + Y^3 Q,
// and the split image of D_k = U_k^dag - I = conj(Y + R^T). The segment
// product is accumulated in the Delta-form W = I + F over up to SPLIT_GROUP
// slices: with E_k = U_k - I = X + R (R by a transpose of R^T through LDS),
//   F <- F + X + R + E_k F,
// where X enters exactly in FP64 and only R and the product E_k F (a split
// product, A operand from the image of D_k) are rounded. A CPU emulation of
// this arithmetic on the four-qubit ensemble (2000 slices, groups of 2 to 250
// slices) kept |dF| at F = 1 within 3e-11 to 9e-11, as with an FP64 chain
// product per slice. The first group of a segment gives T = I + F; a later
// one (segments above SPLIT_GROUP slices) is folded in with one FP64 product
// T <- T + F T.
// PRE: every problem of the launch holds operators prescaled by c dt (uniform dt).
template <int NP, bool PRE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NP > 0 ? PR_SPLIT_SEGMENT_WAVES_REG
                                                                           : PR_SPLIT_SEGMENT_WAVES))) void k_segment_h(
    const ProbDev* __restrict__ probs, const WorkItem* __restrict__ items, const double* __restrict__ u) {
  // F's split scale (||F||_2 <= 2, W being unitary) and what undoes it with D's
  const float sf = 8192.f;
  const double finv = 1.0 / (8192.0 * (double)SPLIT_SD);
  const WorkItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const Ln x = lane_info<1>();
  const int m = it.member;
  const int nc = P.nc;
  // the bound of every slice of the item, 64 slices at a time
  {
    const double n0m = gptr(P.n0)[P.h0pm ? m : 0];
    const double* sc = P.scale + (size_t)m * nc;
    bool over = false;
    for (int kk = it.k0; kk < it.k1; kk += 64) {
      const int k = kk + x.lane;
      double nb = 0;
      if (k < it.k1) {
        const double* uk = u + P.coff + (size_t)k * nc;
        auto term = [&](int j) { return fabs(gptr(uk)[j]) * (gptr(P.ncn)[j] * gptr(sc)[j]); };
        if (P.nblk > 0) {
          // packed blocks: the generator is block diagonal, so its norm is
          // its largest block's (the drift bound n0m covers every block's)
          for (int b = 0; b < P.nblk; b++) {
            double nbb = P.ncg < nc ? term(P.ncg + b) : 0.0;
            for (int j = P.blk_ch[b]; j < P.blk_ch[b + 1]; j++) nbb += term(j);
            nb = fmax(nb, nbb);
          }
        } else {
          for (int j = 0; j < nc; j++) nb = fma(fabs(gptr(uk)[j]), gptr(P.ncn)[j] * gptr(sc)[j], nb);
        }
        nb = (n0m + nb) * (P.prescaled ? 1.0 : slice_dt(P, k));
      }
      over = over || __ballot(k < it.k1 && !(nb <= THETA_SPLIT)) != 0;
    }
    if (threadIdx.x == 0) P.seg_big[(size_t)m * P.nseg + it.seg] = over ? 1 : 0;
    if (over) return;
  }
  double2 *LA, *LB, *zb;
  double* red;
  smem_parts<1>(LA, LB, red, zb);
  float2* L32 = reinterpret_cast<float2*>(LB);
  // the member's drift (own layout) is read from LDS every slice; the tile is
  // also the fold's, which writes F over it
  const double2* H0o = P.H0o + (P.h0pm ? (size_t)m * 256 : 0);
  auto drift_to_lds = [&]() {
    const Blk h = load_blk(H0o, x);
#pragma unroll
    for (int r = 0; r < 4; r++) LA[r * 64 + x.lane] = make_double2(h.re[r], h.im[r]);
  };
  drift_to_lds();
  SplitAsm<NP> A;
  split_asm_load<NP>(P, A, x);
  const double scl = x.lane < nc ? gptr(P.scale)[(size_t)m * nc + x.lane] : 0.0;
  // The item's controls pass through LDS in chunks of ch slices (one global
  // load and wait per chunk; a per-slice global load would wait on the image
  // stores issued before it as well, which share its counter).
  double* LC = reinterpret_cast<double*>(LB + Smem<1>::TILE);
  const int ch = max(1, SPLIT_CTRL_CHUNK / nc);
  auto load_chunk = [&](int kc) {
    const int n = min(ch, it.k1 - kc) * nc;
    const double* src = u + P.coff + (size_t)kc * nc;
    for (int e = x.lane; e < n; e += 64) LC[e] = gptr(src)[e];
    bsync<1>();
  };
  load_chunk(it.k0);
  double un = x.lane < nc ? LC[x.lane] : 0.0;
  unsigned* img = reinterpret_cast<unsigned*>(P.Udm) + (size_t)m * P.nt * 512;
  double2* Tout = P.Tseg + ((size_t)m * P.nseg + it.seg) * 256;
  const f4 diag = {rowof32(x.q, 0) == x.c ? 1.f : 0.f, rowof32(x.q, 1) == x.c ? 1.f : 0.f,
                   rowof32(x.q, 2) == x.c ? 1.f : 0.f, rowof32(x.q, 3) == x.c ? 1.f : 0.f};
  bool first_group = true;  // the segment product T so far is I (later groups: T is in Tout)
  Blk F = zero_blk();       // W - I of the current group (own layout, FP64)
  int g = 0;                // slices in the current group
  bsync<1>();
  for (int k = it.k0; k < it.k1; k++) {
    const double wl = scl * un;
    if (k + 1 < it.k1) {
      const int j = (k + 1 - it.k0) % ch;
      if (j == 0) load_chunk(k + 1);
      if (x.lane < nc) un = LC[j * nc + x.lane];
    }
    Blk H0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double2 v = LA[r * 64 + x.lane];
      H0.re[r] = v.x;
      H0.im[r] = v.y;
    }
    const Blk X = split_asm<NP, PRE>(P, m, k, u, wl, A, H0, x);
    // F enters the product E_k F as it was before this slice; X goes into F at once
    Blk32 f32;
    f32.re = __builtin_convertvector(F.re, f4);
    f32.im = __builtin_convertvector(F.im, f4);
    F.re += X.re;
    F.im += X.im;
    Blk32 y, r;  // Y = X^T, R^T
    const Sp sD = split_exp_slice(X, diag, y, r);
    // own(R) = the transpose of own(R^T), through an FP32 tile
#pragma unroll
    for (int i = 0; i < 4; i++) L32[lidx(0, rowof32(x.q, i), x.c)] = make_float2(r.re[i], r.im[i]);
    // D = U^dag - I = conj(Y + R^T), stored split (unless the walk recomputes
    // it); it is also the A operand of E = D^dag
    if (!P.recompute) store_split(img + (size_t)k * 512, sD, x);
    bsync<1>();
    Blk32 t;  // R, then R + E F
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float2 v = L32[lidx(0, x.c, rowof32(x.q, i))];  // R^T[c][4q + i] = R[4q + i][c]
      t.re[i] = v.x;
      t.im[i] = v.y;
    }
    if (g > 0) {
      const AOp aE{cat(sD.rh, sD.ih), cat(sD.rl, sD.il), cat(negh(sD.ih), sD.rh), cat(negh(sD.il), sD.rl)};
      const Blk32 pf = mm16<false>(aE, bop(split_scaled(f32, sf)));  // 2^13 2^10 E F
      // R + E F in FP32 (E F is FP32-class already), then into F in FP64
      const float fi = (float)finv;
      t.re += fi * pf.re;
      t.im += fi * pf.im;
    }
    F.re += __builtin_convertvector(t.re, d4);
    F.im += __builtin_convertvector(t.im, d4);
    bsync<1>();
    if (++g == SPLIT_GROUP || k + 1 == it.k1) {
      // fold the group in: T = I + F for the first group, T <- T + F T after;
      // T is kept in Tseg between folds (FP64 MFMA layout)
#pragma unroll
      for (int i = 0; i < 4; i++) LA[lidx(0, rowof32(x.q, i), x.c)] = make_double2(F.re[i], F.im[i]);
      bsync<1>();
      Blk T;
      if (first_group) {
        T = eye_blk(x);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const double2 v = LA[lidx(0, rowof(x.q, i), x.c)];  // F in the FP64 MFMA layout
          T.re[i] += v.x;
          T.im[i] += v.y;
        }
      } else {
        T = load_blk(Tout, x);
        const Blk none = zero_blk();
        const Blk FT = gemm_prod<1, N, N>(LA, none, LB, T, x);  // F T (FP64)
        T.re += FT.re;
        T.im += FT.im;
      }
      if (k + 1 == it.k1) {
        // The segment product is unitary; the FP16 matrix cores truncate each
        // group of 8 products toward zero below 2^-24 of the group's largest,
        // which shrinks the products E_k F by a few 1e-8 of their size and
        // leaves T with a Hermitian error of ~1e-9 that moves every member's F
        // at F ~ 1 the same way. One Newton-Schulz step of the polar
        // decomposition in FP64, T <- T (3 I - T^dag T) / 2, takes out the
        // Hermitian part H of T = V (I + H) to O(H^2).
        bsync<1>();
        const Blk TT = gemm_prod<1, C, N>(LA, T, LB, T, x);  // T^dag T (both from registers)
        Blk Z;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          Z.re[i] = -0.5 * TT.re[i];
          Z.im[i] = -0.5 * TT.im[i];
        }
        const Blk I = eye_blk(x);
        Z.re += 1.5 * I.re;
        Blk Ts = T;
        put<1>(LA, Ts, x);
        bsync<1>();
        T = gemm_prod<1, N, N>(LA, Ts, LB, Z, x);  // T (3 I - T^dag T) / 2
      }
      store_blk(Tout, T, x);
      bsync<1>();
      first_group = false;
      g = 0;
      F = zero_blk();
      if (k + 1 < it.k1) {
        drift_to_lds();
        bsync<1>();
      }
    }
  }
  store_exp_counts(P, m, it.seg, 0, 0, 0, 0);
}

// k_gradient for d <= 16, unitary dynamics, split form: the walk M_k = C_k +
// C_k D_k and C_{k+1} = M_k + D_k^dag M_k with both increments on split
// products (D_k from its stored split image, prefetched one slice ahead) and
// C, M accumulated in FP64 (own layout); traces FP64. C's split is scaled per
// work item by a power of two from the largest |entry| of C at the segment
// start (the walk is unitary, so its entries stay within 16 times that). Work
// items that the segment kernel redid in FP64 (P.seg_big, every slice of the
// item) are left to k_gradient<..., REDO>, launched next.
// RECOMP: problems without stored images (P.recompute) rebuild D_k every slice
// with the segment kernel's own calls (split_asm_mem with the group's pass
// count hnp and prescaled flag hpre, then split_exp_slice), so the walk sees
// the bits the image would have held; the drift sits in the LDS tile LB.
template <int V, bool REG, bool RECOMP = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PR_SPLIT_GRADIENT_WAVES))) void k_gradient_h(
    const ProbDev* __restrict__ probs, const WorkItem* __restrict__ items, const double* __restrict__ u, int hnp,
    int hpre) {
  const WorkItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int m = it.member;
  if (P.seg_big[(size_t)m * P.nseg + it.seg]) return;
  const Ln x = lane_info<1>();
  double2 *LA, *LB, *zb;
  double* red;
  smem_parts<1>(LA, LB, red, zb);
  float2* L32 = reinterpret_cast<float2*>(LA);
  const unsigned* img = reinterpret_cast<const unsigned*>(P.Udm) + (size_t)m * P.nt * 512;
  const bool rec = RECOMP && P.recompute;
  const int nc = P.nc;
  const double scl = (RECOMP && x.lane < nc) ? gptr(P.scale)[(size_t)m * nc + x.lane] : 0.0;
  const f4 diag = {rowof32(x.q, 0) == x.c ? 1.f : 0.f, rowof32(x.q, 1) == x.c ? 1.f : 0.f,
                   rowof32(x.q, 2) == x.c ? 1.f : 0.f, rowof32(x.q, 3) == x.c ? 1.f : 0.f};
  // D_k's split image, as stored (hi dwords, lo dwords), rebuilt from X_k
  auto recompute_d = [&](int k, u4v& h, u4v& l) {
    Blk H0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double2 v = LB[r * 64 + x.lane];
      H0.re[r] = v.x;
      H0.im[r] = v.y;
    }
    const double un = x.lane < nc ? gptr(u)[P.coff + (size_t)k * nc + x.lane] : 0.0;
    const Blk X = split_asm_mem(P, m, k, u, scl * un, hnp, hpre != 0, H0, x);
    Blk32 y, r;
    const Sp sD = split_exp_slice(X, diag, y, r);
    h = u4v{sD.rh.w0, sD.rh.w1, sD.ih.w0, sD.ih.w1};
    l = u4v{sD.rl.w0, sD.rl.w1, sD.il.w0, sD.il.w1};
  };
  u4v dh, dl;
  if (!rec) {
    // the first image is in flight while C_s is formed
    dh = load_split_part(img + (size_t)it.k0 * 512, x.lane);
    dl = load_split_part(img + (size_t)it.k0 * 512, 64 + x.lane);
  }
  Blk Es = load_blk(P.Eseg + ((size_t)m * (P.nseg + 1) + it.seg) * 256, x);
  Blk Om = load_blk(P.Om + (size_t)m * 256, x);
  Blk Y = mm<1, N, N>(Es, Om, LA, LB, x);
  put<1>(LA, Es, x);
  bsync<1>();
  Blk Ed;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    double2 v = LA[lidx(0, x.c, rowof(x.q, r))];  // Es[c][q + 4r]
    Ed.re[r] = v.x;
    Ed.im[r] = -v.y;
  }
  bsync<1>();
  const Blk Cs = mm<1, N, N>(Y, Ed, LA, LB, x);
  // C into the own layout (kept in FP64)
  put<1>(LA, Cs, x);
  bsync<1>();
  Blk Cm;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    double2 v = LA[lidx(0, rowof32(x.q, r), x.c)];
    Cm.re[r] = v.x;
    Cm.im[r] = v.y;
  }
  bsync<1>();
  // scale of C's splits: 2^-e with 2^e >= max |C_s entry| (both parts)
  float sc;
  double inv;
  {
    double v = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) v = fmax(v, fmax(fabs(Cm.re[r]), fabs(Cm.im[r])));
    v = max_in_row(v);
    v = fmax(v, __shfl_xor(v, 16));
    v = fmax(v, __shfl_xor(v, 32));
    const int e = v > 0 ? ilogb(v) + 1 : 0;  // v < 2^e
    sc = ldexpf(1.0f, -e);
    inv = ldexp(1.0, e) / (double)SPLIT_SD;  // undoes the scales of C (or M) and of D in a product
  }
  TraceRegs<V> tregs;
  if constexpr (REG) trace_regs_load<V>(P, m, tregs, x);
  // slice lengths: with a uniform dt one scalar load before the walk (a
  // per-slice load would also wait for the image prefetched after it)
  const bool pre = P.prescaled != 0;
  const double dt0 = P.dt0;
  double dtk = pre ? dt0 : gptr(P.dts)[it.k0];
  auto contract_c = [&](int kk, double dtkk) {
    if constexpr (REG)
      contract_regs<V, true>(P, m, kk, dtkk, Cm, LA, tregs, x);
    else
      contract<1, V, true>(P, m, kk, Cm, LA, x);
  };
  if (rec) {
    // the member's drift (own layout) into LB, which the walk leaves alone
    const Blk h = load_blk(P.H0o + (P.h0pm ? (size_t)m * 256 : 0), x);
#pragma unroll
    for (int r = 0; r < 4; r++) LB[r * 64 + x.lane] = make_double2(h.re[r], h.im[r]);
    bsync<1>();
  }
  if (P.after) contract_c(it.k0, dtk);
  for (int k = it.k0; k < it.k1; k++) {
    double dtn = dtk;
    if (!pre && k + 1 < it.k1) dtn = gptr(P.dts)[k + 1];
    if (rec) recompute_d(k, dh, dl);
    // own(C^T), then its split: the A operand of C. With the first-order term
    // after U_k the trace of slice k has just left C_k in the FP64 tile (own
    // positions); otherwise C goes through the tile in FP32.
    Blk32 ct;
    if (P.after) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const double2 v = LA[lidx(0, x.c, rowof32(x.q, r))];  // C[c][4q + r]
        ct.re[r] = (float)v.x;
        ct.im[r] = (float)v.y;
      }
    } else {
      Blk32 c32;
      c32.re = __builtin_convertvector(Cm.re, f4);
      c32.im = __builtin_convertvector(Cm.im, f4);
#pragma unroll
      for (int r = 0; r < 4; r++) L32[lidx(0, rowof32(x.q, r), x.c)] = make_float2(c32.re[r], c32.im[r]);
      bsync<1>();
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const float2 v = L32[lidx(0, x.c, rowof32(x.q, r))];  // C[c][4q + r]
        ct.re[r] = v.x;
        ct.im[r] = v.y;
      }
    }
    const BOp bD{__builtin_bit_cast(h8, dh), __builtin_bit_cast(h8, dl)};
    const Blk32 p1 = mm16<false>(aop_t(split_scaled(ct, sc)), bD);  // sc 2^10 C D
    // A operand of E = U - I = D^dag: own(E^T) = own(conj(D)) = (Re D, -Im D)
    const H4 dih{dh.z, dh.w}, dil{dl.z, dl.w}, drh{dh.x, dh.y}, drl{dl.x, dl.y};
    const AOp aE{bD.h, bD.l, cat(negh(dih), drh), cat(negh(dil), drl)};
    if (!rec && k + 1 < it.k1) {
      dh = load_split_part(img + (size_t)(k + 1) * 512, x.lane);
      dl = load_split_part(img + (size_t)(k + 1) * 512, 64 + x.lane);
    }
    Blk Mk;
    Mk.re = Cm.re + inv * __builtin_convertvector(p1.re, d4);
    Mk.im = Cm.im + inv * __builtin_convertvector(p1.im, d4);
    Blk32 m32;
    m32.re = __builtin_convertvector(Mk.re, f4);
    m32.im = __builtin_convertvector(Mk.im, f4);
    const Blk32 p2 = mm16<false>(aE, bop(split_scaled(m32, sc)));  // 2^10 sc (U - I) M
    Cm.re = Mk.re + inv * __builtin_convertvector(p2.re, d4);
    Cm.im = Mk.im + inv * __builtin_convertvector(p2.im, d4);
    bsync<1>();
    if (P.after) {
      if (k + 1 < it.k1) contract_c(k + 1, dtn);
    } else if constexpr (REG) {
      contract_regs<V, true>(P, m, k, dtk, Mk, LA, tregs, x);
    } else {
      contract<1, V, true>(P, m, k, Mk, LA, x);
    }
    dtk = dtn;
  }
}

}  // namespace pr
