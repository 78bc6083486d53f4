// pulsar_rocm.hip — C-ABI of libpulsar_rocm.so (declared in include/pulsar_rocm.h).
//
// Host side: validation of the problems, conversion of column-major ComplexF64
// matrices to 16x16 tile images, time segmentation, device allocation and the
// launch sequence k_segment -> k_combine -> k_gradient -> k_reduce per
// evaluation. Every HIP call is checked; failures come back as PR_ERR_* codes
// with a message for pr_last_error (the reference's GPU paths instead warn and
// rerun on the CPU, CUDABackend.jl:144-147; that fallback belongs to the
// caller, which sees the status code).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pulsar_rocm.h"
#include "grape_kernels.hpp"
#include "split_kernels.hpp"
#include "big_kernels.hpp"

using cd = std::complex<double>;

// ------------------------------------------------------- static scratch
// The scratch memory of pr_eval_device_static: one array per GPU, held from
// the library's load until the process ends. Host code never uses the name
// `scratch` as an address (it is a host-side stand-in); pr_init looks up the
// array's device address on each device it initialises and keeps it per device.
#define PR_STATIC_SCRATCH_BYTES (64ull << 20)
#define PR_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[PR_STATIC_SCRATCH_BYTES];

namespace pr {

// Byte offsets into the static scratch of one problem's scratch buffers (-1:
// the plan allocates none of that buffer for the problem).
struct StaticOff {
  long long Ust, Udm, big, seg_big, Tseg, Eseg, Om, Bb, Pscr, Fm, Gm, udef, smax, Zseg, Pend, Lend, Qc;
};

// dst[p] = src[p] with every scratch pointer moved into the static scratch at
// base (the plan's own problem records stay as they are).
__global__ void k_bind_static(const ProbDev* __restrict__ src, ProbDev* __restrict__ dst,
                              const StaticOff* __restrict__ off, int n, unsigned char* base) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  ProbDev d = src[p];
  const StaticOff o = off[p];
  auto at = [&](long long x) -> void* { return x < 0 ? nullptr : (void*)(base + x); };
  d.Ust = (double2*)at(o.Ust);
  d.Udm = (float2*)at(o.Udm);
  d.big = (unsigned char*)at(o.big);
  d.seg_big = (int*)at(o.seg_big);
  d.Tseg = (double2*)at(o.Tseg);
  d.Eseg = (double2*)at(o.Eseg);
  d.Om = (double2*)at(o.Om);
  d.Bb = (double2*)at(o.Bb);
  d.Pscr = (double2*)at(o.Pscr);
  d.Fm = (double*)at(o.Fm);
  d.Gm = (double*)at(o.Gm);
  d.udef = (double*)at(o.udef);
  d.smax = (int*)at(o.smax);
  d.Zseg = (double2*)at(o.Zseg);
  d.Pend = (double2*)at(o.Pend);
  d.Lend = (double2*)at(o.Lend);
  d.Qc = (double2*)at(o.Qc);
  dst[p] = d;
}

// zero n doubles (a fill kernel: no memset node in a caller's captured graph)
__global__ void k_zero(double* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

}  // namespace pr

namespace {

unsigned char* g_scratch_ptr[PR_MAX_DEVICES] = {nullptr};

// Lindblad problems without stored images: segments of at most this many
// slices, and a gradient launch over at most lind_window(B) work items (about
// 8192 wavefronts), so that the walk's scratch (forward states and recomputed
// U_k^T, 2 * LIND_RECOMPUTE_SEG slices of NB tiles per work item) stays within
// 8192 * 64 * 4 KiB = 2 GiB whatever the ensemble and the slice count.
constexpr int LIND_RECOMPUTE_SEG = 32;
int lind_window(int B) { return std::max(1, 8192 / (B * B)); }

// segments from which a single-wavefront problem takes the two-level prefix,
// and the group member count below which groups take it
#ifndef PR_PREFIX_MIN_SEG1
#define PR_PREFIX_MIN_SEG1 64
#endif
#ifndef PR_PREFIX_MAX_MEMBERS
#define PR_PREFIX_MAX_MEMBERS 128
#endif

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(PR_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

bool fault_is(const char* what) {
  const char* f = getenv("PULSAR_ROCM_FAULT");
  return f && strcmp(f, what) == 0;
}

int g_device = -1;

int ensure_device() {
  if (g_device >= 0) {
    HIPCHK(hipSetDevice(g_device));
    return PR_OK;
  }
  return pr_init(0);
}

// element (r, c) of a column-major interleaved complex matrix
inline cd at(const double* a, int d, int r, int c) {
  size_t i = (size_t)r + (size_t)c * d;
  return cd(a[2 * i], a[2 * i + 1]);
}

// Tile image of a d x d column-major matrix padded to 16B x 16B (transpose:
// image of M^T). Block (I, J) holds 256 complex in [r][lane] order; lane
// l = c + 16q, register r is element (16I + q + 4r, 16J + c), the FP64 MFMA
// result layout.
// own32: the FP32 MFMA result layout instead (d <= 16): register r of lane
// c + 16q is element (4q + r, c).
inline int lane_row(int l, int r, bool own32) { return own32 ? 4 * (l >> 4) + r : (l >> 4) + 4 * r; }

void tile_image(const double* a, int d, int B, bool transpose, double2* out, bool own32 = false) {
  for (int I = 0; I < B; I++)
    for (int J = 0; J < B; J++) {
      double2* blk = out + (size_t)(I * B + J) * 256;
      for (int r = 0; r < 4; r++)
        for (int l = 0; l < 64; l++) {
          int row = 16 * I + lane_row(l, r, own32), col = 16 * J + (l & 15);
          cd v(0, 0);
          if (row < d && col < d) v = transpose ? at(a, d, col, row) : at(a, d, row, col);
          blk[r * 64 + l] = make_double2(v.real(), v.imag());
        }
    }
}

// tile image of the matrix whose column s is vector s of vecs ([n][d] complex)
void tile_columns(const double* vecs, int n, int d, int B, double2* out) {
  for (int I = 0; I < B; I++)
    for (int J = 0; J < B; J++) {
      double2* blk = out + (size_t)(I * B + J) * 256;
      for (int r = 0; r < 4; r++)
        for (int l = 0; l < 64; l++) {
          int row = 16 * I + (l >> 4) + 4 * r, col = 16 * J + (l & 15);
          double2 v = make_double2(0, 0);
          if (row < d && col < n) v = make_double2(vecs[2 * ((size_t)col * d + row)], vecs[2 * ((size_t)col * d + row) + 1]);
          blk[r * 64 + l] = v;
        }
    }
}

// Register-layout sparse form of the controls: for block (I, J), register r and
// lane l (the element (16I + (l>>4) + 4r, 16J + (l&15))) the nonzero (j, A_j)
// entries at that position, or of A_j^T when transpose is set.
struct Csr {
  std::vector<int> ptr, j;
  std::vector<double2> v;
};

Csr build_csr(const double* Hc, int nc, int d, int B, bool transpose, bool own32 = false) {
  Csr out;
  const int NB = B * B;
  const size_t dd = (size_t)d * d;
  out.ptr.assign((size_t)NB * 256 + 1, 0);
  for (int w = 0; w < NB; w++) {
    const int I = w / B, J = w % B;
    for (int r = 0; r < 4; r++)
      for (int l = 0; l < 64; l++) {
        const size_t pos = ((size_t)w * 4 + r) * 64 + l;
        const int row = 16 * I + lane_row(l, r, own32), col = 16 * J + (l & 15);
        if (row < d && col < d)
          for (int j = 0; j < nc; j++) {
            cd v = transpose ? at(Hc + 2 * dd * j, d, col, row) : at(Hc + 2 * dd * j, d, row, col);
            if (v != cd(0, 0)) {
              out.j.push_back(j);
              out.v.push_back(make_double2(v.real(), v.imag()));
            }
          }
        out.ptr[pos + 1] = (int)out.j.size();
      }
  }
  if (out.j.empty()) {  // keep the device arrays non-empty
    out.j.push_back(0);
    out.v.push_back(make_double2(0, 0));
  }
  return out;
}

// Slot form of a register-layout sparse form (nc <= 64, NB <= 64): block w gets as many
// passes as its fullest position has entries; pass p holds the p-th entry of
// every position of the block, zero where a position has fewer. The index is
// 4 j, the ds_bpermute byte address of the lane that holds control j's weight.
struct Slots {
  std::vector<int> ptr;
  std::vector<int4> idx;
  std::vector<double2> v;
};

Slots build_slots(const Csr& c, int NB) {
  Slots out;
  out.ptr.assign(NB + 1, 0);
  for (int w = 0; w < NB; w++) {
    int passes = 0;
    for (int pos = w * 256; pos < (w + 1) * 256; pos++) passes = std::max(passes, c.ptr[pos + 1] - c.ptr[pos]);
    out.ptr[w + 1] = out.ptr[w] + passes;
    for (int p = 0; p < passes; p++)
      for (int l = 0; l < 64; l++) {
        int id[4];
        for (int r = 0; r < 4; r++) {
          const int pos = (w * 4 + r) * 64 + l, e = c.ptr[pos] + p;
          const bool has = e < c.ptr[pos + 1];
          id[r] = has ? 4 * c.j[e] : 0;
          out.v.push_back(has ? c.v[e] : make_double2(0, 0));
        }
        out.idx.push_back(make_int4(id[0], id[1], id[2], id[3]));
      }
  }
  // v is [pass][lane][r] above; the device reads [pass][r][lane]
  std::vector<double2> t(out.v.size());
  const size_t np = out.idx.size() / 64;
  for (size_t p = 0; p < np; p++)
    for (int l = 0; l < 64; l++)
      for (int r = 0; r < 4; r++) t[(p * 4 + r) * 64 + l] = out.v[(p * 64 + l) * 4 + r];
  out.v.swap(t);
  if (out.idx.empty()) {  // keep the device arrays non-empty
    out.idx.push_back(make_int4(0, 0, 0, 0));
    out.v.push_back(make_double2(0, 0));
  }
  return out;
}

// Trace operand of the gradient: per control j the nonzeros A_j[a][b] with the
// LDS tile index of the element (b, a) of the matrix they multiply.
struct Trace {
  std::vector<int> ptr, idx;
  std::vector<double2> v;
};

Trace build_trace(const double* Hc, int nc, int d) {
  Trace out;
  const int B = (d + 15) / 16;
  const size_t dd = (size_t)d * d;
  out.ptr.push_back(0);
  for (int j = 0; j < nc; j++) {
    for (int b = 0; b < d; b++)
      for (int a = 0; a < d; a++) {
        cd v = at(Hc + 2 * dd * j, d, a, b);
        if (v == cd(0, 0)) continue;
        const int blk = (b / 16) * B + (a / 16);
        out.idx.push_back(blk * 272 + (b % 16) * 17 + (a % 16));
        out.v.push_back(make_double2(v.real(), v.imag()));
      }
    out.ptr.push_back((int)out.idx.size());
  }
  if (out.idx.empty()) {
    out.idx.push_back(0);
    out.v.push_back(make_double2(0, 0));
  }
  return out;
}

// Copy of n column-major complex d x d matrices made exactly Hermitian from
// their upper triangles (row <= column), the part Julia's Hermitian(H) reads
// (the reference exponentiates eigen(Hermitian(H)), Propagators.jl:663).
std::vector<double> hermitian_upper(const double* a, int n, int d) {
  const size_t dd = (size_t)d * d;
  std::vector<double> out(a, a + 2 * dd * n);
  for (int m = 0; m < n; m++) {
    double* h = out.data() + 2 * dd * m;
    for (int c = 0; c < d; c++) {
      h[2 * ((size_t)c + (size_t)c * d) + 1] = 0.0;
      for (int r = c + 1; r < d; r++) {  // lower entry (r, c) from upper (c, r)
        const size_t lo = (size_t)r + (size_t)c * d, up = (size_t)c + (size_t)r * d;
        h[2 * lo] = h[2 * up];
        h[2 * lo + 1] = -h[2 * up + 1];
      }
    }
  }
  return out;
}

double norm1(const double* a, int d) {
  double best = 0;
  for (int c = 0; c < d; c++) {
    double s = 0;
    for (int r = 0; r < d; r++) s += std::abs(at(a, d, r, c));
    best = std::max(best, s);
  }
  return best;
}

double herm_dev(const double* a, int d) {
  double dev = 0;
  for (int c = 0; c < d; c++)
    for (int r = 0; r < d; r++) dev = std::max(dev, std::abs(at(a, d, r, c) - std::conj(at(a, d, c, r))));
  return dev;
}

double unitary_dev(const double* a, int d) {
  double s = 0;
  for (int i = 0; i < d; i++)
    for (int j = 0; j < d; j++) {
      cd acc(0, 0);
      for (int k = 0; k < d; k++) acc += std::conj(at(a, d, k, i)) * at(a, d, k, j);
      if (i == j) acc -= 1.0;
      s += std::norm(acc);
    }
  return std::sqrt(s) / d;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct Plan {
  // the problems the kernels run on: the caller's, or, when small problems are
  // packed (pr_options.pack), the plan's own list of unpacked and packed ones
  int n_probs = 0;
  pr_options opt{};
  std::vector<pr::ProbDev> host_probs;
  std::vector<int> ng;                // n_ctrl * n_t per problem
  std::vector<long long> coff;        // controls offsets (n_probs + 1)
  long long n_params = 0;
  // the caller's problems (controls, F and G of pr_eval and pr_eval_device):
  // the same as above unless packed
  int n_user = 0;
  std::vector<long long> ucoff;       // (n_user + 1)
  long long n_uparams = 0;
  bool packed = false;
  std::vector<int4> umap;             // per caller's problem: (problem, block or -1, first channel, n_ctrl)
  long long* d_uoff = nullptr;
  int* d_unt = nullptr;               // n_t per caller's problem
  int* d_dsrc = nullptr;              // per problem: the caller's problem, or -1 - first entry of d_chan
  int2* d_chan = nullptr;             // per packed channel: (caller's problem, control j, or -1: drift)
  int4* d_umap = nullptr;
  double* d_uw = nullptr;             // per caller's problem: its weight (packed problems run at 1)
  double* d_udev = nullptr;           // controls in the plan's own layout (n_params)
  double* d_Fdev = nullptr;           // (n_probs)
  double* d_Gdev = nullptr;           // (n_params)
  std::vector<void*> allocs;
  pr::ProbDev* d_probs = nullptr;
  long long* d_off = nullptr;
  int* d_first = nullptr;             // per 256 controls: the problem of the first (pr::prob_of)
  int* d_ufirst = nullptr;            // the same over the caller's problems (packed plans)
  double* d_controls = nullptr;
  double* d_F = nullptr;
  double* d_G = nullptr;
  double* d_part = nullptr;  // [red_chunks][n_params] partial member sums
  int red_chunks = 1;        // member chunks of the gradient reduction: min(RED_CHUNKS, largest M)
  double* h_stage = nullptr; // pinned host staging: n_params controls or gradients, then n_probs F
  // segment length of each group relative to the work-per-item rule (the
  // autotuner's choice; 1 otherwise)
  std::vector<double> seg_scale = std::vector<double>(13, 1.0);
  // per group (autotuner's choice; 1 otherwise): 1 lets the d <= 16 kernels
  // hold the trace operand and the split assembly tables in registers, 0
  // makes them read both from memory (fewer registers, more loads)
  std::vector<int> tune_regs = std::vector<int>(13, 1);
  // pr_eval's staging ranges: problems [stage_cut[r], stage_cut[r + 1]) are
  // packed by thread r and copied by one transfer each, event stage_ev[r]
  // marking the range's gradient download
  std::vector<int> stage_cut;
  std::vector<hipEvent_t> stage_ev;
  // pr_eval_device_static: the byte sizes of each problem's scratch buffers,
  // their offsets in the static scratch (udef arrays first, so one fill zeroes
  // them), the reduction partials' offset, the total, and the device copies
  std::vector<pr::StaticOff> static_sz, static_off;
  long long static_part = 0, static_udef_bytes = 0, static_bytes = 0;
  long long static_udev = 0, static_Fdev = 0, static_Gdev = 0;  // packed plans
  bool static_ok = true;  // false when a group keeps per-item scratch outside the records (dim > 64)
  pr::StaticOff* d_static_off = nullptr;
  pr::ProbDev* d_probs_static = nullptr;
  struct Group {
    int B = 0, V = 0;
    int gid = 0;            // index of the group (B, or 8 + B for tracking)
    bool lindblad = false;  // some problem of the group has Lindblad dynamics
    bool mixed = false;     // low-precision exp and gradient walk (d <= 16 only): FP32 MFMA, or
    bool split = false;     // FP16x2 split products (PR_SPLIT16)
    bool prescaled = false; // every problem of the group has operators prescaled by c dt
    int split_np = 0;       // slot passes of the split segment kernel's assembly held in registers
                            // (the group's common pass count, 1..3), or 0: tables read per slice
    bool trace_regs = false;  // mixed walk keeps the trace operand in registers
    bool recompute = false;   // some problem of the group recomputes U_k in the gradient walk
    int window = 0;           // > 0: the gradient walk is launched over at most this many
                              // consecutive work items at a time (Lindblad scratch slots)
    bool track = false;       // tracking-objective problems (k_segment, k_track_scan, k_track_walk)
    bool track_packed = false;  // d <= 16 tracking with n_pairs <= 8: states and co-states in one block
    double2* big_scratch = nullptr;  // d > 64: per work item matrices in HBM
    pr::WorkItem* d_items = nullptr;
    int n_items = 0;
    pr::WorkItem* d_citems = nullptr;  // chunks of the two-level prefix (k_combine_local), if any
    int n_citems = 0;
    int2* d_members = nullptr;
    int n_members = 0;
  };
  std::vector<Group> groups;
  std::vector<double> taylor_deg;
  hipStream_t stream = nullptr;
  // optional per-kernel timing with HIP events on the launch stream
  bool timing = false;
  std::vector<hipEvent_t> ev_free;
  struct Rec {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  hipEvent_t ev_open = nullptr;
  double acc_ms[PR_KERNEL_CLASSES] = {0};
  long long acc_n[PR_KERNEL_CLASSES] = {0};
  // captured launch sequences of one evaluation, keyed by the buffers and
  // stream they were recorded with (replayed with hipGraphLaunch)
  struct Captured {
    const double* u;
    double* F;
    double* G;
    hipStream_t st;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
  bool use_graphs = true;
  ~Plan() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    for (auto& r : pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : ev_free) (void)hipEventDestroy(e);
    for (auto e : stage_ev) (void)hipEventDestroy(e);
    if (ev_open) (void)hipEventDestroy(ev_open);
    for (void* p : allocs) (void)hipFree(p);
    if (h_stage) (void)hipHostFree(h_stage);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <class T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(PR_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return PR_OK;
  }
  template <class T>
  int upload(T** p, const T* src, size_t count) {
    int rc = alloc(p, count);
    if (rc) return rc;
    HIPCHK(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return PR_OK;
  }
};

int validate(const pr_problem& p, int idx) {
  if (p.dim < 1 || p.n_ctrl < 1 || p.n_t < 1 || p.n_members < 1)
    return fail(PR_ERR_ARG, "problem %d: dim, n_ctrl, n_t and n_members must be >= 1", idx);
  if (!p.H0 || !p.Hc) return fail(PR_ERR_ARG, "problem %d: H0 and Hc are required", idx);
  if (p.dynamics != PR_CLOSED && p.dynamics != PR_LINDBLAD) return fail(PR_ERR_ARG, "problem %d: unknown dynamics", idx);
  if (p.dim > 128) return fail(PR_ERR_UNSUPPORTED, "problem %d: dim %d > 128 is not supported by this build", idx, p.dim);
  if (p.dim > 64 && (p.dynamics != PR_CLOSED || p.objective != PR_GATE))
    return fail(PR_ERR_UNSUPPORTED, "problem %d: dim %d > 64 is supported for closed-dynamics gate problems only",
                idx, p.dim);
  if (!p.dt_slices && !(p.dt >= 0.0)) return fail(PR_ERR_ARG, "problem %d: dt must be non-negative", idx);
  if (p.objective == PR_GATE || p.objective == PR_GATE_PROPAGATORS) {
    if (!p.U_target) return fail(PR_ERR_ARG, "problem %d: gate objective needs U_target", idx);
    if (p.dynamics != PR_CLOSED) return fail(PR_ERR_UNSUPPORTED, "problem %d: gate objective needs closed dynamics", idx);
    double dev = unitary_dev(p.U_target, p.dim);
    if (dev > 1e-8) return fail(PR_ERR_ARG, "problem %d: target is not unitary: ||U^dag U - I||/dim = %g > 1e-8", idx, dev);
  } else if (p.objective == PR_PAIRS_SQUARE || p.objective == PR_PAIRS_REAL || p.objective == PR_STATE ||
             p.objective == PR_STATE_PROPAGATORS || p.objective == PR_PAIRS_MODULUS) {
    if (p.n_pairs < 1 || !p.psi_init || !p.psi_targ) return fail(PR_ERR_ARG, "problem %d: pair objective needs states", idx);
    if ((p.objective == PR_STATE || p.objective == PR_STATE_PROPAGATORS) && p.dynamics != PR_CLOSED)
      return fail(PR_ERR_UNSUPPORTED, "problem %d: the state objective needs closed dynamics", idx);
    // lindblad_grad_prefactor has :square and :real only (Lindblad.jl:220-232)
    if (p.objective == PR_PAIRS_MODULUS && p.dynamics != PR_CLOSED)
      return fail(PR_ERR_UNSUPPORTED, "problem %d: the :modulus objective needs closed dynamics", idx);
    if (p.n_pairs > 16 * ((p.dim + 15) / 16))
      return fail(PR_ERR_UNSUPPORTED, "problem %d: n_pairs %d exceeds the padded dimension", idx, p.n_pairs);
  } else if (p.objective == PR_TRACKING) {
    if (p.dynamics != PR_CLOSED)
      return fail(PR_ERR_UNSUPPORTED, "problem %d: the tracking objective needs closed dynamics", idx);
    if (p.n_pairs < 1 || !p.psi_init) return fail(PR_ERR_ARG, "problem %d: tracking objective needs initial states", idx);
    if (p.n_pairs > 16 * ((p.dim + 15) / 16))
      return fail(PR_ERR_UNSUPPORTED, "problem %d: n_pairs %d exceeds the padded dimension", idx, p.n_pairs);
    // GRAPETracking.jl:146-148 (empty tracking) and :177-178 (step range)
    if (p.n_track < 1 || !p.track_step || !p.track_weight || !p.track_target)
      return fail(PR_ERR_ARG, "problem %d: tracking objective needs n_track >= 1 checkpoints", idx);
    for (int c = 0; c < p.n_track; c++) {
      if (p.track_step[c] < 1 || p.track_step[c] > p.n_t)
        return fail(PR_ERR_ARG, "problem %d: TrackingPoint.step=%d out of range [1, %d]", idx, p.track_step[c], p.n_t);
      if (!std::isfinite(p.track_weight[c])) return fail(PR_ERR_ARG, "problem %d: checkpoint %d weight is not finite", idx, c);
    }
  } else {
    return fail(PR_ERR_ARG, "problem %d: unknown objective %d", idx, p.objective);
  }
  if (p.dynamics == PR_CLOSED) {
    size_t dd = (size_t)p.dim * p.dim;
    int nh0 = p.H0_per_member ? p.n_members : 1;
    for (int m = 0; m < nh0; m++) {
      double dev = herm_dev(p.H0 + 2 * dd * m, p.dim);
      if (dev > 1e-10) return fail(PR_ERR_ARG, "problem %d: H0 (member %d) is not Hermitian: max|H - H^dag| = %g", idx, m, dev);
    }
    for (int j = 0; j < p.n_ctrl; j++) {
      double dev = herm_dev(p.Hc + 2 * dd * j, p.dim);
      if (dev > 1e-10) return fail(PR_ERR_ARG, "problem %d: control %d is not Hermitian: max|H - H^dag| = %g", idx, j, dev);
    }
  }
  if (p.dt_slices)
    for (int k = 0; k < p.n_t; k++)
      if (!(p.dt_slices[k] >= 0.0)) return fail(PR_ERR_ARG, "problem %d: dt_slices[%d] must be non-negative", idx, k);
  return PR_OK;
}

template <int B, int V>
int set_grad_attr(size_t bytes) {
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_gradient<B, V, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_gradient<B, V, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_gradient<B, V, false, false, true>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_gradient<B, V, true, false, true>,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return PR_OK;
}

template <int B>
int set_smem_attr() {
  size_t bytes = pr::Smem<B>::max;
  if (bytes > 65536) {
    HIPCHK(hipFuncSetAttribute((const void*)pr::k_segment<B, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    HIPCHK(hipFuncSetAttribute((const void*)pr::k_segment<B, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    HIPCHK(hipFuncSetAttribute((const void*)pr::k_combine<B>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    HIPCHK(hipFuncSetAttribute((const void*)pr::k_combine_local<B>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
    HIPCHK(hipFuncSetAttribute((const void*)pr::k_combine_apply<B>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
    int rc;
    if ((rc = set_grad_attr<B, 2>(bytes)) || (rc = set_grad_attr<B, 4>(bytes)) || (rc = set_grad_attr<B, 8>(bytes)) ||
        (rc = set_grad_attr<B, 16>(bytes)) || (rc = set_grad_attr<B, 32>(bytes)))
      return rc;
    for (const void* f : {(const void*)pr::k_segment<B, true, false, true>, (const void*)pr::k_track_scan<B>,
                          (const void*)pr::k_track_walk<B, 2>, (const void*)pr::k_track_walk<B, 4>,
                          (const void*)pr::k_track_walk<B, 8>, (const void*)pr::k_track_walk<B, 16>,
                          (const void*)pr::k_track_walk<B, 32>,
                          (const void*)pr::k_track_walk<B, 2, false, false, true>,
                          (const void*)pr::k_track_walk<B, 4, false, false, true>,
                          (const void*)pr::k_track_walk<B, 8, false, false, true>,
                          (const void*)pr::k_track_walk<B, 16, false, false, true>,
                          (const void*)pr::k_track_walk<B, 32, false, false, true>})
      HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  }
  return PR_OK;
}

int next_pow2_cap32(int n) {
  int v = 2;
  while (v < n && v < 32) v <<= 1;
  return v;
}

int take_event(Plan& pl, hipEvent_t* e) {
  if (!pl.ev_free.empty()) {
    *e = pl.ev_free.back();
    pl.ev_free.pop_back();
    return PR_OK;
  }
  HIPCHK(hipEventCreate(e));
  return PR_OK;
}

// event pair around one kernel launch of class cls (no-op unless timing is on)
int tbegin(Plan& pl, hipStream_t st) {
  if (!pl.timing) return PR_OK;
  int rc = take_event(pl, &pl.ev_open);
  if (rc) return rc;
  HIPCHK(hipEventRecord(pl.ev_open, st));
  return PR_OK;
}
int tend(Plan& pl, int cls, hipStream_t st) {
  if (!pl.timing) return PR_OK;
  hipEvent_t b;
  int rc = take_event(pl, &b);
  if (rc) return rc;
  HIPCHK(hipEventRecord(b, st));
  pl.pending.push_back({cls, pl.ev_open, b});
  pl.ev_open = nullptr;
  return PR_OK;
}

int launch_gradient_split(Plan& pl, const Plan::Group& g, const double* d_u, hipStream_t st) {
  // the recomputing walk keeps the drift in a second tile
  const size_t sm = g.recompute ? pr::Smem<1>::segment : pr::Smem<1>::gradient_unitary;
  const int hnp = g.split_np, hpre = g.prescaled ? 1 : 0;
  switch (g.V) {
#define PR_HCASE(VV)                                                                                    \
  case VV:                                                                                              \
    if (g.recompute && g.trace_regs)                                                                    \
      hipLaunchKernelGGL((pr::k_gradient_h<VV, true, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, \
                         g.d_items, d_u, hnp, hpre);                                                   \
    else if (g.recompute)                                                                               \
      hipLaunchKernelGGL((pr::k_gradient_h<VV, false, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, \
                         g.d_items, d_u, hnp, hpre);                                                   \
    else if (g.trace_regs)                                                                              \
      hipLaunchKernelGGL((pr::k_gradient_h<VV, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, g.d_items, d_u, \
                         hnp, hpre);                                                                   \
    else                                                                                                \
      hipLaunchKernelGGL((pr::k_gradient_h<VV, false>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, g.d_items, d_u, \
                         hnp, hpre);                                                                   \
    break;
    PR_HCASE(2) PR_HCASE(4) PR_HCASE(8) PR_HCASE(16) PR_HCASE(32)
#undef PR_HCASE
    default:
      return fail(PR_ERR_LAUNCH, "internal: bad reduction width %d", g.V);
  }
  HIPCHK(hipGetLastError());
  // the work items redone in FP64 by the segment kernels walk in FP64
  switch (g.V) {
#define PR_RCASE(VV)                                                                                    \
  case VV:                                                                                              \
    if (g.recompute && g.trace_regs)                                                                    \
      hipLaunchKernelGGL((pr::k_gradient<1, VV, false, true, true, true>), dim3(g.n_items), dim3(64),         \
                         pr::Smem<1>::segment, st, pl.d_probs, g.d_items, d_u);                         \
    else if (g.recompute)                                                                               \
      hipLaunchKernelGGL((pr::k_gradient<1, VV, false, false, true, true>), dim3(g.n_items), dim3(64),        \
                         pr::Smem<1>::segment, st, pl.d_probs, g.d_items, d_u);                         \
    else if (g.trace_regs)                                                                              \
      hipLaunchKernelGGL((pr::k_gradient<1, VV, false, true, false, true>), dim3(g.n_items), dim3(64), sm, st,   \
                         pl.d_probs, g.d_items, d_u);                                                   \
    else                                                                                                \
      hipLaunchKernelGGL((pr::k_gradient<1, VV, false, false, false, true>), dim3(g.n_items), dim3(64), sm, st,  \
                         pl.d_probs, g.d_items, d_u);                                                   \
    break;
    PR_RCASE(2) PR_RCASE(4) PR_RCASE(8) PR_RCASE(16) PR_RCASE(32)
#undef PR_RCASE
  }
  HIPCHK(hipGetLastError());
  return PR_OK;
}

int launch_gradient_mixed(Plan& pl, const Plan::Group& g, const double* d_u, hipStream_t st) {
  if (g.split) return launch_gradient_split(pl, g, d_u, st);
  // the recomputing walk rebuilds D_k in a second tile
  const size_t sm = g.recompute ? pr::Smem<1>::segment : pr::Smem<1>::gradient_unitary;
  switch (g.V) {
#define PR_MCASE(VV)                                                                                    \
  case VV:                                                                                              \
    if (g.recompute && g.trace_regs)                                                                    \
      hipLaunchKernelGGL((pr::k_gradient_mx<VV, true, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, \
                         g.d_items, d_u);                                                              \
    else if (g.recompute)                                                                               \
      hipLaunchKernelGGL((pr::k_gradient_mx<VV, false, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, \
                         g.d_items, d_u);                                                              \
    else if (g.trace_regs)                                                                              \
      hipLaunchKernelGGL((pr::k_gradient_mx<VV, true>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, g.d_items, d_u); \
    else                                                                                                \
      hipLaunchKernelGGL((pr::k_gradient_mx<VV, false>), dim3(g.n_items), dim3(64), sm, st, pl.d_probs, g.d_items, d_u); \
    break;
    PR_MCASE(2) PR_MCASE(4) PR_MCASE(8) PR_MCASE(16) PR_MCASE(32)
#undef PR_MCASE
    default:
      return fail(PR_ERR_LAUNCH, "internal: bad reduction width %d", g.V);
  }
  HIPCHK(hipGetLastError());
  return PR_OK;
}

template <int B>
int launch_gradient(Plan& pl, const Plan::Group& g, const double* d_u, hipStream_t st) {
  if (B == 1 && g.mixed) return launch_gradient_mixed(pl, g, d_u, st);
  size_t sm = g.lindblad ? pr::Smem<B>::gradient_lindblad : pr::Smem<B>::gradient_unitary;
  // the recomputing walk runs the segment kernel's exp (two tiles at B == 1)
  if (g.recompute) sm = std::max(sm, pr::Smem<B>::segment);
  const dim3 blk(64 * B * B);
  if (g.window > 0) {
    // the Lindblad walk over windows of consecutive work items: a problem's
    // items in one launch take distinct scratch slots (slot = item % chunk,
    // chunk >= the window or all of the problem's items)
    for (int i0 = 0; i0 < g.n_items; i0 += g.window) {
      const int n = std::min(g.window, g.n_items - i0);
      switch (g.V) {
#define PR_WCASE(VV)                                                                                          \
  case VV:                                                                                                    \
    if (pl.opt.grad_mode == PR_GRAD_EXACT)                                                                    \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, true, false, true>), dim3(n), blk, sm, st, pl.d_probs,        \
                         g.d_items + i0, d_u);                                                                \
    else                                                                                                      \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, false, false, true>), dim3(n), blk, sm, st, pl.d_probs,       \
                         g.d_items + i0, d_u);                                                                \
    break;
        PR_WCASE(2) PR_WCASE(4) PR_WCASE(8) PR_WCASE(16) PR_WCASE(32)
#undef PR_WCASE
        default:
          return fail(PR_ERR_LAUNCH, "internal: bad reduction width %d", g.V);
      }
      HIPCHK(hipGetLastError());
    }
    return PR_OK;
  }
  switch (g.V) {
#define PR_GCASE(VV)                                                                                        \
  case VV:                                                                                                  \
    if (g.recompute && pl.opt.grad_mode == PR_GRAD_EXACT)                                                   \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, true, false, true>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u); \
    else if (g.recompute)                                                                               \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, false, false, true>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u); \
    else if (pl.opt.grad_mode == PR_GRAD_EXACT)                                                             \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, true>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u); \
    else if (B == 1 && g.trace_regs)                                                                    \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, false, true>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u); \
    else                                                                                                \
      hipLaunchKernelGGL((pr::k_gradient<B, VV, false>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u); \
    break;
    PR_GCASE(2) PR_GCASE(4) PR_GCASE(8) PR_GCASE(16) PR_GCASE(32)
#undef PR_GCASE
    default:
      return fail(PR_ERR_LAUNCH, "internal: bad reduction width %d", g.V);
  }
  HIPCHK(hipGetLastError());
  return PR_OK;
}

template <int B>
int launch_group(Plan& pl, const Plan::Group& g, const double* d_u, bool grad, hipStream_t st) {
  const dim3 blk(64 * B * B);
  int rc;
  if ((rc = tbegin(pl, st))) return rc;
  if (B == 1 && g.mixed) {
    if (g.split) {
      const size_t sm = pr::SPLIT_SEGMENT_SMEM;
#define PR_SEGH(NP, PRE) hipLaunchKernelGGL((pr::k_segment_h<NP, PRE>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u)
      if (g.prescaled) {
        switch (g.split_np) {
          case 1: PR_SEGH(1, true); break;
          case 2: PR_SEGH(2, true); break;
          case 3: PR_SEGH(3, true); break;
          default: PR_SEGH(0, false);
        }
      } else {
        switch (g.split_np) {
          case 1: PR_SEGH(1, false); break;
          case 2: PR_SEGH(2, false); break;
          case 3: PR_SEGH(3, false); break;
          default: PR_SEGH(0, false);
        }
      }
#undef PR_SEGH
    } else
      hipLaunchKernelGGL(pr::k_segment_mx, dim3(g.n_items), blk, pr::Smem<1>::segment, st, pl.d_probs, g.d_items, d_u);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL((pr::k_segment<1, true, true>), dim3(g.n_items), blk, pr::Smem<1>::segment, st, pl.d_probs,
                       g.d_items, d_u);
  } else if (g.lindblad)
    hipLaunchKernelGGL((pr::k_segment<B, false>), dim3(g.n_items), blk, pr::Smem<B>::segment, st, pl.d_probs, g.d_items, d_u);
  else
    hipLaunchKernelGGL((pr::k_segment<B, true>), dim3(g.n_items), blk, pr::Smem<B>::segment, st, pl.d_probs, g.d_items, d_u);
  HIPCHK(hipGetLastError());
  if ((rc = tend(pl, 0, st))) return rc;
  if ((rc = tbegin(pl, st))) return rc;
  if (g.n_citems > 0) {
    hipLaunchKernelGGL((pr::k_combine_local<B>), dim3(g.n_citems), blk, pr::Smem<B>::combine, st, pl.d_probs,
                       g.d_citems);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL((pr::k_combine<B>), dim3(g.n_members), blk, pr::Smem<B>::combine, st, pl.d_probs, g.d_members);
  HIPCHK(hipGetLastError());
  if (g.n_citems > 0) {
    hipLaunchKernelGGL((pr::k_combine_apply<B>), dim3(g.n_items), blk, pr::Smem<B>::combine, st, pl.d_probs,
                       g.d_items);
    HIPCHK(hipGetLastError());
  }
  if ((rc = tend(pl, 1, st))) return rc;
  if (grad) {
    if ((rc = tbegin(pl, st))) return rc;
    if ((rc = launch_gradient<B>(pl, g, d_u, st))) return rc;
    if ((rc = tend(pl, 2, st))) return rc;
  }
  return PR_OK;
}

// Tracking group: the segment kernel stores every U_k^T image and each
// segment's checkpoint operator, a per-member scan chains the segments, then
// every segment walks back in parallel.
template <int B>
int launch_track(Plan& pl, const Plan::Group& g, const double* d_u, hipStream_t st) {
  const dim3 blk(64 * B * B);
  int rc;
  if ((rc = tbegin(pl, st))) return rc;
  hipLaunchKernelGGL((pr::k_segment<B, true, false, true>), dim3(g.n_items), blk, pr::Smem<B>::segment, st, pl.d_probs,
                     g.d_items, d_u);
  HIPCHK(hipGetLastError());
  if ((rc = tend(pl, 0, st))) return rc;
  if ((rc = tbegin(pl, st))) return rc;
  hipLaunchKernelGGL((pr::k_track_scan<B>), dim3(g.n_members), blk, pr::Smem<B>::base, st, pl.d_probs, g.d_members);
  HIPCHK(hipGetLastError());
  if ((rc = tend(pl, 1, st))) return rc;
  if ((rc = tbegin(pl, st))) return rc;
  switch (g.V) {
#define PR_TCASE(VV)                                                                                                 \
  case VV:                                                                                                           \
    if (B == 1 && g.track_packed && g.trace_regs && g.recompute)                                                     \
      hipLaunchKernelGGL((pr::k_track_walk<1, VV, true, true, true>), dim3(g.n_items), blk,                         \
                         pr::Smem<1>::base, st, pl.d_probs, g.d_items, d_u);                                          \
    else if (B == 1 && g.track_packed && g.recompute)                                                                \
      hipLaunchKernelGGL((pr::k_track_walk<1, VV, true, false, true>), dim3(g.n_items), blk,                        \
                         pr::Smem<1>::base, st, pl.d_probs, g.d_items, d_u);                                          \
    else if (g.recompute)                                                                                            \
      hipLaunchKernelGGL((pr::k_track_walk<B, VV, false, false, true>), dim3(g.n_items), blk, pr::Smem<B>::base, st, \
                         pl.d_probs, g.d_items, d_u);                                                                 \
    else if (B == 1 && g.track_packed && g.trace_regs)                                                               \
      hipLaunchKernelGGL((pr::k_track_walk<1, VV, true, true>), dim3(g.n_items), blk,                               \
                         pr::Smem<1>::base, st, pl.d_probs, g.d_items, d_u);                                          \
    else if (B == 1 && g.track_packed)                                                                               \
      hipLaunchKernelGGL((pr::k_track_walk<1, VV, true>), dim3(g.n_items), blk,                                     \
                         pr::Smem<1>::base, st, pl.d_probs, g.d_items, d_u);                                          \
    else                                                                                                             \
      hipLaunchKernelGGL((pr::k_track_walk<B, VV>), dim3(g.n_items), blk, pr::Smem<B>::base, st, pl.d_probs,         \
                         g.d_items, d_u);                                                                            \
    break;
    PR_TCASE(2) PR_TCASE(4) PR_TCASE(8) PR_TCASE(16) PR_TCASE(32)
#undef PR_TCASE
    default:
      return fail(PR_ERR_LAUNCH, "internal: bad reduction width %d", g.V);
  }
  HIPCHK(hipGetLastError());
  return tend(pl, 2, st);
}

template <int B>
int set_big_attr() {
  const int bytes = (int)pr::big::Cfg<B>::smem;
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_segment_big<B>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_combine_big<B>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  HIPCHK(hipFuncSetAttribute((const void*)pr::k_gradient_big<B>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  return PR_OK;
}

template <int B>
int launch_big(Plan& pl, const Plan::Group& g, const double* d_u, bool grad, hipStream_t st) {
  const dim3 blk(64 * pr::big::W);
  const size_t sm = pr::big::Cfg<B>::smem;
  int rc;
  if ((rc = tbegin(pl, st))) return rc;
  hipLaunchKernelGGL((pr::k_segment_big<B>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u, g.big_scratch);
  HIPCHK(hipGetLastError());
  if ((rc = tend(pl, 0, st))) return rc;
  if ((rc = tbegin(pl, st))) return rc;
  hipLaunchKernelGGL((pr::k_combine_big<B>), dim3(g.n_members), blk, sm, st, pl.d_probs, g.d_members, g.big_scratch);
  HIPCHK(hipGetLastError());
  if ((rc = tend(pl, 1, st))) return rc;
  if (grad) {
    if ((rc = tbegin(pl, st))) return rc;
    hipLaunchKernelGGL((pr::k_gradient_big<B>), dim3(g.n_items), blk, sm, st, pl.d_probs, g.d_items, d_u,
                       g.big_scratch);
    HIPCHK(hipGetLastError());
    if ((rc = tend(pl, 2, st))) return rc;
  }
  return PR_OK;
}

// the launches of one dimension group (segment, combine, gradient walk)
int launch_one(Plan& pl, const Plan::Group& g, const double* d_u, bool grad, hipStream_t st) {
  if (g.track) {
    switch (g.B) {
      case 1: return launch_track<1>(pl, g, d_u, st);
      case 2: return launch_track<2>(pl, g, d_u, st);
      case 3: return launch_track<3>(pl, g, d_u, st);
      case 4: return launch_track<4>(pl, g, d_u, st);
      default: return fail(PR_ERR_LAUNCH, "internal: block count %d", g.B);
    }
  }
  switch (g.B) {
    case 1: return launch_group<1>(pl, g, d_u, grad, st);
    case 2: return launch_group<2>(pl, g, d_u, grad, st);
    case 3: return launch_group<3>(pl, g, d_u, grad, st);
    case 4: return launch_group<4>(pl, g, d_u, grad, st);
    case 5: return launch_big<5>(pl, g, d_u, grad, st);
    case 6: return launch_big<6>(pl, g, d_u, grad, st);
    case 7: return launch_big<7>(pl, g, d_u, grad, st);
    case 8: return launch_big<8>(pl, g, d_u, grad, st);
    default: return fail(PR_ERR_LAUNCH, "internal: block count %d", g.B);
  }
}

// packed plans: the caller's controls d_u in the plan's own layout (pl.d_udev)
int pack_controls(Plan& pl, const double* d_u, hipStream_t st) {
  const long long total = pl.n_params;
  hipLaunchKernelGGL(pr::k_pack_controls, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, pl.d_probs,
                     pl.d_off, pl.d_first, total, pl.d_dsrc, pl.d_chan, pl.d_uoff, pl.d_unt, d_u, pl.d_udev);
  HIPCHK(hipGetLastError());
  return PR_OK;
}

int run(Plan& pl, const double* d_u, double* d_F, double* d_G, hipStream_t st) {
  int rc;
  const double* u = d_u;
  double* F = d_F;
  double* G = d_G;
  if (pl.packed) {
    if ((rc = tbegin(pl, st))) return rc;
    if ((rc = pack_controls(pl, d_u, st))) return rc;
    if ((rc = tend(pl, 3, st))) return rc;
    u = pl.d_udev;
    F = pl.d_Fdev;
    G = d_G ? pl.d_Gdev : nullptr;
  }
  for (const auto& g : pl.groups) {
    if ((rc = launch_one(pl, g, u, G != nullptr, st))) return rc;
  }
  if ((rc = tbegin(pl, st))) return rc;
  hipLaunchKernelGGL(pr::k_reduce_F, dim3(pl.n_probs), dim3(64), 0, st, pl.d_probs, pl.n_probs, F);
  HIPCHK(hipGetLastError());
  if (G) {
    long long total = pl.n_params;
    const int nch = pl.red_chunks;
    hipLaunchKernelGGL(pr::k_reduce_G1, dim3((unsigned)((total + 255) / 256), nch), dim3(256), 0, st, pl.d_probs,
                       pl.d_off, pl.d_first, total, nch, nch == 1 ? G : pl.d_part);
    HIPCHK(hipGetLastError());
    if (nch > 1) {
      hipLaunchKernelGGL(pr::k_reduce_G2, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total, nch,
                         pl.d_part, G);
      HIPCHK(hipGetLastError());
    }
  }
  if (pl.packed) {
    hipLaunchKernelGGL(pr::k_unpack_F, dim3((pl.n_user + 255) / 256), dim3(256), 0, st, pl.n_user, pl.d_umap,
                       pl.d_uw, pl.d_probs, F, d_F);
    HIPCHK(hipGetLastError());
    if (d_G) {
      const long long total = pl.n_uparams;
      hipLaunchKernelGGL(pr::k_unpack_G, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total, pl.d_uoff,
                         pl.d_ufirst, pl.d_umap, pl.d_uw, pl.d_probs, pl.d_off, G, d_G);
      HIPCHK(hipGetLastError());
    }
  }
  if ((rc = tend(pl, 3, st))) return rc;
  return PR_OK;
}

// One evaluation as a replayed HIP graph: the first call with a given set of
// buffers and stream captures run()'s launches (no memset or copy nodes), later
// calls replay them. Per-kernel timing records events between launches, so
// timed evaluations launch directly.
int run_graph(Plan& pl, const double* d_u, double* d_F, double* d_G, hipStream_t st) {
  if (pl.timing || !pl.use_graphs) return run(pl, d_u, d_F, d_G, st);
  for (const auto& g : pl.graphs)
    if (g.u == d_u && g.F == d_F && g.G == d_G && g.st == st) {
      HIPCHK(hipGraphLaunch(g.exec, st));
      return PR_OK;
    }
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = run(pl, d_u, d_F, d_G, st);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(PR_ERR_LAUNCH, "hipStreamEndCapture: %s", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(PR_ERR_LAUNCH, "hipGraphInstantiate: %s", hipGetErrorString(ei));
  if (pl.graphs.size() >= 8) {  // bounded cache: drop the oldest
    (void)hipGraphExecDestroy(pl.graphs.front().exec);
    pl.graphs.erase(pl.graphs.begin());
  }
  pl.graphs.push_back({d_u, d_F, d_G, st, exec});
  HIPCHK(hipGraphLaunch(exec, st));
  return PR_OK;
}

// Offsets of every problem's scratch buffers in the static scratch, 256-byte
// aligned: all udef arrays first (the one range that must start at zero), then
// the other buffers problem by problem, then the reduction partials. The device
// copies of the offsets and of the problem records they produce are part of the
// plan, so pr_eval_device_static itself allocates nothing.
int static_layout(Plan& pl) {
  auto align = [](long long x) { return (x + 255) & ~255LL; };
  long long at = 0;
  pl.static_off.assign(pl.n_probs, pr::StaticOff{});
  for (int p = 0; p < pl.n_probs; p++) {
    pl.static_off[p].udef = at;
    at = align(at + pl.static_sz[p].udef);
  }
  pl.static_udef_bytes = at;
  for (int p = 0; p < pl.n_probs; p++) {
    const pr::StaticOff& z = pl.static_sz[p];
    pr::StaticOff& o = pl.static_off[p];
    auto place = [&](long long bytes) -> long long {
      if (bytes == 0) return -1;
      const long long r = at;
      at = align(at + bytes);
      return r;
    };
    o.Ust = place(z.Ust); o.Udm = place(z.Udm); o.big = place(z.big); o.seg_big = place(z.seg_big);
    o.Tseg = place(z.Tseg); o.Eseg = place(z.Eseg); o.Om = place(z.Om); o.Bb = place(z.Bb);
    o.Pscr = place(z.Pscr); o.Fm = place(z.Fm); o.Gm = place(z.Gm); o.smax = place(z.smax);
    o.Zseg = place(z.Zseg); o.Pend = place(z.Pend); o.Lend = place(z.Lend); o.Qc = place(z.Qc);
  }
  pl.static_part = at;
  at = align(at + (long long)sizeof(double) * pl.red_chunks * pl.n_params);
  if (pl.packed) {
    pl.static_udev = at;
    at = align(at + (long long)sizeof(double) * pl.n_params);
    pl.static_Fdev = at;
    at = align(at + (long long)sizeof(double) * pl.n_probs);
    pl.static_Gdev = at;
    at = align(at + (long long)sizeof(double) * pl.n_params);
  }
  pl.static_bytes = at;
  for (const auto& g : pl.groups)
    if (g.big_scratch) pl.static_ok = false;
  int rc;
  if ((rc = pl.upload(&pl.d_static_off, pl.static_off.data(), pl.n_probs))) return rc;
  if ((rc = pl.alloc(&pl.d_probs_static, pl.n_probs))) return rc;
  return PR_OK;
}

// ------------------------------------------------ packing of small problems
// pr_options.pack: problems with one member, dim <= 8 and a uniform slice
// length, of the gate objective or a state-pair one (pairs square / real /
// modulus, state; square and real for Lindblad dynamics), become diagonal
// blocks of one problem of dim <= 16, so that one single-wavefront work item evaluates up to 8 of them
// (a d = 2 problem alone fills 4 of the 256 elements of its 16 x 16 tile).
// - A pack holds problems of one dynamics, objective and dt, at most 16 rows, 16
//   pairs and 64 channels, and runs the longest n_t. Block b's drift is a
//   control channel of its own whose waveform is 1 on the block's own slices
//   and 0 after, where its controls are 0 too: the propagators of those slices
//   are exactly the identity on the block.
// - Exp, products and traces of a block-diagonal matrix act block by block; the
//   kernels differ from an unpacked problem only in the shared scaling
//   exponent and Taylor degree of a slice (the largest of the blocks'), a
//   rounding-level difference.
// - F is per block (k_combine: a gate block's trace, or the sum over its own
//   pairs), and so is a gate pack's adjoint seed; a pair's seed weight is its
//   own already. The pack runs at weight 1 and each problem's weight (its
//   own, or the default 1/n_pairs of the pair objectives) is applied when F
//   and G are scattered back. Control j of block b is channel ch0_b + j, and
//   its gradient row is the problem's. The drift channels come after every
//   block's controls and have no gradient (ncg).
// Packs are filled in order of decreasing n_t, each holding problems whose n_t
// is within 1/16 of its longest, so little padding is run.
struct PackBuild {
  std::vector<pr_problem> dev;               // the plan's problem list
  std::vector<std::vector<double>> store;    // host arrays of the packed problems
  std::vector<std::vector<int>> istore;      // (their checkpoint steps)
  std::vector<int> ncg;                      // per problem: channels with a gradient
  std::vector<std::vector<int>> blk_row;     // per problem: block rows (packed) or empty
  std::vector<std::vector<int>> blk_pair;    // per problem: block pair columns (packed pairs) or empty
  std::vector<std::vector<int>> blk_ch;      // per problem: block control channels (packed) or empty
  std::vector<int> dsrc;                     // per problem: caller's problem, or -1 - first channel in chan
  std::vector<int2> chan;                    // per packed channel: (caller's problem, j or -1)
  std::vector<int4> umap;                    // per caller's problem: (problem, block or -1 or -f, ch0, n_ctrl)
  std::vector<double> uw;                    // per caller's problem: weight applied when unpacking
  std::vector<std::vector<double>> bw;       // per problem: member weights of packed members, or empty
};

// The column-major d x d complex matrix a into the D x D one that starts at
// element base of dst, on rows and columns o .. o + d - 1.
void embed_block(std::vector<double>& dst, size_t base, const double* a, int d, int o, int D) {
  for (int c = 0; c < d; c++)
    for (int r = 0; r < d; r++) {
      const size_t i = (size_t)r + (size_t)c * d, j = base + (size_t)(o + r) + (size_t)(o + c) * D;
      dst[2 * j] = a[2 * i];
      dst[2 * j + 1] = a[2 * i + 1];
    }
}

bool pair_objective(int obj) {
  return obj == PR_PAIRS_SQUARE || obj == PR_PAIRS_REAL || obj == PR_PAIRS_MODULUS || obj == PR_STATE;
}

// closed dynamics: the gate and the pair objectives; Lindblad dynamics (whose
// generators exp and chains treat block by block as well): the pair objectives
// it allows, square and real
bool packable_kind(const pr_problem& q) {
  if (q.dynamics == PR_CLOSED) return q.objective == PR_GATE || pair_objective(q.objective);
  return q.objective == PR_PAIRS_SQUARE || q.objective == PR_PAIRS_REAL;
}

// one-member problems that pack together: those kinds, and tracking problems
// (each block keeps its own checkpoints; F splits by the blocks' initial
// states in k_track_scan)
bool packable(const pr_problem& q) {
  return (packable_kind(q) || q.objective == PR_TRACKING) && q.n_members == 1 && q.dim <= 8 && !q.dt_slices;
}

// Members of an ensemble (M >= 2, dim <= 8, per-slice dt allowed: the members
// share the slices) pack the same way, f to a block problem of
// M' = ceil(M / f) members: block b of member m' is member
// m' f + b, with its drift, control scales and target (or pairs); the
// controls are the ensemble's, f times over; the member weights scale the
// blocks' seeds and F (bw), and the gradient rows of control j in the f blocks
// are summed when G is scattered back. A last member short of f blocks is
// filled with copies of member 0 at weight 0. f = the most blocks of d rows
// (and n_pairs pairs) in 16, at most 64 / n_ctrl and M.
int ens_blocks(const pr_problem& q) {
  if (!packable_kind(q) || q.n_members < 2 || q.dim > 8) return 0;
  int f = std::min({16 / q.dim, 64 / q.n_ctrl, q.n_members});
  if (q.objective != PR_GATE) f = std::min(f, 16 / std::max(1, q.n_pairs));
  return f >= 2 ? f : 0;
}

void pack_members(const pr_problem& q, int p, int f, PackBuild& pk) {
  const int d = q.dim, nc = q.n_ctrl, M = q.n_members, Mp = (M + f - 1) / f, D = f * d, nc2 = f * nc;
  const bool gate = q.objective == PR_GATE;
  const int np = gate ? 0 : q.n_pairs, np2 = f * np;
  const size_t dd = (size_t)d * d, DD = (size_t)D * D;
  const int nh0 = q.H0_per_member ? Mp : 1;
  std::vector<double> H0(2 * DD * nh0, 0.0), Hc(2 * DD * nc2, 0.0), Ut(gate ? 2 * DD : 0, 0.0),
      sc((size_t)Mp * nc2, 1.0), wt(Mp, 1.0), P0(2 * (size_t)D * np2, 0.0), PT(2 * (size_t)D * np2, 0.0);
  std::vector<double> bw((size_t)Mp * f, 0.0);
  auto embed = [&](std::vector<double>& dst, size_t base, const double* a, int o) {
    embed_block(dst, base, a, d, o, D);
  };
  const double w0 = gate ? 1.0 / M : 1.0 / ((double)M * q.n_pairs);
  const int c0 = (int)pk.chan.size();
  for (int b = 0; b < f; b++) {
    const int o = b * d;
    for (int j = 0; j < nc; j++) {
      embed(Hc, DD * (b * nc + j), q.Hc + 2 * dd * j, o);
      pk.chan.push_back(make_int2(p, j));
    }
    if (gate) embed(Ut, 0, q.U_target, o);
    for (int t = 0; t < np; t++)
      for (int r = 0; r < d; r++)
        for (int k = 0; k < 2; k++) {
          P0[2 * ((size_t)(b * np + t) * D + o + r) + k] = q.psi_init[2 * ((size_t)t * d + r) + k];
          PT[2 * ((size_t)(b * np + t) * D + o + r) + k] = q.psi_targ[2 * ((size_t)t * d + r) + k];
        }
    for (int mp = 0; mp < Mp; mp++) {
      const int m = mp * f + b, src = m < M ? m : 0;  // (a filling block copies member 0)
      if (q.H0_per_member || mp == 0)
        embed(H0, q.H0_per_member ? DD * mp : 0, q.H0 + (q.H0_per_member ? 2 * dd * src : 0), o);
      for (int j = 0; j < nc; j++)
        sc[(size_t)mp * nc2 + b * nc + j] = q.ctrl_scale ? q.ctrl_scale[(size_t)src * nc + j] : 1.0;
      bw[(size_t)mp * f + b] = m < M ? (q.weight ? q.weight[m] : w0) : 0.0;
    }
  }
  pr_problem z = q;
  z.dim = D;
  z.n_ctrl = nc2;
  z.n_members = Mp;
  auto keep = [&](std::vector<double>& v) -> const double* {
    pk.store.push_back(std::move(v));
    return pk.store.back().data();
  };
  z.H0 = keep(H0);
  z.Hc = keep(Hc);
  z.ctrl_scale = keep(sc);
  z.weight = keep(wt);
  if (gate) {
    z.U_target = keep(Ut);
  } else {
    z.n_pairs = np2;
    z.psi_init = keep(P0);
    z.psi_targ = keep(PT);
  }
  std::vector<int> rows(f + 1), pcols(f + 1);
  for (int b = 0; b <= f; b++) {
    rows[b] = b * d;
    pcols[b] = b * np;
  }
  pk.umap[p] = make_int4((int)pk.dev.size(), -f, 0, nc);
  pk.dev.push_back(z);
  pk.ncg.push_back(nc2);
  pk.blk_row.push_back(rows);
  pk.blk_pair.push_back(gate ? std::vector<int>() : pcols);
  std::vector<int> chs(f + 1);
  for (int b = 0; b <= f; b++) chs[b] = b * nc;
  pk.blk_ch.push_back(chs);
  pk.bw.push_back(bw);
  pk.dsrc.push_back(-1 - c0);
}

// Tracking ensembles (M >= 2, dim <= 8, one weight for every member) pack f
// members to a block member, f the largest divisor of M with f d <= 16,
// f n_pairs <= 16 and f n_ctrl <= 64: the checkpoint targets become
// blockdiag(rho_c, ..., rho_c), so a packed member's F is the sum of its
// blocks' (the tracking walk needs nothing per block) and the member weight
// carries over; the gradient rows of the f blocks are summed when G is
// scattered back, as for the other packed ensembles.
int track_blocks(const pr_problem& q) {
  if (q.objective != PR_TRACKING || q.n_members < 2 || q.dim > 8) return 0;
  if (q.weight)
    for (int m = 1; m < q.n_members; m++)
      if (q.weight[m] != q.weight[0]) return 0;
  int best = 0;
  for (int f = 2; f <= q.n_members; f++)
    if (q.n_members % f == 0 && f * q.dim <= 16 && f * q.n_pairs <= 16 && f * q.n_ctrl <= 64) best = f;
  return best;
}

void pack_tracking_members(const pr_problem& q, int p, int f, PackBuild& pk) {
  const int d = q.dim, nc = q.n_ctrl, M = q.n_members, Mp = M / f, D = f * d, nc2 = f * nc, np = q.n_pairs;
  const size_t dd = (size_t)d * d, DD = (size_t)D * D;
  const int nh0 = q.H0_per_member ? Mp : 1;
  std::vector<double> H0(2 * DD * nh0, 0.0), Hc(2 * DD * nc2, 0.0), sc((size_t)Mp * nc2, 1.0),
      P0(2 * (size_t)D * f * np, 0.0), TT(2 * DD * q.n_track, 0.0), wt(Mp, q.weight ? q.weight[0] : 0.0);
  auto embed = [&](std::vector<double>& dst, size_t base, const double* a, int o) {
    embed_block(dst, base, a, d, o, D);
  };
  const int c0 = (int)pk.chan.size();
  for (int b = 0; b < f; b++) {
    const int o = b * d;
    for (int j = 0; j < nc; j++) {
      embed(Hc, DD * (b * nc + j), q.Hc + 2 * dd * j, o);
      pk.chan.push_back(make_int2(p, j));
    }
    for (int c = 0; c < q.n_track; c++) embed(TT, DD * c, q.track_target + 2 * dd * c, o);
    for (int t = 0; t < np; t++)
      for (int r = 0; r < d; r++)
        for (int k = 0; k < 2; k++)
          P0[2 * ((size_t)(b * np + t) * D + o + r) + k] = q.psi_init[2 * ((size_t)t * d + r) + k];
    for (int mp = 0; mp < Mp; mp++) {
      const int m = mp * f + b;
      if (q.H0_per_member || mp == 0)
        embed(H0, q.H0_per_member ? DD * mp : 0, q.H0 + (q.H0_per_member ? 2 * dd * m : 0), o);
      for (int j = 0; j < nc; j++)
        sc[(size_t)mp * nc2 + b * nc + j] = q.ctrl_scale ? q.ctrl_scale[(size_t)m * nc + j] : 1.0;
    }
  }
  pr_problem z = q;
  z.dim = D;
  z.n_ctrl = nc2;
  z.n_members = Mp;
  z.n_pairs = f * np;
  z.psi_targ = nullptr;
  auto keep = [&](std::vector<double>& v) -> const double* {
    pk.store.push_back(std::move(v));
    return pk.store.back().data();
  };
  z.H0 = keep(H0);
  z.Hc = keep(Hc);
  z.ctrl_scale = keep(sc);
  z.psi_init = keep(P0);
  z.track_target = keep(TT);
  // (the default weights, 1/(M n_pairs), are the packed ensemble's as well)
  z.weight = q.weight ? keep(wt) : nullptr;
  pk.umap[p] = make_int4((int)pk.dev.size(), -f, 0, nc);
  pk.dev.push_back(z);
  pk.ncg.push_back(nc2);
  pk.blk_row.emplace_back();
  pk.blk_pair.emplace_back();
  pk.blk_ch.emplace_back();
  pk.bw.emplace_back();
  pk.dsrc.push_back(-1 - c0);
}

// false when nothing is packed
bool build_packs(const pr_problem* probs, int n_probs, PackBuild& pk) {
  std::vector<int> cand;
  bool any_ens = false;
  for (int p = 0; p < n_probs; p++) {
    if (packable(probs[p])) cand.push_back(p);
    if (ens_blocks(probs[p]) || track_blocks(probs[p])) any_ens = true;
  }
  if (cand.size() < 2 && !any_ens) return false;
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) {
    if (probs[a].dynamics != probs[b].dynamics) return probs[a].dynamics < probs[b].dynamics;
    if (probs[a].objective != probs[b].objective) return probs[a].objective < probs[b].objective;
    if (probs[a].dt != probs[b].dt) return probs[a].dt < probs[b].dt;
    return probs[a].n_t > probs[b].n_t;
  });
  struct Bin {
    std::vector<int> mem;
    int rows = 0, pairs = 0, chans = 0, nt = 0, obj = 0, dyn = 0;
    double dt = 0;
  };
  std::vector<Bin> bins;
  std::vector<size_t> open;  // indices of bins still taking problems (at most 8)
  for (int p : cand) {
    const pr_problem& q = probs[p];
    const int np = q.objective == PR_GATE ? 0 : q.n_pairs;
    int hit = -1;
    for (size_t i = 0; i < open.size() && hit < 0; i++) {
      const Bin& b = bins[open[i]];
      if (b.dyn == q.dynamics && b.obj == q.objective && b.dt == q.dt && b.rows + q.dim <= 16 && b.pairs + np <= 16 &&
          b.chans + q.n_ctrl + 1 <= 64 && b.nt - q.n_t <= b.nt / 16)
        hit = (int)i;
    }
    if (hit < 0) {
      if (open.size() == 8) open.erase(open.begin());
      bins.emplace_back();
      bins.back().nt = q.n_t;
      bins.back().dt = q.dt;
      bins.back().obj = q.objective;
      bins.back().dyn = q.dynamics;
      open.push_back(bins.size() - 1);
      hit = (int)open.size() - 1;
    }
    Bin& b = bins[open[hit]];
    b.mem.push_back(p);
    b.rows += q.dim;
    b.pairs += np;
    b.chans += q.n_ctrl + 1;
    if (b.rows == 16 || b.pairs == 16) open.erase(open.begin() + hit);
  }
  std::vector<int> pack_of(n_probs, -1);
  int n_packs = 0;
  for (size_t i = 0; i < bins.size(); i++)
    if (bins[i].mem.size() > 1) {
      for (int p : bins[i].mem) pack_of[p] = (int)i;
      n_packs++;
    }
  if (n_packs == 0 && !any_ens) return false;
  pk.umap.assign(n_probs, make_int4(0, -1, 0, 0));
  pk.uw.assign(n_probs, 1.0);
  // unpacked problems and packed ensembles first, in the caller's order, then
  // the packs of single problems
  for (int p = 0; p < n_probs; p++) {
    if (pack_of[p] >= 0) continue;
    if (const int f = ens_blocks(probs[p])) {
      pack_members(probs[p], p, f, pk);
      continue;
    }
    if (const int f = track_blocks(probs[p])) {
      pack_tracking_members(probs[p], p, f, pk);
      continue;
    }
    pk.umap[p] = make_int4((int)pk.dev.size(), -1, 0, probs[p].n_ctrl);
    pk.dev.push_back(probs[p]);
    pk.ncg.push_back(probs[p].n_ctrl);
    pk.blk_row.emplace_back();
    pk.blk_pair.emplace_back();
    pk.blk_ch.emplace_back();
    pk.bw.emplace_back();
    pk.dsrc.push_back(p);
  }
  for (const Bin& bn : bins) {
    if (bn.mem.size() < 2) continue;
    const int dp = (int)pk.dev.size();
    const bool gate = bn.obj == PR_GATE, tracking = bn.obj == PR_TRACKING;
    int D = 0, ncg = 0, npr = 0, nck = 0;
    for (int p : bn.mem) {
      D += probs[p].dim;
      ncg += probs[p].n_ctrl;
      npr += gate ? 0 : probs[p].n_pairs;
      nck += tracking ? probs[p].n_track : 0;
    }
    const int nb = (int)bn.mem.size(), nc = ncg + nb;
    const size_t DD = (size_t)D * D;
    std::vector<double> H0(2 * DD, 0.0), Hc(2 * DD * nc, 0.0), Ut(gate ? 2 * DD : 0, 0.0), sc(nc, 1.0),
        P0(2 * (size_t)D * npr, 0.0), PT(2 * (size_t)D * npr, 0.0), w1(1, 1.0), TT(2 * DD * nck, 0.0),
        TW(nck, 0.0);
    std::vector<int> TS(nck, 1);
    int ck0 = 0;  // checkpoints placed so far
    std::vector<int> rows(1, 0), pcols(1, 0), chs(1, 0);
    auto embed = [&](std::vector<double>& dst, size_t base, const double* a, int d, int o) {
      embed_block(dst, base, a, d, o, D);
    };
    int o = 0, ch0 = 0, s0 = 0;
    for (int b = 0; b < nb; b++) {
      const int p = bn.mem[b];
      const pr_problem& q = probs[p];
      const int d = q.dim;
      const size_t dd = (size_t)d * d;
      for (int j = 0; j < q.n_ctrl; j++) {
        embed(Hc, DD * (ch0 + j), q.Hc + 2 * dd * j, d, o);
        if (q.ctrl_scale) sc[ch0 + j] = q.ctrl_scale[j];
        pk.chan.push_back(make_int2(p, j));
      }
      embed(Hc, DD * (ncg + b), q.H0, d, o);
      if (gate) {
        embed(Ut, 0, q.U_target, d, o);
        pk.uw[p] = q.weight ? q.weight[0] : 1.0;
      } else {
        // pair s of the block: vectors [n_pairs][d] at rows o .. o + d - 1
        // (tracking: the initial states only)
        for (int t = 0; t < q.n_pairs; t++)
          for (int r = 0; r < d; r++)
            for (int k = 0; k < 2; k++) {
              P0[2 * ((size_t)(s0 + t) * D + o + r) + k] = q.psi_init[2 * ((size_t)t * d + r) + k];
              if (!tracking) PT[2 * ((size_t)(s0 + t) * D + o + r) + k] = q.psi_targ[2 * ((size_t)t * d + r) + k];
            }
        s0 += q.n_pairs;
        pk.uw[p] = q.weight ? q.weight[0] : 1.0 / q.n_pairs;
        // the block's checkpoints, their targets on its rows
        for (int c = 0; tracking && c < q.n_track; c++, ck0++) {
          TS[ck0] = q.track_step[c];
          TW[ck0] = q.track_weight[c];
          embed(TT, DD * ck0, q.track_target + 2 * dd * c, d, o);
        }
      }
      pk.umap[p] = make_int4(dp, b, ch0, q.n_ctrl);
      o += d;
      ch0 += q.n_ctrl;
      rows.push_back(o);
      pcols.push_back(s0);
      chs.push_back(ch0);
    }
    for (int b = 0; b < nb; b++) pk.chan.push_back(make_int2(bn.mem[b], -1));
    pr_problem z = probs[bn.mem[0]];
    z.dim = D;
    z.n_ctrl = nc;
    z.n_t = bn.nt;
    z.dt = bn.dt;
    z.dt_slices = nullptr;
    z.n_members = 1;
    z.H0_per_member = 0;
    auto keep = [&](std::vector<double>& v) -> const double* {
      pk.store.push_back(std::move(v));
      return pk.store.back().data();
    };
    z.H0 = keep(H0);
    z.Hc = keep(Hc);
    z.ctrl_scale = keep(sc);
    z.weight = keep(w1);
    if (gate) {
      z.U_target = keep(Ut);
    } else {
      z.n_pairs = npr;
      z.psi_init = keep(P0);
      z.psi_targ = tracking ? nullptr : keep(PT);
    }
    if (tracking) {
      z.n_track = nck;
      pk.istore.push_back(std::move(TS));
      z.track_step = pk.istore.back().data();
      z.track_weight = keep(TW);
      z.track_target = keep(TT);
    }
    pk.dev.push_back(z);
    pk.ncg.push_back(ncg);
    pk.blk_row.push_back(rows);
    pk.blk_pair.push_back(gate ? std::vector<int>() : pcols);
    pk.blk_ch.push_back(chs);
    pk.bw.emplace_back();
    pk.dsrc.push_back(-1 - ((int)pk.chan.size() - nc));
  }
  return true;
}

int build_plan(Plan& pl, const pr_problem* probs, int n_probs, const pr_problem* uprobs, int n_user,
               const PackBuild* pk) {
  pl.n_probs = n_probs;
  pl.host_probs.resize(n_probs);
  pl.static_sz.resize(n_probs);
  pl.coff.assign(n_probs + 1, 0);
  // groups: index B (1..8) by block count; tracking problems, which run their
  // own walk kernel, in groups 8 + B (B <= 4)
  auto group_of = [&](const pr_problem& q) { return (q.objective == PR_TRACKING ? 8 : 0) + (q.dim + 15) / 16; };
  std::vector<std::vector<pr::WorkItem>> items(13);
  std::vector<std::vector<int2>> members(13);
  std::vector<int> vmax(13, 1);
  std::vector<int> tr_nnz_max(13, 0);  // most nonzeros of one control, per group
  std::vector<int> nc_max(13, 0);
  std::vector<double> group_work(13, 0.0);  // member-slices per group
  std::vector<int> group_members(13, 0);
  std::vector<char> group_lind(13, 0);
  for (int p = 0; p < n_probs; p++) {
    const int gi = group_of(probs[p]);
    group_work[gi] += (double)probs[p].n_members * probs[p].n_t;
    group_members[gi] += probs[p].n_members;
    if (probs[p].dynamics != PR_CLOSED) group_lind[gi] = 1;
  }
  // two-level prefix (k_combine_local): groups (B <= 4) of closed dynamics
  // whose members cannot fill the CUs (< 128 of them), for problems of 16 or
  // more segments (64 for single-wavefront products, which are short against
  // the two extra launches), in chunks of about sqrt(S) segments
  std::vector<std::vector<pr::WorkItem>> citems(13);
  // the mixed-precision kernels cover single-wavefront (d <= 16) unitary
  // problems under the reference gradient formula; any other problem of the
  // plan, or a d <= 16 group holding a Lindblad problem, runs in FP64
  bool mixed1 = (pl.opt.precision == PR_MIXED || pl.opt.precision == PR_SPLIT16) &&
                pl.opt.grad_mode == PR_GRAD_REFERENCE;
  for (int p = 0; p < n_probs; p++)
    if (probs[p].dim <= 16 && probs[p].dynamics != PR_CLOSED) mixed1 = false;
  // (tracking problems run FP64 in their own groups, with stored images)
  // U_k images between the passes (store_props): closed-dynamics FP64 problems
  // can recompute U_k in the gradient walk instead (d > 64: from checkpoints
  // every 32 slices at most). Auto (-1) recomputes when all images would take
  // more than half of the free HBM.
  // (a d <= 16 closed problem sharing its group with a Lindblad problem runs
  // the general segment kernel, whose U_k image the recomputation would not
  // reproduce bit for bit: it keeps its images)
  bool lindblad1 = false;
  for (int p = 0; p < n_probs; p++)
    if (probs[p].dim <= 16 && probs[p].dynamics != PR_CLOSED) lindblad1 = true;
  // Lindblad problems (d <= 64) recompute too: their gradient walk keeps the
  // forward states and recomputed U_k of a bounded number of work items at a
  // time (segments of at most LIND_RECOMPUTE_SEG slices, at most
  // lind_window(B) work items per launch).
  // The reduced-precision forms (PR_MIXED, PR_SPLIT16) rebuild their D_k
  // images in the walk as well.
  auto can_recompute = [&](const pr_problem& q) {
    if (q.dynamics != PR_CLOSED) return q.objective != PR_TRACKING && q.dim <= 64;
    if (q.objective == PR_TRACKING) return q.dim <= 64;
    return !(q.dim <= 16 && lindblad1);
  };
  bool recompute = pl.opt.store_props == 0;
  if (pl.opt.store_props < 0) {
    // images, and the forward states of Lindblad problems (as large again)
    double bytes = 0;
    for (int p = 0; p < n_probs; p++) {
      const double nb = (double)((probs[p].dim + 15) / 16) * ((probs[p].dim + 15) / 16);
      bytes += (probs[p].dynamics == PR_CLOSED ? 1.0 : 2.0) * (double)probs[p].n_members * probs[p].n_t * nb * 256 *
               sizeof(double2);
    }
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    recompute = bytes > 0.5 * (double)free_b;
  }
  for (int p = 0; p < n_probs; p++) {
    const pr_problem& q = probs[p];
    pr::ProbDev& D = pl.host_probs[p];
    memset(&D, 0, sizeof D);
    const int d = q.dim, B = (d + 15) / 16, NB = B * B, M = q.n_members, nc = q.n_ctrl, nt = q.n_t;
    const int gi = group_of(q);
    const bool track = q.objective == PR_TRACKING;
    const size_t MB = (size_t)NB * 256, dd = (size_t)d * d;
    D.d = d; D.B = B; D.nc = nc; D.nt = nt; D.M = M; D.dyn = q.dynamics; D.obj = q.objective;
    D.ncg = pk ? pk->ncg[p] : nc;
    if (pk && !pk->blk_row[p].empty()) {
      D.nblk = (int)pk->blk_row[p].size() - 1;
      std::copy(pk->blk_row[p].begin(), pk->blk_row[p].end(), D.blk_row);
      std::copy(pk->blk_pair[p].begin(), pk->blk_pair[p].end(), D.blk_pair);
      std::copy(pk->blk_ch[p].begin(), pk->blk_ch[p].end(), D.blk_ch);
    }
    if (pk && !pk->bw[p].empty()) {
      double* bwp;
      int rcb = pl.upload(&bwp, pk->bw[p].data(), pk->bw[p].size());
      if (rcb) return rcb;
      D.ens = 1;
      D.bw = bwp;
    }
    const bool gate = q.objective == PR_GATE || q.objective == PR_GATE_PROPAGATORS;
    D.npairs = gate ? 0 : q.n_pairs;
    D.h0pm = q.H0_per_member ? 1 : 0;
    D.unitary = q.dynamics == PR_CLOSED ? 1 : 0;
    D.grad_exact = pl.opt.grad_mode == PR_GRAD_EXACT;
    // compute_grape_gradient's first-order term sits after U_k; the MR kernels'
    // in place of it
    D.after = q.dynamics == PR_CLOSED &&
              (q.objective == PR_GATE || q.objective == PR_GATE_PROPAGATORS || q.objective == PR_STATE ||
               q.objective == PR_STATE_PROPAGATORS);
    D.cre = q.dynamics == PR_CLOSED ? 0.0 : 1.0;
    D.cim = q.dynamics == PR_CLOSED ? -1.0 : 0.0;
    // time segmentation: about 8192 wavefronts of work over all problems of
    // the same block count, every work item about the same number of slices
    // (so problems of different lengths do not leave a tail of long items)
    int L;
    if (pl.opt.segment > 0) {
      L = pl.opt.segment;
    } else {
      // (d > 64: workgroups of 16 wavefronts, about 512 of them)
      long long want = B > 4 ? 512 : std::max(1LL, (long long)(8192 / NB));
      const double per_item = std::ceil(group_work[gi] / (double)want);
      L = (int)std::min<double>(std::max<double>(std::round(per_item * pl.seg_scale[gi]), 1.0), (double)nt);
      L = std::max(L, std::min(nt, 4));
      // tracking and Lindblad problems chain their segments one after another
      // in a per-member scan (k_track_scan) or backward pass (k_combine): with
      // few members the rule's short segments would make that chain the
      // evaluation's longest part, so their segments are at least sqrt(n_t)
      // long (a 2000-slice tracking problem of few members: about 1.0 -> 0.33 ms)
      if (track || q.dynamics != PR_CLOSED) L = std::max(L, (int)std::ceil(std::sqrt((double)nt)));
      // d > 64 without stored images: the segment starts E_s are the
      // checkpoints the gradient walk recomputes from, every 32 slices at most;
      // Lindblad likewise, which bounds each work item's scratch slot
      if ((B > 4 || q.dynamics != PR_CLOSED) && recompute && can_recompute(q)) L = std::min(L, LIND_RECOMPUTE_SEG);
    }
    L = std::min(L, nt);
    const int S = (nt + L - 1) / L;
    D.nseg = S; D.seglen = L;
    // operators; closed dynamics use the Hermitian part of the upper triangle
    // (validated above to 1e-10), so generators -i dt H are exactly anti-Hermitian
    const int nh0 = D.h0pm ? M : 1;
    std::vector<double> h0h, hch;
    const double* H0src = q.H0;
    const double* Hcsrc = q.Hc;
    if (q.dynamics == PR_CLOSED) {
      h0h = hermitian_upper(q.H0, nh0, d);
      hch = hermitian_upper(q.Hc, nc, d);
    } else {
      h0h.assign(q.H0, q.H0 + 2 * dd * nh0);
      hch.assign(q.Hc, q.Hc + 2 * dd * nc);
    }
    // with a uniform slice length the assembly operators carry the factor
    // c dt (c = -i closed, 1 Lindblad): X = c dt A0 + sum_j s u_j (c dt A_j)
    D.prescaled = q.dt_slices ? 0 : 1;
    D.dt0 = q.dt;
    if (D.prescaled) {
      const double fr = D.cre * q.dt, fi = D.cim * q.dt;
      for (auto* v : {&h0h, &hch})
        for (size_t i = 0; i < v->size(); i += 2) {
          const double a = (*v)[i], b = (*v)[i + 1];
          (*v)[i] = fr * a - fi * b;
          (*v)[i + 1] = fr * b + fi * a;
        }
    }
    H0src = h0h.data();
    Hcsrc = hch.data();
    std::vector<double2> img(MB * std::max(nh0, nc));
    std::vector<double> nrm(std::max(nh0, nc));
    for (int m = 0; m < nh0; m++) {
      tile_image(H0src + 2 * dd * m, d, B, false, img.data() + MB * m);
      nrm[m] = norm1(H0src + 2 * dd * m, d);
    }
    int rc;
    double2* dp;
    double* dd_;
    if ((rc = pl.upload(&dp, img.data(), MB * nh0))) return rc;
    D.H0 = dp;
    if (B == 1 && mixed1 && !track) {
      for (int m = 0; m < nh0; m++) tile_image(H0src + 2 * dd * m, d, 1, false, img.data() + MB * m, true);
      if ((rc = pl.upload(&dp, img.data(), MB * nh0))) return rc;
      D.H0o = dp;
      Csr co = build_csr(Hcsrc, nc, d, 1, false, true);
      int *ip, *jp;
      double2* vp;
      if ((rc = pl.upload(&ip, co.ptr.data(), co.ptr.size()))) return rc;
      if ((rc = pl.upload(&jp, co.j.data(), co.j.size()))) return rc;
      if ((rc = pl.upload(&vp, co.v.data(), co.v.size()))) return rc;
      D.co_ptr = ip; D.co_j = jp; D.co_v = vp;
      if (nc <= 64) {
        Slots so = build_slots(co, 1);
        int4* sp;
        if ((rc = pl.upload(&sp, so.idx.data(), so.idx.size()))) return rc;
        if ((rc = pl.upload(&vp, so.v.data(), so.v.size()))) return rc;
        D.so_idx = sp; D.so_v = vp;
        std::copy(so.ptr.begin(), so.ptr.end(), D.so_pass);
      }
    }
    if (B == 1 && q.dynamics == PR_CLOSED) {
      // transposed drift images and operator tables: k_segment assembles X^T
      for (int m = 0; m < nh0; m++) tile_image(H0src + 2 * dd * m, d, 1, true, img.data() + MB * m);
      if ((rc = pl.upload(&dp, img.data(), MB * nh0))) return rc;
      D.H0t = dp;
      Csr ct = build_csr(Hcsrc, nc, d, 1, true);
      int *ip, *jp;
      double2* vp;
      if ((rc = pl.upload(&ip, ct.ptr.data(), ct.ptr.size()))) return rc;
      if ((rc = pl.upload(&jp, ct.j.data(), ct.j.size()))) return rc;
      if ((rc = pl.upload(&vp, ct.v.data(), ct.v.size()))) return rc;
      D.ct_ptr = ip; D.ct_j = jp; D.ct_v = vp;
      if (nc <= 64) {
        Slots st = build_slots(ct, 1);
        int4* sp;
        if ((rc = pl.upload(&sp, st.idx.data(), st.idx.size()))) return rc;
        if ((rc = pl.upload(&vp, st.v.data(), st.v.size()))) return rc;
        D.slt_idx = sp; D.slt_v = vp;
        std::copy(st.ptr.begin(), st.ptr.end(), D.slt_pass);
      }
    }
    if ((rc = pl.upload(&dd_, nrm.data(), nh0))) return rc;
    D.n0 = dd_;
    for (int j = 0; j < nc; j++) nrm[j] = norm1(Hcsrc + 2 * dd * j, d);
    if ((rc = pl.upload(&dd_, nrm.data(), nc))) return rc;
    D.ncn = dd_;
    {
      Csr cs = build_csr(Hcsrc, nc, d, B, false);
      int *ip, *jp;
      double2* vp;
      if ((rc = pl.upload(&ip, cs.ptr.data(), cs.ptr.size()))) return rc;
      if ((rc = pl.upload(&jp, cs.j.data(), cs.j.size()))) return rc;
      if ((rc = pl.upload(&vp, cs.v.data(), cs.v.size()))) return rc;
      D.cs_ptr = ip; D.cs_j = jp; D.cs_v = vp;
      if (nc <= 64) {
        Slots sl = build_slots(cs, NB);
        int4* sp;
        if ((rc = pl.upload(&sp, sl.idx.data(), sl.idx.size()))) return rc;
        if ((rc = pl.upload(&vp, sl.v.data(), sl.v.size()))) return rc;
        D.sl_idx = sp; D.sl_v = vp;
        std::copy(sl.ptr.begin(), sl.ptr.end(), D.sl_pass);
      }
      // the gradient contracts the operators as given (Gradients.jl:274-289)
      Trace tr = build_trace(q.Hc, nc, d);
      // the contraction needs Re(c Tr(A_j M)): Im z for closed dynamics
      // (c = -i), Re z for Lindblad (c = 1). Each entry a = A_j[a][b] is stored
      // as (Im a, Re a) or (Re a, -Im a), so that term = v.x Re m + v.y Im m.
      for (auto& v : tr.v)
        v = q.dynamics == PR_CLOSED ? make_double2(v.y, v.x) : make_double2(v.x, -v.y);
      if ((rc = pl.upload(&ip, tr.ptr.data(), tr.ptr.size()))) return rc;
      if ((rc = pl.upload(&jp, tr.idx.data(), tr.idx.size()))) return rc;
      if ((rc = pl.upload(&vp, tr.v.data(), tr.v.size()))) return rc;
      D.tr_ptr = ip; D.tr_idx = jp; D.tr_v = vp;
      for (int j = 0; j < D.ncg; j++) tr_nnz_max[gi] = std::max(tr_nnz_max[gi], tr.ptr[j + 1] - tr.ptr[j]);
    }
    std::vector<double> sc((size_t)M * nc, 1.0);
    if (q.ctrl_scale) std::copy(q.ctrl_scale, q.ctrl_scale + (size_t)M * nc, sc.begin());
    if ((rc = pl.upload(&dd_, sc.data(), sc.size()))) return rc;
    D.scale = dd_;
    std::vector<double> wt(M, gate ? 1.0 / M : 1.0 / ((double)M * q.n_pairs));
    if (q.weight) std::copy(q.weight, q.weight + M, wt.begin());
    if ((rc = pl.upload(&dd_, wt.data(), M))) return rc;
    D.weight = dd_;
    std::vector<double> dts(nt, q.dt);
    if (q.dt_slices) std::copy(q.dt_slices, q.dt_slices + nt, dts.begin());
    if ((rc = pl.upload(&dd_, dts.data(), nt))) return rc;
    D.dts = dd_;
    if (gate) {
      tile_image(q.U_target, d, B, false, img.data());
      if ((rc = pl.upload(&dp, img.data(), MB))) return rc;
      D.Ut = dp;
    } else {
      tile_columns(q.psi_init, q.n_pairs, d, B, img.data());
      if ((rc = pl.upload(&dp, img.data(), MB))) return rc;
      D.Psi0 = dp;
      if (!track) {
        tile_columns(q.psi_targ, q.n_pairs, d, B, img.data());
        if ((rc = pl.upload(&dp, img.data(), MB))) return rc;
        D.PsiT = dp;
      }
    }
    if (track) {
      // one checkpoint matrix per distinct step: C = sum of w (rho + rho^dag) / 2
      std::vector<int> ck(nt, -1);
      std::vector<std::vector<cd>> cm;
      for (int c = 0; c < q.n_track; c++) {
        const int k = q.track_step[c] - 1;
        if (ck[k] < 0) {
          ck[k] = (int)cm.size();
          cm.emplace_back(dd, cd(0, 0));
        }
        const double w = q.track_weight[c];
        const double* t = q.track_target + 2 * dd * c;
        auto& C = cm[ck[k]];
        for (int col = 0; col < d; col++)
          for (int row = 0; row < d; row++)
            C[row + (size_t)col * d] += 0.5 * w * (at(t, d, row, col) + std::conj(at(t, d, col, row)));
      }
      std::vector<double2> cimg(MB * cm.size());
      for (size_t c = 0; c < cm.size(); c++)
        tile_image(reinterpret_cast<const double*>(cm[c].data()), d, B, false, cimg.data() + MB * c);
      int* ip;
      if ((rc = pl.upload(&ip, ck.data(), ck.size()))) return rc;
      D.ck_of = ip;
      if ((rc = pl.upload(&dp, cimg.data(), cimg.size()))) return rc;
      D.Ck = dp;
    }
    // scratch (the byte size of each buffer is recorded for the static layout)
    pr::StaticOff& z = pl.static_sz[p];
    z = pr::StaticOff{0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#define PR_SCRATCH(field, count)                                              \
  do {                                                                        \
    const size_t cnt_ = (size_t)(count);                                      \
    if ((rc = pl.alloc(&D.field, cnt_))) return rc;                           \
    z.field = (long long)(cnt_ * sizeof(*D.field));                           \
  } while (0)
    D.recompute = recompute && can_recompute(q) ? 1 : 0;
    if (!D.recompute) PR_SCRATCH(Ust, (size_t)M * nt * MB);
    if (B == 1 && mixed1 && !track) {
      // FP32 (or split FP16) images of small-generator slices; the FP64 images
      // above hold the others. Without images the split walk rebuilds them.
      if (!D.recompute) PR_SCRATCH(Udm, (size_t)M * nt * 256);
      PR_SCRATCH(big, (size_t)M * nt);
      PR_SCRATCH(seg_big, (size_t)M * S);
    }
    PR_SCRATCH(Tseg, (size_t)M * S * MB);
    PR_SCRATCH(Eseg, (size_t)M * (S + 1) * MB);
    if (B <= 4 && !track && !group_lind[gi] && group_members[gi] < PR_PREFIX_MAX_MEMBERS &&
        S >= (B == 1 ? PR_PREFIX_MIN_SEG1 : 16)) {
      D.ck = (int)std::ceil(std::sqrt((double)S));
      D.nck = (S + D.ck - 1) / D.ck;
      PR_SCRATCH(Qc, (size_t)M * D.nck * MB);
      for (int m = 0; m < M; m++)
        for (int c = 0; c < D.nck; c++) {
          pr::WorkItem w;
          w.prob = p; w.member = m; w.seg = c; w.k0 = c * D.ck; w.k1 = std::min(S, (c + 1) * D.ck);
          citems[gi].push_back(w);
        }
    }
    PR_SCRATCH(Om, (size_t)M * MB);
    if (!D.unitary) {
      PR_SCRATCH(Bb, (size_t)M * S * MB);
      if (D.recompute) {
        // forward states and recomputed U_k^T of lind_window(B) work items
        D.chunk = std::min(M * S, lind_window(B));
        PR_SCRATCH(Pscr, (size_t)D.chunk * 2 * L * MB);
      } else {
        D.chunk = M * S;
        PR_SCRATCH(Pscr, (size_t)M * S * L * MB);
      }
    }
    PR_SCRATCH(Fm, std::max(M, D.nblk));
    PR_SCRATCH(Gm, (size_t)M * nc * nt);
    // (the drift channels of a packed problem have no gradient: their rows stay 0)
    if (D.ncg < nc) HIPCHK(hipMemset(D.Gm, 0, sizeof(double) * M * nc * nt));
    PR_SCRATCH(udef, M);
    HIPCHK(hipMemset(D.udef, 0, sizeof(double) * M));
    PR_SCRATCH(smax, (size_t)4 * M * S);
    if (track) {
      PR_SCRATCH(Zseg, (size_t)M * S * MB);
      PR_SCRATCH(Pend, (size_t)M * S * MB);
      PR_SCRATCH(Lend, (size_t)M * S * MB);
    }
#undef PR_SCRATCH
    D.coff = pl.coff[p];
    pl.coff[p + 1] = pl.coff[p] + (long long)nc * nt;
    for (int m = 0; m < M; m++) {
      members[gi].push_back(make_int2(p, m));
      for (int s = 0; s < S; s++) {
        pr::WorkItem w;
        w.prob = p; w.member = m; w.seg = s; w.k0 = s * L; w.k1 = std::min(nt, (s + 1) * L);
        items[gi].push_back(w);
      }
    }
    vmax[gi] = std::max(vmax[gi], next_pow2_cap32(D.ncg));
    nc_max[gi] = std::max(nc_max[gi], D.ncg);
  }
  pl.n_params = pl.coff[n_probs];
  // the caller's problems
  pl.n_user = n_user;
  pl.ucoff.assign(n_user + 1, 0);
  for (int p = 0; p < n_user; p++) pl.ucoff[p + 1] = pl.ucoff[p] + (long long)uprobs[p].n_ctrl * uprobs[p].n_t;
  pl.n_uparams = pl.ucoff[n_user];
  pl.packed = pk != nullptr;
  int rc;
  if ((rc = pl.upload(&pl.d_probs, pl.host_probs.data(), n_probs))) return rc;
  if ((rc = pl.upload(&pl.d_off, pl.coff.data(), n_probs + 1))) return rc;
  // per 256 controls, the problem of the first one (and n - 1 at the end)
  auto first_table = [](const std::vector<long long>& off, int n) {
    const long long total = off[n], nb = (total + 255) / 256;
    std::vector<int> f(nb + 1, n - 1);
    int p = 0;
    for (long long b = 0; b < nb; b++) {
      while (p + 1 < n && off[p + 1] <= 256 * b) p++;
      f[b] = p;
    }
    return f;
  };
  {
    const std::vector<int> f = first_table(pl.coff, n_probs);
    if ((rc = pl.upload(&pl.d_first, f.data(), f.size()))) return rc;
  }
  if ((rc = pl.alloc(&pl.d_controls, pl.n_uparams))) return rc;
  if ((rc = pl.alloc(&pl.d_F, n_user))) return rc;
  if ((rc = pl.alloc(&pl.d_G, pl.n_uparams))) return rc;
  if (pk) {
    std::vector<int> unt(n_user);
    for (int p = 0; p < n_user; p++) unt[p] = uprobs[p].n_t;
    pl.umap = pk->umap;
    if ((rc = pl.upload(&pl.d_uoff, pl.ucoff.data(), n_user + 1))) return rc;
    const std::vector<int> uf = first_table(pl.ucoff, n_user);
    if ((rc = pl.upload(&pl.d_ufirst, uf.data(), uf.size()))) return rc;
    if ((rc = pl.upload(&pl.d_unt, unt.data(), n_user))) return rc;
    if ((rc = pl.upload(&pl.d_dsrc, pk->dsrc.data(), n_probs))) return rc;
    if ((rc = pl.upload(&pl.d_chan, pk->chan.data(), pk->chan.size()))) return rc;
    if ((rc = pl.upload(&pl.d_umap, pk->umap.data(), n_user))) return rc;
    if ((rc = pl.upload(&pl.d_uw, pk->uw.data(), n_user))) return rc;
    if ((rc = pl.alloc(&pl.d_udev, pl.n_params))) return rc;
    if ((rc = pl.alloc(&pl.d_Fdev, n_probs))) return rc;
    if ((rc = pl.alloc(&pl.d_Gdev, pl.n_params))) return rc;
  }
  {
    int mmax = 1;
    for (int p = 0; p < n_probs; p++) mmax = std::max(mmax, probs[p].n_members);
    pl.red_chunks = std::min(pr::RED_CHUNKS, mmax);
  }
  if ((rc = pl.alloc(&pl.d_part, (size_t)pl.red_chunks * pl.n_params))) return rc;
  {
    void* h = nullptr;
    const hipError_t e = hipHostMalloc(&h, sizeof(double) * (pl.n_uparams + n_user), hipHostMallocDefault);
    if (e != hipSuccess) return fail(PR_ERR_NOMEM, "hipHostMalloc(%lld doubles): %s", pl.n_uparams + n_user,
                                     hipGetErrorString(e));
    pl.h_stage = static_cast<double*>(h);
  }
  for (int gi = 1; gi <= 12; gi++) {
    if (items[gi].empty()) continue;
    const int B = gi > 8 ? gi - 8 : gi;
    Plan::Group g;
    g.B = B;
    g.gid = gi;
    g.track = gi > 8;
    g.track_packed = gi == 9;
    for (int p = 0; p < n_probs; p++)
      if (group_of(probs[p]) == gi && probs[p].n_pairs > 8) g.track_packed = false;
    g.V = vmax[gi];
    g.n_items = (int)items[gi].size();
    g.n_members = (int)members[gi].size();
    for (int p = 0; p < n_probs; p++)
      if (group_of(probs[p]) == gi && probs[p].dynamics == PR_LINDBLAD) g.lindblad = true;
    g.mixed = B == 1 && mixed1 && !g.track;
    g.split = g.mixed && pl.opt.precision == PR_SPLIT16;
    g.prescaled = true;
    for (int p = 0; p < n_probs; p++)
      if (group_of(probs[p]) == gi && !pl.host_probs[p].prescaled) g.prescaled = false;
    if (g.split) {
      int np = -1;
      for (int p = 0; p < n_probs; p++) {
        if (group_of(probs[p]) != gi) continue;
        const pr::ProbDev& D = pl.host_probs[p];
        const int n = D.so_idx ? D.so_pass[1] - D.so_pass[0] : 0;
        np = (np < 0 || np == n) ? n : 0;
      }
      g.split_np = (pl.tune_regs[gi] && np >= 1 && np <= 3) ? np : 0;
    }
    for (int p = 0; p < n_probs; p++)
      if (group_of(probs[p]) == gi && pl.host_probs[p].recompute) {
        g.recompute = true;
        if (!pl.host_probs[p].unitary) g.window = lind_window(B);
      }
    // every control fits one lane group of 64 / V lanes with <= TR_REG nonzeros per lane
    g.trace_regs = pl.tune_regs[gi] && B == 1 && vmax[gi] >= nc_max[gi] &&
                   tr_nnz_max[gi] <= pr::TR_REG * (64 / vmax[gi]);
    if ((rc = pl.upload(&g.d_items, items[gi].data(), items[gi].size()))) return rc;
    g.n_citems = (int)citems[gi].size();
    if (g.n_citems > 0 && (rc = pl.upload(&g.d_citems, citems[gi].data(), citems[gi].size()))) return rc;
    if ((rc = pl.upload(&g.d_members, members[gi].data(), members[gi].size()))) return rc;
    pl.groups.push_back(g);
    if (B > 4) {
      const size_t n = (size_t)std::max(g.n_items, g.n_members) * pr::BIG_SCRATCH * (size_t)(B * B) * 256;
      if ((rc = pl.alloc(&pl.groups.back().big_scratch, n))) return rc;
    }
    switch (B) {
      case 1: rc = set_smem_attr<1>(); break;
      case 2: rc = set_smem_attr<2>(); break;
      case 3: rc = set_smem_attr<3>(); break;
      case 4: rc = set_smem_attr<4>(); break;
      case 5: rc = set_big_attr<5>(); break;
      case 6: rc = set_big_attr<6>(); break;
      case 7: rc = set_big_attr<7>(); break;
      case 8: rc = set_big_attr<8>(); break;
    }
    if (rc) return rc;
  }
  return static_layout(pl);
}

int fill_diag(const Plan& pl, double* diag) {
  for (int p = 0; p < pl.n_user; p++) {
    // (a packed problem reports its pack's values)
    const pr::ProbDev& D = pl.host_probs[pl.packed ? pl.umap[p].x : p];
    std::vector<double> ud(D.M);
    std::vector<int> sm((size_t)4 * D.M * D.nseg);
    HIPCHK(hipMemcpy(ud.data(), D.udef, sizeof(double) * D.M, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sm.data(), D.smax, sizeof(int) * sm.size(), hipMemcpyDeviceToHost));
    double umax = 0;
    for (double v : ud) umax = std::max(umax, v);
    int smax = 0, any9 = 0;
    for (size_t i = 0; i < sm.size(); i += 4) {
      smax = std::max(smax, sm[i]);
      any9 |= sm[i + 1] > 0;
    }
    diag[PR_DIAG_PER_PROBLEM * p + 0] = umax;
    diag[PR_DIAG_PER_PROBLEM * p + 1] = smax;
    diag[PR_DIAG_PER_PROBLEM * p + 2] = any9 ? 9 : 6;
    diag[PR_DIAG_PER_PROBLEM * p + 3] = D.nseg;
  }
  return PR_OK;
}

}  // namespace

// The kernel-shape autotuner (pr_options.segment = PR_SEGMENT_AUTOTUNE): for
// each candidate (a multiple of the work-per-item segment length, times the
// operand placement of the single-wavefront kernels: trace operand and split
// assembly tables in registers or read from memory) a trial plan of the same
// problems is built and its dimension groups are timed one by one (HIP events
// around each group's launches: 1 warm-up and 3 timed evaluations at zero
// controls); every group then takes the candidate it ran fastest with. The
// final plan is built with those choices; the trial plans are freed before it
// is.
int time_groups(Plan& pl, std::vector<double>& ms) {
  ms.assign(13, 0.0);
  HIPCHK(hipMemsetAsync(pl.d_controls, 0, sizeof(double) * pl.n_uparams, pl.stream));
  const double* u = pl.d_controls;
  if (pl.packed) {
    const int rc = pack_controls(pl, pl.d_controls, pl.stream);
    if (rc) return rc;
    u = pl.d_udev;
  }
  const size_t ng = pl.groups.size();
  std::vector<hipEvent_t> ev(2 * ng, nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  int rc = e == hipSuccess ? PR_OK : fail(PR_ERR_DEVICE, "hipEventCreate: %s", hipGetErrorString(e));
  for (int rep = 0; rep < 4 && rc == PR_OK; rep++) {
    for (size_t i = 0; i < ng && rc == PR_OK; i++) {
      if ((e = hipEventRecord(ev[2 * i], pl.stream)) != hipSuccess) break;
      rc = launch_one(pl, pl.groups[i], u, true, pl.stream);
      if (rc == PR_OK && (e = hipEventRecord(ev[2 * i + 1], pl.stream)) != hipSuccess) break;
    }
    if (rc == PR_OK && e == hipSuccess) e = hipStreamSynchronize(pl.stream);
    if (rc == PR_OK && e != hipSuccess) rc = fail(PR_ERR_DEVICE, "autotune timing: %s", hipGetErrorString(e));
    for (size_t i = 0; rep > 0 && i < ng && rc == PR_OK; i++) {
      float t = 0;
      if ((e = hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1])) != hipSuccess)
        rc = fail(PR_ERR_DEVICE, "hipEventElapsedTime: %s", hipGetErrorString(e));
      ms[pl.groups[i].gid] += t;
    }
  }
  for (auto x : ev)
    if (x) (void)hipEventDestroy(x);
  return rc;
}

int autotune_segments(Plan& pl, const pr_problem* probs, int n_probs, const pr_problem* uprobs, int n_user,
                      const PackBuild* pk) {
  const double cand[] = {0.5, 1.0, 2.0, 4.0};
  std::vector<double> best(13, 1e300), scale(13, 1.0), ms;
  std::vector<int> regs(13, 1);
  // the memory placement is only a different kernel for groups of single-
  // wavefront problems (block count 1, plain or tracking)
  bool small = false;
  for (int p = 0; p < n_probs; p++) small = small || probs[p].dim <= 16;
  for (int r = 1; r >= (small ? 0 : 1); r--) {
    for (double c : cand) {
      Plan* t = new Plan;
      t->opt = pl.opt;
      t->use_graphs = false;
      t->seg_scale.assign(13, c);
      t->tune_regs.assign(13, r);
      hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
      int rc = e == hipSuccess ? PR_OK : fail(PR_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
      if (rc == PR_OK) rc = build_plan(*t, probs, n_probs, uprobs, n_user, pk);
      if (rc == PR_OK) rc = time_groups(*t, ms);
      delete t;
      if (rc) return rc;
      for (int gi = 0; gi < 13; gi++)
        if (ms[gi] > 0 && ms[gi] < best[gi]) {
          best[gi] = ms[gi];
          scale[gi] = c;
          regs[gi] = r;
        }
    }
  }
  pl.seg_scale = scale;
  pl.tune_regs = regs;
  return PR_OK;
}

// Every problem's controls packed back to back into the plan's pinned staging
// buffer, so that one host-to-device copy moves them all.
int stage_controls(Plan& pl, const double* const* controls, const char* who) {
  for (int p = 0; p < pl.n_user; p++) {
    if (!controls[p]) return fail(PR_ERR_ARG, "%s: controls[%d] is NULL", who, p);
    memcpy(pl.h_stage + pl.ucoff[p], controls[p], sizeof(double) * (pl.ucoff[p + 1] - pl.ucoff[p]));
  }
  return PR_OK;
}

// Staging ranges of pr_eval: one per host thread, of about equal bytes, at
// least 2 MiB each and at most 8 (a plan of a few MiB is packed by the
// calling thread alone).
int stage_ranges(Plan& pl) {
  if (!pl.stage_cut.empty()) return PR_OK;
  const size_t min_range = (size_t)1 << 18;  // doubles
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int R = (int)std::max<size_t>(1, std::min<size_t>({8, hw, (size_t)pl.n_uparams / min_range}));
  pl.stage_cut.assign(1, 0);
  for (int r = 1; r < R; r++) {
    const long long target = (long long)((double)pl.n_uparams * r / R);
    int p = pl.stage_cut.back();
    while (p < pl.n_user && pl.ucoff[p] < target) p++;
    if (p > pl.stage_cut.back() && p < pl.n_user) pl.stage_cut.push_back(p);
  }
  pl.stage_cut.push_back(pl.n_user);
  pl.stage_ev.assign(pl.stage_cut.size() - 1, nullptr);
  for (auto& e : pl.stage_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return PR_OK;
}

// pr_eval's transfers with host buffers: the ranges' controls are packed in
// parallel and each range is uploaded as soon as it is packed (the next
// ranges are still being packed meanwhile); after the evaluation, each
// range's gradients are downloaded and unpacked by their thread once their
// own download has finished, while later ranges are still in flight.
int staged_eval(Plan& pl, const double* const* controls, double* F, double* const* G) {
  for (int p = 0; p < pl.n_user; p++)
    if (!controls[p]) return fail(PR_ERR_ARG, "pr_eval: controls[%d] is NULL", p);
  int rc = stage_ranges(pl);
  if (rc) return rc;
  const int R = (int)pl.stage_ev.size();
  auto range_of = [&](int r, size_t& off, size_t& len) {
    off = (size_t)pl.ucoff[pl.stage_cut[r]];
    len = (size_t)(pl.ucoff[pl.stage_cut[r + 1]] - pl.ucoff[pl.stage_cut[r]]);
  };
  auto pack = [&](int r) {
    for (int p = pl.stage_cut[r]; p < pl.stage_cut[r + 1]; p++)
      memcpy(pl.h_stage + pl.ucoff[p], controls[p], sizeof(double) * (pl.ucoff[p + 1] - pl.ucoff[p]));
  };
  auto unpack = [&](int r) {
    for (int p = pl.stage_cut[r]; p < pl.stage_cut[r + 1]; p++)
      if (G[p]) memcpy(G[p], pl.h_stage + pl.ucoff[p], sizeof(double) * (pl.ucoff[p + 1] - pl.ucoff[p]));
  };
  std::vector<std::thread> th;
  for (int r = 1; r < R; r++) th.emplace_back([&, r] { pack(r); });
  pack(0);
  hipError_t e = hipSuccess;
  for (int r = 0; r < R; r++) {
    if (r > 0) th[r - 1].join();
    size_t off, len;
    range_of(r, off, len);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pl.d_controls + off, pl.h_stage + off, sizeof(double) * len, hipMemcpyHostToDevice,
                         pl.stream);
  }
  th.clear();
  if (e != hipSuccess) return fail(PR_ERR_DEVICE, "pr_eval: hipMemcpyAsync: %s", hipGetErrorString(e));
  rc = run_graph(pl, pl.d_controls, pl.d_F, G ? pl.d_G : nullptr, pl.stream);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(pl.h_stage + pl.n_uparams, pl.d_F, sizeof(double) * pl.n_user, hipMemcpyDeviceToHost,
                        pl.stream));
  if (G) {
    for (int r = 0; r < R; r++) {
      size_t off, len;
      range_of(r, off, len);
      HIPCHK(hipMemcpyAsync(pl.h_stage + off, pl.d_G + off, sizeof(double) * len, hipMemcpyDeviceToHost, pl.stream));
      HIPCHK(hipEventRecord(pl.stage_ev[r], pl.stream));
    }
    std::vector<hipError_t> we(R, hipSuccess);
    for (int r = 1; r < R; r++)
      th.emplace_back([&, r] {
        if (g_device >= 0) (void)hipSetDevice(g_device);
        we[r] = hipEventSynchronize(pl.stage_ev[r]);
        if (we[r] == hipSuccess) unpack(r);
      });
    we[0] = hipEventSynchronize(pl.stage_ev[0]);
    if (we[0] == hipSuccess) unpack(0);
    for (auto& t : th) t.join();
    for (int r = 0; r < R; r++)
      if (we[r] != hipSuccess) return fail(PR_ERR_DEVICE, "pr_eval: hipEventSynchronize: %s", hipGetErrorString(we[r]));
  }
  HIPCHK(hipStreamSynchronize(pl.stream));
  memcpy(F, pl.h_stage + pl.n_uparams, sizeof(double) * pl.n_user);
  return PR_OK;
}

extern "C" {

int pr_abi_version(void) { return PR_ABI_VERSION; }

int pr_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int pr_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return fail(PR_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(PR_ERR_DEVICE, "device %d out of range (%d visible)", device, n);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(PR_ERR_DEVICE, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  HIPCHK(hipSetDevice(device));
  if (device < PR_MAX_DEVICES && !g_scratch_ptr[device]) {
    // the static scratch's address on this device (each GPU holds its own copy)
    void* ptr = nullptr;
    const hipError_t es = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (es != hipSuccess || !ptr)
      return fail(PR_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(es));
    g_scratch_ptr[device] = static_cast<unsigned char*>(ptr);
  }
  g_device = device;
  return PR_OK;
}

int pr_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return PR_ERR_ARG;
  snprintf(buf, n, "%s", g_err.c_str());
  return PR_OK;
}

int pr_plan_create(const pr_problem* probs, int n_probs, const pr_options* opt, void** plan) {
  if (fault_is("plan")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=plan");
  if (!probs || n_probs < 1 || !plan) return fail(PR_ERR_ARG, "pr_plan_create: need problems and a plan pointer");
  *plan = nullptr;
  pr_options o{PR_FP64, PR_GRAD_REFERENCE, 0, -1, 1};
  if (opt) o = *opt;
  if (o.precision != PR_FP64 && o.precision != PR_MIXED && o.precision != PR_SPLIT16)
    return fail(PR_ERR_ARG, "unknown precision %d", o.precision);
  if (o.grad_mode != PR_GRAD_REFERENCE && o.grad_mode != PR_GRAD_EXACT)
    return fail(PR_ERR_ARG, "unknown grad_mode %d", o.grad_mode);
  if (o.store_props < -1 || o.store_props > 1) return fail(PR_ERR_ARG, "store_props %d is not -1, 0 or 1", o.store_props);
  if (o.segment < PR_SEGMENT_AUTOTUNE) return fail(PR_ERR_ARG, "segment must be >= 0, or PR_SEGMENT_AUTOTUNE");
  if (o.pack != 0 && o.pack != 1) return fail(PR_ERR_ARG, "pack %d is not 0 or 1", o.pack);
  for (int p = 0; p < n_probs; p++) {
    if (probs[p].dim > 64 && o.grad_mode == PR_GRAD_EXACT)
      return fail(PR_ERR_UNSUPPORTED, "problem %d: the exact gradient is available for dim <= 64", p);
    if ((probs[p].objective == PR_TRACKING || probs[p].objective == PR_GATE_PROPAGATORS ||
         probs[p].objective == PR_STATE_PROPAGATORS) &&
        o.grad_mode == PR_GRAD_EXACT)
      return fail(PR_ERR_UNSUPPORTED, "problem %d: objective %d has the reference gradient only", p, probs[p].objective);
    int rc = validate(probs[p], p);
    if (rc) return rc;
  }
  int rc = ensure_device();
  if (rc) return rc;
  Plan* pl = new Plan;
  pl->opt = o;
  const char* ng = getenv("PULSAR_ROCM_NO_GRAPH");
  pl->use_graphs = !(ng && ng[0] && ng[0] != '0');
  hipError_t e = hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete pl;
    return fail(PR_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  PackBuild pk;
  const bool packed = o.pack && build_packs(probs, n_probs, pk);
  const pr_problem* dprobs = packed ? pk.dev.data() : probs;
  const int ndev = packed ? (int)pk.dev.size() : n_probs;
  const PackBuild* pkp = packed ? &pk : nullptr;
  if (o.segment == PR_SEGMENT_AUTOTUNE) {
    pl->opt.segment = 0;
    rc = autotune_segments(*pl, dprobs, ndev, probs, n_probs, pkp);
    if (rc) {
      delete pl;
      return rc;
    }
  }
  rc = build_plan(*pl, dprobs, ndev, probs, n_probs, pkp);
  if (rc) {
    delete pl;
    return rc;
  }
  *plan = pl;
  return PR_OK;
}

int pr_eval(void* plan, const double* const* controls, double* F, double* const* G, double* diag) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!plan || !controls || !F) return fail(PR_ERR_ARG, "pr_eval: plan, controls and F are required");
  Plan& pl = *static_cast<Plan*>(plan);
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  int rc = staged_eval(pl, controls, F, G);
  if (rc) return rc;
  if (diag) return fill_diag(pl, diag);
  return PR_OK;
}

int pr_eval_device(void* plan, const double* d_controls, double* d_F, double* d_G, void* stream) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!plan || !d_controls || !d_F) return fail(PR_ERR_ARG, "pr_eval_device: plan, d_controls and d_F are required");
  Plan& pl = *static_cast<Plan*>(plan);
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : pl.stream;
  return run_graph(pl, d_controls, d_F, d_G, st);
}

int pr_eval_device_static(void* plan, const double* d_controls, double* d_F, double* d_G, void* stream) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!plan || !d_controls || !d_F)
    return fail(PR_ERR_ARG, "pr_eval_device_static: plan, d_controls and d_F are required");
  Plan& pl = *static_cast<Plan*>(plan);
  if (g_device < 0 || g_device >= PR_MAX_DEVICES || !g_scratch_ptr[g_device])
    return fail(PR_ERR_DEVICE, "pr_eval_device_static: pr_init has not set up this device");
  if (!pl.static_ok)
    return fail(PR_ERR_UNSUPPORTED, "pr_eval_device_static: plans with dim > 64 problems keep per-item scratch "
                                    "of their own; use pr_eval_device");
  if (pl.static_bytes > (long long)PR_STATIC_SCRATCH_BYTES)
    return fail(PR_ERR_UNSUPPORTED, "pr_eval_device_static: the plan needs %lld bytes of scratch; the static "
                                    "scratch holds %llu (use store_props = 0, fewer members or pr_eval_device)",
                pl.static_bytes, (unsigned long long)PR_STATIC_SCRATCH_BYTES);
  HIPCHK(hipSetDevice(g_device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : pl.stream;
  unsigned char* base = g_scratch_ptr[g_device];
  // problem records pointing into the static scratch, and the range that must
  // start at zero (the unitarity defects)
  hipLaunchKernelGGL(pr::k_bind_static, dim3((pl.n_probs + 63) / 64), dim3(64), 0, st, pl.d_probs,
                     pl.d_probs_static, pl.d_static_off, pl.n_probs, base);
  HIPCHK(hipGetLastError());
  if (pl.static_udef_bytes > 0) {
    hipLaunchKernelGGL(pr::k_zero, dim3(64), dim3(256), 0, st, reinterpret_cast<double*>(base),
                       pl.static_udef_bytes / (long long)sizeof(double));
    HIPCHK(hipGetLastError());
  }
  // the plan's launch sequence on those records, launched directly (no graph
  // capture of its own, no timing events), so a caller may capture it
  pr::ProbDev* own_probs = pl.d_probs;
  double* own_part = pl.d_part;
  double *own_udev = pl.d_udev, *own_Fdev = pl.d_Fdev, *own_Gdev = pl.d_Gdev;
  const bool own_timing = pl.timing;
  pl.d_probs = pl.d_probs_static;
  pl.d_part = reinterpret_cast<double*>(base + pl.static_part);
  if (pl.packed) {
    pl.d_udev = reinterpret_cast<double*>(base + pl.static_udev);
    pl.d_Fdev = reinterpret_cast<double*>(base + pl.static_Fdev);
    pl.d_Gdev = reinterpret_cast<double*>(base + pl.static_Gdev);
  }
  pl.timing = false;
  const int rc = run(pl, d_controls, d_F, d_G, st);
  pl.d_probs = own_probs;
  pl.d_part = own_part;
  pl.d_udev = own_udev;
  pl.d_Fdev = own_Fdev;
  pl.d_Gdev = own_Gdev;
  pl.timing = own_timing;
  return rc;
}

int pr_plan_buffers(void* plan, double** d_controls, double** d_F, double** d_G) {
  if (!plan || !d_controls || !d_F || !d_G) return fail(PR_ERR_ARG, "pr_plan_buffers: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  *d_controls = pl.d_controls;
  *d_F = pl.d_F;
  *d_G = pl.d_G;
  return PR_OK;
}

int pr_plan_upload_controls(void* plan, const double* const* controls) {
  if (!plan || !controls) return fail(PR_ERR_ARG, "pr_plan_upload_controls: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  int rc = stage_controls(pl, controls, "pr_plan_upload_controls");
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(pl.d_controls, pl.h_stage, sizeof(double) * pl.n_uparams, hipMemcpyHostToDevice, pl.stream));
  HIPCHK(hipStreamSynchronize(pl.stream));
  return PR_OK;
}

int pr_plan_synchronize(void* plan) {
  if (!plan) return fail(PR_ERR_ARG, "pr_plan_synchronize: plan is NULL");
  HIPCHK(hipStreamSynchronize(static_cast<Plan*>(plan)->stream));
  return PR_OK;
}

int pr_plan_set_timing(void* plan, int enable) {
  if (!plan) return fail(PR_ERR_ARG, "pr_plan_set_timing: plan is NULL");
  static_cast<Plan*>(plan)->timing = enable != 0;
  return PR_OK;
}

int pr_plan_kernel_times(void* plan, double* ms, long long* launches) {
  if (!plan || !ms || !launches) return fail(PR_ERR_ARG, "pr_plan_kernel_times: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  for (auto& r : pl.pending) {
    HIPCHK(hipEventSynchronize(r.b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
    pl.acc_ms[r.cls] += t;
    pl.acc_n[r.cls] += 1;
    pl.ev_free.push_back(r.a);
    pl.ev_free.push_back(r.b);
  }
  pl.pending.clear();
  for (int c = 0; c < PR_KERNEL_CLASSES; c++) {
    ms[c] = pl.acc_ms[c];
    launches[c] = pl.acc_n[c];
    pl.acc_ms[c] = 0;
    pl.acc_n[c] = 0;
  }
  return PR_OK;
}

int pr_plan_exp_counts(void* plan, long long* counts) {
  if (!plan || !counts) return fail(PR_ERR_ARG, "pr_plan_exp_counts: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  HIPCHK(hipStreamSynchronize(pl.stream));
  for (int p = 0; p < pl.n_user; p++) {
    for (int i = 0; i < 4; i++) counts[4 * p + i] = 0;
    // a pack's exps are reported by its first problem
    if (pl.packed && pl.umap[p].y > 0) continue;
    const pr::ProbDev& D = pl.host_probs[pl.packed ? pl.umap[p].x : p];
    std::vector<int> sm((size_t)4 * D.M * D.nseg);
    HIPCHK(hipMemcpy(sm.data(), D.smax, sizeof(int) * sm.size(), hipMemcpyDeviceToHost));
    long long n9 = 0, nsq = 0, nredo = 0;
    for (size_t i = 0; i < sm.size(); i += 4) {
      n9 += sm[i + 1];
      nsq += sm[i + 2];
      nredo += sm[i + 3];
    }
    counts[4 * p + 0] = (long long)D.M * D.nt - n9;
    counts[4 * p + 1] = n9;
    counts[4 * p + 2] = nsq;
    counts[4 * p + 3] = nredo;
  }
  return PR_OK;
}

int pr_plan_shapes(void* plan, int* shapes) {
  if (!plan || !shapes) return fail(PR_ERR_ARG, "pr_plan_shapes: NULL argument");
  const Plan& pl = *static_cast<Plan*>(plan);
  auto group_of = [](const pr::ProbDev& D) { return (D.obj == PR_TRACKING ? 8 : 0) + D.B; };
  for (int p = 0; p < pl.n_user; p++) {
    const pr::ProbDev& D = pl.host_probs[pl.packed ? pl.umap[p].x : p];
    int regs = 0, np = 0;
    for (const auto& g : pl.groups)
      if (g.gid == group_of(D)) {
        regs = g.trace_regs ? 1 : 0;
        np = g.split ? g.split_np : 0;
      }
    shapes[4 * p + 0] = D.seglen;
    shapes[4 * p + 1] = D.nseg;
    shapes[4 * p + 2] = regs;
    shapes[4 * p + 3] = np;
  }
  return PR_OK;
}

long long pr_plan_n_params(void* plan) {
  if (!plan) return -1;
  return static_cast<Plan*>(plan)->n_uparams;
}

int pr_grape_optimize(void* fid_plan, void* grad_plan, const double* controls_init, const pr_grape_config* cfg,
                      double* controls_best, double* fidelity_history, double* gnorm_history,
                      pr_grape_result* result) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!fid_plan || !controls_init || !cfg || !controls_best || !fidelity_history || !gnorm_history || !result)
    return fail(PR_ERR_ARG, "pr_grape_optimize: NULL argument");
  Plan& fp = *static_cast<Plan*>(fid_plan);
  Plan* gp = static_cast<Plan*>(grad_plan);
  if (fp.n_user != 1 || (gp && (gp->n_user != 1 || gp->n_uparams != fp.n_uparams)))
    return fail(PR_ERR_ARG, "pr_grape_optimize: plans must hold one problem each, with the same controls");
  if (cfg->max_iter <= 0 || !(cfg->step_size > 0) || !(cfg->min_step_size > 0) ||
      cfg->max_step_size < cfg->step_size || cfg->min_step_size > cfg->step_size)
    return fail(PR_ERR_ARG, "pr_grape_optimize: invalid configuration");
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  const long long n = fp.n_uparams;
  hipStream_t st = fp.stream;
  DevBuf best, part, nrm;
  HIPCHK(hipMalloc(&best.p, sizeof(double) * n));
  HIPCHK(hipMalloc(&part.p, sizeof(double) * pr::NORM_BLOCKS));
  HIPCHK(hipMalloc(&nrm.p, sizeof(double)));
  double* d_u = fp.d_controls;
  double* d_best = static_cast<double*>(best.p);
  HIPCHK(hipMemcpyAsync(d_u, controls_init, sizeof(double) * n, hipMemcpyHostToDevice, st));
  int rc;
  // F at controls d (F only), read back
  auto fid = [&](const double* d, double* F) -> int {
    int r = run_graph(fp, d, fp.d_F, nullptr, st);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(F, fp.d_F, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PR_OK;
  };
  auto copy_best = [&]() -> int {
    HIPCHK(hipMemcpyAsync(d_best, d_u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return PR_OK;
  };
  std::vector<double> fh;
  int n_f = 0, n_g = 0, n_iter = 0, converged = 0, reason = PR_STOP_MAX_ITER;
  double step = cfg->step_size, F = 0, best_F;
  if ((rc = fid(d_u, &F))) return rc;
  n_f++;
  fh.push_back(F);
  best_F = F;
  if ((rc = copy_best())) return rc;
  for (int it = 1; it <= cfg->max_iter; it++) {
    n_iter = it;
    double* d_G;
    if (gp) {
      if ((rc = run_graph(fp, d_u, fp.d_F, nullptr, st))) return rc;
      if ((rc = run_graph(*gp, d_u, gp->d_F, gp->d_G, st))) return rc;
      d_G = gp->d_G;
    } else {
      if ((rc = run_graph(fp, d_u, fp.d_F, fp.d_G, st))) return rc;
      d_G = fp.d_G;
    }
    n_f++;
    n_g++;
    hipLaunchKernelGGL(pr::k_norm2_part, dim3(pr::NORM_BLOCKS), dim3(256), 0, st, d_G, n,
                       static_cast<double*>(part.p));
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(pr::k_norm2_final, dim3(1), dim3(64), 0, st, static_cast<const double*>(part.p),
                       static_cast<double*>(nrm.p));
    HIPCHK(hipGetLastError());
    double g2 = 0;
    HIPCHK(hipMemcpyAsync(&F, fp.d_F, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&g2, nrm.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const double gnorm = std::sqrt(g2);
    fh.push_back(F);
    gnorm_history[it - 1] = gnorm;
    if (F > best_F) {
      best_F = F;
      if ((rc = copy_best())) return rc;
    }
    const size_t nh = fh.size();
    if (cfg->adapt_step_size && nh >= 2) {
      const double delta = fh[nh - 1] - fh[nh - 2];
      double s2 = step;
      if (delta > 0.0)
        s2 = step * 1.05;
      else if (delta < -1e-14)
        s2 = step * 0.5;
      step = std::min(std::max(s2, cfg->min_step_size), cfg->max_step_size);
    }
    hipLaunchKernelGGL(pr::k_grape_step, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_u,
                       static_cast<const double*>(d_G), step, n);
    HIPCHK(hipGetLastError());
    int why = -1;
    if (F >= 1.0 - 1e-12)
      why = PR_STOP_PERFECT;
    else if (gnorm < cfg->gradient_norm_tol)
      why = PR_STOP_GRADIENT;
    else if (nh >= 3 && std::fabs(fh[nh - 1] - fh[nh - 2]) < cfg->convergence_tol)
      why = PR_STOP_FIDELITY;
    if (why >= 0) {
      converged = 1;
      reason = why;
      double Fp = 0;
      if ((rc = fid(d_u, &Fp))) return rc;
      fh.push_back(Fp);
      n_f++;
      if (Fp > best_F) {
        best_F = Fp;
        if ((rc = copy_best())) return rc;
      }
      break;
    }
  }
  double F_final = 0;
  if ((rc = fid(d_best, &F_final))) return rc;
  n_f++;
  HIPCHK(hipMemcpyAsync(controls_best, d_best, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::copy(fh.begin(), fh.end(), fidelity_history);
  result->n_iterations = n_iter;
  result->converged = converged;
  result->reason = reason;
  result->n_fidelity_history = (int)fh.size();
  result->n_fidelity_evaluations = n_f;
  result->n_gradient_evaluations = n_g;
  result->fidelity = F_final;
  result->step_size_final = step;
  return PR_OK;
}

int pr_lbfgs_optimize(void* plan, const double* controls_init, const pr_lbfgs_config* cfg, double* controls_best,
                      double* fidelity_history, double* gnorm_history, pr_grape_result* result) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!plan || !controls_init || !cfg || !controls_best || !fidelity_history || !gnorm_history || !result)
    return fail(PR_ERR_ARG, "pr_lbfgs_optimize: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  if (pl.n_user != 1) return fail(PR_ERR_ARG, "pr_lbfgs_optimize: the plan must hold one problem");
  if (cfg->max_iter <= 0 || !(cfg->step_size > 0) || cfg->lbfgs_memory < 1 || cfg->lbfgs_memory > 64)
    return fail(PR_ERR_ARG, "pr_lbfgs_optimize: invalid configuration");
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  const long long n = pl.n_uparams;
  const int mem = cfg->lbfgs_memory;
  const bool clip = cfg->amplitude_limit > 0;
  hipStream_t st = pl.stream;
  // device vectors: u (plan controls), un, g, gn, d, best, S and Y rows, scalars
  DevBuf vecs, hist, scal;
  HIPCHK(hipMalloc(&vecs.p, sizeof(double) * (size_t)n * 5));
  // mem + 1 rows: a new pair is written to the free row and joins the history
  // only if its curvature is accepted
  const int R = mem + 1;
  HIPCHK(hipMalloc(&hist.p, sizeof(double) * (size_t)n * 2 * R));
  HIPCHK(hipMalloc(&scal.p, sizeof(double) * (pr::NORM_BLOCKS + 4)));
  double* d_u = pl.d_controls;
  double* d_un = static_cast<double*>(vecs.p);
  double* d_g = d_un + n;
  double* d_gn = d_g + n;
  double* d_d = d_gn + n;
  double* d_best = d_d + n;
  double* d_S = static_cast<double*>(hist.p);
  double* d_Y = d_S + (size_t)n * R;
  double* d_part = static_cast<double*>(scal.p);
  double* d_sc = d_part + pr::NORM_BLOCKS;  // [0] |g|^2 or max|d|, [1..2] y.s, s.s
  const dim3 eb((unsigned)((n + 255) / 256)), et(256);
  HIPCHK(hipMemcpyAsync(d_u, controls_init, sizeof(double) * n, hipMemcpyHostToDevice, st));
  int rc;
  auto read = [&](const double* src, double* dst, int cnt) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PR_OK;
  };
  auto fid = [&](const double* x, double* F) -> int {
    int r = run_graph(pl, x, pl.d_F, nullptr, st);
    return r ? r : read(pl.d_F, F, 1);
  };
  // F and the gradient at x into gout
  auto grad = [&](const double* x, double* gout, double* F) -> int {
    int r = run_graph(pl, x, pl.d_F, gout, st);
    return r ? r : read(pl.d_F, F, 1);
  };
  auto trial = [&](double a) -> int {
    hipLaunchKernelGGL(pr::k_lbfgs_trial, eb, et, 0, st, d_u, d_d, a, clip ? cfg->amplitude_limit : 0.0, n, d_un);
    HIPCHK(hipGetLastError());
    return PR_OK;
  };
  std::vector<double> fh;
  int n_f = 0, n_g = 0, n_iter = 0, converged = 0, reason = PR_STOP_MAX_ITER, k = 0, first = 0;
  double F = 0, best_F, a_last = 0;
  if ((rc = fid(d_u, &F))) return rc;
  n_f++;
  fh.push_back(F);
  best_F = F;
  HIPCHK(hipMemcpyAsync(d_best, d_u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  for (int it = 1; it <= cfg->max_iter; it++) {
    n_iter = it;
    double Fg = 0;
    if ((rc = grad(d_u, d_g, &Fg))) return rc;
    n_g++;
    hipLaunchKernelGGL(pr::k_norm2_part, dim3(pr::NORM_BLOCKS), dim3(256), 0, st, d_g, n, d_part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(pr::k_norm2_final, dim3(1), dim3(64), 0, st, d_part, d_sc);
    HIPCHK(hipGetLastError());
    double g2 = 0;
    if ((rc = read(d_sc, &g2, 1))) return rc;
    const double gnorm = std::sqrt(g2);
    gnorm_history[it - 1] = gnorm;
    hipLaunchKernelGGL(pr::k_lbfgs_direction, dim3(1), dim3(pr::LBFGS_THREADS), 0, st, d_g, d_S, d_Y, k, first, R,
                       n, 1e-14, d_d, d_sc);
    HIPCHK(hipGetLastError());
    double dmax = 0;
    if ((rc = read(d_sc, &dmax, 1))) return rc;
    if (dmax < 1e-14) {
      converged = 1;
      reason = PR_STOP_ZERO_DIRECTION;
      break;
    }
    hipLaunchKernelGGL(pr::k_lbfgs_scale, eb, et, 0, st, d_d, d_sc, n);
    HIPCHK(hipGetLastError());
    // backtracking line search on F (GRAPE.jl:1082-1104)
    double a = clip ? cfg->step_size * cfg->amplitude_limit : cfg->step_size, Fn = 0;
    bool ok = false;
    for (int bt = 0; bt < 20; bt++) {
      if ((rc = trial(a)) || (rc = fid(d_un, &Fn))) return rc;
      n_f++;
      if (Fn >= F) {
        ok = true;
        break;
      }
      a *= 0.5;
    }
    if (!ok) {
      if ((rc = trial(a)) || (rc = fid(d_un, &Fn))) return rc;
      n_f++;
    }
    a_last = a;
    if ((rc = grad(d_un, d_gn, &Fg))) return rc;
    n_g++;
    // curvature pair into the free row, kept when y.s > 1e-14 |s|^2 (the
    // oldest pair dropped beyond mem)
    const int slot = (first + k) % R;
    hipLaunchKernelGGL(pr::k_lbfgs_pair, dim3(1), dim3(pr::LBFGS_THREADS), 0, st, d_d, a, d_gn, d_g, n,
                       d_S + (size_t)slot * n, d_Y + (size_t)slot * n, d_sc + 1);
    HIPCHK(hipGetLastError());
    double ysss[2];
    if ((rc = read(d_sc + 1, ysss, 2))) return rc;
    if (ysss[0] > 1e-14 * ysss[1]) {
      if (k < mem)
        k++;
      else
        first = (first + 1) % R;
    }
    HIPCHK(hipMemcpyAsync(d_u, d_un, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    F = Fn;
    fh.push_back(F);
    if (F > best_F) {
      best_F = F;
      HIPCHK(hipMemcpyAsync(d_best, d_u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    }
    const size_t nh = fh.size();
    int why = -1;
    if (gnorm < cfg->gradient_norm_tol)
      why = PR_STOP_GRADIENT;
    else if (F >= 1.0 - 1e-8)
      why = PR_STOP_PERFECT;
    else if (nh >= 2 && std::fabs(fh[nh - 1] - fh[nh - 2]) < cfg->convergence_tol)
      why = PR_STOP_FIDELITY;
    if (why >= 0) {
      converged = 1;
      reason = why;
      break;
    }
  }
  HIPCHK(hipMemcpyAsync(controls_best, d_best, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::copy(fh.begin(), fh.end(), fidelity_history);
  result->n_iterations = n_iter;
  result->converged = converged;
  result->reason = reason;
  result->n_fidelity_history = (int)fh.size();
  result->n_fidelity_evaluations = n_f;
  result->n_gradient_evaluations = n_g;
  result->fidelity = best_F;
  result->step_size_final = a_last;
  return PR_OK;
}

int pr_mr_optimcon(void* plan, const double* controls_init, const pr_mr_config* cfg, double* controls_best,
                   double* fidelity_history, double* gnorm_history, pr_grape_result* result) {
  if (fault_is("eval")) return fail(PR_ERR_INJECTED, "failure injected by PULSAR_ROCM_FAULT=eval");
  if (!plan || !controls_init || !cfg || !controls_best || !fidelity_history || !gnorm_history || !result)
    return fail(PR_ERR_ARG, "pr_mr_optimcon: NULL argument");
  Plan& pl = *static_cast<Plan*>(plan);
  if (pl.n_user != 1) return fail(PR_ERR_ARG, "pr_mr_optimcon: the plan must hold one problem");
  const int f = cfg->tile;
  if ((cfg->method != PR_MR_LBFGS && cfg->method != PR_MR_GRAPE) || cfg->max_iter < 0 || f < 1 ||
      cfg->lbfgs_memory < 1 || cfg->lbfgs_memory > 64 || !(cfg->u_bound >= cfg->l_bound))
    return fail(PR_ERR_ARG, "pr_mr_optimcon: invalid configuration");
  // the caller's problem's n_t and n_ctrl (a plan that packs its members
  // runs a wider problem of the same n_t)
  const int nt0 = pl.host_probs[pl.packed ? pl.umap[0].x : 0].nt;
  const int nc0 = (int)(pl.n_uparams / nt0);
  if (nc0 % f != 0) return fail(PR_ERR_ARG, "pr_mr_optimcon: n_ctrl %d is not a multiple of tile %d", nc0, f);
  if (g_device >= 0) HIPCHK(hipSetDevice(g_device));
  const int nc = nc0 / f;
  const long long n = (long long)nc * nt0;  // waveform entries
  const int mem = cfg->lbfgs_memory, R = mem + 1;
  const double lo = cfg->l_bound, hi = cfg->u_bound, a0 = 0.05 * (hi - lo);
  hipStream_t st = pl.stream;
  // device vectors: w, w_new, G, G_new, d, best, (S, Y) rows, scalars
  DevBuf vecs, hist, scal;
  HIPCHK(hipMalloc(&vecs.p, sizeof(double) * (size_t)n * 6));
  if (cfg->method == PR_MR_LBFGS) HIPCHK(hipMalloc(&hist.p, sizeof(double) * (size_t)n * 2 * R));
  HIPCHK(hipMalloc(&scal.p, sizeof(double) * (pr::NORM_BLOCKS + 4)));
  double* d_w = static_cast<double*>(vecs.p);
  double* d_wn = d_w + n;
  double* d_g = d_wn + n;
  double* d_gn = d_g + n;
  double* d_d = d_gn + n;
  double* d_best = d_d + n;
  double* d_S = static_cast<double*>(hist.p);
  double* d_Y = d_S ? d_S + (size_t)n * R : nullptr;
  double* d_part = static_cast<double*>(scal.p);
  double* d_sc = d_part + pr::NORM_BLOCKS;
  const dim3 eb((unsigned)((n + 255) / 256)), et(256);
  int rc;
  auto read = [&](const double* src, double* dst, int cnt) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return PR_OK;
  };
  // F (and G into gout when it is not NULL) at the waveform x
  auto eval = [&](const double* x, double* gout, double* F) -> int {
    const double* cx = x;
    if (f > 1) {
      hipLaunchKernelGGL(pr::k_mr_tile, eb, et, 0, st, x, n, nc, f, pl.d_controls);
      HIPCHK(hipGetLastError());
      cx = pl.d_controls;
    }
    int r = run_graph(pl, cx, pl.d_F, gout ? (f > 1 ? pl.d_G : gout) : nullptr, st);
    if (r) return r;
    if (gout && f > 1) {
      hipLaunchKernelGGL(pr::k_mr_untile, eb, et, 0, st, pl.d_G, n, nc, f, gout);
      HIPCHK(hipGetLastError());
    }
    return read(pl.d_F, F, 1);
  };
  auto gnorm = [&](const double* g, double* out) -> int {
    hipLaunchKernelGGL(pr::k_norm2_part, dim3(pr::NORM_BLOCKS), dim3(256), 0, st, g, n, d_part);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(pr::k_norm2_final, dim3(1), dim3(64), 0, st, d_part, d_sc);
    HIPCHK(hipGetLastError());
    double g2 = 0;
    int r = read(d_sc, &g2, 1);
    *out = std::sqrt(g2);
    return r;
  };
  auto copy = [&](double* dst, const double* src) -> int {
    HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    return PR_OK;
  };
  // the guess, clipped
  HIPCHK(hipMemcpyAsync(d_wn, controls_init, sizeof(double) * n, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(pr::k_mr_trial, eb, et, 0, st, d_wn, d_wn, 0.0, lo, hi, n, d_w);
  HIPCHK(hipGetLastError());
  std::vector<double> fh, gh;
  double F = 0, gn = 0, best_F;
  if ((rc = eval(d_w, d_g, &F)) || (rc = gnorm(d_g, &gn))) return rc;
  int n_f = 1, n_g = 1, converged = 0, reason = PR_STOP_MAX_ITER, k = 0, first = 0;
  fh.push_back(F);
  gh.push_back(gn);
  best_F = F;
  if ((rc = copy(d_best, d_w))) return rc;
  for (int it = 0; it < cfg->max_iter; it++) {
    double Fn = 0;
    if (cfg->method == PR_MR_LBFGS) {
      hipLaunchKernelGGL(pr::k_lbfgs_direction, dim3(1), dim3(pr::LBFGS_THREADS), 0, st, d_g, d_S, d_Y, k, first,
                         R, n, 0.0, d_d, d_sc);
      HIPCHK(hipGetLastError());
      double dmax = 0;
      if ((rc = read(d_sc, &dmax, 1))) return rc;
      if (dmax < 1e-14) {
        converged = 1;
        reason = PR_STOP_ZERO_DIRECTION;
        break;
      }
      hipLaunchKernelGGL(pr::k_lbfgs_scale, eb, et, 0, st, d_d, d_sc, n);
      HIPCHK(hipGetLastError());
      // Armijo backtracking (F only); w_new ends in d_wn
      double a = a0, Ft = 0;
      bool ok = false;
      for (int h = 1; h <= 20; h++) {
        hipLaunchKernelGGL(pr::k_mr_trial, eb, et, 0, st, d_w, d_d, a, lo, hi, n, d_wn);
        HIPCHK(hipGetLastError());
        if ((rc = eval(d_wn, nullptr, &Ft))) return rc;
        if (Ft >= F - 1e-4 * std::fabs(a)) {
          ok = true;
          n_f += h + 1;
          break;
        }
        a *= 0.5;
      }
      if (!ok) {
        hipLaunchKernelGGL(pr::k_mr_trial, eb, et, 0, st, d_w, d_d, a, lo, hi, n, d_wn);
        HIPCHK(hipGetLastError());
        if ((rc = eval(d_wn, nullptr, &Ft))) return rc;
        if (!(Ft > F) && (rc = copy(d_wn, d_w))) return rc;
        n_f += 21;
      }
      if ((rc = eval(d_wn, d_gn, &Fn))) return rc;
      n_f++;
      n_g++;
      const int slot = (first + k) % R;
      hipLaunchKernelGGL(pr::k_mr_pair, dim3(1), dim3(pr::LBFGS_THREADS), 0, st, d_w, d_wn, d_gn, d_g, n,
                         d_S + (size_t)slot * n, d_Y + (size_t)slot * n, d_sc + 1);
      HIPCHK(hipGetLastError());
      double ysss[2];
      if ((rc = read(d_sc + 1, ysss, 2))) return rc;
      if (ysss[0] > 1e-14 * ysss[1]) {
        if (k < mem)
          k++;
        else
          first = (first + 1) % R;
      }
      if ((rc = copy(d_w, d_wn)) || (rc = copy(d_g, d_gn))) return rc;
    } else {
      hipLaunchKernelGGL(pr::k_absmax, dim3(1), dim3(pr::LBFGS_THREADS), 0, st, d_g, n, d_sc);
      HIPCHK(hipGetLastError());
      double gmax = 0;
      if ((rc = read(d_sc, &gmax, 1))) return rc;
      if (gmax > 1e-14) {
        hipLaunchKernelGGL(pr::k_mr_trial, eb, et, 0, st, d_w, d_g, a0 / gmax, lo, hi, n, d_wn);
        HIPCHK(hipGetLastError());
        if ((rc = copy(d_w, d_wn))) return rc;
      }
      if ((rc = eval(d_w, d_g, &Fn))) return rc;
      n_f++;
      n_g++;
    }
    const double F_prev = F;
    F = Fn;
    if ((rc = gnorm(d_g, &gn))) return rc;
    fh.push_back(F);
    gh.push_back(gn);
    if (F > best_F) {
      best_F = F;
      if ((rc = copy(d_best, d_w))) return rc;
    }
    int why = -1;
    if (F >= 1.0 - 1e-8)
      why = PR_STOP_PERFECT;
    else if (gn < cfg->grad_tol)
      why = PR_STOP_GRADIENT;
    else if (cfg->method == PR_MR_LBFGS && std::fabs(F - F_prev) < 1e-9)
      why = PR_STOP_FIDELITY;
    if (why >= 0) {
      converged = 1;
      reason = why;
      break;
    }
  }
  HIPCHK(hipMemcpyAsync(controls_best, d_best, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::copy(fh.begin(), fh.end(), fidelity_history);
  std::copy(gh.begin(), gh.end(), gnorm_history);
  result->n_iterations = (int)fh.size() - 1;
  result->converged = converged;
  result->reason = reason;
  result->n_fidelity_history = (int)fh.size();
  result->n_fidelity_evaluations = n_f;
  result->n_gradient_evaluations = n_g;
  result->fidelity = best_F;
  result->step_size_final = 0;
  return PR_OK;
}

int pr_plan_destroy(void* plan) {
  if (!plan) return PR_OK;
  delete static_cast<Plan*>(plan);
  return PR_OK;
}

}  // extern "C"
