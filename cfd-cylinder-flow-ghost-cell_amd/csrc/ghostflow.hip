// ghostflow.hip -- MI355X (gfx950) kernels and C ABI of the ghost-cell AB2/CN cylinder step.
//
// Layout in HBM (one case):
//   velocity arrays: [u (ny, nx+1) | v (ny+1, nx)] contiguous, row-major ("vel layout");
//   pressure arrays: (ny, nx) row-major.
//   State arrays are "extended": active faces hold the unknowns, ghost faces hold the
//   quadratic-normal extension, column 0 of u holds the inlet, v walls are zero.
//   Work vectors of the reference's packed space ("q-arrays") use the same layout
//   with zeros on every inactive node, so packing is the identity.
//
// Operators applied to q-arrays are a structured stencil (the cylinder-free channel
// operator) plus "override" rows near the cylinder, taken verbatim from the host-assembled
// sparse operator (D = Dfull E, G = raw_g P / rho, I - dt/2 L, A = D G).  The separable
// rectangle inverse is an orthonormal y-eigenbasis transform (dense FP64 product) followed
// by per-mode tridiagonal solves along x (parallel cyclic reduction in LDS) and the inverse
// transform; the boundary-row (Woodbury) correction evaluates R^{-1}b only at the capacitance
// columns, applies the precomputed S^{-1} and folds the second inverse into one more
// tridiagonal sweep.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <map>
#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <initializer_list>
#include <mutex>
#include "ghostflow.h"

#define DEV __device__ __forceinline__
typedef unsigned long long u64;
// Device memory through the global address space.  Pointers read from a descriptor table are
// generic to the compiler, and generic accesses compile to flat_* instructions, which also count
// against the scalar-load wait counter; G() gives the global_* forms.
template <class T> using gptr = T __attribute__((address_space(1)))*;
template <class T> DEV gptr<T> G(T* p) { return (gptr<T>)p; }

namespace {

constexpr double kEps = 2.220446049250313e-16;
constexpr double kTiny = 2.2250738585072014e-308;

// ------------------------------------------------------------------------------------
// Buffers and scalar slots of a case
// ------------------------------------------------------------------------------------
enum BufId {
  // vel-sized
  B_S0 = 0, B_S1, B_H0, B_H1, B_RHS, B_STAR, B_RES, B_DEFECT, B_PROJ, B_HALF, B_TV0, B_TV1,
  B_CKS, B_CKH,        // checkpoint of the state and the advection history (speculative chunks)
  // p-sized
  B_PH, B_POUT, B_DIVP, B_PHI, B_RESP, B_PHI2, B_TP0, B_TP1,
  B_CKP,               // checkpoint of the half pressure
  B_KA, B_KB, B_KX, B_KR,   // GMRES work vectors (operator output, preconditioned vector, correction, residual)
  B_KV0,               // first GMRES basis vector: the head of the basis block, null until GMRES is selected
  NBUF
};
constexpr int kFirstPBuf = B_PH;

// per-solve slot block: 8 slots
enum { SV_VAL = 0, SV_RES1, SV_RES2, SV_FINAL, SV_ROUND, SV_SSQR, SV_SSQV, SV_ASCALE, SV_N };
enum Slot {
  SL_UMAX = 0, SL_VMAX, SL_DIV0, SL_DIV1, SL_DIV2, SL_COUPLED, SL_NONFINITE, SL_SPARE,
  SL_HU = 8, SL_HV = 16,             // Helmholtz u / v solve blocks
  SL_P0 = 24, SL_P1 = 32, SL_P2 = 40, // pressure solve blocks (projection 1, 2, startup)
  SL_FORCE = 48,                     // 6 doubles (written by one block, stripe 0)
  SL_GM_IT = 54, SL_GM_REF = 57, SL_GM_ST = 60,   // GMRES iterations / refinements / status of pressure solves 0-2 (stripe 0)
  NSLOT = 64
};
// Every slot is striped over NSTRIPE words so that the blocks of a reduction hit different
// addresses (same-address atomics serialise); readers reduce over the stripes.  Stripe s of slot k
// is word s * SLOT_STRIDE + k: the stripes of one slot lie 4 KB apart, so a reduction's
// device-scope atomics spread over the memory channels instead of queueing on the one that holds
// a contiguous 512-byte run, while the NSLOT slots of a stripe stay together for the readers.
constexpr int NSTRIPE = 64;
constexpr int SLOT_STRIDE = 512;
constexpr int NSLOTW = NSTRIPE * SLOT_STRIDE;
static_assert(NSLOT == 64 && NSLOT <= SLOT_STRIDE, "one slot per lane of a wave, all within a stripe");
__host__ __device__ constexpr size_t slot_word(int k, int s) { return (size_t)s * SLOT_STRIDE + k; }

// Two ints read as one 8-byte load (a CSR range, a solid interval)
struct alignas(8) IPair { int x, y; };
DEV IPair load_pair(const IPair* p, int k) {
  const long long v = G(reinterpret_cast<const long long*>(p))[k];
  return IPair{(int)(unsigned)(v & 0xffffffffll), (int)(v >> 32)};
}

struct Override {
  int j0, i0, bh, bw;
  const IPair* map;  // bh*bw: the CSR range (k0, k1) of the cell's override row, or (-1, -1)
  const int* col;
  const double* val;
};

struct GhostTab {
  int n, width;       // rows, entries per row (padded with zero weights)
  const int* node;
  const int* col;     // n*width
  const double* w;    // n*width
};

struct CaseDev {
  int nx, ny, usize, vsize, psize, velsize;
  double dx, dy, nu, rho;
  double* inlet;                // the inlet profile the step uses (u at x = 0)
  const double* inlet_base;     // inflow perturbation: base profile, amp sin(2 pi y / Ly), duration T
  const double* inlet_shape;    //   (pert_T <= 0: no perturbation; the inlet is set by mcf_set_inlet)
  double pert_T;
  const int* lo[3];
  const int* hi[3];
  const IPair* lohi[3];   // the same intervals as (lo, hi) pairs: one load per row
  int sbox[3][4];   // bounding box (j0, j1, i0, i1) of the solid nodes of each field
  Override ov[MCF_N_OPS];
  GhostTab gt[3];
  double* buf[NBUF];
  u64* sl;
  // surface quadrature
  int nsurf;
  const int *pp, *pc, *up, *uc, *vp, *vc;
  const double *pv, *uv, *vv, *nrm, *nw;
  double diameter, uref;
  double* fpart; // per-wave partial force sums (4 per wave), then the probe values
  // pressure probes: at most FORCE_PPW CSR rows over p, evaluated on the step's output pressure
  int nprobe;
  const int *qp, *qc;
  const double* qv;
  double* diag;  // diag ring
  int* clock_step;     // row of the next diagnostics entry
  double* clock_time;  // simulated time of the last completed step
  double* prm;         // per-case step parameters: dt, AB2 alpha = dt/(2 dt_prev), factor = dt/(dt+dt_prev)
};

// One separable solve problem (one rectangle of one case)
struct SolveDesc {
  int kind, yy, xx;
  int par;              // 1: V[yy-1-j][m] = (-1)^m V[j][m] (the parity-split transform applies)
  int off, ld;          // region of the rectangle inside its array: base + off, row stride ld
  int joff, ioff;       // rectangle (j,i) -> array (j+joff, i+ioff) for the solid test
  int field;            // solid-interval field of the array
  double a, b;          // operator = a I + b (lx + ly)
  const double* V;      // yy*yy, V[j*yy+k]
  const double* VT;     // yy*yy, VT[k*yy+j]
  const double* lam;    // yy
  const double* lower;  // xx-1
  const double* diag;   // xx
  const double* upper;  // xx-1
  const double* luL;    // yy*xx: sub-diagonal multipliers of the LU of each mode's x system
  const double* luI;    // yy*xx: inverse pivots
  double* F;            // yy*xx scratch
  // capacitance
  int r, npts;
  const int* pt_j;      // npts (ordered by x column, then y)
  const int* pt_x;      // npts (group order)
  double* firstK;       // npts
  const double* SC;     // r*npts: S^{-1} C, rows grouped by x column, points in group order
  double* y;            // r: the multipliers, rows grouped by x column
  int rg_n;             // number of x columns holding capacitance rows
  const int* rg_ptr;    // rg_n+1: each column's rows (group order)
  const int* rg_x;      // rg_n
  const double* Vsel;   // yy*r: V at the rows' y index, [mode][row in group order]
  const int* gpos;      // r: group position of each capacitance row (caller order)
  const int* lo;        // solid intervals of the array rows
  const int* hi;
  double* const* bufs;  // the case's buffer table (device)
  u64* sl;              // the case's slots
  int slot0;            // first slot of this solve's block
  // the capacitance (Schur) system itself, for the iterative solve and the pair dump
  const double* Sth;    // r*r: S = I + C R^{-1} E with rows and columns in boundary-angle order
  const int* Cptr;      // r+1: C over the capacitance points, CSR in row order
  const int* Ccol;      // point index in group order (as firstK)
  const double* Cval;
  const int* apos;      // r: angle position of each capacitance row
  double* dump_c;       // dump_cap*r: recorded right-hand sides C R^{-1} b (angle order)
  double* dump_y;       // dump_cap*r: recorded multipliers S^{-1} C R^{-1} b (angle order)
  int* dump_tag;        // dump_cap: 4*step + projection
  int* dump_count;
  int dump_cap;
  const int* clock;     // the case's device step counter
  // x-FFT form of the rectangle inverse: the x operator diagonalised by its DCT-IV (xf 1, pressure),
  // DST-IV (xf 2, v) or mirrored-outlet sine basis (xf 3, u) through a complex FFT of length
  // M = xx / 2, and one tridiagonal solve along y per x mode m; xf 0: not available for this problem
  int xf, fm, fns, ftn;
  const int* fst;       // 3 per radix stage: radix r, offsets of its twiddles and of its r roots in ftw
  const double* ftw;    // ftn complex: the stage twiddles w_{rL}^{q j} (q = 1..r-1, j < L) at a stage's
                        // offset + (q-1) L + j, then the roots w_r^e (e < r) of every stage
  const double* fpre;   // M complex: exp(-i pi (4n+1) / (4 xx)); u: xx complex exp(-i pi k / (2 xx))
  const double* fpost;  // M complex: sqrt(2 / xx) exp(-i pi k / xx); u: exp(-2 pi i k / xx)
  const double* yl;     // yy*xx, [j][m]: multiplier of row j of mode m's y system
  const double* yi;     // yy*xx, [j][m]: inverse pivot
  double ye;            // the y systems' off-diagonal b / dy^2
  const double* xwp;    // (point columns) x xx: basis W[x][m] at each distinct x column of the points
  const int* pt_c;      // npts: each point's row of xwp
  const double* xwf;    // rg_n x xx: the basis at the x column of each capacitance row group
  const int* yr_ptr;    // yy+1: capacitance entries of each rectangle row j
  const int* yr_t;      // r: their y index (group order), by row
  const int* yr_g;      // r: their column group (row of xwf)
  int yr_lo, yr_n;      // the band of rows yr_lo .. yr_lo + yr_n - 1 holding capacitance entries
  double* xcb;          // yr_n x xx: the correction's right-hand side c on the band rows
};

DEV int stripe_id() { return (blockIdx.x + 7 * blockIdx.y) & (NSTRIPE - 1); }
DEV double woodbury_target(double vmax) { return fmax(1e-10, 32.0 * kEps * fmax(1.0, vmax)); }

// max that propagates NaN (a NaN anywhere must fail the gates that read the maximum)
DEV double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }
// Cross-lane move of a double through DPP (its two 32-bit halves); lanes without a source lane, and
// rows outside ROW_MASK, read 0
template <int CTRL, int ROW_MASK = 0xf>
DEV double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
DEV double from_lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// Reduction over the 64 lanes of a wave in a fixed order, on the VALU (DPP) instead of LDS
// permutes: within each 16-lane row by shifts of 1, 2, 4 and 8 (lane 15 then holds the row's
// total), rows 0+1 and 2+3 by broadcasting lane 15 into the next row, then rows 0-1 into rows 2-3
// by broadcasting lane 31; lane 63 holds the total, which every lane receives.  Lanes without a
// source contribute 0, the identity of both reductions below.  All 64 lanes must be active.
template <class Op>
DEV double wave_reduce(double v, Op op) {
  v = op(v, dpp_move<0x111>(v));        // row_shr:1
  v = op(v, dpp_move<0x112>(v));        // row_shr:2
  v = op(v, dpp_move<0x114>(v));        // row_shr:4
  v = op(v, dpp_move<0x118>(v));        // row_shr:8
  v = op(v, dpp_move<0x142, 0xa>(v));   // row_bcast:15 into rows 1 and 3
  v = op(v, dpp_move<0x143, 0xc>(v));   // row_bcast:31 into rows 2 and 3
  return from_lane63(v);
}
DEV double wave_max_abs(double v) {
  return wave_reduce(fabs(v), [](double a, double b) { return nanmax(a, b); });
}
DEV double wave_sum(double v) {
  return wave_reduce(v, [](double a, double b) { return a + b; });
}
// Slot value (max over stripes, or sum); must be called by all 64 lanes of a wave.
DEV double slot_max(const u64* sl, int k) {
  const int lane = threadIdx.x & 63;
  return wave_max_abs(__longlong_as_double((long long)G(sl)[slot_word(k, lane)]));
}
DEV double slot_sum(const u64* sl, int k) {
  const int lane = threadIdx.x & 63;
  return wave_sum(__longlong_as_double((long long)G(sl)[slot_word(k, lane)]));
}
// Whether a refinement pass runs: its residual misses the Woodbury target (or is NaN), from the
// slot values over all stripes
DEV bool pred_eval(const u64* sl, int slot0, int res_slot, int val_slot) {
  const double res = slot_max(sl, slot0 + res_slot);
  const double val = slot_max(sl, val_slot);
  return res > woodbury_target(val) || res != res;
}
// The pass's decision as k_decide stored it, one word for slot k at DEC_BASE + k in stripe 0 (past
// the slots, which the step's slot clear leaves alone): a predicated kernel's waves read one word
// instead of each reducing the two slots' 64 stripes, 64 cache lines per load (0.7 us per
// predicated launch; 8.5 us of a transient step's two refinement passes).
constexpr int DEC_BASE = NSLOT;
static_assert(DEC_BASE + NSLOT <= SLOT_STRIDE, "the decision words fit in stripe 0");
// The velocity right-hand side maxima of a step (u, v), striped like the slots, in a pair per step
// parity past the decision words: SL_RHS + 2 parity + component.  A step's advection stage writes
// its pair while the previous step's slot gather may still read and clear the other one (the step
// graph runs that gather beside the next step's first stage).
constexpr int SL_RHS = DEC_BASE + NSLOT;
static_assert(SL_RHS + 4 <= SLOT_STRIDE, "the right-hand side maxima fit in a stripe");
DEV bool pred_refine(const u64* sl, int slot0, int res_slot) { return G(sl)[DEC_BASE + slot0 + res_slot] != 0ull; }
// Decide a refinement pass once its residual is in the slots: slot block k0 (and k1 when >= 0; the
// velocity solve's u and v), residual slot res_slot, right-hand side maxima in slots v0 (and v1);
// one wave per case
__global__ __launch_bounds__(64) void k_decide(const CaseDev* __restrict__ cases, int k0, int k1, int res_slot, int v0,
                                               int v1) {
  const CaseDev& c = cases[blockIdx.y];
  const bool r0 = pred_eval(c.sl, k0, res_slot, v0);
  const bool r1 = k1 >= 0 && pred_eval(c.sl, k1, res_slot, v1);
  if (threadIdx.x != 0) return;
  G(c.sl)[DEC_BASE + k0 + res_slot] = r0 ? 1ull : 0ull;
  if (k1 >= 0) G(c.sl)[DEC_BASE + k1 + res_slot] = r1 ? 1ull : 0ull;
}
// Block statistics of a 256-thread block (every thread calls): the maxima of |v[k]| and the sums of
// sv[q] over the block, one atomic per block and statistic, into dst[k] (and dst2[k] when set) and
// sdst[q].  Every thread stores its values in LDS once; wave 0 then reduces each statistic: lane l
// combines entries l, l + 64, l + 128 and l + 192, then the wave by DPP, in a fixed order.  The
// maxima are taken on the bit patterns of |v|: non-negative doubles order like their patterns and a
// NaN (sign cleared) lies above +inf, so the u64 maximum propagates NaN as the gates require.
constexpr int RED_T = 256;
DEV double max_bits(double a, double b) {
  return (u64)__double_as_longlong(a) > (u64)__double_as_longlong(b) ? a : b;
}
template <int N, int S = 0>
DEV void block_stats_atomic(const double (&v)[N], u64* const (&dst)[N], u64* const (&dst2)[N],
                            const double* sv = nullptr, double* const* sdst = nullptr) {
  __shared__ double red[N + S][RED_T];
  const int t = threadIdx.x, lane = t & 63;
  #pragma unroll
  for (int k = 0; k < N; ++k) red[k][t] = fabs(v[k]);
  #pragma unroll
  for (int q = 0; q < S; ++q) red[N + q][t] = sv[q];
  __syncthreads();
  if (t < 64) {
    const int sid = stripe_id();
    #pragma unroll
    for (int k = 0; k < N; ++k) {
      double m = red[k][lane];
      #pragma unroll
      for (int e = 1; e < RED_T / 64; ++e) m = max_bits(m, red[k][lane + 64 * e]);
      m = wave_reduce(m, [](double a, double b) { return max_bits(a, b); });
      const u64 bits = (u64)__double_as_longlong(m);
      if (lane == 0 && bits != 0) {
        if (dst[k]) atomicMax(dst[k] + slot_word(0, sid), bits);
        if (dst2[k]) atomicMax(dst2[k] + slot_word(0, sid), bits);
      }
    }
    #pragma unroll
    for (int q = 0; q < S; ++q) {
      double m = red[N + q][lane];
      #pragma unroll
      for (int e = 1; e < RED_T / 64; ++e) m += red[N + q][lane + 64 * e];
      m = wave_sum(m);
      if (lane == 0 && m != 0.0) unsafeAtomicAdd(sdst[q] + slot_word(0, sid), m);
    }
  }
}
template <int N>
DEV void block_max_atomic(const double (&v)[N], u64* const (&dst)[N]) {
  u64* const none[N] = {};
  block_stats_atomic<N>(v, dst, none);
}
DEV void block_sum_atomic(double v, double* dst) {
  __shared__ double reds[16];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  double m = wave_sum(v);
  if (lane == 0) reds[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += reds[q];
    if (s != 0.0) unsafeAtomicAdd(dst + slot_word(0, stripe_id()), s);
  }
}

// Solid test: only nodes inside the solid's bounding box read the per-row interval table
DEV bool solid(const CaseDev& c, int field, int j, int i) {
  const int* b = c.sbox[field];
  if (j < b[0] || j > b[1] || i < b[2] || i > b[3]) return false;
  const IPair lh = load_pair(c.lohi[field], j);   // one 8-byte load: the interval's two ends arrive together
  return (lh.x <= i) & (i <= lh.y);
}
DEV bool u_act(const CaseDev& c, int j, int i) { return i >= 1 && !solid(c, 0, j, i); }
DEV bool v_act(const CaseDev& c, int j, int i) { return j >= 1 && j <= c.ny - 1 && !solid(c, 1, j, i); }
DEV bool p_act(const CaseDev& c, int j, int i) { return !solid(c, 2, j, i); }

// The CSR range of the override row of cell (j, i), or (-1, -1): the map holds the range itself, so
// a row's entries are one dependent load away from the lookup
DEV IPair ov_row(const Override& o, int j, int i) {
  const IPair none{-1, -1};
  if (!o.map) return none;
  int a = j - o.j0, b = i - o.i0;
  if (a < 0 || b < 0 || a >= o.bh || b >= o.bw) return none;
  return load_pair(o.map, a * o.bw + b);
}
// Override rows hold up to 17 nonzeros.  A row is read in batches of OV_B entries: the batch's
// column and value loads go out together, then its x loads, so a batch costs two memory round trips
// instead of two per nonzero.  Indices past the row's end are clamped to its last entry (the loads
// stay in the row and need no predicate) and their products are not added; the products are added
// in nonzero order, as a plain loop over the row adds them.
constexpr int OV_B = 8;
template <class P, class F>
DEV void ov_batches(const Override& o, IPair r, P x, F use) {
  const auto col = G(o.col);
  const auto val = G(o.val);
  const int k0 = r.x, k1 = r.y;
  for (int k = k0; k < k1; k += OV_B) {
    int cc[OV_B];
    double vv[OV_B], xv[OV_B];
    #pragma unroll
    for (int u = 0; u < OV_B; ++u) {
      const int kk = min(k + u, k1 - 1);
      cc[u] = col[kk];
      vv[u] = val[kk];
    }
    #pragma unroll
    for (int u = 0; u < OV_B; ++u) xv[u] = x[cc[u]];
    #pragma unroll
    for (int u = 0; u < OV_B; ++u)
      if (k + u < k1) use(vv[u], xv[u]);
  }
}
template <class P>
DEV double ov_dot(const Override& o, IPair r, P x) {
  double acc = 0.0;
  ov_batches(o, r, x, [&](double v, double xv) { acc += v * xv; });
  return acc;
}
template <class P>
DEV void ov_dot_abs(const Override& o, IPair r, P x, double& acc, double& aacc, int& nnz) {
  acc = 0.0; aacc = 0.0; nnz = 0;
  ov_batches(o, r, x, [&](double v, double xv) {
    double t = v * xv;
    acc += t;
    aacc += fabs(v) * fabs(xv);
    nnz += (v != 0.0);
  });
}

// ------------------------------------------------------------------------------------
// Element-wise and stencil kernels run grid-stride over a bounded grid (GSB blocks per case):
// fewer, fuller blocks keep the dispatch and per-block reduction cost small.
// ------------------------------------------------------------------------------------
constexpr int GSB = 4096;

// Step parameters as kernel arguments.  A single case's captured step fixes dt (its graphs are
// rebuilt when dt changes), so its launches carry dt, the AB2 alpha and the pressure factor and no
// kernel waits on a parameter load before its first operand load; dt = 0 (batches, whose cases
// differ, and one-off launches) reads them from the case's parameter array.
struct StepPrm { double dt, alpha, factor; };
DEV double prm_dt(const CaseDev& c, const StepPrm& s) { return s.dt > 0.0 ? s.dt : G(c.prm)[0]; }
DEV double prm_alpha(const CaseDev& c, const StepPrm& s) { return s.dt > 0.0 ? s.alpha : G(c.prm)[1]; }
DEV double prm_factor(const CaseDev& c, const StepPrm& s) { return s.dt > 0.0 ? s.factor : G(c.prm)[2]; }

// The element-wise and stencil stages give each block one contiguous chunk of [lo, hi) (a multiple
// of the block size per block), and the blocks that run on one XCD consecutive chunks.  Workgroups
// go to the 8 XCDs round-robin by launch id, and a stage's local block index bx follows the launch
// id up to a constant shift, so the blocks with equal bx mod 8 share an XCD: its chunks form one
// band of the field, and the neighbour rows a chunk reads are mostly the band's own, already in
// that XCD's L2.  (A grid-stride order spread each block over the whole field: a 5-point stencil
// over 1320 x 246 took 4.2 us per graph node that way and 3.0 us in bands.)  Batched launches
// (one grid row per case, a few hundred blocks each) keep the grid-stride order, which measured
// faster there (64-case ensemble 2.36e9 against 2.27e9 cell-steps/s).
struct Chunk { int beg, end, step; };
DEV Chunk xcd_chunk(int lo, int hi, int bx, int gx) {
  if (gridDim.y > 1) return Chunk{lo + bx * (int)blockDim.x, hi, gx * (int)blockDim.x};
  const int span = gx * (int)blockDim.x;
  const int per = (hi - lo + span - 1) / span * (int)blockDim.x;
  int lb = bx >> 3;
  for (int q = 0; q < (bx & 7); ++q) lb += (gx - q + 7) >> 3;   // blocks with index = q (mod 8)
  const int beg = lo + lb * per;
  return Chunk{beg, min(hi, beg + per), (int)blockDim.x};
}

// Stage 1: advection (flux form), AB2, diffusion and the CN right-hand side
// (numba_kernels.py:17-142, 596-616; time_integration.py:30-93)
struct AdvArgs {
  int q_id, adv_id, p_id, prev_id, now_id, out_id;
  int mode;  // 0: AB2 rhs, 1: projected-midpoint, 2: rhs with explicit = adv(adv_id), 3: out = L q + offset
  StepPrm sp;
  int rhs_slot;  // rhs modes: the slot pair of the step's right-hand side maxima (u, v)
};

// SAME: the advected and the advecting fields are one buffer (the AB2 and midpoint steps), so their
// reads at one index are one load
template <bool SAME>
DEV void adv_body(const CaseDev& c, const AdvArgs& a, int bx, int gx) {
  const int nx = c.nx, ny = c.ny, nx1 = nx + 1;
  const gptr<const double> A = G(c.buf[a.adv_id]);
  const gptr<const double> Q = SAME ? A : G(c.buf[a.q_id]);
  const gptr<const double> Pp = G(c.buf[a.p_id]);
  const gptr<const double> prev = a.prev_id >= 0 ? G(c.buf[a.prev_id]) : nullptr;
  const gptr<double> now = a.now_id >= 0 ? G(c.buf[a.now_id]) : nullptr;
  const gptr<double> out = G(c.buf[a.out_id]);
  const double dx = c.dx, dy = c.dy, idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy);
  const gptr<const double> in = G(c.inlet);
  const gptr<const double> Av = A + c.usize;
  const gptr<const double> Qv = Q + c.usize;
  const double dt = prm_dt(c, a.sp), alpha = prm_alpha(c, a.sp);
  double mu = 0.0, mv = 0.0;
  // Each face loads all of its operands in one batch, at indices clamped into the arrays, and forms
  // its value without branches around the loads (inactive faces compute a value that is then
  // discarded); the override and solid lookups are issued after the batch, so the batch is in
  // flight while they resolve.  Per-term loads under the boundary tests took one round trip per
  // test.  The arithmetic is the stencil's, term for term.
  const Chunk ch = xcd_chunk(0, c.velsize, bx, gx);
  for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
    double res = 0.0, nw = 0.0;
    if (idx < c.usize) {
      const int j = idx / nx1, i = idx - j * nx1;
      const int iw = max(i - 1, 0), ie = min(i + 1, nx), ic = min(i, nx - 1);
      const int js = max(j - 1, 0), jn = min(j + 1, ny - 1);
      // u at (j, i-1 / i / i+1) and (j-1 / j+1, i); v at rows j and j+1 beside the face (v has ny+1 rows)
      const double inj = in[j];
      const double a_c = A[j * nx1 + i], a_w0 = A[j * nx1 + iw], a_e = A[j * nx1 + ie];
      const double a_s = A[js * nx1 + i], a_n = A[jn * nx1 + i];
      const double vs_w = Av[j * nx + iw], vs_e = Av[j * nx + ic];
      const double vn_w = Av[(j + 1) * nx + iw], vn_e = Av[(j + 1) * nx + ic];
      const double q = Q[j * nx1 + i], q_w0 = Q[j * nx1 + iw], q_e = Q[j * nx1 + ie];
      const double q_s = Q[js * nx1 + i], q_n = Q[jn * nx1 + i];
      const double p_e = Pp[j * nx + ic], p_w = Pp[j * nx + iw];
      const double pv = a.mode == 0 ? prev[idx] : 0.0;
      const IPair r = ov_row(c.ov[MCF_OP_GRAD_U], j, i);   // looked up together with the solid test
      const bool act = u_act(c, j, i);
      // u at column 0 is the inlet profile
      const double a_w = (i == 1) ? inj : a_w0, q_w = (i == 1) ? inj : q_w0;
      auto fxx = [&](double l, double h) { double uf = 0.5 * (l + h); return -(uf * uf * dy); };
      auto fxy = [&](bool wall, double vw, double ve, double ulo, double uhi) {
        if (wall) return 0.0;
        double vf = (i == nx) ? vw : 0.5 * (vw + ve);
        double ua = 0.5 * (ulo + uhi);
        return -(vf * ua * dx);
      };
      const double fn = fxy(j + 1 == ny, vn_w, vn_e, a_c, a_n), fs = fxy(j == 0, vs_w, vs_e, a_s, a_c);
      double ru;
      if (i < nx) ru = (fxx(a_c, a_e) - fxx(a_w, a_c)) + (fn - fs);
      else ru = -(a_c * a_c * dy) - fxx(a_w, a_c) + 0.5 * (fn - fs);
      double adv = ru / (dx * dy);
      if (i == nx) adv *= 2.0;
      double lap = (-2.0 * idx2 - ((j == 0 || j == ny - 1) ? 3.0 : 2.0) * idy2) * q
                   + ((i == nx) ? 2.0 : 1.0) * idx2 * q_w;
      if (i < nx) lap += idx2 * q_e;
      if (j < ny - 1) lap += idy2 * q_n;
      if (j > 0) lap += idy2 * q_s;
      const double diff = c.nu * lap;
      double grad = ((i < nx) ? (p_e - p_w) / dx : -2.0 * p_w / dx) / c.rho;
      if (r.x >= 0) grad = ov_dot(c.ov[MCF_OP_GRAD_U], r, Pp);
      const double doff = (i == 1) ? c.nu * inj * idx2 : 0.0;
      if (a.mode == 0) {
        nw = adv;
        const double ex = (1.0 + alpha) * adv - alpha * pv;
        res = q + dt * (ex + 0.5 * diff - grad) + 0.5 * dt * doff;
      } else if (a.mode == 1) {
        nw = adv;
        res = ((adv + diff) - grad) * (0.5 * dt) + q;
      } else if (a.mode == 3) {
        nw = adv;
        res = diff;
      } else {
        res = q + dt * (adv + 0.5 * diff - grad) + 0.5 * dt * doff;
      }
      res = act ? res : 0.0;
      nw = act ? nw : 0.0;
      mu = nanmax(mu, fabs(res));
    } else {
      const int t = idx - c.usize;
      const int j = t / nx, i = t - j * nx;
      // rows clamped for the wall faces j = 0 and j = ny (inactive): v has ny+1 rows, u and p ny
      const int js = max(j - 1, 0), jn = min(j + 1, ny), ju = min(j, ny - 1);
      const int iw = max(i - 1, 0), ie = min(i + 1, nx - 1);
      // v at (j, i-1 / i / i+1) and (j-1 / j+1, i); u at columns i and i+1 of rows j-1 and j
      const double v_c = Av[j * nx + i], v_n = Av[jn * nx + i], v_s = Av[js * nx + i];
      const double v_e = Av[j * nx + ie], v_w = Av[j * nx + iw];
      const double ue_s = A[js * nx1 + i + 1], ue_n = A[ju * nx1 + i + 1];
      const double uw_s = A[js * nx1 + i], uw_n = A[ju * nx1 + i];
      const double q = Qv[j * nx + i], q_n = Qv[jn * nx + i], q_s = Qv[js * nx + i];
      const double q_e = Qv[j * nx + ie], q_w = Qv[j * nx + iw];
      const double p_c = Pp[ju * nx + i], p_s = Pp[js * nx + i];
      const double pv = a.mode == 0 ? prev[idx] : 0.0;
      const IPair r = ov_row(c.ov[MCF_OP_GRAD_V], j, i);
      const bool act = v_act(c, j, i);
      auto fyy = [&](double lo, double hi) { double vf = 0.5 * (lo + hi); return -(vf * vf * dx); };
      double fe;
      if (i + 1 == nx) { double uo = 0.5 * (ue_s + ue_n); fe = -(uo * v_c * dy); }
      else { double uf = 0.5 * (ue_s + ue_n); double va = 0.5 * (v_c + v_e); fe = -(uf * va * dy); }
      double fw = 0.0;
      if (i != 0) { double uf = 0.5 * (uw_s + uw_n); double va = 0.5 * (v_w + v_c); fw = -(uf * va * dy); }
      const double rv = (fe - fw) + (fyy(v_c, v_n) - fyy(v_s, v_c));
      const double adv = rv / (dx * dy);
      double lap = (-((i == 0) ? 3.0 : ((i == nx - 1) ? 1.0 : 2.0)) * idx2 - 2.0 * idy2) * q
                   + idy2 * (q_n + q_s);
      if (i < nx - 1) lap += idx2 * q_e;
      if (i > 0) lap += idx2 * q_w;
      const double diff = c.nu * lap;
      double grad = (p_c - p_s) / dy / c.rho;
      if (r.x >= 0) grad = ov_dot(c.ov[MCF_OP_GRAD_V], r, Pp);
      if (a.mode == 0) {
        nw = adv;
        const double ex = (1.0 + alpha) * adv - alpha * pv;
        res = q + dt * (ex + 0.5 * diff - grad);
      } else if (a.mode == 1) {
        nw = adv;
        res = ((adv + diff) - grad) * (0.5 * dt) + q;
      } else if (a.mode == 3) {
        nw = adv;
        res = diff;
      } else {
        res = q + dt * (adv + 0.5 * diff - grad);
      }
      res = act ? res : 0.0;
      nw = act ? nw : 0.0;
      mv = nanmax(mv, fabs(res));
    }
    if (now) now[idx] = nw;
    out[idx] = res;
  }
  double v2[2] = {mu, mv};
  const bool rhs_mode = a.mode == 0 || a.mode == 2;
  u64* const d2[2] = {rhs_mode ? c.sl + a.rhs_slot : nullptr, rhs_mode ? c.sl + (a.rhs_slot + 1) : nullptr};
  block_max_atomic<2>(v2, d2);
}
__global__ __launch_bounds__(256) void k_adv_rhs(const CaseDev* __restrict__ cases, AdvArgs a) {
  const CaseDev& c = cases[blockIdx.y];
  if (a.q_id == a.adv_id) adv_body<true>(c, a, blockIdx.x, gridDim.x);
  else adv_body<false>(c, a, blockIdx.x, gridDim.x);
}

// Helmholtz residual / defect on vel q-arrays: H x = x - dt/2 L x
// (time_integration.py:71-74, 97-119; numba_kernels.py:634-642)
struct HelmArgs {
  int b_id, x_id, out_id, out2_id;
  int mode;   // 0: residual (out = b - Hx, slot RES_k), 1: final (max only), 2: defect
  int res_slot;  // SV_RES1/SV_RES2/SV_FINAL
  int pred_slot; // -1 always, else run only where that refinement happened
  StepPrm sp;
};

DEV void helm_body(const CaseDev& c, const HelmArgs& a, int bx, int gx) {
  const int nx = c.nx, ny = c.ny, nx1 = nx + 1;
  const bool run_u = a.pred_slot < 0 || pred_refine(c.sl, SL_HU, a.pred_slot);
  const bool run_v = a.pred_slot < 0 || pred_refine(c.sl, SL_HV, a.pred_slot);
  const int lo = run_u ? 0 : c.usize, hi = run_v ? c.velsize : c.usize;
  const Chunk ch = xcd_chunk(lo, hi, bx, gx);
  if (!(run_u || run_v) || ch.beg >= ch.end) return;   // no element: nothing to reduce
  double mu = 0.0, mv = 0.0, xu = 0.0, xv = 0.0;
  {
    const gptr<const double> B = G(c.buf[a.b_id]);
    const gptr<const double> X = G(c.buf[a.x_id]);
    const gptr<const double> Xv = X + c.usize;
    const gptr<double> O = a.out_id >= 0 ? G(c.buf[a.out_id]) : nullptr;
    const gptr<double> O2 = a.out2_id >= 0 ? G(c.buf[a.out2_id]) : nullptr;
    const double idx2 = 1.0 / (c.dx * c.dx), idy2 = 1.0 / (c.dy * c.dy);
    const double h = -0.5 * prm_dt(c, a.sp);
    // a face's stencil operands and right-hand side load in one batch at clamped indices, ahead of
    // the activity test (as in the advection stage); override rows replace the stencil value
    for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
      double res = 0.0, r2 = 0.0;
      const double bv = B[idx];
      if (idx < c.usize) {
        const int j = idx / nx1, i = idx - j * nx1;
        const double x = X[j * nx1 + i], xw = X[j * nx1 + max(i - 1, 0)], xe = X[j * nx1 + min(i + 1, nx)];
        const double xn = X[min(j + 1, ny - 1) * nx1 + i], xs = X[max(j - 1, 0) * nx1 + i];
        const IPair r = ov_row(c.ov[MCF_OP_HELM_U], j, i);   // looked up together with the solid test
        const bool act = u_act(c, j, i);
        double lap = (-2.0 * idx2 - ((j == 0 || j == ny - 1) ? 3.0 : 2.0) * idy2) * x
                     + ((i == nx) ? 2.0 : 1.0) * idx2 * xw;
        if (i < nx) lap += idx2 * xe;
        if (j < ny - 1) lap += idy2 * xn;
        if (j > 0) lap += idy2 * xs;
        double hx = x + h * (c.nu * lap);
        if (act && r.x >= 0) hx = ov_dot(c.ov[MCF_OP_HELM_U], r, X);
        if (act) {
          res = bv - hx;
          if (a.mode == 2) { r2 = x + res; xu = nanmax(xu, fabs(x)); }
        }
        mu = nanmax(mu, fabs(res));
      } else {
        const int t = idx - c.usize;
        const int j = t / nx, i = t - j * nx;
        // rows clamped for the wall faces j = 0 and j = ny (inactive)
        const double x = Xv[j * nx + i], xn = Xv[min(j + 1, ny) * nx + i], xs = Xv[max(j - 1, 0) * nx + i];
        const double xe = Xv[j * nx + min(i + 1, nx - 1)], xw = Xv[j * nx + max(i - 1, 0)];
        const IPair r = ov_row(c.ov[MCF_OP_HELM_V], j, i);
        const bool act = v_act(c, j, i);
        double lap = (-((i == 0) ? 3.0 : ((i == nx - 1) ? 1.0 : 2.0)) * idx2 - 2.0 * idy2) * x
                     + idy2 * (xn + xs);
        if (i < nx - 1) lap += idx2 * xe;
        if (i > 0) lap += idx2 * xw;
        double hx = x + h * (c.nu * lap);
        if (act && r.x >= 0) hx = ov_dot(c.ov[MCF_OP_HELM_V], r, X);
        if (act) {
          res = bv - hx;
          if (a.mode == 2) { r2 = x + res; xv = nanmax(xv, fabs(x)); }
        }
        mv = nanmax(mv, fabs(res));
      }
      if (a.mode == 0 || a.mode == 2) O[idx] = res;
      if (a.mode == 2) O2[idx] = r2;
    }
  }
  if (a.mode == 2) {
    double v2[2] = {xu, xv};
    u64* const d2[2] = {c.sl + SL_UMAX, c.sl + SL_VMAX};
    block_max_atomic<2>(v2, d2);
  } else {
    // mode 3 (lean solve, no refinement): the same residual is the first and the final one
    const double v2[2] = {mu, mv};
    u64* const d2[2] = {run_u ? c.sl + (SL_HU + a.res_slot) : nullptr,
                        run_v ? c.sl + (SL_HV + a.res_slot) : nullptr};
    u64* const f2[2] = {a.mode == 3 ? c.sl + (SL_HU + SV_FINAL) : nullptr,
                        a.mode == 3 ? c.sl + (SL_HV + SV_FINAL) : nullptr};
    block_stats_atomic<2>(v2, d2, f2);
  }
}
__global__ __launch_bounds__(256) void k_helm(const CaseDev* __restrict__ cases, HelmArgs a) {
  helm_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// Divergence of a vel q-array: (D q + div_offset) * scale  (optimized_ghost_cell.py:212-216)
struct DivArgs { int x_id, out_id; double dtmul; int slot; StepPrm sp; };   // dtmul 0: no scaling, else / (dtmul dt)

DEV void div_body(const CaseDev& c, const DivArgs& a, int bx, int gx) {
  const Chunk ch = xcd_chunk(0, c.psize, bx, gx);
  if (ch.beg >= ch.end) return;   // no cell of this case for this block
  const int nx = c.nx, nx1 = nx + 1;
  const gptr<const double> X = G(c.buf[a.x_id]);
  const gptr<const double> Xv = X + c.usize;
  const gptr<double> O = a.out_id >= 0 ? G(c.buf[a.out_id]) : nullptr;
  const gptr<const double> inl = G(c.inlet);
  const double scale = a.dtmul == 0.0 ? 1.0 : 1.0 / (a.dtmul * prm_dt(c, a.sp));
  double m = 0.0;
  for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
    const int j = idx / nx, i = idx - j * nx;
    double d = 0.0;
    // the cell's four faces (and the inlet value) load in one batch ahead of the activity test
    const double inj = inl[j], xl = X[j * nx1 + i], xr = X[j * nx1 + i + 1];
    const double vt = Xv[(j + 1) * nx + i], vb = Xv[j * nx + i];
    const IPair r = ov_row(c.ov[MCF_OP_DIV], j, i);   // looked up together with the solid test
    const bool act = p_act(c, j, i);
    const double ul = (i == 0) ? inj : xl;
    double dc = (xr - ul) / c.dx + (vt - vb) / c.dy;
    if (act && r.x >= 0) {
      const double off = (i == 0) ? -inj / c.dx : 0.0;
      dc = ov_dot(c.ov[MCF_OP_DIV], r, X) + off;
    }
    if (act) d = dc * scale;
    if (O) O[idx] = d;
    m = nanmax(m, fabs(d));
  }
  double v1[1] = {m};
  u64* const d1[1] = {c.sl + a.slot};
  block_max_atomic<1>(v1, d1);
}
__global__ __launch_bounds__(256) void k_div(const CaseDev* __restrict__ cases, DivArgs a) {
  div_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// Projection update out = q - dt G phi (numba_kernels.py:619-623), with the optional
// coupled momentum residual -defect + dt G phi (time_integration.py:137-142)
struct UpdArgs { int x_id, phi_id, out_id, defect_id; double dtmul; StepPrm sp; };   // dt = dtmul * case dt

DEV void upd_body(const CaseDev& c, const UpdArgs& a, int bx, int gx) {
  const Chunk ch = xcd_chunk(0, c.velsize, bx, gx);
  if (ch.beg >= ch.end) return;   // no face of this case for this block
  const int nx = c.nx, nx1 = nx + 1;
  const gptr<const double> X = G(c.buf[a.x_id]);
  const gptr<const double> Pp = G(c.buf[a.phi_id]);
  const gptr<double> O = G(c.buf[a.out_id]);
  const gptr<const double> Df = a.defect_id >= 0 ? G(c.buf[a.defect_id]) : nullptr;
  const double dt = a.dtmul * prm_dt(c, a.sp);
  double m = 0.0;
  const gptr<const double> Dl = Df ? Df : X;   // read unconditionally; used only with a defect
  // a face's pressure pair, old value and defect load in one batch at clamped indices, ahead of
  // the activity test; override rows replace the stencil gradient
  for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
    double out = 0.0, g;
    bool act;
    const double xv = X[idx], dv = Dl[idx];
    if (idx < c.usize) {
      const int j = idx / nx1, i = idx - j * nx1;
      const double pe = Pp[j * nx + min(i, nx - 1)], pw = Pp[j * nx + max(i - 1, 0)];
      const IPair r = ov_row(c.ov[MCF_OP_GRAD_U], j, i);   // looked up together with the solid test
      act = u_act(c, j, i);
      g = ((i < nx) ? (pe - pw) / c.dx : -2.0 * pw / c.dx) / c.rho;
      if (act && r.x >= 0) g = ov_dot(c.ov[MCF_OP_GRAD_U], r, Pp);
    } else {
      const int t = idx - c.usize;
      const int j = t / nx, i = t - j * nx;
      // rows clamped for the wall faces j = 0 and j = ny (inactive)
      const double pc = Pp[min(j, c.ny - 1) * nx + i], ps = Pp[max(j - 1, 0) * nx + i];
      const IPair r = ov_row(c.ov[MCF_OP_GRAD_V], j, i);
      act = v_act(c, j, i);
      g = (pc - ps) / c.dy / c.rho;
      if (act && r.x >= 0) g = ov_dot(c.ov[MCF_OP_GRAD_V], r, Pp);
    }
    if (act) out = xv - dt * g;
    O[idx] = out;
    if (Df && act) m = nanmax(m, fabs(-dv + dt * g));
  }
  if (Df) {
    double v1[1] = {m};
    u64* const d1[1] = {c.sl + SL_COUPLED};
    block_max_atomic<1>(v1, d1);
  }
}
__global__ __launch_bounds__(256) void k_update(const CaseDev* __restrict__ cases, UpdArgs a) {
  upd_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// Pressure residual b - A phi and the monitored-solve statistics (pressure_schur.py:250-273)
struct PresArgs { int b_id, x_id, out_id; int mode; int slot0; int res_slot; int pred_slot; };

// Row (j, i) of the pressure operator A applied to X: (A X)_row, sum |a||x| over the row and the
// row's nonzero count.  Override rows come from their CSR rows; the others are the rectangle's
// 5-point stencil (outlet column Dirichlet, other walls Neumann).
// The stencil's five values load in one batch (a missing neighbour reads the centre, unused) and
// the stencil row is formed for every row, so that the loads go out ahead of the override test.
template <class P>
DEV void p_row(const CaseDev& c, P X, int idx, int j, int i, IPair r, double& ax, double& aax, int& nnz) {
  const int nx = c.nx, ny = c.ny;
  const double x = X[idx];
  const double tw = X[i > 0 ? idx - 1 : idx], te = X[i < nx - 1 ? idx + 1 : idx];
  const double ts = X[j > 0 ? idx - nx : idx], tn = X[j < ny - 1 ? idx + nx : idx];
  const double idx2 = 1.0 / (c.dx * c.dx), idy2 = 1.0 / (c.dy * c.dy);
  const double ox = idx2 / c.rho, oy = idy2 / c.rho;
  const double cx = (i == 0) ? -1.0 : ((i == nx - 1) ? -3.0 : -2.0);
  const double cy = (j == 0 || j == ny - 1) ? -1.0 : -2.0;
  const double dg = (cx * idx2 + cy * idy2) / c.rho;
  ax = dg * x; aax = fabs(dg) * fabs(x); nnz = 1;
  if (i > 0) { ax += ox * tw; aax += ox * fabs(tw); ++nnz; }
  if (i < nx - 1) { ax += ox * te; aax += ox * fabs(te); ++nnz; }
  if (j > 0) { ax += oy * ts; aax += oy * fabs(ts); ++nnz; }
  if (j < ny - 1) { ax += oy * tn; aax += oy * fabs(tn); ++nnz; }
  if (r.x >= 0) ov_dot_abs(c.ov[MCF_OP_PRESSURE], r, X, ax, aax, nnz);
}

DEV void pres_body(const CaseDev& c, const PresArgs& a, int bx, int gx) {
  const Chunk ch = xcd_chunk(0, c.psize, bx, gx);
  if (ch.beg >= ch.end) return;   // no cell of this case for this block
  if (a.pred_slot >= 0 && !pred_refine(c.sl, a.slot0, a.pred_slot)) return;
  const int nx = c.nx;
  const gptr<const double> B = G(c.buf[a.b_id]);
  const gptr<const double> X = G(c.buf[a.x_id]);
  const gptr<double> O = a.mode == 0 ? G(c.buf[a.out_id]) : nullptr;
  double mres = 0.0, mrnd = 0.0, msc = 0.0, ssr = 0.0, ssv = 0.0;
  for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
    const int j = idx / nx, i = idx - j * nx;
    double res = 0.0;
    const IPair orow = ov_row(c.ov[MCF_OP_PRESSURE], j, i);   // looked up together with the solid test
    const bool act = p_act(c, j, i);
    const double b = B[idx];
    double ax, aax; int nnz;
    p_row(c, X, idx, j, i, orow, ax, aax, nnz);   // ahead of the activity test: one batch of loads
    if (act) {
      res = b - ax;
      const double sc = aax + fabs(b);
      mres = nanmax(mres, fabs(res));
      msc = nanmax(msc, sc);
      mrnd = nanmax(mrnd, (2.0 * nnz + 2.0) * kEps * sc);
      ssr += res * res;
      ssv += b * b;
    }
    if (O) O[idx] = res;
  }
  u64* s = c.sl;
  const int b0 = a.slot0;
  if (a.mode == 0) {
    double v1[1] = {mres};
    u64* const d1[1] = {s + (b0 + a.res_slot)};
    block_max_atomic<1>(v1, d1);
  } else {
    // mode 1: final statistics; mode 3 (lean solve, no refinement): first residual and final
    // statistics of the same x in one pass
    const double v3[3] = {mres, mrnd, msc};
    u64* const d3[3] = {s + (b0 + SV_FINAL), s + (b0 + SV_ROUND), s + (b0 + SV_ASCALE)};
    u64* const r3[3] = {a.mode == 3 ? s + (b0 + a.res_slot) : nullptr, nullptr, nullptr};
    const double sums[2] = {ssr, ssv};
    double* const sd[2] = {reinterpret_cast<double*>(s + (b0 + SV_SSQR)),
                           reinterpret_cast<double*>(s + (b0 + SV_SSQV))};
    block_stats_atomic<3, 2>(v3, d3, r3, sums, sd);
  }
}
__global__ __launch_bounds__(256) void k_pres(const CaseDev* __restrict__ cases, PresArgs a) {
  pres_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------------------------------
// Separable rectangle inverse: dense y transform (orthonormal eigenbasis), x tridiagonal
// ------------------------------------------------------------------------------------
struct GemmArgs {
  int forward;   // 1: F = VT * in ; 0: out = V * F (epilogue)
  int in_id, out_id;
  int epi;       // 0 store, 1 masked store, 2 masked accumulate
  int pred_slot; // -1: always
  int slot_shift; // added to the desc's slot block (pressure solve index)
  const int* gate = nullptr;   // when set, the launch runs only while *gate != 0
};

DEV gptr<const double> desc_in(const SolveDesc& d, int id) { return G(G(d.bufs)[id]) + d.off; }
DEV gptr<double> desc_out(const SolveDesc& d, int id) { return G(G(d.bufs)[id]) + d.off; }

typedef double d4 __attribute__((ext_vector_type(4)));

// XCD-aware tile order: workgroups go to the 8 XCDs round-robin by linear id, so each XCD is given
// a contiguous run of tiles (problem slowest, then column panel, row-block fastest) and the A and B
// panels its tiles share stay in that XCD's L2.  A bijection of [0, T).
DEV void xcd_tile(int& z, int& cx, int& ry) {
  const int gx = gridDim.x, gy = gridDim.y, T = gx * gy * gridDim.z;
  const int id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const int xcd = id & 7;
  int t = id >> 3;
  for (int q = 0; q < xcd; ++q) t += (T - q + 7) >> 3;
  z = t / (gx * gy);
  const int rem = t - z * gx * gy;
  cx = rem / gy;
  ry = rem - cx * gy;
}

// One output element of a transform: forward into the scratch F; inverse into the output array
// with the solid mask (epi 0 store, 1 masked store, 2 masked accumulate)
DEV void gemm_put(const SolveDesc& d, const GemmArgs& g, int gr, int gc, double v) {
  if (g.forward) { G(d.F)[(size_t)gr * d.xx + gc] = v; return; }
  const gptr<double> O = desc_out(d, g.out_id);
  const int ai = gc + d.ioff, aj = gr + d.joff;
  const int slo = d.lo ? G(d.lo)[aj] : 1, shi = d.lo ? G(d.hi)[aj] : 0;
  const gptr<double> pp = O + (size_t)gr * d.ld + gc;
  if (g.epi == 0) *pp = v;
  else if (slo <= ai && ai <= shi) *pp = 0.0;
  else *pp = (g.epi == 2) ? (*pp + v) : v;
}
// The inverse's put with its operands loaded ahead: the output row's solid interval (slo, shi) and,
// when accumulating (epi 2), the old value
DEV void gemm_put_pre(const SolveDesc& d, const GemmArgs& g, int gr, int gc, double v, int slo, int shi, double old) {
  const gptr<double> pp = desc_out(d, g.out_id) + (size_t)gr * d.ld + gc;
  const int ai = gc + d.ioff;
  if (g.epi == 0) *pp = v;
  else if (slo <= ai && ai <= shi) *pp = 0.0;
  else *pp = (g.epi == 2) ? (old + v) : v;
}

// Dense FP64 y transform on the matrix cores: C (M x N) = A (M x K) * B (K x N), M = K = yy, N = xx.
// A^T is read in row-major (k, m) order: forward (A = V^T) reads V, inverse (A = V) reads V^T.
// Each block stages KC-deep K chunks of its A and B tiles in LDS; the next chunk is fetched into
// registers while the current one feeds the MFMAs, so a transform costs ceil(K/KC) memory round
// trips.  Fragment layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[4*r + l/16][l%16] in accumulator element r.
// Block tile TM x TN (4 waves in 2 x 2, each (TM/2) x (TN/2) of 16 x 16 MFMA tiles) with a K chunk
// of KC staged in LDS; the autotuner picks the shape per context.  Every shape accumulates each
// output element over k in the same order, so all shapes give identical results.
template <int TM, int TN, int KC>
__global__ __launch_bounds__(256) void k_gemm(const SolveDesc* __restrict__ descs, GemmArgs g) {
  constexpr int MT = TM / 32, NT = TN / 32;
  int z, cx, ry;
  xcd_tile(z, cx, ry);
  const SolveDesc& d = descs[z];
  const int M = d.yy, N = d.xx, K = d.yy;
  const int row0 = ry * TM, col0 = cx * TN;
  if (row0 >= M || col0 >= N) return;
  if (g.gate && *g.gate == 0) return;
  if (g.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + g.slot_shift, g.pred_slot)) return;
  const gptr<const double> AT = G(g.forward ? d.V : d.VT);
  gptr<const double> Bm;
  int ldb;
  if (g.forward) { Bm = desc_in(d, g.in_id); ldb = d.ld; }
  else { Bm = G(d.F); ldb = d.xx; }
  __shared__ double As[KC][TM];
  __shared__ double Bs[KC][TN];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int kq = lane >> 4, l16 = lane & 15;
  constexpr int NA = KC * TM / 256, NB = KC * TN / 256;
  double ra[NA], rb[NB];
  auto gload = [&](int k0) {
    #pragma unroll
    for (int t = 0; t < NA; ++t) {
      const int e = tid + 256 * t, kk = k0 + e / TM, mm = row0 + e % TM;
      ra[t] = (kk < K && mm < M) ? AT[(size_t)kk * M + mm] : 0.0;
    }
    #pragma unroll
    for (int t = 0; t < NB; ++t) {
      const int e = tid + 256 * t, kk = k0 + e / TN, nn = col0 + e % TN;
      rb[t] = (kk < K && nn < N) ? Bm[(size_t)kk * ldb + nn] : 0.0;
    }
  };
  auto sstore = [&]() {
    #pragma unroll
    for (int t = 0; t < NA; ++t) { const int e = tid + 256 * t; As[e / TM][e % TM] = ra[t]; }
    #pragma unroll
    for (int t = 0; t < NB; ++t) { const int e = tid + 256 * t; Bs[e / TN][e % TN] = rb[t]; }
  };
  d4 acc[MT][NT];
  #pragma unroll
  for (int mi = 0; mi < MT; ++mi)
    #pragma unroll
    for (int ni = 0; ni < NT; ++ni) acc[mi][ni] = d4{0.0, 0.0, 0.0, 0.0};
  gload(0);
  for (int k0 = 0; k0 < K; k0 += KC) {
    sstore();
    __syncthreads();
    if (k0 + KC < K) gload(k0 + KC);
    #pragma unroll
    for (int t = 0; t < KC / 4; ++t) {
      const int kl = 4 * t + kq;
      double a[MT];
      #pragma unroll
      for (int mi = 0; mi < MT; ++mi) a[mi] = As[kl][(TM / 2) * wm + 16 * mi + l16];
      #pragma unroll
      for (int ni = 0; ni < NT; ++ni) {
        const double b = Bs[kl][(TN / 2) * wn + 16 * ni + l16];
        #pragma unroll
        for (int mi = 0; mi < MT; ++mi) acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi], b, acc[mi][ni], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  #pragma unroll
  for (int mi = 0; mi < MT; ++mi) {
    const int rbase = row0 + (TM / 2) * wm + 16 * mi + kq;
    #pragma unroll
    for (int ni = 0; ni < NT; ++ni) {
      const int gc = col0 + (TN / 2) * wn + 16 * ni + l16;
      if (gc >= N) continue;
      #pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = rbase + 4 * r;
        if (gr < M) gemm_put(d, g, gr, gc, acc[mi][ni][r]);
      }
    }
  }
}

// Split-K transform (launch variants 3-6): a block of S waves computes a (16 MT) x 32 output tile;
// wave w takes the w-th of S consecutive K slices, holds the whole tile in 2 MT independent
// 16 x 16 accumulators and reads its MFMA operands straight from global memory (the panels are L2
// resident).  A slice's loads are issued together in batches of SK_KB k-groups (one batch covers
// the slice for K <= 4 S SK_KB), so a wave waits for memory once and then runs its MFMAs back to
// back -- no LDS staging and no barrier in the K loop.  The slices' partial tiles are added in LDS
// in slice order.  Both tile heights slice K alike and give identical results for one S; the sum
// order differs from the staged tiles' (and between S = 8 and 4), so the variants agree to
// round-off.  The 16-row tile doubles the number of blocks, which spreads the matrix-core work more
// evenly over the CUs when the 32-row grid holds only a little more than one block per CU; the
// 4-wave blocks (S = 4: half the LDS, twice the K per wave) fit twice as many blocks per CU, for
// grids larger than one round of the 8-wave blocks (the two velocity problems at 1320x246).
constexpr int SK_S = 8, SK_KB = 8;
template <int MT, int S>
__global__ __launch_bounds__(64 * S) void k_gemm_sk(const SolveDesc* __restrict__ descs, GemmArgs g) {
  int z, cx, ry;
  xcd_tile(z, cx, ry);
  const SolveDesc& d = descs[z];
  const int M = d.yy, N = d.xx, K = d.yy;
  const int row0 = ry * (16 * MT), col0 = cx * 32;
  if (row0 >= M || col0 >= N) return;
  if (g.gate && *g.gate == 0) return;
  if (g.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + g.slot_shift, g.pred_slot)) return;
  const gptr<const double> AT = G(g.forward ? d.V : d.VT);
  gptr<const double> Bm;
  int ldb;
  if (g.forward) { Bm = desc_in(d, g.in_id); ldb = d.ld; }
  else { Bm = G(d.F); ldb = d.xx; }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, kq = lane >> 4, l16 = lane & 15;
  const int NG = (K + 3) >> 2, per = (NG + S - 1) / S;
  const int g0 = min(NG, w * per), g1 = min(NG, g0 + per);
  // operand rows / columns past the matrix are read at the last one (in bounds, no predicate): they
  // feed only output rows / columns that are not stored
  int ra[MT];
  #pragma unroll
  for (int mi = 0; mi < MT; ++mi) ra[mi] = min(row0 + 16 * mi + l16, M - 1);
  const int cb[2] = {min(col0 + l16, N - 1), min(col0 + 16 + l16, N - 1)};
  d4 acc[MT][2];
  #pragma unroll
  for (int mi = 0; mi < MT; ++mi)
    #pragma unroll
    for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = d4{0.0, 0.0, 0.0, 0.0};
  // k-group q covers k = 4q .. 4q+3; lane l holds A[m = l%16][k = 4q + l/16] and B[k][n = l%16]
  for (int gb = g0; gb < g1; gb += SK_KB) {
    double a[SK_KB][MT], b[SK_KB][2];
    #pragma unroll
    for (int t = 0; t < SK_KB; ++t) {
      const int kc = min(4 * (gb + t) + kq, K - 1);   // a valid k; past the slice or K it is zeroed below
      const gptr<const double> Ar = AT + kc * M;
      const gptr<const double> Br = Bm + kc * ldb;
      #pragma unroll
      for (int q = 0; q < MT; ++q) a[t][q] = Ar[ra[q]];
      #pragma unroll
      for (int q = 0; q < 2; ++q) b[t][q] = Br[cb[q]];
    }
    // every load of the batch goes out before the first MFMA (the scheduler would otherwise
    // interleave them and expose a memory round trip per k-group); each k-group's MFMAs then wait
    // only for that group's loads, so they run while the later loads arrive
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int t = 0; t < SK_KB; ++t) {
      const int kk = 4 * (gb + t) + kq;
      const bool ok = gb + t < g1 && kk < K;   // k past the slice or past K contributes zero
      #pragma unroll
      for (int mi = 0; mi < MT; ++mi)
        #pragma unroll
        for (int ni = 0; ni < 2; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? a[t][mi] : 0.0, b[t][ni], acc[mi][ni], 0, 0, 0);
    }
  }
  // partial tiles of the slices: red[slice][subtile][element][lane]; wave w then finishes
  // elements E (w / NS) .. E (w / NS) + E - 1 of subtile w % NS
  constexpr int NS = 2 * MT, E = 4 * NS / S;
  static_assert(S >= NS && E * S == 4 * NS, "the waves split the (subtile, element) pairs evenly");
  __shared__ double red[S][NS][4][64];
  #pragma unroll
  for (int mi = 0; mi < MT; ++mi)
    #pragma unroll
    for (int ni = 0; ni < 2; ++ni)
      #pragma unroll
      for (int r = 0; r < 4; ++r) red[w][2 * mi + ni][r][lane] = acc[mi][ni][r];
  __syncthreads();
  const int st = w % NS, mi = st >> 1, ni = st & 1;
  const int gc = col0 + 16 * ni + l16;
  if (gc >= N) return;
  #pragma unroll
  for (int e = 0; e < E; ++e) {
    const int r = E * (w / NS) + e;
    const int gr = row0 + 16 * mi + kq + 4 * r;
    double v = red[0][st][r][lane];
    #pragma unroll
    for (int q = 1; q < S; ++q) v += red[q][st][r][lane];
    if (gr < M) gemm_put(d, g, gr, gc, v);
  }
}

// Parity-split transform (launch variants 8 and 9: S = 8 and 4 waves).  The y eigenbases are DCT-II,
// DST-II and DST-I columns, for which V[yy-1-j][m] = (-1)^m V[j][m] (the library checks this when a
// rectangle is set, and offers these variants only where it holds).  With h = yy/2 mirrored row
// pairs (j, yy-1-j) and, for odd yy, the middle row h:
//   forward  G[m] = sum_{j<h} V[j][m] (F[j] + (-1)^m F[yy-1-j])  (+ V[h][m] F[h] for even m),
//   inverse  P[j] = sum_{m even} V[j][m] G[m],  Q[j] = sum_{m odd} V[j][m] G[m],
//            out[j] = P + Q,  out[yy-1-j] = P - Q  (out[h] = P),
// i.e. two half-size products instead of one full one: half the MFMAs and half the operand loads of
// the dense transform.  Forward: a block computes 16 modes of one parity (m = 2i + p) x 32 columns,
// its S waves splitting the h (+1) row pairs, the B operand formed from a row and its mirror as it
// is loaded.  Inverse: a block computes 16 top rows x 32 columns; waves 0 .. S/2-1 accumulate P over
// the even modes and waves S/2 .. S-1 Q over the odd ones, and the epilogue writes both rows of each
// pair.  Partial tiles are added in LDS in slice order; the results agree with the dense variants to
// round-off.  Mode order in F is the natural one, so the sweeps and capacitance kernels are unchanged.
// KB k-groups per load batch: 4 keeps the forward kernel at 80-96 VGPRs (5-6 waves per SIMD, so a
// velocity launch's 1344 blocks take two rounds instead of three); 8 (140 VGPRs) measured 7.1 against
// 6.45 us per pressure transform and 0.197 against 0.189 ms per step.
template <int S, bool FWD, int KB = 4>
__global__ __launch_bounds__(64 * S) void k_gemm_par(const SolveDesc* __restrict__ descs, GemmArgs g) {
  int z, cx, ry;
  xcd_tile(z, cx, ry);
  const SolveDesc& d = descs[z];
  const int yy = d.yy, N = d.xx, h = yy >> 1, odd = yy & 1;
  const int par = FWD ? (ry & 1) : 0;              // forward: the block's mode parity
  const int i0 = 16 * (FWD ? (ry >> 1) : ry), col0 = cx * 32;
  const int nrow = (FWD && par) ? h : h + odd;     // modes of this parity / top rows with the middle one
  if (i0 >= nrow || col0 >= N) return;
  if (g.gate && *g.gate == 0) return;
  if (g.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + g.slot_shift, g.pred_slot)) return;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, kq = lane >> 4, l16 = lane & 15;
  // this wave's parity half (inverse) and K slice: forward K = row pairs (and the middle row) of the
  // block's parity; inverse K = the modes of the wave's parity
  constexpr int SH = FWD ? S : S / 2;
  const int q = FWD ? par : (w >= SH ? 1 : 0);
  const int ws = FWD ? w : w - q * SH;
  const int K = FWD ? nrow : (q ? h : h + odd);
  const int NG = (K + 3) >> 2, per = (NG + SH - 1) / SH;
  const int g0 = min(NG, ws * per), g1 = min(NG, g0 + per);
  const gptr<const double> AT = G(FWD ? d.V : d.VT);
  gptr<const double> Bm;
  int ldb;
  if (FWD) { Bm = desc_in(d, g.in_id); ldb = d.ld; }
  else { Bm = G(d.F); ldb = d.xx; }
  // rows / columns past the problem are read at the last one (in bounds): they feed only outputs
  // that are not stored
  const int ir = min(i0 + l16, nrow - 1);
  const int am = FWD ? 2 * ir + par : ir;           // A column index: mode (forward) or top row (inverse)
  const int cb[2] = {min(col0 + l16, N - 1), min(col0 + 16 + l16, N - 1)};
  const double sgn = par ? -1.0 : 1.0;
  // the 8 (column half, element) pairs of the tile, E per wave (the epilogue below)
  constexpr int E = 8 / S;
  // inverse: the epilogue's operands -- the solid intervals of the two output rows of each pair and,
  // when accumulating, their old values -- load before the K loop (read in the epilogue, they were
  // one more memory round trip after the reduction); indexes past the problem are clamped, the
  // epilogue stores only inside it
  int elo[E][2], ehi[E][2];
  double eold[E][2];
  if (!FWD) {
    const gptr<const double> O = desc_in(d, g.out_id);
    #pragma unroll
    for (int e = 0; e < E; ++e) {
      const int pr = E * w + e, c = pr >> 2, r = pr & 3;
      const int gc = min(col0 + 16 * c + l16, N - 1);
      const int i = min(i0 + kq + 4 * r, nrow - 1);
      #pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int jr = t == 0 ? i : (i < h ? yy - 1 - i : i);
        const int aj = jr + d.joff;
        elo[e][t] = (g.epi != 0 && d.lo) ? G(d.lo)[aj] : 1;
        ehi[e][t] = (g.epi != 0 && d.lo) ? G(d.hi)[aj] : 0;
        eold[e][t] = g.epi == 2 ? O[(size_t)jr * d.ld + gc] : 0.0;
      }
    }
  }
  d4 acc[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
  for (int gb = g0; gb < g1; gb += KB) {
    double a[KB], b[KB][2], bm[KB][2];
    #pragma unroll
    for (int t = 0; t < KB; ++t) {
      const int kc = min(4 * (gb + t) + kq, K - 1);   // a valid k; past the slice or K it is zeroed below
      // forward: k is row j (mirror yy-1-j); inverse: k is the q-parity mode 2k + q
      const int ka = FWD ? kc : 2 * kc + q;
      a[t] = AT[(size_t)ka * yy + am];
      const gptr<const double> Br = Bm + (size_t)ka * ldb;
      #pragma unroll
      for (int c = 0; c < 2; ++c) b[t][c] = Br[cb[c]];
      if (FWD) {
        const gptr<const double> Bn = Bm + (size_t)(yy - 1 - kc) * ldb;
        #pragma unroll
        for (int c = 0; c < 2; ++c) bm[t][c] = Bn[cb[c]];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int t = 0; t < KB; ++t) {
      const int kk = 4 * (gb + t) + kq;
      const bool ok = gb + t < g1 && kk < K;
      #pragma unroll
      for (int c = 0; c < 2; ++c) {
        // forward: a row and its mirror (the middle row k = h, odd yy, stands alone)
        const double bv = FWD ? (kk < h ? fma(sgn, bm[t][c], b[t][c]) : b[t][c]) : b[t][c];
        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ok ? a[t] : 0.0, bv, acc[c], 0, 0, 0);
      }
    }
  }
  __shared__ double red[S][2][4][64];
  #pragma unroll
  for (int c = 0; c < 2; ++c)
    #pragma unroll
    for (int r = 0; r < 4; ++r) red[w][c][r][lane] = acc[c][r];
  __syncthreads();
  #pragma unroll
  for (int e = 0; e < E; ++e) {
    const int pr = E * w + e, c = pr >> 2, r = pr & 3;
    const int gc = col0 + 16 * c + l16;
    const int i = i0 + kq + 4 * r;
    if (gc >= N || i >= nrow) continue;
    if (FWD) {
      double v = red[0][c][r][lane];
      #pragma unroll
      for (int u = 1; u < S; ++u) v += red[u][c][r][lane];
      gemm_put(d, g, 2 * i + par, gc, v);
    } else {
      double P = red[0][c][r][lane], Q = red[SH][c][r][lane];
      #pragma unroll
      for (int u = 1; u < SH; ++u) { P += red[u][c][r][lane]; Q += red[SH + u][c][r][lane]; }
      if (i < h) {
        gemm_put_pre(d, g, i, gc, P + Q, elo[e][0], ehi[e][0], eold[e][0]);
        gemm_put_pre(d, g, yy - 1 - i, gc, P - Q, elo[e][1], ehi[e][1], eold[e][1]);
      } else {
        gemm_put_pre(d, g, i, gc, P, elo[e][0], ehi[e][0], eold[e][0]);   // the middle row of odd yy
      }
    }
  }
}

struct PcrArgs { int mode; int pred_slot; int slot_shift; const int* gate = nullptr; };

// Cross-lane move through DPP in which lanes without a source lane (and rows outside ROW_MASK) keep
// `old` instead of reading 0
template <int CTRL, int ROW_MASK = 0xf>
DEV double dpp_or(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffll), (int)(unsigned)(b & 0xffffffffll),
                                             CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
template <int L>
DEV double lane_value(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), L);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), L);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// (A, B) <- (A, B) o (the map CTRL brings in): apply the incoming map first, then this lane's;
// lanes with no incoming map compose with the identity (1, 0)
template <int CTRL, int ROW_MASK = 0xf>
DEV void affine_step(double& A, double& B) {
  const double Ai = dpp_or<CTRL, ROW_MASK>(A, 1.0), Bi = dpp_or<CTRL, ROW_MASK>(B, 0.0);
  B = A * Bi + B;
  A = A * Ai;
}
// Inclusive scans of affine maps y -> A y + B over the 64 lanes of a wave on the VALU (DPP), all
// lanes active.  Forward: lane l ends with the maps of lanes 0..l composed in lane order (row
// shifts by 1, 2, 4, 8, then the row broadcasts of lanes 15 and 31).  Reverse: lane l ends with the
// maps of lanes 63..l composed from lane 63 down (row shifts left, then the row totals of the rows
// above, read from lanes 16, 32 and 48).
DEV void wave_affine_fwd(double& A, double& B) {
  affine_step<0x111>(A, B); affine_step<0x112>(A, B); affine_step<0x114>(A, B); affine_step<0x118>(A, B);
  affine_step<0x142, 0xa>(A, B);   // row_bcast:15 into rows 1 and 3
  affine_step<0x143, 0xc>(A, B);   // row_bcast:31 into rows 2 and 3
}
DEV void wave_affine_rev(double& A, double& B) {
  affine_step<0x101>(A, B); affine_step<0x102>(A, B); affine_step<0x104>(A, B); affine_step<0x108>(A, B);
  const double A1 = lane_value<16>(A), B1 = lane_value<16>(B), A2 = lane_value<32>(A), B2 = lane_value<32>(B);
  const double A3 = lane_value<48>(A), B3 = lane_value<48>(B);
  // the rows above row r, composed from row 3 down: row 2 <- R3, row 1 <- R2 o R3, row 0 <- R1 o R2 o R3
  const double e1A = A2 * A3, e1B = A2 * B3 + B2;
  const double e0A = A1 * e1A, e0B = A1 * e1B + B1;
  const int row = (threadIdx.x & 63) >> 4;
  const double EA = row == 0 ? e0A : row == 1 ? e1A : row == 2 ? A3 : 1.0;
  const double EB = row == 0 ? e0B : row == 1 ? e1B : row == 2 ? B3 : 0.0;
  B = A * EB + B;
  A = A * EA;
}

// Block-wide exclusive scan of affine maps y -> A y + B (composition in thread order, or in reverse
// order when REVERSE); returns the value reached from 0 before this thread's chunk.  Waves scan on
// the VALU; the W wave totals are composed through LDS.
template <int W, bool REVERSE>
DEV double block_affine_scan(double A, double B) {
  __shared__ double sa[W], sb[W];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (REVERSE) wave_affine_rev(A, B);
  else wave_affine_fwd(A, B);
  if (lane == (REVERSE ? 0 : 63)) { sa[w] = A; sb[w] = B; }
  __syncthreads();
  // value entering this wave: compose the totals of the waves before it (after it, reversed)
  double vin = 0.0;
  if (!REVERSE) { for (int q = 0; q < w; ++q) vin = sa[q] * vin + sb[q]; }
  else { for (int q = W - 1; q > w; --q) vin = sa[q] * vin + sb[q]; }
  // value after each lane; a lane's entering value is its neighbour's (the wave's edge lane: vin)
  const double after = A * vin + B;
  const int r16 = lane & 15;
  double enter;
  if (!REVERSE) {
    const double inrow = dpp_or<0x111>(after, vin);            // row_shr:1; lane 0 of a row keeps vin
    const double b15 = dpp_or<0x142, 0xa>(after, vin);         // lanes 15, 47 into rows 1, 3
    const double b31 = dpp_or<0x143, 0x4>(after, vin);         // lane 31 into row 2
    enter = (r16 != 0 || lane == 0) ? inrow : ((lane >> 4) == 2 ? b31 : b15);
  } else {
    const double inrow = dpp_or<0x101>(after, vin);            // row_shl:1; lane 15 of a row keeps vin
    const double a16 = lane_value<16>(after), a32 = lane_value<32>(after), a48 = lane_value<48>(after);
    enter = (r16 != 15 || lane == 63) ? inrow : (lane == 15 ? a16 : lane == 31 ? a32 : a48);
  }
  __syncthreads();
  return enter;
}

// Per-mode tridiagonal solve along x with the precomputed LU factors of (a + b lam_k) I + b lx
// (constant super-diagonal b*upper): one block per mode, each thread runs the
// forward and backward recurrences on a contiguous chunk held in registers; chunks are joined
// by block scans of the affine maps.  mode 0: F[k] <- T_k^{-1} F[k];  mode 1: F[k] <- F[k] -
// T_k^{-1} z_k with z_k built from the capacitance correction y (rows grouped by x column).
// Launch shapes (chosen per context by the autotuner): TRI_T threads per mode, up to TRI_C
// points per thread; every shape covers rows up to TRI_MAXX = 4096 points.
constexpr int TRI_MAXX = 4096;
// column groups of capacitance rows per thread in the correction gather: a problem may hold up to
// TRI_GPT x 128 such groups (the narrowest sweep has 128 threads)
constexpr int TRI_GPT = 4, TRI_MAXG = TRI_GPT * 128;
template <int TRI_T>
__global__ __launch_bounds__(TRI_T) void k_tri(const SolveDesc* __restrict__ descs, PcrArgs p) {
  constexpr int TRI_C = TRI_MAXX / TRI_T;
  const SolveDesc& d = descs[blockIdx.y];
  const int k = blockIdx.x;
  if (k >= d.yy) return;
  if (p.gate && *p.gate == 0) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  const int xx = d.xx, yy = d.yy, tid = threadIdx.x;
  const int C = (xx + TRI_T - 1) / TRI_T;
  const int i0 = min(tid * C, xx), n = min(i0 + C, xx) - i0;
  const gptr<double> Frow = G(d.F) + (size_t)k * xx;
  const gptr<const double> Lrow = G(d.luL) + (size_t)k * xx;
  const gptr<const double> Irow = G(d.luI) + (size_t)k * xx;
  const double cu = d.b * G(d.upper)[0];
  __shared__ double zl[TRI_MAXX];
  // the factor rows first: their loads stay in flight while mode 1 gathers its right-hand side
  double r[TRI_C], l[TRI_C], id[TRI_C];
  #pragma unroll
  for (int c = 0; c < TRI_C; ++c) {
    const int i = i0 + c;
    if (c < n) {
      l[c] = Lrow[i];
      id[c] = Irow[i];
      r[c] = (p.mode == 0) ? Frow[i] : 0.0;
    } else { l[c] = 0.0; id[c] = 0.0; r[c] = 0.0; }
  }
  if (p.mode == 1) {
    // z_k = V^T (rows of the capacitance correction scattered by y), dense only on the x columns
    // that hold a correction row.  The rows (in x-column groups) are staged in passes of two per
    // thread: the factors Vsel[k][t] = V[j_t][k] and y[t] of a pass load together into LDS.  The thread of a
    // group (up to TRI_GPT groups per thread) then adds its rows' products in row order, and the
    // group sums are scattered into the zeroed row, which reuses the staging array.
    constexpr int P = 2 * TRI_T;   // rows per pass
    static_assert(2 * P <= TRI_MAXX, "a pass's (V, y) pairs fit the row array");
    const gptr<const int> rg_ptr = G(d.rg_ptr), rg_x = G(d.rg_x);
    const gptr<const double> Vk = G(d.Vsel) + (size_t)k * d.r, yv = G(d.y);
    const int ng = d.rg_n, nr = d.r;
    int t0[TRI_GPT], t1[TRI_GPT], xg[TRI_GPT];
    double sg[TRI_GPT];
    #pragma unroll
    for (int m = 0; m < TRI_GPT; ++m) {
      const int gq = tid + TRI_T * m;
      const bool has = gq < ng;
      t0[m] = has ? rg_ptr[gq] : 0;
      t1[m] = has ? rg_ptr[gq + 1] : 0;
      xg[m] = has ? rg_x[gq] : -1;
      sg[m] = 0.0;
    }
    for (int base = 0; base < nr; base += P) {
      double vr[2], yr[2];
      #pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int t = min(base + tid + TRI_T * u, nr - 1);
        vr[u] = Vk[t];
        yr[u] = yv[t];
      }
      #pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int q = tid + TRI_T * u;
        if (base + q < nr) { zl[2 * q] = vr[u]; zl[2 * q + 1] = yr[u]; }
      }
      __syncthreads();
      #pragma unroll
      for (int m = 0; m < TRI_GPT; ++m) {
        const int a = max(t0[m], base), b = min(t1[m], base + P);
        for (int t = a; t < b; ++t) sg[m] += zl[2 * (t - base)] * zl[2 * (t - base) + 1];
      }
      __syncthreads();
    }
    for (int i = tid; i < xx; i += TRI_T) zl[i] = 0.0;
    __syncthreads();
    #pragma unroll
    for (int m = 0; m < TRI_GPT; ++m)
      if (xg[m] >= 0) zl[xg[m]] = sg[m];
    __syncthreads();
    #pragma unroll
    for (int c = 0; c < TRI_C; ++c)
      if (c < n) r[c] = zl[i0 + c];
  }
  // forward: y_i = r_i - l_i y_{i-1}
  double A = 1.0, B = 0.0;
  #pragma unroll
  for (int c = 0; c < TRI_C; ++c)
    if (c < n) { const double al = -l[c]; A *= al; B = al * B + r[c]; }
  double y = block_affine_scan<TRI_T / 64, false>(A, B);
  #pragma unroll
  for (int c = 0; c < TRI_C; ++c)
    if (c < n) { y = r[c] - l[c] * y; r[c] = y; }
  // backward: x_i = (y_i - cu x_{i+1}) / d_i
  A = 1.0; B = 0.0;
  #pragma unroll
  for (int c = TRI_C - 1; c >= 0; --c)
    if (c < n) {
      const double al = (i0 + c < xx - 1) ? -cu * id[c] : 0.0;
      A *= al; B = al * B + r[c] * id[c];
    }
  double x = block_affine_scan<TRI_T / 64, true>(A, B);
  #pragma unroll
  for (int c = TRI_C - 1; c >= 0; --c)
    if (c < n) {
      const double al = (i0 + c < xx - 1) ? -cu * id[c] : 0.0;
      x = al * x + r[c] * id[c];
      r[c] = x;
    }
  if (p.mode == 0) {
    #pragma unroll
    for (int c = 0; c < TRI_C; ++c)
      if (c < n) Frow[i0 + c] = r[c];
  } else {
    // F <- F - T^{-1} z: the chunk's loads of F all go out before the first store
    double f[TRI_C];
    #pragma unroll
    for (int c = 0; c < TRI_C; ++c) f[c] = (c < n) ? Frow[i0 + c] : 0.0;
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int c = 0; c < TRI_C; ++c)
      if (c < n) Frow[i0 + c] = f[c] - r[c];
  }
}

// first = R^{-1} b at the capacitance points: one wave per point, lanes over the y modes
__global__ __launch_bounds__(256) void k_points(const SolveDesc* __restrict__ descs, PcrArgs p) {
  const SolveDesc& d = descs[blockIdx.y];
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blockIdx.x * 4 >= d.npts) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  if (q >= d.npts) return;
  const int yy = d.yy, xx = d.xx, lane = threadIdx.x & 63;
  const int x = G(d.pt_x)[q];
  const gptr<const double> Vr = G(d.V) + (size_t)G(d.pt_j)[q] * yy;
  const gptr<const double> F = G(d.F);
  // batches of PT_U modes per lane: the batch's loads go out together; products added in mode order
  constexpr int PT_U = 8;
  double s = 0.0;
  for (int k0 = lane; k0 < yy; k0 += 64 * PT_U) {
    double a[PT_U], b[PT_U];
    #pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const int kk = min(k0 + 64 * u, yy - 1);
      a[u] = Vr[kk];
      b[u] = F[(size_t)kk * xx + x];
    }
    #pragma unroll
    for (int u = 0; u < PT_U; ++u)
      if (k0 + 64 * u < yy) s += a[u] * b[u];
  }
  s = wave_sum(s);
  if (lane == 0) G(d.firstK)[q] = s;
}

// y = (S^{-1} C) first: one wave per row of the precomputed r x npts matrix S^{-1} C
__global__ __launch_bounds__(256) void k_capacitance(const SolveDesc* __restrict__ descs, PcrArgs p) {
  const SolveDesc& d = descs[blockIdx.y];
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blockIdx.x * 4 >= d.r) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  if (row >= d.r) return;
  const int lane = threadIdx.x & 63, np = d.npts;
  const gptr<const double> Sr = G(d.SC) + (size_t)row * np;
  const gptr<const double> fk = G(d.firstK);
  // batches of CAP_U points per lane: the batch's loads go out together; products added in point order
  constexpr int CAP_U = 16;
  double s = 0.0;
  for (int q0 = lane; q0 < np; q0 += 64 * CAP_U) {
    double a[CAP_U], b[CAP_U];
    #pragma unroll
    for (int u = 0; u < CAP_U; ++u) {
      const int q = min(q0 + 64 * u, np - 1);
      a[u] = Sr[q];
      b[u] = fk[q];
    }
    __builtin_amdgcn_sched_barrier(0);   // the batch's loads all go out before the first product
    #pragma unroll
    for (int u = 0; u < CAP_U; ++u)
      if (q0 + 64 * u < np) s += a[u] * b[u];
  }
  s = wave_sum(s);
  if (lane == 0) G(d.y)[row] = s;
}

// ------------------------------------------------------------------------------------
// x-FFT rectangle inverse: DCT-IV / DST-IV along x through a complex FFT of half length, tridiagonal
// solves along y per x mode
// ------------------------------------------------------------------------------------
constexpr int XF_MAXM = 2048;   // rows up to 4096 points
constexpr int YS_MAXY = 2048;   // y systems up to 2048 rows
constexpr int XF_MAXR = 64;     // largest radix of a stage
typedef double2 cplx;
DEV cplx cmul(cplx a, cplx b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
DEV cplx cadd(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
DEV cplx csub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }

// One Stockham stage of radix r over the M points in `in`: for j < L, k < m = M / (r L),
// out[j + L (p + r k)] = sum_q w_{rL}^{q j} in[j + L (k + m q)] w_r^{q p}, twiddles and roots in LDS.
// Radix 2 and 4 run one butterfly per thread; other radices one output per thread (each output
// reads its r inputs), which keeps the threads busy when r is large.  In the first stage (L = 1)
// every twiddle is 1.
template <int R>
DEV void fft_butterfly(const cplx* in, cplx* out, int M, int L, const cplx* tw) {
  const int m = M / (R * L), nb = M / R;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const int j = b % L, k = b / L;
    cplx a[R];
    a[0] = in[j + L * k];
    #pragma unroll
    for (int q = 1; q < R; ++q) a[q] = L > 1 ? cmul(in[j + L * (k + m * q)], tw[(q - 1) * L + j]) : in[j + L * (k + m * q)];
    if constexpr (R == 2) {
      out[j + L * (0 + 2 * k)] = cadd(a[0], a[1]);
      out[j + L * (1 + 2 * k)] = csub(a[0], a[1]);
    } else {
      const cplx s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
      const cplx s13 = cadd(a[1], a[3]), d13 = csub(a[1], a[3]);
      const cplx md13 = make_double2(d13.y, -d13.x);   // -i (a1 - a3)
      out[j + L * (0 + 4 * k)] = cadd(s02, s13);
      out[j + L * (1 + 4 * k)] = cadd(d02, md13);
      out[j + L * (2 + 4 * k)] = csub(s02, s13);
      out[j + L * (3 + 4 * k)] = csub(d02, md13);
    }
  }
}
template <int R>
DEV void fft_outputs_r(const cplx* in, cplx* out, int M, int L, const cplx* tw, const cplx* root) {
  const int m = M / (R * L);
  for (int o = threadIdx.x; o < M; o += blockDim.x) {
    const int j = o % L, pk = o / L, pp = pk % R, k = pk / R;
    cplx a[R];
    #pragma unroll
    for (int q = 0; q < R; ++q) a[q] = in[j + L * (k + m * q)];
    if (L > 1) {
      #pragma unroll
      for (int q = 1; q < R; ++q) a[q] = cmul(a[q], tw[(q - 1) * L + j]);
    }
    cplx acc = a[0];
    int e = 0;
    #pragma unroll
    for (int q = 1; q < R; ++q) {
      e += pp;
      if (e >= R) e -= R;
      acc = cadd(acc, cmul(a[q], root[e]));
    }
    out[o] = acc;
  }
}
DEV void fft_outputs(const cplx* in, cplx* out, int M, int L, int r, const cplx* tw, const cplx* root) {
  const int m = M / (r * L);
  for (int o = threadIdx.x; o < M; o += blockDim.x) {
    const int j = o % L, pk = o / L, pp = pk % r, k = pk / r;
    cplx acc = in[j + L * k];
    int e = 0;
    for (int q = 1; q < r; ++q) {
      e += pp;
      if (e >= r) e -= r;
      cplx a = in[j + L * (k + m * q)];
      if (L > 1) a = cmul(a, tw[(q - 1) * L + j]);
      acc = cadd(acc, cmul(a, root[e]));
    }
    out[o] = acc;
  }
}

// Dynamic LDS of k_xdct for a complex length M and t twiddles + roots: two M-point buffers, then
// the stage twiddles and roots
inline size_t xdct_lds(int M, int t) { return (size_t)(2 * M + t) * sizeof(cplx); }
// A bound on the stage tables of length M: the twiddles of the stages number M - 1 (sum of
// (r - 1) L telescopes) and the roots sum(r) <= prod(r) = M.
inline int xdct_tables(int M) { return 2 * M; }
constexpr size_t XF_MAXLDS = 64 * 1024;

// The x transform of each row, one block per row: forward from the input array into F, inverse
// from F into the output array with the solid mask (epilogue as the y transform's), through an
// M-point complex FFT (M = N / 2).
//   xf 1, 2: the orthonormal DCT-IV / DST-IV, its own inverse.  X = DCT-IV(x) from the FFT of
//     z_n = (x_{2n} + i x_{N-1-2n}) exp(-i pi (4n+1)/(4N)): with y_k = sqrt(2/N) exp(-i pi k/N) Z_k,
//     X_{2k} = Re y_k, X_{N-1-2k} = -Im y_k; DST-IV(x)_k = (-1)^k DCT-IV(x reversed)_k.
//   xf 3 (u faces 1..N, mirrored outlet): basis W[i][m] = sin(pi (2m+1)(i+1) / (2N)); forward
//     H = W^{-1} x = (1/N) (-1)^k DCT-III(x reversed)_k, inverse x_i = (1/2) DCT-II((-1)^m H_m)_{N-1-i}
//     (unnormalised DCT-II / DCT-III).  DCT-II of a real row: v_t = x_{2t}, v_{N-1-t} = x_{2t+1},
//     Z = FFT(v_{2n} + i v_{2n+1}), V_k = E_k + e^{-2 pi i k/N} O_k, V_{k+M} = E_k - e^{-2 pi i k/N} O_k
//     with E, O the even / odd parts of Z, and DCT-II_k = 2 Re(V_k e^{-i pi k/(2N)}); DCT-III runs the
//     same steps backwards (an inverse FFT as the conjugate of the FFT of the conjugate).
//     Tables: fpre = e^{-i pi k/(2N)} (k < N), fpost = e^{-2 pi i k/N} (k < M).
// The row, the pre- and post-twiddles and the stage tables are loaded in one batch at the start.
constexpr int XF_U = XF_MAXM / 256;   // points per thread at most
constexpr int XF_MAXS = 16;           // radix stages at most
DEV int vsrc(int t, int N) { return t < (N >> 1) ? 2 * t : 2 * (N - 1 - t) + 1; }   // v_t = x_{vsrc(t)}
template <int U>
__global__ __launch_bounds__(256) void k_xdct(const SolveDesc* __restrict__ descs, GemmArgs g) {
  const SolveDesc& d = descs[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= d.yy) return;
  if (g.gate && *g.gate == 0) return;
  if (g.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + g.slot_shift, g.pred_slot)) return;
  extern __shared__ cplx xs[];
  __shared__ int stp[3 * XF_MAXS];
  const int N = d.xx, M = N >> 1, T = d.ftn, tid = threadIdx.x, xf = d.xf;
  cplx* b0 = xs;
  cplx* b1 = xs + M;
  cplx* tabs = xs + 2 * M;   // twiddles of every stage, then the roots
  const gptr<const double> src = g.forward ? desc_in(d, g.in_id) + (size_t)j * d.ld : G(d.F) + (size_t)j * N;
  const gptr<const double> pre = G(d.fpre), post = G(d.fpost), tab = G(d.ftw);
  // batch: four row values, three complex twiddles per point, the stage tables
  double a0[U], a1[U], a2[U], a3[U], p0r[U], p0i[U], p1r[U], p1i[U], qr[U], qi[U], tr[2 * U], ti[2 * U];
  #pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = min(tid + 256 * u, M - 1);
    if (xf != 3) {
      a0[u] = src[2 * k]; a1[u] = src[N - 1 - 2 * k]; a2[u] = 0.0; a3[u] = 0.0;
      p0r[u] = pre[2 * k]; p0i[u] = pre[2 * k + 1]; p1r[u] = 0.0; p1i[u] = 0.0;
    } else if (g.forward) {
      // X'_k = x_{N-1-k}, X'_{N-k} = x_{k-1}, X'_{k+M} = x_{M-1-k}, X'_{M-k} = x_{M-1+k}
      a0[u] = src[N - 1 - k]; a1[u] = src[max(k - 1, 0)]; a2[u] = src[M - 1 - k]; a3[u] = src[M - 1 + k];
      p0r[u] = pre[2 * k]; p0i[u] = pre[2 * k + 1]; p1r[u] = pre[2 * (k + M)]; p1i[u] = pre[2 * (k + M) + 1];
    } else {
      // v_{2k} and v_{2k+1} of the sign-flipped coefficients
      a0[u] = src[vsrc(2 * k, N)]; a1[u] = src[vsrc(2 * k + 1, N)]; a2[u] = 0.0; a3[u] = 0.0;
      p0r[u] = pre[2 * k]; p0i[u] = pre[2 * k + 1]; p1r[u] = pre[2 * (k + M)]; p1i[u] = pre[2 * (k + M) + 1];
    }
    qr[u] = post[2 * k]; qi[u] = post[2 * k + 1];
  }
  #pragma unroll
  for (int u = 0; u < 2 * U; ++u) {
    const int t = min(tid + 256 * u, T - 1);
    tr[u] = tab[2 * t]; ti[u] = tab[2 * t + 1];
  }
  const int sv = G(d.fst)[min(tid, 3 * d.fns - 1)];
  #pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = tid + 256 * u;
    if (k >= M) continue;
    if (xf == 1) {
      b0[k] = cmul(make_double2(a0[u], a1[u]), make_double2(p0r[u], p0i[u]));
    } else if (xf == 2) {
      b0[k] = cmul(make_double2(a1[u], a0[u]), make_double2(p0r[u], p0i[u]));
    } else if (g.forward) {
      // V_k = (X'_k - i X'_{N-k}) / 2 e^{i pi k/(2N)}, V_{k+M} likewise; Z_k = E + i O with
      // O = (V_k - V_{k+M}) / 2 e^{2 pi i k/N}; the FFT runs on conj(Z)
      const double xnk = k == 0 ? 0.0 : a1[u];
      const cplx vk = cmul(make_double2(0.5 * a0[u], -0.5 * xnk), make_double2(p0r[u], -p0i[u]));
      const cplx vm = cmul(make_double2(0.5 * a2[u], -0.5 * a3[u]), make_double2(p1r[u], -p1i[u]));
      const cplx e = make_double2(0.5 * (vk.x + vm.x), 0.5 * (vk.y + vm.y));
      const cplx o = cmul(make_double2(0.5 * (vk.x - vm.x), 0.5 * (vk.y - vm.y)), make_double2(qr[u], -qi[u]));
      b0[k] = make_double2(e.x - o.y, -(e.y + o.x));
    } else {
      // H~_m = (-1)^m H_m: v_{2k} = H~_{vsrc(2k)}, v_{2k+1} = H~_{vsrc(2k+1)}
      const int s0 = vsrc(2 * k, N), s1 = vsrc(2 * k + 1, N);
      b0[k] = make_double2((s0 & 1) ? -a0[u] : a0[u], (s1 & 1) ? -a1[u] : a1[u]);
    }
  }
  #pragma unroll
  for (int u = 0; u < 2 * U; ++u)
    if (tid + 256 * u < T) tabs[tid + 256 * u] = make_double2(tr[u], ti[u]);
  if (tid < 3 * d.fns) stp[tid] = sv;
  __syncthreads();
  const int* st = stp;
  cplx* in = b0;
  cplx* out = b1;
  int L = 1;
  for (int sg = 0; sg < d.fns; ++sg) {
    const int r = st[3 * sg];
    const cplx* tw = tabs + st[3 * sg + 1];
    const cplx* root = tabs + st[3 * sg + 2];
    switch (r) {
      case 4: fft_butterfly<4>(in, out, M, L, tw); break;
      case 2: fft_butterfly<2>(in, out, M, L, tw); break;
      case 3: fft_outputs_r<3>(in, out, M, L, tw, root); break;
      case 5: fft_outputs_r<5>(in, out, M, L, tw, root); break;
      case 7: fft_outputs_r<7>(in, out, M, L, tw, root); break;
      case 11: fft_outputs_r<11>(in, out, M, L, tw, root); break;
      case 13: fft_outputs_r<13>(in, out, M, L, tw, root); break;
      default: fft_outputs(in, out, M, L, r, tw, root); break;
    }
    __syncthreads();
    cplx* t = in; in = out; out = t;
    L *= r;
  }
  const gptr<double> Fr = G(d.F) + (size_t)j * N;
  const gptr<double> O = desc_out(d, g.out_id) + (size_t)j * d.ld;
  const int aj = j + d.joff;
  const int slo = (!g.forward && d.lo) ? G(d.lo)[aj] : 1, shi = (!g.forward && d.lo) ? G(d.hi)[aj] : 0;
  auto put = [&](int i, double v) {
    if (g.forward) { Fr[i] = v; return; }
    const int ai = i + d.ioff;
    if (g.epi == 0) O[i] = v;
    else if (slo <= ai && ai <= shi) O[i] = 0.0;
    else O[i] = (g.epi == 2) ? (O[i] + v) : v;
  };
  #pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = tid + 256 * u;
    if (k >= M) continue;
    if (xf != 3) {
      const cplx y = cmul(in[k], make_double2(qr[u], qi[u]));
      // DST-IV: (-1)^i at i = 2k is +1, at i = N-1-2k (odd) -1
      put(2 * k, y.x);
      put(N - 1 - 2 * k, xf == 2 ? y.y : -y.y);
    } else if (g.forward) {
      // z = conj(FFT(conj Z)) / M: v_{2k} = Re / M, v_{2k+1} = -Im / M; H_q = 2 (-1)^q v_t at q = vsrc(t)
      const cplx rv = in[k];
      const int q0 = vsrc(2 * k, N), q1 = vsrc(2 * k + 1, N);
      const double s = 2.0 / M;
      put(q0, (q0 & 1) ? -s * rv.x : s * rv.x);
      put(q1, (q1 & 1) ? s * rv.y : -s * rv.y);
    } else {
      // V_k = E + e^{-2 pi i k/N} O, V_{k+M} = E - e^{-2 pi i k/N} O; x_{N-1-k} = Re(V_k e^{-i pi k/(2N)})
      const cplx zk = in[k], zc = in[(M - k) % M];
      const cplx zcc = make_double2(zc.x, -zc.y);
      const cplx e = make_double2(0.5 * (zk.x + zcc.x), 0.5 * (zk.y + zcc.y));
      const cplx dd = make_double2(0.5 * (zk.x - zcc.x), 0.5 * (zk.y - zcc.y));
      const cplx o = make_double2(dd.y, -dd.x);   // (Z_k - conj Z_{M-k}) / (2i)
      const cplx wo = cmul(make_double2(qr[u], qi[u]), o);
      const cplx vk = cadd(e, wo), vm = csub(e, wo);
      put(N - 1 - k, vk.x * p0r[u] - vk.y * p0i[u]);
      put(M - 1 - k, vm.x * p1r[u] - vm.y * p1i[u]);
    }
  }
}

// y tridiagonal solves of the x modes, in place in F (mode 0), or F <- F - T^{-1} c with c the
// x transform of the capacitance correction (mode 1).  A block holds MB = 256 / NSEG consecutive
// modes and splits the rows into NSEG segments of at most C rows; thread (s, q) composes its
// segment's affine maps, the segments' maps are scanned in LDS, and each thread then runs its
// segment with the carried value.  Loads of one row of a block's modes are contiguous.
template <int NSEG, int C>
__global__ __launch_bounds__(256) void k_ysweep(const SolveDesc* __restrict__ descs, PcrArgs p) {
  constexpr int MB = 256 / NSEG;
  const SolveDesc& d = descs[blockIdx.y];
  const int m0 = blockIdx.x * MB;
  if (m0 >= d.xx) return;
  if (p.gate && *p.gate == 0) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  const int xx = d.xx, yy = d.yy, tid = threadIdx.x, q = tid % MB, s = tid / MB;
  const bool live = m0 + q < xx;
  const int m = min(m0 + q, xx - 1);
  const int S = (yy + NSEG - 1) / NSEG;
  const int j0 = min(s * S, yy), n = min(j0 + S, yy) - j0;
  const gptr<double> F = G(d.F);
  const gptr<const double> Lm = G(d.yl), Im = G(d.yi);
  const double e = d.ye;
  // every load at a clamped, in-bounds row, all in one batch; rows past the segment are not used
  double r[C], l[C], id[C];
  #pragma unroll
  for (int c = 0; c < C; ++c) {
    const size_t at = (size_t)min(j0 + c, yy - 1) * xx + m;
    l[c] = Lm[at];
    id[c] = Im[at];
    r[c] = F[at];
  }
  if (p.mode == 1) {
    #pragma unroll
    for (int c = 0; c < C; ++c) r[c] = 0.0;
  }
  __shared__ double sa[NSEG][MB], sb[NSEG][MB], carry[NSEG][MB];
  if (p.mode == 1) {
    // the correction's right-hand side c (k_xcap), non-zero on the band rows only
    const gptr<const double> cb = G(d.xcb);
    const int lo = d.yr_lo, nbnd = d.yr_n;
    #pragma unroll
    for (int c = 0; c < C; ++c) {
      const int b = min(max(j0 + c - lo, 0), nbnd - 1);
      const double v = cb[(size_t)b * xx + m];
      r[c] = (j0 + c >= lo && j0 + c < lo + nbnd) ? v : 0.0;
    }
  }
  // forward: z_j = r_j - l_j z_{j-1}
  double A = 1.0, B = 0.0;
  #pragma unroll
  for (int c = 0; c < C; ++c)
    if (c < n) { A *= -l[c]; B = -l[c] * B + r[c]; }
  sa[s][q] = A; sb[s][q] = B;
  __syncthreads();
  if (s == 0) {
    double z = 0.0;
    for (int t = 0; t < NSEG; ++t) { carry[t][q] = z; z = sa[t][q] * z + sb[t][q]; }
  }
  __syncthreads();
  double z = carry[s][q];
  #pragma unroll
  for (int c = 0; c < C; ++c)
    if (c < n) { z = r[c] - l[c] * z; r[c] = z; }
  __syncthreads();
  // backward: x_j = (z_j - e x_{j+1}) / d_j
  A = 1.0; B = 0.0;
  #pragma unroll
  for (int c = C - 1; c >= 0; --c)
    if (c < n) {
      const double al = (j0 + c < yy - 1) ? -e * id[c] : 0.0;
      A *= al; B = al * B + r[c] * id[c];
    }
  sa[s][q] = A; sb[s][q] = B;
  __syncthreads();
  if (s == 0) {
    double x = 0.0;
    for (int t = NSEG - 1; t >= 0; --t) { carry[t][q] = x; x = sa[t][q] * x + sb[t][q]; }
  }
  __syncthreads();
  double x = carry[s][q];
  #pragma unroll
  for (int c = C - 1; c >= 0; --c)
    if (c < n) {
      const double al = (j0 + c < yy - 1) ? -e * id[c] : 0.0;
      x = al * x + r[c] * id[c];
      r[c] = x;
    }
  if (!live) return;
  if (p.mode == 0) {
    #pragma unroll
    for (int c = 0; c < C; ++c)
      if (c < n) F[(size_t)(j0 + c) * xx + m] = r[c];
  } else {
    double f[C];
    #pragma unroll
    for (int c = 0; c < C; ++c) f[c] = F[(size_t)min(j0 + c, yy - 1) * xx + m];
    __builtin_amdgcn_sched_barrier(0);
    #pragma unroll
    for (int c = 0; c < C; ++c)
      if (c < n) F[(size_t)(j0 + c) * xx + m] = f[c] - r[c];
  }
}

// first = R^{-1} b at the capacitance points in the x-FFT form: sum_m W[x_t][m] F[j_t][m], one
// wave per point, lanes over the x modes
__global__ __launch_bounds__(256) void k_xpoints(const SolveDesc* __restrict__ descs, PcrArgs p) {
  const SolveDesc& d = descs[blockIdx.y];
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blockIdx.x * 4 >= d.npts) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  if (q >= d.npts) return;
  const int xx = d.xx, lane = threadIdx.x & 63;
  const gptr<const double> w = G(d.xwp) + (size_t)G(d.pt_c)[q] * xx;
  const gptr<const double> f = G(d.F) + (size_t)G(d.pt_j)[q] * xx;
  constexpr int PT_U = 8;
  double s = 0.0;
  for (int m0 = lane; m0 < xx; m0 += 64 * PT_U) {
    double a[PT_U], b[PT_U];
    #pragma unroll
    for (int u = 0; u < PT_U; ++u) {
      const int mm = min(m0 + 64 * u, xx - 1);
      a[u] = w[mm];
      b[u] = f[mm];
    }
    #pragma unroll
    for (int u = 0; u < PT_U; ++u)
      if (m0 + 64 * u < xx) s += a[u] * b[u];
  }
  s = wave_sum(s);
  if (lane == 0) G(d.firstK)[q] = s;
}

// y = (S^{-1} C) first and the x transform of the correction E y on the band rows, one block per
// band row j: each entry of the row (a capacitance row of that y row) gets its multiplier as
// k_capacitance forms it (one wave per entry, the same sum order), kept in LDS and stored, and the
// block then writes c[j][m] = sum over the row's entries of Wf[x_e][m] y_e for every x mode m.
constexpr int XC_MAXE = 256;   // capacitance entries of one y row at most
template <int UM>
__global__ __launch_bounds__(256) void k_xcap(const SolveDesc* __restrict__ descs, PcrArgs p) {
  const SolveDesc& d = descs[blockIdx.y];
  const int b = blockIdx.x;
  if (b >= d.yr_n) return;
  if (p.pred_slot >= 0 && !pred_refine(d.sl, d.slot0 + p.slot_shift, p.pred_slot)) return;
  __shared__ double ys[XC_MAXE];
  __shared__ int gs[XC_MAXE];
  const int j = d.yr_lo + b, lane = threadIdx.x & 63, wv = threadIdx.x >> 6, xx = d.xx, np = d.npts;
  const gptr<const int> rp = G(d.yr_ptr), rt = G(d.yr_t), rg = G(d.yr_g);
  const int e0 = rp[j], ne = rp[j + 1] - e0;
  const gptr<const double> fk = G(d.firstK);
  constexpr int CAP_U = 16;
  for (int e = wv; e < ne; e += 4) {
    const int t = rt[e0 + e];
    const gptr<const double> Sr = G(d.SC) + (size_t)t * np;
    double s = 0.0;
    for (int q0 = lane; q0 < np; q0 += 64 * CAP_U) {
      double a[CAP_U], c[CAP_U];
      #pragma unroll
      for (int u = 0; u < CAP_U; ++u) {
        const int q = min(q0 + 64 * u, np - 1);
        a[u] = Sr[q];
        c[u] = fk[q];
      }
      __builtin_amdgcn_sched_barrier(0);
      #pragma unroll
      for (int u = 0; u < CAP_U; ++u)
        if (q0 + 64 * u < np) s += a[u] * c[u];
    }
    s = wave_sum(s);
    if (lane == 0) {
      ys[e] = s;
      gs[e] = rg[e0 + e];
      G(d.y)[t] = s;
    }
  }
  __syncthreads();
  // c[j][m] for the UM modes m = tid + 256 u of this thread: the entries' basis values load in
  // chunks of four entries x UM modes, entries added in row order
  const gptr<const double> wf = G(d.xwf);
  const gptr<double> out = G(d.xcb) + (size_t)b * xx;
  double acc[UM];
  #pragma unroll
  for (int u = 0; u < UM; ++u) acc[u] = 0.0;
  for (int e = 0; e < ne; e += 4) {
    double w[4][UM];
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
      const size_t row = (size_t)gs[min(e + k, ne - 1)] * xx;
      #pragma unroll
      for (int u = 0; u < UM; ++u) w[k][u] = wf[row + min((int)threadIdx.x + 256 * u, xx - 1)];
    }
    #pragma unroll
    for (int k = 0; k < 4; ++k)
      if (e + k < ne) {
        const double yk = ys[e + k];
        #pragma unroll
        for (int u = 0; u < UM; ++u) acc[u] += w[k][u] * yk;
      }
  }
  #pragma unroll
  for (int u = 0; u < UM; ++u)
    if ((int)threadIdx.x + 256 * u < xx) out[threadIdx.x + 256 * u] = acc[u];
}

// Records the pair of a first pressure Schur solve: rhs = C first (first = R^{-1} b at the
// capacitance points) and the multiplier y = S^{-1} rhs, both in boundary-angle order, tagged with
// the step and the projection.  One block per problem; a full ring stops recording.
__global__ __launch_bounds__(256) void k_dump(const SolveDesc* __restrict__ descs, int proj) {
  const SolveDesc& d = descs[blockIdx.y];
  if (!d.dump_y) return;
  __shared__ int slot;
  if (threadIdx.x == 0) slot = *d.dump_count;
  __syncthreads();
  if (slot >= d.dump_cap) return;
  double* c = d.dump_c + (size_t)slot * d.r;
  double* y = d.dump_y + (size_t)slot * d.r;
  for (int m = threadIdx.x; m < d.r; m += blockDim.x) {
    double s = 0.0;
    for (int q = d.Cptr[m]; q < d.Cptr[m + 1]; ++q) s += d.Cval[q] * d.firstK[d.Ccol[q]];
    c[d.apos[m]] = s;
    y[d.apos[m]] = d.y[d.gpos[m]];
  }
  if (threadIdx.x == 0) {
    d.dump_tag[slot] = 4 * (*d.clock) + proj;
    *d.dump_count = slot + 1;
  }
}

// ------------------------------------------------------------------------------------
// Ghost extension fill (unpack: E q + offset), in place, plus max |value| of written nodes
// ------------------------------------------------------------------------------------
struct FillArgs { int x_id; int vel; int inlet; int track; };

DEV void fill_body(const CaseDev& c, const FillArgs& a, int bx, int gx) {
  const gptr<double> X = G(c.buf[a.x_id]);
  int t = bx * blockDim.x + threadIdx.x;
  double mu = 0.0, mv = 0.0;
  auto row = [&](const GhostTab& g, int k, gptr<const double> base) {
    double s = 0.0;
    const gptr<const int> cl = G(g.col) + (size_t)k * g.width;
    const gptr<const double> wt = G(g.w) + (size_t)k * g.width;
    if (g.width == 8) {
      #pragma unroll
      for (int q = 0; q < 8; ++q) s += wt[q] * base[cl[q]];
    } else {
      for (int q = 0; q < g.width; ++q) s += wt[q] * base[cl[q]];
    }
    return s;
  };
  if (a.vel) {
    const GhostTab& gu = c.gt[0];
    const GhostTab& gv = c.gt[1];
    if (t < gu.n) {
      const double s = row(gu, t, X);
      X[G(gu.node)[t]] = s; mu = s;
    } else if ((t -= gu.n) < gv.n) {
      const double s = row(gv, t, X + c.usize);
      X[c.usize + G(gv.node)[t]] = s; mv = s;
    } else if ((t -= gv.n) < c.ny) {
      const double v = a.inlet ? G(c.inlet)[t] : 0.0;
      X[t * (c.nx + 1)] = v; mu = v;
    }
  } else {
    const GhostTab& gp = c.gt[2];
    if (t < gp.n) X[G(gp.node)[t]] = row(gp, t, X);
  }
  if (a.vel && a.track) {
    double v2[2] = {mu, mv};
    u64* const d2[2] = {c.sl + SL_UMAX, c.sl + SL_VMAX};
    block_max_atomic<2>(v2, d2);
  }
}
__global__ __launch_bounds__(256) void k_fill(const CaseDev* __restrict__ cases, FillArgs a) {
  fill_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// The inlet of the coming step from its end time t (main.py:280-286): base (1 + shape (1 - t/T)) while
// t <= T, the base profile afterwards, with shape = amp sin(2 pi y / Ly); products and sums rounded
// one at a time, as the reference evaluates them.  `pending`: the previous step's clock advance has not
// run yet (its tail runs later in this step), so t = (clock + dt) + dt, rounded as the advance and
// then this sum would round it.
__global__ __launch_bounds__(256) void k_inlet(const CaseDev* __restrict__ cases, int pending) {
  const CaseDev& c = cases[blockIdx.y];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= c.ny || !(c.pert_T > 0.0)) return;
  const double dt = G(c.prm)[0];
  double t = *G(c.clock_time) + dt;
  if (pending) t = t + dt;
  const double base = G(c.inlet_base)[j];
  double v = base;
  if (t <= c.pert_T) {
    const double decay = __dsub_rn(1.0, __ddiv_rn(t, c.pert_T));
    v = __dmul_rn(base, __dadd_rn(1.0, __dmul_rn(G(c.inlet_shape)[j], decay)));
  }
  G(c.inlet)[j] = v;
}

// half_pressure = pressure + increment; output = half + factor * increment
struct PressArgs { int p_id, inc_id, half_id, out_id; StepPrm sp; };
DEV void press_body(const CaseDev& c, const PressArgs& a, int bx, int gx) {
  const gptr<const double> P = G(c.buf[a.p_id]);
  const gptr<const double> I = G(c.buf[a.inc_id]);
  const gptr<double> H = G(c.buf[a.half_id]);
  const gptr<double> O = G(c.buf[a.out_id]);
  const double factor = prm_factor(c, a.sp);
  const Chunk ch = xcd_chunk(0, c.psize, bx, gx);
  for (int idx = ch.beg + threadIdx.x; idx < ch.end; idx += ch.step) {
    const double inc = I[idx];
    const double half = P[idx] + inc;
    H[idx] = half;
    O[idx] = half + factor * inc;
  }
}
__global__ __launch_bounds__(256) void k_pressures(const CaseDev* __restrict__ cases, PressArgs a) {
  press_body(cases[blockIdx.y], a, blockIdx.x, gridDim.x);
}

// Independent stages of a step in one launch: the first body's nblk blocks, then the second's,
// then the third's.  Each body keeps its own block count and grid-stride loop, so every element
// is computed exactly as by its own kernel; a launch costs the same few microseconds whatever its
// size, so merging stages that do not read each other's output shortens the step.
// BK_FIN: the previous step's forces into its row and the clock (finalize_forces; one block).
enum BodyKind { BK_NONE = 0, BK_DIV, BK_HELM, BK_PRES, BK_UPD, BK_PRESS, BK_FILL, BK_FIN };
struct Body {
  int kind, nblk;
  union { DivArgs d; HelmArgs h; PresArgs p; UpdArgs u; PressArgs q; FillArgs f; };
};
DEV void finalize_forces(const CaseDev& c);
struct Fused { Body b[3]; int n; };
DEV void run_body(const CaseDev& c, const Body& b, int bx) {
  switch (b.kind) {
    case BK_DIV: div_body(c, b.d, bx, b.nblk); break;
    case BK_HELM: helm_body(c, b.h, bx, b.nblk); break;
    case BK_PRES: pres_body(c, b.p, bx, b.nblk); break;
    case BK_UPD: upd_body(c, b.u, bx, b.nblk); break;
    case BK_PRESS: press_body(c, b.q, bx, b.nblk); break;
    case BK_FILL: fill_body(c, b.f, bx, b.nblk); break;
    case BK_FIN: if (threadIdx.x < 64) finalize_forces(c); break;
    default: break;
  }
}
__global__ __launch_bounds__(256) void k_fused(const CaseDev* __restrict__ cases, Fused f) {
  const CaseDev& c = cases[blockIdx.y];
  int bx = blockIdx.x;
  #pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (q >= f.n) return;
    if (bx < f.b[q].nblk) { run_body(c, f.b[q], bx); return; }
    bx -= f.b[q].nblk;
  }
}
// The combinations the step uses, instantiated with only their own bodies: a kernel's registers
// are those of its largest body, so each keeps the occupancy of its stages' own kernels.
template <int K>
DEV void run_kind(const CaseDev& c, const Body& b, int bx) {
  if constexpr (K == BK_DIV) div_body(c, b.d, bx, b.nblk);
  else if constexpr (K == BK_HELM) helm_body(c, b.h, bx, b.nblk);
  else if constexpr (K == BK_PRES) pres_body(c, b.p, bx, b.nblk);
  else if constexpr (K == BK_UPD) upd_body(c, b.u, bx, b.nblk);
  else if constexpr (K == BK_PRESS) press_body(c, b.q, bx, b.nblk);
  else if constexpr (K == BK_FILL) fill_body(c, b.f, bx, b.nblk);
  else if constexpr (K == BK_FIN) {
    if (threadIdx.x < 64) finalize_forces(c);
  }
}
template <int K0, int K1, int K2>
__global__ __launch_bounds__(256) void k_fused_of(const CaseDev* __restrict__ cases, Fused f) {
  const CaseDev& c = cases[blockIdx.y];
  int bx = blockIdx.x;
  if (bx < f.b[0].nblk) { run_kind<K0>(c, f.b[0], bx); return; }
  bx -= f.b[0].nblk;
  if (K2 == BK_NONE || bx < f.b[1].nblk) { run_kind<K1>(c, f.b[1], bx); return; }
  bx -= f.b[1].nblk;
  run_kind<K2>(c, f.b[2], bx);
}

// Surface forces: 16 lanes per quadrature point, four points per wave.  Each lane takes every 16th
// nonzero of the point's three CSR rows (surface pressure reconstruction, du/dn, dv/dn), and the
// three rows are walked together: their range loads, then their index and value loads, then their
// gathers go out as one batch each (three memory round trips for rows of up to 16 nonzeros, where
// one row after the other took nine).  A lane adds its own nonzeros of each row in row order; the
// lane groups add their products by a butterfly, and each wave writes one partial sum of its four
// points, which force_totals adds in a fixed order (bit-reproducible).
constexpr int FORCE_PPW = 4;      // points per wave
constexpr int FORCE_MAXP = 4096;  // surface points
__host__ __device__ inline int force_blocks(int nsurf) { return (nsurf + FORCE_PPW - 1) / FORCE_PPW; }
struct ForceArgs { int vel_id, p_id; double mu; };
DEV void forces_body(const CaseDev& c, const ForceArgs& a, int unit) {   // one wave, partial sum `unit`
  const gptr<const double> Vel = G(c.buf[a.vel_id]);
  const gptr<const double> Pp = G(c.buf[a.p_id]);
  const int lane = threadIdx.x & 63, sub = lane & 15;
  const int k = unit * FORCE_PPW + (lane >> 4);
  if (unit * FORCE_PPW >= c.nsurf) return;
  if (unit == 0 && c.nprobe > 0) {
    // the probes, one 16-lane group each like a surface point, stored after the force partials
    const int g = lane >> 4;
    double pr = 0.0;
    if (g < c.nprobe) {
      const gptr<const int> qp = G(c.qp), qc = G(c.qc);
      const gptr<const double> qv = G(c.qv);
      for (int t = qp[g] + sub; t < qp[g + 1]; t += 16) pr += qv[t] * Pp[qc[t]];
    }
    for (int o = 8; o > 0; o >>= 1) pr += __shfl_xor(pr, o, 64);
    if (g < c.nprobe && sub == 0) G(c.fpart)[4 * force_blocks(c.nsurf) + g] = pr;
  }
  double ps = 0.0, du = 0.0, dv = 0.0;
  double nxk = 0.0, nyk = 0.0, nwx = 0.0, nwy = 0.0;   // loaded with the row ranges
  if (k < c.nsurf) {
    const gptr<const double> nrm = G(c.nrm), nw = G(c.nw);
    nxk = nrm[2 * k]; nyk = nrm[2 * k + 1]; nwx = nw[2 * k]; nwy = nw[2 * k + 1];
    const gptr<const int> pp = G(c.pp), pc = G(c.pc), up = G(c.up), uc = G(c.uc), vp = G(c.vp), vc = G(c.vc);
    const gptr<const double> pv = G(c.pv), uv = G(c.uv), vv = G(c.vv);
    const int p0 = pp[k], p1 = pp[k + 1], u0 = up[k], u1 = up[k + 1], v0 = vp[k], v1 = vp[k + 1];
    const int len = max(p1 - p0, max(u1 - u0, v1 - v0));
    for (int j = sub; j < len; j += 16) {
      // a nonzero the row does not have reads entry 0 (always present) and is not added: the loads
      // stay unconditional, so none waits for the one before it
      const bool hp = p0 + j < p1, hu = u0 + j < u1, hv = v0 + j < v1;
      const int tp = hp ? p0 + j : 0, tu = hu ? u0 + j : 0, tv = hv ? v0 + j : 0;
      const int ip = pc[tp], iu = uc[tu], iv = vc[tv];
      const double wp = pv[tp], wu = uv[tu], wv = vv[tv];
      const double xp = Pp[ip], xu = Vel[iu], xv = Vel[c.usize + iv];
      if (hp) ps += wp * xp;
      if (hu) du += wu * xu;
      if (hv) dv += wv * xv;
    }
  }
  for (int o = 8; o > 0; o >>= 1) {
    ps += __shfl_xor(ps, o, 64);
    du += __shfl_xor(du, o, 64);
    dv += __shfl_xor(dv, o, 64);
  }
  double s[4] = {0.0, 0.0, 0.0, 0.0};  // pressure x, y; traction x, y
  if (k < c.nsurf && sub == 0) {
    const double dot = du * nxk + dv * nyk;
    s[0] = ps * nwx;
    s[1] = ps * nwy;
    const double mu = a.mu > 0.0 ? a.mu : c.nu * c.rho;
    s[2] = mu * (du + nxk * dot);
    s[3] = mu * (dv + nyk * dot);
  }
  #pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double v = wave_sum(s[q]);
    if (lane == 0) G(c.fpart)[unit * 4 + q] = v;
  }
}

// (Fx, Fy, Fx_p, Fy_p, Fx_v, Fy_v) from k_forces' per-wave partial sums.  Called by all 64 lanes of a
// one-wave block: lane l adds partials l, l + 64, ... in order, then the wave adds the 64 lane sums
// by its DPP reduction (a fixed order), so the totals are bit-reproducible.
DEV void force_totals(const CaseDev& c, double (&vals)[6]) {
  const int lane = threadIdx.x & 63;
  const int nb = force_blocks(c.nsurf);
  const gptr<const double> fp = G(c.fpart);
  double p[4] = {0.0, 0.0, 0.0, 0.0};
  // four partials per lane in flight at a time, added in the same order as one at a time
  for (int b0 = lane; b0 < nb; b0 += 256) {
    double x[4][4];
    #pragma unroll
    for (int i = 0; i < 4; ++i)
      #pragma unroll
      for (int q = 0; q < 4; ++q) x[i][q] = b0 + 64 * i < nb ? fp[(b0 + 64 * i) * 4 + q] : 0.0;
    #pragma unroll
    for (int i = 0; i < 4; ++i)
      if (b0 + 64 * i < nb)
        #pragma unroll
        for (int q = 0; q < 4; ++q) p[q] += x[i][q];
  }
  double t[4];
  #pragma unroll
  for (int q = 0; q < 4; ++q) t[q] = wave_sum(p[q]);
  const double scale = M_PI * c.diameter / c.nsurf;
  const double fxp = -t[0], fyp = -t[1], fxv = t[2] * scale, fyv = t[3] * scale;
  vals[0] = fxp + fxv; vals[1] = fyp + fyv; vals[2] = fxp; vals[3] = fyp; vals[4] = fxv; vals[5] = fyv;
}
__global__ __launch_bounds__(64) void k_force_total(const CaseDev* __restrict__ cases) {
  const CaseDev& c = cases[blockIdx.y];
  double vals[6];
  force_totals(c, vals);
  if (threadIdx.x != 0) return;
  double* f = reinterpret_cast<double*>(c.sl + SL_FORCE);   // stripe 0 of slots SL_FORCE .. SL_FORCE + 5
  for (int q = 0; q < 6; ++q) f[q] = vals[q];
}

// Gather the step's scalars into the diagnostics ring and clear the slots (one wave per case)
// Record of one monitored pressure solve (pressure_schur.py:58-70) from its slot values: the direct
// backend's refinements follow from its first two residuals, GMRES reports its own counts.
struct SolveSlots { double val, r1, r2, fin, rnd, asc, ssr, ssv, it, ref, st; };
__host__ __device__ inline void solve_record(const SolveSlots& s, bool gmres, double* rec) {
  const double t1 = fmax(1e-10, 32.0 * kEps * fmax(1.0, s.val));
  const double target = fmax(fmax(1e-10, 64.0 * kEps * fmax(1.0, s.val)), s.rnd);
  rec[MCF_S_ITERATIONS] = gmres ? s.it : 0.0;
  rec[MCF_S_RES_LINF] = s.fin;
  rec[MCF_S_REL_L2] = sqrt(s.ssr) / fmax(sqrt(s.ssv), kTiny);
  rec[MCF_S_REFINE] = gmres ? s.ref : (double)((s.r1 > t1 ? 1 : 0) + (s.r1 > t1 && s.r2 > t1 ? 1 : 0));
  rec[MCF_S_TARGET] = target;
  rec[MCF_S_ROUND] = s.rnd;
  rec[MCF_S_BACKWARD] = s.fin / fmax(s.asc, kTiny);
  rec[MCF_S_STATUS] = gmres ? s.st : 0.0;
  rec[MCF_S_SECONDS] = 0.0;
  rec[MCF_S_SPARE] = 0.0;
}
// the gate of pressure_schur.py:69-70
__host__ __device__ inline bool solve_ok(const double* rec) {
  return rec[MCF_S_STATUS] == 0.0 && rec[MCF_S_RES_LINF] <= rec[MCF_S_TARGET];
}

// lean: which solves ran without (some of) their refinement launches (LEAN_VEL: the velocity solve
// without any, LEAN_PRES: the pressure solves without any, ONCE_PRES: the pressure solves without
// the second)
enum { LEAN_VEL = 1, LEAN_PRES = 2, LEAN_ALL = 3, ONCE_PRES = 4 };
// The lean masks a speculating run switches between (spec_mask)
constexpr int SPEC_MASKS[4] = {LEAN_ALL, 0, LEAN_VEL | ONCE_PRES, LEAN_VEL};
struct FinArgs { int nproj; int lean; int force_miss; int gmres; int hval; };   // hval: the step's SL_RHS pair
static_assert(MCF_DIAG_LEN == 64, "one diagnostics entry per lane of finalize_slots");
// Slots gathered by finalize_slots: FIN_MAX maxima over the stripes, then FIN_SUM sums
constexpr int FIN_MAX = 41, FIN_SUM = 6, FIN_NQ = FIN_MAX + FIN_SUM;
__constant__ const int kFinSlots[FIN_NQ] = {
    SL_UMAX, SL_VMAX, SL_HU + SV_VAL, SL_HV + SV_VAL, SL_HU + SV_FINAL, SL_HV + SV_FINAL,
    SL_HU + SV_RES1, SL_HU + SV_RES2, SL_HV + SV_RES1, SL_HV + SV_RES2, SL_COUPLED,
    SL_DIV0, SL_DIV1, SL_DIV2,
    SL_P0 + SV_VAL, SL_P0 + SV_RES1, SL_P0 + SV_RES2, SL_P0 + SV_FINAL, SL_P0 + SV_ROUND, SL_P0 + SV_ASCALE,
    SL_P1 + SV_VAL, SL_P1 + SV_RES1, SL_P1 + SV_RES2, SL_P1 + SV_FINAL, SL_P1 + SV_ROUND, SL_P1 + SV_ASCALE,
    SL_P2 + SV_VAL, SL_P2 + SV_RES1, SL_P2 + SV_RES2, SL_P2 + SV_FINAL, SL_P2 + SV_ROUND, SL_P2 + SV_ASCALE,
    SL_GM_IT, SL_GM_IT + 1, SL_GM_IT + 2, SL_GM_REF, SL_GM_REF + 1, SL_GM_REF + 2, SL_GM_ST, SL_GM_ST + 1, SL_GM_ST + 2,
    SL_P0 + SV_SSQR, SL_P0 + SV_SSQV, SL_P1 + SV_SSQR, SL_P1 + SV_SSQV, SL_P2 + SV_SSQR, SL_P2 + SV_SSQV};

// Gather the step's scalars into the diagnostics ring and clear the slots (one wave per case), in
// two parts.  finalize_slots needs only the slots, which are final once the step's last stage has
// run, so it runs as one more block beside the force waves (k_forces, k_adv_tail): every lane
// stores its stripe of every slot in LDS; lane q then reduces slot q over the stripes (maxima of
// |.|, which do not depend on the order, or sums in stripe order); lane 0 evaluates the gates and
// the per-solve records; the row is stored once, one entry per lane, with the force entries zero,
// and the slots are cleared.  A later launch then adds the force partial sums into the forces, C_D
// and C_L, the probes and the finite-force gate, and advances the clock (finalize_forces).  (As one
// kernel after the forces, the whole gather took 7.3 us of a 0.191 ms step.)
DEV void finalize_slots(const CaseDev& c, const FinArgs& a) {   // every thread of the block calls; wave 0 works
  const gptr<const u64> s = G(c.sl);
  const int lane = threadIdx.x;
  const bool w0 = lane < 64;
  // step parameters and clock, loaded together with the slots
  const double dt = G(c.prm)[0];
  const int row = *G(c.clock_step);
  const double time = *G(c.clock_time) + dt;
  __shared__ double raw[FIN_NQ][65];   // padded: lane q reads row q
  __shared__ double v[FIN_NQ];
  __shared__ double d[MCF_DIAG_LEN];
  if (w0) {
    #pragma unroll
    for (int q = 0; q < FIN_NQ; ++q) {
      const int k = q == 2 ? a.hval : q == 3 ? a.hval + 1 : kFinSlots[q];   // the step's right-hand side maxima
      raw[q][lane] = __longlong_as_double((long long)s[slot_word(k, lane)]);
    }
    d[lane] = 0.0;
  }
  __syncthreads();
  if (lane < FIN_NQ) {
    double r = 0.0;
    if (lane < FIN_MAX) {
      // max |.| over the stripes on the bit patterns (NaN above +inf), four independent chains
      u64 m[4] = {0ull, 0ull, 0ull, 0ull};
      #pragma unroll 4
      for (int t = 0; t < 64; t += 4)
        #pragma unroll
        for (int e = 0; e < 4; ++e) {
          const u64 b = (u64)__double_as_longlong(fabs(raw[lane][t + e]));
          m[e] = b > m[e] ? b : m[e];
        }
      const u64 m01 = m[0] > m[1] ? m[0] : m[1], m23 = m[2] > m[3] ? m[2] : m[3];
      r = __longlong_as_double((long long)(m01 > m23 ? m01 : m23));
    } else {
      #pragma unroll 8
      for (int t = 0; t < 64; ++t) r += raw[lane][t];
    }
    v[lane] = r;
  }
  __syncthreads();
  if (lane == 0) {
    const double umax = v[0], vmax = v[1], hu_val = v[2], hv_val = v[3], hu_fin = v[4], hv_fin = v[5];
    const double hu_r1 = v[6], hu_r2 = v[7], hv_r1 = v[8], hv_r2 = v[9], coupled = v[10];
    const double div[3] = {v[11], v[12], v[13]};
    const double maxdiv = div[0];
    const double cfl = dt * (umax / c.dx + vmax / c.dy);
    d[MCF_D_TIME] = time; d[MCF_D_DT] = dt;
    d[MCF_D_MAXDIV] = maxdiv;
    d[MCF_D_CFL] = cfl;
    const double rhs_max = fmax(hu_val, hv_val);
    const double mres = nanmax(hu_fin, hv_fin);
    d[MCF_D_MOMENTUM_RES] = mres;
    d[MCF_D_RHS_MAX] = rhs_max;
    d[MCF_D_COUPLED_RES] = coupled;
    int status = 0;
    if (!(mres <= 1e-10 * fmax(1.0, rhs_max))) status |= 1;
    int refine = 0;
    bool second = false;   // a pressure solve needed the second refinement that ONCE_PRES left out
    double worst = -1.0, relmax = 0.0;
    // one record per pressure solve in call order: the startup projection (block 2) comes first
    for (int q = 0; q < a.nproj; ++q) {
      const int b = a.nproj == 3 ? (q == 0 ? 2 : q - 1) : q;
      const double pfin = v[17 + 6 * b], prnd = v[18 + 6 * b];
      // with one refinement pass the final residual is the one the second pass would have tested
      const bool once = a.lean & ONCE_PRES;
      const double pval = v[14 + 6 * b], pr1 = v[15 + 6 * b], pr2 = once ? pfin : v[16 + 6 * b];
      const double tgt = woodbury_target(pval);
      second = second || (once && (pr1 > tgt || pr1 != pr1) && (pfin > tgt || pfin != pfin));
      const SolveSlots ss{pval, pr1, pr2, pfin, prnd, v[19 + 6 * b],
                          v[FIN_MAX + 2 * b], v[FIN_MAX + 2 * b + 1], v[32 + b], v[35 + b], v[38 + b]};
      double rec[MCF_SOLVE_LEN];
      solve_record(ss, a.gmres != 0, rec);
      for (int k = 0; k < MCF_SOLVE_LEN; ++k) d[MCF_D_SOLVE0 + MCF_SOLVE_LEN * q + k] = rec[k];
      refine += (int)rec[MCF_S_REFINE];
      if (!solve_ok(rec)) status |= 2;
      relmax = fmax(relmax, rec[MCF_S_REL_L2]);
      if (b == 0) d[MCF_D_P_RES1] = pfin;
      if (pfin > worst || pfin != pfin) {
        worst = pfin;
        d[MCF_D_P_RES0] = pfin;
        d[MCF_D_P_TARGET0] = rec[MCF_S_TARGET];
        d[MCF_D_P_ROUND0] = prnd;
        d[MCF_D_P_BACKWARD0] = rec[MCF_S_BACKWARD];
      }
    }
    d[MCF_D_P_REL0] = relmax;
    const double tu = woodbury_target(hu_val), tv = woodbury_target(hv_val);
    const int hrefine = (hu_r1 > tu) + (hu_r1 > tu && hu_r2 > tu) + (hv_r1 > tv) + (hv_r1 > tv && hv_r2 > tv);
    d[MCF_D_P_REFINE] = refine;
    d[MCF_D_H_REFINE] = hrefine;
    d[MCF_D_P_SOLVES] = a.nproj;
    double divg = 0.0;
    for (int q = 0; q < 3; ++q) divg = nanmax(divg, div[q]);
    if (!(divg <= 1e-10)) status |= 4;
    if (!(maxdiv < 1e-10) || !(cfl < 0.5)) status |= 8;
    // a lean step (refinement launches left out) whose first residual called for a refinement of a
    // solve that ran lean: the host replays the chunk with the refinements in place
    if (((a.lean & LEAN_PRES) && refine > 0) || ((a.lean & LEAN_VEL) && hrefine > 0) || second ||
        (a.lean && a.force_miss))
      status |= 32;
    d[MCF_D_STATUS] = status;   // the finite-force gate (16) is added by finalize_forces
  }
  __syncthreads();
  if (!w0) return;
  G(c.diag)[(size_t)row * MCF_DIAG_LEN + lane] = d[lane];
  const gptr<u64> wsl = G(c.sl);
  for (int st = 0; st < NSTRIPE; ++st) wsl[slot_word(lane, st)] = 0ull;   // the used words of every stripe
  if (lane < 2)
    for (int st = 0; st < NSTRIPE; ++st) wsl[slot_word(a.hval + lane, st)] = 0ull;   // and the step's pair
}

// Surface forces (forces_body); with `slots`, the launch's last block instead gathers the step's
// slot values into the diagnostics row (finalize_slots) while the others compute the forces
__global__ __launch_bounds__(64) void k_forces(const CaseDev* __restrict__ cases, ForceArgs a, FinArgs fin, int slots) {
  const CaseDev& c = cases[blockIdx.y];
  if (slots && blockIdx.x == gridDim.x - 1) {
    finalize_slots(c, fin);
    return;
  }
  forces_body(c, a, blockIdx.x);
}

// The step's forces into its diagnostics row (written by finalize_slots in an earlier launch), the
// finite-force gate, and the clock (one wave per case): a kernel of its own after k_forces
// (k_finalize), or one more body of a later fused launch (BK_FIN)
DEV void finalize_forces(const CaseDev& c) {
  const int lane = threadIdx.x & 63;
  const double dt = G(c.prm)[0];
  const int row = *G(c.clock_step);
  const double time = *G(c.clock_time) + dt;
  double f[6];
  force_totals(c, f);   // the step's forces, from the force waves' partial sums
  if (lane != 0) return;
  const gptr<double> dr = G(c.diag) + (size_t)row * MCF_DIAG_LEN;
  const double denom = 0.5 * c.rho * c.uref * c.uref * c.diameter;
  const double cd = f[0] / denom, cl = f[1] / denom;
  dr[MCF_D_FX] = f[0]; dr[MCF_D_FY] = f[1];
  dr[MCF_D_CD] = cd; dr[MCF_D_CL] = cl;
  dr[MCF_D_CD_P] = f[2] / denom; dr[MCF_D_CL_P] = f[3] / denom;
  dr[MCF_D_CD_V] = f[4] / denom; dr[MCF_D_CL_V] = f[5] / denom;
  if (c.nprobe > 0) {
    const gptr<const double> pv = G(c.fpart) + 4 * force_blocks(c.nsurf);
    dr[MCF_D_PROBE0] = pv[0];
    if (c.nprobe > 1) dr[MCF_D_PROBE1] = pv[1];
  }
  if (!isfinite(cd) || !isfinite(cl)) dr[MCF_D_STATUS] = (double)((int)dr[MCF_D_STATUS] | 16);
  *G(c.clock_step) = row + 1;
  *G(c.clock_time) = time;
}
__global__ __launch_bounds__(64) void k_finalize(const CaseDev* __restrict__ cases) {
  finalize_forces(cases[blockIdx.y]);
}

// The first stage of a step with the previous step's slot gather and force waves in its first `tb`
// blocks (one case): block 0 gathers the slots, blocks 1 .. tb - 1 hold four force waves each.  The
// previous step's forces are completed by a later launch of the step (BK_FIN), after the kernel
// boundary has made the partial sums and the row visible: an in-kernel hand-over (count the waves
// in, the last completes) needs a device-scope release per wave, whose L2 write-back cost the
// advection blocks more than the launch it saved.  tb is a multiple of 8, so the advection blocks
// keep their XCD placement.
constexpr int TAIL_WAVES = 4;
__host__ __device__ inline int adv_tail_blocks(int nsurf) {
  return (1 + (force_blocks(nsurf) + TAIL_WAVES - 1) / TAIL_WAVES + 7) / 8 * 8;
}
__global__ __launch_bounds__(256) void k_adv_tail(const CaseDev* __restrict__ cases, AdvArgs a, ForceArgs fa,
                                                  FinArgs fin, int tb) {
  const CaseDev& c = cases[blockIdx.y];
  if ((int)blockIdx.x < tb) {
    if (blockIdx.x == 0) finalize_slots(c, fin);   // the whole block: it synchronizes
    else forces_body(c, fa, TAIL_WAVES * ((int)blockIdx.x - 1) + (int)(threadIdx.x >> 6));
    return;
  }
  if (a.q_id == a.adv_id) adv_body<true>(c, a, blockIdx.x - tb, gridDim.x - tb);
  else adv_body<false>(c, a, blockIdx.x - tb, gridDim.x - tb);
}

// Step parameters of every case of the launch; dt <= 0 keeps each case's own dt
__global__ void k_set_prm(const CaseDev* __restrict__ cases, double dt, double alpha, double factor) {
  const CaseDev& c = cases[blockIdx.y];
  if (threadIdx.x == 0) {
    if (dt > 0.0) c.prm[0] = dt;
    c.prm[1] = alpha;
    c.prm[2] = factor;
  }
}

__global__ void k_set_clock(const CaseDev* __restrict__ cases, double t0) {
  const CaseDev& c = cases[blockIdx.y];
  if (threadIdx.x == 0) { *c.clock_step = 0; *c.clock_time = t0; }
}

// Host view of the slots, one lane per slot (one wave): max |.| over the stripes, the sum for the
// residual sum-of-squares slots, stripe 0 for the forces; the 64 values go to `out`.
__global__ __launch_bounds__(64) void k_slot_values(const CaseDev* __restrict__ cases, double* out) {
  const CaseDev& c = cases[0];
  const int k = threadIdx.x;
  const gptr<const u64> sl = G(c.sl);
  const bool force = k >= SL_FORCE && k < SL_FORCE + 6;
  const bool sum = k >= SL_HU && k < SL_FORCE && ((k - SL_HU) % 8 == SV_SSQR || (k - SL_HU) % 8 == SV_SSQV);
  double v = 0.0;
  if (force) v = __longlong_as_double((long long)sl[slot_word(k, 0)]);
  else
    for (int t = 0; t < NSTRIPE; ++t) {
      const double x = __longlong_as_double((long long)sl[slot_word(k, t)]);
      if (sum) v += x;
      else v = (fabs(x) > v || x != x) ? fabs(x) : v;
    }
  G(out)[k] = v;
}

__global__ void k_clear_slots(const CaseDev* __restrict__ cases) {
  const CaseDev& c = cases[blockIdx.y];
  for (int w = threadIdx.x; w < NSTRIPE * NSLOT; w += blockDim.x) c.sl[slot_word(w % NSLOT, w / NSLOT)] = 0ull;
  for (int w = threadIdx.x; w < NSTRIPE * 4; w += blockDim.x) c.sl[slot_word(SL_RHS + (w & 3), w >> 2)] = 0ull;
}

// Checkpoint of what a steady step reads from earlier steps: the state S[parity], the advection
// history H[1 - parity] and the half pressure (save = 1: copy out, 0: copy back)
__global__ __launch_bounds__(256) void k_ckpt(const CaseDev* __restrict__ cases, int parity, int save) {
  const CaseDev& c = cases[blockIdx.y];
  double* S = c.buf[B_S0 + parity];
  double* H = c.buf[B_H0 + 1 - parity];
  double* P = c.buf[B_PH];
  double* cs = c.buf[B_CKS];
  double* ch = c.buf[B_CKH];
  double* cp = c.buf[B_CKP];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < c.velsize; idx += gridDim.x * blockDim.x) {
    if (save) { cs[idx] = S[idx]; ch[idx] = H[idx]; }
    else { S[idx] = cs[idx]; H[idx] = ch[idx]; }
    if (idx < c.psize) {
      if (save) cp[idx] = P[idx];
      else P[idx] = cp[idx];
    }
  }
}

__global__ __launch_bounds__(256) void k_copy(const CaseDev* __restrict__ cases, int src, int dst, int n) {
  const CaseDev& c = cases[blockIdx.y];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) c.buf[dst][idx] = c.buf[src][idx];
}

__global__ __launch_bounds__(256) void k_divstats(const double* u, const double* v, const unsigned char* mask,
                                                  int nx, int ny, double dx, double dy, double* out) {
  // out: [max, sum, count, core sum, core count] accumulated with atomics
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0, s = 0.0, cnt = 0.0, cs = 0.0, cc = 0.0;
  if (idx < nx * ny) {
    const int j = idx / nx, i = idx - j * nx;
    if (!mask[idx]) {
      const double dv = (u[j * (nx + 1) + i + 1] - u[j * (nx + 1) + i]) / dx + (v[(j + 1) * nx + i] - v[j * nx + i]) / dy;
      m = fabs(dv); s = dv * dv; cnt = 1.0;
      bool core = true;
      if (i > 0 && mask[idx - 1]) core = false;
      if (i < nx - 1 && mask[idx + 1]) core = false;
      if (j > 0 && mask[idx - nx]) core = false;
      if (j < ny - 1 && mask[idx + nx]) core = false;
      if (core) { cs = s; cc = 1.0; }
    }
  }
  double v1[1] = {m};
  u64* const d1[1] = {reinterpret_cast<u64*>(out)};
  block_max_atomic<1>(v1, d1);
  __syncthreads(); block_sum_atomic(s, out + 1);
  __syncthreads(); block_sum_atomic(cnt, out + 2);
  __syncthreads(); block_sum_atomic(cs, out + 3);
  __syncthreads(); block_sum_atomic(cc, out + 4);
}

}  // namespace

// ======================================================================================
// Host side
// ======================================================================================
// Launch shapes of the transform (tile TM x TN, K chunk) and of the tridiagonal sweep (threads per
// mode), per problem class; picked by the autotuner or pinned by the caller.
namespace {

// ------------------------------------------------------------------------------------
// Iterative pressure backend (pressure_schur.py:16-72 with iterative=True): restarted GMRES on the
// pressure operator A, left-preconditioned by the rectangle inverse M (zero-filled rectangle
// right-hand side, separable solve, restricted to the active cells).  The control flow is SciPy's
// gmres, which the reference calls (scipy/sparse/linalg/_isolve/iterative.py, SciPy 1.15): cycles
// of at most `restart` Arnoldi steps; an inner stop when the preconditioned residual |S[col+1]|
// falls to ptol, with ptol adapted after every cycle; Givens rotations as LAPACK's dlartg; the
// outer stop on the true residual ||b - A x||_2 <= max(atol, rtol ||b||_2).  Orthogonalisation is
// classical Gram-Schmidt applied twice (CGS2): each pass is one kernel over all basis vectors
// instead of one dependent launch per vector (modified Gram-Schmidt), with the same stability in
// practice.  Vectors live on the full pressure rectangle and are zero off the active cells.  Sums
// are deterministic: every block writes its partial sums and one workgroup adds them in a fixed
// order.  Basis vectors are stored unnormalised with a scale each (vector k = scale[k] * raw k),
// which saves a normalisation pass per step.
// ------------------------------------------------------------------------------------
constexpr int KR_MAXR = 64;                            // longest restart cycle
constexpr int KR_T = 256, KR_E = 2, KR_CHUNK = KR_T * KR_E;   // vector kernels: elements per block
constexpr int KR_CT = 1024;                            // threads of the control kernel

struct KryCtl {
  // read back by the host after each group of launches
  int active;      // 1 while the current cycle's Arnoldi steps run (gates their launches)
  int done;        // the GMRES call has ended
  int info;        // 0 converged, maxiter otherwise (SciPy's info)
  int inner;       // Arnoldi steps of this call (the reference's callback count)
  int cycle, col, breakdown, ncols;
  // parameters
  int restart, maxiter;
  double rtol, atol_in;
  // state
  double bnrm2, atol, ptol, pmax, presid, rnorm, h0;
  double h[KR_MAXR][KR_MAXR + 1];   // h[col][k]: column col of the Hessenberg matrix
  double giv[KR_MAXR][2];
  double S[KR_MAXR + 1];
  double hk[KR_MAXR + 1];           // first-pass projections of the current column
  double coef[KR_MAXR + 1];         // coefficients of the pending update, on the raw vectors
  double scale[KR_MAXR + 1];
};
constexpr int KR_STAT_INTS = 8;

struct KryArgs {
  KryCtl* ctl;
  double* V;       // restart + 1 raw basis vectors, ldv apart (vector 0 is buffer B_KV0)
  size_t ldv;
  double* part;    // block partial sums, part[q * nblk + block]
  int nblk;
};

// Sum of v over the block into red[q][wave]; kr_flush writes the block's nq sums to the partials.
DEV void kr_wave_sum(double v, int q, double (*red)[KR_T / 64]) {
  const double s = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = s;
}
DEV void kr_flush(double (*red)[KR_T / 64], int nq, const KryArgs& a) {
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += KR_T) {
    double s = 0.0;
    #pragma unroll
    for (int w = 0; w < KR_T / 64; ++w) s += red[q][w];
    a.part[(size_t)q * a.nblk + blockIdx.x] = s;
  }
}
DEV bool kr_active_cell(const CaseDev& c, int idx) {
  const int j = idx / c.nx, i = idx - j * c.nx;
  return p_act(c, j, i);
}

// r = b on the active cells (0 elsewhere), x = 0; partial sums of b^2
__global__ __launch_bounds__(KR_T) void k_kry_begin(const CaseDev* __restrict__ cases, KryArgs a, int b_id, int x_id, int r_id) {
  const CaseDev& c = cases[0];
  __shared__ double red[1][KR_T / 64];
  const double* B = c.buf[b_id];
  double* X = c.buf[x_id];
  double* R = c.buf[r_id];
  double ss = 0.0;
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    if (idx < c.psize) {
      const double b = kr_active_cell(c, idx) ? B[idx] : 0.0;
      R[idx] = b;
      X[idx] = 0.0;
      ss += b * b;
    }
  }
  kr_wave_sum(ss, 0, red);
  kr_flush(red, 1, a);
}

// out = A (scale[col] * raw V[col]) on the active cells
__global__ __launch_bounds__(KR_T) void k_kry_matvec(const CaseDev* __restrict__ cases, KryArgs a, int out_id) {
  const KryCtl& k = *a.ctl;
  if (!k.active) return;
  const CaseDev& c = cases[0];
  const int col = k.col;
  const double sc = k.scale[col];
  const double* X = a.V + (size_t)col * a.ldv;
  double* O = c.buf[out_id];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    if (idx >= c.psize) continue;
    const int j = idx / c.nx, i = idx - j * c.nx;
    double y = 0.0;
    const IPair orow = ov_row(c.ov[MCF_OP_PRESSURE], j, i);
    if (p_act(c, j, i)) {
      double ax, aax; int nnz;
      p_row(c, X, idx, j, i, orow, ax, aax, nnz);
      y = sc * ax;
    }
    O[idx] = y;
  }
}

// First Gram-Schmidt pass: partial sums of raw V[q] . w (q <= col) and of w . w
__global__ __launch_bounds__(KR_T) void k_kry_dots(const CaseDev* __restrict__ cases, KryArgs a, int w_id) {
  const KryCtl& k = *a.ctl;
  if (!k.active) return;
  const CaseDev& c = cases[0];
  __shared__ double red[KR_MAXR + 2][KR_T / 64];
  const int col = k.col, n = c.psize;
  const double* W = c.buf[w_id];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  double w[KR_E];
  double ss = 0.0;
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    w[e] = idx < n ? W[idx] : 0.0;
    ss += w[e] * w[e];
  }
  for (int q = 0; q <= col; ++q) {
    const double* Vq = a.V + (size_t)q * a.ldv;
    double d = 0.0;
    #pragma unroll
    for (int e = 0; e < KR_E; ++e) {
      const int idx = base + e * KR_T;
      if (idx < n) d += Vq[idx] * w[e];
    }
    kr_wave_sum(d, q, red);
  }
  kr_wave_sum(ss, col + 1, red);
  kr_flush(red, col + 2, a);
}

// Gram-Schmidt update w -= sum_q coef[q] raw V[q].  pass 1: store w and the partial sums of the
// second projections V[q] . w; pass 2: store w as raw V[col + 1] with the partial sum of w . w.
__global__ __launch_bounds__(KR_T) void k_kry_orth(const CaseDev* __restrict__ cases, KryArgs a, int w_id, int pass) {
  const KryCtl& k = *a.ctl;
  if (!k.active) return;
  const CaseDev& c = cases[0];
  __shared__ double red[KR_MAXR + 2][KR_T / 64];
  __shared__ double cf[KR_MAXR + 1];
  const int col = k.col, n = c.psize;
  for (int q = threadIdx.x; q <= col; q += KR_T) cf[q] = k.coef[q];
  __syncthreads();
  double* W = c.buf[w_id];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  double w[KR_E];
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    w[e] = idx < n ? W[idx] : 0.0;
  }
  for (int q = 0; q <= col; ++q) {
    const double* Vq = a.V + (size_t)q * a.ldv;
    const double f = cf[q];
    #pragma unroll
    for (int e = 0; e < KR_E; ++e) {
      const int idx = base + e * KR_T;
      if (idx < n) w[e] -= f * Vq[idx];
    }
  }
  if (pass == 1) {
    #pragma unroll
    for (int e = 0; e < KR_E; ++e) {
      const int idx = base + e * KR_T;
      if (idx < n) W[idx] = w[e];
    }
    for (int q = 0; q <= col; ++q) {
      const double* Vq = a.V + (size_t)q * a.ldv;
      double d = 0.0;
      #pragma unroll
      for (int e = 0; e < KR_E; ++e) {
        const int idx = base + e * KR_T;
        if (idx < n) d += Vq[idx] * w[e];
      }
      kr_wave_sum(d, q, red);
    }
    kr_flush(red, col + 1, a);
  } else {
    double* Vn = a.V + (size_t)(col + 1) * a.ldv;
    double ss = 0.0;
    #pragma unroll
    for (int e = 0; e < KR_E; ++e) {
      const int idx = base + e * KR_T;
      if (idx < n) { Vn[idx] = w[e]; ss += w[e] * w[e]; }
    }
    kr_wave_sum(ss, 0, red);
    kr_flush(red, 1, a);
  }
}

// Partial sums of raw V[0] . raw V[0] (a new cycle's first vector)
__global__ __launch_bounds__(KR_T) void k_kry_norm(const CaseDev* __restrict__ cases, KryArgs a) {
  const CaseDev& c = cases[0];
  __shared__ double red[1][KR_T / 64];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  double ss = 0.0;
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    if (idx < c.psize) { const double v = a.V[idx]; ss += v * v; }
  }
  kr_wave_sum(ss, 0, red);
  kr_flush(red, 1, a);
}

// x += sum_{q < ncols} coef[q] raw V[q] (the cycle's update, coef = y * scale)
__global__ __launch_bounds__(KR_T) void k_kry_xupd(const CaseDev* __restrict__ cases, KryArgs a, int x_id) {
  const KryCtl& k = *a.ctl;
  const CaseDev& c = cases[0];
  __shared__ double cf[KR_MAXR + 1];
  const int m = k.ncols;
  for (int q = threadIdx.x; q < m; q += KR_T) cf[q] = k.coef[q];
  __syncthreads();
  double* X = c.buf[x_id];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  double acc[KR_E] = {};
  for (int q = 0; q < m; ++q) {
    const double* Vq = a.V + (size_t)q * a.ldv;
    const double f = cf[q];
    #pragma unroll
    for (int e = 0; e < KR_E; ++e) {
      const int idx = base + e * KR_T;
      if (idx < c.psize) acc[e] += f * Vq[idx];
    }
  }
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    if (idx < c.psize) X[idx] += acc[e];
  }
}

// r = b - A x on the active cells (0 elsewhere); partial sums of r^2
__global__ __launch_bounds__(KR_T) void k_kry_resid(const CaseDev* __restrict__ cases, KryArgs a, int b_id, int x_id, int r_id) {
  const CaseDev& c = cases[0];
  __shared__ double red[1][KR_T / 64];
  const double* B = c.buf[b_id];
  const double* X = c.buf[x_id];
  double* R = c.buf[r_id];
  const int base = blockIdx.x * KR_CHUNK + threadIdx.x;
  double ss = 0.0;
  #pragma unroll
  for (int e = 0; e < KR_E; ++e) {
    const int idx = base + e * KR_T;
    if (idx >= c.psize) continue;
    const int j = idx / c.nx, i = idx - j * c.nx;
    double r = 0.0;
    const IPair orow = ov_row(c.ov[MCF_OP_PRESSURE], j, i);
    if (p_act(c, j, i)) {
      double ax, aax; int nnz;
      p_row(c, X, idx, j, i, orow, ax, aax, nnz);
      r = B[idx] - ax;
    }
    R[idx] = r;
    ss += r * r;
  }
  kr_wave_sum(ss, 0, red);
  kr_flush(red, 1, a);
}

// x += y on the active cells (a refinement's correction)
__global__ __launch_bounds__(256) void k_kry_axpy(const CaseDev* __restrict__ cases, int x_id, int y_id) {
  const CaseDev& c = cases[0];
  double* X = c.buf[x_id];
  const double* Y = c.buf[y_id];
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < c.psize; idx += gridDim.x * blockDim.x)
    if (kr_active_cell(c, idx)) X[idx] += Y[idx];
}

// GMRES counts of pressure solve `psolve` into its slots (stripe 0)
__global__ void k_kry_record(const CaseDev* __restrict__ cases, int psolve, double iters, double refine, double status) {
  const CaseDev& c = cases[0];
  if (threadIdx.x != 0) return;
  c.sl[SL_GM_IT + psolve] = (u64)__double_as_longlong(iters);    // stripe 0
  c.sl[SL_GM_REF + psolve] = (u64)__double_as_longlong(refine);
  c.sl[SL_GM_ST + psolve] = (u64)__double_as_longlong(status);
}

// Plane rotation of LAPACK 3.10's dlartg (E. Anderson, "Algorithm 978: Safe scaling in the Level 1
// BLAS", ACM TOMS 44, 2017): [c s; -s c] [f; g] = [r; 0].
DEV void kr_lartg(double f, double g, double& c, double& s, double& r) {
  const double safmin = 2.2250738585072014e-308, safmax = 1.0 / safmin;
  const double rtmin = sqrt(safmin), rtmax = sqrt(safmax / 2.0);
  if (g == 0.0) { c = 1.0; s = 0.0; r = f; return; }
  if (f == 0.0) { c = 0.0; s = copysign(1.0, g); r = fabs(g); return; }
  const double f1 = fabs(f), g1 = fabs(g);
  if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    const double d = sqrt(f * f + g * g);
    c = f1 / d;
    r = copysign(d, f);
    s = g / r;
  } else {
    const double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
    const double fs = f / u, gs = g / u;
    const double d = sqrt(fs * fs + gs * gs);
    c = fabs(fs) / d;
    r = copysign(d, f);
    s = gs / r;
    r *= u;
  }
}

enum { KP_BEGIN = 0, KP_START_FIRST, KP_START, KP_DOTS1, KP_DOTS2, KP_NORM, KP_SOLVE, KP_RESID };

// Control step of the GMRES call: adds the block partial sums (one wave per sum, blocks in a fixed
// order) and updates the scalar state on one thread.
__global__ __launch_bounds__(KR_CT) void k_kry_ctl(KryArgs a, int phase) {
  KryCtl& k = *a.ctl;
  const bool arnoldi = phase == KP_DOTS1 || phase == KP_DOTS2 || phase == KP_NORM;
  if (arnoldi && !k.active) return;
  __shared__ double red[KR_MAXR + 2];
  const int col = k.col;
  const int nq = phase == KP_DOTS1 ? col + 2 : phase == KP_DOTS2 ? col + 1 : phase == KP_SOLVE ? 0 : 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = wv; q < nq; q += KR_CT / 64) {
    double s = 0.0;
    for (int b = lane; b < a.nblk; b += 64) s += a.part[(size_t)q * a.nblk + b];
    s = wave_sum(s);
    if (lane == 0) red[q] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  switch (phase) {
    case KP_BEGIN: {
      const double bn = sqrt(red[0]);
      k.bnrm2 = bn;
      k.atol = fmax(k.atol_in, k.rtol * bn);
      k.inner = 0; k.cycle = 0; k.info = 0; k.done = 0; k.active = 0; k.breakdown = 0; k.ncols = 0; k.col = 0;
      if (!isfinite(bn)) { k.done = 1; k.info = k.maxiter; }   // a non-finite right-hand side cannot converge
      else if (bn == 0.0 || bn < k.atol) k.done = 1;          // x = 0 is the answer
      break;
    }
    case KP_START_FIRST:
    case KP_START: {
      const double t = sqrt(red[0]);
      if (phase == KP_START_FIRST) { k.ptol = t * fmin(1.0, k.atol / k.bnrm2); k.pmax = 1.0; }
      k.scale[0] = 1.0 / t;
      for (int q = 0; q <= k.restart; ++q) k.S[q] = 0.0;
      k.S[0] = t;
      k.col = 0; k.breakdown = 0; k.active = 1;
      break;
    }
    case KP_DOTS1: {
      for (int q = 0; q <= col; ++q) { k.hk[q] = red[q] * k.scale[q]; k.coef[q] = k.hk[q] * k.scale[q]; }
      k.h0 = sqrt(red[col + 1]);
      break;
    }
    case KP_DOTS2: {
      for (int q = 0; q <= col; ++q) {
        const double d = red[q] * k.scale[q];
        k.h[col][q] = k.hk[q] + d;
        k.coef[q] = d * k.scale[q];
      }
      break;
    }
    case KP_NORM: {
      const double h1 = sqrt(red[0]);
      k.h[col][col + 1] = h1;
      if (h1 <= kEps * k.h0) { k.h[col][col + 1] = 0.0; k.breakdown = 1; }   // exact solution in the span
      else k.scale[col + 1] = 1.0 / h1;
      for (int q = 0; q < col; ++q) {
        const double cq = k.giv[q][0], sq = k.giv[q][1];
        const double n0 = k.h[col][q], n1 = k.h[col][q + 1];
        k.h[col][q] = cq * n0 + sq * n1;
        k.h[col][q + 1] = -sq * n0 + cq * n1;
      }
      double cr, sr, rr;
      kr_lartg(k.h[col][col], k.h[col][col + 1], cr, sr, rr);
      k.giv[col][0] = cr; k.giv[col][1] = sr;
      k.h[col][col] = rr; k.h[col][col + 1] = 0.0;
      const double t = -sr * k.S[col];
      k.S[col] = cr * k.S[col];
      k.S[col + 1] = t;
      k.presid = fabs(t);
      k.inner += 1;
      if (k.presid <= k.ptol || k.breakdown || col + 1 == k.restart) { k.active = 0; k.ncols = col + 1; }
      else k.col = col + 1;
      break;
    }
    case KP_SOLVE: {
      // back substitution with the triangular Hessenberg factor (singular columns skipped), then
      // coef = y * scale for the update of x
      const int m = k.ncols - 1;
      if (k.h[m][m] == 0.0) k.S[m] = 0.0;
      for (int q = 0; q <= m; ++q) k.coef[q] = k.S[q];
      for (int q = m; q > 0; --q) {
        if (k.coef[q] != 0.0) {
          k.coef[q] /= k.h[q][q];
          const double t = k.coef[q];
          for (int p = 0; p < q; ++p) k.coef[p] -= t * k.h[q][p];
        }
      }
      if (k.coef[0] != 0.0) k.coef[0] /= k.h[0][0];
      for (int q = 0; q <= m; ++q) k.coef[q] *= k.scale[q];
      break;
    }
    case KP_RESID: {
      const double rn = sqrt(red[0]);
      k.rnorm = rn;
      k.cycle += 1;
      if (rn <= k.atol) { k.done = 1; k.info = 0; }
      else if (!isfinite(rn) || k.breakdown || k.cycle >= k.maxiter) { k.done = 1; k.info = k.maxiter; }
      else {
        k.pmax = k.presid <= k.ptol ? fmax(kEps, 0.25 * k.pmax) : fmin(1.0, 1.5 * k.pmax);
        k.ptol = k.presid * fmin(k.pmax, k.atol / rn);
      }
      break;
    }
  }
}

}  // namespace

constexpr int N_GEMM_V = 7, N_TRI_V = 3;
// the x-FFT form of the rectangle inverse: transform k_xdct and sweep k_ysweep, a class's pair
constexpr int GV_XFFT = N_GEMM_V, TV_YSWEEP = N_TRI_V;
// the parity-split transform with 8- and 4-wave blocks (numbered after the x-FFT form)
constexpr int GV_PAR8 = GV_XFFT + 1, GV_PAR4 = GV_XFFT + 2, GV_LAST = GV_PAR4;
inline bool gv_par(int v) { return v == GV_PAR8 || v == GV_PAR4; }
constexpr int GEMM_TM[N_GEMM_V] = {32, 32, 64, 32, 16, 32, 16}, GEMM_TN[N_GEMM_V] = {64, 32, 64, 32, 32, 32, 32};
constexpr int TRI_TH[N_TRI_V] = {128, 256, 512};
struct Launch { int gemm_p = 0, gemm_h = 0, tri_p = 1, tri_h = 1; };

// A caller's shape for both classes: transform gemm with sweep width tri; gemm GV_XFFT selects the
// x-FFT pair for a class whose problems all have it, and the 16-row split-K transform with width
// tri for the other class.
void pin_launch(Launch& ln, int gemm, int tri, bool xf_p, bool xf_h) {
  const int fallback = 4;
  ln.gemm_p = gemm == GV_XFFT ? (xf_p ? GV_XFFT : fallback) : gemm;
  ln.gemm_h = gemm == GV_XFFT ? (xf_h ? GV_XFFT : fallback) : gemm;
  ln.tri_p = ln.gemm_p == GV_XFFT ? TV_YSWEEP : tri;
  ln.tri_h = ln.gemm_h == GV_XFFT ? TV_YSWEEP : tri;
}
// The parity-split transform runs where every problem of its class has the reflection parity (known
// once the class's rectangles are set, which for the velocity problems is at the first step; until
// then the request is reported as it stands), and the 16-row split-K transform elsewhere; the
// request itself is kept.
Launch effective_launch(Launch ln, bool par_p, bool par_h) {
  if (gv_par(ln.gemm_p) && !par_p) ln.gemm_p = 4;
  if (gv_par(ln.gemm_h) && !par_h) ln.gemm_h = 4;
  return ln;
}

inline dim3 gemm_grid(int v, int yy, int xx, int count) {
  return dim3((xx + GEMM_TN[v] - 1) / GEMM_TN[v], (yy + GEMM_TM[v] - 1) / GEMM_TM[v], count);
}
void launch_gemm(int v, int yy, int xx, int count, hipStream_t st, const SolveDesc* D, GemmArgs g) {
  if (v == GV_XFFT) {
    const int M = xx / 2;
    const size_t lds = xdct_lds(M, xdct_tables(M));
    const dim3 gx(yy, count);
    if (M <= 256) hipLaunchKernelGGL(k_xdct<1>, gx, dim3(256), lds, st, D, g);
    else if (M <= 512) hipLaunchKernelGGL(k_xdct<2>, gx, dim3(256), lds, st, D, g);
    else if (M <= 768) hipLaunchKernelGGL(k_xdct<3>, gx, dim3(256), lds, st, D, g);
    else hipLaunchKernelGGL(k_xdct<4>, gx, dim3(256), lds, st, D, g);
    return;
  }
  if (gv_par(v)) {
    const int hb = ((yy + 1) / 2 + 15) / 16;   // 16-row blocks of the larger parity / of the top rows
    const dim3 gp((xx + 31) / 32, g.forward ? 2 * hb : hb, count);
    if (v == GV_PAR8) {
      if (g.forward) hipLaunchKernelGGL((k_gemm_par<SK_S, true>), gp, dim3(64 * SK_S), 0, st, D, g);
      else hipLaunchKernelGGL((k_gemm_par<SK_S, false>), gp, dim3(64 * SK_S), 0, st, D, g);
    } else {
      if (g.forward) hipLaunchKernelGGL((k_gemm_par<SK_S / 2, true>), gp, dim3(32 * SK_S), 0, st, D, g);
      else hipLaunchKernelGGL((k_gemm_par<SK_S / 2, false>), gp, dim3(32 * SK_S), 0, st, D, g);
    }
    return;
  }
  const dim3 gg = gemm_grid(v, yy, xx, count);
  switch (v) {
    case 1: hipLaunchKernelGGL((k_gemm<32, 32, 64>), gg, dim3(256), 0, st, D, g); break;
    case 2: hipLaunchKernelGGL((k_gemm<64, 64, 32>), gg, dim3(256), 0, st, D, g); break;
    case 3: hipLaunchKernelGGL((k_gemm_sk<2, SK_S>), gg, dim3(64 * SK_S), 0, st, D, g); break;
    case 4: hipLaunchKernelGGL((k_gemm_sk<1, SK_S>), gg, dim3(64 * SK_S), 0, st, D, g); break;
    case 5: hipLaunchKernelGGL((k_gemm_sk<2, SK_S / 2>), gg, dim3(32 * SK_S), 0, st, D, g); break;
    case 6: hipLaunchKernelGGL((k_gemm_sk<1, SK_S / 2>), gg, dim3(32 * SK_S), 0, st, D, g); break;
    default: hipLaunchKernelGGL((k_gemm<32, 64, 64>), gg, dim3(256), 0, st, D, g); break;
  }
}
// y sweep shape for the largest rectangle height: segments per mode x rows per segment
void launch_ysweep(int yy, int xx, int count, hipStream_t st, const SolveDesc* D, PcrArgs p) {
  if (yy <= 128) hipLaunchKernelGGL((k_ysweep<16, 8>), dim3((xx + 15) / 16, count), dim3(256), 0, st, D, p);
  else if (yy <= 256) hipLaunchKernelGGL((k_ysweep<16, 16>), dim3((xx + 15) / 16, count), dim3(256), 0, st, D, p);
  else if (yy <= 512) hipLaunchKernelGGL((k_ysweep<16, 32>), dim3((xx + 15) / 16, count), dim3(256), 0, st, D, p);
  else if (yy <= 1024) hipLaunchKernelGGL((k_ysweep<32, 32>), dim3((xx + 7) / 8, count), dim3(256), 0, st, D, p);
  else hipLaunchKernelGGL((k_ysweep<64, 32>), dim3((xx + 3) / 4, count), dim3(256), 0, st, D, p);
}
void launch_tri(int v, int yy, int xx, int count, hipStream_t st, const SolveDesc* D, PcrArgs p) {
  if (v == TV_YSWEEP) { launch_ysweep(yy, xx, count, st, D, p); return; }
  switch (v) {
    case 0: hipLaunchKernelGGL(k_tri<128>, dim3(yy, count), dim3(128), 0, st, D, p); break;
    case 2: hipLaunchKernelGGL(k_tri<512>, dim3(yy, count), dim3(512), 0, st, D, p); break;
    default: hipLaunchKernelGGL(k_tri<256>, dim3(yy, count), dim3(256), 0, st, D, p); break;
  }
}

struct StepGraphs;

// GMRES workspace of a handle (allocated when the GMRES pressure backend is selected)
struct Krylov {
  int restart = 50, maxiter = 100;
  double rtol = 1e-12, atol = 1e-12;
  KryCtl* ctl = nullptr;
  double* V = nullptr;            // restart + 1 basis vectors
  double* part = nullptr;         // (KR_MAXR + 2) * nblk partial sums
  size_t ldv = 0;
  int nblk = 0;
  int* stat = nullptr;            // pinned host copy of the status words
  std::vector<double> seconds;    // wall time of each monitored solve since the last read, in call order
  hipGraphExec_t group = nullptr; // captured group of Arnoldi steps
  hipGraph_t group_graph = nullptr;
  int group_gemm = -1, group_tri = -1;   // launch shapes it was captured with
  void drop_group() {
    if (group) (void)hipGraphExecDestroy(group);
    if (group_graph) (void)hipGraphDestroy(group_graph);
    group = nullptr; group_graph = nullptr; group_gemm = group_tri = -1;
  }
};

struct mcf_handle {
  int nx, ny, usize, vsize, psize, velsize;
  double dx, dy, nu, rho;
  hipStream_t stream = nullptr;
  std::string err;
  CaseDev hc{};
  CaseDev* dc = nullptr;
  SolveDesc hs[3]{};      // pressure, u, v
  SolveDesc* ds = nullptr;
  bool has_rect[3] = {false, false, false};
  bool has_schur[3] = {false, false, false};
  bool has_ghost[3] = {false, false, false};
  bool has_solid[3] = {false, false, false};
  bool has_surface = false;
  std::map<std::string, void*> owned;
  std::vector<void*> bufs;
  std::vector<int> pt_pos[3];   // capacitance point (caller order) -> group order
  const int* rc_j[3] = {nullptr, nullptr, nullptr};   // y index of each capacitance row (group order)
  double* inlet = nullptr;
  u64* slots = nullptr;
  double* slot_vals = nullptr;   // NSLOT host-view slot values (k_slot_values)
  double* diag = nullptr;
  int diag_cap = 0;
  double prev_dt = 0.0, time = 0.0;
  bool has_history = false;
  int parity = 0;      // state in S[parity], history in H[parity]
  double* tmp_rect = nullptr;   // scratch for the rectangle-inverse entry point
  size_t tmp_rect_n = 0;
  double* na_scratch = nullptr; // this device's copy of `scratch` (allocation-free rectangle inverse)
  double graph_dt = -1.0;
  bool use_graphs = true;
  int* clock_step = nullptr;
  double* clock_time = nullptr;
  double* prm = nullptr;
  bool prm_valid = false;
  double prm_dt = 0.0, prm_alpha = 0.0, prm_factor = 0.0;
  StepGraphs* graphs = nullptr;
  Launch ln;                 // launch shapes of this handle's solves
  bool tuned = false;        // ln chosen (by the autotuner or the caller)
  StepGraphs* lean_graphs = nullptr;
  StepGraphs* mixed[8] = {};   // other lean masks (LEAN_VEL, LEAN_VEL | ONCE_PRES), built on first use
  int spec = 1;              // speculative lean steps: 0 off, 1 on, 2 on with every lean step flagged
  int replays = 0;           // chunks replayed with the refinements in place
  int spec_chunk = 8;        // steps per speculative chunk: doubles after a clean chunk, back to 8 on a miss
  int spec_full = 0;         // length of the next chunk when it runs in full (0: the next chunk is lean)
  bool spec_vfull = false;   // the last chunk refined a velocity solve: full chunks refine every solve
  bool spec_p2 = false;      // the last chunk refined a pressure solve twice: full chunks allow it
  int pbackend = MCF_PB_DIRECT;   // pressure backend
  Krylov* kry = nullptr;
};

namespace {

int fail(mcf_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHK(h, call)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) return fail(h, 4, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int upload(mcf_handle* h, const std::string& key, const T* src, size_t n, T** dst) {
  auto it = h->owned.find(key);
  if (it != h->owned.end()) { (void)hipFree(it->second); h->owned.erase(it); }
  *dst = nullptr;
  if (n == 0) return 0;
  void* p = nullptr;
  HIPCHK(h, hipMalloc(&p, n * sizeof(T)));
  h->owned[key] = p;
  if (src) HIPCHK(h, hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice));
  else HIPCHK(h, hipMemset(p, 0, n * sizeof(T)));
  *dst = static_cast<T*>(p);
  return 0;
}

void drop_graphs(mcf_handle* h);
int sync_descs(mcf_handle* h) {
  drop_graphs(h);
  HIPCHK(h, hipMemcpy(h->dc, &h->hc, sizeof(CaseDev), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->ds, h->hs, sizeof(h->hs), hipMemcpyHostToDevice));
  return 0;
}

inline dim3 grid1(int n, int nb = 256) { return dim3((n + nb - 1) / nb, 1, 1); }

// Vsel[k][t] = V[j_t][k]: the y eigenbasis at the capacitance rows' y indices, laid out so that the
// correction sweep of mode k reads its r factors contiguously.
__global__ __launch_bounds__(256) void k_vsel(const double* V, const int* rj, int yy, int r, double* out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)yy * r) return;
  const int k = (int)(e / r), t = (int)(e % r);
  out[e] = V[(size_t)rj[t] * yy + k];
}

// (Re)builds a problem's Vsel from its current eigenbasis and capacitance rows; synchronous.
int build_vsel(mcf_handle* h, int kind) {
  SolveDesc& s = h->hs[kind];
  double* dv = nullptr;
  const size_t n = (size_t)s.yy * s.r;
  if (int rc = upload<double>(h, "schur" + std::to_string(kind) + "vs", nullptr, n, &dv)) return rc;
  s.Vsel = dv;
  if (!n) return 0;
  hipLaunchKernelGGL(k_vsel, grid1((int)n), dim3(256), 0, h->stream, s.V, h->rc_j[kind], s.yy, s.r, dv);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}
// grid of the grid-stride kernels
inline dim3 gridS(int n) { return dim3(std::min((n + 255) / 256, GSB), 1, 1); }

int launch_check(mcf_handle* h, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(h, 4, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

int sv_block(int kind, int psolve) {
  if (kind == MCF_U) return SL_HU;
  if (kind == MCF_V) return SL_HV;
  return SL_P0 + 8 * psolve;
}

// Everything a launch needs: the case table, the solve problems and the launch shapes (maxima
// over the cases).  One case and a batch of uneven cases use the same code.
struct Ctx {
  mcf_handle* gm = nullptr;   // set when the pressure solves run GMRES (one case)
  Launch ln;             // launch shapes
  bool dump;             // record the (rhs, multiplier) pairs of first pressure solves
  mcf_handle* err;       // receives error messages
  hipStream_t stream;
  const CaseDev* dc;     // n cases
  const SolveDesc* dsp;  // n pressure problems
  const SolveDesc* dsh;  // 2n Helmholtz problems (u, v per case)
  int n;
  int maxvel, maxp, maxfill, maxsurf, maxny;
  bool pert;             // some case perturbs its inflow: each step sets its inlet first (k_inlet)
  bool noxf_p, noxf_h;   // some pressure / velocity problem has no x-FFT form
  bool nopar_p, nopar_h; // some pressure / velocity problem's basis lacks the reflection parity
  int p_band, h_band;    // most band rows of a problem's capacitance entries (x-FFT form)
  int p_yy, p_xx, p_np, p_r, h_yy, h_xx, h_np, h_r;
  StepPrm sp{};          // a single case's step parameters (dt = 0: read from the case)
  // the velocity solve's right-hand side maxima (a step's SL_RHS pair; an entry point's own slots)
  int hval_u = SL_HU + SV_VAL, hval_v = SL_HV + SV_VAL;
  bool defer = false;    // the caller launches a pending step tail (one handle's steps)
};

void ctx_shapes(Ctx& x, const CaseDev& c, const SolveDesc* hs3) {
  x.maxvel = std::max(x.maxvel, c.velsize);
  x.maxp = std::max(x.maxp, c.psize);
  x.maxfill = std::max(x.maxfill, c.gt[0].n + c.gt[1].n + c.ny);
  x.maxsurf = std::max(x.maxsurf, c.nsurf);
  x.maxny = std::max(x.maxny, c.ny);
  x.pert = x.pert || c.pert_T > 0.0;
  x.noxf_p = x.noxf_p || !hs3[0].xf;
  x.noxf_h = x.noxf_h || !hs3[1].xf || !hs3[2].xf;
  // a rectangle not set yet (yy 0; the velocity ones come with the first step) does not count
  auto nopar = [](const SolveDesc& d) { return d.yy > 0 && !d.par; };
  x.nopar_p = x.nopar_p || nopar(hs3[0]);
  x.nopar_h = x.nopar_h || nopar(hs3[1]) || nopar(hs3[2]);
  x.p_band = std::max(x.p_band, hs3[0].yr_n);
  x.h_band = std::max({x.h_band, hs3[1].yr_n, hs3[2].yr_n});
  x.p_yy = std::max(x.p_yy, hs3[0].yy); x.p_xx = std::max(x.p_xx, hs3[0].xx);
  x.p_np = std::max(x.p_np, hs3[0].npts); x.p_r = std::max(x.p_r, hs3[0].r);
  for (int q = 1; q < 3; ++q) {
    x.h_yy = std::max(x.h_yy, hs3[q].yy); x.h_xx = std::max(x.h_xx, hs3[q].xx);
    x.h_np = std::max(x.h_np, hs3[q].npts); x.h_r = std::max(x.h_r, hs3[q].r);
  }
}

Ctx ctx_of(mcf_handle* h) {
  Ctx x{};
  x.err = h; x.stream = h->stream; x.dc = h->dc; x.dsp = h->ds; x.dsh = h->ds + 1; x.n = 1;
  x.ln = h->ln;
  x.dump = h->hs[0].dump_cap > 0;
  x.gm = h->pbackend == MCF_PB_GMRES ? h : nullptr;
  x.defer = true;
  ctx_shapes(x, h->hc, h->hs);
  x.ln = effective_launch(x.ln, !x.nopar_p, !x.nopar_h);
  return x;
}

// Grid of the grid-stride kernels over `cases` cases.  One case: at most GSB_SINGLE blocks (five
// per CU), which keeps one element per thread for the pressure-sized stages and gives the
// velocity-sized ones about two, so that their waves fit the CUs' wave slots in one round (measured
// 0.6-1.2 us less per velocity-sized launch at 1320x246).  A batch: at most GSB_BATCH blocks per
// case, so that each block pays its prologue (the case's parameters and buffer pointers, several
// dependent loads) over several elements; the batch has parallelism to spare (the 64-case
// ensemble measured 1.668 / 1.681 / 1.646 / 1.586e9 cell-steps/s at 128 / 256 / 512 / 1024).
constexpr int GSB_SINGLE = 1280, GSB_BATCH = 256;
inline dim3 gridSn(int n, int cases) {
  dim3 g = gridS(n);
  g.x = std::min<unsigned>(g.x, cases > 1 ? GSB_BATCH : GSB_SINGLE);
  g.y = cases;
  return g;
}
inline dim3 grid1n(int n, int cases) { dim3 g = grid1(std::max(n, 1)); g.y = cases; return g; }
// The advection stage holds more registers (four waves per SIMD instead of five): one case gets at
// most GSB_ADV blocks, the count that fills the wave slots once (1320x246: 11.7 us against 12.1 us
// at GSB_SINGLE, which leaves a quarter of the blocks for a second round).
constexpr int GSB_ADV = 1024;
inline dim3 gridAdv(int n, int cases) {
  dim3 g = gridSn(n, cases);
  if (cases == 1) g.x = std::min<unsigned>(g.x, GSB_ADV);
  return g;
}

// One Woodbury apply (in -> out): epi 1 store masked, 2 accumulate masked.
// The capacitance points and multipliers of a solve (the x-FFT form's own kernels for GV_XFFT)
void launch_points_cap(const Ctx& x, bool pres, int gv, const SolveDesc* D, int count, int xx, int np, int r,
                       PcrArgs p0) {
  if (gv == GV_XFFT) {
    hipLaunchKernelGGL(k_xpoints, dim3((np + 3) / 4, count), dim3(256), 0, x.stream, D, p0);
    const dim3 gb(pres ? x.p_band : x.h_band, count);
    if (xx <= 512) hipLaunchKernelGGL(k_xcap<2>, gb, dim3(256), 0, x.stream, D, p0);
    else if (xx <= 1024) hipLaunchKernelGGL(k_xcap<4>, gb, dim3(256), 0, x.stream, D, p0);
    else if (xx <= 2048) hipLaunchKernelGGL(k_xcap<8>, gb, dim3(256), 0, x.stream, D, p0);
    else hipLaunchKernelGGL(k_xcap<16>, gb, dim3(256), 0, x.stream, D, p0);
  } else {
    hipLaunchKernelGGL(k_points, dim3((np + 3) / 4, count), dim3(256), 0, x.stream, D, p0);
    hipLaunchKernelGGL(k_capacitance, dim3((r + 3) / 4, count), dim3(256), 0, x.stream, D, p0);
  }
}

int enqueue_apply(const Ctx& x, bool pres, int in_id, int out_id, int epi, int pred_slot, int slot_shift) {
  const SolveDesc* D = pres ? x.dsp : x.dsh;
  const int count = pres ? x.n : 2 * x.n;
  const int yy = pres ? x.p_yy : x.h_yy, xx = pres ? x.p_xx : x.h_xx;
  const int np = pres ? x.p_np : x.h_np, r = pres ? x.p_r : x.h_r;
  const int gv = pres ? x.ln.gemm_p : x.ln.gemm_h, tv = pres ? x.ln.tri_p : x.ln.tri_h;
  GemmArgs g{1, in_id, out_id, 0, pred_slot, slot_shift};
  launch_gemm(gv, yy, xx, count, x.stream, D, g);
  PcrArgs p0{0, pred_slot, slot_shift};
  launch_tri(tv, yy, xx, count, x.stream, D, p0);
  if (r > 0) {
    launch_points_cap(x, pres, gv, D, count, xx, np, r, p0);
    if (x.dump && pres && pred_slot < 0)
      hipLaunchKernelGGL(k_dump, dim3(1, count), dim3(256), 0, x.stream, D, slot_shift / 8);
    PcrArgs p1{1, pred_slot, slot_shift};
    launch_tri(tv, yy, xx, count, x.stream, D, p1);
  }
  GemmArgs gi{0, in_id, out_id, epi, pred_slot, slot_shift};
  launch_gemm(gv, yy, xx, count, x.stream, D, gi);
  return launch_check(x.err, "woodbury apply");
}

int read_slots(mcf_handle* h, std::vector<double>& out);

// ---- stage bodies and merged launches ----
Body body_of(int kind, int nblk) { Body b; std::memset(&b, 0, sizeof b); b.kind = kind; b.nblk = nblk; return b; }
Body body_div(const Ctx& x, DivArgs a) { Body b = body_of(BK_DIV, gridSn(x.maxp, x.n).x); a.sp = x.sp; b.d = a; return b; }
Body body_helm(const Ctx& x, HelmArgs a) { Body b = body_of(BK_HELM, gridSn(x.maxvel, x.n).x); a.sp = x.sp; b.h = a; return b; }
Body body_pres(const Ctx& x, PresArgs a) { Body b = body_of(BK_PRES, gridSn(x.maxp, x.n).x); b.p = a; return b; }
Body body_upd(const Ctx& x, UpdArgs a) { Body b = body_of(BK_UPD, gridSn(x.maxvel, x.n).x); a.sp = x.sp; b.u = a; return b; }
Body body_press(const Ctx& x, PressArgs a) { Body b = body_of(BK_PRESS, gridSn(x.maxp, x.n).x); a.sp = x.sp; b.q = a; return b; }
Body body_fill(const Ctx& x, FillArgs a) { Body b = body_of(BK_FILL, grid1n(x.maxfill, 1).x); b.f = a; return b; }
Body no_body() { return body_of(BK_NONE, 0); }
Body body_fin() { return body_of(BK_FIN, 1); }

// Launch one stage body with its own kernel
void launch_body(const Ctx& x, const Body& b) {
  const dim3 g(b.nblk, x.n), blk(256);
  switch (b.kind) {
    case BK_DIV: hipLaunchKernelGGL(k_div, g, blk, 0, x.stream, x.dc, b.d); break;
    case BK_HELM: hipLaunchKernelGGL(k_helm, g, blk, 0, x.stream, x.dc, b.h); break;
    case BK_PRES: hipLaunchKernelGGL(k_pres, g, blk, 0, x.stream, x.dc, b.p); break;
    case BK_UPD: hipLaunchKernelGGL(k_update, g, blk, 0, x.stream, x.dc, b.u); break;
    case BK_PRESS: hipLaunchKernelGGL(k_pressures, g, blk, 0, x.stream, x.dc, b.q); break;
    case BK_FILL: hipLaunchKernelGGL(k_fill, g, blk, 0, x.stream, x.dc, b.f); break;
    case BK_FIN: hipLaunchKernelGGL(k_finalize, dim3(1, x.n), dim3(64), 0, x.stream, x.dc); break;
    default: break;
  }
}

// Launch the given stages together.  One case: one launch (k_fused); a batch of cases: one launch
// per stage, since the launch cost is shared by the cases and the stages' own kernels keep their
// higher occupancy on the larger grids.
void launch_bodies(const Ctx& x, std::initializer_list<Body> bodies) {
  Fused f;
  std::memset(&f, 0, sizeof f);
  int total = 0;
  for (const Body& b : bodies)
    if (b.kind != BK_NONE) { f.b[f.n++] = b; total += b.nblk; }
  if (f.n == 0) return;
  if (f.n == 1 || x.n > 1) {
    for (int q = 0; q < f.n; ++q) launch_body(x, f.b[q]);
    return;
  }
  const dim3 g(total, x.n), blk(256);
  const int k0 = f.b[0].kind, k1 = f.b[1].kind, k2 = f.n > 2 ? f.b[2].kind : BK_NONE;
  if (k0 == BK_HELM && k1 == BK_DIV && k2 == BK_NONE)
    hipLaunchKernelGGL((k_fused_of<BK_HELM, BK_DIV, BK_NONE>), g, blk, 0, x.stream, x.dc, f);
  else if (k0 == BK_PRES && k1 == BK_UPD && k2 == BK_NONE)
    hipLaunchKernelGGL((k_fused_of<BK_PRES, BK_UPD, BK_NONE>), g, blk, 0, x.stream, x.dc, f);
  else if (k0 == BK_DIV && k1 == BK_HELM && k2 == BK_NONE)
    hipLaunchKernelGGL((k_fused_of<BK_DIV, BK_HELM, BK_NONE>), g, blk, 0, x.stream, x.dc, f);
  else if (k0 == BK_HELM && k1 == BK_DIV && k2 == BK_FIN)
    hipLaunchKernelGGL((k_fused_of<BK_HELM, BK_DIV, BK_FIN>), g, blk, 0, x.stream, x.dc, f);
  else if (k0 == BK_DIV && k1 == BK_PRESS && k2 == BK_FILL)
    hipLaunchKernelGGL((k_fused_of<BK_DIV, BK_PRESS, BK_FILL>), g, blk, 0, x.stream, x.dc, f);
  else
    hipLaunchKernelGGL(k_fused, g, blk, 0, x.stream, x.dc, f);
}

// ---- GMRES pressure backend (host control; the device kernels are above) ----
constexpr int KR_GROUP = 8;   // Arnoldi steps enqueued between reads of the status words

KryArgs kry_args(const Krylov& K) { return KryArgs{K.ctl, K.V, K.ldv, K.part, K.nblk}; }

int kry_read(mcf_handle* h) {
  Krylov& K = *h->kry;
  HIPCHK(h, hipMemcpyAsync(K.stat, K.ctl, KR_STAT_INTS * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// The preconditioner: rectangle inverse of in_id into out_id, zero off the active cells
void kry_precondition(const Ctx& x, int in_id, int out_id, const int* gate) {
  GemmArgs g{1, in_id, out_id, 0, -1, 0, gate};
  launch_gemm(x.ln.gemm_p, x.p_yy, x.p_xx, 1, x.stream, x.dsp, g);
  PcrArgs p{0, -1, 0, gate};
  launch_tri(x.ln.tri_p, x.p_yy, x.p_xx, 1, x.stream, x.dsp, p);
  GemmArgs gi{0, in_id, out_id, 1, -1, 0, gate};
  launch_gemm(x.ln.gemm_p, x.p_yy, x.p_xx, 1, x.stream, x.dsp, gi);
}

// KR_GROUP Arnoldi steps: each reads the column index and the activity flag from the control
// state, so the same launches serve every step of every cycle
void enqueue_arnoldi_group(const Ctx& x, const Krylov& K) {
  const KryArgs a = kry_args(K);
  const dim3 gb(K.nblk), tb(KR_T), gc(1), tc(KR_CT);
  hipStream_t st = x.stream;
  const int* gate = &K.ctl->active;
  for (int s = 0; s < KR_GROUP; ++s) {
    hipLaunchKernelGGL(k_kry_matvec, gb, tb, 0, st, x.dc, a, (int)B_KA);
    kry_precondition(x, B_KA, B_KB, gate);
    hipLaunchKernelGGL(k_kry_dots, gb, tb, 0, st, x.dc, a, (int)B_KB);
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_DOTS1);
    hipLaunchKernelGGL(k_kry_orth, gb, tb, 0, st, x.dc, a, (int)B_KB, 1);
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_DOTS2);
    hipLaunchKernelGGL(k_kry_orth, gb, tb, 0, st, x.dc, a, (int)B_KB, 2);
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_NORM);
  }
}

// One group of Arnoldi steps: a graph captured once per launch shape and replayed (eager launches
// when graphs are off)
int kry_group(const Ctx& x) {
  mcf_handle* h = x.gm;
  Krylov& K = *h->kry;
  if (!h->use_graphs) {
    enqueue_arnoldi_group(x, K);
    return launch_check(h, "gmres arnoldi steps");
  }
  if (!K.group || K.group_gemm != x.ln.gemm_p || K.group_tri != x.ln.tri_p) {
    K.drop_group();
    HIPCHK(h, hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
    enqueue_arnoldi_group(x, K);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(x.stream, &g);
    if (e != hipSuccess) return fail(h, 4, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    K.group_graph = g;
    HIPCHK(h, hipGraphInstantiate(&K.group, g, nullptr, nullptr, 0));
    K.group_gemm = x.ln.gemm_p;
    K.group_tri = x.ln.tri_p;
  }
  HIPCHK(h, hipGraphLaunch(K.group, x.stream));
  return 0;
}

// One GMRES call on the right-hand side b_id from x0 = 0 into x_id (SciPy's gmres with the
// reference's arguments).  The host enqueues the Arnoldi steps of a cycle in groups and reads the
// status words after each group; the device ends the cycle (active = 0) on its inner test or after
// `restart` steps, after which the remaining launches of the group return at once.  At most
// maxiter cycles run.
int kry_call(const Ctx& x, int b_id, int x_id, int& inner, int& info) {
  mcf_handle* h = x.gm;
  Krylov& K = *h->kry;
  const KryArgs a = kry_args(K);
  const dim3 gb(K.nblk), tb(KR_T), gc(1), tc(KR_CT);
  hipStream_t st = x.stream;
  hipLaunchKernelGGL(k_kry_begin, gb, tb, 0, st, x.dc, a, b_id, x_id, (int)B_KR);
  hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_BEGIN);
  if (int rc = kry_read(h)) return rc;
  bool first = true;
  while (!K.stat[1]) {
    kry_precondition(x, B_KR, B_KV0, nullptr);
    hipLaunchKernelGGL(k_kry_norm, gb, tb, 0, st, x.dc, a);
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)(first ? KP_START_FIRST : KP_START));
    first = false;
    for (int s0 = 0; s0 < K.restart; s0 += KR_GROUP) {
      if (int rc = kry_group(x)) return rc;
      if (int rc = kry_read(h)) return rc;
      if (!K.stat[0]) break;
    }
    if (K.stat[0]) return fail(h, 4, "GMRES cycle did not end after restart steps");
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_SOLVE);
    hipLaunchKernelGGL(k_kry_xupd, gb, tb, 0, st, x.dc, a, x_id);
    hipLaunchKernelGGL(k_kry_resid, gb, tb, 0, st, x.dc, a, b_id, x_id, (int)B_KR);
    hipLaunchKernelGGL(k_kry_ctl, gc, tc, 0, st, a, (int)KP_RESID);
    if (int rc = kry_read(h)) return rc;
  }
  inner = K.stat[3];
  info = K.stat[2];
  return launch_check(h, "gmres");
}

// Monitored GMRES pressure solve (pressure_schur.py:33-70): GMRES, then while GMRES succeeded and
// the true residual's max exceeds max(1e-10, 64 eps max(1, max|b|)), at most twice, a GMRES solve
// on that residual added to x; then the final statistics into the solve's slot block and the
// GMRES counts (Arnoldi steps of all calls, refinements, status) into its record slots.
int kry_solve(const Ctx& x, int b_id, int x_id, int psolve) {
  mcf_handle* h = x.gm;
  const int slot0 = SL_P0 + 8 * psolve;
  const auto t0 = std::chrono::steady_clock::now();
  int inner = 0, info = 0;
  if (int rc = kry_call(x, b_id, x_id, inner, info)) return rc;
  int total = inner, status = info, refinements = 0;
  for (int it = 0; it < 2 && !status; ++it) {
    const int res_slot = it == 0 ? SV_RES1 : SV_RES2;
    PresArgs r{b_id, x_id, B_RESP, 0, slot0, res_slot, -1};
    hipLaunchKernelGGL(k_pres, gridSn(x.maxp, 1), dim3(256), 0, x.stream, x.dc, r);
    std::vector<double> sv;
    if (int rc = read_slots(h, sv)) return rc;
    const double target = std::max(1e-10, 64 * kEps * std::max(1.0, sv[slot0 + SV_VAL]));
    if (sv[slot0 + res_slot] <= target) break;
    if (int rc = kry_call(x, B_RESP, B_KX, inner, info)) return rc;
    total += inner;
    status = info;
    ++refinements;
    hipLaunchKernelGGL(k_kry_axpy, gridSn(x.maxp, 1), dim3(256), 0, x.stream, x.dc, x_id, (int)B_KX);
  }
  PresArgs f{b_id, x_id, -1, 1, slot0, SV_FINAL, -1};
  hipLaunchKernelGGL(k_pres, gridSn(x.maxp, 1), dim3(256), 0, x.stream, x.dc, f);
  hipLaunchKernelGGL(k_kry_record, dim3(1), dim3(64), 0, x.stream, x.dc, psolve, (double)total,
                     (double)refinements, (double)status);
  HIPCHK(h, hipStreamSynchronize(x.stream));
  h->kry->seconds.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return launch_check(h, "gmres pressure solve");
}

// Woodbury solve with <= 2 refinements (pressure_schur.py:116-124) and the final residual, with
// `refs` refinement passes launched (each runs only when the residual before it calls for it).  The
// lean form (refs = 0) computes the first and the final residual in one pass; refs = 1 leaves the
// second pass out.  A step whose solve would have needed a pass that was left out is flagged by
// the step's slot gather (finalize_slots) and replayed in full.
// With `tail`, the final-statistics stage is returned there instead of launched, for the caller to
// merge into its next launch.
int enqueue_solve(const Ctx& x, bool pres, int in_id, int out_id, int psolve, int refs = 2,
                  Body* tail = nullptr) {
  if (tail) *tail = no_body();
  if (pres && x.gm) return kry_solve(x, in_id, out_id, psolve);
  const int shift = pres ? 8 * psolve : 0;
  const int slot0 = pres ? SL_P0 + 8 * psolve : SL_HU;
  int rc;
  if ((rc = enqueue_apply(x, pres, in_id, out_id, 1, -1, shift))) return rc;
  if (refs == 0) {
    const Body fin = pres ? body_pres(x, PresArgs{in_id, out_id, -1, 3, slot0, SV_RES1, -1})
                          : body_helm(x, HelmArgs{in_id, out_id, -1, -1, 3, SV_RES1, -1});
    if (tail) *tail = fin;
    else launch_bodies(x, {fin});
    return launch_check(x.err, "lean woodbury solve");
  }
  for (int it = 0; it < refs; ++it) {
    const int res_slot = it == 0 ? SV_RES1 : SV_RES2;
    const int pred = it == 0 ? -1 : SV_RES1;
    // the residual, then the pass's decision (k_decide), which every launch of the pass reads
    if (pres) {
      PresArgs a{in_id, out_id, B_RESP, 0, slot0, res_slot, pred};
      hipLaunchKernelGGL(k_pres, gridSn(x.maxp, x.n), dim3(256), 0, x.stream, x.dc, a);
      hipLaunchKernelGGL(k_decide, dim3(1, x.n), dim3(64), 0, x.stream, x.dc, slot0, -1, res_slot, slot0 + (int)SV_VAL, -1);
      if ((rc = enqueue_apply(x, pres, B_RESP, out_id, 2, res_slot, shift))) return rc;
    } else {
      HelmArgs a{in_id, out_id, B_RES, -1, 0, res_slot, pred};
      a.sp = x.sp;
      hipLaunchKernelGGL(k_helm, gridSn(x.maxvel, x.n), dim3(256), 0, x.stream, x.dc, a);
      hipLaunchKernelGGL(k_decide, dim3(1, x.n), dim3(64), 0, x.stream, x.dc, (int)SL_HU, (int)SL_HV, res_slot,
                         x.hval_u, x.hval_v);
      if ((rc = enqueue_apply(x, pres, B_RES, out_id, 2, res_slot, shift))) return rc;
    }
  }
  const Body fin = pres ? body_pres(x, PresArgs{in_id, out_id, -1, 1, slot0, SV_FINAL, -1})
                        : body_helm(x, HelmArgs{in_id, out_id, -1, -1, 1, SV_FINAL, -1});
  if (tail) *tail = fin;
  else launch_bodies(x, {fin});
  return launch_check(x.err, "woodbury solve");
}

// project(x, dtmul * dt): div -> pressure solve -> update -> divergence gate (optimized_ghost_cell.py:218-229).
// `head` (an independent stage of the caller) is launched with the divergence; the solve's final
// statistics are launched with the update; with `gate`, the divergence gate is returned there for
// the caller to merge into its next launch.
int enqueue_project(const Ctx& x, int x_id, int out_id, int phi_id, double dtmul, int psolve, int defect_id,
                    int refs = 2, const Body* head = nullptr, Body* gate = nullptr, bool prev_fin = false) {
  int rc;
  launch_bodies(x, {head ? *head : no_body(), body_div(x, DivArgs{x_id, B_DIVP, dtmul, SL_P0 + 8 * psolve + SV_VAL}),
                    prev_fin ? body_fin() : no_body()});
  Body fin;
  if ((rc = enqueue_solve(x, true, B_DIVP, phi_id, psolve, refs, &fin))) return rc;
  launch_bodies(x, {fin, body_upd(x, UpdArgs{x_id, phi_id, out_id, defect_id, dtmul})});
  const Body g = body_div(x, DivArgs{out_id, -1, 0.0, SL_DIV0 + psolve});
  if (gate) *gate = g;
  else launch_bodies(x, {g});
  return launch_check(x.err, "project");
}

int check_ready(mcf_handle* h, double dt) {
  for (int k = 0; k < 3; ++k) {
    if (!h->has_rect[k]) return fail(h, 1, "rectangle factor missing");
    if (!h->has_schur[k]) return fail(h, 1, "capacitance data missing");
    if (!h->has_ghost[k] || !h->has_solid[k]) return fail(h, 1, "ghost tables missing");
  }
  if (!h->has_surface) return fail(h, 1, "surface quadrature missing");
  for (int k = 1; k < 3; ++k) {
    const double want = -0.5 * dt * h->nu;
    if (std::fabs(h->hs[k].b - want) > 1e-14 * std::fabs(want)) return fail(h, 1, "Helmholtz factors were built for another dt");
  }
  return 0;
}

// A step's tail: its surface forces and the slot gather into its diagnostics row (k_forces), then
// the forces into the row and the clock (k_finalize), for the step of parity `parity`, whose new
// state is S[1 - parity].
FinArgs step_fin(const Ctx& x, int parity, int nproj, int lean, bool force_miss) {
  return FinArgs{nproj, lean, force_miss ? 1 : 0, x.gm ? 1 : 0, SL_RHS + 2 * parity};
}
ForceArgs step_forces(int parity) { return ForceArgs{B_S0 + 1 - parity, B_POUT, -1.0}; }
void launch_tail(const Ctx& x, int parity, const FinArgs& fin) {
  hipLaunchKernelGGL(k_forces, dim3(force_blocks(x.maxsurf) + 1, x.n), dim3(64), 0, x.stream, x.dc,
                     step_forces(parity), fin, 1);
  hipLaunchKernelGGL(k_finalize, dim3(1, x.n), dim3(64), 0, x.stream, x.dc);
}
// A step's tail reads the new state, the output pressure, the slots (which it clears) and the
// clock; the next step's first stage reads the state and writes only its own SL_RHS pair, so the
// tail's gather and force waves can run in the first stage's launch (k_adv_tail) and its last part
// in the launch before the first pressure solve (BK_FIN).  TAIL_OWN: the step launches its own tail last; TAIL_DEFER: it
// leaves its tail to the caller; TAIL_PREV: it leaves its own and runs the previous step's (parity
// 1 - parity, same lean form) inside its own launches.  Only one case's steady
// steps defer (a GMRES step, which reads the slots back, runs its own); a step that sets the inlet
// first tells k_inlet that the clock advance of the step before is still pending.
enum { TAIL_OWN = 0, TAIL_DEFER = 1, TAIL_PREV = 2 };
bool tail_can_defer(const Ctx& x) { return x.defer && x.n == 1 && !x.gm; }

// One AB2/CN step of every case of the context (time_integration.py:7-145,
// optimized_ghost_cell.py:236-286, main.py:383-415).  Launches only: dt, alpha and the pressure
// factor are read from each case's parameters (k_set_prm); the caller advances the host state.
int enqueue_step_kernels(const Ctx& x, int parity, bool first, int lean = 0, bool force_miss = false,
                         int tail = TAIL_OWN) {
  if (first) lean = 0;
  if (first || !tail_can_defer(x)) tail = TAIL_OWN;
  const int vrefs = (lean & LEAN_VEL) ? 0 : 2;
  const int prefs = (lean & LEAN_PRES) ? 0 : (lean & ONCE_PRES) ? 1 : 2;
  int rc;
  const int S = B_S0 + parity, Snew = B_S0 + 1 - parity;
  const int Hprev = B_H0 + 1 - parity, Hnow = B_H0 + parity;
  const dim3 ga = gridAdv(x.maxvel, x.n);
  int nproj = 2;
  const int rhs = SL_RHS + 2 * parity;
  if (x.pert)
    hipLaunchKernelGGL(k_inlet, dim3((x.maxny + 255) / 256, x.n), dim3(256), 0, x.stream, x.dc, tail == TAIL_PREV ? 1 : 0);
  if (first) {
    // projected midpoint startup (time_integration.py:33-47)
    AdvArgs m{S, S, B_PH, -1, Hnow, B_TV0, 1, x.sp, rhs};
    hipLaunchKernelGGL(k_adv_rhs, ga, dim3(256), 0, x.stream, x.dc, m);
    if ((rc = enqueue_project(x, B_TV0, B_HALF, B_TP0, 0.5, 2, -1))) return rc;
    FillArgs f{B_HALF, 1, 1, 0};
    hipLaunchKernelGGL(k_fill, grid1n(x.maxfill, x.n), dim3(256), 0, x.stream, x.dc, f);
    AdvArgs r{S, B_HALF, B_PH, -1, -1, B_RHS, 2, x.sp, rhs};
    hipLaunchKernelGGL(k_adv_rhs, ga, dim3(256), 0, x.stream, x.dc, r);
    nproj = 3;
  } else if (tail == TAIL_PREV) {
    AdvArgs r{S, S, B_PH, Hprev, Hnow, B_RHS, 0, x.sp, rhs};
    // the advection grid (GSB_ADV blocks: four per CU, one residency) gives up the tail's blocks, so
    // that no advection block waits for a free slot behind the others
    const int tb = adv_tail_blocks(x.maxsurf);
    const unsigned ax = std::max<int>(8, std::min<int>(ga.x, GSB_ADV - tb));
    hipLaunchKernelGGL(k_adv_tail, dim3(ax + tb, 1), dim3(256), 0, x.stream, x.dc, r, step_forces(1 - parity),
                       step_fin(x, 1 - parity, 2, lean, force_miss), tb);
  } else {
    AdvArgs r{S, S, B_PH, Hprev, Hnow, B_RHS, 0, x.sp, rhs};
    hipLaunchKernelGGL(k_adv_rhs, ga, dim3(256), 0, x.stream, x.dc, r);
  }
  Ctx xv = x;   // the velocity solve reads this step's right-hand side maxima
  xv.hval_u = rhs; xv.hval_v = rhs + 1;
  // Stages that do not read each other's output share a launch: the Helmholtz statistics with the
  // first projection's divergence, the first gate with the defect, the second gate with the
  // pressures and the ghost fill of the new state.
  Body hfin, gate1, gate2;
  if ((rc = enqueue_solve(xv, false, B_RHS, B_STAR, 0, vrefs, &hfin))) return rc;
  // the previous step's forces and clock complete before this step's first pressure solve (a
  // recording solve tags its pairs with the clock)
  if ((rc = enqueue_project(x, B_STAR, Snew, B_PHI, 1.0, 0, -1, prefs, &hfin, &gate1, tail == TAIL_PREV))) return rc;
  launch_bodies(x, {gate1, body_helm(x, HelmArgs{B_RHS, Snew, B_DEFECT, B_PROJ, 2, SV_FINAL, -1})});
  if ((rc = enqueue_project(x, B_PROJ, B_PROJ, B_PHI2, 1.0, 1, B_DEFECT, prefs, nullptr, &gate2))) return rc;
  launch_bodies(x, {gate2, body_press(x, PressArgs{B_PH, B_PHI2, B_PH, B_POUT}), body_fill(x, FillArgs{Snew, 1, 1, 1})});
  if (tail == TAIL_OWN) launch_tail(x, parity, step_fin(x, parity, nproj, lean, force_miss));
  return launch_check(x.err, "step");
}

}  // namespace

// Step graphs: one per parity, for steady AB2 steps (history present, unchanged dt).  Where the
// step's tail can be deferred (one case, direct pressure backend), exec[p]
// leaves its own tail pending and exec_prev[p] also runs the pending tail of the previous step
// beside its first stage; the caller launches the last pending tail itself (launch_tail).
struct StepGraphs {
  hipGraph_t graph[2] = {nullptr, nullptr}, graph_prev[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr}, exec_prev[2] = {nullptr, nullptr};
  bool valid = false;
  bool defer = false;   // exec leaves its tail pending
  int lean = 0;
  bool force_miss = false;
  void drop() {
    for (int p = 0; p < 2; ++p) {
      for (hipGraphExec_t* e : {&exec[p], &exec_prev[p]})
        if (*e) { (void)hipGraphExecDestroy(*e); *e = nullptr; }
      for (hipGraph_t* g : {&graph[p], &graph_prev[p]})
        if (*g) { (void)hipGraphDestroy(*g); *g = nullptr; }
    }
    valid = false;
    defer = false;
  }
  static int capture(const Ctx& x, int p, int lean, bool force_miss, int tail, hipGraph_t* graph, hipGraphExec_t* exec) {
    HIPCHK(x.err, hipStreamBeginCapture(x.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step_kernels(x, p, false, lean, force_miss, tail);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(x.stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(x.err, 4, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    *graph = g;
    HIPCHK(x.err, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    // the executable's one-time upload here rather than at its first launch
    HIPCHK(x.err, hipGraphUpload(*exec, x.stream));
    return 0;
  }
  int build(const Ctx& x, int lean_ = 0, bool force_miss_ = false) {
    drop();
    lean = lean_;
    force_miss = force_miss_;
    const bool d = tail_can_defer(x);
    for (int p = 0; p < 2; ++p) {
      if (int rc = capture(x, p, lean, force_miss, d ? TAIL_DEFER : TAIL_OWN, &graph[p], &exec[p])) return rc;
      if (d)
        if (int rc = capture(x, p, lean, force_miss, TAIL_PREV, &graph_prev[p], &exec_prev[p])) return rc;
    }
    defer = d;
    valid = true;
    return 0;
  }
};

namespace {

void advance_host_state(mcf_handle* h, double dt) {
  h->parity = 1 - h->parity;
  h->prev_dt = dt;
  h->has_history = true;
  h->time += dt;
}

void drop_graphs(mcf_handle* h) {
  if (h->graphs) h->graphs->drop();
  if (h->lean_graphs) h->lean_graphs->drop();
  for (StepGraphs* g : h->mixed)
    if (g) g->drop();
  if (h->kry) h->kry->drop_group();
  h->graph_dt = -1.0;
  h->prm_valid = false;
}

// n steps: steps without matching history run eagerly, the rest replay the captured graphs.
// The autotuner: times every launch shape of the transform and of the tridiagonal sweep on the
// context's own problems (forward transforms into the scratch F and in-place sweeps of F, which no
// later kernel reads before rewriting; no live buffer is touched) and keeps the fastest per
// problem class.  The staged transform tiles (0-2) give identical results, and so do the two split-K
// tiles of one block size; the families (staged, 8-wave split-K, 4-wave split-K, x-FFT) and the
// sweep widths differ in round-off, so the choice is kept with a checkpoint (mcf_get_launch /
// mcf_set_launch_classes) for a resumed run to round as the uninterrupted one.
int autotune(const Ctx& x, Launch& out) {
  hipEvent_t e0, e1;
  HIPCHK(x.err, hipEventCreate(&e0));
  HIPCHK(x.err, hipEventCreate(&e1));
  const int reps = 10;
  int rc = 0;
  for (int cls = 0; cls < 2 && !rc; ++cls) {
    const bool pres = cls == 0;
    const SolveDesc* D = pres ? x.dsp : x.dsh;
    const int count = pres ? x.n : 2 * x.n;
    const int yy = pres ? x.p_yy : x.h_yy, xx = pres ? x.p_xx : x.h_xx;
    const GemmArgs g{1, pres ? B_DIVP : B_RHS, 0, 0, -1, 0};
    const PcrArgs p0{0, -1, 0};
    const int np = pres ? x.p_np : x.h_np, r = pres ? x.p_r : x.h_r;
    float best = 1e30f, best_gemm = 1e30f;
    int bg = 0, bt = 1;
    // the dense y transforms and the x sweeps, then (where every problem of the class has it) the
    // x-FFT transform with its y sweep as one pair against the best dense pair
    const bool xf = pres ? !x.noxf_p : !x.noxf_h, par = pres ? !x.nopar_p : !x.nopar_h;
    int gv[N_GEMM_V + 2], ngv = 0;
    for (int k = 0; k < N_GEMM_V; ++k) gv[ngv++] = k;
    if (par) { gv[ngv++] = GV_PAR8; gv[ngv++] = GV_PAR4; }
    const int nv = ngv + N_TRI_V + (xf ? 1 : 0);
    for (int v = 0; v < nv && !rc; ++v) {
      const bool pair = v == ngv + N_TRI_V;
      const bool gemm = v < ngv;
      const int k = gemm ? gv[v] : v - ngv;
      if (v == ngv) { best_gemm = best; best = 1e30f; }
      auto launch = [&]() {
        if (pair) {
          launch_gemm(GV_XFFT, yy, xx, count, x.stream, D, g);
          launch_tri(TV_YSWEEP, yy, xx, count, x.stream, D, p0);
          if (r > 0) launch_points_cap(x, pres, GV_XFFT, D, count, xx, np, r, p0);
        } else if (gemm) {
          launch_gemm(k, yy, xx, count, x.stream, D, g);
        } else {
          launch_tri(k, yy, xx, count, x.stream, D, p0);
        }
      };
      for (int r = 0; r < 2; ++r) launch();
      (void)hipEventRecord(e0, x.stream);
      for (int r = 0; r < reps; ++r) launch();
      (void)hipEventRecord(e1, x.stream);
      if ((rc = launch_check(x.err, "autotune"))) break;
      float ms = 0.f;
      if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(x.err, 4, "autotune timing failed");
        break;
      }
      if (pair) {
        // against the best dense transform and sweep with their capacitance launches
        float dense = best_gemm + best;
        if (r > 0) {
          for (int q = 0; q < 2; ++q) launch_points_cap(x, pres, 0, D, count, xx, np, r, p0);
          (void)hipEventRecord(e0, x.stream);
          for (int q = 0; q < reps; ++q) launch_points_cap(x, pres, 0, D, count, xx, np, r, p0);
          (void)hipEventRecord(e1, x.stream);
          float mc = 0.f;
          if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&mc, e0, e1) != hipSuccess) {
            rc = fail(x.err, 4, "autotune timing failed");
            break;
          }
          dense += mc;
        }
        if (ms < dense) { bg = GV_XFFT; bt = TV_YSWEEP; }
      } else if (ms < best) {
        best = ms;
        if (gemm) bg = k; else bt = k;
      }
    }
    if (pres) { out.gemm_p = bg; out.tri_p = bt; } else { out.gemm_h = bg; out.tri_h = bt; }
  }
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  return rc;
}

int ensure_tuned(mcf_handle* h) {
  if (h->tuned) return 0;
  if (int rc = autotune(ctx_of(h), h->ln)) return rc;
  h->tuned = true;
  return 0;
}

StepGraphs* graphs_for(mcf_handle* h, int lean) {
  StepGraphs*& g = lean == LEAN_ALL ? h->lean_graphs : lean == 0 ? h->graphs : h->mixed[lean];
  if (!g) g = new StepGraphs();
  return g;
}

int enqueue_steps(mcf_handle* h, int n, double dt, int lean = 0) {
  if (int rc = ensure_tuned(h)) return rc;
  Ctx x = ctx_of(h);
  StepGraphs* gs = graphs_for(h, lean);
  // the parity of the step whose tail a deferring graph left pending (-1: none), launched before
  // an eager step and at the end of the call
  int pend = -1;
  auto flush = [&]() {
    if (pend >= 0) launch_tail(x, pend, step_fin(x, pend, 2, gs->lean, gs->force_miss));
    pend = -1;
  };
  // a speculating run switches between the SPEC_MASKS pairs as the transient comes and goes: all of
  // them are captured when the first is needed, so that no later chunk pays a capture
  const bool spec = h->spec > 0 && h->hs[0].dump_cap == 0 && h->pbackend == MCF_PB_DIRECT;
  hipLaunchKernelGGL(k_set_clock, dim3(1, 1), dim3(64), 0, h->stream, h->dc, h->time);
  for (int s = 0; s < n; ++s) {
    const bool steady = h->has_history && h->prev_dt == dt;
    const double alpha = !h->has_history ? 0.0 : dt / (2.0 * h->prev_dt);
    const double factor = !h->has_history ? 1.0 : dt / (dt + h->prev_dt);
    // the launches carry the parameters k_set_prm stores (a captured steady step: dt, 0.5, 0.5)
    x.sp = StepPrm{dt, alpha, factor};
    if (!(h->prm_valid && h->prm_dt == dt && h->prm_alpha == alpha && h->prm_factor == factor)) {
      hipLaunchKernelGGL(k_set_prm, dim3(1, 1), dim3(64), 0, h->stream, h->dc, dt, alpha, factor);
      h->prm_valid = true; h->prm_dt = dt; h->prm_alpha = alpha; h->prm_factor = factor;
    }
    // GMRES steps read the solver state back between launches: they run eagerly
    if (!steady || !h->use_graphs || x.gm) {
      flush();
      if (int rc = enqueue_step_kernels(x, h->parity, !h->has_history, x.gm ? 0 : lean, h->spec == 2)) return rc;
    } else {
      if (h->graph_dt != dt) drop_graphs(h);
      if (!gs->valid) {
        if (int rc = gs->build(x, lean, h->spec == 2)) return rc;
        for (int m : SPEC_MASKS) {
          StepGraphs* g = graphs_for(h, m);
          if (spec && !g->valid)
            if (int rc = g->build(x, m, h->spec == 2)) return rc;
        }
      }
      h->graph_dt = dt;
      HIPCHK(h, hipGraphLaunch(pend >= 0 ? gs->exec_prev[h->parity] : gs->exec[h->parity], h->stream));
      pend = gs->defer ? h->parity : -1;
    }
    advance_host_state(h, dt);
  }
  flush();
  return launch_check(h, "step tail");
}

// Step parameters for a one-off operation at an explicit dt (projection, operator, solve entry points)
void set_prm_once(mcf_handle* h, double dt) {
  hipLaunchKernelGGL(k_set_prm, dim3(1, 1), dim3(64), 0, h->stream, h->dc, dt, 0.0, 1.0);
  h->prm_valid = false;
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
// Scratch of mcf_rect_inverse_noalloc: its solve descriptor, its one-entry buffer table and the
// staged rectangle, carved out of one file-scope device array.  Each GPU holds its own copy until
// the process ends; the host reaches it only through the address hipGetSymbolAddress returns on
// that device (the array's name in host code is a host-side stand-in, not a device address).
constexpr size_t NA_DESC = 0;                      // SolveDesc, from word 0
constexpr size_t NA_TABLE = 128;                   // the buffer table's one pointer
constexpr size_t NA_FIELD = 256;                   // the staged rectangle, 2 KiB in
constexpr size_t NA_MAX_VALUES = (size_t)3 << 18;  // 786,432 values (6 MiB): up to an 800 x 820 rectangle
static_assert(sizeof(SolveDesc) <= (NA_TABLE - NA_DESC) * sizeof(double), "descriptor region too small");
static __device__ __align__(256) double scratch[NA_FIELD + NA_MAX_VALUES];

namespace {

struct HostTables {
  std::vector<int> lo[3], hi[3];
};

std::mutex na_mutex;
std::map<int, double*> na_address;   // device ordinal -> that device's `scratch`

// The current device's address of `scratch`, looked up once per device.
hipError_t na_scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(na_mutex);
  auto it = na_address.find(dev);
  if (it == na_address.end()) {
    double* p = nullptr;
    e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    it = na_address.emplace(dev, p).first;
  }
  *out = it->second;
  return hipSuccess;
}

std::map<const mcf_handle*, HostTables> g_host;

int field_rows(const mcf_handle* h, int field) { return field == MCF_FIELD_V ? h->ny + 1 : h->ny; }
int field_cols(const mcf_handle* h, int field) { return field == MCF_FIELD_U ? h->nx + 1 : h->nx; }

int ensure_diag(mcf_handle* h, int rows) {
  if (rows <= h->diag_cap) return 0;
  int cap = std::max(rows, 2 * h->diag_cap);
  if (h->diag) (void)hipFree(h->diag);
  h->diag = nullptr;
  HIPCHK(h, hipMalloc(&h->diag, (size_t)cap * MCF_DIAG_LEN * sizeof(double)));
  h->diag_cap = cap;
  h->hc.diag = h->diag;
  // kernels reach the ring through the device descriptor, so the captured step graphs stay valid:
  // only the descriptor is copied (no graph is rebuilt inside a long run's first call)
  HIPCHK(h, hipMemcpy(h->dc, &h->hc, sizeof(CaseDev), hipMemcpyHostToDevice));
  return 0;
}

// Clear every slot and set stripe 0 (the first NSLOT words) to `vals`
int upload_slots(mcf_handle* h, const std::vector<double>& vals) {
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipMemcpyAsync(h->slots, vals.data(), NSLOT * sizeof(double), hipMemcpyHostToDevice, h->stream));
  return 0;
}

// Host view of the slots (max |.| over stripes, sums for the sum slots, stripe 0 for the forces),
// reduced on the device so that only NSLOT values cross to the host.
int read_slots(mcf_handle* h, std::vector<double>& out) {
  hipLaunchKernelGGL(k_slot_values, dim3(1), dim3(NSLOT), 0, h->stream, h->dc, h->slot_vals);
  out.assign(NSLOT, 0.0);
  HIPCHK(h, hipMemcpyAsync(out.data(), h->slot_vals, NSLOT * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

// Slot values of pressure solve `psolve` (host view from read_slots)
SolveSlots slots_of(const std::vector<double>& sv, int psolve) {
  const int b = SL_P0 + 8 * psolve;
  return SolveSlots{sv[b + SV_VAL], sv[b + SV_RES1], sv[b + SV_RES2], sv[b + SV_FINAL], sv[b + SV_ROUND],
                    sv[b + SV_ASCALE], sv[b + SV_SSQR], sv[b + SV_SSQV], sv[SL_GM_IT + psolve],
                    sv[SL_GM_REF + psolve], sv[SL_GM_ST + psolve]};
}

int status_error(mcf_handle* h, const double* d, int step) {
  const int st = (int)d[MCF_D_STATUS];
  if (!st) return 0;
  char buf[512];
  if (st & 16) {
    std::snprintf(buf, sizeof buf, "Non-finite force at step %d (t=%g); results not saved", step, d[MCF_D_TIME]);
    return fail(h, 3, buf);
  }
  if (st & 1) std::snprintf(buf, sizeof buf, "CN momentum solve failed at step %d: residual=%g", step, d[MCF_D_MOMENTUM_RES]);
  else if (st & 2) std::snprintf(buf, sizeof buf, "Pressure solve failed at step %d: residual_linf=%g target=%g", step, d[MCF_D_P_RES0], d[MCF_D_P_TARGET0]);
  else if (st & 4) std::snprintf(buf, sizeof buf, "Ghost projection failed at step %d", step);
  else std::snprintf(buf, sizeof buf, "Ghost step failed divergence/CFL gate at step %d: div=%g cfl=%g", step, d[MCF_D_MAXDIV], d[MCF_D_CFL]);
  return fail(h, 2, buf);
}

// The x-FFT form of a rectangle (pressure: DCT-IV, v: DST-IV, u: mirrored-outlet sine basis; xx
// even): the radix plan of the
// complex FFT of length M = xx / 2 (radix 4 first, then the primes up to XF_MAXR), its twiddles, the
// pre- and post-twiddles of the DCT-IV, and per x mode m the LU factors of the y system
// T_m = (a + b mu_m) I + b ly, mu_m = -4 sin^2((m + 1/2) pi / (2 xx)) / dx^2.  T_m or -T_m is an
// M-matrix whose row sums are sums of terms of one sign, so its pivots are formed from the row
// sums (g_0 = s_0, g_j = s_j + c g_{j-1} / d_{j-1}, d_j = g_j + c_j) without the cancellation the
// plain recurrence suffers when both ends are Neumann (ly singular).  Leaves xf = 0 where the form
// does not apply.
int build_xfft(mcf_handle* h, int kind) {
  SolveDesc& s = h->hs[kind];
  s.xf = 0;
  const int yy = s.yy, xx = s.xx;
  if ((xx & 1) || xx / 2 > XF_MAXM || yy > YS_MAXY || yy < 2) return 0;
  const int M = xx / 2;
  std::vector<int> rad;
  int m = M;
  for (int p = 3; m > 1 && p <= XF_MAXR; p += 2)
    while (m % p == 0) { rad.push_back(p); m /= p; }
  if (m != 1 && (m & (m - 1))) return 0;   // a prime factor above XF_MAXR
  // the largest odd radix first (its twiddles are all 1), then the powers of two as radix 4 (and 2)
  std::sort(rad.begin(), rad.end(), std::greater<int>());
  while (m % 4 == 0) { rad.push_back(4); m /= 4; }
  if (m == 2) rad.push_back(2);
  std::vector<int> st;
  std::vector<double> tw, root;
  auto unit = [](std::vector<double>& v, double turns) {   // exp(-2 pi i turns)
    v.push_back(std::cos(2.0 * M_PI * turns));
    v.push_back(-std::sin(2.0 * M_PI * turns));
  };
  long L = 1;
  for (int r : rad) {
    st.push_back(r);
    st.push_back((int)(tw.size() / 2));
    st.push_back((int)(root.size() / 2));   // shifted past the twiddles below
    for (int q = 1; q < r; ++q)
      for (long j = 0; j < L; ++j) unit(tw, (double)(q * j) / (double)(r * L));
    for (int e = 0; e < r; ++e) unit(root, (double)e / r);
    L *= r;
  }
  const int ntw = (int)(tw.size() / 2);
  for (size_t q = 2; q < st.size(); q += 3) st[q] += ntw;
  tw.insert(tw.end(), root.begin(), root.end());
  const int ftn = (int)(tw.size() / 2);
  if ((int)rad.size() > XF_MAXS || ftn > xdct_tables(M) || xdct_lds(M, xdct_tables(M)) > XF_MAXLDS) return 0;
  std::vector<double> pre, post(2 * M);
  if (kind != MCF_U) {
    // DCT-IV / DST-IV: exp(-i pi (4n+1) / (4N)) and sqrt(2/N) exp(-i pi k / N), n, k < M
    pre.resize(2 * M);
    const double scale = std::sqrt(2.0 / xx);
    for (int n = 0; n < M; ++n) {
      const double a = M_PI * (4.0 * n + 1.0) / (4.0 * xx), b = M_PI * n / xx;
      pre[2 * n] = std::cos(a); pre[2 * n + 1] = -std::sin(a);
      post[2 * n] = scale * std::cos(b); post[2 * n + 1] = -scale * std::sin(b);
    }
  } else {
    // u: exp(-i pi k / (2N)) for k < N and exp(-2 pi i k / N) for k < M
    pre.resize(2 * xx);
    for (int k = 0; k < xx; ++k) {
      const double a = M_PI * k / (2.0 * xx);
      pre[2 * k] = std::cos(a); pre[2 * k + 1] = -std::sin(a);
    }
    for (int k = 0; k < M; ++k) {
      const double b = 2.0 * M_PI * k / xx;
      post[2 * k] = std::cos(b); post[2 * k + 1] = -std::sin(b);
    }
  }
  // y systems: ly = tridiag(1, diagy, 1) / dy^2 with the kind's wall ends (lower, upper)
  const double ends[3][2] = {{1.0, 1.0}, {3.0, 3.0}, {2.0, 2.0}};
  const double dy2 = h->dy * h->dy, dx2 = h->dx * h->dx;
  const double a = s.a, b = s.b, sgn = b > 0 ? -1.0 : 1.0, c = std::fabs(b) / dy2;
  std::vector<double> yl((size_t)yy * xx), yi((size_t)yy * xx);
  for (int mm = 0; mm < xx; ++mm) {
    const double sx = std::sin((mm + 0.5) * M_PI / (2.0 * xx));
    const double mu = -4.0 * sx * sx / dx2;
    double g = 0.0, dprev = 0.0;
    for (int j = 0; j < yy; ++j) {
      double end = 0.0;
      if (j == 0) end += (1.0 - ends[kind][0]) / dy2;
      if (j == yy - 1) end += (1.0 - ends[kind][1]) / dy2;
      const double sj = sgn * a + sgn * b * mu + sgn * b * end;   // each term >= 0
      g = j == 0 ? sj : sj + c * g / dprev;
      const double dA = g + (j < yy - 1 ? c : 0.0);
      const double dT = sgn * dA;
      yl[(size_t)j * xx + mm] = j == 0 ? 0.0 : (b / dy2) / (sgn * dprev);
      yi[(size_t)j * xx + mm] = 1.0 / dT;
      dprev = dA;
    }
  }
  const std::string key = "xf" + std::to_string(kind);
  int* dst;
  double *dtw, *dpre, *dpost, *dyl, *dyi;
  if (int rc = upload(h, key + "st", st.data(), st.size(), &dst)) return rc;
  if (int rc = upload(h, key + "tw", tw.data(), tw.size(), &dtw)) return rc;
  if (int rc = upload(h, key + "pre", pre.data(), pre.size(), &dpre)) return rc;
  if (int rc = upload(h, key + "post", post.data(), post.size(), &dpost)) return rc;
  if (int rc = upload(h, key + "yl", yl.data(), yl.size(), &dyl)) return rc;
  if (int rc = upload(h, key + "yi", yi.data(), yi.size(), &dyi)) return rc;
  s.fm = M; s.fns = (int)rad.size(); s.ftn = ftn; s.fst = dst; s.ftw = dtw; s.fpre = dpre; s.fpost = dpost;
  s.yl = dyl; s.yi = dyi; s.ye = b / dy2;
  s.xf = kind == MCF_PRESSURE ? 1 : kind == MCF_V ? 2 : 3;
  return 0;
}

// The x basis of an x-FFT problem at column x and mode m: the inverse transform's coefficient
// W[x][m] (x_i = sum_m W[i][m] H_m), or with `forward` the forward one (H_m = sum_i Wf[i][m] x_i).
// DCT-IV / DST-IV are orthonormal and symmetric (Wf = W); for u, W[i][m] = sin(pi (2m+1)(i+1) / (2N))
// and Wf[i][m] = (2/N) W[i][m], halved at the outlet face i = N-1.
double xbasis(int xf, int xx, int x, int m, bool forward) {
  if (xf == 3) {
    const long long P = ((long long)(2 * m + 1) * (x + 1)) % (4LL * xx);
    const double w = std::sin(M_PI * (double)P / (2.0 * xx));
    return forward ? (2.0 / xx) * (x == xx - 1 ? 0.5 : 1.0) * w : w;
  }
  const long long P = ((long long)(2 * x + 1) * (2 * m + 1)) % (8LL * xx);
  const double t = M_PI * (double)P / (4.0 * xx);
  return std::sqrt(2.0 / xx) * (xf == 1 ? std::cos(t) : std::sin(t));
}

// The capacitance tables of the x-FFT form: the basis at the points' distinct x columns and at
// the capacitance row groups' columns, and the entries (y index, column group) of each row.
int build_xfft_schur(mcf_handle* h, int kind, const std::vector<int>& pt_x, const std::vector<int>& rg_x,
                     const std::vector<int>& rg_ptr, const std::vector<int>& rc_j) {
  SolveDesc& s = h->hs[kind];
  if (!s.xf) return 0;
  const int xx = s.xx, yy = s.yy, npts = (int)pt_x.size(), r = (int)rc_j.size();
  std::vector<int> pc, pt_c(std::max(npts, 1), 0);
  for (int t = 0; t < npts; ++t) {
    if (pc.empty() || pc.back() != pt_x[t]) pc.push_back(pt_x[t]);
    pt_c[t] = (int)pc.size() - 1;
  }
  std::vector<double> wp(std::max<size_t>(pc.size() * xx, 1)), wf(std::max<size_t>(rg_x.size() * xx, 1));
  for (size_t c = 0; c < pc.size(); ++c)
    for (int m = 0; m < xx; ++m) wp[c * xx + m] = xbasis(s.xf, xx, pc[c], m, false);
  for (size_t g = 0; g < rg_x.size(); ++g)
    for (int m = 0; m < xx; ++m) wf[g * xx + m] = xbasis(s.xf, xx, rg_x[g], m, true);
  std::vector<int> cnt(yy + 1, 0), grp(std::max(r, 1), 0);
  for (size_t g = 0; g + 1 < rg_ptr.size(); ++g)
    for (int t = rg_ptr[g]; t < rg_ptr[g + 1]; ++t) grp[t] = (int)g;
  for (int t = 0; t < r; ++t) cnt[rc_j[t] + 1]++;
  for (int j = 0; j < yy; ++j) cnt[j + 1] += cnt[j];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1), et(std::max(r, 1), 0), eg(std::max(r, 1), 0);
  for (int t = 0; t < r; ++t) {
    const int e = fill[rc_j[t]]++;
    et[e] = t; eg[e] = grp[t];
  }
  int lo = yy, hi = -1, most = 0;
  for (int j = 0; j < yy; ++j)
    if (cnt[j + 1] > cnt[j]) { lo = std::min(lo, j); hi = std::max(hi, j); most = std::max(most, cnt[j + 1] - cnt[j]); }
  if (most > XC_MAXE) { s.xf = 0; return 0; }
  const int nband = hi >= lo ? hi - lo + 1 : 0;
  const std::string key = "xs" + std::to_string(kind);
  int *dpc, *drp, *det, *deg;
  double *dwp, *dwf, *dcb;
  if (int rc = upload<double>(h, key + "cb", nullptr, (size_t)std::max(nband, 1) * xx, &dcb)) return rc;
  s.yr_lo = nband ? lo : 0; s.yr_n = nband; s.xcb = dcb;
  if (int rc = upload(h, key + "pc", pt_c.data(), pt_c.size(), &dpc)) return rc;
  if (int rc = upload(h, key + "wp", wp.data(), wp.size(), &dwp)) return rc;
  if (int rc = upload(h, key + "wf", wf.data(), wf.size(), &dwf)) return rc;
  if (int rc = upload(h, key + "rp", cnt.data(), cnt.size(), &drp)) return rc;
  if (int rc = upload(h, key + "et", et.data(), et.size(), &det)) return rc;
  if (int rc = upload(h, key + "eg", eg.data(), eg.size(), &deg)) return rc;
  s.pt_c = dpc; s.xwp = dwp; s.xwf = dwf; s.yr_ptr = drp; s.yr_t = det; s.yr_g = deg;
  return 0;
}

// The rectangle inverse of kind on host rhs -> host out, staged in `field` (yy*xx values) with its
// one-entry buffer table `tb` and descriptor `td` in device memory the caller provides.  Copies and
// launches are ordered on the handle's stream, and the call returns after it drains.
int rect_inverse_on(mcf_handle* h, int kind, const double* rhs, double* out, double* field, double** tb,
                    SolveDesc* td) {
  SolveDesc t = h->hs[kind];
  const size_t n = (size_t)t.yy * t.xx;
  t.bufs = tb; t.off = 0; t.ld = t.xx; t.lo = nullptr; t.hi = nullptr;
  HIPCHK(h, hipMemcpyAsync(tb, &field, sizeof(double*), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(td, &t, sizeof(SolveDesc), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(field, rhs, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  int gv = kind == MCF_PRESSURE ? h->ln.gemm_p : h->ln.gemm_h;
  const int tv = kind == MCF_PRESSURE ? h->ln.tri_p : h->ln.tri_h;
  if (gv_par(gv) && !t.par) gv = 4;
  GemmArgs g{1, 0, 0, 0, -1, 0};
  launch_gemm(gv, t.yy, t.xx, 1, h->stream, td, g);
  PcrArgs p0{0, -1, 0};
  launch_tri(tv, t.yy, t.xx, 1, h->stream, td, p0);
  GemmArgs gi{0, 0, 0, 0, -1, 0};
  launch_gemm(gv, t.yy, t.xx, 1, h->stream, td, gi);
  int rc = launch_check(h, "rect inverse");
  HIPCHK(h, hipMemcpyAsync(out, field, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return rc;
}

}  // namespace

extern "C" {

int mcf_create(int nx, int ny, double dx, double dy, double nu, double rho, mcf_handle** out) {
  if (!out) return 1;
  *out = nullptr;
  if (nx < 4 || ny < 4 || nx > 4096 || !(dx > 0) || !(dy > 0) || !(nu > 0) || !(rho > 0)) return 1;
  mcf_handle* h = new mcf_handle();
  h->nx = nx; h->ny = ny; h->dx = dx; h->dy = dy; h->nu = nu; h->rho = rho;
  h->usize = ny * (nx + 1); h->vsize = (ny + 1) * nx; h->psize = ny * nx;
  h->velsize = h->usize + h->vsize;
  *out = h;
  g_host[h] = HostTables{};
  HIPCHK(h, hipStreamCreate(&h->stream));
  HIPCHK(h, na_scratch_address(&h->na_scratch));
  for (int b = 0; b < NBUF; ++b) {
    if (b == B_KV0) { h->hc.buf[b] = nullptr; continue; }   // set with the GMRES workspace
    void* p = nullptr;
    size_t n = b < kFirstPBuf ? h->velsize : h->psize;
    HIPCHK(h, hipMalloc(&p, n * sizeof(double)));
    HIPCHK(h, hipMemset(p, 0, n * sizeof(double)));
    h->bufs.push_back(p);
    h->hc.buf[b] = static_cast<double*>(p);
  }
  HIPCHK(h, hipMalloc(&h->slots, NSLOTW * sizeof(u64)));
  HIPCHK(h, hipMemset(h->slots, 0, NSLOTW * sizeof(u64)));
  HIPCHK(h, hipMalloc(&h->slot_vals, NSLOT * sizeof(double)));
  HIPCHK(h, hipMalloc(&h->inlet, ny * sizeof(double)));
  HIPCHK(h, hipMemset(h->inlet, 0, ny * sizeof(double)));
  HIPCHK(h, hipMalloc(&h->clock_step, sizeof(int)));
  HIPCHK(h, hipMalloc(&h->clock_time, sizeof(double)));
  HIPCHK(h, hipMemset(h->clock_step, 0, sizeof(int)));
  HIPCHK(h, hipMemset(h->clock_time, 0, sizeof(double)));
  HIPCHK(h, hipMalloc(&h->prm, 4 * sizeof(double)));
  HIPCHK(h, hipMemset(h->prm, 0, 4 * sizeof(double)));
  h->graphs = new StepGraphs();
  h->lean_graphs = new StepGraphs();
  HIPCHK(h, hipMalloc(&h->dc, sizeof(CaseDev)));
  HIPCHK(h, hipMalloc(&h->ds, sizeof(h->hs)));
  CaseDev& c = h->hc;
  c.nx = nx; c.ny = ny; c.usize = h->usize; c.vsize = h->vsize; c.psize = h->psize; c.velsize = h->velsize;
  c.dx = dx; c.dy = dy; c.nu = nu; c.rho = rho;
  c.inlet = h->inlet;
  c.sl = h->slots;
  c.clock_step = h->clock_step;
  c.clock_time = h->clock_time;
  c.prm = h->prm;
  if (const char* e = std::getenv("GHOSTFLOW_NO_GRAPH")) h->use_graphs = (e[0] == '0');
  for (int k = 0; k < 3; ++k) {
    SolveDesc& s = h->hs[k];
    s.kind = k;
    s.bufs = reinterpret_cast<double* const*>(reinterpret_cast<char*>(h->dc) + offsetof(CaseDev, buf));
    s.sl = h->slots;
    s.slot0 = sv_block(k, 0);
  }
  if (int rc = ensure_diag(h, 64)) return rc;
  return sync_descs(h);
}

void mcf_destroy(mcf_handle* h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(h);
  delete h->graphs;
  delete h->lean_graphs;
  for (StepGraphs*& g : h->mixed) { delete g; g = nullptr; }
  h->graphs = nullptr;
  h->lean_graphs = nullptr;
  (void)hipFree(h->clock_step); (void)hipFree(h->clock_time); (void)hipFree(h->prm);
  for (auto& kv : h->owned) (void)hipFree(kv.second);
  for (void* p : h->bufs) (void)hipFree(p);
  (void)hipFree(h->slots); (void)hipFree(h->slot_vals); (void)hipFree(h->inlet); (void)hipFree(h->dc); (void)hipFree(h->ds);
  if (h->diag) (void)hipFree(h->diag);
  if (h->tmp_rect) (void)hipFree(h->tmp_rect);
  if (h->kry) {
    h->kry->drop_group();
    (void)hipFree(h->kry->ctl); (void)hipFree(h->kry->V); (void)hipFree(h->kry->part);
    (void)hipHostFree(h->kry->stat);
    delete h->kry;
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  g_host.erase(h);
  delete h;
}

const char* mcf_last_error(const mcf_handle* h) { return h ? h->err.c_str() : "null handle"; }

int mcf_device_name(char* buf, int len) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 4;
  std::snprintf(buf, len, "%s %s", prop.name, prop.gcnArchName);
  return 0;
}

int mcf_set_inlet(mcf_handle* h, const double* u_inlet) {
  if (!h || !u_inlet) return 1;
  HIPCHK(h, hipMemcpy(h->inlet, u_inlet, h->ny * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int mcf_set_solid_intervals(mcf_handle* h, int field, const int* lo, const int* hi) {
  if (!h || field < 0 || field > 2 || !lo || !hi) return fail(h, 1, "invalid solid intervals");
  const int rows = field_rows(h, field), cols = field_cols(h, field);
  for (int j = 0; j < rows; ++j)
    if (lo[j] < 0 || hi[j] >= cols) return fail(h, 1, "solid interval outside the array");
  int *dlo, *dhi, *dlh;
  std::string k = "solid" + std::to_string(field);
  std::vector<int> lh((size_t)2 * rows);
  for (int j = 0; j < rows; ++j) { lh[2 * j] = lo[j]; lh[2 * j + 1] = hi[j]; }
  if (int rc = upload(h, k + "lo", lo, rows, &dlo)) return rc;
  if (int rc = upload(h, k + "hi", hi, rows, &dhi)) return rc;
  if (int rc = upload(h, k + "lh", lh.data(), lh.size(), &dlh)) return rc;
  h->hc.lo[field] = dlo; h->hc.hi[field] = dhi;
  h->hc.lohi[field] = reinterpret_cast<const IPair*>(dlh);
  int j0 = rows, j1 = -1, i0 = cols, i1 = -1;
  for (int j = 0; j < rows; ++j)
    if (lo[j] <= hi[j]) { j0 = std::min(j0, j); j1 = std::max(j1, j); i0 = std::min(i0, lo[j]); i1 = std::max(i1, hi[j]); }
  h->hc.sbox[field][0] = j0; h->hc.sbox[field][1] = j1; h->hc.sbox[field][2] = i0; h->hc.sbox[field][3] = i1;
  g_host[h].lo[field].assign(lo, lo + rows);
  g_host[h].hi[field].assign(hi, hi + rows);
  for (int q = 0; q < 3; ++q) {
    if (h->hs[q].field == field && h->has_rect[q]) { h->hs[q].lo = dlo; h->hs[q].hi = dhi; }
  }
  h->has_solid[field] = true;
  return sync_descs(h);
}

int mcf_set_ghost(mcf_handle* h, int field, int n, const int* node, const int* ptr, const int* col, const double* w) {
  if (!h || field < 0 || field > 2 || n < 0 || (n && (!node || !ptr || !col || !w))) return fail(h, 1, "invalid ghost table");
  const int size = field == MCF_FIELD_U ? h->usize : (field == MCF_FIELD_V ? h->vsize : h->psize);
  const int nnz = n ? ptr[n] : 0;
  for (int k = 0; k < n; ++k) {
    if (node[k] < 0 || node[k] >= size || ptr[k] > ptr[k + 1]) return fail(h, 1, "ghost node out of range");
  }
  for (int k = 0; k < nnz; ++k) if (col[k] < 0 || col[k] >= size) return fail(h, 1, "ghost column out of range");
  GhostTab& g = h->hc.gt[field];
  std::string key = "ghost" + std::to_string(field);
  int width = 1;
  for (int k = 0; k < n; ++k) width = std::max(width, ptr[k + 1] - ptr[k]);
  std::vector<int> pc((size_t)n * width);
  std::vector<double> pw((size_t)n * width, 0.0);
  for (int k = 0; k < n; ++k)
    for (int q = 0; q < width; ++q) {
      const int t = ptr[k] + q;
      const bool real = t < ptr[k + 1];
      pc[(size_t)k * width + q] = real ? col[t] : col[ptr[k]];
      pw[(size_t)k * width + q] = real ? w[t] : 0.0;
    }
  int *dn, *dcol; double* dw;
  if (int rc = upload(h, key + "n", node, n, &dn)) return rc;
  if (int rc = upload(h, key + "c", pc.data(), pc.size(), &dcol)) return rc;
  if (int rc = upload(h, key + "w", pw.data(), pw.size(), &dw)) return rc;
  g.n = n; g.width = width; g.node = dn; g.col = dcol; g.w = dw;
  h->has_ghost[field] = true;
  return sync_descs(h);
}

int mcf_set_override(mcf_handle* h, int op, int n, const int* row, const int* ptr, const int* col, const double* val) {
  if (!h || op < 0 || op >= MCF_N_OPS || n < 0) return fail(h, 1, "invalid override");
  int rows_field, in_size;
  switch (op) {
    case MCF_OP_GRAD_U: rows_field = MCF_FIELD_U; in_size = h->psize; break;
    case MCF_OP_GRAD_V: rows_field = MCF_FIELD_V; in_size = h->psize; break;
    case MCF_OP_DIV: rows_field = MCF_FIELD_P; in_size = h->velsize; break;
    case MCF_OP_HELM_U: rows_field = MCF_FIELD_U; in_size = h->velsize; break;
    case MCF_OP_HELM_V: rows_field = MCF_FIELD_V; in_size = h->velsize; break;
    default: rows_field = MCF_FIELD_P; in_size = h->psize; break;
  }
  Override& o = h->hc.ov[op];
  o = Override{0, 0, 0, 0, nullptr, nullptr, nullptr};
  if (n == 0) return sync_descs(h);
  const int cols = field_cols(h, rows_field), rows = field_rows(h, rows_field);
  int j0 = 1 << 30, j1 = -1, i0 = 1 << 30, i1 = -1;
  for (int k = 0; k < n; ++k) {
    if (row[k] < 0 || row[k] >= rows * cols || ptr[k] > ptr[k + 1]) return fail(h, 1, "override row out of range");
    const int j = row[k] / cols, i = row[k] % cols;
    j0 = std::min(j0, j); j1 = std::max(j1, j); i0 = std::min(i0, i); i1 = std::max(i1, i);
  }
  const int nnz = ptr[n];
  for (int k = 0; k < nnz; ++k) if (col[k] < 0 || col[k] >= in_size) return fail(h, 1, "override column out of range");
  const int bh = j1 - j0 + 1, bw = i1 - i0 + 1;
  // the map holds each override row's CSR range (k0, k1); (-1, -1) where the stencil row stands
  std::vector<int> map((size_t)2 * bh * bw, -1);
  for (int k = 0; k < n; ++k) {
    const int j = row[k] / cols, i = row[k] % cols;
    const size_t e = (size_t)(j - j0) * bw + (i - i0);
    map[2 * e] = ptr[k];
    map[2 * e + 1] = ptr[k + 1];
  }
  std::string key = "ov" + std::to_string(op);
  int *dm, *dcol; double* dv;
  if (int rc = upload(h, key + "m", map.data(), map.size(), &dm)) return rc;
  if (int rc = upload(h, key + "c", col, nnz, &dcol)) return rc;
  if (int rc = upload(h, key + "v", val, nnz, &dv)) return rc;
  o = Override{j0, i0, bh, bw, reinterpret_cast<const IPair*>(dm), dcol, dv};
  return sync_descs(h);
}

int mcf_set_rectangle(mcf_handle* h, int kind, int yy, int xx, double a, double b, const double* V,
                      const double* lam, const double* lower, const double* diag, const double* upper) {
  if (!h || kind < 0 || kind > 2 || !V || !lam || !lower || !diag || !upper) return fail(h, 1, "invalid rectangle");
  const int wy = kind == MCF_V ? h->ny - 1 : h->ny;
  if (yy != wy || xx != h->nx) return fail(h, 1, "rectangle shape does not match the grid");
  if (xx > TRI_MAXX) return fail(h, 1, "rectangle wider than the tridiagonal kernel supports");
  for (int i = 1; i < xx - 1; ++i)
    if (upper[i] != upper[0]) return fail(h, 1, "the x super-diagonal must be constant");
  SolveDesc& s = h->hs[kind];
  std::string key = "rect" + std::to_string(kind);
  std::vector<double> VT((size_t)yy * yy);
  for (int j = 0; j < yy; ++j)
    for (int k = 0; k < yy; ++k) VT[(size_t)k * yy + j] = V[(size_t)j * yy + k];
  double *dV, *dVT, *dl, *dlo, *ddg, *dup, *dF;
  if (int rc = upload(h, key + "V", V, (size_t)yy * yy, &dV)) return rc;
  if (int rc = upload(h, key + "VT", VT.data(), VT.size(), &dVT)) return rc;
  if (int rc = upload(h, key + "lam", lam, yy, &dl)) return rc;
  if (int rc = upload(h, key + "lo", lower, xx - 1, &dlo)) return rc;
  if (int rc = upload(h, key + "dg", diag, xx, &ddg)) return rc;
  if (int rc = upload(h, key + "up", upper, xx - 1, &dup)) return rc;
  if (int rc = upload<double>(h, key + "F", nullptr, (size_t)yy * xx, &dF)) return rc;
  // LU (no pivoting; the x systems are diagonally dominant) of (a + b lam_k) I + b lx, per mode
  std::vector<double> Lm((size_t)yy * xx), Im((size_t)yy * xx);
  for (int k = 0; k < yy; ++k) {
    const double shift = a + b * lam[k];
    double dprev = shift + b * diag[0];
    Lm[(size_t)k * xx] = 0.0;
    Im[(size_t)k * xx] = 1.0 / dprev;
    for (int i = 1; i < xx; ++i) {
      const double l = (b * lower[i - 1]) / dprev;
      const double di = shift + b * diag[i] - l * (b * upper[i - 1]);
      Lm[(size_t)k * xx + i] = l;
      Im[(size_t)k * xx + i] = 1.0 / di;
      dprev = di;
    }
  }
  double *dL, *dI;
  if (int rc = upload(h, key + "luL", Lm.data(), Lm.size(), &dL)) return rc;
  if (int rc = upload(h, key + "luI", Im.data(), Im.size(), &dI)) return rc;
  s.luL = dL; s.luI = dI;
  // the reflection parity of the basis, V[yy-1-j][m] = (-1)^m V[j][m] (DCT-II, DST-II and DST-I
  // columns have it), to round-off; the parity-split transform is offered only where it holds
  double vmax = 0.0, dev = 0.0;
  for (int j = 0; j < yy; ++j)
    for (int m = 0; m < yy; ++m) {
      const double v = V[(size_t)j * yy + m], r = V[(size_t)(yy - 1 - j) * yy + m];
      vmax = std::max(vmax, std::fabs(v));
      dev = std::max(dev, std::fabs(r - ((m & 1) ? -v : v)));
    }
  s.par = dev <= 1e-12 * vmax ? 1 : 0;
  s.yy = yy; s.xx = xx; s.a = a; s.b = b;
  s.V = dV; s.VT = dVT; s.lam = dl; s.lower = dlo; s.diag = ddg; s.upper = dup; s.F = dF;
  if (h->has_schur[kind])
    if (int rc = build_vsel(h, kind)) return rc;
  if (kind == MCF_PRESSURE) { s.off = 0; s.ld = h->nx; s.joff = 0; s.ioff = 0; s.field = MCF_FIELD_P; }
  else if (kind == MCF_U) { s.off = 1; s.ld = h->nx + 1; s.joff = 0; s.ioff = 1; s.field = MCF_FIELD_U; }
  else { s.off = h->usize + h->nx; s.ld = h->nx; s.joff = 1; s.ioff = 0; s.field = MCF_FIELD_V; }
  s.lo = h->hc.lo[s.field]; s.hi = h->hc.hi[s.field];
  if (int rc = build_xfft(h, kind)) return rc;
  h->has_rect[kind] = true;
  return sync_descs(h);
}

int mcf_set_schur(mcf_handle* h, int kind, int r, const int* rows, int npts, const int* points,
                  const double* SC) {
  if (!h || kind < 0 || kind > 2 || r < 0 || npts < 0 || (r && (!rows || !SC)) || (npts && !points))
    return fail(h, 1, "invalid capacitance data");
  SolveDesc& s = h->hs[kind];
  if (!h->has_rect[kind]) return fail(h, 1, "set the rectangle before its capacitance data");
  if (r > 2048) return fail(h, 1, "Ghost inverse exceeds the bounded boundary-row rank");
  const int yy = s.yy, xx = s.xx, n = yy * xx;
  for (int k = 0; k < r; ++k) if (rows[k] < 0 || rows[k] >= n) return fail(h, 1, "capacitance row out of range");
  for (int k = 0; k < npts; ++k) if (points[k] < 0 || points[k] >= n) return fail(h, 1, "capacitance point out of range");
  // points ordered by (x column, y)
  std::vector<int> order(npts);
  for (int k = 0; k < npts; ++k) order[k] = k;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    int xa = points[a] % xx, xb = points[b] % xx;
    return xa != xb ? xa < xb : points[a] < points[b];
  });
  std::vector<int> pos(npts), pt_j(npts), pt_x(npts);
  for (int t = 0; t < npts; ++t) {
    const int k = order[t];
    pos[k] = t;
    pt_j[t] = points[k] / xx;
    pt_x[t] = points[k] % xx;
  }
  // capacitance rows grouped by x column
  std::vector<int> cnt(xx, 0);
  for (int k = 0; k < r; ++k) cnt[rows[k] % xx]++;
  std::vector<int> rg_ptr, rg_x, start(xx, -1);
  int acc = 0;
  for (int x = 0; x < xx; ++x)
    if (cnt[x]) { rg_x.push_back(x); rg_ptr.push_back(acc); start[x] = acc; acc += cnt[x]; }
  rg_ptr.push_back(acc);
  if ((int)rg_x.size() > TRI_MAXG)
    return fail(h, 1, "capacitance rows lie on more than " + std::to_string(TRI_MAXG) + " x columns");
  // row k of the caller's S^{-1} C becomes row gpos[k] (group order), its columns in point order
  std::vector<int> rc_j(r), gpos(r);
  for (int k = 0; k < r; ++k) {
    const int x = rows[k] % xx, t = start[x]++;
    rc_j[t] = rows[k] / xx; gpos[k] = t;
  }
  std::vector<double> scg((size_t)r * npts);
  for (int m = 0; m < r; ++m)
    for (int k = 0; k < npts; ++k) scg[(size_t)gpos[m] * npts + pos[k]] = SC[(size_t)m * npts + k];
  std::string key = "schur" + std::to_string(kind);
  int *dpj, *dpx, *drp, *drx, *drj, *dgp; double *dS, *dfk, *dy;
  if (int rc = upload(h, key + "pj", pt_j.data(), pt_j.size(), &dpj)) return rc;
  if (int rc = upload(h, key + "px", pt_x.data(), pt_x.size(), &dpx)) return rc;
  if (int rc = upload(h, key + "SC", scg.data(), scg.size(), &dS)) return rc;
  if (int rc = upload(h, key + "rp", rg_ptr.data(), rg_ptr.size(), &drp)) return rc;
  if (int rc = upload(h, key + "rx", rg_x.data(), rg_x.size(), &drx)) return rc;
  if (int rc = upload(h, key + "rj", rc_j.data(), rc_j.size(), &drj)) return rc;
  if (int rc = upload(h, key + "gp", gpos.data(), gpos.size(), &dgp)) return rc;
  if (int rc = upload<double>(h, key + "fk", nullptr, std::max(npts, 1), &dfk)) return rc;
  if (int rc = upload<double>(h, key + "y", nullptr, std::max(r, 1), &dy)) return rc;
  s.r = r; s.npts = npts;
  s.pt_j = dpj; s.pt_x = dpx; s.firstK = dfk; s.SC = dS; s.y = dy;
  s.rg_n = (int)rg_x.size(); s.rg_ptr = drp; s.rg_x = drx; s.gpos = dgp;
  h->rc_j[kind] = drj;
  if (int rc = build_vsel(h, kind)) return rc;
  if (int rc = build_xfft_schur(h, kind, pt_x, rg_x, rg_ptr, rc_j)) return rc;
  // a new capacitance set invalidates the Schur system and any pair dump built on the old one
  s.Sth = nullptr; s.Cptr = nullptr; s.Ccol = nullptr; s.Cval = nullptr; s.apos = nullptr;
  s.dump_c = s.dump_y = nullptr; s.dump_tag = s.dump_count = nullptr; s.dump_cap = 0;
  h->pt_pos[kind] = pos;
  h->has_schur[kind] = true;
  return sync_descs(h);
}

int mcf_set_surface(mcf_handle* h, int n, const int* pp, const int* pc, const double* pv,
                    const int* up, const int* uc, const double* uv, const int* vp, const int* vc,
                    const double* vv, const double* normals, const double* normal_weights,
                    double diameter, double u_ref) {
  if (!h || n <= 0) return fail(h, 1, "invalid surface quadrature");
  const int np = pp[n], nu = up[n], nv = vp[n];
  for (int k = 0; k < np; ++k) if (pc[k] < 0 || pc[k] >= h->psize) return fail(h, 1, "surface p column out of range");
  for (int k = 0; k < nu; ++k) if (uc[k] < 0 || uc[k] >= h->usize) return fail(h, 1, "surface u column out of range");
  for (int k = 0; k < nv; ++k) if (vc[k] < 0 || vc[k] >= h->vsize) return fail(h, 1, "surface v column out of range");
  CaseDev& c = h->hc;
  int *a, *b, *d, *e, *f, *g; double *x, *y, *z, *nr, *nw;
  // k_forces reads entry 0 of every column and value array for the nonzeros a row does not have,
  // so an empty array is stored as one zero entry
  auto upload_nz = [h](const std::string& key, auto* src, int cnt, auto** dst) {
    return cnt > 0 ? upload(h, key, src, (size_t)cnt, dst) : upload(h, key, decltype(src)(nullptr), 1, dst);
  };
  if (int rc = upload(h, "spp", pp, n + 1, &a)) return rc;
  if (int rc = upload_nz("spc", pc, np, &b)) return rc;
  if (int rc = upload_nz("spv", pv, np, &x)) return rc;
  if (int rc = upload(h, "sup", up, n + 1, &d)) return rc;
  if (int rc = upload_nz("suc", uc, nu, &e)) return rc;
  if (int rc = upload_nz("suv", uv, nu, &y)) return rc;
  if (int rc = upload(h, "svp", vp, n + 1, &f)) return rc;
  if (int rc = upload_nz("svc", vc, nv, &g)) return rc;
  if (int rc = upload_nz("svv", vv, nv, &z)) return rc;
  if (int rc = upload(h, "snr", normals, 2 * n, &nr)) return rc;
  if (int rc = upload(h, "snw", normal_weights, 2 * n, &nw)) return rc;
  c.nsurf = n; c.pp = a; c.pc = b; c.pv = x; c.up = d; c.uc = e; c.uv = y; c.vp = f; c.vc = g; c.vv = z;
  c.nrm = nr; c.nw = nw; c.diameter = diameter; c.uref = u_ref;
  if (n > FORCE_MAXP) return fail(h, 1, "too many surface points");
  double* fp;
  if (int rc = upload<double>(h, "sfp", nullptr, (size_t)4 * force_blocks(n) + FORCE_PPW, &fp)) return rc;
  c.fpart = fp;
  h->has_surface = true;
  return sync_descs(h);
}

int mcf_set_inlet_perturbation(mcf_handle* h, const double* base, const double* shape, double duration) {
  if (!h || !std::isfinite(duration) || (duration > 0.0 && (!base || !shape)))
    return fail(h, 1, "invalid inflow perturbation");
  CaseDev& c = h->hc;
  const bool was = c.pert_T > 0.0;
  if (duration > 0.0) {
    for (int j = 0; j < h->ny; ++j)
      if (!std::isfinite(base[j]) || !std::isfinite(shape[j])) return fail(h, 1, "inflow perturbation must be finite");
    double *b, *s;
    if (int rc = upload(h, "ipb", base, h->ny, &b)) return rc;
    if (int rc = upload(h, "ips", shape, h->ny, &s)) return rc;
    c.inlet_base = b; c.inlet_shape = s; c.pert_T = duration;
  } else {
    c.pert_T = 0.0;
  }
  // the step graphs hold k_inlet only while a perturbation is set
  if (was != (c.pert_T > 0.0)) drop_graphs(h);
  return sync_descs(h);
}

int mcf_set_probes(mcf_handle* h, int n, const int* ptr, const int* col, const double* val) {
  if (!h || n < 0 || n > 2 || (n > 0 && (!ptr || !col || !val))) return fail(h, 1, "invalid pressure probes");
  if (!h->has_surface) return fail(h, 1, "set the surface quadrature before the pressure probes");
  CaseDev& c = h->hc;
  c.nprobe = 0;
  if (n == 0) return sync_descs(h);
  if (ptr[0] != 0) return fail(h, 1, "probe rows must start at 0");
  for (int k = 0; k < n; ++k) if (ptr[k + 1] < ptr[k]) return fail(h, 1, "probe row pointers must not decrease");
  for (int t = 0; t < ptr[n]; ++t) if (col[t] < 0 || col[t] >= h->psize) return fail(h, 1, "probe column out of range");
  int *a, *b;
  double* x;
  if (int rc = upload(h, "qpp", ptr, n + 1, &a)) return rc;
  if (int rc = upload(h, "qpc", col, ptr[n], &b)) return rc;
  if (int rc = upload(h, "qpv", val, ptr[n], &x)) return rc;
  c.qp = a; c.qc = b; c.qv = x; c.nprobe = n;
  return sync_descs(h);
}

int mcf_set_state(mcf_handle* h, const double* vel, const double* p, const double* prev_adv,
                  const double* half_p, double prev_dt, double time) {
  if (!h || !vel || !p) return fail(h, 1, "state arrays required");
  if (!h->has_solid[2]) return fail(h, 1, "solid intervals must be set first");
  if (prev_adv && !(prev_dt > 0 && std::isfinite(prev_dt))) return fail(h, 1, "AB2 history requires a finite positive previous timestep");
  h->parity = 0;
  HIPCHK(h, hipMemcpy(h->hc.buf[B_S0], vel, h->velsize * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> pm(p, p + h->psize);
  const HostTables& t = g_host[h];
  for (int j = 0; j < h->ny; ++j)
    for (int i = t.lo[2][j]; i <= t.hi[2][j]; ++i) pm[(size_t)j * h->nx + i] = 0.0;
  HIPCHK(h, hipMemcpy(h->hc.buf[B_PH], half_p ? half_p : pm.data(), h->psize * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->hc.buf[B_POUT], pm.data(), h->psize * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(h->hc.buf[B_TP1], p, h->psize * sizeof(double), hipMemcpyHostToDevice));
  h->has_history = prev_adv != nullptr;
  if (prev_adv) HIPCHK(h, hipMemcpy(h->hc.buf[B_H1], prev_adv, h->velsize * sizeof(double), hipMemcpyHostToDevice));
  h->prev_dt = prev_adv ? prev_dt : 0.0;
  h->time = time;
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->owned.erase("p_raw_valid");
  h->owned["p_raw_valid"] = nullptr;
  return 0;
}

int mcf_get_state(mcf_handle* h, double* vel, double* p, double* prev_adv, double* half_p,
                  double* prev_dt, double* time, int* has_history) {
  if (!h) return 1;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (vel) HIPCHK(h, hipMemcpy(vel, h->hc.buf[B_S0 + h->parity], h->velsize * sizeof(double), hipMemcpyDeviceToHost));
  if (p) {
    if (h->owned.count("p_raw_valid")) {
      HIPCHK(h, hipMemcpy(p, h->hc.buf[B_TP1], h->psize * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      hipLaunchKernelGGL(k_copy, grid1(h->psize), dim3(256), 0, h->stream, h->dc, (int)B_POUT, (int)B_TP1, h->psize);
      FillArgs f{B_TP1, 0, 0, 0};
      hipLaunchKernelGGL(k_fill, grid1(std::max(h->hc.gt[2].n, 1)), dim3(256), 0, h->stream, h->dc, f);
      HIPCHK(h, hipStreamSynchronize(h->stream));
      HIPCHK(h, hipMemcpy(p, h->hc.buf[B_TP1], h->psize * sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  if (prev_adv && h->has_history)
    HIPCHK(h, hipMemcpy(prev_adv, h->hc.buf[B_H0 + 1 - h->parity], h->velsize * sizeof(double), hipMemcpyDeviceToHost));
  if (half_p && h->has_history)
    HIPCHK(h, hipMemcpy(half_p, h->hc.buf[B_PH], h->psize * sizeof(double), hipMemcpyDeviceToHost));
  if (prev_dt) *prev_dt = h->prev_dt;
  if (time) *time = h->time;
  if (has_history) *has_history = h->has_history ? 1 : 0;
  return 0;
}

namespace {
constexpr int SPEC_MIN = 8, SPEC_MAX = 512, SPEC_FULL = 16, SPEC_QUIET = 4;

// Speculation policy after a chunk.  A clean lean chunk doubles the next one.  After a lean chunk
// that missed (and was replayed in full) or a full chunk, the next chunk is lean again, from the
// shortest length, once the chunk's last SPEC_QUIET steps refined nothing; otherwise it is a full
// chunk of SPEC_FULL steps.  Start-up transients refine for tens of steps and steady flow rarely,
// so the full steps end within SPEC_FULL steps of the transient.
void spec_update(int& chunk, int& full, bool lean, bool miss, bool quiet) {
  if (lean && !miss) {
    chunk = std::min(2 * chunk, SPEC_MAX);
    full = 0;
    return;
  }
  chunk = SPEC_MIN;
  full = quiet ? 0 : SPEC_FULL;
}
// In the last SPEC_QUIET of k diagnostics rows: any velocity-solve refinement (vel), any pressure
// solve refined twice (p2).  The tail predicts the next chunk: refinement needs fade through a
// transient, and a start-up step (which runs in full anyway) should not make the chunk after it
// launch every pass.  A step that needs a pass the mask leaves out is flagged and replayed.
void refine_kinds(const double* rows, int k, bool& vel, bool& p2) {
  for (int s = std::max(0, k - SPEC_QUIET); s < k; ++s) {
    const double* r = rows + (size_t)s * MCF_DIAG_LEN;
    vel = vel || r[MCF_D_H_REFINE] > 0;
    for (int q = 0; q < (int)r[MCF_D_P_SOLVES] && q < 3; ++q)
      p2 = p2 || r[MCF_D_SOLVE0 + MCF_SOLVE_LEN * q + MCF_S_REFINE] >= 2;
  }
}
// Lean mask of the next chunk: all solves lean; or, for a full chunk, the velocity solve lean and
// the pressure solves with one refinement pass (start-up transients refine the pressure solves
// once per solve, the velocity solve rarely), each widened to the full form while the last chunk
// needed it
int spec_mask(bool spec, int full, bool vfull, bool p2) {
  return !spec ? 0 : full == 0 ? LEAN_ALL : vfull ? 0 : LEAN_VEL | (p2 ? 0 : ONCE_PRES);
}

// No refinement (pressure or velocity) in the last SPEC_QUIET of k diagnostics rows
bool quiet_tail(const double* rows, int k) {
  for (int s = std::max(0, k - SPEC_QUIET); s < k; ++s) {
    const double* r = rows + (size_t)s * MCF_DIAG_LEN;
    if (r[MCF_D_P_REFINE] > 0 || r[MCF_D_H_REFINE] > 0) return false;
  }
  return true;
}

// k steps on one handle into d (k rows).  Speculative: steady steps leave the refinement launches
// out; if any step of the chunk would have refined, the chunk is replayed from the checkpoint with
// the refinements in place.
// Wall time of the host-driven (GMRES) pressure solves into their records, in call order
void patch_seconds(mcf_handle* h, double* d, int rows) {
  if (!h->kry) return;
  std::vector<double>& sec = h->kry->seconds;
  size_t t = 0;
  for (int s = 0; s < rows; ++s) {
    double* row = d + (size_t)s * MCF_DIAG_LEN;
    const int n = (int)row[MCF_D_P_SOLVES];
    for (int q = 0; q < n && q < 3 && t < sec.size(); ++q)
      row[MCF_D_SOLVE0 + MCF_SOLVE_LEN * q + MCF_S_SECONDS] = sec[t++];
  }
  sec.clear();
}

int run_chunk(mcf_handle* h, int k, double dt, bool spec, int lean, double* d) {
  const int parity0 = h->parity;
  const double prev_dt0 = h->prev_dt, time0 = h->time;
  const bool hist0 = h->has_history;
  const Ctx x = ctx_of(h);
  const size_t bytes = (size_t)k * MCF_DIAG_LEN * sizeof(double);
  if (lean) hipLaunchKernelGGL(k_ckpt, gridSn(x.maxvel, 1), dim3(256), 0, h->stream, h->dc, parity0, 1);
  if (int rc = enqueue_steps(h, k, dt, lean)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(d, h->diag, bytes, hipMemcpyDeviceToHost));
  patch_seconds(h, d, k);
  bool miss = false;
  for (int s = 0; s < k && lean; ++s) miss = miss || ((int)d[(size_t)s * MCF_DIAG_LEN + MCF_D_STATUS] & 32);
  if (miss) {
    h->parity = parity0; h->prev_dt = prev_dt0; h->time = time0; h->has_history = hist0;
    h->prm_valid = false;
    ++h->replays;
    hipLaunchKernelGGL(k_ckpt, gridSn(x.maxvel, 1), dim3(256), 0, h->stream, h->dc, parity0, 0);
    if (int rc = enqueue_steps(h, k, dt, 0)) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(d, h->diag, bytes, hipMemcpyDeviceToHost));
  }
  if (spec) {
    spec_update(h->spec_chunk, h->spec_full, lean == LEAN_ALL, miss, quiet_tail(d, k));
    bool vel = false, p2 = false;
    refine_kinds(d, k, vel, p2);
    h->spec_vfull = vel; h->spec_p2 = p2;
  }
  return 0;
}
}  // namespace

int mcf_run(mcf_handle* h, int n, double dt, double* diag) {
  if (!h || n <= 0 || !(dt > 0) || !std::isfinite(dt)) return fail(h, 1, "n > 0 and a finite positive dt are required");
  if (int rc = check_ready(h, dt)) return rc;
  if (int rc = ensure_diag(h, n)) return rc;
  h->owned.erase("p_raw_valid");
  // no speculation while pairs are dumped (a replayed chunk would record its pairs twice) or on the
  // GMRES backend (its steps have no lean form)
  const bool spec = h->spec > 0 && h->hs[0].dump_cap == 0 && h->pbackend == MCF_PB_DIRECT;
  std::vector<double> d((size_t)n * MCF_DIAG_LEN);
  for (int done = 0; done < n;) {
    const int lean = spec_mask(spec, h->spec_full, h->spec_vfull, h->spec_p2);
    const int k = !spec ? n - done : std::min(lean == LEAN_ALL ? h->spec_chunk : h->spec_full, n - done);
    if (int rc = run_chunk(h, k, dt, spec, lean, d.data() + (size_t)done * MCF_DIAG_LEN)) return rc;
    for (int s = 0; s < k; ++s)
      if (int rc = status_error(h, d.data() + (size_t)(done + s) * MCF_DIAG_LEN, done + s)) {
        if (diag) std::memcpy(diag, d.data(), (size_t)(done + k) * MCF_DIAG_LEN * sizeof(double));
        return rc;
      }
    done += k;
  }
  if (diag) std::memcpy(diag, d.data(), d.size() * sizeof(double));
  return 0;
}
int mcf_step(mcf_handle* h, double dt, double* diag) { return mcf_run(h, 1, dt, diag); }

int mcf_project(mcf_handle* h, const double* q, double dt, double* out, double* phi, double* info4) {
  if (!h || !q || !(dt > 0)) return fail(h, 1, "invalid projection arguments");
  if (!h->has_rect[0] || !h->has_schur[0]) return fail(h, 1, "pressure factor missing");
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TV1], q, h->velsize * sizeof(double), hipMemcpyHostToDevice, h->stream));
  set_prm_once(h, dt);
  if (int rc = enqueue_project(ctx_of(h), B_TV1, B_TV0, B_TP0, 1.0, 0, -1)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (out) HIPCHK(h, hipMemcpy(out, h->hc.buf[B_TV0], h->velsize * sizeof(double), hipMemcpyDeviceToHost));
  if (phi) HIPCHK(h, hipMemcpy(phi, h->hc.buf[B_TP0], h->psize * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> sv;
  if (int rc = read_slots(h, sv)) return rc;
  double rec[MCF_SOLVE_LEN];
  solve_record(slots_of(sv, 0), h->pbackend == MCF_PB_GMRES, rec);
  if (h->kry) h->kry->seconds.clear();
  const double fin = rec[MCF_S_RES_LINF];
  const double div = sv[SL_DIV0];
  if (info4) { info4[0] = fin; info4[1] = rec[MCF_S_REFINE]; info4[2] = div; info4[3] = rec[MCF_S_TARGET]; }
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (!solve_ok(rec)) return fail(h, 2, "Pressure solve failed: residual " + std::to_string(fin) +
                                  " status " + std::to_string((int)rec[MCF_S_STATUS]));
  if (!(div <= 1e-10)) {
    char buf[128];
    std::snprintf(buf, sizeof buf, "Ghost projection failed: divergence=%.6g", div);
    return fail(h, 2, buf);
  }
  return 0;
}

int mcf_extend(mcf_handle* h, const double* q, double* out) {
  if (!h || !q || !out) return 1;
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TV0], q, h->velsize * sizeof(double), hipMemcpyHostToDevice, h->stream));
  FillArgs f{B_TV0, 1, 1, 0};
  const int nf = h->hc.gt[0].n + h->hc.gt[1].n + h->ny;
  hipLaunchKernelGGL(k_fill, grid1(nf), dim3(256), 0, h->stream, h->dc, f);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->hc.buf[B_TV0], h->velsize * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int mcf_forces(mcf_handle* h, const double* vel, const double* p, double mu, double* out6) {
  if (!h || !vel || !p || !out6 || !h->has_surface) return fail(h, 1, "invalid force arguments");
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TV0], vel, h->velsize * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TP0], p, h->psize * sizeof(double), hipMemcpyHostToDevice, h->stream));
  ForceArgs fa{B_TV0, B_TP0, mu};
  hipLaunchKernelGGL(k_forces, dim3(force_blocks(h->hc.nsurf), 1), dim3(64), 0, h->stream, h->dc, fa, FinArgs{}, 0);
  hipLaunchKernelGGL(k_force_total, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<double> sv;
  if (int rc = read_slots(h, sv)) return rc;
  for (int q = 0; q < 6; ++q) out6[q] = sv[SL_FORCE + q];
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int mcf_rect_inverse(mcf_handle* h, int kind, const double* rhs, double* out) {
  if (!h || kind < 0 || kind > 2 || !rhs || !out || !h->has_rect[kind]) return fail(h, 1, "invalid rectangle inverse arguments");
  SolveDesc t = h->hs[kind];
  const size_t n = (size_t)t.yy * t.xx;
  if (h->tmp_rect_n < n) {
    if (h->tmp_rect) (void)hipFree(h->tmp_rect);
    h->tmp_rect = nullptr;
    HIPCHK(h, hipMalloc(&h->tmp_rect, n * sizeof(double)));
    h->tmp_rect_n = n;
  }
  double** tb = nullptr;
  SolveDesc* td = nullptr;
  HIPCHK(h, hipMalloc(&tb, sizeof(double*)));
  HIPCHK(h, hipMalloc(&td, sizeof(SolveDesc)));
  int rc = rect_inverse_on(h, kind, rhs, out, h->tmp_rect, tb, td);
  (void)hipFree(tb); (void)hipFree(td);
  return rc;
}

int mcf_rect_inverse_noalloc(mcf_handle* h, int kind, const double* rhs, double* out) {
  if (!h || kind < 0 || kind > 2 || !rhs || !out || !h->has_rect[kind]) return fail(h, 1, "invalid rectangle inverse arguments");
  if (!h->na_scratch) return fail(h, 4, "allocation-free scratch is not set up on this device");
  const size_t n = (size_t)h->hs[kind].yy * h->hs[kind].xx;
  if (n > NA_MAX_VALUES)
    return fail(h, 1, "rectangle of " + std::to_string(n) + " values exceeds the allocation-free limit of " +
                          std::to_string(NA_MAX_VALUES) + " values (6 MiB); use mcf_rect_inverse");
  double* base = h->na_scratch;
  return rect_inverse_on(h, kind, rhs, out, base + NA_FIELD, reinterpret_cast<double**>(base + NA_TABLE),
                         reinterpret_cast<SolveDesc*>(base + NA_DESC));
}

int mcf_schur_solve(mcf_handle* h, int kind, const double* rhs, double* out, double* info4) {
  if (!h || kind < 0 || kind > 2 || !rhs || !out) return fail(h, 1, "invalid capacitance solve arguments");
  const bool pres = kind == MCF_PRESSURE;
  for (int q = pres ? 0 : 1; q < (pres ? 1 : 3); ++q)
    if (!h->has_rect[q] || !h->has_schur[q]) return fail(h, 1, "factor missing");
  if (!h->has_solid[0] || !h->has_solid[1] || !h->has_solid[2]) return fail(h, 1, "solid intervals missing");
  const int n = pres ? h->psize : h->velsize;
  const int in_id = pres ? B_TP0 : B_TV0, out_id = pres ? B_TP1 : B_TV1;
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[in_id], rhs, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  // max |values| of each block feeds the refinement targets
  std::vector<double> hv(rhs, rhs + n);
  double vm[2] = {0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    const int comp = pres ? 0 : (k < h->usize ? 0 : 1);
    vm[comp] = std::max(vm[comp], std::fabs(hv[k]));
  }
  std::vector<double> s0(NSLOT, 0.0);   // stripe 0 carries the right-hand side maxima
  if (pres) s0[SL_P0 + SV_VAL] = vm[0];
  else { s0[SL_HU + SV_VAL] = vm[0]; s0[SL_HV + SV_VAL] = vm[1]; }
  if (int rc = upload_slots(h, s0)) return rc;
  const double dt = pres ? 1.0 : -2.0 * h->hs[1].b / h->nu;
  set_prm_once(h, dt);
  if (int rc = enqueue_solve(ctx_of(h), pres, in_id, out_id, 0)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->hc.buf[out_id], n * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> sv;
  if (int rc = read_slots(h, sv)) return rc;
  auto sd = [&](int k) { return sv[k]; };
  if (info4) {
    if (pres) {
      const double val = sd(SL_P0 + SV_VAL), t1 = std::max(1e-10, 32 * kEps * std::max(1.0, val));
      const double r1 = sd(SL_P0 + SV_RES1), r2 = sd(SL_P0 + SV_RES2);
      info4[0] = sd(SL_P0 + SV_FINAL); info4[1] = (r1 > t1) + (r1 > t1 && r2 > t1);
      info4[2] = sd(SL_P0 + SV_ROUND); info4[3] = std::max(std::max(1e-10, 64 * kEps * std::max(1.0, val)), info4[2]);
    } else {
      double ref = 0;
      for (int b : {SL_HU, SL_HV}) {
        const double t1 = std::max(1e-10, 32 * kEps * std::max(1.0, sd(b + SV_VAL)));
        const double r1 = sd(b + SV_RES1), r2 = sd(b + SV_RES2);
        ref += (r1 > t1) + (r1 > t1 && r2 > t1);
      }
      info4[0] = std::max(sd(SL_HU + SV_FINAL), sd(SL_HV + SV_FINAL)); info4[1] = ref;
      info4[2] = sd(SL_HU + SV_FINAL); info4[3] = sd(SL_HV + SV_FINAL);
    }
  }
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int mcf_set_pressure_backend(mcf_handle* h, int backend, int restart, int max_cycles, double rtol, double atol) {
  if (!h) return 1;
  if ((backend != MCF_PB_DIRECT && backend != MCF_PB_GMRES) || restart < 1 || restart > KR_MAXR || max_cycles < 1 ||
      !(rtol >= 0) || !(atol >= 0) || !std::isfinite(rtol) || !std::isfinite(atol))
    return fail(h, 1, "invalid pressure backend settings");
  if (h->stream) HIPCHK(h, hipStreamSynchronize(h->stream));
  if (backend == MCF_PB_GMRES) {
    if (!h->kry) h->kry = new Krylov();
    Krylov& K = *h->kry;
    const size_t ldv = ((size_t)h->psize + 63) / 64 * 64;
    if (!K.V || K.restart != restart) {
      if (K.V) (void)hipFree(K.V);
      K.V = nullptr;
      HIPCHK(h, hipMalloc(&K.V, (size_t)(restart + 1) * ldv * sizeof(double)));
      HIPCHK(h, hipMemset(K.V, 0, (size_t)(restart + 1) * ldv * sizeof(double)));
    }
    K.ldv = ldv;
    K.nblk = (h->psize + KR_CHUNK - 1) / KR_CHUNK;
    if (!K.part) HIPCHK(h, hipMalloc(&K.part, (size_t)(KR_MAXR + 2) * K.nblk * sizeof(double)));
    if (!K.ctl) HIPCHK(h, hipMalloc(&K.ctl, sizeof(KryCtl)));
    if (!K.stat) HIPCHK(h, hipHostMalloc(&K.stat, KR_STAT_INTS * sizeof(int)));
    K.restart = restart; K.maxiter = max_cycles; K.rtol = rtol; K.atol = atol;
    KryCtl init{};
    init.restart = restart; init.maxiter = max_cycles; init.rtol = rtol; init.atol_in = atol;
    HIPCHK(h, hipMemcpy(K.ctl, &init, sizeof(KryCtl), hipMemcpyHostToDevice));
    h->hc.buf[B_KV0] = K.V;
  }
  h->pbackend = backend;
  return sync_descs(h);
}

int mcf_get_pressure_backend(mcf_handle* h, int* backend) {
  if (!h || !backend) return 1;
  *backend = h->pbackend;
  return 0;
}

int mcf_pressure_solve(mcf_handle* h, const double* rhs, double* out, double* record) {
  if (!h || !rhs || !out) return fail(h, 1, "invalid pressure solve arguments");
  if (!h->has_rect[0] || !h->has_schur[0] || !h->has_solid[2]) return fail(h, 1, "pressure factor missing");
  // the values on the active cells must be finite (pressure_schur.py:35-36); their max sets the targets
  const HostTables& ht = g_host[h];
  double vmax = 0.0;
  for (int j = 0; j < h->ny; ++j)
    for (int i = 0; i < h->nx; ++i) {
      if (ht.lo[2][j] <= i && i <= ht.hi[2][j]) continue;
      const double v = rhs[(size_t)j * h->nx + i];
      if (!std::isfinite(v)) return fail(h, 1, "Pressure RHS must be finite");
      vmax = std::max(vmax, std::fabs(v));
    }
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TP0], rhs, h->psize * sizeof(double), hipMemcpyHostToDevice, h->stream));
  std::vector<double> s0(NSLOT, 0.0);
  s0[SL_P0 + SV_VAL] = vmax;
  if (int rc = upload_slots(h, s0)) return rc;
  set_prm_once(h, 1.0);
  if (int rc = enqueue_solve(ctx_of(h), true, B_TP0, B_TP1, 0)) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipMemcpy(out, h->hc.buf[B_TP1], h->psize * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<double> sv;
  if (int rc = read_slots(h, sv)) return rc;
  double rec[MCF_SOLVE_LEN];
  solve_record(slots_of(sv, 0), h->pbackend == MCF_PB_GMRES, rec);
  if (h->kry) {
    if (!h->kry->seconds.empty()) rec[MCF_S_SECONDS] = h->kry->seconds.back();
    h->kry->seconds.clear();
  }
  if (record) std::memcpy(record, rec, sizeof rec);
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (!solve_ok(rec)) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "Pressure solve failed: residual_linf=%g target=%g status=%d iterations=%d",
                  rec[MCF_S_RES_LINF], rec[MCF_S_TARGET], (int)rec[MCF_S_STATUS], (int)rec[MCF_S_ITERATIONS]);
    return fail(h, 2, buf);
  }
  return 0;
}

int mcf_apply_operator(mcf_handle* h, int op, const double* in, double* out, double dt) {
  if (!h || !in || !out || op < 0 || op > 4) return fail(h, 1, "invalid operator arguments");
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  const dim3 gv = gridS(h->velsize), gp = gridS(h->psize);
  set_prm_once(h, dt > 0 ? dt : 1.0);
  if (op == MCF_APPLY_GRADIENT) {
    HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TP0], in, h->psize * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemsetAsync(h->hc.buf[B_TV0], 0, h->velsize * sizeof(double), h->stream));
    UpdArgs u{B_TV0, B_TP0, B_TV1, -1, -1.0};
    hipLaunchKernelGGL(k_update, gv, dim3(256), 0, h->stream, h->dc, u);
  } else {
    HIPCHK(h, hipMemcpyAsync(h->hc.buf[B_TV0], in, h->velsize * sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (op == MCF_APPLY_DIVERGENCE) {
      DivArgs d{B_TV0, B_TP0, 0.0, SL_SPARE};
      hipLaunchKernelGGL(k_div, gp, dim3(256), 0, h->stream, h->dc, d);
    } else if (op == MCF_APPLY_HMATVEC) {
      HIPCHK(h, hipMemsetAsync(h->hc.buf[B_TV1], 0, h->velsize * sizeof(double), h->stream));
      HelmArgs a{B_TV1, B_TV0, B_RES, -1, 0, SV_SSQR, -1};
      hipLaunchKernelGGL(k_helm, gv, dim3(256), 0, h->stream, h->dc, a);
    } else {
      FillArgs f{B_TV0, 1, 1, 0};
      const int nf = h->hc.gt[0].n + h->hc.gt[1].n + h->ny;
      hipLaunchKernelGGL(k_fill, grid1(nf), dim3(256), 0, h->stream, h->dc, f);
      HIPCHK(h, hipMemsetAsync(h->hc.buf[B_TP0], 0, h->psize * sizeof(double), h->stream));
      AdvArgs m{B_TV0, B_TV0, B_TP0, -1, B_TV1, B_RES, 3};
      hipLaunchKernelGGL(k_adv_rhs, gridAdv(h->velsize, 1), dim3(256), 0, h->stream, h->dc, m);
    }
  }
  if (int rc = launch_check(h, "apply operator")) return rc;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const double* src = op == MCF_APPLY_DIVERGENCE ? h->hc.buf[B_TP0]
                    : (op == MCF_APPLY_ADVECTION || op == MCF_APPLY_GRADIENT) ? h->hc.buf[B_TV1] : h->hc.buf[B_RES];
  const int n = op == MCF_APPLY_DIVERGENCE ? h->psize : h->velsize;
  HIPCHK(h, hipMemcpy(out, src, n * sizeof(double), hipMemcpyDeviceToHost));
  if (op == MCF_APPLY_HMATVEC) for (int k = 0; k < n; ++k) out[k] = -out[k];
  hipLaunchKernelGGL(k_clear_slots, dim3(1, 1), dim3(64), 0, h->stream, h->dc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int mcf_divergence_stats(mcf_handle* h, const double* vel, const unsigned char* mask, double* out5) {
  if (!h || !vel || !mask || !out5) return 1;
  unsigned char* dm = nullptr;
  double* acc = nullptr;
  HIPCHK(h, hipMalloc(&dm, h->psize));
  HIPCHK(h, hipMalloc(&acc, NSLOTW * sizeof(double)));   // five slots in the slot layout
  HIPCHK(h, hipMemcpy(dm, mask, h->psize, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(acc, 0, NSLOTW * sizeof(double)));
  HIPCHK(h, hipMemcpy(h->hc.buf[B_TV0], vel, h->velsize * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_divstats, grid1(h->psize), dim3(256), 0, h->stream, h->hc.buf[B_TV0],
                     h->hc.buf[B_TV0] + h->usize, dm, h->nx, h->ny, h->dx, h->dy, acc);
  HIPCHK(h, hipStreamSynchronize(h->stream));
  std::vector<double> raw(NSLOTW);
  HIPCHK(h, hipMemcpy(raw.data(), acc, raw.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int q = 0; q < 5; ++q) {
    double v = 0.0;
    for (int t = 0; t < NSTRIPE; ++t) v = q == 0 ? std::max(v, raw[slot_word(q, t)]) : v + raw[slot_word(q, t)];
    out5[q] = v;
  }
  (void)hipFree(dm); (void)hipFree(acc);
  return 0;
}

int mcf_profile(mcf_handle* h, int n, double dt, double* out5) {
  if (!h || n <= 0 || !out5) return 1;
  if (int rc = check_ready(h, dt)) return rc;
  if (int rc = ensure_diag(h, n)) return rc;
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0)); HIPCHK(h, hipEventCreate(&e1));
  h->owned.erase("p_raw_valid");
  // one untimed step settles the history so that the timed steps replay the captured graph; with
  // speculation on, the timed steps are the lean ones (a step that would have refined is counted
  // in out5[4], not replayed: the state is not used after profiling)
  const int lean = h->spec > 0 && h->pbackend == MCF_PB_DIRECT ? LEAN_ALL : 0;
  if (int rc = enqueue_steps(h, 1, dt, lean)) return rc;
  HIPCHK(h, hipEventRecord(e0, h->stream));
  if (int rc = enqueue_steps(h, n, dt, lean)) return rc;
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
  out5[2] = ms / n;
  // dominant kernels, timed back to back on the same stream over the pressure problem
  const SolveDesc& p = h->hs[0];
  const int reps = 50;
  const int gvp = ctx_of(h).ln.gemm_p;
  GemmArgs g{1, B_DIVP, B_DIVP, 0, -1, 0};
  HIPCHK(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r) launch_gemm(gvp, p.yy, p.xx, 1, h->stream, h->ds, g);
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
  out5[0] = ms / reps;
  // the same launches one at a time, each bracketed by its own event pair and completed before the
  // next is enqueued: the kernel's own duration plus the events' timestamp overhead, without the
  // overlap of consecutive dispatches' ramp-down and ramp-up
  {
    double sum = 0.0;
    for (int r = 0; r < reps; ++r) {
      HIPCHK(h, hipEventRecord(e0, h->stream));
      launch_gemm(gvp, p.yy, p.xx, 1, h->stream, h->ds, g);
      HIPCHK(h, hipEventRecord(e1, h->stream));
      HIPCHK(h, hipEventSynchronize(e1));
      HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
      sum += ms;
    }
    out5[3] = sum / reps;
  }
  PcrArgs p0{0, -1, 0};
  HIPCHK(h, hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r)
    launch_tri(h->ln.tri_p, p.yy, p.xx, 1, h->stream, h->ds, p0);
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
  out5[1] = ms / reps;
  out5[4] = 0;
  (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
  std::vector<double> d((size_t)n * MCF_DIAG_LEN);
  HIPCHK(h, hipMemcpy(d.data(), h->diag, d.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int s = 0; s < n; ++s) {
    const double* row = d.data() + (size_t)s * MCF_DIAG_LEN;
    if ((int)row[MCF_D_STATUS] & 32) { out5[4] += 1; continue; }
    if (int rc = status_error(h, row, s)) return rc;
  }
  return launch_check(h, "profile");
}

int mcf_set_launch(mcf_handle* h, int gemm, int tri) {
  if (!h || gemm < -1 || gemm > GV_LAST || tri < -1 || tri >= N_TRI_V || ((gemm < 0) != (tri < 0)))
    return fail(h, 1, "invalid launch shape");
  if (gemm < 0) {
    h->ln = Launch();
    h->tuned = false;
  } else {
    const Ctx x = ctx_of(h);
    pin_launch(h->ln, gemm, tri, !x.noxf_p, !x.noxf_h);
    h->tuned = true;
  }
  drop_graphs(h);
  return 0;
}

int mcf_set_launch_classes(mcf_handle* h, const int* in4) {
  if (!h || !in4) return fail(h, 1, "launch shapes required");
  const Ctx x = ctx_of(h);
  for (int c = 0; c < 2; ++c) {
    const int g = in4[c], t = in4[2 + c];
    const bool xf = c == 0 ? !x.noxf_p : !x.noxf_h;
    const bool tri_ok = t >= 0 && t < N_TRI_V;
    const bool ok = (g == GV_XFFT) ? (t == TV_YSWEEP && xf) : ((gv_par(g) || (g >= 0 && g < N_GEMM_V)) && tri_ok);
    if (!ok) return fail(h, 1, "invalid launch shape for a problem class");
  }
  h->ln.gemm_p = in4[0]; h->ln.gemm_h = in4[1]; h->ln.tri_p = in4[2]; h->ln.tri_h = in4[3];
  h->tuned = true;
  drop_graphs(h);
  return 0;
}

int mcf_get_launch(mcf_handle* h, int* out5) {
  if (!h || !out5) return 1;
  const Launch ln = ctx_of(h).ln;   // the shapes a run takes now
  out5[0] = ln.gemm_p; out5[1] = ln.gemm_h; out5[2] = ln.tri_p; out5[3] = ln.tri_h;
  out5[4] = h->tuned ? 1 : 0;
  return 0;
}

int mcf_set_speculation(mcf_handle* h, int mode) {
  if (!h || mode < 0 || mode > 2) return fail(h, 1, "speculation mode must be 0, 1 or 2");
  if (mode != h->spec) drop_graphs(h);
  h->spec = mode;
  return 0;
}

int mcf_get_speculation(mcf_handle* h, int* out2) {
  if (!h || !out2) return 1;
  out2[0] = h->spec;
  out2[1] = h->replays;
  return 0;
}

}  // extern "C"

// ======================================================================================
// Batched ensemble: many cases of uneven size advance together, one launch per stage
// ======================================================================================
struct mcf_batch {
  std::vector<mcf_handle*> cases;
  std::vector<double> dt;       // each case's time step (its Helmholtz factors were built for it)
  CaseDev* dc = nullptr;
  SolveDesc* ds = nullptr;      // n pressure problems, then 2n Helmholtz problems
  hipStream_t stream = nullptr;
  StepGraphs graphs;
  Ctx ctx{};
  bool built = false;
  Launch ln;                    // launch shapes for the whole batch
  bool tuned = false;
  StepGraphs lean_graphs;       // steady steps without the refinement launches
  StepGraphs mixed[8];          // steady steps with other lean masks (as the handle's)
  int spec = 1;                 // speculative lean steps (as mcf_set_speculation)
  int replays = 0;
  int spec_chunk = 8;           // steps per speculative chunk (as the handle's)
  int spec_full = 0;
  bool spec_vfull = false, spec_p2 = false;
  void drop_graphs() {
    graphs.drop(); lean_graphs.drop();
    for (StepGraphs& g : mixed) g.drop();
  }
};

namespace {

// Order of a batch's transform problems.  The XCD-aware tile order gives each XCD a contiguous
// eighth of the problems, so the problems are arranged in eight consecutive groups of equal count
// and nearly equal transform work (largest first, each into the lightest group with room; ties by
// index): with cases in their given order, uneven sizes left some XCDs 1.4x the mean work.
std::vector<int> xcd_balanced_order(const std::vector<double>& work) {
  const int n = (int)work.size(), G = std::min(8, n);
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work[a] > work[b]; });
  std::vector<std::vector<int>> groups(G);
  std::vector<double> load(G, 0.0);
  for (int i : idx) {
    int best = -1;
    for (int g = 0; g < G; ++g) {
      const int cap = n / G + (g < n % G ? 1 : 0);
      if ((int)groups[g].size() < cap && (best < 0 || load[g] < load[best])) best = g;
    }
    groups[best].push_back(i);
    load[best] += work[i];
  }
  std::vector<int> order;
  for (auto& g : groups) order.insert(order.end(), g.begin(), g.end());
  return order;
}

int batch_build(mcf_batch* b) {
  mcf_handle* e = b->cases[0];
  const int n = (int)b->cases.size();
  std::vector<CaseDev> hc(n);
  std::vector<SolveDesc> hs(3 * n);
  Ctx x{};
  x.err = e; x.stream = b->stream; x.n = n;
  std::vector<double> wp(n), wh(2 * n);
  for (int c = 0; c < n; ++c) {
    const SolveDesc* s3 = b->cases[c]->hs;
    wp[c] = (double)s3[0].yy * s3[0].yy * s3[0].xx;
    for (int q = 0; q < 2; ++q) wh[2 * c + q] = (double)s3[1 + q].yy * s3[1 + q].yy * s3[1 + q].xx;
  }
  const std::vector<int> op = xcd_balanced_order(wp), oh = xcd_balanced_order(wh);
  for (int c = 0; c < n; ++c) {
    mcf_handle* h = b->cases[c];
    hc[c] = h->hc;
    hs[c] = b->cases[op[c]]->hs[0];
    hs[n + c] = b->cases[oh[c] / 2]->hs[1 + oh[c] % 2];
    hs[2 * n + c] = b->cases[oh[n + c] / 2]->hs[1 + oh[n + c] % 2];
    ctx_shapes(x, h->hc, h->hs);
  }
  if (!b->dc) HIPCHK(e, hipMalloc(&b->dc, n * sizeof(CaseDev)));
  if (!b->ds) HIPCHK(e, hipMalloc(&b->ds, 3 * n * sizeof(SolveDesc)));
  HIPCHK(e, hipMemcpy(b->dc, hc.data(), n * sizeof(CaseDev), hipMemcpyHostToDevice));
  HIPCHK(e, hipMemcpy(b->ds, hs.data(), 3 * n * sizeof(SolveDesc), hipMemcpyHostToDevice));
  x.dc = b->dc; x.dsp = b->ds; x.dsh = b->ds + n;
  x.ln = effective_launch(b->ln, !x.nopar_p, !x.nopar_h);
  b->ctx = x;
  b->drop_graphs();
  b->built = true;
  return 0;
}

}  // namespace

extern "C" {

int mcf_batch_create(mcf_handle* const* cases, const double* dts, int n, mcf_batch** out) {
  if (!out) return 1;
  *out = nullptr;
  if (!cases || !dts || n <= 0) return 1;
  mcf_handle* e = cases[0];
  for (int c = 0; c < n; ++c) {
    mcf_handle* h = cases[c];
    if (!h) return 1;
    const double dt = dts[c];
    if (!(dt > 0)) return fail(e, 1, "case " + std::to_string(c) + ": dt must be positive");
    if (int rc = check_ready(h, dt)) { e->err = "case " + std::to_string(c) + ": " + h->err; return rc; }
    if (h->pbackend != MCF_PB_DIRECT)
      return fail(e, 1, "case " + std::to_string(c) + ": the GMRES pressure backend is not batched");
    if (h->parity != cases[0]->parity || h->has_history != cases[0]->has_history)
      return fail(e, 1, "batched cases must share the step parity and history state");
    if (h->has_history && h->prev_dt != dt) return fail(e, 1, "batched cases must continue at their previous dt");
  }
  mcf_batch* b = new mcf_batch();
  b->cases.assign(cases, cases + n);
  b->dt.assign(dts, dts + n);
  if (hipStreamCreate(&b->stream) != hipSuccess) { delete b; return fail(e, 4, "hipStreamCreate failed"); }
  *out = b;
  return 0;
}

void mcf_batch_destroy(mcf_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->drop_graphs();
  (void)hipFree(b->dc); (void)hipFree(b->ds);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* mcf_batch_last_error(const mcf_batch* b) { return b && !b->cases.empty() ? b->cases[0]->err.c_str() : "null batch"; }

int mcf_batch_run(mcf_batch* b, int nsteps, double* diag) {
  if (!b || nsteps <= 0) return 1;
  mcf_handle* e = b->cases[0];
  const int n = (int)b->cases.size();
  bool rings = false;
  for (mcf_handle* h : b->cases) {
    (void)hipStreamSynchronize(h->stream);
    const int cap = h->diag_cap;
    if (int rc = ensure_diag(h, nsteps)) return rc;
    rings = rings || cap != h->diag_cap;
    h->owned.erase("p_raw_valid");
  }
  if (!b->built) {
    if (int rc = batch_build(b)) return rc;
  } else if (rings) {
    // new diagnostics rings: the batch's case descriptors are copied again; its graphs reach them
    // through the same table and stay valid
    std::vector<CaseDev> hc(n);
    for (int c = 0; c < n; ++c) hc[c] = b->cases[c]->hc;
    HIPCHK(e, hipMemcpy(b->dc, hc.data(), n * sizeof(CaseDev), hipMemcpyHostToDevice));
  }
  if (!b->tuned) {
    if (int rc = autotune(b->ctx, b->ln)) return rc;
    b->ctx.ln = effective_launch(b->ln, !b->ctx.nopar_p, !b->ctx.nopar_h);
    b->tuned = true;
    b->drop_graphs();
  }
  const Ctx& x = b->ctx;
  // speculative chunks over the batch, as mcf_run: lean steady steps in chunks of growing size, a
  // chunk replayed with the refinements from the checkpoint if any step of any case would refine
  const bool spec = b->spec > 0;
  struct Host { int parity; double prev_dt, time; bool hist; };
  std::vector<Host> saved(n);
  auto enqueue = [&](int k, int lean) -> int {
    for (int c = 0; c < n; ++c) {
      mcf_handle* h = b->cases[c];
      const int zero = 0;
      HIPCHK(e, hipMemcpy(h->clock_step, &zero, sizeof(int), hipMemcpyHostToDevice));
      HIPCHK(e, hipMemcpy(h->clock_time, &h->time, sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(e, hipMemcpy(h->prm, &b->dt[c], sizeof(double), hipMemcpyHostToDevice));
      h->prm_valid = false;
    }
    StepGraphs& gs = lean == LEAN_ALL ? b->lean_graphs : lean == 0 ? b->graphs : b->mixed[lean];
    for (int s = 0; s < k; ++s) {
      mcf_handle* h0 = b->cases[0];
      const bool first = !h0->has_history;
      hipLaunchKernelGGL(k_set_prm, dim3(1, n), dim3(64), 0, b->stream, b->dc, -1.0, first ? 0.0 : 0.5, first ? 1.0 : 0.5);
      if (first || !e->use_graphs) {
        if (int rc = enqueue_step_kernels(x, h0->parity, first, lean, b->spec == 2)) return rc;
      } else {
        if (!gs.valid) {
          if (int rc = gs.build(x, lean, b->spec == 2)) return rc;
          // every pair the speculation can switch to, captured now (as the handle's)
          for (int m : SPEC_MASKS) {
            StepGraphs& g = m == LEAN_ALL ? b->lean_graphs : m == 0 ? b->graphs : b->mixed[m];
            if (spec && !g.valid)
              if (int rc = g.build(x, m, b->spec == 2)) return rc;
          }
        }
        HIPCHK(e, hipGraphLaunch(gs.exec[h0->parity], b->stream));
      }
      for (int c = 0; c < n; ++c) advance_host_state(b->cases[c], b->dt[c]);
    }
    HIPCHK(e, hipStreamSynchronize(b->stream));
    return 0;
  };
  const size_t per = (size_t)nsteps * MCF_DIAG_LEN;
  std::vector<double> d(per * n);
  // every chunk runs for every case: a case that fails a gate keeps stepping (the caller drops it)
  // and no longer counts for the speculation
  int first_error = 0;
  std::vector<char> dead(n, 0);
  for (int done = 0; done < nsteps;) {
    const int lean = spec_mask(spec, b->spec_full, b->spec_vfull, b->spec_p2);
    const int k = !spec ? nsteps - done : std::min(lean == LEAN_ALL ? b->spec_chunk : b->spec_full, nsteps - done);
    for (int c = 0; c < n; ++c) {
      mcf_handle* h = b->cases[c];
      saved[c] = Host{h->parity, h->prev_dt, h->time, h->has_history};
    }
    if (lean) hipLaunchKernelGGL(k_ckpt, gridSn(x.maxvel, n), dim3(256), 0, b->stream, b->dc, saved[0].parity, 1);
    if (int rc = enqueue(k, lean)) return rc;
    const size_t rows = (size_t)k * MCF_DIAG_LEN;
    auto copy_rows = [&]() -> int {
      for (int c = 0; c < n; ++c)
        HIPCHK(e, hipMemcpy(d.data() + per * c + (size_t)done * MCF_DIAG_LEN, b->cases[c]->diag, rows * sizeof(double),
                            hipMemcpyDeviceToHost));
      return 0;
    };
    if (int rc = copy_rows()) return rc;
    bool miss = false;
    for (int c = 0; c < n && lean; ++c)
      for (int s = 0; s < k && !dead[c]; ++s) miss = miss || ((int)d[per * c + (size_t)(done + s) * MCF_DIAG_LEN + MCF_D_STATUS] & 32);
    if (miss) {
      for (int c = 0; c < n; ++c) {
        mcf_handle* h = b->cases[c];
        h->parity = saved[c].parity; h->prev_dt = saved[c].prev_dt; h->time = saved[c].time; h->has_history = saved[c].hist;
      }
      ++b->replays;
      hipLaunchKernelGGL(k_ckpt, gridSn(x.maxvel, n), dim3(256), 0, b->stream, b->dc, saved[0].parity, 0);
      if (int rc = enqueue(k, 0)) return rc;
      if (int rc = copy_rows()) return rc;
    }
    if (spec) {
      bool quiet = true, vel = false, p2 = false;   // over the members that still count
      for (int c = 0; c < n; ++c) {
        if (dead[c]) continue;
        const double* rows = d.data() + per * c + (size_t)done * MCF_DIAG_LEN;
        quiet = quiet && quiet_tail(rows, k);
        refine_kinds(rows, k, vel, p2);
      }
      spec_update(b->spec_chunk, b->spec_full, lean == LEAN_ALL, miss, quiet);
      b->spec_vfull = vel; b->spec_p2 = p2;
    }
    for (int c = 0; c < n; ++c) {
      mcf_handle* h = b->cases[c];
      for (int s = done; s < done + k && !dead[c]; ++s)
        if (int rc = status_error(h, d.data() + per * c + (size_t)s * MCF_DIAG_LEN, s)) {
          dead[c] = 1;
          if (!first_error) {
            first_error = rc;
            e->err = "case " + std::to_string(c) + ": " + h->err;
          }
        }
    }
    done += k;
  }
  if (diag)
    for (int c = 0; c < n; ++c) std::memcpy(diag + per * c, d.data() + per * c, per * sizeof(double));
  return first_error;
}

int mcf_batch_set_launch(mcf_batch* b, int gemm, int tri) {
  if (!b) return 1;
  mcf_handle* e = b->cases[0];
  if (gemm < -1 || gemm > GV_LAST || tri < -1 || tri >= N_TRI_V || ((gemm < 0) != (tri < 0)))
    return fail(e, 1, "invalid launch shape");
  if (gemm < 0) {
    b->ln = Launch();
    b->tuned = false;
  } else {
    pin_launch(b->ln, gemm, tri, !b->ctx.noxf_p, !b->ctx.noxf_h);
    b->tuned = true;
  }
  b->ctx.ln = effective_launch(b->ln, !b->ctx.nopar_p, !b->ctx.nopar_h);
  b->drop_graphs();
  return 0;
}

int mcf_batch_get_launch(mcf_batch* b, int* out5) {
  if (!b || !out5) return 1;
  const Launch ln = b->built ? effective_launch(b->ln, !b->ctx.nopar_p, !b->ctx.nopar_h) : b->ln;
  out5[0] = ln.gemm_p; out5[1] = ln.gemm_h; out5[2] = ln.tri_p; out5[3] = ln.tri_h;
  out5[4] = b->tuned ? 1 : 0;
  return 0;
}

int mcf_batch_set_speculation(mcf_batch* b, int mode) {
  if (!b) return 1;
  if (mode < 0 || mode > 2) return fail(b->cases[0], 1, "speculation mode must be 0, 1 or 2");
  if (mode != b->spec) b->drop_graphs();
  b->spec = mode;
  return 0;
}

int mcf_batch_get_speculation(mcf_batch* b, int* out2) {
  if (!b || !out2) return 1;
  out2[0] = b->spec;
  out2[1] = b->replays;
  return 0;
}

}  // extern "C"

// ======================================================================================
// The capacitance (Schur) system S y = C R^{-1} b: pair dump and a restarted GMRES on the device
// ======================================================================================
namespace {

constexpr int GM_T = 1024, GM_W = GM_T / 64, GM_MAXM = 160, GM_MAXR = 2048;

struct GmresArgs {
  const double* S;      // r*r row-major
  int r, m, max_iter;
  double rtol;
  const double* b;      // n*r
  const double* x0;     // n*r or null
  double* x;            // n*r
  int* iters;           // n
  double* relres;       // n: final true residual |b - S x| / |b| (2-norm)
  double* V;            // n*(m+1)*r Krylov basis
  double* R;            // n*m*m upper triangle, column-major (column i at R + i*m)
};

// sum over the block (16 waves); every thread returns the same value, summed in a fixed order
DEV double gm_block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < GM_W; ++k) s += red[k];
  return s;
}

// out = S v (v, out in LDS): one wave per row, lanes over the columns
DEV void gm_matvec(const double* S, int r, const double* v, double* out) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = w; i < r; i += GM_W) {
    const double* Si = S + (size_t)i * r;
    double s = 0.0;
    for (int k = lane; k < r; k += 64) s += Si[k] * v[k];
    s = wave_sum(s);
    if (lane == 0) out[i] = s;
  }
  __syncthreads();
}

// out = b - S x in compensated arithmetic (Dot2 of Ogita, Rump & Oishi 2005: exact products by FMA,
// error-free sums), so the restart residual is accurate to about the unit round-off of the
// residual itself rather than of |S| |x|
DEV void gm_two_sum(double a, double b, double& s, double& e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

DEV void gm_residual(const double* S, int r, const double* x, const double* b, double* out) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = w; i < r; i += GM_W) {
    const double* Si = S + (size_t)i * r;
    double s = lane == 0 ? b[i] : 0.0, c = 0.0;
    for (int k = lane; k < r; k += 64) {
      const double p = -Si[k] * x[k];
      const double pe = fma(-Si[k], x[k], -p);
      double t, e;
      gm_two_sum(s, p, t, e);
      s = t;
      c += e + pe;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double so = __shfl_xor(s, off), co = __shfl_xor(c, off);
      double t, e;
      gm_two_sum(s, so, t, e);
      s = t;
      c = c + co + e;
    }
    if (lane == 0) out[i] = s + c;
  }
  __syncthreads();
}

// GMRES(m) with classical Gram-Schmidt applied twice and Givens rotations (Saad & Schultz 1986),
// one workgroup per system; stops when the true residual |b - S x| <= rtol |b|, after max_iter
// Arnoldi steps, or when a restart no longer reduces the true residual by 10%.
__global__ __launch_bounds__(GM_T) void k_gmres(GmresArgs a) {
  extern __shared__ double lds[];
  const int sys = blockIdx.x, r = a.r, m = a.m, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double* xs = lds;            // r
  double* ws = xs + r;         // r
  double* vs = ws + r;         // r
  double* hc = vs + r;         // m+2: current Hessenberg column
  double* tp = hc + m + 2;     // m+2: Gram-Schmidt coefficients of one pass
  double* cs = tp + m + 2;     // m
  double* sn = cs + m;         // m
  double* g = sn + m;          // m+2
  double* red = g + m + 2;     // GM_W
  const double* b = a.b + (size_t)sys * r;
  double* V = a.V + (size_t)sys * (m + 1) * r;
  double* R = a.R + (size_t)sys * m * m;
  for (int i = tid; i < r; i += GM_T) xs[i] = a.x0 ? a.x0[(size_t)sys * r + i] : 0.0;
  double part = 0.0;
  for (int i = tid; i < r; i += GM_T) part += b[i] * b[i];
  const double bnorm = sqrt(gm_block_sum(part, red));
  const double target = a.rtol * bnorm;
  int total = 0, cycles = 0;
  double beta = 0.0, beta_prev = 0.0;
  while (true) {
    __syncthreads();
    gm_residual(a.S, r, xs, b, ws);
    part = 0.0;
    for (int i = tid; i < r; i += GM_T) part += ws[i] * ws[i];
    beta = sqrt(gm_block_sum(part, red));
    if (!(beta > target) || total >= a.max_iter) break;
    // a restart that no longer reduces the true residual has reached the attainable accuracy
    if (cycles > 0 && !(beta < 0.9 * beta_prev)) break;
    beta_prev = beta;
    ++cycles;
    for (int i = tid; i < r; i += GM_T) V[i] = ws[i] / beta;
    if (tid == 0) g[0] = beta;
    int jdone = 0;
    for (int j = 0; j < m; ++j) {
      __syncthreads();
      for (int i = tid; i < r; i += GM_T) vs[i] = V[(size_t)j * r + i];
      __syncthreads();
      gm_matvec(a.S, r, vs, ws);
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = w; i <= j; i += GM_W) {
          const double* Vi = V + (size_t)i * r;
          double d = 0.0;
          for (int k = lane; k < r; k += 64) d += Vi[k] * ws[k];
          d = wave_sum(d);
          if (lane == 0) { tp[i] = d; hc[i] = pass ? hc[i] + d : d; }
        }
        __syncthreads();
        for (int k = tid; k < r; k += GM_T) {
          double acc = ws[k];
          for (int i = 0; i <= j; ++i) acc -= tp[i] * V[(size_t)i * r + k];
          ws[k] = acc;
        }
        __syncthreads();
      }
      part = 0.0;
      for (int i = tid; i < r; i += GM_T) part += ws[i] * ws[i];
      const double hn = sqrt(gm_block_sum(part, red));
      if (hn > 0.0)
        for (int i = tid; i < r; i += GM_T) V[(size_t)(j + 1) * r + i] = ws[i] / hn;
      if (tid == 0) {
        hc[j + 1] = hn;
        for (int i = 0; i < j; ++i) {
          const double t0 = cs[i] * hc[i] + sn[i] * hc[i + 1];
          hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
          hc[i] = t0;
        }
        const double rho = hypot(hc[j], hc[j + 1]);
        cs[j] = rho > 0.0 ? hc[j] / rho : 1.0;
        sn[j] = rho > 0.0 ? hc[j + 1] / rho : 0.0;
        hc[j] = rho;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
      }
      __syncthreads();
      for (int i = tid; i <= j; i += GM_T) R[(size_t)j * m + i] = hc[i];
      ++total;
      jdone = j + 1;
      if (!(fabs(g[j + 1]) > target) || total >= a.max_iter || !(hn > 0.0)) break;
    }
    __syncthreads();
    // back substitution R z = g, column by column (z overwrites g)
    for (int i = jdone - 1; i >= 0; --i) {
      if (tid == 0) g[i] = g[i] / R[(size_t)i * m + i];
      __syncthreads();
      const double zi = g[i];
      for (int k = tid; k < i; k += GM_T) g[k] -= R[(size_t)i * m + k] * zi;
      __syncthreads();
    }
    for (int k = tid; k < r; k += GM_T) {
      double acc = xs[k];
      for (int i = 0; i < jdone; ++i) acc += g[i] * V[(size_t)i * r + k];
      xs[k] = acc;
    }
  }
  for (int i = tid; i < r; i += GM_T) a.x[(size_t)sys * r + i] = xs[i];
  if (tid == 0) {
    a.iters[sys] = total;
    a.relres[sys] = bnorm > 0.0 ? beta / bnorm : 0.0;
  }
}

}  // namespace

extern "C" {

int mcf_set_schur_system(mcf_handle* h, int kind, const double* S, const int* Cptr, const int* Ccol,
                         const double* Cval, const int* angle_order) {
  if (!h || kind < 0 || kind > 2) return fail(h, 1, "invalid Schur system");
  SolveDesc& s = h->hs[kind];
  if (!h->has_schur[kind]) return fail(h, 1, "set the capacitance data before the Schur system");
  const int r = s.r, npts = s.npts;
  if (r == 0) return 0;
  if (!S || !Cptr || !Ccol || !Cval || !angle_order) return fail(h, 1, "invalid Schur system");
  if (Cptr[0] != 0) return fail(h, 1, "C row pointer must start at 0");
  for (int m = 0; m < r; ++m) if (Cptr[m + 1] < Cptr[m]) return fail(h, 1, "C row pointer not monotone");
  const int nnz = Cptr[r];
  std::vector<int> col(nnz);
  for (int q = 0; q < nnz; ++q) {
    if (Ccol[q] < 0 || Ccol[q] >= npts) return fail(h, 1, "C column out of range");
    col[q] = h->pt_pos[kind][Ccol[q]];
  }
  std::vector<int> apos(r, -1);
  for (int t = 0; t < r; ++t) {
    const int m = angle_order[t];
    if (m < 0 || m >= r || apos[m] >= 0) return fail(h, 1, "angle order is not a permutation of the rows");
    apos[m] = t;
  }
  std::vector<double> sth((size_t)r * r);
  for (int t = 0; t < r; ++t)
    for (int u = 0; u < r; ++u) sth[(size_t)t * r + u] = S[(size_t)angle_order[t] * r + angle_order[u]];
  const std::string key = "ssys" + std::to_string(kind);
  double *dS, *dv; int *dp, *dc, *da;
  if (int rc = upload(h, key + "S", sth.data(), sth.size(), &dS)) return rc;
  if (int rc = upload(h, key + "p", Cptr, (size_t)r + 1, &dp)) return rc;
  if (int rc = upload(h, key + "c", col.data(), std::max<size_t>(col.size(), 1), &dc)) return rc;
  if (int rc = upload(h, key + "v", Cval, std::max<size_t>(nnz, 1), &dv)) return rc;
  if (int rc = upload(h, key + "a", apos.data(), apos.size(), &da)) return rc;
  s.Sth = dS; s.Cptr = dp; s.Ccol = dc; s.Cval = dv; s.apos = da; s.clock = h->clock_step;
  return sync_descs(h);
}

int mcf_set_dump(mcf_handle* h, int capacity) {
  if (!h || capacity < 0) return fail(h, 1, "invalid dump capacity");
  SolveDesc& s = h->hs[0];
  if (capacity > 0 && !s.Sth) return fail(h, 1, "set the pressure Schur system before dumping pairs");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const size_t n = capacity > 0 ? (size_t)capacity * std::max(s.r, 1) : 0;
  if (int rc = upload<double>(h, "dumpc", nullptr, n, &s.dump_c)) return rc;
  if (int rc = upload<double>(h, "dumpy", nullptr, n, &s.dump_y)) return rc;
  if (int rc = upload<int>(h, "dumpt", nullptr, (size_t)capacity, &s.dump_tag)) return rc;
  if (int rc = upload<int>(h, "dumpn", nullptr, capacity > 0 ? 1 : 0, &s.dump_count)) return rc;
  s.dump_cap = capacity > 0 && s.r > 0 ? capacity : 0;
  return sync_descs(h);
}

int mcf_get_dump(mcf_handle* h, int max_pairs, double* c, double* y, int* tags, int* count) {
  if (!h || !count || max_pairs < 0) return fail(h, 1, "invalid dump request");
  const SolveDesc& s = h->hs[0];
  *count = 0;
  if (!s.dump_cap) return 0;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  int n = 0;
  HIPCHK(h, hipMemcpy(&n, s.dump_count, sizeof(int), hipMemcpyDeviceToHost));
  n = std::min(n, max_pairs);
  if (n > 0) {
    if (!c || !y || !tags) return fail(h, 1, "invalid dump buffers");
    HIPCHK(h, hipMemcpy(c, s.dump_c, (size_t)n * s.r * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(y, s.dump_y, (size_t)n * s.r * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(tags, s.dump_tag, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
  }
  const int zero = 0;
  HIPCHK(h, hipMemcpy(s.dump_count, &zero, sizeof(int), hipMemcpyHostToDevice));
  *count = n;
  return 0;
}

int mcf_schur_gmres(mcf_handle* h, int kind, int n, const double* rhs, const double* x0, double rtol,
                    int max_iter, int restart, double* y, int* iters, double* relres) {
  if (!h || kind < 0 || kind > 2 || n < 0 || !(rtol > 0) || max_iter < 0) return fail(h, 1, "invalid GMRES request");
  const SolveDesc& s = h->hs[kind];
  if (!s.Sth) return fail(h, 1, "Schur system not set");
  const int r = s.r;
  if (r > GM_MAXR) return fail(h, 1, "Schur system too large for the device GMRES");
  if (n == 0) return 0;
  if (!rhs || !y || !iters || !relres) return fail(h, 1, "invalid GMRES buffers");
  const int m = std::max(1, std::min(restart > 0 ? restart : GM_MAXM, std::min(GM_MAXM, r)));
  const int chunk = 256;
  const size_t lds = sizeof(double) * (3 * (size_t)r + 4 * ((size_t)m + 2) + 2 * (size_t)m + GM_W);
  double *db, *dx0 = nullptr, *dx, *dV, *dR, *drr; int* dit;
  const int nc = std::min(n, chunk);
  HIPCHK(h, hipMalloc(&db, (size_t)nc * r * sizeof(double)));
  if (x0) HIPCHK(h, hipMalloc(&dx0, (size_t)nc * r * sizeof(double)));
  HIPCHK(h, hipMalloc(&dx, (size_t)nc * r * sizeof(double)));
  HIPCHK(h, hipMalloc(&dV, (size_t)nc * (m + 1) * r * sizeof(double)));
  HIPCHK(h, hipMalloc(&dR, (size_t)nc * m * m * sizeof(double)));
  HIPCHK(h, hipMalloc(&drr, (size_t)nc * sizeof(double)));
  HIPCHK(h, hipMalloc(&dit, (size_t)nc * sizeof(int)));
  int rc = 0;
  for (int c0 = 0; c0 < n && !rc; c0 += chunk) {
    const int k = std::min(chunk, n - c0);
    const size_t off = (size_t)c0 * r;
    if (hipMemcpy(db, rhs + off, (size_t)k * r * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        (x0 && hipMemcpy(dx0, x0 + off, (size_t)k * r * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
      rc = fail(h, 4, "GMRES upload failed");
      break;
    }
    GmresArgs a{s.Sth, r, m, max_iter, rtol, db, dx0, dx, dit, drr, dV, dR};
    hipLaunchKernelGGL(k_gmres, dim3(k), dim3(GM_T), lds, h->stream, a);
    if ((rc = launch_check(h, "gmres"))) break;
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(y + off, dx, (size_t)k * r * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(iters + c0, dit, (size_t)k * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(relres + c0, drr, (size_t)k * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(h, 4, "GMRES failed on the device");
  }
  (void)hipFree(db); (void)hipFree(dx0); (void)hipFree(dx); (void)hipFree(dV); (void)hipFree(dR);
  (void)hipFree(drr); (void)hipFree(dit);
  return rc;
}

}  // extern "C"
