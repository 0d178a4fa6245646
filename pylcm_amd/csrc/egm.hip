// DC-EGM with the FUES upper envelope on gfx950: one workgroup (4 waves) per
// (discrete state, discrete action) column, all arithmetic in FP64.
//
// Phases, separated by workgroup barriers (phase 1 is its own kernel,
// egm_euler_kernel, spread over blocks of 64 savings nodes):
//   1. continuation + Euler inversion, one savings node per thread
//      (src/_lcm/egm/continuation.py:323-500,959-1140,1353-1534;
//       src/_lcm/egm/euler.py:18-68; src/_lcm/egm/step_core.py:397-409);
//   2. closed-form constrained candidates and the -inf -> NaN masking
//      (step_core.py:218-266,443-514);
//   3. LDS bitonic sort by (grid asc, value desc, policy desc, index), the
//      order of the reference's stable lexsort (fues.py:241-245);
//   4. coincident-abscissa reduction, one candidate per lane (fues.py:341-424);
//   5. the FUES scan on wave 0: the carry (k, j) is wave-uniform, the
//      same-segment searches test 64 candidates per step with a ballot
//      (fues.py:640-966,1016-1189);
//   6. marginal row u'(policy), publish of V on the state grid with the
//      Hermite read and the constrained floor (step_core.py:517-616).
// Everything the scan touches sits in LDS (4 sorted rows of FP64).
//
// Floating-point contraction is off in this file so that every product and
// sum rounds as in the FP64 restatement the kernel is checked against.

#pragma clang fp contract(off)

#include "common.h"

namespace lcm {
namespace egm {

constexpr int kThreads = 256;
constexpr int kWave = 64;
constexpr int kTapeRegs = LCM_TAPE_MAX_INS + 1;

// opcodes of pylcm_amd/_tape.py
enum : int {
  kAdd = 1, kSub, kMul, kDiv, kPow, kNeg, kExp, kLog, kSqrt, kMax, kMin, kWhere, kLog1p, kExpm1, kAbs
};

__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000LL); }
__device__ __forceinline__ double pinf() { return __longlong_as_double(0x7ff0000000000000LL); }
__device__ __forceinline__ double ninf() { return __longlong_as_double((long long)0xfff0000000000000ULL); }
__device__ __forceinline__ bool isnan_(double x) { return x != x; }
__device__ __forceinline__ bool isfin(double x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool isneginf_(double x) { return x == ninf(); }

// maximum that propagates NaN (jnp.maximum)
__device__ __forceinline__ double nmax(double a, double b) {
  if (isnan_(a) || isnan_(b)) return qnan();
  return a > b ? a : b;
}

// ---------------------------------------------------------------------------
// expression tape: value and forward derivative at x
// ---------------------------------------------------------------------------

__device__ double tape_eval(const short* __restrict__ prog, const double* __restrict__ cst, double x,
                            double* dout) {
  const int n = prog[0];
  const int out = prog[1];
  double rv[kTapeRegs], rd[kTapeRegs];
  rv[0] = x;
  rd[0] = 1.0;
  for (int i = 0; i < n; ++i) {
    const short* in = prog + 4 + 4 * i;
    const int op = in[0], a = in[1], b = in[2], c = in[3];
    double va = a >= 0 ? rv[a] : cst[-a - 1];
    double da = a >= 0 ? rd[a] : 0.0;
    double vb = 0.0, db = 0.0;
    if (op <= kPow || op == kMax || op == kMin) {
      vb = b >= 0 ? rv[b] : cst[-b - 1];
      db = b >= 0 ? rd[b] : 0.0;
    }
    double v = 0.0, d = 0.0;
    switch (op) {
      case kAdd: v = va + vb; d = da + db; break;
      case kSub: v = va - vb; d = da - db; break;
      case kMul: v = va * vb; d = da * vb + va * db; break;
      case kDiv: v = va / vb; d = da / vb - va * db / (vb * vb); break;
      case kPow:
        v = pow(va, vb);
        d = (da != 0.0 ? vb * pow(va, vb - 1.0) * da : 0.0) + (db != 0.0 ? v * log(va) * db : 0.0);
        break;
      case kNeg: v = -va; d = -da; break;
      case kExp: v = exp(va); d = v * da; break;
      case kLog: v = log(va); d = da / va; break;
      case kSqrt: v = sqrt(va); d = da / (2.0 * v); break;
      case kLog1p: v = log1p(va); d = da / (1.0 + va); break;
      case kExpm1: v = expm1(va); d = (v + 1.0) * da; break;
      case kAbs: v = fabs(va); d = (va > 0.0 ? da : (va < 0.0 ? -da : 0.0)); break;
      case kMax:
      case kMin: {
        const bool pick_a = op == kMax ? (va > vb) : (va < vb);
        const bool tie = va == vb;
        v = (isnan_(va) || isnan_(vb)) ? qnan() : ((pick_a || tie) ? va : vb);
        d = tie ? 0.5 * (da + db) : (pick_a ? da : db);
        break;
      }
      case kWhere: {
        const double m = va;  // operand a is the column mask
        const double v1 = b >= 0 ? rv[b] : cst[-b - 1];
        const double d1 = b >= 0 ? rd[b] : 0.0;
        const double v2 = c >= 0 ? rv[c] : cst[-c - 1];
        const double d2 = c >= 0 ? rd[c] : 0.0;
        v = m != 0.0 ? v1 : v2;
        d = m != 0.0 ? d1 : d2;
        break;
      }
      default: v = qnan(); d = qnan(); break;
    }
    rv[i + 1] = v;
    rd[i + 1] = d;
  }
  if (dout) *dout = out >= 0 ? rd[out] : 0.0;
  return out >= 0 ? rv[out] : cst[-out - 1];
}

// ---------------------------------------------------------------------------
// NaN-padded row reads (src/_lcm/egm/interp.py:460-530,889-985,1237-1309)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double between(double xq, double xl, double xu, double fl, double fu,
                                          bool hermite, double sl, double su) {
  const double w = xu - xl;
  const double safe = w == 0.0 ? 1.0 : w;
  const double raw = (xq - xl) / safe;
  const double rel = w == 0.0 ? 1.0 : (raw < 0.0 ? 0.0 : (raw > 1.0 ? 1.0 : raw));
  const double wl = 1.0 - rel;
  double lin = (wl > 0.0 ? wl * fl : 0.0) + (rel > 0.0 ? rel * fu : 0.0);
  const bool fin = isfin(fl) && isfin(fu);
  if (fin && w != 0.0 && raw < 0.0) lin = lin + (raw - rel) * (fu - fl);
  else lin = lin + 0.0;
  if (!hermite) return lin;
  const double df = fu - fl;
  const double sdf = isfin(df) ? df : 0.0;
  const double sec = sdf / safe;
  const double sg = sec > 0.0 ? 1.0 : (sec < 0.0 ? -1.0 : 0.0);
  const double ml = sl * sec > 0.0 ? sg * fmin(fabs(sl), 3.0 * fabs(sec)) : 0.0;
  const double mu = su * sec > 0.0 ? sg * fmin(fabs(su), 3.0 * fabs(sec)) : 0.0;
  const double a = safe * ml - sdf;
  const double b = sdf - safe * mu;
  const bool ok = w > 0.0 && fin && isfin(sl) && isfin(su);
  const double corr = rel * (1.0 - rel) * ((1.0 - rel) * a + rel * b);
  return lin + (ok ? corr : 0.0);
}

// Number of entries of xp[0, valid) that are <= x (searchsorted side="right"),
// found by galloping from a previous answer `hint` (any value in [0, valid]):
// monotone query sequences resolve in a couple of loads.
__device__ int count_le(const double* __restrict__ xp, int valid, double x, int hint) {
  int L = 0, H = valid;
  const int h = hint < 0 ? 0 : (hint > valid ? valid : hint);
  if (h < valid && xp[h] <= x) {  // the answer is above h: gallop right
    L = h + 1;
    for (int step = 1; L < H; step <<= 1) {
      const int m = min(L + step - 1, H - 1);
      if (xp[m] <= x) {
        L = m + 1;
      } else {
        H = m;
        break;
      }
    }
  } else {  // the answer is at most h: gallop left
    H = h;
    if (h > 0) {
      if (xp[h - 1] <= x) return h;
      H = h - 1;
    }
    for (int step = 1; L < H; step <<= 1) {
      const int m = max(H - step, L);
      if (xp[m] > x) {
        H = m;
      } else {
        L = m + 1;
        break;
      }
    }
  }
  while (L < H) {
    const int m = (L + H) >> 1;
    if (xp[m] <= x) L = m + 1;
    else H = m;
  }
  return L;
}

// A starting point for count_le on a weakly increasing row: x's position
// by linear interpolation between the row's ends (count_le returns the same
// answer from any hint; a close one saves most of its dependent loads).
__device__ __forceinline__ int guess_index(const double* __restrict__ xp, int valid, double x) {
  if (valid < 2 || !isfin(x)) return valid >> 1;
  const double lo = xp[0], hi = xp[valid - 1];
  if (!(hi > lo)) return valid >> 1;
  const double f = (x - lo) / (hi - lo);
  if (!(f > 0.0)) return 0;
  if (!(f < 1.0)) return valid;
  return (int)(f * (valid - 1));
}

// Value (Hermite with the marginal row as node slopes) and marginal (linear)
// reads of one carry row at x, sharing one bracket; `hint` is updated.
__device__ __forceinline__ void read_pair(double x, const LcmEgmRow& rw, int valid, int& hint, double& val,
                                          double& mar) {
  if (isnan_(x) || valid == 0) {
    val = qnan();
    mar = qnan();
    return;
  }
  if (valid == 1) {
    val = rw.value[0];
    mar = rw.marg[0];
    return;
  }
  const int u = count_le(rw.grid, valid, x, hint);
  hint = u;
  const int up = u < 1 ? 1 : (u > valid - 1 ? valid - 1 : u);
  const int dn = up - 1;
  const double xl = rw.grid[dn], xu = rw.grid[up];
  const double sl = rw.marg[dn], su = rw.marg[up];
  val = between(x, xl, xu, rw.value[dn], rw.value[up], true, sl, su);
  mar = between(x, xl, xu, sl, su, false, 0.0, 0.0);
}

// ---------------------------------------------------------------------------
// FUES helpers (fues.py:1082-1189)
// ---------------------------------------------------------------------------

// fabs(a / b) > th, as the rounded quotient compares, deciding by a product
// when the ratio is clearly away from th (b != 0).
__device__ __forceinline__ bool abs_ratio_gt(double a, double b, double th) {
  const double aa = fabs(a), tb = th * fabs(b);
  if (tb > 1e-290 && isfin(tb) && isfin(aa)) {
    if (aa > tb * (1.0 + 1e-14)) return true;
    if (aa < tb * (1.0 - 1e-14)) return false;
  }
  return fabs(a / b) > th;
}

__device__ __forceinline__ double slope(double xa, double ya, double xb, double yb) {
  const double dx = xb - xa;
  return dx == 0.0 ? 0.0 : (yb - ya) / dx;
}

__device__ __forceinline__ bool policy_jump(double ga, double pa, double gb, double pb, double th) {
  const double ia = ga - pa, ib = gb - pb;
  if (ga == gb) return ia != ib;
  return fabs(slope(ga, ia, gb, ib)) > th;
}

__device__ __forceinline__ void intersect(double xa, double ya, double sa, double xb, double yb, double sb,
                                          double& x, double& y) {
  const double den = sa - sb;
  if (den == 0.0) {
    x = qnan();
    y = ya + sa * (xa - xa);
  } else {
    x = (yb - ya + sa * xa - sb * xb) / den;
    y = ya + sa * (x - xa);
  }
}

struct LdsRows {
  double* g;
  double* p;
  double* v;
  double* s;
  int* idx;
};

// bitonic comparator: does candidate slot a sort before slot b?
__device__ __forceinline__ bool sort_before(double ka, int ia, double kb, int ib, int n,
                                            const double* __restrict__ cp, const double* __restrict__ cv) {
  if (ka != kb) return ka < kb;
  if (ia >= n || ib >= n) return ia < ib;
  double va = cv[ia], vb = cv[ib];
  va = isnan_(va) ? ninf() : va;
  vb = isnan_(vb) ? ninf() : vb;
  if (va != vb) return va > vb;
  double pa = cp[ia], pb = cp[ib];
  pa = isnan_(pa) ? ninf() : pa;
  pb = isnan_(pb) ? ninf() : pb;
  if (pa != pb) return pa > pb;
  return ia < ib;
}

// First position after (forward) or before (backward) idx, within the window,
// whose candidate is valid and lies on the anchor's segment, or -1; all lanes of
// the wave return the same answer. The forward search serves two anchors in
// one pass over the candidates.

__device__ void find_two_forward(const LdsRows& r, int n, int idx, int n_scan, double th,
                                 double ag1, double ap1, double ag2, double ap2, int& pos1, int& pos2) {
  const int lane = threadIdx.x & (kWave - 1);
  pos1 = -1;
  pos2 = -1;
  const int last = min(n - 1, idx + n_scan);
  for (int base = idx + 1; base <= last; base += kWave) {
    const int pos = base + lane;
    bool ok = pos <= last;
    double g = 0.0, p = 0.0, v = 0.0;
    if (ok) {
      g = r.g[pos];
      p = r.p[pos];
      v = r.v[pos];
      ok = !isnan_(g) && !isnan_(v);
    }
    const unsigned long long m1 = __ballot(ok && pos1 < 0 && !policy_jump(ag1, ap1, g, p, th));
    const unsigned long long m2 = __ballot(ok && pos2 < 0 && !policy_jump(ag2, ap2, g, p, th));
    if (pos1 < 0 && m1) pos1 = base + __ffsll((long long)m1) - 1;
    if (pos2 < 0 && m2) pos2 = base + __ffsll((long long)m2) - 1;
    if (pos1 >= 0 && pos2 >= 0) break;
  }
}

__device__ int find_backward(const LdsRows& r, int idx, int n_scan, double th, double ag, double ap) {
  const int lane = threadIdx.x & (kWave - 1);
  const int first = max(0, idx - n_scan);
  for (int base = idx - 1; base >= first; base -= kWave) {
    const int pos = base - lane;
    bool ok = pos >= first;
    double g = 0.0, p = 0.0, v = 0.0;
    if (ok) {
      g = r.g[pos];
      p = r.p[pos];
      v = r.v[pos];
      ok = !isnan_(g) && !isnan_(v) && !policy_jump(ag, ap, g, p, th);
    }
    const unsigned long long m = __ballot(ok);
    if (m) return base - (__ffsll((long long)m) - 1);
  }
  return -1;
}

// ---------------------------------------------------------------------------
// the column kernel
// ---------------------------------------------------------------------------

// Aggregates one child block of N rows at one savings node: the logsum with
// its softmax-weighted marginal (scale > 0) or the hard max with the first
// argmax's marginal (continuation.py:1317-1351, logsum.py:32-64).
__device__ __forceinline__ void aggregate_rows(const double* vals, const double* mars, int stride, int N,
                                               double scale, double& sv, double& sm) {
  double m = ninf();
  bool any_nan = false;
  int arg = -1;
  for (int d = 0; d < N; ++d) {
    const double val = vals[d * stride];
    if (isnan_(val)) {
      if (!any_nan) arg = d;
      any_nan = true;
    } else if (!any_nan && (arg < 0 || val > m)) {
      m = val;
      arg = d;
    }
  }
  if (arg < 0) arg = 0;
  if (any_nan) m = qnan();
  if (scale > 0.0) {
    const double fm = isneginf_(m) ? 0.0 : m;
    double sum = 0.0, shmax = ninf();
    bool all_masked = true;
    for (int d = 0; d < N; ++d) {
      const double v = vals[d * stride];
      const double sh = (v - fm) / scale;
      sum = sum + exp(sh);
      shmax = (isnan_(sh) || isnan_(shmax)) ? qnan() : (sh > shmax ? sh : shmax);
      all_masked = all_masked && isneginf_(v);
    }
    sv = fm + scale * log(sum);
    // each row's exp(sh - shmax) once, for the sum and for its weight (the
    // sums run in the same order as two separate passes would)
    double e[LCM_EGM_MAX_ROWS];
    double esum = 0.0;
#pragma unroll
    for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d) {
      if (d < N) {
        e[d] = exp((vals[d * stride] - fm) / scale - shmax);
        esum = esum + e[d];
      }
    }
    sm = 0.0;
#pragma unroll
    for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d) {
      if (d < N) {
        const double pd = all_masked ? 0.0 : e[d] / esum;
        sm = sm + pd * mars[d * stride];
      }
    }
  } else {
    sv = m;
    sm = 0.0;
    for (int d = 0; d < N; ++d) sm = sm + (d == arg ? 1.0 : 0.0) * mars[d * stride];
  }
}

// Savings nodes per block of the Euler kernel: 64, 128 or 256 (at most one
// per thread in the aggregation phase), chosen per launch by the grid size.
constexpr int kNodesMin = 64, kNodesMax = 256;
constexpr int kMarginalBlocks = 16;  // workgroups per column for the marginal row (4096 slots)
constexpr int kPublishBlocks = 8;    // workgroups per column for the V row (2048 wealth nodes)

// Phase 1: continuation and Euler inversion. Block (bx, col) takes savings
// nodes [NODES bx, NODES bx + NODES) of column col; its 256 threads read every
// (child row, node) pair in parallel into LDS, then one thread per node aggregates
// over the child's choice rows, nodes and target regimes and inverts the
// Euler equation. Writes candidate slots n_con + j and, for node 0, E[V'] at
// the savings limit (cand[4n]).
template <int NODES>
__global__ void __launch_bounds__(kThreads) egm_euler_kernel(const LcmEgmCol* __restrict__ cols) {
  const LcmEgmCol& C = cols[blockIdx.y];
  const int tid = threadIdx.x;
  const int n_s = C.n_s, n_con = C.n_con, n = n_s + n_con;
  const int j0 = blockIdx.x * NODES;
  if (j0 >= n_s) return;
  const double beta = C.beta;
  const double eps = 2.220446049250313e-16;
  int Q = 0;
  for (int t = 0; t < C.n_targets; ++t)
    if (C.tkind[t] != 0) Q += C.tK[t] * C.tN[t];
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* vals = reinterpret_cast<double*>(lds_raw);  // [Q][NODES]
  double* mars = vals + (size_t)Q * NODES;
  // A work unit is kRun consecutive savings nodes of one child row: the
  // queries rise with the node, so each bracket search gallops from the
  // previous node's answer instead of starting over. Longer runs pay the
  // first search of a run (from an interpolated guess) less often but leave
  // fewer units to spread over the threads: 8 is 5 % faster than 4 on the
  // 64-draw batch's 256-node blocks, 4 on the small single-solve launches.
  constexpr int kRun = NODES >= 256 ? 8 : 4;
  constexpr int kUnitsPerRow = NODES / kRun;
  for (int u = tid; u < Q * kUnitsPerRow; u += kThreads) {
    const int q = u / kUnitsPerRow, jj0 = (u - q * kUnitsPerRow) * kRun;
    if (j0 + jj0 >= n_s) continue;
    int t = 0, base = 0;
    for (; t < C.n_targets; ++t) {
      if (C.tkind[t] == 0) continue;
      const int rows_t = C.tK[t] * C.tN[t];
      if (q < base + rows_t) break;
      base += rows_t;
    }
    const LcmEgmRow& rw = C.tRows[t][q - base];
    const int valid = *rw.n_valid;
    int hint = guess_index(rw.grid, valid, C.tR[t][j0 + jj0]);
    for (int jj = jj0; jj < jj0 + kRun; ++jj) {
      const int j = j0 + jj;
      if (j >= n_s) break;
      double val, mar;
      read_pair(C.tR[t][j], rw, valid, hint, val, mar);
      mar = mar * C.tG[t][j];
      if (isneginf_(val)) mar = 0.0;
      vals[q * NODES + jj] = val;
      mars[q * NODES + jj] = mar;
    }
  }
  __syncthreads();
  const int j = j0 + tid;
  if (tid >= NODES || j >= n_s) return;
  double EV = 0.0, EMU = 0.0;
  int row0 = 0;
  for (int t = 0; t < C.n_targets; ++t) {
    const double pr = C.tprob[t];
    if (C.tkind[t] == 0) {
      EV = EV + (pr > 0.0 ? pr * C.tscalar[t] : pr * 0.0);
      continue;
    }
    const int K = C.tK[t], N = C.tN[t];
    double vs = 0.0, ms = 0.0;
    for (int k = 0; k < K; ++k) {
      const int off = (row0 + k * N) * NODES + tid;
      double sv, sm;
      aggregate_rows(vals + off, mars + off, NODES, N, C.tscale[t], sv, sm);
      const double w = C.tW[t][k];
      vs = vs + zero_safe(w, sv);
      ms = ms + zero_safe(w, sm);
    }
    row0 += K * N;
    EV = EV + (pr > 0.0 ? pr * vs : pr * 0.0);
    EMU = EMU + (pr > 0.0 ? pr * ms : pr * 0.0);
  }
  const double disc = beta * EMU;
  const double target = disc < 0.0 ? qnan() : nmax(disc, eps);
  const double c = tape_eval(C.i_prog, C.i_c, target, nullptr);
  const double Aj = C.A[j];
  const double val = tape_eval(C.u_prog, C.u_c, c, nullptr) + beta * EV;
  double* cg = C.cand;
  cg[n_con + j] = Aj + c;
  cg[n + n_con + j] = c;
  cg[2 * n + n_con + j] = val;
  cg[3 * n + n_con + j] = Aj;
  if (j == 0) cg[4 * n] = EV;
}

// Phases 2-6 of one column (after egm_euler_kernel).
__global__ void __launch_bounds__(kThreads) egm_column_kernel(const LcmEgmCol* __restrict__ cols) {
  const LcmEgmCol& C = cols[blockIdx.x];
  const int tid = threadIdx.x;
  const int n_s = C.n_s, n_con = C.n_con, n = n_s + n_con;
  double* cg = C.cand;
  double* cp = cg + n;
  double* cv = cp + n;
  double* cs = cv + n;
  const double beta = C.beta;
  const double eps = 2.220446049250313e-16;

  __shared__ int sh_live;
  __shared__ int sh_nkept;
  if (tid == 0) sh_live = 0;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int P = kWave;
  while (P < n) P <<= 1;
  const double sh_ev0 = cg[4 * n];
  __syncthreads();

  // ---- 2. constrained candidates, dead masking ------------------------------
  const double limit = C.A[0];
  const double dev0 = beta * sh_ev0;
  {
    const double span = fmin(cg[n_con], C.w_max) - limit;
    // fmin would drop a NaN first endogenous point; the reference's minimum keeps it
    const double span_nan = isnan_(cg[n_con]) ? qnan() : span;
    const bool empty = !(span_nan > 0.0 && isfin(span_nan));
    const double safe = empty ? 1.0 : span_nan;
    for (int k = tid; k < n_con; k += kThreads) {
      const double a = safe * 1e-4 * pow(C.ratio, (double)k);
      const double v = tape_eval(C.u_prog, C.u_c, a, nullptr) + dev0;
      cg[k] = empty ? qnan() : limit + a;
      cp[k] = empty ? qnan() : a;
      cv[k] = empty ? qnan() : v;
      cs[k] = limit;
    }
  }
  __syncthreads();
  {
    int live = 0;
    for (int i = tid; i < n; i += kThreads) {
      if (isneginf_(cv[i])) {
        cg[i] = qnan();
        cp[i] = qnan();
        cv[i] = qnan();
        cs[i] = qnan();
      } else {
        live = 1;
      }
    }
    if (live) atomicOr(&sh_live, 1);
  }
  __syncthreads();

  // ---- 3. sort in LDS ---------------------------------------------------------
  LdsRows r;
  r.g = reinterpret_cast<double*>(lds_raw);
  r.p = r.g + P;
  r.v = r.p + P;
  r.s = r.v + P;
  r.idx = reinterpret_cast<int*>(r.s + P);
  // candidates usually arrive sorted (the endogenous grid rises with the
  // savings node away from kinks): skip the network when they do
  int unsorted = 0;
  for (int i = tid; i + 1 < n; i += kThreads) {
    const double a = cg[i], b = cg[i + 1];
    if (!sort_before(isnan_(a) ? pinf() : a, i, isnan_(b) ? pinf() : b, i + 1, n, cp, cv)) unsorted = 1;
  }
  for (int i = tid; i < P; i += kThreads) {
    const double g = i < n ? cg[i] : pinf();
    r.g[i] = isnan_(g) ? pinf() : g;
    r.idx[i] = i;
  }
  const bool need_sort = __syncthreads_or(unsorted) != 0;
  for (int k = 2; need_sort && k <= P; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int i = tid; i < P; i += kThreads) {
        const int l = i ^ jj;
        if (l > i) {
          const bool asc = (i & k) == 0;
          const double ki = r.g[i], kl = r.g[l];
          const int ii = r.idx[i], il = r.idx[l];
          const bool swap = asc ? sort_before(kl, il, ki, ii, n, cp, cv) : sort_before(ki, ii, kl, il, n, cp, cv);
          if (swap) {
            r.g[i] = kl;
            r.g[l] = ki;
            r.idx[i] = il;
            r.idx[l] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < n; i += kThreads) {
    const int o = r.idx[i];
    r.g[i] = cg[o];
    r.p[i] = cp[o];
    r.v[i] = cv[o];
    r.s[i] = cs[o];
  }
  __syncthreads();

  // ---- 4. coincident-abscissa reduction -------------------------------------
  {
    bool keep_reg[(LCM_EGM_MAX_CAND + kThreads - 1) / kThreads];
#pragma unroll
    for (int q = 0; q < (LCM_EGM_MAX_CAND + kThreads - 1) / kThreads; ++q) {
      const int i = tid + q * kThreads;
      bool keep = false;
      if (i < n) {
        const double gi = r.g[i];
        if (!isnan_(gi)) {
          int s0 = i, e0 = i;
          while (s0 > 0 && r.g[s0 - 1] == gi) --s0;
          while (e0 < n - 1 && r.g[e0 + 1] == gi) ++e0;
          double gm = ninf();
          for (int q2 = s0; q2 <= e0; ++q2) gm = nmax(gm, r.v[q2]);
          const bool is_max = r.v[i] == gm;
          bool collapse = false;
          if (is_max) {
            for (int q2 = s0; q2 < i; ++q2)
              if (r.v[q2] == gm && r.s[q2] == r.s[i]) {
                collapse = true;
                break;
              }
          }
          keep = is_max && !collapse;
        }
      }
      keep_reg[q] = keep;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < (LCM_EGM_MAX_CAND + kThreads - 1) / kThreads; ++q) {
      const int i = tid + q * kThreads;
      if (i < n && !keep_reg[q]) {
        r.g[i] = qnan();
        r.p[i] = qnan();
        r.v[i] = qnan();
        r.s[i] = qnan();
      }
    }
  }
  __syncthreads();

  // ---- 5. FUES scan on wave 0 -------------------------------------------------
  // simple[i]: with k = candidate i-2 and j = candidate i-1 in the carry, step i
  // is a plain accept (no segment switch, not dropped). The scan jumps over runs
  // of such steps, found 64 at a time, whenever the carry is in that state.
  const double th = C.jump_thresh;
  int* simple = r.idx;
  for (int i = tid; i < n; i += kThreads) {
    bool ok = false;
    if (i >= 2) {
      const double Gi = r.g[i], Pi = r.p[i], Vi = r.v[i], Si = r.s[i];
      const double Gj = r.g[i - 1], Pj = r.p[i - 1], Vj = r.v[i - 1], Sj = r.s[i - 1];
      const bool valid = !isnan_(Gi) && !isnan_(Vi);
      const bool sw = (Gj == Gi) ? (Sj != Si) : abs_ratio_gt((Gi - Pi) - (Gj - Pj), Gi - Gj, th);
      bool dropped = !valid || Vi < Vj;
      if (!dropped && Si < Sj) dropped = slope(Gj, Vj, Gi, Vi) < slope(r.g[i - 2], r.v[i - 2], Gj, Vj);
      ok = !sw && !dropped;
    }
    simple[i] = ok ? 1 : 0;
  }
  __syncthreads();
  const int n_pad = C.n_pad;
  double* og = C.out_grid;
  double* op = C.out_policy;
  double* ov = C.out_value;
  if (tid < kWave) {
    const int lane = tid;
    const int n_scan = C.n_scan > 0 ? C.n_scan : max(n - 1, 1);
    double Gk = r.g[0], Pk = r.p[0], Vk = r.v[0];
    double Gj = Gk, Pj = Pk, Vj = Vk, Sj = r.s[0];
    int total = 0;
    int jidx = 0;        // candidate held in j
    bool kcanon = false;  // k holds candidate jidx - 1
    double nG = n > 1 ? r.g[1] : 0.0, nP = n > 1 ? r.p[1] : 0.0, nV = n > 1 ? r.v[1] : 0.0;
    double nS = n > 1 ? r.s[1] : 0.0;
    for (int idx = 1; idx < n; ++idx) {
      if (kcanon && jidx == idx - 1) {
        int end = n;
        for (int base = idx; base < n; base += kWave) {
          const int pos = base + lane;
          const unsigned long long stop = __ballot(pos < n && simple[pos] == 0);
          if (stop) {
            end = base + __ffsll((long long)stop) - 1;
            break;
          }
        }
        const int cnt = end - idx;
        if (cnt > 0) {
          // steps idx .. end-1 each finalise the previous candidate
          for (int l = lane; l < cnt; l += kWave) {
            const int o = total + l;
            if (o < n_pad) {
              og[o] = r.g[idx - 1 + l];
              op[o] = r.p[idx - 1 + l];
              ov[o] = r.v[idx - 1 + l];
            }
          }
          total += cnt;
          Gk = r.g[end - 2]; Pk = r.p[end - 2]; Vk = r.v[end - 2];
          Gj = r.g[end - 1]; Pj = r.p[end - 1]; Vj = r.v[end - 1]; Sj = r.s[end - 1];
          jidx = end - 1;
          idx = end;
          if (idx >= n) break;
          nG = r.g[idx]; nP = r.p[idx]; nV = r.v[idx]; nS = r.s[idx];
        }
      }
      const double Gi = nG, Pi = nP, Vi = nV, Si = nS;
      if (idx + 1 < n) {  // next candidate's loads overlap this step's arithmetic
        nG = r.g[idx + 1];
        nP = r.p[idx + 1];
        nV = r.v[idx + 1];
        nS = r.s[idx + 1];
      }
      const bool valid = !isnan_(Gi) && !isnan_(Vi);
      const bool switches = (Gj == Gi) ? (Sj != Si) : abs_ratio_gt((Gi - Pi) - (Gj - Pj), Gi - Gj, th);
      const bool sav_dec = Si < Sj;
      if (!switches) {
        // Without a segment switch there is no node crossing, no dominated j and
        // no crossing to insert: the step either drops i or finalises j.
        bool dropped = !valid || Vi < Vj;
        if (!dropped && sav_dec) dropped = slope(Gj, Vj, Gi, Vi) < slope(Gk, Vk, Gj, Vj);
        if (!dropped) {
          if (lane == 0 && total < n_pad) {
            og[total] = Gj;
            op[total] = Pj;
            ov[total] = Vj;
          }
          total += 1;
          Gk = Gj; Pk = Pj; Vk = Vj;
          Gj = Gi; Pj = Pi; Vj = Vi; Sj = Si;
          kcanon = jidx == idx - 1;
          jidx = idx;
        }
        continue;
      }
      const double secant = slope(Gj, Vj, Gi, Vi);
      const double grad_before = slope(Gk, Vk, Gj, Vj);
      const bool node_cross = valid && switches && Gi == Gj && Vi == Vj;
      int pj_pos, pf_pos;
      find_two_forward(r, n, idx, n_scan, th, Gj, Pj, Gi, Pi, pj_pos, pf_pos);
      const bool jf = pj_pos >= 0;
      const double jg = jf ? r.g[pj_pos] : 0.0, jp = jf ? r.p[pj_pos] : 0.0, jv = jf ? r.v[pj_pos] : 0.0;
      const double sec_i_jseg = slope(Gi, Vi, jg, jv);
      const bool below_j = switches && jf && ((Gi == jg) ? (Vi < jv) : (secant < sec_i_jseg));
      const bool dropped = (!valid || (Vi < Vj && !switches) || (sav_dec && secant < grad_before) || below_j) &&
                           !node_cross;
      int pb_pos = -1;
      if (pf_pos < 0) pb_pos = find_backward(r, idx, n_scan, th, Gi, Pi);
      const bool partner = pf_pos >= 0 || pb_pos >= 0;
      const int pp = pf_pos >= 0 ? pf_pos : pb_pos;
      const double pg = partner ? r.g[pp] : 0.0, ppol = partner ? r.p[pp] : 0.0, pv = partner ? r.v[pp] : 0.0;
      const double i_slope = slope(Gi, Vi, pg, pv);
      const double i_pslope = slope(Gi, Pi, pg, ppol);
      const bool j_dom = !dropped && switches && partner && secant > i_slope && !node_cross;
      const double pslope_kj = slope(Gk, Pk, Gj, Pj);
      double kg6, kv6;
      intersect(Gj, Vj, grad_before, Gi, Vi, i_slope, kg6, kv6);
      const double kpl6 = Pj + pslope_kj * (kg6 - Gj);
      const double kpr6 = Pi + i_pslope * (kg6 - Gi);
      const bool kink6 = j_dom && kg6 > Gk && kg6 < Gj && kg6 < Gi;
      const double j_slope = jf ? slope(Gj, Vj, jg, jv) : grad_before;
      const double j_pslope = jf ? slope(Gj, Pj, jg, jp) : pslope_kj;
      double kg5, kv5;
      intersect(Gj, Vj, j_slope, Gi, Vi, i_slope, kg5, kv5);
      const double kpl5 = Pj + j_pslope * (kg5 - Gj);
      const double kpr5 = Pi + i_pslope * (kg5 - Gi);
      const bool crossing = !dropped && !j_dom && !node_cross && switches && partner;
      const double tol = 16.0 * eps * fmax(fmax(fabs(Gj), fabs(Gi)), 1.0);
      const bool on_j = crossing && fabs(kg5 - Gj) <= tol;
      const bool on_i = crossing && !on_j && fabs(kg5 - Gi) <= tol;
      const bool kink5 = crossing && !on_j && !on_i && kg5 > Gj && kg5 < Gi;
      const bool plain = !dropped && !j_dom && !(kink5 || on_j || on_i);
      // emission (fues.py:549-637): up to three rows, lanes 0..2 write one each
      int cnt = 0;
      double eg = 0.0, ep = 0.0, ev = 0.0;
      if (kink6) {
        cnt = 2;
        eg = kg6;
        ep = lane == 0 ? kpl6 : kpr6;
        ev = kv6;
      } else if (kink5) {
        cnt = 3;
        eg = lane == 0 ? Gj : kg5;
        ep = lane == 0 ? Pj : (lane == 1 ? kpl5 : kpr5);
        ev = lane == 0 ? Vj : kv5;
      } else if (on_j) {
        cnt = 2;
        eg = Gj;
        ep = lane == 0 ? Pj : kpr5;
        ev = lane == 0 ? Vj : kv5;
      } else if (on_i) {
        cnt = 2;
        eg = lane == 0 ? Gj : Gi;
        ep = lane == 0 ? Pj : kpl5;
        ev = lane == 0 ? Vj : kv5;
      } else if (plain) {
        cnt = 1;
        eg = Gj;
        ep = Pj;
        ev = Vj;
      }
      if (lane < cnt && total + lane < n_pad) {
        og[total + lane] = eg;
        op[total + lane] = ep;
        ov[total + lane] = ev;
      }
      total += cnt;
      if (!dropped) {
        const bool keeps_k = j_dom && !kink6;
        kcanon = !keeps_k && !kink5 && !kink6 && !on_j && jidx == idx - 1;
        jidx = idx;
        if (!keeps_k) {
          if (kink5) {
            Gk = kg5; Pk = kpr5; Vk = kv5;
          } else if (kink6) {
            Gk = kg6; Pk = kpr6; Vk = kv6;
          } else if (on_j) {
            Gk = Gj; Pk = kpr5; Vk = kv5;
          } else {
            Gk = Gj; Pk = Pj; Vk = Vj;
          }
        }
        Gj = Gi; Pj = Pi; Vj = Vi; Sj = Si;
      }
    }
    if (!isnan_(Gj) && !isnan_(Vj)) {
      if (lane == 0 && total < n_pad) {
        og[total] = Gj;
        op[total] = Pj;
        ov[total] = Vj;
      }
      total += 1;
    }
    if (lane == 0) {
      sh_nkept = total;
      C.out_counts[0] = total;
      C.out_counts[1] = min(total, n_pad);
    }
  }
  __syncthreads();

  // the marginal row and the column's V on the wealth grid are filled by
  // egm_marginal_kernel and egm_publish_kernel, many workgroups per column
  if (tid == 0) cg[4 * n + 1] = sh_live == 0 ? 1.0 : 0.0;
}

// Phase 6a of a column: marginal-utility row of the kept points, NaN past them.
__global__ void __launch_bounds__(kThreads) egm_marginal_kernel(const LcmEgmCol* __restrict__ cols) {
  const LcmEgmCol& C = cols[blockIdx.y];
  const int n = C.n_s + C.n_con;
  const double* cg = C.cand;
  const int n_pad = C.n_pad;
  const int n_kept = C.out_counts[0];
  const int n_valid = min(n_kept, n_pad);
  const bool over = n_kept > n_pad;
  const bool all_dead = cg[4 * n + 1] != 0.0;
  double* og = C.out_grid;
  double* op = C.out_policy;
  double* ov = C.out_value;
  double* om = C.out_marg;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n_pad; i += gridDim.x * kThreads) {
    if (i >= n_valid) {
      og[i] = qnan();
      op[i] = qnan();
      ov[i] = qnan();
    }
    const double pol = op[i];
    double mu = qnan();
    if (!isnan_(pol)) tape_eval(C.u_prog, C.u_c, pol, &mu);
    if (isneginf_(ov[i])) mu = 0.0;
    if (over) {
      ov[i] = qnan();
      mu = qnan();
    }
    if (all_dead) {
      ov[i] = ninf();
      mu = 0.0;
    }
    om[i] = mu;
  }
}

// Phase 6b of a column: its value on the wealth grid (interpolated envelope,
// or the constrained branch below the first kept point).
__global__ void __launch_bounds__(kThreads) egm_publish_kernel(const LcmEgmCol* __restrict__ cols) {
  const LcmEgmCol& C = cols[blockIdx.y];
  const int n = C.n_s + C.n_con;
  const double* cg = C.cand;
  const int n_pad = C.n_pad;
  const int n_kept = C.out_counts[0];
  const int n_valid = min(n_kept, n_pad);
  const bool over = n_kept > n_pad;
  const bool all_dead = cg[4 * n + 1] != 0.0;
  const double limit = C.A[0];
  const double dev0 = C.beta * cg[4 * n];
  const double* og = C.out_grid;
  const double* ov = C.out_value;
  const double* om = C.out_marg;
  const double g0 = og[0];
  for (int w = blockIdx.x * kThreads + threadIdx.x; w < C.n_w; w += gridDim.x * kThreads) {
    const double x = C.W[w];
    double vint;
    if (isnan_(x) || n_valid == 0) {
      vint = qnan();
    } else if (n_valid == 1) {
      vint = ov[0];
    } else {
      // a thread's wealth nodes are a grid stride apart: a fresh guess each time
      const int u = count_le(og, n_valid, x, guess_index(og, n_valid, x));
      const int up = u < 1 ? 1 : (u > n_valid - 1 ? n_valid - 1 : u);
      vint = between(x, og[up - 1], og[up], ov[up - 1], ov[up], true, om[up - 1], om[up]);
    }
    const double cf = x - limit;
    const double tiny = 2.2250738585072014e-308;
    const double vcon = cf > 0.0 ? tape_eval(C.u_prog, C.u_c, fmax(cf, tiny), nullptr) + dev0 : ninf();
    double V = (cf > 0.0 && x <= g0) ? vcon : nmax(vint, vcon);
    if (over) V = qnan();
    if (all_dead) V = ninf();
    C.out_V[w] = V;
  }
}

// logsum / max over a discrete-state combo's columns, written in the model dtype
template <typename T>
__global__ void __launch_bounds__(256) egm_aggregate_kernel(const LcmEgmAgg* __restrict__ aggs) {
  const LcmEgmAgg& A = aggs[blockIdx.y];
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= A.n_w) return;
  const double* c = A.cols;
  double m = c[w];
  for (int d = 1; d < A.n_d; ++d) m = nmax(m, c[(size_t)d * A.n_w + w]);
  double V = m;
  if (A.scale > 0.0) {
    const double fm = isneginf_(m) ? 0.0 : m;
    double sum = 0.0;
    for (int d = 0; d < A.n_d; ++d) sum = sum + exp((c[(size_t)d * A.n_w + w] - fm) / A.scale);
    V = fm + A.scale * log(sum);
  }
  static_cast<T*>(A.out)[w] = static_cast<T>(V);
}

__global__ void tape_eval_kernel(const short* prog, const double* c, const double* x, double* v, double* d,
                                 int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double dd = 0.0;
  v[i] = tape_eval(prog, c, x[i], &dd);
  d[i] = dd;
}

}  // namespace egm
}  // namespace lcm

using namespace lcm;

extern "C" size_t lcm_egm_lds_bytes(int n_cand) {
  int P = egm::kWave;
  while (P < n_cand) P <<= 1;
  return (size_t)P * (4 * sizeof(double) + sizeof(int));
}

extern "C" int lcm_egm_launch_ex(lcm_ctx* ctx, const LcmEgmCol* d_cols, int n_cols, size_t lds_bytes,
                                 int max_rows, int max_s, int nodes_per_block) {
  if (!ctx || (!d_cols && n_cols)) return fail(LCM_E_ARG, "lcm_egm_launch: null argument");
  if (nodes_per_block != 0 && nodes_per_block != 64 && nodes_per_block != 128 && nodes_per_block != 256)
    return fail(LCM_E_ARG, "lcm_egm_launch: nodes_per_block must be 0 (chosen per launch), 64, 128 or 256");
  if (n_cols <= 0) return LCM_OK;
  if (n_cols > 65535) return fail(LCM_E_ARG, "lcm_egm_launch: more than 65535 columns in one launch");
  if (lds_bytes > lcm_egm_lds_bytes(LCM_EGM_MAX_CAND))
    return fail(LCM_E_ARG, "lcm_egm_launch: LDS request beyond the candidate limit");
  int dev = 0, max_lds = 0;
  LCM_HIP(hipGetDevice(&dev));
  LCM_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (lds_bytes + 64 > (size_t)max_lds)
    return fail(LCM_E_ARG, "lcm_egm_launch: the candidate rows do not fit in this device's LDS");
  if (max_rows < 0 || max_rows > LCM_EGM_MAX_ROWS_PER_COL || max_s < 1 || max_s > LCM_EGM_MAX_CAND)
    return fail(LCM_E_ARG, "lcm_egm_launch: row or node count out of range");
  // Nodes per Euler block: more per block when the launch has workgroups to
  // spare (the aggregation phase then keeps more of the block's threads busy),
  // 64 when the launch is small (a single solve: more, shorter blocks).
  int nodes = egm::kNodesMin;
  if (nodes_per_block) {
    nodes = nodes_per_block;
    if ((size_t)max_rows * nodes * 2 * sizeof(double) + 64 > (size_t)max_lds)
      return fail(LCM_E_ARG, "lcm_egm_launch: the continuation rows of that many nodes do not fit in LDS");
  } else {
    while (nodes < egm::kNodesMax && (size_t)max_rows * (2 * nodes) * 2 * sizeof(double) + 64 <= (size_t)max_lds &&
           (size_t)n_cols * ((max_s + 2 * nodes - 1) / (2 * nodes)) >= 4096)
      nodes *= 2;
  }
  const size_t lds_euler = (size_t)max_rows * nodes * 2 * sizeof(double);
  const dim3 egrid((max_s + nodes - 1) / nodes, n_cols);
  if (nodes == 256)
    hipLaunchKernelGGL(egm::egm_euler_kernel<256>, egrid, dim3(egm::kThreads), lds_euler, ctx->stream, d_cols);
  else if (nodes == 128)
    hipLaunchKernelGGL(egm::egm_euler_kernel<128>, egrid, dim3(egm::kThreads), lds_euler, ctx->stream, d_cols);
  else
    hipLaunchKernelGGL(egm::egm_euler_kernel<64>, egrid, dim3(egm::kThreads), lds_euler, ctx->stream, d_cols);
  LCM_HIP(hipGetLastError());
  hipLaunchKernelGGL(egm::egm_column_kernel, dim3(n_cols), dim3(egm::kThreads), lds_bytes, ctx->stream, d_cols);
  LCM_HIP(hipGetLastError());
  hipLaunchKernelGGL(egm::egm_marginal_kernel, dim3(egm::kMarginalBlocks, n_cols), dim3(egm::kThreads), 0,
                     ctx->stream, d_cols);
  LCM_HIP(hipGetLastError());
  hipLaunchKernelGGL(egm::egm_publish_kernel, dim3(egm::kPublishBlocks, n_cols), dim3(egm::kThreads), 0,
                     ctx->stream, d_cols);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

extern "C" int lcm_egm_launch(lcm_ctx* ctx, const LcmEgmCol* d_cols, int n_cols, size_t lds_bytes, int max_rows,
                              int max_s) {
  return lcm_egm_launch_ex(ctx, d_cols, n_cols, lds_bytes, max_rows, max_s, 0);
}

extern "C" int lcm_egm_aggregate(lcm_ctx* ctx, int dtype, const LcmEgmAgg* d_aggs, int n_aggs, int max_w) {
  if (!ctx || (!d_aggs && n_aggs)) return fail(LCM_E_ARG, "lcm_egm_aggregate: null argument");
  if (n_aggs <= 0 || max_w <= 0) return LCM_OK;
  if (n_aggs > 65535) return fail(LCM_E_ARG, "lcm_egm_aggregate: too many combos in one launch");
  dim3 grid((max_w + 255) / 256, n_aggs);
  if (dtype == LCM_DTYPE_F64)
    hipLaunchKernelGGL(egm::egm_aggregate_kernel<double>, grid, dim3(256), 0, ctx->stream, d_aggs);
  else if (dtype == LCM_DTYPE_F32)
    hipLaunchKernelGGL(egm::egm_aggregate_kernel<float>, grid, dim3(256), 0, ctx->stream, d_aggs);
  else
    return fail(LCM_E_ARG, "lcm_egm_aggregate: unknown dtype");
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

extern "C" int lcm_abi_sizes_egm(size_t* row, size_t* col, size_t* agg) {
  if (row) *row = sizeof(LcmEgmRow);
  if (col) *col = sizeof(LcmEgmCol);
  if (agg) *agg = sizeof(LcmEgmAgg);
  return LCM_OK;
}

extern "C" int lcm_tape_eval(lcm_ctx* ctx, const short* d_prog, const double* d_c, const double* d_x,
                             double* d_v, double* d_d, int n) {
  if (!ctx || !d_prog || !d_c || ((!d_x || !d_v || !d_d) && n > 0))
    return fail(LCM_E_ARG, "lcm_tape_eval: null argument");
  if (n <= 0) return LCM_OK;
  hipLaunchKernelGGL(egm::tape_eval_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_prog, d_c, d_x,
                     d_v, d_d, n);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}
