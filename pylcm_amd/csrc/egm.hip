// DC-EGM with the FUES upper envelope on gfx950: one workgroup (4 waves) per
// (discrete state, discrete action) column, all arithmetic in FP64.
//
// Phases of egm_column_kernel, separated by workgroup barriers:
//   1. continuation + Euler inversion, one savings node per lane
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
__device__ __forceinline__ bool isfin(double x) { return x - x == 0.0; }
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

// x NaN -> NaN; valid 0 -> NaN; valid 1 -> fp[0]; else the bracket of
// searchsorted(side="right") clipped to [1, valid - 1].
__device__ double read_row(double x, const double* __restrict__ xp, const double* __restrict__ fp,
                           const double* __restrict__ sp, int valid) {
  if (isnan_(x) || valid == 0) return qnan();
  if (valid == 1) return fp[0];
  int lo = 0, hi = valid;  // first index with xp > x
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (xp[m] <= x) lo = m + 1;
    else hi = m;
  }
  int up = lo < 1 ? 1 : (lo > valid - 1 ? valid - 1 : lo);
  const int dn = up - 1;
  if (sp) return between(x, xp[dn], xp[up], fp[dn], fp[up], true, sp[dn], sp[up]);
  return between(x, xp[dn], xp[up], fp[dn], fp[up], false, 0.0, 0.0);
}

// ---------------------------------------------------------------------------
// FUES helpers (fues.py:1082-1189)
// ---------------------------------------------------------------------------

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

  __shared__ double sh_ev0;
  __shared__ int sh_live;
  __shared__ int sh_nkept;
  if (tid == 0) sh_live = 0;

  // ---- 1. continuation and Euler inversion per savings node ----------------
  for (int j = tid; j < n_s; j += kThreads) {
    double EV = 0.0, EMU = 0.0;
    for (int t = 0; t < C.n_targets; ++t) {
      const double pr = C.tprob[t];
      if (C.tkind[t] == 0) {
        EV = EV + (pr > 0.0 ? pr * C.tscalar[t] : pr * 0.0);
        continue;
      }
      const double Rq = C.tR[t][j];
      const double gR = C.tG[t][j];
      const int K = C.tK[t], N = C.tN[t];
      const double scale = C.tscale[t];
      double vs = 0.0, ms = 0.0;
      for (int k = 0; k < K; ++k) {
        const LcmEgmRow* rows = C.tRows[t] + (size_t)k * N;
        double vals[LCM_EGM_MAX_ROWS], mars[LCM_EGM_MAX_ROWS];
        double m = ninf();
        bool any_nan = false;
        int arg = -1;
#pragma unroll
        for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d) {
          if (d < N) {
            const LcmEgmRow& rw = rows[d];
            const int valid = *rw.n_valid;
            const double val = read_row(Rq, rw.grid, rw.value, rw.marg, valid);
            double mar = read_row(Rq, rw.grid, rw.marg, nullptr, valid) * gR;
            if (isneginf_(val)) mar = 0.0;
            vals[d] = val;
            mars[d] = mar;
            if (isnan_(val)) {
              if (!any_nan) arg = d;
              any_nan = true;
            } else if (!any_nan && (arg < 0 || val > m)) {
              m = val;
              arg = d;
            }
          }
        }
        if (arg < 0) arg = 0;
        if (any_nan) m = qnan();
        double sv, sm;
        if (scale > 0.0) {
          const double fm = isneginf_(m) ? 0.0 : m;
          double sum = 0.0, shmax = ninf();
          bool all_masked = true;
#pragma unroll
          for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d) {
            if (d < N) {
              const double sh = (vals[d] - fm) / scale;
              sum = sum + exp(sh);
              shmax = (isnan_(sh) || isnan_(shmax)) ? qnan() : (sh > shmax ? sh : shmax);
              all_masked = all_masked && isneginf_(vals[d]);
            }
          }
          sv = fm + scale * log(sum);
          double esum = 0.0;
#pragma unroll
          for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d)
            if (d < N) esum = esum + exp((vals[d] - fm) / scale - shmax);
          sm = 0.0;
#pragma unroll
          for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d) {
            if (d < N) {
              const double pd = all_masked ? 0.0 : exp((vals[d] - fm) / scale - shmax) / esum;
              sm = sm + pd * mars[d];
            }
          }
        } else {
          sv = m;
          sm = 0.0;
#pragma unroll
          for (int d = 0; d < LCM_EGM_MAX_ROWS; ++d)
            if (d < N) sm = sm + (d == arg ? 1.0 : 0.0) * mars[d];
        }
        const double w = C.tW[t][k];
        vs = vs + zero_safe(w, sv);
        ms = ms + zero_safe(w, sm);
      }
      EV = EV + (pr > 0.0 ? pr * vs : pr * 0.0);
      EMU = EMU + (pr > 0.0 ? pr * ms : pr * 0.0);
    }
    const double disc = beta * EMU;
    const double target = disc < 0.0 ? qnan() : nmax(disc, eps);
    const double c = tape_eval(C.i_prog, C.i_c, target, nullptr);
    const double Aj = C.A[j];
    const double val = tape_eval(C.u_prog, C.u_c, c, nullptr) + beta * EV;
    cg[n_con + j] = Aj + c;
    cp[n_con + j] = c;
    cv[n_con + j] = val;
    cs[n_con + j] = Aj;
    if (j == 0) sh_ev0 = EV;
  }
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
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int P = kWave;
  while (P < n) P <<= 1;
  LdsRows r;
  r.g = reinterpret_cast<double*>(lds_raw);
  r.p = r.g + P;
  r.v = r.p + P;
  r.s = r.v + P;
  r.idx = reinterpret_cast<int*>(r.s + P);
  for (int i = tid; i < P; i += kThreads) {
    const double g = i < n ? cg[i] : pinf();
    r.g[i] = isnan_(g) ? pinf() : g;
    r.idx[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
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
  const int n_pad = C.n_pad;
  double* og = C.out_grid;
  double* op = C.out_policy;
  double* ov = C.out_value;
  if (tid < kWave) {
    const int lane = tid;
    const double th = C.jump_thresh;
    const int n_scan = C.n_scan > 0 ? C.n_scan : max(n - 1, 1);
    double Gk = r.g[0], Pk = r.p[0], Vk = r.v[0];
    double Gj = Gk, Pj = Pk, Vj = Vk, Sj = r.s[0];
    int total = 0;
    for (int idx = 1; idx < n; ++idx) {
      const double Gi = r.g[idx], Pi = r.p[idx], Vi = r.v[idx], Si = r.s[idx];
      const bool valid = !isnan_(Gi) && !isnan_(Vi);
      const bool switches = (Gj == Gi) ? (Sj != Si) : (fabs(slope(Gj, Gj - Pj, Gi, Gi - Pi)) > th);
      const bool node_cross = valid && switches && Gi == Gj && Vi == Vj;
      const double secant = slope(Gj, Vj, Gi, Vi);
      const double grad_before = slope(Gk, Vk, Gj, Vj);
      int pj_pos, pf_pos;
      find_two_forward(r, n, idx, n_scan, th, Gj, Pj, Gi, Pi, pj_pos, pf_pos);
      const bool jf = pj_pos >= 0;
      const double jg = jf ? r.g[pj_pos] : 0.0, jp = jf ? r.p[pj_pos] : 0.0, jv = jf ? r.v[pj_pos] : 0.0;
      const double sec_i_jseg = slope(Gi, Vi, jg, jv);
      const bool below_j = switches && jf && ((Gi == jg) ? (Vi < jv) : (secant < sec_i_jseg));
      const bool sav_dec = Si < Sj;
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

  // ---- 6. marginal row and publish --------------------------------------------
  const int n_kept = sh_nkept;
  const int n_valid = min(n_kept, n_pad);
  const bool over = n_kept > n_pad;
  const bool all_dead = sh_live == 0;
  double* om = C.out_marg;
  for (int i = tid; i < n_pad; i += kThreads) {
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
  __syncthreads();
  const double g0 = og[0];
  for (int w = tid; w < C.n_w; w += kThreads) {
    const double x = C.W[w];
    const double vint = read_row(x, og, ov, om, n_valid);
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

extern "C" int lcm_egm_launch(lcm_ctx* ctx, const LcmEgmCol* d_cols, int n_cols, size_t lds_bytes) {
  if (!ctx || (!d_cols && n_cols)) return fail(LCM_E_ARG, "lcm_egm_launch: null argument");
  if (n_cols <= 0) return LCM_OK;
  if (lds_bytes > lcm_egm_lds_bytes(LCM_EGM_MAX_CAND))
    return fail(LCM_E_ARG, "lcm_egm_launch: LDS request beyond the candidate limit");
  int dev = 0, max_lds = 0;
  LCM_HIP(hipGetDevice(&dev));
  LCM_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (lds_bytes + 64 > (size_t)max_lds)
    return fail(LCM_E_ARG, "lcm_egm_launch: the candidate rows do not fit in this device's LDS");
  hipLaunchKernelGGL(egm::egm_column_kernel, dim3(n_cols), dim3(egm::kThreads), lds_bytes, ctx->stream, d_cols);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
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
