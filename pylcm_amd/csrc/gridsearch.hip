// Fused grid-search period kernel for gfx950.
//
// One launch evaluates a whole backward step for any number of (problem,
// regime) descriptors: for every state it contracts the next-period value rows
// into EV, interpolates EV at the next continuous state of every action, adds
// utility, masks infeasible actions and reduces to the max and the first
// argmax. It stands in for the compiled core `max_Q_over_a(Q_and_F)`
// (src/_lcm/regime_building/max_Q_over_a.py:30-191,
// src/_lcm/regime_building/Q_and_F.py:58-170), the interpolation
// `map_coordinates` (src/_lcm/regime_building/ndimage.py:28-109) on the grid
// coordinates of src/_lcm/grids/coordinates.py:30-167, the expectation
// (Q_and_F.py:1145-1209), `argmax_and_max` (src/_lcm/regime_building/argmax.py:6-69)
// and, with taste shocks, `logsum_and_softmax` (src/_lcm/logsum.py:32-64).
//
// Work decomposition: a workgroup of 256 threads owns TILE consecutive states
// of one (descriptor, discrete state). Lanes map to states; the 256/TILE waves
// that share a state interleave over the actions and are merged through LDS
// at the end (ties keep the smaller flat index, so the result is the first
// argmax whatever the interleave). Per discrete action the block stages the
// contracted EV row (next-period V rows weighted by M) in LDS, then walks the
// actions in chunks of kActChunk. The action loop index is wave-uniform: in
// the common (folded) loop every per-action operand {A', B', U, bounds} is a
// scalar load from the global tables into SGPRs, so the only LDS traffic per
// action is each lane's 8-byte gather of (EV_k, slope_k); the other loops stage
// the chunk's operands {A, B, U, lo} (+ hi, or affine constraint rows) in LDS
// and read them as broadcasts. Block-uniform conditions (continuation kind,
// finite EV row, operand magnitudes) select a specialised action loop outside
// the loop. Blocks are remapped so that consecutive tasks (same problem, same
// V rows) land on the same XCD and share its L2.

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace lcm {

constexpr int kBlock = 256;
constexpr int kXcds = 8;
constexpr int kActChunk = 256;  // actions staged in LDS at a time

// Kernel variants, chosen per launch by the host:
//   kVarFast:    operands staged in LDS, linear next-state grid, constraints
//                in interval form Flo <= w <= Fhi;
//   kVarStaged:  operands staged in LDS, any grid kind, affine constraint rows;
//   kVarWdep:    some table varies with the continuous state: per-lane reads.
constexpr int kVarFast = 0;
constexpr int kVarStaged = 1;
constexpr int kVarWdep = 2;

// The fast variant keeps the EV row as (value, slope) pairs, so node k of the
// row sits at element 2k; the other variants keep it contiguous.
template <int VAR>
constexpr int kEvStride = VAR == kVarFast ? 2 : 1;

// Continuation kinds of a descriptor (block-uniform).
constexpr int kModeInterp = 0;    // stateful targets: interpolate EV
constexpr int kModeTerminal = 1;  // terminal regime: Q = U
constexpr int kModeScalar = 2;    // only stateless targets: Q = U + beta * cs

template <typename T>
struct Grid1D {
  int kind;
  int n;
  T start, step, inv_step, log_start, log_step;
  const T* pts;  // LDS copy for irregular grids
};

template <typename T>
__device__ __forceinline__ T lin_coordinate(const Grid1D<T>& g, T x) {
  // get_linspace_coordinate (coordinates.py:30-44)
  if constexpr (std::is_same<T, double>::value) {
    return (x - g.start) / g.step;
  } else {
    return (x - g.start) * g.inv_step;
  }
}

template <typename T>
__device__ __forceinline__ T coordinate(const Grid1D<T>& g, T x) {
  if (g.kind == LCM_GRID_LIN) {
    return lin_coordinate(g, x);
  } else if (g.kind == LCM_GRID_LOG) {
    // get_logspace_coordinate (coordinates.py:73-114)
    const T cl = (log(x) - g.log_start) / g.log_step;
    const T rlo = floor(cl);
    const T lower = exp(g.log_start + g.log_step * rlo);
    const T upper = exp(g.log_start + g.log_step * (rlo + T(1)));
    return rlo + (x - lower) / (upper - lower);
  } else {
    // get_irreg_coordinate (coordinates.py:117-167): searchsorted(side="right")
    if (g.n < 2) return T(0);
    int lo = 0, hi = g.n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (g.pts[mid] <= x) lo = mid + 1; else hi = mid;
    }
    const int up = min(max(lo, 1), g.n - 1);
    const int dn = up - 1;
    return T(dn) + (x - g.pts[dn]) / (g.pts[up] - g.pts[dn]);
  }
}

__device__ __forceinline__ float clamp_index(float x, float hi) {
  return __builtin_amdgcn_fmed3f(x, 0.0f, hi);  // NaN -> in range; t stays NaN
}
__device__ __forceinline__ double clamp_index(double x, double hi) {
  return fmin(fmax(x, 0.0), hi);
}

// Linear interpolation with linear extrapolation: lower index clipped to
// [0, n-2] (ndimage.py:102-109); with FINITE false a zero corner weight drops
// its term (the row holds -inf somewhere). Node k of the row is row[k * ES].
template <typename T, bool FINITE, int ES>
__device__ __forceinline__ T interp_row(const T* row, int n, T coord) {
  if (n < 2) return row[0];
  const T fl = clamp_index(floor(coord), T(n - 2));
  const int lo = static_cast<int>(fl);
  const T t = coord - fl;
  const T w0 = T(1) - t;
  const T v0 = row[lo * ES];
  const T v1 = row[(lo + 1) * ES];
  if constexpr (FINITE) return w0 * v0 + t * v1;
  return zero_safe(w0, v0) + zero_safe(t, v1);
}

// Running first-argmax of feasible Q: idx < 0 means no feasible action seen;
// a feasible NaN poisons the max (jnp.max propagates NaN).
template <typename T>
struct Running {
  T best;
  int idx;
  bool nan;
};

template <typename T>
__device__ __forceinline__ void offer(Running<T>& r, T q, bool feas, int a) {
  r.nan |= feas & (q != q);
  const bool take = feas & (q == q) & ((q > r.best) | (r.idx < 0));
  r.best = take ? q : r.best;
  r.idx = take ? a : r.idx;
}

// Merge candidate k into r: the larger value wins, ties keep the smaller index.
template <typename T>
__device__ __forceinline__ void merge(Running<T>& r, const Running<T>& k) {
  r.nan |= k.nan;
  if (k.idx >= 0 && (r.idx < 0 || k.best > r.best || (k.best == r.best && k.idx < r.idx))) {
    r.best = k.best;
    r.idx = k.idx;
  }
}

template <typename T>
struct alignas(sizeof(T) * 4) Act {
  T A, B, U, lo;  // lo: lower feasibility bound on w (fast variant)
};

template <typename T>
struct alignas(sizeof(T) * 2) Con {
  T a, b;
};

__device__ __forceinline__ size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

// Merge the SPLIT partial results of each state into the group-0 thread.
template <typename T, int TILE>
__device__ __forceinline__ void block_merge(Running<T>& r, T* red_v, int* red_i, int* red_f,
                                            int tid, int ls, int grp) {
  constexpr int SPLIT = kBlock / TILE;
  if constexpr (SPLIT > 1) {
    __syncthreads();
    red_v[tid] = r.best;
    red_i[tid] = r.idx;
    red_f[tid] = int(r.nan);
    __syncthreads();
    if (grp == 0) {
      for (int k = 1; k < SPLIT; ++k) {
        const int j = ls + k * TILE;
        merge(r, Running<T>{red_v[j], red_i[j], red_f[j] != 0});
      }
    }
  }
}

// One entry of the folded EV row: value at a node and slope to the next.
template <typename T>
struct alignas(sizeof(T) * 2) Pair {
  T v, dv;
};

// a * b + c with one rounding, written out so that every use of an
// expression rounds the same way whatever the compiler's contraction choices.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The accumulator of the EV contraction: FP64 whatever T, so the row is the
// contraction rounded once to T.
template <typename T>
using EvAcc = double;

// Everything the action loop reads (block-uniform except w / valid / wr).
template <typename T>
struct LoopArgs {
  const Act<T>* act;
  const T* his;
  const Con<T>* con;
  const T* ev;
  const T* Ug;
  const T* Ag;
  const T* Bg;
  const Pair<T>* pairs;  // folded EV row (fast variant)
  const T* pm;           // prefix max of the staged lower bounds (fast variant)
  const T* a_uni;        // A' of the chunk when it is one value (fast variant)
  const LcmGsDesc* D;
  Grid1D<T> g;
  T w, beta, cs;
  T w_wave;  // smallest state of this wave (its lane 0; states increase along the lanes)
  int n_constr, strict_mask, g_n, n_c, d, sd, wr, c0, cn, grp;
  bool valid;
};

// Feasibility of one staged action (interval or row constraints).
template <typename T, int VAR>
__device__ __forceinline__ bool staged_feas(const LoopArgs<T>& L, const Act<T>& a, int i) {
  const T w = L.w;
  if constexpr (VAR == kVarFast) {
    return (w >= a.lo) & (w <= L.his[i]);
  } else {
    bool feas = true;
    for (int k = 0; k < L.n_constr; ++k) {
      const Con<T> f = L.con[k * kActChunk + i];
      const T v = f.a * w + f.b;
      feas = feas & (((L.strict_mask >> k) & 1) ? (v > T(0)) : (v >= T(0)));
    }
    return feas;
  }
}

// Q of one staged action for this lane's state.
template <typename T, int VAR, int MODE, bool FINITE>
__device__ __forceinline__ T staged_q(const LoopArgs<T>& L, const Act<T>& a) {
  const T w = L.w;
  if constexpr (MODE == kModeInterp) {
    const T x = a.A * w + a.B;
    T coord;
    if constexpr (VAR == kVarFast) coord = lin_coordinate(L.g, x);
    else coord = coordinate(L.g, x);
    return a.U + L.beta * (L.cs + interp_row<T, FINITE, kEvStride<VAR>>(L.ev, L.g_n, coord));
  } else if constexpr (MODE == kModeTerminal) {
    return a.U;
  } else {
    return a.U + L.beta * L.cs;
  }
}

// The action loop over one staged chunk, specialised on the kernel variant,
// the continuation kind and whether the EV row is finite. Two actions per
// iteration (offered in increasing order, so the first argmax is kept) give
// the scheduler independent LDS reads to overlap.
template <typename T, int VAR, int MODE, bool FINITE, int SPLIT>
__device__ __forceinline__ void staged_loop(const LoopArgs<T>& L, Running<T>& run) {
  for (int i = L.grp; i < L.cn; i += 2 * SPLIT) {
    const int i1 = i + SPLIT;
    const bool has1 = i1 < L.cn;
    const int j1 = has1 ? i1 : i;  // stay inside the staged chunk
    const Act<T> a0 = L.act[i];
    const Act<T> a1 = L.act[j1];
    const bool f0 = L.valid & staged_feas<T, VAR>(L, a0, i);
    const bool f1 = L.valid & has1 & staged_feas<T, VAR>(L, a1, j1);
    if (__ballot(f0 | f1) == 0) continue;  // no lane of this wave can take either action
    const T q0 = staged_q<T, VAR, MODE, FINITE>(L, a0);
    const T q1 = staged_q<T, VAR, MODE, FINITE>(L, a1);
    offer(run, q0, f0, L.d * L.n_c + L.c0 + i);
    offer(run, q1, f1, L.d * L.n_c + L.c0 + i1);
  }
}

// q where f holds, NaN elsewhere, as a select on a value every lane computed
__device__ __forceinline__ float select_nan(bool f, float q) {
  return __builtin_bit_cast(float, f ? __builtin_bit_cast(unsigned, q) : 0x7fc00000u);
}
__device__ __forceinline__ double select_nan(bool f, double q) { return f ? q : __builtin_nan(""); }

// Pointers into global tables read with wave-uniform indices: loads through
// them go to the scalar unit, so per-action operands arrive in SGPRs and cost
// neither LDS bandwidth nor VGPRs.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using UniformPtr = const __attribute__((address_space(4))) T*;
#else
template <typename T>
using UniformPtr = const T*;  // the host pass only type-checks kernel bodies
#endif

// Folded loop of the fast variant for the common case (interval feasibility,
// linear next grid, finite EV row with at least two points). The host folded
// the grid into each action, so the coordinate of the next state is one fma,
// coord = A' * w + B' with A' = A / step and B' = (B - start) / step; the EV row
// is held as pairs (EV_k + cs, EV_{k+1} - EV_k) with the scalar continuation
// cs already added, so Q = U + beta * (EV_lo + t * slope), t = coord - lo, lo
// the floor of coord clamped to [0, g_n - 2]. Every fma is written out, so the
// recovery of a group's maximiser recomputes exactly the loop's values. The
// per-action operands A', B', U and the bounds are read in place from the
// global tables through UniformPtr (scalar loads); UA: A' is one value over
// the chunk, held in a vector register. The block runs this loop only on
// chunks whose magnitudes (|coord| <= 2^22, |EV|, |slope| <= 1e30, U not NaN)
// rule out overflow and NaN; the other chunks take the staged loop. LO / HI
// drop the bound checks a chunk does not need (all lower bounds -inf / all
// upper bounds +inf). The update is `offer`'s first-argmax rule, with NaN as
// the running best while no action is feasible: the first feasible action
// passes !(q <= best) whatever its value, later ones need q > best.
template <typename T, int SPLIT, bool LO, bool HI, bool UA, bool INTERP>
__device__ __forceinline__ void folded_loop(const LoopArgs<T>& L, UniformPtr<T> pa, UniformPtr<T> pb,
                                              UniformPtr<T> pu, UniformPtr<T> plo, UniformPtr<T> phi,
                                              Running<T>& run) {
  const T w = L.w;
  const T top = T(L.g_n - 2);
  const Pair<T>* __restrict__ pr = L.pairs;
  const int base = L.d * L.n_c + L.c0;
  const T beta = L.beta;
  T best = run.idx < 0 ? T(NAN) : run.best;
  int idx = run.idx;
  auto feas_of = [&](int i) -> bool {
    bool f = L.valid;
    if constexpr (LO) f = f & (w >= plo[i]);
    if constexpr (HI) f = f & (w <= phi[i]);
    return f;
  };
  // UA: every A' of the chunk is the one value staged in LDS, read once into a
  // vector register, so the fma takes one scalar operand (B') and needs no copy
  const T a_lds = UA ? L.a_uni[0] : T(0);
  const T cs = L.cs;
  auto q_of = [&](int i) -> T {
    if constexpr (!INTERP) return fma_t(beta, cs, pu[i]);  // a scalar continuation
    const T coord = fma_t(UA ? a_lds : pa[i], w, pb[i]);
    const T fl = clamp_index(T(floor(coord)), top);
    const Pair<T> p = pr[static_cast<int>(fl)];
    return fma_t(beta, fma_t(coord - fl, p.dv, p.v), pu[i]);
  };
  auto update = [&](T q, bool f, int c) {
    const bool take = f & !(q <= best);
    best = take ? q : best;
    idx = take ? c : idx;
  };
  // In the grouped loops the maximiser is kept as (it, slot): the loop
  // position of the last group that improved and the position k in that
  // group, idx = it + k * SPLIT.
  int it = idx, slot = 0;
  if (__builtin_amdgcn_ballot_w64(L.valid) == 0) return;  // no state of this wave is real
  int i = L.grp;
  if constexpr (LO && !HI) {
    // Actions [0, a) have every lower bound up to theirs at most the wave's
    // smallest state, so every lane can take them: no bound checks there.
    // pm is the chunk's prefix maximum of the lower bounds, so a counts the
    // actions with pm <= w_wave: each lane tests a quarter of the chunk and
    // the ballots are counted.
    const int lane = __lane_id();
    int a = 0;
#pragma unroll
    for (int k = 0; k < kActChunk / 64; ++k) {
      const int c = lane + 64 * k;
      a += __builtin_popcountll(__builtin_amdgcn_ballot_w64(c < L.cn && L.pm[c] <= L.w_wave));
    }
    a = __builtin_amdgcn_readfirstlane(a);
    // Every action here is feasible and its Q is not NaN (the folded loop's
    // gates), so a group of eight updates with its maximum alone and remembers
    // only which group improved (the leftovers run in groups of four); the
    // position inside the best group is recovered once, after the loops,
    // from the same Q values recomputed.
    bool open8 = false;
    int it8 = it;
    for (; i + 7 * SPLIT < a; i += 8 * SPLIT) {
      const T q0 = q_of(i);
      const T q1 = q_of(i + SPLIT);
      const T q2 = q_of(i + 2 * SPLIT);
      const T q3 = q_of(i + 3 * SPLIT);
      const T q4 = q_of(i + 4 * SPLIT);
      const T q5 = q_of(i + 5 * SPLIT);
      const T q6 = q_of(i + 6 * SPLIT);
      const T q7 = q_of(i + 7 * SPLIT);
      const T m = fmax(fmax(fmax(q0, q1), fmax(q2, q3)), fmax(fmax(q4, q5), fmax(q6, q7)));
      const bool take = L.valid & !(m <= best);
      best = take ? m : best;
      it8 = take ? base + i : it8;
      open8 |= take;
    }
    bool open = false;
    for (; i + 3 * SPLIT < a; i += 4 * SPLIT) {
      const T q0 = q_of(i);
      const T q1 = q_of(i + SPLIT);
      const T q2 = q_of(i + 2 * SPLIT);
      const T q3 = q_of(i + 3 * SPLIT);
      const T m = fmax(fmax(q0, q1), fmax(q2, q3));
      const bool take = L.valid & !(m <= best);
      best = take ? m : best;
      it = take ? base + i : it;
      open |= take;
    }
    if (open) {
      // the first position of the group that reaches its maximum
      const int ql = it - base;
      const T r0 = q_of(ql);
      const T r1 = q_of(ql + SPLIT);
      const T r2 = q_of(ql + 2 * SPLIT);
      const T r3 = q_of(ql + 3 * SPLIT);
      const T mr = fmax(fmax(r0, r1), fmax(r2, r3));
      slot = r0 == mr ? 0 : (r1 == mr ? 1 : (r2 == mr ? 2 : 3));
      best = mr;  // V from the values the slot was chosen from
    } else if (open8) {
      const int ql = it8 - base;
      const T r0 = q_of(ql);
      const T r1 = q_of(ql + SPLIT);
      const T r2 = q_of(ql + 2 * SPLIT);
      const T r3 = q_of(ql + 3 * SPLIT);
      const T r4 = q_of(ql + 4 * SPLIT);
      const T r5 = q_of(ql + 5 * SPLIT);
      const T r6 = q_of(ql + 6 * SPLIT);
      const T r7 = q_of(ql + 7 * SPLIT);
      const T mr = fmax(fmax(fmax(r0, r1), fmax(r2, r3)), fmax(fmax(r4, r5), fmax(r6, r7)));
      slot = r0 == mr ? 0 : r1 == mr ? 1 : r2 == mr ? 2 : r3 == mr ? 3 : r4 == mr ? 4 : r5 == mr ? 5 : r6 == mr ? 6 : 7;
      it = it8;
      best = mr;
    }
  }
  // the same quad update with the infeasible actions entered as NaN (which
  // the maximum skips); every lane evaluates, the clamped index keeping the
  // read in the row, so there is no branch per action
  bool open_b = false;  // as `open` in the prefix loop
  for (; i + 3 * SPLIT < L.cn; i += 4 * SPLIT) {
    const bool f0 = feas_of(i);
    const bool f1 = feas_of(i + SPLIT);
    const bool f2 = feas_of(i + 2 * SPLIT);
    const bool f3 = feas_of(i + 3 * SPLIT);
    // four ballots (each a compare result already in a scalar mask)
    if ((__builtin_amdgcn_ballot_w64(f0) | __builtin_amdgcn_ballot_w64(f1) | __builtin_amdgcn_ballot_w64(f2) |
         __builtin_amdgcn_ballot_w64(f3)) == 0)
      continue;
    const T q0 = select_nan(f0, q_of(i));
    const T q1 = select_nan(f1, q_of(i + SPLIT));
    const T q2 = select_nan(f2, q_of(i + 2 * SPLIT));
    const T q3 = select_nan(f3, q_of(i + 3 * SPLIT));
    const T m = fmax(fmax(q0, q1), fmax(q2, q3));
    // m is NaN exactly when none of the four is feasible (a feasible Q is
    // never NaN in the folded loop)
    const bool take = (m == m) & !(m <= best);
    best = take ? m : best;
    it = take ? base + i : it;
    open_b |= take;
  }
  if (open_b) {
    // the first feasible position of the best quad that reaches its maximum
    const int ql = it - base;
    const T r0 = select_nan(feas_of(ql), q_of(ql));
    const T r1 = select_nan(feas_of(ql + SPLIT), q_of(ql + SPLIT));
    const T r2 = select_nan(feas_of(ql + 2 * SPLIT), q_of(ql + 2 * SPLIT));
    const T r3 = select_nan(feas_of(ql + 3 * SPLIT), q_of(ql + 3 * SPLIT));
    const T mr = fmax(fmax(r0, r1), fmax(r2, r3));
    slot = r0 == mr ? 0 : (r1 == mr ? 1 : (r2 == mr ? 2 : 3));
    best = mr;  // V from the values the slot was chosen from
  }
  idx = it + slot * SPLIT;
  for (; i < L.cn; i += SPLIT) {
    const bool f = feas_of(i);
    if (__builtin_amdgcn_ballot_w64(f) == 0) continue;
    update(q_of(i), f, base + i);
  }
  run.best = idx < 0 ? T(-INFINITY) : best;
  run.idx = idx;
}

template <typename T, int SPLIT, bool UA, bool INTERP>
__device__ __forceinline__ void folded_bounds(const LoopArgs<T>& L, bool lo, bool hi, UniformPtr<T> pa,
                                              UniformPtr<T> pb, UniformPtr<T> pu, UniformPtr<T> plo,
                                              UniformPtr<T> phi, Running<T>& run) {
  if (lo) {
    if (hi) folded_loop<T, SPLIT, true, true, UA, INTERP>(L, pa, pb, pu, plo, phi, run);
    else folded_loop<T, SPLIT, true, false, UA, INTERP>(L, pa, pb, pu, plo, phi, run);
  } else {
    if (hi) folded_loop<T, SPLIT, false, true, UA, INTERP>(L, pa, pb, pu, plo, phi, run);
    else folded_loop<T, SPLIT, false, false, UA, INTERP>(L, pa, pb, pu, plo, phi, run);
  }
}

// The action loop over one chunk, specialised on the kernel variant, the
// continuation kind and whether the EV row is finite.
template <typename T, int VAR, int MODE, bool FINITE, int SPLIT>
__device__ __forceinline__ void action_loop(const LoopArgs<T>& L, Running<T>& run) {
  if constexpr (VAR != kVarWdep) {
    staged_loop<T, VAR, MODE, FINITE, SPLIT>(L, run);
    return;
  }
  const T w = L.w;
  for (int i = L.grp; i < L.cn; i += SPLIT) {
    const int c = L.c0 + i;
    T Uc, Ac, Bc;
    bool feas = L.valid;
    if constexpr (VAR == kVarFast) {
      const Act<T> a = L.act[i];
      const T hi = L.his[i];
      Uc = a.U;
      Ac = a.A;
      Bc = a.B;
      feas = feas & (w >= a.lo) & (w <= hi);
    } else if constexpr (VAR == kVarStaged) {
      const Act<T> a = L.act[i];
      Uc = a.U;
      Ac = a.A;
      Bc = a.B;
      for (int k = 0; k < L.n_constr; ++k) {
        const Con<T> f = L.con[k * kActChunk + i];
        const T v = f.a * w + f.b;
        feas = feas & (((L.strict_mask >> k) & 1) ? (v > T(0)) : (v >= T(0)));
      }
    } else {
      const LcmGsDesc* D = L.D;
      Uc = L.Ug[c];
      Ac = MODE == kModeInterp ? L.Ag[c] : T(0);
      Bc = MODE == kModeInterp ? L.Bg[c] : T(0);
      for (int k = 0; k < L.n_constr; ++k) {
        const T* Fa = static_cast<const T*>(D->Fa[k]);
        const T* Fb = static_cast<const T*>(D->Fb[k]);
        const int64_t ia = int64_t(L.sd) * D->Fa_ssd[k] + int64_t(L.wr) * D->Fa_sw[k] + int64_t(L.d) * L.n_c + c;
        const int64_t ib = int64_t(L.sd) * D->Fb_ssd[k] + int64_t(L.wr) * D->Fb_sw[k] + int64_t(L.d) * L.n_c + c;
        const T v = Fa[ia] * w + Fb[ib];
        feas = feas & (((L.strict_mask >> k) & 1) ? (v > T(0)) : (v >= T(0)));
      }
    }
    if (__ballot(feas) == 0) continue;  // no lane of this wave can take action c
    T q;
    if constexpr (MODE == kModeInterp) {
      const T x = Ac * w + Bc;
      T coord;
      if constexpr (VAR == kVarFast) coord = lin_coordinate(L.g, x);
      else coord = coordinate(L.g, x);
      q = Uc + L.beta * (L.cs + interp_row<T, FINITE, kEvStride<VAR>>(L.ev, L.g_n, coord));
    } else if constexpr (MODE == kModeTerminal) {
      q = Uc;
    } else {
      q = Uc + L.beta * L.cs;
    }
    offer(run, q, feas, L.d * L.n_c + c);
  }
}

template <typename T, int VAR, int SPLIT>
__device__ __forceinline__ void dispatch_loop(const LoopArgs<T>& L, int mode, bool finite_row,
                                              Running<T>& run) {
  if (mode == kModeInterp) {
    if (finite_row) action_loop<T, VAR, kModeInterp, true, SPLIT>(L, run);
    else action_loop<T, VAR, kModeInterp, false, SPLIT>(L, run);
  } else if (mode == kModeTerminal) {
    action_loop<T, VAR, kModeTerminal, true, SPLIT>(L, run);
  } else {
    action_loop<T, VAR, kModeScalar, true, SPLIT>(L, run);
  }
}

// Waves per SIMD the register allocation must leave room for. DENSE: the
// float32 kernel is compiled into 64 VGPRs (the SGPRs it spills go to VGPR
// lanes), so a SIMD holds 8 waves instead of the 6 its unconstrained 76 VGPRs
// allow: 9 % less time per launch on the 1024-draw batch, whose launches
// fill every CU many times over. Each wave runs longer with fewer registers,
// so a launch too small to fill the CUs is faster with the unconstrained
// kernel; the host picks per launch by its workgroup count (on the 1024-draw
// workload's shapes the two builds break even near 4,400 workgroups, the
// dense one is 6 % faster at 8,600 and 9 % at 33,800, and 10-25 % slower
// below 2,000). The float64 kernel
// needs about 125 VGPRs and is compiled unconstrained only. Both builds run
// the same arithmetic and give the same bits.
template <typename T, bool DENSE>
constexpr int kWavesPerEu = (DENSE && sizeof(T) == 4) ? 8 : 1;

template <typename T, int TILE, int VAR, bool DENSE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<T, DENSE>)))
gs_period_kernel(const LcmGsDesc* __restrict__ descs,
                 const LcmGsTask* __restrict__ tasks, int n_tasks) {
  constexpr bool WDEP = VAR == kVarWdep;
  constexpr bool FAST = VAR == kVarFast;
  constexpr int SPLIT = kBlock / TILE;
  constexpr int CH = kActChunk;
  extern __shared__ __align__(16) unsigned char smem[];

  // XCD-aware remap: hardware dispatches block b to XCD b % 8; give each XCD a
  // contiguous run of tasks. n_tasks is padded to a multiple of 8 by the host.
  const int b = blockIdx.x;
  const int per_xcd = n_tasks / kXcds;
  const int task_id = (b % kXcds) * per_xcd + (b / kXcds);
  // descriptors and task lists are constant while the kernel runs: read them
  // through the scalar cache whatever the kernel stores meanwhile
  const LcmGsTask tk = UniformPtr<LcmGsTask>(tasks)[task_id];
  if (tk.desc < 0) return;  // padding task: whole block exits together

  const UniformPtr<LcmGsDesc> D = UniformPtr<LcmGsDesc>(descs + tk.desc);
  const int tid = threadIdx.x;
  const int ls = tid % TILE;
  // the waves sharing a state interleave over actions; grp is wave-uniform
  const int grp = __builtin_amdgcn_readfirstlane(tid / TILE);
  const int n_w = D->n_w, n_d = D->n_d, n_c = D->n_c, sd = tk.sd;
  const int wi = tk.w0 + ls;
  const bool valid = wi < n_w;
  const int wr = valid ? wi : 0;
  const T w = static_cast<const T*>(D->w_pts)[wr];
  const T w_first = static_cast<const T*>(D->w_pts)[tk.w0];
  const T w_last = static_cast<const T*>(D->w_pts)[min(tk.w0 + TILE, n_w) - 1];
  const T w_abs = fmax(fabs(w_first), fabs(w_last));  // max |w| over the block (w_pts increase)
  const int ls0 = (tid & ~63) % TILE;  // state of this wave's lane 0
  const T w_wave = static_cast<const T*>(D->w_pts)[min(tk.w0 + ls0, n_w - 1)];
  const bool terminal = D->terminal != 0;
  const bool taste = D->has_taste != 0;
  // a stateful continuation needs a next-state grid with at least one point
  const bool interp = !terminal && D->n_targets > 0 && D->g_n > 0;
  const int mode = interp ? kModeInterp : (terminal ? kModeTerminal : kModeScalar);
  const int g_n = interp ? D->g_n : 0;
  const int n_constr = D->n_constr;

  // LDS carve (the host sizes the launch for the largest g_n and n_constr).
  // The first region holds the EV row (and the irregular grid points) while
  // the actions are scanned, and the SPLIT wave groups' partial results when
  // they are merged: every merge starts with a barrier after the last read of
  // the row, and the next row is written after the next barrier
  const size_t g_bytes = align16(size_t(g_n > 0 ? g_n : 1) * sizeof(T));
  const size_t row_bytes = max(2 * g_bytes, size_t(kBlock) * (sizeof(T) + 8));
  // EV row: contiguous, or in the fast variant the .v members of the pairs
  constexpr int ES = kEvStride<VAR>;
  T* ev = reinterpret_cast<T*>(smem);
  Pair<T>* pairs = reinterpret_cast<Pair<T>*>(smem);
  T* gp = reinterpret_cast<T*>(smem + g_bytes);  // irregular grid points (not the fast variant)
  Act<T>* act = reinterpret_cast<Act<T>*>(smem + row_bytes);
  T* his = reinterpret_cast<T*>(smem + row_bytes + CH * sizeof(Act<T>));
  const size_t con_ofs = row_bytes + CH * sizeof(Act<T>) + align16(CH * sizeof(T));
  Con<T>* con = reinterpret_cast<Con<T>*>(smem + con_ofs);  // constraint rows (staged variant)
  T* pm = reinterpret_cast<T*>(smem + con_ofs);              // fast variant: prefix max of lo
  T* wtot = reinterpret_cast<T*>(smem + con_ofs + align16(CH * sizeof(T)));  // its per-wave maxima
  T* red_v = reinterpret_cast<T*>(smem);  // shares the row region (see above)
  int* red_i = reinterpret_cast<int*>(red_v + kBlock);
  int* red_f = red_i + kBlock;

  LoopArgs<T> L;
  L.act = act;
  L.his = his;
  L.con = con;
  L.ev = ev;
  L.pairs = pairs;
  L.pm = pm;
  T* a_uni = wtot + kBlock / 64;  // after the per-wave maxima
  L.a_uni = a_uni;
  L.D = reinterpret_cast<const LcmGsDesc*>(descs + tk.desc);
  L.w = w;
  L.valid = valid;
  L.n_constr = n_constr;
  L.strict_mask = D->strict_mask;
  L.g_n = g_n;
  L.n_c = n_c;
  L.sd = sd;
  L.wr = wr;
  L.grp = grp;
  L.beta = T(D->beta);
  L.w_wave = w_wave;

  Grid1D<T>& g = L.g;
  g.kind = D->g_kind;
  g.n = g_n;
  g.start = T(D->g_start);
  const T g_stop = T(D->g_stop);
  g.step = (g_n > 1) ? (g_stop - g.start) / T(g_n - 1) : T(1);
  g.inv_step = T(1) / g.step;
  g.log_start = (g.kind == LCM_GRID_LOG) ? T(log(T(D->g_start))) : T(0);
  g.log_step = (g.kind == LCM_GRID_LOG && g_n > 1)
                   ? (T(log(g_stop)) - g.log_start) / T(g_n - 1)
                   : T(1);
  g.pts = gp;
  if (interp && g.kind == LCM_GRID_IRREG) {
    const T* src = static_cast<const T*>(D->g_pts);
    for (int k = tid; k < g_n; k += kBlock) gp[k] = src[k];
  }

  const int64_t wofs_u = WDEP ? int64_t(wr) * D->U_sw : 0;
  const int64_t wofs_a = WDEP ? int64_t(wr) * D->A_sw : 0;
  const int64_t wofs_b = WDEP ? int64_t(wr) * D->B_sw : 0;

  Running<T> run{T(-INFINITY), -1, false};
  // taste shocks: running logsum over the discrete actions (max tm, scaled
  // sum ts, NaN flag), so no per-action array stays live in registers
  T tm = T(-INFINITY), ts = T(0);
  bool tnan = false;

  for (int d = 0; d < n_d; ++d) {
    T cs = T(0);
    bool finite_row = true;
    bool row_small = false;  // folded row built and within the folded loop's bounds
    if (!terminal) {
      __syncthreads();  // previous readers of ev are done
      for (int s = 0; s < D->n_scalar; ++s) {
        const T m = static_cast<const T*>(D->ms[s])[int64_t(sd) * n_d + d];
        const T v = static_cast<const T*>(D->Vs[s])[0];
        cs += zero_safe(m, v);
      }
      const bool fold_cs = FAST && mode == kModeInterp;
      // EV[sd, d][k] = sum_r sum_j M_r[sd, d, j] * Vn_r[j, k], accumulated
      // in FP64 and rounded once to T (a few FP64 fmas per node and block)
      int local_nonfinite = 0;
      for (int k = tid; k < g_n; k += kBlock) {
        EvAcc<T> acc = 0.0;
        for (int r = 0; r < D->n_targets; ++r) {
          const int nj = D->n_sd_next[r];
          const T* __restrict__ Mr = static_cast<const T*>(D->M[r]) + (int64_t(sd) * n_d + d) * nj;
          const T* __restrict__ Vr = static_cast<const T*>(D->Vn[r]);
          for (int j = 0; j < nj; ++j) {
            const T m = Mr[j];
            if (m != T(0)) acc = fma_t(double(m), double(Vr[int64_t(j) * g_n + k]), acc);
          }
        }
        // the fast variant's folded loop reads the scalar continuation from
        // the row (EV_k + cs, rounded once)
        const T ev_k = fold_cs ? T(acc + double(cs)) : T(acc);
        ev[k * ES] = ev_k;
        local_nonfinite |= !is_finite(ev_k);
      }
      finite_row = __syncthreads_or(local_nonfinite) == 0;
      if (FAST && mode == kModeInterp && finite_row && g_n >= 2) {
        // slopes of the folded row; the values stay as the other loops read them
        int big = 0;
        for (int k = tid; k < g_n; k += kBlock) {
          const T v = pairs[k].v;
          const T dv = k + 1 < g_n ? pairs[k + 1].v - v : T(0);
          pairs[k].dv = dv;
          big |= !(fabs(v) <= T(1e30)) | !(fabs(dv) <= T(1e30));
        }
        row_small = __syncthreads_or(big) == 0;
      }
    }
    L.cs = (FAST && mode == kModeInterp) ? T(0) : cs;  // folded into the row there
    L.d = d;
    L.Ug = static_cast<const T*>(D->U) + int64_t(sd) * D->U_ssd + wofs_u + int64_t(d) * n_c;
    L.Ag = interp ? static_cast<const T*>(D->A) + int64_t(sd) * D->A_ssd + wofs_a + int64_t(d) * n_c : nullptr;
    L.Bg = interp ? static_cast<const T*>(D->B) + int64_t(sd) * D->B_ssd + wofs_b + int64_t(d) * n_c : nullptr;
    // the host-folded tables of the folded loop (no continuous-state axis there)
    const T* Afg = interp ? static_cast<const T*>(D->Af) + int64_t(sd) * D->A_ssd + int64_t(d) * n_c : nullptr;
    const T* Bfg = interp ? static_cast<const T*>(D->Bf) + int64_t(sd) * D->B_ssd + int64_t(d) * n_c : nullptr;
    // the common cases run the folded loop: interpolation with the grid
    // folded into A, B and the EV row scaled, or a scalar continuation
    const bool folded = FAST && ((mode == kModeInterp && finite_row && row_small) || mode == kModeScalar);

    for (int c0 = 0; c0 < n_c; c0 += CH) {
      const int cn = min(CH, n_c - c0);
      L.c0 = c0;
      L.cn = cn;
      if constexpr (FAST) {
        // skip a chunk no state of this block can take (w_pts increase)
        const T* fc = static_cast<const T*>(D->Fchunk) + int64_t(sd) * D->Fc_ssd +
                      (int64_t(d) * ((n_c + CH - 1) / CH) + c0 / CH) * 2;
        if (fc[0] > w_last || fc[1] < w_first) continue;
      }
      if constexpr (FAST) {
        if (folded) {
          __syncthreads();  // previous chunk consumed
          const int64_t fi = int64_t(sd) * D->Fi_ssd + int64_t(d) * n_c + c0;
          int need_lo = 0, need_hi = 0, unsafe = 0, a_diff = 0;
          const T a_first = interp ? Afg[c0] : T(0);
          T lo_me = T(-INFINITY);  // this thread's lower bound (one action per thread)
          static_assert(CH == kBlock, "one staged action per thread");
          for (int i = tid; i < cn; i += kBlock) {
            const T lo_i = static_cast<const T*>(D->Flo)[fi + i];
            const T hi_i = static_cast<const T*>(D->Fhi)[fi + i];
            need_lo |= !(lo_i == T(-INFINITY));
            need_hi |= !(hi_i == T(INFINITY));
            lo_me = lo_i == lo_i ? lo_i : T(INFINITY);  // a NaN bound is never met
            // the loop reads its operands in place (A' = A / step and
            // B' = (B - start) / step folded by the host, each rounded once, and
            // U); here only the gates: |coord| over the block's states (w_pts
            // increase), Q of a scalar continuation not NaN, and whether A' is
            // one value over the chunk
            const T U = L.Ug[c0 + i];
            if (interp) {
              const T A = Afg[c0 + i], B = Bfg[c0 + i];
              unsafe |= !(fabs(A) * w_abs + fabs(B) <= T(4194304)) | (U != U);
              a_diff |= !(A == a_first);
            } else {
              const T q = fma_t(L.beta, cs, U);
              unsafe |= q != q;
            }
          }
          // inclusive prefix max of the lower bounds: within the wave by
          // shuffles, then across waves through LDS
          const int lane = tid & 63;
#pragma unroll
          for (int off = 1; off < 64; off <<= 1) {
            const T y = __shfl_up(lo_me, off);
            if (lane >= off) lo_me = fmax(lo_me, y);
          }
          if (lane == 63) wtot[tid >> 6] = lo_me;
          const bool lo = __builtin_amdgcn_readfirstlane(__syncthreads_or(need_lo)) != 0;
          for (int k = 0; k < (tid >> 6); ++k) lo_me = fmax(lo_me, wtot[k]);
          if (tid < cn) pm[tid] = lo_me;
          const bool hi = __builtin_amdgcn_readfirstlane(__syncthreads_or(need_hi)) != 0;
          if (__builtin_amdgcn_readfirstlane(__syncthreads_or(unsafe)) == 0) {
            const UniformPtr<T> pa = UniformPtr<T>(interp ? Afg + c0 : nullptr);
            const UniformPtr<T> pb = UniformPtr<T>(interp ? Bfg + c0 : nullptr);
            const UniformPtr<T> pu = UniformPtr<T>(L.Ug + c0);
            const UniformPtr<T> plo = UniformPtr<T>(static_cast<const T*>(D->Flo) + fi);
            const UniformPtr<T> phi = UniformPtr<T>(static_cast<const T*>(D->Fhi) + fi);
            if (interp) {
              if (tid == 0) a_uni[0] = a_first;
              if (__builtin_amdgcn_readfirstlane(__syncthreads_or(a_diff)) == 0)
                folded_bounds<T, SPLIT, true, true>(L, lo, hi, pa, pb, pu, plo, phi, run);
              else
                folded_bounds<T, SPLIT, false, true>(L, lo, hi, pa, pb, pu, plo, phi, run);
            } else {
              folded_bounds<T, SPLIT, false, false>(L, lo, hi, pa, pb, pu, plo, phi, run);
            }
            continue;
          }
          // otherwise the staged chunk below replaces the folded one
        }
      }
      if constexpr (!WDEP) {
        __syncthreads();  // previous chunk consumed
        const int64_t fi = int64_t(sd) * D->Fi_ssd + int64_t(d) * n_c + c0;
        for (int i = tid; i < cn; i += kBlock) {
          Act<T> a;
          a.U = L.Ug[c0 + i];
          a.A = interp ? L.Ag[c0 + i] : T(0);
          a.B = interp ? L.Bg[c0 + i] : T(0);
          a.lo = FAST ? static_cast<const T*>(D->Flo)[fi + i] : T(0);
          act[i] = a;
          if constexpr (FAST) {
            his[i] = static_cast<const T*>(D->Fhi)[fi + i];
          } else {
            for (int k = 0; k < n_constr; ++k) {
              const T* Fa = static_cast<const T*>(D->Fa[k]) + int64_t(sd) * D->Fa_ssd[k] + int64_t(d) * n_c;
              const T* Fb = static_cast<const T*>(D->Fb[k]) + int64_t(sd) * D->Fb_ssd[k] + int64_t(d) * n_c;
              Con<T> f;
              f.a = Fa[c0 + i];
              f.b = Fb[c0 + i];
              con[k * CH + i] = f;
            }
          }
        }
        __syncthreads();
      }
      dispatch_loop<T, VAR, SPLIT>(L, mode, finite_row, run);
    }

    if (taste) {
      // per-discrete-action max over the continuous axis, merged over groups
      block_merge<T, TILE>(run, red_v, red_i, red_f, tid, ls, grp);
      const T qd = run.nan ? T(NAN) : (run.idx >= 0 ? run.best : T(-INFINITY));
      const int id = (run.idx >= 0 && !run.nan) ? run.idx : 0;
      if (grp == 0 && valid) {
        D->pol_out[(int64_t(sd) * n_w + wi) * n_d + d] = id;
        // logsum_and_softmax (logsum.py:32-64), m + scale * log(sum exp((q - m) / scale)),
        // accumulated one discrete action at a time: the sum is rescaled
        // whenever the maximum rises
        const T lam = T(D->taste_scale);
        tnan |= qd != qd;
        if (qd == qd) {
          const T m1 = fmax(tm, qd);
          if (m1 != T(-INFINITY)) ts = ts * exp((tm - m1) / lam) + exp((qd - m1) / lam);
          tm = m1;
        }
      }
      run = Running<T>{T(-INFINITY), -1, false};
    }
  }

  if (taste) {
    if (grp == 0 && valid) {
      const T lam = T(D->taste_scale);
      const T fm = (tm == T(-INFINITY)) ? T(0) : tm;
      T V = fm + lam * log(ts);
      if (tnan) V = T(NAN);
      static_cast<T*>(D->V_out)[int64_t(sd) * n_w + wi] = V;
    }
    return;
  }

  block_merge<T, TILE>(run, red_v, red_i, red_f, tid, ls, grp);
  if (grp == 0 && valid) {
    const int64_t o = int64_t(sd) * n_w + wi;
    static_cast<T*>(D->V_out)[o] = run.nan ? T(NAN) : (run.idx >= 0 ? run.best : T(-INFINITY));
    D->pol_out[o] = (run.idx >= 0 && !run.nan) ? run.idx : 0;
  }
}

template <typename T, int TILE, bool DENSE>
int launch_occ(lcm_ctx* ctx, const LcmGsDesc* d, const LcmGsTask* t, int n, size_t lds, int var) {
  auto kern = var == kVarFast     ? gs_period_kernel<T, TILE, kVarFast, DENSE>
              : var == kVarStaged ? gs_period_kernel<T, TILE, kVarStaged, DENSE>
                                  : gs_period_kernel<T, TILE, kVarWdep, DENSE>;
  if (lds > 64 * 1024) {
    LCM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  }
  hipLaunchKernelGGL(kern, dim3(n), dim3(kBlock), lds, ctx->stream, d, t, n);
  LCM_HIP(hipGetLastError());
  return LCM_OK;
}

template <typename T, int TILE>
int launch_tile(lcm_ctx* ctx, const LcmGsDesc* d, const LcmGsTask* t, int n, size_t lds, int var, bool dense) {
  if constexpr (sizeof(T) == 4) {
    if (dense) return launch_occ<T, TILE, true>(ctx, d, t, n, lds, var);
  }
  return launch_occ<T, TILE, false>(ctx, d, t, n, lds, var);
}

template <typename T>
int launch_t(lcm_ctx* ctx, const LcmGsDesc* d, const LcmGsTask* t, int n, int tile, size_t lds, int var,
             bool dense) {
  switch (tile) {
    case 64: return launch_tile<T, 64>(ctx, d, t, n, lds, var, dense);
    case 128: return launch_tile<T, 128>(ctx, d, t, n, lds, var, dense);
    case 256: return launch_tile<T, 256>(ctx, d, t, n, lds, var, dense);
    default: return fail(LCM_E_ARG, "tile must be 64, 128 or 256");
  }
}

}  // namespace lcm

// Dynamic LDS bytes a launch needs for next-state grids of up to max_g points
// and up to max_constr constraint rows (mirrors the carve in gs_period_kernel).
extern "C" size_t lcm_gs_lds_bytes(int dtype, int max_g, int max_constr) {
  const size_t sz = dtype == LCM_DTYPE_F64 ? 8 : 4;
  const size_t g = (size_t(max_g > 0 ? max_g : 1) * sz + 15) & ~size_t(15);
  const size_t his = (lcm::kActChunk * sz + 15) & ~size_t(15);
  const size_t row = std::max(2 * g, size_t(lcm::kBlock) * (sz + 8));  // EV row, or the merge's partials
  return row + lcm::kActChunk * 4 * sz + 2 * his + 64 + size_t(max_constr) * lcm::kActChunk * 2 * sz + 64;
}

extern "C" int lcm_gs_launch_ex(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                                const LcmGsTask* d_tasks, int n_tasks, int tile,
                                size_t lds_bytes, int variant) {
  using namespace lcm;
  if (!ctx || !d_descs || !d_tasks) return fail(LCM_E_ARG, "lcm_gs_launch: null argument");
  if (n_tasks <= 0 || n_tasks % kXcds != 0)
    return fail(LCM_E_ARG, "lcm_gs_launch: n_tasks must be a positive multiple of 8");
  if (lds_bytes > 160 * 1024) return fail(LCM_E_ARG, "lcm_gs_launch: LDS request above 160 KiB");
  const int occ = variant & (LCM_GS_OCC_DENSE | LCM_GS_OCC_SPARSE);
  variant &= ~(LCM_GS_OCC_DENSE | LCM_GS_OCC_SPARSE);
  if (variant < kVarFast || variant > kVarWdep) return fail(LCM_E_ARG, "lcm_gs_launch: unknown variant");
  if (occ == (LCM_GS_OCC_DENSE | LCM_GS_OCC_SPARSE))
    return fail(LCM_E_ARG, "lcm_gs_launch: both register budgets requested");
  // the 8-waves-per-SIMD build for launches that fill the CUs many times over
  // and whose LDS lets more than the unconstrained build's 6 waves share a
  // SIMD (a workgroup is 4 waves, one per SIMD: the workgroups that fit in a
  // CU's LDS are the waves per SIMD); with fewer the dense build only runs
  // each wave slower
  const size_t lds_per_wg = lds_bytes > 0 ? lds_bytes : 1;
  const bool room = (size_t(160) * 1024) / lds_per_wg > 6;
  const bool dense = occ ? occ == LCM_GS_OCC_DENSE : (n_tasks >= LCM_GS_DENSE_MIN_TASKS && room);
  if (dtype == LCM_DTYPE_F32)
    return launch_t<float>(ctx, d_descs, d_tasks, n_tasks, tile, lds_bytes, variant, dense);
  if (dtype == LCM_DTYPE_F64)
    return launch_t<double>(ctx, d_descs, d_tasks, n_tasks, tile, lds_bytes, variant, false);
  return fail(LCM_E_ARG, "lcm_gs_launch: unknown dtype");
}

// The general variant is correct for every descriptor.
extern "C" int lcm_gs_launch(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs,
                             const LcmGsTask* d_tasks, int n_tasks, int tile,
                             size_t lds_bytes) {
  return lcm_gs_launch_ex(ctx, dtype, d_descs, d_tasks, n_tasks, tile, lds_bytes, lcm::kVarWdep);
}

extern "C" int lcm_gs_serve_launch(lcm_ctx* ctx, int dtype, const LcmGsDesc* d_descs, const LcmGsTask* h_tasks,
                                   LcmGsTask* d_tasks, int n_tasks, int tile, size_t lds_bytes, int variant,
                                   void* ev_start, void* ev_ring, void* const* dst, const void* const* src,
                                   const size_t* bytes, int n_copies, void* ev_end) {
  using namespace lcm;
  if (!ctx || !h_tasks || !d_tasks || (n_copies > 0 && (!dst || !src || !bytes)) || n_copies < 0)
    return fail(LCM_E_ARG, "lcm_gs_serve_launch: null argument");
  if (ev_start) LCM_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev_start), ctx->stream));
  LCM_HIP(hipMemcpyAsync(d_tasks, h_tasks, size_t(n_tasks) * sizeof(LcmGsTask), hipMemcpyHostToDevice,
                         ctx->stream));
  const int rc = lcm_gs_launch_ex(ctx, dtype, d_descs, d_tasks, n_tasks, tile, lds_bytes, variant);
  if (rc != LCM_OK) return rc;
  if (ev_ring) LCM_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev_ring), ctx->stream));
  for (int i = 0; i < n_copies; ++i) {
    if (!dst[i] || !src[i]) return fail(LCM_E_ARG, "lcm_gs_serve_launch: null copy pointer");
    if (bytes[i]) LCM_HIP(hipMemcpyAsync(dst[i], src[i], bytes[i], hipMemcpyDeviceToHost, ctx->stream));
  }
  if (ev_end) LCM_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(ev_end), ctx->stream));
  return LCM_OK;
}
