/*
 * C restatement of one grid-search Bellman step (TEST INFRASTRUCTURE ONLY:
 * the CPU baseline timed by bench.py and a cross-check of oracle/gridsearch.py).
 *
 * Same order of operations as the reference's Q_and_F + max_Q_over_a
 * (src/_lcm/regime_building/Q_and_F.py:466-681,1145-1209;
 *  src/_lcm/regime_building/ndimage.py:28-109;
 *  src/_lcm/regime_building/max_Q_over_a.py:152-160;
 *  src/_lcm/regime_building/argmax.py:6-69; src/_lcm/logsum.py:32-64):
 * for every state-action pair, interpolate each next node's value row at the
 * next continuous state, weight by the node weights, normalise by their sum,
 * weight by the regime probability, divide by the represented mass, add
 * beta * CE to utility, and reduce with the first-argmax rule. Inputs are the
 * tables oracle/gridsearch.py evaluates from the user's functions.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static double zero_safe(double w, double v) {
  if (w == 0.0 && !isfinite(v)) return 0.0;
  return w * v;
}

static double coordinate(int kind, int G, const double* pts, double x) {
  if (G < 2) return 0.0;
  double start = pts[0], stop = pts[G - 1];
  if (kind == 0) { /* linspace */
    double step = (stop - start) / (G - 1);
    return (x - start) / step;
  }
  if (kind == 1) { /* logspace */
    double sl = log(start), el = log(stop);
    double step = (el - sl) / (G - 1);
    double cl = (log(x) - sl) / step;
    double rlo = floor(cl);
    double lo = exp(sl + step * rlo), hi = exp(sl + step * (rlo + 1));
    return rlo + (x - lo) / (hi - lo);
  }
  /* irregular: searchsorted(side="right") */
  int a = 0, b = G;
  while (a < b) {
    int m = (a + b) / 2;
    if (pts[m] <= x) a = m + 1; else b = m;
  }
  int up = a < 1 ? 1 : (a > G - 1 ? G - 1 : a);
  int dn = up - 1;
  return dn + (x - pts[dn]) / (pts[up] - pts[dn]);
}

static double interp(const double* row, int G, double coord) {
  if (G < 2) return row[0];
  double fl = floor(coord);
  if (!(fl >= 0)) fl = 0; /* NaN -> 0 as well */
  if (fl > G - 2) fl = G - 2;
  int lo = (int)fl;
  double t = coord - fl;
  return zero_safe(1.0 - t, row[lo]) + zero_safe(t, row[lo + 1]);
}

/*
 * n_sd, n_w, n_d, n_c : state = sd*n_w + w ; action = d*n_c + c
 * U, wn : [S*A] ; F : [S*A] (1 = feasible)
 * n_t stateful targets: V_r [K_r*G], W_r [n_sd*n_d*K_r] node weights,
 *                       p_r [n_sd*n_d] regime probabilities
 * n_s stateless targets: v_s value, q_s [n_sd*n_d] probabilities
 * terminal: Q = U. taste > 0: logsum over d of the per-d max.
 * Vout [S]; pol [S] or [S*n_d] with taste shocks.
 * pcmp / qcmp (optional, NULL to skip): for a policy to compare against (same
 * layout as pol), qcmp[2 * slot] receives this step's Q at that policy's action
 * and qcmp[2 * slot + 1] the best Q of the slot, so a caller can tell a near-tie
 * from a genuine argmax mismatch.
 */
int gs_step_ref(int n_sd, int n_w, int n_d, int n_c, const double* U, const uint8_t* F,
                const double* wn, int terminal, int n_t, const double* const* Vt,
                const int* Kt, const double* const* Wt, const double* const* pt, int n_s,
                const double* vs, const double* const* qs, int gkind, int G,
                const double* gpts, double beta, double taste, double* Vout, int32_t* pol,
                const int32_t* pcmp, double* qcmp, int threads) {
  const long S = (long)n_sd * n_w;
  const long A = (long)n_d * n_c;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (long s = 0; s < S; ++s) {
    const int sd = (int)(s / n_w);
    double best = -INFINITY;
    int bidx = 0, seen = 0, nan = 0;
    double qc_best[16];
    int qc_idx[16];
    for (int d = 0; d < n_d; ++d) {
      double mass = 0.0, ce0 = 0.0;
      int neg = 0;
      const long sdd = (long)sd * n_d + d;
      if (!terminal) {
        for (int r = 0; r < n_s; ++r) {
          double p = qs[r][sdd];
          mass += p;
          neg |= (p < 0);
          ce0 += zero_safe(p, vs[r]);
        }
        for (int r = 0; r < n_t; ++r) {
          double p = pt[r][sdd];
          mass += p;
          neg |= (p < 0);
        }
      }
      double dbest = -INFINITY;
      int dseen = 0, dnan = 0, didx = 0;
      const long cmp_slot = taste > 0 ? s * n_d + d : s;
      const int cmp_idx = pcmp ? pcmp[cmp_slot] : -1;
      if (qcmp && (taste > 0 || d == 0)) qcmp[2 * cmp_slot] = -INFINITY;
      for (int c = 0; c < n_c; ++c) {
        const long i = s * A + (long)d * n_c + c;
        if (!F[i]) continue;
        double q;
        if (terminal) {
          q = U[i];
        } else {
          double ce = ce0;
          if (n_t > 0) {
            const double coord = coordinate(gkind, G, gpts, wn[i]);
            for (int r = 0; r < n_t; ++r) {
              const int K = Kt[r];
              const double* W = Wt[r] + sdd * K;
              double num = 0.0, wsum = 0.0;
              for (int k = 0; k < K; ++k) {
                num += zero_safe(W[k], interp(Vt[r] + (long)k * G, G, coord));
                wsum += W[k];
              }
              const double E = num / (wsum > 0 ? wsum : 1.0);
              ce += zero_safe(pt[r][sdd], E);
            }
          }
          const int ok = fabs(mass - 1.0) <= 1e-3 && !neg;
          ce = ce / (ok ? mass : NAN);
          q = U[i] + beta * ce;
        }
        if (qcmp && d * n_c + c == cmp_idx) qcmp[2 * cmp_slot] = q;
        if (q != q) { dnan = 1; continue; }
        if (!dseen || q > dbest) { dbest = q; didx = d * n_c + c; dseen = 1; }
      }
      if (qcmp && taste > 0) qcmp[2 * cmp_slot + 1] = dnan ? NAN : (dseen ? dbest : -INFINITY);
      if (taste > 0) {
        qc_best[d] = dnan ? NAN : (dseen ? dbest : -INFINITY);
        qc_idx[d] = (dseen && !dnan) ? didx : 0;
      } else {
        nan |= dnan;
        if (dseen && (!seen || dbest > best)) { best = dbest; bidx = didx; seen = 1; }
      }
    }
    if (taste > 0) {
      double m = -INFINITY;
      int has_nan = 0;
      for (int d = 0; d < n_d; ++d) { has_nan |= qc_best[d] != qc_best[d]; if (qc_best[d] > m) m = qc_best[d]; }
      double fm = (m == -INFINITY) ? 0.0 : m, sum = 0.0;
      for (int d = 0; d < n_d; ++d) sum += exp((qc_best[d] - fm) / taste);
      Vout[s] = has_nan ? NAN : fm + taste * log(sum);
      for (int d = 0; d < n_d; ++d) pol[s * n_d + d] = qc_idx[d];
    } else {
      Vout[s] = nan ? NAN : (seen ? best : -INFINITY);
      if (qcmp) qcmp[2 * s + 1] = Vout[s];
      pol[s] = (seen && !nan) ? bidx : 0;
    }
  }
  return 0;
}
