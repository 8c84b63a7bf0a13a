// Differential evolution driver (DE/rand/1/bin, synchronous generations).
//
// Replaces BoostDESolver::minimize (src/optimization/solver.cpp:113-155): the
// box is supplied by the caller (the reference derives it at :118-125), the
// initial guess is clamped into the population as member 0 (:138-141), every
// trial is judged by the exact cost |F|^2 (:143-146). The generator is a
// counter-based splitmix64 stream so the same seed reproduces the same
// population on any device (restated by the CPU oracle's or_de):
//   member m > 0 initial value k: unit(seed, 0, m*D + k)
//   generation g picks r1, r2, r3 by rejection: unit(seed, 2g, m*1024 + t)
//   jrand and the crossover draws: unit(seed, 2g+1, m*(D+1) + k), k = D for jrand
#include "common.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace fsa {

int fs_population_impl(Context &c, const double *X, int np, double *cost);
int surrogate_screen(Context &c, const fs_de_opts *o, int gen, const std::vector<double> &pop,
                     const std::vector<double> &cost, const double *lo, const double *hi,
                     std::vector<double> &trial);
void surrogate_observe(Context &c, const fs_de_opts *o, const std::vector<double> &trial,
                       const std::vector<double> &tcost, const double *lo, const double *hi);
void surrogate_begin(Context &c, const fs_de_opts *o, const std::vector<double> &pop,
                     const std::vector<double> &cost, const double *lo, const double *hi);
void surrogate_maybe_train(Context &c, const fs_de_opts *o, int gen);
double rng_unit(uint64_t seed, uint64_t stream, uint64_t index);
void de_trial(uint64_t seed, int g, int m, int np, int D, double F, double CR, const double *pop,
              const double *lo, const double *hi, double *out, uint64_t stream_base);

namespace {

uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

} // namespace

double rng_unit(uint64_t seed, uint64_t stream, uint64_t index) {
  const uint64_t u = splitmix64(splitmix64(seed ^ splitmix64(stream)) + index);
  return double(u >> 11) * 0x1.0p-53;
}

int de_pick(uint64_t seed, uint64_t stream, int m, int np, int &t, int e0, int e1, int e2) {
  for (;;) {
    const int r = int(rng_unit(seed, stream, uint64_t(m) * 1024u + uint64_t(t)) * np);
    ++t;
    if (r != e0 && r != e1 && r != e2 && r < np) return r;
  }
}

// trial vector of member m in generation g (the reference rand/1/bin step)
void de_trial(uint64_t seed, int g, int m, int np, int D, double F, double CR, const double *pop,
              const double *lo, const double *hi, double *out, uint64_t stream_base) {
  int t = 0;
  const uint64_t ps = 2u * uint64_t(g) + stream_base;
  const int r1 = de_pick(seed, ps, m, np, t, m, -1, -1);
  const int r2 = de_pick(seed, ps, m, np, t, m, r1, -1);
  const int r3 = de_pick(seed, ps, m, np, t, m, r1, r2);
  const uint64_t cs = ps + 1u;
  const int jr = int(rng_unit(seed, cs, uint64_t(m) * (D + 1) + D) * D);
  for (int k = 0; k < D; ++k) {
    double v;
    if (rng_unit(seed, cs, uint64_t(m) * (D + 1) + k) < CR || k == jr) {
      v = pop[size_t(r1) * D + k] + F * (pop[size_t(r2) * D + k] - pop[size_t(r3) * D + k]);
      if (v < lo[k]) v = lo[k];
      if (v > hi[k]) v = hi[k];
    } else {
      v = pop[size_t(m) * D + k];
    }
    out[k] = v;
  }
}

int de_run(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
           double *best, double *hist, int *gens) {
  const int D = c.D, NP = o->np;
  FS_REQUIRE(NP >= 4, FS_ERR_ARG, "DE needs a population of at least 4");
  FS_REQUIRE(o->generations >= 0, FS_ERR_ARG, "negative generation count");
  std::vector<double> pop(size_t(NP) * D), trial(size_t(NP) * D), cost(NP), tcost(NP);
  for (int k = 0; k < D; ++k) pop[k] = std::min(std::max(x[k], lo[k]), hi[k]);
  for (int m = 1; m < NP; ++m)
    for (int k = 0; k < D; ++k)
      pop[size_t(m) * D + k] = lo[k] + rng_unit(o->seed, 0, uint64_t(m) * D + k) * (hi[k] - lo[k]);
  fs_population_impl(c, pop.data(), NP, cost.data());
  if (o->surrogate) surrogate_begin(c, o, pop, cost, lo, hi);
  int g = 1;
  double b = std::numeric_limits<double>::infinity();
  for (int m = 0; m < NP; ++m) b = std::min(b, cost[m]);
  for (; g <= o->generations; ++g) {
    if (o->surrogate) {
      surrogate_maybe_train(c, o, g);
      surrogate_screen(c, o, g, pop, cost, lo, hi, trial);
    } else {
      for (int m = 0; m < NP; ++m)
        de_trial(o->seed, g, m, NP, D, o->F, o->CR, pop.data(), lo, hi, trial.data() + size_t(m) * D,
                 0);
    }
    fs_population_impl(c, trial.data(), NP, tcost.data());
    if (o->surrogate) surrogate_observe(c, o, trial, tcost, lo, hi);
    b = std::numeric_limits<double>::infinity();
    for (int m = 0; m < NP; ++m) {
      if (tcost[m] <= cost[m]) {
        cost[m] = tcost[m];
        std::copy(trial.begin() + size_t(m) * D, trial.begin() + size_t(m + 1) * D,
                  pop.begin() + size_t(m) * D);
      }
      b = std::min(b, cost[m]);
    }
    if (hist) hist[g - 1] = b;
    if (o->target_cost > 0.0 && b <= o->target_cost) {
      ++g;
      break;
    }
  }
  *gens = g - 1;
  int bm = 0;
  for (int m = 1; m < NP; ++m)
    if (cost[m] < cost[bm]) bm = m;
  std::copy(pop.begin() + size_t(bm) * D, pop.begin() + size_t(bm + 1) * D, x);
  if (best) *best = cost[bm];
  return FS_OK;
}

// ---------------------------------------------------------------------------
// MLP surrogate screening
//   training set: every exactly evaluated (x, cost) pair, x mapped to [-1, 1]
//   over the box, target log(cost) standardised; retrained every
//   retrain_every generations (minibatch Adam, batches drawn uniformly from
//   all pairs, gathered on the device); screening draws `oversample` trial
//   vectors per member (candidate 0 is the plain DE trial), ranks them by the
//   predicted cost and sends each member's best one to the exact evaluation;
//   when the Kendall tau between predicted and exact costs of a screened
//   generation falls below rank_tau_min the driver proposes plain trials until
//   the next retrain.
// ---------------------------------------------------------------------------
namespace {

__global__ void k_gather_rows(int B, int D, const int32_t *__restrict__ idx,
                              const float *__restrict__ X, const float *__restrict__ t,
                              float *__restrict__ bx, float *__restrict__ bt) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(B) * D) return;
  const int b = int(e / D), k = int(e % D);
  const int64_t r = idx[b];
  bx[e] = X[r * D + k];
  if (k == 0) bt[b] = t[r];
}

} // namespace

struct SurrogateState {
  Mlp mlp;
  bool ready = false, fallback = false;
  std::vector<float> hX, ht; // normalised inputs, log costs
  DBuf<float> dX, dt, bX, bt, cand, pred;
  DBuf<int32_t> idx;
  int64_t on_device = 0;
  double t_mean = 0.0, t_std = 1.0;
  std::vector<float> chosen_pred;
  uint64_t rng_ctr = 0;
  int retrains = 0, fallbacks = 0, screened = 0;
};

void destroy_surrogate(SurrogateState *s) { delete s; }

namespace {

void normalise(const double *x, const double *lo, const double *hi, int D, float *out) {
  for (int k = 0; k < D; ++k) {
    const double w = hi[k] - lo[k];
    out[k] = w > 0 ? float(2.0 * (x[k] - lo[k]) / w - 1.0) : 0.0f;
  }
}

void add_samples(SurrogateState &S, const double *X, const double *cost, int n, int D,
                 const double *lo, const double *hi) {
  const size_t o = S.hX.size();
  S.hX.resize(o + size_t(n) * D);
  for (int m = 0; m < n; ++m) {
    normalise(X + size_t(m) * D, lo, hi, D, S.hX.data() + o + size_t(m) * D);
    S.ht.push_back(float(std::log(std::max(cost[m], 1e-300))));
  }
}

void train(Context &c, SurrogateState &S, int D, int hidden, uint64_t seed) {
  const int64_t n = int64_t(S.ht.size());
  if (S.mlp.D == 0) {
    mlp_init(S.mlp, D, hidden);
    // Glorot-style normal init from the counter RNG (Box-Muller)
    std::vector<float> p(S.mlp.n_params(), 0.0f);
    auto normal = [&](uint64_t i) {
      const double u1 = std::max(rng_unit(seed, 7777, 2 * i), 1e-300);
      const double u2 = rng_unit(seed, 7777, 2 * i + 1);
      return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    size_t o = 0;
    for (size_t k = 0; k < size_t(hidden) * D; ++k) p[o++] = float(normal(o) / std::sqrt(double(D)));
    o += hidden;
    for (size_t k = 0; k < size_t(hidden) * hidden; ++k)
      p[o] = float(normal(o) / std::sqrt(double(hidden))), ++o;
    o += hidden;
    for (int k = 0; k < hidden; ++k) p[o] = float(normal(o) / std::sqrt(double(hidden))), ++o;
    FS_HIP(hipMemcpyAsync(S.mlp.P.p, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice,
                          c.stream));
    FS_HIP(hipMemsetAsync(S.mlp.Mo.p, 0, sizeof(float) * p.size(), c.stream));
    FS_HIP(hipMemsetAsync(S.mlp.Vo.p, 0, sizeof(float) * p.size(), c.stream));
  }
  // standardise the targets over all pairs
  double mean = 0.0, sq = 0.0;
  for (float v : S.ht) mean += v;
  mean /= double(n);
  for (float v : S.ht) sq += (v - mean) * (v - mean);
  S.t_mean = mean;
  S.t_std = std::sqrt(sq / double(std::max<int64_t>(n - 1, 1))) + 1e-12;
  std::vector<float> ts(S.ht.size());
  for (size_t k = 0; k < ts.size(); ++k) ts[k] = float((S.ht[k] - mean) / S.t_std);
  S.dX.alloc(S.hX.size());
  S.dt.alloc(ts.size());
  FS_HIP(hipMemcpyAsync(S.dX.p, S.hX.data(), sizeof(float) * S.hX.size(), hipMemcpyHostToDevice,
                        c.stream));
  FS_HIP(hipMemcpyAsync(S.dt.p, ts.data(), sizeof(float) * ts.size(), hipMemcpyHostToDevice,
                        c.stream));
  const int B = int(std::min<int64_t>(n, 2048));
  const int steps = 200;
  S.bX.alloc(size_t(B) * D);
  S.bt.alloc(B);
  S.idx.alloc(B);
  std::vector<int32_t> idx(B);
  for (int st = 0; st < steps; ++st) {
    for (int b = 0; b < B; ++b)
      idx[b] = int32_t(rng_unit(seed, 9999, S.rng_ctr++) * double(n)) % int32_t(n);
    FS_HIP(hipMemcpyAsync(S.idx.p, idx.data(), sizeof(int32_t) * B, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_gather_rows, dim3(unsigned((int64_t(B) * D + 255) / 256)), dim3(256), 0,
                       c.stream, B, D, S.idx.p, S.dX.p, S.dt.p, S.bX.p, S.bt.p);
    FS_HIP(hipGetLastError());
    mlp_backward(S.mlp, c.stream, B, S.bX.p, S.bt.p);
    mlp_adam(S.mlp, c.stream, 1e-3f);
    // the index buffer is rewritten next step: keep host and device in order
    FS_HIP(hipStreamSynchronize(c.stream));
  }
  S.ready = true;
  S.fallback = false;
  ++S.retrains;
}

double kendall_tau(const std::vector<float> &a, const std::vector<double> &b, int nmax) {
  const int n = std::min<int>(int(a.size()), nmax);
  long conc = 0, disc = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const double s = double(a[i] - a[j]) * (b[i] - b[j]);
      if (s > 0) ++conc;
      else if (s < 0) ++disc;
    }
  const long tot = conc + disc;
  return tot ? double(conc - disc) / double(tot) : 1.0;
}

} // namespace

int surrogate_screen(Context &c, const fs_de_opts *o, int gen, const std::vector<double> &pop,
                     const std::vector<double> &cost, const double *lo, const double *hi,
                     std::vector<double> &trial) {
  (void)cost;
  const int D = c.D, NP = o->np;
  SurrogateState &S = *c.surr;
  for (int m = 0; m < NP; ++m)
    de_trial(o->seed, gen, m, NP, D, o->F, o->CR, pop.data(), lo, hi, trial.data() + size_t(m) * D, 0);
  S.chosen_pred.clear();
  const int O = std::max(1, o->oversample);
  if (!S.ready || S.fallback || O == 1) return FS_OK;
  std::vector<double> cands(size_t(O) * NP * D);
  std::copy(trial.begin(), trial.end(), cands.begin());
  for (int r = 1; r < O; ++r)
    for (int m = 0; m < NP; ++m)
      de_trial(o->seed, gen, m, NP, D, o->F, o->CR, pop.data(), lo, hi,
               cands.data() + (size_t(r) * NP + m) * D, uint64_t(r) << 32);
  std::vector<float> xn(cands.size());
  for (size_t q = 0; q < size_t(O) * NP; ++q) normalise(cands.data() + q * D, lo, hi, D, xn.data() + q * D);
  S.cand.alloc(xn.size());
  S.pred.alloc(size_t(O) * NP);
  FS_HIP(hipMemcpyAsync(S.cand.p, xn.data(), sizeof(float) * xn.size(), hipMemcpyHostToDevice,
                        c.stream));
  mlp_forward(S.mlp, c.stream, O * NP, S.cand.p, S.pred.p);
  std::vector<float> pr(size_t(O) * NP);
  FS_HIP(hipMemcpyAsync(pr.data(), S.pred.p, sizeof(float) * pr.size(), hipMemcpyDeviceToHost,
                        c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
  S.chosen_pred.resize(NP);
  for (int m = 0; m < NP; ++m) {
    int best = 0;
    for (int r = 1; r < O; ++r)
      if (pr[size_t(r) * NP + m] < pr[size_t(best) * NP + m]) best = r;
    S.chosen_pred[m] = pr[size_t(best) * NP + m];
    if (best != 0)
      std::copy(cands.begin() + (size_t(best) * NP + m) * D,
                cands.begin() + (size_t(best) * NP + m + 1) * D, trial.begin() + size_t(m) * D);
  }
  ++S.screened;
  return FS_OK;
}

void surrogate_observe(Context &c, const fs_de_opts *o, const std::vector<double> &trial,
                       const std::vector<double> &tcost, const double *lo, const double *hi) {
  SurrogateState &S = *c.surr;
  add_samples(S, trial.data(), tcost.data(), o->np, c.D, lo, hi);
  if (!S.chosen_pred.empty()) {
    std::vector<double> lc(tcost.size());
    for (size_t k = 0; k < lc.size(); ++k) lc[k] = std::log(std::max(tcost[k], 1e-300));
    const double tau = kendall_tau(S.chosen_pred, lc, 512);
    if (tau < o->rank_tau_min) {
      S.fallback = true;
      ++S.fallbacks;
    }
  }
}

void surrogate_begin(Context &c, const fs_de_opts *o, const std::vector<double> &pop,
                     const std::vector<double> &cost, const double *lo, const double *hi) {
  destroy_surrogate(c.surr);
  c.surr = new SurrogateState();
  add_samples(*c.surr, pop.data(), cost.data(), o->np, c.D, lo, hi);
}

void surrogate_maybe_train(Context &c, const fs_de_opts *o, int gen) {
  SurrogateState &S = *c.surr;
  const int every = std::max(1, o->retrain_every);
  if ((gen - 1) % every == 0 && int64_t(S.ht.size()) >= 2 * int64_t(o->np))
    train(c, S, c.D, o->hidden > 0 ? o->hidden : 256, o->seed);
}

} // namespace fsa
