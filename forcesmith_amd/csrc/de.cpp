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
#include <chrono>
#include <cmath>
#include <cstdio>
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

int de_run_device(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
                  double *best, double *hist, int *gens);
void record_de_stats(Context &c, int gens, bool screened_run);

int de_run(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
           double *best, double *hist, int *gens) {
  if (population_supported(c)) return de_run_device(c, o, lo, hi, x, best, hist, gens);
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
  record_de_stats(c, *gens, o->surrogate != 0);
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
//   over the box, target log(cost) standardised; the pairs stay on the device
//   (appended each generation); retrained every retrain_every generations
//   (minibatch Adam, batches drawn uniformly from all pairs by a counter RNG
//   on the device); screening draws `oversample` trial
//   vectors per member (candidate 0 is the plain DE trial), ranks them by the
//   predicted cost and sends each member's best one to the exact evaluation;
//   when the Kendall tau between predicted and exact costs of a screened
//   generation falls below rank_tau_min the driver proposes plain trials until
//   the next retrain.
// ---------------------------------------------------------------------------
namespace {

__device__ inline uint64_t smix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// minibatch of B rows drawn uniformly (counter RNG, the host stream of
// rng_unit) from the n stored samples; targets standardised on the fly
__global__ void k_gather_rows(int B, int D, int64_t n, uint64_t seed, uint64_t ctr,
                              const float *__restrict__ X, const float *__restrict__ lt, float mean,
                              float inv_std, float *__restrict__ bx, float *__restrict__ bt) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(B) * D) return;
  const int b = int(e / D), k = int(e % D);
  const uint64_t u = smix(smix(seed ^ smix(9999ull)) + ctr + uint64_t(b));
  int64_t r = int64_t(double(u >> 11) * 0x1.0p-53 * double(n));
  if (r >= n) r = n - 1;
  bx[e] = X[r * D + k];
  if (k == 0) bt[b] = (lt[r] - mean) * inv_std;
}

} // namespace

struct SurrogateState {
  Mlp mlp;
  bool ready = false, fallback = false;
  // every exactly evaluated pair, resident on the device: inputs mapped to
  // [-1, 1] over the box and log costs; the host keeps the log costs only
  DBuf<float> dX, dlt, bX, bt, cand, pred;
  int64_t n = 0, cap = 0;
  std::vector<float> ht;
  std::vector<float> stage;
  double t_mean = 0.0, t_std = 1.0;
  std::vector<float> chosen_pred;
  uint64_t rng_ctr = 0;
  int retrains = 0, fallbacks = 0, screened = 0;
  float first_loss = 0.0f, last_loss = 0.0f;
};

void destroy_surrogate(SurrogateState *s) { delete s; }

// statistics of the run that just ended (the surrogate state of a plain run
// is that of an earlier call, so it is read only when this run screened)
void record_de_stats(Context &c, int gens, bool screened_run) {
  fs_de_stats st{};
  st.generations = gens;
  if (screened_run && c.surr) {
    st.retrains = c.surr->retrains;
    st.screened = c.surr->screened;
    st.fallbacks = c.surr->fallbacks;
    st.first_loss = c.surr->first_loss;
    st.last_loss = c.surr->last_loss;
  }
  c.de_stats = st;
}

namespace {

void normalise(const double *x, const double *lo, const double *hi, int D, float *out) {
  for (int k = 0; k < D; ++k) {
    const double w = hi[k] - lo[k];
    out[k] = w > 0 ? float(2.0 * (x[k] - lo[k]) / w - 1.0) : 0.0f;
  }
}

void add_samples(Context &c, SurrogateState &S, const double *X, const double *cost, int n, int D,
                 const double *lo, const double *hi) {
  FS_REQUIRE(S.n + n <= S.cap, FS_ERR_STATE, "surrogate sample store is full");
  S.stage.resize(size_t(n) * D);
  std::vector<float> lt(n);
  for (int m = 0; m < n; ++m) {
    normalise(X + size_t(m) * D, lo, hi, D, S.stage.data() + size_t(m) * D);
    lt[m] = float(std::log(std::max(cost[m], 1e-300)));
    S.ht.push_back(lt[m]);
  }
  FS_HIP(hipMemcpyAsync(S.dX.p + size_t(S.n) * D, S.stage.data(), sizeof(float) * S.stage.size(),
                        hipMemcpyHostToDevice, c.stream));
  FS_HIP(hipMemcpyAsync(S.dlt.p + S.n, lt.data(), sizeof(float) * n, hipMemcpyHostToDevice, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream)); // the staging buffers are reused
  S.n += n;
}

void train(Context &c, SurrogateState &S, int D, int hidden, uint64_t seed) {
  const int64_t n = S.n;
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
  const int B = int(std::min<int64_t>(n, 2048));
  const int steps = 200;
  S.bX.alloc(size_t(B) * D);
  S.bt.alloc(B);
  for (int st = 0; st < steps; ++st) {
    hipLaunchKernelGGL(k_gather_rows, dim3(unsigned((int64_t(B) * D + 255) / 256)), dim3(256), 0,
                       c.stream, B, D, n, seed, S.rng_ctr, S.dX.p, S.dlt.p, float(S.t_mean),
                       float(1.0 / S.t_std), S.bX.p, S.bt.p);
    FS_HIP(hipGetLastError());
    S.rng_ctr += uint64_t(B);
    mlp_backward(S.mlp, c.stream, B, S.bX.p, S.bt.p);
    if (st == 0 && S.retrains == 0)
      FS_HIP(hipMemcpyAsync(&S.first_loss, S.mlp.loss.p, sizeof(float), hipMemcpyDeviceToHost,
                            c.stream));
    mlp_adam(S.mlp, c.stream, 1e-3f);
  }
  FS_HIP(hipMemcpyAsync(&S.last_loss, S.mlp.loss.p, sizeof(float), hipMemcpyDeviceToHost, c.stream));
  FS_HIP(hipStreamSynchronize(c.stream));
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
  add_samples(c, S, trial.data(), tcost.data(), o->np, c.D, lo, hi);
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
  SurrogateState &S = *c.surr;
  S.cap = int64_t(o->np) * (int64_t(std::max(o->generations, 0)) + 1);
  S.dX.alloc(size_t(S.cap) * c.D);
  S.dlt.alloc(size_t(S.cap));
  add_samples(c, S, pop.data(), cost.data(), o->np, c.D, lo, hi);
}

void surrogate_maybe_train(Context &c, const fs_de_opts *o, int gen) {
  SurrogateState &S = *c.surr;
  const int every = std::max(1, o->retrain_every);
  if ((gen - 1) % every == 0 && S.n >= 2 * int64_t(o->np))
    train(c, S, c.D, o->hidden > 0 ? o->hidden : 256, o->seed);
}


// ---------------------------------------------------------------------------
// Device-resident DE (EAM, mixed or FP64): population, trials, costs and
// selection never leave the GPU. A generation is
//   k_de_trials -> population_launch -> k_de_select -> k_de_best
// with the generation index in device memory, so the plain-mode generation
// is captured once into a hipGraph and replayed; the host reads the best-cost
// history back every few generations (early stop on the target cost).
// Arithmetic and RNG are those of de_trial / rng_unit above (FP contraction
// off), so the device and host drivers produce the same trials.
// ---------------------------------------------------------------------------
namespace {

__device__ inline uint64_t dsmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double dunit(uint64_t seed, uint64_t stream, uint64_t index) {
  const uint64_t u = dsmix(dsmix(seed ^ dsmix(stream)) + index);
  return double(u >> 11) * 0x1.0p-53;
}

__device__ inline int dpick(uint64_t seed, uint64_t stream, int m, int np, int &t, int e0, int e1,
                            int e2) {
  for (;;) {
    const int r = int(dunit(seed, stream, uint64_t(m) * 1024u + uint64_t(t)) * np);
    ++t;
    if (r != e0 && r != e1 && r != e2 && r < np) return r;
  }
}

__global__ void k_de_init(int NP, int D, uint64_t seed, const double *__restrict__ x0,
                          const double *__restrict__ lo, const double *__restrict__ hi,
                          double *__restrict__ pop) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= int64_t(NP) * D) return;
  const int m = int(e / D), k = int(e % D);
  double v;
  if (m == 0) v = fmin(fmax(x0[k], lo[k]), hi[k]);
  else v = lo[k] + dunit(seed, 0, uint64_t(m) * D + k) * (hi[k] - lo[k]);
  pop[e] = v;
}

// trial r of member m (r = 0: the plain trial; r > 0: oversampled candidates
// on the stream base r << 32); one thread per (member, candidate)
__global__ void k_de_trials(int NP, int D, int O, uint64_t seed, double F, double CR,
                            const int32_t *__restrict__ gen, const double *__restrict__ pop,
                            const double *__restrict__ lo, const double *__restrict__ hi,
                            double *__restrict__ out) {
  const int64_t id = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id >= int64_t(NP) * O) return;
  const int r = int(id / NP), m = int(id % NP);
  const int g = *gen;
  int t = 0;
  const uint64_t ps = 2u * uint64_t(g) + (uint64_t(r) << 32);
  const int r1 = dpick(seed, ps, m, NP, t, m, -1, -1);
  const int r2 = dpick(seed, ps, m, NP, t, m, r1, -1);
  const int r3 = dpick(seed, ps, m, NP, t, m, r1, r2);
  const uint64_t cs = ps + 1u;
  const int jr = int(dunit(seed, cs, uint64_t(m) * (D + 1) + D) * D);
  double *o = out + (size_t(r) * NP + m) * D;
  for (int k = 0; k < D; ++k) {
    double v;
    if (dunit(seed, cs, uint64_t(m) * (D + 1) + k) < CR || k == jr) {
      v = pop[size_t(r1) * D + k] + F * (pop[size_t(r2) * D + k] - pop[size_t(r3) * D + k]);
      if (v < lo[k]) v = lo[k];
      if (v > hi[k]) v = hi[k];
    } else {
      v = pop[size_t(m) * D + k];
    }
    o[k] = v;
  }
}

// greedy selection (the reference keeps the trial when its cost is not worse)
__global__ void k_de_select(int NP, int D, const double *__restrict__ trial,
                            const double *__restrict__ tcost, double *__restrict__ pop,
                            double *__restrict__ cost) {
  const int m = blockIdx.x;
  const bool take = tcost[m] <= cost[m];
  if (!take) return;
  for (int k = threadIdx.x; k < D; k += blockDim.x) pop[size_t(m) * D + k] = trial[size_t(m) * D + k];
  if (threadIdx.x == 0) cost[m] = tcost[m];
}

// best cost of the population into hist[g - 1] and the next generation index
__global__ void k_de_best(int NP, const double *__restrict__ cost, int32_t *__restrict__ gen,
                          double *__restrict__ hist) {
  __shared__ double sv[256];
  double b = INFINITY;
  for (int m = threadIdx.x; m < NP; m += blockDim.x) b = fmin(b, cost[m]);
  sv[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sv[threadIdx.x] = fmin(sv[threadIdx.x], sv[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int g = *gen;
    hist[g - 1] = sv[0];
    *gen = g + 1;
  }
}

// screened trial per member: the candidate with the lowest predicted cost
__global__ void k_de_pick(int NP, int D, int O, const double *__restrict__ cands,
                          const float *__restrict__ pred, double *__restrict__ trial,
                          float *__restrict__ chosen) {
  const int m = blockIdx.x;
  int best = 0;
  for (int r = 1; r < O; ++r)
    if (pred[size_t(r) * NP + m] < pred[size_t(best) * NP + m]) best = r;
  const double *src = cands + (size_t(best) * NP + m) * D;
  for (int k = threadIdx.x; k < D; k += blockDim.x) trial[size_t(m) * D + k] = src[k];
  if (threadIdx.x == 0) chosen[m] = pred[size_t(best) * NP + m];
}

// inputs mapped to [-1, 1] over the box (normalise() above)
__global__ void k_de_normalise(int64_t n, int D, const double *__restrict__ X,
                               const double *__restrict__ lo, const double *__restrict__ hi,
                               float *__restrict__ out) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * D) return;
  const int k = int(e % D);
  const double w = hi[k] - lo[k];
  out[e] = w > 0 ? float(2.0 * (X[e] - lo[k]) / w - 1.0) : 0.0f;
}

__global__ void k_de_logcost(int n, const double *__restrict__ c, float *__restrict__ out) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) out[m] = float(log(fmax(c[m], 1e-300)));
}

inline unsigned nblk(int64_t n, int b) { return unsigned((n + b - 1) / b); }

struct DeDev {
  DBuf<double> pop, trial, cand, cost, tcost, lo, hi, x0, hist;
  DBuf<int32_t> gen;
  DBuf<float> chosen;
};

} // namespace

int de_run_device(Context &c, const fs_de_opts *o, const double *lo, const double *hi, double *x,
                  double *best, double *hist, int *gens) {
  const int D = c.D, NP = o->np, G = std::max(o->generations, 0);
  FS_REQUIRE(NP >= 4, FS_ERR_ARG, "DE needs a population of at least 4");
  const int O = o->surrogate ? std::max(1, o->oversample) : 1;
  hipStream_t st = c.stream;
  DeDev d;
  d.pop.alloc(size_t(NP) * D);
  d.trial.alloc(size_t(NP) * D);
  d.cand.alloc(size_t(O) * NP * D);
  d.cost.alloc(NP);
  d.tcost.alloc(NP);
  d.lo.alloc(D);
  d.hi.alloc(D);
  d.x0.alloc(D);
  d.hist.alloc(std::max(G, 1));
  d.gen.alloc(1);
  d.chosen.alloc(NP);
  FS_HIP(hipMemcpyAsync(d.lo.p, lo, 8 * D, hipMemcpyHostToDevice, st));
  FS_HIP(hipMemcpyAsync(d.hi.p, hi, 8 * D, hipMemcpyHostToDevice, st));
  FS_HIP(hipMemcpyAsync(d.x0.p, x, 8 * D, hipMemcpyHostToDevice, st));
  const int32_t one = 1;
  FS_HIP(hipMemcpyAsync(d.gen.p, &one, 4, hipMemcpyHostToDevice, st));
  population_prepare(c, NP);
  hipLaunchKernelGGL(k_de_init, dim3(nblk(int64_t(NP) * D, 256)), dim3(256), 0, st, NP, D, o->seed,
                     d.x0.p, d.lo.p, d.hi.p, d.pop.p);
  FS_HIP(hipGetLastError());
  population_tune_launch(c, d.pop.p, NP, d.cost.p);
  FS_HIP(hipStreamSynchronize(st));

  auto generation = [&](bool screened) {
    if (!screened) {
      hipLaunchKernelGGL(k_de_trials, dim3(nblk(NP, 64)), dim3(64), 0, st, NP, D, 1, o->seed, o->F,
                         o->CR, d.gen.p, d.pop.p, d.lo.p, d.hi.p, d.trial.p);
      FS_HIP(hipGetLastError());
    }
    population_launch(c, d.trial.p, NP, d.tcost.p);
    hipLaunchKernelGGL(k_de_select, dim3(NP), dim3(64), 0, st, NP, D, d.trial.p, d.tcost.p, d.pop.p,
                       d.cost.p);
    FS_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_de_best, dim3(1), dim3(256), 0, st, NP, d.cost.p, d.gen.p, d.hist.p);
    FS_HIP(hipGetLastError());
  };

  std::vector<double> h(std::max(G, 1));
  int done = 0;
  const auto t_start = std::chrono::steady_clock::now();
  auto progress = [&](int gdone) {
    if (o->progress_every <= 0 || gdone % o->progress_every != 0) return;
    double b = 0.0;
    FS_HIP(hipMemcpyAsync(&b, d.hist.p + (gdone - 1), 8, hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    std::fprintf(stderr, "de: generation %d best %.9e elapsed %.1f s\n", gdone, b, el);
    std::fflush(stderr);
  };
  const int chunk = 16; // generations between history reads (early-stop checks)
  auto reached = [&](int upto) {
    if (o->target_cost <= 0.0) return -1;
    FS_HIP(hipMemcpyAsync(h.data(), d.hist.p, 8 * size_t(upto), hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    for (int g = 0; g < upto; ++g)
      if (h[g] <= o->target_cost) return g + 1;
    return -1;
  };
  int stop = -1;
  if (!o->surrogate) {
    // one generation captured as a graph, replayed G times
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (G > 0) {
      FS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        generation(false);
      } catch (...) {
        (void)hipStreamEndCapture(st, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
      }
      FS_HIP(hipStreamEndCapture(st, &graph));
      FS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    try {
      while (done < G && stop < 0) {
        const int n = std::min(chunk, G - done);
        for (int k = 0; k < n; ++k) {
          FS_HIP(hipGraphLaunch(exec, st));
          progress(done + k + 1);
        }
        done += n;
        stop = reached(done);
      }
    } catch (...) {
      (void)hipStreamSynchronize(st);
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      throw;
    }
    // the replays are only queued (without a target cost nothing above waits
    // for them): the executable graph and its kernel arguments must outlive
    // every queued launch, so wait before destroying it
    FS_HIP(hipStreamSynchronize(st));
    if (exec) FS_HIP(hipGraphExecDestroy(exec));
    if (graph) FS_HIP(hipGraphDestroy(graph));
  } else {
    // screened generations: host decisions (retrain, rank-error fallback)
    // between device steps
    std::vector<double> popv(size_t(NP) * D), costv(NP);
    FS_HIP(hipMemcpyAsync(popv.data(), d.pop.p, 8 * popv.size(), hipMemcpyDeviceToHost, st));
    FS_HIP(hipMemcpyAsync(costv.data(), d.cost.p, 8 * NP, hipMemcpyDeviceToHost, st));
    FS_HIP(hipStreamSynchronize(st));
    surrogate_begin(c, o, popv, costv, lo, hi);
    SurrogateState &S = *c.surr;
    std::vector<float> chosen(NP);
    std::vector<double> tc(NP);
    while (done < G && stop < 0) {
      const int g = done + 1;
      surrogate_maybe_train(c, o, g);
      const bool screen = S.ready && !S.fallback && O > 1;
      hipLaunchKernelGGL(k_de_trials, dim3(nblk(int64_t(NP) * (screen ? O : 1), 64)), dim3(64), 0, st,
                         NP, D, screen ? O : 1, o->seed, o->F, o->CR, d.gen.p, d.pop.p, d.lo.p, d.hi.p,
                         screen ? d.cand.p : d.trial.p);
      FS_HIP(hipGetLastError());
      if (screen) {
        S.cand.alloc(size_t(O) * NP * D);
        S.pred.alloc(size_t(O) * NP);
        hipLaunchKernelGGL(k_de_normalise, dim3(nblk(int64_t(O) * NP * D, 256)), dim3(256), 0, st,
                           int64_t(O) * NP, D, d.cand.p, d.lo.p, d.hi.p, S.cand.p);
        FS_HIP(hipGetLastError());
        mlp_forward(S.mlp, st, O * NP, S.cand.p, S.pred.p);
        hipLaunchKernelGGL(k_de_pick, dim3(NP), dim3(64), 0, st, NP, D, O, d.cand.p, S.pred.p,
                           d.trial.p, d.chosen.p);
        FS_HIP(hipGetLastError());
        ++S.screened;
      }
      generation(true);
      // observe: append the exactly judged trials to the training store
      FS_REQUIRE(S.n + NP <= S.cap, FS_ERR_STATE, "surrogate sample store is full");
      hipLaunchKernelGGL(k_de_normalise, dim3(nblk(int64_t(NP) * D, 256)), dim3(256), 0, st,
                         int64_t(NP), D, d.trial.p, d.lo.p, d.hi.p, S.dX.p + size_t(S.n) * D);
      FS_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_de_logcost, dim3(nblk(NP, 256)), dim3(256), 0, st, NP, d.tcost.p,
                         S.dlt.p + S.n);
      FS_HIP(hipGetLastError());
      FS_HIP(hipMemcpyAsync(tc.data(), d.tcost.p, 8 * NP, hipMemcpyDeviceToHost, st));
      if (screen)
        FS_HIP(hipMemcpyAsync(chosen.data(), d.chosen.p, 4 * NP, hipMemcpyDeviceToHost, st));
      FS_HIP(hipStreamSynchronize(st));
      S.n += NP;
      for (int m = 0; m < NP; ++m) S.ht.push_back(float(std::log(std::max(tc[m], 1e-300))));
      if (screen) {
        std::vector<double> lc(NP);
        for (int m = 0; m < NP; ++m) lc[m] = std::log(std::max(tc[m], 1e-300));
        if (kendall_tau(chosen, lc, 512) < o->rank_tau_min) {
          S.fallback = true;
          ++S.fallbacks;
        }
      }
      ++done;
      progress(done);
      if (o->target_cost > 0.0) {
        double b = 0.0;
        FS_HIP(hipMemcpyAsync(&b, d.hist.p + (done - 1), 8, hipMemcpyDeviceToHost, st));
        FS_HIP(hipStreamSynchronize(st));
        if (b <= o->target_cost) stop = done;
      }
    }
  }
  const int ran = stop > 0 ? stop : done;
  *gens = ran;
  record_de_stats(c, ran, o->surrogate != 0);
  // best member
  std::vector<double> costv(NP);
  FS_HIP(hipMemcpyAsync(costv.data(), d.cost.p, 8 * NP, hipMemcpyDeviceToHost, st));
  FS_HIP(hipMemcpyAsync(h.data(), d.hist.p, 8 * size_t(std::max(done, 1)), hipMemcpyDeviceToHost, st));
  FS_HIP(hipStreamSynchronize(st));
  int bm = 0;
  for (int m = 1; m < NP; ++m)
    if (costv[m] < costv[bm]) bm = m;
  FS_HIP(hipMemcpyAsync(x, d.pop.p + size_t(bm) * D, 8 * D, hipMemcpyDeviceToHost, st));
  FS_HIP(hipStreamSynchronize(st));
  if (best) *best = costv[bm];
  if (hist)
    for (int g = 0; g < ran; ++g) hist[g] = h[g];
  return FS_OK;
}

} // namespace fsa
