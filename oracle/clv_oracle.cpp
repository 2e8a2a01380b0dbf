// clv_oracle.cpp -- FP64 CPU oracle of the complex-Langevin update.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the timed CPU
// baseline.  Never linked into libclv.so.  Every function cites the reference
// file:line whose behaviour it restates (paths relative to the reference
// repository DoAmann/CL_var_dt).
#include "clv_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kTwoPi = 6.283185307179586;  // nearest double to 2*pi

// ---------------------------------------------------------------------------
// Philox4x32-10, from the published description (Salmon et al., SC'11):
// round: (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)),
// key bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------
inline void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = uint32_t(p1);
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = uint32_t(p0);
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Engine noise map: counter = (site, pair | stream<<8, step_lo, step_hi),
// key = seed; uniforms u = (x + 1/2) 2^-32; Box-Muller (1958) on (u0,u1) and
// (u2,u3).  Replaces curandGenerateNormalDouble (Propagator_without_cutoff.cu:432).
inline void normals4(uint64_t seed, uint32_t stream, uint64_t step, uint32_t site,
                     uint32_t pair, double out[4]) {
  uint32_t c[4] = {site, pair | (stream << 8), uint32_t(step), uint32_t(step >> 32)};
  philox10(c, uint32_t(seed), uint32_t(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = (double(c[0]) + 0.5) * s, u1 = (double(c[1]) + 0.5) * s;
  const double u2 = (double(c[2]) + 0.5) * s, u3 = (double(c[3]) + 0.5) * s;
  const double r0 = std::sqrt(-2.0 * std::log(u0));
  const double r1 = std::sqrt(-2.0 * std::log(u2));
  const double t0 = kTwoPi * u1, t1 = kTwoPi * u3;
  out[0] = r0 * std::cos(t0);
  out[1] = r0 * std::sin(t0);
  out[2] = r1 * std::cos(t1);
  out[3] = r1 * std::sin(t1);
}

// ---------------------------------------------------------------------------
// Mixed-radix Stockham autosort FFT (Van Loan, Computational Frameworks for
// the FFT, ch. 1.7), natural order in and out.  Stage with radix R and span
// p: for i < n/R, k = i mod p, x_r = in[i + r n/R] w^{rk/(pR)}, DFT_R,
// out[(i-k)R + k + r p] = X_r.  Unnormalised; sign as cuFFT's exponent.
// ---------------------------------------------------------------------------
struct FftPlan {
  int n = 0;
  int sign = 0;
  std::vector<int> radix;
  std::vector<cd> w;  // w[m] = exp(sign 2 pi i m / n)
};

FftPlan make_plan(int n, int sign) {
  FftPlan pl;
  pl.n = n;
  pl.sign = sign;
  int m = n;
  while (m % 4 == 0) { pl.radix.push_back(4); m /= 4; }
  while (m % 2 == 0) { pl.radix.push_back(2); m /= 2; }
  for (int f = 3; m > 1; f += 2)
    while (m % f == 0) { pl.radix.push_back(f); m /= f; }
  pl.w.resize(n);
  for (int k = 0; k < n; ++k) {
    const double a = kTwoPi * double(k) / double(n);
    pl.w[k] = cd(std::cos(a), sign * std::sin(a));
  }
  return pl;
}

void fft_exec(const FftPlan &pl, cd *a, cd *work) {
  const int n = pl.n;
  if (n <= 1) return;
  cd *in = a, *out = work;
  int p = 1;
  cd x[64], y[64];
  for (int R : pl.radix) {
    const int nr = n / R;
    const int step = n / (p * R);  // twiddle index unit for w^{1/(pR)}
    const int wr = n / R;          // index unit for the radix-R root
    for (int i = 0; i < nr; ++i) {
      const int k = i % p;
      for (int r = 0; r < R; ++r) {
        cd v = in[i + r * nr];
        if (r && k) v *= pl.w[(r * k * step) % n];
        x[r] = v;
      }
      for (int q = 0; q < R; ++q) {
        cd s = x[0];
        for (int r = 1; r < R; ++r) s += x[r] * pl.w[((r * q) % R) * wr];
        y[q] = s;
      }
      const int j = (i - k) * R + k;
      for (int r = 0; r < R; ++r) out[j + r * p] = y[r];
    }
    std::swap(in, out);
    p *= R;
  }
  if (in != a) std::memcpy(a, in, sizeof(cd) * n);
}

// Rank-d transform of one plane (X,Y,Z), z fastest: the batch member of
// cufftPlanMany(rank, {X[,Y[,Z]]}, ..., CUFFT_Z2Z, 4*Nt) (Lattice.cu:135-156).
struct Plane3 {
  int nx, ny, nz;
  FftPlan px, py, pz;
};

void fft_plane(const Plane3 &pl, cd *a, std::vector<cd> &line, std::vector<cd> &work) {
  const int X = pl.nx, Y = pl.ny, Z = pl.nz;
  if (Z > 1)
    for (int xy = 0; xy < X * Y; ++xy) fft_exec(pl.pz, a + size_t(xy) * Z, work.data());
  if (Y > 1)
    for (int x = 0; x < X; ++x)
      for (int z = 0; z < Z; ++z) {
        for (int y = 0; y < Y; ++y) line[y] = a[(size_t(x) * Y + y) * Z + z];
        fft_exec(pl.py, line.data(), work.data());
        for (int y = 0; y < Y; ++y) a[(size_t(x) * Y + y) * Z + z] = line[y];
      }
  if (X > 1)
    for (int yz = 0; yz < Y * Z; ++yz) {
      for (int x = 0; x < X; ++x) line[x] = a[size_t(x) * Y * Z + yz];
      fft_exec(pl.px, line.data(), work.data());
      for (int x = 0; x < X; ++x) a[size_t(x) * Y * Z + yz] = line[x];
    }
}

// p of compute_kinetic_part (Propagator_without_cutoff.cu:169-194).
inline double mom(int j, int n) {
  if (j < n / 2 || n == 1) return kTwoPi * double(j) / double(n);
  return -kTwoPi * double(n - j) / double(n);
}

}  // namespace

struct orc_ctx {
  clv_shape sh;
  clv_params p;
  int nthreads;
  int X, Y, Z, Nt;
  size_t Vs, S;  // spatial volume, sites per chain
  std::vector<cd> psi, psib, kin, kinb, drift, driftb;
  Plane3 fwd, inv;  // sign -1 / +1 plans
  std::vector<double> p2;  // -p^2 multiplier per spatial site (k-space)
  clv_step_stats st;
  // observable sums (reference accumulators, Observables_without_cutoff.cu:722-770)
  int64_t count;
  double nsum[4];
  double esum;
  std::vector<double> F[3];
};

namespace {

inline size_t idx(const orc_ctx *o, int c, int t, size_t s) {
  // ind() of Lattice.cu:2-5 with my_map's periodic wrap (Helpers.cu:1-15)
  if (t < 0) t += o->Nt;
  if (t >= o->Nt) t -= o->Nt;
  return (size_t(c) * o->Nt + t) * o->Vs + s;
}

// kin = (1/(a_x V_s)) F^{-s}[-p^2 F^{s}[f]] with s = +1 for psi (fft() is
// CUFFT_INVERSE on a non-conjugated lattice) and -1 for psibar, Lattice.cu:158-187
// and Propagator_without_cutoff.cu:410-424.
void kinetic_field(orc_ctx *o, const std::vector<cd> &f, std::vector<cd> &out, bool conj) {
  out = f;
  const Plane3 &first = conj ? o->fwd : o->inv;
  const Plane3 &second = conj ? o->inv : o->fwd;
  const double norm = 1.0 / double(o->p.a_x * double(o->Vs));
  const int nplanes = 4 * o->Nt;
#pragma omp parallel num_threads(o->nthreads)
  {
    std::vector<cd> line(std::max({o->X, o->Y, o->Z})), work(std::max({o->X, o->Y, o->Z}));
#pragma omp for schedule(static)
    for (int pl = 0; pl < nplanes; ++pl) {
      cd *a = out.data() + size_t(pl) * o->Vs;
      fft_plane(first, a, line, work);
      for (size_t s = 0; s < o->Vs; ++s) a[s] = -o->p2[s] * a[s];
      fft_plane(second, a, line, work);
      for (size_t s = 0; s < o->Vs; ++s) a[s] = norm * a[s];
    }
  }
}

void compute_kinetic(orc_ctx *o) {
  kinetic_field(o, o->psi, o->kin, false);
  kinetic_field(o, o->psib, o->kinb, true);
}

// calculate_drift_without_noise (Propagator_without_cutoff.cu:1-129) plus the
// spin term -c1 S_i (zero for c1 = 0) and the conj-drift switch.
void compute_drift(orc_ctx *o) {
  const clv_params &p = o->p;
  const int Nt = o->Nt;
  const size_t Vs = o->Vs;
  const double sfac[4] = {p.mu + p.P + p.Q, p.mu, p.mu, p.mu - p.P + p.Q};
  static const int ta[4] = {1, 0, 3, 2}, tb[4] = {3, 2, 1, 0}, tc[4] = {2, 3, 0, 1};
  static const int opp[4] = {3, 2, 1, 0};
#pragma omp parallel for collapse(2) schedule(static) num_threads(o->nthreads)
  for (int t = 0; t < Nt; ++t)
    for (long long sl = 0; sl < (long long)Vs; ++sl) {
      const size_t s = size_t(sl);
      cd pm[4], p0[4], pp[4], b0[4], bp[4];
      for (int c = 0; c < 4; ++c) {
        pm[c] = o->psi[idx(o, c, t - 1, s)];
        p0[c] = o->psi[idx(o, c, t, s)];
        pp[c] = o->psi[idx(o, c, t + 1, s)];
        b0[c] = o->psib[idx(o, c, t, s)];
        bp[c] = o->psib[idx(o, c, t + 1, s)];
      }
      cd Nm[4], Np[4];
      for (int c = 0; c < 4; ++c) {
        Nm[c] = pm[c] * b0[c];   // N_ij (:24-27)
        Np[c] = p0[c] * bp[c];   // N_ij_conjug (:29-32)
      }
      // spin bilinears F_a of the time-split pair (psibar(t), psi(t-1)) and
      // (psibar(t+1), psi(t)), F_a = tr(Psibar sigma_a Psi), holomorphic form
      auto Fs = [](const cd *b, const cd *q, cd F[3]) {
        F[0] = b[2] * q[0] + b[0] * q[2] + b[3] * q[1] + b[1] * q[3];
        F[1] = cd(0, 1) * (b[2] * q[0] - b[0] * q[2] + b[3] * q[1] - b[1] * q[3]);
        F[2] = b[0] * q[0] + b[1] * q[1] - b[2] * q[2] - b[3] * q[3];
      };
      cd Fm[3], Fp[3];
      Fs(b0, pm, Fm);
      Fs(bp, p0, Fp);
      // dF_a/dpsibar_i on psi(t-1) and dF_a/dpsi_i on psibar(t+1)
      const cd I(0, 1);
      cd Sm[4], Sp[4];
      {
        const cd *q = pm;
        const cd dFx[4] = {q[2], q[3], q[0], q[1]};
        const cd dFy[4] = {-I * q[2], -I * q[3], I * q[0], I * q[1]};
        const cd dFz[4] = {q[0], q[1], -q[2], -q[3]};
        for (int i = 0; i < 4; ++i) Sm[i] = Fm[0] * dFx[i] + Fm[1] * dFy[i] + Fm[2] * dFz[i];
      }
      {
        const cd *b = bp;
        const cd dFx[4] = {b[2], b[3], b[0], b[1]};
        const cd dFy[4] = {I * b[2], I * b[3], -I * b[0], -I * b[1]};
        const cd dFz[4] = {b[0], b[1], -b[2], -b[3]};
        for (int i = 0; i < 4; ++i) Sp[i] = Fp[0] * dFx[i] + Fp[1] * dFy[i] + Fp[2] * dFz[i];
      }
      for (int i = 0; i < 4; ++i) {
        // tempm_i, tempp_i (:35-57)
        cd sm = 0.0, sp = 0.0;
        for (int j = 0; j < 4; ++j)
          if (j != opp[i]) { sm += Nm[j]; sp += Np[j]; }
        const cd tm = sm * pm[i] + (pm[ta[i]] * b0[tb[i]]) * pm[tc[i]];
        const cd tp = sp * bp[i] + (bp[ta[i]] * p0[tb[i]]) * bp[tc[i]];
        const cd km = o->kin[idx(o, i, t - 1, s)];
        const cd kp = o->kinb[idx(o, i, t + 1, s)];
        const double af = p.a_x * p.factor;
        // drift (:64-69 etc.)
        cd term_param = sfac[i] * pm[i];
        cd term_inter = -2.0 * p.G * tm - p.c1 * Sm[i];
        o->drift[idx(o, i, t, s)] = af * (pm[i] - p0[i] + p.eps * (term_param + term_inter + km));
        // conjugate drift (:71-76); psi(t+1) in the reference mode
        const cd xq = (p.conj_drift_mode == CLV_CONJ_DRIFT_ACTION) ? bp[i] : pp[i];
        term_param = sfac[i] * xq;
        term_inter = -2.0 * p.G * tp - p.c1 * Sp[i];
        o->driftb[idx(o, i, t, s)] = af * (bp[i] - b0[i] + p.eps * (term_param + term_inter + kp));
      }
    }
}

// compute_max_and_mean_drift + get_max_and_mean_drift (:231-265, 351-380):
// |D| over component 0 only, mean divided by 4*V_s*Nt (:353, 375).  Fixed order:
// per-tau partials summed in tau order.
void max_mean(const orc_ctx *o, double *mx, double *mean) {
  const int Nt = o->Nt;
  std::vector<double> pmax(Nt, 0.0), psum(Nt, 0.0);
#pragma omp parallel for schedule(static) num_threads(o->nthreads)
  for (int t = 0; t < Nt; ++t) {
    double m = 0.0, s = 0.0;
    for (size_t sl = 0; sl < o->Vs; ++sl) {
      const cd d = o->drift[idx(o, 0, t, sl)];
      const double a = std::hypot(d.real(), d.imag());
      m = std::max(m, a);
      s += a;
    }
    pmax[t] = m;
    psum[t] = s;
  }
  double m = 0.0, s = 0.0;
  for (int t = 0; t < Nt; ++t) { m = std::max(m, pmax[t]); s += psum[t]; }
  *mx = m;
  *mean = s / double(4 * o->S);
}

// multiply_by_timestep + add_noise + operator+= (:432-441, Lattice.cu:32-36):
// psi += (dt D + eta), psibar += (dt Dbar + conj eta), eta ~ N(0, dt) per part.
void update(orc_ctx *o, double dt) {
  const double sq = std::sqrt(dt);
  const bool noise = o->p.noise_on != 0;
  const uint64_t step = uint64_t(o->st.steps);
  const uint64_t seed = o->p.seed;
  const int Nt = o->Nt;
#pragma omp parallel for collapse(2) schedule(static) num_threads(o->nthreads)
  for (int t = 0; t < Nt; ++t)
    for (long long sl = 0; sl < (long long)o->Vs; ++sl) {
      const size_t s = size_t(sl);
      const uint32_t site = uint32_t(size_t(t) * o->Vs + s);
      double nz[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (noise) {
        normals4(seed, 0, step, site, 0, nz);
        normals4(seed, 0, step, site, 1, nz + 4);
      }
      for (int c = 0; c < 4; ++c) {
        const size_t i = idx(o, c, t, s);
        const cd eta(sq * nz[2 * c], sq * nz[2 * c + 1]);
        cd d = o->drift[i] * dt;
        cd db = o->driftb[i] * dt;
        if (noise) { d += eta; db += std::conj(eta); }
        o->psi[i] += d;
        o->psib[i] += db;
      }
    }
}

}  // namespace

extern "C" {

// compute_new_dt (Propagator_without_cutoff.cu:382-404): DT*mean/max (DT when
// max <= 0), clamped to [dt_min, dt_max], then blended (1-alpha) DT + alpha dt.
double orc_new_dt(double mean_drift, double max_drift, const clv_params *p, int32_t *clamp) {
  double dt = (max_drift <= 0.0) ? p->DT : (mean_drift / max_drift) * p->DT;
  int32_t c = 0;
  if (dt < p->dt_min) { dt = p->dt_min; c = -1; }
  if (dt > p->dt_max) { dt = p->dt_max; c = 1; }
  if (clamp) *clamp = c;
  return (1 - p->alpha) * p->DT + p->alpha * dt;
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox10(c, key[0], key[1]);
  std::memcpy(out, c, sizeof(c));
}

void orc_normals(uint64_t seed, uint32_t stream, uint64_t step, uint32_t site,
                 uint32_t pair, double out[4]) {
  normals4(seed, stream, step, site, pair, out);
}

void orc_fft1d(double *data, int32_t n, int32_t sign) {
  FftPlan pl = make_plan(n, sign);
  std::vector<cd> work(n);
  fft_exec(pl, reinterpret_cast<cd *>(data), work.data());
}

void orc_fft3d(double *data, int32_t nx, int32_t ny, int32_t nz, int32_t sign) {
  Plane3 pl{nx, ny, nz, make_plan(nx, sign), make_plan(ny, sign), make_plan(nz, sign)};
  const int m = std::max({nx, ny, nz});
  std::vector<cd> line(m), work(m);
  fft_plane(pl, reinterpret_cast<cd *>(data), line, work);
}

orc_ctx *orc_create(const clv_shape *shape, const clv_params *p, int32_t nthreads) {
  orc_ctx *o = new orc_ctx();
  o->sh = *shape;
  o->p = *p;
  o->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  o->X = shape->nx; o->Y = shape->ny; o->Z = shape->nz; o->Nt = shape->nt;
  o->Vs = size_t(o->X) * o->Y * o->Z;
  o->S = o->Vs * o->Nt;
  const size_t L = 4 * o->S;
  o->psi.assign(L, 0.0); o->psib.assign(L, 0.0);
  o->kin.assign(L, 0.0); o->kinb.assign(L, 0.0);
  o->drift.assign(L, 0.0); o->driftb.assign(L, 0.0);
  o->fwd = Plane3{o->X, o->Y, o->Z, make_plan(o->X, -1), make_plan(o->Y, -1), make_plan(o->Z, -1)};
  o->inv = Plane3{o->X, o->Y, o->Z, make_plan(o->X, +1), make_plan(o->Y, +1), make_plan(o->Z, +1)};
  orc_set_params(o, p);
  std::memset(&o->st, 0, sizeof(o->st));
  o->st.dt = p->DT;
  orc_reset_observables(o);
  return o;
}

void orc_destroy(orc_ctx *o) { delete o; }

void orc_set_params(orc_ctx *o, const clv_params *p) {
  o->p = *p;
  o->p2.assign(o->Vs, 0.0);
  for (int x = 0; x < o->X; ++x)
    for (int y = 0; y < o->Y; ++y)
      for (int z = 0; z < o->Z; ++z) {
        const double px = mom(x, o->X), py = mom(y, o->Y), pz = mom(z, o->Z);
        // p2 = px*px (:196: p_x only) unless the full Laplacian is asked for
        const double p2 = (p->kin_dims == 3) ? px * px + py * py + pz * pz : px * px;
        o->p2[(size_t(x) * o->Y + y) * o->Z + z] = p2;
      }
}

void orc_set_fields(orc_ctx *o, const double *psi, const double *psibar) {
  std::memcpy(static_cast<void *>(o->psi.data()), psi, sizeof(cd) * o->psi.size());
  std::memcpy(static_cast<void *>(o->psib.data()), psibar, sizeof(cd) * o->psib.size());
}

void orc_get_fields(const orc_ctx *o, double *psi, double *psibar) {
  if (psi) std::memcpy(psi, o->psi.data(), sizeof(cd) * o->psi.size());
  if (psibar) std::memcpy(psibar, o->psib.data(), sizeof(cd) * o->psib.size());
}

// psi_c = sqrt(n0) + sqrt(var) (g_re + i g_im) from the init stream (stream 1,
// step 0); psibar = conj(psi).  (New: the reference starts from set_value,
// Lattice.cu:50-81.)
void orc_init_fields(orc_ctx *o, double n0, double var) {
  const double a = std::sqrt(n0), sd = std::sqrt(var);
  for (size_t site = 0; site < o->S; ++site) {
    double nz[8];
    normals4(o->p.seed, 1, 0, uint32_t(site), 0, nz);
    normals4(o->p.seed, 1, 0, uint32_t(site), 1, nz + 4);
    for (int c = 0; c < 4; ++c) {
      const cd v(a + sd * nz[2 * c], sd * nz[2 * c + 1]);
      o->psi[size_t(c) * o->S + site] = v;
      o->psib[size_t(c) * o->S + site] = std::conj(v);
    }
  }
}

void orc_kinetic(orc_ctx *o, double *kin, double *kinbar) {
  compute_kinetic(o);
  if (kin) std::memcpy(kin, o->kin.data(), sizeof(cd) * o->kin.size());
  if (kinbar) std::memcpy(kinbar, o->kinb.data(), sizeof(cd) * o->kinb.size());
}

void orc_drift(orc_ctx *o, double *drift, double *driftbar) {
  compute_kinetic(o);
  compute_drift(o);
  if (drift) std::memcpy(drift, o->drift.data(), sizeof(cd) * o->drift.size());
  if (driftbar) std::memcpy(driftbar, o->driftb.data(), sizeof(cd) * o->driftb.size());
}

void orc_max_mean(orc_ctx *o, double *max_drift, double *mean_drift) {
  compute_kinetic(o);
  compute_drift(o);
  max_mean(o, max_drift, mean_drift);
}

void orc_apply_update(orc_ctx *o, double dt) {
  compute_kinetic(o);
  compute_drift(o);
  update(o, dt);
  o->st.dt = dt;
  o->st.t_langevin += dt;
  o->st.steps += 1;
}

void orc_propagate(orc_ctx *o, int64_t nsteps, int32_t eval_every) {
  for (int64_t n = 0; n < nsteps; ++n) {
    compute_kinetic(o);
    compute_drift(o);
    double mx, mean;
    max_mean(o, &mx, &mean);
    int32_t clamp = 0;
    const double dt = orc_new_dt(mean, mx, &o->p, &clamp);
    if (clamp < 0) o->st.clamp_low++;
    if (clamp > 0) o->st.clamp_high++;
    if (!std::isfinite(mx) || !std::isfinite(mean)) o->st.nonfinite = 1;
    o->st.max_drift = mx;
    o->st.mean_drift = mean;
    const int64_t i = o->st.steps;
    update(o, dt);
    o->st.dt = dt;
    o->st.t_langevin += dt;
    o->st.steps += 1;
    if (eval_every > 0 && i % eval_every == 0) orc_evaluate(o);
  }
}

void orc_get_step_stats(const orc_ctx *o, clv_step_stats *out) { *out = o->st; }

// One evaluation's sums.  n_c in k-space exactly as evaluate_n_* after
// psi->fft() (CUFFT_INVERSE) and psi_conjugate->fft() (CUFFT_FORWARD)
// (Observables_without_cutoff.cu:35-101, 1310-1382), the same n_c in real
// space, F_a(x) (:456-488), and the energy estimator
//   h(x,t) = sum_i [-psibar_i(t+1) K_i(t) - (s_i - mu) psibar_i(t+1) psi_i(t)]
//          + G tr[(Psibar(t+1) Psi(t))^2] + (c1/2) sum_a F_a(t+1,t)^2.
static void eval_sums(orc_ctx *o, double nk[4], double nr[4], double *e,
                      std::vector<double> *F) {
  const int Nt = o->Nt;
  const size_t Vs = o->Vs;
  const double ax = o->p.a_x;
  // k-space copies
  std::vector<cd> ph = o->psi, pbh = o->psib;
#pragma omp parallel num_threads(o->nthreads)
  {
    std::vector<cd> line(std::max({o->X, o->Y, o->Z})), work(line.size());
#pragma omp for schedule(static)
    for (int pl = 0; pl < 4 * Nt; ++pl) {
      fft_plane(o->inv, ph.data() + size_t(pl) * Vs, line, work);
      fft_plane(o->fwd, pbh.data() + size_t(pl) * Vs, line, work);
    }
  }
  kinetic_field(o, o->psi, o->kin, false);
  const double sz[4] = {o->p.P + o->p.Q, 0.0, 0.0, -o->p.P + o->p.Q};
  for (int c = 0; c < 4; ++c) { nk[c] = 0.0; nr[c] = 0.0; }
  double esum = 0.0;
  for (int t = 0; t < Nt; ++t) {
    for (int c = 0; c < 4; ++c) {
      double sk = 0.0, sr = 0.0;
      for (size_t s = 0; s < Vs; ++s) {
        sk += (pbh[idx(o, c, t + 1, s)] * ph[idx(o, c, t, s)]).real() * ax;
        sr += (o->psib[idx(o, c, t + 1, s)] * o->psi[idx(o, c, t, s)]).real() * ax;
      }
      nk[c] += sk;
      nr[c] += sr;
    }
    for (size_t s = 0; s < Vs; ++s) {
      cd b[4], q[4], k[4];
      for (int c = 0; c < 4; ++c) {
        b[c] = o->psib[idx(o, c, t + 1, s)];
        q[c] = o->psi[idx(o, c, t, s)];
        k[c] = o->kin[idx(o, c, t, s)];
      }
      cd h = 0.0;
      cd nn = 0.0;
      for (int c = 0; c < 4; ++c) {
        h += -b[c] * k[c] - sz[c] * (b[c] * q[c]);
        nn += b[c] * q[c];
      }
      const cd detb = b[0] * b[3] - b[1] * b[2];
      const cd detq = q[0] * q[3] - q[1] * q[2];
      h += o->p.G * (nn * nn - 2.0 * detb * detq);
      const cd Fx = b[2] * q[0] + b[0] * q[2] + b[3] * q[1] + b[1] * q[3];
      const cd Fy = cd(0, 1) * (b[2] * q[0] - b[0] * q[2] + b[3] * q[1] - b[1] * q[3]);
      const cd Fz = b[0] * q[0] + b[1] * q[1] - b[2] * q[2] - b[3] * q[3];
      h += 0.5 * o->p.c1 * (Fx * Fx + Fy * Fy + Fz * Fz);
      esum += h.real() * ax;
      if (F) {
        F[0][s] += Fx.real();
        F[1][s] += Fy.real();
        F[2][s] += Fz.real();
      }
    }
  }
  *e = esum;
}

void orc_evaluate(orc_ctx *o) {
  double nk[4], nr[4], e;
  eval_sums(o, nk, nr, &e, o->F);
  for (int c = 0; c < 4; ++c) o->nsum[c] += nk[c];
  o->esum += e;
  o->count += 1;
}

void orc_observable_sums(orc_ctx *o, double out[9]) {
  double nk[4], nr[4], e;
  eval_sums(o, nk, nr, &e, nullptr);
  for (int c = 0; c < 4; ++c) { out[c] = nk[c]; out[4 + c] = nr[c]; }
  out[8] = e;
}

// write_n_* / write_F_* (Observables_without_cutoff.cu:1527-1547, 1614-1634)
void orc_read_observables(const orc_ctx *o, clv_observables *out, double *Fx,
                          double *Fy, double *Fz) {
  std::memset(out, 0, sizeof(*out));
  out->count = o->count;
  if (o->count == 0) return;
  const double Nt = double(o->Nt), cnt = double(o->count), Vs = double(o->Vs);
  double tot = 0.0;
  for (int c = 0; c < 4; ++c) {
    out->n_comp[c] = o->nsum[c] / Nt / cnt / Vs;
    tot += o->nsum[c];
  }
  out->n_tot = tot / Nt / cnt / Vs;
  out->energy = o->esum / Nt / cnt;
  double *dst[3] = {Fx, Fy, Fz};
  double tots[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (size_t s = 0; s < o->Vs; ++s) {
      const double v = o->F[a][s] / Nt / cnt;
      tots[a] += v;
      if (dst[a]) dst[a][s] = v;
    }
  out->Fx_tot = tots[0];
  out->Fy_tot = tots[1];
  out->Fz_tot = tots[2];
  out->density = out->n_tot / Vs;
  out->Sz = out->Fz_tot / Vs;
}

void orc_reset_observables(orc_ctx *o) {
  o->count = 0;
  for (int c = 0; c < 4; ++c) o->nsum[c] = 0.0;
  o->esum = 0.0;
  for (int a = 0; a < 3; ++a) o->F[a].assign(o->Vs, 0.0);
}

}  // extern "C"
