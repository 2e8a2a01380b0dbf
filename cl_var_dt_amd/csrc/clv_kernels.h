// clv_kernels.h -- kernels of one Langevin step on a bucket of equal-shape
// chains, and the host-side launch records.  Included by clv_engine.hip only.
//
// Internal layout of one field (psi or psibar) of one chain: site-major with
// the 4 components innermost.  Site (t, r, x), r = y*Z + z the row index, sits
// at ((t*R + r)*X + x); element (site, c) at site*4 + c.  A row of X sites is
// therefore contiguous (coalesced row loads for the x-transform) and one
// site's 4 components are one 32-byte (FP32) or 64-byte (FP64) load.
#pragma once

#include "clv_device.h"

namespace clv {

template <typename T>
struct BucketArgs {
  C2<T> *buf0_psi, *buf0_psib;  // ping-pong buffers; per-chain parity selects
  C2<T> *buf1_psi, *buf1_psib;
  const ChainParams *prm;
  ChainState *st;               // controller state read by this launch
  ChainState *st_next;          // written by the fused kernel (== st otherwise)
  const C2<T> *tw;              // 2X row-transform twiddles (RowFFT order)
  const double *p2;             // p2[k] = -p_x(k)^2 / X (FP64, from the host)
  C2<T> *kt_psi, *kt_psib;      // transverse kinetic term (kin_dims = 3), or null
  double *part_rd;              // [chain][Nt][ntiles][2] (max, sum) of the current fields
  double *part_wr;              // partials written by this launch
  double *obs_stage;            // [chain][nchunks][Vs][3]
  double *obs_spart;            // [chain][nchunks][ntiles][5]
  double *obs_F;                // [chain][Vs][3]
  double *obs_acc;              // [chain][5]  (n_0..n_3, energy)
  int64_t *obs_count;           // [chain]
  int64_t chain_stride;         // complex elements per chain field (4 S)
  int64_t S, Vs;
  int R, Nt, rows, ntiles, nchunks, tc;
  int tcs;                      // tau slices per drift-stats workgroup
  int fsets;                    // fused kernel: LDS line sets (2: alternate by tau parity)
  int fbatch;                   // fused kernel: tau slices per batch of next-step K_0 lines (1, 2, 4)
  int fblocks;                  // fused kernel: tau blocks (one workgroup each) per (tile, chain)
  int tstride;                  // drift stats: > 0 processes one slice per workgroup, at blockIdx.y * tstride
  int chain0;                  // first chain of the launch
  int eval_every;               // obs kernels: evaluate chains with cur_step % eval_every == 0;
                                // < 0: each chain's own ChainParams::eval_every; 0: always
};

// tau slices per transform batch of k_obs.  One slice of X >= 64 is at most
// 16 lines and leaves transform levels idle; two fill them but hold twice the
// fields in registers.  Measured on MI355X, two are faster for FP32 at
// X >= 96 (6-15%) and FP64 at X = 64, 96, 256 (3-23%); FP32 at X = 64 is 3%
// slower, and FP64 at X = 128, 192 drops to one wave per SIMD (33% slower).
__host__ __device__ constexpr int obs_slices(int complex_bytes, int X) {
  return complex_bytes == 8 ? (X >= 96 ? 2 : 1) : ((X == 64 || X == 96 || X == 256) ? 2 : 1);
}

// Does this chain evaluate its observables after the step just applied?
template <typename T>
__device__ __forceinline__ bool evaluates(const BucketArgs<T> &A, const ChainParams &P,
                                          const ChainState &st) {
  const int ev = A.eval_every < 0 ? P.eval_every : A.eval_every;
  if (A.eval_every < 0 && (ev <= 0 || st.cur_step < P.eval_from)) return false;
  return ev <= 0 || (st.cur_step % ev) == 0;
}

template <typename T>
__device__ __forceinline__ void load4(const C2<T> *f, int64_t site, C2<T> v[4]) {
  if constexpr (sizeof(T) == 4) {
    const float4 *p = reinterpret_cast<const float4 *>(f + site * 4);
    const float4 a = p[0], b = p[1];
    v[0] = {a.x, a.y};
    v[1] = {a.z, a.w};
    v[2] = {b.x, b.y};
    v[3] = {b.z, b.w};
  } else {
    const double2 *p = reinterpret_cast<const double2 *>(f + site * 4);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double2 a = p[c];
      v[c] = {a.x, a.y};
    }
  }
}

template <typename T>
__device__ __forceinline__ C2<T> load1(const C2<T> *f, int64_t site, int c) {
  return f[site * 4 + c];
}

template <typename T>
__device__ __forceinline__ void store4(C2<T> *f, int64_t site, const C2<T> v[4]) {
  if constexpr (sizeof(T) == 4) {
    float4 *p = reinterpret_cast<float4 *>(f + site * 4);
    p[0] = make_float4(v[0].x, v[0].y, v[1].x, v[1].y);
    p[1] = make_float4(v[2].x, v[2].y, v[3].x, v[3].y);
  } else {
    double2 *p = reinterpret_cast<double2 *>(f + site * 4);
#pragma unroll
    for (int c = 0; c < 4; ++c) p[c] = make_double2(v[c].x, v[c].y);
  }
}

// Per-chain coefficients of the drift in the field precision, with
// a_x * factor * eps folded in (the kinetic multiplier table carries the same
// factor, so K' below is a_x factor eps K).
template <typename T>
struct Coef {
  T af;       // a_x factor
  T cs[4];    // a_x factor eps s_i,  s = (mu+P+Q, mu, mu, mu-P+Q)
  T kA, kB;   // a_x factor eps (2G + c1),  a_x factor eps (2G + 2 c1)
  T kK;       // a_x factor eps (weight of the unscaled transverse kinetic term)
  int conj_drift;
};

template <typename T>
__device__ __forceinline__ Coef<T> coef(const ChainParams &p) {
  Coef<T> c;
  const double e = p.a_x * p.factor * p.eps;
  c.af = T(p.a_x * p.factor);
  c.cs[0] = T(e * (p.mu + p.P + p.Q));
  c.cs[1] = T(e * p.mu);
  c.cs[2] = T(e * p.mu);
  c.cs[3] = T(e * (p.mu - p.P + p.Q));
  c.kA = T(e * (2.0 * p.G + p.c1));
  c.kB = T(e * (2.0 * p.G + 2.0 * p.c1));
  c.kK = T(e);
  c.conj_drift = p.conj_drift_mode;
  return c;
}

// Does this chain add the transverse kinetic term (kin_dims = 3 on a lattice
// with Y or Z > 1)?
template <typename T>
__device__ __forceinline__ bool has_kt(const BucketArgs<T> &A, const ChainParams &p) {
  return A.kt_psi != nullptr && p.kin_dims == 3;
}

// k + kK * kt
template <typename T>
__device__ __forceinline__ C2<T> add_kt(C2<T> k, T kK, C2<T> kt) {
  return {k.x + kK * kt.x, k.y + kK * kt.y};
}

// Multiplier-table scale of the drift kernels: a_x factor eps / a_x.
__device__ __forceinline__ double drift_table_scale(const ChainParams &p) { return p.factor * p.eps; }

// sum_j a_j b_j
template <typename T>
__device__ __forceinline__ C2<T> cdot4(const C2<T> a[4], const C2<T> b[4]) {
  return {a[0].x * b[0].x - a[0].y * b[0].y + a[1].x * b[1].x - a[1].y * b[1].y +
              a[2].x * b[2].x - a[2].y * b[2].y + a[3].x * b[3].x - a[3].y * b[3].y,
          a[0].x * b[0].y + a[0].y * b[0].x + a[1].x * b[1].y + a[1].y * b[1].x +
              a[2].x * b[2].y + a[2].y * b[2].x + a[3].x * b[3].y + a[3].y * b[3].x};
}

// det [[a0, a1], [a2, a3]] = a0 a3 - a1 a2
template <typename T>
__device__ __forceinline__ C2<T> cdet(const C2<T> a[4]) {
  return {a[0].x * a[3].x - a[0].y * a[3].y - (a[1].x * a[2].x - a[1].y * a[2].y),
          a[0].x * a[3].y + a[0].y * a[3].x - (a[1].x * a[2].y + a[1].y * a[2].x)};
}

// One component of a drift: af (f1 - f0) + kin + c f1' + B cof, with
// f1' = f1 except in the psibar equation of the reference conj-drift mode.  Written
// as one chain of FMAs (the complex products' negations are free operand
// modifiers), six instructions per real part.
template <typename T>
__device__ __forceinline__ C2<T> drift_term(T af, C2<T> f1, C2<T> f0, C2<T> kin, C2<T> c, C2<T> f1c,
                                            C2<T> B, C2<T> cof) {
  T x = fmad(af, f1.x - f0.x, kin.x);
  x = fmad(c.x, f1c.x, x);
  x = fmad(-c.y, f1c.y, x);
  x = fmad(B.x, cof.x, x);
  x = fmad(-B.y, cof.y, x);
  T y = fmad(af, f1.y - f0.y, kin.y);
  y = fmad(c.x, f1c.y, y);
  y = fmad(c.y, f1c.x, y);
  y = fmad(B.x, cof.y, y);
  y = fmad(B.y, cof.x, y);
  return {x, y};
}

// psi + (dt D + eta) and psibar + (dt Db + conj eta) at one site (add_noise,
// multiply_by_timestep and operator+=, Propagator_without_cutoff.cu:131-157,
// 434-441); noise nz = (re, im) of eta per component.
template <typename T>
__device__ __forceinline__ void update_site(T dt, const C2<T> p0[4], const C2<T> b0[4], const C2<T> D[4],
                                            const C2<T> Db[4], const T nz[8], C2<T> o[4], C2<T> ob[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const T ex = nz[2 * c], ey = nz[2 * c + 1];
    o[c] = {p0[c].x + fmad(dt, D[c].x, ex), p0[c].y + fmad(dt, D[c].y, ey)};
    ob[c] = {b0[c].x + fmad(dt, Db[c].x, ex), b0[c].y + fmad(dt, Db[c].y, -ey)};
  }
}

// Drift of psi and psibar at one site, all four components
// (calculate_drift_without_noise, Propagator_without_cutoff.cu:1-129, with
// factor = f) plus the spin term -c1 S_i and the conj-drift switch:
//   D_i  = a_x f [psi_i(t-1) - psi_i(t) + eps (s_i psi_i(t-1) - 2G T-_i - c1 S-_i + K_i(t-1))]
//   Db_i = a_x f [psib_i(t+1) - psib_i(t) + eps (s_i x_i - 2G T+_i - c1 S+_i + Kb_i(t+1))]
// with x = psi(t+1) (the reference's :71,88,105,122) or psibar(t+1).  The interaction is
// evaluated through the invariants of the 2x2 matrices Psi = [[psi_0, psi_1],
// [psi_2, psi_3]] and Psibar on the time-split pair (Psi(t-1), Psibar(t)):
// with n_j = psi_j psib_j, N = sum_j n_j, det = det Psi and the cofactors
// cof = (psib_3, -psib_2, -psib_1, psib_0) = d(det Psibar)/d psib_i,
//   T_i = (sum_{j != opp(i)} n_j) psi_i + psi_a psib_b psi_c = N psi_i - det cof_i
//   S_i = F_z (sigma_z psi)_i + F_-+ (sigma_+- psi)_i    = N psi_i - 2 det cof_i
// (the second from the Fierz identity sum_a F_a^2 = N^2 - 4 det Psi det Psibar),
// so -2G T_i - c1 S_i = -(2G + c1) N psi_i + (2G + 2 c1) det cof_i.  The
// psibar equation is the same with the roles of Psi(t) and Psibar(t+1) swapped.
//   pm = psi(t-1), p0 = psi(t), pp = psi(t+1), b0 = psibar(t), bp = psibar(t+1)
//   km = K'(t-1), kp = Kb'(t+1) (already scaled by a_x f eps)
// kA N of a time-split pair (a, b): N = sum_j a_j b_j.
template <typename T>
__device__ __forceinline__ C2<T> drift_an(const Coef<T> &k, const C2<T> a[4], const C2<T> b[4]) {
  const C2<T> N = cdot4(a, b);
  return {k.kA * N.x, k.kA * N.y};
}

// The psi drift D (all components) from psi(t-1) = pm, psi(t) = p0, its
// kinetic term km = K'(t-1), psibar(t) = b0 and anm = kA N of the pair (pm, b0).
template <typename T>
__device__ __forceinline__ void drift_psi(const Coef<T> &k, const C2<T> pm[4], const C2<T> p0[4],
                                          const C2<T> km[4], const C2<T> b0[4], C2<T> anm,
                                          C2<T> D[4]) {
  const C2<T> det = cdet(pm);
  const C2<T> B = {k.kB * det.x, k.kB * det.y};
  const T ay = -anm.y, ax = anm.x;
  const C2<T> cof[4] = {b0[3], {-b0[2].x, -b0[2].y}, {-b0[1].x, -b0[1].y}, b0[0]};
#pragma unroll
  for (int i = 0; i < 4; ++i)
    D[i] = drift_term(k.af, pm[i], p0[i], km[i], C2<T>{k.cs[i] - ax, ay}, pm[i], B, cof[i]);
}

// The psibar drift Db (all components) from psibar(t+1) = bp, psibar(t) = b0,
// its kinetic term kp = Kb'(t+1), psi(t) = p0, psi(t+1) = pp (the reference
// conj-drift mode) and anp = kA N of the pair (p0, bp).
template <typename T>
__device__ __forceinline__ void drift_psib(const Coef<T> &k, const C2<T> p0[4], const C2<T> pp[4],
                                           const C2<T> b0[4], const C2<T> bp[4],
                                           const C2<T> kp[4], C2<T> anp, C2<T> Db[4]) {
  const C2<T> det = cdet(bp);
  const C2<T> B = {k.kB * det.x, k.kB * det.y};
  const T ax = anp.x, ay = anp.y;
  const C2<T> cof[4] = {p0[3], {-p0[2].x, -p0[2].y}, {-p0[1].x, -p0[1].y}, p0[0]};
  if (k.conj_drift == 1) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      Db[i] = drift_term(k.af, bp[i], b0[i], kp[i], C2<T>{k.cs[i] - ax, -ay}, bp[i], B, cof[i]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      // s_i psi(t+1) - A psibar(t+1)
      const C2<T> kin = {kp[i].x + k.cs[i] * pp[i].x, kp[i].y + k.cs[i] * pp[i].y};
      Db[i] = drift_term(k.af, bp[i], b0[i], kin, C2<T>{-ax, -ay}, bp[i], B, cof[i]);
    }
  }
}

template <typename T>
__device__ __forceinline__ void drift_site(const Coef<T> &k, const C2<T> pm[4], const C2<T> p0[4],
                                           const C2<T> pp[4], const C2<T> b0[4], const C2<T> bp[4],
                                           const C2<T> km[4], const C2<T> kp[4], C2<T> D[4],
                                           C2<T> Db[4]) {
  drift_psi(k, pm, p0, km, b0, drift_an(k, pm, b0), D);
  drift_psib(k, p0, pp, b0, bp, kp, drift_an(k, p0, bp), Db);
}

// Component 0 of the psi drift only (what the max/mean reduction reads),
// in two parts: everything but the kinetic term, and then + K'_0(t-1) as the
// last operation, so the fused march can form the first part at one tau and
// add the kinetic term after a later, batched transform with the same bits.
template <typename T>
__device__ __forceinline__ C2<T> drift0_local(const Coef<T> &k, const C2<T> pm[4], C2<T> p00,
                                              const C2<T> b0[4]) {
  const C2<T> N = cdot4(pm, b0), det = cdet(pm);
  const C2<T> B = {k.kB * det.x, k.kB * det.y};
  const C2<T> c = {k.cs[0] - k.kA * N.x, -(k.kA * N.y)};
  const C2<T> f1 = pm[0], cof = b0[3];
  T x = k.af * (f1.x - p00.x);
  x = fmad(c.x, f1.x, x);
  x = fmad(-c.y, f1.y, x);
  x = fmad(B.x, cof.x, x);
  x = fmad(-B.y, cof.y, x);
  T y = k.af * (f1.y - p00.y);
  y = fmad(c.x, f1.y, y);
  y = fmad(c.y, f1.x, y);
  y = fmad(B.x, cof.y, y);
  y = fmad(B.y, cof.x, y);
  return {x, y};
}

template <typename T>
__device__ __forceinline__ C2<T> drift0_add_kin(C2<T> local, C2<T> km0) {
  return {local.x + km0.x, local.y + km0.y};
}

template <typename T>
__device__ __forceinline__ C2<T> drift0_site(const Coef<T> &k, const C2<T> pm[4], C2<T> p00,
                                             const C2<T> b0[4], C2<T> km0) {
  return drift0_add_kin(drift0_local(k, pm, p00, b0), km0);
}

// |D_0| in FP64 (the reduction's input).  With FP32 fields the square sum
// is exact in FP64 and the hardware square root (v_sqrt_f64, not correctly
// rounded but far inside FP32's precision) replaces the correctly rounded
// sequence; FP64 fields keep the latter.
template <typename T>
__device__ __forceinline__ double drift_abs(C2<T> d) {
  const double dx = double(d.x), dy = double(d.y);
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_sqrt(dx * dx + dy * dy);
  else return sqrt(dx * dx + dy * dy);
}

// Shared-memory carve-up of the row-transform kernels: twiddles, the
// kinetic multiplier table, `nlines` padded lines, reduction scratch.
template <typename T, int X, int N1 = Split<X>::N1>
struct Smem {
  using F = RowFFT<T, X, N1>;
  C2<T> *tw;
  T *mult;
  C2<T> *buf;
  double *red;
  __device__ Smem(char *base, int nlines) {
    tw = reinterpret_cast<C2<T> *>(base);
    mult = reinterpret_cast<T *>(tw + tw_entries(X, N1));
    buf = reinterpret_cast<C2<T> *>(mult + X + (X & 1));
    red = reinterpret_cast<double *>(buf + nlines * F::PAD);
  }
  // cell of site x in line l
  __device__ __forceinline__ C2<T> &at(int l, int x) { return buf[l * F::PAD + F::pos(x)]; }
  // first byte after the reduction scratch (extra per-kernel LDS)
  __device__ __forceinline__ C2<T> *red_end() { return reinterpret_cast<C2<T> *>(red + 64); }
  static size_t bytes(int nlines) {
    return sizeof(C2<T>) * tw_entries(X, N1) + sizeof(T) * (X + (X & 1)) + sizeof(C2<T>) * nlines * F::PAD +
           sizeof(double) * 64;
  }
};

// Twiddles and the kinetic multiplier m(k) = scale * (-p_x(k)^2 / X) into LDS:
// scale = 1/a_x gives K (compute_kinetic_part + normalize, Propagator_without_cutoff.cu:160-203,
// 423-424); the drift kernels use drift_table_scale.
template <typename T, int X, int N1>
__device__ __forceinline__ void load_tables(Smem<T, X, N1> &sm, const C2<T> *tw, const double *p2,
                                            double scale, int tid, int nthr) {
  for (int e = tid; e < tw_entries(X, N1); e += nthr) sm.tw[e] = tw[e];
  for (int e = tid; e < X; e += nthr) sm.mult[e] = T(p2[e] * scale);
}

// Fixed-order (max, sum) of n per-tile partials, by one full wave: lane l
// folds tiles l, l+64, ... in order, then a fixed xor tree.
__device__ __forceinline__ void reduce_partials(const double *pp, int n, int lane, double *m,
                                                double *s) {
  double mm = 0.0, ss = 0.0;
  const double2 *p2 = reinterpret_cast<const double2 *>(pp);
  int i = lane;
  // eight loads in flight, then the same sequential fold per lane
  for (; i + 7 * 64 < n; i += 8 * 64) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = p2[i + u * 64];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      mm = fmax(mm, v[u].x);
      ss += v[u].y;
    }
  }
  for (; i < n; i += 64) {
    mm = fmax(mm, p2[i].x);
    ss += p2[i].y;
  }
  *m = wave_max(mm);
  *s = wave_sum(ss);
}

// (max, sum) of one tile's nthr values vb[0..nthr) in a fixed order, by one
// wave: lane l folds vb[l], vb[l+64], ... in order, then the wave trees.
// Valid in lane 0.  The drift-statistics partials of both schedules come from
// here, so only one wave per tile and tau runs the cross-lane trees.
__device__ __forceinline__ void tile_max_sum(const double *vb, int nthr, int lane, double *m,
                                             double *s) {
  double mm = vb[lane], ss = vb[lane];
  for (int i = lane + 64; i < nthr; i += 64) {
    const double v = vb[i];
    mm = fmax(mm, v);
    ss += v;
  }
  *m = wave_max(mm);
  *s = wave_sum(ss);
}

}  // namespace clv
