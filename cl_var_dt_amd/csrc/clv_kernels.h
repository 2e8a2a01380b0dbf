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
  const C2<T> *tw;              // tw[m] = exp(-2 pi i m / X)
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
  int chain0;                   // first chain of the launch
  int eval_every;               // obs kernels: evaluate chains with cur_step % eval_every == 0;
                                // < 0: each chain's own ChainParams::eval_every; 0: always
};

// Does this chain evaluate its observables after the step just applied?
template <typename T>
__device__ __forceinline__ bool evaluates(const BucketArgs<T> &A, const ChainParams &P,
                                          const ChainState &st) {
  const int ev = A.eval_every < 0 ? P.eval_every : A.eval_every;
  if (A.eval_every < 0 && (ev <= 0 || st.cur_step < P.eval_from)) return false;
  return ev <= 0 || (st.cur_step % ev) == 0;
}

// p of compute_kinetic_part (Propagator_without_cutoff.cu:169-176).
__device__ __forceinline__ double mom(int j, int n) {
  if (j < n / 2 || n == 1) return 6.283185307179586 * double(j) / double(n);
  return -6.283185307179586 * double(n - j) / double(n);
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

// Per-chain coefficients in the field precision.
template <typename T>
struct Coef {
  T s[4];   // mu+P+Q, mu, mu, mu-P+Q
  T G2, c1, eps, af;
  int q3;
};

template <typename T>
__device__ __forceinline__ Coef<T> coef(const ChainParams &p) {
  Coef<T> c;
  c.s[0] = T(p.mu + p.P + p.Q);
  c.s[1] = T(p.mu);
  c.s[2] = T(p.mu);
  c.s[3] = T(p.mu - p.P + p.Q);
  c.G2 = T(2.0 * p.G);
  c.c1 = T(p.c1);
  c.eps = T(p.eps);
  c.af = T(p.a_x * p.factor);
  c.q3 = p.q3_mode;
  return c;
}

// Three-term sums of N_j without the opposite component (tempm_i / tempp_i,
// Propagator_without_cutoff.cu:35-57), in the reference's order.
template <typename T>
__device__ __forceinline__ C2<T> nsum3(int i, const C2<T> N[4]) {
  switch (i) {
    case 0: return cadd(cadd(N[0], N[1]), N[2]);
    case 1: return cadd(cadd(N[0], N[1]), N[3]);
    case 2: return cadd(cadd(N[0], N[2]), N[3]);
    default: return cadd(cadd(N[1], N[2]), N[3]);
  }
}

// Drift of psi and psibar at one site, all four components
// (calculate_drift_without_noise, Propagator_without_cutoff.cu:1-129) plus
// the spin term -c1 S_i and the Q3 switch.
//   pm = psi(t-1), p0 = psi(t), pp = psi(t+1), b0 = psibar(t), bp = psibar(t+1)
//   km = K(t-1), kp = Kbar(t+1)
template <typename T>
__device__ __forceinline__ void drift_site(const Coef<T> &k, const C2<T> pm[4], const C2<T> p0[4],
                                           const C2<T> pp[4], const C2<T> b0[4], const C2<T> bp[4],
                                           const C2<T> km[4], const C2<T> kp[4], C2<T> D[4],
                                           C2<T> Db[4]) {
  C2<T> Nm[4], Np[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    Nm[c] = cmul(pm[c], b0[c]);
    Np[c] = cmul(p0[c], bp[c]);
  }
  // triples (a,b,c)_i: 0:(1,3,2) 1:(0,2,3) 2:(3,1,0) 3:(2,0,1)
  constexpr int ta[4] = {1, 0, 3, 2}, tb[4] = {3, 2, 1, 0}, tcx[4] = {2, 3, 0, 1};
  C2<T> Sm[4] = {}, Sp[4] = {};
  if (k.c1 != T(0)) {
    // F+- = Fx +- i Fy, Fz of (psibar(t), psi(t-1)) and (psibar(t+1), psi(t))
    const C2<T> Fpm = cscale(T(2), cadd(cmul(b0[0], pm[2]), cmul(b0[1], pm[3])));
    const C2<T> Fmm = cscale(T(2), cadd(cmul(b0[2], pm[0]), cmul(b0[3], pm[1])));
    const C2<T> Fzm = csub(cadd(Nm[0], Nm[1]), cadd(Nm[2], Nm[3]));
    Sm[0] = cadd(cmul(Fmm, pm[2]), cmul(Fzm, pm[0]));
    Sm[1] = cadd(cmul(Fmm, pm[3]), cmul(Fzm, pm[1]));
    Sm[2] = csub(cmul(Fpm, pm[0]), cmul(Fzm, pm[2]));
    Sm[3] = csub(cmul(Fpm, pm[1]), cmul(Fzm, pm[3]));
    const C2<T> Fpp = cscale(T(2), cadd(cmul(bp[0], p0[2]), cmul(bp[1], p0[3])));
    const C2<T> Fmp = cscale(T(2), cadd(cmul(bp[2], p0[0]), cmul(bp[3], p0[1])));
    const C2<T> Fzp = csub(cadd(Np[0], Np[1]), cadd(Np[2], Np[3]));
    Sp[0] = cadd(cmul(Fpp, bp[2]), cmul(Fzp, bp[0]));
    Sp[1] = cadd(cmul(Fpp, bp[3]), cmul(Fzp, bp[1]));
    Sp[2] = csub(cmul(Fmp, bp[0]), cmul(Fzp, bp[2]));
    Sp[3] = csub(cmul(Fmp, bp[1]), cmul(Fzp, bp[3]));
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const C2<T> tm = cadd(cmul(nsum3(i, Nm), pm[i]), cmul(cmul(pm[ta[i]], b0[tb[i]]), pm[tcx[i]]));
    const C2<T> tp = cadd(cmul(nsum3(i, Np), bp[i]), cmul(cmul(bp[ta[i]], p0[tb[i]]), bp[tcx[i]]));
    // psi drift
    C2<T> inter = csub(cscale(-k.G2, tm), cscale(k.c1, Sm[i]));
    C2<T> inner = cadd(cadd(cscale(k.s[i], pm[i]), inter), km[i]);
    D[i] = cscale(k.af, cadd(csub(pm[i], p0[i]), cscale(k.eps, inner)));
    // psibar drift; Q3: psi(t+1) multiplies s_i in the reference mode
    const C2<T> xq = (k.q3 == 1) ? bp[i] : pp[i];
    inter = csub(cscale(-k.G2, tp), cscale(k.c1, Sp[i]));
    inner = cadd(cadd(cscale(k.s[i], xq), inter), kp[i]);
    Db[i] = cscale(k.af, cadd(csub(bp[i], b0[i]), cscale(k.eps, inner)));
  }
}

// Component 0 of the psi drift only (what the max/mean reduction reads, Q2).
template <typename T>
__device__ __forceinline__ C2<T> drift0_site(const Coef<T> &k, const C2<T> pm[4], C2<T> p00,
                                             const C2<T> b0[4], C2<T> km0) {
  C2<T> Nm[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) Nm[c] = cmul(pm[c], b0[c]);
  C2<T> S0 = {T(0), T(0)};
  if (k.c1 != T(0)) {
    const C2<T> Fmm = cscale(T(2), cadd(cmul(b0[2], pm[0]), cmul(b0[3], pm[1])));
    const C2<T> Fzm = csub(cadd(Nm[0], Nm[1]), cadd(Nm[2], Nm[3]));
    S0 = cadd(cmul(Fmm, pm[2]), cmul(Fzm, pm[0]));
  }
  const C2<T> tm = cadd(cmul(nsum3(0, Nm), pm[0]), cmul(cmul(pm[1], b0[3]), pm[2]));
  const C2<T> inter = csub(cscale(-k.G2, tm), cscale(k.c1, S0));
  const C2<T> inner = cadd(cadd(cscale(k.s[0], pm[0]), inter), km0);
  return cscale(k.af, cadd(csub(pm[0], p00), cscale(k.eps, inner)));
}

// Shared-memory carve-up of the row-transform kernels: twiddles, the
// kinetic multiplier table, `nlines` padded lines, reduction scratch.
template <typename T, int X>
struct Smem {
  using F = RowFFT<T, X>;
  C2<T> *tw;
  T *mult;
  C2<T> *buf;
  double *red;
  __device__ Smem(char *base, int nlines) {
    tw = reinterpret_cast<C2<T> *>(base);
    mult = reinterpret_cast<T *>(tw + X);
    buf = reinterpret_cast<C2<T> *>(mult + X + (X & 1));
    red = reinterpret_cast<double *>(buf + nlines * F::PAD);
  }
  // cell of site x in line l
  __device__ __forceinline__ C2<T> &at(int l, int x) { return buf[l * F::PAD + F::pos(x)]; }
  // first byte after the reduction scratch (extra per-kernel LDS)
  __device__ __forceinline__ C2<T> *red_end() { return reinterpret_cast<C2<T> *>(red + 64); }
  static size_t bytes(int nlines) {
    return sizeof(C2<T>) * X + sizeof(T) * (X + (X & 1)) + sizeof(C2<T>) * nlines * F::PAD +
           sizeof(double) * 64;
  }
};

template <typename T, int X>
__device__ __forceinline__ void load_tables(Smem<T, X> &sm, const C2<T> *tw, double a_x, int tid,
                                            int nthr) {
  const double norm = 1.0 / (a_x * double(X));
  for (int e = tid; e < X; e += nthr) {
    sm.tw[e] = tw[e];
    const double p = mom(e, X);
    sm.mult[e] = T(-(p * p) * norm);
  }
}

// Fixed-order (max, sum) of n per-tile partials, by one full wave: lane l
// folds tiles l, l+64, ... in order, then a fixed xor tree.
__device__ __forceinline__ void reduce_partials(const double *pp, int n, int lane, double *m,
                                                double *s) {
  double mm = 0.0, ss = 0.0;
  for (int i = lane; i < n; i += 64) {
    mm = fmax(mm, pp[2 * i]);
    ss += pp[2 * i + 1];
  }
  *m = wave_max(mm);
  *s = wave_sum(ss);
}

// Workgroup (max, sum) of v in a fixed order; valid in thread 0.  Uses red[0..63].
__device__ __forceinline__ void block_max_sum(double v, double *red, int tid, int nthr, double *m,
                                              double *s) {
  double mm = wave_max(v), ss = wave_sum(v);
  const int w = tid >> 6, nw = (nthr + 63) >> 6;
  if ((tid & 63) == 0) {
    red[w] = mm;
    red[32 + w] = ss;
  }
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < nw; ++q) {
      mm = fmax(mm, red[q]);
      ss += red[32 + q];
    }
  }
  *m = mm;
  *s = ss;
}

}  // namespace clv
