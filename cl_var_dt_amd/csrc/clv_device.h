// clv_device.h -- device building blocks of the gfx950 complex-Langevin engine:
// complex arithmetic, Philox4x32-10 + Box-Muller noise, the adaptive-dt
// controller, wave/workgroup reductions, and the LDS Stockham FFT that
// produces the kinetic term of a row of X sites.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clv {

// ---------------------------------------------------------------- complex --
template <typename T>
struct C2 {
  T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(T s, C2<T> a) { return {s * a.x, s * a.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) { return {a.x, -a.y}; }
// a * (s i), s = +-1
template <int S, typename T>
__device__ __forceinline__ C2<T> cmul_si(C2<T> a) { return {T(-S) * a.y, T(S) * a.x}; }

// ------------------------------------------------------- per-chain records --
// Per-chain parameters as the kernels read them (clv_params, converted).
struct ChainParams {
  double mu, P, G, c1, Q, eps, a_x, DT, dt_min, dt_max, alpha, factor;
  uint64_t seed;
  int32_t q3_mode, noise_on, active, pad;
};

// Per-chain controller state, device resident.
struct ChainState {
  double dt;          // dt of the step being / last applied
  double max_drift;   // max |D_0| that produced dt
  double mean_drift;  // sum |D_0| / (4 V_s Nt)
  double t_langevin;  // sum of dt
  int64_t cur_step;   // index of the step being / last applied (noise counter)
  int64_t steps;      // number of steps taken
  int64_t clamp_low, clamp_high;
  int32_t nonfinite;
  int32_t parity;     // buffer holding the current fields
  int32_t in_parity;  // buffer the step being applied reads
  int32_t pad[3];
};

// ----------------------------------------------------------------- Philox --
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Each round uses two
// 32x32->64 products (v_mad_u64_u32).
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c.y ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c.w ^ k1;
    c = make_uint4(n0, uint32_t(p1), n2, uint32_t(p0));
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Two complex N(0,1) draws for (seed, stream, step, site, pair):
// counter (site, pair | stream << 8, step_lo, step_hi), uniforms
// (x + 1/2) 2^-32, Box-Muller (1958) on (u0,u1) and (u2,u3).
__device__ __forceinline__ void normals4(double out[4], uint64_t seed, uint32_t stream,
                                         uint64_t step, uint32_t site, uint32_t pair) {
  const uint4 r = philox10(make_uint4(site, pair | (stream << 8), uint32_t(step), uint32_t(step >> 32)),
                           uint32_t(seed), uint32_t(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = (double(r.x) + 0.5) * s, u1 = (double(r.y) + 0.5) * s;
  const double u2 = (double(r.z) + 0.5) * s, u3 = (double(r.w) + 0.5) * s;
  const double r0 = sqrt(-2.0 * log(u0));
  const double r1 = sqrt(-2.0 * log(u2));
  double s0, c0, s1, c1;
  sincos(6.283185307179586 * u1, &s0, &c0);
  sincos(6.283185307179586 * u3, &s1, &c1);
  out[0] = r0 * c0;
  out[1] = r0 * s0;
  out[2] = r1 * c1;
  out[3] = r1 * s1;
}

__device__ __forceinline__ void normals4(float out[4], uint64_t seed, uint32_t stream,
                                         uint64_t step, uint32_t site, uint32_t pair) {
  const uint4 r = philox10(make_uint4(site, pair | (stream << 8), uint32_t(step), uint32_t(step >> 32)),
                           uint32_t(seed), uint32_t(seed >> 32));
  const float s = 2.3283064365386963e-10f;
  const float u0 = float(r.x) * s + 1.16415321826934814e-10f;
  const float u1 = float(r.y) * s + 1.16415321826934814e-10f;
  const float u2 = float(r.z) * s + 1.16415321826934814e-10f;
  const float u3 = float(r.w) * s + 1.16415321826934814e-10f;
  const float r0 = sqrtf(-2.0f * logf(u0));
  const float r1 = sqrtf(-2.0f * logf(u2));
  float s0, c0, s1, c1;
  sincosf(6.28318530717958648f * u1, &s0, &c0);
  sincosf(6.28318530717958648f * u3, &s1, &c1);
  out[0] = r0 * c0;
  out[1] = r0 * s0;
  out[2] = r1 * c1;
  out[3] = r1 * s1;
}

// ------------------------------------------------------------- controller --
// Propagator::compute_new_dt (Propagator_without_cutoff.cu:382-404), with
// explicit IEEE operations so the device and host results are identical.
__device__ __forceinline__ double new_dt(double mean, double mx, const ChainParams &p, int *clamp) {
  double dt = (mx <= 0.0) ? p.DT : __dmul_rn(__ddiv_rn(mean, mx), p.DT);
  int c = 0;
  if (dt < p.dt_min) { dt = p.dt_min; c = -1; }
  if (dt > p.dt_max) { dt = p.dt_max; c = 1; }
  *clamp = c;
  return __dadd_rn(__dmul_rn(__dsub_rn(1.0, p.alpha), p.DT), __dmul_rn(p.alpha, dt));
}

// ------------------------------------------------------------- reductions --
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// --------------------------------------------------------------- FFT plan --
// Radix schedule of the row transform of length X (Stockham stages).
template <int X> struct Plan;
template <> struct Plan<4>   { static constexpr int NS = 1, R0 = 4, R1 = 1, R2 = 1; };
template <> struct Plan<8>   { static constexpr int NS = 1, R0 = 8, R1 = 1, R2 = 1; };
template <> struct Plan<16>  { static constexpr int NS = 2, R0 = 4, R1 = 4, R2 = 1; };
template <> struct Plan<24>  { static constexpr int NS = 2, R0 = 8, R1 = 3, R2 = 1; };
template <> struct Plan<32>  { static constexpr int NS = 2, R0 = 8, R1 = 4, R2 = 1; };
template <> struct Plan<48>  { static constexpr int NS = 3, R0 = 4, R1 = 4, R2 = 3; };
template <> struct Plan<64>  { static constexpr int NS = 2, R0 = 8, R1 = 8, R2 = 1; };
template <> struct Plan<96>  { static constexpr int NS = 3, R0 = 8, R1 = 4, R2 = 3; };
template <> struct Plan<128> { static constexpr int NS = 3, R0 = 8, R1 = 4, R2 = 4; };
template <> struct Plan<192> { static constexpr int NS = 3, R0 = 8, R1 = 8, R2 = 3; };
template <> struct Plan<256> { static constexpr int NS = 3, R0 = 8, R1 = 8, R2 = 4; };

// Small DFTs, y_q = sum_r x_r exp(S 2 pi i r q / R).
template <int S, typename T>
__device__ __forceinline__ void dft2(C2<T> *x) {
  const C2<T> a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}
template <int S, typename T>
__device__ __forceinline__ void dft3(C2<T> *x) {
  const T h = T(0.86602540378443864676);  // sqrt(3)/2
  const C2<T> t = cadd(x[1], x[2]);
  const C2<T> d = csub(x[1], x[2]);
  const C2<T> u = cmul_si<S>(cscale(h, d));
  const C2<T> m = {x[0].x - T(0.5) * t.x, x[0].y - T(0.5) * t.y};
  x[0] = cadd(x[0], t);
  x[1] = cadd(m, u);
  x[2] = csub(m, u);
}
template <int S, typename T>
__device__ __forceinline__ void dft4(C2<T> *x) {
  const C2<T> t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
  const C2<T> t2 = cadd(x[1], x[3]), t3 = cmul_si<S>(csub(x[1], x[3]));
  x[0] = cadd(t0, t2);
  x[2] = csub(t0, t2);
  x[1] = cadd(t1, t3);
  x[3] = csub(t1, t3);
}
template <int S, typename T>
__device__ __forceinline__ void dft8(C2<T> *x) {
  C2<T> e[4] = {x[0], x[2], x[4], x[6]};
  C2<T> o[4] = {x[1], x[3], x[5], x[7]};
  dft4<S>(e);
  dft4<S>(o);
  const T r = T(0.70710678118654752440);
  // w8^k o_k, w8 = (1 + S i)/sqrt(2)
  const C2<T> o1 = {r * (o[1].x - T(S) * o[1].y), r * (o[1].y + T(S) * o[1].x)};
  const C2<T> o2 = cmul_si<S>(o[2]);
  const C2<T> o3 = {r * (-o[3].x - T(S) * o[3].y), r * (-o[3].y + T(S) * o[3].x)};
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}
template <int R, int S, typename T>
__device__ __forceinline__ void dft(C2<T> *x) {
  if constexpr (R == 2) dft2<S>(x);
  else if constexpr (R == 3) dft3<S>(x);
  else if constexpr (R == 4) dft4<S>(x);
  else if constexpr (R == 8) dft8<S>(x);
}

// One Stockham stage over NL lines of length X held in LDS (line l at
// src + l*X): for i < X/R, k = i mod P, x_r = src[i + r X/R] w^{S r k X/(P R)},
// DFT_R, dst[(i-k) R + k + r P] = y_r.  tw[m] = exp(-2 pi i m / X).
template <typename T, int X, int R, int P, int S>
__device__ __forceinline__ void fft_stage(const C2<T> *__restrict__ src, C2<T> *__restrict__ dst,
                                          const C2<T> *__restrict__ tw, int nlines, int tid,
                                          int nthr) {
  constexpr int NB = X / R;
  const int total = nlines * NB;
  for (int b = tid; b < total; b += nthr) {
    const int line = b / NB;
    const int i = b - line * NB;
    const int k = i % P;
    const C2<T> *s = src + line * X;
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      C2<T> a = s[i + r * NB];
      if (P > 1 && r > 0) {
        const C2<T> w = tw[(r * k * (X / (P * R))) % X];
        a = cmul(a, S < 0 ? w : cconj(w));
      }
      v[r] = a;
    }
    dft<R, S>(v);
    C2<T> *d = dst + line * X + (i - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * P] = v[r];
  }
}

template <typename T, int X, int S>
__device__ __forceinline__ C2<T> *fft_lines(C2<T> *a, C2<T> *b, const C2<T> *tw, int nlines,
                                            int tid, int nthr) {
  using PL = Plan<X>;
  fft_stage<T, X, PL::R0, 1, S>(a, b, tw, nlines, tid, nthr);
  __syncthreads();
  if constexpr (PL::NS == 1) {
    return b;
  } else {
    fft_stage<T, X, PL::R1, PL::R0, S>(b, a, tw, nlines, tid, nthr);
    __syncthreads();
    if constexpr (PL::NS == 2) {
      return a;
    } else {
      fft_stage<T, X, PL::R2, PL::R0 * PL::R1, S>(a, b, tw, nlines, tid, nthr);
      __syncthreads();
      return b;
    }
  }
}

// Kinetic term of NL lines held in LDS buffer a (b is scratch):
// K = F^{+}[m(k) F^{-}[line]] with m(k) = -p_x(k)^2 / (a_x X).  Returns the
// buffer holding K (natural order).  Caller syncs before the call.
template <typename T, int X>
__device__ __forceinline__ C2<T> *kinetic_lines(C2<T> *a, C2<T> *b, const C2<T> *tw,
                                                const T *mult, int nlines, int tid, int nthr) {
  C2<T> *f = fft_lines<T, X, -1>(a, b, tw, nlines, tid, nthr);
  for (int e = tid; e < nlines * X; e += nthr) {
    const T m = mult[e % X];
    f[e] = cscale(m, f[e]);
  }
  __syncthreads();
  C2<T> *g = (f == a) ? b : a;
  return fft_lines<T, X, +1>(f, g, tw, nlines, tid, nthr);
}

}  // namespace clv
