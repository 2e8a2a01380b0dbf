// clv_device.h -- device building blocks of the gfx950 complex-Langevin engine:
// complex arithmetic, Philox4x32-10 + Box-Muller noise, the adaptive-dt
// controller, wave/workgroup reductions, and the two-level LDS/register FFT
// that produces the kinetic term of a row of X sites.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clv_math64.h"

namespace clv {

// ---------------------------------------------------------------- complex --
template <typename T>
struct alignas(2 * sizeof(T)) C2 {
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

// a * b + c with one rounding, in the field precision
template <typename T>
__device__ __forceinline__ T fmad(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// ------------------------------------------------------- per-chain records --
// Per-chain parameters as the kernels read them (clv_params, converted).
struct ChainParams {
  double mu, P, G, c1, Q, eps, a_x, DT, dt_min, dt_max, alpha, factor;
  uint64_t seed;
  int32_t conj_drift_mode, noise_on, active;
  int32_t kin_dims;    // 1: p^2 = p_x^2; 3: full Laplacian (transverse term in kt)
  int32_t eval_every;  // per-chain observable cadence (serving); 0 never
  int32_t pad;
  int64_t eval_from;   // first step index a per-chain cadence may evaluate
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
// a ^ b ^ c in one instruction: gfx950's three-input bitwise op with the
// truth table of XOR3 (0x96).
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Each round uses two
// 32x32->64 products (v_mad_u64_u32) and two three-input XORs.
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c.x;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c.z;
    const uint32_t n0 = xor3(uint32_t(p1 >> 32), c.y, k0);
    const uint32_t n2 = xor3(uint32_t(p0 >> 32), c.w, k1);
    c = make_uint4(n0, uint32_t(p1), n2, uint32_t(p0));
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Two complex N(0, sd^2) draws for (seed, stream, step, site, pair):
// counter (site, pair | stream << 8, step_lo, step_hi), uniforms
// (x + 1/2) 2^-32, Box-Muller (1958) on (u0,u1) and (u2,u3), the standard
// deviation sd folded into the radius.  The logarithm and the sine/cosine
// are the uniform-range routines of clv_math64.h (within 2 ulp).
__device__ __forceinline__ void normals4(double out[4], uint64_t seed, uint32_t stream,
                                         uint64_t step, uint32_t site, uint32_t pair,
                                         double sd = 1.0) {
  const uint4 r = philox10(make_uint4(site, pair | (stream << 8), uint32_t(step), uint32_t(step >> 32)),
                           uint32_t(seed), uint32_t(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = (double(r.x) + 0.5) * s, u1 = (double(r.y) + 0.5) * s;
  const double u2 = (double(r.z) + 0.5) * s, u3 = (double(r.w) + 0.5) * s;
  const double r0 = sd * sqrt(-2.0 * log_unit(u0));
  const double r1 = sd * sqrt(-2.0 * log_unit(u2));
  double s0, c0, s1, c1;
  sincos_2pi(u1, &s0, &c0);
  sincos_2pi(u3, &s1, &c1);
  out[0] = r0 * c0;
  out[1] = r0 * s0;
  out[2] = r1 * c1;
  out[3] = r1 * s1;
}

__device__ __forceinline__ void normals4(float out[4], uint64_t seed, uint32_t stream,
                                         uint64_t step, uint32_t site, uint32_t pair,
                                         float sd = 1.0f) {
  const uint4 r = philox10(make_uint4(site, pair | (stream << 8), uint32_t(step), uint32_t(step >> 32)),
                           uint32_t(seed), uint32_t(seed >> 32));
  // FP32 uniforms from the same 32-bit words as the FP64 path; the hardware
  // transcendentals: v_log_f32 is log2, v_sin/v_cos take revolutions.
  const float s = 2.3283064365386963e-10f;
  const float h = 1.16415321826934814e-10f;
  const float u0 = fmaf(float(r.x), s, h), u1 = fmaf(float(r.y), s, h);
  const float u2 = fmaf(float(r.z), s, h), u3 = fmaf(float(r.w), s, h);
  const float r0 = sd * __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u0));
  const float r1 = sd * __builtin_amdgcn_sqrtf(-1.38629436111989061f * __builtin_amdgcn_logf(u2));
  out[0] = r0 * __builtin_amdgcn_cosf(u1);
  out[1] = r0 * __builtin_amdgcn_sinf(u1);
  out[2] = r1 * __builtin_amdgcn_cosf(u3);
  out[3] = r1 * __builtin_amdgcn_sinf(u3);
}

// ------------------------------------------------------------- controller --
// Propagator::compute_new_dt (Propagator_without_cutoff.cu:382-404).  No FMA
// contraction, so the device and host (clv_compute_new_dt) results are identical.
__device__ __forceinline__ double new_dt(double mean, double mx, const ChainParams &p, int *clamp) {
#pragma clang fp contract(off)
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

// ------------------------------------------------------------ row FFT ----
// Exact twiddle constants of the small DFTs: cos / sin (2 pi m / R).
__host__ __device__ constexpr double kcos(int m, int R) {
  constexpr double c16[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                              0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                              -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977,
                              0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674};
  constexpr double c12[12] = {1.0, 0.86602540378443865, 0.5, 0.0, -0.5, -0.86602540378443865,
                              -1.0, -0.86602540378443865, -0.5, 0.0, 0.5, 0.86602540378443865};
  constexpr double c6[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
  m = ((m % R) + R) % R;
  return R == 16 ? c16[m] : R == 12 ? c12[m] : R == 8 ? c16[2 * m] : R == 6 ? c6[m] : R == 4 ? c16[4 * m] : 1.0;
}
__host__ __device__ constexpr double ksin(int m, int R) {
  // sin(2 pi m / R) = cos(2 pi (m - R/4) / R) for R divisible by 4
  constexpr double s6[6] = {0.0, 0.86602540378443865, 0.86602540378443865, 0.0,
                            -0.86602540378443865, -0.86602540378443865};
  return R == 6 ? s6[((m % 6) + 6) % 6] : kcos(m - R / 4, R);
}

// Small in-register DFTs, y_q = sum_r x_r exp(S 2 pi i r q / R).
template <int S, typename T>
__device__ __forceinline__ void dft2(C2<T> *x) {
  const C2<T> a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}
template <int S, typename T>
__device__ __forceinline__ void dft3(C2<T> *x) {
  const T h = T(0.86602540378443865);  // sqrt(3)/2
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
template <int R, int S, typename T>
__device__ __forceinline__ void dft(C2<T> *x);

// a + g b and a + g (S i) b for a constant g: an add where g = +-1, else
// one FMA per real part.
template <typename T>
__device__ __forceinline__ C2<T> caxpy(C2<T> a, double g, C2<T> b) {
  if (g == 1.0) return cadd(a, b);
  if (g == -1.0) return csub(a, b);
  return {fmad(T(g), b.x, a.x), fmad(T(g), b.y, a.y)};
}
template <int S, typename T>
__device__ __forceinline__ C2<T> caxpy_i(C2<T> a, double g, C2<T> b) {
  const double h = S * g;
  if (h == 1.0) return {a.x - b.y, a.y + b.x};
  if (h == -1.0) return {a.x + b.y, a.y - b.x};
  return {fmad(T(-h), b.y, a.x), fmad(T(h), b.x, a.y)};
}

// x exp(S 2 pi i m / R) = g (x u) with x u cheap: quarter turns (g = 1, u one
// of 1, -1, i, -i: no arithmetic), odd eighth turns (g = 1/sqrt 2,
// u = +-1 +- i: two adds), other angles (u = 1 + i tan with g = cos, or
// cot + i with g = sin, whichever ratio is at most 1: two FMAs).  Returns x u
// and sets *g; the caller folds g into the FMAs of the next butterflies
// (Linzer and Feig 1993, "Modified FFTs for fused multiply-add
// architectures"), so an inner twiddle costs 0 or 2 instructions, not 4.
template <int S, typename T>
__device__ __forceinline__ C2<T> twiddle_unit(C2<T> x, int m, int R, double *g) {
  const double c = kcos(m, R), s = S * ksin(m, R);
  *g = 1.0;
  if (s == 0.0) return c > 0.0 ? x : C2<T>{-x.x, -x.y};
  if (c == 0.0) return s > 0.0 ? C2<T>{-x.y, x.x} : C2<T>{x.y, -x.x};
  const double ac = c < 0.0 ? -c : c, as = s < 0.0 ? -s : s;
  if (ac == as) {
    *g = ac;
    const C2<T> xc = c > 0.0 ? x : C2<T>{-x.x, -x.y};  // x sign(c)
    const C2<T> xs = s > 0.0 ? x : C2<T>{-x.x, -x.y};  // x sign(s)
    return {xc.x - xs.y, xc.y + xs.x};
  }
  if (ac >= as) {
    *g = c;
    const T t = T(s / c);
    return {fmad(-x.y, t, x.x), fmad(x.x, t, x.y)};
  }
  *g = s;
  const T t = T(c / s);
  return {fmad(x.x, t, -x.y), fmad(x.y, t, x.x)};
}

// DFT_B of a_n = g_n b_n (g_0 = 1) given b in x: the scales ride on the
// butterflies' FMAs, so this costs what the unscaled DFT_B costs.
template <int S, typename T>
__device__ __forceinline__ void dft2_scaled(C2<T> *x, const double *g) {
  const C2<T> a = x[0], b = x[1];
  x[0] = caxpy(a, g[1], b);
  x[1] = caxpy(a, -g[1], b);
}
template <int S, typename T>
__device__ __forceinline__ void dft3_scaled(C2<T> *x, const double *g) {
  const double r = g[2] / g[1], h = 0.86602540378443865;  // sqrt(3)/2
  const C2<T> t = caxpy(x[1], r, x[2]);   // (a_1 + a_2) / g_1
  const C2<T> d = caxpy(x[1], -r, x[2]);  // (a_1 - a_2) / g_1
  const C2<T> m = caxpy(x[0], -0.5 * g[1], t);
  x[0] = caxpy(x[0], g[1], t);
  x[1] = caxpy_i<S>(m, h * g[1], d);
  x[2] = caxpy_i<S>(m, -h * g[1], d);
}
template <int S, typename T>
__device__ __forceinline__ void dft4_scaled(C2<T> *x, const double *g) {
  const double r = g[3] / g[1];
  const C2<T> s02 = caxpy(x[0], g[2], x[2]), d02 = caxpy(x[0], -g[2], x[2]);
  const C2<T> s13 = caxpy(x[1], r, x[3]), d13 = caxpy(x[1], -r, x[3]);  // / g_1
  x[0] = caxpy(s02, g[1], s13);
  x[2] = caxpy(s02, -g[1], s13);
  x[1] = caxpy_i<S>(d02, g[1], d13);
  x[3] = caxpy_i<S>(d02, -g[1], d13);
}
template <int B, int S, typename T>
__device__ __forceinline__ void dft_scaled(C2<T> *x, const double *g) {
  if constexpr (B == 2) dft2_scaled<S>(x, g);
  else if constexpr (B == 3) dft3_scaled<S>(x, g);
  else dft4_scaled<S>(x, g);
}

// R = A*B: n = n2 + B n1, k = k1 + A k2 (decimation as in Cooley-Tukey, 1965).
// The inner twiddles W_R^{n2 k1} enter as twiddle_unit and scaled DFT_B.
template <int A, int B, int S, typename T>
__device__ __forceinline__ void dft_split(C2<T> *x) {
  static_assert(B >= 2 && B <= 4, "second factor of a split DFT");
  C2<T> t[A * B];
  double g[A * B];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    C2<T> v[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) v[n1] = x[n2 + B * n1];
    dft<A, S>(v);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1)
      t[n2 * A + k1] = twiddle_unit<S>(v[k1], n2 * k1, A * B, &g[n2 * A + k1]);
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    C2<T> v[B];
    double gk[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
      v[n2] = t[n2 * A + k1];
      gk[n2] = g[n2 * A + k1];
    }
    dft_scaled<B, S>(v, gk);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = v[k2];
  }
}

template <int R, int S, typename T>
__device__ __forceinline__ void dft(C2<T> *x) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<S>(x);
  } else if constexpr (R == 3) {
    dft3<S>(x);
  } else if constexpr (R == 4) {
    dft4<S>(x);
  } else if constexpr (R == 6) {
    dft_split<3, 2, S>(x);
  } else if constexpr (R == 8) {
    dft_split<4, 2, S>(x);
  } else if constexpr (R == 12) {
    dft_split<4, 3, S>(x);
  } else if constexpr (R == 16) {
    dft_split<4, 4, S>(x);
  }
}

// Row length X = N1 * N2 of the two-level transform: the default split per X
// (RowFFT takes any N1 dividing X with N1 and X / N1 in the small DFTs above;
// the engine instantiates the default and one alternative per length, and
// its autotuner picks between them).
template <int X> struct Split;
template <> struct Split<4>   { static constexpr int N1 = 2, N2 = 2; };
template <> struct Split<8>   { static constexpr int N1 = 4, N2 = 2; };
template <> struct Split<16>  { static constexpr int N1 = 4, N2 = 4; };
template <> struct Split<24>  { static constexpr int N1 = 6, N2 = 4; };
template <> struct Split<32>  { static constexpr int N1 = 8, N2 = 4; };
template <> struct Split<48>  { static constexpr int N1 = 8, N2 = 6; };
template <> struct Split<64>  { static constexpr int N1 = 8, N2 = 8; };
template <> struct Split<96>  { static constexpr int N1 = 12, N2 = 8; };
template <> struct Split<128> { static constexpr int N1 = 16, N2 = 8; };
template <> struct Split<192> { static constexpr int N1 = 16, N2 = 12; };
template <> struct Split<256> { static constexpr int N1 = 16, N2 = 16; };

// Kinetic term of rows of X sites held in LDS, in place, by a two-level
// (four-step) FFT: x[n2 + N2 n1] -> DFT_N1 over n1, twiddle W_X^{n2 k1} ->
// DFT_N2 over n2 gives X[k1 + N1 k2]; multiply by m(k) = -p_x(k)^2/(a_x X);
// the inverse runs the same two levels backwards and leaves x in natural
// order.  Each level reads and writes the same LDS cells, so one buffer
// suffices; a line is padded by one cell per N2 so that both access patterns
// are free of bank conflicts.
// Cells per line: at least X + X/N2; for X >= 32 also congruent to N2 modulo
// 32, so that the lanes of a first-level task (consecutive n2, consecutive
// lines) fall in distinct 8-byte LDS bank pairs.  Short lines keep the
// minimum: padding them to that congruence would multiply their LDS footprint.
__host__ __device__ constexpr int row_pad(int X, int N2) {
  int p = X + X / N2;
  if (X < 32) return p;
  while (p % 32 != N2 % 32) ++p;
  return p;
}

// Where the forward twiddles W_X^{n2 k1} are applied.  With short middle-level
// lines (N2 <= 12) in the middle level, from the table entry of the inverse
// twiddle W_X^{-m2 k1} at m2 = n2, conjugated: the first level then holds
// only its N1 values, which keeps the fused march's register peak low.  With
// N2 = 16 after the first level's DFT (a forward table of its own), where
// the middle level's 2 N2 values and multipliers would need more registers.
__host__ __device__ constexpr bool mid_twiddles(int N2) { return N2 <= 12; }
// complex entries of the twiddle table of a split X = N1 * N2
__host__ __device__ constexpr int tw_entries(int X, int N1) { return mid_twiddles(X / N1) ? X : 2 * X; }

template <typename T, int X, int N1_ = Split<X>::N1>
struct RowFFT {
  static constexpr int N1 = N1_, N2 = X / N1_;
  static_assert(N1 * N2 == X, "N1 must divide the row length");
  static constexpr int PAD = row_pad(X, N2);  // LDS cells per line
  static constexpr bool MID = mid_twiddles(N2);
  static constexpr int INV = MID ? 0 : X;     // first inverse twiddle in the table
  __device__ static __forceinline__ int pos(int p) { return p + p / N2; }

  // lines l < nl at buf + l*PAD; mult[k] = m(k); twiddles in the order each
  // level reads them, so the lanes of one access hit consecutive cells:
  //   tw[INV + m2*N1 + k1] = exp(+2 pi i m2 k1 / X), and without MID also
  //   tw[k1*N2 + n2] = exp(-2 pi i n2 k1 / X) (with MID its conjugate is used).
  // Caller syncs before (data written) and reads after the trailing sync.
  __device__ static __forceinline__ void kinetic(C2<T> *buf, const C2<T> *tw, const T *mult,
                                                 int nl, int tid, int nthr) {
    for (int task = tid; task < nl * N2; task += nthr) {
      const int l = task / N2, n2 = task - (task / N2) * N2;
      C2<T> *L = buf + l * PAD;
      C2<T> v[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) v[n1] = L[pos(n2 + N2 * n1)];
      dft<N1, -1>(v);
      if constexpr (!MID) {
#pragma unroll
        for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], tw[k1 * N2 + n2]);
      }
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) L[pos(n2 + N2 * k1)] = v[k1];
    }
    __syncthreads();
    for (int task = tid; task < nl * N1; task += nthr) {
      const int l = task / N1, k1 = task - (task / N1) * N1;
      C2<T> *L = buf + l * PAD;
      C2<T> v[N2], w[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) v[n2] = L[pos(n2 + N2 * k1)];
      if constexpr (MID) {
#pragma unroll
        for (int n2 = 1; n2 < N2; ++n2) {
          w[n2] = tw[n2 * N1 + k1];
          v[n2] = cmul(v[n2], cconj(w[n2]));
        }
      }
      dft<N2, -1>(v);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) v[k2] = cscale(mult[k1 + N1 * k2], v[k2]);
      dft<N2, +1>(v);
#pragma unroll
      for (int m2 = 1; m2 < N2; ++m2) v[m2] = cmul(v[m2], MID ? w[m2] : tw[INV + m2 * N1 + k1]);
#pragma unroll
      for (int m2 = 0; m2 < N2; ++m2) L[pos(m2 + N2 * k1)] = v[m2];
    }
    __syncthreads();
    for (int task = tid; task < nl * N2; task += nthr) {
      const int l = task / N2, m2 = task - (task / N2) * N2;
      C2<T> *L = buf + l * PAD;
      C2<T> v[N1];
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) v[k1] = L[pos(m2 + N2 * k1)];
      dft<N1, +1>(v);
#pragma unroll
      for (int m1 = 0; m1 < N1; ++m1) L[pos(m2 + N2 * m1)] = v[m1];
    }
    __syncthreads();
  }
};

}  // namespace clv
