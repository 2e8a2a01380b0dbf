// clv_engine.hip -- gfx950 kernels of the complex-Langevin step and the C ABI
// of include/clv.h.
//
// One Langevin step of a bucket of equal-shape chains (two-pass schedule):
//   k_drift_stats  K_0(t-1) by an LDS row transform, D_0, per-tile max/sum of |D_0|
//   k_controller   fixed-order per-chain reduction of the tiles, dt (device) or
//                  stats for the host (plain method), bookkeeping
//   k_update       K(t-1), Kbar(t+1) for all components, D and Dbar, Philox
//                  noise in registers, psi/psibar written to the other buffer
// Observables (every eval_every steps): k_obs (per tau-chunk) + k_obs_finish.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/clv.h"
#include "clv_internal.h"
#include "clv_kernels.h"

using namespace clv;
using namespace clvi;

// Waves per SIMD the fused kernel is compiled for (register budget 512/n),
// FP32 fields; the FP64 kernel holds twice the data and is built for 2.
constexpr int kFusedWavesPerEU = 3;

// ============================================================== kernels ====

// D_0 at every site of TC tau-slices of `rows` rows; writes each slice's tile
// (max |D_0|, sum |D_0|) -- compute_max_and_mean_drift, Propagator_without_cutoff.cu:231-265,
// in FP64 and without atomics (tile_max_sum, one wave per slice).
template <typename T, int X, int N1>
__global__ __launch_bounds__(256) void k_drift_stats(BucketArgs<T> A) {
  extern __shared__ __align__(16) char smem_raw[];
  const int chain = A.chain0 + blockIdx.z;
  const ChainParams &P = A.prm[chain];
  if (!P.active) return;
  // TC slices from t0; with a stride, one slice per workgroup (the first tau
  // of each block of the tau-blocked fused march)
  const int TC = A.tstride ? 1 : A.tcs;
  const int tile = blockIdx.x, t0 = blockIdx.y * (A.tstride ? A.tstride : TC);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int j = tid / X, x = tid - (tid / X) * X;
  const bool has_site = j < A.rows;
  const int rows = A.rows;
  Smem<T, X, N1> sm(smem_raw, TC * rows);
  double *vb = reinterpret_cast<double *>(sm.red_end());  // [TC][nthr] |D_0|
  load_tables<T, X, N1>(sm, A.tw, A.p2, drift_table_scale(P), tid, nthr);
  const int par = A.st[chain].parity;  // current buffer
  const C2<T> *psi = (par ? A.buf1_psi : A.buf0_psi) + chain * A.chain_stride;
  const C2<T> *psib = (par ? A.buf1_psib : A.buf0_psib) + chain * A.chain_stride;
  const int r = tile * rows + j;
  // psi(t-1) of the TC slices, each slice read once: its component 0 feeds
  // the transform batch and the whole site stays in registers for the drift
  constexpr int TCMAX = 8;
  C2<T> pm[TCMAX][4];
  if (has_site) {
#pragma unroll
    for (int q = 0; q < TCMAX; ++q) {
      if (q >= TC) break;
      const int t = t0 + q;
      const int tm1 = (t == 0) ? A.Nt - 1 : t - 1;
      load4(psi, (int64_t(tm1) * A.R + r) * X + x, pm[q]);
      sm.at(q * rows + j, x) = pm[q][0];
    }
  }
  __syncthreads();
  RowFFT<T, X, N1>::kinetic(sm.buf, sm.tw, sm.mult, TC * rows, tid, nthr);
  const Coef<T> k = coef<T>(P);
#pragma unroll
  for (int q = 0; q < TCMAX; ++q) {
    if (q >= TC) break;
    const int t = t0 + q;
    double v = 0.0;
    if (has_site) {
      C2<T> b0[4];
      load4(psib, (int64_t(t) * A.R + r) * X + x, b0);
      const C2<T> p00 = (q + 1 < TC) ? pm[q + 1][0] : load1(psi, (int64_t(t) * A.R + r) * X + x, 0);
      C2<T> km0 = sm.at(q * rows + j, x);
      if (has_kt(A, P)) {
        const int tm1 = (t == 0) ? A.Nt - 1 : t - 1;
        km0 = add_kt(km0, k.kK, load1(A.kt_psi + chain * A.chain_stride, (int64_t(tm1) * A.R + r) * X + x, 0));
      }
      v = drift_abs(drift0_site(k, pm[q], p00, b0, km0));
    }
    vb[q * nthr + tid] = v;
  }
  __syncthreads();
  const int w = tid >> 6, nw = nthr >> 6, lane = tid & 63;
  for (int q = w; q < TC; q += nw) {
    double m, s;
    tile_max_sum(vb + q * nthr, nthr, lane, &m, &s);
    if (lane == 0) {
      double *o = A.part_wr + ((int64_t(chain) * A.Nt + t0 + q) * A.ntiles + tile) * 2;
      o[0] = m;
      o[1] = s;
    }
  }
}

// Per-chain controller: one wave per chain reduces the tiles' partials in a
// fixed order.  mode 0: dt on the device (compute_new_dt, :382-404) and commit;
// mode 1: only publish max/mean (the plain method reads them on the host);
// mode 2: commit a dt chosen on the host (clamp code in host_clamp).
template <typename T>
__global__ __launch_bounds__(64) void k_controller(BucketArgs<T> A, int mode, double host_dt,
                                                   int host_clamp) {
  const int chain = A.chain0 + blockIdx.x;
  const ChainParams &P = A.prm[chain];
  if (!P.active) return;
  ChainState &st = A.st[chain];
  const int lane = threadIdx.x;
  if (mode != 2) {
    const int n = A.Nt * A.ntiles;
    double m, s;
    reduce_partials(A.part_rd + int64_t(chain) * n * 2, n, lane, &m, &s);
    if (lane == 0) {
      const double mean = s / double(4 * A.S);
      st.max_drift = m;
      st.mean_drift = mean;
      if (!isfinite(m) || !isfinite(s)) st.nonfinite = 1;
      if (mode == 0) {
        int clamp = 0;
        const double dt = new_dt(mean, m, P, &clamp);
        st.dt = dt;
        st.cur_step = st.steps;
        st.steps += 1;
        st.t_langevin += dt;
        if (clamp < 0) st.clamp_low += 1;
        if (clamp > 0) st.clamp_high += 1;
        st.in_parity = st.parity;  // this step reads the current buffer
        st.parity = 1 - st.parity; // and its output becomes current
      }
    }
  } else if (lane == 0) {
    st.dt = host_dt;
    st.cur_step = st.steps;
    st.steps += 1;
    st.t_langevin += host_dt;
    if (host_clamp < 0) st.clamp_low += 1;
    if (host_clamp > 0) st.clamp_high += 1;
    st.in_parity = st.parity;
    st.parity = 1 - st.parity;
  }
}

// The update of one tau-slice of `rows` rows: multiply_by_timestep + add_noise
// + fields += drift (Propagator_without_cutoff.cu:131-157, 432-441) with the
// drift recomputed in registers and the noise drawn from Philox in registers.
template <typename T, int X, int N1>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 5 : 2) void k_update(BucketArgs<T> A) {
  extern __shared__ __align__(16) char smem_raw[];
  const int chain = A.chain0 + blockIdx.z;
  const ChainParams &P = A.prm[chain];
  if (!P.active) return;
  const int tile = blockIdx.x, t = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int j = tid / X, x = tid - (tid / X) * X;
  const bool has_site = j < A.rows;
  Smem<T, X, N1> sm(smem_raw, 8 * A.rows);
  load_tables<T, X, N1>(sm, A.tw, A.p2, drift_table_scale(P), tid, nthr);
  const ChainState &st = A.st[chain];
  const int par = st.in_parity;
  const double dt = st.dt;
  const uint64_t step = uint64_t(st.cur_step);
  const C2<T> *psi = (par ? A.buf1_psi : A.buf0_psi) + chain * A.chain_stride;
  const C2<T> *psib = (par ? A.buf1_psib : A.buf0_psib) + chain * A.chain_stride;
  C2<T> *opsi = (par ? A.buf0_psi : A.buf1_psi) + chain * A.chain_stride;
  C2<T> *opsib = (par ? A.buf0_psib : A.buf1_psib) + chain * A.chain_stride;
  const int rows = A.rows;
  const int r = tile * rows + j;
  const int tm1 = (t == 0) ? A.Nt - 1 : t - 1;
  const int tp1 = (t == A.Nt - 1) ? 0 : t + 1;
  C2<T> pm[4], p0[4], pp[4], b0[4], bp[4], km[4], kp[4];
  const int64_t s0 = (int64_t(t) * A.R + r) * X + x;
  const int64_t sp1 = (int64_t(tp1) * A.R + r) * X + x;
  // K(t-1) of psi and Kbar(t+1) of psibar: one batch of 8 row sets
  if (has_site) {
    load4(psi, (int64_t(tm1) * A.R + r) * X + x, pm);
    load4(psib, sp1, bp);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      sm.at(c * rows + j, x) = pm[c];
      sm.at((4 + c) * rows + j, x) = bp[c];
    }
  }
  __syncthreads();
  RowFFT<T, X, N1>::kinetic(sm.buf, sm.tw, sm.mult, 8 * rows, tid, nthr);
  if (!has_site) return;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    km[c] = sm.at(c * rows + j, x);
    kp[c] = sm.at((4 + c) * rows + j, x);
  }
  load4(psi, s0, p0);
  load4(psi, sp1, pp);
  load4(psib, s0, b0);
  const Coef<T> k = coef<T>(P);
  if (has_kt(A, P)) {
    C2<T> ktm[4], ktp[4];
    load4(A.kt_psi + chain * A.chain_stride, (int64_t(tm1) * A.R + r) * X + x, ktm);
    load4(A.kt_psib + chain * A.chain_stride, sp1, ktp);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      km[c] = add_kt(km[c], k.kK, ktm[c]);
      kp[c] = add_kt(kp[c], k.kK, ktp[c]);
    }
  }
  C2<T> D[4], Db[4];
  drift_site(k, pm, p0, pp, b0, bp, km, kp, D, Db);
  T nz[8] = {};
  if (P.noise_on) {
    const uint32_t site = uint32_t((int64_t(t) * X + x) * A.R + r);  // reference site order
    const T sq = T(sqrt(dt));
    normals4(nz, P.seed, 0u, step, site, 0u, sq);
    normals4(nz + 4, P.seed, 0u, step, site, 1u, sq);
  }
  const T dtT = T(dt);
  C2<T> o[4], ob[4];
  update_site(dtT, p0, b0, D, Db, nz, o, ob);
  store4(opsi, s0, o);
  store4(opsib, s0, ob);
}

// Fused single-pass step (schedule 2).  One workgroup owns `rows` rows of one
// chain for all Nt slices and marches along tau:
//   1. dt of this step from the partials of the current fields (the same
//      fixed-order reduction as k_controller; workgroup 0 commits the state),
//   2. per tau: K(t) and Kbar(t+1) of the old fields in one LDS batch of 8
//      row sets, drift + Philox noise + update, and the non-kinetic part of
//      D_0 of the NEW fields at t (next step's |D_0| partials).  The kinetic
//      part, K_0 of the new psi_0(t-1), is one more row set per tau; those are
//      transformed `fbatch` taus at a time (parked in LDS until then), so most
//      taus transform exactly 8 row sets, which fill the waves of every
//      transform level.
// Old fields are read from the current buffer and the new ones written to the
// other (as the two-pass update does), so the wrap-around neighbours of
// tau = Nt-1 are still the old tau = 0.  Reads and writes the state once per
// step; results are bit-identical to the two-pass schedule for every fbatch
// and line-set count (the same per-line transforms and drift0_site split).
template <typename T, int X, int N1>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? kFusedWavesPerEU : 2) void k_fused(BucketArgs<T> A) {
  extern __shared__ __align__(16) char smem_raw[];
  const int chain = A.chain0 + blockIdx.y;
  const int tile = blockIdx.x;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const ChainParams &P = A.prm[chain];
  const ChainState &st = A.st[chain];
  // tau block of this workgroup: [t0, t0 + Lb)
  const int Lb = A.Nt / A.fblocks, t0 = int(blockIdx.z) * Lb, t1 = t0 + Lb;
  if (!P.active) {
    if (tile == 0 && t0 == 0 && tid == 0) A.st_next[chain] = st;
    return;
  }
  const int rows = A.rows, Nt = A.Nt, B = A.fbatch;
  // line sets of 8 + B row sets: 8 for the old fields' transforms, B slots for
  // the parked next-step K_0 inputs (used in the set of the batch taus)
  const int setw = (8 + B) * rows;
  Smem<T, X, N1> sm(smem_raw, A.fsets * setw);
  C2<T> *lpark = sm.red_end();  // [B-1][nthr] parked non-kinetic D_0 parts
  double *vb = reinterpret_cast<double *>(lpark + (B - 1) * nthr);  // [max(B,2)][nthr] |D_0|
  using RF = RowFFT<T, X, N1>;
  load_tables<T, X, N1>(sm, A.tw, A.p2, drift_table_scale(P), tid, nthr);
  if (tid < 64) {
    const int n = Nt * A.ntiles;
    double m, s;
    reduce_partials(A.part_rd + int64_t(chain) * n * 2, n, tid, &m, &s);
    if (tid == 0) {
      const double mean = s / double(4 * A.S);
      int clamp = 0;
      const double dt = new_dt(mean, m, P, &clamp);
      sm.red[0] = dt;
      if (tile == 0 && t0 == 0) {
        ChainState ns = st;
        ns.dt = dt;
        ns.max_drift = m;
        ns.mean_drift = mean;
        if (!isfinite(m) || !isfinite(s)) ns.nonfinite = 1;
        ns.cur_step = st.steps;
        ns.steps = st.steps + 1;
        ns.t_langevin = st.t_langevin + dt;
        if (clamp < 0) ns.clamp_low += 1;
        if (clamp > 0) ns.clamp_high += 1;
        ns.in_parity = st.parity;
        ns.parity = 1 - st.parity;
        A.st_next[chain] = ns;
      }
    }
  }
  __syncthreads();
  const double dt = sm.red[0];
  const uint64_t step = uint64_t(st.steps);
  const int j = tid / X, x = tid - (tid / X) * X;
  const bool has_site = j < rows;
  const int r = tile * rows + j;
  const int par = st.parity;
  const C2<T> *psi = (par ? A.buf1_psi : A.buf0_psi) + chain * A.chain_stride;
  const C2<T> *psib = (par ? A.buf1_psib : A.buf0_psib) + chain * A.chain_stride;
  C2<T> *opsi = (par ? A.buf0_psi : A.buf1_psi) + chain * A.chain_stride;
  C2<T> *opsib = (par ? A.buf0_psib : A.buf1_psib) + chain * A.chain_stride;
  const Coef<T> k = coef<T>(P);
  const T dtT = T(dt);
  const T sq = T(sqrt(dt));
  auto site_at = [&](int t) { return (int64_t(t) * A.R + r) * X + x; };
  // In a line set the 8 row-set lines of row j are adjacent (line 8 j + c),
  // so a thread's cells sit at fixed offsets from one address; the parked
  // K_0 slots follow them (line 8 rows + q rows + j).
  C2<T> *const cell = sm.buf + (has_site ? 8 * j : 0) * RF::PAD + RF::pos(x);
  // Registers carried along tau (old = fields of this step, new = updated):
  //   dc = the psi drift D(t) (formed in the previous iteration, right after
  //   the transform that gave K(t-1)), nprev = new psi(t-1),
  //   p0 = old psi(t), b0 = old psibar(t), bp = old psibar(t+1).
  // Each iteration loads old psi(t+1) (needed after the transform) and
  // prefetches old psibar(t+2) (the next iteration's transform input), so
  // every slice is read from memory once and the loads overlap the work
  // (issuing the prefetch after the transform instead made the step 5%
  // slower).  Carrying D(t) instead of psi(t-1), K(t-1) and kA N keeps the
  // register footprint small.
  C2<T> dc[4], nprev[4], p0[4], b0[4], bp[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    nprev[c] = {T(0), T(0)};
    dc[c] = {T(0), T(0)};
  }
  {
    C2<T> pm[4];
    if (has_site) {
      load4(psi, site_at(t0 == 0 ? Nt - 1 : t0 - 1), pm);
      load4(psi, site_at(t0), p0);
      load4(psib, site_at(t0), b0);
      load4(psib, site_at((t0 + 1) % Nt), bp);
#pragma unroll
      for (int c = 0; c < 4; ++c) sm.at(c * rows + j, x) = pm[c];
    }
    __syncthreads();
    RF::kinetic(sm.buf, sm.tw, sm.mult, 4 * rows, tid, nthr);
    if (has_site) {
      C2<T> km[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) km[c] = sm.at(c * rows + j, x);
      drift_psi(k, pm, p0, km, b0, drift_an(k, pm, b0), dc);
    }
  }
  __syncthreads();  // the LDS lines are rewritten below
  // With two LDS sets the transform lines alternate by tau parity, so a tau's
  // lines can be written while slower waves still read the previous tau's
  // results (one set: a barrier ends each tau).  Taus are counted from the
  // block's first, u = t - t0.  The batch taus (u % B == B-1) are odd for
  // B > 1, so the parked slots sit at the end of set (B > 1) * (fsets - 1).
  // Each tau's |D_0| go to LDS slot u % max(B, 2); after the next barrier
  // one wave per tau reduces them (tile_max_sum) and writes the partial, the
  // B taus of a batch spread over the waves.  The block's first tau has no
  // |D_0| here (it needs the new psi of the tau before, another block's):
  // with one block the march ends with it, else k_drift_stats adds it.
  const int w = tid >> 6, nw = nthr >> 6, lane = tid & 63;
  const int rmask = (B < 2 ? 2 : B) - 1;
  const int park = (B > 1 ? A.fsets - 1 : 0) * setw + 8 * rows;  // first parked slot (B > 1)
  auto put_v = [&](int u, double v) { vb[(u & rmask) * nthr + tid] = v; };
  auto partial = [&](int u, int slot) {  // one whole wave, after a barrier
    double m, s;
    tile_max_sum(vb + slot * nthr, nthr, lane, &m, &s);
    if (lane == 0) {
      double *o = A.part_wr + ((int64_t(chain) * Nt + t0 + u) * A.ntiles + tile) * 2;
      o[0] = m;
      o[1] = s;
    }
  };
  // the partials of batch tau ub: taus max(1, ub-B+1) .. ub, q-th by wave q % nw
  auto reduce_batch = [&](int ub) {
    for (int q = w; q < B; q += nw) {
      const int u = ub - B + 1 + q;
      if (u >= 1) partial(u, u & rmask);
    }
  };
  for (int t = t0; t < t1; ++t) {
    const int u = t - t0;
    const int tp1 = (t == Nt - 1) ? 0 : t + 1;
    const int L0 = (u & (A.fsets - 1)) * setw;  // first line of this tau's set
    const bool batch = u > 0 && (u & (B - 1)) == B - 1;
    C2<T> pp[4], bn[4];
    if (has_site) {
      load4(psi, site_at(tp1), pp);
      if (t + 1 < t1) load4(psib, site_at((t + 2) % Nt), bn);
      C2<T> *const cl = cell + L0 * RF::PAD;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        cl[c * RF::PAD] = p0[c];
        cl[(4 + c) * RF::PAD] = bp[c];
      }
      // K_0 input of tau = t (new psi_0(t-1)); at a batch tau into its own
      // slot of this set, otherwise parked after this tau's first barrier
      if (batch) sm.at(L0 + (8 + B - 1) * rows + j, x) = nprev[0];
    }
    __syncthreads();
    if (u > 1 && ((u - 1) & (B - 1)) == B - 1) reduce_batch(u - 1);
    if (has_site && !batch && u > 0) sm.at(park + (u & (B - 1)) * rows + j, x) = nprev[0];
    RF::kinetic(sm.buf + L0 * RF::PAD, sm.tw, sm.mult, (batch ? 8 + B : 8) * rows, tid, nthr);
    // the parked taus of this batch
    if (batch) {
      for (int q = 0; q < B - 1; ++q) {
        const int uq = u - (B - 1) + q;
        if (uq < 1) continue;
        double v = 0.0;
        if (has_site) v = drift_abs(drift0_add_kin(lpark[q * nthr + tid], sm.at(L0 + (8 + q) * rows + j, x)));
        put_v(uq, v);
      }
    }
    double v_batch = 0.0;
    if (has_site) {
      const C2<T> *const cl = cell + L0 * RF::PAD;
      T nz[8] = {};
      if (P.noise_on) {
        const uint32_t site = uint32_t((int64_t(t) * X + x) * A.R + r);
        normals4(nz, P.seed, 0u, step, site, 0u, sq);
        normals4(nz + 4, P.seed, 0u, step, site, 1u, sq);
      }
      // psibar drift of this site (K of psibar(t+1) from this transform);
      // kA N of the pair (psi(t), psibar(t+1)) is also the psi pair of t+1
      const C2<T> an = drift_an(k, p0, bp);
      C2<T> o[4], ob[4];
      {
        C2<T> kp[4], Db[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) kp[c] = cl[(4 + c) * RF::PAD];
        drift_psib(k, p0, pp, b0, bp, kp, an, Db);
        update_site(dtT, p0, b0, dc, Db, nz, o, ob);
      }
      const int64_t s0 = site_at(t);
      store4(opsi, s0, o);
      store4(opsib, s0, ob);

      // the psi drift of t+1 from K of psi(t) (this transform)
      if (t + 1 < t1) {
        C2<T> kn[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) kn[c] = cl[c * RF::PAD];
        drift_psi(k, p0, pp, kn, bp, an, dc);
      }
      if (u > 0) {
        const C2<T> l0 = drift0_local(k, nprev, o[0], ob);
        if (batch) v_batch = drift_abs(drift0_add_kin(l0, sm.at(L0 + (8 + B - 1) * rows + j, x)));
        else lpark[(u & (B - 1)) * nthr + tid] = l0;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        p0[c] = pp[c];
        b0[c] = bp[c];
        bp[c] = bn[c];
        nprev[c] = o[c];
      }
    }
    if (batch) put_v(u, v_batch);
    if (A.fsets == 1) __syncthreads();  // the one set is rewritten by the next tau
  }
  __syncthreads();
  const int ul = Lb - 1;  // last tau of the march
  const bool last_batch = ul > 0 && (ul & (B - 1)) == B - 1;
  if (last_batch) reduce_batch(ul);
  // K_0 lines still to transform: the taus parked after the last batch
  // (slots 0 .. np-1; slot 0 unused before the first batch) and, with one
  // block, tau = 0 (new psi_0(Nt-1)) in the slot after them.  B = 1 uses line 0.
  const int np = (last_batch || ul == 0) ? 0 : (ul & (B - 1)) + 1;
  const int nz0 = (A.fblocks == 1) ? 1 : 0;  // the tau = 0 item
  if (np + nz0 == 0) return;                 // uniform: nothing pending
  const int pb = (B > 1) ? park : 0;
  // |D_0| of the new tau = 0 (needs the new psi(Nt-1)); this thread wrote
  // the new tau = 0 values of its site at t = 0
  C2<T> n0k[4], nb0k[4];
  if (has_site && nz0) {
    load4(opsi, site_at(0), n0k);
    load4(opsib, site_at(0), nb0k);
    sm.at(pb + np * rows + j, x) = nprev[0];
  }
  __syncthreads();
  RF::kinetic(sm.buf + pb * RF::PAD, sm.tw, sm.mult, (np + nz0) * rows, tid, nthr);
  // the pending taus in slots 0 .. np-1 (the slot of u0 + q is q), tau = 0
  // in slot np; items i = 0 .. np + nz0 - 1, item i reduced by wave i % nw
  const int u0 = ul - (np - 1);
  for (int q = 0; q < np; ++q) {
    if (u0 + q < 1) continue;
    double v = 0.0;
    if (has_site) v = drift_abs(drift0_add_kin(lpark[q * nthr + tid], sm.at(pb + q * rows + j, x)));
    vb[q * nthr + tid] = v;
  }
  if (nz0) {
    double v0 = 0.0;
    if (has_site) v0 = drift_abs(drift0_site(k, nprev, n0k[0], nb0k, sm.at(pb + np * rows + j, x)));
    vb[np * nthr + tid] = v0;
  }
  __syncthreads();
  for (int i = w; i < np + nz0; i += nw) {
    if (i < np && u0 + i < 1) continue;
    partial(i < np ? u0 + i : 0, i);
  }
}

// Observables of the current fields over one tau-chunk: n_c (real space,
// equal to the k-space sums of evaluate_n_* by Parseval), F_a(x)
// (evaluate_spin_density_realspace, Observables_without_cutoff.cu:456-488)
// and the energy density h on the time-split pair (psibar(t+1), psi(t)).
// S = obs_slices(sizeof(C2<T>), X) slices per transform batch (two fill the
// transform levels that one slice of a long row leaves mostly idle, with half
// the barriers per slice).  Each thread still adds its slices in tau order,
// so S does not change the result.
template <typename T, int X, int N1>
__global__ __launch_bounds__(256) void k_obs(BucketArgs<T> A) {
  constexpr int S = obs_slices(int(sizeof(C2<T>)), X);
  extern __shared__ __align__(16) char smem_raw[];
  const int chain = A.chain0 + blockIdx.z;
  const ChainParams &P = A.prm[chain];
  if (!P.active) return;
  const ChainState &st = A.st[chain];
  if (!evaluates(A, P, st)) return;
  const int tile = blockIdx.x, chunk = blockIdx.y;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int j = tid / X, x = tid - (tid / X) * X;
  const bool has_site = j < A.rows;
  const int rows = A.rows, Nt = A.Nt, tc = A.tc;
  Smem<T, X, N1> sm(smem_raw, 4 * S * rows);
  load_tables<T, X, N1>(sm, A.tw, A.p2, 1.0 / P.a_x, tid, nthr);
  const int par = st.parity;
  const C2<T> *psi = (par ? A.buf1_psi : A.buf0_psi) + chain * A.chain_stride;
  const C2<T> *psib = (par ? A.buf1_psib : A.buf0_psib) + chain * A.chain_stride;
  const int r = tile * rows + j;
  const double ax = P.a_x, G = P.G, c1 = P.c1;
  const double sz[4] = {P.P + P.Q, 0.0, 0.0, -P.P + P.Q};
  // interaction part of h through the matrix invariants of the pair
  // (psibar(t+1), psi(t)): G tr[(Psibar^T Psi)^2] + (c1/2) sum_a F_a^2
  //   = (G + c1/2) N^2 - (2G + 2c1) det Psibar det Psi   (N = sum_c psib_c psi_c)
  const double gN = G + 0.5 * c1, gD = 2.0 * G + 2.0 * c1;
  double n[4] = {0, 0, 0, 0}, e = 0.0, F[3] = {0, 0, 0};
  auto site_at = [&](int t) { return (int64_t(t) * A.R + r) * X + x; };
  // psi(t) and psibar(t+1) of the slices t .. t+cnt-1
  auto load_pair = [&](int t, int cnt, C2<T> qd[S][4], C2<T> bd[S][4]) {
#pragma unroll
    for (int u = 0; u < S; ++u) {
      if (u >= cnt) break;
      const int tu = t + u;
      load4(psi, site_at(tu), qd[u]);
      load4(psib, site_at(tu + 1 == Nt ? 0 : tu + 1), bd[u]);
    }
  };
  // a batch's fields are loaded one batch ahead, overlapping the transform;
  // slice u's 4 lines of row j are adjacent (line 4 rows u + 4 j + c)
  constexpr int PAD = RowFFT<T, X, N1>::PAD;
  C2<T> *const cell = sm.buf + (has_site ? 4 * j : 0) * PAD + RowFFT<T, X, N1>::pos(x);
  C2<T> q[S][4], b[S][4];
  if (has_site) load_pair(chunk * tc, tc < S ? tc : S, q, b);
  for (int tt = 0; tt < tc; tt += S) {
    const int t = chunk * tc + tt, cnt = tc - tt < S ? tc - tt : S;
    C2<T> qn[S][4], bn[S][4];
    if (has_site) {
#pragma unroll
      for (int u = 0; u < S; ++u) {
        if (u >= cnt) break;
#pragma unroll
        for (int c = 0; c < 4; ++c) cell[(4 * rows * u + c) * PAD] = q[u][c];
      }
      if (tt + S < tc) load_pair(t + S, tc - tt - S < S ? tc - tt - S : S, qn, bn);
    }
    __syncthreads();
    RowFFT<T, X, N1>::kinetic(sm.buf, sm.tw, sm.mult, 4 * cnt * rows, tid, nthr);
    if (has_site) {
#pragma unroll
      for (int u = 0; u < S; ++u) {
        if (u >= cnt) break;
        C2<double> bb[4], qq[4], kk[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const C2<T> kc = cell[(4 * rows * u + c) * PAD];
          bb[c] = {double(b[u][c].x), double(b[u][c].y)};
          qq[c] = {double(q[u][c].x), double(q[u][c].y)};
          kk[c] = {double(kc.x), double(kc.y)};
          if (has_kt(A, P)) {
            const C2<T> kt = load1(A.kt_psi + chain * A.chain_stride, site_at(t + u), c);
            kk[c].x += double(kt.x);
            kk[c].y += double(kt.y);
          }
        }
        // only Re h is accumulated
        C2<double> nn = {0.0, 0.0};
        double hr = 0.0, bq_re[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const C2<double> bq = cmul(bb[c], qq[c]);
          bq_re[c] = bq.x;
          n[c] += bq.x * ax;
          nn = cadd(nn, bq);
          hr -= (bb[c].x * kk[c].x - bb[c].y * kk[c].y) + sz[c] * bq.x;
        }
        const C2<double> detb = csub(cmul(bb[0], bb[3]), cmul(bb[1], bb[2]));
        const C2<double> detq = csub(cmul(qq[0], qq[3]), cmul(qq[1], qq[2]));
        hr += gN * (nn.x * nn.x - nn.y * nn.y) - gD * (detb.x * detq.x - detb.y * detq.y);
        e += hr * ax;
        // Re F_x, Re F_y = -Im(psib_2 psi_0 + psib_3 psi_1 - psib_0 psi_2 - psib_1 psi_3), Re F_z
        F[0] += (bb[2].x * qq[0].x - bb[2].y * qq[0].y) + (bb[0].x * qq[2].x - bb[0].y * qq[2].y) +
                (bb[3].x * qq[1].x - bb[3].y * qq[1].y) + (bb[1].x * qq[3].x - bb[1].y * qq[3].y);
        F[1] -= (bb[2].x * qq[0].y + bb[2].y * qq[0].x) + (bb[3].x * qq[1].y + bb[3].y * qq[1].x) -
                (bb[0].x * qq[2].y + bb[0].y * qq[2].x) - (bb[1].x * qq[3].y + bb[1].y * qq[3].x);
        F[2] += (bq_re[0] + bq_re[1]) - (bq_re[2] + bq_re[3]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < S; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        q[u][c] = qn[u][c];
        b[u][c] = bn[u][c];
      }
  }
  if (has_site) {
    const int64_t vs = int64_t(x) * A.R + r;  // reference spatial order (x, y, z)
    double *o = A.obs_stage + ((int64_t(chain) * A.nchunks + chunk) * A.Vs + vs) * 3;
    o[0] = F[0];
    o[1] = F[1];
    o[2] = F[2];
  }
  double vals[5] = {n[0], n[1], n[2], n[3], e};
  const int w = tid >> 6, nw = (nthr + 63) >> 6;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const double v = wave_sum(vals[q]);
    if ((tid & 63) == 0) sm.red[q * 8 + w] = v;
  }
  __syncthreads();
  if (tid < 5) {
    double s = 0.0;
    for (int q = 0; q < nw; ++q) s += sm.red[tid * 8 + q];
    A.obs_spart[((int64_t(chain) * A.nchunks + chunk) * A.ntiles + tile) * 5 + tid] = s;
  }
}

// Transverse part of the kinetic term for kin_dims = 3, the full Laplacian
// (the reference keeps p_x^2 only, Propagator_without_cutoff.cu:196).  The
// spectral second derivative is a sum of 1-D ones,
//   K = K_x + K_y + K_z,  K_d = (1/a_x) F_d^-1[-p_d^2 F_d psi] / n_d,
// so the x part stays in the row transform of the drift kernels and this
// kernel adds K_d along one axis d in {y, z}, for both fields of the current
// buffer, into kt (unscaled; the drift kernels weight it by a_x f eps).  A
// workgroup transforms, for one tau and one position on the other transverse
// axis, the 4 components of W consecutive x-lines along d: 4W lines of L.
// Site of line position l: ((tau R + base*bmul + l*rstride) X + x0 + xl).
template <typename T>
struct TransArgs {
  const C2<T> *tw;  // RowFFT twiddles of length L
  const double *p2; // -p_d(k)^2 / L
  int X, W, bmul, rstride, nbase;
  int accumulate;   // add to kt (second axis) instead of overwriting it
  int gate;         // only chains that evaluate observables after this step
};

template <typename T, int L>
__global__ __launch_bounds__(256) void k_transverse(BucketArgs<T> A, TransArgs<T> G) {
  extern __shared__ __align__(16) char smem_raw[];
  const int chain = A.chain0 + blockIdx.z;
  const ChainParams &P = A.prm[chain];
  if (!P.active || !has_kt(A, P)) return;
  const ChainState &st = A.st[chain];
  if (G.gate && !evaluates(A, P, st)) return;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int W = G.W, X = G.X, nxb = X / W;
  const int xb = blockIdx.x % nxb, base = blockIdx.x / nxb, t = blockIdx.y;
  Smem<T, L> sm(smem_raw, 4 * W);
  load_tables<T, L>(sm, G.tw, G.p2, 1.0 / P.a_x, tid, nthr);
  const int par = st.parity;  // current fields
  const int64_t row0 = int64_t(t) * A.R + int64_t(base) * G.bmul;
  for (int f = 0; f < 2; ++f) {
    const C2<T> *src = (f == 0 ? (par ? A.buf1_psi : A.buf0_psi) : (par ? A.buf1_psib : A.buf0_psib)) +
                       chain * A.chain_stride;
    C2<T> *dst = (f == 0 ? A.kt_psi : A.kt_psib) + chain * A.chain_stride;
    for (int e = tid; e < W * L; e += nthr) {
      const int xl = e % W, l = e / W;
      C2<T> v[4];
      load4(src, (row0 + int64_t(l) * G.rstride) * X + xb * W + xl, v);
#pragma unroll
      for (int c = 0; c < 4; ++c) sm.at(c * W + xl, l) = v[c];
    }
    __syncthreads();
    RowFFT<T, L>::kinetic(sm.buf, sm.tw, sm.mult, 4 * W, tid, nthr);
    for (int e = tid; e < W * L; e += nthr) {
      const int xl = e % W, l = e / W;
      const int64_t site = (row0 + int64_t(l) * G.rstride) * X + xb * W + xl;
      C2<T> v[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] = sm.at(c * W + xl, l);
      if (G.accumulate) {
        C2<T> o[4];
        load4(dst, site, o);
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] = cadd(o[c], v[c]);
      }
      store4(dst, site, v);
    }
    __syncthreads();  // the lines are rewritten by the next field
  }
}

// Fixed-order accumulation of the chunk results into the running sums
// (the accumulators of Observables::evaluate, :1274-1306, 1322-1382).
template <typename T>
__global__ __launch_bounds__(256) void k_obs_finish(BucketArgs<T> A) {
  const int chain = A.chain0 + blockIdx.y;
  const ChainParams &P = A.prm[chain];
  if (!P.active) return;
  const ChainState &st = A.st[chain];
  if (!evaluates(A, P, st)) return;
  const int64_t vs = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (vs < A.Vs) {
    // the site's three values of each chunk together, chunks in order
    double s[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < A.nchunks; ++c) {
      const double *o = A.obs_stage + ((int64_t(chain) * A.nchunks + c) * A.Vs + vs) * 3;
#pragma unroll
      for (int a = 0; a < 3; ++a) s[a] += o[a];
    }
    double *F = A.obs_F + (int64_t(chain) * A.Vs + vs) * 3;
#pragma unroll
    for (int a = 0; a < 3; ++a) F[a] += s[a];
  }
  if (blockIdx.x == 0 && threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int n = A.nchunks * A.ntiles;
    const double *sp = A.obs_spart + int64_t(chain) * n * 5;
    for (int q = 0; q < 5; ++q) {
      double s = 0.0;
      for (int i = lane; i < n; i += 64) s += sp[i * 5 + q];
      s = wave_sum(s);
      if (lane == 0) A.obs_acc[chain * 5 + q] += s;
    }
    if (lane == 0) A.obs_count[chain] += 1;
  }
}

// psi_c = sqrt(n0) + sqrt(var) (g_re + i g_im), psibar = conj(psi); g from
// the chain's Philox key, stream 1, step 0, at the reference site index.
template <typename T>
__global__ __launch_bounds__(256) void k_init(BucketArgs<T> A, double n0, double var) {
  const int chain = A.chain0 + blockIdx.y;
  const ChainParams &P = A.prm[chain];
  const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= A.S) return;
  const int par = A.st[chain].parity;
  C2<T> *psi = (par ? A.buf1_psi : A.buf0_psi) + chain * A.chain_stride;
  C2<T> *psib = (par ? A.buf1_psib : A.buf0_psib) + chain * A.chain_stride;
  // internal site s = (t*R + r)*X + x
  const int64_t X = A.S / (int64_t(A.Nt) * A.R);
  const int64_t x = s % X, tr = s / X;
  const int64_t r = tr % A.R, t = tr / A.R;
  const uint32_t site = uint32_t((t * X + x) * A.R + r);
  double g[8];
  normals4(g, P.seed, 1u, 0ull, site, 0u);
  normals4(g + 4, P.seed, 1u, 0ull, site, 1u);
  const double a = sqrt(n0), sd = sqrt(var);
  C2<T> v[4], vb[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    v[c] = {T(a + sd * g[2 * c]), T(sd * g[2 * c + 1])};
    vb[c] = cconj(v[c]);
  }
  store4(psi, s, v);
  store4(psib, s, vb);
}

// ========================================================== host side =====
// ========================================================== host side =====
namespace {

const char *kVersion = "clv 0.2 (gfx950)";
thread_local std::string g_create_error = "";

template <typename T>
BucketArgs<T> make_args(const Bucket &b, int chain0, int eval_every, bool fused = false) {
  BucketArgs<T> A;
  A.buf0_psi = static_cast<C2<T> *>(b.psi[0]);
  A.buf0_psib = static_cast<C2<T> *>(b.psib[0]);
  A.buf1_psi = static_cast<C2<T> *>(b.psi[1]);
  A.buf1_psib = static_cast<C2<T> *>(b.psib[1]);
  A.prm = b.d_prm;
  const int slot = b.epoch & 1;
  A.st = b.d_st + slot * b.nch;
  A.st_next = fused ? b.d_st + (slot ^ 1) * b.nch : A.st;
  A.tw = static_cast<const C2<T> *>(b.d_tw);
  A.p2 = b.d_p2;
  A.kt_psi = b.kt_on ? static_cast<C2<T> *>(b.kt[0]) : nullptr;
  A.kt_psib = b.kt_on ? static_cast<C2<T> *>(b.kt[1]) : nullptr;
  A.part_rd = b.d_part + slot * b.part_slot;
  A.part_wr = fused ? b.d_part + (slot ^ 1) * b.part_slot : A.part_rd;
  A.obs_stage = b.d_stage;
  A.obs_spart = b.d_spart;
  A.obs_F = b.d_F;
  A.obs_acc = b.d_acc;
  A.obs_count = b.d_count;
  A.chain_stride = 4 * b.S;
  A.S = b.S;
  A.Vs = b.Vs;
  A.R = b.R;
  A.Nt = b.Nt;
  A.rows = b.rows;
  A.ntiles = b.ntiles;
  A.nchunks = b.nchunks;
  A.tc = b.tc;
  A.tcs = b.tcs;
  A.fsets = b.fused_sets;
  A.fbatch = b.fused_batch;
  A.fblocks = b.fused_blocks;
  A.tstride = 0;
  A.chain0 = chain0;
  A.eval_every = eval_every;
  return A;
}

// ---- launchers, dispatched on X at run time ----
// which: 0 drift_stats, 1 update, 2 obs, 3 fused, 4 drift stats of the fused
// march's block boundaries
template <typename T, int X, int N1>
void launch_row_kernels(const Bucket &b, hipStream_t s, int which, int chain0, int nch,
                        int eval_every) {
  const dim3 block(b.nthr);
  if (which == 0) {
    BucketArgs<T> A = make_args<T>(b, chain0, eval_every);
    const size_t sh = Smem<T, X, N1>::bytes(b.tcs * b.rows) + sizeof(double) * b.tcs * b.nthr;
    hipLaunchKernelGGL((k_drift_stats<T, X, N1>), dim3(b.ntiles, b.Nt / b.tcs, nch), block, sh, s, A);
  } else if (which == 1) {
    BucketArgs<T> A = make_args<T>(b, chain0, eval_every);
    const size_t sh = Smem<T, X, N1>::bytes(8 * b.rows);
    hipLaunchKernelGGL((k_update<T, X, N1>), dim3(b.ntiles, b.Nt, nch), block, sh, s, A);
  } else if (which == 2) {
    BucketArgs<T> A = make_args<T>(b, chain0, eval_every);
    const size_t sh = Smem<T, X, N1>::bytes(4 * obs_slices(int(sizeof(C2<T>)), X) * b.rows);
    hipLaunchKernelGGL((k_obs<T, X, N1>), dim3(b.ntiles, b.nchunks, nch), block, sh, s, A);
  } else if (which == 3) {
    BucketArgs<T> A = make_args<T>(b, chain0, eval_every, true);
    hipLaunchKernelGGL((k_fused<T, X, N1>), dim3(b.ntiles, nch, b.fused_blocks), block, b.fused_lds, s, A);
  } else {
    // |D_0| partials of the first tau of each block of a tau-blocked fused
    // march (the new fields: launched after the state slots swap)
    BucketArgs<T> A = make_args<T>(b, chain0, eval_every);
    A.tstride = b.Nt / b.fused_blocks;
    const size_t sh = Smem<T, X, N1>::bytes(b.rows) + sizeof(double) * b.nthr;
    hipLaunchKernelGGL((k_drift_stats<T, X, N1>), dim3(b.ntiles, b.fused_blocks, nch), block, sh, s, A);
  }
}

#define CLV_SWITCH_X(X_, CALL)                 \
  switch (X_) {                                \
    case 4: CALL(4); break;                    \
    case 8: CALL(8); break;                    \
    case 16: CALL(16); break;                  \
    case 24: CALL(24); break;                  \
    case 32: CALL(32); break;                  \
    case 48: CALL(48); break;                  \
    case 64: CALL(64); break;                  \
    case 96: CALL(96); break;                  \
    case 128: CALL(128); break;                \
    case 192: CALL(192); break;                \
    case 256: CALL(256); break;                \
    default: break;                            \
  }

// Transverse kinetic term of chains [chain0, chain0+nch), one launch per axis.
template <typename T, int L>
void launch_transverse_axis(const Bucket &b, const Bucket::KtAxis &a, hipStream_t s, int chain0,
                            int nch, int accumulate, int gate, int ev) {
  const BucketArgs<T> A = make_args<T>(b, chain0, ev);
  TransArgs<T> G;
  G.tw = static_cast<const C2<T> *>(a.tw);
  G.p2 = a.p2;
  G.X = b.X;
  G.W = a.W;
  G.bmul = a.bmul;
  G.rstride = a.rstride;
  G.nbase = a.nbase;
  G.accumulate = accumulate;
  G.gate = gate;
  const size_t sh = Smem<T, L>::bytes(4 * a.W);
  hipLaunchKernelGGL((k_transverse<T, L>), dim3((b.X / a.W) * a.nbase, b.Nt, nch), dim3(256), sh, s,
                     A, G);
}

template <typename T>
void dispatch_transverse(const Bucket &b, hipStream_t s, int chain0, int nch, int gate, int ev) {
  for (size_t i = 0; i < b.kt_axes.size(); ++i) {
    const Bucket::KtAxis &a = b.kt_axes[i];
    const int acc = i > 0 ? 1 : 0;
#define CLV_TR(LL) launch_transverse_axis<T, LL>(b, a, s, chain0, nch, acc, gate, ev)
    CLV_SWITCH_X(a.L, CLV_TR)
#undef CLV_TR
  }
}

// The row-transform splits instantiated: (X, N1) with X = N1 * N2, the
// default of Split<X> first, then one alternative per length where the small
// DFTs allow it.
#define CLV_SWITCH_SPLIT(X_, N1_, CALL)                              \
  switch ((X_) * 1000 + (N1_)) {                                    \
    case 4002: CALL(4, 2); break;                                   \
    case 8004: CALL(8, 4); break;                                   \
    case 16004: CALL(16, 4); break;                                 \
    case 16008: CALL(16, 8); break;                                 \
    case 24006: CALL(24, 6); break;                                 \
    case 24004: CALL(24, 4); break;                                 \
    case 32008: CALL(32, 8); break;                                 \
    case 32004: CALL(32, 4); break;                                 \
    case 48008: CALL(48, 8); break;                                 \
    case 48006: CALL(48, 6); break;                                 \
    case 64008: CALL(64, 8); break;                                 \
    case 64016: CALL(64, 16); break;                                \
    case 96012: CALL(96, 12); break;                                \
    case 96008: CALL(96, 8); break;                                 \
    case 128016: CALL(128, 16); break;                              \
    case 128008: CALL(128, 8); break;                               \
    case 192016: CALL(192, 16); break;                              \
    case 192012: CALL(192, 12); break;                              \
    case 256016: CALL(256, 16); break;                              \
    default: break;                                                 \
  }

template <typename T>
void dispatch_row(const Bucket &b, hipStream_t s, int which, int chain0, int nch, int ev) {
#define CLV_ROW(XX, NN) launch_row_kernels<T, XX, NN>(b, s, which, chain0, nch, ev)
  CLV_SWITCH_SPLIT(b.X, b.n1, CLV_ROW)
#undef CLV_ROW
}

template <typename T>
void launch_controller(const Bucket &b, hipStream_t s, int chain0, int nch, int mode, double dt,
                       int clamp) {
  BucketArgs<T> A = make_args<T>(b, chain0, 0);
  hipLaunchKernelGGL((k_controller<T>), dim3(nch), dim3(64), 0, s, A, mode, dt, clamp);
}

template <typename T>
void launch_obs_finish(const Bucket &b, hipStream_t s, int chain0, int nch, int ev) {
  BucketArgs<T> A = make_args<T>(b, chain0, ev);
  const int gx = int((b.Vs + 255) / 256);
  hipLaunchKernelGGL((k_obs_finish<T>), dim3(gx, nch), dim3(256), 0, s, A);
}

template <typename T>
void launch_init(const Bucket &b, hipStream_t s, int chain0, double n0, double var) {
  BucketArgs<T> A = make_args<T>(b, chain0, 0);
  const int gx = int((b.S + 255) / 256);
  hipLaunchKernelGGL((k_init<T>), dim3(gx, 1), dim3(256), 0, s, A, n0, var);
}

hipEvent_t take_event(clv_ctx *c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch `fn` on the context's stream; with profiling on, bracket it with
// events of kind `kind`.
template <class F>
void timed(clv_ctx *c, int kind, F &&fn) {
  if (!c->profiling) {
    fn();
    return;
  }
  hipEvent_t a = take_event(c), b = take_event(c);
  if (a) (void)hipEventRecord(a, c->stream);
  fn();
  if (b) (void)hipEventRecord(b, c->stream);
  if (a && b) c->pending.push_back({kind, a, b});
}

void drain_events(clv_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->kt_ms[p.kind] += ms;
      c->kt_n[p.kind] += 1;
    }
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
}

// precision-dispatching wrappers
void row_kernel(clv_ctx *c, const Bucket &b, int which, int chain0, int nch, int ev) {
  static const int kinds[5] = {0, 2, 3, 5, 0};
  timed(c, kinds[which], [&] {
    if (c->precision == CLV_FP32) dispatch_row<float>(b, c->stream, which, chain0, nch, ev);
    else dispatch_row<double>(b, c->stream, which, chain0, nch, ev);
  });
}
// gate: only chains that evaluate observables after this step (eval cadence ev)
void transverse(clv_ctx *c, const Bucket &b, int chain0, int nch, int gate, int ev) {
  if (!b.kt_on) return;
  timed(c, 6, [&] {
    if (c->precision == CLV_FP32) dispatch_transverse<float>(b, c->stream, chain0, nch, gate, ev);
    else dispatch_transverse<double>(b, c->stream, chain0, nch, gate, ev);
  });
}
void controller(clv_ctx *c, const Bucket &b, int chain0, int nch, int mode, double dt, int cl) {
  timed(c, 1, [&] {
    if (c->precision == CLV_FP32) launch_controller<float>(b, c->stream, chain0, nch, mode, dt, cl);
    else launch_controller<double>(b, c->stream, chain0, nch, mode, dt, cl);
  });
}
void obs_finish(clv_ctx *c, const Bucket &b, int chain0, int nch, int ev) {
  timed(c, 4, [&] {
    if (c->precision == CLV_FP32) launch_obs_finish<float>(b, c->stream, chain0, nch, ev);
    else launch_obs_finish<double>(b, c->stream, chain0, nch, ev);
  });
}

// N2 of the default split of the two-level row transform (Split<X> in clv_device.h)
int default_n2(int X) {
  switch (X) {
    case 4: case 8: return 2;
    case 16: case 24: case 32: return 4;
    case 48: return 6;
    case 64: case 96: case 128: return 8;
    case 192: return 12;
    default: return 16;
  }
}
int default_n1(int X) { return X / default_n2(X); }

// Splits N1 of length X instantiated in CLV_SWITCH_SPLIT, the default first.
std::vector<int> split_options(int X) {
  std::vector<int> o = {default_n1(X)};
  switch (X) {
    case 16: o.push_back(8); break;
    case 24: o.push_back(4); break;
    case 32: o.push_back(4); break;
    case 48: o.push_back(6); break;
    case 64: o.push_back(16); break;
    case 96: o.push_back(8); break;
    case 128: o.push_back(8); break;
    case 192: o.push_back(12); break;
    default: break;
  }
  return o;
}

// LDS bytes of the fused kernel with `sets` line sets of (8 + batch) row sets
// each, plus the parked non-kinetic D_0 parts of batch - 1 taus per thread
// and |D_0| of max(batch, 2) taus per thread for the partials.
size_t fused_lds_bytes(const Bucket &b, int sets, int batch, int n1) {
  return size_t(b.esz) * tw_entries(b.X, n1) + size_t(b.esz / 2) * (b.X + (b.X & 1)) +
         size_t(b.esz) * (8 + batch) * sets * b.rows * row_pad(b.X, b.X / n1) +
         sizeof(double) * 64 + size_t(b.esz) * (batch - 1) * b.nthr +
         sizeof(double) * (batch < 2 ? 2 : batch) * b.nthr;
}
size_t fused_lds_bytes(const Bucket &b, int sets, int batch) { return fused_lds_bytes(b, sets, batch, b.n1); }

// LDS bytes of the two-pass update kernel (8 row sets) with `rows` rows per
// tile and split n1: the largest of the two-pass and observable kernels.
size_t update_lds_bytes(const Bucket &b, int rows, int n1) {
  return size_t(b.esz) * tw_entries(b.X, n1) + size_t(b.esz / 2) * (b.X + (b.X & 1)) +
         size_t(b.esz) * 8 * rows * row_pad(b.X, b.X / n1) + sizeof(double) * 64;
}

// rows per workgroup: the largest divisor of R with X*rows <= max_sites
int pick_rows(int X, int R, int max_sites) {
  int best = 1;
  for (int q = 1; q <= R && q * X <= max_sites; ++q)
    if (R % q == 0) best = q;
  return best;
}

// Sites per workgroup tile.  Fixed: the tile sets the partial-sum
// granularity, so results must not depend on anything but the shape.
constexpr int kTileSites = 256;

int dalloc(clv_ctx *ctx, Bucket &b, void **p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();  // not left for a later launch check to report
    return fail(ctx, CLV_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                        "): " + hipGetErrorString(e));
  }
  b.allocs.push_back(*p);
  b.bytes += int64_t(bytes);
  ctx->bytes += int64_t(bytes);
  return CLV_OK;
}

int check_chain(clv_ctx *ctx, int chain) {
  if (!ctx) return CLV_ERR_ARG;
  if (chain < 0 || chain >= int(ctx->where.size()))
    return fail(ctx, CLV_ERR_NOTFOUND, "chain index " + std::to_string(chain) + " out of range");
  return CLV_OK;
}

// Does any active chain of the bucket evaluate after its next step?
bool bucket_needs_eval(const Bucket &b, int ev) {
  if (ev == 0) return false;
  for (int i = 0; i < b.nch; ++i) {
    if (!b.host_active[i]) continue;
    const int e = ev < 0 ? b.host_eval[i] : ev;
    if (ev < 0 && b.host_steps[i] < b.host_eval_from[i]) continue;
    if (e > 0 && b.host_steps[i] % e == 0) return true;
  }
  return false;
}

}  // namespace

// ================================================ internal helpers (clvi) ==
namespace clvi {

int fail(clv_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  else g_create_error = msg;
  return code;
}

bool supported_x(int X) {
  switch (X) {
    case 4: case 8: case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192:
    case 256:
      return true;
    default:
      return false;
  }
}

int64_t chain_bytes(const clv_shape &s, int precision) {
  const int64_t esz = precision == CLV_FP32 ? 8 : 16;
  const int64_t S = int64_t(s.nx) * s.ny * s.nz * s.nt;
  const int64_t Vs = int64_t(s.nx) * s.ny * s.nz;
  const int64_t R = int64_t(s.ny) * s.nz;
  // An upper bound of what setup_bucket allocates per chain: 2 ping-pong
  // copies of psi and psibar, F accumulators and their per-chunk staging
  // (at most Nt chunks), the two slots of (max, sum) partials and the
  // observable partials (at most R tiles per tau), parameters and state.
  return 4 * S * 4 * esz + 3 * Vs * 8 * (1 + s.nt) + 72 * int64_t(s.nt) * R + 4096;
}

ChainParams to_device_params(const clv_params &p, int eval_every, int active, int64_t eval_from) {
  ChainParams d;
  d.mu = p.mu; d.P = p.P; d.G = p.G; d.c1 = p.c1; d.Q = p.Q; d.eps = p.eps; d.a_x = p.a_x;
  d.DT = p.DT; d.dt_min = p.dt_min; d.dt_max = p.dt_max; d.alpha = p.alpha; d.factor = p.factor;
  d.seed = p.seed; d.conj_drift_mode = p.conj_drift_mode; d.noise_on = p.noise_on; d.active = active;
  d.kin_dims = p.kin_dims;
  d.eval_every = eval_every;
  d.pad = 0;
  d.eval_from = eval_from;
  return d;
}

// Row-transform tables of lines of length L, computed in FP64 on the host:
// twiddles in the order RowFFT reads them ([INV + m2*N1 + k1] = exp(+2 pi i m2 k1 / L),
// and without mid_twiddles [k1*N2 + n2] = exp(-2 pi i n2 k1 / L), INV = L) in the
// field precision, and
// -p(k)^2 / L with p of compute_kinetic_part (Propagator_without_cutoff.cu:169-176).
void row_tables_host(int L, int n1, int precision, std::vector<char> *tw_bytes, std::vector<double> *p2) {
  const int n1s = n1, n2s = L / n1;
  const bool mid = mid_twiddles(n2s);
  const int inv = mid ? 0 : L;
  std::vector<double> tw(2 * tw_entries(L, n1));
  for (int k1 = 0; k1 < n1s; ++k1)
    for (int n2 = 0; n2 < n2s; ++n2) {
      const double a = 6.283185307179586 * double((n2 * k1) % L) / double(L);
      if (!mid) {
        tw[2 * (k1 * n2s + n2)] = std::cos(a);
        tw[2 * (k1 * n2s + n2) + 1] = -std::sin(a);
      }
      tw[2 * (inv + n2 * n1s + k1)] = std::cos(a);
      tw[2 * (inv + n2 * n1s + k1) + 1] = std::sin(a);
    }
  p2->assign(L, 0.0);
  for (int k = 0; k < L; ++k) {
    const double p = (k < L / 2 || L == 1) ? 6.283185307179586 * double(k) / double(L)
                                           : -6.283185307179586 * double(L - k) / double(L);
    (*p2)[k] = -(p * p) / double(L);
  }
  if (precision == CLV_FP32) {
    std::vector<float> twf(tw.begin(), tw.end());
    tw_bytes->assign(reinterpret_cast<const char *>(twf.data()),
                     reinterpret_cast<const char *>(twf.data() + twf.size()));
  } else {
    tw_bytes->assign(reinterpret_cast<const char *>(tw.data()),
                     reinterpret_cast<const char *>(tw.data() + tw.size()));
  }
}

int upload_row_tables(clv_ctx *ctx, int L, int n1, void *d_tw, double *d_p2) {
  std::vector<char> tw;
  std::vector<double> p2;
  row_tables_host(L, n1, ctx->precision, &tw, &p2);
  HIPCHK(ctx, hipMemcpy(d_p2, p2.data(), sizeof(double) * L, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(d_tw, tw.data(), tw.size(), hipMemcpyHostToDevice));
  return CLV_OK;
}

// kin_dims = 3 on this bucket: buffers of the transverse kinetic term and the
// line tables of the y and z axes longer than 1.
int enable_kt(clv_ctx *ctx, Bucket &b) {
  if (b.kt_on) return CLV_OK;
  const int Y = b.shape.ny, Z = b.shape.nz;
  std::vector<Bucket::KtAxis> axes;
  if (Y > 1) axes.push_back({Y, 1, Z, Z, 1});         // lines along y: r = z + y*Z
  if (Z > 1) axes.push_back({Z, Z, 1, Y, 1});         // lines along z: r = y*Z + z
  if (axes.empty()) return CLV_OK;                     // 1-D: p_y = p_z = 0
  const size_t field = size_t(b.nch) * size_t(b.S) * 4 * b.esz;
  int rc;
  for (int f = 0; f < 2; ++f)
    if ((rc = dalloc(ctx, b, &b.kt[f], field))) return rc;
  const int budget = (b.esz == 8) ? 4096 : 2048;       // complex values per workgroup
  for (Bucket::KtAxis &a : axes) {
    if ((rc = dalloc(ctx, b, &a.tw, 2 * b.esz * a.L))) return rc;
    if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&a.p2), sizeof(double) * a.L))) return rc;
    if ((rc = upload_row_tables(ctx, a.L, default_n1(a.L), a.tw, a.p2))) return rc;
    a.W = 1;  // x-lines per workgroup: the largest divisor of X within the budget
    for (int w = 1; w <= b.X; ++w)
      if (b.X % w == 0 && 4 * w * a.L <= budget) a.W = w;
  }
  b.kt_axes = axes;
  b.kt_on = true;
  drop_graphs(ctx);
  return CLV_OK;
}

// Tiling of a bucket of b.shape (rows and tiles per tau-slice, threads per
// workgroup, observable tau-chunks, drift-statistics tau slices, fused kernel
// options and schedule); no device work.
void plan_bucket(Bucket &b, int precision) {
  const clv_shape &sh = b.shape;
  b.X = sh.nx;
  b.R = sh.ny * sh.nz;
  b.Nt = sh.nt;
  b.Vs = int64_t(sh.nx) * sh.ny * sh.nz;
  b.S = b.Vs * sh.nt;
  b.esz = (precision == CLV_FP32) ? 8 : 16;
  b.n1 = default_n1(b.X);
  // FP64 rows of up to 64 sites in tiles of 128 sites (two waves): at two
  // waves/SIMD the smaller workgroups wait less at their barriers.  Measured
  // at 64x64x1x64: fused step 3% and two-pass step 5-10% faster than in tiles
  // of 256; at 32x32x1x32 1-4%; longer FP64 rows keep 256 (one 128-site row
  // per workgroup made the fused step at 128x128x1x32 16% slower).
  const int tile_sites = (precision == CLV_FP64 && b.X <= 64) ? kTileSites / 2 : kTileSites;
  b.rows = pick_rows(b.X, b.R, std::max(tile_sites, b.X));
  // the update kernel stages 8 lines per row: keep its LDS within 64 KB (the
  // row tile follows the default split only, so no split changes the tiling)
  while (b.rows > 1 && update_lds_bytes(b, b.rows, b.n1) > 65536) {
    int q = b.rows - 1;
    while (b.R % q != 0) --q;
    b.rows = q;
  }
  b.ntiles = b.R / b.rows;
  b.nthr = ((b.X * b.rows + 63) / 64) * 64;
  b.tc = 1;
  for (int q = 1; q <= 8; ++q)
    if (b.Nt % q == 0) b.tc = q;
  b.nchunks = b.Nt / b.tc;
  b.tcs = 1;  // tau slices per drift-stats workgroup: fill the transform batch
  for (int q = 1; q <= 4 && q * b.rows * b.X <= 1024; ++q)
    if (b.Nt % q == 0) b.tcs = q;
  b.part_slot = size_t(2) * b.nch * b.Nt * b.ntiles;
  b.fused_sets = b.fused_sets_max = (fused_lds_bytes(b, 2, 1) <= 65536) ? 2 : 1;
  b.fused_batch = 1;
  // Untuned default: one line set with K_0 lines in batches of 4 (else 2)
  // taus.  Measured on MI355X it is the fastest option or within 3% of it
  // from 16^3 to 256^2 rows (up to 16% ahead of unbatched).
  for (int B = 4; B >= 2; B /= 2)
    if (b.fused_batch == 1 && B <= b.Nt && fused_lds_bytes(b, 1, B) <= 65536) {
      b.fused_sets = 1;
      b.fused_batch = B;
    }
  b.fused_lds = fused_lds_bytes(b, b.fused_sets, b.fused_batch);
  if (b.fused_lds > 65536) b.fused_lds = 0;
  auto_plan(b, b.nch);
}

// Switch bucket b to row split n1 (0: the default): new twiddle tables, the
// fused kernel's LDS options recomputed (fewer line sets or a smaller K_0
// batch where the new line padding needs it), the fused schedule dropped
// where it no longer fits.  Results change at the rounding level only (the
// transform's arithmetic is regrouped), as a different cuFFT plan would.
int set_row_split(clv_ctx *ctx, Bucket &b, int n1) {
  if (n1 == 0) n1 = default_n1(b.X);
  bool known = false;
  for (int o : split_options(b.X)) known |= (o == n1);
  if (!known)
    return fail(ctx, CLV_ERR_ARG, "row split N1=" + std::to_string(n1) + " is not available for X=" +
                                      std::to_string(b.X));
  if (update_lds_bytes(b, b.rows, n1) > 65536)
    return fail(ctx, CLV_ERR_ARG, "row split N1=" + std::to_string(n1) + " does not fit LDS for shape " +
                                      std::to_string(b.X) + "x" + std::to_string(b.R) + "x" + std::to_string(b.Nt));
  if (n1 == b.n1) return CLV_OK;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // the tables may be in use
  int rc = upload_row_tables(ctx, b.X, n1, b.d_tw, b.d_p2);
  if (rc) return rc;
  b.n1 = n1;
  b.fused_sets_max = (fused_lds_bytes(b, 2, 1) <= 65536) ? 2 : 1;
  if (b.fused_sets > b.fused_sets_max) b.fused_sets = b.fused_sets_max;
  while (b.fused_batch > 1 && fused_lds_bytes(b, b.fused_sets, b.fused_batch) > 65536) b.fused_batch /= 2;
  b.fused_lds = fused_lds_bytes(b, b.fused_sets, b.fused_batch);
  if (b.fused_lds > 65536) b.fused_lds = 0;
  if (b.schedule == 2 && b.fused_lds == 0) b.schedule = 1;
  b.primed = false;
  drop_graphs(ctx);
  return CLV_OK;
}

int setup_bucket(clv_ctx *ctx, Bucket &b) {
  plan_bucket(b, ctx->precision);
  const size_t field = size_t(b.nch) * size_t(b.S) * 4 * b.esz;
  int rc;
  for (int k = 0; k < 2; ++k) {
    if ((rc = dalloc(ctx, b, &b.psi[k], field))) return rc;
    if ((rc = dalloc(ctx, b, &b.psib[k], field))) return rc;
  }
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_prm), sizeof(ChainParams) * b.nch))) return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_st), 2 * sizeof(ChainState) * b.nch))) return rc;
  if ((rc = dalloc(ctx, b, &b.d_tw, 2 * b.esz * b.X))) return rc;  // any split
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_p2), sizeof(double) * b.X))) return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_part), sizeof(double) * 2 * b.part_slot)))
    return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_stage),
                   sizeof(double) * 3 * b.nch * b.nchunks * b.Vs)))
    return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_spart),
                   sizeof(double) * 5 * b.nch * b.nchunks * b.ntiles)))
    return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_F), sizeof(double) * 3 * b.nch * b.Vs))) return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_acc), sizeof(double) * 5 * b.nch))) return rc;
  if ((rc = dalloc(ctx, b, reinterpret_cast<void **>(&b.d_count), sizeof(int64_t) * b.nch))) return rc;
  if ((rc = upload_row_tables(ctx, b.X, b.n1, b.d_tw, b.d_p2))) return rc;
  // Zeroed on the context's own stream and waited for: a null-stream memset
  // is not ordered with the context's non-blocking stream, so a field
  // initialisation launched right after could run first and be overwritten.
  HIPCHK(ctx, hipMemsetAsync(b.d_st, 0, 2 * sizeof(ChainState) * b.nch, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_prm, 0, sizeof(ChainParams) * b.nch, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_F, 0, sizeof(double) * 3 * b.nch * b.Vs, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_acc, 0, sizeof(double) * 5 * b.nch, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_count, 0, sizeof(int64_t) * b.nch, ctx->stream));
  for (int k = 0; k < 2; ++k) {
    HIPCHK(ctx, hipMemsetAsync(b.psi[k], 0, field, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(b.psib[k], 0, field, ctx->stream));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  b.host_steps.assign(b.nch, 0);
  b.host_eval.assign(b.nch, 0);
  b.host_eval_from.assign(b.nch, 0);
  b.host_active.assign(b.nch, 1);
  return CLV_OK;
}

void free_bucket(clv_ctx *ctx, Bucket &b) {
  for (void *p : b.allocs) (void)hipFree(p);
  b.allocs.clear();
  ctx->bytes -= b.bytes;
  b.bytes = 0;
}

void count_step(Bucket &b, int64_t n) {
  for (int i = 0; i < b.nch; ++i)
    if (b.host_active[i]) b.host_steps[i] += n;
}

void step_bucket(clv_ctx *ctx, Bucket &b, int eval_every) {
  if (uses_fused(b)) {
    if (!b.primed) row_kernel(ctx, b, 0, 0, b.nch, 0);  // partials of the current fields
    row_kernel(ctx, b, 3, 0, b.nch, 0);
    b.epoch ^= 1;  // new state and next step's partials live in the other slot
    if (b.fused_blocks > 1) row_kernel(ctx, b, 4, 0, b.nch, 0);
    b.primed = true;
  } else {
    transverse(ctx, b, 0, b.nch, 0, 0);
    row_kernel(ctx, b, 0, 0, b.nch, 0);
    controller(ctx, b, 0, b.nch, 0, 0.0, 0);
    row_kernel(ctx, b, 1, 0, b.nch, 0);
    b.primed = false;
  }
  const bool ev = bucket_needs_eval(b, eval_every);
  count_step(b, 1);
  if (ev) {
    transverse(ctx, b, 0, b.nch, 1, eval_every);
    row_kernel(ctx, b, 2, 0, b.nch, eval_every);
    obs_finish(ctx, b, 0, b.nch, eval_every);
  }
}

void enqueue_step_gated(clv_ctx *ctx, Bucket &b, int eval_every, bool obs) {
  if (uses_fused(b)) {
    row_kernel(ctx, b, 3, 0, b.nch, 0);
    b.epoch ^= 1;
    if (b.fused_blocks > 1) row_kernel(ctx, b, 4, 0, b.nch, 0);
  } else {
    transverse(ctx, b, 0, b.nch, 0, 0);
    row_kernel(ctx, b, 0, 0, b.nch, 0);
    controller(ctx, b, 0, b.nch, 0, 0.0, 0);
    row_kernel(ctx, b, 1, 0, b.nch, 0);
  }
  if (eval_every != 0 && obs) {
    transverse(ctx, b, 0, b.nch, 1, eval_every);
    row_kernel(ctx, b, 2, 0, b.nch, eval_every);
    obs_finish(ctx, b, 0, b.nch, eval_every);
  }
}

void drop_graphs(clv_ctx *ctx) {
  for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
}

int read_state(clv_ctx *ctx, const Bucket &b, int local, ChainState *out) {
  HIPCHK(ctx, hipMemcpyAsync(out, b.d_st + (b.epoch & 1) * b.nch + local, sizeof(ChainState),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CLV_OK;
}

void stats_from_state(const ChainState &st, clv_step_stats *out) {
  out->dt = st.dt;
  out->max_drift = st.max_drift;
  out->mean_drift = st.mean_drift;
  out->t_langevin = st.t_langevin;
  out->steps = st.steps;
  out->clamp_low = st.clamp_low;
  out->clamp_high = st.clamp_high;
  out->nonfinite = st.nonfinite;
  out->reserved = 0;
}

// Give slot `local` a fresh chain: parameters, zero controller state and
// accumulators (both state slots), Gaussian initial fields, active.  All
// operations are ordered on the context's stream, so a bucket that is being
// stepped can admit a chain between two of its steps.
int chain_admit(clv_ctx *ctx, Bucket &b, int local, const clv_params &p, int eval_every,
                int64_t eval_from, double n0, double var) {
  ChainParams d = to_device_params(p, eval_every, 1, eval_from);
  HIPCHK(ctx, hipMemcpyAsync(b.d_prm + local, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream));
  ChainState z;
  std::memset(&z, 0, sizeof(z));
  for (int slot = 0; slot < 2; ++slot)
    HIPCHK(ctx, hipMemcpyAsync(b.d_st + slot * b.nch + local, &z, sizeof(z), hipMemcpyHostToDevice,
                               ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_F + 3 * b.Vs * local, 0, sizeof(double) * 3 * b.Vs, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_acc + 5 * local, 0, sizeof(double) * 5, ctx->stream));
  HIPCHK(ctx, hipMemsetAsync(b.d_count + local, 0, sizeof(int64_t), ctx->stream));
  if (ctx->precision == CLV_FP32) launch_init<float>(b, ctx->stream, local, n0, var);
  else launch_init<double>(b, ctx->stream, local, n0, var);
  HIPCHK(ctx, hipGetLastError());
  // the host records must stay valid until the copies ran
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  b.host_steps[local] = 0;
  b.host_eval[local] = eval_every;
  b.host_eval_from[local] = eval_from;
  b.host_active[local] = 1;
  b.primed = false;
  return CLV_OK;
}

void chain_evaluate(clv_ctx *ctx, Bucket &b, int local) {
  transverse(ctx, b, local, 1, 0, 0);
  row_kernel(ctx, b, 2, local, 1, 0);
  obs_finish(ctx, b, local, 1, 0);
}

int chain_set_active(clv_ctx *ctx, Bucket &b, int local, const clv_params &p, int eval_every,
                     bool active) {
  ChainParams d = to_device_params(p, eval_every, active ? 1 : 0, b.host_eval_from[local]);
  HIPCHK(ctx, hipMemcpyAsync(b.d_prm + local, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  b.host_active[local] = active ? 1 : 0;
  return CLV_OK;
}

int chains_deactivate(clv_ctx *ctx, Bucket &b, const std::vector<int> &locals) {
  if (locals.empty()) return CLV_OK;
  // zeroed records: active = 0, which is all the kernels read of a free slot
  std::vector<ChainParams> z(locals.size());
  std::memset(z.data(), 0, sizeof(ChainParams) * z.size());
  for (size_t i = 0; i < locals.size(); ++i)
    HIPCHK(ctx, hipMemcpyAsync(b.d_prm + locals[i], &z[i], sizeof(ChainParams), hipMemcpyHostToDevice,
                               ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // z must outlive the copies
  for (int l : locals) b.host_active[l] = 0;
  return CLV_OK;
}

int chain_read_observables(clv_ctx *ctx, Bucket &b, int local, clv_observables *out, double *Fx,
                           double *Fy, double *Fz) {
  double acc[5];
  int64_t count = 0;
  std::vector<double> F(3 * b.Vs);
  HIPCHK(ctx, hipMemcpyAsync(acc, b.d_acc + 5 * local, sizeof(acc), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(&count, b.d_count + local, sizeof(count), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipMemcpyAsync(F.data(), b.d_F + 3 * b.Vs * local, sizeof(double) * F.size(),
                             hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::memset(out, 0, sizeof(*out));
  out->count = count;
  if (count == 0) return CLV_OK;
  const double Nt = double(b.Nt), cnt = double(count), Vs = double(b.Vs);
  double tot = 0.0;
  for (int c = 0; c < 4; ++c) {
    // real-space sums; the reference's k-space sums are V_s times larger and
    // write_n_* divides by V_s (Observables_without_cutoff.cu:1527-1547)
    out->n_comp[c] = acc[c] / Nt / cnt;
    tot += acc[c];
  }
  out->n_tot = tot / Nt / cnt;
  out->energy = acc[4] / Nt / cnt;
  double *dst[3] = {Fx, Fy, Fz};
  double tots[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (int64_t s = 0; s < b.Vs; ++s) {
      const double v = F[3 * s + a] / Nt / cnt;
      tots[a] += v;
      if (dst[a]) dst[a][s] = v;
    }
  out->Fx_tot = tots[0];
  out->Fy_tot = tots[1];
  out->Fz_tot = tots[2];
  out->density = out->n_tot / Vs;
  out->Sz = out->Fz_tot / Vs;
  return CLV_OK;
}

}  // namespace clvi

// =============================================================== C ABI ====
extern "C" {

const char *clv_version(void) { return kVersion; }

const char *clv_last_error(const clv_ctx *ctx) {
  if (ctx) return ctx->err.c_str();
  return g_create_error.c_str();
}

void clv_default_params(clv_params *p) {
  std::memset(p, 0, sizeof(*p));
  p->mu = 0.0;
  p->P = 0.0;
  p->G = 0.1;      // Parameters.h:15
  p->c1 = 0.0;
  p->Q = 0.0;      // Parameters.h:24
  p->eps = 0.01;   // Parameters.h:13
  p->a_x = 1.0;    // Parameters.h:14
  p->DT = 0.002;   // Parameters.h:29
  p->dt_min = 1e-8;
  p->dt_max = 1e-2;
  p->alpha = 0.9;
  p->factor = 1.0;
  p->seed = 0;
  p->kin_dims = 1;
  p->conj_drift_mode = CLV_CONJ_DRIFT_REFERENCE;
  p->noise_on = 1;
}

double clv_compute_new_dt(double mean_drift, double max_drift, const clv_params *p,
                          int32_t *clamp) {
  double dt = (max_drift <= 0.0) ? p->DT : (mean_drift / max_drift) * p->DT;
  int32_t c = 0;
  if (dt < p->dt_min) { dt = p->dt_min; c = -1; }
  if (dt > p->dt_max) { dt = p->dt_max; c = 1; }
  if (clamp) *clamp = c;
  return (1 - p->alpha) * p->DT + p->alpha * dt;
}

int clv_create(const clv_shape *shapes, int32_t nchains, int32_t precision, int32_t device,
               clv_ctx **out) {
  if (!out || !shapes || nchains <= 0) return fail(nullptr, CLV_ERR_ARG, "clv_create: bad arguments");
  *out = nullptr;
  if (precision != CLV_FP32 && precision != CLV_FP64)
    return fail(nullptr, CLV_ERR_ARG, "clv_create: precision must be CLV_FP32 or CLV_FP64");
  for (int i = 0; i < nchains; ++i) {
    const clv_shape &s = shapes[i];
    if (s.nx <= 0 || s.ny <= 0 || s.nz <= 0 || s.nt < 2)
      return fail(nullptr, CLV_ERR_ARG, "clv_create: chain " + std::to_string(i) + " has a bad shape");
    if (!supported_x(s.nx))
      return fail(nullptr, CLV_ERR_ARG,
                  "clv_create: X=" + std::to_string(s.nx) +
                      " not supported (row transform lengths 4,8,16,24,32,48,64,96,128,192,256)");
    if (int64_t(s.nx) * s.ny * s.nz * s.nt >= (int64_t(1) << 31))
      return fail(nullptr, CLV_ERR_ARG, "clv_create: lattice too large");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, CLV_ERR_HIP, "clv_create: no HIP device visible");
  if (device < 0 || device >= ndev) return fail(nullptr, CLV_ERR_ARG, "clv_create: bad device index");
  std::unique_ptr<clv_ctx> ctx(new clv_ctx());
  ctx->device = device;
  ctx->precision = precision;
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, CLV_ERR_HIP, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(nullptr, CLV_ERR_HIP, "hipStreamCreate failed");
  // group chains into buckets of equal shape
  std::map<std::tuple<int, int, int, int>, int> key2b;
  ctx->where.resize(nchains);
  for (int i = 0; i < nchains; ++i) {
    const clv_shape &s = shapes[i];
    auto key = std::make_tuple(s.nx, s.ny, s.nz, s.nt);
    auto it = key2b.find(key);
    int bi;
    if (it == key2b.end()) {
      bi = int(ctx->buckets.size());
      key2b[key] = bi;
      ctx->buckets.emplace_back();
      ctx->buckets.back().shape = s;
    } else {
      bi = it->second;
    }
    Bucket &b = ctx->buckets[bi];
    ctx->where[i] = {bi, b.nch};
    b.ids.push_back(i);
    b.nch += 1;
  }
  for (Bucket &b : ctx->buckets) {
    int rc = setup_bucket(ctx.get(), b);
    if (rc) {
      g_create_error = ctx->err;
      clv_destroy(ctx.release());
      return rc;
    }
  }
  ctx->params.resize(nchains);
  for (int i = 0; i < nchains; ++i) {
    clv_default_params(&ctx->params[i]);
    ctx->params[i].seed = uint64_t(i);
    int rc = clv_set_params(ctx.get(), i, &ctx->params[i]);
    if (rc) {
      g_create_error = ctx->err;
      clv_destroy(ctx.release());
      return rc;
    }
  }
  *out = ctx.release();
  return CLV_OK;
}

int clv_destroy(clv_ctx *ctx) {
  if (!ctx) return CLV_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drain_events(ctx);
  drop_graphs(ctx);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  for (Bucket &b : ctx->buckets) free_bucket(ctx, b);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CLV_OK;
}

int clv_num_chains(const clv_ctx *ctx) { return ctx ? int(ctx->where.size()) : 0; }
int64_t clv_device_bytes(const clv_ctx *ctx) { return ctx ? ctx->bytes : 0; }

int clv_set_params(clv_ctx *ctx, int32_t chain, const clv_params *p) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!p) return fail(ctx, CLV_ERR_ARG, "clv_set_params: NULL params");
  if (p->kin_dims != 1 && p->kin_dims != 3)
    return fail(ctx, CLV_ERR_ARG,
                "clv_set_params: kin_dims must be 1 (p^2 = p_x^2, the reference's kinetic term) "
                "or 3 (p_x^2 + p_y^2 + p_z^2)");
  if (p->conj_drift_mode != CLV_CONJ_DRIFT_REFERENCE && p->conj_drift_mode != CLV_CONJ_DRIFT_ACTION)
    return fail(ctx, CLV_ERR_ARG, "clv_set_params: bad conj_drift_mode");
  if (!(p->a_x > 0.0) || !(p->DT > 0.0) || !(p->eps > 0.0))
    return fail(ctx, CLV_ERR_ARG, "clv_set_params: a_x, DT and eps must be positive");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  if (p->kin_dims == 3) {
    if ((b.shape.ny > 1 && !supported_x(b.shape.ny)) || (b.shape.nz > 1 && !supported_x(b.shape.nz)))
      return fail(ctx, CLV_ERR_ARG,
                  "clv_set_params: kin_dims = 3 needs Y and Z in {1, 4, 8, 16, 24, 32, 48, 64, 96, "
                  "128, 192, 256}");
    if ((rc = enable_kt(ctx, b))) return rc;
  }
  ctx->params[chain] = *p;
  ChainParams d = to_device_params(*p, b.host_eval[local], b.host_active[local],
                                   b.host_eval_from[local]);
  HIPCHK(ctx, hipMemcpyAsync(b.d_prm + local, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  // The fused kernel left next step's |D_0| partials computed with the old
  // parameters (coefficients and kinetic table scale): recompute them.
  b.primed = false;
  return CLV_OK;
}

int clv_set_fields(clv_ctx *ctx, int32_t chain, const double *psi, const double *psibar) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!psi || !psibar) return fail(ctx, CLV_ERR_ARG, "clv_set_fields: NULL field");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  ChainState st;
  if ((rc = read_state(ctx, b, local, &st))) return rc;
  const int par = st.parity;
  const int X = b.X, R = b.R, Nt = b.Nt;
  const size_t n = size_t(b.S) * 4;
  for (int f = 0; f < 2; ++f) {
    const double *src = f ? psibar : psi;
    std::vector<char> host(n * b.esz);
    for (int c = 0; c < 4; ++c)
      for (int t = 0; t < Nt; ++t)
        for (int x = 0; x < X; ++x)
          for (int r = 0; r < R; ++r) {
            const size_t ref = ((size_t(c) * Nt + t) * X + x) * R + r;
            const size_t in = ((size_t(t) * R + r) * X + x) * 4 + c;
            if (ctx->precision == CLV_FP32) {
              float *d = reinterpret_cast<float *>(host.data());
              d[2 * in] = float(src[2 * ref]);
              d[2 * in + 1] = float(src[2 * ref + 1]);
            } else {
              double *d = reinterpret_cast<double *>(host.data());
              d[2 * in] = src[2 * ref];
              d[2 * in + 1] = src[2 * ref + 1];
            }
          }
    void *dst = static_cast<char *>(f ? b.psib[par] : b.psi[par]) + size_t(local) * n * b.esz;
    HIPCHK(ctx, hipMemcpy(dst, host.data(), n * b.esz, hipMemcpyHostToDevice));
  }
  b.primed = false;
  return CLV_OK;
}

int clv_get_fields(clv_ctx *ctx, int32_t chain, double *psi, double *psibar) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  ChainState st;
  if ((rc = read_state(ctx, b, local, &st))) return rc;
  const int par = st.parity;
  const int X = b.X, R = b.R, Nt = b.Nt;
  const size_t n = size_t(b.S) * 4;
  for (int f = 0; f < 2; ++f) {
    double *dst = f ? psibar : psi;
    if (!dst) continue;
    std::vector<char> host(n * b.esz);
    const void *src = static_cast<const char *>(f ? b.psib[par] : b.psi[par]) + size_t(local) * n * b.esz;
    HIPCHK(ctx, hipMemcpy(host.data(), src, n * b.esz, hipMemcpyDeviceToHost));
    for (int c = 0; c < 4; ++c)
      for (int t = 0; t < Nt; ++t)
        for (int x = 0; x < X; ++x)
          for (int r = 0; r < R; ++r) {
            const size_t ref = ((size_t(c) * Nt + t) * X + x) * R + r;
            const size_t in = ((size_t(t) * R + r) * X + x) * 4 + c;
            if (ctx->precision == CLV_FP32) {
              const float *s = reinterpret_cast<const float *>(host.data());
              dst[2 * ref] = s[2 * in];
              dst[2 * ref + 1] = s[2 * in + 1];
            } else {
              const double *s = reinterpret_cast<const double *>(host.data());
              dst[2 * ref] = s[2 * in];
              dst[2 * ref + 1] = s[2 * in + 1];
            }
          }
  }
  return CLV_OK;
}

int clv_init_fields(clv_ctx *ctx, int32_t chain, double n0, double var) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (n0 < 0 || var < 0) return fail(ctx, CLV_ERR_ARG, "clv_init_fields: n0 and var must be >= 0");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  if (ctx->precision == CLV_FP32) launch_init<float>(b, ctx->stream, local, n0, var);
  else launch_init<double>(b, ctx->stream, local, n0, var);
  b.primed = false;
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CLV_OK;
}

// The step count of every active chain of b when they are all equal, else -1.
static int64_t common_steps(const Bucket &b) {
  int64_t s = -1;
  for (int i = 0; i < b.nch; ++i) {
    if (!b.host_active[i]) continue;
    if (s >= 0 && b.host_steps[i] != s) return -1;
    s = b.host_steps[i];
  }
  return s < 0 ? 0 : s;
}

// Captured graph of an even number of steps of every bucket (even so that the
// fused schedule's alternating state slots return to the captured parity).
// With a uniform cadence (eval_every > 0, at most 32) and every bucket's
// chains at one step count, the graph spans a cadence period and launches
// the observable kernels only after the steps that evaluate (the phase is
// part of the key); otherwise two steps with the observable kernels after
// each, gated per chain on the device.  *steps: steps per replay.
static int graph_for(clv_ctx *ctx, int eval_every, hipGraphExec_t *out, int *steps) {
  bool phased = eval_every > 0 && eval_every <= 32;
  std::vector<int64_t> phase;
  for (const Bucket &b : ctx->buckets) {
    const int64_t s = common_steps(b);
    if (s < 0) phased = false;
    phase.push_back(s < 0 ? 0 : s % (eval_every > 0 ? eval_every : 1));
  }
  const int G = phased ? (eval_every % 2 == 0 ? eval_every : 2 * eval_every) : 2;
  std::string key = std::to_string(eval_every) + (phased ? "p" : "g");
  for (size_t i = 0; i < ctx->buckets.size(); ++i) {
    const Bucket &b = ctx->buckets[i];
    key += "/" + std::to_string(b.schedule) + "." + std::to_string(b.fused_blocks) + ":" +
           std::to_string(b.epoch & 1);
    if (phased) key += ":" + std::to_string(phase[i]);
  }
  *steps = G;
  for (auto &g : ctx->graphs)
    if (g.key == key) {
      *out = g.exec;
      return CLV_OK;
    }
  std::vector<int> epochs;
  for (const Bucket &b : ctx->buckets) epochs.push_back(b.epoch);
  // With several buckets each one's steps go on a branch of their own, forked
  // from the context's stream and joined back, so a replay lets the buckets'
  // kernels overlap (the buckets own disjoint memory; each branch keeps its
  // bucket's steps in order).  Uneven 3-D batch of 256 chains in 9 buckets:
  // 5.22 against 5.56 ms per step, the same results.  The side streams only
  // record the capture and are released after it.
  const hipStream_t origin = ctx->stream;
  const size_t nb = ctx->buckets.size();
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> evs;
  bool fork = nb > 1;
  if (fork) {
    side.assign(nb, nullptr);
    evs.assign(nb + 1, nullptr);
    for (auto &s : side) fork = fork && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
    for (auto &e : evs) fork = fork && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
  }
  auto release = [&] {
    for (hipStream_t s : side)
      if (s) (void)hipStreamDestroy(s);
    for (hipEvent_t e : evs)
      if (e) (void)hipEventDestroy(e);
  };
  if (hipStreamBeginCapture(origin, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    release();
    return fail(ctx, CLV_ERR_HIP, "hipStreamBeginCapture failed");
  }
  if (fork) (void)hipEventRecord(evs[nb], origin);
  for (size_t i = 0; i < nb; ++i) {
    if (fork) {
      (void)hipStreamWaitEvent(side[i], evs[nb], 0);
      ctx->stream = side[i];
    }
    for (int n = 0; n < G; ++n)
      enqueue_step_gated(ctx, ctx->buckets[i], eval_every,
                         !phased || (phase[i] + n) % eval_every == 0);
    if (fork) {
      (void)hipEventRecord(evs[i], side[i]);
      ctx->stream = origin;
      (void)hipStreamWaitEvent(origin, evs[i], 0);
    }
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(origin, &graph);
  release();
  for (size_t i = 0; i < ctx->buckets.size(); ++i) ctx->buckets[i].epoch = epochs[i];
  if (ec != hipSuccess) return fail(ctx, CLV_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(ctx, CLV_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  ctx->graphs.push_back({key, exec});
  *out = exec;
  return CLV_OK;
}

int clv_propagate(clv_ctx *ctx, int64_t nsteps, int32_t eval_every) {
  if (!ctx) return CLV_ERR_ARG;
  if (nsteps < 0) return fail(ctx, CLV_ERR_ARG, "clv_propagate: nsteps < 0");
  (void)hipSetDevice(ctx->device);
  if (!ctx->use_graphs || ctx->profiling || nsteps < 4) {
    for (int64_t n = 0; n < nsteps; ++n)
      for (Bucket &b : ctx->buckets) step_bucket(ctx, b, eval_every);
    HIPCHK(ctx, hipGetLastError());
    return CLV_OK;
  }
  // fused buckets need the partials of their current fields before replay
  for (Bucket &b : ctx->buckets)
    if (uses_fused(b) && !b.primed) {
      row_kernel(ctx, b, 0, 0, b.nch, 0);
      b.primed = true;
    }
  hipGraphExec_t exec = nullptr;
  int G = 2;
  int rc = graph_for(ctx, eval_every, &exec, &G);
  if (rc) return rc;
  const int64_t reps = nsteps / G;
  for (int64_t k = 0; k < reps; ++k) HIPCHK(ctx, hipGraphLaunch(exec, ctx->stream));
  for (Bucket &b : ctx->buckets) {
    count_step(b, G * reps);
    if (!(uses_fused(b))) b.primed = false;
  }
  for (int64_t n = G * reps; n < nsteps; ++n)
    for (Bucket &b : ctx->buckets) step_bucket(ctx, b, eval_every);
  HIPCHK(ctx, hipGetLastError());
  return CLV_OK;
}

int clv_prepare_graph(clv_ctx *ctx, int32_t eval_every) {
  if (!ctx) return CLV_ERR_ARG;
  if (!ctx->use_graphs || ctx->profiling) return CLV_OK;
  (void)hipSetDevice(ctx->device);
  hipGraphExec_t exec = nullptr;
  int G = 2;
  return graph_for(ctx, eval_every, &exec, &G);
}

int clv_drift_stats(clv_ctx *ctx, int32_t chain, double *max_drift, double *mean_drift) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  transverse(ctx, b, local, 1, 0, 0);
  row_kernel(ctx, b, 0, local, 1, 0);
  controller(ctx, b, local, 1, 1, 0.0, 0);
  HIPCHK(ctx, hipGetLastError());
  ChainState st;
  if ((rc = read_state(ctx, b, local, &st))) return rc;
  if (max_drift) *max_drift = st.max_drift;
  if (mean_drift) *mean_drift = st.mean_drift;
  return CLV_OK;
}

int clv_apply_update(clv_ctx *ctx, int32_t chain, double dt) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!(dt >= 0.0)) return fail(ctx, CLV_ERR_ARG, "clv_apply_update: dt must be >= 0");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  transverse(ctx, b, local, 1, 0, 0);  // of the fields this update reads
  controller(ctx, b, local, 1, 2, dt, 0);
  row_kernel(ctx, b, 1, local, 1, 0);
  b.host_steps[local] += 1;
  b.primed = false;
  HIPCHK(ctx, hipGetLastError());
  return CLV_OK;
}

int clv_propagate_plain(clv_ctx *ctx, int32_t chain, int64_t nsteps, int32_t eval_every) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  const clv_params &p = ctx->params[chain];
  ChainState *pinned = nullptr;
  HIPCHK(ctx, hipHostMalloc(reinterpret_cast<void **>(&pinned), sizeof(ChainState)));
  for (int64_t n = 0; n < nsteps; ++n) {
    // drift + max/mean on the device, blocking copy to the host
    transverse(ctx, b, local, 1, 0, 0);
    row_kernel(ctx, b, 0, local, 1, 0);
    controller(ctx, b, local, 1, 1, 0.0, 0);
    hipError_t e = hipMemcpyAsync(pinned, b.d_st + (b.epoch & 1) * b.nch + local, sizeof(ChainState),
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipHostFree(pinned);
      return fail(ctx, CLV_ERR_HIP, std::string("plain step: ") + hipGetErrorString(e));
    }
    // dt on the host (compute_new_dt), passed by value to the update
    int32_t clamp = 0;
    const double dt = clv_compute_new_dt(pinned->mean_drift, pinned->max_drift, &p, &clamp);
    controller(ctx, b, local, 1, 2, dt, clamp);
    row_kernel(ctx, b, 1, local, 1, 0);
    b.primed = false;
    const bool ev = eval_every > 0 && (b.host_steps[local] % eval_every == 0);
    b.host_steps[local] += 1;
    if (ev) {
      transverse(ctx, b, local, 1, 1, eval_every);
      row_kernel(ctx, b, 2, local, 1, eval_every);
      obs_finish(ctx, b, local, 1, eval_every);
    }
  }
  (void)hipHostFree(pinned);
  HIPCHK(ctx, hipGetLastError());
  return CLV_OK;
}

int clv_get_step_stats(clv_ctx *ctx, int32_t chain, clv_step_stats *out) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!out) return fail(ctx, CLV_ERR_ARG, "clv_get_step_stats: NULL out");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  ChainState st;
  if ((rc = read_state(ctx, ctx->buckets[bi], local, &st))) return rc;
  stats_from_state(st, out);
  return CLV_OK;
}

int clv_set_step_stats(clv_ctx *ctx, int32_t chain, const clv_step_stats *in) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!in || in->steps < 0) return fail(ctx, CLV_ERR_ARG, "clv_set_step_stats: bad stats");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  ChainState st;
  if ((rc = read_state(ctx, b, local, &st))) return rc;
  st.dt = in->dt;
  st.max_drift = in->max_drift;
  st.mean_drift = in->mean_drift;
  st.t_langevin = in->t_langevin;
  st.steps = in->steps;
  st.cur_step = in->steps > 0 ? in->steps - 1 : 0;
  st.clamp_low = in->clamp_low;
  st.clamp_high = in->clamp_high;
  st.nonfinite = in->nonfinite;
  HIPCHK(ctx, hipMemcpyAsync(b.d_st + (b.epoch & 1) * b.nch + local, &st, sizeof(st),
                             hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  b.host_steps[local] = in->steps;
  b.primed = false;
  return CLV_OK;
}

int clv_synchronize(clv_ctx *ctx) {
  if (!ctx) return CLV_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CLV_OK;
}

int clv_evaluate(clv_ctx *ctx, int32_t chain) {
  if (!ctx) return CLV_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  if (chain < 0) {
    for (Bucket &b : ctx->buckets) {
      transverse(ctx, b, 0, b.nch, 0, 0);
      row_kernel(ctx, b, 2, 0, b.nch, 0);
      obs_finish(ctx, b, 0, b.nch, 0);
    }
  } else {
    int rc = check_chain(ctx, chain);
    if (rc) return rc;
    const auto [bi, local] = ctx->where[chain];
    Bucket &b = ctx->buckets[bi];
    transverse(ctx, b, local, 1, 0, 0);
    row_kernel(ctx, b, 2, local, 1, 0);
    obs_finish(ctx, b, local, 1, 0);
  }
  HIPCHK(ctx, hipGetLastError());
  return CLV_OK;
}

int clv_read_observables(clv_ctx *ctx, int32_t chain, clv_observables *out, double *Fx,
                         double *Fy, double *Fz) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (!out) return fail(ctx, CLV_ERR_ARG, "clv_read_observables: NULL out");
  (void)hipSetDevice(ctx->device);
  const auto [bi, local] = ctx->where[chain];
  Bucket &b = ctx->buckets[bi];
  return chain_read_observables(ctx, b, local, out, Fx, Fy, Fz);
}

int clv_reset_observables(clv_ctx *ctx, int32_t chain) {
  if (!ctx) return CLV_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  auto reset = [&](Bucket &b, int local, int n) -> int {
    HIPCHK(ctx, hipMemsetAsync(b.d_F + 3 * b.Vs * local, 0, sizeof(double) * 3 * b.Vs * n, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(b.d_acc + 5 * local, 0, sizeof(double) * 5 * n, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(b.d_count + local, 0, sizeof(int64_t) * n, ctx->stream));
    return CLV_OK;
  };
  int rc;
  if (chain < 0) {
    for (Bucket &b : ctx->buckets)
      if ((rc = reset(b, 0, b.nch))) return rc;
  } else {
    if ((rc = check_chain(ctx, chain))) return rc;
    const auto [bi, local] = ctx->where[chain];
    if ((rc = reset(ctx->buckets[bi], local, 1))) return rc;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CLV_OK;
}

}  // extern "C"

// ---- autotuner: schedule per shape bucket, timed on the live state ----
namespace {

std::string device_name(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  return std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + "cu";
}

std::string tune_key(const std::string &dev, const Bucket &b, int precision) {
  std::ostringstream k;
  k << dev << "|" << b.shape.nx << "x" << b.shape.ny << "x" << b.shape.nz << "x" << b.shape.nt
    << "|" << (precision == CLV_FP32 ? "fp32" : "fp64");
  return k.str();
}

// One tuning choice.  Within one row split all of them give identical
// results: the schedules are bit-identical, the reduction tile (tcs tau slices
// per drift-statistics workgroup) does not change the per-(tau, tile)
// partials, the fused kernel's LDS line sets only change its barriers, and
// its K_0 batch only groups the same per-line transforms differently.  The
// row split n1 regroups the transform's arithmetic (rounding-level changes,
// as cufftPlanMany's own plan choice, Lattice.cu:135-156).  The row tile is
// fixed per shape: it sets the partial-sum granularity.
struct TuneChoice {
  int schedule = 1, tcs = 1, fsets = 1, fbatch = 1, n1 = 0, fblocks = 1;
};

int json_int(const std::string &line, const char *field, int dflt) {
  const std::string k = std::string("\"") + field + "\": ";
  const size_t p = line.find(k);
  return p == std::string::npos ? dflt : std::atoi(line.c_str() + p + k.size());
}

// cache lines: {"key": "...", "schedule": S, "tcs": T, "fsets": F, "fbatch": B, "n1": N,
//               "fblocks": K, "ms": best, ..., "candidates": [[schedule, tcs, fsets, fbatch, n1, fblocks, ms], ...]}
std::map<std::string, TuneChoice> read_tune_cache(const char *path) {
  std::map<std::string, TuneChoice> out;
  if (!path) return out;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t k0 = line.find("\"key\": \"");
    if (k0 == std::string::npos || line.find("\"schedule\": ") == std::string::npos) continue;
    const size_t k1 = line.find('"', k0 + 8);
    TuneChoice c;
    c.schedule = json_int(line, "schedule", 1);
    c.tcs = json_int(line, "tcs", 0);
    c.fsets = json_int(line, "fsets", 0);
    c.fbatch = json_int(line, "fbatch", 1);
    c.n1 = json_int(line, "n1", 0);
    c.fblocks = json_int(line, "fblocks", 1);
    out[line.substr(k0 + 8, k1 - k0 - 8)] = c;
  }
  return out;
}

// Device copy of everything a step changes in a bucket.
struct Snapshot {
  std::vector<std::pair<void *, void *>> saved;  // (device location, copy)
  std::vector<size_t> sizes;
  std::vector<int64_t> steps;
  int epoch = 0;
  bool primed = false;
};

int take_snapshot(clv_ctx *ctx, Bucket &b, Snapshot &sn) {
  const size_t field = size_t(b.nch) * size_t(b.S) * 4 * b.esz;
  std::vector<std::pair<void *, size_t>> parts = {
      {b.psi[0], field}, {b.psi[1], field}, {b.psib[0], field}, {b.psib[1], field},
      {b.d_st, 2 * sizeof(ChainState) * b.nch}, {b.d_part, sizeof(double) * 2 * b.part_slot},
      {b.d_F, sizeof(double) * 3 * b.nch * b.Vs}, {b.d_acc, sizeof(double) * 5 * b.nch},
      {b.d_count, sizeof(int64_t) * b.nch}};
  for (auto &pr : parts) {
    void *copy = nullptr;
    if (hipMalloc(&copy, pr.second) != hipSuccess) {
      (void)hipGetLastError();
      for (auto &sv : sn.saved) (void)hipFree(sv.second);  // the copies made so far
      sn.saved.clear();
      sn.sizes.clear();
      return fail(ctx, CLV_ERR_NOMEM, "autotune: no room for a state snapshot");
    }
    HIPCHK(ctx, hipMemcpyAsync(copy, pr.first, pr.second, hipMemcpyDeviceToDevice, ctx->stream));
    sn.saved.push_back({pr.first, copy});
    sn.sizes.push_back(pr.second);
  }
  sn.steps = b.host_steps;
  sn.epoch = b.epoch;
  sn.primed = b.primed;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return CLV_OK;
}

int restore_snapshot(clv_ctx *ctx, Bucket &b, Snapshot &sn, bool release) {
  for (size_t i = 0; i < sn.saved.size(); ++i)
    HIPCHK(ctx, hipMemcpyAsync(sn.saved[i].first, sn.saved[i].second, sn.sizes[i],
                               hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  b.host_steps = sn.steps;
  b.epoch = sn.epoch;
  b.primed = sn.primed;
  if (release) {
    for (auto &pr : sn.saved) (void)hipFree(pr.second);
    sn.saved.clear();
  }
  return CLV_OK;
}

// Valid choices for bucket b: every instantiated row split whose kernels fit
// 64 KB of LDS at the bucket's row tile, and for each the reduction tiles that
// divide Nt and fit LDS, and the fused line sets and K_0 batches (of at most
// Nt) that fit 64 KB, each with one tau block and, where the chains give
// fewer than kFusedMinWaves waves, with the untimed block count.
std::vector<TuneChoice> tune_choices(const Bucket &b) {
  std::vector<TuneChoice> out;
  for (int n1 : split_options(b.X)) {
    if (update_lds_bytes(b, b.rows, n1) > 65536) continue;
    const size_t line = size_t(b.esz) * row_pad(b.X, b.X / n1);
    for (int q = 1; q <= 4; ++q)
      if (b.Nt % q == 0 && q * b.rows * b.X <= 1024 && line * q * b.rows <= 48 * 1024)
        out.push_back({1, q, 1, 1, n1});
    if (!b.kt_on)
      for (int f = 1; f <= 2; ++f)
        for (int B = 1; B <= 4; B *= 2)
          if ((B == 1 || B <= b.Nt) && fused_lds_bytes(b, f, B, n1) <= 65536) {
            out.push_back({2, b.tcs, f, B, n1, 1});
            const int nb = fused_blocks_for(b, b.nch);
            if (nb > 1) out.push_back({2, b.tcs, f, B, n1, nb});
          }
  }
  return out;
}

bool choice_valid(const Bucket &b, const TuneChoice &c) {
  for (const TuneChoice &v : tune_choices(b))
    if (v.n1 == c.n1 && v.schedule == c.schedule && (c.schedule == 2 || v.tcs == c.tcs) &&
        (c.schedule == 1 || (v.fsets == c.fsets && v.fbatch == c.fbatch && v.fblocks == c.fblocks)))
      return true;
  return false;
}

int apply_choice(clv_ctx *ctx, Bucket &b, const TuneChoice &c) {
  if (c.n1 != b.n1) {
    int rc = set_row_split(ctx, b, c.n1);
    if (rc) return rc;
  }
  b.schedule = c.schedule;
  if (c.schedule == 1) {
    b.tcs = c.tcs;
  } else {
    b.fused_sets = c.fsets;
    b.fused_batch = c.fbatch;
    b.fused_blocks = c.fblocks;
    b.fused_lds = fused_lds_bytes(b, c.fsets, c.fbatch);
  }
  b.primed = false;
  return CLV_OK;
}

// ms per step of bucket b under choice c (eager steps, no observables)
int time_choice(clv_ctx *ctx, Bucket &b, const TuneChoice &c, int nsteps, double *ms) {
  if (int rc = apply_choice(ctx, b, c)) return rc;
  for (int n = 0; n < 2; ++n) step_bucket(ctx, b, 0);
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
  for (int n = 0; n < nsteps; ++n) step_bucket(ctx, b, 0);
  HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
  HIPCHK(ctx, hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(ctx, hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = double(t) / nsteps;
  return CLV_OK;
}

}  // namespace

extern "C" {

int clv_autotune(clv_ctx *ctx, const char *cache_path) {
  if (!ctx) return CLV_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const std::string dev = device_name(ctx->device);
  const std::map<std::string, TuneChoice> cache = read_tune_cache(cache_path);
  const bool prof = ctx->profiling;
  ctx->profiling = false;
  int rc = CLV_OK;
  for (Bucket &b : ctx->buckets) {
    const std::string key = tune_key(dev, b, ctx->precision) + "|" + std::to_string(b.nch) + "ch";
    auto hit = cache.find(key);
    if (hit != cache.end() && choice_valid(b, hit->second)) {
      if ((rc = apply_choice(ctx, b, hit->second))) break;
      continue;
    }
    const std::vector<TuneChoice> cand = tune_choices(b);
    if (cand.size() <= 1) {
      if (!cand.empty() && (rc = apply_choice(ctx, b, cand[0]))) break;
      continue;
    }
    Snapshot sn;
    if ((rc = take_snapshot(ctx, b, sn))) break;
    std::vector<double> ms(cand.size(), 0.0);
    const int n = 8;
    for (size_t q = 0; q < cand.size() && !rc; ++q) {
      if (q > 0) rc = restore_snapshot(ctx, b, sn, false);
      if (!rc) rc = time_choice(ctx, b, cand[q], n, &ms[q]);
    }
    const int rc2 = restore_snapshot(ctx, b, sn, true);
    if (rc || rc2) {
      rc = rc ? rc : rc2;
      break;
    }
    size_t best = 0;
    for (size_t q = 1; q < cand.size(); ++q)
      if (ms[q] < ms[best]) best = q;
    if ((rc = apply_choice(ctx, b, cand[best]))) break;
    if (cache_path) {
      double two = 0.0, fused = 0.0;  // best of each schedule
      for (size_t q = 0; q < cand.size(); ++q) {
        double &m = cand[q].schedule == 1 ? two : fused;
        if (m == 0.0 || ms[q] < m) m = ms[q];
      }
      std::ofstream out(cache_path, std::ios::app);
      out << "{\"key\": \"" << key << "\", \"schedule\": " << cand[best].schedule
          << ", \"tcs\": " << b.tcs << ", \"fsets\": " << b.fused_sets << ", \"fbatch\": " << b.fused_batch
          << ", \"n1\": " << b.n1 << ", \"fblocks\": " << b.fused_blocks
          << ", \"ms\": " << ms[best]
          << ", \"ms_two_pass\": " << two << ", \"ms_fused\": " << fused << ", \"candidates\": [";
      for (size_t q = 0; q < cand.size(); ++q)
        out << (q ? ", " : "") << "[" << cand[q].schedule << ", " << cand[q].tcs << ", " << cand[q].fsets
            << ", " << cand[q].fbatch << ", " << cand[q].n1 << ", " << cand[q].fblocks << ", " << ms[q] << "]";
      out << "]}\n";
    }
  }
  ctx->profiling = prof;
  drop_graphs(ctx);
  return rc;
}

int clv_set_schedule(clv_ctx *ctx, int32_t schedule) {
  if (!ctx) return CLV_ERR_ARG;
  if (schedule < 0 || schedule > 2) return fail(ctx, CLV_ERR_ARG, "clv_set_schedule: bad schedule");
  for (Bucket &b : ctx->buckets) {
    if (schedule == 2 && b.fused_lds == 0)
      return fail(ctx, CLV_ERR_ARG, "clv_set_schedule: the fused schedule does not fit LDS for shape " +
                                        std::to_string(b.X) + "x" + std::to_string(b.R) + "x" +
                                        std::to_string(b.Nt) + " at this precision");
    if (schedule == 2 && b.kt_on)
      return fail(ctx, CLV_ERR_ARG, "clv_set_schedule: chains with kin_dims = 3 run the two-pass schedule");
  }
  for (Bucket &b : ctx->buckets) {
    if (schedule == 0) auto_plan(b, b.nch);
    else b.schedule = schedule;
  }
  drop_graphs(ctx);
  return CLV_OK;
}

int clv_set_fused_blocks(clv_ctx *ctx, int32_t blocks) {
  if (!ctx) return CLV_ERR_ARG;
  if (blocks < 0) return fail(ctx, CLV_ERR_ARG, "clv_set_fused_blocks: blocks < 0");
  for (const Bucket &b : ctx->buckets)
    if (blocks > 1 && (b.Nt % blocks != 0 || b.Nt / blocks < 2))
      return fail(ctx, CLV_ERR_ARG, "clv_set_fused_blocks: " + std::to_string(blocks) +
                                        " blocks do not divide Nt = " + std::to_string(b.Nt) +
                                        " into blocks of at least 2 slices");
  for (Bucket &b : ctx->buckets) {
    const int nb = (blocks > 0) ? blocks : fused_blocks_for(b, b.nch);
    b.fused_blocks = nb > 0 ? nb : 1;
  }
  drop_graphs(ctx);
  return CLV_OK;
}

int clv_set_fused_options(clv_ctx *ctx, int32_t line_sets, int32_t batch) {
  if (!ctx) return CLV_ERR_ARG;
  if (line_sets < 0 || line_sets > 2 || !(batch == 0 || batch == 1 || batch == 2 || batch == 4))
    return fail(ctx, CLV_ERR_ARG, "clv_set_fused_options: line_sets must be 0, 1 or 2 and batch 0, 1, 2 or 4");
  for (const Bucket &b : ctx->buckets) {
    if (b.fused_lds == 0) continue;  // no fused kernel for this shape
    const int f = line_sets ? line_sets : b.fused_sets, B = batch ? batch : b.fused_batch;
    if (f > b.fused_sets_max || fused_lds_bytes(b, f, B) > 65536)
      return fail(ctx, CLV_ERR_ARG, "clv_set_fused_options: " + std::to_string(f) + " line sets with batch " +
                                        std::to_string(B) + " do not fit LDS for shape " + std::to_string(b.X) +
                                        "x" + std::to_string(b.R) + "x" + std::to_string(b.Nt));
  }
  for (Bucket &b : ctx->buckets) {
    if (b.fused_lds == 0) continue;
    if (line_sets) b.fused_sets = line_sets;
    if (batch) b.fused_batch = batch;
    b.fused_lds = fused_lds_bytes(b, b.fused_sets, b.fused_batch);
  }
  drop_graphs(ctx);
  return CLV_OK;
}

int clv_set_graphs(clv_ctx *ctx, int32_t on) {
  if (!ctx) return CLV_ERR_ARG;
  ctx->use_graphs = on != 0;
  if (!ctx->use_graphs) drop_graphs(ctx);
  return CLV_OK;
}

int clv_set_profiling(clv_ctx *ctx, int32_t on) {
  if (!ctx) return CLV_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  drain_events(ctx);
  ctx->profiling = on != 0;
  for (int k = 0; k < 7; ++k) {
    ctx->kt_ms[k] = 0.0;
    ctx->kt_n[k] = 0;
  }
  return CLV_OK;
}

int clv_kernel_time(clv_ctx *ctx, int32_t kind, double *total_ms, int64_t *launches) {
  if (!ctx) return CLV_ERR_ARG;
  if (kind < 0 || kind > 6) return fail(ctx, CLV_ERR_ARG, "clv_kernel_time: bad kind");
  (void)hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  drain_events(ctx);
  if (total_ms) *total_ms = ctx->kt_ms[kind];
  if (launches) *launches = ctx->kt_n[kind];
  return CLV_OK;
}

int clv_set_row_split(clv_ctx *ctx, int32_t chain, int32_t n1) {
  int rc = check_chain(ctx, chain);
  if (rc) return rc;
  if (n1 < 0) return fail(ctx, CLV_ERR_ARG, "clv_set_row_split: n1 < 0");
  (void)hipSetDevice(ctx->device);
  return set_row_split(ctx, ctx->buckets[ctx->where[chain].first], n1);
}

int clv_get_row_split(const clv_ctx *ctx, int32_t chain) {
  if (!ctx || chain < 0 || chain >= int(ctx->where.size())) return CLV_ERR_ARG;
  return ctx->buckets[ctx->where[chain].first].n1;
}

int clv_get_schedule(const clv_ctx *ctx, int32_t chain) {
  if (!ctx || chain < 0 || chain >= int(ctx->where.size())) return CLV_ERR_ARG;
  return uses_fused(ctx->buckets[ctx->where[chain].first]) ? 2 : 1;
}

}  // extern "C"

// ====================================================== allocation-free step ==
// clv_step_noalloc: one step of one chain in caller-owned device buffers with
// no device allocation.  Everything its kernels need beyond the fields -- the
// chain's parameter record and controller state, the row-transform tables and
// the per-(tau, tile) (max, sum) partials of |D_0| -- is carved out of
// `scratch`, a file-scope device array of kScratchBytes that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel or a copy: it looks up each device's address of
// the array once (hipGetSymbolAddress, with that device current) and uses
// that pointer.
constexpr size_t kScratchBytes = size_t(1) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // device address of `scratch`, by device index

// This device's address of `scratch` (the device is current), looked up on the
// first call on the device.
int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (int(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(nullptr, CLV_ERR_HIP, std::string("clv_step_noalloc: hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return CLV_OK;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

bool overlaps(const void *a, const void *b, size_t n) {
  const char *x = static_cast<const char *>(a), *y = static_cast<const char *>(b);
  return x < y + n && y < x + n;
}

}  // namespace

extern "C" int clv_step_noalloc(const clv_shape *shape, int32_t precision, const clv_params *p,
                                const void *psi, const void *psibar, void *psi_out, void *psibar_out,
                                clv_step_stats *stats, int32_t device, void *stream) {
#define NOALLOC_CHK(expr)                                                                          \
  do {                                                                                             \
    const hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                        \
      (void)hipGetLastError();                                                                     \
      (void)hipStreamSynchronize(static_cast<hipStream_t>(stream)); /* host records in use */     \
      return fail(nullptr, CLV_ERR_HIP, std::string("clv_step_noalloc: ") + #expr + ": " +         \
                                            hipGetErrorString(e_));                                \
    }                                                                                              \
  } while (0)
  if (!shape || !p || !psi || !psibar || !psi_out || !psibar_out || !stats)
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: NULL argument");
  if (precision != CLV_FP32 && precision != CLV_FP64)
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: precision must be CLV_FP32 or CLV_FP64");
  const clv_shape &s = *shape;
  if (s.nx <= 0 || s.ny <= 0 || s.nz <= 0 || s.nt < 2 || !supported_x(s.nx) ||
      int64_t(s.nx) * s.ny * s.nz * s.nt >= (int64_t(1) << 31))
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: bad or unsupported shape");
  if (p->kin_dims != 1)
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: kin_dims must be 1 (the transverse term needs a buffer)");
  if (p->conj_drift_mode != CLV_CONJ_DRIFT_REFERENCE && p->conj_drift_mode != CLV_CONJ_DRIFT_ACTION)
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: bad conj_drift_mode");
  if (!(p->a_x > 0.0) || !(p->DT > 0.0) || !(p->eps > 0.0))
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: a_x, DT and eps must be positive");
  if (stats->steps < 0) return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: stats->steps < 0");
  Bucket b;
  b.shape = s;
  b.nch = 1;
  plan_bucket(b, precision);
  // the scratch carve-up: parameters, two state slots, tables, partials
  size_t off = 0;
  auto take = [&](size_t n) {
    const size_t o = off;
    off = align256(off + n);
    return o;
  };
  const size_t o_prm = take(sizeof(ChainParams));
  const size_t o_st = take(2 * sizeof(ChainState));
  const size_t o_tw = take(size_t(b.esz) * tw_entries(b.X, b.n1));
  const size_t o_p2 = take(sizeof(double) * b.X);
  const size_t o_part = take(sizeof(double) * b.part_slot);
  if (off > kScratchBytes)
    return fail(nullptr, CLV_ERR_ARG,
                "clv_step_noalloc: shape " + std::to_string(s.nx) + "x" + std::to_string(s.ny) + "x" +
                    std::to_string(s.nz) + "x" + std::to_string(s.nt) + " needs " + std::to_string(off) +
                    " bytes of scratch, more than its " + std::to_string(kScratchBytes) +
                    " (Nt x row tiles x 16 bytes of partials)");
  const size_t fbytes = size_t(b.S) * 4 * b.esz;
  const void *outs[2] = {psi_out, psibar_out};
  for (const void *o : outs)
    if (overlaps(o, psi, fbytes) || overlaps(o, psibar, fbytes))
      return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: output fields overlap the input fields");
  if (overlaps(psi_out, psibar_out, fbytes))
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: psi_out and psibar_out overlap");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    return fail(nullptr, CLV_ERR_ARG, "clv_step_noalloc: bad device index");
  }
  NOALLOC_CHK(hipSetDevice(device));
  void *scr = nullptr;
  if (int rc = scratch_address(device, &scr)) return rc;
  char *base = static_cast<char *>(scr);
  b.d_prm = reinterpret_cast<ChainParams *>(base + o_prm);
  b.d_st = reinterpret_cast<ChainState *>(base + o_st);
  b.d_tw = base + o_tw;
  b.d_p2 = reinterpret_cast<double *>(base + o_p2);
  b.d_part = reinterpret_cast<double *>(base + o_part);
  b.psi[0] = const_cast<void *>(psi);
  b.psib[0] = const_cast<void *>(psibar);
  b.psi[1] = psi_out;
  b.psib[1] = psibar_out;
  b.schedule = 1;
  b.epoch = 0;
  // host records, kept alive until the stream is synchronized below
  const ChainParams hp = to_device_params(*p, 0, 1);
  ChainState st;
  std::memset(&st, 0, sizeof(st));
  st.dt = stats->dt;
  st.max_drift = stats->max_drift;
  st.mean_drift = stats->mean_drift;
  st.t_langevin = stats->t_langevin;
  st.steps = stats->steps;
  st.cur_step = stats->steps > 0 ? stats->steps - 1 : 0;
  st.clamp_low = stats->clamp_low;
  st.clamp_high = stats->clamp_high;
  st.nonfinite = stats->nonfinite;
  st.parity = 0;  // the inputs are buffer 0, the outputs buffer 1
  std::vector<char> tw;
  std::vector<double> p2;
  row_tables_host(b.X, b.n1, precision, &tw, &p2);
  clv_ctx ctx;  // launch record only: no device memory, the caller's stream
  ctx.device = device;
  ctx.precision = precision;
  ctx.stream = static_cast<hipStream_t>(stream);
  ctx.use_graphs = false;
  NOALLOC_CHK(hipMemcpyAsync(b.d_prm, &hp, sizeof(hp), hipMemcpyHostToDevice, ctx.stream));
  NOALLOC_CHK(hipMemcpyAsync(b.d_st, &st, sizeof(st), hipMemcpyHostToDevice, ctx.stream));
  NOALLOC_CHK(hipMemcpyAsync(b.d_tw, tw.data(), tw.size(), hipMemcpyHostToDevice, ctx.stream));
  NOALLOC_CHK(hipMemcpyAsync(b.d_p2, p2.data(), sizeof(double) * p2.size(), hipMemcpyHostToDevice, ctx.stream));
  // the two-pass step of clv_propagate: partials, device dt, update into buffer 1
  row_kernel(&ctx, b, 0, 0, 1, 0);
  controller(&ctx, b, 0, 1, 0, 0.0, 0);
  row_kernel(&ctx, b, 1, 0, 1, 0);
  NOALLOC_CHK(hipGetLastError());
  ChainState out;
  NOALLOC_CHK(hipMemcpyAsync(&out, b.d_st, sizeof(out), hipMemcpyDeviceToHost, ctx.stream));
  NOALLOC_CHK(hipStreamSynchronize(ctx.stream));
  ctx.stream = nullptr;  // the caller's
  stats_from_state(out, stats);
  return CLV_OK;
#undef NOALLOC_CHK
}
