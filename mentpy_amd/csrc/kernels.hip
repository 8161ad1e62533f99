// HIP kernels (gfx950) of the measurement-pattern simulator and their C-ABI
// launchers.
//
// Execution model: one simulation per lane.  The live state of a simulation
// (2^n_slots complex64 amplitudes for SV, 4^n_slots entries for DM) lives in
// VGPRs for the whole run; a step is an in-place pass over amplitude pairs
// (quartets for DM) whose register indices are compile-time constants because
// the measured slot rotates with period n_slots and the step loop is unrolled
// by n_slots.  Per-step control (angle index, append/CZ flags, CZ neighbour
// mask) is wave-uniform and read with scalar loads, so a wave of 64
// simulations of one pattern runs with no divergence and no cross-lane
// traffic; for periodic plans (grids in column order, chains) the CZ signs
// are compile-time constants instead.  HBM sees only the angles, the input
// states and the outputs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>

#include "plan.h"

namespace mbqc {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// flip the sign of both components when sbit == 0x80000000
__device__ __forceinline__ float2 flip(float2 v, uint32_t sbit) {
  return make_float2(__uint_as_float(__float_as_uint(v.x) ^ sbit),
                     __uint_as_float(__float_as_uint(v.y) ^ sbit));
}
// CZ sign of slot-index j against the neighbour mask: parity(j & mask) << 31
__device__ __forceinline__ uint32_t czbit(uint32_t j, uint32_t mask) {
  return uint32_t(__builtin_popcount(j & mask) & 1) << 31;
}

// Round-to-nearest-even integer of v (|v| < 2^22) by the 1.5 * 2^23 shifter:
// fl(v + 1.5 * 2^23) has a unit last mantissa bit, so its lowest bit is the
// parity of rint(v).  With the shifter added in an FMA, v = a * c is rounded
// once (no rounding of a * c first).
constexpr float kShifter = 12582912.0f;      // 1.5 * 2^23

// sin and cos of x for |x| up to a few thousand: x = k pi + r with a two-term
// Cody-Waite split of pi (hi part the nearest float), |r| <= pi/2, then
// polynomials fitted for minimum maximum error on [-pi/2, pi/2] (odd degree 9
// for sin, even degree 10 for cos; fitted in double by linear programming on
// 4,000 Chebyshev points, approximation error below 5e-10, float evaluation
// error about 1e-7), and the sign (-1)^k on both.
__device__ __forceinline__ void sincos_poly(float x, float* sn, float* cs) {
  const float kb = fmaf(x, 0.318309886183790671538f, kShifter);
  const float k = kb - kShifter;
  float r = fmaf(-k, 3.14159274101257324219f, x);
  r = fmaf(-k, -8.74227800037248e-8f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.601407004e-6f, -1.980598900e-4f);
  ps = fmaf(r2, ps, 8.332983591e-3f);
  ps = fmaf(r2, ps, -1.666665375e-1f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.650938882e-7f, 2.479057366e-5f);
  pc = fmaf(r2, pc, -1.388906501e-3f);
  pc = fmaf(r2, pc, 4.166670144e-2f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  const uint32_t sgn = __float_as_uint(kb) << 31;
  *sn = __uint_as_float(__float_as_uint(sr) ^ sgn);
  *cs = __uint_as_float(__float_as_uint(cr) ^ sgn);
}

// The reference's global phase, prod_k sgn(cos(theta_k/2)) e^{i theta_k/2}
// (the factor its amplitude-sum partial trace introduces,
// mentpy/calculator/state_ops.py:68-73).  With m_k = rint(theta_k / 2 pi),
// theta_k/2 - pi m_k lies in [-pi/2, pi/2], so sgn(cos(theta_k/2)) = (-1)^m_k
// and the product is exp(i (sum_k theta_k / 2 - pi sum_k m_k)).  The angles
// are summed in double (full-rate on this GPU) and the integers m_k in int32,
// and the phase is formed once at the end.
// m_k must be exact for every float32 angle, also the ones next to an odd
// multiple of pi where cos(theta/2) changes sign (float32(pi) itself lies
// 8.7e-8 above pi: m = 1, phase -i).  A float32 product with fl32(1/2pi),
// 4e-8 too small relative to 1/2pi, rounds those down, so m_k is rounded in
// double: fma(theta, fl64(1/2pi), 1.5 * 2^52) rounds the exact product to an
// integer (ties to even) held in the low word of its bits, with a relative
// error of 1.1e-16 that can only matter within 1e-13 of a half-integer;
// no float32 angle below 2^12 comes that close (the nearest lies more than
// 1e-9 away, checked exhaustively on the CPU).
constexpr double kShifter64 = 6755399441055744.0;   // 1.5 * 2^52
struct RefPhase {
  double th_sum = 0.0;
  int32_t m_sum = 0;
  __device__ __forceinline__ void add(float th) {
    const double d = double(th);
    const double mb = fma(d, 0.15915494309189533577, kShifter64);
    m_sum += int32_t(uint32_t(__double_as_longlong(mb)));
    th_sum += d;
  }
  // add(th) where `on`, nothing elsewhere
  __device__ __forceinline__ void add_if(float th, bool on) {
    const double d = on ? double(th) : 0.0;
    const double mb = fma(d, 0.15915494309189533577, kShifter64);
    m_sum += int32_t(uint32_t(__double_as_longlong(mb)));
    th_sum += d;
  }
  __device__ __forceinline__ float2 value() const {
    constexpr double kPi = 3.14159265358979323846;
    double a = 0.5 * th_sum - kPi * double(m_sum);
    a -= 2.0 * kPi * rint(a * (0.5 / kPi));
    double sd, cd;
    sincos(a, &sd, &cd);
    return make_float2(float(cd), float(sd));
  }
};

// b = e^{-i theta}; with REF the step's reference phase is accumulated too.
template <bool REF>
__device__ __forceinline__ float2 step_coeff(float th, RefPhase* ph) {
  if constexpr (REF) ph->add(th);
  float s, c;
  sincos_poly(th, &s, &c);
  return make_float2(c, -s);
}

// ---------------------------------------------------------------- SV step
// Measure slot Q at XY angle theta (b = e^{-i theta}) and put the next
// scheduled qubit (|+> with its CZs) into slot Q.  With x / y the halves of the
// state where slot Q is 0 / 1:  u0 = x + b y  is the new Q=0 half and
// alpha u0 + beta x  the new Q=1 half, where the per-step, per-pair
// coefficients (host-computed) fold in the CZ signs, the CZ with the measured
// qubit and the "no append" case (alpha = beta = 0).  The 1/sqrt(2) factors of
// the |+> and of the projector are left out (tracked as an exponent).
template <int NS, int Q>
__device__ __forceinline__ void sv_step(float2 (&t)[1 << NS], float2 b, const float* coef) {
  constexpr int H = 1 << NS;
  constexpr int QB = 1 << Q;
  int i = 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & QB) continue;
    const float al = coef[2 * i], be = coef[2 * i + 1];
    ++i;
    const float2 x = t[j], y = t[j | QB];
    float2 u0;
    u0.x = fmaf(-b.y, y.y, fmaf(b.x, y.x, x.x));
    u0.y = fmaf(b.x, y.y, fmaf(b.y, y.x, x.y));
    t[j] = u0;
    t[j | QB] = make_float2(fmaf(al, u0.x, be * x.x), fmaf(al, u0.y, be * x.y));
  }
}

// sv_step of a periodic pattern (every step appends a CZ neighbour of the
// measured qubit; the appended qubit's other neighbours are
// periodic_nmask(NS, NS - 1 - Q), a compile-time mask): the new Q=1 half is
// sign (2 x - u0) with sign the parity of j & NMASK, the same arithmetic as
// alpha u0 + beta x with alpha = -sign, beta = 2 sign, with the signs folded
// into the instructions instead of read from the coefficient table.
template <int NS, int Q, uint32_t NMASK>
__device__ __forceinline__ void sv_step_periodic(float2 (&t)[1 << NS], float2 b) {
  constexpr int H = 1 << NS;
  constexpr int QB = 1 << Q;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & QB) continue;
    const float2 x = t[j], y = t[j | QB];
    float2 u0;
    u0.x = fmaf(-b.y, y.y, fmaf(b.x, y.x, x.x));
    u0.y = fmaf(b.x, y.y, fmaf(b.y, y.x, x.y));
    t[j] = u0;
    if (__builtin_popcount(uint32_t(j) & NMASK) & 1)
      t[j | QB] = make_float2(fmaf(-2.0f, x.x, u0.x), fmaf(-2.0f, x.y, u0.y));
    else
      t[j | QB] = make_float2(fmaf(2.0f, x.x, -u0.x), fmaf(2.0f, x.y, -u0.y));
  }
}

// Job description shared by the SV / DM / adjoint kernels (passed by value
// or read from global memory by the persistent kernel).
struct Job {
  const int32_t* tab;       // packed device tables of the plan
  int32_t off_in, off_sign, off_out, off_svres, off_dmres, e_sv, e_dm;
  int32_t ns, n_steps, n_out, live_mask, off_coef;
  int32_t off_aidx, off_run;  // contiguous angle indices, stretch ends
  // 1: no step has a fixed angle and every angle offset angle_index *
  // angle_stride_a fits in 32 bits (the statevector kernels' short fetch)
  int32_t fast_fetch;
  const float* angles;
  int64_t angle_stride;     // elements between angle vectors (0: shared)
  int64_t angle_stride_a;   // elements between the angles of one vector
  const float2* psi;
  int64_t in_stride;
  float2* out;
  float2* phase_out;
  int64_t S, total;
};

// Identity modifier of the per-step projection coefficient b = e^{-i theta}.
struct NoModify {
  __device__ __forceinline__ float2 operator()(int, float2 b) const { return b; }
};

// Pair coefficients of step k: row (steps[k].flags >> 16) of the plan's
// coefficient table.  Long patterns store each distinct row once (a grid has
// one per slot), so their rows stay within the scalar cache; short ones keep
// one row per step (row = k), and kernels built with PER_STEP address them
// without reading the step flags first.
template <int NS, bool PER_STEP = false>
__device__ __forceinline__ const float* coef_row(const float* coef, const mbqc_step* steps, int k) {
  if constexpr (PER_STEP) return coef + (k << NS);
  return coef + (int(steps[k].flags >> 16) << NS);
}

// Lanes that share an angle vector in aligned groups of G = kShareGroup<NS>
// consecutive lanes (S a multiple of G inputs per vector) can share the sine,
// cosine and reference-phase work: lane j of a group evaluates step j and the
// values are broadcast inside the group (DPP quad permutations for G = 4, a
// ds_swizzle for G = 8, a DPP row broadcast for G = 16; no memory access).
// Windows of 2 and 3 slots share over 16 lanes, which evaluate all the steps
// of one angle fetch (8 or 6) at once.
__host__ __device__ constexpr int share_group(int ns) { return ns <= 3 ? 16 : (ns == 4 ? 4 : 8); }
template <int NS>
constexpr int kShareGroup = share_group(NS);

template <int G, int SRC>
__device__ __forceinline__ float group_bcast(float v) {
  static_assert(SRC < G, "source lane inside the group");
  if constexpr (G == 16) {
    // row_newbcast: lane SRC of each 16-lane row to the whole row
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x150 + SRC, 0xf, 0xf, false));
  } else if constexpr (G == 2) {
    constexpr int ctrl = SRC | (SRC << 2) | ((2 + SRC) << 4) | ((2 + SRC) << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, false));
  } else if constexpr (G == 4) {
    constexpr int ctrl = SRC | (SRC << 2) | (SRC << 4) | (SRC << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xf, 0xf, false));
  } else {
    // bit mode: lane' = (lane & and_mask) | or_mask within each 32-lane half
    constexpr int pattern = (0x20 - G) | (SRC << 5);
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), pattern));
  }
}

// Block [k0, k0 + NS): lane j of the group evaluates step j (and, with REF,
// adds its reference phase: the group's lanes hold disjoint parts of the sum,
// reduced by group_reduce_phase at the end).
template <int NS, bool REF, int... SUBS>
__device__ __forceinline__ void shared_block_coeffs(const float (&th)[NS], float2 (&bb)[NS],
                                                    RefPhase* ph, int k0, int n_steps,
                                                    std::integer_sequence<int, SUBS...>) {
  constexpr int G = kShareGroup<NS>;
  static_assert(NS >= 2 && NS <= G, "one step per lane of the group");
  const int j = int(threadIdx.x) & (G - 1);
  float mine = th[0];
#pragma unroll
  for (int sub = 1; sub < NS; ++sub) mine = (j == sub) ? th[sub] : mine;
  float sn, cs;
  sincos_poly(mine, &sn, &cs);
  if constexpr (REF) ph->add_if(mine, j < NS && k0 + j < n_steps);
  ((bb[SUBS] = make_float2(group_bcast<G, SUBS>(cs), -group_bcast<G, SUBS>(sn))), ...);
}

// Sum of the reference-phase parts over the G lanes of a sharing group.
template <int G>
__device__ __forceinline__ void group_reduce_phase(RefPhase& p) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    p.m_sum += __shfl_xor(p.m_sum, m, 64);
    p.th_sum += __shfl_xor(p.th_sum, m, 64);
  }
}

template <int NS, bool REF, bool FULL, bool PER_STEP, bool SH, bool PERIODIC, int SUB,
          typename Mod>
__device__ __forceinline__ void sv_sub(float2 (&t)[1 << NS], const mbqc_step* steps,
                                       const float* coef, int k0, int n_steps, const float (&th)[NS],
                                       const float2 (&bb)[NS], RefPhase* ph, Mod& mod) {
  const int k = k0 + SUB;
  if (FULL || k < n_steps) {
    float2 b0;
    if constexpr (SH) {
      b0 = bb[SUB];   // reference phase added where the coefficients were evaluated
    } else {
      b0 = step_coeff<REF>(th[SUB], ph);
    }
    if constexpr (PERIODIC)
      sv_step_periodic<NS, NS - 1 - SUB, periodic_nmask(NS, SUB)>(t, mod(k, b0));
    else
      sv_step<NS, NS - 1 - SUB>(t, mod(k, b0), coef_row<NS, PER_STEP>(coef, steps, k));
  }
}

template <int NS, bool REF, bool FULL, bool PER_STEP, bool SH, bool PERIODIC, typename Mod,
          int... SUBS>
__device__ __forceinline__ void sv_block(float2 (&t)[1 << NS], const mbqc_step* steps,
                                         const float* coef, int k0, int n_steps,
                                         const float (&th)[NS], RefPhase* ph, Mod& mod,
                                         std::integer_sequence<int, SUBS...> seq) {
  float2 bb[NS];
  if constexpr (SH) shared_block_coeffs<NS, REF>(th, bb, ph, k0, n_steps, seq);
  (sv_sub<NS, REF, FULL, PER_STEP, SH, PERIODIC, SUBS>(t, steps, coef, k0, n_steps, th, bb, ph,
                                                       mod),
   ...);
}

// Block B of an angle fetch whose coefficients the group's lanes evaluated at
// once (lane m of the 16-lane row holds step m's sine and cosine).
template <int NS, bool REF, bool PER_STEP, bool PERIODIC, int B, typename OnBlock, typename Mod,
          int... SUBS>
__device__ __forceinline__ void sv_fetched_block(float2 (&t)[1 << NS], const mbqc_step* steps,
                                                 const float* coef, int blk, int n_steps, float cs,
                                                 float sn, RefPhase* ph, OnBlock& on_block,
                                                 Mod& mod, std::integer_sequence<int, SUBS...>) {
  float2 bb[NS];
  ((bb[SUBS] = make_float2(group_bcast<16, B * NS + SUBS>(cs), -group_bcast<16, B * NS + SUBS>(sn))),
   ...);
  const float th_unused[NS] = {};
  on_block(blk + B, t);
  (sv_sub<NS, REF, true, PER_STEP, true, PERIODIC, SUBS>(t, steps, coef, (blk + B) * NS, n_steps,
                                                         th_unused, bb, ph, mod),
   ...);
}

// The K full blocks of one angle fetch, sharing over 16 lanes: lane m of the
// row evaluates step m of the fetch (and adds its reference phase).
template <int NS, int K, bool REF, bool PER_STEP, bool PERIODIC, typename OnBlock, typename Mod,
          int... BS>
__device__ __forceinline__ void sv_fetched_blocks(float2 (&t)[1 << NS], const mbqc_step* steps,
                                                  const float* coef, int blk, int n_steps,
                                                  const float (&th)[K * NS], RefPhase* ph,
                                                  OnBlock& on_block, Mod& mod,
                                                  std::integer_sequence<int, BS...>) {
  static_assert(K * NS <= 16, "one step per lane of the row");
  const int j = int(threadIdx.x) & 15;
  float mine = th[0];
#pragma unroll
  for (int m = 1; m < K * NS; ++m) mine = (j == m) ? th[m] : mine;
  float sn, cs;
  sincos_poly(mine, &sn, &cs);
  if constexpr (REF) ph->add_if(mine, j < K * NS);
  (sv_fetched_block<NS, REF, PER_STEP, PERIODIC, BS>(t, steps, coef, blk, n_steps, cs, sn, ph,
                                                     on_block, mod,
                                                     std::make_integer_sequence<int, NS>{}),
   ...);
}

// Angles of steps [k0, k0 + M).  Steps past the end (n_steps >= 1) read the
// last step's angle, so every load is issued from a valid address (their
// values are never used).  Every step issues exactly one load, from its
// angle-vector entry or from its stored fixed angle (the address is selected,
// not the value): with a fixed number of loads in flight, the wait before a
// block's math covers only that block's angles, not loads issued for later
// blocks.
// The angle indices come from the plan's contiguous, padded index table
// (aidx, M <= kAngleIdxPad entries past the last step at most), so they are
// adjacent scalar loads.
// Jobs with Job::fast_fetch (no fixed angle, 32-bit angle offsets: the usual
// case) take one uniform branch per fetch to loads whose addresses are a
// 32-bit scalar product added to the lane's angle vector; otherwise each step
// chooses its source (a uniform branch and a 64-bit product per step).
template <int M>
__device__ __forceinline__ void fetch_step_angles(const mbqc_step* steps, const int32_t* aidx,
                                                  int k0, int n_steps, const float* ang,
                                                  int64_t sa, bool fast, float (&th)[M]) {
  static_assert(M <= kAngleIdxPad, "index table padding");
  // angles live in device memory: a global-typed pointer keeps the loads
  // global (a descriptor read from memory otherwise leaves them flat)
  using gfloat = const __attribute__((address_space(1))) float;
  gfloat* gang = (gfloat*)ang;
  int ai[M];
#pragma unroll
  for (int m = 0; m < M; ++m) ai[m] = aidx[k0 + m];
  if (fast) {
#pragma unroll
    for (int m = 0; m < M; ++m) th[m] = gang[uint32_t(ai[m]) * uint32_t(sa)];
    return;
  }
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int k = k0 + m < n_steps ? k0 + m : n_steps - 1;
    gfloat* src = ai[m] >= 0 ? gang + int64_t(ai[m]) * sa : (gfloat*)&steps[k].fixed_angle;
    th[m] = *src;
  }
}

// Blocks whose angles one fetch brings in: small windows do little work per
// block, so their loads are issued several blocks ahead of their use.
template <int NS>
constexpr int kFetchBlocks = NS <= 2 ? 4 : (NS == 3 ? 2 : 1);

// Run blocks [blk0, blk1) of NS steps: full blocks of NS unrolled steps, with
// the angles of the next kFetchBlocks blocks in flight, then the partial tail
// block.  `on_block(blk, t)` runs at the start of every block (after its
// rescale).  The full blocks run in stretches that end before the next
// rescaled block, so the inner loop is straight-line code whose loop-carried
// state needs no register copies (a rescale branch inside it cost 32 moves
// per block).  PERIODIC: the full blocks take the compile-time coefficients
// of a periodic plan (Plan::periodic), the tail block the table.
template <int NS, bool REF, bool PER_STEP, bool SH = false, bool PERIODIC = false,
          bool SETTLE = true, typename OnBlock, typename Mod = NoModify>
__device__ __forceinline__ void sv_run_blocks(float2 (&t)[1 << NS], const Job& J, const float* ang,
                                              RefPhase* ph, int blk0, int blk1, OnBlock on_block,
                                              Mod&& mod = Mod()) {
  constexpr int H = 1 << NS;
  constexpr int K = kFetchBlocks<NS>;
  using Seq = std::make_integer_sequence<int, NS>;
  if (blk0 >= blk1) return;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* run_end = J.tab + J.off_run;
  const int32_t* aidx = J.tab + J.off_aidx;
  const float* coef = reinterpret_cast<const float*>(J.tab + J.off_coef);
  const int n = J.n_steps;
  const int nfull = n / NS;
  const int full_end = nfull < blk1 ? nfull : blk1;
  // th_next holds the angles of blocks blk .. blk + K - 1
  float th_next[K * NS];
  const int64_t sa = J.angle_stride_a;
  const bool fast = J.fast_fetch != 0;
  fetch_step_angles<K * NS>(steps, aidx, blk0 * NS, n, ang, sa, fast, th_next);
  int blk = blk0;
  while (blk < full_end) {
    // SETTLE: the angles in flight enter the block loop through an empty
    // asm that reads and writes them.  The compiler places its memory-counter
    // waits per loop header, and angles arriving there straight from loads
    // (the first fetch) would put waits inside the loop body, where they also
    // wait for the loads just issued for the next fetch; with the asm the
    // wait is once per stretch of blocks (4x20 forward: 0.344 -> 0.332 and
    // 0.348 -> 0.343 ms per step on two boxes).  The batched kernel, which
    // starts a stretch per simulation, runs without it.
    if constexpr (SETTLE) {
#pragma unroll
      for (int m = 0; m < K * NS; ++m) asm("" : "+v"(th_next[m]));
    }
    const int re = run_end[blk];
    const int end = re < full_end ? re : full_end;
    if (resc[blk]) {
#pragma unroll
      for (int j = 0; j < H; ++j) t[j] = cscale(t[j], 0x1p-32f);
    }
    for (; blk + K <= end; blk += K) {
      float th[K * NS];
#pragma unroll
      for (int m = 0; m < K * NS; ++m) th[m] = th_next[m];
      fetch_step_angles<K * NS>(steps, aidx, (blk + K) * NS, n, ang, sa, fast, th_next);
      if constexpr (SH && kShareGroup<NS> == 16) {
        sv_fetched_blocks<NS, K, REF, PER_STEP, PERIODIC>(t, steps, coef, blk, n, th, ph, on_block,
                                                          mod, std::make_integer_sequence<int, K>{});
      } else {
#pragma unroll
        for (int b = 0; b < K; ++b) {
          float tb[NS];
#pragma unroll
          for (int s2 = 0; s2 < NS; ++s2) tb[s2] = th[b * NS + s2];
          on_block(blk + b, t);
          sv_block<NS, REF, true, PER_STEP, SH, PERIODIC>(t, steps, coef, (blk + b) * NS, n, tb,
                                                          ph, mod, Seq{});
        }
      }
    }
    if constexpr (K > 1) {
      if (blk < end) {
        // fewer than K blocks left in the stretch: run them from th_next,
        // then fetch from the stretch's end
#pragma unroll
        for (int b = 0; b < K - 1; ++b) {
          if (blk + b < end) {
            float tb[NS];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) tb[s2] = th_next[b * NS + s2];
            on_block(blk + b, t);
            sv_block<NS, REF, true, PER_STEP, SH, PERIODIC>(t, steps, coef, (blk + b) * NS, n, tb,
                                                            ph, mod, Seq{});
          }
        }
        blk = end;
        fetch_step_angles<K * NS>(steps, aidx, blk * NS, n, ang, sa, fast, th_next);
      }
    }
  }
  if (blk < blk1) {
    // the partial tail block (blk == nfull)
    if (resc[blk]) {
#pragma unroll
      for (int j = 0; j < H; ++j) t[j] = cscale(t[j], 0x1p-32f);
    }
    float tb[NS];
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) tb[s2] = th_next[s2];
    on_block(blk, t);
    sv_block<NS, REF, false, PER_STEP, SH, false>(t, steps, coef, blk * NS, n, tb, ph, mod, Seq{});
  }
}

// All steps of the pattern.
template <int NS, bool REF, bool PERIODIC = false, typename OnBlock, typename Mod = NoModify>
__device__ __forceinline__ void sv_run_steps(float2 (&t)[1 << NS], const Job& J, const float* ang,
                                             RefPhase* ph, OnBlock on_block, Mod&& mod = Mod()) {
  sv_run_blocks<NS, REF, false, false, PERIODIC>(t, J, ang, ph, 0, (J.n_steps + NS - 1) / NS,
                                                 on_block, static_cast<Mod&&>(mod));
}

template <int NS>
__device__ __forceinline__ void load_block_angles(const mbqc_step* steps, int k0, int n_steps,
                                                  const float* ang, int64_t sa, float (&th)[NS]) {
#pragma unroll
  for (int sub = 0; sub < NS; ++sub) {
    const int k = k0 + sub;
    th[sub] = 0.0f;
    if (k < n_steps) {
      const int ai = steps[k].angle_index;
      th[sub] = ai >= 0 ? ang[int64_t(ai) * sa] : steps[k].fixed_angle;
    }
  }
}

template <int NS>
__device__ __forceinline__ void sv_load_input(const Job& J, const float2* psi, float2 (&t)[1 << NS]) {
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
#pragma unroll
  for (int j = 0; j < (1 << NS); ++j) {
    const float2 v = psi[in_index[j]];
    t[j] = flip(v, sign[j] ? 0x80000000u : 0u);
  }
}

__device__ __forceinline__ float sv_final_scale(int e) {
  float s = ldexpf(1.0f, -(e >> 1));
  return (e & 1) ? s * 0.70710678118654752440f : s;
}

// Makes this wave's LDS writes visible to its own later LDS reads: waits for
// them to complete and keeps the compiler from moving LDS accesses across.
// (Waves only exchange data through their own LDS tile, so no workgroup
// barrier is needed, and persistent waves with different trip counts can use
// it.)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0), other counters unchanged
  __builtin_amdgcn_wave_barrier();
}

// Output rows of a wave of 64 consecutive simulations (row r = simulation
// g0 + r, 2^n_out complex entries, slot index j -> column out_pos[j], times f)
// through a per-wave LDS tile of 16 columns: each pass writes columns
// [c0, c0 + 16) of all 64 rows, so a store instruction covers four 128-byte
// row segments instead of one entry in each of 64 rows.  Used for outputs of
// 32 amplitudes (five live slots); rows r >= rows_valid are not written.
// The tiles are dynamic LDS (kColumnTile entries per wave of the workgroup),
// so a one-wave workgroup reserves one tile.
constexpr int kColumnTile = 64 * 17;

// IDENT: slot index j is column j (out_pos unused, f = 1).
template <int NS, bool IDENT = false>
__device__ __forceinline__ void wave_store_rows(const float2 (&t)[1 << NS], const int32_t* out_pos,
                                                float2 f, float2* obase, int n_out, int rows_valid) {
  constexpr int H = 1 << NS;
  constexpr int CW = H < 16 ? H : 16;
  constexpr int LOG_CW = NS < 4 ? NS : 4;
  static_assert(64 * (CW + 1) <= kColumnTile, "tile size");
  extern __shared__ float2 column_tiles[];
  const int lane = threadIdx.x & 63;
  float2* sw = column_tiles + (threadIdx.x >> 6) * kColumnTile;
  const int dout = 1 << n_out;
  const int cs = n_out < LOG_CW ? n_out : LOG_CW;   // log2 of the columns per pass
  const int cw = 1 << cs;
  for (int c0 = 0; c0 < dout; c0 += cw) {
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int p = (IDENT ? j : out_pos[j]) - c0;
      if (p >= 0 && p < cw) sw[lane * (CW + 1) + p] = IDENT ? t[j] : cmul(t[j], f);
    }
    wave_lds_sync();
    for (int e = lane; e < 64 * cw; e += 64) {
      const int r = e >> cs, c = e & (cw - 1);
      if (r < rows_valid) obase[int64_t(r) * dout + c0 + c] = sw[r * (CW + 1) + c];
    }
    wave_lds_sync();
  }
}

// How a wave's output rows are stored.
enum StoreMode {
  kStoreDirect = 0,   // each lane writes its own row
  kStoreBlock = 1,    // all rows of the workgroup's waves staged at once (outputs of <= 16
                      // amplitudes; every wave of the workgroup must take this path)
  kStoreColumns = 2,  // wave_store_rows (any wave on its own)
};

// One SV simulation on this lane (g = b*S + s); the wave covers the 64
// consecutive simulations g - lane .. g - lane + 63, of which the first
// rows_valid are written (kStoreDirect: active lanes write).  Inactive lanes
// run on simulation 0.
template <int NS, bool REF, int STORE, bool PER_STEP = false, bool SH = false,
          bool PERIODIC = false, bool SETTLE = true>
__device__ __forceinline__ void sv_sim(const Job& J, int64_t g, bool active, int rows_valid,
                                       uint32_t flags) {
  constexpr int H = 1 << NS;
  const int64_t gg = active ? g : 0;
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  float2 t[H];
  sv_load_input<NS>(J, J.psi + si * J.in_stride, t);
  RefPhase rp;
  sv_run_blocks<NS, REF, PER_STEP, SH, PERIODIC, SETTLE>(t, J, ang, &rp, 0,
                                                         (J.n_steps + NS - 1) / NS,
                                       [](int, const float2 (&)[H]) {});
  // with shared coefficients each lane of a group added the phases of its own steps
  if constexpr (REF && SH) group_reduce_phase<kShareGroup<NS>>(rp);
  const float2 ph = REF ? rp.value() : make_float2(1.0f, 0.0f);
  // output gather: slot-index j -> output index out_pos[j] (dead slots skipped)
  const int32_t* out_pos = J.tab + J.off_out;
  float scale = sv_final_scale(J.e_sv);
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float n2 = 0.0f;
#pragma unroll
    for (int j = 0; j < H; ++j)
      if (out_pos[j] >= 0) n2 = fmaf(t[j].x, t[j].x, fmaf(t[j].y, t[j].y, n2));
    scale = 1.0f / sqrtf(n2);
  }
  const float2 f = cscale(ph, scale);
  const int lane = threadIdx.x & 63;
  const int64_t g0 = g - lane;
  if constexpr (STORE == kStoreBlock) {
    static_assert(H <= 16, "block staging supports at most 16 amplitudes");
    __shared__ float2 stage[4][64 * (H + 1)];
    const int dout = 1 << J.n_out;
    float2* sw = stage[threadIdx.x >> 6];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int p = out_pos[j];
      if (p >= 0) sw[lane * (dout + 1) + p] = cmul(t[j], f);
    }
    __syncthreads();
    float2* obase = J.out + g0 * dout;
    const int shift = J.n_out;
    for (int e = lane; e < 64 * dout; e += 64) {
      const int r = e >> shift, c = e & (dout - 1);
      if (r < rows_valid) obase[e] = sw[r * (dout + 1) + c];
    }
  } else if constexpr (STORE == kStoreColumns) {
    wave_store_rows<NS>(t, out_pos, f, J.out + g0 * (int64_t(1) << J.n_out), J.n_out, rows_valid);
  } else {
    if (active) {
      float2* o = J.out + g * (int64_t(1) << J.n_out);
#pragma unroll
      for (int j = 0; j < H; ++j) {
        const int p = out_pos[j];
        if (p >= 0) o[p] = cmul(t[j], f);
      }
    }
  }
  if (active && (flags & MBQC_FLAG_WRITE_PHASE) && J.phase_out) J.phase_out[g] = ph;
}

// Output store per live-window size (measured on 1M-simulation batches):
// block staging up to 16 amplitudes; column passes for 32 (5x6 grid:
// 0.44 -> 0.20 ms, 5x30: 1.07 -> 0.75 ms); direct stores for 64 (staging
// was slower there).
template <int NS>
constexpr int kForwardStore = NS <= 4 ? kStoreBlock : (NS == 5 ? kStoreColumns : kStoreDirect);

template <int NS, bool REF, bool PER_STEP>
__global__ __launch_bounds__(256) void k_sv_forward(Job J, uint32_t flags) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t g0 = g - (threadIdx.x & 63);
  const int rows = int(J.total - g0 < 64 ? J.total - g0 : 64);
  sv_sim<NS, REF, kForwardStore<NS>, PER_STEP>(J, g, g < J.total, rows, flags);
}

// k_sv_forward for periodic plans (grids in column order, chains).
template <int NS, bool REF>
__global__ __launch_bounds__(256) void k_sv_forward_periodic(Job J, uint32_t flags) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t g0 = g - (threadIdx.x & 63);
  const int rows = int(J.total - g0 < 64 ? J.total - g0 : 64);
  sv_sim<NS, REF, kForwardStore<NS>, false, false, true>(J, g, g < J.total, rows, flags);
}

// ---------------------------------------------------------------- DM step
// Quartet form of  rho -> Tr_Q[(D^dag P_theta) CZ (rho (x) |+><+|) CZ]:
// with B_ab the blocks of rho where slot Q is a (row) / b (column), and
// gb = gamma e^{-i theta} (gamma = coherence damping of the measured qubit),
//   P = B00 + B11, M = B00 - B11, U = conj(gb) B01, V = gb B10,
//   r00 = P+U+V, r01 = M-U+V, r10 = M+U-V, r11 = P-U-V   (CZ to the measured qubit)
//   r_cd = P+U+V                                       (no CZ to the measured qubit)
// new[(c,j),(d,k)] = sign_j^c sign_k^d r_cd.  The factor 1/2 of |+><+| and
// the division by the outcome probability (1/2) are tracked as an exponent.
// Hermitian H x H matrix in registers, stored as its upper triangle
// (row-major, a <= b); an entry below the diagonal is the conjugate of its
// mirror.  Every index is a compile-time constant in the unrolled loops, so
// the choice between the stored entry and its conjugate costs nothing.
template <int H>
struct HermReg {
  static constexpr int N = H * (H + 1) / 2;
  float2 u[N];
  static constexpr int idx(int a, int b) { return a * H - (a * (a - 1)) / 2 + (b - a); }
  __device__ __forceinline__ float2 get(int a, int b) const {
    return a <= b ? u[idx(a, b)] : cconj(u[idx(b, a)]);
  }
  __device__ __forceinline__ void set(int a, int b, float2 v) {
    if (a <= b)
      u[idx(a, b)] = v;
    else
      u[idx(b, a)] = cconj(v);
  }
};

// One DM step on slot Q (the quartet update of the measured qubit with the
// appended |+><+| and its CZ signs), on the quartets (j, k) with j <= k only:
// the others are their conjugate transposes.
template <int NS, int Q>
__device__ __forceinline__ void dm_step(HermReg<1 << NS>& r, float2 gb, uint32_t flags,
                                        uint32_t nmask) {
  constexpr int H = 1 << NS;
  constexpr int QB = 1 << Q;
  const float2 gbc = cconj(gb);
  const bool app = flags & 1u, sneg = flags & 2u;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & QB) continue;
#pragma unroll
    for (int k = j; k < H; ++k) {
      if (k & QB) continue;
      const float2 B00 = r.get(j, k), B01 = r.get(j, k | QB);
      const float2 B10 = r.get(j | QB, k), B11 = r.get(j | QB, k | QB);
      const float2 U = cmul(gbc, B01), V = cmul(gb, B10);
      const float2 P = cadd(B00, B11);
      const float2 r00 = cadd(P, cadd(U, V));
      float2 n01, n10, n11;
      if (app) {
        const uint32_t sj = czbit(j, nmask), sk = czbit(k, nmask);
        if (sneg) {
          const float2 M = csub(B00, B11);
          n01 = flip(cadd(M, csub(V, U)), sk);
          n10 = flip(cadd(M, csub(U, V)), sj);
          n11 = flip(csub(P, cadd(U, V)), sj ^ sk);
        } else {
          n01 = flip(r00, sk);
          n10 = flip(r00, sj);
          n11 = flip(r00, sj ^ sk);
        }
      } else {
        n01 = n10 = n11 = make_float2(0.f, 0.f);
      }
      r.set(j, k, r00);
      r.set(j, k | QB, n01);
      if (j != k) r.set(j | QB, k, n10);   // for j == k it is the conjugate of n01
      r.set(j | QB, k | QB, n11);
    }
  }
}

// Single-qubit noise of the density-matrix path.  Every supported channel is
// phase covariant: on a qubit's 2x2 blocks it maps the diagonal pair
// (r00, r11) -> (d00 r00 + d01 r11, d10 r00 + d11 r11) and scales both
// off-diagonal blocks by the real factor coh.  A measured qubit only shows
// the trace (preserved) and coh to its XY projection, so its noise (and the
// readout flip) folds into gamma = coh (1 - 2 p_readout); output qubits get the
// full map at the end.
struct DmNoise {
  float gamma;
  float d00, d01, d10, d11, coh;
  int32_t out_noise;  // 0: the output map is the identity
  float flip;         // 1 - 2 p_readout
};

template <int NS, int M>
__device__ __forceinline__ void dm_channel(HermReg<1 << NS>& r, const DmNoise& nz) {
  constexpr int H = 1 << NS;
  constexpr int MB = 1 << M;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & MB) continue;
#pragma unroll
    for (int k = j; k < H; ++k) {
      if (k & MB) continue;
      const float2 a = r.get(j, k), d = r.get(j | MB, k | MB);
      r.set(j, k, make_float2(fmaf(nz.d00, a.x, nz.d01 * d.x), fmaf(nz.d00, a.y, nz.d01 * d.y)));
      r.set(j | MB, k | MB,
            make_float2(fmaf(nz.d10, a.x, nz.d11 * d.x), fmaf(nz.d10, a.y, nz.d11 * d.y)));
      r.set(j, k | MB, cscale(r.get(j, k | MB), nz.coh));
      if (j != k) r.set(j | MB, k, cscale(r.get(j | MB, k), nz.coh));
    }
  }
}

// PERIODIC: the step flags of a periodic plan (append, CZ to the measured
// qubit, neighbour mask periodic_nmask) as compile-time constants, so the CZ
// signs of the quartets fold into the arithmetic.
template <int NS, bool PERIODIC, int SUB>
__device__ __forceinline__ void dm_sub(HermReg<1 << NS>& r, const mbqc_step* steps, int k0,
                                       int n_steps, const float (&th)[NS], float gamma) {
  const int k = k0 + SUB;
  if (k < n_steps) {
    const float2 b = step_coeff<false>(th[SUB], nullptr);
    if constexpr (PERIODIC)
      dm_step<NS, NS - 1 - SUB>(r, cscale(b, gamma), 3u, periodic_nmask(NS, SUB));
    else
      dm_step<NS, NS - 1 - SUB>(r, cscale(b, gamma), steps[k].flags, steps[k].nmask);
  }
}

template <int NS, bool PERIODIC, int... SUBS>
__device__ __forceinline__ void dm_block(HermReg<1 << NS>& r, const mbqc_step* steps, int k0,
                                         int n_steps, const float (&th)[NS], float gamma,
                                         std::integer_sequence<int, SUBS...>) {
  (dm_sub<NS, PERIODIC, SUBS>(r, steps, k0, n_steps, th, gamma), ...);
}

template <int NS, int... MS>
__device__ __forceinline__ void dm_output_noise(HermReg<1 << NS>& r, uint32_t live,
                                                const DmNoise& nz, std::integer_sequence<int, MS...>) {
  ((live & (1u << MS) ? dm_channel<NS, MS>(r, nz) : void()), ...);
}

template <int NS, bool PERIODIC>
__global__ __launch_bounds__(256) void k_dm_forward(Job J, DmNoise nz, uint32_t flags) {
  constexpr int H = 1 << NS;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = g < J.total;
  const int64_t gg = active ? g : 0;
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* resc = J.tab + J.off_dmres;
  float2 v[H];
  sv_load_input<NS>(J, J.psi + si * J.in_stride, v);
  HermReg<H> r;
#pragma unroll
  for (int j = 0; j < H; ++j)
#pragma unroll
    for (int k = j; k < H; ++k) r.set(j, k, cmul(v[j], cconj(v[k])));
  const int n = J.n_steps;
  const int nblk = (n + NS - 1) / NS;
  for (int blk = 0; blk < nblk; ++blk) {
    if (resc[blk]) {
#pragma unroll
      for (int e = 0; e < HermReg<H>::N; ++e) r.u[e] = cscale(r.u[e], 0x1p-32f);
    }
    float th[NS];
    load_block_angles<NS>(steps, blk * NS, n, ang, J.angle_stride_a, th);
    dm_block<NS, PERIODIC>(r, steps, blk * NS, n, th, nz.gamma,
                           std::make_integer_sequence<int, NS>{});
  }
  if (nz.out_noise)
    dm_output_noise<NS>(r, uint32_t(J.live_mask), nz, std::make_integer_sequence<int, NS>{});
  const int32_t* out_pos = J.tab + J.off_out;
  float scale = ldexpf(1.0f, -J.e_dm);
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float tr = 0.0f;
#pragma unroll
    for (int j = 0; j < H; ++j)
      if (out_pos[j] >= 0) tr += r.get(j, j).x;
    scale = 1.0f / tr;
  }
  if (active) {
    const int64_t d = int64_t(1) << J.n_out;
    float2* o = J.out + g * d * d;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int pj = out_pos[j];
      if (pj < 0) continue;
#pragma unroll
      for (int k = 0; k < H; ++k) {
        const int pk = out_pos[k];
        if (pk >= 0) o[int64_t(pj) * d + pk] = cscale(r.get(j, k), scale);
      }
    }
  }
}

// ------------------------------------------------------- sampled outcomes
// 32 pseudo-random bits per (seed, simulation, block of 32 steps): the
// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014) applied to a Weyl-sequence key.
__device__ __forceinline__ uint32_t draw32(uint64_t seed, uint64_t sim, uint64_t block) {
  uint64_t z = seed + sim * 0x9E3779B97F4A7C15ull + (block + 1) * 0xD1B54A32D192ED03ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return uint32_t(z ^ (z >> 31));
}

// Per-lane Pauli frame and outcome draw of the sampled mode.  Before step k,
// bit 0 of fx / fz is the X / Z frame of schedule[k]: the node is measured at
// (-1)^x theta, the drawn projection m' (probability 1/2 each) is onto
// |+/- theta'>, and the physical outcome is m' xor z.  An outcome 1 flips the
// frames of the flow successor and of its later neighbours (xmask/zmask).
// Frames are 32-bit: a plan that samples applies every CZ edge within its
// window of at most 7 positions, so every offset is below 14 (the host
// checks < 32).
struct SampleMod {
  uint32_t fx = 0, fz = 0;
  uint64_t seed = 0, sim = 0;
  const uint64_t* xm;
  const uint64_t* zm;
  int32_t* out_words;   // this simulation's outcome words (null for inactive lanes)
  int n;
  uint32_t rbits = 0, word = 0;
  __device__ __forceinline__ float2 operator()(int k, float2 b) {
    const int kb = k & 31;
    if (kb == 0) rbits = draw32(seed, sim, uint64_t(k >> 5));
    const uint32_t mprime = (rbits >> kb) & 1u;
    const uint32_t xk = fx & 1u, zk = fz & 1u;
    const uint32_t mk = mprime ^ zk;
    word |= mk << kb;
    if (kb == 31 || k == n - 1) {
      if (out_words) out_words[k >> 5] = int32_t(word);
      word = 0;
    }
    const uint32_t sel = 0u - mk;
    fx = (fx ^ (uint32_t(xm[k]) & sel)) >> 1;
    fz = (fz ^ (uint32_t(zm[k]) & sel)) >> 1;
    float2 bb = xk ? cconj(b) : b;
    return mprime ? make_float2(-bb.x, -bb.y) : bb;
  }
};

template <int NS, bool PERIODIC>
__global__ __launch_bounds__(256) void k_sv_sample(Job J, const uint64_t* __restrict__ xm,
                                                   const uint64_t* __restrict__ zm,
                                                   const int32_t* __restrict__ out_off,
                                                   int32_t* __restrict__ outcomes, int32_t words,
                                                   uint64_t seed) {
  constexpr int H = 1 << NS;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = g < J.total;
  const int64_t gg = active ? g : 0;
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  float2 t[H];
  sv_load_input<NS>(J, J.psi + si * J.in_stride, t);
  RefPhase ph;
  SampleMod mod;
  mod.seed = seed;
  mod.sim = uint64_t(gg);
  mod.xm = xm;
  mod.zm = zm;
  mod.n = J.n_steps;
  int32_t* row = outcomes + gg * (words + 2);
  mod.out_words = active ? row : nullptr;
  sv_run_steps<NS, false, PERIODIC>(t, J, ang, &ph, [](int, const float2 (&)[H]) {}, mod);
  // residual frame of the output qubits, in output-index bit order
  uint32_t xo = 0, zo = 0;
  for (int m = 0; m < J.n_out; ++m) {
    const int off = out_off[m];
    xo |= ((mod.fx >> off) & 1u) << (J.n_out - 1 - m);
    zo |= ((mod.fz >> off) & 1u) << (J.n_out - 1 - m);
  }
  // all output positions first: the outcome stores above keep the compiler
  // from moving these table loads across the output stores below
  const int32_t* out_pos = J.tab + J.off_out;
  int pos[H];
#pragma unroll
  for (int j = 0; j < H; ++j) pos[j] = out_pos[j];
  float n2 = 0.0f;
#pragma unroll
  for (int j = 0; j < H; ++j)
    if (pos[j] >= 0) n2 = fmaf(t[j].x, t[j].x, fmaf(t[j].y, t[j].y, n2));
  const float scale = 1.0f / sqrtf(n2);
  // the output frame Z^z X^x applied while storing:
  // new[o] = (-1)^{popcount(o & z)} old[o ^ x], i.e. old[p] lands at p ^ x
  if constexpr (H <= 16) {
    // outputs of up to 16 amplitudes leave through LDS as contiguous rows
    // (the workgroup's waves cover consecutive simulations)
    __shared__ float2 stage[4][64 * 17];
    const int lane = threadIdx.x & 63;
    const int dout = 1 << J.n_out;
    float2* sw = stage[threadIdx.x >> 6];
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int p = pos[j];
      if (p >= 0) {
        const uint32_t q = uint32_t(p) ^ xo;
        sw[lane * (dout + 1) + int(q)] = flip(cscale(t[j], scale), czbit(q, zo));
      }
    }
    __syncthreads();
    const int64_t g0 = g - lane;
    const int rows = int(J.total - g0 < 64 ? J.total - g0 : 64);
    float2* obase = J.out + g0 * dout;
    for (int e = lane; e < 64 * dout; e += 64) {
      const int r = e >> J.n_out, c = e & (dout - 1);
      if (r < rows) obase[e] = sw[r * (dout + 1) + c];
    }
  } else if (active) {
    float2* o = J.out + g * (int64_t(1) << J.n_out);
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int p = pos[j];
      if (p >= 0) {
        const uint32_t q = uint32_t(p) ^ xo;
        o[q] = flip(cscale(t[j], scale), czbit(q, zo));
      }
    }
  }
  if (active) {
    row[words] = int32_t(xo);
    row[words + 1] = int32_t(zo);
  }
}

// ------------------------------------------------------- LDS-resident kernels
// Larger live windows (SV with 7..12 slots, DM with 4..6 slots): one
// workgroup per simulation keeps the state in LDS; the measured slot is a
// runtime value, each thread updates a strided subset of the amplitude pairs
// (DM: quartets) and the workgroup synchronises between steps.
// Workgroups are dealt to the chip's kXcds XCDs round-robin (workgroup b runs
// on XCD b % kXcds), each XCD with its own L2.  xcd_block maps workgroup b of
// n to a logical block so that every XCD takes a contiguous run of logical
// blocks (a bijection of [0, n)): neighbouring simulations, which share cache
// lines of the angle table, then meet in one L2 instead of each XCD fetching
// the line.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned n) {
  const unsigned q = n / kXcds, r = n % kXcds, x = b % kXcds;
  return x * q + (x < r ? x : r) + b / kXcds;
}

__device__ __forceinline__ int insert_zero(int v, int q) {
  return ((v >> q) << (q + 1)) | (v & ((1 << q) - 1));
}

// Sum over the T threads of a workgroup (T a power of two); a workgroup of
// fewer than 64 threads only exchanges among its own lanes.
template <int T>
__device__ __forceinline__ float block_sum(float v, float* red) {
  constexpr int first = T >= 64 ? 32 : T / 2;
  for (int m = first; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  if constexpr (T > 64) {
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    v = 0.0f;
#pragma unroll
    for (int i = 0; i < T / 64; ++i) v += red[i];
  }
  return v;
}

// Angles of the workgroup's simulation for steps [k0, k0 + kAngleChunk), read
// by all threads into LDS at once, so a step does not wait on its own global
// load (the caller synchronises before and after).
constexpr int kAngleChunk = 32;

template <int T>
__device__ __forceinline__ void stage_angles(float* th_s, const mbqc_step* steps, const float* ang,
                                             int64_t sa, int k0, int n_steps, int tid) {
  for (int i = tid; i < kAngleChunk; i += T) {
    const int k = k0 + i;
    if (k < n_steps) {
      const mbqc_step st = steps[k];
      th_s[i] = st.angle_index >= 0 ? ang[int64_t(st.angle_index) * sa] : st.fixed_angle;
    }
  }
}

// The projection coefficients gamma e^{-i theta} of steps [k0, k0 + kAngleChunk),
// computed once per step by the workgroup's threads in parallel and kept in
// LDS, so the sine and cosine are not on every step's dependent chain (the
// caller synchronises before and after).
template <int T>
__device__ __forceinline__ void stage_coeffs(float2* gb_s, const mbqc_step* steps, const float* ang,
                                             int64_t sa, int k0, int n_steps, float gamma, int tid) {
  for (int i = tid; i < kAngleChunk; i += T) {
    const int k = k0 + i;
    if (k < n_steps) {
      const mbqc_step st = steps[k];
      const float th = st.angle_index >= 0 ? ang[int64_t(st.angle_index) * sa] : st.fixed_angle;
      gb_s[i] = cscale(step_coeff<false>(th, nullptr), gamma);
    }
  }
}

// One step on an LDS-resident statevector of H amplitudes, the pairs split
// over the T threads of the workgroup (caller synchronises before and after).
template <int T>
__device__ __forceinline__ void sv_lds_step(float2* st, int H, const mbqc_step& step, float2 b,
                                            int tid) {
  const int q = int((step.flags >> 8) & 0xffu);
  const int QB = 1 << q;
  const float s = (step.flags & 2u) ? -1.0f : 1.0f;
  const float app = (step.flags & 1u) ? 1.0f : 0.0f;
  for (int i = tid; i < H / 2; i += T) {
    const int j = insert_zero(i, q);
    const float2 x = st[j], y = st[j | QB];
    const float c = (__builtin_popcount(unsigned(j) & step.nmask) & 1) ? -app : app;
    const float al = c * s, be = c * (1.0f - s);
    float2 u0;
    u0.x = fmaf(-b.y, y.y, fmaf(b.x, y.x, x.x));
    u0.y = fmaf(b.x, y.y, fmaf(b.y, y.x, x.y));
    st[j] = u0;
    st[j | QB] = make_float2(fmaf(al, u0.x, be * x.x), fmaf(al, u0.y, be * x.y));
  }
}

// One step on an LDS-resident density matrix (H x H) measuring slot q, one
// quartet of the measured qubit per thread iteration; gb = gamma e^{-i theta}.
// Called with a compile-time q (block sub-steps), the quartet addresses depend
// on the thread only.
template <int T>
__device__ __forceinline__ void dm_lds_step(float2* r, int H, int q, uint32_t flags,
                                            uint32_t nmask, float2 gb, int tid) {
  const int HQ = H / 2;
  const float2 gbc = cconj(gb);
  const int QB = 1 << q;
  const bool app = flags & 1u, sneg = flags & 2u;
  for (int i = tid; i < HQ * HQ; i += T) {
    const int j = insert_zero(i / HQ, q), kk = insert_zero(i % HQ, q);
    const float2 B00 = r[j * H + kk], B01 = r[j * H + (kk | QB)];
    const float2 B10 = r[(j | QB) * H + kk], B11 = r[(j | QB) * H + (kk | QB)];
    const float2 U = cmul(gbc, B01), V = cmul(gb, B10);
    const float2 P = cadd(B00, B11);
    const float2 r00 = cadd(P, cadd(U, V));
    if (app) {
      const uint32_t sj = czbit(j, nmask), sk = czbit(kk, nmask);
      if (sneg) {
        const float2 M = csub(B00, B11);
        r[j * H + kk] = r00;
        r[j * H + (kk | QB)] = flip(cadd(M, csub(V, U)), sk);
        r[(j | QB) * H + kk] = flip(cadd(M, csub(U, V)), sj);
        r[(j | QB) * H + (kk | QB)] = flip(csub(P, cadd(U, V)), sj ^ sk);
      } else {
        r[j * H + kk] = r00;
        r[j * H + (kk | QB)] = flip(r00, sk);
        r[(j | QB) * H + kk] = flip(r00, sj);
        r[(j | QB) * H + (kk | QB)] = flip(r00, sj ^ sk);
      }
    } else {
      r[j * H + kk] = r00;
      r[j * H + (kk | QB)] = make_float2(0.f, 0.f);
      r[(j | QB) * H + kk] = make_float2(0.f, 0.f);
      r[(j | QB) * H + (kk | QB)] = make_float2(0.f, 0.f);
    }
  }
}

// Simulation g of job J on the workgroup's LDS state st (2^NS amplitudes),
// with red (4 floats) and th_s (kAngleChunk floats) as LDS scratch.
template <int NS, int T>
__device__ __forceinline__ void sv_lds_run(const Job& J, int64_t g, uint32_t flags, float2* st,
                                           float* red, float* th_s, int tid) {
  constexpr int H = 1 << NS;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float2* psi = J.psi + si * J.in_stride;
  for (int j = tid; j < H; j += T) st[j] = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
  RefPhase rp;
  const bool ref = flags & MBQC_FLAG_REF_PHASE;
  for (int k = 0; k < J.n_steps; ++k) {
    if (k % NS == 0 && resc[k / NS]) {
      __syncthreads();
      for (int j = tid; j < H; j += T) st[j] = cscale(st[j], 0x1p-32f);
    }
    if (k % kAngleChunk == 0) {
      __syncthreads();
      stage_angles<T>(th_s, steps, ang, J.angle_stride_a, k, J.n_steps, tid);
      __syncthreads();
    }
    const mbqc_step step = steps[k];
    const float th = th_s[k % kAngleChunk];
    const float2 b = ref ? step_coeff<true>(th, &rp) : step_coeff<false>(th, nullptr);
    __syncthreads();
    sv_lds_step<T>(st, H, step, b, tid);
  }
  __syncthreads();
  const float2 ph = ref ? rp.value() : make_float2(1.0f, 0.0f);
  float scale = sv_final_scale(J.e_sv);
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float n2 = 0.0f;
    for (int j = tid; j < H; j += T)
      if (out_pos[j] >= 0) n2 = fmaf(st[j].x, st[j].x, fmaf(st[j].y, st[j].y, n2));
    n2 = block_sum<T>(n2, red);
    scale = 1.0f / sqrtf(n2);
  }
  const float2 f = cscale(ph, scale);
  float2* o = J.out + g * (int64_t(1) << J.n_out);
  for (int j = tid; j < H; j += T) {
    const int p = out_pos[j];
    if (p >= 0) o[p] = cmul(st[j], f);
  }
  if (tid == 0 && (flags & MBQC_FLAG_WRITE_PHASE) && J.phase_out) J.phase_out[g] = ph;
}

template <int NS, int T>
__global__ __launch_bounds__(T) void k_sv_forward_lds(Job J, uint32_t flags) {
  __shared__ float2 st[1 << NS];
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  sv_lds_run<NS, T>(J, xcd_block(blockIdx.x, gridDim.x), flags, st, red, th_s, threadIdx.x);
}

// Windows above the LDS: the same simulation on statevectors in a global
// workspace, ws[(g - g0) * 2^NS + j] for simulations g = g0 + blockIdx (each
// workgroup its own slice, synchronised by the barrier as in LDS).
template <int NS>
__global__ __launch_bounds__(256) void k_sv_forward_global(Job J, uint32_t flags,
                                                           float2* __restrict__ ws, int64_t g0) {
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  const int64_t l = xcd_block(blockIdx.x, gridDim.x);
  sv_lds_run<NS, 256>(J, g0 + l, flags, ws + (l << NS), red, th_s, threadIdx.x);
}

// Sampled outcomes for windows of 7..14 slots: k_sv_sample's frame and draw
// arithmetic (SampleMod, evaluated identically by every thread of the
// workgroup, so the frame is uniform) on the LDS-resident state of
// k_sv_forward_lds.  Raw phase; the output is renormalised and leaves with the
// output frame applied, as in k_sv_sample.
template <int NS, int T>
__device__ __forceinline__ void sample_run(const Job& J, const uint64_t* __restrict__ xm,
                                           const uint64_t* __restrict__ zm,
                                           const int32_t* __restrict__ out_off,
                                           int32_t* __restrict__ outcomes, int32_t words,
                                           uint64_t seed, float2* st, float* red, float* th_s,
                                           int64_t g, int tid) {
  constexpr int H = 1 << NS;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float2* psi = J.psi + si * J.in_stride;
  for (int j = tid; j < H; j += T) st[j] = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
  int32_t* row = outcomes + g * (words + 2);
  SampleMod mod;
  mod.seed = seed;
  mod.sim = uint64_t(g);
  mod.xm = xm;
  mod.zm = zm;
  mod.n = J.n_steps;
  mod.out_words = tid == 0 ? row : nullptr;
  for (int k = 0; k < J.n_steps; ++k) {
    if (k % NS == 0 && resc[k / NS]) {
      __syncthreads();
      for (int j = tid; j < H; j += T) st[j] = cscale(st[j], 0x1p-32f);
    }
    if (k % kAngleChunk == 0) {
      __syncthreads();
      stage_angles<T>(th_s, steps, ang, J.angle_stride_a, k, J.n_steps, tid);
      __syncthreads();
    }
    const mbqc_step step = steps[k];
    const float2 b = mod(k, step_coeff<false>(th_s[k % kAngleChunk], nullptr));
    __syncthreads();
    sv_lds_step<T>(st, H, step, b, tid);
  }
  __syncthreads();
  uint32_t xo = 0, zo = 0;
  for (int m = 0; m < J.n_out; ++m) {
    const int off = out_off[m];
    xo |= ((mod.fx >> off) & 1u) << (J.n_out - 1 - m);
    zo |= ((mod.fz >> off) & 1u) << (J.n_out - 1 - m);
  }
  float n2 = 0.0f;
  for (int j = tid; j < H; j += T)
    if (out_pos[j] >= 0) n2 = fmaf(st[j].x, st[j].x, fmaf(st[j].y, st[j].y, n2));
  n2 = block_sum<T>(n2, red);
  const float scale = 1.0f / sqrtf(n2);
  float2* o = J.out + g * (int64_t(1) << J.n_out);
  for (int j = tid; j < H; j += T) {
    const int p = out_pos[j];
    if (p >= 0) {
      const uint32_t q = uint32_t(p) ^ xo;
      o[q] = flip(cscale(st[j], scale), czbit(q, zo));
    }
  }
  if (tid == 0) {
    row[words] = int32_t(xo);
    row[words + 1] = int32_t(zo);
  }
}

template <int NS, int T>
__global__ __launch_bounds__(T) void k_sv_sample_lds(Job J, const uint64_t* __restrict__ xm,
                                                     const uint64_t* __restrict__ zm,
                                                     const int32_t* __restrict__ out_off,
                                                     int32_t* __restrict__ outcomes, int32_t words,
                                                     uint64_t seed) {
  __shared__ float2 st[1 << NS];
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  sample_run<NS, T>(J, xm, zm, out_off, outcomes, words, seed, st, red, th_s,
                    xcd_block(blockIdx.x, gridDim.x), threadIdx.x);
}

// Windows above the LDS: the same with the state in a global workspace.
template <int NS>
__global__ __launch_bounds__(256) void k_sv_sample_global(Job J, const uint64_t* __restrict__ xm,
                                                          const uint64_t* __restrict__ zm,
                                                          const int32_t* __restrict__ out_off,
                                                          int32_t* __restrict__ outcomes,
                                                          int32_t words, uint64_t seed,
                                                          float2* __restrict__ ws, int64_t g0) {
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  const int64_t l = xcd_block(blockIdx.x, gridDim.x);
  sample_run<NS, 256>(J, xm, zm, out_off, outcomes, words, seed, ws + (l << NS), red, th_s, g0 + l,
                      threadIdx.x);
}

// Sub-step SUB of the block starting at step k0 (measured slot NS - 1 - SUB);
// c0 = k0's position in the staged coefficient chunk.
template <int NS, int T, bool PERIODIC, int SUB>
__device__ __forceinline__ void dm_lds_sub(float2* r, const mbqc_step* steps, const float2* gb_s,
                                           int k0, int c0, int n, int tid) {
  const int k = k0 + SUB;
  if (k < n) {
    const float2 gb = gb_s[c0 + SUB];
    if constexpr (PERIODIC) {
      __syncthreads();
      dm_lds_step<T>(r, 1 << NS, NS - 1 - SUB, 3u, periodic_nmask(NS, SUB), gb, tid);
    } else {
      const mbqc_step st = steps[k];
      __syncthreads();
      dm_lds_step<T>(r, 1 << NS, NS - 1 - SUB, st.flags, st.nmask, gb, tid);
    }
  }
}

template <int NS, int T, bool PERIODIC, int... SUBS>
__device__ __forceinline__ void dm_lds_block(float2* r, const mbqc_step* steps, const float2* gb_s,
                                             int k0, int c0, int n, int tid,
                                             std::integer_sequence<int, SUBS...>) {
  (dm_lds_sub<NS, T, PERIODIC, SUBS>(r, steps, gb_s, k0, c0, n, tid), ...);
}

// Density matrix in LDS: T threads per simulation (one quartet each per step)
// and G simulations per workgroup, so that small windows (T < 64) still fill
// whole waves.  Steps run in blocks of NS with the measured slot fixed per
// sub-step; coefficients are staged in chunks of a multiple of NS steps.
template <int NS, int T, int G, bool PERIODIC>
__global__ __launch_bounds__(T * G) void k_dm_forward_lds(Job J, DmNoise nz, uint32_t flags) {
  constexpr int H = 1 << NS;
  constexpr int HQ = H / 2;
  __shared__ float2 rs[G][H * H];
  __shared__ float2 gb_all[G][kAngleChunk];
  __shared__ float red[4];
  const int gs = threadIdx.x / T, tid = threadIdx.x % T;
  const int64_t g = int64_t(xcd_block(blockIdx.x, gridDim.x)) * G + gs;
  const bool active = g < J.total;
  const int64_t gg = active ? g : J.total - 1;
  float2* r = rs[gs];
  float2* gb_s = gb_all[gs];
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_dmres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float2* psi = J.psi + si * J.in_stride;
  for (int e = tid; e < H * H; e += T) {
    const int j = e / H, k = e % H;
    const float2 vj = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
    const float2 vk = flip(psi[in_index[k]], sign[k] ? 0x80000000u : 0u);
    r[e] = cmul(vj, cconj(vk));
  }
  constexpr int CH = (kAngleChunk / NS) * NS;
  const int n = J.n_steps;
  for (int k0 = 0; k0 < n; k0 += NS) {
    if (resc[k0 / NS]) {
      __syncthreads();
      for (int e = tid; e < H * H; e += T) r[e] = cscale(r[e], 0x1p-32f);
    }
    const int c0 = k0 % CH;
    if (c0 == 0) {
      __syncthreads();
      stage_coeffs<T>(gb_s, steps, ang, J.angle_stride_a, k0, n, nz.gamma, tid);
      __syncthreads();
    }
    dm_lds_block<NS, T, PERIODIC>(r, steps, gb_s, k0, c0, n, tid,
                                  std::make_integer_sequence<int, NS>{});
  }
  // noise of every output qubit
  for (int m = 0; m < NS; ++m) {
    if (!(uint32_t(J.live_mask) & (1u << m)) || !nz.out_noise) continue;
    const int MB = 1 << m;
    __syncthreads();
    for (int i = tid; i < HQ * HQ; i += T) {
      const int j = insert_zero(i / HQ, m), kk = insert_zero(i % HQ, m);
      const float2 a = r[j * H + kk], d = r[(j | MB) * H + (kk | MB)];
      r[j * H + kk] = make_float2(fmaf(nz.d00, a.x, nz.d01 * d.x), fmaf(nz.d00, a.y, nz.d01 * d.y));
      r[(j | MB) * H + (kk | MB)] =
          make_float2(fmaf(nz.d10, a.x, nz.d11 * d.x), fmaf(nz.d10, a.y, nz.d11 * d.y));
      r[j * H + (kk | MB)] = cscale(r[j * H + (kk | MB)], nz.coh);
      r[(j | MB) * H + kk] = cscale(r[(j | MB) * H + kk], nz.coh);
    }
  }
  __syncthreads();
  float scale = ldexpf(1.0f, -J.e_dm);
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float tr = 0.0f;
    for (int j = tid; j < H; j += T)
      if (out_pos[j] >= 0) tr += r[j * H + j].x;
    tr = block_sum<T>(tr, red);      // T < 64: sums within this simulation's lanes
    scale = 1.0f / tr;
  }
  if (!active) return;
  const int64_t d = int64_t(1) << J.n_out;
  float2* o = J.out + g * d * d;
  for (int e = tid; e < H * H; e += T) {
    const int pj = out_pos[e / H], pk = out_pos[e % H];
    if (pj >= 0 && pk >= 0) o[int64_t(pj) * d + pk] = cscale(r[e], scale);
  }
}

// ------------------------------------------- DM, one entry per lane
// Value of lane (lane ^ MASK) in the same wave: DPP quad permutations for
// MASK 1 and 2, ds_swizzle (no LDS allocation, no barrier) within 32-lane
// halves, a bpermute shuffle across them.
template <int MASK>
__device__ __forceinline__ float lane_xor(float v) {
  if constexpr (MASK == 1) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xf, 0xf, false));
  } else if constexpr (MASK == 2) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xf, 0xf, false));
  } else if constexpr (MASK < 32) {
    // bit mode: lane' = ((lane & 0x1f) ^ MASK) within each 32-lane half
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), (MASK << 10) | 0x1f));
  } else {
    return __shfl_xor(v, MASK, 64);
  }
}
template <int MASK>
__device__ __forceinline__ float2 lane_xor2(float2 v) {
  return make_float2(lane_xor<MASK>(v.x), lane_xor<MASK>(v.y));
}

// One step of the density matrix held one entry per lane: lane entry
// (a, b) of an H x H matrix, H = 2^NS, lanes numbered a * H + b within the
// simulation's group of H^2 lanes.  The quartet of measured slot Q that holds
// the entry is {(a,b), (a,b^Q), (a^Q,b), (a^Q,b^Q)}: column, row and diagonal
// partners are lanes xor Q, xor Q<<NS and both.  Same update as dm_lds_step
// (quartet form of Tr_Q[(D^dag P) CZ (rho (x) |+><+|) CZ], gb = gamma e^{-i theta}).
template <int NS, int Q>
__device__ __forceinline__ float2 dm_wave_step(float2 x, int a, int b, float2 gb, uint32_t flags,
                                               uint32_t nmask) {
  constexpr int QB = 1 << Q;
  const float2 c = lane_xor2<QB>(x);           // (a, b^Q)
  const float2 r = lane_xor2<(QB << NS)>(x);   // (a^Q, b)
  const float2 d = lane_xor2<QB>(r);           // (a^Q, b^Q)
  const bool ra = (a >> Q) & 1, rb = (b >> Q) & 1;
  // the quartet's entries by position
  const float2 b00 = ra ? (rb ? d : r) : (rb ? c : x);
  const float2 b01 = ra ? (rb ? r : d) : (rb ? x : c);
  const float2 b10 = ra ? (rb ? c : x) : (rb ? d : r);
  const float2 b11 = ra ? (rb ? x : c) : (rb ? r : d);
  const float2 U = cmul(cconj(gb), b01), V = cmul(gb, b10);
  const float2 P = cadd(b00, b11);
  const float2 r00 = cadd(P, cadd(U, V));
  if (!(flags & 1u)) return (ra | rb) ? make_float2(0.f, 0.f) : r00;
  // nmask holds no bit of slot Q, so the row / column CZ signs are those of a / b
  const uint32_t sa = czbit(uint32_t(a), nmask), sb = czbit(uint32_t(b), nmask);
  if (flags & 2u) {
    const float2 M = csub(b00, b11);
    if (!ra && !rb) return r00;
    if (!ra) return flip(cadd(M, csub(V, U)), sb);
    if (!rb) return flip(cadd(M, csub(U, V)), sa);
    return flip(csub(P, cadd(U, V)), sa ^ sb);
  }
  return flip(r00, (ra ? sa : 0u) ^ (rb ? sb : 0u));
}

// Coefficient of chunk entry i of this simulation's group (lanes gbase .. gbase+E-1).
template <int E>
__device__ __forceinline__ float2 group_coeff(float2 cf, int gbase, int i) {
  if constexpr (E == 64) {
    return make_float2(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(cf.x), i)),
                       __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cf.y), i)));
  } else {
    return make_float2(__shfl(cf.x, gbase + i, 64), __shfl(cf.y, gbase + i, 64));
  }
}

template <int NS, bool PERIODIC, int SUB>
__device__ __forceinline__ void dm_wave_sub(float2& x, int a, int b, const mbqc_step* steps,
                                            float2 cf, int gbase, int k0, int c0, int n) {
  const int k = k0 + SUB;
  if (k < n) {
    const float2 gb = group_coeff<(1 << (2 * NS))>(cf, gbase, c0 + SUB);
    if constexpr (PERIODIC) {
      x = dm_wave_step<NS, NS - 1 - SUB>(x, a, b, gb, 3u, periodic_nmask(NS, SUB));
    } else {
      const mbqc_step st = steps[k];
      x = dm_wave_step<NS, NS - 1 - SUB>(x, a, b, gb, st.flags, st.nmask);
    }
  }
}

template <int NS, bool PERIODIC, int... SUBS>
__device__ __forceinline__ void dm_wave_block(float2& x, int a, int b, const mbqc_step* steps,
                                              float2 cf, int gbase, int k0, int c0, int n,
                                              std::integer_sequence<int, SUBS...>) {
  (dm_wave_sub<NS, PERIODIC, SUBS>(x, a, b, steps, cf, gbase, k0, c0, n), ...);
}

// Output channel of slot M on the entry held by this lane (as dm_channel).
template <int NS, int M>
__device__ __forceinline__ void dm_wave_channel(float2& x, int a, int b, const DmNoise& nz,
                                                uint32_t live) {
  constexpr int MB = 1 << M;
  if (!(live & MB)) return;
  const float2 p = lane_xor2<(MB | (MB << NS))>(x);   // (a^M, b^M)
  const bool am = (a >> M) & 1, bm = (b >> M) & 1;
  if (am != bm) {
    x = cscale(x, nz.coh);
  } else {
    const float2 A = am ? p : x, D = am ? x : p;
    const float w0 = am ? nz.d10 : nz.d00, w1 = am ? nz.d11 : nz.d01;
    x = make_float2(fmaf(w0, A.x, w1 * D.x), fmaf(w0, A.y, w1 * D.y));
  }
}

template <int NS, int... MS>
__device__ __forceinline__ void dm_wave_channels(float2& x, int a, int b, const DmNoise& nz,
                                                 uint32_t live, std::integer_sequence<int, MS...>) {
  (dm_wave_channel<NS, MS>(x, a, b, nz, live), ...);
}

// Density matrix with one entry per lane (windows of at most 3 slots: 64
// entries, so one simulation per wave at 3 slots, 4 at 2 slots, 16 at 1).
// The step's quartet partners are cross-lane moves at compile-time distances
// (the measured slot is fixed per sub-step of a block), so a step needs no LDS
// round trip and no barrier; a 4096-simulation batch of 3-slot windows runs
// 4096 waves, four per SIMD.  Each lane of a simulation's group computes one
// step's coefficient gamma e^{-i theta} per chunk of steps, broadcast to the
// group step by step.
template <int NS, bool PERIODIC>
__global__ __launch_bounds__(256) void k_dm_forward_wave(Job J, DmNoise nz, uint32_t flags) {
  constexpr int H = 1 << NS, E = H * H;
  static_assert(E <= 64, "one entry per lane");
  constexpr int SPW = 64 / E;                 // simulations per wave
  constexpr int CH = (E / NS) * NS;           // steps per coefficient chunk
  const int lane = threadIdx.x & 63;
  const int64_t wave = int64_t(xcd_block(blockIdx.x, gridDim.x)) * (blockDim.x >> 6) +
                       (threadIdx.x >> 6);
  const int e = lane & (E - 1), gbase = lane & ~(E - 1);
  const int a = e >> NS, b = e & (H - 1);
  const int64_t g = wave * SPW + lane / E;
  const bool active = g < J.total;
  const int64_t gg = active ? g : J.total - 1;
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_dmres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float2* psi = J.psi + si * J.in_stride;
  float2 x;
  {
    const float2 va = flip(psi[in_index[a]], sign[a] ? 0x80000000u : 0u);
    const float2 vb = flip(psi[in_index[b]], sign[b] ? 0x80000000u : 0u);
    x = cmul(va, cconj(vb));
  }
  const int n = J.n_steps;
  float2 cf = make_float2(0.f, 0.f);
  for (int k0 = 0; k0 < n; k0 += NS) {
    if (resc[k0 / NS]) x = cscale(x, 0x1p-32f);
    const int c0 = k0 % CH;
    if (c0 == 0) {
      // this lane's share of the chunk: the coefficient of step k0 + e
      const int k = k0 + e;
      float th = 0.f;
      if (e < CH && k < n) {
        const mbqc_step st = steps[k];
        th = st.angle_index >= 0 ? ang[int64_t(st.angle_index) * J.angle_stride_a] : st.fixed_angle;
      }
      cf = cscale(step_coeff<false>(th, nullptr), nz.gamma);
    }
    dm_wave_block<NS, PERIODIC>(x, a, b, steps, cf, gbase, k0, c0, n,
                                std::make_integer_sequence<int, NS>{});
  }
  if (nz.out_noise)
    dm_wave_channels<NS>(x, a, b, nz, uint32_t(J.live_mask), std::make_integer_sequence<int, NS>{});
  float scale = ldexpf(1.0f, -J.e_dm);
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float tr = (a == b && out_pos[a] >= 0) ? x.x : 0.0f;
    // sum over the E lanes of the group (xor distances below E stay inside it)
    if constexpr (E > 1) tr += lane_xor<1>(tr);
    if constexpr (E > 2) tr += lane_xor<2>(tr);
    if constexpr (E > 4) tr += lane_xor<4>(tr);
    if constexpr (E > 8) tr += lane_xor<8>(tr);
    if constexpr (E > 16) tr += lane_xor<16>(tr);
    if constexpr (E > 32) tr += lane_xor<32>(tr);
    scale = 1.0f / tr;
  }
  if (!active) return;
  const int pj = out_pos[a], pk = out_pos[b];
  if (pj >= 0 && pk >= 0) {
    const int64_t d = int64_t(1) << J.n_out;
    J.out[g * d * d + int64_t(pj) * d + pk] = cscale(x, scale);
  }
}

// ---------------------------------------------------------------- adjoint
// Forward pass of the adjoint: raw phase, unnormalized output, one checkpoint
// of the live state at the start of every block of NS steps (after the
// block's rescale), stored column-fastest for coalesced access.
template <int NS, bool PERIODIC>
// ckpt is __restrict__ so the checkpoint stores inside the step loop do not
// count as possible writes to the plan tables (which keeps table reads scalar).
// A checkpoint is kept every `ck` blocks (1 or 2; 2 halves the checkpoint
// traffic, which bounds large batches, for one extra block of recompute in
// the backward sweep).
__global__ __launch_bounds__(256) void k_sv_adj_forward(Job J, float2* __restrict__ ckpt, int ck) {
  constexpr int H = 1 << NS;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = g < J.total;
  const int64_t gg = active ? g : 0;
  const int64_t bi = gg / J.S, si = gg - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  float2 t[H];
  sv_load_input<NS>(J, J.psi + si * J.in_stride, t);
  RefPhase ph;
  const int64_t total = J.total;
  sv_run_steps<NS, false, PERIODIC>(t, J, ang, &ph, [&](int blk, const float2 (&tt)[H]) {
    if (active && (blk & (ck - 1)) == 0) {
      const int64_t c = blk >> (ck - 1);
#pragma unroll
      for (int j = 0; j < H; ++j) ckpt[(c * H + j) * total + g] = tt[j];
    }
  });
  // the wave's output rows leave through its LDS column tile (dynamic LDS of
  // one tile per wave)
  const int32_t* out_pos = J.tab + J.off_out;
  const float scale = sv_final_scale(J.e_sv);
  const int64_t g0 = g - (threadIdx.x & 63);
  const int rows = int(total - g0 < 64 ? total - g0 : 64);
  wave_store_rows<NS>(t, out_pos, make_float2(scale, 0.0f), J.out + g0 * (int64_t(1) << J.n_out),
                      J.n_out, rows);
}

// Adjoint of one SV step (slot Q).  Forward: u0 = x + b y, new0 = u0,
// new1 = alpha u0 + beta x.  Given the cotangent lam of the new state
// (dL = Re sum conj(lam) d(new)):  h = lam0 + alpha lam1 (cotangent of u0),
//   lam_x = h + beta lam1,  lam_y = conj(b) h,  dL/dtheta += Im(b sum y conj(h)).
template <int NS, int Q>
__device__ __forceinline__ float sv_adj_step(float2 (&lam)[1 << NS], const float2 (&y)[1 << (NS - 1)],
                                             float2 b, const float* coef) {
  constexpr int H = 1 << NS;
  constexpr int QB = 1 << Q;
  const float2 bc = cconj(b);
  float ar = 0.f, ai = 0.f;
  int i = 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & QB) continue;
    const float al = coef[2 * i], be = coef[2 * i + 1];
    const float2 l0 = lam[j], l1 = lam[j | QB];
    const float2 h = make_float2(fmaf(al, l1.x, l0.x), fmaf(al, l1.y, l0.y));
    const float2 yy = y[i++];
    ar = fmaf(yy.x, h.x, fmaf(yy.y, h.y, ar));
    ai = fmaf(yy.y, h.x, fmaf(-yy.x, h.y, ai));
    lam[j] = make_float2(fmaf(be, l1.x, h.x), fmaf(be, l1.y, h.y));
    lam[j | QB] = cmul(bc, h);
  }
  // Im(b * acc)
  return fmaf(b.x, ai, b.y * ar);
}

// sv_adj_step of a periodic step (alpha = -sign, beta = 2 sign with sign the
// parity of j & NMASK, as in sv_step_periodic): the same arithmetic.
template <int NS, int Q, uint32_t NMASK>
__device__ __forceinline__ float sv_adj_step_periodic(float2 (&lam)[1 << NS],
                                                      const float2 (&y)[1 << (NS - 1)], float2 b) {
  constexpr int H = 1 << NS;
  constexpr int QB = 1 << Q;
  const float2 bc = cconj(b);
  float ar = 0.f, ai = 0.f;
  int i = 0;
#pragma unroll
  for (int j = 0; j < H; ++j) {
    if (j & QB) continue;
    const float sg = (__builtin_popcount(uint32_t(j) & NMASK) & 1) ? -1.0f : 1.0f;
    const float2 l0 = lam[j], l1 = lam[j | QB];
    const float2 h = make_float2(fmaf(-sg, l1.x, l0.x), fmaf(-sg, l1.y, l0.y));
    const float2 yy = y[i++];
    ar = fmaf(yy.x, h.x, fmaf(yy.y, h.y, ar));
    ai = fmaf(yy.y, h.x, fmaf(-yy.x, h.y, ai));
    lam[j] = make_float2(fmaf(2.0f * sg, l1.x, h.x), fmaf(2.0f * sg, l1.y, h.y));
    lam[j | QB] = cmul(bc, h);
  }
  return fmaf(b.x, ai, b.y * ar);
}

template <int NS, int Q>
__device__ __forceinline__ void take_y(const float2 (&t)[1 << NS], float2 (&y)[1 << (NS - 1)]) {
  constexpr int QB = 1 << Q;
  int i = 0;
#pragma unroll
  for (int j = 0; j < (1 << NS); ++j) {
    if (j & QB) continue;
    y[i++] = t[j | QB];
  }
}

struct AdjCtx {
  float* gpart;     // [B][P][n_angles]
  int32_t ck;       // blocks per checkpoint (1 or 2)
  int64_t C;        // columns per parameter set
  int32_t P;        // partial sums per parameter set
  int32_t red;      // lanes reduced together (1, pow2 <= 64)
  int32_t n_angles;
};

template <int NS, bool PERIODIC, int SUB>
__device__ __forceinline__ void adj_recompute(float2 (&t)[1 << NS], float2 (&ys)[NS][1 << (NS - 1)],
                                              float2 (&bs)[NS], const mbqc_step* steps,
                                              const float* coef, int k0, int n,
                                              const float (&th)[NS]) {
  const int k = k0 + SUB;
  if (PERIODIC || k < n) {   // a periodic plan has only full blocks
    bs[SUB] = step_coeff<false>(th[SUB], nullptr);
    take_y<NS, NS - 1 - SUB>(t, ys[SUB]);
    if constexpr (SUB + 1 < NS) {
      if constexpr (PERIODIC)
        sv_step_periodic<NS, NS - 1 - SUB, periodic_nmask(NS, SUB)>(t, bs[SUB]);
      else
        sv_step<NS, NS - 1 - SUB>(t, bs[SUB], coef_row<NS>(coef, steps, k));
    }
  }
}

// Sum of v over aligned groups of `red` lanes (red a power of two <= 64),
// every lane of a group ending with the group's sum.  Groups up to 16 lanes
// use DPP moves (xor 1 and xor 2 quad permutations, then the 8- and 16-lane
// mirrors, each pairing one half of the group with the other); wider groups
// finish with cross-lane shuffles.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}

__device__ __forceinline__ float lane_group_sum(float v, int red) {
  // one uniform branch for the usual groups of 16 or more lanes
  if (red >= 16) {
    v += dpp_move<0xB1>(v);                 // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);                 // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);                // row_half_mirror
    v += dpp_move<0x140>(v);                // row_mirror
    if (red >= 32) {
      v += __shfl_xor(v, 16, 64);
      if (red >= 64) v += __shfl_xor(v, 32, 64);
    }
  } else {
    if (red >= 2) v += dpp_move<0xB1>(v);
    if (red >= 4) v += dpp_move<0x4E>(v);
    if (red >= 8) v += dpp_move<0x141>(v);
  }
  return v;
}

template <int NS, bool PERIODIC, int SUB>
__device__ __forceinline__ void adj_back(float2 (&lam)[1 << NS], const float2 (&ys)[NS][1 << (NS - 1)],
                                         const float2 (&bs)[NS], const mbqc_step* steps,
                                         const float* coef, int k0, int n, const AdjCtx& A,
                                         float* prow, bool writer) {
  const int k = k0 + SUB;
  if (PERIODIC || k < n) {
    float d;
    if constexpr (PERIODIC)
      d = sv_adj_step_periodic<NS, NS - 1 - SUB, periodic_nmask(NS, SUB)>(lam, ys[SUB], bs[SUB]);
    else
      d = sv_adj_step<NS, NS - 1 - SUB>(lam, ys[SUB], bs[SUB], coef_row<NS>(coef, steps, k));
    const int ai = steps[k].angle_index;
    if (ai >= 0) {
      d = lane_group_sum(d, A.red);
      if (writer) prow[ai] = d;
    }
  }
}

template <int NS, bool PERIODIC, int... SUBS>
__device__ __forceinline__ void adj_block(float2 (&t)[1 << NS], float2 (&lam)[1 << NS],
                                          const mbqc_step* steps, const float* coef, int k0, int n,
                                          const float (&th)[NS], const AdjCtx& A, float* prow,
                                          bool writer, std::integer_sequence<int, SUBS...>) {
  float2 ys[NS][1 << (NS - 1)];
  float2 bs[NS];
  (adj_recompute<NS, PERIODIC, SUBS>(t, ys, bs, steps, coef, k0, n, th), ...);
  (adj_back<NS, PERIODIC, NS - 1 - SUBS>(lam, ys, bs, steps, coef, k0, n, A, prow, writer), ...);
}

template <int NS, bool PERIODIC>
// The pointers are __restrict__ (and the partial sums come in as their own
// argument): the per-step partial-sum stores would otherwise count as possible
// writes to the plan tables, turning every table read into a vector load.
__global__ __launch_bounds__(64) void k_sv_adj_backward(Job J, const float2* __restrict__ ckpt,
                                                        const float2* __restrict__ seed, AdjCtx A,
                                                        float* __restrict__ gpart) {
  constexpr int H = 1 << NS;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = g < J.total;
  const int64_t gg = active ? g : 0;
  const int64_t bi = gg / A.C, ci = gg - bi * A.C;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float* coef = reinterpret_cast<const float*>(J.tab + J.off_coef);
  const float scale = sv_final_scale(J.e_sv);
  float2 lam[H];
  const float2* sd = seed + gg * (int64_t(1) << J.n_out);
  // every lane loads (inactive lanes simulation 0's seed) and a select zeroes
  // what is not used: loads under a lane guard each wait for their own round trip
#pragma unroll
  for (int j = 0; j < H; ++j) {
    const int p = out_pos[j];
    const float2 v = cscale(sd[p >= 0 ? p : 0], scale);
    lam[j] = (active && p >= 0) ? v : make_float2(0.f, 0.f);
  }
  const int n = J.n_steps;
  const int nblk = (n + NS - 1) / NS;
  // this lane's partial-sum row and whether it writes it (first lane of each
  // group of A.red columns), fixed for the whole sweep
  const int64_t prow_i = A.P == 1 ? 0 : (A.red == 64 ? ci / 64 : ci);
  float* prow = gpart + (bi * A.P + prow_i) * A.n_angles;
  const bool writer = active && (A.red == 1 || (ci & (A.red - 1)) == 0);
  // inactive lanes recompute simulation 0 (their sums stay in their own,
  // inactive lane groups and are never written)
  const int ck = A.ck;
  const int ngrp = (nblk + ck - 1) / ck;
  for (int grp = ngrp - 1; grp >= 0; --grp) {
    float2 s0[H];
    const float2* cb = ckpt + int64_t(grp) * H * J.total + gg;
#pragma unroll
    for (int j = 0; j < H; ++j) s0[j] = cb[int64_t(j) * J.total];
    const int b0 = grp * ck;
    float th0[NS];
    load_block_angles<NS>(steps, b0 * NS, n, ang, J.angle_stride_a, th0);
    if (ck == 2 && b0 + 1 < nblk) {
      // the state at block b0 + 1: block b0 (a full block) from the
      // checkpoint, then b0 + 1's rescale
      float2 t[H];
#pragma unroll
      for (int j = 0; j < H; ++j) t[j] = s0[j];
      RefPhase ph;
      NoModify nomod;
      sv_block<NS, false, true, false, false, PERIODIC>(t, steps, coef, b0 * NS, n, th0, &ph,
                                                        nomod, std::make_integer_sequence<int, NS>{});
      if (resc[b0 + 1]) {
#pragma unroll
        for (int j = 0; j < H; ++j) t[j] = cscale(t[j], 0x1p-32f);
      }
      float th1[NS];
      load_block_angles<NS>(steps, (b0 + 1) * NS, n, ang, J.angle_stride_a, th1);
      adj_block<NS, PERIODIC>(t, lam, steps, coef, (b0 + 1) * NS, n, th1, A, prow, writer,
                              std::make_integer_sequence<int, NS>{});
      if (resc[b0 + 1]) {
#pragma unroll
        for (int j = 0; j < H; ++j) lam[j] = cscale(lam[j], 0x1p-32f);
      }
    }
    adj_block<NS, PERIODIC>(s0, lam, steps, coef, b0 * NS, n, th0, A, prow, writer,
                            std::make_integer_sequence<int, NS>{});
    if (resc[b0]) {
#pragma unroll
      for (int j = 0; j < H; ++j) lam[j] = cscale(lam[j], 0x1p-32f);
    }
  }
}

// ---------------------------------------------------- adjoint, LDS-resident
// Windows of 6..kMaxAdjLdsSlots live slots: one workgroup per simulation in
// the forward, the state in LDS (as k_sv_forward_lds, raw phase), a
// checkpoint of the state at the start of every ck-th block (after the
// block's rescale) stored contiguously per simulation:
// ckpt[(c * total + g) * 2^NS + j].
constexpr int kMaxAdjLdsSlots = 10;

template <int NS, int T>
__device__ __forceinline__ void adj_forward_run(const Job& J, float2* __restrict__ ckpt, int ck,
                                                float2* st, float* th_s, int64_t g, int tid) {
  constexpr int H = 1 << NS;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float2* psi = J.psi + si * J.in_stride;
  for (int j = tid; j < H; j += T) st[j] = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
  for (int k = 0; k < J.n_steps; ++k) {
    if (k % NS == 0) {
      const int blk = k / NS;
      if (resc[blk]) {
        __syncthreads();
        for (int j = tid; j < H; j += T) st[j] = cscale(st[j], 0x1p-32f);
      }
      if ((blk & (ck - 1)) == 0) {
        __syncthreads();
        float2* cp = ckpt + (int64_t(blk >> (ck - 1)) * J.total + g) * H;
        for (int j = tid; j < H; j += T) cp[j] = st[j];
      }
    }
    if (k % kAngleChunk == 0) {
      __syncthreads();
      stage_angles<T>(th_s, steps, ang, J.angle_stride_a, k, J.n_steps, tid);
      __syncthreads();
    }
    const mbqc_step step = steps[k];
    const float2 b = step_coeff<false>(th_s[k % kAngleChunk], nullptr);
    __syncthreads();
    sv_lds_step<T>(st, H, step, b, tid);
  }
  __syncthreads();
  const float scale = sv_final_scale(J.e_sv);
  float2* o = J.out + g * (int64_t(1) << J.n_out);
  for (int j = tid; j < H; j += T) {
    const int p = out_pos[j];
    if (p >= 0) o[p] = cscale(st[j], scale);
  }
}

template <int NS, int T>
__global__ __launch_bounds__(T) void k_sv_adj_forward_lds(Job J, float2* __restrict__ ckpt, int ck) {
  __shared__ float2 st[1 << NS];
  __shared__ float th_s[kAngleChunk];
  adj_forward_run<NS, T>(J, ckpt, ck, st, th_s, xcd_block(blockIdx.x, gridDim.x), threadIdx.x);
}

// Windows above the LDS: the same with the state in a global workspace
// (ws[(g - g0) * 2^NS + j] for simulations g = g0 + blockIdx).
template <int NS>
__global__ __launch_bounds__(256) void k_sv_adj_forward_global(Job J, float2* __restrict__ ckpt,
                                                               int ck, float2* __restrict__ ws,
                                                               int64_t g0) {
  __shared__ float th_s[kAngleChunk];
  const int64_t l = xcd_block(blockIdx.x, gridDim.x);
  adj_forward_run<NS, 256>(J, ckpt, ck, ws + (l << NS), th_s, g0 + l, threadIdx.x);
}

// The angle of step k of this simulation (its angle-vector entry or its fixed angle).
__device__ __forceinline__ float step_angle(const mbqc_step& st, const float* ang, int64_t sa) {
  return st.angle_index >= 0 ? ang[int64_t(st.angle_index) * sa] : st.fixed_angle;
}

// Block [k0, k0 + NS) of the LDS adjoint: recompute the block forward from
// `cur` (its state at the block start), keeping every step's measured half
// (the pre-step amplitudes with slot Q = 1) in ys, then run the adjoint steps
// in reverse on lam, adding each angle's contribution to acc (one writer).
// Adjoint of new0 = u0 = x + b y, new1 = al u0 + be x (al, be from the step
// flags as in sv_lds_step): h = lam0 + al lam1, lam_x = h + be lam1,
// lam_y = conj(b) h, dL/dtheta += Im(b sum y conj(h)).
template <int NS, int T>
__device__ void adj_block_lds(float2* cur, float2* lam, float2* ys, float* red, float* acc,
                              const mbqc_step* steps, const float* ang, int64_t sa, int k0, int n,
                              int tid) {
  constexpr int H = 1 << NS, HP = H / 2;
  const int kend = k0 + NS < n ? k0 + NS : n;
  for (int k = k0; k < kend; ++k) {
    const mbqc_step st = steps[k];
    const int q = int((st.flags >> 8) & 0xffu), QB = 1 << q;
    __syncthreads();
    for (int i = tid; i < HP; i += T) ys[(k - k0) * HP + i] = cur[insert_zero(i, q) | QB];
    if (k + 1 < kend) {
      const float2 b = step_coeff<false>(step_angle(st, ang, sa), nullptr);
      __syncthreads();
      sv_lds_step<T>(cur, H, st, b, tid);
    }
  }
  for (int k = kend - 1; k >= k0; --k) {
    const mbqc_step st = steps[k];
    const int q = int((st.flags >> 8) & 0xffu), QB = 1 << q;
    const float s = (st.flags & 2u) ? -1.0f : 1.0f;
    const float app = (st.flags & 1u) ? 1.0f : 0.0f;
    const float2 b = step_coeff<false>(step_angle(st, ang, sa), nullptr);
    const float2 bc = cconj(b);
    float ar = 0.f, ai = 0.f;
    __syncthreads();
    for (int i = tid; i < HP; i += T) {
      const int j = insert_zero(i, q);
      const float c = (__builtin_popcount(unsigned(j) & st.nmask) & 1) ? -app : app;
      const float al = c * s, be = c * (1.0f - s);
      const float2 l0 = lam[j], l1 = lam[j | QB];
      const float2 h = make_float2(fmaf(al, l1.x, l0.x), fmaf(al, l1.y, l0.y));
      const float2 yy = ys[(k - k0) * HP + i];
      ar = fmaf(yy.x, h.x, fmaf(yy.y, h.y, ar));
      ai = fmaf(yy.y, h.x, fmaf(-yy.x, h.y, ai));
      lam[j] = make_float2(fmaf(be, l1.x, h.x), fmaf(be, l1.y, h.y));
      lam[j | QB] = cmul(bc, h);
    }
    float d = fmaf(b.x, ai, b.y * ar);
    d = block_sum<T>(d, red);
    if (tid == 0 && st.angle_index >= 0) acc[st.angle_index] += d;
  }
}

// Backward sweep of the LDS adjoint: one workgroup per (parameter set,
// partial-sum row), walking its A.red columns (all of them when A.red == C)
// one after the other; the per-angle sums stay in LDS and the row is stored
// once.  Dynamic LDS: lam, cur (2^NS each), ys (NS x 2^(NS-1)), acc (n_angles).
template <int NS, int T>
__device__ __forceinline__ void adj_backward_run(const Job& J, const float2* __restrict__ ckpt,
                                                 const float2* __restrict__ seed, const AdjCtx& A,
                                                 float* __restrict__ gpart, float2* lam, float2* cur,
                                                 float2* ys, float* acc, float* red, int64_t row,
                                                 int tid) {
  constexpr int H = 1 << NS;
  const int64_t bi = row / A.P, r = row - bi * A.P;   // row = bi * P + r
  const int ncol = A.red;
  const float* ang = J.angles + bi * J.angle_stride;
  const int64_t sa = J.angle_stride_a;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* resc = J.tab + J.off_svres;
  const int32_t* out_pos = J.tab + J.off_out;
  const float scale = sv_final_scale(J.e_sv);
  const int n = J.n_steps;
  const int nblk = (n + NS - 1) / NS;
  const int ck = A.ck;
  const int ngrp = (nblk + ck - 1) / ck;
  for (int a = tid; a < A.n_angles; a += T) acc[a] = 0.0f;
  for (int cc = 0; cc < ncol; ++cc) {
    const int64_t g = bi * A.C + r * ncol + cc;   // simulation (set bi, column)
    const float2* sd = seed + g * (int64_t(1) << J.n_out);
    __syncthreads();
    for (int j = tid; j < H; j += T) {
      const int p = out_pos[j];
      lam[j] = p >= 0 ? cscale(sd[p], scale) : make_float2(0.f, 0.f);
    }
    for (int grp = ngrp - 1; grp >= 0; --grp) {
      const float2* cp = ckpt + (int64_t(grp) * J.total + g) * H;
      const int b0 = grp * ck;
      if (ck == 2 && b0 + 1 < nblk) {
        // the state at block b0 + 1: block b0 from the checkpoint, then b0 + 1's rescale
        __syncthreads();
        for (int j = tid; j < H; j += T) cur[j] = cp[j];
        for (int k = b0 * NS; k < (b0 + 1) * NS; ++k) {
          const mbqc_step st = steps[k];
          const float2 b = step_coeff<false>(step_angle(st, ang, sa), nullptr);
          __syncthreads();
          sv_lds_step<T>(cur, H, st, b, tid);
        }
        if (resc[b0 + 1]) {
          __syncthreads();
          for (int j = tid; j < H; j += T) cur[j] = cscale(cur[j], 0x1p-32f);
        }
        adj_block_lds<NS, T>(cur, lam, ys, red, acc, steps, ang, sa, (b0 + 1) * NS, n, tid);
        if (resc[b0 + 1]) {
          __syncthreads();
          for (int j = tid; j < H; j += T) lam[j] = cscale(lam[j], 0x1p-32f);
        }
      }
      __syncthreads();
      for (int j = tid; j < H; j += T) cur[j] = cp[j];
      adj_block_lds<NS, T>(cur, lam, ys, red, acc, steps, ang, sa, b0 * NS, n, tid);
      if (resc[b0]) {
        __syncthreads();
        for (int j = tid; j < H; j += T) lam[j] = cscale(lam[j], 0x1p-32f);
      }
    }
  }
  __syncthreads();
  float* prow = gpart + row * A.n_angles;
  for (int a = tid; a < A.n_angles; a += T) prow[a] = acc[a];
}

template <int NS, int T>
__global__ __launch_bounds__(T) void k_sv_adj_backward_lds(Job J, const float2* __restrict__ ckpt,
                                                           const float2* __restrict__ seed, AdjCtx A,
                                                           float* __restrict__ gpart) {
  constexpr int H = 1 << NS, HP = H / 2;
  extern __shared__ float2 dyn[];
  __shared__ float red[4];
  float2* lam = dyn;
  float2* cur = lam + H;
  float2* ys = cur + H;
  float* acc = reinterpret_cast<float*>(ys + NS * HP);
  adj_backward_run<NS, T>(J, ckpt, seed, A, gpart, lam, cur, ys, acc, red,
                          xcd_block(blockIdx.x, gridDim.x), threadIdx.x);
}

// Windows above the LDS: lam, cur and the block's measured halves in a global
// workspace slice per workgroup (ws[(row - r0) * (2 + NS/2) * 2^NS ...]), the
// per-angle sums in dynamic LDS.
template <int NS>
__global__ __launch_bounds__(256) void k_sv_adj_backward_global(Job J, const float2* __restrict__ ckpt,
                                                                const float2* __restrict__ seed,
                                                                AdjCtx A, float* __restrict__ gpart,
                                                                float2* __restrict__ ws, int64_t r0) {
  constexpr int H = 1 << NS, HP = H / 2;
  extern __shared__ float acc_dyn[];
  __shared__ float red[4];
  const int64_t l = xcd_block(blockIdx.x, gridDim.x);
  float2* lam = ws + l * int64_t(2 * H + NS * HP);
  float2* cur = lam + H;
  float2* ys = cur + H;
  adj_backward_run<NS, 256>(J, ckpt, seed, A, gpart, lam, cur, ys, acc_dyn, red, r0 + l,
                            threadIdx.x);
}

// ---------------------------------------------------------------- segmented adjoint
// Transfer-matrix adjoint cut in time.  The blocks of the pattern are split
// into segments p = 0..n-1 at block boundaries; segment p's transfer matrix
// T_p (from the window state at its first block, before that block's rescale,
// to the window state at the next segment's first block) is computed for all
// segments at once, one window basis column per lane (segment 0 starts from
// the input basis states instead, and the last segment ends with the output
// gather and the final scale).  A small per-set kernel chains them:
//   S_1 = T_0,  S_{p+1} = T_p S_p,  T = T_{n-1} S_{n-1},
// and, given the seed G = dL/dconj(T), propagates the cotangents back:
//   L_n = G,  G_p = L_{p+1} S_p^H,  L_p = T_p^H L_{p+1},  G_0 = L_1,
// so that dL = Re sum_p tr(G_p^H dT_p).  The backward sweep of segment p then
// runs from its checkpoints with the columns of G_p as seeds.  With n segments
// a parameter set occupies n times as many lanes, each for 1/n of the steps.
constexpr int kMaxSegments = 16;

struct SegInfo {
  int32_t n;                        // segments
  int32_t cols0, cols;              // columns of segment 0 (2^n_in) and of the others (2^NS)
  int32_t rows_last;                // rows of the last segment's matrix (2^n_out)
  int64_t B;                        // parameter sets
  int32_t blk[kMaxSegments + 1];    // first block of each segment; blk[n] = number of blocks
  int64_t wave[kMaxSegments + 1];   // first 64-lane wave of each segment; wave[n] = total
  int64_t ckpt[kMaxSegments];       // float2 offsets into the workspace: checkpoints,
  int64_t tmat[kMaxSegments];       //   T_p (columns of rows_p entries),
  int64_t smat[kMaxSegments];       //   S_p for p >= 2 (columns of 2^NS entries),
  int64_t gmat[kMaxSegments];       //   G_p (columns of rows_p entries)
};

__device__ __forceinline__ int seg_of_wave(const SegInfo& G, int64_t w) {
  int p = 0;
  for (int q = 1; q < G.n; ++q) p = (w >= G.wave[q]) ? q : p;
  return p;
}

__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 acc) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y)));
}
// acc + conj(a) b
__device__ __forceinline__ float2 cfma_conj(float2 a, float2 b, float2 acc) {
  return make_float2(fmaf(a.x, b.x, fmaf(a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(-a.y, b.x, acc.y)));
}

// One lane per (segment, set, column); one wave per workgroup, every wave
// inside one segment.  `ckpt` and `ws` are the same workspace: the checkpoint
// stores go through their own restrict pointer, so the compiler can prove
// they never touch the plan tables and keeps the table reads on scalar loads.
template <int NS, bool PERIODIC>
__global__ __launch_bounds__(64) void k_sv_seg_forward(Job J, SegInfo G, float2* __restrict__ ckpt,
                                                       float2* __restrict__ ws) {
  constexpr int H = 1 << NS;
  const int p = seg_of_wave(G, blockIdx.x);
  const int C = p == 0 ? G.cols0 : G.cols;
  const int64_t Np = (G.wave[p + 1] - G.wave[p]) * 64;
  const int64_t l = (int64_t(blockIdx.x) - G.wave[p]) * 64 + threadIdx.x;
  const bool active = l < G.B * C;
  const int64_t ll = active ? l : 0;
  const int64_t bi = ll / C;
  const int c = int(ll - bi * C);
  const float* ang = J.angles + bi * J.angle_stride;
  float2 t[H];
  if (p == 0) {
    const int32_t* in_index = J.tab + J.off_in;
    const int32_t* sign = J.tab + J.off_sign;
#pragma unroll
    for (int j = 0; j < H; ++j)
      t[j] = make_float2(in_index[j] == c ? (sign[j] ? -1.0f : 1.0f) : 0.0f, 0.0f);
  } else {
#pragma unroll
    for (int j = 0; j < H; ++j) t[j] = make_float2(j == c ? 1.0f : 0.0f, 0.0f);
  }
  RefPhase ph;
  float2* ck = ckpt + G.ckpt[p];
  const int b0 = G.blk[p];
  sv_run_blocks<NS, false, false, false, PERIODIC>(t, J, ang, &ph, b0, G.blk[p + 1],
                                                   [&](int blk, const float2 (&tt)[H]) {
    // every lane stores (inactive lanes, l < Np, into the padding of the
    // segment's checkpoint rows): a store branch that can be skipped would
    // make the wait for the next block's angle loads wait for these stores
#pragma unroll
    for (int j = 0; j < H; ++j) ck[(int64_t(blk - b0) * H + j) * Np + l] = tt[j];
  });
  // the wave's rows (64 consecutive lanes of the segment) leave through its
  // LDS column tile
  const int64_t l0 = l - threadIdx.x;
  const int rows = int(G.B * C - l0 < 64 ? G.B * C - l0 : 64);
  if (p + 1 < G.n) {
    wave_store_rows<NS, true>(t, nullptr, make_float2(1.f, 0.f), ws + G.tmat[p] + l0 * H, NS, rows);
  } else {
    const int32_t* out_pos = J.tab + J.off_out;
    const float scale = sv_final_scale(J.e_sv);
    wave_store_rows<NS>(t, out_pos, make_float2(scale, 0.f), ws + G.tmat[p] + l0 * G.rows_last,
                        J.n_out, rows);
  }
}

// The chain kernels hold every matrix in LDS as 2^NS x 2^NS, column-major
// with a column stride of 2^NS + 1 (lanes that walk a row hit distinct
// banks), zero-padded past its rows and columns, so all products run over the
// full, unrolled dimension.  Each thread owns kPer entries e = tid + 256 i; the
// next step's matrices are fetched into registers while the current product
// runs, so the chain waits on global memory only once.
template <int NS>
struct SegTile {
  static constexpr int H = 1 << NS;
  static constexpr int LD = H + 1;
  static constexpr int kPer = (H * H + 255) / 256;
  float2 v[kPer];
  __device__ __forceinline__ void fetch(const float2* src, int rows, int cols) {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = int(threadIdx.x) + 256 * i;
      const int c = e >> NS, r = e & (H - 1);
      v[i] = (e < H * H && r < rows && c < cols) ? src[c * rows + r] : make_float2(0.f, 0.f);
    }
  }
  __device__ __forceinline__ void put(float2* lds) const {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = int(threadIdx.x) + 256 * i;
      if (e < H * H) lds[(e >> NS) * LD + (e & (H - 1))] = v[i];
    }
  }
};

// Per parameter set (one 256-thread workgroup): S_{p+1} = T_p S_p and
// T = T_{n-1} S_{n-1}.
template <int NS>
__global__ __launch_bounds__(256) void k_seg_chain_forward(SegInfo G, float2* __restrict__ ws,
                                                           float2* __restrict__ out) {
  constexpr int H = 1 << NS;
  constexpr int kPer = SegTile<NS>::kPer;
  constexpr int LD = SegTile<NS>::LD;
  __shared__ float2 Sm[H * LD], Tm[H * LD];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int C0 = G.cols0;
  auto rows_of = [&](int p) { return p + 1 == G.n ? G.rows_last : H; };
  SegTile<NS> s0, tn;
  s0.fetch(ws + G.tmat[0] + b * C0 * H, H, C0);
  tn.fetch(ws + G.tmat[1] + b * H * rows_of(1), rows_of(1), H);
  s0.put(Sm);
  for (int p = 1; p < G.n; ++p) {
    const bool last = p + 1 == G.n;
    const int R = rows_of(p);
    __syncthreads();
    tn.put(Tm);
    __syncthreads();
    if (!last) tn.fetch(ws + G.tmat[p + 1] + b * H * rows_of(p + 1), rows_of(p + 1), H);
    float2 acc[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = tid + 256 * i;
      const int c = (e >> NS) & (H - 1), r = e & (H - 1);
      float2 a = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < H; ++k) a = cfma(Tm[k * LD + r], Sm[c * LD + k], a);
      acc[i] = a;
    }
    __syncthreads();
    float2* dst = last ? out + b * C0 * R : ws + G.smat[p + 1] + b * C0 * H;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = tid + 256 * i;
      if (e < H * H) {
        const int c = e >> NS, r = e & (H - 1);
        if (!last) Sm[c * LD + r] = acc[i];
        if (r < R && c < C0) dst[c * R + r] = acc[i];
      }
    }
  }
}

// Per parameter set: the segment seeds G_p from the output seed G.
template <int NS>
__global__ __launch_bounds__(256) void k_seg_chain_backward(SegInfo G, float2* __restrict__ ws,
                                                            const float2* __restrict__ seed) {
  constexpr int H = 1 << NS;
  constexpr int kPer = SegTile<NS>::kPer;
  constexpr int LD = SegTile<NS>::LD;
  __shared__ float2 Lm[H * LD], Sm[H * LD], Tm[H * LD];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int C0 = G.cols0;
  auto rows_of = [&](int p) { return p + 1 == G.n ? G.rows_last : H; };
  // S_1 = T_0
  auto s_src = [&](int p) { return (p == 1 ? ws + G.tmat[0] : ws + G.smat[p]) + b * C0 * H; };
  SegTile<NS> l0, sn, tn;
  l0.fetch(seed + b * C0 * G.rows_last, G.rows_last, C0);
  sn.fetch(s_src(G.n - 1), H, C0);
  tn.fetch(ws + G.tmat[G.n - 1] + b * H * G.rows_last, G.rows_last, H);
  l0.put(Lm);
  for (int p = G.n - 1; p >= 1; --p) {
    const int R = rows_of(p);                         // rows of L and of T_p
    __syncthreads();
    sn.put(Sm);
    tn.put(Tm);
    __syncthreads();
    if (p > 1) {
      sn.fetch(s_src(p - 1), H, C0);
      tn.fetch(ws + G.tmat[p - 1] + b * H * H, H, H);
    }
    // G_p[r][j] = sum_c L[r][c] conj(S_p[j][c])
    float2* gp = ws + G.gmat[p] + b * H * R;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = tid + 256 * i;
      const int j = (e >> NS) & (H - 1), r = e & (H - 1);
      float2 a = make_float2(0.f, 0.f);
#pragma unroll
      for (int c = 0; c < H; ++c) a = cfma_conj(Sm[c * LD + j], Lm[c * LD + r], a);
      if (e < H * H && r < R) gp[j * R + r] = a;
    }
    // L_p[k][c] = sum_r conj(T_p[r][k]) L[r][c]
    float2 acc[kPer];
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = tid + 256 * i;
      const int c = (e >> NS) & (H - 1), k = e & (H - 1);
      float2 a = make_float2(0.f, 0.f);
#pragma unroll
      for (int r = 0; r < H; ++r) a = cfma_conj(Tm[k * LD + r], Lm[c * LD + r], a);
      acc[i] = a;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int e = tid + 256 * i;
      if (e < H * H) Lm[(e >> NS) * LD + (e & (H - 1))] = acc[i];
    }
  }
  __syncthreads();
  float2* g0 = ws + G.gmat[0] + b * C0 * H;
  for (int e = tid; e < C0 * H; e += 256) g0[e] = Lm[(e >> NS) * LD + (e & (H - 1))];
}

template <int NS, bool PERIODIC>
__global__ __launch_bounds__(64) void k_sv_seg_backward(Job J, SegInfo G, const float2* __restrict__ ws,
                                                        int32_t n_angles, float* __restrict__ grad) {
  constexpr int H = 1 << NS;
  const int p = seg_of_wave(G, blockIdx.x);
  const int C = p == 0 ? G.cols0 : G.cols;
  const int64_t Np = (G.wave[p + 1] - G.wave[p]) * 64;
  const int64_t l = (int64_t(blockIdx.x) - G.wave[p]) * 64 + threadIdx.x;
  const bool active = l < G.B * C;
  const int64_t ll = active ? l : 0;
  const int64_t bi = ll / C;
  const int c = int(ll - bi * C);
  const float* ang = J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* resc = J.tab + J.off_svres;
  const float* coef = reinterpret_cast<const float*>(J.tab + J.off_coef);
  // The seeds are loaded by every lane (inactive lanes read lane 0's) and
  // zeroed by a select: loads under a lane guard each wait for their own
  // round trip.
  float2 lam[H];
  if (p + 1 == G.n) {
    const int32_t* out_pos = J.tab + J.off_out;
    const float scale = sv_final_scale(J.e_sv);
    const float2* sd = ws + G.gmat[p] + ll * G.rows_last;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const int q = out_pos[j];
      const float2 v = cscale(sd[q >= 0 ? q : 0], scale);
      lam[j] = (active && q >= 0) ? v : make_float2(0.f, 0.f);
    }
  } else {
    const float2* sd = ws + G.gmat[p] + ll * H;
#pragma unroll
    for (int j = 0; j < H; ++j) {
      const float2 v = sd[j];
      lam[j] = active ? v : make_float2(0.f, 0.f);
    }
  }
  AdjCtx A{};
  A.red = C;
  float* prow = grad + bi * n_angles;
  const bool writer = active && c == 0;
  const float2* ck = ws + G.ckpt[p];
  const int n = J.n_steps;
  const int b0 = G.blk[p];
  // inactive lanes recompute lane 0 of the segment (see k_sv_adj_backward)
  for (int blk = G.blk[p + 1] - 1; blk >= b0; --blk) {
    float2 t[H];
    const float2* cb = ck + int64_t(blk - b0) * H * Np + ll;
#pragma unroll
    for (int j = 0; j < H; ++j) t[j] = cb[int64_t(j) * Np];
    float th[NS];
    load_block_angles<NS>(steps, blk * NS, n, ang, J.angle_stride_a, th);
    adj_block<NS, PERIODIC>(t, lam, steps, coef, blk * NS, n, th, A, prow, writer,
                            std::make_integer_sequence<int, NS>{});
    if (resc[blk]) {
#pragma unroll
      for (int j = 0; j < H; ++j) lam[j] = cscale(lam[j], 0x1p-32f);
    }
  }
}

// ---------------------------------------------------------------- MB-QML loss
// Infidelity loss and adjoint seed of the MB-QML cost from the transfer
// matrices, one fused pass (the Gram form of 1 - mean_s |<phi_s|T psi_s>|^2,
// mentpy/calculator/borrows.py:4-6 with the tutorial loss,
// docs/tutorials/intro-to-mbqml.rst:35-54):
//   z_b = M t_b,  loss_b = 1 - Re(t_b^H z_b) / S,  seed_b = -2 z_b / S,
// with t_b = vec(T_b) (K = 2^(|I|+|O|) entries) and the data Gram matrix M
// given transposed (MT[k][n] = M[n][k]).  A workgroup covers kGramSets
// parameter sets x 64 output indices n; its 256 threads are 64 n x 4 groups
// of kGramSets/4 sets; the sets' t rows are staged in LDS, MT is read
// coalesced along n.  Loss partial sums go through LDS and one atomic-free
// second pass (k_gram_loss) adds them per set.
constexpr int kGramSets = 16;

__global__ __launch_bounds__(256) void k_gram_seed(const float2* __restrict__ t,
                                                   const float2* __restrict__ MT, int64_t B, int K,
                                                   float inv_S, float2* __restrict__ seed,
                                                   float* __restrict__ lpart) {
  extern __shared__ float2 ts[];                       // [kGramSets][K]
  constexpr int G = kGramSets / 4;
  const int tid = threadIdx.x;
  const int nl = tid & 63, grp = tid >> 6;
  const int n = blockIdx.x * 64 + nl;
  const int64_t b0 = int64_t(blockIdx.y) * kGramSets;
  for (int e = tid; e < kGramSets * K; e += 256) {
    const int sb = e / K, k = e - sb * K;
    ts[e] = (b0 + sb < B) ? t[(b0 + sb) * K + k] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  // lanes past K (K < 64) stay for the wave-wide loss reduction below
  const bool live = n < K;
  const int nn = live ? n : 0;
  float2 acc[G];
#pragma unroll
  for (int g = 0; g < G; ++g) acc[g] = make_float2(0.f, 0.f);
  const float2* tg = ts + grp * G * K;
  // K is a power of two >= 4: chunks of up to 16 loads of MT are issued
  // together before their FMAs, so the loop does not wait once per k
  constexpr int U = 16;
  const int u = K < U ? K : U;
  for (int k0 = 0; k0 < K; k0 += u) {
    float2 m[U];
#pragma unroll
    for (int i = 0; i < U; ++i)
      if (i < u) m[i] = MT[int64_t(k0 + i) * K + nn];
#pragma unroll
    for (int i = 0; i < U; ++i) {
      if (i < u) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const float2 x = tg[g * K + k0 + i];
          acc[g].x = fmaf(x.x, m[i].x, fmaf(-x.y, m[i].y, acc[g].x));
          acc[g].y = fmaf(x.x, m[i].y, fmaf(x.y, m[i].x, acc[g].y));
        }
      }
    }
  }
  const float sc = -2.0f * inv_S;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t b = b0 + grp * G + g;
    if (b < B) {
      if (live) seed[b * K + n] = cscale(acc[g], sc);
      // Re(conj(t[b,n]) z[b,n]), one partial per (set, 64-wide n block)
      const float2 x = tg[g * K + nn];
      float part = live ? fmaf(x.x, acc[g].x, x.y * acc[g].y) : 0.0f;
      for (int m2 = 32; m2 >= 1; m2 >>= 1) part += __shfl_xor(part, m2, 64);
      if (nl == 0) lpart[b * gridDim.x + blockIdx.x] = part;
    }
  }
}

// The same on the matrix cores, for K a multiple of 16: one wave per 16 x 16
// tile of Z (16 parameter sets x 16 indices n), complex products as four real
// v_mfma_f32_16x16x4f32 per k-chunk of 4 (Zr += Tr Mr' - Ti Mi', Zi += Tr Mi' +
// Ti Mr').  Operand layout of the 16x16x4 f32 MFMA: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; the accumulator holds D[4(l/16)+r][l%16], r = 0..3.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_gram_seed_mfma(const float2* __restrict__ t,
                                                       const float2* __restrict__ MT, int64_t B,
                                                       int K, float inv_S,
                                                       float2* __restrict__ seed,
                                                       float* __restrict__ lpart) {
  const int l = threadIdx.x;
  const int li = l & 15, lk = l >> 4;
  const int n0 = blockIdx.x * 16;
  const int64_t b0 = int64_t(blockIdx.y) * 16;
  const int64_t ba = b0 + li;                        // A row of this lane
  const bool arow = ba < B;
  const float2* ta = t + (arow ? ba : 0) * K + lk;
  const float2* mb = MT + int64_t(lk) * K + n0 + li;
  // Two accumulator pairs (even / odd k-chunks) keep consecutive MFMAs
  // independent; the next U chunks are loaded while the current ones run
  // (one wave per SIMD at the usual sizes, so nothing else hides the loads).
  f32x4 zr0 = {0.f, 0.f, 0.f, 0.f}, zi0 = {0.f, 0.f, 0.f, 0.f};
  f32x4 zr1 = {0.f, 0.f, 0.f, 0.f}, zi1 = {0.f, 0.f, 0.f, 0.f};
  constexpr int U = 8;                               // k-chunks of 4 loaded together
  float2 an[U], bn[U];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + 4 * u;
      const bool in = k < K;
      an[u] = (in && arow) ? ta[k] : make_float2(0.f, 0.f);
      bn[u] = in ? mb[int64_t(k) * K] : make_float2(0.f, 0.f);
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < K; k0 += 4 * U) {
    float2 a[U], b[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = an[u];
      b[u] = bn[u];
    }
    if (k0 + 4 * U < K) fetch(k0 + 4 * U);
#pragma unroll
    for (int u = 0; u < U; u += 2) {
      zr0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].x, zr0, 0, 0, 0);
      zi0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].x, b[u].y, zi0, 0, 0, 0);
      zr1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u + 1].x, b[u + 1].x, zr1, 0, 0, 0);
      zi1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u + 1].x, b[u + 1].y, zi1, 0, 0, 0);
      zr0 = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[u].y, b[u].y, zr0, 0, 0, 0);
      zi0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u].y, b[u].x, zi0, 0, 0, 0);
      zr1 = __builtin_amdgcn_mfma_f32_16x16x4f32(-a[u + 1].y, b[u + 1].y, zr1, 0, 0, 0);
      zi1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u + 1].y, b[u + 1].x, zi1, 0, 0, 0);
    }
  }
  f32x4 zr, zi;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    zr[r] = zr0[r] + zr1[r];
    zi[r] = zi0[r] + zi1[r];
  }
  const float sc = -2.0f * inv_S;
  const int n = n0 + li;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t b = b0 + 4 * lk + r;              // uniform within each 16-lane group
    float part = 0.0f;
    if (b < B) {
      const float2 z = make_float2(zr[r], zi[r]);
      seed[b * K + n] = cscale(z, sc);
      const float2 x = t[b * K + n];
      part = fmaf(x.x, z.x, x.y * z.y);
    }
    // sum over the 16 lanes (indices n) of this row
    for (int m = 8; m >= 1; m >>= 1) part += __shfl_xor(part, m, 64);
    if (li == 0 && b < B) lpart[b * gridDim.x + blockIdx.x] = part;
  }
}

// Loss and seed from t and a precomputed z = M t (library GEMM path): one
// workgroup per parameter set, loss = 1 - Re(t^H z) / S, seed = -2 z / S.
__global__ __launch_bounds__(256) void k_gram_finish(const float2* __restrict__ t,
                                                     const float2* __restrict__ z, int K, float inv_S,
                                                     float* __restrict__ loss,
                                                     float2* __restrict__ seed) {
  __shared__ float red[4];
  const int64_t b = blockIdx.x;
  const float2* tb = t + b * K;
  const float2* zb = z + b * K;
  float2* sb = seed + b * K;
  const float sc = -2.0f * inv_S;
  float part = 0.0f;
  for (int k = threadIdx.x; k < K; k += 256) {
    const float2 x = tb[k], y = zb[k];
    part = fmaf(x.x, y.x, fmaf(x.y, y.y, part));
    sb[k] = cscale(y, sc);
  }
  part = block_sum<256>(part, red);
  if (threadIdx.x == 0) loss[b] = 1.0f - part * inv_S;
}

// MB-QML loss (and seed) from propagated inputs: for parameter set b (one
// workgroup), f_bs = sum_o conj(Y[s][o]) out[b][s][o],
// loss[b] = 1 - (1/S) sum_s |f_bs|^2 and, if seed is given,
// seed[b][s][o] = -2/S f_bs Y[s][o] (dL/d conj(out)).  G lanes (G = min(O, 64),
// a power of two) share a row s, so each load instruction of a wave reads
// contiguous rows; f is summed over the G lanes with shuffles inside the
// group (all lanes of a group walk the same rows).
__global__ __launch_bounds__(256) void k_overlap_loss(const float2* __restrict__ out,
                                                      const float2* __restrict__ Y, int S, int O,
                                                      int G, float inv_S, float* __restrict__ loss,
                                                      float2* __restrict__ seed) {
  __shared__ float red[4];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int o0 = tid & (G - 1), r0 = tid / G, rows = 256 / G;
  const float sc = -2.0f * inv_S;
  float acc = 0.0f;
  for (int s = r0; s < S; s += rows) {
    const float2* ob = out + (b * S + s) * int64_t(O);
    const float2* y = Y + int64_t(s) * O;
    float2 f = make_float2(0.f, 0.f);
    for (int o = o0; o < O; o += G) {
      const float2 v = ob[o], w = y[o];
      f.x = fmaf(w.x, v.x, fmaf(w.y, v.y, f.x));
      f.y = fmaf(w.x, v.y, fmaf(-w.y, v.x, f.y));
    }
    for (int m = G >> 1; m >= 1; m >>= 1) {
      f.x += __shfl_xor(f.x, m, 64);
      f.y += __shfl_xor(f.y, m, 64);
    }
    if (o0 == 0) acc = fmaf(f.x, f.x, fmaf(f.y, f.y, acc));
    if (seed) {
      float2* sb = seed + (b * S + s) * int64_t(O);
      const float2 fs = cscale(f, sc);
      for (int o = o0; o < O; o += G) sb[o] = cmul(fs, y[o]);
    }
  }
  acc = block_sum<256>(acc, red);
  if (tid == 0) loss[b] = 1.0f - acc * inv_S;
}

__global__ void k_gram_loss(const float* __restrict__ lpart, int64_t B, int nblk, float inv_S,
                            float* __restrict__ loss) {
  const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.0f;
  for (int i = 0; i < nblk; ++i) s += lpart[b * nblk + i];
  loss[b] = 1.0f - s * inv_S;
}

// ---------------------------------------------------------------- segment chains on MFMA
// The chain products of the segmented adjoint for 16-amplitude windows as
// 16x16x16 complex products on the matrix cores, one wave per parameter set:
// four real v_mfma_f32_16x16x4f32 per k-chunk of 4 (re*re - im*im, re*im +
// im*re).  Operands come from LDS tiles (column-major, stride 17, zero-padded
// past their rows/columns): lane l reads A[l%16][k0 + l/16] and
// B[k0 + l/16][l%16], and holds D[4(l/16) + i][l%16] of the result.  Every
// tile access walks 16 consecutive rows or a stride-17 row, so the reads are
// free of bank conflicts.
constexpr int kChainLd = 17;

// acc += A * B for one k-chunk; a = A[row][kk], b = B[kk][col] of this lane
__device__ __forceinline__ void cmfma_chunk(f32x4& dr, f32x4& di, float2 a, float2 b) {
  dr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.x, dr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.y, di, 0, 0, 0);
  dr = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.y, b.y, dr, 0, 0, 0);
  di = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.x, di, 0, 0, 0);
}

// 16x16 tile from global (column-major, `rows` x `cols` entries, ld = rows)
// into LDS (stride 17, zero-padded), by the 64 lanes of the wave.
__device__ __forceinline__ void chain_tile_load(float2* lds, const float2* src, int rows, int cols) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = l + 64 * i;
    const int c = e >> 4, r = e & 15;
    lds[c * kChainLd + r] = (r < rows && c < cols) ? src[c * rows + r] : make_float2(0.f, 0.f);
  }
}

// The same tile into four registers per lane (entries l + 64 i), so the
// next step's tiles load while the current product runs.
struct ChainTileRegs {
  float2 v[4];
  __device__ __forceinline__ void load(const float2* src, int rows, int cols) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = l + 64 * i;
      const int c = e >> 4, r = e & 15;
      v[i] = (r < rows && c < cols) ? src[c * rows + r] : make_float2(0.f, 0.f);
    }
  }
  __device__ __forceinline__ void put(float2* lds) const {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = l + 64 * i;
      lds[(e >> 4) * kChainLd + (e & 15)] = v[i];
    }
  }
};

// D tile of this lane into LDS (column-major) and, for rows < rows and
// columns < cols, into global memory (ld = rows).
__device__ __forceinline__ void chain_tile_store(float2* lds, float2* dst, const f32x4& dr,
                                                 const f32x4& di, int rows, int cols) {
  const int l = threadIdx.x & 63;
  const int c = l & 15, g = l >> 4;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = 4 * g + i;
    const float2 v = make_float2(dr[i], di[i]);
    if (lds) lds[c * kChainLd + r] = v;
    if (dst && r < rows && c < cols) dst[c * rows + r] = v;
  }
}

// Per parameter set (one wave): S_{p+1} = T_p S_p and T = T_{n-1} S_{n-1}
// (16-amplitude windows).  (Fetching every tile up front measured slower than
// one step's tiles at a time.)
__global__ __launch_bounds__(64) void k_seg_chain_forward_mfma(SegInfo G, float2* __restrict__ ws,
                                                               float2* __restrict__ out) {
  __shared__ float2 Sm[16 * kChainLd], Tm[16 * kChainLd];
  const int64_t b = blockIdx.x;
  const int l = threadIdx.x;
  const int C0 = G.cols0;
  auto rows_of = [&](int p) { return p + 1 == G.n ? G.rows_last : 16; };
  ChainTileRegs tn;
  tn.load(ws + G.tmat[1] + b * 16 * rows_of(1), rows_of(1), 16);
  chain_tile_load(Sm, ws + G.tmat[0] + b * C0 * 16, 16, C0);
  for (int p = 1; p < G.n; ++p) {
    const bool last = p + 1 == G.n;
    const int R = rows_of(p);
    tn.put(Tm);
    // the next step's tile loads while this product runs
    if (!last) tn.load(ws + G.tmat[p + 1] + b * 16 * rows_of(p + 1), rows_of(p + 1), 16);
    wave_lds_sync();
    f32x4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      const int kk = k0 + (l >> 4), rc = l & 15;
      cmfma_chunk(dr, di, Tm[kk * kChainLd + rc], Sm[rc * kChainLd + kk]);
    }
    wave_lds_sync();
    if (last)
      chain_tile_store(nullptr, out + b * C0 * R, dr, di, R, C0);
    else
      chain_tile_store(Sm, ws + G.smat[p + 1] + b * C0 * 16, dr, di, 16, C0);
  }
}

// Per parameter set (one wave): the segment seeds G_p = L_{p+1} S_p^H with
// L_p = T_p^H L_{p+1}, L_n = G (16-amplitude windows).
__global__ __launch_bounds__(64) void k_seg_chain_backward_mfma(SegInfo G, float2* __restrict__ ws,
                                                                const float2* __restrict__ seed) {
  __shared__ float2 Lm[16 * kChainLd], Sm[16 * kChainLd], Tm[16 * kChainLd];
  const int64_t b = blockIdx.x;
  const int l = threadIdx.x;
  const int C0 = G.cols0;
  int R = G.rows_last;                                   // rows of L
  auto s_src = [&](int p) { return (p == 1 ? ws + G.tmat[0] : ws + G.smat[p]) + b * C0 * 16; };
  ChainTileRegs sn, tn;
  sn.load(s_src(G.n - 1), 16, C0);
  tn.load(ws + G.tmat[G.n - 1] + b * 16 * R, R, 16);
  chain_tile_load(Lm, seed + b * C0 * R, R, C0);
  const int kk0 = l >> 4, rc = l & 15;
  for (int p = G.n - 1; p >= 1; --p) {
    sn.put(Sm);
    tn.put(Tm);
    // the next step's tiles load while these products run
    if (p > 1) {
      sn.load(s_src(p - 1), 16, C0);
      tn.load(ws + G.tmat[p - 1] + b * 16 * 16, 16, 16);
    }
    wave_lds_sync();
    // G_p[r][j] = sum_c L[r][c] conj(S_p[j][c]):  A = L, B[c][j] = conj(S_p[j][c])
    f32x4 gr = {0.f, 0.f, 0.f, 0.f}, gi = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      const int c = k0 + kk0;
      cmfma_chunk(gr, gi, Lm[c * kChainLd + rc], cconj(Sm[c * kChainLd + rc]));
    }
    // L_p[k][c] = sum_r conj(T_p[r][k]) L[r][c]:  A[k][r] = conj(T_p[r][k]), B = L
    f32x4 lr = {0.f, 0.f, 0.f, 0.f}, li = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < 16; k0 += 4) {
      const int r = k0 + kk0;
      cmfma_chunk(lr, li, cconj(Tm[rc * kChainLd + r]), Lm[rc * kChainLd + r]);
    }
    wave_lds_sync();
    chain_tile_store(nullptr, ws + G.gmat[p] + b * 16 * R, gr, gi, R, 16);
    chain_tile_store(Lm, nullptr, lr, li, 16, 16);
    wave_lds_sync();
    R = 16;
  }
  float2* g0 = ws + G.gmat[0] + b * C0 * 16;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int e = l + 64 * i;
    const int c = e >> 4, r = e & 15;
    if (c < C0) g0[c * 16 + r] = Lm[c * kChainLd + r];
  }
}

// ---------------------------------------------------------------- optimizer
// Adam update of mentpy's AdamOptimizer.step (mentpy/optimizers/adam.py:54-66)
// for a batch of parameter vectors, in place on the device:
//   m = b1 m + (1-b1) g,  v = b2 v + (1-b2) g^2,
//   x -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).
// With b1 = b2 = 0 and eps = 0 it is not SGD; SGD (mentpy/optimizers/sgd.py:49-59)
// has its own kernel: v = momentum v - lr g; x += v (or the Nesterov form).
__global__ void k_adam(float* __restrict__ x, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       const int32_t* __restrict__ step) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = float(*step + 1);
  const float c1 = 1.0f / (1.0f - powf(b1, t));
  const float c2 = 1.0f / (1.0f - powf(b2, t));
  const float gi = g[i];
  const float mi = fmaf(b1, m[i], (1.0f - b1) * gi);
  const float vi = fmaf(b2, v[i], (1.0f - b2) * gi * gi);
  m[i] = mi;
  v[i] = vi;
  x[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_count(int32_t* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ void k_sgd(float* __restrict__ x, const float* __restrict__ g, float* __restrict__ v,
                      int64_t n, float lr, float momentum, int nesterov) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float vi = fmaf(momentum, v[i], -lr * gi);
  v[i] = vi;
  x[i] += nesterov ? fmaf(momentum, vi, -lr * gi) : vi;
}

// ------------------------------------------------------- step-by-step windows
// Window index r -> live register index j, the appended qubit's bit a and the
// sign of its CZs (see WindowMap).
__device__ __forceinline__ int window_source(const WindowMap& m, int r, uint32_t* sbit) {
  int j = 0, a = 0;
  for (int i = 0; i < m.W; ++i) {
    const int bit = (r >> (m.W - 1 - i)) & 1;
    const int s = m.slot[i];
    if (s >= 0)
      j |= bit << s;
    else
      a = bit;
  }
  *sbit = a ? czbit(uint32_t(r), m.nb) : 0u;
  return j;
}

// Reference window after k_steps measurements, statevector: one workgroup per
// simulation, live state in dynamic LDS (2^n_slots amplitudes), step j at the
// caller's angle step_angles[b * step_stride + j].
// The body, on the statevector st (LDS or a global-memory slice) of
// simulation g, T threads, red (4 floats of LDS) as scratch.
template <int T>
__device__ __forceinline__ void sv_window_run(float2* st, float* red, const Job& J,
                                              const float* __restrict__ step_angles,
                                              int64_t step_stride, int32_t k_steps,
                                              const WindowMap& map, uint32_t flags, int64_t g,
                                              int tid) {
  const int ns = J.ns, H = 1 << ns;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = step_angles + bi * step_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_svres;
  const float2* psi = J.psi + si * J.in_stride;
  for (int j = tid; j < H; j += T) st[j] = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
  RefPhase rp;
  const bool ref = flags & MBQC_FLAG_REF_PHASE;
  for (int k = 0; k < k_steps; ++k) {
    if (k % ns == 0 && resc[k / ns]) {
      __syncthreads();
      for (int j = tid; j < H; j += T) st[j] = cscale(st[j], 0x1p-32f);
    }
    const mbqc_step step = steps[k];
    const float2 b = ref ? step_coeff<true>(ang[k], &rp) : step_coeff<false>(ang[k], nullptr);
    __syncthreads();
    sv_lds_step<T>(st, H, step, b, tid);
  }
  __syncthreads();
  const float2 ph = ref ? rp.value() : make_float2(1.0f, 0.0f);
  float n2 = 0.0f;
  for (int j = tid; j < H; j += T) n2 = fmaf(st[j].x, st[j].x, fmaf(st[j].y, st[j].y, n2));
  n2 = block_sum<T>(n2, red);
  bool appended = false;
  for (int i = 0; i < map.W; ++i) appended |= map.slot[i] < 0;
  const float2 f = cscale(ph, (appended ? 0.70710678118654752440f : 1.0f) / sqrtf(n2));
  const int D = 1 << map.W;
  float2* o = J.out + g * int64_t(D);
  for (int r = tid; r < D; r += T) {
    uint32_t sb;
    const int j = window_source(map, r, &sb);
    o[r] = flip(cmul(st[j], f), sb);
  }
}

__global__ __launch_bounds__(64) void k_sv_window(Job J, const float* __restrict__ step_angles,
                                                  int64_t step_stride, int32_t k_steps,
                                                  WindowMap map, uint32_t flags) {
  extern __shared__ float2 st[];
  __shared__ float red[4];
  sv_window_run<64>(st, red, J, step_angles, step_stride, k_steps, map, flags, blockIdx.x,
                    threadIdx.x);
}

// Windows above the LDS: the same on statevectors in a global workspace
// (ws[(g - g0) * 2^n_slots + j], one slice per workgroup).
__global__ __launch_bounds__(256) void k_sv_window_global(Job J, const float* __restrict__ step_angles,
                                                          int64_t step_stride, int32_t k_steps,
                                                          WindowMap map, uint32_t flags,
                                                          float2* __restrict__ ws, int64_t g0) {
  __shared__ float red[4];
  const int64_t l = blockIdx.x;
  sv_window_run<256>(ws + (l << J.ns), red, J, step_angles, step_stride, k_steps, map, flags, g0 + l,
                     threadIdx.x);
}

// The same for the density matrix (state in dynamic LDS, 4^n_slots entries);
// the measured qubits' noise is applied as in the full run, output noise is not.
template <int T>
__device__ __forceinline__ void dm_window_run(float2* r, float* red, const Job& J,
                                              const float* __restrict__ step_angles,
                                              int64_t step_stride, int32_t k_steps,
                                              const WindowMap& map, const DmNoise& nz, int64_t g,
                                              int tid) {
  const int ns = J.ns, H = 1 << ns;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = step_angles + bi * step_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const int32_t* resc = J.tab + J.off_dmres;
  const float2* psi = J.psi + si * J.in_stride;
  for (int e = tid; e < H * H; e += T) {
    const int j = e / H, k = e % H;
    const float2 vj = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
    const float2 vk = flip(psi[in_index[k]], sign[k] ? 0x80000000u : 0u);
    r[e] = cmul(vj, cconj(vk));
  }
  for (int k = 0; k < k_steps; ++k) {
    if (k % ns == 0 && resc[k / ns]) {
      __syncthreads();
      for (int e = tid; e < H * H; e += T) r[e] = cscale(r[e], 0x1p-32f);
    }
    const mbqc_step step = steps[k];
    const float2 gb = cscale(step_coeff<false>(ang[k], nullptr), nz.gamma);
    __syncthreads();
    dm_lds_step<T>(r, H, int((step.flags >> 8) & 0xffu), step.flags, step.nmask, gb, tid);
  }
  __syncthreads();
  float tr = 0.0f;
  for (int j = tid; j < H; j += T) tr += r[j * H + j].x;
  tr = block_sum<T>(tr, red);
  bool appended = false;
  for (int i = 0; i < map.W; ++i) appended |= map.slot[i] < 0;
  const float scale = (appended ? 0.5f : 1.0f) / tr;
  const int D = 1 << map.W;
  float2* o = J.out + g * int64_t(D) * D;
  for (int e = tid; e < D * D; e += T) {
    const int row = e / D, col = e % D;
    uint32_t sr, sc;
    const int jr = window_source(map, row, &sr), jc = window_source(map, col, &sc);
    o[e] = flip(cscale(r[jr * H + jc], scale), sr ^ sc);
  }
}

__global__ __launch_bounds__(64) void k_dm_window(Job J, const float* __restrict__ step_angles,
                                                  int64_t step_stride, int32_t k_steps,
                                                  WindowMap map, DmNoise nz) {
  extern __shared__ float2 r[];
  __shared__ float red[4];
  dm_window_run<64>(r, red, J, step_angles, step_stride, k_steps, map, nz, blockIdx.x, threadIdx.x);
}

// Windows above the LDS: the same on density matrices in a global workspace.
__global__ __launch_bounds__(256) void k_dm_window_global(Job J, const float* __restrict__ step_angles,
                                                          int64_t step_stride, int32_t k_steps,
                                                          WindowMap map, DmNoise nz,
                                                          float2* __restrict__ ws, int64_t g0) {
  __shared__ float red[4];
  const int64_t l = blockIdx.x;
  dm_window_run<256>(ws + (l << (2 * J.ns)), red, J, step_angles, step_stride, k_steps, map, nz,
                     g0 + l, threadIdx.x);
}

// ------------------------------------------------- measurements in any plane
// The reference's density-matrix step in every plane of Ment.matrix
// (mentpy/simulators/np_simulator_dm.py:307-346, mentpy/operators/ment.py:218-253)
// with the step's outcome probability evaluated on the device: outcome 0
// unless its probability is below 1e-4 (then 1), the state divided by the
// outcome probability, and Z-plane nodes traced out without projection (the
// reference's mode="expectation").  The +1 outcome of a node projects on
// (I + n.sigma)/2 with the Bloch vector n of its plane at angle theta
// (XY: (c, s, 0), XZ: (c, 0, s), YZ: (0, c, s)).  With the measured qubit's
// noise folded in (populations (d00 B00 + d01 B11, d10 B00 + d11 B11),
// coherences times coh) and the readout flip f = 1 - 2 p_m, outcome o (sign
// g = +-1) maps the quartet blocks B_ab of the measured slot to
//   P = W0 B00 + W1 B11,  M = W0 B00 - W1 B11,  V = C B10,  U = conj(C) B01,
//   W0 = (1 + f g n_z) d00 + (1 - f g n_z) d10,
//   W1 = (1 + f g n_z) d01 + (1 - f g n_z) d11,  C = f g coh (n_x - i n_y),
// the same quartet form as the XY kernels (which have W0 = W1 = 1).  The
// probability of outcome 0 (noise included, readout flip not) is
//   p0 = (W0' t0 + W1' t1 + 2 coh (n_x x + n_y y)) / (2 (t0 + t1))
// with W0', W1' at f g = 1, t0 / t1 the traces of B00 / B11 and x + iy that of
// B10 (zero when the appended qubit's CZ with the measured one comes first);
// the mixed post-selected probability is p_m + f p_o.  Every step divides
// by it (and by the trace and the appended |+><+|'s factor 2), so the state
// keeps unit trace and needs no growth tracking.
__device__ __forceinline__ void plane_bloch(uint32_t plane, float th, float* nx, float* ny,
                                            float* nz) {
  float s, c;
  sincos_poly(th, &s, &c);
  *nx = plane == MBQC_PLANE_XY || plane == MBQC_PLANE_XZ ? c : 0.0f;
  *ny = plane == MBQC_PLANE_XY ? s : (plane == MBQC_PLANE_YZ ? c : 0.0f);
  *nz = plane == MBQC_PLANE_XZ || plane == MBQC_PLANE_YZ ? s : (plane == MBQC_PLANE_Z ? 1.0f : 0.0f);
}

// dm_lds_step with the population weights W0, W1 and coherence coefficient C
// of any plane (see above).
template <int T>
__device__ __forceinline__ void dm_lds_step_general(float2* r, int H, int q, uint32_t flags,
                                                    uint32_t nmask, float2 C, float w0, float w1,
                                                    int tid) {
  const int HQ = H / 2;
  const float2 Cc = cconj(C);
  const int QB = 1 << q;
  const bool app = flags & 1u, sneg = flags & 2u;
  for (int i = tid; i < HQ * HQ; i += T) {
    const int j = insert_zero(i / HQ, q), kk = insert_zero(i % HQ, q);
    const float2 B00 = r[j * H + kk], B01 = r[j * H + (kk | QB)];
    const float2 B10 = r[(j | QB) * H + kk], B11 = r[(j | QB) * H + (kk | QB)];
    const float2 U = cmul(Cc, B01), V = cmul(C, B10);
    const float2 P = make_float2(fmaf(w0, B00.x, w1 * B11.x), fmaf(w0, B00.y, w1 * B11.y));
    const float2 r00 = cadd(P, cadd(U, V));
    if (app) {
      const uint32_t sj = czbit(j, nmask), sk = czbit(kk, nmask);
      if (sneg) {
        const float2 M = make_float2(fmaf(w0, B00.x, -w1 * B11.x), fmaf(w0, B00.y, -w1 * B11.y));
        r[j * H + kk] = r00;
        r[j * H + (kk | QB)] = flip(cadd(M, csub(V, U)), sk);
        r[(j | QB) * H + kk] = flip(cadd(M, csub(U, V)), sj);
        r[(j | QB) * H + (kk | QB)] = flip(csub(P, cadd(U, V)), sj ^ sk);
      } else {
        r[j * H + kk] = r00;
        r[j * H + (kk | QB)] = flip(r00, sk);
        r[(j | QB) * H + kk] = flip(r00, sj);
        r[(j | QB) * H + (kk | QB)] = flip(r00, sj ^ sk);
      }
    } else {
      r[j * H + kk] = r00;
      r[j * H + (kk | QB)] = make_float2(0.f, 0.f);
      r[(j | QB) * H + kk] = make_float2(0.f, 0.f);
      r[(j | QB) * H + (kk | QB)] = make_float2(0.f, 0.f);
    }
  }
}

// One simulation per 64-thread workgroup, density matrix in dynamic LDS
// (4^n_slots entries, n_slots <= 6), k_steps steps.  Forward mode
// (window_mode = 0): angles from the angle table, output noise, output in
// out_pos order.  Window mode: per-step angles step_angles[b * step_stride + k]
// and the reference's window after k_steps measurements (as k_dm_window).
// If `outcomes` is not null, step k's outcome (0 or 1; for a Z node the
// expectation p1 = 1 - p0) goes to outcomes[g * k_steps + k].
// The body, on a density matrix r (LDS or a global-memory slice) of
// simulation g, with T threads and red (4 floats of LDS) as scratch.
template <int T>
__device__ __forceinline__ void dm_general_run(float2* r, float* red, const Job& J,
                                               const float* __restrict__ step_angles,
                                               int64_t step_stride, int32_t k_steps,
                                               const WindowMap& map, int32_t window_mode,
                                               const DmNoise& nz, uint32_t flags,
                                               float* __restrict__ outcomes, int64_t g, int tid) {
  const int ns = J.ns, H = 1 << ns, HQ = H / 2;
  const int64_t bi = g / J.S, si = g - bi * J.S;
  const float* ang = window_mode ? step_angles + bi * step_stride : J.angles + bi * J.angle_stride;
  const mbqc_step* steps = reinterpret_cast<const mbqc_step*>(J.tab);
  const int32_t* in_index = J.tab + J.off_in;
  const int32_t* sign = J.tab + J.off_sign;
  const float2* psi = J.psi + si * J.in_stride;
  for (int e = tid; e < H * H; e += T) {
    const int j = e / H, k = e % H;
    const float2 vj = flip(psi[in_index[j]], sign[j] ? 0x80000000u : 0u);
    const float2 vk = flip(psi[in_index[k]], sign[k] ? 0x80000000u : 0u);
    r[e] = cmul(vj, cconj(vk));
  }
  const float pm = 0.5f * (1.0f - nz.flip);
  for (int k = 0; k < k_steps; ++k) {
    const mbqc_step step = steps[k];
    const int q = int((step.flags >> 8) & 0xffu);
    const uint32_t plane = (step.flags >> 2) & 3u;
    const int QB = 1 << q;
    float th = 0.0f;
    if (plane != MBQC_PLANE_Z) {
      if (window_mode) th = ang[k];
      else th = step.angle_index >= 0 ? ang[int64_t(step.angle_index) * J.angle_stride_a]
                                      : step.fixed_angle;
    }
    __syncthreads();
    // traces of the measured slot's diagonal blocks
    float t0 = 0.f, t1 = 0.f, x = 0.f, y = 0.f;
    for (int a = tid; a < HQ; a += T) {
      const int j = insert_zero(a, q);
      t0 += r[j * H + j].x;
      t1 += r[(j | QB) * H + (j | QB)].x;
      const float2 b10 = r[(j | QB) * H + j];
      x += b10.x;
      y += b10.y;
    }
    t0 = block_sum<T>(t0, red);
    t1 = block_sum<T>(t1, red);
    x = block_sum<T>(x, red);
    y = block_sum<T>(y, red);
    float nx, ny, nzz;
    plane_bloch(plane, th, &nx, &ny, &nzz);
    const float tt = t0 + t1;
    const float w0p = (1.0f + nzz) * nz.d00 + (1.0f - nzz) * nz.d10;
    const float w1p = (1.0f + nzz) * nz.d01 + (1.0f - nzz) * nz.d11;
    // a CZ between the appended |+> and the measured qubit before the
    // measurement (flags bit 1) dephases the measured qubit: its coherences
    // then have zero trace and do not enter the probability
    const float coh_term = (step.flags & 2u) ? 0.0f : 2.0f * nz.coh * (nx * x + ny * y);
    const float p0 = (w0p * t0 + w1p * t1 + coh_term) / (2.0f * tt);
    float w0, w1, outcome, denom;
    float2 C;
    if (plane == MBQC_PLANE_Z) {
      w0 = w1 = 1.0f;
      C = make_float2(0.f, 0.f);
      outcome = 1.0f - p0;
      denom = tt;
    } else {
      const bool one = p0 < 1e-4f;
      const float gf = one ? -nz.flip : nz.flip;
      w0 = (1.0f + gf * nzz) * nz.d00 + (1.0f - gf * nzz) * nz.d10;
      w1 = (1.0f + gf * nzz) * nz.d01 + (1.0f - gf * nzz) * nz.d11;
      C = make_float2(gf * nz.coh * nx, -gf * nz.coh * ny);
      outcome = one ? 1.0f : 0.0f;
      const float po = one ? 1.0f - p0 : p0;
      denom = 2.0f * (pm + nz.flip * po) * tt;
    }
    const float sc = 1.0f / ((step.flags & 1u) ? 2.0f * denom : denom);
    w0 *= sc;
    w1 *= sc;
    C = cscale(C, sc);
    if (outcomes && tid == 0)
      outcomes[g * int64_t(window_mode ? k_steps : J.n_steps) + k] = outcome;
    __syncthreads();
    dm_lds_step_general<T>(r, H, q, step.flags, step.nmask, C, w0, w1, tid);
  }
  __syncthreads();
  if (window_mode) {
    float tr = 0.0f;
    for (int j = tid; j < H; j += T) tr += r[j * H + j].x;
    tr = block_sum<T>(tr, red);
    bool appended = false;
    for (int i = 0; i < map.W; ++i) appended |= map.slot[i] < 0;
    const float scale = (appended ? 0.5f : 1.0f) / tr;
    const int D = 1 << map.W;
    float2* o = J.out + g * int64_t(D) * D;
    for (int e = tid; e < D * D; e += T) {
      const int row = e / D, col = e % D;
      uint32_t sr, sc2;
      const int jr = window_source(map, row, &sr), jc = window_source(map, col, &sc2);
      o[e] = flip(cscale(r[jr * H + jc], scale), sr ^ sc2);
    }
    return;
  }
  // noise of every output qubit
  for (int m = 0; m < ns; ++m) {
    if (!(uint32_t(J.live_mask) & (1u << m)) || !nz.out_noise) continue;
    const int MB = 1 << m;
    __syncthreads();
    for (int i = tid; i < HQ * HQ; i += T) {
      const int j = insert_zero(i / HQ, m), kk = insert_zero(i % HQ, m);
      const float2 a = r[j * H + kk], d = r[(j | MB) * H + (kk | MB)];
      r[j * H + kk] = make_float2(fmaf(nz.d00, a.x, nz.d01 * d.x), fmaf(nz.d00, a.y, nz.d01 * d.y));
      r[(j | MB) * H + (kk | MB)] =
          make_float2(fmaf(nz.d10, a.x, nz.d11 * d.x), fmaf(nz.d10, a.y, nz.d11 * d.y));
      r[j * H + (kk | MB)] = cscale(r[j * H + (kk | MB)], nz.coh);
      r[(j | MB) * H + kk] = cscale(r[(j | MB) * H + kk], nz.coh);
    }
  }
  __syncthreads();
  const int32_t* out_pos = J.tab + J.off_out;
  float scale = 1.0f;
  if (flags & MBQC_FLAG_RENORMALIZE) {
    float tr = 0.0f;
    for (int j = tid; j < H; j += T)
      if (out_pos[j] >= 0) tr += r[j * H + j].x;
    tr = block_sum<T>(tr, red);
    scale = 1.0f / tr;
  }
  const int64_t d = int64_t(1) << J.n_out;
  float2* o = J.out + g * d * d;
  for (int e = tid; e < H * H; e += T) {
    const int pj = out_pos[e / H], pk = out_pos[e % H];
    if (pj >= 0 && pk >= 0) o[int64_t(pj) * d + pk] = cscale(r[e], scale);
  }
}

__global__ __launch_bounds__(64) void k_dm_general(Job J, const float* __restrict__ step_angles,
                                                   int64_t step_stride, int32_t k_steps,
                                                   WindowMap map, int32_t window_mode, DmNoise nz,
                                                   uint32_t flags, float* __restrict__ outcomes) {
  extern __shared__ float2 r[];
  __shared__ float red[4];
  dm_general_run<64>(r, red, J, step_angles, step_stride, k_steps, map, window_mode, nz, flags,
                     outcomes, xcd_block(blockIdx.x, gridDim.x), threadIdx.x);
}

// Windows above the LDS: the same body on density matrices in a global
// workspace, ws[(g - g0) * 4^n_slots + e] for simulations g0 + blockIdx (the
// workgroup's own slice; its threads synchronise through the barrier as in
// LDS).
__global__ __launch_bounds__(256) void k_dm_general_global(Job J, const float* __restrict__ step_angles,
                                                           int64_t step_stride, int32_t k_steps,
                                                           WindowMap map, int32_t window_mode,
                                                           DmNoise nz, uint32_t flags,
                                                           float* __restrict__ outcomes,
                                                           float2* __restrict__ ws, int64_t g0) {
  __shared__ float red[4];
  const int64_t l = xcd_block(blockIdx.x, gridDim.x);
  dm_general_run<256>(ws + l * (int64_t(1) << (2 * J.ns)), red, J, step_angles, step_stride, k_steps,
                      map, window_mode, nz, flags, outcomes, g0 + l, threadIdx.x);
}


// ---------------------------------------------------------------- batched
struct BatchItem {
  int32_t job;
  int32_t count;   // simulations in this item
  int64_t first;   // first simulation index within the job
};

// Persistent kernel over the work items of one live-window size NS: every
// wave pulls items (runs of simulations of one job) from its group's device
// counter until none are left.  Items are sorted by decreasing cost on the
// host, so long patterns start first.  One kernel per NS keeps each group at
// the register footprint (and so the occupancy) of its own window size.
// A Job read from device memory (rather than passed as a kernel argument)
// holds generic pointers: loads and stores through them compile to flat
// instructions, and the plan-table loads could not be proven free of the
// kernel's own stores, so they would all be per-lane vector loads.  The same
// layout with the pointers typed as global memory, and the read-only plan
// tables as constant memory, restores global vector instructions and scalar
// table loads.
struct GlobalJob {
  const __attribute__((address_space(4))) int32_t* tab;
  int32_t off_in, off_sign, off_out, off_svres, off_dmres, e_sv, e_dm;
  int32_t ns, n_steps, n_out, live_mask, off_coef;
  int32_t off_aidx, off_run;
  int32_t fast_fetch;
  const __attribute__((address_space(1))) float* angles;
  int64_t angle_stride;
  int64_t angle_stride_a;
  const __attribute__((address_space(1))) float2* psi;
  int64_t in_stride;
  __attribute__((address_space(1))) float2* out;
  __attribute__((address_space(1))) float2* phase_out;
  int64_t S, total;
};
static_assert(sizeof(GlobalJob) == sizeof(Job), "GlobalJob mirrors Job");

__device__ __forceinline__ Job load_job(const Job* jobs, int i) {
  const GlobalJob g = reinterpret_cast<const GlobalJob*>(jobs)[i];
  Job j;
  j.tab = (const int32_t*)g.tab;
  j.off_in = g.off_in;
  j.off_sign = g.off_sign;
  j.off_out = g.off_out;
  j.off_svres = g.off_svres;
  j.off_dmres = g.off_dmres;
  j.e_sv = g.e_sv;
  j.e_dm = g.e_dm;
  j.ns = g.ns;
  j.n_steps = g.n_steps;
  j.n_out = g.n_out;
  j.live_mask = g.live_mask;
  j.off_coef = g.off_coef;
  j.off_aidx = g.off_aidx;
  j.off_run = g.off_run;
  j.fast_fetch = g.fast_fetch;
  j.angles = (const float*)g.angles;
  j.angle_stride = g.angle_stride;
  j.angle_stride_a = g.angle_stride_a;
  j.psi = (const float2*)g.psi;
  j.in_stride = g.in_stride;
  j.out = (float2*)g.out;
  j.phase_out = (float2*)g.phase_out;
  j.S = g.S;
  j.total = g.total;
  return j;
}

// Output stores of the batched kernel: direct (staging the 32-amplitude rows
// through LDS tiles measured slower here: 3.54e9 -> 3.27e9 sims/s on the
// 64-grid batch, whose window-size groups run concurrently).
template <int NS>
constexpr int kBatchStore = kStoreDirect;

template <bool REF, int NS, bool SH, bool PERIODIC>
__global__ __launch_bounds__(256) void k_batch(const Job* __restrict__ jobs,
                                               const BatchItem* __restrict__ items, int64_t n_items,
                                               int* counter, uint32_t flags) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const BatchItem I = items[it];
    const Job J = load_job(jobs, I.job);
    for (int off = 0; off < I.count; off += 64) {
      const int64_t g = I.first + off + lane;
      const int64_t left = J.total - (I.first + off);
      const int rows = int(std::min<int64_t>(std::min<int64_t>(I.count - off, left), 64));
      // direct stores: staging tiles would cost the concurrently running
      // window-size groups LDS occupancy (measured slower)
      sv_sim<NS, REF, kBatchStore<NS>, false, SH, PERIODIC, false>(
          J, g, (off + lane) < I.count && g < J.total, rows, flags);
    }
  }
}

// Batched groups of 7..14 live slots: a persistent workgroup per resident
// slot pulls items from the group's counter and runs their simulations one
// after the other with the state in LDS (sv_lds_run, as k_sv_forward_lds).
// Every thread reads the item index from LDS, so all leave the loop together
// once the counter passes n_items.
template <int NS, int T>
__global__ __launch_bounds__(T) void k_batch_lds(const Job* __restrict__ jobs,
                                                 const BatchItem* __restrict__ items,
                                                 int64_t n_items, int* counter, uint32_t flags) {
  __shared__ float2 st[1 << NS];
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  __shared__ int item_s;
  const int tid = threadIdx.x;
  for (;;) {
    __syncthreads();   // the previous item's last reads of item_s and st are done
    if (tid == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= n_items) break;
    const BatchItem I = items[it];
    const Job J = load_job(jobs, I.job);
    const int64_t end = std::min<int64_t>(I.first + I.count, J.total);
    for (int64_t g = I.first; g < end; ++g) {
      __syncthreads();
      sv_lds_run<NS, T>(J, g, flags, st, red, th_s, tid);
    }
  }
}

// Groups above the LDS (15..20 slots): the same with each persistent
// workgroup's state in its own slice of a global workspace,
// ws[blockIdx * 2^NS + j].
template <int NS>
__global__ __launch_bounds__(256) void k_batch_global(const Job* __restrict__ jobs,
                                                      const BatchItem* __restrict__ items,
                                                      int64_t n_items, int* counter, uint32_t flags,
                                                      float2* __restrict__ ws) {
  __shared__ float red[4];
  __shared__ float th_s[kAngleChunk];
  __shared__ int item_s;
  const int tid = threadIdx.x;
  float2* st = ws + (int64_t(blockIdx.x) << NS);
  for (;;) {
    __syncthreads();   // the previous item's last reads of item_s and st are done
    if (tid == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int it = item_s;
    if (it >= n_items) break;
    const BatchItem I = items[it];
    const Job J = load_job(jobs, I.job);
    const int64_t end = std::min<int64_t>(I.first + I.count, J.total);
    for (int64_t g = I.first; g < end; ++g) {
      __syncthreads();
      sv_lds_run<NS, 256>(J, g, flags, st, red, th_s, tid);
    }
  }
}

// group counters live kCounterStride ints apart (separate cache lines)
constexpr int kCounterStride = 64;

__global__ void k_reset_counters(int* counter, int n) {
  const int i = threadIdx.x;
  if (blockIdx.x == 0 && i < n) counter[i * kCounterStride] = 0;
}

}  // namespace mbqc

// ==================================================================== C ABI
using namespace mbqc;
// Static device scratch of mbqc_gram_loss_seed_c64_noalloc (loss partials):
// 1 MiB on every device the library runs on, held until the process ends.
// Calls that overlap on one device share it and race.
constexpr int64_t kScratchFloats = int64_t(1) << 18;
static __device__ __align__(256) float scratch[kScratchFloats];


namespace {

constexpr int kMaxSvSlots = 6;       // register-resident SV kernels
constexpr int kMaxBatchSlots = 20;   // batched groups (LDS above kMaxSvSlots, global above 14)
// below this many simulations the DM path keeps each density matrix in LDS
// (one quartet per thread, several simulations per workgroup); the two
// kernels take the same time at 32,768 simulations of the 3x12 grid
constexpr int64_t kDmLaneKernelMinSims = 32768;
// up to this many simulations (by live-window size 0..3) the DM path holds
// one density-matrix entry per lane (k_dm_forward_wave).  Measured per launch
// (wave / LDS / one-simulation-per-lane kernels, µs): 3 slots at 1,024
// simulations 9.8 / 10.4 / 38.6 and at 4,096 16.1 / 12.4 / 39.1; 2 slots at
// 16,384 10.2 / 10.5 / 12.4 and at 32,768 15.9 / 13.1 / 13.1; 1 slot at 32,768
// 4.8 / - / 5.1.
constexpr int64_t kDmWaveKernelMaxSims[4] = {0, 32768, 16384, 1024};
constexpr int kMaxSvLdsSlots = 14;   // LDS-resident SV kernel (2^14 amplitudes: 128 KB)
constexpr int kMaxDmSlots = 3;       // register-resident DM kernel
constexpr int kMaxDmLdsSlots = 7;    // LDS-resident DM kernel (4^7 entries: 128 KB)
// Above the LDS the state lives in a stream-ordered workspace (hipMallocAsync
// on the caller's stream, freed on it after the launch), at most
// kGlobalWorkspaceBytes per launch (larger batches run in chunks).
constexpr int kMaxSvGlobalSlots = 20;   // 2^20 amplitudes: 8 MiB per simulation
constexpr int kMaxAdjGlobalSlots = 16;  // adjoint: (2 + 16/2) x 2^16 amplitudes per row, 5 MiB
constexpr int kMaxDmGlobalSlots = 10;   // 4^10 entries: 8 MiB per simulation
constexpr size_t kGlobalWorkspaceBytes = size_t(4) << 30;

// Stream-ordered workspace of the launches beyond the LDS.  Under stream
// capture a HIP graph records it as allocation and free nodes, and replays of
// such a graph were measured to hand the kernels a stale workspace (the first
// replay right, later ones wrong), so these launches refuse capture.
hipError_t global_ws_alloc(void** ws, size_t bytes, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cs);
  if (e != hipSuccess) return e;
  if (cs != hipStreamCaptureStatusNone) return hipErrorStreamCaptureUnsupported;
  return hipMallocAsync(ws, bytes, st);
}

int hip_fail(hipError_t e, const char* what);

int global_fail(hipError_t e, const char* what) {
  if (e == hipErrorStreamCaptureUnsupported)
    return fail(MBQC_ERR_NOT_IMPLEMENTED,
                std::string(what) + ": live windows beyond the LDS use a workspace allocated per "
                                    "call and cannot be captured in a HIP graph");
  return hip_fail(e, what);
}

// Launch `kernel_chunk(ws, g0, n)` over the job's simulations in chunks whose
// states (state_bytes each) fit the workspace.
template <typename F>
hipError_t run_global_chunks(const Job& J, size_t state_bytes, hipStream_t st, F&& kernel_chunk) {
  const int64_t per = std::max<int64_t>(
      1, std::min<int64_t>(int64_t(kGlobalWorkspaceBytes / state_bytes), 0x7fffffff));
  const int64_t n0 = std::min<int64_t>(per, J.total);
  void* ws = nullptr;
  hipError_t e = global_ws_alloc(&ws, size_t(n0) * state_bytes, st);
  if (e != hipSuccess) return e;
  for (int64_t g0 = 0; g0 < J.total && e == hipSuccess; g0 += n0) {
    kernel_chunk(static_cast<float2*>(ws), g0, std::min<int64_t>(n0, J.total - g0));
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(ws, st);
  return e != hipSuccess ? e : ef;
}

template <int NS>
hipError_t launch_sv_global(const Job& J, uint32_t flags, hipStream_t st) {
  return run_global_chunks(J, sizeof(float2) << NS, st, [&](float2* ws, int64_t g0, int64_t n) {
    hipLaunchKernelGGL((k_sv_forward_global<NS>), dim3(unsigned(n)), dim3(256), 0, st, J, flags, ws,
                       g0);
  });
}

hipError_t launch_dm_global(const Job& J, const DmNoise& nz, uint32_t flags, float* outcomes,
                            hipStream_t st) {
  const WindowMap none{};
  return run_global_chunks(J, sizeof(float2) << (2 * J.ns), st,
                           [&](float2* ws, int64_t g0, int64_t n) {
    hipLaunchKernelGGL(k_dm_general_global, dim3(unsigned(n)), dim3(256), 0, st, J,
                       static_cast<const float*>(nullptr), int64_t(0), int32_t(J.n_steps), none, 0,
                       nz, flags, outcomes, ws, g0);
  });
}
constexpr int kMaxAdjSlots = 5;

int hip_fail(hipError_t e, const char* what) {
  return fail(MBQC_ERR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

// Dynamic LDS above 64 KB (windows of 14 SV / 7 DM slots take 128 KB of the
// 160 KB per workgroup) is requested from the runtime for the kernel first.
int allow_lds(const void* kernel, size_t bytes) {
  if (bytes <= 65536) return MBQC_OK;
  hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "hipFuncSetAttribute(dynamic LDS)");
}

// Upload (once per device) and return the plan's packed tables.
int device_tables(Plan* p, const int32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(MBQC_ERR_RUNTIME, "device id out of range");
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->dev_tab[dev]) {
    std::vector<int32_t> t = pack_device_tables(*p);
    void* d = nullptr;
    e = hipMalloc(&d, t.size() * sizeof(int32_t));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(plan tables)");
    e = hipMemcpy(d, t.data(), t.size() * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return hip_fail(e, "hipMemcpy(plan tables)");
    }
    p->dev_tab[dev] = d;
  }
  *out = static_cast<const int32_t*>(p->dev_tab[dev]);
  return MBQC_OK;
}

// Upload at plan creation (current device; nothing to do on a host without
// a GPU, where launches fail later anyway).
int upload_at_create(Plan* p) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();   // leave no error behind for a later check
    return MBQC_OK;
  }
  const int32_t* tab = nullptr;
  return device_tables(p, &tab);
}

struct RegisterUploadHook {
  RegisterUploadHook() { g_plan_upload_hook = upload_at_create; }
} register_upload_hook;

// The statevector kernels measure in the XY plane only, as the reference's
// statevector backend does (np_simulator_sv.py:54-59, 329-332).
int check_xy_plan(const Plan* p) {
  if (p->general_planes)
    return fail(MBQC_ERR_VALUE, "the pattern has XZ, YZ or Z measurements, but only XY plane is "
                                "supported by the statevector kernels");
  return MBQC_OK;
}

// Adjoint sweeps write one gradient entry per step, so an angle shared by two
// measured nodes would keep only one node's contribution.
int check_adjoint_plan(const Plan* p) {
  int rc = check_xy_plan(p);
  if (rc != MBQC_OK) return rc;
  if (p->ns > kMaxAdjGlobalSlots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "adjoint kernels support live windows of at most " +
                                              std::to_string(kMaxAdjGlobalSlots) + " qubits");
  if (p->shared_angles)
    return fail(MBQC_ERR_NOT_IMPLEMENTED,
                "adjoint gradients need every angle-vector entry to belong to at most one "
                "measured node");
  if (!p->rotating)
    return fail(MBQC_ERR_NOT_IMPLEMENTED,
                "adjoint gradients need the nodes measured in schedule order");
  return MBQC_OK;
}

int make_job(Plan* p, const float* angles, int64_t angle_stride, int64_t angle_stride_a,
             const void* psi, int64_t in_stride, void* out, int64_t B, int64_t S, Job* J) {
  if (B < 0 || S <= 0) return fail(MBQC_ERR_VALUE, "batch sizes must be B >= 0, S >= 1");
  if (B > 0) {   // an empty batch launches nothing and may pass null pointers
    if (!angles && p->n_angles > 0) return fail(MBQC_ERR_VALUE, "angles pointer is null");
    if (!psi) return fail(MBQC_ERR_VALUE, "input pointer is null");
    if (!out) return fail(MBQC_ERR_VALUE, "output pointer is null");
  }
  if (angle_stride < 0 || angle_stride_a < 0 || in_stride < 0)
    return fail(MBQC_ERR_VALUE, "negative stride");
  if (p->n_angles > 1 && angle_stride_a == 0)
    return fail(MBQC_ERR_VALUE, "angle element stride must be positive");
  if (in_stride > 0 && in_stride < (int64_t(1) << p->n_in))
    return fail(MBQC_ERR_VALUE, "input stride shorter than the input state");
  const int32_t* tab = nullptr;
  int rc = device_tables(p, &tab);
  if (rc != MBQC_OK) return rc;
  TableOffsets o = table_offsets(*p);
  Job j{};
  j.tab = tab;
  j.off_in = o.in_index;
  j.off_sign = o.init_sign;
  j.off_out = o.out_pos;
  j.off_svres = o.sv_rescale;
  j.off_aidx = o.angle_idx;
  j.off_run = o.sv_run_end;
  j.off_dmres = o.dm_rescale;
  j.off_coef = o.coef;
  j.e_sv = p->e_sv;
  j.e_dm = p->e_dm;
  j.ns = p->ns;
  j.n_steps = p->n_steps;
  j.n_out = p->n_out;
  uint32_t live = 0;
  for (int s : p->out_slot) live |= 1u << s;
  j.live_mask = int32_t(live);
  j.angles = angles;
  j.angle_stride = angle_stride;
  j.angle_stride_a = angle_stride_a;
  bool fixed = false;
  for (const mbqc_step& st : p->steps) fixed = fixed || st.angle_index < 0;
  j.fast_fetch = !fixed && p->n_angles > 0 &&
                 uint64_t(p->n_angles - 1) * uint64_t(angle_stride_a) <= 0xffffffffull;
  j.psi = static_cast<const float2*>(psi);
  j.in_stride = in_stride;
  j.out = static_cast<float2*>(out);
  j.phase_out = nullptr;
  j.S = S;
  j.total = B * S;
  *J = j;
  return MBQC_OK;
}

template <int NS>
hipError_t launch_sv(const Job& J, bool per_step, bool periodic, uint32_t flags, hipStream_t st) {
  const int threads = 256;
  const int64_t blocks = (J.total + threads - 1) / threads;
  const size_t lds = kForwardStore<NS> == kStoreColumns ? (threads / 64) * kColumnTile * sizeof(float2) : 0;
  const bool ref = flags & MBQC_FLAG_REF_PHASE;
  if (periodic) {
    if (ref)
      hipLaunchKernelGGL((k_sv_forward_periodic<NS, true>), dim3(blocks), dim3(threads), lds, st, J,
                         flags);
    else
      hipLaunchKernelGGL((k_sv_forward_periodic<NS, false>), dim3(blocks), dim3(threads), lds, st, J,
                         flags);
  } else if (per_step) {
    if (ref)
      hipLaunchKernelGGL((k_sv_forward<NS, true, true>), dim3(blocks), dim3(threads), lds, st, J, flags);
    else
      hipLaunchKernelGGL((k_sv_forward<NS, false, true>), dim3(blocks), dim3(threads), lds, st, J, flags);
  } else {
    if (ref)
      hipLaunchKernelGGL((k_sv_forward<NS, true, false>), dim3(blocks), dim3(threads), lds, st, J, flags);
    else
      hipLaunchKernelGGL((k_sv_forward<NS, false, false>), dim3(blocks), dim3(threads), lds, st, J, flags);
  }
  return hipGetLastError();
}

// Phase-covariant single-qubit channel as (coh, a, c): off-diagonal blocks
// times coh, r00 - r11 -> a (r00 - r11) + c (r00 + r11), trace kept.
struct Covariant {
  double coh = 1.0, a = 1.0, c = 0.0;
};
// apply `second` after `first`
Covariant then(const Covariant& first, const Covariant& second) {
  return {first.coh * second.coh, first.a * second.a, second.a * first.c + second.c};
}

// Compose the channels of mbqc_noise in the order depolarizing, amplitude
// damping, phase damping, phase flip, generalized amplitude damping, each with
// its Kraus operators' effect on (coherence, population difference):
//   DepolarizingChannel(p): coh = a = 1 - 4p/3
//   AmplitudeDamping(g):    coh = sqrt(1-g), a = 1-g, c = g
//   PhaseDamping(g):        coh = sqrt(1-g)
//   PhaseFlip(p):           coh = 1 - 2p
//   GeneralizedAmplitudeDamping(g, p): coh = sqrt(1-g), a = 1-g, c = g (2p-1)
DmNoise compose_noise(const mbqc_noise& n) {
  Covariant ch;
  ch = then(ch, {1.0 - 4.0 * n.p_depol / 3.0, 1.0 - 4.0 * n.p_depol / 3.0, 0.0});
  ch = then(ch, {std::sqrt(1.0 - n.amp_damping), 1.0 - n.amp_damping, double(n.amp_damping)});
  ch = then(ch, {std::sqrt(1.0 - n.phase_damping), 1.0, 0.0});
  ch = then(ch, {1.0 - 2.0 * n.phase_flip, 1.0, 0.0});
  ch = then(ch, {std::sqrt(1.0 - n.gad_gamma), 1.0 - n.gad_gamma,
                 double(n.gad_gamma) * (2.0 * n.gad_p - 1.0)});
  DmNoise nz;
  nz.gamma = float(ch.coh * (1.0 - 2.0 * n.p_readout));
  // r00' = (P' + M')/2, r11' = (P' - M')/2 with P' = P, M' = a M + c P
  nz.d00 = float((1.0 + ch.c + ch.a) / 2.0);
  nz.d01 = float((1.0 + ch.c - ch.a) / 2.0);
  nz.d10 = float((1.0 - ch.c - ch.a) / 2.0);
  nz.d11 = float((1.0 - ch.c + ch.a) / 2.0);
  nz.coh = float(ch.coh);
  nz.out_noise = (ch.coh != 1.0 || ch.a != 1.0 || ch.c != 0.0) ? 1 : 0;
  nz.flip = float(1.0 - 2.0 * n.p_readout);
  return nz;
}

template <int NS>
hipError_t launch_dm(const Job& J, const DmNoise& nz, uint32_t flags, bool periodic,
                     hipStream_t st) {
  const int threads = 256;
  const int64_t blocks = (J.total + threads - 1) / threads;
  if (periodic)
    hipLaunchKernelGGL((k_dm_forward<NS, true>), dim3(blocks), dim3(threads), 0, st, J, nz, flags);
  else
    hipLaunchKernelGGL((k_dm_forward<NS, false>), dim3(blocks), dim3(threads), 0, st, J, nz, flags);
  return hipGetLastError();
}

// Blocks per checkpoint of the unsegmented adjoint: from 65,536 columns the
// sweeps are bound by the checkpoint traffic (one 2^NS-amplitude state per
// column and block, written once and read once), so every second block is
// kept and the backward sweep recomputes the other from it.
inline int adj_ckpt_blocks(int64_t n_columns) { return n_columns >= 65536 ? 2 : 1; }

template <int NS>
hipError_t launch_adj_fwd(const Job& J, float2* ckpt, int ck, bool periodic, hipStream_t st) {
  // one wave per workgroup: at small batches the waves spread over all CUs
  // instead of sharing a few CUs' scalar units and instruction issue
  const int threads = 64;
  const int64_t blocks = (J.total + threads - 1) / threads;
  const size_t lds = (threads / 64) * kColumnTile * sizeof(float2);
  if (periodic)
    hipLaunchKernelGGL((k_sv_adj_forward<NS, true>), dim3(blocks), dim3(threads), lds, st, J, ckpt, ck);
  else
    hipLaunchKernelGGL((k_sv_adj_forward<NS, false>), dim3(blocks), dim3(threads), lds, st, J, ckpt, ck);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_adj_bwd(const Job& J, const float2* ckpt, const float2* seed, const AdjCtx& A,
                          bool periodic, hipStream_t st) {
  const int threads = 64;
  const int64_t blocks = (J.total + threads - 1) / threads;
  if (periodic)
    hipLaunchKernelGGL((k_sv_adj_backward<NS, true>), dim3(blocks), dim3(threads), 0, st, J, ckpt,
                       seed, A, A.gpart);
  else
    hipLaunchKernelGGL((k_sv_adj_backward<NS, false>), dim3(blocks), dim3(threads), 0, st, J, ckpt,
                       seed, A, A.gpart);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_adj_fwd_lds(const Job& J, float2* ckpt, int ck, hipStream_t st) {
  constexpr int T = (1 << (NS - 1)) < 256 ? (1 << (NS - 1)) : 256;
  hipLaunchKernelGGL((k_sv_adj_forward_lds<NS, T>), dim3(unsigned(J.total)), dim3(T), 0, st, J, ckpt,
                     ck);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_adj_fwd_global(const Job& J, float2* ckpt, int ck, hipStream_t st) {
  return run_global_chunks(J, sizeof(float2) << NS, st, [&](float2* ws, int64_t g0, int64_t n) {
    hipLaunchKernelGGL((k_sv_adj_forward_global<NS>), dim3(unsigned(n)), dim3(256), 0, st, J, ckpt,
                       ck, ws, g0);
  });
}

// Backward sweep above the LDS: per partial-sum row a global slice of
// (2 + NS/2) 2^NS amplitudes (allocated and freed in stream order, at most
// kGlobalWorkspaceBytes per launch, rows in chunks) and n_angles floats of
// dynamic LDS.
template <int NS>
hipError_t launch_adj_bwd_global(const Job& J, const float2* ckpt, const float2* seed,
                                 const AdjCtx& A, hipStream_t st) {
  constexpr int64_t H = int64_t(1) << NS;
  const size_t row_bytes = size_t(2 * H + NS * (H / 2)) * sizeof(float2);
  const size_t lds = size_t(A.n_angles) * sizeof(float);
  const void* kern = reinterpret_cast<const void*>(&k_sv_adj_backward_global<NS>);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  const int64_t rows = (J.total / A.C) * A.P;
  const int64_t per = std::max<int64_t>(
      1, std::min<int64_t>(int64_t(kGlobalWorkspaceBytes / row_bytes), 0x7fffffff));
  const int64_t n0 = std::min<int64_t>(per, rows);
  void* ws = nullptr;
  hipError_t e = global_ws_alloc(&ws, size_t(n0) * row_bytes, st);
  if (e != hipSuccess) return e;
  for (int64_t r0 = 0; r0 < rows && e == hipSuccess; r0 += n0) {
    const int64_t n = std::min<int64_t>(n0, rows - r0);
    hipLaunchKernelGGL((k_sv_adj_backward_global<NS>), dim3(unsigned(n)), dim3(256), lds, st, J,
                       ckpt, seed, A, A.gpart, static_cast<float2*>(ws), r0);
    e = hipGetLastError();
  }
  const hipError_t ef = hipFreeAsync(ws, st);
  return e != hipSuccess ? e : ef;
}

// Dynamic LDS of k_sv_adj_backward_lds: lam and cur (2^ns each), a block's
// measured halves (ns x 2^(ns-1)) and the per-angle sums.
inline size_t adj_lds_bytes(int ns, int n_angles) {
  const size_t H = size_t(1) << ns;
  return (2 * H + size_t(ns) * (H / 2)) * sizeof(float2) + size_t(n_angles) * sizeof(float);
}
constexpr size_t kMaxDynLds = 160 * 1024 - 256;   // gfx950: 160 KiB per workgroup, less static LDS

template <int NS>
hipError_t launch_adj_bwd_lds(const Job& J, const float2* ckpt, const float2* seed, const AdjCtx& A,
                              hipStream_t st) {
  constexpr int T = (1 << (NS - 1)) < 256 ? (1 << (NS - 1)) : 256;
  const size_t lds = adj_lds_bytes(NS, A.n_angles);
  const void* kern = reinterpret_cast<const void*>(&k_sv_adj_backward_lds<NS, T>);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  const int64_t rows = (J.total / A.C) * A.P;
  hipLaunchKernelGGL((k_sv_adj_backward_lds<NS, T>), dim3(unsigned(rows)), dim3(T), lds, st, J, ckpt,
                     seed, A, A.gpart);
  return hipGetLastError();
}

// Segment layout of the time-cut adjoint for B parameter sets and n segments;
// returns the workspace size in bytes (0 when n is out of range).
size_t seg_layout(const Plan& p, int64_t B, int n, SegInfo* G) {
  const int nblk = p.n_blocks();
  if (n < 2 || n > kMaxSegments || n > nblk || B < 0) return 0;
  SegInfo g{};
  const int64_t H = int64_t(1) << p.ns;
  g.n = n;
  g.cols0 = 1 << p.n_in;
  g.cols = int32_t(H);
  g.rows_last = 1 << p.n_out;
  g.B = B;
  auto align = [](int64_t x) { return (x + 63) & ~int64_t(63); };
  int64_t off = 0, wave = 0;
  for (int q = 0; q <= n; ++q) g.blk[q] = int32_t((int64_t(q) * nblk) / n);
  for (int q = 0; q < n; ++q) {
    const int64_t C = q == 0 ? g.cols0 : H;
    const int64_t rows = q + 1 == n ? g.rows_last : H;
    const int64_t waves = (B * C + 63) / 64;
    g.wave[q] = wave;
    wave += waves;
    g.ckpt[q] = off;
    off = align(off + int64_t(g.blk[q + 1] - g.blk[q]) * H * waves * 64);
    g.tmat[q] = off;
    off = align(off + B * C * rows);
    g.gmat[q] = off;
    off = align(off + B * C * rows);
    g.smat[q] = off;
    if (q >= 2) off = align(off + B * g.cols0 * H);
  }
  g.wave[n] = wave;
  if (G) *G = g;
  return size_t(off) * sizeof(float2);
}

template <int NS>
hipError_t launch_seg_forward(const Job& J, const SegInfo& G, float2* ws, float2* out, bool periodic,
                              hipStream_t st) {
  if (G.wave[G.n] > 0) {
    if (periodic)
      hipLaunchKernelGGL((k_sv_seg_forward<NS, true>), dim3(G.wave[G.n]), dim3(64),
                         kColumnTile * sizeof(float2), st, J, G, ws, ws);
    else
      hipLaunchKernelGGL((k_sv_seg_forward<NS, false>), dim3(G.wave[G.n]), dim3(64),
                         kColumnTile * sizeof(float2), st, J, G, ws, ws);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && G.B > 0) {
    if constexpr (NS == 4)
      hipLaunchKernelGGL(k_seg_chain_forward_mfma, dim3(G.B), dim3(64), 0, st, G, ws, out);
    else
      hipLaunchKernelGGL((k_seg_chain_forward<NS>), dim3(G.B), dim3(256), 0, st, G, ws, out);
  }
  return e == hipSuccess ? hipGetLastError() : e;
}

template <int NS>
hipError_t launch_seg_backward(const Job& J, const SegInfo& G, float2* ws, const float2* seed,
                               int32_t n_angles, float* grad, bool periodic, hipStream_t st) {
  if (G.B > 0) {
    if constexpr (NS == 4)
      hipLaunchKernelGGL(k_seg_chain_backward_mfma, dim3(G.B), dim3(64), 0, st, G, ws, seed);
    else
      hipLaunchKernelGGL((k_seg_chain_backward<NS>), dim3(G.B), dim3(256), 0, st, G, ws, seed);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && G.wave[G.n] > 0) {
    if (periodic)
      hipLaunchKernelGGL((k_sv_seg_backward<NS, true>), dim3(G.wave[G.n]), dim3(64), 0, st, J, G,
                         ws, n_angles, grad);
    else
      hipLaunchKernelGGL((k_sv_seg_backward<NS, false>), dim3(G.wave[G.n]), dim3(64), 0, st, J, G,
                         ws, n_angles, grad);
  }
  return e == hipSuccess ? hipGetLastError() : e;
}

struct BatchGroupLaunch {
  dim3 grid, block;
  hipStream_t stream;
  const Job* jobs;
  const BatchItem* items;
  int64_t n_items;
  int* counter;
  uint32_t flags;
};

// Lanes g = b*S + s of one angle vector form aligned groups of
// share_group(NS) lanes when every job's S is a multiple of it (work items
// start at multiples of 64 and B*S is then a multiple of the group too).
// `periodic`: every plan of the group is periodic.
template <int NS, bool PER>
void launch_batch_group_p(const BatchGroupLaunch& L, bool ref, bool share) {
  const size_t lds =
      kBatchStore<NS> == kStoreColumns ? (L.block.x / 64) * kColumnTile * sizeof(float2) : 0;
#define MBQC_BATCH_LAUNCH(R, SH)                                                                     \
  hipLaunchKernelGGL((k_batch<R, NS, SH, PER>), L.grid, L.block, lds, L.stream, L.jobs, L.items, \
                     L.n_items, L.counter, L.flags)
  if constexpr (NS >= 2) {
    if (share) {
      if (ref) MBQC_BATCH_LAUNCH(true, true);
      else MBQC_BATCH_LAUNCH(false, true);
      return;
    }
  }
  if (ref) MBQC_BATCH_LAUNCH(true, false);
  else MBQC_BATCH_LAUNCH(false, false);
#undef MBQC_BATCH_LAUNCH
}

template <int NS>
void launch_batch_group(const BatchGroupLaunch& L, bool ref, bool share, bool periodic) {
  if (periodic) launch_batch_group_p<NS, true>(L, ref, share);
  else launch_batch_group_p<NS, false>(L, ref, share);
}

template <int NS>
hipError_t launch_sample(const Job& J, const int32_t* tab, const TableOffsets& o, int32_t* outcomes,
                         int32_t words, uint64_t seed, bool periodic, hipStream_t st) {
  const int threads = 256;
  const int64_t blocks = (J.total + threads - 1) / threads;
  const uint64_t* xm = reinterpret_cast<const uint64_t*>(tab + o.xmask);
  const uint64_t* zm = reinterpret_cast<const uint64_t*>(tab + o.zmask);
  if (periodic)
    hipLaunchKernelGGL((k_sv_sample<NS, true>), dim3(blocks), dim3(threads), 0, st, J, xm, zm,
                       tab + o.out_offset, outcomes, words, seed);
  else
    hipLaunchKernelGGL((k_sv_sample<NS, false>), dim3(blocks), dim3(threads), 0, st, J, xm, zm,
                       tab + o.out_offset, outcomes, words, seed);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_sv_lds(const Job& J, uint32_t flags, hipStream_t st) {
  constexpr int T = (1 << (NS - 1)) < 256 ? (1 << (NS - 1)) : 256;
  hipLaunchKernelGGL((k_sv_forward_lds<NS, T>), dim3(J.total), dim3(T), 0, st, J, flags);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_sample_global(const Job& J, const int32_t* tab, const TableOffsets& o,
                                int32_t* outcomes, int32_t words, uint64_t seed, hipStream_t st) {
  const uint64_t* xm = reinterpret_cast<const uint64_t*>(tab + o.xmask);
  const uint64_t* zm = reinterpret_cast<const uint64_t*>(tab + o.zmask);
  return run_global_chunks(J, sizeof(float2) << NS, st, [&](float2* ws, int64_t g0, int64_t n) {
    hipLaunchKernelGGL((k_sv_sample_global<NS>), dim3(unsigned(n)), dim3(256), 0, st, J, xm, zm,
                       tab + o.out_offset, outcomes, words, seed, ws, g0);
  });
}

template <int NS>
hipError_t launch_sample_lds(const Job& J, const int32_t* tab, const TableOffsets& o,
                             int32_t* outcomes, int32_t words, uint64_t seed, hipStream_t st) {
  constexpr int T = (1 << (NS - 1)) < 256 ? (1 << (NS - 1)) : 256;
  const uint64_t* xm = reinterpret_cast<const uint64_t*>(tab + o.xmask);
  const uint64_t* zm = reinterpret_cast<const uint64_t*>(tab + o.zmask);
  hipLaunchKernelGGL((k_sv_sample_lds<NS, T>), dim3(unsigned(J.total)), dim3(T), 0, st, J, xm, zm,
                     tab + o.out_offset, outcomes, words, seed);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_dm_wave(const Job& J, const DmNoise& nz, uint32_t flags, bool periodic,
                          hipStream_t st) {
  constexpr int SPW = 64 >> (2 * NS);               // simulations per wave
  const int64_t waves = (J.total + SPW - 1) / SPW;
  const int64_t blocks = (waves + 3) / 4;           // 4 waves per workgroup
  if (periodic)
    hipLaunchKernelGGL((k_dm_forward_wave<NS, true>), dim3(unsigned(blocks)), dim3(256), 0, st, J,
                       nz, flags);
  else
    hipLaunchKernelGGL((k_dm_forward_wave<NS, false>), dim3(unsigned(blocks)), dim3(256), 0, st, J,
                       nz, flags);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_dm_lds(const Job& J, const DmNoise& nz, uint32_t flags, bool periodic,
                         hipStream_t st) {
  constexpr int Q = 1 << (2 * (NS - 1));          // quartets per step
  constexpr int T = Q < 256 ? Q : 256;            // threads per simulation
  constexpr int G = T < 64 ? 64 / T : 1;          // simulations per workgroup
  const int64_t blocks = (J.total + G - 1) / G;
  if (periodic)
    hipLaunchKernelGGL((k_dm_forward_lds<NS, T, G, true>), dim3(blocks), dim3(T * G), 0, st, J, nz,
                       flags);
  else
    hipLaunchKernelGGL((k_dm_forward_lds<NS, T, G, false>), dim3(blocks), dim3(T * G), 0, st, J, nz,
                       flags);
  return hipGetLastError();
}

int seg_setup(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
              int64_t angle_stride_a, const void* buf, int64_t B, int32_t n_segments,
              void* workspace, size_t workspace_bytes, Plan** pp, Job* J, SegInfo* G) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  int rc0 = check_adjoint_plan(p);
  if (rc0 != MBQC_OK) return rc0;
  if (p->ns > kMaxAdjSlots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "the segmented adjoint supports live windows of at most " +
                                              std::to_string(kMaxAdjSlots) + " qubits");
  if (n_segments < 2 || n_segments > std::min(kMaxSegments, p->n_blocks()))
    return fail(MBQC_ERR_VALUE, "n_segments must be in 2.." +
                                    std::to_string(std::min(kMaxSegments, p->n_blocks())));
  const size_t need = seg_layout(*p, B, n_segments, G);
  if (!workspace || workspace_bytes < need) return fail(MBQC_ERR_VALUE, "segmented workspace too small");
  // the input pointer is unused (columns are basis states); pass buf as a valid pointer
  int rc = make_job(p, angles, angle_stride, angle_stride_a, buf, 0, const_cast<void*>(buf), B, 1, J);
  if (rc != MBQC_OK) return rc;
  *pp = p;
  return MBQC_OK;
}

// Validated noise parameters, composed for the kernels.
int dm_noise(const mbqc_noise* noise, DmNoise* nz) {
  const mbqc_noise zero{};
  const mbqc_noise& n = noise ? *noise : zero;
  const float vals[7] = {n.p_depol, n.p_readout, n.amp_damping, n.phase_damping,
                         n.phase_flip, n.gad_gamma, n.gad_p};
  for (float v : vals)
    if (!(v >= 0.f && v <= 1.f)) return fail(MBQC_ERR_VALUE, "noise parameters must lie in [0, 1]");
  *nz = compose_noise(n);
  return MBQC_OK;
}

int dm_forward(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
               int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
               float* outcomes, int64_t B, int64_t S, const mbqc_noise* noise, uint32_t flags,
               void* stream) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  if (p->ns > kMaxDmGlobalSlots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "density-matrix kernels support live windows of at most " +
                                              std::to_string(kMaxDmGlobalSlots) + " qubits");
  DmNoise nz;
  int rc = dm_noise(noise, &nz);
  if (rc != MBQC_OK) return rc;
  Job J;
  rc = make_job(p, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, B, S, &J);
  if (rc != MBQC_OK) return rc;
  if (J.total == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  if (p->ns > kMaxDmLdsSlots) {
    // density matrices beyond the LDS: the general kernel on a global workspace
    e = launch_dm_global(J, nz, flags, outcomes, st);
    return e == hipSuccess ? MBQC_OK : global_fail(e, "dm_forward launch (global workspace)");
  }
  // other planes, per-step probabilities or a measurement order off the
  // slot rotation: the kernel that reads each step's slot and plane
  if (p->general_planes || !p->rotating || (flags & MBQC_FLAG_STEP_PROBS)) {
    if (J.total > 0x7fffffff) return fail(MBQC_ERR_VALUE, "too many simulations for one launch");
    WindowMap none{};
    const size_t lds = sizeof(float2) << (2 * p->ns);
    if (int rl = allow_lds(reinterpret_cast<const void*>(&k_dm_general), lds); rl != MBQC_OK)
      return rl;
    hipLaunchKernelGGL(k_dm_general, dim3(unsigned(J.total)), dim3(64), lds, st, J,
                       static_cast<const float*>(nullptr), int64_t(0), int32_t(p->n_steps), none, 0,
                       nz, flags, outcomes);
    e = hipGetLastError();
    return e == hipSuccess ? MBQC_OK : hip_fail(e, "dm_forward launch");
  }
  // Small batches of small windows leave the one-simulation-per-lane kernel
  // with a few waves on a few CUs; there the one-entry-per-lane kernel (up to
  // kDmWaveKernelMaxSims[ns] simulations) or the LDS kernel spread the work
  // over the whole chip.
  const bool small = J.total < kDmLaneKernelMinSims;
  const bool wave = p->ns <= 3 && J.total <= kDmWaveKernelMaxSims[p->ns];
  const bool pe = p->periodic && !(flags & MBQC_FLAG_TABLE_STEPS);
  switch (p->ns) {
    case 1: e = wave ? launch_dm_wave<1>(J, nz, flags, pe, st) : launch_dm<1>(J, nz, flags, pe, st); break;
    case 2: e = wave ? launch_dm_wave<2>(J, nz, flags, pe, st) : launch_dm<2>(J, nz, flags, pe, st); break;
    case 3:
      e = wave ? launch_dm_wave<3>(J, nz, flags, pe, st)
               : (small ? launch_dm_lds<3>(J, nz, flags, pe, st) : launch_dm<3>(J, nz, flags, pe, st));
      break;
    case 4: e = launch_dm_lds<4>(J, nz, flags, pe, st); break;
    case 5: e = launch_dm_lds<5>(J, nz, flags, pe, st); break;
    case 6: e = launch_dm_lds<6>(J, nz, flags, pe, st); break;
    case 7: e = launch_dm_lds<7>(J, nz, flags, pe, st); break;
  }
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "dm_forward launch");
}

// Loss partials of the Gram loss: one per (parameter set, tile of indices n).
int64_t gram_partials(int64_t B, int32_t K) {
  return B * int64_t(K % 16 == 0 ? K / 16 : (K + 63) / 64);
}

int gram_loss_seed(const void* t, const void* MT, int64_t B, int32_t K, int32_t S, float* loss,
                   void* seed, float* workspace, void* stream) {
  if (!t || !MT || !loss || !seed) return fail(MBQC_ERR_VALUE, "null pointer");
  if (B < 0 || K <= 0 || S <= 0) return fail(MBQC_ERR_VALUE, "sizes must be B >= 0, K >= 1, S >= 1");
  if (int64_t(kGramSets) * K * int64_t(sizeof(float2)) > 64 * 1024)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "Gram loss supports K <= 512");
  if (B == 0) return MBQC_OK;
  if ((B + 15) / 16 > 65535) return fail(MBQC_ERR_VALUE, "too many parameter sets");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float inv_S = 1.0f / float(S);
  int nblk;
  if (K % 16 == 0) {
    nblk = K / 16;
    hipLaunchKernelGGL(k_gram_seed_mfma, dim3(nblk, unsigned((B + 15) / 16)), dim3(64), 0, st,
                       static_cast<const float2*>(t), static_cast<const float2*>(MT), B, K, inv_S,
                       static_cast<float2*>(seed), workspace);
  } else {
    nblk = (K + 63) / 64;
    const size_t lds = size_t(kGramSets) * K * sizeof(float2);
    hipLaunchKernelGGL(k_gram_seed, dim3(nblk, unsigned((B + kGramSets - 1) / kGramSets)),
                       dim3(256), lds, st, static_cast<const float2*>(t),
                       static_cast<const float2*>(MT), B, K, inv_S, static_cast<float2*>(seed),
                       workspace);
  }
  hipLaunchKernelGGL(k_gram_loss, dim3(unsigned((B + 255) / 256)), dim3(256), 0, st, workspace, B,
                     nblk, inv_S, loss);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "gram loss launch");
}

// Address of this device's copy of `scratch`, looked up once per device with
// that device current (the array's name is only a host-side stand-in).
int scratch_ptr(float** out) {
  static std::mutex mu;
  static float* ptr[kMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(MBQC_ERR_RUNTIME, "device id out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    ptr[dev] = static_cast<float*>(p);
  }
  *out = ptr[dev];
  return MBQC_OK;
}

// Groups above kMaxSvSlots: one persistent workgroup of T threads (T / 64
// waves) per resident slot, `waves` / (T / 64) of them, at least one and at
// most one per item.
template <int NS>
void launch_batch_lds_group(int64_t waves, const BatchGroupLaunch& L) {
  constexpr int T = (1 << (NS - 1)) < 256 ? (1 << (NS - 1)) : 256;
  const int64_t wgs = std::max<int64_t>(1, std::min<int64_t>(waves / (T / 64), L.n_items));
  hipLaunchKernelGGL((k_batch_lds<NS, T>), dim3(unsigned(wgs)), dim3(T), 0, L.stream, L.jobs,
                     L.items, L.n_items, L.counter, L.flags);
}

// Groups above the LDS: `waves` / 4 persistent 256-thread workgroups (at least
// one, at most one per item, at most kGlobalWorkspaceBytes of states), each
// with its slice of a workspace allocated and freed in stream order.
template <int NS>
hipError_t launch_batch_global_group(int64_t waves, const BatchGroupLaunch& L, void* ws,
                                     int64_t ws_wgs) {
  // the batch's own workspace (allocated with the batch, so that a launch
  // allocates nothing and can be captured in a HIP graph): one state per
  // resident workgroup
  const int64_t wgs =
      std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(waves / 4, L.n_items), ws_wgs));
  hipLaunchKernelGGL((k_batch_global<NS>), dim3(unsigned(wgs)), dim3(256), 0, L.stream, L.jobs,
                     L.items, L.n_items, L.counter, L.flags, static_cast<float2*>(ws));
  return hipGetLastError();
}

// Resident workgroups of a batch group of 15..20 slots (one per CU); the batch
// allocates its workspace for this many states per group at creation.
constexpr int64_t kBatchGlobalWgs = 256;

}  // namespace

struct mbqc_batch {
  int max_ns = 1;
  int64_t group_begin[kMaxBatchSlots + 1] = {};  // items of live-window size ns
  int64_t group_count[kMaxBatchSlots + 1] = {};
  std::vector<Job> jobs;
  std::vector<BatchItem> items;
  Job* d_jobs = nullptr;
  BatchItem* d_items = nullptr;
  int* d_counter = nullptr;
  int device = -1;
  // side streams and events that run the window-size groups concurrently
  // (fork from and join back into the caller's stream, so the pattern is
  // also valid under graph capture)
  hipStream_t side[kMaxBatchSlots + 1] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t join[kMaxBatchSlots + 1] = {};
  // every job of the group has S a multiple of share_group(ns)
  bool share[kMaxBatchSlots + 1] = {};
  // every plan of the group is periodic (Plan::periodic)
  bool periodic[kMaxBatchSlots + 1] = {};
  // groups of 15..20 slots: the states of up to gws_wgs[ns] resident workgroups
  void* gws[kMaxBatchSlots + 1] = {};
  int64_t gws_wgs[kMaxBatchSlots + 1] = {};
};

// One launch of a batch: the counter reset, then one persistent kernel per
// live-window size ns with group_waves[ns] > 0 resident wavefronts (capped at
// the group's item count); with several groups each runs on its own stream,
// forked from and joined into `stream`.  Groups given no waves are skipped
// (their outputs are not written).
static int batch_launch(mbqc_batch* bt, const int32_t* group_waves, int32_t wg_waves,
                        uint32_t flags, void* stream) {
  if (wg_waves < 1 || wg_waves > 4) return fail(MBQC_ERR_VALUE, "wg_waves must be 1..4");
  int launched = 0;
  for (int ns = 1; ns <= kMaxBatchSlots; ++ns) {
    if (group_waves[ns] == 0) continue;
    if (group_waves[ns] < wg_waves) return fail(MBQC_ERR_VALUE, "waves must be at least wg_waves");
    ++launched;
  }
  if (bt->items.empty() || launched == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, bt->d_counter, kMaxBatchSlots + 1);
  const int threads = 64 * wg_waves;
  const bool ref = flags & MBQC_FLAG_REF_PHASE;
  const bool concurrent = bt->fork != nullptr && launched > 1;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && concurrent) e = hipEventRecord(bt->fork, st);
  // largest windows first
  for (int ns = kMaxBatchSlots; ns >= 1 && e == hipSuccess; --ns) {
    const int64_t n = bt->group_count[ns];
    if (n == 0 || group_waves[ns] == 0) continue;
    hipStream_t gs = st;
    if (concurrent) {
      gs = bt->side[ns];
      e = hipStreamWaitEvent(gs, bt->fork, 0);
      if (e != hipSuccess) break;
    }
    // no more resident waves than the group has items
    const int64_t wv =
        std::min<int64_t>(group_waves[ns], ((n + wg_waves - 1) / wg_waves) * wg_waves);
    const int blocks = int(wv / wg_waves);
    const BatchItem* items = bt->d_items + bt->group_begin[ns];
    int* counter = bt->d_counter + ns * kCounterStride;
    const BatchGroupLaunch L{dim3(blocks), dim3(threads), gs, bt->d_jobs, items, n, counter, flags};
    const bool sh = bt->share[ns];
    const bool pe = bt->periodic[ns] && !(flags & MBQC_FLAG_TABLE_STEPS);
    switch (ns) {
      case 1: launch_batch_group<1>(L, ref, false, pe); break;
      case 2: launch_batch_group<2>(L, ref, sh, pe); break;
      case 3: launch_batch_group<3>(L, ref, sh, pe); break;
      case 4: launch_batch_group<4>(L, ref, sh, pe); break;
      case 5: launch_batch_group<5>(L, ref, sh, pe); break;
      case 6: launch_batch_group<6>(L, ref, sh, pe); break;
      case 7: launch_batch_lds_group<7>(group_waves[ns], L); break;
      case 8: launch_batch_lds_group<8>(group_waves[ns], L); break;
      case 9: launch_batch_lds_group<9>(group_waves[ns], L); break;
      case 10: launch_batch_lds_group<10>(group_waves[ns], L); break;
      case 11: launch_batch_lds_group<11>(group_waves[ns], L); break;
      case 12: launch_batch_lds_group<12>(group_waves[ns], L); break;
      case 13: launch_batch_lds_group<13>(group_waves[ns], L); break;
      case 14: launch_batch_lds_group<14>(group_waves[ns], L); break;
      case 15: e = launch_batch_global_group<15>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
      case 16: e = launch_batch_global_group<16>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
      case 17: e = launch_batch_global_group<17>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
      case 18: e = launch_batch_global_group<18>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
      case 19: e = launch_batch_global_group<19>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
      default: e = launch_batch_global_group<20>(group_waves[ns], L, bt->gws[ns], bt->gws_wgs[ns]); break;
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && concurrent) {
      e = hipEventRecord(bt->join[ns], gs);
      if (e == hipSuccess) e = hipStreamWaitEvent(st, bt->join[ns], 0);
    }
  }
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "batch launch");
}

extern "C" {

int mbqc_plan_upload(const mbqc_plan* plan) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  const int32_t* tab = nullptr;
  return device_tables(const_cast<Plan*>(reinterpret_cast<const Plan*>(plan)), &tab);
}

void mbqc_plan_destroy(mbqc_plan* plan) {
  if (!plan) return;
  Plan* p = reinterpret_cast<Plan*>(plan);
  for (int d = 0; d < kMaxDevices; ++d)
    if (p->dev_tab[d]) (void)hipFree(p->dev_tab[d]);
  delete p;
}

int mbqc_sv_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                        int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                        void* phase_out, int64_t B, int64_t S, uint32_t flags, void* stream) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  if (p->ns > kMaxSvGlobalSlots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "statevector kernels support live windows of at most " +
                                              std::to_string(kMaxSvGlobalSlots) + " qubits");
  int rc0 = check_xy_plan(p);
  if (rc0 != MBQC_OK) return rc0;
  Job J;
  int rc = make_job(p, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, B, S, &J);
  if (rc != MBQC_OK) return rc;
  if ((flags & MBQC_FLAG_WRITE_PHASE) && !phase_out && B > 0)
    return fail(MBQC_ERR_VALUE, "phase_out is null");
  J.phase_out = static_cast<float2*>(phase_out);
  if (J.total == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  const bool periodic = p->periodic && !(flags & MBQC_FLAG_TABLE_STEPS);
  if (p->ns > kMaxSvLdsSlots) {
    // statevectors beyond the LDS: the LDS kernel's simulation on a global workspace
    switch (p->ns) {
      case 15: e = launch_sv_global<15>(J, flags, st); break;
      case 16: e = launch_sv_global<16>(J, flags, st); break;
      case 17: e = launch_sv_global<17>(J, flags, st); break;
      case 18: e = launch_sv_global<18>(J, flags, st); break;
      case 19: e = launch_sv_global<19>(J, flags, st); break;
      case 20: e = launch_sv_global<20>(J, flags, st); break;
    }
    return e == hipSuccess ? MBQC_OK : global_fail(e, "sv_forward launch (global workspace)");
  }
  if (!p->rotating) {
    // measured slots off the rotation: the LDS kernel reads them per step
    switch (p->ns) {
      case 1: e = launch_sv_lds<1>(J, flags, st); break;
      case 2: e = launch_sv_lds<2>(J, flags, st); break;
      case 3: e = launch_sv_lds<3>(J, flags, st); break;
      case 4: e = launch_sv_lds<4>(J, flags, st); break;
      case 5: e = launch_sv_lds<5>(J, flags, st); break;
      case 6: e = launch_sv_lds<6>(J, flags, st); break;
      case 7: e = launch_sv_lds<7>(J, flags, st); break;
      case 8: e = launch_sv_lds<8>(J, flags, st); break;
      case 9: e = launch_sv_lds<9>(J, flags, st); break;
      case 10: e = launch_sv_lds<10>(J, flags, st); break;
      case 11: e = launch_sv_lds<11>(J, flags, st); break;
      case 12: e = launch_sv_lds<12>(J, flags, st); break;
      case 13: e = launch_sv_lds<13>(J, flags, st); break;
      case 14: e = launch_sv_lds<14>(J, flags, st); break;
    }
    return e == hipSuccess ? MBQC_OK : hip_fail(e, "sv_forward launch");
  }
  switch (p->ns) {
    case 1: e = launch_sv<1>(J, p->coef_per_step, periodic, flags, st); break;
    case 2: e = launch_sv<2>(J, p->coef_per_step, periodic, flags, st); break;
    case 3: e = launch_sv<3>(J, p->coef_per_step, periodic, flags, st); break;
    case 4: e = launch_sv<4>(J, p->coef_per_step, periodic, flags, st); break;
    case 5: e = launch_sv<5>(J, p->coef_per_step, periodic, flags, st); break;
    case 6: e = launch_sv<6>(J, p->coef_per_step, periodic, flags, st); break;
    case 7: e = launch_sv_lds<7>(J, flags, st); break;
    case 8: e = launch_sv_lds<8>(J, flags, st); break;
    case 9: e = launch_sv_lds<9>(J, flags, st); break;
    case 10: e = launch_sv_lds<10>(J, flags, st); break;
    case 11: e = launch_sv_lds<11>(J, flags, st); break;
    case 12: e = launch_sv_lds<12>(J, flags, st); break;
    case 13: e = launch_sv_lds<13>(J, flags, st); break;
    case 14: e = launch_sv_lds<14>(J, flags, st); break;
  }
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "sv_forward launch");
}

int mbqc_sv_sample_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                       int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                       int32_t* outcomes, int64_t B, int64_t S, uint64_t seed, void* stream) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  int rc0 = check_xy_plan(p);
  if (rc0 != MBQC_OK) return rc0;
  if (!p->can_sample)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "sampled outcomes unavailable: " + p->sample_reason);
  if (!outcomes && B > 0) return fail(MBQC_ERR_VALUE, "outcomes pointer is null");
  Job J;
  int rc = make_job(p, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, B, S, &J);
  if (rc != MBQC_OK) return rc;
  if (J.total == 0) return MBQC_OK;
  const int32_t* tab = J.tab;
  TableOffsets o = table_offsets(*p);
  const int32_t words = (p->n_steps + 31) / 32;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  switch (p->ns) {
    case 1: e = launch_sample<1>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 2: e = launch_sample<2>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 3: e = launch_sample<3>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 4: e = launch_sample<4>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 5: e = launch_sample<5>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 6: e = launch_sample<6>(J, tab, o, outcomes, words, seed, p->periodic, st); break;
    case 7: e = launch_sample_lds<7>(J, tab, o, outcomes, words, seed, st); break;
    case 8: e = launch_sample_lds<8>(J, tab, o, outcomes, words, seed, st); break;
    case 9: e = launch_sample_lds<9>(J, tab, o, outcomes, words, seed, st); break;
    case 10: e = launch_sample_lds<10>(J, tab, o, outcomes, words, seed, st); break;
    case 11: e = launch_sample_lds<11>(J, tab, o, outcomes, words, seed, st); break;
    case 12: e = launch_sample_lds<12>(J, tab, o, outcomes, words, seed, st); break;
    case 13: e = launch_sample_lds<13>(J, tab, o, outcomes, words, seed, st); break;
    case 14: e = launch_sample_lds<14>(J, tab, o, outcomes, words, seed, st); break;
    case 15: e = launch_sample_global<15>(J, tab, o, outcomes, words, seed, st); break;
    case 16: e = launch_sample_global<16>(J, tab, o, outcomes, words, seed, st); break;
    case 17: e = launch_sample_global<17>(J, tab, o, outcomes, words, seed, st); break;
    case 18: e = launch_sample_global<18>(J, tab, o, outcomes, words, seed, st); break;
    case 19: e = launch_sample_global<19>(J, tab, o, outcomes, words, seed, st); break;
    case 20: e = launch_sample_global<20>(J, tab, o, outcomes, words, seed, st); break;
    default: return fail(MBQC_ERR_NOT_IMPLEMENTED, "sampled outcomes need n_slots <= 20");
  }
  return e == hipSuccess ? MBQC_OK : global_fail(e, "sv_sample launch");
}

// Shared set-up of the window entry points.
static int window_job(const mbqc_plan* plan, const float* step_angles, int64_t step_stride,
                      const void* psi_in, int64_t in_stride, int32_t k, void* out, int64_t B,
                      int64_t S, int max_slots, Job* J, WindowMap* map, Plan** pp) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  if (p->ns > max_slots)
    return fail(MBQC_ERR_NOT_IMPLEMENTED, "window snapshots support live windows of at most " +
                                              std::to_string(max_slots) + " qubits");
  if (k > 0 && !step_angles) return fail(MBQC_ERR_VALUE, "step angles pointer is null");
  if (step_stride < 0) return fail(MBQC_ERR_VALUE, "negative stride");
  int rc = window_map(*p, k, map);
  if (rc != MBQC_OK) return rc;
  // with k == 0 no angle is read; the input pointer stands in for the
  // (unused) angle pointer that make_job requires
  rc = make_job(p, step_angles ? step_angles : reinterpret_cast<const float*>(psi_in), 0, 1, psi_in,
                in_stride, out, B, S, J);
  *pp = p;
  return rc;
}

int mbqc_sv_window_c64(const mbqc_plan* plan, const float* step_angles, int64_t step_stride,
                       const void* psi_in, int64_t in_stride, int32_t k, void* out, int64_t B,
                       int64_t S, uint32_t flags, void* stream) {
  Job J;
  WindowMap map;
  Plan* p = nullptr;
  int rc = window_job(plan, step_angles, step_stride, psi_in, in_stride, k, out, B, S,
                      kMaxSvGlobalSlots, &J, &map, &p);
  if (rc != MBQC_OK) return rc;
  rc = check_xy_plan(p);
  if (rc != MBQC_OK) return rc;
  if (J.total == 0) return MBQC_OK;
  hipStream_t gst = static_cast<hipStream_t>(stream);
  if (p->ns > kMaxSvLdsSlots) {
    hipError_t e = run_global_chunks(J, sizeof(float2) << p->ns, gst,
                                     [&](float2* ws, int64_t g0, int64_t n) {
      hipLaunchKernelGGL(k_sv_window_global, dim3(unsigned(n)), dim3(256), 0, gst, J, step_angles,
                         step_stride, k, map, flags, ws, g0);
    });
    return e == hipSuccess ? MBQC_OK : global_fail(e, "sv_window launch (global workspace)");
  }
  const size_t lds = sizeof(float2) << p->ns;
  if (int rl = allow_lds(reinterpret_cast<const void*>(&k_sv_window), lds); rl != MBQC_OK) return rl;
  hipLaunchKernelGGL(k_sv_window, dim3(J.total), dim3(64), lds, static_cast<hipStream_t>(stream), J,
                     step_angles, step_stride, k, map, flags);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "sv_window launch");
}

int mbqc_dm_window_c64(const mbqc_plan* plan, const float* step_angles, int64_t step_stride,
                       const void* psi_in, int64_t in_stride, int32_t k, void* out, int64_t B,
                       int64_t S, const mbqc_noise* noise, float* outcomes, void* stream) {
  Job J;
  WindowMap map;
  Plan* p = nullptr;
  int rc = window_job(plan, step_angles, step_stride, psi_in, in_stride, k, out, B, S,
                      kMaxDmGlobalSlots, &J, &map, &p);
  if (rc != MBQC_OK) return rc;
  DmNoise nz;
  rc = dm_noise(noise, &nz);
  if (rc != MBQC_OK) return rc;
  if (J.total == 0) return MBQC_OK;
  if (p->ns > kMaxDmLdsSlots) {
    hipStream_t gst = static_cast<hipStream_t>(stream);
    const bool general = p->general_planes || outcomes;
    hipError_t e = run_global_chunks(J, sizeof(float2) << (2 * p->ns), gst,
                                     [&](float2* ws, int64_t g0, int64_t n) {
      if (general)
        hipLaunchKernelGGL(k_dm_general_global, dim3(unsigned(n)), dim3(256), 0, gst, J,
                           step_angles, step_stride, k, map, 1, nz, 0u, outcomes, ws, g0);
      else
        hipLaunchKernelGGL(k_dm_window_global, dim3(unsigned(n)), dim3(256), 0, gst, J,
                           step_angles, step_stride, k, map, nz, ws, g0);
    });
    return e == hipSuccess ? MBQC_OK : global_fail(e, "dm_window launch (global workspace)");
  }
  if (J.total > 0x7fffffff) return fail(MBQC_ERR_VALUE, "too many simulations for one launch");
  const size_t lds = sizeof(float2) << (2 * p->ns);
  rc = allow_lds(p->general_planes || outcomes ? reinterpret_cast<const void*>(&k_dm_general)
                                                : reinterpret_cast<const void*>(&k_dm_window),
                 lds);
  if (rc != MBQC_OK) return rc;
  if (p->general_planes || outcomes)
    hipLaunchKernelGGL(k_dm_general, dim3(unsigned(J.total)), dim3(64), lds,
                       static_cast<hipStream_t>(stream), J, step_angles, step_stride, k, map, 1, nz,
                       0u, outcomes);
  else
    hipLaunchKernelGGL(k_dm_window, dim3(unsigned(J.total)), dim3(64), lds,
                       static_cast<hipStream_t>(stream), J, step_angles, step_stride, k, map, nz);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "dm_window launch");
}

int mbqc_dm_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                        int64_t angle_stride_a, const void* psi_in, int64_t in_stride, void* out,
                        int64_t B, int64_t S, const mbqc_noise* noise, uint32_t flags, void* stream) {
  return dm_forward(plan, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, nullptr, B,
                    S, noise, flags, stream);
}

int mbqc_dm_forward_outcomes_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                 int64_t angle_stride_a, const void* psi_in, int64_t in_stride,
                                 void* out, float* outcomes, int64_t B, int64_t S,
                                 const mbqc_noise* noise, uint32_t flags, void* stream) {
  if (!outcomes && B > 0) return fail(MBQC_ERR_VALUE, "outcomes pointer is null");
  return dm_forward(plan, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, outcomes, B,
                    S, noise, flags | MBQC_FLAG_STEP_PROBS, stream);
}

size_t mbqc_sv_adjoint_workspace_bytes(const mbqc_plan* plan, int64_t n_columns) {
  if (!plan || n_columns < 0) return 0;
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  const int ck = adj_ckpt_blocks(n_columns);
  return size_t((p->n_blocks() + ck - 1) / ck) * (size_t(1) << p->ns) * size_t(n_columns) * 8u;
}

int32_t mbqc_sv_segments_auto(const mbqc_plan* plan, int64_t B) {
  if (!plan || B <= 0) return 1;
  const Plan* p = reinterpret_cast<const Plan*>(plan);
  if (p->ns > kMaxAdjSlots) return 1;   // larger windows: the unsegmented LDS sweeps
  // Measured on the 4x16 and 4x20 grids (16 lanes per set): the best count
  // falls from 4 segments at 16,384 lanes to 1 at 131,072 and more, about as
  // 4 sqrt(16384 / lanes); more segments add chain work per set.
  const double lanes = double(B) * double(int64_t(1) << p->ns);
  int64_t n = std::llround(4.0 * std::sqrt(16384.0 / lanes));
  n = std::min<int64_t>(n, std::min(kMaxSegments, p->n_blocks()));
  return int32_t(std::max<int64_t>(n, 1));
}

size_t mbqc_sv_segmented_workspace_bytes(const mbqc_plan* plan, int64_t B, int32_t n_segments) {
  if (!plan) return 0;
  return seg_layout(*reinterpret_cast<const Plan*>(plan), B, n_segments, nullptr);
}


int mbqc_sv_segmented_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                  int64_t angle_stride_a, void* out, int64_t B, int32_t n_segments,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  Plan* p = nullptr;
  Job J;
  SegInfo G;
  int rc = seg_setup(plan, angles, angle_stride, angle_stride_a, out, B, n_segments, workspace,
                     workspace_bytes, &p, &J, &G);
  if (rc != MBQC_OK) return rc;
  if (B == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  float2* ws = static_cast<float2*>(workspace);
  float2* o = static_cast<float2*>(out);
  hipError_t e = hipSuccess;
  switch (p->ns) {
    case 1: e = launch_seg_forward<1>(J, G, ws, o, p->periodic, st); break;
    case 2: e = launch_seg_forward<2>(J, G, ws, o, p->periodic, st); break;
    case 3: e = launch_seg_forward<3>(J, G, ws, o, p->periodic, st); break;
    case 4: e = launch_seg_forward<4>(J, G, ws, o, p->periodic, st); break;
    case 5: e = launch_seg_forward<5>(J, G, ws, o, p->periodic, st); break;
  }
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "segmented forward launch");
}

int mbqc_sv_segmented_backward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                   int64_t angle_stride_a, const void* seed, int64_t B,
                                   int32_t n_segments, void* workspace, size_t workspace_bytes,
                                   float* grad, void* stream) {
  if (!seed || !grad) return fail(MBQC_ERR_VALUE, "null seed or gradient pointer");
  Plan* p = nullptr;
  Job J;
  SegInfo G;
  int rc = seg_setup(plan, angles, angle_stride, angle_stride_a, seed, B, n_segments, workspace,
                     workspace_bytes, &p, &J, &G);
  if (rc != MBQC_OK) return rc;
  if (B == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  float2* ws = static_cast<float2*>(workspace);
  const float2* sd = static_cast<const float2*>(seed);
  hipError_t e = hipSuccess;
  switch (p->ns) {
    case 1: e = launch_seg_backward<1>(J, G, ws, sd, p->n_angles, grad, p->periodic, st); break;
    case 2: e = launch_seg_backward<2>(J, G, ws, sd, p->n_angles, grad, p->periodic, st); break;
    case 3: e = launch_seg_backward<3>(J, G, ws, sd, p->n_angles, grad, p->periodic, st); break;
    case 4: e = launch_seg_backward<4>(J, G, ws, sd, p->n_angles, grad, p->periodic, st); break;
    case 5: e = launch_seg_backward<5>(J, G, ws, sd, p->n_angles, grad, p->periodic, st); break;
  }
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "segmented backward launch");
}

int32_t mbqc_sv_adjoint_partials(int64_t C) {
  if (C <= 0) return 1;
  if (C <= 64 && (C & (C - 1)) == 0) return 1;
  if (C % 64 == 0) return int32_t(C / 64);
  return int32_t(C);
}

int mbqc_sv_adjoint_forward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                int64_t angle_stride_a, const void* psi_in, int64_t in_stride,
                                void* out, int64_t B, int64_t C, void* ckpt, size_t ckpt_bytes,
                                void* stream) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  int rc0 = check_adjoint_plan(p);
  if (rc0 != MBQC_OK) return rc0;
  Job J;
  int rc = make_job(p, angles, angle_stride, angle_stride_a, psi_in, in_stride, out, B, C, &J);
  if (rc != MBQC_OK) return rc;
  if (!ckpt || ckpt_bytes < mbqc_sv_adjoint_workspace_bytes(plan, B * C))
    return fail(MBQC_ERR_VALUE, "checkpoint workspace too small");
  if (J.total == 0) return MBQC_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  float2* ck = static_cast<float2*>(ckpt);
  hipError_t e = hipSuccess;
  switch (p->ns) {
    case 1: e = launch_adj_fwd<1>(J, ck, adj_ckpt_blocks(B * C), p->periodic, st); break;
    case 2: e = launch_adj_fwd<2>(J, ck, adj_ckpt_blocks(B * C), p->periodic, st); break;
    case 3: e = launch_adj_fwd<3>(J, ck, adj_ckpt_blocks(B * C), p->periodic, st); break;
    case 4: e = launch_adj_fwd<4>(J, ck, adj_ckpt_blocks(B * C), p->periodic, st); break;
    case 5: e = launch_adj_fwd<5>(J, ck, adj_ckpt_blocks(B * C), p->periodic, st); break;
    case 6: e = launch_adj_fwd_lds<6>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 7: e = launch_adj_fwd_lds<7>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 8: e = launch_adj_fwd_lds<8>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 9: e = launch_adj_fwd_lds<9>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 10: e = launch_adj_fwd_lds<10>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 11: e = launch_adj_fwd_global<11>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 12: e = launch_adj_fwd_global<12>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 13: e = launch_adj_fwd_global<13>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 14: e = launch_adj_fwd_global<14>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 15: e = launch_adj_fwd_global<15>(J, ck, adj_ckpt_blocks(B * C), st); break;
    case 16: e = launch_adj_fwd_global<16>(J, ck, adj_ckpt_blocks(B * C), st); break;
  }
  return e == hipSuccess ? MBQC_OK : global_fail(e, "adjoint forward launch");
}

int mbqc_sv_adjoint_backward_c64(const mbqc_plan* plan, const float* angles, int64_t angle_stride,
                                 int64_t angle_stride_a, const void* seed, int64_t B, int64_t C,
                                 const void* ckpt, size_t ckpt_bytes, float* grad_partial,
                                 void* stream) {
  if (!plan) return fail(MBQC_ERR_VALUE, "null plan");
  Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(plan));
  int rc0 = check_adjoint_plan(p);
  if (rc0 != MBQC_OK) return rc0;
  if (!seed || !grad_partial) return fail(MBQC_ERR_VALUE, "null seed or gradient pointer");
  if (!ckpt || ckpt_bytes < mbqc_sv_adjoint_workspace_bytes(plan, B * C))
    return fail(MBQC_ERR_VALUE, "checkpoint workspace too small");
  Job J;
  // psi is unused by the backward sweep; pass the seed as a valid pointer
  int rc = make_job(p, angles, angle_stride, angle_stride_a, seed, 0, const_cast<void*>(seed), B, C,
                    &J);
  if (rc != MBQC_OK) return rc;
  if (J.total == 0) return MBQC_OK;
  AdjCtx A{};
  A.gpart = grad_partial;
  A.C = C;
  A.P = mbqc_sv_adjoint_partials(C);
  A.red = (C <= 64 && (C & (C - 1)) == 0) ? int32_t(C) : (C % 64 == 0 ? 64 : 1);
  A.n_angles = p->n_angles;
  A.ck = adj_ckpt_blocks(B * C);
  if (p->ns > kMaxAdjSlots &&
      (p->ns > kMaxAdjLdsSlots ? size_t(p->n_angles) * sizeof(float)
                               : adj_lds_bytes(p->ns, p->n_angles)) > kMaxDynLds)
    return fail(MBQC_ERR_NOT_IMPLEMENTED,
                "adjoint of a " + std::to_string(p->ns) + "-slot window with " +
                    std::to_string(p->n_angles) + " angles exceeds the workgroup's LDS");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float2* ck = static_cast<const float2*>(ckpt);
  const float2* sd = static_cast<const float2*>(seed);
  hipError_t e = hipSuccess;
  switch (p->ns) {
    case 1: e = launch_adj_bwd<1>(J, ck, sd, A, p->periodic, st); break;
    case 2: e = launch_adj_bwd<2>(J, ck, sd, A, p->periodic, st); break;
    case 3: e = launch_adj_bwd<3>(J, ck, sd, A, p->periodic, st); break;
    case 4: e = launch_adj_bwd<4>(J, ck, sd, A, p->periodic, st); break;
    case 5: e = launch_adj_bwd<5>(J, ck, sd, A, p->periodic, st); break;
    case 6: e = launch_adj_bwd_lds<6>(J, ck, sd, A, st); break;
    case 7: e = launch_adj_bwd_lds<7>(J, ck, sd, A, st); break;
    case 8: e = launch_adj_bwd_lds<8>(J, ck, sd, A, st); break;
    case 9: e = launch_adj_bwd_lds<9>(J, ck, sd, A, st); break;
    case 10: e = launch_adj_bwd_lds<10>(J, ck, sd, A, st); break;
    case 11: e = launch_adj_bwd_global<11>(J, ck, sd, A, st); break;
    case 12: e = launch_adj_bwd_global<12>(J, ck, sd, A, st); break;
    case 13: e = launch_adj_bwd_global<13>(J, ck, sd, A, st); break;
    case 14: e = launch_adj_bwd_global<14>(J, ck, sd, A, st); break;
    case 15: e = launch_adj_bwd_global<15>(J, ck, sd, A, st); break;
    case 16: e = launch_adj_bwd_global<16>(J, ck, sd, A, st); break;
  }
  return e == hipSuccess ? MBQC_OK : global_fail(e, "adjoint backward launch");
}

int mbqc_batch_create(const mbqc_batch_job* jobs, int32_t n_jobs, int32_t sims_per_item,
                      mbqc_batch** out) {
  if (!out) return fail(MBQC_ERR_VALUE, "out is null");
  *out = nullptr;
  if (!jobs || n_jobs <= 0) return fail(MBQC_ERR_VALUE, "no jobs");
  if (sims_per_item <= 0 || sims_per_item % 64 != 0)
    return fail(MBQC_ERR_VALUE, "sims_per_item must be a positive multiple of 64");
  auto* bt = new mbqc_batch();
  for (bool& sh : bt->share) sh = true;
  for (bool& pe : bt->periodic) pe = true;
  struct Key {
    int ns;
    int64_t cost;
    size_t index;
  };
  std::vector<Key> order;
  for (int i = 0; i < n_jobs; ++i) {
    const mbqc_batch_job& bj = jobs[i];
    if (!bj.plan) {
      delete bt;
      return fail(MBQC_ERR_VALUE, "job has a null plan");
    }
    Plan* p = const_cast<Plan*>(reinterpret_cast<const Plan*>(bj.plan));
    if (p->ns > kMaxBatchSlots) {
      delete bt;
      return fail(MBQC_ERR_NOT_IMPLEMENTED, "batched kernel supports live windows of at most " +
                                                std::to_string(kMaxBatchSlots) + " qubits");
    }
    if (int rc0 = check_xy_plan(p); rc0 != MBQC_OK) {
      delete bt;
      return rc0;
    }
    if (!p->rotating) {
      delete bt;
      return fail(MBQC_ERR_NOT_IMPLEMENTED,
                  "batched launches need the nodes measured in schedule order");
    }
    Job J;
    int rc = make_job(p, bj.angles, bj.angle_stride, bj.angle_stride_a, bj.psi_in, bj.in_stride,
                      bj.out, bj.B, bj.S, &J);
    if (rc != MBQC_OK) {
      delete bt;
      return rc;
    }
    bt->jobs.push_back(J);
    bt->max_ns = std::max(bt->max_ns, p->ns);
    if (p->ns < 2 || J.S % share_group(p->ns) != 0) bt->share[p->ns] = false;
    if (!p->periodic) bt->periodic[p->ns] = false;
    const int64_t cost = int64_t(p->n_steps) << p->ns;
    for (int64_t f = 0; f < J.total; f += sims_per_item) {
      BatchItem it{};
      it.job = i;
      it.first = f;
      it.count = int32_t(std::min<int64_t>(sims_per_item, J.total - f));
      order.push_back({p->ns, cost, bt->items.size()});
      bt->items.push_back(it);
    }
  }
  // grouped by live-window size; longest patterns first within a group
  // (stable within a job)
  std::stable_sort(order.begin(), order.end(), [](const Key& a, const Key& b) {
    return a.ns != b.ns ? a.ns > b.ns : a.cost > b.cost;
  });
  std::vector<BatchItem> sorted;
  sorted.reserve(order.size());
  for (size_t k = 0; k < order.size(); ++k) {
    const int ns = order[k].ns;
    if (bt->group_count[ns] == 0) bt->group_begin[ns] = int64_t(k);
    ++bt->group_count[ns];
    sorted.push_back(bt->items[order[k].index]);
  }
  bt->items.swap(sorted);
  hipError_t e = hipGetDevice(&bt->device);
  if (e == hipSuccess) e = hipMalloc(&bt->d_jobs, sizeof(Job) * bt->jobs.size());
  if (e == hipSuccess && !bt->items.empty())
    e = hipMalloc(&bt->d_items, sizeof(BatchItem) * bt->items.size());
  if (e == hipSuccess) e = hipMalloc(&bt->d_counter, sizeof(int) * kCounterStride * (kMaxBatchSlots + 1));
  for (int ns = kMaxSvLdsSlots + 1; ns <= kMaxBatchSlots && e == hipSuccess; ++ns) {
    if (bt->group_count[ns] == 0) continue;
    const size_t state_bytes = sizeof(float2) << ns;
    bt->gws_wgs[ns] = std::min<int64_t>(std::min<int64_t>(bt->group_count[ns], kBatchGlobalWgs),
                                        int64_t(kGlobalWorkspaceBytes / state_bytes));
    e = hipMalloc(&bt->gws[ns], size_t(bt->gws_wgs[ns]) * state_bytes);
  }
  int n_groups = 0;
  for (int ns = 1; ns <= kMaxBatchSlots; ++ns) n_groups += bt->group_count[ns] > 0;
  if (e == hipSuccess && n_groups > 1) {
    e = hipEventCreateWithFlags(&bt->fork, hipEventDisableTiming);
    for (int ns = 1; ns <= kMaxBatchSlots && e == hipSuccess; ++ns) {
      if (bt->group_count[ns] == 0) continue;
      e = hipStreamCreateWithFlags(&bt->side[ns], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&bt->join[ns], hipEventDisableTiming);
    }
  }
  if (e == hipSuccess)
    e = hipMemcpy(bt->d_jobs, bt->jobs.data(), sizeof(Job) * bt->jobs.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && !bt->items.empty())
    e = hipMemcpy(bt->d_items, bt->items.data(), sizeof(BatchItem) * bt->items.size(),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    mbqc_batch_destroy(bt);
    return hip_fail(e, "batch allocation");
  }
  *out = bt;
  return MBQC_OK;
}

void mbqc_batch_destroy(mbqc_batch* bt) {
  if (!bt) return;
  if (bt->d_jobs) (void)hipFree(bt->d_jobs);
  if (bt->d_items) (void)hipFree(bt->d_items);
  if (bt->d_counter) (void)hipFree(bt->d_counter);
  for (int ns = 0; ns <= kMaxBatchSlots; ++ns) {
    if (bt->gws[ns]) (void)hipFree(bt->gws[ns]);
    if (bt->side[ns]) (void)hipStreamDestroy(bt->side[ns]);
    if (bt->join[ns]) (void)hipEventDestroy(bt->join[ns]);
  }
  if (bt->fork) (void)hipEventDestroy(bt->fork);
  delete bt;
}

int64_t mbqc_batch_items(const mbqc_batch* bt) { return bt ? int64_t(bt->items.size()) : -1; }

int mbqc_gram_finish_c64(const void* t, const void* z, int64_t B, int32_t K, int32_t S, float* loss,
                         void* seed, void* stream) {
  if (!t || !z || !loss || !seed) return fail(MBQC_ERR_VALUE, "null pointer");
  if (B < 0 || K <= 0 || S <= 0) return fail(MBQC_ERR_VALUE, "sizes must be B >= 0, K >= 1, S >= 1");
  if (B == 0) return MBQC_OK;
  if (B > 0x7fffffff) return fail(MBQC_ERR_VALUE, "too many parameter sets");
  hipLaunchKernelGGL(k_gram_finish, dim3(unsigned(B)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(t), static_cast<const float2*>(z), K, 1.0f / float(S),
                     loss, static_cast<float2*>(seed));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "gram finish launch");
}

int mbqc_overlap_loss_seed_c64(const void* out, const void* targets, int64_t B, int32_t S,
                               int32_t O, float* loss, void* seed, void* stream) {
  if (B < 0 || S <= 0 || O <= 0 || (O & (O - 1)) != 0)
    return fail(MBQC_ERR_VALUE, "sizes must be B >= 0, S >= 1 and O a power of two");
  if (B == 0) return MBQC_OK;
  if (!out || !targets || !loss) return fail(MBQC_ERR_VALUE, "null pointer");
  if (B > 0x7fffffff) return fail(MBQC_ERR_VALUE, "too many parameter sets");
  const int G = O < 64 ? O : 64;
  hipLaunchKernelGGL(k_overlap_loss, dim3(unsigned(B)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const float2*>(out), static_cast<const float2*>(targets), S, O, G,
                     1.0f / float(S), loss, static_cast<float2*>(seed));
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "overlap loss launch");
}

size_t mbqc_gram_workspace_bytes(int64_t B, int32_t K) {
  if (B < 0 || K <= 0) return 0;
  return size_t(B) * size_t((K + 15) / 16) * sizeof(float);
}

int mbqc_gram_loss_seed_c64(const void* t, const void* MT, int64_t B, int32_t K, int32_t S,
                            float* loss, void* seed, float* workspace, void* stream) {
  if (!workspace) return fail(MBQC_ERR_VALUE, "null pointer");
  return gram_loss_seed(t, MT, B, K, S, loss, seed, workspace, stream);
}

int mbqc_gram_loss_seed_c64_noalloc(const void* t, const void* MT, int64_t B, int32_t K,
                                    int32_t S, float* loss, void* seed, void* stream) {
  if (B < 0 || K <= 0) return fail(MBQC_ERR_VALUE, "sizes must be B >= 0, K >= 1, S >= 1");
  if (gram_partials(B, K) > kScratchFloats)
    return fail(MBQC_ERR_VALUE, "B x " + std::to_string(gram_partials(1, K)) + " = " +
                                    std::to_string(gram_partials(B, K)) +
                                    " loss partials exceed the static scratch of " +
                                    std::to_string(kScratchFloats) + " floats");
  float* ws = nullptr;
  int rc = scratch_ptr(&ws);
  if (rc != MBQC_OK) return rc;
  return gram_loss_seed(t, MT, B, K, S, loss, seed, ws, stream);
}

int mbqc_adam_step_f32(float* x, const float* grad, float* m, float* v, int64_t n, float lr, float b1,
                       float b2, float eps, int32_t* step, void* stream) {
  if (!x || !grad || !m || !v || !step) return fail(MBQC_ERR_VALUE, "null pointer");
  if (n < 0) return fail(MBQC_ERR_VALUE, "negative size");
  if (n == 0) return MBQC_OK;
  const int threads = 256;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_adam, dim3((n + threads - 1) / threads), dim3(threads), 0, st, x, grad, m, v, n,
                     lr, b1, b2, eps, step);
  hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, st, step);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "adam launch");
}

int mbqc_sgd_step_f32(float* x, const float* grad, float* v, int64_t n, float lr, float momentum,
                      int32_t nesterov, void* stream) {
  if (!x || !grad || !v) return fail(MBQC_ERR_VALUE, "null pointer");
  if (n < 0) return fail(MBQC_ERR_VALUE, "negative size");
  if (n == 0) return MBQC_OK;
  const int threads = 256;
  hipLaunchKernelGGL(k_sgd, dim3((n + threads - 1) / threads), dim3(threads), 0,
                     static_cast<hipStream_t>(stream), x, grad, v, n, lr, momentum, nesterov);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MBQC_OK : hip_fail(e, "sgd launch");
}

int mbqc_batch_launch(mbqc_batch* bt, int32_t waves, int32_t wg_waves, uint32_t flags, void* stream) {
  if (!bt) return fail(MBQC_ERR_VALUE, "null batch");
  int32_t gw[kMaxBatchSlots + 1];
  for (int ns = 0; ns <= kMaxBatchSlots; ++ns) gw[ns] = bt->group_count[ns] > 0 ? waves : 0;
  return batch_launch(bt, gw, wg_waves, flags, stream);
}

int mbqc_batch_launch_groups(mbqc_batch* bt, const int32_t* group_waves, int32_t n_entries,
                             int32_t wg_waves, uint32_t flags, void* stream) {
  if (!bt) return fail(MBQC_ERR_VALUE, "null batch");
  if (!group_waves) return fail(MBQC_ERR_VALUE, "group_waves is null");
  if (n_entries < 1 || n_entries > kMaxBatchSlots + 1)
    return fail(MBQC_ERR_VALUE, "group_waves must have 1.." + std::to_string(kMaxBatchSlots + 1) +
                                    " entries");
  int32_t gw[kMaxBatchSlots + 1] = {};
  for (int ns = 0; ns < n_entries; ++ns) {
    if (group_waves[ns] < 0) return fail(MBQC_ERR_VALUE, "negative wave count");
    gw[ns] = bt->group_count[ns] > 0 ? group_waves[ns] : 0;
  }
  return batch_launch(bt, gw, wg_waves, flags, stream);
}

int64_t mbqc_batch_group_items(const mbqc_batch* bt, int32_t n_slots) {
  if (!bt || n_slots < 0 || n_slots > kMaxBatchSlots) return -1;
  return bt->group_count[n_slots];
}

}  // extern "C"
