// Compile-time specialised Stockham FFTs for the pencil and row kernels.
//
// The length N, the radix plan (greedy 16, 8, 4, 2, then 3, 5, 7), the number
// of transforms per workgroup CNT and the thread count NT are template
// parameters, so every index, stride and loop bound is a constant.  A
// transform is a chain of stages; the first stage reads through a caller
// functor (typically straight from HBM) and the last stage writes through a
// caller functor (typically straight to HBM, or a spectral multiply into
// LDS), with LDS buffers only between stages.  Stage s (radix R, stride NS =
// product of the earlier radices), butterfly i in [0, N/R):
//   v[r] = x[i + r N/R] * W_N^(sign r k N/(NS R)),  k = i % NS
//   y[(i/NS) NS R + k + r NS] = DFT_R(v)[r]
#pragma once
#include <type_traits>

#include "kw_fft.hpp"

namespace kwt {

// Radix plans.  Plan 0: greedy 16, 8, 4, 2, then 3, 5, 7, with a 16 that
// would be the last stage of a longer transform split 4 x 4 (256 = 16 x 4 x 4:
// the last stage's radix sets the registers each thread holds between a C2R
// and its R2C, and a radix-16 last stage costs 16 of them).  Plan 1: the same
// with radix 8 at most (more stages, fewer registers per butterfly).
constexpr int next_radix(int m, int pl = 0) {
  if (pl == 1)
    return (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5
         : (m % 7 == 0) ? 7 : 0;
  return (m == 16) ? 4 : (m % 16 == 0) ? 16 : (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2
       : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : (m % 7 == 0) ? 7 : 0;
}
constexpr int num_stages(int n, int pl = 0) {
  int s = 0;
  while (n > 1) {
    const int r = next_radix(n, pl);
    if (r == 0) return -1;
    n /= r;
    ++s;
  }
  return s;
}
constexpr int radix_of(int n, int s, int pl = 0) {
  for (int i = 0; i < s; ++i) n /= next_radix(n, pl);
  return next_radix(n, pl);
}
constexpr int ns_of(int n, int s, int pl = 0) {
  int p = 1;
  for (int i = 0; i < s; ++i) {
    const int r = next_radix(n, pl);
    p *= r;
    n /= r;
  }
  return p;
}

// The same radices in reverse order (a valid Stockham plan too).  Stage 0 of
// the reversed plan reads, per work item, exactly the positions stage S-1 of
// the forward plan writes (both have N / R_{S-1} butterflies of stride
// N / R_{S-1}), so a forward transform's outputs can feed an inverse
// transform from registers.
constexpr int radix_rev(int n, int s, int pl = 0) { return radix_of(n, num_stages(n, pl) - 1 - s, pl); }
constexpr int ns_rev(int n, int s, int pl = 0) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= radix_rev(n, i, pl);
  return p;
}
template <bool REV>
constexpr int plan_radix(int n, int s, int pl = 0) { return REV ? radix_rev(n, s, pl) : radix_of(n, s, pl); }
template <bool REV>
constexpr int plan_ns(int n, int s, int pl = 0) { return REV ? ns_rev(n, s, pl) : ns_of(n, s, pl); }

// Loaders take (t, a) or (t, a, slot), slot = it * R + r as for stores.
template <typename Load>
__device__ __forceinline__ float2 call_load(Load& ld, int t, int a, int slot) {
  if constexpr (std::is_invocable_v<Load&, int, int, int>) {
    return ld(t, a, slot);
  } else {
    (void)slot;
    return ld(t, a);
  }
}

// radix-16 as 4 x 4 with internal twiddles W16^(n2 k1)
__device__ __forceinline__ void bf16(float2* v, float sign) {
  const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;  // cos/sin(pi/8)
  const float r2 = 0.70710678118654752440f;
  float2 y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 q[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
    bf4(q, sign);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = q[k1];
  }
  // W16^m, m = n2*k1 in {1,2,3,2,4,6,3,6,9}
  y[1][1] = cmul(y[1][1], make_float2(c1, sign * s1));
  y[1][2] = cmul(y[1][2], make_float2(r2, sign * r2));
  y[1][3] = cmul(y[1][3], make_float2(s1, sign * c1));
  y[2][1] = cmul(y[2][1], make_float2(r2, sign * r2));
  y[2][2] = cmul_si(y[2][2], sign);
  y[2][3] = cmul(y[2][3], make_float2(-r2, sign * r2));
  y[3][1] = cmul(y[3][1], make_float2(s1, sign * c1));
  y[3][2] = cmul(y[3][2], make_float2(-r2, sign * r2));
  y[3][3] = cmul(y[3][3], make_float2(-c1, -sign * s1));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 q[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    bf4(q, sign);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = q[k2];
  }
}

template <int R>
__device__ __forceinline__ void bfly(float2* v, float sign) {
  if constexpr (R == 2) bf2(v);
  else if constexpr (R == 3) bf3(v, sign);
  else if constexpr (R == 4) bf4(v, sign);
  else if constexpr (R == 5) bf5(v, sign);
  else if constexpr (R == 7) bf7(v, sign);
  else if constexpr (R == 8) bf8(v, sign);
  else bf16(v, sign);
}

// LDS layouts with a pad every 16 positions (spreads the stride-R Stockham
// writes over banks).
template <int N>
struct Pad {
  static constexpr int P = N + N / 16;
  __device__ static constexpr int at(int a) { return a + (a >> 4); }
};

// LDS accessors map (transform t, position a) to an element.  ptr(t, a) and
// off(c) let a stage address R butterfly inputs or outputs from one base with
// constant offsets (the ds_read / ds_write immediate) instead of one computed
// address each: with the position padded every 16 (PADA), at(x + c) =
// at(x) + at(c) holds for every x when c is a multiple of 16, and for x a
// multiple of 16 whatever c.
// pencils: transform index t fastest (coalesced columns), position a slowest
template <int N, int CNT>
struct PencilLds {
  float2* p;
  __device__ float2& operator()(int t, int a) const { return p[Pad<N>::at(a) * CNT + t]; }
  __device__ float2* ptr(int t, int a) const { return p + Pad<N>::at(a) * CNT + t; }
  static constexpr int off(int c) { return (c + c / 16) * CNT; }
  static constexpr bool PADA = true;
  static constexpr int size = Pad<N>::P * CNT;
};
// rows: position a fastest
template <int N, int CNT>
struct RowLds {
  float2* p;
  __device__ float2& operator()(int t, int a) const { return p[t * Pad<N>::P + Pad<N>::at(a)]; }
  __device__ float2* ptr(int t, int a) const { return p + t * Pad<N>::P + Pad<N>::at(a); }
  static constexpr int off(int c) { return c + c / 16; }
  static constexpr bool PADA = true;
  static constexpr int size = Pad<N>::P * CNT;
};

// Reads and writes of an LDS accessor between two stages (the stages
// recognise them and use base + constant offsets where the padding allows).
template <typename Lds>
struct LdsRead {
  Lds L;
  __device__ float2 operator()(int t, int a) const { return L(t, a); }
};
template <typename Lds>
struct LdsWrite {
  Lds L;
  __device__ void operator()(int t, int a, float2 v, int) const { L(t, a) = v; }
};
template <typename F>
struct LdsOf {
  static constexpr int kind = 0;
  static constexpr bool offs(int, bool) { return false; }
};
template <typename Lds>
struct LdsOf<LdsRead<Lds>> {
  static constexpr int kind = 1;
  // positions x + r * step for every r from one base: step a multiple of 16
  static constexpr bool offs(int step, bool) { return !Lds::PADA || step % 16 == 0; }
};
template <typename Lds>
struct LdsOf<LdsWrite<Lds>> {
  static constexpr int kind = 2;
  // positions x + r * step, r < R: also exact when x is a multiple of 16
  static constexpr bool offs(int step, bool x16) { return !Lds::PADA || step % 16 == 0 || x16; }
};

// R butterfly values at positions a0 + r * step (loads) ...
template <int R, int STEP, bool X16, typename Load>
__device__ __forceinline__ void load_r(Load& ld, int t, int a0, int slot0, float2* v) {
  if constexpr (LdsOf<Load>::kind == 1 && LdsOf<Load>::offs(STEP, X16)) {
    const float2* b = ld.L.ptr(t, a0);
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = b[decltype(ld.L)::off(r * STEP)];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = call_load(ld, t, a0 + r * STEP, slot0 + r);
  }
}
// ... and stores
template <int R, int STEP, bool X16, typename Store>
__device__ __forceinline__ void store_r(Store& st, int t, int a0, int slot0, const float2* v) {
  if constexpr (LdsOf<Store>::kind == 2 && LdsOf<Store>::offs(STEP, X16)) {
    float2* b = st.L.ptr(t, a0);
#pragma unroll
    for (int r = 0; r < R; ++r) b[decltype(st.L)::off(r * STEP)] = v[r];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) st(t, a0 + r * STEP, v[r], slot0 + r);
  }
}

// One stage over `cnt` transforms (cnt <= CMAX; TFAST requires cnt == CMAX).
// TFAST: work item w -> (t = w % CMAX, i = w / CMAX), else (t = w / NB, i = w % NB).
// Store functors receive (t, a, value, slot) with slot = it * R + r, a
// compile-time constant after unrolling (lets callers keep outputs in registers).
template <int N, int S, int CMAX, int NT, bool TFAST, typename Load, typename Store>
__device__ __forceinline__ void stage(const float2* tw, float sign, int cnt, Load& ld, Store& st) {
  constexpr int R = radix_of(N, S);
  constexpr int NS = ns_of(N, S);
  constexpr int NB = N / R;
  constexpr int TMAX = NB * CMAX;
  constexpr int TWSTEP = N / (NS * R);
  constexpr int ITER = (TMAX + NT - 1) / NT;
  const int total = NB * cnt;
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (w >= total) break;
    int t, i;
    if constexpr (TFAST) {
      t = w % CMAX;
      i = w / CMAX;
    } else {
      t = w / NB;
      i = w % NB;
    }
    const int k = i % NS;
    float2 v[R];
    load_r<R, NB, false>(ld, t, i, it * R, v);
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 wv = tw[r * k * TWSTEP];
        wv.y *= -sign;
        v[r] = cmul(v[r], wv);
      }
    }
    bfly<R>(v, sign);
    const int j = (i / NS) * NS * R + k;
    store_r<R, NS, (NS == 1 && R % 16 == 0)>(st, t, j, it * R, v);
  }
}

// Full transform: stage 0 loads through `first`, the last stage stores through
// `last`, intermediate stages ping-pong between LDS accessors A and B (B is
// only touched for three or more stages).  Intermediate stages are separated
// by barriers; there is no trailing barrier.
template <int N, int CMAX, int NT, bool TFAST, int S = 0, typename Load, typename Store, typename Lds>
__device__ __forceinline__ void run(const float2* tw, float sign, int cnt, Load& first, Store& last, Lds A, Lds B) {
  constexpr int NSTAGES = num_stages(N);
  static_assert(NSTAGES > 0, "length must factor into 2, 3, 5, 7");
  if constexpr (NSTAGES == 1) {
    stage<N, 0, CMAX, NT, TFAST>(tw, sign, cnt, first, last);
  } else if constexpr (S == 0) {
    LdsWrite<Lds> st{A};
    stage<N, 0, CMAX, NT, TFAST>(tw, sign, cnt, first, st);
    __syncthreads();
    run<N, CMAX, NT, TFAST, 1>(tw, sign, cnt, first, last, A, B);
  } else if constexpr (S == NSTAGES - 1) {
    LdsRead<Lds> ld{A};
    stage<N, S, CMAX, NT, TFAST>(tw, sign, cnt, ld, last);
  } else {
    LdsRead<Lds> ld{A};
    LdsWrite<Lds> st{B};
    stage<N, S, CMAX, NT, TFAST>(tw, sign, cnt, ld, st);
    __syncthreads();
    run<N, CMAX, NT, TFAST, S + 1>(tw, sign, cnt, first, last, B, A);
  }
}

// One stage over all CMAX transforms that may write the buffer it reads: every
// work item's inputs are loaded into registers, then a barrier, then the
// butterflies and stores (a Stockham stage writes other positions than it
// reads, so without the barrier a store could overwrite an unread input).
// BAR false drops that barrier where no such hazard exists (see run_ip's flags).
template <int N, int S, int CMAX, int NT, bool TFAST, bool REV = false, int PL = 0, bool BAR = true, typename Load,
          typename Store>
__device__ __forceinline__ void stage_ip(const float2* tw, float sign, Load& ld, Store& st) {
  constexpr int R = plan_radix<REV>(N, S, PL);
  constexpr int NS = plan_ns<REV>(N, S, PL);
  constexpr int NB = N / R;
  constexpr int TMAX = NB * CMAX;
  constexpr int TWSTEP = N / (NS * R);
  constexpr int ITER = (TMAX + NT - 1) / NT;
  float2 v[ITER][R];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (TMAX % NT == 0 || w < TMAX) {
      const int t = TFAST ? w % CMAX : w / NB;
      const int i = TFAST ? w / CMAX : w % NB;
      load_r<R, NB, false>(ld, t, i, it * R, v[it]);
    }
  }
  if constexpr (BAR) __syncthreads();
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (TMAX % NT == 0 || w < TMAX) {
      const int t = TFAST ? w % CMAX : w / NB;
      const int i = TFAST ? w / CMAX : w % NB;
      const int k = i % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 wv = tw[r * k * TWSTEP];
          wv.y *= -sign;
          v[it][r] = cmul(v[it][r], wv);
        }
      }
      bfly<R>(v[it], sign);
    }
  }
  // stores after every butterfly: a global store between two iterations'
  // twiddle loads would make each wait for the other's round trip
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (TMAX % NT == 0 || w < TMAX) {
      const int t = TFAST ? w % CMAX : w / NB;
      const int i = TFAST ? w / CMAX : w % NB;
      const int k = i % NS;
      const int j = (i / NS) * NS * R + k;
      store_r<R, NS, (NS == 1 && R % 16 == 0)>(st, t, j, it * R, v[it]);
    }
  }
}

// Caller guarantees for run_ip (flags FL): kFree0 -- no thread reads or
// writes L when the transform starts, so the first stage may store into L
// without waiting; kNoLdsLast -- `last` does not write LDS, so the final stage
// needs no barrier between its LDS reads and its stores.  Without them every
// stage waits for all reads of the buffer before its stores overwrite it.
constexpr unsigned kFree0 = 1, kNoLdsLast = 2;

// Full transform of all CMAX transforms in one LDS buffer L (in place):
// stage 0 loads through `first`, the last stage stores through `last`, the
// others read and write L.  Stages are separated by barriers; there is no
// trailing barrier.
template <int N, int CMAX, int NT, bool TFAST, bool REV = false, int PL = 0, unsigned FL = 0, int S = 0,
          typename Load, typename Store, typename Lds>
__device__ __forceinline__ void run_ip(const float2* tw, float sign, Load& first, Store& last, Lds L) {
  constexpr int NSTAGES = num_stages(N, PL);
  static_assert(NSTAGES > 0, "length must factor into 2, 3, 5, 7");
  if constexpr (NSTAGES == 1) {
    stage_ip<N, 0, CMAX, NT, TFAST, REV, PL, !(FL & (kFree0 | kNoLdsLast))>(tw, sign, first, last);
  } else if constexpr (S == 0) {
    LdsWrite<Lds> st{L};
    stage_ip<N, 0, CMAX, NT, TFAST, REV, PL, !(FL & kFree0)>(tw, sign, first, st);
    __syncthreads();
    run_ip<N, CMAX, NT, TFAST, REV, PL, FL, 1>(tw, sign, first, last, L);
  } else if constexpr (S == NSTAGES - 1) {
    LdsRead<Lds> ld{L};
    stage_ip<N, S, CMAX, NT, TFAST, REV, PL, !(FL & kNoLdsLast)>(tw, sign, ld, last);
  } else {
    LdsRead<Lds> ld{L};
    LdsWrite<Lds> st{L};
    stage_ip<N, S, CMAX, NT, TFAST, REV, PL>(tw, sign, ld, st);
    __syncthreads();
    run_ip<N, CMAX, NT, TFAST, REV, PL, FL, S + 1>(tw, sign, first, last, L);
  }
}

// In-place stage and transform over a runtime number of transforms cnt <=
// CMAX, work items transform-major (t = w / NB), as stage()/run() distribute
// them: the same slot numbering, so last_pos() maps the last stage's slots.
template <int N, int S, int CMAX, int NT, bool BAR = true, typename Load, typename Store>
__device__ __forceinline__ void stage_ipc(const float2* tw, float sign, int cnt, Load& ld, Store& st) {
  constexpr int R = radix_of(N, S);
  constexpr int NS = ns_of(N, S);
  constexpr int NB = N / R;
  constexpr int TMAX = NB * CMAX;
  constexpr int TWSTEP = N / (NS * R);
  constexpr int ITER = (TMAX + NT - 1) / NT;
  const int total = NB * cnt;
  float2 v[ITER][R];
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (w < total) {
      const int t = w / NB, i = w % NB;
      load_r<R, NB, false>(ld, t, i, it * R, v[it]);
    }
  }
  if constexpr (BAR) __syncthreads();
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (w < total) {
      const int i = w % NB;
      const int k = i % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 wv = tw[r * k * TWSTEP];
          wv.y *= -sign;
          v[it][r] = cmul(v[it][r], wv);
        }
      }
      bfly<R>(v[it], sign);
    }
  }
#pragma unroll
  for (int it = 0; it < ITER; ++it) {
    const int w = it * NT + (int)threadIdx.x;
    if (w < total) {
      const int t = w / NB, i = w % NB;
      const int k = i % NS;
      const int j = (i / NS) * NS * R + k;
      store_r<R, NS, (NS == 1 && R % 16 == 0)>(st, t, j, it * R, v[it]);
    }
  }
}

template <int N, int CMAX, int NT, unsigned FL = 0, int S = 0, typename Load, typename Store, typename Lds>
__device__ __forceinline__ void run_ipc(const float2* tw, float sign, int cnt, Load& first, Store& last, Lds L) {
  constexpr int NSTAGES = num_stages(N);
  static_assert(NSTAGES > 0, "length must factor into 2, 3, 5, 7");
  if constexpr (NSTAGES == 1) {
    stage_ipc<N, 0, CMAX, NT, !(FL & (kFree0 | kNoLdsLast))>(tw, sign, cnt, first, last);
  } else if constexpr (S == 0) {
    LdsWrite<Lds> st{L};
    stage_ipc<N, 0, CMAX, NT, !(FL & kFree0)>(tw, sign, cnt, first, st);
    __syncthreads();
    run_ipc<N, CMAX, NT, FL, 1>(tw, sign, cnt, first, last, L);
  } else if constexpr (S == NSTAGES - 1) {
    LdsRead<Lds> ld{L};
    stage_ipc<N, S, CMAX, NT, !(FL & kNoLdsLast)>(tw, sign, cnt, ld, last);
  } else {
    LdsRead<Lds> ld{L};
    LdsWrite<Lds> st{L};
    stage_ipc<N, S, CMAX, NT>(tw, sign, cnt, ld, st);
    __syncthreads();
    run_ipc<N, CMAX, NT, FL, S + 1>(tw, sign, cnt, first, last, L);
  }
}

// Registers a thread holds between a forward run_ip (last stage stores by
// slot) and the reversed-plan inverse (first stage loads by slot).
template <int N, int CMAX, int NT, int PL = 0>
constexpr int handoff_slots() {
  constexpr int R = radix_of(N, num_stages(N, PL) - 1, PL);
  return ((N / R) * CMAX + NT - 1) / NT * R;
}

// Number of output slots a thread owns in the last stage, and the (t, a)
// position of a slot (false if the thread has no work item for it).
template <int N, int CMAX, int NT>
constexpr int last_slots() {
  constexpr int S = num_stages(N) - 1;
  constexpr int R = radix_of(N, S);
  return ((N / R) * CMAX + NT - 1) / NT * R;
}

template <int N, int CMAX, int NT, bool TFAST>
__device__ __forceinline__ bool last_pos(int cnt, int slot, int& t, int& a) {
  constexpr int S = num_stages(N) - 1;
  constexpr int R = radix_of(N, S);
  constexpr int NS = ns_of(N, S);
  constexpr int NB = N / R;
  const int it = slot / R, r = slot % R;
  const int w = it * NT + (int)threadIdx.x;
  if (w >= NB * cnt) return false;
  int i;
  if constexpr (TFAST) {
    t = w % CMAX;
    i = w / CMAX;
  } else {
    t = w / NB;
    i = w % NB;
  }
  const int k = i % NS;
  a = (i / NS) * NS * R + k + r * NS;
  return true;
}

// Block index remap so that consecutive logical tiles land on the same XCD
// (workgroups are dealt round-robin over the 8 XCDs, each with its own L2).
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
  const int per = nb / 8;
  if (b >= per * 8) return b;
  return (b % 8) * per + b / 8;
}

}  // namespace kwt
