// Length-specialised x-row kernels (gfx950): the elementwise physics of the
// time step fused between the x-axis C2R and R2C of whole rows, for Nx with a
// compiled instantiation (other Nx use the generic kernels of kw_kernels.hip).
//
// Two real rows share one complex transform (z = a + ib).  The C2R's first
// stage reads the half spectra straight from HBM (Hermitian extension on the
// fly); its last stage leaves each thread holding a fixed set of output
// positions ("slots") in registers, the same set for every transform of the
// block, so the three velocity-divergence components of a point meet in
// registers for the density/EOS update without an LDS round trip.  Updated
// real rows go to LDS once: sensors gather from there and the R2C reads them.
// Statement references are to k-Wave/kspaceFirstOrder3D.m.
#include "kw_fft_t.hpp"
#include "kw_kernels.hpp"

namespace kw {

template <int NX>
struct RowCfg {
  static constexpr int ST = kwt::num_stages(NX);
  static constexpr int PMAX = (2048 / NX) > 16 ? 16 : ((2048 / NX) < 1 ? 1 : 2048 / NX);
  static constexpr int NT = 256;
  static constexpr int SLOTS = kwt::last_slots<NX, PMAX, NT>();
  static constexpr int PADN = kwt::Pad<NX>::P;
  static constexpr int NXH = NX / 2 + 1;
  static constexpr int NWORK = ST >= 3 ? 2 : 1;  // FFT ping-pong buffers
};

// Row-major LDS accessor with a runtime number of rows (pairs).
template <int NX>
struct Rows {
  float2* p;
  __device__ float2& operator()(int t, int a) const { return p[t * kwt::Pad<NX>::P + kwt::Pad<NX>::at(a)]; }
};

template <int NX>
struct RowBufs {
  Rows<NX> A, B, U, U2;
  int* slot;
  __device__ RowBufs(float2* smem, int P, bool two_u) {
    using C = RowCfg<NX>;
    const int sz = P * C::PADN;
    A.p = smem;
    B.p = smem + (C::NWORK == 2 ? sz : 0);
    U.p = smem + C::NWORK * sz;
    U2.p = U.p + (two_u ? sz : 0);
    slot = reinterpret_cast<int*>(U2.p + sz);
  }
};

// C2R of the half spectra of rows (r0 + 2t, r0 + 2t + 1) from HBM; the last
// stage hands each output to `sink(t, a, value, slot)`.
template <int NX, typename Sink>
__device__ __forceinline__ void c2r_rows(const Grid& g, const float2* __restrict__ G, int r0, int P,
                                         const float2* tw, RowBufs<NX>& b, Sink& sink) {
  using C = RowCfg<NX>;
  auto ld = [&](int t, int a) {
    const int ra = r0 + 2 * t, rb = ra + 1;
    const bool mirror = 2 * a > NX;
    const int k = mirror ? NX - a : a;
    float2 A = ra < g.rows ? G[(long long)ra * C::NXH + k] : make_float2(0.f, 0.f);
    float2 B = rb < g.rows ? G[(long long)rb * C::NXH + k] : make_float2(0.f, 0.f);
    if (k == 0 || 2 * k == NX) {
      A.y = 0.f;
      B.y = 0.f;
    }
    if (mirror) {
      A.y = -A.y;
      B.y = -B.y;
    }
    return make_float2(A.x - B.y, A.y + B.x);
  };
  kwt::run<NX, C::PMAX, C::NT, false>(tw, 1.f, P, ld, sink, b.A, b.B);
}

// R2C of the packed real rows in `src` (natural order) to the half spectra of
// rows (r0 + 2t, r0 + 2t + 1) in HBM.  Ends with the stores (no barrier);
// `src` is overwritten with the packed spectrum.
template <int NX>
__device__ __forceinline__ void r2c_rows(const Grid& g, Rows<NX> src, float2* __restrict__ G, int r0, int P,
                                         const float2* tw, RowBufs<NX>& b) {
  using C = RowCfg<NX>;
  auto ld = [&](int t, int a) { return src(t, a); };
  auto st = [&](int t, int a, float2 v, int) { src(t, a) = v; };
  kwt::run<NX, C::PMAX, C::NT, false>(tw, -1.f, P, ld, st, b.A, b.B);
  __syncthreads();
  for (int e = threadIdx.x; e < P * C::NXH; e += C::NT) {
    const int p = e / C::NXH, k = e - p * C::NXH;
    const int ra = r0 + 2 * p, rb = ra + 1;
    const float2 Z = src(p, k);
    const float2 Zm = src(p, k == 0 ? 0 : NX - k);
    if (ra < g.rows) G[(long long)ra * C::NXH + k] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (rb < g.rows) G[(long long)rb * C::NXH + k] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

// Sensor gather (extractSensorData.m:61-140) from packed real rows in LDS.
template <int NX>
__device__ void record_rows(const Sens& S, Rows<NX> U, const Grid& g, int r0, int P) {
  if (!S.record || S.ns == 0) return;
  const int rend = min(r0 + 2 * P, g.rows);
  const int j0 = S.row_ptr[r0], j1 = S.row_ptr[rend];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / NX);
    const int x = (int)(lin - (long long)r * NX);
    const int lr = r - r0;
    const float2 w = U(lr >> 1, x);
    const float v = (lr & 1) ? w.y : w.x;
    if (S.raw) S.raw[S.t_rec * S.ns + j] = v;
    if (S.vmax) S.vmax[j] = S.first ? v : fmaxf(S.vmax[j], v);
    if (S.vmin) S.vmin[j] = S.first ? v : fminf(S.vmin[j], v);
    if (S.rms_sum) {
      double s = S.first ? 0.0 : S.rms_sum[2 * j];
      double c = S.first ? 0.0 : S.rms_sum[2 * j + 1];
      const double y = (double)v * (double)v - c;
      const double t = s + y;
      c = (t - s) - y;
      S.rms_sum[2 * j] = t;
      S.rms_sum[2 * j + 1] = c;
    }
  }
}

__device__ __forceinline__ void field_rec(const Sens& S, long long idx, float v) {
  if (S.record) {
    if (S.max_all) S.max_all[idx] = S.first ? v : fmaxf(S.max_all[idx], v);
    if (S.min_all) S.min_all[idx] = S.first ? v : fminf(S.min_all[idx], v);
  }
  if (S.final_field) S.final_field[idx] = v;
}

template <int NX>
__device__ void slot_map(const Src& S, int* slot, const Grid& g, int r0, int P) {
  for (int e = threadIdx.x; e < 2 * P * NX; e += blockDim.x) slot[e] = -1;
  __syncthreads();
  const int rend = min(r0 + 2 * P, g.rows);
  const int j0 = S.row_ptr[r0], j1 = S.row_ptr[rend];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / NX);
    slot[(r - r0) * NX + (int)(lin - (long long)r * NX)] = j;
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ T pick3(int i, T a, T b, T c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// ---------------------------------------------------------------------------
// velocity (3D.m:790-801), velocity sources (3D.m:804-884), u sensors, R2C.
// ---------------------------------------------------------------------------
struct VelArgsT {
  float2* G[3];
  float* u[3];
  Src src[3];
  Sens sens[3];
  int init;
};

template <int NX>
__global__ void __launch_bounds__(256) k_xvel_t(Grid g, const float2* __restrict__ tw, Medium m, VelArgsT A,
                                                int P) {
  using C = RowCfg<NX>;
  extern __shared__ float2 smem[];
  RowBufs<NX> b(smem, P, false);
  const int i = blockIdx.y;
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  const Src S = pick3(i, A.src[0], A.src[1], A.src[2]);
  const Sens SE = pick3(i, A.sens[0], A.sens[1], A.sens[2]);
  float2* Gi = pick3(i, A.G[0], A.G[1], A.G[2]);
  float* ui = pick3(i, A.u[0], A.u[1], A.u[2]);
  const SF rsg = pick3(i, m.rsg[0], m.rsg[1], m.rsg[2]);
  const float* pmlsg = pick3(i, m.pml_sg[0], m.pml_sg[1], m.pml_sg[2]);
  const bool src_on = !A.init && S.active;

  float2 sadd[C::SLOTS];
  if (src_on && S.mode == 2 && S.many) {
    auto sink = [&](int, int, float2 v, int s) { sadd[s] = v; };
    c2r_rows<NX>(g, S.spec, r0, P, tw, b, sink);
    __syncthreads();
  }
  if (src_on && S.mode != 2) slot_map<NX>(S, b.slot, g, r0, P);
  float2 grad[C::SLOTS];
  auto sink = [&](int, int, float2 v, int s) { grad[s] = v; };
  c2r_rows<NX>(g, Gi, r0, P, tw, b, sink);

#pragma unroll
  for (int s = 0; s < C::SLOTS; ++s) {
    int t, a;
    if (!kwt::last_pos<NX, C::PMAX, C::NT, false>(P, s, t, a)) continue;
    float un[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * t + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * NX + a;
      const float gv = h ? grad[s].y : grad[s].x;
      const float rs = rsg.at(idx);
      float v;
      if (A.init) {
        v = 0.5f * m.dt * rs * gv;
      } else {
        const float pf = i == 0 ? pmlsg[a] : (i == 1 ? pmlsg[r % g.Ny] : pmlsg[r / g.Ny]);
        v = pf * (pf * ui[idx] - m.dt * rs * gv);
        if (src_on) {
          if (S.mode == 2) {
            v += S.many ? (h ? sadd[s].y : sadd[s].x) : S.sval * S.pattern[idx];
          } else {
            const int j = b.slot[(2 * t + h) * NX + a];
            if (j >= 0) {
              const float sv = S.sig[S.many ? j : 0];
              v = (S.mode == 0) ? sv : v + sv;
            }
          }
        }
      }
      un[h] = v;
      ui[idx] = v;
      field_rec(SE, idx, v);
    }
    b.U(t, a) = make_float2(un[0], un[1]);
  }
  __syncthreads();
  record_rows<NX>(SE, b.U, g, r0, P);
  if (A.init) return;
  r2c_rows<NX>(g, b.U, Gi, r0, P, tw, b);
}

// ---------------------------------------------------------------------------
// density (3D.m:910-925), pressure source (3D.m:928-960), EOS (3D.m:963-1017).
// ---------------------------------------------------------------------------
struct DensArgsT {
  float2* G[3];
  float* rho[3];
  Src psrc;
  float* srho;
  float2* pk;
  Sens ps;
};

template <int NX>
__global__ void __launch_bounds__(256) k_xdens_t(Grid g, const float2* __restrict__ tw, Medium m, DensArgsT A,
                                                 int P) {
  using C = RowCfg<NX>;
  extern __shared__ float2 smem[];
  const bool absorb = m.absorbing == 1;
  RowBufs<NX> b(smem, P, absorb);
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  const Src& S = A.psrc;

  float2 sadd[C::SLOTS];
  if (S.active && S.mode == 2 && S.many) {
    auto sink = [&](int, int, float2 v, int s) { sadd[s] = v; };
    c2r_rows<NX>(g, S.spec, r0, P, tw, b, sink);
    __syncthreads();
  }
  if (S.active && S.mode != 2) slot_map<NX>(S, b.slot, g, r0, P);
  float2 du[3][C::SLOTS];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c < g.dim) {
      auto sink = [&](int, int, float2 v, int s) { du[c][s] = v; };
      c2r_rows<NX>(g, A.G[c], r0, P, tw, b, sink);
      __syncthreads();
    } else {
#pragma unroll
      for (int s = 0; s < C::SLOTS; ++s) du[c][s] = make_float2(0.f, 0.f);
    }
  }

#pragma unroll
  for (int s = 0; s < C::SLOTS; ++s) {
    int t, a;
    if (!kwt::last_pos<NX, C::PMAX, C::NT, false>(P, s, t, a)) continue;
    float outA[2] = {0.f, 0.f}, outC[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * t + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * NX + a;
      const float r0v = m.rho0.at(idx);
      float rold[3], dui[3];
      float srho_old = 0.f, sdu = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rold[c] = c < g.dim ? A.rho[c][idx] : 0.f;
        dui[c] = h ? du[c][s].y : du[c][s].x;
        srho_old += rold[c];
        sdu += dui[c];
      }
      const float fac = m.nonlinear ? 2.f * srho_old + r0v : r0v;
      const int j = (S.active && S.mode != 2) ? b.slot[(2 * t + h) * NX + a] : -1;
      float srho = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c >= g.dim) break;
        const float pf = c == 0 ? m.pml[0][a] : (c == 1 ? m.pml[1][r % g.Ny] : m.pml[2][r / g.Ny]);
        float rn = pf * (pf * rold[c] - m.dt * fac * dui[c]);
        if (S.active) {
          if (S.mode == 2) {
            rn += S.many ? (h ? sadd[s].y : sadd[s].x) : S.sval * S.pattern[idx];
          } else if (j >= 0) {
            const float sv = S.sig[S.many ? j : 0];
            rn = (S.mode == 0) ? sv : rn + sv;
          }
        }
        A.rho[c][idx] = rn;
        srho += rn;
      }
      if (absorb) {
        A.srho[idx] = srho;
        outA[h] = r0v * sdu;
        outC[h] = srho;
      } else {
        const float c0v = m.c0.at(idx);
        float pv = srho;
        if (m.absorbing == 2) pv += m.tau.at(idx) * r0v * sdu;
        if (m.nonlinear) pv += m.bona.at(idx) * srho * srho / (2.f * r0v);
        pv *= c0v * c0v;
        outA[h] = pv;
        field_rec(A.ps, idx, pv);
      }
    }
    b.U(t, a) = make_float2(outA[0], outA[1]);
    if (absorb) b.U2(t, a) = make_float2(outC[0], outC[1]);
  }
  __syncthreads();
  if (absorb) {
    r2c_rows<NX>(g, b.U, A.G[0], r0, P, tw, b);
    __syncthreads();
    r2c_rows<NX>(g, b.U2, A.G[1], r0, P, tw, b);
  } else {
    record_rows<NX>(A.ps, b.U, g, r0, P);
    r2c_rows<NX>(g, b.U, A.pk, r0, P, tw, b);
  }
}

// ---------------------------------------------------------------------------
// power-law absorbing EOS (3D.m:970-975, 990-996)
// ---------------------------------------------------------------------------
template <int NX>
__global__ void __launch_bounds__(256) k_xpres_t(Grid g, const float2* __restrict__ tw, Medium m,
                                                 const float2* __restrict__ GA, const float2* __restrict__ GB,
                                                 const float* __restrict__ srho, float2* pk, Sens ps, int P) {
  using C = RowCfg<NX>;
  extern __shared__ float2 smem[];
  RowBufs<NX> b(smem, P, false);
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  float2 va[C::SLOTS], vb[C::SLOTS];
  {
    auto sink = [&](int, int, float2 v, int s) { va[s] = v; };
    c2r_rows<NX>(g, GA, r0, P, tw, b, sink);
  }
  __syncthreads();
  {
    auto sink = [&](int, int, float2 v, int s) { vb[s] = v; };
    c2r_rows<NX>(g, GB, r0, P, tw, b, sink);
  }
#pragma unroll
  for (int s = 0; s < C::SLOTS; ++s) {
    int t, a;
    if (!kwt::last_pos<NX, C::PMAX, C::NT, false>(P, s, t, a)) continue;
    float out[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * t + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * NX + a;
      const float sr = srho[idx];
      const float c0v = m.c0.at(idx);
      float pv = sr + m.tau.at(idx) * (h ? va[s].y : va[s].x) - m.eta.at(idx) * (h ? vb[s].y : vb[s].x);
      if (m.nonlinear) pv += m.bona.at(idx) * sr * sr / (2.f * m.rho0.at(idx));
      pv *= c0v * c0v;
      out[h] = pv;
      field_rec(ps, idx, pv);
    }
    b.U(t, a) = make_float2(out[0], out[1]);
  }
  __syncthreads();
  record_rows<NX>(ps, b.U, g, r0, P);
  r2c_rows<NX>(g, b.U, pk, r0, P, tw, b);
}

// initial pressure (3D.m:1024-1028)
template <int NX>
__global__ void __launch_bounds__(256) k_xinit_t(Grid g, const float2* __restrict__ tw, Medium m,
                                                 const float* __restrict__ p0, float* rx, float* ry, float* rz,
                                                 float2* pk, Sens ps, int P) {
  extern __shared__ float2 smem[];
  RowBufs<NX> b(smem, P, false);
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  for (int e = threadIdx.x; e < P * NX; e += RowCfg<NX>::NT) {
    const int t = e / NX, a = e - t * NX;
    float out[2] = {0.f, 0.f};
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * t + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * NX + a;
      const float pv = p0[idx];
      const float c0v = m.c0.at(idx);
      const float rv = pv / ((float)g.dim * c0v * c0v);
      rx[idx] = rv;
      ry[idx] = rv;
      if (g.dim == 3) rz[idx] = rv;
      out[h] = pv;
      field_rec(ps, idx, pv);
    }
    b.U(t, a) = make_float2(out[0], out[1]);
  }
  __syncthreads();
  record_rows<NX>(ps, b.U, g, r0, P);
  r2c_rows<NX>(g, b.U, pk, r0, P, tw, b);
}

// sparse source field -> x R2C (3D.m:815-822); unit scatters 1.0
template <int NX>
__global__ void __launch_bounds__(256) k_xsrc_t(Grid g, const float2* __restrict__ tw, Src S, int unit, float2* out,
                                                int P) {
  extern __shared__ float2 smem[];
  RowBufs<NX> b(smem, P, false);
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  for (int e = threadIdx.x; e < P * NX; e += RowCfg<NX>::NT) b.U(e / NX, e % NX) = make_float2(0.f, 0.f);
  __syncthreads();
  const int rend = min(r0 + 2 * P, g.rows);
  for (int j = S.row_ptr[r0] + threadIdx.x; j < S.row_ptr[rend]; j += RowCfg<NX>::NT) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / NX);
    const int lr = r - r0, x = (int)(lin - (long long)r * NX);
    const float v = unit ? 1.f : S.sig[S.many ? j : 0];
    float2& w = b.U(lr >> 1, x);
    if (lr & 1)
      w.y = v;
    else
      w.x = v;
  }
  __syncthreads();
  r2c_rows<NX>(g, b.U, out, r0, P, tw, b);
}

// plain x C2R into a real field
template <int NX>
__global__ void __launch_bounds__(256) k_xc2r_t(Grid g, const float2* __restrict__ tw, const float2* __restrict__ in,
                                                float* out, int P) {
  extern __shared__ float2 smem[];
  RowBufs<NX> b(smem, P, false);
  const int r0 = kwt::xcd_swizzle(blockIdx.x, gridDim.x) * 2 * P;
  auto sink = [&](int t, int a, float2 v, int) {
    const int ra = r0 + 2 * t;
    if (ra < g.rows) out[(long long)ra * NX + a] = v.x;
    if (ra + 1 < g.rows) out[(long long)(ra + 1) * NX + a] = v.y;
  };
  c2r_rows<NX>(g, in, r0, P, tw, b, sink);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#define KW_ROW_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(384) X(512) X(768)

template <int NX>
static size_t row_smem(int P, bool two_u, bool slot) {
  using C = RowCfg<NX>;
  size_t s = (size_t)(C::NWORK + 1 + (two_u ? 1 : 0)) * P * C::PADN * sizeof(float2);
  if (slot) s += (size_t)2 * P * NX * sizeof(int);
  return s;
}

// pairs per workgroup: up to PMAX, fewer when that leaves under ~512 workgroups
// or the LDS footprint would pass 64 KiB
template <int NX>
static int row_pairs(const Grid& g, bool two_u, bool slot) {
  using C = RowCfg<NX>;
  int P = C::PMAX;
  while (P > 1 && (g.rows + 2 * P - 1) / (2 * P) < 512) P /= 2;
  while (P > 1 && row_smem<NX>(P, two_u, slot) > 65536) P /= 2;
  return P;
}

static inline unsigned row_blocks(const Grid& g, int P) { return (unsigned)((g.rows + 2 * P - 1) / (2 * P)); }

bool row_supported(int nx) {
  switch (nx) {
#define KW_CASE(n) case n:
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    return true;
    default:
      return false;
  }
}

template <int NX>
static void xvel(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G, float* const* u,
                 const Src* src, const Sens* sens, int init) {
  VelArgsT A;
  bool slot = false;
  for (int c = 0; c < 3; ++c) {
    A.G[c] = c < g.dim ? G[c] : nullptr;
    A.u[c] = c < g.dim ? u[c] : nullptr;
    A.src[c] = src[c];
    A.sens[c] = sens[c];
    slot = slot || (src[c].active && src[c].mode != 2);
  }
  A.init = init;
  const int P = row_pairs<NX>(g, false, slot);
  hipLaunchKernelGGL(k_xvel_t<NX>, dim3(row_blocks(g, P), g.dim), dim3(256), row_smem<NX>(P, false, slot), s, g, tw,
                     m, A, P);
}

template <int NX>
static void xdens(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G,
                  float* const* rho, const Src& psrc, float* srho, float2* pk, const Sens& ps) {
  DensArgsT A;
  for (int c = 0; c < 3; ++c) {
    A.G[c] = c < g.dim ? G[c] : nullptr;
    A.rho[c] = c < g.dim ? rho[c] : nullptr;
  }
  A.psrc = psrc;
  A.srho = srho;
  A.pk = pk;
  A.ps = ps;
  const bool two = m.absorbing == 1, slot = psrc.active && psrc.mode != 2;
  const int P = row_pairs<NX>(g, two, slot);
  hipLaunchKernelGGL(k_xdens_t<NX>, dim3(row_blocks(g, P)), dim3(256), row_smem<NX>(P, two, slot), s, g, tw, m, A, P);
}

template <int NX>
static void xpres(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float2* GA, const float2* GB,
                  const float* srho, float2* pk, const Sens& ps) {
  const int P = row_pairs<NX>(g, false, false);
  hipLaunchKernelGGL(k_xpres_t<NX>, dim3(row_blocks(g, P)), dim3(256), row_smem<NX>(P, false, false), s, g, tw, m,
                     GA, GB, srho, pk, ps, P);
}

template <int NX>
static void xinit(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float* p0,
                  float* const* rho, float2* pk, const Sens& ps) {
  const int P = row_pairs<NX>(g, false, false);
  hipLaunchKernelGGL(k_xinit_t<NX>, dim3(row_blocks(g, P)), dim3(256), row_smem<NX>(P, false, false), s, g, tw, m,
                     p0, rho[0], rho[1], g.dim == 3 ? rho[2] : nullptr, pk, ps, P);
}

template <int NX>
static void xsrc(hipStream_t s, const Grid& g, const float2* tw, const Src& src, int unit, float2* out) {
  const int P = row_pairs<NX>(g, false, false);
  hipLaunchKernelGGL(k_xsrc_t<NX>, dim3(row_blocks(g, P)), dim3(256), row_smem<NX>(P, false, false), s, g, tw, src,
                     unit, out, P);
}

template <int NX>
static void xc2r(hipStream_t s, const Grid& g, const float2* tw, const float2* in, float* out) {
  const int P = row_pairs<NX>(g, false, false);
  hipLaunchKernelGGL(k_xc2r_t<NX>, dim3(row_blocks(g, P)), dim3(256), row_smem<NX>(P, false, false), s, g, tw, in,
                     out, P);
}

bool launch_x_velocity_fast(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G,
                            float* const* u, const Src* src, const Sens* sens, int init) {
  switch (g.Nx) {
#define KW_CASE(n)                                  \
  case n:                                           \
    xvel<n>(s, g, tw, m, G, u, src, sens, init);    \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_x_density_fast(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G,
                           float* const* rho, const Src& psrc, float* srho, float2* pk, const Sens& ps) {
  switch (g.Nx) {
#define KW_CASE(n)                                       \
  case n:                                                \
    xdens<n>(s, g, tw, m, G, rho, psrc, srho, pk, ps);   \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_x_pressure_fast(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float2* GA,
                            const float2* GB, const float* srho, float2* pk, const Sens& ps) {
  switch (g.Nx) {
#define KW_CASE(n)                                 \
  case n:                                          \
    xpres<n>(s, g, tw, m, GA, GB, srho, pk, ps);   \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_x_init_fast(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float* p0,
                        float* const* rho, float2* pk, const Sens& ps) {
  switch (g.Nx) {
#define KW_CASE(n)                              \
  case n:                                       \
    xinit<n>(s, g, tw, m, p0, rho, pk, ps);     \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_x_source_fast(hipStream_t s, const Grid& g, const float2* tw, const Src& src, int unit, float2* out) {
  switch (g.Nx) {
#define KW_CASE(n)                          \
  case n:                                   \
    xsrc<n>(s, g, tw, src, unit, out);      \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_x_c2r_fast(hipStream_t s, const Grid& g, const float2* tw, const float2* in, float* out) {
  switch (g.Nx) {
#define KW_CASE(n)                  \
  case n:                           \
    xc2r<n>(s, g, tw, in, out);     \
    return true;
    KW_ROW_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

}  // namespace kw
