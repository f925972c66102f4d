// HIP kernels (gfx950) for the k-space pseudospectral time loop.
//
// One time step (3D, power-law absorbing) is this sequence of passes; 2D has
// no y pass because y is then the last axis:
//   y fwd(Pk) -> last[Pk -> G_i = D+_i kappa P] -> y inv(G) -> x_velocity
//   -> y fwd(G) -> last[G_i -> D-_i kappa G_i] -> y inv(G) -> x_density
//   -> y fwd(G0,G1) -> last[nabla1, nabla2] -> y inv -> x_pressure
// The x kernels hold all elementwise physics (PML, density, equation of
// state, sources, sensors) between the x-axis C2R and R2C of their rows.
// Statement references are to k-Wave/kspaceFirstOrder3D.m.
#include <math.h>

#include <algorithm>

#include "kw_kernels.hpp"

namespace kw {

#define KW_MAXE 4  // elements owned per thread in the x kernels

// ---------------------------------------------------------------------------
// x-row helpers: two real rows (ra = r0+2p, rb = ra+1) share one complex row.
// ---------------------------------------------------------------------------

// Half spectra of rows ra, rb -> packed full spectrum Z = A + iB (Hermitian
// extension; DC and Nyquist imaginary parts dropped as a C2R does).
__device__ __forceinline__ void load_c2r_pairs(const float2* __restrict__ G, float2* buf, const Grid& g,
                                               int r0, int P) {
  const int Nx = g.Nx, Nxh = g.Nxh;
  for (int e = threadIdx.x; e < P * Nxh; e += blockDim.x) {
    const int p = e / Nxh, k = e - p * Nxh;
    const int ra = r0 + 2 * p, rb = ra + 1;
    float2 A = ra < g.rows ? G[(long long)ra * Nxh + k] : make_float2(0.f, 0.f);
    float2 B = rb < g.rows ? G[(long long)rb * Nxh + k] : make_float2(0.f, 0.f);
    const bool self = (k == 0) || (2 * k == Nx);
    if (self) {
      A.y = 0.f;
      B.y = 0.f;
    }
    buf[p * Nx + k] = make_float2(A.x - B.y, A.y + B.x);
    if (!self) buf[p * Nx + Nx - k] = make_float2(A.x + B.y, B.x - A.y);
  }
}

// Packed spectrum Z of rows (a + ib) -> half spectra A (row ra), B (row rb).
__device__ __forceinline__ void store_r2c_pairs(const float2* buf, float2* __restrict__ G, const Grid& g,
                                                int r0, int P) {
  const int Nx = g.Nx, Nxh = g.Nxh;
  for (int e = threadIdx.x; e < P * Nxh; e += blockDim.x) {
    const int p = e / Nxh, k = e - p * Nxh;
    const int ra = r0 + 2 * p, rb = ra + 1;
    const float2 Z = buf[p * Nx + k];
    const float2 Zm = buf[p * Nx + (k == 0 ? 0 : Nx - k)];
    if (ra < g.rows) G[(long long)ra * Nxh + k] = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
    if (rb < g.rows) G[(long long)rb * Nxh + k] = make_float2(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  }
}

__device__ __forceinline__ void load_twiddles(float2* tw, const float2* __restrict__ tw_g, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) tw[j] = tw_g[j];
}

// Gather sensor values of the block's rows from the packed real rows in buf
// (extractSensorData.m:61-140: raw, running max/min, rms).
__device__ void record_points(const Sens& S, const float2* buf, const Grid& g, int r0, int P) {
  if (!S.record || S.ns == 0) return;
  const int rend = min(r0 + 2 * P, g.rows);
  const int j0 = S.row_ptr[r0], j1 = S.row_ptr[rend];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / g.Nx);
    const int x = (int)(lin - (long long)r * g.Nx);
    const int lr = r - r0;
    const float2 w = buf[(lr >> 1) * g.Nx + x];
    const float v = (lr & 1) ? w.y : w.x;
    if (S.raw) S.raw[S.t_rec * S.ns + j] = v;
    if (S.vmax) S.vmax[j] = S.first ? v : fmaxf(S.vmax[j], v);
    if (S.vmin) S.vmin[j] = S.first ? v : fminf(S.vmin[j], v);
    if (S.rms_sum) {
      // compensated (Kahan) FP64 accumulation of v^2
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

// Full-grid running max/min and final-field write for one element.
__device__ __forceinline__ void record_field(const Sens& S, long long idx, float v) {
  if (S.record) {
    if (S.max_all) S.max_all[idx] = S.first ? v : fmaxf(S.max_all[idx], v);
    if (S.min_all) S.min_all[idx] = S.first ? v : fminf(S.min_all[idx], v);
  }
  if (S.final_field) S.final_field[idx] = v;
}

// Marks the block's sparse source points in an LDS slot map (-1 = none).
__device__ void build_slot_map(const Src& S, int* slot, const Grid& g, int r0, int P) {
  const int n = 2 * P * g.Nx;
  for (int e = threadIdx.x; e < n; e += blockDim.x) slot[e] = -1;
  __syncthreads();
  const int rend = min(r0 + 2 * P, g.rows);
  const int j0 = S.row_ptr[r0], j1 = S.row_ptr[rend];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / g.Nx);
    const int x = (int)(lin - (long long)r * g.Nx);
    slot[(r - r0) * g.Nx + x] = j;
  }
  __syncthreads();
}


// ---------------------------------------------------------------------------
// velocity update (3D.m:790-801), velocity sources (3D.m:804-884), u sensors,
// then the x R2C of the new velocity for the divergence.
// blockIdx.y = component.  init: p0 start (3D.m:1029-1033), u = dt/2 rho^-1 grad.
// ---------------------------------------------------------------------------
struct VelArgs {
  float2* G[3];
  float* u[3];
  Src src[3];
  Sens sens[3];
  int init;
};

__global__ void __launch_bounds__(256) k_x_velocity(Grid g, FftAxis ax, const float2* __restrict__ tw_g,
                                                    Medium m, VelArgs A, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  int* slot = reinterpret_cast<int*>(bufB + P * Nx);
  const int i = blockIdx.y;
  const int r0 = blockIdx.x * 2 * P;
  // select this component's arguments with uniform branches (no dynamic
  // indexing into kernel-argument arrays)
  const Src S = i == 0 ? A.src[0] : (i == 1 ? A.src[1] : A.src[2]);
  const Sens SE = i == 0 ? A.sens[0] : (i == 1 ? A.sens[1] : A.sens[2]);
  float2* Gi = i == 0 ? A.G[0] : (i == 1 ? A.G[1] : A.G[2]);
  float* ui = i == 0 ? A.u[0] : (i == 1 ? A.u[1] : A.u[2]);
  const SF rsg = i == 0 ? m.rsg[0] : (i == 1 ? m.rsg[1] : m.rsg[2]);
  const float* pmlsg = i == 0 ? m.pml_sg[0] : (i == 1 ? m.pml_sg[1] : m.pml_sg[2]);
  const bool src_on = !A.init && S.active;
  load_twiddles(tw, tw_g, Nx);

  float2 sadd[KW_MAXE];
  if (src_on && S.mode == 2 && S.many) {
    load_c2r_pairs(S.spec, bufA, g, r0, P);
    __syncthreads();
    const float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
#pragma unroll
    for (int k = 0; k < KW_MAXE; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      sadd[k] = e < P * Nx ? res[e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  if (src_on && S.mode != 2) build_slot_map(S, slot, g, r0, P);

  load_c2r_pairs(Gi, bufA, g, r0, P);
  __syncthreads();
  float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
  float2* other = res == bufA ? bufB : bufA;

#pragma unroll
  for (int k = 0; k < KW_MAXE; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e >= P * Nx) break;
    const int p = e / Nx, x = e - p * Nx;
    const float2 grad = res[e];
    float un[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * Nx + x;
      const float gv = h ? grad.y : grad.x;
      const float rs = rsg.at(idx);
      float v;
      if (A.init) {
        v = 0.5f * m.dt * rs * gv;
      } else {
        const float pf = i == 0 ? pmlsg[x] : (i == 1 ? pmlsg[r % g.Ny] : pmlsg[r / g.Ny]);
        v = pf * (pf * ui[idx] - m.dt * rs * gv);
        if (src_on) {
          if (S.mode == 2) {
            v += S.many ? (h ? sadd[k].y : sadd[k].x) : S.sval * S.pattern[idx];
          } else {
            const int j = slot[(2 * p + h) * Nx + x];
            if (j >= 0) {
              const float sv = S.delay ? S.sig[S.delay[j]] : S.sig[S.many ? j : 0];
              v = (S.mode == 0) ? sv : v + sv;
            }
          }
        }
      }
      un[h] = v;
      ui[idx] = v;
      record_field(SE, idx, v);
    }
    res[e] = make_float2(un[0], un[1]);
  }
  __syncthreads();
  record_points(SE, res, g, r0, P);
  if (A.init) return;
  const float2* spec = lds_fft<false>(res, other, tw, ax, P, Nx, -1.f);
  store_r2c_pairs(spec, Gi, g, r0, P);
}

// ---------------------------------------------------------------------------
// density update (3D.m:910-925), pressure source (3D.m:928-960) and the
// equation of state (3D.m:963-1017).  Lossless/stokes: p is formed here and
// recorded, then its x R2C goes to pk.  Power-law absorbing: writes the x R2C
// of rho0*sum(du) and of sum(rho) for the fractional Laplacians, and sum(rho).
// ---------------------------------------------------------------------------
struct DensArgs {
  float2* G[3];
  float* rho[3];
  Src psrc;
  float* srho;
  float2* pk;
  Sens ps;
};

__global__ void __launch_bounds__(256) k_x_density(Grid g, FftAxis ax, const float2* __restrict__ tw_g,
                                                   Medium m, DensArgs A, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  float2* bufC = bufB + P * Nx;
  int* slot = reinterpret_cast<int*>(bufC + P * Nx);
  const int r0 = blockIdx.x * 2 * P;
  const Src& S = A.psrc;
  load_twiddles(tw, tw_g, Nx);

  float2 sadd[KW_MAXE];
  if (S.active && S.mode == 2 && S.many) {
    load_c2r_pairs(S.spec, bufA, g, r0, P);
    __syncthreads();
    const float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
#pragma unroll
    for (int k = 0; k < KW_MAXE; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      sadd[k] = e < P * Nx ? res[e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  if (S.active && S.mode != 2) build_slot_map(S, slot, g, r0, P);

  float2 du[3][KW_MAXE];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= g.dim) {
#pragma unroll
      for (int k = 0; k < KW_MAXE; ++k) du[c][k] = make_float2(0.f, 0.f);
      continue;
    }
    load_c2r_pairs(A.G[c], bufA, g, r0, P);
    __syncthreads();
    const float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
#pragma unroll
    for (int k = 0; k < KW_MAXE; ++k) {
      const int e = threadIdx.x + k * blockDim.x;
      du[c][k] = e < P * Nx ? res[e] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }

  const bool absorb = m.absorbing == 1;
#pragma unroll
  for (int k = 0; k < KW_MAXE; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e >= P * Nx) break;
    const int p = e / Nx, x = e - p * Nx;
    float outA[2] = {0.f, 0.f}, outC[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * Nx + x;
      const float r0v = m.rho0.at(idx);
      float rold[3], dui[3];
      float srho_old = 0.f, sdu = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        rold[c] = c < g.dim ? A.rho[c][idx] : 0.f;
        dui[c] = h ? du[c][k].y : du[c][k].x;
        srho_old += rold[c];
        sdu += dui[c];
      }
      const float fac = m.nonlinear ? 2.f * srho_old + r0v : r0v;
      int j = -1;
      if (S.active && S.mode != 2) j = slot[(2 * p + h) * Nx + x];
      float srho = 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (c >= g.dim) break;
        const float pf = c == 0 ? m.pml[0][x] : (c == 1 ? m.pml[1][r % g.Ny] : m.pml[2][r / g.Ny]);
        float rn = pf * (pf * rold[c] - m.dt * fac * dui[c]);
        if (S.active) {
          if (S.mode == 2) {
            rn += S.many ? (h ? sadd[k].y : sadd[k].x) : S.sval * S.pattern[idx];
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
        record_field(A.ps, idx, pv);
      }
    }
    bufA[e] = make_float2(outA[0], outA[1]);
    if (absorb) bufC[e] = make_float2(outC[0], outC[1]);
  }
  __syncthreads();
  if (absorb) {
    const float2* spec = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, -1.f);
    store_r2c_pairs(spec, A.G[0], g, r0, P);
    __syncthreads();
    spec = lds_fft<false>(bufC, bufB, tw, ax, P, Nx, -1.f);
    store_r2c_pairs(spec, A.G[1], g, r0, P);
  } else {
    record_points(A.ps, bufA, g, r0, P);
    const float2* spec = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, -1.f);
    store_r2c_pairs(spec, A.pk, g, r0, P);
  }
}

// ---------------------------------------------------------------------------
// power-law absorbing equation of state (3D.m:970-975, 990-996):
// p = c0^2 (sum rho + tau*IFFT(nabla1 FFT(rho0 sum du)) - eta*IFFT(nabla2 FFT(sum rho))
//            [+ B/A sum(rho)^2 / (2 rho0)])
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_x_pressure(Grid g, FftAxis ax, const float2* __restrict__ tw_g,
                                                    Medium m, const float2* __restrict__ GA,
                                                    const float2* __restrict__ GB, const float* __restrict__ srho,
                                                    float2* pk, Sens ps, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  const int r0 = blockIdx.x * 2 * P;
  load_twiddles(tw, tw_g, Nx);
  float2 va[KW_MAXE], vb[KW_MAXE];
  load_c2r_pairs(GA, bufA, g, r0, P);
  __syncthreads();
  const float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
#pragma unroll
  for (int k = 0; k < KW_MAXE; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    va[k] = e < P * Nx ? res[e] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  load_c2r_pairs(GB, bufA, g, r0, P);
  __syncthreads();
  res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
#pragma unroll
  for (int k = 0; k < KW_MAXE; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    vb[k] = e < P * Nx ? res[e] : make_float2(0.f, 0.f);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < KW_MAXE; ++k) {
    const int e = threadIdx.x + k * blockDim.x;
    if (e >= P * Nx) break;
    const int p = e / Nx, x = e - p * Nx;
    float out[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * Nx + x;
      const float s = srho[idx];
      const float c0v = m.c0.at(idx);
      float pv = s + m.tau.at(idx) * (h ? va[k].y : va[k].x) - m.eta.at(idx) * (h ? vb[k].y : vb[k].x);
      if (m.nonlinear) pv += m.bona.at(idx) * s * s / (2.f * m.rho0.at(idx));
      pv *= c0v * c0v;
      out[h] = pv;
      record_field(ps, idx, pv);
    }
    bufA[e] = make_float2(out[0], out[1]);
  }
  __syncthreads();
  record_points(ps, bufA, g, r0, P);
  const float2* spec = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, -1.f);
  store_r2c_pairs(spec, pk, g, r0, P);
}

// ---------------------------------------------------------------------------
// initial pressure (3D.m:1024-1028): p = p0, rho_i = p0 / (dim c0^2); record p;
// x R2C of p0 into pk.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_x_init(Grid g, FftAxis ax, const float2* __restrict__ tw_g, Medium m,
                                                const float* __restrict__ p0, float* rho0x, float* rho0y,
                                                float* rho0z, float2* pk, Sens ps, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  const int r0 = blockIdx.x * 2 * P;
  load_twiddles(tw, tw_g, Nx);
  for (int e = threadIdx.x; e < P * Nx; e += blockDim.x) {
    const int p = e / Nx, x = e - p * Nx;
    float out[2] = {0.f, 0.f};
    for (int h = 0; h < 2; ++h) {
      const int r = r0 + 2 * p + h;
      if (r >= g.rows) continue;
      const long long idx = (long long)r * Nx + x;
      const float pv = p0[idx];
      const float c0v = m.c0.at(idx);
      const float rv = pv / ((float)g.dim * c0v * c0v);
      rho0x[idx] = rv;
      rho0y[idx] = rv;
      if (g.dim == 3) rho0z[idx] = rv;
      out[h] = pv;
      record_field(ps, idx, pv);
    }
    bufA[e] = make_float2(out[0], out[1]);
  }
  __syncthreads();
  record_points(ps, bufA, g, r0, P);
  const float2* spec = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, -1.f);
  store_r2c_pairs(spec, pk, g, r0, P);
}

// ---------------------------------------------------------------------------
// sparse source field -> x R2C (for 'additive' sources: 3D.m:815-822).
// unit != 0 scatters 1.0 at every source point (used to build the pattern).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_x_source(Grid g, FftAxis ax, const float2* __restrict__ tw_g, Src S,
                                                  int unit, float2* out, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  const int r0 = blockIdx.x * 2 * P;
  load_twiddles(tw, tw_g, Nx);
  for (int e = threadIdx.x; e < P * Nx; e += blockDim.x) bufA[e] = make_float2(0.f, 0.f);
  __syncthreads();
  const int rend = min(r0 + 2 * P, g.rows);
  const int j0 = S.row_ptr[r0], j1 = S.row_ptr[rend];
  for (int j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
    const long long lin = S.lin[j];
    const int r = (int)(lin / Nx);
    const int x = (int)(lin - (long long)r * Nx);
    const int lr = r - r0;
    const float v = unit ? 1.f : S.sig[S.many ? j : 0];
    float2* w = &bufA[(lr >> 1) * Nx + x];
    if (lr & 1)
      w->y = v;
    else
      w->x = v;
  }
  __syncthreads();
  const float2* spec = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, -1.f);
  store_r2c_pairs(spec, out, g, r0, P);
}

// plain x C2R of an X-domain array into a real field
__global__ void __launch_bounds__(256) k_x_c2r(Grid g, FftAxis ax, const float2* __restrict__ tw_g,
                                               const float2* __restrict__ in, float* out, int P) {
  extern __shared__ float2 smem[];
  const int Nx = g.Nx;
  float2* tw = smem;
  float2* bufA = tw + Nx;
  float2* bufB = bufA + P * Nx;
  const int r0 = blockIdx.x * 2 * P;
  load_twiddles(tw, tw_g, Nx);
  load_c2r_pairs(in, bufA, g, r0, P);
  __syncthreads();
  const float2* res = lds_fft<false>(bufA, bufB, tw, ax, P, Nx, 1.f);
  for (int e = threadIdx.x; e < P * Nx; e += blockDim.x) {
    const int p = e / Nx, x = e - p * Nx;
    const int ra = r0 + 2 * p;
    if (ra < g.rows) out[(long long)ra * Nx + x] = res[e].x;
    if (ra + 1 < g.rows) out[(long long)(ra + 1) * Nx + x] = res[e].y;
  }
}

// ---------------------------------------------------------------------------
// y pass (3D): complex FFT along y of X-domain arrays, tiles of tx kx columns
// at one z.  blockIdx = (kx tile, z, field).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ypass(Grid g, FftAxis ax, const float2* __restrict__ tw_g, Fields3 F,
                                               float sign, int TX) {
  extern __shared__ float2 smem[];
  const int Ny = g.Ny, Nxh = g.Nxh;
  float2* tw = smem;
  float2* bufA = tw + Ny;
  float2* bufB = bufA + Ny * TX;
  float2* f = F.f[blockIdx.z];
  const int kx0 = blockIdx.x * TX;
  const long long base = (long long)blockIdx.y * Ny * Nxh + kx0;
  load_twiddles(tw, tw_g, Ny);
  for (int e = threadIdx.x; e < Ny * TX; e += blockDim.x) {
    const int a = e / TX, t = e - a * TX;
    bufA[e] = (kx0 + t < Nxh) ? f[base + (long long)a * Nxh + t] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* res = lds_fft<true>(bufA, bufB, tw, ax, TX, 0, sign);
  for (int e = threadIdx.x; e < Ny * TX; e += blockDim.x) {
    const int a = e / TX, t = e - a * TX;
    if (kx0 + t < Nxh) f[base + (long long)a * Nxh + t] = res[e];
  }
}

// ---------------------------------------------------------------------------
// last axis (z in 3D, y in 2D): forward FFT, spectral operator, inverse FFT.
// Pencils are indexed by q over the contiguous (kx[,y]) plane, so a tile of
// tx consecutive q is one coalesced segment per axis position.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_last(Grid g, FftAxis ax, const float2* __restrict__ tw_g, LastArgs A,
                                              SpecTables T, int TX) {
  extern __shared__ float2 smem[];
  const int n = ax.n;
  const int Q = g.dim == 3 ? g.Nxh * g.Ny : g.Nxh;
  float2* tw = smem;
  float2* b0 = tw + n;
  float2* b1 = b0 + n * TX;
  float2* b2 = b1 + n * TX;
  const int q0 = blockIdx.x * TX;
  load_twiddles(tw, tw_g, n);
  int cur = -1;
  float2* spec = b0;
  float2* w1 = b1;
  for (int o = 0; o < A.nout; ++o) {
    const SpecOut so = A.outs[o];
    if (so.src != cur) {
      const float2* in = A.in[so.src];
      for (int e = threadIdx.x; e < n * TX; e += blockDim.x) {
        const int a = e / TX, t = e - a * TX;
        const int q = q0 + t;
        b0[e] = q < Q ? in[(long long)a * Q + q] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      spec = lds_fft<true>(b0, b1, tw, ax, TX, 0, -1.f);
      w1 = spec == b0 ? b1 : b0;
      cur = so.src;
    }
    for (int e = threadIdx.x; e < n * TX; e += blockDim.x) {
      const int a = e / TX, t = e - a * TX;
      const int q = q0 + t;
      float2 v = make_float2(0.f, 0.f);
      if (q < Q) {
        const long long sidx = (long long)a * Q + q;
        v = spec[e];
        int kx, ky, kz;
        if (g.dim == 3) {
          ky = q / g.Nxh;
          kx = q - ky * g.Nxh;
          kz = a;
        } else {
          kx = q;
          ky = a;
          kz = 0;
        }
        switch (so.op) {
          case OP_DPOS_K:
          case OP_DNEG_K: {
            const float2* dv = so.op == OP_DPOS_K ? T.dpos[so.axis] : T.dneg[so.axis];
            const int kk = so.axis == 0 ? kx : (so.axis == 1 ? ky : kz);
            v = cscale(cmul(dv[kk], v), T.kappa[sidx]);
          } break;
          case OP_SHIFT: {
            const int kk = so.axis == 0 ? kx : (so.axis == 1 ? ky : kz);
            v = cmul(T.shift[so.axis][kk], v);
          } break;
          case OP_NABLA1: v = cscale(v, T.nabla1[sidx]); break;
          case OP_NABLA2: v = cscale(v, T.nabla2[sidx]); break;
          case OP_SKAPPA: v = cscale(v, T.skappa[sidx]); break;
          default: break;
        }
        v = cscale(v, A.scale);
      }
      w1[e] = v;
    }
    __syncthreads();
    const float2* res = lds_fft<true>(w1, b2, tw, ax, TX, 0, 1.f);
    float2* out = A.out[o];
    for (int e = threadIdx.x; e < n * TX; e += blockDim.x) {
      const int a = e / TX, t = e - a * TX;
      const int q = q0 + t;
      if (q < Q) out[(long long)a * Q + q] = res[e];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// FP64 precompute kernels
// ---------------------------------------------------------------------------
__global__ void k_twiddles(float2* tw, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[m] = make_float2((float)c, (float)-s);
}

// FFT-order wavenumber of bin j for an n-point axis with spacing d
// (ifftshift of kWaveGrid.makeDim, kWaveGrid.m:779-797).
__device__ __forceinline__ double kvec(int j, int n, double d) {
  const int pos = n - n / 2;  // bins [0, pos) are non-negative
  const double nx = (j < pos) ? (double)j / n : (double)(j - n) / n;
  return 2.0 * M_PI / d * nx;
}

// ddx_k_shift_pos/neg = i k exp(+-i k d/2)  (3D.m:627-634)
// shift: x_shift_neg = exp(-i k d/2) (kspaceFirstOrder_createStorageVariables.m:87-99)
__global__ void k_dvec(float2* dpos, float2* dneg, float2* shift, int nfull, int nout, double d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nout) return;
  const double k = kvec(j, nfull, d);
  double s, c;
  sincos(k * d / 2.0, &s, &c);
  dpos[j] = make_float2((float)(-k * s), (float)(k * c));
  dneg[j] = make_float2((float)(k * s), (float)(k * c));
  // real(ifft(...)) drops the imaginary Nyquist term i*U_N of a real u, so the
  // shift is 0 there (keeps the product Hermitian for the half-spectrum path)
  if (shift) shift[j] = (nfull % 2 == 0 && 2 * j == nfull) ? make_float2(0.f, 0.f) : make_float2((float)c, (float)(-s));
}

__global__ void k_gather(const float* __restrict__ field, const long long* __restrict__ lin, int n,
                         float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = field[lin[j]];
}

// getPML.m:50-85
// one_sided: the radial axis of axisymmetric grids (getPML.m:73-77)
__global__ void k_pml(float* pml, int n, double d, double dt, double c, int size, double alpha, int staggered,
                      int one_sided) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double v = 1.0;
  const double off = staggered ? 0.5 : 0.0;
  if (j < size && !one_sided) {
    const double x = j + 1 + off;
    const double r = (x - size - 1) / (0.0 - size);
    v = exp(-alpha * (c / d) * r * r * r * r * dt / 2.0);
  } else if (j >= n - size) {
    const double x = j - (n - size) + 1 + off;
    const double r = x / size;
    v = exp(-alpha * (c / d) * r * r * r * r * dt / 2.0);
  }
  pml[j] = (float)v;
}

// kappa = sinc(c_ref k dt/2), source_kappa = cos(c_ref k dt/2) (3D.m:651-655);
// nabla1 = k^(y-2), nabla2 = k^(y-1), Inf -> 0 (createAbsorptionVariables.m:34-48)
__global__ void k_spectral_tables(Grid g, OpParams p, float* kappa, float* skappa, float* nabla1,
                                  float* nabla2) {
  const long long nspec = (long long)g.Nxs * g.Ny * g.Nz;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nspec) return;
  int kxi, kyi, kzi;
  if (g.kxm) {  // [kx][z][y]
    kxi = (int)(id / ((long long)g.Ny * g.Nz));
    const long long rest = id - (long long)kxi * g.Ny * g.Nz;
    kzi = (int)(rest / g.Ny);
    kyi = (int)(rest - (long long)kzi * g.Ny);
  } else {      // [z][y][kx], row stride Nxs
    kxi = (int)(id % g.Nxs);
    const long long rest = id / g.Nxs;
    kyi = (int)(rest % g.Ny);
    kzi = (int)(rest / g.Ny);
    if (kxi >= g.Nxh) {  // row padding: never read
      if (kappa) kappa[id] = 0.f;
      if (skappa) skappa[id] = 0.f;
      if (nabla1) nabla1[id] = 0.f;
      if (nabla2) nabla2[id] = 0.f;
      return;
    }
  }
  const double kx = kvec(kxi, g.Nx, p.dx[0]);
  const double ky = kvec(kyi, g.Ny, p.dx[1]);
  const double kz = g.dim == 3 ? kvec(kzi, g.Nz, p.dx[2]) : 0.0;
  const double k = sqrt(kx * kx + ky * ky + kz * kz);
  const double arg = p.c_ref * k * p.dt / 2.0;
  if (kappa) kappa[id] = (float)(arg == 0.0 ? 1.0 : sin(arg) / arg);
  if (skappa) skappa[id] = (float)cos(arg);
  if (nabla1) {
    const double v = pow(k, p.alpha_power - 2.0);
    nabla1[id] = (float)(isinf(v) ? 0.0 : v);
  }
  if (nabla2) {
    const double v = pow(k, p.alpha_power - 1.0);
    nabla2[id] = (float)(isinf(v) ? 0.0 : v);
  }
}

// absorb_tau/eta (createAbsorptionVariables.m:31-48, 83-90), alpha in dB/(MHz^y cm)
__global__ void k_absorb_coeffs(long long n, const float* c0, float c0s, const float* alpha, float alphas,
                                double y, int absorbing, float* tau, float* eta) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = c0 ? (double)c0[i] : (double)c0s;
  const double a = alpha ? (double)alpha[i] : (double)alphas;
  const double yy = absorbing == 2 ? 2.0 : y;
  const double anp = 100.0 * a * pow(1e-6 / (2.0 * M_PI), yy) / (20.0 * log10(exp(1.0)));
  if (absorbing == 2) {
    tau[i] = (float)(-2.0 * anp * c);
  } else {
    tau[i] = (float)(-2.0 * anp * pow(c, yy - 1.0));
    eta[i] = (float)(2.0 * anp * pow(c, yy) * tan(M_PI * yy / 2.0));
  }
}

__global__ void k_reciprocal(long long n, const float* in, float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)(1.0 / (double)in[i]);
}

__global__ void k_finish_rms(long long n, const double* sum, double count, float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)sqrt(sum[2 * i] / count);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long n, int b) { return (unsigned)((n + b - 1) / b); }

bool fast_length(int n) {
  switch (n) {
#define KW_CASE(n) case n:
    KW_FAST_SIZES(KW_CASE)
#undef KW_CASE
    return true;
    default:
      return false;
  }
}

// dispatch to the length-specialised kernels instantiated for `nt` threads
#define KW_BY_NT(nt, fn, ...)                                   \
  ((nt) == 64 ? fn<64>(__VA_ARGS__)                             \
              : (nt) == 128 ? fn<128>(__VA_ARGS__) : (nt) == 256 ? fn<256>(__VA_ARGS__) : false)

// pencils also have 512-thread instantiations
#define KW_BY_NT_PENCIL(nt, fn, ...) \
  ((nt) == 1024 ? fn<1024>(__VA_ARGS__) : (nt) == 512 ? fn<512>(__VA_ARGS__) : KW_BY_NT(nt, fn, __VA_ARGS__))

static bool ypass_fast(hipStream_t s, const Grid& g, int n, const float2* tw, float2* const* f, int nf, float sign,
                       int nt, int plan) {
  return KW_BY_NT_PENCIL(nt, launch_ypass_fast_t, s, g, n, tw, f, nf, sign, plan);
}

static bool last_fast(hipStream_t s, const Grid& g, int n, const float2* tw, const LastArgs& a, const SpecTables& T,
                      int nt, int plan) {
  return KW_BY_NT_PENCIL(nt, launch_last_fast_t, s, g, n, tw, a, T, plan);
}

static size_t x_smem(const Grid& g, int P, int nbuf, bool slot) {
  size_t s = (size_t)g.Nx * sizeof(float2) + (size_t)nbuf * P * g.Nx * sizeof(float2);
  if (slot) s += (size_t)2 * P * g.Nx * sizeof(int);
  return s;
}

static inline unsigned x_blocks(const Grid& g, int P) { return (unsigned)((g.rows + 2 * P - 1) / (2 * P)); }

void launch_ypass(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, float2* const* f, int nf,
                  float sign, TileLaunch tl) {
  if (ypass_fast(s, g, ax.n, tw, f, nf, sign, tl.nt, tl.plan)) return;
  Fields3 F{{f[0], nf > 1 ? f[1] : nullptr, nf > 2 ? f[2] : nullptr}};
  dim3 grid((g.Nxh + tl.tx - 1) / tl.tx, g.Nz, nf);
  size_t sm = (size_t)g.Ny * sizeof(float2) * (1 + 2 * tl.tx);
  hipLaunchKernelGGL(k_ypass, grid, dim3(tl.threads), sm, s, g, ax, tw, F, sign, tl.tx);
}

void launch_last(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const LastArgs& a,
                 const SpecTables& T, TileLaunch tl) {
  if (last_fast(s, g, ax.n, tw, a, T, tl.nt, tl.plan)) return;
  const int Q = g.dim == 3 ? g.Nxh * g.Ny : g.Nxh;
  dim3 grid((Q + tl.tx - 1) / tl.tx);
  size_t sm = (size_t)ax.n * sizeof(float2) * (1 + 3 * tl.tx);
  hipLaunchKernelGGL(k_last, grid, dim3(tl.threads), sm, s, g, ax, tw, a, T, tl.tx);
}

// ---- batched launches (length-specialised kernels only) ----------------------
void BatchSink::kernel(const void* fn, dim3 grid, dim3 block, size_t smem, void** args) {
  if (err != hipSuccess) return;
  if (!graph) {
    err = hipLaunchKernel(fn, grid, block, args, smem, stream);
    return;
  }
  hipKernelNodeParams p{};
  p.func = const_cast<void*>(fn);
  p.gridDim = grid;
  p.blockDim = block;
  p.sharedMemBytes = (unsigned)smem;
  p.kernelParams = args;
  p.extra = nullptr;
  hipGraphNode_t node = nullptr;
  err = hipGraphAddKernelNode(&node, graph, last ? &last : nullptr, last ? 1 : 0, &p);
  if (err == hipSuccess) last = node;
}

bool launch_x_batch(BatchSink& s, const Grid& g, int which, const void* jobs, const int* live, int nb, bool two,
                    bool slot, XLaunch xl) {
  return KW_BY_NT(xl.nt, launch_x_batch_fast_t, s, g, which, jobs, live, nb, two, slot);
}

bool launch_ypass_batch(BatchSink& s, const Grid& g, int n, const YpassJob* J, const int* live, int nb, int nf,
                        float sign, TileLaunch tl) {
  return KW_BY_NT_PENCIL(tl.nt, launch_ypass_batch_fast_t, s, g, n, J, live, nb, nf, sign, tl.plan);
}

bool launch_last_batch(BatchSink& s, const Grid& g, int n, const LastJob* J, const int* live, int nb, int nout,
                       TileLaunch tl) {
  return KW_BY_NT_PENCIL(tl.nt, launch_last_batch_fast_t, s, g, n, J, live, nb, nout, tl.plan);
}

// Step prologue of a batch: advances each job's time index and rewrites the
// per-step fields of its x-kernel arguments as the single-simulation host
// loop does before each step (source activity, signal row and single-signal
// value; sensor record flags and column; p_final at the last step); live[j]
// is 0 once job j has run its Nt steps.
__global__ void k_batch_clock(BatchClock* __restrict__ clk, XVelJob* __restrict__ vel, XDensJob* __restrict__ dens,
                              XPresJob* __restrict__ pres, int* __restrict__ live, int nb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  BatchClock& c = clk[j];
  const long long t = c.t + 1;
  const int on = t <= c.Nt;
  live[j] = on;
  if (!on) return;
  c.t = t;
  // pointers in the tables are offsets from each table array (PtrEncode)
  auto src = [&](Src& S, int k, const PtrEncode& enc) {
    S.active = c.flag[k] >= (unsigned long long)t;
    S.sig = nullptr;
    S.sval = 0.f;
    if (S.active) {
      const size_t row = (size_t)(t - 1) * c.nseries[k];
      S.sig = c.sig[k] + row;
      S.sval = c.single[k] ? c.sig[k][row] : 0.f;
      enc(S.sig);
    }
  };
  const bool rec = t >= c.start;
  auto sens = [&](Sens& S, bool pressure, const PtrEncode& enc) {
    S.record = rec;
    S.first = rec && t == c.start;
    S.t_rec = rec ? t - c.start : 0;
    if (pressure) {
      S.final_field = t == c.Nt ? c.p_final : nullptr;
      enc(S.final_field);
    }
  };
  const PtrEncode ev{reinterpret_cast<const char*>(vel)}, ed{reinterpret_cast<const char*>(dens)},
      ep{reinterpret_cast<const char*>(pres)};
  for (int k = 0; k < 3; ++k) {
    src(vel[j].A.src[k], k, ev);
    sens(vel[j].A.sens[k], false, ev);
  }
  src(dens[j].A.psrc, 3, ed);
  sens(dens[j].A.ps, true, ed);
  sens(pres[j].ps, true, ep);
}

void launch_batch_clock(BatchSink& s, BatchClock* clk, XVelJob* vel, XDensJob* dens, XPresJob* pres, int* live,
                        int nb) {
  void* args[] = {&clk, &vel, &dens, &pres, &live, &nb};
  s.kernel(reinterpret_cast<const void*>(&k_batch_clock), dim3((nb + 63) / 64), dim3(64), 0, args);
}

void launch_x_velocity(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                       float2* const* G, float* const* u, const Src* src, const Sens* sens, int init,
                       XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_velocity_fast_t, s, g, tw, m, G, u, src, sens, init)) return;
  VelArgs A;
  for (int c = 0; c < 3; ++c) {
    A.G[c] = c < g.dim ? G[c] : nullptr;
    A.u[c] = c < g.dim ? u[c] : nullptr;
    A.src[c] = src[c];
    A.sens[c] = sens[c];
  }
  A.init = init;
  dim3 grid(x_blocks(g, xl.pairs), g.dim);
  hipLaunchKernelGGL(k_x_velocity, grid, dim3(xl.threads), x_smem(g, xl.pairs, 2, true), s, g, ax, tw, m, A,
                     xl.pairs);
}

void launch_x_density(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                      float2* const* G, float* const* rho, const Src& psrc, float* srho_out, float2* pk,
                      const Sens& ps, XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_density_fast_t, s, g, tw, m, G, rho, psrc, srho_out, pk, ps)) return;
  DensArgs A;
  for (int c = 0; c < 3; ++c) {
    A.G[c] = c < g.dim ? G[c] : nullptr;
    A.rho[c] = c < g.dim ? rho[c] : nullptr;
  }
  A.psrc = psrc;
  A.srho = srho_out;
  A.pk = pk;
  A.ps = ps;
  hipLaunchKernelGGL(k_x_density, dim3(x_blocks(g, xl.pairs)), dim3(xl.threads), x_smem(g, xl.pairs, 3, true), s,
                     g, ax, tw, m, A, xl.pairs);
}

void launch_x_pressure(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                       const float2* GA, const float2* GB, const float* srho, float2* pk, const Sens& ps,
                       XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_pressure_fast_t, s, g, tw, m, GA, GB, srho, pk, ps)) return;
  hipLaunchKernelGGL(k_x_pressure, dim3(x_blocks(g, xl.pairs)), dim3(xl.threads), x_smem(g, xl.pairs, 2, false),
                     s, g, ax, tw, m, GA, GB, srho, pk, ps, xl.pairs);
}

void launch_x_init(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                   const float* p0, float* const* rho, float2* pk, const Sens& ps, XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_init_fast_t, s, g, tw, m, p0, rho, pk, ps)) return;
  hipLaunchKernelGGL(k_x_init, dim3(x_blocks(g, xl.pairs)), dim3(xl.threads), x_smem(g, xl.pairs, 2, false), s,
                     g, ax, tw, m, p0, rho[0], rho[1], g.dim == 3 ? rho[2] : nullptr, pk, ps, xl.pairs);
}

void launch_x_source(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Src& src, int unit,
                     float2* out, XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_source_fast_t, s, g, tw, src, unit, out)) return;
  hipLaunchKernelGGL(k_x_source, dim3(x_blocks(g, xl.pairs)), dim3(xl.threads), x_smem(g, xl.pairs, 2, false), s,
                     g, ax, tw, src, unit, out, xl.pairs);
}

void launch_x_c2r(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const float2* in, float* out,
                  XLaunch xl) {
  if (KW_BY_NT(xl.nt, launch_x_c2r_fast_t, s, g, tw, in, out, (const float2*)nullptr)) return;
  hipLaunchKernelGGL(k_x_c2r, dim3(x_blocks(g, xl.pairs)), dim3(xl.threads), x_smem(g, xl.pairs, 2, false), s, g,
                     ax, tw, in, out, xl.pairs);
}

bool launch_x_c2r_mult(hipStream_t s, const Grid& g, const float2* tw, const float2* in, float* out,
                       const float2* mult, XLaunch xl) {
  return KW_BY_NT(xl.nt, launch_x_c2r_fast_t, s, g, tw, in, out, mult);
}

bool launch_x_r2c(hipStream_t s, const Grid& g, const float2* tw, const float* in, float2* out, XLaunch xl) {
  return KW_BY_NT(xl.nt, launch_x_r2c_fast_t, s, g, tw, in, out);
}

void launch_gather(hipStream_t s, const float* field, const long long* lin, int n, float* out) {
  if (n > 0) hipLaunchKernelGGL(k_gather, dim3(nblk(n, 256)), dim3(256), 0, s, field, lin, n, out);
}

void launch_twiddles(hipStream_t s, float2* tw, int n) {
  hipLaunchKernelGGL(k_twiddles, dim3(nblk(n, 256)), dim3(256), 0, s, tw, n);
}

void launch_dvec(hipStream_t s, float2* dpos, float2* dneg, float2* shift, int nfull, int nout, double d) {
  hipLaunchKernelGGL(k_dvec, dim3(nblk(nout, 256)), dim3(256), 0, s, dpos, dneg, shift, nfull, nout, d);
}

void launch_pml(hipStream_t s, float* pml, int n, double d, double dt, double c, int size, double alpha,
                int staggered, int one_sided) {
  hipLaunchKernelGGL(k_pml, dim3(nblk(n, 256)), dim3(256), 0, s, pml, n, d, dt, c, size, alpha, staggered,
                     one_sided);
}

void launch_spectral_tables(hipStream_t s, const Grid& g, const OpParams& p, float* kappa, float* skappa,
                            float* nabla1, float* nabla2) {
  const long long n = (long long)g.Nxs * g.Ny * g.Nz;
  hipLaunchKernelGGL(k_spectral_tables, dim3(nblk(n, 256)), dim3(256), 0, s, g, p, kappa, skappa, nabla1, nabla2);
}

void launch_absorb_coeffs(hipStream_t s, long long n, const float* c0, float c0s, const float* alpha, float alphas,
                          double y, int absorbing, float* tau, float* eta) {
  hipLaunchKernelGGL(k_absorb_coeffs, dim3(nblk(n, 256)), dim3(256), 0, s, n, c0, c0s, alpha, alphas, y, absorbing,
                     tau, eta);
}

void launch_reciprocal(hipStream_t s, long long n, const float* in, float* out) {
  hipLaunchKernelGGL(k_reciprocal, dim3(nblk(n, 256)), dim3(256), 0, s, n, in, out);
}

// plain vector copy (counter calibration: FETCH_SIZE/WRITE_SIZE of a known byte count)
__global__ void k_copy(const float4* __restrict__ in, float4* __restrict__ out, long long n4) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long long)gridDim.x * blockDim.x)
    out[i] = in[i];
}

void launch_copy(hipStream_t s, const void* in, void* out, size_t bytes) {
  const long long n4 = (long long)(bytes / 16);
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, (const float4*)in, (float4*)out, n4);
}

// Streaming copy with U 16-byte loads in flight per thread (all loads of an
// iteration issue before its stores), optionally with non-temporal hints:
// the measured device copy ceiling.
typedef float kw_f4 __attribute__((ext_vector_type(4)));
template <int U, bool NTMP>
__global__ void __launch_bounds__(256) k_copy_u(const kw_f4* __restrict__ in, kw_f4* __restrict__ out, long long n4) {
  const long long stride = (long long)gridDim.x * 256 * U;
  for (long long base = (long long)blockIdx.x * 256 * U + threadIdx.x; base < n4; base += stride) {
    kw_f4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * 256;
      if (i < n4) v[u] = NTMP ? __builtin_nontemporal_load(&in[i]) : in[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long i = base + (long long)u * 256;
      if (i < n4) {
        if (NTMP)
          __builtin_nontemporal_store(v[u], &out[i]);
        else
          out[i] = v[u];
      }
    }
  }
}

bool launch_copy_variant(hipStream_t s, const void* in, void* out, size_t bytes, int variant, int blocks) {
  const long long n4 = (long long)(bytes / 16);
  const kw_f4* i4 = (const kw_f4*)in;
  kw_f4* o4 = (kw_f4*)out;
  dim3 g(blocks > 0 ? blocks : 2048), b(256);
  switch (variant) {
    case 0: hipLaunchKernelGGL(k_copy, g, b, 0, s, (const float4*)in, (float4*)out, n4); return true;
    case 1: hipLaunchKernelGGL((k_copy_u<4, false>), g, b, 0, s, i4, o4, n4); return true;
    case 2: hipLaunchKernelGGL((k_copy_u<8, false>), g, b, 0, s, i4, o4, n4); return true;
    case 3: hipLaunchKernelGGL((k_copy_u<4, true>), g, b, 0, s, i4, o4, n4); return true;
    case 4: hipLaunchKernelGGL((k_copy_u<8, true>), g, b, 0, s, i4, o4, n4); return true;
    default: return false;
  }
}

// zero fill (used instead of hipMemsetAsync)
__global__ void k_zero(kw_f4* __restrict__ p4, long long n4, unsigned char* __restrict__ tail, int ntail) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) p4[i] = kw_f4{0.f, 0.f, 0.f, 0.f};
  if (i < ntail) tail[i] = 0;
}

void launch_zero(hipStream_t s, void* p, size_t bytes) {
  if (!bytes) return;
  // device allocations here are 256-byte aligned; the tail covers sizes not a multiple of 16
  const long long n4 = (long long)(bytes / 16);
  const int ntail = (int)(bytes - (size_t)n4 * 16);
  const long long n = std::max<long long>(n4, ntail);
  hipLaunchKernelGGL(k_zero, dim3(nblk(n, 256)), dim3(256), 0, s, (kw_f4*)p, n4,
                     (unsigned char*)p + (size_t)n4 * 16, ntail);
}

void launch_finish_rms(hipStream_t s, long long n, const double* sum, double count, float* out) {
  hipLaunchKernelGGL(k_finish_rms, dim3(nblk(n, 256)), dim3(256), 0, s, n, sum, count, out);
}

}  // namespace kw
