// Axisymmetric radial pass (kspaceFirstOrderAS.m with WSWA symmetry, the one
// the input file selects: kspaceFirstOrder_inputChecking.m:1400-1411).
//
// The grid is 2-D with x axial and y radial (y index 0 on the axis).  The
// x-row kernels run unchanged on it; this pass replaces the y transform of
// the plain 2-D loop.  Per x-spectrum column (one kx) every field is mirrored
// to a 4 Ny radial period with the symmetry of its grid (AS.m:713-717 for p
// and ux, 953-962 for uy and uy / y, 884-888 for the uy source) and
// transformed on the expanded grid kWaveGrid(Nx, dx, 4 Ny, dy) (AS.m:488-521).
// Every mirrored sequence e satisfies e[j + 2Ny] = -e[j], so only the odd
// frequencies k = 2m + 1 of the 4 Ny grid are non-zero, and
//   E[2m+1] = 2 sum_{j<2Ny} (e[j] w^j) W_2Ny^(jm),   w = exp(-2 pi i / 4Ny):
// a 2 Ny FFT of the pre-twiddled first half.  The inverse is the conjugate
// post-twiddle of a 2 Ny inverse FFT; the first Ny outputs are kept (the
// trim of AS.m:729-730).  The operators are evaluated at those odd
// frequencies: kappa on the expanded grid, ddy_k * y_shift_pos for dp/dy,
// kappa * y_shift_neg * (ddy_k * F(uy) + F(uy / y_sg)) for the radial
// divergence (AS.m:982-986), and the x operators of the plain loop.
#include <algorithm>
#include <cmath>

#include "kw_kernels.hpp"

namespace kw {

namespace {

// FFT-order wavenumber of bin j of an n-point axis (ifftshift of makeDim)
__device__ __forceinline__ double as_kvec(int j, int n, double d) {
  const int pos = n - n / 2;
  const double nx = (j < pos) ? (double)j / n : (double)(j - n) / n;
  return 2.0 * M_PI / d * nx;
}

template <typename T>
__device__ __forceinline__ T sel3(int i, T a, T b, T c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// FP64 operator tables at the odd frequencies 2m+1 of the 4 Ny grid:
// kappa / source_kappa [m][kx] (AS.m:514-516), i ky exp(+-i ky dy/2),
// exp(-i ky dy/2) (AS.m:508-510), 1 / y_vec_sg (AS.m:571).
__global__ void k_as_tables(int Nx, int Nxh, int Ny, double dx, double dy, double dt, double c_ref, float* kappa,
                            float* skappa, float2* ydpos, float2* ydneg, float2* yshn, float* ysg_inv) {
  const int L = 2 * Ny;
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id < (long long)Nxh * L) {
    const int m = (int)(id / Nxh);
    const int kx = (int)(id - (long long)m * Nxh);
    const double kxv = as_kvec(kx, Nx, dx);
    const double kyv = as_kvec(2 * m + 1, 4 * Ny, dy);
    const double a = c_ref * sqrt(kxv * kxv + kyv * kyv) * dt / 2.0;
    kappa[id] = (float)(a == 0.0 ? 1.0 : sin(a) / a);
    if (skappa) skappa[id] = (float)cos(a);
  }
  if (id < L) {
    const double k = as_kvec(2 * (int)id + 1, 4 * Ny, dy);
    double s, c;
    sincos(k * dy / 2.0, &s, &c);
    ydpos[id] = make_float2((float)(-k * s), (float)(k * c));
    ydneg[id] = make_float2((float)(k * s), (float)(k * c));
    yshn[id] = make_float2((float)c, (float)(-s));
  }
  if (id < Ny) ysg_inv[id] = (float)(1.0 / (((double)id + 0.5) * dy));
}

// Mirror kinds of the radial sequence: WS-WA (p, ux, p and ux sources),
// HA-HS (uy, uy source), HS-HA (uy / y_sg, only inside the radial divergence).
enum { MIR_WSWA = 0, MIR_HAHS = 1, MIR_HSHA = 2 };

// One workgroup per (tile of TX consecutive kx, output).  LDS: three
// [2Ny][TX] complex buffers (two for the ping-pong FFT, a third for the second
// forward transform of the radial divergence).
__global__ void __launch_bounds__(256) k_as_radial(Grid g, FftAxis ax, AsTables T, LastArgs A, int TX, float scale) {
  extern __shared__ float2 smem[];
  const int Ny = g.Ny, L = 2 * Ny, Nxh = g.Nxh;
  const int o = blockIdx.y;
  const int kx0 = blockIdx.x * TX;
  const SpecOut so = sel3(o, A.outs[0], A.outs[1], A.outs[2]);
  const float2* in = sel3(so.src, (const float2*)A.in[0], (const float2*)A.in[1], (const float2*)A.in[2]);
  float2* out = sel3(o, A.out[0], A.out[1], A.out[2]);
  const bool divy = so.op == OP_DNEG_K && so.axis == 1;
  const int kind = (divy || (so.op == OP_SKAPPA && so.axis == 1)) ? MIR_HAHS : MIR_WSWA;
  const int n = L * TX;
  float2* b0 = smem;
  float2* b1 = smem + n;
  float2* b2 = smem + 2 * n;

  auto load = [&](float2* dst, int k) {
    for (int w = threadIdx.x; w < n; w += blockDim.x) {
      const int t = w % TX, j = w / TX;
      const int kx = kx0 + t;
      float2 v = make_float2(0.f, 0.f);
      int y;
      float sg = 1.f;
      if (j < Ny) {
        y = j;
      } else if (k == MIR_WSWA) {
        y = j == Ny ? -1 : 2 * Ny - j;
        sg = -1.f;
      } else {
        y = 2 * Ny - 1 - j;
        sg = k == MIR_HAHS ? 1.f : -1.f;
      }
      if (kx < Nxh && y >= 0) {
        v = in[(long long)y * Nxh + kx];
        const float f = k == MIR_HSHA ? sg * T.ysg_inv[y] : sg;
        v = cscale(v, f);
      }
      dst[j * TX + t] = cmul(v, T.tw4[j]);
    }
  };

  load(b0, kind);
  __syncthreads();
  float2* F1 = lds_fft<true>(b0, b1, T.tw2, ax, TX, 0, -1.f);
  float2* fr = F1 == b0 ? b1 : b0;
  float2* F2 = nullptr;
  if (divy) {
    load(b2, MIR_HSHA);
    __syncthreads();
    F2 = lds_fft<true>(b2, fr, T.tw2, ax, TX, 0, -1.f);
    if (F2 == fr) fr = b2;
  }
  for (int w = threadIdx.x; w < n; w += blockDim.x) {
    const int t = w % TX, m = w / TX;
    const int kx = kx0 + t;
    float2 v = F1[w];
    if (kx < Nxh) {
      const long long ti = (long long)m * Nxh + kx;
      if (so.op == OP_SKAPPA) {
        v = cscale(v, T.skappa[ti]);
      } else if (so.op == OP_DPOS_K) {
        v = cscale(cmul(so.axis == 0 ? T.dposx[kx] : T.ydpos[m], v), T.kappa[ti]);
      } else if (so.op == OP_DNEG_K && so.axis == 0) {
        v = cscale(cmul(T.dnegx[kx], v), T.kappa[ti]);
      } else if (divy) {
        v = cscale(cadd(cmul(T.ydneg[m], v), cmul(T.yshn[m], F2[w])), T.kappa[ti]);
      }
    } else {
      v = make_float2(0.f, 0.f);
    }
    F1[w] = v;
  }
  __syncthreads();
  const float2* R = lds_fft<true>(F1, fr, T.tw2, ax, TX, 0, 1.f);
  for (int w = threadIdx.x; w < Ny * TX; w += blockDim.x) {
    const int t = w % TX, j = w / TX;
    const int kx = kx0 + t;
    if (kx < Nxh) out[(long long)j * Nxh + kx] = cscale(cmul(cconj(T.tw4[j]), R[w]), scale);
  }
}

int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

void launch_as_tables(hipStream_t s, const Grid& g, double dx, double dy, double dt, double c_ref, const AsTables& T) {
  const long long n = std::max<long long>((long long)g.Nxh * 2 * g.Ny, 2LL * g.Ny);
  hipLaunchKernelGGL(k_as_tables, dim3(nblk(n, 256)), dim3(256), 0, s, g.Nx, g.Nxh, g.Ny, dx, dy, dt, c_ref,
                     const_cast<float*>(T.kappa), const_cast<float*>(T.skappa), const_cast<float2*>(T.ydpos),
                     const_cast<float2*>(T.ydneg), const_cast<float2*>(T.yshn), const_cast<float*>(T.ysg_inv));
}

int as_radial_tile(int Ny, size_t lds_max, size_t* lds) {
  int tx = 16;
  while (tx > 1 && 3ull * 2 * Ny * tx * sizeof(float2) > 65536) tx /= 2;
  const size_t b = 3ull * 2 * Ny * tx * sizeof(float2);
  if (b > lds_max) return 0;
  *lds = b;
  return tx;
}

void launch_as_radial(hipStream_t s, const Grid& g, const FftAxis& ax, const AsTables& T, const LastArgs& a, int tx,
                      size_t lds) {
  if (lds > 65536) {
    // once per process; a function-local static initialiser is thread-safe
    static const bool attr = ((void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_as_radial),
                                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024),
                              true);
    (void)attr;
  }
  // normalised inverse of the expanded transform: 2 / (4 Ny) from the odd-
  // frequency fold, 1 / Nx for the unnormalised x C2R that follows
  const float scale = (float)(1.0 / ((double)g.Nx * 2.0 * g.Ny));
  dim3 grid((g.Nxh + tx - 1) / tx, a.nout);
  hipLaunchKernelGGL(k_as_radial, grid, dim3(256), lds, s, g, ax, T, a, tx, scale);
}

}  // namespace kw
