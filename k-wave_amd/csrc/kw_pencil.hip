// Length-specialised pencil passes (gfx950): the y pass and the last-axis
// forward-FFT * spectral operator * inverse-FFT pass, for axis lengths with a
// compiled instantiation.  Other lengths use the generic kernels of
// kw_kernels.hip.  Pencils are tiled over a flattened column index q so every
// tile but the last is full; the spectral multiply is fused into the first
// stage of each inverse transform, so the forward spectrum stays in LDS and
// is reused by every output that derives from the same input.
#include "kw_fft_t.hpp"
#include "kw_kernels.hpp"

namespace kw {

template <int N>
struct PencilCfg {
  static constexpr int ST = kwt::num_stages(N);
  static constexpr int NBUF = ST >= 3 ? 3 : 2;  // last-axis LDS buffers
  static constexpr int tx() {
    int t = 32;
    while (t > 1 && (long)NBUF * t * kwt::Pad<N>::P * 8 > 65536) t /= 2;
    return t > 16 ? 16 : t;
  }
  static constexpr int TX = tx();
  static constexpr int NT = 256;
  static_assert(NT % TX == 0, "pencil count must divide the thread count");
};

template <typename T>
__device__ __forceinline__ T sel3(int i, T a, T b, T c) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// y pass: columns q = kx + Nxh * z (flattened), pencils along y (stride Nxh).
template <int N>
__global__ void __launch_bounds__(256) k_ypass_t(Grid g, const float2* __restrict__ tw, Fields3 F, float sign) {
  using C = PencilCfg<N>;
  using L = kwt::PencilLds<N, C::TX>;
  extern __shared__ float2 smem[];
  L A{smem}, B{smem + L::size};
  float2* f = sel3((int)blockIdx.y, F.f[0], F.f[1], F.f[2]);
  const int tile = kwt::xcd_swizzle(blockIdx.x, gridDim.x);
  const int q = tile * C::TX + (int)threadIdx.x % C::TX;
  const bool valid = q < g.Nxh * g.Nz;
  long long off = 0;
  if (valid) {
    const int z = q / g.Nxh;
    off = (long long)z * g.Ny * g.Nxh + (q - z * g.Nxh);
  }
  const long long stride = g.Nxh;
  auto ld = [&](int, int a) { return valid ? f[off + a * stride] : make_float2(0.f, 0.f); };
  auto st = [&](int, int a, float2 v, int) {
    if (valid) f[off + a * stride] = v;
  };
  kwt::run<N, C::TX, C::NT, true>(tw, sign, C::TX, ld, st, A, B);
}

// last axis (z in 3-D, y in 2-D): columns q over the contiguous (kx[, y]) plane.
// One output per workgroup (the logical block index interleaves outputs of
// the same tile, which the XCD remap keeps on one XCD, so an input shared by
// several outputs is re-read from L2): forward FFT of the output's source into
// LDS, then the inverse FFT whose first stage applies the spectral operator.
template <int N>
__global__ void __launch_bounds__(256) k_last_t(Grid g, const float2* __restrict__ tw, LastArgs A, SpecTables T) {
  using C = PencilCfg<N>;
  using L = kwt::PencilLds<N, C::TX>;
  extern __shared__ float2 smem[];
  L S{smem}, W1{smem + L::size}, W2{smem + (C::NBUF == 3 ? 2 : 1) * L::size};
  const int Q = g.dim == 3 ? g.Nxh * g.Ny : g.Nxh;
  const int logical = kwt::xcd_swizzle(blockIdx.x, gridDim.x);
  const int o = logical % A.nout;
  const int tile = logical / A.nout;
  const int q = tile * C::TX + (int)threadIdx.x % C::TX;
  const bool valid = q < Q;
  int kx = q, ky = 0;
  if (g.dim == 3) {
    ky = q / g.Nxh;
    kx = q - ky * g.Nxh;
  }
  const SpecOut so = sel3(o, A.outs[0], A.outs[1], A.outs[2]);
  const float2* in = sel3(so.src, (const float2*)A.in[0], (const float2*)A.in[1], (const float2*)A.in[2]);
  auto ld = [&](int, int a) { return valid ? in[(long long)a * Q + q] : make_float2(0.f, 0.f); };
  auto st = [&](int t, int a, float2 v, int) { S(t, a) = v; };
  kwt::run<N, C::TX, C::NT, true>(tw, -1.f, C::TX, ld, st, W1, W2);
  __syncthreads();

  const float2* dv = nullptr;
  int kk = 0;
  if (so.op == OP_DPOS_K || so.op == OP_DNEG_K) {
    dv = so.op == OP_DPOS_K ? sel3(so.axis, T.dpos[0], T.dpos[1], T.dpos[2])
                            : sel3(so.axis, T.dneg[0], T.dneg[1], T.dneg[2]);
    // the axis along the pencil (z in 3-D, y in 2-D) is indexed by the position a
    kk = so.axis == 0 ? kx : ((g.dim == 3 && so.axis == 1) ? ky : -1);
  }
  const float* tab = so.op == OP_NABLA1 ? T.nabla1
                     : (so.op == OP_NABLA2 ? T.nabla2 : (so.op == OP_SKAPPA ? T.skappa : T.kappa));
  const float scale = A.scale;
  const int op = so.op;
  auto ldm = [&](int t, int a) {
    float2 v = S(t, a);
    if (!valid) return make_float2(0.f, 0.f);
    const long long sidx = (long long)a * Q + q;
    if (op == OP_DPOS_K || op == OP_DNEG_K) {
      const float2 d = dv[kk >= 0 ? kk : a];
      v = cscale(cmul(d, v), tab[sidx] * scale);
    } else if (op == OP_SCALE) {
      v = cscale(v, scale);
    } else {
      v = cscale(v, tab[sidx] * scale);
    }
    return v;
  };
  float2* out = sel3(o, A.out[0], A.out[1], A.out[2]);
  auto sto = [&](int, int a, float2 v, int) {
    if (valid) out[(long long)a * Q + q] = v;
  };
  kwt::run<N, C::TX, C::NT, true>(tw, 1.f, C::TX, ldm, sto, W1, W2);
}

#define KW_PENCIL_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(384) X(512) X(768)

bool pencil_supported(int n) {
  switch (n) {
#define KW_CASE(n) case n:
    KW_PENCIL_SIZES(KW_CASE)
#undef KW_CASE
    return true;
    default:
      return false;
  }
}

template <int N>
static void ypass_t(hipStream_t s, const Grid& g, const float2* tw, const Fields3& F, int nf, float sign) {
  using C = PencilCfg<N>;
  const int cols = g.Nxh * g.Nz;
  dim3 grid((cols + C::TX - 1) / C::TX, nf);
  const size_t sm = (size_t)(C::ST >= 3 ? 2 : (C::ST == 2 ? 1 : 0)) * kwt::PencilLds<N, C::TX>::size * sizeof(float2);
  hipLaunchKernelGGL(k_ypass_t<N>, grid, dim3(C::NT), sm, s, g, tw, F, sign);
}

template <int N>
static void last_t(hipStream_t s, const Grid& g, const float2* tw, const LastArgs& a, const SpecTables& T) {
  using C = PencilCfg<N>;
  const int Q = g.dim == 3 ? g.Nxh * g.Ny : g.Nxh;
  const size_t sm = (size_t)C::NBUF * kwt::PencilLds<N, C::TX>::size * sizeof(float2);
  hipLaunchKernelGGL(k_last_t<N>, dim3(((Q + C::TX - 1) / C::TX) * a.nout), dim3(C::NT), sm, s, g, tw, a, T);
}

bool launch_ypass_fast(hipStream_t s, const Grid& g, int n, const float2* tw, float2* const* f, int nf, float sign) {
  Fields3 F{{f[0], nf > 1 ? f[1] : f[0], nf > 2 ? f[2] : f[0]}};
  switch (n) {
#define KW_CASE(n)                         \
  case n:                                  \
    ypass_t<n>(s, g, tw, F, nf, sign);     \
    return true;
    KW_PENCIL_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

bool launch_last_fast(hipStream_t s, const Grid& g, int n, const float2* tw, const LastArgs& a, const SpecTables& T) {
  switch (n) {
#define KW_CASE(n)               \
  case n:                        \
    last_t<n>(s, g, tw, a, T);   \
    return true;
    KW_PENCIL_SIZES(KW_CASE)
#undef KW_CASE
    default:
      return false;
  }
}

}  // namespace kw
