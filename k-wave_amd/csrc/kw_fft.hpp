// Mixed-radix (2/3/4/5/7/8) Stockham FFT primitives that run on tiles held
// in LDS.  Used by every transform pass of the time loop: the x-row kernels
// (real rows packed two to a complex row), the y pass and the last-axis
// spectral kernel.
//
// Stage s with radix R and stride Ns (product of the earlier radices), for a
// length-n transform, butterfly i in [0, n/R):
//   k = i % Ns,  v[r] = x[i + r*n/R] * W_n^(sign * r*k*n/(Ns*R)),
//   V = DFT_R(v),  y[(i/Ns)*Ns*R + k + r*Ns] = V[r]
// which leaves the output in natural order (autosort).  Twiddles come from a
// table W_n^m = exp(-2*pi*i*m/n), m in [0, n), built in double precision.
#pragma once
#include <hip/hip_runtime.h>

#define KW_MAX_STAGES 16

struct FftAxis {
  int n;                       // transform length
  int nstages;
  int radix[KW_MAX_STAGES];    // product == n, each in {2,3,4,5,7,8}
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by (sign * i)
__device__ __forceinline__ float2 cmul_si(float2 a, float sign) { return make_float2(-sign * a.y, sign * a.x); }

// ---- radix butterflies (sign = -1 forward, +1 inverse) --------------------

__device__ __forceinline__ void bf2(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

__device__ __forceinline__ void bf3(float2* v, float sign) {
  const float h = 0.86602540378443864676f;  // sqrt(3)/2
  float2 s = cadd(v[1], v[2]);
  float2 d = csub(v[1], v[2]);
  float2 t1 = make_float2(v[0].x - 0.5f * s.x, v[0].y - 0.5f * s.y);
  float2 t2 = cmul_si(cscale(d, h), sign);
  v[0] = cadd(v[0], s);
  v[1] = cadd(t1, t2);
  v[2] = csub(t1, t2);
}

__device__ __forceinline__ void bf4(float2* v, float sign) {
  float2 a = cadd(v[0], v[2]);
  float2 b = csub(v[0], v[2]);
  float2 c = cadd(v[1], v[3]);
  float2 d = cmul_si(csub(v[1], v[3]), sign);
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

__device__ __forceinline__ void bf8(float2* v, float sign) {
  const float r = 0.70710678118654752440f;  // 1/sqrt(2)
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  bf4(e, sign);
  bf4(o, sign);
  // W8^q = exp(sign*2*pi*i*q/8)
  float2 o1 = cmul(o[1], make_float2(r, sign * r));
  float2 o2 = cmul_si(o[2], sign);
  float2 o3 = cmul(o[3], make_float2(-r, sign * r));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// Odd prime radix R (5 or 7) from the symmetric-pair form of the DFT:
// V_q = v0 + sum_{m=1}^{(R-1)/2} [ (v_m + v_{R-m}) cos(2 pi q m / R)
//                                 + sign*i (v_m - v_{R-m}) sin(2 pi q m / R) ]
template <int R>
__device__ __forceinline__ void bf_odd(float2* v, float sign, const float* c, const float* s) {
  constexpr int H = (R - 1) / 2;
  float2 sp[H], sm[H];
#pragma unroll
  for (int m = 1; m <= H; ++m) {
    sp[m - 1] = cadd(v[m], v[R - m]);
    sm[m - 1] = csub(v[m], v[R - m]);
  }
  float2 v0 = v[0];
  float2 dc = v0;
#pragma unroll
  for (int m = 0; m < H; ++m) dc = cadd(dc, sp[m]);
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    float2 re = v0, im = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      const int qm = (q * m) % R;
      re = make_float2(re.x + c[qm] * sp[m - 1].x, re.y + c[qm] * sp[m - 1].y);
      im = make_float2(im.x + s[qm] * sm[m - 1].x, im.y + s[qm] * sm[m - 1].y);
    }
    float2 t = cmul_si(im, sign);
    v[q] = cadd(re, t);
    v[R - q] = csub(re, t);
  }
  v[0] = dc;
}

__device__ __forceinline__ void bf5(float2* v, float sign) {
  // cos/sin(2 pi j / 5), j = 0..4
  const float c[5] = {1.0f, 0.30901699437494742410f, -0.80901699437494742410f,
                      -0.80901699437494742410f, 0.30901699437494742410f};
  const float s[5] = {0.0f, 0.95105651629515357212f, 0.58778525229247312917f,
                      -0.58778525229247312917f, -0.95105651629515357212f};
  bf_odd<5>(v, sign, c, s);
}

__device__ __forceinline__ void bf7(float2* v, float sign) {
  // cos/sin(2 pi j / 7), j = 0..6
  const float c[7] = {1.0f, 0.62348980185873353053f, -0.22252093395631440429f,
                      -0.90096886790241912624f, -0.90096886790241912624f,
                      -0.22252093395631440429f, 0.62348980185873353053f};
  const float s[7] = {0.0f, 0.78183148246802980871f, 0.97492791218182360702f,
                      0.43388373911755812048f, -0.43388373911755812048f,
                      -0.97492791218182360702f, -0.78183148246802980871f};
  bf_odd<7>(v, sign, c, s);
}

__device__ __forceinline__ void butterfly(float2* v, int R, float sign) {
  switch (R) {
    case 2: bf2(v); break;
    case 3: bf3(v, sign); break;
    case 4: bf4(v, sign); break;
    case 5: bf5(v, sign); break;
    case 7: bf7(v, sign); break;
    default: bf8(v, sign); break;
  }
}

// Runs `cnt` length-ax.n transforms on LDS data, ping-ponging between src and
// dst.  Element a of transform t is at
//   TFAST ? (a * cnt + t)           (transform index fastest: y/z tiles)
//         : (t * tstride + a)       (position fastest: x rows)
// `tw` is the twiddle table (LDS or global).  sign = -1 forward, +1 inverse.
// Returns the buffer holding the result.  Ends with a barrier.
template <bool TFAST>
__device__ float2* lds_fft(float2* src, float2* dst, const float2* tw, const FftAxis& ax,
                           int cnt, int tstride, float sign) {
  const int n = ax.n;
  int Ns = 1;
  for (int st = 0; st < ax.nstages; ++st) {
    const int R = ax.radix[st];
    const int nb = n / R;
    const int total = cnt * nb;
    const int twstep = n / (Ns * R);
    for (int w = threadIdx.x; w < total; w += blockDim.x) {
      int t, i;
      if (TFAST) {
        t = w % cnt;
        i = w / cnt;
      } else {
        t = w / nb;
        i = w - t * nb;
      }
      const int k = i % Ns;
      float2 v[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (r < R) {
          const int a = i + r * nb;
          v[r] = TFAST ? src[a * cnt + t] : src[t * tstride + a];
        }
      }
      if (k != 0) {
#pragma unroll
        for (int r = 1; r < 8; ++r) {
          if (r < R) {
            float2 wv = tw[r * k * twstep];
            wv.y *= -sign;  // table holds exp(-i...); sign=+1 conjugates
            v[r] = cmul(v[r], wv);
          }
        }
      }
      butterfly(v, R, sign);
      const int j = (i / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        if (r < R) {
          const int a = j + r * Ns;
          if (TFAST)
            dst[a * cnt + t] = v[r];
          else
            dst[t * tstride + a] = v[r];
        }
      }
    }
    __syncthreads();
    float2* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= R;
  }
  return src;
}
