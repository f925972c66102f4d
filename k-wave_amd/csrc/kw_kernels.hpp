// Argument blocks shared by the HIP kernels (kw_kernels.hip) and the host
// orchestration (kw_sim.cpp).  Plain structs passed by value as kernel
// arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kw_fft.hpp"

namespace kw {

// Grid geometry.  Real fields are C-order (Nz, Ny, Nx): idx = x + Nx*(y + Ny*z).
// x-spectra ("X domain", after the x R2C) are complex (Nz, Ny, Nxh), Nxh = Nx/2+1.
struct Grid {
  int Nx, Ny, Nz, Nxh, dim;
  int rows;  // Ny * Nz
  long long N;
  // x-domain spectra layout: 0 row-major [z][y][kx] (kx fastest),
  // 1 kx-major [kx][z][y] (plane pass, y fastest)
  int kxm;
  // x-row kernels' target number of workgroups per launch (0: 1024 or
  // KW_ROW_WGS): fewer row pairs per workgroup until a launch has this many;
  // a launch-geometry choice of the autotuner, the results do not depend on it
  int row_wgs;
  // row stride of row-major spectra (>= Nxh): padded to a multiple of 8
  // complex values (64 bytes) on 3-D grids whose passes all have
  // length-specialised kernels, so every pencil tile's row segment is whole
  // 64-byte sectors; positions kx >= Nxh are never read or written
  int Nxs;
};

// Scalar-or-field value: field pointer (nullptr for a scalar) + scalar.
struct SF {
  const float* f;
  float s;
  __device__ __forceinline__ float at(long long i) const { return f ? f[i] : s; }
};

struct Medium {
  SF c0, rho0, rsg[3], bona, tau, eta;
  const float* pml[3];     // non-staggered PML, length N_axis
  const float* pml_sg[3];  // staggered PML
  float dt;
  int nonlinear;
  int absorbing;  // 0 lossless, 1 power law, 2 stokes
};

// Source for one quantity (a velocity component, or the pressure).
struct Src {
  int active;        // signal still running at this step
  int mode;          // 0 dirichlet, 1 additive-no-correction, 2 additive
  int many;          // one signal per point
  const float* sig;  // (Nt_src, nseries) for this step's row already offset
  const int* row_ptr;       // CSR over rows (rows+1), source points sorted by index
  const long long* lin;     // 0-based linear index of each source point
  const float* pattern;  // additive, single signal: corrected unit pattern (N)
  const float2* spec;    // additive, many signals: X-domain corrected source
  float sval;            // additive single signal: value at this step
  const int* delay;      // transducer: point j takes sig[delay[j]] (sig offset to this step)
};

// Sensor recording for one quantity.
struct Sens {
  int record;   // record at this step
  int first;    // first recorded step (initialise max/min)
  long long t_rec;
  int ns;
  const int* row_ptr;    // CSR over rows (rows+1), sensor points sorted by index
  const long long* lin;  // 0-based linear index of each sensor point
  float* raw;       // (Nt_rec, ns)
  float* vmax;
  float* vmin;
  double* rms_sum;  // compensated sum of squares (sum, compensation pairs)
  float* max_all;
  float* min_all;
  float* final_field;  // write the field (p_final) when non-null
};

// OP_SHIFT: the half-cell shift exp(-i k d/2) along one axis, no kappa
// (u_non_staggered, extractSensorData.m:39-55)
enum SpecOp { OP_SCALE = 0, OP_DPOS_K = 1, OP_DNEG_K = 2, OP_NABLA1 = 3, OP_NABLA2 = 4, OP_SKAPPA = 5, OP_SHIFT = 6 };

struct SpecOut {
  int src;   // input index
  int op;    // SpecOp
  int axis;  // derivative axis for OP_DPOS_K / OP_DNEG_K
};

struct SpecTables {
  const float2* dpos[3];
  const float2* dneg[3];
  const float2* shift[3];  // exp(-i k d/2) per axis (u_non_staggered only)
  const float* kappa;
  const float* skappa;
  const float* nabla1;
  const float* nabla2;
};

struct LastArgs {
  float2* in[3];
  float2* out[3];
  int nin, nout;
  SpecOut outs[3];
  float scale;
  // last-axis pass, 3-D: outputs 0 and 1 are the x and y derivatives (with
  // kappa) of the same input; their multipliers depend on kx and ky only, so
  // they commute with the pencil's transforms and one workgroup makes both
  // from a single inverse transform of kappa times the spectrum
  int pair01;
};

struct Fields3 {
  float2* f[3];
};

// Arguments of the x-row velocity and density kernels (kw_xrow.inc).
struct VelArgsT {
  float2* G[3];
  float* u[3];
  Src src[3];
  Sens sens[3];
  int init;
};
struct DensArgsT {
  float2* G[3];
  float* rho[3];
  Src psrc;
  float* srho;
  float2* pk;
  Sens ps;
};

// Batched launches: one simulation per blockIdx.z, each kernel's arguments
// in a device table of one entry per simulation (all of one grid shape,
// physics branch and launch configuration).  live[j] == 0 skips job j.
struct XVelJob {
  Grid g;
  const float2* tw;
  Medium m;
  VelArgsT A;
};
struct XDensJob {
  Grid g;
  const float2* tw;
  Medium m;
  DensArgsT A;
};
struct XPresJob {
  Grid g;
  const float2* tw;
  Medium m;
  const float2* GA;
  const float2* GB;
  const float* srho;
  float2* pk;
  Sens ps;
};
struct PlaneJob {
  Grid g;
  const float2* twY;
  const float2* twZ;
  LastArgs A;
  SpecTables T;
};
struct YpassJob {
  Grid g;
  const float2* tw;
  Fields3 F;
};
struct LastJob {
  Grid g;
  const float2* tw;
  LastArgs A;
  SpecTables T;
};

// Pointer fields of the batch job tables hold byte offsets from the start
// of the table array that contains them (null stays null): a batched kernel
// rebases them on its table argument, so the compiler knows they address
// global memory (pointers loaded from memory are otherwise generic and
// every access through them a flat access).  visit_ptrs applies a functor to
// every pointer field of an argument block.
struct PtrEncode {
  const char* base;
  template <typename T>
  __host__ __device__ void operator()(T*& p) const {
    p = p ? (T*)(reinterpret_cast<const char*>(p) - base) : nullptr;
  }
};
struct PtrDecode {
  const char* base;
  template <typename T>
  __host__ __device__ void operator()(T*& p) const {
    p = p ? (T*)(base + (long long)(size_t)p) : nullptr;
  }
};
template <typename F>
__host__ __device__ inline void visit_ptrs(SF& a, const F& f) { f(a.f); }
template <typename F>
__host__ __device__ inline void visit_ptrs(Src& s, const F& f) {
  f(s.sig);
  f(s.row_ptr);
  f(s.lin);
  f(s.pattern);
  f(s.spec);
  f(s.delay);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(Sens& s, const F& f) {
  f(s.row_ptr);
  f(s.lin);
  f(s.raw);
  f(s.vmax);
  f(s.vmin);
  f(s.rms_sum);
  f(s.max_all);
  f(s.min_all);
  f(s.final_field);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(Medium& m, const F& f) {
  visit_ptrs(m.c0, f);
  visit_ptrs(m.rho0, f);
  visit_ptrs(m.bona, f);
  visit_ptrs(m.tau, f);
  visit_ptrs(m.eta, f);
  for (int i = 0; i < 3; ++i) {
    visit_ptrs(m.rsg[i], f);
    f(m.pml[i]);
    f(m.pml_sg[i]);
  }
}
template <typename F>
__host__ __device__ inline void visit_ptrs(VelArgsT& a, const F& f) {
  for (int i = 0; i < 3; ++i) {
    f(a.G[i]);
    f(a.u[i]);
    visit_ptrs(a.src[i], f);
    visit_ptrs(a.sens[i], f);
  }
}
template <typename F>
__host__ __device__ inline void visit_ptrs(DensArgsT& a, const F& f) {
  for (int i = 0; i < 3; ++i) {
    f(a.G[i]);
    f(a.rho[i]);
  }
  visit_ptrs(a.psrc, f);
  f(a.srho);
  f(a.pk);
  visit_ptrs(a.ps, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(LastArgs& a, const F& f) {
  for (int i = 0; i < 3; ++i) {
    f(a.in[i]);
    f(a.out[i]);
  }
}
template <typename F>
__host__ __device__ inline void visit_ptrs(SpecTables& t, const F& f) {
  for (int i = 0; i < 3; ++i) {
    f(t.dpos[i]);
    f(t.dneg[i]);
    f(t.shift[i]);
  }
  f(t.kappa);
  f(t.skappa);
  f(t.nabla1);
  f(t.nabla2);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(Fields3& a, const F& f) {
  for (int i = 0; i < 3; ++i) f(a.f[i]);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(XVelJob& j, const F& f) {
  f(j.tw);
  visit_ptrs(j.m, f);
  visit_ptrs(j.A, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(XDensJob& j, const F& f) {
  f(j.tw);
  visit_ptrs(j.m, f);
  visit_ptrs(j.A, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(XPresJob& j, const F& f) {
  f(j.tw);
  visit_ptrs(j.m, f);
  f(j.GA);
  f(j.GB);
  f(j.srho);
  f(j.pk);
  visit_ptrs(j.ps, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(PlaneJob& j, const F& f) {
  f(j.twY);
  f(j.twZ);
  visit_ptrs(j.A, f);
  visit_ptrs(j.T, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(YpassJob& j, const F& f) {
  f(j.tw);
  visit_ptrs(j.F, f);
}
template <typename F>
__host__ __device__ inline void visit_ptrs(LastJob& j, const F& f) {
  f(j.tw);
  visit_ptrs(j.A, f);
  visit_ptrs(j.T, f);
}

// Where the batched launches go: launched on `stream`, or (graph set)
// appended to an explicitly built hipGraph as a chain of kernel nodes (no
// stream capture, so other threads' HIP calls cannot invalidate it).
struct BatchSink {
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
  void kernel(const void* fn, dim3 grid, dim3 block, size_t smem, void** args);
};

// Per-job time index of a batch and what the step prologue derives the
// per-step kernel arguments from (sources: ux, uy, uz, p).
struct BatchClock {
  long long t;  // steps done; advanced by the prologue
  long long Nt, start;
  const float* sig[4];  // signal bases (flag rows of nseries values)
  unsigned long long flag[4], nseries[4];
  int single[4];  // the step's value is passed as Src::sval
  float* p_final;
};

// Host-side launch wrappers (kw_kernels.hip).
struct XLaunch {
  int pairs;    // packed row pairs per workgroup (generic kernels)
  int threads;  // workgroup size (generic kernels)
  int nt;       // workgroup size of the length-specialised kernels (64/128/256; 0 = generic only)
};
struct TileLaunch {
  int tx;       // pencils per workgroup (generic kernels)
  int threads;
  int nt;       // workgroup size of the length-specialised kernels (64/128/256/512; 0 = generic only)
  int plan;     // radix plan of the length-specialised kernels (kwt::next_radix; 1 only for lengths >= 160)
};

// Axis lengths with compiled length-specialised kernels (kw_pencil_nt*.hip,
// kw_xrow_nt*.hip, instantiated for 64-, 128- and 256-thread workgroups; the
// pencils also for 512).
#define KW_FAST_SIZES(X) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(384) X(512) X(768)
bool fast_length(int n);

// plane pass (kw_plane.hip): LDS bytes for a (Ny, Nz) plane, 0 if none
size_t plane_lds_bytes(int ny, int nz);
// its workgroup size when none is requested (KW_PLANE_NT or the per-size default)
int default_plane_nt(int ny, int nz);
// nt: workgroup size 256, 512 or 1024; 0 picks the per-size default
bool launch_plane(hipStream_t s, const Grid& g, const float2* twY, const float2* twZ, const LastArgs& a,
                  const SpecTables& T, int nt = 0);

bool launch_plane_batch(BatchSink& s, const Grid& g, const PlaneJob* J, const int* live, int nb, int nout,
                        int nt = 0);

template <int NT>
bool launch_ypass_fast_t(hipStream_t s, const Grid& g, int n, const float2* tw, float2* const* f, int nf, float sign,
                         int plan);
template <int NT>
bool launch_last_fast_t(hipStream_t s, const Grid& g, int n, const float2* tw, const LastArgs& a, const SpecTables& T,
                        int plan);
template <int NT>
bool launch_ypass_batch_fast_t(BatchSink& s, const Grid& g, int n, const YpassJob* J, const int* live, int nb,
                               int nf, float sign, int plan);
template <int NT>
bool launch_last_batch_fast_t(BatchSink& s, const Grid& g, int n, const LastJob* J, const int* live, int nb, int nout,
                              int plan);
template <int NT>
bool launch_x_velocity_fast_t(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G,
                              float* const* u, const Src* src, const Sens* sens, int init);
template <int NT>
bool launch_x_density_fast_t(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, float2* const* G,
                             float* const* rho, const Src& psrc, float* srho, float2* pk, const Sens& ps);
template <int NT>
bool launch_x_pressure_fast_t(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float2* GA,
                              const float2* GB, const float* srho, float2* pk, const Sens& ps);
template <int NT>
bool launch_x_init_fast_t(hipStream_t s, const Grid& g, const float2* tw, const Medium& m, const float* p0,
                          float* const* rho, float2* pk, const Sens& ps);
template <int NT>
bool launch_x_source_fast_t(hipStream_t s, const Grid& g, const float2* tw, const Src& src, int unit, float2* out);
template <int NT>
bool launch_x_c2r_fast_t(hipStream_t s, const Grid& g, const float2* tw, const float2* in, float* out,
                         const float2* mult);
template <int NT>
bool launch_x_batch_fast_t(BatchSink& s, const Grid& g, int which, const void* jobs, const int* live, int nb,
                           bool two, bool slot);
template <int NT>
bool launch_x_r2c_fast_t(hipStream_t s, const Grid& g, const float2* tw, const float* in, float2* out);

void launch_ypass(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, float2* const* f,
                  int nf, float sign, TileLaunch tl);
void launch_last(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const LastArgs& a,
                 const SpecTables& T, TileLaunch tl);
void launch_x_velocity(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                       float2* const* G, float* const* u, const Src* src, const Sens* sens, int init,
                       XLaunch xl);
void launch_x_density(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                      float2* const* G, float* const* rho, const Src& psrc, float* srho_out, float2* pk,
                      const Sens& ps, XLaunch xl);
void launch_x_pressure(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                       const float2* GA, const float2* GB, const float* srho, float2* pk, const Sens& ps,
                       XLaunch xl);
void launch_x_init(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Medium& m,
                   const float* p0, float* const* rho, float2* pk, const Sens& ps, XLaunch xl);
void launch_x_source(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const Src& src,
                     int unit, float2* out, XLaunch xl);
void launch_x_c2r(hipStream_t s, const Grid& g, const FftAxis& ax, const float2* tw, const float2* in,
                  float* out, XLaunch xl);
// x C2R of (mult[kx] .* in) into a real field, and x R2C of a real field
// (specialised row lengths only; false when Nx has no compiled kernels)
bool launch_x_c2r_mult(hipStream_t s, const Grid& g, const float2* tw, const float2* in, float* out,
                       const float2* mult, XLaunch xl);
bool launch_x_r2c(hipStream_t s, const Grid& g, const float2* tw, const float* in, float2* out, XLaunch xl);
// out[j] = field[lin[j]], j < n
void launch_gather(hipStream_t s, const float* field, const long long* lin, int n, float* out);

// batched forms (one simulation per blockIdx.z; false: no length-specialised
// kernel for this launch configuration).  which: 0 velocity, 1 density, 2
// absorbing EOS.
bool launch_x_batch(BatchSink& s, const Grid& g, int which, const void* jobs, const int* live, int nb, bool two,
                    bool slot, XLaunch xl);
bool launch_ypass_batch(BatchSink& s, const Grid& g, int n, const YpassJob* J, const int* live, int nb, int nf,
                        float sign, TileLaunch tl);
bool launch_last_batch(BatchSink& s, const Grid& g, int n, const LastJob* J, const int* live, int nb, int nout,
                       TileLaunch tl);
void launch_batch_clock(BatchSink& s, BatchClock* clk, XVelJob* vel, XDensJob* dens, XPresJob* pres, int* live,
                        int nb);

// Axisymmetric radial pass (kw_axisym.hip): tables at the odd frequencies
// 2m+1 (m < 2 Ny) of the 4 Ny radially expanded grid
struct AsTables {
  const float* kappa;    // [m][kx]
  const float* skappa;   // [m][kx], additive sources only
  const float2* ydpos;   // i ky exp(+i ky dy/2)
  const float2* ydneg;   // i ky exp(-i ky dy/2)
  const float2* yshn;    // exp(-i ky dy/2)
  const float* ysg_inv;  // 1 / y_vec_sg, Ny values
  const float2* tw4;     // exp(-2 pi i j / 4Ny)
  const float2* tw2;     // FFT twiddles of length 2 Ny
  const float2* dposx;   // x operators of the plain loop (half spectrum)
  const float2* dnegx;
};
void launch_as_tables(hipStream_t s, const Grid& g, double dx, double dy, double dt, double c_ref, const AsTables& T);
// pencils per workgroup for the radial pass (0 if 2 Ny does not fit lds_max); *lds = bytes
int as_radial_tile(int Ny, size_t lds_max, size_t* lds);
// out_o = the radial-spectral operator of a.outs[o] on the x spectra a.in (OP_DPOS_K,
// OP_DNEG_K on axis 0 / 1, OP_SKAPPA with axis 1 for the uy source), normalised
void launch_as_radial(hipStream_t s, const Grid& g, const FftAxis& ax, const AsTables& T, const LastArgs& a, int tx,
                      size_t lds);

// FP64 precompute
struct OpParams {
  double dx[3];
  double dt, c_ref;
  double alpha_power;
  int absorbing;
};
void launch_twiddles(hipStream_t s, float2* tw, int n);
void launch_dvec(hipStream_t s, float2* dpos, float2* dneg, float2* shift, int nfull, int nout, double d);
// one_sided: the layer at the far end only (the radial axis of axisymmetric grids)
void launch_pml(hipStream_t s, float* pml, int n, double d, double dt, double c, int size, double alpha,
                int staggered, int one_sided = 0);
void launch_spectral_tables(hipStream_t s, const Grid& g, const OpParams& p, float* kappa, float* skappa,
                            float* nabla1, float* nabla2);
void launch_absorb_coeffs(hipStream_t s, long long n, const float* c0, float c0s, const float* alpha,
                          float alphas, double y, int absorbing, float* tau, float* eta);
void launch_reciprocal(hipStream_t s, long long n, const float* in, float* out);
void launch_finish_rms(hipStream_t s, long long n, const double* sum, double count, float* out);
void launch_copy(hipStream_t s, const void* in, void* out, size_t bytes);
// variant 0: launch_copy's kernel; 1/2: 4/8 loads in flight per thread; 3/4: the same
// with non-temporal loads and stores.  blocks: grid size (256-thread workgroups)
bool launch_copy_variant(hipStream_t s, const void* in, void* out, size_t bytes, int variant, int blocks);
void launch_zero(hipStream_t s, void* p, size_t bytes);

}  // namespace kw
