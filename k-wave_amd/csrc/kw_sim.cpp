// Host orchestration behind the C ABI of include/kwave_hip.h: validates the
// input-file dataset set, builds FFT plans and launch geometry, runs the
// FP64 operator precompute on the device, and sequences the per-step passes
// of kw_kernels.hip.  Statement references are to k-Wave/kspaceFirstOrder3D.m.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <utility>
#include <string>
#include <vector>

#include "../../include/kwave_hip.h"
#include "kw_kernels.hpp"

using namespace kw;

namespace {

thread_local std::string g_err;

kw_status fail(kw_status st, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return st;
}

}  // namespace

// shared with the serving loop (kw_server.cpp), which reports through the
// same thread-local message
namespace kw {
void set_last_error(const char* msg) { g_err = msg ? msg : ""; }
}  // namespace kw

namespace {

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIPCHK(x)                                   \
  do {                                              \
    hipError_t _e = (x);                            \
    if (_e != hipSuccess) throw HipError{_e, #x};   \
  } while (0)

struct OomError {
  size_t bytes;
};
struct ArenaFull {
  size_t need, cap;
};

constexpr size_t kAllocAlign = 256;
constexpr size_t alloc_round(size_t bytes) {
  return ((bytes > 16 ? bytes : 16) + kAllocAlign - 1) / kAllocAlign * kAllocAlign;
}

// Factor n into radices {8,4,2,3,5,7}; false if another prime divides n.
bool make_axis(int n, FftAxis* ax) {
  ax->n = n;
  ax->nstages = 0;
  int m = n;
  auto push = [&](int r) {
    if (ax->nstages >= KW_MAX_STAGES) return false;
    ax->radix[ax->nstages++] = r;
    m /= r;
    return true;
  };
  while (m % 8 == 0 && m >= 8)
    if (!push(8)) return false;
  while (m % 4 == 0)
    if (!push(4)) return false;
  while (m % 2 == 0)
    if (!push(2)) return false;
  for (int r : {3, 5, 7})
    while (m % r == 0)
      if (!push(r)) return false;
  if (ax->nstages == 0) {  // n == 1: identity transform
    ax->radix[0] = 1;
  }
  return m == 1;
}

enum KernelKind { KK_YPASS, KK_LAST, KK_XVEL, KK_XDENS, KK_XPRES, KK_XINIT, KK_XSRC, KK_PLANE, KK_RADIAL, KK_COUNT };
const char* kKindNames[KK_COUNT] = {"y_pass", "last_axis", "x_velocity", "x_density",
                                    "x_pressure", "x_init", "x_source", "yz_plane", "radial"};

struct SrcData {
  int* delay = nullptr;  // transducer: per-point 0-based delay into the signal
  uint64_t flag = 0;  // signal length
  int mode = 0, many = 0;
  uint64_t nseries = 0, npoints = 0;
  float* sig = nullptr;          // device (flag, nseries)
  int* row_ptr = nullptr;        // device (rows+1)
  long long* lin = nullptr;      // device (npoints)
  float* pattern = nullptr;      // device N (additive, single signal)
  float2* spec = nullptr;        // device X-domain (additive, many)
  std::vector<float> host_sig;   // for the single-signal value per step
};

}  // namespace

struct kw_sim {
  Grid g{};
  int device = 0;
  long long Nt = 0, t = 0, start = 1;
  uint64_t rec = 0;
  hipStream_t stream = nullptr;
  hipStream_t last = nullptr;  // stream of the latest kw_sim_run
  FftAxis axX{}, axY{}, axL{};
  float2 *twX = nullptr, *twY = nullptr, *twL = nullptr;
  XLaunch xl{};
  TileLaunch tly{}, tll{};
  int plane_nt = 0;  // 0: per-size default
  // axisymmetric grids: the radial pass replaces the y transform
  int axisym = 0;
  FftAxis axR{};  // length 2 Ny
  AsTables at{};
  int as_tx = 0;
  size_t as_lds = 0;
  // state
  float* u[3] = {};
  float* rho[3] = {};
  float* srho = nullptr;
  float2* Pk = nullptr;
  float2* G[3] = {};
  float* p0 = nullptr;
  // medium / operators
  Medium m{};
  SpecTables T{};
  // device allocations: stream-ordered pool allocations, or (arena_base set)
  // bump allocations inside a caller-provided region that is never freed here
  std::vector<std::pair<void*, size_t>> owned;
  size_t bytes_now = 0, bytes_peak = 0;
  char* arena_base = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  SrcData psrc, usrc[3];
  uint64_t u_flags[3] = {};
  // sensor
  int ns = 0;
  int* s_row_ptr = nullptr;
  long long* s_lin = nullptr;
  float *p_raw = nullptr, *p_max = nullptr, *p_min = nullptr, *p_max_all = nullptr, *p_min_all = nullptr,
        *p_final = nullptr;
  double* p_rms = nullptr;
  float *u_raw[3] = {}, *u_max[3] = {}, *u_min[3] = {}, *u_max_all[3] = {}, *u_min_all[3] = {};
  // u_non_staggered: per-axis series, a spectrum and a real field of scratch,
  // and the x shift pre-divided by Nx (the x C2R is unnormalised)
  float* u_ns[3] = {};
  float2* ns_spec = nullptr;
  float* ns_field = nullptr;
  float2* ns_xmult = nullptr;
  double* u_rms[3] = {};
  long long nrec_total = 0;  // Nt - start + 1
  bool own_stream = true;  // false: the caller's stream (kw_sim_run_noalloc)
  uint64_t serial = 0;     // unique per handle in the process (batch membership)
  // profiling: an event pair per launch, resolved when the totals are read
  bool profiling = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  struct EvPair {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<EvPair> prof_ev;
  size_t prof_used = 0;
  double prof_ms[KK_COUNT] = {};
  uint64_t prof_n[KK_COUNT] = {};

  template <typename T_>
  T_* alloc(size_t count) {
    void* p = nullptr;
    // every allocation is counted rounded up to kAllocAlign (the footprint
    // formula uses the same rounding)
    const size_t bytes = alloc_round(count * sizeof(T_));
    if (arena_base) {
      if (arena_used + bytes > arena_cap) throw ArenaFull{arena_used + bytes, arena_cap};
      p = arena_base + arena_used;
      arena_used += bytes;
    } else {
      // stream-ordered: no device-wide synchronisation, so simulations created
      // or destroyed while others run do not stall them
      hipError_t e = hipMallocAsync(&p, bytes, stream);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw OomError{bytes};
      }
    }
    owned.push_back({p, bytes});
    bytes_now += bytes;
    bytes_peak = std::max(bytes_peak, bytes_now);
    return static_cast<T_*>(p);
  }
  // Frees a set-up buffer once the kernels that read it are queued (pool
  // allocations only; an arena keeps it until the end of the call).
  void release(const void* p) {
    if (arena_base || !p) return;
    for (size_t i = 0; i < owned.size(); ++i)
      if (owned[i].first == p) {
        HIPCHK(hipFreeAsync(owned[i].first, stream));
        bytes_now -= owned[i].second;
        owned.erase(owned.begin() + (long)i);
        return;
      }
  }
  template <typename T_>
  T_* upload(const T_* host, size_t count) {
    T_* d = alloc<T_>(count);
    // host buffers may be temporaries of the caller: wait for this stream only
    if (count) {
      HIPCHK(hipMemcpyAsync(d, host, count * sizeof(T_), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    } else {
      zero(d, sizeof(T_));
    }
    return d;
  }
  // a fill kernel rather than hipMemsetAsync (the same code path serves
  // callers that may capture the stream into a graph)
  void zero(void* p, size_t bytes) { launch_zero(stream, p, bytes); }

  long long nspec() const { return (long long)g.Nxs * g.Ny * g.Nz; }
};

namespace {

void free_all(kw_sim* s) {
  // after the last work queued on the stream the handle last ran on
  hipStream_t st = s->last ? s->last : s->stream;
  if (st && st != s->stream && s->stream) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess) {
      (void)hipEventRecord(e, st);
      (void)hipStreamWaitEvent(s->stream, e, 0);
      (void)hipEventDestroy(e);
    }
  }
  if (!s->arena_base)
    for (const auto& p : s->owned) (void)hipFreeAsync(p.first, s->stream ? s->stream : st);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  s->owned.clear();
  s->bytes_now = 0;
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  for (const auto& e : s->prof_ev) {
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  s->prof_ev.clear();
  if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
}

// Launch geometry for the x kernels: P row pairs per workgroup with
// P*Nx <= KW_MAXE * threads, aiming at >= 1024 workgroups.
XLaunch choose_x(const Grid& g) {
  const int maxe = 4;
  int pmax = std::max(1, (maxe * 256) / g.Nx);
  int p = std::max(1, std::min(pmax, g.rows / 2048));
  p = std::min(p, std::max(1, (g.rows + 1) / 2));
  int threads = ((p * g.Nx / 2 + 63) / 64) * 64;
  threads = std::max(64, std::min(256, threads));
  while (p * g.Nx > maxe * threads && threads < 256) threads += 64;
  return XLaunch{p, threads, 0};
}

// Pencil tile width: split `width` columns evenly into tiles of <= 16 that
// keep `nbuf` LDS buffers of n*tx complex values within 64 KiB.
TileLaunch choose_tile(int width, int n, int nbuf) {
  const size_t cap = 65536;
  int txmax = 16;
  while (txmax > 1 && (size_t)n * sizeof(float2) * (1 + (size_t)nbuf * txmax) > cap) --txmax;
  int ntiles = (width + txmax - 1) / txmax;
  int tx = (width + ntiles - 1) / ntiles;
  return TileLaunch{tx, 256, 0, 0};
}

// Thread count for the length-specialised kernels: 64, 128 or 256, or 0 to
// use the generic kernels; an environment variable overrides the default.
int env_nt(const char* name, int def, bool allow512 = false) {
  const char* v = std::getenv(name);
  if (!v || !*v) return def;
  const int n = std::atoi(v);
  return (n == 0 || n == 64 || n == 128 || n == 256 || (allow512 && (n == 512 || n == 1024))) ? n : def;
}

// Radix plan of a pencil length from the environment (0 or 1; plan 1 is
// compiled for lengths >= 160 only), default 0.
int env_plan(const char* name, int n) {
  const char* v = std::getenv(name);
  if (!v || !*v) return 0;
  return (std::atoi(v) == 1 && n >= 160) ? 1 : 0;
}

// Layout of the x-domain spectra for a description: kx-major (plane pass)
// when the (y, z) plane of one kx fits a workgroup's LDS and the x rows have
// specialised kernels; otherwise row-major with the row stride Nxs (Grid::Nxs),
// padded to a multiple of 8 complex values on 3-D grids whose passes all
// have specialised kernels when that adds at most 6 % to the spectra
// (KW_PAD_SPECTRA=0 never pads, =1 always may).  kw_sim_create and the
// footprint formula both take it from here.
struct SpecLayout {
  int kxm, nxs;
};
SpecLayout spectra_layout(const kw_sim_desc* d, int lds_max) {
  const int Nx = (int)d->Nx, Ny = (int)d->Ny, Nz = (int)d->Nz;
  const int Nxh = Nx / 2 + 1, dim = Nz > 1 ? 3 : 2;
  const int nlast = dim == 3 ? Nz : Ny;
  const bool fast_rows = fast_length(Nx) && env_nt("KW_ROW_NT", 256) != 0;
  const int pencil_nt = env_nt("KW_PENCIL_NT", 256, true);
  const int last_nt = env_nt("KW_LAST_NT", std::getenv("KW_PENCIL_NT") ? pencil_nt : 256, true);
  const bool fast_y = dim == 3 && fast_length(Ny) && pencil_nt != 0;
  const bool fast_last = fast_length(nlast) && last_nt != 0;
  const size_t need = dim == 3 ? plane_lds_bytes(Ny, Nz) : 0;
  const char* pe = std::getenv("KW_PLANE");
  const bool allow = !(pe && pe[0] == '0');
  SpecLayout L{(allow && need > 0 && need <= (size_t)lds_max && fast_rows) ? 1 : 0, Nxh};
  const char* pad = std::getenv("KW_PAD_SPECTRA");
  if (dim == 3 && !d->axisymmetric_flag && !L.kxm && fast_rows && fast_y && fast_last && !(pad && pad[0] == '0')) {
    const int nxs = (Nxh + 7) / 8 * 8;
    if ((pad && pad[0] == '1') || (long long)(nxs - Nxh) * 100 <= 6LL * Nxh) L.nxs = nxs;
  }
  return L;
}

// The device's largest workgroup LDS allocation (gfx950: 160 KiB if the
// device cannot be queried).
int device_lds_max(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0) {
    (void)hipGetLastError();
    v = 160 * 1024;
  }
  return v;
}

// The x-row kernels' default workgroup target (KW_ROW_WGS or 1024).
long long row_wg_default() {
  const char* e = std::getenv("KW_ROW_WGS");
  const long long n = e ? std::atoll(e) : 0;
  return n >= 64 ? n : 1024LL;
}

size_t x_smem_bytes(const Grid& g, int P, int nbuf, bool slot) {
  size_t s = (size_t)g.Nx * sizeof(float2) + (size_t)nbuf * P * g.Nx * sizeof(float2);
  if (slot) s += (size_t)2 * P * g.Nx * sizeof(int);
  return s;
}

// CSR of sorted 0-based linear indices over rows of length Nx.
void build_rows(const std::vector<long long>& lin, const Grid& g, std::vector<int>& row_ptr) {
  row_ptr.assign(g.rows + 1, 0);
  size_t j = 0;
  for (int r = 0; r <= g.rows; ++r) {
    const long long lim = (long long)r * g.Nx;
    while (j < lin.size() && lin[j] < lim) ++j;
    row_ptr[r] = (int)j;
  }
}

kw_status load_indices(const uint64_t* idx, uint64_t n, long long N, const char* what,
                       std::vector<long long>& out) {
  out.resize(n);
  for (uint64_t j = 0; j < n; ++j) {
    if (idx[j] < 1 || (long long)idx[j] > N)
      return fail(KW_ERR_INVALID, "%s[%llu] = %llu outside 1..%lld", what, (unsigned long long)j,
                  (unsigned long long)idx[j], N);
    out[j] = (long long)idx[j] - 1;
    if (j > 0 && out[j] <= out[j - 1])
      return fail(KW_ERR_UNSUPPORTED, "%s must be strictly increasing (binary-mask find() order)", what);
  }
  return KW_OK;
}

SF make_sf(kw_sim* s, const kw_real_array& a) {
  SF r{nullptr, 0.f};
  if (a.count == 1) {
    r.s = a.data[0];
  } else {
    r.f = s->upload(a.data, (size_t)a.count);
  }
  return r;
}

kw_status check_array(const kw_real_array& a, long long N, const char* name, bool required) {
  if (!a.data || a.count == 0) {
    if (required) return fail(KW_ERR_INVALID, "%s missing", name);
    return KW_OK;
  }
  if (a.count != 1 && (long long)a.count != N)
    return fail(KW_ERR_INVALID, "%s has %llu values; expected 1 or %lld", name, (unsigned long long)a.count, N);
  return KW_OK;
}

// ---- per-kernel timing ------------------------------------------------------
// Adds the durations of the recorded launches to the per-kind totals.
void prof_resolve(kw_sim* s) {
  if (!s->prof_used) return;
  HIPCHK(hipEventSynchronize(s->prof_ev[s->prof_used - 1].b));
  for (size_t i = 0; i < s->prof_used; ++i) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->prof_ev[i].a, s->prof_ev[i].b));
    s->prof_ms[s->prof_ev[i].kind] += ms;
    s->prof_n[s->prof_ev[i].kind] += 1;
  }
  s->prof_used = 0;
}

// Brackets one launch with a pair of events on its stream (no host wait).
struct Prof {
  kw_sim* s;
  hipStream_t st;
  size_t i = 0;
  Prof(kw_sim* s_, hipStream_t st_, int k) : s(s_), st(st_) {
    if (!s->profiling) return;
    if (s->prof_used == 4096) prof_resolve(s);  // bounded event pool
    if (s->prof_used == s->prof_ev.size()) {
      kw_sim::EvPair e{nullptr, nullptr, 0};
      HIPCHK(hipEventCreate(&e.a));
      HIPCHK(hipEventCreate(&e.b));
      s->prof_ev.push_back(e);
    }
    i = s->prof_used++;
    s->prof_ev[i].kind = k;
    HIPCHK(hipEventRecord(s->prof_ev[i].a, st));
  }
  ~Prof() noexcept(false) {
    if (s->profiling) HIPCHK(hipEventRecord(s->prof_ev[i].b, st));
  }
};

// ---- transform helpers ------------------------------------------------------
void ypass(kw_sim* s, hipStream_t st, float2* const* f, int nf, float sign) {
  if (s->g.dim != 3) return;
  Prof pr(s, st, KK_YPASS);
  launch_ypass(st, s->g, s->axY, s->twY, f, nf, sign, s->tly);
}

void last(kw_sim* s, hipStream_t st, const LastArgs& a) {
  Prof pr(s, st, KK_LAST);
  launch_last(st, s->g, s->axL, s->twL, a, s->T, s->tll);
}

// Spectral operators on x-domain spectra: out_o = ifft_yz(op_o .* fft_yz(in_src(o))).
// kx-major spectra: one plane pass; otherwise y pass over the inputs, the
// last-axis pass, y pass over the outputs.
void spectral(kw_sim* s, hipStream_t st, const LastArgs& a) {
  if (s->axisym) {
    Prof pr(s, st, KK_RADIAL);
    launch_as_radial(st, s->g, s->axR, s->at, a, s->as_tx, s->as_lds);
    return;
  }
  if (s->g.kxm) {
    Prof pr(s, st, KK_PLANE);
    if (!launch_plane(st, s->g, s->twY, s->twL, a, s->T, s->plane_nt))
      throw HipError{hipErrorInvalidConfiguration, "launch_plane"};
    return;
  }
  ypass(s, st, a.in, a.nin, -1.f);
  last(s, st, a);
  ypass(s, st, a.out, a.nout, 1.f);
}

// Corrected additive source: real(ifft(source_kappa .* fft(scatter))) in X
// domain, (3D.m:815-822); unit scatters 1.0 (pattern build).  comp: the
// source's component (the radial mirror of an axisymmetric uy source differs).
void source_spectrum(kw_sim* s, hipStream_t st, const Src& src, int unit, float2* out, int comp = 0) {
  {
    Prof pr(s, st, KK_XSRC);
    launch_x_source(st, s->g, s->axX, s->twX, src, unit, out, s->xl);
  }
  LastArgs a{};
  a.in[0] = out;
  a.out[0] = out;
  a.nin = 1;
  a.nout = 1;
  a.outs[0] = SpecOut{0, OP_SKAPPA, comp};
  a.scale = (float)(1.0 / (double)s->g.N);
  spectral(s, st, a);
}

Src make_src(const kw_sim* s, const SrcData& d, long long t) {
  Src r{};
  r.active = d.flag >= (uint64_t)t;
  r.mode = d.mode;
  r.many = d.many;
  r.row_ptr = d.row_ptr;
  r.lin = d.lin;
  r.pattern = d.pattern;
  r.spec = d.spec;
  r.delay = d.delay;
  if (r.active) {
    r.sig = d.sig + (size_t)(t - 1) * d.nseries;
    r.sval = d.host_sig.empty() ? 0.f : d.host_sig[(size_t)(t - 1) * d.nseries];
  }
  (void)s;
  return r;
}

Sens make_sens_p(const kw_sim* s, long long t, bool final_step) {
  Sens r{};
  const bool rec = t >= s->start;
  r.record = rec;
  r.first = rec && (t == s->start);
  r.t_rec = rec ? t - s->start : 0;
  r.ns = s->ns;
  r.row_ptr = s->s_row_ptr;
  r.lin = s->s_lin;
  r.raw = s->p_raw;
  r.vmax = s->p_max;
  r.vmin = s->p_min;
  r.rms_sum = s->p_rms;
  r.max_all = s->p_max_all;
  r.min_all = s->p_min_all;
  r.final_field = final_step ? s->p_final : nullptr;
  return r;
}

Sens make_sens_u(const kw_sim* s, long long t, int a) {
  Sens r{};
  const bool rec = t >= s->start;
  r.record = rec;
  r.first = rec && (t == s->start);
  r.t_rec = rec ? t - s->start : 0;
  r.ns = s->ns;
  r.row_ptr = s->s_row_ptr;
  r.lin = s->s_lin;
  r.raw = s->u_raw[a];
  r.vmax = s->u_max[a];
  r.vmin = s->u_min[a];
  r.rms_sum = s->u_rms[a];
  r.max_all = s->u_max_all[a];
  r.min_all = s->u_min_all[a];
  r.final_field = nullptr;
  return r;
}

// One time step t (1-based) of kspaceFirstOrder3D.m:767-1206.
// u_non_staggered at the sensor points (extractSensorData.m:39-55): per axis,
// real(ifft(shift_neg .* fft(u, [], axis), [], axis)).  x: C2R_x(shift_x / Nx .*
// R2C_x(u)); y, z: the spectral pass with OP_SHIFT along the axis (the other
// transforms cancel), then C2R_x; each gathered at the sensor points.
void record_non_staggered(kw_sim* s, hipStream_t st, long long t) {
  if (!s->u_ns[0] || t < s->start || s->ns == 0) return;
  const Grid& g = s->g;
  const long long row = t - s->start;
  for (int a = 0; a < g.dim; ++a) {
    if (!launch_x_r2c(st, g, s->twX, s->u[a], s->ns_spec, s->xl))
      throw HipError{hipErrorInvalidConfiguration, "x R2C for u_non_staggered"};
    const float2* mult = nullptr;
    if (a == 0) {
      mult = s->ns_xmult;
    } else {
      LastArgs la{};
      la.in[0] = la.out[0] = s->ns_spec;
      la.nin = la.nout = 1;
      la.outs[0] = SpecOut{0, OP_SHIFT, a};
      la.scale = (float)(1.0 / (double)g.N);
      spectral(s, st, la);
    }
    if (!launch_x_c2r_mult(st, g, s->twX, s->ns_spec, s->ns_field, mult, s->xl))
      throw HipError{hipErrorInvalidConfiguration, "x C2R for u_non_staggered"};
    launch_gather(st, s->ns_field, s->s_lin, s->ns, s->u_ns[a] + (size_t)row * s->ns);
  }
}

// The gradient's x and y outputs share one inverse transform in the
// specialised last-axis pass of 3-D row-major grids (LastArgs::pair01).
int gradient_pair(const kw_sim* s) {
  return (s->g.dim == 3 && !s->g.kxm && !s->axisym && s->tll.nt) ? 1 : 0;
}

void step(kw_sim* s, hipStream_t st, long long t, bool final_step) {
  const Grid& g = s->g;
  const int dim = g.dim;
  const float invN = (float)(1.0 / (double)g.N);
  Sens ps = make_sens_p(s, t, final_step);
  Src nosrc[3] = {};
  Sens us[3] = {};
  for (int a = 0; a < dim; ++a) us[a] = make_sens_u(s, t, a);

  if (t == 1 && s->p0) {
    // p0 start (3D.m:1021-1035): p = p0, rho_i = p0/(dim c0^2),
    // u = dt/2 rho_sg^-1 real(ifft(ddx_k_shift_pos kappa fft(p0)))
    {
      Prof pr(s, st, KK_XINIT);
      launch_x_init(st, g, s->axX, s->twX, s->m, s->p0, s->rho, s->Pk, ps, s->xl);
    }
    LastArgs a{};
    a.in[0] = s->Pk;
    a.nin = 1;
    a.nout = dim;
    for (int c = 0; c < dim; ++c) {
      a.out[c] = s->G[c];
      a.outs[c] = SpecOut{0, OP_DPOS_K, c};
    }
    a.scale = invN;
    a.pair01 = gradient_pair(s);
    spectral(s, st, a);
    {
      Prof pr(s, st, KK_XVEL);
      launch_x_velocity(st, g, s->axX, s->twX, s->m, s->G, s->u, nosrc, us, 1, s->xl);
    }
    // restore the x spectrum of p0 for the next step
    Sens none{};
    none.row_ptr = s->s_row_ptr;
    {
      Prof pr(s, st, KK_XINIT);
      launch_x_init(st, g, s->axX, s->twX, s->m, s->p0, s->rho, s->Pk, none, s->xl);
    }
    record_non_staggered(s, st, t);
    return;
  }

  // ---- velocity: u = pml_sg (pml_sg u - dt rho_sg^-1 real(ifft(D+ kappa p_k))) (3D.m:790-801)
  {
    LastArgs a{};
    a.in[0] = s->Pk;
    a.nin = 1;
    a.nout = dim;
    for (int c = 0; c < dim; ++c) {
      a.out[c] = s->G[c];
      a.outs[c] = SpecOut{0, OP_DPOS_K, c};
    }
    a.scale = invN;
    a.pair01 = gradient_pair(s);
    spectral(s, st, a);
  }
  Src us_src[3] = {};
  for (int c = 0; c < dim; ++c) {
    us_src[c] = make_src(s, s->usrc[c], t);
    if (us_src[c].active && us_src[c].mode == 2 && us_src[c].many)
      source_spectrum(s, st, us_src[c], 0, s->usrc[c].spec, c);
  }
  {
    Prof pr(s, st, KK_XVEL);
    launch_x_velocity(st, g, s->axX, s->twX, s->m, s->G, s->u, us_src, us, 0, s->xl);
  }

  // ---- divergence: du_i = real(ifft(D-_i kappa fft(u_i))) (3D.m:905-907)
  {
    LastArgs a{};
    a.nin = dim;
    a.nout = dim;
    for (int c = 0; c < dim; ++c) {
      a.in[c] = s->G[c];
      a.out[c] = s->G[c];
      a.outs[c] = SpecOut{c, OP_DNEG_K, c};
    }
    a.scale = invN;
    spectral(s, st, a);
  }

  // ---- density, pressure source, equation of state (3D.m:910-1017)
  Src psrc = make_src(s, s->psrc, t);
  if (psrc.active && psrc.mode == 2 && psrc.many) source_spectrum(s, st, psrc, 0, s->psrc.spec);
  {
    Prof pr(s, st, KK_XDENS);
    launch_x_density(st, g, s->axX, s->twX, s->m, s->G, s->rho, psrc, s->srho, s->Pk, ps, s->xl);
  }
  if (s->m.absorbing == 1) {
    LastArgs a{};
    a.nin = 2;
    a.nout = 2;
    a.in[0] = a.out[0] = s->G[0];
    a.in[1] = a.out[1] = s->G[1];
    a.outs[0] = SpecOut{0, OP_NABLA1, 0};
    a.outs[1] = SpecOut{1, OP_NABLA2, 0};
    a.scale = invN;
    spectral(s, st, a);
    Prof pr(s, st, KK_XPRES);
    launch_x_pressure(st, g, s->axX, s->twX, s->m, s->G[0], s->G[1], s->srho, s->Pk, ps, s->xl);
  }
  record_non_staggered(s, st, t);
}

kw_status setup_source(kw_sim* s, SrcData& d, uint64_t flag, uint64_t mode, uint64_t many, uint64_t npoints,
                       const uint64_t* index, const float* input, const char* name,
                       const std::vector<long long>* shared_lin, const std::vector<int>* shared_rows,
                       int* shared_row_dev, long long* shared_lin_dev, int comp = 0) {
  d.flag = flag;
  if (!flag) return KW_OK;
  if (mode > 2) return fail(KW_ERR_INVALID, "%s mode %llu not in 0..2", name, (unsigned long long)mode);
  if (!input) return fail(KW_ERR_INVALID, "%s input missing", name);
  d.mode = (int)mode;
  d.many = many ? 1 : 0;
  d.npoints = npoints;
  d.nseries = d.many ? npoints : 1;
  (void)index;
  d.row_ptr = shared_row_dev;
  d.lin = shared_lin_dev;
  (void)shared_lin;
  (void)shared_rows;
  d.sig = s->upload(input, (size_t)(flag * d.nseries));
  if (!d.many) d.host_sig.assign(input, input + flag);
  if (d.mode == 2) {
    if (d.many) {
      d.spec = s->alloc<float2>((size_t)s->nspec());
    } else {
      // corrected unit pattern real(ifft(source_kappa fft(mask))) (3D.m:815-822)
      float2* tmp = s->alloc<float2>((size_t)s->nspec());
      Src u{};
      u.row_ptr = d.row_ptr;
      u.lin = d.lin;
      source_spectrum(s, s->stream, u, 1, tmp, comp);
      d.pattern = s->alloc<float>((size_t)s->g.N);
      launch_x_c2r(s->stream, s->g, s->axX, s->twX, tmp, d.pattern, s->xl);
      s->release(tmp);
    }
  }
  return KW_OK;
}

}  // namespace

// =============================================================================
// C ABI
// =============================================================================
// ---- checkpoint state: the time index, the fields carried between steps
// (u, rho, the spectrum of p) and every running output.  Medium, operators
// and sources are rebuilt from the input at restart.
namespace {
struct StateBuf {
  void* p;
  size_t bytes;
};
std::vector<StateBuf> state_buffers(const kw_sim* s) {
  std::vector<StateBuf> v;
  const size_t N = (size_t)s->g.N, ns = (size_t)std::max(s->ns, 1), nrec = (size_t)s->nrec_total;
  auto add = [&](const void* p, size_t b) {
    if (p) v.push_back({const_cast<void*>(p), b});
  };
  for (int a = 0; a < s->g.dim; ++a) {
    add(s->u[a], N * 4);
    add(s->rho[a], N * 4);
  }
  add(s->Pk, (size_t)s->nspec() * 8);
  add(s->p_raw, ns * nrec * 4);
  add(s->p_max, ns * 4);
  add(s->p_min, ns * 4);
  add(s->p_rms, ns * 16);
  add(s->p_max_all, N * 4);
  add(s->p_min_all, N * 4);
  add(s->p_final, N * 4);
  for (int a = 0; a < s->g.dim; ++a) {
    add(s->u_raw[a], ns * nrec * 4);
    add(s->u_max[a], ns * 4);
    add(s->u_min[a], ns * 4);
    add(s->u_rms[a], ns * 16);
    add(s->u_max_all[a], N * 4);
    add(s->u_min_all[a], N * 4);
    add(s->u_ns[a], ns * nrec * 4);
  }
  return v;
}
constexpr uint64_t kCkptMagic = 0x31305450434b574bULL;  // "KWCKPT01"
constexpr size_t kCkptHeader = 12;                      // uint64 words
void ckpt_header(const kw_sim* s, uint64_t* h) {
  h[0] = kCkptMagic;
  h[1] = (uint64_t)s->g.Nx;
  h[2] = (uint64_t)s->g.Ny;
  h[3] = (uint64_t)s->g.Nz;
  h[4] = (uint64_t)s->Nt;
  h[5] = (uint64_t)s->t;
  h[6] = (uint64_t)s->ns;
  h[7] = s->rec;
  h[8] = (uint64_t)s->start;
  h[9] = (uint64_t)s->g.kxm;
  h[10] = (uint64_t)state_buffers(s).size();
  h[11] = 0;
}
}  // namespace


// ---- autotuning of launch geometry: workgroup sizes of the row, pencil and
// plane kernels, timed over a few steps per candidate; results are cached per
// grid shape and physics branch for the process (the serving loop tunes each
// shape once).
namespace {
struct TuneKey {
  int Nx, Ny, Nz, kxm, absorbing, nonlinear, psrc, usrc;
  bool operator<(const TuneKey& o) const {
    return std::tie(Nx, Ny, Nz, kxm, absorbing, nonlinear, psrc, usrc) <
           std::tie(o.Nx, o.Ny, o.Nz, o.kxm, o.absorbing, o.nonlinear, o.psrc, o.usrc);
  }
};
struct TuneCfg {
  int row, pencil_y, pencil_last, plane, plan_y, plan_last, row_wgs;
  double ms;
};
// One entry per key; the first caller tunes while holding the entry's lock, so
// concurrent callers of the same shape wait for that result instead of tuning
// (and overwriting the cache) themselves.
struct TuneEntry {
  std::mutex mu;
  bool done = false;
  TuneCfg cfg{};
};
std::mutex g_tune_mu;
std::map<TuneKey, std::shared_ptr<TuneEntry>> g_tune_cache;

TuneKey tune_key(const kw_sim* s) {
  return TuneKey{s->g.Nx,        s->g.Ny, s->g.Nz, s->g.kxm, s->m.absorbing, s->m.nonlinear,
                 (int)(s->psrc.flag > 0), (int)(s->usrc[0].flag + s->usrc[1].flag + s->usrc[2].flag > 0)};
}

void apply_cfg(kw_sim* s, const TuneCfg& c) {
  if (s->xl.nt) s->xl.nt = c.row;
  if (s->tly.nt) s->tly.nt = c.pencil_y;
  if (s->tll.nt) s->tll.nt = c.pencil_last;
  s->plane_nt = c.plane;
  s->tly.plan = c.plan_y;
  s->tll.plan = c.plan_last;
  s->g.row_wgs = c.row_wgs;
}

TuneCfg current_cfg(const kw_sim* s) {
  return TuneCfg{s->xl.nt, s->tly.nt, s->tll.nt, s->plane_nt, s->tly.plan, s->tll.plan, s->g.row_wgs, 1e30};
}

// back to step 0: the fields carried between steps and every output cleared
void reset_state(kw_sim* s) {
  for (const StateBuf& b : state_buffers(s)) s->zero(b.p, b.bytes);
  HIPCHK(hipStreamSynchronize(s->stream));
  s->t = 0;
}
}  // namespace


// Device bytes kw_sim_create allocates for a description, following its
// allocation sequence (same order, same rounding, same early releases):
// *peak with stream-ordered pool allocations (set-up buffers released),
// *total when nothing is released (an arena).  False on an invalid grid.
static bool footprint_plan(const kw_sim_desc* d, uint64_t* peak, uint64_t* total) {
  *peak = *total = 0;
  if (!d || d->Nx == 0 || d->Ny == 0 || d->Nz == 0) return false;
  uint64_t now = 0;
  auto A = [&](uint64_t bytes) {
    const uint64_t b = alloc_round((size_t)bytes);
    now += b;
    *total += b;
    *peak = std::max(*peak, now);
  };
  auto R = [&](uint64_t bytes) { now -= alloc_round((size_t)bytes); };
  const uint64_t N = d->Nx * d->Ny * d->Nz;
  const int dim = d->Nz > 1 ? 3 : 2;
  const uint64_t Nxh = d->Nx / 2 + 1;
  const uint64_t nspec = (uint64_t)spectra_layout(d, device_lds_max(d->device)).nxs * d->Ny * d->Nz;
  const uint64_t rows = d->Ny * d->Nz;
  const bool want_ns = (d->record_flags & KW_REC_U_NON_STAGGERED) != 0;
  const bool additive = (d->p_source_flag && d->p_source_mode == 2) ||
                        ((d->ux_source_flag || d->uy_source_flag || d->uz_source_flag) && d->u_source_mode == 2);
  const uint64_t nfull[3] = {d->Nx, d->Ny, d->Nz};
  const uint64_t nout[3] = {Nxh, d->Ny, d->Nz};
  auto field = [&](const kw_real_array& a) { return a.count > 1; };
  // twiddles, derivative / shift vectors, u_non_staggered scratch
  A(d->Nx * 8);
  A(d->Ny * 8);
  if (dim == 3) A(d->Nz * 8);
  for (int a = 0; a < dim; ++a) {
    A(nout[a] * 8);
    A(nout[a] * 8);
    if (want_ns) A(nout[a] * 8);
  }
  if (want_ns) {
    A(Nxh * 8);
    A(nspec * 8);
    A(N * 4);
  }
  // spectral tables, PML
  A(nspec * 4);
  if (additive) A(nspec * 4);
  if (d->absorbing_flag == 1) {
    A(nspec * 4);
    A(nspec * 4);
  }
  for (int a = 0; a < dim; ++a) {
    A(nfull[a] * 4);
    A(nfull[a] * 4);
  }
  if (d->axisymmetric_flag) {
    const uint64_t L = 2 * d->Ny;
    A(2 * L * 8);
    A(L * 8);
    A(Nxh * L * 4);
    if (additive) A(Nxh * L * 4);
    for (int k = 0; k < 3; ++k) A(L * 8);
    A(d->Ny * 4);
  }
  // medium
  if (field(d->c0)) A(N * 4);
  if (field(d->rho0)) A(N * 4);
  const kw_real_array* sg[3] = {&d->rho0_sgx, &d->rho0_sgy, &d->rho0_sgz};
  for (int a = 0; a < dim; ++a)
    if (field(*sg[a])) {
      A(N * 4);
      A(N * 4);
      R(N * 4);
    }
  if (d->nonlinear_flag && field(d->BonA)) A(N * 4);
  if (d->absorbing_flag) {
    const bool fld = field(d->c0) || field(d->alpha_coeff);
    if (field(d->alpha_coeff)) A(N * 4);
    A((fld ? N : 1) * 4);
    A((fld ? N : 1) * 4);
    if (field(d->alpha_coeff)) R(N * 4);
  }
  // state
  for (int a = 0; a < dim; ++a) {
    A(N * 4);
    A(N * 4);
    A(nspec * 8);
  }
  A(nspec * 8);
  if (d->absorbing_flag == 1) A(N * 4);
  if (d->p0_source_input) A(N * 4);
  // sensor and outputs
  const uint64_t ns = d->sensor_npoints, nsz = std::max<uint64_t>(ns, 1);
  const uint64_t start = d->record_start_index ? d->record_start_index : 1;
  const uint64_t nrec = d->Nt >= start ? d->Nt - start + 1 : 0;
  const uint64_t rf = d->record_flags;
  A((rows + 1) * 4);
  A(nsz * 8);
  if (rf & KW_REC_P) A(nsz * nrec * 4);
  if (rf & KW_REC_P_MAX) A(nsz * 4);
  if (rf & KW_REC_P_MIN) A(nsz * 4);
  if (rf & KW_REC_P_RMS) A(2 * nsz * 8);
  if (rf & KW_REC_P_MAX_ALL) A(N * 4);
  if (rf & KW_REC_P_MIN_ALL) A(N * 4);
  if (rf & KW_REC_P_FINAL) A(N * 4);
  for (int a = 0; a < dim; ++a) {
    if (rf & KW_REC_U) A(nsz * nrec * 4);
    if (rf & KW_REC_U_MAX) A(nsz * 4);
    if (rf & KW_REC_U_MIN) A(nsz * 4);
    if (rf & KW_REC_U_RMS) A(2 * nsz * 8);
    if (rf & KW_REC_U_MAX_ALL) A(N * 4);
    if (rf & KW_REC_U_MIN_ALL) A(N * 4);
    if (rf & KW_REC_U_NON_STAGGERED) A(nsz * nrec * 4);
  }
  // sources (setup_source)
  auto src = [&](uint64_t flag, uint64_t mode, uint64_t many, uint64_t np) {
    A(flag * (many ? np : 1) * 4);
    if (mode == 2) {
      if (many) {
        A(nspec * 8);
      } else {
        A(nspec * 8);
        A(N * 4);
        R(nspec * 8);
      }
    }
  };
  if (d->p_source_flag) {
    A((rows + 1) * 4);
    A(d->p_source_npoints * 8);
    src(d->p_source_flag, d->p_source_mode, d->p_source_many, d->p_source_npoints);
  }
  if (d->transducer_source_flag) {
    A((rows + 1) * 4);
    A(d->u_source_npoints * 8);
    A(d->transducer_source_length * 4);
    A(d->u_source_npoints * 4);
  }
  const uint64_t uf[3] = {d->ux_source_flag, d->uy_source_flag, dim == 3 ? d->uz_source_flag : 0};
  if (uf[0] || uf[1] || uf[2]) {
    A((rows + 1) * 4);
    A(d->u_source_npoints * 8);
    bool shared = false;
    for (int a = 0; a < dim; ++a) {
      if (!uf[a]) continue;
      if (shared && d->u_source_mode == 2 && !d->u_source_many && !d->axisymmetric_flag) {
        A(uf[a] * 4);
        continue;
      }
      src(uf[a], d->u_source_mode, d->u_source_many, d->u_source_npoints);
      shared = d->u_source_mode == 2 && !d->u_source_many;
    }
  }
  return true;
}

extern "C" {

const char* kw_last_error(void) { return g_err.c_str(); }

uint64_t kw_sim_footprint_bytes(const kw_sim_desc* d) {
  uint64_t peak = 0, total = 0;
  footprint_plan(d, &peak, &total);
  return peak;
}

// kw_sim_create with every device buffer either allocated from the stream-ordered
// pool (arena == nullptr) or placed in [arena, arena + arena_cap); ext_stream:
// run on the caller's stream instead of creating one.
static kw_status create_impl(const kw_sim_desc* d, char* arena, size_t arena_cap, bool use_ext,
                             hipStream_t ext_stream, kw_sim** out) {
  if (!d || !out) return fail(KW_ERR_INVALID, "null argument");
  *out = nullptr;
  if (d->abi_version != KW_ABI_VERSION)
    return fail(KW_ERR_INVALID, "abi_version %u, library is %d", d->abi_version, KW_ABI_VERSION);
  if (d->Nx < 1 || d->Ny < 1 || d->Nz < 1 || d->Nt < 1)
    return fail(KW_ERR_INVALID, "grid %llux%llux%llu, Nt %llu", (unsigned long long)d->Nx,
                (unsigned long long)d->Ny, (unsigned long long)d->Nz, (unsigned long long)d->Nt);
  if (d->Nx < 2 || d->Ny < 2) return fail(KW_ERR_UNSUPPORTED, "1-D grids are not supported");
  if (d->Nx > 1024) return fail(KW_ERR_UNSUPPORTED, "Nx = %llu exceeds 1024", (unsigned long long)d->Nx);
  if (d->Ny > 8192 || d->Nz > 8192) return fail(KW_ERR_UNSUPPORTED, "Ny/Nz exceed 8192");
  const long long N = (long long)(d->Nx * d->Ny * d->Nz);
  if ((long long)(d->Ny * d->Nz) >= (1LL << 31)) return fail(KW_ERR_UNSUPPORTED, "too many rows");
  if (!(d->dx > 0 && d->dy > 0 && d->dt > 0 && d->c_ref > 0))
    return fail(KW_ERR_INVALID, "dx, dy, dt and c_ref must be positive");
  if (d->Nz > 1 && !(d->dz > 0)) return fail(KW_ERR_INVALID, "dz must be positive");

  kw_status st;
  if ((st = check_array(d->c0, N, "c0", true))) return st;
  if ((st = check_array(d->rho0, N, "rho0", true))) return st;
  if ((st = check_array(d->rho0_sgx, N, "rho0_sgx", true))) return st;
  if ((st = check_array(d->rho0_sgy, N, "rho0_sgy", true))) return st;
  if (d->Nz > 1 && (st = check_array(d->rho0_sgz, N, "rho0_sgz", true))) return st;
  if (d->nonlinear_flag && (st = check_array(d->BonA, N, "BonA", true))) return st;
  if (d->absorbing_flag > 2) return fail(KW_ERR_INVALID, "absorbing_flag %llu", (unsigned long long)d->absorbing_flag);
  if (d->absorbing_flag && (st = check_array(d->alpha_coeff, N, "alpha_coeff", true))) return st;
  if (d->absorbing_flag == 1 && std::fabs(d->alpha_power - 1.0) < 1e-12)
    return fail(KW_ERR_INVALID, "alpha_power = 1 is not supported by the power-law absorption model");

  if (d->axisymmetric_flag) {
    if (d->Nz > 1) return fail(KW_ERR_INVALID, "axisymmetric inputs are 2-D (x axial, y radial)");
    if (d->absorbing_flag == 1)
      return fail(KW_ERR_INVALID, "the axisymmetric code supports stokes absorption only (absorbing_flag 2)");
    if (d->record_flags & KW_REC_U_NON_STAGGERED)
      return fail(KW_ERR_UNSUPPORTED, "u_non_staggered is not supported for axisymmetric inputs");
  }
  // host-side checks of everything the kernels index with, before any HIP call
  const uint64_t psz_chk[3] = {d->pml_x_size, d->pml_y_size, d->Nz > 1 ? d->pml_z_size : 0};
  const uint64_t n_chk[3] = {d->Nx, d->Ny, d->Nz};
  for (int a = 0; a < 3; ++a)
    // the radial PML of an axisymmetric grid is one-sided (inputChecking.m:1378-1379)
    if ((d->axisymmetric_flag && a == 1 ? psz_chk[a] : 2 * psz_chk[a]) > n_chk[a])
      return fail(KW_ERR_INVALID, "PML size %llu too large for an axis of %llu points",
                  (unsigned long long)psz_chk[a], (unsigned long long)n_chk[a]);
  if (d->sensor_npoints && !d->sensor_mask_index) return fail(KW_ERR_INVALID, "sensor_mask_index missing");
  std::vector<long long> slin, plin, ulin;
  if ((st = load_indices(d->sensor_mask_index, d->sensor_npoints, N, "sensor_mask_index", slin))) return st;
  if (d->p_source_flag) {
    if (d->p_source_mode > 2) return fail(KW_ERR_INVALID, "p_source_mode %llu", (unsigned long long)d->p_source_mode);
    if (!d->p_source_index || !d->p_source_input || d->p_source_npoints == 0)
      return fail(KW_ERR_INVALID, "p source: index or input missing");
    if ((st = load_indices(d->p_source_index, d->p_source_npoints, N, "p_source_index", plin))) return st;
  }
  const uint64_t uflag_chk[3] = {d->ux_source_flag, d->uy_source_flag, d->Nz > 1 ? d->uz_source_flag : 0};
  const float* uin_chk[3] = {d->ux_source_input, d->uy_source_input, d->uz_source_input};
  if (uflag_chk[0] || uflag_chk[1] || uflag_chk[2]) {
    if (d->u_source_mode > 2) return fail(KW_ERR_INVALID, "u_source_mode %llu", (unsigned long long)d->u_source_mode);
    if (!d->u_source_index || d->u_source_npoints == 0) return fail(KW_ERR_INVALID, "u source: index missing");
    for (int a = 0; a < 3; ++a)
      if (uflag_chk[a] && !uin_chk[a]) return fail(KW_ERR_INVALID, "u source: input of component %d missing", a);
    if ((st = load_indices(d->u_source_index, d->u_source_npoints, N, "u_source_index", ulin))) return st;
  }
  std::vector<int> tdelay;  // transducer: 0-based delays
  if (d->transducer_source_flag) {
    if (d->Nz < 2) return fail(KW_ERR_INVALID, "transducer sources are 3-D only (inputChecking.m:1160)");
    if (uflag_chk[0] || uflag_chk[1] || uflag_chk[2])
      return fail(KW_ERR_INVALID, "a transducer source shares u_source_index and excludes ux/uy/uz sources");
    if (!d->u_source_index || d->u_source_npoints == 0 || !d->delay_mask || !d->transducer_source_input)
      return fail(KW_ERR_INVALID, "transducer source: u_source_index, delay_mask or input missing");
    if ((st = load_indices(d->u_source_index, d->u_source_npoints, N, "u_source_index", ulin))) return st;
    tdelay.resize(d->u_source_npoints);
    for (uint64_t j = 0; j < d->u_source_npoints; ++j) {
      const uint64_t dm = d->delay_mask[j];
      if (dm < 1 || dm - 1 + d->transducer_source_flag > d->transducer_source_length)
        return fail(KW_ERR_INVALID, "delay_mask[%llu] = %llu: signal of %llu samples too short for %llu steps",
                    (unsigned long long)j, (unsigned long long)dm, (unsigned long long)d->transducer_source_length,
                    (unsigned long long)d->transducer_source_flag);
      tdelay[j] = (int)(dm - 1);
    }
  }
  kw_sim* s = new kw_sim();
  {
    static std::atomic<uint64_t> next_serial{0};
    s->serial = ++next_serial;
  }
  try {
    Grid& g = s->g;
    g.Nx = (int)d->Nx;
    g.Ny = (int)d->Ny;
    g.Nz = (int)d->Nz;
    g.Nxh = g.Nx / 2 + 1;
    g.Nxs = g.Nxh;
    g.row_wgs = (int)row_wg_default();
    g.dim = d->Nz > 1 ? 3 : 2;
    g.rows = g.Ny * g.Nz;
    g.N = N;
    s->Nt = (long long)d->Nt;
    s->start = d->record_start_index ? (long long)d->record_start_index : 1;
    if (s->start > s->Nt) {
      delete s;
      return fail(KW_ERR_INVALID, "record_start_index %lld beyond Nt %lld", s->start, s->Nt);
    }
    s->rec = d->record_flags;
    s->device = d->device;
    if (!make_axis(g.Nx, &s->axX) || !make_axis(g.Ny, &s->axY) ||
        (g.dim == 3 && !make_axis(g.Nz, &s->axL))) {
      delete s;
      return fail(KW_ERR_UNSUPPORTED, "grid %dx%dx%d: every dimension must factor into 2, 3, 5 and 7", g.Nx, g.Ny,
                  g.Nz);
    }
    if (g.dim == 2) s->axL = s->axY;
    s->xl = choose_x(g);
    if (g.dim == 3) s->tly = choose_tile(g.Nxh, g.Ny, 2);
    s->tll = choose_tile(g.dim == 3 ? g.Nxh * g.Ny : g.Nxh, s->axL.n, 3);
    // workgroup size of the length-specialised kernels (0: generic kernels)
    const int pencil_nt = env_nt("KW_PENCIL_NT", 256, true);
    // last-axis pencils of 384 points and more: 512-thread workgroups (half the
    // registers per thread, twice the waves per CU; measured 12-21 % faster);
    // 768 points: 1024 threads, whose tile is 16 pencils wide (128-byte row
    // segments, one workgroup per CU; 768^3: 3.90 -> 3.75 ms per launch)
    const int last_nt = env_nt("KW_LAST_NT", std::getenv("KW_PENCIL_NT") ? pencil_nt
                                                 : (s->axL.n == 768 ? 1024 : s->axL.n >= 384 ? 512 : 256), true);
    // 256-thread rows from 256 points up and at 160 (whose halved row pairs
    // leave 256 threads five slots each: 160x160x192 step 0.547 -> 0.507 ms)
    int row_nt = env_nt("KW_ROW_NT", (g.Nx >= 256 || g.Nx == 160) ? 256 : 128);
    s->xl.nt = fast_length(g.Nx) ? row_nt : 0;
    s->tly.nt = (g.dim == 3 && fast_length(g.Ny)) ? pencil_nt : 0;
    s->tll.nt = fast_length(s->axL.n) ? last_nt : 0;
    // radix plans of the pencils (kwt::next_radix; plan 1 exists for lengths >= 160)
    s->tly.plan = env_plan("KW_YPASS_PLAN", g.Ny);
    s->tll.plan = env_plan("KW_LAST_PLAN", s->axL.n);
    if ((d->record_flags & KW_REC_U_NON_STAGGERED) && !s->xl.nt) {
      delete s;
      return fail(KW_ERR_UNSUPPORTED, "u_non_staggered needs Nx among the compiled row lengths "
                                      "(32, 48, 64, 96, 128, 160, 192, 256, 384, 512, 768)");
    }
    if (x_smem_bytes(g, s->xl.pairs, 3, true) > 65536) {
      delete s;
      return fail(KW_ERR_UNSUPPORTED, "x-row tile does not fit in LDS");
    }

    HIPCHK(hipSetDevice(s->device));
    // kx-major spectra and the plane pass when the (y, z) plane of one kx
    // fits a workgroup's LDS and the x rows have specialised kernels
    {
      int lds_max = 0;
      HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
      const SpecLayout L = spectra_layout(d, lds_max);
      g.kxm = L.kxm;
      g.Nxs = L.nxs;
      // padded rows need every pass specialised (the generic kernels use Nxh rows)
      if (g.Nxs != g.Nxh && !(s->xl.nt && s->tly.nt && s->tll.nt))
        throw HipError{hipErrorInvalidConfiguration, "spectra_layout"};
      if (d->axisymmetric_flag) {
        s->axisym = 1;
        s->as_tx = make_axis(2 * g.Ny, &s->axR) ? as_radial_tile(g.Ny, (size_t)lds_max, &s->as_lds) : 0;
        if (!s->as_tx) {
          delete s;
          return fail(KW_ERR_UNSUPPORTED, "axisymmetric radial pass: Ny = %d does not fit the workgroup's LDS", g.Ny);
        }
      }
      // kx-major rows stage their spectra through LDS: 256-thread row kernels
      // (measured fastest there; otherwise 256 for Nx >= 256 and 128 below)
      if (g.kxm && !std::getenv("KW_ROW_NT")) s->xl.nt = 256;
    }
    if (use_ext) {
      s->stream = ext_stream;
      s->own_stream = false;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    s->arena_base = arena;
    s->arena_cap = arena ? arena_cap : 0;
    hipStream_t st0 = s->stream;

    // twiddle tables (FP64 -> FP32)
    s->twX = s->alloc<float2>(g.Nx);
    launch_twiddles(st0, s->twX, g.Nx);
    s->twY = s->alloc<float2>(g.Ny);
    launch_twiddles(st0, s->twY, g.Ny);
    if (g.dim == 3) {
      s->twL = s->alloc<float2>(g.Nz);
      launch_twiddles(st0, s->twL, g.Nz);
    } else {
      s->twL = s->twY;
    }

    // derivative / shift vectors (3D.m:627-634), spectral tables (3D.m:651-655)
    const double dxs[3] = {d->dx, d->dy, g.dim == 3 ? d->dz : 1.0};
    const int nfull[3] = {g.Nx, g.Ny, g.Nz};
    const int nout[3] = {g.Nxh, g.Ny, g.Nz};
    const bool want_ns = (d->record_flags & KW_REC_U_NON_STAGGERED) != 0;
    for (int a = 0; a < g.dim; ++a) {
      float2* dp = s->alloc<float2>(nout[a]);
      float2* dn = s->alloc<float2>(nout[a]);
      float2* sh = want_ns ? s->alloc<float2>(nout[a]) : nullptr;
      launch_dvec(st0, dp, dn, sh, nfull[a], nout[a], dxs[a]);
      s->T.dpos[a] = dp;
      s->T.dneg[a] = dn;
      s->T.shift[a] = sh;
    }
    if (want_ns) {
      // x shift / Nx for the x C2R (the y and z shifts go through the spectral pass, 1/N folded in)
      std::vector<float2> hx((size_t)g.Nxh);
      HIPCHK(hipMemcpyAsync(hx.data(), s->T.shift[0], hx.size() * sizeof(float2), hipMemcpyDeviceToHost, st0));
      HIPCHK(hipStreamSynchronize(st0));
      for (auto& v : hx) {
        v.x /= (float)g.Nx;
        v.y /= (float)g.Nx;
      }
      s->ns_xmult = s->upload(hx.data(), hx.size());
      s->ns_spec = s->alloc<float2>((size_t)s->nspec());
      s->ns_field = s->alloc<float>((size_t)N);
    }
    OpParams op{};
    op.dx[0] = dxs[0];
    op.dx[1] = dxs[1];
    op.dx[2] = dxs[2];
    op.dt = d->dt;
    op.c_ref = d->c_ref;
    op.alpha_power = d->alpha_power;
    op.absorbing = (int)d->absorbing_flag;
    float* kappa = s->alloc<float>((size_t)s->nspec());
    const bool additive = (d->p_source_flag && d->p_source_mode == 2) ||
                          ((d->ux_source_flag || d->uy_source_flag || d->uz_source_flag) && d->u_source_mode == 2);
    float* skappa = additive ? s->alloc<float>((size_t)s->nspec()) : nullptr;
    float* nab1 = d->absorbing_flag == 1 ? s->alloc<float>((size_t)s->nspec()) : nullptr;
    float* nab2 = d->absorbing_flag == 1 ? s->alloc<float>((size_t)s->nspec()) : nullptr;
    launch_spectral_tables(st0, g, op, kappa, skappa, nab1, nab2);
    s->T.kappa = kappa;
    s->T.skappa = skappa;
    s->T.nabla1 = nab1;
    s->T.nabla2 = nab2;

    // PML (3D.m:617-622)
    const uint64_t psz[3] = {d->pml_x_size, d->pml_y_size, d->pml_z_size};
    const double pal[3] = {d->pml_x_alpha, d->pml_y_alpha, d->pml_z_alpha};
    for (int a = 0; a < g.dim; ++a) {
      float* pm = s->alloc<float>(nfull[a]);
      float* pms = s->alloc<float>(nfull[a]);
      // axisymmetric: radial layer at the outer edge only (kspaceFirstOrderAS.m:470-471)
      const int one_sided = s->axisym && a == 1;
      launch_pml(st0, pm, nfull[a], dxs[a], d->dt, d->c_ref, (int)psz[a], pal[a], 0, one_sided);
      launch_pml(st0, pms, nfull[a], dxs[a], d->dt, d->c_ref, (int)psz[a], pal[a], 1, one_sided);
      s->m.pml[a] = pm;
      s->m.pml_sg[a] = pms;
    }

    // axisymmetric radial-pass tables (kspaceFirstOrderAS.m:488-521, 571), FP64 -> FP32
    if (s->axisym) {
      const int L = 2 * g.Ny;
      float2* tw4 = s->alloc<float2>((size_t)2 * L);
      float2* tw2 = s->alloc<float2>((size_t)L);
      launch_twiddles(st0, tw4, 2 * L);
      launch_twiddles(st0, tw2, L);
      AsTables& at = s->at;
      at.tw4 = tw4;
      at.tw2 = tw2;
      at.kappa = s->alloc<float>((size_t)g.Nxh * L);
      at.skappa = additive ? s->alloc<float>((size_t)g.Nxh * L) : nullptr;
      at.ydpos = s->alloc<float2>((size_t)L);
      at.ydneg = s->alloc<float2>((size_t)L);
      at.yshn = s->alloc<float2>((size_t)L);
      at.ysg_inv = s->alloc<float>((size_t)g.Ny);
      at.dposx = s->T.dpos[0];
      at.dnegx = s->T.dneg[0];
      launch_as_tables(st0, g, d->dx, d->dy, d->dt, d->c_ref, at);
    }

    // medium
    Medium& m = s->m;
    m.dt = (float)d->dt;
    m.nonlinear = d->nonlinear_flag ? 1 : 0;
    m.absorbing = (int)d->absorbing_flag;
    m.c0 = make_sf(s, d->c0);
    m.rho0 = make_sf(s, d->rho0);
    const kw_real_array* sg[3] = {&d->rho0_sgx, &d->rho0_sgy, &d->rho0_sgz};
    for (int a = 0; a < g.dim; ++a) {
      if (sg[a]->count == 1) {
        m.rsg[a].s = (float)(1.0 / (double)sg[a]->data[0]);
      } else {
        float* tmp = s->upload(sg[a]->data, (size_t)N);
        float* inv = s->alloc<float>((size_t)N);
        launch_reciprocal(st0, N, tmp, inv);
        s->release(tmp);
        m.rsg[a].f = inv;
      }
    }
    if (m.nonlinear) m.bona = make_sf(s, d->BonA);
    if (m.absorbing) {
      const bool field = d->c0.count > 1 || d->alpha_coeff.count > 1;
      const float* c0d = m.c0.f;
      const float* ad = d->alpha_coeff.count > 1 ? s->upload(d->alpha_coeff.data, (size_t)N) : nullptr;
      const long long n = field ? N : 1;
      float* tau = s->alloc<float>((size_t)n);
      float* eta = s->alloc<float>((size_t)n);
      launch_absorb_coeffs(st0, n, field ? c0d : nullptr, d->c0.data[0], ad, d->alpha_coeff.data[0],
                           d->alpha_power, (int)d->absorbing_flag, tau, eta);
      s->release(ad);
      if (field) {
        m.tau.f = tau;
        m.eta.f = eta;
      } else {
        float h[2];
        HIPCHK(hipMemcpyAsync(&h[0], tau, sizeof(float), hipMemcpyDeviceToHost, st0));
        HIPCHK(hipMemcpyAsync(&h[1], eta, sizeof(float), hipMemcpyDeviceToHost, st0));
        HIPCHK(hipStreamSynchronize(st0));
        m.tau.s = h[0];
        m.eta.s = m.absorbing == 1 ? h[1] : 0.f;
      }
    }

    // state
    for (int a = 0; a < g.dim; ++a) {
      s->u[a] = s->alloc<float>((size_t)N);
      s->rho[a] = s->alloc<float>((size_t)N);
      s->G[a] = s->alloc<float2>((size_t)s->nspec());
      s->zero(s->u[a], (size_t)N * 4);
      s->zero(s->rho[a], (size_t)N * 4);
    }
    if (g.dim == 2) {  // the absorbing EOS uses G[0], G[1]; keep G[2] valid
      s->G[2] = nullptr;
    }
    s->Pk = s->alloc<float2>((size_t)s->nspec());
    s->zero(s->Pk, (size_t)s->nspec() * sizeof(float2));
    if (m.absorbing == 1) s->srho = s->alloc<float>((size_t)N);
    if (d->p0_source_input) s->p0 = s->upload(d->p0_source_input, (size_t)N);

    // sensor
    s->ns = (int)d->sensor_npoints;
    std::vector<int> srows;
    build_rows(slin, g, srows);
    s->s_row_ptr = s->upload(srows.data(), srows.size());
    s->s_lin = s->upload(slin.data(), slin.size());  // (no sensor points: one zeroed entry)
    s->nrec_total = s->Nt - s->start + 1;
    const size_t nsz = std::max(s->ns, 1);
    const uint64_t rf = s->rec;
    // time series start zeroed: a fetch before the last step reads zeros for
    // the steps still to come, not whatever the memory held
    if (rf & KW_REC_P) {
      s->p_raw = s->alloc<float>(nsz * (size_t)s->nrec_total);
      s->zero(s->p_raw, nsz * (size_t)s->nrec_total * 4);
    }
    if (rf & KW_REC_P_MAX) s->p_max = s->alloc<float>(nsz);
    if (rf & KW_REC_P_MIN) s->p_min = s->alloc<float>(nsz);
    if (rf & KW_REC_P_RMS) s->p_rms = s->alloc<double>(2 * nsz);
    if (rf & KW_REC_P_MAX_ALL) s->p_max_all = s->alloc<float>((size_t)N);
    if (rf & KW_REC_P_MIN_ALL) s->p_min_all = s->alloc<float>((size_t)N);
    if (rf & KW_REC_P_FINAL) s->p_final = s->alloc<float>((size_t)N);
    for (int a = 0; a < g.dim; ++a) {
      if (rf & KW_REC_U) {
        s->u_raw[a] = s->alloc<float>(nsz * (size_t)s->nrec_total);
        s->zero(s->u_raw[a], nsz * (size_t)s->nrec_total * 4);
      }
      if (rf & KW_REC_U_MAX) s->u_max[a] = s->alloc<float>(nsz);
      if (rf & KW_REC_U_MIN) s->u_min[a] = s->alloc<float>(nsz);
      if (rf & KW_REC_U_RMS) s->u_rms[a] = s->alloc<double>(2 * nsz);
      if (rf & KW_REC_U_MAX_ALL) s->u_max_all[a] = s->alloc<float>((size_t)N);
      if (rf & KW_REC_U_MIN_ALL) s->u_min_all[a] = s->alloc<float>((size_t)N);
      if (rf & KW_REC_U_NON_STAGGERED) {
        s->u_ns[a] = s->alloc<float>(nsz * (size_t)s->nrec_total);
        s->zero(s->u_ns[a], nsz * (size_t)s->nrec_total * 4);
      }
    }

    // sources (point lists share the row CSR construction)
    if (d->p_source_flag) {
      const std::vector<long long>& lin = plin;
      std::vector<int> rows;
      build_rows(lin, g, rows);
      int* rd = s->upload(rows.data(), rows.size());
      long long* ld = s->upload(lin.data(), lin.size());
      if ((st = setup_source(s, s->psrc, d->p_source_flag, d->p_source_mode, d->p_source_many, d->p_source_npoints,
                             d->p_source_index, d->p_source_input, "p_source", nullptr, nullptr, rd, ld))) {
        free_all(s);
        delete s;
        return st;
      }
    }
    if (d->transducer_source_flag) {
      // ux additive-no-correction with a per-point offset into one signal (3D.m:889-902)
      std::vector<int> rows;
      build_rows(ulin, g, rows);
      SrcData& sd = s->usrc[0];
      sd.flag = d->transducer_source_flag;
      sd.mode = 1;
      sd.many = 0;
      sd.npoints = d->u_source_npoints;
      sd.nseries = 1;
      sd.row_ptr = s->upload(rows.data(), rows.size());
      sd.lin = s->upload(ulin.data(), ulin.size());
      sd.sig = s->upload(d->transducer_source_input, (size_t)d->transducer_source_length);
      sd.delay = s->upload(tdelay.data(), tdelay.size());
    }
    const uint64_t uf[3] = {d->ux_source_flag, d->uy_source_flag, g.dim == 3 ? d->uz_source_flag : 0};
    const float* uin[3] = {d->ux_source_input, d->uy_source_input, d->uz_source_input};
    if (uf[0] || uf[1] || uf[2]) {
      const std::vector<long long>& lin = ulin;
      std::vector<int> rows;
      build_rows(lin, g, rows);
      int* rd = s->upload(rows.data(), rows.size());
      long long* ld = s->upload(lin.data(), lin.size());
      float* shared_pattern = nullptr;
      for (int a = 0; a < g.dim; ++a) {
        if (!uf[a]) continue;
        SrcData& sd = s->usrc[a];
        // (axisymmetric: ux and uy patterns differ, their radial mirrors do)
        if (shared_pattern && d->u_source_mode == 2 && !d->u_source_many && !s->axisym) {
          // the corrected pattern depends only on the (shared) positions
          sd.flag = uf[a];
          sd.mode = 2;
          sd.many = 0;
          sd.npoints = d->u_source_npoints;
          sd.nseries = 1;
          sd.row_ptr = rd;
          sd.lin = ld;
          sd.sig = s->upload(uin[a], (size_t)uf[a]);
          sd.host_sig.assign(uin[a], uin[a] + uf[a]);
          sd.pattern = shared_pattern;
          continue;
        }
        if ((st = setup_source(s, sd, uf[a], d->u_source_mode, d->u_source_many, d->u_source_npoints,
                               d->u_source_index, uin[a], "u_source", nullptr, nullptr, rd, ld, a))) {
          free_all(s);
          delete s;
          return st;
        }
        shared_pattern = sd.pattern;
      }
    }
    HIPCHK(hipStreamSynchronize(st0));
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    free_all(s);
    delete s;
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  } catch (const OomError& e) {
    free_all(s);
    delete s;
    return fail(KW_ERR_OOM, "device allocation of %zu bytes failed", e.bytes);
  } catch (const ArenaFull& e) {
    free_all(s);
    delete s;
    return fail(KW_ERR_UNSUPPORTED, "simulation needs more than the %zu-byte scratch region (%zu bytes)", e.cap,
                e.need);
  }
  *out = s;
  return KW_OK;
}

kw_status kw_sim_create(const kw_sim_desc* d, kw_sim** out) {
  return create_impl(d, nullptr, 0, false, nullptr, out);
}

kw_status kw_sim_memory(const kw_sim* s, uint64_t* allocated, uint64_t* peak) {
  if (!s) return fail(KW_ERR_INVALID, "null handle");
  if (allocated) *allocated = s->bytes_now;
  if (peak) *peak = s->bytes_peak;
  return KW_OK;
}

kw_status kw_sim_run(kw_sim* s, uint64_t nsteps, void* stream) {
  if (!s) return fail(KW_ERR_INVALID, "null handle");
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  s->last = st;
  try {
    HIPCHK(hipSetDevice(s->device));
    const long long end = std::min<long long>(s->Nt, s->t + (long long)nsteps);
    for (long long t = s->t + 1; t <= end; ++t) step(s, st, t, t == end);
    s->t = end;
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

kw_status kw_sim_sync(kw_sim* s) {
  if (!s) return fail(KW_ERR_INVALID, "null handle");
  hipError_t e = hipStreamSynchronize(s->last ? s->last : s->stream);
  if (e != hipSuccess) return fail(KW_ERR_DEVICE, "HIP error %d (%s)", (int)e, hipGetErrorString(e));
  return KW_OK;
}

uint64_t kw_sim_steps_done(const kw_sim* s) { return s ? (uint64_t)s->t : 0; }

void* kw_sim_stream(kw_sim* s) { return s ? (void*)s->stream : nullptr; }

static const void* output_ptr(const kw_sim* s, int id, size_t* bytes, bool* is_rms) {
  *is_rms = false;
  const size_t N = (size_t)s->g.N;
  const size_t ns = (size_t)s->ns;
  const size_t nrec = (size_t)s->nrec_total;
  if (id < 0 || id >= KW_OUT_COUNT) return nullptr;
  if (id >= KW_OUT_UX) {
    const int base = (id - KW_OUT_UX) / 3 * 3 + KW_OUT_UX;
    const int a = id - base;
    if (a >= s->g.dim) return nullptr;
    switch (base) {
      case KW_OUT_UX: *bytes = ns * nrec * 4; return s->u_raw[a];
      case KW_OUT_UX_MAX: *bytes = ns * 4; return s->u_max[a];
      case KW_OUT_UX_MIN: *bytes = ns * 4; return s->u_min[a];
      case KW_OUT_UX_RMS: *bytes = ns * 4; *is_rms = true; return s->u_rms[a];
      case KW_OUT_UX_MAX_ALL: *bytes = N * 4; return s->u_max_all[a];
      case KW_OUT_UX_MIN_ALL: *bytes = N * 4; return s->u_min_all[a];
      case KW_OUT_UX_FINAL: *bytes = N * 4; return (s->rec & KW_REC_U_FINAL) ? s->u[a] : nullptr;
      case KW_OUT_UX_NON_STAGGERED: *bytes = ns * nrec * 4; return s->u_ns[a];
    }
    return nullptr;
  }
  switch (id) {
    case KW_OUT_P: *bytes = ns * nrec * 4; return s->p_raw;
    case KW_OUT_P_MAX: *bytes = ns * 4; return s->p_max;
    case KW_OUT_P_MIN: *bytes = ns * 4; return s->p_min;
    case KW_OUT_P_RMS: *bytes = ns * 4; *is_rms = true; return s->p_rms;
    case KW_OUT_P_MAX_ALL: *bytes = N * 4; return s->p_max_all;
    case KW_OUT_P_MIN_ALL: *bytes = N * 4; return s->p_min_all;
    case KW_OUT_P_FINAL: *bytes = N * 4; return s->p_final;
  }
  return nullptr;
}

size_t kw_sim_output_bytes(const kw_sim* s, int id) {
  if (!s) return 0;
  size_t b = 0;
  bool rms;
  return output_ptr(s, id, &b, &rms) ? b : 0;
}

kw_status kw_sim_fetch(kw_sim* s, int id, void* dst, size_t bytes) {
  if (!s || !dst) return fail(KW_ERR_INVALID, "null argument");
  size_t need = 0;
  bool rms = false;
  const void* src = output_ptr(s, id, &need, &rms);
  if (!src) return fail(KW_ERR_INVALID, "output %d was not recorded", id);
  if (bytes != need) return fail(KW_ERR_INVALID, "output %d needs %zu bytes, got %zu", id, need, bytes);
  try {
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->last ? s->last : s->stream));
    if (rms) {
      const long long nrec_done = std::max<long long>(0, s->t - s->start + 1);
      std::vector<double> h(2 * (size_t)s->ns);
      HIPCHK(hipMemcpy(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost));
      float* o = static_cast<float*>(dst);
      for (int j = 0; j < s->ns; ++j)
        o[j] = nrec_done > 0 ? (float)std::sqrt(h[2 * (size_t)j] / (double)nrec_done) : 0.f;
    } else {
      HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    }
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

kw_status kw_sim_destroy(kw_sim* s) {
  if (!s) return KW_OK;
  (void)hipSetDevice(s->device);
  free_all(s);
  delete s;
  return KW_OK;
}

size_t kw_sim_checkpoint_bytes(const kw_sim* s) {
  if (!s) return 0;
  size_t b = kCkptHeader * 8;
  for (const StateBuf& x : state_buffers(s)) b += x.bytes;
  return b;
}

kw_status kw_sim_checkpoint_save(kw_sim* s, void* host, size_t bytes) {
  if (!s || !host) return fail(KW_ERR_INVALID, "null argument");
  if (bytes != kw_sim_checkpoint_bytes(s)) return fail(KW_ERR_INVALID, "checkpoint needs %zu bytes", kw_sim_checkpoint_bytes(s));
  try {
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->last ? s->last : s->stream));
    uint64_t* h = static_cast<uint64_t*>(host);
    ckpt_header(s, h);
    char* p = static_cast<char*>(host) + kCkptHeader * 8;
    for (const StateBuf& x : state_buffers(s)) {
      HIPCHK(hipMemcpy(p, x.p, x.bytes, hipMemcpyDeviceToHost));
      p += x.bytes;
    }
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

kw_status kw_sim_checkpoint_load(kw_sim* s, const void* host, size_t bytes) {
  if (!s || !host) return fail(KW_ERR_INVALID, "null argument");
  if (bytes != kw_sim_checkpoint_bytes(s)) return fail(KW_ERR_INVALID, "checkpoint size does not match this simulation");
  const uint64_t* h = static_cast<const uint64_t*>(host);
  uint64_t want[kCkptHeader];
  ckpt_header(s, want);
  if (h[0] != kCkptMagic) return fail(KW_ERR_INVALID, "not a checkpoint of this solver");
  for (size_t i = 1; i < kCkptHeader; ++i)
    if (i != 5 && h[i] != want[i])
      return fail(KW_ERR_INVALID, "checkpoint was written for a different simulation (header word %zu)", i);
  if (h[5] > (uint64_t)s->Nt) return fail(KW_ERR_INVALID, "checkpoint time index beyond Nt");
  try {
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->last ? s->last : s->stream));
    const char* p = static_cast<const char*>(host) + kCkptHeader * 8;
    for (const StateBuf& x : state_buffers(s)) {
      HIPCHK(hipMemcpy(x.p, p, x.bytes, hipMemcpyHostToDevice));
      p += x.bytes;
    }
    s->t = (long long)h[5];
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

kw_status kw_sim_autotune(kw_sim* s, int trial_steps, int* row_nt, int* pencil_nt, int* plane_nt) {
  if (!s) return fail(KW_ERR_INVALID, "null handle");
  if (trial_steps < 1) trial_steps = 4;
  try {
    HIPCHK(hipSetDevice(s->device));
    const TuneKey key = tune_key(s);
    TuneCfg best = current_cfg(s);
    std::shared_ptr<TuneEntry> ent;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      std::shared_ptr<TuneEntry>& slot = g_tune_cache[key];
      if (!slot) slot = std::make_shared<TuneEntry>();
      ent = slot;
    }
    std::lock_guard<std::mutex> klk(ent->mu);
    if (ent->done) best = ent->cfg;
    if (!ent->done) {
      if (!s->ev0) HIPCHK(hipEventCreate(&s->ev0));
      if (!s->ev1) HIPCHK(hipEventCreate(&s->ev1));
      const long long steps = std::min<long long>(s->Nt, trial_steps + 1);
      auto trial = [&](TuneCfg c) {
        apply_cfg(s, c);
        reset_state(s);
        step(s, s->stream, 1, false);  // first step (code objects, caches)
        HIPCHK(hipEventRecord(s->ev0, s->stream));
        for (long long t = 2; t <= steps; ++t) step(s, s->stream, t, false);
        HIPCHK(hipEventRecord(s->ev1, s->stream));
        HIPCHK(hipEventSynchronize(s->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        c.ms = ms / std::max<long long>(1, steps - 1);
        if (c.ms < best.ms) best = c;
      };
      // coordinate descent from the defaults: row workgroup size, then
      // (y pass, last axis) workgroup sizes x radix plans, then the plane pass
      best.ms = 1e30;
      trial(best);
      if (s->xl.nt) {
        for (int r : (s->g.Nx <= 96 ? std::vector<int>{64, 128, 256} : std::vector<int>{128, 256})) {
          TuneCfg c = best;
          c.row = r;
          trial(c);
        }
        // row pairs per workgroup, through the launch's workgroup target
        for (int w : {512, 2048, 4096}) {
          TuneCfg c = best;
          c.row_wgs = w;
          trial(c);
        }
      }
      // (axisymmetric grids: the radial pass replaces the pencils)
      if (!s->g.kxm && !s->axisym && (s->tly.nt || s->tll.nt)) {
        const std::vector<std::pair<int, int>> nts{{128, 128}, {256, 256}, {256, 512}, {512, 512}, {256, 1024}};
        const bool py = s->g.dim == 3 && s->axY.n >= 160, pl = s->axL.n >= 160;
        for (const auto& pe : nts)
          for (int a = 0; a <= (py ? 1 : 0); ++a)
            for (int b = 0; b <= (pl ? 1 : 0); ++b) {
              TuneCfg c = best;
              c.pencil_y = pe.first;
              c.pencil_last = pe.second;
              c.plan_y = a;
              c.plan_last = b;
              trial(c);
            }
      }
      if (s->g.kxm)
        for (int pn : (s->g.Ny * s->g.Nz < 16384 ? std::vector<int>{256, 512, 1024} : std::vector<int>{512, 1024})) {
          TuneCfg c = best;
          c.plane = pn;
          trial(c);
        }
      ent->cfg = best;
      ent->done = true;
    }
    apply_cfg(s, best);
    reset_state(s);
    if (row_nt) *row_nt = s->xl.nt;
    if (pencil_nt) *pencil_nt = s->tly.nt ? s->tly.nt : s->tll.nt;
    if (plane_nt) *plane_nt = s->plane_nt;
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

kw_status kw_sim_launch_config(const kw_sim* s, kw_launch_config* c) {
  if (!s || !c) return fail(KW_ERR_INVALID, "null argument");
  c->row_nt = s->xl.nt;
  c->ypass_nt = s->g.dim == 3 && !s->g.kxm && !s->axisym ? s->tly.nt : 0;
  c->last_nt = !s->g.kxm && !s->axisym ? s->tll.nt : 0;
  c->plane_nt = s->g.kxm ? (s->plane_nt ? s->plane_nt : default_plane_nt(s->g.Ny, s->g.Nz)) : 0;
  c->ypass_plan = c->ypass_nt ? s->tly.plan : 0;
  c->last_plan = c->last_nt ? s->tll.plan : 0;
  c->kx_major = s->g.kxm;
  c->row_wg_target = s->xl.nt ? s->g.row_wgs : 0;
  return KW_OK;
}

kw_status kw_sim_set_profiling(kw_sim* s, int enable) {
  if (!s) return fail(KW_ERR_INVALID, "null handle");
  s->profiling = enable != 0;
  s->prof_used = 0;  // launches recorded but not yet read are dropped
  for (int k = 0; k < KK_COUNT; ++k) {
    s->prof_ms[k] = 0.0;
    s->prof_n[k] = 0;
  }
  return KW_OK;
}

int kw_sim_profile_get(kw_sim* s, const char** names, double* total_ms, uint64_t* launches, int cap) {
  if (!s) return 0;
  try {
    HIPCHK(hipSetDevice(s->device));
    prof_resolve(s);
  } catch (const HipError& e) {
    fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
    return -1;
  }
  for (int k = 0; k < KK_COUNT && k < cap; ++k) {
    if (names) names[k] = kKindNames[k];
    if (total_ms) total_ms[k] = s->prof_ms[k];
    if (launches) launches[k] = s->prof_n[k];
  }
  return KK_COUNT;
}

kw_status kw_diag_copy(uint64_t bytes, int device, double* ms) { return kw_diag_copy_ex(bytes, device, 0, 0, ms); }

kw_status kw_diag_copy_ex(uint64_t bytes, int device, int variant, int blocks, double* ms) {
  if (variant < 0 || variant > 4 || blocks < 0 || bytes < 16)
    return fail(KW_ERR_INVALID, "copy variant %d, %d blocks, %llu bytes", variant, blocks, (unsigned long long)bytes);
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  kw_status st = KW_OK;
  try {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipMalloc(&a, bytes));
    HIPCHK(hipMalloc(&b, bytes));
    HIPCHK(hipMemset(a, 0, bytes));
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch_copy_variant(nullptr, a, b, (size_t)bytes, variant, blocks);  // code-object load, cold caches
    HIPCHK(hipEventRecord(e0, nullptr));
    const int reps = 10;
    for (int r = 0; r < reps; ++r)
      launch_copy_variant(nullptr, (r & 1) ? b : a, (r & 1) ? a : b, (size_t)bytes, variant, blocks);
    HIPCHK(hipEventRecord(e1, nullptr));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t / reps;
  } catch (const HipError& e) {
    st = fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return st;
}

}  // extern "C"

// =============================================================================
// Batched steps: simulations of one grid shape, physics branch and launch
// configuration advance together, each kernel of the step launched once for
// the whole batch (one simulation per blockIdx.z) with per-simulation
// arguments in device tables.  A step prologue kernel advances every
// simulation's time index on the device, so the launches of a quantum of
// steps do not change between quanta and are captured once in a hipGraph,
// replayed while the batch's membership is unchanged.  Each simulation runs
// the kernel bodies, launch shapes and statement order of its own
// single-simulation step, so its results are the same bit for bit.
// =============================================================================
namespace {

// The batch key (empty: the simulation's steps cannot be batched: 2-D and
// axisymmetric grids, u_non_staggered, additive sources with one signal per
// point, or lengths without specialised kernels).
std::string batch_key_of(const kw_sim* s) {
  const Grid& g = s->g;
  if (g.dim != 3 || s->axisym || s->u_ns[0]) return std::string();
  if (s->psrc.flag && s->psrc.mode == 2 && s->psrc.many) return std::string();
  for (int c = 0; c < 3; ++c)
    if (s->usrc[c].flag && s->usrc[c].mode == 2 && s->usrc[c].many) return std::string();
  if (!s->xl.nt) return std::string();
  if (!g.kxm && (!s->tly.nt || !s->tll.nt)) return std::string();
  char buf[256];
  std::snprintf(buf, sizeof buf, "%dx%dx%d/%d kxm%d abs%d nl%d row%d/%d y%d/%d/%d last%d/%d/%d plane%d dev%d", g.Nx,
                g.Ny, g.Nz, g.Nxs, g.kxm, s->m.absorbing, s->m.nonlinear, s->xl.nt, g.row_wgs, s->tly.nt, s->tly.plan,
                s->tly.tx, s->tll.nt, s->tll.plan, s->tll.tx, s->plane_nt, s->device);
  return std::string(buf);
}

}  // namespace

namespace kw {

struct SimBatch {
  std::string key;
  int device = 0;
  // membership the device tables were built for, and each member's step
  // count they assume
  std::vector<kw_sim*> members;
  std::vector<uint64_t> serials;
  std::vector<long long> t_expect;
  // device tables (one allocation)
  char* table = nullptr;
  size_t table_cap = 0;
  hipStream_t table_stream = nullptr;
  BatchClock* clk = nullptr;
  int* live = nullptr;
  XVelJob* vel = nullptr;
  XDensJob* dens = nullptr;
  XPresJob* pres = nullptr;
  PlaneJob* plane[3] = {};
  YpassJob* yin[3] = {};
  LastJob* lastj[3] = {};
  YpassJob* yout[3] = {};
  int nstage = 0;
  int stage_nin[3] = {}, stage_nout[3] = {}, stage_pair[3] = {};
  bool slot_u = false, slot_p = false;
  // host copies of the tables (kept until the next rebuild)
  std::vector<char> host;
  // captured quantum
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  bool warmed = false;  // one direct quantum ran for this membership
};

}  // namespace kw

namespace {

size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// The spectral stages of a step (kspaceFirstOrder3D.m:790-801, 905-907,
// 975-976/1001-1002), as step() builds them.
int step_stages(const kw_sim* s, LastArgs* st) {
  const Grid& g = s->g;
  const float invN = (float)(1.0 / (double)g.N);
  int n = 0;
  {
    LastArgs a{};
    a.in[0] = s->Pk;
    a.nin = 1;
    a.nout = g.dim;
    for (int c = 0; c < g.dim; ++c) {
      a.out[c] = s->G[c];
      a.outs[c] = SpecOut{0, OP_DPOS_K, c};
    }
    a.scale = invN;
    a.pair01 = gradient_pair(s);
    st[n++] = a;
  }
  {
    LastArgs a{};
    a.nin = g.dim;
    a.nout = g.dim;
    for (int c = 0; c < g.dim; ++c) {
      a.in[c] = s->G[c];
      a.out[c] = s->G[c];
      a.outs[c] = SpecOut{c, OP_DNEG_K, c};
    }
    a.scale = invN;
    st[n++] = a;
  }
  if (s->m.absorbing == 1) {
    LastArgs a{};
    a.nin = 2;
    a.nout = 2;
    a.in[0] = a.out[0] = s->G[0];
    a.in[1] = a.out[1] = s->G[1];
    a.outs[0] = SpecOut{0, OP_NABLA1, 0};
    a.outs[1] = SpecOut{1, OP_NABLA2, 0};
    a.scale = invN;
    st[n++] = a;
  }
  return n;
}

Fields3 fields_of(float2* const* f, int n) { return Fields3{{f[0], n > 1 ? f[1] : f[0], n > 2 ? f[2] : f[0]}}; }

// Builds the device tables for the current members (host t as given).
void batch_build(SimBatch* b, const std::vector<kw_sim*>& sims, hipStream_t st) {
  const int nb = (int)sims.size();
  const kw_sim* s0 = sims[0];
  LastArgs stages0[3];
  b->nstage = step_stages(s0, stages0);
  for (int k = 0; k < b->nstage; ++k) {
    b->stage_nin[k] = stages0[k].nin;
    b->stage_nout[k] = stages0[k].nout;
    b->stage_pair[k] = stages0[k].pair01;
  }
  const bool kxm = s0->g.kxm != 0;
  // layout
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const size_t o_clk = take(sizeof(BatchClock) * nb), o_live = take(sizeof(int) * nb);
  const size_t o_vel = take(sizeof(XVelJob) * nb), o_dens = take(sizeof(XDensJob) * nb);
  const size_t o_pres = take(sizeof(XPresJob) * nb);
  size_t o_plane[3] = {}, o_yin[3] = {}, o_last[3] = {}, o_yout[3] = {};
  for (int k = 0; k < b->nstage; ++k) {
    if (kxm) {
      o_plane[k] = take(sizeof(PlaneJob) * nb);
    } else {
      o_yin[k] = take(sizeof(YpassJob) * nb);
      o_last[k] = take(sizeof(LastJob) * nb);
      o_yout[k] = take(sizeof(YpassJob) * nb);
    }
  }
  const size_t bytes = off;
  if (bytes > b->table_cap) {
    // stream-ordered: no device-wide synchronisation while other batches run
    if (b->table) HIPCHK(hipFreeAsync(b->table, st));
    b->table = nullptr;
    const size_t cap = std::max(bytes, b->table_cap + b->table_cap / 2);
    HIPCHK(hipMallocAsync((void**)&b->table, cap, st));
    b->table_cap = cap;
    b->table_stream = st;
  }
  b->host.assign(bytes, 0);
  char* h = b->host.data();
  auto* clk = reinterpret_cast<BatchClock*>(h + o_clk);
  auto* vel = reinterpret_cast<XVelJob*>(h + o_vel);
  auto* dens = reinterpret_cast<XDensJob*>(h + o_dens);
  auto* pres = reinterpret_cast<XPresJob*>(h + o_pres);
  b->slot_u = b->slot_p = false;
  for (int j = 0; j < nb; ++j) {
    kw_sim* s = sims[j];
    const long long t1 = s->t + 1;  // the next step (the prologue rewrites the per-step fields)
    BatchClock& c = clk[j];
    c.t = s->t;
    c.Nt = s->Nt;
    c.start = s->start;
    const SrcData* sd[4] = {&s->usrc[0], &s->usrc[1], &s->usrc[2], &s->psrc};
    for (int k = 0; k < 4; ++k) {
      c.sig[k] = sd[k]->sig;
      c.flag[k] = sd[k]->flag;
      c.nseries[k] = sd[k]->nseries;
      c.single[k] = sd[k]->host_sig.empty() ? 0 : 1;
    }
    c.p_final = s->p_final;
    Sens us[3] = {};
    for (int a = 0; a < s->g.dim; ++a) us[a] = make_sens_u(s, t1, a);
    XVelJob& v = vel[j];
    v.g = s->g;
    v.tw = s->twX;
    v.m = s->m;
    for (int a = 0; a < 3; ++a) {
      v.A.G[a] = a < s->g.dim ? s->G[a] : nullptr;
      v.A.u[a] = a < s->g.dim ? s->u[a] : nullptr;
      v.A.src[a] = make_src(s, s->usrc[a], t1);
      v.A.sens[a] = us[a];
      if (s->usrc[a].flag && s->usrc[a].mode != 2) b->slot_u = true;
    }
    v.A.init = 0;
    const Sens ps = make_sens_p(s, t1, t1 == s->Nt);
    XDensJob& d = dens[j];
    d.g = s->g;
    d.tw = s->twX;
    d.m = s->m;
    for (int a = 0; a < 3; ++a) {
      d.A.G[a] = a < s->g.dim ? s->G[a] : nullptr;
      d.A.rho[a] = a < s->g.dim ? s->rho[a] : nullptr;
    }
    d.A.psrc = make_src(s, s->psrc, t1);
    if (s->psrc.flag && s->psrc.mode != 2) b->slot_p = true;
    d.A.srho = s->srho;
    d.A.pk = s->Pk;
    d.A.ps = ps;
    XPresJob& pr = pres[j];
    pr.g = s->g;
    pr.tw = s->twX;
    pr.m = s->m;
    pr.GA = s->G[0];
    pr.GB = s->G[1];
    pr.srho = s->srho;
    pr.pk = s->Pk;
    pr.ps = ps;
    LastArgs stages[3];
    step_stages(s, stages);
    for (int k = 0; k < b->nstage; ++k) {
      if (kxm) {
        reinterpret_cast<PlaneJob*>(h + o_plane[k])[j] = PlaneJob{s->g, s->twY, s->twL, stages[k], s->T};
      } else {
        reinterpret_cast<YpassJob*>(h + o_yin[k])[j] = YpassJob{s->g, s->twY, fields_of(stages[k].in, stages[k].nin)};
        reinterpret_cast<LastJob*>(h + o_last[k])[j] = LastJob{s->g, s->twL, stages[k], s->T};
        reinterpret_cast<YpassJob*>(h + o_yout[k])[j] =
            YpassJob{s->g, s->twY, fields_of(stages[k].out, stages[k].nout)};
      }
    }
  }
  char* dv = b->table;
  // pointers of the kernels' argument tables as offsets from their array
  // (the batched kernels rebase them; the step prologue keeps the encoding)
  for (int j = 0; j < nb; ++j) {
    visit_ptrs(vel[j], PtrEncode{dv + o_vel});
    visit_ptrs(dens[j], PtrEncode{dv + o_dens});
    visit_ptrs(pres[j], PtrEncode{dv + o_pres});
    for (int k = 0; k < b->nstage; ++k) {
      if (kxm) {
        visit_ptrs(reinterpret_cast<PlaneJob*>(h + o_plane[k])[j], PtrEncode{dv + o_plane[k]});
      } else {
        visit_ptrs(reinterpret_cast<YpassJob*>(h + o_yin[k])[j], PtrEncode{dv + o_yin[k]});
        visit_ptrs(reinterpret_cast<LastJob*>(h + o_last[k])[j], PtrEncode{dv + o_last[k]});
        visit_ptrs(reinterpret_cast<YpassJob*>(h + o_yout[k])[j], PtrEncode{dv + o_yout[k]});
      }
    }
  }
  b->clk = reinterpret_cast<BatchClock*>(dv + o_clk);
  b->live = reinterpret_cast<int*>(dv + o_live);
  b->vel = reinterpret_cast<XVelJob*>(dv + o_vel);
  b->dens = reinterpret_cast<XDensJob*>(dv + o_dens);
  b->pres = reinterpret_cast<XPresJob*>(dv + o_pres);
  for (int k = 0; k < 3; ++k) {
    b->plane[k] = k < b->nstage && kxm ? reinterpret_cast<PlaneJob*>(dv + o_plane[k]) : nullptr;
    b->yin[k] = k < b->nstage && !kxm ? reinterpret_cast<YpassJob*>(dv + o_yin[k]) : nullptr;
    b->lastj[k] = k < b->nstage && !kxm ? reinterpret_cast<LastJob*>(dv + o_last[k]) : nullptr;
    b->yout[k] = k < b->nstage && !kxm ? reinterpret_cast<YpassJob*>(dv + o_yout[k]) : nullptr;
  }
  HIPCHK(hipMemcpyAsync(b->table, h, bytes, hipMemcpyHostToDevice, st));
  b->members = sims;
  b->serials.resize(sims.size());
  b->t_expect.resize(sims.size());
  for (size_t j = 0; j < sims.size(); ++j) {
    b->serials[j] = sims[j]->serial;
    b->t_expect[j] = sims[j]->t;
  }
}

// One step of the batch: the prologue, then step()'s launch sequence for t >= 2
// (direct launches, or kernel nodes of the quantum's graph).
void batch_step(SimBatch* b, BatchSink& sk) {
  const kw_sim* s0 = b->members[0];
  const Grid& g = s0->g;
  const int nb = (int)b->members.size();
  launch_batch_clock(sk, b->clk, b->vel, b->dens, b->pres, b->live, nb);
  auto stage = [&](int k) {
    if (g.kxm) {
      if (!launch_plane_batch(sk, g, b->plane[k], b->live, nb, b->stage_nout[k], s0->plane_nt))
        throw HipError{hipErrorInvalidConfiguration, "batched plane pass"};
      return;
    }
    if (!launch_ypass_batch(sk, g, s0->axY.n, b->yin[k], b->live, nb, b->stage_nin[k], -1.f, s0->tly) ||
        !launch_last_batch(sk, g, s0->axL.n, b->lastj[k], b->live, nb, b->stage_nout[k] - b->stage_pair[k],
                           s0->tll) ||
        !launch_ypass_batch(sk, g, s0->axY.n, b->yout[k], b->live, nb, b->stage_nout[k], 1.f, s0->tly))
      throw HipError{hipErrorInvalidConfiguration, "batched pencil pass"};
  };
  stage(0);
  if (!launch_x_batch(sk, g, 0, b->vel, b->live, nb, false, b->slot_u, s0->xl))
    throw HipError{hipErrorInvalidConfiguration, "batched x velocity"};
  stage(1);
  const bool two = s0->m.absorbing == 1;
  if (!launch_x_batch(sk, g, 1, b->dens, b->live, nb, two, b->slot_p, s0->xl))
    throw HipError{hipErrorInvalidConfiguration, "batched x density"};
  if (two) {
    stage(2);
    if (!launch_x_batch(sk, g, 2, b->pres, b->live, nb, false, false, s0->xl))
      throw HipError{hipErrorInvalidConfiguration, "batched x pressure"};
  }
  HIPCHK(sk.err);
}

void batch_drop_graph(SimBatch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->graph_steps = 0;
}

}  // namespace

namespace kw {

std::string sim_batch_key(const kw_sim* s) { return s ? batch_key_of(s) : std::string(); }

SimBatch* batch_create() { return new SimBatch(); }

void batch_destroy(SimBatch* b) {
  if (!b) return;
  batch_drop_graph(b);
  if (b->table) {
    (void)hipFreeAsync(b->table, b->table_stream);
    (void)hipStreamSynchronize(b->table_stream);
  }
  delete b;
}

// Advances every simulation of `sims` (one batch key, none finished) by up to
// `nsteps` steps each (clamped to its Nt) on `st`.  The first step of a
// simulation with an initial pressure runs on its own first.  use_graph:
// replay a captured quantum once a membership has run one quantum directly.
// Asynchronous; a finished simulation's outputs are complete once `st` is.
kw_status batch_run(SimBatch* b, const std::vector<kw_sim*>& sims, int nsteps, hipStream_t st, bool use_graph) {
  if (!b || sims.empty() || nsteps < 1) return KW_OK;
  try {
    HIPCHK(hipSetDevice(sims[0]->device));
    const std::string key = batch_key_of(sims[0]);
    if (key.empty()) return fail(KW_ERR_UNSUPPORTED, "simulation cannot run batched");
    for (kw_sim* s : sims) {
      if (batch_key_of(s) != key) return fail(KW_ERR_INVALID, "batch members differ in shape or configuration");
      if (s->t >= s->Nt) return fail(KW_ERR_INVALID, "batch member has finished");
    }
    // the p0 start (3D.m:1021-1035) is a step of its own
    for (kw_sim* s : sims)
      if (s->t == 0 && s->p0) {
        s->last = st;
        step(s, st, 1, s->Nt == 1);
        s->t = 1;
      }
    std::vector<kw_sim*> live;
    for (kw_sim* s : sims)
      if (s->t < s->Nt) live.push_back(s);
    if (live.empty()) return KW_OK;
    bool same = live == b->members;
    for (size_t j = 0; same && j < live.size(); ++j)
      same = live[j]->serial == b->serials[j] && live[j]->t == b->t_expect[j];
    if (!same) {
      batch_drop_graph(b);
      b->warmed = false;
      batch_build(b, live, st);
    }
    long long q = 0;
    for (kw_sim* s : live) q = std::max<long long>(q, std::min<long long>(nsteps, s->Nt - s->t));
    const int steps = (int)q;
    if (use_graph && b->warmed) {
      if (!b->exec || b->graph_steps != steps) {
        // the quantum as an explicitly built chain of kernel nodes
        batch_drop_graph(b);
        hipGraph_t graph = nullptr;
        HIPCHK(hipGraphCreate(&graph, 0));
        BatchSink sk;
        sk.graph = graph;
        try {
          for (int i = 0; i < steps; ++i) batch_step(b, sk);
        } catch (...) {
          (void)hipGraphDestroy(graph);
          throw;
        }
        const hipError_t e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(e);
        b->graph_steps = steps;
      }
      HIPCHK(hipGraphLaunch(b->exec, st));
    } else {
      BatchSink sk;
      sk.stream = st;
      for (int i = 0; i < steps; ++i) batch_step(b, sk);
      b->warmed = true;
    }
    for (size_t j = 0; j < live.size(); ++j) {
      kw_sim* s = live[j];
      s->last = st;
      s->t = std::min<long long>(s->Nt, s->t + steps);
      b->t_expect[j] = s->t;
    }
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return fail(KW_ERR_DEVICE, "HIP error %d (%s) in %s", (int)e.e, hipGetErrorString(e.e), e.what);
  }
  return KW_OK;
}

// After the stream a simulation last ran on has been synchronised: later
// calls on the handle wait on its own stream again.
void sim_detach(kw_sim* s) {
  if (s) s->last = nullptr;
}

// For kw_noalloc.hip: a simulation whose every device buffer lies in
// [arena, arena + cap), run on `stream` (which may be the null stream).
kw_status sim_create_in(const kw_sim_desc* d, void* arena, size_t cap, hipStream_t stream, kw_sim** out) {
  return create_impl(d, static_cast<char*>(arena), cap, true, stream, out);
}
// Bytes such a simulation occupies in the arena (nothing is released there); 0 if invalid.
uint64_t sim_arena_bytes(const kw_sim_desc* d) {
  uint64_t peak = 0, total = 0;
  return footprint_plan(d, &peak, &total) ? total : 0;
}
kw_status set_error(kw_status st, const char* msg) {
  g_err = msg ? msg : "";
  return st;
}
}  // namespace kw
