// kw_sim_run_noalloc: one whole simulation (set-up, Nt steps, output copies)
// with every device buffer inside a fixed, statically allocated scratch
// region, for callers that cannot call hipMalloc.  It runs the kernels of
// kw_sim_create + kw_sim_run + kw_sim_fetch in the same order with the same
// launch shapes, so its outputs equal theirs bit for bit.
//
// The region is `scratch` below, a file-scope __device__ array: every GPU the
// library runs on holds one copy of it until the process ends.  Host code
// never passes the array's name to a kernel or a copy (on the host it is only
// a stand-in for the symbol); the device address of each GPU's copy comes
// from hipGetSymbolAddress, once per device, and is kept in g_scratch.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/kwave_hip.h"

namespace kw {
kw_status sim_create_in(const kw_sim_desc* d, void* arena, size_t cap, hipStream_t stream, kw_sim** out);
uint64_t sim_arena_bytes(const kw_sim_desc* d);
kw_status set_error(kw_status st, const char* msg);
}  // namespace kw

namespace {

constexpr size_t kScratchBytes = (size_t)KW_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
  std::mutex mu;       // one call at a time per device within the process
  bool ready = false;  // ptr looked up
  void* ptr = nullptr; // device address of this device's copy of `scratch`
};
DeviceScratch g_scratch[kMaxDevices];

kw_status fail_msg(kw_status st, const char* fmt, unsigned long long a = 0, unsigned long long b = 0) {
  char buf[256];
  std::snprintf(buf, sizeof buf, fmt, a, b);
  return kw::set_error(st, buf);
}

// Looks up this device's copy of the region (the device must be current).
kw_status scratch_ptr(int device, void** out) {
  DeviceScratch& ds = g_scratch[device];
  if (!ds.ready) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail_msg(KW_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed on device %llu (HIP error %llu)",
                      (unsigned long long)device, (unsigned long long)e);
    }
    ds.ptr = p;
    ds.ready = true;
  }
  *out = ds.ptr;
  return KW_OK;
}

}  // namespace

extern "C" {

kw_status kw_sim_run_noalloc(const kw_sim_desc* desc, void* stream, const kw_output_buffer* outputs,
                             int noutputs) {
  if (!desc) return kw::set_error(KW_ERR_INVALID, "null description");
  if (noutputs < 0 || (noutputs > 0 && !outputs)) return kw::set_error(KW_ERR_INVALID, "bad output list");
  const int dev = desc->device;
  if (dev < 0 || dev >= kMaxDevices) return fail_msg(KW_ERR_INVALID, "device %llu out of range", (unsigned long long)dev);
  const uint64_t need = kw::sim_arena_bytes(desc);
  if (need == 0) return kw::set_error(KW_ERR_INVALID, "invalid grid");
  if (need > kScratchBytes)
    return fail_msg(KW_ERR_UNSUPPORTED,
                    "simulation needs %llu device bytes; kw_sim_run_noalloc's scratch region holds %llu",
                    (unsigned long long)need, (unsigned long long)kScratchBytes);
  for (int i = 0; i < noutputs; ++i)
    if (!outputs[i].host_dst) return kw::set_error(KW_ERR_INVALID, "null output buffer");

  DeviceScratch& ds = g_scratch[dev];
  std::lock_guard<std::mutex> lk(ds.mu);
  const hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail_msg(KW_ERR_DEVICE, "hipSetDevice(%llu) failed", (unsigned long long)dev);
  }
  void* region = nullptr;
  kw_status st = scratch_ptr(dev, &region);
  if (st) return st;

  kw_sim* sim = nullptr;
  st = kw::sim_create_in(desc, region, kScratchBytes, (hipStream_t)stream, &sim);
  if (st) return st;
  st = kw_sim_run(sim, desc->Nt, stream);
  for (int i = 0; i < noutputs && st == KW_OK; ++i)
    st = kw_sim_fetch(sim, outputs[i].output_id, outputs[i].host_dst, outputs[i].bytes);
  // the region is free for the next call only once this call's work is done
  const hipError_t es = hipStreamSynchronize((hipStream_t)stream);
  if (es != hipSuccess && st == KW_OK) st = fail_msg(KW_ERR_DEVICE, "stream synchronisation failed (HIP error %llu)",
                                                     (unsigned long long)es);
  // frees nothing inside the region (and no stream: it is the caller's)
  const kw_status st2 = kw_sim_destroy(sim);
  return st ? st : st2;
}

}  // extern "C"
