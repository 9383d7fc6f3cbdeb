// ss_noalloc.hip -- raft_ss_radon_noalloc: the slant-stack Radon with its workspace in a
// static device array, for callers that cannot allocate (graph capture, serving loops).
//
// The array is a file-scope __device__ symbol.  Its name in host code is a host-side
// stand-in, not a device address, so the host never passes it to a kernel or a copy:
// the first call on each device (with that device current) looks up the device
// address with hipGetSymbolAddress and keeps it per device; every later use takes
// that address.  Each GPU holds its own copy of the array at its own address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/raft_noalloc.h"
#include "ss_internal.h"

static __device__ __align__(256) unsigned char scratch[RAFT_SS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be the current device.
int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices)
    return ss_set_error(RAFT_SS_EINVAL, "device %d outside the %d the allocation-free entry tracks", device,
                        kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return ss_set_error(RAFT_SS_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return 0;
}

}  // namespace

extern "C" int raft_ss_radon_noalloc(const raft_ss_plan* p, const float* d_image, float* d_out, int angle_major,
                                     int j0, int j1, int nslices, void* stream_) {
  int rc = check_plan(p, nslices);  // also: p's device is the current one
  if (rc) return rc;
  if ((rc = check_range(p, j0, j1))) return rc;
  if (!d_image || !d_out) return ss_set_error(RAFT_SS_EINVAL, "null device pointer");
  SSDevPlan dp = p->dev();
  // the plan's slab choice when an earlier call made one (this entry never times)
  hipStream_t stream = as_stream(stream_);
  if ((rc = ss_plan_radon_rows(p, j0, j1, nslices, stream, false, &dp))) return rc;
  const size_t layout = (size_t)nslices * p->g.n * p->g.pitch;
  const size_t layout_bytes = 2 * layout * sizeof(float2);
  const size_t need = layout_bytes + ss_radon_scratch(dp, j0, j1, nslices);
  if (need > (size_t)RAFT_SS_NOALLOC_SCRATCH_BYTES)
    return ss_set_error(RAFT_SS_EINVAL,
                        "radon_noalloc: %d slice(s) of %d x %d need %zu bytes of workspace, more than the %zu of "
                        "the static scratch array",
                        nslices, p->g.n, p->g.n, need, (size_t)RAFT_SS_NOALLOC_SCRATCH_BYTES);
  unsigned char* base = nullptr;
  if ((rc = scratch_address(p->device, &base))) return rc;
  float2* gA = reinterpret_cast<float2*>(base);
  float2* gB = gA + layout;
  float* slabs = reinterpret_cast<float*>(base + layout_bytes);
  HIP_TRY(ss_launch_gather_layouts(dp, d_image, gA, gB, nslices, stream));
  SSEmRatioArgs none{nullptr, nullptr};
  HIP_TRY(ss_launch_radon(dp, gA, gB, j0, j1, nslices, angle_major ? SS_OUT_SINOT : SS_OUT_SINO, d_out, none, slabs,
                          stream));
  return 0;
}
