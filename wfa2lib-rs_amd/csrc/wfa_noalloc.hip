// wfa_noalloc.hip — wfa_align_batch_noalloc: the batch engine for callers
// that cannot allocate device memory.
//
// Every buffer the batch engine would allocate (work-queue counters, retry
// lists, work order, wavefront rings, CIGAR history pool, headers and chunk
// tables) is carved from one fixed device array, `scratch`, of
// WFA_NOALLOC_SCRATCH_BYTES per GPU. The array is a file-scope __device__
// variable: the runtime gives every GPU that loads the library its own copy,
// which lives until the process ends. Host code never passes the array's
// name to a kernel (on the host it is only a stand-in for the symbol): the
// first call on a device looks up that device's address with
// hipGetSymbolAddress and keeps it, and every buffer is an offset from it.
//
// A call runs exactly the kernels wfa_engine_align_batch runs, with the
// same results; it only plans its launches within the array (fewer
// resident slots when the array is the limit, and WFA_STATUS_OOM for a pair
// whose history still does not fit after the wider passes, as with an
// engine under wfa_engine_set_scratch_budget). It calls no hipMalloc,
// hipHostMalloc, hipFree or hipEventCreate.
//
// All calls on one GPU share that GPU's array and every call leaves its
// contents behind. Calls are serialized per GPU by a lock, and each returns
// only after its kernels have finished, so within one process two calls
// never use the array at the same time; work a caller issued on another
// stream before the call is not waited for and does not touch the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wfa_device.h"
#include "wfa_engine.h"

static __device__ __align__(256) uint8_t scratch[WFA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

// One engine per device, built by the first call on that device without
// allocating: its buffers live in the device's copy of `scratch`.
struct DeviceSlot {
  std::once_flag once;
  int err = WFA_OK;
  wfa_engine_t engine;
  int32_t host_counter[16] = {};
};

DeviceSlot g_slots[kMaxDevices];

int init_slot(DeviceSlot& s, int device) {
  wfa_engine_t& e = s.engine;
  e.device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return WFA_ERR_DEVICE;
  e.num_cus = prop.multiProcessorCount;
  e.lds_max = prop.sharedMemPerBlock;
  // this device's copy of the array (the device is current here)
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr)
    return WFA_ERR_DEVICE;
  e.arena = static_cast<uint8_t*>(ptr);
  e.arena_cap = WFA_NOALLOC_SCRATCH_BYTES;
  // the work-queue counters sit at the start of the array; the launches
  // reset them with a kernel before use
  e.d_counters = reinterpret_cast<int32_t*>(e.arena);
  e.arena_top = 256;
  e.h_pinned = s.host_counter;
  e.max_len_limit = WFA_NOALLOC_MAX_LENGTH;
  // timing events stay null: the engine then skips them
  return WFA_OK;
}

}  // namespace

extern "C" int32_t wfa_align_batch_noalloc(const wfa_attr_t* attr, const wfa_batch_t* batch,
                                           void* hip_stream) {
  if (!attr || !batch) return WFA_ERR_INVALID_ARGUMENT;
  if (batch->num_pairs > WFA_NOALLOC_MAX_PAIRS || batch->max_length > WFA_NOALLOC_MAX_LENGTH)
    return WFA_ERR_LIMIT;
  // BiWFA keeps its own growing buffers: not in this entry point
  if (attr->memory_mode == WFA_MEMORY_ULTRALOW) return WFA_ERR_LIMIT;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return WFA_ERR_DEVICE;
  if (device < 0 || device >= kMaxDevices) return WFA_ERR_LIMIT;
  DeviceSlot& s = g_slots[device];
  std::call_once(s.once, [&]() { s.err = init_slot(s, device); });
  if (s.err != WFA_OK) return s.err;
  return wfahost::align_batch_locked(&s.engine, attr, batch, (hipStream_t)hip_stream);
}
