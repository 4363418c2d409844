// kg_static.hip -- kg_schedule_step_static: one Schedule::step of a flockers
// Field2D whose agents live in caller-owned device arrays, with no device
// allocation.
//
// Every buffer beyond the caller's inputs and outputs (the binned read
// buffer, the flock kernel's output, the bin counts, offsets and cursors,
// the scan partials, the per-tile flags, the error word, the step counter,
// the work-list prefix and the replica descriptor) is carved out of
// `scratch`, a file-scope device array of kScratchBytes.  Each GPU holds its
// own copy of it until the process ends.  Its device address is looked up
// once per device with hipGetSymbolAddress; host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// device symbol).
//
// Calls share their GPU's array: two calls that overlap on one GPU (on
// different streams or host threads) race on it.  Everything a call reads
// before writing is cleared by the call itself (bin counts, error word, step
// counter) with fill kernels, so the previous call's contents never leak.
#include <mutex>

#include "kg_common.hpp"

using namespace kg;

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
constexpr size_t kAlign = 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once).
int scratch_address(unsigned char** out) {
  int dev = 0;
  KG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(KG_E_INVALID, "device %d beyond the %d devices supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    KG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(KG_E_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return KG_OK;
}

// Bump carving of the scratch array; with base == nullptr it only counts.
struct Carver {
  unsigned char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t n) {
    used = (used + kAlign - 1) / kAlign * kAlign;
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += (n ? n : 1) * sizeof(T);
    return p;
  }
  Buf agents(size_t n) {
    Buf b;
    b.x = take<float>(n);
    b.y = take<float>(n);
    b.hx = take<float>(n);
    b.hy = take<float>(n);
    b.id = take<uint32_t>(n);
    return b;
  }
};

struct Carve {
  Buf R, W;
  uint32_t *count, *offset, *offset2, *cursor, *partial, *gflag, *err, *prefix;
  uint64_t* gstep;
  RepDev* rep;
};

size_t carve(unsigned char* base, uint32_t n, const Geom& g, Carve* c) {
  Carver k{base};
  const uint32_t nb = g.nbins;
  const uint32_t nscan = (nb + kScanTile - 1) / kScanTile;
  c->R = k.agents(n);
  c->W = k.agents(n);
  c->count = k.take<uint32_t>(nb);
  c->offset = k.take<uint32_t>(nb + 1);
  c->offset2 = k.take<uint32_t>(nb + 1);
  c->cursor = k.take<uint32_t>(nb);
  c->partial = k.take<uint32_t>(nscan + 1);
  c->gflag = k.take<uint32_t>(g.ntiles + 2);
  c->err = k.take<uint32_t>(1);
  c->prefix = k.take<uint32_t>(4);
  c->gstep = k.take<uint64_t>(1);
  c->rep = k.take<RepDev>(1);
  return k.used;
}

// Work-list prefix {0, tiles, 0, agents}, error word and step counter.
__global__ void k_static_setup(uint32_t* prefix, uint32_t ntiles, uint32_t n, uint32_t* err,
                               uint64_t* gstep) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    prefix[0] = 0u;
    prefix[1] = ntiles;
    prefix[2] = 0u;
    prefix[3] = n;
    *err = 0u;
    *gstep = 0ull;
  }
}

}  // namespace

extern "C" int kg_schedule_step_static(const kg_field_desc* desc, const kg_flock_params* p,
                                       uint64_t step, const uint32_t* ids, const float* x,
                                       const float* y, const float* hx, const float* hy,
                                       uint32_t n, uint32_t* out_ids, float* out_x,
                                       float* out_y, float* out_hx, float* out_hy,
                                       void* stream) {
  clear_error();
  if (!desc || !p) return set_error(KG_E_INVALID, "desc and params must not be NULL");
  if (desc->capacity == 0 || desc->capacity > (1u << 31))
    return set_error(KG_E_INVALID, "capacity must be in [1, 2^31]");
  if (n > desc->capacity)
    return set_error(KG_E_CAPACITY, "%u agents exceed the capacity %u", n, desc->capacity);
  if ((desc->bin_rows & (desc->bin_rows - 1)) || (desc->tile_cols & (desc->tile_cols - 1)) ||
      (desc->tile_rows & (desc->tile_rows - 1)))
    return set_error(KG_E_INVALID, "bin_rows, tile_cols and tile_rows must be 0 or powers of two");
  std::string why;
  if (!params_valid(*p, &why)) return set_error(KG_E_INVALID, "%s", why.c_str());
  // the same geometry kg_field_create makes from this description
  Geom g{};
  if (!make_geom_rt(desc->width, desc->height, desc->discretization, desc->toroidal,
                    desc->capacity, desc->tile_cols, desc->tile_rows, desc->bin_rows, &g, &why,
                    true))
    return set_error(KG_E_INVALID, "%s", why.c_str());
  Carve c{};
  const size_t need = carve(nullptr, n, g, &c);
  if (need > kScratchBytes)
    return set_error(KG_E_CAPACITY,
                     "%u agents on a %u-bin grid need %zu scratch bytes; the static scratch "
                     "holds %zu",
                     n, g.nbins, need, kScratchBytes);
  if (n == 0) return KG_OK;
  if (!ids || !x || !y || !hx || !hy || !out_ids || !out_x || !out_y || !out_hx || !out_hy)
    return set_error(KG_E_INVALID, "agent arrays must not be NULL");
  unsigned char* base = nullptr;
  if (int rc = scratch_address(&base)) return rc;
  carve(base, n, g, &c);

  hipStream_t s = static_cast<hipStream_t>(stream);
  RepDev h{};
  h.g = g;
  set_rep_params(&h, *p);
  h.bin_base = 0;
  h.step_base = step;
  const int lds_cap = clamp_lds_cap(pick_lds_cap(g, desc->capacity));

  WorkList wl;
  wl.reps = c.rep;
  wl.nreps = 1;
  wl.tile_prefix = c.prefix;
  wl.ntiles = g.ntiles;
  wl.app_prefix = c.prefix + 2;
  wl.napp = n;
  // the kernels take the input as a Buf; the append phases only read it
  Buf A;
  A.x = const_cast<float*>(x);
  A.y = const_cast<float*>(y);
  A.hx = const_cast<float*>(hx);
  A.hy = const_cast<float*>(hy);
  A.id = const_cast<uint32_t*>(ids);
  Buf out;
  out.x = out_x;
  out.y = out_y;
  out.hx = out_hx;
  out.hy = out_hy;
  out.id = out_ids;
  ScanState ss;
  ss.partial = c.partial;

  hipLaunchKernelGGL(k_static_setup, dim3(1), dim3(64), 0, s, c.prefix, g.ntiles, n, c.err,
                     c.gstep);
  KG_HIP(hipGetLastError());
  KG_HIP(launch_set_rep(c.rep, h, s));
  KG_HIP(launch_fill_u32(c.count, 0u, g.nbins, s));
  // set_object_location of every agent, then the lazy_update that makes them
  // the read buffer (field_2d.rs:838-846, 905-921)
  KG_HIP(launch_count_append(wl, A, c.count, c.err, s));
  KG_HIP(launch_scan(c.count, c.offset, c.cursor, g.nbins, ss, nullptr, s));
  KG_HIP(launch_scatter_append(wl, A, c.cursor, c.R, s));
  KG_HIP(launch_sort_bins(wl, c.R, c.offset, c.W, s));
  uint32_t e = 0;
  KG_HIP(hipMemcpyAsync(&e, c.err, sizeof(e), hipMemcpyDeviceToHost, s));
  KG_HIP(hipStreamSynchronize(s));
  if (e)  // the read buffer holds fewer than n agents: stop before the step
    return set_error(KG_E_RANGE, "an input location discretizes outside the bag grid");
  // Bird::step of every agent at schedule step `step` (schedule.rs:395-408),
  // then Flocker::update's lazy_update into the caller's output
  KG_HIP(launch_flock(wl, c.R, c.offset, c.W, c.count, c.gstep, c.err, lds_cap, c.gflag, n, s));
  KG_HIP(launch_scan(c.count, c.offset2, c.cursor, g.nbins, ss, nullptr, s));
  KG_HIP(launch_scatter_aligned(wl, c.R, c.W, c.offset, c.cursor, out, s));
  KG_HIP(launch_sort_bins(wl, out, c.offset2, c.W, s));
  KG_HIP(hipMemcpyAsync(&e, c.err, sizeof(e), hipMemcpyDeviceToHost, s));
  KG_HIP(hipStreamSynchronize(s));
  if (e)
    return set_error(KG_E_RANGE,
                     "a location discretizes outside the bag grid; the output is incomplete");
  return KG_OK;
}
