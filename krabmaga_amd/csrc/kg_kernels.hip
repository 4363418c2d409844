// kg_kernels.hip -- HIP kernels (gfx950, wave64) of the agent-step path.
//
//   init      Flocker::init on the device            tests/model/flockers/state.rs:41-56
//   count     bin histogram of appended agents       field_2d.rs:838-846 (set_object_location)
//   scan      exclusive prefix over bin counts       (bag layout, field_2d.rs:307)
//   scatter   agents to their bins (rebinning)       field_2d.rs:905-921 (lazy_update)
//   gather    rebinning of a flock step by tile      field_2d.rs:905-921 (lazy_update)
//   flock     fused neighbour query + Bird::step     field_2d.rs:386-440, bird.rs:38-154
//             + bin histogram of the new positions
//   query/... single-replica API queries             field_2d.rs:472-808
#include "kg_phases.hpp"
#include "kg_rt.hpp"

namespace kg {

// ---------------------------------------------------------------------------
// replica descriptor update (stream ordered)

__global__ void k_set_rep(RepDev* dst, RepDev v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

hipError_t launch_set_rep(RepDev* dst, const RepDev& v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_rep, dim3(1), dim3(64), 0, s, dst, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// init

__global__ void k_init_uniform(Buf A, uint32_t base, uint32_t n, float w, float h, uint64_t seed,
                               int heading_mode) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x, y, hx, hy;
  init_agent(i, w, h, seed, heading_mode, &x, &y, &hx, &hy);
  A.x[base + i] = x;
  A.y[base + i] = y;
  A.hx[base + i] = hx;
  A.hy[base + i] = hy;
  A.id[base + i] = i;
}

hipError_t launch_init_uniform(Buf A, uint32_t base, uint32_t n, float w, float h, uint64_t seed,
                               int heading_mode, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_init_uniform, dim3((n + 255) / 256), dim3(256), 0, s, A, base, n, w, h, seed,
                     heading_mode);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// count / scatter of appended agents (set_object_location and init)

__global__ void k_count_append(const RepDev* __restrict__ reps, int nreps,
                               const uint32_t* __restrict__ app_prefix, uint32_t napp, Buf A,
                               uint32_t* __restrict__ count, uint32_t* __restrict__ err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < napp) count_append_one(reps, nreps, app_prefix, A, i, count, err);
}

hipError_t launch_count_append(const WorkList& wl, Buf A, uint32_t* count, uint32_t* err,
                               hipStream_t s) {
  if (wl.napp == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_append, dim3((wl.napp + 255) / 256), dim3(256), 0, s, wl.reps,
                     wl.nreps, wl.app_prefix, wl.napp, A, count, err);
  return hipGetLastError();
}

__global__ void k_scatter_append(const RepDev* __restrict__ reps, int nreps,
                                 const uint32_t* __restrict__ app_prefix, uint32_t napp, Buf A,
                                 uint32_t* __restrict__ cursor, Buf R2) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < napp) scatter_append_one(reps, nreps, app_prefix, A, i, cursor, R2);
}

hipError_t launch_scatter_append(const WorkList& wl, Buf A, uint32_t* cursor, Buf R2,
                                 hipStream_t s) {
  if (wl.napp == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_append, dim3((wl.napp + 255) / 256), dim3(256), 0, s, wl.reps,
                     wl.nreps, wl.app_prefix, wl.napp, A, cursor, R2);
  return hipGetLastError();
}

// Rebinning of the write buffer the flock kernel produced: it is aligned with
// the current read buffer, so work item (replica, tile) moves exactly the
// agents of that tile's old range.  fill_base null: cursor[b] is bin b's next
// free slot; else cursor[b] is bin b's count of agents still to place, and
// they go to fill_base[b] + count - 1 downwards (the count ends at zero).
__device__ __forceinline__ void scatter_aligned_tiles(const RepDev* __restrict__ reps, int nreps,
                                                      const uint32_t* __restrict__ tile_prefix,
                                                      Buf R, Buf W,
                                                      const uint32_t* __restrict__ offset_cur,
                                                      uint32_t* __restrict__ cursor, Buf R2,
                                                      const uint32_t* __restrict__ fill_base,
                                                      uint32_t nitems) {
  for (uint32_t bi = blockIdx.x; bi < nitems; bi += gridDim.x) {
    const uint32_t item = xcd_remap(bi, nitems);
    const TileRange tr = tile_range(reps, nreps, tile_prefix, item, offset_cur);
    const RepDev& rp = reps[tr.r];
    const Geom& g = rp.g;
    const uint32_t lane = threadIdx.x & (kWave - 1);
    // The read buffer is in bin order and birds move little, so neighbouring
    // lanes mostly share a destination bin: one cursor atomic per run of
    // equal bins in a wave (the bin sort that follows fixes the order inside
    // a bin).  Every lane takes part in every round for the wave operations.
    for (uint32_t i0 = tr.lo; i0 < tr.hi; i0 += blockDim.x) {
      const uint32_t i = i0 + threadIdx.x;
      bool valid = i < tr.hi;
      float x = 0.0f, y = 0.0f;
      uint32_t bin = 0xFFFFFFFFu;
      if (valid) {
        x = W.x[i];
        y = W.y[i];
        int32_t cx, cy;
        discretize(g, x, y, &cx, &cy);
        valid = cell_valid(g, cx, cy);  // an invalid cell was reported by the flock kernel
        if (valid) bin = rp.bin_base + bin_of(g, cx, cy);
      }
      const uint32_t prev = __shfl_up(bin, 1, kWave);
      const bool head = valid && (lane == 0 || prev != bin);
      const uint64_t heads = __ballot(head);
      const uint64_t vmask = __ballot(valid);
      const uint64_t upto = lane == kWave - 1 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
      const uint64_t below = heads & upto;
      const int h = below ? 63 - __clzll((long long)below) : 0;
      // the run of head h ends at the next head
      const uint64_t after = heads & ~upto;
      const int nh = after ? __ffsll((long long)after) - 1 : kWave;
      const uint64_t from_h = ~((1ull << h) - 1ull);
      const uint64_t to_nh = nh == kWave ? ~0ull : ((1ull << nh) - 1ull);
      uint32_t base = 0;
      if (head) {
        const uint32_t run = (uint32_t)__popcll(vmask & from_h & to_nh);
        base = fill_base ? fill_base[bin] + atomicSub(&cursor[bin], run) - run
                         : atomicAdd(&cursor[bin], run);
      }
      base = __shfl(base, h, kWave);
      if (valid) {
        // base plus the valid lanes of the run before this one
        const uint32_t pos = base + (uint32_t)__popcll(vmask & from_h & (upto >> 1));
        R2.x[pos] = x;
        R2.y[pos] = y;
        R2.hx[pos] = W.hx[i];
        R2.hy[pos] = W.hy[i];
        R2.id[pos] = R.id[i];
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_scatter_aligned(const RepDev* __restrict__ reps, int nreps,
                                                         const uint32_t* __restrict__ tile_prefix,
                                                         Buf R, Buf W,
                                                         const uint32_t* __restrict__ offset_cur,
                                                         uint32_t* __restrict__ cursor, Buf R2,
                                                         uint32_t nitems) {
  scatter_aligned_tiles(reps, nreps, tile_prefix, R, W, offset_cur, cursor, R2, nullptr,
                        nitems);
}

// Work-groups of the gated fallback: in most steps they only read the gate
// (a launch costs in proportion to its work-groups), and in a step they redo
// they loop over the tiles.
constexpr uint32_t kGatedGrid = 256;

hipError_t launch_scatter_aligned(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                                  uint32_t* cursor, Buf R2, hipStream_t s) {
  if (wl.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_aligned, dim3(wl.ntiles), dim3(256), 0, s, wl.reps, wl.nreps,
                     wl.tile_prefix, R, W, offset_cur, cursor, R2, wl.ntiles);
  return hipGetLastError();
}

// The gather rebinning: one work-group per tile, phases in kg_phases.hpp.
// Each thread keeps its candidates in registers from load to store, so a
// bird is read once and written once; a tile with more candidates (or new
// birds) than kGatherCap sends the step to the scatter and the bin sort.
// With sc.count the work-group also does the scan's part for its tile: its
// bins' new offsets follow from the earlier scan tiles' sums and the counts
// of the earlier bins of its own scan tile (read from this step's counts,
// which nothing clears in this step), and it writes them to offset_new.
__global__ __launch_bounds__(kGatherThreads) void k_rebin_gather(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, Buf R,
    Buf W, const uint32_t* __restrict__ offset_cur, uint32_t* __restrict__ offset_new, Outbox ob,
    Buf R2, Gate gate, GatherScan sc) {
  __shared__ uint32_t s_warp[kGatherThreads / kWave];
  __shared__ uint32_t s_bstart[kGatherMaxBins + 1];
  __shared__ uint32_t s_fill[kGatherMaxBins];
  __shared__ uint64_t s_key[kGatherCap];
  __shared__ uint32_t s_sitem[kGatherSources], s_scnt[kGatherSources];
  __shared__ uint32_t s_fail;
  const uint32_t t = threadIdx.x;
  const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
  const int r = find_rep(tile_prefix, nreps, item);
  const RepDev& rp = reps[r];
  const Geom& g = rp.g;
  const uint32_t tile = item - tile_prefix[r];
  const uint32_t nbt = 1u << (g.tx_log2 + g.tr_log2);
  const uint32_t tb0 = rp.bin_base + tile * nbt;
  const uint32_t lo = offset_cur[tb0], nown = offset_cur[tb0 + nbt] - lo;
  if (gate.bar && item == 0 && t == 0) *gate.bar = 0u;  // for the fallback's grid wait
  uint32_t nlo;
  if (sc.count) {
    // before the tile: the earlier scan tiles' sums and the earlier bins of
    // its own scan tile; one bin per thread (nbt <= kGatherThreads)
    const uint32_t q = tb0 / (uint32_t)kScanTile;
    uint32_t part = 0;
    if (!sc.prefixed)
      for (uint32_t i = t; i < q; i += kGatherThreads) part += sc.tile_sum[i];
    else if (t == 0)
      part = sc.tile_sum[q];
    for (uint32_t b = q * (uint32_t)kScanTile + t; b < tb0; b += kGatherThreads)
      part += sc.count[b];
    block_exclusive_scan<kGatherThreads>(part, s_warp, &nlo);
    const uint32_t c = t < nbt ? sc.count[tb0 + t] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan<kGatherThreads>(c, s_warp, &tot);
    if (t < nbt) {
      s_bstart[t] = ex;
      offset_new[tb0 + t] = nlo + ex;
      if (sc.cursor) sc.cursor[tb0 + t] = nlo + ex;
      sc.count_next[tb0 + t] = 0u;
    }
    if (tile == (uint32_t)g.ntiles - 1u) {
      // bins between this replica's last and the next replica's first (on
      // the next scan tile) hold nobody: they start where this tile ends
      const uint32_t pad_end = min(sc.nbins, (tb0 + nbt + (uint32_t)kScanTile - 1u) /
                                                 (uint32_t)kScanTile * (uint32_t)kScanTile);
      for (uint32_t b = tb0 + nbt + t; b < pad_end; b += kGatherThreads) offset_new[b] = nlo + tot;
    }
    if (t == 0) {
      s_bstart[nbt] = tot;
      if (item == sc.last_item) offset_new[sc.nbins] = nlo + tot;
      if (tb0 % (uint32_t)kScanTile == 0) sc.tile_sum_next[q] = 0u;
      if (item == 0) {
        *sc.gstep += 1;
        if (gate.clear) *gate.clear = 0u;
      }
    }
  } else {
    nlo = offset_new[tb0];
    for (uint32_t b = t; b <= nbt; b += kGatherThreads) s_bstart[b] = offset_new[tb0 + b] - nlo;
  }
  if (t < kGatherSources) {
    const int32_t sl = gather_source(g, tile, t);
    const uint32_t it = sl >= 0 ? tile_prefix[r] + (uint32_t)sl : 0u;
    s_sitem[t] = it;
    s_scnt[t] = sl >= 0 ? ob.n[it] : 0u;
  }
  if (t == 0) s_fail = 0u;
  __syncthreads();
  uint32_t ncand = nown;
  for (uint32_t q = 0; q < kGatherSources; ++q) ncand += s_scnt[q];
  if (t == 0 && (ncand > kGatherCap || s_bstart[nbt] > kGatherCap)) s_fail = 1u;
  if (t < kGatherSources && s_scnt[t] > kOutboxCap) s_fail = 1u;
  for (uint32_t b = t; b < nbt; b += kGatherThreads) {
    if (s_bstart[b + 1] - s_bstart[b] > kGatherBinMax) s_fail = 1u;
    s_fill[b] = 0u;
  }
  __syncthreads();
  // s_fail is read only between barriers, so every thread takes the same path
  if (!s_fail) {
    // per candidate: (bin << 16 | slot in the bin) or ~0, and the bird
    uint32_t cbs[kGatherRegs], cid[kGatherRegs];
    float cx[kGatherRegs], cy[kGatherRegs], chx[kGatherRegs], chy[kGatherRegs];
#pragma unroll
    for (uint32_t u = 0; u < kGatherRegs; ++u) {
      cbs[u] = ~0u;
      const uint32_t k = t + u * kGatherThreads;
      if (k < ncand) {
        const uint32_t src = gather_src(ob, s_sitem, s_scnt, lo, nown, k);
        cx[u] = W.x[src];
        cy[u] = W.y[src];
        chx[u] = W.hx[src];
        chy[u] = W.hy[src];
        cid[u] = R.id[src];
        int32_t row;
        const uint32_t b = gather_bin(g, tile, cx[u], cy[u], &row);
        if (b < nbt) {
          const uint32_t slot = atom_add(&s_fill[b], 1u);
          if (slot < s_bstart[b + 1] - s_bstart[b]) {
            s_key[s_bstart[b] + slot] = ((uint64_t)(uint32_t)row << 32) | cid[u];
            cbs[u] = (b << 16) | slot;
          } else {
            s_fail = 1u;
          }
        }
      }
    }
    __syncthreads();
    for (uint32_t b = t; b < nbt; b += kGatherThreads)
      if (s_fill[b] != s_bstart[b + 1] - s_bstart[b]) s_fail = 1u;
    __syncthreads();
    if (!s_fail) {
#pragma unroll
      for (uint32_t u = 0; u < kGatherRegs; ++u) {
        if (cbs[u] == ~0u) continue;
        const uint32_t b = cbs[u] >> 16;
        const uint32_t e = s_bstart[b] + (cbs[u] & 0xFFFFu);
        const uint32_t pos = nlo + s_bstart[b] + gather_rank(s_key, s_bstart[b], s_bstart[b + 1], e,
                                                             s_key[e]);
        R2.x[pos] = cx[u];
        R2.y[pos] = cy[u];
        R2.hx[pos] = chx[u];
        R2.hy[pos] = chy[u];
        R2.id[pos] = cid[u];
      }
    }
  }
  if (t == 0 && s_fail) *gate.flag = 1u;
}

hipError_t launch_rebin_gather(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                               uint32_t* offset_new, const Outbox& ob, Buf R2, const Gate& gate,
                               const GatherScan& sc, hipStream_t s) {
  if (wl.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rebin_gather, dim3(wl.ntiles), dim3(kGatherThreads), 0, s, wl.reps,
                     wl.nreps, wl.tile_prefix, R, W, offset_cur, offset_new, ob, R2, gate, sc);
  return hipGetLastError();
}

__global__ void k_materialize(Buf R, Buf W, uint32_t n, Buf A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  A.x[i] = W.x[i];
  A.y[i] = W.y[i];
  A.hx[i] = W.hx[i];
  A.hy[i] = W.hy[i];
  A.id[i] = R.id[i];
}

hipError_t launch_materialize(Buf R, Buf W, uint32_t n, Buf A, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_materialize, dim3((n + 255) / 256), dim3(256), 0, s, R, W, n, A);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// scan: exclusive prefix over the bin counts in two launches (tile sums,
// then tile scans, each tile adding up the sums of the tiles before it), or,
// past kScanInlineTiles tiles (where those sums would grow with the square
// of the tile count), three: a one-block scan of the tile sums between them.
// No work-group waits on another, so nothing depends on inter-work-group
// ordering.  The last launch also zeroes the counts (for the next
// histogram), fills the scatter cursors, clears the step's gate flag and
// advances the schedule-step counter.

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ count,
                                                              uint32_t n,
                                                              uint32_t* __restrict__ partial) {
  __shared__ uint32_t s_warp[kScanThreads / kWave];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) sum += count[base + k];
  uint32_t total;
  block_exclusive_scan<kScanThreads>(sum, s_warp, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One block: partial[0..m) becomes its exclusive prefix, in segments of
// kPrefixSeg entries.  A segment is loaded coalesced (thread t takes entries
// t, t + 1024, ...) into LDS, each thread scans a run of kPrefixRun
// consecutive entries from there (the padding word after every 32 keeps the
// runs' reads on distinct banks), one block scan gives the runs' bases, and
// the prefixes go back out coalesced.  The next segment's loads are issued
// before the current one is scanned, so one round trip to memory overlaps
// the LDS work.
constexpr uint32_t kPrefixRun = 8;
constexpr uint32_t kPrefixSeg = 1024 * kPrefixRun;

__device__ __forceinline__ uint32_t prefix_pad(uint32_t i) { return i + (i >> 5); }

__global__ __launch_bounds__(1024) void k_scan_partials(uint32_t* __restrict__ partial,
                                                        uint32_t m) {
  __shared__ uint32_t s_warp[1024 / kWave];
  __shared__ uint32_t s_seg[kPrefixSeg + kPrefixSeg / 32];
  const uint32_t t = threadIdx.x;
  uint32_t nxt[kPrefixRun];
#pragma unroll
  for (uint32_t u = 0; u < kPrefixRun; ++u) {
    const uint32_t i = u * 1024u + t;
    nxt[u] = i < m ? partial[i] : 0u;
  }
  uint32_t carry = 0;
  for (uint32_t base = 0; base < m; base += kPrefixSeg) {
#pragma unroll
    for (uint32_t u = 0; u < kPrefixRun; ++u) s_seg[prefix_pad(u * 1024u + t)] = nxt[u];
    const uint32_t nb = base + kPrefixSeg;
#pragma unroll
    for (uint32_t u = 0; u < kPrefixRun; ++u) {
      const uint32_t i = nb + u * 1024u + t;
      nxt[u] = i < m ? partial[i] : 0u;
    }
    __syncthreads();
    uint32_t v[kPrefixRun], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPrefixRun; ++k) {
      v[k] = s_seg[prefix_pad(t * kPrefixRun + k)];
      sum += v[k];
    }
    uint32_t total;
    uint32_t run = carry + block_exclusive_scan<1024>(sum, s_warp, &total);
#pragma unroll
    for (uint32_t k = 0; k < kPrefixRun; ++k) {
      s_seg[prefix_pad(t * kPrefixRun + k)] = run;
      run += v[k];
    }
    carry += total;
    __syncthreads();
#pragma unroll
    for (uint32_t u = 0; u < kPrefixRun; ++u) {
      const uint32_t i = u * 1024u + t;
      if (base + i < m) partial[base + i] = s_seg[prefix_pad(i)];
    }
    __syncthreads();
  }
}

// prefixed: partial[] already holds each tile's exclusive prefix.
__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t* __restrict__ count,
                                                            uint32_t* __restrict__ offset,
                                                            uint32_t* __restrict__ cursor, uint32_t n,
                                                            const uint32_t* __restrict__ partial,
                                                            int prefixed, uint64_t* gstep,
                                                            uint32_t* gate) {
  __shared__ uint32_t s_warp[kScanThreads / kWave];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  // the tiles before this one (at most kScanInlineTiles sums, cached)
  uint32_t before = 0;
  if (!prefixed)
    for (uint32_t q = threadIdx.x; q < blockIdx.x; q += kScanThreads) before += partial[q];
  else if (threadIdx.x == 0)
    before = partial[blockIdx.x];
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? count[base + k] : 0u;
    sum += v[k];
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) count[base + k] = 0u;
  uint32_t tile_before;
  block_exclusive_scan<kScanThreads>(before, s_warp, &tile_before);
  uint32_t total;
  uint32_t run = tile_before + block_exclusive_scan<kScanThreads>(sum, s_warp, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) {
      offset[base + k] = run;
      cursor[base + k] = run;
    }
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) offset[n] = tile_before + total;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (gstep) *gstep += 1;
    if (gate) *gate = 0u;
  }
}

hipError_t launch_scan_prefix(uint32_t* sums, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, sums, n);
  return hipGetLastError();
}

hipError_t launch_scan(uint32_t* count, uint32_t* offset, uint32_t* cursor, uint32_t nbins,
                       const ScanState& ss, uint64_t* gstep, hipStream_t s, uint32_t* gate) {
  const uint32_t ntiles = (nbins + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntiles), dim3(kScanThreads), 0, s, count, nbins,
                     ss.partial);
  const int prefixed = ntiles > kScanInlineTiles ? 1 : 0;
  if (prefixed)
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, ss.partial, ntiles);
  hipLaunchKernelGGL(k_scan_down, dim3(ntiles), dim3(kScanThreads), 0, s, count, offset, cursor,
                     nbins, ss.partial, prefixed, gstep, gate);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// flock: fused neighbour query + Bird::step, one block per (replica, tile);
// phase bodies in kg_phases.hpp.

size_t flock_lds_bytes(int cap) { return flock_smem_bytes(cap); }

// A tile whose staged set exceeds the LDS capacity runs in chunks of own
// columns (each with its halo) one after another; a tile with an own column
// whose staged items alone exceed it takes the global path: it sets its flag
// in `gflag`, and k_flock_global then steps its birds one thread per bird, so
// a crowded tile is spread over the whole GPU instead of one work-group.
// Such a tile also stamps gflag[ntiles] with the step (step + 1, so a zeroed
// word matches no step), which lets k_flock_global return at once in a step
// without one, and counts itself in gflag[ntiles + 1].  With kInlineGlobal
// (no k_flock_global launch) the work-group steps its tile's birds itself
// (compiled apart: the global walk's registers would cost the LDS path
// occupancy).  With tile_sum (the scan's per-scan-tile sums, tiles of at
// most kGatherMaxBins bins) the LDS path counts its stayers in LDS and adds
// them to the bins and the sum once at its end.
template <bool kInlineGlobal>
__global__ __launch_bounds__(kFlockThreads) void k_flock(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, Buf R,
    const uint32_t* __restrict__ offset, Buf W, uint32_t* __restrict__ count_next,
    const uint64_t* __restrict__ gstep, uint32_t* __restrict__ err, int lds_cap,
    uint32_t* __restrict__ gflag, Outbox ob, uint32_t* __restrict__ tile_sum) {
  extern __shared__ __align__(16) uint8_t smem_raw[];
  __shared__ uint32_t s_seg_start[3 * kMaxStageCols];
  __shared__ uint32_t s_seg_len[3 * kMaxStageCols];
  __shared__ uint32_t s_col_base[kMaxStageCols + 1];
  __shared__ uint32_t s_warp[kFlockThreads / kWave];
  __shared__ uint32_t s_nown;
  __shared__ uint32_t s_outn;
  __shared__ uint32_t s_hist[kGatherMaxBins];

  const int t = threadIdx.x;
  const uint32_t item = xcd_remap(blockIdx.x, gridDim.x);
  const FlockTile ft = flock_tile(reps, nreps, tile_prefix, item, *gstep);
  const RepDev& rp = reps[ft.r];
  const uint32_t nbt = ft.tb1 - ft.tb0;
  if (tile_sum)
    for (uint32_t b = t; b < nbt; b += kFlockThreads) s_hist[b] = 0u;

  bool use_lds = rp.lds_ok != 0 && ft.ncols <= kMaxStageCols;
  uint32_t L = 0;
  if (use_lds) {
    const uint32_t len =
        t < ft.ncols ? flock_segments(rp, ft, t, offset, s_seg_start, s_seg_len) : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan<kFlockThreads>(len, s_warp, &total);
    if (t < ft.ncols) s_col_base[t] = ex;
    if (t == 0) {
      s_col_base[ft.ncols] = total;
      s_nown = 0;
      s_outn = 0;
    }
    L = total;
    __syncthreads();
  }
  const int32_t own_lo = ft.c0 - ft.sc_lo, own_hi = ft.c1 - ft.sc_lo;
  if (use_lds && L > (uint32_t)lds_cap) {
    // chunks need every own column to fit on its own (the same on all
    // lanes); a tile of many chunks is better spread over the GPU
    use_lds = L <= kFlockChunkFactor * (uint32_t)lds_cap;
    for (int32_t c = own_lo; use_lds && c < own_hi; ++c)
      if (flock_chunk_size(rp, ft, s_col_base, c, c + 1) > (uint32_t)lds_cap) use_lds = false;
  }
  if (t == 0) {
    gflag[item] = use_lds ? 0u : 1u;
    if (!use_lds) {
      gflag[gridDim.x] = (uint32_t)*gstep + 1u;
      atomicAdd(&gflag[gridDim.x + 1], 1u);
      if (ob.n) ob.n[item] = 0u;  // the global path lists the leavers
    }
  }
  if (use_lds) {
    const FlockSinks lv{ob, item - ft.tile, &s_outn, tile_sum ? s_hist : nullptr, tile_sum};
    const FlockSmem S = flock_carve(smem_raw, lds_cap);
    for (int32_t oa = own_lo; oa < own_hi;) {
      const int32_t ob = L <= (uint32_t)lds_cap
                             ? own_hi
                             : flock_chunk_end(rp, ft, s_col_base, oa, own_hi, (uint32_t)lds_cap);
      int32_t sa, sb;
      flock_chunk_cols(rp, ft, oa, ob, &sa, &sb);
      const uint32_t base = s_col_base[sa], n = s_col_base[sb] - base;
      for (uint32_t k = t; k < n; k += kFlockThreads) {
        flock_stage(rp, ft, base + k, k, R, s_seg_start, s_seg_len, s_col_base, S);
        flock_own(ft, k, oa, ob, S, &s_nown);
      }
      __syncthreads();
      const uint32_t nown = s_nown;
      for (uint32_t o = t; o < nown; o += kFlockThreads)
        flock_agent_lds(rp, ft, S.own[o], s_col_base, base, S, W, count_next, err, lv);
      oa = ob;
      if (oa < own_hi) {
        // the next chunk reuses the LDS
        __syncthreads();
        if (t == 0) s_nown = 0;
        __syncthreads();
      }
    }
    if (ob.n || tile_sum) __syncthreads();
    if (ob.n && t == 0) ob.n[item] = s_outn;
    if (tile_sum) {
      // the stayers into their bins; all of them into the one scan tile
      // that holds the tile's bins, one atomic per wave
      uint32_t h = 0;
      for (uint32_t b = t; b < nbt; b += kFlockThreads) {
        const uint32_t v = s_hist[b];
        if (v) atomicAdd(&count_next[ft.tb0 + b], v);
        h += v;
      }
      for (int off = kWave / 2; off > 0; off >>= 1) h += __shfl_xor(h, off, kWave);
      if ((t & (kWave - 1)) == 0 && h) atomicAdd(&tile_sum[ft.tb0 / (uint32_t)kScanTile], h);
    }
  } else if (kInlineGlobal) {
    __syncthreads();  // the outbox count is cleared
    const uint32_t lo = offset[ft.tb0], hi = offset[ft.tb1];
    const FlockSinks lv{ob, item - ft.tile, nullptr, nullptr, tile_sum};
    for (uint32_t gi = lo + (uint32_t)t; gi < hi; gi += kFlockThreads)
      flock_agent_global(rp, ft.step, gi, R, offset, W, count_next, err, lv);
  }
}

// The birds of flagged tiles, one thread per read-buffer slot.  The read
// buffer holds the replicas one after another (replica r at
// app_prefix[r], its tiles one after another), each tile's bins in order.
// kLeavers: outboxes are filled (compiled apart: the listing costs the walk
// registers).  kStride: each work-group takes every gridDim.x-th run of
// kFlockThreads birds (a grid much smaller than the runs is cheap to launch
// in the steps no tile takes this path; the loop costs registers), else
// work-group i takes run i.
template <bool kLeavers, bool kStride>
__global__ __launch_bounds__(kFlockThreads) void k_flock_global(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix,
    const uint32_t* __restrict__ app_prefix, Buf R, const uint32_t* __restrict__ offset, Buf W,
    uint32_t* __restrict__ count_next, const uint64_t* __restrict__ gstep,
    uint32_t* __restrict__ err, const uint32_t* __restrict__ gflag, uint32_t ntiles,
    uint32_t n, Outbox ob, uint32_t* __restrict__ tile_sum) {
  __shared__ uint32_t s_any;
  // no tile took the global path in this step
  if (gflag[ntiles] != (uint32_t)*gstep + 1u) return;
  // global tile of read-buffer slot q, or false for a bird outside its grid
  auto tile_at = [&](uint32_t q, uint32_t* t) {
    const int r = nreps == 1 ? 0 : find_rep(app_prefix, nreps, q);
    const Geom& g = reps[r].g;
    int32_t cx, cy;
    discretize(g, R.x[q], R.y[q], &cx, &cy);
    if (!cell_valid(g, cx, cy)) return false;
    *t = tile_prefix[r] + tile_of(g, cx, cy);
    return true;
  };
  auto one_run = [&](uint32_t run) {
    const uint32_t g0 = run * kFlockThreads;
    const uint32_t gi = g0 + threadIdx.x;
    // the run's birds belong to the tiles from its first bird's to its last
    // bird's: skip the run unless one of those took the global path
    if (threadIdx.x == 0) {
      const uint32_t g1 = min(n, g0 + kFlockThreads) - 1u;
      uint32_t t0 = 0, t1 = 0, any = 1u;  // an out-of-grid bird: let every thread look
      if (tile_at(g0, &t0) && tile_at(g1, &t1) && t1 - t0 < kFlockThreads) {
        any = 0u;
        for (uint32_t t = t0; t <= t1; ++t) any |= gflag[t];
      }
      s_any = any;
    }
    __syncthreads();
    const uint32_t any = s_any;
    if (kStride) __syncthreads();  // s_any is rewritten for the next run
    uint32_t t;
    if (!any || gi >= n || !tile_at(gi, &t) || !gflag[t]) return;
    const int r = nreps == 1 ? 0 : find_rep(app_prefix, nreps, gi);
    const RepDev& rp = reps[r];
    flock_agent_global(rp, rp.step_base + *gstep, gi, R, offset, W, count_next, err,
                       FlockSinks{kLeavers ? ob : Outbox{}, tile_prefix[r], nullptr, nullptr,
                                  tile_sum});
  };
  if (kStride) {
    const uint32_t nruns = (n + kFlockThreads - 1) / kFlockThreads;
    for (uint32_t run = blockIdx.x; run < nruns; run += gridDim.x) one_run(run);
  } else {
    one_run(blockIdx.x);
  }
}

hipError_t launch_flock(const WorkList& wl, Buf R, const uint32_t* offset, Buf W,
                        uint32_t* count_next, const uint64_t* gstep, uint32_t* err, int lds_cap,
                        uint32_t* gflag, uint32_t nagents, hipStream_t s, const Outbox& ob,
                        bool inline_global, uint32_t* tile_sum, uint32_t global_grid) {
  if (wl.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(inline_global ? k_flock<true> : k_flock<false>, dim3(wl.ntiles),
                     dim3(kFlockThreads), flock_lds_bytes(lds_cap), s, wl.reps, wl.nreps,
                     wl.tile_prefix, R, offset, W, count_next, gstep, err, lds_cap, gflag, ob,
                     tile_sum);
  const uint32_t nruns = (nagents + kFlockThreads - 1) / kFlockThreads;
  if (nagents && !inline_global) {
    const bool stride = global_grid && global_grid < nruns;
    auto kern = stride ? (ob.n ? k_flock_global<true, true> : k_flock_global<false, true>)
                       : (ob.n ? k_flock_global<true, false> : k_flock_global<false, false>);
    hipLaunchKernelGGL(kern, dim3(stride ? global_grid : nruns), dim3(kFlockThreads), 0, s,
                       wl.reps, wl.nreps, wl.tile_prefix, wl.app_prefix, R, offset, W, count_next,
                       gstep, err, gflag, wl.ntiles, nagents, ob, tile_sum);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bin sort: every bin of a freshly scattered read buffer into (cell y, id)
// order, ties by position; one block per (replica, tile), the tile's bins
// taken 256 at a time, one per thread, so every bin's size is read in one
// round trip.  A bin of at most kSortSmall entries is sorted by its own lane
// (sort_small_bin); a bin of up to 64 entries by a quarter wave, four bins
// side by side (sort_seg_bins); a bin of up to four times a wave's width by
// one wave -- all in registers, rewritten in place (every lane has loaded its
// entries before any lane stores); larger bins are ranked by the whole
// block against keys staged in LDS, written to the scratch buffer T at the
// same indices and copied back.

constexpr int kSortThreads = 256;
constexpr int kSortSlots = 4;                              // entries per lane on the wave path
constexpr uint32_t kSortWaveMax = kSortSlots * kWave;      // largest bin one wave sorts
constexpr uint32_t kSortChunk = 2048;  // keys staged per pass for the large bins

__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t lane) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)lane);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src, kWave);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

constexpr uint32_t kSortSeg = 16;                           // lanes per bin on the segment path
constexpr uint32_t kSortSegMax = kSortSeg * kSortSlots;     // largest bin a segment sorts

// Four bins at once, one per 16-lane segment of the wave: the segment of
// `lane` ranks bin [s, s + m) (m <= kSortSegMax; m = 0 for an idle
// segment), entry sl + 16 u in slot u of lane sl of the segment.
__device__ __forceinline__ void sort_seg_bins(const Geom& g, Buf R, uint32_t s, uint32_t m,
                                              uint32_t lane) {
  const uint32_t sl = lane & (kSortSeg - 1);
  const uint32_t seg0 = lane & ~(kSortSeg - 1);
  uint32_t idv[kSortSlots], rk[kSortSlots];
  float xv[kSortSlots], yv[kSortSlots], hxv[kSortSlots], hyv[kSortSlots];
  uint64_t kv[kSortSlots];
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = sl + (uint32_t)u * kSortSeg;
    idv[u] = 0;
    yv[u] = 0.0f;
    if (e < m) {
      idv[u] = R.id[s + e];
      yv[u] = R.y[s + e];
    }
    kv[u] = e < m ? bin_key(g, yv[u], idv[u]) : ~0ull;
    rk[u] = 0;
  }
  // out of order: some entry's predecessor (lane sl - 1 of the same slot,
  // or lane 15 of the slot before for sl = 0) has a larger key
  bool inv = false;
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = sl + (uint32_t)u * kSortSeg;
    const uint64_t left = shfl64(kv[u], sl > 0 ? lane - 1 : lane);
    const uint64_t wrap = shfl64(kv[u > 0 ? u - 1 : 0], seg0 + kSortSeg - 1);
    const uint64_t prev = sl > 0 ? left : wrap;
    if (e > 0 && e < m && prev > kv[u]) inv = true;
  }
  const uint64_t segmask = 0xFFFFull << seg0;
  const bool todo = (__ballot(inv) & segmask) != 0;
  if (!__any(todo)) return;
  uint32_t mmax = 0;
#pragma unroll
  for (uint32_t q = 0; q < kWave; q += kSortSeg)
    mmax = max(mmax, (uint32_t)__builtin_amdgcn_readlane((int)(todo ? m : 0u), (int)q));
  for (uint32_t j = 0; j < mmax; ++j) {
    const uint32_t ju = j / kSortSeg;  // the same on every lane
    const uint64_t src = ju == 0 ? kv[0] : (ju == 1 ? kv[1] : (ju == 2 ? kv[2] : kv[3]));
    const uint64_t kj = shfl64(src, seg0 + (j & (kSortSeg - 1)));
#pragma unroll
    for (int u = 0; u < kSortSlots; ++u) {
      const uint32_t e = sl + (uint32_t)u * kSortSeg;
      rk[u] += (kj < kv[u]) || (kj == kv[u] && j < e);
    }
  }
  if (!todo) return;
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = sl + (uint32_t)u * kSortSeg;
    xv[u] = hxv[u] = hyv[u] = 0.0f;
    if (e < m) {
      xv[u] = R.x[s + e];
      hxv[u] = R.hx[s + e];
      hyv[u] = R.hy[s + e];
    }
  }
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = sl + (uint32_t)u * kSortSeg;
    if (e < m) {
      R.id[s + rk[u]] = idv[u];
      R.x[s + rk[u]] = xv[u];
      R.y[s + rk[u]] = yv[u];
      R.hx[s + rk[u]] = hxv[u];
      R.hy[s + rk[u]] = hyv[u];
    }
  }
}

// One wave ranks bin [s, s + m) (kSortSegMax < m <= kSortWaveMax): entry
// lane + 64 u in slot u of the lane's registers.
__device__ __forceinline__ void sort_wave_bin(const Geom& g, Buf R, uint32_t s, uint32_t m,
                                              uint32_t lane) {
  uint32_t idv[kSortSlots], rk[kSortSlots];
  float xv[kSortSlots], yv[kSortSlots], hxv[kSortSlots], hyv[kSortSlots];
  uint64_t kv[kSortSlots];
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = lane + (uint32_t)u * kWave;
    idv[u] = 0;
    yv[u] = 0.0f;
    if (e < m) {
      idv[u] = R.id[s + e];
      yv[u] = R.y[s + e];
    }
    kv[u] = e < m ? bin_key(g, yv[u], idv[u]) : ~0ull;
    rk[u] = 0;
  }
  // already in order (the scatter keeps most bins so): nothing to write.
  // Entry e is out of order if its predecessor e - 1 has a larger key; the
  // predecessor of slot u's entry is lane - 1 of slot u, or lane 63 of slot
  // u - 1 for lane 0.
  bool inv = false;
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = lane + (uint32_t)u * kWave;
    const uint64_t up = __shfl_up(kv[u], 1, kWave);
    const uint64_t wrap = u > 0 ? readlane64(kv[u > 0 ? u - 1 : 0], kWave - 1) : 0ull;
    const uint64_t prev = lane == 0 ? wrap : up;
    if (e > 0 && e < m && prev > kv[u]) inv = true;
  }
  if (!__any(inv)) return;
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = lane + (uint32_t)u * kWave;
    xv[u] = hxv[u] = hyv[u] = 0.0f;
    if (e < m) {
      xv[u] = R.x[s + e];
      hxv[u] = R.hx[s + e];
      hyv[u] = R.hy[s + e];
    }
  }
  for (uint32_t j = 0; j < m; ++j) {
    const uint32_t ju = j / kWave;  // the same on every lane
    const uint64_t src = ju == 0 ? kv[0] : (ju == 1 ? kv[1] : (ju == 2 ? kv[2] : kv[3]));
    const uint64_t kj = readlane64(src, j & (kWave - 1));
#pragma unroll
    for (int u = 0; u < kSortSlots; ++u) {
      const uint32_t e = lane + (uint32_t)u * kWave;
      rk[u] += (kj < kv[u]) || (kj == kv[u] && j < e);
    }
  }
#pragma unroll
  for (int u = 0; u < kSortSlots; ++u) {
    const uint32_t e = lane + (uint32_t)u * kWave;
    if (e < m) {
      R.id[s + rk[u]] = idv[u];
      R.x[s + rk[u]] = xv[u];
      R.y[s + rk[u]] = yv[u];
      R.hx[s + rk[u]] = hxv[u];
      R.hy[s + rk[u]] = hyv[u];
    }
  }
}

__device__ __forceinline__ void sort_bins_tiles(const RepDev* __restrict__ reps, int nreps,
                                                const uint32_t* __restrict__ tile_prefix, Buf R,
                                                const uint32_t* __restrict__ offset, Buf T,
                                                uint32_t nitems) {
  __shared__ uint64_t s_key[kSortChunk];
  __shared__ uint32_t s_seg[kSortThreads];
  __shared__ uint32_t s_mid[kSortThreads];
  __shared__ uint32_t s_big[kSortThreads];
  __shared__ uint32_t s_nseg, s_nmid, s_nbig;
  const int t = threadIdx.x;
  const uint32_t lane = (uint32_t)t & (kWave - 1);
  const uint32_t wid = (uint32_t)t / kWave;
  for (uint32_t bi = blockIdx.x; bi < nitems; bi += gridDim.x) {
    const uint32_t item = xcd_remap(bi, nitems);
    const int r = find_rep(tile_prefix, nreps, item);
    const Geom g = reps[r].g;
    const uint32_t nb = 1u << (g.tx_log2 + g.tr_log2);
    const uint32_t b0 = reps[r].bin_base + ((item - tile_prefix[r]) << (g.tx_log2 + g.tr_log2));

    for (uint32_t k0 = 0; k0 < nb; k0 += kSortThreads) {
      if (t == 0) {
        s_nseg = 0u;
        s_nmid = 0u;
        s_nbig = 0u;
      }
      __syncthreads();
      if (k0 + (uint32_t)t < nb) {
        const uint32_t b = b0 + k0 + (uint32_t)t;
        const uint32_t s = offset[b], m = offset[b + 1] - s;
        if (m >= 2u && m <= kSortSmall) sort_small_bin(g, R, s, m);
        else if (m > kSortSmall && m <= kSortSegMax) s_seg[atomicAdd(&s_nseg, 1u)] = b;
        else if (m > kSortSegMax && m <= kSortWaveMax) s_mid[atomicAdd(&s_nmid, 1u)] = b;
        else if (m > kSortWaveMax) s_big[atomicAdd(&s_nbig, 1u)] = b;
      }
      __syncthreads();
      // bins of up to 64: a 16-lane segment each, four per wave
      const uint32_t nseg = s_nseg;
      constexpr uint32_t kSegsPerWave = kWave / kSortSeg;
      for (uint32_t i0 = wid * kSegsPerWave; i0 < nseg; i0 += kSortThreads / kSortSeg) {
        const uint32_t i = i0 + lane / kSortSeg;
        uint32_t s = 0, m = 0;
        if (i < nseg) {
          s = offset[s_seg[i]];
          m = offset[s_seg[i] + 1] - s;
        }
        sort_seg_bins(g, R, s, m, lane);
      }
      // middle bins: one wave each
      const uint32_t nmid = s_nmid;
      for (uint32_t i = wid; i < nmid; i += kSortThreads / kWave) {
        const uint32_t s = offset[s_mid[i]];
        sort_wave_bin(g, R, s, offset[s_mid[i] + 1] - s, lane);
      }
      // larger bins: the whole block
      const uint32_t nbig = s_nbig;
      for (uint32_t i = 0; i < nbig; ++i) {
        const uint32_t s = offset[s_big[i]];
        const uint32_t m = offset[s_big[i] + 1] - s;
        const bool staged_once = m <= kSortChunk;
        if (staged_once) {
          for (uint32_t j = t; j < m; j += kSortThreads) s_key[j] = bin_key(g, R.y[s + j], R.id[s + j]);
          __syncthreads();
        }
        for (uint32_t e0 = 0; e0 < m; e0 += kSortThreads) {
          const uint32_t e = e0 + (uint32_t)t;
          const bool act = e < m;
          const uint64_t key = act ? bin_key(g, R.y[s + e], R.id[s + e]) : 0ull;
          uint32_t rank = 0;
          for (uint32_t c0 = 0; c0 < m; c0 += kSortChunk) {
            const uint32_t n = m - c0 < kSortChunk ? m - c0 : kSortChunk;
            if (!staged_once) {
              __syncthreads();
              for (uint32_t j = t; j < n; j += kSortThreads)
                s_key[j] = bin_key(g, R.y[s + c0 + j], R.id[s + c0 + j]);
              __syncthreads();
            }
            if (act)
              for (uint32_t j = 0; j < n; ++j) {
                const uint64_t kj = s_key[j];
                rank += (kj < key) || (kj == key && c0 + j < e);
              }
          }
          if (act) {
            T.id[s + rank] = R.id[s + e];
            T.x[s + rank] = R.x[s + e];
            T.y[s + rank] = R.y[s + e];
            T.hx[s + rank] = R.hx[s + e];
            T.hy[s + rank] = R.hy[s + e];
          }
        }
        __syncthreads();
        for (uint32_t j = t; j < m; j += kSortThreads) {
          R.id[s + j] = T.id[s + j];
          R.x[s + j] = T.x[s + j];
          R.y[s + j] = T.y[s + j];
          R.hx[s + j] = T.hx[s + j];
          R.hy[s + j] = T.hy[s + j];
        }
        __syncthreads();
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kSortThreads) void k_sort_bins(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, Buf R,
    const uint32_t* __restrict__ offset, Buf T, uint32_t nitems) {
  sort_bins_tiles(reps, nreps, tile_prefix, R, offset, T, nitems);
}

// Waits until all `nblocks` work-groups of the grid have arrived at `bar`
// (zeroed before the launch), with every write the grid made before it
// visible to every work-group after it.  The fallback's 256 work-groups are
// all resident at once on this GPU; the wait still ends after kGridWaitTicks
// of the 100 MHz wall clock, returning false, so no launch can hang on it.
constexpr uint64_t kGridWaitTicks = 10000000;  // 100 ms

__device__ bool grid_wait(uint32_t* bar, uint32_t nblocks) {
  __shared__ uint32_t s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t0 = wall_clock64();
    uint32_t ok = 1u;
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < nblocks) {
      if (wall_clock64() - t0 > kGridWaitTicks) {
        ok = 0u;
        break;
      }
      __builtin_amdgcn_s_sleep(4);
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0u;
}

// The rebinning a gathered step could not do, in one launch: the scatter of
// every tile, a grid-wide wait, then the bin sort of every tile (T: the
// sort's scratch).  Returns at once unless the gate is set.  count_down:
// cursor holds this step's bin counts, counted down from offset_new (the
// summed steps' gather writes no cursors).  A wait that times out sets bit
// 2 of err and leaves the bins unsorted.
__global__ __launch_bounds__(256) void k_rebin_fallback(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, Buf R,
    Buf W, const uint32_t* __restrict__ offset_cur, uint32_t* __restrict__ cursor, Buf R2,
    const uint32_t* __restrict__ offset_new, Buf T, Gate gate, uint32_t* __restrict__ err,
    int count_down, uint32_t nitems) {
  if (*gate.flag == 0u) return;  // the gather rebinned this step
  if (blockIdx.x == 0 && threadIdx.x == 0) *gate.redone += 1u;  // steps redone
  scatter_aligned_tiles(reps, nreps, tile_prefix, R, W, offset_cur, cursor, R2,
                        count_down ? offset_new : nullptr, nitems);
  if (!grid_wait(gate.bar, gridDim.x)) {
    if (threadIdx.x == 0) atomicOr(err, 2u);
    return;
  }
  sort_bins_tiles(reps, nreps, tile_prefix, R2, offset_new, T, nitems);
}

hipError_t launch_rebin_fallback(const WorkList& wl, Buf R, Buf W, const uint32_t* offset_cur,
                                 uint32_t* cursor, Buf R2, const uint32_t* offset_new, Buf T,
                                 const Gate& gate, uint32_t* err, bool count_down,
                                 hipStream_t s) {
  if (wl.ntiles == 0) return hipSuccess;
  if (!gate.flag || !gate.redone || !gate.bar || !err) return hipErrorInvalidValue;
  const uint32_t grid = wl.ntiles > kGatedGrid ? kGatedGrid : wl.ntiles;
  hipLaunchKernelGGL(k_rebin_fallback, dim3(grid), dim3(256), 0, s, wl.reps, wl.nreps,
                     wl.tile_prefix, R, W, offset_cur, cursor, R2, offset_new, T, gate, err,
                     count_down ? 1 : 0, wl.ntiles);
  return hipGetLastError();
}

hipError_t launch_sort_bins(const WorkList& wl, Buf R, const uint32_t* offset, Buf T,
                            hipStream_t s) {
  if (wl.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_bins, dim3(wl.ntiles), dim3(kSortThreads), 0, s, wl.reps, wl.nreps,
                     wl.tile_prefix, R, offset, T, wl.ntiles);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// single-replica API queries

__global__ void k_query(const RepDev* __restrict__ rep, Buf R, const uint32_t* __restrict__ offset,
                        float x, float y, float dist, float T, int relax, uint32_t* out,
                        uint32_t cap, uint32_t* nout) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Geom g = rep->g;
  uint32_t n = 0;
  query_global(g, R, offset + rep->bin_base, x, y, dist, T, relax != 0, [&](uint32_t e) {
    if (n < cap) out[n] = R.id[e];
    ++n;
  });
  *nout = n;
}

hipError_t launch_query(const RepDev* rep, Buf R, const uint32_t* offset, float x, float y,
                        float dist, float T, int relax, uint32_t* out, uint32_t cap,
                        uint32_t* nout, hipStream_t s) {
  hipLaunchKernelGGL(k_query, dim3(1), dim3(64), 0, s, rep, R, offset, x, y, dist, T, relax, out,
                     cap, nout);
  return hipGetLastError();
}

__global__ void k_neighbor_count(const RepDev* __restrict__ rep, Buf R,
                                 const uint32_t* __restrict__ offset, uint32_t n, float dist,
                                 float T, uint32_t* counts) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geom g = rep->g;
  uint32_t c = 0;
  query_global(g, R, offset + rep->bin_base, R.x[i], R.y[i], dist, T, false,
               [&](uint32_t) { ++c; });
  counts[i] = c;
}

hipError_t launch_neighbor_count(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                                 float dist, float T, uint32_t* counts, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_neighbor_count, dim3((n + 127) / 128), dim3(128), 0, s, rep, R, offset, n,
                     dist, T, counts);
  return hipGetLastError();
}

__global__ void k_neighbor_fill(const RepDev* __restrict__ rep, Buf R,
                                const uint32_t* __restrict__ offset, uint32_t n, float dist,
                                float T, const uint32_t* __restrict__ row_ptr, uint32_t* ids) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geom g = rep->g;
  uint32_t k = row_ptr[i];
  query_global(g, R, offset + rep->bin_base, R.x[i], R.y[i], dist, T, false,
               [&](uint32_t e) { ids[k++] = R.id[e]; });
}

hipError_t launch_neighbor_fill(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                                float dist, float T, const uint32_t* row_ptr, uint32_t* ids,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_neighbor_fill, dim3((n + 127) / 128), dim3(128), 0, s, rep, R, offset, n,
                     dist, T, row_ptr, ids);
  return hipGetLastError();
}

// get_objects on the read buffer: the bag of (x, y) in ascending id.
__global__ void k_objects_read(const RepDev* __restrict__ rep, Buf R,
                               const uint32_t* __restrict__ offset, float x, float y,
                               uint32_t* out, uint32_t cap, uint32_t* nout) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Geom g = rep->g;
  int32_t bx, by;
  discretize(g, x, y, &bx, &by);
  uint32_t n = 0;
  if (cell_valid(g, bx, by)) {
    const uint32_t b = rep->bin_base + bin_of(g, bx, by);
    const uint32_t s = offset[b], e = offset[b + 1];
    int64_t last = -1;
    for (;;) {
      uint32_t best = 0xFFFFFFFFu;
      bool any = false;
      for (uint32_t q = s; q < e; ++q) {
        const uint32_t qid = R.id[q];
        if ((int64_t)qid <= last || (any && qid >= best)) continue;
        int32_t qx, qy;
        discretize(g, R.x[q], R.y[q], &qx, &qy);
        if (qx != bx || qy != by) continue;
        best = qid;
        any = true;
      }
      if (!any) break;
      for (uint32_t q = s; q < e; ++q) {
        if (R.id[q] != best) continue;
        int32_t qx, qy;
        discretize(g, R.x[q], R.y[q], &qx, &qy);
        if (qx != bx || qy != by) continue;
        if (n < cap) out[n] = best;
        ++n;
      }
      last = best;
    }
  }
  *nout = n;
}

hipError_t launch_objects_read(const RepDev* rep, Buf R, const uint32_t* offset, float x, float y,
                               uint32_t* out, uint32_t cap, uint32_t* nout, hipStream_t s) {
  hipLaunchKernelGGL(k_objects_read, dim3(1), dim3(64), 0, s, rep, R, offset, x, y, out, cap, nout);
  return hipGetLastError();
}

// get_objects_unbuffered: the agents of buffer B (n entries) in the bag of
// (x, y), in B's order (insertion order) or, with by_id, in ascending id.
__global__ void k_objects_write(const RepDev* __restrict__ rep, Buf B, uint32_t n, int by_id,
                                float x, float y, uint32_t* out, uint32_t cap, uint32_t* nout) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Geom g = rep->g;
  int32_t bx, by;
  discretize(g, x, y, &bx, &by);
  uint32_t m = 0;
  if (!by_id) {
    for (uint32_t q = 0; q < n; ++q) {
      int32_t qx, qy;
      discretize(g, B.x[q], B.y[q], &qx, &qy);
      if (qx != bx || qy != by) continue;
      if (m < cap) out[m] = B.id[q];
      ++m;
    }
  } else {
    int64_t last = -1;
    for (;;) {
      uint32_t best = 0xFFFFFFFFu;
      bool any = false;
      for (uint32_t q = 0; q < n; ++q) {
        const uint32_t qid = B.id[q];
        if ((int64_t)qid <= last || (any && qid >= best)) continue;
        int32_t qx, qy;
        discretize(g, B.x[q], B.y[q], &qx, &qy);
        if (qx != bx || qy != by) continue;
        best = qid;
        any = true;
      }
      if (!any) break;
      for (uint32_t q = 0; q < n; ++q) {
        if (B.id[q] != best) continue;
        int32_t qx, qy;
        discretize(g, B.x[q], B.y[q], &qx, &qy);
        if (qx != bx || qy != by) continue;
        if (m < cap) out[m] = best;
        ++m;
      }
      last = best;
    }
  }
  *nout = m;
}

hipError_t launch_objects_write(const RepDev* rep, Buf B, uint32_t n, int by_id, float x, float y,
                                uint32_t* out, uint32_t cap, uint32_t* nout, hipStream_t s) {
  hipLaunchKernelGGL(k_objects_write, dim3(1), dim3(64), 0, s, rep, B, n, by_id, x, y, out, cap,
                     nout);
  return hipGetLastError();
}

// Number of distinct occupied cells of the read buffer: an agent counts if
// no other agent of its cell precedes it in (id, index) order.
__global__ void k_count_cells(const RepDev* __restrict__ rep, Buf R,
                              const uint32_t* __restrict__ offset, uint32_t n,
                              uint32_t* distinct) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geom g = rep->g;
  int32_t cx, cy;
  discretize(g, R.x[i], R.y[i], &cx, &cy);
  const uint32_t b = rep->bin_base + bin_of(g, cx, cy);
  const uint32_t s = offset[b], e = offset[b + 1];
  const uint32_t id = R.id[i];
  for (uint32_t q = s; q < e; ++q) {
    if (q == i) continue;
    const uint32_t qid = R.id[q];
    if (qid > id || (qid == id && q > i)) continue;
    int32_t qx, qy;
    discretize(g, R.x[q], R.y[q], &qx, &qy);
    if (qx == cx && qy == cy) return;
  }
  atomicAdd(distinct, 1u);
}

hipError_t launch_count_cells(const RepDev* rep, Buf R, const uint32_t* offset, uint32_t n,
                              uint32_t* distinct, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_cells, dim3((n + 255) / 256), dim3(256), 0, s, rep, R, offset, n,
                     distinct);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// empty bags (field_2d.rs:727-800): an occupancy bitmap over the bag grid in
// the reference's bag index order (cx * dh + cy), 32 bags per word; the
// empty bags of each word counted, scanned, then emitted in index order.

__global__ void k_mark_bags(const RepDev* __restrict__ rep, Buf R, uint32_t n,
                            uint32_t* __restrict__ occ) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Geom g = rep->g;
  int32_t cx, cy;
  discretize(g, R.x[i], R.y[i], &cx, &cy);
  if (!cell_valid(g, cx, cy)) return;
  const uint64_t idx = (uint64_t)cx * (uint64_t)g.dh + (uint64_t)cy;
  atomicOr(&occ[idx >> 5], 1u << (idx & 31u));
}

__device__ __forceinline__ uint32_t empty_mask(const uint32_t* occ, uint32_t w, uint64_t ncells) {
  const uint64_t lo = (uint64_t)w << 5;
  const uint32_t valid = ncells - lo >= 32u ? 0xFFFFFFFFu : ((1u << (uint32_t)(ncells - lo)) - 1u);
  return ~occ[w] & valid;
}

__global__ void k_count_empty(const uint32_t* __restrict__ occ, uint32_t nwords, uint64_t ncells,
                              uint32_t* __restrict__ cnt) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < nwords) cnt[w] = (uint32_t)__popc(empty_mask(occ, w, ncells));
}

// not_discretize (field_2d.rs:345-353): (cx as f32 * d, cy as f32 * d)
__global__ void k_emit_empty(const RepDev* __restrict__ rep, const uint32_t* __restrict__ occ,
                             uint32_t nwords, uint64_t ncells, const uint32_t* __restrict__ off,
                             float* __restrict__ x, float* __restrict__ y) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwords) return;
  const Geom g = rep->g;
  uint32_t m = empty_mask(occ, w, ncells);
  uint32_t pos = off[w];
  while (m) {
    const uint32_t b = (uint32_t)__ffs(m) - 1u;
    const uint64_t idx = ((uint64_t)w << 5) + b;
    x[pos] = (float)(int32_t)(idx / (uint64_t)g.dh) * g.d;
    y[pos] = (float)(int32_t)(idx % (uint64_t)g.dh) * g.d;
    ++pos;
    m &= m - 1u;
  }
}

hipError_t launch_mark_bags(const RepDev* rep, Buf R, uint32_t n, uint32_t* occ, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mark_bags, dim3((n + 255) / 256), dim3(256), 0, s, rep, R, n, occ);
  return hipGetLastError();
}

hipError_t launch_count_empty(const uint32_t* occ, uint32_t nwords, uint64_t ncells,
                              uint32_t* cnt, hipStream_t s) {
  if (nwords == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_empty, dim3((nwords + 255) / 256), dim3(256), 0, s, occ, nwords,
                     ncells, cnt);
  return hipGetLastError();
}

hipError_t launch_emit_empty(const RepDev* rep, const uint32_t* occ, uint32_t nwords,
                             uint64_t ncells, const uint32_t* off, float* x, float* y,
                             hipStream_t s) {
  if (nwords == 0) return hipSuccess;
  hipLaunchKernelGGL(k_emit_empty, dim3((nwords + 255) / 256), dim3(256), 0, s, rep, occ, nwords,
                     ncells, off, x, y);
  return hipGetLastError();
}

// remove_object_location on an insertion-ordered buffer: stable compaction
// dropping every entry with this id in the bag of (x, y).
__global__ void k_remove(const RepDev* __restrict__ rep, Buf A, uint32_t n, uint32_t id, float x,
                         float y, uint32_t* n_after) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Geom g = rep->g;
  int32_t bx, by;
  discretize(g, x, y, &bx, &by);
  uint32_t k = 0;
  for (uint32_t q = 0; q < n; ++q) {
    bool drop = false;
    if (A.id[q] == id) {
      int32_t qx, qy;
      discretize(g, A.x[q], A.y[q], &qx, &qy);
      drop = (qx == bx && qy == by);
    }
    if (drop) continue;
    if (k != q) {
      A.x[k] = A.x[q];
      A.y[k] = A.y[q];
      A.hx[k] = A.hx[q];
      A.hy[k] = A.hy[q];
      A.id[k] = A.id[q];
    }
    ++k;
  }
  *n_after = k;
}

hipError_t launch_remove(const RepDev* rep, Buf A, uint32_t n, uint32_t id, float x, float y,
                         uint32_t* n_after, hipStream_t s) {
  hipLaunchKernelGGL(k_remove, dim3(1), dim3(64), 0, s, rep, A, n, id, x, y, n_after);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// tile-resident step (phase bodies in kg_rt.hpp)

size_t rt_lds_bytes(const RtCaps& c) { return rt_smem_bytes(c); }

// Output phase of a tile (after its stayers are classified and counted per
// bin in LDS): sort the stayers by (bin, cell y, id), write them and the new
// bin offsets.
template <int kT>
__device__ __forceinline__ void rt_output(const RepDev& rp, const RtTile& rt, const RtDev& D,
                                          const RtMeta& me, uint32_t bo_own, int nb_tile,
                                          uint32_t nstay, RtSmem S, uint32_t* s_warp) {
  const int t = threadIdx.x;
  const int pout = rt.pin ^ 1;
  // exclusive scan of the nb_tile bin counts, 4 per thread
  uint32_t v[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int b = 4 * t + k;
    v[k] = b < nb_tile ? S.hist[b] : 0u;
    sum += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan<kT>(sum, s_warp, &tot);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int b = 4 * t + k;
    if (b < nb_tile) {
      S.bbase[b] = run;
      S.bcur[b] = run;
    }
    run += v[k];
  }
  if (t == 0) S.bbase[nb_tile] = nstay;
  __syncthreads();
  for (uint32_t s = t; s < nstay; s += kT) {
    const uint32_t pos = atomicAdd(&S.bcur[S.obin[s]], 1u);
    S.operm[pos] = (uint16_t)s;
  }
  __syncthreads();
  const uint32_t scap = me.scap;
  for (uint32_t s = t; s < nstay; s += kT) rt_write_stayer(rp, rt, D, me, s, S);
  for (int b = t; b <= nb_tile; b += kT)
    D.boff[pout][bo_own + b] = S.bbase[b] < scap ? S.bbase[b] : scap;
  if (t == 0) D.cnt[pout][rt.gt] = nstay < scap ? nstay : scap;
}

// One schedule step of the tiles of class `cls`: cls 0 covers every tile
// (other classes' tiles return at once), further classes the tiles in
// `list`.  `mode` is kRtModeClass (tiles that outgrow the capacities defer),
// kRtModeFinal (they take the global path) or kRtModeDeferred (block i runs
// the i-th tile deferred in this step, with `caps` of the last class).
template <int kT>
__global__ __launch_bounds__(kT) void k_rt_step(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, RtDev D,
    RtCaps caps, uint32_t koff, int cls, const uint32_t* __restrict__ list, int mode) {
  extern __shared__ __align__(16) uint8_t smem_raw[];
  __shared__ uint32_t s_seg_start[3 * kMaxStageCols];
  __shared__ uint32_t s_seg_len[3 * kMaxStageCols];
  __shared__ uint32_t s_col_base[kMaxStageCols + 1];
  __shared__ uint32_t s_src_base[10];
  __shared__ uint32_t s_src_pre[11];
  __shared__ uint32_t s_warp[kT / kWave];
  __shared__ uint32_t s_ne, s_nown, s_nstay, s_fail, s_defer;

  const int t = threadIdx.x;
  // schedule step: launch koff of a batch that starts at *gstep
  const uint64_t g = *D.gstep + koff;
  const int par = (int)(g & 1u);
  uint32_t item;
  const bool deferred = mode == kRtModeDeferred;
  if (deferred) {
    // block 0 clears the list that the next step's class launches fill
    const uint32_t n = D.ovf[par][0];
    if (blockIdx.x == 0 && t == 0) {
      D.ovf[par ^ 1][0] = 0u;
      atomicMax(&D.flags[1], n);
    }
    if (blockIdx.x >= n || blockIdx.x >= kRtOvfMax) return;
    item = D.ovf[par][1 + blockIdx.x];
  } else {
    item = cls ? list[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
  }
  const RtTile rt = rt_tile(reps, nreps, tile_prefix, item, g);
  const RtMeta me = D.meta[rt.gt];
  if (!deferred && (int)me.cls != cls) return;
  const RepDev& rp = reps[rt.r];
  if (rp.dead) {
    // a retired run: no work, but the engine-wide spill counter of step
    // g + 2 is still tile 0's to clear
    if (t == 0 && rt.gt == 0) *D.scnt[rt.ph_reset] = 0u;
    return;
  }
  const RtSmem S = rt_carve(smem_raw, caps);
  const int nb_tile = rp.g.tx * rp.g.tr;
  const int pout = rt.pin ^ 1;
  const uint32_t bo_own = rt_boff_at(rp, rt, rt.gt);
  const RtSegs sg = rt_segs(rp.g, rt);
  if (t == 0) {
    s_ne = 0;
    s_nown = 0;
    s_nstay = 0;
    s_fail = 0;
    s_defer = 0;
  }

  const bool lds_ok = rp.lds_ok != 0 && rt_fits_tables(rt, sg, caps) && nb_tile <= caps.nb;
  if (lds_ok)
    for (int b = t; b <= nb_tile; b += kT) S.hist[b] = 0u;
  uint32_t len = 0;
  // waves 0-2: the staged columns' runs; wave 3: the extras' sources
  if (lds_ok && t < rt.ncols) len = rt_segments(rp, rt, D, t, s_seg_start, s_seg_len);
  const int ks = t - (kT - kWave);
  if (ks >= 0 && ks < 10) {
    uint32_t b;
    const uint32_t n = rt_extra_count(rp, rt, D, ks, &b);
    s_src_base[ks] = b;
    s_src_pre[ks + 1] = n;
  }
  uint32_t total;
  const uint32_t ex = block_exclusive_scan<kT>(len, s_warp, &total);
  if (lds_ok && t < rt.ncols) s_col_base[t] = ex;
  if (t == 0) {
    s_col_base[rt.ncols] = total;
    s_src_pre[0] = 0;
    for (int k = 0; k < 10; ++k) s_src_pre[k + 1] += s_src_pre[k];
  }
  __syncthreads();
  bool global_path = !lds_ok || total > (uint32_t)caps.runs;
  uint32_t nown = 0, ne = 0;
  if (!global_path) {
    const uint32_t next = s_src_pre[10];
    if (!rt_fetch(rp, rt, D, t, t + kT, total, t, next, s_seg_start, s_seg_len, s_col_base,
                  s_src_base, s_src_pre, caps, &s_ne, &s_nown, S))
      s_fail = 1u;
    for (uint32_t k = t + 2 * kT; k < total; k += kT)
      rt_stage(rp, rt, D, k, s_seg_start, s_seg_len, s_col_base, caps.own, &s_nown, S);
    for (uint32_t j = t + kT; j < next; j += kT)
      if (!rt_gather_extra(rp, rt, D, j, s_src_base, s_src_pre, caps.extras, &s_ne, caps.own,
                           &s_nown, S))
        s_fail = 1u;
    const uint32_t nbt = (uint32_t)(rt.ncols * sg.nbc);
    for (uint32_t q = t; q < nbt; q += kT) rt_bintab(rp, rt, D, sg, q, s_seg_len, S);
    __syncthreads();
    nown = s_nown;
    ne = s_ne;
    global_path = s_fail != 0 || nown > (uint32_t)caps.own;
  }
  if (global_path && mode == kRtModeClass && lds_ok) {
    // outgrew its class's capacities: hand the tile to the step's launch
    // for deferred tiles (nothing of it has been written yet)
    if (t == 0) {
      const uint32_t slot = atomicAdd(&D.ovf[par][0], 1u);
      if (slot < kRtOvfMax) {
        D.ovf[par][1 + slot] = rt.gt;
        s_defer = 1u;
      }
      atom_or(D.flags, kRtFlagDefer);
    }
    __syncthreads();
    if (s_defer) return;
  }
  if (!global_path) {
    // extras sorted by (column, row, id) and their column starts
    for (uint32_t j = t; j < ne; j += kT) rt_rank_extra(j, ne, S);
    for (int c = t; c <= rt.ncols; c += kT) rt_ecbase(c, ne, S);
    __syncthreads();
    // queries and Bird::step, then classification of the thread's own
    // results (movers are pushed at once); no barrier between the two
    for (uint32_t o = t; o < nown; o += kT) rt_agent_lds(rp, rt, sg, o, s_col_base, S);
    for (uint32_t o = t; o < nown; o += kT) rt_classify(rp, rt, D, o, &s_nstay, S);
    __syncthreads();
    rt_output<kT>(rp, rt, D, me, bo_own, nb_tile, s_nstay, S, s_warp);
  } else {
    if (t == 0) atom_or(D.flags, kRtFlagGlobal);
    rt_tile_global(rp, rt, D, (uint32_t)t, kT);
    for (int b = t; b <= nb_tile; b += kT) D.boff[pout][bo_own + b] = 0u;
    if (t == 0) D.cnt[pout][rt.gt] = 0u;
  }

  if (t == 0) {
    // counters of step g + 2: nothing reads or writes them during step g
    D.icnt[rt.ph_reset][rt.gt] = 0u;
    if (rt.gt == 0) *D.scnt[rt.ph_reset] = 0u;
  }
}

// Entry into the tile-resident layout: each tile of class `cls` takes its
// agents from the bin-grouped read buffer Ra (schedule step load_g's output)
// and writes them as the stayers of parity (load_g + 1) & 1.
__global__ __launch_bounds__(kRtThreads) void k_rt_load(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, RtDev D,
    RtCaps caps, int cls, const uint32_t* __restrict__ list, Buf Ra,
    const uint32_t* __restrict__ offset_a, uint64_t load_g) {
  extern __shared__ __align__(16) uint8_t smem_raw[];
  __shared__ uint32_t s_warp[kRtThreads / kWave];
  __shared__ uint32_t s_nstay;
  const int t = threadIdx.x;
  const uint32_t item = cls ? list[blockIdx.x] : xcd_remap(blockIdx.x, gridDim.x);
  const RtTile rt = rt_tile(reps, nreps, tile_prefix, item, load_g);
  const RtMeta me = D.meta[rt.gt];
  if ((int)me.cls != cls) return;
  const RepDev& rp = reps[rt.r];
  const RtSmem S = rt_carve(smem_raw, caps);
  const int nb_tile = rp.g.tx * rp.g.tr;
  const int pout = rt.pin ^ 1;
  const uint32_t bo_own = rt_boff_at(rp, rt, rt.gt);
  const uint32_t tile = rt.gt - rt.tbase;
  const uint32_t b0 = rp.bin_base + (tile << (rp.g.tx_log2 + rp.g.tr_log2));
  const uint32_t lo = offset_a[b0], hi = offset_a[b0 + (uint32_t)nb_tile];
  const uint32_t nown = hi - lo;
  if (t == 0) s_nstay = 0;
  if (nown > (uint32_t)caps.own || nb_tile > caps.nb) {
    for (uint32_t i = lo + t; i < hi; i += kRtThreads) {
      int32_t cx, cy;
      discretize(rp.g, Ra.x[i], Ra.y[i], &cx, &cy);
      rt_push(rp, rt, D, rt.tbase + tile_of(rp.g, cx, cy), Ra.x[i], Ra.y[i], Ra.hx[i], Ra.hy[i],
              Ra.id[i]);
    }
    for (int b = t; b <= nb_tile; b += kRtThreads) D.boff[pout][bo_own + b] = 0u;
    if (t == 0) D.cnt[pout][rt.gt] = 0u;
    return;
  }
  for (int b = t; b <= nb_tile; b += kRtThreads) S.hist[b] = 0u;
  __syncthreads();
  for (uint32_t o = t; o < nown; o += kRtThreads) {
    S.rx[o] = Ra.x[lo + o];
    S.ry[o] = Ra.y[lo + o];
    S.rhx[o] = Ra.hx[lo + o];
    S.rhy[o] = Ra.hy[lo + o];
    S.rid[o] = Ra.id[lo + o];
    rt_classify(rp, rt, D, o, &s_nstay, S);
  }
  __syncthreads();
  rt_output<kRtThreads>(rp, rt, D, me, bo_own, nb_tile, s_nstay, S, s_warp);
}

hipError_t launch_rt_step(const WorkList& wl, const RtDev& D, const RtCaps& caps, uint32_t koff,
                          int cls, const uint32_t* list, uint32_t nlist, int mode, hipStream_t s) {
  const uint32_t grid = mode == kRtModeDeferred ? kRtOvfMax : (cls ? nlist : wl.ntiles);
  if (grid == 0) return hipSuccess;
  const size_t lds = rt_lds_bytes(caps);
  if (caps.threads == 1024)
    hipLaunchKernelGGL(k_rt_step<1024>, dim3(grid), dim3(1024), lds, s, wl.reps, wl.nreps,
                       wl.tile_prefix, D, caps, koff, cls, list, mode);
  else if (caps.threads == 512)
    hipLaunchKernelGGL(k_rt_step<512>, dim3(grid), dim3(512), lds, s, wl.reps, wl.nreps,
                       wl.tile_prefix, D, caps, koff, cls, list, mode);
  else
    hipLaunchKernelGGL(k_rt_step<256>, dim3(grid), dim3(256), lds, s, wl.reps, wl.nreps,
                       wl.tile_prefix, D, caps, koff, cls, list, mode);
  return hipGetLastError();
}

hipError_t launch_rt_load(const WorkList& wl, const RtDev& D, const RtCaps& caps, int cls,
                          const uint32_t* list, uint32_t nlist, Buf Ra, const uint32_t* offset_a,
                          uint64_t load_g, hipStream_t s) {
  const uint32_t grid = cls ? nlist : wl.ntiles;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rt_load, dim3(grid), dim3(kRtThreads), rt_lds_bytes(caps), s, wl.reps,
                     wl.nreps, wl.tile_prefix, D, caps, cls, list, Ra, offset_a, load_g);
  return hipGetLastError();
}

// Writes the agents of tiles [tile0, tile0 + gridDim.x) of the tile-resident
// layout (stayers, inbox entries, and -- for a replica's first tile -- its
// spill entries, all of the current parity) into the append buffer A: replica
// r's agents land in [app_base, app_base + n) at rep_cursor[r] (one atomic per
// work-group, so every replica's range is contiguous).
__global__ __launch_bounds__(256) void k_rt_unload(
    const RepDev* __restrict__ reps, int nreps, const uint32_t* __restrict__ tile_prefix, RtDev D,
    Buf A, uint32_t* __restrict__ rep_cursor, uint32_t tile0) {
  __shared__ uint32_t s_nsp, s_base, s_slot;
  const uint32_t item = tile0 + blockIdx.x;
  const uint64_t g = *D.gstep;
  const RtTile rt = rt_tile(reps, nreps, tile_prefix, item, g);
  const RepDev& rp = reps[rt.r];
  if (rp.dead) return;
  const uint32_t t = threadIdx.x;
  const uint32_t ns = D.cnt[rt.pin][rt.gt];
  uint32_t ib;
  const uint32_t ni = rt_extra_count(rp, rt, D, 4, &ib);
  uint32_t sb;
  const bool first = rt.gt == rt.tbase;
  const uint32_t nsl = first ? rt_extra_count(rp, rt, D, 9, &sb) : 0u;
  if (t == 0) {
    s_nsp = 0;
    s_slot = 0;
  }
  __syncthreads();
  for (uint32_t q = t; q < nsl; q += blockDim.x)
    if (D.spill_rep[rt.pin][q] == (uint32_t)rt.r) atomicAdd(&s_nsp, 1u);
  __syncthreads();
  const uint32_t n = ns + ni + s_nsp;
  if (t == 0) s_base = rp.app_base + atomicAdd(&rep_cursor[rt.r], n);
  __syncthreads();
  const uint32_t base = s_base;
  auto put = [&](const Buf& B, uint32_t gi, uint32_t k) {
    A.x[base + k] = B.x[gi];
    A.y[base + k] = B.y[gi];
    A.hx[base + k] = B.hx[gi];
    A.hy[base + k] = B.hy[gi];
    A.id[base + k] = B.id[gi];
  };
  const uint32_t sbase = D.meta[rt.gt].sbase;
  for (uint32_t o = t; o < ns; o += blockDim.x) put(D.stay[rt.pin], sbase + o, o);
  for (uint32_t o = t; o < ni; o += blockDim.x) put(D.inbox[rt.pin], ib + o, ns + o);
  for (uint32_t q = t; q < nsl; q += blockDim.x)
    if (D.spill_rep[rt.pin][q] == (uint32_t)rt.r) put(D.spill[rt.pin], q, ns + ni + atomicAdd(&s_slot, 1u));
}

hipError_t launch_rt_unload(const WorkList& wl, const RtDev& D, Buf A, uint32_t* rep_cursor,
                            uint32_t tile0, uint32_t ntiles, hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_rt_unload, dim3(ntiles), dim3(256), 0, s, wl.reps, wl.nreps,
                     wl.tile_prefix, D, A, rep_cursor, tile0);
  return hipGetLastError();
}

// count[t] = agents of tile t in the bin-grouped layout (a replica's tiles
// are runs of tx * tr consecutive bins from its bin_base).
__global__ void k_tile_counts(const RepDev* __restrict__ reps, int nreps,
                              const uint32_t* __restrict__ tile_prefix, uint32_t nt,
                              const uint32_t* __restrict__ offset, uint32_t* __restrict__ count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int r = find_rep(tile_prefix, nreps, t);
  const Geom& g = reps[r].g;
  const uint32_t nbt = (uint32_t)(g.tx * g.tr);
  const uint32_t b0 = reps[r].bin_base + (t - tile_prefix[r]) * nbt;
  count[t] = offset[b0 + nbt] - offset[b0];
}

hipError_t launch_tile_counts(const WorkList& wl, const uint32_t* offset, uint32_t* count,
                              hipStream_t s) {
  if (wl.ntiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_counts, dim3((wl.ntiles + 255) / 256), dim3(256), 0, s, wl.reps,
                     wl.nreps, wl.tile_prefix, wl.ntiles, offset, count);
  return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

}  // namespace kg

namespace kg {

__global__ void k_set_u64(uint64_t* p, uint64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t launch_set_u64(uint64_t* p, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

__global__ void k_add_u64(uint64_t* p, uint64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

hipError_t launch_add_u64(uint64_t* p, uint64_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_add_u64, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

}  // namespace kg
