// The per-frame pixel regrouping of the reflection pass (A8): a stable device-wide sort of the
// pixels by their int16 group id, the run heads of the sorted ids, and the <= P-pixel blocks of each
// run.  Written without a library sort or scan and without memset operations: every counter and
// offset is written by the kernels below, so the whole sequence can be captured into a HIP graph and
// replayed, and none of it synchronises the host (except snisr_sort_by_key, whose reference form
// returns the group count to the host).
//
// Sort: least-significant-digit radix sort with two 8-bit digit passes over key ^ 0x8000 (that
// orders signed int16 ascending).  A pass is three kernels: per-chunk digit histograms, one
// workgroup turning them into every (digit, chunk) start, and a stable scatter in which each
// workgroup walks its chunk in 256-item tiles in order; an item's rank among the equal digits before
// it is found with eight wavefront ballots (one per digit bit).
#include "snisr_common.hpp"

namespace snisr {

constexpr int kRsThreads = 256;             // one item per thread per tile
constexpr int kRsMinTiles = 4;              // tiles per chunk (at least)
constexpr int64_t kRsMaxChunks = 2048;      // chunks (histogram rows) at most
constexpr int kRsDigits = 256;
constexpr int64_t kMaxGroups = 65536;       // distinct int16 keys

struct RsPlan {
  int64_t chunk, n_chunks;
};

// chunk = items per workgroup: 4 tiles of 256, or more so that there are at most 2048 chunks
inline RsPlan rs_plan(int64_t n) {
  int64_t tiles = (n + (int64_t)kRsThreads * kRsMaxChunks - 1) / ((int64_t)kRsThreads * kRsMaxChunks);
  if (tiles < kRsMinTiles) tiles = kRsMinTiles;
  const int64_t chunk = tiles * kRsThreads;
  return {chunk, n > 0 ? (n + chunk - 1) / chunk : 0};
}

__device__ __forceinline__ int rs_digit(int16_t k, int shift) { return (((uint32_t)(uint16_t)k ^ 0x8000u) >> shift) & 0xFF; }

// Lanes of the wavefront that are valid and hold digit d (8 ballots, one per digit bit).
__device__ __forceinline__ uint64_t rs_match(int d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1;
    const uint64_t v = __ballot(bit);
    m &= bit ? v : ~v;
  }
  return m;
}

__device__ __forceinline__ uint64_t lanes_below(int lane) { return (lane ? (~0ull >> (64 - lane)) : 0ull); }

// hist[c * 256 + d] = items of chunk c with digit d
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const int16_t* __restrict__ keys, int64_t n, int shift,
                                                       int64_t chunk, int32_t* __restrict__ hist) {
  __shared__ int32_t h[kRsDigits];
  const int t = threadIdx.x, lane = t & 63;
  h[t] = 0;
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * chunk, e = s + chunk < n ? s + chunk : n;
  for (int64_t i0 = s; i0 < e; i0 += kRsThreads) {
    const int64_t i = i0 + t;
    const bool valid = i < e;
    const int d = valid ? rs_digit(keys[i], shift) : 0;
    const uint64_t m = rs_match(d, valid);
    if (valid && (m & lanes_below(lane)) == 0) atomicAdd(&h[d], __popcll(m));  // one add per digit and wave
  }
  __syncthreads();
  hist[(int64_t)blockIdx.x * kRsDigits + t] = h[t];
}

// One workgroup: hist[c][d] becomes the number of items with digit d in chunks before c, and
// dbase[d] the number of items with a smaller digit.
__global__ __launch_bounds__(kRsDigits) void k_rs_scan(int32_t* __restrict__ hist, int64_t n_chunks,
                                                      int32_t* __restrict__ dbase) {
  __shared__ int32_t s[kRsDigits];
  const int d = threadIdx.x;
  int32_t run = 0;
#pragma unroll 8
  for (int64_t c = 0; c < n_chunks; ++c) {
    const int32_t v = hist[c * kRsDigits + d];
    hist[c * kRsDigits + d] = run;
    run += v;
  }
  s[d] = run;
  __syncthreads();
  // inclusive Hillis-Steele scan of the 256 digit totals
  for (int off = 1; off < kRsDigits; off <<= 1) {
    const int32_t add = d >= off ? s[d - off] : 0;
    __syncthreads();
    s[d] += add;
    __syncthreads();
  }
  dbase[d] = s[d] - run;
}

// Stable scatter of chunk blockIdx.x: items keep their order within a digit.  vin == NULL: the
// values are the item indices.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const int16_t* __restrict__ kin, const int32_t* __restrict__ vin,
                                                          int16_t* __restrict__ kout, int32_t* __restrict__ vout, int64_t n,
                                                          int shift, int64_t chunk, const int32_t* __restrict__ hist,
                                                          const int32_t* __restrict__ dbase) {
  __shared__ int32_t base[kRsDigits];
  __shared__ int32_t wcnt[kRsThreads / 64][kRsDigits];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  base[t] = hist[(int64_t)blockIdx.x * kRsDigits + t] + dbase[t];
  const int64_t s = (int64_t)blockIdx.x * chunk, e = s + chunk < n ? s + chunk : n;
  for (int64_t i0 = s; i0 < e; i0 += kRsThreads) {
#pragma unroll
    for (int q = 0; q < kRsThreads / 64; ++q) wcnt[q][t] = 0;   // each thread clears its own digit column
    __syncthreads();
    const int64_t i = i0 + t;
    const bool valid = i < e;
    const int16_t k = valid ? kin[i] : (int16_t)0;
    const int32_t v = valid ? (vin ? vin[i] : (int32_t)i) : 0;
    const int d = rs_digit(k, shift);
    const uint64_t m = rs_match(d, valid);
    const int r = __popcll(m & lanes_below(lane));
    if (valid && r == 0) wcnt[w][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int32_t pos = base[d] + r;
      for (int q = 0; q < w; ++q) pos += wcnt[q][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    int32_t add = 0;
#pragma unroll
    for (int q = 0; q < kRsThreads / 64; ++q) add += wcnt[q][t];
    base[t] += add;   // read back by other threads only after the next tile's first barrier
  }
}

// Run heads of the sorted keys.  Per chunk: the number of positions i with i == 0 or
// keys[i] != keys[i-1].
__device__ __forceinline__ bool is_head(const int16_t* keys, int64_t i) { return i == 0 || keys[i] != keys[i - 1]; }

__global__ __launch_bounds__(kRsThreads) void k_heads_count(const int16_t* __restrict__ keys, int64_t n, int64_t chunk,
                                                           int32_t* __restrict__ part) {
  __shared__ int32_t wsum[kRsThreads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int64_t s = (int64_t)blockIdx.x * chunk, e = s + chunk < n ? s + chunk : n;
  int32_t c = 0;
  for (int64_t i = s + t; i < e; i += kRsThreads) c += is_head(keys, i) ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane == 0) wsum[w] = c;
  __syncthreads();
  if (t == 0) {
    int32_t tot = 0;
    for (int q = 0; q < kRsThreads / 64; ++q) tot += wsum[q];
    part[blockIdx.x] = tot;
  }
}

// One workgroup of 1024 threads: exclusive scan of v[0..len) in place, the total to *total.
// len_dev (optional) replaces len with a device count (clamped to len).
__global__ __launch_bounds__(1024) void k_scan_excl(int32_t* __restrict__ v, int64_t len, const int32_t* __restrict__ len_dev,
                                                   int32_t* __restrict__ total) {
  __shared__ int32_t wsum[16];
  __shared__ int32_t carry;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int64_t L = len;
  if (len_dev) {
    const int64_t d = *len_dev;
    L = d < len ? d : len;
  }
  if (t == 0) carry = 0;
  __syncthreads();
  for (int64_t i0 = 0; i0 < L; i0 += 1024) {
    const int64_t i = i0 + t;
    const int32_t x = i < L ? v[i] : 0;
    int32_t inc = x;   // inclusive wavefront scan
    for (int off = 1; off < 64; off <<= 1) {
      const int32_t y = __shfl_up(inc, off);
      if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int32_t before = carry;
    for (int q = 0; q < w; ++q) before += wsum[q];
    if (i < L) v[i] = before + inc - x;
    __syncthreads();
    if (t == 1023) carry = before + inc;
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// heads[part[c] + rank within the chunk] = i for every run head i of chunk c
__global__ __launch_bounds__(kRsThreads) void k_heads_write(const int16_t* __restrict__ keys, int64_t n, int64_t chunk,
                                                           const int32_t* __restrict__ part, int32_t* __restrict__ heads) {
  __shared__ int32_t wcnt[kRsThreads / 64];
  __shared__ int32_t base;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) base = part[blockIdx.x];
  const int64_t s = (int64_t)blockIdx.x * chunk, e = s + chunk < n ? s + chunk : n;
  for (int64_t i0 = s; i0 < e; i0 += kRsThreads) {
    const int64_t i = i0 + t;
    const bool f = i < e && is_head(keys, i);
    const uint64_t m = __ballot(f);
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    int32_t pos = base + __popcll(m & lanes_below(lane));
    for (int q = 0; q < w; ++q) pos += wcnt[q];
    if (f) heads[pos] = (int32_t)i;
    int32_t add = 0;   // read before the barrier: the next tile's counts overwrite wcnt
    if (t == 0)
      for (int q = 0; q < kRsThreads / 64; ++q) add += wcnt[q];
    __syncthreads();
    if (t == 0) base += add;   // read by the other threads only after the next tile's barrier
  }
}

// blk[r] = blocks of run r (ceil(run length / per_block)) for r < m, then (by k_scan_excl) offsets
__global__ __launch_bounds__(256) void k_run_blocks(const int32_t* __restrict__ heads, const int32_t* __restrict__ m,
                                                    int64_t n, int64_t cap, int32_t per_block, int32_t* __restrict__ blk) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t M = *m;
  if (r >= cap || r >= M) return;
  const int64_t end = r + 1 < M ? heads[r + 1] : n;
  blk[r] = (int32_t)((end - heads[r] + per_block - 1) / per_block);
}

// One thread per output block b <= max_blocks: the run of b (binary search over the block offsets),
// its group and first sorted position (the run's pixels split into blocks whose sizes differ by at most
// one, larger first, as assign_pixels does); entries past the last block get group -1 and start n, and
// block_starts[max_blocks] = n.
__global__ __launch_bounds__(256) void k_emit_blocks(const int16_t* __restrict__ keys, const int32_t* __restrict__ heads,
                                                     const int32_t* __restrict__ m, const int32_t* __restrict__ off,
                                                     const int32_t* __restrict__ total_blocks, int64_t n,
                                                     int64_t max_blocks, int16_t* __restrict__ block_net,
                                                     int32_t* __restrict__ block_starts, int32_t* __restrict__ n_blocks) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > max_blocks) return;
  const int64_t M = *m;
  const int64_t total = *total_blocks < max_blocks ? *total_blocks : max_blocks;
  if (b == 0) *n_blocks = (int32_t)total;
  if (b >= total) {
    if (b < max_blocks) block_net[b] = (int16_t)-1;
    block_starts[b] = (int32_t)n;
    return;
  }
  int64_t lo = 0, hi = M;   // off[lo] <= b < off[hi] (off[M] = total)
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid;
  }
  const int64_t r = lo;
  const int32_t h = heads[r];
  const int32_t cnt = (int32_t)((r + 1 < M ? heads[r + 1] : n) - h);
  const int32_t nb = (r + 1 < M ? off[r + 1] : (int32_t)*total_blocks) - off[r];
  const int32_t base = cnt / nb, add = cnt - base * nb;
  const int32_t j = (int32_t)(b - off[r]);
  block_net[b] = keys[h];
  block_starts[b] = h + j * base + (j < add ? j : add);
}

// sort_by_key's run compaction: keys[r] = sorted[heads[r]] and starts[r] = starts_in[heads[r]] for
// r < m; keys past m keep the sorted order.
__global__ __launch_bounds__(256) void k_sbk_compact(const int16_t* __restrict__ sorted, const int32_t* __restrict__ heads,
                                                     const int32_t* __restrict__ m, const int32_t* __restrict__ starts_in,
                                                     int64_t n, int16_t* __restrict__ keys, int32_t* __restrict__ starts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t M = *m;
  if (i < M) {
    keys[i] = sorted[heads[i]];
    starts[i] = starts_in[heads[i]];
  } else {
    keys[i] = sorted[i];
  }
}

__global__ __launch_bounds__(256) void k_copy_i32(const int32_t* __restrict__ a, int64_t n, int32_t* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ---- host side

static int64_t up256(int64_t v) { return (v + 255) / 256 * 256; }

// Workspace carve-up shared by both entry points.
struct RsWork {
  int16_t *k_tmp, *k_sorted;
  int32_t *v_tmp, *hist, *dbase, *part, *m, *heads, *off, *total;
};

static int64_t rs_work_bytes(int64_t n) {
  const RsPlan p = rs_plan(n);
  const int64_t g = n < kMaxGroups ? n : kMaxGroups;
  return 2 * up256(2 * n) + up256(4 * n) + up256(4 * p.n_chunks * kRsDigits) + up256(4 * kRsDigits) +
         up256(4 * (p.n_chunks + 1)) + up256(4) + up256(4 * (g + 1)) + up256(4 * (g + 1)) + up256(4);
}

static RsWork rs_carve(void* ws, int64_t n) {
  const RsPlan p = rs_plan(n);
  const int64_t g = n < kMaxGroups ? n : kMaxGroups;
  uint8_t* q = (uint8_t*)ws;
  RsWork w;
  auto take = [&](int64_t bytes) { uint8_t* r = q; q += up256(bytes); return r; };
  w.k_tmp = (int16_t*)take(2 * n);
  w.k_sorted = (int16_t*)take(2 * n);
  w.v_tmp = (int32_t*)take(4 * n);
  w.hist = (int32_t*)take(4 * p.n_chunks * kRsDigits);
  w.dbase = (int32_t*)take(4 * kRsDigits);
  w.part = (int32_t*)take(4 * (p.n_chunks + 1));
  w.m = (int32_t*)take(4);
  w.heads = (int32_t*)take(4 * (g + 1));
  w.off = (int32_t*)take(4 * (g + 1));
  w.total = (int32_t*)take(4);
  return w;
}

// Stable sort of keys[0..n) (values = vin, or the indices when vin is NULL) into k_out / v_out, then
// the run heads (w.heads[0..*w.m)).  n > 0.
static int rs_sort_and_heads(const int16_t* keys, const int32_t* vin, int64_t n, int16_t* k_out, int32_t* v_out,
                             const RsWork& w, hipStream_t st) {
  const RsPlan p = rs_plan(n);
  const dim3 gc((unsigned)p.n_chunks), bt(kRsThreads);
  // pass 1: low byte, keys -> tmp
  hipLaunchKernelGGL(k_rs_hist, gc, bt, 0, st, keys, n, 0, p.chunk, w.hist);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(kRsDigits), 0, st, w.hist, p.n_chunks, w.dbase);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_rs_scatter, gc, bt, 0, st, keys, vin, w.k_tmp, w.v_tmp, n, 0, p.chunk, w.hist, w.dbase);
  SNISR_LAUNCHED();
  // pass 2: high byte, tmp -> out
  hipLaunchKernelGGL(k_rs_hist, gc, bt, 0, st, w.k_tmp, n, 8, p.chunk, w.hist);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_rs_scan, dim3(1), dim3(kRsDigits), 0, st, w.hist, p.n_chunks, w.dbase);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_rs_scatter, gc, bt, 0, st, w.k_tmp, w.v_tmp, k_out, v_out, n, 8, p.chunk, w.hist, w.dbase);
  SNISR_LAUNCHED();
  // run heads of the sorted keys
  hipLaunchKernelGGL(k_heads_count, gc, bt, 0, st, k_out, n, p.chunk, w.part);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, w.part, p.n_chunks, (const int32_t*)nullptr, w.m);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_heads_write, gc, bt, 0, st, k_out, n, p.chunk, w.part, w.heads);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

static bool stream_capturing(hipStream_t st, bool* failed) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const hipError_t e = hipStreamIsCapturing(st, &cs);
  *failed = e != hipSuccess;
  if (e != hipSuccess) (void)hipGetLastError();
  return cs != hipStreamCaptureStatusNone;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_sort_by_key_workspace_size(int64_t n) {
  if (n < 0 || n >= (int64_t(1) << 31)) return -1;
  return rs_work_bytes(n) + up256(4 * n);
}

// A8 sort_by_key (rendering/src/reorder.cu:11-23): stable ascending sort of keys i16[n] carrying
// values i32[n], then the run heads: keys[0..m) = distinct keys, starts[0..m) = starts[] at the head
// of each run (the reference passes starts = arange(n), giving each run's first position).  Returns
// m in *n_unique (device-to-host copy, synchronises the stream as the reference's int return does), so
// it cannot run on a capturing stream: that is refused before any launch.
int snisr_sort_by_key(int16_t* keys, int32_t* values, int32_t* starts, int64_t n, void* workspace,
                      int64_t workspace_bytes, int32_t* n_unique, void* stream) {
  SNISR_REQUIRE(n_unique, SNISR_ERR_ARG, "sort_by_key: n_unique is NULL");
  SNISR_REQUIRE(n >= 0 && n < (int64_t(1) << 31), SNISR_ERR_ARG, "sort_by_key: bad size");
  if (n == 0) {
    *n_unique = 0;
    return SNISR_OK;
  }
  SNISR_REQUIRE(keys && values && starts && workspace, SNISR_ERR_ARG, "sort_by_key: NULL pointer");
  const int64_t need = snisr_sort_by_key_workspace_size(n);
  SNISR_REQUIRE(workspace_bytes >= need, SNISR_ERR_CAPACITY, "sort_by_key: workspace too small (%lld < %lld bytes)",
                (long long)workspace_bytes, (long long)need);
  const hipStream_t st = (hipStream_t)stream;
  bool failed = false;
  const bool capturing = stream_capturing(st, &failed);
  SNISR_REQUIRE(!failed, SNISR_ERR_DEVICE, "sort_by_key: hipStreamIsCapturing failed");
  SNISR_REQUIRE(!capturing, SNISR_ERR_UNSUPPORTED,
                "sort_by_key: returns the group count to the host, so it cannot run on a capturing stream "
                "(use snisr_regroup_pixels)");
  const RsWork w = rs_carve(workspace, n);
  int32_t* starts_copy = (int32_t*)((uint8_t*)workspace + rs_work_bytes(n));
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  // values are sorted into place (pass 2 reads the workspace copy); the sorted keys stay in k_sorted
  int rc = rs_sort_and_heads(keys, values, n, w.k_sorted, values, w, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_copy_i32, g, b, 0, st, starts, n, starts_copy);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_sbk_compact, g, b, 0, st, w.k_sorted, w.heads, w.m, starts_copy, n, keys, starts);
  SNISR_LAUNCHED();
  SNISR_HIP(hipMemcpyAsync(n_unique, w.m, 4, hipMemcpyDeviceToHost, st));
  SNISR_HIP(hipStreamSynchronize(st));
  return SNISR_OK;
}

int64_t snisr_regroup_max_blocks(int64_t n, int32_t pixels_per_block) {
  if (n < 0 || pixels_per_block <= 0) return -1;
  // sum over groups of ceil(count / P) <= n / P + (distinct groups), and at most 65,536 i16 groups
  return (n + pixels_per_block - 1) / pixels_per_block + (n < kMaxGroups ? n : kMaxGroups);
}

int64_t snisr_regroup_workspace_size(int64_t n) {
  if (n < 0 || n >= (int64_t(1) << 31)) return -1;
  return rs_work_bytes(n);
}

// The regrouping of FRender.render (sort_by_key, the host-side block_starts and assign_pixels:
// rendering/src/__init__.py:272-291, reorder.cu:11-94) on the device with no host synchronisation.
// The launch sequence depends only on n and max_blocks, and uses no memset, so it may be captured.
int snisr_regroup_pixels(const int16_t* net_idx, int64_t n, int32_t pixels_per_block, int64_t max_blocks,
                         int32_t* query, int16_t* block_net, int32_t* block_starts, int32_t* n_blocks,
                         void* workspace, int64_t workspace_bytes, void* stream) {
  SNISR_REQUIRE(n >= 0 && n < (int64_t(1) << 31), SNISR_ERR_ARG, "regroup_pixels: bad size");
  SNISR_REQUIRE(pixels_per_block > 0, SNISR_ERR_ARG, "regroup_pixels: pixels_per_block must be positive");
  SNISR_REQUIRE(max_blocks >= 0 && max_blocks < (int64_t(1) << 31), SNISR_ERR_ARG, "regroup_pixels: bad max_blocks");
  SNISR_REQUIRE(block_starts && n_blocks && workspace && (max_blocks == 0 || block_net), SNISR_ERR_ARG,
                "regroup_pixels: NULL pointer");
  SNISR_REQUIRE(n == 0 || (net_idx && query), SNISR_ERR_ARG, "regroup_pixels: NULL pointer");
  const int64_t mb_need = snisr_regroup_max_blocks(n, pixels_per_block);
  SNISR_REQUIRE(max_blocks >= mb_need, SNISR_ERR_CAPACITY,
                "regroup_pixels: max_blocks %lld below snisr_regroup_max_blocks(n, P) = %lld", (long long)max_blocks,
                (long long)mb_need);
  const int64_t need = snisr_regroup_workspace_size(n);
  SNISR_REQUIRE(need > 0 && workspace_bytes >= need, SNISR_ERR_CAPACITY,
                "regroup_pixels: workspace too small (%lld < %lld bytes)", (long long)workspace_bytes, (long long)need);
  const hipStream_t st = (hipStream_t)stream;
  const RsWork w = rs_carve(workspace, n);
  const int64_t g = n < kMaxGroups ? n : kMaxGroups;
  const dim3 b256(256);
  if (n > 0) {
    int rc = rs_sort_and_heads(net_idx, nullptr, n, w.k_sorted, query, w, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_run_blocks, dim3((unsigned)((g + 255) / 256)), b256, 0, st, w.heads, w.m, n, g,
                       pixels_per_block, w.off);
    SNISR_LAUNCHED();
  }
  // exclusive offsets of the runs' blocks; with n == 0 this writes m = 0 and a total of 0
  if (n == 0) {
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, w.part, (int64_t)0, (const int32_t*)nullptr, w.m);
    SNISR_LAUNCHED();
  }
  hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, w.off, g, (const int32_t*)w.m, w.total);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_emit_blocks, dim3((unsigned)((max_blocks + 1 + 255) / 256)), b256, 0, st, w.k_sorted, w.heads,
                     w.m, w.off, w.total, n, max_blocks, block_net, block_starts, n_blocks);
  SNISR_LAUNCHED();
  return SNISR_OK;
}

}  // extern "C"
