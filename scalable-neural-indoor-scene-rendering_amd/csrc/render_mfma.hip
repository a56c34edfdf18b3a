// MFMA render path: sigma-only march that counts, scan, march that emits per-sample records
// (contiguous per ray), per-tile MLP on v_mfma_f32_32x32x16_{f16,bf16} with FP32 accumulation,
// and an ordered per-ray reduction (deterministic pixels).
//
// MLP data flow (one wavefront = 64 records = two 32-sample column blocks):
//   D[32 out x 32 samples] = W[32 out x 16 k] * X[16 k x 32 samples]  per MFMA.
//   A operand (weights): lane l holds W[row l%32][k slots 8h..8h+7], h = l/32, 16 bytes per lane,
//   packed per (out-block, k-chunk) as 1 KB fragments so that a wavefront loads one fragment with
//   one coalesced 16-byte load per lane, straight from L2 (the per-tile weights are reused by every
//   wavefront of that tile).
//   B operand (activations): lane l holds X[k slots 8h..8h+7][sample l%32].
//   D: lane l holds rows 8(i/4) + 4h + (i%4), i = 0..15, of sample l%32.  After bias + ReLU these
//   16 values become the next layer's B operand for the two k-chunks (2*ob, 2*ob+1) without any
//   data movement; the next layer's weight columns are permuted to match (pack_input_index).
#include <atomic>
#include <vector>

#include "snisr_common.hpp"
#include "snisr_device.hpp"
#include "mfma_ops.hpp"
#include "train_dw.hpp"
#include "train_mma.hpp"

namespace snisr {


// ------------------------------------------------------------------------- packed layout
__host__ __device__ constexpr int nkc_of(int in) { return (in + 15) / 16; }
__host__ __device__ constexpr int nob_of(int out) { return (out + 31) / 32; }
__host__ __device__ constexpr int layer_in(int in0, int wd, int l) { return l == 0 ? in0 : wd; }
__host__ __device__ constexpr int layer_out(int wd, int L, int l) { return l == L - 1 ? 3 : wd; }
__host__ __device__ constexpr int64_t layer_bytes(int in0, int wd, int L, int l) {
  return (int64_t)nob_of(layer_out(wd, L, l)) * (nkc_of(layer_in(in0, wd, l)) * 1024 + 128);
}
__host__ __device__ constexpr int64_t layer_offset(int in0, int wd, int L, int l) {
  int64_t o = 0;
  for (int m = 0; m < l; ++m) o += layer_bytes(in0, wd, L, m);
  return o;
}
__host__ __device__ constexpr int64_t tile_bytes(int in0, int wd, int L) { return layer_offset(in0, wd, L, L); }


__device__ __forceinline__ uint16_t to_bits(float x, int dtype) {
  if (dtype == SNISR_BF16) {
    const __bf16 b = (__bf16)x;
    return *reinterpret_cast<const uint16_t*>(&b);
  }
  const _Float16 h = (_Float16)x;
  return *reinterpret_cast<const uint16_t*>(&h);
}

// one thread per packed 16-bit weight or per bias float of one layer of one tile
__global__ void k_pack_mfma_layer(const float* __restrict__ src, int64_t src_stride, int64_t src_off,
                                  uint8_t* __restrict__ dst, int64_t dst_stride, int64_t dst_off, int l, int in,
                                  int out, int dtype, int n_tiles) {
  const int nkc = nkc_of(in), nob = nob_of(out);
  const int64_t nw = (int64_t)nob * nkc * 512, nbias = (int64_t)nob * 32;
  const int64_t per_tile = nw + nbias;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= per_tile * n_tiles) return;
  const int64_t tile = g / per_tile;
  const int64_t r = g - tile * per_tile;
  const float* W = src + tile * src_stride + src_off;
  const float* b = W + (int64_t)out * in;
  uint8_t* D = dst + tile * dst_stride + dst_off;
  if (r < nw) {
    const int frag = (int)(r / 512), within = (int)(r % 512);
    const int ob = frag / nkc, kc = frag % nkc;
    const int lane = within / 8, e = within % 8;
    const int h = lane / 32, row = 32 * ob + lane % 32;
    const int x = pack_input_index(l, kc, h, e);
    const float v = (row < out && x < in) ? W[(int64_t)row * in + x] : 0.0f;
    reinterpret_cast<uint16_t*>(D)[r] = to_bits(v, dtype);
  } else {
    const int q = (int)(r - nw);  // (ob*2 + h)*16 + i
    const int ob = q / 32, h = (q / 16) % 2, i = q % 16;
    const int row = 32 * ob + 8 * (i / 4) + 4 * h + (i % 4);
    reinterpret_cast<float*>(D + (int64_t)nob * nkc * 1024)[q] = row < out ? b[row] : 0.0f;
  }
}

int64_t mfma_pack_bytes(const snisr_scene* s, int dtype) {
  if (!s || s->mlp_layers < 2 || s->mlp_width % 32 != 0 || s->mlp_in <= 0) {
    set_error("MFMA packing needs >= 2 layers and a hidden width that is a multiple of 32");
    return -1;
  }
  if (dtype != SNISR_F16 && dtype != SNISR_BF16) {
    set_error("MFMA packing dtype must be SNISR_F16 or SNISR_BF16");
    return -1;
  }
  return (int64_t)s->n_tiles * tile_bytes(s->mlp_in, s->mlp_width, s->mlp_layers);
}

int mfma_pack(const snisr_scene* s, int dtype, void* packed, void* stream) {
  if (mfma_pack_bytes(s, dtype) < 0) return SNISR_ERR_SHAPE;
  SNISR_REQUIRE(aligned(packed, 256), SNISR_ERR_ALIGN, "pack_mlp: packed buffer must be 256-byte aligned");
  const int in0 = s->mlp_in, wd = s->mlp_width, L = s->mlp_layers;
  const int64_t tb = tile_bytes(in0, wd, L);
  int64_t src_off = 0;
  for (int l = 0; l < L; ++l) {
    const int in = layer_in(in0, wd, l), out = layer_out(wd, L, l);
    const int64_t per_tile = (int64_t)nob_of(out) * nkc_of(in) * 512 + nob_of(out) * 32;
    const int64_t total = per_tile * s->n_tiles;
    hipLaunchKernelGGL(k_pack_mfma_layer, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, s->mlp_params, s->mlp_stride, src_off, (uint8_t*)packed, tb,
                       layer_offset(in0, wd, L, l), l, in, out, dtype, s->n_tiles);
    SNISR_LAUNCHED();
    src_off += (int64_t)out * in + out;
  }
  return SNISR_OK;
}

// ------------------------------------------------------------------------- workspace
struct Workspace {
  int64_t* total;       // header: samples emitted in total
  int32_t* overflow;    // header: 1 if a ray's records did not fit
  int32_t* counts;      // [n]
  int64_t* offsets;     // [n+1]
  int64_t* block_sums;  // [nblocks]
  float4* rec_f;        // [cap] voxel coordinate + weight   (ray-major)
  int2* rec_ti;         // [cap] tile, ray                    (ray-major)
  float4* rgbw;         // [cap] weighted colour of every record (ray-major)
  int32_t* perm;        // [cap + 64 T] record index in tile order, tile buckets 64-aligned, -1 = padding
  int32_t* tile_count;  // [T]
  int64_t* tile_start;  // [T+1] 64-aligned bucket starts
  int64_t* tile_cursor; // [T]
  uint16_t* enc;        // [n][16] per-ray view encoding in the MFMA operand type (15 values + 0)
  float4* rec_b;        // [cap] training: interval, T after the sample, clamped sigma, 0
  float4* crec;         // [cap] training: unweighted colour of every record
  float* dsig;          // [cap] training: dL/dsigma of every record
  float* act;           // training: activations/deltas per sorted record (train_act_floats each)
  float* frag;          // training: per-tile FP32 MFMA weight fragments (forward and transposed)
  int64_t cap;
  int32_t n_tiles;
};

constexpr int kScanItems = 8, kScanBlock = 256, kScanTile = kScanItems * kScanBlock;
constexpr int kMaxTiles = 4096;  // tile histogram in LDS

static inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// schedule keys: first occupied tile along the ray (T when none) x 8 direction octants, or the tile
// alone when that would exceed the LDS histogram
constexpr int kMaxBuckets = 8192;

static int64_t layout(int64_t n, int64_t cap, int32_t T, uint8_t* base, Workspace* w, int64_t train_floats = 0,
                      int64_t frag_floats = 0) {
  const int64_t nblocks = (n + kScanTile - 1) / kScanTile;
  int64_t off = 256;
  const int64_t o_counts = off; off = align_up(off + 4 * n, 256);
  const int64_t o_off = off; off = align_up(off + 8 * (n + 1), 256);
  const int64_t o_bs = off; off = align_up(off + 8 * (nblocks + 1), 256);
  const int64_t o_tc = off; off = align_up(off + 4 * (int64_t)T, 256);
  const int64_t o_ts = off; off = align_up(off + 8 * ((int64_t)T + 1), 256);
  const int64_t o_cu = off; off = align_up(off + 8 * (int64_t)T, 256);
  const int64_t o_enc = off; off = align_up(off + 32 * n, 256);
  const int64_t o_pm = off; off = align_up(off + 4 * (cap + 64 * (int64_t)T), 256);
  const int64_t o_rf = off; off = align_up(off + 16 * cap, 256);
  const int64_t o_ti = off; off = align_up(off + 8 * cap, 256);
  const int64_t o_rgb = off; off = align_up(off + 16 * cap, 256);
  int64_t o_rb = 0, o_cr = 0, o_ds = 0, o_act = 0, o_fr = 0;
  if (train_floats > 0) {
    o_fr = off; off = align_up(off + 4 * frag_floats, 256);
    o_rb = off; off = align_up(off + 16 * cap, 256);
    o_cr = off; off = align_up(off + 16 * cap, 256);
    o_ds = off; off = align_up(off + 4 * cap, 256);
    o_act = off; off = align_up(off + 4 * train_floats * (cap + 64 * (int64_t)T), 256);
  }
  if (w && base) {
    w->total = reinterpret_cast<int64_t*>(base);
    w->overflow = reinterpret_cast<int32_t*>(base + 8);
    w->counts = reinterpret_cast<int32_t*>(base + o_counts);
    w->offsets = reinterpret_cast<int64_t*>(base + o_off);
    w->block_sums = reinterpret_cast<int64_t*>(base + o_bs);
    w->tile_count = reinterpret_cast<int32_t*>(base + o_tc);
    w->tile_start = reinterpret_cast<int64_t*>(base + o_ts);
    w->tile_cursor = reinterpret_cast<int64_t*>(base + o_cu);
    w->enc = reinterpret_cast<uint16_t*>(base + o_enc);
    w->perm = reinterpret_cast<int32_t*>(base + o_pm);
    w->rec_f = reinterpret_cast<float4*>(base + o_rf);
    w->rec_ti = reinterpret_cast<int2*>(base + o_ti);
    w->rgbw = reinterpret_cast<float4*>(base + o_rgb);
    w->rec_b = train_floats > 0 ? reinterpret_cast<float4*>(base + o_rb) : nullptr;
    w->crec = train_floats > 0 ? reinterpret_cast<float4*>(base + o_cr) : nullptr;
    w->dsig = train_floats > 0 ? reinterpret_cast<float*>(base + o_ds) : nullptr;
    w->act = train_floats > 0 ? reinterpret_cast<float*>(base + o_act) : nullptr;
    w->frag = train_floats > 0 ? reinterpret_cast<float*>(base + o_fr) : nullptr;
    w->cap = cap;
    w->n_tiles = T;
  }
  return off;
}

static int64_t default_capacity(int64_t n) { return 48 * n; }

// Render workspace: one pass of the march writes every ray's records at fixed slots, ray i owning
// slots [i * cpr, (i + 1) * cpr) (cpr = record capacity per ray); no count pass, no scan.  A ray's
// records come in runs of one tile each (the ray walks its tiles in order); the ray's first four
// runs (tile, length, in order) sit in one 16-byte entry of runs4, any further run r at slot r of
// its rec_t region, so the tile counting sort works per run, not per record, reads the runs of
// consecutive rays with coalesced loads, and writes each run's slots contiguously into the tile's
// bucket of perm.  The MLP writes
// each record's weighted colour back into its slot of rec_f and the reduction sums every ray's
// slots in sample order.
struct RenderWs {
  int64_t* total;         // header: samples in total
  int32_t* overflow;      // header: 1 if a ray had more than cpr samples
  int32_t* max_count;     // header: largest per-ray sample count
  int32_t* counts;        // [n]
  uint16_t* enc;          // [n][16] per-ray view encoding in the MFMA operand type (15 values + 0)
  int32_t* order;         // [n] ray schedule: thread -> ray
  int32_t* ray_key;       // [n]
  int32_t* bucket_count;  // [n_buckets]
  int64_t* bucket_cursor; // [n_buckets]
  int32_t* tile_count;    // [T]
  int64_t* tile_start;    // [T+1] 64-aligned bucket starts
  int64_t* tile_cursor;   // [T]
  float4* rec_f;          // [n * cpr] voxel coordinate + weight; after the MLP: weighted colour
  int32_t* rec_t;         // [n * cpr] run r >= 4 of a ray (per-tile runs in sample order) at its slot r
  int4* runs4;            // [n] the ray's runs 0..3: tile | length << 16 each
  int32_t* perm;          // [scap] record slots in tile order, 64-aligned buckets, -1 = padding
  int32_t* skip;          // 1 when the sorted samples exceed scap (render skipped, rgb NaN)
  unsigned long long* cursor;  // per chunk: rays handed out by the queue-fed march
  int64_t n;
  int64_t scap;           // sorted positions: sample capacity + 64 per tile of bucket padding
  int32_t cpr;
  int32_t n_tiles;
  int32_t n_buckets;
  int32_t chunks;         // ray chunks with their own tile sort (tile_*, perm and skip per chunk)
};

// The view of ray chunk c = rays [r0, r0 + nc) of a chunked workspace: its records, counts and view
// encodings at local ray indices 0 .. nc-1, and its own tile arrays, sorted positions and skip flag,
// so the tile sort, the MLP and the reduction run on it unchanged.
static RenderWs chunk_view(const RenderWs& W, int c, int64_t r0, int64_t nc) {
  RenderWs v = W;
  const int64_t T = W.n_tiles;
  v.counts = W.counts + r0;
  v.enc = W.enc + 16 * r0;
  v.rec_f = W.rec_f + r0 * W.cpr;
  v.rec_t = W.rec_t + r0 * W.cpr;
  v.runs4 = W.runs4 + r0;
  v.tile_count = W.tile_count + c * T;
  v.tile_start = W.tile_start + c * (T + 1);
  v.tile_cursor = W.tile_cursor + c * T;
  v.perm = W.perm + c * W.scap;
  v.skip = W.skip + c;
  v.cursor = W.cursor + c;
  v.n = nc;
  v.chunks = 1;
  return v;
}

constexpr int32_t kDefaultRecordsPerRay = 64;
constexpr int64_t kDefaultSamplesPerRay = 16;  // default sample capacity per ray of the sorted arrays
constexpr int64_t kRenderBytesPerSlot = 16 + 4;

constexpr int64_t kSortedBytes = 4;  // bytes per tile-sorted position (perm)
constexpr int32_t kMaxChunks = 64;   // ray chunks of a pipelined render

static int64_t render_layout(int64_t n, int64_t cpr, int64_t samples, int32_t T, uint8_t* base, RenderWs* w,
                             int32_t chunks = 1) {
  int64_t off = 256;
  const int32_t nbk = kMaxBuckets;   // the largest key of any schedule mode
  const int64_t k = chunks > 0 ? chunks : 1;
  const int64_t slots = n * cpr, sorted = samples + 64 * (int64_t)T;   // per chunk
  const int64_t o_skip = off; off = align_up(off + 4 * k, 256);
  const int64_t o_qc = off; off = align_up(off + 8 * k, 256);
  const int64_t o_cnt = off; off = align_up(off + 4 * n, 256);
  const int64_t o_enc = off; off = align_up(off + 32 * n, 256);
  const int64_t o_ord = off; off = align_up(off + 4 * n, 256);
  const int64_t o_key = off; off = align_up(off + 4 * n, 256);
  const int64_t o_bkc = off; off = align_up(off + 4 * (int64_t)nbk, 256);
  const int64_t o_bku = off; off = align_up(off + 8 * (int64_t)nbk, 256);
  const int64_t o_tc = off; off = align_up(off + 4 * k * (int64_t)T, 256);
  const int64_t o_ts = off; off = align_up(off + 8 * k * ((int64_t)T + 1), 256);
  const int64_t o_cu = off; off = align_up(off + 8 * k * (int64_t)T, 256);
  const int64_t o_rf = off; off = align_up(off + 16 * slots, 256);
  const int64_t o_rt = off; off = align_up(off + 4 * slots, 256);
  const int64_t o_r4 = off; off = align_up(off + 16 * n, 256);
  const int64_t o_pm = off; off = align_up(off + 4 * k * sorted, 256);
  if (w && base) {
    w->total = reinterpret_cast<int64_t*>(base);
    w->overflow = reinterpret_cast<int32_t*>(base + 8);
    w->max_count = reinterpret_cast<int32_t*>(base + 12);
    w->skip = reinterpret_cast<int32_t*>(base + o_skip);
    w->cursor = reinterpret_cast<unsigned long long*>(base + o_qc);
    w->scap = sorted;
    w->counts = reinterpret_cast<int32_t*>(base + o_cnt);
    w->enc = reinterpret_cast<uint16_t*>(base + o_enc);
    w->order = reinterpret_cast<int32_t*>(base + o_ord);
    w->ray_key = reinterpret_cast<int32_t*>(base + o_key);
    w->bucket_count = reinterpret_cast<int32_t*>(base + o_bkc);
    w->bucket_cursor = reinterpret_cast<int64_t*>(base + o_bku);
    w->tile_count = reinterpret_cast<int32_t*>(base + o_tc);
    w->tile_start = reinterpret_cast<int64_t*>(base + o_ts);
    w->tile_cursor = reinterpret_cast<int64_t*>(base + o_cu);
    w->rec_f = reinterpret_cast<float4*>(base + o_rf);
    w->rec_t = reinterpret_cast<int32_t*>(base + o_rt);
    w->runs4 = reinterpret_cast<int4*>(base + o_r4);
    w->perm = reinterpret_cast<int32_t*>(base + o_pm);
    w->n = n;
    w->cpr = (int32_t)cpr;
    w->n_tiles = T;
    w->n_buckets = nbk;
    w->chunks = (int32_t)k;
  }
  return off;
}

// ------------------------------------------------------------------------- march passes
template <int MODE>  // 0: count, 1: emit records, 2: emit records + training extras
__global__ __launch_bounds__(64) void k_march_pass(snisr_scene S, snisr_march_params P, snisr_rays R,
                                                   snisr_render_out O, Workspace W, int packet_w, int enc_dtype,
                                                   const int32_t* __restrict__ order) {
  const int64_t tid = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (tid >= R.n) return;
  const int64_t i = order ? (int64_t)order[tid] : ray_of_thread(tid, R.n, packet_w);
  const float* og = R.origins + i * R.origin_stride;
  const V3 origin = v3(og[0], og[1], og[2]);
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  const float t_hit = R.t_hit ? R.t_hit[i] : __builtin_inff();
  const int32_t* vis = R.visited ? R.visited + i * P.max_tiles : nullptr;
  MarchResult res;
  if (MODE > 0) {
    const int64_t base = W.offsets[i];
    int64_t k = base;
    bool over = false;
    march_ray_any(packet_w > 0 && !order, S, P, origin, d, t_hit, vis, nullptr, 0,
              [&](int tile, const V3& f, const Corners&, float w, float sig, float interval, float T_after) {
                if (k < W.cap) {
                  W.rec_f[k] = make_float4(f.x, f.y, f.z, w);
                  W.rec_ti[k] = make_int2(tile, (int)i);
                  if (MODE == 2) W.rec_b[k] = make_float4(interval, T_after, sig, 0.0f);
                } else {
                  over = true;
                }
                ++k;
              },
              res);
    if (over) *W.overflow = 1;
  } else {
    int32_t* bricks = O.bricks ? O.bricks + i * O.brick_cap : nullptr;
    march_ray_any(packet_w > 0 && !order, S, P, origin, d, t_hit, vis, bricks, O.brick_cap, [&](int, const V3&, const Corners&, float, float, float, float) {}, res);
    if (bricks)
      for (int k = res.n_bricks; k < O.brick_cap; ++k) bricks[k] = -1;
    W.counts[i] = res.n_samples;
    float enc[15];
    view_encoding(d, enc);
    uint16_t e16[16];
#pragma unroll
    for (int k = 0; k < 15; ++k) e16[k] = to_bits(enc[k], enc_dtype);
    e16[15] = 0;
    uint4* dst = reinterpret_cast<uint4*>(W.enc + i * 16);
    dst[0] = *reinterpret_cast<const uint4*>(e16);
    dst[1] = *reinterpret_cast<const uint4*>(e16 + 8);
    if (O.T) O.T[i] = res.T;
    if (O.n_samples) O.n_samples[i] = res.n_samples;
    if (O.depth) O.depth[i] = res.depth;
  }
}

// ------------------------------------------------------------------------- exclusive scan
__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// block-level exclusive scan of one value per thread; returns the block total via *total
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total) {
  __shared__ int64_t wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t incl = wave_incl_scan(v);
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  int64_t before = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kScanBlock / 64; ++k) {
    if (k < wid) before += wsum[k];
    all += wsum[k];
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_local(const int32_t* __restrict__ counts, int64_t n,
                                                           int64_t* __restrict__ offsets, int64_t* __restrict__ bsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? counts[base + k] : 0;
    s += v[k];
  }
  int64_t total;
  int64_t run = block_excl_scan(s, &total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) offsets[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(int64_t* __restrict__ bsum, int64_t nblocks,
                                                            int64_t* __restrict__ offsets, int64_t n,
                                                            int64_t* __restrict__ total_out,
                                                            int32_t* __restrict__ overflow,
                                                            int32_t* __restrict__ tile_count, int n_tiles) {
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x) tile_count[t] = 0;
  int64_t carry = 0;
  for (int64_t start = 0; start < nblocks; start += kScanBlock) {
    const int64_t j = start + threadIdx.x;
    const int64_t v = j < nblocks ? bsum[j] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (j < nblocks) bsum[j] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) {
    offsets[n] = carry;
    *total_out = carry;
    *overflow = 0;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(int64_t* __restrict__ offsets, int64_t n,
                                                         const int64_t* __restrict__ bsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const int64_t add = bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < n) offsets[base + k] += add;
}


// 8 consecutive features of one corner voxel as floats
template <bool HALF>
__device__ __forceinline__ void load8(const void* __restrict__ feats, int64_t off, uint32_t m, float* v) {
  if (HALF) {
    FeatLoad<true>::load(feats, off, m, v);
  } else {
    FeatLoad<false>::load(feats, off, m, v);
    FeatLoad<false>::load(feats, off + 4, m, v + 4);
  }
}

// ------------------------------------------------------------------------- tile sort
// Counting sort of the ray-major records by tile into 64-aligned buckets (perm holds record
// indices; padding slots are -1).  The order inside a bucket is not deterministic, but every
// record's result goes to its own ray-major slot and does not depend on its neighbours, so the
// pixels are.
constexpr int kSortBlock = 1024, kSortGrid = 512;

__device__ __forceinline__ void record_range(int64_t total, int64_t& b, int64_t& e) {
  const int64_t per = (total + gridDim.x - 1) / gridDim.x;
  b = (int64_t)blockIdx.x * per;
  e = b + per < total ? b + per : total;
}

__global__ __launch_bounds__(kSortBlock) void k_tile_hist(Workspace W) {
  __shared__ int hist[kMaxTiles];
  const int T = W.n_tiles;
  for (int t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const int64_t total = *W.total < W.cap ? *W.total : W.cap;
  int64_t b, e;
  record_range(total, b, e);
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) atomicAdd(&hist[W.rec_ti[k].x], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += blockDim.x)
    if (hist[t]) atomicAdd(&W.tile_count[t], hist[t]);
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan(Workspace W) {
  const int T = W.n_tiles;
  int64_t carry = 0;
  for (int start = 0; start < T; start += kScanBlock) {
    const int t = start + threadIdx.x;
    const int64_t v = t < T ? (int64_t)((W.tile_count[t] + 63) & ~63) : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (t < T) {
      W.tile_start[t] = carry + ex;
      W.tile_cursor[t] = carry + ex;
    }
    carry += tot;
  }
  if (threadIdx.x == 0) W.tile_start[T] = carry;
  __syncthreads();
  // padding slots at the end of every bucket
  for (int t = 0; t < T; ++t) {
    const int64_t pb = W.tile_start[t] + W.tile_count[t], pe = W.tile_start[t + 1];
    for (int64_t k = pb + threadIdx.x; k < pe; k += blockDim.x) W.perm[k] = -1;
  }
}

__global__ __launch_bounds__(kSortBlock) void k_tile_scatter(Workspace W) {
  __shared__ int hist[kMaxTiles];
  __shared__ int64_t base[kMaxTiles];
  const int T = W.n_tiles;
  for (int t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  const int64_t total = *W.total < W.cap ? *W.total : W.cap;
  int64_t b, e;
  record_range(total, b, e);
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) atomicAdd(&hist[W.rec_ti[k].x], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += blockDim.x) {
    base[t] = hist[t] ? atomicAdd((unsigned long long*)&W.tile_cursor[t], (unsigned long long)hist[t]) : 0;
    hist[t] = 0;
  }
  __syncthreads();
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    const int t = W.rec_ti[k].x;
    const int r = atomicAdd(&hist[t], 1);
    W.perm[base[t] + r] = (int32_t)k;
  }
}




// ------------------------------------------------------------------------- ray schedule
// Orders the rays so that the wavefronts of the march walk the same bricks together: a counting sort
// (LDS histograms, one scan, atomic bucket cursors) by a key of
//   mode 1: (first occupied tile, direction octant);
//   mode 2: (first occupied tile, direction octant, origin cell of a c^3 grid over the scene box);
//   mode 3: (first occupied tile, origin cell of a c^3 grid over the scene box);
// c is the largest cell count per axis whose buckets fit the LDS histogram.  Only the thread -> ray
// map changes; every ray's records and pixel are the same as without the schedule.
struct KeyShape {
  int32_t mode, cells, n_buckets;
};

static KeyShape key_shape(int mode, int32_t T) {
  KeyShape k{mode, 1, T + 1};
  if (mode == 1) {
    k.n_buckets = 8 * (T + 1) <= kMaxBuckets ? 8 * (T + 1) : T + 1;
  } else if (mode >= 2) {
    const int per = kMaxBuckets / ((T + 1) * (mode == 2 ? 8 : 1));
    int c = 1;
    while ((c + 1) * (c + 1) * (c + 1) <= per) ++c;
    k.cells = c;
    k.n_buckets = (T + 1) * (mode == 2 ? 8 : 1) * c * c * c;
    if (mode == 2 && per < 1) k = KeyShape{1, 1, T + 1};   // too many tiles for octants: first tile only
  }
  return k;
}

__global__ __launch_bounds__(256) void k_ray_key(snisr_scene S, snisr_rays R, RenderWs W, KeyShape K) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < W.n_buckets) W.bucket_count[i] = 0;
  if (i >= R.n) return;
  const float* og = R.origins + i * R.origin_stride;
  const V3 o = v3(og[0], og[1], og[2]);
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  TileWalker w;
  w.init(S, o, d, 1);
  int t = w.next();
  if (t < 0) t = W.n_tiles;
  const int oct = (d.x < 0.0f ? 1 : 0) + (d.y < 0.0f ? 2 : 0) + (d.z < 0.0f ? 4 : 0);
  int key = t;
  if (K.mode == 1) {
    key = W.n_buckets == 8 * (W.n_tiles + 1) ? 8 * t + oct : t;
  } else if (K.mode >= 2) {
    const int c = K.cells;
    int cell = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float ext = (float)S.tile_shape[a] * S.tile_size;
      int q = (int)((get(o, a) - S.scene_min[a]) / ext * (float)c);
      q = q < 0 ? 0 : (q >= c ? c - 1 : q);
      cell = cell * c + q;
    }
    key = K.mode == 2 ? (t * 8 + oct) * c * c * c + cell : t * c * c * c + cell;
  }
  W.ray_key[i] = key;
}

__device__ __forceinline__ void ray_range(int64_t n, int64_t& b, int64_t& e) {
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  b = (int64_t)blockIdx.x * per;
  e = b + per < n ? b + per : n;
}

__global__ __launch_bounds__(kSortBlock) void k_key_hist(RenderWs W, int64_t n) {
  __shared__ int hist[kMaxBuckets];
  const int nb = W.n_buckets;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  int64_t b, e;
  ray_range(n, b, e);
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) atomicAdd(&hist[W.ray_key[k]], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += blockDim.x)
    if (hist[t]) atomicAdd(&W.bucket_count[t], hist[t]);
}

__global__ __launch_bounds__(kScanBlock) void k_key_scan(RenderWs W) {
  const int nb = W.n_buckets;
  int64_t carry = 0;
  for (int start = 0; start < nb; start += kScanBlock) {
    const int t = start + threadIdx.x;
    const int64_t v = t < nb ? (int64_t)W.bucket_count[t] : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (t < nb) W.bucket_cursor[t] = carry + ex;
    carry += tot;
  }
}

__global__ __launch_bounds__(kSortBlock) void k_key_scatter(RenderWs W, int64_t n) {
  __shared__ int hist[kMaxBuckets];
  __shared__ int64_t base[kMaxBuckets];
  const int nb = W.n_buckets;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  int64_t b, e;
  ray_range(n, b, e);
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) atomicAdd(&hist[W.ray_key[k]], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    base[t] = hist[t] ? atomicAdd((unsigned long long*)&W.bucket_cursor[t], (unsigned long long)hist[t]) : 0;
    hist[t] = 0;
  }
  __syncthreads();
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    const int t = W.ray_key[k];
    const int r = atomicAdd(&hist[t], 1);
    W.order[base[t] + r] = (int32_t)k;
  }
}

// ------------------------------------------------------------------------- render: one march pass
__global__ void k_render_begin(RenderWs W) {
  if (threadIdx.x == 0) {
    *W.total = 0;
    *W.overflow = 0;
    *W.max_count = 0;
  }
  for (int c = threadIdx.x; c < W.chunks; c += blockDim.x) {
    W.skip[c] = 0;
    W.cursor[c] = 0;
  }
  for (int t = threadIdx.x; t < W.chunks * W.n_tiles; t += blockDim.x) W.tile_count[t] = 0;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// The runs of one ray as a march emits them (tile | length << 16, in sample order): runs 0..3 stay
// in registers and go out as one 16-byte store at the ray's end, later ones straight to the ray's
// rec_t slots.
struct RunList {
  int v0, v1, v2, v3, n;
  __device__ __forceinline__ void reset() {
    v0 = v1 = v2 = v3 = 0;
    n = 0;
  }
  __device__ __forceinline__ void add(int v, int32_t* rt) {
    if (n == 0) v0 = v;
    else if (n == 1) v1 = v;
    else if (n == 2) v2 = v;
    else if (n == 3) v3 = v;
    else rt[n] = v;
    ++n;
  }
  __device__ __forceinline__ void store(int4* dst) const { *dst = make_int4(v0, v1, v2, v3); }
};

// One wavefront per 64 rays: march each ray once, writing its records (voxel coordinate + weight,
// tile) at its own slots, its sample count, T, depth and view encoding; per-wave totals feed the
// header (sample total, largest count, overflow when a ray has more samples than slots).
// Camera rays in 8x8 pixel packets (one packet per wavefront), marched with the nested tile / brick /
// sample loops; arbitrary rays go through k_march_queue.
__global__ __launch_bounds__(64) void k_march_render(snisr_scene S, snisr_march_params P, snisr_rays R,
                                                     snisr_render_out O, RenderWs W, int packet_w, int enc_dtype,
                                                     const int32_t* __restrict__ order, int64_t tid0, int64_t tid1) {
  // Records are staged per lane in LDS and written 8 at a time (128 bytes, one cache line when the
  // slot count is a multiple of 8): lines written piecemeal by many rays in flight would leave L2
  // partly written, several times over.
  __shared__ float4 rbuf[8][64];
  const int lane = threadIdx.x;
  const int64_t tid = tid0 + (int64_t)blockIdx.x * 64 + threadIdx.x;
  int ns = 0;
  if (tid < tid1) {
    const int64_t i = order ? (int64_t)order[tid] : ray_of_thread(tid, R.n, packet_w);
    const float* og = R.origins + i * R.origin_stride;
    const V3 origin = v3(og[0], og[1], og[2]);
    const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
    const float t_hit = R.t_hit ? R.t_hit[i] : __builtin_inff();
    const int32_t* vis = R.visited ? R.visited + i * P.max_tiles : nullptr;
    const int cpr = W.cpr;
    const int64_t slot0 = i * (int64_t)cpr;
    int k = 0, run_tile = -1, run_start = 0;
    RunList rl;
    rl.reset();
    MarchResult res;
    int32_t* bricks = O.bricks ? O.bricks + i * O.brick_cap : nullptr;
    // this lane's staged records 0..cnt-1 go to slots base.. of its ray
    auto flush = [&](int base, int cnt) {
      float4* dst = W.rec_f + slot0 + base;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < cnt) dst[j] = rbuf[j][lane];
    };
    auto on_sample = [&](int tile, const V3& f, const Corners&, float w, float, float, float) {
                if (k < cpr) {
                  if (tile != run_tile) {
                    if (run_tile >= 0) rl.add(run_tile | ((k - run_start) << 16), W.rec_t + slot0);
                    run_tile = tile;
                    run_start = k;
                  }
                  const int j = k & 7;
                  rbuf[j][lane] = make_float4(f.x, f.y, f.z, w);
                  if (j == 7) flush(k - 7, 8);
                }
                ++k;
              };
    march_ray(S, P, origin, d, t_hit, vis, bricks, O.brick_cap, on_sample, res);
    const int nrec = min(k, cpr);
    if (nrec & 7) flush(nrec & ~7, nrec & 7);
    if (run_tile >= 0) rl.add(run_tile | ((nrec - run_start) << 16), W.rec_t + slot0);
    rl.store(W.runs4 + i);
    if (bricks)
      for (int b = res.n_bricks; b < O.brick_cap; ++b) bricks[b] = -1;
    W.counts[i] = res.n_samples;
    float enc[15];
    view_encoding(d, enc);
    uint16_t e16[16];
#pragma unroll
    for (int q = 0; q < 15; ++q) e16[q] = to_bits(enc[q], enc_dtype);
    e16[15] = 0;
    uint4* dst = reinterpret_cast<uint4*>(W.enc + i * 16);
    dst[0] = *reinterpret_cast<const uint4*>(e16);
    dst[1] = *reinterpret_cast<const uint4*>(e16 + 8);
    if (O.T) O.T[i] = res.T;
    if (O.n_samples) O.n_samples[i] = res.n_samples;
    if (O.depth) O.depth[i] = res.depth;
    ns = res.n_samples;
  }
  const int64_t sum = wave_sum_i64((int64_t)ns);
  const int mx = wave_max_i32(ns);
  if (threadIdx.x == 0) {
    if (sum) atomicAdd((unsigned long long*)W.total, (unsigned long long)sum);
    if (mx) atomicMax(W.max_count, mx);
    if (mx > W.cpr) *W.overflow = 1;
  }
}

// 15-value view encoding of every ray, in the MLP's operand type (the packet march writes it itself)
__global__ __launch_bounds__(256) void k_ray_enc(snisr_rays R, RenderWs W, int enc_dtype, int64_t r0, int64_t r1) {
  const int64_t i = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= r1) return;
  const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
  float enc[15];
  view_encoding(d, enc);
  uint16_t e16[16];
#pragma unroll
  for (int q = 0; q < 15; ++q) e16[q] = to_bits(enc[q], enc_dtype);
  e16[15] = 0;
  uint4* dst = reinterpret_cast<uint4*>(W.enc + i * 16);
  dst[0] = *reinterpret_cast<const uint4*>(e16);
  dst[1] = *reinterpret_cast<const uint4*>(e16 + 8);
}

// The march of arbitrary rays (no pixel packets, or a ray schedule) as a queue: a wavefront's lanes
// take rays (thread indices tid0 .. tid1 - 1, ray = order[tid] or tid) from a per-chunk counter and,
// when a quarter of the lanes have finished theirs, take new ones together, so that a wavefront is
// not held by its longest ray while most lanes idle.  Each ray is marched exactly as by
// march_ray_flat (same steps, same records), only the lanes that run it change.  The grid is sized
// to the device's resident wavefronts; every wavefront ends once the counter passes tid1 and its
// lanes' rays are done.
constexpr int kRefillQuorum = 4;
__global__ __launch_bounds__(64) void k_march_queue(snisr_scene S, snisr_march_params P, snisr_rays R,
                                                    snisr_render_out O, RenderWs W, const int32_t* __restrict__ order,
                                                    int64_t tid0, int64_t tid1) {
  __shared__ float4 rbuf[8][64];   // records staged per lane, written 8 at a time (see k_march_render)
  const int lane = threadIdx.x;
  const int cpr = W.cpr;
  FlatMarch m;
  m.phase = FlatMarch::kDone;
  bool has = false, more = true;
  int64_t i = 0, slot0 = 0;
  int kr = 0, run_tile = -1, run_start = 0;
  RunList rl;
  rl.reset();
  int32_t* bricks = nullptr;
  int64_t sum = 0;
  int mx = 0;
  auto flush = [&](int base, int cnt) {
    float4* dst = W.rec_f + slot0 + base;
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (j < cnt) dst[j] = rbuf[j][lane];
  };
  auto on_sample = [&](int tile, const V3& f, const Corners&, float w, float, float, float) {
    if (kr < cpr) {
      if (tile != run_tile) {
        if (run_tile >= 0) rl.add(run_tile | ((kr - run_start) << 16), W.rec_t + slot0);
        run_tile = tile;
        run_start = kr;
      }
      const int j = kr & 7;
      rbuf[j][lane] = make_float4(f.x, f.y, f.z, w);
      if (j == 7) flush(kr - 7, 8);
    }
    ++kr;
  };
  while (true) {
    if (has && m.phase == FlatMarch::kDone) {  // finish this lane's ray
      const int nrec = min(kr, cpr);
      if (nrec & 7) flush(nrec & ~7, nrec & 7);
      if (run_tile >= 0) rl.add(run_tile | ((nrec - run_start) << 16), W.rec_t + slot0);
      rl.store(W.runs4 + i);
      if (bricks)
        for (int b = m.nbr; b < O.brick_cap; ++b) bricks[b] = -1;
      W.counts[i] = m.nsamp;
      if (O.T) O.T[i] = m.T;
      if (O.n_samples) O.n_samples[i] = m.nsamp;
      if (O.depth) O.depth[i] = m.depth;
      sum += m.nsamp;
      mx = m.nsamp > mx ? m.nsamp : mx;
      has = false;
    }
    const uint64_t idle = __builtin_amdgcn_ballot_w64(!has && more);
    const uint64_t busy = __builtin_amdgcn_ballot_w64(has);
    if (idle && (busy == 0 || kRefillQuorum * __builtin_popcountll(idle) >= 64)) {
      const int first = (int)__builtin_ctzll(idle);
      unsigned long long base = 0;
      if (lane == first) base = atomicAdd(W.cursor, (unsigned long long)__builtin_popcountll(idle));
      base = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(base >> 32), first) << 32) |
             (uint32_t)__shfl((int)(uint32_t)base, first);
      if (!has && more) {
        const int64_t t = tid0 + (int64_t)base + __builtin_popcountll(idle & ((1ull << lane) - 1));
        if (t < tid1) {
          i = order ? (int64_t)order[t] : t;
          const float* og = R.origins + i * R.origin_stride;
          const V3 d = normalize_dir(v3(R.dirs[3 * i], R.dirs[3 * i + 1], R.dirs[3 * i + 2]));
          m.init(S, P, v3(og[0], og[1], og[2]), d, R.t_hit ? R.t_hit[i] : __builtin_inff(),
                 R.visited ? R.visited + i * P.max_tiles : nullptr);
          slot0 = i * (int64_t)cpr;
          kr = 0;
          run_tile = -1;
          run_start = 0;
          rl.reset();
          bricks = O.bricks ? O.bricks + i * O.brick_cap : nullptr;
          has = true;
        } else {
          more = false;
        }
      }
    }
    if (__builtin_amdgcn_ballot_w64(has) == 0) break;
    if (has) {
      const bool turn = tile_quorum(m.phase == FlatMarch::kTile);
      m.step(S, P, turn, bricks, O.brick_cap, on_sample);
    }
  }
  const int64_t wsum = wave_sum_i64(sum);
  const int wmx = wave_max_i32(mx);
  if (lane == 0) {
    if (wsum) atomicAdd((unsigned long long*)W.total, (unsigned long long)wsum);
    if (wmx) atomicMax(W.max_count, wmx);
    if (wmx > W.cpr) *W.overflow = 1;
  }
}

// rays [b, e) of this workgroup
__device__ __forceinline__ void wg_ray_range(int64_t n, int64_t& b, int64_t& e) {
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  b = (int64_t)blockIdx.x * per;
  e = b + per < n ? b + per : n;
}

// Calls f(tile, first record, length) for the runs of ray i, in order: runs 0..3 from the ray's
// runs4 entry (one 16-byte load, coalesced over consecutive rays), later ones from its rec_t slots.
template <class F>
__device__ __forceinline__ void for_each_run(const RenderWs& W, int64_t i, F f) {
  const int cpr = W.cpr;
  const int c = min(W.counts[i], cpr);
  if (c <= 0) return;
  const int4 q = W.runs4[i];
  int k = 0;
  auto run = [&](int v) {  // false once the runs cover the ray's records
    const int len = v >> 16;
    if (len <= 0) return false;  // unreachable for runs written by the march; keeps the loop finite
    f(v & 0xffff, k, len);
    k += len;
    return k < c;
  };
  if (!run(q.x) || !run(q.y) || !run(q.z) || !run(q.w)) return;
  const int32_t* rt = W.rec_t + i * (int64_t)cpr;
  for (int r = 4; r < cpr; ++r)
    if (!run(rt[r])) return;
}

// per-tile record counts of this workgroup's rays into LDS, one atomic per tile run
__device__ __forceinline__ void rtile_count_block(const RenderWs& W, int* hist, int64_t b, int64_t e) {
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x)
    for_each_run(W, i, [&](int t, int, int len) { atomicAdd(&hist[t], len); });
}

__global__ __launch_bounds__(kSortBlock) void k_rtile_hist(RenderWs W) {
  __shared__ int hist[kMaxTiles];
  const int T = W.n_tiles;
  for (int t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0;
  __syncthreads();
  int64_t b, e;
  wg_ray_range(W.n, b, e);
  rtile_count_block(W, hist, b, e);
  __syncthreads();
  for (int t = threadIdx.x; t < T; t += blockDim.x)
    if (hist[t]) atomicAdd(&W.tile_count[t], hist[t]);
}

__global__ __launch_bounds__(kScanBlock) void k_rtile_scan(RenderWs W) {
  const int T = W.n_tiles;
  int64_t carry = 0;
  for (int start = 0; start < T; start += kScanBlock) {
    const int t = start + threadIdx.x;
    const int64_t v = t < T ? (int64_t)((W.tile_count[t] + 63) & ~63) : 0;
    int64_t tot;
    const int64_t ex = block_excl_scan(v, &tot);
    if (t < T) {
      W.tile_start[t] = carry + ex;
      W.tile_cursor[t] = carry + ex;
    }
    carry += tot;
  }
  if (threadIdx.x == 0) W.tile_start[T] = carry;
  if (carry > W.scap) {  // more samples than sorted positions: skip the rest of the render
    if (threadIdx.x == 0) {
      *W.skip = 1;
      *W.overflow = 1;
    }
    return;
  }
  __syncthreads();
  // padding slots at the end of every bucket
  for (int t = 0; t < T; ++t) {
    const int64_t pb = W.tile_start[t] + W.tile_count[t], pe = W.tile_start[t + 1];
    for (int64_t k = pb + threadIdx.x; k < pe; k += blockDim.x) W.perm[k] = -1;
  }
}

// Scatter of the records into the tile buckets, in chunks of kSortBlock rays (one per thread).  Each
// chunk reserves its records' positions per tile with one global atomic per tile, stages the
// record indices in LDS grouped by tile, and then writes each tile's range as consecutive words:
// written straight from the runs, a wavefront's 4-byte stores would land in 64 places at once.  A
// chunk whose records exceed the staging buffer (or more than kSortBlock tiles) writes directly.
constexpr int kScatterLds = 60 * 1024;  // dynamic LDS of k_rtile_scatter (with the static part, under 64 KB)
// dst[j] = v0 + j for j < len: single words up to a 16-byte boundary, then four words per store
__device__ __forceinline__ void write_run(int32_t* dst, int32_t v0, int len) {
  int j = 0;
  for (; j < len && (reinterpret_cast<uintptr_t>(dst + j) & 15); ++j) dst[j] = v0 + j;
  for (; j + 4 <= len; j += 4) *reinterpret_cast<int4*>(dst + j) = make_int4(v0 + j, v0 + j + 1, v0 + j + 2, v0 + j + 3);
  for (; j < len; ++j) dst[j] = v0 + j;
}
__global__ __launch_bounds__(kSortBlock) void k_rtile_scatter(RenderWs W) {
  extern __shared__ __align__(16) unsigned char sc_lds[];
  __shared__ int wsum[kSortBlock / 64];
  __shared__ int chunk_total;
  if (*W.skip) return;
  const int T = W.n_tiles;
  int64_t* base = reinterpret_cast<int64_t*>(sc_lds);                 // [T]
  int* hist = reinterpret_cast<int*>(sc_lds + 8 * (size_t)T);         // [T]  counts, then cursors
  int* lofs = hist + T;                                                // [T + 1] local offsets
  int32_t* stage = lofs + T + 1;                                       // the rest
  const int n_stage = (int)((kScatterLds - 16 * (size_t)T - 4) / 4);
  const bool can_stage = T <= kSortBlock;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int64_t b, e;
  wg_ray_range(W.n, b, e);
  for (int64_t c0 = b; c0 < e; c0 += kSortBlock) {
    const int64_t i = c0 + threadIdx.x;
    const bool live = i < e;
    for (int t = threadIdx.x; t < T; t += blockDim.x) hist[t] = 0;
    __syncthreads();
    if (live) for_each_run(W, i, [&](int t, int, int len) { atomicAdd(&hist[t], len); });
    __syncthreads();
    if (can_stage) {  // exclusive offsets of the tiles inside the chunk (one tile per thread)
      const int v = threadIdx.x < T ? hist[threadIdx.x] : 0;
      const int incl = (int)wave_incl_scan((int64_t)v);
      if (lane == 63) wsum[wid] = incl;
      __syncthreads();
      int before = 0, all = 0;
#pragma unroll
      for (int k = 0; k < kSortBlock / 64; ++k) {
        before += k < wid ? wsum[k] : 0;
        all += wsum[k];
      }
      if (threadIdx.x < T) lofs[threadIdx.x] = before + incl - v;
      if (threadIdx.x == 0) {
        lofs[T] = all;
        chunk_total = all;
      }
    }
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
      base[t] = hist[t] ? atomicAdd((unsigned long long*)&W.tile_cursor[t], (unsigned long long)hist[t]) : 0;
      hist[t] = 0;
    }
    __syncthreads();
    const int total = can_stage ? chunk_total : 0;
    const bool staged = can_stage && total <= n_stage;
    if (live) {
      const int64_t slot0 = i * (int64_t)W.cpr;
      for_each_run(W, i, [&](int t, int k, int len) {
        const int p = atomicAdd(&hist[t], len);
        const int32_t v0 = (int32_t)(slot0 + k);
        if (staged) write_run(stage + lofs[t] + p, v0, len);
        else write_run(W.perm + base[t] + p, v0, len);
      });
    }
    __syncthreads();
    if (staged) {
      // the tile whose range holds q is the last tile starting at or before q; q grows along a
      // thread's loop, so the search starts at the previous q's tile: a few steps forward, then a
      // binary search over the rest when the tile is farther
      int lo = 0;
      for (int q = threadIdx.x; q < total; q += blockDim.x) {
        for (int step = 0; step < 4 && lo + 1 < T && lofs[lo + 1] <= q; ++step) ++lo;
        if (lo + 1 < T && lofs[lo + 1] <= q) {
          int hi = T - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (lofs[mid] <= q) lo = mid;
            else hi = mid - 1;
          }
        }
        W.perm[base[lo] + (q - lofs[lo])] = stage[q];
      }
    }
    __syncthreads();  // hist, lofs and stage are reused by the next chunk
  }
}

// per-ray sum of the weighted colours (NaN for a ray whose samples exceeded its slots): four lanes
// per ray, lane j summing records j, j + 4, ... in order, so that a wavefront reads 16 rays' records
// as 64-byte pieces; the four partial sums are then added as (s0 + s1) + (s2 + s3).  The order is
// fixed, so the result is deterministic.
__global__ __launch_bounds__(256) void k_render_reduce(RenderWs W, float* __restrict__ rgb) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
  const int j = threadIdx.x & 3;
  const bool live = i < W.n;  // every lane takes part in the shuffles below
  const int c = live ? W.counts[i] : 0;
  const bool bad = live && (c > W.cpr || *W.skip);
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  if (live && !bad) {
    const float4* v = W.rec_f + i * (int64_t)W.cpr;
    for (int k = j; k < c; k += 4) {
      const float4 q = v[k];
      c0 = c0 + q.x;
      c1 = c1 + q.y;
      c2 = c2 + q.z;
    }
  }
#pragma unroll
  for (int m = 1; m <= 2; m <<= 1) {
    c0 = c0 + __shfl_xor(c0, m);
    c1 = c1 + __shfl_xor(c1, m);
    c2 = c2 + __shfl_xor(c2, m);
  }
  if (live && j == 0) {
    const float nan = __builtin_nanf("");
    rgb[3 * i] = bad ? nan : c0;
    rgb[3 * i + 1] = bad ? nan : c1;
    rgb[3 * i + 2] = bad ? nan : c2;
  }
}

// Eight trilinear corner weights (corner order x0y0z0, x1y0z0, ... as in tri_combine) as half2
// broadcasts, 0 for empty corners; the slots of empty corners are redirected to slot 0 so every
// load stays in bounds (the scene always holds at least one brick).
__device__ __forceinline__ void corner_weights_h2(Corners& c, h2 (&w2)[8]) {
  const float wx[2] = {1.0f - c.wx, c.wx}, wy[2] = {1.0f - c.wy, c.wy}, wz[2] = {1.0f - c.wz, c.wz};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float w = c.slot[q] >= 0 ? wx[q & 1] * wy[(q >> 1) & 1] * wz[q >> 2] : 0.0f;
    const _Float16 hw = (_Float16)w;
    w2[q] = h2{hw, hw};
    c.slot[q] = c.slot[q] >= 0 ? c.slot[q] : 0;
  }
}

// 8 consecutive FP16 features interpolated as sum_q w_q v_q with packed half FMAs
// (cp[q]: the corner rows, already offset to this lane's features; off: further bytes, a constant)
template <class M>
__device__ __forceinline__ typename M::vec interp8_h2(const uint8_t* const (&cp)[8], const h2 (&w2)[8], int off) {
  h2 acc[4];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint4 raw = *reinterpret_cast<const uint4*>(cp[q] + off);
    const h2* v = reinterpret_cast<const h2*>(&raw);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = q == 0 ? w2[0] * v[j] : __builtin_elementwise_fma(w2[q], v[j], acc[j]);
  }
  float f[8];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    f[2 * j] = (float)acc[j][0];
    f[2 * j + 1] = (float)acc[j][1];
  }
  return M::cvt8(f);
}

template <>
__device__ __forceinline__ f16x8 interp8_h2<Mfma<SNISR_F16>>(const uint8_t* const (&cp)[8], const h2 (&w2)[8],
                                                             int off) {
  h2 acc[4];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const uint4 raw = *reinterpret_cast<const uint4*>(cp[q] + off);
    const h2* v = reinterpret_cast<const h2*>(&raw);
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = q == 0 ? w2[0] * v[j] : __builtin_elementwise_fma(w2[q], v[j], acc[j]);
  }
  f16x8 r;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    r[2 * j] = acc[j][0];
    r[2 * j + 1] = acc[j][1];
  }
  return r;
}


// ------------------------------------------------------------------------- MFMA MLP
// Persistent: one 512-thread workgroup per CU; workgroup g owns an equal contiguous span of the
// 64-record chunks of the tile-sorted order.  For every tile inside its span it stages that tile's
// packed weights (~86 KB for width 128 x 4) in LDS once, then its 8 wavefronts process the span's
// chunks of that tile, each MFMA reading its A fragment from LDS (ds_read_b128, conflict-free).
// Per layer the accumulators of one 32-row out-block live at a time (32 VGPRs); the layer's input
// and output activations stay in registers in the B-operand layout.
template <int F, int WD, int L, bool HALF, int DT>
__global__ __launch_bounds__(512) void k_mlp_mfma(snisr_scene S, snisr_rays R, RenderWs W) {
  using M = Mfma<DT>;
  using vec = typename M::vec;
  constexpr int IN = F + 15;
  constexpr int NKC0 = nkc_of(IN);
  constexpr int NFC = F / 16;
  constexpr int NOB = WD / 32;
  constexpr int NKH = WD / 16;
  constexpr int64_t TB = tile_bytes(IN, WD, L);
  static_assert(F % 16 == 0 && WD % 32 == 0 && NKC0 == NFC + 1, "shape");
  extern __shared__ __align__(16) uint8_t lds_w[];

  const int lane = threadIdx.x & 63, h = lane >> 5, s = lane & 31;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int T = W.n_tiles;
  const int64_t nchunks = (*W.skip ? 0 : W.tile_start[T]) >> 6;
  const int64_t per = (nchunks + gridDim.x - 1) / gridDim.x;
  int64_t cb0 = (int64_t)blockIdx.x * per;
  const int64_t cend = cb0 + per < nchunks ? cb0 + per : nchunks;
  const uint8_t* __restrict__ gw = reinterpret_cast<const uint8_t*>(S.mlp_packed);

  while (cb0 < cend) {
    // tile of chunk cb0 (buckets are sorted by tile id)
    int lo = 0, hi = T;  // tile_start[lo]/64 <= cb0 < tile_start[hi]/64
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((W.tile_start[mid] >> 6) <= cb0) lo = mid; else hi = mid;
    }
    const int t = lo;
    const int64_t seg_end = (W.tile_start[t + 1] >> 6) < cend ? (W.tile_start[t + 1] >> 6) : cend;
    const int32_t tile_base = S.tile_brick_base[t], tile_res = S.tile_res[t];
    // stage the tile's weights
    __syncthreads();
    {
      const uint4* src = reinterpret_cast<const uint4*>(gw + (int64_t)t * TB);
      uint4* dst = reinterpret_cast<uint4*>(lds_w);
      for (int k = threadIdx.x; k < (int)(TB / 16); k += blockDim.x) dst[k] = src[k];
    }
    __syncthreads();

    // records of this wave's first chunk; every later chunk's perm entries and records are
    // loaded one chunk ahead (perm at the top of the previous iteration, the records after its
    // gather), so that only the brick-table and feature loads are exposed
    // lane (h, s) owns record s of column block h, i.e. perm entry 64 c + lane of chunk c: it loads
    // that record, resolves its corners and writes its colour; the corners of both blocks reach
    // every lane through v_permlane32_swap.  The next chunk's perm entry and record are loaded one
    // chunk ahead (perm at the top of the previous iteration, the record after its gather), so
    // that only the brick-table and feature loads are exposed.
    int32_t ko = -1;
    float4 ro = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (cb0 + wv < seg_end) {
      ko = W.perm[(cb0 + wv) * 64 + lane];
      if (ko >= 0) ro = W.rec_f[ko];
    }
    for (int64_t c = cb0 + wv; c < seg_end; c += nwv) {
      const int32_t nko = c + nwv < seg_end ? W.perm[(c + nwv) * 64 + lane] : -1;
      // ---- gather: lane (h, s) builds its half of the input column of sample s of both blocks
      int32_t ray[2], valid[2];
      swap_halves(ko >= 0 ? ko / W.cpr : 0, ray[0], ray[1]);
      swap_halves(ko >= 0 ? 1 : 0, valid[0], valid[1]);
      vec xin[2][NKC0];
      if constexpr (HALF) {
        // branch-free for both column blocks (padding records read slot 0 with zero weights), so the
        // feature loads of the two blocks are in flight together
        Corners co;
        find_corners(S.brick_map + tile_base, tile_res, v3(ro.x, ro.y, ro.z), co);
        h2 wo[8];
        corner_weights_h2(co, wo);
        const _Float16 keep = ko >= 0 ? (_Float16)1.0f : (_Float16)0.0f;
        Corners cn[2];
        h2 w2[2][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          swap_halves(co.slot[q], cn[0].slot[q], cn[1].slot[q]);
          const h2 wq = wo[q] * h2{keep, keep};
          uint32_t b0, b1;
          swap_halves(__builtin_bit_cast(uint32_t, wq), b0, b1);
          w2[0][q] = __builtin_bit_cast(h2, b0);
          w2[1][q] = __builtin_bit_cast(h2, b1);
        }
        // one 64-bit row address per corner (this lane's 8 features of k-chunk 0); the k-chunks
        // follow at constant offsets of 32 bytes
        // (slots are non-negative here: corner_weights_h2 redirected the empty ones to slot 0)
        const uint8_t* cp[2][8];
        const uint8_t* base_h = reinterpret_cast<const uint8_t*>(S.feats) + 16 * h;
#pragma unroll
        for (int cbk = 0; cbk < 2; ++cbk)
#pragma unroll
          for (int q = 0; q < 8; ++q) cp[cbk][q] = base_h + (uint64_t)(uint32_t)cn[cbk].slot[q] * (uint64_t)(2 * F);
#pragma unroll
        for (int kc = 0; kc < NFC; ++kc)
#pragma unroll
          for (int cbk = 0; cbk < 2; ++cbk) xin[cbk][kc] = interp8_h2<M>(cp[cbk], w2[cbk], 32 * kc);
      } else {
        // FP32 features: both records in every lane, each block's corners resolved in turn
        float4 rr[2];
        swap_halves(ro.x, rr[0].x, rr[1].x);
        swap_halves(ro.y, rr[0].y, rr[1].y);
        swap_halves(ro.z, rr[0].z, rr[1].z);
#pragma unroll
        for (int cbk = 0; cbk < 2; ++cbk) {
          float xv[NFC][8];
#pragma unroll
          for (int kc = 0; kc < NFC; ++kc)
#pragma unroll
            for (int e = 0; e < 8; ++e) xv[kc][e] = 0.0f;
          if (valid[cbk]) {
            Corners cn;
            find_corners(S.brick_map + tile_base, tile_res, v3(rr[cbk].x, rr[cbk].y, rr[cbk].z), cn);
            uint32_t cmask[8];  // ahead of the loads: the masking statements bound scheduling regions
#pragma unroll
            for (int q = 0; q < 8; ++q) cmask[q] = corner_mask(cn.slot[q] >= 0);
#pragma unroll
            for (int kc = 0; kc < NFC; ++kc) {
              float v[8][8];
#pragma unroll
              for (int q = 0; q < 8; ++q)  // unconditional loads, empty corners masked to zero
                load8<false>(S.feats, (int64_t)(cn.slot[q] >= 0 ? cn.slot[q] : 0) * F + 16 * kc + 8 * h, cmask[q], v[q]);
#pragma unroll
              for (int e = 0; e < 8; ++e) {
                const float qv[8] = {v[0][e], v[1][e], v[2][e], v[3][e], v[4][e], v[5][e], v[6][e], v[7][e]};
                xv[kc][e] = tri_combine(qv, cn.wx, cn.wy, cn.wz);
              }
            }
          }
#pragma unroll
          for (int kc = 0; kc < NFC; ++kc) xin[cbk][kc] = M::cvt8(xv[kc]);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      // view encoding: precomputed per ray in the operand type by the count pass
#pragma unroll
      for (int cbk = 0; cbk < 2; ++cbk) {
        const vec e = load_frag<vec>(reinterpret_cast<const uint8_t*>(W.enc + (int64_t)ray[cbk] * 16 + 8 * h));
        float z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        xin[cbk][NFC] = valid[cbk] ? e : M::cvt8(z);
      }
      const float4 nro = nko >= 0 ? W.rec_f[nko] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      __builtin_amdgcn_sched_barrier(0);

      // ---- MLP, A fragments from LDS
      vec xa[2][NKH], xb[2][NKH];
      {
        const uint8_t* Lp = lds_w + layer_offset(IN, WD, L, 0);
        mma_layer<M, NKC0, NOB>(Lp, reinterpret_cast<const float*>(Lp + NOB * NKC0 * 1024), lane, h, xin[0], xin[1],
                                xa[0], xa[1]);
      }
#pragma unroll
      for (int l = 1; l < L - 1; ++l) {
        auto& src = (l & 1) ? xa : xb;
        auto& dst = (l & 1) ? xb : xa;
        const uint8_t* Lp = lds_w + layer_offset(IN, WD, L, l);
        mma_layer<M, NKH, NOB>(Lp, reinterpret_cast<const float*>(Lp + NOB * NKH * 1024), lane, h, src[0], src[1],
                               dst[0], dst[1]);
      }
      {
        auto& src = ((L - 2) & 1) ? xb : xa;  // output of layer L-2
        const uint8_t* Lp = lds_w + layer_offset(IN, WD, L, L - 1);
        const float* bias = reinterpret_cast<const float*>(Lp + NKH * 1024);
        f32x16 acc0 = bias_block(bias, 0, h), acc1 = acc0;
        mma_block<M, NKH>(Lp, lane, src[0], src[1], acc0, acc1);
        // rows 0..2 of the output block sit in elements 0..2 of the lanes with h == 0, for both blocks;
        // one swap per row hands block 1's to the lanes with h == 1, so every lane finishes its own record
        float o[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          float unused;
          swap_lanes(acc0[i], acc1[i], o[i], unused);
        }
        if (ko >= 0)
          W.rec_f[ko] = make_float4(ro.w * sigmoid_hw(o[0]), ro.w * sigmoid_hw(o[1]), ro.w * sigmoid_hw(o[2]), 0.0f);
      }
      ko = nko;
      ro = nro;
    }
    cb0 = seg_end;
  }
}

// ------------------------------------------------------------------------- ordered reduction
__global__ __launch_bounds__(256) void k_reduce(Workspace W, int64_t n, float* __restrict__ rgb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = W.offsets[i], e = W.offsets[i + 1];
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  if (e > W.cap) {
    c0 = c1 = c2 = __builtin_nanf("");
  } else {
    for (int64_t k = b; k < e; ++k) {
      const float4 v = W.rgbw[k];
      c0 = c0 + v.x;
      c1 = c1 + v.y;
      c2 = c2 + v.z;
    }
  }
  rgb[3 * i] = c0;
  rgb[3 * i + 1] = c1;
  rgb[3 * i + 2] = c2;
}


// ------------------------------------------------------------------------- training (FP32)
// Forward: same count/scan/emit (+ interval, T after, sigma per record) and tile sort as the render,
// then the per-record FP32 MLP on FP32 MFMA (k_train_mma).  Backward: a per-ray reverse pass gives
// dL/dc and dL/dsigma of every record; k_train_mma recomputes the MLP forward of every record
// (segment recomputation), writes each layer's input rows and output-delta rows, backpropagates to
// the features and scatter-adds the trilinear adjoint into the feature and density gradients; a
// reduction over each tile's rows gives dW = sum delta (x) input and db = sum delta (k_train_dw).
__host__ __device__ constexpr int in_pad_of(int in0, int wd, int l) { return l == 0 ? ((in0 + 3) & ~3) : wd; }
__host__ __device__ constexpr int out_pad_of(int wd, int L, int l) { return l == L - 1 ? 4 : wd; }
__host__ __device__ constexpr int64_t train_act_floats(int in0, int wd, int L) {
  int64_t s = 0;
  for (int l = 0; l < L; ++l) s += in_pad_of(in0, wd, l) + out_pad_of(wd, L, l);
  return s;
}
__host__ __device__ constexpr int64_t act_a_off(int in0, int wd, int L, int l) {
  int64_t s = 0;
  for (int m = 0; m < l; ++m) s += in_pad_of(in0, wd, m);
  return s;
}
__host__ __device__ constexpr int64_t act_d_off(int in0, int wd, int L, int l) {
  int64_t s = 0;
  for (int m = 0; m < L; ++m) s += in_pad_of(in0, wd, m);
  for (int m = 0; m < l; ++m) s += out_pad_of(wd, L, m);
  return s;
}
// float offset of layer l's W[out][in] block in the mlp_params layout
__host__ __device__ constexpr int64_t param_w_off(int in0, int wd, int L, int l) {
  int64_t s = 0;
  for (int m = 0; m < l; ++m) {
    const int i = m == 0 ? in0 : wd, o = m == L - 1 ? 3 : wd;
    s += (int64_t)o * i + o;
  }
  return s;
}

// dL/dsigma and dL/dc of every record of ray i, in reverse sample order
__global__ __launch_bounds__(256) void k_train_comp_bwd(Workspace W, int64_t n, const float* __restrict__ g) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = W.offsets[i], e = W.offsets[i + 1];
  if (e > W.cap) return;  // overflowed in the forward: no records
  const float g0 = g[3 * i], g1 = g[3 * i + 1], g2 = g[3 * i + 2];
  float suffix = 0.0f;
  for (int64_t k = e - 1; k >= b; --k) {
    const float4 c = W.crec[k];
    const float4 rf = W.rec_f[k];
    const float4 rb = W.rec_b[k];
    const float cg = c.x * g0 + c.y * g1 + c.z * g2;
    float ds = rb.x * (rb.y * cg - suffix);
    suffix = suffix + rf.w * cg;
    if (!(rb.z > 0.0f)) ds = 0.0f;
    W.dsig[k] = ds;
    W.rgbw[k] = make_float4(rf.w * g0, rf.w * g1, rf.w * g2, 0.0f);
  }
}

template <int IN, int WD, int L>
// two waves per SIMD at least: the 256-register cap keeps the accumulators in VGPRs (AGPR
// accumulators were copied in and out for every pair and row tile)
__global__ __launch_bounds__(256, 2) void k_train_dw(Workspace W, int64_t mlp_stride, float* __restrict__ dmlp) {
  static_assert(WD <= 128 && (IN + 3) / 4 * 4 <= 128, "rows fit the LDS pitch");
  __shared__ __align__(16) float Ds[kDwRows * kDwPitch];
  __shared__ __align__(16) float As[kDwRows * kDwPitch];
  const int T = W.n_tiles;
  const int tile = blockIdx.y / L, l = blockIdx.y % L;
  const int64_t NJ = W.cap + 64 * (int64_t)T;
  const float* A = W.act + NJ * act_a_off(IN, WD, L, l);
  const float* D = W.act + NJ * act_d_off(IN, WD, L, l);
  const int64_t j0 = W.tile_start[tile], cnt = W.tile_count[tile];
  const int64_t per = ((cnt + gridDim.x - 1) / gridDim.x + kDwRows - 1) / kDwRows * kDwRows;
  const int64_t jb = j0 + (int64_t)blockIdx.x * per;
  const int64_t je = jb + per < j0 + cnt ? jb + per : j0 + cnt;
  float* G = dmlp + (int64_t)tile * mlp_stride + param_w_off(IN, WD, L, l);
  constexpr int IP0 = in_pad_of(IN, WD, 0);
  if (l == 0) dw_layer<IP0, WD, IN, WD, true>(A, D, jb, je, G, Ds, As);
  else if (l == L - 1) dw_layer<WD, 4, WD, 3, true>(A, D, jb, je, G, Ds, As);
  else dw_layer<WD, WD, WD, WD, true>(A, D, jb, je, G, Ds, As);
}

// The FP32 MFMA training MLP helpers (fragment layout, packing, product chains) are in train_mma.hpp.

template <int F, int WD, int L, bool HALF, bool BWD>
__global__ __launch_bounds__(256, 2) void k_train_mma(snisr_scene S, snisr_rays R, Workspace W, float* __restrict__ dsigma,
                                                   float* __restrict__ dfeats) {
  constexpr int IN = F + 15, NS0 = tr_ns0(IN), NB = WD / 32, NSH = WD / 2;
  static_assert(F % 8 == 0 && F <= 32, "feature rows must fill whole 8-row groups of the first block");
  static_assert(WD % 32 == 0 && WD <= 128 && L >= 2 && L <= 8, "MLP shape (one 64-bit ReLU mask per layer)");
  constexpr int64_t STRIDE = tr_stride(IN, WD, L);
  constexpr int kTrLdsPerWave = 32 * (F + 1) + 2 * 32 * 8 + 32;
  __shared__ float tr_lds[BWD ? 4 * kTrLdsPerWave : 1];
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  const int T = W.n_tiles;
  const int64_t NJ = W.cap + 64 * (int64_t)T;
  const int64_t j = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32 + col;
  int32_t k = -1;
  if (j < W.tile_start[T]) k = W.perm[j];
  const bool valid = k >= 0;
  const uint64_t vm = __ballot(valid);
  if (vm == 0) return;  // a chunk of padding only (no barriers below)
  float4 rf = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  int2 ti = make_int2(0, 0);
  if (valid) {
    rf = W.rec_f[k];
    ti = W.rec_ti[k];
  }
  const int tile = __shfl(ti.x, (int)__builtin_ctzll(vm));
  Corners cn;
  if (valid) {
    find_corners(S.brick_map + S.tile_brick_base[tile], S.tile_res[tile], v3(rf.x, rf.y, rf.z), cn);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) cn.slot[q] = -1;
    cn.wx = cn.wy = cn.wz = 0.0f;
  }
  uint32_t cmask[8];  // computed ahead of the loads: the masking statements bound scheduling regions
#pragma unroll
  for (int q = 0; q < 8; ++q) cmask[q] = corner_mask(cn.slot[q] >= 0);
  // ---- layer-0 inputs: rows tr_row(s, h)
  float enc[15];
  {
    V3 d = v3(1.0f, 0.0f, 0.0f);
    if (valid) d = normalize_dir(v3(R.dirs[3 * (int64_t)ti.y], R.dirs[3 * (int64_t)ti.y + 1], R.dirs[3 * (int64_t)ti.y + 2]));
    view_encoding(d, enc);
  }
  float x0[NS0];
#pragma unroll
  for (int g = 0; g < NS0 / 4; ++g) {
    const int base = 32 * (g / 4) + 8 * (g % 4);  // first row of the 8-row group (lane half 0)
    if (base < F) {
      float v[8][4];
#pragma unroll
      // unconditional loads (empty corners read slot 0, which exists whenever a lane of the
      // chunk is valid, and are zeroed after), so the eight loads are in flight together
      for (int q = 0; q < 8; ++q) {
        const int64_t off = (int64_t)(cn.slot[q] >= 0 ? cn.slot[q] : 0) * F + base + 4 * h;
        const uint32_t m = cmask[q];
        if (HALF) {
          uint2 raw = *reinterpret_cast<const uint2*>(reinterpret_cast<const _Float16*>(S.feats) + off);
          raw.x &= m;
          raw.y &= m;
          const _Float16* hv = reinterpret_cast<const _Float16*>(&raw);
#pragma unroll
          for (int t = 0; t < 4; ++t) v[q][t] = (float)hv[t];
        } else {
          FeatLoad<false>::load(S.feats, off, m, v[q]);
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float qv[8] = {v[0][t], v[1][t], v[2][t], v[3][t], v[4][t], v[5][t], v[6][t], v[7][t]};
        x0[4 * g + t] = tri_combine(qv, cn.wx, cn.wy, cn.wz);
      }
    } else {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int e0 = base - F + t, e1 = e0 + 4;  // view-encoding index for lane half 0 / 1
        const float v0 = e0 < 15 ? enc[e0 < 15 ? e0 : 0] : 0.0f;
        const float v1 = e1 < 15 ? enc[e1 < 15 ? e1 : 0] : 0.0f;
        x0[4 * g + t] = h ? v1 : v0;
      }
    }
  }
  const float* P = W.frag + (int64_t)tile * STRIDE;
  float* act = W.act;
  if (BWD && valid) {
    // rows in the lane layout of the weight-gradient reduction: group g of this wavefront's 32
    // records is one contiguous 1 KB store
    float* A0 = act + NJ * act_a_off(IN, WD, L, 0) + (j - col) * in_pad_of(IN, WD, 0) + 4 * lane;
#pragma unroll
    for (int g = 0; g < NS0 / 4; ++g) store4(A0 + 256 * g, x0 + 4 * g);
  }
  // ---- forward: hidden layers (ReLU), masks of every hidden input for the backward
  float a[NB * 16];
  uint64_t mask[L];
  mma_blocks<NS0, NB>(
      P + tr_fwd_off(IN, WD, L, 0), x0, lane, [&](int ob) { return bias_block(P + tr_bias_off(IN, WD, L, 0), ob, h); },
      [&](int ob, const f32x16& acc) {
#pragma unroll
        for (int r = 0; r < 16; ++r) a[ob * 16 + r] = acc[r] > 0.0f ? acc[r] : 0.0f;
      });
#pragma unroll
  for (int l = 1; l < L; ++l) {
    mask[l] = 0;
#pragma unroll
    for (int r = 0; r < NB * 16; ++r) mask[l] |= (uint64_t)(a[r] > 0.0f) << (r & 63);
    if (BWD && valid) {
      float* Al = act + NJ * act_a_off(IN, WD, L, l) + (j - col) * WD + 4 * lane;
#pragma unroll
      for (int g = 0; g < NB * 4; ++g) store4(Al + 256 * g, a + 4 * g);
    }
    if (l == L - 1) break;
    float b[NB * 16];
    mma_blocks<NSH, NB>(
        P + tr_fwd_off(IN, WD, L, l), a, lane, [&](int ob) { return bias_block(P + tr_bias_off(IN, WD, L, l), ob, h); },
        [&](int ob, const f32x16& acc) {
#pragma unroll
          for (int r = 0; r < 16; ++r) b[ob * 16 + r] = acc[r] > 0.0f ? acc[r] : 0.0f;
        });
#pragma unroll
    for (int r = 0; r < NB * 16; ++r) a[r] = b[r];
  }
  // ---- output layer: rows 0..2 (lane half 0, registers 0..2)
  f32x16 z = bias_block(P + tr_bias_off(IN, WD, L, L - 1), 0, h);
  z = mma_f32_chain<NSH>(P + tr_fwd_off(IN, WD, L, L - 1), a, z, lane);
  if (!BWD) {
    if (valid && h == 0) {
      const float c0 = sigmoid_cw(z[0]), c1 = sigmoid_cw(z[1]), c2 = sigmoid_cw(z[2]);
      W.crec[k] = make_float4(c0, c1, c2, 0.0f);
      W.rgbw[k] = make_float4(rf.w * c0, rf.w * c1, rf.w * c2, 0.0f);
    }
    return;
  }
  // ---- output delta from dL/dc (written into rgbw by the compositing adjoint)
  float dout[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (valid && h == 0) {
    const float4 dc = W.rgbw[k];
    const float c0 = sigmoid_cw(z[0]), c1 = sigmoid_cw(z[1]), c2 = sigmoid_cw(z[2]);
    dout[0] = dc.x * c0 * (1.0f - c0);
    dout[1] = dc.y * c1 * (1.0f - c1);
    dout[2] = dc.z * c2 * (1.0f - c2);
    store4(act + NJ * act_d_off(IN, WD, L, L - 1) + j * 4, dout);
  }
  // ---- backpropagate: d_{l-1} = mask_l (W_l^T d_l)
  float d[NB * 16];
  mma_blocks<4, NB>(
      P + tr_bwd_off(IN, WD, L, L - 1), dout, lane, [&](int) { return f32x16{}; },
      [&](int ib, const f32x16& acc) {
#pragma unroll
        for (int r = 0; r < 16; ++r) d[ib * 16 + r] = ((mask[L - 1] >> ((ib * 16 + r) & 63)) & 1) ? acc[r] : 0.0f;
      });
#pragma unroll
  for (int l = L - 2; l >= 0; --l) {
    if (valid) {
      float* Dl = act + NJ * act_d_off(IN, WD, L, l) + (j - col) * WD + 4 * lane;
#pragma unroll
      for (int g = 0; g < NB * 4; ++g) store4(Dl + 256 * g, d + 4 * g);
    }
    if (l == 0) break;
    float e[NB * 16];
    mma_blocks<NSH, NB>(
        P + tr_bwd_off(IN, WD, L, l), d, lane, [&](int) { return f32x16{}; },
        [&](int ib, const f32x16& acc) {
#pragma unroll
          for (int r = 0; r < 16; ++r) e[ib * 16 + r] = ((mask[l] >> ((ib * 16 + r) & 63)) & 1) ? acc[r] : 0.0f;
        });
#pragma unroll
    for (int r = 0; r < NB * 16; ++r) d[r] = e[r];
  }
  // ---- dL/d(features): rows 0..F-1 of W_0^T d_0
  f32x16 df = {};
  df = mma_f32_chain<NSH>(P + tr_bwd_off(IN, WD, L, 0), d, df, lane);
  // ---- trilinear adjoint.  The wave's 32 records go through LDS so that each atomic instruction
  // covers whole feature rows (64 / F records x F contiguous features) instead of 64 scattered
  // values; padding records have no corners.
  float* dfl = tr_lds + (threadIdx.x >> 6) * kTrLdsPerWave;  // [32][F + 1] feature gradients
  float* wl = dfl + 32 * (F + 1);                             // [32][8] corner weights
  int* sl = reinterpret_cast<int*>(wl + 32 * 8);              // [32][8] corner slots
  float* gl = wl + 2 * 32 * 8;                                // [32] dL/dsigma
#pragma unroll
  for (int g = 0; g < F / 8; ++g)
#pragma unroll
    for (int t = 0; t < 4; ++t) dfl[col * (F + 1) + 8 * g + 4 * h + t] = df[4 * g + t];
  if (h == 0) {
    const float wx[2] = {1.0f - cn.wx, cn.wx}, wy[2] = {1.0f - cn.wy, cn.wy}, wz[2] = {1.0f - cn.wz, cn.wz};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      wl[col * 8 + q] = wx[q & 1] * wy[(q >> 1) & 1] * wz[q >> 2];
      sl[col * 8 + q] = cn.slot[q];
    }
    gl[col] = valid ? W.dsig[k] : 0.0f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // Consecutive records of a tile bucket are mostly consecutive samples of one ray (0.2-voxel
  // steps), which share their cell: each lane group walks its segment of records in order and keeps
  // a running sum per corner, adding it with one atomic only when that corner's slot changes.  This
  // cuts the atomics, which leave L2 for device memory, by about the run length.
  constexpr int RPI = 64 / F, SEG = 32 / RPI;  // lane groups, records per group
  static_assert(RPI * SEG == 32, "the lane groups cover the 32 records");
  const int fr = lane % F, rr = lane / F;
  if (rr < RPI) {  // F = 24: lanes 48..63 form no group
    float run[8];
    int cur[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      run[q] = 0.0f;
      cur[q] = -1;
    }
#pragma unroll 2
    for (int j = 0; j < SEG; ++j) {
      const int rec = rr * SEG + j;
      const float g = dfl[rec * (F + 1) + fr];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int slot = sl[rec * 8 + q];
        const float v = wl[rec * 8 + q] * g;
        if (slot != cur[q]) {
          if (cur[q] >= 0) atomicAdd(&dfeats[(int64_t)cur[q] * F + fr], run[q]);
          cur[q] = slot;
          run[q] = v;
        } else {
          run[q] += v;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (cur[q] >= 0) atomicAdd(&dfeats[(int64_t)cur[q] * F + fr], run[q]);
  }
  {  // density: lane = 8 segments of 4 records x 8 corners, the same merging along the segment
    const int q = lane & 7, seg = lane >> 3;
    float run = 0.0f;
    int cur = -1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int rec = seg * 4 + j;
      const int slot = sl[rec * 8 + q];
      const float v = wl[rec * 8 + q] * gl[rec];
      if (slot != cur) {
        if (cur >= 0 && run != 0.0f) atomicAdd(&dsigma[cur], run);
        cur = slot;
        run = v;
      } else {
        run += v;
      }
    }
    if (cur >= 0 && run != 0.0f) atomicAdd(&dsigma[cur], run);
  }
}

__global__ void k_fill_zero(float* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

int packet_width(const snisr_rays& R);  // render_fp32.hip

// compute units of the current device (a per-device cache of a constant attribute)
// compute units of the current device (cached per device)
int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int v = cus[dev].load(std::memory_order_relaxed);
  if (v <= 0) {
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

// resident wavefronts per CU of the queue-fed march (its grid covers the device once), per device
static int march_queue_waves_per_cu() {
  static std::atomic<int> per[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 8;
  int v = per[dev].load(std::memory_order_relaxed);
  if (v <= 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, k_march_queue, 64, 0) != hipSuccess || v <= 0) v = 8;
    per[dev].store(v, std::memory_order_relaxed);
  }
  return v;
}

}  // namespace snisr

using namespace snisr;

extern "C" {

int64_t snisr_render_mfma_workspace_size(const snisr_scene* scene, int64_t n, int64_t records_per_ray,
                                         int64_t sample_capacity) {
  const snisr_tuning tn{0, 512, kSortGrid, (int32_t)(records_per_ray <= 4096 ? records_per_ray : -1), 1};
  return snisr_render_mfma_workspace_size_tuned(scene, n, &tn, sample_capacity);
}

int64_t snisr_render_mfma_workspace_size_tuned(const snisr_scene* scene, int64_t n, const snisr_tuning* tuning,
                                               int64_t sample_capacity) {
  if (!scene || !tuning || n < 0 || tuning->records_per_ray < 0 || sample_capacity < 0 || scene->mlp_in <= 0) {
    set_error("render_mfma_workspace_size: bad arguments");
    return -1;
  }
  const int64_t cpr = tuning->records_per_ray ? tuning->records_per_ray : kDefaultRecordsPerRay;
  const int64_t cap = sample_capacity ? sample_capacity : kDefaultSamplesPerRay * n;
  return render_layout(n, cpr, cap, scene->n_tiles, nullptr, nullptr, tuning->chunks > 1 ? tuning->chunks : 1);
}

int snisr_render_mfma(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                      const snisr_render_out* out, void* workspace, int64_t workspace_bytes, void* stream) {
  return snisr_render_mfma_tuned(scene, params, rays, out, nullptr, workspace, workspace_bytes, stream);
}

}  // extern "C"

namespace snisr {

// The second stream of a chunked render and its events, per host thread and device: a chunked
// render forks the stream it is given onto this one and joins it back before returning, so it can
// be captured into a HIP graph like an unchunked one.
struct ChunkStreams {
  hipStream_t aux = nullptr;
  std::vector<hipEvent_t> ev;
};

static int chunk_streams(int need_events, ChunkStreams** out) {
  thread_local ChunkStreams per_dev[64];
  int dev = 0;
  SNISR_HIP(hipGetDevice(&dev));
  SNISR_REQUIRE(dev >= 0 && dev < 64, SNISR_ERR_DEVICE, "render_mfma: device %d beyond 64", dev);
  ChunkStreams& cs = per_dev[dev];
  if (!cs.aux) SNISR_HIP(hipStreamCreateWithFlags(&cs.aux, hipStreamNonBlocking));
  while ((int)cs.ev.size() < need_events) {
    hipEvent_t e;
    SNISR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    cs.ev.push_back(e);
  }
  *out = &cs;
  return SNISR_OK;
}

// chunk c of k over n rays: whole 8-row packet bands when the march runs 8x8 packets (band of
// 8 * image_width rays), else whole 64-ray wavefronts
static void chunk_range(int64_t n, int64_t unit, int c, int k, int64_t* r0, int64_t* r1) {
  const int64_t units = (n + unit - 1) / unit;
  const int64_t u0 = units * c / k, u1 = units * (c + 1) / k;
  *r0 = u0 * unit < n ? u0 * unit : n;
  *r1 = u1 * unit < n ? u1 * unit : n;
}

}  // namespace snisr

extern "C" {

int snisr_render_mfma_tuned(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                            const snisr_render_out* out, const snisr_tuning* tuning, void* workspace,
                            int64_t workspace_bytes, void* stream) {
  const snisr_tuning tn = tuning ? *tuning : snisr_tuning{0, 512, kSortGrid, 0, 1};
  SNISR_REQUIRE(tn.records_per_ray >= 0 && tn.records_per_ray <= 4096, SNISR_ERR_ARG,
                "render_mfma: records_per_ray out of range");
  SNISR_REQUIRE(tn.ray_order >= 0 && tn.ray_order <= 3, SNISR_ERR_ARG, "render_mfma: ray_order must be 0..3");
  SNISR_REQUIRE(tn.mlp_threads == 256 || tn.mlp_threads == 512, SNISR_ERR_ARG,
                "render_mfma: mlp_threads must be 256 or 512");
  SNISR_REQUIRE(tn.sort_grid >= 1 && tn.sort_grid <= 65536, SNISR_ERR_ARG, "render_mfma: sort_grid out of range");
  SNISR_REQUIRE(tn.chunks >= 0 && tn.chunks <= kMaxChunks, SNISR_ERR_ARG, "render_mfma: chunks must be 0..%d",
                kMaxChunks);
  SNISR_REQUIRE(tn.chunks <= 1 || tn.ray_order == 0, SNISR_ERR_ARG,
                "render_mfma: chunks > 1 needs ray_order 0 (chunks are contiguous ray ranges)");
  SNISR_REQUIRE(out && out->rgb, SNISR_ERR_ARG, "render_mfma: out.rgb is required");
  if (int e = check_scene(scene, true)) return e;
  if (int e = check_params(params)) return e;
  if (int e = check_rays(rays, params)) return e;
  SNISR_REQUIRE(scene->mlp_packed, SNISR_ERR_ARG, "render_mfma: scene.mlp_packed is NULL (snisr_pack_mlp)");
  SNISR_REQUIRE(aligned(scene->mlp_packed, 256), SNISR_ERR_ALIGN, "render_mfma: mlp_packed must be 256-byte aligned");
  SNISR_REQUIRE(workspace && aligned(workspace, 256), SNISR_ERR_ARG, "render_mfma: workspace must be 256-byte aligned");
  SNISR_REQUIRE(!out->bricks || out->brick_cap > 0, SNISR_ERR_ARG, "render_mfma: bricks needs brick_cap > 0");
  const int64_t n = rays->n;
  if (n == 0) return SNISR_OK;
  SNISR_REQUIRE(scene->n_tiles <= kMaxTiles, SNISR_ERR_SHAPE, "render_mfma: at most %d tiles", kMaxTiles);
  const int pw = packet_width(*rays);
  // ray chunks: contiguous ranges of whole packet bands (or wavefronts), at most one per unit
  const int64_t unit = pw > 0 ? 8 * (int64_t)pw : 64;
  const int64_t units = (n + unit - 1) / unit;
  const int k = (int)(tn.chunks > 1 ? (tn.chunks < units ? tn.chunks : units) : 1);
  // records per ray from the schedule (default 64); every remaining byte goes to the sorted arrays,
  // split evenly between the chunks
  int64_t cpr = tn.records_per_ray > 0 ? tn.records_per_ray : kDefaultRecordsPerRay;
  if (n * cpr >= (int64_t(1) << 31)) cpr = ((int64_t(1) << 31) - 1) / n;  // slot ids are 32-bit
  SNISR_REQUIRE(cpr > 0, SNISR_ERR_CAPACITY, "render_mfma: batch too large for 32-bit record slots");
  const int64_t fixed = render_layout(n, cpr, 0, scene->n_tiles, nullptr, nullptr, k);
  SNISR_REQUIRE(workspace_bytes > fixed, SNISR_ERR_CAPACITY,
                "render_mfma: workspace too small for %lld records per ray", (long long)cpr);
  int64_t samples = (workspace_bytes - fixed) / (kSortedBytes * k);
  while (samples > 0 && render_layout(n, cpr, samples, scene->n_tiles, nullptr, nullptr, k) > workspace_bytes)
    --samples;
  SNISR_REQUIRE(samples > 0, SNISR_ERR_CAPACITY, "render_mfma: workspace too small for the sorted samples");
  RenderWs W;
  render_layout(n, cpr, samples, scene->n_tiles, (uint8_t*)workspace, &W, k);

  const int F = scene->feat_dim, WD = scene->mlp_width, L = scene->mlp_layers;
  const bool half = scene->feat_dtype == SNISR_F16;
  const int dt = scene->mlp_packed_dtype;
  hipStream_t st = (hipStream_t)stream;
  const dim3 sort_grid((unsigned)tn.sort_grid), sort_block(kSortBlock);

  // the shaded half of the render on one chunk's view: tile sort, MFMA MLP, per-ray reduction
  auto launch_march = [&](int64_t r0, int64_t r1, const int32_t* order, int c) -> int {
    const int64_t units = (r1 - r0 + 63) / 64;
    if (pw > 0 && !order) {
      hipLaunchKernelGGL(k_march_render, dim3((unsigned)units), dim3(64), 0, st, *scene, *params, *rays, *out,
                         W, pw, dt, order, r0, r1);
      SNISR_LAUNCHED();
      return SNISR_OK;
    }
    hipLaunchKernelGGL(k_ray_enc, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), 0, st, *rays, W, dt, r0, r1);
    SNISR_LAUNCHED();
    RenderWs Wq = W;
    Wq.cursor = W.cursor + c;
    const int64_t resident = (int64_t)device_cus() * march_queue_waves_per_cu();
    hipLaunchKernelGGL(k_march_queue, dim3((unsigned)(units < resident ? units : resident)), dim3(64), 0, st,
                       *scene, *params, *rays, *out, Wq, order, r0, r1);
    SNISR_LAUNCHED();
    return SNISR_OK;
  };
  auto launch_mlp = [&](const RenderWs& V, hipStream_t s) -> int {
    // one workgroup per CU (its tile weights fill the LDS), and for large batches several per CU in
    // turn, each over a shorter span of chunks, so that the spans' unequal costs even out: about one
    // workgroup per 1024 sorted chunks per CU, at most 8 (C2: 1; C5: 0.2 ms and C4: 0.4 ms faster
    // with 4-8 than with 1, C2 0.5-3% slower with more than 1)
    const int64_t cus = device_cus();
    int64_t per_cu = V.scap / (64 * 1024 * cus);
    per_cu = per_cu < 1 ? 1 : (per_cu > 8 ? 8 : per_cu);
    const dim3 grid((unsigned)(cus * per_cu)), block((unsigned)tn.mlp_threads);
    bool launched = false;
#define SNISR_MFMA_CASE(F_, W_, L_, H_, D_)                                                        \
    if (!launched && F == F_ && WD == W_ && L == L_ && half == H_ && dt == D_) {                    \
      auto kern = k_mlp_mfma<F_, W_, L_, H_, D_>;                                                   \
      const size_t lds = (size_t)tile_bytes(F_ + 15, W_, L_);                                       \
      static std::atomic<bool> lds_set{false};                                                      \
      if (!lds_set.load()) {                                                                        \
        SNISR_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        lds_set.store(true);                                                                        \
      }                                                                                             \
      hipLaunchKernelGGL(kern, grid, block, lds, s, *scene, *rays, V);                              \
      launched = true;                                                                              \
    }
    SNISR_MFMA_CASE(32, 128, 4, true, SNISR_F16)
    SNISR_MFMA_CASE(32, 128, 4, true, SNISR_BF16)
    SNISR_MFMA_CASE(16, 64, 3, false, SNISR_F16)
    SNISR_MFMA_CASE(16, 64, 3, false, SNISR_BF16)
    SNISR_MFMA_CASE(32, 128, 4, false, SNISR_F16)
#undef SNISR_MFMA_CASE
    if (!launched) {
      set_error("render_mfma: unsupported shape F=%d width=%d layers=%d feat_dtype=%d packed=%d", F, WD, L,
                scene->feat_dtype, dt);
      return SNISR_ERR_SHAPE;
    }
    SNISR_LAUNCHED();
    return SNISR_OK;
  };
  auto launch_shade = [&](const RenderWs& V, float* rgb, hipStream_t s, bool stages) -> int {
    hipLaunchKernelGGL(k_rtile_hist, sort_grid, sort_block, 0, s, V);
    SNISR_LAUNCHED();
    hipLaunchKernelGGL(k_rtile_scan, dim3(1), dim3(kScanBlock), 0, s, V);
    SNISR_LAUNCHED();
    if (stages) SNISR_HIP(stage_event(out, 3, s));
    hipLaunchKernelGGL(k_rtile_scatter, sort_grid, sort_block, kScatterLds, s, V);
    SNISR_LAUNCHED();
    if (stages) SNISR_HIP(stage_event(out, 4, s));
    if (int e = launch_mlp(V, s)) return e;
    if (stages) SNISR_HIP(stage_event(out, 5, s));
    hipLaunchKernelGGL(k_render_reduce, dim3((unsigned)((4 * V.n + 255) / 256)), dim3(256), 0, s, V, rgb);
    SNISR_LAUNCHED();
    return SNISR_OK;
  };

  SNISR_HIP(stage_event(out, 0, st));
  hipLaunchKernelGGL(k_render_begin, dim3(1), dim3(256), 0, st, W);
  SNISR_LAUNCHED();
  const int32_t* order = nullptr;
  if (tn.ray_order >= 1) {
    const KeyShape ks = key_shape(tn.ray_order, scene->n_tiles);
    W.n_buckets = ks.n_buckets;
    const int64_t kn = n > W.n_buckets ? n : W.n_buckets;
    hipLaunchKernelGGL(k_ray_key, dim3((unsigned)((kn + 255) / 256)), dim3(256), 0, st, *scene, *rays, W, ks);
    SNISR_LAUNCHED();
    hipLaunchKernelGGL(k_key_hist, sort_grid, sort_block, 0, st, W, n);
    SNISR_LAUNCHED();
    hipLaunchKernelGGL(k_key_scan, dim3(1), dim3(kScanBlock), 0, st, W);
    SNISR_LAUNCHED();
    hipLaunchKernelGGL(k_key_scatter, sort_grid, sort_block, 0, st, W, n);
    SNISR_LAUNCHED();
    order = W.order;
  }
  SNISR_HIP(stage_event(out, 1, st));
  if (k == 1) {
    if (int e = launch_march(0, n, order, 0)) return e;
    SNISR_HIP(stage_event(out, 2, st));
    if (int e = launch_shade(W, out->rgb, st, true)) return e;
  } else {
    // march chunk c on the given stream while the second stream sorts and shades chunk c - 1:
    // the march is bound by dependent brick-table loads and the MLP by matrix-core issue, so
    // the two share the CUs
    ChunkStreams* cs = nullptr;
    if (int e = chunk_streams(k + 1, &cs)) return e;
    for (int c = 0; c < k; ++c) {
      int64_t r0, r1;
      chunk_range(n, unit, c, k, &r0, &r1);
      if (int e = launch_march(r0, r1, nullptr, c)) return e;
      SNISR_HIP(hipEventRecord(cs->ev[c], st));
      SNISR_HIP(hipStreamWaitEvent(cs->aux, cs->ev[c], 0));
      if (int e = launch_shade(chunk_view(W, c, r0, r1 - r0), out->rgb + 3 * r0, cs->aux, false)) return e;
    }
    SNISR_HIP(hipEventRecord(cs->ev[k], cs->aux));
    SNISR_HIP(hipStreamWaitEvent(st, cs->ev[k], 0));
    // the stages overlap: 2..5 are all taken at the join
    for (int q = 2; q <= 5; ++q) SNISR_HIP(stage_event(out, q, st));
  }
  SNISR_HIP(stage_event(out, 6, st));
  return SNISR_OK;
}


// record splits per (tile, layer) of the weight-gradient reduction: about sixteen workgroups per CU
// (two are resident at a time), so that the layers' unequal costs even out over many short
// workgroups (C3: 4 per CU 3.63 ms, 8: 3.05, 16: 2.95, 32: 2.94, 64: 3.01)
static int dw_splits(int64_t tile_layers) {
  const int64_t s = 16 * (int64_t)device_cus() / (tile_layers > 0 ? tile_layers : 1);
  return (int)(s < 1 ? 1 : (s > 128 ? 128 : s));
}

static int64_t train_frag_floats(const snisr_scene* s) {
  return (int64_t)s->n_tiles * tr_stride(s->mlp_in, s->mlp_width, s->mlp_layers);
}

int64_t snisr_train_workspace_size(const snisr_scene* scene, int64_t n, int64_t sample_capacity) {
  if (!scene || n < 0 || sample_capacity < 0) {
    set_error("train_workspace_size: bad arguments");
    return -1;
  }
  const int64_t cap = sample_capacity ? sample_capacity : default_capacity(n);
  return layout(n, cap, scene->n_tiles, nullptr, nullptr,
                train_act_floats(scene->mlp_in, scene->mlp_width, scene->mlp_layers), train_frag_floats(scene));
}

static int train_capacity(const snisr_scene* scene, int64_t n, int64_t bytes, int64_t* cap_out) {
  const int64_t tf = train_act_floats(scene->mlp_in, scene->mlp_width, scene->mlp_layers);
  const int64_t ff = train_frag_floats(scene);
  const int64_t fixed = layout(n, 0, scene->n_tiles, nullptr, nullptr, tf, ff);
  SNISR_REQUIRE(bytes >= fixed, SNISR_ERR_CAPACITY, "train: workspace too small");
  const int64_t per = 44 + 36 + 4 * tf;  // render record arrays + training arrays + activation rows
  int64_t cap = (bytes - fixed - 8 * 256) / per;
  while (cap > 0 && layout(n, cap, scene->n_tiles, nullptr, nullptr, tf, ff) > bytes) --cap;
  SNISR_REQUIRE(cap > 0, SNISR_ERR_CAPACITY, "train: workspace too small");
  *cap_out = cap;
  return SNISR_OK;
}

#define SNISR_TRAIN_SHAPES(X) X(16, 32, 3, false) X(16, 64, 3, false) X(32, 128, 4, false) X(32, 128, 4, true)

// Training forward: the render of snisr_render_fp32 (same traversal and samples) that keeps the
// per-sample records in the workspace for snisr_train_backward.  Replaces the forward of the
// autograd chain of Tile.train_batch (training/scripts/tile/tile.py:631-820).
int snisr_train_forward(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                        const snisr_render_out* out, void* workspace, int64_t workspace_bytes, void* stream) {
  SNISR_REQUIRE(out && out->rgb, SNISR_ERR_ARG, "train_forward: out.rgb is required");
  if (int e = check_scene(scene, true)) return e;
  if (int e = check_params(params)) return e;
  if (int e = check_rays(rays, params)) return e;
  SNISR_REQUIRE(scene->mlp_params, SNISR_ERR_ARG, "train_forward: scene.mlp_params is NULL");
  SNISR_REQUIRE(workspace && aligned(workspace, 256), SNISR_ERR_ARG, "train_forward: workspace must be 256-byte aligned");
  SNISR_REQUIRE(scene->n_tiles <= kMaxTiles, SNISR_ERR_SHAPE, "train_forward: at most %d tiles", kMaxTiles);
  const int64_t n = rays->n;
  if (n == 0) return SNISR_OK;
  int64_t cap = 0;
  if (int e = train_capacity(scene, n, workspace_bytes, &cap)) return e;
  Workspace W;
  layout(n, cap, scene->n_tiles, (uint8_t*)workspace, &W,
         train_act_floats(scene->mlp_in, scene->mlp_width, scene->mlp_layers), train_frag_floats(scene));
  const int F = scene->feat_dim, WD = scene->mlp_width, L = scene->mlp_layers;
  const bool half = scene->feat_dtype == SNISR_F16;
  hipStream_t st = (hipStream_t)stream;
  const int pw = packet_width(*rays);
  const dim3 mgrid((unsigned)((n + 63) / 64));
  const int64_t nblocks = (n + kScanTile - 1) / kScanTile;
  bool ok = false;
#define SNISR_CHECK_SHAPE(F_, W_, L_, H_) ok = ok || (F == F_ && WD == W_ && L == L_ && half == H_);
  SNISR_TRAIN_SHAPES(SNISR_CHECK_SHAPE)
#undef SNISR_CHECK_SHAPE
  SNISR_REQUIRE(ok, SNISR_ERR_SHAPE, "train_forward: unsupported shape F=%d width=%d layers=%d", F, WD, L);
  SNISR_HIP(stage_event(out, 0, st));
  // no ray schedule: sorting arbitrary training rays by (first tile, octant) made the march slower
  const int32_t* order = nullptr;
  hipLaunchKernelGGL(k_march_pass<0>, mgrid, dim3(64), 0, st, *scene, *params, *rays, *out, W, pw, (int)SNISR_F16,
                     order);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nblocks), dim3(kScanBlock), 0, st, W.counts, n, W.offsets,
                     W.block_sums);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanBlock), 0, st, W.block_sums, nblocks, W.offsets, n, W.total,
                     W.overflow, W.tile_count, W.n_tiles);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nblocks), dim3(kScanBlock), 0, st, W.offsets, n, W.block_sums);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_march_pass<2>, mgrid, dim3(64), 0, st, *scene, *params, *rays, *out, W, pw, (int)SNISR_F16,
                     order);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_tile_hist, dim3(kSortGrid), dim3(kSortBlock), 0, st, W);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(kScanBlock), 0, st, W);
  SNISR_LAUNCHED();
  hipLaunchKernelGGL(k_tile_scatter, dim3(kSortGrid), dim3(kSortBlock), 0, st, W);
  SNISR_LAUNCHED();
  SNISR_HIP(stage_event(out, 1, st));
  {  // this step's weights as FP32 MFMA fragments (the backward reads the same copy)
    const int64_t tot = train_frag_floats(scene);
    hipLaunchKernelGGL(k_pack_train_frag, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, scene->mlp_params,
                       scene->mlp_stride, W.frag, scene->mlp_in, WD, L, scene->n_tiles, 3);
    SNISR_LAUNCHED();
  }
  const dim3 jgrid((unsigned)((cap + 64 * (int64_t)scene->n_tiles + 127) / 128));
#define SNISR_FWD(F_, W_, L_, H_)                                                                      \
  if (F == F_ && WD == W_ && L == L_ && half == H_)                                                      \
    hipLaunchKernelGGL((k_train_mma<F_, W_, L_, H_, false>), jgrid, dim3(256), 0, st, *scene, *rays, W, \
                       (float*)nullptr, (float*)nullptr);
  SNISR_TRAIN_SHAPES(SNISR_FWD)
#undef SNISR_FWD
  SNISR_LAUNCHED();
  SNISR_HIP(stage_event(out, 2, st));
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, W, n, out->rgb);
  SNISR_LAUNCHED();
  SNISR_HIP(stage_event(out, 3, st));
  return SNISR_OK;
}

// Training backward: gradients of L = sum_i <grad_rgb_i, rgb_i> for the rays of the preceding
// snisr_train_forward on the same workspace (which it reads).  dsigma f32[NB,512], dfeats
// f32[NB,512,F] and dmlp f32[T,P] (mlp_params layout) are overwritten.  Hand-written adjoint of
// the compositing, trilinear gather and MLP that torch autograd provides in the reference
// (training/scripts/tile/tile.py:780-785).  stage_events (optional, 4): before the compositing
// adjoint, after it, after the MLP/trilinear adjoint, after the weight-gradient reduction.
int snisr_train_backward(const snisr_scene* scene, const snisr_march_params* params, const snisr_rays* rays,
                         const float* grad_rgb, float* dsigma, float* dfeats, float* dmlp, void* workspace,
                         int64_t workspace_bytes, void* const* stage_events, int32_t n_stage_events, void* stream) {
  if (int e = check_scene(scene, true)) return e;
  if (int e = check_params(params)) return e;
  if (int e = check_rays(rays, params)) return e;
  SNISR_REQUIRE(grad_rgb && dsigma && dfeats && dmlp, SNISR_ERR_ARG, "train_backward: NULL gradient pointer");
  SNISR_REQUIRE(workspace && aligned(workspace, 256), SNISR_ERR_ARG, "train_backward: workspace");
  const int64_t n = rays->n;
  hipStream_t st = (hipStream_t)stream;
  snisr_render_out ev{};
  ev.stage_events = stage_events;
  ev.n_stage_events = n_stage_events;
  const int F = scene->feat_dim, WD = scene->mlp_width, L = scene->mlp_layers;
  const bool half = scene->feat_dtype == SNISR_F16;
  const int64_t nsig = (int64_t)scene->n_bricks * 512, nfeat = nsig * F, nmlp = (int64_t)scene->n_tiles * scene->mlp_stride;
  SNISR_HIP(stage_event(&ev, 0, st));
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nsig + 255) / 256)), dim3(256), 0, st, dsigma, nsig);
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nfeat + 255) / 256)), dim3(256), 0, st, dfeats, nfeat);
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nmlp + 255) / 256)), dim3(256), 0, st, dmlp, nmlp);
  SNISR_LAUNCHED();
  if (n == 0) return SNISR_OK;
  int64_t cap = 0;
  if (int e = train_capacity(scene, n, workspace_bytes, &cap)) return e;
  Workspace W;
  layout(n, cap, scene->n_tiles, (uint8_t*)workspace, &W,
         train_act_floats(scene->mlp_in, scene->mlp_width, scene->mlp_layers), train_frag_floats(scene));
  hipLaunchKernelGGL(k_train_comp_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, W, n, grad_rgb);
  SNISR_LAUNCHED();
  SNISR_HIP(stage_event(&ev, 1, st));
  const dim3 jgrid((unsigned)((cap + 64 * (int64_t)scene->n_tiles + 127) / 128));
  bool ok = false;
#define SNISR_BWD(F_, W_, L_, H_)                                                                            \
  if (F == F_ && WD == W_ && L == L_ && half == H_) {                                                        \
    hipLaunchKernelGGL((k_train_mma<F_, W_, L_, H_, true>), jgrid, dim3(256), 0, st, *scene, *rays, W, dsigma, \
                       dfeats);                                                                              \
    SNISR_LAUNCHED();                                                                                       \
    SNISR_HIP(stage_event(&ev, 2, st));                                                                     \
    const dim3 g((unsigned)dw_splits(scene->n_tiles * L_), (unsigned)(scene->n_tiles * L_));               \
    hipLaunchKernelGGL((k_train_dw<F_ + 15, W_, L_>), g, dim3(256), 0, st, W, scene->mlp_stride, dmlp);     \
    SNISR_LAUNCHED();                                                                                       \
    ok = true;                                                                                              \
  }
  SNISR_TRAIN_SHAPES(SNISR_BWD)
#undef SNISR_BWD
  SNISR_REQUIRE(ok, SNISR_ERR_SHAPE, "train_backward: unsupported shape F=%d width=%d layers=%d", F, WD, L);
  SNISR_HIP(stage_event(&ev, 3, st));
  return SNISR_OK;
}

// Total samples of the last snisr_render_mfma on this workspace and whether any ray overflowed
// the record capacity (synchronises the stream).
int snisr_render_mfma_stats(const void* workspace, int64_t* total, int32_t* overflow, int32_t* max_count,
                            void* stream) {
  SNISR_REQUIRE(workspace && total && overflow, SNISR_ERR_ARG, "render_mfma_stats: NULL pointer");
  int64_t hdr[2];
  SNISR_HIP(hipMemcpyAsync(hdr, workspace, 16, hipMemcpyDeviceToHost, (hipStream_t)stream));
  SNISR_HIP(hipStreamSynchronize((hipStream_t)stream));
  *total = hdr[0];
  *overflow = (int32_t)(hdr[1] & 0xffffffff);
  if (max_count) *max_count = (int32_t)((uint64_t)hdr[1] >> 32);
  return SNISR_OK;
}

}  // extern "C"
