// LBVH Barnes-Hut on CDNA4: Morton codes -> LDS radix sort -> Karras radix
// tree -> level-synchronous bottom-up moments -> escape links -> stackless or
// wavefront-cooperative traversal.
//
// Tree layout follows the reference's NodesArray indexing
// (src/solvers/barnes_hut.cpp:309-400): leaf k is the body at sorted position
// k, the internal node of split i is n + i, the root is n.
//
// Parity notes:
//  * Morton codes use the reference's FP64 normalisation and clamp
//    (src/solvers/bh_specialization.cpp:38-46): bit-identical.
//  * The sort is a stable LSD radix sort, i.e. the (code, index) order.
//  * The Karras split uses the reference's tie-break on sorted positions
//    (barnes_hut.cpp:340-342) and its search sequence: identical topology.
//  * Moments: M = m_R + m_L, COM = (c_R m_R + c_L m_L) / M in FP64 without
//    contraction (barnes_hut.cpp:145-150). Each node depends only on its two
//    children, so the bottom-up order cannot change the result: node masses
//    and centres of mass are bit-identical to the recursive unite().
//  * Traversal: the reference visits right children first (it pushes left,
//    then right, barnes_hut.cpp:241-256). Opening goes to the right child and
//    every accepted/skipped node continues at its escape link, which is the
//    next node of that same order, so the FP64 traversal accumulates in the
//    reference's order and is bit-identical to get_node_body (:202-261).
#include <utility>

#include "common.hpp"
#include "device.hpp"

namespace aster {

constexpr int kRsThreads = 256;
constexpr int kRsTile = 2048;  // keys per block (8 per thread)
// Depth bound of the radix tree: an internal node's split bit strictly
// increases with depth over 30 code bits + 32 tie-break bits, so depth <= 63.
constexpr int kMaxLevels = 128;
// Traversal groups (and the moment kernel's per-wavefront subtrees): subtrees
// of at most kGroup leaves whose parent has more.
constexpr int kGroup = 64;

// ------------------------------------------------------------ Morton -------

__device__ __forceinline__ uint32_t part1by2(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t cell_of(double p, double b, uint32_t* bad) {
#pragma clang fp contract(off)
  const double q = (p + b) / (2 * b);
  const double s = q * (1 << 10);
  // The reference converts without a lower clamp (undefined for q < 0): such
  // inputs are reported instead of guessed.
  if (!(q >= 0.0) || !(s < 4294967296.0)) *bad = 1u;
  const uint32_t c = (uint32_t)s;
  return c < 1023u ? c : 1023u;
}

// Also packs {x, y, z, m} of each body into one 32-byte record, so the leaf
// initialisation gathers one record through the sort permutation instead of
// four scattered planes.
__global__ __launch_bounds__(256) void k_morton(int n, const double* __restrict__ x, const double* __restrict__ m,
                                                const double* __restrict__ box, uint32_t* __restrict__ codes,
                                                uint32_t* __restrict__ idx, double4* __restrict__ xm,
                                                uint32_t* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t bad = 0;
  const double px = x[i], py = x[n + i], pz = x[2 * n + i];
  xm[i] = make_double4(px, py, pz, m[i]);
  const uint32_t ix = cell_of(px, box[0], &bad);
  const uint32_t iy = cell_of(py, box[1], &bad);
  const uint32_t iz = cell_of(pz, box[2], &bad);
  codes[i] = (part1by2(iz) << 2) | (part1by2(iy) << 1) | part1by2(ix);
  idx[i] = (uint32_t)i;
  if (bad) atomicOr(flags, (uint32_t)FLAG_OUT_OF_BOX);
}

// -------------------------------------------------------- radix sort ------

// LSD radix sort of (Morton code, body) pairs, 4 passes of 8 bits. A block
// handles a tile of kRsPer * THREADS keys. Per pass, k_rs_hist writes each
// tile's digit counts (hist[d * stride + b]); the scatter needs every digit's
// start, i.e. the keys of smaller digits in all tiles plus those of the same
// digit in earlier tiles. With at most kRsRawMaxBlocks tiles the scatter
// forms these starts itself from the raw counts (two launches per pass); with
// more, an exclusive scan of the count matrix runs between the two kernels.
constexpr int kRsPer = 8;
constexpr int kRsRawMaxBlocks = 128;

__host__ __device__ constexpr int rs_stride(int nblocks) { return (nblocks + 3) & ~3; }

template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_rs_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                     uint32_t* __restrict__ hist, int stride) {
  constexpr int TILE = kRsPer * THREADS, H = THREADS / 256;  // H copies of the counters
  static_assert(THREADS % 256 == 0, "one counter per thread and copy");
  __shared__ uint32_t cnt[H][256];
  const int t = threadIdx.x;
  const int base = blockIdx.x * TILE;
  // every load of the thread is issued before the first is used (clamped
  // addresses, no branch around a load), so the tile costs one memory
  // round trip rather than kRsPer
  uint32_t key[kRsPer];
#pragma unroll
  for (int r = 0; r < kRsPer; ++r) key[r] = keys[min(base + r * THREADS + t, n - 1)];
  cnt[t >> 8][t & 255] = 0;
  __syncthreads();
  uint32_t* mine = cnt[(t >> 6) % H];  // wavefront w counts into copy w mod H
#pragma unroll
  for (int r = 0; r < kRsPer; ++r)
    if (base + r * THREADS + t < n) atomicAdd(&mine[(key[r] >> shift) & 255u], 1u);
  __syncthreads();
  if (t < 256) {
    uint32_t s = 0;
#pragma unroll
    for (int h = 0; h < H; ++h) s += cnt[h][t];
    hist[(size_t)t * stride + blockIdx.x] = s;
  }
}

// Exclusive prefix sum of a u32 array, in place, in two passes over chunks of
// kScanChunk entries: chunk totals, then each chunk scanned and offset by the
// sum of the totals before it (every block re-reduces that short prefix, so
// no block waits on another).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanChunk = kScanThreads * kScanPer;

__device__ __forceinline__ uint32_t block_sum_256(uint32_t v, uint32_t* ws) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  return ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in, int count,
                                                              uint32_t* __restrict__ part) {
  __shared__ uint32_t ws[4];
  const int base = blockIdx.x * kScanChunk;
  // all kScanPer loads issued before the first is used (clamped addresses)
  uint32_t v[kScanPer];
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) v[r] = in[min(base + r * kScanThreads + (int)threadIdx.x, count - 1)];
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) s += base + r * kScanThreads + (int)threadIdx.x < count ? v[r] : 0u;
  s = block_sum_256(s, ws);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Thread t owns entries [base + 16 t, base + 16 t + 16) of its chunk. With
// `emit` (0/1 flags in data) it also compacts: emit[prefix] = position of
// every set flag, and emit[total] = count.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t* __restrict__ data, int count,
                                                             const uint32_t* __restrict__ part,
                                                             uint32_t* __restrict__ total,
                                                             int32_t* __restrict__ emit) {
  __shared__ uint32_t ws[4];
  __shared__ uint32_t wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  // the thread's entries are loaded first (clamped addresses), so their
  // latency overlaps the sum of the earlier chunks' totals
  const int base = blockIdx.x * kScanChunk + kScanPer * t;
  uint32_t v[kScanPer];
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) v[r] = data[min(base + r, count - 1)];
  uint32_t pre = 0;
  for (int b = t; b < (int)blockIdx.x; b += kScanThreads) pre += part[b];
  pre = block_sum_256(pre, ws);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    if (base + r >= count) v[r] = 0u;
    s += v[r];
  }
  // inclusive scan of the per-thread sums across the wave, then the waves
  uint32_t inc = s;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(inc, off, 64);
    if (lane >= off) inc += y;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t run = pre + inc - s;
  for (int q = 0; q < w; ++q) run += wsum[q];
#pragma unroll
  for (int r = 0; r < kScanPer; ++r) {
    if (base + r < count) {
      data[base + r] = run;
      if (emit && v[r]) emit[run] = base + r;
    }
    run += v[r];
  }
  if (total && base < count && base + kScanPer >= count) {
    *total = run;
    if (emit) emit[run] = count;
  }
}

// Stable scatter. Each wavefront owns 512 consecutive keys of the tile (8
// rounds of 64) and ranks its keys among equal digits with its own running
// counters in LDS: rank = the wavefront's earlier keys of the digit
// (wcnt[w][d]) + earlier lanes of the round (ballot match mask). LDS
// operations of one wavefront execute in order, so the rounds need no
// workgroup barrier. Then the tile's digit totals give each digit's start in
// the tile (lpre) and each wavefront's offset inside it; the tile is
// reordered by digit in LDS and written out with consecutive lanes on
// consecutive addresses within each digit's run.
// RAW: hist holds the raw tile counts and the block forms its digit starts:
// Q = THREADS / 256 threads per digit sum the digit's row (16-byte loads),
// then a scan over the 256 digit totals. Otherwise hist holds the starts
// (the exclusive scan of the count matrix).
template <int THREADS, bool RAW>
__global__ __launch_bounds__(THREADS) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                        int n, int shift, const uint32_t* __restrict__ hist,
                                                        int nblocks, int stride) {
  constexpr int W = THREADS / 64;       // wavefronts per tile
  constexpr int TILE = kRsPer * THREADS;
  constexpr int C = TILE / W;           // keys per wavefront
  constexpr int R = C / 64;             // rounds per wavefront
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t wcnt[W][256];
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t lpre[256];
  __shared__ uint32_t skey[TILE], sval[TILE];
  const int t = threadIdx.x;
  const int lane = t & 63, w = t >> 6;
  const int blk = (int)blockIdx.x;
  // This wavefront's keys and values, loaded first (clamped addresses, no
  // branch around a load) so that their latency overlaps the digit starts.
  const int base = blk * TILE + w * C;
  uint32_t key[R], val[R], pos[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = min(base + r * 64 + lane, n - 1);
    key[r] = kin[p];
    val[r] = vin[p];
  }
  if constexpr (RAW) {
    constexpr int Q = THREADS / 256;
    // 16-byte loads per thread for the most tiles the raw form takes, all
    // issued before the sums (entries past nblocks re-read entry 0 and count
    // nothing)
    constexpr int KMAX = (kRsRawMaxBlocks / 4 + Q - 1) / Q;
    const int d = t / Q, q = t - d * Q;
    const uint4* __restrict__ row = reinterpret_cast<const uint4*>(hist + (size_t)d * stride);
    uint4 rv[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int i = q + k * Q;
      rv[k] = row[4 * i < nblocks ? i : 0];
    }
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t e[4] = {rv[k].x, rv[k].y, rv[k].z, rv[k].w};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int b = 4 * (q + k * Q) + u;
        const uint32_t c = b < nblocks ? e[u] : 0u;
        tot += c;
        pre += b < blk ? c : 0u;
      }
    }
#pragma unroll
    for (int off = 1; off < Q; off <<= 1) {  // the Q threads of a digit are adjacent lanes
      pre += __shfl_xor(pre, off, 64);
      tot += __shfl_xor(tot, off, 64);
    }
    if (q == 0) {
      gbase[d] = pre;
      lpre[d] = tot;
    }
    __syncthreads();
    uint32_t v = 0, inc = 0;
    if (t < 256) {
      v = lpre[t];
      inc = v;
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
      }
      if (lane == 63) wsum[w] = inc;
    }
    __syncthreads();
    if (t < 256) {
      uint32_t p = inc - v;
      for (int q2 = 0; q2 < w; ++q2) p += wsum[q2];
      gbase[t] += p;
    }
  } else {
    if (t < 256) gbase[t] = hist[(size_t)t * stride + blk];
  }
  for (int i = t; i < W * 256; i += THREADS) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool valid = base + r * 64 + lane < n;
    const uint32_t d = (key[r] >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(m & lt);
    const uint32_t prev = wcnt[w][d];
    wave_sync();  // every lane has read its digit's count before the leaders update it
    if (valid && rank == 0) wcnt[w][d] = prev + (uint32_t)__popcll(m);
    wave_sync();
    pos[r] = valid ? prev + rank : 0xFFFFFFFFu;  // rank among this wavefront's keys of digit d
  }
  __syncthreads();
  // digit t: its total in the tile, each wavefront's offset inside it, and
  // the exclusive prefix of the totals over the digits
  uint32_t tot = 0, inc = 0;
  if (t < 256) {
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const uint32_t c = wcnt[q][t];
      wcnt[q][t] = tot;
      tot += c;
    }
    inc = tot;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(inc, off, 64);
      if (lane >= off) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
  }
  __syncthreads();
  if (t < 256) {
    uint32_t pre = inc - tot;
    for (int q = 0; q < w; ++q) pre += wsum[q];
    lpre[t] = pre;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (pos[r] != 0xFFFFFFFFu) {
      const uint32_t d = (key[r] >> shift) & 255u;
      const uint32_t slot = lpre[d] + wcnt[w][d] + pos[r];
      skey[slot] = key[r];
      sval[slot] = val[r];
    }
  }
  __syncthreads();
  const int valid_here = min(TILE, n - blk * TILE);
  for (int j = t; j < valid_here; j += THREADS) {
    const uint32_t k = skey[j];
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t g = gbase[d] + ((uint32_t)j - lpre[d]);
    kout[g] = k;
    vout[g] = sval[j];
  }
}

// ------------------------------------------------------- Karras build ------

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t a = codes[i], b = codes[j];
  if (a == b) return 32 + __clz((uint32_t)(i ^ j));
  return __clz(a ^ b);
}

// Thread k initialises leaf k (body perm[k]) and, for k < n-1, the internal
// node n+k with its sorted-position range. Also writes parent links. VEL
// also gathers the velocities (only the 1PN traversal reads node velocities).
template <bool VEL>
__global__ __launch_bounds__(256) void k_build(int n, const uint32_t* __restrict__ codes,
                                               const uint32_t* __restrict__ perm, const double4* __restrict__ xm,
                                               const double* __restrict__ v,
                                               int32_t* __restrict__ left, int32_t* __restrict__ right,
                                               int32_t* __restrict__ parent, int2* __restrict__ grange,
                                               double4* __restrict__ nmc, double4* __restrict__ nvel,
                                               float4* __restrict__ nbox, float4* __restrict__ nrec,
                                               float4* __restrict__ gsh, int32_t* __restrict__ splitnode,
                                               uint32_t* __restrict__ flags, int leaf_records) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  {
    const uint32_t b = perm[k];
    const double4 pm = xm[b];
    const double px = pm.x, py = pm.y, pz = pm.z, mb = pm.w;
    nmc[k] = pm;
    if (VEL) nvel[k] = make_double4(v[b], v[n + b], v[2 * n + b], 0.0);
    const float fx = (float)px, fy = (float)py, fz = (float)pz;
    const float4 r = make_float4(fx, fy, fz, (float)mb);
    if (leaf_records) {  // the moment passes form a leaf's box from nmc themselves
      const float4 p = make_float4(fx, fy, fz, 0.f);
      nbox[2 * k] = p;
      nbox[2 * k + 1] = p;
      nrec[k] = r;
    }
    gsh[k] = r;
    left[k] = -1;
    right[k] = -1;
  }
  if (k == 0) parent[n] = -1;
  if (k >= n - 1) return;
  const int i = k;
  const int d_left = (i == 0) ? -1 : kdelta(codes, n, i, i - 1);
  const int d_right = kdelta(codes, n, i, i + 1);
  const int d = (d_right > d_left) ? 1 : -1;
  const int d_min = (d == 1) ? d_left : d_right;
  int l_max = 2;
  while (kdelta(codes, n, i, i + l_max * d) > d_min) l_max *= 2;
  int l = 0;
  for (int t = l_max / 2; t >= 1; t /= 2)
    if (kdelta(codes, n, i, i + (l + t) * d) > d_min) l += t;
  const int j = i + l * d;
  const int d_node = kdelta(codes, n, i, j);
  int s = 0;
  for (int t = (l + 1) / 2; t >= 1; t = (t + 1) / 2) {
    if (kdelta(codes, n, i, i + (s + t) * d) > d_node) s += t;
    if (t == 1) break;
  }
  const int gamma = i + s * d + min(d, 0);
  const int lo = min(i, j), hi = max(i, j);
  const int L = (lo == gamma) ? gamma : n + gamma;
  const int R = (hi == gamma + 1) ? gamma + 1 : n + gamma + 1;
  if (gamma < 0 || gamma > n - 2 || R >= 2 * n - 1) {  // cannot happen for a sorted key array
    atomicOr(flags, (uint32_t)FLAG_TREE);
    return;
  }
  left[n + i] = L;
  right[n + i] = R;
  parent[L] = n + i;
  parent[R] = n + i;
  grange[i] = make_int2(lo, hi);
  splitnode[gamma] = n + i;  // every split position belongs to exactly one internal node
}

// Depth of every node and its escape link in the right-first preorder (the
// next node after its subtree: the left sibling if it is a right child, else
// its parent's escape), for the per-body traversals and the reference MAC.
__global__ __launch_bounds__(256) void k_depth(int n, const int32_t* __restrict__ left,
                                               const int32_t* __restrict__ right, const int32_t* __restrict__ parent,
                                               int32_t* __restrict__ depth, int32_t* __restrict__ escape,
                                               uint32_t* __restrict__ flags) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = 2 * n - 1;
  if (node >= nn) return;
  int dep = 0, esc = -1, cur = node;
  for (;;) {
    const int p = parent[cur];
    if (p < 0) break;
    if (dep >= kMaxLevels - 1) {
      atomicOr(flags, (uint32_t)FLAG_TREE);
      break;
    }
    if (esc < 0 && right[p] == cur) esc = left[p];
    ++dep;
    cur = p;
  }
  depth[node] = dep;
  escape[node] = esc;
}

// Depths of the nodes above the traversal groups, whose moments the band
// kernels form level by level. They are exactly the internal nodes whose
// split lies at a boundary between consecutive groups (a node above the
// groups splits between two of them, and each boundary is one node's
// split), so thread t < ngroups - 1 takes boundary t, walks up for the
// node's depth and writes entry t of the compact (depth, node) list. Thread
// ngroups - 1 + g walks up from group
// g's root instead: with the deepest climb of the group's leaves
// (k_group_flags) that gives the group's deepest internal node. The block's
// largest internal-node depth goes to dmax[block].
__device__ __forceinline__ int walk_depth(int node, const int32_t* __restrict__ parent, uint32_t* __restrict__ flags) {
  int d = 0;
  for (int p = parent[node]; p >= 0; p = parent[p]) {
    if (++d >= kMaxLevels - 1) {
      atomicOr(flags, (uint32_t)FLAG_TREE);
      break;
    }
  }
  return d;
}

__global__ __launch_bounds__(256) void k_above_census(int n, const int32_t* __restrict__ gstart,
                                                      const uint32_t* __restrict__ gcount,
                                                      const int32_t* __restrict__ splitnode,
                                                      const int32_t* __restrict__ groot,
                                                      const int32_t* __restrict__ climb,
                                                      const int32_t* __restrict__ parent, int32_t* __restrict__ depth,
                                                      uint32_t* __restrict__ cdep, uint32_t* __restrict__ cnode,
                                                      int32_t* __restrict__ dmax, uint32_t* __restrict__ gcnt,
                                                      uint32_t* __restrict__ flags) {
  __shared__ int wmax[4];
  if (blockIdx.x == 0) gcnt[threadIdx.x] = 0u;  // the 2 kMaxLevels = 256 census counters and cursors
  const int ng = (int)*gcount;
  int mx = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * ng - 1; t += gridDim.x * blockDim.x) {
    if (t < ng - 1) {
      const int node = splitnode[gstart[t + 1] - 1];
      const int d = walk_depth(node, parent, flags);
      depth[node] = d;
      cdep[t] = (uint32_t)d;  // one entry per node above the groups
      cnode[t] = (uint32_t)node;
      mx = max(mx, d);
    } else {
      const int g = t - (ng - 1);
      const int lo = gstart[g], hi = gstart[g + 1] - 1;
      // a group has at most kGroup leaves: all their loads are issued at once
      // (indices clamped to the group) rather than one round trip per leaf
      int cm = 0;
#pragma unroll
      for (int k = 0; k < kGroup; ++k) cm = max(cm, climb[min(lo + k, hi)]);
      if (cm > 0) mx = max(mx, walk_depth(groot[lo], parent, flags) + cm - 1);
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) dmax[blockIdx.x] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

// The census list ordered by level, in two passes over tiles of kRsTile
// entries (8 per thread, 256 threads). k_census_count: per-tile depth counts
// in LDS, added into gcnt[0..127]. k_census_place: every tile derives the
// level offsets from gcnt, reserves a range of each level for its entries
// with one atomic per depth on the cursors gcnt[128..255], and writes its
// nodes there; tile 0 also writes levels[] (offsets, level count, largest
// depth over the census blocks' maxima). The order inside a level is
// arbitrary: each node's moments depend on its children only. Lanes of a
// wavefront with equal depths (few distinct values, many nodes each) are
// counted with one LDS atomic per distinct depth.
static_assert(kMaxLevels == 128, "depth_match compares 7 bits; the census keeps 2 x 128 counters");
__device__ __forceinline__ uint64_t depth_match(bool valid, uint32_t d, uint64_t lt, uint32_t& rank) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 7; ++b) {  // depth < kMaxLevels = 2^7
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  rank = (uint32_t)__popcll(m & lt);
  return m;
}

__global__ __launch_bounds__(kRsThreads) void k_census_count(const uint32_t* __restrict__ gcount,
                                                             const uint32_t* __restrict__ cdep,
                                                             uint32_t* __restrict__ gcnt) {
  __shared__ uint32_t lcnt[kMaxLevels];
  const int t = threadIdx.x, lane = t & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int cnt = (int)*gcount - 1;  // one node above the groups per boundary between groups
  const int base = blockIdx.x * kRsTile;
  if (base >= cnt) return;  // uniform over the block
  if (t < kMaxLevels) lcnt[t] = 0u;
  __syncthreads();
  uint32_t d[kRsTile / kRsThreads];
#pragma unroll
  for (int r = 0; r < kRsTile / kRsThreads; ++r) {
    const int i = base + r * kRsThreads + t;
    d[r] = i < cnt ? cdep[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kRsTile / kRsThreads; ++r) {
    const bool valid = base + r * kRsThreads + t < cnt;
    uint32_t rank;
    const uint64_t m = depth_match(valid, d[r], lt, rank);
    if (valid && rank == 0) atomicAdd(&lcnt[d[r]], (uint32_t)__popcll(m));
  }
  __syncthreads();
  if (t < kMaxLevels && lcnt[t]) atomicAdd(&gcnt[t], lcnt[t]);
}

__global__ __launch_bounds__(kRsThreads) void k_census_place(const uint32_t* __restrict__ gcount,
                                                             const uint32_t* __restrict__ cdep,
                                                             const uint32_t* __restrict__ cnode,
                                                             uint32_t* __restrict__ gcnt,
                                                             const int32_t* __restrict__ dmax, int cblocks,
                                                             int32_t* __restrict__ bucket,
                                                             int32_t* __restrict__ levels) {
  constexpr int R = kRsTile / kRsThreads;
  __shared__ uint32_t off[kMaxLevels], lcnt[kMaxLevels], lbase[kMaxLevels];
  __shared__ int wmax[kRsThreads / 64];
  const int t = threadIdx.x, lane = t & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int cnt = (int)*gcount - 1;
  const int base = blockIdx.x * kRsTile;
  if (blockIdx.x > 0 && base >= cnt) return;  // uniform over the block; tile 0 always writes levels[]
  if (t < kMaxLevels) lcnt[t] = 0u;
  if (t == 0) {
    uint32_t run = 0;
    for (int q = 0; q < kMaxLevels; ++q) {
      off[q] = run;
      run += gcnt[q];
    }
  }
  __syncthreads();
  uint32_t d[R], node[R], pos[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = base + r * kRsThreads + t;
    d[r] = i < cnt ? cdep[i] : 0u;
    node[r] = i < cnt ? cnode[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool valid = base + r * kRsThreads + t < cnt;
    uint32_t rank, first = 0;
    const uint64_t m = depth_match(valid, d[r], lt, rank);
    if (valid && rank == 0) first = atomicAdd(&lcnt[d[r]], (uint32_t)__popcll(m));
    first = __shfl(first, (int)__builtin_ctzll(m | (1ull << lane)), 64);  // from the group's lowest lane
    pos[r] = first + rank;
  }
  __syncthreads();
  if (t < kMaxLevels && lcnt[t]) lbase[t] = off[t] + atomicAdd(&gcnt[kMaxLevels + t], lcnt[t]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (base + r * kRsThreads + t < cnt) bucket[lbase[d[r]] + pos[r]] = (int32_t)node[r];
  if (blockIdx.x == 0) {
    int mx = 0;
    for (int i = t; i < cblocks; i += kRsThreads) mx = max(mx, dmax[i]);
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if (lane == 0) wmax[t >> 6] = mx;
    __syncthreads();
    if (t < kMaxLevels) levels[t] = (int32_t)off[t];
    if (t == 0) {
      int nlev = 0;
      for (int q = 0; q < kMaxLevels; ++q)
        if (gcnt[q]) nlev = q + 1;
      levels[kMaxLevels + 1] = nlev;
      levels[kMaxLevels + 2] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
    }
  }
}

// A node's moments: FP64 centre of mass and mass (and centre-of-mass velocity),
// FP32 bounding box.
struct MomRec {
  double4 com, vel;
  float4 lo, hi;
};

// The group traversal reads, for each internal node, its two children's
// records, node ids and thresholds from slots kid_slot(node, 0/1): opening a
// node costs one load of 64 contiguous bytes. The node that forms its moments
// writes its children's slots; the root has the extra slot 2 (n - 1).
__host__ __device__ inline int kid_slot(int n, int parent, int side) { return 2 * (parent - n) + side; }

// Geometric opening threshold of a node from its FP32 box and record: accept
// when d^2 > crit2 = (l / theta + delta)^2, l = longest AABB side, delta =
// |COM - AABB centre| (a leaf's degenerate box gives 0).
__device__ __forceinline__ float geo_crit2(const float4& lo, const float4& hi, const float4& r, double theta) {
#pragma clang fp contract(off)
  const float l = fmaxf(hi.x - lo.x, fmaxf(hi.y - lo.y, hi.z - lo.z));
  const float cx = 0.5f * (hi.x + lo.x);
  const float cy = 0.5f * (hi.y + lo.y);
  const float cz = 0.5f * (hi.z + lo.z);
  const float ex = r.x - cx, ey = r.y - cy, ez = r.z - cz;
  const float delta = sqrtf(ex * ex + ey * ey + ez * ez);
  const float q = l / (float)theta + delta;
  return q * q;
}

__device__ __forceinline__ float4 node_rec(const double4& com) {
  return make_float4((float)com.x, (float)com.y, (float)com.z, (float)com.w);
}

// The group traversal's child slot (see k_links): record, node id, threshold.
__device__ __forceinline__ void put_kid(float4* __restrict__ kids, int slot, int node, const double4& com,
                                        const float4& lo, const float4& hi, double theta) {
  const float4 r = node_rec(com);
  kids[2 * slot] = r;
  kids[2 * slot + 1] = make_float4(__int_as_float(node), geo_crit2(lo, hi, r, theta), 0.f, 0.f);
}

// A leaf's box is its position rounded to FP32 (what k_build would store).
__device__ __forceinline__ float4 leaf_box(const double4& com) {
  return make_float4((float)com.x, (float)com.y, (float)com.z, 0.f);
}

template <bool VEL>
__device__ __forceinline__ MomRec load_moments(int n, int node, const double4* __restrict__ nmc,
                                               const double4* __restrict__ nvel, const float4* __restrict__ nbox) {
  MomRec r;
  r.com = nmc[node];
  r.vel = VEL ? nvel[node] : make_double4(0.0, 0.0, 0.0, 0.0);
  if (node < n) {
    r.lo = r.hi = leaf_box(r.com);
  } else {
    r.lo = nbox[2 * node];
    r.hi = nbox[2 * node + 1];
  }
  return r;
}

template <bool VEL>
__device__ __forceinline__ MomRec unite_moments(const MomRec& cl, const MomRec& cr) {
#pragma clang fp contract(off)
  const double mR = cr.com.w, mL = cl.com.w;
  const double total = mR + mL;
  MomRec me;
  me.com = make_double4(0.0, 0.0, 0.0, 0.0);
  me.vel = make_double4(0.0, 0.0, 0.0, 0.0);
  if (total > 0) {
    me.com.w = total;
    me.com.x = (cr.com.x * mR + cl.com.x * mL) / total;
    me.com.y = (cr.com.y * mR + cl.com.y * mL) / total;
    me.com.z = (cr.com.z * mR + cl.com.z * mL) / total;
    if (VEL) {
      me.vel.x = (cr.vel.x * mR + cl.vel.x * mL) / total;
      me.vel.y = (cr.vel.y * mR + cl.vel.y * mL) / total;
      me.vel.z = (cr.vel.z * mR + cl.vel.z * mL) / total;
    }
  }
  me.lo = make_float4(fminf(cl.lo.x, cr.lo.x), fminf(cl.lo.y, cr.lo.y), fminf(cl.lo.z, cr.lo.z), 0.f);
  me.hi = make_float4(fmaxf(cl.hi.x, cr.hi.x), fmaxf(cl.hi.y, cr.hi.y), fmaxf(cl.hi.z, cr.hi.z), 0.f);
  return me;
}

template <bool VEL>
__device__ __forceinline__ MomRec shfl_xor_rec(const MomRec& r, int m) {
  MomRec o;
  o.com = make_double4(__shfl_xor(r.com.x, m, 64), __shfl_xor(r.com.y, m, 64), __shfl_xor(r.com.z, m, 64),
                       __shfl_xor(r.com.w, m, 64));
  o.vel = VEL ? make_double4(__shfl_xor(r.vel.x, m, 64), __shfl_xor(r.vel.y, m, 64), __shfl_xor(r.vel.z, m, 64), 0.0)
              : make_double4(0.0, 0.0, 0.0, 0.0);
  o.lo = make_float4(__shfl_xor(r.lo.x, m, 64), __shfl_xor(r.lo.y, m, 64), __shfl_xor(r.lo.z, m, 64), 0.f);
  o.hi = make_float4(__shfl_xor(r.hi.x, m, 64), __shfl_xor(r.hi.y, m, 64), __shfl_xor(r.hi.z, m, 64), 0.f);
  return o;
}

// Band of up to H <= 6 levels above the groups, formed by a team of 2^H lanes
// per node at the band's top depth. Lane s of a team stands for slot s of the
// band's bottom level: a node at band level l covers the slots
// [k 2^(H-l), (k+1) 2^(H-l)) and is handled by the first of them, its leader.
// Going down, the leaders of the nodes formed here load their children and
// leaf ranges (one round of loads per level; the ranges tell which children
// lie above the groups) and the other lanes of the node take theirs by
// shuffle. Then the leader of every child that is already complete (group
// roots, leaves, nodes of the band below) loads its record, all in one round.
// Going up, a leader unites its own record (its node's left child) with the
// one held half its span away (the right child). H levels cost H + 2
// dependent load rounds (the band top, H levels of children, the records)
// and one launch; a team is at most one wavefront, so
// no barrier is needed. Bottom-up order is what unite (barnes_hut.cpp:137-152)
// needs: kernel boundaries order the bands.
template <bool VEL, int H>
__global__ __launch_bounds__(256) void k_moments_band(int n, const int32_t* __restrict__ bucket, int count,
                                                      const int32_t* __restrict__ left,
                                                      const int32_t* __restrict__ right,
                                                      const int2* __restrict__ grange, double4* __restrict__ nmc,
                                                      double4* __restrict__ nvel, float4* __restrict__ nbox,
                                                      float4* __restrict__ nrec, double theta,
                                                      float4* __restrict__ kids) {
  static_assert(H >= 1 && H <= 6, "a band's team is at most one wavefront");
  constexpr int S = 1 << H;
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int team = g >> H;
  if (team >= count) return;  // teams are aligned to S lanes and leave whole
  const int lane = threadIdx.x & 63;
  const int s = g & (S - 1);
  int ids[H];            // node covering this slot at band level l
  uint32_t inmask = 0;   // bit l: that node is formed here
  int cur = bucket[team];
  bool cin = true;
  int clev = 0;  // band level at which the complete node covering this slot appears
#pragma unroll
  for (int l = 0; l < H; ++l) {
    const int span = S >> l;
    ids[l] = cur;
    if (cin) inmask |= 1u << l;
    int L = 0, R = 0, fl = 0;
    if (cin && (s & (span - 1)) == 0) {
      L = left[cur];
      R = right[cur];
      const int2 r = grange[cur - n];  // children cover [r.x, gamma] and [gamma + 1, r.y]
      const int gamma = L < n ? L : L - n;
      if (l + 1 < H)
        fl = (L >= n && gamma - r.x + 1 > kGroup ? 1 : 0) | (R >= n && r.y - gamma > kGroup ? 2 : 0);
    }
    const int src = lane & ~(span - 1);
    L = __shfl(L, src, 64);
    R = __shfl(R, src, 64);
    fl = __shfl(fl, src, 64);
    if (cin) {
      const bool rt = (s & (span >> 1)) != 0;
      cur = rt ? R : L;
      cin = ((rt ? fl >> 1 : fl) & 1) != 0;
      clev = l + 1;
    }
  }
  MomRec rec{};
  if ((s & ((S >> clev) - 1)) == 0) rec = load_moments<VEL>(n, cur, nmc, nvel, nbox);
  int cid = cur;  // node whose record this lane holds
#pragma unroll
  for (int l = H - 1; l >= 0; --l) {
    const int half = S >> (l + 1);
    const MomRec o = shfl_xor_rec<VEL>(rec, half);
    const int oid = __shfl_xor(cid, half, 64);
    if (((inmask >> l) & 1u) && (s & (2 * half - 1)) == 0) {
      const int node = ids[l];
      const MomRec me = unite_moments<VEL>(rec, o);
      nmc[node] = me.com;
      if (VEL) nvel[node] = me.vel;
      nbox[2 * node] = me.lo;
      nbox[2 * node + 1] = me.hi;
      if (!kids) nrec[node] = node_rec(me.com);
      if (kids) {  // this node's child slots (and the root's own slot)
        put_kid(kids, kid_slot(n, node, 0), cid, rec.com, rec.lo, rec.hi, theta);
        put_kid(kids, kid_slot(n, node, 1), oid, o.com, o.lo, o.hi, theta);
        if (node == n) put_kid(kids, 2 * (n - 1), n, me.com, me.lo, me.hi, theta);
      }
      rec = me;
      cid = node;
    }
  }
}

// The top levels above the groups (at most kTopNodes nodes in all) in one
// workgroup: thread p owns the p-th node of the depth-sorted list. Its
// children's ids and leaf ranges, then the records of the children that are
// already complete (formed by the bands below, group roots, leaves), are
// loaded up front; the levels are then formed bottom-up through LDS, one
// barrier per level, instead of one band launch per 3 levels.
constexpr int kTopNodes = 512;
constexpr int kTopMaxLevels = 32;
struct TopLevels {
  int nlev;                    // levels 0 .. nlev - 1
  int off[kTopMaxLevels + 1];  // their offsets in the depth-sorted list
};

template <bool VEL>
__global__ __launch_bounds__(kTopNodes) void k_moments_top(int n, const int32_t* __restrict__ bucket, TopLevels tl,
                                                           const int32_t* __restrict__ left,
                                                           const int32_t* __restrict__ right,
                                                           const int2* __restrict__ grange, double4* __restrict__ nmc,
                                                           double4* __restrict__ nvel, float4* __restrict__ nbox,
                                                           float4* __restrict__ nrec, double theta,
                                                           float4* __restrict__ kids, int32_t* __restrict__ topslot) {
  __shared__ double4 scom[kTopNodes];
  __shared__ double4 svel[VEL ? kTopNodes : 1];
  __shared__ float4 slo[kTopNodes], shi[kTopNodes];
  const int p = threadIdx.x;
  const bool mine = p < tl.off[tl.nlev];
  int node = n, L = 0, R = 0, lev = 0;
  bool inL = false, inR = false;
  if (mine) {
    node = bucket[p];
    while (tl.off[lev + 1] <= p) ++lev;
    L = left[node];
    R = right[node];
    const int2 r = grange[node - n];  // children cover [r.x, gamma] and [gamma + 1, r.y]
    const int gamma = L < n ? L : L - n;
    const bool below = lev + 1 < tl.nlev;  // children at a level formed here
    inL = below && L >= n && gamma - r.x + 1 > kGroup;
    inR = below && R >= n && r.y - gamma > kGroup;
    topslot[node - n] = p;
  }
  __syncthreads();
  MomRec cl{}, cr{};
  int sl = -1, sr = -1;
  if (mine) {
    if (inL) sl = topslot[L - n];
    else cl = load_moments<VEL>(n, L, nmc, nvel, nbox);
    if (inR) sr = topslot[R - n];
    else cr = load_moments<VEL>(n, R, nmc, nvel, nbox);
  }
  for (int d = tl.nlev - 1; d >= 0; --d) {
    if (mine && lev == d) {
      if (sl >= 0) {
        cl.com = scom[sl];
        if (VEL) cl.vel = svel[sl];
        cl.lo = slo[sl];
        cl.hi = shi[sl];
      }
      if (sr >= 0) {
        cr.com = scom[sr];
        if (VEL) cr.vel = svel[sr];
        cr.lo = slo[sr];
        cr.hi = shi[sr];
      }
      const MomRec me = unite_moments<VEL>(cl, cr);
      scom[p] = me.com;
      if (VEL) svel[p] = me.vel;
      slo[p] = me.lo;
      shi[p] = me.hi;
      nmc[node] = me.com;
      if (VEL) nvel[node] = me.vel;
      nbox[2 * node] = me.lo;
      nbox[2 * node + 1] = me.hi;
      if (!kids) nrec[node] = node_rec(me.com);
      if (kids) {  // this node's child slots (and the root's own slot)
        put_kid(kids, kid_slot(n, node, 0), L, cl.com, cl.lo, cl.hi, theta);
        put_kid(kids, kid_slot(n, node, 1), R, cr.com, cr.lo, cr.hi, theta);
        if (node == n) put_kid(kids, 2 * (n - 1), n, me.com, me.lo, me.hi, theta);
      }
    }
    __syncthreads();
  }
}

// Moments of every internal node inside a traversal group, one wavefront per
// group. Internal node n+i's range starts or ends at i (Karras), so the
// hi - lo internal nodes of a group with leaf range [lo, hi] are n+lo ..
// n+hi-1 when the group's root is n+lo, and n+lo+1 .. n+hi when it is n+hi.
// Lane t owns node n+base+t; the lanes sweep bottom-up, a node being ready
// once both children are (leaves always are). Finished records stay in LDS
// for the parents in the same wavefront; leaves come from k_build. Same
// arithmetic as unite_moments, so the results are the same bits.
template <bool VEL>
__global__ __launch_bounds__(256) void k_group_moments(int n, const int32_t* __restrict__ gstart,
                                                       const uint32_t* __restrict__ gcount,
                                                       const int2* __restrict__ grange,
                                                       const int32_t* __restrict__ left,
                                                       const int32_t* __restrict__ right, double4* __restrict__ nmc,
                                                       double4* __restrict__ nvel, float4* __restrict__ nbox,
                                                       float4* __restrict__ nrec, double theta,
                                                       float4* __restrict__ kids) {
#pragma clang fp contract(off)
  __shared__ double4 smc[4][64];
  __shared__ double4 svel[VEL ? 4 : 1][64];
  __shared__ float4 slo[4][64], shi[4][64];
  __shared__ int sdone[4][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ngroups = (int)*gcount;
  const int stride = (int)gridDim.x * 4;
  for (int g = blockIdx.x * 4 + wl; g < ngroups; g += stride) {
    const int lo = gstart[g], hi = gstart[g + 1] - 1;
    const int nint = hi - lo;  // internal nodes of the group
    if (nint == 0) continue;   // a single leaf (uniform across the wavefront)
    bool root_lo = lo <= n - 2;
    if (root_lo) {
      const int2 r = grange[lo];
      root_lo = r.x == lo && r.y == hi;
    }
    const int base = root_lo ? lo : lo + 1;
    const bool mine = lane < nint;
    const int node = n + base + lane;
    int L = 0, R = 0;
    // leaf children are loaded before the sweep (one round trip); internal
    // children come from LDS as their lanes finish
    double4 cL = make_double4(0.0, 0.0, 0.0, 0.0), cR = cL, vL = cL, vR = cL;
    float4 loL = make_float4(0.f, 0.f, 0.f, 0.f), hiL = loL, loR = loL, hiR = loL;
    if (mine) {
      L = left[node];
      R = right[node];
      if (L < n) {
        cL = nmc[L];
        if (VEL) vL = nvel[L];
        loL = hiL = leaf_box(cL);
      }
      if (R < n) {
        cR = nmc[R];
        if (VEL) vR = nvel[R];
        loR = hiR = leaf_box(cR);
      }
    }
    bool done = !mine;
    sdone[wl][lane] = done ? 1 : 0;
    wave_sync();
    double4 com = make_double4(0.0, 0.0, 0.0, 0.0), vel = com;
    float4 blo = loL, bhi = loL;
    for (int it = 0; it < 64; ++it) {
      bool ready = false;
      if (!done) {
        const bool rl = L < n || sdone[wl][L - n - base];
        const bool rr = R < n || sdone[wl][R - n - base];
        ready = rl && rr;
      }
      if (ready) {
        if (L >= n) {
          cL = smc[wl][L - n - base];
          if (VEL) vL = svel[wl][L - n - base];
          loL = slo[wl][L - n - base];
          hiL = shi[wl][L - n - base];
        }
        if (R >= n) {
          cR = smc[wl][R - n - base];
          if (VEL) vR = svel[wl][R - n - base];
          loR = slo[wl][R - n - base];
          hiR = shi[wl][R - n - base];
        }
        const double mR = cR.w, mL = cL.w;
        const double total = mR + mL;
        if (total > 0) {
          com.w = total;
          com.x = (cR.x * mR + cL.x * mL) / total;
          com.y = (cR.y * mR + cL.y * mL) / total;
          com.z = (cR.z * mR + cL.z * mL) / total;
          if (VEL) {
            vel.x = (vR.x * mR + vL.x * mL) / total;
            vel.y = (vR.y * mR + vL.y * mL) / total;
            vel.z = (vR.z * mR + vL.z * mL) / total;
          }
        }
        blo = make_float4(fminf(loL.x, loR.x), fminf(loL.y, loR.y), fminf(loL.z, loR.z), 0.f);
        bhi = make_float4(fmaxf(hiL.x, hiR.x), fmaxf(hiL.y, hiR.y), fmaxf(hiL.z, hiR.z), 0.f);
        smc[wl][lane] = com;
        if (VEL) svel[wl][lane] = vel;
        slo[wl][lane] = blo;
        shi[wl][lane] = bhi;
        done = true;
      }
      wave_sync();
      if (ready) sdone[wl][lane] = 1;
      wave_sync();
      if (__ballot(!done) == 0ull) break;
    }
    if (mine) {
      nmc[node] = com;
      if (VEL) nvel[node] = vel;
      // With child slots (the group traversal) nobody reads nrec, nor the
      // box of a node below its group's root.
      if (!kids || node == n + (root_lo ? lo : hi)) {
        nbox[2 * node] = blo;
        nbox[2 * node + 1] = bhi;
      }
      if (!kids) nrec[node] = node_rec(com);
      if (kids) {  // this node's child slots (and the root's own slot)
        put_kid(kids, kid_slot(n, node, 0), L, cL, loL, hiL, theta);
        put_kid(kids, kid_slot(n, node, 1), R, cR, loR, hiR, theta);
        if (node == n) put_kid(kids, 2 * (n - 1), n, com, blo, bhi, theta);
      }
    }
  }
}

// Per node: traversal record = right child (or -1), escape link, opening
// threshold crit2 (accept when d^2 > crit2) and, for leaves, the body.
// mac 0 (parity): crit2 = (s / theta)^2 with s = |box| 2^-depth.
// mac 1 (geometric): crit2 = (l / theta + delta)^2, l = longest AABB side,
//                    delta = |COM - AABB centre|.
__global__ __launch_bounds__(256) void k_links(int n, const int32_t* __restrict__ right,
                                               const int32_t* __restrict__ escape, const int32_t* __restrict__ depth,
                                               const float4* __restrict__ nbox, const float4* __restrict__ nrec,
                                               const double* __restrict__ boxd, double theta, int mac,
                                               int4* __restrict__ nlink, const uint32_t* __restrict__ perm) {
  const int node = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = 2 * n - 1;
  if (node >= nn) return;
  float crit2;
  if (mac == 0) {
    const int dep = depth[node];
    const double s = ldexp(boxd[3], -dep);
    const double q = s / theta;
    crit2 = (float)(q * q);
  } else {
    crit2 = geo_crit2(nbox[2 * node], nbox[2 * node + 1], nrec[node], theta);
  }
  const int body = node < n ? (int)perm[node] : -1;
  nlink[node] = make_int4(node < n ? -1 : right[node], escape[node], __float_as_int(crit2), body);
}

// --------------------------------------------------------- traversals ------

// FP32 per-body stackless traversal (lanes = consecutive sorted bodies).
// Post-Newtonian terms of the FP32 traversals: for every node a body takes,
// the pairwise 1PN / 2PN / 2.5PN terms of compute_gravitational
// (tool_chain_impl.cpp:126-136) in FP64, with the node's FP64 centre of mass,
// mass and mass-weighted velocity as get_node_body takes them
// (barnes_hut.cpp:202-261; nodes at distance 0 skipped, :235), times 1/m_i,
// added to the FP32 Newtonian sums.
struct NodePn {
  const double4* nmc;   // FP64 {centre of mass, mass} per node (leaf k: body perm[k])
  const double4* nvel;  // FP64 mass-weighted velocity per node
  double G, ic, soft;
  int pnf;
};

// The terms from a node (record nc, velocity nv) on the body of leaf record
// pme / velocity vme.
__device__ __forceinline__ void node_pn_rec(const NodePn& p, const double4& nc, const double4& nv,
                                            const double4& pme, const double4& vme, double inv_m, double& ax,
                                            double& ay, double& az) {
#pragma clang fp contract(off)
  const double dx = nc.x - pme.x, dy = nc.y - pme.y, dz = nc.z - pme.z;
  const double d2 = dx * dx + dy * dy + dz * dz;
  if (d2 == 0.0) return;
  if (p.pnf == ASTER_PN1) {
    // the pairwise 1PN term alone: get_pn1 (tool_chain_impl.cpp:110-118) as an
    // acceleration, n = (x_i - x_node)/r, grouped for fused multiply-adds:
    //   G m2/r^2 [n (-v1^2 - 2 v2^2 + 4 v1.v2 + 3/2 (n.v2)^2 + (5 G m1 + 4 G m2)/r)
    //             + (v1 - v2)(4 n.v1 - 3 n.v2)] / c^2
    const double inv = rsqrt64(d2 + p.soft);
    const double nx = -dx * inv, ny = -dy * inv, nz = -dz * inv;
    const double nv1 = fma(nx, vme.x, fma(ny, vme.y, nz * vme.z));
    const double nv2 = fma(nx, nv.x, fma(ny, nv.y, nz * nv.z));
    const double v11 = fma(vme.x, vme.x, fma(vme.y, vme.y, vme.z * vme.z));
    const double v12 = fma(vme.x, nv.x, fma(vme.y, nv.y, vme.z * nv.z));
    const double v22 = fma(nv.x, nv.x, fma(nv.y, nv.y, nv.z * nv.z));
    double sc = fma(4.0, v12, -v11);
    sc = fma(-2.0, v22, sc);
    sc = fma(1.5 * nv2, nv2, sc);
    sc = fma(fma(4.0 * p.G, nc.w, 5.0 * p.G * pme.w), inv, sc);
    const double t = fma(4.0, nv1, -3.0 * nv2);
    const double k = p.G * nc.w * (inv * inv) * (p.ic * p.ic);
    ax = fma(fma(nx, sc, (vme.x - nv.x) * t), k, ax);
    ay = fma(fma(ny, sc, (vme.y - nv.y) * t), k, ay);
    az = fma(fma(nz, sc, (vme.z - nv.z) * t), k, az);
    return;
  }
  const double inv_r = 1.0 / sqrt(d2 + p.soft);
  double fx = 0.0, fy = 0.0, fz = 0.0;
  pn_add64(p.pnf, pme.w, nc.w, vme.x, vme.y, vme.z, nv.x, nv.y, nv.z, inv_r, -(dx * inv_r), -(dy * inv_r),
           -(dz * inv_r), p.G, p.ic, fx, fy, fz);
  ax += fx * inv_m;
  ay += fy * inv_m;
  az += fz * inv_m;
}

__device__ __forceinline__ void node_pn(const NodePn& p, int node, const double4& pme, const double4& vme,
                                        double inv_m, double& ax, double& ay, double& az) {
  node_pn_rec(p, p.nmc[node], p.nvel[node], pme, vme, inv_m, ax, ay, az);
}

template <bool MASK, bool PN>
__global__ __launch_bounds__(256) void k_bh_f32(int n, const float4* __restrict__ gsh, const float4* __restrict__ nrec,
                                                const int4* __restrict__ nlink, float eps2, float msc, double G,
                                                double* __restrict__ a, unsigned long long* __restrict__ counters,
                                                uint32_t* __restrict__ flags, NodePn pnp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < n;
  const float4 me = gsh[live ? k : n - 1];
  float ax = 0.f, ay = 0.f, az = 0.f;
  double4 pme = make_double4(0.0, 0.0, 0.0, 1.0), vme = make_double4(0.0, 0.0, 0.0, 0.0);
  double pax = 0.0, pay = 0.0, paz = 0.0, inv_m = 1.0;
  if (PN) {
    pme = pnp.nmc[live ? k : n - 1];
    vme = pnp.nvel[live ? k : n - 1];
    inv_m = 1.0 / pme.w;
  }
  uint32_t inter = 0, visits = 0;
  int node = live ? n : -1;
  const uint32_t limit = 4u * (uint32_t)n + 16u;
  while (node >= 0) {
    if (visits > limit) {
      atomicOr(flags, (uint32_t)FLAG_TREE);
      break;
    }
    const float4 r = nrec[node];
    const int4 L = nlink[node];
    ++visits;
    const float dx = r.x - me.x, dy = r.y - me.y, dz = r.z - me.z;
    const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (d2 == 0.f) {
      node = L.y;
      continue;
    }
    if (node < n || d2 > __int_as_float(L.z)) {
      const float r2 = d2 + eps2;
      float inv = __builtin_amdgcn_rsqf(r2);
      if (MASK) inv = (r2 > 0.f) ? inv : 0.f;
      const float f = ((r.w * msc) * inv) * (inv * inv);
      ax = fmaf(f, dx, ax);
      ay = fmaf(f, dy, ay);
      az = fmaf(f, dz, az);
      if (PN) node_pn(pnp, node, pme, vme, inv_m, pax, pay, paz);
      ++inter;
      node = L.y;
    } else {
      node = L.x;
    }
  }
  if (live) {
    const int b = nlink[k].w;
    a[b] = G * (double)ax + pax;
    a[n + b] = G * (double)ay + pay;
    a[2 * n + b] = G * (double)az + paz;
  }
  unsigned long long ci = inter, cv = visits;
  for (int off = 32; off > 0; off >>= 1) {
    ci += __shfl_xor(ci, off, 64);
    cv += __shfl_xor(cv, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counters[BHC_INTERACTIONS], ci);
    atomicAdd(&counters[BHC_VISITS], cv);
  }
}

// FP64 per-body traversal in the reference's order and arithmetic
// (get_node_body, barnes_hut.cpp:202-261).
template <bool PN>
__global__ __launch_bounds__(256) void k_bh_f64(int n, const double* __restrict__ x, const double* __restrict__ v,
                                                const double* __restrict__ m, const double4* __restrict__ nmc,
                                                const double4* __restrict__ nvel, const int4* __restrict__ nlink, const int32_t* __restrict__ depth,
                                                const double* __restrict__ boxd, double theta, double G, double ic,
                                                int pnf,
                                                double soft, double* __restrict__ a,
                                                unsigned long long* __restrict__ counters,
                                                uint32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = k < n;
  const int b = nlink[live ? k : n - 1].w;
  const double pix = x[b], piy = x[n + b], piz = x[2 * n + b];
  double vix = 0.0, viy = 0.0, viz = 0.0;
  if (PN) {
    vix = v[b];
    viy = v[n + b];
    viz = v[2 * n + b];
  }
  const double mi = m[b];
  const double inv_m = 1.0 / mi;
  const double theta2 = theta * theta;
  const double bd = boxd[3];
  double ax = 0.0, ay = 0.0, az = 0.0;
  uint32_t inter = 0, visits = 0;
  int node = live ? n : -1;
  const uint32_t limit = 4u * (uint32_t)n + 16u;
  while (node >= 0) {
    if (visits > limit) {
      atomicOr(flags, (uint32_t)FLAG_TREE);
      break;
    }
    const int4 L = nlink[node];
    ++visits;
    const double4 nc = nmc[node];
    const double cx = nc.x, cy = nc.y, cz = nc.z;
    const double dx = pix - cx, dy = piy - cy, dz = piz - cz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 == 0.0) {
      node = L.y;
      continue;
    }
    const double s = ldexp(bd, -depth[node]);
    if ((s * s < d2 * theta2) || node < n) {
      double vnx = 0.0, vny = 0.0, vnz = 0.0;
      if (PN) {
        const double4 nv = nvel[node];
        vnx = nv.x;
        vny = nv.y;
        vnz = nv.z;
      }
      double fx, fy, fz;
      grav64(mi, nc.w, vix, viy, viz, vnx, vny, vnz, pix, piy, piz, cx, cy, cz, G, ic, soft, PN ? pnf : 0, fx, fy,
             fz);
      ax += fx * inv_m;
      ay += fy * inv_m;
      az += fz * inv_m;
      ++inter;
      node = L.y;
    } else {
      node = L.x;
    }
  }
  if (live) {
    a[b] = ax;
    a[n + b] = ay;
    a[2 * n + b] = az;
  }
  unsigned long long ci = inter, cv = visits;
  for (int off = 32; off > 0; off >>= 1) {
    ci += __shfl_xor(ci, off, 64);
    cv += __shfl_xor(cv, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counters[BHC_INTERACTIONS], ci);
    atomicAdd(&counters[BHC_VISITS], cv);
  }
}

// ------------------------------------------------------------- groups ------
//
// The wavefront traversal works on groups of bodies that share one walk. A
// group is a subtree of the radix tree with at most kGroup leaves whose
// parent has more: its bodies are contiguous in Morton order and lie in one
// Morton cell (the subtree's common code prefix), so the group box is compact
// wherever the curve jumps. Groups partition the sorted bodies.

// Leaf k climbs while the parent still has at most kGroup leaves; k starts a
// group when it is the first position of the subtree it stops at.
// The climb's length (the leaf's depth below its group's root) and, at the
// group's first position, the root itself are kept for the depth census.
__global__ __launch_bounds__(256) void k_group_flags(int n, const int32_t* __restrict__ parent,
                                                     const int2* __restrict__ grange, uint32_t* __restrict__ flag,
                                                     int32_t* __restrict__ climb, int32_t* __restrict__ groot) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int v = k, lo = k, steps = 0;
  for (;;) {
    const int p = parent[v];
    if (p < 0) break;
    const int2 r = grange[p - n];
    if (r.y - r.x + 1 > kGroup) break;
    v = p;
    lo = r.x;
    ++steps;
  }
  flag[k] = (lo == k) ? 1u : 0u;
  climb[k] = steps;
  if (lo == k) groot[k] = v;
}


// Wavefront-cooperative traversal, one group per wavefront (lanes = the
// group's bodies, at most 64). The wavefront walks the tree once for the
// whole group with a shared LDS stack of opened nodes, testing up to 64
// candidate nodes per iteration: lane 2 i + s takes child s of the i-th popped
// node from that node's child slots (record, id and threshold of both
// children in 64 contiguous bytes), so a batch costs one load round trip.
// A node is accepted when the distance from its centre of
// mass to the group's bounding box exceeds its opening radius (the geometric
// test, so it holds for every body of the group); leaves are always accepted.
// Accepted nodes are appended to an LDS interaction list that the lanes then
// evaluate in pieces (below; 12 VALU + 1 rsq per entry and lane), FP32 sums
// per list flush, Kahan-compensated across flushes.
// Stack bound: each iteration pops the top min(sp, 32) entries and pushes
// their opened children (at most 64) in order, so the stack stays sorted by
// depth from bottom to top and a pop only ever leaves behind part of one
// pushed batch: at most 32 entries per depth below a top batch of at most 64.
// With internal nodes at depths 0..L-1 that is fewer than 32 L + 64 entries;
// the stack is sized from L of the tree at hand (dynamic LDS).
// Scheduling: the 8 XCDs take workgroups round-robin, so workgroup b works on
// the groups of chunk (b mod 8), consecutive in Morton order, and neighbouring
// groups (which open the same nodes) share an L2.
// List evaluation in pieces: the g bodies of a group would leave 64 - g lanes
// idle if each body took one lane through the whole list (groups average ~42
// bodies). Instead each body's list is cut into q pieces of equal length
// (lam = ceil(m / q), the list padded with zero-mass entries, which add exactly
// zero force), and the g q pieces are spread over P passes of the 64 lanes:
// slot j = 64 p + lane of pass p takes piece j mod q of body j / q (slots past
// g q repeat piece 0 of the last body and are discarded). Every lane of a pass
// runs the same lam iterations, so the loop stays uniform. After the passes
// the partial sums go to LDS and body b adds its pieces b q .. b q + q - 1 in
// that order (deterministic). (P, q) depends only on g and is chosen, for a
// list of ~96 entries, to minimise P (lam + 1) with P <= 3 and q <= 64; g = 64
// gives P = q = 1, g = 42 gives P = 2, q = 3 (98% of the lanes busy instead
// of 66%). Lane use over a 2^20-body tree, from the rsq count: 90.5% (73%
// with a power-of-two number of lanes per body of a group of <= 32).
// The list holds up to kGrpList entries (flushed once it has more than
// kGrpList - 64, since one round adds at most 64) plus the padding (< q <= 64
// entries); the same LDS holds the P 64 partial sums after the passes.
// 256-entry lists, a 4th pass and a smaller stack (more waves per CU) each
// measured within 1%: the kernel is VALU-issue bound.
constexpr int kGrpList = 128;
constexpr int kGrpListPad = 64;  // room for that padding
constexpr int kPieceMaxPasses = 3;
static_assert(kPieceMaxPasses * 64 <= kGrpList + kGrpListPad, "partial sums must fit the list LDS");

__host__ __device__ constexpr int grp_stack_cap(int levels) { return 32 * levels + 64; }

struct PieceTable {
  uint8_t passes[65];
  uint8_t pieces[65];
};

constexpr PieceTable make_piece_table() {
  PieceTable t{};
  for (int g = 1; g <= 64; ++g) {
    int best = 1 << 30, bp = 1, bq = 1;
    for (int p = 1; p <= kPieceMaxPasses; ++p) {
      int q = (64 * p) / g;
      if (q > 64) q = 64;
      const int cost = p * ((96 + q - 1) / q + 1);
      if (cost < best) {
        best = cost;
        bp = p;
        bq = q;
      }
    }
    t.passes[g] = (uint8_t)bp;
    t.pieces[g] = (uint8_t)bq;
  }
  return t;
}

__constant__ PieceTable kPieces = make_piece_table();

// One piece: lam list entries from `p` against the body at `me`.
template <bool MASK>
__device__ __forceinline__ void group_piece_sum(const float4* __restrict__ p, int lam, const float4& me,
                                                float eps2, float& tx, float& ty, float& tz) {
#pragma unroll 4
  for (int e = 0; e < lam; ++e) {
    const float4 q = p[e];
    const float dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
    float r2 = fmaf(dx, dx, eps2);
    r2 = fmaf(dy, dy, r2);
    r2 = fmaf(dz, dz, r2);
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (r2 > 0.f) ? inv : 0.f;
    const float f = (q.w * inv) * (inv * inv);
    tx = fmaf(f, dx, tx);
    ty = fmaf(f, dy, ty);
    tz = fmaf(f, dz, tz);
  }
}

// Post-Newtonian terms of the group traversal (PN): with every list flush each
// body's lane also adds, in FP64, the terms of each list entry (NodePn). The
// list keeps its node ids in LDS for this.
template <bool MASK, bool PN>
__global__ __launch_bounds__(64) void k_bh_group(int n, const int32_t* __restrict__ gstart,
                                                 const uint32_t* __restrict__ gcount, const float4* __restrict__ gsh,
                                                 const float4* __restrict__ kids, const uint32_t* __restrict__ perm,
                                                 float eps2, float msc, double G,
                                                 double* __restrict__ a, unsigned long long* __restrict__ counters,
                                                 uint32_t* __restrict__ flags, int stack_cap, NodePn pnp) {
  extern __shared__ float4 grp_lds[];
  float4* lst = grp_lds;                                              // kGrpList + kGrpListPad entries
  int* stk = reinterpret_cast<int*>(grp_lds + kGrpList + kGrpListPad);  // stack_cap entries
  int* lid = stk + stack_cap;  // PN: node id of each list entry (kGrpList entries)
  // PN: the entries' FP64 node records, gathered once per flush (kGrpList each)
  double4* pnc = reinterpret_cast<double4*>(grp_lds + kGrpList + kGrpListPad + (stack_cap + kGrpList + 3) / 4);
  double4* pnv = pnc + kGrpList;
  const int lane = threadIdx.x;
  const int ngroups = (int)*gcount;
  const int per_xcd = (ngroups + 7) >> 3;
  const int xcd = blockIdx.x & 7;
  const int stride = (int)(gridDim.x >> 3);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  unsigned long long tot_inter = 0, tot_visits = 0;
  for (int i = (int)(blockIdx.x >> 3); i < per_xcd; i += stride) {
    const int g = xcd * per_xcd + i;
    if (g >= ngroups) break;
    const int lo = gstart[g], hi = gstart[g + 1] - 1;
    const int gsz = __builtin_amdgcn_readfirstlane(hi - lo + 1);
    const int k = lo + lane;
    const bool live = k <= hi;
    const float4 me = gsh[live ? k : hi];
    // PN: this body's FP64 position, mass and velocity (leaf k's records)
    double4 pme = make_double4(0.0, 0.0, 0.0, 1.0), vme = make_double4(0.0, 0.0, 0.0, 0.0);
    double pax = 0.0, pay = 0.0, paz = 0.0;
    if (PN) {
      pme = pnp.nmc[live ? k : hi];
      vme = pnp.nvel[live ? k : hi];
    }
    // this lane's piece in each pass
    const int npass = __builtin_amdgcn_readfirstlane((int)kPieces.passes[gsz]);
    const int npc = __builtin_amdgcn_readfirstlane((int)kPieces.pieces[gsz]);
    float4 ppos[kPieceMaxPasses];
    int pidx[kPieceMaxPasses];
#pragma unroll
    for (int p = 0; p < kPieceMaxPasses; ++p) {
      const int j = 64 * p + lane;
      int b = j / npc, pc = j - b * npc;
      if (b >= gsz) {
        b = gsz - 1;
        pc = 0;
      }
      ppos[p] = make_float4(__shfl(me.x, b, 64), __shfl(me.y, b, 64), __shfl(me.z, b, 64), 0.f);
      pidx[p] = pc;
    }
    float blo[3] = {me.x, me.y, me.z}, bhi[3] = {me.x, me.y, me.z};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      for (int off = 32; off > 0; off >>= 1) {
        blo[c] = fminf(blo[c], __shfl_xor(blo[c], off, 64));
        bhi[c] = fmaxf(bhi[c], __shfl_xor(bhi[c], off, 64));
      }
      // every lane holds the same value; saying so keeps the list count and
      // the root test scalar
      blo[c] = uniform_f32(blo[c]);
      bhi[c] = uniform_f32(bhi[c]);
    }
    float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    unsigned long long listed = 0, visits = 1;
    int sp = 0, cnt = 0;
    auto dmin2 = [&](const float4& r) {
      const float ex = fmaxf(fmaxf(blo[0] - r.x, r.x - bhi[0]), 0.f);
      const float ey = fmaxf(fmaxf(blo[1] - r.y, r.y - bhi[1]), 0.f);
      const float ez = fmaxf(fmaxf(blo[2] - r.z, r.z - bhi[2]), 0.f);
      return fmaf(ex, ex, fmaf(ey, ey, ez * ez));
    };
    auto flush = [&]() {
      const int m = __builtin_amdgcn_readfirstlane(cnt);
      const int lam = __builtin_amdgcn_readfirstlane((m + npc - 1) / npc);
      const int mp = npc * lam;  // m <= mp < m + npc <= m + 64
      if (lane < mp - m) lst[m + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
      wave_sync();
      float px[kPieceMaxPasses], py[kPieceMaxPasses], pz[kPieceMaxPasses];
#pragma unroll
      for (int p = 0; p < kPieceMaxPasses; ++p) {
        px[p] = py[p] = pz[p] = 0.f;
        if (p < npass) group_piece_sum<MASK>(lst + pidx[p] * lam, lam, ppos[p], eps2, px[p], py[p], pz[p]);
      }
      wave_sync();  // every lane is done with the list: its LDS takes the partial sums
#pragma unroll
      for (int p = 0; p < kPieceMaxPasses; ++p)
        if (p < npass) lst[64 * p + lane] = make_float4(px[p], py[p], pz[p], 0.f);
      wave_sync();
      if (live) {
        float tx = 0.f, ty = 0.f, tz = 0.f;
        const float4* __restrict__ mine = lst + lane * npc;
        for (int c = 0; c < npc; ++c) {
          const float4 v = mine[c];
          tx += v.x;
          ty += v.y;
          tz += v.z;
        }
        kahan_add32(sx, cx, tx);
        kahan_add32(sy, cy, ty);
        kahan_add32(sz, cz, tz);
      }
      if (PN) {
        // gather the list's node records into LDS (one round per 64 entries),
        // then every body's lane reads them as broadcasts
        for (int e = lane; e < m; e += 64) {
          const int node = lid[e];
          pnc[e] = pnp.nmc[node];
          pnv[e] = pnp.nvel[node];
        }
        wave_sync();
        if (live) {
          const double inv_m = 1.0 / pme.w;
          for (int e = 0; e < m; ++e) node_pn_rec(pnp, pnc[e], pnv[e], pme, vme, inv_m, pax, pay, paz);
        }
      }
      wave_sync();  // the partial sums are read before the list is refilled
      listed += (unsigned long long)cnt;
      cnt = 0;
    };
    {
      const int root = n, rs = 2 * (n - 1);
      const float4 r = kids[2 * rs];
      if (dmin2(r) > kids[2 * rs + 1].y) {
        if (lane == 0) {
          lst[0] = make_float4(r.x, r.y, r.z, r.w * msc);
          if (PN) lid[0] = root;
        }
        cnt = 1;
      } else {
        if (lane == 0) stk[0] = root;
        sp = 1;
      }
    }
    wave_sync();
    // The children of the next batch (the top min(sp, 32) stack entries) are
    // loaded before a pending flush evaluates the list, so the load's latency
    // overlaps that arithmetic. Lane 2 i + s takes child s of entry i: the two
    // children of a node are 64 contiguous bytes. Lanes past the batch load a
    // copy of its last entry and are masked off; every lane loads, so the
    // loop carries one definition of the record and needs no branches.
    int take = 0;
    float4 nr;
    float2 na;
    // With an empty stack (the walk is over) it loads the root's children.
    auto fetch = [&]() {
      take = min(sp, 32);
      const int e = max(min(lane >> 1, take - 1), 0);
      const int opened = sp > 0 ? stk[sp - take + e] : n;
      const int slot = kid_slot(n, opened, lane & 1);
      nr = kids[2 * slot];
      na = *reinterpret_cast<const float2*>(kids + 2 * slot + 1);
    };
    fetch();
    bool overflow = false;
    while (sp > 0) {
      const float4 r = nr;
      const int node = __float_as_int(na.x);
      // both terms evaluated (no short circuit): the id and the threshold
      // arrive with the record in one round trip
      const bool take_it = (node < n) | (dmin2(r) > na.y);
      const bool mine = lane < 2 * take;
      const bool acc = mine & take_it, opn = mine & !take_it;
      sp -= take;
      visits += 2 * take;
      wave_sync();
      const uint64_t am = __ballot(acc), om = __ballot(opn);
      const int nopen = (int)__popcll(om);
      // cannot happen for a valid tree (see bound above); the walk then ends
      // here (no early exit from the loop, which keeps its control flow simple)
      const bool over = sp + nopen > stack_cap;
      if (over) {
        if (lane == 0) atomicOr(flags, (uint32_t)FLAG_TREE);
        overflow = true;
      }
      // list entries carry the scaled mass (one multiply per entry, not per pair)
      if (acc) {
        const int e = cnt + (int)__popcll(am & lt);
        lst[e] = make_float4(r.x, r.y, r.z, r.w * msc);
        if (PN) lid[e] = node;
      }
      if (opn && !over) stk[sp + __popcll(om & lt)] = node;
      cnt += (int)__popcll(am);
      sp = over ? 0 : sp + nopen;
      wave_sync();
      fetch();
      if (cnt > kGrpList - 64) {
        flush();
        wave_sync();
      }
    }
    if (cnt > 0) flush();
    wave_sync();
    if (live) {
      const int b = (int)perm[k];
      a[b] = G * ((double)sx - (double)cx) + pax;
      a[n + b] = G * ((double)sy - (double)cy) + pay;
      a[2 * n + b] = G * ((double)sz - (double)cz) + paz;
    }
    // every body's own leaf is in its group's list (distance 0, no force)
    const unsigned long long nlive = (unsigned long long)(hi - lo + 1);
    tot_inter += listed * nlive - nlive;
    tot_visits += visits;
    if (overflow) break;
  }
  if (lane == 0 && (tot_inter | tot_visits)) {
    atomicAdd(&counters[BHC_INTERACTIONS], tot_inter);
    atomicAdd(&counters[BHC_VISITS], tot_visits);
  }
}

__global__ void k_zero_counters(unsigned long long* p) {
  if (threadIdx.x < BHC_COUNT) p[threadIdx.x] = 0ull;
}

// --------------------------------------------------------------- launch ----

static int reserve_sort(aster_ctx* c) {
  const size_t n = (size_t)c->n;
  const size_t nblocks = (n + kRsTile - 1) / kRsTile;  // the most tiles any pass layout uses
  ASTER_HIP(c, c->codes.reserve(n));
  ASTER_HIP(c, c->idx.reserve(n));
  ASTER_HIP(c, c->codes2.reserve(n));
  ASTER_HIP(c, c->idx2.reserve(n));
  ASTER_HIP(c, c->xm4.reserve(n));
  ASTER_HIP(c, c->hist.reserve(256 * (size_t)rs_stride((int)nblocks)));
  ASTER_HIP(c, c->scanpart.reserve((n + kScanChunk - 1) / kScanChunk + 1));
  return ASTER_OK;
}

// In-place exclusive scan of data[0, count); the total goes to *total when
// given (a device pointer).
static int device_scan(aster_ctx* c, uint32_t* data, int count, uint32_t* total, int32_t* emit = nullptr) {
  const int chunks = (count + kScanChunk - 1) / kScanChunk;
  if (chunks == 0) return ASTER_OK;
  ASTER_HIP(c, c->scanpart.reserve((size_t)chunks));
  hipLaunchKernelGGL(k_scan_reduce, dim3(chunks), dim3(kScanThreads), 0, c->stream, data, count, c->scanpart.p);
  hipLaunchKernelGGL(k_scan_apply, dim3(chunks), dim3(kScanThreads), 0, c->stream, data, count, c->scanpart.p,
                     total, emit);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

static int reserve_tree(aster_ctx* c, bool full) {
  const size_t n = (size_t)c->n, nn = 2 * n - 1;
  int rc = reserve_sort(c);
  if (rc) return rc;
  ASTER_HIP(c, c->left.reserve(nn));
  ASTER_HIP(c, c->right.reserve(nn));
  ASTER_HIP(c, c->parent.reserve(nn));
  ASTER_HIP(c, c->escape.reserve(nn));
  ASTER_HIP(c, c->depth.reserve(nn));
  ASTER_HIP(c, c->levels.reserve(kMaxLevels + 3));
  ASTER_HIP(c, c->bucket.reserve(n));
  ASTER_HIP(c, c->splitnode.reserve(n));
  ASTER_HIP(c, c->groot.reserve(n));
  ASTER_HIP(c, c->climb.reserve(n));
  ASTER_HIP(c, c->topslot.reserve(n));
  ASTER_HIP(c, c->grange.reserve(n));
  ASTER_HIP(c, c->nmc.reserve(nn));
  ASTER_HIP(c, c->nbox.reserve(2 * nn));
  ASTER_HIP(c, c->nrec.reserve(nn));
  if (full) ASTER_HIP(c, c->nlink.reserve(nn));
  else ASTER_HIP(c, c->kids.reserve(2 * nn));  // 2 float4 per child slot; 2 (n - 1) + 1 slots
  ASTER_HIP(c, c->gsh.reserve(n));
  if (!c->bh_counters.p) {
    // Counters accumulate over evaluations until aster_bh_stats reads them.
    ASTER_HIP(c, c->bh_counters.reserve(BHC_COUNT));
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, c->stream, c->bh_counters.p);
    ASTER_HIP(c, hipGetLastError());
  }
  return ASTER_OK;
}

static int morton_codes(aster_ctx* c, bool user_box) {
  phase_begin(c, ASTER_PHASE_MORTON);
  int rc = launch_box(c, user_box);
  if (rc) return rc;
  const int n = (int)c->n;
  hipLaunchKernelGGL(k_morton, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->x.p, c->m.p, c->boxd.p,
                     c->codes.p, c->idx.p, c->xm4.p, c->flags.p);
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_MORTON);
  c->tree_valid = false;
  return ASTER_OK;
}

int bh_morton_only(aster_ctx* c) {
  int rc = reserve_sort(c);
  if (rc) return rc;
  return morton_codes(c, true);
}

// Morton codes of the current state, stably sorted: c->codes holds the sorted
// codes and c->idx the body at each sorted position. user_box = false always
// uses the symmetric box of the data.
int morton_sort(aster_ctx* c, bool user_box) {
  int rc = reserve_sort(c);
  if (rc) return rc;
  if ((rc = morton_codes(c, user_box))) return rc;
  const int n = (int)c->n;
  uint32_t *kin = c->codes.p, *vin = c->idx.p, *kout = c->codes2.p, *vout = c->idx2.p;
  // The smaller tile (4096 or 8192 keys) that needs at most kRsRawMaxBlocks
  // tiles; beyond 2^20 keys, 2048-key tiles and the scan. (2048-key tiles
  // with the offsets formed in the scatter were slower at 2^18 keys: each
  // block then sums twice the count matrix with half the threads.)
  int threads = 0;
  for (int th = 512; th <= 1024 && !threads; th *= 2)
    if ((n + kRsPer * th - 1) / (kRsPer * th) <= kRsRawMaxBlocks) threads = th;
  const bool raw = threads != 0;
  if (!raw) threads = kRsThreads;
  const int tile = kRsPer * threads;
  const int nblocks = (n + tile - 1) / tile;
  const int stride = raw ? rs_stride(nblocks) : nblocks;  // the scan runs over 256 * nblocks contiguous counts
  phase_begin(c, ASTER_PHASE_SORT);
  for (int shift = 0; shift < 32; shift += 8) {
    const dim3 grid(nblocks), block(threads);
    switch (threads) {
      case 256:
        hipLaunchKernelGGL(k_rs_hist<256>, grid, block, 0, c->stream, kin, n, shift, c->hist.p, stride);
        break;
      case 512:
        hipLaunchKernelGGL(k_rs_hist<512>, grid, block, 0, c->stream, kin, n, shift, c->hist.p, stride);
        break;
      default:
        hipLaunchKernelGGL(k_rs_hist<1024>, grid, block, 0, c->stream, kin, n, shift, c->hist.p, stride);
        break;
    }
    if (!raw) {
      if ((rc = device_scan(c, c->hist.p, 256 * nblocks, nullptr))) return rc;
      hipLaunchKernelGGL((k_rs_scatter<kRsThreads, false>), grid, block, 0, c->stream, kin, vin, kout, vout, n, shift,
                         c->hist.p, nblocks, stride);
    } else if (threads == 512) {
      hipLaunchKernelGGL((k_rs_scatter<512, true>), grid, block, 0, c->stream, kin, vin, kout, vout, n, shift,
                         c->hist.p, nblocks, stride);
    } else {
      hipLaunchKernelGGL((k_rs_scatter<1024, true>), grid, block, 0, c->stream, kin, vin, kout, vout, n, shift,
                         c->hist.p, nblocks, stride);
    }
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_SORT);
  // Four passes: sorted codes are back in c->codes, permutation in c->idx.
  return ASTER_OK;
}

// Groups for the wavefront traversal (subtrees of at most kGroup leaves whose
// parent has more), as first sorted positions gstart[0..ngroups] in Morton
// order; the count stays on the device.
static int build_groups(aster_ctx* c) {
  const int n = (int)c->n;
  ASTER_HIP(c, c->gscan.reserve((size_t)n));
  ASTER_HIP(c, c->gstart.reserve((size_t)n + 1));
  ASTER_HIP(c, c->gcount.reserve(1));
  hipLaunchKernelGGL(k_group_flags, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->parent.p, c->grange.p,
                     c->gscan.p, c->climb.p, c->groot.p);
  // the scan of the start flags also writes the group starts and the sentinel n
  int rc = device_scan(c, c->gscan.p, n, c->gcount.p, c->gstart.p);
  if (rc) return rc;
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

constexpr int kBand = 6;  // levels per moments launch above the groups (a team of 2^6 lanes)

static int launch_moments_band(aster_ctx* c, bool vel, int h, int off, int cnt, float4* kids) {
  const int n = (int)c->n;
  if (h < 1 || h > kBand) return fail(c, ASTER_ERR_STATE, "moments band height out of range");
  const int64_t threads = (int64_t)cnt << h;
  if (threads > INT32_MAX) return fail(c, ASTER_ERR_STATE, "moments band too wide");
  const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
  const int32_t* b = c->bucket.p + off;
#define ASTER_BAND(V, H)                                                                                          \
  hipLaunchKernelGGL((k_moments_band<V, H>), grid, block, 0, c->stream, n, b, cnt, c->left.p, c->right.p,        \
                     c->grange.p, c->nmc.p, c->nvel.p, c->nbox.p, c->nrec.p, c->P.theta, kids)
#define ASTER_BAND_H(H)                 \
  case H:                               \
    if (vel) ASTER_BAND(true, H);       \
    else ASTER_BAND(false, H);          \
    break;
  switch (h) {
    ASTER_BAND_H(1)
    ASTER_BAND_H(2)
    ASTER_BAND_H(3)
    ASTER_BAND_H(4)
    ASTER_BAND_H(5)
    ASTER_BAND_H(6)
  }
#undef ASTER_BAND_H
#undef ASTER_BAND
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

// full: also the depth and escape link of every node (the per-body
// traversals and the reference MAC need them; the group traversal does not).
static int build_tree(aster_ctx* c, int mac, bool vel, bool full) {
  const int n = (int)c->n;
  int rc = reserve_tree(c, full);
  if (rc) return rc;
  float4* kids = full ? nullptr : c->kids.p;  // the group traversal's child slots
  if ((rc = morton_sort(c, true))) return rc;
  phase_begin(c, ASTER_PHASE_BUILD);
  if (vel) ASTER_HIP(c, c->nvel.reserve((size_t)2 * n - 1));
  auto build = vel ? k_build<true> : k_build<false>;
  hipLaunchKernelGGL(build, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->codes.p, c->idx.p, c->xm4.p,
                     c->v.p, c->left.p, c->right.p, c->parent.p, c->grange.p, c->nmc.p, c->nvel.p, c->nbox.p,
                     c->nrec.p, c->gsh.p, c->splitnode.p, c->flags.p, full ? 1 : 0);
  ASTER_HIP(c, hipGetLastError());
  // traversal groups (also the per-wavefront subtrees of the moment pass)
  if ((rc = build_groups(c))) return rc;
  phase_end(c, ASTER_PHASE_BUILD);
  const int nn = 2 * n - 1;
  phase_begin(c, ASTER_PHASE_MOMENTS);
  // Depth census of the nodes above the groups: their depths (walks from the
  // group boundaries, and from the group roots for the largest depth) as a
  // compact (depth, node) list, then two passes list them level by level
  // (depth < 128). The sort scratch of the Morton pass holds the list.
  const int cblocks = (int)std::min<int64_t>((2 * (int64_t)n / 32 + 2 + 255) / 256, 4096);
  ASTER_HIP(c, c->dmax.reserve((size_t)cblocks));
  ASTER_HIP(c, c->censusctr.reserve(2 * kMaxLevels));
  hipLaunchKernelGGL(k_above_census, dim3(cblocks), dim3(256), 0, c->stream, n, c->gstart.p, c->gcount.p,
                     c->splitnode.p, c->groot.p, c->climb.p, c->parent.p, c->depth.p, c->codes2.p, c->idx2.p,
                     c->dmax.p, c->censusctr.p, c->flags.p);
  const int ctiles = (n + kRsTile - 1) / kRsTile;  // the list has fewer than n entries
  hipLaunchKernelGGL(k_census_count, dim3(ctiles), dim3(kRsThreads), 0, c->stream, c->gcount.p, c->codes2.p,
                     c->censusctr.p);
  hipLaunchKernelGGL(k_census_place, dim3(ctiles), dim3(kRsThreads), 0, c->stream, c->gcount.p, c->codes2.p,
                     c->idx2.p, c->censusctr.p, c->dmax.p, cblocks, c->bucket.p, c->levels.p);
  if (full)
    hipLaunchKernelGGL(k_depth, dim3((nn + 255) / 256), dim3(256), 0, c->stream, n, c->left.p, c->right.p,
                       c->parent.p, c->depth.p, c->escape.p, c->flags.p);
  ASTER_HIP(c, hipGetLastError());
  // The level bounds and the largest depth go to pinned host memory
  // asynchronously; the host waits for them only after the group kernel below
  // is queued, so the device keeps working meanwhile.
  if (!c->host_words) {
    ASTER_HIP(c, hipHostMalloc((void**)&c->host_words, (kMaxLevels + 3) * sizeof(int32_t), hipHostMallocDefault));
    ASTER_HIP(c, hipEventCreateWithFlags(&c->tree_ev, hipEventDisableTiming));
  }
  ASTER_HIP(c, hipMemcpyAsync(c->host_words, c->levels.p, (kMaxLevels + 3) * sizeof(int32_t), hipMemcpyDeviceToHost,
                              c->stream));
  ASTER_HIP(c, hipEventRecord(c->tree_ev, c->stream));
  // every node inside a group: one wavefront per group, bottom-up in LDS
  const int gblocks = ((n / 32 + 1) + 3) / 4;
  auto gmom = vel ? k_group_moments<true> : k_group_moments<false>;
  hipLaunchKernelGGL(gmom, dim3(gblocks), dim3(256), 0, c->stream, n, c->gstart.p, c->gcount.p, c->grange.p, c->left.p,
                     c->right.p, c->nmc.p, c->nvel.p, c->nbox.p, c->nrec.p, c->P.theta, kids);
  ASTER_HIP(c, hipGetLastError());
  ASTER_HIP(c, hipEventSynchronize(c->tree_ev));
  const int32_t* lv = c->host_words;
  const int nlev = lv[kMaxLevels + 1];  // depths above the groups
  const int maxdep = lv[kMaxLevels + 2];
  if (nlev < 0 || nlev >= kMaxLevels || maxdep < nlev - 1 || maxdep >= kMaxLevels - 1 || (nlev > 0 && lv[0] != 0))
    return fail(c, ASTER_ERR_STATE, "radix tree depth census is inconsistent");
  c->tree_levels = maxdep + 1;
  // the nodes above the groups: the top levels holding at most kTopNodes
  // nodes in one workgroup, the rest in bands of up to kBand levels, deepest
  // band first (the children of a band's deepest nodes are group roots,
  // leaves or nodes of the band below)
  int ntop = 0;
  while (ntop < nlev && ntop < kTopMaxLevels && lv[ntop + 1] <= kTopNodes) ++ntop;
  for (int d = nlev - 1; d >= ntop;) {
    const int top = d - kBand + 1 > ntop ? d - kBand + 1 : ntop;
    const int h = d - top + 1;
    const int off = lv[top], cnt = lv[top + 1] - off;
    if (cnt <= 0) return fail(c, ASTER_ERR_STATE, "radix tree depth census has an empty level");
    if ((rc = launch_moments_band(c, vel, h, off, cnt, kids))) return rc;
    d = top - 1;
  }
  if (ntop > 0) {
    TopLevels tl;
    tl.nlev = ntop;
    for (int d = 0; d <= ntop; ++d) tl.off[d] = lv[d];
    auto topk = vel ? k_moments_top<true> : k_moments_top<false>;
    hipLaunchKernelGGL(topk, dim3(1), dim3(kTopNodes), 0, c->stream, n, c->bucket.p, tl, c->left.p, c->right.p,
                       c->grange.p, c->nmc.p, c->nvel.p, c->nbox.p, c->nrec.p, c->P.theta, kids, c->topslot.p);
  }
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_MOMENTS);
  phase_begin(c, ASTER_PHASE_LINKS);
  if (full) {  // the group traversal's child slots were written with the moments
    hipLaunchKernelGGL(k_links, dim3((nn + 255) / 256), dim3(256), 0, c->stream, n, c->right.p, c->escape.p,
                       c->depth.p, c->nbox.p, c->nrec.p, c->boxd.p, c->P.theta, mac, c->nlink.p, c->idx.p);
    ASTER_HIP(c, hipGetLastError());
  }
  phase_end(c, ASTER_PHASE_LINKS);
  c->tree_valid = true;
  c->tree_mac = mac;
  return ASTER_OK;
}

// Barnes-Hut over every body of the context's state.
static int bh_forces_all(aster_ctx* c, int method) {
  const int n = (int)c->n;
  const int pnf = c->P.pn_flags & ASTER_PN_TERMS;
  const bool pn = pnf != 0;
  const bool f64 = method == ASTER_BH_PARITY && c->P.precision == ASTER_FP64;
  if (n == 1) {
    ASTER_HIP(c, hipMemsetAsync(c->a.p, 0, 3 * sizeof(double), c->stream));
    return ASTER_OK;
  }
  const int mac = (method == ASTER_BH_PARITY) ? 0 : 1;
  Fp32Scales sc{};
  int rc;
  if (!f64 && (rc = fp32_scales(c, &sc))) return rc;
  if ((rc = build_tree(c, mac, pn, method != ASTER_BH_GROUP))) return rc;
  phase_begin(c, ASTER_PHASE_TRAVERSE);
  const double ic = 1.0 / c->P.c;
  if (f64) {
    if (pn)
      hipLaunchKernelGGL((k_bh_f64<true>), dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->m.p,
                         c->nmc.p, c->nvel.p, c->nlink.p, c->depth.p, c->boxd.p, c->P.theta, c->P.G, ic, pnf,
                         c->P.softening, c->a.p, c->bh_counters.p, c->flags.p);
    else
      hipLaunchKernelGGL((k_bh_f64<false>), dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->m.p,
                         c->nmc.p, c->nvel.p, c->nlink.p, c->depth.p, c->boxd.p, c->P.theta, c->P.G, ic, pnf,
                         c->P.softening, c->a.p, c->bh_counters.p, c->flags.p);
  } else {
    const float eps2 = sc.eps2;
    const bool mask = sc.mask;
    if (method == ASTER_BH_GROUP) {
      // One workgroup per group while groups average more than 32 bodies
      // (the usual case); workgroups loop when there are more groups. The
      // count is a multiple of 8 for the XCD-interleaved group order.
      const int blocks = ((n / 32 + 1) + 7) & ~7;
      const dim3 grid(blocks), block(64);
      const int cap = grp_stack_cap(c->tree_levels);
      // PN: node ids, then (16-byte aligned) FP64 records and velocities of the list
      const size_t lds = (kGrpList + kGrpListPad) * sizeof(float4) +
                         (pn ? (size_t)(cap + kGrpList + 3) / 4 * sizeof(float4) + 2 * kGrpList * sizeof(double4)
                             : (size_t)cap * sizeof(int));
      const NodePn gp{c->nmc.p, c->nvel.p, c->P.G, ic, c->P.softening, pnf};
      auto kern = pn ? (mask ? k_bh_group<true, true> : k_bh_group<false, true>)
                     : (mask ? k_bh_group<true, false> : k_bh_group<false, false>);
      hipLaunchKernelGGL(kern, grid, block, lds, c->stream, n, c->gstart.p, c->gcount.p, c->gsh.p, c->kids.p,
                         c->idx.p, eps2, sc.msc, sc.gsc, c->a.p, c->bh_counters.p, c->flags.p, cap, gp);
    } else {
      const NodePn gp{c->nmc.p, c->nvel.p, c->P.G, ic, c->P.softening, pnf};
      auto kern = pn ? (mask ? k_bh_f32<true, true> : k_bh_f32<false, true>)
                     : (mask ? k_bh_f32<true, false> : k_bh_f32<false, false>);
      hipLaunchKernelGGL(kern, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->gsh.p, c->nrec.p, c->nlink.p,
                         eps2, sc.msc, sc.gsc, c->a.p, c->bh_counters.p, c->flags.p, gp);
    }
  }
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_TRAVERSE);
  return ASTER_OK;
}

// Solver::set_bounds on Barnes-Hut (Barnes_Hut::make_tree and compute_forces,
// src/solvers/barnes_hut.cpp:309-322,414-417): the tree holds the bodies
// [lo, hi) only, and only those bodies get new accelerations; the others keep
// theirs. (The reference indexes the leaves from 0 instead of lower and takes
// the root at the full body count, so its own bounded tree is only right for
// the full range; this follows what the bounds mean.) The range is copied to
// contiguous sub-state planes, evaluated as a state of hi - lo bodies, and
// the accelerations copied back.
static int bh_forces_range(aster_ctx* c, int method, int lo, int hi) {
  const int64_t n = c->n, r = (int64_t)hi - lo;
  ASTER_HIP(c, c->sub_x.reserve((size_t)3 * r));
  ASTER_HIP(c, c->sub_v.reserve((size_t)3 * r));
  ASTER_HIP(c, c->sub_a.reserve((size_t)3 * r));
  ASTER_HIP(c, c->sub_m.reserve((size_t)r));
  for (int q = 0; q < 3; ++q) {
    ASTER_HIP(c, hipMemcpyAsync(c->sub_x.p + q * r, c->x.p + q * n + lo, r * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
    ASTER_HIP(c, hipMemcpyAsync(c->sub_v.p + q * r, c->v.p + q * n + lo, r * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
  }
  ASTER_HIP(c, hipMemcpyAsync(c->sub_m.p, c->m.p + lo, r * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  // evaluate on the sub-state, then put the full state back whatever happened
  auto swap_state = [c]() {
    std::swap(c->x, c->sub_x);
    std::swap(c->v, c->sub_v);
    std::swap(c->a, c->sub_a);
    std::swap(c->m, c->sub_m);
  };
  const double mt = c->mass_total;
  const bool mt_valid = c->mass_total_valid;
  swap_state();
  c->n = r;
  c->npad = round_up(r, kApBlock);
  c->mass_total_valid = false;
  int rc = bh_forces_all(c, method);
  swap_state();
  c->n = n;
  c->npad = round_up(n, kApBlock);
  c->mass_total = mt;
  c->mass_total_valid = mt_valid;
  c->tree_valid = false;  // the tree holds the sub-state's bodies
  if (rc) return rc;
  for (int q = 0; q < 3; ++q)
    ASTER_HIP(c, hipMemcpyAsync(c->a.p + q * n + lo, c->sub_a.p + q * r, r * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
  return ASTER_OK;
}

int bh_forces(aster_ctx* c, int method) {
  const int n = (int)c->n;
  const int pnf = c->P.pn_flags & ASTER_PN_TERMS;
  const bool f64 = method == ASTER_BH_PARITY && c->P.precision == ASTER_FP64;
  if (pnf & ASTER_EIH) return fail(c, ASTER_ERR_INVALID, "ASTER_EIH needs ASTER_ALLPAIRS_FP64");
  if (!(c->P.theta >= 0.0)) return fail(c, ASTER_ERR_INVALID, "theta must be >= 0");
  const int lo = c->P.lower < 0 ? 0 : c->P.lower;
  const int hi = (c->P.upper < 0 || c->P.upper > n) ? n : c->P.upper;
  if (lo > hi) return fail(c, ASTER_ERR_INVALID, "body bounds: lower > upper");
  if (lo == 0 && hi == n) return bh_forces_all(c, method);
  if (lo == hi) return ASTER_OK;  // no bodies in range: nothing is evaluated (the reference warns and returns)
  return bh_forces_range(c, method, lo, hi);
}

}  // namespace aster
