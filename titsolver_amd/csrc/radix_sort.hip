// Stable LSD radix sort of (cell key, particle slot) pairs, 8- or 11-bit
// digits (two passes of 11 bits for keys of 17 to 22 bits, the cell keys of a
// 2M-particle tank, instead of three of 8).
//
// The sorted pairs end in the input arrays (an odd pass count copies them
// back), so no buffer changes roles and a captured step can hold the sort
// whether or not it runs (every kernel returns at once while ctl[0] is 0).
//
// Each pass is three launches:
//   hist    one workgroup per 4096-key tile counts its digits in LDS;
//   scan    one workgroup per digit turns the digit-major tile counts into
//           exclusive tile offsets and the digit total;
//   scatter each tile ranks its keys stably in LDS and writes them out.
// Ranks inside a wave come from matching digits with one ballot per digit
// bit (64-lane masks), so equal digits of a nearly sorted key stream cost one LDS update
// per wave instead of one atomic per lane.  Within a tile the order is wave,
// then round, then lane, which is the input order: the sort is stable, so
// particles of one cell keep their previous relative order and the result is
// deterministic.
#include "internal.hpp"

namespace tit_amd {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kWaveSpan = 64 * kItems;

// The lanes of the wave (among `valid`) whose digit equals this lane's.
template <int BITS>
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t peers = valid;
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

__device__ __forceinline__ uint64_t lanemask_lt(int lane) {
  return (1ull << lane) - 1ull;
}

template <int BITS>
__global__ __launch_bounds__(kThreads) void k_radix_hist(
    const uint32_t* __restrict__ keys, uint32_t n, int shift,
    uint32_t* __restrict__ hist, uint32_t ntiles, const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  constexpr uint32_t R = 1u << BITS;
  __shared__ uint32_t h[R];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t d = tid; d < R; d += kThreads) h[d] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTile + wave * kWaveSpan;
  uint32_t key[kItems];  // all loads issued before the first match
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * 64 + lane;
    key[r] = i < n ? keys[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * 64 + lane;
    const bool ok = i < n;
    const uint64_t valid = __ballot(ok);
    if (valid == 0) break;  // wave-uniform
    const uint32_t d = (key[r] >> shift) & (R - 1u);
    const uint64_t peers = match_digit<BITS>(d, valid);
    if (ok && (peers & lanemask_lt(lane)) == 0) {
      atomicAdd(&h[d], static_cast<uint32_t>(__popcll(peers)));
    }
  }
  __syncthreads();
  for (uint32_t d = tid; d < R; d += kThreads) hist[d * ntiles + blockIdx.x] = h[d];
}

__device__ __forceinline__ uint32_t block_inclusive_scan256(uint32_t v,
                                                            uint32_t* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t o = tid >= off ? sh[tid - off] : 0u;
    __syncthreads();
    sh[tid] += o;
    __syncthreads();
  }
  return sh[tid];
}

// One workgroup per digit.
__global__ __launch_bounds__(256) void k_radix_digit_scan(
    uint32_t* __restrict__ hist, uint32_t ntiles, uint32_t* __restrict__ tot,
    const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t sh[256];
  uint32_t* row = hist + static_cast<size_t>(blockIdx.x) * ntiles;
  const int tid = threadIdx.x;
  const uint32_t per = (ntiles + 255u) / 256u;
  const uint32_t b0 = min(tid * per, ntiles), b1 = min(b0 + per, ntiles);
  uint32_t s = 0;
  for (uint32_t k = b0; k < b1; ++k) s += row[k];
  const uint32_t incl = block_inclusive_scan256(s, sh);
  uint32_t excl = incl - s;
  for (uint32_t k = b0; k < b1; ++k) {
    const uint32_t t = row[k];
    row[k] = excl;
    excl += t;
  }
  if (tid == 255) tot[blockIdx.x] = incl;
}

template <int BITS>
__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    uint32_t n, int shift, const uint32_t* __restrict__ hist,
    const uint32_t* __restrict__ tot, uint32_t ntiles, const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  constexpr uint32_t R = 1u << BITS;
  constexpr uint32_t kPer = R / kThreads;  // digits per thread in the digit scans
  __shared__ uint32_t wcnt[kWaves][R];
  __shared__ uint32_t sh[256];
  __shared__ uint32_t tile_base[R];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint32_t d = tid; d < R; d += kThreads) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][d] = 0;
  }
  // global exclusive digit offsets; thread tid owns digits [tid kPer, (tid + 1) kPer)
  {
    uint32_t t_d[kPer], s = 0;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      t_d[q] = tot[tid * kPer + q];
      s += t_d[q];
    }
    uint32_t ex = block_inclusive_scan256(s, sh) - s;
#pragma unroll
    for (uint32_t q = 0; q < kPer; ++q) {
      const uint32_t d = tid * kPer + q;
      tile_base[d] = ex + hist[d * ntiles + blockIdx.x];
      ex += t_d[q];
    }
  }
  __syncthreads();

  uint32_t key[kItems], val[kItems], rank[kItems];
  const uint32_t base = blockIdx.x * kTile + wave * kWaveSpan;
#pragma unroll
  for (int r = 0; r < kItems; ++r) {  // all loads issued before the first match
    const uint32_t i = base + r * 64 + lane;
    key[r] = i < n ? keys_in[i] : 0u;
    val[r] = i < n ? (vals_in ? vals_in[i] : i) : 0u;
  }
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * 64 + lane;
    const bool ok = i < n;
    const uint64_t valid = __ballot(ok);
    const uint32_t k = key[r];
    const uint32_t d = (k >> shift) & (R - 1u);
    const uint64_t peers = match_digit<BITS>(d, valid);
    const uint64_t below = peers & lanemask_lt(lane);
    const uint32_t prior = wcnt[wave][d];
    rank[r] = prior + static_cast<uint32_t>(__popcll(below));
    if (ok && below == 0) {
      wcnt[wave][d] = prior + static_cast<uint32_t>(__popcll(peers));
    }
  }
  __syncthreads();
  for (uint32_t d = tid; d < R; d += kThreads) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t t = wcnt[w][d];
      wcnt[w][d] = s;
      s += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const uint32_t i = base + r * 64 + lane;
    if (i < n) {
      const uint32_t d = (key[r] >> shift) & (R - 1u);
      const uint32_t pos = tile_base[d] + wcnt[wave][d] + rank[r];
      keys_out[pos] = key[r];
      vals_out[pos] = val[r];
    }
  }
}

__global__ __launch_bounds__(kThreads) void k_copy_pairs(
    uint32_t n, const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va,
    uint32_t* __restrict__ kb, uint32_t* __restrict__ vb, const uint32_t* __restrict__ ctl) {
  if (!ctl[0]) return;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    kb[i] = ka[i];
    vb[i] = va[i];
  }
}

}  // namespace

uint32_t radix_sort_tiles(uint32_t n) {
  return (n + kTile - 1) / kTile;
}

void radix_sort_pairs(hipStream_t stream, Profiler& prof, uint32_t n,
                      int key_bits, uint32_t* keys, uint32_t* keys_alt,
                      uint32_t* vals, uint32_t* vals_alt,
                      const SortScratch& scratch, const uint32_t* ctl) {
  if (n == 0) return;
  const uint32_t ntiles = radix_sort_tiles(n);
  // digit width: two passes of 11 bits where they cover the key, else 8
  const int bits = (key_bits > 16 && key_bits <= 22) || key_bits > 24 ? 11 : 8;
  int passes = (key_bits + bits - 1) / bits;
  if (passes < 1) passes = 1;
  uint32_t *kin = keys, *kout = keys_alt, *vin = vals, *vout = vals_alt;
  for (int p = 0; p < passes; ++p) {
    const int shift = bits * p;
    const uint32_t* vsrc = p == 0 ? nullptr : vin;
    {
      ProfScope ps(prof, stream, "radix_hist");
      if (bits == 11) {
        k_radix_hist<11><<<ntiles, kThreads, 0, stream>>>(kin, n, shift, scratch.hist, ntiles,
                                                           ctl);
      } else {
        k_radix_hist<8><<<ntiles, kThreads, 0, stream>>>(kin, n, shift, scratch.hist, ntiles,
                                                          ctl);
      }
    }
    {
      ProfScope ps(prof, stream, "radix_digit_scan");
      k_radix_digit_scan<<<1u << bits, 256, 0, stream>>>(scratch.hist, ntiles, scratch.tot,
                                                          ctl);
    }
    {
      ProfScope ps(prof, stream, "radix_scatter");
      if (bits == 11) {
        k_radix_scatter<11><<<ntiles, kThreads, 0, stream>>>(kin, vsrc, kout, vout, n, shift,
                                                              scratch.hist, scratch.tot,
                                                              ntiles, ctl);
      } else {
        k_radix_scatter<8><<<ntiles, kThreads, 0, stream>>>(kin, vsrc, kout, vout, n, shift,
                                                             scratch.hist, scratch.tot,
                                                             ntiles, ctl);
      }
    }
    uint32_t* t = kin;
    kin = kout;
    kout = t;
    t = vin;
    vin = vout;
    vout = t;
  }
  if (passes & 1) {  // the result is in the alternate arrays: copy it back
    ProfScope ps(prof, stream, "radix_copy");
    k_copy_pairs<<<(n + kThreads - 1) / kThreads, kThreads, 0, stream>>>(n, kin, vin, keys,
                                                                        vals, ctl);
  }
}

}  // namespace tit_amd
