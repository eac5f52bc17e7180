// Fused SSPRK stage kernel: continuity + momentum pair sums, EOS, update.
//
// Work unit: a chunk of at most chunk_targets consecutive (cell-sorted)
// particles of one cell column of one tank, one particle per lane.  After
// every cell rebuild the chunks are recut on the device so that each column's
// particles split into equal chunks; a persistent grid then walks the chunk
// list in XCD-major bands (launch index b runs on XCD b % 8 and takes chunks
// from that XCD's eighth of the list, so each XCD's L2 holds a contiguous
// band of columns).
//
// 2D: the FP64 positions of the chunk's candidates (the 5 neighbouring
// columns, the chunk's rows +- 2) are staged in LDS.  Each lane walks its own
// 5 x 5 cell stencil one column range at a time: the exact FP64 predicate
// (four LDS loads per batch) builds a 64-bit acceptance mask per batch, queued
// in registers; then the FP32 pair terms run over the accepted candidates of all
// batches in one loop, reading their pair records (velocity, m/rho,
// rho - rho0 and p/rho^2, cs) from L2 one candidate ahead of the arithmetic.
// Chunks whose candidates exceed the LDS capacity, and 3D, take positions
// from global memory too.  Every path visits candidates in the same order and
// rounds identically, so all schedules produce the same bits.
#include "stage_common.hpp"

namespace tit_amd {

namespace {

// 2D chunks with LDS-staged positions; chunks that do not fit go to the
// overflow list.  18 KB of LDS and at most 128 VGPRs: four waves per SIMD.
template <int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT, int CAP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_stage_lds(
    StageArgs A) {
  __shared__ double2 s_xy[CAP];
  __shared__ uint32_t s_coff[kCols2][kRowsMax + 1];
  __shared__ uint32_t s_gb[kCols2], s_lb[kCols2 + 1];
  __shared__ int s_fits;
  __shared__ uint32_t s_gbase;
  const int tid = threadIdx.x;
  const Band band = xcd_band(A.counts[0]);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const ChunkInfo c = chunk_info<2>(A, q);
    const int zf = static_cast<int>(A.key[c.t0] - c.col_row);
    const int zl = static_cast<int>(A.key[c.t1 - 1] - c.col_row);
    const int zlo = max(zf - kStencil, 0), zhi = min(zl + kStencil, c.nz - 1);
    const int nrows = zhi - zlo + 1;
    __syncthreads();  // the previous chunk's LDS reads are done
    const bool rows_fit = nrows <= kRowsMax;
    if (rows_fit) {
      for (int r = tid; r < kCols2 * (nrows + 1); r += NT) {
        const int k = r / (nrows + 1), row = r % (nrows + 1);
        const int ixk = c.cx - kStencil + k;
        s_coff[k][row] =
            (ixk >= 0 && ixk < c.T.nc[0])
                ? A.cell_off[c.T.cell_off + static_cast<uint32_t>(ixk) * c.nz + zlo + row]
                : 0u;
      }
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t lb = 0, g0 = ~0u, g1 = 0;
      if (rows_fit) {
        for (int k = 0; k < kCols2; ++k) {
          s_gb[k] = s_coff[k][0];
          s_lb[k] = lb;
          lb += s_coff[k][nrows] - s_coff[k][0];
          const int ixk = c.cx - kStencil + k;
          if (ixk >= 0 && ixk < c.T.nc[0]) {
            g0 = min(g0, s_coff[k][0]);
            g1 = max(g1, s_coff[k][nrows]);
          }
        }
        s_lb[kCols2] = lb;
        s_gbase = g0;
      }
      s_fits = rows_fit && lb <= A.lds_cap && g1 - g0 < kWordSpan;
      if (!s_fits) A.overflow[atomicAdd(A.counts + 1, 1u)] = q;
    }
    __syncthreads();
    if (!s_fits) continue;
    const uint32_t total = s_lb[kCols2];
    for (uint32_t l = tid; l < total; l += NT) {
      int k = 0;
      while (l >= s_lb[k + 1]) ++k;
      const uint32_t j = s_gb[k] + (l - s_lb[k]);
      s_xy[l] = make_double2(A.in.x[0][j], A.in.x[1][j]);
    }
    __syncthreads();
    const LdsPos ps{s_xy, s_coff, s_gb, s_lb, zlo, c.nz, c.cx, c.T.nc[0], s_gbase};
    chunk_targets<2, KERNEL, FLAVOUR, KIND, FINAL>(A, c, ps, red);
  }
  red.flush(A);
}

// Chunks with positions from global memory: 3D, and 2D overflow chunks.
template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_stage_global(
    StageArgs A, const uint32_t* list, const uint32_t* count) {
  __shared__ uint32_t s_wm[NT / 64][kWordsG * 64];
  __shared__ uint32_t s_wj[NT / 64][kWordsG * 64];
  const Band band = xcd_band(*count);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const ChunkInfo c = chunk_info<DIM>(A, list ? list[q] : q);
    const GlobalPos<DIM> ps{c.T,  A.in, A.cell_off, c.cx, c.cy, s_wm[threadIdx.x >> 6],
                            s_wj[threadIdx.x >> 6]};
    chunk_targets<DIM, KERNEL, FLAVOUR, KIND, FINAL>(A, c, ps, red);
  }
  red.flush(A);
}

// ---- 3D block chunks ------------------------------------------------------------

struct BlockChunk {
  uint32_t tank;
  DevTank T;
  int bx, by;    // first cell column of the block
  int zs, ze;    // the chunk's rows [zs, ze)
  uint32_t n;    // targets
};

__device__ __forceinline__ BlockChunk block_chunk(const StageArgs& A, uint32_t q) {
  const uint4 ch = A.bchunks[q];
  const uint4 blk = A.blocks[ch.x];
  BlockChunk c;
  c.tank = blk.x;
  c.T = A.tanks[c.tank];
  c.bx = static_cast<int>(blk.y);
  c.by = static_cast<int>(blk.z);
  c.zs = static_cast<int>(ch.y);
  c.ze = static_cast<int>(ch.z);
  c.n = ch.w;
  return c;
}

// Global index of the first cell of column (ix, iy) of tank T.
__device__ __forceinline__ uint32_t column_row3(const DevTank& T, int ix, int iy) {
  return T.cell_off + (static_cast<uint32_t>(ix) * T.nc[1] + static_cast<uint32_t>(iy)) *
                          static_cast<uint32_t>(T.nc[2]);
}

// Target t of a block chunk (s_tb: prefix of the targets over the block's
// kBlk3 x kBlk3 columns): its particle index and column.
__device__ __forceinline__ void block_target(const BlockChunk& c, const uint32_t* s_tb,
                                             const uint32_t* s_first, uint32_t t, uint32_t& i,
                                             int& cx, int& cy) {
  int q = 0;
  while (q + 1 < kBlk3 * kBlk3 && t >= s_tb[q + 1]) ++q;
  i = s_first[q] + (t - s_tb[q]);
  cx = c.bx + q / kBlk3;
  cy = c.by + q % kBlk3;
}

// 3D chunks with LDS-staged positions: the block's window (kWin3 x kWin3
// columns, the chunk's rows +- 2) in LDS, one target per thread.  Chunks whose
// window has more rows or candidates than the LDS holds, or more targets
// than threads, go to the overflow list.
template <int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT, int CAP>
__global__ __launch_bounds__(NT) void k_stage_lds3(StageArgs A) {
  constexpr int NW = kWin3 * kWin3;
  constexpr int NB = kBlk3 * kBlk3;
  __shared__ double s_x[CAP], s_y[CAP], s_z[CAP];
  __shared__ uint32_t s_coff[NW][kRows3 + 1];
  __shared__ uint32_t s_lb[NW + 1], s_delta[NW];
  __shared__ uint32_t s_tb[NB + 1], s_first[NB];
  __shared__ int s_fits;
  const int tid = threadIdx.x;
  const Band band = xcd_band(A.counts[2]);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const BlockChunk c = block_chunk(A, q);
    const int nz = c.T.nc[2];
    const int zlo = max(c.zs - kStencil, 0), zhi = min(c.ze - 1 + kStencil, nz - 1);
    const int nrows = zhi - zlo + 1;
    const int wx0 = c.bx - kStencil, wy0 = c.by - kStencil;
    __syncthreads();  // the previous chunk's LDS reads are done
    const bool rows_fit = nrows <= kRows3;
    if (rows_fit) {
      for (int r = tid; r < NW * (nrows + 1); r += NT) {
        const int kw = r / (nrows + 1), row = r % (nrows + 1);
        const int ix = wx0 + kw / kWin3, iy = wy0 + kw % kWin3;
        const bool in = ix >= 0 && ix < c.T.nc[0] && iy >= 0 && iy < c.T.nc[1];
        s_coff[kw][row] = in ? A.cell_off[column_row3(c.T, ix, iy) + zlo + row] : 0u;
      }
    }
    __syncthreads();
    static_assert(NW == 64, "one lane of the first wave per window column");
    if (tid < 64) {
      // wave 0: LDS offsets of the window columns (lane = column) and the
      // targets of the block's columns (lanes 0..NB-1), by shuffle scans
      uint32_t len = 0, nb = 0, first = 0;
      if (rows_fit) {
        len = s_coff[tid][nrows] - s_coff[tid][0];
        if (tid < NB) {
          const int kw = (kStencil + tid / kBlk3) * kWin3 + kStencil + tid % kBlk3;
          first = s_coff[kw][c.zs - zlo];
          nb = s_coff[kw][c.ze - zlo] - first;
        }
      }
      uint32_t sl = len, sb = nb;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t yl = __shfl_up(sl, off), yb = __shfl_up(sb, off);
        if (tid >= off) {
          sl += yl;
          sb += yb;
        }
      }
      s_lb[tid] = sl - len;
      s_delta[tid] = s_coff[tid][0] - (sl - len);
      if (tid < NB) {
        s_tb[tid] = sb - nb;
        s_first[tid] = first;
      }
      const uint32_t lb = __shfl(sl, 63), nt = __shfl(sb, 63);
      if (tid == 0) {
        s_lb[NW] = lb;
        s_tb[NB] = nt;
        s_fits = rows_fit && lb <= A.lds_cap && nt <= static_cast<uint32_t>(NT);
        if (!s_fits) A.boverflow[atomicAdd(A.counts + 3, 1u)] = q;
      }
    }
    __syncthreads();
    if (!s_fits) continue;
    // each window column's candidates by 32 lanes (a column holds about 30)
    for (int r = tid; r < NW * 32; r += NT) {
      const int kw = r >> 5;
      const uint32_t e = s_lb[kw + 1], d = s_delta[kw];
      for (uint32_t l = s_lb[kw] + (r & 31); l < e; l += 32) {
        const uint32_t j = l + d;
        s_x[l] = A.in.x[0][j];
        s_y[l] = A.in.x[1][j];
        s_z[l] = A.in.x[2][j];
      }
    }
    __syncthreads();
    const LdsPos3 ps{s_x, s_y, s_z, s_coff, s_lb, s_delta, zlo, nz, wx0, wy0,
                     c.T.nc[0], c.T.nc[1]};
    const uint32_t t = static_cast<uint32_t>(tid);
    const bool in_range = t < s_tb[NB];
    uint32_t i = 0;
    int cx = c.bx, cy = c.by;
    if (in_range) block_target(c, s_tb, s_first, t, i, cx, cy);
    float vmax;
    unsigned long long dtmin;
    target_step<3, KERNEL, FLAVOUR, KIND, FINAL>(A, c.T, c.tank, cx, cy,
                                                 column_row3(c.T, cx, cy), i, in_range, ps,
                                                 vmax, dtmin);
    reduce_chunk<KIND, FINAL>(A, c.tank, vmax, dtmin, red);
  }
  red.flush(A);
}

// The block chunks that did not fit, positions from global memory, in rounds
// of NT targets.
template <int KERNEL, int FLAVOUR, int KIND, bool FINAL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_stage_block_global(
    StageArgs A) {
  constexpr int NB = kBlk3 * kBlk3;
  __shared__ uint32_t s_wm[NT / 64][kWordsG * 64];
  __shared__ uint32_t s_wj[NT / 64][kWordsG * 64];
  __shared__ uint32_t s_tb[NB + 1], s_first[NB];
  const Band band = xcd_band(A.counts[3]);
  WaveReduce red;
  for (uint32_t q = band.lo; q < band.hi; q += band.step) {
    const BlockChunk c = block_chunk(A, A.boverflow[q]);
    __syncthreads();  // the previous chunk's table reads are done
    if (threadIdx.x == 0) {
      uint32_t nt = 0;
      for (int b = 0; b < NB; ++b) {
        const int ix = c.bx + b / kBlk3, iy = c.by + b % kBlk3;
        uint32_t f = 0, e = 0;
        if (ix < c.T.nc[0] && iy < c.T.nc[1]) {
          const uint32_t row = column_row3(c.T, ix, iy);
          f = A.cell_off[row + c.zs];
          e = A.cell_off[row + c.ze];
        }
        s_tb[b] = nt;
        s_first[b] = f;
        nt += e - f;
      }
      s_tb[NB] = nt;
    }
    __syncthreads();
    for (uint32_t base = 0; base < s_tb[NB]; base += NT) {
      const uint32_t t = base + threadIdx.x;
      const bool in_range = t < s_tb[NB];
      uint32_t i = 0;
      int cx = c.bx, cy = c.by;
      if (in_range) block_target(c, s_tb, s_first, t, i, cx, cy);
      const GlobalPos<3> ps{c.T, A.in, A.cell_off, cx, cy, s_wm[threadIdx.x >> 6],
                            s_wj[threadIdx.x >> 6]};
      float vmax;
      unsigned long long dtmin;
      target_step<3, KERNEL, FLAVOUR, KIND, FINAL>(A, c.T, c.tank, cx, cy,
                                                   column_row3(c.T, cx, cy), i, in_range, ps,
                                                   vmax, dtmin);
      reduce_chunk<KIND, FINAL>(A, c.tank, vmax, dtmin, red);
    }
  }
  red.flush(A);
}

// Block chunks of one block (one 64-lane workgroup per block): the targets of
// each row over the block's columns, then lane 0 cuts the rows greedily into
// chunks of at most kStage3Threads targets spanning at most kChunkRows3 rows
// (a single row with more targets forms a chunk of its own), skipping empty
// rows at a chunk's ends.  With emit, the chunks are written from the
// block's scanned offset; otherwise the count goes to bcount.
__global__ __launch_bounds__(64) void k_block_chunks(StageArgs A, int emit, const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t s_t[64];
  const uint32_t b = blockIdx.x;
  const uint4 blk = A.blocks[b];
  const DevTank T = A.tanks[blk.x];
  const int nz = T.nc[2];
  const int bx = static_cast<int>(blk.y), by = static_cast<int>(blk.z);
  const int lane = threadIdx.x;
  int start = -1, last = -1;
  uint32_t cur = 0, count = 0;
  const uint32_t base = emit ? A.bcount[b] : 0u;
  auto close = [&]() {
    if (emit) A.bchunks[base + count] = make_uint4(b, start, last + 1, cur);
    ++count;
  };
  for (int z0 = 0; z0 < nz; z0 += 64) {
    const int z = z0 + lane;
    uint32_t tz = 0;
    if (z < nz) {
      for (int q = 0; q < kBlk3 * kBlk3; ++q) {
        const int ix = bx + q / kBlk3, iy = by + q % kBlk3;
        if (ix < T.nc[0] && iy < T.nc[1]) {
          const uint32_t row = column_row3(T, ix, iy);
          tz += A.cell_off[row + z + 1] - A.cell_off[row + z];
        }
      }
    }
    s_t[lane] = tz;
    __syncthreads();
    if (lane == 0) {
      const int rows = min(64, nz - z0);
      for (int r = 0; r < rows; ++r) {
        const uint32_t t = s_t[r];
        if (t == 0) continue;
        const int zz = z0 + r;
        if (start >= 0 && (cur + t > static_cast<uint32_t>(kStage3Threads) ||
                           zz - start + 1 > kChunkRows3)) {
          close();
          start = -1;
        }
        if (start < 0) {
          start = zz;
          cur = 0;
        }
        cur += t;
        last = zz;
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    if (start >= 0) close();
    if (!emit) A.bcount[b] = count;
  }
}

// ---- chunk list (after every cell rebuild) ----------------------------------

constexpr int kScanBlock = 1024;

// The chunks of a column, in order: the column's particles in
// ceil(n / chunk_targets) equal parts, each cut further into pieces whose
// targets span at most kChunkRows cells (a piece starting at particle p ends
// before the first particle with key >= key[p] + kChunkRows).  A column gains
// at most (nz - 1) / kChunkRows cuts: each cut starts a piece kChunkRows
// cells or more above the last, and the parts' spans do not overlap.
struct ColumnParts {
  uint32_t beg, n, nch;
};

__device__ __forceinline__ ColumnParts column_parts(const StageArgs& A, uint32_t col_i) {
  const uint4 col = A.cols[col_i];
  const DevTank& T = A.tanks[col.x];
  const int nz = T.nc[A.P.dim - 1];
  ColumnParts cp;
  cp.beg = A.cell_off[col.w];
  cp.n = A.cell_off[col.w + nz] - cp.beg;
  cp.nch = (cp.n + A.chunk_targets - 1) / A.chunk_targets;
  return cp;
}

// The pieces of part k of a column: emit(i, p0, p1) for its i-th piece;
// returns their count.
template <class F>
__device__ __forceinline__ uint32_t part_pieces(const StageArgs& A, const ColumnParts& cp,
                                                uint32_t k, F&& emit) {
  const uint32_t p0 = cp.beg + static_cast<uint32_t>((static_cast<uint64_t>(cp.n) * k) / cp.nch);
  const uint32_t p1 =
      cp.beg + static_cast<uint32_t>((static_cast<uint64_t>(cp.n) * (k + 1)) / cp.nch);
  uint32_t p = p0, count = 0;
  while (p < p1) {
    const uint32_t lim = A.key[p] + static_cast<uint32_t>(kChunkRows);
    uint32_t q = p1;
    if (A.key[p1 - 1] >= lim) {
      // the first particle in (p, p1) with key >= lim (keys ascend)
      uint32_t lo = p + 1, hi = p1 - 1;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (A.key[mid] >= lim) {
          hi = mid;
        } else {
          lo = mid + 1;
        }
      }
      q = lo;
    }
    emit(count, p, q);
    ++count;
    p = q;
  }
  return count;
}

// Pieces of the column's parts that lane takes (parts lane, lane + 64, ...).
__device__ __forceinline__ uint32_t lane_pieces(const StageArgs& A, const ColumnParts& cp,
                                                int lane) {
  uint32_t cnt = 0;
  for (uint32_t k = lane; k < cp.nch; k += 64) {
    cnt += part_pieces(A, cp, k, [](uint32_t, uint32_t, uint32_t) {});
  }
  return cnt;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Exclusive scan over a 1024-thread block: wave shuffles, then the 16 wave
// totals.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wsum,
                                                    uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
      const uint32_t t = wsum[w];
      wsum[w] = s;
      s += t;
    }
    wsum[kScanBlock / 64] = s;
  }
  __syncthreads();
  const uint32_t r = wsum[wave] + x - v;
  *total = wsum[kScanBlock / 64];
  __syncthreads();
  return r;
}

// Chunk list in column order, in three passes, one wave per column (its
// lanes take the column's parts in turn): chunks per column and their
// per-workgroup totals; an exclusive scan of the totals (one workgroup, also
// writes the chunk count); each column's chunks at its scanned offset.
constexpr uint32_t kColsPerBlock = kScanBlock / 64;

__global__ __launch_bounds__(kScanBlock) void k_chunk_count(StageArgs A, uint32_t* block_tot,
                                                           const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wtot[kColsPerBlock];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * kColsPerBlock + wave;
  uint32_t cnt = 0;
  if (c < A.n_cols) cnt = lane_pieces(A, column_parts(A, c), lane);
  cnt = wave_sum_u32(cnt);
  if (lane == 0) wtot[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < kColsPerBlock; ++w) t += wtot[w];
    block_tot[blockIdx.x] = t;
  }
}

// Exclusive scan of n_blocks counts in place (one workgroup); out[0] = the
// total (chunks), out[1] = 0 (overflow chunks, appended by the stage kernel).
__global__ __launch_bounds__(kScanBlock) void k_chunk_base(uint32_t* block_tot, uint32_t n_blocks,
                                                          uint32_t* out, const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wsum[kScanBlock / 64 + 1];
  uint32_t base_off = 0;
  for (uint32_t base = 0; base < n_blocks; base += kScanBlock) {
    const uint32_t b = base + threadIdx.x;
    const uint32_t v = b < n_blocks ? block_tot[b] : 0u;
    uint32_t tot;
    const uint32_t off = base_off + block_excl_scan(v, wsum, &tot);
    if (b < n_blocks) block_tot[b] = off;
    base_off += tot;
  }
  if (threadIdx.x == 0) {
    out[0] = base_off;
    out[1] = 0;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_chunk_emit(StageArgs A,
                                                          const uint32_t* block_tot,
                                                          const uint32_t* ctl) {
  if (!ctl[0]) return;
  __shared__ uint32_t wtot[kColsPerBlock];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t c = blockIdx.x * kColsPerBlock + wave;
  // the counts again (the offsets need every earlier column's)
  ColumnParts cp{0u, 0u, 0u};
  uint32_t cnt = 0;
  if (c < A.n_cols) {
    cp = column_parts(A, c);
    cnt = lane_pieces(A, cp, lane);
  }
  cnt = wave_sum_u32(cnt);
  if (lane == 0) wtot[wave] = cnt;
  __syncthreads();
  if (c >= A.n_cols) return;
  uint32_t off = block_tot[blockIdx.x];
  for (int w = 0; w < wave; ++w) off += wtot[w];
  // parts in rounds of 64, each part's pieces at its exclusive prefix
  for (uint32_t k0 = 0; k0 < cp.nch; k0 += 64) {
    const uint32_t k = k0 + static_cast<uint32_t>(lane);
    const uint32_t np =
        k < cp.nch ? part_pieces(A, cp, k, [](uint32_t, uint32_t, uint32_t) {}) : 0u;
    uint32_t x = np;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    const uint32_t base = off + x - np;
    if (k < cp.nch) {
      (void)part_pieces(A, cp, k, [&](uint32_t i, uint32_t p0, uint32_t p1) {
        A.chunks[base + i] = make_uint4(c, p0, p1, 0u);
      });
    }
    off += __shfl(x, 63);
  }
}

// Clears an overflow count (counts[1] of the pair given).
__global__ void k_reset_overflow(uint32_t* counts) { counts[1] = 0u; }

// ---- dispatch -------------------------------------------------------------------

constexpr int kThreads = kChunkMax;

template <int DIM, int KERNEL, int FLAVOUR, int KIND, bool FINAL>
void launch_kind(hipStream_t s, const StageArgs& a) {
  if constexpr (DIM == 2) {
    k_reset_overflow<<<1, 1, 0, s>>>(a.counts);
    k_stage_lds<KERNEL, FLAVOUR, KIND, FINAL, kThreads, kStageLdsCap>
        <<<a.grid_lds, kThreads, 0, s>>>(a);
    k_stage_global<2, KERNEL, FLAVOUR, KIND, FINAL, kThreads>
        <<<a.grid_global, kThreads, 0, s>>>(a, a.overflow, a.counts + 1);
  } else if (a.lds_cap > 0 && a.n_blocks > 0) {
    k_reset_overflow<<<1, 1, 0, s>>>(a.counts + 2);
    k_stage_lds3<KERNEL, FLAVOUR, KIND, FINAL, kStage3Threads, kStage3LdsCap>
        <<<a.grid_lds3, kStage3Threads, 0, s>>>(a);
    k_stage_block_global<KERNEL, FLAVOUR, KIND, FINAL, kThreads>
        <<<a.grid_block_global, kThreads, 0, s>>>(a);
  } else {
    k_stage_global<3, KERNEL, FLAVOUR, KIND, FINAL, kThreads>
        <<<a.grid_global, kThreads, 0, s>>>(a, nullptr, a.counts);
  }
}

template <int DIM, int KERNEL, int FLAVOUR>
void dispatch_kind(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (kind == kRhsOnly) {
    launch_kind<DIM, KERNEL, FLAVOUR, kRhsOnly, false>(s, a);
  } else if (kind == kEuler) {
    if (final) {
      launch_kind<DIM, KERNEL, FLAVOUR, kEuler, true>(s, a);
    } else {
      launch_kind<DIM, KERNEL, FLAVOUR, kEuler, false>(s, a);
    }
  } else {
    if (final) {
      launch_kind<DIM, KERNEL, FLAVOUR, kCombine, true>(s, a);
    } else {
      launch_kind<DIM, KERNEL, FLAVOUR, kCombine, false>(s, a);
    }
  }
}

template <int DIM, int KERNEL>
void dispatch_flavour(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.flavour == kFlavourCfg) {
    dispatch_kind<DIM, KERNEL, kFlavourCfg>(s, a, kind, final);
  } else if (a.P.flavour == kFlavourWall) {
    dispatch_kind<DIM, KERNEL, kFlavourWall>(s, a, kind, final);
  } else if (a.P.flavour == kFlavourCfgM) {
    dispatch_kind<DIM, KERNEL, kFlavourCfgM>(s, a, kind, final);
  } else {
    dispatch_kind<DIM, KERNEL, kFlavourRef>(s, a, kind, final);
  }
}

template <int DIM>
void dispatch_kernel(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.P.kernel == kKernelC4) {
    dispatch_flavour<DIM, kKernelC4>(s, a, kind, final);
  } else {
    dispatch_flavour<DIM, kKernelC2>(s, a, kind, final);
  }
}

template <class F>
int occupancy(F* f, int threads = kThreads) {
  int n = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(f),
                                                     threads, 0);
  return std::max(n, 1);
}

}  // namespace

uint32_t stage_lds_capacity(int dim) { return dim == 2 ? kStageLdsCap : kStage3LdsCap; }

void stage_grids(const DevParams& P, int n_cu, StageGrids* g) {
  int bl = 1, bg, b3 = 1, bb = 1;
  if (P.dim == 2) {
    bl = occupancy(k_stage_lds<kKernelC2, kFlavourCfg, kCombine, false, kThreads, kStageLdsCap>);
    bg = occupancy(k_stage_global<2, kKernelC2, kFlavourCfg, kCombine, false, kThreads>);
  } else {
    bg = occupancy(k_stage_global<3, kKernelC2, kFlavourCfg, kCombine, false, kThreads>);
    b3 = occupancy(k_stage_lds3<kKernelC2, kFlavourCfg, kCombine, false, kStage3Threads,
                                kStage3LdsCap>,
                   kStage3Threads);
    bb = occupancy(k_stage_block_global<kKernelC2, kFlavourCfg, kCombine, false, kThreads>);
  }
  // multiples of 8: every XCD owns the same number of slots
  auto grid = [&](int per_cu) { return 8u * static_cast<uint32_t>((per_cu * n_cu + 7) / 8); };
  g->lds = grid(bl);
  g->global = grid(bg);
  g->lds3 = grid(b3);
  g->block_global = grid(bb);
}

void launch_build_chunks(hipStream_t s, const StageArgs& a, uint32_t* block_tot,
                         const uint32_t* ctl) {
  const uint32_t nb = (a.n_cols + kColsPerBlock - 1) / kColsPerBlock;
  if (nb == 0) return;
  k_chunk_count<<<nb, kScanBlock, 0, s>>>(a, block_tot, ctl);
  k_chunk_base<<<1, kScanBlock, 0, s>>>(block_tot, nb, a.counts, ctl);
  k_chunk_emit<<<nb, kScanBlock, 0, s>>>(a, block_tot, ctl);
  if (a.P.dim == 3 && a.n_blocks > 0) {
    k_block_chunks<<<a.n_blocks, 64, 0, s>>>(a, 0, ctl);
    k_chunk_base<<<1, kScanBlock, 0, s>>>(a.bcount, a.n_blocks, a.counts + 2, ctl);
    k_block_chunks<<<a.n_blocks, 64, 0, s>>>(a, 1, ctl);
  }
}

void launch_stage(hipStream_t s, const StageArgs& a, int kind, bool final) {
  if (a.n == 0 || a.n_cols == 0) return;
  if (a.P.dim == 2) {
    dispatch_kernel<2>(s, a, kind, final);
  } else {
    dispatch_kernel<3>(s, a, kind, final);
  }
}

}  // namespace tit_amd
