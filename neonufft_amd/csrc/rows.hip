// Row kernels of the pruned FFT for 2D / 3D plans whose fine grid has a
// power-of-two x extent n0 >= 32: the x-direction FFT runs in LDS on pencils
// of whole x rows, joined to the step before it (type 1) or after it (type 2),
// so the fine grid is not written, read back and transposed in between.
//
//  - type 1: reduce_rows_fft_kernel sums the spreading tiles that cover each
//    cell of the pencil (same sums as the tile reduction: every cell receives
//    its own bin's tile and the covering halo of the neighbour bins' tiles,
//    each subproblem of each), transforms the rows along x in LDS and writes
//    only the retained x modes, transposed, to the pruned intermediate
//    h[(z N0 + x') n1 + y];
//  - type 2: expand_rows_ifft_kernel reads the retained x modes of the pencil
//    from h, zero-pads them to n0 in LDS, transforms along x and writes whole
//    fine-grid rows.
// In 3D with power-of-two y and z extents, the y and z passes also run in
// LDS here (y_keep / z_out for type 1, z_in / y_expand for type 2), so no
// rocFFT pass, transpose or separate deconvolution pass is left.
// The results equal those of the full transform the reference runs
// (src/neonufft/gpu/plan_impl.hpp:91-136: spread, in-place FFT, downsample;
// upsample, in-place FFT, interpolate), up to rounding.
#include "lds_fft.hpp"
#include "mode_index.hpp"

namespace nnf {

namespace {

template <typename T>
constexpr int kRowThreads = sizeof(T) == 4 ? 512 : 256;

// Moves `total` items global -> LDS: each lane issues the loads of up to
// kBatch items (clamped to a valid item, so none sits behind a branch)
// before it stores any; a loop that stores each load before issuing the next
// waits on every load in turn.
template <int NT, typename Load, typename Store>
__device__ inline void batched_copy(int total, Load load, Store store) {
  constexpr int kBatch = 8;
  for (int i0 = static_cast<int>(threadIdx.x); i0 < total; i0 += NT * kBatch) {
    decltype(load(0)) v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = load(min(i0 + u * NT, total - 1));
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (i0 + u * NT < total) store(i0 + u * NT, v[u]);
    }
  }
}

// Pencil of workgroup blockIdx.x: rows [y0, y0 + rk) of plane z. Pencils
// next to each other in y and z (which read the same halo tiles) go to one
// XCD and so meet in its L2.
struct Pencil {
  int z, y0, rk;
};

__device__ inline Pencil pencil_of(int rows, int n1) {
  const int groups = (n1 + rows - 1) / rows;
  const int b = static_cast<int>(xcd_remap(blockIdx.x, gridDim.x));
  Pencil p;
  p.z = b / groups;
  p.y0 = (b - p.z * groups) * rows;
  p.rk = min(rows, n1 - p.y0);
  return p;
}

template <typename T, int DIM>
__global__ void __launch_bounds__(kRowThreads<T>)
    reduce_rows_fft_kernel(RowFftJob<T> job, int ns) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const BinGeom& g = job.g;
  const int n0 = g.n[0], n1 = g.n[1];
  const Pencil pc = pencil_of(job.rows, n1);
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + job.rows * job.pitch;
  // subproblem ranges of the bins the pencil's cells can meet: every x bin,
  // the y bins of its rows and their neighbours, the z bin of its plane and
  // its neighbours (periodic); entry (jx, jy, jz) at jx + nb0 (jy + ny jz)
  uint32_t* rng = reinterpret_cast<uint32_t*>(fim + job.rows * job.pitch);
  const int h = ns / 2;
  const int nb0 = g.nbins[0];
  const int by_lo = pc.y0 / g.bin[1] - 1;
  const int ny = (pc.y0 + pc.rk - 1) / g.bin[1] + 2 - by_lo;  // <= (rows - 1) / bin + 4
  const int bz_lo = DIM >= 3 ? pc.z / g.bin[2] - 1 : 0;
  const int nz = DIM >= 3 ? 3 : 1;
  const int nent = nb0 * ny * nz;
  for (int e = threadIdx.x; e < nent; e += NT) {
    const int jx = e % nb0;
    const int jy = (e / nb0) % ny;
    const int jz = e / (nb0 * ny);
    int by = (by_lo + jy) % g.nbins[1];
    by = by < 0 ? by + g.nbins[1] : by;
    int bz = 0;
    if constexpr (DIM >= 3) {
      bz = (bz_lo + jz) % g.nbins[2];
      bz = bz < 0 ? bz + g.nbins[2] : bz;
    }
    const int bin = jx + nb0 * (by + g.nbins[1] * bz);
    rng[2 * e] = job.sub_start[bin];
    rng[2 * e + 1] = job.sub_start[bin + 1];
  }
  __syncthreads();

  // Cover of a coordinate along dimension d: own bin (as its entry index
  // along d in the range table) and position in that bin's tile, and the one
  // neighbour whose halo also covers it (entry -1 if none) with its position.
  struct Cover {
    int jown, pown, joth, poth;
  };
  auto cover = [&](int c, int d, int lo) {
    Cover cv;
    int bd;
    const int li = div_rem(c, g.bin[d], 1.0f / static_cast<float>(g.bin[d]), &bd);
    const int w = min(g.bin[d], g.n[d] - bd * g.bin[d]);
    const int nb = g.nbins[d];
    // entry of bin b along d: (b - lo) mod nb (lo = first bin of the table)
    auto entry = [&](int bb) {
      bb = bb < 0 ? bb + nb : (bb >= nb ? bb - nb : bb);
      int j = bb - lo;
      j = j < 0 ? j + nb : (j >= nb ? j - nb : j);
      return j;
    };
    cv.jown = entry(bd);
    cv.pown = li + h;
    cv.joth = -1;
    cv.poth = 0;
    if (li < ns - h) {
      const int bp = bd == 0 ? nb - 1 : bd - 1;
      cv.joth = entry(bd - 1);
      cv.poth = li + min(g.bin[d], g.n[d] - bp * g.bin[d]) + h;
    } else if (li >= w - h) {
      cv.joth = entry(bd + 1);
      cv.poth = li - w + h;
    }
    return cv;
  };
  // Per row of the pencil and per covering tile pair (own / neighbour y bin
  // x own / neighbour z bin): the offset of the row's tile row inside a tile
  // and the range-table entry of x bin 0 (the x bin adds to it), or -1 when
  // that neighbour does not cover the row.
  constexpr int NQ = DIM >= 3 ? 4 : 2;
  __shared__ int combo[16][NQ][2];
  if (static_cast<int>(threadIdx.x) < pc.rk * NQ) {
    const int r = static_cast<int>(threadIdx.x) / NQ;
    const int q = static_cast<int>(threadIdx.x) - r * NQ;
    const int oy = DIM >= 3 ? q >> 1 : q;
    const int oz = DIM >= 3 ? q & 1 : 0;
    const Cover cy = cover(pc.y0 + r, 1, by_lo);
    Cover cz{0, 0, -1, 0};
    if constexpr (DIM >= 3) cz = cover(pc.z, 2, bz_lo);
    const int jy = oy ? cy.joth : cy.jown, py = oy ? cy.poth : cy.pown;
    const int jz = oz ? cz.joth : cz.jown, pz = oz ? cz.poth : cz.pown;
    const bool ok = jy >= 0 && jz >= 0;
    combo[r][q][0] = ok ? g.tile[0] * (py + g.tile[1] * pz) : -1;
    combo[r][q][1] = ok ? nb0 * (jy + ny * jz) : 0;
  }
  for (int i = threadIdx.x; i < 2 * job.rows * job.pitch; i += NT) fre[i] = T(0);
  __syncthreads();

  // Tile rows into the pencil, one x bin per item: the row segment of x bin
  // bx's tiles (every subproblem, own and covering-neighbour y / z tiles)
  // is summed and added to grid cells bx bin0 - ns/2 + j (mod n0),
  // j < bin width + ns. Consecutive lanes take consecutive j, so tile rows
  // are read coalesced; the first subproblem of every covering tile is
  // loaded before any is waited for. Segments of bins of one parity do not
  // overlap, so the bins are added in two passes (three with an odd bin
  // count: the last bin, whose wrapped segment meets bin 0's, alone).
  const std::size_t tc = tile_cells(g);
  const int t0 = g.tile[0];
  const int bin0 = g.bin[0];
  const float rt0 = 1.0f / static_cast<float>(t0);
  const int nphase = (nb0 & 1) ? 3 : 2;
  for (int phase = 0; phase < nphase; ++phase) {
    // bins of this phase: bx = phase, phase + 2, ... (the odd last bin: phase 2)
    const int last = (nb0 & 1) ? nb0 - 1 : nb0;
    const int nbp = phase == 2 ? 1 : (last - phase + 1) / 2;
    const float rnbp = 1.0f / static_cast<float>(nbp);
    const int items = pc.rk * nbp * t0;
    // two items per lane per step, so that every tile load of both is in
    // flight before the first wait
    constexpr int U = 2;
    for (int i0 = threadIdx.x; i0 < items; i0 += U * NT) {
      int r[U], x[U];
      bool live[U];
      std::size_t o[U][NQ];
      uint32_t s0[U][NQ], s1[U][NQ];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int i = min(i0 + u * NT, items - 1);
        int rest;
        const int j = div_rem(i, t0, rt0, &rest);
        const int bxi = div_rem(rest, nbp, rnbp, &r[u]);
        const int bx = phase == 2 ? nb0 - 1 : phase + 2 * bxi;
        const int ext = min(bin0, n0 - bx * bin0) + ns;
        live[u] = i0 + u * NT < items && j < ext;
        int xx = bx * bin0 - h + j;
        x[u] = xx < 0 ? xx + n0 : (xx >= n0 ? xx - n0 : xx);
        // element offsets of the first subproblem's tile row segment; a
        // pair that does not cover the row (its entry base is 0, so the
        // range read stays in the table), a bin without subproblems or a
        // dead item reads element 0 (the tile buffer always holds at least
        // one tile) and is then dropped
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int off = combo[r[u]][q][0];
          const int e = bx + combo[r[u]][q][1];
          const uint32_t lo = rng[2 * e], hi = rng[2 * e + 1];
          s0[u][q] = lo;
          s1[u][q] = off >= 0 && live[u] ? hi : lo;
          o[u][q] = lo < s1[u][q]
                        ? static_cast<std::size_t>(lo) * tc + static_cast<std::size_t>(off + j)
                        : std::size_t{0};
        }
      }
      const Cplx<T>* __restrict__ tiles = job.tiles;
      Cplx<T> v[U][NQ];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int q = 0; q < NQ; ++q) v[u][q] = tiles[o[u][q]];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        T ar = T(0), ai = T(0);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const bool has = s0[u][q] < s1[u][q];
          ar += has ? v[u][q].x : T(0);
          ai += has ? v[u][q].y : T(0);
        }
        // further subproblems of bins split by their point count
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          for (uint32_t sp = s0[u][q] + 1; sp < s1[u][q]; ++sp) {
            const Cplx<T> w = tiles[o[u][q] + static_cast<std::size_t>(sp - s0[u][q]) * tc];
            ar += w.x;
            ai += w.y;
          }
        }
        if (live[u]) {
          fre[r[u] * job.pitch + x[u]] += ar;
          fim[r[u] * job.pitch + x[u]] += ai;
        }
      }
    }
    __syncthreads();
  }

  lds_fft_x<T, NT, false>(fre, fim, pc.rk, job.pitch, job.log2n0, T(1), job.tw);

  // retained mode xp of row r sits at position bitrev(f); lanes take r
  // fastest so that consecutive lanes write consecutive y of one mode
  const int N0 = job.N0;
  const int P = N0 - N0 / 2;
  const int sh = 32 - job.log2n0;
  const int lr = job.log2rows;
  Cplx<T>* __restrict__ hz = job.h + static_cast<std::size_t>(pc.z) * N0 * n1;
  for (int i = threadIdx.x; i < (N0 << lr); i += NT) {
    const int r = i & (job.rows - 1);
    const int xp = i >> lr;
    if (r >= pc.rk) continue;
    const int f = xp < P ? xp : xp + n0 - N0;
    const int pos = r * job.pitch + static_cast<int>(__brev(static_cast<unsigned>(f)) >> sh);
    hz[static_cast<std::size_t>(xp) * n1 + pc.y0 + r] = Cplx<T>{fre[pos], fim[pos]};
  }
}

template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) expand_rows_ifft_kernel(RowFftJob<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int n0 = job.g.n[0], n1 = job.g.n[1];
  const Pencil pc = pencil_of(job.rows, n1);
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + job.rows * job.pitch;
  for (int i = threadIdx.x; i < 2 * job.rows * job.pitch; i += NT) fre[i] = T(0);
  __syncthreads();
  // retained x modes of the pencil's rows at their fine positions
  const int N0 = job.N0;
  const int P = N0 - N0 / 2;
  const int lr = job.log2rows;
  const Cplx<T>* __restrict__ hz = job.h + static_cast<std::size_t>(pc.z) * N0 * n1;
  // items (r fastest, then x'); rows r >= rk repeat row rk - 1 and are dropped
  batched_copy<NT>(
      N0 << lr,
      [&](int i) {
        return hz[static_cast<std::size_t>(i >> lr) * n1 + pc.y0 + min(i & (job.rows - 1), pc.rk - 1)];
      },
      [&](int i, const Cplx<T>& v) {
        const int r = i & (job.rows - 1);
        if (r >= pc.rk) return;
        const int xp = i >> lr;
        const int o = r * job.pitch + (xp < P ? xp : xp + n0 - N0);
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();

  lds_fft_x<T, NT, false>(fre, fim, pc.rk, job.pitch, job.log2n0, T(1), job.tw);

  const int sh = 32 - job.log2n0;
  Cplx<T>* __restrict__ grid =
      job.grid + (static_cast<std::size_t>(pc.z) * n1 + pc.y0) * static_cast<std::size_t>(n0);
  for (int i = threadIdx.x; i < pc.rk * n0; i += NT) {
    const int r = i >> job.log2n0;
    const int x = i & (n0 - 1);
    const int pos = r * job.pitch + static_cast<int>(__brev(static_cast<unsigned>(x)) >> sh);
    grid[static_cast<std::size_t>(r) * n0 + x] = Cplx<T>{fre[pos], fim[pos]};
  }
}

// ---------------------------------------------------------------------------
// y and z passes of the 3D pruned FFT (AxisFftJob). Each workgroup loads
// `rows` lines of one axis into LDS rows (natural order, zero padding where
// modes are absent), transforms them with the same decimation-in-frequency
// FFT as the x rows (output at bit-reversed positions) and writes the next
// layout directly, so the rocFFT pass, the transpose and the deconvolution
// pass become one read and one write of the data.
// ---------------------------------------------------------------------------

// Type 1: lines y of h at one x' and rows z0 .. z0 + rk; the retained y' go
// to k[(y' N0 + x') n2 + z] (consecutive lanes: consecutive z of one y').
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) y_keep_kernel(AxisFftJob<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows_y, n1 = job.n1, n2 = job.n2, N0 = job.N0;
  const int pitch = n1 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int b = static_cast<int>(blockIdx.x);
  const int xp = b % N0;
  const int z0 = (b / N0) * R;
  const int rk = min(R, n2 - z0);
  const Cplx<T>* __restrict__ src = job.h + static_cast<std::size_t>(xp) * n1;
  const std::size_t zs = static_cast<std::size_t>(N0) * n1;
  const int l1 = job.log2n1;
  batched_copy<NT>(
      rk << l1,
      [&](int i) { return src[static_cast<std::size_t>(z0 + (i >> l1)) * zs + (i & (n1 - 1))]; },
      [&](int i, const Cplx<T>& v) {
        const int o = (i >> l1) * pitch + (i & (n1 - 1));
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, job.log2n1, T(1), job.tw);
  const int sh = 32 - job.log2n1;
  const int lr = job.log2rows_y;
  Cplx<T>* __restrict__ dst = job.k + static_cast<std::size_t>(xp) * n2 + z0;
  const std::size_t ys = static_cast<std::size_t>(N0) * n2;
  for (int i = threadIdx.x; i < (job.N1 << lr); i += NT) {
    const int r = i & (R - 1), yp = i >> lr;
    if (r >= rk) continue;
    const int f = ret_fine(yp, n1, job.N1);
    const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(f)) >> sh);
    dst[static_cast<std::size_t>(yp) * ys + r] = Cplx<T>{fre[pos], fim[pos]};
  }
}

// Type 1: lines z of k at one y' and x' = x0 .. x0 + rk; the retained z'
// deconvolved into the caller's modes (consecutive lanes: consecutive x').
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) z_out_kernel(AxisFftJob<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows_z, n2 = job.n2, N0 = job.N0, N1 = job.N1, N2 = job.N2;
  const int pitch = n2 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int groups = (N0 + R - 1) / R;
  const int b = static_cast<int>(blockIdx.x);
  const int yp = b / groups;
  const int x0 = (b - yp * groups) * R;
  const int rk = min(R, N0 - x0);
  const Cplx<T>* __restrict__ src = job.k + (static_cast<std::size_t>(yp) * N0 + x0) * n2;
  const int l2 = job.log2n2;
  batched_copy<NT>(
      rk << l2, [&](int i) { return src[i]; },
      [&](int i, const Cplx<T>& v) {
        const int o = (i >> l2) * pitch + (i & (n2 - 1));
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, job.log2n2, T(1), job.tw);
  const int sh = 32 - job.log2n2;
  const int lr = job.log2rows_z;
  const int km = ret_mode(yp, N1);
  const T fm = job.fs[1][km < 0 ? -km : km];
  const long long om = static_cast<long long>(order_index(km, N1, job.order_fft)) * job.st[1];
  Cplx<T>* __restrict__ out = job.modes_out;
  for (int i = threadIdx.x; i < (N2 << lr); i += NT) {
    const int r = i & (R - 1), lp = i >> lr;
    if (r >= rk) continue;
    const int kx = ret_mode(x0 + r, N0), kl = ret_mode(lp, N2);
    const T fac = job.fs[0][kx < 0 ? -kx : kx] * fm * job.fs[2][kl < 0 ? -kl : kl];
    const int f = ret_fine(lp, n2, N2);
    const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(f)) >> sh);
    out[static_cast<long long>(order_index(kx, N0, job.order_fft)) * job.st[0] + om +
        static_cast<long long>(order_index(kl, N2, job.order_fft)) * job.st[2]] =
        Cplx<T>{fre[pos] * fac, fim[pos] * fac};
  }
}

// 2D type 1: lines y of transform b's h at x' = x0 .. x0 + rk; the retained
// y' deconvolved into that transform's modes (consecutive lanes: consecutive
// x'). Neighbouring x' groups of a transform map to one XCD, so the pieces
// of output lines they write (rows x 8 bytes each) meet in one L2.
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) y_out2_kernel(Y2Job<T> job) {
  constexpr int NT = kRowThreads<T>;
  static_assert(NT % 16 == 0, "workgroup rows (at most 16) must divide the thread count");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows, n1 = job.n1, N0 = job.N0, N1 = job.N1;
  const int pitch = n1 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int groups = (N0 + R - 1) / R;
  const int lb = static_cast<int>(
      xcd_remap(blockIdx.x, static_cast<uint32_t>(job.count) * static_cast<uint32_t>(groups)));
  const int b = lb / groups;
  const int x0 = (lb - b * groups) * R;
  const int rk = min(R, N0 - x0);
  const Cplx<T>* __restrict__ src =
      job.h + b * job.member_stride + static_cast<long long>(x0) * n1;
  const int l1 = job.log2n1;
  batched_copy<NT>(
      rk << l1, [&](int i) { return src[i]; },
      [&](int i, const Cplx<T>& v) {
        const int o = (i >> l1) * pitch + (i & (n1 - 1));
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, l1, T(1), job.tw);
  const int sh = 32 - l1;
  const int lr = job.log2rows;
  Cplx<T>* __restrict__ out = job.outs ? job.outs[b] : job.single;
  // a lane keeps its line r (NT is a multiple of R) and steps through y'
  const int r = static_cast<int>(threadIdx.x) & (R - 1);
  if (r < rk) {
    const int kx = ret_mode(x0 + r, N0);
    const T fx = job.fs0[kx < 0 ? -kx : kx];
    Cplx<T>* __restrict__ orow =
        out + static_cast<long long>(order_index(kx, N0, job.order_fft)) * job.st0;
    const int rp = r * pitch;
    for (int lp = static_cast<int>(threadIdx.x) >> lr; lp < N1; lp += NT >> lr) {
      const int ky = ret_mode(lp, N1);
      const T fac = fx * job.fs1[ky < 0 ? -ky : ky];
      const int f = ret_fine(lp, n1, N1);
      const int pos = rp + static_cast<int>(__brev(static_cast<unsigned>(f)) >> sh);
      orow[static_cast<long long>(order_index(ky, N1, job.order_fft)) * job.st1] =
          Cplx<T>{fre[pos] * fac, fim[pos] * fac};
    }
  }
}

// 2D type 2: transform b's modes of x' = x0 .. x0 + rk, deconvolved, at their
// y positions of zero-padded lines; the transformed lines go to h. Same
// workgroup map as y_out2_kernel.
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) y_in2_kernel(Y2Job<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows, n1 = job.n1, N0 = job.N0, N1 = job.N1;
  const int pitch = n1 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int groups = (N0 + R - 1) / R;
  const int lb = static_cast<int>(
      xcd_remap(blockIdx.x, static_cast<uint32_t>(job.count) * static_cast<uint32_t>(groups)));
  const int b = lb / groups;
  const int x0 = (lb - b * groups) * R;
  const int rk = min(R, N0 - x0);
  for (int i = threadIdx.x; i < 2 * R * pitch; i += NT) fre[i] = T(0);
  __syncthreads();
  const int lr = job.log2rows;
  const Cplx<T>* __restrict__ in = job.ins ? job.ins[b] : job.single_in;
  // items (r fastest, then y'); rows r >= rk repeat row rk - 1 and are dropped
  batched_copy<NT>(
      N1 << lr,
      [&](int i) {
        const int r = min(i & (R - 1), rk - 1), lp = i >> lr;
        const int kx = ret_mode(x0 + r, N0), ky = ret_mode(lp, N1);
        const T fac = job.fs0[kx < 0 ? -kx : kx] * job.fs1[ky < 0 ? -ky : ky];
        const Cplx<T> a = in[static_cast<long long>(order_index(kx, N0, job.order_fft)) * job.st0 +
                             static_cast<long long>(order_index(ky, N1, job.order_fft)) * job.st1];
        return Cplx<T>{a.x * fac, a.y * fac};
      },
      [&](int i, const Cplx<T>& v) {
        const int r = i & (R - 1), lp = i >> lr;
        if (r >= rk) return;
        const int o = r * pitch + ret_fine(lp, n1, N1);
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  const int l1 = job.log2n1;
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, l1, T(1), job.tw);
  const int sh = 32 - l1;
  Cplx<T>* __restrict__ dst = job.h + b * job.member_stride + static_cast<long long>(x0) * n1;
  for (int i = threadIdx.x; i < (rk << l1); i += NT) {
    const int r = i >> l1, y = i & (n1 - 1);
    const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(y)) >> sh);
    dst[static_cast<long long>(r) * n1 + y] = Cplx<T>{fre[pos], fim[pos]};
  }
}

// Type 2: the caller's modes of one y' and x' = x0 .. x0 + rk, deconvolved,
// at their z positions of zero-padded lines; the transformed lines go to k.
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) z_in_kernel(AxisFftJob<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows_z, n2 = job.n2, N0 = job.N0, N1 = job.N1, N2 = job.N2;
  const int pitch = n2 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int groups = (N0 + R - 1) / R;
  const int b = static_cast<int>(blockIdx.x);
  const int yp = b / groups;
  const int x0 = (b - yp * groups) * R;
  const int rk = min(R, N0 - x0);
  for (int i = threadIdx.x; i < 2 * R * pitch; i += NT) fre[i] = T(0);
  __syncthreads();
  const int lr = job.log2rows_z;
  const int km = ret_mode(yp, N1);
  const T fm = job.fs[1][km < 0 ? -km : km];
  const long long om = static_cast<long long>(order_index(km, N1, job.order_fft)) * job.st[1];
  const Cplx<T>* __restrict__ in = job.modes_in;
  // items (r fastest, then z'); rows r >= rk repeat row rk - 1 and are dropped
  batched_copy<NT>(
      N2 << lr,
      [&](int i) {
        const int r = min(i & (R - 1), rk - 1), lp = i >> lr;
        const int kx = ret_mode(x0 + r, N0), kl = ret_mode(lp, N2);
        const T fac = fm * job.fs[0][kx < 0 ? -kx : kx] * job.fs[2][kl < 0 ? -kl : kl];
        const Cplx<T> a =
            in[static_cast<long long>(order_index(kx, N0, job.order_fft)) * job.st[0] + om +
               static_cast<long long>(order_index(kl, N2, job.order_fft)) * job.st[2]];
        return Cplx<T>{a.x * fac, a.y * fac};
      },
      [&](int i, const Cplx<T>& v) {
        const int r = i & (R - 1), lp = i >> lr;
        if (r >= rk) return;
        const int o = r * pitch + ret_fine(lp, n2, N2);
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, job.log2n2, T(1), job.tw);
  const int sh = 32 - job.log2n2;
  Cplx<T>* __restrict__ dst = job.k + (static_cast<std::size_t>(yp) * N0 + x0) * n2;
  for (int i = threadIdx.x; i < (rk << job.log2n2); i += NT) {
    const int r = i >> job.log2n2, z = i & (n2 - 1);
    const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(z)) >> sh);
    dst[static_cast<std::size_t>(r) * n2 + z] = Cplx<T>{fre[pos], fim[pos]};
  }
}

// Type 2: the retained y' of k at one x' and rows z0 .. z0 + rk at their y
// positions of zero-padded lines; the transformed lines go to h.
template <typename T>
__global__ void __launch_bounds__(kRowThreads<T>) y_expand_kernel(AxisFftJob<T> job) {
  constexpr int NT = kRowThreads<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int R = job.rows_y, n1 = job.n1, n2 = job.n2, N0 = job.N0;
  const int pitch = n1 + 8;
  T* fre = reinterpret_cast<T*>(smem_raw);
  T* fim = fre + R * pitch;
  const int b = static_cast<int>(blockIdx.x);
  const int xp = b % N0;
  const int z0 = (b / N0) * R;
  const int rk = min(R, n2 - z0);
  for (int i = threadIdx.x; i < 2 * R * pitch; i += NT) fre[i] = T(0);
  __syncthreads();
  const int lr = job.log2rows_y;
  const Cplx<T>* __restrict__ src = job.k + static_cast<std::size_t>(xp) * n2 + z0;
  const std::size_t ys = static_cast<std::size_t>(N0) * n2;
  // items (r fastest, then y'); rows r >= rk repeat row rk - 1 and are dropped
  batched_copy<NT>(
      job.N1 << lr,
      [&](int i) {
        return src[static_cast<std::size_t>(i >> lr) * ys + min(i & (R - 1), rk - 1)];
      },
      [&](int i, const Cplx<T>& v) {
        const int r = i & (R - 1);
        if (r >= rk) return;
        const int o = r * pitch + ret_fine(i >> lr, n1, job.N1);
        fre[o] = v.x;
        fim[o] = v.y;
      });
  __syncthreads();
  lds_fft_x<T, NT, false>(fre, fim, rk, pitch, job.log2n1, T(1), job.tw);
  const int sh = 32 - job.log2n1;
  Cplx<T>* __restrict__ dst = job.h + static_cast<std::size_t>(xp) * n1;
  const std::size_t zs = static_cast<std::size_t>(N0) * n1;
  for (int i = threadIdx.x; i < (rk << job.log2n1); i += NT) {
    const int r = i >> job.log2n1, y = i & (n1 - 1);
    const int pos = r * pitch + static_cast<int>(__brev(static_cast<unsigned>(y)) >> sh);
    dst[static_cast<std::size_t>(z0 + r) * zs + y] = Cplx<T>{fre[pos], fim[pos]};
  }
}

}  // namespace

template <typename T, int DIM>
void launch_reduce_rows_fft(const RowFftJob<T>& job, int ns, hipStream_t stream) {
  const int groups = (job.g.n[1] + job.rows - 1) / job.rows;
  const std::size_t blocks = static_cast<std::size_t>(groups) * job.g.n[2];
  const std::size_t lds =
      row_fft_lds_bytes(job.rows, job.pitch, sizeof(T), row_fft_rng_entries(job.g, job.rows, DIM));
  hipLaunchKernelGGL((reduce_rows_fft_kernel<T, DIM>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kRowThreads<T>), lds, stream, job, ns);
  check_launch("reduce_rows_fft_kernel");
}

template <typename T>
void launch_expand_rows_ifft(const RowFftJob<T>& job, hipStream_t stream) {
  const int groups = (job.g.n[1] + job.rows - 1) / job.rows;
  const std::size_t blocks = static_cast<std::size_t>(groups) * job.g.n[2];
  const std::size_t lds = row_fft_lds_bytes(job.rows, job.pitch, sizeof(T), 0);
  hipLaunchKernelGGL((expand_rows_ifft_kernel<T>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kRowThreads<T>), lds, stream, job);
  check_launch("expand_rows_ifft_kernel");
}

template void launch_reduce_rows_fft<float, 2>(const RowFftJob<float>&, int, hipStream_t);
template void launch_reduce_rows_fft<float, 3>(const RowFftJob<float>&, int, hipStream_t);
template void launch_reduce_rows_fft<double, 2>(const RowFftJob<double>&, int, hipStream_t);
template void launch_reduce_rows_fft<double, 3>(const RowFftJob<double>&, int, hipStream_t);
template void launch_expand_rows_ifft<float>(const RowFftJob<float>&, hipStream_t);
template void launch_expand_rows_ifft<double>(const RowFftJob<double>&, hipStream_t);

template <typename T>
void launch_axis_fft_type1(const AxisFftJob<T>& job, hipStream_t stream) {
  const std::size_t by = static_cast<std::size_t>(job.N0) * ((job.n2 + job.rows_y - 1) / job.rows_y);
  const std::size_t bz = static_cast<std::size_t>(job.N1) * ((job.N0 + job.rows_z - 1) / job.rows_z);
  if (by == 0 || bz == 0) return;
  hipLaunchKernelGGL((y_keep_kernel<T>), dim3(static_cast<unsigned>(by)), dim3(kRowThreads<T>),
                     row_fft_lds_bytes(job.rows_y, job.n1 + 8, sizeof(T), 0), stream, job);
  check_launch("y_keep_kernel");
  hipLaunchKernelGGL((z_out_kernel<T>), dim3(static_cast<unsigned>(bz)), dim3(kRowThreads<T>),
                     row_fft_lds_bytes(job.rows_z, job.n2 + 8, sizeof(T), 0), stream, job);
  check_launch("z_out_kernel");
}

template <typename T>
void launch_axis_fft_type2(const AxisFftJob<T>& job, hipStream_t stream) {
  const std::size_t by = static_cast<std::size_t>(job.N0) * ((job.n2 + job.rows_y - 1) / job.rows_y);
  const std::size_t bz = static_cast<std::size_t>(job.N1) * ((job.N0 + job.rows_z - 1) / job.rows_z);
  if (by == 0 || bz == 0) return;
  hipLaunchKernelGGL((z_in_kernel<T>), dim3(static_cast<unsigned>(bz)), dim3(kRowThreads<T>),
                     row_fft_lds_bytes(job.rows_z, job.n2 + 8, sizeof(T), 0), stream, job);
  check_launch("z_in_kernel");
  hipLaunchKernelGGL((y_expand_kernel<T>), dim3(static_cast<unsigned>(by)), dim3(kRowThreads<T>),
                     row_fft_lds_bytes(job.rows_y, job.n1 + 8, sizeof(T), 0), stream, job);
  check_launch("y_expand_kernel");
}

template <typename T>
void launch_y_out2(const Y2Job<T>& job, hipStream_t stream) {
  const std::size_t blocks =
      static_cast<std::size_t>(job.count) * ((job.N0 + job.rows - 1) / job.rows);
  if (blocks == 0 || job.N1 == 0) return;
  if (blocks > 0xFFFFFFFFull) throw neonufft::InputError("Neonufft:: y pass too large");
  hipLaunchKernelGGL((y_out2_kernel<T>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kRowThreads<T>), row_fft_lds_bytes(job.rows, job.n1 + 8, sizeof(T), 0),
                     stream, job);
  check_launch("y_out2_kernel");
}

template <typename T>
void launch_y_in2(const Y2Job<T>& job, hipStream_t stream) {
  const std::size_t blocks =
      static_cast<std::size_t>(job.count) * ((job.N0 + job.rows - 1) / job.rows);
  if (blocks == 0) return;
  if (blocks > 0xFFFFFFFFull) throw neonufft::InputError("Neonufft:: y pass too large");
  hipLaunchKernelGGL((y_in2_kernel<T>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kRowThreads<T>), row_fft_lds_bytes(job.rows, job.n1 + 8, sizeof(T), 0),
                     stream, job);
  check_launch("y_in2_kernel");
}

template void launch_y_out2<float>(const Y2Job<float>&, hipStream_t);
template void launch_y_out2<double>(const Y2Job<double>&, hipStream_t);
template void launch_y_in2<float>(const Y2Job<float>&, hipStream_t);
template void launch_y_in2<double>(const Y2Job<double>&, hipStream_t);
template void launch_axis_fft_type1<float>(const AxisFftJob<float>&, hipStream_t);
template void launch_axis_fft_type1<double>(const AxisFftJob<double>&, hipStream_t);
template void launch_axis_fft_type2<float>(const AxisFftJob<float>&, hipStream_t);
template void launch_axis_fft_type2<double>(const AxisFftJob<double>&, hipStream_t);

}  // namespace nnf
