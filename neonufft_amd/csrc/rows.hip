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
// The results equal those of the full transform the reference runs
// (src/neonufft/gpu/plan_impl.hpp:91-136: spread, in-place FFT, downsample;
// upsample, in-place FFT, interpolate), up to rounding.
#include "lds_fft.hpp"

namespace nnf {

namespace {

template <typename T>
constexpr int kRowThreads = sizeof(T) == 4 ? 512 : 256;

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

  // Per dimension of a cell: own bin (as its entry index along that
  // dimension of the range table) and position in its tile, and the one
  // neighbour whose halo also covers it (entry -1 if none) with its position.
  struct Cover {
    int jown, pown, joth, poth;
  };
  const float rb0 = 1.0f / static_cast<float>(g.bin[0]);
  const float rb1 = 1.0f / static_cast<float>(g.bin[1]);
  auto cover = [&](int c, int d, float rb, int lo) {
    Cover cv;
    int bd;
    const int li = div_rem(c, g.bin[d], rb, &bd);
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
  const std::size_t tc = tile_cells(g);
  Cover cz{0, 0, -1, 0};
  if constexpr (DIM >= 3) cz = cover(pc.z, 2, 1.0f / static_cast<float>(g.bin[2]), bz_lo);

  // Threads take x fastest; with n0 <= NT several rows run side by side and
  // each thread keeps one x for the whole pencil.
  const int xthreads = min(NT, n0);
  const int rpar = NT / xthreads;
  const int tx = static_cast<int>(threadIdx.x) % xthreads;
  const int tr = static_cast<int>(threadIdx.x) / xthreads;
  const Cover cx0 = cover(tx, 0, rb0, 0);
  for (int r = tr; r < pc.rk; r += rpar) {
    const Cover cy = cover(pc.y0 + r, 1, rb1, by_lo);
    for (int x = tx; x < n0; x += xthreads) {
      const Cover cx = x == tx ? cx0 : cover(x, 0, rb0, 0);
      constexpr int kCombos = DIM >= 3 ? 8 : 4;
      Cplx<T> v[kCombos];
      int idx[kCombos];
      std::size_t tpos[kCombos];
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        const bool ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const int jx = ox ? cx.joth : cx.jown;
        const int jy = oy ? cy.joth : cy.jown;
        const int jz = oz ? cz.joth : cz.jown;
        idx[k] = (jx < 0 || jy < 0 || jz < 0) ? -1 : jx + nb0 * (jy + ny * jz);
        const int p0 = ox ? cx.poth : cx.pown;
        const int p1 = oy ? cy.poth : cy.pown;
        const int p2 = DIM >= 3 ? (oz ? cz.poth : cz.pown) : 0;
        tpos[k] = static_cast<std::size_t>(p0 + g.tile[0] * (p1 + g.tile[1] * p2));
      }
      // first subproblem of every covering bin: independent loads
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        v[k] = Cplx<T>{T(0), T(0)};
        if (idx[k] >= 0 && rng[2 * idx[k] + 1] > rng[2 * idx[k]]) {
          v[k] = job.tiles[tpos[k] + static_cast<std::size_t>(rng[2 * idx[k]]) * tc];
        }
      }
      T ar = T(0), ai = T(0);
#pragma unroll
      for (int k = 0; k < kCombos; ++k) {
        ar += v[k].x;
        ai += v[k].y;
        if (idx[k] >= 0) {  // bins split into several subproblems
          const uint32_t s1 = rng[2 * idx[k] + 1];
          for (uint32_t s = rng[2 * idx[k]] + 1; s < s1; ++s) {
            const Cplx<T> w = job.tiles[tpos[k] + static_cast<std::size_t>(s) * tc];
            ar += w.x;
            ai += w.y;
          }
        }
      }
      fre[r * job.pitch + x] = ar;
      fim[r * job.pitch + x] = ai;
    }
  }
  __syncthreads();

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
  for (int i = threadIdx.x; i < (N0 << lr); i += NT) {
    const int r = i & (job.rows - 1);
    const int xp = i >> lr;
    if (r >= pc.rk) continue;
    const int f = xp < P ? xp : xp + n0 - N0;
    const Cplx<T> v = hz[static_cast<std::size_t>(xp) * n1 + pc.y0 + r];
    fre[r * job.pitch + f] = v.x;
    fim[r * job.pitch + f] = v.y;
  }
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

}  // namespace nnf
