// mps_kernels.h -- gfx950 kernels of the explicit MPS step.
//
// Every neighbour loop visits candidates in the reference's stencil order
// (cell z, then y, then x: kernel.cu:1038-1040) and, inside a cell, in
// ascending particle id, one lane per home particle.  Sums therefore run in
// exactly the order of the CPU restatement, and with -ffp-contract=off and
// correctly rounded division and sqrt the results are bit-identical to it.
// Because x is the fastest key, the three cells jx-1..jx+1 of one (jy, jz)
// row are contiguous in a sorted set, so a 3-D stencil is 9 ranges.
#pragma once
#include "mps_device.h"

namespace mps {

__device__ inline int wave_alloc(int* counter) {
  // One slot per active lane, one atomic per wavefront.
  const unsigned long long mask = __ballot(1);
  const int leader = __ffsll((long long)mask) - 1;
  const int lane = __lane_id();
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader);
  return base + __popcll(mask & ((1ull << lane) - 1ull));
}

__global__ void k_fill_int(int* a, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

template <class T>
__global__ void k_fill(T* a, int n, T v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}

// ---------------------------------------------------------------------------
// Exclusive scan of int counts (cell occupancy -> cell_start).
// Tile = kBlock threads x kScanIPT items.
constexpr int kScanIPT = 8;
constexpr int kScanTile = kBlock * kScanIPT;

__device__ inline int block_exclusive_scan(int v, int* sh, int& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int x = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  total = sh[kBlock - 1];
  int excl = sh[t] - v;
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int* in, int n, int* part) {
  __shared__ int sh[kBlock];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanIPT;
  int s = 0;
  for (int k = 0; k < kScanIPT; k++)
    if (base + k < n) s += in[base + k];
  int total;
  block_exclusive_scan(s, sh, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// out[i] = sum(in[0..i)) + offset of the tile; out[n] = total when last.
__global__ __launch_bounds__(kBlock) void k_scan_down(const int* in, int n, const int* part_excl,
                                                     int* out) {
  __shared__ int sh[kBlock];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanIPT;
  int v[kScanIPT];
  int s = 0;
  for (int k = 0; k < kScanIPT; k++) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  int total;
  int run = block_exclusive_scan(s, sh, total) + (part_excl ? part_excl[blockIdx.x] : 0);
  for (int k = 0; k < kScanIPT; k++) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (base <= n && n < base + kScanIPT) out[n] = run;  // grand total
}

// ---------------------------------------------------------------------------
// near[k] = 1 when any entry of the set with cell_start cs lies in the
// 27-bucket (9 in 2-D) stencil of bucket k.  Lets a particle skip a stencil
// pass over a set that is empty around it.
template <class real, int DIM>
__global__ void k_cell_near(const int* __restrict__ cs, Cst<real> c, uint8_t* __restrict__ near) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= c.ncell) return;
  const int iz = DIM == 3 ? k / c.nbxy : 0;
  const int rem = k - iz * c.nbxy;
  const int iy = rem / c.nb[0], ix = rem - iy * c.nb[0];
  if (ix < 1 || ix > c.nb[0] - 2 || iy < 1 || iy > c.nb[1] - 2 ||
      (DIM == 3 && (iz < 1 || iz > c.nb[2] - 2))) {
    near[k] = 0;  // no particle is ever assigned to a border bucket
    return;
  }
  bool any = false;
  const int zlo = DIM == 3 ? iz - 1 : 0, zhi = DIM == 3 ? iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = iy - 1; jy <= iy + 1; jy++)
      any = any || cs[cell_key(ix - 1, jy, jz, c)] != cs[cell_key(ix + 1, jy, jz, c) + 1];
  near[k] = any ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Bucket build: key + count, scatter, rank inside the cell by id, gather.

template <class real, int DIM>
__global__ void k_bkt_key(const Vec<real, DIM>* __restrict__ pos, const int8_t* __restrict__ typ,
                          int n_static, const int* __restrict__ n_dyn, Cst<real> c,
                          int* __restrict__ key, int* __restrict__ slot, int* __restrict__ cnt) {
  const int n = n_dyn ? *n_dyn : n_static;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int k;
    if (typ && typ[i] == kGST) {
      k = c.ncell;
    } else {
      Cell<real, DIM> q = cell_of<real, DIM>(pos[i], c);
      k = cell_key(q.ix, q.iy, q.iz, c);
    }
    key[i] = k;
    slot[i] = atomicAdd(&cnt[k], 1);
  }
}

// Element count: n_static, or mult * (*n_dyn) when n_dyn is given.
__global__ void k_bkt_scatter(const int* __restrict__ key, const int* __restrict__ slot,
                              const int* __restrict__ cs, int n_static, const int* __restrict__ n_dyn,
                              int mult, int* __restrict__ tmp) {
  const int n = n_dyn ? *n_dyn * mult : n_static;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int k = key[i];
    if (k >= 0) tmp[cs[k] + slot[i]] = i;
  }
}

// Final position of an element = start of its cell + number of elements of
// the same cell with a smaller id.  The deleted tail (key == ncell) keeps the
// scatter order: its order is never observed.
__global__ void k_bkt_rank(const int* __restrict__ tmp, const int* __restrict__ key,
                           const int* __restrict__ id, const int* __restrict__ cs, int n_static,
                           const int* __restrict__ n_dyn, int mult, int ncell,
                           int* __restrict__ perm) {
  const int n = n_dyn ? *n_dyn * mult : n_static;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int i = tmp[s];
    const int k = key[i];
    if (k == ncell) {
      perm[s] = i;
      continue;
    }
    const int b = cs[k], e = cs[k + 1];
    const int me = id[i];
    int rank = 0;
    for (int t = b; t < e; t++) rank += id[tmp[t]] < me;
    perm[b + rank] = i;
  }
}

template <class real, int DIM>
struct FluidArrays {
  Vec<real, DIM>* pos;
  Vec<real, DIM>* vel;
  Vec<real, DIM>* acc;  // may be null
  real* prs;
  real* pav;
  real* nd;
  int8_t* typ;
  int* id;
};

template <class real, int DIM>
__global__ void k_gather_fluid(const int* __restrict__ perm, int n, FluidArrays<real, DIM> src,
                               FluidArrays<real, DIM> dst) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int i = perm[s];
    dst.pos[s] = src.pos[i];
    dst.vel[s] = src.vel[i];
    if (src.acc) dst.acc[s] = src.acc[i];
    dst.prs[s] = src.prs[i];
    dst.pav[s] = src.pav[i];
    dst.nd[s] = src.nd[i];
    dst.typ[s] = src.typ[i];
    dst.id[s] = src.id[i];
  }
}

template <class real, int DIM>
__global__ void k_gather_mirror(const int* __restrict__ perm, const int* __restrict__ n_dyn,
                                const Vec<real, DIM>* __restrict__ spos,
                                const real* __restrict__ sprs, const int* __restrict__ sid,
                                Vec<real, DIM>* __restrict__ dpos, real* __restrict__ dprs,
                                int* __restrict__ did) {
  const int n = *n_dyn;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int i = perm[s];
    dpos[s] = spos[i];
    dprs[s] = sprs[i];
    did[s] = sid[i];
  }
}

// ---------------------------------------------------------------------------
// Mirror generation + cull: d_GenMRR_nonslip (kernel.cu:720-893) followed by
// d_MRRinout (kernel.cu:932-985) for each generated mirror; survivors are
// appended to the raw mirror set.

// Nearest wall particle within r of a mirror; the first met wins a tie
// (kernel.cu:965-971).  A candidate can only pass `dist <= r && dist <
// walldist` when d2 <= r2_hi and d2 <= d2 of the current best (sqrt is
// correctly rounded, hence monotone), so sqrt runs only for those.
template <class real, int DIM>
__device__ inline bool mirror_culled(real mx, real my, real mz, const Vec<real, DIM>* __restrict__ wpos,
                                     const Vec<real, DIM>* __restrict__ wnrm,
                                     const int* __restrict__ wcs, const Cst<real>& c) {
  Vec<real, DIM> m = mkvec<real, DIM>(mx, my, mz);
  Cell<real, DIM> q = cell_of<real, DIM>(m, c);
  bool found = false;
  real walldist = real(10) * c.r;
  real lim = c.r2_hi;
  real wpx = 0, wpy = 0, wpz = 0, wx = 0, wy = 0, wz = 0;
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = wcs[cell_key(q.ix - 1, jy, jz, c)], e = wcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for (int j = b; j < e; j++) {
        const Vec<real, DIM> w = wpos[j];
        const real jx = w.x, jy2 = w.y, jz2 = vz(w);
        const real d2 = norm2<real, DIM>(mx - jx, my - jy2, mz - jz2);
        if (!(d2 <= lim)) continue;
        real dist = sqrt(d2);
        if (dist <= c.r && dist < walldist) {
          lim = d2;
          found = true;
          walldist = dist;
          const Vec<real, DIM> n = wnrm[j];
          wx = n.x;
          wy = n.y;
          wz = vz(n);
          wpx = jx;
          wpy = jy2;
          wpz = jz2;
        }
      }
    }
  real inout = dot3<real, DIM>(mx - wpx, my - wpy, mz - wpz, wx, wy, wz);
  return found && inout >= real(0);
}

// Plane mirror of (px,py,pz) across the plane of wall w with normal n
// (kernel.cu:815-836); false when the particle lies behind the plane.
template <class real, int DIM>
__device__ inline bool plane_mirror(real px, real py, real pz, const Vec<real, DIM>& w,
                                    const Vec<real, DIM>& n, real& mx, real& my, real& mz) {
  const real wpx = w.x, wpy = w.y, wpz = vz(w);
  const real nx = n.x, ny = n.y, nz = vz(n);
  const real PWx = px - wpx, PWy = py - wpy, PWz = pz - wpz;
  const real PW_WLL = dot3<real, DIM>(PWx, PWy, PWz, nx, ny, nz);
  if (PW_WLL < real(0)) return false;
  const real distance = PW_WLL / sqrt(norm2<real, DIM>(nx, ny, nz));
  mx = wpx + PWx - real(2) * distance * nx;
  my = wpy + PWy - real(2) * distance * ny;
  mz = DIM == 3 ? wpz + PWz - real(2) * distance * nz : real(0);
  return true;
}

template <class real, int DIM>
__device__ inline void emit_mirror(real mx, real my, real mz, real prs, int id,
                                   Vec<real, DIM>* __restrict__ mpos, real* __restrict__ mprs,
                                   int* __restrict__ mid, int* __restrict__ mcount, int mcap) {
  const int slot = wave_alloc(mcount);
  if (slot < mcap) {
    mpos[slot] = mkvec<real, DIM>(mx, my, mz);
    mprs[slot] = prs;
    mid[slot] = id;
  }
}

// ---------------------------------------------------------------------------
// Wave-parallel mirror generation (at most 64 wall-normal classes).
//
// k_mirror_near lists the live fluid particles whose stencil holds any wall.
// k_mirror_wave then gives each listed particle one wavefront:
//  * the lanes split the wall candidates of the stencil; in the reference's
//    scan order (z, y, x cells, then id) the candidates appear in increasing
//    wall-array index, so "the first wall met of a normal within r"
//    (kernel.cu:758-775) is the smallest index of that class within r: an
//    LDS atomicMin per class.  Slot k of the reference's FromWLL registry is
//    the class with the k-th smallest first index.  Edge_unique
//    (kernel.cu:777-795) is the same over Edge walls.
//  * each plane mirror and the edge mirror is culled by d_MRRinout's
//    nearest-wall test (kernel.cu:944-983): the lanes split the candidates,
//    keep the first minimum of their share, and a wave reduction picks the
//    smallest (dist, index), which is the reference's first-met minimum.
//  * slot k of listed particle t goes to raw entry 10*t + k (id -1 if empty).

constexpr int kInf = 0x7fffffff;

template <class real, int DIM>
__global__ __launch_bounds__(kBlock) void k_mirror_near(const Vec<real, DIM>* __restrict__ fpos,
                                                       const int* __restrict__ fcs, int n_alloc,
                                                       const uint8_t* __restrict__ wnear, Cst<real> c,
                                                       int* __restrict__ near_list,
                                                       int* __restrict__ near_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc || i >= fcs[c.ncell]) return;
  const Cell<real, DIM> q = cell_of<real, DIM>(fpos[i], c);
  if (wnear[cell_key(q.ix, q.iy, q.iz, c)]) near_list[wave_alloc(near_count)] = i;
}

template <class real>
struct BestWall {
  real dist;
  int j;
};

// Wave-wide nearest wall within r of m (first met wins ties); j = -1 if none.
template <class real, int DIM>
__device__ inline BestWall<real> wave_nearest_wall(real mx, real my, real mz,
                                                   const Vec<real, DIM>* __restrict__ wpos,
                                                   const int* __restrict__ wcs, const Cst<real>& c,
                                                   int lane) {
  const Cell<real, DIM> q = cell_of<real, DIM>(mkvec<real, DIM>(mx, my, mz), c);
  real best = real(10) * c.r;
  real lim = c.r2_hi;
  int bj = -1;
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = wcs[cell_key(q.ix - 1, jy, jz, c)], e = wcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for (int j = b + lane; j < e; j += 64) {
        const Vec<real, DIM> w = wpos[j];
        const real d2 = norm2<real, DIM>(mx - w.x, my - w.y, mz - vz(w));
        if (!(d2 <= lim)) continue;
        const real dist = sqrt(d2);
        if (dist <= c.r && dist < best) {
          best = dist;
          bj = j;
          lim = d2;
        }
      }
    }
  // lexicographic (dist, j) minimum over the wave; lanes without a hit hold j = -1
  for (int off = 32; off >= 1; off >>= 1) {
    const real od = __shfl_xor(best, off);
    const int oj = __shfl_xor(bj, off);
    const bool take = oj >= 0 && (bj < 0 || od < best || (od == best && oj < bj));
    if (take) {
      best = od;
      bj = oj;
    }
  }
  return {best, bj};
}

template <class real, int DIM>
__global__ __launch_bounds__(64) void k_mirror_wave(
    const Vec<real, DIM>* __restrict__ fpos, const real* __restrict__ fprs,
    const int* __restrict__ fid, const int* __restrict__ near_list,
    const int* __restrict__ near_count, const Vec<real, DIM>* __restrict__ wpos,
    const Vec<real, DIM>* __restrict__ wnrm, const uint8_t* __restrict__ wcls,
    const int* __restrict__ wcs, Cst<real> c, Vec<real, DIM>* __restrict__ mpos,
    real* __restrict__ mprs, int* __restrict__ mid) {
  __shared__ int first_reg[64], first_edge[64], sorted_reg[kNumMRR], sorted_edge[kNumMRR];
  __shared__ Vec<real, DIM> s_pos[kNumMRR];
  __shared__ real s_prs[kNumMRR];
  __shared__ int s_id[kNumMRR];
  const int lane = threadIdx.x;
  const int nnear = *near_count;
  for (int t = blockIdx.x; t < nnear; t += gridDim.x) {
    const int i = near_list[t];
    const Vec<real, DIM> p = fpos[i];
    const real px = p.x, py = p.y, pz = vz(p);
    const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
    first_reg[lane] = kInf;
    first_edge[lane] = kInf;
    if (lane < kNumMRR) s_id[lane] = -1;
    __syncthreads();
    const real r2_2 = real(1) * c.r2;
    const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
    for (int jz = zlo; jz <= zhi; jz++)
      for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
        const int b = wcs[cell_key(q.ix - 1, jy, jz, c)], e = wcs[cell_key(q.ix + 1, jy, jz, c) + 1];
        for (int j = b + lane; j < e; j += 64) {
          const Vec<real, DIM> w = wpos[j];
          if (norm2<real, DIM>(px - w.x, py - w.y, pz - vz(w)) <= r2_2) {
            const int cl = wcls[j];
            atomicMin(&first_reg[cl & 63], j);
            if (cl & 0x80) atomicMin(&first_edge[cl & 63], j);
          }
        }
      }
    __syncthreads();
    // registry order: rank of each class's first index
    const int fr = first_reg[lane], fe = first_edge[lane];
    int rr = 0, re_ = 0;
    for (int k = 0; k < 64; k++) {
      rr += first_reg[k] < fr;
      re_ += first_edge[k] < fe;
    }
    if (fr != kInf && rr < kNumMRR) sorted_reg[rr] = fr;
    if (fe != kInf && re_ < kNumMRR) sorted_edge[re_] = fe;
    const int nreg = min(__popcll(__ballot(fr != kInf)), kNumMRR);
    const int ne = min(__popcll(__ballot(fe != kInf)), kNumMRR);
    __syncthreads();
    const int base_id = fid[i] * kNumMRR;
    const real prs_i = fprs[i];
    for (int k = 0; k < nreg; k++) {  // plane mirrors, kernel.cu:811-838
      const int F = sorted_reg[k];
      real mx, my, mz;
      if (!plane_mirror<real, DIM>(px, py, pz, wpos[F], wnrm[F], mx, my, mz)) continue;
      const BestWall<real> bw = wave_nearest_wall<real, DIM>(mx, my, mz, wpos, wcs, c, lane);
      bool culled = false;
      if (bw.j >= 0) {
        const Vec<real, DIM> w = wpos[bw.j], n = wnrm[bw.j];
        culled = dot3<real, DIM>(mx - w.x, my - w.y, mz - vz(w), n.x, n.y, vz(n)) >= real(0);
      }
      if (!culled && lane == 0) {
        s_pos[k] = mkvec<real, DIM>(mx, my, mz);
        s_prs[k] = prs_i;
        s_id[k] = base_id + k;
      }
    }
    if (ne >= 2 && nreg < kNumMRR) {  // edge mirror, kernel.cu:840-888
      real sx = 0, sy = 0, sz = 0;
      for (int k = 0; k < ne; k++) {
        const Vec<real, DIM> n = wnrm[sorted_edge[k]];
        sx += n.x;
        sy += n.y;
        sz += vz(n);
      }
      const Vec<real, DIM> w = wpos[sorted_edge[ne - 1]];
      const real wpx = w.x, wpy = w.y, wpz = vz(w);
      const real PWx = px - wpx, PWy = py - wpy, PWz = pz - wpz;
      const real abs_s = sqrt(norm2<real, DIM>(sx, sy, sz));
      sx /= abs_s;
      sy /= abs_s;
      sz /= abs_s;
      const real PW_WLL = dot3<real, DIM>(PWx, PWy, PWz, sx, sy, sz);
      if (PW_WLL >= real(0)) {
        const real distance = PW_WLL / sqrt(norm2<real, DIM>(sx, sy, sz));
        const real mx = wpx + PWx - real(2) * distance * sx;
        const real my = wpy + PWy - real(2) * distance * sy;
        const real mz = DIM == 3 ? wpz + PWz - real(2) * distance * sz : real(0);
        const BestWall<real> bw = wave_nearest_wall<real, DIM>(mx, my, mz, wpos, wcs, c, lane);
        bool culled = false;
        if (bw.j >= 0) {
          const Vec<real, DIM> w2 = wpos[bw.j], n = wnrm[bw.j];
          culled = dot3<real, DIM>(mx - w2.x, my - w2.y, mz - vz(w2), n.x, n.y, vz(n)) >= real(0);
        }
        if (!culled && lane == 0) {
          s_pos[nreg] = mkvec<real, DIM>(mx, my, mz);
          s_prs[nreg] = prs_i;
          s_id[nreg] = base_id + nreg;
        }
      }
    }
    __syncthreads();
    if (lane < kNumMRR) {
      const size_t o = (size_t)t * kNumMRR + lane;
      mid[o] = s_id[lane];
      if (s_id[lane] >= 0) {
        mpos[o] = s_pos[lane];
        mprs[o] = s_prs[lane];
      }
    }
    __syncthreads();
  }
}

// Bucket key of a raw mirror slot; empty slots (id < 0) get key -1 and are
// neither counted nor scattered.
template <class real, int DIM>
__global__ void k_mirror_key(const Vec<real, DIM>* __restrict__ pos, const int* __restrict__ id,
                             const int* __restrict__ n_dyn, int mult, Cst<real> c,
                             int* __restrict__ key, int* __restrict__ slot, int* __restrict__ cnt) {
  const int n = *n_dyn * mult;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (id[i] < 0) {
      key[i] = -1;
      continue;
    }
    Cell<real, DIM> q = cell_of<real, DIM>(pos[i], c);
    const int k = cell_key(q.ix, q.iy, q.iz, c);
    key[i] = k;
    slot[i] = atomicAdd(&cnt[k], 1);
  }
}

// General mirror generation (any number of distinct wall normals).
template <class real, int DIM>
__global__ __launch_bounds__(kBlock) void k_mirror_gen(
    const Vec<real, DIM>* __restrict__ fpos, const real* __restrict__ fprs,
    const int* __restrict__ fid, const int* __restrict__ fcs, int n_alloc,
    const Vec<real, DIM>* __restrict__ wpos, const Vec<real, DIM>* __restrict__ wnrm,
    const int8_t* __restrict__ wedge, const int* __restrict__ wcs, Cst<real> c,
    Vec<real, DIM>* __restrict__ mpos, real* __restrict__ mprs, int* __restrict__ mid,
    int* __restrict__ mcount, int mcap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc) return;
  const int n_live = fcs[c.ncell];
  if (i >= n_live) return;
  const Vec<real, DIM> p = fpos[i];
  const real px = p.x, py = p.y, pz = vz(p);
  const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
  const real r2_2 = real(1) * c.r2;
  int reg[kNumMRR], edg[kNumMRR];
  int nreg = 0, ne = 0;
  bool wexist = false, eexist = false;
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = wcs[cell_key(q.ix - 1, jy, jz, c)], e = wcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for (int j = b; j < e; j++) {
        const Vec<real, DIM> w = wpos[j];
        const real d2 = norm2<real, DIM>(px - w.x, py - w.y, pz - vz(w));
        if (d2 <= r2_2) {
          wexist = true;
          const Vec<real, DIM> n = wnrm[j];
          bool dup = false;
          for (int k = 0; k < nreg; k++) {
            const Vec<real, DIM> nk = wnrm[reg[k]];
            if (nk.x == n.x && nk.y == n.y && vz(nk) == vz(n)) dup = true;
          }
          if (!dup && nreg < kNumMRR) reg[nreg++] = j;
          if (wedge[j] == kEdge) {
            eexist = true;
            bool edup = false;
            for (int k = 0; k < ne; k++) {
              const Vec<real, DIM> nk = wnrm[edg[k]];
              if (nk.x == n.x && nk.y == n.y && vz(nk) == vz(n)) edup = true;
            }
            if (!edup && ne < kNumMRR) edg[ne++] = j;
          }
        }
      }
    }
  if (!wexist) return;
  const real prs_i = fprs[i];
  const int base_id = fid[i] * kNumMRR;
  for (int k = 0; k < nreg; k++) {  // plane mirrors, kernel.cu:811-838
    const Vec<real, DIM> w = wpos[reg[k]];
    const Vec<real, DIM> n = wnrm[reg[k]];
    const real wpx = w.x, wpy = w.y, wpz = vz(w);
    const real nx = n.x, ny = n.y, nz = vz(n);
    const real PWx = px - wpx, PWy = py - wpy, PWz = pz - wpz;
    const real PW_WLL = dot3<real, DIM>(PWx, PWy, PWz, nx, ny, nz);
    if (PW_WLL < real(0)) continue;
    const real distance = PW_WLL / sqrt(norm2<real, DIM>(nx, ny, nz));
    const real mx = wpx + PWx - real(2) * distance * nx;
    const real my = wpy + PWy - real(2) * distance * ny;
    const real mz = DIM == 3 ? wpz + PWz - real(2) * distance * nz : real(0);
    if (mirror_culled<real, DIM>(mx, my, mz, wpos, wnrm, wcs, c)) continue;
    const int slot = wave_alloc(mcount);
    if (slot < mcap) {
      mpos[slot] = mkvec<real, DIM>(mx, my, mz);
      mprs[slot] = prs_i;
      mid[slot] = base_id + k;
    }
  }
  if (eexist && ne >= 2 && nreg < kNumMRR) {  // edge mirror, kernel.cu:840-888
    real wpx = 0, wpy = 0, wpz = 0, sx = 0, sy = 0, sz = 0, PWx = 0, PWy = 0, PWz = 0;
    for (int k = 0; k < ne; k++) {
      const Vec<real, DIM> w = wpos[edg[k]];
      const Vec<real, DIM> n = wnrm[edg[k]];
      wpx = w.x;
      wpy = w.y;
      wpz = vz(w);
      sx += n.x;
      sy += n.y;
      sz += vz(n);
      PWx = px - wpx;
      PWy = py - wpy;
      PWz = pz - wpz;
    }
    const real abs_s = sqrt(norm2<real, DIM>(sx, sy, sz));
    sx /= abs_s;
    sy /= abs_s;
    sz /= abs_s;
    const real PW_WLL = dot3<real, DIM>(PWx, PWy, PWz, sx, sy, sz);
    if (PW_WLL < real(0)) return;
    const real distance = PW_WLL / sqrt(norm2<real, DIM>(sx, sy, sz));
    const real mx = wpx + PWx - real(2) * distance * sx;
    const real my = wpy + PWy - real(2) * distance * sy;
    const real mz = DIM == 3 ? wpz + PWz - real(2) * distance * sz : real(0);
    if (mirror_culled<real, DIM>(mx, my, mz, wpos, wnrm, wcs, c)) return;
    const int slot = wave_alloc(mcount);
    if (slot < mcap) {
      mpos[slot] = mkvec<real, DIM>(mx, my, mz);
      mprs[slot] = prs_i;
      mid[slot] = base_id + nreg;
    }
  }
}

// ---------------------------------------------------------------------------
// d_ChkPcl (kernel.cu:481-513) on a particle's new state.  Returns false
// when the particle is deleted (then every field is zeroed).
template <class real, int DIM>
__device__ inline bool chk_pcl(real& x, real& y, real& z, real& vx, real& vy, real& vz_,
                               const Cst<real>& c) {
  bool in = x < c.hi[0] && x > c.lo[0] && y < c.hi[1] && y > c.lo[1];
  if constexpr (DIM == 3) in = in && z < c.hi[2] && z > c.lo[2];
  if (in) {
    real U = sqrt(norm2<real, DIM>(vx, vy, vz_));
    if (U > c.umax) {
      vx *= c.umax / U;
      vy *= c.umax / U;
      vz_ *= c.umax / U;
    }
    return true;
  }
  x = y = z = vx = vy = vz_ = real(0);
  return false;
}

// ---------------------------------------------------------------------------
// Visits, in ascending order, the candidates j of [b, e) that pass test(j)
// and calls body(j) on each.  The cheap test runs over a chunk of 64
// candidates first and the hits are kept in a bit mask; the expensive body
// then runs over the hits only.  Without this, the lanes of a wavefront
// (each with its own hits) would together take the expensive branch for
// almost every candidate.  The visiting order is unchanged.
template <class T, class Body>
__device__ inline void for_hits(int b, int e, T&& test, Body&& body) {
  for (int base = b; base < e; base += 64) {
    const int n = min(64, e - base);
    unsigned long long mask = 0;
    for (int t = 0; t < n; t++)
      if (test(base + t)) mask |= 1ull << t;
    while (mask) {
      const int t = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      body(base + t);
    }
  }
}

// ---------------------------------------------------------------------------
// d_VscTrm (kernel.cu:1013-1119), optionally fused with d_UpPcl1
// (kernel.cu:1138-1149).
template <class real, int DIM, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_vsc(
    const Vec<real, DIM>* __restrict__ pos, const Vec<real, DIM>* __restrict__ vel,
    const int* __restrict__ fcs, int n_alloc, const Vec<real, DIM>* __restrict__ mpos,
    const int* __restrict__ mcs, const uint8_t* __restrict__ mnear, Cst<real> c,
    Vec<real, DIM>* __restrict__ acc_out, Vec<real, DIM>* __restrict__ pos_out,
    Vec<real, DIM>* __restrict__ vel_out, int8_t* __restrict__ typ, real* __restrict__ prs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc) return;
  const int n_live = fcs[c.ncell];
  if (i >= n_live) {
    if constexpr (FUSE) {
      pos_out[i] = pos[i];
      vel_out[i] = vel[i];
    }
    return;
  }
  const Vec<real, DIM> p = pos[i], v = vel[i];
  const real px = p.x, py = p.y, pz = vz(p);
  const real vx = v.x, vy = v.y, vvz = vz(v);
  real ax = 0, ay = 0, az = 0;
  const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  const real l0 = c.l0, r = c.r;
  const real r2_hi = c.r2_hi;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = fcs[cell_key(q.ix - 1, jy, jz, c)], e = fcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            return j != i && norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) <= r2_hi;
          },
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
            const real dist = sqrt(norm2<real, DIM>(qx, qy, qz));
            if (dist <= r) {
              const Vec<real, DIM> vj = vel[j];
              ax += real(2) * (vj.x - vx) * l0 / dist / dist / dist;
              ay += real(2) * (vj.y - vy) * l0 / dist / dist / dist;
              if constexpr (DIM == 3) az += real(2) * (vj.z - vvz) * l0 / dist / dist / dist;
            }
          });
    }
  if (mnear[cell_key(q.ix, q.iy, q.iz, c)])
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = mcs[cell_key(q.ix - 1, jy, jz, c)], e = mcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            return norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) <= r2_hi;
          },
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
            const real dist = sqrt(norm2<real, DIM>(qx, qy, qz));
            if (dist <= r) {
              ax += real(2) * (real(0) - vx) * l0 / dist / dist / dist;
              ay += real(2) * (real(0) - vy) * l0 / dist / dist / dist;
              if constexpr (DIM == 3) az += real(2) * (real(0) - vvz) * l0 / dist / dist / dist;
            }
          });
    }
  const real accx = c.n0knm * ax + c.g[0];
  const real accy = c.n0knm * ay + c.g[1];
  const real accz = DIM == 3 ? c.n0knm * az + c.g[2] : real(0);
  if constexpr (!FUSE) {
    acc_out[i] = mkvec<real, DIM>(accx, accy, accz);
  } else {
    real nvx = vx + accx * c.dt, nvy = vy + accy * c.dt, nvz = DIM == 3 ? vvz + accz * c.dt : real(0);
    real nx = px + nvx * c.dt, ny = py + nvy * c.dt, nz = DIM == 3 ? pz + nvz * c.dt : real(0);
    if (!chk_pcl<real, DIM>(nx, ny, nz, nvx, nvy, nvz, c)) {
      typ[i] = kGST;
      prs[i] = real(0);
    }
    pos_out[i] = mkvec<real, DIM>(nx, ny, nz);
    vel_out[i] = mkvec<real, DIM>(nvx, nvy, nvz);
  }
}

// d_UpPcl1 alone (phase mode): v += a dt; x += v dt; a = 0; ChkPcl.
template <class real, int DIM>
__global__ void k_up1(Vec<real, DIM>* __restrict__ pos, Vec<real, DIM>* __restrict__ vel,
                      Vec<real, DIM>* __restrict__ acc, int8_t* __restrict__ typ,
                      real* __restrict__ prs, int n, Cst<real> c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || typ[i] != kFLD) return;
  const Vec<real, DIM> a = acc[i], p = pos[i], v = vel[i];
  real nvx = v.x + a.x * c.dt, nvy = v.y + a.y * c.dt, nvz = DIM == 3 ? vz(v) + vz(a) * c.dt : real(0);
  real nx = p.x + nvx * c.dt, ny = p.y + nvy * c.dt, nz = DIM == 3 ? vz(p) + nvz * c.dt : real(0);
  acc[i] = mkvec<real, DIM>(0, 0, 0);
  if (!chk_pcl<real, DIM>(nx, ny, nz, nvx, nvy, nvz, c)) {
    typ[i] = kGST;
    prs[i] = real(0);
  }
  pos[i] = mkvec<real, DIM>(nx, ny, nz);
  vel[i] = mkvec<real, DIM>(nvx, nvy, nvz);
}

// ---------------------------------------------------------------------------
// d_ChkCol + d_UpChkCol (kernel.cu:1168-1287).  Jacobi: reads vel, writes
// vel_out (and acc_out in phase mode).  Inside each cell the fluid
// particles come before the wall particles, as in the restated lists.
template <class real, int DIM, bool PHASE>
__global__ __launch_bounds__(kBlock) void k_col(
    const Vec<real, DIM>* __restrict__ pos, const Vec<real, DIM>* __restrict__ vel,
    const int* __restrict__ fcs, int n_alloc, const Vec<real, DIM>* __restrict__ wpos,
    const Vec<real, DIM>* __restrict__ wvel, const int* __restrict__ wcs,
    const Vec<real, DIM>* __restrict__ mpos, const int* __restrict__ mcs, Cst<real> c,
    Vec<real, DIM>* __restrict__ vel_out, Vec<real, DIM>* __restrict__ acc_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc) return;
  const int n_live = fcs[c.ncell];
  if (i >= n_live) {
    vel_out[i] = vel[i];
    return;
  }
  const Vec<real, DIM> p = pos[i], v = vel[i];
  const real px = p.x, py = p.y, pz = vz(p);
  const real vx = v.x, vy = v.y, vvz = vz(v);
  real v2x = vx, v2y = vy, v2z = vvz;
  const real mi = c.mass;
  const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++)
      for (int jx = q.ix - 1; jx <= q.ix + 1; jx++) {
        const int ck = cell_key(jx, jy, jz, c);
        for (int pass = 0; pass < 2; pass++) {
          const int b = pass == 0 ? fcs[ck] : wcs[ck];
          const int e = pass == 0 ? fcs[ck + 1] : wcs[ck + 1];
          const Vec<real, DIM>* __restrict__ P = pass == 0 ? pos : wpos;
          const Vec<real, DIM>* __restrict__ V = pass == 0 ? vel : wvel;
          const real lim2 = pass == 0 ? c.rlim2 : c.rlim_wall2;
          const int self = pass == 0 ? i : -1;
          for_hits(
              b, e,
              [&](int j) {
                const Vec<real, DIM> pj = P[j];
                return j != self && norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) < lim2;
              },
              [&](int j) {
                const Vec<real, DIM> pj = P[j];
                const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
                const real dist2 = norm2<real, DIM>(qx, qy, qz);
                const Vec<real, DIM> vj = V[j];
                real fDT = dot3<real, DIM>(vx - vj.x, vy - vj.y, vvz - vz(vj), qx, qy, qz);
                if (fDT > real(0)) {
                  const real mj = c.mass;
                  fDT *= c.col * mj / (mi + mj) / dist2;
                  v2x -= qx * fDT;
                  v2y -= qy * fDT;
                  if constexpr (DIM == 3) v2z -= qz * fDT;
                }
              });
        }
      }
  const real mj = c.mass;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = mcs[cell_key(q.ix - 1, jy, jz, c)], e = mcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            return norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) < c.rlim2;
          },
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
            const real dist2 = norm2<real, DIM>(qx, qy, qz);
            real fDT = dot3<real, DIM>(vx - real(0), vy - real(0), vvz - real(0), qx, qy, qz);
            if (fDT > real(0)) {
              fDT *= c.col * mj / (mi + mj) / dist2;
              v2x -= qx * fDT;
              v2y -= qy * fDT;
              if constexpr (DIM == 3) v2z -= qz * fDT;
            }
          });
    }
  const Vec<real, DIM> out = mkvec<real, DIM>(v2x, v2y, v2z);
  vel_out[i] = out;
  if constexpr (PHASE) acc_out[i] = out;
}

// ---------------------------------------------------------------------------
// d_MkPrs (kernel.cu:1306-1411).  Writes prs and the number density nd.
template <class real, int DIM>
__global__ __launch_bounds__(kBlock) void k_prs(const Vec<real, DIM>* __restrict__ pos,
                                                 const int* __restrict__ fcs, int n_alloc,
                                                 const Vec<real, DIM>* __restrict__ mpos,
                                                 const int* __restrict__ mcs, Cst<real> c,
                                                 real* __restrict__ prs, real* __restrict__ nd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc) return;
  const int n_live = fcs[c.ncell];
  if (i >= n_live) return;
  const Vec<real, DIM> p = pos[i];
  const real px = p.x, py = p.y, pz = vz(p);
  const real rp = c.rp;
  real ni = 0;
  const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = fcs[cell_key(q.ix - 1, jy, jz, c)], e = fcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            return j != i && norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) <= c.rp2_hi;
          },
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            const real dist = sqrt(norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz));
            if (dist < rp) ni += (rp / dist) - (dist / rp);
          });
    }
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = mcs[cell_key(q.ix - 1, jy, jz, c)], e = mcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            return norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) <= c.rp2_hi;
          },
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            const real dist = sqrt(norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz));
            if (dist < rp) ni += (rp / dist) - (dist / rp);
          });
    }
  real pressure = real(ni > c.n0_grad) * (ni - c.n0_grad) * c.prs_coef * c.mass;
  if (pressure <= c.pmax && pressure >= real(0)) {
  } else {
    if (pressure > c.pmax) pressure = c.pmax;
    else pressure = real(0);
  }
  prs[i] = pressure;
  nd[i] = ni;
}

// ---------------------------------------------------------------------------
// d_PrsGrdTrm, active version (kernel.cu:1651-1794), optionally fused with
// d_UpPcl2 (kernel.cu:1815-1829).  The fused form writes pos/vel/prs to the
// other buffer (neighbours still read the old prs) and adds prs to pav.
template <class real, int DIM, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_grad(
    const Vec<real, DIM>* __restrict__ pos, const Vec<real, DIM>* __restrict__ vel,
    const real* __restrict__ prs, const int* __restrict__ fcs, int n_alloc,
    const Vec<real, DIM>* __restrict__ mpos, const real* __restrict__ mprs,
    const int* __restrict__ mcs, Cst<real> c, Vec<real, DIM>* __restrict__ acc_out,
    Vec<real, DIM>* __restrict__ pos_out, Vec<real, DIM>* __restrict__ vel_out,
    real* __restrict__ prs_out, real* __restrict__ pav, int8_t* __restrict__ typ) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_alloc) return;
  const int n_live = fcs[c.ncell];
  if (i >= n_live) {
    if constexpr (FUSE) {
      pos_out[i] = pos[i];
      vel_out[i] = vel[i];
      prs_out[i] = prs[i];
      pav[i] += prs[i];
    }
    return;
  }
  const Vec<real, DIM> p = pos[i];
  const real px = p.x, py = p.y, pz = vz(p);
  const real rp = c.rp, rp2 = c.rp2;
  const real prs_i = prs[i];
  real Prs_min = prs_i;
  real ax = 0, ay = 0, az = 0;
  const Cell<real, DIM> q = cell_of<real, DIM>(p, c);
  const int zlo = DIM == 3 ? q.iz - 1 : 0, zhi = DIM == 3 ? q.iz + 1 : 0;
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = fcs[cell_key(q.ix - 1, jy, jz, c)], e = fcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for (int j = b; j < e; j++) {
        if (j == i) continue;
        const Vec<real, DIM> pj = pos[j];
        const real dist2 = norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz);
        if (dist2 < rp2) {
          const real pr = prs[j];
          if (pr < Prs_min) Prs_min = pr;
        }
      }
    }
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = fcs[cell_key(q.ix - 1, jy, jz, c)], e = fcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            return j != i && norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) < rp2;
          },
          [&](int j) {
            const Vec<real, DIM> pj = pos[j];
            const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
            const real dist2 = norm2<real, DIM>(qx, qy, qz);
            const real dist = sqrt(dist2);
            const real w = ((rp / dist) - (dist / rp)) / dist2;
            const real d = prs[j] - Prs_min;
            ax += d * w * qx;
            ay += d * w * qy;
            if constexpr (DIM == 3) az += d * w * qz;
          });
    }
  for (int jz = zlo; jz <= zhi; jz++)
    for (int jy = q.iy - 1; jy <= q.iy + 1; jy++) {
      const int b = mcs[cell_key(q.ix - 1, jy, jz, c)], e = mcs[cell_key(q.ix + 1, jy, jz, c) + 1];
      for_hits(
          b, e,
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            return norm2<real, DIM>(pj.x - px, pj.y - py, vz(pj) - pz) < rp2;
          },
          [&](int j) {
            const Vec<real, DIM> pj = mpos[j];
            const real qx = pj.x - px, qy = pj.y - py, qz = vz(pj) - pz;
            const real dist2 = norm2<real, DIM>(qx, qy, qz);
            const real dist = sqrt(dist2);
            const real w = ((rp / dist) - (dist / rp)) / dist2;
            const real d = mprs[j] - Prs_min;
            ax += d * w * qx;
            ay += d * w * qy;
            if constexpr (DIM == 3) az += d * w * qz;
          });
    }
  const real accx = c.invro * ax * c.grad_a;
  const real accy = c.invro * ay * c.grad_a;
  const real accz = DIM == 3 ? c.invro * az * c.grad_a : real(0);
  if constexpr (!FUSE) {
    acc_out[i] = mkvec<real, DIM>(accx, accy, accz);
  } else {
    const Vec<real, DIM> v = vel[i];
    real nvx = v.x + accx * c.dt, nvy = v.y + accy * c.dt;
    real nvz = DIM == 3 ? vz(v) + accz * c.dt : real(0);
    real nx = px + accx * c.dt * c.dt, ny = py + accy * c.dt * c.dt;
    real nz = DIM == 3 ? pz + accz * c.dt * c.dt : real(0);
    pav[i] += prs_i;
    real po = prs_i;
    if (!chk_pcl<real, DIM>(nx, ny, nz, nvx, nvy, nvz, c)) {
      typ[i] = kGST;
      po = real(0);
    }
    pos_out[i] = mkvec<real, DIM>(nx, ny, nz);
    vel_out[i] = mkvec<real, DIM>(nvx, nvy, nvz);
    prs_out[i] = po;
  }
}

// d_UpPcl2 alone (phase mode).
template <class real, int DIM>
__global__ void k_up2(Vec<real, DIM>* __restrict__ pos, Vec<real, DIM>* __restrict__ vel,
                      Vec<real, DIM>* __restrict__ acc, int8_t* __restrict__ typ,
                      real* __restrict__ prs, real* __restrict__ pav, int n, Cst<real> c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (typ[i] == kFLD) {
    const Vec<real, DIM> a = acc[i], p = pos[i], v = vel[i];
    real nvx = v.x + a.x * c.dt, nvy = v.y + a.y * c.dt;
    real nvz = DIM == 3 ? vz(v) + vz(a) * c.dt : real(0);
    real nx = p.x + a.x * c.dt * c.dt, ny = p.y + a.y * c.dt * c.dt;
    real nz = DIM == 3 ? vz(p) + vz(a) * c.dt * c.dt : real(0);
    acc[i] = mkvec<real, DIM>(0, 0, 0);
    pav[i] += prs[i];
    if (!chk_pcl<real, DIM>(nx, ny, nz, nvx, nvy, nvz, c)) {
      typ[i] = kGST;
      prs[i] = real(0);
    }
    pos[i] = mkvec<real, DIM>(nx, ny, nz);
    vel[i] = mkvec<real, DIM>(nvx, nvy, nvz);
  } else {
    pav[i] += prs[i];
  }
}

// pav += prs for static walls whose pressure is not zero.
template <class real>
__global__ void k_wall_pav(real* __restrict__ pav, const real* __restrict__ prs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pav[i] += prs[i];
}

// Arithmetic probe: the device's correctly-rounded sqrt and division, used by
// the tests to show they agree bit for bit with the host's.
template <class real>
__global__ void k_probe(const real* a, const real* b, real* s, real* q, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    s[i] = sqrt(a[i]);
    q[i] = a[i] / b[i];
  }
}

}  // namespace mps
