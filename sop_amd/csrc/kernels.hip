// kernels.hip -- gfx950 kernels of the SOP overdamped hot path.
//
//   k_bbox / k_cell_count / k_cell_scan / k_cell_scatter / k_cell_rank /
//   k_pl_cells / k_nl_build / k_shell_filter   list rebuilds: per-system open
//                cell grid, counting sort of beads into cells (histogram, DPP
//                wave scan, scatter, deterministic in-cell ranking), 27-cell
//                stencil with the exact FP32 predicates and ballot/mbcnt wave
//                compaction, pair-list rows filtered from shell rows and packed
//                per tile.  Replaces the reference's flag kernels + CUDPP
//                split/scan over the O(N^2) master list and its pair-list
//                filter (sop.cu:704-826, 835-920, 928-1262, 1456-1620,
//                1827-2608).
//   k_step       fused FENE + i,i+2 soft sphere + native LJ + non-native
//                repulsion gather (no atomics), group reduction over the lanes
//                of a bead, Philox4x32-10 random force, overdamped move
//                (energy.cu:81-167,228-604; sop.cu:487-520).
//   k_energy / k_sys_reduce   FP64 reductions of FP32 pair energies, virial,
//                native contacts and Rg^2 (energy.cu:194-442, sop.cu:528-611).
//
// Every kernel walks a device-resident work table with a grid-stride loop, so
// one launch covers any batch of systems and a captured graph stays valid when
// the batch changes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

#include "sop_dev.h"

namespace sopk {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int w_type(float w) { return __float_as_int(w) & 3; }
__device__ __forceinline__ int w_chain(float w) { return __float_as_int(w) >> 2; }

__device__ __forceinline__ const float4 *buf_in(const DevPtrs &p, int parity) {
  return (parity & 1) ? p.pos1 : p.pos0;
}
__device__ __forceinline__ float4 *buf_out(const DevPtrs &p, int parity) {
  return (parity & 1) ? p.pos0 : p.pos1;
}

// Exact FP32 predicate pieces, same operation order and rounding as the
// reference's single-precision list kernels; this file is compiled with
// -ffp-contract=off and correctly rounded division/sqrt.
__device__ __forceinline__ float mi_exact(float d, float boxl) {
  if (boxl > 0.0f) {
    float q = d / boxl;
    float r = rintf(q);
    float pr = boxl * r;
    d = d - pr;
  }
  return d;
}

__device__ __forceinline__ float dist2_exact(float4 a, float4 b, float boxl) {
  float dx = mi_exact(b.x - a.x, boxl);
  float dy = mi_exact(b.y - a.y, boxl);
  float dz = mi_exact(b.z - a.z, boxl);
  float p = dx * dx;
  float q = dy * dy;
  float r = dz * dz;
  return (p + q) + r;
}

__device__ __forceinline__ float cutoff_of(double factor, float absval, float base) {
  return factor > 0.0 ? (float)(factor * (double)base) : absval;
}

// Fast minimum image for forces (not used by any list predicate).
__device__ __forceinline__ float mi_fast(float d, float boxl, float inv_boxl) {
  return d - boxl * rintf(d * inv_boxl);
}

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }

// DPP lane moves.  Controls: quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E,
// row_half_mirror = 0x141, row_mirror = 0x140, row_shr:n = 0x110+n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over aligned groups of W lanes; every lane of a group receives the sum.
template <int W>
__device__ __forceinline__ float group_sum(float v) {
  static_assert(W == 4 || W == 8 || W == 16 || W == 32 || W == 64, "group width");
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  if (W >= 8) v += dpp_f<0x141>(v);
  if (W >= 16) v += dpp_f<0x140>(v);
  if (W >= 32) {
    // xor 16 inside each 32-lane half: ds_swizzle bit mode, and=0x1F, xor=0x10
    v += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
  }
  if (W >= 64) v += __shfl_xor(v, 32);
  return v;
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Inclusive prefix sum across the 64 lanes of a wave with DPP: Hillis-Steele
// inside each 16-lane row (row_shr 1,2,4,8), then row_bcast15 / row_bcast31
// carry row totals into the rows above.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);
  return v;
}

// Maximum over the wave's lanes (all lanes active).
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// Number of set bits of mask in lanes below this one (ballot prefix).
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 product per multiplier (v_mad_u64_u32) instead of
    // separate low and high multiplies
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
    c = make_uint4((uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// 24-bit uniform in (0,1): odd multiples of 2^-24, exact in float.
__device__ __forceinline__ float u24(uint32_t r) {
  return (float)((r >> 8) | 1u) * 5.9604644775390625e-08f;
}

// Three standard normals for (seed, bead, step): Box-Muller on Philox output.
__device__ __forceinline__ void normals3(uint32_t k0, uint32_t k1, uint32_t bead, int64_t step,
                                         float &n0, float &n1, float &n2) {
  const uint4 r = philox10(make_uint4(bead, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0u), k0, k1);
  const float ra = sqrtf(-2.0f * logf(u24(r.x)));
  const float rb = sqrtf(-2.0f * logf(u24(r.z)));
  float sa, ca, sb, cb;
  sincospif(2.0f * u24(r.y), &sa, &ca);
  sincospif(2.0f * u24(r.w), &sb, &cb);
  n0 = ra * ca;
  n1 = ra * sa;
  n2 = rb * cb;
}

__device__ __forceinline__ bool live(const SysDev &S, int kstep) { return S.step + kstep < S.step_end; }
__device__ __forceinline__ bool due_nl(const SysDev &S, int kstep) {
  return live(S, kstep) && ((S.step + kstep) % S.n_nl) == 0;
}
__device__ __forceinline__ bool due_pl(const SysDev &S, int kstep) {
  return live(S, kstep) && ((S.step + kstep) % S.n_pl) == 0;
}

__device__ __forceinline__ int cell_index(const SysDev &S, float4 x) {
  int cx = (int)floorf((x.x - S.org[0]) * S.inv_edge[0]);
  int cy = (int)floorf((x.y - S.org[1]) * S.inv_edge[1]);
  int cz = (int)floorf((x.z - S.org[2]) * S.inv_edge[2]);
  cx = min(max(cx, 0), S.dims[0] - 1);
  cy = min(max(cy, 0), S.dims[1] - 1);
  cz = min(max(cz, 0), S.dims[2] - 1);
  return (cz * S.dims[1] + cy) * S.dims[0] + cx;
}

// ------------------------------------------------------------ list rebuilds
//
// The neighbor list is defined by the snapshot pos_ref taken at its last
// rebuild: NL = master pairs with d2(pos_ref) < rc_nl^2, and the pair list is
// PL = NL pairs with d2(x) < rc_pl^2.  When every bead has moved less than
// skin/2 = (rc_nl - rc_pl)/2 since the snapshot, every pair with
// d(x) < rc_pl has d(pos_ref) < rc_pl + 2 max|dx| < rc_nl, so PL equals the
// master pairs with d2(x) < rc_pl^2: it is built straight from a cell grid of
// edge >= rc_pl (BUILD_DIRECT) and the neighbor list need not be written.
// Otherwise the neighbor list is materialized from pos_ref and filtered
// (BUILD_NL).  Both give the same set as the reference's two-level scheme.
//
// The direct build writes shell rows: master pairs with d(x) < rc_pl + shell,
// snapshot pos_shell = x, and the pair list is those rows filtered at x.  At a
// later direct-capable rebuild where every bead has moved less than shell/2
// since pos_shell, every pair with d(x) < rc_pl has d(pos_shell) < rc_pl +
// shell, so filtering the same rows again gives the same pair list without
// the cell grid (BUILD_SHELL).

// Systems whose rebuild sorts beads into a cell grid.
__device__ __forceinline__ bool grid_build(const SysDev &S) {
  return S.build != BUILD_NONE && S.build != BUILD_SHELL;
}

__device__ __forceinline__ const float4 *build_src(const DevPtrs &p, const SysDev &S, int kstep) {
  return S.build == BUILD_DIRECT ? buf_in(p, S.parity + kstep) : p.pos_ref;
}

// Per system: snapshot and build decision.
constexpr int SKIN_THREADS = 1024;
__global__ __launch_bounds__(SKIN_THREADS) void k_skin(DevPtrs p, int kstep, int flags) {
  constexpr int NW = SKIN_THREADS / 64;
  __shared__ float red[2][NW];
  __shared__ int s_build;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  if (blockIdx.x == 0 && tid == 0) *p.gsort = 0;  // k_bbox raises it (first kernel of a rebuild)
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const int s = p.syslist[q];
    SysDev &S = p.sys[s];
    const bool mat = (flags & MATERIALIZE_NL) != 0;
    const bool nl = !mat && ((flags & FORCE_NL) || due_nl(S, kstep));
    const bool pl = !mat && ((flags & FORCE_PL) || due_pl(S, kstep));
    // system constants in registers (the snapshot stores below may alias S)
    const int bo = S.bead_off, n = S.n;
    const float4 *x = buf_in(p, S.parity + kstep);
    const bool shell_chk = pl && S.sl_valid;
    float d2max = 0.0f, d2sh = 0.0f;
    if (nl || pl) {
      // four beads per pass with no branch on the bead count, so the loads of
      // a pass issue together (a bead past the end repeats the last one:
      // the same snapshot store, the same displacement)
      for (int b0 = tid; b0 < n; b0 += 4 * SKIN_THREADS) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int i = bo + min(b0 + u * SKIN_THREADS, n - 1);
          const float4 xi = x[i];
          if (nl) {
            p.pos_ref[i] = xi;
          } else {
            const float4 r = p.pos_ref[i];
            const float dx = xi.x - r.x, dy = xi.y - r.y, dz = xi.z - r.z;
            d2max = fmaxf(d2max, dx * dx + dy * dy + dz * dz);
          }
          if (shell_chk) {
            const float4 q = p.pos_shell[i];
            const float ex = xi.x - q.x, ey = xi.y - q.y, ez = xi.z - q.z;
            d2sh = fmaxf(d2sh, ex * ex + ey * ey + ez * ez);
          }
        }
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      d2max = fmaxf(d2max, __shfl_xor(d2max, o));
      d2sh = fmaxf(d2sh, __shfl_xor(d2sh, o));
    }
    if (lane == 0) {
      red[0][wave] = d2max;
      red[1][wave] = d2sh;
    }
    __syncthreads();
    if (tid == 0) {
      float m0 = 0.0f, m1 = 0.0f;
      for (int w = 0; w < NW; ++w) {
        m0 = fmaxf(m0, red[0][w]);
        m1 = fmaxf(m1, red[1][w]);
      }
      const float dmax = sqrtf(m0), dsh = sqrtf(m1);
      int build = BUILD_NONE;
      if (mat) {
        build = S.nl_valid ? BUILD_NONE : BUILD_NL_ONLY;
      } else {
        if (nl) S.nl_valid = 0;
        if (pl) {
          if (S.skin > 0.0f && (nl || 2.0f * dmax < S.skin))
            build = shell_chk && 2.0f * dsh < S.shell_ok ? BUILD_SHELL : BUILD_DIRECT;
          else
            build = BUILD_NL;
        } else if (flags & FORCE_NL) {
          build = BUILD_NL_ONLY;
        }
      }
      S.build = build;
      if (build == BUILD_DIRECT) S.sl_valid = 1;
      if (build == BUILD_NL) S.sl_valid = 0;  // k_nl_build writes its rows into the shell-row slots
      if (build == BUILD_DIRECT || build == BUILD_SHELL || build == BUILD_NL)
        ++S.nbuild[build == BUILD_DIRECT ? 0 : (build == BUILD_SHELL ? 1 : 2)];
      s_build = build;
    }
    __syncthreads();
    if (s_build == BUILD_DIRECT)  // the shell rows about to be built belong to these coordinates
      for (int b = tid; b < n; b += SKIN_THREADS) p.pos_shell[bo + b] = x[bo + b];
    __syncthreads();
  }
}

// Per system: bounding box -> open cell grid; zero the cell histogram.
__global__ __launch_bounds__(256) void k_bbox(DevPtrs p, int kstep) {
  __shared__ float red[6][4];
  __shared__ int s_ncells;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const int s = p.syslist[q];
    SysDev &S = p.sys[s];
    if (!grid_build(S)) continue;
    const float4 *src = build_src(p, S, kstep);
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    // four beads per pass with no branch on the count, so their loads overlap
    // (a bead past the end repeats the last one; min and max are unchanged)
    const int nb = S.n, bo = S.bead_off;
    for (int b0 = tid; b0 < nb; b0 += 4 * (int)blockDim.x) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float4 x = src[bo + min(b0 + u * (int)blockDim.x, nb - 1)];
        mn[0] = fminf(mn[0], x.x); mx[0] = fmaxf(mx[0], x.x);
        mn[1] = fminf(mn[1], x.y); mx[1] = fmaxf(mx[1], x.y);
        mn[2] = fminf(mn[2], x.z); mx[2] = fmaxf(mx[2], x.z);
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        mn[d] = fminf(mn[d], __shfl_xor(mn[d], o));
        mx[d] = fmaxf(mx[d], __shfl_xor(mx[d], o));
      }
    }
    if (lane == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) { red[d][wave] = mn[d]; red[3 + d][wave] = mx[d]; }
    }
    __syncthreads();
    if (tid == 0) {
      const float min_edge = S.build == BUILD_DIRECT ? S.cell_min_edge_sl : S.cell_min_edge;
      float ext_max = 0.0f;
      int nc = 1;
      for (int d = 0; d < 3; ++d) {
        float a = red[d][0], b = red[3 + d][0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { a = fminf(a, red[d][w]); b = fmaxf(b, red[3 + d][w]); }
        const float ext = b - a;
        // Cells of the minimum edge, one more than fit in the extent so the
        // grid covers it; when that exceeds dmax, dmax cells stretched over
        // the extent (edge ext/dmax >= min_edge).  Stretching every cell to
        // ext/floor(ext/min_edge) instead would make the 27-cell stencil up
        // to 8x the volume it needs to be.
        int dims = 1;
        float inv = 0.0f;
        if (ext > 0.0f) {
          dims = (int)floorf(ext / min_edge) + 1;
          if (dims <= S.dmax) {
            inv = 1.0f / min_edge;
          } else {
            dims = S.dmax;
            inv = (float)dims / ext;
          }
        }
        S.org[d] = a;
        S.inv_edge[d] = inv;
        S.dims[d] = dims;
        nc *= dims;
        ext_max = fmaxf(ext_max, ext);
      }
      S.ncells = nc;
      S.sort_lds = S.n <= SORT_N_MAX && nc <= SORT_C_MAX;
      if (!S.sort_lds) atomicOr(p.gsort, 1);
      if (S.boxl > 0.0f) {
        const int ok = ext_max < 0.5f * S.boxl - S.wrap_margin;
        S.wrap_free = ok;
        if (!ok) {
          atomicOr(&S.err, ERR_PBC);
          S.sl_valid = 0;
        }
      } else {
        S.wrap_free = 1;
      }
      s_ncells = nc;
    }
    __syncthreads();
    const int nc = s_ncells;
    for (int c = tid; c < nc; c += blockDim.x) p.cell_cnt[S.cell_base + c] = 0;
    __syncthreads();
  }
}

// Per bead: cell id and histogram.
__global__ __launch_bounds__(64) void k_cell_count(DevPtrs p, int kstep) {
  if (*p.gsort == 0) return;  // every grid build of the batch sorts in LDS (k_bbox)
  const int nwork = *p.nwork;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int t = p.work[w];
    if (t < 0) continue;
    const SysDev &S = p.sys[p.tile_sys[t]];
    if (!grid_build(S) || S.sort_lds) continue;
    const int i = t * TILE + threadIdx.x;
    if (i < S.bead_off + S.n) {
      const int c = cell_index(S, build_src(p, S, kstep)[i]);
      p.cell_of[i] = c;
      atomicAdd(&p.cell_cnt[S.cell_base + c], 1);
    }
  }
}

// Per system: exclusive scan of the histogram (DPP wave scan + wave totals).
__global__ __launch_bounds__(256) void k_cell_scan(DevPtrs p) {
  if (*p.gsort == 0) return;  // every grid build of the batch sorts in LDS (k_bbox)
  __shared__ int wsum[4], wused[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const int s = p.syslist[q];
    SysDev &S = p.sys[s];
    if (!grid_build(S) || S.sort_lds) continue;
    const int nc = S.ncells, cb = S.cell_base;
    int carry = 0, used = 0;
    for (int c0 = 0; c0 < nc; c0 += 256) {
      const int c = c0 + tid;
      const int v = c < nc ? p.cell_cnt[cb + c] : 0;
      const int incl = wave_incl_scan(v);
      if (lane == 63) wsum[wave] = incl;
      const int wu = __popcll(__ballot(v > 0));
      if (lane == 0) wused[wave] = wu;
      __syncthreads();
      int off = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        off += (w < wave) ? wsum[w] : 0;
        tot += wsum[w];
      }
      if (c < nc) {
        p.cell_start[cb + c] = carry + off + incl - v;
        p.cell_cursor[cb + c] = 0;
      }
      carry += tot;
      used += wused[0] + wused[1] + wused[2] + wused[3];
      __syncthreads();
    }
    if (tid == 0) {
      S.ncells_used = used;
      p.cell_start[cb + nc] = carry;
      if (carry != S.n) atomicOr(&S.err, ERR_INTERNAL);
    }
  }
}

// Per bead: scatter into its cell's segment (arbitrary in-cell order).
__global__ __launch_bounds__(64) void k_cell_scatter(DevPtrs p) {
  if (*p.gsort == 0) return;  // every grid build of the batch sorts in LDS (k_bbox)
  const int nwork = *p.nwork;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int t = p.work[w];
    if (t < 0) continue;
    const SysDev &S = p.sys[p.tile_sys[t]];
    if (!grid_build(S) || S.sort_lds) continue;
    const int i = t * TILE + threadIdx.x;
    if (i < S.bead_off + S.n) {
      const int c = p.cell_of[i];
      const int slot = atomicAdd(&p.cell_cursor[S.cell_base + c], 1);
      const int pos = p.cell_start[S.cell_base + c] + slot;
      if (pos >= 0 && pos < S.n) p.cell_tmp[S.bead_off + pos] = i;
    }
  }
}

// Per bead: rank among the beads of its cell by index -> sorted segment, with
// a cell-ordered copy of the coordinates.  The in-cell order, and with it every
// list, is independent of the order of the atomics above.
__global__ __launch_bounds__(64) void k_cell_rank(DevPtrs p, int kstep) {
  if (*p.gsort == 0) return;  // every grid build of the batch sorts in LDS (k_bbox)
  const int nwork = *p.nwork;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int t = p.work[w];
    if (t < 0) continue;
    const SysDev &S = p.sys[p.tile_sys[t]];
    if (!grid_build(S) || S.sort_lds) continue;
    const int i = t * TILE + threadIdx.x;
    if (i < S.bead_off + S.n) {
      const int c = p.cell_of[i];
      const int st = p.cell_start[S.cell_base + c], en = p.cell_start[S.cell_base + c + 1];
      int rank = 0;
      if (st >= 0 && en <= S.n) {
        for (int m = st; m < en; ++m) rank += p.cell_tmp[S.bead_off + m] < i;
        const int pos = st + rank;
        if (pos < S.n) {
          p.cell_pos[i] = pos;
          p.nat_sorted[S.bead_off + pos] = make_int2(p.nat_start[i], p.nat_cnt[i]);
          p.cell_sorted[S.bead_off + pos] = i;
          p.pos_sorted[S.bead_off + pos] = build_src(p, S, kstep)[i];
        }
      }
    }
  }
}

__device__ __forceinline__ bool native_pair(const int32_t *nj, int nn, int j) {
  int lo = 0, hi = nn;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (nj[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return lo < nn && nj[lo] == j;
}

// Native partners of the current bead, as a wave-private LDS bitmap keyed by
// (local index mod 65536); a set bit is confirmed by the exact search unless
// the system is small enough for the key to be the index itself.
constexpr int NATBITS_WORDS = 2048;

__device__ __forceinline__ void natbits_mark(uint32_t *bits, const int32_t *natj, int nn, int bead_off, bool set) {
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < nn; k += 64) {
    const int jl = (natj[k] - bead_off) & 0xFFFF;
    if (set) atomicOr(&bits[jl >> 5], 1u << (jl & 31));
    else bits[jl >> 5] = 0u;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Candidate scan of the 27 cells around xi as 9 rows of 3 contiguous cells,
// reading the cell-ordered copy of the coordinates.  Non-native master pairs
// with d2 below the cutoff of their type pair (rcA: tj = 1, rcB: tj = 2) are
// appended to out[0..cap) with ballot/mbcnt
// compaction; returns the count (which may exceed cap).
template <typename Extra, typename Emit>
__device__ __forceinline__ int scan_cells(const DevPtrs &p, const SysDev &S, int i, float4 xi,
                                          float rcA, float rcB, int ns, int nn, uint32_t *bits, Extra extra,
                                          Emit emit) {
  const int lane = threadIdx.x & 63;
  const int wi = __float_as_int(xi.w), ci = wi >> 2;
  const float boxl = S.boxl;
  const int cx = min(max((int)floorf((xi.x - S.org[0]) * S.inv_edge[0]), 0), S.dims[0] - 1);
  const int cy = min(max((int)floorf((xi.y - S.org[1]) * S.inv_edge[1]), 0), S.dims[1] - 1);
  const int cz = min(max((int)floorf((xi.z - S.org[2]) * S.inv_edge[2]), 0), S.dims[2] - 1);
  const int x0 = max(cx - 1, 0), x1 = min(cx + 1, S.dims[0] - 1);
  const int32_t *natj = p.nat_j + ns;
  const bool exact_key = S.n <= 65536;
  natbits_mark(bits, natj, nn, S.bead_off, true);
  int cnt = 0;
  for (int zz = max(cz - 1, 0); zz <= min(cz + 1, S.dims[2] - 1); ++zz)
    for (int yy = max(cy - 1, 0); yy <= min(cy + 1, S.dims[1] - 1); ++yy) {
      const int row = (zz * S.dims[1] + yy) * S.dims[0];
      const int st = max(p.cell_start[S.cell_base + row + x0], 0);
      const int en = min(p.cell_start[S.cell_base + row + x1 + 1], S.n);
      for (int m0 = st; m0 < en; m0 += 64) {
        const int m = m0 + lane;
        bool pred = false;
        int j = 0;
        if (m < en) {
          j = p.cell_sorted[S.bead_off + m];
          const float4 xj = p.pos_sorted[S.bead_off + m];
          const int wj = __float_as_int(xj.w);
          const int dj = j - i;
          const bool excl = (j == i) || (((wj >> 2) == ci) && dj >= -2 && dj <= 2) ||
                            j < S.bead_off || j >= S.bead_off + S.n;
          if (!excl && dist2_exact(xi, xj, boxl) < ((wj & 3) == 2 ? rcB : rcA)) {
            const int jl = (j - S.bead_off) & 0xFFFF;
            bool nat = (bits[jl >> 5] >> (jl & 31)) & 1u;
            if (nat && !exact_key) nat = native_pair(natj, nn, j);
            pred = !nat && extra(j, wj);
          }
        }
        const uint64_t mk = __ballot(pred);
        if (pred) emit(cnt + lanes_below(mk), j);
        cnt += __popcll(mk);
      }
    }
  natbits_mark(bits, natj, nn, S.bead_off, false);
  return cnt;
}

// Per system (one 1024-thread block) whose sort fits LDS: the whole counting
// sort of the global kernels above in LDS -- cell of each bead, histogram,
// exclusive scan, member lists, and the rank of each bead among its cell's
// members by index, so the in-cell order is the index order -- then the
// cell-ordered arrays written out in one pass.
constexpr int SORT_THREADS = 1024;

__global__ __launch_bounds__(SORT_THREADS) void k_cell_sort_lds(DevPtrs p, int kstep) {
  __shared__ int s_cell[SORT_N_MAX];
  __shared__ int s_mem[SORT_N_MAX];
  __shared__ int s_cnt[SORT_C_MAX];  // counts, then end offsets
  __shared__ int s_wsum[SORT_THREADS / 64], s_wused[SORT_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const int s = p.syslist[q];
    SysDev &S = p.sys[s];
    if (!grid_build(S) || !S.sort_lds) continue;
    const int n = S.n, nc = S.ncells, bo = S.bead_off, cb = S.cell_base;
    const float4 *src = build_src(p, S, kstep);
    for (int c = tid; c < nc; c += SORT_THREADS) s_cnt[c] = 0;
    // this thread's beads (at most SB): every load and cell index before the
    // first store (a load issued after a store would wait for it)
    constexpr int SB = SORT_N_MAX / SORT_THREADS;
    float4 xs[SB];
    int cs[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      xs[u] = src[bo + min(tid + u * SORT_THREADS, n - 1)];
      cs[u] = cell_index(S, xs[u]);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int b = tid + u * SORT_THREADS;
      if (b < n) {
        s_cell[b] = cs[u];
        p.cell_of[bo + b] = cs[u];
        atomicAdd(&s_cnt[cs[u]], 1);
      }
    }
    __syncthreads();
    // exclusive scan of the counts: 4 consecutive cells per thread
    constexpr int PER = SORT_C_MAX / SORT_THREADS;
    int v[PER], sum = 0, used = 0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int c = tid * PER + u;
      v[u] = c < nc ? s_cnt[c] : 0;
      sum += v[u];
      used += v[u] > 0;
    }
    const int incl = wave_incl_scan(sum);
    const int wu = wave_sum_i(used);
    if (lane == 63) s_wsum[wave] = incl;
    if (lane == 0) s_wused[wave] = wu;
    __syncthreads();
    int off = 0, tot_used = 0;
    for (int w = 0; w < SORT_THREADS / 64; ++w) {
      off += w < wave ? s_wsum[w] : 0;
      tot_used += s_wused[w];
    }
    int run = off + incl - sum;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int c = tid * PER + u;
      if (c < nc) {
        p.cell_start[cb + c] = run;
        s_cnt[c] = run;  // cursor, ends at the cell's end offset
      }
      run += v[u];
    }
    if (tid == 0) {
      p.cell_start[cb + nc] = n;
      S.ncells_used = tot_used;
    }
    __syncthreads();
    for (int b = tid; b < n; b += SORT_THREADS) s_mem[atomicAdd(&s_cnt[s_cell[b]], 1)] = b;
    __syncthreads();
    int ns[SB], nn[SB];
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int i = bo + min(tid + u * SORT_THREADS, n - 1);
      ns[u] = p.nat_start[i];
      nn[u] = p.nat_cnt[i];
    }
#pragma unroll
    for (int u = 0; u < SB; ++u) {
      const int b = tid + u * SORT_THREADS;
      if (b < n) {
        const int c = s_cell[b];
        const int st = c > 0 ? s_cnt[c - 1] : 0, en = s_cnt[c];
        int r = 0;
        for (int m = st; m < en; ++m) r += s_mem[m] < b;
        const int pos = st + r;
        const int i = bo + b;
        p.cell_pos[i] = pos;
        p.cell_sorted[bo + pos] = i;
        p.pos_sorted[bo + pos] = xs[u];
        p.nat_sorted[bo + pos] = make_int2(ns[u], nn[u]);
      }
    }
    __syncthreads();
  }
}

// Per batch: exclusive scan of the home-cell counts of the systems that build
// their pair list directly (one block).
__global__ __launch_bounds__(256) void k_cellwork(DevPtrs p) {
  __shared__ int wsum[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  int carry = 0;
  bool nl = false;  // some system materializes its neighbor list at this rebuild
  for (int q0 = 0; q0 < nsys; q0 += 256) {
    const int q = q0 + tid;
    int v = 0;
    if (q < nsys) {
      const SysDev &S = p.sys[p.syslist[q]];
      v = S.build == BUILD_DIRECT ? S.ncells_used : 0;
      nl |= S.build == BUILD_NL || S.build == BUILD_NL_ONLY;
    }
    const int incl = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      off += (w < wave) ? wsum[w] : 0;
      tot += wsum[w];
    }
    if (q < nsys) p.cellwork[q] = carry + off + incl - v;
    carry += tot;
    __syncthreads();
  }
  const int any_nl = __syncthreads_or(nl);
  if (tid == 0) {
    p.cellwork[nsys] = carry;
    *p.ncellwork = carry;
    *p.nlbuild = any_nl;
  }
}

// Per system of the batch: list its home cells for the direct build.
__global__ __launch_bounds__(256) void k_cellexpand(DevPtrs p) {
  __shared__ int wsum[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const SysDev &S = p.sys[p.syslist[q]];
    if (S.build != BUILD_DIRECT) continue;
    const int off = p.cellwork[q], nc = S.ncells, cb = S.cell_base;
    int base = 0;
    for (int c0 = 0; c0 < nc; c0 += 256) {
      const int c = c0 + tid;
      const bool used = c < nc && p.cell_start[cb + c + 1] > p.cell_start[cb + c];
      const uint64_t m = __ballot(used);
      if (lane == 0) wsum[wave] = __popcll(m);
      __syncthreads();
      int o = base;
      for (int w = 0; w < wave; ++w) o += wsum[w];
      if (used) p.cellitems[off + o + lanes_below(m)] = make_int2(q, c);
      base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
      __syncthreads();
    }
  }
}

// Block per home cell: the shell rows of the cell's beads straight from the
// shell-cutoff grid.  The bounds of the 9 neighbor rows and all their
// candidates are staged in LDS together (CAND_MAX at a time); every home bead
// of the cell tests them.  Each bead's list comes out in row order, then
// cell-sorted order: the order of a per-bead scan of the same cells.
constexpr int HOME_MAX = 64;   // home beads per pass
constexpr int CAND_MAX = 1024; // candidates staged per pass (and bits of the exclusion bitmaps)
constexpr int EXCL_WORDS = CAND_MAX / 32;

// Pair-list rows are padded with the bead's own index up to a multiple of
// ROW_PAD entries, so the step kernel reads whole chunks without a count
// test: a self entry has d2 = 0, which the repulsion guard turns into zero
// force.  pl_cap is a multiple of 16, so a padded row fits its capacity.
constexpr int ROW_PAD = 4;
// 16-bit pair-list entries per step-kernel chunk load (4: 8 bytes; 8: 16
// bytes, which needs ROW_PAD 8): four entries per lane and load leave fewer
// lanes idle at the end of a row than eight
#ifndef STEP_CHUNK16
#define STEP_CHUNK16 4
#endif
static_assert(ROW_PAD % STEP_CHUNK16 == 0 && ROW_PAD % 4 == 0, "rows must hold whole chunks");
// LDS of one workgroup on gfx950, and the shell filter's row staging: per
// wave, 16 beads of a round x FILTER_SROW 16-bit entries (32-bit entries take
// twice the bytes; longer rows are completed by a second pass)
constexpr int LDS_BYTES_MAX = 160 * 1024;
constexpr int FILTER_SROW = 144;
constexpr int FILTER_STAGE_BYTES = STEP_WAVES * 16 * FILTER_SROW * 2;
// Buffer resource of `bytes` bytes at `base` (wave-uniform): loads through
// it take a 32-bit offset, and a load past the end returns zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, uint32_t bytes) {
  const uint64_t a = (uint64_t)base;
  const uint64_t u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(a >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  return __builtin_amdgcn_make_buffer_rsrc((void *)u, (short)0, (int)__builtin_amdgcn_readfirstlane((int)bytes),
                                           0x00020000);
}

__device__ __forceinline__ uint4 load_b128(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
  return make_uint4(v[0], v[1], v[2], v[3]);
}

__device__ __forceinline__ int2 load_i2(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
  return make_int2((int)v[0], (int)v[1]);
}

// Output row of one bead as a raw buffer of pl_cap entries: stores past the
// row's end are dropped by the buffer bounds check (the count still grows,
// which the caller reports as an overflow).  The caller's row address and
// size are wave-uniform; readfirstlane puts the descriptor in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(void *row, int bytes) {
  const uint64_t a = (uint64_t)row;
  const uint64_t u = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(a >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  return __builtin_amdgcn_make_buffer_rsrc((void *)u, (short)0, __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}

// The fast form of the home-bead scan: the staged candidates are padded to a
// multiple of 128 with coordinates no pair can reach, and the home bead's
// bitmap over candidate slots already excludes itself, its chain neighbours
// i +- 1, 2 and its native partners, so a candidate costs its loads, the
// distance, one bit test and, when kept, one buffer store.  No minimum image:
// a rebuild whose extent reaches boxl/2 - margin raises ERR_PBC, so every
// separation here is below boxl/2, where the minimum image is the identity.
// ONE: every bead has the home bead's type (one cutoff rcA).
template <bool ONE, bool IDX16>
__device__ __forceinline__ int pl_scan_fast(const float4 *s_cand, const int *s_cj, int ncp, float4 xi, float rcA,
                                            float rcB, const uint32_t *excl, __amdgpu_buffer_rsrc_t row, int bo,
                                            int lane, int cnt) {
  const uint32_t lbit = 1u << (lane & 31);
  for (int k0 = 0; k0 < ncp; k0 += 128) {
    int jj[2];
    float4 xj[2];
    uint32_t ew[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int ci = k0 + u * 64 + lane;
      jj[u] = s_cj[ci];
      xj[u] = s_cand[ci];
      ew[u] = excl[ci >> 5];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const float dx = xj[u].x - xi.x, dy = xj[u].y - xi.y, dz = xj[u].z - xi.z;
      const float d2 = (dx * dx + dy * dy) + dz * dz;
      const float rc = ONE ? rcA : ((__float_as_int(xj[u].w) & 3) == 2 ? rcB : rcA);
      const bool keep = (d2 < rc) & ((ew[u] & lbit) == 0u);
      const uint64_t mk = __ballot(keep);
      if (keep) {
        const int o = cnt + lanes_below(mk);
        if (IDX16) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(jj[u] - bo), row, o * 2, 0, 0);
        else __builtin_amdgcn_raw_buffer_store_b32((unsigned int)jj[u], row, o * 4, 0, 0);
      }
      cnt += __popcll(mk);
    }
  }
  return cnt;
}

__global__ __launch_bounds__(256) void k_pl_cells(DevPtrs p, int kstep) {
  __shared__ float4 s_home[HOME_MAX];
  __shared__ int s_hj[HOME_MAX], s_hcnt[HOME_MAX], s_hnn[HOME_MAX], s_hns[HOME_MAX];
  __shared__ uint32_t s_excl[HOME_MAX][EXCL_WORDS];  // per home bead: candidate slots it must skip
  __shared__ float4 s_cand[CAND_MAX];
  __shared__ int s_cj[CAND_MAX];
  __shared__ int s_rst[9], s_roff[10], s_hrange[2];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int total = *p.ncellwork;
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    const int2 item = p.cellitems[g];
    const int s = p.syslist[item.x];
    const int c = item.y;
    const SysDev &S = p.sys[s];
    const int bo = S.bead_off, cb = S.cell_base, n = S.n;
    const int pl_cap = S.pl_cap;
    const int64_t pl_base = S.pl_base;
    const float rc2_11 = S.rc2_sl_rep[0], rc2_12 = S.rc2_sl_rep[1], rc2_21 = S.rc2_sl_rep[2],
                rc2_22 = S.rc2_sl_rep[3];
    const int d0 = S.dims[0], d1 = S.dims[1], d2 = S.dims[2];
    const int cx = c % d0, cy = (c / d0) % d1, cz = c / (d0 * d1);
    // bounds of the home cell and of the 3x3 neighbor rows (x-range cx-1..cx+1)
    if (tid < 9) {
      const int zz = cz + tid / 3 - 1, yy = cy + tid % 3 - 1;
      int st = 0, len = 0;
      if (zz >= 0 && zz < d2 && yy >= 0 && yy < d1) {
        const int row = (zz * d1 + yy) * d0;
        st = max(p.cell_start[cb + row + max(cx - 1, 0)], 0);
        len = max(min(p.cell_start[cb + row + min(cx + 1, d0 - 1) + 1], n) - st, 0);
      }
      s_rst[tid] = st;
      s_roff[tid + 1] = len;
    } else if (tid < 11) {
      s_hrange[tid - 9] = min(max(p.cell_start[cb + c + (tid - 9)], 0), n);
    }
    __syncthreads();
    if (tid == 0) {
      s_roff[0] = 0;
      for (int r = 1; r <= 9; ++r) s_roff[r] += s_roff[r - 1];
    }
    __syncthreads();
    const int hs = s_hrange[0], he = s_hrange[1];
    const int T = s_roff[9];
    if (hs >= he) {
      __syncthreads();  // everyone has read the bounds before the next cell rewrites them
      continue;
    }
    const bool exact = T > CAND_MAX;  // candidates staged in passes of CAND_MAX
    // stage the candidates [b0, b0 + nc), padded to a multiple of 128 with
    // coordinates no pair can reach
    auto stage = [&](int b0, int nc) {
      // CAND_MAX / 256 slots per thread, all loads before the LDS writes
      // (slots past nc load the last candidate and store the padding)
      constexpr int CS = CAND_MAX / 256;
      float4 xc[CS];
      int jc[CS];
#pragma unroll
      for (int u = 0; u < CS; ++u) {
        const int idx = b0 + min(tid + 256 * u, nc - 1);
        int r = 0;
#pragma unroll
        for (int q = 1; q < 9; ++q) r += s_roff[q] <= idx;
        const int m = s_rst[r] + (idx - s_roff[r]);
        xc[u] = p.pos_sorted[bo + m];
        jc[u] = p.cell_sorted[bo + m];
      }
      const int ncp = (nc + 127) & ~127;
#pragma unroll
      for (int u = 0; u < CS; ++u) {
        const int ci = tid + 256 * u;
        if (ci < nc) {
          s_cand[ci] = xc[u];
          s_cj[ci] = jc[u];
        } else if (ci < ncp) {
          s_cand[ci] = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
          s_cj[ci] = -1;
        }
      }
    };
    // the first home pass's beads load beside the candidate staging
    int2 ni0 = make_int2(0, 0);
    float4 xh0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    int jh0 = 0;
    if (tid < min(HOME_MAX, he - hs)) {
      ni0 = p.nat_sorted[bo + hs + tid];
      xh0 = p.pos_sorted[bo + hs + tid];
      jh0 = p.cell_sorted[bo + hs + tid];
    }
    if (!exact) stage(0, T);
    for (int h0 = hs; h0 < he; h0 += HOME_MAX) {
      const int nh = min(HOME_MAX, he - h0);
      if (tid < nh) {
        if (h0 != hs) {
          ni0 = p.nat_sorted[bo + h0 + tid];
          xh0 = p.pos_sorted[bo + h0 + tid];
          jh0 = p.cell_sorted[bo + h0 + tid];
        }
        s_home[tid] = xh0;
        s_hj[tid] = jh0;
        s_hcnt[tid] = 0;
        s_hns[tid] = ni0.x;
        s_hnn[tid] = ni0.y;
      }
      // candidate passes of CAND_MAX staged slots (one pass when they all fit,
      // staged above for every home pass)
      for (int b0 = 0; b0 < T; b0 += CAND_MAX) {
        const int nc = min(CAND_MAX, T - b0);
        if (exact) stage(b0, nc);
        if (tid < nh) {
#pragma unroll
          for (int w = 0; w < EXCL_WORDS; ++w) s_excl[tid][w] = 0u;
        }
        __syncthreads();
        // exclusion bitmaps over this pass's slots: the slot of each native
        // partner and of the bead itself and its chain neighbours i +- 1, 2,
        // from their cells in this rebuild's grid (the row) and sorted
        // positions (the offset)
        // four home beads per pass of a wave, 16 lanes each, so the dependent
        // loads of several beads are in flight together
        for (int t0 = 0; wave + 4 * t0 < nh; t0 += 4) {
          const int h = wave + 4 * (t0 + (lane >> 4));
          if (h >= nh) continue;
          const int ji = s_hj[h], ns = s_hns[h], nn = s_hnn[h];
          const int chi = __float_as_int(s_home[h].w) >> 2;
          for (int k = lane & 15; k < nn + 5; k += 16) {
            int cn, pos, jn;
            if (k < nn) {
              const int jp = p.nat_j[ns + k];  // native partner: its cell and place in the cell order
              cn = p.cell_of[jp];
              pos = p.cell_pos[jp];
              jn = -1;
            } else {
              jn = ji + (k - nn - 2);
              if (jn < bo || jn >= bo + n) continue;
              cn = p.cell_of[jn];
              pos = p.cell_pos[jn];
            }
            const int nx = cn % d0, ny = (cn / d0) % d1, nz = cn / (d0 * d1);
            if (abs(nx - cx) > 1 || abs(ny - cy) > 1 || abs(nz - cz) > 1) continue;
            const int r = (nz - cz + 1) * 3 + (ny - cy + 1);
            const int slot = s_roff[r] + (pos - s_rst[r]);
            if (slot < s_roff[r] || slot >= s_roff[r + 1]) continue;
            const int ls = slot - b0;  // staged position in this pass
            if (ls < 0 || ls >= nc) continue;
            if (jn >= 0 && (s_cj[ls] != jn || (__float_as_int(s_cand[ls].w) >> 2) != chi)) continue;
            atomicOr(&s_excl[h][ls >> 5], 1u << (ls & 31));
          }
        }
        __syncthreads();
        for (int h = wave; h < nh; h += 4) {
          const float4 xi = s_home[h];
          const int ji = s_hj[h];
          const int wi = __float_as_int(xi.w);
          const float rcA = (wi & 3) == 2 ? rc2_21 : rc2_11, rcB = (wi & 3) == 2 ? rc2_22 : rc2_12;
          int32_t *out = p.slrep_j + pl_base + (int64_t)(ji - bo) * pl_cap;
          uint16_t *out16 = S.idx16 ? (uint16_t *)(p.slrep_j + pl_base) + (int64_t)(ji - bo) * pl_cap : nullptr;
          const int ncp = (nc + 127) & ~127;
          const bool one = S.type_mask != 3;
          const float rc1 = S.type_mask == 2 ? rcB : rcA;
          void *row = S.idx16 ? (void *)out16 : (void *)out;
          const __amdgpu_buffer_rsrc_t rr = row_rsrc(row, pl_cap * (S.idx16 ? 2 : 4));
          const int c0 = s_hcnt[h];
          int cnt;
          if (one) cnt = S.idx16 ? pl_scan_fast<true, true>(s_cand, s_cj, ncp, xi, rc1, rc1, s_excl[h], rr, bo, lane, c0)
                                 : pl_scan_fast<true, false>(s_cand, s_cj, ncp, xi, rc1, rc1, s_excl[h], rr, bo, lane, c0);
          else cnt = S.idx16 ? pl_scan_fast<false, true>(s_cand, s_cj, ncp, xi, rcA, rcB, s_excl[h], rr, bo, lane, c0)
                             : pl_scan_fast<false, false>(s_cand, s_cj, ncp, xi, rcA, rcB, s_excl[h], rr, bo, lane, c0);
          if (lane == 0) s_hcnt[h] = cnt;
        }
        __syncthreads();  // the next pass restages and clears the bitmaps
      }
      __syncthreads();
      if (tid < nh) {
        const int j = s_hj[tid], cr = s_hcnt[tid];
        p.slrep_cnt[j] = min(cr, pl_cap);
        if (cr > pl_cap) {
          atomicOr(&p.sys[s].err, ERR_PL_OVF);
          p.sys[s].sl_valid = 0;  // truncated rows must not be filtered again
        }
      }
      if (wave == 0) {  // the longest row of this pass, for the host's capacity check
        const int cr = lane < nh ? s_hcnt[lane] : 0;
        const int mx = wave_max(cr);
        if (lane == 0) atomicMax(&p.sys[s].row_max, mx);
      }
      __syncthreads();
    }
  }
}

// Wave per bead, from the cell grid of the neighbor-list snapshot pos_ref.
// BUILD_NL_ONLY: materialize the neighbor list (pairs within rc_nl at
// pos_ref) into the neighbor-list scratch.  BUILD_NL: the pair list of the
// reference's two-level scheme, neighbor-list pairs within rc_pl at x, in one
// pass without materializing the neighbor list: the native pair list in each
// bead's native slots, the non-native rows in the shell-row slots (the shell
// filter packs them into the pair list), which therefore stop holding shell
// rows (k_skin clears sl_valid).
__global__ __launch_bounds__(256) void k_nl_build(DevPtrs p, int kstep) {
  if (*p.nlbuild == 0) return;  // no system of the batch needs it (k_cellwork)
  __shared__ uint32_t s_bits[4][NATBITS_WORDS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = lane; k < NATBITS_WORDS; k += 64) s_bits[wave][k] = 0u;
  __builtin_amdgcn_wave_barrier();
  const int nwork = *p.nwork;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int t = p.work[w];
    if (t < 0) continue;
    const int s = p.tile_sys[t];
    const SysDev &S = p.sys[s];
    if (S.build != BUILD_NL && S.build != BUILD_NL_ONLY) continue;
    const bool plb = S.build == BUILD_NL;
    const float4 *ref = p.pos_ref;
    const float4 *pin = buf_in(p, S.parity + kstep);
    const float boxl = S.boxl;
    const int bend = min(t * TILE + TILE, S.bead_off + S.n);
    for (int q = 0; q < TILE / 4; ++q) {
      const int i = t * TILE + wave * (TILE / 4) + q;
      if (i >= bend) break;
      const float4 xi = ref[i];
      const float4 xn = pin[i];
      const int ti = __float_as_int(xi.w) & 3;
      const int ns = p.nat_start[i], nn = p.nat_cnt[i];
      int cnt = 0;
      for (int k0 = 0; k0 < nn; k0 += 64) {
        const int k = k0 + lane;
        bool pred = false;
        int j = 0;
        float r0 = 0.0f;
        if (k < nn) {
          j = p.nat_j[ns + k];
          r0 = p.nat_r0[ns + k];
          const float rc = cutoff_of(S.nl_factor, S.nl_abs, r0);
          pred = dist2_exact(xi, ref[j], boxl) < rc * rc;
          if (plb && pred) {
            const float rp = cutoff_of(S.pl_factor, S.pl_abs, r0);
            pred = dist2_exact(xn, pin[j], boxl) < rp * rp;
          }
        }
        const uint64_t m = __ballot(pred);
        if (pred) {
          const int o = ns + cnt + lanes_below(m);
          if (plb) {
            p.platt[o] = make_int2(j, __float_as_int(r0));
          } else {
            p.nlatt_j[o] = j;
            p.nlatt_r0[o] = r0;
          }
        }
        cnt += __popcll(m);
      }
      const float rcA = S.rc2_nl_rep[(ti - 1) * 2], rcB = S.rc2_nl_rep[(ti - 1) * 2 + 1];
      if (plb) {
        if (lane == 0) {
          p.platt_cnt[i] = cnt;
          p.platt_off[i] = ns;  // in the bead's own native slots
        }
        const float pA = S.rc2_pl_rep[(ti - 1) * 2], pB = S.rc2_pl_rep[(ti - 1) * 2 + 1];
        const int64_t row = (int64_t)(i - S.bead_off) * S.pl_cap;  // the bead's shell-row slot
        uint16_t *r16 = (uint16_t *)(p.slrep_j + S.pl_base) + row;
        int32_t *r32 = p.slrep_j + S.pl_base + row;
        const int cap = S.pl_cap, bo = S.bead_off;
        const bool i16 = S.idx16 != 0;
        const int cr = scan_cells(
            p, S, i, xi, rcA, rcB, ns, nn, s_bits[wave],
            [&](int j, int wj) { return dist2_exact(xn, pin[j], boxl) < ((wj & 3) == 2 ? pB : pA); },
            [&](int o, int j) {
              if (o < cap) {
                if (i16) r16[o] = (uint16_t)(j - bo);
                else r32[o] = j;
              }
            });
        if (lane == 0) {
          p.slrep_cnt[i] = min(cr, cap);
          if (cr > cap) atomicOr(&p.sys[s].err, ERR_PL_OVF);
          atomicMax(&p.sys[s].row_max, cr);
        }
      } else {
        if (lane == 0) p.nlatt_cnt[i] = cnt;
        int32_t *out = p.nlrep_j + S.nl_base + (int64_t)(i - S.bead_off) * S.nl_cap;
        const int cap = S.nl_cap;
        const int cr = scan_cells(p, S, i, xi, rcA, rcB, ns, nn, s_bits[wave], [](int, int) { return true; },
                                  [&](int o, int j) {
                                    if (o < cap) out[o] = j;
                                  });
        if (lane == 0) {
          p.nlrep_cnt[i] = min(cr, cap);
          if (cr > cap) atomicOr(&p.sys[s].err, ERR_NL_OVF);
          p.sys[s].nl_valid = 1;
        }
      }
    }
  }
}

// Per tile of a system whose pair list is rebuilt from its neighbor list: the
// order in which the step kernel's lane groups take the tile's beads, by
// decreasing row length (non-native + native), ties by index, so the beads
// that one round of lane groups computes together run loops of similar
// length (grid and shell builds: the shell filter computes the same order).
// Results do not depend on the order: each bead is computed by one group the
// same way.
__global__ __launch_bounds__(256) void k_tile_order(DevPtrs p) {
  if (*p.nlbuild == 0) return;  // shell and grid builds: the shell filter orders their tiles
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwork = *p.nwork;
  for (int w = blockIdx.x * 4 + wave; w < nwork; w += gridDim.x * 4) {
    const int t = p.work[w];
    if (t < 0) continue;
    const SysDev &S = p.sys[p.tile_sys[t]];
    if (S.build != BUILD_NL) continue;
    const int b0 = t * TILE, i = b0 + lane;
    // grid and shell builds: shell-row counts (the pair lists are filtered
    // after this); neighbor-list builds: the previous pair-list counts
    const int nrep = S.build == BUILD_NL ? p.plrep_cnt[i] : p.slrep_cnt[i];
    const int key = i < S.bead_off + S.n ? nrep + p.platt_cnt[i] : -1;
    int rank = 0;
    for (int j = 0; j < TILE; ++j) {
      const int kj = __shfl(key, j);
      rank += (kj > key) | ((kj == key) & (j < lane));
    }
    p.tile_ord[b0 + rank] = (uint8_t)lane;
  }
}

// ----------------------------------------------- fused force + integrate

// Partner coordinates come from the LDS copy of the item's arena range
// (LDS = true, base = first staged bead) or from global memory.
template <bool LDS>
__device__ __forceinline__ float4 partner(const float4 *__restrict__ pin, const float4 *s_pos, int base, int j) {
  return LDS ? s_pos[j - base] : pin[j];
}

// One wave, one tile of 64 beads.  LPB lanes per bead gather its partners and
// reduce with DPP; GPW = 64 / LPB beads per round, LPB rounds.  After each
// round, lane m picks up the force of bead m of the tile from the group that
// computed it, so the random force and the move run on all 64 lanes, one bead
// per lane, with no block barrier.  ONE: every bead of the system has the
// same type, so the force-field constants are wave-uniform (scalar registers).
template <int LPB, int MODE, bool LDS, bool IDX16, bool ONE>
__device__ __forceinline__ void wave_tile(const DevPtrs &p, const SysDev &S, int s, int t, int kstep,
                                          const float4 *__restrict__ pin, float4 *__restrict__ pout,
                                          const float4 *s_pos, int lds0, int r0, int r1) {
  constexpr int GPW = 64 / LPB;
  const int lane = threadIdx.x & 63;
  const int g = lane / LPB, l = lane % LPB;
  const int bead_off = S.bead_off, n = S.n;
  const int b0 = t * TILE;
  const int bend = min(b0 + TILE, bead_off + n);
  // No minimum image: a state whose extent reaches boxl/2 - margin is refused
  // at each rebuild (ERR_PBC), so every separation here is below boxl/2.
  const int64_t step = S.step + kstep;
  const int q1 = S.type_mask == 2 ? 3 : 0;  // type-pair index of a one-type system
  // slot q of the tile's order (k_tile_order) is computed in round q / GPW by
  // group q % GPW; lane m's own bead b0 + m sits in slot qm
  const int ordm = p.tile_ord[b0 + lane];
  const int qm = __builtin_amdgcn_ds_permute(ordm << 2, lane);
  const int src = (qm % GPW) * LPB, rm = qm / GPW;
  // per-bead list heads of the whole tile in one coalesced load each; a
  // round picks its beads' values up by shuffles
  // (native pair-list length | pair-list length << 16: both below 2^16, which
  // the host checks)
  int t_ns = 0, t_cn = 0, t_po = 0;
  float t_br = 0.0f;  // r0 of bond (b0 + lane, b0 + lane + 1)
  // (from the bead's list head: rows from the tile's region start, native
  // lists from the native slots of the tile's first bead)
  const int nat0 = p.nat_start[b0], reg0 = (b0 - bead_off) * S.pl_cap;
  if (b0 + lane < bend) {
    const uint2 hd = p.plhead[b0 + lane];
    t_cn = (int)hd.x;
    t_po = reg0 + ROW_PAD * (int)(hd.y & 0xFFFFu);
    t_ns = nat0 + (int)(hd.y >> 16);
    t_br = p.bond_r0[b0 + lane];
  }
  const float br_before = b0 > bead_off ? p.bond_r0[b0 - 1] : 0.0f;  // bond (b0 - 1, b0)
  // the non-native rows in chunks of E entries: STEP_CHUNK16 16-bit entries
  // (8 or 4: one 16- or 8-byte load) or 4 32-bit ones
  constexpr int E = IDX16 ? STEP_CHUNK16 : 4;
  constexpr uint32_t ESZ = IDX16 ? 2u : 4u;
  constexpr uint32_t CB = (uint32_t)E * ESZ;  // chunk bytes
  using chunk_t = uint4;
  auto load_chunk = [&](__amdgpu_buffer_rsrc_t r, uint32_t off) -> chunk_t {
    if constexpr (CB == 16) {
      return load_b128(r, off);
    } else {
      const int2 v = load_i2(r, off);
      return make_uint4((uint32_t)v.x, (uint32_t)v.y, 0u, 0u);
    }
  };
  // buffer resources in scalar registers: this system's pair-list segment
  // (rows packed per tile, plrep_off) and the native pair list, addressed by
  // 32-bit offsets
  const __amdgpu_buffer_rsrc_t rrow =
      buf_rsrc(p.plrep_j + S.pl_base, (uint32_t)S.n * (uint32_t)S.pl_cap * ESZ);
  const __amdgpu_buffer_rsrc_t rnat = buf_rsrc(p.platt, 0xFFFFFFFFu);
  float mfx = 0.0f, mfy = 0.0f, mfz = 0.0f;  // force on bead b0 + lane
  for (int r = r0; r < r1; ++r) {
    const int li = __shfl(ordm, r * GPW + g);
    const int i = b0 + li;
    const int ns = __shfl(t_ns, li), cn = __shfl(t_cn, li), po = __shfl(t_po, li);
    const int nc = cn & 0xFFFF, nr = cn >> 16;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    if (i < bend) {
      // first loads of the native and non-native loops, issued before any
      // gather: the lane's first row chunk and first two native entries
      // ({partner, bits of r0})
      const uint32_t rowoff = (uint32_t)po * ESZ;
      const int nch = (nr + E - 1) / E;
      chunk_t q = {};
      if (l < nch) q = load_chunk(rrow, rowoff + CB * (uint32_t)l);
      int2 na = make_int2(0, 0), nb2 = make_int2(0, 0);
      if (l < nc) na = load_i2(rnat, 8u * (uint32_t)(ns + l));
      if (l + LPB < nc) nb2 = load_i2(rnat, 8u * (uint32_t)(ns + l + LPB));
      const float4 xi = partner<LDS>(pin, s_pos, lds0, i);
      const int wi = __float_as_int(xi.w);
      const int row = ONE ? q1 : ((wi & 3) - 1) * 2;
      // native LJ over the native pair list (the first two entries per lane loaded above)
      {
        // c = -12 eps u q6 (q6 - 1) with u = 1/d2, q6 = (r0^2 u)^3, as
        // fma(t, q6, -t) with t = -12 eps u q6
        const float eA = -12.0f * S.eps_att[row], eB = ONE ? eA : -12.0f * S.eps_att[row + 1];
        for (int k = l; k < nc; k += LPB) {
          int2 en;
          if (k == l) en = na;
          else if (k == l + LPB) en = nb2;
          else en = load_i2(rnat, 8u * (uint32_t)(ns + k));
          const int j = en.x;
          const float r0 = __int_as_float(en.y);
          const float4 xj = partner<LDS>(pin, s_pos, lds0, j);
          float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
          const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const float eps = ONE ? eA : (w_type(xj.w) == 2 ? eB : eA);
          const float r02 = r0 * r0;
          const float ir2 = rcpf(d2);
          const float q2 = r02 * ir2;
          const float q6 = q2 * q2 * q2;
          const float t = eps * ir2 * q6;
          float c = fmaf(t, q6, -t);
          c = d2 > 1.0e-7f * r02 ? c : 0.0f;
          fx = fmaf(c, dx, fx);
          fy = fmaf(c, dy, fy);
          fz = fmaf(c, dz, fz);
        }
      }
      // non-native repulsion: lane l takes chunks l, l + LPB, ... of the row,
      // one load per E pairs, the next chunk's load issued before the current
      // chunk's gathers; the row's padding entries are the bead itself and
      // contribute nothing (d2 = 0 fails the guard)
      {
        // c = u q6 (m12 q6 + m6) with u = 1/d2, q6 = (sigma^2 u)^3, as
        // u^4 (a u^3 + b) with a = m12 sigma^12, b = m6 sigma^6
        const float aA = S.rep_a[row], aB = ONE ? aA : S.rep_a[row + 1];
        const float bA = S.rep_b[row], bB = ONE ? bA : S.rep_b[row + 1];
        const float gA = S.rep_guard[row], gB = ONE ? gA : S.rep_guard[row + 1];
        auto rep = [&](float4 xj) {
          float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
          const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
          const bool tb = !ONE && w_type(xj.w) == 2;
          const float u = rcpf(d2);
          const float u2 = u * u;
          float c = (u2 * u2) * fmaf(tb ? aB : aA, u2 * u, tb ? bB : bA);
          c = d2 > (tb ? gB : gA) ? c : 0.0f;
          fx = fmaf(c, dx, fx);
          fy = fmaf(c, dy, fy);
          fz = fmaf(c, dz, fz);
        };
#pragma unroll 2
        for (int c = l; c < nch; c += LPB) {
          const chunk_t cur = q;
          // the next chunk's load without a condition: past the row it reads
          // the rows after it, or zeros past the segment (buffer range check)
          q = load_chunk(rrow, rowoff + CB * (uint32_t)(c + LPB));
          uint32_t ww[E];
          if constexpr (IDX16) {
            ww[0] = cur.x & 0xFFFFu; ww[1] = cur.x >> 16; ww[2] = cur.y & 0xFFFFu; ww[3] = cur.y >> 16;
            if constexpr (E == 8) {
              ww[4 % E] = cur.z & 0xFFFFu; ww[5 % E] = cur.z >> 16; ww[6 % E] = cur.w & 0xFFFFu; ww[7 % E] = cur.w >> 16;
            }
          } else {
            ww[0] = cur.x; ww[1] = cur.y; ww[2] = cur.z; ww[3] = cur.w;
          }
#pragma unroll
          for (int e = 0; e < E; ++e) rep(partner<LDS>(pin, s_pos, lds0, IDX16 ? bead_off + (int)ww[e] : (int)ww[e]));
        }
      }
    }
    fx = group_sum<LPB>(fx);
    fy = group_sum<LPB>(fy);
    fz = group_sum<LPB>(fz);
    const float vx = __shfl(fx, src), vy = __shfl(fy, src), vz = __shfl(fz, src);
    if (rm == r) {
      mfx = vx;
      mfy = vy;
      mfz = vz;
    }
  }
  // Chain terms, each pair once, over the tile's lanes: lane m computes the
  // FENE bond (i, i+1) and the i,i+2 soft sphere of its bead i = b0 + m (the
  // force on i; the partner takes the opposite), and takes the partner halves
  // of the bond (i-1, i) and the pair (i-2, i) from lanes m-1 and m-2.  The
  // pairs that start before the tile come from lanes 0 and 1 themselves.
  {
    const int i = b0 + lane;
    const int send = bead_off + n;
    float4 xc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < bend) xc = partner<LDS>(pin, s_pos, lds0, i);
    // force on a from the pair (a, b) of chain offset off (1: FENE bond of
    // length r0, 2: soft sphere); zero across chains
    auto chain_pair = [&](float4 xa, int ja, int off, float r0, float &ox, float &oy, float &oz) {
      ox = oy = oz = 0.0f;
      const int jb = ja + off;
      if (ja < bead_off || jb >= send) return;
      const float4 xb = partner<LDS>(pin, s_pos, lds0, jb);
      if ((__float_as_int(xb.w) >> 2) != (__float_as_int(xa.w) >> 2)) return;
      const float dx = xb.x - xa.x, dy = xb.y - xa.y, dz = xb.z - xa.z;
      const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      float c;
      if (off == 1) {
        const float d = sqrtf(d2);
        const float dev = d - r0;
        const float gg = 1.0f - dev * dev * S.inv_R0sq;
        c = S.k_bnd * dev * rcpf(d * gg);
        if (gg <= 0.0f) atomicOr(&p.sys[s].err, ERR_BOND);
      } else {
        const float ir2 = rcpf(d2);
        const float ir4 = ir2 * ir2;
        c = -S.ss_f6 * (ir4 * ir4);
      }
      ox = c * dx;
      oy = c * dy;
      oz = c * dz;
    };
    float bx, by, bz, sx, sy, sz;
    chain_pair(xc, i < bend ? i : send, 1, t_br, bx, by, bz);
    chain_pair(xc, i < bend ? i : send, 2, 0.0f, sx, sy, sz);
    // the partner halves from lanes m-1 (bond) and m-2 (soft sphere)
    const int m1 = lane >= 1 ? lane - 1 : 0, m2 = lane >= 2 ? lane - 2 : 0;
    float pbx = __shfl(bx, m1), pby = __shfl(by, m1), pbz = __shfl(bz, m1);
    float psx = __shfl(sx, m2), psy = __shfl(sy, m2), psz = __shfl(sz, m2);
    if (lane < 2 && i < bend) {  // pairs that start before the tile
      float ex, ey, ez;
      const float4 xp2 = i - 2 >= bead_off ? partner<LDS>(pin, s_pos, lds0, i - 2) : xc;
      chain_pair(xp2, i - 2 >= bead_off ? i - 2 : send, 2, 0.0f, ex, ey, ez);
      psx = ex; psy = ey; psz = ez;
      if (lane == 0) {
        const float4 xp1 = i - 1 >= bead_off ? partner<LDS>(pin, s_pos, lds0, i - 1) : xc;
        chain_pair(xp1, i - 1 >= bead_off ? i - 1 : send, 1, br_before, ex, ey, ez);
        pbx = ex; pby = ey; pbz = ez;
      }
    }
    mfx += (bx + sx) - (pbx + psx);
    mfy += (by + sy) - (pby + psy);
    mfz += (bz + sz) - (pbz + psz);
  }
  const int i = b0 + lane;
  if (i >= bend) return;
  if (rm < r0 || rm >= r1) return;  // the other wave of a split tile computed this bead
  if (MODE == MODE_FORCE) {
    p.fout[i] = make_float4(mfx, mfy, mfz, 0.0f);
    return;
  }
  float n0, n1, n2;
  normals3(S.seed_lo, S.seed_hi, (uint32_t)(i - bead_off), step, n0, n1, n2);
  const float var = S.var;
  const float ftx = mfx + n0 * var, fty = mfy + n1 * var, ftz = mfz + n2 * var;
  if (MODE == MODE_FORCE_NOISE) {
    p.fout[i] = make_float4(ftx, fty, ftz, 0.0f);
    return;
  }
  float4 x = partner<LDS>(pin, s_pos, lds0, i);
  if (S.dyn == 0) {
    // overdamped (sop.cu:487-520): x += (h/zeta) F
    const float a5 = S.a5;
    x.x = x.x + a5 * ftx;
    x.y = x.y + a5 * fty;
    x.z = x.z + a5 * ftz;
  } else {
    // underdamped (sop.cu:421-484): the reference's velocity update after its
    // force evaluation, v = a3 incr + a4 F, happens here at the start of the
    // next step, then incr = a1 v + a2 F, x += incr
    const float4 st = p.vel[i];
    float vx2 = st.x, vy2 = st.y, vz2 = st.z;
    if (__float_as_int(st.w) == 1) {
      vx2 = S.a3 * st.x + S.a4 * ftx;
      vy2 = S.a3 * st.y + S.a4 * fty;
      vz2 = S.a3 * st.z + S.a4 * ftz;
    }
    const float ix = S.a1 * vx2 + S.a2 * ftx, iy = S.a1 * vy2 + S.a2 * fty, iz = S.a1 * vz2 + S.a2 * ftz;
    x.x = x.x + ix;
    x.y = x.y + iy;
    x.z = x.z + iz;
    p.vel[i] = S.vupd ? make_float4(ix, iy, iz, __int_as_float(1)) : make_float4(vx2, vy2, vz2, 0.0f);
  }
  pout[i] = x;
  if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z))) atomicOr(&p.sys[s].err, ERR_NAN);
}

// Stage an item's arena beads [z, z + w) (whole tiles) into LDS.  Wave v
// copies tiles v, v + STEP_WAVES, ...: lane j first looks up the coordinate
// buffer of tile v + STEP_WAVES * j (one dependent chain per wave instead of
// one per bead), then each batch of copies issues all its loads before its
// LDS writes.
__device__ __forceinline__ void stage_item(const DevPtrs &p, int z, int w, int kstep, float4 *s_pos, int wave,
                                           int lane) {
  const int t0 = z / TILE, nt = w / TILE;
  const int mt = wave + STEP_WAVES * lane;
  int par = 0;
  if (mt < nt) par = p.sys[p.tile_sys[t0 + mt]].parity + kstep;
  constexpr int JB = 5;
  if (wave >= nt) return;
  const int jl = (nt - 1 - wave) / STEP_WAVES;  // this wave's last tile is wave + STEP_WAVES * jl
  for (int j0 = 0; j0 <= jl; j0 += JB) {
    // no conditions, so every load issues before the first write: past its
    // last tile the wave copies that tile again (same data, same buffer)
    float4 v[JB];
    int k[JB];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      const int jj = min(j0 + j, jl);
      k[j] = wave + STEP_WAVES * jj;
      const int pj = __shfl(par, jj);
      v[j] = buf_in(p, pj)[z + k[j] * TILE + lane];
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) s_pos[k[j] * TILE + lane] = v[j];
  }
}

// Work items of one launch: LDS items (tier 0), whole-CU LDS items (tier 1),
// or the items that gather from global memory (KLDS = false).
template <bool KLDS, bool FILTER>
__device__ __forceinline__ void item_range(const DevPtrs &p, int tier, int &w0, int &w1) {
  const int n0 = FILTER ? *p.nfitems : *p.nitems, n1 = FILTER ? *p.nfitems_big : *p.nitems_big;
  if (!KLDS) {
    w0 = n0 + n1;
    w1 = w0 + (FILTER ? *p.nfitems_far : *p.nitems_far);
  } else if (tier == 0) {
    w0 = 0;
    w1 = n0;
  } else {
    w0 = n0;
    w1 = n0 + n1;
  }
}

// Fused step over work items {first tile, tiles | waves per tile << 8, first
// staged bead or -1, staged beads}: with s waves per tile, waves s k .. s k +
// s - 1 of the block share tile first + k, each taking LPB / s consecutive
// rounds (tiles * s <= STEP_WAVES).  The
// LDS kernel (KLDS) takes the items whose systems fit in LDS and first copies
// the staged arena beads (whole systems, each from its own coordinate buffer)
// into dynamic LDS, so every partner gather is an LDS read; the other kernel
// takes the items that gather from global memory (items[nitems .. nitems +
// nitems_far)).  The result is identical either way.  Both are launched each
// step, so each has its own register budget.
template <int LPB, int MODE, bool KLDS>
__device__ __forceinline__ void step_items(const DevPtrs &p, int kstep, float4 *s_pos, int tier) {
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: system data in scalar registers
  int w0, w1;
  item_range<KLDS, false>(p, tier, w0, w1);
  for (int w = w0 + blockIdx.x; w < w1; w += gridDim.x) {
    const int4 it = p.items[w];
    if (it.x < 0) continue;
    if (KLDS) {
      __syncthreads();  // the previous item's reads are done
      stage_item(p, it.z, it.w, kstep, s_pos, wave, tid & 63);
      __syncthreads();
    }
    // split tiles: waves ns*k .. ns*k + ns - 1 share tile k of the item, each
    // taking LPB / ns consecutive rounds
    const int ns = __builtin_amdgcn_readfirstlane(item_split(it.y));
    const int tw = wave / ns;
    const int r0 = (wave % ns) * (LPB / ns), r1 = r0 + LPB / ns;
    if (tw >= item_tiles(it.y)) continue;
    const int t = it.x + tw;
    const int s = __builtin_amdgcn_readfirstlane(p.tile_sys[t]);
    const SysDev &S = p.sys[s];
    const int par = S.parity + kstep;
    const float4 *__restrict__ pin = buf_in(p, par);
    float4 *__restrict__ pout = buf_out(p, par);
    if (MODE == MODE_STEP && S.step + kstep >= S.step_end) {
      // finished inside this block of steps: carry the coordinates over
      const int i = t * TILE + (tid & 63);
      if (r0 == 0 && i < S.bead_off + S.n) pout[i] = pin[i];
      continue;
    }
    const int lds0 = KLDS ? it.z : 0;
    if (S.type_mask != 3) {
      if (S.idx16) wave_tile<LPB, MODE, KLDS, true, true>(p, S, s, t, kstep, pin, pout, s_pos, lds0, r0, r1);
      else wave_tile<LPB, MODE, KLDS, false, true>(p, S, s, t, kstep, pin, pout, s_pos, lds0, r0, r1);
    } else {
      if (S.idx16) wave_tile<LPB, MODE, KLDS, true, false>(p, S, s, t, kstep, pin, pout, s_pos, lds0, r0, r1);
      else wave_tile<LPB, MODE, KLDS, false, false>(p, S, s, t, kstep, pin, pout, s_pos, lds0, r0, r1);
    }
  }
}

// Pair list of one tile's beads = their shell rows filtered at x with the
// reference's exact predicate (d2 < rc_pl^2, squares summed in order without
// contraction), for grid and shell builds, and the native pair list from the
// native CSR.  Neighbor-list builds: k_nl_build has left the filtered rows in
// the shell-row slots and written the native pair list; the rows are packed
// here as they are.  4 lanes per bead, 16 beads per round: lane l
// takes 4-entry chunks l, l + 4, ... of the row, kept entries are compacted
// per chunk position, so a row comes out in (chunk pass, position, lane)
// order.  The rows are packed per tile (plrep_off): a round's row starts
// follow from its kept counts, so a round first stages up to FILTER_SROW
// entries per row in LDS (STAGE) and writes them out once the starts are
// known; the entries past FILTER_SROW, and every entry without staging, come
// from a second pass over the round's rows, which repeats the first's
// compaction order.  Rows are padded to ROW_PAD entries for the step kernel.
template <bool LDS, bool IDX16, bool STAGE, bool ONE>
__device__ __forceinline__ void filter_tile(const DevPtrs &p, const SysDev &S, int s, int t,
                                            const float4 *__restrict__ pin, const float4 *s_pos, int lds0,
                                            void *s_stage) {
  using ent_t = typename std::conditional<IDX16, uint16_t, uint32_t>::type;
  using chunk_t = typename std::conditional<IDX16, uint2, uint4>::type;
  const int lane = threadIdx.x & 63, g = lane >> 2, l = lane & 3;
  const uint32_t below = (1u << l) - 1u;
  // system constants in registers: the stores below may alias SysDev as far
  // as the compiler knows, and a reload would wait for every earlier store
  const int bo = S.bead_off, pl_cap = S.pl_cap;
  const int64_t pl_base = S.pl_base;
  const int bend = min(t * TILE + TILE, bo + S.n);
  const bool nlb = S.build == BUILD_NL;  // neighbor-list build: rows already filtered (k_nl_build)
  // cutoffs as wave-uniform values (a per-lane indexed load would be a vector
  // load, and waiting for it would also wait for the earlier stores)
  const float rc11 = S.rc2_pl_rep[0], rc12 = S.rc2_pl_rep[1], rc21 = S.rc2_pl_rep[2], rc22 = S.rc2_pl_rep[3];
  // row lengths of the whole tile in one load, picked up by shuffles; beads
  // in the tile's order (k_tile_order), slot q = 16 r + g
  const bool tin = t * TILE + lane < bend;
  const int t_nr = tin ? p.slrep_cnt[t * TILE + lane] : 0;
  const int t_ns = tin ? p.nat_start[t * TILE + lane] : 0, t_nn = tin ? p.nat_cnt[t * TILE + lane] : 0;
  // the tile's bead order for this filter and the step kernel (k_tile_order's
  // key: shell-row length + the native pair-list length of the last rebuild,
  // decreasing, ties by index), read before this pass rewrites platt_cnt.
  // Computed at grid builds only: shell builds keep the order of the grid
  // build that made their shell rows (row lengths change little between
  // them), and neighbor-list builds the order k_tile_order gave them.  The
  // order only groups beads into rounds; no result depends on it.
  int ordm;
  if (nlb || S.build == BUILD_SHELL) {
    ordm = p.tile_ord[t * TILE + lane];
  } else {
    const int key = tin ? t_nr + p.platt_cnt[t * TILE + lane] : -1;
    int rank = 0;
    for (int j = 0; j < TILE; ++j) {
      const int kj = __shfl(key, j);
      rank += (kj > key) | ((kj == key) & (j < lane));
    }
    p.tile_ord[t * TILE + rank] = (uint8_t)lane;
    ordm = __builtin_amdgcn_ds_permute(rank << 2, lane);  // slot lane holds bead ordm
  }
  const double plf = S.pl_factor;
  const float pla = S.pl_abs;
  ent_t *const seg = reinterpret_cast<ent_t *>(p.plrep_j + pl_base);
  // this tile's region of the segment, and the rows packed so far; the same
  // for the native pair list, whose region is the tile's native CSR range
  const int region = (t * TILE - bo) * pl_cap;
  int packed = 0;
  const int nat_region = __shfl(t_ns, 0);
  int nat_packed = 0;
  for (int r = 0; r < 4; ++r) {
    const int li = __shfl(ordm, r * 16 + g);
    const int i = t * TILE + li;
    const bool act = i < bend;
    const int self = act ? i : bend - 1;  // a bead of this tile: safe to gather
    const int nr = __shfl(t_nr, li);
    // shuffles outside any condition: a lane that a condition turns off does
    // not provide its value to the lanes reading from it
    const int ns = __shfl(t_ns, li), nn_all = __shfl(t_nn, li);
    const int nn = act && !nlb ? nn_all : 0;
    // the bead's first NB natives per lane, loaded with the row's first batch
    constexpr int NB = 4;
    int nj[NB];
    float nr0[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int k = l + 4 * b;
      nj[b] = self;
      nr0[b] = 0.0f;
      if (k < nn) {
        nj[b] = p.nat_j[ns + k];
        nr0[b] = p.nat_r0[ns + k];
      }
    }
    const float4 xi = partner<LDS>(pin, s_pos, lds0, self);
    const bool t2 = (__float_as_int(xi.w) & 3) == 2;
    // the cutoffs by the partner's type (one-type systems: the one of their type)
    const float rcA = ONE ? (t2 ? rc22 : rc11) : (t2 ? rc21 : rc11), rcB = t2 ? rc22 : rc12;
    // entry value of the bead itself, and entry value -> staged position (LDS)
    // or arena index
    const int vself = IDX16 ? self - bo : self;
    const int vpos = LDS ? (IDX16 ? bo - lds0 : -lds0) : (IDX16 ? bo : 0);
    const chunk_t *src = reinterpret_cast<const chunk_t *>(
        reinterpret_cast<const ent_t *>(p.slrep_j + pl_base) + (int64_t)(self - bo) * pl_cap);
    ent_t *s_row = STAGE ? reinterpret_cast<ent_t *>(s_stage) + g * FILTER_SROW : nullptr;  // this group's staged row
    // One pass over the first m entries of the row (m = 0: the group takes
    // no part): kept entries in (chunk batch, position, lane) order, put(o, j)
    // for the o-th.  Chunks in batches of CB per lane, all loads of a batch
    // issued before its stores (a load issued after a store waits for the
    // store as well).  The order does not depend on CB.
    auto pass = [&](int m, auto put) -> int {
      constexpr int CB = IDX16 ? 8 : 4;  // 64 bytes of row loads in flight per lane
      const int nch = (m + 3) >> 2;
      const int clast = max(nch - 1, 0);
      int cnt = 0;
      for (int m0 = 0; __ballot(l + 4 * m0 < nch) != 0; m0 += CB) {
        chunk_t ch[CB];
#pragma unroll
        for (int b = 0; b < CB; ++b) ch[b] = src[min(l + 4 * (m0 + b), clast)];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
          const int c = l + 4 * (m0 + b);
          if (__ballot(c < nch) == 0) break;
          // entry values as stored (16-bit local or 32-bit arena indices);
          // an entry past the row's end reads the bead itself
          int v[4];
          if constexpr (IDX16) {
            v[0] = (int)(ch[b].x & 0xFFFFu); v[1] = (int)(ch[b].x >> 16);
            v[2] = (int)(ch[b].y & 0xFFFFu); v[3] = (int)(ch[b].y >> 16);
          } else {
            v[0] = (int)ch[b].x; v[1] = (int)ch[b].y; v[2] = (int)ch[b].z; v[3] = (int)ch[b].w;
          }
          bool ok[4];
          float4 xj[4];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            ok[e] = 4 * c + e < m;
            const int q = (ok[e] ? v[e] : vself) + vpos;
            xj[e] = LDS ? s_pos[q] : pin[q];
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            // the reference's predicate without the minimum image, which is the
            // identity here (a grid build whose extent reaches boxl/2 - margin
            // raises ERR_PBC, and beads move less than the margin before the next)
            const float dx = xj[e].x - xi.x, dy = xj[e].y - xi.y, dz = xj[e].z - xi.z;
            const float rc = ONE ? rcA : (w_type(xj[e].w) == 2 ? rcB : rcA);
            const bool keep = ok[e] & (nlb | (((dx * dx + dy * dy) + dz * dz) < rc));
            const uint32_t bits = (uint32_t)(__builtin_amdgcn_ballot_w64(keep) >> (4 * g)) & 0xFu;
            if (keep) put(cnt + __popc(bits & below), v[e]);
            cnt += __popc(bits);
          }
        }
      }
      return cnt;
    };
    const int mrow = act ? nr : 0;
    int cnt;
    if constexpr (STAGE) {
      cnt = pass(mrow, [&](int o, int v) {
        if (o < FILTER_SROW) s_row[o] = (ent_t)v;
      });
    } else {
      cnt = pass(mrow, [](int, int) {});
    }
    // native pair list: natives within their pair cutoff, in native order,
    // packed per tile like the rows (platt_off): the keep decisions first,
    // kept as one bit per entry of the lane (up to 32), then the entries at
    // the round's starts.  A batch past the first is loaded again for the
    // writes; a bead with more than 128 natives (4 lanes x 32) has its
    // decisions made again.
    int h_ncnt = 0, h_noff = 0;  // the bead's native pair-list length and offset in the tile's slots
    if (nlb) {
      if (act) {
        h_ncnt = p.platt_cnt[i];
        h_noff = p.platt_off[i] - nat_region;
      }
    } else {
      const bool multi = __ballot(nn > 4 * NB) != 0, huge = __ballot(nn > 4 * 32) != 0;
      uint32_t km = 0;  // keep bits: bit m0 + b for the lane's entry l + 4 (m0 + b)
      auto nat_pass = [&](bool wr, int o0) -> int {
        int ncnt = 0;
        for (int m0 = 0; __ballot(l + 4 * m0 < nn) != 0; m0 += NB) {
          if (m0 > 0 || (wr && multi)) {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              const int k = l + 4 * (m0 + b);
              nj[b] = self;
              nr0[b] = 0.0f;
              if (k < nn) {
                nj[b] = p.nat_j[ns + k];
                nr0[b] = p.nat_r0[ns + k];
              }
            }
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            bool keep;
            const int bit = m0 + b;
            if (wr && !huge) {
              keep = (km >> bit) & 1u;
            } else {
              const int k = l + 4 * (m0 + b);
              const float4 xj = partner<LDS>(pin, s_pos, lds0, nj[b]);
              const float rc = cutoff_of(plf, pla, nr0[b]);
              // the reference's predicate without the minimum image (see above)
              const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
              keep = (k < nn) & (((dx * dx + dy * dy) + dz * dz) < rc * rc);
              if (!wr && bit < 32 && keep) km |= 1u << bit;
            }
            const uint32_t bits = (uint32_t)(__ballot(keep) >> (4 * g)) & 0xFu;
            if (wr && keep) p.platt[o0 + ncnt + __popc(bits & below)] = make_int2(nj[b], __float_as_int(nr0[b]));
            ncnt += __popc(bits);
          }
        }
        return ncnt;
      };
      const int ncnt = nat_pass(false, 0);
      const int nincl = wave_incl_scan(l == 0 ? ncnt : 0);
      const int nstart = nat_region + nat_packed + nincl - ncnt;
      nat_packed += __shfl(nincl, 63);
      nat_pass(true, nstart);
      if (act && l == 0) {
        p.platt_cnt[i] = ncnt;
        p.platt_off[i] = nstart;
      }
      h_ncnt = ncnt;
      h_noff = nstart - nat_region;
    }
    // row starts: the round's rows follow the rows packed so far, in slot order
    const int c = act ? min(cnt, pl_cap) : 0;
    const int len = (c + ROW_PAD - 1) & ~(ROW_PAD - 1);
    const int incl = wave_incl_scan(l == 0 ? len : 0);
    const int start = region + packed + incl - len;
    packed += __shfl(incl, 63);
    ent_t *dst = seg + start;
    if (act && l == 0) {
      p.plrep_cnt[i] = c;
      p.plrep_off[i] = start;
      p.plhead[i] = make_uint2((uint32_t)h_ncnt | ((uint32_t)c << 16),
                               (uint32_t)((start - region) / ROW_PAD) | ((uint32_t)h_noff << 16));
      if (cnt > pl_cap) atomicOr(&p.sys[s].err, ERR_PL_OVF);
    }
    // pad to a multiple of ROW_PAD with the bead itself (its d2 = 0 fails the
    // step kernel's repulsion guard)
    const ent_t me = (ent_t)(IDX16 ? i - bo : i);
    for (int k = c + l; k < len; k += 4) {
      if (STAGE && k < FILTER_SROW) s_row[k] = me;
      else dst[k] = me;
    }
    if constexpr (STAGE) {
      // the staged part of each row to global memory in 8-byte chunks (row
      // starts are multiples of ROW_PAD entries from a 64-byte aligned segment)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      constexpr int PER8 = 8 / (int)sizeof(ent_t);
      const int nst = min(len, FILTER_SROW) / PER8;
      const uint2 *srow2 = reinterpret_cast<const uint2 *>(s_row);
      uint2 *drow2 = reinterpret_cast<uint2 *>(dst);
      for (int k = l; k < nst; k += 4) drow2[k] = srow2[k];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // the entries not staged: a second pass over the rows that have them
    const int lo = STAGE ? FILTER_SROW : 0;
    if (__ballot(c > lo) != 0) {
      pass(c > lo ? mrow : 0, [&](int o, int v) {
        if (o >= lo && o < c) dst[o] = (ent_t)v;
      });
    }
  }
}

// Shell filter over its own work items (the step kernel's LDS staging, one
// wave per tile): wave w of the block takes tile first + w of its item.
// stage_off >= 0: rows are staged in LDS at this byte offset (FILTER_SROW
// entries per bead of a round, per wave), for systems whose entries fit
// stage_bytes (16-bit entries need FILTER_STAGE_BYTES, 32-bit twice that).
template <bool KLDS>
__global__ __launch_bounds__(STEP_THREADS) void k_shell_filter(DevPtrs p, int kstep, int tier, int stage_off,
                                                               int stage_bytes, int builds) {
  extern __shared__ float4 s_pos[];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  char *s_stage = reinterpret_cast<char *>(s_pos) + stage_off;
  int w0, w1;
  item_range<KLDS, true>(p, tier, w0, w1);
  for (int w = w0 + blockIdx.x; w < w1; w += gridDim.x) {
    const int4 it = p.fitems[w];
    if (it.x < 0) continue;
    // one wave per tile (the tile order it computes first must come from one
    // wave); a tile is filtered when its system's build is one of `builds`
    // (bit per build kind).  Every wave reads the whole item's decisions, so
    // an item with nothing to filter is skipped by the whole block (no
    // staging) without a block-wide vote.
    const int lane = tid & 63;
    bool mine = false;
    if (lane < item_tiles(it.y)) mine = ((builds >> p.sys[p.tile_sys[it.x + lane]].build) & 1) != 0;
    const uint64_t tiles_mine = __ballot(mine);
    if (tiles_mine == 0) continue;
    if (KLDS) {
      __syncthreads();  // the previous item's reads are done
      stage_item(p, it.z, it.w, kstep, s_pos, wave, lane);
      __syncthreads();
    }
    if (((tiles_mine >> wave) & 1) == 0) continue;
    const int t = it.x + wave;
    const int s = __builtin_amdgcn_readfirstlane(p.tile_sys[t]);
    const SysDev &S = p.sys[s];
    const float4 *pin = buf_in(p, S.parity + kstep);
    const int lds0 = KLDS ? it.z : 0;
    // each wave owns stage_bytes / STEP_WAVES bytes whatever the entry width
    // of the system it filters: the waves of a global-gather block move on to
    // their next item without a barrier, so one wave may stage 16-bit rows
    // while another stages 32-bit ones
    char *wst = s_stage + wave * (stage_bytes / STEP_WAVES);
    const bool st16 = stage_off >= 0 && stage_bytes >= FILTER_STAGE_BYTES;
    const bool st32 = stage_off >= 0 && stage_bytes >= 2 * FILTER_STAGE_BYTES;
    // one-type systems: every partner has the bead's type (one cutoff)
    const bool one = S.type_mask != 3;
    if (S.idx16) {
      if (st16) {
        if (one) filter_tile<KLDS, true, true, true>(p, S, s, t, pin, s_pos, lds0, wst);
        else filter_tile<KLDS, true, true, false>(p, S, s, t, pin, s_pos, lds0, wst);
      } else {
        filter_tile<KLDS, true, false, false>(p, S, s, t, pin, s_pos, lds0, nullptr);
      }
    } else {
      if (st32) filter_tile<KLDS, false, true, false>(p, S, s, t, pin, s_pos, lds0, wst);
      else filter_tile<KLDS, false, false, false>(p, S, s, t, pin, s_pos, lds0, nullptr);
    }
  }
}

// 8 waves per SIMD: two blocks per compute unit
template <int LPB, int MODE>
__global__ __launch_bounds__(STEP_THREADS, 8) void k_step(DevPtrs p, int kstep, int tier) {
  extern __shared__ float4 s_pos[];
  step_items<LPB, MODE, true>(p, kstep, s_pos, tier);
}

template <int LPB, int MODE>
__global__ __launch_bounds__(STEP_THREADS) void k_step_far(DevPtrs p, int kstep) {
  step_items<LPB, MODE, false>(p, kstep, nullptr, 0);
}

// ------------------------------------------------ energies / observables

// 16 lanes per bead; FP32 pair energies, FP64 accumulation; pairs are seen
// from both beads and weighted 1/2.  Native contacts (d/r0 < 1.25 over all
// natives) are counted from the lower index.
__global__ __launch_bounds__(256) void k_energy(DevPtrs p) {
  constexpr int LPB = 16, GPW = 4, ROUNDS = TILE / 16;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane / LPB, l = lane % LPB;
  const int nwork = *p.nwork;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const int t = p.work[w];
    if (t < 0) continue;
    const int s = p.tile_sys[t];
    const SysDev &S = p.sys[s];
    const int bead_off = S.bead_off, n = S.n;
    const int b0 = t * TILE;
    const int bend = min(b0 + TILE, bead_off + n);
    const float4 *pin = buf_in(p, S.parity);
    const float boxl = S.boxl, inv_boxl = S.inv_boxl;
    const bool wrapf = S.wrap_free != 0 || boxl <= 0.0f;
    for (int r = 0; r < ROUNDS; ++r) {
      const int i = b0 + r * (4 * GPW) + wave * GPW + g;
      double eb = 0.0, ea = 0.0, en = 0.0, er = 0.0, vir = 0.0, con = 0.0;
      if (i < bend) {
        const float4 xi = pin[i];
        const int wi = __float_as_int(xi.w);
        const int row = ((wi & 3) - 1) * 2;
        if (l < 4) {
          const int off = l < 2 ? l - 2 : l - 1;
          const int j = i + off;
          if (j >= bead_off && j < bead_off + n) {
            const float4 xj = pin[j];
            if ((__float_as_int(xj.w) >> 2) == (wi >> 2)) {
              float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
              if (!wrapf) {
                dx = mi_fast(dx, boxl, inv_boxl);
                dy = mi_fast(dy, boxl, inv_boxl);
                dz = mi_fast(dz, boxl, inv_boxl);
              }
              const float d2 = dx * dx + dy * dy + dz * dz;
              if (off == 1 || off == -1) {
                const float r0 = p.bond_r0[off < 0 ? j : i];
                const float d = sqrtf(d2);
                const float dev = d - r0;
                const float gg = 1.0f - dev * dev * S.inv_R0sq;
                eb += 0.5 * (double)(-S.half_k_R0sq * log1pf(-dev * dev * S.inv_R0sq));
                const float c = S.k_bnd * dev / (d * gg);
                vir += 0.5 * (double)(-c * d2);
              } else {
                const float ir2 = 1.0f / d2;
                const float ir6 = ir2 * ir2 * ir2;
                ea += 0.5 * (double)(S.ss_e6 * ir6);
                vir += 0.5 * (double)(S.ss_f6 * ir6);
              }
            }
          }
        }
        const int ns = p.nat_start[i], po = p.platt_off[i];
        const int nc = p.platt_cnt[i];
        for (int k = l; k < nc; k += LPB) {
          const int2 pe = p.platt[po + k];
          const int j = pe.x;
          const float r0 = __int_as_float(pe.y);
          const float4 xj = pin[j];
          float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
          if (!wrapf) {
            dx = mi_fast(dx, boxl, inv_boxl);
            dy = mi_fast(dy, boxl, inv_boxl);
            dz = mi_fast(dz, boxl, inv_boxl);
          }
          const float d2 = dx * dx + dy * dy + dz * dz;
          const float eps = w_type(xj.w) == 2 ? S.eps_att[row + 1] : S.eps_att[row];
          const float r02 = r0 * r0;
          const float q2 = r02 / d2;
          const float q6 = q2 * q2 * q2;
          en += 0.5 * (double)(eps * (q6 * q6 - 2.0f * q6));
          if (d2 > 1.0e-7f * r02) vir += 0.5 * (double)(12.0f * eps * q6 * (q6 - 1.0f));
        }
        // contacts over all natives, counted once (partner index above i)
        const int nn = p.nat_cnt[i];
        for (int k = l; k < nn; k += LPB) {
          const int j = p.nat_j[ns + k];
          if (j > i) {
            const float d = sqrtf(dist2_exact(xi, pin[j], boxl));
            if (d / p.nat_r0[ns + k] < 1.25f) con += 1.0;
          }
        }
        const int64_t pb = S.pl_base + p.plrep_off[i];
        const int nr = p.plrep_cnt[i];
        for (int k = l; k < nr; k += LPB) {
          const int j = S.idx16 ? bead_off + ((const uint16_t *)(p.plrep_j + S.pl_base))[pb - S.pl_base + k]
                                : p.plrep_j[pb + k];
          const float4 xj = pin[j];
          float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
          if (!wrapf) {
            dx = mi_fast(dx, boxl, inv_boxl);
            dy = mi_fast(dy, boxl, inv_boxl);
            dz = mi_fast(dz, boxl, inv_boxl);
          }
          const float d2 = dx * dx + dy * dy + dz * dz;
          const bool tb = w_type(xj.w) == 2;
          const float q2 = (tb ? S.sig2_rep[row + 1] : S.sig2_rep[row]) / d2;
          const float q6 = q2 * q2 * q2;
          const float ep = tb ? S.eps_rep[row + 1] : S.eps_rep[row];
          er += 0.5 * (double)(ep * (S.c12 * q6 * q6 + S.c6 * q6));
          if (d2 > (tb ? S.rep_guard[row + 1] : S.rep_guard[row]))
            vir += 0.5 * (double)(-((tb ? S.rep_m12[row + 1] : S.rep_m12[row]) * q6 +
                                    (tb ? S.rep_m6[row + 1] : S.rep_m6[row])) * q6);
        }
      }
#pragma unroll
      for (int o = LPB / 2; o >= 1; o >>= 1) {
        eb += __shfl_xor(eb, o);
        ea += __shfl_xor(ea, o);
        en += __shfl_xor(en, o);
        er += __shfl_xor(er, o);
        vir += __shfl_xor(vir, o);
        con += __shfl_xor(con, o);
      }
      if (l == 0 && i < bend) {
        double *o = p.epart + (int64_t)(i - p.epart_base) * EPART_N;
        o[0] = eb; o[1] = ea; o[2] = en; o[3] = er; o[4] = vir; o[5] = con;
      }
    }
  }
}

// Per system, fixed-order reduction of the bead partials plus Rg^2 from the
// centroid form (1/N) sum |r|^2 - |r_cm|^2, identical to the reference's
// pair sum / N^2 (sop.cu:564-580).
__global__ __launch_bounds__(256) void k_sys_reduce(DevPtrs p) {
  __shared__ double sh[4][10];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nsys = *p.nsys;
  for (int q = blockIdx.x; q < nsys; q += gridDim.x) {
    const int s = p.syslist[q];
    const SysDev &S = p.sys[s];
    const float4 *pin = buf_in(p, S.parity);
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = 0.0;
    for (int b = tid; b < S.n; b += blockDim.x) {
      const int i = S.bead_off + b;
      const double *e = p.epart + (int64_t)(i - p.epart_base) * EPART_N;
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k] += e[k];
      const float4 x = pin[i];
      acc[6] += x.x;
      acc[7] += x.y;
      acc[8] += x.z;
      acc[9] += (double)x.x * x.x + (double)x.y * x.y + (double)x.z * x.z;
    }
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = wave_sum_d(acc[k]);
    if (lane == 0)
      for (int k = 0; k < 10; ++k) sh[wave][k] = acc[k];
    __syncthreads();
    if (tid == 0) {
      double tot[10];
      for (int k = 0; k < 10; ++k) tot[k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
      double *o = p.sysout + (int64_t)(s - p.sysout_base) * SYSOUT_N;
      o[0] = tot[0]; o[1] = tot[1]; o[2] = tot[2]; o[3] = tot[3];
      o[4] = tot[0] + tot[1] + tot[2] + tot[3];
      o[5] = tot[4];
      o[6] = tot[5];
      const double inv = 1.0 / (double)S.n;
      const double cx = tot[6] * inv, cy = tot[7] * inv, cz = tot[8] * inv;
      o[7] = tot[9] * inv - (cx * cx + cy * cy + cz * cz);
      o[8] = tot[6]; o[9] = tot[7]; o[10] = tot[8];
      o[11] = (double)S.n;
    }
    __syncthreads();
  }
}

// Per system: open a run of nsteps (the step kernel moves a system while
// step < step_end).
__global__ void k_begin(DevPtrs p, int64_t nsteps) {
  const int nsys = *p.nsys;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nsys; q += gridDim.x * blockDim.x) {
    SysDev &S = p.sys[p.syslist[q]];
    S.step_end = S.step + nsteps;
  }
}

// One system's shell rows and packed pair-list rows moved to a segment of
// another row capacity (wave per tile): each row keeps its entries and its
// place in the tile's order, so the lists, and every result computed from
// them, are unchanged.  The two segments do not overlap.
template <typename T>
__device__ __forceinline__ void relayout_tile(const DevPtrs &p, int bo, int n, int t, int64_t ob, int oc, int64_t nb,
                                              int nc) {
  const int lane = threadIdx.x & 63;
  const T *so = reinterpret_cast<const T *>(p.slrep_j + ob);
  T *sn = reinterpret_cast<T *>(p.slrep_j + nb);
  const T *po = reinterpret_cast<const T *>(p.plrep_j + ob);
  T *pn = reinterpret_cast<T *>(p.plrep_j + nb);
  const int shift = t * TILE * (nc - oc);  // the tile's region moves with the capacity
  for (int q = 0; q < TILE; ++q) {
    const int li = t * TILE + q;
    if (li >= n) break;
    const int i = bo + li;
    const int cs = p.slrep_cnt[i];
    for (int k = lane; k < cs; k += 64) sn[(int64_t)li * nc + k] = so[(int64_t)li * oc + k];
    const int off = p.plrep_off[i];
    const int len = (p.plrep_cnt[i] + ROW_PAD - 1) & ~(ROW_PAD - 1);
    for (int k = lane; k < len; k += 64) pn[off + shift + k] = po[off + k];
    if (lane == 0) p.plrep_off[i] = off + shift;
  }
}

__global__ __launch_bounds__(64) void k_relayout(DevPtrs p, int bo, int n, int idx16, int64_t ob, int oc, int64_t nb,
                                                 int nc) {
  const int t = blockIdx.x;
  if (idx16) relayout_tile<uint16_t>(p, bo, n, t, ob, oc, nb, nc);
  else relayout_tile<uint32_t>(p, bo, n, t, ob, oc, nb, nc);
}

// Per system: advance the step counter after a block of kernel steps.
__global__ void k_advance(DevPtrs p, int nsteps) {
  const int nsys = *p.nsys;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nsys; q += gridDim.x * blockDim.x) {
    SysDev &S = p.sys[p.syslist[q]];
    const int64_t st = S.step + nsteps;
    S.step = st < S.step_end ? st : (S.step_end > S.step ? S.step_end : S.step);
    S.parity = (S.parity + nsteps) & 1;
  }
}

// ------------------------------------------------------------ self test
// Lane-layout checks of the DPP / swizzle helpers against shuffles, and of
// the device normal generator; run by the GPU tests before the engine.
__global__ void k_selftest(int *out, float *fout) {
  const int lane = threadIdx.x;
  const int v = (lane * 7 + 3) % 11;
  int ref = 0;
  for (int k = 0; k <= lane; ++k) ref += (k * 7 + 3) % 11;
  const int sc = wave_incl_scan(v);
  int bad = sc != ref;
  const float fv = (float)((lane * 5) % 13);
  float r4 = 0, r8 = 0, r16 = 0, r32 = 0, r64 = 0;
  for (int k = 0; k < 64; ++k) {
    const float x = (float)((k * 5) % 13);
    if (k / 4 == lane / 4) r4 += x;
    if (k / 8 == lane / 8) r8 += x;
    if (k / 16 == lane / 16) r16 += x;
    if (k / 32 == lane / 32) r32 += x;
    r64 += x;
  }
  bad |= (group_sum<4>(fv) != r4) << 1;
  bad |= (group_sum<8>(fv) != r8) << 2;
  bad |= (group_sum<16>(fv) != r16) << 3;
  bad |= (group_sum<32>(fv) != r32) << 4;
  bad |= (group_sum<64>(fv) != r64) << 5;
  const uint64_t m = __ballot(lane % 3 == 0);
  int below = 0;
  for (int k = 0; k < lane; ++k) below += (k % 3 == 0);
  bad |= (lanes_below(m) != below) << 6;
  out[lane] = bad;
  float n0, n1, n2;
  normals3(0x12345678u, 0x9abcdef0u, (uint32_t)lane, 1000 + lane, n0, n1, n2);
  fout[3 * lane] = n0;
  fout[3 * lane + 1] = n1;
  fout[3 * lane + 2] = n2;
}

// --------------------------------------------------------------- launchers

#define SOPK_LAUNCH_CHECK() hipGetLastError()

// The whole-CU and global-gather launches of the step kernel and of the shell
// filter go on a side stream beside the two-blocks-per-unit LDS launch when
// both sides have work: fork (side waits for st) and join (st waits for
// side).  Under graph capture these become two graph branches.
static bool fork_wanted(const StepFork *fk, const StepGrids &sg) {
  return fk && sg.grid > 0 && (sg.grid_big > 0 || sg.grid_far > 0);
}

static hipError_t fork_begin(const StepFork *fk, hipStream_t st) {
  hipError_t h = hipEventRecord(fk->fork, st);
  if (h == hipSuccess) h = hipStreamWaitEvent(fk->side, fk->fork, 0);
  return h;
}

static hipError_t fork_end(const StepFork *fk, hipStream_t st) {
  hipError_t h = hipEventRecord(fk->join, fk->side);
  if (h == hipSuccess) h = hipStreamWaitEvent(st, fk->join, 0);
  return h;
}

// The shell filter's launches (LDS items, whole-CU items, global-gather
// items) on one stream, for the systems whose build is one of `builds`.
static void launch_filter(const DevPtrs &p, const StepGrids &sg, int kstep, int builds, hipStream_t st) {
  // row staging after the staged positions where LDS has room for it
  const bool stage0 = sg.lds + FILTER_STAGE_BYTES <= LDS_BYTES_MAX;
  if (sg.grid > 0)
    hipLaunchKernelGGL(k_shell_filter<true>, dim3(sg.grid), dim3(STEP_THREADS),
                       sg.lds + (stage0 ? FILTER_STAGE_BYTES : 0), st, p, kstep, 0, stage0 ? (int)sg.lds : -1,
                       stage0 ? FILTER_STAGE_BYTES : 0, builds);
  if (sg.grid_big > 0)
    hipLaunchKernelGGL(k_shell_filter<true>, dim3(sg.grid_big), dim3(STEP_THREADS), sg.lds_big, st, p, kstep, 1,
                       -1, 0, builds);
  // global-gather items stage no coordinates: room for 32-bit rows
  if (sg.grid_far > 0)
    hipLaunchKernelGGL(k_shell_filter<false>, dim3(sg.grid_far), dim3(STEP_THREADS), 2 * FILTER_STAGE_BYTES, st,
                       p, kstep, 0, 0, 2 * FILTER_STAGE_BYTES, builds);
}

// One list-rebuild point for the current batch: snapshot / decision, cell
// grid and counting sort, then whichever of the direct pair-list build or the
// neighbor-list materialization + filter each system needs.
hipError_t launch_lists(const DevPtrs &p, const ListGrids &lg, int kstep, int flags, hipStream_t st,
                        const StepFork *fk) {
  const int grid_tiles = lg.grid_tiles, grid_sys = lg.grid_sys;
  hipLaunchKernelGGL(k_skin, dim3(grid_sys), dim3(SKIN_THREADS), 0, st, p, kstep, flags);
  hipLaunchKernelGGL(k_bbox, dim3(grid_sys), dim3(256), 0, st, p, kstep);
  // Systems whose pair list comes from their shell rows need nothing from
  // the cell grid: with a second stream (fk) their filter runs beside the
  // grid chain, and the filter of the grid-built systems follows the chain on
  // the second stream.  The two filters take disjoint systems.
  constexpr int ALL = (1 << BUILD_DIRECT) | (1 << BUILD_SHELL) | (1 << BUILD_NL);
  constexpr int SHELL_ONLY = 1 << BUILD_SHELL;
  hipStream_t cs = st;  // the grid chain's stream
  if (fk) {
    const hipError_t h = fork_begin(fk, st);
    if (h != hipSuccess) return h;
    cs = fk->side;
  }
  hipLaunchKernelGGL(k_cell_sort_lds, dim3(grid_sys), dim3(SORT_THREADS), 0, cs, p, kstep);
  hipLaunchKernelGGL(k_cell_count, dim3(grid_tiles), dim3(64), 0, cs, p, kstep);
  hipLaunchKernelGGL(k_cell_scan, dim3(grid_sys), dim3(256), 0, cs, p);
  hipLaunchKernelGGL(k_cell_scatter, dim3(grid_tiles), dim3(64), 0, cs, p);
  hipLaunchKernelGGL(k_cell_rank, dim3(grid_tiles), dim3(64), 0, cs, p, kstep);
  hipLaunchKernelGGL(k_cellwork, dim3(1), dim3(256), 0, cs, p);
  hipLaunchKernelGGL(k_cellexpand, dim3(grid_sys), dim3(256), 0, cs, p);
  hipLaunchKernelGGL(k_pl_cells, dim3(grid_tiles), dim3(256), 0, cs, p, kstep);
  hipLaunchKernelGGL(k_tile_order, dim3(grid_tiles), dim3(256), 0, cs, p);
  // neighbor-list builds: their rows before the shell filter packs them
  hipLaunchKernelGGL(k_nl_build, dim3(grid_tiles), dim3(256), 0, cs, p, kstep);
  if (fk) {
    launch_filter(p, lg.fg, kstep, ALL & ~SHELL_ONLY, cs);
    launch_filter(p, lg.fg, kstep, SHELL_ONLY, st);
    const hipError_t h = fork_end(fk, st);
    if (h != hipSuccess) return h;
  } else {
    launch_filter(p, lg.fg, kstep, ALL, st);
  }
  return SOPK_LAUNCH_CHECK();
}

template <int MODE>
static hipError_t launch_step_mode(const DevPtrs &p, int lpb, const StepGrids &sg, int kstep, hipStream_t st,
                                   const StepFork *fk) {
  const bool fork = fork_wanted(fk, sg);
  if (fork) {
    const hipError_t h = fork_begin(fk, st);
    if (h != hipSuccess) return h;
  }
  const hipStream_t far_st = fork ? fk->side : st;
  for (int tier = 0; tier < 2; ++tier) {
    const int grid = tier ? sg.grid_big : sg.grid;
    const size_t lds = tier ? sg.lds_big : sg.lds;
    hipStream_t ts = tier ? far_st : st;
    if (grid > 0) switch (lpb) {
      case 4: hipLaunchKernelGGL((k_step<4, MODE>), dim3(grid), dim3(STEP_THREADS), lds, ts, p, kstep, tier); break;
      case 8: hipLaunchKernelGGL((k_step<8, MODE>), dim3(grid), dim3(STEP_THREADS), lds, ts, p, kstep, tier); break;
      case 32: hipLaunchKernelGGL((k_step<32, MODE>), dim3(grid), dim3(STEP_THREADS), lds, ts, p, kstep, tier); break;
      default: hipLaunchKernelGGL((k_step<16, MODE>), dim3(grid), dim3(STEP_THREADS), lds, ts, p, kstep, tier); break;
    }
  }
  const int grid_far = sg.grid_far;
  if (grid_far > 0) switch (lpb) {
    case 4: hipLaunchKernelGGL((k_step_far<4, MODE>), dim3(grid_far), dim3(STEP_THREADS), 0, far_st, p, kstep); break;
    case 8: hipLaunchKernelGGL((k_step_far<8, MODE>), dim3(grid_far), dim3(STEP_THREADS), 0, far_st, p, kstep); break;
    case 32: hipLaunchKernelGGL((k_step_far<32, MODE>), dim3(grid_far), dim3(STEP_THREADS), 0, far_st, p, kstep); break;
    default: hipLaunchKernelGGL((k_step_far<16, MODE>), dim3(grid_far), dim3(STEP_THREADS), 0, far_st, p, kstep); break;
  }
  hipError_t h = SOPK_LAUNCH_CHECK();
  if (fork && h == hipSuccess) h = fork_end(fk, st);
  return h;
}

hipError_t launch_step(const DevPtrs &p, int mode, int lpb, const StepGrids &sg, int kstep, hipStream_t st,
                       const StepFork *fk) {
  if (mode == MODE_FORCE) return launch_step_mode<MODE_FORCE>(p, lpb, sg, kstep, st, fk);
  if (mode == MODE_FORCE_NOISE) return launch_step_mode<MODE_FORCE_NOISE>(p, lpb, sg, kstep, st, fk);
  return launch_step_mode<MODE_STEP>(p, lpb, sg, kstep, st, fk);
}

// Allow the step kernels dynamic LDS up to `bytes`.
hipError_t step_lds_limit(size_t bytes) {
  hipError_t h = hipSuccess;
#define SOPK_LDS_ATTR(L, M)                                                                                      \
  if (h == hipSuccess)                                                                                           \
    h = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_step<L, M>), hipFuncAttributeMaxDynamicSharedMemorySize, \
                            (int)bytes);
  SOPK_LDS_ATTR(4, 0) SOPK_LDS_ATTR(8, 0) SOPK_LDS_ATTR(16, 0) SOPK_LDS_ATTR(32, 0)
  SOPK_LDS_ATTR(4, 1) SOPK_LDS_ATTR(8, 1) SOPK_LDS_ATTR(16, 1) SOPK_LDS_ATTR(32, 1)
  SOPK_LDS_ATTR(4, 2) SOPK_LDS_ATTR(8, 2) SOPK_LDS_ATTR(16, 2) SOPK_LDS_ATTR(32, 2)
#undef SOPK_LDS_ATTR
  return h;
}

// Allow the shell filter the dynamic LDS its launches take (a compute unit's
// whole LDS for the LDS items, row staging for the global-gather ones).
hipError_t filter_lds_init() {
  hipError_t h = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_shell_filter<true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES_MAX);
  if (h == hipSuccess)
    h = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_shell_filter<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)(2 * FILTER_STAGE_BYTES));
  return h;
}

hipError_t launch_relayout(const DevPtrs &p, int bo, int n, int idx16, int64_t old_base, int old_cap,
                           int64_t new_base, int new_cap, hipStream_t st) {
  hipLaunchKernelGGL(k_relayout, dim3((n + TILE - 1) / TILE), dim3(64), 0, st, p, bo, n, idx16, old_base, old_cap,
                     new_base, new_cap);
  return SOPK_LAUNCH_CHECK();
}

hipError_t launch_energy(const DevPtrs &p, int grid_tiles, int grid_sys, hipStream_t st) {
  hipLaunchKernelGGL(k_energy, dim3(grid_tiles), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_sys_reduce, dim3(grid_sys), dim3(256), 0, st, p);
  return SOPK_LAUNCH_CHECK();
}

hipError_t launch_advance(const DevPtrs &p, int nsteps, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(4), dim3(256), 0, st, p, nsteps);
  return SOPK_LAUNCH_CHECK();
}

hipError_t launch_begin(const DevPtrs &p, int64_t nsteps, hipStream_t st) {
  hipLaunchKernelGGL(k_begin, dim3(4), dim3(256), 0, st, p, nsteps);
  return SOPK_LAUNCH_CHECK();
}

// ------------------------------------------ allocation-free energy entry point
//
// The work table, counters and energy partials of sop_energy_eval_noalloc
// live in this file-scope array instead of memory the library allocates.
// Host code never passes the array's name to a kernel or a copy (in host
// code it stands for no device address): the address of each device's copy
// comes from hipGetSymbolAddress, once per device, and only that address is
// used.  Every call on one device shares that device's copy: calls that
// overlap on one device race.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device ordinal -> address of its copy of scratch
}  // namespace

// The device address of `device`'s copy of scratch; the device must be current.
hipError_t noalloc_scratch(int device, void **ptr) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  const hipError_t h = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (h != hipSuccess) return h;
  g_scratch_ptr[device] = p;
  *ptr = p;
  return hipSuccess;
}

// Work table of one system (tiles t0 .. t0 + nt - 1) and its counters, in
// the scratch that p.work / p.nwork / p.nsys / p.syslist point into.
__global__ void k_noalloc_setup(DevPtrs p, int t0, int nt, int slot) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) p.work[k] = t0 + k;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *p.nwork = nt;
    *p.nsys = 1;
    p.syslist[0] = slot;
  }
}

hipError_t launch_noalloc_energy(const DevPtrs &p, int t0, int nt, int slot, int grid_tiles, hipStream_t st) {
  hipLaunchKernelGGL(k_noalloc_setup, dim3(std::max(1, std::min(64, (nt + 255) / 256))), dim3(256), 0, st, p, t0,
                     nt, slot);
  hipLaunchKernelGGL(k_energy, dim3(grid_tiles), dim3(256), 0, st, p);
  hipLaunchKernelGGL(k_sys_reduce, dim3(1), dim3(256), 0, st, p);
  return SOPK_LAUNCH_CHECK();
}

hipError_t launch_selftest(int *out, float *fout, hipStream_t st) {
  hipLaunchKernelGGL(k_selftest, dim3(1), dim3(64), 0, st, out, fout);
  return SOPK_LAUNCH_CHECK();
}

} // namespace sopk
