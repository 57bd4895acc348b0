// kernels_coarse.hip -- coarse-grid correction of the relaxation (opt-in,
// rgbdflow_params.coarse_levels; not in the reference).
//
// Coarse level l+1 aggregates the 2x2 pixels (2I..2I+1, 2J..2J+1) of level l
// (one row or column where a size is odd).  Its system has the relaxed
// system's form -- f32 records of 5-point systems with a 3x3 block per pixel
// -- so the tiled and plain relaxation kernels run on it unchanged:
//   g_right(I,J), g_down(I,J): the fine edge weights (xy and z alike) that
//     cross to the right / lower aggregate, summed;
//   edges: g / 2, the crossing weights' mean;
//   block: the children's blocks summed, plus 2 e on the diagonal for every
//     edge inside the aggregate, plus 0.5 * (g_left + g_up + g_right +
//     g_down) -- the children's data terms, less exactly the coarse edges.
// Every sum starts from zero in f32, in the order the comments give, so the
// records equal the oracle's bit for bit.  One V-cycle: for each level, the
// finer level's residual summed over each aggregate into the coarse rhs and
// the coarse unknowns zeroed (k_cg_restrict), relaxation; back up, more
// relaxation and the coarse unknowns added to their fine pixels
// (k_cg_prolong); the levels that fit in LDS together (kCgFusePx pixels) do
// their part of the V-cycle in one workgroup per pair (k_cg_bottom).  A
// coarse level's SolveCtl::outer_done says, for the current V-cycle, that the
// relaxed level has stopped: then nothing is changed.
#include <hip/hip_runtime.h>

#include "records.hpp"
#include "relax_pixel.hpp"

namespace rgbdflow {

namespace {

// The 13 values of pixel l's record (any storage).
__device__ __forceinline__ void coefs(const float4* base, int storage, size_t N, size_t l, float (&c)[kCoef]) {
  float y[3];
  load_record(base, storage, N, l, c, y);
}

// The four edge weights (eRxy, eRz, eDxy, eDz) of pixel l: the first part
// plane of an f32 record, else the decoded record.
__device__ __forceinline__ void edges(const float4* base, int storage, size_t N, size_t l, float (&e)[4]) {
  if (storage == kStoreF32) {
    const float4 q = gload(base + l);
    e[0] = q.x;
    e[1] = q.y;
    e[2] = q.z;
    e[3] = q.w;
    return;
  }
  float c[kCoef];
  coefs(base, storage, N, l, c);
  e[0] = c[kERxy];
  e[1] = c[kERz];
  e[2] = c[kEDxy];
  e[3] = c[kEDz];
}

// An aggregate's children (2I,2J), (2I,2J+1), (2I+1,2J), (2I+1,2J+1) with
// their records (zero where a child is past the level's edge) and the edge
// weights reaching them from outside: the right edges of column 2J-1 (rows
// 2I, 2I+1) and the down edges of row 2I-1 (columns 2J, 2J+1), zero where
// those pixels do not exist.
struct Aggregate {
  bool r1, c1;         // second row / column present
  size_t p00;          // pixel index of (2I, 2J)
  float a[4][kCoef];
  float left[2][2];    // [row][xy, z]: right edges of (2I + row, 2J - 1)
  float up[2][2];      // [column][xy, z]: down edges of (2I - 1, 2J + column)
};

__device__ __forceinline__ void load_aggregate(const LevelView::Finer& f, int I, int J, Aggregate& g) {
  const size_t N = (size_t)f.W * f.H;
  const int i0 = 2 * I, j0 = 2 * J;
  g.r1 = i0 + 1 < f.H;
  g.c1 = j0 + 1 < f.W;
  g.p00 = (size_t)i0 * f.W + j0;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const bool has = ((d & 1) ? g.c1 : true) && ((d & 2) ? g.r1 : true);
    if (has) {
      coefs(f.coef, f.storage, N, g.p00 + (d >> 1) * f.W + (d & 1), g.a[d]);
    } else {
#pragma unroll
      for (int k = 0; k < kCoef; k++) g.a[d][k] = 0.f;
    }
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    float e[4] = {0.f, 0.f, 0.f, 0.f};
    if (j0 > 0 && (r == 0 || g.r1)) edges(f.coef, f.storage, N, g.p00 + r * f.W - 1, e);
    g.left[r][0] = e[0];
    g.left[r][1] = e[1];
    float u[4] = {0.f, 0.f, 0.f, 0.f};
    if (i0 > 0 && (r == 0 || g.c1)) edges(f.coef, f.storage, N, g.p00 - f.W + r, u);
    g.up[r][0] = u[2];
    g.up[r][1] = u[3];
  }
}

// The coarse rhs of an aggregate: each child's residual r - sum_n e_n x_n -
// B x for its three unknowns, in f32 in the order left, up, right, down, then
// the block's row (missing neighbours skipped), summed over the children in
// order.  X: the unknowns of rows i0-1 .. i0+2, columns j0-1 .. j0+2 (zero
// outside the level; the corners are not read); W, H: the finer level.
__device__ __forceinline__ void aggregate_rhs(const Aggregate& g, const float4 (&X)[4][4], int W, int H, int i0,
                                              int j0, float (&s)[3]) {
  s[0] = s[1] = s[2] = 0.f;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int dr = d >> 1, dc = d & 1, i = i0 + dr, j = j0 + dc;
    if (i >= H || j >= W) continue;
    const float(&c)[kCoef] = g.a[d];
    const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
    // the left / up edges: a sibling's right / down edges, or the outside ones
    const float lxy = dc ? g.a[d - 1][kERxy] : g.left[dr][0], lz = dc ? g.a[d - 1][kERz] : g.left[dr][1];
    const float uxy = dr ? g.a[d - 2][kEDxy] : g.up[dc][0], uz = dr ? g.a[d - 2][kEDz] : g.up[dc][1];
    const float4 o = X[1 + dr][1 + dc], L = X[1 + dr][dc], U = X[dr][1 + dc], R = X[1 + dr][2 + dc],
                 D = X[2 + dr][1 + dc];
    const float xo[3] = {o.x, o.y, o.z}, xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z},
                xr[3] = {R.x, R.y, R.z}, xd[3] = {D.x, D.y, D.z};
    const float B[3][3] = {{c[kBuu], c[kBuv], c[kBuw]}, {c[kBuv], c[kBvv], c[kBvw]}, {c[kBuw], c[kBvw], c[kBww]}};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      float a = c[kRu + k];
      if (hl) a -= (k < 2 ? lxy : lz) * xl[k];
      if (hu) a -= (k < 2 ? uxy : uz) * xu[k];
      if (hr) a -= (k < 2 ? c[kERxy] : c[kERz]) * xr[k];
      if (hd) a -= (k < 2 ? c[kEDxy] : c[kEDz]) * xd[k];
      a -= B[k][0] * xo[0];
      a -= B[k][1] * xo[1];
      a -= B[k][2] * xo[2];
      s[k] += a;
    }
  }
}

// The coarse pixel of a block's thread, or false past the level.
__device__ __forceinline__ bool coarse_pixel(const Views& w, int blk, const LevelView*& v, int& I, int& J) {
  const int item = work_item(w, blk);
  v = &w.items[item];
  const int L = (blk - w.start[item]) * kBlock + threadIdx.x;
  if (L >= v->N) return false;
  I = L / v->W;
  J = L - I * v->W;
  return true;
}

}  // namespace

// The coarse operator of one level from the next finer one (rhs zero); runs
// after each assembly of the relaxed level's system.
__global__ void __launch_bounds__(kBlock) k_cg_operator(Views w) {
  const LevelView* v;
  int I, J;
  if (!coarse_pixel(w, xcd_block(blockIdx.x, gridDim.x), v, I, J)) return;
  const LevelView::Finer& f = v->finer;
  if (f.top_ctl->outer_done) return;
  const int Wc = v->W, Hc = v->H;
  Aggregate g;
  load_aggregate(f, I, J, g);
  const bool r1 = g.r1, c1 = g.c1;
  // g_right: the right edges of the children in column 2J+1; g_down: the
  // down edges of the children in row 2I+1 (zero at the level's right /
  // lower edge); g_left, g_up: the left / upper aggregates' g_right, g_down
  float gr[2], gd[2], gl[2], gu[2];
#pragma unroll
  for (int z = 0; z < 2; z++) {
    gr[z] = 0.f;
    if (J + 1 < Wc) {
      gr[z] += g.a[1][kERxy + z];
      if (r1) gr[z] += g.a[3][kERxy + z];
    }
    gd[z] = 0.f;
    if (I + 1 < Hc) {
      gd[z] += g.a[2][kEDxy + z];
      if (c1) gd[z] += g.a[3][kEDxy + z];
    }
    gl[z] = 0.f;
    gl[z] += g.left[0][z];
    if (r1) gl[z] += g.left[1][z];
    gu[z] = 0.f;
    gu[z] += g.up[0][z];
    if (c1) gu[z] += g.up[1][z];
  }
  float c[kCoef];
  float around[2];
#pragma unroll
  for (int z = 0; z < 2; z++) {
    float t = 0.f;
    if (J > 0) t += gl[z];
    if (I > 0) t += gu[z];
    t += gr[z];
    t += gd[z];
    around[z] = t;
  }
  c[kERxy] = 0.5f * gr[0];
  c[kERz] = 0.5f * gr[1];
  c[kEDxy] = 0.5f * gd[0];
  c[kEDz] = 0.5f * gd[1];
  const int blk[6] = {kBuu, kBvv, kBww, kBuv, kBuw, kBvw};
#pragma unroll
  for (int b = 0; b < 6; b++) {
    const int k = blk[b];
    float s = 0.f;
    s += g.a[0][k];
    if (c1) s += g.a[1][k];
    if (r1) s += g.a[2][k];
    if (r1 && c1) s += g.a[3][k];
    if (b < 3) {
      // Buu, Bvv: the xy edges; Bww: the z edges.  Inner edges: right of
      // (2I,2J), right of (2I+1,2J), down of (2I,2J), down of (2I,2J+1)
      const int er = b < 2 ? kERxy : kERz, ed = b < 2 ? kEDxy : kEDz;
      if (c1) s += 2.f * g.a[0][er];
      if (c1 && r1) s += 2.f * g.a[2][er];
      if (r1) s += 2.f * g.a[0][ed];
      if (r1 && c1) s += 2.f * g.a[1][ed];
      s += 0.5f * around[b < 2 ? 0 : 1];
    }
    c[k] = s;
  }
  c[kRu] = c[kRv] = c[kRw] = 0.f;
  store_record(v->coef, kStoreF32, (size_t)v->N, (size_t)(I * Wc + J), c);
}

// The coarse rhs: the finer level's residuals at its current unknowns
// (finer_alt: in du_alt) summed over each aggregate, children in order; the
// coarse unknowns zeroed in du.  slot_base / n_prev_checks: the relaxed
// level's stop test (as its relaxation launches make it).
__global__ void __launch_bounds__(kBlock) k_cg_restrict(Views w, int finer_alt, int slot_base, int n_prev_checks) {
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const LevelView* v;
  int I, J;
  const bool in = coarse_pixel(w, blk, v, I, J);
  const LevelView::Finer& f = v->finer;
  const bool stopped = f.first ? (f.ctl->outer_done || relax_stopped(f.ctl, slot_base, n_prev_checks))
                               : f.ctl->outer_done != 0;
  if (blk == w.start[work_item(w, blk)] && threadIdx.x == 0) v->ctl->outer_done = stopped ? 1 : 0;
  if (!in || stopped) return;
  const float4* x = finer_alt ? f.du_alt : f.du;
  Aggregate g;
  load_aggregate(f, I, J, g);
  const int W = f.W, H = f.H, i0 = 2 * I, j0 = 2 * J;
  // the unknowns the children's stencils reach: rows i0-1 .. i0+2, columns
  // j0-1 .. j0+2 (not the corners), zero outside the level
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 X[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = i0 - 1 + r, j = j0 - 1 + q;
      const bool corner = (r == 0 || r == 3) && (q == 0 || q == 3);
      X[r][q] = (!corner && i >= 0 && i < H && j >= 0 && j < W) ? gload(x + (size_t)i * W + j) : zero;
    }
  float s[3];
  aggregate_rhs(g, X, W, H, i0, j0, s);
  const size_t L = (size_t)I * v->W + J, Nc = (size_t)v->N;
  // rhs slots of an f32 record: part 2 .z, .w and part 3 .x
  float* p2 = reinterpret_cast<float*>(v->coef + 2 * Nc + L);
  float* p3 = reinterpret_cast<float*>(v->coef + 3 * Nc + L);
  gp(p2)[2] = s[0];
  gp(p2)[3] = s[1];
  gp(p3)[0] = s[2];
  gstore(v->du + L, make_float4(0.f, 0.f, 0.f, 0.f));
}

// x_finer += P x_coarse: each coarse pixel's unknowns (alt: in du_alt) added
// to its fine pixels' (finer_alt: in the finer level's du_alt).
__global__ void __launch_bounds__(kBlock) k_cg_prolong(Views w, int alt, int finer_alt) {
  const LevelView* v;
  int I, J;
  if (!coarse_pixel(w, xcd_block(blockIdx.x, gridDim.x), v, I, J)) return;
  if (v->ctl->outer_done) return;
  const LevelView::Finer& f = v->finer;
  const float4 xc = gload((alt ? v->du_alt : v->du) + I * v->W + J);
  float4* x = finer_alt ? f.du_alt : f.du;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
    if (i < f.H && j < f.W) {
      float4 o = gload(x + i * f.W + j);
      o.x += xc.x;
      o.y += xc.y;
      o.z += xc.z;
      gstore(x + i * f.W + j, o);
    }
  }
}

// The coarse levels of an item from the first whose remaining levels fit
// kCgFusePx pixels down to the last, in one workgroup with every level in
// LDS: the operators (made by k_cg_operator) loaded, the first level's rhs
// restricted from the finer level in HBM, then that part of the V-cycle --
// restriction, relaxation (the plain kernel's pixel update, one colour pass
// per barrier) and prolongation -- in LDS, and the correction added to the
// finer level's unknowns in HBM.  The same arithmetic in the same order as the
// per-level kernels, so the results are theirs.
__global__ void __launch_bounds__(kCgBottomThreads) k_cg_bottom(Views w, int alt_mask, int slot_base,
                                                                 int n_prev_checks, int sweeps) {
  __shared__ float Rs[kCoef][kCgFusePx];  // records: the Coef values per pixel
  __shared__ float Xs[3][kCgFusePx];      // unknowns
  __shared__ int lw[kMaxCg], lh[kMaxCg], lo[kMaxCg];
  constexpr int NT = kCgBottomThreads;
  const int tid = threadIdx.x;
  const LevelView& v = w.items[blockIdx.x];
  const LevelView::Finer& f = v.finer;
  const bool stopped = f.first ? (f.ctl->outer_done || relax_stopped(f.ctl, slot_base, n_prev_checks))
                               : f.ctl->outer_done != 0;
  if (stopped) return;
  const int nlev = f.fused;
  if (tid == 0) {
    int W = v.W, H = v.H, o = 0;
    for (int m = 0; m < nlev; m++) {
      lw[m] = W;
      lh[m] = H;
      lo[m] = o;
      o += W * H;
      W = (W + 1) / 2;
      H = (H + 1) / 2;
    }
  }
  __syncthreads();
  // the operators: level m's records follow level m-1's in the coarse buffers
  {
    const float4* g = v.coef;
    for (int m = 0; m < nlev; m++) {
      const int N = lw[m] * lh[m], o = lo[m];
      for (int p = tid; p < N; p += NT) {
        float c[kCoef];
        coefs(g, kStoreF32, (size_t)N, (size_t)p, c);
#pragma unroll
        for (int k = 0; k < kRu; k++) Rs[k][o + p] = c[k];
      }
      g += (size_t)kRec * N;
    }
  }
  // the first level's rhs from the finer level's residual (in HBM)
  {
    const float4* x = ((alt_mask >> f.level) & 1) ? f.du_alt : f.du;
    const int W0 = lw[0], N0 = W0 * lh[0];
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = tid; p < N0; p += NT) {
      const int I = p / W0, J = p - I * W0, i0 = 2 * I, j0 = 2 * J;
      Aggregate g;
      load_aggregate(f, I, J, g);
      float4 X[4][4];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int i = i0 - 1 + r, j = j0 - 1 + q;
          const bool corner = (r == 0 || r == 3) && (q == 0 || q == 3);
          X[r][q] = (!corner && i >= 0 && i < f.H && j >= 0 && j < f.W) ? gload(x + (size_t)i * f.W + j) : zero;
        }
      float sr[3];
      aggregate_rhs(g, X, f.W, f.H, i0, j0, sr);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        Rs[kRu + k][p] = sr[k];
        Xs[k][p] = 0.f;
      }
    }
  }
  __syncthreads();
  // `sweeps` red-black iterations of level m, the plain kernel's update
  auto relax = [&](int m) {
    const int W = lw[m], H = lh[m], N = W * H, o = lo[m];
    for (int it = 0; it < sweeps; it++)
      for (int colour = 0; colour < 2; colour++) {
        for (int p = tid; p < N; p += NT) {
          const int i = p / W, j = p - i * W;
          if (((i + j) & 1) != colour) continue;
          const int q = o + p;
          const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
          PixCoef c;
          c.lxy = hl ? Rs[kERxy][q - 1] : 0.f;
          c.lz = hl ? Rs[kERz][q - 1] : 0.f;
          c.uxy = hu ? Rs[kEDxy][q - W] : 0.f;
          c.uz = hu ? Rs[kEDz][q - W] : 0.f;
          c.rxy = Rs[kERxy][q];
          c.rz = Rs[kERz][q];
          c.dxy = Rs[kEDxy][q];
          c.dz = Rs[kEDz][q];
          c.Buu = Rs[kBuu][q];
          c.Bvv = Rs[kBvv][q];
          c.Bww = Rs[kBww][q];
          c.Buv = Rs[kBuv][q];
          c.Buw = Rs[kBuw][q];
          c.Bvw = Rs[kBvw][q];
          c.ru = Rs[kRu][q];
          c.rv = Rs[kRv][q];
          c.rw = Rs[kRw][q];
          float xo[3], xl[3], xu[3], xr[3], xd[3], xn[3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            xo[k] = Xs[k][q];
            xl[k] = hl ? Xs[k][q - 1] : 0.f;
            xu[k] = hu ? Xs[k][q - W] : 0.f;
            xr[k] = hr ? Xs[k][q + 1] : 0.f;
            xd[k] = hd ? Xs[k][q + W] : 0.f;
          }
          if (v.pixel_gs == 2) relax_update_block(c, v.omega, v.omega_c, hl, hu, hr, hd, xl, xu, xr, xd, xo, xn);
          else relax_update(c, v.omega, v.omega_c, hl, hu, hr, hd, xl, xu, xr, xd, xo, xn, v.pixel_gs == 1);
#pragma unroll
          for (int k = 0; k < 3; k++) Xs[k][q] = xn[k];
        }
        __syncthreads();
      }
  };
  relax(0);
  for (int m = 1; m < nlev; m++) {
    // level m's rhs from level m-1's residual, in LDS
    const int Wf = lw[m - 1], Hf = lh[m - 1], of = lo[m - 1], Wc = lw[m], Nc = Wc * lh[m], oc = lo[m];
    for (int p = tid; p < Nc; p += NT) {
      const int I = p / Wc, J = p - I * Wc, i0 = 2 * I, j0 = 2 * J;
      Aggregate g;
      g.r1 = i0 + 1 < Hf;
      g.c1 = j0 + 1 < Wf;
#pragma unroll
      for (int d = 0; d < 4; d++) {
        const bool has = ((d & 1) ? g.c1 : true) && ((d & 2) ? g.r1 : true);
        const int q = of + (i0 + (d >> 1)) * Wf + j0 + (d & 1);
#pragma unroll
        for (int k = 0; k < kCoef; k++) g.a[d][k] = has ? Rs[k][q] : 0.f;
      }
#pragma unroll
      for (int r = 0; r < 2; r++) {
        const bool hl = j0 > 0 && (r == 0 || g.r1), hu = i0 > 0 && (r == 0 || g.c1);
        const int ql = of + (i0 + r) * Wf + j0 - 1, qu = of + (i0 - 1) * Wf + j0 + r;
        g.left[r][0] = hl ? Rs[kERxy][ql] : 0.f;
        g.left[r][1] = hl ? Rs[kERz][ql] : 0.f;
        g.up[r][0] = hu ? Rs[kEDxy][qu] : 0.f;
        g.up[r][1] = hu ? Rs[kEDz][qu] : 0.f;
      }
      float4 X[4][4];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int i = i0 - 1 + r, j = j0 - 1 + q;
          const bool corner = (r == 0 || r == 3) && (q == 0 || q == 3);
          const bool ok = !corner && i >= 0 && i < Hf && j >= 0 && j < Wf;
          const int t = of + i * Wf + j;
          X[r][q] = ok ? make_float4(Xs[0][t], Xs[1][t], Xs[2][t], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
      float sr[3];
      aggregate_rhs(g, X, Wf, Hf, i0, j0, sr);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        Rs[kRu + k][oc + p] = sr[k];
        Xs[k][oc + p] = 0.f;
      }
    }
    __syncthreads();
    relax(m);
  }
  for (int m = nlev - 1; m >= 0; m--) {
    relax(m);
    const int Wc = lw[m], oc = lo[m];
    if (m > 0) {
      // level m-1's unknowns += its aggregates' (in LDS)
      const int Wf = lw[m - 1], Nf = Wf * lh[m - 1], of = lo[m - 1];
      for (int p = tid; p < Nf; p += NT) {
        const int i = p / Wf, j = p - i * Wf, t = oc + (i >> 1) * Wc + (j >> 1);
#pragma unroll
        for (int k = 0; k < 3; k++) Xs[k][of + p] += Xs[k][t];
      }
      __syncthreads();
    } else {
      // the finer level's unknowns in HBM
      float4* x = ((alt_mask >> f.level) & 1) ? f.du_alt : f.du;
      const int N0 = Wc * lh[0];
      for (int p = tid; p < N0; p += NT) {
        const int I = p / Wc, J = p - I * Wc;
        const float xc[3] = {Xs[0][p], Xs[1][p], Xs[2][p]};
#pragma unroll
        for (int d = 0; d < 4; d++) {
          const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
          if (i < f.H && j < f.W) {
            float4 o = gload(x + (size_t)i * f.W + j);
            o.x += xc[0];
            o.y += xc[1];
            o.z += xc[2];
            gstore(x + (size_t)i * f.W + j, o);
          }
        }
      }
    }
  }
}

void launch_cg_bottom(hipStream_t s, const Views& v, int alt_mask, int slot_base, int n_prev_checks, int sweeps) {
  if (v.n > 0)
    hipLaunchKernelGGL(k_cg_bottom, dim3(v.n), dim3(kCgBottomThreads), 0, s, v, alt_mask, slot_base, n_prev_checks,
                       sweeps);
}

void launch_cg_operator(hipStream_t s, const Views& v, int blocks) {
  if (blocks > 0) hipLaunchKernelGGL(k_cg_operator, dim3(blocks), dim3(kBlock), 0, s, v);
}
void launch_cg_restrict(hipStream_t s, const Views& v, int blocks, int finer_alt, int slot_base, int n_prev_checks) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_cg_restrict, dim3(blocks), dim3(kBlock), 0, s, v, finer_alt, slot_base, n_prev_checks);
}
void launch_cg_prolong(hipStream_t s, const Views& v, int blocks, int alt, int finer_alt) {
  if (blocks > 0) hipLaunchKernelGGL(k_cg_prolong, dim3(blocks), dim3(kBlock), 0, s, v, alt, finer_alt);
}

}  // namespace rgbdflow
