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
// (k_cg_prolong).  A coarse level's SolveCtl::outer_done says, for the current
// V-cycle, that the relaxed level has stopped: then nothing is changed.
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

// g_right (down = false) or g_down (down = true) of coarse pixel (I, J), xy
// (z = 0) and z (z = 1) weights: rows 2I, 2I+1 of column 2J+1's right edges,
// or columns 2J, 2J+1 of row 2I+1's down edges; zero at the level's edge.
__device__ __forceinline__ void cross(const LevelView::Finer& f, int Wc, int Hc, int I, int J, bool down,
                                      float (&g)[2]) {
  const size_t N = (size_t)f.W * f.H;
  g[0] = g[1] = 0.f;
  float a[kCoef];
  if (!down) {
    if (J + 1 >= Wc) return;
    const size_t p = (size_t)(2 * I) * f.W + 2 * J + 1;
    coefs(f.coef, f.storage, N, p, a);
    g[0] += a[kERxy];
    g[1] += a[kERz];
    if (2 * I + 1 < f.H) {
      coefs(f.coef, f.storage, N, p + f.W, a);
      g[0] += a[kERxy];
      g[1] += a[kERz];
    }
  } else {
    if (I + 1 >= Hc) return;
    const size_t p = (size_t)(2 * I + 1) * f.W + 2 * J;
    coefs(f.coef, f.storage, N, p, a);
    g[0] += a[kEDxy];
    g[1] += a[kEDz];
    if (2 * J + 1 < f.W) {
      coefs(f.coef, f.storage, N, p + 1, a);
      g[0] += a[kEDxy];
      g[1] += a[kEDz];
    }
  }
}

// The residual r - sum_n e_n x_n - B x of a pixel of the finer level for its
// three unknowns, each in f32 in the order left, up, right, down, then the
// block's row (missing neighbours skipped).
__device__ __forceinline__ void residual3(const LevelView::Finer& f, const float4* x, int i, int j, float (&res)[3]) {
  const int W = f.W, H = f.H, N = W * H, l = i * W + j;
  const bool hl = j > 0, hu = i > 0, hr = j < W - 1, hd = i < H - 1;
  const PixCoef c = load_coef(f.coef, f.storage, W, N, l, hl, hu);
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  const float4 o = gload(x + l), L = hl ? gload(x + l - 1) : zero, U = hu ? gload(x + l - W) : zero,
               R = hr ? gload(x + l + 1) : zero, D = hd ? gload(x + l + W) : zero;
  const float xo[3] = {o.x, o.y, o.z}, xl[3] = {L.x, L.y, L.z}, xu[3] = {U.x, U.y, U.z}, xr[3] = {R.x, R.y, R.z},
              xd[3] = {D.x, D.y, D.z};
  const float B[3][3] = {{c.Buu, c.Buv, c.Buw}, {c.Buv, c.Bvv, c.Bvw}, {c.Buw, c.Bvw, c.Bww}};
  const float rhs[3] = {c.ru, c.rv, c.rw};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    float a = rhs[k];
    if (hl) a -= (k < 2 ? c.lxy : c.lz) * xl[k];
    if (hu) a -= (k < 2 ? c.uxy : c.uz) * xu[k];
    if (hr) a -= (k < 2 ? c.rxy : c.rz) * xr[k];
    if (hd) a -= (k < 2 ? c.dxy : c.dz) * xd[k];
    a -= B[k][0] * xo[0];
    a -= B[k][1] * xo[1];
    a -= B[k][2] * xo[2];
    res[k] = a;
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
  const size_t N = (size_t)f.W * f.H;
  const int i0 = 2 * I, j0 = 2 * J;
  const bool r1 = i0 + 1 < f.H, c1 = j0 + 1 < f.W;
  const size_t p00 = (size_t)i0 * f.W + j0;
  // children (2I,2J), (2I,2J+1), (2I+1,2J), (2I+1,2J+1)
  float a00[kCoef], a01[kCoef] = {}, a10[kCoef] = {}, a11[kCoef] = {};
  coefs(f.coef, f.storage, N, p00, a00);
  if (c1) coefs(f.coef, f.storage, N, p00 + 1, a01);
  if (r1) coefs(f.coef, f.storage, N, p00 + f.W, a10);
  if (r1 && c1) coefs(f.coef, f.storage, N, p00 + f.W + 1, a11);
  float gr[2], gd[2], gl[2] = {0.f, 0.f}, gu[2] = {0.f, 0.f};
  cross(f, Wc, Hc, I, J, false, gr);
  cross(f, Wc, Hc, I, J, true, gd);
  if (J > 0) cross(f, Wc, Hc, I, J - 1, false, gl);
  if (I > 0) cross(f, Wc, Hc, I - 1, J, true, gu);
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
    s += a00[k];
    if (c1) s += a01[k];
    if (r1) s += a10[k];
    if (r1 && c1) s += a11[k];
    if (b < 3) {
      // Buu, Bvv: the xy edges; Bww: the z edges.  Inner edges: right of
      // (2I,2J), right of (2I+1,2J), down of (2I,2J), down of (2I,2J+1)
      const int er = b < 2 ? kERxy : kERz, ed = b < 2 ? kEDxy : kEDz;
      if (c1) s += 2.f * a00[er];
      if (c1 && r1) s += 2.f * a10[er];
      if (r1) s += 2.f * a00[ed];
      if (r1 && c1) s += 2.f * a01[ed];
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
  float s[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const int i = 2 * I + (d >> 1), j = 2 * J + (d & 1);
    if (i < f.H && j < f.W) {
      float r[3];
      residual3(f, x, i, j, r);
      s[0] += r[0];
      s[1] += r[1];
      s[2] += r[2];
    }
  }
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
