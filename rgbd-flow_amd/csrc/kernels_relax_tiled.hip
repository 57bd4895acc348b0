// kernels_relax_tiled.hip -- red-black relaxation with several iterations per
// launch (temporal blocking in LDS) on gfx950.
//
// The same iteration as k_relax_plain (every pixel update is relax_update on
// the same operands), scheduled differently: each workgroup owns an interior
// tile of the level and loads the tile plus a halo of 2*T pixels into LDS,
// then runs T red-black iterations (2*T colour passes, one barrier each).
// After colour pass s the values are exact at distance >= s from the loaded
// region's cut edges, so the interior (distance >= 2*T) is exact after T
// iterations and is written back; the halo's results are discarded.  Image
// borders need no halo, so a tile that covers the whole level runs a whole
// check interval (25 iterations) per launch.
//
// Layout: the region is stored as horizontal pixel pairs (global columns
// 2k, 2k+1); each pair holds one pixel of each colour, so a colour pass
// touches exactly one pixel per pair and reads the other colour's planes at
// pair offsets {k-1, k, k+1} in the same row and k in the rows above/below:
// consecutive lanes read consecutive LDS words.  Each thread keeps the
// coefficients of the pixels it owns in registers for all T iterations.
#include "internal.hpp"
#include "relax_pixel.hpp"

#include <algorithm>

namespace rgbdflow {

namespace {

// One colour pass over the pairs a thread owns; colour C is a compile-time
// constant so every per-pixel array is indexed statically (kept in VGPRs).
template <int P, int C>
__device__ __forceinline__ void tiled_pass(float* X, int cap, int PR, int RH, const PixCoef (&cf)[P][2],
                                           const int (&qy)[P], const int (&qk)[P], const int (&par)[P],
                                           const unsigned (&fl)[P][2], bool last, float& change) {
  constexpr int O = 1 - C;  // the colour read
#pragma unroll
  for (int s = 0; s < P; s++) {
    const unsigned f = fl[s][C];
    if (!(f & 16u)) continue;  // no pixel of this colour
    const int y = qy[s], kk = qk[s], e = par[s];
    const int q = y * PR + kk;
    // same-row neighbours of the other colour: red reads pairs (k-1+e, k+e),
    // black reads (k-e, k+1-e); the rows above/below read pair k
    const int kl = C == 0 ? kk - 1 + e : kk - e;
    const int kr = C == 0 ? kk + e : kk + 1 - e;
    const int ql = y * PR + min(max(kl, 0), PR - 1);
    const int qr = y * PR + min(max(kr, 0), PR - 1);
    const int qu = max(y - 1, 0) * PR + kk;
    const int qd = min(y + 1, RH - 1) * PR + kk;
    float xo[3], xl[3], xu[3], xr[3], xd[3], xn[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float* Po = X + (O * 3 + k) * cap;
      xo[k] = X[(C * 3 + k) * cap + q];
      xl[k] = Po[ql];
      xu[k] = Po[qu];
      xr[k] = Po[qr];
      xd[k] = Po[qd];
    }
    relax_update(cf[s][C], f & 1u, f & 2u, f & 4u, f & 8u, xl, xu, xr, xd, xo, xn);
#pragma unroll
    for (int k = 0; k < 3; k++) X[(C * 3 + k) * cap + q] = xn[k];
    if (last && (f & 32u))
      change = fmaxf(change, fmaxf(fmaxf(fabsf(xn[0] - xo[0]), fabsf(xn[1] - xo[1])), fabsf(xn[2] - xo[2])));
  }
}

// Reads du from one buffer and writes the tile interiors to the other (flip
// selects which): a workgroup's halo overlaps other workgroups' interiors,
// which must still hold the values from before this launch.  A launch after an
// early exit copies its interior, so the host knows which buffer holds du
// after every launch.  Block b works on tile b - start[item] of its item.
template <int NT, int P>
__global__ void __launch_bounds__(NT) k_relax_tiled(Views wl, int flip, int n_iter, int slot_base,
                                                    int n_prev_checks, int check_slot) {
  extern __shared__ float X[];
  const int item = work_item(wl, blockIdx.x);
  const LevelView& v = wl.items[item];
  if (v.ctl->outer_done) return;
  const TileSched ts = v.ts;
  const float* __restrict__ src = flip ? v.du_alt : v.du;
  float* __restrict__ dst = flip ? v.du : v.du_alt;
  const int W = v.W, H = v.H, N = v.N;
  const int tb = blockIdx.x - wl.start[item];
  const int bx = tb % ts.tiles_x, by = tb / ts.tiles_x;
  const int ox = bx * ts.TW, oy = by * ts.TH;
  const int ex = min(ox + ts.TW, W), ey = min(oy + ts.TH, H);
  if (relax_stopped(v.ctl, slot_base, n_prev_checks)) {
    for (int t = threadIdx.x; t < (ex - ox) * (ey - oy); t += NT) {
      const int l = (oy + t / (ex - ox)) * W + ox + t % (ex - ox);
#pragma unroll
      for (int k = 0; k < 3; k++) dst[(size_t)k * N + l] = src[(size_t)k * N + l];
    }
    return;
  }
  const int h = ts.single ? 0 : 2 * n_iter;
  const int rx0 = max(0, ox - h) & ~1, ry0 = max(0, oy - h);
  const int rx1 = min(W, ex + h), ry1 = min(H, ey + h);
  const int PR = (rx1 - rx0 + 1) >> 1, RH = ry1 - ry0;
  const int np = PR * RH;
  const int cap = ts.cap;  // plane (colour c, component k) at X + (c * 3 + k) * cap

  PixCoef cf[P][2];
  int qy[P], qk[P], par[P];
  unsigned fl[P][2];  // bits: 1 left, 2 up, 4 right, 8 down exist; 16 pixel valid; 32 interior
#pragma unroll
  for (int s = 0; s < P; s++) {
    const int q = threadIdx.x + s * NT;
    const int y = q / max(PR, 1), kk = q - y * PR;
    qy[s] = y;
    qk[s] = kk;
    const int gy = ry0 + y;
    const int e = (gy + rx0) & 1;  // 0: the pair's first pixel is red
    par[s] = e;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int gc = rx0 + 2 * kk + (c == 0 ? e : 1 - e);
      const bool valid = q < np && gc < W;
      const bool hl = gc > 0, hu = gy > 0;
      unsigned f = (hl ? 1u : 0u) | (hu ? 2u : 0u) | (gc < W - 1 ? 4u : 0u) | (gy < H - 1 ? 8u : 0u);
      if (valid) f |= 16u;
      if (valid && gc >= ox && gc < ex && gy >= oy && gy < ey) f |= 32u;
      fl[s][c] = f;
      const int l = gy * W + gc;
      if (valid) {
        cf[s][c] = load_coef(v.coef, N, W, l, hl, hu);
#pragma unroll
        for (int k = 0; k < 3; k++) X[(c * 3 + k) * cap + q] = src[(size_t)k * N + l];
      } else {
        cf[s][c] = PixCoef{};
      }
    }
  }

  float change = 0.f;
  for (int it = 0; it < n_iter; it++) {
    const bool last = it == n_iter - 1;
    __syncthreads();
    tiled_pass<P, 0>(X, cap, PR, RH, cf, qy, qk, par, fl, last, change);
    __syncthreads();
    tiled_pass<P, 1>(X, cap, PR, RH, cf, qy, qk, par, fl, last, change);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < P; s++) {
    const int q = qy[s] * PR + qk[s];
    const int gy = ry0 + qy[s];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (!(fl[s][c] & 32u)) continue;
      const int gc = rx0 + 2 * qk[s] + (c == 0 ? par[s] : 1 - par[s]);
      const int l = gy * W + gc;
#pragma unroll
      for (int k = 0; k < 3; k++) dst[(size_t)k * N + l] = X[(c * 3 + k) * cap + q];
    }
  }
  if (check_slot >= 0) {
    const float m = wave_max(change);
    if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(&v.ctl->maxdiff[check_slot], __float_as_uint(m));
  }
  if (tb == 0 && threadIdx.x == 0) v.ctl->sweeps_exec += n_iter;
}

// Copies du_alt into du for every item (after an odd number of tiled launches).
__global__ void __launch_bounds__(kBlock) k_copy_du(Views wl) {
  const int item = work_item(wl, blockIdx.x);
  const LevelView& v = wl.items[item];
  if (v.ctl->outer_done) return;
  const int l = (blockIdx.x - wl.start[item]) * kBlock + threadIdx.x;
  if (l >= v.N) return;
  for (int k = 0; k < 3; k++) v.du[(size_t)k * v.N + l] = v.du_alt[(size_t)k * v.N + l];
}

struct TiledConfig {
  int nt, p;
  void (*fn)(Views, int, int, int, int, int);
};

#define RGBDFLOW_TILED(NT, P) TiledConfig{NT, P, k_relax_tiled<NT, P>}
const TiledConfig kConfigs[] = {RGBDFLOW_TILED(256, 1), RGBDFLOW_TILED(512, 1), RGBDFLOW_TILED(1024, 1),
                                RGBDFLOW_TILED(256, 2), RGBDFLOW_TILED(512, 2)};
#undef RGBDFLOW_TILED

}  // namespace

int tiled_config_count() { return (int)(sizeof(kConfigs) / sizeof(kConfigs[0])); }
void tiled_config_shape(int idx, int* nt, int* p) {
  *nt = kConfigs[idx].nt;
  *p = kConfigs[idx].p;
}

size_t tiled_config_lds(int idx) { return (size_t)6 * kConfigs[idx].nt * kConfigs[idx].p * sizeof(float); }

__global__ void k_copy_f32(const float* src, size_t n, float* dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// A well-conditioned synthetic system for timing schedules: unit-scale
// diagonal, weak couplings, small rhs; du = 0.
__global__ void k_fill_test_system(float* coef, float* du, int W, int H) {
  const int N = W * H;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < N; l += gridDim.x * blockDim.x) {
    const int j = l % W, i = l / W;
    coef[kERxy * (size_t)N + l] = j < W - 1 ? 0.1f : 0.f;
    coef[kERz * (size_t)N + l] = j < W - 1 ? 0.1f : 0.f;
    coef[kEDxy * (size_t)N + l] = i < H - 1 ? 0.1f : 0.f;
    coef[kEDz * (size_t)N + l] = i < H - 1 ? 0.1f : 0.f;
    coef[kBuu * (size_t)N + l] = -1.f;
    coef[kBvv * (size_t)N + l] = -1.f;
    coef[kBww * (size_t)N + l] = -1.f;
    coef[kBuv * (size_t)N + l] = 0.01f;
    coef[kBuw * (size_t)N + l] = 0.01f;
    coef[kBvw * (size_t)N + l] = 0.01f;
    coef[kRu * (size_t)N + l] = 1e-3f * (l % 7);
    coef[kRv * (size_t)N + l] = 1e-3f * (l % 5);
    coef[kRw * (size_t)N + l] = 1e-3f * (l % 3);
    du[l] = du[N + l] = du[2 * N + l] = 0.f;
  }
}

void launch_fill_test_system(hipStream_t s, float* coef, float* du, int W, int H) {
  int nb = (int)std::min<size_t>(((size_t)W * H + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_test_system, dim3(nb), dim3(256), 0, s, coef, du, W, H);
}

void launch_copy_du(hipStream_t s, const Views& v, int blocks) {
  hipLaunchKernelGGL(k_copy_du, dim3(blocks), dim3(kBlock), 0, s, v);
}

void launch_relax_tiled(hipStream_t s, const Views& v, int blocks, int config, int cap, int flip, int n_iter,
                        int slot_base, int n_prev_checks, int check_slot) {
  const TiledConfig& c = kConfigs[config];
  const size_t lds = (size_t)6 * cap * sizeof(float);
  static bool raised[16] = {};
  if (lds > 65536 && !raised[config]) {
    hipFuncSetAttribute(reinterpret_cast<const void*>(c.fn), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    raised[config] = true;
  }
  hipLaunchKernelGGL(c.fn, dim3(blocks), dim3(c.nt), lds, s, v, flip, n_iter, slot_base, n_prev_checks, check_slot);
}

}  // namespace rgbdflow
