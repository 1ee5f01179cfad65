// awfl_pert.hip -- FP32 perturbation-form split sweeps with Kahan-compensated
// updates, and their C ABI (awfl_*_f32 in include/awfl_dycore.h).
//
// Same kernel organisation as the FP64 sweeps (awfl_dycore.hip): a workgroup
// per x-line staged in LDS for the x sweep, one lane per x-column walking a
// segment for y and z.  Storage is two FP32 arrays per state in the FP64
// layout: the value s and its Kahan compensation c; the represented value is
// s - c.  Each sweep reads (s, c) and writes (s', c') with
//   y = dt*tend - c;  t = s + y;  c' = (t - s) - y;  s' = t.
// FP64 is used only on the host: background profiles, the pressure of each
// background rho*theta (the equation-of-state reference) and, through
// awfl_state_sums on the unpacked state, conservation sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "awfl_cell_pert.hpp"
#include "awfl_internal.hpp"

using namespace awfl;
using namespace awfl::pert;

namespace {

struct GridF {
  int nx, ny, nz, hs;
  int64_t sy, sz, sl;
  int bc[3];
  int weno_s, weno_w, sim2d;
  const PointBg* bgL;      // [nz]       x/y: level cell averages
  const PointBg* bgG;      // [nz*ngll]  z: GLL points
  const FaceBgP* fbL;      // [nz]       x/y faces
  const FaceBgP* fbF;      // [nz+1]     z faces
};

__device__ __forceinline__ void kahan(float s, float c, float inc, float& so, float& co) {
  float y = inc - c;
  float t = s + y;
  co = (t - s) - y;
  so = t;
}

template <int ORD, int NGLL>
__global__ __launch_bounds__(256) void k_sweep_x_f32(const float* __restrict__ in, const float* __restrict__ cin,
                                                     float* __restrict__ out, float* __restrict__ cout, GridF G,
                                                     SweepKF K) {
  constexpr int hs = (ORD - 1) / 2;
  extern __shared__ float smf[];
  const int nx = G.nx;
  const int W = nx + 2 * hs;
  const unsigned line = tile_index();
  const int j = line % G.ny;
  const int k = line / G.ny;
  float* s_st = smf;
  float* s_L = smf + 5 * W;
  float* s_R = s_L + 5 * nx;
  float* s_F = smf;                    // reuses the staging region
  const int64_t row = (int64_t)(hs + k) * G.sz + (int64_t)(hs + j) * G.sy + hs;
  const bool per = G.bc[0] == AWFL_BC_PERIODIC;
  for (int c = threadIdx.x; c < W; c += blockDim.x) {
    int ci = c - hs, src;
    bool zn = false;
    if (per) {
      src = ci % nx;
      if (src < 0) src += nx;
    } else if (ci < 0) {
      src = 0; zn = true;
    } else if (ci >= nx) {
      src = nx - 1; zn = true;
    } else {
      src = ci;
    }
#pragma unroll
    for (int l = 0; l < 5; l++) {
      // the stencils see s; the compensation c (below 1/2 ulp of s) enters
      // only the compensated update
      float v = in[l * G.sl + row + src];
      s_st[l * W + c] = (zn && l == idU) ? 0.0f : v;
    }
  }
  __syncthreads();
  CellBgP<NGLL> bg;
  const PointBg b = G.bgL[k];
  sfor<NGLL>([&](auto ii) { bg.pt[ii] = b; });
  bg.wall_bot = bg.wall_top = false;
  for (int c = threadIdx.x; c < nx; c += blockDim.x) {
    float st[5][ORD];
    sfor<5>([&](auto l) { sfor<ORD>([&](auto s) { st[l][s] = s_st[l * W + c + s]; }); });
    float eL[5], eR[5], sa;
    cell_edges_pert<ORD, NGLL, 0>(st, bg, G.weno_s, G.weno_w, K, eL, eR, sa);
#pragma unroll
    for (int l = 0; l < 5; l++) {
      s_L[l * nx + c] = eL[l];
      s_R[l * nx + c] = eR[l];
    }
  }
  __syncthreads();
  const FaceBgP fb = G.fbL[k];
  const int nfaces = per ? nx : nx + 1;
  for (int f = threadIdx.x; f < nfaces; f += blockDim.x) {
    int a, bb;
    bool aL = false, bR = false;
    if (per) {
      a = f == 0 ? nx - 1 : f - 1;
      bb = f;
    } else if (f == 0) {
      a = 0; bb = 0; aL = true;
    } else if (f == nx) {
      a = nx - 1; bb = nx - 1; bR = true;
    } else {
      a = f - 1; bb = f;
    }
    float qL[5], qR[5], F[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      qL[l] = aL ? s_L[l * nx + a] : s_R[l * nx + a];
      qR[l] = bR ? s_R[l * nx + bb] : s_L[l * nx + bb];
    }
    riemann_pert<0>(qL, qR, fb, F);
#pragma unroll
    for (int l = 0; l < 5; l++) s_F[l * (nx + 1) + f] = F[l];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nx; c += blockDim.x) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int cp = (per && c + 1 == nx) ? 0 : c + 1;
      float td = -(s_F[l * (nx + 1) + cp] - s_F[l * (nx + 1) + c]) * K.rd;
      if (G.sim2d && l == idV) td = 0.0f;
      const int64_t o = l * G.sl + row + c;
      kahan(in[o], cin[o], K.dt * td, out[o], cout[o]);   // s again: an L2 hit, and LDS stays small
    }
  }
}

template <int ORD, int NGLL, int DIR>
__global__ __launch_bounds__(64) void k_sweep_col_f32(const float* __restrict__ in, const float* __restrict__ cin,
                                                      float* __restrict__ out, float* __restrict__ cout, GridF G,
                                                      SweepKF K, int seglen) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  // a y walk takes the segments of one level before the next level (measured
  // faster for the FP32 walk; the FP64 walk showed no difference)
  const dim3 tb = tile_coords<DIR == 1>();
  const int i = tb.x * blockDim.x + threadIdx.x;
  if (i >= G.nx) return;
  const int o = tb.y;
  const int n = DIR == 1 ? G.ny : G.nz;
  const int64_t sa = DIR == 1 ? G.sy : G.sz;
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int64_t base = DIR == 1 ? (int64_t)(hs + o) * G.sz + (int64_t)hs * G.sy + hs + i
                                : (int64_t)hs * G.sz + (int64_t)(hs + o) * G.sy + hs + i;
  const int s0 = tb.z * seglen;
  const int s1 = min(n, s0 + seglen);
  auto ld = [&](int l, int a) -> float {
    int src;
    bool zn = false;
    if (per) {
      if (n > hs) {                    // the walk reaches at most hs + 1 cells past either end
        src = a < 0 ? a + n : (a >= n ? a - n : a);
      } else {
        src = a % n;
        if (src < 0) src += n;
      }
    } else if (a < 0) {
      src = 0; zn = true;
    } else if (a >= n) {
      src = n - 1; zn = true;
    } else {
      src = a;
    }
    const int64_t idx = l * G.sl + base + (int64_t)src * sa;
    float v = in[idx];                 // s; c enters only the update
    return (zn && l == NM) ? 0.0f : v;
  };
  // the y sweep's background is its level's, loaded once
  const PointBg bgl = DIR == 1 ? G.bgL[o] : PointBg{};
  const FaceBgP fbl = DIR == 1 ? G.fbL[o] : FaceBgP{};
  auto bg_of = [&](int kc, CellBgP<NGLL>& bg) {
    if constexpr (DIR == 1) {
      sfor<NGLL>([&](auto ii) { bg.pt[ii] = bgl; });
      bg.wall_bot = bg.wall_top = false;
    } else {
      sfor<NGLL>([&](auto ii) { bg.pt[ii] = G.bgG[kc * NGLL + ii]; });
      bg.wall_bot = !per && kc == 0;
      bg.wall_top = !per && kc == n - 1;
    }
  };
  auto face_bg = [&](int f) -> FaceBgP { return DIR == 1 ? fbl : G.fbF[f]; };
  // compensated update of one cell into registers, and its store
  auto update = [&](const float (&F0)[5], const float (&F1)[5], float src, const float (&sc)[5],
                    const float (&cc)[5], float (&so)[5], float (&co)[5]) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      float td = -(F1[l] - F0[l]) * K.rd;
      if (DIR == 2 && l == idW) td = src + td;
      if (DIR == 2 && G.sim2d && l == idV) td = 0.0f;
      kahan(sc[l], cc[l], K.dt * td, so[l], co[l]);
    }
  };
  auto put = [&](int c, const float (&so)[5], const float (&co)[5]) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int64_t idx = l * G.sl + base + (int64_t)c * sa;
      out[idx] = so[l];
      cout[idx] = co[l];
    }
  };

  float w[5][ORD];
  sfor<5>([&](auto l) { sfor<ORD>([&](auto s) { w[l][s] = ld(l, s0 - 1 - hs + s); }); });
  float prevR[5], prevF[5], prevSrc = 0.0f;
  if (per || s0 > 0) {
    CellBgP<NGLL> bg;
    bg_of(((s0 - 1) % n + n) % n, bg);
    float eL[5], sa_;
    cell_edges_pert<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, prevR, sa_);
  }
  // software pipeline: the stencil cell of the next iteration and the
  // compensation of the next update are loaded one iteration ahead, and the
  // updated cell of an iteration is stored at the top of the next one, ahead
  // of that iteration's loads: a wait for the loads then never waits for
  // stores issued just before it
  float nxt[5], cnx[5], cprev[5], ps[5], pc[5];
  sfor<5>([&](auto l) {
    nxt[l] = ld(l, s0 + hs);
    cnx[l] = cin[l * G.sl + base + (int64_t)(((s0 - 1) % n + n) % n) * sa];
    prevF[l] = 0.0f;
    ps[l] = pc[l] = 0.0f;
  });
  for (int v = s0; v < s1; v++) {
    if (v > s0 + 1) put(v - 2, ps, pc);
    // the window is centred on cell v-1, the cell updated in this iteration
    float prevS[5];
    sfor<5>([&](auto l) {
      prevS[l] = w[l][hs];
      cprev[l] = cnx[l];
      sfor<ORD - 1>([&](auto s) { w[l][s] = w[l][s + 1]; });
      w[l][ORD - 1] = nxt[l];
      nxt[l] = ld(l, v + 1 + hs);
      cnx[l] = cin[l * G.sl + base + (int64_t)v * sa];
    });
    CellBgP<NGLL> bg;
    bg_of(v, bg);
    float eL[5], eR[5], src;
    cell_edges_pert<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, eR, src);
    float F[5];
    const FaceBgP fb = face_bg(v);
    if (!per && v == 0) riemann_pert<DIR>(eL, eL, fb, F);
    else riemann_pert<DIR>(prevR, eL, fb, F);
    // cell v-1 (for v == s0 a value that is never stored)
    update(prevF, F, prevSrc, prevS, cprev, ps, pc);
#pragma unroll
    for (int l = 0; l < 5; l++) {
      prevF[l] = F[l];
      prevR[l] = eR[l];
    }
    prevSrc = src;
  }
  if (s1 > s0 + 1) put(s1 - 2, ps, pc);
  float F[5], lastS[5];
  sfor<5>([&](auto l) { lastS[l] = w[l][hs]; });   // cell s1-1
  const FaceBgP fb = face_bg(s1);
  if (!per && s1 == n) {
    riemann_pert<DIR>(prevR, prevR, fb, F);
  } else {
    sfor<5>([&](auto l) {
      sfor<ORD - 1>([&](auto s) { w[l][s] = w[l][s + 1]; });
      w[l][ORD - 1] = nxt[l];          // cell s1 + hs, loaded by the last iteration
    });
    CellBgP<NGLL> bg;
    bg_of(s1 % n, bg);
    float eL[5], eR[5], sa_;
    cell_edges_pert<ORD, NGLL, DIR>(w, bg, G.weno_s, G.weno_w, K, eL, eR, sa_);
    riemann_pert<DIR>(prevR, eL, fb, F);
  }
  update(prevF, F, prevSrc, lastS, cnx, ps, pc);
  put(s1 - 1, ps, pc);
}

__global__ void k_to_f32(const double* __restrict__ d, float* __restrict__ s, float* __restrict__ c, int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    float v = (float)d[t];
    s[t] = v;
    c[t] = (float)((double)v - d[t]);   // represented value s - c
  }
}

__global__ void k_from_f32(const float* __restrict__ s, const float* __restrict__ c, double* __restrict__ d,
                           int64_t n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    d[t] = (double)s[t] - (double)c[t];
}

struct VariantF {
  int ord, ngll;
  void (*sweep_x)(const float*, const float*, float*, float*, GridF, SweepKF);
  void (*sweep_y)(const float*, const float*, float*, float*, GridF, SweepKF, int);
  void (*sweep_z)(const float*, const float*, float*, float*, GridF, SweepKF, int);
};

#define AWFL_VARIANT_F(O, N) \
  VariantF { O, N, k_sweep_x_f32<O, N>, k_sweep_col_f32<O, N, 1>, k_sweep_col_f32<O, N, 2> }

const VariantF kVariantsF[] = {AWFL_VARIANT_F(3, 2), AWFL_VARIANT_F(5, 3)};

const VariantF* find_variant_f(int ord, int ngll) {
  for (const VariantF& v : kVariantsF)
    if (v.ord == ord && v.ngll == ngll) return &v;
  return nullptr;
}

constexpr double GAMMA_D = 1.40027894002789400278940027894;
constexpr double C0_D = 27.5629410929725921310572974482;

// FP32 background tables from the FP64 profiles (host), and buffers
int ensure_f32(awfl_ctx* c) {
  if (!find_variant_f(c->cfg.ord, c->cfg.ngll))
    return fail(c, AWFL_E_UNSUPPORTED, "FP32 perturbation form: (ord, ngll) not compiled in");
  if (c->d_bgf) return AWFL_OK;
  int nz = c->cfg.nz, ngll = c->cfg.ngll, hs = c->hs;
  std::vector<PointBg> bgL(nz), bgG((size_t)nz * ngll);
  std::vector<FaceBgP> fbL(nz), fbF(nz + 1);
  auto point = [](double r, double rt, double p) {
    PointBg b;
    b.r = (float)r; b.rt = (float)rt; b.p = (float)p; b.rrt = (float)(1.0 / rt);
    return b;
  };
  for (int k = 0; k < nz; k++) {
    double r = c->hy[0][hs + k], rt = c->hy[3][hs + k];
    double p = C0_D * std::pow(rt, GAMMA_D);     // pressure of the background rho*theta itself
    bgL[k] = point(r, rt, p);
    FaceBgP f;
    f.r = (float)r; f.rt = (float)rt; f.p = (float)p; f.rrt = (float)(1.0 / rt);
    double bq = r * rt;
    f.bq = (float)bq; f.rq = (float)(1.0 / bq); f.pq = (float)(C0_D * std::pow(bq, GAMMA_D));
    fbL[k] = f;
    for (int kk = 0; kk < ngll; kk++) {
      size_t q = (size_t)k * ngll + kk;
      bgG[q] = point(c->hy[4][q], c->hy[7][q], c->hy[5][q]);
    }
  }
  for (int f = 0; f <= nz; f++) {
    size_t q = f < nz ? (size_t)f * ngll : (size_t)(nz - 1) * ngll + ngll - 1;
    FaceBgP fb;
    fb.r = (float)c->hy[4][q]; fb.rt = (float)c->hy[7][q]; fb.p = (float)c->hy[5][q];
    fb.rrt = (float)(1.0 / c->hy[7][q]);
    fb.bq = fb.rq = fb.pq = 0.0f;
    fbF[f] = fb;
  }
  size_t bytes = sizeof(PointBg) * (bgL.size() + bgG.size()) + sizeof(FaceBgP) * (fbL.size() + fbF.size());
  char* d = nullptr;
  HIP_TRY(c, hipMalloc(&d, bytes));
  size_t off = 0;
  auto put = [&](const void* src, size_t n) -> hipError_t {
    hipError_t e = hipMemcpy(d + off, src, n, hipMemcpyHostToDevice);
    off += n;
    return e;
  };
  HIP_TRY(c, put(bgL.data(), sizeof(PointBg) * bgL.size()));
  HIP_TRY(c, put(bgG.data(), sizeof(PointBg) * bgG.size()));
  HIP_TRY(c, put(fbL.data(), sizeof(FaceBgP) * fbL.size()));
  HIP_TRY(c, put(fbF.data(), sizeof(FaceBgP) * fbF.size()));
  c->d_bgf = d;
  return AWFL_OK;
}

GridF grid_f(const awfl_ctx* c) {
  GridF F{};
  Grid G = c->grid();
  F.nx = G.nx; F.ny = G.ny; F.nz = G.nz; F.hs = G.hs;
  F.sy = G.sy; F.sz = G.sz; F.sl = G.sl;
  F.bc[0] = G.bc[0]; F.bc[1] = G.bc[1]; F.bc[2] = G.bc[2];
  F.weno_s = G.weno_s; F.weno_w = G.weno_w; F.sim2d = G.sim2d;
  int nz = c->cfg.nz, ngll = c->cfg.ngll;
  const char* d = static_cast<const char*>(c->d_bgf);
  F.bgL = reinterpret_cast<const PointBg*>(d);
  F.bgG = F.bgL + nz;
  F.fbL = reinterpret_cast<const FaceBgP*>(F.bgG + (size_t)nz * ngll);
  F.fbF = F.fbL + nz;
  return F;
}

SweepKF sweep_constants_f(const awfl_ctx* c, double dt, int dir) {
  SweepKF K{};
  double dsw = dir == 0 ? c->dx : dir == 1 ? c->dy : c->dz;
  K.dt = (float)dt;
  K.rd = (float)(1.0 / dsw);
  double dtmult = dt;
  for (int kt = 0; kt < KMAX; kt++) {
    K.ndt[kt] = (float)(-1.0 / (dsw * (kt + 1.0)));
    if (kt >= 1) {
      K.tav[kt] = (float)(dtmult / (kt + 1.0));
      dtmult *= dt;
    }
  }
  const int ng = c->cfg.ngll;
  for (int kt = 0; kt + 1 < ng && kt < PK_NA - 1; kt++)
    for (int s = 0; s < ng; s++)
      for (int ii = 0; ii < ng; ii++) {
        double D = ng == 2 ? awfl_deriv_2[s][ii] : awfl_deriv_3[s][ii];
        K.dn[kt][s][ii] = (float)(D * (-1.0 / (dsw * (kt + 1.0))));
      }
  return K;
}

int launch_sweep_f32(awfl_ctx* c, const float* s, const float* cs, float* so, float* co, double dt, int dir) {
  GridF F = grid_f(c);
  SweepKF K = sweep_constants_f(c, dt, dir);
  const VariantF* v = find_variant_f(c->cfg.ord, c->cfg.ngll);
  if (dir == 0) {
    int nx = c->cfg.nx;
    int threads = line_threads(nx);
    size_t lds = sizeof(float) * (size_t)(5 * (nx + 2 * c->hs) + 10 * nx);
    if (lds > 64 * 1024)
      HIP_TRY(c, hipFuncSetAttribute((const void*)v->sweep_x, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    v->sweep_x<<<c->cfg.ny * c->cfg.nz, threads, lds, c->stream>>>(s, cs, so, co, F, K);
    LAUNCH_CHECK(c);
    return AWFL_OK;
  }
  int n = dir == 1 ? c->cfg.ny : c->cfg.nz;
  int other = dir == 1 ? c->cfg.nz : c->cfg.ny;
  int bx = (c->cfg.nx + 63) / 64;
  // segments as in the FP64 sweeps: at most 32 cells, down to 16 for ~4
  // waves per SIMD
  int64_t layer = (int64_t)bx * other;
  int nseg = (int)std::max<int64_t>(1, std::min<int64_t>(std::max<int64_t>((4096 + layer - 1) / layer, (n + 31) / 32),
                                                           std::max(1, n / 16)));
  int seglen = (n + nseg - 1) / nseg;
  nseg = (n + seglen - 1) / seglen;
  dim3 grid(bx, other, nseg);
  auto kern = dir == 1 ? v->sweep_y : v->sweep_z;
  kern<<<grid, 64, 0, c->stream>>>(s, cs, so, co, F, K, seglen);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int grid_stride_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace

bool awfl_detail::has_f32_variant(int ord, int ngll) { return find_variant_f(ord, ngll) != nullptr; }

extern "C" {

int awfl_state_to_f32(awfl_ctx* c, const double* d64, float* s32, float* c32) {
  if (!c) return AWFL_E_ARG;
  if (!d64 || !s32 || !c32) return fail(c, AWFL_E_ARG, "null buffer");
  int64_t n = c->state_elems();
  k_to_f32<<<grid_stride_blocks(n), 256, 0, c->stream>>>(d64, s32, c32, n);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_state_from_f32(awfl_ctx* c, const float* s32, const float* c32, double* d64) {
  if (!c) return AWFL_E_ARG;
  if (!d64 || !s32 || !c32) return fail(c, AWFL_E_ARG, "null buffer");
  int64_t n = c->state_elems();
  k_from_f32<<<grid_stride_blocks(n), 256, 0, c->stream>>>(s32, c32, d64, n);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_sweep_f32(awfl_ctx* c, const float* s_in, const float* c_in, float* s_out, float* c_out, double dt,
                   int dir) {
  if (!c) return AWFL_E_ARG;
  if (!s_in || !c_in || !s_out || !c_out) return fail(c, AWFL_E_ARG, "null buffer");
  if (s_in == s_out || c_in == c_out) return fail(c, AWFL_E_ARG, "sweeps run out of place");
  if (dir < 0 || dir > 2) return fail(c, AWFL_E_ARG, "dir must be 0, 1 or 2");
  int rc = ensure_f32(c);
  if (rc != AWFL_OK) return rc;
  if (dir == 1 && c->sim2d) {
    HIP_TRY(c, hipMemcpyAsync(s_out, s_in, sizeof(float) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c_out, c_in, sizeof(float) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    return AWFL_OK;
  }
  return launch_sweep_f32(c, s_in, c_in, s_out, c_out, dt, dir);
}

int awfl_time_step_f32(awfl_ctx* c, float* s, float* cs, double dt) {
  if (!c) return AWFL_E_ARG;
  if (!s || !cs) return fail(c, AWFL_E_ARG, "null state");
  int rc = ensure_f32(c);
  if (rc != AWFL_OK) return rc;
  for (int b = 0; b < 4; b++)
    if (!c->f_buf[b]) HIP_TRY(c, hipMalloc(&c->f_buf[b], sizeof(float) * c->state_elems()));
  int dirs[3], nd = 0;
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  for (int k = 0; k < 3; k++) {
    int d = c->dimSwitch ? fwd[k] : bwd[k];
    if (d == 1 && c->sim2d) continue;
    dirs[nd++] = d;
  }
  const float *cur = s, *curc = cs;
  int pair = 0;
  for (int k = 0; k < nd; k++) {
    float *dst, *dstc;
    if (k == nd - 1) {
      dst = s;
      dstc = cs;
    } else {
      dst = c->f_buf[2 * pair];
      dstc = c->f_buf[2 * pair + 1];
      pair ^= 1;
    }
    rc = launch_sweep_f32(c, cur, curc, dst, dstc, dt, dirs[k]);
    if (rc != AWFL_OK) return rc;
    cur = dst;
    curc = dstc;
  }
  c->dimSwitch = !c->dimSwitch;
  return AWFL_OK;
}

}  // extern "C"
