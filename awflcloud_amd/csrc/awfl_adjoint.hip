// awfl_adjoint.hip -- tangent-linear and discrete-adjoint FP64 split sweeps,
// the adjoint of a time step, and checkpointed forward/adjoint runs (C ABI
// awfl_sweep_tangent / awfl_sweep_adjoint / awfl_time_step_adjoint /
// awfl_forward_ckpt / awfl_adjoint_ckpt in include/awfl_dycore.h).
//
// One split sweep is out = in + dt * T(in) (awfl_sweep).  Its adjoint
// adj_in = adj_out + J^T (dt adj_out) runs as one fused kernel per line
// along the sweep direction (k_adj_line: forward, face adjoints, transform
// and WENO reverses and the gather, all in LDS).  Lines too long for LDS
// (more than about 580 cells at WENO5) run as five kernels over per-cell
// scratch in HBM:
//   k_edges      cell forward, one thread per cell: WENO values at the GLL
//                points [5 ngll/cell], then time-averaged edge states and
//                z source                                                [11/cell]
//   k_face_adj   face adjoints: flux adjoint dt/d (adj_out[right] -
//                adj_out[left]) pulled back through the Riemann flux to the
//                two edge states the face used                           [10/cell]
//   k_dt_adj     transform reverse, one thread per cell: recompute the
//                transform levels from the GLL values and pull the edge and
//                source adjoints back to them                     [5 ngll/cell]
//   k_weno_adj   WENO reverse, one thread per (cell, variable): GLL-value
//                adjoints back to the stencil                      [5 ord/cell]
//   k_gather     adj_in[c] = adj_out[c] + sum of the stencil adjoints of
//                every cell whose stencil holds c (periodic wrap, or the
//                wall clamp with its zeroed normal momentum)
// The tangent-linear sweep evaluates the same forward with Dual numbers
// (k_edges<Dual>, then k_tlm_update); it exists to test the adjoint by the
// dot-product identity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "awfl_adjoint.hpp"
#include "awfl_internal.hpp"

using namespace awfl;
using awfl::adj::Dual;

namespace {

constexpr int NE = 11;          // edge record: eL[5], eR[5], z source
constexpr int NEB = 10;         // edge adjoints: eLb[5], eRb[5]

struct Pos {
  int i, j, k;
};

__device__ __forceinline__ int64_t sidx(const Grid& G, int l, const Pos& p) {
  return l * G.sl + (int64_t)(G.hs + p.k) * G.sz + (int64_t)(G.hs + p.j) * G.sy + G.hs + p.i;
}
__device__ __forceinline__ int64_t cidx(const Grid& G, const Pos& p) {
  return ((int64_t)p.k * G.ny + p.j) * G.nx + p.i;
}
__device__ __forceinline__ int64_t ncell(const Grid& G) { return (int64_t)G.nx * G.ny * G.nz; }
// kernels over cells launch a grid (ceil(nx/64), ny, nz) of 64-lane blocks:
// the cell comes from the block and lane indices without any division
__device__ __forceinline__ bool cell_of_launch(const Grid& G, Pos& p, int64_t& t) {
  p.i = blockIdx.x * blockDim.x + threadIdx.x;
  p.j = blockIdx.y;
  p.k = blockIdx.z;
  t = ((int64_t)p.k * G.ny + p.j) * G.nx + p.i;
  return p.i < G.nx;
}
template <int DIR>
__device__ __forceinline__ int along(const Pos& p) {
  return DIR == 0 ? p.i : DIR == 1 ? p.j : p.k;
}
template <int DIR>
__device__ __forceinline__ Pos moved(Pos p, int a) {
  if (DIR == 0) p.i = a;
  else if (DIR == 1) p.j = a;
  else p.k = a;
  return p;
}
template <int DIR>
__device__ __forceinline__ int axis_len(const Grid& G) {
  return DIR == 0 ? G.nx : DIR == 1 ? G.ny : G.nz;
}
// logical index a along an axis of n cells -> stored cell (periodic wrap or
// wall clamp); zn marks a clamped slot, whose normal momentum is zero
__device__ __forceinline__ int map_axis(int a, int n, bool per, bool& zn) {
  zn = false;
  if (per) {                    // a is within a few cells of [0, n)
    while (a < 0) a += n;
    while (a >= n) a -= n;
    return a;
  }
  if (a < 0) {
    zn = true;
    return 0;
  }
  if (a >= n) {
    zn = true;
    return n - 1;
  }
  return a;
}

template <int NGLL, int DIR>
__device__ __forceinline__ CellBg<NGLL> cell_bg(const Grid& G, int k) {
  CellBg<NGLL> bg{};
  if constexpr (DIR == 2) {
    sfor<NGLL>([&](auto ii) {
      bg.rG[ii] = G.hyDensGLL[k * NGLL + ii];
      bg.rtG[ii] = G.hyDensThetaGLL[k * NGLL + ii];
      bg.pG[ii] = G.hyPressureGLL[k * NGLL + ii];
      bg.ptG[ii] = G.prG[k * NGLL + ii];
    });
    const bool per = G.bc[2] == AWFL_BC_PERIODIC;
    bg.wall_bot = !per && k == 0;
    bg.wall_top = !per && k == G.nz - 1;
  } else {
    bg.r = G.hyDensCells[G.hs + k];
    bg.rt = G.hyDensThetaCells[G.hs + k];
    bg.pt = G.prT[k];
    bg.wall_bot = bg.wall_top = false;
  }
  return bg;
}

// power references of face f of the line through level k, as the forward
// sweeps use them: x the level's (rho theta), y (rho rho*theta) and the
// level's (rho theta), z the face's GLL value
template <int DIR>
__device__ __forceinline__ FaceRef face_ref(const Grid& G, int k, int f) {
  FaceRef fr;
  if constexpr (DIR == 0) {
    fr.rt = fr.q5 = G.prT[k];
  } else if constexpr (DIR == 1) {
    fr.rt = G.prRT[k];
    fr.q5 = G.prT[k];
  } else {
    fr.rt = fr.q5 = G.prF[f];
  }
  return fr;
}

// background pressure subtracted by the z flux at face f (S:1628-1629)
template <int NGLL>
__device__ __forceinline__ double face_pressure(const Grid& G, int f) {
  return f < G.nz ? G.hyPressureGLL[f * NGLL] : G.hyPressureGLL[(G.nz - 1) * NGLL + NGLL - 1];
}

template <int ORD, int DIR>
__device__ __forceinline__ void load_stencil(const double* __restrict__ st, const Grid& G, const Pos& p,
                                             double (&w)[5][ORD]) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int a0 = along<DIR>(p);
  sfor<ORD>([&](auto s) {
    bool zn;
    Pos q = moved<DIR>(p, map_axis(a0 - hs + s, n, per, zn));
    sfor<5>([&](auto l) { w[l][s] = (zn && l == NM) ? 0.0 : st[sidx(G, l, q)]; });
  });
}

template <int ORD, int DIR>
__device__ __forceinline__ void load_stencil(const double* __restrict__ st, const double* __restrict__ dst,
                                             const Grid& G, const Pos& p, Dual (&w)[5][ORD]) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int a0 = along<DIR>(p);
  sfor<ORD>([&](auto s) {
    bool zn;
    Pos q = moved<DIR>(p, map_axis(a0 - hs + s, n, per, zn));
    sfor<5>([&](auto l) {
      const int64_t o = sidx(G, l, q);
      w[l][s] = (zn && l == NM) ? Dual(0.0) : Dual(st[o], dst[o]);
    });
  });
}

// cell forward -> edge record (value; and tangent for Dual); the double
// variant also keeps the WENO values at the GLL points, Q0[l*NGLL+ii][cell]
template <int ORD, int NGLL, int DIR, class T>
__global__ __launch_bounds__(64) void k_edges(const double* __restrict__ st, const double* __restrict__ dst, Grid G,
                                              SweepK K, double* __restrict__ E, double* __restrict__ Ed,
                                              double* __restrict__ Q0) {
  Pos p;
  int64_t t;
  if (!cell_of_launch(G, p, t)) return;
  const int64_t nc = ncell(G);
  T w[5][ORD];
  if constexpr (std::is_same<T, double>::value) load_stencil<ORD, DIR>(st, G, p, w);
  else load_stencil<ORD, DIR>(st, dst, G, p, w);
  const CellBg<NGLL> bg = cell_bg<NGLL, DIR>(G, p.k);
  adj::Tape<NGLL, T> tp;
  T q0[5][NGLL], eL[5], eR[5], sa;
  adj::cell_recon<ORD, NGLL>(w, G.weno_s, G.weno_w, q0);
  if constexpr (std::is_same<T, double>::value)
    sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { Q0[(l * NGLL + ii) * nc + t] = q0[l][ii]; }); });
  adj::cell_dt<NGLL, DIR>(q0, bg, K, tp, eL, eR, sa);
  sfor<5>([&](auto l) {
    E[l * nc + t] = adj::val(eL[l]);
    E[(5 + l) * nc + t] = adj::val(eR[l]);
  });
  E[10 * nc + t] = adj::val(sa);
  if constexpr (!std::is_same<T, double>::value) {
    sfor<5>([&](auto l) {
      Ed[l * nc + t] = eL[l].d;
      Ed[(5 + l) * nc + t] = eR[l].d;
    });
    Ed[10 * nc + t] = sa.d;
  }
}

// the two edge states of face f (between cells f-1 and f along DIR) of the
// line through p: interior/periodic (eR[f-1], eL[f]); wall faces take both
// sides from the one adjacent cell
template <int DIR>
__device__ __forceinline__ void face_states(const double* __restrict__ E, const Grid& G, const Pos& p, int f,
                                            int64_t& cl, int64_t& cr, int& offl, int& offr) {
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  if (!per && f == 0) {
    cl = cr = cidx(G, moved<DIR>(p, 0));
    offl = offr = 0;                    // eL of cell 0 on both sides
  } else if (!per && f == n) {
    cl = cr = cidx(G, moved<DIR>(p, n - 1));
    offl = offr = 5;                    // eR of cell n-1 on both sides
  } else {
    int a = f - 1;
    if (a < 0) a += n;
    int b = f == n ? 0 : f;
    cl = cidx(G, moved<DIR>(p, a));
    cr = cidx(G, moved<DIR>(p, b));
    offl = 5;
    offr = 0;
  }
}

// Face adjoints.  Thread of cell p handles the face at its low side, and the
// last cell of a wall line also the face at its high side.
template <int NGLL, int DIR>
__global__ __launch_bounds__(64) void k_face_adj(const double* __restrict__ E, const double* __restrict__ adj_out,
                                                 Grid G, SweepK K, double* __restrict__ EB) {
  constexpr int NM = 1 + DIR;
  Pos p;
  int64_t t;
  if (!cell_of_launch(G, p, t)) return;
  const int64_t nc = ncell(G);
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int a = along<DIR>(p);
  const double c = K.dt * K.rd;
  auto do_face = [&](int f) {
    int64_t cl, cr;
    int offl, offr;
    face_states<DIR>(E, G, p, f, cl, cr, offl, offr);
    double L[5], R[5], Fb[5], Lb[5], Rb[5];
    sfor<5>([&](auto l) {
      L[l] = E[(offl + l) * nc + cl];
      R[l] = E[(offr + l) * nc + cr];
      Lb[l] = Rb[l] = 0.0;
    });
    // out[left] -= c F, out[right] += c F (the cell that lies on each side)
    const bool has_l = per || f > 0, has_r = per || f < n;
    Pos pl = moved<DIR>(p, f == 0 ? n - 1 : f - 1), pr = moved<DIR>(p, f == n ? 0 : f);
    sfor<5>([&](auto l) {
      double v = 0.0;
      if (has_r) v += adj_out[sidx(G, l, pr)];
      if (has_l) v -= adj_out[sidx(G, l, pl)];
      Fb[l] = (G.sim2d && l == idV) ? 0.0 : c * v;
    });
    adj::riemann_adj<DIR>(L, R, face_ref<DIR>(G, p.k, f), Fb, Lb, Rb);
    if (!per && f == 0) {
      sfor<5>([&](auto l) { EB[l * nc + cr] = Lb[l] + Rb[l]; });
    } else if (!per && f == n) {
      sfor<5>([&](auto l) { EB[(5 + l) * nc + cl] = Lb[l] + Rb[l]; });
    } else {
      sfor<5>([&](auto l) {
        EB[(5 + l) * nc + cl] = Lb[l];
        EB[l * nc + cr] = Rb[l];
      });
    }
    (void)NM;
  };
  do_face(a);
  if (!per && a == n - 1) do_face(n);
}

// Transform reverse: edge and source adjoints -> GLL-value adjoints
// Q0B[l*NGLL+ii][cell]
template <int NGLL, int DIR>
__global__ __launch_bounds__(64) void k_dt_adj(const double* __restrict__ Q0, const double* __restrict__ EB,
                                               const double* __restrict__ adj_out, Grid G, SweepK K,
                                               double* __restrict__ Q0B) {
  Pos p;
  int64_t t;
  if (!cell_of_launch(G, p, t)) return;
  const int64_t nc = ncell(G);
  double q0[5][NGLL], q0b[5][NGLL], eLb[5], eRb[5];
  sfor<5>([&](auto l) {
    sfor<NGLL>([&](auto ii) { q0[l][ii] = Q0[(l * NGLL + ii) * nc + t]; });
    eLb[l] = EB[l * nc + t];
    eRb[l] = EB[(5 + l) * nc + t];
  });
  const CellBg<NGLL> bg = cell_bg<NGLL, DIR>(G, p.k);
  // z: the source enters rho*w as out += dt * src
  const double srcb = DIR == 2 ? K.dt * adj_out[sidx(G, idW, p)] : 0.0;
  adj::cell_dt_adj<NGLL, DIR>(q0, bg, K, eLb, eRb, srcb, q0b);
  sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { Q0B[(l * NGLL + ii) * nc + t] = q0b[l][ii]; }); });
}

// WENO reverse, one thread per (variable, cell): GLL-value adjoints ->
// stencil adjoints STB[l*ORD+s][cell]
template <int ORD, int NGLL, int DIR>
__global__ __launch_bounds__(64) void k_weno_adj(const double* __restrict__ st, const double* __restrict__ Q0B,
                                                  Grid G, double* __restrict__ STB) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  // grid (ceil(nx/64), 5 ny, nz): the variable is blockIdx.y / ny
  const int64_t nc = ncell(G);
  const int l = blockIdx.y / G.ny;
  Pos p;
  p.i = blockIdx.x * blockDim.x + threadIdx.x;
  p.j = blockIdx.y - l * G.ny;
  p.k = blockIdx.z;
  if (p.i >= G.nx) return;
  const int64_t t = ((int64_t)p.k * G.ny + p.j) * G.nx + p.i;
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int a0 = along<DIR>(p);
  double u[ORD], gb[NGLL], ub[ORD];
  sfor<ORD>([&](auto s) {
    bool zn;
    Pos q = moved<DIR>(p, map_axis(a0 - hs + s, n, per, zn));
    u[s] = (zn && l == NM) ? 0.0 : st[sidx(G, l, q)];
    ub[s] = 0.0;
  });
  sfor<NGLL>([&](auto ii) { gb[ii] = Q0B[(l * NGLL + ii) * nc + t]; });
  adj::reconstruct_adj<ORD, NGLL>(u, (l == idR || l == idT) ? G.weno_s : G.weno_w, gb, ub);
  sfor<ORD>([&](auto s) { STB[(l * ORD + s) * nc + t] = ub[s]; });
}

// ---------------------------------------------------------------------------
// Sweep adjoint fused into one kernel: one work-group per line along DIR
// with the line and every intermediate in LDS (nothing per cell crosses
// HBM).  x lines are contiguous; y and z lines are strided, and the
// XCD-aware line order keeps neighbouring lines (which share every cache
// line they touch) on one XCD at about the same time.  LDS (doubles): the
// state line [5][n+2hs], a region [5 NGLL][n] for the WENO values of each
// cell and then, in place, their adjoints q0b, and a region [10][n] for the
// edge states, replaced in place by the edge-state adjoints (every slot is
// read and written by the one face thread that owns it); the two regions
// then hold the stencil adjoints of all five variables (adj_row); then the
// line's constants and adj_out along the line [5][n], staged once.
// ---------------------------------------------------------------------------
// WENO reverses per thread side by side in the line kernel
template <int ORD>
constexpr int adj_pv() {
  return 2 * ORD <= 10 ? 2 : 1;
}
// row of the stencil adjoint of variable l, slot s, in the line kernel's row
// region [s_q | s_e] of 5 NGLL + 10 rows: variable after variable from the
// first edge row on, wrapping to row 0
template <int ORD, int NGLL>
__host__ __device__ constexpr int adj_row(int l, int s) {
  return (5 * NGLL + l * ORD + s) % (5 * NGLL + 10);
}
// every stencil adjoint has a row of its own, and none lands on the q0b rows
// (l NGLL .. l NGLL + NGLL - 1) of a variable whose WENO reverse runs in a
// later group of PV
template <int ORD, int NGLL, int PV>
constexpr bool adj_rows_ok() {
  if (5 * ORD > 5 * NGLL + 10) return false;
  for (int l = 0; l < 5; ++l) {
    const int next = (l / PV + 1) * PV < 5 ? (l / PV + 1) * PV : 5;
    for (int s = 0; s < ORD; ++s) {
      const int r = adj_row<ORD, NGLL>(l, s);
      if (r >= next * NGLL && r < 5 * NGLL) return false;
    }
  }
  return true;
}
template <int ORD, int NGLL>
constexpr int adj_line_rows() {
  return 5 * NGLL + 10;
}
constexpr int LINE_K = 10;

template <int ORD, int NGLL, int DIR>
__global__ __launch_bounds__(256) void k_adj_line(const double* __restrict__ in, const double* __restrict__ adj_out,
                                                  double* __restrict__ adj_in, Grid G, SweepK K) {
  static_assert(ORD <= 10, "stencil adjoints of one variable fit the edge region");
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  extern __shared__ double sm[];
  const int n = axis_len<DIR>(G);
  const int W = n + 2 * hs;
  const int bd = line_threads(n);     // blockDim.x of the launch
  // line -> its first interior cell, the stride along DIR and the level k of
  // an x or y line
  const unsigned line = tile_index();
  int64_t base, stride;
  int kl = 0;
  if constexpr (DIR == 0) {
    const int j = line % G.ny;
    kl = line / G.ny;
    base = (int64_t)(hs + kl) * G.sz + (int64_t)(hs + j) * G.sy + hs;
    stride = 1;
  } else if constexpr (DIR == 1) {
    const int i = line % G.nx;
    kl = line / G.nx;
    base = (int64_t)(hs + kl) * G.sz + (int64_t)hs * G.sy + hs + i;
    stride = G.sy;
  } else {
    const int i = line % G.nx;
    const int j = line / G.nx;
    base = (int64_t)hs * G.sz + (int64_t)(hs + j) * G.sy + hs + i;
    stride = G.sz;
  }
  double* s_st = sm;                 // [5][W]
  double* s_q = s_st + 5 * W;        // [5 NGLL][n]: q0, then q0b
  double* s_e = s_q + 5 * NGLL * n;  // [10][n]: eL|eR, then eLb|eRb, then stencil adjoints
  double* s_k = s_e + 10 * n;        // [LINE_K]: background and power references of an x or y line
  double* s_ao = s_k + LINE_K;       // [5][n]: adj_out along the line
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  // background of cell c: the line's level for x and y (read from LDS, so
  // that these per-line values live in vector registers only where they are
  // used), the cell's own for z
  auto bg_of = [&](int c) {
    if constexpr (DIR == 2) {
      return cell_bg<NGLL, DIR>(G, c);
    } else {
      CellBg<NGLL> bg{};
      bg.r = s_k[0];
      bg.rt = s_k[1];
      bg.pt = PowRef{s_k[2], s_k[3], s_k[4], s_k[5]};
      bg.wall_bot = bg.wall_top = false;
      return bg;
    }
  };
  auto fref_of = [&](int f) {
    if constexpr (DIR == 2) {
      return face_ref<DIR>(G, kl, f);
    } else {
      FaceRef fr;
      fr.q5 = PowRef{s_k[2], s_k[3], s_k[4], s_k[5]};
      fr.rt = DIR == 0 ? fr.q5 : PowRef{s_k[6], s_k[7], s_k[8], s_k[9]};
      return fr;
    }
  };
  // 1. stage the line (boundary rule at load), adj_out along it and an x or
  // y line's constants.  Every global load of a thread is issued before its
  // first LDS store, so the stage waits for memory once (each thread holds
  // at most one halo cell, and one cell of the line when n <= bd).
  {
    const int t = threadIdx.x;
    const bool halo = t < 2 * hs;
    const int hc = t < hs ? t : n + t;  // padded slot of halo cell t
    double hv[5];
    if (halo) {
      bool zn;
      const int src = map_axis(hc - hs, n, per, zn);
      sfor<5>([&](auto l) { hv[l] = (zn && l == NM) ? 0.0 : in[l * G.sl + base + src * stride]; });
    }
    double kv[LINE_K];
    if constexpr (DIR != 2) {
      if (t == 0) {
        const CellBg<NGLL> bg = cell_bg<NGLL, DIR>(G, kl);
        const FaceRef fr = face_ref<DIR>(G, kl, 0);
        kv[0] = bg.r;
        kv[1] = bg.rt;
        kv[2] = bg.pt.h, kv[3] = bg.pt.l, kv[4] = bg.pt.rb, kv[5] = bg.pt.b;
        kv[6] = fr.rt.h, kv[7] = fr.rt.l, kv[8] = fr.rt.rb, kv[9] = fr.rt.b;
      }
    }
    for (int c = t; c < n; c += bd) {
      double v[5], a[5];
      sfor<5>([&](auto l) {
        v[l] = in[l * G.sl + base + c * stride];
        a[l] = adj_out[l * G.sl + base + c * stride];
      });
      sfor<5>([&](auto l) {
        s_st[l * W + hs + c] = v[l];
        s_ao[l * n + c] = a[l];
      });
    }
    if (halo) sfor<5>([&](auto l) { s_st[l * W + hc] = hv[l]; });
    if constexpr (DIR != 2) {
      if (t == 0) sfor<LINE_K>([&](auto k) { s_k[k] = kv[k]; });
    }
  }
  __syncthreads();
  // 2. cell forward: WENO values (independent chains the scheduler
  // interleaves: two variables at a time in x, all five in y and z, as
  // measured fastest) and edge states
  auto dow = [&](int l) { return (l == idR || l == idT) ? G.weno_s : G.weno_w; };
  auto recon = [&](int c, int l0, auto nv) {
    constexpr int NV = decltype(nv)::value;
    double w[NV][ORD], g[NV][NGLL];
    sfor<NV>([&](auto v) { sfor<ORD>([&](auto s) { w[v][s] = s_st[(l0 + v) * W + c + s]; }); });
    bool all = true;
    sfor<NV>([&](auto v) { all = all && dow(l0 + v); });
    if (all) sfor<NV>([&](auto v) { reconstruct<ORD, NGLL>(w[v], true, g[v]); });
    else sfor<NV>([&](auto v) { reconstruct<ORD, NGLL>(w[v], dow(l0 + v), g[v]); });
    sfor<NV>([&](auto v) { sfor<NGLL>([&](auto ii) { s_q[((l0 + v) * NGLL + ii) * n + c] = g[v][ii]; }); });
  };
  for (int c = threadIdx.x; c < n; c += bd) {
    if constexpr (DIR == 0) {
#pragma unroll 1
      for (int l0 = 0; l0 < 4; l0 += 2) recon(c, l0, std::integral_constant<int, 2>{});
      recon(c, 4, std::integral_constant<int, 1>{});
    } else {
      recon(c, 0, std::integral_constant<int, 5>{});
    }
    double q0[5][NGLL], eL[5], eR[5], sa;
    sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { q0[l][ii] = s_q[(l * NGLL + ii) * n + c]; }); });
    adj::Tape<NGLL, double> tp;
    adj::cell_dt<NGLL, DIR>(q0, bg_of(c), K, tp, eL, eR, sa);
    sfor<5>([&](auto l) {
      s_e[l * n + c] = eL[l];
      s_e[(5 + l) * n + c] = eR[l];
    });
  }
  __syncthreads();
  // 3. faces: flux adjoint -> edge-state adjoints, over the edge states.  A
  // z line with walls has n + 1 faces for n threads: thread 0 takes both wall
  // faces through the short wall form (riemann_adj_zwall), the others one
  // interior face each, all in one pass.
  const double cdr = K.dt * K.rd;
  const bool zwall = DIR == 2 && !per;
  const int nfaces = (per || zwall) ? n : n + 1;
  for (int f = threadIdx.x; f < nfaces; f += bd) {
    if (zwall && f == 0) {
      sfor<2>([&](auto side) {         // face 0, then face n
        const int sl = side == 0 ? 0 : 5 * n + n - 1;            // the boundary cell's edge slot
        const int ca = side == 0 ? 0 : n - 1;
        double L[5], Fb[5], Lb[5];
        sfor<5>([&](auto l) {
          L[l] = s_e[sl + l * n];
          const double v = side == 0 ? s_ao[l * n + ca] : -s_ao[l * n + ca];
          Fb[l] = (G.sim2d && l == idV) ? 0.0 : cdr * v;
          Lb[l] = 0.0;
        });
        adj::riemann_adj_zwall(L, fref_of(side == 0 ? 0 : n), Fb, Lb);
        sfor<5>([&](auto l) { s_e[sl + l * n] = Lb[l]; });
      });
      continue;
    }
    const bool wl = !per && f == 0, wr = !per && f == n;
    const int a = wl ? 0 : (f == 0 ? n - 1 : f - 1);      // cell on the low side
    const int b = wr ? n - 1 : (f == n ? 0 : f);          // cell on the high side
    // slots this face owns: the low cell's high edge and the high cell's
    // low edge (at a wall, the one edge of the boundary cell)
    const int sL = wl ? 0 : (5 * n + a);                  // + l * n
    const int sR = wr ? (5 * n + n - 1) : b;
    double L[5], R[5], Fb[5], Lb[5], Rb[5];
    sfor<5>([&](auto l) {
      L[l] = s_e[sL + l * n];
      R[l] = s_e[sR + l * n];
      double v = 0.0;
      if (!wr) v += s_ao[l * n + b];
      if (!wl) v -= s_ao[l * n + a];
      Fb[l] = (G.sim2d && l == idV) ? 0.0 : cdr * v;
      Lb[l] = Rb[l] = 0.0;
    });
    adj::riemann_adj<DIR>(L, R, fref_of(f), Fb, Lb, Rb);
    if (wl || wr) {
      sfor<5>([&](auto l) { s_e[sL + l * n] = Lb[l] + Rb[l]; });
    } else {
      sfor<5>([&](auto l) {
        s_e[sL + l * n] = Lb[l];
        s_e[sR + l * n] = Rb[l];
      });
    }
  }
  __syncthreads();
  // 4a. transform reverse: q0b over q0 (each cell's own column); in z the
  // source enters rho*w as out += dt * src
  for (int c = threadIdx.x; c < n; c += bd) {
    double q0[5][NGLL], q0b[5][NGLL], eLb[5], eRb[5];
    sfor<5>([&](auto l) {
      sfor<NGLL>([&](auto ii) { q0[l][ii] = s_q[(l * NGLL + ii) * n + c]; });
      eLb[l] = s_e[l * n + c];
      eRb[l] = s_e[(5 + l) * n + c];
    });
    const double srcb = DIR == 2 ? K.dt * s_ao[idW * n + c] : 0.0;
    adj::cell_dt_adj<NGLL, DIR>(q0, bg_of(c), K, eLb, eRb, srcb, q0b);
    sfor<5>([&](auto l) { sfor<NGLL>([&](auto ii) { s_q[(l * NGLL + ii) * n + c] = q0b[l][ii]; }); });
  }
  // 4b. WENO reverse in each cell's own column, PV variables at a time side
  // by side (independent chains the scheduler interleaves).  The stencil
  // adjoints of every variable go to the row region [s_q | s_e] (5 NGLL + 10
  // rows; adj_row), in rows that hold no q0b of a later group, so one
  // barrier and one gather of the stencil adjoints that point at each cell
  // finish the sweep.
  static_assert(adj_rows_ok<ORD, NGLL, adj_pv<ORD>()>(), "stencil adjoints fit the row region");
  constexpr int PV = adj_pv<ORD>();
  double* s_r = s_q;
  auto weno_rev = [&](auto l0c, auto nv) {
    constexpr int l0 = decltype(l0c)::value;
    constexpr int NV = decltype(nv)::value;
    bool all = true;
    sfor<NV>([&](auto v) { all = all && dow(l0 + v); });
    for (int c = threadIdx.x; c < n; c += bd) {
      double u[NV][ORD], ub[NV][ORD], gb[NV][NGLL];
      sfor<NV>([&](auto v) {
        sfor<ORD>([&](auto s) {
          u[v][s] = s_st[(l0 + v) * W + c + s];
          ub[v][s] = 0.0;
        });
        sfor<NGLL>([&](auto ii) { gb[v][ii] = s_q[((l0 + v) * NGLL + ii) * n + c]; });
      });
      if (all) {
        sfor<NV>([&](auto v) { adj::reconstruct_adj<ORD, NGLL>(u[v], true, gb[v], ub[v]); });
      } else {
        sfor<NV>([&](auto v) { adj::reconstruct_adj<ORD, NGLL>(u[v], dow(l0 + v), gb[v], ub[v]); });
      }
      sfor<NV>([&](auto v) {
        sfor<ORD>([&](auto s) { s_r[adj_row<ORD, NGLL>(l0 + v, s) * n + c] = ub[v][s]; });
      });
    }
  };
  sfor<5 / PV>([&](auto g) { weno_rev(std::integral_constant<int, g * PV>{}, std::integral_constant<int, PV>{}); });
  if constexpr (5 % PV != 0)
    weno_rev(std::integral_constant<int, 5 - 5 % PV>{}, std::integral_constant<int, 5 % PV>{});
  __syncthreads();
  // src = c + hs - s is the cell whose stencil slot s is c: a periodic line
  // wraps it (by one period when n >= hs, else by the remainder), a wall
  // drops it when it falls outside the line.  At a wall the stencil slots
  // past the end read the clamped end cell; the normal momentum there is
  // zero, so it passes nothing back.
  // MODE 0: periodic with n < hs (wrap by the remainder), 1: periodic with
  // n >= hs (one period at most, every slot in the line), 2: walls
  auto gather = [&](int l0, auto nv, auto mode) {
    constexpr int NV = decltype(nv)::value;
    constexpr int MODE = decltype(mode)::value;
    for (int c = threadIdx.x; c < n; c += bd) {
      sfor<NV>([&](auto v) {
        const int l = l0 + v;
        double acc = s_ao[l * n + c];
        sfor<ORD>([&](auto s) {
          if constexpr (s == hs) {
            acc += s_r[adj_row<ORD, NGLL>(l, s) * n + c];
          } else {
            int src = c + hs - s;
            if constexpr (MODE == 0) {
              src %= n;
              src += src < 0 ? n : 0;
            } else if constexpr (MODE == 1) {
              src += src < 0 ? n : 0;
              src -= src >= n ? n : 0;
            }
            if constexpr (MODE != 2) {
              acc += s_r[adj_row<ORD, NGLL>(l, s) * n + src];
            } else {
              const bool ok = src >= 0 && src < n;
              const double val = s_r[adj_row<ORD, NGLL>(l, s) * n + (ok ? src : c)];
              acc += ok ? val : 0.0;
            }
          }
        });
        if (MODE == 2 && l != NM) {
          if (c == 0)                    // slot s of cell cc reads index cc - hs + s < 0
            sfor<hs>([&](auto cc) {
              sfor<ORD>([&](auto s) {
                if constexpr (cc - hs + s < 0) {
                  if (cc < n) acc += s_r[adj_row<ORD, NGLL>(l, s) * n + cc];
                }
              });
            });
          if (c == n - 1)                // slot s of cell n-1-m reads index >= n
            sfor<hs>([&](auto m) {
              const int cc = n - 1 - m;
              sfor<ORD>([&](auto s) {
                if constexpr (s >= hs + 1 + m) {
                  if (cc >= 0) acc += s_r[adj_row<ORD, NGLL>(l, s) * n + cc];
                }
              });
            });
        }
        adj_in[l * G.sl + base + c * stride] = acc;
      });
    }
  };
  if (!per) gather(0, std::integral_constant<int, 5>{}, std::integral_constant<int, 2>{});
  else if (n < hs) gather(0, std::integral_constant<int, 5>{}, std::integral_constant<int, 0>{});
  else gather(0, std::integral_constant<int, 5>{}, std::integral_constant<int, 1>{});
  // the halo cells of an x line (y and z halos: k_zero_halo)
  if constexpr (DIR == 0) {
    for (int c = threadIdx.x; c < 2 * hs; c += bd) {
      const int64_t o = c < hs ? base - hs + c : base + n + (c - hs);
      sfor<5>([&](auto l) { adj_in[l * G.sl + o] = 0.0; });
    }
  }
}

// zero the halo of a padded array, one thread per halo cell: the 2hs z
// planes, the 2hs y rows of every interior level and, with xh, the 2hs x
// cells of every interior row (x line kernels write those themselves)
__host__ __device__ inline int64_t halo_cells(const Grid& G, bool xh) {
  const int64_t nxh = G.nx + 2 * G.hs, nyh = G.ny + 2 * G.hs, h2 = 2 * G.hs;
  return h2 * nyh * nxh + (int64_t)G.nz * h2 * nxh + (xh ? (int64_t)G.nz * G.ny * h2 : 0);
}

__global__ __launch_bounds__(256) void k_zero_halo(Grid G, bool xh, double* __restrict__ a) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= halo_cells(G, xh)) return;
  const int hs = G.hs, h2 = 2 * hs;
  const int64_t nxh = G.nx + h2, nyh = G.ny + h2;
  const int64_t nA = h2 * nyh * nxh, nB = (int64_t)G.nz * h2 * nxh;
  int64_t kp, jp, ip;                  // padded coordinates
  if (e < nA) {
    const int64_t plane = e / (nyh * nxh), r = e % (nyh * nxh);
    kp = plane < hs ? plane : G.nz + plane;
    jp = r / nxh;
    ip = r % nxh;
  } else if (e < nA + nB) {
    const int64_t p = e - nA, r = p % (h2 * nxh), row = r / nxh;
    kp = hs + p / (h2 * nxh);
    jp = row < hs ? row : G.ny + row;
    ip = r % nxh;
  } else {
    const int64_t p = e - nA - nB, r = p % ((int64_t)G.ny * h2), col = r % h2;
    kp = hs + p / ((int64_t)G.ny * h2);
    jp = hs + r / h2;
    ip = col < hs ? col : G.nx + col;
  }
  const int64_t t = kp * G.sz + jp * G.sy + ip;
  sfor<5>([&](auto l) { a[l * G.sl + t] = 0.0; });
}

// adj_in over the padded layout: zero in the halo, in the interior the
// identity part plus every stencil adjoint that points at the cell
template <int ORD, int DIR>
__global__ __launch_bounds__(64) void k_gather(const double* __restrict__ STB, const double* __restrict__ adj_out,
                                                Grid G, double* __restrict__ adj_in) {
  constexpr int hs = (ORD - 1) / 2;
  constexpr int NM = 1 + DIR;
  // grid (ceil((nx+2hs)/64), ny+2hs, nz+2hs) over the padded layout
  const int ih = blockIdx.x * blockDim.x + threadIdx.x;
  if (ih >= G.nx + 2 * G.hs) return;
  const int64_t t = (int64_t)blockIdx.z * G.sz + (int64_t)blockIdx.y * G.sy + ih;
  const int i = ih - G.hs, j = (int)blockIdx.y - G.hs, k = (int)blockIdx.z - G.hs;
  if (i < 0 || i >= G.nx || j < 0 || j >= G.ny || k < 0 || k >= G.nz) {
    sfor<5>([&](auto l) { adj_in[l * G.sl + t] = 0.0; });
    return;
  }
  const int64_t nc = ncell(G);
  const Pos p{i, j, k};
  const int n = axis_len<DIR>(G);
  const bool per = G.bc[DIR] == AWFL_BC_PERIODIC;
  const int a = along<DIR>(p);
  double acc[5];
  sfor<5>([&](auto l) { acc[l] = adj_out[l * G.sl + t]; });
  sfor<ORD>([&](auto s) {
    // the cell whose stencil slot s is this cell
    int src = a + hs - s;
    bool ok = true;
    if (per) {
      src %= n;
      if (src < 0) src += n;
    } else {
      ok = src >= 0 && src < n;
    }
    if (ok) {
      const int64_t c = cidx(G, moved<DIR>(p, src));
      sfor<5>([&](auto l) { acc[l] += STB[(l * ORD + s) * nc + c]; });
    }
  });
  if (!per) {
    // clamped slots: below the first cell (cells 0..hs-1) and above the last
    if (a == 0) {
      sfor<hs>([&](auto cc) {
        sfor<ORD>([&](auto s) {
          if constexpr (cc - hs + s < 0) {
            if (cc < n) {
              const int64_t c = cidx(G, moved<DIR>(p, cc));
              sfor<5>([&](auto l) {
                if constexpr (l != NM) acc[l] += STB[(l * ORD + s) * nc + c];
              });
            }
          }
        });
      });
    }
    if (a == n - 1) {
      sfor<hs>([&](auto m) {
        const int cc = n - 1 - m;       // cells n-hs..n-1
        sfor<ORD>([&](auto s) {
          // slot s of cell cc lies at cc - hs + s >= n  <=>  s >= hs + 1 + m
          if constexpr (s >= hs + 1 + m) {
            if (cc >= 0) {
              const int64_t c = cidx(G, moved<DIR>(p, cc));
              sfor<5>([&](auto l) {
                if constexpr (l != NM) acc[l] += STB[(l * ORD + s) * nc + c];
              });
            }
          }
        });
      });
    }
  }
  sfor<5>([&](auto l) { adj_in[l * G.sl + t] = acc[l]; });
}

// tangent-linear update over the padded layout (halo zero)
template <int NGLL, int DIR>
__global__ __launch_bounds__(64) void k_tlm_update(const double* __restrict__ in, const double* __restrict__ din,
                                                    const double* __restrict__ E, const double* __restrict__ Ed,
                                                    Grid G, SweepK K, double* __restrict__ out,
                                                    double* __restrict__ dout) {
  constexpr int NM = 1 + DIR;
  const int ih = blockIdx.x * blockDim.x + threadIdx.x;
  if (ih >= G.nx + 2 * G.hs) return;
  const int64_t t = (int64_t)blockIdx.z * G.sz + (int64_t)blockIdx.y * G.sy + ih;
  const int i = ih - G.hs, j = (int)blockIdx.y - G.hs, k = (int)blockIdx.z - G.hs;
  if (i < 0 || i >= G.nx || j < 0 || j >= G.ny || k < 0 || k >= G.nz) {
    sfor<5>([&](auto l) { out[l * G.sl + t] = dout[l * G.sl + t] = 0.0; });
    return;
  }
  const int64_t nc = ncell(G);
  const Pos p{i, j, k};
  const int a = along<DIR>(p);
  auto flux = [&](int f, Dual (&F)[5]) {
    int64_t cl, cr;
    int offl, offr;
    face_states<DIR>(E, G, p, f, cl, cr, offl, offr);
    Dual L[5], R[5];
    sfor<5>([&](auto l) {
      L[l] = Dual(E[(offl + l) * nc + cl], Ed[(offl + l) * nc + cl]);
      R[l] = Dual(E[(offr + l) * nc + cr], Ed[(offr + l) * nc + cr]);
    });
    adj::riemann_fwd<DIR>(L, R, face_ref<DIR>(G, p.k, f), DIR == 2 ? face_pressure<NGLL>(G, f) : 0.0, F);
  };
  Dual F0[5], F1[5];
  flux(a, F0);
  flux(a + 1, F1);
  const int64_t c = cidx(G, p);
  sfor<5>([&](auto l) {
    Dual td = (F1[l] - F0[l]) * Dual(-K.rd);
    if constexpr (DIR == 2 && l == idW) td = Dual(E[10 * nc + c], Ed[10 * nc + c]) + td;
    if (G.sim2d && l == idV) td = Dual(0.0);
    out[l * G.sl + t] = in[l * G.sl + t] + K.dt * td.v;
    dout[l * G.sl + t] = din[l * G.sl + t] + K.dt * td.d;
  });
  (void)NM;
}

struct VariantA {
  int ord, ngll;
  void (*edges[3])(const double*, const double*, Grid, SweepK, double*, double*, double*);
  void (*edges_d[3])(const double*, const double*, Grid, SweepK, double*, double*, double*);
  void (*face[3])(const double*, const double*, Grid, SweepK, double*);
  void (*dt[3])(const double*, const double*, const double*, Grid, SweepK, double*);
  void (*weno[3])(const double*, const double*, Grid, double*);
  void (*gather[3])(const double*, const double*, Grid, double*);
  void (*tlm[3])(const double*, const double*, const double*, const double*, Grid, SweepK, double*, double*);
  void (*line[3])(const double*, const double*, double*, Grid, SweepK);
  int line_rows;
};

#define AWFL_VARIANT_A(O, N)                                                                        \
  VariantA {                                                                                        \
    O, N, {k_edges<O, N, 0, double>, k_edges<O, N, 1, double>, k_edges<O, N, 2, double>},           \
        {k_edges<O, N, 0, Dual>, k_edges<O, N, 1, Dual>, k_edges<O, N, 2, Dual>},                   \
        {k_face_adj<N, 0>, k_face_adj<N, 1>, k_face_adj<N, 2>},                                     \
        {k_dt_adj<N, 0>, k_dt_adj<N, 1>, k_dt_adj<N, 2>},                                           \
        {k_weno_adj<O, N, 0>, k_weno_adj<O, N, 1>, k_weno_adj<O, N, 2>},                            \
        {k_gather<O, 0>, k_gather<O, 1>, k_gather<O, 2>},                                           \
        {k_tlm_update<N, 0>, k_tlm_update<N, 1>, k_tlm_update<N, 2>},                               \
        {k_adj_line<O, N, 0>, k_adj_line<O, N, 1>, k_adj_line<O, N, 2>}, adj_line_rows<O, N>(),       \
  }

const VariantA kVariantsA[] = {AWFL_VARIANT_A(3, 2), AWFL_VARIANT_A(5, 3)};

const VariantA* find_variant_a(int ord, int ngll) {
  for (const VariantA& v : kVariantsA)
    if (v.ord == ord && v.ngll == ngll) return &v;
  return nullptr;
}

int check_variant(awfl_ctx* c) {
  if (!find_variant_a(c->cfg.ord, c->cfg.ngll))
    return fail(c, AWFL_E_UNSUPPORTED, "adjoint: (ord, ngll) not compiled in");
  return AWFL_OK;
}

// per-cell scratch of the multi-kernel sweeps (tangent; adjoint lines too
// long for LDS)
int ensure_cell_scratch(awfl_ctx* c) {
  int rc = check_variant(c);
  if (rc != AWFL_OK) return rc;
  const int64_t nc = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  const int per_cell = NE + NEB + 5 * c->cfg.ord + 10 * c->cfg.ngll;
  if (!c->a_scr) HIP_TRY(c, hipMalloc(&c->a_scr, sizeof(double) * nc * per_cell));
  return AWFL_OK;
}

// second adjoint state of the step reversal
int ensure_tmp(awfl_ctx* c) {
  int rc = check_variant(c);
  if (rc != AWFL_OK) return rc;
  if (!c->a_tmp) HIP_TRY(c, hipMalloc(&c->a_tmp, sizeof(double) * c->state_elems()));
  return AWFL_OK;
}

// LDS bytes of the fused line kernel along dir (0 if it does not fit)
size_t line_lds(const awfl_ctx* c, const VariantA* v, int dir) {
  const int n = dir == 0 ? c->cfg.nx : dir == 1 ? c->cfg.ny : c->cfg.nz;
  // state line, row regions, line constants, adj_out along the line
  const size_t b = sizeof(double) * (size_t)(5 * (n + 2 * c->hs) + v->line_rows * n + LINE_K + 5 * n);
  return b <= 160 * 1024 ? b : 0;
}

// launch shapes: cells (ceil(nx/64), ny, nz), (cell, variable) pairs
// (ceil(nx/64), 5 ny, nz), padded layout (ceil((nx+2hs)/64), ny+2hs, nz+2hs)
dim3 cells_grid(const awfl_ctx* c) { return dim3((c->cfg.nx + 63) / 64, c->cfg.ny, c->cfg.nz); }
dim3 cellvar_grid(const awfl_ctx* c) { return dim3((c->cfg.nx + 63) / 64, 5 * c->cfg.ny, c->cfg.nz); }
dim3 padded_grid(const awfl_ctx* c) {
  return dim3((c->cfg.nx + 2 * c->hs + 63) / 64, c->cfg.ny + 2 * c->hs, c->cfg.nz + 2 * c->hs);
}

// adj_in = adj_out + J(in)^T (dt adj_out) for one sweep
// zero_halo: leave the halo of adj_in zero (one extra kernel; x lines write
// their own x halo cells).  Chains of adjoint sweeps read only interior
// cells, so they zero the halo once at the end instead.
int sweep_adjoint(awfl_ctx* c, const double* in, const double* adj_out, double* adj_in, double dt, int dir,
                  bool zero_halo = true) {
  if (dir == 1 && c->sim2d) {
    HIP_TRY(c, hipMemcpyAsync(adj_in, adj_out, sizeof(double) * c->state_elems(), hipMemcpyDeviceToDevice,
                              c->stream));
    return AWFL_OK;
  }
  int rc = check_variant(c);
  if (rc != AWFL_OK) return rc;
  const VariantA* v = find_variant_a(c->cfg.ord, c->cfg.ngll);
  const Grid G = c->grid();
  const SweepK K = sweep_constants(c, dt, dir);
  const int64_t nc = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  const int ng = c->cfg.ngll;
  if (const size_t lds = line_lds(c, v, dir)) {
    const int n = dir == 0 ? c->cfg.nx : dir == 1 ? c->cfg.ny : c->cfg.nz;
    const unsigned lines = (unsigned)(nc / n);
    if (lds > 64 * 1024)
      HIP_TRY(c, hipFuncSetAttribute((const void*)v->line[dir], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int threads = line_threads(n);
    v->line[dir]<<<lines, threads, lds, c->stream>>>(in, adj_out, adj_in, G, K);
    LAUNCH_CHECK(c);
    if (zero_halo) {
      const int64_t nh = halo_cells(G, dir != 0);
      k_zero_halo<<<(unsigned)((nh + 255) / 256), 256, 0, c->stream>>>(G, dir != 0, adj_in);
      LAUNCH_CHECK(c);
    }
    return AWFL_OK;
  }
  rc = ensure_cell_scratch(c);
  if (rc != AWFL_OK) return rc;
  double* E = c->a_scr;
  double* EB = E + NE * nc;
  double* STB = EB + NEB * nc;
  double* Q0 = STB + (int64_t)5 * c->cfg.ord * nc;
  double* Q0B = Q0 + (int64_t)5 * ng * nc;
  v->edges[dir]<<<cells_grid(c), 64, 0, c->stream>>>(in, nullptr, G, K, E, nullptr, Q0);
  LAUNCH_CHECK(c);
  v->face[dir]<<<cells_grid(c), 64, 0, c->stream>>>(E, adj_out, G, K, EB);
  LAUNCH_CHECK(c);
  v->dt[dir]<<<cells_grid(c), 64, 0, c->stream>>>(Q0, EB, adj_out, G, K, Q0B);
  LAUNCH_CHECK(c);
  v->weno[dir]<<<cellvar_grid(c), 64, 0, c->stream>>>(in, Q0B, G, STB);
  LAUNCH_CHECK(c);
  v->gather[dir]<<<padded_grid(c), 64, 0, c->stream>>>(STB, adj_out, G, adj_in);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int split_dirs(const awfl_ctx* c, int ds, int (&dirs)[3]) {
  static const int fwd[3] = {0, 1, 2}, bwd[3] = {2, 1, 0};
  int nd = 0;
  for (int s = 0; s < 3; s++) {
    int d = ds ? fwd[s] : bwd[s];
    if (d == 1 && c->sim2d) continue;
    dirs[nd++] = d;
  }
  return nd;
}

int ensure_states(awfl_ctx* c, int64_t nstates) {
  if (c->a_states_n >= nstates) return AWFL_OK;
  c->a_rec_ckpt = nullptr;
  if (c->a_states) HIP_TRY(c, hipFree(c->a_states));
  c->a_states = nullptr;
  c->a_states_n = 0;
  HIP_TRY(c, hipMalloc(&c->a_states, sizeof(double) * c->state_elems() * nstates));
  c->a_states_n = nstates;
  return AWFL_OK;
}

// reverse of the steps [n0, n1) whose split inputs are in `states`; adj is
// updated in place (ping-pong through a_tmp)
int reverse_steps(awfl_ctx* c, const double* states, int n0, int n1, int ds0, double dt, double* adj) {
  const int64_t el = c->state_elems();
  double* cur = adj;
  double* oth = c->a_tmp;
  int dirs[3];
  for (int n = n1 - 1; n >= n0; n--) {
    const int ds = (n % 2 == 0) ? ds0 : !ds0;
    const int nd = split_dirs(c, ds, dirs);
    for (int k = nd - 1; k >= 0; k--) {
      const double* in = states + (int64_t)((n - n0) * nd + k) * el;
      int rc = sweep_adjoint(c, in, cur, oth, dt, dirs[k], false);
      if (rc != AWFL_OK) return rc;
      std::swap(cur, oth);
    }
  }
  if (cur != adj) HIP_TRY(c, hipMemcpyAsync(adj, cur, sizeof(double) * el, hipMemcpyDeviceToDevice, c->stream));
  if (n1 > n0) {
    const Grid G = c->grid();
    const int64_t nh = halo_cells(G, true);
    k_zero_halo<<<(unsigned)((nh + 255) / 256), 256, 0, c->stream>>>(G, true, adj);
    LAUNCH_CHECK(c);
  }
  return AWFL_OK;
}

// split inputs of steps [n0, n1) into `states`, from the state before step n0
int record_into(awfl_ctx* c, double* states, const double* start, int n0, int n1, int ds0, double dt) {
  const int64_t el = c->state_elems();
  int dirs[3];
  const int nd = split_dirs(c, 1, dirs);
  HIP_TRY(c, hipMemcpyAsync(states, start, sizeof(double) * el, hipMemcpyDeviceToDevice, c->stream));
  const int64_t last = (int64_t)(n1 - n0) * nd - 1;
  for (int n = n0; n < n1; n++) {
    const int ds = (n % 2 == 0) ? ds0 : !ds0;
    split_dirs(c, ds, dirs);
    for (int k = 0; k < nd; k++) {
      const int64_t idx = (int64_t)(n - n0) * nd + k;
      if (idx == last) break;
      int rc = awfl_sweep(c, states + idx * el, states + (idx + 1) * el, dt, dirs[k]);
      if (rc != AWFL_OK) return rc;
    }
  }
  return AWFL_OK;
}

// the same into c->a_states, on the context's stream
int record_steps(awfl_ctx* c, const double* start, int n0, int n1, int ds0, double dt) {
  c->a_rec_ckpt = nullptr;             // a_states is rewritten
  int dirs[3];
  const int nd = split_dirs(c, 1, dirs);
  int rc = ensure_states(c, (int64_t)(n1 - n0) * nd);
  if (rc != AWFL_OK) return rc;
  return record_into(c, c->a_states, start, n0, n1, ds0, dt);
}

// the second split-input buffer, the recomputation stream and its events
int ensure_pipeline(awfl_ctx* c, int64_t nstates) {
  if (c->a_states2_n < nstates) {
    c->a_rec_ckpt = nullptr;
    if (c->a_states2) HIP_TRY(c, hipFree(c->a_states2));
    c->a_states2 = nullptr;
    c->a_states2_n = 0;
    HIP_TRY(c, hipMalloc(&c->a_states2, sizeof(double) * c->state_elems() * nstates));
    c->a_states2_n = nstates;
  }
  if (!c->a_s2) HIP_TRY(c, hipStreamCreateWithFlags(&c->a_s2, hipStreamNonBlocking));
  for (hipEvent_t& e : c->a_ev)
    if (!e) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return AWFL_OK;
}

// launches of the context go to another stream while it lives
struct StreamSwap {
  awfl_ctx* c;
  hipStream_t saved;
  StreamSwap(awfl_ctx* ctx, hipStream_t s) : c(ctx), saved(ctx->stream) { c->stream = s; }
  ~StreamSwap() { c->stream = saved; }
};

}  // namespace

extern "C" {


int awfl_sweep_tangent(awfl_ctx* c, const double* d_in, const double* d_din, double* d_out, double* d_dout,
                       double dt, int dir) {
  if (!c) return AWFL_E_ARG;
  if (!d_in || !d_din || !d_out || !d_dout) return fail(c, AWFL_E_ARG, "null buffer");
  if (d_in == d_out || d_din == d_dout) return fail(c, AWFL_E_ARG, "tangent sweeps run out of place");
  if (dir < 0 || dir > 2) return fail(c, AWFL_E_ARG, "dir must be 0, 1 or 2");
  if (dir == 1 && c->sim2d) {
    HIP_TRY(c, hipMemcpyAsync(d_out, d_in, sizeof(double) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(d_dout, d_din, sizeof(double) * c->state_elems(), hipMemcpyDeviceToDevice, c->stream));
    return AWFL_OK;
  }
  int rc = ensure_cell_scratch(c);
  if (rc != AWFL_OK) return rc;
  const VariantA* v = find_variant_a(c->cfg.ord, c->cfg.ngll);
  const Grid G = c->grid();
  const SweepK K = sweep_constants(c, dt, dir);
  const int64_t nc = (int64_t)c->cfg.nx * c->cfg.ny * c->cfg.nz;
  double* E = c->a_scr;
  double* Ed = E + NE * nc;            // fits: NE <= NEB + 5*ord
  v->edges_d[dir]<<<cells_grid(c), 64, 0, c->stream>>>(d_in, d_din, G, K, E, Ed, nullptr);
  LAUNCH_CHECK(c);
  v->tlm[dir]<<<padded_grid(c), 64, 0, c->stream>>>(d_in, d_din, E, Ed, G, K, d_out, d_dout);
  LAUNCH_CHECK(c);
  return AWFL_OK;
}

int awfl_sweep_adjoint(awfl_ctx* c, const double* d_in, const double* d_adj_out, double* d_adj_in, double dt,
                       int dir) {
  if (!c) return AWFL_E_ARG;
  if (!d_in || !d_adj_out || !d_adj_in) return fail(c, AWFL_E_ARG, "null buffer");
  if (d_adj_out == d_adj_in) return fail(c, AWFL_E_ARG, "adjoint sweeps run out of place");
  if (dir < 0 || dir > 2) return fail(c, AWFL_E_ARG, "dir must be 0, 1 or 2");
  return sweep_adjoint(c, d_in, d_adj_out, d_adj_in, dt, dir);
}

int awfl_time_step_adjoint(awfl_ctx* c, const double* d_state_before, double* d_adj, double dt, int dim_switch) {
  if (!c) return AWFL_E_ARG;
  if (!d_state_before || !d_adj) return fail(c, AWFL_E_ARG, "null buffer");
  int rc = ensure_tmp(c);
  if (rc != AWFL_OK) return rc;
  rc = record_steps(c, d_state_before, 0, 1, dim_switch ? 1 : 0, dt);
  if (rc != AWFL_OK) return rc;
  return reverse_steps(c, c->a_states, 0, 1, dim_switch ? 1 : 0, dt, d_adj);
}

int awfl_forward_ckpt(awfl_ctx* c, double* d_state, double dt, int nsteps, int every, double* d_ckpt) {
  if (!c) return AWFL_E_ARG;
  if (!d_state || !d_ckpt) return fail(c, AWFL_E_ARG, "null buffer");
  if (nsteps < 0 || every < 1) return fail(c, AWFL_E_ARG, "nsteps >= 0 and every >= 1");
  const int64_t el = c->state_elems();
  // the last two segments run as the split sweeps the adjoint would
  // recompute, keeping their inputs for awfl_adjoint_ckpt on the same
  // checkpoints (the last in a_states, the one before in a_states2: the two
  // buffers the adjoint pipelines through anyway); the earlier segments run
  // as fused steps
  const int ds0 = c->dimSwitch ? 1 : 0;
  const int nseg = (nsteps + every - 1) / every;
  const int nkeep = std::min(2, nseg);
  const int nfused = (nseg - nkeep) * every;
  for (int n = 0; n < nfused; n++) {
    if (n % every == 0)
      HIP_TRY(c, hipMemcpyAsync(d_ckpt + (int64_t)(n / every) * el, d_state, sizeof(double) * el,
                                hipMemcpyDeviceToDevice, c->stream));
    int rc = awfl_time_step(c, d_state, dt);
    if (rc != AWFL_OK) return rc;
  }
  if (nsteps == 0) return AWFL_OK;
  // room for whole segments, as the adjoint needs, so that it keeps these
  int dirs[3];
  const int nd = split_dirs(c, 1, dirs);
  int rc = ensure_states(c, (int64_t)every * nd);
  if (rc == AWFL_OK && nkeep == 2) rc = ensure_pipeline(c, (int64_t)every * nd);
  if (rc != AWFL_OK) return rc;
  c->a_rec_ckpt = nullptr;             // the kept inputs are rewritten
  for (int m = nseg - nkeep; m < nseg; m++) {
    const int n0 = m * every, n1 = std::min(nsteps, n0 + every);
    double* buf = m == nseg - 1 ? c->a_states : c->a_states2;
    HIP_TRY(c, hipMemcpyAsync(d_ckpt + (int64_t)m * el, d_state, sizeof(double) * el, hipMemcpyDeviceToDevice,
                              c->stream));
    rc = record_into(c, buf, d_state, n0, n1, ds0, dt);
    if (rc != AWFL_OK) return rc;
    // the last sweep of the segment's last step, from its recorded input into the state
    split_dirs(c, ((n1 - 1) % 2 == 0) ? ds0 : !ds0, dirs);
    const int64_t last = (int64_t)(n1 - n0) * nd - 1;
    rc = awfl_sweep(c, buf + last * el, d_state, dt, dirs[nd - 1]);
    if (rc != AWFL_OK) return rc;
  }
  c->dimSwitch = (nsteps % 2 == 0) ? ds0 : !ds0;   // as nsteps fused steps leave it
  c->a_rec_ckpt = d_ckpt;
  c->a_rec_dt = dt;
  c->a_rec_nsteps = nsteps;
  c->a_rec_every = every;
  c->a_rec_ds0 = ds0;
  c->a_rec_kept = nkeep;
  return AWFL_OK;
}

int awfl_adjoint_ckpt(awfl_ctx* c, const double* d_ckpt, double dt, int nsteps, int every, int dim_switch0,
                      double* d_adj) {
  if (!c) return AWFL_E_ARG;
  if (!d_ckpt || !d_adj) return fail(c, AWFL_E_ARG, "null buffer");
  if (nsteps < 0 || every < 1) return fail(c, AWFL_E_ARG, "nsteps >= 0 and every >= 1");
  int rc = ensure_tmp(c);
  if (rc != AWFL_OK) return rc;
  const int64_t el = c->state_elems();
  const int ds0 = dim_switch0 ? 1 : 0;
  const int nseg = (nsteps + every - 1) / every;
  // the last segment's split inputs are still in place when the forward run
  // on these checkpoints was awfl_forward_ckpt on this context with the same
  // arguments, on the same stream, and nothing has rewritten them since.
  // They serve one adjoint only: a later adjoint on a buffer at the same
  // address (reused by the caller for other contents) recomputes.
  const bool kept = c->a_rec_ckpt == d_ckpt && c->a_rec_dt == dt && c->a_rec_nsteps == nsteps &&
                    c->a_rec_every == every && c->a_rec_ds0 == ds0;
  c->a_rec_ckpt = nullptr;
  if (nseg == 0) return AWFL_OK;
  int dirs[3];
  const int nd = split_dirs(c, 1, dirs);
  const int64_t nst = (int64_t)every * nd;
  rc = ensure_states(c, nst);            // keeps a_states when it is large enough
  if (rc == AWFL_OK && nseg > 1) rc = ensure_pipeline(c, nst);
  if (rc != AWFL_OK) return rc;
  // segment m's split inputs are in buf[m % 2 == (nseg - 1) % 2 ? 0 : 1]; the
  // last segment's in a_states and the one before in a_states2 (kept from
  // the forward run, or recomputed here).  While the sweep adjoints of
  // segment m run on the context's stream, segment m-1 is recomputed into
  // the other buffer on a_s2: after everything enqueued so far on the
  // context's stream (the reverse of segment m+1 read that buffer last), and
  // the reverse of segment m-1 waits for it.
  const int nkept = kept ? c->a_rec_kept : 0;
  double* buf[2] = {c->a_states, c->a_states2};
  int cur = 0;
  if (nkept < 1) {
    const int m = nseg - 1;
    rc = record_into(c, buf[cur], d_ckpt + (int64_t)m * el, m * every, std::min(nsteps, m * every + every), ds0, dt);
    if (rc != AWFL_OK) return rc;
  }
  for (int m = nseg - 1; m >= 0; m--) {
    const int n0 = m * every, n1 = std::min(nsteps, n0 + every);
    const bool recompute = m > 0 && m - 1 < nseg - nkept;   // segment m-1 not kept
    if (recompute) {
      HIP_TRY(c, hipEventRecord(c->a_ev[0], c->stream));
      HIP_TRY(c, hipStreamWaitEvent(c->a_s2, c->a_ev[0], 0));
      {
        StreamSwap sw(c, c->a_s2);
        rc = record_into(c, buf[1 - cur], d_ckpt + (int64_t)(m - 1) * el, n0 - every, n0, ds0, dt);
      }
      if (rc != AWFL_OK) return rc;
      HIP_TRY(c, hipEventRecord(c->a_ev[1], c->a_s2));
    }
    rc = reverse_steps(c, buf[cur], n0, n1, ds0, dt, d_adj);
    if (rc != AWFL_OK) return rc;
    if (recompute) HIP_TRY(c, hipStreamWaitEvent(c->stream, c->a_ev[1], 0));
    cur = 1 - cur;
  }
  return AWFL_OK;
}

}  // extern "C"
