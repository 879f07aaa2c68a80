// Fused finite-volume stencil ("atom") builder for one matrix row.
//
// One row = equation d1 of one cell.  The row's atom is the dof x 3x3x3 block
// A[d2][sx][sy][sz] of coefficients on variable d2 of the neighbour at slot
// offset (sx-1, sy-1, sz-1) (transiflow/Discretization.py:71-102).  It is
// built in registers from the coordinate vectors and the state, term by term,
// in the same floating-point order as the reference's whole-grid atom sums:
//
//   linear part     Discretization.py:229-317 (and the term builders :740-1084)
//   nonlinear part  Discretization.py:319-365,1178-1459
//   rhs/jacobian    atomF = Fnl + L ;  atomJ = (Jnl + Fnl) + L      (:385-415)
//   boundaries      BoundaryConditions.py, as an op list applied in call order
//   apply           F = sum_{sz,sy,sx,d2} A * x_pad + frc             (:439-472)
//   CSR             Discretization.py:474-524 + CrsMatrix.compress
//
// d1 is a template parameter, so every atom entry that no term of that
// equation touches is a compile-time zero.  The file is compiled with
// -ffp-contract=off (no FMA, so products and sums round exactly as numpy's),
// -fno-signed-zeros and -ffinite-math-only: those two only let the compiler
// fold x+0 -> x and 0*x -> 0 on the constant-zero entries; nothing is
// reassociated, so every nonzero entry is computed by the same sequence of
// IEEE operations as the reference and the results agree bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace tf {

constexpr int MAX_BC_OPS = 32;
enum : int { BC_WALL = 0, BC_FORCE = 1 };

// One BoundaryConditions primitive.  WALL is no_slip (sign -1) or free_slip
// (sign +1) at face (axis, side); FORCE is _constant_forcing_<face> on
// variable ``var`` with forcing constant fc and atom constant ac
// (moving lids, fixed temperatures and heat/salinity fluxes reduce to it).
struct BcOp {
  int kind, axis, side, var;
  double sign, fc, ac;
};

struct Grid {
  int n[3];
  int periodic[3];
  const double* coord[3];  // device coordinate vectors, length n+3 each
};

struct Physics {
  double lin_scale;  // 1/(Re sqrt(Gr))
  double T_scale;    // 1/(Pr sqrt(Gr))
  double S_scale;    // 1/(Le Pr sqrt(Gr))
  double rbp_coef;   // Bi/(Bi+1)  (Rayleigh-Benard Perturbation only)
  int rbp;
  int nonlinear;     // 0 when Re == 0 and there is no T/S (Discretization.py:333-335)
  int nops;
  const BcOp* ops;   // device array, nops entries, applied in order
};

constexpr int SL(int sx, int sy, int sz) { return (sx * 3 + sy) * 3 + sz; }
// slot s along axis AX, centre along the others
template <int AX>
constexpr int SLA(int s) {
  return AX == 0 ? SL(s, 1, 1) : AX == 1 ? SL(1, s, 1) : SL(1, 1, s);
}
// slot s along AX, (t0, t1) along the two other axes in ascending order
template <int AX>
constexpr int SLT(int s, int t0, int t1) {
  return AX == 0 ? SL(s, t0, t1) : AX == 1 ? SL(t0, s, t1) : SL(t0, t1, s);
}
// slot sa along A, sb along B, centre along the third
constexpr int SL2(int A, int sa, int B, int sb) {
  return SL(A == 0 ? sa : B == 0 ? sb : 1, A == 1 ? sa : B == 1 ? sb : 1,
            A == 2 ? sa : B == 2 ? sb : 1);
}

// (r0, r1, r2) argument permutations of the reference's 1D helpers
// (e.g. v_yy calls _u_xx(atom, j, i, k, y, x, z)).
template <int A> struct Role;
template <> struct Role<0> { static constexpr int r0 = 0, r1 = 1, r2 = 2; };
template <> struct Role<1> { static constexpr int r0 = 1, r1 = 0, r2 = 2; };
template <> struct Role<2> { static constexpr int r0 = 2, r1 = 1, r2 = 0; };

template <int DOF>
struct Cell {
  int idx[3];
  int n[3];
  int periodic[3];
  const double* coord[3];
  const double* __restrict__ state;

  // coordinate vector entry idx+off with Python negative-index wrap
  __device__ __forceinline__ double X(int axis, int off) const {
    int t = idx[axis] + off;
    if (t < 0) t += n[axis] + 3;
    return coord[axis][t];
  }
  __device__ __forceinline__ double diff(int axis, int lo, int hi) const {
    return X(axis, hi) - X(axis, lo);
  }
  __device__ __forceinline__ double hdiff(int axis, int lo, int hi) const {
    return (X(axis, hi) - X(axis, lo)) / 2.0;
  }

  // Padded state (utils.create_padded_state_mtx, utils.py:62-133): component
  // c of the cell at offset (rx, ry, rz).  Ghosts are zero unless the axis is
  // periodic; on a non-periodic axis the wall-normal velocity of the last
  // interior cell is zero.
  __device__ __forceinline__ double P(int c, int rx, int ry, int rz) const {
    int p[3] = {idx[0] + 1 + rx, idx[1] + 1 + ry, idx[2] + 1 + rz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (p[a] == 0) {
        if (!periodic[a]) return 0.0;
        p[a] = n[a];
      } else if (p[a] == n[a] + 1) {
        if (!periodic[a]) return 0.0;
        p[a] = 1;
      }
    }
    if (c < 3 && !periodic[c < 3 ? c : 0] && p[c < 3 ? c : 0] == n[c < 3 ? c : 0]) return 0.0;
    size_t cell = ((size_t)(p[2] - 1) * n[1] + (p[1] - 1)) * n[0] + (p[0] - 1);
    return state[cell * DOF + c];
  }
};

// ---------------------------------------------------------------------------
// Linear part
// ---------------------------------------------------------------------------

template <int D1, int EQ, int VAR, int AX, int DOF>
__device__ __forceinline__ void add_line(double* acc, double a0, double a1, double a2) {
  if constexpr (D1 == EQ) {
    acc[VAR * 27 + SLA<AX>(0)] += a0;
    acc[VAR * 27 + SLA<AX>(1)] += a1;
    acc[VAR * 27 + SLA<AX>(2)] += a2;
  }
}

// _u_xx (Discretization.py:741-754)
template <int R, int DOF>
__device__ __forceinline__ void face_lap(const Cell<DOF>& c, double& a0, double& a1, double& a2) {
  constexpr int r0 = Role<R>::r0, r1 = Role<R>::r1, r2 = Role<R>::r2;
  double dx = c.diff(r0, -1, 0), dxp1 = c.diff(r0, 0, 1);
  double dy = c.diff(r1, -1, 0), dz = c.diff(r2, -1, 0);
  a0 = 1.0 / dx * dy * dz;
  a2 = 1.0 / dxp1 * dy * dz;
  a1 = -a0 - a2;
}

// _u_yy (Discretization.py:778-791), role given explicitly
template <int R, int DOF>
__device__ __forceinline__ void centre_lap_y(const Cell<DOF>& c, double& a0, double& a1, double& a2) {
  constexpr int r0 = Role<R>::r0, r1 = Role<R>::r1, r2 = Role<R>::r2;
  double dy = c.hdiff(r1, -2, 0), dyp1 = c.hdiff(r1, -1, 1);
  double dx = c.hdiff(r0, -1, 1), dz = c.diff(r2, -1, 0);
  a0 = 1.0 / dy * dx * dz;
  a2 = 1.0 / dyp1 * dx * dz;
  a1 = -a0 - a2;
}

// _u_zz (Discretization.py:815-828)
template <int R, int DOF>
__device__ __forceinline__ void centre_lap_z(const Cell<DOF>& c, double& a0, double& a1, double& a2) {
  constexpr int r0 = Role<R>::r0, r1 = Role<R>::r1, r2 = Role<R>::r2;
  double dz = c.hdiff(r2, -2, 0), dzp1 = c.hdiff(r2, -1, 1);
  double dx = c.hdiff(r0, -1, 1), dy = c.diff(r1, -1, 0);
  a0 = 1.0 / dz * dx * dy;
  a2 = 1.0 / dzp1 * dx * dy;
  a1 = -a0 - a2;
}

// _C_xx (Discretization.py:852-865)
template <int R, int DOF>
__device__ __forceinline__ void scalar_lap(const Cell<DOF>& c, double& a0, double& a1, double& a2) {
  constexpr int r0 = Role<R>::r0, r1 = Role<R>::r1, r2 = Role<R>::r2;
  double dx = c.hdiff(r0, -2, 0), dxp1 = c.hdiff(r0, -1, 1);
  double dy = c.diff(r1, -1, 0), dz = c.diff(r2, -1, 0);
  a0 = 1.0 / dx * dy * dz;
  a2 = 1.0 / dxp1 * dy * dz;
  a1 = -a0 - a2;
}

// dy*dz of _forward_u_x / _backward_u_x for role R
template <int R, int DOF>
__device__ __forceinline__ double face_area(const Cell<DOF>& c) {
  return c.diff(Role<R>::r1, -1, 0) * c.diff(Role<R>::r2, -1, 0);
}

// dx*dy*dz/2 of _forward_average_x / _backward_average_x for role R
template <int R, int DOF>
__device__ __forceinline__ double half_volume(const Cell<DOF>& c) {
  return c.hdiff(Role<R>::r0, -1, 1) * c.diff(Role<R>::r1, -1, 0) * c.diff(Role<R>::r2, -1, 0) / 2.0;
}

// Velocity Laplacian terms of component V along axis AX (V's own axis ->
// _u_xx form; the others -> _u_yy/_u_zz forms with V's role).
template <int D1, int V, int AX, int DOF>
__device__ __forceinline__ void vel_lap(const Cell<DOF>& c, double* lap) {
  if constexpr (D1 == V) {
    double a0, a1, a2;
    if constexpr (AX == V) {
      face_lap<V>(c, a0, a1, a2);
    } else if constexpr ((V == 0 && AX == 1) || (V == 1 && AX == 0) || (V == 2 && AX == 1)) {
      centre_lap_y<V>(c, a0, a1, a2);
    } else {
      centre_lap_z<V>(c, a0, a1, a2);
    }
    add_line<D1, V, V, AX, DOF>(lap, a0, a1, a2);
  }
}

template <int D1, int VAR, int AX, int DOF>
__device__ __forceinline__ void scal_lap(const Cell<DOF>& c, double* lap) {
  if constexpr (D1 == VAR) {
    double a0, a1, a2;
    scalar_lap<AX>(c, a0, a1, a2);
    add_line<D1, VAR, VAR, AX, DOF>(lap, a0, a1, a2);
  }
}

template <int DIM, int DOF, int BUOY, int D1>
__device__ __forceinline__ void linear_row(const Cell<DOF>& c, const Physics& ph, double* L) {
  constexpr int E = DOF * 27;
  double lap[E] = {};
  double grad[E] = {};
  double dv[E] = {};
  // (u_xx + u_yy [+ u_zz]) + (v_xx + v_yy [+ v_zz]) [+ w_xx + w_yy + w_zz]
  vel_lap<D1, 0, 0, DOF>(c, lap);
  vel_lap<D1, 0, 1, DOF>(c, lap);
  if constexpr (DIM == 3) vel_lap<D1, 0, 2, DOF>(c, lap);
  vel_lap<D1, 1, 0, DOF>(c, lap);
  vel_lap<D1, 1, 1, DOF>(c, lap);
  if constexpr (DIM == 3) {
    vel_lap<D1, 1, 2, DOF>(c, lap);
    vel_lap<D1, 2, 0, DOF>(c, lap);
    vel_lap<D1, 2, 1, DOF>(c, lap);
    vel_lap<D1, 2, 2, DOF>(c, lap);
  }
  // pressure gradient p_x + p_y [+ p_z]: forward difference in equation a
  {
    if constexpr (D1 == 0) { double a2 = face_area<0>(c); add_line<D1, 0, DIM, 0, DOF>(grad, 0.0, -a2, a2); }
    if constexpr (D1 == 1) { double a2 = face_area<1>(c); add_line<D1, 1, DIM, 1, DOF>(grad, 0.0, -a2, a2); }
    if constexpr (DIM == 3 && D1 == 2) { double a2 = face_area<2>(c); add_line<D1, 2, DIM, 2, DOF>(grad, 0.0, -a2, a2); }
  }
  // divergence u_x + v_y [+ w_z]: backward difference in the continuity row
  if constexpr (D1 == DIM) {
    double a;
    a = face_area<0>(c); add_line<D1, DIM, 0, 0, DOF>(dv, -a, a, 0.0);
    a = face_area<1>(c); add_line<D1, DIM, 1, 1, DOF>(dv, -a, a, 0.0);
    if constexpr (DIM == 3) { a = face_area<2>(c); add_line<D1, DIM, 2, 2, DOF>(dv, -a, a, 0.0); }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) L[e] = ph.lin_scale * lap[e] - grad[e] + dv[e];

  if constexpr (DOF > DIM + 1) {  // temperature
    constexpr int T = DIM + 1;
    double tl[E] = {};
    double fa[E] = {};
    scal_lap<D1, T, 0, DOF>(c, tl);
    scal_lap<D1, T, 1, DOF>(c, tl);
    if constexpr (DIM == 3) scal_lap<D1, T, 2, DOF>(c, tl);
    if constexpr (D1 == BUOY) { double a = half_volume<BUOY>(c); add_line<D1, BUOY, T, BUOY, DOF>(fa, 0.0, a, a); }
#pragma unroll
    for (int e = 0; e < E; ++e) L[e] = L[e] + ph.T_scale * tl[e];
#pragma unroll
    for (int e = 0; e < E; ++e) L[e] = L[e] + fa[e];
  }
  if constexpr (DOF > DIM + 2) {  // salinity
    constexpr int S = DIM + 2;
    double sl[E] = {};
    double fa[E] = {};
    scal_lap<D1, S, 0, DOF>(c, sl);
    scal_lap<D1, S, 1, DOF>(c, sl);
    if constexpr (DIM == 3) scal_lap<D1, S, 2, DOF>(c, sl);
    if constexpr (D1 == BUOY) { double a = half_volume<BUOY>(c); add_line<D1, BUOY, S, BUOY, DOF>(fa, 0.0, a, a); }
#pragma unroll
    for (int e = 0; e < E; ++e) L[e] = L[e] + ph.S_scale * sl[e];
#pragma unroll
    for (int e = 0; e < E; ++e) L[e] = L[e] - fa[e];
  }
  if constexpr (DOF > DIM + 1 && D1 == DIM + 1) {  // Rayleigh-Benard perturbation
    if (ph.rbp) {
      double ba[E] = {};
      double a = half_volume<BUOY>(c);
      add_line<D1, DIM + 1, BUOY, BUOY, DOF>(ba, a, a, 0.0);
#pragma unroll
      for (int e = 0; e < E; ++e) L[e] = L[e] + ph.rbp_coef * ba[e];
    }
  }
}

// ---------------------------------------------------------------------------
// Nonlinear part
// ---------------------------------------------------------------------------

template <int A>
__device__ __forceinline__ void unit(int& x, int& y, int& z, int s) {
  x = A == 0 ? s : 0;
  y = A == 1 ? s : 0;
  z = A == 2 ? s : 0;
}

template <int A, int DOF>
__device__ __forceinline__ double Pe(const Cell<DOF>& c, int comp, int s) {
  int x, y, z;
  unit<A>(x, y, z, s);
  return c.P(comp, x, y, z);
}

template <int A, int B, int DOF>
__device__ __forceinline__ double Pe2(const Cell<DOF>& c, int comp, int sa, int sb) {
  int x = (A == 0 ? sa : 0) + (B == 0 ? sb : 0);
  int y = (A == 1 ? sa : 0) + (B == 1 ? sb : 0);
  int z = (A == 2 ? sa : 0) + (B == 2 ? sb : 0);
  return c.P(comp, x, y, z);
}

template <int D1, int A, int DOF>
__device__ __forceinline__ void sub_line2(double* acc, int var, double v_lo, double v_hi) {
  // acc[var, slots 0:2 along A] -= v_lo ; acc[var, slots 1:3 along A] -= v_hi
  acc[var * 27 + SLA<A>(0)] -= v_lo;
  acc[var * 27 + SLA<A>(1)] -= v_lo;
  acc[var * 27 + SLA<A>(1)] -= v_hi;
  acc[var * 27 + SLA<A>(2)] -= v_hi;
}

// u_u_x / v_v_y / w_w_z (Discretization.py:1268-1279,1351-1362,1434-1445)
template <int D1, int A, int DOF>
__device__ __forceinline__ void self_adv(const Cell<DOF>& c, double* Fn, double* Jn) {
  if constexpr (D1 == A) {
    double area = face_area<A>(c);
    double a1 = -area, a2 = area;
    double p_lo = Pe<A>(c, A, -1), p_0 = Pe<A>(c, A, 0), p_hi = Pe<A>(c, A, 1);
    double avg_lo = 0.5 * p_lo + 0.5 * p_0;
    double avg_hi = 0.5 * p_0 + 0.5 * p_hi;
    double v1 = a1 * avg_lo * 1.0 / 2.0;
    double v2 = a2 * avg_hi * 1.0 / 2.0;
    sub_line2<D1, A, DOF>(Fn, A, v1, v2);
    sub_line2<D1, A, DOF>(Jn, A, v1, v2);
  }
}

// Transport of velocity B by velocity A: u_v_x (0,1), v_u_y (1,0), u_w_x (0,2),
// v_w_y (1,2), w_u_z (2,0), w_v_z (2,1)  (Discretization.py:1281-1311,1335-1349,1364-1378,1402-1432)
template <int D1, int A, int B, int DOF>
__device__ __forceinline__ void cross_adv(const Cell<DOF>& c, double* Fn, double* Jn) {
  if constexpr (D1 == B) {
    constexpr int third = 3 - A - B;
    double area = c.hdiff(B, -1, 1) * c.diff(third, -1, 0);
    double a1 = area, a0 = -area;
    // _weighted_average along B at this cell (Discretization.py:1212-1221)
    double dxb = (c.X(B, 1) - c.X(B, -1)) / 2.0;
    double w0 = 1.0 / 2.0 * (c.X(B, 0) - c.X(B, -1)) / dxb;
    double w1 = 1.0 / 2.0 * (c.X(B, 1) - c.X(B, 0)) / dxb;
    // velocity A weighted-averaged to the B location, on the lower and upper A face
    double wav_lo = w0 * Pe2<A, B>(c, A, -1, 0) + w1 * Pe2<A, B>(c, A, -1, 1);
    double wav_hi = w0 * Pe2<A, B>(c, A, 0, 0) + w1 * Pe2<A, B>(c, A, 0, 1);
    double f0 = a0 * wav_lo * 1.0 / 2.0;
    double f1 = a1 * wav_hi * 1.0 / 2.0;
    sub_line2<D1, A, DOF>(Fn, B, f0, f1);
    // velocity B averaged along A
    double pb_lo = Pe<A>(c, B, -1), pb_0 = Pe<A>(c, B, 0), pb_hi = Pe<A>(c, B, 1);
    double pav_lo = 0.5 * pb_lo + 0.5 * pb_0;
    double pav_hi = 0.5 * pb_0 + 0.5 * pb_hi;
    double t0 = a0 * pav_lo, t1 = a1 * pav_hi;
    Jn[A * 27 + SL2(A, 0, B, 1)] -= t0 * w0;
    Jn[A * 27 + SL2(A, 0, B, 2)] -= t0 * w1;
    Jn[A * 27 + SL2(A, 1, B, 1)] -= t1 * w0;
    Jn[A * 27 + SL2(A, 1, B, 2)] -= t1 * w1;
  }
}

// u_C_x / v_C_y / w_C_z (Discretization.py:1313-1325,1380-1392,1447-1459)
template <int D1, int A, int VAR, int DOF>
__device__ __forceinline__ void scalar_adv(const Cell<DOF>& c, double* Fn, double* Jn) {
  if constexpr (D1 == VAR) {
    double a1 = face_area<A>(c);
    double a0 = -a1;
    double v0 = a0 * Pe<A>(c, A, -1) * 1.0 / 2.0;
    double v1 = a1 * Pe<A>(c, A, 0) * 1.0 / 2.0;
    sub_line2<D1, A, DOF>(Fn, VAR, v0, v1);
    double q_lo = Pe<A>(c, VAR, -1), q_0 = Pe<A>(c, VAR, 0), q_hi = Pe<A>(c, VAR, 1);
    double avg_lo = 0.5 * q_lo + 0.5 * q_0;
    double avg_hi = 0.5 * q_0 + 0.5 * q_hi;
    Jn[A * 27 + SLA<A>(0)] -= a0 * avg_lo;
    Jn[A * 27 + SLA<A>(1)] -= a1 * avg_hi;
  }
}

template <int DIM, int DOF, int D1>
__device__ __forceinline__ void nonlinear_row(const Cell<DOF>& c, double* Fn, double* Jn) {
  self_adv<D1, 0, DOF>(c, Fn, Jn);
  cross_adv<D1, 0, 1, DOF>(c, Fn, Jn);
  cross_adv<D1, 1, 0, DOF>(c, Fn, Jn);
  self_adv<D1, 1, DOF>(c, Fn, Jn);
  if constexpr (DIM > 2) {
    cross_adv<D1, 0, 2, DOF>(c, Fn, Jn);
    cross_adv<D1, 1, 2, DOF>(c, Fn, Jn);
    cross_adv<D1, 2, 0, DOF>(c, Fn, Jn);
    cross_adv<D1, 2, 1, DOF>(c, Fn, Jn);
    self_adv<D1, 2, DOF>(c, Fn, Jn);
  }
  if constexpr (DOF > DIM + 1) {
    scalar_adv<D1, 0, DIM + 1, DOF>(c, Fn, Jn);
    scalar_adv<D1, 1, DIM + 1, DOF>(c, Fn, Jn);
    if constexpr (DIM > 2) scalar_adv<D1, 2, DIM + 1, DOF>(c, Fn, Jn);
  }
  if constexpr (DOF > DIM + 2) {
    scalar_adv<D1, 0, DIM + 2, DOF>(c, Fn, Jn);
    scalar_adv<D1, 1, DIM + 2, DOF>(c, Fn, Jn);
    if constexpr (DIM > 2) scalar_adv<D1, 2, DIM + 2, DOF>(c, Fn, Jn);
  }
}

// ---------------------------------------------------------------------------
// Boundary conditions (BoundaryConditions.py:55-233,471-544)
// ---------------------------------------------------------------------------

template <int D1, int AX, int SIDE, int DOF>
__device__ __forceinline__ void bc_wall(const Cell<DOF>& c, double* A, double& frc, double sign) {
  const int i = c.idx[AX], n = c.n[AX];
  if constexpr (SIDE > 0) {
    if (i == n - 1) {
#pragma unroll
      for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
          for (int t1 = 0; t1 < 3; ++t1)
            A[d2 * 27 + SLT<AX>(1, t0, t1)] = A[d2 * 27 + SLT<AX>(1, t0, t1)] + sign * A[d2 * 27 + SLT<AX>(2, t0, t1)];
#pragma unroll
      for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
        for (int t1 = 0; t1 < 3; ++t1) A[AX * 27 + SLT<AX>(1, t0, t1)] = 0.0;
      if constexpr (D1 == AX) {
#pragma unroll
        for (int e = 0; e < DOF * 27; ++e) A[e] = 0.0;
      }
#pragma unroll
      for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
          for (int t1 = 0; t1 < 3; ++t1) A[d2 * 27 + SLT<AX>(2, t0, t1)] = 0.0;
      if constexpr (D1 == AX) A[AX * 27 + SL(1, 1, 1)] = -1.0;
      // the reference clears component 1 on the top face (BoundaryConditions.py:131,221)
      if constexpr (D1 == (AX < 2 ? AX : 1)) frc = 0.0;
    }
    if constexpr (D1 == AX) {
      if (i == n - 2) {
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
          for (int t1 = 0; t1 < 3; ++t1) A[AX * 27 + SLT<AX>(2, t0, t1)] = 0.0;
      }
    }
  } else {
    if (i == 0) {
#pragma unroll
      for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
        for (int t1 = 0; t1 < 3; ++t1) A[AX * 27 + SLT<AX>(0, t0, t1)] = 0.0;
#pragma unroll
      for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
          for (int t1 = 0; t1 < 3; ++t1)
            A[d2 * 27 + SLT<AX>(1, t0, t1)] = A[d2 * 27 + SLT<AX>(1, t0, t1)] + sign * A[d2 * 27 + SLT<AX>(0, t0, t1)];
#pragma unroll
      for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
          for (int t1 = 0; t1 < 3; ++t1) A[d2 * 27 + SLT<AX>(0, t0, t1)] = 0.0;
    }
  }
}

template <int D1, int AX, int SIDE, int VAR, int DOF>
__device__ __forceinline__ void bc_force(const Cell<DOF>& c, double* A, double& frc, double fc, double ac) {
  const int i = c.idx[AX], n = c.n[AX];
  constexpr int out = SIDE > 0 ? 2 : 0;
  if (i == (SIDE > 0 ? n - 1 : 0)) {
    if constexpr (D1 == VAR) {
      // forcing: tangential slots, higher axis outermost (BoundaryConditions.py:471-478)
      double acc = 0.0;
#pragma unroll
      for (int t1 = 0; t1 < 3; ++t1)
#pragma unroll
        for (int t0 = 0; t0 < 3; ++t0) acc = acc + A[VAR * 27 + SLT<AX>(out, t0, t1)] * fc;
      frc = frc + acc;
    }
#pragma unroll
    for (int t0 = 0; t0 < 3; ++t0)
#pragma unroll
      for (int t1 = 0; t1 < 3; ++t1) {
        A[VAR * 27 + SLT<AX>(1, t0, t1)] = A[VAR * 27 + SLT<AX>(1, t0, t1)] + ac * A[VAR * 27 + SLT<AX>(out, t0, t1)];
        A[VAR * 27 + SLT<AX>(out, t0, t1)] = 0.0;
      }
  }
}

template <int D1, int AX, int SIDE, int DOF>
__device__ __forceinline__ void bc_force_var(const Cell<DOF>& c, double* A, double& frc, const BcOp& op) {
  switch (op.var) {
    case 0: bc_force<D1, AX, SIDE, 0, DOF>(c, A, frc, op.fc, op.ac); break;
    case 1: bc_force<D1, AX, SIDE, 1, DOF>(c, A, frc, op.fc, op.ac); break;
    case 2: bc_force<D1, AX, SIDE, 2, DOF>(c, A, frc, op.fc, op.ac); break;
    case 3: if constexpr (DOF > 3) bc_force<D1, AX, SIDE, 3, DOF>(c, A, frc, op.fc, op.ac); break;
    case 4: if constexpr (DOF > 4) bc_force<D1, AX, SIDE, 4, DOF>(c, A, frc, op.fc, op.ac); break;
    default: break;
  }
}

template <int D1, int DOF>
__device__ __forceinline__ void apply_bcs(const Cell<DOF>& c, const Physics& ph, double* A, double& frc) {
  for (int o = 0; o < ph.nops; ++o) {
    const BcOp op = ph.ops[o];
    const int face = op.axis * 2 + (op.side > 0 ? 1 : 0);
    if (op.kind == BC_WALL) {
      switch (face) {
        case 0: bc_wall<D1, 0, -1, DOF>(c, A, frc, op.sign); break;
        case 1: bc_wall<D1, 0, 1, DOF>(c, A, frc, op.sign); break;
        case 2: bc_wall<D1, 1, -1, DOF>(c, A, frc, op.sign); break;
        case 3: bc_wall<D1, 1, 1, DOF>(c, A, frc, op.sign); break;
        case 4: bc_wall<D1, 2, -1, DOF>(c, A, frc, op.sign); break;
        case 5: bc_wall<D1, 2, 1, DOF>(c, A, frc, op.sign); break;
        default: break;
      }
    } else {
      switch (face) {
        case 0: bc_force_var<D1, 0, -1, DOF>(c, A, frc, op); break;
        case 1: bc_force_var<D1, 0, 1, DOF>(c, A, frc, op); break;
        case 2: bc_force_var<D1, 1, -1, DOF>(c, A, frc, op); break;
        case 3: bc_force_var<D1, 1, 1, DOF>(c, A, frc, op); break;
        case 4: bc_force_var<D1, 2, -1, DOF>(c, A, frc, op); break;
        case 5: bc_force_var<D1, 2, 1, DOF>(c, A, frc, op); break;
        default: break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Row assembly
// ---------------------------------------------------------------------------

enum : int { ATOM_F = 0, ATOM_J = 1 };

// Final atom of row D1 after the boundary conditions; frc is the row's forcing.
template <int DIM, int DOF, int BUOY, int D1, int KIND>
__device__ __forceinline__ void build_row(const Cell<DOF>& c, const Physics& ph, double* A, double& frc) {
  constexpr int E = DOF * 27;
  double L[E];
  linear_row<DIM, DOF, BUOY, D1>(c, ph, L);
  double Fn[E] = {};
  double Jn[E] = {};
  if (ph.nonlinear) nonlinear_row<DIM, DOF, D1>(c, Fn, Jn);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if constexpr (KIND == ATOM_F)
      A[e] = Fn[e] + L[e];
    else
      A[e] = (Jn[e] + Fn[e]) + L[e];
  }
  frc = 0.0;
  apply_bcs<D1, DOF>(c, ph, A, frc);
}

// F row: sum over (sz, sy, sx, d2) in the reference order, then + frc.
template <int DOF>
__device__ __forceinline__ double apply_row(const Cell<DOF>& c, const double* A, double frc) {
  double out = 0.0;
#pragma unroll
  for (int sz = 0; sz < 3; ++sz)
#pragma unroll
    for (int sy = 0; sy < 3; ++sy)
#pragma unroll
      for (int sx = 0; sx < 3; ++sx)
#pragma unroll
        for (int d2 = 0; d2 < DOF; ++d2) out = out + A[d2 * 27 + SL(sx, sy, sz)] * c.P(d2, sx - 1, sy - 1, sz - 1);
  return out + frc;
}

constexpr double DROP = 1e-14;

// CSR emission of one row (assemble_jacobian + CrsMatrix.compress).
// Columns wrap modulo each axis length.  Entries are visited in the
// reference's appearance order (d2, sx, sy, sz) for duplicate summation.
// Fast path: if the kept candidates already have strictly increasing
// columns in (sz, sy, sx, d2) order (every interior cell), they are emitted
// in that order.  Otherwise a selection pass emits the distinct columns in
// ascending order, summing duplicates in appearance order.
template <int DOF, bool WRITE>
__device__ __forceinline__ int emit_row(const Cell<DOF>& c, const double* A, int* __restrict__ col_out,
                                        double* __restrict__ val_out) {
  int cx[3], cy[3], cz[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    int t = c.idx[0] + s - 1;
    cx[s] = t < 0 ? t + c.n[0] : (t >= c.n[0] ? t - c.n[0] : t);
    t = c.idx[1] + s - 1;
    cy[s] = t < 0 ? t + c.n[1] : (t >= c.n[1] ? t - c.n[1] : t);
    t = c.idx[2] + s - 1;
    cz[s] = t < 0 ? t + c.n[2] : (t >= c.n[2] ? t - c.n[2] : t);
  }
  const int sx_stride = DOF, sy_stride = c.n[0] * DOF, sz_stride = c.n[0] * c.n[1] * DOF;
  auto colof = [&](int d2, int sx, int sy, int sz) {
    return cz[sz] * sz_stride + cy[sy] * sy_stride + cx[sx] * sx_stride + d2;
  };

  bool fast = true;
  int last = -1;
#pragma unroll
  for (int sz = 0; sz < 3; ++sz)
#pragma unroll
    for (int sy = 0; sy < 3; ++sy)
#pragma unroll
      for (int sx = 0; sx < 3; ++sx)
#pragma unroll
        for (int d2 = 0; d2 < DOF; ++d2) {
          double v = A[d2 * 27 + SL(sx, sy, sz)];
          if (fabs(v) > DROP) {
            int cc = colof(d2, sx, sy, sz);
            fast = fast && cc > last;
            last = cc;
          }
        }

  int count = 0;
  if (fast) {
#pragma unroll
    for (int sz = 0; sz < 3; ++sz)
#pragma unroll
      for (int sy = 0; sy < 3; ++sy)
#pragma unroll
        for (int sx = 0; sx < 3; ++sx)
#pragma unroll
          for (int d2 = 0; d2 < DOF; ++d2) {
            double v = A[d2 * 27 + SL(sx, sy, sz)];
            if (fabs(v) > DROP) {
              if constexpr (WRITE) {
                col_out[count] = colof(d2, sx, sy, sz);
                val_out[count] = v;
              }
              ++count;
            }
          }
    return count;
  }

  int prev = -1;
  for (int guard = 0; guard < DOF * 27; ++guard) {
    int cur = INT_MAX;
#pragma unroll
    for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
      for (int s = 0; s < 27; ++s) {
        double v = A[d2 * 27 + s];
        if (fabs(v) > DROP) {
          int cc = colof(d2, s / 9, (s / 3) % 3, s % 3);
          if (cc > prev && cc < cur) cur = cc;
        }
      }
    if (cur == INT_MAX) break;
    double acc = 0.0;
#pragma unroll
    for (int d2 = 0; d2 < DOF; ++d2)
#pragma unroll
      for (int s = 0; s < 27; ++s) {
        double v = A[d2 * 27 + s];
        if (fabs(v) > DROP && colof(d2, s / 9, (s / 3) % 3, s % 3) == cur) acc = acc + v;
      }
    if (fabs(acc) > DROP) {
      if constexpr (WRITE) {
        col_out[count] = cur;
        val_out[count] = acc;
      }
      ++count;
    }
    prev = cur;
  }
  return count;
}

}  // namespace tf
