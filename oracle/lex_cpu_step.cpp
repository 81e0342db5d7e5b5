// CPU restatement of the forward WS-RK3 step in C++ with OpenMP: the all-core CPU
// baseline of the compressible benchmark.  TEST INFRASTRUCTURE: only tests/ and the
// cpu_baseline leg of bench.py load it; the product path never does.
//
// Scope: solver_opt 2 (explicit acoustic substeps), turb_opt 1 (Smagorinsky frozen from
// stage 0), no Coriolis, damping, radiation or surface fluxes, theta, qv and n_passive
// passive scalars.  Each routine restates the NumPy oracle function of the same name in
// oracle/lex_oracle.py, which cites the reference lines it follows:
//   padding3_array / padding3u_array / padding3v_array / set_w_bc
//                                   one_step_integration.py:950-985, boundary_conditions.py:104-109
//   weno5 / weno3 / vertical faces  advection.py:305-516
//   get_divergence(2)               advection.py:200-236
//   advection_scalar                advection.py:9-34
//   advection_u / _v / _w           advection.py:37-149, :561-1281
//   advection_pi                    advection.py:152-197
//   calculate_buoyancy, theta_rho   pressure_equations.py:123-172
//   pressure_gradient_force         pressure_gradient_coriolis.py:7-37
//   compute_deformation / _nm / _k / compute_smag   turbulence.py:9-111, :298-446, :474-504
//   sound_rk_step                   one_step_integration.py:519-556, :638-675, :757-794
//   rk_sub_step / ws_rk3_big_step   one_step_integration.py:16-447, one_sprint_functions.py:10-30
// Expressions keep the oracle's evaluation order (the file is built with
// -ffp-contract=off), so the FP64 instantiation reproduces the FP64 oracle to rounding
// of the library log / sqrt / cbrt, and the FP32 one is another FP32 evaluation of the
// same arithmetic.  Arrays use the reference's padded C-order layout (x, y, z), ghost
// widths (3, 3, 1); the base state is taken as z-columns (it is horizontally uniform).
#include <omp.h>

#include <cmath>
#include <cstddef>
#include <cstring>
#include <utility>
#include <vector>

namespace {

constexpr double kRd = 287.04, kRv = 461.5, kCp = 1004.64, kG = 9.80616;
constexpr double kCv = kCp - kRd, kReps = kRv / kRd, kRepsm1 = kRv / kRd - 1.0;

// storage left uninitialised: every array is written in full by the pass that makes it
template <class T>
struct Buf {
  T* p = nullptr;
  size_t n = 0;
  Buf() = default;
  explicit Buf(size_t m) : p(m ? new T[m] : nullptr), n(m) {}
  Buf(const Buf& o) : Buf(o.n) { if (n) std::memcpy(p, o.p, n * sizeof(T)); }
  Buf(Buf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  Buf& operator=(Buf o) noexcept { std::swap(p, o.p); std::swap(n, o.n); return *this; }
  ~Buf() { delete[] p; }
  T* data() { return p; }
  const T* data() const { return p; }
  size_t size() const { return n; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
};

template <class T>
struct A3 {   // dense 3-D array, C order
  int n0 = 0, n1 = 0, n2 = 0;
  Buf<T> d;
  A3() = default;
  A3(int a, int b, int c) : n0(a), n1(b), n2(c), d((size_t)a * b * c) {}
  T& operator()(int i, int j, int k) { return d[((size_t)i * n1 + j) * n2 + k]; }
  T operator()(int i, int j, int k) const { return d[((size_t)i * n1 + j) * n2 + k]; }
};

template <class F>
void par2(int n0, int n1, F f) {   // f(i, j) over a 2-D index range, all cores
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < n0; ++i)
    for (int j = 0; j < n1; ++j) f(i, j);
}

inline int wrap(int i, int n) { return ((i % n) + n) % n; }

template <class T>
struct Ctx {
  int nx, ny, nz;
  T dx, dy, dz;
  double dxyz;                         // dx dy dz in double (compute_k's grid scale)
  std::vector<T> rho, th0, qv0, pi0;   // base columns, padded z (nz + 2)
};

// ---- halos -------------------------------------------------------------------
// interior (nx, ny, nz) -> padded scalar (nx+6, ny+6, nz+2)
template <class T>
A3<T> pad_c(const Ctx<T>& c, const A3<T>& in) {
  A3<T> o(c.nx + 6, c.ny + 6, c.nz + 2);
  par2(o.n0, o.n1, [&](int I, int J) {
    const int i = wrap(I - 3, c.nx), j = wrap(J - 3, c.ny);
    for (int K = 0; K < c.nz + 2; ++K) o(I, J, K) = in(i, j, K == 0 ? 0 : (K == c.nz + 1 ? c.nz - 1 : K - 1));
  });
  return o;
}
// interior (nx+1, ny, nz) -> padded u (nx+7, ny+6, nz+2): x ghosts skip the duplicate face
template <class T>
A3<T> pad_u(const Ctx<T>& c, const A3<T>& in) {
  A3<T> o(c.nx + 7, c.ny + 6, c.nz + 2);
  par2(o.n0, o.n1, [&](int I, int J) {
    int i = I - 3;
    if (i < 0) i += c.nx;
    if (i > c.nx) i -= c.nx;
    const int j = wrap(J - 3, c.ny);
    for (int K = 0; K < c.nz + 2; ++K) o(I, J, K) = in(i, j, K == 0 ? 0 : (K == c.nz + 1 ? c.nz - 1 : K - 1));
  });
  return o;
}
template <class T>
A3<T> pad_v(const Ctx<T>& c, const A3<T>& in) {
  A3<T> o(c.nx + 6, c.ny + 7, c.nz + 2);
  par2(o.n0, o.n1, [&](int I, int J) {
    const int i = wrap(I - 3, c.nx);
    int j = J - 3;
    if (j < 0) j += c.ny;
    if (j > c.ny) j -= c.ny;
    for (int K = 0; K < c.nz + 2; ++K) o(I, J, K) = in(i, j, K == 0 ? 0 : (K == c.nz + 1 ? c.nz - 1 : K - 1));
  });
  return o;
}
// interior (nx, ny, nz+1) -> padded w (nx+6, ny+6, nz+3) with set_w_bc
template <class T>
A3<T> pad_w(const Ctx<T>& c, const A3<T>& in) {
  A3<T> o(c.nx + 6, c.ny + 6, c.nz + 3);
  par2(o.n0, o.n1, [&](int I, int J) {
    const int i = wrap(I - 3, c.nx), j = wrap(J - 3, c.ny);
    for (int K = 0; K < c.nz + 3; ++K) {
      const bool zero = K <= 1 || K >= c.nz + 1;
      o(I, J, K) = zero ? T(0) : in(i, j, K - 1);
    }
  });
  return o;
}

// ---- WENO ----------------------------------------------------------------------
template <class T>
inline T weno5(T a, T b, T c, T d, T e, T weps) {
  const T k1312 = T(13.0 / 12.0), q = T(0.25);
  const T t1 = a - T(2.0) * b + c, t2 = a - T(4.0) * b + T(3.0) * c;
  const T b1 = k1312 * (t1 * t1) + q * (t2 * t2);
  const T t3 = b - T(2.0) * c + d, t4 = b - d;
  const T b2 = k1312 * (t3 * t3) + q * (t4 * t4);
  const T t5 = c - T(2.0) * d + e, t6 = T(3.0) * c - T(4.0) * d + e;
  const T b3 = k1312 * (t5 * t5) + q * (t6 * t6);
  const T tau = std::fabs(b1 - b3);
  const T cap = T(1.0e30);
  const T r1 = std::fmin(cap, tau / (b1 + weps)), r2 = std::fmin(cap, tau / (b2 + weps)),
          r3 = std::fmin(cap, tau / (b3 + weps));
  const T w1 = T(0.1) * (T(1.0) + r1 * r1), w2 = T(0.6) * (T(1.0) + r2 * r2), w3 = T(0.3) * (T(1.0) + r3 * r3);
  const T q1 = T(2.0 / 6.0) * a + T(-7.0 / 6.0) * b + T(11.0 / 6.0) * c;
  const T q2 = T(-1.0 / 6.0) * b + T(5.0 / 6.0) * c + T(2.0 / 6.0) * d;
  const T q3 = T(2.0 / 6.0) * c + T(5.0 / 6.0) * d + T(-1.0 / 6.0) * e;
  return (w1 * q1 + w2 * q2 + w3 * q3) / (w1 + w2 + w3);
}
template <class T>
inline T weno3(T a, T b, T c, T weps) {
  const T b1 = (a - b) * (a - b), b2 = (b - c) * (b - c);
  const T tau = std::fabs(b1 - b2);
  const T r1 = std::fmin(T(1.0e30), tau / (b1 + weps)), r2 = std::fmin(T(1.0e30), tau / (b2 + weps));
  const T w1 = T(1.0 / 3.0) * (T(1.0) + r1 * r1), w2 = T(2.0 / 3.0) * (T(1.0) + r2 * r2);
  return (w1 * (T(-1.0 / 2.0) * a + T(3.0 / 2.0) * b) + w2 * (T(1.0 / 2.0) * b + T(1.0 / 2.0) * c)) / (w1 + w2);
}
// The oracle evaluates both upwind reconstructions of a face and selects one by the sign
// of the mass flux (lax.select); only the selected one is evaluated here (the same value).
// Horizontal face between s[m-1] and s[m] from s[m-3..m+2]: mass x face value.
template <class T>
inline T hflux(T mass, T sm3, T sm2, T sm1, T s0, T sp1, T sp2, T weps) {
  return mass >= T(0) ? mass * weno5(sm3, sm2, sm1, s0, sp1, weps) : mass * weno5(sp2, sp1, s0, sm1, sm2, weps);
}
// the vertical table on a padded column p[0..L-1] at face f (between p[f] and p[f+1])
template <class T>
inline T vflux(T mass, const T* p, int L, int f, T weps) {
  T r = T(0);
  if (mass >= T(0)) {
    if (f == 1) r = T(0.5) * (p[1] + p[2]);
    else if (f == 2) r = weno3(p[1], p[2], p[3], weps);
    else if (f == L - 3) r = weno3(p[L - 4], p[L - 3], p[L - 2], weps);
    else if (f >= 3 && f <= L - 4) r = weno5(p[f - 2], p[f - 1], p[f], p[f + 1], p[f + 2], weps);
  } else {
    if (f == 1) r = weno3(p[3], p[2], p[1], weps);
    else if (f == L - 4) r = weno3(p[L - 2], p[L - 3], p[L - 4], weps);
    else if (f == L - 3) r = T(0.5) * (p[L - 2] + p[L - 3]);
    else if (f >= 2 && f <= L - 5) r = weno5(p[f + 3], p[f + 2], p[f + 1], p[f], p[f - 1], weps);
  }
  return mass * r;
}

// ---- divergences -----------------------------------------------------------------
// div(rho0 u) on the whole padded scalar array (get_divergence)
template <class T>
A3<T> divergence(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w) {
  A3<T> o(c.nx + 6, c.ny + 6, c.nz + 2);
  par2(o.n0, o.n1, [&](int I, int J) {
    for (int K = 0; K < c.nz + 2; ++K) {
      const T r = c.rho[K];   // 0.5 (r + r) of a horizontally uniform column
      const T rzl = (K == 0) ? T(0) : T(0.5) * (c.rho[K - 1] + c.rho[K]);
      const T rzh = (K == c.nz + 1) ? T(0) : T(0.5) * (c.rho[K] + c.rho[K + 1]);
      o(I, J, K) = (r * u(I + 1, J, K) - r * u(I, J, K)) / c.dx + (r * v(I, J + 1, K) - r * v(I, J, K)) / c.dy +
                   (rzh * w(I, J, K + 1) - rzl * w(I, J, K)) / c.dz;
    }
  });
  return o;
}
template <class T>
A3<T> divergence2(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w) {
  A3<T> o(c.nx + 6, c.ny + 6, c.nz + 2);
  par2(o.n0, o.n1, [&](int I, int J) {
    for (int K = 0; K < c.nz + 2; ++K)
      o(I, J, K) = (u(I + 1, J, K) - u(I, J, K)) / c.dx + (v(I, J + 1, K) - v(I, J, K)) / c.dy +
                   (w(I, J, K + 1) - w(I, J, K)) / c.dz;
  });
  return o;
}

// ---- scalar advection (advection_scalar, zero surface flux) ----------------------
template <class T>
A3<T> advection_scalar(const Ctx<T>& c, const A3<T>& s, const A3<T>& u, const A3<T>& v, const A3<T>& w, T weps,
                       const A3<T>& fdiv) {
  const int nx = c.nx, ny = c.ny, nz = c.nz, L = nz + 2;
  A3<T> fx(nx + 1, ny, nz), fy(nx, ny + 1, nz), fz(nx, ny, nz + 1), out(nx, ny, nz);
  par2(nx + 1, ny, [&](int g, int j) {
    for (int k = 0; k < nz; ++k) {
      const int J = j + 3, K = k + 1;
      const T m = T(0.5) * (c.rho[K] + c.rho[K]) * u(g + 3, J, K);
      fx(g, j, k) = hflux(m, s(g, J, K), s(g + 1, J, K), s(g + 2, J, K), s(g + 3, J, K), s(g + 4, J, K), s(g + 5, J, K), weps);
    }
  });
  par2(nx, ny + 1, [&](int i, int h) {
    for (int k = 0; k < nz; ++k) {
      const int I = i + 3, K = k + 1;
      const T m = T(0.5) * (c.rho[K] + c.rho[K]) * v(I, h + 3, K);
      fy(i, h, k) = hflux(m, s(I, h, K), s(I, h + 1, K), s(I, h + 2, K), s(I, h + 3, K), s(I, h + 4, K), s(I, h + 5, K), weps);
    }
  });
  par2(nx, ny, [&](int i, int j) {
    const T* p = &s.d[((size_t)(i + 3) * s.n1 + (j + 3)) * s.n2];
    for (int f = 0; f <= nz; ++f) {
      const T m = T(0.5) * (c.rho[f] + c.rho[f + 1]) * w(i + 3, j + 3, f + 1);
      fz(i, j, f) = (f == 0) ? T(0) : vflux(m, p, L, f, weps);
    }
  });
  par2(nx, ny, [&](int i, int j) {
    for (int k = 0; k < nz; ++k) {
      const T conv = -((fx(i + 1, j, k) - fx(i, j, k)) / c.dx + (fy(i, j + 1, k) - fy(i, j, k)) / c.dy +
                       (fz(i, j, k + 1) - fz(i, j, k)) / c.dz);
      out(i, j, k) = (conv + s(i + 3, j + 3, k + 1) * fdiv(i + 3, j + 3, k + 1)) / c.rho[k + 1];
    }
  });
  return out;
}

// ---- momentum advection ------------------------------------------------------------
template <class T>
A3<T> advection_u(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w, T weps, const A3<T>& fd) {
  const int nx = c.nx, ny = c.ny, nz = c.nz, L = nz + 2;
  A3<T> fz(nx + 1, ny, nz + 1), fx(nx + 2, ny, nz), fy(nx + 1, ny + 1, nz), out(nx + 1, ny, nz);
  par2(nx + 1, ny, [&](int iu, int j) {   // vertical_flux_u2 at u face P = iu + 3
    const int P = iu + 3, J = j + 3;
    const T* p = &u.d[((size_t)P * u.n1 + J) * u.n2];
    for (int f = 0; f <= nz; ++f) {
      const T r8 = T(0.5) * (c.rho[f] + c.rho[f + 1]);
      const T m0 = r8 * w(P - 1, J, f + 1), m1 = r8 * w(P, J, f + 1);
      fz(iu, j, f) = (f == 0) ? T(0) : vflux(T(0.5) * (m0 + m1), p, L, f, weps);
    }
  });
  par2(nx + 2, ny, [&](int g, int j) {   // x flux at scalar cell g + 2
    for (int k = 0; k < nz; ++k) {
      const int J = j + 3, K = k + 1;
      const T m = T(0.5) * (u(g + 2, J, K) + u(g + 3, J, K)) * c.rho[K];
      fx(g, j, k) = hflux(m, u(g, J, K), u(g + 1, J, K), u(g + 2, J, K), u(g + 3, J, K), u(g + 4, J, K), u(g + 5, J, K), weps);
    }
  });
  par2(nx + 1, ny + 1, [&](int iu, int h) {   // y flux at the corner (u face P, v face h + 3)
    const int P = iu + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T r = T(0.5) * (c.rho[K] + c.rho[K]);
      const T m0 = r * v(P - 1, h + 3, K), m1 = r * v(P, h + 3, K);
      fy(iu, h, k) = hflux(T(0.5) * (m0 + m1), u(P, h, K), u(P, h + 1, K), u(P, h + 2, K), u(P, h + 3, K), u(P, h + 4, K), u(P, h + 5, K), weps);
    }
  });
  par2(nx + 1, ny, [&](int iu, int j) {
    for (int k = 0; k < nz; ++k) {
      const int P = iu + 3, J = j + 3, K = k + 1;
      const T conv = -((fx(iu + 1, j, k) - fx(iu, j, k)) / c.dx + (fy(iu, j + 1, k) - fy(iu, j, k)) / c.dy +
                       (fz(iu, j, k + 1) - fz(iu, j, k)) / c.dz);
      const T div8u = T(0.5) * (fd(P - 1, J, K) + fd(P, J, K));
      const T rho8u = T(0.5) * (c.rho[K] + c.rho[K]);
      out(iu, j, k) = (conv + u(P, J, K) * div8u) / rho8u;
    }
  });
  return out;
}
template <class T>
A3<T> advection_v(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w, T weps, const A3<T>& fd) {
  const int nx = c.nx, ny = c.ny, nz = c.nz, L = nz + 2;
  A3<T> fz(nx, ny + 1, nz + 1), fx(nx + 1, ny + 1, nz), fy(nx, ny + 2, nz), out(nx, ny + 1, nz);
  par2(nx, ny + 1, [&](int i, int iv) {
    const int I = i + 3, Q = iv + 3;
    const T* p = &v.d[((size_t)I * v.n1 + Q) * v.n2];
    for (int f = 0; f <= nz; ++f) {
      const T r8 = T(0.5) * (c.rho[f] + c.rho[f + 1]);
      const T m0 = r8 * w(I, Q - 1, f + 1), m1 = r8 * w(I, Q, f + 1);
      fz(i, iv, f) = (f == 0) ? T(0) : vflux(T(0.5) * (m0 + m1), p, L, f, weps);
    }
  });
  par2(nx + 1, ny + 1, [&](int g, int iv) {   // x flux at the corner (u face g + 3, v face Q)
    const int Q = iv + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T r = T(0.5) * (c.rho[K] + c.rho[K]);
      const T m0 = r * u(g + 3, Q - 1, K), m1 = r * u(g + 3, Q, K);
      fx(g, iv, k) = hflux(T(0.5) * (m0 + m1), v(g, Q, K), v(g + 1, Q, K), v(g + 2, Q, K), v(g + 3, Q, K), v(g + 4, Q, K), v(g + 5, Q, K), weps);
    }
  });
  par2(nx, ny + 2, [&](int i, int h) {   // y flux at scalar cell h + 2
    const int I = i + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T m = T(0.5) * (v(I, h + 2, K) + v(I, h + 3, K)) * c.rho[K];
      fy(i, h, k) = hflux(m, v(I, h, K), v(I, h + 1, K), v(I, h + 2, K), v(I, h + 3, K), v(I, h + 4, K), v(I, h + 5, K), weps);
    }
  });
  par2(nx, ny + 1, [&](int i, int iv) {
    for (int k = 0; k < nz; ++k) {
      const int I = i + 3, Q = iv + 3, K = k + 1;
      const T conv = -((fx(i + 1, iv, k) - fx(i, iv, k)) / c.dx + (fy(i, iv + 1, k) - fy(i, iv, k)) / c.dy +
                       (fz(i, iv, k + 1) - fz(i, iv, k)) / c.dz);
      const T div8v = T(0.5) * (fd(I, Q - 1, K) + fd(I, Q, K));
      const T rho8v = T(0.5) * (c.rho[K] + c.rho[K]);
      out(i, iv, k) = (conv + v(I, Q, K) * div8v) / rho8v;
    }
  });
  return out;
}
template <class T>
A3<T> advection_w(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w, T weps, const A3<T>& fd) {
  const int nx = c.nx, ny = c.ny, nz = c.nz, Lw = nz + 3;
  A3<T> fz(nx, ny, nz + 2), fx(nx + 1, ny, nz + 1), fy(nx, ny + 1, nz + 1), out(nx, ny, nz + 1);
  par2(nx, ny, [&](int i, int j) {   // vertical flux at scalar level f
    const int I = i + 3, J = j + 3;
    const T* p = &w.d[((size_t)I * w.n1 + J) * w.n2];
    for (int f = 0; f <= nz + 1; ++f) {
      const T m = T(0.5) * (w(I, J, f) + w(I, J, f + 1)) * c.rho[f];
      fz(i, j, f) = vflux(m, p, Lw, f, weps);
    }
  });
  par2(nx + 1, ny, [&](int g, int j) {   // x flux at (u face g + 3, w level m + 1)
    const int J = j + 3;
    for (int m = 0; m <= nz; ++m) {
      const int K = m + 1;
      const T mu0 = T(0.5) * (c.rho[m] + c.rho[m]) * u(g + 3, J, m);
      const T mu1 = T(0.5) * (c.rho[m + 1] + c.rho[m + 1]) * u(g + 3, J, m + 1);
      fx(g, j, m) = hflux(T(0.5) * (mu0 + mu1), w(g, J, K), w(g + 1, J, K), w(g + 2, J, K), w(g + 3, J, K), w(g + 4, J, K), w(g + 5, J, K), weps);
    }
  });
  par2(nx, ny + 1, [&](int i, int h) {
    const int I = i + 3;
    for (int m = 0; m <= nz; ++m) {
      const int K = m + 1;
      const T mv0 = T(0.5) * (c.rho[m] + c.rho[m]) * v(I, h + 3, m);
      const T mv1 = T(0.5) * (c.rho[m + 1] + c.rho[m + 1]) * v(I, h + 3, m + 1);
      fy(i, h, m) = hflux(T(0.5) * (mv0 + mv1), w(I, h, K), w(I, h + 1, K), w(I, h + 2, K), w(I, h + 3, K), w(I, h + 4, K), w(I, h + 5, K), weps);
    }
  });
  par2(nx, ny, [&](int i, int j) {
    const int I = i + 3, J = j + 3;
    for (int m = 0; m <= nz; ++m) {
      const T conv = -((fx(i + 1, j, m) - fx(i, j, m)) / c.dx + (fy(i, j + 1, m) - fy(i, j, m)) / c.dy +
                       (fz(i, j, m + 1) - fz(i, j, m)) / c.dz);
      const T div8w = T(0.5) * (fd(I, J, m) + fd(I, J, m + 1));
      const T rho8w = T(0.5) * (c.rho[m] + c.rho[m + 1]);
      T t = (conv + w(I, J, m + 1) * div8w) / rho8w;
      if (m == 0 || m == nz) t = T(0);
      out(i, j, m) = t;
    }
  });
  return out;
}

// ---- low-order pi' advection (advection_pi, cc1 = cc2 = 0.5) ----------------------
template <class T>
A3<T> advection_pi(const Ctx<T>& c, const A3<T>& pip, const A3<T>& u, const A3<T>& v, const A3<T>& w) {
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  A3<T> out(nx, ny, nz);
  const T cc1 = T(0.5), cc2 = T(0.5);
  par2(nx, ny, [&](int i, int j) {
    const int I = i + 3, J = j + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T udx0 = u(I, J, K) * ((pip(I, J, K) - pip(I - 1, J, K)) / c.dx);
      const T udx1 = u(I + 1, J, K) * ((pip(I + 1, J, K) - pip(I, J, K)) / c.dx);
      const T vdy0 = v(I, J, K) * ((pip(I, J, K) - pip(I, J - 1, K)) / c.dy);
      const T vdy1 = v(I, J + 1, K) * ((pip(I, J + 1, K) - pip(I, J, K)) / c.dy);
      // w levels k (below) and k + 1 (above): padded w levels K and K + 1
      auto wdz = [&](int m) -> T {   // m = interior w level 0..nz
        if (m == 0 || m == nz) return T(0);
        const T dpi0 = (c.pi0[m + 1] - c.pi0[m]) / c.dz;
        const T dpip = (pip(I, J, m + 1) - pip(I, J, m)) / c.dz;
        return w(I, J, m + 1) * (dpi0 + dpip);
      };
      const T r8lo = T(0.5) * (c.rho[k] + c.rho[k + 1]), r8hi = T(0.5) * (c.rho[k + 1] + c.rho[k + 2]);
      out(i, j, k) = -(T(0.5) * (udx1 + udx0) + T(0.5) * (vdy1 + vdy0) +
                       (cc1 * wdz(k) * r8lo + cc2 * wdz(k + 1) * r8hi) / c.rho[K]);
    }
  });
  return out;
}

// ---- buoyancy, theta_rho, pressure gradient ---------------------------------------
template <class T>
inline T theta_rho(T th, T qv) { return th * (T(1.0) + T(kReps) * qv) / (T(1.0) + qv); }

template <class T>
A3<T> b8w_of(const Ctx<T>& c, const A3<T>& th, const A3<T>& qv) {   // interpolate_scalar2w of the buoyancy
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  A3<T> out(nx, ny, nz + 1);
  par2(nx, ny, [&](int i, int j) {
    std::vector<T> b(nz);
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T tr0 = c.th0[K] * (T(1.0) + T(kReps) * c.qv0[K]) / (T(1.0) + c.qv0[K]);
      const T tr = theta_rho(th(i + 3, j + 3, K), qv(i + 3, j + 3, K));
      b[k] = T(kG) * (tr - tr0) / tr0;
    }
    out(i, j, 0) = T(1.875) * b[0] + T(-1.25) * b[1] + T(0.375) * b[2];
    for (int m = 1; m < nz; ++m) out(i, j, m) = T(0.5) * (b[m - 1] + b[m]);
    out(i, j, nz) = T(1.875) * b[nz - 1] + T(-1.25) * b[nz - 2] + T(0.375) * b[nz - 3];
  });
  return out;
}

// ---- Smagorinsky (compute_deformation -> compute_nm -> compute_k -> compute_smag) ---
template <class T>
struct Sgs {
  A3<T> su, sv, sw;
  std::vector<A3<T>> ss;
};

template <class T>
inline T ex_lo(T a, T b, T c) { return T(1.875) * a + T(-1.25) * b + T(0.375) * c; }

template <class T>
Sgs<T> smagorinsky(const Ctx<T>& c, const A3<T>& u, const A3<T>& v, const A3<T>& w, const A3<T>& th, const A3<T>& qv,
                   const std::vector<const A3<T>*>& scal, T cs, T prandtl) {
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  A3<T> s11(nx, ny, nz), s22(nx, ny, nz), s33(nx, ny, nz), s12(nx, ny, nz), s13(nx, ny, nz), s23(nx, ny, nz),
      df(nx, ny, nz), n2(nx, ny, nz);
  auto e12 = [&](int P, int Q, int K) {   // corner (u face P, v face Q)
    const T dudy = (u(P, Q, K) - u(P, Q - 1, K)) / c.dy;
    const T dvdx = (v(P, Q, K) - v(P - 1, Q, K)) / c.dx;
    return T(0.5) * (dudy + dvdx);
  };
  auto e13 = [&](int P, int J, int L) {   // (u face P, w level L = 2..nz)
    const T dudz = (u(P, J, L) - u(P, J, L - 1)) / c.dz;
    const T dwdx = (w(P, J, L) - w(P - 1, J, L)) / c.dx;
    return T(0.5) * (dudz + dwdx);
  };
  auto e23 = [&](int I, int Q, int L) {
    const T dvdz = (v(I, Q, L) - v(I, Q, L - 1)) / c.dz;
    const T dwdy = (w(I, Q, L) - w(I, Q - 1, L)) / c.dy;
    return T(0.5) * (dvdz + dwdy);
  };
  par2(nx, ny, [&](int i, int j) {
    const int I = i + 3, J = j + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T r = c.rho[K];
      const T a11 = r * (u(I + 1, J, K) - u(I, J, K)) / c.dx;
      const T a22 = r * (v(I, J + 1, K) - v(I, J, K)) / c.dy;
      const T a33 = r * (w(I, J, K + 1) - w(I, J, K)) / c.dz;
      const T a12 = r * (T(0.25) * (e12(I, J, K) + e12(I + 1, J + 1, K) + e12(I, J + 1, K) + e12(I + 1, J, K)));
      T a13, a23;
      if (k == 0 || k == nz - 1) {
        const bool lo = k == 0;
        auto ex = [&](auto fe, int A, int B) {
          return lo ? ex_lo(fe(A, B, 2), fe(A, B, 3), fe(A, B, 4)) : ex_lo(fe(A, B, nz), fe(A, B, nz - 1), fe(A, B, nz - 2));
        };
        a13 = T(0.5) * (ex(e13, I + 1, J) + ex(e13, I, J));
        a23 = T(0.5) * (ex(e23, I, J + 1) + ex(e23, I, J));
      } else {
        a13 = T(0.25) * (e13(I, J, K) + e13(I + 1, J, K + 1) + e13(I, J, K + 1) + e13(I + 1, J, K));
        a23 = T(0.25) * (e23(I, J, K) + e23(I, J + 1, K + 1) + e23(I, J, K + 1) + e23(I, J + 1, K));
      }
      a13 = r * a13;
      a23 = r * a23;
      s11(i, j, k) = a11;
      s22(i, j, k) = a22;
      s33(i, j, k) = a33;
      s12(i, j, k) = a12;
      s13(i, j, k) = a13;
      s23(i, j, k) = a23;
      df(i, j, k) = (T(2.0) * (a11 * a11 + a22 * a22 + a33 * a33) + T(4.0) * (a12 * a12 + a13 * a13 + a23 * a23)) / (r * r);
    }
    // N^2 (compute_nm)
    std::vector<T> n8(nz - 1);
    for (int L = 2; L <= nz; ++L) {
      const T ta = th(I, J, L) * (T(1.0) + T(kRepsm1) * qv(I, J, L));
      const T tb = th(I, J, L - 1) * (T(1.0) + T(kRepsm1) * qv(I, J, L - 1));
      n8[L - 2] = std::log(ta / tb) / c.dz * T(kG);
    }
    n2(i, j, 0) = ex_lo(n8[0], n8[1], n8[2]);
    for (int k = 1; k < nz - 1; ++k) n2(i, j, k) = T(0.5) * (n8[k - 1] + n8[k]);
    n2(i, j, nz - 1) = ex_lo(n8[nz - 2], n8[nz - 3], n8[nz - 4]);
  });
  // compute_k on padded arrays (the coefficients' halos are the periodic images)
  const T gs = std::cbrt((T)c.dxyz);
  A3<T> km(nx, ny, nz), kh(nx, ny, nz);
  par2(nx, ny, [&](int i, int j) {
    for (int k = 0; k < nz; ++k) {
      T s2n2 = df(i, j, k) - n2(i, j, k) / prandtl;
      s2n2 = s2n2 > T(0.0) ? s2n2 : T(0.0);
      const T cg = cs * gs;
      km(i, j, k) = (cg * cg) * std::sqrt(s2n2);
      kh(i, j, k) = km(i, j, k) / prandtl;
    }
  });
  const A3<T> P11 = pad_c(c, s11), P22 = pad_c(c, s22), P33 = pad_c(c, s33), P12 = pad_c(c, s12),
              P13 = pad_c(c, s13), P23 = pad_c(c, s23), Km = pad_c(c, km), Kh = pad_c(c, kh);
  auto tau = [&](const A3<T>& s, int I, int J, int K) { return T(-2.0) * Km(I, J, K) * s(I, J, K); };
  Sgs<T> o;
  o.su = A3<T>(nx + 1, ny, nz);
  o.sv = A3<T>(nx, ny + 1, nz);
  o.sw = A3<T>(nx, ny, nz + 1);
  // t13 / t23 at w level m (interpolate_scalar2w_0 over the interior levels)
  auto t13w = [&](int I, int J, int m) {
    return (m == 0 || m == nz) ? T(0) : T(0.5) * (tau(P13, I, J, m) + tau(P13, I, J, m + 1));
  };
  auto t23w = [&](int I, int J, int m) {
    return (m == 0 || m == nz) ? T(0) : T(0.5) * (tau(P23, I, J, m) + tau(P23, I, J, m + 1));
  };
  par2(nx + 1, ny, [&](int iu, int j) {
    const int P = iu + 3, J = j + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T d11 = (tau(P11, P, J, K) - tau(P11, P - 1, J, K)) / c.dx;
      auto d12y = [&](int X, int Y) { return (tau(P12, X, Y, K) - tau(P12, X, Y - 1, K)) / c.dy; };
      const T d12 = T(0.25) * (d12y(P - 1, J) + d12y(P, J + 1) + d12y(P - 1, J + 1) + d12y(P, J));
      auto d13s = [&](int X) { return (t13w(X, J, k + 1) - t13w(X, J, k)) / c.dz; };
      const T d13 = T(0.5) * (d13s(P - 1) + d13s(P));
      o.su(iu, j, k) = -(d11 + d12 + d13) / (T(0.5) * (c.rho[K] + c.rho[K]));
    }
  });
  par2(nx, ny + 1, [&](int i, int iv) {
    const int I = i + 3, Q = iv + 3;
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      const T d22 = (tau(P22, I, Q, K) - tau(P22, I, Q - 1, K)) / c.dy;
      auto d12x = [&](int X, int Y) { return (tau(P12, X, Y, K) - tau(P12, X - 1, Y, K)) / c.dx; };
      const T d21 = T(0.25) * (d12x(I, Q - 1) + d12x(I + 1, Q) + d12x(I, Q) + d12x(I + 1, Q - 1));
      auto d23s = [&](int Y) { return (t23w(I, Y, k + 1) - t23w(I, Y, k)) / c.dz; };
      const T d23 = T(0.5) * (d23s(Q - 1) + d23s(Q));
      o.sv(i, iv, k) = -(d21 + d22 + d23) / (T(0.5) * (c.rho[K] + c.rho[K]));
    }
  });
  par2(nx, ny, [&](int i, int j) {
    const int I = i + 3, J = j + 3;
    std::vector<T> d13x(nz), d23y(nz), rr(nz);
    for (int k = 0; k < nz; ++k) {
      const int K = k + 1;
      d13x[k] = (tau(P13, I + 1, J, K) - tau(P13, I - 1, J, K)) / (T(2.0) * c.dx);
      d23y[k] = (tau(P23, I, J + 1, K) - tau(P23, I, J - 1, K)) / (T(2.0) * c.dy);
      rr[k] = c.rho[K];
    }
    for (int m = 0; m <= nz; ++m) {
      const T d33 = (tau(P33, I, J, m + 1) - tau(P33, I, J, m)) / c.dz;
      const T x13 = (m == 0 || m == nz) ? T(0) : T(0.5) * (d13x[m - 1] + d13x[m]);
      const T y23 = (m == 0 || m == nz) ? T(0) : T(0.5) * (d23y[m - 1] + d23y[m]);
      const T r8w = (m == 0) ? ex_lo(rr[0], rr[1], rr[2])
                             : (m == nz ? ex_lo(rr[nz - 1], rr[nz - 2], rr[nz - 3]) : T(0.5) * (rr[m - 1] + rr[m]));
      T s = -(x13 + y23 + d33) / r8w;
      if (m == 0 || m == nz) s = T(0);
      o.sw(i, j, m) = s;
    }
  });
  // scalar SGS diffusion (_scalar_sgs with kh on every face)
  for (const A3<T>* sp : scal) {
    const A3<T>& s = *sp;
    A3<T> out(nx, ny, nz);
    par2(nx, ny, [&](int i, int j) {
      const int I = i + 3, J = j + 3;
      std::vector<T> rr(nz), kk(nz);
      for (int k = 0; k < nz; ++k) {
        rr[k] = c.rho[k + 1];
        kk[k] = Kh(I, J, k + 1);
      }
      auto r8w = [&](int m, const std::vector<T>& a) {
        return (m == 0) ? ex_lo(a[0], a[1], a[2]) : (m == nz ? ex_lo(a[nz - 1], a[nz - 2], a[nz - 3]) : T(0.5) * (a[m - 1] + a[m]));
      };
      for (int k = 0; k < nz; ++k) {
        const int K = k + 1;
        auto tx = [&](int X) {   // x face X (between cells X - 1 and X)
          const T r8u = T(0.5) * (c.rho[K] + c.rho[K]);
          const T k8u = T(0.5) * (Kh(X, J, K) + Kh(X - 1, J, K));
          const T dx = (s(X, J, K) - s(X - 1, J, K)) / c.dx;
          return -r8u * k8u * dx;
        };
        auto ty = [&](int Y) {
          const T r8v = T(0.5) * (c.rho[K] + c.rho[K]);
          const T k8v = T(0.5) * (Kh(I, Y, K) + Kh(I, Y - 1, K));
          const T dy = (s(I, Y, K) - s(I, Y - 1, K)) / c.dy;
          return -r8v * k8v * dy;
        };
        auto tz = [&](int m) {   // interior w level m (0..nz)
          if (m == 0 || m == nz) return T(0);
          const T dz = (s(I, J, m + 1) - s(I, J, m)) / c.dz;
          return -r8w(m, rr) * r8w(m, kk) * dz;
        };
        out(i, j, k) = -((tx(I + 1) - tx(I)) / c.dx + (ty(J + 1) - ty(J)) / c.dy + (tz(k + 1) - tz(k)) / c.dz) / c.rho[K];
      }
    });
    o.ss.push_back(std::move(out));
  }
  return o;
}

// ---- the step -----------------------------------------------------------------------
template <class T>
struct State {
  A3<T> th, u, v, w, pip, qv;
  std::vector<A3<T>> pas;
};

template <class T>
A3<T> interior_c(const Ctx<T>& c, const A3<T>& a) {
  A3<T> o(c.nx, c.ny, c.nz);
  par2(c.nx, c.ny, [&](int i, int j) {
    for (int k = 0; k < c.nz; ++k) o(i, j, k) = a(i + 3, j + 3, k + 1);
  });
  return o;
}

template <class T>
State<T> sound_rk_step(const Ctx<T>& c, const State<T>& x0, const A3<T>& pip_star, const A3<T>& ru, const A3<T>& rv,
                       const A3<T>& rw, const A3<T>& trho, int n, T dts) {
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  A3<T> u = x0.u, v = x0.v, w = x0.w, p = x0.pip;
  const T cr = T(-kRd / kCv), h = T(-kCp * 0.5);
  for (int s = 0; s < n; ++s) {
    A3<T> ui(nx + 1, ny, nz), vi(nx, ny + 1, nz), wi(nx, ny, nz + 1);
    par2(nx + 1, ny, [&](int iu, int j) {
      const int P = iu + 3, J = j + 3;
      for (int k = 0; k < nz; ++k) {
        const int K = k + 1;
        const T pu = h * (trho(P, J, K) + trho(P - 1, J, K)) * ((p(P, J, K) - p(P - 1, J, K)) / c.dx);
        ui(iu, j, k) = u(P, J, K) + (pu + ru(iu, j, k)) * dts;
      }
    });
    par2(nx, ny + 1, [&](int i, int iv) {
      const int I = i + 3, Q = iv + 3;
      for (int k = 0; k < nz; ++k) {
        const int K = k + 1;
        const T pv = h * (trho(I, Q, K) + trho(I, Q - 1, K)) * ((p(I, Q, K) - p(I, Q - 1, K)) / c.dy);
        vi(i, iv, k) = v(I, Q, K) + (pv + rv(i, iv, k)) * dts;
      }
    });
    par2(nx, ny, [&](int i, int j) {
      const int I = i + 3, J = j + 3;
      for (int m = 0; m <= nz; ++m) {
        T pw = h * (trho(I, J, m + 1) + trho(I, J, m)) * ((p(I, J, m + 1) - p(I, J, m)) / c.dz);
        if (m == 0 || m == nz) pw = T(0);
        wi(i, j, m) = w(I, J, m + 1) + (pw + rw(i, j, m)) * dts;
      }
    });
    u = pad_u(c, ui);
    v = pad_v(c, vi);
    w = pad_w(c, wi);
    const A3<T> div = divergence2(c, u, v, w);
    const A3<T> apt = advection_pi(c, pip_star, u, v, w);
    A3<T> pn(nx, ny, nz);
    par2(nx, ny, [&](int i, int j) {
      for (int k = 0; k < nz; ++k) {
        const int I = i + 3, J = j + 3, K = k + 1;
        const T piadv = c.pi0[K] + pip_star(I, J, K);
        const T dpdt = cr * (piadv * div(I, J, K)) + apt(i, j, k);
        pn(i, j, k) = p(I, J, K) + T(1.1) * dpdt * dts;
      }
    });
    p = pad_c(c, pn);
  }
  State<T> o;
  o.u = std::move(u);
  o.v = std::move(v);
  o.w = std::move(w);
  o.pip = std::move(p);
  return o;
}

template <class T>
A3<T> update_scalar(const Ctx<T>& c, const A3<T>& s0, const A3<T>& s, const A3<T>& u, const A3<T>& v,
                    const A3<T>& w, const A3<T>& fdiv, T weps, T dt, const A3<T>* sgs) {
  const A3<T> tend = advection_scalar(c, s, u, v, w, weps, fdiv);
  A3<T> in(c.nx, c.ny, c.nz);
  par2(c.nx, c.ny, [&](int i, int j) {
    for (int k = 0; k < c.nz; ++k) {
      T x = s0(i + 3, j + 3, k + 1) + tend(i, j, k) * dt;
      if (sgs) x = x + (*sgs)(i, j, k) * dt;
      in(i, j, k) = x;
    }
  });
  return pad_c(c, in);
}

template <class T>
State<T> stage(const Ctx<T>& c, int st, const State<T>& x0, const State<T>& x, double dt_stage, double dt,
               int n_sound, const Sgs<T>& sgs) {
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  const T WEPS_MOM = T(1.0e-16), WEPS_TH = T(1.0e-17), WEPS_QV = T(1.0e-18);
  const A3<T> fdiv = divergence(c, x.u, x.v, x.w);
  const A3<T> b8w = b8w_of(c, x.th, x.qv);
  A3<T> ru = advection_u(c, x.u, x.v, x.w, WEPS_MOM, fdiv);
  A3<T> rv = advection_v(c, x.u, x.v, x.w, WEPS_MOM, fdiv);
  A3<T> rw = advection_w(c, x.u, x.v, x.w, WEPS_MOM, fdiv);
  // rhs = adv + (Coriolis 0) + b8w + frozen SGS
  par2(nx + 1, ny, [&](int i, int j) {
    for (int k = 0; k < nz; ++k) ru(i, j, k) = (ru(i, j, k) + T(0)) + sgs.su(i, j, k);
  });
  par2(nx, ny + 1, [&](int i, int j) {
    for (int k = 0; k < nz; ++k) rv(i, j, k) = (rv(i, j, k) - T(0)) + sgs.sv(i, j, k);
  });
  par2(nx, ny, [&](int i, int j) {
    for (int m = 0; m <= nz; ++m) rw(i, j, m) = (rw(i, j, m) + b8w(i, j, m)) + sgs.sw(i, j, m);
  });
  A3<T> trho(nx + 6, ny + 6, nz + 2);
  par2(nx + 6, ny + 6, [&](int I, int J) {
    for (int K = 0; K < nz + 2; ++K) trho(I, J, K) = theta_rho(x.th(I, J, K), x.qv(I, J, K));
  });
  int n;
  double dts;
  if (st == 0) {
    n = 1;
    dts = dt_stage;
  } else {
    dts = dt / n_sound;
    n = (int)std::nearbyint(dt_stage / dts);   // Python round: half to even (the FE default mode)
  }
  State<T> o = sound_rk_step(c, x0, x.pip, ru, rv, rw, trho, n, (T)dts);
  const T dtT = (T)dt_stage;
  o.th = update_scalar(c, x0.th, x.th, x.u, x.v, x.w, fdiv, WEPS_TH, dtT, &sgs.ss[0]);
  o.qv = update_scalar(c, x0.qv, x.qv, x.u, x.v, x.w, fdiv, WEPS_QV, dtT, &sgs.ss[1]);
  for (size_t q = 0; q < x.pas.size(); ++q)
    o.pas.push_back(update_scalar(c, x0.pas[q], x.pas[q], x.u, x.v, x.w, fdiv, WEPS_QV, dtT, &sgs.ss[2 + q]));
  return o;
}

template <class T>
State<T> big_step(const Ctx<T>& c, const State<T>& x0, double dt, int n_sound, T cs, T prandtl) {
  std::vector<const A3<T>*> scal = {&x0.th, &x0.qv};
  for (const auto& a : x0.pas) scal.push_back(&a);
  const Sgs<T> sgs = smagorinsky(c, x0.u, x0.v, x0.w, x0.th, x0.qv, scal, cs, prandtl);
  State<T> x1 = stage(c, 0, x0, x0, dt / 3.0, dt, n_sound, sgs);
  State<T> x2 = stage(c, 1, x0, x1, dt / 2.0, dt, n_sound, sgs);
  return stage(c, 2, x0, x2, dt, dt, n_sound, sgs);
}

template <class T>
int run(int nx, int ny, int nz, double dx, double dy, double dz, double dt, int n_sound, double cs, double prandtl,
        int n_passive, const T* rho0, const T* theta0, const T* qv0, const T* pi0, T* const* state, int nsteps) {
  if (nx < 4 || ny < 4 || nz < 5 || n_passive < 0 || n_passive > 8 || nsteps < 0 || n_sound < 1) return -1;
  Ctx<T> c;
  c.nx = nx;
  c.ny = ny;
  c.nz = nz;
  c.dx = (T)dx;
  c.dy = (T)dy;
  c.dz = (T)dz;
  c.dxyz = dx * dy * dz;
  c.rho.assign(rho0, rho0 + nz + 2);
  c.th0.assign(theta0, theta0 + nz + 2);
  c.qv0.assign(qv0, qv0 + nz + 2);
  c.pi0.assign(pi0, pi0 + nz + 2);
  auto load = [&](const T* p, int a, int b, int d) {
    A3<T> o(a, b, d);
    std::memcpy(o.d.data(), p, o.d.size() * sizeof(T));
    return o;
  };
  State<T> x;
  x.th = load(state[0], nx + 6, ny + 6, nz + 2);
  x.u = load(state[1], nx + 7, ny + 6, nz + 2);
  x.v = load(state[2], nx + 6, ny + 7, nz + 2);
  x.w = load(state[3], nx + 6, ny + 6, nz + 3);
  x.pip = load(state[4], nx + 6, ny + 6, nz + 2);
  x.qv = load(state[5], nx + 6, ny + 6, nz + 2);
  for (int q = 0; q < n_passive; ++q) x.pas.push_back(load(state[6 + q], nx + 6, ny + 6, nz + 2));
  for (int s = 0; s < nsteps; ++s) x = big_step(c, x, dt, n_sound, (T)cs, (T)prandtl);
  auto save = [&](const A3<T>& a, T* p) { std::memcpy(p, a.d.data(), a.d.size() * sizeof(T)); };
  save(x.th, state[0]);
  save(x.u, state[1]);
  save(x.v, state[2]);
  save(x.w, state[3]);
  save(x.pip, state[4]);
  save(x.qv, state[5]);
  for (int q = 0; q < n_passive; ++q) save(x.pas[q], state[6 + q]);
  return 0;
}

}  // namespace

extern "C" {
// OpenMP threads the steps run on
int lex_cpu_threads() { return omp_get_max_threads(); }

// nsteps WS-RK3 steps of the padded state in place (state = theta, u, v, w, pip, qv,
// then n_passive scalars; base columns of length nz + 2).  Returns 0, or -1 for a bad
// argument.  The thread count is OpenMP's (OMP_NUM_THREADS).
int lex_cpu_rk3_steps_f32(int nx, int ny, int nz, double dx, double dy, double dz, double dt, int n_sound, double cs,
                          double prandtl, int n_passive, const float* rho0, const float* theta0, const float* qv0,
                          const float* pi0, float* const* state, int nsteps) {
  return run<float>(nx, ny, nz, dx, dy, dz, dt, n_sound, cs, prandtl, n_passive, rho0, theta0, qv0, pi0, state, nsteps);
}
int lex_cpu_rk3_steps_f64(int nx, int ny, int nz, double dx, double dy, double dz, double dt, int n_sound, double cs,
                          double prandtl, int n_passive, const double* rho0, const double* theta0, const double* qv0,
                          const double* pi0, double* const* state, int nsteps) {
  return run<double>(nx, ny, nz, dx, dy, dz, dt, n_sound, cs, prandtl, n_passive, rho0, theta0, qv0, pi0, state, nsteps);
}
}
