// ORACLE — test infrastructure only. Serial restatement of the reference's
// finite-volume kernel bodies on the SIMPLEC path (src/kutta/kernels/fv_*.kern.h),
// one function per kernel with the reference's parameter order, evaluated in
// the reference's operation order with -ffp-contract=off. oracle_fv_each()
// runs one of them over [0, n) as the serial CPU backend does
// (runtime/cpu/cpu_device.cpp:64-68).
#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <utility>

#include "oracle.hpp"

namespace oracle::fv {
namespace {

// kernels/prelude.h:66-74 (CPU branch: std::fmax/fmin/sqrt/fabs/fma).
inline float kmax(float a, float b) { return std::fmax(a, b); }
inline float kmin(float a, float b) { return std::fmin(a, b); }
inline float ksqrt(float a) { return std::sqrt(a); }
inline float kabs(float a) { return std::fabs(a); }

using I = const int*;
using F = const float*;
using W = float*;
using WI = int*;

// ---------------------------------------------------------- fv_flow.kern.h
// :5-32. Boundary kinds: 0 wall, 1 inlet, 2 outlet with prescribed backflow,
// 3 symmetry, 4 pressure outlet with normal backflow.
void fv_velocity_boundary(int i, I owner, I kind, F sx, F sy, F sz, F u, F v, F w, F pu, F pv,
                          F pw, F flux, W bu, W bv, W bw, WI scalar_kind, int internal, float rho) {
  const int f = i + internal, c = owner[f], k = kind[i];
  float vx = pu[i], vy = pv[i], vz = pw[i];
  const float area2 = sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f];
  scalar_kind[i] = 0;
  if ((k == 2 || k == 4) && flux[f] >= 0.0f) {
    vx = u[c];
    vy = v[c];
    vz = w[c];
    scalar_kind[i] = 2;
  }
  if (k == 4 && flux[f] < 0.0f) {
    const float area = ksqrt(area2), speed = flux[f] / (rho * area);
    vx = speed * (sx[f] / area);
    vy = speed * (sy[f] / area);
    vz = speed * (sz[f] / area);
  }
  if (k == 3) {
    vx = u[c];
    vy = v[c];
    vz = w[c];
    scalar_kind[i] = 2;
  }
  if (k == 0 || k == 3) {
    const float un = (vx * sx[f] + vy * sy[f] + vz * sz[f]) / area2;
    vx -= un * sx[f];
    vy -= un * sy[f];
    vz -= un * sz[f];
  }
  bu[i] = vx;
  bv[i] = vy;
  bw[i] = vz;
}

// :34-41
void fv_pressure_boundary(int i, I owner, I kind, F rx, F ry, F rz, F p, F gx, F gy, F gz,
                          F prescribed, W boundary, int internal) {
  const int f = i + internal, c = owner[f];
  const float dx = rx[f], dy = ry[f], dz = rz[f];
  if (kind[i] == 2 || kind[i] == 4) boundary[i] = prescribed[i];
  else boundary[i] = p[c] + gx[c] * dx + gy[c] * dy + gz[c] * dz;
}

// :43-65
void fv_pressure_faces(int i, I owner, I neighbour, F px, F py, F pz, F nx, F ny, F nz, F p,
                       F gx, F gy, F gz, F boundary, W face, W jump, int internal) {
  const int c = owner[i];
  const float side_p = p[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
  if (i >= internal) {
    face[i] = boundary[i - internal];
    jump[i] = side_p - boundary[i - internal];
    return;
  }
  const int n = neighbour[i];
  const float side_n = p[n] + gx[n] * nx[i] + gy[n] * ny[i] + gz[n] * nz[i];
  face[i] = 0.5f * (side_p + side_n);
  const float dx = px[i] - nx[i], dy = py[i] - ny[i], dz = pz[i] - nz[i];
  const float mx = 0.5f * (px[i] + nx[i]), my = 0.5f * (py[i] + ny[i]), mz = 0.5f * (pz[i] + nz[i]);
  const float d2 = dx * dx + dy * dy + dz * dz, m2 = mx * mx + my * my + mz * mz;
  const float trust = kmax(0.0f, kmin(1.0f, (4.0f * d2 - m2) / (3.0f * d2)));
  const float link =
      p[c] - p[n] + 0.5f * ((gx[c] + gx[n]) * dx + (gy[c] + gy[n]) * dy + (gz[c] + gz[n]) * dz);
  jump[i] = trust == 1.0f ? side_p - side_n : link + trust * ((side_p - side_n) - link);
}

// :67-78
void fv_momentum_source(int i, I offsets, I faces, F sign, F sx, F sy, F sz, F face_pressure,
                        F volume, F body_u, F body_v, F body_w, W su, W sv, W sw) {
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const float pf = sign[j] * face_pressure[f];
    fx += pf * sx[f];
    fy += pf * sy[f];
    fz += pf * sz[f];
  }
  su[i] = body_u[i] - fx / volume[i];
  sv[i] = body_v[i] - fy / volume[i];
  sw[i] = body_w[i] - fz / volume[i];
}

// :83-109
void fv_momentum_faces(int i, I owner, I neighbour, I kind, F px, F py, F pz, F nx, F ny, F nz,
                       F u, F v, F w, F ux, F uy, F uz, F vx, F vy, F vz, F wx, F wy, F wz,
                       F diagonal, F bu, F bv, F bw, W fu, W fv, W fw, int internal) {
  if (i >= internal) {
    const int b = i - internal, c = owner[i];
    const bool outlet = kind[b] == 2 || kind[b] == 4;
    fu[i] = outlet ? u[c] : bu[b];
    fv[i] = outlet ? v[c] : bv[b];
    fw[i] = outlet ? w[c] : bw[b];
    return;
  }
  const int c = owner[i], n = neighbour[i];
  const float a = diagonal[c] / (diagonal[c] + diagonal[n]);
  const float rx = a * px[i] + (1.0f - a) * nx[i];
  const float ry = a * py[i] + (1.0f - a) * ny[i];
  const float rz = a * pz[i] + (1.0f - a) * nz[i];
  fu[i] = a * u[c] + (1.0f - a) * u[n] + 0.5f * ((ux[c] + ux[n]) * rx + (uy[c] + uy[n]) * ry + (uz[c] + uz[n]) * rz);
  fv[i] = a * v[c] + (1.0f - a) * v[n] + 0.5f * ((vx[c] + vx[n]) * rx + (vy[c] + vy[n]) * ry + (vz[c] + vz[n]) * rz);
  fw[i] = a * w[c] + (1.0f - a) * w[n] + 0.5f * ((wx[c] + wx[n]) * rx + (wy[c] + wy[n]) * ry + (wz[c] + wz[n]) * rz);
}

// :111-116
void fv_convective_outflow(int i, I offsets, I faces, F sign, F flux, W outflow) {
  float total = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) total += kmax(0.0f, sign[j] * flux[faces[j]]);
  outflow[i] = total;
}

// :122-133
void fv_peclet_side_velocity(int i, I owner, I neighbour, F px, F py, F pz, F nx, F ny, F nz, F q,
                             F gx, F gy, F gz, F weights, F diagonal, F outflow, W face) {
  const int c = owner[i], n = neighbour[i];
  const float a = weights[c] / (weights[c] + weights[n]);
  const float beta = kmin(1.0f, (outflow[c] + outflow[n]) / kmax(diagonal[c] + diagonal[n], 1.0e-20f));
  const float side = a * (q[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i]) +
                     (1.0f - a) * (q[n] + gx[n] * nx[i] + gy[n] * ny[i] + gz[n] * nz[i]);
  face[i] += beta * (side - face[i]);
}

// :139-162
void fv_rhie_chow(int i, I owner, I neighbour, I kind, F sx, F sy, F sz, F diffusion, F volume,
                  F diagonal, F inverse, F u, F v, F w, F pressure_jump, W flux, W coefficient,
                  int internal, float rho) {
  const int c = owner[i];
  const float ux = u[i], uy = v[i], uz = w[i];
  float rc = volume[c] / kmax(diagonal[c], 1.0e-20f), dc = volume[c] * inverse[c];
  const float jump = pressure_jump[i];
  if (i < internal) {
    const int n = neighbour[i];
    rc = 0.5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
    dc = 0.5f * (dc + volume[n] * inverse[n]);
  } else {
    const int k = kind[i - internal];
    if (k == 0 || k == 3) {
      flux[i] = 0.0f;
      coefficient[i] = 0.0f;
      return;
    }
    if (k == 1) {
      flux[i] = rho * (ux * sx[i] + uy * sy[i] + uz * sz[i]);
      coefficient[i] = 0.0f;
      return;
    }
  }
  flux[i] = rho * (ux * sx[i] + uy * sy[i] + uz * sz[i]) + rho * rc * diffusion[i] * jump;
  coefficient[i] = rho * dc * diffusion[i];
}

// :185-190
void fv_correct_flux(int i, I owner, I neighbour, F correction, F coefficient, W flux, int internal) {
  const float pc = correction[owner[i]];
  const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
  flux[i] += coefficient[i] * (pc - pn);
}

// :192-202
void fv_correct_velocity(int i, F volume, F inverse, F diagonal, F gx, F gy, F gz, F correction,
                         W u, W v, W w, W p, float alpha) {
  const float d = volume[i] * inverse[i];
  u[i] -= d * gx[i];
  v[i] -= d * gy[i];
  w[i] -= d * gz[i];
  p[i] += kmin(alpha, 0.7f * diagonal[i] * inverse[i]) * correction[i];
}

// :204-215
void fv_monitor_cell(int i, I offsets, I faces, F sign, F flux, F u, F v, F w, F diagonal,
                     F residual_u, F residual_v, F residual_w, W continuity, W scale, W speed, W ru,
                     W rv, W rw) {
  float mass = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) mass += sign[j] * flux[faces[j]];
  const float s = ksqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
  continuity[i] = kabs(mass);
  scale[i] = kabs(diagonal[i]) * s;
  speed[i] = s;
  ru[i] = kabs(residual_u[i]);
  rv[i] = kabs(residual_v[i]);
  rw[i] = kabs(residual_w[i]);
}

// :217-219
void fv_speed(int i, F u, F v, F w, W speed) { speed[i] = ksqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]); }

// :221-226
void fv_continuity(int i, I offsets, I faces, F sign, F flux, W result) {
  float mass = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) mass += sign[j] * flux[faces[j]];
  result[i] = kabs(mass);
}

// :228-233
void fv_wall_pressure_force(int i, F selected, F sx, F sy, F sz, F face_pressure, W x, W y, W z,
                            int internal, float reference) {
  const int f = i + internal;
  const float pf = selected[i] * (face_pressure[f] - reference);
  x[i] = pf * sx[f];
  y[i] = pf * sy[f];
  z[i] = pf * sz[f];
}

// :235-257
void fv_wall_viscous_force(int i, I owner, F selected, F sx, F sy, F sz, F rx, F ry, F rz, F u, F v,
                           F w, F ux, F uy, F uz, F vx, F vy, F vz, F wx, F wy, F wz, F bu, F bv,
                           F bw, W x, W y, W z, int internal, float viscosity) {
  const int f = i + internal, c = owner[f];
  const float area = ksqrt(sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f]);
  const float ex = sx[f] / area, ey = sy[f] / area, ez = sz[f] / area;
  const float dn = rx[f] * ex + ry[f] * ey + rz[f] * ez;
  const float du = (bu[i] - u[c] - ux[c] * rx[f] - uy[c] * ry[f] - uz[c] * rz[f]) / dn;
  const float dv = (bv[i] - v[c] - vx[c] * rx[f] - vy[c] * ry[f] - vz[c] * rz[f]) / dn;
  const float dw = (bw[i] - w[c] - wx[c] * rx[f] - wy[c] * ry[f] - wz[c] * rz[f]) / dn;
  // Wall-corrected velocity gradient rows.
  const float g00 = ux[c] + du * ex, g01 = uy[c] + du * ey, g02 = uz[c] + du * ez;
  const float g10 = vx[c] + dv * ex, g11 = vy[c] + dv * ey, g12 = vz[c] + dv * ez;
  const float g20 = wx[c] + dw * ex, g21 = wy[c] + dw * ey, g22 = wz[c] + dw * ez;
  const float s = -viscosity * selected[i];
  x[i] = s * (2.0f * g00 * sx[f] + (g01 + g10) * sy[f] + (g02 + g20) * sz[f]);
  y[i] = s * ((g10 + g01) * sx[f] + 2.0f * g11 * sy[f] + (g12 + g21) * sz[f]);
  z[i] = s * ((g20 + g02) * sx[f] + (g21 + g12) * sy[f] + 2.0f * g22 * sz[f]);
}

// :261-266
void fv_fluent_momentum_scale(int i, F ax, F ay, F az, F volume, F speed, W x, W y, W z, float rho,
                              float dt) {
  const float t = dt > 0.0f ? rho * volume[i] / dt : 0.0f;
  x[i] = 2.0f * (ax[i] + t) * speed[i];
  y[i] = 2.0f * (ay[i] + t) * speed[i];
  z[i] = 2.0f * (az[i] + t) * speed[i];
}

// :268-276
void fv_physical_time_terms(int i, F volume, F old1, F old2, W diagonal, W rhs, float rho_dt,
                            float a0, float a1, float a2, int add) {
  const float s = rho_dt * volume[i];
  const float d0 = add ? diagonal[i] : 0.0f;
  const float b0 = add ? rhs[i] : 0.0f;
  diagonal[i] = d0 + s * a0;
  rhs[i] = b0 + s * (a1 * old1[i] + a2 * old2[i]);
}

// :278-281
void fv_face_time_defect(int i, F mass, F sx, F sy, F sz, F ux, F uy, F uz, W defect, float rho) {
  defect[i] = mass[i] - rho * (sx[i] * ux[i] + sy[i] * uy[i] + sz[i] * uz[i]);
}

// :283-295
void fv_face_time_flux(int i, I owner, I neighbour, I kind, F volume, F diagonal, F defect1,
                       F defect2, W mass, int internal, float rho_dt, float a1, float a2) {
  const int c = owner[i];
  float d = volume[c] / kmax(diagonal[c], 1.0e-20f);
  if (i < internal) {
    const int n = neighbour[i];
    d = 0.5f * (d + volume[n] / kmax(diagonal[n], 1.0e-20f));
  } else {
    const int k = kind[i - internal];
    if (k == 0 || k == 1 || k == 3) return;
  }
  mass[i] += rho_dt * d * (a1 * defect1[i] + a2 * defect2[i]);
}

// :297-299
void fv_cell_mass_error(int i, F imbalance, F volume, W error, float dt_over_rho) {
  error[i] = dt_over_rho * kabs(imbalance[i]) / volume[i];
}

// ------------------------------------------------------ fv_gradient.kern.h
// Pair helpers, kernels/prelude.h:102-133: TwoSum, FMA-carried product-add,
// a gradient trace of a pair, and the difference of two pairs.
struct Pair2 {
  float hi, lo;
};
inline Pair2 two_sum(float a, float b) {
  const float s = a + b, t = s - a;
  return {s, (a - (s - t)) + (b - t)};
}
inline Pair2 product_add(Pair2 acc, float x, float y) {
  const float p = x * y;
  const Pair2 s = two_sum(acc.hi, p);
  return two_sum(s.hi, (acc.lo + s.lo) + std::fma(x, y, -p));
}
inline Pair2 trace(float p, float low, float gx, float gy, float gz, float lx, float ly, float lz,
                   float x, float y, float z) {
  Pair2 v = two_sum(p, low);
  v = product_add(v, gx, x);
  v = product_add(v, lx, x);
  v = product_add(v, gy, y);
  v = product_add(v, ly, y);
  v = product_add(v, gz, z);
  v = product_add(v, lz, z);
  return v;
}
inline float pair_delta(Pair2 a, Pair2 b) {
  const Pair2 d = two_sum(a.hi, -b.hi);
  return d.hi + (d.lo + (a.lo - b.lo));
}
// Adds w * (d.hi + d.lo) to a pair accumulator, high word first.
inline Pair2 add_weighted(Pair2 acc, float w, Pair2 d) {
  return product_add(product_add(acc, w, d.hi), w, d.lo);
}

// Neighbour value of entry j: a cell, or boundary face -(other+1).
inline float across(int other, F cell, F boundary) { return other < 0 ? boundary[-other - 1] : cell[other]; }

// :6-23. inverse holds the symmetric 3x3 as six planes xx xy xz yy yz zz.
void fv_gradient(int i, I offsets, I adjacent, F dx, F dy, F dz, F fx, F fy, F fz, F inverse,
                 F cell, F boundary, W gx, W gy, W gz, int n) {
  float bx = 0.0f, by = 0.0f, bz = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int o = adjacent[j];
    const float d = across(o, cell, boundary) - cell[i];
    const float wx = o < 0 ? fx[-o - 1] : dx[j];
    const float wy = o < 0 ? fy[-o - 1] : dy[j];
    const float wz = o < 0 ? fz[-o - 1] : dz[j];
    bx += wx * d;
    by += wy * d;
    bz += wz * d;
  }
  gx[i] = inverse[i] * bx + inverse[i + n] * by + inverse[i + 2 * n] * bz;
  gy[i] = inverse[i + n] * bx + inverse[i + 3 * n] * by + inverse[i + 4 * n] * bz;
  gz[i] = inverse[i + 2 * n] * bx + inverse[i + 4 * n] * by + inverse[i + 5 * n] * bz;
}

// :28-56
void fv_paired_gradient(int i, I offsets, I adjacent, F dx, F dy, F dz, F fx, F fy, F fz,
                        F inverse, F cell, F low, F boundary, F boundary_low, W gx, W gy, W gz,
                        W lx, W ly, W lz, int n) {
  Pair2 bx{0.0f, 0.0f}, by{0.0f, 0.0f}, bz{0.0f, 0.0f};
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int o = adjacent[j];
    const float hi = across(o, cell, boundary);
    const float lo = o < 0 ? boundary_low[-o - 1] : low[o];
    Pair2 d = two_sum(hi, -cell[i]);
    d = product_add(d, lo, 1.0f);
    d = product_add(d, -low[i], 1.0f);
    const float wx = o < 0 ? fx[-o - 1] : dx[j];
    const float wy = o < 0 ? fy[-o - 1] : dy[j];
    const float wz = o < 0 ? fz[-o - 1] : dz[j];
    bx = add_weighted(bx, wx, d);
    by = add_weighted(by, wy, d);
    bz = add_weighted(bz, wz, d);
  }
  const float xx = inverse[i], xy = inverse[i + n], xz = inverse[i + 2 * n];
  const float yy = inverse[i + 3 * n], yz = inverse[i + 4 * n], zz = inverse[i + 5 * n];
  Pair2 x{0.0f, 0.0f}, y{0.0f, 0.0f}, z{0.0f, 0.0f};
  x = add_weighted(add_weighted(add_weighted(x, xx, bx), xy, by), xz, bz);
  y = add_weighted(add_weighted(add_weighted(y, xy, bx), yy, by), yz, bz);
  z = add_weighted(add_weighted(add_weighted(z, xz, bx), yz, by), zz, bz);
  gx[i] = x.hi;
  gy[i] = y.hi;
  gz[i] = z.hi;
  lx[i] = x.lo;
  ly[i] = y.lo;
  lz[i] = z.lo;
}

// :64-68. eps2 < 0: Barth-Jespersen; otherwise Venkatakrishnan's smooth form.
inline float limit_face(float d, float delta, float eps2) {
  if (eps2 < 0.0f) return delta / d;
  const float a = delta * delta + eps2, b = d * delta;
  return (a + 2.0f * b) / (a + 2.0f * d * d + b);
}

// :72-74
void fv_range_shift(int i, F field, W above, W below, float bound) {
  above[i] = field[i] + bound;
  below[i] = bound - field[i];
}

// Increment grad . r to the face of entry j, r on the cell's own side.
inline float face_increment(float gx, float gy, float gz, int j, I faces, F sign, F rx, F ry, F rz,
                            F nx, F ny, F nz) {
  const int f = faces[j];
  const bool owner_side = sign[j] > 0.0f;
  return gx * (owner_side ? rx[f] : nx[f]) + gy * (owner_side ? ry[f] : ny[f]) +
         gz * (owner_side ? rz[f] : nz[f]);
}
// Smallest factor over the faces, clamped to [0, 1], given extremal rooms.
inline float limit_cell(int i, float gx, float gy, float gz, float up, float down, I offsets,
                        I faces, F sign, F rx, F ry, F rz, F nx, F ny, F nz, float eps2) {
  float psi = 1.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const float d = face_increment(gx, gy, gz, j, faces, sign, rx, ry, rz, nx, ny, nz);
    if (d > 0.0f) psi = kmin(psi, limit_face(d, up, eps2));
    if (d < 0.0f) psi = kmin(psi, limit_face(d, down, eps2));
  }
  return kmax(0.0f, kmin(1.0f, psi));
}

// :76-94
void fv_limiter(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry, F rz, F nx, F ny, F nz,
                F cell, F boundary, F gx, F gy, F gz, W limiter, float eps2) {
  float lo = cell[i], hi = cell[i];
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const float v = across(adjacent[j], cell, boundary);
    lo = kmin(lo, v);
    hi = kmax(hi, v);
  }
  limiter[i] = limit_cell(i, gx[i], gy[i], gz[i], hi - cell[i], lo - cell[i], offsets, faces,
                          sign, rx, ry, rz, nx, ny, nz, eps2);
}

// :103-123
void fv_paired_limiter(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry, F rz, F nx, F ny,
                       F nz, F cell, F low, F boundary, I kind, F gx, F gy, F gz, W limiter,
                       float eps2) {
  float lo = 0.0f, hi = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int o = adjacent[j];
    const float v = across(o, cell, boundary);
    const float l = o < 0 ? (kind[-o - 1] == 2 ? low[i] : 0.0f) : low[o];
    const float d = (v - cell[i]) + (l - low[i]);
    lo = kmin(lo, d);
    hi = kmax(hi, d);
  }
  limiter[i] = limit_cell(i, gx[i], gy[i], gz[i], hi, lo, offsets, faces, sign, rx, ry, rz, nx,
                          ny, nz, eps2);
}

// :127-145
void fv_paired_gradient_limiter(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry, F rz,
                                F nx, F ny, F nz, F cell, F low, F boundary, F boundary_low, F gx,
                                F gy, F gz, F lx, F ly, F lz, W limiter, float eps2) {
  float lo = 0.0f, hi = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int o = adjacent[j];
    const float v = across(o, cell, boundary);
    const float l = o < 0 ? boundary_low[-o - 1] : low[o];
    const float d = pair_delta({v, l}, {cell[i], low[i]});
    lo = kmin(lo, d);
    hi = kmax(hi, d);
  }
  limiter[i] = limit_cell(i, gx[i] + lx[i], gy[i] + ly[i], gz[i] + lz[i], hi, lo, offsets, faces,
                          sign, rx, ry, rz, nx, ny, nz, eps2);
}

// :147-166
void fv_vector_gradient(int i, I offsets, I adjacent, F dx, F dy, F dz, F fx, F fy, F fz,
                        F inverse, F u, F v, F w, F bu, F bv, F bw, W ux, W uy, W uz, W vx, W vy,
                        W vz, W wx, W wy, W wz, int n) {
  float au = 0.0f, bu_ = 0.0f, cu = 0.0f, av = 0.0f, bv_ = 0.0f, cv = 0.0f, aw = 0.0f, bw_ = 0.0f,
        cw = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int o = adjacent[j];
    const float du = across(o, u, bu) - u[i];
    const float dv = across(o, v, bv) - v[i];
    const float dw = across(o, w, bw) - w[i];
    const float rx = o < 0 ? fx[-o - 1] : dx[j];
    const float ry = o < 0 ? fy[-o - 1] : dy[j];
    const float rz = o < 0 ? fz[-o - 1] : dz[j];
    au += rx * du;
    bu_ += ry * du;
    cu += rz * du;
    av += rx * dv;
    bv_ += ry * dv;
    cv += rz * dv;
    aw += rx * dw;
    bw_ += ry * dw;
    cw += rz * dw;
  }
  const float xx = inverse[i], xy = inverse[i + n], xz = inverse[i + 2 * n];
  const float yy = inverse[i + 3 * n], yz = inverse[i + 4 * n], zz = inverse[i + 5 * n];
  ux[i] = xx * au + xy * bu_ + xz * cu;
  uy[i] = xy * au + yy * bu_ + yz * cu;
  uz[i] = xz * au + yz * bu_ + zz * cu;
  vx[i] = xx * av + xy * bv_ + xz * cv;
  vy[i] = xy * av + yy * bv_ + yz * cv;
  vz[i] = xz * av + yz * bv_ + zz * cv;
  wx[i] = xx * aw + xy * bw_ + xz * cw;
  wy[i] = xy * aw + yy * bw_ + yz * cw;
  wz[i] = xz * aw + yz * bw_ + zz * cw;
}

// :168-171
void fv_scale_gradient(int i, F gx, F gy, F gz, F limiter, W x, W y, W z) {
  x[i] = limiter[i] * gx[i];
  y[i] = limiter[i] * gy[i];
  z[i] = limiter[i] * gz[i];
}

// :173-180
void fv_limiter_scale(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry, F rz, F nx, F ny,
                      F nz, F cell, F boundary, F gx, F gy, F gz, W limiter, W x, W y, W z,
                      float eps2) {
  fv_limiter(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, boundary, gx, gy, gz,
             limiter, eps2);
  fv_scale_gradient(i, gx, gy, gz, limiter, x, y, z);
}

// :182-189
void fv_paired_limiter_scale(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry, F rz, F nx,
                             F ny, F nz, F cell, F low, F boundary, I kind, F gx, F gy, F gz,
                             W limiter, W x, W y, W z, float eps2) {
  fv_paired_limiter(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, low, boundary,
                    kind, gx, gy, gz, limiter, eps2);
  fv_scale_gradient(i, gx, gy, gz, limiter, x, y, z);
}

// :191-199
void fv_paired_gradient_limiter_scale(int i, I offsets, I adjacent, I faces, F sign, F rx, F ry,
                                      F rz, F nx, F ny, F nz, F cell, F low, F boundary,
                                      F boundary_low, F gx, F gy, F gz, F lx, F ly, F lz,
                                      W limiter, W x, W y, W z, float eps2) {
  fv_paired_gradient_limiter(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, low,
                             boundary, boundary_low, gx, gy, gz, lx, ly, lz, limiter, eps2);
  fv_scale_gradient(i, gx, gy, gz, limiter, x, y, z);
}

// Half the inverse-weighted jump moments added to the gradient (:218-220).
inline void add_half_jump(int i, int n, F inverse, F gx, F gy, F gz, float bx, float by, float bz,
                          W x, W y, W z) {
  x[i] = gx[i] + 0.5f * (inverse[i] * bx + inverse[i + n] * by + inverse[i + 2 * n] * bz);
  y[i] = gy[i] + 0.5f * (inverse[i + n] * bx + inverse[i + 3 * n] * by + inverse[i + 4 * n] * bz);
  z[i] = gz[i] + 0.5f * (inverse[i + 2 * n] * bx + inverse[i + 4 * n] * by + inverse[i + 5 * n] * bz);
}

// :202-221
void fv_jump_gradient(int i, I offsets, I faces, I adjacent, F sign, F dx, F dy, F dz, F inverse,
                      F px, F py, F pz, F nx, F ny, F nz, F cell, F boundary, F gx, F gy, F gz,
                      W x, W y, W z, int n) {
  float bx = 0.0f, by = 0.0f, bz = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j], o = adjacent[j];
    const bool owner_side = sign[j] > 0.0f;
    const float own = cell[i] + gx[i] * (owner_side ? px[f] : nx[f]) +
                      gy[i] * (owner_side ? py[f] : ny[f]) + gz[i] * (owner_side ? pz[f] : nz[f]);
    float other = 0.0f;
    if (o < 0) other = boundary[-o - 1];
    else
      other = cell[o] + gx[o] * (owner_side ? nx[f] : px[f]) + gy[o] * (owner_side ? ny[f] : py[f]) +
              gz[o] * (owner_side ? nz[f] : pz[f]);
    const float jump = other - own;
    bx += dx[j] * jump;
    by += dy[j] * jump;
    bz += dz[j] * jump;
  }
  add_half_jump(i, n, inverse, gx, gy, gz, bx, by, bz, x, y, z);
}

// :227-240
void fv_face_jump(int i, I owner, I neighbour, F px, F py, F pz, F nx, F ny, F nz, F cell,
                  F boundary, F gx, F gy, F gz, W forward, W reverse, int internal) {
  const int c = owner[i];
  const float own = cell[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
  float other = 0.0f;
  if (i < internal) {
    const int nb = neighbour[i];
    other = cell[nb] + gx[nb] * nx[i] + gy[nb] * ny[i] + gz[nb] * nz[i];
    reverse[i] = own - other;
  } else {
    other = boundary[i - internal];
  }
  forward[i] = other - own;
}

// :242-254
void fv_gather_jump_gradient(int i, I offsets, I faces, F sign, F dx, F dy, F dz, F inverse,
                             F forward, F reverse, F gx, F gy, F gz, W x, W y, W z, int n) {
  float bx = 0.0f, by = 0.0f, bz = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const float jump = sign[j] > 0.0f ? forward[f] : reverse[f];
    bx += dx[j] * jump;
    by += dy[j] * jump;
    bz += dz[j] * jump;
  }
  add_half_jump(i, n, inverse, gx, gy, gz, bx, by, bz, x, y, z);
}

// :258-292
void fv_paired_jump_gradient(int i, I offsets, I faces, I adjacent, F sign, F dx, F dy, F dz,
                             F inverse, F px, F py, F pz, F nx, F ny, F nz, F cell, F low,
                             F boundary, F boundary_low, F gx, F gy, F gz, F lx, F ly, F lz, W x,
                             W y, W z, int n) {
  Pair2 bx{0.0f, 0.0f}, by{0.0f, 0.0f}, bz{0.0f, 0.0f};
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j], o = adjacent[j];
    const bool owner_side = sign[j] > 0.0f;
    const Pair2 own = trace(cell[i], low[i], gx[i], gy[i], gz[i], lx[i], ly[i], lz[i],
                            owner_side ? px[f] : nx[f], owner_side ? py[f] : ny[f],
                            owner_side ? pz[f] : nz[f]);
    Pair2 other{0.0f, 0.0f};
    if (o < 0) other = {boundary[-o - 1], boundary_low[-o - 1]};
    else
      other = trace(cell[o], low[o], gx[o], gy[o], gz[o], lx[o], ly[o], lz[o],
                    owner_side ? nx[f] : px[f], owner_side ? ny[f] : py[f],
                    owner_side ? nz[f] : pz[f]);
    Pair2 jump = two_sum(other.hi, -own.hi);
    jump = product_add(jump, other.lo, 1.0f);
    jump = product_add(jump, -own.lo, 1.0f);
    bx = add_weighted(bx, dx[j], jump);
    by = add_weighted(by, dy[j], jump);
    bz = add_weighted(bz, dz[j], jump);
  }
  const float xx = .5f * inverse[i], xy = .5f * inverse[i + n], xz = .5f * inverse[i + 2 * n];
  const float yy = .5f * inverse[i + 3 * n], yz = .5f * inverse[i + 4 * n], zz = .5f * inverse[i + 5 * n];
  Pair2 a = two_sum(gx[i], lx[i]), b = two_sum(gy[i], ly[i]), c = two_sum(gz[i], lz[i]);
  a = add_weighted(add_weighted(add_weighted(a, xx, bx), xy, by), xz, bz);
  b = add_weighted(add_weighted(add_weighted(b, xy, bx), yy, by), yz, bz);
  c = add_weighted(add_weighted(add_weighted(c, xz, bx), yz, by), zz, bz);
  x[i] = a.hi;
  y[i] = b.hi;
  z[i] = c.hi;
}

// :296-299. Scatters stored inverse words back to their cells.
void fv_lsq_restore_inverse(int i, I cells, F values, W inverse, int count, int n) {
  for (int k = 0; k < 6; ++k) inverse[cells[i] + k * n] = values[i + k * count];
}

// ----------------------------------------------------- fv_transport.kern.h
// :3-10
void fv_diffusion_correction_bound(int i, I owner, I neighbour, I kind, F field, F boundary,
                                   F diffusion, W bound, int internal) {
  const float own = field[owner[i]];
  float other = own;
  if (i < internal) other = field[neighbour[i]];
  else if (kind[i - internal] == 0) other = boundary[i - internal];
  bound[i] = diffusion[i] * kabs(other - own);
}

// :12-35
void fv_transport_faces(int i, I owner, I neighbour, F weight, F px, F py, F pz, F nx, F ny, F nz,
                        F tx, F ty, F tz, F gx, F gy, F gz, F limiter, F flux, F mu, I kind,
                        W correction, int internal, int second_order, int limited_diffusion) {
  const int c = owner[i];
  float ax = gx[c], ay = gy[c], az = gz[c];
  float convection = 0.0f;
  const float m = flux[i];
  if (i < internal) {
    const int n = neighbour[i];
    const float w = weight[i];
    ax = w * ax + (1.0f - w) * gx[n];
    ay = w * ay + (1.0f - w) * gy[n];
    az = w * az + (1.0f - w) * gz[n];
    const bool forward = m >= 0.0f;
    const int up = forward ? c : n;
    const float rx = forward ? px[i] : nx[i], ry = forward ? py[i] : ny[i], rz = forward ? pz[i] : nz[i];
    convection = limiter[up] * (gx[up] * rx + gy[up] * ry + gz[up] * rz) * m;
  } else {
    const int k = kind[i - internal];
    if (k != 0) ax = ay = az = 0.0f;
    if (m > 0.0f && k != 2) convection = limiter[c] * (gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i]) * m;
  }
  float diffusion = ax * tx[i] + ay * ty[i] + az * tz[i];
  if (limited_diffusion) diffusion = kmax(-correction[i], kmin(correction[i], diffusion));
  correction[i] = mu[i] * diffusion - (second_order != 0 ? convection : 0.0f);
}

// :37-112. Row i of a transport matrix: off-diagonals -D + min(F, 0) into the
// row's own entries, an upward-rounded diagonal, boundary and deferred
// correction terms on the right-hand side (optionally carried as a float
// pair), the relaxed/pseudo-time diagonal, and the SIMPLEC consistent inverse.
void fv_transport_assemble(int i, I offsets, I faces, F sign, I adjacent, I entry, I rows, I diag,
                           F diffusion, F flux, F mu, I kind, F boundary, F correction, F source,
                           F volume, F old, F sink, F extra_diagonal, F extra_rhs, W values,
                           W rhs, W rhs_low, W diagonal, W diagonal_low, W consistent_inverse,
                           int internal, float rho, float dt, float relax, int compensated) {
  const float sink_volume = sink[i] * volume[i];
  UpperSum ap = kupper_add({sink_volume, 0.0f}, extra_diagonal[i]);
  float b = source[i] * volume[i] + extra_rhs[i];
  float carry = 0.0f;
  if (compensated) {
    const float product = source[i] * volume[i];
    const Pair2 start = two_sum(product, extra_rhs[i]);
    b = start.hi;
    carry = start.lo + std::fma(source[i], volume[i], -product);
  }
  for (int j = offsets[i]; j < offsets[i + 1]; ++j)
    if (adjacent[j] >= 0) values[entry[j]] = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j], o = adjacent[j];
    const float mass = sign[j] * flux[f];
    float d = mu[f] * diffusion[f];
    ap = kupper_add(ap, mass);
    if (o >= 0) {
      values[entry[j]] += -d + kmin(mass, 0.0f);
    } else {
      const int bf = f - internal;
      if (kind[bf] != 0) d = 0.0f;
      ap = kupper_add(ap, d);
      ap = kupper_add(ap, -kmin(mass, 0.0f));
      if (compensated) {
        const Pair2 factor = two_sum(d, -kmin(mass, 0.0f));
        const float product = factor.hi * boundary[bf];
        const Pair2 sum = two_sum(b, product);
        carry += sum.lo + std::fma(factor.hi, boundary[bf], -product) + factor.lo * boundary[bf];
        b = sum.hi;
      } else {
        b += (d - kmin(mass, 0.0f)) * boundary[bf];
      }
    }
    const float deferred = sign[j] * correction[f];
    if (compensated) {
      const Pair2 sum = two_sum(b, deferred);
      carry += sum.lo;
      b = sum.hi;
    } else {
      b += deferred;
    }
  }
  for (int j = rows[i]; j < rows[i + 1]; ++j)
    if (j != diag[i]) ap = kupper_add(ap, -values[j]);
  const float physical = kupper_value(ap);
  diagonal[i] = physical;
  if (compensated) {
    const Pair2 gap = two_sum(ap.sum, -physical);
    diagonal_low[i] = gap.hi + (gap.lo + ap.error + std::fma(sink[i], volume[i], -sink_volume));
  }
  const float pseudo_time = dt > 0.0f ? rho * volume[i] / dt : 0.0f;
  const float relaxed = kadd_up(physical, pseudo_time) / relax;
  if (compensated) {
    const Pair2 factor = two_sum(relaxed, -physical);
    const float product = factor.hi * old[i];
    const Pair2 sum = two_sum(b, product);
    carry += sum.lo + std::fma(factor.hi, old[i], -product) + factor.lo * old[i];
    b = sum.hi;
  } else {
    b += (relaxed - physical) * old[i];
  }
  values[diag[i]] = relaxed;
  rhs[i] = b;
  if (compensated) rhs_low[i] = carry;
  UpperSum denominator{relaxed, 0.0f};
  for (int j = rows[i]; j < rows[i + 1]; ++j)
    if (j != diag[i]) denominator = kupper_add(denominator, values[j]);
  const float consistent = kupper_value(denominator);
  consistent_inverse[i] = 1.0f / kmax(consistent > 0.0f ? consistent : relaxed, 1.0e-20f);
}

// :115-124
void fv_fixed_rows(int i, I rows, I diag, I mask, F fixed, W values, W rhs) {
  if (mask[i] == 0) return;
  const float scale = values[diag[i]];
  for (int j = rows[i]; j < rows[i + 1]; ++j) values[j] = 0.0f;
  values[diag[i]] = scale;
  rhs[i] = scale * fixed[i];
}

// :126-128
void fv_floor(int i, W field, float floor) { field[i] = kmax(field[i], floor); }

// :130-135
void fv_scalar_steady_measure(int i, F field, F diagonal, F residual, I excluded, W absolute,
                              W scale, int has_excluded) {
  const bool omit = has_excluded && excluded[i] != 0;
  absolute[i] = omit ? 0.0f : kabs(residual[i]);
  scale[i] = omit ? 0.0f : kabs(diagonal[i] * field[i]);
}

// :137-139
void fv_scalar_set_fixed(int i, I fixed, F value, W field) {
  if (fixed[i]) field[i] = value[i];
}

// :141-148
void fv_scalar_monitor(int i, F field, F diagonal, F residual, I excluded, F volume, W absolute,
                       W scale, W fluent_scale, float rho, float dt) {
  const bool omit = excluded[i] != 0;
  absolute[i] = omit ? 0.0f : kabs(residual[i]);
  scale[i] = omit ? 0.0f : kabs(diagonal[i] * field[i]);
  const float pseudo_time = dt > 0.0f ? rho * volume[i] / dt : 0.0f;
  fluent_scale[i] = omit ? 0.0f : kabs((diagonal[i] + pseudo_time) * field[i]);
}

// -------------------------------------------------------- fv_paired.kern.h
// Pair additions with hysteresis, kernels/prelude.h:133-171.
inline Pair2 pair_add(float hi, float lo, float d, float slack) {
  const float low = lo + d;
  if (kabs(low) <= slack) return {hi, low};
  const Pair2 s = two_sum(hi, d);
  return two_sum(s.hi, s.lo + lo);
}
// Renormalised hi + lo + d, and its offset from the held high word.
inline Pair2 pair_sum3(float hi, float lo, float d, float& offset) {
  const Pair2 base = two_sum(hi, lo), increment = two_sum(base.lo, d);
  Pair2 value = two_sum(base.hi, increment.hi);
  value = two_sum(value.hi, value.lo + increment.lo);
  const Pair2 difference = two_sum(value.hi, -hi);
  offset = difference.hi + (difference.lo + value.lo);
  return value;
}
inline Pair2 pair_add_relative(float hi, float lo, float d, float relative_slack) {
  const Pair2 base = two_sum(hi, lo), increment = two_sum(base.lo, d);
  Pair2 value = two_sum(base.hi, increment.hi);
  value = two_sum(value.hi, value.lo + increment.lo);
  const float slack = relative_slack * kmax(kabs(base.hi), kabs(value.hi));
  const Pair2 difference = two_sum(value.hi, -hi);
  const float offset = difference.hi + (difference.lo + value.lo);
  if (kabs(offset) <= slack) return {hi, offset};
  return value;
}
inline bool same_sign(float a, float b) {
  return (a > 0.0f && b > 0.0f) || (a < 0.0f && b < 0.0f) || (a == 0.0f && b == 0.0f);
}
inline Pair2 pair_add_flux(float hi, float lo, float d, float slack) {
  float offset;
  const Pair2 value = pair_sum3(hi, lo, d, offset);
  if (same_sign(hi, value.hi) && kabs(offset) <= slack) return {hi, offset};
  return value;
}
// Centroid-link jump fade factor, as in fv_pressure_faces.
inline float link_trust(float dx, float dy, float dz, float mx, float my, float mz) {
  const float d2 = dx * dx + dy * dy + dz * dz, m2 = mx * mx + my * my + mz * mz;
  return kmax(0.0f, kmin(1.0f, (4.0f * d2 - m2) / (3.0f * d2)));
}

// :5-33
void fv_compensated_pressure_jump(int i, I owner, I neighbour, F px, F py, F pz, F nx, F ny, F nz,
                                  F p, F gx, F gy, F gz, F boundary, W jump, int internal) {
  const int c = owner[i];
  Pair2 full = two_sum(p[c], i < internal ? -p[neighbour[i]] : -boundary[i - internal]);
  full = product_add(full, gx[c], px[i]);
  full = product_add(full, gy[c], py[i]);
  full = product_add(full, gz[c], pz[i]);
  if (i >= internal) {
    jump[i] = full.hi + full.lo;
    return;
  }
  const int n = neighbour[i];
  full = product_add(full, -gx[n], nx[i]);
  full = product_add(full, -gy[n], ny[i]);
  full = product_add(full, -gz[n], nz[i]);
  const float side = full.hi + full.lo;
  const float trust = link_trust(px[i] - nx[i], py[i] - ny[i], pz[i] - nz[i], 0.5f * (px[i] + nx[i]),
                                 0.5f * (py[i] + ny[i]), 0.5f * (pz[i] + nz[i]));
  if (trust == 1.0f) {
    jump[i] = side;
    return;
  }
  Pair2 link = two_sum(p[c], -p[n]);
  const float r[3][2] = {{px[i], nx[i]}, {py[i], ny[i]}, {pz[i], nz[i]}};
  F g[3] = {gx, gy, gz};
  for (int k = 0; k < 3; ++k) {
    link = product_add(link, g[k][c], 0.5f * r[k][0]);
    link = product_add(link, g[k][n], 0.5f * r[k][0]);
    link = product_add(link, g[k][c], -0.5f * r[k][1]);
    link = product_add(link, g[k][n], -0.5f * r[k][1]);
  }
  const float shared = link.hi + link.lo;
  jump[i] = shared + trust * (side - shared);
}

// :39-44
inline float trace_difference(Pair2 a, Pair2 b) {
  Pair2 d = two_sum(a.hi, -b.hi);
  d = product_add(d, a.lo, 1.0f);
  d = product_add(d, -b.lo, 1.0f);
  return d.hi + d.lo;
}
// :46-65
void fv_paired_pressure_jump(int i, I owner, I neighbour, F px, F py, F pz, F nx, F ny, F nz, F p,
                             F low, F gx, F gy, F gz, F lx, F ly, F lz, F boundary,
                             F boundary_low, W jump, int internal) {
  const int c = owner[i];
  const Pair2 own = trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[i], py[i], pz[i]);
  if (i >= internal) {
    jump[i] = trace_difference(own, {boundary[i - internal], boundary_low[i - internal]});
    return;
  }
  const int n = neighbour[i];
  const Pair2 next = trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], nx[i], ny[i], nz[i]);
  const float side = trace_difference(own, next);
  const float dx = px[i] - nx[i], dy = py[i] - ny[i], dz = pz[i] - nz[i];
  const float trust = link_trust(dx, dy, dz, 0.5f * (px[i] + nx[i]), 0.5f * (py[i] + ny[i]),
                                 0.5f * (pz[i] + nz[i]));
  if (trust == 1.0f) {
    jump[i] = side;
    return;
  }
  const Pair2 a = trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], .5f * dx, .5f * dy, .5f * dz);
  const Pair2 b = trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], -.5f * dx, -.5f * dy, -.5f * dz);
  const float shared = trace_difference(a, b);
  jump[i] = shared + trust * (side - shared);
}

// :69-85
void fv_compensated_rhie_chow(int i, I owner, I neighbour, I kind, F sx, F sy, F sz, F diffusion,
                              F volume, F diagonal, F u, F v, F w, F lu, F lv, F lw, F jump,
                              W flux, W low, int internal, float rho) {
  const bool boundary = i >= internal;
  if (boundary && (kind[i - internal] == 0 || kind[i - internal] == 3)) {
    flux[i] = 0.0f;
    low[i] = 0.0f;
    return;
  }
  const int c = owner[i];
  float rc = volume[c] / kmax(diagonal[c], 1.0e-20f);
  if (!boundary) {
    const int n = neighbour[i];
    rc = .5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
  }
  Pair2 value{0.0f, 0.0f};
  value = product_add(value, u[i], sx[i]);
  value = product_add(value, lu[i], sx[i]);
  value = product_add(value, v[i], sy[i]);
  value = product_add(value, lv[i], sy[i]);
  value = product_add(value, w[i], sz[i]);
  value = product_add(value, lw[i], sz[i]);
  if (!boundary || kind[i - internal] != 1) value = product_add(value, rc * diffusion[i], jump[i]);
  const Pair2 scaled = add_weighted({0.0f, 0.0f}, rho, value);
  flux[i] = scaled.hi;
  low[i] = scaled.lo;
}

// :89-97
void fv_paired_pressure_boundary_low(int i, I owner, I kind, F px, F py, F pz, F p, F low, F gx,
                                     F gy, F gz, F lx, F ly, F lz, F boundary, W boundary_low,
                                     int internal) {
  if (kind[i] == 2 || kind[i] == 4) {
    boundary_low[i] = 0.0f;
    return;
  }
  const int f = i + internal, c = owner[f];
  Pair2 value = trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[f], py[f], pz[f]);
  value = product_add(value, -boundary[i], 1.0f);
  boundary_low[i] = value.hi + value.lo;
}

// :98-128
void fv_paired_pressure_faces(int i, I owner, I neighbour, I kind, I poor, F px, F py, F pz, F nx,
                              F ny, F nz, F p, F low, F gx, F gy, F gz, F lx, F ly, F lz,
                              F boundary, F boundary_low, F diagonal, W face, W face_low,
                              int internal, int standard) {
  const int c = owner[i];
  Pair2 value = two_sum(p[c], low[c]);
  if (i < internal) {
    const int n = neighbour[i];
    if (standard) {
      const float wc = 1.0f / kmax(diagonal[c], 1.0e-20f), wn = 1.0f / kmax(diagonal[n], 1.0e-20f);
      Pair2 numerator{0.0f, 0.0f};
      numerator = product_add(numerator, wc, p[c]);
      numerator = product_add(numerator, wc, low[c]);
      numerator = product_add(numerator, wn, p[n]);
      numerator = product_add(numerator, wn, low[n]);
      const Pair2 denominator = two_sum(wc, wn);
      const float q = numerator.hi / denominator.hi;
      numerator = product_add(numerator, -q, denominator.hi);
      numerator = product_add(numerator, -q, denominator.lo);
      value = two_sum(q, (numerator.hi + numerator.lo) / denominator.hi);
    } else {
      const bool plain = poor[c] || poor[n];
      const Pair2 own = plain ? two_sum(p[c], low[c])
                              : trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[i], py[i], pz[i]);
      const Pair2 next = plain ? two_sum(p[n], low[n])
                               : trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], nx[i], ny[i], nz[i]);
      value = two_sum(0.5f * own.hi, 0.5f * next.hi);
      value = two_sum(value.hi, value.lo + 0.5f * own.lo + 0.5f * next.lo);
    }
  } else {
    const int b = i - internal;
    if (kind[b] == 2 || kind[b] == 4) value = two_sum(boundary[b], boundary_low[b]);
    else if (!standard && !poor[c])
      value = trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[i], py[i], pz[i]);
  }
  face[i] = value.hi;
  face_low[i] = value.lo;
}

// :140-152
void fv_paired_momentum_residual(int i, I rows, I cols, F values, I diag, F physical,
                                 F physical_low, F hi, F lo, F pressure_gradient_lo, F volume,
                                 F rhs, F rhs_low, W r) {
  float sum = rhs[i], error = rhs_low[i];
  for (int j = rows[i]; j < rows[i + 1]; ++j) {
    const float a = values[j], x = hi[cols[j]], product = a * x;
    const Pair2 s = two_sum(sum, -product);
    sum = s.hi;
    error += s.lo - std::fma(a, x, -product) - a * lo[cols[j]];
  }
  error += (values[diag[i]] - physical[i]) * lo[i] - volume[i] * pressure_gradient_lo[i];
  error -= physical_low[i] * hi[i] + physical_low[i] * lo[i];
  r[i] = sum + error;
}

// :156-170
void fv_compensated_momentum_residual(int i, I rows, I cols, F values, I diag, F physical,
                                      F physical_low, F hi, F lo, F pressure_gradient_lo,
                                      F volume, F rhs, F rhs_low, W r) {
  Pair2 value = two_sum(rhs[i], rhs_low[i]);
  for (int j = rows[i]; j < rows[i + 1]; ++j) {
    value = product_add(value, -values[j], hi[cols[j]]);
    const float low_coefficient = j == diag[i] ? physical[i] : values[j];
    value = product_add(value, -low_coefficient, lo[cols[j]]);
  }
  value = product_add(value, -physical_low[i], hi[i]);
  value = product_add(value, -physical_low[i], lo[i]);
  value = product_add(value, -volume[i], pressure_gradient_lo[i]);
  r[i] = value.hi + value.lo;
}

// :172-174
void fv_paired_accumulate(int i, F delta, W hi, W lo, float slack) {
  const Pair2 x = pair_add(hi[i], lo[i], delta[i], slack);
  hi[i] = x.hi;
  lo[i] = x.lo;
}

// :179-196
void fv_compensated_pressure_force(int i, I offsets, I faces, F sign, F sx, F sy, F sz, F volume,
                                   F pressure, F low, W x, W y, W z) {
  float a = 0.0f, b = 0.0f, c = 0.0f;
  Pair2 ax{0.0f, 0.0f}, ay{0.0f, 0.0f}, az{0.0f, 0.0f};
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const float hp = sign[j] * pressure[f], lp = sign[j] * low[f];
    a += hp * sx[f];
    b += hp * sy[f];
    c += hp * sz[f];
    ax = product_add(product_add(ax, hp, sx[f]), lp, sx[f]);
    ay = product_add(product_add(ay, hp, sy[f]), lp, sy[f]);
    az = product_add(product_add(az, hp, sz[f]), lp, sz[f]);
  }
  const float vol = volume[i];
  ax = product_add(ax, -a / vol, vol);
  ay = product_add(ay, -b / vol, vol);
  az = product_add(az, -c / vol, vol);
  x[i] = (ax.hi + ax.lo) / vol;
  y[i] = (ay.hi + ay.lo) / vol;
  z[i] = (az.hi + az.lo) / vol;
}

// :200-218
void fv_paired_velocity_boundary(int i, I owner, I kind, F sx, F sy, F sz, F u, F v, F w, F flux,
                                 F flux_low, W bu, W bv, W bw, WI scalar_kind, int internal,
                                 float rho) {
  const int f = i + internal, c = owner[f], k = kind[i];
  float vx = 0.0f, vy = 0.0f, vz = 0.0f;
  const float area2 = sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f];
  scalar_kind[i] = 0;
  if (((k == 2 || k == 4) && flux[f] >= 0.0f) || k == 3) {
    vx = u[c];
    vy = v[c];
    vz = w[c];
    scalar_kind[i] = 2;
  }
  if (k == 4 && flux[f] < 0.0f) {
    const float factor = flux_low[f] / (rho * area2);
    vx = factor * sx[f];
    vy = factor * sy[f];
    vz = factor * sz[f];
  }
  if (k == 3) {
    const float un = (vx * sx[f] + vy * sy[f] + vz * sz[f]) / area2;
    vx -= un * sx[f];
    vy -= un * sy[f];
    vz -= un * sz[f];
  }
  bu[i] = vx;
  bv[i] = vy;
  bw[i] = vz;
}

// :220-236
void fv_paired_advection(int i, I owner, I neighbour, I kind, F px, F py, F pz, F nx, F ny, F nz,
                         F field, F boundary, F gx, F gy, F gz, F limiter, F flux, F flux_low,
                         W correction, int internal, int second_order) {
  const int c = owner[i];
  int up = c;
  float value = field[c], rx = px[i], ry = py[i], rz = pz[i];
  bool reconstruct = second_order != 0;
  if (i < internal) {
    if (flux[i] < 0.0f) {
      up = neighbour[i];
      rx = nx[i];
      ry = ny[i];
      rz = nz[i];
    }
    value = field[up];
  } else {
    if (flux[i] < 0.0f) value = boundary[i - internal];
    if (flux[i] <= 0.0f || kind[i - internal] == 2) reconstruct = false;
  }
  if (reconstruct) value += limiter[up] * (gx[up] * rx + gy[up] * ry + gz[up] * rz);
  correction[i] = flux_low[i] * value;
}

// :238-247
void fv_paired_face_residual(int i, I offsets, I faces, F sign, F secondary, F advection, F stress,
                             F boundary_rhs, W residual, int turbulent) {
  float sum = residual[i], error = boundary_rhs[i];
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const float value = sign[j] * (secondary[f] - advection[f] + (turbulent ? stress[f] : 0.0f));
    const Pair2 s = two_sum(sum, value);
    sum = s.hi;
    error += s.lo;
  }
  residual[i] = sum + error;
}

// :259-290
void fv_paired_flux(int i, I owner, I neighbour, I kind, F sx, F sy, F sz, F px, F py, F pz, F nx,
                    F ny, F nz, F weights, F diagonal, F volume, F diffusion, F u, F v, F w, F p,
                    F gx, F gy, F gz, F previous, W flux, W flux_lo, int internal, float rho,
                    float slack_per_area) {
  const int c = owner[i];
  float rc = volume[c] / kmax(diagonal[c], 1.0e-20f), low = 0.0f;
  if (i < internal) {
    const int n = neighbour[i];
    const float a = weights[c] / (weights[c] + weights[n]);
    const float ux = a * u[c] + (1.0f - a) * u[n];
    const float uy = a * v[c] + (1.0f - a) * v[n];
    const float uz = a * w[c] + (1.0f - a) * w[n];
    rc = 0.5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
    const float dx = px[i] - nx[i], dy = py[i] - ny[i], dz = pz[i] - nz[i];
    const float jump = p[c] - p[n] + 0.5f * ((gx[c] + gx[n]) * dx + (gy[c] + gy[n]) * dy + (gz[c] + gz[n]) * dz);
    low = rho * (ux * sx[i] + uy * sy[i] + uz * sz[i]) + rho * rc * diffusion[i] * jump;
  } else if (kind[i - internal] == 2 || kind[i - internal] == 4) {
    const float jump = p[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
    low = rho * (u[c] * sx[i] + v[c] * sy[i] + w[c] * sz[i]) + rho * rc * diffusion[i] * jump;
  }
  const float slack = slack_per_area * ksqrt(sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i]);
  const Pair2 x = two_sum(flux[i], -previous[i]);
  const float offset = x.hi + (x.lo + low);
  if (kabs(offset) <= slack) {
    flux[i] = previous[i];
    flux_lo[i] = offset;
  } else {
    flux_lo[i] = low;
  }
}

// :292-303
void fv_paired_flux_offset(int i, F sx, F sy, F sz, F previous, W flux, W low, float slack_per_area) {
  const float slack = slack_per_area * ksqrt(sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i]);
  const Pair2 total = two_sum(flux[i], low[i]);
  const Pair2 x = two_sum(total.hi, -previous[i]);
  const float offset = x.hi + (x.lo + total.lo);
  if (same_sign(previous[i], total.hi) && kabs(offset) <= slack) {
    flux[i] = previous[i];
    low[i] = offset;
  } else {
    flux[i] = total.hi;
    low[i] = total.lo;
  }
}

// :308-318
void fv_paired_continuity(int i, I offsets, I faces, F sign, F flux, F flux_lo, W imbalance,
                          W magnitude) {
  float sum = 0.0f, error = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const Pair2 s = two_sum(sum, sign[j] * flux[f]);
    sum = s.hi;
    error += s.lo + sign[j] * flux_lo[f];
  }
  const float mass = sum + error;
  imbalance[i] = mass;
  magnitude[i] = kabs(mass);
}

// :321-323
void fv_paired_pressure_rhs(int i, F imbalance, W rhs, int reference) {
  rhs[i] = i == reference ? 0.0f : -imbalance[i];
}

// :326-333
void fv_paired_correct_flux(int i, I owner, I neighbour, F sx, F sy, F sz, F correction,
                            F coefficient, W flux, W flux_lo, int internal, float slack_per_area) {
  const float pc = correction[owner[i]];
  const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
  const float slack = slack_per_area * ksqrt(sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i]);
  const Pair2 x = pair_add(flux[i], flux_lo[i], coefficient[i] * (pc - pn), slack);
  flux[i] = x.hi;
  flux_lo[i] = x.lo;
}

// :336-347
void fv_paired_correct_velocity(int i, F volume, F inverse, F diagonal, F gx, F gy, F gz,
                                F correction, W u, W v, W w, W p, W u_lo, W v_lo, W w_lo, W p_lo,
                                float alpha, float velocity_slack, float pressure_slack) {
  const float d = volume[i] * inverse[i];
  Pair2 x = pair_add(u[i], u_lo[i], -d * gx[i], velocity_slack);
  u[i] = x.hi;
  u_lo[i] = x.lo;
  x = pair_add(v[i], v_lo[i], -d * gy[i], velocity_slack);
  v[i] = x.hi;
  v_lo[i] = x.lo;
  x = pair_add(w[i], w_lo[i], -d * gz[i], velocity_slack);
  w[i] = x.hi;
  w_lo[i] = x.lo;
  const float pressure_alpha = kmin(alpha, 0.7f * diagonal[i] * inverse[i]);
  x = pair_add(p[i], p_lo[i], pressure_alpha * correction[i], pressure_slack);
  p[i] = x.hi;
  p_lo[i] = x.lo;
}

// :349-351
void fv_paired_value(int i, F hi, F lo, W value) { value[i] = hi[i] + lo[i]; }
// :352-355
void fv_paired_vector_values(int i, F x, F y, F z, F lx, F ly, F lz, W u, W v, W w) {
  u[i] = x[i] + lx[i];
  v[i] = y[i] + ly[i];
  w[i] = z[i] + lz[i];
}

// :356-363
void fv_compensated_correct_flux(int i, I owner, I neighbour, F sx, F sy, F sz, F correction,
                                 F coefficient, W flux, W flux_lo, int internal,
                                 float slack_per_area) {
  const int c = owner[i];
  const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
  const float slack = slack_per_area * ksqrt(sx[i] * sx[i] + sy[i] * sy[i] + sz[i] * sz[i]);
  const Pair2 x = pair_add_flux(flux[i], flux_lo[i], coefficient[i] * (correction[c] - pn), slack);
  flux[i] = x.hi;
  flux_lo[i] = x.lo;
}

// :365-377
void fv_compensated_correct_velocity(int i, F volume, F inverse, F diagonal, F gx, F gy, F gz,
                                     F correction, W u, W v, W w, W p, W u_lo, W v_lo, W w_lo,
                                     W p_lo, float alpha, float velocity_slack,
                                     float pressure_slack) {
  const float d = volume[i] * inverse[i];
  Pair2 x = pair_add(u[i], u_lo[i], -d * gx[i], velocity_slack);
  u[i] = x.hi;
  u_lo[i] = x.lo;
  x = pair_add(v[i], v_lo[i], -d * gy[i], velocity_slack);
  v[i] = x.hi;
  v_lo[i] = x.lo;
  x = pair_add(w[i], w_lo[i], -d * gz[i], velocity_slack);
  w[i] = x.hi;
  w_lo[i] = x.lo;
  const float pressure_alpha = kmin(alpha, 0.7f * diagonal[i] * inverse[i]);
  x = pair_add_relative(p[i], p_lo[i], pressure_alpha * correction[i], pressure_slack);
  p[i] = x.hi;
  p_lo[i] = x.lo;
}

// ---------------------------------------------------------- fv_rans.kern.h
// :4-10
void fv_backflow_pressure(int i, I owner, I total, F mass, F u, F v, F w, W boundary, int internal,
                          float rho) {
  const int f = i + internal, c = owner[f];
  if (total[i] != 0 && mass[f] < 0.0f) boundary[i] -= 0.5f * rho * (u[c] * u[c] + v[c] * v[c] + w[c] * w[c]);
}

// :12-17
void fv_effective_viscosity(int i, I owner, I neighbour, I kind, F weight, F mut, W mu,
                            int internal, float molecular) {
  float turbulent = mut[owner[i]];
  if (i < internal) turbulent = weight[i] * turbulent + (1.0f - weight[i]) * mut[neighbour[i]];
  mu[i] = i >= internal && kind[i - internal] == 0 ? 0.0f : molecular + turbulent;
}

// :21-37
void fv_transpose_stress(int i, I owner, I neighbour, I kind, F weight, F mu, F sx, F sy, F sz,
                         F ux, F uy, F uz, F vx, F vy, F vz, F wx, F wy, F wz, W x, W y, W z,
                         int internal) {
  const int c = owner[i];
  float g[9] = {ux[c], uy[c], uz[c], vx[c], vy[c], vz[c], wx[c], wy[c], wz[c]};
  if (i < internal) {
    const int n = neighbour[i];
    const float t = weight[i], s = 1.0f - t;
    const float h[9] = {ux[n], uy[n], uz[n], vx[n], vy[n], vz[n], wx[n], wy[n], wz[n]};
    for (int k = 0; k < 9; ++k) g[k] = t * g[k] + s * h[k];
  }
  const float m = i >= internal && kind[i - internal] == 3 ? 0.0f : mu[i];
  x[i] = m * (g[0] * sx[i] + g[3] * sy[i] + g[6] * sz[i]);
  y[i] = m * (g[1] * sx[i] + g[4] * sy[i] + g[7] * sz[i]);
  z[i] = m * (g[2] * sx[i] + g[5] * sy[i] + g[8] * sz[i]);
}

// :38-46
void fv_stress_source(int i, I offsets, I faces, F sign, F volume, F x, F y, F z, W su, W sv, W sw) {
  float a = 0.0f, b = 0.0f, c = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    a += sign[j] * x[f];
    b += sign[j] * y[f];
    c += sign[j] * z[f];
  }
  su[i] += a / volume[i];
  sv[i] += b / volume[i];
  sw[i] += c / volume[i];
}

// :47-72. Implicit symmetry and wall-function terms of one velocity component.
void fv_boundary_momentum(int i, I offsets, I faces, I kind, F sx, F sy, F sz, F coefficient, F mu,
                          F diffusion, F u, F v, F w, F bu, F bv, F bw, W diagonal, W rhs,
                          int internal, int component, int wall_function) {
  float ap = 0.0f, b = 0.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    if (f < internal) continue;
    const int bf = f - internal, k = kind[bf];
    if (k != 3 && !(k == 0 && wall_function)) continue;
    const float area = ksqrt(sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f]);
    const float ex = sx[f] / area, ey = sy[f] / area, ez = sz[f] / area;
    const float nc = component == 0 ? ex : (component == 1 ? ey : ez);
    const float uc = component == 0 ? u[i] : (component == 1 ? v[i] : w[i]);
    const float wc = component == 0 ? bu[bf] : (component == 1 ? bv[bf] : bw[bf]);
    if (k == 3) {
      const float d = mu[f] * diffusion[f] * nc * nc;
      ap += d;
      b -= mu[f] * diffusion[f] * nc * (ex * u[i] + ey * v[i] + ez * w[i] - nc * uc);
      continue;
    }
    const float c = area * coefficient[bf];
    ap += c * (1.0f - nc * nc);
    b += c * (wc - nc * (ex * bu[bf] + ey * bv[bf] + ez * bw[bf]) + nc * (ex * u[i] + ey * v[i] + ez * w[i] - nc * uc));
  }
  diagonal[i] = ap;
  rhs[i] = b;
}

// :75-78
void fv_momentum_diagonal(int i, F ax, F ay, F az, F ix, F iy, F iz, W physical, W inverse) {
  physical[i] = (ax[i] + ay[i] + az[i]) / 3.0f;
  inverse[i] = (ix[i] + iy[i] + iz[i]) / 3.0f;
}

// :79-83
void fv_sst_force(int i, F selected, F sx, F sy, F sz, F tx, F ty, F tz, W x, W y, W z, int internal) {
  const int f = i + internal;
  const float a = -selected[i] * ksqrt(sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f]);
  x[i] = a * tx[i];
  y[i] = a * ty[i];
  z[i] = a * tz[i];
}

// :86-101. Turbulence boundary values: spec 0 direct k/omega, 1 intensity and
// length scale, 2 intensity and viscosity ratio.
void fv_turbulence_boundary(int i, I owner, I kind, I spec, F a, F b, F mass, F bu, F bv, F bw,
                            F k, F omega, F wall_omega, W bk, W bo, WI transport_kind,
                            int internal, float rho, float mu) {
  const int f = i + internal, c = owner[f], type = kind[i];
  float kinetic = k[c], dissipation = omega[c];
  transport_kind[i] = 1;
  if (type == 1 || ((type == 2 || type == 4) && mass[f] < 0.0f)) {
    kinetic = a[i];
    dissipation = b[i];
    if (spec[i] != 0) {
      kinetic = 1.5f * a[i] * a[i] * (bu[i] * bu[i] + bv[i] * bv[i] + bw[i] * bw[i]);
      dissipation = spec[i] == 1 ? ksqrt(kinetic) / (0.5477225575f * b[i]) : rho * kinetic / (mu * b[i]);
    }
    transport_kind[i] = 0;
  }
  bk[i] = kinetic;
  bo[i] = type == 0 ? wall_omega[i] : dissipation;
}

// -------------------------------------------------------------- sst.kern.h
// Coefficient blends and limiters, :5-38. Transcendentals are the C++
// library's float overloads, as in the reference's CPU build (prelude.h:59-64).
inline float sst_inv_sigma_k(float f1) { return f1 / 1.176f + (1.0f - f1); }
inline float sst_inv_sigma_w(float f1) { return f1 / 2.0f + (1.0f - f1) / 1.168f; }
inline float sst_beta(float f1) { return 0.075f * f1 + 0.0828f * (1.0f - f1); }
inline float sst_gamma(float f1, int set_2003) {
  const float g1 = set_2003 ? 5.0f / 9.0f : 0.075f / 0.09f - 0.41f * 0.41f / (2.0f * 0.3f);
  const float g2 = set_2003 ? 0.44f : 0.0828f / 0.09f - 0.41f * 0.41f / (1.168f * 0.3f);
  return f1 * g1 + (1.0f - f1) * g2;
}
inline float sst_alpha_star(float ret, int low_re) { return low_re ? 1.0f - 0.976f / (1.0f + ret / 6.0f) : 1.0f; }
inline float sst_beta_star(float ret, int low_re) {
  if (!low_re) return 0.09f;
  const float r = kmin(ret / 8.0f, 100.0f), r4 = r * r * r * r;
  return 0.09f * ((4.0f / 15.0f) + r4) / (1.0f + r4);
}
inline float sst_power_blend(float a, float b, float p) {
  const float hi = kmax(a, b), lo = kmin(a, b);
  return hi == 0.0f ? 0.0f : hi * std::pow(1.0f + std::pow(lo / hi, p), 1.0f / p);
}
inline float sst_finite_ratio(float numerator, float denominator) {
  return kmax(-1e38f, kmin(numerator / denominator, 1e38f));
}

// :40-50
void sst_rates(int i, F ux, F uy, F uz, F vx, F vy, F vz, F wx, F wy, F wz, W strain, W vorticity) {
  float xy = uy[i] + vx[i], xz = uz[i] + wx[i], yz = vz[i] + wy[i];
  strain[i] = ksqrt(2.0f * (ux[i] * ux[i] + vy[i] * vy[i] + wz[i] * wz[i]) + xy * xy + xz * xz + yz * yz);
  xy = uy[i] - vx[i];
  xz = uz[i] - wx[i];
  yz = vz[i] - wy[i];
  vorticity[i] = ksqrt(xy * xy + xz * xz + yz * yz);
}

// :52-70
void sst_blending(int i, F rho, F mu, F k, F omega, F distance, F kx, F ky, F kz, F ox, F oy, F oz,
                  W f1, W f2, W cdkw, float cd_floor, float k_floor, float omega_floor) {
  const float ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor), y = distance[i];
  const float cd = sst_finite_ratio(2.0f * rho[i] * (kx[i] * ox[i] + ky[i] * oy[i] + kz[i] * oz[i]), 1.168f * w);
  const float a = ksqrt(ke) / (0.09f * w * y);
  const float b = 500.0f * (mu[i] / rho[i]) / (y * y * w);
  const float c = 4.0f * rho[i] * ke / (1.168f * kmax(cd, cd_floor) * y * y);
  const float p1 = kmin(kmin(kmax(a, b), c), 10.0f);
  const float p2 = kmin(kmax(2.0f * a, b), 10.0f);
  f1[i] = std::tanh(p1 * p1 * p1 * p1);
  f2[i] = std::tanh(p2 * p2);
  cdkw[i] = cd;
}

// :72-88
void sst_viscosity(int i, F rho, F mu, F k, F omega, F strain, F vorticity, F f1, F f2, W mut,
                   W gamma_k, W gamma_w, int low_re, int use_vorticity, float max_ratio,
                   float k_floor, float omega_floor) {
  const float ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor);
  const float ret = (rho[i] / mu[i]) * (ke / w);
  const float alpha = sst_alpha_star(ret, low_re);
  const float rate = use_vorticity ? vorticity[i] : strain[i];
  float mt = rho[i] * ke / kmax(w / alpha, rate * f2[i] / 0.31f);
  mt = kmin(mt, max_ratio * mu[i]);
  mut[i] = mt;
  gamma_k[i] = mu[i] + mt * sst_inv_sigma_k(f1[i]);
  gamma_w[i] = mu[i] + mt * sst_inv_sigma_w(f1[i]);
}

// :90-103
void sst_blending_viscosity(int i, F rho, F mu, F k, F omega, F distance, F kx, F ky, F kz, F ox,
                            F oy, F oz, F strain, F vorticity, W f1, W f2, W cdkw, W mut,
                            W gamma_k, W gamma_w, float cd_floor, float k_floor, float omega_floor,
                            int low_re, int use_vorticity, float max_ratio) {
  sst_blending(i, rho, mu, k, omega, distance, kx, ky, kz, ox, oy, oz, f1, f2, cdkw, cd_floor, k_floor,
               omega_floor);
  sst_viscosity(i, rho, mu, k, omega, strain, vorticity, f1, f2, mut, gamma_k, gamma_w, low_re,
                use_vorticity, max_ratio, k_floor, omega_floor);
}

// :105-151. Blended wall law on wall face i (boundary index).
void sst_wall(int i, I owner, I wall, int n_internal, F ax, F ay, F az, F ux, F uy, F uz, F wall_ux,
              F wall_uy, F wall_uz, F dy, F rho, F mu, F k, F omega, W tau, W shear_x, W shear_y,
              W shear_z, W u_star, W u_tau, W y_plus, W y_plus_utau, W production,
              W shear_coefficient, float k_floor, float omega_floor) {
  tau[i] = shear_x[i] = shear_y[i] = shear_z[i] = 0.0f;
  u_star[i] = u_tau[i] = y_plus[i] = y_plus_utau[i] = 0.0f;
  production[i] = 0.0f;
  shear_coefficient[i] = 0.0f;
  if (!wall[i]) return;
  const int f = n_internal + i, c = owner[f];
  const float r = rho[c], m = mu[c], d = dy[i], ke = kmax(k[c], k_floor);
  const float w = kmax(omega[c], omega_floor);
  const float length = ksqrt(ax[f] * ax[f] + ay[f] * ay[f] + az[f] * az[f]);
  const float inv = 1.0f / length;
  const float nx = inv * ax[f], ny = inv * ay[f], nz = inv * az[f];
  const float rx = ux[c] - wall_ux[i], ry = uy[c] - wall_uy[i], rz = uz[c] - wall_uz[i];
  const float along = rx * nx + ry * ny + rz * nz;
  const float tx = rx - along * nx, ty = ry - along * ny, tz = rz - along * nz;
  const float speed = ksqrt(tx * tx + ty * ty + tz * tz);
  const float us = ksqrt((m / r) * speed / d + 0.3f * ke);
  const float yp = (r / m) * us * d;
  const float log_up = std::log(9.793f * kmax(yp, 0.2f)) / 0.4187f;
  const float ratio = kmin(yp, log_up) / kmax(yp, log_up);
  const float blend = std::pow(1.0f + ratio * ratio * ratio * ratio, 0.25f);
  const float coefficient = (yp <= log_up ? m / d : r * us / log_up) * blend;
  const float stress = coefficient * speed;
  shear_coefficient[i] = coefficient;
  const float ut = ksqrt(stress / r);
  tau[i] = stress;
  u_star[i] = us;
  u_tau[i] = ut;
  y_plus[i] = yp;
  y_plus_utau[i] = (r / m) * ut * d;
  if (speed > 0.0f) {
    shear_x[i] = (-coefficient) * tx;
    shear_y[i] = (-coefficient) * ty;
    shear_z[i] = (-coefficient) * tz;
  }
  const float p_lam = (r * ke / w) * (stress / m) * (stress / m);
  const float p_log = yp > 0.0f ? (stress * stress / m) / (0.4187f * yp) : 0.0f;
  const float hi = kmax(p_lam, p_log), lo = kmin(p_lam, p_log);
  production[i] = hi > 0.0f ? lo / (1.0f + lo / hi) : 0.0f;
}

// :153-169
void sst_wall_omega(int i, I owner, I wall, int n_internal, F dy, F rho, F mu, F u_star,
                    W fixed_omega, W face_omega) {
  fixed_omega[i] = 0.0f;
  face_omega[i] = 0.0f;
  if (!wall[i]) return;
  const int c = owner[n_internal + i];
  const float m = mu[c], r = rho[c], d = dy[i], us = u_star[i];
  const float w_lam = 2.0f * (m / r) / (0.072f * d * d);
  const float w_log = us / (0.3f * 0.4187f * d);
  fixed_omega[i] = sst_power_blend(w_lam, w_log, 1.3f);
  face_omega[i] = sst_power_blend(3.0f * w_lam, 2.0f * w_log, 1.3f);
}

// :171-190
void sst_wall_cells(int i, I offsets, I faces, I wall, F dy, F face_omega, F face_production,
                    int n_internal, WI fixed_mask, WI selected_face, W fixed_omega, W production) {
  int selected = -1;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int b = faces[j] - n_internal;
    if (b >= 0 && wall[b] &&
        (selected < 0 || dy[b] < dy[selected] || (dy[b] == dy[selected] && b < selected)))
      selected = b;
  }
  selected_face[i] = selected;
  fixed_mask[i] = selected >= 0 ? 1 : 0;
  fixed_omega[i] = selected >= 0 ? face_omega[selected] : 0.0f;
  production[i] = selected >= 0 ? face_production[selected] : 0.0f;
}

// :192-236
void sst_sources(int i, F rho, F mu, F k, F omega, F strain, F vorticity, F f1, F cdkw, F mut,
                 I wall_mask, F wall_pk, W k_source, W k_sink, W omega_source, W omega_sink,
                 W production_k, int low_re, int gamma_2003, int production_mode, int limiter,
                 int limit_omega, int wall_override, float clip, int limit_cross, float cross_clip,
                 float k_floor, float omega_floor, float omega_molecular_ratio) {
  const float r = rho[i], ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor);
  const float ret = (r / mu[i]) * (ke / w);
  const float beta_star = sst_beta_star(ret, low_re);
  const float rate2 = production_mode == 1 ? strain[i] * vorticity[i]
                      : (production_mode == 2 ? vorticity[i] * vorticity[i] : strain[i] * strain[i]);
  float raw_pk = mut[i] * rate2;
  if (wall_override && wall_mask[i]) raw_pk = wall_pk[i];
  const float cap = clip * r * beta_star * ke * w;
  const float pk = limiter ? kmin(raw_pk, cap) : raw_pk;
  const float pw_k = limiter && limit_omega ? pk : raw_pk;
  float gamma = sst_gamma(f1[i], gamma_2003);
  if (low_re) gamma *= 1.0f - (8.0f / 9.0f) / (1.0f + ret / 2.95f);
  float pw = mut[i] > 0.0f ? gamma * r * (pw_k / mut[i]) : gamma * r * rate2;
  if (omega_molecular_ratio > 0.0f) pw = gamma * r * (pw_k / (mut[i] + omega_molecular_ratio * mu[i]));
  float cross = (1.0f - f1[i]) * cdkw[i];
  if (limit_cross) {
    const float bound = (cross_clip * r * sst_beta(f1[i]) * w) * w;
    cross = kmax(-bound, kmin(cross, bound));
  }
  k_source[i] = pk;
  k_sink[i] = r * beta_star * w;
  const float destruction = r * sst_beta(f1[i]) * w;
  omega_source[i] = pw + kmax(cross, 0.0f) + destruction * w;
  omega_sink[i] = 2.0f * destruction + sst_finite_ratio(kmax(-cross, 0.0f), w);
  production_k[i] = pk;
}

// :238-253
void sst_face_diffusion(int i, I owner, I neighbour, F weight, I wall, F mu, F gamma_k, F gamma_w,
                        int n_internal, W face_k, W face_w) {
  const int c = owner[i];
  if (i < n_internal) {
    const int n = neighbour[i];
    const float a = weight[i];
    face_k[i] = a * gamma_k[c] + (1.0f - a) * gamma_k[n];
    face_w[i] = a * gamma_w[c] + (1.0f - a) * gamma_w[n];
  } else {
    const bool on_wall = wall[i - n_internal] != 0;
    face_k[i] = on_wall ? 0.0f : gamma_k[c];
    face_w[i] = on_wall ? mu[c] : gamma_w[c];
  }
}

// :255-259
void sst_sustain(int i, F rho, F f1, W k_source, W omega_source, float ambient_k, float ambient_omega) {
  k_source[i] += 0.09f * rho[i] * ambient_omega * ambient_k;
  omega_source[i] += sst_beta(f1[i]) * rho[i] * ambient_omega * ambient_omega;
}

// ------------------------------------------------- fv_fluent_schemes.kern.h
// :10-12. SIMPLE's correction coefficient: the inverse relaxed diagonal.
void fv_simple_inverse(int i, I diag, F values, W inverse) {
  inverse[i] = 1.0f / kmax(values[diag[i]], 1.0e-20f);
}

// :28-39. Interior faces weigh the cell pressures by inverse momentum
// diagonals; unprescribed boundary faces (not kinds 2, 4) take the cell value.
void fv_standard_pressure_faces(int i, I owner, I neighbour, I kind, F p, F diagonal, W face,
                                W boundary, int internal) {
  const int c = owner[i];
  if (i < internal) {
    const int nb = neighbour[i];
    const float wc = 1.0f / kmax(diagonal[c], 1.0e-20f);
    const float wn = 1.0f / kmax(diagonal[nb], 1.0e-20f);
    face[i] = (p[c] * wc + p[nb] * wn) / (wc + wn);
    return;
  }
  const int b = i - internal;
  if (kind[b] == 2 || kind[b] == 4) return;
  face[i] = p[c];
  boundary[b] = p[c];
}

// :47-50. Unprescribed boundary values hold the owner's pressure.
void fv_cell_pressure_boundary(int i, I owner, I kind, F p, W boundary, int internal) {
  if (kind[i] != 2 && kind[i] != 4) boundary[i] = p[owner[i + internal]];
}

// :54-59. Wall force per boundary face to the fluid's wall stress.
void fv_force_to_shear(int i, F mask, F ax, F ay, F az, F fx, F fy, F fz, W sx, W sy, W sz,
                       int internal) {
  const int f = i + internal;
  const float a = ksqrt(ax[f] * ax[f] + ay[f] * ay[f] + az[f] * az[f]);
  const float scale = mask[i] > 0.0f ? -1.0f / a : 0.0f;
  sx[i] = scale * fx[i];
  sy[i] = scale * fy[i];
  sz[i] = scale * fz[i];
}

// :70-84. A flagged cell's side of a face pressure is its cell value.
void fv_poor_cell_pressure_faces(int i, I owner, I neighbour, I kind, I poor, F px, F py, F pz,
                                 F nx, F ny, F nz, F p, F gx, F gy, F gz, W face, W boundary,
                                 int internal) {
  const int c = owner[i];
  if (i >= internal) {
    const int b = i - internal;
    if (poor[c] == 0 || kind[b] == 2 || kind[b] == 4) return;
    face[i] = p[c];
    boundary[b] = p[c];
    return;
  }
  const int nb = neighbour[i];
  if (poor[c] == 0 && poor[nb] == 0) return;
  const float pp = poor[c] != 0 ? p[c] : p[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
  const float pn = poor[nb] != 0 ? p[nb] : p[nb] + gx[nb] * nx[i] + gy[nb] * ny[i] + gz[nb] * nz[i];
  face[i] = 0.5f * (pp + pn);
}

// :93-97. Mixed-order interior faces use both cell pressures.
void fv_poor_pressure_interfaces(int i, I owner, I neighbour, I poor, F p, W face) {
  const int c = owner[i], nb = neighbour[i];
  if (poor[c] != poor[nb]) face[i] = 0.5f * (p[c] + p[nb]);
}

// ---------------------------------------------------------------- dispatch
template <class T>
T from_arg(void* const* a, std::size_t k) {
  if constexpr (std::is_pointer_v<T>) return static_cast<T>(a[k]);
  else {
    T v;
    std::memcpy(&v, a[k], sizeof v);
    return v;
  }
}
template <class... P, std::size_t... K>
void run(void (*body)(int, P...), int n, void* const* a, std::index_sequence<K...>) {
  const std::tuple<P...> args{from_arg<P>(a, K)...};
  for (int i = 0; i < n; ++i) std::apply([&](P... p) { body(i, p...); }, args);
}
template <class T>
char code() {
  if constexpr (std::is_pointer_v<T>) return 'b';
  else if constexpr (std::is_same_v<T, float>) return 'f';
  else if constexpr (std::is_same_v<T, unsigned>) return 'u';
  else return 'i';
}
struct Entry {
  void (*run)(int, void* const*);
  std::string signature;
};
template <class... P>
Entry entry(void (*)(int, P...), void (*runner)(int, void* const*)) {
  return {runner, std::string{code<P>()...}};
}
#define KERNEL(name)                                                                     \
  {                                                                                      \
    #name, entry(&name, [](int n, void* const* a) {                                      \
      run(&name, n, a, std::make_index_sequence<std::tuple_size_v<decltype(args_of(&name))>>{}); \
    })                                                                                   \
  }
template <class... P>
std::tuple<P...> args_of(void (*)(int, P...));

const std::map<std::string, Entry>& table() {
  static const std::map<std::string, Entry> t = {
      KERNEL(fv_velocity_boundary), KERNEL(fv_pressure_boundary), KERNEL(fv_pressure_faces),
      KERNEL(fv_momentum_source), KERNEL(fv_momentum_faces), KERNEL(fv_convective_outflow),
      KERNEL(fv_peclet_side_velocity), KERNEL(fv_rhie_chow), KERNEL(fv_correct_flux),
      KERNEL(fv_correct_velocity), KERNEL(fv_monitor_cell), KERNEL(fv_speed),
      KERNEL(fv_continuity), KERNEL(fv_wall_pressure_force), KERNEL(fv_wall_viscous_force),
      KERNEL(fv_fluent_momentum_scale), KERNEL(fv_physical_time_terms),
      KERNEL(fv_face_time_defect), KERNEL(fv_face_time_flux), KERNEL(fv_cell_mass_error),
      KERNEL(fv_gradient), KERNEL(fv_paired_gradient), KERNEL(fv_range_shift),
      KERNEL(fv_limiter), KERNEL(fv_paired_limiter), KERNEL(fv_paired_gradient_limiter),
      KERNEL(fv_vector_gradient), KERNEL(fv_scale_gradient), KERNEL(fv_limiter_scale),
      KERNEL(fv_paired_limiter_scale), KERNEL(fv_paired_gradient_limiter_scale),
      KERNEL(fv_jump_gradient), KERNEL(fv_face_jump), KERNEL(fv_gather_jump_gradient),
      KERNEL(fv_paired_jump_gradient), KERNEL(fv_lsq_restore_inverse),
      KERNEL(fv_diffusion_correction_bound), KERNEL(fv_transport_faces),
      KERNEL(fv_transport_assemble), KERNEL(fv_fixed_rows), KERNEL(fv_floor),
      KERNEL(fv_scalar_steady_measure), KERNEL(fv_scalar_set_fixed), KERNEL(fv_scalar_monitor),
      KERNEL(fv_compensated_pressure_jump), KERNEL(fv_paired_pressure_jump),
      KERNEL(fv_compensated_rhie_chow), KERNEL(fv_paired_pressure_boundary_low),
      KERNEL(fv_paired_pressure_faces), KERNEL(fv_paired_momentum_residual),
      KERNEL(fv_compensated_momentum_residual), KERNEL(fv_paired_accumulate),
      KERNEL(fv_compensated_pressure_force), KERNEL(fv_paired_velocity_boundary),
      KERNEL(fv_paired_advection), KERNEL(fv_paired_face_residual), KERNEL(fv_paired_flux),
      KERNEL(fv_paired_flux_offset), KERNEL(fv_paired_continuity), KERNEL(fv_paired_pressure_rhs),
      KERNEL(fv_paired_correct_flux), KERNEL(fv_paired_correct_velocity), KERNEL(fv_paired_value),
      KERNEL(fv_paired_vector_values), KERNEL(fv_compensated_correct_flux),
      KERNEL(fv_compensated_correct_velocity),
      KERNEL(fv_backflow_pressure), KERNEL(fv_effective_viscosity), KERNEL(fv_transpose_stress),
      KERNEL(fv_stress_source), KERNEL(fv_boundary_momentum), KERNEL(fv_momentum_diagonal),
      KERNEL(fv_sst_force), KERNEL(fv_turbulence_boundary),
      KERNEL(sst_rates), KERNEL(sst_blending), KERNEL(sst_viscosity), KERNEL(sst_blending_viscosity),
      KERNEL(sst_wall), KERNEL(sst_wall_omega), KERNEL(sst_wall_cells), KERNEL(sst_sources),
      KERNEL(sst_face_diffusion), KERNEL(sst_sustain),
      KERNEL(fv_simple_inverse), KERNEL(fv_standard_pressure_faces),
      KERNEL(fv_cell_pressure_boundary), KERNEL(fv_force_to_shear),
      KERNEL(fv_poor_cell_pressure_faces), KERNEL(fv_poor_pressure_interfaces),
  };
  return t;
}

}  // namespace
}  // namespace oracle::fv

extern "C" {
// Runs kernel `name` for i in [0, n); args[k] points at the k-th parameter
// (a buffer, or a 4-byte scalar). Returns 0, or 1 for an unknown name.
int oracle_fv_each(const char* name, int n, void* const* args) {
  const auto& t = oracle::fv::table();
  const auto it = t.find(name);
  if (it == t.end()) return 1;
  it->second.run(n, args);
  return 0;
}
// Parameter kinds of kernel `name` ('b' buffer, 'f' float, 'i' int), or NULL.
const char* oracle_fv_signature(const char* name) {
  const auto& t = oracle::fv::table();
  const auto it = t.find(name);
  return it == t.end() ? nullptr : it->second.signature.c_str();
}
// Number of kernels and the name of the k-th, for enumeration.
int oracle_fv_count() { return int(oracle::fv::table().size()); }
const char* oracle_fv_name(int k) {
  const auto& t = oracle::fv::table();
  auto it = t.begin();
  for (int j = 0; j < k && it != t.end(); ++j) ++it;
  return it == t.end() ? nullptr : it->first.c_str();
}
}
