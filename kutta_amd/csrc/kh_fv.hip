// Finite-volume kernels of the SIMPLEC pressure path for gfx950. One lane
// owns one cell and gathers over its faces; each lane writes only its own CSR
// row, so no atomics are needed and results are deterministic.
#include <cstddef>
#include <type_traits>
#include <utility>

#include "kh_device_math.hpp"

namespace kh {
namespace {

__device__ __forceinline__ int cell_index() { return int(xcd_block() * kBlock + threadIdx.x); }

// kernels/fv_flow.kern.h:164-183 — pressure-correction M-matrix row: clear the
// row's neighbour entries, subtract each face coefficient from its (possibly
// shared) entry, eliminate the reference cell's row and column, and build the
// diagonal as the upward-rounded sum of the stored off-diagonals plus boundary
// anchors; b = -sum sign * flux.
__global__ void __launch_bounds__(kBlock) k_fv_pressure_assemble(int n,
    const int* __restrict__ offsets, const int* __restrict__ faces,
    const float* __restrict__ sign, const int* __restrict__ adjacent,
    const int* __restrict__ entry, const int* __restrict__ rows, const int* __restrict__ diag,
    const float* __restrict__ coefficient, const float* __restrict__ flux,
    float* __restrict__ values, float* __restrict__ rhs, int reference) {
  const int i = cell_index();
  if (i >= n) return;
  const int begin = offsets[i], end = offsets[i + 1];
  for (int j = begin; j < end; ++j)
    if (adjacent[j] >= 0) values[entry[j]] = 0.0f;
  UpperSum ap{0.0f, 0.0f};
  float b = 0.0f;
  for (int j = begin; j < end; ++j) {
    const int f = faces[j], nb = adjacent[j];
    const float d = coefficient[f];
    b -= sign[j] * flux[f];
    if (nb < 0 || nb == reference) ap = upper_add(ap, d);
    if (nb >= 0 && nb != reference && i != reference) values[entry[j]] -= d;
  }
  const int di = diag[i];
  for (int j = rows[i]; j < rows[i + 1]; ++j)
    if (j != di) ap = upper_add(ap, -values[j]);
  values[di] = i == reference ? 1.0f : upper_value(ap);
  rhs[i] = i == reference ? 0.0f : b;
}

// ------------------------------------------------- one lane per face or cell
// Each body below is a functor whose operator()(i, ...) takes the reference
// kernel's parameters in order; k_each runs it for i in [0, n) and the table
// derives the argument unpacking and signature from the operator's type.
using CI = const int* __restrict__;
using CF = const float* __restrict__;
using WF = float* __restrict__;
using WI = int* __restrict__;

// kernels/fv_flow.kern.h:5-32. Boundary kinds: 0 wall, 1 velocity inlet,
// 2 outlet with prescribed backflow, 3 symmetry, 4 pressure outlet.
struct FvVelocityBoundary {
  __device__ void operator()(int i, CI owner, CI kind, CF sx, CF sy, CF sz, CF u, CF v, CF w,
                             CF pu, CF pv, CF pw, CF flux, WF bu, WF bv, WF bw, WI scalar_kind,
                             int internal, float rho) const {
    const int f = i + internal, c = owner[f], k = kind[i];
    const float ax = sx[f], ay = sy[f], az = sz[f], m = flux[f];
    const float area2 = ax * ax + ay * ay + az * az;
    float vx = pu[i], vy = pv[i], vz = pw[i];
    int sk = 0;
    if ((k == 2 || k == 4) && m >= 0.0f) {
      vx = u[c]; vy = v[c]; vz = w[c];
      sk = 2;
    }
    if (k == 4 && m < 0.0f) {
      const float area = sqrtf(area2), speed = m / (rho * area);
      vx = speed * (ax / area); vy = speed * (ay / area); vz = speed * (az / area);
    }
    if (k == 3) {
      vx = u[c]; vy = v[c]; vz = w[c];
      sk = 2;
    }
    if (k == 0 || k == 3) {
      const float un = (vx * ax + vy * ay + vz * az) / area2;
      vx -= un * ax; vy -= un * ay; vz -= un * az;
    }
    scalar_kind[i] = sk;
    bu[i] = vx; bv[i] = vy; bw[i] = vz;
  }
};
// :34-41
struct FvPressureBoundary {
  __device__ void operator()(int i, CI owner, CI kind, CF rx, CF ry, CF rz, CF p, CF gx, CF gy,
                             CF gz, CF prescribed, WF boundary, int internal) const {
    const int f = i + internal, k = kind[i];
    if (k == 2 || k == 4) {
      boundary[i] = prescribed[i];
    } else {
      const int c = owner[f];
      boundary[i] = p[c] + gx[c] * rx[f] + gy[c] * ry[f] + gz[c] * rz[f];
    }
  }
};
// :43-65
struct FvPressureFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CF px, CF py, CF pz, CF nx, CF ny,
                             CF nz, CF p, CF gx, CF gy, CF gz, CF boundary, WF face, WF jump,
                             int internal) const {
    const int c = owner[i];
    const float ox = px[i], oy = py[i], oz = pz[i];
    const float sp = p[c] + gx[c] * ox + gy[c] * oy + gz[c] * oz;
    if (i >= internal) {
      const float pb = boundary[i - internal];
      face[i] = pb;
      jump[i] = sp - pb;
      return;
    }
    const int n = neighbour[i];
    const float qx = nx[i], qy = ny[i], qz = nz[i];
    const float sn = p[n] + gx[n] * qx + gy[n] * qy + gz[n] * qz;
    face[i] = 0.5f * (sp + sn);
    const float dx = ox - qx, dy = oy - qy, dz = oz - qz;
    const float mx = 0.5f * (ox + qx), my = 0.5f * (oy + qy), mz = 0.5f * (oz + qz);
    const float d2 = dx * dx + dy * dy + dz * dz, m2 = mx * mx + my * my + mz * mz;
    const float trust = fmaxf(0.0f, fminf(1.0f, (4.0f * d2 - m2) / (3.0f * d2)));
    const float link = p[c] - p[n] + 0.5f * ((gx[c] + gx[n]) * dx + (gy[c] + gy[n]) * dy + (gz[c] + gz[n]) * dz);
    jump[i] = trust == 1.0f ? sp - sn : link + trust * ((sp - sn) - link);
  }
};
// :67-78
struct FvMomentumSource {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF sx, CF sy, CF sz,
                             CF face_pressure, CF volume, CF body_u, CF body_v, CF body_w, WF su,
                             WF sv, WF sw) const {
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j];
      const float pf = sign[j] * face_pressure[f];
      fx += pf * sx[f]; fy += pf * sy[f]; fz += pf * sz[f];
    }
    const float vol = volume[i];
    su[i] = body_u[i] - fx / vol; sv[i] = body_v[i] - fy / vol; sw[i] = body_w[i] - fz / vol;
  }
};
// :83-109
struct FvMomentumFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF px, CF py, CF pz, CF nx,
                             CF ny, CF nz, CF u, CF v, CF w, CF ux, CF uy, CF uz, CF vx, CF vy,
                             CF vz, CF wx, CF wy, CF wz, CF diagonal, CF bu, CF bv, CF bw, WF fu,
                             WF fv, WF fw, int internal) const {
    const int c = owner[i];
    if (i >= internal) {
      const int b = i - internal, k = kind[b];
      if (k == 2 || k == 4) {
        fu[i] = u[c]; fv[i] = v[c]; fw[i] = w[c];
      } else {
        fu[i] = bu[b]; fv[i] = bv[b]; fw[i] = bw[b];
      }
      return;
    }
    const int n = neighbour[i];
    const float a = diagonal[c] / (diagonal[c] + diagonal[n]), b = 1.0f - a;
    const float rx = a * px[i] + b * nx[i], ry = a * py[i] + b * ny[i], rz = a * pz[i] + b * nz[i];
    fu[i] = a * u[c] + b * u[n] + 0.5f * ((ux[c] + ux[n]) * rx + (uy[c] + uy[n]) * ry + (uz[c] + uz[n]) * rz);
    fv[i] = a * v[c] + b * v[n] + 0.5f * ((vx[c] + vx[n]) * rx + (vy[c] + vy[n]) * ry + (vz[c] + vz[n]) * rz);
    fw[i] = a * w[c] + b * w[n] + 0.5f * ((wx[c] + wx[n]) * rx + (wy[c] + wy[n]) * ry + (wz[c] + wz[n]) * rz);
  }
};
// :111-116
struct FvConvectiveOutflow {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF flux, WF outflow) const {
    float total = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) total += fmaxf(0.0f, sign[j] * flux[faces[j]]);
    outflow[i] = total;
  }
};
// :122-133
struct FvPecletSideVelocity {
  __device__ void operator()(int i, CI owner, CI neighbour, CF px, CF py, CF pz, CF nx, CF ny,
                             CF nz, CF q, CF gx, CF gy, CF gz, CF weights, CF diagonal,
                             CF outflow, WF face) const {
    const int c = owner[i], n = neighbour[i];
    const float a = weights[c] / (weights[c] + weights[n]);
    const float beta = fminf(1.0f, (outflow[c] + outflow[n]) / fmaxf(diagonal[c] + diagonal[n], 1.0e-20f));
    const float side = a * (q[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i]) +
                       (1.0f - a) * (q[n] + gx[n] * nx[i] + gy[n] * ny[i] + gz[n] * nz[i]);
    const float now = face[i];
    face[i] = now + beta * (side - now);
  }
};
// :139-162
struct FvRhieChow {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF sx, CF sy, CF sz,
                             CF diffusion, CF volume, CF diagonal, CF inverse, CF u, CF v, CF w,
                             CF pressure_jump, WF flux, WF coefficient, int internal,
                             float rho) const {
    const int c = owner[i];
    float rc = volume[c] / fmaxf(diagonal[c], 1.0e-20f), dc = volume[c] * inverse[c];
    if (i < internal) {
      const int n = neighbour[i];
      rc = 0.5f * (rc + volume[n] / fmaxf(diagonal[n], 1.0e-20f));
      dc = 0.5f * (dc + volume[n] * inverse[n]);
    } else {
      const int k = kind[i - internal];
      if (k == 0 || k == 3) {
        flux[i] = 0.0f; coefficient[i] = 0.0f;
        return;
      }
      if (k == 1) {
        flux[i] = rho * (u[i] * sx[i] + v[i] * sy[i] + w[i] * sz[i]);
        coefficient[i] = 0.0f;
        return;
      }
    }
    const float d = diffusion[i];
    flux[i] = rho * (u[i] * sx[i] + v[i] * sy[i] + w[i] * sz[i]) + rho * rc * d * pressure_jump[i];
    coefficient[i] = rho * dc * d;
  }
};
// :185-190
struct FvCorrectFlux {
  __device__ void operator()(int i, CI owner, CI neighbour, CF correction, CF coefficient,
                             WF flux, int internal) const {
    const float pc = correction[owner[i]];
    const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
    flux[i] += coefficient[i] * (pc - pn);
  }
};
// :192-202
struct FvCorrectVelocity {
  __device__ void operator()(int i, CF volume, CF inverse, CF diagonal, CF gx, CF gy, CF gz,
                             CF correction, WF u, WF v, WF w, WF p, float alpha) const {
    const float inv = inverse[i], d = volume[i] * inv;
    u[i] -= d * gx[i]; v[i] -= d * gy[i]; w[i] -= d * gz[i];
    p[i] += fminf(alpha, 0.7f * diagonal[i] * inv) * correction[i];
  }
};
// :204-215
struct FvMonitorCell {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF flux, CF u, CF v, CF w,
                             CF diagonal, CF residual_u, CF residual_v, CF residual_w,
                             WF continuity, WF scale, WF speed, WF ru, WF rv, WF rw) const {
    float mass = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) mass += sign[j] * flux[faces[j]];
    const float s = sqrtf(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
    continuity[i] = fabsf(mass);
    scale[i] = fabsf(diagonal[i]) * s;
    speed[i] = s;
    ru[i] = fabsf(residual_u[i]); rv[i] = fabsf(residual_v[i]); rw[i] = fabsf(residual_w[i]);
  }
};
// :217-219
struct FvSpeed {
  __device__ void operator()(int i, CF u, CF v, CF w, WF speed) const {
    speed[i] = sqrtf(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]);
  }
};
// :221-226
struct FvContinuity {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF flux, WF result) const {
    float mass = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) mass += sign[j] * flux[faces[j]];
    result[i] = fabsf(mass);
  }
};
// :228-233
struct FvWallPressureForce {
  __device__ void operator()(int i, CF selected, CF sx, CF sy, CF sz, CF face_pressure, WF x,
                             WF y, WF z, int internal, float reference) const {
    const int f = i + internal;
    const float pf = selected[i] * (face_pressure[f] - reference);
    x[i] = pf * sx[f]; y[i] = pf * sy[f]; z[i] = pf * sz[f];
  }
};
// :235-257
struct FvWallViscousForce {
  __device__ void operator()(int i, CI owner, CF selected, CF sx, CF sy, CF sz, CF rx, CF ry,
                             CF rz, CF u, CF v, CF w, CF ux, CF uy, CF uz, CF vx, CF vy, CF vz,
                             CF wx, CF wy, CF wz, CF bu, CF bv, CF bw, WF x, WF y, WF z,
                             int internal, float viscosity) const {
    const int f = i + internal, c = owner[f];
    const float ax = sx[f], ay = sy[f], az = sz[f], dx = rx[f], dy = ry[f], dz = rz[f];
    const float area = sqrtf(ax * ax + ay * ay + az * az);
    const float ex = ax / area, ey = ay / area, ez = az / area;
    const float dn = dx * ex + dy * ey + dz * ez;
    const float du = (bu[i] - u[c] - ux[c] * dx - uy[c] * dy - uz[c] * dz) / dn;
    const float dv = (bv[i] - v[c] - vx[c] * dx - vy[c] * dy - vz[c] * dz) / dn;
    const float dw = (bw[i] - w[c] - wx[c] * dx - wy[c] * dy - wz[c] * dz) / dn;
    const float g00 = ux[c] + du * ex, g01 = uy[c] + du * ey, g02 = uz[c] + du * ez;
    const float g10 = vx[c] + dv * ex, g11 = vy[c] + dv * ey, g12 = vz[c] + dv * ez;
    const float g20 = wx[c] + dw * ex, g21 = wy[c] + dw * ey, g22 = wz[c] + dw * ez;
    const float s = -viscosity * selected[i];
    x[i] = s * (2.0f * g00 * ax + (g01 + g10) * ay + (g02 + g20) * az);
    y[i] = s * ((g10 + g01) * ax + 2.0f * g11 * ay + (g12 + g21) * az);
    z[i] = s * ((g20 + g02) * ax + (g21 + g12) * ay + 2.0f * g22 * az);
  }
};
// :261-266
struct FvFluentMomentumScale {
  __device__ void operator()(int i, CF ax, CF ay, CF az, CF volume, CF speed, WF x, WF y, WF z,
                             float rho, float dt) const {
    const float t = dt > 0.0f ? rho * volume[i] / dt : 0.0f, s = speed[i];
    x[i] = 2.0f * (ax[i] + t) * s; y[i] = 2.0f * (ay[i] + t) * s; z[i] = 2.0f * (az[i] + t) * s;
  }
};
// :268-276
struct FvPhysicalTimeTerms {
  __device__ void operator()(int i, CF volume, CF old1, CF old2, WF diagonal, WF rhs,
                             float rho_dt, float a0, float a1, float a2, int add) const {
    const float s = rho_dt * volume[i];
    const float d0 = add ? diagonal[i] : 0.0f, b0 = add ? rhs[i] : 0.0f;
    diagonal[i] = d0 + s * a0;
    rhs[i] = b0 + s * (a1 * old1[i] + a2 * old2[i]);
  }
};
// :278-281
struct FvFaceTimeDefect {
  __device__ void operator()(int i, CF mass, CF sx, CF sy, CF sz, CF ux, CF uy, CF uz, WF defect,
                             float rho) const {
    defect[i] = mass[i] - rho * (sx[i] * ux[i] + sy[i] * uy[i] + sz[i] * uz[i]);
  }
};
// :283-295
struct FvFaceTimeFlux {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF volume, CF diagonal,
                             CF defect1, CF defect2, WF mass, int internal, float rho_dt,
                             float a1, float a2) const {
    const int c = owner[i];
    float d = volume[c] / fmaxf(diagonal[c], 1.0e-20f);
    if (i < internal) {
      const int n = neighbour[i];
      d = 0.5f * (d + volume[n] / fmaxf(diagonal[n], 1.0e-20f));
    } else {
      const int k = kind[i - internal];
      if (k == 0 || k == 1 || k == 3) return;
    }
    mass[i] += rho_dt * d * (a1 * defect1[i] + a2 * defect2[i]);
  }
};
// :297-299
struct FvCellMassError {
  __device__ void operator()(int i, CF imbalance, CF volume, WF error, float dt_over_rho) const {
    error[i] = dt_over_rho * fabsf(imbalance[i]) / volume[i];
  }
};

// ---------------------------------------------------- fv_gradient.kern.h
// Float pair arithmetic of kernels/prelude.h:102-133: TwoSum, the product-add
// that carries the product's rounding error through FMA, the gradient trace
// of a pair and the difference of two pairs.
struct F2 {
  float hi, lo;
};
__device__ __forceinline__ F2 two_sum(float a, float b) {
  const float s = a + b, t = s - a;
  return {s, (a - (s - t)) + (b - t)};
}
__device__ __forceinline__ F2 product_add(F2 acc, float x, float y) {
  const float p = x * y;
  const F2 s = two_sum(acc.hi, p);
  return two_sum(s.hi, (acc.lo + s.lo) + __builtin_fmaf(x, y, -p));
}
__device__ __forceinline__ F2 add_weighted(F2 acc, float w, F2 d) {
  return product_add(product_add(acc, w, d.hi), w, d.lo);
}
__device__ __forceinline__ F2 pair_trace(float p, float low, float gx, float gy, float gz, float lx,
                                         float ly, float lz, float x, float y, float z) {
  F2 v = two_sum(p, low);
  v = product_add(v, gx, x);
  v = product_add(v, lx, x);
  v = product_add(v, gy, y);
  v = product_add(v, ly, y);
  v = product_add(v, gz, z);
  return product_add(v, lz, z);
}
__device__ __forceinline__ float pair_delta(F2 a, F2 b) {
  const F2 d = two_sum(a.hi, -b.hi);
  return d.hi + (d.lo + (a.lo - b.lo));
}
// Value across cell-face entry j: cell `other`, or boundary face -(other+1).
__device__ __forceinline__ float across(int other, CF cell, CF boundary) {
  return other < 0 ? boundary[-other - 1] : cell[other];
}
// The symmetric 3x3 least-squares inverse of cell i: planes xx xy xz yy yz zz.
struct Sym3 {
  float xx, xy, xz, yy, yz, zz;
  __device__ Sym3(CF inverse, int i, int n)
      : xx(inverse[i]), xy(inverse[i + n]), xz(inverse[i + 2 * n]), yy(inverse[i + 3 * n]),
        yz(inverse[i + 4 * n]), zz(inverse[i + 5 * n]) {}
};

// :6-23
struct FvGradient {
  __device__ void operator()(int i, CI offsets, CI adjacent, CF dx, CF dy, CF dz, CF fx, CF fy,
                             CF fz, CF inverse, CF cell, CF boundary, WF gx, WF gy, WF gz,
                             int n) const {
    const float own = cell[i];
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int o = adjacent[j];
      const float d = across(o, cell, boundary) - own;
      const float wx = o < 0 ? fx[-o - 1] : dx[j];
      const float wy = o < 0 ? fy[-o - 1] : dy[j];
      const float wz = o < 0 ? fz[-o - 1] : dz[j];
      bx += wx * d; by += wy * d; bz += wz * d;
    }
    const Sym3 m(inverse, i, n);
    gx[i] = m.xx * bx + m.xy * by + m.xz * bz;
    gy[i] = m.xy * bx + m.yy * by + m.yz * bz;
    gz[i] = m.xz * bx + m.yz * by + m.zz * bz;
  }
};
// :28-56
struct FvPairedGradient {
  __device__ void operator()(int i, CI offsets, CI adjacent, CF dx, CF dy, CF dz, CF fx, CF fy,
                             CF fz, CF inverse, CF cell, CF low, CF boundary, CF boundary_low,
                             WF gx, WF gy, WF gz, WF lx, WF ly, WF lz, int n) const {
    const float own = cell[i], own_low = low[i];
    F2 bx{0.0f, 0.0f}, by{0.0f, 0.0f}, bz{0.0f, 0.0f};
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int o = adjacent[j];
      const float hi = across(o, cell, boundary), lo = o < 0 ? boundary_low[-o - 1] : low[o];
      F2 d = two_sum(hi, -own);
      d = product_add(d, lo, 1.0f);
      d = product_add(d, -own_low, 1.0f);
      const float wx = o < 0 ? fx[-o - 1] : dx[j];
      const float wy = o < 0 ? fy[-o - 1] : dy[j];
      const float wz = o < 0 ? fz[-o - 1] : dz[j];
      bx = add_weighted(bx, wx, d); by = add_weighted(by, wy, d); bz = add_weighted(bz, wz, d);
    }
    const Sym3 m(inverse, i, n);
    const F2 zero{0.0f, 0.0f};
    const F2 x = add_weighted(add_weighted(add_weighted(zero, m.xx, bx), m.xy, by), m.xz, bz);
    const F2 y = add_weighted(add_weighted(add_weighted(zero, m.xy, bx), m.yy, by), m.yz, bz);
    const F2 z = add_weighted(add_weighted(add_weighted(zero, m.xz, bx), m.yz, by), m.zz, bz);
    gx[i] = x.hi; gy[i] = y.hi; gz[i] = z.hi;
    lx[i] = x.lo; ly[i] = y.lo; lz[i] = z.lo;
  }
};
// :64-68
__device__ __forceinline__ float limit_face(float d, float delta, float eps2) {
  if (eps2 < 0.0f) return delta / d;
  const float a = delta * delta + eps2, b = d * delta;
  return (a + 2.0f * b) / (a + 2.0f * d * d + b);
}
// :72-74
struct FvRangeShift {
  __device__ void operator()(int i, CF field, WF above, WF below, float bound) const {
    const float v = field[i];
    above[i] = v + bound;
    below[i] = bound - v;
  }
};
// Face increments and the smallest admissible factor over a cell's faces.
__device__ __forceinline__ float limit_cell(int i, float gx, float gy, float gz, float up,
                                            float down, CI offsets, CI faces, CF sign, CF rx,
                                            CF ry, CF rz, CF nx, CF ny, CF nz, float eps2) {
  float psi = 1.0f;
  for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
    const int f = faces[j];
    const bool own = sign[j] > 0.0f;
    const float d = gx * (own ? rx[f] : nx[f]) + gy * (own ? ry[f] : ny[f]) + gz * (own ? rz[f] : nz[f]);
    if (d > 0.0f) psi = fminf(psi, limit_face(d, up, eps2));
    if (d < 0.0f) psi = fminf(psi, limit_face(d, down, eps2));
  }
  return fmaxf(0.0f, fminf(1.0f, psi));
}
// :76-94
struct FvLimiter {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF boundary, CF gx, CF gy,
                             CF gz, WF limiter, float eps2) const {
    const float own = cell[i];
    float lo = own, hi = own;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const float v = across(adjacent[j], cell, boundary);
      lo = fminf(lo, v); hi = fmaxf(hi, v);
    }
    limiter[i] = limit_cell(i, gx[i], gy[i], gz[i], hi - own, lo - own, offsets, faces, sign, rx,
                            ry, rz, nx, ny, nz, eps2);
  }
};
// :103-123
struct FvPairedLimiter {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF low, CF boundary, CI kind,
                             CF gx, CF gy, CF gz, WF limiter, float eps2) const {
    const float own = cell[i], own_low = low[i];
    float lo = 0.0f, hi = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int o = adjacent[j];
      const float v = across(o, cell, boundary);
      const float l = o < 0 ? (kind[-o - 1] == 2 ? own_low : 0.0f) : low[o];
      const float d = (v - own) + (l - own_low);
      lo = fminf(lo, d); hi = fmaxf(hi, d);
    }
    limiter[i] = limit_cell(i, gx[i], gy[i], gz[i], hi, lo, offsets, faces, sign, rx, ry, rz, nx,
                            ny, nz, eps2);
  }
};
// :127-145
struct FvPairedGradientLimiter {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF low, CF boundary,
                             CF boundary_low, CF gx, CF gy, CF gz, CF lx, CF ly, CF lz,
                             WF limiter, float eps2) const {
    const F2 own{cell[i], low[i]};
    float lo = 0.0f, hi = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int o = adjacent[j];
      const F2 v{across(o, cell, boundary), o < 0 ? boundary_low[-o - 1] : low[o]};
      const float d = pair_delta(v, own);
      lo = fminf(lo, d); hi = fmaxf(hi, d);
    }
    limiter[i] = limit_cell(i, gx[i] + lx[i], gy[i] + ly[i], gz[i] + lz[i], hi, lo, offsets,
                            faces, sign, rx, ry, rz, nx, ny, nz, eps2);
  }
};
// :147-166
struct FvVectorGradient {
  __device__ void operator()(int i, CI offsets, CI adjacent, CF dx, CF dy, CF dz, CF fx, CF fy,
                             CF fz, CF inverse, CF u, CF v, CF w, CF bu, CF bv, CF bw, WF ux,
                             WF uy, WF uz, WF vx, WF vy, WF vz, WF wx, WF wy, WF wz,
                             int n) const {
    const float u0 = u[i], v0 = v[i], w0 = w[i];
    float m[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int o = adjacent[j];
      const float du = across(o, u, bu) - u0, dv = across(o, v, bv) - v0, dw = across(o, w, bw) - w0;
      const float rx = o < 0 ? fx[-o - 1] : dx[j];
      const float ry = o < 0 ? fy[-o - 1] : dy[j];
      const float rz = o < 0 ? fz[-o - 1] : dz[j];
      m[0] += rx * du; m[1] += ry * du; m[2] += rz * du;
      m[3] += rx * dv; m[4] += ry * dv; m[5] += rz * dv;
      m[6] += rx * dw; m[7] += ry * dw; m[8] += rz * dw;
    }
    const Sym3 s(inverse, i, n);
    ux[i] = s.xx * m[0] + s.xy * m[1] + s.xz * m[2];
    uy[i] = s.xy * m[0] + s.yy * m[1] + s.yz * m[2];
    uz[i] = s.xz * m[0] + s.yz * m[1] + s.zz * m[2];
    vx[i] = s.xx * m[3] + s.xy * m[4] + s.xz * m[5];
    vy[i] = s.xy * m[3] + s.yy * m[4] + s.yz * m[5];
    vz[i] = s.xz * m[3] + s.yz * m[4] + s.zz * m[5];
    wx[i] = s.xx * m[6] + s.xy * m[7] + s.xz * m[8];
    wy[i] = s.xy * m[6] + s.yy * m[7] + s.yz * m[8];
    wz[i] = s.xz * m[6] + s.yz * m[7] + s.zz * m[8];
  }
};
// :168-171
struct FvScaleGradient {
  __device__ void operator()(int i, CF gx, CF gy, CF gz, CF limiter, WF x, WF y, WF z) const {
    const float l = limiter[i];
    x[i] = l * gx[i]; y[i] = l * gy[i]; z[i] = l * gz[i];
  }
};
// :173-180
struct FvLimiterScale {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF boundary, CF gx, CF gy,
                             CF gz, WF limiter, WF x, WF y, WF z, float eps2) const {
    FvLimiter{}(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, boundary, gx, gy,
                gz, limiter, eps2);
    FvScaleGradient{}(i, gx, gy, gz, limiter, x, y, z);
  }
};
// :182-189
struct FvPairedLimiterScale {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF low, CF boundary, CI kind,
                             CF gx, CF gy, CF gz, WF limiter, WF x, WF y, WF z,
                             float eps2) const {
    FvPairedLimiter{}(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, low,
                      boundary, kind, gx, gy, gz, limiter, eps2);
    FvScaleGradient{}(i, gx, gy, gz, limiter, x, y, z);
  }
};
// :191-199
struct FvPairedGradientLimiterScale {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF low, CF boundary,
                             CF boundary_low, CF gx, CF gy, CF gz, CF lx, CF ly, CF lz,
                             WF limiter, WF x, WF y, WF z, float eps2) const {
    FvPairedGradientLimiter{}(i, offsets, adjacent, faces, sign, rx, ry, rz, nx, ny, nz, cell, low,
                              boundary, boundary_low, gx, gy, gz, lx, ly, lz, limiter, eps2);
    FvScaleGradient{}(i, gx, gy, gz, limiter, x, y, z);
  }
};
// g + (1/2) M b for the jump moments b (:218-220, :251-253).
__device__ __forceinline__ void half_jump_update(int i, int n, CF inverse, CF gx, CF gy, CF gz,
                                                 float bx, float by, float bz, WF x, WF y, WF z) {
  const Sym3 m(inverse, i, n);
  x[i] = gx[i] + 0.5f * (m.xx * bx + m.xy * by + m.xz * bz);
  y[i] = gy[i] + 0.5f * (m.xy * bx + m.yy * by + m.yz * bz);
  z[i] = gz[i] + 0.5f * (m.xz * bx + m.yz * by + m.zz * bz);
}
// :202-221
struct FvJumpGradient {
  __device__ void operator()(int i, CI offsets, CI faces, CI adjacent, CF sign, CF dx, CF dy,
                             CF dz, CF inverse, CF px, CF py, CF pz, CF nx, CF ny, CF nz,
                             CF cell, CF boundary, CF gx, CF gy, CF gz, WF x, WF y, WF z,
                             int n) const {
    const float c0 = cell[i], g0x = gx[i], g0y = gy[i], g0z = gz[i];
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j], o = adjacent[j];
      const bool side = sign[j] > 0.0f;
      const float own = c0 + g0x * (side ? px[f] : nx[f]) + g0y * (side ? py[f] : ny[f]) +
                        g0z * (side ? pz[f] : nz[f]);
      const float other = o < 0 ? boundary[-o - 1]
                                : cell[o] + gx[o] * (side ? nx[f] : px[f]) +
                                      gy[o] * (side ? ny[f] : py[f]) + gz[o] * (side ? nz[f] : pz[f]);
      const float jump = other - own;
      bx += dx[j] * jump; by += dy[j] * jump; bz += dz[j] * jump;
    }
    half_jump_update(i, n, inverse, gx, gy, gz, bx, by, bz, x, y, z);
  }
};
// :227-240
struct FvFaceJump {
  __device__ void operator()(int i, CI owner, CI neighbour, CF px, CF py, CF pz, CF nx, CF ny,
                             CF nz, CF cell, CF boundary, CF gx, CF gy, CF gz, WF forward,
                             WF reverse, int internal) const {
    const int c = owner[i];
    const float own = cell[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
    float other;
    if (i < internal) {
      const int nb = neighbour[i];
      other = cell[nb] + gx[nb] * nx[i] + gy[nb] * ny[i] + gz[nb] * nz[i];
      reverse[i] = own - other;
    } else {
      other = boundary[i - internal];
    }
    forward[i] = other - own;
  }
};
// :242-254
struct FvGatherJumpGradient {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF dx, CF dy, CF dz,
                             CF inverse, CF forward, CF reverse, CF gx, CF gy, CF gz, WF x, WF y,
                             WF z, int n) const {
    float bx = 0.0f, by = 0.0f, bz = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j];
      const float jump = sign[j] > 0.0f ? forward[f] : reverse[f];
      bx += dx[j] * jump; by += dy[j] * jump; bz += dz[j] * jump;
    }
    half_jump_update(i, n, inverse, gx, gy, gz, bx, by, bz, x, y, z);
  }
};
// :258-292
struct FvPairedJumpGradient {
  __device__ void operator()(int i, CI offsets, CI faces, CI adjacent, CF sign, CF dx, CF dy,
                             CF dz, CF inverse, CF px, CF py, CF pz, CF nx, CF ny, CF nz,
                             CF cell, CF low, CF boundary, CF boundary_low, CF gx, CF gy, CF gz,
                             CF lx, CF ly, CF lz, WF x, WF y, WF z, int n) const {
    F2 bx{0.0f, 0.0f}, by{0.0f, 0.0f}, bz{0.0f, 0.0f};
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j], o = adjacent[j];
      const bool side = sign[j] > 0.0f;
      const F2 own = pair_trace(cell[i], low[i], gx[i], gy[i], gz[i], lx[i], ly[i], lz[i],
                                side ? px[f] : nx[f], side ? py[f] : ny[f], side ? pz[f] : nz[f]);
      F2 other{0.0f, 0.0f};
      if (o < 0) other = {boundary[-o - 1], boundary_low[-o - 1]};
      else
        other = pair_trace(cell[o], low[o], gx[o], gy[o], gz[o], lx[o], ly[o], lz[o],
                           side ? nx[f] : px[f], side ? ny[f] : py[f], side ? nz[f] : pz[f]);
      F2 jump = two_sum(other.hi, -own.hi);
      jump = product_add(jump, other.lo, 1.0f);
      jump = product_add(jump, -own.lo, 1.0f);
      bx = add_weighted(bx, dx[j], jump); by = add_weighted(by, dy[j], jump);
      bz = add_weighted(bz, dz[j], jump);
    }
    const Sym3 m(inverse, i, n);
    const float xx = .5f * m.xx, xy = .5f * m.xy, xz = .5f * m.xz;
    const float yy = .5f * m.yy, yz = .5f * m.yz, zz = .5f * m.zz;
    F2 a = two_sum(gx[i], lx[i]), b = two_sum(gy[i], ly[i]), c = two_sum(gz[i], lz[i]);
    a = add_weighted(add_weighted(add_weighted(a, xx, bx), xy, by), xz, bz);
    b = add_weighted(add_weighted(add_weighted(b, xy, bx), yy, by), yz, bz);
    c = add_weighted(add_weighted(add_weighted(c, xz, bx), yz, by), zz, bz);
    x[i] = a.hi; y[i] = b.hi; z[i] = c.hi;
  }
};
// :296-299. One lane per stored cell; the cells are distinct.
struct FvLsqRestoreInverse {
  __device__ void operator()(int i, CI cells, CF values, WF inverse, int count, int n) const {
    const int c = cells[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) inverse[c + k * n] = values[i + k * count];
  }
};

template <class Body, class... P>
__global__ void __launch_bounds__(kBlock) k_each(int n, P... p) {
  const int i = cell_index();
  if (i < n) Body{}(i, p...);
}
template <class T>
T from_arg(const kh_arg* a, std::size_t k) {
  if constexpr (std::is_pointer_v<T>) return static_cast<T>(a[k].value.ptr);
  else if constexpr (std::is_same_v<T, float>) return a[k].value.f32;
  else if constexpr (std::is_same_v<T, unsigned>) return a[k].value.u32;
  else return a[k].value.i32;
}
template <class T>
constexpr char kind_code() {
  if constexpr (std::is_pointer_v<T>) return 'b';
  else if constexpr (std::is_same_v<T, float>) return 'f';
  else if constexpr (std::is_same_v<T, unsigned>) return 'u';
  else return 'i';
}
template <class Body, class... P, std::size_t... K>
void launch_impl(kh_device* d, int n, const kh_arg* a, void (Body::*)(int, P...) const,
                 std::index_sequence<K...>) {
  k_each<Body, P...><<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, from_arg<P>(a, K)...);
}
template <class Body, class... P>
constexpr std::size_t arity(void (Body::*)(int, P...) const) { return sizeof...(P); }
template <class Body>
void launch_body(kh_device* d, int n, const kh_arg* a) {
  using Op = decltype(&Body::operator());
  launch_impl(d, n, a, Op{}, std::make_index_sequence<arity(Op{})>{});
}
template <class Body, class... P>
const char* signature_impl(void (Body::*)(int, P...) const) {
  static const char s[] = {kind_code<P>()..., '\0'};
  return s;
}
template <class Body>
const char* signature_of() {
  return signature_impl(decltype(&Body::operator()){});
}

#define KB(k) static_cast<const int*>(buf(a, k))
#define KF(k) static_cast<float*>(buf(a, k))
#define KCF(k) static_cast<const float*>(buf(a, k))
#define GRID(n) dim3(blocks_for(n)), dim3(kBlock), 0, d->stream

void l_fv_pressure_assemble(kh_device* d, int n, const kh_arg* a) {
  k_fv_pressure_assemble<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KB(3), KB(4), KB(5), KB(6), KCF(7),
                                      KCF(8), KF(9), KF(10), i32(a, 11));
}

}  // namespace

const std::vector<KernelEntry>& fv_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_FV_PRESSURE_ASSEMBLE, "fv_pressure_assemble", "bbbbbbbbbbbi", l_fv_pressure_assemble},
#define FV(id, name, Body) {id, name, signature_of<Body>(), launch_body<Body>}
      FV(KH_FV_VELOCITY_BOUNDARY, "fv_velocity_boundary", FvVelocityBoundary),
      FV(KH_FV_PRESSURE_BOUNDARY, "fv_pressure_boundary", FvPressureBoundary),
      FV(KH_FV_PRESSURE_FACES, "fv_pressure_faces", FvPressureFaces),
      FV(KH_FV_MOMENTUM_SOURCE, "fv_momentum_source", FvMomentumSource),
      FV(KH_FV_MOMENTUM_FACES, "fv_momentum_faces", FvMomentumFaces),
      FV(KH_FV_CONVECTIVE_OUTFLOW, "fv_convective_outflow", FvConvectiveOutflow),
      FV(KH_FV_PECLET_SIDE_VELOCITY, "fv_peclet_side_velocity", FvPecletSideVelocity),
      FV(KH_FV_RHIE_CHOW, "fv_rhie_chow", FvRhieChow),
      FV(KH_FV_CORRECT_FLUX, "fv_correct_flux", FvCorrectFlux),
      FV(KH_FV_CORRECT_VELOCITY, "fv_correct_velocity", FvCorrectVelocity),
      FV(KH_FV_MONITOR_CELL, "fv_monitor_cell", FvMonitorCell),
      FV(KH_FV_SPEED, "fv_speed", FvSpeed),
      FV(KH_FV_CONTINUITY, "fv_continuity", FvContinuity),
      FV(KH_FV_WALL_PRESSURE_FORCE, "fv_wall_pressure_force", FvWallPressureForce),
      FV(KH_FV_WALL_VISCOUS_FORCE, "fv_wall_viscous_force", FvWallViscousForce),
      FV(KH_FV_FLUENT_MOMENTUM_SCALE, "fv_fluent_momentum_scale", FvFluentMomentumScale),
      FV(KH_FV_PHYSICAL_TIME_TERMS, "fv_physical_time_terms", FvPhysicalTimeTerms),
      FV(KH_FV_FACE_TIME_DEFECT, "fv_face_time_defect", FvFaceTimeDefect),
      FV(KH_FV_FACE_TIME_FLUX, "fv_face_time_flux", FvFaceTimeFlux),
      FV(KH_FV_CELL_MASS_ERROR, "fv_cell_mass_error", FvCellMassError),
      FV(KH_FV_GRADIENT, "fv_gradient", FvGradient),
      FV(KH_FV_PAIRED_GRADIENT, "fv_paired_gradient", FvPairedGradient),
      FV(KH_FV_RANGE_SHIFT, "fv_range_shift", FvRangeShift),
      FV(KH_FV_LIMITER, "fv_limiter", FvLimiter),
      FV(KH_FV_PAIRED_LIMITER, "fv_paired_limiter", FvPairedLimiter),
      FV(KH_FV_PAIRED_GRADIENT_LIMITER, "fv_paired_gradient_limiter", FvPairedGradientLimiter),
      FV(KH_FV_VECTOR_GRADIENT, "fv_vector_gradient", FvVectorGradient),
      FV(KH_FV_SCALE_GRADIENT, "fv_scale_gradient", FvScaleGradient),
      FV(KH_FV_LIMITER_SCALE, "fv_limiter_scale", FvLimiterScale),
      FV(KH_FV_PAIRED_LIMITER_SCALE, "fv_paired_limiter_scale", FvPairedLimiterScale),
      FV(KH_FV_PAIRED_GRADIENT_LIMITER_SCALE, "fv_paired_gradient_limiter_scale",
         FvPairedGradientLimiterScale),
      FV(KH_FV_JUMP_GRADIENT, "fv_jump_gradient", FvJumpGradient),
      FV(KH_FV_FACE_JUMP, "fv_face_jump", FvFaceJump),
      FV(KH_FV_GATHER_JUMP_GRADIENT, "fv_gather_jump_gradient", FvGatherJumpGradient),
      FV(KH_FV_PAIRED_JUMP_GRADIENT, "fv_paired_jump_gradient", FvPairedJumpGradient),
      FV(KH_FV_LSQ_RESTORE_INVERSE, "fv_lsq_restore_inverse", FvLsqRestoreInverse),
#undef FV
  };
  return table;
}

}  // namespace kh
