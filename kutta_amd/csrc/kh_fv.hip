// Finite-volume kernels of the SIMPLEC pressure path for gfx950. One lane
// owns one cell and gathers over its faces; each lane writes only its own CSR
// row, so no atomics are needed and results are deterministic.
#include <cstddef>
#include <type_traits>
#include <utility>

#include <hip/hip_ext.h>

#include "kh_device_math.hpp"

namespace kh {
namespace {

__device__ __forceinline__ int cell_index() { return int(xcd_block() * kBlock + threadIdx.x); }

// ------------------------------------------------- one lane per face or cell
// Each body below is a functor whose operator()(i, ...) takes the reference
// kernel's parameters in order; k_each runs it for i in [0, n) and the table
// derives the argument unpacking and signature from the operator's type.
using CI = const int* __restrict__;
using CF = const float* __restrict__;
using WF = float* __restrict__;

// The same value read through one selected address (one load, no branch).
__device__ __forceinline__ float across_load(int other, CF cell, CF boundary) {
  return *(other < 0 ? boundary + (-other - 1) : cell + other);
}
// A face-indexed value or, for a boundary neighbour o < 0, boundary slot -o-1.
__device__ __forceinline__ float face_or_boundary(int other, int j, CF per_face, CF per_boundary) {
  return *(other < 0 ? per_boundary + (-other - 1) : per_face + j);
}
// Runs use(j, v) for the faces (or row entries) j of [first, last) in order,
// where v = load(j) holds the values a face reads. The loads of up to K faces
// are issued before any is used, so a cell pays about one memory round trip
// per chunk instead of one per face; the arithmetic is the face-by-face
// loop's. A full chunk loads without conditions (one straight run of loads;
// six covers a hexahedron's faces), a partial one (a tetrahedron's four)
// loads only its faces.
constexpr int kFaceChunk = 8;  // the transport assembly's chunk
template <int K = 6, bool Exact = true, class Load, class Use>
__device__ __forceinline__ void each_face(int first, int last, Load load, Use use) {
  for (int j0 = first; j0 < last; j0 += K) {
    decltype(load(first)) v[K];
    if (!Exact) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = load(min(j0 + k, last - 1));
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (j0 + k < last) use(j0 + k, v[k]);
    } else if (last - j0 >= K) {
#pragma unroll
      for (int k = 0; k < K; ++k) v[k] = load(j0 + k);
#pragma unroll
      for (int k = 0; k < K; ++k) use(j0 + k, v[k]);
    } else {
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (j0 + k < last) v[k] = load(j0 + k);
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (j0 + k < last) use(j0 + k, v[k]);
    }
  }
}
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
    const float trust = kmax(0.0f, kmin(1.0f, (4.0f * d2 - m2) / (3.0f * d2)));
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
    struct Face {
      float s, p, x, y, z;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j];
          return Face{sign[j], face_pressure[f], sx[f], sy[f], sz[f]};
        },
        [&](int, const Face& q) {
          const float pf = q.s * q.p;
          fx += pf * q.x; fy += pf * q.y; fz += pf * q.z;
        });
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
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) total += kmax(0.0f, sign[j] * flux[faces[j]]);
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
    const float beta = kmin(1.0f, (outflow[c] + outflow[n]) / kmax(diagonal[c] + diagonal[n], 1.0e-20f));
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
    float rc = volume[c] / kmax(diagonal[c], 1.0e-20f), dc = volume[c] * inverse[c];
    if (i < internal) {
      const int n = neighbour[i];
      rc = 0.5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
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
    p[i] += kmin(alpha, 0.7f * diagonal[i] * inv) * correction[i];
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
    struct Face {
      float v, wx, wy, wz;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int o = adjacent[j];
          return Face{across_load(o, cell, boundary), face_or_boundary(o, j, dx, fx), face_or_boundary(o, j, dy, fy),
                      face_or_boundary(o, j, dz, fz)};
        },
        [&](int, const Face& q) {
          const float d = q.v - own;
          bx += q.wx * d; by += q.wy * d; bz += q.wz * d;
        });
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
    struct Face {
      float hi, lo, wx, wy, wz;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int o = adjacent[j];
          return Face{across_load(o, cell, boundary), across_load(o, low, boundary_low),
                      face_or_boundary(o, j, dx, fx), face_or_boundary(o, j, dy, fy), face_or_boundary(o, j, dz, fz)};
        },
        [&](int, const Face& q) {
          F2 d = two_sum(q.hi, -own);
          d = product_add(d, q.lo, 1.0f);
          d = product_add(d, -own_low, 1.0f);
          bx = add_weighted(bx, q.wx, d); by = add_weighted(by, q.wy, d); bz = add_weighted(bz, q.wz, d);
        });
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
  struct Face {
    float x, y, z;
  };
  each_face<6, false>(
      offsets[i], offsets[i + 1],
      [&](int j) {
        const int f = faces[j];
        const bool own = sign[j] > 0.0f;
        return Face{*((own ? rx : nx) + f), *((own ? ry : ny) + f), *((own ? rz : nz) + f)};
      },
      [&](int, const Face& q) {
        const float d = gx * q.x + gy * q.y + gz * q.z;
        if (d > 0.0f) psi = kmin(psi, limit_face(d, up, eps2));
        if (d < 0.0f) psi = kmin(psi, limit_face(d, down, eps2));
      });
  return kmax(0.0f, kmin(1.0f, psi));
}
// :76-94
struct FvLimiter {
  __device__ void operator()(int i, CI offsets, CI adjacent, CI faces, CF sign, CF rx, CF ry,
                             CF rz, CF nx, CF ny, CF nz, CF cell, CF boundary, CF gx, CF gy,
                             CF gz, WF limiter, float eps2) const {
    const float own = cell[i];
    float lo = own, hi = own;
    each_face<6, false>(
        offsets[i], offsets[i + 1], [&](int j) { return across_load(adjacent[j], cell, boundary); },
        [&](int, float v) {
          lo = kmin(lo, v); hi = kmax(hi, v);
        });
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
    struct Face {
      float v, l;
    };
    each_face<6, false>(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int o = adjacent[j];
          return Face{across_load(o, cell, boundary), o < 0 ? (kind[-o - 1] == 2 ? own_low : 0.0f) : low[o]};
        },
        [&](int, const Face& q) {
          const float d = (q.v - own) + (q.l - own_low);
          lo = kmin(lo, d); hi = kmax(hi, d);
        });
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
    each_face<6, false>(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int o = adjacent[j];
          return F2{across_load(o, cell, boundary), across_load(o, low, boundary_low)};
        },
        [&](int, const F2& v) {
          const float d = pair_delta(v, own);
          lo = kmin(lo, d); hi = kmax(hi, d);
        });
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
    struct Face {
      float u, v, w, rx, ry, rz;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int o = adjacent[j];
          return Face{across_load(o, u, bu), across_load(o, v, bv), across_load(o, w, bw),
                      face_or_boundary(o, j, dx, fx), face_or_boundary(o, j, dy, fy), face_or_boundary(o, j, dz, fz)};
        },
        [&](int, const Face& q) {
          const float du = q.u - u0, dv = q.v - v0, dw = q.w - w0;
          m[0] += q.rx * du; m[1] += q.ry * du; m[2] += q.rz * du;
          m[3] += q.rx * dv; m[4] += q.ry * dv; m[5] += q.rz * dv;
          m[6] += q.rx * dw; m[7] += q.ry * dw; m[8] += q.rz * dw;
        });
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
    // A boundary face reads the boundary value through the neighbour slot and
    // this cell's gradient as a stand-in (unused) for the neighbour's.
    struct Face {
      int o;
      float v, ogx, ogy, ogz, ax, ay, az, bx, by, bz, dx, dy, dz;
    };
    each_face<4>(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j], o = adjacent[j], oc = o < 0 ? i : o;
          const bool side = sign[j] > 0.0f;
          return Face{o, across_load(o, cell, boundary), gx[oc], gy[oc], gz[oc],
                      *((side ? px : nx) + f), *((side ? py : ny) + f), *((side ? pz : nz) + f),
                      *((side ? nx : px) + f), *((side ? ny : py) + f), *((side ? nz : pz) + f),
                      dx[j], dy[j], dz[j]};
        },
        [&](int, const Face& q) {
          const float own = c0 + g0x * q.ax + g0y * q.ay + g0z * q.az;
          const float other = q.o < 0 ? q.v : q.v + q.ogx * q.bx + q.ogy * q.by + q.ogz * q.bz;
          const float jump = other - own;
          bx += q.dx * jump; by += q.dy * jump; bz += q.dz * jump;
        });
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
    struct Face {
      float jump, dx, dy, dz;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) { return Face{*((sign[j] > 0.0f ? forward : reverse) + faces[j]), dx[j], dy[j], dz[j]}; },
        [&](int, const Face& q) {
          bx += q.dx * q.jump; by += q.dy * q.jump; bz += q.dz * q.jump;
        });
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
    const float c0 = cell[i], l0 = low[i], g0x = gx[i], g0y = gy[i], g0z = gz[i];
    const float h0x = lx[i], h0y = ly[i], h0z = lz[i];
    // As FvJumpGradient: a boundary face reads the boundary pair through the
    // neighbour slots and this cell's gradients as unused stand-ins.
    struct Face {
      int o;
      float v, vl, ogx, ogy, ogz, olx, oly, olz, ax, ay, az, bx, by, bz, dx, dy, dz;
    };
    each_face<4>(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j], o = adjacent[j], oc = o < 0 ? i : o;
          const bool side = sign[j] > 0.0f;
          return Face{o, across_load(o, cell, boundary), across_load(o, low, boundary_low), gx[oc], gy[oc], gz[oc],
                      lx[oc], ly[oc], lz[oc], *((side ? px : nx) + f), *((side ? py : ny) + f),
                      *((side ? pz : nz) + f), *((side ? nx : px) + f), *((side ? ny : py) + f),
                      *((side ? nz : pz) + f), dx[j], dy[j], dz[j]};
        },
        [&](int, const Face& q) {
          const F2 own = pair_trace(c0, l0, g0x, g0y, g0z, h0x, h0y, h0z, q.ax, q.ay, q.az);
          F2 other{0.0f, 0.0f};
          if (q.o < 0) other = {q.v, q.vl};
          else other = pair_trace(q.v, q.vl, q.ogx, q.ogy, q.ogz, q.olx, q.oly, q.olz, q.bx, q.by, q.bz);
          F2 jump = two_sum(other.hi, -own.hi);
          jump = product_add(jump, other.lo, 1.0f);
          jump = product_add(jump, -own.lo, 1.0f);
          bx = add_weighted(bx, q.dx, jump); by = add_weighted(by, q.dy, jump);
          bz = add_weighted(bz, q.dz, jump);
        });
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

// --------------------------------------------------- fv_transport.kern.h
// :3-10
struct FvDiffusionCorrectionBound {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF field, CF boundary,
                             CF diffusion, WF bound, int internal) const {
    const float own = field[owner[i]];
    float other = own;
    if (i < internal) other = field[neighbour[i]];
    else if (kind[i - internal] == 0) other = boundary[i - internal];
    bound[i] = diffusion[i] * fabsf(other - own);
  }
};
// :12-35. Deferred correction of face i: the non-orthogonal diffusion part
// from the face-weighted gradient, minus the limited second-order upwind part.
struct FvTransportFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CF weight, CF px, CF py, CF pz, CF nx,
                             CF ny, CF nz, CF tx, CF ty, CF tz, CF gx, CF gy, CF gz, CF limiter,
                             CF flux, CF mu, CI kind, WF correction, int internal,
                             int second_order, int limited_diffusion) const {
    const int c = owner[i];
    const float m = flux[i];
    float ax = gx[c], ay = gy[c], az = gz[c], convection = 0.0f;
    if (i < internal) {
      const int n = neighbour[i];
      const float w = weight[i], v = 1.0f - w;
      ax = w * ax + v * gx[n]; ay = w * ay + v * gy[n]; az = w * az + v * gz[n];
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
    if (limited_diffusion) {
      const float bound = correction[i];
      diffusion = kmax(-bound, kmin(bound, diffusion));
    }
    correction[i] = mu[i] * diffusion - (second_order != 0 ? convection : 0.0f);
  }
};
// The transport row's storage during assembly: the matrix itself, or, for
// rows of at most kSlots entries, registers (a slot is read from memory only
// if no face of the cell writes it, and every slot is stored once at the
// end). Either way each entry sees the same additions in the same order.
struct MemoryRow {
  float* values;
  int begin, end, di;
  __device__ void clear(int e) { values[e] = 0.0f; }
  __device__ void fill_untouched() {}
  __device__ void add(int e, float v) { values[e] += v; }
  template <class F>
  __device__ void each_off_diagonal(F f) const {
    for (int j = begin; j < end; ++j)
      if (j != di) f(values[j]);
  }
  __device__ void set_diagonal(float v) { values[di] = v; }
  __device__ void store() {}
};
struct RegisterRow {
  static constexpr int kSlots = 16;
  float* values;
  int begin, len, dk;
  unsigned touched = 0;
  float slot[kSlots] = {};
  __device__ void clear(int e) {
    const int k = e - begin;
    if (k >= 0 && k < kSlots) touched |= 1u << k;
  }
  __device__ void fill_untouched() {
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if (s < len && !((touched >> s) & 1u)) slot[s] = values[begin + s];
  }
  // Every slot takes an addition, -0.0 (the exact identity of IEEE addition)
  // for all but the entry's own: a compare-and-branch chain is folded back
  // into a dynamically indexed store, which would put the row in scratch.
  __device__ void add(int e, float v) {
    const int k = e - begin;
#pragma unroll
    for (int s = 0; s < kSlots; ++s) slot[s] += s == k ? v : -0.0f;
  }
  template <class F>
  __device__ void each_off_diagonal(F f) const {
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if (s < len && s != dk) f(slot[s]);
  }
  __device__ void set_diagonal(float v) {
#pragma unroll
    for (int s = 0; s < kSlots; ++s) slot[s] = s == dk ? v : slot[s];
  }
  __device__ void store() {
#pragma unroll
    for (int s = 0; s < kSlots; ++s)
      if (s < len) values[begin + s] = slot[s];
  }
};

// kernels/fv_flow.kern.h:164-183 — pressure-correction M-matrix row: clear the
// row's neighbour entries, subtract each face coefficient from its (possibly
// shared) entry, eliminate the reference cell's row and column, and build the
// diagonal as the upward-rounded sum of the stored off-diagonals plus boundary
// anchors; b = -sum sign * flux. As the transport assembly, the row is held in
// registers for rows of at most RegisterRow::kSlots entries (every entry sees
// the same operations in the same order: a cleared entry is 0, and x - d is
// x + (-d)) and the faces' values are loaded a chunk at a time.
struct PressureAssembly {
  template <class Row>
  __device__ __forceinline__ static void assemble(Row& row, int i, int begin, int end, CI faces, CF sign,
                                                  CI adjacent, CI entry, CF coefficient, CF flux,
                                                  WF rhs, int reference) {
    for (int j = begin; j < end; ++j)
      if (adjacent[j] >= 0) row.clear(entry[j]);
    row.fill_untouched();
    UpperSum ap{0.0f, 0.0f};
    float b = 0.0f;
    struct Face {
      int nb, e;
      float s, d, m;
    };
    each_face(
        begin, end,
        [&](int j) {
          const int f = faces[j];
          return Face{adjacent[j], entry[j], sign[j], coefficient[f], flux[f]};
        },
        [&](int, const Face& q) {
          b -= q.s * q.m;
          if (q.nb < 0 || q.nb == reference) ap = upper_add(ap, q.d);
          if (q.nb >= 0 && q.nb != reference && i != reference) row.add(q.e, -q.d);
        });
    row.each_off_diagonal([&](float v) { ap = upper_add(ap, -v); });
    row.set_diagonal(i == reference ? 1.0f : upper_value(ap));
    row.store();
    rhs[i] = i == reference ? 0.0f : b;
  }
};
__global__ void __launch_bounds__(kBlock) k_fv_pressure_assemble(int n,
    const int* __restrict__ offsets, const int* __restrict__ faces,
    const float* __restrict__ sign, const int* __restrict__ adjacent,
    const int* __restrict__ entry, const int* __restrict__ rows, const int* __restrict__ diag,
    const float* __restrict__ coefficient, const float* __restrict__ flux,
    float* __restrict__ values, float* __restrict__ rhs, int reference) {
  const int i = cell_index();
  if (i >= n) return;
  const int begin = offsets[i], end = offsets[i + 1];
  const int r0 = rows[i], r1 = rows[i + 1], di = diag[i];
  if (r1 - r0 <= RegisterRow::kSlots) {
    RegisterRow row{values, r0, r1 - r0, di - r0};
    PressureAssembly::assemble(row, i, begin, end, faces, sign, adjacent, entry, coefficient, flux, rhs, reference);
  } else {
    MemoryRow row{values, r0, r1, di};
    PressureAssembly::assemble(row, i, begin, end, faces, sign, adjacent, entry, coefficient, flux, rhs, reference);
  }
}

// :37-112. One transport-matrix row per lane, written into the row's own
// entries only: off-diagonals -D + min(F, 0); the diagonal as the upward-
// rounded sum of |off-diagonals|, net outflow, boundary anchors and sinks;
// the right-hand side optionally carried as a (hi, lo) pair; relaxation with
// the pseudo-time term; and the SIMPLEC consistent inverse.
struct FvTransportAssemble {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CI adjacent, CI entry, CI rows,
                             CI diag, CF diffusion, CF flux, CF mu, CI kind, CF boundary,
                             CF correction, CF source, CF volume, CF old, CF sink,
                             CF extra_diagonal, CF extra_rhs, WF values, WF rhs, WF rhs_low,
                             WF diagonal, WF diagonal_low, WF consistent_inverse, int internal,
                             float rho, float dt, float relax, int compensated) const {
    const int begin = rows[i], end = rows[i + 1], di = diag[i];
    if (end - begin <= RegisterRow::kSlots) {
      RegisterRow row{values, begin, end - begin, di - begin};
      assemble(row, i, offsets, faces, sign, adjacent, entry, diffusion, flux, mu, kind, boundary, correction,
               source, volume, old, sink, extra_diagonal, extra_rhs, rhs, rhs_low, diagonal, diagonal_low,
               consistent_inverse, internal, rho, dt, relax, compensated);
    } else {
      MemoryRow row{values, begin, end, di};
      assemble(row, i, offsets, faces, sign, adjacent, entry, diffusion, flux, mu, kind, boundary, correction,
               source, volume, old, sink, extra_diagonal, extra_rhs, rhs, rhs_low, diagonal, diagonal_low,
               consistent_inverse, internal, rho, dt, relax, compensated);
    }
  }
  template <class Row>
  __device__ __forceinline__ static void assemble(Row& row, int i, CI offsets, CI faces, CF sign, CI adjacent, CI entry,
                                  CF diffusion, CF flux, CF mu, CI kind, CF boundary, CF correction, CF source,
                                  CF volume, CF old, CF sink, CF extra_diagonal, CF extra_rhs, WF rhs,
                                  WF rhs_low, WF diagonal, WF diagonal_low, WF consistent_inverse, int internal,
                                  float rho, float dt, float relax, int compensated) {
    const float vol = volume[i], sink_volume = sink[i] * vol;
    UpperSum ap = upper_add({sink_volume, 0.0f}, extra_diagonal[i]);
    float b = source[i] * vol + extra_rhs[i], carry = 0.0f;
    if (compensated) {
      const float product = source[i] * vol;
      const F2 start = two_sum(product, extra_rhs[i]);
      b = start.hi;
      carry = start.lo + __builtin_fmaf(source[i], vol, -product);
    }
    const int first = offsets[i], last = offsets[i + 1];
    for (int j = first; j < last; ++j)
      if (adjacent[j] >= 0) row.clear(entry[j]);
    row.fill_untouched();
    // Faces a chunk at a time: the chunk's face ids, neighbours, signs and
    // entries are loaded first, then the gathered face values, then the
    // faces are summed in order, the arithmetic of the face-by-face loop.
    for (int j0 = first; j0 < last; j0 += kFaceChunk) {
      int fc[kFaceChunk], oc[kFaceChunk], ec[kFaceChunk];
      float sc[kFaceChunk], flc[kFaceChunk], mc[kFaceChunk], dc[kFaceChunk], cc[kFaceChunk];
#pragma unroll
      for (int k = 0; k < kFaceChunk; ++k)
        if (j0 + k < last) {
          fc[k] = faces[j0 + k];
          oc[k] = adjacent[j0 + k];
          sc[k] = sign[j0 + k];
          ec[k] = entry[j0 + k];
        }
#pragma unroll
      for (int k = 0; k < kFaceChunk; ++k)
        if (j0 + k < last) {
          flc[k] = flux[fc[k]];
          mc[k] = mu[fc[k]];
          dc[k] = diffusion[fc[k]];
          cc[k] = correction[fc[k]];
        }
#pragma unroll
      for (int k = 0; k < kFaceChunk; ++k) {
        if (j0 + k >= last) break;
        const int f = fc[k], o = oc[k];
        const float mass = sc[k] * flc[k], inflow = kmin(mass, 0.0f);
        float d = mc[k] * dc[k];
        ap = upper_add(ap, mass);
        if (o >= 0) {
          row.add(ec[k], -d + inflow);
        } else {
          const int bf = f - internal;
          if (kind[bf] != 0) d = 0.0f;
          ap = upper_add(ap, d);
          ap = upper_add(ap, -inflow);
          const float value = boundary[bf];
          if (compensated) {
            const F2 factor = two_sum(d, -inflow);
            const float product = factor.hi * value;
            const F2 sum = two_sum(b, product);
            carry += sum.lo + __builtin_fmaf(factor.hi, value, -product) + factor.lo * value;
            b = sum.hi;
          } else {
            b += (d - inflow) * value;
          }
        }
        const float deferred = sc[k] * cc[k];
        if (compensated) {
          const F2 sum = two_sum(b, deferred);
          carry += sum.lo;
          b = sum.hi;
        } else {
          b += deferred;
        }
      }
    }
    row.each_off_diagonal([&](float v) { ap = upper_add(ap, -v); });
    const float physical = upper_value(ap);
    diagonal[i] = physical;
    if (compensated) {
      const F2 gap = two_sum(ap.sum, -physical);
      diagonal_low[i] = gap.hi + (gap.lo + ap.error + __builtin_fmaf(sink[i], vol, -sink_volume));
    }
    const float pseudo_time = dt > 0.0f ? rho * vol / dt : 0.0f;
    const float relaxed = add_up(physical, pseudo_time) / relax;
    const float previous = old[i];
    if (compensated) {
      const F2 factor = two_sum(relaxed, -physical);
      const float product = factor.hi * previous;
      const F2 sum = two_sum(b, product);
      carry += sum.lo + __builtin_fmaf(factor.hi, previous, -product) + factor.lo * previous;
      b = sum.hi;
    } else {
      b += (relaxed - physical) * previous;
    }
    row.set_diagonal(relaxed);
    rhs[i] = b;
    if (compensated) rhs_low[i] = carry;
    UpperSum denominator{relaxed, 0.0f};
    row.each_off_diagonal([&](float v) { denominator = upper_add(denominator, v); });
    row.store();
    const float consistent = upper_value(denominator);
    consistent_inverse[i] = 1.0f / kmax(consistent > 0.0f ? consistent : relaxed, 1.0e-20f);
  }
};
// :115-124
struct FvFixedRows {
  __device__ void operator()(int i, CI rows, CI diag, CI mask, CF fixed, WF values, WF rhs) const {
    if (mask[i] == 0) return;
    const int di = diag[i];
    const float scale = values[di];
    for (int j = rows[i]; j < rows[i + 1]; ++j) values[j] = 0.0f;
    values[di] = scale;
    rhs[i] = scale * fixed[i];
  }
};
// :126-128
struct FvFloor {
  __device__ void operator()(int i, WF field, float floor) const { field[i] = kmax(field[i], floor); }
};
// :130-135
struct FvScalarSteadyMeasure {
  __device__ void operator()(int i, CF field, CF diagonal, CF residual, CI excluded, WF absolute,
                             WF scale, int has_excluded) const {
    const bool omit = has_excluded && excluded[i] != 0;
    absolute[i] = omit ? 0.0f : fabsf(residual[i]);
    scale[i] = omit ? 0.0f : fabsf(diagonal[i] * field[i]);
  }
};
// :137-139
struct FvScalarSetFixed {
  __device__ void operator()(int i, CI fixed, CF value, WF field) const {
    if (fixed[i]) field[i] = value[i];
  }
};
// :141-148
struct FvScalarMonitor {
  __device__ void operator()(int i, CF field, CF diagonal, CF residual, CI excluded, CF volume,
                             WF absolute, WF scale, WF fluent_scale, float rho, float dt) const {
    const bool omit = excluded[i] != 0;
    const float d = diagonal[i], v = field[i];
    absolute[i] = omit ? 0.0f : fabsf(residual[i]);
    scale[i] = omit ? 0.0f : fabsf(d * v);
    const float pseudo_time = dt > 0.0f ? rho * volume[i] / dt : 0.0f;
    fluent_scale[i] = omit ? 0.0f : fabsf((d + pseudo_time) * v);
  }
};

// ------------------------------------------------------ fv_paired.kern.h
// Pair additions with hysteresis (kernels/prelude.h:133-171): the high word
// stays put while the low word absorbs increments within `slack`.
__device__ __forceinline__ F2 pair_add(float hi, float lo, float d, float slack) {
  const float low = lo + d;
  if (fabsf(low) <= slack) return {hi, low};
  const F2 s = two_sum(hi, d);
  return two_sum(s.hi, s.lo + lo);
}
__device__ __forceinline__ F2 pair_sum3(float hi, float lo, float d) {
  const F2 base = two_sum(hi, lo), increment = two_sum(base.lo, d);
  const F2 value = two_sum(base.hi, increment.hi);
  return two_sum(value.hi, value.lo + increment.lo);
}
__device__ __forceinline__ float held_offset(F2 value, float hi) {
  const F2 difference = two_sum(value.hi, -hi);
  return difference.hi + (difference.lo + value.lo);
}
__device__ __forceinline__ F2 pair_add_relative(float hi, float lo, float d, float relative_slack) {
  const F2 base = two_sum(hi, lo);
  const F2 value = pair_sum3(hi, lo, d);
  const float slack = relative_slack * kmax(fabsf(base.hi), fabsf(value.hi));
  const float offset = held_offset(value, hi);
  if (fabsf(offset) <= slack) return {hi, offset};
  return value;
}
__device__ __forceinline__ bool same_sign(float a, float b) {
  return (a > 0.0f && b > 0.0f) || (a < 0.0f && b < 0.0f) || (a == 0.0f && b == 0.0f);
}
__device__ __forceinline__ F2 pair_add_flux(float hi, float lo, float d, float slack) {
  const F2 value = pair_sum3(hi, lo, d);
  const float offset = held_offset(value, hi);
  if (same_sign(hi, value.hi) && fabsf(offset) <= slack) return {hi, offset};
  return value;
}
__device__ __forceinline__ float link_trust(float dx, float dy, float dz, float mx, float my,
                                            float mz) {
  const float d2 = dx * dx + dy * dy + dz * dz, m2 = mx * mx + my * my + mz * mz;
  return kmax(0.0f, kmin(1.0f, (4.0f * d2 - m2) / (3.0f * d2)));
}
__device__ __forceinline__ float trace_difference(F2 a, F2 b) {
  F2 d = two_sum(a.hi, -b.hi);
  d = product_add(d, a.lo, 1.0f);
  d = product_add(d, -b.lo, 1.0f);
  return d.hi + d.lo;
}
__device__ __forceinline__ float face_slack(float slack_per_area, float sx, float sy, float sz) {
  return slack_per_area * sqrtf(sx * sx + sy * sy + sz * sz);
}

// :5-33
struct FvCompensatedPressureJump {
  __device__ void operator()(int i, CI owner, CI neighbour, CF px, CF py, CF pz, CF nx, CF ny,
                             CF nz, CF p, CF gx, CF gy, CF gz, CF boundary, WF jump,
                             int internal) const {
    const int c = owner[i];
    const float ox = px[i], oy = py[i], oz = pz[i];
    const bool face = i < internal;
    const int n = face ? neighbour[i] : 0;
    F2 full = two_sum(p[c], face ? -p[n] : -boundary[i - internal]);
    full = product_add(full, gx[c], ox);
    full = product_add(full, gy[c], oy);
    full = product_add(full, gz[c], oz);
    if (!face) {
      jump[i] = full.hi + full.lo;
      return;
    }
    const float qx = nx[i], qy = ny[i], qz = nz[i];
    full = product_add(full, -gx[n], qx);
    full = product_add(full, -gy[n], qy);
    full = product_add(full, -gz[n], qz);
    const float side = full.hi + full.lo;
    const float trust = link_trust(ox - qx, oy - qy, oz - qz, 0.5f * (ox + qx), 0.5f * (oy + qy), 0.5f * (oz + qz));
    if (trust == 1.0f) {
      jump[i] = side;
      return;
    }
    F2 link = two_sum(p[c], -p[n]);
    link = product_add(link, gx[c], 0.5f * ox); link = product_add(link, gx[n], 0.5f * ox);
    link = product_add(link, gx[c], -0.5f * qx); link = product_add(link, gx[n], -0.5f * qx);
    link = product_add(link, gy[c], 0.5f * oy); link = product_add(link, gy[n], 0.5f * oy);
    link = product_add(link, gy[c], -0.5f * qy); link = product_add(link, gy[n], -0.5f * qy);
    link = product_add(link, gz[c], 0.5f * oz); link = product_add(link, gz[n], 0.5f * oz);
    link = product_add(link, gz[c], -0.5f * qz); link = product_add(link, gz[n], -0.5f * qz);
    const float shared = link.hi + link.lo;
    jump[i] = shared + trust * (side - shared);
  }
};
// :46-65
struct FvPairedPressureJump {
  __device__ void operator()(int i, CI owner, CI neighbour, CF px, CF py, CF pz, CF nx, CF ny,
                             CF nz, CF p, CF low, CF gx, CF gy, CF gz, CF lx, CF ly, CF lz,
                             CF boundary, CF boundary_low, WF jump, int internal) const {
    const int c = owner[i];
    const float ox = px[i], oy = py[i], oz = pz[i];
    const F2 own = pair_trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], ox, oy, oz);
    if (i >= internal) {
      jump[i] = trace_difference(own, {boundary[i - internal], boundary_low[i - internal]});
      return;
    }
    const int n = neighbour[i];
    const float qx = nx[i], qy = ny[i], qz = nz[i];
    const F2 next = pair_trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], qx, qy, qz);
    const float side = trace_difference(own, next);
    const float dx = ox - qx, dy = oy - qy, dz = oz - qz;
    const float trust = link_trust(dx, dy, dz, 0.5f * (ox + qx), 0.5f * (oy + qy), 0.5f * (oz + qz));
    if (trust == 1.0f) {
      jump[i] = side;
      return;
    }
    const F2 a = pair_trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], .5f * dx, .5f * dy, .5f * dz);
    const F2 b = pair_trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], -.5f * dx, -.5f * dy, -.5f * dz);
    const float shared = trace_difference(a, b);
    jump[i] = shared + trust * (side - shared);
  }
};
// :69-85
struct FvCompensatedRhieChow {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF sx, CF sy, CF sz,
                             CF diffusion, CF volume, CF diagonal, CF u, CF v, CF w, CF lu, CF lv,
                             CF lw, CF jump, WF flux, WF low, int internal, float rho) const {
    const bool boundary = i >= internal;
    const int k = boundary ? kind[i - internal] : -1;
    if (k == 0 || k == 3) {
      flux[i] = 0.0f; low[i] = 0.0f;
      return;
    }
    const int c = owner[i];
    float rc = volume[c] / kmax(diagonal[c], 1.0e-20f);
    if (!boundary) {
      const int n = neighbour[i];
      rc = .5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
    }
    const float ax = sx[i], ay = sy[i], az = sz[i];
    F2 value{0.0f, 0.0f};
    value = product_add(value, u[i], ax); value = product_add(value, lu[i], ax);
    value = product_add(value, v[i], ay); value = product_add(value, lv[i], ay);
    value = product_add(value, w[i], az); value = product_add(value, lw[i], az);
    if (k != 1) value = product_add(value, rc * diffusion[i], jump[i]);
    const F2 scaled = add_weighted({0.0f, 0.0f}, rho, value);
    flux[i] = scaled.hi;
    low[i] = scaled.lo;
  }
};
// :89-97
struct FvPairedPressureBoundaryLow {
  __device__ void operator()(int i, CI owner, CI kind, CF px, CF py, CF pz, CF p, CF low, CF gx,
                             CF gy, CF gz, CF lx, CF ly, CF lz, CF boundary, WF boundary_low,
                             int internal) const {
    const int k = kind[i];
    if (k == 2 || k == 4) {
      boundary_low[i] = 0.0f;
      return;
    }
    const int f = i + internal, c = owner[f];
    F2 value = pair_trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[f], py[f], pz[f]);
    value = product_add(value, -boundary[i], 1.0f);
    boundary_low[i] = value.hi + value.lo;
  }
};
// :98-128
struct FvPairedPressureFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CI poor, CF px, CF py, CF pz,
                             CF nx, CF ny, CF nz, CF p, CF low, CF gx, CF gy, CF gz, CF lx,
                             CF ly, CF lz, CF boundary, CF boundary_low, CF diagonal, WF face,
                             WF face_low, int internal, int standard) const {
    const int c = owner[i];
    F2 value = two_sum(p[c], low[c]);
    if (i < internal) {
      const int n = neighbour[i];
      if (standard) {
        const float wc = 1.0f / kmax(diagonal[c], 1.0e-20f), wn = 1.0f / kmax(diagonal[n], 1.0e-20f);
        F2 num{0.0f, 0.0f};
        num = product_add(num, wc, p[c]); num = product_add(num, wc, low[c]);
        num = product_add(num, wn, p[n]); num = product_add(num, wn, low[n]);
        const F2 den = two_sum(wc, wn);
        const float q = num.hi / den.hi;
        num = product_add(num, -q, den.hi);
        num = product_add(num, -q, den.lo);
        value = two_sum(q, (num.hi + num.lo) / den.hi);
      } else {
        const bool plain = poor[c] || poor[n];
        const F2 own = plain ? two_sum(p[c], low[c])
                             : pair_trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[i], py[i], pz[i]);
        const F2 next = plain ? two_sum(p[n], low[n])
                              : pair_trace(p[n], low[n], gx[n], gy[n], gz[n], lx[n], ly[n], lz[n], nx[i], ny[i], nz[i]);
        value = two_sum(0.5f * own.hi, 0.5f * next.hi);
        value = two_sum(value.hi, value.lo + 0.5f * own.lo + 0.5f * next.lo);
      }
    } else {
      const int b = i - internal, k = kind[b];
      if (k == 2 || k == 4) value = two_sum(boundary[b], boundary_low[b]);
      else if (!standard && !poor[c])
        value = pair_trace(p[c], low[c], gx[c], gy[c], gz[c], lx[c], ly[c], lz[c], px[i], py[i], pz[i]);
    }
    face[i] = value.hi;
    face_low[i] = value.lo;
  }
};
// :140-152
struct FvPairedMomentumResidual {
  __device__ void operator()(int i, CI rows, CI cols, CF values, CI diag, CF physical,
                             CF physical_low, CF hi, CF lo, CF pressure_gradient_lo, CF volume,
                             CF rhs, CF rhs_low, WF r) const {
    float sum = rhs[i], error = rhs_low[i];
    struct Entry {
      float a, x, l;
    };
    each_face(
        rows[i], rows[i + 1],
        [&](int j) {
          const int col = cols[j];
          return Entry{values[j], hi[col], lo[col]};
        },
        [&](int, const Entry& q) {
          const float product = q.a * q.x;
          const F2 s = two_sum(sum, -product);
          sum = s.hi;
          error += s.lo - __builtin_fmaf(q.a, q.x, -product) - q.a * q.l;
        });
    const float pl = physical_low[i], li = lo[i];
    error += (values[diag[i]] - physical[i]) * li - volume[i] * pressure_gradient_lo[i];
    error -= pl * hi[i] + pl * li;
    r[i] = sum + error;
  }
};
// :156-170
struct FvCompensatedMomentumResidual {
  __device__ void operator()(int i, CI rows, CI cols, CF values, CI diag, CF physical,
                             CF physical_low, CF hi, CF lo, CF pressure_gradient_lo, CF volume,
                             CF rhs, CF rhs_low, WF r) const {
    F2 value = two_sum(rhs[i], rhs_low[i]);
    const int di = diag[i];
    const float own_physical = physical[i];
    struct Entry {
      float a, x, l;
    };
    each_face(
        rows[i], rows[i + 1],
        [&](int j) {
          const int col = cols[j];
          return Entry{values[j], hi[col], lo[col]};
        },
        [&](int j, const Entry& q) {
          value = product_add(value, -q.a, q.x);
          const float low_coefficient = j == di ? own_physical : q.a;
          value = product_add(value, -low_coefficient, q.l);
        });
    value = product_add(value, -physical_low[i], hi[i]);
    value = product_add(value, -physical_low[i], lo[i]);
    value = product_add(value, -volume[i], pressure_gradient_lo[i]);
    r[i] = value.hi + value.lo;
  }
};
// :172-174
struct FvPairedAccumulate {
  __device__ void operator()(int i, CF delta, WF hi, WF lo, float slack) const {
    const F2 x = pair_add(hi[i], lo[i], delta[i], slack);
    hi[i] = x.hi; lo[i] = x.lo;
  }
};
// :179-196
struct FvCompensatedPressureForce {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF sx, CF sy, CF sz, CF volume,
                             CF pressure, CF low, WF x, WF y, WF z) const {
    float a = 0.0f, b = 0.0f, c = 0.0f;
    F2 ax{0.0f, 0.0f}, ay{0.0f, 0.0f}, az{0.0f, 0.0f};
    struct Face {
      float s, p, l, x, y, z;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j];
          return Face{sign[j], pressure[f], low[f], sx[f], sy[f], sz[f]};
        },
        [&](int, const Face& q) {
          const float hp = q.s * q.p, lp = q.s * q.l;
          const float fx = q.x, fy = q.y, fz = q.z;
          a += hp * fx; b += hp * fy; c += hp * fz;
          ax = product_add(product_add(ax, hp, fx), lp, fx);
          ay = product_add(product_add(ay, hp, fy), lp, fy);
          az = product_add(product_add(az, hp, fz), lp, fz);
        });
    const float vol = volume[i];
    ax = product_add(ax, -a / vol, vol);
    ay = product_add(ay, -b / vol, vol);
    az = product_add(az, -c / vol, vol);
    x[i] = (ax.hi + ax.lo) / vol; y[i] = (ay.hi + ay.lo) / vol; z[i] = (az.hi + az.lo) / vol;
  }
};
// :200-218
struct FvPairedVelocityBoundary {
  __device__ void operator()(int i, CI owner, CI kind, CF sx, CF sy, CF sz, CF u, CF v, CF w,
                             CF flux, CF flux_low, WF bu, WF bv, WF bw, WI scalar_kind,
                             int internal, float rho) const {
    const int f = i + internal, c = owner[f], k = kind[i];
    const float ax = sx[f], ay = sy[f], az = sz[f], m = flux[f];
    const float area2 = ax * ax + ay * ay + az * az;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f;
    int sk = 0;
    if (((k == 2 || k == 4) && m >= 0.0f) || k == 3) {
      vx = u[c]; vy = v[c]; vz = w[c];
      sk = 2;
    }
    if (k == 4 && m < 0.0f) {
      const float factor = flux_low[f] / (rho * area2);
      vx = factor * ax; vy = factor * ay; vz = factor * az;
    }
    if (k == 3) {
      const float un = (vx * ax + vy * ay + vz * az) / area2;
      vx -= un * ax; vy -= un * ay; vz -= un * az;
    }
    scalar_kind[i] = sk;
    bu[i] = vx; bv[i] = vy; bw[i] = vz;
  }
};
// :220-236
struct FvPairedAdvection {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF px, CF py, CF pz, CF nx,
                             CF ny, CF nz, CF field, CF boundary, CF gx, CF gy, CF gz, CF limiter,
                             CF flux, CF flux_low, WF correction, int internal,
                             int second_order) const {
    const int c = owner[i];
    const float m = flux[i];
    int up = c;
    float value = field[c], rx = px[i], ry = py[i], rz = pz[i];
    bool reconstruct = second_order != 0;
    if (i < internal) {
      if (m < 0.0f) {
        up = neighbour[i];
        rx = nx[i]; ry = ny[i]; rz = nz[i];
      }
      value = field[up];
    } else {
      if (m < 0.0f) value = boundary[i - internal];
      if (m <= 0.0f || kind[i - internal] == 2) reconstruct = false;
    }
    if (reconstruct) value += limiter[up] * (gx[up] * rx + gy[up] * ry + gz[up] * rz);
    correction[i] = flux_low[i] * value;
  }
};
// :238-247
struct FvPairedFaceResidual {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF secondary, CF advection,
                             CF stress, CF boundary_rhs, WF residual, int turbulent) const {
    float sum = residual[i], error = boundary_rhs[i];
    struct Face {
      float s, a, b, c;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j];
          return Face{sign[j], secondary[f], advection[f], turbulent ? stress[f] : 0.0f};
        },
        [&](int, const Face& q) {
          const float value = q.s * (q.a - q.b + q.c);
          const F2 t = two_sum(sum, value);
          sum = t.hi;
          error += t.lo;
        });
    residual[i] = sum + error;
  }
};
// :259-290
struct FvPairedFlux {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF sx, CF sy, CF sz, CF px,
                             CF py, CF pz, CF nx, CF ny, CF nz, CF weights, CF diagonal,
                             CF volume, CF diffusion, CF u, CF v, CF w, CF p, CF gx, CF gy, CF gz,
                             CF previous, WF flux, WF flux_lo, int internal, float rho,
                             float slack_per_area) const {
    const int c = owner[i];
    const float ax = sx[i], ay = sy[i], az = sz[i];
    float rc = volume[c] / kmax(diagonal[c], 1.0e-20f), low = 0.0f;
    if (i < internal) {
      const int n = neighbour[i];
      const float a = weights[c] / (weights[c] + weights[n]), b = 1.0f - a;
      const float ux = a * u[c] + b * u[n], uy = a * v[c] + b * v[n], uz = a * w[c] + b * w[n];
      rc = 0.5f * (rc + volume[n] / kmax(diagonal[n], 1.0e-20f));
      const float dx = px[i] - nx[i], dy = py[i] - ny[i], dz = pz[i] - nz[i];
      const float jump = p[c] - p[n] + 0.5f * ((gx[c] + gx[n]) * dx + (gy[c] + gy[n]) * dy + (gz[c] + gz[n]) * dz);
      low = rho * (ux * ax + uy * ay + uz * az) + rho * rc * diffusion[i] * jump;
    } else {
      const int k = kind[i - internal];
      if (k == 2 || k == 4) {
        const float jump = p[c] + gx[c] * px[i] + gy[c] * py[i] + gz[c] * pz[i];
        low = rho * (u[c] * ax + v[c] * ay + w[c] * az) + rho * rc * diffusion[i] * jump;
      }
    }
    const float slack = face_slack(slack_per_area, ax, ay, az);
    const float held = previous[i];
    const F2 x = two_sum(flux[i], -held);
    const float offset = x.hi + (x.lo + low);
    if (fabsf(offset) <= slack) {
      flux[i] = held;
      flux_lo[i] = offset;
    } else {
      flux_lo[i] = low;
    }
  }
};
// :292-303
struct FvPairedFluxOffset {
  __device__ void operator()(int i, CF sx, CF sy, CF sz, CF previous, WF flux, WF low,
                             float slack_per_area) const {
    const float slack = face_slack(slack_per_area, sx[i], sy[i], sz[i]);
    const float held = previous[i];
    const F2 total = two_sum(flux[i], low[i]);
    const F2 x = two_sum(total.hi, -held);
    const float offset = x.hi + (x.lo + total.lo);
    if (same_sign(held, total.hi) && fabsf(offset) <= slack) {
      flux[i] = held; low[i] = offset;
    } else {
      flux[i] = total.hi; low[i] = total.lo;
    }
  }
};
// :308-318
struct FvPairedContinuity {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF flux, CF flux_lo,
                             WF imbalance, WF magnitude) const {
    float sum = 0.0f, error = 0.0f;
    struct Face {
      float s, hi, lo;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j];
          return Face{sign[j], flux[f], flux_lo[f]};
        },
        [&](int, const Face& q) {
          const F2 t = two_sum(sum, q.s * q.hi);
          sum = t.hi;
          error += t.lo + q.s * q.lo;
        });
    const float mass = sum + error;
    imbalance[i] = mass;
    magnitude[i] = fabsf(mass);
  }
};
// :321-323
struct FvPairedPressureRhs {
  __device__ void operator()(int i, CF imbalance, WF rhs, int reference) const {
    rhs[i] = i == reference ? 0.0f : -imbalance[i];
  }
};
// :326-333
struct FvPairedCorrectFlux {
  __device__ void operator()(int i, CI owner, CI neighbour, CF sx, CF sy, CF sz, CF correction,
                             CF coefficient, WF flux, WF flux_lo, int internal,
                             float slack_per_area) const {
    const float pc = correction[owner[i]];
    const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
    const float slack = face_slack(slack_per_area, sx[i], sy[i], sz[i]);
    const F2 x = pair_add(flux[i], flux_lo[i], coefficient[i] * (pc - pn), slack);
    flux[i] = x.hi; flux_lo[i] = x.lo;
  }
};
// Velocity and pressure pair corrections of fv_paired_correct_velocity
// (:336-347) and fv_compensated_correct_velocity (:365-377), which differ
// only in the pressure pair's hysteresis rule.
template <bool Relative>
__device__ __forceinline__ void correct_velocity_pairs(int i, CF volume, CF inverse, CF diagonal,
                                                       CF gx, CF gy, CF gz, CF correction, WF u,
                                                       WF v, WF w, WF p, WF u_lo, WF v_lo,
                                                       WF w_lo, WF p_lo, float alpha,
                                                       float velocity_slack,
                                                       float pressure_slack) {
  const float inv = inverse[i], d = volume[i] * inv;
  F2 x = pair_add(u[i], u_lo[i], -d * gx[i], velocity_slack);
  u[i] = x.hi; u_lo[i] = x.lo;
  x = pair_add(v[i], v_lo[i], -d * gy[i], velocity_slack);
  v[i] = x.hi; v_lo[i] = x.lo;
  x = pair_add(w[i], w_lo[i], -d * gz[i], velocity_slack);
  w[i] = x.hi; w_lo[i] = x.lo;
  const float step = kmin(alpha, 0.7f * diagonal[i] * inv) * correction[i];
  x = Relative ? pair_add_relative(p[i], p_lo[i], step, pressure_slack)
               : pair_add(p[i], p_lo[i], step, pressure_slack);
  p[i] = x.hi; p_lo[i] = x.lo;
}
struct FvPairedCorrectVelocity {
  __device__ void operator()(int i, CF volume, CF inverse, CF diagonal, CF gx, CF gy, CF gz,
                             CF correction, WF u, WF v, WF w, WF p, WF u_lo, WF v_lo, WF w_lo,
                             WF p_lo, float alpha, float velocity_slack,
                             float pressure_slack) const {
    correct_velocity_pairs<false>(i, volume, inverse, diagonal, gx, gy, gz, correction, u, v, w, p,
                                  u_lo, v_lo, w_lo, p_lo, alpha, velocity_slack, pressure_slack);
  }
};
// :349-351
struct FvPairedValue {
  __device__ void operator()(int i, CF hi, CF lo, WF value) const { value[i] = hi[i] + lo[i]; }
};
// :352-355
struct FvPairedVectorValues {
  __device__ void operator()(int i, CF x, CF y, CF z, CF lx, CF ly, CF lz, WF u, WF v,
                             WF w) const {
    u[i] = x[i] + lx[i]; v[i] = y[i] + ly[i]; w[i] = z[i] + lz[i];
  }
};
// :356-363
struct FvCompensatedCorrectFlux {
  __device__ void operator()(int i, CI owner, CI neighbour, CF sx, CF sy, CF sz, CF correction,
                             CF coefficient, WF flux, WF flux_lo, int internal,
                             float slack_per_area) const {
    const float pn = i < internal ? correction[neighbour[i]] : 0.0f;
    const float slack = face_slack(slack_per_area, sx[i], sy[i], sz[i]);
    const F2 x = pair_add_flux(flux[i], flux_lo[i], coefficient[i] * (correction[owner[i]] - pn), slack);
    flux[i] = x.hi; flux_lo[i] = x.lo;
  }
};
struct FvCompensatedCorrectVelocity {
  __device__ void operator()(int i, CF volume, CF inverse, CF diagonal, CF gx, CF gy, CF gz,
                             CF correction, WF u, WF v, WF w, WF p, WF u_lo, WF v_lo, WF w_lo,
                             WF p_lo, float alpha, float velocity_slack,
                             float pressure_slack) const {
    correct_velocity_pairs<true>(i, volume, inverse, diagonal, gx, gy, gz, correction, u, v, w, p,
                                 u_lo, v_lo, w_lo, p_lo, alpha, velocity_slack, pressure_slack);
  }
};

// -------------------------------------------------------- fv_rans.kern.h
// :4-10
struct FvBackflowPressure {
  __device__ void operator()(int i, CI owner, CI total, CF mass, CF u, CF v, CF w, WF boundary,
                             int internal, float rho) const {
    const int f = i + internal;
    if (total[i] != 0 && mass[f] < 0.0f) {
      const int c = owner[f];
      boundary[i] -= 0.5f * rho * (u[c] * u[c] + v[c] * v[c] + w[c] * w[c]);
    }
  }
};
// :12-17
struct FvEffectiveViscosity {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF weight, CF mut, WF mu,
                             int internal, float molecular) const {
    float turbulent = mut[owner[i]];
    if (i < internal) {
      const float t = weight[i];
      turbulent = t * turbulent + (1.0f - t) * mut[neighbour[i]];
    }
    mu[i] = i >= internal && kind[i - internal] == 0 ? 0.0f : molecular + turbulent;
  }
};
// :21-37
struct FvTransposeStress {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF weight, CF mu, CF sx,
                             CF sy, CF sz, CF ux, CF uy, CF uz, CF vx, CF vy, CF vz, CF wx, CF wy,
                             CF wz, WF x, WF y, WF z, int internal) const {
    const int c = owner[i];
    float a = ux[c], b = uy[c], d = uz[c], e = vx[c], g = vy[c], h = vz[c], j = wx[c], k = wy[c], l = wz[c];
    if (i < internal) {
      const int n = neighbour[i];
      const float t = weight[i], s = 1.0f - t;
      a = t * a + s * ux[n]; b = t * b + s * uy[n]; d = t * d + s * uz[n];
      e = t * e + s * vx[n]; g = t * g + s * vy[n]; h = t * h + s * vz[n];
      j = t * j + s * wx[n]; k = t * k + s * wy[n]; l = t * l + s * wz[n];
    }
    const float m = i >= internal && kind[i - internal] == 3 ? 0.0f : mu[i];
    const float ax = sx[i], ay = sy[i], az = sz[i];
    x[i] = m * (a * ax + e * ay + j * az);
    y[i] = m * (b * ax + g * ay + k * az);
    z[i] = m * (d * ax + h * ay + l * az);
  }
};
// :38-46
struct FvStressSource {
  __device__ void operator()(int i, CI offsets, CI faces, CF sign, CF volume, CF x, CF y, CF z,
                             WF su, WF sv, WF sw) const {
    float a = 0.0f, b = 0.0f, c = 0.0f;
    struct Face {
      float s, x, y, z;
    };
    each_face(
        offsets[i], offsets[i + 1],
        [&](int j) {
          const int f = faces[j];
          return Face{sign[j], x[f], y[f], z[f]};
        },
        [&](int, const Face& q) {
          a += q.s * q.x; b += q.s * q.y; c += q.s * q.z;
        });
    const float vol = volume[i];
    su[i] += a / vol; sv[i] += b / vol; sw[i] += c / vol;
  }
};
// :47-72
struct FvBoundaryMomentum {
  __device__ void operator()(int i, CI offsets, CI faces, CI kind, CF sx, CF sy, CF sz,
                             CF coefficient, CF mu, CF diffusion, CF u, CF v, CF w, CF bu, CF bv,
                             CF bw, WF diagonal, WF rhs, int internal, int component,
                             int wall_function) const {
    float ap = 0.0f, b = 0.0f;
    const float u0 = u[i], v0 = v[i], w0 = w[i];
    const float uc = component == 0 ? u0 : (component == 1 ? v0 : w0);
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int f = faces[j];
      if (f < internal) continue;
      const int bf = f - internal, k = kind[bf];
      if (k != 3 && !(k == 0 && wall_function)) continue;
      const float ax = sx[f], ay = sy[f], az = sz[f];
      const float area = sqrtf(ax * ax + ay * ay + az * az);
      const float ex = ax / area, ey = ay / area, ez = az / area;
      const float nc = component == 0 ? ex : (component == 1 ? ey : ez);
      if (k == 3) {
        const float md = mu[f] * diffusion[f];
        ap += md * nc * nc;
        b -= md * nc * (ex * u0 + ey * v0 + ez * w0 - nc * uc);
        continue;
      }
      const float wx = bu[bf], wy = bv[bf], wz = bw[bf];
      const float wc = component == 0 ? wx : (component == 1 ? wy : wz);
      const float c = area * coefficient[bf];
      ap += c * (1.0f - nc * nc);
      b += c * (wc - nc * (ex * wx + ey * wy + ez * wz) + nc * (ex * u0 + ey * v0 + ez * w0 - nc * uc));
    }
    diagonal[i] = ap;
    rhs[i] = b;
  }
};
// :75-78
struct FvMomentumDiagonal {
  __device__ void operator()(int i, CF ax, CF ay, CF az, CF ix, CF iy, CF iz, WF physical,
                             WF inverse) const {
    physical[i] = (ax[i] + ay[i] + az[i]) / 3.0f;
    inverse[i] = (ix[i] + iy[i] + iz[i]) / 3.0f;
  }
};
// :79-83
struct FvSstForce {
  __device__ void operator()(int i, CF selected, CF sx, CF sy, CF sz, CF tx, CF ty, CF tz, WF x,
                             WF y, WF z, int internal) const {
    const int f = i + internal;
    const float a = -selected[i] * sqrtf(sx[f] * sx[f] + sy[f] * sy[f] + sz[f] * sz[f]);
    x[i] = a * tx[i]; y[i] = a * ty[i]; z[i] = a * tz[i];
  }
};
// :86-101
struct FvTurbulenceBoundary {
  __device__ void operator()(int i, CI owner, CI kind, CI spec, CF a, CF b, CF mass, CF bu, CF bv,
                             CF bw, CF k, CF omega, CF wall_omega, WF bk, WF bo,
                             WI transport_kind, int internal, float rho, float mu) const {
    const int f = i + internal, c = owner[f], type = kind[i];
    float kinetic = k[c], dissipation = omega[c];
    int tk = 1;
    if (type == 1 || ((type == 2 || type == 4) && mass[f] < 0.0f)) {
      const float ai = a[i], bi = b[i];
      kinetic = ai;
      dissipation = bi;
      const int sp = spec[i];
      if (sp != 0) {
        const float x = bu[i], y = bv[i], z = bw[i];
        kinetic = 1.5f * ai * ai * (x * x + y * y + z * z);
        dissipation = sp == 1 ? sqrtf(kinetic) / (0.5477225575f * bi) : rho * kinetic / (mu * bi);
      }
      tk = 0;
    }
    transport_kind[i] = tk;
    bk[i] = kinetic;
    bo[i] = type == 0 ? wall_omega[i] : dissipation;
  }
};

// ------------------------------------------------------------ sst.kern.h
// k-omega SST coefficient blends and limiters (:5-38). tanh, log and pow come
// from the device math library, so those results agree with the serial
// reference to a few ulps rather than bit for bit.
__device__ __forceinline__ float sst_inv_sigma_k(float f1) { return f1 / 1.176f + (1.0f - f1); }
__device__ __forceinline__ float sst_inv_sigma_w(float f1) { return f1 / 2.0f + (1.0f - f1) / 1.168f; }
__device__ __forceinline__ float sst_beta(float f1) { return 0.075f * f1 + 0.0828f * (1.0f - f1); }
__device__ __forceinline__ float sst_gamma(float f1, int set_2003) {
  const float g1 = set_2003 ? 5.0f / 9.0f : 0.075f / 0.09f - 0.41f * 0.41f / (2.0f * 0.3f);
  const float g2 = set_2003 ? 0.44f : 0.0828f / 0.09f - 0.41f * 0.41f / (1.168f * 0.3f);
  return f1 * g1 + (1.0f - f1) * g2;
}
__device__ __forceinline__ float sst_alpha_star(float ret, int low_re) {
  return low_re ? 1.0f - 0.976f / (1.0f + ret / 6.0f) : 1.0f;
}
__device__ __forceinline__ float sst_beta_star(float ret, int low_re) {
  if (!low_re) return 0.09f;
  const float r = kmin(ret / 8.0f, 100.0f), r4 = r * r * r * r;
  return 0.09f * ((4.0f / 15.0f) + r4) / (1.0f + r4);
}
__device__ __forceinline__ float sst_power_blend(float a, float b, float p) {
  const float hi = kmax(a, b), lo = kmin(a, b);
  return hi == 0.0f ? 0.0f : hi * powf(1.0f + powf(lo / hi, p), 1.0f / p);
}
__device__ __forceinline__ float sst_finite_ratio(float numerator, float denominator) {
  return kmax(-1e38f, kmin(numerator / denominator, 1e38f));
}
// :40-50
struct SstRates {
  __device__ void operator()(int i, CF ux, CF uy, CF uz, CF vx, CF vy, CF vz, CF wx, CF wy, CF wz,
                             WF strain, WF vorticity) const {
    const float a = ux[i], b = uy[i], c = uz[i], d = vx[i], e = vy[i], f = vz[i], g = wx[i], h = wy[i], l = wz[i];
    const float sxy = b + d, sxz = c + g, syz = f + h;
    strain[i] = sqrtf(2.0f * (a * a + e * e + l * l) + sxy * sxy + sxz * sxz + syz * syz);
    const float wxy = b - d, wxz = c - g, wyz = f - h;
    vorticity[i] = sqrtf(wxy * wxy + wxz * wxz + wyz * wyz);
  }
};
// :52-70
struct SstBlending {
  __device__ void operator()(int i, CF rho, CF mu, CF k, CF omega, CF distance, CF kx, CF ky, CF kz,
                             CF ox, CF oy, CF oz, WF f1, WF f2, WF cdkw, float cd_floor,
                             float k_floor, float omega_floor) const {
    const float r = rho[i], ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor), y = distance[i];
    const float cd = sst_finite_ratio(2.0f * r * (kx[i] * ox[i] + ky[i] * oy[i] + kz[i] * oz[i]), 1.168f * w);
    const float a = sqrtf(ke) / (0.09f * w * y);
    const float b = 500.0f * (mu[i] / r) / (y * y * w);
    const float c = 4.0f * r * ke / (1.168f * kmax(cd, cd_floor) * y * y);
    const float p1 = kmin(kmin(kmax(a, b), c), 10.0f);
    const float p2 = kmin(kmax(2.0f * a, b), 10.0f);
    f1[i] = tanhf(p1 * p1 * p1 * p1);
    f2[i] = tanhf(p2 * p2);
    cdkw[i] = cd;
  }
};
// :72-88
struct SstViscosity {
  __device__ void operator()(int i, CF rho, CF mu, CF k, CF omega, CF strain, CF vorticity, CF f1,
                             CF f2, WF mut, WF gamma_k, WF gamma_w, int low_re, int use_vorticity,
                             float max_ratio, float k_floor, float omega_floor) const {
    const float r = rho[i], m = mu[i];
    const float ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor);
    const float ret = (r / m) * (ke / w);
    const float alpha = sst_alpha_star(ret, low_re);
    const float rate = use_vorticity ? vorticity[i] : strain[i];
    float mt = r * ke / kmax(w / alpha, rate * f2[i] / 0.31f);
    mt = kmin(mt, max_ratio * m);
    mut[i] = mt;
    const float blend = f1[i];
    gamma_k[i] = m + mt * sst_inv_sigma_k(blend);
    gamma_w[i] = m + mt * sst_inv_sigma_w(blend);
  }
};
// :90-103
struct SstBlendingViscosity {
  __device__ void operator()(int i, CF rho, CF mu, CF k, CF omega, CF distance, CF kx, CF ky, CF kz,
                             CF ox, CF oy, CF oz, CF strain, CF vorticity, WF f1, WF f2, WF cdkw,
                             WF mut, WF gamma_k, WF gamma_w, float cd_floor, float k_floor,
                             float omega_floor, int low_re, int use_vorticity,
                             float max_ratio) const {
    SstBlending{}(i, rho, mu, k, omega, distance, kx, ky, kz, ox, oy, oz, f1, f2, cdkw, cd_floor,
                  k_floor, omega_floor);
    SstViscosity{}(i, rho, mu, k, omega, strain, vorticity, f1, f2, mut, gamma_k, gamma_w, low_re,
                   use_vorticity, max_ratio, k_floor, omega_floor);
  }
};
// :105-151
struct SstWall {
  __device__ void operator()(int i, CI owner, CI wall, int n_internal, CF ax, CF ay, CF az, CF ux,
                             CF uy, CF uz, CF wall_ux, CF wall_uy, CF wall_uz, CF dy, CF rho,
                             CF mu, CF k, CF omega, WF tau, WF shear_x, WF shear_y, WF shear_z,
                             WF u_star, WF u_tau, WF y_plus, WF y_plus_utau, WF production,
                             WF shear_coefficient, float k_floor, float omega_floor) const {
    float t_tau = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, t_us = 0.0f, t_ut = 0.0f, t_yp = 0.0f;
    float t_ypu = 0.0f, t_prod = 0.0f, t_coef = 0.0f;
    if (wall[i]) {
      const int f = n_internal + i, c = owner[f];
      const float r = rho[c], m = mu[c], d = dy[i], ke = kmax(k[c], k_floor);
      const float w = kmax(omega[c], omega_floor);
      const float fx = ax[f], fy = ay[f], fz = az[f];
      const float inv = 1.0f / sqrtf(fx * fx + fy * fy + fz * fz);
      const float nx = inv * fx, ny = inv * fy, nz = inv * fz;
      const float rx = ux[c] - wall_ux[i], ry = uy[c] - wall_uy[i], rz = uz[c] - wall_uz[i];
      const float along = rx * nx + ry * ny + rz * nz;
      const float tx = rx - along * nx, ty = ry - along * ny, tz = rz - along * nz;
      const float speed = sqrtf(tx * tx + ty * ty + tz * tz);
      const float us = sqrtf((m / r) * speed / d + 0.3f * ke);
      const float yp = (r / m) * us * d;
      const float log_up = logf(9.793f * kmax(yp, 0.2f)) / 0.4187f;
      const float ratio = kmin(yp, log_up) / kmax(yp, log_up);
      const float blend = powf(1.0f + ratio * ratio * ratio * ratio, 0.25f);
      const float coefficient = (yp <= log_up ? m / d : r * us / log_up) * blend;
      const float stress = coefficient * speed;
      const float ut = sqrtf(stress / r);
      t_coef = coefficient; t_tau = stress; t_us = us; t_ut = ut; t_yp = yp;
      t_ypu = (r / m) * ut * d;
      if (speed > 0.0f) {
        sx = (-coefficient) * tx; sy = (-coefficient) * ty; sz = (-coefficient) * tz;
      }
      const float p_lam = (r * ke / w) * (stress / m) * (stress / m);
      const float p_log = yp > 0.0f ? (stress * stress / m) / (0.4187f * yp) : 0.0f;
      const float hi = kmax(p_lam, p_log), lo = kmin(p_lam, p_log);
      t_prod = hi > 0.0f ? lo / (1.0f + lo / hi) : 0.0f;
    }
    tau[i] = t_tau; shear_x[i] = sx; shear_y[i] = sy; shear_z[i] = sz;
    u_star[i] = t_us; u_tau[i] = t_ut; y_plus[i] = t_yp; y_plus_utau[i] = t_ypu;
    production[i] = t_prod; shear_coefficient[i] = t_coef;
  }
};
// :153-169
struct SstWallOmega {
  __device__ void operator()(int i, CI owner, CI wall, int n_internal, CF dy, CF rho, CF mu,
                             CF u_star, WF fixed_omega, WF face_omega) const {
    float fixed = 0.0f, face = 0.0f;
    if (wall[i]) {
      const int c = owner[n_internal + i];
      const float m = mu[c], r = rho[c], d = dy[i];
      const float w_lam = 2.0f * (m / r) / (0.072f * d * d);
      const float w_log = u_star[i] / (0.3f * 0.4187f * d);
      fixed = sst_power_blend(w_lam, w_log, 1.3f);
      face = sst_power_blend(3.0f * w_lam, 2.0f * w_log, 1.3f);
    }
    fixed_omega[i] = fixed;
    face_omega[i] = face;
  }
};
// :171-190
struct SstWallCells {
  __device__ void operator()(int i, CI offsets, CI faces, CI wall, CF dy, CF face_omega,
                             CF face_production, int n_internal, WI fixed_mask, WI selected_face,
                             WF fixed_omega, WF production) const {
    int selected = -1;
    float best = 0.0f;
    for (int j = offsets[i]; j < offsets[i + 1]; ++j) {
      const int b = faces[j] - n_internal;
      if (b >= 0 && wall[b]) {
        const float d = dy[b];
        if (selected < 0 || d < best || (d == best && b < selected)) {
          selected = b;
          best = d;
        }
      }
    }
    selected_face[i] = selected;
    fixed_mask[i] = selected >= 0 ? 1 : 0;
    fixed_omega[i] = selected >= 0 ? face_omega[selected] : 0.0f;
    production[i] = selected >= 0 ? face_production[selected] : 0.0f;
  }
};
// :192-236
struct SstSources {
  __device__ void operator()(int i, CF rho, CF mu, CF k, CF omega, CF strain, CF vorticity, CF f1,
                             CF cdkw, CF mut, CI wall_mask, CF wall_pk, WF k_source, WF k_sink,
                             WF omega_source, WF omega_sink, WF production_k, int low_re,
                             int gamma_2003, int production_mode, int limiter, int limit_omega,
                             int wall_override, float clip, int limit_cross, float cross_clip,
                             float k_floor, float omega_floor, float omega_molecular_ratio) const {
    const float r = rho[i], m = mu[i], ke = kmax(k[i], k_floor), w = kmax(omega[i], omega_floor);
    const float ret = (r / m) * (ke / w);
    const float beta_star = sst_beta_star(ret, low_re);
    const float s = strain[i], v = vorticity[i];
    const float rate2 = production_mode == 1 ? s * v : (production_mode == 2 ? v * v : s * s);
    const float mt = mut[i], blend = f1[i];
    float raw_pk = mt * rate2;
    if (wall_override && wall_mask[i]) raw_pk = wall_pk[i];
    const float cap = clip * r * beta_star * ke * w;
    const float pk = limiter ? kmin(raw_pk, cap) : raw_pk;
    const float pw_k = limiter && limit_omega ? pk : raw_pk;
    float gamma = sst_gamma(blend, gamma_2003);
    if (low_re) gamma *= 1.0f - (8.0f / 9.0f) / (1.0f + ret / 2.95f);
    float pw = mt > 0.0f ? gamma * r * (pw_k / mt) : gamma * r * rate2;
    if (omega_molecular_ratio > 0.0f) pw = gamma * r * (pw_k / (mt + omega_molecular_ratio * m));
    float cross = (1.0f - blend) * cdkw[i];
    if (limit_cross) {
      const float bound = (cross_clip * r * sst_beta(blend) * w) * w;
      cross = kmax(-bound, kmin(cross, bound));
    }
    k_source[i] = pk;
    k_sink[i] = r * beta_star * w;
    const float destruction = r * sst_beta(blend) * w;
    omega_source[i] = pw + kmax(cross, 0.0f) + destruction * w;
    omega_sink[i] = 2.0f * destruction + sst_finite_ratio(kmax(-cross, 0.0f), w);
    production_k[i] = pk;
  }
};
// :238-253
struct SstFaceDiffusion {
  __device__ void operator()(int i, CI owner, CI neighbour, CF weight, CI wall, CF mu,
                             CF gamma_k, CF gamma_w, int n_internal, WF face_k,
                             WF face_w) const {
    const int c = owner[i];
    if (i < n_internal) {
      const int n = neighbour[i];
      const float a = weight[i], b = 1.0f - a;
      face_k[i] = a * gamma_k[c] + b * gamma_k[n];
      face_w[i] = a * gamma_w[c] + b * gamma_w[n];
    } else {
      const bool on_wall = wall[i - n_internal] != 0;
      face_k[i] = on_wall ? 0.0f : gamma_k[c];
      face_w[i] = on_wall ? mu[c] : gamma_w[c];
    }
  }
};
// :255-259
struct SstSustain {
  __device__ void operator()(int i, CF rho, CF f1, WF k_source, WF omega_source, float ambient_k,
                             float ambient_omega) const {
    const float r = rho[i];
    k_source[i] += 0.09f * r * ambient_omega * ambient_k;
    omega_source[i] += sst_beta(f1[i]) * r * ambient_omega * ambient_omega;
  }
};

// ------------------------------------------------------------------------
// kernels/fv_fluent_schemes.kern.h
// :10-12. SIMPLE's correction coefficient, the inverse relaxed diagonal.
struct FvSimpleInverse {
  __device__ void operator()(int i, CI diag, CF values, WF inverse) const {
    inverse[i] = 1.0f / kmax(values[diag[i]], 1.0e-20f);
  }
};
// :28-39. Inverse-diagonal weighted face pressure; unprescribed boundary
// faces (kinds other than 2 and 4) hold the owner's pressure.
struct FvStandardPressureFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CF p, CF diagonal, WF face,
                             WF boundary, int internal) const {
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
};
// :47-50
struct FvCellPressureBoundary {
  __device__ void operator()(int i, CI owner, CI kind, CF p, WF boundary, int internal) const {
    if (kind[i] != 2 && kind[i] != 4) boundary[i] = p[owner[i + internal]];
  }
};
// :54-59
struct FvForceToShear {
  __device__ void operator()(int i, CF mask, CF ax, CF ay, CF az, CF fx, CF fy, CF fz, WF sx, WF sy,
                             WF sz, int internal) const {
    const int f = i + internal;
    const float a = sqrtf(ax[f] * ax[f] + ay[f] * ay[f] + az[f] * az[f]);
    const float scale = mask[i] > 0.0f ? -1.0f / a : 0.0f;
    sx[i] = scale * fx[i];
    sy[i] = scale * fy[i];
    sz[i] = scale * fz[i];
  }
};
// :70-84
struct FvPoorCellPressureFaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CI kind, CI poor, CF px, CF py, CF pz,
                             CF nx, CF ny, CF nz, CF p, CF gx, CF gy, CF gz, WF face, WF boundary,
                             int internal) const {
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
};
// :93-97
struct FvPoorPressureInterfaces {
  __device__ void operator()(int i, CI owner, CI neighbour, CI poor, CF p, WF face) const {
    const int c = owner[i], nb = neighbour[i];
    if (poor[c] != poor[nb]) face[i] = 0.5f * (p[c] + p[nb]);
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
  if (d->packet_start) {
    hipExtLaunchKernelGGL(k_each<Body, P...>, dim3(blocks_for(n)), dim3(kBlock), 0, d->stream,
                          d->packet_start, d->packet_stop, 0, n, from_arg<P>(a, K)...);
  } else {
    k_each<Body, P...><<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, from_arg<P>(a, K)...);
  }
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

// Exact FP64 distance from each cell centroid to the nearest wall triangle
// (fv/device_mesh.cpp:18-157). The reference searches a BVH that skips a node
// when the distance to its bounding box exceeds the best distance so far; the
// minimum it finds is the minimum over every triangle. Here each thread owns
// one cell. The triangles come in groups of 32 with a bounding box per group
// (the same box test as the BVH's nodes). A thread first evaluates the group
// whose box is nearest, to start from a short distance. The workgroup then
// streams the triangles through LDS in tiles of four groups, skipping a tile
// no thread needs and a group that no thread of a wavefront needs. The
// arithmetic is the reference's, operation for operation, without
// contraction.
struct WallPoint {
  double x, y, z;
};
__device__ inline WallPoint wp_sub(WallPoint a, WallPoint b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline WallPoint wp_add(WallPoint a, WallPoint b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline WallPoint wp_scale(WallPoint a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline double wp_dot(WallPoint a, WallPoint b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline WallPoint wp_cross(WallPoint a, WallPoint b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline double wp_segment2(WallPoint p, WallPoint a, WallPoint b) {
  const WallPoint ab = wp_sub(b, a);
  const double length2 = wp_dot(ab, ab);
  double t = 0;
  if (length2 != 0) {
    t = wp_dot(wp_sub(p, a), ab) / length2;
    t = t < 0.0 ? 0.0 : (1.0 < t ? 1.0 : t);
  }
  const WallPoint r = wp_sub(p, wp_add(a, wp_scale(ab, t)));
  return wp_dot(r, r);
}
__device__ inline double wp_triangle2(WallPoint p, WallPoint a, WallPoint b, WallPoint c) {
  const WallPoint ab = wp_sub(b, a), ac = wp_sub(c, a), ap = wp_sub(p, a);
  const WallPoint nrm = wp_cross(ab, ac);
  const double nn = wp_dot(nrm, nrm);
  if (nn > 0) {
    const double h = wp_dot(ap, nrm);
    const WallPoint q = wp_sub(p, wp_scale(nrm, h / nn));
    if (wp_dot(wp_cross(ab, wp_sub(q, a)), nrm) >= 0 && wp_dot(wp_cross(wp_sub(c, b), wp_sub(q, b)), nrm) >= 0 &&
        wp_dot(wp_cross(wp_sub(a, c), wp_sub(q, c)), nrm) >= 0)
      return h * h / nn;
  }
  double best = wp_segment2(p, a, b);
  const double e1 = wp_segment2(p, b, c), e2 = wp_segment2(p, c, a);
  if (e1 < best) best = e1;
  if (e2 < best) best = e2;
  return best;
}
__device__ inline double wp_box2(WallPoint p, const double* b) {
  const double lx = b[0] - p.x, ly = b[1] - p.y, lz = b[2] - p.z;
  const double hx = p.x - b[3], hy = p.y - b[4], hz = p.z - b[5];
  const WallPoint d{fmax(fmax(lx, 0.0), hx), fmax(fmax(ly, 0.0), hy), fmax(fmax(lz, 0.0), hz)};
  return wp_dot(d, d);
}
__device__ inline double wp_at(const double* t, WallPoint p) {
  return wp_triangle2(p, {t[0], t[1], t[2]}, {t[3], t[4], t[5]}, {t[6], t[7], t[8]});
}
constexpr int kWallGroup = 32;             // triangles per bounding box
constexpr int kWallTile = 4 * kWallGroup;  // triangles per LDS tile (9 KiB)
__global__ void __launch_bounds__(kBlock) k_wall_distance(int n, const double* __restrict__ centroid,
                                                          const double* __restrict__ triangle,
                                                          const double* __restrict__ bounds, int count,
                                                          float* __restrict__ out) {
  __shared__ double tile[kWallTile * 9];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool live = i < n;
  const int groups = (count + kWallGroup - 1) / kWallGroup;
  WallPoint p{0, 0, 0};
  if (live) p = {centroid[3 * std::size_t(i)], centroid[3 * std::size_t(i) + 1], centroid[3 * std::size_t(i) + 2]};
  double best = __builtin_inf();
  if (live && groups > 0) {
    int nearest = 0;
    double box = wp_box2(p, bounds);
    for (int g = 1; g < groups; ++g) {
      const double d2 = wp_box2(p, bounds + 6 * std::size_t(g));
      if (d2 < box) box = d2, nearest = g;
    }
    const int end = min(count, (nearest + 1) * kWallGroup);
    for (int j = nearest * kWallGroup; j < end; ++j) {
      const double d2 = wp_at(triangle + 9 * std::size_t(j), p);
      if (d2 < best) best = d2;
    }
  }
  for (int base = 0; base < count; base += kWallTile) {
    const int m = min(kWallTile, count - base);
    const int g0 = base / kWallGroup, gm = (m + kWallGroup - 1) / kWallGroup;
    int need = 0;
    for (int g = 0; g < gm && live && !need; ++g) need = !(wp_box2(p, bounds + 6 * std::size_t(g0 + g)) > best);
    if (!__syncthreads_or(need)) continue;  // no thread of the workgroup needs this tile
    for (int k = threadIdx.x; k < 9 * m; k += kBlock) tile[k] = triangle[9 * std::size_t(base) + k];
    __syncthreads();
    for (int g = 0; g < gm; ++g) {
      const bool want = live && !(wp_box2(p, bounds + 6 * std::size_t(g0 + g)) > best);
      if (!__any(want)) continue;
      if (want) {
        const int end = min(m, (g + 1) * kWallGroup);
        for (int j = g * kWallGroup; j < end; ++j) {
          const double d2 = wp_at(tile + 9 * j, p);
          if (d2 < best) best = d2;
        }
      }
    }
    __syncthreads();  // the tile is read before the next one is loaded
  }
  if (live) out[i] = float(sqrt(best));
}
// n = cells; centroids, triangles, group bounds, triangle count, out
void l_wall_distance(kh_device* d, int n, const kh_arg* a) {
  const int count = i32(a, 3);
  if (count < 0) throw std::runtime_error("wall_distance: negative triangle count");
  k_wall_distance<<<GRID(n)>>>(n, static_cast<const double*>(buf(a, 0)), static_cast<const double*>(buf(a, 1)),
                               static_cast<const double*>(buf(a, 2)), count, static_cast<float*>(buf(a, 4)));
}

void l_fv_pressure_assemble(kh_device* d, int n, const kh_arg* a) {
  k_fv_pressure_assemble<<<GRID(n)>>>(n, KB(0), KB(1), KCF(2), KB(3), KB(4), KB(5), KB(6), KCF(7),
                                      KCF(8), KF(9), KF(10), i32(a, 11));
}

}  // namespace

const std::vector<KernelEntry>& fv_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_FV_PRESSURE_ASSEMBLE, "fv_pressure_assemble", "bbbbbbbbbbbi", l_fv_pressure_assemble},
      {KH_WALL_DISTANCE, "wall_distance", "bbbib", l_wall_distance},
#define FV(id, name, Body) {id, name, signature_of<Body>(), launch_body<Body>, true}
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
      FV(KH_FV_DIFFUSION_CORRECTION_BOUND, "fv_diffusion_correction_bound", FvDiffusionCorrectionBound),
      FV(KH_FV_TRANSPORT_FACES, "fv_transport_faces", FvTransportFaces),
      FV(KH_FV_TRANSPORT_ASSEMBLE, "fv_transport_assemble", FvTransportAssemble),
      FV(KH_FV_FIXED_ROWS, "fv_fixed_rows", FvFixedRows),
      FV(KH_FV_FLOOR, "fv_floor", FvFloor),
      FV(KH_FV_SCALAR_STEADY_MEASURE, "fv_scalar_steady_measure", FvScalarSteadyMeasure),
      FV(KH_FV_SCALAR_SET_FIXED, "fv_scalar_set_fixed", FvScalarSetFixed),
      FV(KH_FV_SCALAR_MONITOR, "fv_scalar_monitor", FvScalarMonitor),
      FV(KH_FV_COMPENSATED_PRESSURE_JUMP, "fv_compensated_pressure_jump", FvCompensatedPressureJump),
      FV(KH_FV_PAIRED_PRESSURE_JUMP, "fv_paired_pressure_jump", FvPairedPressureJump),
      FV(KH_FV_COMPENSATED_RHIE_CHOW, "fv_compensated_rhie_chow", FvCompensatedRhieChow),
      FV(KH_FV_PAIRED_PRESSURE_BOUNDARY_LOW, "fv_paired_pressure_boundary_low", FvPairedPressureBoundaryLow),
      FV(KH_FV_PAIRED_PRESSURE_FACES, "fv_paired_pressure_faces", FvPairedPressureFaces),
      FV(KH_FV_PAIRED_MOMENTUM_RESIDUAL, "fv_paired_momentum_residual", FvPairedMomentumResidual),
      FV(KH_FV_COMPENSATED_MOMENTUM_RESIDUAL, "fv_compensated_momentum_residual", FvCompensatedMomentumResidual),
      FV(KH_FV_PAIRED_ACCUMULATE, "fv_paired_accumulate", FvPairedAccumulate),
      FV(KH_FV_COMPENSATED_PRESSURE_FORCE, "fv_compensated_pressure_force", FvCompensatedPressureForce),
      FV(KH_FV_PAIRED_VELOCITY_BOUNDARY, "fv_paired_velocity_boundary", FvPairedVelocityBoundary),
      FV(KH_FV_PAIRED_ADVECTION, "fv_paired_advection", FvPairedAdvection),
      FV(KH_FV_PAIRED_FACE_RESIDUAL, "fv_paired_face_residual", FvPairedFaceResidual),
      FV(KH_FV_PAIRED_FLUX, "fv_paired_flux", FvPairedFlux),
      FV(KH_FV_PAIRED_FLUX_OFFSET, "fv_paired_flux_offset", FvPairedFluxOffset),
      FV(KH_FV_PAIRED_CONTINUITY, "fv_paired_continuity", FvPairedContinuity),
      FV(KH_FV_PAIRED_PRESSURE_RHS, "fv_paired_pressure_rhs", FvPairedPressureRhs),
      FV(KH_FV_PAIRED_CORRECT_FLUX, "fv_paired_correct_flux", FvPairedCorrectFlux),
      FV(KH_FV_PAIRED_CORRECT_VELOCITY, "fv_paired_correct_velocity", FvPairedCorrectVelocity),
      FV(KH_FV_PAIRED_VALUE, "fv_paired_value", FvPairedValue),
      FV(KH_FV_PAIRED_VECTOR_VALUES, "fv_paired_vector_values", FvPairedVectorValues),
      FV(KH_FV_COMPENSATED_CORRECT_FLUX, "fv_compensated_correct_flux", FvCompensatedCorrectFlux),
      FV(KH_FV_COMPENSATED_CORRECT_VELOCITY, "fv_compensated_correct_velocity", FvCompensatedCorrectVelocity),
      FV(KH_FV_BACKFLOW_PRESSURE, "fv_backflow_pressure", FvBackflowPressure),
      FV(KH_FV_EFFECTIVE_VISCOSITY, "fv_effective_viscosity", FvEffectiveViscosity),
      FV(KH_FV_TRANSPOSE_STRESS, "fv_transpose_stress", FvTransposeStress),
      FV(KH_FV_STRESS_SOURCE, "fv_stress_source", FvStressSource),
      FV(KH_FV_BOUNDARY_MOMENTUM, "fv_boundary_momentum", FvBoundaryMomentum),
      FV(KH_FV_MOMENTUM_DIAGONAL, "fv_momentum_diagonal", FvMomentumDiagonal),
      FV(KH_FV_SST_FORCE, "fv_sst_force", FvSstForce),
      FV(KH_FV_TURBULENCE_BOUNDARY, "fv_turbulence_boundary", FvTurbulenceBoundary),
      FV(KH_SST_RATES, "sst_rates", SstRates),
      FV(KH_SST_BLENDING, "sst_blending", SstBlending),
      FV(KH_SST_VISCOSITY, "sst_viscosity", SstViscosity),
      FV(KH_SST_BLENDING_VISCOSITY, "sst_blending_viscosity", SstBlendingViscosity),
      FV(KH_SST_WALL, "sst_wall", SstWall),
      FV(KH_SST_WALL_OMEGA, "sst_wall_omega", SstWallOmega),
      FV(KH_SST_WALL_CELLS, "sst_wall_cells", SstWallCells),
      FV(KH_SST_SOURCES, "sst_sources", SstSources),
      FV(KH_SST_FACE_DIFFUSION, "sst_face_diffusion", SstFaceDiffusion),
      FV(KH_SST_SUSTAIN, "sst_sustain", SstSustain),
      FV(KH_FV_SIMPLE_INVERSE, "fv_simple_inverse", FvSimpleInverse),
      FV(KH_FV_STANDARD_PRESSURE_FACES, "fv_standard_pressure_faces", FvStandardPressureFaces),
      FV(KH_FV_CELL_PRESSURE_BOUNDARY, "fv_cell_pressure_boundary", FvCellPressureBoundary),
      FV(KH_FV_FORCE_TO_SHEAR, "fv_force_to_shear", FvForceToShear),
      FV(KH_FV_POOR_CELL_PRESSURE_FACES, "fv_poor_cell_pressure_faces", FvPoorCellPressureFaces),
      FV(KH_FV_POOR_PRESSURE_INTERFACES, "fv_poor_pressure_interfaces", FvPoorPressureInterfaces),
#undef FV
  };
  return table;
}

}  // namespace kh
