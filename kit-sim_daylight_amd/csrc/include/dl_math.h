// SPDX-License-Identifier: Apache-2.0 AND LicenseRef-Commons-Clause-1.0
// Radiance notice: the Perez tables, tcos and glass formulas here follow Radiance 5.4 (gendaymtx.c, tcos.c, m_glass.c),
// Copyright (c) 1990 to 2022 The Regents of the University of California, through LBNL; Radiance Software License 2.0,
// reproduced with its conditions and disclaimer in NOTICES.md.
// The per-ray and per-step arithmetic of the trace-and-accumulate path, as __host__ __device__ functions: the HIP kernels
// run them, and host code that must agree with a kernel bit for bit (the scene's padding, the sun table) calls the same
// functions. Everything here is compiled without floating-point contraction (-ffp-contract=off); fma is written out where
// it is meant.
//
// Ray rules follow reference src/trace_dev.cu and src/params.h (cited per function). The triangle test is the watertight
// test of Woop, Benthin & Wald (JCGT 2(1), 2013); CUDA's fast rsqrtf / sincospif become 1 / sqrtf and a Taylor sincospi.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DL_HD __host__ __device__ __forceinline__

namespace dlm {

constexpr double kPi = 3.14159265358979323846;

DL_HD float as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
DL_HD uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }
DL_HD float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // rounded per operation

// offset_ray (trace_dev.cu:16-25): 256 ulps along n per axis, or 2^-16 n within 1/32 m of the origin.
DL_HD float offset_axis(float p, float n) {
  const int o = (int)(256.f * n);
  const float stepped = as_float((uint32_t)((int32_t)as_uint(p) + (p < 0.f ? -o : o)));
  return fabsf(p) < (1.f / 32.f) ? p + (1.f / 65536.f) * n : stepped;
}
DL_HD float3 offset_ray(float3 p, float3 n) { return make_float3(offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z)); }

// pcg3d (Jarzynski & Olano 2020, JCGT 9(3)) and its [0, 1) float (params.h:64-71)
DL_HD uint3 pcg3d(uint32_t x, uint32_t y, uint32_t z) {
  x = x * 1664525u + 1013904223u; y = y * 1664525u + 1013904223u; z = z * 1664525u + 1013904223u;
  x += y * z; y += z * x; z += x * y;
  x ^= x >> 16; y ^= y >> 16; z ^= z >> 16;
  x += y * z; y += z * x; z += x * y;
  return make_uint3(x, y, z);
}
DL_HD float u01(uint32_t v) { return (float)(v >> 8) * (1.f / 16777216.f); }

// sin(pi x), cos(pi x): quarter-turn reduction (exact) and Taylor polynomials of degree 9 / 10 by fma
DL_HD void sincospi(float x, float* s, float* c) {
  const float q = rintf(x * 2.f);
  const float r = x - q * 0.5f;
  const int k = (int)((long long)q & 3);
  const float r2 = r * r;
  float ps = 8.21458866e-2f;
  ps = fmaf(ps, r2, -5.99264528e-1f);
  ps = fmaf(ps, r2, 2.55016404e+0f);
  ps = fmaf(ps, r2, -5.16771278e+0f);
  ps = fmaf(ps, r2, 3.14159274e+0f);
  const float sr = ps * r;
  float pc = -2.58068914e-2f;
  pc = fmaf(pc, r2, 2.35330630e-1f);
  pc = fmaf(pc, r2, -1.33526277e+0f);
  pc = fmaf(pc, r2, 4.05871213e+0f);
  pc = fmaf(pc, r2, -4.93480220e+0f);
  const float cr = fmaf(pc, r2, 1.f);
  if (k == 0) { *s = sr; *c = cr; }
  else if (k == 1) { *s = cr; *c = -sr; }
  else if (k == 2) { *s = -sr; *c = -cr; }
  else { *s = -cr; *c = sr; }
}

// view_cell (params.h:77-85): cell d (column col) of a field, jittered by (element, d, side) or its centre; field frame
// x right, y forward, z up. rsLo / rsHi: the row's sines of elevation.
DL_HD float3 view_cell(uint32_t d, uint32_t col, float rsLo, float rsHi, float az0, float dAz, bool jitter, uint32_t element,
                        uint32_t side) {
  float a = 0.5f, s = 0.5f * (rsLo + rsHi);
  if (jitter) {
    const uint3 h = pcg3d(element, d, side);
    a = u01(h.x);
    s = rsLo + u01(h.y) * (rsHi - rsLo);
  }
  float sa, ca;
  sincospi((az0 + ((float)col + a) * dAz) * (1.f / 180.f), &sa, &ca);
  const float cc = sqrtf(fmaxf(0.f, 1.f - s * s));
  return make_float3(cc * sa, cc * ca, s);
}

// ------------------------------------------------------------------------------------------------ the triangle test
// Per-ray constants of the watertight test: the dominant axis kz (x > y > z on ties), kx / ky swapped when d[kz] < 0
// (winding kept), the shear S. Box test: inverse direction and the ray origin moved by the scene's padding toward the
// near / far plane of each slab.
struct Ray {
  float3 o, d;
  int kx, ky, kz;
  float Sx, Sy, Sz;
  float3 inv, oNear, oFar;
};
DL_HD float comp(float3 v, int k) { return k == 0 ? v.x : k == 1 ? v.y : v.z; }
// the origin's part of the setup: the slab origins moved by the box padding towards the near / far planes
DL_HD void ray_origin(Ray& r, float3 o, float pad) {
  r.o = o;
  r.oNear = make_float3(r.inv.x >= 0.f ? o.x + pad : o.x - pad, r.inv.y >= 0.f ? o.y + pad : o.y - pad, r.inv.z >= 0.f ? o.z + pad : o.z - pad);
  r.oFar = make_float3(r.inv.x >= 0.f ? o.x - pad : o.x + pad, r.inv.y >= 0.f ? o.y - pad : o.y + pad, r.inv.z >= 0.f ? o.z - pad : o.z + pad);
}
// the direction's part of the setup (shared by every ray along d)
DL_HD void ray_dir(Ray& r, float3 d) {
  r.d = d;
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  r.kz = ax >= ay ? (ax >= az ? 0 : 2) : (ay >= az ? 1 : 2);
  r.kx = r.kz == 2 ? 0 : r.kz + 1;
  r.ky = r.kx == 2 ? 0 : r.kx + 1;
  const float dz = comp(d, r.kz);
  if (dz < 0.f) { const int t = r.kx; r.kx = r.ky; r.ky = t; }
  r.Sx = comp(d, r.kx) / dz;
  r.Sy = comp(d, r.ky) / dz;
  r.Sz = 1.f / dz;
  const float dx = d.x != 0.f ? d.x : ((as_uint(d.x) >> 31) ? -1e-30f : 1e-30f);
  const float dy = d.y != 0.f ? d.y : ((as_uint(d.y) >> 31) ? -1e-30f : 1e-30f);
  const float dzz = d.z != 0.f ? d.z : ((as_uint(d.z) >> 31) ? -1e-30f : 1e-30f);
  r.inv = make_float3(1.f / dx, 1.f / dy, 1.f / dzz);
}
DL_HD void ray_setup(Ray& r, float3 o, float3 d, float pad) {
  ray_dir(r, d);
  ray_origin(r, o, pad);
}
// A direction's part as two float4 (a table built once per direction set): inv xyz, Sx | Sy, Sz, axes kx | ky << 2 |
// kz << 4 as bits, 0. dir_unpack gives back exactly what ray_dir computed.
DL_HD void dir_pack(const Ray& r, float4& a, float4& b) {
  a = make_float4(r.inv.x, r.inv.y, r.inv.z, r.Sx);
  b = make_float4(r.Sy, r.Sz, as_float(uint32_t(r.kx | r.ky << 2 | r.kz << 4)), 0.f);
}
DL_HD void dir_unpack(Ray& r, float3 d, float4 a, float4 b) {
  r.d = d;
  r.inv = make_float3(a.x, a.y, a.z);
  r.Sx = a.w; r.Sy = b.x; r.Sz = b.y;
  const uint32_t k = as_uint(b.z);
  r.kx = int(k & 3u); r.ky = int((k >> 2) & 3u); r.kz = int((k >> 4) & 3u);
}

// Hit on [tmin, tmax]: t and front (det > 0: the ray meets the side the winding's normal faces). cull drops back faces.
// bc (optional): the hit's barycentric weights of v1 and v2 (V / det, W / det).
DL_HD bool tri_hit(const Ray& r, float3 v0, float3 v1, float3 v2, float tmin, float tmax, bool cull, float& tOut, bool& front,
                   float* bc = nullptr) {
  const float3 A = make_float3(v0.x - r.o.x, v0.y - r.o.y, v0.z - r.o.z);
  const float3 B = make_float3(v1.x - r.o.x, v1.y - r.o.y, v1.z - r.o.z);
  const float3 C = make_float3(v2.x - r.o.x, v2.y - r.o.y, v2.z - r.o.z);
  const float Az0 = comp(A, r.kz), Bz0 = comp(B, r.kz), Cz0 = comp(C, r.kz);
  const float Ax = comp(A, r.kx) - r.Sx * Az0, Ay = comp(A, r.ky) - r.Sy * Az0;
  const float Bx = comp(B, r.kx) - r.Sx * Bz0, By = comp(B, r.ky) - r.Sy * Bz0;
  const float Cx = comp(C, r.kx) - r.Sx * Cz0, Cy = comp(C, r.ky) - r.Sy * Cz0;
  float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
  bool edge = false;
  if (U == 0.f || V == 0.f || W == 0.f) {   // on an edge in float: decide with exact products in double
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    edge = U == 0.f || V == 0.f || W == 0.f;
  }
  if ((U < 0.f || V < 0.f || W < 0.f) && (U > 0.f || V > 0.f || W > 0.f)) return false;
  const float det = U + V + W;
  if (det == 0.f) return false;
  if (cull && det < 0.f) return false;
  if (edge) {
    // the ray meets an edge (or a vertex): a top-left rule gives it to one triangle of those that share it, as a
    // hardware triangle test does (each edge's direction in the winding, turned by the facing: two triangles on either
    // side of an edge see it in opposite directions)
    const float sg = det > 0.f ? 1.f : -1.f;
    const auto mine = [sg](float ex, float ey) { ex *= sg; ey *= sg; return ey > 0.f || (ey == 0.f && ex > 0.f); };
    if ((U == 0.f && !mine(Cx - Bx, Cy - By)) || (V == 0.f && !mine(Ax - Cx, Ay - Cy)) || (W == 0.f && !mine(Bx - Ax, By - Ay))) return false;
  }
  const float T = U * (r.Sz * Az0) + V * (r.Sz * Bz0) + W * (r.Sz * Cz0);
  const float t = T / det;
  if (!(t >= tmin && t <= tmax)) return false;
  tOut = t;
  front = det > 0.f;
  if (bc) { bc[0] = V / det; bc[1] = W / det; }
  return true;
}

// own_plane (trace_dev.cu:63-66): the hit's triangle is not degenerate, its plane passes within 2^-22 x its extent of
// the origin, and it faces away from n (the sensor's side)
DL_HD bool own_plane(float3 d, float3 n, float t, bool front, float3 v0, float3 v1, float3 v2) {
  const float3 e1 = make_float3(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z), e2 = make_float3(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z);
  const float3 ng = make_float3(e1.y * e2.z - e1.z * e2.y, e1.z * e2.x - e1.x * e2.z, e1.x * e2.y - e1.y * e2.x);
  const float l2 = dot3(ng, ng);
  if (!(l2 > 0.f)) return false;
  const float s = front ? 1.f / sqrtf(l2) : -(1.f / sqrtf(l2));
  const float3 hn = make_float3(s * ng.x, s * ng.y, s * ng.z);
  const float ext = fmaxf(fmaxf(fmaxf(fabsf(e1.x), fabsf(e1.y)), fmaxf(fabsf(e1.z), fabsf(e2.x))), fmaxf(fabsf(e2.y), fabsf(e2.z)));
  return t * fabsf(dot3(d, hn)) < ext * 2.3841858e-7f && dot3(n, hn) <= 0.f;
}
// own_hit (trace_dev.cu:60-72): the hit lies under 1/32 m above the sensor's plane (the gate), and own_plane.
DL_HD bool own_hit(float3 d, float3 n, float t, bool front, float3 v0, float3 v1, float3 v2) {
  return t * dot3(n, d) < 1.f / 32.f && own_plane(d, n, t, front, v0, v1, v2);
}

// ------------------------------------------------------------------------------------------------ transmission
// Slot flags (reference src/params.h:14-19; include/dl_device.h DLG_SLOT_*)
constexpr uint32_t kSlotGlass = 1, kSlotThin = 2, kSlotInward = 4, kSlotTwin = 8;

// A transmittance byte back as the fraction it stands for: b / 255.f, correctly rounded (the quotient in double rounds to
// the same float: b / 255 repeats an 8-bit pattern, so it never lies within 2^-53 of a float midpoint).
DL_HD float byte_T(uint32_t b) { return float(double(b) / 255.0); }
// The byte a transmittance is stored as (trace_dev.cu:141): round to nearest of T x 255 in float.
DL_HD uint32_t T_byte(float T) { return uint32_t(rintf(T * 255.f)); }

// Glass transmissivity tn from the transmittance at normal incidence Tn (Radiance rad5R4 src/rt/m_glass.c, the form of
// reference src/glass.h:11-13), at least 0.
DL_HD double tn_from_Tn(double Tn) {
  return fmax(0., (sqrt(.8402528435 + .0072522239 * Tn * Tn) - .9166530661) / .0036261119 / Tn);
}

// Transmittance T of a glass pane at incidence cosine c (0..1) for transmissivity tn (reference src/glass.h:17-25, after
// m_glass.c): refractive index 1.52, the cosine inside c2 by Snell's law, the path's transmissivity tn^(1 / c2), the
// Fresnel reflectances of the s and p polarisations at each face and the pane's internal reflections summed,
// (1 - r)^2 d / (1 - r^2 d^2), averaged over the two. In double; the reference's float formula agrees to a few float ulps.
// The reflectance R likewise: the first face's Fresnel reflection plus what the pane's internal reflections send back,
// rs (1 + (1 - 2 rs) d^2) / (1 - rs^2 d^2) per polarisation, averaged.
DL_HD void glass_TR(double c, double tn, float& T, float& R) {
  const double n = 1.52;
  const double c2 = sqrt((1. - 1. / (n * n)) + c * c / (n * n));
  const double d = pow(tn, 1. / c2), d2 = d * d;
  double rs = (c - n * c2) / (c + n * c2), rp = (c2 - n * c) / (c2 + n * c);
  rs *= rs;
  rp *= rp;
  T = float(.5 * (1. - rs) * (1. - rs) * d / (1. - rs * rs * d2) + .5 * (1. - rp) * (1. - rp) * d / (1. - rp * rp * d2));
  R = float(.5 * rs * (1. + (1. - 2. * rs) * d2) / (1. - rs * rs * d2) + .5 * rp * (1. + (1. - 2. * rp) * d2) / (1. - rp * rp * d2));
}
DL_HD float glass_T(double c, double tn) {
  float T, R;
  glass_TR(c, tn, T, R);
  return T;
}

// What a translucent slot's triangle passes (reference __anyhit__tr, trace_dev.cu:149-165): 1 on a thin solid's exit
// face (the back face, or the front face when it is wound inward), Radiance glass at the hit's incidence (the geometric
// normal of v0 v1 v2 in double from the float edges, against the unit direction d), else tau.
DL_HD float pass_factor(float tau, float tn, uint32_t flags, bool front, float3 v0, float3 v1, float3 v2, float3 d) {
  if ((flags & kSlotThin) && front == ((flags & kSlotInward) != 0)) return 1.f;
  if (!(flags & kSlotGlass)) return tau;
  const double ax = v1.x - v0.x, ay = v1.y - v0.y, az = v1.z - v0.z;
  const double bx = v2.x - v0.x, by = v2.y - v0.y, bz = v2.z - v0.z;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double l = sqrt(nx * nx + ny * ny + nz * nz);
  if (!(l > 0.)) return 0.f;
  return glass_T(fabs(nx * d.x + ny * d.y + nz * d.z) / l, tn);
}

// ------------------------------------------------------------------------------------------------ reflected light
// 32-bit reversal (the Hammersley point's radical inverse)
DL_HD uint32_t brev32(uint32_t x) {
  x = (x >> 16) | (x << 16);
  x = ((x & 0xff00ff00u) >> 8) | ((x & 0x00ff00ffu) << 8);
  x = ((x & 0xf0f0f0f0u) >> 4) | ((x & 0x0f0f0f0fu) << 4);
  x = ((x & 0xccccccccu) >> 2) | ((x & 0x33333333u) << 2);
  return ((x & 0xaaaaaaaau) >> 1) | ((x & 0x55555555u) << 1);
}
// A cosine-weighted direction about unit n (trace_dev.cu:235-242): sin^2 theta = u, azimuth 2 pi v, in the orthonormal
// basis of Duff et al. 2017 (JCGT 6(1), "Building an orthonormal basis, revisited").
DL_HD float3 cosine_dir(float3 n, float u, float v) {
  const float sg = copysignf(1.f, n.z), a = -1.f / (sg + n.z), b = n.x * n.y * a;
  const float3 t = make_float3(1.f + sg * n.x * n.x * a, sg * b, -sg * n.x), w = make_float3(b, sg + n.y * n.y * a, -n.y);
  float s, c;
  sincospi(2.f * v, &s, &c);
  const float r = sqrtf(u), z = sqrtf(fmaxf(0.f, 1.f - u)), rc = r * c, rs = r * s;
  return make_float3((rc * t.x + rs * w.x) + z * n.x, (rc * t.y + rs * w.y) + z * n.y, (rc * t.z + rs * w.z) + z * n.z);
}
// The bin of an escaping direction (trace_dev.cu:244-258, Radiance reinhart.cal's rbin): 0 below the horizon (the
// ground), else 1 + its Reinhart patch of sky density mf in gendaymtx order, the direction turned back to true north
// (cosN, sinN of --north); alpha the patch rows' height (radians); the zenith cap is the last bin. The angles in double.
DL_HD uint32_t sky_bin(float3 d, uint32_t mf, float alpha, uint32_t nBins, float cosN, float sinN) {
  if (!(d.z >= 0.f)) return 0;
  const uint32_t tnaz[7] = {30, 30, 24, 24, 18, 12, 6};   // Tregenza patches per row (x mf rows of x mf patches)
  const double el = asin(double(fminf(d.z, 1.f)));
  uint32_t r = uint32_t(el / double(alpha));
  if (r >= 7 * mf) return nBins - 1;
  const uint32_t band = r / mf, n = tnaz[band] * mf;
  uint32_t b = 1 + (r % mf) * n;
  for (uint32_t q = 0; q < band; ++q) b += tnaz[q] * mf * mf;
  double az = atan2(double(d.x * cosN + d.y * sinN), double(d.y * cosN - d.x * sinN));   // clockwise from north
  if (az < 0.) az += 2. * kPi;
  return b + uint32_t(az * (double(n) / (2. * kPi)) + .5) % n;
}

// ------------------------------------------------------------------------------------------------ the sun (FP64)
DL_HD int month_days(int m, bool leap) {
  return m == 2 ? (leap ? 29 : 28) : (m == 4 || m == 6 || m == 9 || m == 11) ? 30 : 31;
}
DL_HD double pos_mod(double a, double m) { const double r = fmod(a, m); return r < 0 ? r + m : r; }
// x >= 0 to two decimals, exactly as printing "%.2f" and reading it back: the exact binary value rounded half to even.
DL_HD double round_2dp(double x) {
  const double p = x * 100.0, e = fma(x, 100.0, -p);   // x * 100 = p + e exactly
  const double n = floor(p), d = (p - n) - 0.5;          // both exact
  double k = n;
  if (d > -e) k = n + 1;
  else if (d == -e) k = fmod(n, 2.0) == 0.0 ? n : n + 1;
  return k / 100.0;
}
// NOAA / Meeus solar position with the Ladybug Sunpath conventions (reference src/solar.cpp:66-105): degrees, azimuth
// clockwise from north, up = refracted altitude >= 0; the vector towards the sun (true north).
struct SunPos { double alt, azi; bool up; float3 vec; };
DL_HD double rad_s(double d) { return d * kPi / 180; }   // solar.cpp's order: (d pi) / 180
DL_HD double deg_s(double r) { return r * 180 / kPi; }
DL_HD float3 altaz_vec(double alt, double azi) {           // sky.cpp rad_vector: cos alt is reused
  const double ca = cos(alt);
  return make_float3((float)(ca * sin(azi)), (float)(ca * cos(azi)), (float)sin(alt));
}
DL_HD SunPos sun_noaa(double lat, double lon, double tz, int month, int day, double hourFrac, bool leap) {
  int hour = (int)hourFrac, minute = (int)llround((hourFrac - hour) * 60);
  if (minute == 60) { ++hour; minute = 0; }
  int doy = day;
  for (int m = 1; m < month; ++m) doy += month_days(m, leap);
  const int y = leap ? 2015 : 2016;                         // the year before 2016 / 2017, Excel serial from 1900
  const int base = ((leap ? 2016 : 2017) - 1900) * 365 + (y / 4 - y / 100 + y / 400) - (1899 / 4 - 1899 / 100 + 1899 / 400) + doy + 1;
  const double jd = base + 2415018.5 + round_2dp((minute + hour * 60) / 1440.0) - tz / 24;
  const double jc = (jd - 2451545) / 36525;
  const double L0 = pos_mod(280.46646 + jc * (36000.76983 + jc * 0.0003032), 360);
  const double M = 357.52911 + jc * (35999.05029 - 0.0001537 * jc);
  const double ecc = 0.016708634 - jc * (0.000042037 + 0.0000001267 * jc);
  const double C = sin(rad_s(M)) * (1.914602 - jc * (0.004817 + 0.000014 * jc)) + sin(rad_s(2 * M)) * (0.019993 - 0.000101 * jc) +
                   sin(rad_s(3 * M)) * 0.000289;
  const double om = rad_s(125.04 - 1934.136 * jc);
  const double lam = rad_s(L0 + C - 0.00569 - 0.00478 * sin(om));
  const double eps0 = 23 + (26 + (21.448 - jc * (46.815 + jc * (0.00059 - jc * 0.001813))) / 60) / 60;
  const double eps = rad_s(eps0 + 0.00256 * cos(om));
  const double dec = asin(sin(eps) * sin(lam));
  double yy = tan(eps / 2);
  yy *= yy;
  const double eot = 4 * deg_s(yy * sin(2 * rad_s(L0)) - 2 * ecc * sin(rad_s(M)) + 4 * ecc * yy * sin(rad_s(M)) * cos(2 * rad_s(L0)) -
                               0.5 * yy * yy * sin(4 * rad_s(L0)) - 1.25 * ecc * ecc * sin(2 * rad_s(M)));
  const double tst = pos_mod(hourFrac * 60 + eot + 4 * lon - 60 * tz, 1440);
  const double H = tst < 0 ? tst / 4 + 180 : tst / 4 - 180;
  const double la = rad_s(lat);
  const double zen = acos(sin(la) * sin(dec) + cos(la) * cos(dec) * cos(rad_s(H)));
  const double alt = 90 - deg_s(zen), t = tan(rad_s(alt));
  double r;
  if (alt > 85) r = 0;
  else if (alt > 5) r = 58.1 / t - 0.07 / (t * t * t) + 0.000086 / (t * t * t * t * t);
  else if (alt > -0.575) r = 1735 + alt * (-518.2 + alt * (103.4 + alt * (-12.79 + alt * 0.711)));
  else r = -20.772 / t;
  SunPos s;
  s.alt = alt + r / 3600;
  const double a = (sin(la) * cos(zen) - sin(dec)) / (cos(la) * sin(zen));
  if (a >= -1 && a <= 1) s.azi = pos_mod(H > 0 ? deg_s(acos(a)) + 180 : 540 - deg_s(acos(a)), 360);
  else s.azi = a > 0 ? 180 : 0;
  s.up = s.alt >= 0;
  s.vec = altaz_vec(rad_s(s.alt), rad_s(s.azi));
  return s;
}
// counterclockwise about +z by deg (sky.cpp rotate_z; degrees to radians as d (pi / 180))
DL_HD float3 rotate_z(float3 v, double deg) {
  const double t = deg * (kPi / 180.), c = cos(t), s = sin(t);
  return make_float3((float)(v.x * c - v.y * s), (float)(v.x * s + v.y * c), v.z);
}

// ------------------------------------------------------------------------------------------------ the Perez sky (FP64)
// The published coefficient tables (Perez, Seals & Michalsky 1993, Table 1; Perez et al. 1990, Table 4) and the
// gendaymtx clamps, as reference src/sky.cpp:297-474 ports them from Radiance.
struct PerezStep {
  int cat;            // clearness category 0..7
  double zen;         // sun zenith, clamped to [3, 90] degrees (radians)
  double c[5];        // a .. e
  double difRad, dirRad, difVis, dirVis;   // diffuse / direct illuminance, -O1 (x 179) and -O0 (efficacy)
  float gRad, gVis;   // ground row value before the ground reflectance (float, as gendaymtx's array)
  bool night;
};
DL_HD PerezStep perez_step(int jd, double alt, double dni, double dhi) {
  PerezStep q{};
  q.night = dni + dhi <= 1e-4;
  if (q.night) return q;
  const double pwc = exp(0.07 * 11.0 - 0.075);
  double zen;
  if (alt <= 0.0) zen = 90.0 * (kPi / 180.);
  else if (alt >= 87.0 * (kPi / 180.)) zen = 3.0 * (kPi / 180.);
  else zen = 90.0 * (kPi / 180.) - alt;
  q.zen = zen;
  const double da = (jd - 1.0) * (2.0 * kPi / 365.0);
  const double ecc = 1.00011 + 0.034221 * cos(da) + 0.00128 * sin(da) + 0.000719 * cos(2.0 * da) + 0.000077 * sin(2.0 * da);
  const double am = 1.0 / (cos(zen) + 0.15 * pow(93.885 - zen * (180. / kPi), -1.253));
  double bright = dhi * am / (1367.0 * ecc);
  const double z3 = zen * zen * zen;
  double clear = ((dhi + dni) / dhi + 1.041 * z3) / (1.0 + 1.041 * z3);
  if (clear > 11.9) clear = 11.9;
  if (bright < 0.01) bright = 0.01;
  const double lo[8] = {1.000, 1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200};
  const double hi[8] = {1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200, 12.01};
  int cat = 0;
  while (cat < 8 && !(clear >= lo[cat] && clear < hi[cat])) ++cat;
  if (cat > 7) cat = 7;
  q.cat = cat;
  // Perez 1990 Table 4 (diffuse, direct) a b c d
  const double eD[8][4] = {{97.24, -0.46, 12.00, -8.91},   {107.22, 1.15, 0.59, -3.95},  {104.97, 2.96, -5.53, -8.77},
                           {102.39, 5.59, -13.95, -13.90}, {100.71, 5.94, -22.75, -23.74}, {106.42, 3.83, -36.15, -28.83},
                           {141.88, 1.90, -53.24, -14.03}, {152.23, 0.35, -45.27, -7.98}};
  const double eB[8][4] = {{57.20, -4.55, -2.98, 117.12}, {98.99, -3.46, -1.21, 12.38},  {109.83, -4.90, -1.71, -8.81},
                           {110.34, -5.84, -1.99, -4.56}, {106.36, -3.97, -1.75, -6.16}, {107.19, -1.25, -1.51, -26.73},
                           {105.75, 0.77, -1.26, -34.44}, {101.18, 1.58, -1.10, -8.29}};
  q.difVis = dhi * (eD[cat][0] + eD[cat][1] * pwc + eD[cat][2] * cos(zen) + eD[cat][3] * log(bright));
  const double r = eB[cat][0] + eB[cat][1] * pwc + eB[cat][2] * exp(5.73 * zen - 5.0) + eB[cat][3] * bright;
  q.dirVis = dni * (r > 0.0 ? r : 0.0);
  q.difRad = dhi * 179.;
  q.dirRad = dni * 179.;
  float g = (float)q.difRad;
  if (alt > 0) g = (float)(g + q.dirRad * sin(alt));
  q.gRad = (float)(g * (1. / kPi / 179.));
  g = (float)q.difVis;
  if (alt > 0) g = (float)(g + q.dirVis * sin(alt));
  q.gVis = (float)(g * (1. / kPi / 179.));
  // Perez 1993 Table 1: a..e as x0 + x1 z + delta (x2 + x3 z)
  const double X[8][5][4] = {
      {{1.3525, -0.2576, -0.2690, -1.4366}, {-0.7670, 0.0007, 1.2734, -0.1233}, {2.8000, 0.6004, 1.2375, 1.0000},
       {1.8734, 0.6297, 0.9738, 0.2809}, {0.0356, -0.1246, -0.5718, 0.9938}},
      {{-1.2219, -0.7730, 1.4148, 1.1016}, {-0.2054, 0.0367, -3.9128, 0.9156}, {6.9750, 0.1774, 6.4477, -0.1239},
       {-1.5798, -0.5081, -1.7812, 0.1080}, {0.2624, 0.0672, -0.2190, -0.4285}},
      {{-1.1000, -0.2515, 0.8952, 0.0156}, {0.2782, -0.1812, -4.5000, 1.1766}, {24.7219, -13.0812, -37.7000, 34.8438},
       {-5.0000, 1.5218, 3.9229, -2.6204}, {-0.0156, 0.1597, 0.4199, -0.5562}},
      {{-0.5484, -0.6654, -0.2672, 0.7117}, {0.7234, -0.6219, -5.6812, 2.6297}, {33.3389, -18.3000, -62.2500, 52.0781},
       {-3.5000, 0.0016, 1.1477, 0.1062}, {0.4659, -0.3296, -0.0876, -0.0329}},
      {{-0.6000, -0.3566, -2.5000, 2.3250}, {0.2937, 0.0496, -5.6812, 1.8415}, {21.0000, -4.7656, -21.5906, 7.2492},
       {-3.5000, -0.1554, 1.4062, 0.3988}, {0.0032, 0.0766, -0.0656, -0.1294}},
      {{-1.0156, -0.3670, 1.0078, 1.4051}, {0.2875, -0.5328, -3.8500, 3.3750}, {14.0000, -0.9999, -7.1406, 7.5469},
       {-3.4000, -0.1078, -1.0750, 1.5702}, {-0.0672, 0.4016, 0.3017, -0.4844}},
      {{-1.0000, 0.0211, 0.5025, -0.5119}, {-0.3000, 0.1922, 0.7023, -1.6317}, {19.0000, -5.0000, 1.2438, -1.9094},
       {-4.0000, 0.0250, 0.3844, 0.2656}, {1.0468, -0.3788, -2.4517, 1.4656}},
      {{-1.0500, 0.0289, 0.4260, 0.3590}, {-0.3250, 0.1156, 0.7781, 0.0025}, {31.0625, -14.5000, -46.1148, 55.3750},
       {-7.2312, 0.4050, 13.3500, 0.6234}, {1.5000, -0.6426, 1.8564, 0.5636}}};
  double delta = bright;
  if (clear > 1.065 && clear < 2.8 && delta < 0.2) delta = 0.2;
  for (int i = 0; i < 5; ++i) q.c[i] = X[cat][i][0] + X[cat][i][1] * zen + delta * (X[cat][i][2] + X[cat][i][3] * zen);
  if (cat == 0) {
    q.c[2] = exp(pow(delta * (X[0][2][0] + X[0][2][1] * zen), X[0][2][2])) - X[0][2][3];
    q.c[3] = -exp(delta * (X[0][3][0] + X[0][3][1] * zen)) + X[0][3][2] + delta * X[0][3][3];
  }
  return q;
}
// relative luminance of the patch at (alt, azi) (float angles, as gendaymtx keeps them) under step q's sky
DL_HD float perez_patch(const PerezStep& q, float palt, float pazi, double azimuth) {
  const double dAz = fabs(pazi - azimuth);
  const double zp = kPi * 0.5 - palt;
  const double gam = acos(cos(q.zen) * cos(zp) + sin(q.zen) * sin(zp) * cos(dAz));
  float L = (float)((1.0 + q.c[0] * exp(q.c[1] / cos(zp))) * (1.0 + q.c[2] * exp(q.c[3] * gam) + q.c[4] * cos(gam) * cos(gam)));
  return L < 0 ? 0.f : L;
}
// bright() of a patch after ComputeSky's scaling: float(L sf) x skycolor per channel, then the CIE weights in double
DL_HD double sky_bright(float L, double sf) {
  const float l = (float)(L * sf);
  const float r = l * .96f, g = l * 1.004f, b = l * 1.118f;
  return 0.265074126 * r + 0.670114631 * g + 0.064811243 * b;
}
// bright() of a grey value v (r = g = b, as AddDirect adds the sun): the CIE weights summed in channel order (run.cpp:27)
DL_HD double sun_bright(float v) { return 0.265074126 * v + 0.670114631 * v + 0.064811243 * v; }
DL_HD double ground_bright(float g, float grefl) {
  const float v = g * grefl;
  return 0.265074126 * v + 0.670114631 * v + 0.064811243 * v;
}

}  // namespace dlm
