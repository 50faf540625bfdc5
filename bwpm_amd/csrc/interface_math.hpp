// FP64 interface physics shared by host set-up code and device table builders:
// longitudinal wavenumber with the grazing rule, Fresnel coefficients (Convention A), the 2×2
// transmission and reflection matrices on (E_x, E_y), and the operator table of wende2023 table 1.
//
// Semantics follow the reference (cpp/src/interface.cpp:8-52, cpp/src/solver.cpp:17-24,73-106);
// the expressions keep the reference's association order and are compiled without FMA
// contraction so that the FP64 values round like the reference's host code.
#pragma once

#include "common.hpp"

namespace bwpm {

using zd = cx<double>;

inline constexpr double kGrazingTolerance = 1e-12;  // relative, cpp/include/bwpm/interface.hpp:17

BWPM_HD double zabs(zd a) { return hypot(a.x, a.y); }

// Smith's complex division.
BWPM_HD zd zdiv(zd a, zd b) {
#pragma clang fp contract(off)
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x;
    const double den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  const double r = b.x / b.y;
  const double den = b.x * r + b.y;
  return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
BWPM_HD zd zdiv(zd a, double s) { return {a.x / s, a.y / s}; }

// Principal square root.
BWPM_HD zd zsqrt(zd z) {
#pragma clang fp contract(off)
  if (z.x == 0.0 && z.y == 0.0) return {0.0, z.y};
  const double t = sqrt(0.5 * (fabs(z.x) + hypot(z.x, z.y)));
  if (z.x >= 0.0) return {t, z.y / (2.0 * t)};
  return {fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

BWPM_HD zd zexp(zd z) {
  const double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return {e * c, e * s};
}

BWPM_HD zd zmul(zd a, zd b) {
#pragma clang fp contract(off)
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
BWPM_HD zd zmul(zd a, double s) {
#pragma clang fp contract(off)
  return {a.x * s, a.y * s};
}
BWPM_HD zd zadd(zd a, zd b) { return {a.x + b.x, a.y + b.y}; }
BWPM_HD zd zsub(zd a, zd b) { return {a.x - b.x, a.y - b.y}; }

// k = 2π·(m·(1/(n·d))) in FFT order, rounded like 2π·numpy.fft.fftfreq (solver.cpp:17-24).
inline double fftfreq(long long i, long long n, double d) {
  const long long m = i < (n + 1) / 2 ? i : i - n;
  const double val = 1.0 / (static_cast<double>(n) * d);
  return 2.0 * 3.14159265358979323846 * (static_cast<double>(m) * val);
}

// k_z = sqrt(ñ²k₀² − k⊥²), Im ≥ 0 (Re ≥ 0 if Im = 0); |ñ²k₀² − k⊥²| ≤ 1e-12·|ñ²k₀²| counts as
// exactly grazing; |k_z| < 1e-12·k₀ → i·1e-12·k₀ (interface.cpp:8-18).
BWPM_HD zd kz_of(zd n, double k0, double kp2) {
#pragma clang fp contract(off)
  const zd n2k2 = zmul(zmul(n, n), k0 * k0);
  zd q{n2k2.x - kp2, n2k2.y};
  if (zabs(q) <= kGrazingTolerance * zabs(n2k2)) q = zd{0.0, 0.0};
  zd kz = zsqrt(q);
  if (kz.y < 0.0 || (kz.y == 0.0 && kz.x < 0.0)) kz = zd{-kz.x, -kz.y};
  const double tiny = 1e-12 * k0;
  if (zabs(kz) < tiny) kz = zd{0.0, tiny};
  return kz;
}

// Propagating, non-grazing test of the Born machinery (solver.cpp:233).
BWPM_HD bool born_propagating(zd host, double k0, double kp2) {
#pragma clang fp contract(off)
  const zd h2k2 = zmul(zmul(zmul(host, host), k0), k0);
  const double kabs2 = zabs(h2k2);
  return kp2 < kabs2 && zabs(zd{h2k2.x - kp2, h2k2.y}) > kGrazingTolerance * kabs2;
}

struct Fresnel {
  zd r_s, r_p, t_s, t_p, kz_in, kz_out;
};

BWPM_HD Fresnel fresnel(double kx, double ky, zd n1, zd n2, double k0) {
#pragma clang fp contract(off)
  const double kp2 = kx * kx + ky * ky;
  const zd kz1 = kz_of(n1, k0, kp2);
  const zd kz2 = kz_of(n2, k0, kp2);
  const zd q1 = zdiv(kz1, k0), q2 = zdiv(kz2, k0);
  const zd n22q1 = zmul(zmul(n2, n2), q1);
  const zd n11q2 = zmul(zmul(n1, n1), q2);
  const zd den_p = zadd(n22q1, n11q2);
  const zd sum = zadd(q1, q2);
  Fresnel f;
  f.r_s = zdiv(zsub(q1, q2), sum);
  f.r_p = zdiv(zsub(n22q1, n11q2), den_p);
  f.t_s = zdiv(zmul(q1, 2.0), sum);
  f.t_p = zdiv(zmul(zmul(zmul(n1, 2.0), n2), q1), den_p);
  f.kz_in = kz1;
  f.kz_out = kz2;
  return f;
}

// Symmetric 2×2 matrix [[m00, m01], [m01, m11]] (the transmission and reflection matrices are
// symmetric, so m10 is not stored).
struct Mat2 {
  zd m00, m01, m11;
};

// T of wende2023 eqs. (2), (3) without propagation factor; k⊥ = 0 → t_s·I (interface.cpp:30-40).
BWPM_HD Mat2 transmission(double kx, double ky, zd n1, zd n2, double k0) {
#pragma clang fp contract(off)
  const Fresnel c = fresnel(kx, ky, n1, n2, k0);
  const double kp2 = kx * kx + ky * ky;
  if (kp2 == 0.0) return {c.t_s, zd{0.0, 0.0}, c.t_s};
  const zd kr = zmul(n1, k0), ks = zmul(n2, k0);
  const zd gamma = zadd(zmul(c.kz_in, c.kz_out), zdiv(zmul(c.kz_out, kp2), c.kz_in));
  const zd scale = zdiv(zd{1.0, 0.0}, zmul(zmul(kr, kp2), ks));
  const zd trs = zmul(zmul(c.t_s, kr), ks);
  const zd off = zmul(zmul(zmul(zsub(zmul(c.t_p, gamma), trs), kx), ky), scale);
  const zd m00 = zmul(zadd(zmul(zmul(zmul(c.t_p, kx), kx), gamma),
                           zmul(zmul(zmul(zmul(c.t_s, ky), ky), kr), ks)), scale);
  const zd m11 = zmul(zadd(zmul(zmul(zmul(c.t_p, ky), ky), gamma),
                           zmul(zmul(zmul(zmul(c.t_s, kx), kx), kr), ks)), scale);
  return {m00, off, m11};
}

// R of wende2023 eqs. (8), (9) without propagation factor; k⊥ = 0 → r_s·I (interface.cpp:42-52).
BWPM_HD Mat2 reflection(double kx, double ky, zd n1, zd n2, double k0) {
#pragma clang fp contract(off)
  const Fresnel c = fresnel(kx, ky, n1, n2, k0);
  const double kp2 = kx * kx + ky * ky;
  if (kp2 == 0.0) return {c.r_s, zd{0.0, 0.0}, c.r_s};
  const zd kr = zmul(n1, k0);
  const zd gamma = zadd(zmul(c.kz_in, c.kz_in), zd{kp2, 0.0});
  const zd scale = zdiv(zd{1.0, 0.0}, zmul(zmul(kr, kp2), kr));
  const zd rkk = zmul(zmul(c.r_s, kr), kr);
  const zd sum = zadd(zmul(c.r_p, gamma), rkk);
  const zd off = zmul(zmul(zmul(zd{-sum.x, -sum.y}, kx), ky), scale);
  const zd nrp{-c.r_p.x, -c.r_p.y};
  const zd m00 = zmul(zadd(zmul(zmul(zmul(nrp, kx), kx), gamma),
                           zmul(zmul(zmul(zmul(c.r_s, ky), ky), kr), kr)), scale);
  const zd m11 = zmul(zadd(zmul(zmul(zmul(nrp, ky), ky), gamma),
                           zmul(zmul(zmul(zmul(c.r_s, kx), kx), kr), kr)), scale);
  return {m00, off, m11};
}

// Operator kinds of wende2023 table 1 (cpp/include/bwpm/solver.hpp:140).
enum Kind : int { kFwdT = 0, kFwdR = 1, kBwdT = 2, kBwdR = 3 };

// exp(i·f·k_z·dz) with f = 1 or 2.
BWPM_HD zd propagation(zd kz, double dz, double f) {
#pragma clang fp contract(off)
  const zd ikz{-f * kz.y, f * kz.x};
  return zexp(zd{ikz.x * dz, ikz.y * dz});
}

// The operator matrix M_{kind, r, s}(k) of table 1 (solver.cpp:73-106).
BWPM_HD Mat2 operator_matrix(int kind, double kx, double ky, zd nr, zd ns, double k0, double dz) {
#pragma clang fp contract(off)
  const double kp2 = kx * kx + ky * ky;
  Mat2 a;
  zd phase{1.0, 0.0};
  switch (kind) {
    case kFwdT:
      a = transmission(kx, ky, nr, ns, k0);
      phase = propagation(kz_of(ns, k0, kp2), dz, 1.0);
      break;
    case kFwdR:
      a = reflection(kx, ky, nr, ns, k0);
      break;
    case kBwdT:
      a = transmission(kx, ky, nr, ns, k0);
      phase = propagation(kz_of(nr, k0, kp2), dz, 1.0);
      break;
    default:
      a = reflection(kx, ky, nr, ns, k0);
      phase = propagation(kz_of(nr, k0, kp2), dz, 2.0);
  }
  return {zmul(a.m00, phase), zmul(a.m01, phase), zmul(a.m11, phase)};
}

}  // namespace bwpm
