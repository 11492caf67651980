// One-qubit matrices as a real rotation between diagonals (host code).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>

namespace aqs_hip {

// U = diag(g0, g1) R diag(1, z) with R = [[c, -s], [s, c]] real,
// c = |u00| and s = |u10| (both >= 0).
struct RotationForm {
    double c = 1, s = 0;
    std::complex<double> g0 = 1, g1 = 1, z = 1;
};

// False when s = 0 or when the form does not reproduce U to 1e-12 (it holds
// when u00 u11 / (u01 u10) is real and negative: every unitary, and scaled
// unitaries; FP32 rounding of a complex one can break it).
inline bool factor_rotation(std::complex<double> u00, std::complex<double> u01, std::complex<double> u10,
                            std::complex<double> u11, RotationForm& f) {
    using cd = std::complex<double>;
    const double c = std::abs(u00), s = std::abs(u10);
    if (s == 0) return false;
    cd g0, g1 = u10 / s, z(1);
    if (c >= s) {
        g0 = u00 / c;
        z = u11 / (g1 * c);
    } else if (c > 0) {
        g0 = u00 / c;
        z = -u01 / (g0 * s);
    } else {
        g0 = -u01 / s;
    }
    const double err = std::max(std::max(std::abs(g0 * c - u00), std::abs(-g0 * s * z - u01)),
                                std::max(std::abs(g1 * s - u10), std::abs(g1 * c * z - u11)));
    if (!(err <= 1e-12 * std::max(c, s))) return false;
    f.c = c;
    f.s = s;
    f.g0 = g0;
    f.g1 = g1;
    f.z = z;
    return true;
}

// The rotation (c, s) as its norm times a unit rotation applied by three
// shears, p -= t q, q += su p, p -= t q, with t = su / (1 + cu) <= 1 (cu,
// su >= 0).  A unitary's norm, 1 up to rounding, is taken as exactly 1.
inline void rotation_shears(double c, double s, double& t, double& su, double& norm) {
    const double nr = std::sqrt(c * c + s * s);
    norm = std::abs(nr - 1.0) < 1e-15 ? 1.0 : nr;
    const double cu = c / norm;
    su = s / norm;
    t = su / (1.0 + cu);
}

}  // namespace aqs_hip
