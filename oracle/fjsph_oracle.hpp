/*
 * FJSPH oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * An FP64 CPU restatement of the reference's per-step hot path, written from
 * the reference source (file:line cited per function).  It is the checker for
 * the HIP engine: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The engine never links or calls it.
 *
 * The reference cannot be compiled here (Eigen, nanoflann and nlohmann/json
 * are not vendored), so the two third-party algorithms it relies on are
 * restated from their published sources:
 *   - nanoflann (>= 1.5, unpinned): exact (eps = 0) radius search,
 *     RadiusResultSet::addPoint keeps dist < radius, with
 *     L2_Simple dist = sum_k (q_k - p_k)^2 in component order.
 *     Lists here are returned sorted by index; the reference returns tree
 *     order, which changes FP64 sums only in the last bits.
 *   - Eigen 3.3/3.4 (README.md:31, unpinned): ColPivHouseholderQR
 *     (pivoting on the largest remaining column norm, rank rule
 *     |R_ii| > maxpivot * n * eps, inverse = P R^-1 Q^T) and
 *     SelfAdjointEigenSolver::computeDirect (closed-form 2x2 / trigonometric
 *     3x3 on the lower triangle, shifted by trace/n and scaled to [-1,1]).
 *
 * Pinning status of each function is listed in DESIGN.md-independent form in
 * the header of tests/test_oracle.py; what no check pins is called
 * "parity unpinned" there.
 *
 * Deliberate restatement choices (each keeps the reference's single-thread
 * semantics):
 *   - npd in dSPH_PreStep is racy in the reference (src/Shifting.cpp:16-68);
 *     here npd = sum_i sum_{j != i, b_j > PISTON} W_ij / end.
 *   - find_timestep constraints whose denominator is zero (mu = 0, sig = 0,
 *     curve = 0) are skipped, which is what IEEE arithmetic gives (+inf never
 *     wins a min) and what -ffast-math leaves undefined.
 *   - Boundary_DBC scatters into RV_[j] for fluid j beyond the block end
 *     (src/Resid.cpp:84,107), out of bounds; the in-bounds part is empty, so
 *     boundary acc is 0 and Rrho keeps its value.
 */
#ifndef FJSPH_ORACLE_HPP
#define FJSPH_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <stdexcept>
#include <vector>
#include <set>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/fjsph_engine.h"

namespace orc {

static constexpr double kPi = 3.14159265358979323846;
static constexpr double kEps = std::numeric_limits<double>::epsilon();
static constexpr double kDefaultVal = 9999999.0;  /* src/VarDefs.h:195 */
static constexpr long kNoCell = -3;               /* src/VarDefs.h:197 */

/* ---------------------------------------------------------------- vectors */

template <int D>
struct Vec {
    double c[D];
    static Vec zero() { Vec r; for (int k = 0; k < D; ++k) r.c[k] = 0.0; return r; }
    double& operator[](int k) { return c[k]; }
    double operator[](int k) const { return c[k]; }
};
template <int D> inline Vec<D> operator+(const Vec<D>& a, const Vec<D>& b) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
template <int D> inline Vec<D> operator-(const Vec<D>& a, const Vec<D>& b) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] - b.c[k]; return r; }
template <int D> inline Vec<D> operator-(const Vec<D>& a) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = -a.c[k]; return r; }
template <int D> inline Vec<D> operator*(double s, const Vec<D>& a) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = s * a.c[k]; return r; }
template <int D> inline Vec<D> operator*(const Vec<D>& a, double s) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] * s; return r; }
template <int D> inline Vec<D> operator/(const Vec<D>& a, double s) { Vec<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] / s; return r; }
template <int D> inline Vec<D>& operator+=(Vec<D>& a, const Vec<D>& b) { for (int k = 0; k < D; ++k) a.c[k] += b.c[k]; return a; }
template <int D> inline Vec<D>& operator-=(Vec<D>& a, const Vec<D>& b) { for (int k = 0; k < D; ++k) a.c[k] -= b.c[k]; return a; }
template <int D> inline double dot(const Vec<D>& a, const Vec<D>& b) { double s = a.c[0] * b.c[0]; for (int k = 1; k < D; ++k) s += a.c[k] * b.c[k]; return s; }
template <int D> inline double sqnorm(const Vec<D>& a) { return dot(a, a); }
template <int D> inline double norm(const Vec<D>& a) { return std::sqrt(sqnorm(a)); }
/* Eigen MatrixBase::normalized(): a zero vector is returned unchanged. */
template <int D> inline Vec<D> normalized(const Vec<D>& a)
{
    double z = sqnorm(a);
    if (z > 0.0) return a / std::sqrt(z);
    return a;
}

template <int D>
struct Mat {
    double a[D][D];
    static Mat zero() { Mat m; for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) m.a[i][j] = 0.0; return m; }
    static Mat identity() { Mat m = zero(); for (int i = 0; i < D; ++i) m.a[i][i] = 1.0; return m; }
};
template <int D> inline Vec<D> operator*(const Mat<D>& m, const Vec<D>& v)
{
    Vec<D> r;
    for (int i = 0; i < D; ++i) { double s = m.a[i][0] * v.c[0]; for (int k = 1; k < D; ++k) s += m.a[i][k] * v.c[k]; r.c[i] = s; }
    return r;
}

/* --------------------------------------------------------- Eigen restatement */

/* ColPivHouseholderQR<Matrix<double,D,D>> followed by isInvertible() and
 * inverse(); returns false (and leaves inv untouched) when not invertible. */
template <int D>
bool colpiv_qr_inverse(const Mat<D>& A, Mat<D>& inv)
{
    double qr[D][D];
    for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) qr[i][j] = A.a[i][j];
    double hcoeff[D];
    int transp[D];
    double normDirect[D], normUpdated[D];
    double maxColNorm = 0.0;
    for (int j = 0; j < D; ++j) {
        double s = 0.0;
        for (int i = 0; i < D; ++i) s += qr[i][j] * qr[i][j];
        normDirect[j] = std::sqrt(s);
        normUpdated[j] = normDirect[j];
        maxColNorm = std::max(maxColNorm, normDirect[j]);
    }
    const double threshold_helper = (maxColNorm * kEps) * (maxColNorm * kEps) / double(D);
    const double norm_downdate_threshold = std::sqrt(kEps);
    int nonzero_pivots = D;
    double maxpivot = 0.0;
    for (int k = 0; k < D; ++k) {
        int big = k;
        double bigv = normUpdated[k];
        for (int j = k + 1; j < D; ++j) if (normUpdated[j] > bigv) { bigv = normUpdated[j]; big = j; }
        const double big_sq = bigv * bigv;
        if (nonzero_pivots == D && big_sq < threshold_helper * double(D - k)) nonzero_pivots = k;
        transp[k] = big;
        if (big != k) {
            for (int i = 0; i < D; ++i) std::swap(qr[i][k], qr[i][big]);
            std::swap(normUpdated[k], normUpdated[big]);
            std::swap(normDirect[k], normDirect[big]);
        }
        /* makeHouseholderInPlace on qr[k..D-1][k] */
        double tailSq = 0.0;
        for (int i = k + 1; i < D; ++i) tailSq += qr[i][k] * qr[i][k];
        const double c0 = qr[k][k];
        double tau, beta;
        if (tailSq <= std::numeric_limits<double>::min()) {
            tau = 0.0;
            beta = c0;
            for (int i = k + 1; i < D; ++i) qr[i][k] = 0.0;
        } else {
            beta = std::sqrt(c0 * c0 + tailSq);
            if (c0 >= 0.0) beta = -beta;
            for (int i = k + 1; i < D; ++i) qr[i][k] = qr[i][k] / (c0 - beta);
            tau = (beta - c0) / beta;
        }
        hcoeff[k] = tau;
        qr[k][k] = beta;
        if (std::fabs(beta) > maxpivot) maxpivot = std::fabs(beta);
        /* applyHouseholderOnTheLeft to columns k+1.. of rows k.. */
        if (tau != 0.0) {
            for (int j = k + 1; j < D; ++j) {
                double tmp = 0.0;
                for (int i = k + 1; i < D; ++i) tmp += qr[i][k] * qr[i][j];
                tmp += qr[k][j];
                qr[k][j] -= tau * tmp;
                for (int i = k + 1; i < D; ++i) qr[i][j] -= tau * qr[i][k] * tmp;
            }
        }
        for (int j = k + 1; j < D; ++j) {
            if (normUpdated[j] != 0.0) {
                double temp = std::fabs(qr[k][j]) / normUpdated[j];
                temp = (1.0 + temp) * (1.0 - temp);
                temp = temp < 0.0 ? 0.0 : temp;
                const double ratio = normUpdated[j] / normDirect[j];
                const double temp2 = temp * ratio * ratio;
                if (temp2 <= norm_downdate_threshold) {
                    double s = 0.0;
                    for (int i = k + 1; i < D; ++i) s += qr[i][j] * qr[i][j];
                    normDirect[j] = std::sqrt(s);
                    normUpdated[j] = normDirect[j];
                } else {
                    normUpdated[j] *= std::sqrt(temp);
                }
            }
        }
    }
    const double thr = std::fabs(maxpivot) * (double(D) * kEps);
    int rank = 0;
    for (int i = 0; i < nonzero_pivots; ++i) rank += (std::fabs(qr[i][i]) > thr) ? 1 : 0;
    if (rank != D) return false;

    /* inverse = solve(I): c = Q^T I, R c = c, rows permuted back. */
    double c[D][D];
    for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) c[i][j] = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k < D; ++k) {
        const double tau = hcoeff[k];
        if (tau == 0.0) continue;
        for (int j = 0; j < D; ++j) {
            double tmp = c[k][j];
            for (int i = k + 1; i < D; ++i) tmp += qr[i][k] * c[i][j];
            c[k][j] -= tau * tmp;
            for (int i = k + 1; i < D; ++i) c[i][j] -= tau * qr[i][k] * tmp;
        }
    }
    for (int j = 0; j < D; ++j) {
        for (int i = D - 1; i >= 0; --i) {
            double s = c[i][j];
            for (int k = i + 1; k < D; ++k) s -= qr[i][k] * c[k][j];
            c[i][j] = s / qr[i][i];
        }
    }
    int perm[D];
    for (int i = 0; i < D; ++i) perm[i] = i;
    for (int k = 0; k < D; ++k) std::swap(perm[k], perm[transp[k]]);
    for (int i = 0; i < D; ++i) for (int j = 0; j < D; ++j) inv.a[perm[i]][j] = c[i][j];
    return true;
}

/* SelfAdjointEigenSolver::computeDirect, smallest eigenvalue, lower triangle. */
inline double min_eig_direct(const Mat<2>& m)
{
    const double shift = (m.a[0][0] + m.a[1][1]) / 2.0;
    double s00 = m.a[0][0] - shift, s11 = m.a[1][1] - shift, s10 = m.a[1][0];
    double scale = std::max(std::fabs(s00), std::max(std::fabs(s11), std::fabs(s10)));
    if (scale > 0.0) { s00 /= scale; s11 /= scale; s10 /= scale; }
    const double t0 = 0.5 * std::sqrt((s00 - s11) * (s00 - s11) + 4.0 * (s10 * s10));
    const double t1 = 0.5 * (s00 + s11);
    double r0 = t1 - t0;
    return r0 * scale + shift;
}

inline double min_eig_direct(const Mat<3>& mat)
{
    const double shift = (mat.a[0][0] + mat.a[1][1] + mat.a[2][2]) / 3.0;
    double m[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m[i][j] = (i >= j) ? mat.a[i][j] : mat.a[j][i];
    for (int i = 0; i < 3; ++i) m[i][i] -= shift;
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) scale = std::max(scale, std::fabs(m[i][j]));
    if (scale > 0.0) for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] /= scale;
    const double s_inv3 = 1.0 / 3.0;
    const double s_sqrt3 = std::sqrt(3.0);
    const double c0 = m[0][0] * m[1][1] * m[2][2] + 2.0 * m[1][0] * m[2][0] * m[2][1] - m[0][0] * m[2][1] * m[2][1]
                      - m[1][1] * m[2][0] * m[2][0] - m[2][2] * m[1][0] * m[1][0];
    const double c1 = m[0][0] * m[1][1] - m[1][0] * m[1][0] + m[0][0] * m[2][2] - m[2][0] * m[2][0]
                      + m[1][1] * m[2][2] - m[2][1] * m[2][1];
    const double c2 = m[0][0] + m[1][1] + m[2][2];
    const double c2_over_3 = c2 * s_inv3;
    double a_over_3 = (c2 * c2_over_3 - c1) * s_inv3;
    a_over_3 = std::max(a_over_3, 0.0);
    const double half_b = 0.5 * (c0 + c2_over_3 * (2.0 * c2_over_3 * c2_over_3 - c1));
    double q = a_over_3 * a_over_3 * a_over_3 - half_b * half_b;
    q = std::max(q, 0.0);
    const double rho = std::sqrt(a_over_3);
    const double theta = std::atan2(std::sqrt(q), half_b) * s_inv3;
    const double cos_theta = std::cos(theta);
    const double sin_theta = std::sin(theta);
    const double r0 = c2_over_3 - rho * (cos_theta + s_sqrt3 * sin_theta);
    const double r1 = c2_over_3 - rho * (cos_theta - s_sqrt3 * sin_theta);
    const double r2 = c2_over_3 + 2.0 * rho * cos_theta;
    /* minCoeff of the rescaled eigenvalues */
    double e0 = r0 * scale + shift, e1 = r1 * scale + shift, e2 = r2 * scale + shift;
    return std::min(e0, std::min(e1, e2));
}

/* ------------------------------------------------------------ settings */

/* get_n_full (src/Geometry.cpp:282-308): lattice points from -2(H+dx) to
 * 2(H+dx) by repeated addition of dx, counted inside 4H^2 (strict). */
inline double get_n_full(int dim, double dx, double H)
{
    const double sr = 4.0 * H * H;
    long count = 0;
    for (double x = -2.0 * (H + dx); x <= 2.0 * (H + dx); x += dx)
        for (double y = -2.0 * (H + dx); y <= 2.0 * (H + dx); y += dx) {
            if (dim == 3) {
                for (double z = -2.0 * (H + dx); z <= 2.0 * (H + dx); z += dx) {
                    double d = (0.0 + x * x) + y * y;
                    d = d + z * z;
                    if (d < sr) ++count;
                }
            } else {
                double d = (0.0 + x * x) + y * y;
                if (d < sr) ++count;
            }
        }
    return double(count);
}

/* Set_Values (src/IO.cpp:26-128) and AERO::GetYcoef (src/Var.h:244-266). */
inline fjsph_derived derive(const fjsph_settings& s)
{
    fjsph_derived d;
    std::memset(&d, 0, sizeof(d));
    const int D = s.dim;
    d.B = s.rho_rest * std::pow(s.speed_sound, 2) / s.gam;
    double rho_pipe;
    if (s.eos == FJSPH_COLE_EOS) rho_pipe = s.rho_rest * std::pow(((s.press_pipe - s.press_back) / d.B) + 1.0, 1.0 / s.gam);
    else rho_pipe = (s.press_pipe - s.press_back) / (s.speed_sound * s.speed_sound) + s.rho_rest;
    d.rho_max = s.rho_max;
    d.rho_min = s.rho_min;
    if (s.rho_max == 1500 && s.rho_min == 500) {
        d.rho_max = s.rho_rest * (1.0 + s.rho_var * 0.01);
        d.rho_min = s.rho_rest * (1.0 - s.rho_var * 0.01);
    }
    d.dx = s.particle_step * std::pow(rho_pipe / s.rho_rest, 1.0 / D);
    d.sim_mass = s.rho_rest * std::pow(s.particle_step, D);
    d.bnd_mass = d.sim_mass;
    d.sos = std::sqrt(s.temp_g * s.R_g * s.gamma_g);
    d.i_sos_sq = 1.0 / (d.sos * d.sos);
    d.H = s.H_fac * s.particle_step;
    d.H_sq = d.H * d.H;
    d.sr = 4 * d.H_sq;
    d.dsph_cont = 2.0 * s.dsph_delta * d.H * s.speed_sound;
    d.nu = s.mu / s.rho_rest;
    if (D == 2) d.W_correc = 7.0 / (4.0 * kPi * d.H * d.H);
    else d.W_correc = (21 / (16 * kPi * d.H * d.H * d.H));
    d.W_dx = (std::pow(1 - 0.5 * s.particle_step / d.H, 4)) * (2 * s.particle_step / d.H + 1) * d.W_correc;
    /* GetYcoef(fluid, particle_step) */
    const double diam = s.particle_step;
    if (D == 3) {
        d.L = diam * std::cbrt(3.0 / (4.0 * kPi));
        d.A_sphere = kPi * d.L * d.L;
        d.A_plate = diam * diam;
    } else {
        d.L = diam / std::sqrt(kPi);
        d.A_sphere = 2 * d.L;
        d.A_plate = diam;
    }
    d.td = (2.0 * s.rho_rest * std::pow(d.L, D - 1)) / (s.Cd * s.mu);
    d.omega = std::sqrt((s.Ck * s.sig) / (s.rho_rest * std::pow(d.L, D)) - 1.0 / std::pow(d.td, 2.0));
    d.tmax = -2.0 * (std::atan(std::sqrt(std::pow(d.td * d.omega, 2.0) + 1) + d.td * d.omega) - kPi) / d.omega;
    d.Cdef = 1.0 - std::exp(-d.tmax / d.td) * (std::cos(d.omega * d.tmax) + 1 / (d.omega * d.td) * std::sin(d.omega * d.tmax));
    d.ycoef = 0.5 * d.Cdef * (s.Cf / (s.Ck * s.Cb)) * (s.rho_g * d.L) / s.sig;
    d.n_full = get_n_full(D, s.particle_step, d.H);
    d.i_n_full = 1.0 / d.n_full;
    d.interp_fac = 1.0 / s.i_interp_fac;
    if (D == 3) d.A_plate = s.particle_step * s.particle_step;
    else d.A_plate = s.particle_step;
    d.lam_st = (6.0 / 81.0 * std::pow((2.0 * d.H), 3.0) / std::pow(kPi, 4.0) *
                (9.0 / 4.0 * std::pow(kPi, 3.0) - 6.0 * kPi - 4.0));
    return d;
}

/* FLUID::get_pressure / get_density (src/Var.h:203-236). */
inline double get_pressure(const fjsph_settings& s, const fjsph_derived& d, double rho)
{
    if (s.eos == FJSPH_COLE_EOS) return d.B * (std::pow(rho / s.rho_rest, s.gam) - 1) + s.press_back;
    return s.speed_sound * s.speed_sound * (rho - s.rho_rest) + s.press_back;
}
inline double get_density(const fjsph_settings& s, const fjsph_derived& d, double press)
{
    if (s.eos == FJSPH_COLE_EOS) return s.rho_rest * std::pow(((press - s.press_back) / d.B) + 1.0, 1.0 / s.gam);
    return (press - s.press_back) / (s.speed_sound * s.speed_sound) + s.rho_rest;
}

/* ------------------------------------------------------------- particles */

/* The SPHPart fields the hot path uses (src/Var.h:600-641). */
template <int D>
struct Part {
    int64_t part_id = 0;
    long cellID = kNoCell;
    uint32_t b = 0, surf = 0, surfzone = 0, internal = 0, n_failed = 0;
    Vec<D> xi = Vec<D>::zero(), v = Vec<D>::zero(), acc = Vec<D>::zero(), Af = Vec<D>::zero(), aVisc = Vec<D>::zero();
    double Rrho = 0, rho = 0, p = 0, m = 0, curve = 0, norm_curve = 0, woccl = 0, pDist = 0, deltaD = 0;
    Vec<D> cellV = Vec<D>::zero();
    double cellP = 0, cellRho = 0;
    Mat<D> L = Mat<D>::zero();
    Vec<D> gradRho = Vec<D>::zero(), norm = Vec<D>::zero();
    double colourG = 0, colour = 0, lam = 0, lam_nb = 0, kernsum = 0;
    Vec<D> vPert = Vec<D>::zero();
};

typedef std::vector<std::vector<std::pair<int64_t, double>>> OUTL;

/* Neighbour search: every j with sum_k (x_ik - x_jk)^2 < sr, computed in
 * component order in FP64 (nanoflann L2_Simple + RadiusResultSet, eps = 0,
 * src/Neighbours.cpp:15-47, src/Var.h:892-899).  Self included. A uniform
 * grid of edge sqrt(sr) restricts the candidates; it does not change the set.
 * Lists are sorted by index. */
template <int D>
OUTL find_neighbours(const std::vector<Part<D>>& pts, double sr, int nthreads)
{
    const size_t n = pts.size();
    OUTL out(n);
    if (n == 0) return out;
    double lo[D], hi[D];
    for (int k = 0; k < D; ++k) { lo[k] = pts[0].xi[k]; hi[k] = lo[k]; }
    for (const auto& p : pts)
        for (int k = 0; k < D; ++k) { lo[k] = std::min(lo[k], p.xi[k]); hi[k] = std::max(hi[k], p.xi[k]); }
    const double cs = std::sqrt(sr) * 1.0000001 + 1e-300;
    long dims[D];
    for (int k = 0; k < D; ++k) dims[k] = long(std::floor((hi[k] - lo[k]) / cs)) + 1;
    long ncell = 1;
    for (int k = 0; k < D; ++k) ncell *= dims[k];
    std::vector<long> cell(n);
    auto cidx = [&](const Vec<D>& x, long* ci) {
        for (int k = 0; k < D; ++k) {
            long c = long(std::floor((x[k] - lo[k]) / cs));
            ci[k] = std::min(std::max(c, 0L), dims[k] - 1);
        }
    };
    for (size_t i = 0; i < n; ++i) {
        long ci[D];
        cidx(pts[i].xi, ci);
        long lin = ci[D - 1];
        for (int k = D - 2; k >= 0; --k) lin = lin * dims[k] + ci[k];
        cell[i] = lin;
    }
    std::vector<long> start(ncell + 1, 0);
    for (size_t i = 0; i < n; ++i) start[cell[i] + 1]++;
    for (long c = 0; c < ncell; ++c) start[c + 1] += start[c];
    std::vector<int64_t> sorted(n);
    {
        std::vector<long> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; ++i) sorted[fill[cell[i]]++] = int64_t(i);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long i = 0; i < long(n); ++i) {
        long ci[D];
        cidx(pts[i].xi, ci);
        auto& lst = out[i];
        lst.reserve(D == 3 ? 250 : 47);
        long lo_c[D], hi_c[D];
        for (int k = 0; k < D; ++k) { lo_c[k] = std::max(ci[k] - 1, 0L); hi_c[k] = std::min(ci[k] + 1, dims[k] - 1); }
        long cz0 = D == 3 ? lo_c[D - 1] : 0, cz1 = D == 3 ? hi_c[D - 1] : 0;
        for (long cz = cz0; cz <= cz1; ++cz)
            for (long cy = lo_c[1]; cy <= hi_c[1]; ++cy)
                for (long cx = lo_c[0]; cx <= hi_c[0]; ++cx) {
                    long lin = (D == 3) ? ((cz * dims[1] + cy) * dims[0] + cx) : (cy * dims[0] + cx);
                    for (long s = start[lin]; s < start[lin + 1]; ++s) {
                        const int64_t j = sorted[s];
                        double dist = 0.0;
                        for (int k = 0; k < D; ++k) {
                            const double diff = pts[i].xi[k] - pts[j].xi[k];
                            dist = dist + diff * diff;
                        }
                        if (dist < sr) lst.emplace_back(j, dist);
                    }
                }
        std::sort(lst.begin(), lst.end(), [](const std::pair<int64_t, double>& a, const std::pair<int64_t, double>& b) { return a.first < b.first; });
    }
    return out;
}

/* Exhaustive O(N^2) search with the same predicate (used to pin the grid). */
template <int D>
OUTL brute_neighbours(const std::vector<Part<D>>& pts, double sr)
{
    const size_t n = pts.size();
    OUTL out(n);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            double dist = 0.0;
            for (int k = 0; k < D; ++k) {
                const double diff = pts[i].xi[k] - pts[j].xi[k];
                dist = dist + diff * diff;
            }
            if (dist < sr) out[i].emplace_back(int64_t(j), dist);
        }
    return out;
}

/* ------------------------------------------------------ pointwise kernels */

/* Wendland C2 (src/Kernel.h:37-45); not clipped at q >= 2. */
inline double Kernel(double dist, double H, double W_correc)
{
    return (std::pow(1 - 0.5 * dist / H, 4)) * (2 * dist / H + 1) * W_correc;
}
/* src/Kernel.h:48-61 */
template <int D>
inline Vec<D> GradK(const Vec<D>& Rij, double dist, double H, double W_correc)
{
    if (dist / H < 1e-12) return Vec<D>::zero();
    const double pw = std::pow(1 - 0.5 * dist / H, 3);
    Vec<D> r;
    for (int k = 0; k < D; ++k) r.c[k] = 5.0 * (Rij.c[k] / (H * H)) * pw * W_correc;
    return r;
}
/* pairwise_ST_fac / pairwise_ST (src/Kernel.h:88-113) */
inline double pairwise_ST_fac(uint32_t bA, uint32_t bB)
{
    if (bA == FJSPH_BOUND || bB == FJSPH_BOUND) {
        const double contang = 0.5 * kPi * 7.0 / 9.0;
        return (1.0 + 0.5 * std::cos(contang));
    }
    return 1.0;
}
template <int D>
inline Vec<D> pairwise_ST(const Vec<D>& Rji, double r, double h, double npdm2, double pi3o4, uint32_t bA, uint32_t bB)
{
    const double fac = pairwise_ST_fac(bA, bB);
    const double s = -npdm2 * fac * std::cos(pi3o4 * r / h);
    Vec<D> out;
    for (int k = 0; k < D; ++k) out.c[k] = s * (Rji.c[k] / r);
    return out;
}
/* BasePos (src/Kernel.h:153-157) */
template <int D>
inline Vec<D> BasePos(const Part<D>& pi, const Part<D>& pj, const Vec<D>& gradK)
{
    const double s = (pi.p / (pi.rho * pi.rho) + pj.p / (pj.rho * pj.rho));
    Vec<D> r;
    for (int k = 0; k < D; ++k) r.c[k] = pj.m * gradK.c[k] * s;
    return r;
}
/* Continuity_dSPH, both overloads (src/Kernel.h:200-214) */
template <int D>
inline double Continuity_dSPH_grad(const Vec<D>& Rji, double idist2, const Vec<D>& Grad, double volj,
                                   const Vec<D>& gRho_i, const Vec<D>& gRho_j, const Part<D>& pi, const Part<D>& pj)
{
    return volj * ((pj.rho - pi.rho) + 0.5 * dot(gRho_i + gRho_j, Rji)) * dot(Rji, Grad) * idist2;
}
template <int D>
inline double Continuity_dSPH_plain(const Vec<D>& Rji, double idist2, const Vec<D>& Grad, double volj,
                                    const Part<D>& pi, const Part<D>& pj)
{
    return volj * (pj.rho - pi.rho) * dot(Rji, Grad) * idist2;
}
/* ArtVisc (src/Kernel.h:217-244) */
template <int D>
inline Vec<D> ArtVisc(const fjsph_settings& s, const fjsph_derived& d, const Part<D>& pi, const Part<D>& pj,
                      const Vec<D>& Rji, const Vec<D>& Vji, double idist2, const Vec<D>& gradK)
{
    const double vdotr = dot(Vji, Rji);
    if (vdotr > 0.0) return Vec<D>::zero();
    const double muij = d.H * vdotr * idist2;
    const double rhoij = 0.5 * (pi.rho + pj.rho);
    const double cbar = 0.5 * (std::sqrt((d.B * s.gam) / pi.rho) + std::sqrt((d.B * s.gam) / pj.rho));
    const double visc_alpha = s.visc_alpha;
    Vec<D> r;
    for (int k = 0; k < D; ++k) r.c[k] = gradK.c[k] * visc_alpha * cbar * muij / rhoij;
    return r;
}
/* Viscosity (src/Kernel.h:262-269) */
template <int D>
inline Vec<D> Viscosity(double nu, const Part<D>& pi, const Part<D>& pj, const Vec<D>& Rji, const Vec<D>& Vji,
                        double idist2, const Vec<D>& gradK)
{
    const double s = nu * (pi.rho + pj.rho) / (pi.rho * pj.rho) * (dot(Rji, gradK)) * idist2;
    Vec<D> r;
    for (int k = 0; k < D; ++k) r.c[k] = s * Vji.c[k];
    return r;
}
/* ALE terms (src/Kernel.h:179-195) */
template <int D>
inline Vec<D> ALEMomentum(const Part<D>& pi, const Part<D>& pj, double Vj, const Vec<D>& g)
{
    const double a = dot(pj.vPert, g), b = dot(pi.vPert, g), c = dot(pj.vPert - pi.vPert, g);
    Vec<D> r;
    for (int k = 0; k < D; ++k) r.c[k] = ((pj.v.c[k] * a + pi.v.c[k] * b) - pi.v.c[k] * c) * Vj;
    return r;
}
template <int D>
inline double ALEContinuity(const Part<D>& pi, const Part<D>& pj, double Vj, const Vec<D>& g)
{
    return dot((pj.v + pj.vPert) - (pi.v + pi.vPert), g) * Vj;
}
template <int D>
inline double ALECont2ndterm(const Part<D>& pi, const Part<D>& pj, double Vj, const Vec<D>& g)
{
    return dot(pj.rho * pj.vPert + pi.rho * pi.vPert, g) * Vj;
}

/* GetCd (src/Aero.h:10-18) */
inline double GetCd(double Re)
{
    return (1.0 + 0.197 * std::pow(Re, 0.63) + 2.6e-04 * std::pow(Re, 1.38)) * (24.0 / (Re + 0.00001));
}

/* gissler_force (src/Aero.h:37-98) */
template <int D>
inline Vec<D> gissler_force(const fjsph_settings& s, const fjsph_derived& d, const Vec<D>& Vdiff, double rho,
                            double press, double mass, double lam, double nnieghb, double woccl)
{
    const double Vn = norm(Vdiff);
    const double Re = 2.0 * rho * Vn * d.L / s.mu_g;
    double frac2;
    if (s.use_lam) frac2 = std::min(d.interp_fac * lam, 1.0);
    else frac2 = std::min(d.interp_fac * nnieghb * d.i_n_full, 1.0);
    const double frac1 = (1.0 - frac2);
    const double Cds = GetCd(Re);
    double Cdl, Adrop;
    if (s.use_TAB_def) {
        double ymax = sqnorm(Vdiff) * d.ycoef;
        if (ymax > 1.0) ymax = 1.0;
        Cdl = Cds * (1 + 2.632 * ymax);
        if (D == 3) Adrop = kPi * std::pow((d.L + s.Cb * d.L * ymax), 2);
        else Adrop = d.A_sphere + 2 * (s.Cb * d.L * ymax);
    } else {
        Cdl = Cds;
        Adrop = d.A_sphere;
    }
    const double Cdi = frac1 * Cdl + frac2;
    const double Aunocc = (frac1 * Adrop + frac2 * d.A_plate);
    const double Ai = (1.0 - woccl) * Aunocc;
    Vec<D> r;
    for (int k = 0; k < D; ++k)
        r.c[k] = 0.5 * Vn * Vdiff.c[k] / (d.sos * d.sos) * s.gamma_g * press * Cdi * Ai / mass;
    return r;
}

/* induced_pressure (src/Aero.h:106-202) */
template <int D>
inline Vec<D> induced_pressure(const fjsph_settings& s, const fjsph_derived& d, const Vec<D>& Vdiff, const Vec<D>& nrm,
                               double lam, double nnieghb, const Part<D>& pi)
{
    const Vec<D> nn = normalized(nrm);
    const double theta = std::fabs(std::acos(dot(-nn, normalized(Vdiff))));
    double Cp_s, Cp_p, Cp_b, Cp_tot;
    if (theta < 2.4455) Cp_s = 1.0 - (2.25) * std::pow(std::sin(theta), 2.0);
    else Cp_s = 0.075;
    if (theta < 1.570797) Cp_p = std::cos(theta);
    else if (theta < 1.9918) Cp_p = -std::pow(std::cos(6.0 * theta + 0.5 * kPi), 1.5);
    else if (theta < 2.0838) Cp_p = 5.5836 * theta - 11.5601;
    else Cp_p = 0.075;
    if (theta < 0.7854) Cp_b = 1.0;
    else if (theta < 1.570797) Cp_b = 0.5 * (std::cos(4.0 * theta - kPi) + 1.0);
    else Cp_b = 0.0;
    const double normCurve = pi.norm_curve;
    const double fac1 = 0.25, ifac1 = 1 / fac1;
    if (normCurve < -fac1) Cp_tot = Cp_b;
    else if (normCurve < 0.0) { const double frac = (normCurve + fac1) * ifac1; Cp_tot = frac * Cp_b + (1.0 - frac) * Cp_p; }
    else if (normCurve < fac1) { const double frac = (normCurve)*ifac1; Cp_tot = frac * Cp_p + (1.0 - frac) * Cp_s; }
    else Cp_tot = Cp_s;
    const double Plocali = 0.5 * sqnorm(Vdiff) / (d.sos * d.sos) * s.gamma_g * pi.cellP * Cp_tot;
    const double Pi = Plocali;
    const double Re = pi.cellRho * norm(Vdiff) * d.L / s.mu_g;
    const double Cdi = GetCd(Re);
    const double Vn = norm(Vdiff);
    const double Vnorm = dot(Vdiff, nn);
    const Vec<D> Vpar = Vdiff - Vnorm * nn;
    const double Vpn = norm(Vpar);
    const double Re_par = pi.cellRho * Vpn * d.L / s.mu_g;
    const double Cf = 0.027 / std::pow(Re_par + 1e-6, 1.0 / 7.0);
    double frac1;
    if (s.use_lam) frac1 = std::min(d.interp_fac * lam, 1.0);
    else frac1 = std::min(d.interp_fac * nnieghb * d.i_n_full, 1.0);
    Vec<D> r;
    for (int k = 0; k < D; ++k) {
        const double acc_drop = 0.5 * Vdiff.c[k] * Vn / (d.sos * d.sos) * s.gamma_g * pi.cellP * (kPi * d.L * d.L * 0.25) * Cdi / pi.m;
        const double acc_kern = -Pi * d.A_plate * nn.c[k] / pi.m;
        const double acc_skin = 0.5 * Vpn * Vpar.c[k] / (d.sos * d.sos) * s.gamma_g * pi.cellP * Cf * d.A_plate / pi.m;
        r.c[k] = (frac1 * (acc_kern + acc_skin) + (1.0 - frac1) * acc_drop);
    }
    return r;
}

/* CalcAeroAcc (src/Aero.h:204-263) */
template <int D>
inline Vec<D> CalcAeroAcc(const fjsph_settings& s, const fjsph_derived& d, const Part<D>& pi, const Vec<D>& Vdiff,
                          const Vec<D>& nrm, double lam, double nneigh)
{
    switch (s.aero_case) {
    case FJSPH_GISSLER:
        return gissler_force<D>(s, d, Vdiff, pi.cellRho, pi.cellP, pi.m, lam, nneigh, pi.woccl);
    case FJSPH_INDUCED_PRESSURE:
        return induced_pressure<D>(s, d, Vdiff, nrm, lam, nneigh, pi);
    case FJSPH_SKIN_FRIC: {
        const Vec<D> nn = normalized(nrm);
        const double Vnorm = dot(Vdiff, nn);
        if (Vnorm > 0.001) {
            const double Re = s.rho_g * norm(Vdiff) * d.L / s.mu_g;
            const Vec<D> Vpar = Vdiff - std::fabs(Vnorm) * nn;
            const double Cf = 0.027 / std::pow(Re, 1.0 / 7.0);
            const double frac2 = std::min(1.5 * lam, 1.0);
            const double frac1 = (1.0 - frac2);
            const double Cdi = GetCd(Re);
            const double Vpn = norm(Vpar), Vn = norm(Vdiff);
            Vec<D> r;
            for (int k = 0; k < D; ++k) {
                const double acc_press = 0.5 * s.rho_g * Vnorm * Vnorm * d.A_plate * nn.c[k] / pi.m;
                const double acc_skin = 0.5 * s.rho_g * Vpn * Cf * d.A_plate * Vpar.c[k] / pi.m;
                const double acc_drop = 0.5 * s.rho_g * Vn * Vdiff.c[k] * (kPi * d.L * d.L / 4) * Cdi / pi.m;
                r.c[k] = frac2 * (acc_press + acc_skin) + frac1 * acc_drop;
            }
            return r;
        }
        return Vec<D>::zero();
    }
    default:
        return Vec<D>::zero();
    }
}

/* ---------------------------------------------------------------- blocks */

struct Block {
    fjsph_block b;
    std::vector<double> times;
    std::vector<double> vels; /* n_times * 3 */
    /* planes (bound_block, src/Var.h:858-866), unset at default_val */
    fjsph_block_planes pl = unset_planes();
    std::vector<size_t> back;                 /* inlet: back particle of each column  */
    std::vector<std::vector<size_t>> buffer;  /* inlet: buffer particles of each column */

    static fjsph_block_planes unset_planes()
    {
        fjsph_block_planes q{};
        for (int a = 0; a < 3; ++a) q.delete_norm[a] = q.insert_norm[a] = q.aero_norm[a] = FJSPH_NO_PLANE;
        q.delconst = q.insconst = q.aeroconst = FJSPH_NO_PLANE;
        return q;
    }
};

template <int D>
inline double dot_raw(const Vec<D>& x, const double* n)
{
    double s = 0.0;
    for (int a = 0; a < D; ++a) s += x[a] * n[a];
    return s;
}

/* --------------------------------------------------------------- the sim */

/* ----------------------------------------- aerodynamic mesh (row a14)
 * MESH (src/Var.h:396-451) as flat arrays: faces are vertex lists (CSR),
 * leftright[f] = (left cell, right cell or a negative boundary marker),
 * cFaces[c] the faces of cell c (CSR); cell centre, velocity, pressure and
 * density from Average_Point_Data_to_Cell. */
struct MeshData {
    int dim = 3;
    std::vector<double> verts;                 /* nverts * dim */
    std::vector<int64_t> face_off, face_v;     /* CSR faces -> vertices */
    std::vector<int64_t> face_l, face_r;       /* leftright */
    std::vector<int64_t> cell_off, cell_f;     /* CSR cells -> faces */
    std::vector<double> cCentre, cVel, cP, cRho;
    size_t ncells() const { return cell_off.empty() ? 0 : cell_off.size() - 1; }
    const double* vert(int64_t v) const { return &verts[size_t(v) * dim]; }
};

/* Determinant of the 4x4 matrix with rows (p_k, 1), k = 0..3, by Laplace
 * expansion along the first two columns (2x2 minors of columns 0-1 times
 * complementary 2x2 minors of columns 2-3), the structure of Eigen 3.4's
 * fixed-size 4x4 determinant.  The reference only uses its sign. */
inline double det4_rows(const double* p0, const double* p1, const double* p2, const double* p3)
{
    const double* r[4] = {p0, p1, p2, p3};
    auto m = [&](int i, int j) { return j < 3 ? r[i][j] : 1.0; };
    auto d2 = [&](int i0, int i1) { return m(i0, 0) * m(i1, 1) - m(i1, 0) * m(i0, 1); };
    const double d2_01 = d2(0, 1), d2_02 = d2(0, 2), d2_03 = d2(0, 3);
    const double d2_12 = d2(1, 2), d2_13 = d2(1, 3), d2_23 = d2(2, 3);
    auto d3 = [&](int i0, double a0, int i1, double a1, int i2, double a2) {
        return m(i0, 2) * a0 + (-m(i1, 2) * a1 + m(i2, 2) * a2);
    };
    const double d3_0 = d3(1, d2_23, 2, d2_13, 3, d2_12);
    const double d3_1 = d3(0, d2_23, 2, d2_03, 3, d2_02);
    const double d3_2 = d3(0, d2_13, 1, d2_03, 3, d2_01);
    const double d3_3 = d3(0, d2_12, 1, d2_02, 2, d2_01);
    return (-m(0, 3) * d3_0 + m(1, 3) * d3_1) + (-m(2, 3) * d3_2 + m(3, 3) * d3_3);
}

/* Crossings3D (src/Geometry.cpp:485-575): does the segment testp-rayp cross
 * the face?  Signed volumes of tetrahedra; uses face vertices 0, 1, 2 for the
 * plane and the edges (last, 0), (0, 1), (1, 2). */
inline int crossings3d(const MeshData& M, int64_t f, const double* testp, const double* rayp)
{
    const int64_t* fv = &M.face_v[size_t(M.face_off[size_t(f)])];
    const int64_t nv = M.face_off[size_t(f) + 1] - M.face_off[size_t(f)];
    const int flag1 = det4_rows(testp, M.vert(fv[0]), M.vert(fv[1]), M.vert(fv[2])) < 0.0;
    const int flag2 = det4_rows(rayp, M.vert(fv[0]), M.vert(fv[1]), M.vert(fv[2])) < 0.0;
    if (flag1 == flag2) return 0;
    const double* vtx0 = M.vert(fv[nv - 1]);
    const double* vtx1 = M.vert(fv[0]);
    const int flag3 = det4_rows(testp, vtx0, vtx1, rayp) < 0.0;
    for (int ii = 1; ii < 3; ++ii) {
        vtx0 = vtx1;
        vtx1 = M.vert(fv[ii]);
        const int flag4 = det4_rows(testp, vtx0, vtx1, rayp) < 0.0;
        if (flag4 != flag3) return 0;
    }
    return 1;
}

/* Crossings2D (src/Geometry.cpp:354-395): does the +x ray from the point
 * cross the edge? */
inline int crossings2d(const MeshData& M, int64_t f, const double* point)
{
    const int64_t* fv = &M.face_v[size_t(M.face_off[size_t(f)])];
    const double tx = point[0], ty = point[1];
    const double* vtx0 = M.vert(fv[0]);
    const double* vtx1 = M.vert(fv[1]);
    const int yflag0 = vtx0[1] >= ty, yflag1 = vtx1[1] >= ty;
    if (yflag0 != yflag1)
        if (((vtx1[1] - ty) * (vtx1[0] - vtx0[0]) >= (vtx1[0] - tx) * (vtx1[1] - vtx0[1])) == yflag1) return 1;
    return 0;
}

/* CheckCell (src/Containment.cpp:385-421): parity of the face crossings of
 * the +x ray (3D: to testp + 1e5 x), stopping after the second hit. */
inline int check_cell(const MeshData& M, int64_t cell, const double* testp)
{
    double rayp[3] = {testp[0] + 1e+5, testp[1], M.dim == 3 ? testp[2] : 0.0};
    int line_flag = 0, inside_flag = 0;
    for (int64_t k = M.cell_off[size_t(cell)]; k < M.cell_off[size_t(cell) + 1]; ++k) {
        const int64_t f = M.cell_f[size_t(k)];
        const int hit = M.dim == 3 ? crossings3d(M, f, testp, rayp) : crossings2d(M, f, testp);
        if (hit) {
            inside_flag = !inside_flag;
            if (line_flag) break;
            line_flag = 1;
        }
    }
    return inside_flag;
}

/* Average_Point_Data_to_Cell (src/CDFIO.cpp:112-155): over the cell's unique
 * vertices (ascending index), compensated (Kahan) sums divided by their
 * count, for the centre, velocity, pressure and density. */
inline void average_point_data_to_cell(MeshData& M, const double* vvel, const double* vp, const double* vrho)
{
    const size_t nc = M.ncells();
    const int D = M.dim;
    M.cCentre.assign(nc * D, 0.0);
    M.cVel.assign(nc * D, 0.0);
    M.cP.assign(nc, 0.0);
    M.cRho.assign(nc, 0.0);
    auto kahan = [](const std::set<int64_t>& idx, auto&& val) {
        double sum = 0.0, c = 0.0;
        for (int64_t ii : idx) {
            const double y = val(ii) - c;
            const double t = sum + y;
            c = (t - sum) - y;
            sum = t;
        }
        return sum;
    };
    for (size_t ii = 0; ii < nc; ++ii) {
        std::set<int64_t> elem;
        for (int64_t k = M.cell_off[ii]; k < M.cell_off[ii + 1]; ++k) {
            const int64_t f = M.cell_f[size_t(k)];
            for (int64_t q = M.face_off[size_t(f)]; q < M.face_off[size_t(f) + 1]; ++q) elem.insert(M.face_v[size_t(q)]);
        }
        const double nv = double(elem.size());
        for (int a = 0; a < D; ++a) {
            M.cCentre[ii * D + a] = kahan(elem, [&](int64_t v) { return M.verts[size_t(v) * D + a]; }) / nv;
            M.cVel[ii * D + a] = kahan(elem, [&](int64_t v) { return vvel[size_t(v) * D + a]; }) / nv;
        }
        M.cP[ii] = kahan(elem, [&](int64_t v) { return vp[size_t(v)]; }) / nv;
        M.cRho[ii] = kahan(elem, [&](int64_t v) { return vrho[size_t(v)]; }) / nv;
    }
}

/* The k nearest cell centres of a point, nearest first (the result order
 * of nanoflann's KNNResultSet; equal distances by cell index), by brute
 * force. */
inline std::vector<int64_t> knn_cells(const MeshData& M, const double* p, size_t k)
{
    const size_t nc = M.ncells();
    std::vector<std::pair<double, int64_t>> dd(nc);
    for (size_t c = 0; c < nc; ++c) {
        double d2 = 0.0;
        for (int a = 0; a < M.dim; ++a) {
            const double t = M.cCentre[c * M.dim + a] - p[a];
            d2 += t * t;
        }
        dd[c] = {d2, int64_t(c)};
    }
    k = std::min(k, nc);
    std::partial_sort(dd.begin(), dd.begin() + long(k), dd.end());
    std::vector<int64_t> out(k);
    for (size_t i = 0; i < k; ++i) out[i] = dd[i].second;
    return out;
}

/* FindCell's 500-nearest list as the reference iterates it: the nearest
 * centres first, then, with fewer cells than k, the zero-initialised tail
 * of its result vector (cell 0). */
inline std::vector<int64_t> knn_cells_padded(const MeshData& M, const double* p, size_t k)
{
    std::vector<int64_t> out = knn_cells(M, p, k);
    out.resize(k, 0);
    return out;
}

/* get_line_intersection (src/Geometry.cpp:312-341): does the segment
 * p1-cellC cross the edge?  One-sided: denominators below machine epsilon
 * (negative ones included) report no crossing. */
inline int line_intersection(const MeshData& M, int64_t f, const double* p1, const double* cellC)
{
    const int64_t* fv = &M.face_v[size_t(M.face_off[size_t(f)])];
    const double* e1 = M.vert(fv[0]);
    const double* e2 = M.vert(fv[1]);
    const double s0 = cellC[0] - p1[0], s1 = cellC[1] - p1[1];
    const double r0 = e2[0] - e1[0], r1 = e2[1] - e1[1];
    const double denom = (-r0 * s1 + s0 * r1);
    if (denom < std::numeric_limits<double>::epsilon()) return 0;
    const double u = (-s1 * (p1[0] - e1[0]) + s0 * (p1[1] - e1[1])) / denom;
    const double t = (r0 * (p1[1] - e1[1]) - r1 * (p1[0] - e1[0])) / denom;
    return (u > 0 && u < 1 && t > 0 && t < 1) ? 1 : 0;
}

/* FindCell's boundary test for a point in no cell (src/Containment.cpp:
 * 734-778): for each listed cell in order, the segment from the point to
 * the cell's centre against the cell's boundary faces (right < 0) in face
 * order; a hit flips `cross`; a wall face (-1) sets `wall`, a far-field
 * face (-2) sets `far`, and either ends that cell's face loop. */
struct BoundaryTest {
    int cross = 0;
    bool wall = false, far = false;
};
inline BoundaryTest boundary_test(const MeshData& M, const double* testp, const std::vector<int64_t>& cand)
{
    BoundaryTest r;
    for (int64_t idx : cand) {
        const double* rayp = &M.cCentre[size_t(idx) * size_t(M.dim)];
        for (int64_t k = M.cell_off[size_t(idx)]; k < M.cell_off[size_t(idx) + 1]; ++k) {
            const int64_t f = M.cell_f[size_t(k)];
            const int64_t right = M.face_r[size_t(f)];
            if (right >= 0) continue;
            const int ints = M.dim == 3 ? crossings3d(M, f, testp, rayp) : line_intersection(M, f, testp, rayp);
            if (!ints) continue;
            r.cross = !r.cross;
            if (right == -1) {
                r.wall = true;
                break;
            } else if (right == -2) {
                r.far = true;
                break;
            }
        }
    }
    return r;
}

/* FirstCell's branch for a particle leaving the pipe in none of its
 * nearest cells (src/Containment.cpp:466-573): for each listed cell in
 * order, the segment from the particle's position before this test to the
 * cell's centre against each of the cell's boundary faces in face order.
 * Every hit flips `cross` (no face loop ends early); a wall face (-1)
 * reflects the velocity off the face and moves the particle onto the
 * face's plane (Eigen's cross product and normalized(), the plane through
 * the face's second vertex); a far-field face (-2) marks the particle for
 * deletion. */
struct FirstCellResult {
    int cross = 0;
    bool del = false;
};
template <int D>
inline FirstCellResult first_cell_boundary(const MeshData& M, const std::vector<int64_t>& cand, Vec<D>& xi,
                                           Vec<D>& v)
{
    FirstCellResult r;
    Vec<D> testp = xi;
    for (int64_t idx : cand) {
        const double* rayp = &M.cCentre[size_t(idx) * size_t(M.dim)];
        for (int64_t k = M.cell_off[size_t(idx)]; k < M.cell_off[size_t(idx) + 1]; ++k) {
            const int64_t f = M.cell_f[size_t(k)];
            const int64_t right = M.face_r[size_t(f)];
            if (right >= 0) continue;
            const int ints = D == 3 ? crossings3d(M, f, testp.c, rayp) : line_intersection(M, f, testp.c, rayp);
            if (!ints) continue;
            r.cross = !r.cross;
            if (right == -1) {
                const int64_t* fv = &M.face_v[size_t(M.face_off[size_t(f)])];
                Vec<D> nrm;
                if (D == 3) {
                    double r1[3], r2[3];
                    for (int a = 0; a < 3; ++a) {
                        r1[a] = M.vert(fv[1])[a] - M.vert(fv[0])[a];
                        r2[a] = M.vert(fv[2])[a] - M.vert(fv[0])[a];
                    }
                    nrm.c[0] = r1[1] * r2[2] - r1[2] * r2[1];
                    nrm.c[1] = r1[2] * r2[0] - r1[0] * r2[2];
                    nrm.c[D - 1] = r1[0] * r2[1] - r1[1] * r2[0];
                } else {
                    const double r0 = M.vert(fv[1])[0] - M.vert(fv[0])[0];
                    const double r1 = M.vert(fv[1])[1] - M.vert(fv[0])[1];
                    nrm.c[0] = -r1;
                    nrm.c[1] = r0;
                }
                nrm = normalized(nrm);
                v = v - (2 * dot(v, nrm)) * nrm;
                Vec<D> p1;
                for (int a = 0; a < D; ++a) p1.c[a] = M.vert(fv[1])[a];
                const double plane = dot(nrm, p1);
                const double dist = (plane - dot(xi, nrm)) / dot(nrm, nrm);
                xi = xi + dist * nrm;
            } else if (right == -2) {
                r.del = true;
            }
        }
    }
    return r;
}

#include "ipt_oracle.hpp"

/* MESH::maxlength as the TAU readers set it (Place_Edges, Place_Faces,
 * src/CDFIO.cpp:880-929, 1130-1220): the longest 2D edge, or the longest
 * triangle edge / longer diagonal of the other faces in 3D, times 4 (2D) or
 * 5 (3D); computed here for every mesh. */
inline double mesh_max_length(const MeshData& M)
{
    const int D = M.dim;
    auto dist = [&](int64_t a, int64_t b) {
        double s = 0.0;
        for (int k = 0; k < D; ++k) {
            const double d = M.vert(a)[k] - M.vert(b)[k];
            s += d * d;
        }
        return std::sqrt(s);
    };
    double mx = 0.0;
    for (size_t f = 0; f + 1 < M.face_off.size(); ++f) {
        const int64_t* fv = &M.face_v[size_t(M.face_off[f])];
        const int64_t nv = M.face_off[f + 1] - M.face_off[f];
        double e = 0.0;
        if (D == 2) e = dist(fv[0], fv[1]);
        else if (nv == 3) e = std::max(dist(fv[0], fv[1]), std::max(dist(fv[0], fv[2]), dist(fv[1], fv[2])));
        else if (nv >= 4) e = std::max(dist(fv[0], fv[2]), dist(fv[1], fv[3]));
        mx = std::max(mx, e);
    }
    return mx * (D == 2 ? 4.0 : 5.0);
}

template <int D>
struct Sim {
    fjsph_settings s;
    fjsph_derived d;
    /* implicit particle tracking (IPT_SETT) and the ends of the tracks */
    fjsph_ipt_settings ipt{0, 2, 9999999.0, 0.6, 5, 1000000};
    std::vector<IptRecord> ipt_records;
    int64_t ipt_success = 0, ipt_failed = 0;

    IptParams ipt_params() const
    {
        IptParams p{};
        for (int a = 0; a < 3; ++a) p.grav[a] = s.grav[a];
        p.mu_g = s.mu_g;
        p.rho_l = s.rho_rest;
        p.relax = ipt.relax;
        p.max_x = ipt.max_x;
        p.max_len = mesh_max_length(mesh);
        /* Set_Values (src/IO.cpp:126-127) */
        p.diam = std::pow((6.0 * d.sim_mass) / (3.14159265358979323846 * s.rho_rest), 1.0 / 3.0);
        p.area = 3.14159265358979323846 * p.diam * p.diam / 4.0;
        p.eq_order = ipt.eq_order;
        p.n_relax = ipt.n_relax;
        p.max_subits = s.max_subits;
        p.max_steps = ipt.max_steps;
        return p;
    }
    std::vector<Part<D>> pn, pnp1;
    OUTL outlist;
    std::vector<Block> blocks;
    size_t bound_points = 0, total_points = 0;
    size_t max_points = 9999999; /* svar.max_points (src/Var.h:369) */
    int nthreads = 1;
    int serial_dissipation = 1; /* src/Shifting.cpp:131 has no pragma */

    /* INTEG_SETT state */
    double current_time = 0.0, last_frame_time = 0.0, delta_t = 2e-10, cfl = 1.0;
    unsigned n_stable = 0, n_unstable = 0;
    unsigned current_frame = 0; /* INTEG_SETT::current_frame (src/Var.h:71) */
    size_t part_id_next = 0;    /* svar.part_id (src/Var.h:365), the next particle id */
    /* Integrator members (src/Integration.h:52-67) */
    double safe_dt = 0.0;
    double maxf = kEps, maxAf = kEps, maxRho_pc = kEps, maxRhoi = kEps, maxdrho = kEps, minST = 9999999.0, maxU = kEps, maxShift = kEps;
    size_t start_index = 0, end_index = 0;
    unsigned iteration = 0;
    int rollbacks = 0;
    double last_npd = 1.0;
    std::vector<Vec<D>> xih;
    MeshData mesh;
    bool have_mesh = false;
    int64_t mesh_lost = 0; /* aero particles in no cell at the last FindCell */
    int64_t mesh_deleted = 0, mesh_internal = 0, mesh_failed = 0; /* boundary-test outcomes since upload */
    int64_t first_cell_reflected = 0, first_cell_deleted = 0; /* FirstCell's boundary branch since upload */

    /* FindCell (src/Containment.cpp:579-820) for the aero particles of pnp1
     * (FREE with lam_nb <= lam_cutoff; the others get no cell and keep their
     * cell values): the previous cell (ipt_n_failed reset), then the 5 and
     * the 500 nearest cell centres (ipt_n_failed and internal reset).  A
     * particle whose previous cell is unset (negative) skips the first test.
     * With `classify`, a particle in no cell gets the boundary test: a wall
     * crossing sets internal, a far-field crossing condemns it, and with no
     * net crossing ipt_n_failed counts up or, past 10, condemns it.  Returns
     * the condemned indices, ascending, each once (the reference's list can
     * hold an index twice). */
    std::vector<size_t> find_cell(size_t start, size_t end, bool classify)
    {
        int64_t lost = 0, nint = 0, nfail = 0;
        std::vector<char> del(end, 0);
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : lost, nint, nfail)
        for (long ii = long(start); ii < long(end); ++ii) {
            Part<D>& q = pnp1[ii];
            if (q.b != FJSPH_FREE || q.lam_nb > s.lam_cutoff) {
                q.cellID = kNoCell;
                continue;
            }
            const double* testp = q.xi.c;
            auto take = [&](int64_t cell) {
                q.cellID = long(cell);
                for (int a = 0; a < D; ++a) q.cellV.c[a] = mesh.cVel[size_t(cell) * D + a];
                q.cellP = mesh.cP[size_t(cell)];
                q.cellRho = mesh.cRho[size_t(cell)];
            };
            if (q.cellID >= 0 && size_t(q.cellID) < mesh.ncells() && check_cell(mesh, q.cellID, testp)) {
                take(q.cellID);
                q.n_failed = 0;
                continue;
            }
            bool found = false;
            for (size_t k : {size_t(5), size_t(500)}) {
                for (int64_t cell : knn_cells(mesh, testp, k))
                    if (check_cell(mesh, cell, testp)) {
                        take(cell);
                        q.n_failed = 0;
                        q.internal = 0;
                        found = true;
                        break;
                    }
                if (found) break;
            }
            if (found) continue;
            ++lost;
            if (!classify) continue;
            const BoundaryTest r = boundary_test(mesh, testp, knn_cells_padded(mesh, testp, 500));
            if (r.wall) {
                q.internal = 1;
                ++nint;
            }
            bool condemned = r.far;
            if (r.cross == 0) {
                ++nfail;
                if (q.n_failed > 10) condemned = true;
                else q.n_failed++;
            }
            if (condemned) del[size_t(ii)] = 1;
        }
        mesh_lost = lost;
        mesh_internal += nint;
        mesh_failed += nfail;
        std::vector<size_t> out;
        for (size_t ii = start; ii < end; ++ii)
            if (del[ii]) out.push_back(ii);
        return out;
    }

    /* get_aero_velocity's deletion (src/Resid.cpp:483-519): the condemned
     * particles leave pn and pnp1, then the neighbour lists and
     * dSPH_PreStep are made again.  Block ranges and inlet index lists
     * shift by the deletions below them (the reference leaves the ranges
     * and shifts the lists by all of them). */
    void delete_lost(const std::vector<size_t>& del, double& npd)
    {
        if (del.empty()) return;
        for (auto it = del.rbegin(); it != del.rend(); ++it) {
            pnp1.erase(pnp1.begin() + long(*it));
            pn.erase(pn.begin() + long(*it));
        }
        auto below = [&](size_t x) { return size_t(std::lower_bound(del.begin(), del.end(), x) - del.begin()); };
        for (Block& bl : blocks) {
            const int64_t f = bl.b.first, l = bl.b.last;
            bl.b.first -= int64_t(below(size_t(f)));
            bl.b.last -= int64_t(below(size_t(l)));
            for (size_t& x : bl.back) x -= below(x);
            for (auto& col : bl.buffer)
                for (size_t& x : col) x -= below(x);
        }
        total_points -= del.size();
        end_index -= del.size();
        mesh_deleted += int64_t(del.size());
        update_neighbours();
        dSPH_PreStep(total_points, npd);
    }

    /* -------------------------------------------------------- neighbours */
    void update_neighbours() { outlist = find_neighbours<D>(pnp1, d.sr, nthreads); }

    /* ---------------------------------------------- dSPH_PreStep (a3)
     * src/Shifting.cpp:12-123 */
    void dSPH_PreStep(size_t end, double& npd)
    {
        std::vector<double> npd_part(end, 0.0);
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = 0; ii < long(end); ++ii) {
            Mat<D> Lmat_ = Mat<D>::zero(), Lmat_nb_ = Mat<D>::zero();
            Vec<D> gradRho_ = Vec<D>::zero(), norm_ = Vec<D>::zero();
            double kernsum_ = 0.0, colour_ = 0.0, npd_i = 0.0;
            const Part<D>& pi = pnp1[ii];
            for (const auto& jj : outlist[ii]) {
                if (int64_t(ii) == jj.first) { kernsum_ += d.W_correc; continue; }
                const Part<D>& pj = pnp1[jj.first];
                const Vec<D> Rji = pj.xi - pi.xi;
                const double r = std::sqrt(jj.second);
                const double volj = pj.m / pj.rho;
                const double kern = Kernel(r, d.H, d.W_correc);
                const Vec<D> Grad = GradK<D>(-Rji, r, d.H, d.W_correc);
                for (int a = 0; a < D; ++a)
                    for (int b2 = 0; b2 < D; ++b2) Lmat_.a[a][b2] -= volj * Rji.c[a] * Grad.c[b2];
                gradRho_ += volj * (pj.rho - pi.rho) * Grad;
                if (pj.b > FJSPH_PISTON) {
                    for (int a = 0; a < D; ++a)
                        for (int b2 = 0; b2 < D; ++b2) Lmat_nb_.a[a][b2] -= volj * Rji.c[a] * Grad.c[b2];
                    norm_ += volj * Grad;
                    kernsum_ += kern;
                    colour_ += volj * kern;
                    npd_i += kern;
                }
            }
            npd_part[ii] = npd_i;
            Mat<D> Linv = Mat<D>::identity();
            Mat<D> tmp;
            if (colpiv_qr_inverse<D>(Lmat_, tmp)) Linv = tmp;
            gradRho_ = Linv * gradRho_;
            Part<D>& po = pnp1[ii];
            if (pi.b == FJSPH_BOUND) {
                po.lam = 1.0;
                po.lam_nb = 1.0;
            } else {
                po.lam = min_eig_direct(Lmat_);
                po.lam_nb = min_eig_direct(Lmat_nb_);
            }
            po.L = Linv;
            po.gradRho = gradRho_;
            po.norm = Linv * norm_;
            if (po.lam > 0.7) po.colourG = 2;
            else po.colourG = 2.0 * std::max(1.0, 1.0 / (2.0 * colour_));
            po.kernsum = kernsum_;
            po.colour = colour_;
        }
        double sum = 0.0;
        for (size_t ii = 0; ii < end; ++ii) sum += npd_part[ii];
        npd = sum / double(end);
    }

    /* ------------------------------------- get_aero_velocity (constVel)
     * src/Resid.cpp:569-611 */
    void get_aero_velocity(size_t start, size_t end, double& npd)
    {
        if (s.aero_source == FJSPH_MESH_INFL) {
            if (have_mesh) delete_lost(find_cell(start, end, true), npd);
            return;
        }
        if (s.aero_source != FJSPH_CONST_VEL) return;
        Vec<D> vinf;
        for (int k = 0; k < D; ++k) vinf.c[k] = s.v_inf[k];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ii++) {
            bool cond = s.use_lam ? (pnp1[ii].lam_nb < s.lam_cutoff && pnp1[ii].b == FJSPH_FREE)
                                  : (double(outlist[ii].size()) * d.i_n_full < s.lam_cutoff && pnp1[ii].b == FJSPH_FREE);
            if (cond) { pnp1[ii].cellV = vinf; pnp1[ii].cellID = 1; }
            else { pnp1[ii].cellV = Vec<D>::zero(); pnp1[ii].cellID = kNoCell; }
        }
    }

    /* ------------------------------------------- Detect_Surface (a4)
     * src/Geometry.cpp:14-280 */
    void Detect_Surface(size_t start, size_t end)
    {
        std::vector<Vec<D>> norms(end, Vec<D>::zero());
        const double h = 1.33 * s.particle_step;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            Vec<D> nrm = Vec<D>::zero();
            Part<D>& pi = pnp1[ii];
            if (pi.b < FJSPH_PIPE) {
                pi.surf = 0;
                pi.woccl = 1;
            } else if (pi.lam_nb < 0.2) {
                pi.surf = 1;
            } else if (pi.lam_nb < 0.75) {
                const Vec<D> xi = pi.xi;
                const Vec<D> pointT = xi + h * normalized(pi.norm);
                uint32_t surf = 1;
                for (const auto& jj : outlist[ii]) {
                    if (jj.first == ii) continue;
                    const Part<D>& pj = pnp1[jj.first];
                    const Vec<D> Rij = pi.xi - pj.xi;
                    const Vec<D> x_jT = pj.xi - pointT;
                    const double r = std::sqrt(jj.second);
                    if (r >= std::sqrt(2.0) * h) {
                        if (norm(x_jT) < h) { surf = 0; break; }
                    } else {
                        if (D == 2) {
                            Vec<D> tau;
                            tau.c[0] = pi.norm.c[1];
                            tau.c[1] = -pi.norm.c[0];
                            if ((std::fabs(dot(normalized(pi.norm), x_jT)) + std::fabs(dot(normalized(tau), x_jT))) < h) { surf = 0; break; }
                        } else {
                            if (std::acos(dot(normalized(pi.norm), (-Rij) / r)) < kPi / 4.0) { surf = 0; break; }
                        }
                    }
                }
                pi.surf = surf;
            } else {
                pi.surf = 0;
            }
            if (pi.lam > 0.7) {
                for (const auto& jj : outlist[ii]) {
                    if (jj.first == ii) continue;
                    const Part<D>& pj = pnp1[jj.first];
                    const Vec<D> Rij = pi.xi - pj.xi;
                    const double r = std::sqrt(jj.second);
                    const double volj = pj.m / pj.rho;
                    nrm += volj * (pj.lam - pi.lam) * (GradK<D>(Rij, r, d.H, d.W_correc));
                }
            } else {
                for (const auto& jj : outlist[ii]) {
                    if (jj.first == ii) continue;
                    const Part<D>& pj = pnp1[jj.first];
                    const Vec<D> Rij = pi.xi - pj.xi;
                    const double r = std::sqrt(jj.second);
                    const double volj = pj.m / pj.rho;
                    nrm += volj * pj.lam * (GradK<D>(Rij, r, d.H, d.W_correc));
                }
            }
            nrm = pi.L * nrm;
            if (norm(nrm) > 0.1 * pi.lam / d.H) norms[ii] = normalized(nrm);
        }
        Vec<D> vinf;
        for (int k = 0; k < D; ++k) vinf.c[k] = s.v_inf[k];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            double woccl_ = 0.0;
            Vec<D> Vdiff = Vec<D>::zero();
            Part<D>& pi = pnp1[ii];
            if (pi.cellID != -1 && pi.b == FJSPH_FREE && (s.aero_case == FJSPH_GISSLER)) {
                if (s.aero_source == FJSPH_MESH_INFL) Vdiff = pi.cellV - pi.v;
                else Vdiff = vinf - pi.v;
            }
            double curve = 0.0;
            for (const auto& jj : outlist[ii]) {
                if (jj.first == ii) continue;
                const Part<D>& pj = pnp1[jj.first];
                const Vec<D> Rij = pj.xi - pi.xi;
                const double r = std::sqrt(jj.second);
                const double volj = pj.m / pj.rho;
                const Vec<D>& nj = (size_t(jj.first) < end) ? norms[jj.first] : zero_;
                if (norm(norms[ii]) > 0 && norm(nj) > 0) {
                    curve += volj * dot(pi.L * (nj - norms[ii]), GradK<D>(Rij, r, d.H, d.W_correc));
                }
                if (pi.b == FJSPH_FREE && s.aero_case == FJSPH_GISSLER) {
                    const double frac = -dot(Rij, Vdiff) / (norm(Vdiff) * r);
                    if (frac > woccl_) woccl_ = frac;
                }
            }
            if (pi.lam_nb < s.lam_cutoff) pi.woccl = std::max(0.0, std::min(woccl_, 1.0));
            else pi.woccl = 1.0;
            pi.norm = norms[ii];
            pi.curve = curve;
            pi.norm_curve = d.dx * curve;
            pi.pDist = pi.lam;
        }
        if (s.ale) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (long ii = long(start); ii < long(end); ++ii) {
                pnp1[ii].surfzone = 0;
                for (const auto& jj : outlist[ii]) {
                    if (pnp1[jj.first].surf == 1) { pnp1[ii].surfzone = 1; break; }
                }
            }
        }
    }
    Vec<D> zero_ = Vec<D>::zero();

    /* --------------------------------------- dissipation_terms (a5)
     * src/Shifting.cpp:126-186 (serial in the reference) */
    void dissipation_terms(size_t start, size_t end)
    {
        auto body = [&](long ii) {
            const Part<D>& pi = pnp1[ii];
            Vec<D> artViscI = Vec<D>::zero();
            double Rrhod = 0.0;
            for (const auto& jj : outlist[ii]) {
                if (jj.first == ii) continue;
                const Part<D>& pj = pnp1[jj.first];
                const Vec<D> Rji = pj.xi - pi.xi;
                const Vec<D> Vji = pj.v - pi.v;
                const double rr = jj.second;
                const double r = std::sqrt(rr);
                const double idist2 = 1.0 / (rr + 0.0001 * d.H_sq);
                const double volj = pj.m / pj.rho;
                const Vec<D> gradK = GradK<D>(Rji, r, d.H, d.W_correc);
                if (pj.b > FJSPH_PISTON) artViscI += pj.m * ArtVisc<D>(s, d, pi, pj, Rji, Vji, idist2, gradK);
                if (pj.b > FJSPH_PISTON) Rrhod += Continuity_dSPH_grad<D>(Rji, idist2, gradK, volj, pi.gradRho, pj.gradRho, pi, pj);
                else Rrhod += Continuity_dSPH_plain<D>(Rji, idist2, gradK, volj, pi, pj);
            }
            pnp1[ii].aVisc = artViscI;
            pnp1[ii].deltaD = d.dsph_cont * Rrhod;
        };
        if (serial_dissipation) {
            for (long ii = long(start); ii < long(end); ++ii) body(ii);
        } else {
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (long ii = long(start); ii < long(end); ++ii) body(ii);
        }
    }

    /* ------------------------------------------- particle_shift (a6)
     * src/Shifting.cpp:189-290 (ALE build) */
    void particle_shift(size_t start, size_t end)
    {
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            if (pnp1[ii].lam_nb < 0.55 || pnp1[ii].b == FJSPH_BUFFER) { pnp1[ii].vPert = Vec<D>::zero(); continue; }
            const Part<D>& pi = pnp1[ii];
            Vec<D> deltaU = Vec<D>::zero();
            double maxUij = 0.0;
            double woccl = 0.0;
            for (const auto& jj : outlist[ii]) {
                if (jj.first == ii) continue;
                const Part<D>& pj = pnp1[jj.first];
                const Vec<D> Rji = pj.xi - pi.xi;
                const double r = std::sqrt(jj.second);
                const double volj = pj.m / pj.rho;
                const double kern = Kernel(r, d.H, d.W_correc);
                const Vec<D> gradK = GradK<D>(Rji, r, d.H, d.W_correc);
                deltaU += (1.0 + 0.2 * std::pow(kern / d.W_dx, 4.0)) * gradK * volj;
                if (pj.b > FJSPH_PISTON) {
                    const double theta = std::acos(dot(normalized(pi.norm), normalized(pj.norm)));
                    if (theta > woccl) woccl = theta;
                }
                const double du = norm(pj.v - pi.v);
                if (du > maxUij) maxUij = du;
            }
            deltaU = deltaU * (-2.0 * d.H * norm(pi.v));
            deltaU = std::min(norm(deltaU), std::min(maxUij / 2.0, s.max_shift_vel)) * normalized(deltaU);
            const Vec<D> nrm = -normalized(pnp1[ii].norm);
            if (pi.surfzone == 0 && pnp1[ii].lam_nb > 0.55) {
                pnp1[ii].vPert = deltaU;
            } else {
                if (woccl < kPi / 12.0) {
                    Vec<D> out;
                    for (int a = 0; a < D; ++a) {
                        double sacc = 0.0;
                        for (int b2 = 0; b2 < D; ++b2) {
                            const double shift = ((a == b2) ? 1.0 : 0.0) - nrm.c[a] * nrm.c[b2];
                            sacc += shift * deltaU.c[b2];
                        }
                        out.c[a] = sacc;
                    }
                    pnp1[ii].vPert = out;
                } else {
                    pnp1[ii].vPert = Vec<D>::zero();
                }
            }
        }
    }

    /* ------------------------------------------- boundary solvers (a9) */
    /* Get_Boundary_Pressure (src/Resid.cpp:21-76) */
    void Get_Boundary_Pressure(size_t start, size_t end, std::vector<Part<D>>& st)
    {
        std::vector<double> pressure(end - start, 0.0);
        Vec<D> grav;
        for (int k = 0; k < D; ++k) grav.c[k] = s.grav[k];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            const Part<D>& pi = st[ii];
            int isNearSurface = 0;
            double kernsum = 0.0, pkern = 0.0;
            Vec<D> acckern = Vec<D>::zero();
            for (const auto& jj : outlist[ii]) {
                const Part<D>& pj = st[jj.first];
                if (pj.b > FJSPH_PISTON) {
                    const Vec<D> Rji = pi.xi - pj.xi;
                    const double rr = jj.second;
                    const double r = std::sqrt(rr);
                    const double volj = pj.m / pj.rho;
                    const double kern = volj * Kernel(r, d.H, d.W_correc);
                    kernsum += kern;
                    pkern += pj.p * kern;
                    acckern += kern * pj.rho * Rji;
                    if (pj.surfzone) isNearSurface = 1;
                }
            }
            if (kernsum > 0.0) {
                const double p = (pkern + dot(grav - pi.acc, acckern)) / kernsum;
                if (isNearSurface) pressure[ii - start] = std::max(0.0, p);
                else pressure[ii - start] = p;
            }
        }
        for (size_t ii = start; ii < end; ++ii) {
            st[ii].p = pressure[ii - start];
            st[ii].rho = get_density(s, d, pressure[ii - start]);
        }
    }
    /* Boundary_DBC (src/Resid.cpp:78-117): every write of the reference
     * lands at a fluid index beyond the block (out of bounds), so the
     * defined effect is acc = 0 for the block. */
    void Boundary_DBC(size_t start, size_t end, std::vector<Part<D>>& st)
    {
        for (size_t ii = start; ii < end; ++ii) st[ii].acc = Vec<D>::zero();
    }
    /* Boundary_Ghost (src/Resid.cpp:122-160) */
    void Boundary_Ghost(size_t start, size_t end, std::vector<Part<D>>& st, std::vector<int>& near_inlet)
    {
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            const Part<D>& pi = st[ii];
            double Rrhoi = 0.0;
            near_inlet[ii] = 1;
            for (const auto& jj : outlist[ii]) {
                const Part<D>& pj = st[jj.first];
                if (ii == jj.first) continue;
                if (pj.b == FJSPH_PIPE || pj.b == FJSPH_FREE) near_inlet[ii] = 0;
                const Vec<D> Rji = pj.xi - pi.xi;
                const Vec<D> Vji = pj.v - pi.v;
                const double rr = jj.second;
                const double r = std::sqrt(rr);
                const double volj = pj.m / pj.rho;
                const Vec<D> gradK = GradK<D>(Rji, r, d.H, d.W_correc);
                Rrhoi -= volj * dot(Vji, gradK);
            }
            st[ii].Rrho = Rrhoi * pi.rho;
        }
    }
    /* Set_No_Slip (src/Resid.cpp:162-186) */
    void Set_No_Slip(size_t start, size_t end, std::vector<Part<D>>& st)
    {
        for (size_t ii = start; ii < end; ++ii) {
            Vec<D> velsum = Vec<D>::zero();
            double kernsum = 0.0;
            for (const auto& jj : outlist[ii]) {
                const Part<D>& pj = st[jj.first];
                if (int64_t(ii) == jj.first || pj.b <= FJSPH_PISTON) continue;
                const double r = std::sqrt(jj.second);
                const double kern = Kernel(r, d.H, d.W_correc);
                kernsum += kern;
                velsum += pj.v * kern;
            }
            if (kernsum > 0.0) st[ii].v = 2.0 * st[ii].v - velsum / kernsum;
        }
    }

    /* ----------------------------------------- get_acc_and_Rrho (a7, a8)
     * src/Resid.cpp:243-469 (PAIRWISE surface tension, frozen dissipation) */
    void get_acc_and_Rrho_on_i(const std::vector<std::pair<int64_t, double>>& outlist_i, const std::vector<Part<D>>& st,
                               double npdm2, double pi3o4, const Vec<D>& grav_vec, const Part<D>& pi,
                               Vec<D>& acc_i, Vec<D>& acc_aero_i, double& Rrho_i)
    {
        Vec<D> acc_ = Vec<D>::zero(), acc_ale_ = Vec<D>::zero();
        double Rrhoc_ = 0.0, Rrho_ = 0.0;
        Vec<D> visc_ = Vec<D>::zero(), surf_t_ = Vec<D>::zero();
        if (pi.cellID != -1) {
            const Vec<D> V_diff = pi.cellV - pi.v;
            const Vec<D> aero = CalcAeroAcc<D>(s, d, pi, V_diff, pi.norm, pi.lam_nb, double(outlist_i.size()));
            acc_ += aero;
            acc_aero_i = aero;
        }
        for (const auto& jj : outlist_i) {
            const Part<D>& pj = st[jj.first];
            if (pi.part_id == pj.part_id) continue;
            const Vec<D> Rji = pj.xi - pi.xi;
            const Vec<D> Vji = pj.v - pi.v;
            const double rr = jj.second;
            const double r = std::sqrt(rr);
            const double idist2 = 1.0 / (rr + 0.001 * d.H_sq);
            const double volj = pj.m / pj.rho;
            const Vec<D> gradK = GradK<D>(Rji, r, d.H, d.W_correc);
            const Vec<D> contrib = BasePos<D>(pi, pj, gradK);
            const Vec<D> visc = Viscosity<D>(d.nu, pi, pj, Rji, Vji, idist2, gradK);
            visc_ += pj.m * visc;
            if (!s.ale || pi.surfzone == 1) surf_t_ += pairwise_ST<D>(Rji, r, d.H, npdm2, pi3o4, pi.b, pj.b);
            acc_ -= contrib;
            if (s.ale) {
                acc_ale_ += ALEMomentum<D>(pi, pj, volj, gradK);
                Rrho_ -= ALEContinuity<D>(pi, pj, volj, gradK);
                Rrhoc_ += ALECont2ndterm<D>(pi, pj, volj, gradK);
            } else {
                Rrho_ -= volj * dot(Vji, gradK);
            }
        }
        if (s.ale) {
            acc_i = (acc_ + acc_ale_ + pi.aVisc + visc_ + surf_t_ / pi.m + grav_vec);
            Rrho_i = Rrho_ * pi.rho + Rrhoc_ + pi.deltaD;
        } else {
            acc_i = (acc_ + pi.aVisc + visc_ + surf_t_ / pi.m + grav_vec);
            Rrho_i = Rrho_ * pi.rho + pi.deltaD;
        }
    }

    void get_acc_and_Rrho(double npd, std::vector<Part<D>>& st)
    {
        const size_t start = bound_points, end = total_points;
        const double lam = d.lam_st;
        const double npdm2 = (0.5 * s.sig / lam) / (npd * npd);
        const double pi3o4 = 3.0 * kPi / 4.0;
        Vec<D> grav;
        for (int k = 0; k < D; ++k) grav.c[k] = s.grav[k];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (long ii = long(start); ii < long(end); ++ii) {
            const Part<D> pi = st[ii];
            Vec<D> acc, Af = st[ii].Af;
            double Rrho;
            get_acc_and_Rrho_on_i(outlist[ii], st, npdm2, pi3o4, grav, pi, acc, Af, Rrho);
            st[ii].acc = acc;
            st[ii].Af = Af;
            st[ii].Rrho = Rrho;
        }
    }

    /* -------------------------------------------- Newmark-Beta (a10)
     * src/Newmark_Beta.cpp:10-331 */
    Vec<D> block_velocity(const Block& bl, bool rk) const
    {
        Vec<D> vel = Vec<D>::zero();
        if (bl.b.n_times != 0) {
            for (int t = 0; t < bl.b.n_times; t++) {
                bool cond = rk ? (bl.times[t] > current_time) : (current_time > bl.times[t]);
                if (cond) for (int k = 0; k < D; ++k) vel.c[k] = bl.vels[t * 3 + k];
            }
        } else {
            for (int k = 0; k < D; ++k) vel.c[k] = bl.b.vel[k];
        }
        return vel;
    }

    void boundary_update(std::vector<Part<D>>& st, std::vector<int>& near_inlet, bool rk)
    {
        for (const Block& bl : blocks) {
            if (bl.b.is_fluid) continue;
            const size_t f = size_t(bl.b.first), l = size_t(bl.b.last);
            const Vec<D> vel = block_velocity(bl, rk);
            for (size_t jj = f; jj < l; jj++) st[jj].v = vel;
            if (bl.b.no_slip) Set_No_Slip(f, l, st);
            switch (bl.b.bound_solver) {
            case FJSPH_DBC: Boundary_DBC(f, l, st); break;
            case FJSPH_PRESSURE_G: Get_Boundary_Pressure(f, l, st); break;
            case FJSPH_GHOST: Boundary_Ghost(f, l, st, near_inlet); break;
            default: break;
            }
        }
    }

    void Do_NB_Iter(double npd)
    {
        std::vector<int> near_inlet(bound_points, 0);
        const double gamma_t1 = 0.5, gamma_t2 = 1 - gamma_t1;
        const double beta_t1 = 0.25, beta_t2 = 0.5 * (1 - 2 * beta_t1);
        boundary_update(pnp1, near_inlet, false);
        get_acc_and_Rrho(npd, pnp1);
        const double dt = delta_t, dt2 = dt * dt;
        for (const Block& bl : blocks) {
            if (bl.b.is_fluid) continue;
            const size_t f = size_t(bl.b.first), l = size_t(bl.b.last);
            if (bl.b.bound_solver == FJSPH_DBC) {
                for (size_t ii = f; ii < l; ++ii) {
                    const double rho = std::max(d.rho_min, std::min(d.rho_max, pn[ii].rho + dt * (gamma_t1 * pnp1[ii].Rrho + gamma_t2 * pn[ii].Rrho)));
                    pnp1[ii].rho = rho;
                    pnp1[ii].p = get_pressure(s, d, rho);
                }
            } else if (bl.b.bound_solver == FJSPH_GHOST) {
                for (size_t ii = f; ii < l; ++ii) {
                    if (near_inlet[ii]) {
                        const double rho = std::max(s.rho_rest, std::min(d.rho_max, pn[ii].rho + dt * (gamma_t1 * pnp1[ii].Rrho + gamma_t2 * pn[ii].Rrho)));
                        pnp1[ii].rho = rho;
                        pnp1[ii].p = get_pressure(s, d, rho);
                        pnp1[ii].Rrho = std::fmax(0.0, pnp1[ii].Rrho);
                    } else {
                        const double rho = std::max(d.rho_min, std::min(d.rho_max, pn[ii].rho + dt * (gamma_t1 * pnp1[ii].Rrho + gamma_t2 * pn[ii].Rrho)));
                        pnp1[ii].rho = rho;
                        pnp1[ii].p = get_pressure(s, d, rho);
                    }
                }
            }
        }
        for (const Block& bl : blocks) {
            if (!bl.b.is_fluid) continue;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (long ii = long(bl.b.first); ii < long(bl.b.last); ++ii) {
                if (pnp1[ii].b > FJSPH_BUFFER && pnp1[ii].b != FJSPH_OUTLET) {
                    if (s.ale) {
                        pnp1[ii].xi = pn[ii].xi + dt * (pn[ii].v + pnp1[ii].vPert) + dt2 * (beta_t1 * pnp1[ii].acc + beta_t2 * pn[ii].acc);
                    } else {
                        pnp1[ii].xi = pn[ii].xi + dt * pn[ii].v + dt2 * (beta_t2 * pn[ii].acc + beta_t1 * pnp1[ii].acc);
                    }
                    pnp1[ii].v = pn[ii].v + dt * (gamma_t1 * pnp1[ii].acc + gamma_t2 * pn[ii].acc);
                    const double rho = std::max(d.rho_min, std::min(d.rho_max, pn[ii].rho + dt * (gamma_t1 * pnp1[ii].Rrho + gamma_t2 * pn[ii].Rrho)));
                    pnp1[ii].rho = rho;
                    pnp1[ii].p = get_pressure(s, d, rho);
                } else if (pnp1[ii].b == FJSPH_OUTLET) {
                    pnp1[ii].xi = pn[ii].xi + dt * pnp1[ii].v;
                }
            }
            /* the inlet buffer (src/Newmark_Beta.cpp:243-292), after the
             * fluid update of the block */
            if (!bl.pl.inlet) continue;
            if (bl.pl.fixed_vel_or_dynamic == 1) {
                Vec<D> unorm;
                double len = 0.0;
                for (int a = 0; a < D; ++a) len += bl.pl.insert_norm[a] * bl.pl.insert_norm[a];
                len = std::sqrt(len);
                for (int a = 0; a < D; ++a) unorm[a] = bl.pl.insert_norm[a] / len;
                for (size_t c = 0; c < bl.back.size(); ++c) {
                    const Part<D>& bk = pnp1[bl.back[c]];
                    for (size_t jj = 0; jj < bl.buffer[c].size(); ++jj) {
                        Part<D>& q = pnp1[bl.buffer[c][jj]];
                        q.xi = bk.xi - (d.dx * (double(jj) + 1.0)) * unorm;
                        q.v = bk.v;
                        q.rho = bk.rho;
                        q.p = bk.p;
                    }
                }
            } else {
                for (size_t c = 0; c < bl.back.size(); ++c)
                    for (size_t jj = 0; jj < bl.buffer[c].size(); ++jj) {
                        const size_t id = bl.buffer[c][jj];
                        const double rho = std::max(d.rho_min, std::min(d.rho_max, pn[id].rho + dt * (gamma_t1 * pnp1[id].Rrho + gamma_t2 * pn[id].Rrho)));
                        pnp1[id].rho = rho;
                        pnp1[id].p = get_pressure(s, d, rho);
                        pnp1[id].xi = pn[id].xi + dt * pn[id].v;
                    }
            }
        }
    }

    double log_error_of(const std::vector<Vec<D>>& ref) const
    {
        const size_t start = start_index, end = end_index;
        double errsum = 0.0;
        for (size_t ii = start; ii < end; ++ii) errsum += sqnorm(pnp1[ii].xi - ref[ii - start]);
        return std::log10(std::sqrt(errsum / double(end - start)));
    }

    /* Check_Error (src/Newmark_Beta.cpp:10-52) */
    int Check_Error(double& rms_error, double& logbase)
    {
        const double log_error = log_error_of(xih);
        if (iteration == 0) logbase = log_error;
        rms_error = log_error - logbase;
        if (iteration > unsigned(s.max_subits)) {
            if (rms_error > 0.0) {
                pnp1 = pn;
                update_neighbours();
                delta_t = 0.5 * delta_t;
                iteration = 0;
                rms_error = 0.0;
                rollbacks++;
                return -1;
            }
            return 1;
        }
        return 0;
    }

    void save_xih()
    {
        xih.resize(end_index - start_index);
        for (size_t ii = start_index; ii < end_index; ++ii) xih[ii - start_index] = pnp1[ii].xi;
    }

    double Newmark_Beta_loop(double npd, double& logbase)
    {
        double rms_error = 0.0;
        while (rms_error > s.min_residual) {
            save_xih();
            Do_NB_Iter(npd);
            const int errstate = Check_Error(rms_error, logbase);
            if (errstate == 0) iteration++;
            else if (errstate == 1) break;
        }
        return rms_error;
    }

    /* ----------------------------------------------- Runge-Kutta (a11)
     * src/Runge_Kutta.cpp:10-515 */
    double Check_RK_Error(double logbase, const std::vector<Part<D>>& a, const std::vector<Part<D>>& b) const
    {
        double errsum = 0.0;
        for (size_t ii = start_index; ii < end_index; ++ii) errsum += sqnorm(b[ii].xi - a[ii].xi);
        return std::log10(std::sqrt(errsum / double(end_index - start_index))) - logbase;
    }

    void rk_boundary(std::vector<Part<D>>& st, const std::vector<Part<D>>& part_n, double dtf,
                     const std::vector<Part<D>>* s1, const std::vector<Part<D>>* s2, const std::vector<Part<D>>* s3)
    {
        for (const Block& bl : blocks) {
            if (bl.b.is_fluid) continue;
            const size_t f = size_t(bl.b.first), l = size_t(bl.b.last);
            const Vec<D> vel = block_velocity(bl, true);
            for (size_t jj = f; jj < l; jj++) st[jj].v = vel;
            if (bl.b.no_slip) Set_No_Slip(f, l, st);
            std::vector<int> near_inlet(bound_points, 0);
            switch (bl.b.bound_solver) {
            case FJSPH_DBC:
                Boundary_DBC(f, l, st);
                for (size_t ii = f; ii < l; ++ii) {
                    double drho = s1 ? (part_n[ii].Rrho + 2.0 * (*s1)[ii].Rrho + 2.0 * (*s2)[ii].Rrho + (*s3)[ii].Rrho) : st[ii].Rrho;
                    const double rho = std::max(d.rho_min, std::min(d.rho_max, part_n[ii].rho + (s1 ? (dtf / 6.0) : dtf) * drho));
                    st[ii].rho = rho;
                    st[ii].p = get_pressure(s, d, rho);
                    if (s1) st[ii].Rrho = (*s3)[ii].Rrho;
                }
                break;
            case FJSPH_PRESSURE_G: Get_Boundary_Pressure(f, l, st); break;
            case FJSPH_GHOST:
                Boundary_Ghost(f, l, st, near_inlet);
                for (size_t ii = f; ii < l; ++ii) {
                    double drho = s1 ? (part_n[ii].Rrho + 2.0 * (*s1)[ii].Rrho + 2.0 * (*s2)[ii].Rrho + (*s3)[ii].Rrho) : st[ii].Rrho;
                    const double fac = s1 ? (dtf / 6.0) : dtf;
                    if (!near_inlet[ii]) {
                        const double rho = std::max(d.rho_min, std::min(d.rho_max, part_n[ii].rho + fac * drho));
                        st[ii].rho = rho;
                        st[ii].p = get_pressure(s, d, rho);
                    } else {
                        const double rho = std::max(s.rho_rest, std::min(d.rho_max, part_n[ii].rho + fac * drho));
                        st[ii].rho = rho;
                        st[ii].p = get_pressure(s, d, rho);
                        st[ii].Rrho = std::fmax(0.0, st[ii].Rrho);
                    }
                }
                break;
            default: break;
            }
        }
    }

    /* The inlet buffer of a Runge-Kutta stage (src/Runge_Kutta.cpp:171-229,
     * 389-447), after the block's fluid update.  fixed_vel_or_dynamic = 1:
     * off the column's back particle of this stage.  Otherwise the density
     * from part_n and the stage's rate of the particle itself (dt Rrho, or
     * dt/6 of the four rates in the last stage, s1..s3 given) and the
     * position x_n + dt v_n.  The reference's branch for the latter reads the
     * buffer list of block `jj` (the buffer depth) and, in the last stage,
     * the rates of particle `ii` (the column number); the block's own list
     * and the particle's own rates are used here. */
    void rk_buffer(const Block& bl, std::vector<Part<D>>& st, const std::vector<Part<D>>& part_n, double dt,
                   const std::vector<Part<D>>* s1, const std::vector<Part<D>>* s2, const std::vector<Part<D>>* s3)
    {
        if (!bl.pl.inlet) return;
        if (bl.pl.fixed_vel_or_dynamic == 1) {
            Vec<D> unorm;
            double len = 0.0;
            for (int a = 0; a < D; ++a) len += bl.pl.insert_norm[a] * bl.pl.insert_norm[a];
            len = std::sqrt(len);
            for (int a = 0; a < D; ++a) unorm[a] = bl.pl.insert_norm[a] / len;
            for (size_t c = 0; c < bl.back.size(); ++c) {
                const Part<D>& bk = st[bl.back[c]];
                for (size_t jj = 0; jj < bl.buffer[c].size(); ++jj) {
                    Part<D>& q = st[bl.buffer[c][jj]];
                    q.xi = bk.xi - (d.dx * (double(jj) + 1.0)) * unorm;
                    q.v = bk.v;
                    q.rho = bk.rho;
                    q.p = bk.p;
                }
            }
            return;
        }
        for (size_t c = 0; c < bl.back.size(); ++c)
            for (size_t jj = 0; jj < bl.buffer[c].size(); ++jj) {
                const size_t id = bl.buffer[c][jj];
                const double dr = s1 ? (dt / 6.0) * (part_n[id].Rrho + 2.0 * (*s1)[id].Rrho + 2.0 * (*s2)[id].Rrho + (*s3)[id].Rrho)
                                     : dt * st[id].Rrho;
                const double rho = std::max(d.rho_min, std::min(d.rho_max, part_n[id].rho + dr));
                st[id].rho = rho;
                st[id].p = get_pressure(s, d, rho);
                st[id].xi = part_n[id].xi + dt * part_n[id].v;
            }
    }

    std::vector<Part<D>> rk_intermediate(double npd, const std::vector<Part<D>>& part_n, const std::vector<Part<D>>& part_prev, double dt_inter)
    {
        std::vector<Part<D>> part_inter = part_prev;
        rk_boundary(part_inter, part_n, dt_inter, nullptr, nullptr, nullptr);
        get_acc_and_Rrho(npd, part_inter);
        for (const Block& bl : blocks) {
            if (!bl.b.is_fluid) continue;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (long ii = long(bl.b.first); ii < long(bl.b.last); ++ii) {
                Part<D>& q = part_inter[ii];
                if (q.b > FJSPH_BUFFER) {
                    if (s.ale) q.xi = part_n[ii].xi + dt_inter * (q.v + q.vPert);
                    else q.xi = part_n[ii].xi + dt_inter * q.v;
                    q.v = part_n[ii].v + dt_inter * q.acc;
                    const double rho = std::max(d.rho_min, std::min(d.rho_max, part_n[ii].rho + dt_inter * q.Rrho));
                    q.rho = rho;
                    q.p = get_pressure(s, d, rho);
                }
            }
            rk_buffer(bl, part_inter, part_n, dt_inter, nullptr, nullptr, nullptr);
        }
        return part_inter;
    }

    std::vector<Part<D>> rk_final(double npd, double dt, const std::vector<Part<D>>& part_n, const std::vector<Part<D>>& st_1,
                                  const std::vector<Part<D>>& st_2, const std::vector<Part<D>>& st_3)
    {
        std::vector<Part<D>> part_np1 = st_3;
        rk_boundary(part_np1, part_n, dt, &st_1, &st_2, &st_3);
        get_acc_and_Rrho(npd, part_np1);
        for (const Block& bl : blocks) {
            if (!bl.b.is_fluid) continue;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (long ii = long(bl.b.first); ii < long(bl.b.last); ++ii) {
                Part<D>& q = part_np1[ii];
                if (q.b > FJSPH_BUFFER && q.b != FJSPH_OUTLET) {
                    if (s.ale)
                        q.xi = part_n[ii].xi + (dt / 6.0) * ((part_n[ii].v + part_n[ii].vPert) + 2.0 * (st_1[ii].v + st_1[ii].vPert) +
                                                             2.0 * (st_2[ii].v + st_2[ii].vPert) + (st_3[ii].v + st_3[ii].vPert));
                    else
                        q.xi = part_n[ii].xi + (dt / 6.0) * (part_n[ii].v + 2.0 * st_1[ii].v + 2.0 * st_2[ii].v + st_3[ii].v);
                    q.v = part_n[ii].v + (dt / 6.0) * (part_n[ii].acc + 2.0 * st_1[ii].acc + 2.0 * st_2[ii].acc + st_3[ii].acc);
                    const double rho = std::max(d.rho_min, std::min(d.rho_max, part_n[ii].rho + (dt / 6.0) * (part_n[ii].Rrho + 2.0 * st_1[ii].Rrho + 2.0 * st_2[ii].Rrho + st_3[ii].Rrho)));
                    q.rho = rho;
                    q.p = get_pressure(s, d, rho);
                } else if (q.b == FJSPH_OUTLET) {
                    q.xi = part_n[ii].xi + dt * q.v;
                }
            }
            rk_buffer(bl, part_np1, part_n, dt, &st_1, &st_2, &st_3);
        }
        return part_np1;
    }

    /* ------------------------------------------- find_timestep (a12)
     * src/Integration.cpp:370-443 */
    double find_timestep()
    {
        maxf = kEps; maxAf = kEps; maxRhoi = kEps; maxdrho = kEps; maxU = kEps; minST = 9999999.0; maxShift = kEps;
        for (size_t ii = start_index; ii < end_index; ++ii) {
            const Part<D>& q = pnp1[ii];
            maxf = std::max(maxf, norm(q.acc));
            maxAf = std::max(maxAf, norm(q.Af));
            maxdrho = std::max(maxdrho, std::fabs(q.Rrho));
            maxRhoi = std::max(maxRhoi, std::fabs(q.rho - s.rho_rest));
            const double den = (2.0 * kPi * s.sig * std::fabs(q.curve));
            if (den != 0.0) minST = std::min(minST, std::sqrt(q.rho * d.dx * d.dx / den));
            maxU = std::max(maxU, norm(q.v));
            if (s.ale) maxShift = std::max(maxShift, norm(q.vPert));
        }
        maxRho_pc = 100 * maxRhoi / s.rho_rest;
        double mn = 0.25 * std::sqrt(d.H / maxf);
        mn = std::min(mn, 2 * d.H / (maxU));
        if (s.mu != 0.0) mn = std::min(mn, 0.125 * d.H_sq * s.rho_rest / s.mu);
        mn = std::min(mn, 0.067 * minST);
        mn = std::min(mn, 0.5 * std::sqrt(d.H / maxdrho));
        mn = std::min(mn, 1.5 * d.H / s.speed_sound);
        safe_dt = 0.75 * mn;
        double dt = cfl * safe_dt;
        if (dt < s.delta_t_min) dt = s.delta_t_min;
        else if (dt > s.delta_t_max) dt = s.delta_t_max;
        if (dt > last_frame_time + s.frame_time_interval - current_time)
            dt = last_frame_time + s.frame_time_interval - current_time + s.delta_t_min;
        return dt;
    }

    /* Pre-passes shared by both halves of integrate_no_update. */
    void prepasses(double& npd)
    {
        dSPH_PreStep(total_points, npd);
        get_aero_velocity(start_index, end_index, npd);
        Detect_Surface(start_index, end_index);
        dissipation_terms(start_index, end_index);
        if (s.ale) particle_shift(start_index, end_index);
        Check_Pipe_Outlet(npd);
    }

    /* Check_Pipe_Outlet (src/Containment.cpp:822-890): PIPE particles past
     * the aero plane of their block become FREE; with the mesh source they
     * get their cell (FirstCell; the lookup of every aero particle finds
     * the cached cell of the others again).  The deletion of a particle
     * FirstCell places in no cell is not made (counted by the lookup). */
    void Check_Pipe_Outlet(double& npd)
    {
        bool any = false;
        std::vector<size_t> fresh;
        for (const Block& bl : blocks) {
            if (!bl.b.is_fluid) continue;
            for (size_t ii = size_t(bl.b.first); ii < size_t(bl.b.last); ++ii)
                if (pnp1[ii].b == FJSPH_PIPE && dot_raw<D>(pnp1[ii].xi, bl.pl.aero_norm) > bl.pl.aeroconst) {
                    pnp1[ii].b = FJSPH_FREE;
                    any = true;
                    if (pnp1[ii].lam_nb < s.lam_cutoff) fresh.push_back(ii);
                }
        }
        if (!(any && s.aero_source == FJSPH_MESH_INFL && have_mesh)) return;
        (void)find_cell(start_index, end_index, false);
        /* FirstCell for the new aero particles in no cell: the K nearest
         * cells (150 in 3D, 20 in 2D, padded with cell 0 as the reference's
         * result vector is), then the boundary branch */
        const size_t K = D == 3 ? 150 : 20;
        std::vector<size_t> del;
        for (size_t ii : fresh) {
            Part<D>& q = pnp1[ii];
            if (q.cellID >= 0 && size_t(q.cellID) < mesh.ncells() && check_cell(mesh, q.cellID, q.xi.c)) continue;
            const std::vector<int64_t> cand = knn_cells_padded(mesh, q.xi.c, K);
            bool found = false;
            for (int64_t cell : cand)
                if (check_cell(mesh, cell, q.xi.c)) {
                    q.cellID = long(cell);
                    for (int a = 0; a < D; ++a) q.cellV.c[a] = mesh.cVel[size_t(cell) * D + a];
                    q.cellP = mesh.cP[size_t(cell)];
                    q.cellRho = mesh.cRho[size_t(cell)];
                    found = true;
                    break;
                }
            if (found) continue;
            const FirstCellResult r = first_cell_boundary<D>(mesh, cand, q.xi, q.v);
            if (r.cross == 0)
                throw std::runtime_error("FirstCell: first containing cell for a particle leaving the pipe not found");
            first_cell_reflected += 1;
            if (r.del) del.push_back(ii);
        }
        /* the reference erases the condemned particles from pnp1 alone and
         * keeps its neighbour lists; here they leave both states and the
         * lists and dSPH_PreStep are made again, as after FindCell */
        first_cell_deleted += int64_t(del.size());
        delete_lost(del, npd);
    }

    /* integrate_no_update (src/Integration.cpp:27-107) */
    double integrate_no_update()
    {
        start_index = bound_points;
        end_index = total_points;
        iteration = 0;
        rollbacks = 0;
        double rms_error = 0.0, logbase = 0.0, npd = 1.0;
        delta_t = find_timestep();
        update_neighbours();
        prepasses(npd);
        save_xih();
        if (s.integrator == FJSPH_RUNGE_KUTTA) {
            std::vector<Part<D>> st1 = rk_intermediate(npd, pn, pn, 0.5 * delta_t);
            pnp1 = st1;
            logbase = Check_RK_Error(0.0, pn, pnp1);
        } else {
            Do_NB_Iter(npd);
            double rms_dummy = 0.0;
            (void)Check_Error(rms_dummy, logbase);
            iteration++;
        }
        update_neighbours();
        prepasses(npd);
        if (s.integrator == FJSPH_RUNGE_KUTTA) {
            std::vector<Part<D>> st_1 = pnp1;
            std::vector<Part<D>> st_2 = rk_intermediate(npd, pn, st_1, 0.5 * delta_t);
            std::vector<Part<D>> st_3 = rk_intermediate(npd, pn, st_2, delta_t);
            pnp1 = rk_final(npd, delta_t, pn, st_1, st_2, st_3);
            rms_error = Check_RK_Error(logbase, st_3, pnp1);
        } else {
            rms_error = Newmark_Beta_loop(npd, logbase);
        }
        last_npd = npd;
        return rms_error;
    }

    /* update_buffer_region (src/shapes/inlet.cpp:578-640): a back particle
     * past the insertion plane becomes PIPE, the first buffer particle
     * becomes the back one, the buffer moves up, and while under max_points
     * a new BUFFER particle is inserted at the block's end, made from the
     * last buffer particle as SPHPart(x, pj, BUFFER, id) does
     * (src/Var.h:550-592), dx behind it along the insertion normal. */
    size_t update_buffer_region()
    {
        size_t nAdd = 0;
        for (size_t bi = 0; bi < blocks.size(); ++bi) {
            Block& block = blocks[bi];
            if (!block.b.is_fluid || !block.pl.inlet) continue;
            for (size_t ii = 0; ii < block.back.size(); ++ii) {
                const size_t pID = block.back[ii];
                if (!(dot_raw<D>(pnp1[pID].xi, block.pl.insert_norm) > block.pl.insconst)) continue;
                pnp1[pID].b = FJSPH_PIPE;
                pnp1[block.buffer[ii][0]].b = FJSPH_BACK;
                block.back[ii] = block.buffer[ii][0];
                for (size_t jj = 0; jj + 1 < block.buffer[ii].size(); ++jj) block.buffer[ii][jj] = block.buffer[ii][jj + 1];
                if (total_points < max_points) {
                    const Part<D>& pj = pnp1[block.buffer[ii].back()];
                    Part<D> q;
                    for (int a = 0; a < D; ++a) q.xi[a] = pj.xi[a] - d.dx * block.pl.insert_norm[a];
                    q.v = pj.v;
                    q.rho = pj.rho;
                    q.p = pj.p;
                    q.m = pj.m;
                    q.cellP = pj.cellP;
                    q.cellRho = pj.cellRho;
                    q.cellID = kNoCell;
                    q.b = FJSPH_BUFFER;
                    q.part_id = int64_t(part_id_next++);
                    const size_t at = size_t(block.b.last);
                    pnp1.insert(pnp1.begin() + long(at), q);
                    block.buffer[ii].back() = at;
                    block.b.last++;
                    for (size_t jj = bi + 1; jj < blocks.size(); ++jj) {
                        blocks[jj].b.first++;
                        blocks[jj].b.last++;
                    }
                    total_points++;
                    end_index++;
                    nAdd++;
                }
            }
        }
        return nAdd;
    }

    /* update_data (src/Integration.cpp:109-226): inlets, density floor,
     * deletion plane (IPT conversion, out of scope, not made), pn = pnp1. */
    void update_data()
    {
        update_buffer_region();
        for (size_t ii = start_index; ii < end_index; ++ii)
            if (pnp1[ii].rho < 0.0001) pnp1[ii].rho = s.rho_rest;
        std::vector<size_t> to_del;
        std::vector<size_t> n_del(blocks.size(), 0);
        for (size_t bi = 0; bi < blocks.size(); ++bi) {
            const Block& bl = blocks[bi];
            if (!bl.b.is_fluid || bl.pl.delconst == FJSPH_NO_PLANE) continue;
            for (size_t ii = size_t(bl.b.first); ii < size_t(bl.b.last); ++ii)
                if (dot_raw<D>(pnp1[ii].xi, bl.pl.delete_norm) > bl.pl.delconst) {
                    to_del.push_back(ii);
                    n_del[bi]++;
                }
        }
        if (!to_del.empty()) {
            std::sort(to_del.begin(), to_del.end());
            /* IPT hand-off (src/Integration.cpp:150-168): IPTPart(pnp1[ii],
             * current_time, ipt_diam, ipt_area), tracked to its end */
            if (ipt.using_ipt && s.aero_source == FJSPH_MESH_INFL && have_mesh) {
                const IptParams P = ipt_params();
                for (size_t ii : to_del) {
                    IptP<D> q;
                    q.pid = pnp1[ii].part_id;
                    q.t = current_time;
                    q.cellID = pnp1[ii].cellID;
                    q.cellV = pnp1[ii].cellV;
                    q.cellRho = pnp1[ii].cellRho;
                    q.v = pnp1[ii].v;
                    q.xi = pnp1[ii].xi;
                    q.d = P.diam;
                    q.A = P.area;
                    const IptRecord r = ipt_integrate<D>(mesh, P, q);
                    ipt_records.push_back(r);
                    if (r.status == IPT_SUCCESS) ipt_success++;
                    else if (r.status == IPT_FAILED) ipt_failed++;
                }
            }
            for (auto it = to_del.rbegin(); it != to_del.rend(); ++it) {
                pnp1.erase(pnp1.begin() + long(*it));
                total_points--;
            }
            size_t delshift = 0;
            for (size_t bi = 0; bi < blocks.size(); ++bi) {
                Block& bl = blocks[bi];
                if (!bl.b.is_fluid) continue;
                bl.b.first -= int64_t(delshift);
                delshift += n_del[bi];
                bl.b.last -= int64_t(delshift);
                for (size_t& x : bl.back) x -= delshift;
                for (auto& col : bl.buffer)
                    for (size_t& x : col) x -= delshift;
            }
        }
        pn = pnp1;
    }

    /* integrate (src/Integration.cpp:233-303) */
    double integrate()
    {
        const double step_error = integrate_no_update();
        update_data();
        current_time += delta_t;
        if (step_error > s.min_residual || maxRho_pc > s.rho_max_iter) {
            if (step_error > 0.6 * s.min_residual) {
                cfl = std::max(s.cfl_min, cfl - s.cfl_step);
                n_unstable = 0;
            } else if (n_unstable > unsigned(s.n_unstable_limit)) {
                cfl = std::max(s.cfl_min, cfl - s.cfl_step);
                n_unstable = 0;
            } else
                n_unstable++;
        } else
            n_unstable = 0;
        if (iteration < s.subits_factor * s.max_subits && n_unstable == 0) {
            if (n_stable > unsigned(s.n_stable_limit)) {
                cfl = std::min(s.cfl_max, cfl + s.cfl_step);
                n_stable = 0;
            } else
                n_stable++;
        } else
            n_stable = 0;
        return step_error;
    }
};

}  // namespace orc

#endif
