/*
 * Implicit particle tracking (IPT) of particles leaving the SPH zone: one
 * thread per particle marches it from cell to cell of the gas mesh until it
 * leaves the mesh, passes x = max_x, fails, or reaches the step cap.
 *
 *   GetCd (src/Aero.h:10-18), IPT::AeroForce / BFD1 / BFD2 / Integrate
 *   (src/IPT.cpp:829-1106), FindFace and CheckCellFace
 *   (src/Containment.cpp:893-1079), Cross_Plane, RayNormalIntersection and
 *   MollerTrumbore (src/Geometry.cpp:396-480 in 2D, 576-745 in 3D).
 *
 * The operation order is the reference's (this file is compiled without
 * contraction into fused multiply-adds).  Tracks are independent, so the
 * kernel is latency bound by its cell walks: tracks diverge, and the
 * per-thread state (three particle states) lives in registers.  Guards
 * where the reference would index out of range: a particle without a cell,
 * or a crossing with no face nearer than the initial distance 1e6, stops as
 * failed.  The step cap (no counterpart in the reference) bounds every
 * thread's loop.
 */
#include "device_common.hpp"

namespace fj {
namespace {

constexpr double kEpsD = 2.220446049250313080847e-16; /* DBL_EPSILON, MEPSILON (src/VarDefs.h:52) */

template <int D>
struct Vd {
    double c[D];
};

template <int D> __device__ inline Vd<D> vzero() { Vd<D> r; for (int k = 0; k < D; ++k) r.c[k] = 0.0; return r; }
template <int D> __device__ inline Vd<D> operator+(const Vd<D>& a, const Vd<D>& b) { Vd<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] + b.c[k]; return r; }
template <int D> __device__ inline Vd<D> operator-(const Vd<D>& a, const Vd<D>& b) { Vd<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] - b.c[k]; return r; }
template <int D> __device__ inline Vd<D> operator*(double s, const Vd<D>& a) { Vd<D> r; for (int k = 0; k < D; ++k) r.c[k] = s * a.c[k]; return r; }
template <int D> __device__ inline Vd<D> operator/(const Vd<D>& a, double s) { Vd<D> r; for (int k = 0; k < D; ++k) r.c[k] = a.c[k] / s; return r; }
template <int D> __device__ inline double vdot(const Vd<D>& a, const Vd<D>& b) { double s = a.c[0] * b.c[0]; for (int k = 1; k < D; ++k) s += a.c[k] * b.c[k]; return s; }
template <int D> __device__ inline double vnorm(const Vd<D>& a) { return sqrt(vdot(a, a)); }
template <int D> __device__ inline Vd<D> vnormalized(const Vd<D>& a)
{
    const double z = vdot(a, a);
    return z > 0.0 ? a / sqrt(z) : a;
}

template <int D>
struct IptP {
    Vd<D> xi, v, cellV, faceV;
    double t, dt, cellRho, faceRho, acc, d;
    int32_t cellID, faceID, going;
};

__device__ inline double get_cd(double Re)
{
    return (1.0 + 0.197 * pow(Re, 0.63) + 2.6e-04 * pow(Re, 1.38)) * (24.0 / (Re + 0.00001));
}

template <int D>
__device__ double ipt_aero(const Vd<D>& vdiff, const IptP<D>& p, const IptParams& P)
{
    const double Re = 2.0 * p.faceRho * vnorm(vdiff) * p.d / P.mu_g;
    const double Cd = get_cd(Re);
    return vnorm(vdiff) * (3.0 * Cd * p.faceRho) / (4.0 * p.d * P.rho_l);
}

template <int D>
__device__ Vd<D> gvec(const IptParams& P)
{
    Vd<D> g;
    for (int a = 0; a < D; ++a) g.c[a] = P.grav[a];
    return g;
}

template <int D>
__device__ void bfd1(const IptParams& P, double dt, const IptP<D>& pn, IptP<D>& p1)
{
    const Vd<D> vdiff = p1.faceV - p1.v;
    const double res = ipt_aero<D>(vdiff, p1, P);
    p1.acc = res;
    p1.v = (pn.v + (dt * res) * p1.faceV + dt * gvec<D>(P)) / (1.0 + dt * res);
    p1.xi = pn.xi + (0.5 * dt) * (p1.v + pn.v);
}

template <int D>
__device__ void bfd2(const IptParams& P, double dt, double dtm1, const IptP<D>& pm1, const IptP<D>& pn, IptP<D>& p1)
{
    const Vd<D> vdiff = p1.faceV - p1.v;
    const double res = ipt_aero<D>(vdiff, p1, P);
    p1.acc = res;
    const Vd<D> inner = (dt * dtm1) * (res * p1.faceV + gvec<D>(P)) + (dt + dtm1) * pn.v;
    const Vd<D> num = (dt + dtm1) * inner - (dt * dt) * pm1.v;
    p1.v = num / (dtm1 * ((2 * dt + dtm1) + (dt * (dt + dtm1)) * res));
    p1.xi = pn.xi + (0.5 * p1.dt) * (p1.v + pn.v);
}

__device__ inline const double* mv(const DevMesh& M, int32_t v) { return M.verts + (int64_t)v * 3; }

__device__ inline bool cross_plane2(const DevMesh& M, int32_t f, const double* tp, const double* rp)
{
    const double* a = mv(M, M.face_v[M.face_off[f]]);
    const double* b = mv(M, M.face_v[M.face_off[f] + 1]);
    const double v1 = tp[0] * (a[1] - b[1]) - tp[1] * (a[0] - b[0]) + (a[0] * b[1] - b[0] * a[1]);
    const double v2 = rp[0] * (a[1] - b[1]) - rp[1] * (a[0] - b[0]) + (a[0] * b[1] - b[0] * a[1]);
    return (v1 < 0.0) != (v2 < 0.0);
}

__device__ inline bool cross_plane3(const DevMesh& M, int32_t f, const double* tp, const double* rp)
{
    const int32_t o = M.face_off[f];
    const double* a = mv(M, M.face_v[o]);
    const double* b = mv(M, M.face_v[o + 1]);
    const double* c = mv(M, M.face_v[o + 2]);
    return (det4_rows(tp, a, b, c) < 0.0) != (det4_rows(rp, a, b, c) < 0.0);
}

__device__ inline Vd<3> ld3(const double* p)
{
    Vd<3> r;
    r.c[0] = p[0];
    r.c[1] = p[1];
    r.c[2] = p[2];
    return r;
}

__device__ inline Vd<3> cross3(const Vd<3>& a, const Vd<3>& b)
{
    Vd<3> r;
    r.c[0] = a.c[1] * b.c[2] - a.c[2] * b.c[1];
    r.c[1] = a.c[2] * b.c[0] - a.c[0] * b.c[2];
    r.c[2] = a.c[0] * b.c[1] - a.c[1] * b.c[0];
    return r;
}

/* MollerTrumbore as the reference writes it (its second edge is v2 - v1) */
__device__ bool moller(const Vd<3>& v0, const Vd<3>& v1, const Vd<3>& v2, const Vd<3>& o, const Vd<3>& r, double& dt,
                       double& denom)
{
    const Vd<3> e1 = v1 - v0, e2 = v2 - v1;
    const Vd<3> h = cross3(r, e2);
    const double a = vdot(e1, h);
    if (a < kEpsD && a > -kEpsD) {
        dt = kEpsD;
        denom = kEpsD;
        return false;
    }
    const double f = 1.0 / a;
    const Vd<3> s = o - v0;
    const double u = f * vdot(s, h);
    if (u < 0.0 || u > 1.0) {
        dt = 9999999;
        denom = -1;
        return false;
    }
    const Vd<3> q = cross3(s, e1);
    const double v = f * vdot(r, q);
    if (v < 0.0 || u + v > 1.0) {
        dt = 9999999;
        denom = -1;
        return false;
    }
    dt = f * vdot(e2, q);
    denom = dt > kEpsD ? 1.0 : -1.0;
    return true;
}

template <int D>
__device__ void ray_normal(const DevMesh& M, const Vd<D>& o, const Vd<D>& r, int32_t f, int32_t cell, double& dt,
                           double& denom)
{
    const int32_t off = M.face_off[f], nv = M.face_off[f + 1] - off;
    if constexpr (D == 2) {
        const double* a = mv(M, M.face_v[off]);
        const double* b = mv(M, M.face_v[off + 1]);
        Vd<2> nrm, fc, cc;
        nrm.c[0] = a[1] - b[1];
        nrm.c[1] = b[0] - a[0];
        for (int k = 0; k < 2; ++k) {
            fc.c[k] = (a[k] + b[k]) / 2.0;
            cc.c[k] = M.cCentre[(int64_t)cell * 3 + k];
        }
        if (vdot(nrm, fc - cc) < 0) nrm = -1.0 * nrm;
        Vd<2> tp = vzero<2>();
        for (int32_t i = 0; i < nv; ++i) {
            const double* q = mv(M, M.face_v[off + i]);
            Vd<2> w;
            for (int k = 0; k < 2; ++k) w.c[k] = q[k] - o.c[k];
            if (vdot(w, w) > vdot(tp, tp)) tp = w;
        }
        denom = vdot(r, nrm);
        dt = vdot(tp, nrm) / denom;
    } else {
        const Vd<3> p0 = ld3(mv(M, M.face_v[off])), p1 = ld3(mv(M, M.face_v[off + 1])), p2 = ld3(mv(M, M.face_v[off + 2]));
        if (nv == 3) {
            moller(p0, p1, p2, o, r, dt, denom);
        } else if (!moller(p0, p1, p2, o, r, dt, denom)) {
            moller(p0, ld3(mv(M, M.face_v[off + 3])), p2, o, r, dt, denom);
        }
    }
}

template <int D>
__device__ double find_face(const DevMesh& M, const IptP<D>& pn, IptP<D>& p1)
{
    if (pn.cellID < 0 || pn.cellID >= M.ncells) {
        p1.xi = pn.xi;
        p1.v = pn.v;
        p1.going = 0;
        return 0.0;
    }
    Vd<D> tv = 0.5 * (p1.v + pn.v);
    if (vnorm(tv) < 1e-10) tv = p1.cellV;
    const Vd<D> nd = vnormalized(tv);
    const Vd<D> tp = pn.xi - 1e3 * nd, rp = pn.xi + 1e3 * nd;
    double tpa[3] = {0, 0, 0}, rpa[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        tpa[k] = tp.c[k];
        rpa[k] = rp.c[k];
    }
    int32_t next = pn.faceID;
    double mindist = 1e6;
    int hit = 0;
    for (int32_t k = M.cell_off[pn.cellID]; k < M.cell_off[pn.cellID + 1]; ++k) {
        const int32_t f = M.cell_f[k];
        const bool crosses = D == 2 ? cross_plane2(M, f, tpa, rpa) : cross_plane3(M, f, tpa, rpa);
        if (!crosses || f == pn.faceID) continue;
        double dt = 0.0, denom = 0.0;
        ray_normal<D>(M, pn.xi, tv, f, pn.cellID, dt, denom);
        if (denom > 0) {
            if (dt < mindist) {
                next = f;
                mindist = dt > kEpsD ? dt : kEpsD;
            }
            hit = 1;
        }
    }
    if (hit == 0 || next < 0) {
        p1.xi = pn.xi;
        p1.v = pn.v;
        p1.going = 0;
        return 0.0;
    }
    p1.faceID = next;
    const int32_t l = M.face_lr[2 * (int64_t)next], r = M.face_lr[2 * (int64_t)next + 1];
    const int32_t nc = l == pn.cellID ? r : l;
    p1.cellID = nc;
    if (nc > -1) {
        for (int a = 0; a < D; ++a) p1.cellV.c[a] = M.cVel[(int64_t)nc * D + a];
        p1.cellRho = M.cRho[nc];
    }
    return mindist;
}

template <int D>
__device__ inline void face_props(IptP<D>& p1, const IptP<D>& pn)
{
    p1.faceV = 0.5 * (p1.cellV + pn.cellV);
    p1.faceRho = 0.5 * (p1.cellRho + pn.cellRho);
}

template <int D>
__device__ void relax_loop(const DevMesh& M, const IptParams& P, const IptP<D>& pm1, const IptP<D>& pn, IptP<D>& p1,
                           bool first)
{
    int iter = 0;
    double error = 1.0;
    while ((error > -7.0 || iter < 3) && iter < P.max_subits) {
        const double dt = p1.dt;
        if (P.eq_order == 2) bfd2<D>(P, dt, (first || iter == 0) ? p1.dt : pn.dt, pm1, pn, p1);
        else bfd1<D>(P, dt, pn, p1);
        const double dtt = find_face<D>(M, pn, p1);
        if (first || iter > P.n_relax) p1.dt = (1.0 - P.relax) * dtt + P.relax * p1.dt;
        else p1.dt = dtt;
        face_props<D>(p1, pn);
        error = log10(fabs(p1.dt - dt));
        iter++;
    }
}

template <int D>
__global__ __launch_bounds__(64) void k_ipt_track(DevMesh M, IptParams P, int64_t n, const IptStart* __restrict__ in,
                                                  fjsph_ipt_record* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const IptStart s = in[i];
    IptP<D> p1;
    for (int a = 0; a < D; ++a) {
        p1.xi.c[a] = s.x[a];
        p1.v.c[a] = s.v[a];
        p1.cellV.c[a] = s.cellV[a];
    }
    p1.faceV = vzero<D>();
    p1.t = s.t;
    p1.dt = 0.0;
    p1.cellRho = s.cellRho;
    p1.faceRho = 0.0;
    p1.acc = 0.0;
    p1.d = P.diam;
    p1.cellID = s.cell;
    p1.faceID = -1;
    p1.going = 1;
    IptP<D> pm1 = p1, pn = p1;
    int32_t steps = 0, status = 0;
    /* first step (src/IPT.cpp:879-960) */
    p1.dt = find_face<D>(M, pn, p1);
    face_props<D>(p1, pn);
    relax_loop<D>(M, P, pm1, pn, p1, true);
    for (;;) {
        ++steps;
        if (p1.going == 0 || vnorm(p1.v) > 1e3 || vnorm(p1.xi - pn.xi) > P.max_len) {
            status = FJSPH_IPT_FAILED;
            break;
        }
        if (p1.cellID < 0 || p1.xi.c[0] > P.max_x) {
            status = FJSPH_IPT_SUCCESS;
            break;
        }
        if (P.eq_order == 2) pm1 = pn;
        pn = p1;
        p1.t += p1.dt;
        if (steps >= P.max_steps) {
            status = FJSPH_IPT_STOPPED;
            break;
        }
        /* next cell (src/IPT.cpp:990-1104) */
        if (p1.d < 1.0e-5) {
            p1.v = p1.cellV;
        } else if (p1.d < 100.0e-6) {
            p1.v = (1.0 - (p1.d - 1e-05) / (9e-5)) * p1.cellV + ((p1.d - 1e-05) / (9e-5)) * p1.v;
        }
        p1.dt = find_face<D>(M, pn, p1);
        face_props<D>(p1, pn);
        relax_loop<D>(M, P, pm1, pn, p1, false);
    }
    fjsph_ipt_record r;
    r.part_id = s.part_id;
    r.status = status;
    r.steps = steps;
    r.cell = p1.cellID;
    r.face = p1.faceID;
    r.t0 = s.t;
    r.t = p1.t;
    for (int a = 0; a < 3; ++a) {
        r.x0[a] = a < D ? s.x[a] : 0.0;
        r.x[a] = a < D ? p1.xi.c[a] : 0.0;
        r.v[a] = a < D ? p1.v.c[a] : 0.0;
    }
    out[i] = r;
}

/* IPTPart(SPHPart, time, diam, area) (src/Var.h:734-762): position,
 * velocity, cell and its gas values, particle id; everything else starts
 * from zero. */
template <int D>
__global__ __launch_bounds__(TPB) void k_ipt_gather(DevState s, const int32_t* __restrict__ slots, int64_t n, double t,
                                                    IptStart* __restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)TPB + threadIdx.x;
    if (i >= n) return;
    const int64_t sl = slots[i];
    IptStart q;
    for (int a = 0; a < 3; ++a) {
        q.x[a] = a < D ? s.x[sl * D + a] : 0.0;
        q.v[a] = a < D ? s.v[sl * D + a] : 0.0;
        q.cellV[a] = a < D ? s.cellV[sl * D + a] : 0.0;
    }
    q.cellRho = s.cellRho[sl];
    q.t = t;
    q.part_id = s.pid[sl];
    q.cell = s.cellID[sl];
    q.pad = 0;
    out[i] = q;
}

}  // namespace

void launch_ipt_gather(int dim, DevState s, const int32_t* slots, int64_t n, double t, IptStart* out, hipStream_t st)
{
    if (n <= 0) return;
    const dim3 g(unsigned((n + TPB - 1) / TPB)), b(TPB);
    if (dim == 2) hipLaunchKernelGGL(k_ipt_gather<2>, g, b, 0, st, s, slots, n, t, out);
    else hipLaunchKernelGGL(k_ipt_gather<3>, g, b, 0, st, s, slots, n, t, out);
}

void launch_ipt_track(int dim, const DevMesh& m, const IptParams& p, int64_t n, const IptStart* in,
                      fjsph_ipt_record* out, hipStream_t st)
{
    if (n <= 0) return;
    const dim3 g(unsigned((n + 63) / 64)), b(64);
    if (dim == 2) hipLaunchKernelGGL(k_ipt_track<2>, g, b, 0, st, m, p, n, in, out);
    else hipLaunchKernelGGL(k_ipt_track<3>, g, b, 0, st, m, p, n, in, out);
}

}  // namespace fj
