/*
 * Implicit particle tracking (IPT) restated on the CPU -- TEST
 * INFRASTRUCTURE ONLY (the oracle library links it; the engine does not).
 *
 *   GetCd (src/Aero.h:10-18), IPT::AeroForce / BFD1 / BFD2 / Integrate
 *   (src/IPT.cpp:829-1106), FindFace and CheckCellFace
 *   (src/Containment.cpp:893-1079), Cross_Plane, RayNormalIntersection and
 *   MollerTrumbore (src/Geometry.cpp:396-480 in 2D, 576-745 in 3D), the
 *   IPTPart constructor from an SPH particle (src/Var.h:734-762).
 *
 * Written in the reference's operation order (scalar products first,
 * component sums left to right, no fused multiply-add).  Guards where the
 * reference would index out of range: a particle without a cell, or a
 * crossing whose nearest face lies beyond the initial distance 1e6,
 * stops as failed.  A cap on the number of cell-to-cell steps ends a track
 * as "stopped" (the reference has none).
 */
/* Included by fjsph_oracle.hpp inside namespace orc, after the mesh
 * helpers (MeshData, det4_rows) and before the simulation, whose
 * update_data hands particles over to it. */

struct IptParams {
    double grav[3];
    double mu_g, rho_l, relax, max_x, max_len, diam, area;
    int eq_order, n_relax, max_subits;
    int64_t max_steps;
};

enum { IPT_SUCCESS = 1, IPT_FAILED = 2, IPT_STOPPED = 3 };

template <int D>
struct IptP {
    int64_t pid = 0;
    int going = 1, failed = 0;
    double t = 0, dt = 0;
    int64_t faceID = -1, cellID = -3;
    Vec<D> faceV = Vec<D>::zero(), cellV = Vec<D>::zero(), v = Vec<D>::zero(), xi = Vec<D>::zero();
    double faceRho = 0, cellRho = 0, acc = 0, relax = 1, d = 0, A = 0;
};

struct IptRecord {
    int64_t pid;
    int32_t status, steps, cell, face;
    double t0, t, x0[3], x[3], v[3];
};

inline double get_cd(double Re)
{
    return (1.0 + 0.197 * std::pow(Re, 0.63) + 2.6e-04 * std::pow(Re, 1.38)) * (24.0 / (Re + 0.00001));
}

template <int D>
double ipt_aero(const Vec<D>& vdiff, const IptP<D>& p, double mu_g, double rho_l)
{
    const double Re = 2.0 * p.faceRho * norm(vdiff) * p.d / mu_g;
    const double Cd = get_cd(Re);
    return norm(vdiff) * (3.0 * Cd * p.faceRho) / (4.0 * p.d * rho_l);
}

template <int D>
Vec<D> gvec(const IptParams& P)
{
    Vec<D> g;
    for (int a = 0; a < D; ++a) g.c[a] = P.grav[a];
    return g;
}

template <int D>
void bfd1(const IptParams& P, double dt, const IptP<D>& pn, IptP<D>& p1)
{
    const Vec<D> vdiff = p1.faceV - p1.v;
    const double res = ipt_aero<D>(vdiff, p1, P.mu_g, P.rho_l);
    p1.acc = res;
    p1.v = (pn.v + (dt * res) * p1.faceV + dt * gvec<D>(P)) / (1.0 + dt * res);
    p1.xi = pn.xi + (0.5 * dt) * (p1.v + pn.v);
}

template <int D>
void bfd2(const IptParams& P, double dt, double dtm1, const IptP<D>& pm1, const IptP<D>& pn, IptP<D>& p1)
{
    const Vec<D> vdiff = p1.faceV - p1.v;
    const double res = ipt_aero<D>(vdiff, p1, P.mu_g, P.rho_l);
    p1.acc = res;
    const Vec<D> inner = (dt * dtm1) * (res * p1.faceV + gvec<D>(P)) + (dt + dtm1) * pn.v;
    const Vec<D> num = (dt + dtm1) * inner - (dt * dt) * pm1.v;
    p1.v = num / (dtm1 * ((2 * dt + dtm1) + (dt * (dt + dtm1)) * res));
    p1.xi = pn.xi + (0.5 * p1.dt) * (p1.v + pn.v);
}

inline bool cross_plane2(const MeshData& M, int64_t f, const double* tp, const double* rp)
{
    const double* a = M.vert(M.face_v[size_t(M.face_off[size_t(f)])]);
    const double* b = M.vert(M.face_v[size_t(M.face_off[size_t(f)]) + 1]);
    auto vol = [&](const double* q) {
        return q[0] * (a[1] - b[1]) - q[1] * (a[0] - b[0]) + (a[0] * b[1] - b[0] * a[1]);
    };
    return (vol(tp) < 0.0) != (vol(rp) < 0.0);
}

inline bool cross_plane3(const MeshData& M, int64_t f, const double* tp, const double* rp)
{
    const int64_t o = M.face_off[size_t(f)];
    const double* a = M.vert(M.face_v[size_t(o)]);
    const double* b = M.vert(M.face_v[size_t(o) + 1]);
    const double* c = M.vert(M.face_v[size_t(o) + 2]);
    return (det4_rows(tp, a, b, c) < 0.0) != (det4_rows(rp, a, b, c) < 0.0);
}

inline Vec<3> v3(const double* p)
{
    Vec<3> r;
    r.c[0] = p[0];
    r.c[1] = p[1];
    r.c[2] = p[2];
    return r;
}

inline Vec<3> cross3(const Vec<3>& a, const Vec<3>& b)
{
    Vec<3> r;
    r.c[0] = a.c[1] * b.c[2] - a.c[2] * b.c[1];
    r.c[1] = a.c[2] * b.c[0] - a.c[0] * b.c[2];
    r.c[2] = a.c[0] * b.c[1] - a.c[1] * b.c[0];
    return r;
}

/* MollerTrumbore as the reference writes it (its second edge is v2 - v1) */
inline bool moller(const Vec<3>& v0, const Vec<3>& v1, const Vec<3>& v2, const Vec<3>& o, const Vec<3>& r, double& dt,
                   double& denom)
{
    const Vec<3> e1 = v1 - v0, e2 = v2 - v1;
    const Vec<3> h = cross3(r, e2);
    const double a = dot(e1, h);
    if (a < kEps && a > -kEps) {
        dt = kEps;
        denom = kEps;
        return false;
    }
    const double f = 1.0 / a;
    const Vec<3> s = o - v0;
    const double u = f * dot(s, h);
    if (u < 0.0 || u > 1.0) {
        dt = 9999999;
        denom = -1;
        return false;
    }
    const Vec<3> q = cross3(s, e1);
    const double v = f * dot(r, q);
    if (v < 0.0 || u + v > 1.0) {
        dt = 9999999;
        denom = -1;
        return false;
    }
    dt = f * dot(e2, q);
    denom = dt > kEps ? 1.0 : -1.0;
    return true;
}

template <int D>
void ray_normal(const MeshData& M, const Vec<D>& o, const Vec<D>& r, int64_t f, int64_t cell, double& dt, double& denom)
{
    const int64_t off = M.face_off[size_t(f)], nv = M.face_off[size_t(f) + 1] - off;
    if constexpr (D == 2) {
        const double* a = M.vert(M.face_v[size_t(off)]);
        const double* b = M.vert(M.face_v[size_t(off) + 1]);
        Vec<D> nrm, fc, cc;
        nrm.c[0] = a[1] - b[1];
        nrm.c[1] = b[0] - a[0];
        for (int k = 0; k < 2; ++k) {
            fc.c[k] = (a[k] + b[k]) / 2.0;
            cc.c[k] = M.cCentre[size_t(cell) * 2 + k];
        }
        if (dot(nrm, fc - cc) < 0) nrm = -1.0 * nrm;
        Vec<D> tp = Vec<D>::zero();
        for (int64_t i = 0; i < nv; ++i) {
            Vec<D> w;
            const double* q = M.vert(M.face_v[size_t(off + i)]);
            for (int k = 0; k < 2; ++k) w.c[k] = q[k] - o.c[k];
            if (sqnorm(w) > sqnorm(tp)) tp = w;
        }
        denom = dot(r, nrm);
        dt = dot(tp, nrm) / denom;
    } else {
        Vec<3> o3, r3;
        for (int k = 0; k < 3; ++k) {
            o3.c[k] = o.c[k];
            r3.c[k] = r.c[k];
        }
        const Vec<3> p0 = v3(M.vert(M.face_v[size_t(off)])), p1 = v3(M.vert(M.face_v[size_t(off) + 1])),
                     p2 = v3(M.vert(M.face_v[size_t(off) + 2]));
        if (nv == 3) {
            moller(p0, p1, p2, o3, r3, dt, denom);
        } else if (!moller(p0, p1, p2, o3, r3, dt, denom)) {
            moller(p0, v3(M.vert(M.face_v[size_t(off) + 3])), p2, o3, r3, dt, denom);
        }
    }
}

template <int D>
double find_face(const MeshData& M, const IptP<D>& pn, IptP<D>& p1)
{
    if (pn.cellID < 0 || pn.cellID >= int64_t(M.ncells())) {
        p1.xi = pn.xi;
        p1.v = pn.v;
        p1.going = 0;
        return 0.0;
    }
    Vec<D> tv = 0.5 * (p1.v + pn.v);
    if (norm(tv) < 1e-10) tv = p1.cellV;
    const Vec<D> nd = normalized(tv);
    const Vec<D> tp = pn.xi - 1e3 * nd, rp = pn.xi + 1e3 * nd;
    double tpa[3] = {0, 0, 0}, rpa[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        tpa[k] = tp.c[k];
        rpa[k] = rp.c[k];
    }
    int64_t next = pn.faceID;
    double mindist = 1e6;
    int hit = 0;
    for (int64_t k = M.cell_off[size_t(pn.cellID)]; k < M.cell_off[size_t(pn.cellID) + 1]; ++k) {
        const int64_t f = M.cell_f[size_t(k)];
        const bool crosses = D == 2 ? cross_plane2(M, f, tpa, rpa) : cross_plane3(M, f, tpa, rpa);
        if (!crosses || f == pn.faceID) continue;
        double dt = 0.0, denom = 0.0;
        ray_normal<D>(M, pn.xi, tv, f, pn.cellID, dt, denom);
        if (denom > 0) {
            if (dt < mindist) {
                next = f;
                mindist = dt > kEps ? dt : kEps;
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
    const int64_t nc = M.face_l[size_t(next)] == pn.cellID ? M.face_r[size_t(next)] : M.face_l[size_t(next)];
    p1.cellID = nc;
    if (nc > -1) {
        for (int a = 0; a < D; ++a) p1.cellV.c[a] = M.cVel[size_t(nc) * D + a];
        p1.cellRho = M.cRho[size_t(nc)];
    }
    return mindist;
}

template <int D>
void face_props(IptP<D>& p1, const IptP<D>& pn)
{
    p1.faceV = 0.5 * (p1.cellV + pn.cellV);
    p1.faceRho = 0.5 * (p1.cellRho + pn.cellRho);
}

/* IPT::Integrate (src/IPT.cpp:870-1106) for one particle; the record holds
 * the particle where its track ends (Terminate_Particle's last entry). */
template <int D>
IptRecord ipt_integrate(const MeshData& M, const IptParams& P, IptP<D> p1)
{
    IptRecord rec{};
    rec.pid = p1.pid;
    rec.t0 = p1.t;
    for (int a = 0; a < D; ++a) rec.x0[a] = p1.xi.c[a];
    IptP<D> pm1 = p1, pn = p1;
    const int max_iter = P.max_subits;
    const int min_iter = 3;
    int64_t steps = 0;
    auto finish = [&](int status) {
        rec.status = status;
        rec.steps = int32_t(steps);
        rec.t = p1.t;
        rec.cell = int32_t(p1.cellID);
        rec.face = int32_t(p1.faceID);
        for (int a = 0; a < D; ++a) {
            rec.x[a] = p1.xi.c[a];
            rec.v[a] = p1.v.c[a];
        }
        return rec;
    };
    auto terminated = [&](int& status) {
        if (p1.going == 0 || norm(p1.v) > 1e3 || norm(p1.xi - pn.xi) > P.max_len) {
            status = IPT_FAILED;
            return true;
        }
        if (p1.cellID < 0 || p1.xi.c[0] > P.max_x) {
            status = IPT_SUCCESS;
            return true;
        }
        return false;
    };
    auto march = [&]() {
        if (P.eq_order == 2) pm1 = pn;
        pn = p1;
        p1.t += p1.dt;
    };
    /* first step */
    p1.dt = find_face<D>(M, pn, p1);
    face_props<D>(p1, pn);
    int iter = 0;
    double error = 1.0;
    while ((error > -7.0 || iter < min_iter) && iter < max_iter) {
        const double dt = p1.dt;
        if (P.eq_order == 2) bfd2<D>(P, dt, p1.dt, pm1, pn, p1);
        else bfd1<D>(P, dt, pn, p1);
        const double dtt = find_face<D>(M, pn, p1);
        p1.dt = (1.0 - P.relax) * dtt + P.relax * p1.dt;
        face_props<D>(p1, pn);
        error = std::log10(std::fabs(p1.dt - dt));
        iter++;
    }
    ++steps;
    int status = 0;
    if (terminated(status)) return finish(status);
    march();
    for (;;) {
        if (steps >= P.max_steps) return finish(IPT_STOPPED);
        if (p1.d < 1.0e-5) {
            p1.v = p1.cellV;
        } else if (p1.d < 100.0e-6) {
            p1.v = (1.0 - (p1.d - 1e-05) / (9e-5)) * p1.cellV + ((p1.d - 1e-05) / (9e-5)) * p1.v;
        }
        p1.dt = find_face<D>(M, pn, p1);
        face_props<D>(p1, pn);
        iter = 0;
        error = 1.0;
        while ((error > -7.0 || iter < min_iter) && iter < max_iter) {
            const double dt = p1.dt;
            if (P.eq_order == 2) bfd2<D>(P, dt, iter > 0 ? pn.dt : p1.dt, pm1, pn, p1);
            else bfd1<D>(P, dt, pn, p1);
            const double dtt = find_face<D>(M, pn, p1);
            if (iter > P.n_relax) p1.dt = (1.0 - P.relax) * dtt + P.relax * p1.dt;
            else p1.dt = dtt;
            face_props<D>(p1, pn);
            error = std::log10(std::fabs(p1.dt - dt));
            iter++;
        }
        ++steps;
        if (terminated(status)) return finish(status);
        march();
    }
}

