// C ABI of the projection solver's operators (include/dgh.h): the dof-field convection div(U, U) with the LF
// scheme, the nodal gradient and divergence, and the field combinations of its time integration.
#include <hip/hip_runtime.h>

#include <vector>

#include "dgh.h"
#include "dgh_internal.hpp"
#include "ins.hpp"

using namespace dg;

namespace {

// Pr Vq and Ps Vq (Np x Np): the weak derivative of a nodal flux interpolated to the cubature points
bool ensure_ins(dgh_ctx c, RefDev* rd) {
    if (rd->insSr) return true;
    const RefElem& re = rd->re;
    const int Np = re.Np, Nq = re.Nq;
    std::vector<double> Sr(size_t(Np) * Np), Ss(size_t(Np) * Np);
    for (int n = 0; n < Np; ++n)
        for (int j = 0; j < Np; ++j) {
            long double a = 0.0L, b = 0.0L;
            for (int q = 0; q < Nq; ++q) {
                a += (long double)re.Pr[size_t(n) * Nq + q] * re.Vq[size_t(q) * Np + j];
                b += (long double)re.Ps[size_t(n) * Nq + q] * re.Vq[size_t(q) * Np + j];
            }
            Sr[size_t(n) * Np + j] = double(a);
            Ss[size_t(n) * Np + j] = double(b);
        }
    return c->upload(&rd->insSr, Sr) && c->upload(&rd->insSs, Ss) && c->upload(&rd->insLift, re.Lift) &&
           c->upload(&rd->insDr, re.Dr) && c->upload(&rd->insDs, re.Ds);
}

InsArgs ins_args(MeshDev* m) {
    InsArgs a;
    a.K = m->host.K;
    a.geo = m->geo64;
    a.conn = m->conn;
    a.fmask = m->ref->fmask;
    a.bcq = (m->bcq && m->bc_nvar == 2) ? m->bcq : nullptr;
    a.Sr = m->ref->insSr;
    a.Ss = m->ref->insSs;
    a.Lift = m->ref->insLift;
    a.Vf = m->ref->Vf64;
    a.Dr = m->ref->insDr;
    a.Ds = m->ref->insDs;
    return a;
}

// a field of nvar variables with the mesh's size (K and order), on any mesh
FieldDev* sized_field(dgh_ctx c, MeshDev* m, int h, int nvar, const char* what) {
    FieldDev* f = c->field(h);
    if (!f) return nullptr;
    if (f->nvar != nvar || f->K != m->host.K || f->Np != m->ref->re.Np) {
        c->fail(DGH_ERR_ARG, "%s: field %d must have %d variable(s) on a mesh of K = %d, Np = %d", what, h, nvar,
                m->host.K, m->ref->re.Np);
        return nullptr;
    }
    return f;
}

int deriv(dgh_ctx c, int mh, int u, int out, int mode, const char* what) {
    if (!c) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    FieldDev* fu = sized_field(c, m, u, mode == 0 ? 1 : 2, what);
    FieldDev* fo = sized_field(c, m, out, mode == 0 ? 2 : 1, what);
    if (!fu || !fo) return c->last;
    (void)hipSetDevice(c->device);
    if (!ensure_ins(c, m->ref)) return c->last;
    return c->hip(launch_nodal_deriv(m->ref->re.N, ins_args(m), mode, fu->d, fo->d, c->stream), what) ? DGH_OK
                                                                                                      : c->last;
}

}  // namespace

extern "C" {

int dgh_ins_convection(dgh_ctx c, int mh, int U, int out) {
    if (!c) return DGH_ERR_ARG;
    MeshDev* m = c->mesh(mh);
    if (!m) return c->last;
    if (m->ref->re.N > 6) return c->fail(DGH_ERR_UNSUPPORTED, "dgh_ins_convection: orders 1..6"), DGH_ERR_UNSUPPORTED;
    FieldDev* fu = sized_field(c, m, U, 2, "dgh_ins_convection");
    FieldDev* fo = sized_field(c, m, out, 2, "dgh_ins_convection");
    if (!fu || !fo) return c->last;
    if (U == out) return c->fail(DGH_ERR_ARG, "dgh_ins_convection: U and out must differ"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (!ensure_ins(c, m->ref)) return c->last;
    return c->hip(launch_ins_convection(m->ref->re.N, ins_args(m), fu->d, fo->d, c->stream), "ins convection")
               ? DGH_OK
               : c->last;
}

int dgh_nodal_grad(dgh_ctx c, int mh, int u, int out) {
    if (c && u == out) return c->fail(DGH_ERR_ARG, "dgh_nodal_grad: u and out must differ"), DGH_ERR_ARG;
    return deriv(c, mh, u, out, 0, "dgh_nodal_grad");
}

int dgh_nodal_div(dgh_ctx c, int mh, int u, int out) {
    if (c && u == out) return c->fail(DGH_ERR_ARG, "dgh_nodal_div: u and out must differ"), DGH_ERR_ARG;
    return deriv(c, mh, u, out, 1, "dgh_nodal_div");
}

int dgh_field_lincomb(dgh_ctx c, int out, int n, const int* fields, const double* coefs) {
    if (!c) return DGH_ERR_ARG;
    if (n < 1 || n > kLinCombMax || !fields || !coefs)
        return c->fail(DGH_ERR_ARG, "dgh_field_lincomb: 1..%d terms", kLinCombMax), DGH_ERR_ARG;
    FieldDev* fo = c->field(out);
    if (!fo) return c->last;
    LinComb lc;
    lc.n = n;
    for (int t = 0; t < n; ++t) {
        FieldDev* f = c->field(fields[t]);
        if (!f) return c->last;
        if (f->n != fo->n)
            return c->fail(DGH_ERR_ARG, "dgh_field_lincomb: field %d differs in size", fields[t]), DGH_ERR_ARG;
        lc.c[t] = coefs[t];
        lc.x[t] = f->d;
    }
    (void)hipSetDevice(c->device);
    return c->hip(launch_lincomb(lc, fo->n, fo->d, c->ncu, c->stream), "lincomb") ? DGH_OK : c->last;
}

int dgh_field_copy_var(dgh_ctx c, int dst, int dvar, int src, int svar) {
    if (!c) return DGH_ERR_ARG;
    FieldDev* a = c->field(dst);
    FieldDev* b = c->field(src);
    if (!a || !b) return c->last;
    if (a->K != b->K || a->Np != b->Np || dvar < 0 || dvar >= a->nvar || svar < 0 || svar >= b->nvar ||
        (dst == src && dvar == svar))
        return c->fail(DGH_ERR_ARG, "dgh_field_copy_var: fields of different sizes or bad variable"), DGH_ERR_ARG;
    (void)hipSetDevice(c->device);
    return c->hip(launch_copy_var(a->K, a->Np, b->d, b->nvar, svar, a->d, a->nvar, dvar, c->stream), "copy var")
               ? DGH_OK
               : c->last;
}

}  // extern "C"
