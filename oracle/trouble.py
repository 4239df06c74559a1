"""Troubled elements on the oracle's mesh: evolve_myln with its trouble
scores, DG/FV switches and redone substeps or steps, restated from the
reference (src/main/evolve/evolve.c:25-176, RungeKutta.c:19-344,
trouble.c:27-400) over the oracle's RHS, interpolation and mesh.

TEST INFRASTRUCTURE ONLY: the checker for the device's trouble machinery.
The product (nmesh_amd, libnmx.so) never imports this module.

State u is the oracle's flat [var][node] layout.  Switches keep every
element's n, so the node offsets never change and the registers keep their
layout.  The TROUBLE function is indicator(call, mesh, u, u_p) -> one value
in {1, 0, -1} per element; call counts the scoring calls from 0.
"""
import numpy as np

from oracle.orc import trouble_score

NOTROUBLES = 11    # evolve.h:60
NSUB = {"RK4": 4, "sspRK3": 3, "Euler": 1}


class OTroubled:
    def __init__(self, mesh, k, method, indicator, redo="substep", numflux="upwind"):
        self.m = mesh
        self.k = list(k)
        self.method = method
        self.indicator = indicator
        self.substep = redo == "substep"
        self.numflux = numflux
        self.score = np.zeros(mesh.nelem, dtype=np.int32)
        self.calls = 0
        self.redos = 0
        self.switches = 0
        self.log = []      # per scoring call: (u, u_p, scores after it)

    # -- pieces -----------------------------------------------------------
    def rhs(self, mesh, w, t):
        return mesh.rhs(w, t, self.k, self.numflux)

    def offsets(self):
        info = np.array([self.m.elm_info(e) for e in range(self.m.nelem)])
        return info[:, 3], info[:, 2] ** 3

    def blocks(self, e):
        """flat indices of element e's nodes, all variables"""
        off, np_ = self.offsets()
        nn = self.m.nnodes
        return (np.arange(5)[:, None] * nn + off[e] + np.arange(np_[e])[None, :]).ravel()

    def set_score(self, u, up):
        """evolve_set_trouble_score_mesh (trouble.c:112-127)"""
        tr = np.asarray(self.indicator(self.calls, self.m, u, up), dtype=np.int32)
        self.calls += 1
        s = trouble_score(self.score, tr)
        self.log.append((u.copy(), up.copy(), self.score.copy()))
        return s

    def switch(self, fv, regs):
        """p-refinement to uniform or LGL points at the same n
        (evolve_switch_troubled_nodes_mesh / _nontroubled_, trouble.c:
        186-330): the mesh switched, each register interpolated"""
        old = self.m
        out = []
        for r in regs:
            new, rn = old.adapted(r, fv=np.asarray(fv, dtype=np.int32))
            out.append(rn)
        self.m = new
        self.switches += 1
        return out

    def redo(self, rfac, u, up, w, r, t):
        """evolve_trouble_redo_u_step_mesh (RungeKutta.c:223-330): each
        element with a positive score takes its substep back, becomes FV,
        gets r again from its interpolated w against its neighbours' w as
        before any switch, and redoes the substep"""
        M0 = self.m
        fv0 = M0.fv_flags()
        troubled = [e for e in range(M0.nelem) if self.score[e] > 0]
        for e in troubled:
            ix = self.blocks(e)
            u[ix] -= rfac * r[ix]
        rnew = r.copy()
        for e in troubled:
            fe = fv0.copy()
            fe[e] = 1
            Me, we = M0.adapted(w, fv=fe)
            re = self.rhs(Me, we, t)
            ix = self.blocks(e)
            rnew[ix] = re[ix]
        fv = fv0.copy()
        fv[troubled] = 1
        u, up, w = self.switch(fv, [u, up, w])
        for e in troubled:
            ix = self.blocks(e)
            u[ix] += rfac * rnew[ix]
        self.redos += 1
        return u, up, w, rnew

    def maybe_redo(self, redo, rfac, u, up, w, r, t):
        if redo and self.set_score(u, up) > 0:
            return self.redo(rfac, u, up, w, r, t)
        return u, up, w, r

    # -- Evolve_mesh (RungeKutta.c:19-188) ------------------------------------
    def evolve_mesh(self, u, t, dt, redo):
        up = u.copy()
        w = u.copy()
        if self.method == "RK4":
            r = self.rhs(self.m, w, t)
            u = w + dt / 6. * r
            u, up, w, r = self.maybe_redo(redo, dt / 6., u, up, w, r, t)
            w = up + dt / 2. * r
            r = self.rhs(self.m, w, t + 0.5 * dt)
            u = u + dt / 3. * r
            u, up, w, r = self.maybe_redo(redo, dt / 3., u, up, w, r, t + 0.5 * dt)
            w = up + dt / 2. * r
            r = self.rhs(self.m, w, t + 0.5 * dt)
            u = u + dt / 3. * r
            u, up, w, r = self.maybe_redo(redo, dt / 3., u, up, w, r, t + 0.5 * dt)
            w = up + dt * r
            r = self.rhs(self.m, w, t + dt)
            u = u + dt / 6. * r
            u, up, w, r = self.maybe_redo(redo, dt / 6., u, up, w, r, t + dt)
        elif self.method == "sspRK3":
            r = self.rhs(self.m, w, t)
            u = w + dt / 6. * r
            u, up, w, r = self.maybe_redo(redo, dt / 6., u, up, w, r, t)
            w = up + dt * r
            r = self.rhs(self.m, w, t + dt)
            u = u + dt / 6. * r
            u, up, w, r = self.maybe_redo(redo, dt / 6., u, up, w, r, t + dt)
            w = w + dt * r
            w = 0.75 * up + 0.25 * w
            r = self.rhs(self.m, w, t + 0.5 * dt)
            u = u + dt * 2. / 3. * r
            u, up, w, r = self.maybe_redo(redo, dt * 2. / 3., u, up, w, r, t + 0.5 * dt)
        elif self.method == "Euler":
            r = self.rhs(self.m, w, t)
            u = w + dt * r
            u, up, w, r = self.maybe_redo(redo, dt, u, up, w, r, t)
        else:
            raise ValueError(self.method)
        return u, up

    # -- evolve_myln (evolve.c:25-176) -----------------------------------------
    def step(self, u, t, dt):
        """one step from time t; returns the new state"""
        notroubles = NOTROUBLES * (NSUB[self.method] if self.substep else 1)
        u1, up = self.evolve_mesh(u, t, dt, self.substep)
        if self.substep:
            score = trouble_score(self.score)
        else:
            score = self.set_score(u1, up)
            if score > 0:
                # evolve_prepare_do_over_mesh: u = u_p, troubled elements
                # to FV, the step again
                fv = self.m.fv_flags()
                fv[self.score > 0] = 1
                (u0,) = self.switch(fv, [up])
                self.redos += 1
                u1, _ = self.evolve_mesh(u0, t, dt, False)
        if score > 0 or score <= -notroubles:
            # evolve_switch_nontroubled_nodes_mesh
            fv = self.m.fv_flags()
            back = (fv != 0) & (self.score <= -notroubles)
            if back.any():
                fv[back] = 0
                (u1,) = self.switch(fv, [u1])
        return u1
