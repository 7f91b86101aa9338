"""CPU model of the brick layout's AA-pattern stepping, used by the tests to check the layout design (parity
forms, redirected writes at boundaries, the side copies for Bouzidi q >= 1/2, and the conversions) against the oracle.

It uses the same host layout arrays as the HIP kernels (cerebrosim_amd.bricks.build) and the same rules as
csrc/k_bricks.hip; addressing goes through the reference neighbour table and slot_of_cell (the builder verifies
that these agree with the brick neighbour table). The collision is a vectorised FP64 statement of lbm.py:183-240.

storage="fp16" models the FP16 form: a slot holds float16(1024 (P - feq2(writer's moments)) / w) of the population
P it carries, feq2 / w = drho - 1.5 |u|^2 + 3 c.u + 4.5 (c.u)^2; a fluid link i of n carries P_i(n - c_i), a
missing one the bounced P_opp(i)(n), each decoded with its writer's moments of the previous step.
"""
import numpy as np

from cerebrosim_amd import lattice as L

C = L.C.astype(np.float64)
W = L.W
OPP = L.OPP


def h3(c, u):
    t = 1.0 / 3.0
    return (13.5 * ((c[0] ** 2 - t) * (c[1] * u[0] ** 2 * u[1] + c[2] * u[0] ** 2 * u[2])
                    + (c[1] ** 2 - t) * (c[0] * u[0] * u[1] ** 2 + c[2] * u[1] ** 2 * u[2])
                    + (c[2] ** 2 - t) * (c[0] * u[0] * u[2] ** 2 + c[1] * u[1] * u[2] ** 2))
            + 27.0 * c[0] * c[1] * c[2] * u[0] * u[1] * u[2])


class AAModel:
    def __init__(self, lat, layout, nu, force=(0, 0, 0), p_out=(0.0,), bouzidi=True, storage="fp64"):
        self.fp16 = storage == "fp16"
        self.bounced_dir = True
        self.lat, self.Lay = lat, layout
        self.nbr = lat["nbr"]
        self.q = lat["q"].astype(np.float64)
        self.uin = lat["uin"].astype(np.float64)
        self.N = self.nbr.shape[1]
        self.slot = layout["slot_of_cell"].astype(np.int64)
        from cerebrosim_amd import bricks
        self.slots = bricks.slot_table().astype(np.int64)
        self.S = np.zeros((layout["n_slots"] // 64 * 27 * 64,))
        self.side = np.zeros((27, self.N))
        self.om = np.full(self.N, 1.0 / (3 * nu + 0.5))
        self.F = np.asarray(force, float)
        self.p_out = np.asarray(p_out, np.float32).astype(float)
        self.bouzidi = bouzidi
        self.vel = np.zeros((self.N, 3))
        self.rho = np.zeros(self.N)
        self.inl = (self.nbr == L.INLET).any(0)
        self.fluid = self.nbr >= 0
        self.sm = (self.nbr == L.WALL) & (self.q >= 0.5) & self.fluid[OPP]

    def addr(self, cells, k):
        s = self.slot[cells]
        return (s // 64) * 27 * 64 + self.slots[k][s % 64]

    def loc(self, i, par):
        """Address of P_i(n) for all n in the state read by a step of parity `par`."""
        n = np.arange(self.N)
        io = OPP[i]
        a = self.addr(n, io)
        if par == 0 and i != 0:
            m = self.fluid[io]
            a[m] = self.addr(self.nbr[io, m], i)
        return a

    # FP16 form: population value <-> stored value for direction d with the writer cells' moments
    def _eq(self, d, cells):
        u = self.vel[cells]
        cu = u @ C[d]
        return self.rho[cells] - 1.5 * (u ** 2).sum(1) + cu * (3 + 4.5 * cu)

    def encode(self, v, d, cells):
        if not self.fp16:
            return v
        return ((v / W[d] - self._eq(d, cells)) * 1024.0).astype(np.float32).astype(np.float16).astype(np.float64)

    def decode(self, h, d, cells):
        if not self.fp16:
            return h
        return W[d] * (h / 1024.0 + self._eq(d, cells))

    def set_f(self, f, par):
        n = np.arange(self.N)
        for i in range(27):
            self.S[self.loc(i, par)] = self.encode(f[i], i, n)
        self.side[self.sm] = f[self.sm]

    def get_f(self, par):
        n = np.arange(self.N)
        return np.stack([self.decode(self.S[self.loc(i, par)], i, n) for i in range(27)])

    def step(self, par, drive=1.0):
        N, n = self.N, np.arange(self.N)
        raw = np.empty((27, N))
        for i in range(27):
            a = self.addr(n, i)
            if par == 1 and i != 0:
                m = self.fluid[i]
                a[m] = self.addr(self.nbr[i, m], OPP[i])
            raw[i] = self.S[a]
            if self.fp16:
                fl = self.fluid[i] if i != 0 else np.ones(N, bool)
                wr = np.where(fl, self.nbr[i], n) if i != 0 else n
                dec_i = self.decode(raw[i], i, wr)
                # bounced_dir False plants a mistake for the tests: decode a bounced population with direction i
                dec_o = self.decode(raw[i], OPP[i] if self.bounced_dir else i, n)
                raw[i] = np.where(fl, dec_i, dec_o)
        f = raw.copy()
        up = self.vel
        upsq = (up ** 2).sum(1)
        for i in range(1, 27):
            io = OPP[i]
            miss = ~self.fluid[i]
            code = self.nbr[i]
            fo = raw[i]
            v = f[i]
            wall = miss & (code == L.WALL)
            v = np.where(wall, fo, v)
            if self.bouzidi:
                q = self.q[i]
                lo = wall & (q < 0.5) & self.fluid[io]
                v = np.where(lo, 2 * q * fo + (1 - 2 * q) * raw[io], v)
                hi = wall & (q >= 0.5)
                pself = np.where(self.sm[i], self.side[i], raw[io])
                v = np.where(hi, fo / (2 * q) + (2 * q - 1) / (2 * q) * pself, v)
            inlet = miss & (code == L.INLET)
            v = np.where(inlet, fo + 6 * W[i] * (self.uin * drive) @ C[i], v)
            outl = miss & (code <= L.OUTLET0)
            k = np.clip(L.OUTLET0 - code, 0, len(self.p_out) - 1)
            cu = up @ C[i]
            v = np.where(outl, -fo + 2 * W[i] * (3 * self.p_out[k] + 4.5 * cu * cu - 1.5 * upsq), v)
            f[i] = v
        post, u, dr = self.collide(f, drive)
        self.rho = dr
        self.vel = u
        for i in range(27):
            io = OPP[i]
            a = self.addr(n, io)
            if par == 1 and i != 0:
                m = self.fluid[io]
                a[m] = self.addr(self.nbr[io, m], i)
            self.S[a] = self.encode(post[i], i, n)
        self.side[self.sm] = post[self.sm]

    def ddx(self, ax, comp):
        up = self.vel
        out = np.empty(self.N)
        i1, i2 = 2 * ax + 1, 2 * ax + 2
        u0 = up[:, comp]
        h1, v1, h2, v2 = np.ones(self.N), u0.copy(), np.ones(self.N), u0.copy()
        for i, h, v in ((i1, h1, v1), (i2, h2, v2)):
            s = self.nbr[i]
            m = s >= 0
            v[m] = up[s[m], comp]
            w = s == L.WALL
            h[w] = np.maximum(self.q[i][w], 0.1)
            v[w] = 0.0
        out[:] = (h1 ** 2 * (v2 - u0) + h2 ** 2 * (u0 - v1)) / (h1 * h2 * (h1 + h2))
        return out

    def collide(self, f, drive):
        om = self.om
        F = self.F * drive
        dr = f.sum(0)
        u = (C.T @ f).T + F / 2
        usq = (u ** 2).sum(1)
        uF = u @ F
        g = np.stack([self.ddx(a, a) for a in range(3)], 1)
        E = 3 / om[:, None] * u ** 2 * g
        cu = u @ C.T                                           # (N, 27)
        feq2 = W[None] * (dr[:, None] + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq[:, None])
        P = np.einsum("ia,ib,ni->nab", C, C, f.T - feq2)
        P[:, [0, 1, 2], [0, 1, 2]] -= E
        trP = np.trace(P, axis1=1, axis2=2) / 3
        cF = F @ C.T
        cPc = np.einsum("ia,nab,ib->ni", C, P, C)
        h = np.stack([h3(C[i], u.T) for i in range(27)], 1)
        feq = W[None] * (dr[:, None] + 3 * cu + 4.5 * cu ** 2 - 1.5 * usq[:, None] + h)
        fneq = W[None] * (4.5 * (cPc - trP[:, None]) - 1.5 * cF[None])
        src = W[None] * (3 * cF[None] + 9 * cu * cF[None] - 3 * uF[:, None])
        post = (feq + (1 - om[:, None]) * fneq + (1 - 0.5 * om[:, None]) * src).T
        # inlet cells: equilibrium at the prescribed velocity
        ut = self.uin * drive
        utsq = (ut ** 2).sum(1)
        cut = ut @ C.T
        ht = np.stack([h3(C[i], ut.T) for i in range(27)], 1)
        eq_in = (W[None] * (dr[:, None] + 3 * cut + 4.5 * cut ** 2 - 1.5 * utsq[:, None] + ht)).T
        post[:, self.inl] = eq_in[:, self.inl]
        u[self.inl] = ut[self.inl]
        return post, u, dr
