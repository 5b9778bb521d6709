"""Differentiable FP64 restatement of the time loop in PyTorch, used to check
the engine's adjoint gradient with autograd on tiny 3D grids.

TEST INFRASTRUCTURE ONLY (imported by tests/).  Each step follows the
operations of reference code/driver.f90:355-450 in the order the engine
performs them:
  E update        code/fdtd.f90:78-147 (curl H, regions I / structure / III),
                  z-UPML code/boundary_II.f90:84-99;
  max-side wraps  code/boundary.f90:193-348 (E(n+1) = exp(+i ak) E(1));
  H update        code/fdtd.f90:182-195, z-UPML code/boundary_II.f90:170-183,
                  from the E of this step before its TF/SF correction;
  min-side wraps  H(0) = exp(-i ak) H(n);
  TF/SF           code/tfsf.f90:76-119 (3D grids) with the incident field of
                  code/tfsf.f90:327-350, on the cells a later step reads;
  DFT             code/dataprocess.f90:138-150, emg += dt F exp(i w n');
  flux            code/dataprocess.f90:241-277.
Only 3D grids (nx, ny > 1) and the Gaussian plane-wave source are covered.
The flux of this restatement is checked against the oracle before its
gradient is trusted.
"""
import numpy as np
import torch


def _pml_tables(setup):
    """Per-plane UPML coefficients (nz+2, 3) for every z, zero outside the
    slabs (the engine's table of reference code/boundary_II.f90:86-99,
    :170-183 subexpressions)."""
    p = setup.params
    nz = p.nz
    dtq2 = (p.dt / p.q) * (p.dt / p.q)
    names = ["fp", "fm", "cd", "gd", "gb", "ce", "ge", "gh"]
    t = {k: np.zeros((nz + 2, 3)) for k in names}
    for iz in range(1, nz + 1):
        if iz <= p.pmlzi:
            fp = setup.fpzmn[3 * (iz - 1):3 * iz]
            fm = setup.fmzmn[3 * (iz - 1):3 * iz]
            ep = p.eps[0]
        elif iz >= p.pmlzf:
            fp = setup.fpzmx[3 * (iz - p.pmlzf):3 * (iz - p.pmlzf + 1)]
            fm = setup.fmzmx[3 * (iz - p.pmlzf):3 * (iz - p.pmlzf + 1)]
            ep = p.eps[2]
        else:
            continue
        for c in range(3):
            pd, pe = (c + 1) % 3, (c + 2) % 3
            t["fp"][iz, c] = fp[c]
            t["fm"][iz, c] = fm[c]
            t["cd"][iz, c] = fm[pd] / fp[pd]
            t["gd"][iz, c] = (2.0 * ep) / fp[pd]
            t["gb"][iz, c] = ((2.0 * ep) / fp[pd]) * dtq2
            t["ce"][iz, c] = fm[pe] / fp[pe]
            t["ge"][iz, c] = 1.0 / (p.eps0 * ep * fp[pe])
            t["gh"][iz, c] = 1.0 / (p.mu0 * fp[pe])
    return t


def _bloch(ak, sign):
    re = 0.0 * ak[0] - 1.0 * ak[1]
    im = 0.0 * ak[1] + 1.0 * ak[0]
    if sign < 0:
        re, im = -re, -im
    m = np.exp(re)
    return complex(m * np.cos(im), m * np.sin(im))


class TorchFDTD:
    """The time loop of one setup as a function of the structure's eps_inv."""

    def __init__(self, setup):
        p = setup.params
        if p.nx == 1 or p.ny == 1:
            raise ValueError("the PyTorch restatement covers 3D grids only")
        self.setup = setup
        self.p = p
        self.nx, self.ny, self.nz = p.nx, p.ny, p.nz
        self.dtcq2 = ((p.dt * p.c0) / p.q) * ((p.dt * p.c0) / p.q)
        self.e02 = p.eps0 * p.eps0
        pml = _pml_tables(setup)
        z = slice(1, p.nz + 1)

        def col(a):  # (3, nz, 1, 1)
            return torch.tensor(a[z].T.copy(), dtype=torch.float64).reshape(3, p.nz, 1, 1)

        self.pml = {k: col(v) for k, v in pml.items()}
        zz = np.arange(1, p.nz + 1)
        self.upml = torch.tensor((zz <= p.pmlzi) | (zz >= p.pmlzf)).reshape(1, p.nz, 1, 1)
        a1 = np.array([p.epsa_inv[0][i][i] for i in range(3)])
        a3 = np.array([p.epsa_inv[1][i][i] for i in range(3)])
        base = np.zeros((3, p.nz, 1, 1), dtype=np.complex128)
        for c in range(3):
            base[c, zz < p.hi, 0, 0] = a1[c]
            base[c, zz > p.hf, 0, 0] = a3[c]
        self.abase = torch.tensor(base)
        self.f = {k: _bloch(np.array(p.ak[i]), s) for k, i, s in
                  [("xp", 0, 1), ("yp", 1, 1), ("zp", 2, 1), ("xm", 0, -1), ("ym", 1, -1), ("zm", 2, -1)]}
        self.fw = np.array(setup.fw, dtype=np.float64)
        # TF/SF constants (reference code/tfsf.f90:262-290)
        t = {}
        t["ko"] = (2.0 * p.pi) / (p.wlpump * p.f_abohr)
        t["dk"] = p.gpump * p.q
        if p.iz_min > p.hi:
            eps_ms, t["sign"] = p.eps[2], -1.0
        else:
            eps_ms, t["sign"] = p.eps[0], 1.0
        t["sq"] = np.sqrt(eps_ms)
        t["c0dt"] = p.c0 * p.dt
        t["amp"] = p.eamp * p.q
        t["hamp"] = t["amp"] * ((p.dt * p.c0) / p.q)
        self.t = t

    # incident values (comps {1,3,5} for i_pol=2, {2,4,6} for i_pol=1) at
    # ix (array) and plane iz, reference code/tfsf.f90:301-350
    def _incident(self, it, ix, iz):
        p, t = self.p, self.t
        c0dt_it = t["c0dt"] * float(it)
        A = (p.q * (ix - p.ixo).astype(np.float64)) * p.ux + (p.q * (float(iz) - p.zopump)) * p.uz
        kzz = t["ko"] * ((t["sq"] * t["sign"]) * A - c0dt_it)
        tt = (A - c0dt_it / t["sq"]) / t["dk"]
        env = np.exp(-(tt * tt))
        cs, sn = np.cos(kzz), np.sin(kzz)
        us = p.uz * t["sign"]
        if p.i_pol == 2:
            ex = ((t["sign"] * cs) * env) * p.uz + 1j * (((t["sign"] * sn) * env) * p.uz)
            ez = -((cs * env) * p.ux) + 1j * (-((sn * env) * p.ux))
            hy = (us * ex.real - p.ux * ez.real) + 1j * (us * ex.imag - p.ux * ez.imag)
            if p.cosine == 1:
                ex, ez, hy = ex.real + 0j, ez.real + 0j, hy.real + 0j
            return [p.win * (t["amp"] * ex.real) + 1j * (p.win * (t["amp"] * ex.imag)),
                    p.win * (t["amp"] * ez.real) + 1j * (p.win * (t["amp"] * ez.imag)),
                    p.win * (t["hamp"] * hy.real) + 1j * (p.win * (t["hamp"] * hy.imag))]
        hx = -(((t["sign"] * cs) * env) * p.uz) + 1j * (-(((t["sign"] * sn) * env) * p.uz))
        hz = (cs * env) * p.ux + 1j * ((sn * env) * p.ux)
        ey = (p.ux * hz.real - us * hx.real) + 1j * (p.ux * hz.imag - us * hx.imag)
        if p.cosine == 1:
            ey, hz, hx = ey.real + 0j, hz.real + 0j, hx.real + 0j
        return [p.win * (t["amp"] * ey.real) + 1j * (p.win * (t["amp"] * ey.imag)),
                p.win * (t["hamp"] * hx.real) + 1j * (p.win * (t["hamp"] * hx.imag)),
                p.win * (t["hamp"] * hz.real) + 1j * (p.win * (t["hamp"] * hz.imag))]

    def _tfsf(self, it):
        """Constant corrections of E (interior) and H (interior and the ghost
        line a later step reads) after step it, as full-size tensors."""
        p = self.p
        nx, ny, nz = self.nx, self.ny, self.nz
        ce = np.zeros((3, nz + 2, ny + 2, nx + 2), dtype=np.complex128)
        ch = np.zeros_like(ce)
        if not p.em_source:
            return None, None
        ix = np.arange(0, nx + 1)
        if p.i_pol == 2:
            de = p.eps_in[0] * self._incident(it, ix, p.iz_min - 1)[2]     # Ex += eps_in*inc5
            ce[0, p.iz_min, 1:ny + 1, 1:nx + 1] = de[1:][None, :]
            dh = self.dtcq2 * self._incident(it, ix, p.iz_min)[0]          # Hy += dtcq2*inc1
            ch[1, p.iz_min - 1, 1:ny + 1, 0:nx + 1] = dh[None, :]
        else:
            de = p.eps_in[0] * self._incident(it, ix, p.iz_min - 1)[1]     # Ey -= eps_in*inc4
            ce[1, p.iz_min, 1:ny + 1, 1:nx + 1] = -de[1:][None, :]
            dh = self.dtcq2 * self._incident(it, ix, p.iz_min)[0]          # Hx -= dtcq2*inc2
            ch[0, p.iz_min - 1, 0:ny + 1, 1:nx + 1] = -dh[1:][None, :]
        return torch.tensor(ce), torch.tensor(ch)

    def _faces(self, nobox):
        p = self.p
        b = p.box[nobox - 1]
        bxi, bxf, byi, byf, bzi, bzf = [int(v) for v in b]
        X = slice(bxi, bxf + 1)
        Y = slice(byi, byf + 1)
        Z = slice(bzi, bzf + 1)
        # (pair, index tuple of the first and second face), comps (A, B, C, D)
        return [((Z, Y, bxi), (Z, Y, bxf), (1, 5, 2, 4), 0),
                ((Z, byi, X), (Z, byf, X), (2, 3, 0, 5), 1),
                ((bzi, Y, X), (bzf, Y, X), (0, 4, 1, 3), 2)]

    def run(self, w_struct, nobox=1):
        """Jabs (nof, 6) of box nobox after nt steps, as a differentiable
        function of w_struct, the complex eps_inv on the structure cells
        (hf-hi+1, ny, nx)."""
        p = self.p
        nx, ny, nz = self.nx, self.ny, self.nz
        P = self.pml
        cdt = torch.complex128
        A = torch.cat([torch.zeros(3, p.hi - 1, ny, nx, dtype=cdt),
                       w_struct.to(cdt).unsqueeze(0).expand(3, -1, -1, -1),
                       torch.zeros(3, nz - p.hf, ny, nx, dtype=cdt)], dim=1) + self.abase
        E = torch.zeros(3, nz, ny, nx, dtype=cdt)
        Hf = torch.zeros(3, nz + 2, ny + 2, nx + 2, dtype=cdt)   # H with ghosts
        D = torch.zeros(3, nz, ny, nx, dtype=cdt)
        B = torch.zeros(3, nz, ny, nx, dtype=cdt)
        f = self.f
        faces = self._faces(nobox)
        acc = [[None, None] for _ in range(3)]
        fw = torch.tensor(self.fw)
        I = (slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1))
        for it in range(1, p.nt + 1):
            # E update from H of the previous step (ghosts included)
            H1, H2, H3 = Hf[0], Hf[1], Hf[2]
            c0 = ((H3[I] - H3[1:nz + 1, 0:ny, 1:nx + 1]) - H2[I]) + H2[0:nz, 1:ny + 1, 1:nx + 1]
            c1 = ((H1[I] - H1[0:nz, 1:ny + 1, 1:nx + 1]) - H3[I]) + H3[1:nz + 1, 1:ny + 1, 0:nx]
            c2 = ((H2[I] - H2[1:nz + 1, 1:ny + 1, 0:nx]) - H1[I]) + H1[1:nz + 1, 0:ny, 1:nx + 1]
            c = torch.stack([c0, c1, c2])
            Dn = P["cd"] * D + P["gd"] * (self.e02 * c)
            e_pml = P["ce"] * E + P["ge"] * (P["fp"] * Dn - P["fm"] * D)
            e_blk = E + A * c
            En = torch.where(self.upml, e_pml, e_blk)
            D = torch.where(self.upml, Dn, D)
            # max-side images of E (the H update reads them)
            Ef = torch.cat([torch.zeros(3, nz, ny, 1, dtype=cdt), En, f["xp"] * En[..., :1]], dim=3)
            Ef = torch.cat([torch.zeros(3, nz, 1, nx + 2, dtype=cdt), Ef, f["yp"] * Ef[:, :, :1, :]], dim=2)
            Ef = torch.cat([torch.zeros(3, 1, ny + 2, nx + 2, dtype=cdt), Ef, f["zp"] * Ef[:, :1]], dim=1)
            E1, E2, E3 = Ef[0], Ef[1], Ef[2]
            d0 = ((E3[1:nz + 1, 2:ny + 2, 1:nx + 1] - E3[I]) - E2[2:nz + 2, 1:ny + 1, 1:nx + 1]) + E2[I]
            d1 = ((E1[2:nz + 2, 1:ny + 1, 1:nx + 1] - E1[I]) - E3[1:nz + 1, 1:ny + 1, 2:nx + 2]) + E3[I]
            d2 = ((E2[1:nz + 1, 1:ny + 1, 2:nx + 2] - E2[I]) - E1[1:nz + 1, 2:ny + 2, 1:nx + 1]) + E1[I]
            d = torch.stack([d0, d1, d2])
            Hi = Hf[:, 1:nz + 1, 1:ny + 1, 1:nx + 1]
            Bn = P["cd"] * B - P["gb"] * d
            h_pml = P["ce"] * Hi + P["gh"] * (P["fp"] * Bn - P["fm"] * B)
            h_blk = Hi - self.dtcq2 * d
            Hn = torch.where(self.upml, h_pml, h_blk)
            B = torch.where(self.upml, Bn, B)
            # min-side ghosts of H
            Hf = torch.cat([f["xm"] * Hn[..., -1:], Hn, torch.zeros(3, nz, ny, 1, dtype=cdt)], dim=3)
            Hf = torch.cat([f["ym"] * Hf[:, :, -1:, :], Hf, torch.zeros(3, nz, 1, nx + 2, dtype=cdt)], dim=2)
            Hf = torch.cat([f["zm"] * Hf[:, -1:], Hf, torch.zeros(3, 1, ny + 2, nx + 2, dtype=cdt)], dim=1)
            # the y ghost row copies row ny of the x-extended array, whose own
            # ghost columns are not read by the E update; likewise for z
            ce_, ch_ = self._tfsf(it)
            if ce_ is not None:
                En = En + ce_[:, 1:nz + 1, 1:ny + 1, 1:nx + 1]
                Hf = Hf + ch_
            E = En
            if it >= p.to_fourier and it % p.sample_fourier == 0:
                n = it + 1 - p.to_fourier
                ph = torch.polar(torch.ones_like(fw), fw * float(n))   # exp(i w n)
                Ef_s = torch.cat([torch.zeros(3, nz, ny, 1, dtype=cdt), E, torch.zeros(3, nz, ny, 1, dtype=cdt)],
                                 dim=3)
                Ef_s = torch.nn.functional.pad(Ef_s, (0, 0, 1, 1, 1, 1))
                F6 = torch.cat([Ef_s, Hf])  # (6, nz+2, ny+2, nx+2)
                for first, second, comps, j in faces:
                    for s, ix in enumerate((first, second)):
                        v = F6[(slice(None),) + ix].reshape(6, -1)
                        term = (p.dt * v).unsqueeze(0) * ph.reshape(-1, 1, 1)
                        acc[j][s] = term if acc[j][s] is None else acc[j][s] + term
        sflux = (p.eps0 * p.q) / (2.0 * p.dt)
        out = torch.zeros(p.nof, 6, dtype=torch.float64)
        cols = []
        for first, second, comps, j in faces:
            i1, i2, i3, i4 = comps
            for s in range(2):
                a = acc[j][s]
                if a is None:
                    cols.append(torch.zeros(p.nof, dtype=torch.float64))
                    continue
                v = (a[:, i1] * a[:, i2].conj()).real - (a[:, i3] * a[:, i4].conj()).real
                cols.append((sflux if s else -sflux) * v.sum(dim=1))
        out = torch.stack(cols, dim=1)
        return out
