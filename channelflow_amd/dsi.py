"""Batch clients of the DNS: channelflow's dynamical-systems interface for
periodic orbits by multiple shooting, with the integrations batched.

channelflow's Newton-Krylov search for periodic orbits evaluates
    G_i = sigma f^T(u_i) - u_{i+1}      (i = 0..n-1, u_n = u_0, sigma on the last shot)
(MultishootingDSI::eval, nsolver/multiShootingDSI.cpp:70-99, over cfDSI::eval,
channelflow/cfdsi.cpp:63-90, with f^T = the DNS integration over time T,
cfdsi.cpp:693-774) and, for every Krylov direction dx, the finite-difference
Jacobian product (G(x + eps dx) - G(x)) / eps with eps = epsDx / |dx|
(DSI::Jacobian, nsolver/dsi.cpp:66-88; multiShootingDSI.cpp:101-124). The
reference runs each of those integrations as its own DNS, one after another.
Here every integration of a call (the n shots and the n shots of each
perturbed copy) is one member of a single Batch, so they advance together.

Scope: fixed dt (the reference's f^T after TimeStep::adjust_for_T), no
Poincare section, no phase conditions, T fixed. The vector form of a field
(field2vector, channelflow/flowfield.cpp:4448-4563) is provided for the step
size eps, which the reference takes from the vector's 2-norm.
"""
import numpy as np

from .timestep import TimeStep


def field2vector_size(Nx, Ny, Nz):
    """Length of field2vector's vector (flowfield.cpp:4448-4479)."""
    Kx, Kz = Nx // 3 - 1, Nz // 3 - 1
    per_mode = 2 * (Ny - 2) + 2 * (Ny - 4)
    return 2 * (Ny - 2) + Kx * per_mode + Kz * per_mode + 2 * Kx * Kz * per_mode


def field2vector(u):
    """The linearly independent real coefficients of a divergence-free no-slip
    velocity field, in the reference's order (flowfield.cpp:4481-4563, after
    Gibson, Halcrow & Cvitanovic, table 1). u: spectral velocity, shape
    (3, Ny, Nx, Nz/2+1) complex, indexed [i, ny, mx, mz]."""
    d = np.asarray(u)
    _, Ny, Nx, _ = d.shape
    Nz = 2 * (d.shape[3] - 1)
    Kx, Kz = Nx // 3 - 1, Nz // 3 - 1
    mx = lambda kx: kx if kx >= 0 else Nx + kx  # noqa: E731
    out = []
    out.append(d[0, 2:Ny, 0, 0].real)
    out.append(d[2, 2:Ny, 0, 0].real)

    def reim(a):  # Re, Im interleaved over ny
        r = np.empty(2 * a.size)
        r[0::2], r[1::2] = a.real, a.imag
        return r

    for kx in range(1, Kx + 1):
        out.append(reim(d[2, 2:Ny, mx(kx), 0]))
        out.append(reim(d[0, 3:Ny - 1, mx(kx), 0]))
    for kz in range(1, Kz + 1):
        out.append(reim(d[0, 2:Ny, 0, kz]))
        out.append(reim(d[2, 3:Ny - 1, 0, kz]))
    for kx in range(-Kx, Kx + 1):
        if kx == 0:
            continue
        for kz in range(1, Kz + 1):
            out.append(reim(d[0, 2:Ny, mx(kx), kz]))
            out.append(reim(d[2, 3:Ny - 1, mx(kx), kz]))
    v = np.concatenate(out)
    assert v.size == field2vector_size(Nx, Ny, Nz)
    return v


def field2vector_norm2(u):
    """Squared 2-norm of field2vector(u), summed over the same coefficients
    without building the vector."""
    d = np.asarray(u)
    _, Ny, Nx, _ = d.shape
    Nz = 2 * (d.shape[3] - 1)
    Kx, Kz = Nx // 3 - 1, Nz // 3 - 1
    sq = lambda a: float(np.sum(a.real * a.real + a.imag * a.imag))  # noqa: E731
    s = float(np.sum(d[0, 2:Ny, 0, 0].real ** 2) + np.sum(d[2, 2:Ny, 0, 0].real ** 2))
    if Kx > 0:
        s += sq(d[2, 2:Ny, 1:Kx + 1, 0]) + sq(d[0, 3:Ny - 1, 1:Kx + 1, 0])
    if Kz > 0:
        s += sq(d[0, 2:Ny, 0, 1:Kz + 1]) + sq(d[2, 3:Ny - 1, 0, 1:Kz + 1])
    if Kx > 0 and Kz > 0:
        mxs = np.r_[Nx - Kx:Nx, 1:Kx + 1]
        s += sq(d[0, 2:Ny][:, mxs, 1:Kz + 1]) + sq(d[2, 3:Ny - 1][:, mxs, 1:Kz + 1])
    return s


def fd_step(dx_fields, epsDx):
    """eps of DSI::Jacobian (dsi.cpp:68-69) for the multi-shot direction whose
    per-shot fields are dx_fields: the 2-norm of the stacked vectors."""
    mag = np.sqrt(sum(field2vector_norm2(f) for f in dx_fields))
    return 1.0 if mag < epsDx else epsDx / mag


class MultiShooting:
    """Multiple-shooting residual and Jacobian products of plane channel flow,
    every integration of a call batched.

    flags: DNSFlags of the integrations (flags.dt is the requested step; as in
    cfDSI's f^T, TimeStep::adjust_for_T fits it to T). With flags.variabledt
    each integration adapts its own dt to its CFL number every dT, as f^T does
    (cfdsi.cpp:733-767): a first step probes the CFL and restarts from u with
    the adjusted dt, then every dT the CFL at its start decides a reset_dt
    after it. T: the integration time of each shot. symmetry: None (identity)
    or a callable applied to the last shot's f^T(u). Tnormalize: divide G by T
    (cfdsi.cpp:80-81).
    """

    def __init__(self, Nx, Ny, Nz, Lx, Lz, flags, T, a=-1.0, b=1.0, symmetry=None, Tnormalize=False):
        from . import DNSFlags  # noqa: F401  (flags is a DNSFlags)
        if T <= 0:
            raise ValueError("MultiShooting: T must be positive")
        self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b = Nx, Ny, Nz, Lx, Lz, a, b
        ts = TimeStep(flags=flags)
        ts.adjust_for_T(T)
        self.T = float(T)
        self.timestep = ts
        self.variable = bool(flags.variabledt)
        self.nsteps = ts.N() * ts.n()  # steps per integration at fixed dt
        self.flags = flags.copy(dt=ts.dt())
        self.symmetry = symmetry
        self.Tnormalize = bool(Tnormalize)
        self.fcount = 0  # integrations run, as the reference's fcount
        self.dts = []    # final dt of each integration of the last call

    # -- integrations ------------------------------------------------------
    def integrate(self, fields):
        """f^T(u) for every velocity field in `fields` (spectral arrays), all
        advanced together in one Batch."""
        import copy
        from . import Batch, FlowField
        batch = Batch()
        ids = []
        for u in fields:
            fu = FlowField(self.Nx, self.Ny, self.Nz, 3, self.Lx, self.Lz, self.a, self.b, data=u)
            fq = FlowField(self.Nx, self.Ny, self.Nz, 1, self.Lx, self.Lz, self.a, self.b)
            ids.append(batch.add([fu, fq], self.flags))
        self.fcount += len(fields)
        if not self.variable:
            batch.advance(self.nsteps)
            self.dts = [self.flags.dt] * len(ids)
            return [batch.state(m)[0] for m in ids]
        # cfdsi.cpp:733-737: one probe step, dt adjusted to its CFL, restart from u
        dts = [copy.deepcopy(self.timestep) for _ in ids]
        batch.advance(1)
        for m, dt, u in zip(ids, dts, fields):
            dt.adjust(batch.CFL(m))
            batch.reset_dt(m, dt.dt())
            batch.set_state(m, [FlowField(self.Nx, self.Ny, self.Nz, 3, self.Lx, self.Lz, self.a, self.b, data=u)])
        # cfdsi.cpp:747-767: N intervals dT, each n = dT/dt steps of the member's own dt
        for _ in range(self.timestep.N()):
            cfl = {m: batch.CFL(m) for m in ids}
            batch.advance_members({m: dt.n() for m, dt in zip(ids, dts)})
            for m, dt in zip(ids, dts):
                if dt.adjust(cfl[m]):
                    batch.reset_dt(m, dt.dt())
        self.dts = [dt.dt() for dt in dts]
        return [batch.state(m)[0] for m in ids]

    def _residual(self, shots, finals):
        n = len(shots)
        G = []
        for i in range(n):
            fu = finals[i]
            if i == n - 1 and self.symmetry is not None:
                fu = self.symmetry(fu)
            g = fu - shots[(i + 1) % n]
            if self.Tnormalize:
                g = g * (1.0 / self.T)
            G.append(g)
        return G

    # -- DSI operations ----------------------------------------------------
    def eval(self, shots):
        """G(x) for the shots u_0..u_{n-1} (multiShootingDSI.cpp:70-86)."""
        return self._residual(shots, self.integrate(shots))

    def eval_with_jacobians(self, shots, directions, epsDx=1e-7, centdiff=False):
        """G(x) and the Jacobian products DG dx_k for every direction k (each a
        list of per-shot fields), all n (1 + K) integrations (n (1 + 2K) with
        centred differences) in one Batch. Returns (G, [DG dx_k], [eps_k])."""
        n = len(shots)
        work = list(shots)
        eps, pert = [], []  # per direction: eps, and the perturbed shots (x + eps dx; centred: +/- eps/2 dx)
        for dx in directions:
            if len(dx) != n:
                raise ValueError("a direction needs one field per shot")
            e = fd_step(dx, epsDx)
            eps.append(e)
            if centdiff:
                xs = ([shots[i] + (0.5 * e) * dx[i] for i in range(n)], [shots[i] - (0.5 * e) * dx[i] for i in range(n)])
            else:
                xs = ([shots[i] + e * dx[i] for i in range(n)],)
            pert.append(xs)
            for x in xs:
                work += x
        finals = self.integrate(work)
        G = self._residual(shots, finals[:n])
        out = []
        p0 = n
        for k, xs in enumerate(pert):
            e = eps[k]
            Gs = []
            for x in xs:
                Gs.append(self._residual(x, finals[p0:p0 + n]))
                p0 += n
            if centdiff:
                out.append([(1 / e) * (Gs[0][i] - Gs[1][i]) for i in range(n)])
            else:
                out.append([(1 / e) * (Gs[0][i] - G[i]) for i in range(n)])
        return G, out, eps
