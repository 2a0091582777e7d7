"""Batch client of the DNS: edge tracking's attractor tests and bisection.

channelflow's edge tracking (programs/edgetracking.cpp) decides whether a
field falls back to the laminar state or becomes turbulent by integrating it
until a cross-flow-energy or L2 threshold is crossed (CalcAttractor,
edgetracking.cpp:767-812, with attractor(), :836-893), and narrows the edge
between a laminar uL and a turbulent uH by bisection (Bisection, :363-459).
With several trajectories per bisection round (nBisecTrajectories > 1) the
reference tries the blends u = (1 - alpha_j) uL + alpha_j uH, j = 1, 2, ...,
one after another until one is not laminar. Each blend is an independent
integration, so here all of a round's blends advance together in one Batch;
the first such j, and so the new uL and uH, are the same.

With flags.variabledt each trajectory adapts its own dt to its CFL number
after every time unit, as CalcAttractor does (edgetracking.cpp:794-801); no
energy files or saved fields.
"""
import numpy as np

from .fields import l2norm2

LAMINAR, TURBULENT, NO_ATTRACTOR = "laminar", "turbulent", "neither laminar nor turbulent"


class AttractorFlags:
    """AttractorFlags (edgetracking.cpp:126-136, 814-822)."""

    def __init__(self, EcfH=0.0, EcfL=0.0, L2H=0.0, L2L=0.0, normToHeight=False, tMaxDecay=0):
        self.EcfH, self.EcfL, self.L2H, self.L2L = float(EcfH), float(EcfL), float(L2H), float(L2L)
        self.normToHeight, self.tMaxDecay = bool(normToHeight), int(tMaxDecay)


def ecf(u, a=-1.0, b=1.0):
    """Ecf(u) = L2Norm_uvw(u, false, true, true)^2 (diffops.cpp:3885-3931): the
    unnormalised y-integral of |v|^2 + |w|^2 summed over the live Fourier modes."""
    return (b - a) * l2norm2(np.asarray(u)[1:3])


def l2norm(u):
    """L2Norm(u) (diffops.cpp:415-487, normalised)."""
    return float(np.sqrt(l2norm2(np.asarray(u))))


def rel_l2_diff(u1, u2):
    """RelL2Diff = L2Dist(u1, u2) / (L2Norm(u1) + L2Norm(u2)) (edgetracking.cpp:110)."""
    return float(np.sqrt(l2norm2(np.asarray(u1) - np.asarray(u2)))) / (l2norm(u1) + l2norm(u2))


def classify(u, tDecay, lastEcf, aflags, a=-1.0, b=1.0):
    """attractor(u, tDecay, lastEcf, aflags) (edgetracking.cpp:836-893).
    Returns (attractor, new tDecay)."""
    useEcf = aflags.EcfH > 1e-12 and aflags.EcfL > 1e-12
    useL2 = aflags.L2H > 1e-12 and aflags.L2L > 1e-12
    if not (useEcf or useL2):
        raise ValueError("No valid attractor settings specified")
    lam = turb = True
    e = ecf(u, a, b) if (useEcf or aflags.tMaxDecay > 0) else 0.0
    if useEcf:
        val = e / (b - a) if aflags.normToHeight else e
        if val < aflags.EcfH:
            turb = False
        if val > aflags.EcfL:
            lam = False
    if useL2:
        val = l2norm(u)
        if aflags.normToHeight:
            val /= np.sqrt(b - a)
        if val < aflags.L2H:
            turb = False
        if val > aflags.L2L:
            lam = False
    if aflags.tMaxDecay > 0:
        tDecay = tDecay + 1 if e <= lastEcf else 0
        if tDecay > aflags.tMaxDecay:
            lam = True
    if turb and lam:
        raise ValueError("FlowField is both laminar and turbulent")
    return (TURBULENT if turb else LAMINAR if lam else NO_ATTRACTOR), tDecay


class EdgeTracker:
    """Attractor tests and bisection rounds of edge tracking over one grid,
    every integration of a call in one Batch.

    flags: DNSFlags of the integrations; a time unit of the attractor test is
    dt.n() = round(dT / dt) steps, as TimeStep gives it, and with
    flags.variabledt every trajectory's dt follows its CFL number. tMin, tMax:
    time units before the first test and at most (tMinAttractor,
    tMaxAttractor of the reference).
    """

    def __init__(self, Nx, Ny, Nz, Lx, Lz, flags, aflags, tMin=0, tMax=10000, a=-1.0, b=1.0):
        from .timestep import TimeStep
        self.Nx, self.Ny, self.Nz, self.Lx, self.Lz, self.a, self.b = Nx, Ny, Nz, Lx, Lz, a, b
        ts = TimeStep(flags=flags)
        self.timestep = ts
        self.nstep = ts.n()
        self.flags = flags.copy(dt=ts.dt())
        self.aflags = aflags
        self.tMin, self.tMax = int(tMin), int(tMax)
        self.integrations = 0

    def _fields(self, u):
        from . import FlowField
        return (FlowField(self.Nx, self.Ny, self.Nz, 3, self.Lx, self.Lz, self.a, self.b, data=u),
                FlowField(self.Nx, self.Ny, self.Nz, 1, self.Lx, self.Lz, self.a, self.b))

    def attractors(self, us):
        """CalcAttractor for every initial field of `us` (edgetracking.cpp:
        767-812): the initial pressure from PressureSolver, then steps of one
        time unit until the field is laminar or turbulent or tMax is reached.
        Returns [(attractor, time units, final u)]."""
        import copy
        from . import Batch, PressureSolver
        batch = Batch()
        live = {}
        for k, u in enumerate(us):
            fu, fq = self._fields(u)
            PressureSolver(fu, flags=self.flags).solve(fq, fu)
            mid = batch.add([fu, fq], self.flags)
            live[mid] = dict(k=k, t=0, tDecay=0, u=np.array(u, dtype=np.complex128), result=NO_ATTRACTOR,
                             dt=copy.deepcopy(self.timestep))
        self.integrations += len(us)
        out = [None] * len(us)
        while live:
            for st in live.values():
                st["t"] += 1
                st["lastEcf"] = ecf(st["u"], self.a, self.b)
            batch.advance_members({mid: st["dt"].n() for mid, st in live.items()})
            for mid in list(live):
                st = live[mid]
                if st["dt"].variable() and st["dt"].adjust(batch.CFL(mid)):  # edgetracking.cpp:797-801
                    batch.reset_dt(mid, st["dt"].dt())
                st["u"] = batch.state(mid)[0]
                if st["t"] > self.tMin:
                    st["result"], st["tDecay"] = classify(st["u"], st["tDecay"], st["lastEcf"], self.aflags,
                                                          self.a, self.b)
                if st["result"] != NO_ATTRACTOR or st["t"] >= self.tMax:
                    out[st["k"]] = (st["result"], st["t"], st["u"])
                    batch.remove(mid)
                    del live[mid]
        return out

    def bisection_round(self, uL, uH, nTraj=1):
        """One pass of Bisection's loop body (edgetracking.cpp:376-446).
        nTraj = 1: the midpoint (uH + uL)/2 replaces uL or uH. nTraj > 1: the
        blends alpha_j = j/(nTraj + 1), j = 1..nTraj (then up to nTraj + 11),
        are integrated together; with j* the first one that is not laminar,
        uH becomes blend j* and, for j* >= 3, uL becomes blend j* - 2. Returns
        (uL, uH, RelL2Diff(uL, uH))."""
        uL = np.asarray(uL, dtype=np.complex128)
        uH = np.asarray(uH, dtype=np.complex128)
        if nTraj == 1:
            uT = uH.copy()
            uT += uL
            uT *= 0.5
            a = self.attractors([uT])[0][0]
            if a == TURBULENT:
                uH = uT
            elif a == LAMINAR:
                uL = uT
            else:
                raise RuntimeError("Bisection did not converge")
            return uL, uH, rel_l2_diff(uL, uH)
        alphaStep = 1.0 / (nTraj + 1)

        def blend(j):  # uT = ((1 - alpha)/alpha uL + uH) alpha, in the reference's order
            alpha = j * alphaStep
            uT = uL.copy()
            uT *= (1.0 - alpha) / alpha
            uT += uH
            uT *= alpha
            return uT

        # the reference tries j = 1, 2, ... while the blend is laminar, at most
        # up to j = nTraj + 11, and takes the first one that is not
        jstar = None
        for js in (range(1, nTraj + 1), range(nTraj + 1, nTraj + 12)):
            res = self.attractors([blend(j) for j in js])
            for j, (a, _, _) in zip(js, res):
                if a != LAMINAR:
                    jstar = j
                    break
            if jstar is not None:
                break
        if jstar is None:
            raise RuntimeError("Did not find turbulent state")
        uT = blend(jstar)
        if jstar >= 3:
            uL = blend(jstar - 2)
        uH = uT
        return uL, uH, rel_l2_diff(uL, uH)

    def bisection(self, uL, uH, epsilon=1e-6, nTraj=1, max_rounds=100):
        """Bisection (edgetracking.cpp:363-459): rounds until RelL2Diff(uL, uH)
        <= epsilon. Returns (uL, uH, rounds)."""
        l2d = rel_l2_diff(uL, uH)
        rounds = 0
        while l2d > epsilon and rounds < max_rounds:
            uL, uH, l2d = self.bisection_round(uL, uH, nTraj)
            rounds += 1
        return uL, uH, rounds
