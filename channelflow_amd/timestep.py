"""Variable time-step control (channelflow's TimeStep, dnsflags.h:149-198,
dnsflags.cpp:736-970), host side.

dt is kept in [dtmin, dtmax] and, when `variable`, steered so that the CFL
number returned by DNS.CFL stays in [CFLmin, CFLmax]; dt always divides the
reporting interval dT into n whole steps. Invalid bounds raise ValueError where
the reference exits.
"""


def _iround(x):  # round to the nearest integer, halves away from zero (cfbasics/mathdefs.h:208)
    return int(x + 0.5) if x > 0.0 else int(x - 0.5)


_INT_MAX = 2 ** 31 - 1


class TimeStep:
    def __init__(self, dt=None, dtmin=0.0, dtmax=0.0, dT=0.0, CFLmin=0.0, CFLmax=0.0, variable=True, flags=None):
        """TimeStep(dt, dtmin, dtmax, dT, CFLmin, CFLmax, variable) or
        TimeStep(flags=DNSFlags) (dnsflags.cpp:749-797)."""
        if flags is not None:
            dt, dtmin, dtmax, dT = flags.dt, flags.dtmin, flags.dtmax, flags.dT
            CFLmin, CFLmax, variable = flags.CFLmin, flags.CFLmax, flags.variabledt
        self.n_, self.N_, self.T_ = 0, 0, 0.0
        if dt is None:  # the empty TimeStep (dnsflags.cpp:736-747)
            self.dt_ = self.dtmin_ = self.dtmax_ = self.dT_ = 0.0
            self.CFLmin_ = self.CFL_ = self.CFLmax_ = 0.0
            self.variable_ = False
            return
        self.dt_, self.dtmin_, self.dtmax_, self.dT_ = float(dt), float(dtmin), float(dtmax), float(dT)
        self.CFLmin_, self.CFLmax_, self.variable_ = float(CFLmin), float(CFLmax), bool(variable)
        self.CFL_ = (self.CFLmax_ + self.CFLmin_) / 2  # meaningful after the first adjust
        if dtmin < 0 or dt < dtmin or dtmax < dt:
            raise ValueError(f"TimeStep: need 0 <= dtmin <= dt <= dtmax (dtmin={dtmin}, dt={dt}, dtmax={dtmax})")
        if CFLmin < 0 or CFLmax < CFLmin:
            raise ValueError(f"TimeStep: need 0 <= CFLmin <= CFLmax (CFLmin={CFLmin}, CFLmax={CFLmax})")
        if dT < dtmin:
            raise ValueError(f"TimeStep: need dT >= dtmin (dT={dT}, dtmin={dtmin})")
        # dt becomes a whole divisor of dT, then is moved into [dtmin, dtmax]
        self.n_ = max(_iround(self.dT_ / self.dt_), 1)
        self.dt_ = self.dT_ / self.n_
        while self.dt_ < self.dtmin_ and self.n_ >= 2 and self.dT_ != 0:
            self.n_ -= 1
            self.dt_ = self.dT_ / self.n_
        while self.dt_ > self.dtmax_ and self.n_ <= _INT_MAX and self.dT_ != 0:
            self.n_ += 1
            self.dt_ = self.dT_ / self.n_

    def _retarget(self, n):
        """dT / n moved into [dtmin, dtmax]: (n, dt)."""
        dt = self.dT_ / n
        while dt < self.dtmin_ and dt < self.dT_:
            n -= 1
            dt = self.dT_ / n
        while dt > self.dtmax_ and n <= _INT_MAX:
            n += 1
            dt = self.dT_ / n
        return n, dt

    def adjust(self, CFL, a_max=None, verbose=False):
        """adjust(CFL): if variable and CFL is outside (CFLmin, CFLmax), move dt so
        that CFL lands mid-range. adjust(a, a_max): bring a (any quantity
        proportional to dt) below a_max. Returns whether dt changed."""
        if a_max is not None:
            if self.variable_ and CFL >= a_max:
                return self.adjustToDesired(CFL, a_max)
            return False
        self.CFL_ = CFL
        if self.variable_ and (CFL <= self.CFLmin_ or CFL >= self.CFLmax_):
            return self.adjustToMiddle(CFL)
        return False

    def adjustToMiddle(self, CFL, verbose=False):
        if self.dtmin_ == self.dtmax_ or self.dT_ == 0.0:
            return False
        n, dt = self._retarget(max(_iround(2 * self.n_ * CFL / (self.CFLmax_ + self.CFLmin_)), 1))
        CFL *= dt / self.dt_
        if n == self.n_:
            return False
        self.n_, self.dt_, self.CFL_ = n, dt, CFL
        return True

    def adjustToDesired(self, a, a_des, verbose=False):
        if self.dtmin_ == self.dtmax_ or self.dT_ == 0.0:
            return False
        n, dt = self._retarget(max(_iround(self.n_ * a / a_des), 1))
        if n == self.n_:
            return False
        self.n_, self.dt_ = n, dt
        return True

    def adjust_for_T(self, T, verbose=False):
        """Make N dT = T and n dt = dT exactly (dnsflags.cpp:912-957)."""
        self.T_ = T
        if T < 0:
            raise ValueError("TimeStep.adjust_for_T: T < 0")
        if T == 0:
            changed = self.dt_ != 0
            self.dt_, self.n_, self.dT_, self.T_ = 0.0, 0, 0.0, 0.0
            return changed
        N = max(_iround(T / self.dT_), 1)
        dT = T / N
        n = max(_iround(dT / self.dt_), 1)
        dt = dT / n
        while dt < self.dtmin_ and n > 2 and dT != 0:
            n -= 1
            dt = dT / n
        while dt > self.dtmax_ and n <= _INT_MAX and dT != 0:
            n += 1
            dt = dT / n
        CFL = dt * self.CFL_ / self.dt_
        changed = dt != self.dt_
        self.n_, self.N_, self.dt_, self.dT_, self.CFL_ = n, N, dt, dT, CFL
        return changed

    def n(self):
        return self.n_

    def N(self):
        return self.N_

    def dt(self):
        return self.dt_

    def dtmin(self):
        return self.dtmin_

    def dtmax(self):
        return self.dtmax_

    def dT(self):
        return self.dT_

    def T(self):
        return self.T_

    def CFL(self):
        return self.CFL_

    def CFLmin(self):
        return self.CFLmin_

    def CFLmax(self):
        return self.CFLmax_

    def variable(self):
        return self.variable_

    def __float__(self):  # operator Real: dT / n
        return self.dT_ / self.n_
