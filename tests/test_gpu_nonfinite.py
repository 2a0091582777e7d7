"""Non-finite states while stepping (channelflow/cfdsi.cpp:707-710, 769-774:
the reference stops when L2Norm(u) or L2Norm(f^T(u)) is not finite). An
explicit step far past the stability limit (dt = 2 at 16x17x16, norm 0.3)
drives the velocity to NaN/inf within a few steps; advance then reports
CFG_ERR_NONFINITE, in a Batch it names the member, and a well-behaved member
of the same Batch is unaffected."""
import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu
LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5
FLAGS = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
             nu=1 / 400, timestepping="SBDF3", initstepping="SMRK2", nonlinearity="Rotational",
             dealiasing="DealiasXZ")
G = (16, 17, 16)


def fields(u):
    return [cfa.FlowField(*G, 3, LX_C, LZ_C, data=u), cfa.FlowField(*G, 1, LX_C, LZ_C)]


def test_dns_reports_nonfinite(gpu):
    u = random_field(*G, LX_C, LZ_C, norm=0.3, decay=0.5, seed=5)
    dns = cfa.DNS(fields(u), cfa.DNSFlags(**dict(FLAGS, dt=2.0)))
    with pytest.raises(cfa.CfgError) as e:
        dns.advance_resident(60)
    assert e.value.code == cfa.ERR_NONFINITE
    ug, _ = dns.state()
    assert not np.all(np.isfinite(ug))


def test_batch_names_nonfinite_member(gpu):
    u = random_field(*G, LX_C, LZ_C, norm=0.3, decay=0.5, seed=5)
    b = cfa.Batch()
    good = b.add(fields(0.01 * u), cfa.DNSFlags(**dict(FLAGS, dt=0.02)))
    bad = b.add(fields(u), cfa.DNSFlags(**dict(FLAGS, dt=2.0)))
    with pytest.raises(cfa.CfgError) as e:
        b.advance(60)
    assert e.value.code == cfa.ERR_NONFINITE
    assert f"member(s) {bad}" in str(e.value)
    ug, _ = b.state(good)
    assert np.all(np.isfinite(ug))
    ref = cfa.DNS(fields(0.01 * u), cfa.DNSFlags(**dict(FLAGS, dt=0.02)))
    ref.advance_resident(60)
    ur, _ = ref.state()
    assert np.array_equal(ug, ur)
