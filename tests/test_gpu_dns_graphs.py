"""hipGraph replay of a single DNS's repeating steps (cfg_dns_set_graphs) against
eager launches: bit for bit, through the SMRK2 start, the SBDF3 steps and a
reset_dt (DNS::reset_dt, dns.cpp:180-185: new tau data, the initialisation
steps again, so new launch sequences to capture), for the skew-symmetric
and rotational forms."""
import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.fields import random_field

pytestmark = pytest.mark.gpu
LX_C = 2 * np.pi / 1.14
LZ_C = 2 * np.pi / 2.5
G = (32, 33, 32)


@pytest.mark.parametrize("nonlinearity", ["SkewSymmetric", "Rotational"])
def test_dns_graph_replay_matches_eager(gpu, nonlinearity):
    fl = dict(baseflow="LaminarBase", ulowerwall=-1.0, uupperwall=1.0, constraint="BulkVelocity", Ubulk=0.0,
              nu=1 / 400, dt=0.02, timestepping="SBDF3", initstepping="SMRK2", nonlinearity=nonlinearity,
              dealiasing="DealiasXZ")
    u0 = random_field(*G, LX_C, LZ_C, norm=0.1, decay=0.3, seed=11)
    out = []
    for graphs in (False, True):
        d = cfa.DNS([cfa.FlowField(*G, 3, LX_C, LZ_C, data=u0), cfa.FlowField(*G, 1, LX_C, LZ_C)],
                    cfa.DNSFlags(**fl))
        d.set_graphs(graphs)
        d.advance_resident(15)
        d.reset_dt(0.01)
        d.advance_resident(12)
        out.append(d.state())
    (ue, qe), (ug, qg) = out
    print(nonlinearity, "graphs vs eager max|diff|", float(np.max(np.abs(ue - ug))), flush=True)
    assert np.array_equal(ue, ug) and np.array_equal(qe, qg)
