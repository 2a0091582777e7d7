"""The C3 ensemble to its 1000-step horizon.

64 plane Couette members at Re 400 (24 x 32x33x32, 24 x 48x49x48,
16 x 64x65x64; skew-symmetric, bulk velocity, SBDF3 with dt = 0.02 after the
SMRK2 start; perturbation norms log-uniform in [1e-3, 0.3], as the bench
builds them) run together in one Batch for 1000 steps, the horizon the
workload is quoted for, 25 steps per advance. Batch.advance checks every
member for non-finite values after its steps (CFG_ERR_NONFINITE, and
Batch.finite names the members), as the reference's batch client does
(channelflow/cfdsi.cpp:769-774); such members are reported and removed, and
the others go on.

At this fixed dt one member does not survive: member 49 (64x65x64, norm
0.27) has a CFL number that climbs past 0.7 and SBDF3 diverges between steps
150 and 175. The oracle does the same (tests/golden/c3_member49_cfl.json:
its CFL every 25 steps and the step after which it is no longer finite), so
the divergence belongs to the scheme at dt = 0.02. The GPU member must follow
the oracle's CFL history and diverge in the same 25 steps; every other member
must stay finite for 1000 steps, with a perturbation of the order of the wall
velocity and a clock at 1000 dt.
"""
import json
import os
import sys

import numpy as np
import pytest

import channelflow_amd as cfa
from channelflow_amd.fields import l2norm2

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import bench  # noqa: E402

pytestmark = pytest.mark.gpu


def test_c3_ensemble_1000_steps(gpu):
    pin = json.load(open(os.path.join(HERE, "golden", "c3_member49_cfl.json")))
    members = bench.c3_members()
    b = cfa.Batch()
    ids, n0 = {}, {}
    for i, m in enumerate(members):
        u = bench.member_field(m)
        n0[i] = float(np.sqrt(l2norm2(u)))
        ids[i] = b.add(bench.flow_fields(cfa, m["grid"], u), cfa.DNSFlags(**m["flags"]))
    chunk = pin["chunk"]
    cfl49, dropped = [], {}
    for step in range(chunk, 1001, chunk):
        try:
            b.advance(chunk)
        except cfa.CfgError as e:
            assert e.code == cfa.ERR_NONFINITE
            for i, k in list(ids.items()):
                if not b.finite(k):
                    dropped[i] = step
                    b.remove(k)
                    del ids[i]
            print(f"step {step}: not finite, removed members {sorted(i for i, s in dropped.items() if s == step)}",
                  flush=True)
        if pin["member"] in ids:
            cfl49.append(b.CFL(ids[pin["member"]]))
    assert dropped == {pin["member"]: pin["nonfinite_after_step"]}
    ref = np.array(pin["cfl"])
    got = np.array(cfl49[:len(ref)])
    print("member 49 CFL, GPU vs oracle:", np.round(got, 4), np.round(ref, 4), flush=True)
    assert len(cfl49) == len(ref) and np.allclose(got, ref, rtol=1e-6, atol=0)
    worst = 0.0
    for i, k in ids.items():
        u, q = b.state(k)
        assert np.all(np.isfinite(u)) and np.all(np.isfinite(q))
        n = float(np.sqrt(l2norm2(u)))
        worst = max(worst, n)
        assert abs(b.time(k) - 1000 * members[i]["flags"]["dt"]) < 1e-9
        assert n < 1.0, (members[i]["grid"], i, n0[i], n)
    print(f"1000 steps: {len(ids)} of {len(members)} members finite, largest perturbation norm {worst:.3f}",
          flush=True)
