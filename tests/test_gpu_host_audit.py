"""The host KKT audits of points the device returns, against the reference's own audits text
for text (tests/test_host_audit.py explains the dump): the restarted engine's answers (FP64
and FP32 with FP64 refinement) on the structured LPs, in the reference's output convention
(x, row_dual), under the status the engine reports and with the objective it reports."""
import os

import numpy as np
import pytest

import lp_cases
import oracle as O
import test_host_audit as H
import vajra_pdhg as vp

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not O.have_ref(), reason="the reference library (oracle/_ref) is not built")
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_points_audited_as_reference(engine, tmp_path, prec):
    compared = 0
    for s, m, n in H.STRUCTURED:
        lp = lp_cases.structured_feasible_lp(s, m, n)
        mps = os.path.join(tmp_path, f"f{s}.mps")
        lp_cases.write_mps(lp, mps)
        lp.offset = 0.0   # the model read from the MPS file has no offset
        print(f"[case] device point audit {prec} structured {s} {m}x{n}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-8, max_iterations=200000))
        x, rd, y = engine.download()
        assert r.status == 1, r.as_dict()
        obj = float(np.dot(lp.c, x))
        for tol in (1e-6, 1e-9):
            pt = os.path.join(tmp_path, f"f{s}.pt")
            H.write_point(pt, r.status, obj, x, rd)
            mine, ref = H.dumps(mps, pt, tol, str(tmp_path))
            assert mine == ref, (s, tol)
            compared += 1
    assert compared == 2 * len(H.STRUCTURED)
