"""The parameter-scan workload at its real shape: 16 independent wakes on
each of 128^2, 192^2, 256^2 and 384^2 (4 ppc, dxi = 0.02; drivers with
sigma_r ~ U[0.2, 0.8], sigma_xi ~ U[0.5, 1.5], nb ~ U[0.5, 8] from the seed),
each size one batched context with graph replay, over the first 64 layers.

  * one wake per size (a different index per size) equals a single-wake
    context of the same driver bitwise (same tile shapes; residuals, Ex, psi);
  * the 128^2 wake also matches the FP64 oracle (rel-L2 <= 1e-9).
"""
import numpy as np
import pytest

from tertia_amd.params import Params, parameter_scan

pytestmark = pytest.mark.gpu

NS = 64


def sweep(p, graph, choice=None):
    from tertia_amd.mesh import Mesh
    m = Mesh(p, history=("Ex", "Psi"))
    try:
        if choice is None:
            m.autotune(1)
        else:
            m.set_gemm_choice(choice)
        m.enable_graph(graph)
        res = m.MoveAllSplitLayers()
        chosen = m.gemm_choice()
        out = []
        for w in range(p.nwake):
            m.select_wake(w)
            out.append({f: np.stack([m.history(f, s) for s in range(p.nslices)])
                        for f in ("Ex", "Psi")})
    finally:
        m.close()
    return res, out, chosen


@pytest.mark.timeout(600)
def test_parameter_scan_workload(oracle_mod):
    groups = parameter_scan(nslices=NS)
    for gi, pb in enumerate(groups):
        res_b, out_b, choice = sweep(pb, graph=True)
        assert res_b.shape == (16, NS) and np.all(np.isfinite(res_b))
        w = (5 + 3 * gi) % 16
        ps = Params(ny=pb.ny, nz=pb.nz, ly=pb.ly, lz=pb.lz, hx=pb.hx, nslices=NS,
                    bunches=list(pb.wakes[w]))
        res_s, out_s, _ = sweep(ps, graph=False, choice=choice)
        same = np.array_equal(res_s, res_b[w])
        print("scan %d^2 wake %d: batch == single %s" % (pb.ny, w, same), flush=True)
        assert same
        for f in ("Ex", "Psi"):
            assert np.array_equal(out_s[0][f], out_b[w][f]), (pb.ny, w, f)
        if pb.ny == 128:
            o = oracle_mod.Oracle(ps, use_fft=True, nthreads=16, history=("Ex", "Psi"))
            try:
                st, _ = o.sweep()
                assert st == 0
                for f in ("Ex", "Psi"):
                    R = np.stack([o.history(f, s) for s in range(NS)])
                    e = float(np.linalg.norm(out_b[w][f] - R) / np.linalg.norm(R))
                    print("scan 128^2 wake %d vs oracle %s %.1e" % (w, f, e), flush=True)
                    assert e <= 1e-9, (f, e)
            finally:
                o.close()
