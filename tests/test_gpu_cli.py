"""The dvc and strain programs end to end on the GPU against the oracle: the .disp rows
(processing order, labels, statuses, objmin, displacements; Core/dvc.cpp:312-320,
Core/InputRead.cpp:1131-1196), the .stat summary, and the strain CSV files
(Core/strain.cpp:595-656) of a run on that .disp.  Printed values carry 6 significant
digits (objmin, strains) or 6 decimals (displacements); tolerances are the parity
bar plus that rounding."""
import numpy as np
import pytest

import cli_util as U
from digitalvolumecorrelation_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bins():
    return U.ensure_built()


def test_dvc_and_strain_programs(tmp_path, bins, oracle):
    case = synth.translation_u8(size=72, n=4, spacing=8, npts=1000)
    p = dict(case.params, subvol_size=21, disp_max=4)
    start = (36.0, 36.0, 36.0)
    d = U.write_run(tmp_path, case.ref, case.tar, case.points, p, extra={"starting_point": "36.0 36.0 36.0"})
    r = U.run([bins / "dvc", "dvc_in"], d, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "pt/sec average" in r.stdout

    order, neigh = oracle.sort_order_neighbors(case.points, start=np.array(start))
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **p)
    ores, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nthreads=8)
    lab, xyz, st, obj, uvw, rows = U.read_disp(d / "result.disp")
    assert np.array_equal(lab, order + 1)
    assert np.array_equal(st, ores["status"][order])
    assert np.abs(xyz - case.points[order]).max() <= 5e-6 * np.abs(case.points).max()
    good = st == 0
    o = ores[order]
    assert np.abs(uvw[good] - o["par"][good, :3]).max() <= 1e-3 + 5e-7
    assert (np.abs(obj[good] - o["obj"][good]) / o["obj"][good]).max() <= 1e-5 + 5e-6
    assert np.all(uvw[~good] == 0.0) and np.all(obj[~good] == 0.0)
    sta = (d / "result.stat").read_text()
    assert f"number successful = {int(good.sum())}" in sta and "pt/sec" in sta
    assert "neighbour tables: full sort per point" in sta

    r = U.run([bins / "strain", "-sw", "27", "-A", "result.disp"], d, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    _, sneigh = oracle.sort_order_neighbors(xyz, start=xyz[0])
    assert np.array_equal(U.read_sort(d / "result.sort.csv"), sneigh)
    os_ = oracle.strain(xyz, uvw, st, obj, sneigh, ndp=27)
    for suffix, key in ((".Lstr.csv", "L"), (".Estr.csv", "E"), (".dgrd.csv", "DG")):
        lines = (d / f"result-sw27{suffix}").read_text().splitlines()
        assert lines[0].startswith("n,x,y,z,u_fit,v_fit,w_fit,pts_in_sw,sw_radius,")
        vals = np.array([[float(v) for v in ln.split(",")] for ln in lines[1:]])
        assert np.array_equal(vals[:, 0].astype(int), lab)
        assert np.array_equal(vals[:, 7].astype(int), os_["npts"])
        for got, want in ((vals[:, 9:18], os_[key]), (vals[:, 4:7], os_["disp"]), (vals[:, 8], os_["swr"])):
            scale = np.maximum(np.abs(want), 1e-12)
            assert np.all(np.abs(got - want) <= 1e-5 * np.abs(want).max() + 6e-6 * scale)
