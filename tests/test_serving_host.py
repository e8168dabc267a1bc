"""Host-side logic of batches and the sweep server (no GPU): the seeded
workload generators and the per-request hysteresis control."""
import numpy as np

from cupymag_amd.serving.server import SweepJob, estimate_bytes, multigrid_ok
from cupymag_amd.serving.workloads import batch_workload, initial_m, sweep_requests


def test_batch_workload_distribution():
    a = batch_workload(64, seed=4)
    b = batch_workload(64, seed=4)
    assert [x["n"] for x in a] == [x["n"] for x in b]
    ns = np.array([x["n"] for x in a])
    assert ns.min() >= 16 and ns.max() <= 96
    # log-uniform: the median is near the geometric mean of the range
    assert 25 <= np.median(ns) <= 60
    for k, x in enumerate(a):
        c = x["cfg"]
        assert c["physics"]["magnetoelastic_coupling"] == (k % 2 == 0)
        r = c["grid"]["tet_mesh"]["defect_parameters"]
        assert 2.0 <= r <= 0.3 * x["n"]
        f = c["physics"]["external_field"]
        assert np.linalg.norm([f["hext1"], f["hext2"], f["hext3"]]) <= 1e5


def test_sweep_requests_poisson_and_fields():
    reqs = sweep_requests(400, rate=20.0, seed=5, points=41)
    t = np.array([r["arrival"] for r in reqs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 20.0) < 0.01  # mean inter-arrival 1/rate
    for r in reqs[:20]:
        assert 20 <= r["n"] <= 128
        f = r["fields"]
        assert f.shape == (41, 3)
        np.testing.assert_allclose(np.linalg.norm(f[0]), 1e5)
        np.testing.assert_allclose(f[0], -f[-1])
        np.testing.assert_allclose(f[20], 0.0, atol=1e-9)
        R = np.array(r["cfg"]["grid"]["tet_mesh"]["rotation_matrix"])
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) > 0


def test_initial_m_direction():
    d = np.array([-1.0, 2.0, 0.5])
    m = initial_m(1000, 3, max_deg=5.0, direction=d)
    np.testing.assert_allclose(np.linalg.norm(m, axis=1), 1.0, atol=1e-12)
    cosang = m @ (d / np.linalg.norm(d))
    assert np.all(cosang >= np.cos(np.deg2rad(5.0)) - 1e-12)
    mx = initial_m(10, 3)
    assert np.all(mx[:, 0] > 0.99)


class _P:
    ms = 1.2e6
    ME = False
    LLG_accuracy_factor = 1e-3
    use_init_factor = 100.0


class _Rep:
    def __init__(self, sumsq):
        self.sumsq = sumsq
        self.avg_m = np.array([1.0, 0.0, 0.0])


class _CtlRep:
    def __init__(self, point_done=False, state=0, avg=(1.0, 0.0, 0.0)):
        self.point_done, self.state, self.avg_m = point_done, state, np.array(avg)


def test_sweep_job_bookkeeping_from_the_device_control():
    """The per-point rule runs on the device (sweep_ctl.h, pinned in
    test_host.py); the job records each completed point with its field,
    average magnetization, step count and latency, and ends on state 2."""
    req = dict(rid=0, fields=np.array([[-1e5, 0, 0], [0, 0, 0], [1e5, 0, 0]]), steps_per_point=150)
    job = SweepJob(req, None, _P, n_dof=1000, est=1, t_arrival=0.0, t_admit=0.0)
    job.t_point = 0.0
    for k in range(100):
        assert not job.advance(_CtlRep(), float(k))
    assert not job.advance(_CtlRep(True, 0, (0.9, 0, 0)), 100.0)
    assert job.point == 1 and job.points[0][2] == 101 and job.point_lat == [100.0]
    np.testing.assert_array_equal(job.points[0][0], [-1e5, 0, 0])
    for k in range(149):
        job.advance(_CtlRep(), 0.0)
    job.advance(_CtlRep(True), 0.0)
    assert job.point == 2 and job.points[1][2] == 150
    for k in range(100):
        job.advance(_CtlRep(), 0.0)
    assert job.advance(_CtlRep(True, 2), 0.0)
    assert len(job.points) == 3 and job.steps == 101 + 150 + 101


def test_estimate_grows_with_size_and_coupling():
    assert estimate_bytes(1000, True) > estimate_bytes(1000, False)
    assert estimate_bytes(128 ** 3, True) < 30e9  # a 128^3 member with coupling fits many times in 288 GB


def test_multigrid_rule_and_estimate():
    assert multigrid_ok(9) and multigrid_ok(22) and multigrid_ok(87) and multigrid_ok(128)
    assert not multigrid_ok(8) and not multigrid_ok(7)
    for me in (False, True):
        assert estimate_bytes(64 ** 3, me, 64) > estimate_bytes(64 ** 3, me)
    # 16^3 -> 8^3 with coupling: the dense FP32 inverses take 4 (512^2 + 1536^2) bytes
    assert estimate_bytes(16 ** 3, True, 16) - estimate_bytes(16 ** 3, True) >= 4 * (512 ** 2 + 1536 ** 2)


def test_multigrid_estimate_per_mode():
    """FP64 hierarchies (fast64: FP64 level vectors and dense inverses, the
    FP64 level-0 stiffness copy of the floating hierarchies) take more than
    the mixed mode's FP32/FP16 ones; without multigrid the mode changes
    nothing."""
    for me in (False, True):
        assert estimate_bytes(64 ** 3, me, 64, "fast64") > estimate_bytes(64 ** 3, me, 64, "mixed")
        assert estimate_bytes(64 ** 3, me, None, "fast64") == estimate_bytes(64 ** 3, me, None, "mixed")
    # 16^3 -> 8^3 with coupling in FP64: the dense inverses take 8 (512^2 + 1536^2) bytes
    assert estimate_bytes(16 ** 3, True, 16, "fast64") - estimate_bytes(16 ** 3, True) >= 8 * (512 ** 2 + 1536 ** 2)
