"""The resident TGV2 solver (one cooperative launch that keeps every 64x20
tile's state on chip and exchanges tile edges through tagged words) against
the launch sequence (two iterations per launch): the same states bit for
bit after psw_tgv_iterate, the same stop iteration, energy and state after
psw_tgv_solve, a state that resumes on the launch sequence, frames with
partial tiles, a single tile and 1..8 sources; and the failure path of its
bounded waits, forced from the host with a zero spin budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TGV_KW = dict(lambda_=1.0, alpha0=2.0, alpha1=1.5, tau=0.02, sigma=1.0, beta=0.0, gamma=1.0)
KEYS = ("u", "u1x", "u1y", "px", "py", "qx", "qy", "qz", "r")
SPECS = [
    {"seed": 21, "W": 130, "H": 53, "N": 3, "P": 16},     # 3 x 3 tiles, partial in x and y
    {"seed": 22, "W": 64, "H": 10, "N": 2, "P": 8},       # one tile: no neighbours
    {"seed": 23, "W": 200, "H": 120, "N": 8, "P": 24},    # 8 sources
    {"seed": 24, "W": 96, "H": 31, "N": 1, "P": 9},
    {"seed": 25, "W": 640, "H": 480, "N": 8, "P": 32},    # 240 tiles, the full-size frame
]


def _frame(spec, prm):
    from psw_amd import batch
    f = batch.Frame(spec, znear=0.5, zfar=6.0, winsize=7)
    f.sweep.build_homographies()
    f.sweep.run(fast=True)
    f.linearise(prm)
    return f


def _same(sa, sb, what):
    for k in KEYS:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"{what}: {k}")


@pytest.mark.parametrize("spec", SPECS, ids=lambda s: f"{s['W']}x{s['H']}x{s['N']}")
def test_resident_iterate_matches_launches(gpu, monkeypatch, spec):
    from psw_amd import stages
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frame(spec, prm), _frame(spec, prm)
    print(f"[case] resident iterate {spec}", flush=True)
    for n in (4, 7, 10, 23):
        monkeypatch.setenv("PSW_TGV_RESIDENT", "0")
        a.tgv.iterate(n, prm)
        monkeypatch.setenv("PSW_TGV_RESIDENT", "1")
        b.tgv.iterate(n, prm)
        _same(a.tgv.get_state(), b.tgv.get_state(), f"after {n} more iterations")
    # u_{k-1} of the resident end state: three more iterations on the launch
    # sequence (fewer than the resident minimum) from both
    a.tgv.iterate(3, prm)
    b.tgv.iterate(3, prm)
    _same(a.tgv.get_state(), b.tgv.get_state(), "resumed on the launch sequence")
    assert a.tgv.energy(prm) == b.tgv.energy(prm)


@pytest.mark.parametrize("spec", [SPECS[0], SPECS[2], SPECS[4]], ids=lambda s: f"{s['W']}x{s['H']}x{s['N']}")
@pytest.mark.parametrize("check_every", [10, 7])
def test_resident_solve_matches_launches(gpu, monkeypatch, spec, check_every):
    """E* from a long run; the resident solve stops at the same iteration with
    the same energy and state; a solve that reaches max_iters too."""
    from psw_amd import stages
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frame(spec, prm), _frame(spec, prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "0")
    a.tgv.iterate(3000, prm)
    e_star = a.tgv.energy(prm)
    print(f"[case] resident solve {spec} check_every {check_every}: E* {e_star:.9e}", flush=True)
    res = {}
    for f, on in ((a, "0"), (b, "1")):
        monkeypatch.setenv("PSW_TGV_RESIDENT", on)
        f.linearise(prm)
        it, e = f.tgv.solve(3000, check_every, 1e-3, e_star, prm)
        res[on] = (it, e, f.tgv.get_state())
        f.linearise(prm)
        it2, e2 = f.tgv.solve(95, check_every, 1e-12, e_star * 0.5, prm)   # never at the gap
        res[on + "max"] = (it2, e2, f.tgv.get_state())
    print("stops (launches, resident):", res["0"][:2], res["1"][:2], flush=True)
    assert res["0"][0] < 3000 and res["0"][0] % check_every == 0
    assert res["0"][:2] == res["1"][:2]
    _same(res["0"][2], res["1"][2], "stop state")
    assert res["0max"][0] == res["1max"][0] == 95
    assert res["0max"][1] == res["1max"][1]
    _same(res["0max"][2], res["1max"][2], "state at max_iters")


def test_resident_called_off_leaves_the_state(gpu, monkeypatch):
    """With no start budget the launch may be called off before any wait on
    a neighbour (when a tile checks before the last one has started); then
    nothing is written and the launch sequence runs instead.  Either way the
    results are the launch sequence's."""
    from psw_amd import stages
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frame(SPECS[4], prm), _frame(SPECS[4], prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "0")
    a.tgv.iterate(30, prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "1")
    monkeypatch.setenv("PSW_TGV_RESIDENT_START_SPIN", "0")
    print("[case] resident iterate with a zero start budget", flush=True)
    b.tgv.iterate(30, prm)
    _same(a.tgv.get_state(), b.tgv.get_state(), "zero start budget")


def test_resident_wait_failure_is_reported(gpu, monkeypatch):
    """With no spin budget every wait fails at once: the launch still ends,
    the call reports the failure, and the solver works again afterwards."""
    from psw_amd import stages
    from psw_amd._lib import PSW_ERR_DEVICE, PswError
    prm = stages.tgv_params(**TGV_KW)
    a, b = _frame(SPECS[0], prm), _frame(SPECS[0], prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "1")
    monkeypatch.setenv("PSW_TGV_RESIDENT_SPIN", "0")
    print("[case] resident solve with a zero spin budget", flush=True)
    with pytest.raises(PswError) as ei:
        b.tgv.solve(200, 10, 1e-3, 1.0, prm)
    assert ei.value.status == PSW_ERR_DEVICE and "resident" in str(ei.value)
    with pytest.raises(PswError):
        b.tgv.iterate(10, prm)
    monkeypatch.delenv("PSW_TGV_RESIDENT_SPIN")
    b.linearise(prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "0")
    a.tgv.iterate(40, prm)
    monkeypatch.setenv("PSW_TGV_RESIDENT", "1")
    b.tgv.iterate(40, prm)
    _same(a.tgv.get_state(), b.tgv.get_state(), "after the failed launches")
