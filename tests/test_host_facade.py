"""The C++ facade (the reference's class surface) driven by the reference's
main loop (apc_524_project_amd/host/run_main.cpp, src/main.cpp:114-287).

CPU tests cover the input checks and the boundary-factory errors, which fail
before any device call.  GPU tests run the driver in a child process, per
phase (the reference's call sequence) and fused (one device call per step),
and compare its history.dat and final particles with the oracle stepping the
same case: particles bit-exact, history to 1e-12 relative (the device sums
in a different order than the oracle).
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle_ctypes import EXTERNAL, PERIODIC, REFLECTING, Case, Oracle, Wave

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN_MAIN = os.path.join(ROOT, "apc_524_project_amd", "emo_run_main")

BC_NAMES_F = {PERIODIC: "periodic", EXTERNAL: "external"}
BC_NAMES_P = {PERIODIC: "periodic", REFLECTING: "reflecting"}


def need_binary():
    if not os.path.exists(RUN_MAIN):
        pytest.fail(f"{RUN_MAIN} is not built (python -m apc_524_project_amd.build)")


def settings_text(case: Case, nt: int, dt: float, precision: int = 0) -> str:
    f = lambda v: " ".join(repr(float(x)) for x in v)
    i = lambda v: " ".join(str(int(x)) for x in v)
    lines = [
        f"nCell = {i(case.n)}", f"xyz0 = {f(case.x0)}", f"Lxyz = {f(case.L)}",
        f"nt = {nt}", "t0 = 0.0", f"dt = {dt!r}",
        f"relativity = {case.relativity}", f"nspecies = {case.nspecies}",
        f"nparticles_tot = {case.nparticles}", f"dens_phys = {case.dens_phys!r}",
        f"mass_ratio = {f(case.mass_ratio)}", f"charge_ratio = {f(case.charge_ratio)}",
        f"dens_frac = {f(case.dens_frac)}", f"temp = {f(case.temp)}",
        f"isTestParticle = {i(case.is_test)}",
        f"E0 = {f(case.E0)}", f"B0 = {f(case.B0)}",
        "fields_init = " + ("poisson" if case.poisson else "constant"),
        "parts_bound = " + " ".join(BC_NAMES_P[b] for b in case.pbc),
        "fields_bound = " + " ".join(BC_NAMES_F[b] for b in case.fbc),
        f"seed = {case.seed}", f"precision = {precision}",
    ]
    if case.waves:
        w = case.waves
        lines += [f"nwaves = {len(w)}", "inSide = " + i(x.side for x in w), "inPolE = " + i(x.pol for x in w),
                  "peakamps = " + f(x.amp for x in w), "omegas = " + f(x.omega for x in w),
                  "phases = " + f(x.phase for x in w), "invWidths = " + f(x.inv_width for x in w),
                  "delays = " + f(x.delay for x in w)]
    return "\n".join(lines) + "\n"


def run_driver(tmp_path, text, *extra, check=True):
    need_binary()
    s = tmp_path / "settings.txt"
    s.write_text(text)
    out = tmp_path / "out"
    out.mkdir(exist_ok=True)
    r = subprocess.run([RUN_MAIN, str(s), str(out), *extra], capture_output=True, text=True, timeout=120)
    if check:
        assert r.returncode == 0, r.stderr
    return r, out


def read_particles(path):
    b = path.read_bytes()
    (n,) = struct.unpack_from("<q", b, 0)
    off = 8
    x = np.frombuffer(b, np.float64, 3 * n, off).reshape(n, 3)
    off += 24 * n
    v = np.frombuffer(b, np.float64, 3 * n, off).reshape(n, 3)
    off += 24 * n
    g = np.frombuffer(b, np.float64, n, off)
    off += 8 * n
    ids = np.frombuffer(b, np.int64, n, off)
    return x, v, g, ids


def oracle_history(case, nt, dt, nstep_sort=None):
    """The oracle stepping the case; with nstep_sort, the reference loop's
    SortParticles before the push every nstep_sort steps (src/main.cpp:225-226),
    which fixes the within-cell order the first deposit sums in."""
    o = Oracle(case)
    assert o.load() == 0
    o.initialize()
    rows = []
    t = 0.0
    for ti in range(nt):
        if nstep_sort and ti % nstep_sort == 0:
            o.sort()
        if ti % 10 == 0:
            d = o.diag()
            rows.append(("%e" % t, d[:7]))
        o.step(t, dt)
        t += dt
    return o, rows


def parse_history(path):
    lines = path.read_text().splitlines()
    assert lines[0] == ("[1] time [2] part mom x [3] part mom y [4] part mom z [5] part Energy "
                        "[6] B Energy [7] E Energy [8] Total E")
    rows = []
    for ln in lines[1:]:
        parts = ln.split()
        assert len(parts) == 8
        rows.append((parts[0], np.array([float(x) for x in parts[1:]])))
    return rows


# ---- CPU: input handling and factory errors (no device call) ----------------

def test_unknown_setting_is_rejected(tmp_path):
    r, _ = run_driver(tmp_path, "nCell = 4 4 4\nbogus = 1\n", check=False)
    assert r.returncode == 1
    assert "unknown setting bogus" in r.stderr


def test_unknown_particle_boundary_message(tmp_path):
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=64,
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC))
    text = settings_text(case, 1, 0.1).replace("parts_bound = periodic periodic",
                                               "parts_bound = absorbing absorbing")
    r, _ = run_driver(tmp_path, text, check=False)
    assert r.returncode == 1
    # src/boundaries/particle_bc_factory.hpp: the reference's message
    assert "Unable to find particle boundary absorbing" in r.stderr


def test_inconsistent_periodic_dimension_fails_input_check(tmp_path):
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=64,
                pbc=(REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC, PERIODIC))
    r, _ = run_driver(tmp_path, settings_text(case, 1, 0.1), check=False)
    assert r.returncode == 1
    assert "input check failed" in r.stderr


# ---- GPU: the driver loop against the oracle --------------------------------

CASES = {
    "periodic_vay": Case(n=(6, 5, 7), L=(6.0, 2.5, 3.5), nparticles=1680, temp=(5000.0, 5000.0),
                         dens_phys=3.0e9, relativity=1, B0=(0.0, 0.0, 0.01), E0=(0.001, 0.0, 0.0)),
    "lightbc_boris": Case(n=(10, 4, 5), L=(10.0, 4.0, 5.0), nparticles=800, temp=(501.8, 501.8),
                          dens_phys=2.0e9, relativity=0,
                          fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                          pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                          waves=[Wave(side=-1, pol=2, amp=0.2, omega=0.5, phase=10.0, inv_width=0.1,
                                      delay=3.0)]),
}


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [False, True], ids=["per_phase", "fused"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_driver_matches_oracle(tmp_path, name, fused):
    case = CASES[name]
    nt = 23
    dt = case.dt_cfl()
    print(f"[facade] {name} fused={fused}", flush=True)
    _, out = run_driver(tmp_path, settings_text(case, nt, dt), *(["--fused"] if fused else []))
    # the per-phase loop sorts before the push every nstep_sort = 100 steps
    o, rows = oracle_history(case, nt, dt, nstep_sort=None if fused else 100)
    x, v, g, ids = read_particles(out / "particles.bin")
    p = o.particles().by_id()
    assert np.array_equal(ids, p.id)
    assert np.array_equal(x, p.x)
    assert np.array_equal(v, p.v)
    assert np.array_equal(g, p.gamma)
    hist = parse_history(out / "history.dat")
    assert len(hist) == len(rows) == 3
    for (ts, a), (to, b) in zip(hist, rows):
        assert ts == to
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
        assert a[6] == pytest.approx(a[3] + a[4] + a[5], rel=1e-15)


@pytest.mark.gpu
def test_driver_fp32_runs_and_tracks_oracle(tmp_path):
    case = CASES["periodic_vay"]
    nt = 23
    dt = case.dt_cfl()
    _, out = run_driver(tmp_path, settings_text(case, nt, dt, precision=1))
    o, rows = oracle_history(case, nt, dt)
    x, v, g, ids = read_particles(out / "particles.bin")
    p = o.particles().by_id()
    assert np.array_equal(ids, p.id)
    assert np.linalg.norm(x - p.x) / np.linalg.norm(p.x) < 1e-6
    hist = parse_history(out / "history.dat")
    for (ts, a), (to, b) in zip(hist, rows):
        assert ts == to
        np.testing.assert_allclose(a[:4], b[:4], rtol=1e-4)
