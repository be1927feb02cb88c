"""The Fortran driver (shock-turbulence_amd/fortran): the reference's
subroutine names over the C-ABI, run as the reference's march-mode-1 main
program, against the oracle's march on the same initial condition."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from fields import relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "shock-turbulence_amd", "fortran", "st_run")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built (run build())")


@needs_exe
def test_driver_reports_missing_input(tmp_path):
    r = subprocess.run([EXE, str(tmp_path / "absent.nml")], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "cannot open" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_driver_march_matches_oracle(tmp_path):
    st_amd = pytest.importorskip("st_amd")
    n, nsteps, seed = 32, 5, 4242
    k0, urms, mt = 4.0, 1.0, 0.3
    nml = tmp_path / "run.nml"
    nml.write_text("&run nx=%d, ny=%d, nz=%d, nsteps=%d, seed=%d, k0=%g, urms=%g, mt=%g, "
                   "re_lambda=40, pr=0.7, hyper_param=1, qout='q.bin' /\n" % (n, n, n, nsteps, seed, k0, urms, mt))
    print("fortran driver", n, nsteps, flush=True)
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("step")]
    assert len(lines) == nsteps + 1
    Qg = np.fromfile(tmp_path / "q.bin", dtype=np.float64).reshape((n, n, n, 5), order="F")
    # the reference's log: a header and one line per step, 12 value columns
    # after n in formats 10 and 11 (src/main.f:109-116, :316-317)
    log = (tmp_path / "output").read_text().splitlines()
    assert len(log) == nsteps + 2
    assert log[0].split() == ["n", "t", "dt", "CFL_resv", "KE", "MaxDil", "Enstrophy", "MomX", "MomY", "MomZ",
                              "Mass", "Energy", "Entropy"]
    assert all(len(l.split()) == 13 for l in log[1:])
    assert len(log[1]) == 1 + 5 + 12 * 13

    re = st_amd.re_for_taylor(40.0, k0, urms, mt)
    o = O.Oracle(n, n, n, Re=re, Pr=0.7, hyper_param=1)
    Q0 = st_amd.ic_isotropic(n, n, n, k0=k0, urms=urms, mt=mt, seed=seed)
    Qc, t_c, st_c, dts = o.march(Q0, nsteps)
    assert relerr(Qg, Qc).max() < 1e-10
    last = [float(x) for x in lines[-1][2:]]
    assert int(lines[-1][1]) == st_c == nsteps
    # dt follows the state (1e-10 above); the time is the sum of the dts
    assert abs(last[0] - t_c) <= 1e-10 * t_c
    assert abs(last[1] - dts[-1]) <= 1e-10 * dts[-1]
    maxdil, ens = o.compute_grad(Qc)
    d = o.diagnostics(Qc)
    want = [d[3], maxdil, ens, d[0], d[1], d[2], d[4], d[5], d[6]]
    np.testing.assert_allclose(last[3:], want, rtol=1e-10, atol=1e-13)


@needs_exe
@pytest.mark.gpu
def test_driver_march_mode_2(tmp_path):
    """Constant time step (src/main.f:158-184): the Tstep of the start-up
    Compute_Tstep is kept, and the log's CFL column is cfl_resv Tstep^2."""
    st_amd = pytest.importorskip("st_amd")
    n, nsteps, seed = 32, 4, 77
    nml = tmp_path / "run.nml"
    nml.write_text("&run nx=%d, ny=%d, nz=%d, nsteps=%d, seed=%d, k0=4, urms=1, mt=0.3, re_lambda=40, pr=0.7, "
                   "hyper_param=1, march_mode=2, qout='q.bin' /\n" % (n, n, n, nsteps, seed))
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [[float(x) for x in l.split()[1:]] for l in r.stdout.splitlines() if l.startswith("step")]
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    o = O.Oracle(n, n, n, Re=re, Pr=0.7, hyper_param=1)
    Q = st_amd.ic_isotropic(n, n, n, k0=4.0, urms=1.0, mt=0.3, seed=seed)
    _, rv = o.cfl(Q)
    dt = o.tstep(Q, rv, 0)
    t = 0.0
    for k in range(nsteps):
        _, rv = o.cfl(Q)
        Q, t, _ = o.lddrk4(Q, dt, time=t, step=k)
        assert abs(lines[k + 1][2] - dt) == 0.0
        assert abs(lines[k + 1][3] - rv * dt * dt) <= 1e-11 * rv * dt * dt
    Qg = np.fromfile(tmp_path / "q.bin", dtype=np.float64).reshape((n, n, n, 5), order="F")
    assert relerr(Qg, Q).max() < 1e-10


@needs_exe
def test_driver_rejects_wall_clock_modes(tmp_path):
    nml = tmp_path / "run.nml"
    nml.write_text("&run nx=32, ny=32, nz=32, nsteps=2, march_mode=3 /\n")
    r = subprocess.run([EXE, str(nml)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "march_mode 3" in r.stdout


@needs_exe
@pytest.mark.gpu
def test_driver_deck_restart_resume(tmp_path):
    """A reference-format deck and restart files: a run interrupted after its
    restart at step 2 and resumed from that file (as src/main.f:69-71 reads
    <prefix>.restart.0) ends with the same bits as the uninterrupted run."""
    st_amd = pytest.importorskip("st_amd")
    from test_io_formats import DECK
    n = 32
    deck = DECK.replace("458.9d0              Re", "200d0                Re") \
               .replace("200                  Nsteps", "4                    Nsteps") \
               .replace("3.1415926535897931d0 period_z", "6.283185307179586d0  period_z")
    for d in ("full", "part"):
        (tmp_path / d / "restart").mkdir(parents=True)
        (tmp_path / d / "cit.input.1").write_text(deck)
    # both start from the same restart.0 (a seeded IC written in the reference format)
    Q0 = st_amd.ic_isotropic(n, n, n, k0=4.0, urms=1.0, mt=0.3, seed=99)
    for d in ("full", "part"):
        st_amd.restart_write(tmp_path / d / "restart" / "iso.restart.0", Q0, 0.0, 0)
    (tmp_path / "full" / "run.nml").write_text("&run nx=%d, ny=%d, nz=%d, deck='cit.input.1', restart_every=2, "
                                               "qout='q.bin' /\n" % (n, n, n))
    r = subprocess.run([EXE, "run.nml"], cwd=tmp_path / "full", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    Q2, t2, s2 = st_amd.restart_read(tmp_path / "full" / "restart" / "iso.restart.1", n, n, n)
    assert s2 == 2 and t2 > 0
    # the resumed run: 2 steps from the step-2 restart
    st_amd.restart_write(tmp_path / "part" / "restart" / "iso.restart.0", Q2, t2, s2)
    (tmp_path / "part" / "run.nml").write_text("&run nx=%d, ny=%d, nz=%d, deck='cit.input.1', nsteps=2, "
                                               "qout='q.bin' /\n" % (n, n, n))
    # nsteps in the namelist is overridden by the deck's Nsteps = 4; use a deck with 2
    (tmp_path / "part" / "cit.input.1").write_text(deck.replace("4                    Nsteps",
                                                                "2                    Nsteps"))
    r = subprocess.run([EXE, "run.nml"], cwd=tmp_path / "part", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    a = np.fromfile(tmp_path / "full" / "q.bin", dtype=np.float64)
    b = np.fromfile(tmp_path / "part" / "q.bin", dtype=np.float64)
    assert np.array_equal(a, b)
    Q4, t4, s4 = st_amd.restart_read(tmp_path / "full" / "restart" / "iso.restart.2", n, n, n)
    assert s4 == 4 and np.array_equal(Q4.ravel(order="F"), a)
