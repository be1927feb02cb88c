"""On-disk formats (CPU, no GPU): the restart file of Write_Restart /
Read_Restart (src/out.f:493-692) and the input deck of Define_Parameters
(src/define_periodic_sf.f:20-80).  The expected bytes are built here with
struct/numpy from the reference's write sequence, independently of the C
code."""
import struct

import numpy as np
import pytest

import st_amd


def test_restart_bytes_and_round_trip(tmp_path):
    nx, ny, nz = 6, 4, 5
    rng = np.random.default_rng(0)
    Q = np.asfortranarray(rng.normal(size=(ny, nz, nx, 5)))
    p = tmp_path / "run.restart.1"
    st_amd.restart_write(p, Q, 1.25, 37)
    raw = p.read_bytes()
    # rank 0: gsizes(4) = (Ny, Nz, Nx, 5), time, step (src/out.f:538-545);
    # data at disp = 5*intSize + realSize = 28 (:560), Fortran order
    head = struct.pack("=4id i", ny, nz, nx, 5, 1.25, 37)
    assert len(head) == 28
    assert raw == head + Q.tobytes(order="F")
    Q2, t, st = st_amd.restart_read(p, nx, ny, nz)
    assert t == 1.25 and st == 37 and np.array_equal(Q2, Q)
    with pytest.raises(st_amd.StError):
        st_amd.restart_read(p, nx, nz, ny)                 # other extents
    with pytest.raises(st_amd.StError):
        st_amd.restart_read(tmp_path / "missing", nx, ny, nz)


DECK = """Input/Output path and filename:
vis                  vis_path
restart              restart_path
stats                stats_path
probe                probe_path
source               source_path
iso                  prefix

Grid parameters:
1                    Grid_X
1                    Grid_Y
0d0                  x_start
0.5d0                y_start
0d0                  z_start
6.283185307179586d0  period_x
6.283185307179586d0  period_y
3.1415926535897931d0 period_z

Flow parameters:
458.9d0              Re
0.7d0                Pr
1d0                  Sc
1.4d0                gamma
0.76d0               vis_exp
0.0d0                bulk_ratio
1d0                  T_ratio
1                    hyper_param
0                    part_param

Time advancement parameters:
1                    march_mode
0.125d-1             Tstep
0.5d0                cfl_cal
200                  Nsteps
4200d0               run_time
1                    comp_stats

Input/Output parameters:
1                    out_grid
1                    out_init
1                    out_incr_x
1                    out_incr_y
1                    out_incr_z
50                   out_restart
100                  out_vis
1000                 out_probe
1000                 out_stats
1000                 out_source
"""


def test_deck(tmp_path):
    p = tmp_path / "cit.input.1"
    p.write_text(DECK)
    d = st_amd.deck_read(p)
    assert d["restart_path"] == "restart" and d["prefix"] == "iso"
    assert d["y_start"] == 0.5 and d["period_z"] == 3.1415926535897931
    assert d["Re"] == 458.9 and d["Pr"] == 0.7 and d["gamma"] == 1.4 and d["vis_exp"] == 0.76
    assert d["hyper_param"] == 1 and d["march_mode"] == 1
    assert d["Tstep"] == 0.0125 and d["cfl_cal"] == 0.5 and d["Nsteps"] == 200
    assert d["out_restart"] == 50 and d["out_source"] == 1000
    p2 = tmp_path / "short"
    p2.write_text(DECK[: DECK.index("Time advancement")])
    with pytest.raises(st_amd.StError):
        st_amd.deck_read(p2)
