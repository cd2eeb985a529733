"""The drop-in readmodel.py: parameters, shot lists, node split, file names and
model-file format as the reference's readmodel.py defines them (fixtures
below are written from reading readmodel.py:1-386 and pyAcoustic.py:55-114)."""
import numpy as np
import pytest

import readmodel as rm


def test_shot_list_order_and_node_split():
    sg = rm.ShotReceiver(sxbeg=1, sybeg=1, szbeg=1, jsx=10, jsy=10, jsz=1, sxend=21, syend=11, szend=1,
                         dgx=10.0, dgy=10.0, dgt=0.002, shotfolder="./shots/")
    expect = [(1, 1, 1, 10, 10, 1), (1, 11, 1, 10, 10, 1), (11, 1, 1, 10, 10, 1),
              (11, 11, 1, 10, 10, 1), (21, 1, 1, 10, 10, 1), (21, 11, 1, 10, 10, 1)]
    assert sg.shotlist == expect and sg.nshot == 6
    assert sg.nodetask == [expect]
    assert sg.allocateNode(nnode=4) == [expect[0::4], expect[1::4], expect[2::4], expect[3::4]]
    assert sg.nodetask[1] == [(1, 11, 1, 10, 10, 1), (21, 11, 1, 10, 10, 1)]


def test_example_driver_parameters():
    """pyAcoustic.py:55-111: the example's objects construct and print."""
    m = rm.Model(nx=201, ny=201, nz=201, dx=10, dy=10, dz=10, vel_file='./model/vel.bin',
                 rho_file='./model/den.bin', epsilon_file='./model/epsilon.bin', delta_file='./model/delta.bin',
                 theta_file='./model/theta.bin', psi_file='./model/psi.bin')
    sg = rm.ShotReceiver(sxbeg=101, sybeg=101, szbeg=101, jsx=10, jsy=10, jsz=1, sxend=101, syend=101, szend=101,
                         dgx=10.0, dgy=10.0, dgt=0.002, shotfolder='./shot_record/')
    assert sg.shotlist == [(101, 101, 101, 10, 10, 1)]
    fd = rm.Acoustic3dtti(model=m, sg=sg, nt=1001, dt=0.001, fpeak=15.0, wtype=1, snap_folder='./snapshots/',
                          snap_interval=200, savesnap=True, nodei=0, cut_directwave=False)
    text = str(fd)
    assert "Stencil Acoustic tti 3D" in text and "nt : 1001, dt : 0.001" in text and "snap interval : 200" in text
    assert fd.shot_name((51, 51, 51, 10, 10, 1)) == "./shot_record/51.51.bin"
    # 201 receivers x 201 x 501 samples (dgt/dt = 2)
    assert fd.expected_shot_samples() == 201 * 201 * 501


def test_receiver_spacing_assertions():
    m = rm.Model(10, 10, 10, 10, 10, 10, vel_file="v")
    bad = rm.ShotReceiver(1, 1, 1, 1, 1, 1, 1, 1, 1, 15.0, 10.0, 0.002, "./")
    with pytest.raises(AssertionError):
        rm.Acoustic3d2order(m, bad, 100, 0.001, 20.0, "./", 10)


def test_model_file_format_roundtrip(tmp_path):
    m = rm.Model(nx=5, ny=4, nz=3, dx=10, dy=10, dz=10, vel_file=str(tmp_path / "v.bin"),
                 rho_file=str(tmp_path / "r.bin"), epsilon_file=str(tmp_path / "e.bin"),
                 delta_file=str(tmp_path / "d.bin"))
    m.generate_model(vel=2500, rho=1100, epsilon=0.0, delta=0.05)
    raw = np.fromfile(tmp_path / "v.bin", dtype="<f4")
    assert raw.size == 60 and np.all(raw == 2500)
    assert np.all(np.fromfile(tmp_path / "e.bin", dtype="<f4") == 0.0)   # zero-valued file is written
    a = np.arange(60, dtype=np.float32).reshape(5, 4, 3)                   # (nx, ny, nz)
    m.write_array(str(tmp_path / "a.bin"), a)
    raw = np.fromfile(tmp_path / "a.bin", dtype=np.float32)
    # file order: iz + nz*ix + nz*nx*iy  (reference iso:249-255)
    ix, iy, iz = 3, 2, 1
    assert raw[iz + 3 * ix + 3 * 5 * iy] == a[ix, iy, iz]
    assert np.array_equal(m.read_model(str(tmp_path / "a.bin")), a)
