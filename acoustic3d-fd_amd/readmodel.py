"""Drop-in replacement of the reference's readmodel.py (sh39o/Acoustic3d-fd).

Same classes, constructor keywords, methods, file formats and ctypes call
sequence as readmodel.py:1-386; the four operator classes call the legacy
entry points of libacoustic3d.so (include/acoustic3d.h), which run on one
MI355X GPU. Differences, each a fix of a reference defect:

- Model.write_model writes the whole volume in one call (the reference packs
  one float per struct.pack, readmodel.py:26-28); generate_model writes every
  file whose path is set, including zero-valued epsilon/delta/theta/psi (the
  reference skips zeros, :52-59, leaving a VTI/TTI run without its files).
- Model.read_model works (the reference swaps struct.unpack's arguments, :70)
  and returns the (nx, ny, nz) array the reference intended.
- Acoustic3d2order passes snap_interval and cut_directwave in the order the C
  entry point declares them (the reference swaps them, :336-337; the
  2nd-order entry point ignores both, so results are unchanged).
- available_GPU / parallel_run: one GPU. parallel_run runs this node's shots
  on device `device_i`, `streams` at a time (each shot has its own HIP
  stream); nvidia-smi is not used.
- run(..., skip_existing) resumes a survey: a shot whose file already holds
  the expected number of samples is skipped (the reference stubs this with
  `if 0:`, :224).
"""
import os
import sys

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libacoustic3d.so")


class Model:
    def __init__(self, nx, ny, nz, dx, dy, dz,
                 vel_file=None, rho_file=None,
                 delta_file=None, epsilon_file=None,
                 theta_file=None, psi_file=None):
        self.nx = nx
        self.ny = ny
        self.nz = nz
        self.dx = dx
        self.dy = dy
        self.dz = dz
        self.rho_file = rho_file
        self.vel_file = vel_file
        self.delta_file = delta_file
        self.epsilon_file = epsilon_file
        self.theta_file = theta_file
        self.psi_file = psi_file
        self.shape = (self.nx, self.ny, self.nz)

    def write_model(self, filename, value):
        """Homogeneous model file: nx*ny*nz native float32 values (z fastest)."""
        if filename is None:
            print(filename, "cannot be None")
            raise NameError
        np.full(self.nx * self.ny * self.nz, value, dtype=np.float32).tofile(filename)

    def write_array(self, filename, array):
        """Writes an (nx, ny, nz) array in the model-file order (z fastest, then x, then y)."""
        a = np.asarray(array, dtype=np.float32)
        if a.shape != (self.nx, self.ny, self.nz):
            raise ValueError(f"array shape {a.shape} != {(self.nx, self.ny, self.nz)}")
        np.ascontiguousarray(a.transpose(1, 0, 2)).tofile(filename)

    def write_rho(self, value):
        self.write_model(self.rho_file, value)

    def write_vel(self, value):
        self.write_model(self.vel_file, value)

    def write_epsilon(self, value):
        self.write_model(self.epsilon_file, value)

    def write_delta(self, value):
        self.write_model(self.delta_file, value)

    def write_theta(self, value):
        self.write_model(self.theta_file, value)

    def write_psi(self, value):
        self.write_model(self.psi_file, value)

    def generate_model(self, vel=3000, rho=1200, epsilon=0.0, delta=0.0, theta=0.0, psi=0.0):
        self.write_vel(vel)
        if self.rho_file is not None:
            self.write_rho(rho)
        for path, value, writer in ((self.epsilon_file, epsilon, self.write_epsilon),
                                    (self.delta_file, delta, self.write_delta),
                                    (self.theta_file, theta, self.write_theta),
                                    (self.psi_file, psi, self.write_psi)):
            if path is not None:
                writer(value)

    def read_model(self, filename):
        """Returns res[ix, iy, iz] from a model file (z fastest, then x, then y)."""
        a = np.fromfile(filename, dtype=np.float32, count=self.nx * self.ny * self.nz)
        return np.ascontiguousarray(a.reshape(self.ny, self.nx, self.nz).transpose(1, 0, 2))

    def __str__(self):
        res = "*" * 30 + '\n'
        res += "the size of model is:\n"
        res += "nx, ny, nz = {}, {}, {}\n".format(self.nx, self.ny, self.nz)
        res += "dx, dy, dz = {}, {}, {}\n".format(self.dx, self.dy, self.dz)
        res += "*" * 30 + '\n'
        res += "velocity file : " + self.vel_file + '\n'
        if self.rho_file:
            res += "density file : " + self.rho_file + '\n'
        if self.delta_file:
            res += "delta file : " + self.delta_file + '\n'
        if self.epsilon_file:
            res += "epsilon file: " + self.epsilon_file + '\n'
        res += "*" * 30 + '\n'
        return res


class ShotReceiver:

    def __init__(self, sxbeg, sybeg, szbeg, jsx, jsy, jsz, sxend, syend, szend, dgx, dgy, dgt, shotfolder):
        self.sxbeg = sxbeg
        self.sybeg = sybeg
        self.szbeg = szbeg
        self.jsx = jsx
        self.jsy = jsy
        self.jsz = jsz
        self.sxend = sxend
        self.syend = syend
        self.szend = szend
        self.shotfile = shotfolder
        self.shotlist = self.generateShotList()
        self.nshot = len(self.shotlist)
        self.nodetask = self.allocateNode()
        self.receiver_dx = dgx
        self.receiver_dy = dgy
        self.receiver_dt = dgt

    def generateShotList(self):
        """(sx, sy, sz, jsx, jsy, jsz) for x outer, y middle, z inner (readmodel.py:112-119)."""
        shotlist = []
        for isx in range(self.sxbeg, self.sxend + 1, self.jsx):
            for isy in range(self.sybeg, self.syend + 1, self.jsy):
                for isz in range(self.szbeg, self.szend + 1, self.jsz):
                    shotlist.append((isx, isy, isz, self.jsx, self.jsy, self.jsz))
        self.shotlist = shotlist
        return shotlist

    def allocateNode(self, nnode=1):
        """Round-robin split of the shot list over nnode nodes (readmodel.py:121-126)."""
        nodetask = []
        for nodei in range(nnode):
            nodetask.append(self.shotlist[nodei::nnode])
        self.nodetask = nodetask
        return nodetask

    def __str__(self):
        res = '*' * 30 + '\n'
        res += "the shot begin: {}, {}, {}\n".format(self.sxbeg, self.sybeg, self.szbeg)
        res += "the shot inverval: {}, {}, {}\n".format(self.jsx, self.jsy, self.jsz)
        res += "the geophone interval: {}, {}\n".format(self.receiver_dx, self.receiver_dy)
        res += "the geophone sampling time: {}\n".format(self.receiver_dt)
        res += "shot shotfolder is " + self.shotfile + '\n'
        res += "*" * 30 + '\n'
        return res


class FD:
    def __init__(self, model, sg, nt, dt, fpeak, snap_folder, snap_interval, wtype=1, device_i=0, savesnap=False,
                 nodei=0,
                 cut_directwave=False):
        self.model = model
        self.nt = nt
        self.dt = dt
        self.sg = sg
        self.fpeak = fpeak
        self.snap_folder = snap_folder
        self.snap_interval = snap_interval
        self.kernel = None
        self.savesnap = savesnap
        self.nodei = nodei
        self.cut_directwave = cut_directwave
        self.wtype = wtype
        self.device_i = device_i
        self.lib_path = LIB_PATH
        assert (self.sg.receiver_dx % self.model.dx == 0)
        assert (self.sg.receiver_dy % self.model.dy == 0)
        assert (self.sg.receiver_dt % self.dt == 0)

    def __str__(self):
        res = "the Model Property: \n"
        res += str(self.model)
        res += "the Receiver and Shot setting:\n"
        res += str(self.sg)
        res += "the dominant frequency: {}\n".format(self.fpeak)
        res += "nt : {}, dt : {}\n".format(self.nt, self.dt)
        res += "save snapshot : {}\n".format(self.savesnap)
        if self.savesnap:
            res += "snap interval : {}\n".format(self.snap_interval)
            res += "snap folder : {}\n".format(self.snap_folder)
        res += "cut direct wave : {}\n".format(self.cut_directwave)
        return res

    def available_GPU(self):
        """This build runs every shot on one GPU: the configured device."""
        return self.device_i

    def load_kernel(self):
        from ctypes import CDLL, c_int
        try:
            self.kernel = CDLL(self.lib_path)
            self.kernel.cuda_3dfd_tti.restype = c_int
        except OSError:
            print("cannot open " + self.lib_path)
            raise

    def shot_name(self, ishot):
        return self.sg.shotfile + '{0}.{1}.bin'.format(str(ishot[0]), str(ishot[1]))

    def expected_shot_samples(self):
        jx = int(np.float32(self.sg.receiver_dx) / np.float32(self.model.dx))
        jy = int(np.float32(self.sg.receiver_dy) / np.float32(self.model.dy))
        jt = int(np.float32(self.sg.receiver_dt) / np.float32(self.dt))
        n = lambda a, j: (a + j - 1) // j  # noqa: E731
        return n(self.model.nx, jx) * n(self.model.ny, jy) * n(self.nt, jt)

    def run(self, ishot, device, skip_existing=False):
        if self.kernel is None:
            self.load_kernel()
        if device is None:
            device = self.available_GPU()
        print("device: {}, shot: ".format(device), ishot)
        shotname = self.shot_name(ishot)
        if skip_existing and os.path.exists(shotname) and \
                os.path.getsize(shotname) == 4 * self.expected_shot_samples():
            print(shotname + " exists, skip")
            return 0
        return self._call(shotname, ishot, device)

    def parallel_run(self, streams=1):
        """Runs this node's task list on one GPU, `streams` shots at a time."""
        taskList = list(self.sg.nodetask[self.nodei])
        print("there are {} task on this node".format(len(taskList)))
        if self.kernel is None:
            self.load_kernel()
        if streams <= 1:
            for task in taskList:
                self.run(task, self.device_i)
            return
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=streams) as pool:
            list(pool.map(lambda t: self.run(t, self.device_i), taskList))

    def _common(self, ishot):
        from ctypes import c_float, c_int
        return ([c_int(0), c_int(1),
                 c_int(self.model.nx), c_int(self.model.ny), c_int(self.model.nz),
                 c_float(self.model.dx), c_float(self.model.dy), c_float(self.model.dz),
                 c_int(ishot[0]), c_int(ishot[1]), c_int(ishot[2]),
                 c_int(ishot[3]), c_int(ishot[4]), c_int(ishot[5]),
                 c_float(self.sg.receiver_dx), c_float(self.sg.receiver_dy), c_float(self.sg.receiver_dt)])


def _s(x):
    from ctypes import c_char_p
    return c_char_p(bytes(x, 'utf-8'))


class Acoustic3d1order(FD):

    def _call(self, shotname, ishot, device):
        from ctypes import c_bool, c_float, c_int
        return self.kernel.cuda_3dfd_1order(
            _s(self.model.vel_file), _s(self.model.rho_file), _s(self.snap_folder), _s(shotname),
            *self._common(ishot),
            c_int(self.nt), c_float(self.dt), c_int(self.wtype),
            c_float(self.fpeak), c_bool(self.savesnap), c_bool(self.cut_directwave),
            c_int(self.snap_interval),
            c_int(device))

    def __str__(self):
        res = "Stencil Acoustic 3D first order\n"
        res += "*" * 30 + '\n'
        res += super().__str__()
        return res


class Acoustic3dvti(FD):

    def __str__(self):
        res = "Stencil Acoustic vti 3D \n"
        res += "*" * 30 + '\n'
        res += super().__str__()
        return res

    def _call(self, shotname, ishot, device):
        from ctypes import c_bool, c_float, c_int
        return self.kernel.cuda_3dfd_vti(
            _s(self.model.vel_file), _s(self.model.rho_file), _s(self.model.epsilon_file),
            _s(self.model.delta_file), _s(self.snap_folder), _s(shotname),
            *self._common(ishot),
            c_int(self.nt), c_float(self.dt),
            c_float(self.fpeak), c_bool(self.savesnap), c_bool(self.cut_directwave),
            c_int(self.snap_interval),
            c_int(device))


class Acoustic3d2order(FD):

    def __str__(self):
        res = "Stencil Acoustic 3D second order\n"
        res += "*" * 30 + '\n'
        res += str(super().__str__())
        return res

    def _call(self, shotname, ishot, device):
        from ctypes import c_bool, c_float, c_int
        return self.kernel.cuda_3dfd(
            _s(self.model.vel_file), _s(self.snap_folder), _s(shotname),
            *self._common(ishot),
            c_int(self.nt), c_int(300), c_float(self.dt),
            c_float(self.fpeak), c_bool(self.savesnap),
            c_bool(self.cut_directwave), c_int(self.snap_interval),
            c_int(device))


class Acoustic3dtti(FD):

    def __str__(self):
        res = "Stencil Acoustic tti 3D \n"
        res += "*" * 30 + '\n'
        res += super().__str__()
        return res

    def _call(self, shotname, ishot, device):
        from ctypes import c_bool, c_float, c_int
        rho = self.model.rho_file if self.model.rho_file is not None else ""
        return self.kernel.cuda_3dfd_tti(
            _s(self.model.vel_file), _s(rho), _s(self.model.epsilon_file),
            _s(self.model.delta_file), _s(self.model.theta_file), _s(self.model.psi_file),
            _s(self.snap_folder), _s(shotname),
            *self._common(ishot),
            c_int(self.nt), c_float(self.dt),
            c_float(self.fpeak), c_bool(self.savesnap), c_bool(self.cut_directwave),
            c_int(self.snap_interval),
            c_int(device))


if __name__ == "__main__":
    print(__doc__, file=sys.stderr)
