"""ART diffuse-photon transfer on one GPU (host side): the reference's call
sequence init_hd_param -> setup_diffuse_photon -> init_diffuse_gpu ->
send_diffuse_data -> calc_diffuse_photon_radiation
(src/argot_mpi.c:74-78, src/step_radiation_tree.c:321-330) through the
C-ABI of libargot_hip.so, on synthetic inputs."""
import ctypes as C

import numpy as np

from . import abi
from . import radiation as R
from ._lib import check, lib

UENETOK = 1.0                # run_param.uenetok of these runs: prev_uene * uenetok * wmol = T
XHYDROGEN, YHELIUM = 0.755, 0.245   # chemistry.h:23-24


def electron_fraction(w):
    return (w.fHII + w.helium_fact * (w.fHeII + 2.0 * w.fHeIII)).astype(np.float32)


def wmol(felec):
    """WMOL(c) with helium and no H2 (chemistry.h:69), rounded to FP32."""
    return (1.0 / ((1.0 + felec.astype(np.float64)) * XHYDROGEN + 0.25 * YHELIUM)).astype(
        np.float32)


def prev_uene(w):
    """Specific energy that gives T = w.temperature through the reference's
    temper = prev_uene * uenetok * wmol."""
    return (w.temperature / (UENETOK * wmol(electron_fraction(w)).astype(np.float64))).astype(
        np.float32)


def chem_array(w):
    """[ncell, 7]: felec, fHI, fHII, fHeI, fHeII, fHeIII, prev_uene."""
    cols = [electron_fraction(w), w.fHI, w.fHII, w.fHeI, w.fHeII, w.fHeIII, prev_uene(w)]
    return np.stack([np.asarray(c, np.float32).reshape(-1) for c in cols], axis=1)


def fill_art_mesh(w):
    mesh = R.fill_mesh(w)
    v = abi.mesh_view(mesh)
    pc = v["prev_chem"]
    pc["felec"] = electron_fraction(w).reshape(-1)
    v["prev_chem"] = pc
    v["prev_uene"] = prev_uene(w).reshape(-1)
    return mesh


class DiffuseStep:
    """One GPU, one rank: the diffuse transfer of a mesh."""

    def __init__(self, w, nside=8, heatcool=True, batch=0, kahan=False, exp64=False):
        L = lib()
        self.w = w
        self.nside, self.heatcool, self.batch, self.kahan = nside, heatcool, batch, kahan
        self.exp64 = exp64
        self._settings()
        self.freq, self.cs = R.freq_and_csect()
        self.run = R.run_param_for(w, self.freq, self.cs)
        self.run.uenetok = UENETOK
        self.mesh = fill_art_mesh(w)
        self.cmem = abi.CudaMemSpace()
        self.cpar = abi.CudaParam()
        L.argot_hip_set_helium_fact(w.helium_fact)
        L.init_gpu(self.mesh, C.byref(self.cmem), C.byref(self.cpar), C.byref(self.run))
        check("init_gpu")
        L.send_mesh_data(self.mesh, C.byref(self.cmem), C.byref(self.cpar), C.byref(self.run))
        check("send_mesh_data")
        self.hd = abi.HostDiffuseParam()
        L.init_hd_param(C.byref(self.hd))
        check("init_hd_param")
        L.setup_diffuse_photon(C.byref(self.hd), C.byref(self.run))
        check("setup_diffuse_photon")
        self.cd = (abi.CudaDiffuseParam * 4)()
        L.init_diffuse_gpu(self.cd, C.byref(self.cpar), C.byref(self.run))
        check("init_diffuse_gpu")
        L.send_diffuse_data(C.byref(self.hd), self.cd, C.byref(self.cpar))
        check("send_diffuse_data")

    def _settings(self):
        L = lib()
        L.argot_hip_set_diffuse_nside(self.nside)
        L.argot_hip_set_diffuse_heatcool(1 if self.heatcool else 0)
        L.argot_hip_set_diffuse_batch(self.batch)
        L.argot_hip_set_diffuse_kahan(1 if self.kahan else 0)
        L.argot_hip_set_diffuse_exp64(1 if self.exp64 else 0)

    def angles(self):
        n = 12 * self.nside ** 2
        a = self.hd.angle
        ovr = np.array([[a[i].xovr, a[i].yovr, a[i].zovr] for i in range(n)], np.float32)
        base = np.array([a[i].base_id for i in range(n)], np.int16)
        corner = np.array([a[i].corner_id for i in range(n)], np.int16)
        sf = self.hd.step_fact[0]
        fact = np.array([sf.HI[0], sf.HI[1], sf.HeI[0], sf.HeI[1], sf.HeII[0], sf.HeII[1]],
                        np.float32)
        return ovr, base, corner, np.array(self.hd.corner_id_num[:]), fact

    def step(self):
        """Rates zeroed, then one diffuse pass; returns its statistics."""
        L = lib()
        self._settings()
        L.argot_hip_set_helium_fact(self.w.helium_fact)
        L.zero_out_photoion_rate_cuda(C.byref(self.cmem), C.byref(self.cpar), C.byref(self.run))
        check("zero_out_photoion_rate_cuda")
        L.calc_diffuse_photon_radiation(self.mesh, C.byref(self.run), C.byref(self.cmem),
                                        C.byref(self.cpar), C.byref(self.hd), self.cd, None)
        check("calc_diffuse_photon_radiation")
        st = abi.DiffuseStats()
        L.argot_hip_diffuse_stats(C.byref(st))
        return {k: getattr(st, k) for k, _ in abi.DiffuseStats._fields_}

    def rates(self):
        L = lib()
        L.recv_mesh_data(self.mesh, C.byref(self.cmem), C.byref(self.cpar), C.byref(self.run))
        check("recv_mesh_data")
        pc = abi.mesh_view(self.mesh)["prev_chem"]
        return np.stack([pc[f] for f in R.RATE_FIELDS], axis=1).astype(np.float32)

    def close(self):
        L = lib()
        L.free_diffuse_gpu(self.cd, C.byref(self.cpar))
        L.free_hd_param(C.byref(self.hd))
        if self.cmem.mesh_dev:
            L.argot_hip_release_gpu(C.byref(self.cmem))
