"""ctypes binding of include/plum_host.h (libplum_host.so): run.in reader and
the host Monte Carlo driver over the GPU engine."""
import ctypes as C
import os

import numpy as np

from . import engine as E

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplum_host.so")
HOST_SYMBOLS = ("plum_runin_parse", "plum_sim_create", "plum_sim_run", "plum_sim_energy",
                "plum_sim_positions", "plum_sim_observables", "plum_sim_destroy",
                "plum_host_last_error", "plum_synth_count", "plum_synth_system",
                "plum_sim_gc_stats", "plum_sim_widom", "plum_sim_run_device",
                "plum_sim_density_profile", "plum_sim_write_last", "plum_sim_pressure_sample",
                "plum_sim_run_output", "plum_sim_gc_prescreen", "plum_sim_gc_prescreen_stats",
                "plum_slit_forces")


class RunInSummary(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int), ("n_mols", C.c_int), ("n_types", C.c_int),
        ("npbc", C.c_int), ("use_pair", C.c_int), ("use_ewald", C.c_int), ("use_bond", C.c_int),
        ("use_ext", C.c_int), ("use_gc", C.c_int), ("dipole_correction", C.c_int),
        ("box", C.c_double * 3), ("beta", C.c_double), ("move_size", C.c_double),
        ("move_prob", C.c_double * 5), ("steps", C.c_longlong),
        ("lj_cutoff", C.c_double), ("lB", C.c_double), ("alpha", C.c_double),
        ("k_bond", C.c_double), ("r0", C.c_double), ("rigid_bond", C.c_double),
        ("wall_cut", C.c_double), ("phantom", C.c_int), ("grafted", C.c_int),
    ]


class SynthSpec(C.Structure):
    _fields_ = [
        ("n_chains", C.c_int), ("chain_len", C.c_int), ("n_counterions", C.c_int),
        ("n_salt_pairs", C.c_int), ("bead_q", C.c_double), ("counterion_q", C.c_double),
        ("box", C.c_double * 3), ("npbc", C.c_int), ("min_sep", C.c_double), ("r0", C.c_double),
        ("graft_z", C.c_double), ("z_lo", C.c_double), ("z_hi", C.c_double),
        ("seed", C.c_ulonglong),
    ]


class HostError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HostError("libplum_host.so is not built; run __graft_entry__.build()")
        E.lib()  # the engine library first (libplum_host links against it)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.plum_runin_parse.argtypes = [C.c_char_p, C.POINTER(RunInSummary)]
        L.plum_sim_create.argtypes = [C.c_char_p, C.c_uint, C.POINTER(vp)]
        L.plum_sim_run.argtypes = [vp, C.c_longlong, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.plum_sim_energy.argtypes = [vp, C.POINTER(E.Terms), C.POINTER(E.Terms)]
        L.plum_sim_positions.argtypes = [vp, C.POINTER(C.c_double), C.c_int]
        L.plum_sim_observables.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.plum_sim_destroy.argtypes = [vp]
        L.plum_sim_gc_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.plum_sim_widom.argtypes = [vp, C.POINTER(C.c_double)]
        L.plum_sim_gc_prescreen.argtypes = [vp, C.c_int, C.c_double]
        L.plum_sim_gc_prescreen_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_longlong)]
        L.plum_sim_density_profile.argtypes = [vp, C.c_double, C.POINTER(C.c_double), C.c_int,
                                               C.POINTER(C.c_int)]
        L.plum_sim_write_last.argtypes = [vp, C.c_char_p]
        L.plum_sim_run_output.argtypes = [vp, C.c_longlong, C.c_char_p, C.c_int]
        L.plum_sim_run_device.argtypes = [vp, C.c_longlong, C.POINTER(C.c_longlong),
                                          C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
        L.plum_sim_pressure_sample.argtypes = [vp, C.POINTER(E.VpSample), C.POINTER(C.c_double),
                                               C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        L.plum_host_last_error.restype = C.c_char_p
        L.plum_synth_count.argtypes = [C.POINTER(SynthSpec), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.plum_synth_system.argtypes = [C.POINTER(SynthSpec), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise HostError("plum host error %d: %s" % (rc, lib().plum_host_last_error().decode()))


def slit_forces(sysdef, info, phantom):
    """Force sums of one call of the wall-force slit pressure
    (ForceField::CalcPressureForceLJELSlit, pressure.cc:404-470) for the
    configuration of `sysdef`, phantom wall sites first: LJ on ions / chains /
    the other wall, then the Ewald z forces likewise.  `info` carries the
    Ewald box, cutoffs and images (plum_info fields: box_e, vol_e,
    real_cutoff, real_cell, repl_cutoff, repl_cell)."""
    import numpy as np
    L = lib()
    if not getattr(L, "_slit_bound", False):
        L.plum_slit_forces.argtypes = [C.POINTER(E.Params), C.POINTER(E.Info), C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L._slit_bound = True
    p = E.make_params(sysdef)
    e = E.Info()
    for k in ("box_e", "real_cell", "repl_cell"):
        for d in range(3):
            getattr(e, k)[d] = getattr(info, k)[d]
    for k in ("vol_e", "real_cutoff", "repl_cutoff"):
        setattr(e, k, getattr(info, k))
    xyz = np.ascontiguousarray(sysdef["xyz"], dtype=np.float64).reshape(-1)
    q = np.ascontiguousarray(sysdef["q"], dtype=np.float64)
    ty = np.ascontiguousarray(sysdef["type"], dtype=np.int32)
    ms = np.ascontiguousarray(sysdef["mol_start"], dtype=np.int32)
    f = np.zeros(6)
    _check(L.plum_slit_forces(C.byref(p), C.byref(e), len(q), xyz.ctypes.data, q.ctypes.data, ty.ctypes.data,
                              len(ms) - 1, ms.ctypes.data, int(phantom), f.ctypes.data))
    return f


def parse_runin(path):
    s = RunInSummary()
    _check(lib().plum_runin_parse(path.encode(), C.byref(s)))
    return s


class HostSimulation:
    def __init__(self, run_in_path, seed):
        h = C.c_void_p()
        _check(lib().plum_sim_create(run_in_path.encode(), seed, C.byref(h)))
        self.h = h

    def run(self, steps):
        att = (C.c_longlong * 5)()
        acc = (C.c_longlong * 5)()
        _check(lib().plum_sim_run(self.h, steps, att, acc))
        return list(att), list(acc)

    def energies(self):
        r, f = E.Terms(), E.Terms()
        _check(lib().plum_sim_energy(self.h, C.byref(r), C.byref(f)))
        return r.as_dict(), f.as_dict()

    def positions(self, n):
        out = np.zeros(3 * n)
        _check(lib().plum_sim_positions(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out.reshape(-1, 3)

    def run_device(self, steps):
        """Steps with the translational moves on the device loop; returns
        (attempted, accepted, [seconds in the device loop, in exchanges])."""
        att = (C.c_longlong * 5)()
        acc = (C.c_longlong * 5)()
        sec = (C.c_double * 2)()
        _check(lib().plum_sim_run_device(self.h, steps, att, acc, sec))
        return list(att), list(acc), list(sec)

    def density_profile(self, res=1.0):
        """rho(z) of chain beads, cations, anions: array (bins, 3)."""
        nb = C.c_int()
        _check(lib().plum_sim_density_profile(self.h, res, None, 0, C.byref(nb)))
        out = np.zeros(3 * nb.value)
        _check(lib().plum_sim_density_profile(self.h, res,
                                              out.ctypes.data_as(C.POINTER(C.c_double)),
                                              nb.value, C.byref(nb)))
        return out.reshape(-1, 3)

    def write_last(self, prefix):
        _check(lib().plum_sim_write_last(self.h, prefix.encode()))

    def run_output(self, steps, prefix, device=False):
        """Simulation::Run with the reference's sampling and output files
        (<prefix>_stat.dat, _traj_{p,c,a,n}.xyz, _lastcrd.dat(2), _lasttop.dat,
        _lastrhoZ.dat(2))."""
        _check(lib().plum_sim_run_output(self.h, int(steps), prefix.encode(), 1 if device else 0))

    def pressure_sample(self):
        """One volume-perturbation pressure sample: (sample dict, Yethiraj
        and de Miguel averages so far, number of samples)."""
        v = E.VpSample()
        py, pd, n = C.c_double(), C.c_double(), C.c_longlong()
        _check(lib().plum_sim_pressure_sample(self.h, C.byref(v), C.byref(py), C.byref(pd),
                                              C.byref(n)))
        return v.as_dict(), py.value, pd.value, n.value

    def gc_stats(self):
        """(attempted [ins, del], accepted [ins, del], n_beads, n_mols)"""
        att = (C.c_longlong * 2)()
        acc = (C.c_longlong * 2)()
        nb, nm = C.c_int(), C.c_int()
        _check(lib().plum_sim_gc_stats(self.h, att, acc, C.byref(nb), C.byref(nm)))
        return list(att), list(acc), nb.value, nm.value

    def gc_prescreen(self, mode, log_theta=0.0):
        """Two-stage exchange (plum_sim_gc_prescreen): 0 the reference's, 1 the
        reference's recording ln(W / w_1), 2 two-stage with theta = e^log_theta."""
        _check(lib().plum_sim_gc_prescreen(self.h, int(mode), float(log_theta)))

    def gc_prescreen_stats(self):
        """(first-stage rejections [ins, del], passes [ins, del], mean ln(W / w_1),
        ln mean W / w_1, growths recorded)."""
        rej = (C.c_longlong * 2)()
        pas = (C.c_longlong * 2)()
        mean = C.c_double()
        lmean = C.c_double()
        n = C.c_longlong()
        _check(lib().plum_sim_gc_prescreen_stats(self.h, rej, pas, C.byref(mean), C.byref(lmean),
                                                 C.byref(n)))
        return list(rej), list(pas), mean.value, lmean.value, n.value

    def widom(self):
        w = C.c_double()
        _check(lib().plum_sim_widom(self.h, C.byref(w)))
        return w.value

    def observables(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().plum_sim_observables(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "h", None):
            lib().plum_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synth_system(n_chains, chain_len, n_counterions, n_salt_pairs, box, seed, bead_q=-1.0,
                 counterion_q=1.0, npbc=3, min_sep=0.8, r0=1.0, graft_z=-1.0, z_range=None):
    """Seeded chains + ions built by the native generator.  Returns
    (xyz (N,3), q, type, mol_start)."""
    L = lib()
    sp = SynthSpec(n_chains=n_chains, chain_len=chain_len, n_counterions=n_counterions,
                   n_salt_pairs=n_salt_pairs, bead_q=bead_q, counterion_q=counterion_q,
                   npbc=npbc, min_sep=min_sep, r0=r0, graft_z=graft_z,
                   z_lo=z_range[0] if z_range else 0.0, z_hi=z_range[1] if z_range else 0.0,
                   seed=seed)
    for d in range(3):
        sp.box[d] = box[d]
    nb, nm = C.c_int(), C.c_int()
    rc = L.plum_synth_count(C.byref(sp), C.byref(nb), C.byref(nm))
    if rc != 0:
        raise HostError("plum_synth_count failed (%d)" % rc)
    xyz = np.zeros(3 * nb.value)
    q = np.zeros(nb.value)
    ty = np.zeros(nb.value, dtype=np.int32)
    ms = np.zeros(nm.value + 1, dtype=np.int32)
    rc = L.plum_synth_system(C.byref(sp), xyz.ctypes.data_as(C.POINTER(C.c_double)),
                             q.ctypes.data_as(C.POINTER(C.c_double)),
                             ty.ctypes.data_as(C.POINTER(C.c_int)),
                             ms.ctypes.data_as(C.POINTER(C.c_int)))
    if rc != 0:
        raise HostError("plum_synth_system failed (%d): too dense for the overlap rule?" % rc)
    return xyz.reshape(-1, 3), q, ty, ms
