"""ctypes binding of the C ABI in include/plum_gpu.h (libplum_gpu.so).

This is the product path: every call goes to the HIP engine.  There is no
CPU fallback; if the shared library or a GPU is missing, the calls raise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libplum_gpu.so")
MAX_TYPES = 16
TERMS = ("lj", "real", "kspace", "self", "dipole", "bond", "wall")

PLUM_FP64 = 0
PLUM_FP32_PAIR = 1

# every symbol include/plum_gpu.h declares
ABI_SYMBOLS = (
    "plum_create", "plum_destroy", "plum_get_info", "plum_energy_full",
    "plum_energy_difference", "plum_finalize", "plum_totals", "plum_resync",
    "plum_get_positions", "plum_last_error", "plum_mc_run", "plum_set_profiling",
    "plum_kernel_times", "plum_batch_create", "plum_batch_destroy", "plum_batch_mc_run",
    "plum_batch_tune", "plum_batch_get_tuning", "plum_batch_set_tuning", "plum_mc_get_tuning",
    "plum_mc_set_tuning", "plum_surrogate_enable", "plum_surrogate_disable",
    "plum_surrogate_stats_get", "plum_surrogate_get_weights", "plum_surrogate_set_weights",
    "plum_surrogate_train", "plum_surrogate_forward", "plum_surrogate_gradient",
    "plum_surrogate_debug_slot", "plum_max_mol_len", "plum_cbmc_trials", "plum_cbmc_grow",
    "plum_mc_record", "plum_mc_get_records", "plum_vp_sample", "plum_cbmc_trials_noalloc",
)
SURROGATE_PARAMS = 64 * 128 + 128 + 2 * (128 * 128 + 128) + 128 + 1
DESCR = 64


class PlumError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("box", C.c_double * 3), ("npbc", C.c_int), ("beta", C.c_double),
        ("use_pair", C.c_int), ("use_ewald", C.c_int), ("use_bond", C.c_int),
        ("use_ext", C.c_int),
        ("lj_cutoff", C.c_double), ("n_types", C.c_int),
        ("lj_sigma", C.c_double * MAX_TYPES), ("lj_eps", C.c_double * MAX_TYPES),
        ("lB", C.c_double), ("dielectric", C.c_double), ("alpha", C.c_double),
        ("dipole_correction", C.c_int),
        ("k_bond", C.c_double), ("r0", C.c_double),
        ("wall_cut", C.c_double), ("wall_sigma_w", C.c_double), ("wall_eps_w", C.c_double),
        ("wall_sigma", C.c_double * MAX_TYPES), ("wall_eps", C.c_double * MAX_TYPES),
        ("type_kind", C.c_int * MAX_TYPES),
        ("precision", C.c_int), ("dipole_exact", C.c_int), ("cell_div", C.c_int),
    ]


class Terms(C.Structure):
    _fields_ = [(t, C.c_double) for t in TERMS]

    def as_dict(self):
        return {t: getattr(self, t) for t in TERMS}


class Info(C.Structure):
    _fields_ = [
        ("box_e", C.c_double * 3), ("vol_e", C.c_double), ("real_cutoff", C.c_double),
        ("real_cell", C.c_int * 3), ("repl_cutoff", C.c_double), ("repl_cell", C.c_int * 3),
        ("n_k_full", C.c_int), ("n_k_half", C.c_int), ("n_cells", C.c_int * 3),
        ("cell_cap", C.c_int), ("n_beads", C.c_int), ("n_mols", C.c_int),
        ("max_mol_len", C.c_int),
    ]


class MCParams(C.Structure):
    _fields_ = [
        ("move_prob", C.c_double * 5), ("move_size", C.c_double), ("phantom", C.c_int),
        ("grafted", C.c_int), ("bond_len", C.c_double), ("vary_bond", C.c_int),
        ("seed", C.c_ulonglong), ("resync_every", C.c_int),
    ]


class MCStats(C.Structure):
    _fields_ = [
        ("attempted", C.c_longlong * 5), ("accepted", C.c_longlong * 5),
        ("moves", C.c_longlong), ("seconds", C.c_double), ("totals", Terms),
    ]


class BatchTuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int), ("split_div", C.c_int),
                ("batches_per_graph", C.c_int), ("pmax", C.c_int), ("sweep_us", C.c_double),
                ("cell_div", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [
        ("moves", C.c_longlong), ("sweeps", C.c_longlong), ("attempted", C.c_longlong * 5),
        ("accepted", C.c_longlong * 5), ("overflow_replicas", C.c_longlong),
        ("seconds", C.c_double),
    ]


class SurrogateParams(C.Structure):
    _fields_ = [("screen", C.c_int), ("fill", C.c_int), ("train_every", C.c_int),
                ("train_samples", C.c_int), ("ring", C.c_int), ("lr", C.c_double),
                ("r_d", C.c_double), ("target_clip", C.c_double), ("seed", C.c_ulonglong)]


class CbmcRequest(C.Structure):
    _fields_ = [("n_trials", C.c_int), ("grown", C.c_int), ("exclude_lo", C.c_int),
                ("exclude_hi", C.c_int), ("type_chain", C.c_int), ("type_ion", C.c_int),
                ("charged", C.c_int), ("q_chain", C.c_double), ("q_ion", C.c_double)]


class GrowthRequest(C.Structure):
    _fields_ = [("mode", C.c_int), ("chain_len", C.c_int), ("n_trials", C.c_int),
                ("exclude_lo", C.c_int), ("exclude_hi", C.c_int), ("type_chain", C.c_int),
                ("type_ion", C.c_int), ("charged", C.c_int), ("q_chain", C.c_double),
                ("q_ion", C.c_double)]


class Decision(C.Structure):
    _fields_ = [("move", C.c_ulonglong), ("mol", C.c_int), ("n_moved", C.c_int),
                ("move_type", C.c_int), ("accepted", C.c_int), ("evaluated", C.c_int),
                ("early", C.c_int), ("first_stage_pass", C.c_int), ("lo", C.c_int),
                ("hi", C.c_int), ("dE", Terms), ("total", C.c_double),
                ("dE_surrogate", C.c_double)]


class VpSample(C.Structure):
    _fields_ = [("dU", C.c_double), ("el", C.c_double), ("lj", C.c_double), ("wall", C.c_double),
                ("bond", C.c_double), ("dipole", C.c_double), ("n_mol", C.c_int),
                ("vol", C.c_double), ("area_dz", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SurrogateStats(C.Structure):
    _fields_ = [("s1_rejected", C.c_longlong), ("s1_passed", C.c_longlong),
                ("train_evals", C.c_longlong), ("stale_cuts", C.c_longlong),
                ("samples", C.c_longlong), ("train_steps", C.c_longlong),
                ("last_loss", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib():
    """Load libplum_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PlumError("libplum_gpu.so is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.plum_create.argtypes = [C.POINTER(Params), C.c_int, dp, dp, ip, C.c_int, ip,
                                  C.POINTER(vp)]
        L.plum_destroy.argtypes = [vp]
        L.plum_get_info.argtypes = [vp, C.POINTER(Info)]
        L.plum_energy_full.argtypes = [vp, C.POINTER(Terms)]
        L.plum_energy_difference.argtypes = [vp, C.c_int, C.c_int, ip, dp, C.POINTER(Terms), dp]
        L.plum_finalize.argtypes = [vp, C.c_int]
        L.plum_totals.argtypes = [vp, C.POINTER(Terms)]
        L.plum_resync.argtypes = [vp, C.POINTER(Terms)]
        L.plum_get_positions.argtypes = [vp, dp]
        L.plum_last_error.restype = C.c_char_p
        L.plum_mc_run.argtypes = [vp, C.POINTER(MCParams), C.c_longlong, C.POINTER(MCStats)]
        L.plum_set_profiling.argtypes = [vp, C.c_int]
        L.plum_kernel_times.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), dp,
                                        C.POINTER(C.c_longlong)]
        L.plum_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.plum_batch_destroy.argtypes = [vp]
        L.plum_batch_mc_run.argtypes = [vp, C.POINTER(MCParams), C.c_longlong, C.c_int,
                                        C.POINTER(BatchStats), C.POINTER(MCStats)]
        L.plum_batch_tune.argtypes = [vp, C.POINTER(MCParams), C.c_int, C.POINTER(BatchTuning)]
        L.plum_batch_get_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
        L.plum_batch_set_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
        L.plum_mc_get_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
        L.plum_mc_set_tuning.argtypes = [vp, C.POINTER(BatchTuning)]
        fp = C.POINTER(C.c_float)
        L.plum_surrogate_enable.argtypes = [vp, C.POINTER(SurrogateParams)]
        L.plum_surrogate_disable.argtypes = [vp]
        L.plum_surrogate_stats_get.argtypes = [vp, C.POINTER(SurrogateStats)]
        L.plum_surrogate_get_weights.argtypes = [vp, fp, C.c_int]
        L.plum_surrogate_set_weights.argtypes = [vp, fp, C.c_int]
        L.plum_surrogate_train.argtypes = [vp, C.c_int, dp]
        L.plum_surrogate_forward.argtypes = [vp, fp, C.c_int, C.c_int, fp]
        L.plum_surrogate_gradient.argtypes = [vp, fp, dp, ip, C.c_int, dp, fp]
        L.plum_surrogate_debug_slot.argtypes = [vp, C.c_int, fp, dp, ip]
        L.plum_max_mol_len.argtypes = [vp]
        L.plum_cbmc_trials.argtypes = [vp, C.POINTER(CbmcRequest), dp, dp, dp]
        L.plum_cbmc_grow.argtypes = [vp, C.POINTER(GrowthRequest), dp, dp, dp, dp, dp, dp,
                                     C.POINTER(C.c_int), dp]
        L.plum_vp_sample.argtypes = [vp, C.c_int, C.POINTER(VpSample)]
        L.plum_cbmc_trials_noalloc.argtypes = [vp, C.POINTER(CbmcRequest), dp, dp, dp]
        L.plum_mc_record.argtypes = [vp, C.c_int]
        L.plum_mc_get_records.argtypes = [vp, C.c_int, C.POINTER(Decision), ip, dp,
                                          C.POINTER(C.c_longlong)]
        for f in ABI_SYMBOLS:
            if f != "plum_last_error":
                getattr(L, f).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise PlumError("plum error %d: %s" % (rc, lib().plum_last_error().decode()))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_params(sysdef, precision=PLUM_FP64, dipole_exact=0, cell_div=0):
    p = Params()
    for i in range(3):
        p.box[i] = sysdef["box"][i]
    p.npbc = sysdef["npbc"]
    p.beta = sysdef.get("beta", 1.0)
    p.use_pair = int(sysdef.get("use_pair", 1))
    p.use_ewald = int(sysdef.get("use_ewald", 1))
    p.use_bond = int(sysdef.get("use_bond", 1))
    p.use_ext = int(sysdef.get("use_ext", 0))
    p.lj_cutoff = sysdef.get("lj_cutoff", 2.5)
    types = sysdef["types"]
    p.n_types = len(types)
    for t, td in enumerate(types):
        p.lj_sigma[t] = td.get("sigma", 0.0)
        p.lj_eps[t] = td.get("eps", 0.0)
        p.wall_sigma[t] = td.get("wall_sigma", 0.0)
        p.wall_eps[t] = td.get("wall_eps", 0.0)
        p.type_kind[t] = {"L": 1, "R": 2}.get(td["name"], 0)
    p.lB = sysdef.get("lB", 1.0)
    p.dielectric = sysdef.get("dielectric", 80.0)
    p.alpha = sysdef.get("alpha", 0.05)
    p.dipole_correction = int(sysdef.get("dipole", 0))
    p.k_bond = sysdef.get("k_bond", 400.0)
    p.r0 = sysdef.get("r0", 1.0)
    p.wall_cut = sysdef.get("wall_cut", -1.0)
    p.wall_sigma_w = sysdef.get("wall_sigma_w", 1.0)
    p.wall_eps_w = sysdef.get("wall_eps_w", 1.0)
    p.precision = precision
    p.dipole_exact = dipole_exact
    p.cell_div = cell_div
    return p


def mc_params(move_prob, move_size=1.0, seed=1, phantom=0, grafted=0, bond_len=1.0,
              vary_bond=True, resync_every=0, out=None):
    mp = MCParams() if out is None else out
    for i in range(5):
        mp.move_prob[i] = move_prob[i]
    mp.move_size = move_size
    mp.phantom = phantom
    mp.grafted = grafted
    mp.bond_len = bond_len
    mp.vary_bond = int(vary_bond)
    mp.seed = seed
    mp.resync_every = resync_every
    return mp


class Engine:
    """One system on the GPU, driven through the C ABI."""

    def __init__(self, sysdef, precision=PLUM_FP64, dipole_exact=0, cell_div=0):
        L = lib()
        self.params = make_params(sysdef, precision, dipole_exact, cell_div)
        xyz = np.ascontiguousarray(sysdef["xyz"], dtype=np.float64).reshape(-1)
        q = np.ascontiguousarray(sysdef["q"], dtype=np.float64)
        ty = np.ascontiguousarray(sysdef["type"], dtype=np.int32)
        ms = np.ascontiguousarray(sysdef["mol_start"], dtype=np.int32)
        self.n = len(q)
        self.mol_start = ms
        h = C.c_void_p()
        _check(L.plum_create(C.byref(self.params), self.n, _dp(xyz), _dp(q), _ip(ty),
                             len(ms) - 1, _ip(ms), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().plum_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = Info()
        _check(lib().plum_get_info(self.h, C.byref(i)))
        return i

    def energy_full(self):
        t = Terms()
        _check(lib().plum_energy_full(self.h, C.byref(t)))
        return t.as_dict()

    def delta(self, mol, idx, trial_xyz):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        tr = np.ascontiguousarray(trial_xyz, dtype=np.float64).reshape(-1)
        t = Terms()
        tot = C.c_double()
        _check(lib().plum_energy_difference(self.h, mol, len(idx), _ip(idx), _dp(tr),
                                            C.byref(t), C.byref(tot)))
        return tot.value, t.as_dict()

    def finalize(self, accept):
        _check(lib().plum_finalize(self.h, int(accept)))

    def totals(self):
        t = Terms()
        _check(lib().plum_totals(self.h, C.byref(t)))
        return t.as_dict()

    def resync(self):
        t = Terms()
        _check(lib().plum_resync(self.h, C.byref(t)))
        return t.as_dict()

    def positions(self):
        out = np.zeros(3 * self.n)
        _check(lib().plum_get_positions(self.h, _dp(out)))
        return out.reshape(-1, 3)

    def mc_run(self, n_moves, move_prob, move_size=1.0, seed=1, phantom=0, grafted=0,
               bond_len=1.0, vary_bond=True, resync_every=0):
        mp = mc_params(move_prob, move_size, seed, phantom, grafted, bond_len, vary_bond,
                       resync_every)
        st = MCStats()
        _check(lib().plum_mc_run(self.h, C.byref(mp), n_moves, C.byref(st)))
        return st

    def mc_record(self, capacity):
        """Arms a record of the next `capacity` decided proposals of mc_run
        (0 disarms)."""
        _check(lib().plum_mc_record(self.h, int(capacity)))

    def mc_records(self, cap):
        """Recorded decisions in order: a list of dicts (move, mol, type,
        accepted, evaluated, early, pass, dE per term, total, dE_s, idx,
        trial (n, 3)) and the number of proposals decided since arming."""
        m = self.max_mol_len()
        out = (Decision * max(cap, 1))()
        idx = np.zeros(max(cap, 1) * m, dtype=np.int32)
        xyz = np.zeros(max(cap, 1) * m * 3)
        nd = C.c_longlong()
        n = lib().plum_mc_get_records(self.h, cap, out, _ip(idx), _dp(xyz), C.byref(nd))
        if n < 0:
            _check(n)
        recs = []
        for k in range(n):
            o = out[k]
            nm = o.n_moved
            recs.append({
                "move": o.move, "mol": o.mol, "n": nm, "type": o.move_type,
                "accepted": bool(o.accepted), "evaluated": bool(o.evaluated),
                "early": bool(o.early), "pass": bool(o.first_stage_pass), "lo": o.lo,
                "hi": o.hi, "dE": o.dE.as_dict(), "total": o.total,
                "dE_s": o.dE_surrogate,
                "idx": idx[k * m:k * m + nm].copy(),
                "trial": xyz[3 * k * m:3 * (k * m + nm)].reshape(-1, 3).copy(),
            })
        return recs, nd.value

    def vp_sample(self, phantom=0):
        """One volume perturbation of the pressure estimator (plum_vp_sample)."""
        o = VpSample()
        _check(lib().plum_vp_sample(self.h, int(phantom), C.byref(o)))
        return o.as_dict()

    def set_profiling(self, on):
        _check(lib().plum_set_profiling(self.h, int(on)))

    # ---- delayed acceptance with the learned surrogate
    def surrogate_enable(self, screen=32, fill=4, train_every=16, train_samples=64, ring=1024,
                         lr=1e-3, r_d=4.0, target_clip=20.0, seed=1):
        p = SurrogateParams(screen=screen, fill=fill, train_every=train_every,
                            train_samples=train_samples, ring=ring, lr=lr, r_d=r_d,
                            target_clip=target_clip, seed=seed)
        _check(lib().plum_surrogate_enable(self.h, C.byref(p)))

    def surrogate_disable(self):
        _check(lib().plum_surrogate_disable(self.h))

    def surrogate_stats(self):
        st = SurrogateStats()
        _check(lib().plum_surrogate_stats_get(self.h, C.byref(st)))
        return st.as_dict()

    def surrogate_weights(self, w=None):
        fp = C.POINTER(C.c_float)
        if w is None:
            out = np.zeros(SURROGATE_PARAMS, dtype=np.float32)
            _check(lib().plum_surrogate_get_weights(self.h, out.ctypes.data_as(fp), len(out)))
            return out
        w = np.ascontiguousarray(w, dtype=np.float32)
        _check(lib().plum_surrogate_set_weights(self.h, w.ctypes.data_as(fp), len(w)))

    def surrogate_train(self, steps):
        loss = C.c_double()
        _check(lib().plum_surrogate_train(self.h, steps, C.byref(loss)))
        return loss.value

    def cbmc_trials(self, trials, type_chain, q_chain, grown=None, ions=None, type_ion=0,
                    q_ion=0.0, grown_ions=None, exclude=None):
        """Energies of K trial positions of one configurational-bias growth
        step (ForceField::BeadsEnergy, cbmc.cc:5-151): trials (K, 3) chain
        beads, ions (K, 3) their counterions for a charged chain, grown (cur,
        3) the chain beads placed so far and grown_ions their counterions,
        exclude = (lo, hi) a bead index range left out of the system."""
        t = np.ascontiguousarray(trials, dtype=np.float64).reshape(-1, 3)
        K = len(t)
        ch = ions is not None
        tri = np.concatenate([t, np.ascontiguousarray(ions, dtype=np.float64).reshape(K, 3)]) if ch else t
        g = np.zeros((0, 3)) if grown is None else np.asarray(grown, dtype=np.float64).reshape(-1, 3)
        cur = len(g)
        if ch and cur:
            g = np.concatenate([g, np.asarray(grown_ions, dtype=np.float64).reshape(cur, 3)])
        g = np.ascontiguousarray(g) if cur else np.zeros(3)
        lo, hi = exclude if exclude is not None else (0, -1)
        rq = CbmcRequest(K, cur, lo, hi, type_chain, type_ion, int(ch), q_chain, q_ion)
        tri = np.ascontiguousarray(tri)
        out = np.zeros(K)
        _check(lib().plum_cbmc_trials(self.h, C.byref(rq), _dp(tri), _dp(g), _dp(out)))
        return out

    def cbmc_grow(self, mode, L, K, type_chain, q_chain, first=None, disp=None, ions=None,
                  sel=None, actual=None, type_ion=0, q_ion=0.0, charged=True, exclude=None):
        """One whole configurational-bias growth (plum_cbmc_grow): mode 0
        insertion, 1 deletion, 2 test insertion.  disp / ions (L, K, 3) and
        sel (L,) hold the draws of steps 1..L-1 (row 0 unused); first (2, 3)
        the first bead and its counterion; actual (2 L, 3) the chain being
        deleted then its counterions.  Returns (W, stop_step, chain (2 L, 3))."""
        def arr(a, n):
            return np.zeros(n) if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        f = arr(first, 6)
        dsp = arr(disp, 3 * L * K)
        ion = arr(ions, 3 * L * K)
        sl = arr(sel, L)
        act = arr(actual, 6 * L)
        lo, hi = exclude if exclude is not None else (0, -1)
        rq = GrowthRequest(mode, L, K, lo, hi, type_chain, type_ion, int(charged), q_chain,
                           q_ion if charged else 0.0)
        w = C.c_double()
        stop = C.c_int()
        chain = np.zeros(6 * L)
        _check(lib().plum_cbmc_grow(self.h, C.byref(rq), _dp(f), _dp(dsp), _dp(ion), _dp(sl),
                                    _dp(act), C.byref(w), C.byref(stop), _dp(chain)))
        return w.value, stop.value, chain.reshape(-1, 3)

    def surrogate_forward(self, x, bf16=True):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, DESCR)
        out = np.zeros(len(x), dtype=np.float32)
        _check(lib().plum_surrogate_forward(self.h, x.ctypes.data_as(fp), len(x), int(bf16),
                                            out.ctypes.data_as(fp)))
        return out

    def surrogate_gradient(self, x, target, n):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float64)
        n = np.ascontiguousarray(n, dtype=np.int32)
        grad = np.zeros(SURROGATE_PARAMS, dtype=np.float32)
        loss = C.c_double()
        _check(lib().plum_surrogate_gradient(self.h, x.ctypes.data_as(fp), _dp(target), _ip(n),
                                             len(target), C.byref(loss), grad.ctypes.data_as(fp)))
        return loss.value, grad

    def max_mol_len(self):
        return lib().plum_max_mol_len(self.h)

    def surrogate_debug_slot(self, slot):
        m = self.max_mol_len()
        rows = np.zeros((2 * m, DESCR), dtype=np.float32)
        pos = np.zeros((2 * m, 4))
        n = C.c_int()
        _check(lib().plum_surrogate_debug_slot(self.h, slot, rows.ctypes.data_as(C.POINTER(C.c_float)),
                                               _dp(pos), C.byref(n)))
        return rows, pos, n.value

    def tuning(self):
        t = BatchTuning()
        _check(lib().plum_mc_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        """Launch shape of mc_run: blocks_per_cu, split_div, batches_per_graph,
        pmax (proposals per speculative batch)."""
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().plum_mc_set_tuning(self.h, C.byref(t)))

    def kernel_times(self, cap=32):
        names = (C.c_char_p * cap)()
        avg = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        n = lib().plum_kernel_times(self.h, cap, names, _dp(avg),
                                    cnt.ctypes.data_as(C.POINTER(C.c_longlong)))
        if n < 0:
            _check(n)
        return {names[i].decode(): (avg[i], int(cnt[i])) for i in range(n)}


class Batch:
    """Replicas (Engine objects, same precision) advanced together by the
    batched kernels.  The engines must stay alive while the batch exists."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        _check(lib().plum_batch_create(arr, len(self.engines), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().plum_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _mc_array(self, mc):
        """mc: one dict of mc_params keywords per replica."""
        arr = (MCParams * len(self.engines))()
        for r, kw in enumerate(mc):
            mc_params(out=arr[r], **kw)
        return arr

    def mc_run(self, moves_per_replica, mc, use_graph=True, per_replica=False):
        arr = self._mc_array(mc)
        st = BatchStats()
        per = (MCStats * len(self.engines))() if per_replica else None
        _check(lib().plum_batch_mc_run(self.h, arr, moves_per_replica, int(use_graph),
                                       C.byref(st), per))
        return (st, per) if per_replica else st

    def tune(self, mc, moves_per_trial=16):
        arr = self._mc_array(mc)
        t = BatchTuning()
        _check(lib().plum_batch_tune(self.h, arr, moves_per_trial, C.byref(t)))
        return t

    def tuning(self):
        t = BatchTuning()
        _check(lib().plum_batch_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().plum_batch_set_tuning(self.h, C.byref(t)))
