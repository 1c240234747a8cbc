"""partmc_amd -- MI355X (gfx950) particle-resolved Monte Carlo coagulation.

Host-side Python mirror of the C ABI in ``include/partmc_coag.h``.  The
compute path is the HIP library ``libpartmc_coag.so``; there is no CPU
fallback: constructing :class:`CoagGPU` without the built library raises.

The entry point that mirrors the reference is :meth:`CoagGPU.mc_coag`, the
counterpart of ``mc_coag(coag_kernel_type, env_state, aero_data, aero_state,
del_t, tot_n_samp, tot_n_coag)`` (``src/coagulation.F90:43-96``).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpartmc_coag.so")

MAX_COMP = 20
MODE_EXACT, MODE_PARALLEL = 0, 1
PREC_FP64, PREC_FP32_RECHECK = 0, 1
KERNEL_SEDI, KERNEL_ADDITIVE, KERNEL_CONSTANT, KERNEL_BROWN, KERNEL_ZERO = 1, 2, 3, 4, 5
WEIGHT_NONE, WEIGHT_POWER, WEIGHT_MFA = 1, 2, 3

_d, _i, _i64, _u32, _u64, _p = (ctypes.c_double, ctypes.c_int, ctypes.c_int64,
                                ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p)

# name -> (restype, argtypes); the complete exported ABI
ABI = {
    "pmc_gpu_create": (_i, [_p, _i, _i, _i, _p, _i]),
    "pmc_gpu_destroy": (_i, [_p]),
    "pmc_gpu_last_error": (ctypes.c_char_p, []),
    "pmc_gpu_set_env": (_i, [_p, _i, _d, _d, _d]),
    "pmc_gpu_set_weights": (_i, [_p, _i, _i, _i, _p, _p, _p]),
    "pmc_gpu_set_rng": (_i, [_p, _u64, _u32]),
    "pmc_gpu_set_next_id": (_i, [_p, _i, _i64, _i64, _i64]),
    "pmc_gpu_get_next_id": (_i, [_p, _i, _p]),
    "pmc_gpu_reserve": (_i, [_p, _p]),
    "pmc_gpu_get_part_steps": (_i, [_p, _p]),
    "pmc_gpu_get_path_stats": (_i, [_p, _p, _p, _p, _p]),
    "pmc_gpu_get_arrival_stats": (_i, [_p, _p, _p]),
    "pmc_gpu_get_accel_counts": (_i, [_p, _p, _p]),
    "pmc_gpu_get_step_lists": (_i, [_p, _i, _p]),
    "pmc_gpu_record": (_i, [_p, _i, _i64, _i]),
    "pmc_gpu_replay": (_i, [_p, _i, _p, _i64, _i]),
    "pmc_gpu_get_trace": (_i, [_p, _i, _p, _i64, _p, _p]),
    "pmc_gpu_trace_off": (_i, [_p, _i]),
    "pmc_gpu_size_dist": (_i, [_p, _i, _i, _i, _d, _d, _p]),
    "pmc_gpu_size_dist_noalloc": (_i, [_p, _i, _i, _i, _d, _d, _p]),
    "pmc_gpu_set_n_part_ideal": (_i, [_p, _i, _d]),
    "pmc_gpu_clear_removed": (_i, [_p, _i]),
    "pmc_gpu_dilute": (_i, [_p, _p]),
    "pmc_gpu_particle_loss": (_i, [_p, _i, _d, _p]),
    "pmc_gpu_scale_weights": (_i, [_p, _i, _d]),
    "pmc_gpu_append": (_i, [_p, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "pmc_gpu_rebalance": (_i, [_p, _i, _i]),
    "pmc_gpu_get_weights": (_i, [_p, _i, _i, _p, _p, _p, _p, _p]),
    "pmc_gpu_get_error_stats": (_i, [_p, _p]),
    "pmc_gpu_set_timing": (_i, [_p, _i]),
    "pmc_gpu_get_kernel_time": (_i, [_p, ctypes.c_char_p, _p, _p]),
    "pmc_gpu_upload": (_i, [_p, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "pmc_gpu_set_bin_grid": (_i, [_p, _i, _i, _d, _d]),
    "pmc_gpu_get_n_part": (_i, [_p, _i, _p]),
    "pmc_gpu_download": (_i, [_p, _i, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p]),
    "pmc_gpu_set_mode": (_i, [_p, _i, _i, _d]),
    "pmc_gpu_mc_coag": (_i, [_p, _i, _d, _p, _p]),
    "pmc_gpu_mc_coag_async": (_i, [_p, _i, _d]),
    "pmc_gpu_sync": (_i, [_p]),
    "pmc_gpu_get_counters": (_i, [_p, _p, _p, _p, _p, _p, _p]),
    "pmc_gpu_reset_counters": (_i, [_p]),
    "pmc_gpu_set_validation": (_i, [_p, _i]),
    "pmc_gpu_get_grid": (_i, [_p, _i, _p, _p, _p, _p]),
    "pmc_gpu_get_lists": (_i, [_p, _i, _p, _p]),
    "pmc_gpu_get_kbounds": (_i, [_p, _i, _p, _p]),
    "pmc_gpu_get_removed": (_i, [_p, _i, _i, _p, _p, _p, _p]),
    "pmc_gpu_eval_kernel": (_i, [_p, _i, _i, _p, _p, _d, _d, _d, _p, _p]),
    "pmc_gpu_rng_draw": (_i, [_p, _u64, _u32, _u32, _u32, _u32, _u32, _i, _d, _d, _i, _p]),
}

_lib = None


class CoagError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m partmc_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(path)
        for name, (res, args) in ABI.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(code: int):
    if code != 0:
        raise CoagError(code, load().pmc_gpu_last_error().decode())


class CoagGPU:
    """A handle over ``n_box`` independent particle boxes on one GPU."""

    def __init__(self, n_box: int, density: Sequence[float], device: int = 0,
                 max_bin: int = 256):
        L = load()
        self.density = np.ascontiguousarray(density, dtype=np.float64)
        self.n_spec = len(self.density)
        self.n_box = n_box
        h = ctypes.c_void_p()
        _check(L.pmc_gpu_create(ctypes.byref(h), device, n_box, self.n_spec,
                                _ptr(self.density), max_bin))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().pmc_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration -------------------------------------------------
    def set_env(self, box, temp, pressure, beta=1000.0):
        _check(load().pmc_gpu_set_env(self._h, box, temp, pressure, beta))

    def set_weights(self, box, wtype, magnitude, exponent):
        t = np.asfortranarray(np.atleast_2d(wtype), dtype=np.int32)
        m = np.asfortranarray(np.atleast_2d(magnitude), dtype=np.float64)
        e = np.asfortranarray(np.atleast_2d(exponent), dtype=np.float64)
        ng, nc = t.shape
        _check(load().pmc_gpu_set_weights(self._h, box, ng, nc,
                                          _ptr(np.ascontiguousarray(t.ravel(order="F"))),
                                          _ptr(np.ascontiguousarray(m.ravel(order="F"))),
                                          _ptr(np.ascontiguousarray(e.ravel(order="F")))))

    def set_rng(self, seed: int, step: int = 0):
        _check(load().pmc_gpu_set_rng(self._h, seed, step))

    def set_next_id(self, box, next_id, stride=1, offset=0):
        _check(load().pmc_gpu_set_next_id(self._h, box, next_id, stride, offset))

    def next_id(self, box) -> int:
        v = np.zeros(1, dtype=np.int64)
        _check(load().pmc_gpu_get_next_id(self._h, box, _ptr(v)))
        return int(v[0])

    def set_mode(self, mode=MODE_EXACT, precision=PREC_FP64, band=1e-5):
        _check(load().pmc_gpu_set_mode(self._h, mode, precision, band))

    def set_validation(self, on: bool):
        _check(load().pmc_gpu_set_validation(self._h, int(on)))

    # -- particles -----------------------------------------------------
    def upload(self, box, vol, ids, group, klass, n_primary=None, capacity=0,
               n_comp=None, comp_source=None, comp_time=None, least=None, greatest=None):
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(-1, self.n_spec)
        n = vol.shape[0]
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        g = np.ascontiguousarray(group, dtype=np.int32)
        c = np.ascontiguousarray(klass, dtype=np.int32)
        npr = None if n_primary is None else np.ascontiguousarray(n_primary, dtype=np.int32)
        nco = None if n_comp is None else np.ascontiguousarray(n_comp, dtype=np.int32)
        cs = None if comp_source is None else np.ascontiguousarray(comp_source, dtype=np.int32)
        ct = None if comp_time is None else np.ascontiguousarray(comp_time, dtype=np.float64)
        le = None if least is None else np.ascontiguousarray(least, dtype=np.float64)
        gr = None if greatest is None else np.ascontiguousarray(greatest, dtype=np.float64)
        _check(load().pmc_gpu_upload(self._h, box, n, capacity, _ptr(vol), _ptr(ids), _ptr(g),
                                     _ptr(c), _ptr(npr), _ptr(nco), _ptr(cs), _ptr(ct),
                                     _ptr(le), _ptr(gr)))

    def save_state(self, box, filename, species_names=None, record_removals=True):
        """Write the box as aero_state_output_netcdf lays it out (NetCDF
        classic; see partmc_amd.state_io)."""
        from .state_io import write_aero_state
        d = self.download(box)
        wt, wm, we = self.weights(box)
        r = (3.0 * d["vol"].sum(axis=1) / 4.0 / np.pi) ** (1.0 / 3.0)
        inv = np.zeros(len(r))
        for gi in range(wt.shape[0]):
            ci = d["klass"] - 1
            w = wm[gi, ci] * np.where(wt[gi, ci] == 1, 1.0, r ** we[gi, ci])
            inv += 1.0 / w
        d["num_conc"] = 1.0 / inv if len(r) else inv
        rem = self.removed(box) if record_removals else None
        write_aero_state(filename, d, self.density, (wt, wm, we), species_names, rem)

    def load_state(self, box, filename, capacity=0):
        """Restart the box from a file written by save_state, as
        aero_state_input_netcdf reads one: volumes are mass / density."""
        from .state_io import read_aero_state
        s = read_aero_state(filename)
        self.set_weights(box, *s["weights"])
        if "mass" not in s:
            self.upload(box, np.zeros((0, self.n_spec)), np.zeros(0, np.int64), np.zeros(0),
                        np.zeros(0), capacity=capacity)
            return s
        vol = s["mass"] / self.density[None, :]
        self.upload(box, vol, s["id"], s["group"], s["klass"], s["n_primary"], capacity,
                    s["n_comp"], s["comp_source"], s["comp_time"], s["least"], s["greatest"])
        return s

    def set_bin_grid(self, box, n_bin, r_min, r_max):
        _check(load().pmc_gpu_set_bin_grid(self._h, box, n_bin, r_min, r_max))

    def n_part(self, box) -> int:
        v = np.zeros(1, dtype=np.int32)
        _check(load().pmc_gpu_get_n_part(self._h, box, _ptr(v)))
        return int(v[0])

    def download(self, box):
        n = self.n_part(box)
        m = max(n, 1)
        vol = np.zeros((m, self.n_spec))
        ids = np.zeros(m, dtype=np.int64)
        g = np.zeros(m, dtype=np.int32)
        c = np.zeros(m, dtype=np.int32)
        npr = np.zeros(m, dtype=np.int32)
        nco = np.zeros(m, dtype=np.int32)
        cs = np.zeros((m, MAX_COMP), dtype=np.int32)
        ct = np.zeros((m, MAX_COMP))
        le = np.zeros(m)
        gr = np.zeros(m)
        cnt = np.zeros(1, dtype=np.int32)
        _check(load().pmc_gpu_download(self._h, box, m, _ptr(cnt), _ptr(vol), _ptr(ids), _ptr(g),
                                       _ptr(c), _ptr(npr), _ptr(nco), _ptr(cs), _ptr(ct),
                                       _ptr(le), _ptr(gr)))
        n = int(cnt[0])
        return dict(vol=vol[:n], id=ids[:n], group=g[:n], klass=c[:n], n_primary=npr[:n],
                    n_comp=nco[:n], comp_source=cs[:n], comp_time=ct[:n], least=le[:n],
                    greatest=gr[:n])

    def download_ids_vols(self, box):
        """The box's particle ids and species volumes only (n, n_spec)."""
        n = self.n_part(box)
        m = max(n, 1)
        vol = np.zeros((m, self.n_spec))
        ids = np.zeros(m, dtype=np.int64)
        cnt = np.zeros(1, dtype=np.int32)
        _check(load().pmc_gpu_download(self._h, box, m, _ptr(cnt), _ptr(vol), _ptr(ids), None,
                                       None, None, None, None, None, None, None))
        n = int(cnt[0])
        return ids[:n], vol[:n]

    # -- the step --------------------------------------------------------
    def mc_coag(self, coag_kernel_type: int, del_t: float):
        """One coagulation timestep for all boxes; returns per-box
        (tot_n_samp, tot_n_coag) arrays."""
        ns = np.zeros(self.n_box, dtype=np.int64)
        nc = np.zeros(self.n_box, dtype=np.int64)
        _check(load().pmc_gpu_mc_coag(self._h, coag_kernel_type, del_t, _ptr(ns), _ptr(nc)))
        return ns, nc

    def mc_coag_async(self, coag_kernel_type: int, del_t: float):
        _check(load().pmc_gpu_mc_coag_async(self._h, coag_kernel_type, del_t))

    def sync(self):
        _check(load().pmc_gpu_sync(self._h))

    def counters(self):
        arr = {k: np.zeros(self.n_box, dtype=np.int64)
               for k in ("n_samp", "n_coag", "n_warn", "n_recheck", "n_flip")}
        err = np.zeros(self.n_box)
        _check(load().pmc_gpu_get_counters(self._h, *[_ptr(arr[k]) for k in
                                                      ("n_samp", "n_coag", "n_warn",
                                                       "n_recheck", "n_flip")], _ptr(err)))
        arr["max_rel_err"] = err
        return arr

    def reset_counters(self):
        _check(load().pmc_gpu_reset_counters(self._h))

    def error_stats(self):
        out = np.zeros(self.n_box)
        _check(load().pmc_gpu_get_error_stats(self._h, _ptr(out)))
        return {"max_p_err": out}

    def part_steps(self):
        out = np.zeros(self.n_box, dtype=np.int64)
        _check(load().pmc_gpu_get_part_steps(self._h, _ptr(out)))
        return out

    def dilute(self, sample_prob):
        """Dilution out of every box: particle leaves with sample_prob[box]."""
        p = np.ascontiguousarray(sample_prob, dtype=np.float64).reshape(self.n_box)
        _check(load().pmc_gpu_dilute(self._h, _ptr(p)))

    def particle_loss(self, loss_function, del_t, height=None):
        """scenario_particle_loss with the reference's loss function code."""
        h = None if height is None else np.ascontiguousarray(height, dtype=np.float64).reshape(self.n_box)
        _check(load().pmc_gpu_particle_loss(self._h, loss_function, del_t, _ptr(h)))

    def scale_weights(self, box, factor):
        _check(load().pmc_gpu_scale_weights(self._h, box, float(factor)))

    def append(self, box, vol, group, klass, ids=None, n_primary=None):
        """aero_state_add_particle for each row of vol; ids None: new ids."""
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(-1, self.n_spec)
        n = vol.shape[0]
        g = np.ascontiguousarray(group, dtype=np.int32)
        c = np.ascontiguousarray(klass, dtype=np.int32)
        i = None if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        npr = None if n_primary is None else np.ascontiguousarray(n_primary, dtype=np.int32)
        _check(load().pmc_gpu_append(self._h, box, n, _ptr(vol), _ptr(i), _ptr(g), _ptr(c),
                                     _ptr(npr), None, None, None, None, None))

    def clear_removed(self, box):
        _check(load().pmc_gpu_clear_removed(self._h, box))

    def set_n_part_ideal(self, box, n_part):
        _check(load().pmc_gpu_set_n_part_ideal(self._h, box, float(n_part)))

    def rebalance(self, allow_doubling=True, allow_halving=True):
        """aero_state_rebalance on every box."""
        _check(load().pmc_gpu_rebalance(self._h, int(allow_doubling), int(allow_halving)))

    def weights(self, box):
        """(type, magnitude, exponent), each of shape (n_group, n_class)."""
        n = 64
        ng = np.zeros(1, dtype=np.int32)
        nc = np.zeros(1, dtype=np.int32)
        t = np.zeros(n, dtype=np.int32)
        m = np.zeros(n)
        e = np.zeros(n)
        _check(load().pmc_gpu_get_weights(self._h, box, n, _ptr(ng), _ptr(nc), _ptr(t), _ptr(m),
                                          _ptr(e)))
        g, c = int(ng[0]), int(nc[0])
        shape = lambda a: a[: g * c].reshape(c, g).T.copy()
        return shape(t), shape(m), shape(e)

    def size_dist(self, box, n_bin, d_min, d_max, quantity="num"):
        """bin_grid_histogram_1d of the box's diameters weighted by number
        concentration (times mass or volume), per ln D, on a log grid."""
        q = {"num": 0, "mass": 1, "vol": 2}[quantity]
        out = np.zeros(n_bin)
        _check(load().pmc_gpu_size_dist(self._h, box, q, n_bin, d_min, d_max, _ptr(out)))
        return out

    def size_dist_noalloc(self, box, n_bin, d_min, d_max, d_hist, quantity="num"):
        """size_dist into caller-owned device memory (the integer address
        d_hist of n_bin doubles), enqueued on the handle's stream without
        any allocation; sync() waits for it."""
        q = {"num": 0, "mass": 1, "vol": 2}[quantity]
        _check(load().pmc_gpu_size_dist_noalloc(self._h, box, q, n_bin, d_min, d_max,
                                                ctypes.c_void_p(int(d_hist))))

    def record(self, box, capacity, ref_samplers=False):
        """Exact mode: write every uniform the box draws, in call order."""
        _check(load().pmc_gpu_record(self._h, box, int(capacity), int(bool(ref_samplers))))

    def replay(self, box, uniforms, ref_samplers=False):
        """Exact mode: draw the box's uniforms from `uniforms`, in order."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        _check(load().pmc_gpu_replay(self._h, box, _ptr(u), len(u), int(bool(ref_samplers))))

    def trace(self, box, max_n=0):
        """(uniforms used, recorded values up to max_n, ran out)."""
        n = np.zeros(1, dtype=np.int64)
        ro = np.zeros(1, dtype=np.int32)
        out = np.zeros(max(int(max_n), 1))
        _check(load().pmc_gpu_get_trace(self._h, box, _ptr(out), int(max_n), _ptr(n), _ptr(ro)))
        used = int(n[0])
        return used, out[: min(used, int(max_n))].copy(), bool(ro[0])

    def trace_off(self, box):
        _check(load().pmc_gpu_trace_off(self._h, box))

    def arrival_stats(self):
        """Per box: particles that joined a list within a step, and those of
        them the list's reserved arrival entries could not hold."""
        arr = np.zeros(self.n_box, dtype=np.int64)
        over = np.zeros(self.n_box, dtype=np.int64)
        _check(load().pmc_gpu_get_arrival_stats(self._h, _ptr(arr), _ptr(over)))
        return {"n_arrivals": arr, "n_overflow": over}

    def step_lists(self, box, n_class=1):
        """The parallel mode's list sizes at the start of its last step,
        (n_class, n_bin)."""
        e, _, _ = self.grid(box)
        nb = len(e) - 1
        counts = np.zeros(max(nb * n_class, 1), dtype=np.int32)
        _check(load().pmc_gpu_get_step_lists(self._h, box, _ptr(counts)))
        return counts[: nb * n_class].reshape(n_class, nb)

    def accel_counts(self):
        """Per box: samples and coagulations of the per-particle path."""
        ns = np.zeros(self.n_box, dtype=np.int64)
        nc = np.zeros(self.n_box, dtype=np.int64)
        _check(load().pmc_gpu_get_accel_counts(self._h, _ptr(ns), _ptr(nc)))
        return {"n_samp": ns, "n_coag": nc}

    def path_stats(self):
        """Per box: pairs on the per-particle path and dropped draws; for
        the handle: claim rounds run and tests finished in the sequential
        tail."""
        acc = np.zeros(self.n_box, dtype=np.int64)
        drop = np.zeros(self.n_box, dtype=np.int64)
        rounds = np.zeros(4, dtype=np.int64)
        tail = np.zeros((self.n_box, 2), dtype=np.int64)
        _check(load().pmc_gpu_get_path_stats(self._h, _ptr(acc), _ptr(drop), _ptr(rounds),
                                             _ptr(tail)))
        return {"n_accel_pairs": acc, "n_dropped": drop, "n_rounds": int(rounds[0]),
                "n_tail": int(rounds[1]), "n_tail_rounds": int(rounds[2]),
                "acc_samples": int(rounds[3]),
                "box_tail_tests": tail[:, 0].copy(), "box_tail_rounds": tail[:, 1].copy()}

    def reserve(self, capacity):
        cap = np.ascontiguousarray(capacity, dtype=np.int64)
        _check(load().pmc_gpu_reserve(self._h, _ptr(cap)))

    def set_timing(self, on: bool):
        _check(load().pmc_gpu_set_timing(self._h, int(on)))

    def kernel_time(self, name: str):
        ms = np.zeros(1)
        n = np.zeros(1, dtype=np.int64)
        _check(load().pmc_gpu_get_kernel_time(self._h, name.encode(), _ptr(ms), _ptr(n)))
        return float(ms[0]), int(n[0])

    # -- inspection ------------------------------------------------------
    def grid(self, box, max_bin=1024):
        nb = np.zeros(1, dtype=np.int32)
        e = np.zeros(max_bin + 1)
        c = np.zeros(max_bin)
        w = np.zeros(max_bin)
        _check(load().pmc_gpu_get_grid(self._h, box, _ptr(nb), _ptr(e), _ptr(c), _ptr(w)))
        n = int(nb[0])
        return e[: n + 1], c[:n], w[:n]

    def lists(self, box, n_class=1):
        e, _, _ = self.grid(box)
        nb = len(e) - 1
        counts = np.zeros(max(nb * n_class, 1), dtype=np.int32)
        _check(load().pmc_gpu_get_lists(self._h, box, _ptr(counts), None))
        entries = np.zeros(max(int(counts.sum()), 1), dtype=np.int32)
        _check(load().pmc_gpu_get_lists(self._h, box, _ptr(counts), _ptr(entries)))
        return counts[: nb * n_class].reshape(n_class, nb), entries[: int(counts.sum())]

    def kbounds(self, box):
        e, _, _ = self.grid(box)
        nb = len(e) - 1
        kmin = np.zeros(nb * nb)
        kmax = np.zeros(nb * nb)
        _check(load().pmc_gpu_get_kbounds(self._h, box, _ptr(kmin), _ptr(kmax)))
        return kmin.reshape(nb, nb).T.copy(), kmax.reshape(nb, nb).T.copy()

    def removed(self, box, max_n=1 << 20):
        n = np.zeros(1, dtype=np.int32)
        ids = np.zeros(max_n, dtype=np.int64)
        act = np.zeros(max_n, dtype=np.int32)
        oth = np.zeros(max_n, dtype=np.int64)
        _check(load().pmc_gpu_get_removed(self._h, box, max_n, _ptr(n), _ptr(ids), _ptr(act),
                                          _ptr(oth)))
        k = int(n[0])
        return ids[:k], act[:k], oth[:k]

    def eval_kernel(self, kernel_type, vol1, vol2, temp=288.0, pressure=1e5, beta=1000.0):
        v1 = np.ascontiguousarray(vol1, dtype=np.float64).reshape(-1, self.n_spec)
        v2 = np.ascontiguousarray(vol2, dtype=np.float64).reshape(-1, self.n_spec)
        n = v1.shape[0]
        k64 = np.zeros(n)
        k32 = np.zeros(n, dtype=np.float32)
        _check(load().pmc_gpu_eval_kernel(self._h, kernel_type, n, _ptr(v1), _ptr(v2), temp,
                                          pressure, beta, _ptr(k64), _ptr(k32)))
        return k64, k32

    def rng_draw(self, seed, box, step, site, ordinal, kind, which, p1=0.0, p2=0.0, n=1):
        out = np.zeros(n)
        _check(load().pmc_gpu_rng_draw(self._h, seed, box, step, site, ordinal, kind, which,
                                       p1, p2, n, _ptr(out)))
        return out

    # -- convenience -----------------------------------------------------
    def load_population(self, box, pop, capacity=0, next_id=None):
        """Upload a :class:`partmc_amd.populations.Population` into ``box``."""
        n = pop.n_part
        self.set_env(box, pop.temp, pop.pressure, pop.beta)
        self.set_weights(box, [[pop.weight_type]], [[pop.weight_magnitude]],
                         [[pop.weight_exponent]])
        self.upload(box, pop.vol, pop.ids(), np.ones(n, dtype=np.int32),
                    np.ones(n, dtype=np.int32), capacity=capacity)
        self.set_next_id(box, n + 1 if next_id is None else next_id)
        if pop.grid is not None:
            self.set_bin_grid(box, *pop.grid)
