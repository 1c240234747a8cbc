"""CPU oracle for PartMC's mc_coag -- TEST INFRASTRUCTURE ONLY.

A ctypes wrapper over ``libcoag_oracle.so`` (built from ``coag_oracle.c`` by
``oracle/Makefile``).  The oracle is the checker for the HIP implementation;
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  Nothing in ``partmc_amd`` imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcoag_oracle.so")

MAX_COMP = 20
KERNEL_SEDI, KERNEL_ADDITIVE, KERNEL_CONSTANT, KERNEL_BROWN, KERNEL_ZERO = 1, 2, 3, 4, 5
WEIGHT_NONE, WEIGHT_POWER, WEIGHT_MFA = 1, 2, 3
RNG_POISSON, RNG_TEST, RNG_MERGE = 1, 2, 3

_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_philox4x32_10": (None, [_p, _p, _p]),
            "orc_stream_open": (None, [_p, _u64, _u32, _u32, _u32, _u32, _u32]),
            "orc_stream_uniform": (_d, [_p]),
            "orc_rand_int": (_i, [_p, _i]),
            "orc_rand_poisson": (_i, [_p, _d]),
            "orc_rand_binomial": (_i, [_p, _i, _d]),
            "orc_rand_normal": (_d, [_p, _d, _d]),
            "orc_prob_round": (_i, [_p, _d]),
            "orc_sphere_vol2rad": (_d, [_d]),
            "orc_sphere_rad2vol": (_d, [_d]),
            "orc_powi": (_d, [_d, _i]),
            "orc_fast_pow": (_d, [_d, _d]),
            "orc_interp_linear_disc": (_d, [_d, _d, _i, _i]),
            "orc_linspace_find": (_i, [_d, _d, _i, _d]),
            "orc_logspace_find": (_i, [_d, _d, _i, _d]),
            "orc_logspace": (None, [_d, _d, _i, _p]),
            "orc_bin_grid_make_log": (None, [_i, _d, _d, _p, _p, _p]),
            "orc_bin_grid_find": (_i, [_i, _p, _d]),
            "orc_student_t_95_coeff": (_d, [_i]),
            "orc_fall_g": (None, [_d, _p]),
            "orc_effic": (_i, [_d, _d, _p]),
            "orc_brown_terms": (None, [_d, _d, _d, _d, _p]),
            "orc_kernel_brown_helper": (_d, [_d, _d, _d, _d, _d, _d]),
            "orc_kernel_sedi_helper": (_i, [_d, _d, _p]),
            "orc_kernel": (_i, [_i, _i, _p, _p, _p, _d, _d, _d, _p]),
            "orc_kernel_minmax": (_i, [_i, _i, _p, _d, _d, _d, _d, _d, _p, _p]),
            "orc_est_k_minmax_binned": (_i, [_i, _i, _p, _i, _p, _d, _d, _d, _p, _p]),
            "orc_num_conc_at_radius": (_d, [_i, _i, _p, _p, _p, _i, _d]),
            "orc_coag_num_conc_factor": (_d, [_i, _i, _p, _p, _p, _d, _d, _i, _i, _i]),
            "orc_coag_dest_class": (_i, [_i, _i, _p, _p, _p, _p, _i, _i, _i, _i]),
            "orc_max_coag_num_conc_factor": (_d, [_i, _i, _p, _p, _p, _p, _i, _i, _i, _i, _i]),
            "orc_bessi1": (_d, [_d]),
            "orc_soln_additive_exp": (None, [_i, _p, _d, _d, _d, _d, _p, _p]),
            "orc_soln_constant_exp": (None, [_i, _p, _d, _d, _d, _p, _p]),
            "orc_histogram_1d": (None, [_i, _p, _p, _i, _p, _p, _p]),
            "orc_sample_pmf": (None, [_u64, _i, _d, _d, _i64, _i, _p]),
            "orc_box_create": (_p, [_i, _p, _i, _i, _p, _p, _p]),
            "orc_box_destroy": (None, [_p]),
            "orc_box_set_env": (None, [_p, _d, _d, _d]),
            "orc_box_set_rng": (None, [_p, _u64, _u32]),
            "orc_box_set_next_id": (None, [_p, _i64]),
            "orc_box_get_next_id": (_i64, [_p]),
            "orc_box_sort_with_grid": (_i, [_p, _i, _d, _d]),
            "orc_box_add_particles": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p]),
            "orc_box_n_part": (_i, [_p]),
            "orc_box_get_particles": (None, [_p, _p, _p, _p, _p, _p, _p, _p, _p]),
            "orc_box_n_bin": (_i, [_p]),
            "orc_box_get_grid": (None, [_p, _p, _p, _p]),
            "orc_box_get_lists": (None, [_p, _p, _p]),
            "orc_box_get_kmax": (None, [_p, _p, _p]),
            "orc_box_n_info": (_i, [_p]),
            "orc_box_get_info": (None, [_p, _p, _p, _p]),
            "orc_box_clear_info": (None, [_p]),
            "orc_box_n_warn_kmax": (_i64, [_p]),
            "orc_box_n_accel": (_i64, [_p]),
            "orc_box_regridded": (_i, [_p]),
            "orc_mc_coag": (_i, [_p, _i, _d, _u32, _p, _p]),
            "orc_box_double": (_i, [_p, _i, _i]),
            "orc_box_halve": (_i, [_p, _i, _i, _u32, _u32]),
            "orc_box_rebalance": (_i, [_p, _i, _i, _p, _u32]),
            "orc_box_get_weights": (None, [_p, _p]),
            "orc_box_scale_weights": (None, [_p, _d]),
            "orc_loss_rate_dry_dep": (_d, [_d, _d, _d, _d, _d]),
            "orc_loss_rate": (_i, [_i, _d, _d, _d, _d, _d, _p]),
            "orc_box_dilute": (_i, [_p, _d, _u32]),
            "orc_box_particle_loss": (_i, [_p, _i, _d, _d, _u32]),
            "orc_box_rand_group": (_i, [_p, _u64, _u32, _i, _d]),
            "orc_box_set_tape": (_i, [_p, _i, _p, _i64, _i]),
            "orc_box_tape_state": (_i64, [_p, _p, _i64, _p]),
            "orc_box_coagulate_components": (_i, [_p, _u64, _u32, _i, _i, _p, _i, _i, _p, _p,
                                                  _p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Stream(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint32 * 2), ("ctr", ctypes.c_uint32 * 4),
                ("buf", ctypes.c_uint32 * 4), ("have", ctypes.c_int),
                ("block", ctypes.c_uint32), ("tape", ctypes.c_void_p)]

    def __init__(self, seed, box, step, site, ordinal, kind):
        super().__init__()
        lib().orc_stream_open(ctypes.byref(self), seed, box, step, site, ordinal, kind)

    def uniform(self):
        return lib().orc_stream_uniform(ctypes.byref(self))

    def rand_int(self, n):
        return lib().orc_rand_int(ctypes.byref(self), n)

    def poisson(self, mean):
        return lib().orc_rand_poisson(ctypes.byref(self), mean)

    def binomial(self, n, p):
        return lib().orc_rand_binomial(ctypes.byref(self), n, p)

    def normal(self, mean, sd):
        return lib().orc_rand_normal(ctypes.byref(self), mean, sd)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def sample_pmf(seed, which, p1, p2, n, kmax):
    counts = np.zeros(kmax + 1, dtype=np.int64)
    lib().orc_sample_pmf(seed, which, p1, p2, n, kmax, _ptr(counts))
    return counts


def bin_grid_make_log(n_bin, rmin, rmax):
    e = np.zeros(n_bin + 1)
    c = np.zeros(n_bin)
    w = np.zeros(n_bin)
    lib().orc_bin_grid_make_log(n_bin, rmin, rmax, _ptr(e), _ptr(c), _ptr(w))
    return e, c, w


def histogram_1d(edges, widths, x, w):
    """bin_grid_histogram_1d (src/bin_grid.F90:225-248) on a log grid."""
    e = np.ascontiguousarray(edges, dtype=np.float64)
    wd = np.ascontiguousarray(widths, dtype=np.float64)
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ws = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros(len(wd))
    lib().orc_histogram_1d(len(wd), _ptr(e), _ptr(wd), len(xs), _ptr(xs), _ptr(ws), _ptr(out))
    return out


def loss_rate(loss_type, vol, density, temp=288.0, pressure=1e5, height=1000.0):
    """scenario_loss_rate (src/scenario.F90:398-437)."""
    r = np.zeros(1)
    err = lib().orc_loss_rate(loss_type, vol, density, temp, pressure, height, _ptr(r))
    if err:
        raise RuntimeError(f"oracle loss_rate error {err}")
    return float(r[0])


LOSS_NONE, LOSS_CONSTANT, LOSS_VOLUME, LOSS_DRYDEP, LOSS_CHAMBER = 1, 2, 3, 4, 5


def kernel(kind, density, vol1, vol2, temp=288.0, pressure=1e5, beta=1000.0):
    dens = np.ascontiguousarray(density, dtype=np.float64)
    v1 = np.ascontiguousarray(vol1, dtype=np.float64)
    v2 = np.ascontiguousarray(vol2, dtype=np.float64)
    k = np.zeros(1)
    err = lib().orc_kernel(kind, len(dens), _ptr(dens), _ptr(v1), _ptr(v2),
                           temp, pressure, beta, _ptr(k))
    if err:
        raise RuntimeError(f"oracle kernel error {err}")
    return float(k[0])


def brown_terms(vol, density, temp=288.0, pressure=1e5):
    """kernel_brown_helper's air and particle terms: air mean free path,
    radius, Knudsen number, Cunningham factor, diffusivity, mean speed
    squared, particle mean free path and delta squared (SI units)."""
    out = np.zeros(8)
    lib().orc_brown_terms(vol, density, temp, pressure, _ptr(out))
    keys = ("air_mfp", "radius", "knudsen", "cunningham", "diffusivity", "speed_sq",
            "mfp", "delta_sq")
    return dict(zip(keys, out))


def est_k_minmax(kind, density, edges, temp=288.0, pressure=1e5, beta=1000.0):
    dens = np.ascontiguousarray(density, dtype=np.float64)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    n = len(e) - 1
    kmin = np.zeros(n * n)
    kmax = np.zeros(n * n)
    err = lib().orc_est_k_minmax_binned(kind, len(dens), _ptr(dens), n, _ptr(e),
                                        temp, pressure, beta, _ptr(kmin), _ptr(kmax))
    if err:
        raise RuntimeError(f"oracle error {err}")
    # index (b1-1) + n*(b2-1): column-major n x n
    return kmin.reshape(n, n).T.copy(), kmax.reshape(n, n).T.copy()


class Weights:
    """aero_weight_array: arrays of shape (n_group, n_class)."""

    def __init__(self, wtype, magnitude, exponent):
        self.wtype = np.ascontiguousarray(np.atleast_2d(wtype), dtype=np.int32)
        self.mag = np.ascontiguousarray(np.atleast_2d(magnitude), dtype=np.float64)
        self.exp = np.ascontiguousarray(np.atleast_2d(exponent), dtype=np.float64)
        self.n_group, self.n_class = self.wtype.shape
        # Fortran order g + n_group*c
        self._t = np.asfortranarray(self.wtype).ravel(order="F").astype(np.int32)
        self._m = np.asfortranarray(self.mag).ravel(order="F").astype(np.float64)
        self._e = np.asfortranarray(self.exp).ravel(order="F").astype(np.float64)

    def args(self):
        return (self.n_group, self.n_class, _ptr(self._t), _ptr(self._m), _ptr(self._e))

    def num_conc_at_radius(self, klass, r):
        return lib().orc_num_conc_at_radius(*self.args(), klass, r)

    def coag_num_conc_factor(self, i_r, j_r, ic, jc, ijc):
        return lib().orc_coag_num_conc_factor(*self.args(), i_r, j_r, ic, jc, ijc)

    def dest_class(self, centers, ib, jb, ic, jc):
        c = np.ascontiguousarray(centers, dtype=np.float64)
        return lib().orc_coag_dest_class(*self.args(), _ptr(c), ib, jb, ic, jc)

    def max_factor(self, edges, ib, jb, ic, jc, ijc):
        e = np.ascontiguousarray(edges, dtype=np.float64)
        return lib().orc_max_coag_num_conc_factor(*self.args(), _ptr(e), ib, jb, ic, jc, ijc)


class Box:
    """The oracle's aero_state for one box."""

    def __init__(self, density, weights: Weights, temp=288.0, pressure=1e5,
                 beta=1000.0, seed=0, box_index=0):
        self.density = np.ascontiguousarray(density, dtype=np.float64)
        self.n_spec = len(self.density)
        self.weights = weights
        self._h = lib().orc_box_create(self.n_spec, _ptr(self.density), *weights.args())
        if not self._h:
            raise RuntimeError("orc_box_create failed")
        lib().orc_box_set_env(self._h, temp, pressure, beta)
        lib().orc_box_set_rng(self._h, seed, box_index)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_box_destroy(h)
            self._h = None

    def set_env(self, temp, pressure, beta):
        lib().orc_box_set_env(self._h, temp, pressure, beta)

    @property
    def next_id(self):
        return lib().orc_box_get_next_id(self._h)

    @next_id.setter
    def next_id(self, v):
        lib().orc_box_set_next_id(self._h, v)

    def add_particles(self, vol, ids, group, klass, n_primary=None, n_comp=None,
                      comp_source=None, comp_time=None):
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(-1, self.n_spec)
        n = vol.shape[0]
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        g = np.ascontiguousarray(group, dtype=np.int32)
        c = np.ascontiguousarray(klass, dtype=np.int32)
        npr = None if n_primary is None else np.ascontiguousarray(n_primary, dtype=np.int32)
        comps = (None, None, None)
        if n_comp is not None:
            comps = (np.ascontiguousarray(n_comp, dtype=np.int32),
                     np.ascontiguousarray(comp_source, dtype=np.int32).reshape(n, MAX_COMP),
                     np.ascontiguousarray(comp_time, dtype=np.float64).reshape(n, MAX_COMP))
        err = lib().orc_box_add_particles(self._h, n, _ptr(vol), _ptr(ids), _ptr(g), _ptr(c),
                                          None if npr is None else _ptr(npr),
                                          *[None if a is None else _ptr(a) for a in comps])
        if err:
            raise RuntimeError(f"oracle add error {err}")

    def sort_with_grid(self, n_bin, rmin, rmax):
        err = lib().orc_box_sort_with_grid(self._h, n_bin, rmin, rmax)
        if err:
            raise RuntimeError(f"oracle sort error {err}")

    @property
    def n_part(self):
        return lib().orc_box_n_part(self._h)

    @property
    def n_bin(self):
        return lib().orc_box_n_bin(self._h)

    def particles(self):
        n = self.n_part
        vol = np.zeros((n, self.n_spec))
        ids = np.zeros(n, dtype=np.int64)
        g = np.zeros(n, dtype=np.int32)
        c = np.zeros(n, dtype=np.int32)
        npr = np.zeros(n, dtype=np.int32)
        nc = np.zeros(n, dtype=np.int32)
        cs = np.zeros((n, MAX_COMP), dtype=np.int32)
        ct = np.zeros((n, MAX_COMP), dtype=np.float64)
        lib().orc_box_get_particles(self._h, _ptr(vol), _ptr(ids), _ptr(g), _ptr(c),
                                    _ptr(npr), _ptr(nc), _ptr(cs), _ptr(ct))
        return dict(vol=vol, id=ids, group=g, klass=c, n_primary=npr, n_comp=nc,
                    comp_source=cs, comp_time=ct)

    def grid(self):
        n = self.n_bin
        e = np.zeros(n + 1)
        c = np.zeros(max(n, 1))
        w = np.zeros(max(n, 1))
        lib().orc_box_get_grid(self._h, _ptr(e), _ptr(c), _ptr(w))
        return e, c[:n], w[:n]

    def lists(self):
        """Per (class, bin) entry lists of 0-based particle indices."""
        n_bin = self.n_bin
        n_class = self.weights.n_class
        counts = np.zeros(n_bin * n_class, dtype=np.int32)
        entries = np.zeros(max(self.n_part, 1), dtype=np.int32)
        lib().orc_box_get_lists(self._h, _ptr(counts), _ptr(entries))
        return counts.reshape(n_class, n_bin), entries[: int(counts.sum())]

    def kmax(self):
        n = self.n_bin
        kmin = np.zeros(n * n)
        kmax = np.zeros(n * n)
        lib().orc_box_get_kmax(self._h, _ptr(kmin), _ptr(kmax))
        return kmin.reshape(n, n).T.copy(), kmax.reshape(n, n).T.copy()

    def info(self):
        n = lib().orc_box_n_info(self._h)
        ids = np.zeros(n, dtype=np.int64)
        act = np.zeros(n, dtype=np.int32)
        oth = np.zeros(n, dtype=np.int64)
        if n:
            lib().orc_box_get_info(self._h, _ptr(ids), _ptr(act), _ptr(oth))
        return ids, act, oth

    def clear_info(self):
        lib().orc_box_clear_info(self._h)

    @property
    def n_warn_kmax(self):
        return lib().orc_box_n_warn_kmax(self._h)

    @property
    def n_accel(self):
        """Bin pairs that took the per-particle (accelerated) path."""
        return lib().orc_box_n_accel(self._h)

    @property
    def regridded(self):
        return bool(lib().orc_box_regridded(self._h))

    def rebalance(self, n_part_ideal, step, allow_doubling=True, allow_halving=True):
        """aero_state_rebalance; n_part_ideal is a scalar split evenly over
        the (group, class) pairs, as aero_state_set_n_part_ideal does."""
        ng, nc = self.weights.n_group, self.weights.n_class
        ideal = np.full(ng * nc, float(n_part_ideal) / (ng * nc))
        err = lib().orc_box_rebalance(self._h, int(allow_doubling), int(allow_halving),
                                      _ptr(ideal), step)
        if err:
            raise RuntimeError(f"oracle rebalance error {err}")

    def double(self, g=1, c=1):
        lib().orc_box_double(self._h, g, c)

    def halve(self, step, site, g=1, c=1):
        lib().orc_box_halve(self._h, g, c, step, site)

    def scale_weights(self, factor):
        lib().orc_box_scale_weights(self._h, factor)

    def dilute(self, sample_prob, step):
        err = lib().orc_box_dilute(self._h, sample_prob, step)
        if err:
            raise RuntimeError(f"oracle dilute error {err}")

    def particle_loss(self, loss_type, del_t, height, step):
        err = lib().orc_box_particle_loss(self._h, loss_type, del_t, height, step)
        if err:
            raise RuntimeError(f"oracle particle_loss error {err}")

    def weight_magnitudes(self):
        ng, nc = self.weights.n_group, self.weights.n_class
        m = np.zeros(ng * nc)
        lib().orc_box_get_weights(self._h, _ptr(m))
        return m.reshape(nc, ng).T.copy()

    def rand_group(self, seed, ordinal, klass, radius):
        """aero_weight_array_rand_group on stream (seed, ordinal)."""
        return lib().orc_box_rand_group(self._h, seed, ordinal, klass, radius)

    def coagulate_components(self, seed, ordinal, src1, np1, src2, np2):
        """The merged particle's component sources from aero_particle_coagulate
        of two particles with component sources src1, src2 and np1, np2
        primary parts, on stream (seed, ordinal)."""
        s1 = np.ascontiguousarray(src1, dtype=np.int32)
        s2 = np.ascontiguousarray(src2, dtype=np.int32)
        n = np.zeros(1, dtype=np.int32)
        out = np.zeros(MAX_COMP, dtype=np.int32)
        err = lib().orc_box_coagulate_components(self._h, seed, ordinal, len(s1), np1,
                                                 _ptr(s1), len(s2), np2, _ptr(s2), _ptr(n),
                                                 _ptr(out))
        if err:
            raise RuntimeError(f"oracle coagulate error {err}")
        return out[: int(n[0])].copy()

    def record(self, capacity, ref_samplers=False):
        """Write every uniform the next mc_coag calls draw, in call order."""
        if lib().orc_box_set_tape(self._h, 1, None, int(capacity), int(bool(ref_samplers))):
            raise RuntimeError("oracle record")

    def replay(self, uniforms, ref_samplers=False):
        """Draw the next mc_coag calls' uniforms from `uniforms`, in order."""
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        self._tape = u
        if lib().orc_box_set_tape(self._h, 2, _ptr(u), len(u), int(bool(ref_samplers))):
            raise RuntimeError("oracle replay")

    def trace(self, max_n=0):
        """(uniforms used, recorded values up to max_n, ran out)."""
        out = np.zeros(max(int(max_n), 1))
        ro = np.zeros(1, dtype=np.int32)
        used = lib().orc_box_tape_state(self._h, _ptr(out), int(max_n), _ptr(ro))
        return int(used), out[: min(int(used), int(max_n))].copy(), bool(ro[0])

    def trace_off(self):
        lib().orc_box_set_tape(self._h, 0, None, 0, 0)

    def mc_coag(self, kernel_type, del_t, step):
        ns = np.zeros(1, dtype=np.int64)
        nc = np.zeros(1, dtype=np.int64)
        err = lib().orc_mc_coag(self._h, kernel_type, del_t, step, _ptr(ns), _ptr(nc))
        if err:
            raise RuntimeError(f"oracle mc_coag error {err}")
        return int(ns[0]), int(nc[0])
