"""The on-disk state (partmc_amd.state_io) against the layout of
aero_state_output_netcdf / aero_state_input_netcdf (src/aero_state.F90:
2633-3286) and aero_weight_array_output_netcdf (src/aero_weight_array.F90:
540-700): names, dimension order, units and the component index rules; and
a write/read round trip."""
import numpy as np
from scipy.io import netcdf_file

from partmc_amd.state_io import MAX_COMP, read_aero_state, write_aero_state


def synthetic_state(n=60, n_spec=3, seed=2):
    rng = np.random.default_rng(seed)
    n_comp = rng.integers(1, MAX_COMP + 1, size=n).astype(np.int32)
    cs = np.zeros((n, MAX_COMP), np.int32)
    ct = np.zeros((n, MAX_COMP))
    for i, k in enumerate(n_comp):
        cs[i, :k] = rng.integers(1, 9, size=k)
        ct[i, :k] = rng.uniform(0, 3600, size=k)
    return dict(vol=10 ** rng.uniform(-24, -18, size=(n, n_spec)),
                id=rng.integers(1, 2 ** 40, size=n).astype(np.int64),
                group=rng.integers(1, 3, size=n).astype(np.int32),
                klass=np.ones(n, np.int32), n_primary=rng.integers(1, 50, size=n).astype(np.int32),
                n_comp=n_comp, comp_source=cs, comp_time=ct,
                least=ct[:, 0].copy(), greatest=ct.max(axis=1), num_conc=rng.uniform(1, 2, n))


def test_layout_follows_the_reference(tmp_path):
    st = synthetic_state()
    dens = np.array([1800.0, 1000.0, 1500.0])
    w = ([[2], [1]], [[1e9], [2e9]], [[-3.0], [0.0]])
    path = tmp_path / "state.nc"
    write_aero_state(path, st, dens, w, ["SO4", "H2O", "OC"],
                     removed=([5, 6], [2, 1], [7, 0]))
    assert open(path, "rb").read(4) == b"CDF\x02"  # NetCDF classic, 64-bit offset
    with netcdf_file(path, "r", mmap=False) as f:
        v = f.variables
        # Fortran (aero_particle, aero_species) is C order (aero_species, aero_particle)
        assert v["aero_particle_mass"].dimensions == ("aero_species", "aero_particle")
        assert v["aero_particle_mass"].unit == b"kg"
        assert v["weight_type"].dimensions == ("aero_weight_class", "aero_weight_group")
        assert v["weight_magnitude"].unit == b"m^{-3}"
        assert v["aero_species"].names == b"SO4,H2O,OC"
        np.testing.assert_allclose(v["aero_particle_mass"][:].T, st["vol"] * dens, rtol=0)
        # components: 1-based start indices, cumulative, one row per component
        start = v["aero_component_start_ind"][:]
        assert start[0] == 1
        np.testing.assert_array_equal(np.diff(start), st["n_comp"][:-1])
        pn = v["aero_component_particle_num"][:]
        assert len(pn) == st["n_comp"].sum() and pn[0] == 1 and pn[-1] == len(st["id"])
        assert v["aero_removed_action"][:].tolist() == [2, 1]


def test_round_trip(tmp_path):
    st = synthetic_state()
    dens = np.array([1800.0, 1000.0, 1500.0])
    w = ([[2, 1]], [[1e9, 3e8]], [[-1.5, 0.0]])
    path = tmp_path / "state.nc"
    write_aero_state(path, st, dens, w, removed=([], [], []))
    r = read_aero_state(path)
    # vol * density / density: within one rounding
    np.testing.assert_allclose(r["mass"] / dens, st["vol"], rtol=2.3e-16)
    for k in ("id", "group", "klass", "n_primary", "n_comp", "comp_source", "comp_time",
              "least", "greatest", "num_conc"):
        np.testing.assert_array_equal(r[k], st[k], err_msg=k)
    for a, b in zip(r["weights"], w):
        np.testing.assert_array_equal(a, np.asarray(b))
    assert all(len(x) == 0 for x in r["removed"])  # empty log: one AERO_INFO_NONE entry


def test_empty_box(tmp_path):
    st = synthetic_state(n=0)
    path = tmp_path / "empty.nc"
    write_aero_state(path, {**st, "vol": np.zeros((0, 3))}, [1.0, 1.0, 1.0],
                     ([[1]], [[1.0]], [[0.0]]))
    r = read_aero_state(path)
    assert "mass" not in r and r["species_names"] == ["species_1", "species_2", "species_3"]
