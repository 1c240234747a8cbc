"""On-disk aerosol state: the variables of aero_state_output_netcdf /
aero_state_input_netcdf (src/aero_state.F90:2633-3286) with the weight
array of aero_weight_array_output_netcdf (src/aero_weight_array.F90:
540-700) and the species dimension of aero_data_netcdf_dim_aero_species
(src/aero_data.F90:680-710).

The reference writes NetCDF-4 (src/netcdf.F90:76).  No NetCDF-4 writer is
available here, so the file is NetCDF classic (64-bit offset) written with
scipy: the same dimension names, variable names, shapes, units and
descriptions, readable by nf90_open.  Classic NetCDF has no 64-bit integer
type, so the particle ids (NF90_INT64 in the reference) are stored as
doubles, exact below 2**53; nf90_get_var converts them to integer(kind=8)
on read.  Dimension order follows the reference's Fortran arrays: a Fortran
variable with dimensions (aero_particle, aero_species) is stored in C order
as (aero_species, aero_particle).
"""
from __future__ import annotations

import numpy as np
from scipy.io import netcdf_file

MAX_COMP = 20
INFO_DESC = ("valid is 0 (invalid entry), 1 (removed due to dilution), 2 (removed due to "
             "coagulation -- combined particle ID is in \\c aero_removed_other_id), 3 (removed "
             "due to populating halving), or 4 (removed due to weighting changes")


def _var(f, name, dtype, dims, data, **attrs):
    v = f.createVariable(name, dtype, dims)
    v[...] = data
    for k, a in attrs.items():
        setattr(v, k, a)
    return v


def write_aero_state(filename, state, density, weights, species_names=None, removed=None):
    """Write one box's state.

    state: dict as returned by CoagGPU.download (vol (n, n_spec), id, group,
    klass, n_primary, n_comp, comp_source, comp_time, least, greatest);
    weights: (type, magnitude, exponent), each (n_group, n_class);
    removed: optional (id, action, other_id) of the removal log."""
    vol = np.asarray(state["vol"], dtype=np.float64)
    n, n_spec = vol.shape
    density = np.asarray(density, dtype=np.float64)
    wtype, wmag, wexp = (np.atleast_2d(np.asarray(a)) for a in weights)
    n_group, n_class = wtype.shape
    names = species_names or [f"species_{i + 1}" for i in range(n_spec)]
    with netcdf_file(filename, "w", version=2) as f:
        # aero_weight_array_output_netcdf
        f.createDimension("aero_weight_group", n_group)
        f.createDimension("aero_weight_class", n_class)
        _var(f, "aero_weight_group", "i4", ("aero_weight_group",), np.arange(1, n_group + 1),
             description="dummy dimension variable (no useful value)")
        _var(f, "aero_weight_class", "i4", ("aero_weight_class",), np.arange(1, n_class + 1),
             description="dummy dimension variable (no useful value)")
        wdims = ("aero_weight_class", "aero_weight_group")
        _var(f, "weight_type", "i4", wdims, wtype.T.astype(np.int32),
             description="type of each aerosol weighting function: 0 = invalid, "
             "1 = none (w(D) = 1), 2 = power (w(D) = (D/D_0)^alpha), "
             "3 = MFA (mass flow) (w(D) = (D/D_0)^(-3))")
        _var(f, "weight_magnitude", "f8", wdims, wmag.T.astype(np.float64), unit="m^{-3}",
             description="magnitude for each weighting function")
        _var(f, "weight_exponent", "f8", wdims, wexp.T.astype(np.float64), unit="1",
             description="exponent alpha for the power weight_type, "
             "set to -3 for MFA, and zero otherwise")
        # aero_data_netcdf_dim_aero_species
        f.createDimension("aero_species", n_spec)
        _var(f, "aero_species", "i4", ("aero_species",), np.arange(1, n_spec + 1),
             names=",".join(names),
             description="dummy dimension variable (no useful value) - read species names "
             "as comma-separated values from the 'names' attribute")
        if n > 0:
            f.createDimension("aero_particle", n)
            _var(f, "aero_particle", "i4", ("aero_particle",), np.arange(1, n + 1),
                 description="dummy dimension variable (no useful value)")
            n_comp = np.asarray(state["n_comp"], dtype=np.int32)
            tot = int(n_comp.sum())
            f.createDimension("aero_components", max(tot, 1))
            _var(f, "aero_components", "i4", ("aero_components",),
                 np.arange(1, max(tot, 1) + 1),
                 description="dummy dimension variable (no useful value)")
            start = np.concatenate([[1], 1 + np.cumsum(n_comp)[:-1]]).astype(np.int32)
            part_num = np.repeat(np.arange(1, n + 1), n_comp).astype(np.int32)
            cs = np.asarray(state["comp_source"]).reshape(n, MAX_COMP)
            ct = np.asarray(state["comp_time"]).reshape(n, MAX_COMP)
            mask = np.arange(MAX_COMP)[None, :] < n_comp[:, None]
            _var(f, "aero_particle_mass", "f8", ("aero_species", "aero_particle"),
                 (vol * density[None, :]).T, unit="kg",
                 long_name="constituent masses of each aerosol particle")
            _var(f, "aero_component_len", "i4", ("aero_particle",), n_comp,
                 long_name="number of aero_components for each aerosol particle")
            _var(f, "aero_component_start_ind", "i4", ("aero_particle",), start,
                 long_name="start index of aero_component for each aerosol particle")
            _var(f, "aero_component_particle_num", "i4", ("aero_components",),
                 part_num if tot else [0],
                 long_name="associated aerosol particle number for each component")
            _var(f, "aero_component_source_num", "i4", ("aero_components",),
                 cs[mask].astype(np.int32) if tot else [0],
                 long_name="associated source number for each component")
            _var(f, "aero_component_create_time", "f8", ("aero_components",),
                 ct[mask] if tot else [0.0],
                 long_name="associated creation time for each component")
            _var(f, "aero_particle_weight_group", "i4", ("aero_particle",), state["group"],
                 long_name="weight group number of each aerosol particle")
            _var(f, "aero_particle_weight_class", "i4", ("aero_particle",), state["klass"],
                 long_name="weight class number of each aerosol particle")
            _var(f, "aero_water_hyst_leg", "i4", ("aero_particle",), np.zeros(n, np.int32),
                 long_name="leg of the water hysteresis curve leg of each aerosol particle")
            _var(f, "aero_num_conc", "f8", ("aero_particle",), state.get("num_conc", np.zeros(n)),
                 unit="m^{-3}", long_name="number concentration for each particle")
            _var(f, "aero_id", "f8", ("aero_particle",), np.asarray(state["id"], np.float64),
                 long_name="unique ID number of each aerosol particle")
            _var(f, "aero_frozen", "i4", ("aero_particle",), np.zeros(n, np.int32),
                 long_name="flag indicating ice phase of each aerosol particle")
            _var(f, "aero_imf_temperature", "f8", ("aero_particle",), np.zeros(n),
                 long_name="immersion freezing temperature (Singular)")
            _var(f, "aero_ice_density", "f8", ("aero_particle",), np.full(n, np.nan),
                 long_name="Ice density if the particle nucleates to ice, -9999 indicates "
                 "the particle is not an ice.")
            _var(f, "aero_ice_shape_phi", "f8", ("aero_particle",), np.full(n, np.nan),
                 long_name="Ice shape parameter phi")
            _var(f, "aero_least_create_time", "f8", ("aero_particle",), state["least"], unit="s",
                 long_name="least creation time of each aerosol particle")
            _var(f, "aero_greatest_create_time", "f8", ("aero_particle",), state["greatest"],
                 unit="s", long_name="greatest creation time of each aerosol particle")
            _var(f, "aero_particle_n_primary_parts", "i4", ("aero_particle",),
                 state["n_primary"], unit="1",
                 long_name="number of primary particles that contributeto each particle.")
        if removed is not None:
            rid, ract, roth = (np.asarray(a) for a in removed)
            m = len(rid)
            f.createDimension("aero_removed", max(m, 1))
            _var(f, "aero_removed", "i4", ("aero_removed",), np.arange(1, max(m, 1) + 1),
                 description="dummy dimension variable (no useful value)")
            _var(f, "aero_removed_id", "f8", ("aero_removed",),
                 rid.astype(np.float64) if m else [0.0], long_name="ID of removed particles")
            _var(f, "aero_removed_action", "i4", ("aero_removed",),
                 ract.astype(np.int32) if m else [0], long_name="reason for particle removal",
                 description=INFO_DESC)
            _var(f, "aero_removed_other_id", "f8", ("aero_removed",),
                 roth.astype(np.float64) if m else [0.0],
                 long_name="ID of other particle involved in removal")


def read_aero_state(filename):
    """The state written by write_aero_state (or the aero_state variables of
    a reference output file in NetCDF classic): volumes from the masses and
    the species densities implied by them are not stored, so vol is
    mass / density with density given back by the caller."""
    with netcdf_file(filename, "r", mmap=False) as f:
        v = f.variables
        out = {"weights": (np.asarray(v["weight_type"][:]).T.copy(),
                           np.asarray(v["weight_magnitude"][:]).T.copy(),
                           np.asarray(v["weight_exponent"][:]).T.copy()),
               "species_names": v["aero_species"].names.decode().split(",")}
        if "aero_particle" in f.dimensions:
            mass = np.asarray(v["aero_particle_mass"][:]).T.copy()
            n = mass.shape[0]
            n_comp = np.asarray(v["aero_component_len"][:], dtype=np.int32)
            start = np.asarray(v["aero_component_start_ind"][:], dtype=np.int64) - 1
            src = np.asarray(v["aero_component_source_num"][:], dtype=np.int32)
            ctime = np.asarray(v["aero_component_create_time"][:], dtype=np.float64)
            cs = np.zeros((n, MAX_COMP), dtype=np.int32)
            ct = np.zeros((n, MAX_COMP))
            for i in range(n):
                k = n_comp[i]
                cs[i, :k] = src[start[i]:start[i] + k]
                ct[i, :k] = ctime[start[i]:start[i] + k]
            out.update(mass=mass, id=np.asarray(v["aero_id"][:]).astype(np.int64),
                       group=np.asarray(v["aero_particle_weight_group"][:], dtype=np.int32),
                       klass=np.asarray(v["aero_particle_weight_class"][:], dtype=np.int32),
                       n_primary=np.asarray(v["aero_particle_n_primary_parts"][:], dtype=np.int32),
                       n_comp=n_comp, comp_source=cs, comp_time=ct,
                       least=np.asarray(v["aero_least_create_time"][:], dtype=np.float64),
                       greatest=np.asarray(v["aero_greatest_create_time"][:], dtype=np.float64),
                       num_conc=np.asarray(v["aero_num_conc"][:], dtype=np.float64))
        if "aero_removed" in f.dimensions:
            act = np.asarray(v["aero_removed_action"][:], dtype=np.int32)
            keep = act != 0  # a single AERO_INFO_NONE entry stands for an empty log
            out["removed"] = (np.asarray(v["aero_removed_id"][:]).astype(np.int64)[keep],
                              act[keep],
                              np.asarray(v["aero_removed_other_id"][:]).astype(np.int64)[keep])
    return out
