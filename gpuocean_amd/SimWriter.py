"""netCDF output of a CDKLM16 run (utils/SimWriter.py).

``SimNetCDFWriter`` writes the dimensions, coordinate variables, global
attributes and the time series of eta, hu, hv that the reference writes, so
its reader and ``CDKLM16.fromfilename`` find what they expect.  The file is
classic netCDF with 64-bit offsets, written through ``scipy.io.netcdf_file``
(the netCDF4/HDF5 library the reference uses is not available here; every
netCDF-4 reader also reads classic files).  The run is kept in memory and the
file rewritten after every record, so it is complete whenever a record has
been written.  Classic netCDF has no
compression, so the reference's ``zlib=True`` has no counterpart.  Masked
depths (land) are written as the netCDF default fill value with a
``_FillValue`` attribute, as netCDF4 writes masked arrays.
"""

import datetime
import os
import subprocess

import numpy as np
from scipy.io import netcdf_file

NC_FILL_FLOAT = np.float32(9.969209968386869e36)  # netCDF's default fill value for floats


def _filled(a):
    a = np.ma.asarray(a)
    return np.ma.filled(a.astype(np.float32), NC_FILL_FLOAT), bool(np.ma.is_masked(a))


class SimNetCDFWriter:
    """Writes the initial state at construction and one record per
    writeTimestep (utils/SimWriter.py:35-509)."""

    def __init__(self, sim, super_dir_name=None, filename=None, num_layers=1, staggered_grid=False,
                 ignore_ghostcells=False, offset_x=0, offset_y=0):
        if staggered_grid:
            raise NotImplementedError("staggered grids (FBL/CTCS) are not part of gpuocean_amd")
        if num_layers != 1:
            raise NotImplementedError("two-layer output is not part of gpuocean_amd")
        self.ignore_ghostcells = ignore_ghostcells
        self.staggered_grid = staggered_grid
        self.num_layers = num_layers
        self.timestamp = datetime.datetime.now().strftime("%Y_%m_%d-%H_%M_%S")
        self.timestamp_short = datetime.datetime.now().strftime("%Y_%m_%d")
        try:
            self.git_hash = str.strip(str(subprocess.check_output(['git', 'rev-parse', 'HEAD'],
                                                                  stderr=subprocess.DEVNULL)))
        except Exception:
            self.git_hash = "git info missing..."
        self.simulator_short = str(sim.__class__.__name__)
        self.super_dir_name = super_dir_name if super_dir_name is not None else os.getcwd()
        if not filename:
            self.dir_name = os.path.join(self.super_dir_name, "netcdf_" + self.timestamp_short)
            self.output_file_name = os.path.join(self.dir_name, self.simulator_short + "_" + self.timestamp + ".nc")
        elif os.path.split(filename)[0] == "":
            self.dir_name = os.path.join(self.super_dir_name, "netcdf_" + self.timestamp_short)
            self.output_file_name = os.path.join(self.dir_name, self.simulator_short + "_" + filename + ".nc")
        else:
            self.output_file_name = filename + ".nc"
            self.dir_name = os.path.dirname(filename)
        self.output_file_name = self.output_file_name.replace('.nc.nc', '.nc')
        if getattr(sim, "comm", None) is not None:
            self.output_file_name = self.output_file_name.replace(
                '.nc', '_' + str(sim.comm.rank) + '_' + str(sim.local_particle_id) + '.nc')
        if self.dir_name:
            os.makedirs(self.dir_name, exist_ok=True)

        nx, ny, dx, dy = sim.nx, sim.ny, sim.dx, sim.dy
        gx = gy = 2   # ghost cells on every side
        Hi, Hm = sim.downloadBathymetry()
        # The whole run is kept here and the file rewritten after every
        # record, so it is a complete netCDF file at all times.
        self.attrs = {
            "Conventions": "CF-1.4",
            "git_hash": self.git_hash,
            "ignore_ghostcells": str(ignore_ghostcells),
            "num_layers": np.int32(num_layers),
            "staggered_grid": str(staggered_grid),
            "simulator_short": self.simulator_short,
            "boundary_conditions": str(sim.boundary_conditions),
            # sponge widths printed as NumPy 1.x printed its integers
            "boundary_conditions_sponge_mr": str({k: int(v) for k, v in sim.boundary_conditions.getSponge().items()}),
            "time_integrator": np.int32(sim.rk_order),
            "minmod_theta": np.float32(sim.theta),
            "coriolis_force": np.float32(sim.f) if np.asarray(sim.f).size == 1 else np.float32(0.0),
            "coriolis_beta": np.float32(sim.coriolis_beta),
            "y_zero_reference_cell": np.float32(sim.y_zero_reference_cell - 2),
            "eddy_viscosity_coefficient": np.float32(getattr(sim, "A", 0.0) or 0.0),
            "g": np.float32(sim.g),
            "nx": np.int32(nx), "ny": np.int32(ny),
            "dx": np.float32(dx), "dy": np.float32(dy), "dt": np.float32(sim.dt),
            "auto_dt": "False",
            "bottom_friction_r": np.float32(sim.r),
            "ghost_cells_north": np.int32(gy), "ghost_cells_east": np.int32(gx),
            "ghost_cells_south": np.int32(gy), "ghost_cells_west": np.int32(gx),
        }
        src = getattr(getattr(sim, "wind_stress", None), "source_filename", None)
        if src is not None:
            self.attrs["wind_stress_source"] = str(src)
        X, Y = (nx, ny) if ignore_ghostcells else (nx + 2 * gx, ny + 2 * gy)
        self.dims = [('x', X), ('y', Y), ('x_Hi', nx + 2 * gx + 1), ('y_Hi', ny + 2 * gy + 1)]
        axis = lambda name: {"standard_name": "projection_" + name + "_coordinate", "axis": name.upper(),
                             "units": 'meter'}
        if not ignore_ghostcells:
            xs = np.linspace(-gx * dx + dx / 2.0, (nx + gx) * dx - dx / 2.0, nx + 2 * gx)
            ys = np.linspace(-gy * dy + dy / 2.0, (ny + gy) * dy - dy / 2.0, ny + 2 * gy)
        else:
            xs = np.linspace(offset_x, nx * dx, nx)
            ys = np.linspace(offset_y, ny * dy, ny)
        proj = {"grid_mapping_name": 'polar_stereographic',
                "scale_factor_at_projection_origin": 0.9330127018922193,
                "straight_vertical_longitude_from_pole": 70.0, "latitude_of_projection_origin": 90.0,
                "earth_radius": 6371000.0,
                "proj4": '+proj=stere +lat_0=90 +lon_0=70 +lat_ts=60 +units=m +a=6.371e+06 +e=0 +no_defs'}
        # name -> (type code, dimensions, data, attributes)
        self.static = {
            'x': ('f', ('x',), xs, axis("x")),
            'y': ('f', ('y',), ys, axis("y")),
            'x_Hi': ('f', ('x_Hi',), np.linspace(-gx * dx, (nx + gx) * dx, nx + 2 * gx + 1), axis("x")),
            'y_Hi': ('f', ('y_Hi',), np.linspace(-gy * dy, (ny + gy) * dy, ny + 2 * gy + 1), axis("y")),
            'projection_stere': ('i', (), np.int32(0), proj),
            'land_binary_mask': ('f', ('y', 'x'), np.zeros((Y, X), np.float32),
                                 {"standard_name": 'land_binary_mask', "units": '1'}),
        }
        for name, arr, dims in (('Hm', Hm, ('y', 'x')), ('Hi', Hi, ('y_Hi', 'x_Hi'))):
            data, masked = _filled(arr)
            if name == 'Hm' and ignore_ghostcells:
                data = data[gy:-gy, gx:-gx]
            attrs = {"standard_name": 'water_surface_reference_datum_altitude', "grid_mapping": 'projection_stere',
                     "coordinates": 'y x', "units": 'meter'}
            if masked:
                attrs["_FillValue"] = NC_FILL_FLOAT
            self.static[name] = ('f', dims, data, attrs)
        self.record_attrs = {
            'time': {"units": 'seconds since 1970-01-01 00:00:00'},
            'eta': {"standard_name": 'water_surface_height_above_reference_datum', "grid_mapping": 'projection_stere',
                    "coordinates": 'y x', "units": 'meter'},
            'hu': {"standard_name": 'x_sea_water_velocity', "grid_mapping": 'projection_stere',
                   "coordinates": 'y_hu x_hu', "units": 'meter second-1'},
            'hv': {"standard_name": 'y_sea_water_velocity', "grid_mapping": 'projection_stere',
                   "coordinates": 'y_hv x_hv', "units": 'meter second-1'},
        }
        self.records = {'time': [], 'eta': [], 'hu': [], 'hv': []}
        self.ncfile = True   # open until close()
        self.i = 0
        self.writeTimestep(sim)

    def _write_file(self):
        nc = netcdf_file(self.output_file_name, 'w', version=2)
        for k, v in self.attrs.items():
            setattr(nc, k, v)
        nc.createDimension('time', len(self.records['time']))
        for name, size in self.dims:
            nc.createDimension(name, size)

        def put(name, code, dims, data, attrs):
            v = nc.createVariable(name, code, dims)
            for k, a in attrs.items():
                setattr(v, k, a)
            v[...] = data

        put('time', 'f', ('time',), np.asarray(self.records['time'], np.float32), self.record_attrs['time'])
        for name, (code, dims, data, attrs) in self.static.items():
            put(name, code, dims, data, attrs)
        for name in ('eta', 'hu', 'hv'):
            put(name, 'f', ('time', 'y', 'x'), np.stack(self.records[name]).astype(np.float32),
                self.record_attrs[name])
        nc.close()

    def writeModelError(self, sim):
        """Model-error parameters as global attributes (utils/SimWriter.py:388-420)."""
        self.attrs["model_time_step"] = np.float32(sim.model_time_step)
        has = sim.model_error is not None
        self.attrs["has_model_error"] = str(has)
        if has:
            self.attrs["model_error_name"] = sim.model_error.__class__.__name__
            self.attrs["small_scale_perturbation_amplitude"] = np.float32(sim.model_error.soar_q0)
            self.attrs["small_scale_perturbation_interpolation_factor"] = \
                np.int32(sim.model_error.interpolation_factor)
        self._write_file()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_value, traceback):
        self.close()

    def close(self):
        self.ncfile = None

    def writeTimestep(self, sim):
        """One record of (time, eta, hu, hv) from the simulator's latest state."""
        eta, hu, hv = sim.download(interior_domain_only=self.ignore_ghostcells)
        self.write(sim.t, eta, hu, hv)

    def write(self, t, eta, hu, hv, eta2=None, hu2=None, hv2=None):
        """One record of given fields (already shaped as the file's x, y)."""
        if self.ncfile is None:
            raise RuntimeError(f"{self.output_file_name} is closed")
        self.records['time'].append(np.float32(t))
        for name, a in (('eta', eta), ('hu', hu), ('hv', hv)):
            self.records[name].append(np.array(a, np.float32))
        self.i += 1
        self._write_file()
