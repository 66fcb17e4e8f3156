"""Reader of CDKLM16 netCDF output (utils/SimReader.py).

``SimNetCDFReader`` reads classic netCDF files (those ``SimNetCDFWriter``
writes, and any netCDF-3 file with the reference's variables) through
``scipy.io.netcdf_file``.  Files in the HDF5-based netCDF-4 format, which the
reference's netCDF4 library writes by default, need that library, which is not
available here; opening one raises an error that says so.  Numeric global
attributes come back as NumPy scalars and text ones as str, as netCDF4's
``getncattr`` returns them; variables with a ``_FillValue`` come back masked
where they hold it.
"""

import numpy as np
from scipy.io import netcdf_file

from . import Common


def _check_classic(filename):
    with open(filename, 'rb') as f:
        head = f.read(8)
    if head[:3] == b'CDF':
        return
    if head[:8] == b'\x89HDF\r\n\x1a\n':
        raise RuntimeError(f"{filename} is a netCDF-4 (HDF5) file: reading it needs the netCDF4 package, which is "
                           "not available; write it as classic netCDF (e.g. nccopy -k classic) to read it here")
    raise RuntimeError(f"{filename} is not a netCDF file")


class SimNetCDFReader:
    """Attributes, time series and bathymetry of one run (utils/SimReader.py:37-170)."""

    def __init__(self, filename, ignore_ghostcells=True):
        _check_classic(filename)
        self.filename = filename
        self.ignore_ghostcells = ignore_ghostcells
        self.ncfile = netcdf_file(filename, 'r', mmap=False)
        self.ghostCells = [int(self.get('ghost_cells_north')), int(self.get('ghost_cells_east')),
                           int(self.get('ghost_cells_south')), int(self.get('ghost_cells_west'))]
        self.staggered_grid = str(self.get('staggered_grid')) == 'True'
        self.text_font_size = 12

    def close(self):
        if self.ncfile is not None:
            self.ncfile.close()
            self.ncfile = None

    def get(self, attr):
        """A global attribute, or "not found"."""
        if attr not in self.ncfile._attributes:
            return "not found"
        v = self.ncfile._attributes[attr]
        if isinstance(v, bytes):
            return v.decode()
        if isinstance(v, np.ndarray):
            return v[0] if v.size == 1 else v
        return v

    def has(self, fieldname):
        return fieldname in self.ncfile._attributes or fieldname in self.ncfile.variables

    def printVariables(self):
        for var in self.ncfile.variables:
            print(var)

    def printAttributes(self):
        for attr in self.ncfile._attributes:
            print(attr + "\t--> " + str(self.get(attr)))

    def _var(self, name, index=None):
        v = self.ncfile.variables[name]
        data = np.array(v[:] if index is None else v[index])
        fill = v._attributes.get('_FillValue')
        if fill is not None:
            fill = np.asarray(fill).reshape(-1)[0]
            return np.ma.masked_equal(data, fill)
        return data

    def getNumTimeSteps(self):
        return self.ncfile.variables['time'].shape[0]

    def getBC(self):
        return Common.BoundaryConditions.fromstring(self.get("boundary_conditions"))

    def getTimes(self):
        return self._var('time')

    def getLastTimeStep(self):
        return self.getTimeStep(-1)

    def _interior(self, a):
        g = self.ghostCells
        return a[g[2]:-g[0], g[3]:-g[1]]

    def getTimeStep(self, index):
        eta, hu, hv = (self._var(k, index) for k in ('eta', 'hu', 'hv'))
        if self.ignore_ghostcells:
            eta, hu, hv = self._interior(eta), self._interior(hu), self._interior(hv)
        return eta, hu, hv, np.float32(self.ncfile.variables['time'][index])

    def getH(self):
        if self.has('H') and 'H' in self.ncfile.variables:
            return self._var('H')
        return self._var('Hm' if self.staggered_grid else 'Hi')

    def getHm(self):
        return self._var('Hm')

    def getStateAtTime(self, time):
        time = np.round(time)
        times = self.ncfile.variables['time']
        for i in range(times.shape[0]):
            if time == times[i]:
                return self.getStateAtTimeStep(i)
        raise RuntimeError('Time ' + str(time) + ' not in NetCDF file ' + self.filename)

    def getStateAtTimeStep(self, index, etaOnly=False):
        t = self.ncfile.variables['time'][index]
        eta = self._var('eta', index)
        if self.ignore_ghostcells:
            eta = self._interior(eta)
        if etaOnly:
            return eta, t
        hu, hv = self._var('hu', index), self._var('hv', index)
        if self.ignore_ghostcells:
            hu, hv = self._interior(hu), self._interior(hv)
        return eta, hu, hv, t

    def getEtaAtTimeStep(self, index):
        return self.getStateAtTimeStep(index, etaOnly=True)

    def getAxis(self):
        x, y = self._var('x'), self._var('y')
        if self.ignore_ghostcells:
            x = x[self.ghostCells[2]:-self.ghostCells[0]]
            y = y[self.ghostCells[3]:-self.ghostCells[1]]
        return x, y

    def getEtaXSlice(self, t, y):
        y_index = int(y) + int(self.get('ghost_cells_south'))
        return self._var('eta', t)[y_index, self.ghostCells[3]:-self.ghostCells[1]]
