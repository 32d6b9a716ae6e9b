"""Hydrometeor descriptor table, the interface of
python/pyPamtra/descriptorFile.py (class pamDescriptorFile, :12-175).

A table row has 20 columns (descriptorFile.py:19-20); ``data4D`` replaces a
column by a (x, y, z, hydro) field (``add4D``, :125-136).
"""

from __future__ import annotations

import csv

import numpy as np
from numpy.lib import recfunctions as rfn

missingNumber = -9999.
missingIntNumber = int(missingNumber)

NAMES = ["hydro_name", "as_ratio", "liq_ice", "rho_ms", "a_ms", "b_ms", "alpha_as", "beta_as",
         "moment_in", "nbin", "dist_name", "p_1", "p_2", "p_3", "p_4", "d_1", "d_2",
         "scat_name", "vel_size_mod", "canting"]
TYPES = ["U15", float, int, float, float, float, float, float, int, int, "U15", float, float,
         float, float, float, float, "U60", "U30", float]


class pamDescriptorFile(object):
    """Descriptor table of a pyPamtra object (descriptorFile.py:12-28)."""

    def __init__(self, parent):
        self.names = np.array(NAMES)
        self.types = list(TYPES)
        self.data = np.recarray((0,), dtype=list(zip(self.names, self.types)))
        self.data4D = dict()
        self.nhydro = 0
        self.parent = parent
        self.dataFullSpec = dict()
        self.fs_nbin = 0

    def readFile(self, fname):
        """Reads an ASCII descriptor file: space separated, '!' comments,
        Fortran 'd' exponents and quoted strings (descriptorFile.py:31-58)."""
        with open(fname, "r") as f:
            for row in csv.reader(f, delimiter=" ", skipinitialspace=True):
                if not row or row[0].startswith("!"):
                    continue
                row = [r for r in row if r != ""]
                assert len(row) == len(self.names)
                vals = []
                for item, typ in zip(row, self.types):
                    if typ is float:
                        vals.append(item.replace("d", "e").replace("D", "e"))
                    elif typ in ("U15", "S15"):
                        vals.append(item.replace("'", "").replace('"', ""))
                    else:
                        vals.append(item)
                self.addHydrometeor(vals)

    def writeFile(self, fname):
        """Writes the table as space separated ASCII (descriptorFile.py:60-71)."""
        assert len(self.names) == len(self.data.dtype.names)
        with open(fname, "w") as f:
            w = csv.writer(f, delimiter=" ")
            for row in self.data:
                w.writerow(list(row))

    def addHydrometeor(self, hydroTuple):
        """Appends one hydrometeor row (descriptorFile.py:73-95) and grows the
        parent's hydro_q/hydro_n/hydro_reff arrays by a missing column."""
        assert hydroTuple[0] not in self.data["hydro_name"]
        assert len(self.dataFullSpec) == 0
        row = np.array(tuple(hydroTuple), dtype=list(zip(self.names, self.types)))
        self.data = np.append(self.data, row).view(np.recarray)
        self.nhydro += 1
        par = self.parent
        par._shape4D = (par.p["ngridx"], par.p["ngridy"], par.p["max_nlyrs"], self.nhydro)
        for key in ["hydro_q", "hydro_reff", "hydro_n"]:
            if key in par.p:
                par.p[key] = np.concatenate(
                    [par.p[key], np.ones(par._shape3D + (1,)) * missingNumber], axis=-1)

    def removeHydrometeor(self, hydroName):
        """Removes a row by name, or all rows with "all" (descriptorFile.py:97-123)."""
        assert len(self.dataFullSpec) == 0
        if hydroName == "all":
            self.__init__(self.parent)
            return
        idx = np.where(self.data["hydro_name"] == hydroName)[0]
        if len(idx) == 0:
            raise ValueError("Did not find " + hydroName)
        i = int(idx[0])
        self.data = np.delete(self.data, i).view(np.recarray)
        self.nhydro -= 1
        par = self.parent
        par._shape4D = (par.p["ngridx"], par.p["ngridy"], par.p["max_nlyrs"], self.nhydro)
        for key in ["hydro_q", "hydro_reff", "hydro_n"]:
            if key in par.p:
                par.p[key] = np.delete(par.p[key], i, axis=-1)

    def add4D(self, key, arr):
        """Replaces a descriptor column by an (x, y, z, hydro) field
        (descriptorFile.py:125-136)."""
        self.data = rfn.drop_fields(self.data, [key], usemask=False, asrecarray=True)
        self.data4D[key] = np.asarray(arr).reshape(self.parent._shape4D)

    def remove4D(self, key, val):
        """Restores a 1D column from a 4D field (descriptorFile.py:138-158)."""
        dtype = self.types[list(self.names).index(key)]
        self.data = rfn.append_fields(self.data, key, val, dtypes=dtype, usemask=False,
                                      asrecarray=True)
        del self.data4D[key]

    def addFullSpectra(self):
        """Creates the full-spectrum arrays (x, y, z, hydrometeor, bin) of
        nbin bins, nbin = the descriptor's (equal) bin count; d_bound_ds has
        nbin + 1 edges (descriptorFile.py:173-236).  Fill them, then set
        nmlSet["hydro_fullspec"] = True: the given bins replace the size
        distributions.  The self-similar Rayleigh-Gans parameters default
        to kappa 0.19, beta 0.23, gamma 5/3, zeta 1."""
        assert len(self.data) > 0
        assert np.min(self.data["nbin"]) == np.max(self.data["nbin"])
        self.fs_nbin = int(np.max(self.data["nbin"]))
        par = self.parent
        s4 = (par.p["ngridx"], par.p["ngridy"], par.p["max_nlyrs"], self.nhydro)
        par._shape5Dplus = s4 + (self.fs_nbin + 1,)
        par._shape5D = s4 + (self.fs_nbin,)
        for key in ("rho_ds", "d_ds", "n_ds", "mass_ds", "area_ds", "as_ratio", "canting",
                    "fallvelocity"):
            self.dataFullSpec[key] = np.zeros(par._shape5D)
        self.dataFullSpec["d_bound_ds"] = np.zeros(par._shape5Dplus)
        for key, v in (("rg_kappa_ds", 0.19), ("rg_beta_ds", 0.23), ("rg_gamma_ds", 5. / 3.),
                       ("rg_zeta_ds", 1.)):
            self.dataFullSpec[key] = np.full(par._shape5D, v)
