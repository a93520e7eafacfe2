"""Particle catalogue and box geometry
(reference: ``triumvirate.catalogue.ParticleCatalogue``).

Columns are kept as float64 numpy arrays (x, y, z, nz, ws, wc).  The
geometry operations restate the reference's arithmetic:

* ``compute_los``: unit vectors, zero norm replaced by 1
  (catalogue.py:437-458);
* ``centre``: shift by mean(ref bounds) - L/2 (catalogue.py:490-545);
* ``pad``: shift so the reference minimum sits at the pad
  (catalogue.py:547-646);
* ``periodise``: ``(x + L/2 - (max + min)/2) % L`` (catalogue.py:648-676);
* ``offset_coords``: subtract the origin (catalogue.py:678-690).
"""

import warnings

import numpy as np


def _header_names(filepath, format):
    """Column names from a first line '# x y z ...' of a plain-text file
    whose format is not 'ascii.no_header'."""
    if format == 'ascii.no_header':
        return None
    with open(filepath) as f:
        first = f.readline().strip()
    if not first.startswith('#'):
        return None
    names = first.lstrip('#').replace(',', ' ').split()
    return names or None


class ParticleCatalogue:
    """Particle catalogue.

    Parameters
    ----------
    x, y, z : array of float
        Cartesian coordinates.
    nz : float or array of float, optional
        Redshift-dependent background number density (default 0, with a
        warning, as the reference).
    ws, wc : float or array of float, optional
        Sample and clustering weights (default 1).
    """

    def __init__(self, x, y, z, nz=None, ws=1., wc=1., logger=None):
        self._logger = logger
        x = np.array(x, dtype=np.float64, copy=True).ravel()
        n = len(x)
        if nz is None:
            warnings.warn(
                "Catalogue 'nz' field is None and thus set to zero, "
                "which may raise errors in some computations.")
            nz = 0.

        def col(v):
            a = np.array(v, dtype=np.float64)
            if a.ndim == 0:
                return np.full(n, float(a))
            a = a.ravel().copy()
            if len(a) != n:
                raise ValueError(
                    "Inconsistent particle data dimensions (source=extdata).")
            return a

        self._pdata = {
            'x': x, 'y': col(y), 'z': col(z),
            'nz': col(nz), 'ws': col(ws), 'wc': col(wc),
        }
        # Weight columns built from the default 1 are kept read-only, so
        # that they stay identically 1 and need not be sent to the engine.
        self._unit = {}
        for key, v in (('ws', ws), ('wc', wc)):
            self._set_unit(key, v)
        self._source = f'extdata:{id(self._pdata)}'
        self._offset = np.zeros(3)
        self._calc_bounds()
        self.ntotal = n
        self.wtotal = (self._pdata['ws'] * self._pdata['wc']).sum()
        self.wstotal = self._pdata['ws'].sum()

    @classmethod
    def read_from_file(cls, filepath, reader='astropy', names=None,
                       format='ascii.no_header', name_mapping=None,
                       table_kwargs=None, file_kwargs=None, logger=None):
        """Read a catalogue file (catalogue.py:150-388).

        Plain-text formats ('ascii...', 'text') are read with numpy:
        `names` names the columns in file order (default x, y, z, nz,
        ws, wc for as many columns as the file holds; a header line
        starting with '#' may name them instead).  The 'binary' format
        reads fixed-size records whose field names and types `names`
        gives as a structured dtype (a list of (name, type) pairs), as
        nbodykit's binary reader does.  `name_mapping` maps the default
        names to file names.  FITS and HDF5 need astropy or nbodykit,
        which this package does not depend on, and raise.  The catalogue
        source becomes ``extfile:<filepath>``.
        """
        if reader.lower() not in ('astropy', 'nbodykit'):
            raise ValueError(
                f"Unsupported catalogue file reader: {reader}. "
                "Possible options: {'astropy', 'nbodykit'}.")
        if format == 'binary':
            if names is None:
                raise ValueError(
                    "Binary catalogues need `names` as a structured dtype.")
            rec = np.fromfile(filepath, dtype=np.dtype(list(names)),
                              **(file_kwargs or {}))
            cols = {name: np.asarray(rec[name], dtype=np.float64)
                    for name in rec.dtype.names}
        elif format.startswith('ascii') or format == 'text':
            data = np.loadtxt(filepath, ndmin=2, **(table_kwargs or {}))
            if names is None:
                names = _header_names(filepath, format) or \
                    ('x', 'y', 'z', 'nz', 'ws', 'wc')[:data.shape[1]]
            names = list(names)
            if len(names) != data.shape[1]:
                raise ValueError(
                    f"{len(names)} column names for {data.shape[1]} columns "
                    f"in {filepath}.")
            cols = {name: data[:, i] for i, name in enumerate(names)}
        else:
            raise ValueError(
                f"Unsupported catalogue file format without astropy: {format}.")
        for name, alt in (name_mapping or {}).items():
            cols[name] = cols.pop(alt)
        for ax in ('x', 'y', 'z'):
            if ax not in cols:
                raise ValueError(f"Mandatory field {ax} is missing.")
        if 'nz' not in cols:
            warnings.warn(
                "Catalogue 'nz' field is not provided and thus set to zero, "
                "which may raise errors in some computations.")
            cols['nz'] = 0.
        for wname in ('ws', 'wc'):
            if wname not in cols:
                warnings.warn(
                    f"Catalogue '{wname}' field is not provided, "
                    "so is set to unity.")
        self = cls(cols['x'], cols['y'], cols['z'], nz=cols['nz'],
                   ws=cols.get('ws', 1.), wc=cols.get('wc', 1.),
                   logger=logger)
        self._source = f'extfile:{filepath}'
        return self

    def write_attrs_as_header(self, catalogue_ref=None):
        """Catalogue source, size and extents as header lines
        (catalogue.py:803-865); with `catalogue_ref` (the random-source
        catalogue) both are written, as data and random."""
        def lines(cat, source_label, size_label, extent_label):
            b = cat.bounds
            return [
                f"{source_label}: {cat._source}",
                "{}: ntotal = {:d}, wtotal = {:.3f}, wstotal = {:.3f}".format(
                    size_label, cat.ntotal, cat.wtotal, cat.wstotal),
                "{}: ([{:.3f}, {:.3f}], [{:.3f}, {:.3f}], [{:.3f}, {:.3f}])"
                .format(extent_label, *b['x'], *b['y'], *b['z']),
            ]
        if catalogue_ref is None:
            text = lines(self, "Catalogue source", "Catalogue size",
                         "Catalogue particle extents")
        else:
            text = lines(self, "Data catalogue source", "Data catalogue size",
                         "Data-source particle extents") + \
                lines(catalogue_ref, "Random catalogue source",
                      "Random catalogue size",
                      "Random-source particle extents")
        return "\n".join(text)

    def __len__(self):
        return self.ntotal

    def __getitem__(self, key):
        return self._pdata[key]

    def _set_unit(self, key, value):
        unit = np.ndim(value) == 0 and float(value) == 1.
        self._unit[key] = unit
        self._pdata[key].flags.writeable = not unit

    def __setitem__(self, key, value):
        self._pdata[key] = np.broadcast_to(
            np.asarray(value, dtype=np.float64), (self.ntotal,)).copy()
        if key in ('ws', 'wc'):
            self._set_unit(key, value)
        if key in ('x', 'y', 'z'):
            self._calc_bounds()
        if key in ('ws', 'wc'):
            self.wtotal = (self._pdata['ws'] * self._pdata['wc']).sum()
            self.wstotal = self._pdata['ws'].sum()

    def __str__(self):
        return f"ParticleCatalogue(source={self._source})"

    def _calc_bounds(self):
        self.bounds = {
            ax: (self._pdata[ax].min(), self._pdata[ax].max())
            for ax in ('x', 'y', 'z')
        }

    def _check_bounds_in_boxsize(self, boxsize):
        return [ax for i, ax in enumerate('xyz')
                if abs(self.bounds[ax][1] - self.bounds[ax][0]) > boxsize[i]]

    def compute_los(self):
        """Unit line-of-sight vectors, shape (N, 3)."""
        norm = np.sqrt(self._pdata['x']**2 + self._pdata['y']**2
                       + self._pdata['z']**2)
        norm[norm == 0.] = 1.
        return np.ascontiguousarray(np.transpose([
            self._pdata['x'] / norm,
            self._pdata['y'] / norm,
            self._pdata['z'] / norm,
        ]))

    def compute_mean_density(self, volume=None, boxsize=None):
        if np.isscalar(boxsize):
            boxsize = [boxsize, boxsize, boxsize]
        volume = volume or np.prod(boxsize)
        self._pdata['nz'] = np.full(self.ntotal, self.ntotal / volume)

    def offset_coords(self, origin):
        for ax, c in zip('xyz', origin):
            self._pdata[ax] = self._pdata[ax] - c
        self._calc_bounds()
        # Total displacement applied since construction: the original
        # coordinates are the current ones plus this vector.
        self._offset = self._offset + np.asarray(origin, dtype=np.float64)

    def centre(self, boxsize, catalogue_ref=None):
        if np.isscalar(boxsize):
            boxsize = [boxsize, boxsize, boxsize]
        ref = self if catalogue_ref is None else catalogue_ref
        if self._check_bounds_in_boxsize(boxsize):
            warnings.warn(
                "Catalogue extent exceeds the box size. Some particles may "
                "lie outside the box after centring.")
        origin = np.array([
            np.mean(ref.bounds[ax]) - boxsize[i] / 2.
            for i, ax in enumerate('xyz')
        ])
        self.offset_coords(origin)
        if catalogue_ref is not None:
            catalogue_ref.offset_coords(origin)

    def pad(self, boxsize, ngrid=None, boxsize_pad=None, ngrid_pad=None,
            catalogue_ref=None):
        if boxsize_pad is None and ngrid_pad is None:
            warnings.warn(
                "`boxsize_pad` and `ngrid_pad` are both None. "
                "No padding is applied to the catalogue.")
            return
        if boxsize_pad is not None and ngrid_pad is not None:
            raise ValueError(
                "Conflicting padding as `boxsize_pad` and `ngrid_pad` "
                "are both set (not None).")
        if np.isscalar(boxsize):
            boxsize = [boxsize, boxsize, boxsize]
        ref = self if catalogue_ref is None else catalogue_ref
        origin = np.array([ref.bounds[ax][0] for ax in 'xyz'])
        if boxsize_pad:
            origin -= np.multiply(boxsize_pad, boxsize)
        if ngrid_pad:
            origin -= np.multiply(ngrid_pad, np.divide(boxsize, ngrid))
        self.offset_coords(origin)
        if catalogue_ref is not None:
            catalogue_ref.offset_coords(origin)

    def periodise(self, boxsize):
        if np.isscalar(boxsize):
            boxsize = [boxsize, boxsize, boxsize]
        for i, ax in enumerate('xyz'):
            self._pdata[ax] = (
                self._pdata[ax] + boxsize[i] / 2.
                - (self.bounds[ax][1] + self.bounds[ax][0]) / 2.
            ) % boxsize[i]
        self._calc_bounds()

    def columns(self):
        """(x, y, z, nz, ws, wc) float64 arrays."""
        p = self._pdata
        return p['x'], p['y'], p['z'], p['nz'], p['ws'], p['wc']

    def engine_columns(self):
        """As `columns`, with None for a weight column that is identically 1
        (the engine takes an absent column as unit weights and then neither
        reads nor transfers it)."""
        p = self._pdata
        return (p['x'], p['y'], p['z'], p['nz'],
                None if self._unit['ws'] else p['ws'],
                None if self._unit['wc'] else p['wc'])
