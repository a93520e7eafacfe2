"""Binning of coordinate space (reference: ``triumvirate.dataobjs.Binning``).

The edge arithmetic restates ``Binning::compute_binning``
(src/triumvirate/src/dataobjs.cpp:134-250) operation by operation in IEEE
double precision: linear edges are ``bin_min + dbin * i`` with the last edge
set to ``bin_max`` exactly; centres are ``edge_left + dbin / 2``.
"""

import math

from .parameters import InvalidParameterError


class Binning:
    """Coordinate binning.

    Parameters
    ----------
    space : {'fourier', 'config'}
    scheme : {'lin', 'log', 'linpad', 'logpad', 'custom'}
    bin_min, bin_max : float, optional
    num_bins : int, optional
    """

    def __init__(self, space, scheme, bin_min=None, bin_max=None,
                 num_bins=None, boxsize_max=None, ngrid_min=None):
        if space not in ('fourier', 'config'):
            raise InvalidParameterError(f"Invalid binning space: '{space}'.")
        self.space = space
        self.scheme = scheme
        # Padding scales (dataobjs.cpp:43-53); only used by the padded
        # schemes.
        self.nbin_pad = 5
        if boxsize_max is not None and ngrid_min is not None:
            self.dbin_pad_config = (1 + 5.e-3) * boxsize_max / ngrid_min
            self.dbin_pad_fourier = (1 + 5.e-3) * (2 * math.pi) / boxsize_max
        else:
            self.dbin_pad_config = 10.
            self.dbin_pad_fourier = 1.e-3
        self.bin_edges, self.bin_centres, self.bin_widths = [], [], []
        if bin_min is not None and bin_max is not None and num_bins is not None:
            self.set_bins(bin_min, bin_max, num_bins)

    @classmethod
    def from_parameter_set(cls, paramset):
        box = paramset['boxsize']
        ngrid = paramset['ngrid']
        kw = {}
        if all(box.get(a) is not None for a in 'xyz') and \
                all(ngrid.get(a) is not None for a in 'xyz'):
            kw['boxsize_max'] = max(float(box[a]) for a in 'xyz')
            kw['ngrid_min'] = min(int(ngrid[a]) for a in 'xyz')
        b = cls(paramset['space'], paramset['binning'], **kw)
        b.set_bins(paramset['range'][0], paramset['range'][1],
                   paramset['num_bins'])
        return b

    def set_bins(self, bin_min, bin_max, num_bins):
        if bin_min < 0.:
            raise InvalidParameterError("Bin range must be non-negative.")
        if num_bins <= 0:
            raise InvalidParameterError("Number of bins must be positive.")
        self.bin_min = float(bin_min)
        self.bin_max = float(bin_max)
        self.num_bins = int(num_bins)
        self._compute_binning()

    def set_custom_bins(self, bin_edges):
        edges = [float(e) for e in bin_edges]
        self.scheme = 'custom'
        self.bin_min, self.bin_max = edges[0], edges[-1]
        self.num_bins = len(edges) - 1
        self.bin_edges = edges
        self.bin_centres = [(edges[i] + edges[i + 1]) / 2.
                            for i in range(self.num_bins)]
        self.bin_widths = [edges[i + 1] - edges[i]
                           for i in range(self.num_bins)]

    def _compute_binning(self):
        edges, centres, widths = [], [], []
        nb, lo, hi = self.num_bins, self.bin_min, self.bin_max
        pad = (self.dbin_pad_fourier if self.space == 'fourier'
               else self.dbin_pad_config)
        if self.scheme == 'lin':
            dbin = (hi - lo) / float(nb)
            for i in range(nb):
                left = lo + dbin * i
                edges.append(left)
                centres.append(left + dbin / 2.)
                widths.append(dbin)
            edges.append(hi)
        elif self.scheme == 'log':
            if lo == 0.:
                raise InvalidParameterError(
                    "Cannot use logarithmic binning when the lowest edge "
                    "is zero.")
            dln = (math.log(hi) - math.log(lo)) / float(nb)
            for i in range(nb):
                left = lo * math.exp(dln * i)
                right = lo * math.exp(dln * (i + 1))
                edges.append(left)
                centres.append((left + right) / 2.)
                widths.append(right - left)
            edges.append(hi)
        elif self.scheme in ('linpad', 'logpad'):
            for i in range(self.nbin_pad):
                left = pad * i
                edges.append(left)
                centres.append(left + pad / 2.)
                widths.append(pad)
            bmin = pad * self.nbin_pad
            nrest = nb - self.nbin_pad
            if self.scheme == 'linpad':
                dbin = (hi - bmin) / float(nrest)
                for i in range(self.nbin_pad, nb):
                    left = bmin + dbin * (i - self.nbin_pad)
                    edges.append(left)
                    centres.append(left + dbin / 2.)
                    widths.append(dbin)
            else:
                dln = (math.log(hi) - math.log(bmin)) / float(nrest)
                for i in range(self.nbin_pad, nb):
                    left = bmin * math.exp(dln * (i - self.nbin_pad))
                    right = bmin * math.exp(dln * (i - self.nbin_pad + 1))
                    edges.append(left)
                    centres.append((left + right) / 2.)
                    widths.append(right - left)
            edges.append(hi)
        else:
            raise InvalidParameterError(
                f"Unrecognised/unsupported binning `scheme`: {self.scheme}.")
        self.bin_edges, self.bin_centres, self.bin_widths = \
            edges, centres, widths

    def __repr__(self):
        return (f"Binning(space={self.space!r}, scheme={self.scheme!r}, "
                f"bin_min={self.bin_min}, bin_max={self.bin_max}, "
                f"num_bins={self.num_bins})")
