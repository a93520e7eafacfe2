"""On-disk formats of measurements (reference threept.py:158-384,
twopt.py:158-320, catalogue.py:803-865).

A measurement is saved as a tab-separated text table with a '#'-prefixed
header (catalogue attributes, sampling parameters, normalisation factors,
column names), or as an .npz archive holding the result arrays plus the
header string.  File names follow the reference: ``pk{ELL}``,
``xi{ELL}``, ``xiw{ELL}`` for two-point statistics, ``bk{l1l2L}_{form}``,
``zeta...``, ``zetaw...`` (``_wa{ij}`` for wide-angle windows) for
three-point statistics, each followed by the output tag.
"""

from pathlib import Path

import numpy as np


def _interlace_text(paramset):
    # The validated parameter set holds interlacing as a bool.
    return str(bool(paramset['interlace']))


def _three_point_form_tag(paramset):
    form = paramset['form']
    if form == 'diag':
        return '_diag'
    if form == 'off-diag':
        return '_offdiag{:d}'.format(int(paramset['idx_bin']))
    if form == 'row':
        return '_row{:d}'.format(int(paramset['idx_bin']))
    return '_full'


def measurement_filename(paramset):
    """File stem of a measurement (threept.py:158-198, twopt.py:158-193)."""
    tag = (paramset['tags'] or {}).get('output') or ''
    stype = paramset['statistic_type']
    deg = paramset['degrees']
    if stype in ('powspec', '2pcf', '2pcf-win'):
        stem = {'powspec': 'pk', '2pcf': 'xi', '2pcf-win': 'xiw'}[stype]
        return '{}{:d}{}'.format(stem, int(deg['ELL']), tag)
    multipole = '{}{}{}'.format(deg['ell1'], deg['ell2'], deg['ELL'])
    form = _three_point_form_tag(paramset)
    if stype == 'bispec':
        return 'bk{}{}{}'.format(multipole, form, tag)
    if stype == '3pcf':
        return 'zeta{}{}{}'.format(multipole, form, tag)
    if stype == '3pcf-win':
        return 'zetaw{}{}{}'.format(multipole, form, tag)
    if stype == '3pcf-win-wa':
        wa = '_wa{i}{j}'.format(**paramset['wa_orders'])
        return 'zetaw{}{}{}{}'.format(multipole, form, wa, tag)
    raise ValueError(
        "`paramset` 'statistic_type' does not correspond to a recognised "
        "statistic.")


def _column_names(paramset):
    deg = paramset['degrees']
    if paramset['npoint'] == '2pt':
        ell = int(deg['ELL'])
        if paramset['space'] == 'fourier':
            return ['k_cen', 'k_eff', 'nmodes', f'Re{{pk{ell}_raw}}',
                    f'Im{{pk{ell}_raw}}', f'Re{{pk{ell}_shot}}',
                    f'Im{{pk{ell}_shot}}']
        return ['r_cen', 'r_eff', 'npairs', f'Re{{xi{ell}}}', f'Im{{xi{ell}}}']
    mp = '{}{}{}'.format(deg['ell1'], deg['ell2'], deg['ELL'])
    if paramset['space'] == 'fourier':
        return ['k1_cen', 'k1_eff', 'nmodes_1', 'k2_cen', 'k2_eff', 'nmodes_2',
                f'Re{{bk{mp}_raw}}', f'Im{{bk{mp}_raw}}', f'Re{{bk{mp}_shot}}',
                f'Im{{bk{mp}_shot}}']
    return ['r1_cen', 'r1_eff', 'npairs_1', 'r2_cen', 'r2_eff', 'npairs_2',
            f'Re{{zeta{mp}_raw}}', f'Im{{zeta{mp}_raw}}', f'Re{{zeta{mp}_shot}}',
            f'Im{{zeta{mp}_shot}}']


def measurement_header(paramset, norm_part, norm_mesh, norm_meshes=0.):
    """Sampling parameters, normalisation factors and column names
    (threept.py:256-327, twopt.py:195-275)."""
    conv = paramset['norm_convention']
    norm = {'none': 1., 'particle': norm_part, 'mesh': norm_mesh,
            'mesh-mixed': norm_meshes}[conv]
    box = [paramset['boxsize'][a] for a in 'xyz']
    ngrid = [paramset['ngrid'][a] for a in 'xyz']
    mixed = ('mesh-mixed' if paramset['npoint'] == '2pt'
             else 'mesh-mixed; n/a')
    lines = [
        'Box size: [{:.3f}, {:.3f}, {:.3f}]'.format(*box),
        'Box alignment: {}'.format(paramset['alignment']),
        'Mesh number: [{:d}, {:d}, {:d}]'.format(*ngrid),
        'Mesh assignment and interlacing: {}, {}'.format(
            paramset['assignment'], _interlace_text(paramset)),
        'Normalisation factor: {:.9e} ({})'.format(norm, conv),
        'Normalisation factor alternatives: {:.9e} (particle), {:.9e} (mesh), '
        '{:.9e} ({})'.format(norm_part, norm_mesh, norm_meshes, mixed),
        ', '.join('[{:d}] {}'.format(i, c)
                  for i, c in enumerate(_column_names(paramset))),
    ]
    return '\n'.join(lines)


def measurement_datatab(results, paramset):
    """Column-major data table (threept.py:330-384, twopt.py:277-320)."""
    if paramset['npoint'] == '2pt':
        if paramset['space'] == 'fourier':
            cols = [results['kbin'], results['keff'], results['nmodes'],
                    results['pk_raw'].real, results['pk_raw'].imag,
                    results['pk_shot'].real, results['pk_shot'].imag]
        else:
            cols = [results['rbin'], results['reff'], results['npairs'],
                    results['xi'].real, results['xi'].imag]
        return np.transpose(cols)
    if paramset['space'] == 'fourier':
        cols = [results['k1_bin'], results['k1_eff'], results['nmodes_1'],
                results['k2_bin'], results['k2_eff'], results['nmodes_2'],
                results['bk_raw'].real, results['bk_raw'].imag,
                results['bk_shot'].real, results['bk_shot'].imag]
    else:
        cols = [results['r1_bin'], results['r1_eff'], results['npairs_1'],
                results['r2_bin'], results['r2_eff'], results['npairs_2'],
                results['zeta_raw'].real, results['zeta_raw'].imag,
                results['zeta_shot'].real, results['zeta_shot'].imag]
    return np.transpose(cols)


def _text_format(paramset, ncols):
    if paramset['npoint'] == '2pt':
        return '\t'.join(['%.9e'] * 2 + ['%10d'] + ['% .9e'] * (ncols - 3))
    return '\t'.join((['%.9e'] * 2 + ['%10d']) * 2 + ['% .9e'] * (ncols - 6))


def save_measurements(results, paramset, header, save):
    """Write `results` as '.txt' or '.npz' into the parameter set's
    measurement directory (the working directory if unset); returns the
    path written.  For '.npz' the header is also added to `results`, as
    the reference does (threept.py:727-731)."""
    fmt = str(save).lower()
    directory = (paramset['directories'] or {}).get('measurements') or ''
    stem = measurement_filename(paramset)
    if fmt == '.txt':
        path = Path(directory, stem).with_suffix('.txt')
        tab = measurement_datatab(results, paramset)
        np.savetxt(path, tab, fmt=_text_format(paramset, tab.shape[-1]),
                   header=header, delimiter='\t')
        return path
    if fmt.endswith('.npz'):
        results['header'] = header
        path = Path(directory, stem).with_suffix('.npz')
        np.savez(path, **{k: v for k, v in results.items()
                          if not k.startswith('_')})
        return path
    raise ValueError(f"Unrecognised save format for measurements: {save}.")


def record(results, paramset, catalogue_header, norm_factors, save,
           logger=None):
    """Save a driver's `results` when `save` is '.txt' or '.npz'.

    `catalogue_header` is the catalogues' ``write_attrs_as_header`` text and
    `norm_factors` the (particle, mesh[, mixed-mesh]) normalisations; the
    mixed-mesh one is 0 where it does not apply.
    """
    if not save:
        return None
    part, mesh, meshes = (tuple(norm_factors) + (0.,))[:3]
    header = '\n'.join([catalogue_header,
                        measurement_header(paramset, part, mesh, meshes)])
    path = save_measurements(results, paramset, header, save)
    if logger:
        logger.info("Measurements saved to %s.", path)
    return path
